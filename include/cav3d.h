/*
 * cav3d — MI355X (gfx950) implementation of the pseudo-time iteration of the
 * all-speed two-phase solver: RHS residual (MUSCL/MLP, IAPWS-IF97 EOS,
 * preconditioned Roe / Roe-M / AUSMPW+ / AUSM+-up, viscous flux, k-omega
 * SST with DES variants, Merkle / Kunz / Singhal / Schnerr-Sauer / Lee mass
 * transfer) plus the implicit LU-SGS update, with every grid block resident on
 * one GPU and block-interface halos filled by device copy kernels.
 *
 * C ABI: plain pointers and sizes.  Every call returns 0 on success and a
 * negative code on error; cav3d_last_error() gives the message.  Host arrays
 * use the reference's Fortran layouts (component fastest, column-major,
 * explicit lower bounds) and stay owned by the caller; device buffers are
 * owned by the context.  A context is used from one host thread at a time.
 *
 * Reference interfaces replaced (kwonhwi1/3dcode):
 *   cav3d_timeinteg      class(t_update)%timeinteg -> lusgs
 *                        src/update_module.f90:37-52 (interface), :455-733
 *                        (called at src/solve_module.f90:135)
 *   cav3d_calrhs         t_rhs%calrhs + getres/geticav/getitt/getomega_cut
 *                        src/rhs_module.f90:260-683
 *   cav3d_setbc          t_bc%setbc src/bc_module.f90:1235-1396
 *   cav3d_residual       t_resi%setqres + t_resi%residual
 *                        src/residual_module.f90:94-186
 *   cav3d_unsteady_shift t_solve%unsteadyupdate src/solve_module.f90:155-183
 *   cav3d_put_state /    t_variable getpv/getdv/gettv/getqq + setters
 *   cav3d_get_state      src/variable_module.f90:16-21, 69-77
 *   cav3d_set_metrics    t_grid getcx/getex/gettx/getgrd src/grid_module.f90:37-39
 *   cav3d_integrals      t_post vof fields + clcd_writing
 *                        src/postvariable_module.f90:282-295, 507-633
 *   cav3d_export /       t_variable%export_variable src/variable_module.f90:108-182,
 *   cav3d_import         t_ini_restart%restart src/initial_module.f90:245-389
 *   cav3d_set_grid /     t_grid%construct: calnormal, calvolume, calydns
 *   cav3d_build_metrics  src/grid_module.f90:75-332, 350-1145
 *   cav3d_get_metrics    t_grid getcx/getex/gettx/getgrd
 *   cav3d_create         t_update%construct + t_bc%construct (BC faces, 12
 *                        edges, 8 corners, interfaces)
 *                        src/update_module.f90:77-178, src/bc_module.f90:71-1188
 *   cav3d_eos            t_eos%deteos (deteos_prop / deteosonly)
 *                        src/eos_module.f90:1120-1234, evaluated on the GPU
 *   cav3d_eos_host       the same equation of state on the host (set-up use)
 *   cav3d_rhs_vjp        no reference counterpart (discrete adjoint of calrhs)
 */
#ifndef CAV3D_H
#define CAV3D_H

#ifdef __cplusplus
extern "C" {
#endif

/* physical boundary kinds (reference bcname strings, src/bc_module.f90:221-335) */
enum {
    CAV3D_BC_WALL = 1,              /* 'BCWall'  (viscous / kw / inviscid by iturb) */
    CAV3D_BC_INFLOW_SUBSONIC = 2,   /* 'BCInflowSubsonic' */
    CAV3D_BC_OUTFLOW_SUBSONIC = 3,  /* 'BCOutflowSubsonic' */
    CAV3D_BC_FARFIELD = 4,          /* 'BCFarfield' */
    CAV3D_BC_SYMMETRY = 5,          /* 'BCSymmetryPlane' */
    CAV3D_BC_INFLOW_SUPERSONIC = 6, /* 'BCInflowSupersonic' */
    CAV3D_BC_OUTFLOW_SUPERSONIC = 7,/* 'BCOutflowSupersonic' / 'BCExtrapolate' */
    CAV3D_BC_DEGENERATE = 8,        /* 'BCDegeneratePoint' / 'BCDegenerateLine' */
    CAV3D_BC_INFLOW = 9,            /* 'BCInflow'  (no-op, as in the reference) */
    CAV3D_BC_OUTFLOW = 10           /* 'BCOutflow' (no-op, as in the reference) */
};
/* A boundary range of the reference's 'Surface' family (the force-integration
 * surface of src/postvariable_module.f90:541) carries this flag OR'ed into its
 * kind; the kind itself is kind & CAV3D_BC_KIND_MASK. */
enum { CAV3D_BC_KIND_MASK = 0xff, CAV3D_BC_SURFACE = 0x100 };

enum { CAV3D_FP64 = 0, CAV3D_MIXED = 1 };

/* Upper clip of the Merkle rates.  src/cav_module.f90:151-152 clips
 * r_c <= rho*pv(5) and r_v <= rho*(1-pv(5)-pv(6)), where pv(5) is the
 * temperature; with T > 1 K the second bound is negative, so every cell gets
 * r_v = rho*(1-T-y1) and a cavitating run diverges within a few iterations.
 * AS_WRITTEN reproduces that bit for bit; SPECIES clips on the mass
 * fractions, r_c <= rho*y1 and r_v <= rho*(1-y1-y2). */
enum { CAV3D_CLIP_AS_WRITTEN = 0, CAV3D_CLIP_SPECIES = 1 };

/* Run parameters: the config_module fields the path reads
 * (src/config_module.f90:7-28, 90-199), integer codes unchanged. */
typedef struct {
    int iturb;      /* -3 Euler, -2 laminar, 0 k-omega SST */
    int nscheme;    /* 1 preconditioned Roe, 2 Roe-M, 3 AUSMPW+, 4 AUSM+-up */
    int nmuscl;     /* 0 first order, 1 MUSCL-TVD, 2 MLP */
    int nlim;       /* limiter 0..10 */
    int local;      /* -1 fixed, 0 global-min, 1 local time step */
    int prec;       /* 0 none, 1 steady, 2 unsteady preconditioning */
    int ncav;       /* 0 none, 1 Merkle, 2 Kunz, 3 Singhal, 4 Schnerr-Sauer, 5 Lee */
    int nsteady;    /* 0 steady, 1 BDF1, 2 BDF2 dual time */
    int precision;  /* CAV3D_FP64 or CAV3D_MIXED */
    int cav_clip;   /* CAV3D_CLIP_AS_WRITTEN or CAV3D_CLIP_SPECIES */
    int ides;       /* SST length scale: 0 RANS, 1 DES, 2 DDES, 3 IDDES, 4 S-IDDES */
    int reserved;
    double cfl, pref, uref, str, dt_phy, c_v, c_c;
    double dp_ref;  /* 2/(rho_ref*uref^2)  (src/cav_module.f90:81) */
    double t_ref;   /* uref/l_chord        (src/cav_module.f90:82) */
    double t_crit;  /* 647.096 for water   (src/cav_module.f90:86-95) */
    double bc_pv[9];    /* boundary reference state bcinfo%pv, bc_pv[0] absolute */
    double bc_pressure; /* bcinfo%pressure */
    double emutref;     /* initial eddy viscosity */
} cav3d_params;

/* One grid block.  imax/jmax/kmax are vertex counts (CGNS convention,
 * src/grid_module.f90:107-114); cells are 2..imax etc. */
typedef struct {
    int imax, jmax, kmax;
    int nbc;
    const int *bc;   /* nbc x 7: kind, istart[3], iend[3] as grid_module gives them */
    int ncon;
    const int *con;  /* ncon x 22: donor block, istart[3], iend[3], istart_donor[3],
                        iend_donor[3], transmat[3][3] row-major: the connectinfo
                        after src/bc_module.f90:355-488 */
} cav3d_block;

typedef struct cav3d_ctx cav3d_ctx;

/* Blocks need imax, jmax, kmax >= 4 (3 cells per direction). */
int cav3d_create(const cav3d_params *params, int nblocks, const cav3d_block *blocks, cav3d_ctx **out);

/* cx(3,1:imax,1:jmax+1,1:kmax+1), ex(3,1:imax+1,1:jmax,1:kmax+1),
 * tx(3,1:imax+1,1:jmax+1,1:kmax), grd(ngrd,1:imax+1,1:jmax+1,1:kmax+1),
 * ngrd = 4 (laminar/Euler) or 5 (turbulent). */
int cav3d_set_metrics(cav3d_ctx *ctx, int b, const double *cx, const double *ex,
                      const double *tx, const double *grd);

/* Grid metrics computed on the device from the vertex coordinates, in place
 * of cav3d_set_metrics.  x(3,imax,jmax,kmax): the CGNS vertices of block b
 * (t_grid%x, src/grid_module.f90:118-130).  cav3d_build_metrics, once every
 * block has its coordinates, runs t_grid's calnormal, calvolume and (for
 * iturb = 0) calydns over all blocks (src/grid_module.f90:328-330, 350-1145):
 * face vectors, six-pyramid volumes, centroids, the wall distance to the
 * nearest BCWall face centroid of any block, the ghost fills and the
 * interface exchanges.  cav3d_get_metrics returns what the context holds, in
 * the cav3d_set_metrics layouts; any output may be NULL. */
int cav3d_set_grid(cav3d_ctx *ctx, int b, const double *x);
int cav3d_build_metrics(cav3d_ctx *ctx);
int cav3d_get_metrics(cav3d_ctx *ctx, int b, double *cx, double *ex, double *tx, double *grd);

/* Integrals of the current state.  out[10]: vapour volume and gas volume
 * (sum over interior cells of rho*y1/rho_v and rho*y2/rho_g times the cell
 * volume, the vof fields of src/postvariable_module.f90:282-295), the
 * pressure force fx_p, fy_p, fz_p and the viscous force fx_v, fy_v, fz_v over
 * the CAV3D_BC_SURFACE ranges scaled by 2/(rhoref*uref^2)/area, and cl, cd at
 * angle of attack aoa [rad] (clcd_writing, src/postvariable_module.f90:507-633).
 * Surface forces need the vertex coordinates (cav3d_set_grid). */
int cav3d_integrals(cav3d_ctx *ctx, double rhoref, double uref, double area, double aoa, double *out);

/* cav3d_integrals without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Like cav3d_integrals it copies its
 * results to the host and synchronises the context's stream, so it is not
 * for code under stream capture.  Same arguments and
 * results as cav3d_integrals (bit for bit); its partial sums, segment table
 * and final sums live in one fixed 2 MiB device array per GPU instead of
 * buffers allocated per call.  Limits: at most 64 CAV3D_BC_SURFACE ranges and
 * 2*ceil(cells/256) + 6*ceil(surface faces/256) <= 261568 partial sums
 * (about 33 M cells); beyond them it returns an error.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int cav3d_integrals_noalloc(cav3d_ctx *ctx, double rhoref, double uref, double area, double aoa, double *out);

/* Restart file of every block, the reference's out_<iter>.dat layout
 * (export_variable, src/variable_module.f90:108-182): per block, int32 nt_phy,
 * int32 nt, float64 time, pv(npv,-1:imax+3,..), emut(-1:imax+3,..) when
 * turbulent, qq(npv,2,2:imax,..) when dual-time, rcav(2,2:imax,..) when
 * cavitating, native byte order, no record markers.  cav3d_import reads it as
 * the restart reader does (src/initial_module.f90:245-389): interior pv,
 * emut, qq and rcav; dv and tv(1:2) from the EOS; ghosts are left to the next
 * boundary fill; nps / nts / timeprev returned as the reader advances them. */
int cav3d_export(cav3d_ctx *ctx, const char *path, int nt_phy, int nt, double time);
int cav3d_import(cav3d_ctx *ctx, const char *path, int *nps, int *nts, double *timeprev);

/* Discrete adjoint (vector-Jacobian product) of the residual cav3d_calrhs
 * returns.  No counterpart in the reference, which has no adjoint.
 * The differentiated map is res(pv): every block's ghosted pv, with the cell
 * properties (rho, h, c^2, EOS derivatives, mu, kappa) re-derived from pv by
 * the EOS, and with the eddy viscosity, the grid metrics, the BDF history qq
 * and the ghost layers held fixed (ghost cells are inputs like the interior).
 *   res_bar   every block's res_bar in cav3d_calrhs's res layout
 *             (n, 2:imax, 2:jmax, 2:kmax), blocks concatenated in order
 *   pv_bar    out: (d res/d pv)^T res_bar in the pv layout
 *             (n, -1:imax+3, -1:jmax+3, -1:kmax+3), blocks concatenated
 *   scal_bar  out (may be NULL): [0] d(res_bar . res)/d(p_ref) at fixed
 *             gauge pv, [1] d(res_bar . res)/d(u_ref) (low-Mach
 *             preconditioning speed)
 * FP64 throughout, also in a CAV3D_MIXED context (it is the FP64 residual's
 * VJP).  Every option cav3d_create accepts is differentiated.
 * Deterministic (fixed-order gathers, no atomics). */
int cav3d_rhs_vjp(cav3d_ctx *ctx, const double *res_bar, double *pv_bar, double *scal_bar);

/* cav3d_rhs_vjp with the interior cells' eddy viscosity a function of the
 * state instead of held: mu_t = the SST eddy viscosity of the cell's rho,
 * mu, k, omega and its Green-Gauss velocity gradients over the 7-point
 * stencil (t_eddy%calmut, src/eddy_module.f90:135-187), as cav3d_eddy
 * computes it.  The ghost cells' eddy viscosity stays an input held fixed.
 * k-omega SST (iturb 0) only.  No counterpart in the reference. */
int cav3d_rhs_vjp_eddy(cav3d_ctx *ctx, const double *res_bar, double *pv_bar, double *scal_bar);

/* t_eddy%calmut (src/eddy_module.f90:135-187): the interior cells' eddy
 * viscosity from the current pv and cell properties (the step computes it
 * after every update).  k-omega SST only; does not synchronise. */
int cav3d_eddy(cav3d_ctx *ctx);

/* Derivatives of res_bar . res with respect to the run parameters the
 * residual reads, with the ghost layers refilled by the boundary fill
 * (cav3d_setbc) from the same interior.  No counterpart in the reference.
 *   par_bar  out, 14 values: [0] p_ref (at fixed gauge pv), [1] u_ref,
 *            [2] dp_ref, [3] t_ref (the mass-transfer scales), [4..12]
 *            bc_pv[0..8], [13] bc_pressure (the boundary state)
 * The characteristic boundary fills evaluate the ghost cells' EOS at
 * bc_pv[0] plus the gauge pressure, where the residual's cells use p_ref: a
 * case whose bc_pv[0] equals p_ref (the reference's reference state) gets
 * the derivative of a change that moves both as par_bar[0] + par_bar[4];
 * either one alone is not the derivative of the filled map.
 * The ghost layers must hold the boundary fill of the interior when it is
 * called; they hold it again on return.  The total derivative with respect
 * to a quantity the parameters are functions of (a cavitation number, an
 * inflow speed) is par_bar . d(parameters)/d(quantity). */
int cav3d_rhs_vjp_params(cav3d_ctx *ctx, const double *res_bar, double *par_bar);

/* The steady-adjoint objective of the cavitating channel and its gradient:
 *   J = w_vap * (vapour volume: sum over interior cells of rho y1/rho_v vol)
 *     + w_fx  * (x-force on the CAV3D_BC_SURFACE ranges in N: the gauge
 *                pressure and wall-shear integrands of clcd_writing,
 *                src/postvariable_module.f90:507-633, unscaled)
 *   pv_bar   out (may be NULL): dJ/dpv in the pv layout, blocks concatenated
 *            (the eddy viscosity held fixed; ghost entries zero)
 *   dJ_dpref out (may be NULL): dJ/d(p_ref) at fixed gauge pv
 * No counterpart in the reference.  The force needs the vertex
 * coordinates (cav3d_set_grid) when w_fx != 0. */
int cav3d_objective(cav3d_ctx *ctx, double w_vap, double w_fx, double *J, double *pv_bar, double *dJ_dpref);

/* pv(npv,-1:imax+3,-1:jmax+3,-1:kmax+3), dv(18,...), tv(ntv,...),
 * qq(npv,2,2:imax,2:jmax,2:kmax).  dv/tv may be NULL: then dv and tv(1:2)
 * are evaluated from pv on the device and tv(3) is set to emutref. qq may be
 * NULL: then both time levels are set from the state (initial_module set_others). */
int cav3d_put_state(cav3d_ctx *ctx, int b, const double *pv, const double *dv,
                    const double *tv, const double *qq);
int cav3d_get_state(cav3d_ctx *ctx, int b, double *pv, double *dv, double *tv, double *qq);

/* One LU-SGS pseudo-time step of every block.  time may be NULL (no host sync). */
int cav3d_timeinteg(cav3d_ctx *ctx, int nt_phy, int nt, double timeprev, double *time);

/* Fill all ghost cells (physical BCs, interfaces, edges, corners). */
int cav3d_setbc(cav3d_ctx *ctx);

/* RHS of block b from the current state (ghosts as they are): res(npv,2:imax,..),
 * icav(4,..), itt(4,..), omega_cut(..); any output may be NULL. */
int cav3d_calrhs(cav3d_ctx *ctx, double time, int b, double *res, double *icav,
                 double *itt, double *omega_cut);

/* Local time steps of block b (computed from the current state). */
int cav3d_timestep(cav3d_ctx *ctx, int b, double *dt);

/* Norms of the last cav3d_timeinteg: l2[npv] = sqrt(sum (dpv)^2) over all
 * interior cells, ijk_max[4] = {block, i, j, k} of the first strict maximum
 * of |dp| (> 1e-12) in k-j-i order, resmax its value. */
int cav3d_residual(cav3d_ctx *ctx, double *l2, int *ijk_max, double *resmax);

/* qq(2) = qq(1); qq(1) = conservative state (dual time). */
int cav3d_unsteady_shift(cav3d_ctx *ctx);

/* dv(18,n), tv(2,n) of n states (absolute p, T, y1, y2) on the device / host. */
int cav3d_eos(cav3d_ctx *ctx, int n, const double *p, const double *t, const double *y1,
              const double *y2, double *dv, double *tv);
int cav3d_eos_host(int n, const double *p, const double *t, const double *y1, const double *y2,
                   double *dv, double *tv);

/* Device time, in milliseconds, of the last cav3d_timeinteg's kernels by phase
 * (HIP events on the context's stream, recorded while enabled):
 * out[0] timestep, [1] boundary, [2] rhs, [3] lusgs, [4] update+eddy, [5] total,
 * [6] forward sweep kernel, [7] backward sweep kernel; [8] the kernels of the
 * last cav3d_rhs_vjp, [9] those of the last cav3d_calrhs.  out holds 10 doubles. */
int cav3d_profile(cav3d_ctx *ctx, int enable, double *out);

int cav3d_synchronize(cav3d_ctx *ctx);
const char *cav3d_last_error(void);
void cav3d_destroy(cav3d_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

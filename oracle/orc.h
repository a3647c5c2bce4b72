/*
 * ORACLE — test infrastructure only (see orc_eos.c header).
 *
 * Shared declarations of the plain-C FP64 restatement of the reference
 * pseudo-time step (t_lusgs%timeinteg, src/update_module.f90:455-733).
 * Arrays use the reference's Fortran layouts (component fastest,
 * column-major, explicit lower bounds), e.g. pv(npv,-1:imax+3,-1:jmax+3,-1:kmax+3).
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double rho, h, drdp, drdt, dhdp, dhdt; } orc_eos2;
typedef struct { double vis, cond; } orc_prop2;

static inline double fmax_f(double a, double b) { return (b > a) ? b : a; }
static inline double fmin_f(double a, double b) { return (b < a) ? b : a; }

double orc_powi(double x, int m);
double orc_h2o_pww(double t);
double orc_h2o_tww(double p);
void orc_iapws97_l(double p, double t, orc_eos2 *e);
void orc_iapws97_v(double p, double t, orc_eos2 *e);
void orc_iapws97_l_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr);
void orc_iapws97_v_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr);
void orc_ideal(double p, double t, orc_eos2 *e);
void orc_ideal_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr);
void orc_deteos_simple(double p, double t, double y1, double y2, double *dv);
void orc_deteos_prop(double p, double t, double y1, double y2, double *dv, double *tv);
extern int orc_alt_snd2_count;

/* Physical-BC kinds accepted from the host (reference bcname strings). */
enum {
    ORC_BC_WALL = 1, ORC_BC_INFLOW_SUBSONIC = 2, ORC_BC_OUTFLOW_SUBSONIC = 3,
    ORC_BC_FARFIELD = 4, ORC_BC_SYMMETRY = 5, ORC_BC_INFLOW_SUPERSONIC = 6,
    ORC_BC_OUTFLOW_SUPERSONIC = 7, ORC_BC_DEGENERATE = 8, ORC_BC_INFLOW = 9,
    ORC_BC_OUTFLOW = 10
};
/* flag OR'ed into a kind: the range belongs to the 'Surface' family */
enum { ORC_BC_SURFACE = 0x100 };

/* Run configuration (the reference's config_module fields that the path reads). */
typedef struct {
    int npv, ntv, ndv, ngrd;
    int iturb, ides, nscheme, nmuscl, nlim, timemethod, local, prec, ncav, nsteady;
    int nprint;
    double cfl, pref, uref, str, dt_phy, c_v, c_c, dp_ref, t_ref, t_crit;
    /* reference boundary state bcinfo%pv (pv(1) absolute), bcinfo%pressure, emutref */
    double bc_pv[9];
    double bc_pressure;
    double emutref;
    /* cavitation-rate clip: 0 as written (temperature, src/cav_module.f90:151-152),
       1 on the mass fractions (y1 and 1-y1-y2) */
    int cav_clip;
} orc_cfg;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_new(const orc_cfg *cfg, int nblocks, const int *dims);
void orc_free(orc_ctx *c);
int orc_set_grid(orc_ctx *c, int b, const double *cx, const double *ex, const double *tx, const double *grd);
int orc_set_bcs(orc_ctx *c, int b, int nbc, const int *desc);
int orc_set_conns(orc_ctx *c, int b, int ncon, const int *desc);
int orc_finalize(orc_ctx *c);
int orc_put_state(orc_ctx *c, int b, const double *pv, const double *dv, const double *tv, const double *qq);
int orc_get_state(orc_ctx *c, int b, double *pv, double *dv, double *tv, double *qq);
int orc_timestep(orc_ctx *c, int nt_phy, int nt, double timeprev, double *time);
int orc_timeinteg(orc_ctx *c, int nt_phy, int nt, double timeprev, double *time);
int orc_get_bcinfo(orc_ctx *c, int b, int *out);
int orc_setbc(orc_ctx *c);
int orc_calrhs(orc_ctx *c, int b, double *res, double *icav, double *itt, double *omega_cut);
int orc_get_dt(orc_ctx *c, int b, double *dt);
int orc_setqres(orc_ctx *c);
int orc_residual(orc_ctx *c, double *l2, int *ijkmax, double *resmax);
int orc_unsteady_shift(orc_ctx *c);
void orc_set_threads(int n);
int orc_integrals(orc_ctx *c, const double *const *x, double rhoref, double uref, double area, double aoa,
                  double *out);

/* One block for orc_grid_metrics: vertex coordinates in, metrics out
 * (Fortran layouts of t_grid: x(3,2:imax+1,2:jmax+1,2:kmax+1) holding the
 * CGNS vertices, cx(3,1:imax,1:jmax+1,1:kmax+1), ex(3,1:imax+1,1:jmax,1:kmax+1),
 * tx(3,1:imax+1,1:jmax+1,1:kmax), grd(ngrd,1:imax+1,1:jmax+1,1:kmax+1)). */
typedef struct {
    int imax, jmax, kmax;
    int nbc, ncon;
    const double *x;
    const int *bc;     /* nbc x 7 {kind, istart[3], iend[3]} */
    const int *con;    /* ncon x 22, as orc_set_conns takes them */
    double *cx, *ex, *tx, *grd;
} orc_gblock;
int orc_grid_metrics(int nb, orc_gblock *blk, int ngrd);

#ifdef __cplusplus
}
#endif
#endif

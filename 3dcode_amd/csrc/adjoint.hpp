// Discrete adjoint of the residual: cav3d_rhs_vjp.
//
// The operator is the residual cav3d_calrhs evaluates (src/rhs_module.f90:
// 260-643: inviscid and viscous face-flux differences, the mass-transfer and SST
// sources, the BDF term), taken as a function of every block's ghosted
// primitive state pv, with the cell properties re-derived from pv by the EOS
// (rho, h, c^2, the EOS derivatives, mu, kappa: deteos_prop) and with the
// eddy viscosity, the grid metrics, the BDF history and the ghost layers
// held fixed (a ghost cell is an input like an interior cell).  It returns
//   pv_bar = (d res / d pv)^T res_bar
// and the two scalar derivatives of res_bar . res with respect to p_ref and
// u_ref.  The reference has no adjoint; this is the VJP of its residual.
//
// Method: reverse accumulation over faces and cells, local derivatives in
// forward mode.  Each face recomputes its reconstruction and flux (nothing is
// kept from a forward pass) and takes the gradient of
//   phi_f = g_f . F_f,   g_f = res_bar(right cell) - res_bar(left cell)
// through dual-number passes (Dn<W>: W inputs seeded per pass) of the same
// templated physics the residual kernels run, split where the input count is
// smallest:
//   flux core       the two face states, rho and h (c^2) of each side, the
//                   Roe-averaged c^2, [the 18 sensor pressures], u_ref
//   averaged c^2    the ten quantities it is built from
//   face EOS        (p, T, y1, y2) of each side
//   clamps          each clamped value comes from the reconstruction, the
//                   cell or a constant (recorded in the forward recompute)
//   reconstruction  each variable's stencil (38 points under MLP, 6 under TVD)
// The per-face bars of every stencil point go to a buffer and each cell
// gathers its own in a fixed order: the result is deterministic, with no
// atomics.  The cell sources are handled per cell the same way, and each
// cell's property bars return to (p, T, y1, y2) through one 4-lane dual pass
// of the EOS.  Arithmetic is FP64 throughout.
//
// d/d(p_ref) is the sum of pv_bar over the pressure component: the residual
// reads gauge pressure only as p + p_ref or in differences (the pressure
// floor included), so shifting p_ref shifts it as shifting every gauge
// pressure does.  d/d(u_ref) is carried on one extra lane of the flux-core
// passes (the low-Mach preconditioning is the only reader).
//
// Included once, from cav3d.hip, after the context definition.

#include "adjoint_point.hpp"

namespace {

struct AD {
    const double *rb;       // res_bar [npv][NC]
    double *fb[3];          // bars of a face's stencil points [NSLOT][npv][NF]
    double *rsb[3];         // a face's reconstruction seeds [4][npv][NF]: left, right, left cell, right cell
    double *snb[3];         // a face's sensor-pressure bars [18][NF] (Roe-M, AUSMPW+)
    double *db[3];          // cell-property bars of a face's two cells [2][ND][NF]
    double *ub[3];          // d(phi_f)/d(u_ref) [NF]
    double *cb;             // cell-source bars on the 7-point stencil [7][npv][NC]
    double *cdb;            // cell-source property bars [ND][NC]
    double *csb;            // cell-source d(phi_c)/d(u_ref, dp_ref, t_ref) [3][NC] (the mass-transfer scales)
    double *pb;             // pv_bar [npv][NG]
    double *gdb;            // a ghosted cell's gathered property bars [ND][NG], for the EOS pass
    double *mtb[3];         // eddy-viscosity bars of a face's two cells [2][NF] (dmut)
    unsigned long long mask;   // stencil points a face writes (bit ll-1)
    int dmut;               // 1: the interior cells' eddy viscosity is the SST function of
                            // the state (differentiated); 0: held fixed
};

__device__ __forceinline__ bool interior_cell(const BlkD &B, int i, int j, int k)
{
    return i >= 2 && i <= B.I && j >= 2 && j <= B.J && k >= 2 && k <= B.K;
}

__device__ __forceinline__ double rbar(const DS &S, const AD &A, const BlkD &B, int i, int j, int k, int n)
{
    return interior_cell(B, i, j, k) ? A.rb[(long)n * S.NC + cidx(B, i, j, k)] : 0.0;
}

// (normal, first transverse, second transverse) offset of stencil point ll
// from the face's left cell; the same table as stencil_off
__device__ __forceinline__ void stencil_ijk(int ll, int &dn, int &de, int &dt)
{
    int q = ll - 1;
    if (q < 18) { dn = q & 1; de = (q >> 1) % 3 - 1; dt = q / 6 - 1; }
    else if (q < 27) { q -= 18; dn = -1; de = q % 3 - 1; dt = q / 3 - 1; }
    else if (q < 36) { q -= 27; dn = 2; de = q % 3 - 1; dt = q / 3 - 1; }
    else { dn = q == 36 ? -2 : 3; de = 0; dt = 0; }
}

// does the face of direction D with left cell (i,j,k) exist
__device__ __forceinline__ bool face_exists(const BlkD &B, int D, int i, int j, int k)
{
    const int lo[3] = {D == 0 ? 1 : 2, D == 1 ? 1 : 2, D == 2 ? 1 : 2};
    return i >= lo[0] && i <= B.I && j >= lo[1] && j <= B.J && k >= lo[2] && k <= B.K;
}

// Inviscid part of a face up to its reconstruction: the seeds of the
// reconstruction (k_face_rvjp takes them to the stencil points), the two
// cells' property bars and the u_ref derivative.
template <int D, bool SENSOR>
__global__ __launch_bounds__(128) void k_face_vjp(DS S, AD A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NF[D]) return;
    int i, j, k;
    const BlkD &B = face_of<D>(S, t, i, j, k);
    const Par &P = S.P;
    const int npv = S.npv;
    const long NF = S.NF[D];
    const long g9 = gidx(B, i, j, k);
    const long sN = D == 0 ? 1 : (D == 1 ? (long)B.nx : (long)B.nx * B.ny);
    const long gq = g9 + sN;
    const int ip = i + (D == 0), jp = j + (D == 1), kp = k + (D == 2);
    double g[NPV];
    bool any = false;
    for (int n = 0; n < NPV; n++) {
        g[n] = n < npv ? rbar(S, A, B, ip, jp, kp, n) - rbar(S, A, B, i, j, k, n) : 0.0;
        any = any || g[n] != 0.0;
    }
    double *rsb = A.rsb[D], *db = A.db[D];
    if (!any) {
        for (int m = 0; m < 4 * npv; m++) rsb[m * NF + t] = 0.0;
        if (SENSOR) for (int q = 0; q < 18; q++) A.snb[D][q * NF + t] = 0.0;
        for (int m = 0; m < 2 * ND; m++) db[m * NF + t] = 0.0;
        A.ub[D][t] = 0.0;
        return;
    }
    const double *f0 = S.pv + g9;
    long off[NSLOT];
    for (int ll = 1; ll <= NSLOT; ll++) off[ll - 1] = stencil_off<D>(B, ll);
    double n3[3];
    face_vec(S, B, D, i, j, k, n3);
    const double cellp[6] = {S.dc[D_RHO * S.NG + g9], S.dc[D_H * S.NG + g9], S.dc[D_C2 * S.NG + g9],
                             S.dc[D_RHO * S.NG + gq], S.dc[D_H * S.NG + gq], S.dc[D_C2 * S.NG + gq]};
    double ub, xlb[NPV], xrb[NPV], c9b[NPV], c10b[NPV], senb[18];
    face_adj_core<SENSOR>(P, npv, f0, S.NG, off, n3, cellp, g, xlb, xrb, c9b, c10b, senb, db + t, NF, ub);
    for (int n = 0; n < npv; n++) {
        rsb[(0 * npv + n) * NF + t] = xlb[n];
        rsb[(1 * npv + n) * NF + t] = xrb[n];
        rsb[(2 * npv + n) * NF + t] = c9b[n];
        rsb[(3 * npv + n) * NF + t] = c10b[n];
    }
    if (SENSOR) for (int q = 0; q < 18; q++) A.snb[D][q * NF + t] = senb[q];
    A.ub[D][t] = ub;
}

// The reconstruction part of a face's inviscid adjoint, one thread per face
// and variable: the seeds k_face_vjp left, back to the face's stencil points.
template <int D, bool SENSOR>
__global__ __launch_bounds__(256) void k_face_rvjp(DS S, AD A)
{
    const long NF = S.NF[D];
    long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= (long)S.npv * NF) return;
    const int n = (int)(u / NF);
    const long t = u - (long)n * NF;
    int i, j, k;
    const BlkD &B = face_of<D>(S, t, i, j, k);
    const int npv = S.npv;
    const double *rsb = A.rsb[D];
    const double *f0 = S.pv + (long)n * S.NG + gidx(B, i, j, k);
    const double *senb = nullptr;
    double sn18[18];
    if (SENSOR && n == 0) {
        for (int q = 0; q < 18; q++) sn18[q] = A.snb[D][q * NF + t];
        senb = sn18;
    }
    face_adj_recon(S.P, n, [&](int ll) -> double { return f0[stencil_off<D>(B, ll)]; }, rsb[(0 * npv + n) * NF + t],
                   rsb[(1 * npv + n) * NF + t], rsb[(2 * npv + n) * NF + t], rsb[(3 * npv + n) * NF + t], senb, A.mask,
                   A.fb[D] + (long)n * NF + t, (long)npv * NF);
}

// Viscous part of a face, added to the inviscid bars (stencil points 1..18,
// the two cells' mu, kappa and, under SST, h_T and rho).
template <int D>
__global__ __launch_bounds__(128) void k_vface_vjp(DS S, AD A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NF[D]) return;
    int i, j, k;
    const BlkD &B = face_of<D>(S, t, i, j, k);
    const Par &P = S.P;
    const int npv = S.npv;
    const long NF = S.NF[D];
    const long g9 = gidx(B, i, j, k);
    const long sN = D == 0 ? 1 : (D == 1 ? (long)B.nx : (long)B.nx * B.ny);
    const long gq = g9 + sN;
    int ip = i, jp = j, kp = k;
    if (D == 0) ip++; else if (D == 1) jp++; else kp++;
    double g[NPV];
    bool any = false;
    for (int n = 0; n < NPV; n++) {
        g[n] = n < npv ? rbar(S, A, B, i, j, k, n) - rbar(S, A, B, ip, jp, kp, n) : 0.0;
        any = any || g[n] != 0.0;
    }
    const bool turb = P.iturb == 0;
    const bool dm = A.dmut && turb;
    if (!any) {
        if (dm) { A.mtb[D][t] = 0.0; A.mtb[D][NF + t] = 0.0; }
        return;
    }
    const int nv = turb ? 6 : 4;
    const int vars[6] = {1, 2, 3, 4, 7, 8};
    auto xs = [&](int ll, int n) -> double { return S.pv[(long)n * S.NG + g9 + stencil_off<D>(B, ll)]; };
    // inputs: the normal / transverse differences, the face averages, the cell properties
    double g3[6][3], cav[5];
    for (int q = 0; q < nv; q++) {
        const int n = vars[q];
        g3[q][0] = xs(10, n) - xs(9, n);
        g3[q][1] = 0.0625 * (-xs(1, n) - xs(2, n) + xs(5, n) + xs(6, n) + 2.0 * (-xs(7, n) - xs(8, n) + xs(11, n) + xs(12, n))
                             - xs(13, n) - xs(14, n) + xs(17, n) + xs(18, n));
        g3[q][2] = 0.0625 * (-xs(1, n) - xs(2, n) + xs(13, n) + xs(14, n) + 2.0 * (-xs(3, n) - xs(4, n) + xs(15, n) + xs(16, n))
                             - xs(5, n) - xs(6, n) + xs(17, n) + xs(18, n));
    }
    const int ncv = turb ? 5 : 3;
    const int cvar[5] = {1, 2, 3, 7, 8};
    for (int q = 0; q < ncv; q++) cav[q] = 0.5 * (xs(9, cvar[q]) + xs(10, cvar[q]));
    double dl[ND], dr[ND];
    for (int m = 0; m < ND; m++) { dl[m] = S.dc[m * S.NG + g9]; dr[m] = S.dc[m * S.NG + gq]; }
    const double ml = S.mut ? S.mut[g9] : 0.0, mr = S.mut ? S.mut[gq] : 0.0;
    // metrics (as k_vface)
    double n3[3], e4[4][3], t4[4][3];
    face_vec(S, B, D, i, j, k, n3);
    const int E = (D + 1) % 3, T = (D + 2) % 3;
    int p0[3] = {i, j, k}, p1[3] = {ip, jp, kp};
    int a0[3], a1[3];
    for (int q = 0; q < 3; q++) { a0[q] = p0[q]; a1[q] = p1[q]; }
    a0[E]--; a1[E]--;
    face_vec(S, B, E, a0[0], a0[1], a0[2], e4[0]);
    face_vec(S, B, E, a1[0], a1[1], a1[2], e4[1]);
    face_vec(S, B, E, p0[0], p0[1], p0[2], e4[2]);
    face_vec(S, B, E, p1[0], p1[1], p1[2], e4[3]);
    for (int q = 0; q < 3; q++) { a0[q] = p0[q]; a1[q] = p1[q]; }
    a0[T]--; a1[T]--;
    face_vec(S, B, T, a0[0], a0[1], a0[2], t4[0]);
    if (D == 0) {
        face_vec(S, B, T, a1[0], a1[1], a1[2], t4[1]);
        face_vec(S, B, T, p0[0], p0[1], p0[2], t4[2]);
    } else {
        face_vec(S, B, T, p0[0], p0[1], p0[2], t4[1]);
        face_vec(S, B, T, a1[0], a1[1], a1[2], t4[2]);
    }
    face_vec(S, B, T, p1[0], p1[1], p1[2], t4[3]);
    const long wl = widx(B, i, j, k), wr = widx(B, ip, jp, kp);
    const double gl[2] = {S.vol[wl], S.ydn ? S.ydn[wl] : 1.0}, gr[2] = {S.vol[wr], S.ydn ? S.ydn[wr] : 1.0};
    // z = {g3[nv][3], cav[ncv], muL, kapL, muR, kapR, [hTL, hTR, rhoL, rhoR, [mutL, mutR]]}
    constexpr int NZ = 18 + 5 + 8 + 2;
    const int og = 3 * nv, oc = og + ncv;
    const int nz = oc + (turb ? (dm ? 10 : 8) : 4);
    double zx[NZ], zb[NZ];
    for (int q = 0; q < nv; q++) for (int c = 0; c < 3; c++) zx[3 * q + c] = g3[q][c];
    for (int q = 0; q < ncv; q++) zx[og + q] = cav[q];
    zx[oc] = dl[D_MU]; zx[oc + 1] = dl[D_KAP]; zx[oc + 2] = dr[D_MU]; zx[oc + 3] = dr[D_KAP];
    if (turb) { zx[oc + 4] = dl[D_HT]; zx[oc + 5] = dr[D_HT]; zx[oc + 6] = dl[D_RHO]; zx[oc + 7] = dr[D_RHO]; }
    if (dm) { zx[oc + 8] = ml; zx[oc + 9] = mr; }
    visc_adj(P, npv, nv, ncv, nz, zx, n3, e4, t4, gl, gr, dl, dr, ml, mr, dm ? oc + 8 : -1, g, zb);
    if (dm) { A.mtb[D][t] = zb[oc + 8]; A.mtb[D][NF + t] = zb[oc + 9]; }
    // the differences and averages are linear in the stencil
    double *fb = A.fb[D];
    constexpr double C1[18] = {-1, -1, 0, 0, 1, 1, -2, -2, 0, 0, 2, 2, -1, -1, 0, 0, 1, 1};
    constexpr double C2[18] = {-1, -1, -2, -2, -1, -1, 0, 0, 0, 0, 0, 0, 1, 1, 2, 2, 1, 1};
    for (int q = 0; q < nv; q++) {
        const int n = vars[q];
        for (int ll = 1; ll <= 18; ll++) {
            double v = 0.0625 * (C1[ll - 1] * zb[3 * q + 1] + C2[ll - 1] * zb[3 * q + 2]);
            if (ll == 9) v -= zb[3 * q];
            if (ll == 10) v += zb[3 * q];
            if (v != 0.0) fb[((long)(ll - 1) * npv + n) * NF + t] += v;
        }
    }
    for (int q = 0; q < ncv; q++) {
        const int n = cvar[q];
        fb[((long)8 * npv + n) * NF + t] += 0.5 * zb[og + q];
        fb[((long)9 * npv + n) * NF + t] += 0.5 * zb[og + q];
    }
    double *db = A.db[D];
    db[D_MU * NF + t] += zb[oc];
    db[D_KAP * NF + t] += zb[oc + 1];
    db[(ND + D_MU) * NF + t] += zb[oc + 2];
    db[(ND + D_KAP) * NF + t] += zb[oc + 3];
    if (turb) {
        db[D_HT * NF + t] += zb[oc + 4];
        db[(ND + D_HT) * NF + t] += zb[oc + 5];
        db[D_RHO * NF + t] += zb[oc + 6];
        db[(ND + D_RHO) * NF + t] += zb[oc + 7];
    }
}

// the 7-point stencil of the cell sources (x7 of k_combine)
__device__ __forceinline__ void src_off(int q, int &di, int &dj, int &dk)
{
    di = q == 0 ? -1 : (q == 2 ? 1 : 0);
    dj = q == 3 ? -1 : (q == 4 ? 1 : 0);
    dk = q == 5 ? -1 : (q == 6 ? 1 : 0);
}

// Cell sources of one interior cell: the mass transfer, SST, BDF.  MODELS:
// the instantiation with the Kunz / Singhal / Schnerr-Sauer / Lee sources
// (a full EOS of the cell) and the DES length scales, as k_combine<true>.
template <bool MODELS>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_cell_vjp(DS S, AD A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    const Par &P = S.P;
    const int npv = S.npv;
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    const long g = gidx(B, i, j, k);
    double rb[NPV];
    for (int n = 0; n < NPV; n++) rb[n] = n < npv ? A.rb[(long)n * S.NC + t] : 0.0;
    const bool sst = P.iturb == 0, cav_ = P.ncav >= 1, bdf = P.nsteady > 0;
    // z = {pc[npv], rho, h, mu, the Green-Gauss gradients of u, v, w, k, omega (15),
    //      the scalars the mass-transfer rates read: u_ref (Singhal), dp_ref
    //      (Merkle, Kunz's evaporation), t_ref (Merkle, Kunz), [the eddy viscosity]}
    constexpr int NZ = NPV + 3 + 15 + 3 + 1;
    const int od = npv, og = npv + 3;
    const int nz0 = (cav_ || bdf || sst) ? (sst ? og + 15 : og) : 0;
    const bool sc = nz0 && (P.ncav == 1 || P.ncav == 2 || P.ncav == 3);
    const int osc = nz0;                                   // u_ref, dp_ref, t_ref lanes
    const bool dm = A.dmut && sst;                         // the eddy viscosity is a state function
    const int omu = sc ? nz0 + 3 : nz0;                    // its lane
    const int nz = omu + (dm ? 1 : 0);
    double zx[NZ], zb[NZ];
    for (int q = 0; q < NZ; q++) zx[q] = zb[q] = 0.0;
    const int gvar[5] = {1, 2, 3, 7, 8};
    double d[ND];
    for (int m = 0; m < ND; m++) d[m] = S.dc[m * S.NG + g];
    double x7[7][NPV];
    for (int q = 0; q < 7; q++) {
        int di, dj, dk;
        src_off(q, di, dj, dk);
        const long gn = gidx(B, i + di, j + dj, k + dk);
        for (int n = 0; n < NPV; n++) x7[q][n] = n < npv ? S.pv[(long)n * S.NG + gn] : 0.0;
    }
    const double vol = S.vol[widx(B, i, j, k)];
    double f6[6][3];
    cell_faces(S, B, i, j, k, f6);
    const double iv = 1.0 / vol;
    if (nz) {
        for (int n = 0; n < npv; n++) zx[n] = x7[1][n];
        zx[od] = d[D_RHO]; zx[od + 1] = d[D_H]; zx[od + 2] = d[D_MU];
        if (sst)
            for (int v = 0; v < 5; v++) gg_grad<double>(x7, gvar[v], f6, iv, zx + og + 3 * v);
        if (sc) { zx[osc] = P.uref; zx[osc + 1] = P.dp_ref; zx[osc + 2] = P.t_ref; }
        const double ydn = sst ? S.ydn[widx(B, i, j, k)] : 1.0;
        const double mut = S.mut ? S.mut[g] : 0.0;
        if (dm) zx[omu] = mut;
        double q1[NPV], q2[NPV];
        for (int n = 0; n < NPV; n++) {
            q1[n] = (bdf && n < npv) ? S.qq[(long)n * S.NC + t] : 0.0;
            q2[n] = (bdf && n < npv) ? S.qq[(long)(npv + n) * S.NC + t] : 0.0;
        }
        constexpr int W = 2;
        using T = Dn<W>;
        dn_grad<W>(nz, [&](int q) -> double { return zx[q]; }, [&](auto in, int base) -> T {
            Par Pu = P;
            auto lane_of = [&](int q) { const int l = q - base; return (sc && l >= 0 && l < W) ? l + 1 : 0; };
            Pu.uref_seed = lane_of(osc);
            Pu.dpref_seed = lane_of(osc + 1);
            Pu.tref_seed = lane_of(osc + 2);
            T pc[NPV], dt[ND];
            for (int n = 0; n < NPV; n++) pc[n] = n < npv ? in(n) : T(0.0);
            for (int m = 0; m < ND; m++) dt[m] = T(d[m]);
            dt[D_RHO] = in(od); dt[D_H] = in(od + 1); dt[D_MU] = in(od + 2);
            T phi(0.0);
            if (cav_) {
                T rc, rv, ic[4];
                if constexpr (MODELS) {
                    if (P.ncav >= 2) {      // on the cell's full EOS, as k_combine<true>
                        T dd[ND], full[18];
                        mixture_eos<T>(pc[0] + P.pref, pc[4], pc[5], pc[6], dd, false, full);
                        cav_model<T>(Pu, pc, full, vol, rc, rv, ic);
                    } else {
                        merkle<T>(Pu, pc, dt, vol, rc, rv, ic);
                    }
                } else {
                    merkle<T>(Pu, pc, dt, vol, rc, rv, ic);
                }
                phi = phi + rb[5] * (rv - rc);
            }
            if (sst) {
                T gr[5][3];
                for (int v = 0; v < 5; v++)
                    for (int c = 0; c < 3; c++) gr[v][c] = in(og + 3 * v + c);
                T src[2], it[4], oc;
                sst_source_g<MODELS, T>(P, gr[0], gr[1], gr[2], gr[3], gr[4], pc[7], pc[8], f6, vol, ydn, dt,
                                     dm ? in(omu) : T(mut), src, it, oc);
                phi = phi + rb[7] * src[0] + rb[8] * src[1];
            }
            if (bdf) {
                T fx[NPV];
                fx[0] = dt[D_RHO];
                fx[1] = dt[D_RHO] * pc[1];
                fx[2] = dt[D_RHO] * pc[2];
                fx[3] = dt[D_RHO] * pc[3];
                fx[4] = dt[D_RHO] * (dt[D_H] + 0.5 * (pc[1] * pc[1] + pc[2] * pc[2] + pc[3] * pc[3])) - (pc[0] + P.pref);
                for (int n = 5; n < NPV; n++) fx[n] = n < npv ? dt[D_RHO] * pc[n] : T(0.0);
                for (int n = 0; n < npv; n++) {
                    T v = (double)P.nsteady * (P.bdf_b * fx[n] - q1[n]) + 0.5 * ((double)P.nsteady - 1.0) * q2[n];
                    phi = phi - rb[n] * (v * vol / P.dt_phy);
                }
            }
            return phi;
        }, zb);
        if (dm) {
            // the cell's eddy-viscosity bar: its sources' share and its six
            // viscous faces' (the left cell of its + faces, the right cell of
            // its - faces), back to rho, mu, k, omega and the velocity
            // gradients through mu_t = sst_eddy_g (src/eddy_module.f90:135-187)
            double mb = zb[omu];
            for (int D = 0; D < 3; D++) {
                mb += A.mtb[D][fidx(B, D, i, j, k)];
                mb += A.mtb[D][S.NF[D] + fidx(B, D, i - (D == 0), j - (D == 1), k - (D == 2))];
            }
            if (mb != 0.0) {
                double ex[13], eb[13];
                ex[0] = d[D_RHO]; ex[1] = d[D_MU]; ex[2] = x7[1][7]; ex[3] = x7[1][8];
                for (int q = 0; q < 9; q++) ex[4 + q] = zx[og + q];
                dn_grad<4>(13, [&](int q) -> double { return ex[q]; }, [&](auto in, int) -> Dn<4> {
                    const Dn<4> gu[3] = {in(4), in(5), in(6)}, gv[3] = {in(7), in(8), in(9)}, gw[3] = {in(10), in(11), in(12)};
                    return sst_eddy_g<Dn<4>>(in(0), in(1), in(2), in(3), gu, gv, gw, ydn);
                }, eb);
                zb[od] += mb * eb[0];
                zb[od + 2] += mb * eb[1];
                zb[7] += mb * eb[2];
                zb[8] += mb * eb[3];
                for (int q = 0; q < 9; q++) zb[og + q] += mb * eb[4 + q];
            }
        }
    }
    // outputs: centre (q = 1) and neighbour bars, the cell's property bars;
    // a gradient component g[c] = iv * sum_q w_q[c] x_q with the Green-Gauss
    // weights w_q = -f6[0]/2, (face sum)/2, +f6[1]/2, -f6[2]/2, +f6[3]/2,
    // -f6[4]/2, +f6[5]/2 for q = 0..6
    double wq[7][3];
    for (int c = 0; c < 3; c++) {
        wq[0][c] = -0.5 * f6[0][c] * iv;
        wq[2][c] = 0.5 * f6[1][c] * iv;
        wq[3][c] = -0.5 * f6[2][c] * iv;
        wq[4][c] = 0.5 * f6[3][c] * iv;
        wq[5][c] = -0.5 * f6[4][c] * iv;
        wq[6][c] = 0.5 * f6[5][c] * iv;
        wq[1][c] = 0.5 * (f6[1][c] + f6[3][c] + f6[5][c] - f6[0][c] - f6[2][c] - f6[4][c]) * iv;
    }
    for (int q = 0; q < 7; q++)
        for (int n = 0; n < npv; n++) {
            double v = q == 1 ? zb[n] : 0.0;
            if (sst)
                for (int gv = 0; gv < 5; gv++)
                    if (gvar[gv] == n)
                        v += wq[q][0] * zb[og + 3 * gv] + wq[q][1] * zb[og + 3 * gv + 1] + wq[q][2] * zb[og + 3 * gv + 2];
            A.cb[((long)q * npv + n) * S.NC + t] = v;
        }
    for (int m = 0; m < ND; m++) A.cdb[(long)m * S.NC + t] = 0.0;
    A.cdb[(long)D_RHO * S.NC + t] = zb[od];
    A.cdb[(long)D_H * S.NC + t] = zb[od + 1];
    A.cdb[(long)D_MU * S.NC + t] = zb[od + 2];
    for (int q = 0; q < 3; q++) A.csb[(long)q * S.NC + t] = sc ? zb[osc + q] : 0.0;
}

// pv_bar of every ghosted cell: its bars from the faces whose stencils hold
// it and from the sources of its neighbours, in a fixed order, and its
// gathered property bars (k_adj_eos takes those through the EOS: the dual
// pass needs the registers this memory-bound gather does not).
__global__ __launch_bounds__(256) void k_adj_gather(DS S, AD A)
{
    long gl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gl >= S.NG) return;
    const BlkD &B = S.b[find_blk(S, gl, 1)];
    const Par &P = S.P;
    const int npv = S.npv;
    const long r = gl - B.g0;
    const int i = (int)(r % B.nx) - 1, j = (int)((r / B.nx) % B.ny) - 1, k = (int)(r / ((long)B.nx * B.ny)) - 1;
    double acc[NPV], dcb[ND];
    for (int n = 0; n < NPV; n++) acc[n] = 0.0;
    for (int m = 0; m < ND; m++) dcb[m] = 0.0;
    const int c3[3] = {i, j, k};
    for (int D = 0; D < 3; D++) {
        const int E = (D + 1) % 3, T = (D + 2) % 3;
        const long NF = S.NF[D];
        for (int ll = 1; ll <= NSLOT; ll++) {
            if (!((A.mask >> (ll - 1)) & 1ull)) continue;
            int dn, de, dt;
            stencil_ijk(ll, dn, de, dt);
            int f[3];
            f[D] = c3[D] - dn; f[E] = c3[E] - de; f[T] = c3[T] - dt;
            if (!face_exists(B, D, f[0], f[1], f[2])) continue;
            const long fi = fidx(B, D, f[0], f[1], f[2]);
            for (int n = 0; n < npv; n++) acc[n] += A.fb[D][((long)(ll - 1) * npv + n) * NF + fi];
        }
        if (face_exists(B, D, i, j, k)) {
            const long fi = fidx(B, D, i, j, k);
            for (int m = 0; m < ND; m++) dcb[m] += A.db[D][(long)m * NF + fi];
        }
        const int im = i - (D == 0), jm = j - (D == 1), km = k - (D == 2);
        if (face_exists(B, D, im, jm, km)) {
            const long fi = fidx(B, D, im, jm, km);
            for (int m = 0; m < ND; m++) dcb[m] += A.db[D][(long)(ND + m) * NF + fi];
        }
    }
    for (int q = 0; q < 7; q++) {
        int di, dj, dk;
        src_off(q, di, dj, dk);
        const int si = i - di, sj = j - dj, sk = k - dk;
        if (!interior_cell(B, si, sj, sk)) continue;
        const long ct = cidx(B, si, sj, sk);
        for (int n = 0; n < npv; n++) acc[n] += A.cb[((long)q * npv + n) * S.NC + ct];
    }
    if (interior_cell(B, i, j, k)) {
        const long ct = cidx(B, i, j, k);
        for (int m = 0; m < ND; m++) dcb[m] += A.cdb[(long)m * S.NC + ct];
    }
    for (int n = 0; n < npv; n++) A.pb[(long)n * S.NG + gl] = acc[n];
    for (int m = 0; m < ND; m++) A.gdb[(long)m * S.NG + gl] = dcb[m];
}

// the gathered property bars of a ghosted cell back to (p, T, y1, y2),
// added to the pv_bar k_adj_gather left
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_adj_eos(DS S, AD A)
{
    long gl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gl >= S.NG) return;
    const Par &P = S.P;
    double dcb[ND];
    bool any = false;
    for (int m = 0; m < ND; m++) {
        dcb[m] = A.gdb[(long)m * S.NG + gl];
        any = any || dcb[m] != 0.0;
    }
    if (!any) return;
    double x4[4] = {S.pv[gl] + P.pref, S.pv[4 * S.NG + gl], S.pv[5 * S.NG + gl], S.pv[6 * S.NG + gl]}, gx[4];
    eos_adj(P.iturb != -3, x4, dcb, gx);
    A.pb[gl] += gx[0];
    A.pb[4 * S.NG + gl] += gx[1];
    A.pb[5 * S.NG + gl] += gx[2];
    A.pb[6 * S.NG + gl] += gx[3];
}

// deterministic sum of n doubles in two stages: workgroup b sums its strided
// share (q = b*256 + t + k*256*gridDim) by a fixed tree into part[b]; then
// one workgroup sums the parts the same way.  The order never depends on
// timing, so the result is reproducible bit for bit.
constexpr int SUM_WG = 512;
__global__ __launch_bounds__(256) void k_sum_fixed(const double *a, long n, double *part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += 256L * gridDim.x) s += a[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// deterministic sum over q of w[q] * (a[q] - b[q]) * inv_h2, strided as k_sum_fixed
__global__ __launch_bounds__(256) void k_dotdiff_fixed(const double *w, const double *a, const double *b, long n,
                                                       double inv_h2, double *part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += 256L * gridDim.x) s += w[q] * ((a[q] - b[q]) * inv_h2);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int v = 128; v > 0; v >>= 1) {
        if ((int)threadIdx.x < v) sh[threadIdx.x] += sh[threadIdx.x + v];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// stencil points a face's bars cover under the configuration
unsigned long long adj_mask(const Par &P)
{
    unsigned long long m = 0;
    auto set = [&](int ll) { m |= 1ull << (ll - 1); };
    set(9); set(10);
    if (P.nmuscl == 1) { set(23); set(32); set(37); set(38); }
    if (P.nmuscl == 2) for (int ll = 1; ll <= NSLOT; ll++) set(ll);
    if (P.nscheme == 2 || P.nscheme == 3 || P.iturb >= -2) for (int ll = 1; ll <= 18; ll++) set(ll);
    return m;
}

int adj_alloc(cav3d_ctx *c)
{
    if (c->adj_ready) return 0;
    DS &S = c->S;
    const int npv = c->npv;
    int rc;
    if ((rc = dalloc(c, &c->a_rb, (long)npv * S.NC)) || (rc = dalloc(c, &c->a_cb, 7L * npv * S.NC)) ||
        (rc = dalloc(c, &c->a_cdb, (long)ND * S.NC)) || (rc = dalloc(c, &c->a_csb, 3L * S.NC)) || (rc = dalloc(c, &c->a_pb, (long)npv * S.NG)) ||
        (rc = dalloc(c, &c->a_gdb, (long)ND * S.NG)) || (rc = dalloc(c, &c->a_sum, 16 + 16L * SUM_WG)))
        return rc;
    for (int d = 0; d < 3; d++)
        if ((rc = dalloc(c, &c->a_fb[d], (long)NSLOT * npv * S.NF[d])) || (rc = dalloc(c, &c->a_db[d], 2L * ND * S.NF[d])) ||
            (rc = dalloc(c, &c->a_rsb[d], 4L * npv * S.NF[d])) ||
            (rc = dalloc(c, &c->a_snb[d], (S.P.nscheme == 2 || S.P.nscheme == 3) ? 18L * S.NF[d] : 1L)) ||
            (rc = dalloc(c, &c->a_ub[d], S.NF[d])) || (rc = dalloc(c, &c->a_mtb[d], 2L * S.NF[d])))
            return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    c->adj_ready = true;
    return 0;
}

int launch_vjp(cav3d_ctx *c, int dmut = 0)
{
    DS &S = c->S;
    AD A;
    A.rb = c->a_rb;
    A.dmut = dmut;
    for (int d = 0; d < 3; d++) {
        A.fb[d] = c->a_fb[d]; A.db[d] = c->a_db[d]; A.ub[d] = c->a_ub[d]; A.mtb[d] = c->a_mtb[d];
        A.rsb[d] = c->a_rsb[d]; A.snb[d] = c->a_snb[d];
    }
    A.cb = c->a_cb; A.cdb = c->a_cdb; A.csb = c->a_csb; A.pb = c->a_pb; A.gdb = c->a_gdb;
    A.mask = adj_mask(S.P);
    const bool sensor = S.P.nscheme == 2 || S.P.nscheme == 3;
    const int npv = S.npv;
    // the three directions write their own face bars: i on st, j and k on
    // the side streams, joined before the cell sources and the gather
    hipStream_t sd[3] = {c->st, c->side[0], c->side[1]};
    hipEventRecord(c->fork, c->st);
    hipStreamWaitEvent(c->side[0], c->fork, 0);
    hipStreamWaitEvent(c->side[1], c->fork, 0);
#define CAV3D_VJP_DIR(D)                                                                                             \
    do {                                                                                                             \
        if (sensor) {                                                                                                \
            hipLaunchKernelGGL((k_face_vjp<D, true>), dim3(nblk(S.NF[D], 128)), dim3(128), 0, sd[D], S, A);          \
            hipLaunchKernelGGL((k_face_rvjp<D, true>), dim3(nblk(npv * S.NF[D], 256)), dim3(256), 0, sd[D], S, A);   \
        } else {                                                                                                     \
            hipLaunchKernelGGL((k_face_vjp<D, false>), dim3(nblk(S.NF[D], 128)), dim3(128), 0, sd[D], S, A);         \
            hipLaunchKernelGGL((k_face_rvjp<D, false>), dim3(nblk(npv * S.NF[D], 256)), dim3(256), 0, sd[D], S, A);  \
        }                                                                                                            \
        if (S.P.iturb >= -2) hipLaunchKernelGGL((k_vface_vjp<D>), dim3(nblk(S.NF[D], 128)), dim3(128), 0, sd[D], S, A); \
    } while (0)
    CAV3D_VJP_DIR(1);
    CAV3D_VJP_DIR(2);
    CAV3D_VJP_DIR(0);
#undef CAV3D_VJP_DIR
    hipEventRecord(c->join[0], c->side[0]);
    hipEventRecord(c->join[1], c->side[1]);
    hipStreamWaitEvent(c->st, c->join[0], 0);
    hipStreamWaitEvent(c->st, c->join[1], 0);
    // the cell sources read every direction's eddy-viscosity bars
    if (S.P.ides > 0 || S.P.ncav >= 2) hipLaunchKernelGGL(k_cell_vjp<true>, dim3(nblk(S.NC, 128)), dim3(128), 0, c->st, S, A);
    else hipLaunchKernelGGL(k_cell_vjp<false>, dim3(nblk(S.NC, 128)), dim3(128), 0, c->st, S, A);
    hipLaunchKernelGGL(k_adj_gather, dim3(nblk(S.NG, 256)), dim3(256), 0, c->st, S, A);
    hipLaunchKernelGGL(k_adj_eos, dim3(nblk(S.NG, 128)), dim3(128), 0, c->st, S, A);
    return 0;
}

// every block's res_bar (the calrhs res layout, blocks concatenated) into a_rb
int upload_res_bar(cav3d_ctx *c, const double *res_bar)
{
    int rc;
    if ((rc = adj_alloc(c))) return rc;
    DS &S = c->S;
    const int npv = c->npv;
    long maxn = 0;
    for (int b = 0; b < c->nb; b++) maxn = std::max(maxn, (long)npv * c->blk[b].ng);
    if ((rc = ensure_stage(c, maxn))) return rc;
    long off = 0;
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        HIPCHK(hipMemcpyAsync(c->stage, res_bar + off, (long)npv * B.nc * sizeof(double), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_cell_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage, npv, S.NC,
                           B.c0, c->a_rb, B.nc);
        HIPCHK(hipStreamSynchronize(c->st));
        off += (long)npv * B.nc;
    }
    return 0;
}

// the fixed-order sum of n values at src into a_sum[slot]
void sum_into(cav3d_ctx *c, const double *src, long n, int slot)
{
    double *part = c->a_sum + 16 + (long)SUM_WG * slot;
    hipLaunchKernelGGL(k_sum_fixed, dim3(SUM_WG), dim3(256), 0, c->st, src, n, part);
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, c->st, (const double *)part, (long)SUM_WG, c->a_sum + slot);
}

}  // namespace

extern "C" {

static int rhs_vjp_run(cav3d_ctx *c, const double *res_bar, double *pv_bar, double *scal_bar, int dmut, const char *name)
{
    if (!c || !res_bar || !pv_bar) return fail(std::string("bad argument to ") + name, -1);
    int rc;
    if ((rc = upload_res_bar(c, res_bar))) return rc;
    DS &S = c->S;
    const int npv = c->npv;
    long off = 0;
    if (c->profile) hipEventRecord(c->ev[10], c->st);
    launch_vjp(c, dmut);
    if (c->profile) {
        hipEventRecord(c->ev[11], c->st);
        hipEventSynchronize(c->ev[11]);
        float ms;
        hipEventElapsedTime(&ms, c->ev[10], c->ev[11]);
        c->prof_ms[8] = ms;
    }
    if ((rc = check_launch(name))) return rc;
    off = 0;
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, (const double *)c->a_pb, npv, S.NG, B.g0,
                           c->stage, B.ng);
        HIPCHK(hipMemcpyAsync(pv_bar + off, c->stage, (long)npv * B.ng * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        off += (long)npv * B.ng;
    }
    if (scal_bar) {
        const double *src[5] = {c->a_pb, c->a_ub[0], c->a_ub[1], c->a_ub[2], c->a_csb};
        const long len[5] = {S.NG, S.NF[0], S.NF[1], S.NF[2], S.NC};
        for (int q = 0; q < 5; q++) sum_into(c, src[q], len[q], q);
        double h[5];
        HIPCHK(hipMemcpyAsync(h, c->a_sum, sizeof(h), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        scal_bar[0] = h[0];
        scal_bar[1] = ((h[1] + h[2]) + h[3]) + h[4];
    }
    return check_launch(name);
}

int cav3d_rhs_vjp(cav3d_ctx *c, const double *res_bar, double *pv_bar, double *scal_bar)
{
    return rhs_vjp_run(c, res_bar, pv_bar, scal_bar, 0, "cav3d_rhs_vjp");
}

int cav3d_rhs_vjp_eddy(cav3d_ctx *c, const double *res_bar, double *pv_bar, double *scal_bar)
{
    if (c && c->prm.iturb != 0) return fail("cav3d_rhs_vjp_eddy: the eddy viscosity is a state function only under k-omega SST", -1);
    return rhs_vjp_run(c, res_bar, pv_bar, scal_bar, 1, "cav3d_rhs_vjp_eddy");
}

int cav3d_eddy(cav3d_ctx *c)
{
    if (!c) return fail("bad argument to cav3d_eddy", -1);
    if (c->prm.iturb != 0) return fail("cav3d_eddy: the eddy viscosity is computed only under k-omega SST", -1);
    const DS &S = c->S;
    hipLaunchKernelGGL(k_eddy, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S);
    return check_launch("cav3d_eddy");
}

int cav3d_rhs_vjp_params(cav3d_ctx *c, const double *res_bar, double *par_bar)
{
    if (!c || !res_bar || !par_bar) return fail("bad argument to cav3d_rhs_vjp_params", -1);
    int rc;
    if ((rc = upload_res_bar(c, res_bar))) return rc;
    DS &S = c->S;
    const int npv = c->npv;
    if (!c->a_save) {
        if ((rc = dalloc(c, &c->a_save, (long)(2 * npv + ND + 1) * S.NG))) return rc;
    }
    launch_vjp(c);
    // the partials at fixed ghost cells
    sum_into(c, c->a_pb, S.NG, 0);
    sum_into(c, c->a_ub[0], S.NF[0], 1);
    sum_into(c, c->a_ub[1], S.NF[1], 2);
    sum_into(c, c->a_ub[2], S.NF[2], 3);
    for (int q = 0; q < 3; q++) sum_into(c, c->a_csb + (long)q * S.NC, S.NC, 4 + q);
    // the boundary fill's share: for a fixed interior every ghost cell is an
    // affine function of the boundary state bc_pv, bc_pressure (the mirror,
    // surface and neighbour cells and the in/out mode do not move), so the
    // product of the ghost-cell bars with its derivative is a central
    // difference of the fill, exact up to rounding.  The characteristic
    // fills also read u_ref (the preconditioned speed of sound at the
    // surface cell), smoothly where the surface cell's normal speed is not
    // u_ref: the eleventh difference.  The fills also read the surface
    // cells' density and speed of sound, which follow p_ref through the EOS:
    // the twelfth difference refills from the interior properties
    // re-evaluated at p_ref +- h.  The fills evaluate the ghost EOS at
    // bc_pv[0] + gauge pressure, so p_ref is exact together with bc_pv[0]
    // when the two are equal (a change of p_ref that moves the boundary
    // state's reference pressure with it).
    double *save_pv = c->a_save, *save_dc = save_pv + (long)npv * S.NG, *save_mut = save_dc + (long)ND * S.NG;
    double *plus = save_mut + S.NG;
    HIPCHK(hipMemcpyAsync(save_pv, S.pv, (long)npv * S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(save_dc, S.dc, (long)ND * S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    if (S.mut) HIPCHK(hipMemcpyAsync(save_mut, S.mut, S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    const Par P0 = S.P;
    double hs[12], share_u = 0.0, share_p = 0.0;
    for (int m = 0; m < 12; m++) {
        double &theta = (m < 9) ? S.P.bc_pv[m] : (m == 9 ? S.P.bc_pressure : (m == 10 ? S.P.uref : S.P.pref));
        const double v0 = theta;
        const double h = v0 != 0.0 ? 1.0e-4 * fabs(v0) : 1.0e-6;
        hs[m] = h;
        theta = v0 + h;
        if (m == 11) hipLaunchKernelGGL(k_eos_cells, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S, 1, 0);
        launch_setbc(c);
        HIPCHK(hipMemcpyAsync(plus, S.pv, (long)npv * S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
        theta = v0 - h;
        if (m == 11) hipLaunchKernelGGL(k_eos_cells, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S, 1, 0);
        launch_setbc(c);
        theta = v0;
        double *part = c->a_sum + 16 + (long)SUM_WG * (7 + (m % 8));
        hipLaunchKernelGGL(k_dotdiff_fixed, dim3(SUM_WG), dim3(256), 0, c->st, (const double *)c->a_pb, (const double *)plus,
                           (const double *)S.pv, (long)npv * S.NG, 1.0 / (2.0 * h), part);
        hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, c->st, (const double *)part, (long)SUM_WG, c->a_sum + 7 + m % 8);
        if (m == 7 || m == 11) {
            // a_sum[7..14] hold up to eight results: copy them out before reuse
            double hm[8];
            HIPCHK(hipMemcpyAsync(hm, c->a_sum + 7, sizeof(hm), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            const int first = m == 7 ? 0 : 8;
            for (int q = first; q <= m; q++) {
                if (q < 10) par_bar[4 + q] = hm[q % 8];
                else if (q == 10) share_u = hm[q % 8];
                else share_p = hm[q % 8];
            }
        }
    }
    S.P = P0;
    HIPCHK(hipMemcpyAsync(S.pv, save_pv, (long)npv * S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(S.dc, save_dc, (long)ND * S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    if (S.mut) HIPCHK(hipMemcpyAsync(S.mut, save_mut, S.NG * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    double h[7];
    HIPCHK(hipMemcpyAsync(h, c->a_sum, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    par_bar[0] = h[0] + share_p;
    par_bar[1] = (((h[1] + h[2]) + h[3]) + h[4]) + share_u;
    par_bar[2] = h[5];
    par_bar[3] = h[6];
    return check_launch("cav3d_rhs_vjp_params");
}

int cav3d_objective(cav3d_ctx *c, double w_vap, double w_fx, double *J, double *pv_bar, double *dJ_dpref)
{
    if (!c || !J) return fail("bad argument to cav3d_objective", -1);
    DS &S = c->S;
    long nface;
    bool ok;
    std::vector<FSeg> segs = surface_segments(c, nface, ok);
    if (!ok) return fail("cav3d_objective: a surface range is not a block face inside the cells", -2);
    if (nface > 0 && w_fx != 0.0 && !c->xv) return fail("cav3d_objective: the x-force needs the vertex coordinates (cav3d_set_grid)", -1);
    int rc;
    if ((rc = adj_alloc(c))) return rc;
    if (c->a_obj_n < S.NC + nface) {
        if (c->a_obj) { HIPCHK(hipStreamSynchronize(c->st)); HIPCHK(hipFree(c->a_obj)); c->a_obj = nullptr; }
        HIPCHK(hipMalloc((void **)&c->a_obj, 2 * (S.NC + nface) * sizeof(double)));
        c->a_obj_n = S.NC + nface;
    }
    double *val = c->a_obj, *dpr = c->a_obj + c->a_obj_n;
    // pv_bar: zero on the ghost layers, the cells' shares, then each range's faces
    hipLaunchKernelGGL(k_fill, dim3(nblk((long)c->npv * S.NG, 256)), dim3(256), 0, c->st, c->a_pb, 0.0, (long)c->npv * S.NG);
    hipLaunchKernelGGL(k_obj_cells, dim3(nblk(S.NC, POST_WG)), dim3(POST_WG), 0, c->st, S, (const double *)S.vol, w_vap,
                       c->a_pb, val, dpr);
    long off = S.NC;
    if (w_fx != 0.0)
        for (const FSeg &sg : segs) {
            hipLaunchKernelGGL(k_obj_faces, dim3(nblk(sg.n, POST_WG)), dim3(POST_WG), 0, c->st, S, sg, (const double *)c->xv,
                               c->NV, c->v0[sg.b], (const double *)c->xc, w_fx, c->a_pb, val + off, dpr + off);
            off += sg.n;
        }
    sum_into(c, val, off, 0);
    sum_into(c, dpr, off, 1);
    double h[2];
    HIPCHK(hipMemcpyAsync(h, c->a_sum, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *J = h[0];
    if (dJ_dpref) *dJ_dpref = h[1];
    if (pv_bar) {
        const int npv = c->npv;
        long maxn = 0;
        for (int b = 0; b < c->nb; b++) maxn = std::max(maxn, (long)npv * c->blk[b].ng);
        if ((rc = ensure_stage(c, maxn))) return rc;
        long o = 0;
        for (int b = 0; b < c->nb; b++) {
            const BlkD &B = c->blk[b];
            hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, (const double *)c->a_pb, npv, S.NG,
                               B.g0, c->stage, B.ng);
            HIPCHK(hipMemcpyAsync(pv_bar + o, c->stage, (long)npv * B.ng * sizeof(double), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            o += (long)npv * B.ng;
        }
    }
    return check_launch("cav3d_objective");
}

}  // extern "C"

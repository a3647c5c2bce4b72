// Host build of the library's pointwise physics (3dcode_amd/csrc/phys.hpp)
// with C entry points, so CPU tests can compare each routine with the
// oracle on identical inputs.  Test infrastructure only.
#include "../../3dcode_amd/csrc/phys.hpp"
#include "../../3dcode_amd/csrc/adjoint_point.hpp"

using namespace cav;

static Par mkpar(const int *iv, const double *dv)
{
    Par P{};
    P.npv = iv[0]; P.ntv = iv[1]; P.iturb = iv[2]; P.nscheme = iv[3]; P.nmuscl = iv[4]; P.nlim = iv[5];
    P.local = iv[6]; P.prec = iv[7]; P.ncav = iv[8]; P.nsteady = iv[9]; P.cav_clip = iv[10]; P.ides = iv[11];
    P.cfl = dv[0]; P.pref = dv[1]; P.uref = dv[2]; P.str = dv[3]; P.dt_phy = dv[4]; P.c_v = dv[5];
    P.c_c = dv[6]; P.dp_ref = dv[7]; P.t_ref = dv[8]; P.t_crit = dv[9];
    for (int q = 0; q < 9; q++) P.bc_pv[q] = dv[10 + q];
    P.bc_pressure = dv[19]; P.emutref = dv[20]; P.bdf_b = dv[21];
    return P;
}

static CellJ mkcell(const double *pv, const double *d, double mut, double vol)
{
    CellJ c;
    for (int n = 0; n < NPV; n++) c.pv[n] = pv[n];
    for (int m = 0; m < ND; m++) c.d[m] = d[m];
    c.mut = mut; c.vol = vol;
    return c;
}

extern "C" {

void pc_liquid(double p, double t, double *o)
{
    E2<double> e = liquid<double>(p, t);
    o[0] = e.rho; o[1] = e.h; o[2] = e.drdp; o[3] = e.drdt; o[4] = e.dhdp; o[5] = e.dhdt;
}

void pc_vapour(double p, double t, double *o)
{
    E2<double> e = vapour<double>(p, t, sat_p(t));
    o[0] = e.rho; o[1] = e.h; o[2] = e.drdp; o[3] = e.drdt; o[4] = e.dhdp; o[5] = e.dhdt;
}

void pc_mixture(double p, double t, double y1, double y2, double *full, double *tv)
{
    double d[ND];
    mixture_eos<double>(p, t, y1, y2, d, true, full);
    tv[0] = d[D_MU]; tv[1] = d[D_KAP];
}

// x: 38 x 9 (row ll-1, column n); fx: inviscid flux; g3: 6 x 3
void pc_face(const int *iv, const double *dv, const double *x, const double *n3, const double *rh9,
             const double *rh10, double *fx, double *xl, double *xr, double *g3)
{
    Par P = mkpar(iv, dv);
    double c9[NPV], c10[NPV], g[6][3] = {}, sd[18];
    for (int ll = 1; ll <= 18; ll++) sd[ll - 1] = x[(ll - 1) * 9];
    face_recon<double>(P, P.npv, [&](int ll, int n) { return x[(ll - 1) * 9 + n]; }, c9, c10, xl, xr, g);
    if (P.nscheme == 2 || P.nscheme == 3) inviscid_face<double, true>(P, P.npv, n3, c9, c10, xl, xr, rh9, rh10, fx, sd);
    else inviscid_face<double, false>(P, P.npv, n3, c9, c10, xl, xr, rh9, rh10, fx);
    for (int q = 0; q < 18; q++) g3[q] = g[q / 3][q % 3];
}

// the same face in FP32 (the mixed-precision kernels' arithmetic)
void pc_face32(const int *iv, const double *dv, const double *x, const double *n3, const double *rh9,
               const double *rh10, double *fx)
{
    Par P = mkpar(iv, dv);
    float c9[NPV], c10[NPV], xl[NPV], xr[NPV], g[6][3] = {}, sd[18], f[NPV];
    for (int ll = 1; ll <= 18; ll++) sd[ll - 1] = float(x[(ll - 1) * 9]);
    face_recon<float>(P, P.npv, [&](int ll, int n) { return float(x[(ll - 1) * 9 + n]); }, c9, c10, xl, xr, g);
    const float n3f[3] = {float(n3[0]), float(n3[1]), float(n3[2])};
    const float r9[3] = {float(rh9[0]), float(rh9[1]), float(rh9[2])}, r10[3] = {float(rh10[0]), float(rh10[1]), float(rh10[2])};
    if (P.nscheme == 2 || P.nscheme == 3) inviscid_face<float, true>(P, P.npv, n3f, c9, c10, xl, xr, r9, r10, f, sd);
    else inviscid_face<float, false>(P, P.npv, n3f, c9, c10, xl, xr, r9, r10, f);
    for (int n = 0; n < P.npv; n++) fx[n] = double(f[n]);
}

void pc_visc(const int *iv, const double *dv, const double *x, const double *n3, const double *e4,
             const double *t4, const double *gl, const double *gr, const double *dl, const double *dr,
             double mutl, double mutr, double *fv)
{
    Par P = mkpar(iv, dv);
    double c9[NPV], c10[NPV], xl[NPV], xr[NPV], g[6][3] = {};
    face_recon<double>(P, P.npv, [&](int ll, int n) { return x[(ll - 1) * 9 + n]; }, c9, c10, xl, xr, g);
    double e[4][3], t[4][3];
    for (int q = 0; q < 12; q++) { e[q / 3][q % 3] = e4[q]; t[q / 3][q % 3] = t4[q]; }
    double cav[5] = {0.5 * (c9[1] + c10[1]), 0.5 * (c9[2] + c10[2]), 0.5 * (c9[3] + c10[3]), 0.0, 0.0};
    if (P.npv > 7) { cav[3] = 0.5 * (c9[7] + c10[7]); cav[4] = 0.5 * (c9[8] + c10[8]); }
    visc_flux<double>(P, n3, e, t, gl, gr, g, cav, dl, dr, mutl, mutr, fv);
}

double pc_localdt(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
                  const double *f6)
{
    Par P = mkpar(iv, dv);
    double f[6][3];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    return local_dt(P, mkcell(pv, d, mut, vol), f);
}

void pc_getx(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
             const double *f6, double dt, const double *cc, const double *tt, const double *res, double *x)
{
    Par P = mkpar(iv, dv);
    double f[6][3];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    getx(P, P.npv, mkcell(pv, d, mut, vol), f, dt, cc, tt, res, x);
}

// y = A x with A the split Jacobian of the given sign
void pc_jac(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
            const double *n3, int sign, const double *x, double *y)
{
    Par P = mkpar(iv, dv);
    for (int n = 0; n < NPV; n++) y[n] = 0.0;
    jac_apply(P, P.npv, mkcell(pv, d, mut, vol), n3, sign, x, y, 1.0);
}

// the same product through the branch-free row coefficients the sweeps use
void pc_jac_coef(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
                 const double *n3, int sign, double sgn, const double *x, double *y)
{
    Par P = mkpar(iv, dv);
    CellJ c = mkcell(pv, d, mut, vol);
    JacCell jc = jac_cell(P, c);
    JacFace jf = jac_face(P, c, n3);
    for (int n = 0; n < NPV; n++) {
        double cf[8];
        jac_row_coef(P.npv, jac_nb(c.pv, c.d, n3, jc, jf, n, P.npv), sign, n, sgn, cf);
        jac_row_apply(cf, x, x[n], y[n]);
    }
}

void pc_jac_rows(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
                 const double *n3, int sign, double sgn, const double *x, double *y)
{
    Par P = mkpar(iv, dv);
    jac_apply(P, P.npv, mkcell(pv, d, mut, vol), n3, sign, x, y, sgn);
}

// getx through the row-selecting form the sweeps use
void pc_getx_sel(const int *iv, const double *dv, const double *pv, const double *d, double mut, double vol,
                 const double *f6, double dt, const double *cc, const double *tt, const double *res, double *x)
{
    Par P = mkpar(iv, dv);
    double f[6][3], g[NGX];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    CellJ c = mkcell(pv, d, mut, vol);
    getx_prep(P, P.npv, c, f, dt, cc, tt, g);
    for (int n = 0; n < NPV; n++) if (n < P.npv)
        x[n] = getx_apply_sel([&](int q) { return g[q]; }, c.pv, c.d[D_RY1], c.d[D_RT], res, n, res[n], c.pv[n]);
}

void pc_merkle(const int *iv, const double *dv, const double *pc, const double *d, double vol, double *out)
{
    Par P = mkpar(iv, dv);
    double rc, rv, ic[4];
    merkle(P, pc, d, vol, rc, rv, ic);
    out[0] = rc; out[1] = rv;
    for (int q = 0; q < 4; q++) out[2 + q] = ic[q];
}

void pc_cav_model(const int *iv, const double *dv, const double *pc, const double *full, double vol, double *out)
{
    Par P = mkpar(iv, dv);
    double rc, rv, ic[4];
    cav_model(P, pc, full, vol, rc, rv, ic);
    out[0] = rc; out[1] = rv;
    for (int q = 0; q < 4; q++) out[2 + q] = ic[q];
}

// rv - rc of the mass-transfer model (ncav 1-5) on the cell's own EOS, and its
// gradient over (p, u, v, w, T, y1, y2) and u_ref by the dual passes the
// adjoint kernel runs
void pc_cav_adj(const int *iv, const double *dv, const double *pv, double *val, double *grad)
{
    // ncav 1 is Merkle on the compact properties, 2..5 the models on dv(18)
    Par P = mkpar(iv, dv);
    using T = Dn<1>;
    for (int q = 0; q < 8; q++) {
        T pc[NPV];
        for (int n = 0; n < NPV; n++) pc[n] = T(pv[n]);
        Par Pu = P;
        if (q < 7) pc[q] = T::seed(pv[q], 0);
        else Pu.uref_seed = 1;
        T dd[ND], full[18], rc, rv, ic[4];
        mixture_eos<T>(pc[0] + P.pref, pc[4], pc[5], pc[6], dd, false, full);
        if (P.ncav == 1) merkle<T>(Pu, pc, dd, 1.0, rc, rv, ic);
        else cav_model<T>(Pu, pc, full, 1.0, rc, rv, ic);
        val[0] = (rv - rc).v;
        grad[q] = (rv - rc).d[0];
    }
}

double pc_sat_t(double p) { return sat_t(p); }
double pc_h2o_sigma(double t) { return h2o_sigma(t); }

void pc_sst_source(const double *x7, const double *f6, double vol, double ydns, const double *d, double mut,
                   double *out)
{
    double x[7][NPV], f[6][3];
    for (int q = 0; q < 63; q++) x[q / 9][q % 9] = x7[q];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    double src[2], itt[4], oc;
    Par P{};
    sst_source<false>(P, x, f, vol, ydns, d, mut, src, itt, oc);
    out[0] = src[0]; out[1] = src[1];
    for (int q = 0; q < 4; q++) out[2 + q] = itt[q];
    out[6] = oc;
}

void pc_sst_source_des(const int *iv, const double *dv, const double *x7, const double *f6, double vol, double ydns, const double *d, double mut,
                   double *out)
{
    double x[7][NPV], f[6][3];
    for (int q = 0; q < 63; q++) x[q / 9][q % 9] = x7[q];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    double src[2], itt[4], oc;
    Par P = mkpar(iv, dv);
    sst_source<true>(P, x, f, vol, ydns, d, mut, src, itt, oc);
    out[0] = src[0]; out[1] = src[1];
    for (int q = 0; q < 4; q++) out[2 + q] = itt[q];
    out[6] = oc;
}

double pc_sst_eddy(const double *x7, const double *f6, double vol, double ydns, const double *d)
{
    double x[7][NPV], f[6][3];
    for (int q = 0; q < 63; q++) x[q / 9][q % 9] = x7[q];
    for (int q = 0; q < 18; q++) f[q / 3][q % 3] = f6[q];
    return sst_eddy(x, f, vol, ydns, d);
}


// gradient of g . F of one face (the adjoint's inviscid part) on the 38x9
// stencil x[(ll-1)*9 + n]; rh9/rh10 = {rho, h, c^2} of the two cells
void pc_face_adj(const int *iv, const double *dv, const double *x, const double *n3, const double *rh9,
                 const double *rh10, const double *g, double *sbo, double *dcb, double *ub)
{
    Par P = mkpar(iv, dv);
    long off[NSLOT];
    for (int ll = 1; ll <= NSLOT; ll++) off[ll - 1] = (long)(ll - 1) * 9;
    const double cellp[6] = {rh9[0], rh9[1], rh9[2], rh10[0], rh10[1], rh10[2]};
    const unsigned long long all = (1ull << NSLOT) - 1;
    if (P.nscheme == 2 || P.nscheme == 3) face_adj<true>(P, P.npv, x, 1, off, n3, cellp, g, all, sbo, 1, NSLOT, dcb, 1, *ub);
    else face_adj<false>(P, P.npv, x, 1, off, n3, cellp, g, all, sbo, 1, NSLOT, dcb, 1, *ub);
}

// gradient of phi = sum_m w[m] d[m] of the cell properties with respect to (p, T, y1, y2)
void pc_eos_adj(int prop, const double *x4, const double *w, double *gx)
{
    eos_adj(prop != 0, x4, w, gx);
}
}

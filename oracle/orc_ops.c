/*
 * ORACLE — test infrastructure only (see orc_eos.c header).
 *
 * Pointwise operators of the reference pseudo-time step, restated in C in
 * the reference's operation order:
 *   MUSCL-TVD / MLP + limiters   src/muscl_module.f90:104-413
 *   preconditioned Roe            src/flux_module.f90:198-276, 643-695
 *   AUSM+-up                      src/flux_module.f90:548-641
 *   viscous flux + SST F1         src/vsflux_module.f90:63-99, 182-425
 *   Roe-M, AUSMPW+                src/flux_module.f90:278-418, 420-546
 *   Merkle mass transfer          src/cav_module.f90:67-97, 136-171
 *   Kunz, Singhal, Schnerr-Sauer,
 *   Lee mass transfer             src/cav_module.f90:173-356
 *   k-omega SST source            src/turbsource_module.f90:269-414
 *   DES/DDES/IDDES/S-IDDES length src/turbsource_module.f90:416-576
 *   SST eddy viscosity            src/eddy_module.f90:135-187
 *   split flux Jacobians          src/jacobian_module.f90:182-465
 *   diagonal inverse getx         src/lhs_module.f90:346-534
 * Stencil arrays are x[ll][n] with ll,n 0-based (= Fortran x(ll+1,n+1)).
 *
 * Pins (tests/test_oracle_closed_forms.py, test_oracle_invariants.py,
 * test_phys_host.py): flux consistency and antisymmetry, exactness of every
 * limiter on linear data, the viscous stress of linear fields, the SST
 * source and eddy viscosity at their limits, the five mass-transfer models'
 * rates from their publications (no non-condensable gas) and their
 * linearisations against finite differences, the DES family's lengths and
 * shielding functions from their publications, the Jacobian splitting, getx
 * against an assembled inverse.  Parity unpinned: the mass-transfer rates'
 * non-condensable-gas (y2 > 0) terms, which follow the reference's source
 * text only.
 */
#include "orc_int.h"
#include <math.h>

/* ------------------------------------------------------------------ */
/* preconditioning: flux/jacobian/lhs/timestep getsndp2 variants       */
/* (src/flux_module.f90:643-671, jacobian :437-465, lhs :536-564)      */
double orc_sndp2(const orc_cfg *c, double snd2, double uuu2)
{
    switch (c->prec) {
    case 0: return snd2;
    case 1: return fmin_f(snd2, fmax_f(uuu2, c->uref * c->uref));
    default: return fmin_f(snd2, fmax_f(fmax_f(uuu2, c->uref * c->uref), c->str * c->str));
    }
}

/* ------------------------------------------------------------------ */
/* limiters, src/muscl_module.f90:292-413                              */
typedef struct { double r[2], r1[2], r2[2], alp[2]; } mstate;

static double dmin3(double a, double b, double c) { return fmin_f(fmin_f(a, b), c); }
static double dmax3(double a, double b, double c) { return fmax_f(fmax_f(a, b), c); }

static double limiter(int nlim, const mstate *m, int n)
{
    double r = m->r[n], alp = m->alp[n], c;
    switch (nlim) {
    case 0: return 1.0;
    case 1: return fmax_f(0.0, fmin_f(1.0, r));
    case 2: return dmax3(0.0, fmin_f(1.0, alp * r), fmin_f(alp, r));
    case 3: c = (r + fabs(r)) / (1.0 + fabs(r)); return fmax_f(0.0, dmin3(alp, alp * r, c));
    case 4: c = (1.0 + r) / 2.0; return fmax_f(0.0, dmin3(alp, alp * r, c));
    case 5: c = (r * r + r) / (1.0 + r * r); return dmax3(0.0, fmin_f(c * r, 1.0), fmin_f(r, c));
    case 6: c = fmin_f(1.5, alp); return dmax3(0.0, fmin_f(c * r, 1.0), fmin_f(r, c));
    case 7: c = (1.0 + 2.0 * r) / 3.0; return fmax_f(0.0, dmin3(alp, alp * r, c));
    case 8: c = (-(2.0 * m->r1[n]) + 11.0 + 24.0 * r - 3.0 * m->r2[n]) / 30.0;
            return fmax_f(0.0, dmin3(alp, alp * r, c));
    case 9: return (1.0 + 2.0 * r) / 3.0;
    case 10: return (-(2.0 * m->r1[n]) + 11.0 + 24.0 * r - 3.0 * m->r2[n]) / 30.0;
    }
    return 1.0;
}

#define X(ll, k) x[(ll) - 1][(k) - 1]

/* shared ratio set of tvd/mlp, src/muscl_module.f90:113-127, 172-186 */
static void ratios(const double (*x)[ORC_NPV], int k, mstate *m,
                   double *dq, double *dqm, double *dqp, double *dqpp)
{
    *dq = X(10, k) - X(9, k);
    *dqm = X(9, k) - X(23, k);
    double dqm1 = X(23, k) - X(37, k);
    *dqp = X(32, k) - X(10, k);
    double dqp1 = X(38, k) - X(32, k);
    const double eps = 1.0e-30;
    double dqmm = 1.0 / (*dqm + eps);
    m->r[0] = *dq * dqmm;
    m->r1[0] = dqm1 * dqmm;
    m->r2[0] = *dqp * dqmm;
    *dqpp = 1.0 / (*dqp + eps);
    m->r[1] = *dq * *dqpp;
    m->r1[1] = dqp1 * *dqpp;
    m->r2[1] = *dqm * *dqpp;
}

/* positivity / fraction / turbulence clamps, src/muscl_module.f90:136-158 */
static void clamps(const double (*x)[ORC_NPV], int npv, double pref, double *xl, double *xr)
{
    xl[0] = fmax_f(-pref + 1.0e1, xl[0]);
    xr[0] = fmax_f(-pref + 1.0e1, xr[0]);
    if ((xl[5] < 0.0) || (xl[5] > 1.0)) xl[5] = X(9, 6);
    if ((xr[5] < 0.0) || (xr[5] > 1.0)) xr[5] = X(10, 6);
    if ((xl[6] < 0.0) || (xl[6] > 1.0)) xl[6] = X(9, 7);
    if ((xr[6] < 0.0) || (xr[6] > 1.0)) xr[6] = X(10, 7);
    if ((xl[5] + xl[6]) > 1.0) { xl[5] = X(9, 6); xl[6] = X(9, 7); }
    if ((xr[5] + xr[6]) > 1.0) { xr[5] = X(10, 6); xr[6] = X(10, 7); }
    if (npv > 7) {
        if ((xl[7] < 0.0) && (X(9, 8) > 0.0)) xl[7] = X(9, 8);
        if ((xr[7] < 0.0) && (X(10, 8) > 0.0)) xr[7] = X(10, 8);
        if ((xl[8] < 0.0) && (X(9, 9) > 0.0)) xl[8] = X(9, 9);
        if ((xr[8] < 0.0) && (X(10, 9) > 0.0)) xr[8] = X(10, 9);
    }
}

/* src/muscl_module.f90:104-160 */
void orc_tvd(const orc_cfg *c, const double (*x)[ORC_NPV], double *xl, double *xr)
{
    mstate m;
    for (int k = 1; k <= c->npv; k++) {
        double dq, dqm, dqp, dqpp;
        ratios(x, k, &m, &dq, &dqm, &dqp, &dqpp);
        m.alp[0] = 2.0;
        m.alp[1] = 2.0;
        xl[k - 1] = X(9, k) + 0.5 * limiter(c->nlim, &m, 0) * dqm;
        xr[k - 1] = X(10, k) - 0.5 * limiter(c->nlim, &m, 1) * dqp;
    }
    clamps(x, c->npv, c->pref, xl, xr);
}

static double minlist(const double (*x)[ORC_NPV], int k, const int *ll, int n)
{
    double v = X(ll[0], k);
    for (int q = 1; q < n; q++) v = fmin_f(v, X(ll[q], k));
    return v;
}
static double maxlist(const double (*x)[ORC_NPV], int k, const int *ll, int n)
{
    double v = X(ll[0], k);
    for (int q = 1; q < n; q++) v = fmax_f(v, X(ll[q], k));
    return v;
}

/* index sets of the MLP min/max searches, src/muscl_module.f90:194-236 */
static const int S_L27[27] = {1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16,17,18,19,20,21,22,23,24,25,26,27};
static const int S_R27[27] = {1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16,17,18,28,29,30,31,32,33,34,35,36};
static const int S_18[18] = {1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16,17,18};
static const int S_L18[18] = {1,11,3,13,5,15,7,17,9,19,20,21,22,23,24,25,26,27};
static const int S_R18[18] = {10,2,12,4,14,6,16,8,18,28,29,30,31,32,33,34,35,36};

/* src/muscl_module.f90:162-290; flags (if non-NULL) receive per variable
 * bit0/1: left/right transverse-ratio switch, bit2/3: left/right sign branch. */
void orc_mlp(const orc_cfg *c, const double (*x)[ORC_NPV], double *xl, double *xr, int *flags)
{
    const double eps = 1.0e-30, eps2 = 1.0e-3;
    mstate m;
    for (int k = 1; k <= c->npv; k++) {
        double dq, dqm, dqp, dqpp, qmin, qmax;
        int fl = 0;
        ratios(x, k, &m, &dq, &dqm, &dqp, &dqpp);
        double rxy_l = fabs((X(11, k) - X(7, k)) / (X(10, k) - X(23, k) + eps));
        double rxz_l = fabs((X(15, k) - X(3, k)) / (X(10, k) - X(23, k) + eps));
        double rxy_r = fabs((X(12, k) - X(8, k)) / (X(32, k) - X(9, k) + eps));
        double rxz_r = fabs((X(16, k) - X(4, k)) / (X(32, k) - X(9, k) + eps));
        if ((rxy_l < eps2) && (rxz_l < eps2)) {
            fl |= 1;
            qmin = minlist(x, k, S_L27, 27);
            qmax = maxlist(x, k, S_L27, 27);
        } else {
            if ((X(10, k) - X(23, k)) > 0.0) {
                fl |= 4;
                qmax = maxlist(x, k, S_18, 18);
                qmin = minlist(x, k, S_L18, 18);
            } else {
                qmin = minlist(x, k, S_18, 18);
                qmax = maxlist(x, k, S_L18, 18);
            }
        }
        double qml = fmin_f(fabs(qmax - X(9, k)), fabs(qmin - X(9, k)));
        if ((rxy_r < eps2) && (rxz_r < eps2)) {
            fl |= 2;
            qmin = minlist(x, k, S_R27, 27);
            qmax = maxlist(x, k, S_R27, 27);
        } else {
            if ((X(32, k) - X(9, k)) > 0.0) {
                fl |= 8;
                qmax = maxlist(x, k, S_R18, 18);
                qmin = minlist(x, k, S_18, 18);
            } else {
                qmin = minlist(x, k, S_R18, 18);
                qmax = maxlist(x, k, S_18, 18);
            }
        }
        double qmr = fmin_f(fabs(qmax - X(10, k)), fabs(qmin - X(10, k)));
        if (fabs(dq) <= eps) {
            m.alp[0] = 2.0 * fmax_f(1.0, m.r[0]) / (fabs(dq) + eps) / (1.0 + rxy_l + rxz_l) * qml;
            m.alp[1] = 2.0 * fmax_f(1.0, dqp / eps) * fabs(dqpp) / (1.0 + rxy_r + rxz_r) * qmr;
        } else {
            m.alp[0] = 2.0 * fmax_f(1.0, m.r[0]) / (fabs(dq) + eps) / (1.0 + rxy_l + rxz_l) * qml;
            m.alp[1] = 2.0 * fmax_f(1.0, 1.0 / m.r[1]) * fabs(dqpp) / (1.0 + rxy_r + rxz_r) * qmr;
        }
        m.alp[0] = fmax_f(1.0, fmin_f(2.0, m.alp[0]));
        m.alp[1] = fmax_f(1.0, fmin_f(2.0, m.alp[1]));
        xl[k - 1] = X(9, k) + 0.5 * limiter(c->nlim, &m, 0) * dqm;
        xr[k - 1] = X(10, k) - 0.5 * limiter(c->nlim, &m, 1) * dqp;
        if (flags) flags[k - 1] = fl;
    }
    clamps(x, c->npv, c->pref, xl, xr);
}

/* ------------------------------------------------------------------ */
/* inviscid fluxes                                                     */
static double enthalpy(const double *pv, const double *dv)
{
    return dv[1] + 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
}

/* src/flux_module.f90:198-276 */
void orc_roe(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
             const double *dvl, const double *dvr, double *fx)
{
    int npv = c->npv;
    double dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == 0.0) { for (int k = 0; k < npv; k++) fx[k] = 0.0; return; }
    double nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    double uull = nx * pvl[1] + ny * pvl[2] + nz * pvl[3];
    double uurr = nx * pvr[1] + ny * pvr[2] + nz * pvr[3];
    double ravg[ORC_NPV], rdv[18], df[ORC_NPV];
    double ravg_d = 1.0 / (sqrt(dvl[0]) + sqrt(dvr[0]));
    ravg[0] = 0.5 * (pvl[0] + pvr[0]);
    ravg[4] = 0.5 * (pvl[4] + pvr[4]);
    for (int k = 1; k <= 3; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    for (int k = 5; k < npv; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    double hl = enthalpy(pvl, dvl), hr = enthalpy(pvr, dvr);
    double ravg_ht = (sqrt(dvl[0]) * hl + sqrt(dvr[0]) * hr) * ravg_d;
    orc_deteos_simple(ravg[0] + c->pref, ravg[4], ravg[5], ravg[6], rdv);
    double u2 = ravg[1] * ravg[1] + ravg[2] * ravg[2] + ravg[3] * ravg[3];
    rdv[0] = sqrt(dvl[0] * dvr[0]);
    double uuu = nx * ravg[1] + ny * ravg[2] + nz * ravg[3];
    double sndp2 = orc_sndp2(c, rdv[5], u2);
    double q = sndp2 / rdv[5];
    double uup = 0.5 * (1.0 + q) * uuu;
    double omq = 1.0 - q;
    double ddd = 0.5 * sqrt(omq * omq * (uuu * uuu) + 4.0 * sndp2);
    double c_star = 0.5 * (fabs(uup + ddd) + fabs(uup - ddd));
    double m_star = 0.5 * (fabs(uup + ddd) - fabs(uup - ddd)) / ddd;
    double du = m_star * (uurr - uull)
              + (c_star - sndp2 / rdv[5] * fabs(uuu) - 0.5 * (1.0 - sndp2 / rdv[5]) * uuu * m_star)
                * (pvr[0] - pvl[0]) / rdv[0] / sndp2;
    double dp = m_star * (pvr[0] - pvl[0])
              + (c_star - fabs(uuu) + 0.5 * (1.0 - sndp2 / rdv[5]) * uuu * m_star) * rdv[0] * (uurr - uull);
    double au = fabs(uuu);
    df[0] = au * (dvr[0] - dvl[0]) + du * rdv[0];
    df[1] = au * (dvr[0] * pvr[1] - dvl[0] * pvl[1]) + du * rdv[0] * ravg[1] + dp * nx;
    df[2] = au * (dvr[0] * pvr[2] - dvl[0] * pvl[2]) + du * rdv[0] * ravg[2] + dp * ny;
    df[3] = au * (dvr[0] * pvr[3] - dvl[0] * pvl[3]) + du * rdv[0] * ravg[3] + dp * nz;
    df[4] = au * (dvr[0] * hr - pvr[0] - dvl[0] * hl + pvl[0]) + du * rdv[0] * ravg_ht + dp * uuu;
    for (int k = 5; k < npv; k++)
        df[k] = au * (dvr[0] * pvr[k] - dvl[0] * pvl[k]) + du * rdv[0] * ravg[k];
    double pl = pvl[0] + c->pref, pr = pvr[0] + c->pref;
    fx[0] = 0.5 * (dvl[0] * uull + dvr[0] * uurr - df[0]) * dl;
    fx[1] = 0.5 * (dvl[0] * uull * pvl[1] + nx * pl + dvr[0] * uurr * pvr[1] + nx * pr - df[1]) * dl;
    fx[2] = 0.5 * (dvl[0] * uull * pvl[2] + ny * pl + dvr[0] * uurr * pvr[2] + ny * pr - df[2]) * dl;
    fx[3] = 0.5 * (dvl[0] * uull * pvl[3] + nz * pl + dvr[0] * uurr * pvr[3] + nz * pr - df[3]) * dl;
    fx[4] = 0.5 * (dvl[0] * uull * hl + dvr[0] * uurr * hr - df[4]) * dl;
    for (int k = 5; k < npv; k++)
        fx[k] = 0.5 * (dvl[0] * uull * pvl[k] + dvr[0] * uurr * pvr[k] - df[k]) * dl;
}

/* getsndp2_du: no preconditioning, or the steady form for prec 1 and 2
   (src/flux_module.f90:107-116) */
static double sndp2_du(const orc_cfg *c, double snd2, double uuu2)
{
    if (c->prec == 0) return snd2;
    return fmin_f(snd2, fmax_f(uuu2, c->uref * c->uref));
}

/* the lifted stencil pressures and their two ratio minima of Roe-M and
   AUSMPW+ (src/flux_module.f90:375-386, 499-508); sd = x(1:18,1) */
static void sensor(const orc_cfg *c, const double *dvl, const double *dvr, const double *sd, double *m910, double *mall)
{
    double s[19];
    for (int k = 1; k <= 18; k++) s[k] = sd[k - 1] + c->pref + 0.1 * fmin_f(dvl[0] * dvl[5], dvr[0] * dvr[5]);
    double r[18] = {s[9] / s[10], s[10] / s[9], s[11] / s[9], s[9] / s[11], s[9] / s[7], s[7] / s[9],
                    s[9] / s[3], s[3] / s[9], s[9] / s[15], s[15] / s[9], s[10] / s[12], s[12] / s[10],
                    s[10] / s[8], s[8] / s[10], s[4] / s[10], s[10] / s[4], s[10] / s[16], s[16] / s[10]};
    *m910 = fmin_f(r[0], r[1]);
    double m = r[0];
    for (int k = 1; k < 18; k++) m = fmin_f(m, r[k]);
    *mall = m;
}

/* Roe-M, src/flux_module.f90:278-418 */
void orc_roem(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
              const double *dvl, const double *dvr, const double *sd, double *fx)
{
    int npv = c->npv;
    double dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == 0.0) { for (int k = 0; k < npv; k++) fx[k] = 0.0; return; }
    double nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    double uull = nx * pvl[1] + ny * pvl[2] + nz * pvl[3];
    double uurr = nx * pvr[1] + ny * pvr[2] + nz * pvr[3];
    double ravg[ORC_NPV], rdv[18];
    double ravg_d = 1.0 / (sqrt(dvl[0]) + sqrt(dvr[0]));
    ravg[0] = 0.5 * (pvl[0] + pvr[0]);
    ravg[4] = 0.5 * (pvl[4] + pvr[4]);
    for (int k = 1; k <= 3; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    for (int k = 5; k < npv; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    double hl = enthalpy(pvl, dvl), hr = enthalpy(pvr, dvr);
    double ravg_ht = (sqrt(dvl[0]) * hl + sqrt(dvr[0]) * hr) * ravg_d;
    orc_deteos_simple(ravg[0] + c->pref, ravg[4], ravg[5], ravg[6], rdv);
    double u2 = ravg[1] * ravg[1] + ravg[2] * ravg[2] + ravg[3] * ravg[3];
    rdv[0] = sqrt(dvl[0] * dvr[0]);
    double uuu = nx * ravg[1] + ny * ravg[2] + nz * ravg[3];
    double am2rmid_du = sndp2_du(c, rdv[5], u2) / rdv[5];
    double am2rmid = orc_sndp2(c, rdv[5], u2) / rdv[5];
    double fmid = sqrt(am2rmid) * (2.0 - sqrt(am2rmid));
    double fmid_du = sqrt(am2rmid_du) * (2.0 - sqrt(am2rmid_du));
    double amid = sqrt(rdv[5]);
    double b1 = fmax_f(fmax_f(0.0, uuu + amid), uurr + amid);
    double b2 = fmin_f(fmin_f(0.0, uuu - amid), uull - amid);
    double b1_du = fmax_f(fmax_f(0.0, uuu + fmid_du * amid), uurr + fmid_du * amid);
    double b2_du = fmin_f(fmin_f(0.0, uuu - fmid_du * amid), uull - fmid_du * amid);
    double m_tilde = copysign(1.0, uuu) * fmin_f(fabs(uuu) / amid, 1.0);
    double m_tilde_du = copysign(1.0, uuu) * fmin_f(fabs(uuu) / (fmid_du * amid), 1.0);
    double dqp[ORC_NPV], dq[ORC_NPV], fl[ORC_NPV], fr[ORC_NPV], bdq[ORC_NPV];
    for (int k = 0; k < npv; k++) dqp[k] = pvr[k] - pvl[k];
    dq[0] = dvr[0] - dvl[0];
    for (int k = 1; k <= 3; k++) dq[k] = dvr[0] * pvr[k] - dvl[0] * pvl[k];
    dq[4] = dvr[0] * hr - dvl[0] * hl;
    for (int k = 5; k < npv; k++) dq[k] = dvr[0] * pvr[k] - dvl[0] * pvl[k];
    const double nn[3] = {nx, ny, nz};
    fl[0] = dvl[0] * uull;
    fr[0] = dvr[0] * uurr;
    for (int k = 1; k <= 3; k++) {
        fl[k] = dvl[0] * uull * pvl[k] + nn[k - 1] * (pvl[0] + c->pref);
        fr[k] = dvr[0] * uurr * pvr[k] + nn[k - 1] * (pvr[0] + c->pref);
    }
    fl[4] = dvl[0] * uull * hl;
    fr[4] = dvr[0] * uurr * hr;
    for (int k = 5; k < npv; k++) { fl[k] = dvl[0] * uull * pvl[k]; fr[k] = dvr[0] * uurr * pvr[k]; }
    double m910, mall;
    sensor(c, dvl, dvr, sd, &m910, &mall);
    double ff = 1.0 - mall, gg = 1.0 - m910;
    ff = (uuu != 0.0) ? pow(fabs(uuu) / amid, ff) : 1.0;
    gg = (uuu != 0.0) ? pow(fabs(uuu) / amid, gg) : 1.0;
    double add = 1.0 + fabs(m_tilde), aaa = 1.0 + fabs(m_tilde_du);
    bdq[0] = dq[0] - ff * dqp[0] / rdv[5] / fmid;
    for (int k = 1; k <= 3; k++) bdq[k] = bdq[0] * ravg[k] + rdv[0] * dqp[k];
    bdq[4] = bdq[0] * ravg_ht + rdv[0] * (hr - hl);
    for (int k = 5; k < npv; k++) bdq[k] = bdq[0] * ravg[k] + rdv[0] * dqp[k];
    for (int k = 0; k < npv; k++)
        fx[k] = (b1 * fl[k] - b2 * fr[k] + b1 * b2 * (dq[k] - gg * bdq[k] / add)) / (b1 - b2) * dl;
    for (int k = 1; k <= 3; k++)
        fx[k] = fx[k] + gg * b1_du * b2_du * rdv[0] * (uurr - uull) * nn[k - 1] / aaa / (b1_du - b2_du) * dl;
}

/* AUSMPW+, src/flux_module.f90:420-546 */
void orc_ausmpwp(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
                 const double *dvl, const double *dvr, const double *sd, double *fx)
{
    const double beta = 0.125, ku = 1.0;
    int npv = c->npv;
    double dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == 0.0) { for (int k = 0; k < npv; k++) fx[k] = 0.0; return; }
    double nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    double uull = nx * pvl[1] + ny * pvl[2] + nz * pvl[3];
    double uurr = nx * pvr[1] + ny * pvr[2] + nz * pvr[3];
    double ravg[ORC_NPV], rdv[18];
    double ravg_d = 1.0 / (sqrt(dvl[0]) + sqrt(dvr[0]));
    ravg[0] = 0.5 * (pvl[0] + pvr[0]);
    ravg[4] = 0.5 * (pvl[4] + pvr[4]);
    for (int k = 1; k <= 3; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    for (int k = 5; k < npv; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    orc_deteos_simple(ravg[0] + c->pref, ravg[4], ravg[5], ravg[6], rdv);
    rdv[0] = sqrt(dvl[0] * dvr[0]);
    double amid = sqrt(rdv[5]);
    double zmr = uurr / amid, zml = uull / amid;
    double am2mid = ravg[1] * ravg[1] + ravg[2] * ravg[2] + ravg[3] * ravg[3];
    double am2rmid1 = sndp2_du(c, rdv[5], am2mid) / rdv[5];
    double am2rmid = orc_sndp2(c, rdv[5], am2mid) / rdv[5];
    double fmid = sqrt(am2rmid) * (2.0 - sqrt(am2rmid));
    double fmid1 = sqrt(am2rmid1) * (2.0 - sqrt(am2rmid1));
    double alpha = 3.0 / 16.0 * (-4.0 + 5.0 * (fmid1 * fmid1));
    alpha = fmax_f(fmin_f(alpha, 3.0 / 16.0), -3.0 / 4.0);
    double zmmr, pmr, zmpl, ppl;
    if (fabs(zmr) >= 1.0) {
        zmmr = 0.5 * (zmr - fabs(zmr));
        pmr = 0.5 * (1.0 - zmr / fabs(zmr));
    } else {
        zmmr = -(0.25 * ((zmr - 1.0) * (zmr - 1.0))) - beta * ((zmr * zmr - 1.0) * (zmr * zmr - 1.0));
        pmr = 0.25 * ((zmr - 1.0) * (zmr - 1.0)) * (2.0 + zmr) - alpha * zmr * ((zmr * zmr - 1.0) * (zmr * zmr - 1.0));
    }
    if (fabs(zml) >= 1.0) {
        zmpl = 0.5 * (zml + fabs(zml));
        ppl = 0.5 * (1.0 + zml / fabs(zml));
    } else {
        zmpl = 0.25 * ((zml + 1.0) * (zml + 1.0)) + beta * ((zml * zml - 1.0) * (zml * zml - 1.0));
        ppl = 0.25 * ((zml + 1.0) * (zml + 1.0)) * (2.0 - zml) + alpha * zml * ((zml * zml - 1.0) * (zml * zml - 1.0));
    }
    double zmid = zmpl + zmmr;
    double pmid = ppl * (pvl[0] + c->pref) + pmr * (pvr[0] + c->pref) - 2.0 * ku * ppl * pmr * rdv[0] * fmid1 * amid * (uurr - uull);
    double m910, mall;
    sensor(c, dvl, dvr, sd, &m910, &mall);
    double ww1 = 1.0 - (m910 * m910) * m910;
    double ww2 = 1.0 - mall * mall;
    double ww = fmax_f(ww1, ww2);
    double pmt = rdv[0] * rdv[5];
    double pwl = ((pvl[0] + c->pref + rdv[0] * rdv[5]) / pmt - 1.0) * (1.0 - ww) / fmid;
    double pwr = ((pvr[0] + c->pref + rdv[0] * rdv[5]) / pmt - 1.0) * (1.0 - ww) / fmid;
    double zmpl1, zmmr1;
    if (zmid >= 0.0) {
        pwl = pwl * rdv[0] / dvl[0];
        pwr = pwr * rdv[0] / dvl[0];
        zmpl1 = zmpl + zmmr * ((1.0 - ww) * (1.0 + pwr) - pwl);
        zmmr1 = zmmr * ww * (1.0 + pwr);
    } else {
        pwl = pwl * rdv[0] / dvr[0];
        pwr = pwr * rdv[0] / dvr[0];
        zmpl1 = zmpl * ww * (1.0 + pwl);
        zmmr1 = zmmr + zmpl * ((1.0 - ww) * (1.0 + pwl) - pwr);
    }
    double hl = enthalpy(pvl, dvl), hr = enthalpy(pvr, dvr);
    fx[0] = (zmpl1 * amid * dvl[0] + zmmr1 * amid * dvr[0]) * dl;
    fx[1] = (zmpl1 * amid * dvl[0] * pvl[1] + zmmr1 * amid * dvr[0] * pvr[1] + pmid * nx) * dl;
    fx[2] = (zmpl1 * amid * dvl[0] * pvl[2] + zmmr1 * amid * dvr[0] * pvr[2] + pmid * ny) * dl;
    fx[3] = (zmpl1 * amid * dvl[0] * pvl[3] + zmmr1 * amid * dvr[0] * pvr[3] + pmid * nz) * dl;
    fx[4] = (zmpl1 * amid * dvl[0] * hl + zmmr1 * amid * dvr[0] * hr) * dl;
    for (int k = 5; k < npv; k++) fx[k] = (zmpl1 * amid * dvl[0] * pvl[k] + zmmr1 * amid * dvr[0] * pvr[k]) * dl;
}

/* src/flux_module.f90:548-641; *flag (if non-NULL) = 1 when zmid > 0 */
void orc_ausmpup(const orc_cfg *c, const double *n3, const double *pvl, const double *pvr,
                 const double *dvl, const double *dvr, double *fx, int *flag)
{
    const double beta = 0.125, kp = 0.25, ku = 1.0;
    int npv = c->npv;
    double dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == 0.0) { for (int k = 0; k < npv; k++) fx[k] = 0.0; return; }
    double nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    double uull = nx * pvl[1] + ny * pvl[2] + nz * pvl[3];
    double uurr = nx * pvr[1] + ny * pvr[2] + nz * pvr[3];
    double ravg[ORC_NPV], rdv[18];
    double ravg_d = 1.0 / (sqrt(dvl[0]) + sqrt(dvr[0]));
    ravg[0] = 0.5 * (pvl[0] + pvr[0]);
    ravg[4] = 0.5 * (pvl[4] + pvr[4]);
    for (int k = 1; k <= 3; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    for (int k = 5; k < npv; k++) ravg[k] = (sqrt(dvl[0]) * pvl[k] + sqrt(dvr[0]) * pvr[k]) * ravg_d;
    orc_deteos_simple(ravg[0] + c->pref, ravg[4], ravg[5], ravg[6], rdv);
    rdv[0] = sqrt(dvl[0] * dvr[0]);
    double amid = sqrt(rdv[5]);
    double zmr = uurr / amid, zml = uull / amid;
    double un = nx * ravg[1] + ny * ravg[2] + nz * ravg[3];
    double am2mid = un * un;
    double am2rmid = orc_sndp2(c, rdv[5], am2mid) / rdv[5];
    double fmid = sqrt(am2rmid) * (2.0 - sqrt(am2rmid));
    double alpha = 0.1875 * (-4.0 + 5.0 * (fmid * fmid));
    alpha = fmax_f(fmin_f(alpha, 0.1875), -0.75);
    double zmmr, pmr, zmpl, ppl;
    if (fabs(zmr) >= 1.0) {
        zmmr = 0.5 * (zmr - fabs(zmr));
        pmr = 0.5 * (1.0 - zmr / fabs(zmr));
    } else {
        double t1 = (zmr - 1.0) * (zmr - 1.0), t2 = (zmr * zmr - 1.0) * (zmr * zmr - 1.0);
        zmmr = -(0.25 * t1) - beta * t2;
        pmr = 0.25 * t1 * (2.0 + zmr) - alpha * zmr * t2;
    }
    if (fabs(zml) >= 1.0) {
        zmpl = 0.5 * (zml + fabs(zml));
        ppl = 0.5 * (1.0 + zml / fabs(zml));
    } else {
        double t1 = (zml + 1.0) * (zml + 1.0), t2 = (zml * zml - 1.0) * (zml * zml - 1.0);
        zmpl = 0.25 * t1 + beta * t2;
        ppl = 0.25 * t1 * (2.0 - zml) + alpha * zml * t2;
    }
    double zmid = zmpl + zmmr - kp * fmax_f(1.0 - am2mid / rdv[5], 0.0) * (pvr[0] - pvl[0]) / rdv[0] / rdv[5] / fmid;
    double pmid = ppl * (pvl[0] + c->pref) + pmr * (pvr[0] + c->pref)
                - 2.0 * ku * ppl * pmr * rdv[0] * fmid * amid * (uurr - uull);
    double zmpl1, zmmr1;
    if (zmid > 0.0) { zmpl1 = zmid; zmmr1 = 0.0; if (flag) *flag = 1; }
    else { zmmr1 = zmid; zmpl1 = 0.0; if (flag) *flag = 0; }
    double hl = enthalpy(pvl, dvl), hr = enthalpy(pvr, dvr);
    fx[0] = (zmpl1 * amid * dvl[0] + zmmr1 * amid * dvr[0]) * dl;
    fx[1] = (zmpl1 * amid * dvl[0] * pvl[1] + zmmr1 * amid * dvr[0] * pvr[1] + pmid * nx) * dl;
    fx[2] = (zmpl1 * amid * dvl[0] * pvl[2] + zmmr1 * amid * dvr[0] * pvr[2] + pmid * ny) * dl;
    fx[3] = (zmpl1 * amid * dvl[0] * pvl[3] + zmmr1 * amid * dvr[0] * pvr[3] + pmid * nz) * dl;
    fx[4] = (zmpl1 * amid * dvl[0] * hl + zmmr1 * amid * dvr[0] * hr) * dl;
    for (int k = 5; k < npv; k++)
        fx[k] = (zmpl1 * amid * dvl[0] * pvl[k] + zmmr1 * amid * dvr[0] * pvr[k]) * dl;
}

/* ------------------------------------------------------------------ */
/* viscous flux, src/vsflux_module.f90:182-425                         */
#define TD(n) (0.0625 * (-X(1, n) - X(2, n) + X(13, n) + X(14, n) \
                + 2.0 * (-X(3, n) - X(4, n) + X(15, n) + X(16, n)) \
                - X(5, n) - X(6, n) + X(17, n) + X(18, n)))
#define ED(n) (0.0625 * (-X(1, n) - X(2, n) + X(5, n) + X(6, n) \
                + 2.0 * (-X(7, n) - X(8, n) + X(11, n) + X(12, n)) \
                - X(13, n) - X(14, n) + X(17, n) + X(18, n)))

void orc_vsflux(const orc_cfg *c, const double *nx, const double *const e[4], const double *const t[4],
                const double *grdl, const double *grdr, const double (*x)[ORC_NPV],
                const double *dvl, const double *dvr, const double *tvl, const double *tvr, double *fx)
{
    const double c23 = 2.0 / 3.0;
    int turb = (c->iturb == 0);
    double ex = 0.25 * (e[0][0] + e[1][0] + e[2][0] + e[3][0]);
    double ey = 0.25 * (e[0][1] + e[1][1] + e[2][1] + e[3][1]);
    double ez = 0.25 * (e[0][2] + e[1][2] + e[2][2] + e[3][2]);
    double tx = 0.25 * (t[0][0] + t[1][0] + t[2][0] + t[3][0]);
    double ty = 0.25 * (t[0][1] + t[1][1] + t[2][1] + t[3][1]);
    double tz = 0.25 * (t[0][2] + t[1][2] + t[2][2] + t[3][2]);
    double vol = 2.0 / (grdl[0] + grdr[0]);
    double vis = 0.5 * (tvl[0] + tvr[0]);
    double con = 0.5 * (tvl[1] + tvr[1]);
    double emut = 0.0, tcon = 0.0;
    if (turb) {
        emut = 0.5 * (tvl[2] + tvr[2]);
        tcon = 0.5 * (tvl[2] * dvl[11] + tvr[2] * dvr[11]) / 0.9;
    }
    double uc = 0.5 * (X(9, 2) + X(10, 2));
    double vc = 0.5 * (X(9, 3) + X(10, 3));
    double wc = 0.5 * (X(9, 4) + X(10, 4));
    double dudn = X(10, 2) - X(9, 2), dvdn = X(10, 3) - X(9, 3), dwdn = X(10, 4) - X(9, 4), dtdn = X(10, 5) - X(9, 5);
    double dude = ED(2), dvde = ED(3), dwde = ED(4), dtde = ED(5);
    double dudt = TD(2), dvdt = TD(3), dwdt = TD(4), dtdt = TD(5);
    double dudx = (dudn * nx[0] + dude * ex + dudt * tx) * vol;
    double dvdx = (dvdn * nx[0] + dvde * ex + dvdt * tx) * vol;
    double dwdx = (dwdn * nx[0] + dwde * ex + dwdt * tx) * vol;
    double dtdx = (dtdn * nx[0] + dtde * ex + dtdt * tx) * vol;
    double dudy = (dudn * nx[1] + dude * ey + dudt * ty) * vol;
    double dvdy = (dvdn * nx[1] + dvde * ey + dvdt * ty) * vol;
    double dwdy = (dwdn * nx[1] + dwde * ey + dwdt * ty) * vol;
    double dtdy = (dtdn * nx[1] + dtde * ey + dtdt * ty) * vol;
    double dudz = (dudn * nx[2] + dude * ez + dudt * tz) * vol;
    double dvdz = (dvdn * nx[2] + dvde * ez + dvdt * tz) * vol;
    double dwdz = (dwdn * nx[2] + dwde * ez + dwdt * tz) * vol;
    double dtdz = (dtdn * nx[2] + dtde * ez + dtdt * tz) * vol;
    double mu = turb ? (vis + emut) : vis;
    double kap = turb ? (con + tcon) : con;
    double div = dudx + dvdy + dwdz;
    double tauxx = mu * (2.0 * dudx - c23 * div);
    double tauyy = mu * (2.0 * dvdy - c23 * div);
    double tauzz = mu * (2.0 * dwdz - c23 * div);
    double tauxy = mu * (dudy + dvdx);
    double tauxz = mu * (dudz + dwdx);
    double tauyz = mu * (dvdz + dwdy);
    fx[0] = 0.0;
    fx[1] = tauxx * nx[0] + tauxy * nx[1] + tauxz * nx[2];
    fx[2] = tauxy * nx[0] + tauyy * nx[1] + tauyz * nx[2];
    fx[3] = tauxz * nx[0] + tauyz * nx[1] + tauzz * nx[2];
    fx[4] = nx[0] * (uc * tauxx + vc * tauxy + wc * tauxz + kap * dtdx)
          + nx[1] * (uc * tauxy + vc * tauyy + wc * tauyz + kap * dtdy)
          + nx[2] * (uc * tauxz + vc * tauyz + wc * tauzz + kap * dtdz);
    fx[5] = 0.0;
    fx[6] = 0.0;
    if (turb) {
        double dkdn = X(10, 8) - X(9, 8), dodn = X(10, 9) - X(9, 9);
        double dkde = ED(8), dode = ED(9), dkdt = TD(8), dodt = TD(9);
        double dkdx = (dkdn * nx[0] + dkde * ex + dkdt * tx) * vol;
        double dodx = (dodn * nx[0] + dode * ex + dodt * tx) * vol;
        double dkdy = (dkdn * nx[1] + dkde * ey + dkdt * ty) * vol;
        double dody = (dodn * nx[1] + dode * ey + dodt * ty) * vol;
        double dkdz = (dkdn * nx[2] + dkde * ez + dkdt * tz) * vol;
        double dodz = (dodn * nx[2] + dode * ez + dodt * tz) * vol;
        /* SST F1 with face-averaged quantities, src/vsflux_module.f90:400-425 */
        double pkw = dkdx * dodx + dkdy * dody + dkdz * dodz;
        double rho = 0.5 * (dvl[0] + dvr[0]);
        double ydns = 0.5 * (grdl[4] + grdr[4]);
        double visf = 0.5 * (tvl[0] + tvr[0]);
        double kk = 0.5 * (X(9, 8) + X(10, 8));
        double oo = 0.5 * (X(9, 9) + X(10, 9));
        double term0 = 1.712 * rho / oo * pkw;
        double cdkw = fmax_f(term0, 1.0e-10);
        double term1 = sqrt(kk) / (0.09 * oo * ydns);
        double term2 = 500.0 * visf / rho / (oo * (ydns * ydns));
        double term3 = (3.424 * rho * kk) / (cdkw * (ydns * ydns));
        double arg2 = fmax_f(term1, term2);
        double arg1 = fmin_f(arg2, term3);
        double a2 = arg1 * arg1;
        double bigf = tanh(a2 * a2);
        double tdk = bigf * 0.85 + (1.0 - bigf) * 1.0;
        double tdo = bigf * 0.5 + (1.0 - bigf) * 0.856;
        fx[7] = (vis + emut * tdk) * (dkdx * nx[0] + dkdy * nx[1] + dkdz * nx[2]);
        fx[8] = (vis + emut * tdo) * (dodx * nx[0] + dody * nx[1] + dodz * nx[2]);
    }
}
#undef TD
#undef ED

/* ------------------------------------------------------------------ */
/* The other mass-transfer models, src/cav_module.f90:173-357.  Common:
 * nothing at or above the critical temperature; the rate clip as the
 * Merkle model (as written on the temperature, or on the mass fractions);
 * the coupling vector flipped by the sign of its y1 entry.  Constant integer
 * powers are the products gfortran expands them into (x**2 = x*x,
 * x**3 = (x*x)*x, x**4 = (x*x)*(x*x), x**5 = ((x*x)*x)*(x*x)); real powers
 * are run-time pow. */
static void cav_clip(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, double *r_c, double *r_v)
{
    if (c->cav_clip == 0) {
        *r_c = fmin_f(*r_c, dv[0] * X(2, 5));
        *r_v = fmin_f(*r_v, dv[0] * (1.0 - X(2, 5) - X(2, 6)));
    } else {
        *r_c = fmin_f(*r_c, dv[0] * X(2, 6));
        *r_v = fmin_f(*r_v, dv[0] * (1.0 - X(2, 6) - X(2, 7)));
    }
}

static void cav_sign(double *icav)
{
    double lam = copysign(1.0, icav[2]);
    double s = 1.0 * (1.0 - 0.5 * (1.0 - lam)) - 0.5 * (1.0 - lam);
    for (int q = 0; q < 4; q++) icav[q] = icav[q] * s;
}

/* src/eos_module.f90:2430-2438 — surface tension of water against its vapour */
double orc_h2o_sigma(double t)
{
    double t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t3 * t2;
    return 0.0881130429998226 + 0.0000616475914044782 * t - 4.4860167748009e-7 * t2 + 2.53682468718936e-10 * t3
           - 2.62005668908913e-13 * t4 + 2.97878955533757e-16 * t5;
}

/* Kunz, src/cav_module.f90:173-211 */
void orc_kunz(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
              double *r_c_out, double *r_v_out, double *icav)
{
    *r_c_out = 0.0; *r_v_out = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (X(2, 5) >= c->t_crit) return;
    double pww = orc_h2o_pww(X(2, 5));
    double rho1 = 1.0 / dv[0];
    double av = dv[0] * X(2, 6) / dv[3];
    double ag = dv[0] * X(2, 7) / dv[4];
    double al = 1.0 - av - ag;
    double r_c = c->c_c * dv[3] * av * ((al + ag) * (al + ag)) * c->t_ref;
    double r_v = c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
    cav_clip(c, x, dv, &r_c, &r_v);
    *r_c_out = r_c * grd[0];
    *r_v_out = r_v * grd[0];
    double av1 = 1.0 - 4.0 * av + 3.0 * (av * av);
    icav[0] = -(c->c_c * c->t_ref * (dv[6] * X(2, 6) * av1 + 2.0 * (av * av) * dv[14] * (al + ag)));
    icav[1] = -(c->c_c * c->t_ref * (dv[7] * X(2, 6) * av1 + 2.0 * (av * av) * dv[15] * (al + ag)));
    icav[2] = -(c->c_c * c->t_ref * (dv[8] * X(2, 6) + dv[0]) * av1);
    icav[3] = -(c->c_c * c->t_ref * dv[9] * X(2, 6) * av1);
    if (r_v != 0.0) {
        icav[0] = r_v * dv[6] * rho1 - c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) * c->dp_ref * c->t_ref;
        icav[1] = r_v * dv[7] * rho1;
        icav[2] = r_v * dv[8] * rho1 - c->c_v * dv[0] * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
        icav[3] = r_v * dv[9] * rho1 - c->c_v * dv[0] * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
    }
    cav_sign(icav);
}

/* Singhal, src/cav_module.f90:213-260 */
void orc_singhal(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                 double *r_c_out, double *r_v_out, double *icav)
{
    *r_c_out = 0.0; *r_v_out = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (X(2, 5) >= c->t_crit) return;
    double pww = orc_h2o_pww(X(2, 5));
    double sigma = orc_h2o_sigma(X(2, 5));
    double rho1 = 1.0 / dv[2], rho2 = 1.0 / dv[3];
    double kc = c->c_c * c->uref / sigma * dv[2] * dv[3], kv = c->c_v * c->uref / sigma * dv[2] * dv[3];
    double r_c = kc * sqrt(2.0 / 3.0 * fmax_f(X(2, 1) + c->pref - pww, 0.0) * rho1) * X(2, 6);
    double r_v = kv * sqrt(2.0 / 3.0 * fmax_f(pww - X(2, 1) - c->pref, 0.0) * rho1) * (1.0 - X(2, 6) - X(2, 7));
    cav_clip(c, x, dv, &r_c, &r_v);
    *r_c_out = r_c * grd[0];
    *r_v_out = r_v * grd[0];
    if (r_c != 0.0) {
        icav[0] = -(r_c * (dv[14] * rho2 + 0.5 * dv[16] * rho1))
                  - 0.5 * c->c_c * c->uref / sigma * dv[2] * dv[3] * sqrt(2.0 / 3.0 * rho1) * X(2, 6)
                        * pow(X(2, 1) + c->pref - pww, -0.5);
        icav[1] = -(r_c * (dv[15] * rho2 + 0.5 * dv[17] * rho1));
        icav[2] = -(kc * sqrt(2.0 / 3.0 * fmax_f(X(2, 1) + c->pref - pww, 0.0) * rho1));
        icav[3] = 0.0;
    }
    if (r_v != 0.0) {
        icav[0] = r_v * (dv[14] * rho2 + 0.5 * dv[16] * rho1)
                  - 0.5 * c->c_v * c->uref / sigma * dv[2] * dv[3] * sqrt(2.0 / 3.0 * rho1) * (1.0 - X(2, 6) - X(2, 7))
                        * pow(pww - X(2, 1) - c->pref, -0.5);
        icav[1] = r_v * (dv[15] * rho2 + 0.5 * dv[17] * rho1);
        icav[2] = -(kv * sqrt(2.0 / 3.0 * fmax_f(pww - X(2, 1) - c->pref, 0.0) * rho1));
        icav[3] = -(kv * sqrt(2.0 / 3.0 * fmax_f(pww - X(2, 1) - c->pref, 0.0) * rho1));
    }
    cav_sign(icav);
}

/* Schnerr-Sauer, src/cav_module.f90:262-316 (bubble number density 1e13) */
void orc_schnerr_sauer(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                       double *r_c_out, double *r_v_out, double *icav)
{
    *r_c_out = 0.0; *r_v_out = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (X(2, 5) >= c->t_crit) return;
    const double pi = 4.0 * atan(1.0), eta = 1.0e13;
    double pww = orc_h2o_pww(X(2, 5));
    double rho1 = 1.0 / dv[0], rho3 = 1.0 / dv[2], rho4 = 1.0 / dv[3];
    double y1 = X(2, 6), y2 = X(2, 7), yl = 1.0 - X(2, 6) - X(2, 7);
    double kb = pow(4.0 * pi * eta / 3.0, 1.0 / 3.0), k23 = sqrt(2.0 / 3.0);
    double r_c = c->c_c * 3.0 * kb * k23 * dv[0] * pow(dv[3], 1.0 / 3.0) * pow(dv[2], -5.0 / 6.0)
                 * pow(y1, 2.0 / 3.0) * pow(yl, 4.0 / 3.0) * sqrt(fmax_f(X(2, 1) + c->pref - pww, 0.0));
    double r_v = c->c_v * 3.0 * kb * k23 * dv[0] * pow(dv[3], 1.0 / 3.0) * pow(dv[2], -5.0 / 6.0)
                 * pow(y1 + y2, 2.0 / 3.0) * pow(yl, 4.0 / 3.0) * sqrt(fmax_f(pww - X(2, 1) - c->pref, 0.0));
    cav_clip(c, x, dv, &r_c, &r_v);
    *r_c_out = r_c * grd[0];
    *r_v_out = r_v * grd[0];
    if (r_c != 0.0) {
        icav[0] = -(r_c * (dv[6] * rho1 + dv[14] * rho4 / 3.0 - 5.0 * dv[16] * rho3 / 6.0 + 0.5 / (X(2, 1) + c->pref - pww)));
        icav[1] = -(r_c * (dv[7] * rho1 + dv[15] * rho4 / 3.0 - 5.0 * dv[17] * rho3 / 6.0));
        icav[2] = -(r_c * (dv[8] * rho1 + 2.0 / (3.0 * y1) - 4.0 / (3.0 * yl)));
        icav[3] = -(r_c * (dv[9] * rho1 - 4.0 / (3.0 * yl)));
    }
    if (r_v != 0.0) {
        icav[0] = r_v * (dv[6] * rho1 + dv[14] * rho4 / 3.0 - 5.0 * dv[16] * rho3 / 6.0 - 0.5 / (pww - X(2, 1) - c->pref));
        icav[1] = r_v * (dv[7] * rho1 + dv[15] * rho4 / 3.0 - 5.0 * dv[17] * rho3 / 6.0);
        icav[2] = r_v * (dv[8] * rho1 + 2.0 / (3.0 * (y1 + y2)) - 4.0 / (3.0 * yl));
        icav[3] = r_v * (dv[9] * rho1 + 2.0 / (3.0 * (y1 + y2)) - 4.0 / (3.0 * yl));
    }
    cav_sign(icav);
}

/* Lee, src/cav_module.f90:318-356 */
void orc_lee(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
             double *r_c_out, double *r_v_out, double *icav)
{
    *r_c_out = 0.0; *r_v_out = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (X(2, 5) >= c->t_crit) return;
    double tww = orc_h2o_tww(X(2, 1) + c->pref);
    double rho1 = 1.0 / dv[0];
    double r_c = c->c_c * dv[0] * X(2, 6) * fmax_f(tww - X(2, 5), 0.0) / tww;
    double r_v = c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) * fmax_f(X(2, 5) - tww, 0.0) / tww;
    cav_clip(c, x, dv, &r_c, &r_v);
    *r_c_out = r_c * grd[0];
    *r_v_out = r_v * grd[0];
    if (r_c != 0.0) {
        icav[0] = -(r_c * dv[6] * rho1);
        icav[1] = -(r_c * dv[7] * rho1) + c->c_c * dv[0] * X(2, 6) / tww;
        icav[2] = -(r_c * dv[8] * rho1) - c->c_c * dv[0] * fmax_f(tww - X(2, 5), 0.0) / tww;
        icav[3] = -(r_c * dv[9] * rho1);
    }
    if (r_v != 0.0) {
        icav[0] = r_v * dv[6] * rho1;
        icav[1] = r_v * dv[7] * rho1 + c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) / tww;
        icav[2] = r_v * dv[8] * rho1 - c->c_v * dv[0] * fmax_f(X(2, 5) - tww, 0.0) / tww;
        icav[3] = r_v * dv[9] * rho1 - c->c_v * dv[0] * fmax_f(X(2, 5) - tww, 0.0) / tww;
    }
    cav_sign(icav);
}

/* the model of config ncav (src/rhs_module.f90:98-110) */
void orc_cavsource(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                   double *r_c, double *r_v, double *icav)
{
    switch (c->ncav) {
    case 2: orc_kunz(c, x, dv, grd, r_c, r_v, icav); break;
    case 3: orc_singhal(c, x, dv, grd, r_c, r_v, icav); break;
    case 4: orc_schnerr_sauer(c, x, dv, grd, r_c, r_v, icav); break;
    case 5: orc_lee(c, x, dv, grd, r_c, r_v, icav); break;
    default: orc_merkle(c, x, dv, grd, r_c, r_v, icav); break;
    }
}

/* ------------------------------------------------------------------ */
/* Merkle mass transfer, src/cav_module.f90:136-171                    */
void orc_merkle(const orc_cfg *c, const double (*x)[ORC_NPV], const double *dv, const double *grd,
                double *r_c_out, double *r_v_out, double *icav)
{
    *r_c_out = 0.0; *r_v_out = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (X(2, 5) >= c->t_crit) return;
    double pww = orc_h2o_pww(X(2, 5));
    double rho1 = 1.0 / dv[0];
    double r_c = c->c_c * dv[0] * X(2, 6) * fmax_f(X(2, 1) + c->pref - pww, 0.0) * c->dp_ref * c->t_ref;
    double r_v = c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
    if (c->cav_clip == 0) {   /* as written: index 5 is the temperature */
        r_c = fmin_f(r_c, dv[0] * X(2, 5));
        r_v = fmin_f(r_v, dv[0] * (1.0 - X(2, 5) - X(2, 6)));
    } else {                  /* on the mass fractions */
        r_c = fmin_f(r_c, dv[0] * X(2, 6));
        r_v = fmin_f(r_v, dv[0] * (1.0 - X(2, 6) - X(2, 7)));
    }
    *r_c_out = r_c * grd[0];
    *r_v_out = r_v * grd[0];
    if (r_c != 0.0) {
        icav[0] = -(r_c * dv[6] * rho1) - c->c_c * dv[0] * X(2, 6) * c->dp_ref * c->t_ref;
        icav[1] = -(r_c * dv[7] * rho1);
        icav[2] = -(r_c * dv[8] * rho1) - c->c_c * dv[0] * fmax_f(X(2, 1) + c->pref - pww, 0.0) * c->dp_ref * c->t_ref;
        icav[3] = -(r_c * dv[9] * rho1);
    }
    if (r_v != 0.0) {
        icav[0] = r_v * dv[6] * rho1 - c->c_v * dv[0] * (1.0 - X(2, 6) - X(2, 7)) * c->dp_ref * c->t_ref;
        icav[1] = r_v * dv[7] * rho1;
        icav[2] = r_v * dv[8] * rho1 - c->c_v * dv[0] * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
        icav[3] = r_v * dv[9] * rho1 - c->c_v * dv[0] * fmax_f(pww - X(2, 1) - c->pref, 0.0) * c->dp_ref * c->t_ref;
    }
    double lam = copysign(1.0, icav[2]);
    double s = 1.0 * (1.0 - 0.5 * (1.0 - lam)) - 0.5 * (1.0 - lam);
    for (int q = 0; q < 4; q++) icav[q] = icav[q] * s;
}

/* ------------------------------------------------------------------ */
/* 7-point Green-Gauss gradient of variable n at a cell, shared form of
 * src/turbsource_module.f90:285-339 and src/eddy_module.f90:145-175     */
static void gg(const double (*x)[ORC_NPV], int n, const double *const f[6], double vol, double *g)
{
    double a1 = 0.5 * (X(1, n) + X(2, n));
    double a2 = 0.5 * (X(3, n) + X(2, n));
    double a3 = 0.5 * (X(4, n) + X(2, n));
    double a4 = 0.5 * (X(5, n) + X(2, n));
    double a5 = 0.5 * (X(6, n) + X(2, n));
    double a6 = 0.5 * (X(7, n) + X(2, n));
    for (int m = 0; m < 3; m++)
        g[m] = (a2 * f[1][m] + a4 * f[3][m] + a6 * f[5][m] - a1 * f[0][m] - a3 * f[2][m] - a5 * f[4][m]) * vol;
}

/* calturblength, src/turbsource_module.f90:406-576: no_des, sst_des,
   sst_ddes, sst_iddes, sst_s_iddes (ides 0..4).  f: the cell's face
   vectors cx1, cx2, ex1, ex2, tx1, tx2. */
static double turb_length(int ides, const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                          const double *tv, double bigf, double strain_sq, double vorticity_sq)
{
    double tl_rans = sqrt(X(2, 8)) / 9.0 / X(2, 9) * 100.0;
    if (ides == 0) return tl_rans;
    double gsp[3];
    for (int q = 0; q < 3; q++) {
        const double *a = f[2 * q], *b = f[2 * q + 1];
        gsp[q] = 0.5 * (sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) + sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
    }
    for (int q = 0; q < 3; q++) gsp[q] = grd[0] / gsp[q];
    double hmax = fmax_f(fmax_f(gsp[0], gsp[1]), gsp[2]);
    if (ides == 1) {
        double tl_les = 0.61 * hmax;
        return fmin_f(tl_rans, tl_les);
    }
    double cdes = 0.78 * bigf + 0.61 * (1.0 - bigf);
    if (ides == 2) {
        double tl_les = cdes * hmax;
        double rd_num = tv[0] + tv[2];
        double rd_den = 0.41 * 0.41 * (grd[4] * grd[4]) * sqrt(0.5 * (strain_sq + vorticity_sq));
        double rd = rd_num / rd_den;
        double y = 20.0 * rd;
        double fd = 1.0 - tanh((y * y) * y);
        return tl_rans - fd * fmax_f(0.0, tl_rans - tl_les);
    }
    double tl_les = cdes * fmin_f(0.15 * fmax_f(grd[4], hmax), hmax);
    double alpha = 0.25 - grd[4] / hmax;
    double fb = fmin_f(2.0 * exp(-9.0 * (alpha * alpha)), 1.0);
    double rdt_den = 0.41 * 0.41 * (grd[4] * grd[4]) * sqrt(0.5 * (strain_sq + vorticity_sq));
    double rdt = tv[2] / rdt_den;
    double y = 20.0 * rdt;
    double fdt = 1.0 - tanh((y * y) * y);
    double fd = fmax_f(1.0 - fdt, fb);
    if (ides == 4) return fd * tl_rans + (1.0 - fd) * tl_les;
    double rdl = tv[0] / rdt_den;
    double z = 25.0 * rdl, z2 = z * z, z3 = z2 * z, z5 = z3 * z2;
    double fl = tanh(z5 * z5);
    double w = 3.4969 * rdt;
    double ft = tanh((w * w) * w);
    double fe2 = 1.0 - fmax_f(ft, fl);
    double fe1 = (alpha >= 0.0) ? 2.0 * exp(-11.09 * (alpha * alpha)) : 2.0 * exp(-9.0 * (alpha * alpha));
    double fe = fe2 * fmax_f(fe1 - 1.0, 0.0);
    return fd * (1.0 + fe) * tl_rans + (1.0 - fd) * tl_les;
}

void orc_kwsst_source_des(int ides, const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                          const double *dv, const double *tv, double *source, double *itt, double *omega_cut);

/* src/turbsource_module.f90:269-414 (ides = 0) */
void orc_kwsst_source(const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                      const double *dv, const double *tv, double *source, double *itt, double *omega_cut)
{
    orc_kwsst_source_des(0, f, grd, x, dv, tv, source, itt, omega_cut);
}

/* src/turbsource_module.f90:269-383 with the length scale of ides */
void orc_kwsst_source_des(int ides, const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                          const double *dv, const double *tv, double *source, double *itt, double *omega_cut)
{
    const double c23 = 2.0 / 3.0, c59 = 5.0 / 9.0;
    double vol = 1.0 / grd[0];
    double gk[3], go[3], gu[3], gv[3], gw[3];
    gg(x, 8, f, vol, gk);
    gg(x, 9, f, vol, go);
    gg(x, 2, f, vol, gu);
    gg(x, 3, f, vol, gv);
    gg(x, 4, f, vol, gw);
    double dudx = gu[0], dudy = gu[1], dudz = gu[2];
    double dvdx = gv[0], dvdy = gv[1], dvdz = gv[2];
    double dwdx = gw[0], dwdy = gw[1], dwdz = gw[2];
    double div = dudx + dvdy + dwdz;
    double strain_sq = 2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (dudy + dvdx) * (dudy + dvdx)
                     + (dudz + dwdx) * (dudz + dwdx) + (dvdz + dwdy) * (dvdz + dwdy);
    double k = X(2, 8), om = X(2, 9);
    double prod = tv[2] * (strain_sq - c23 * (div * div)) - c23 * dv[0] * k * div;
    prod = fmin_f(prod, 0.9 * dv[0] * k * om);
    *omega_cut = sqrt(strain_sq - c23 * (div * div));
    double pkw = gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2];
    /* calbigf, src/turbsource_module.f90:385-404 */
    double term0 = 1.712 * dv[0] / om * pkw;
    double cdkw = fmax_f(term0, 1.0e-10);
    double term1 = sqrt(k) / (0.09 * om * grd[4]);
    double term2 = 500.0 * tv[0] / dv[0] / (om * (grd[4] * grd[4]));
    double term3 = (3.424 * dv[0] * k) / (cdkw * (grd[4] * grd[4]));
    double arg1 = fmin_f(fmax_f(term1, term2), term3);
    double a2 = arg1 * arg1;
    double bigf = tanh(a2 * a2);
    double talpha = bigf * c59 + (1.0 - bigf) * 0.44;
    double tbeta = bigf * 0.075 + (1.0 - bigf) * 0.0828;
    double ki = 1.0 / k, oi = 1.0 / om, ei = 1.0 / tv[2];
    double vorticity_sq = (dudy - dvdx) * (dudy - dvdx) + (dudz - dwdx) * (dudz - dwdx) + (dvdz - dwdy) * (dvdz - dwdy);
    double tl = turb_length(ides, f, grd, x, tv, bigf, strain_sq, vorticity_sq);
    double dest = dv[0] * pow(k, 1.5) / tl;
    source[0] = (prod - dest) * grd[0];
    source[1] = (talpha * dv[0] * ei * prod - tbeta * dv[0] * (om * om) + 1.712 * (1.0 - bigf) * dv[0] * oi * pkw) * grd[0];
    double dpdk = prod * ki;
    double dpdo = -((prod + c23 * dv[0] * k * div) * oi);
    itt[0] = dpdk - dest * ki;
    itt[1] = dpdo - dest * oi;
    itt[2] = 0.0;
    itt[3] = talpha * dv[0] * ei * dpdo + talpha * dv[0] * ei * prod * oi - 2.0 * tbeta * dv[0] * om
           - 1.712 * (1.0 - bigf) * dv[0] * (oi * oi) * pkw;
    double t = itt[0] + itt[3];
    double d = itt[0] * itt[3] - itt[1] * itt[2];
    double lam1 = 0.5 * t + sqrt(fmax_f(0.25 * (t * t) - d, 0.0));
    double lam2 = 0.5 * t - sqrt(fmax_f(0.25 * (t * t) - d, 0.0));
    lam1 = copysign(1.0, lam1);
    lam2 = copysign(1.0, lam2);
    double s1 = 1.0 * (1.0 - 0.5 * (1.0 - lam1)) - 0.5 * (1.0 - lam1);
    double s2 = 1.0 * (1.0 - 0.5 * (1.0 - lam2)) - 0.5 * (1.0 - lam2);
    itt[0] *= s1; itt[1] *= s1;
    itt[2] *= s2; itt[3] *= s2;
}

/* src/eddy_module.f90:135-187 */
double orc_kwsst_eddy(const double *const f[6], const double *grd, const double (*x)[ORC_NPV],
                      const double *dv, const double *tv)
{
    double vol = 1.0 / grd[0];
    double gu[3], gv[3], gw[3];
    gg(x, 2, f, vol, gu);
    gg(x, 3, f, vol, gv);
    gg(x, 4, f, vol, gw);
    double k = X(2, 8), om = X(2, 9);
    double term1 = sqrt(k) / (0.09 * om * grd[4]);
    double term2 = 500.0 * tv[0] / dv[0] / (om * (grd[4] * grd[4]));
    double arg2 = fmax_f(2.0 * term1, term2);
    double bigf2 = tanh(arg2 * arg2);
    double sijsij = gu[0] * gu[0] + gv[1] * gv[1] + gw[2] * gw[2]
                  + 0.5 * ((gu[1] + gv[0]) * (gu[1] + gv[0]) + (gu[2] + gw[0]) * (gu[2] + gw[0])
                           + (gv[2] + gw[1]) * (gv[2] + gw[1]));
    double sss = sqrt(2.0 * sijsij);
    double term4 = 0.31 * om;
    double term5 = sss * bigf2;
    double term6 = fmax_f(term4, term5);
    return 0.31 * dv[0] * k / term6;
}

/* ------------------------------------------------------------------ */
/* eigenvis of jacobian/timestep (no metric factor),
 * src/jacobian_module.f90:405-435, src/timestep_module.f90:348-377    */
double orc_eigenvis(const orc_cfg *c, const double *dv, const double *tv)
{
    if (c->iturb == -3) return 0.0;
    if (c->iturb == -2)
        return fmax_f(dv[6] * tv[1] / (dv[7] + dv[11] * dv[6] * dv[0] - dv[10] * dv[7] * dv[0]),
                      4.0 / 3.0 * tv[0] / dv[0]);
    return fmax_f(dv[6] * (tv[1] + dv[11] * tv[2] / 0.9) / (dv[7] + dv[11] * dv[6] * dv[0] - dv[10] * dv[7] * dv[0]),
                  4.0 / 3.0 * (tv[0] + tv[2]) / dv[0]);
}

/* src/jacobian_module.f90:182-403; a is npv x npv, column-major a[(l)*npv + n] = a(n+1,l+1) */
void orc_caljac(const orc_cfg *c, const double *n3, const double *grd, const double *pv,
                const double *dv, const double *tv, int sign, double *a)
{
    int npv = c->npv;
    const double cappa = 1.05;
    double ds = n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2];
    double u = n3[0] * pv[1] + n3[1] * pv[2] + n3[2] * pv[3];
    double sndp2 = orc_sndp2(c, dv[5], pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
    double up = 0.5 * (1.0 + sndp2 / dv[5]) * u;
    double omq = 1.0 - sndp2 / dv[5];
    double d = 0.5 * sqrt((u * u) * (omq * omq) + 4.0 * sndp2 * ds);
    double lamda = (double)sign * cappa * (fabs(up) + d);
    double vis = (double)sign * 2.0 * orc_eigenvis(c, dv, tv) * ds / grd[0];
    double beta = 1.0 / sndp2 - 1.0 / dv[5] + dv[6];
    double a1 = u * dv[6], a2 = n3[0] * dv[0], a3 = n3[1] * dv[0], a4 = n3[2] * dv[0];
    double a5 = u * dv[7], a6 = u * dv[8], a7 = u * dv[9];
    double g1 = lamda * beta, g5 = lamda * dv[7], g6 = lamda * dv[8], g7 = lamda * dv[9];
    double ge1 = vis * dv[6], ge5 = vis * dv[7], ge6 = vis * dv[8], ge7 = vis * dv[9];
    double uuu = u + lamda + vis;
    double h = enthalpy(pv, dv);
#define A(n, l) a[((l) - 1) * npv + (n) - 1]
    for (int q = 0; q < npv * npv; q++) a[q] = 0.0;
    A(1, 1) = a1 + g1 + ge1; A(1, 2) = a2; A(1, 3) = a3; A(1, 4) = a4;
    A(1, 5) = a5 + g5 + ge5; A(1, 6) = a6 + g6 + ge6; A(1, 7) = a7 + g7 + ge7;
    double r1[7];
    for (int l = 1; l <= 7; l++) r1[l - 1] = A(1, l);
    for (int l = 1; l <= 7; l++) {
        A(2, l) = pv[1] * r1[l - 1];
        A(3, l) = pv[2] * r1[l - 1];
        A(4, l) = pv[3] * r1[l - 1];
    }
    A(2, 1) = pv[1] * r1[0] + n3[0];
    A(2, 2) = pv[1] * r1[1] + dv[0] * uuu;
    A(3, 1) = pv[2] * r1[0] + n3[1];
    A(3, 3) = pv[2] * r1[2] + dv[0] * uuu;
    A(4, 1) = pv[3] * r1[0] + n3[2];
    A(4, 4) = pv[3] * r1[3] + dv[0] * uuu;
    A(5, 1) = dv[0] * uuu * dv[10] + h * r1[0] - lamda - vis;
    A(5, 2) = dv[0] * uuu * pv[1] + h * r1[1];
    A(5, 3) = dv[0] * uuu * pv[2] + h * r1[2];
    A(5, 4) = dv[0] * uuu * pv[3] + h * r1[3];
    A(5, 5) = dv[0] * uuu * dv[11] + h * r1[4];
    A(5, 6) = dv[0] * uuu * dv[12] + h * r1[5];
    A(5, 7) = dv[0] * uuu * dv[13] + h * r1[6];
    for (int n = 6; n <= npv; n++) {
        for (int l = 1; l <= 7; l++) A(n, l) = pv[n - 1] * r1[l - 1];
    }
    A(6, 6) = pv[5] * r1[5] + dv[0] * uuu;
    A(7, 7) = pv[6] * r1[6] + dv[0] * uuu;
    if (npv > 7) {
        A(8, 8) = uuu * dv[0];
        A(9, 9) = uuu * dv[0];
    }
    for (int q = 0; q < npv * npv; q++) a[q] = a[q] * 0.5;
#undef A
}

/* ------------------------------------------------------------------ */
/* lhs eigen + eigenvis with metric factor, src/lhs_module.f90:459-534  */
static double lhs_eigen(const orc_cfg *c, const double *const f[6], const double *pv, const double *dv)
{
    const double cappa = 1.05;
    double uv2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
    double sndp2 = orc_sndp2(c, dv[5], uv2);
    double lamda = 0.0;
    for (int d = 0; d < 3; d++) {
        const double *f1 = f[2 * d], *f2 = f[2 * d + 1];
        double ox = 0.5 * (f1[0] + f2[0]), oy = 0.5 * (f1[1] + f2[1]), oz = 0.5 * (f1[2] + f2[2]);
        double ds = ox * ox + oy * oy + oz * oz;
        double u = ox * pv[1] + oy * pv[2] + oz * pv[3];
        double up = 0.5 * (1.0 + sndp2 / dv[5]) * u;
        double omq = 1.0 - sndp2 / dv[5];
        double dd = 0.5 * sqrt((u * u) * (omq * omq) + 4.0 * sndp2 * ds);
        if (d == 0) lamda = cappa * (fabs(up) + dd);
        else lamda = lamda + cappa * (fabs(up) + dd);
    }
    return lamda;
}

static double lhs_eigenvis(const orc_cfg *c, const double *const f[6], const double *dv, const double *tv)
{
    double ev = orc_eigenvis(c, dv, tv);
    if (c->iturb == -3) return 0.0;
    double s = 0.0;
    for (int d = 0; d < 3; d++)
        for (int m = 0; m < 3; m++) {
            double q = f[2 * d][m] + f[2 * d + 1][m];
            s = (d == 0 && m == 0) ? 0.25 * (q * q) : s + 0.25 * (q * q);
        }
    return ev * s;
}

/* src/lhs_module.f90:346-457: x = getx(res) */
void orc_getx(const orc_cfg *c, const double *const f[6], const double *grd, const double *pv,
              const double *dv, const double *tv, double dt, const double *cc, const double *tt4,
              const double *res, double *x)
{
    int npv = c->npv;
    double lb = (c->nsteady == 1) ? 1.0 : (c->nsteady == 2 ? 1.5 * 0.5 : 0.0);
    double a = 1.0 / (grd[0] / dt + lhs_eigen(c, f, pv, dv));
    double b = ((double)c->nsteady * lb * grd[0] / c->dt_phy + 2.0 * lhs_eigenvis(c, f, dv, tv) / grd[0]) * a;
    double cv[4], tv4[4];
    for (int q = 0; q < 4; q++) { cv[q] = cc[q] * grd[0] * a; tv4[q] = tt4[q] * grd[0] * a; }
    double b1 = 1.0 + b;
    double i_b1 = 1.0 / (b1 * dv[0]);
    double uv2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
    double sp = orc_sndp2(c, dv[5], uv2);
    double ss = dv[0] * dv[11] * (1.0 / sp + b / dv[5]);
    double rr = 1.0 / sp - 1.0 / dv[5] + b1 * dv[6];
    double hst = enthalpy(pv, dv) - uv2;
    double cross = dv[13] * dv[8] - dv[12] * dv[9];
    double cross1 = dv[12] * dv[7] - dv[11] * dv[8];
    double cross2 = dv[13] * dv[7] - dv[11] * dv[9];
    double dhdp1 = 1.0 - dv[10] * dv[0];
    double ssy1 = dv[12] * dv[0] * rr + b1 * dv[8] * dhdp1;
    double ssy2 = dv[13] * dv[0] * rr + b1 * dv[9] * dhdp1;
    double ssh = dv[0] * rr * hst - b1 * dhdp1 * dv[0];
    double aaa1 = dv[8] * hst + dv[12] * dv[0] - cross * pv[6];
    double aaa2 = dv[7] * hst + dv[11] * dv[0] - cross2 * pv[6];
    double uuu = res[1] * pv[1] + res[2] * pv[2] + res[3] * pv[3] - res[4];
    double coeff1 = res[0] * aaa1 + dv[8] * uuu + cross * res[6];
    double coeff2 = res[0] * aaa2 + dv[7] * uuu + cross2 * res[6];
    double coeff3 = res[0] * (ssh - ssy2 * pv[6]) + dv[0] * rr * uuu + res[6] * ssy2;
    double mm = b1 * dv[0] * (ss + cv[0] * cross1) - cv[1] * ssy1 + cv[2] * ss;
    mm = 1.0 / mm;
    x[0] = (cv[3] * cross1 * (res[0] * pv[6] - res[6]) - cv[1] * coeff1 + cv[2] * coeff2
            + dv[0] * b1 * (coeff2 + cross1 * res[5] - cross1 * pv[5] * res[0])) * mm;
    x[1] = (res[1] - res[0] * pv[1]) * i_b1;
    x[2] = (res[2] - res[0] * pv[2]) * i_b1;
    x[3] = (res[3] - res[0] * pv[3]) * i_b1;
    x[4] = (cv[3] * i_b1 * ssy1 * (res[6] - res[0] * pv[6]) + cv[0] * coeff1 - cv[2] * i_b1 * coeff3
            - (coeff3 - ssy1 * pv[5] * res[0] + res[5] * ssy1)) * mm;
    x[5] = (ss * (res[5] - res[0] * pv[5]) - cv[3] * ss * i_b1 * (res[6] - res[0] * pv[6]) - cv[0] * coeff2
            + cv[1] * i_b1 * coeff3) * mm;
    x[6] = (res[6] - res[0] * pv[6]) * i_b1;
    if (npv > 7) {
        double t = 1.0 / (tv4[1] * tv4[2] - (tv4[0] + b1 * dv[0]) * (tv4[3] + b1 * dv[0]));
        x[7] = ((res[8] - res[0] * pv[8]) * tv4[1] - (b1 * dv[0] + tv4[3]) * (res[7] - res[0] * pv[7])) * t;
        x[8] = ((res[7] - res[0] * pv[7]) * tv4[2] - (b1 * dv[0] + tv4[0]) * (res[8] - res[0] * pv[8])) * t;
    }
    for (int q = 0; q < npv; q++) x[q] = x[q] * a;
}

/* ------------------------------------------------------------------ */
/* local time step of one cell, src/timestep_module.f90:156-198        */
double orc_localdt(const orc_cfg *c, const double *const f[6], const double *grd, const double *pv,
                   const double *dv, const double *tv)
{
    double c1 = 0.5 * (f[0][0] + f[1][0]), c2 = 0.5 * (f[0][1] + f[1][1]), c3 = 0.5 * (f[0][2] + f[1][2]);
    double e1 = 0.5 * (f[2][0] + f[3][0]), e2 = 0.5 * (f[2][1] + f[3][1]), e3 = 0.5 * (f[2][2] + f[3][2]);
    double t1 = 0.5 * (f[4][0] + f[5][0]), t2 = 0.5 * (f[4][1] + f[5][1]), t3 = 0.5 * (f[4][2] + f[5][2]);
    double uc = fabs(c1 * pv[1] + c2 * pv[2] + c3 * pv[3]);
    double vc = fabs(e1 * pv[1] + e2 * pv[2] + e3 * pv[3]);
    double wc = fabs(t1 * pv[1] + t2 * pv[2] + t3 * pv[3]);
    double s1 = c1 * c1 + c2 * c2 + c3 * c3;
    double s2 = e1 * e1 + e2 * e2 + e3 * e3;
    double s3 = t1 * t1 + t2 * t2 + t3 * t3;
    double uv2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
    double sndp2 = orc_sndp2(c, dv[5], uv2);
    double q = sndp2 / dv[5], omq = 1.0 - sndp2 / dv[5];
    double up = uc * (1.0 + q);
    double d = sqrt((uc * uc) * (omq * omq) + 4.0 * sndp2 * s1);
    double eigenx = 0.5 * (up + d);
    up = vc * (1.0 + q);
    d = sqrt((vc * vc) * (omq * omq) + 4.0 * sndp2 * s2);
    double eigeny = 0.5 * (up + d);
    up = wc * (1.0 + q);
    d = sqrt((wc * wc) * (omq * omq) + 4.0 * sndp2 * s3);
    double eigenz = 0.5 * (up + d);
    double eigenvis = orc_eigenvis(c, dv, tv) * (s1 + s2 + s3) / grd[0];
    return c->cfl * grd[0] / (eigenx + eigeny + eigenz + 4.0 * eigenvis);
}

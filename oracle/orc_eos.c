/*
 * ORACLE — test infrastructure only.  Nothing in the product path may
 * include, link or call this file; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Plain-C FP64 restatement of the reference water/vapour/gas equation of
 * state, operation by operation, so that it can be compiled with
 * -ffp-contract=off and compared against the HIP kernels.
 *
 *   iapws97_l      src/eos_module.f90:1331-1374  (IF97 region 1, liquid)
 *   iapws97_v      src/eos_module.f90:1376-1461  (IF97 region 2 / metastable)
 *   *_prop         src/eos_module.f90:1463-1659  (IAPS-85 mu, kappa)
 *   ideal(_prop)   src/eos_module.f90:1236-1282
 *   h2o_pww/tww    src/eos_module.f90:2266-2279, 2340-2354
 *   deteos_simple  src/eos_module.f90:1064-1118
 *   deteos_prop    src/eos_module.f90:1176-1234
 *
 * The coefficient tables are the published IAPWS-IF97 (2007) and IAPS-85
 * values; the reference stores the same numbers (src/eos_module.f90:536-720).
 *
 * Five constants in the reference are default-REAL (single precision)
 * Fortran literals and are therefore rounded to float before use; they are
 * reproduced here as (double)NNNf.
 *
 * Pins: IF97 verification tables 5, 15, 18, 33 (region 4), closed-form
 * folded constants and mixture-rule identities are checked in
 * tests/test_oracle_eos.py.
 */
#include "orc.h"
#include <math.h>

/* libgcc __powidf2: the algorithm gfortran uses for real**integer with a
 * run-time exponent (binary exponentiation, reciprocal for m<0). */
double orc_powi(double x, int m)
{
    unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* ---------------- IF97 region 1 (IAPWS-IF97 table 2) ---------------- */
static const int R1_I[34] = {0,0,0,0,0,0,0,0,1,1,1,1,1,1,2,2,2,2,2,3,3,3,4,4,4,5,8,8,21,23,29,30,31,32};
static const int R1_J[34] = {-2,-1,0,1,2,3,4,5,-9,-7,-1,0,1,3,-3,0,1,3,17,-4,0,6,-5,-2,10,-8,-11,-6,-29,-31,-38,-39,-40,-41};
static const double R1_N[34] = {
     0.14632971213167e0, -0.84548187169114e0, -0.37563603672040e1,  0.33855169168385e1,
    -0.95791963387872e0,  0.15772038513228e0, -0.16616417199501e-1, 0.81214629983568e-3,
     0.28319080123804e-3,-0.60706301565874e-3,-0.18990068218419e-1,-0.32529748770505e-1,
    -0.21841717175414e-1,-0.52838357969930e-4,-0.47184321073267e-3,-0.30001780793026e-3,
     0.47661393906987e-4,-0.44141845330846e-5,-0.72694996297594e-15,-0.31679644845054e-4,
    -0.28270797985312e-5,-0.85205128120103e-9,-0.22425281908000e-5,-0.65171222895601e-6,
    -0.14341729937924e-12,-0.40516996860117e-6,-0.12734301741641e-8,-0.17424871230634e-9,
    -0.68762131295531e-18, 0.14478307828521e-19, 0.26335781662795e-22,-0.11947622640071e-22,
     0.18228094581404e-23,-0.93537087292458e-25};

/* ---------------- IF97 region 2 ideal part (table 10) ---------------- */
static const int R2_J0[9] = {0,1,-5,-4,-3,-2,-1,2,3};
static const double R2_N0[9] = {
    -0.96927686500217e1, 0.10086655968018e2, -0.56087911283020e-2, 0.71452738081455e-1,
    -0.40710498223928e0, 0.14240819171444e1, -0.43839511319450e1, -0.28408632460772e0,
     0.21268463753307e-1};
/* metastable-vapour ideal part: n1, n2 re-fitted (IF97 eq. 18) */
static const double R2M_N0[9] = {
    -0.96937268393049e1, 0.10087275970006e2, -0.56087911283020e-2, 0.71452738081455e-1,
    -0.40710498223928e0, 0.14240819171444e1, -0.43839511319450e1, -0.28408632460772e0,
     0.21268463753307e-1};
/* ---------------- IF97 region 2 residual part (table 11) ---------------- */
static const int R2_IR[43] = {1,1,1,1,1,2,2,2,2,2,3,3,3,3,3,4,4,4,5,6,6,6,7,7,7,8,8,9,10,10,10,16,16,18,20,20,20,21,22,23,24,24,24};
static const int R2_JR[43] = {0,1,2,3,6,1,2,4,7,36,0,1,3,6,35,1,2,3,7,3,16,35,0,11,25,8,36,13,4,10,14,29,50,57,20,35,48,21,53,39,26,40,58};
static const double R2_NR[43] = {
    -0.17731742473213e-2,-0.17834862292358e-1,-0.45996013696365e-1,-0.57581259083432e-1,
    -0.50325278727930e-1,-0.33032641670203e-4,-0.18948987516315e-3,-0.39392777243355e-2,
    -0.43797295650573e-1,-0.26674547914087e-4, 0.20481737692309e-7, 0.43870667284435e-6,
    -0.32277677238570e-4,-0.15033924542148e-2,-0.40668253562649e-1,-0.78847309559367e-9,
     0.12790717852285e-7, 0.48225372718507e-6, 0.22922076337661e-5,-0.16714766451061e-10,
    -0.21171472321355e-2,-0.23895741934104e2,-0.59059564324270e-17,-0.12621808899101e-5,
    -0.38946842435739e-1, 0.11256211360459e-10,-0.82311340897998e1, 0.19809712802088e-7,
     0.10406965210174e-18,-0.10234747095929e-12,-0.10018179379511e-8,-0.80882908646985e-10,
     0.10693031879409e0,-0.33662250574171e0, 0.89185845355421e-24, 0.30629316876232e-12,
    -0.42002467698208e-5,-0.59056029685639e-25, 0.37826947613457e-5,-0.12768608934681e-14,
     0.73087610595061e-28, 0.55414715350778e-16,-0.94369707241210e-6};
/* ---------------- metastable-vapour residual part (table 16) ---------------- */
static const int R2M_IR[13] = {1,1,1,1,2,2,2,3,3,4,4,5,5};
static const int R2M_JR[13] = {0,2,5,11,1,7,16,4,16,7,10,9,10};
static const double R2M_NR[13] = {
    -0.73362260186506e-2,-0.88223831943146e-1,-0.72334555213245e-1,-0.40813178534455e-2,
     0.20097803380207e-2,-0.53045921898642e-1,-0.76190409086970e-2,-0.63498037657313e-2,
    -0.86043093028588e-1, 0.75321581522770e-2,-0.79238375446139e-2,-0.22888160778447e-3,
    -0.26456501482810e-2};

/* ---------------- IAPS-85 transport coefficients ---------------- */
static const double TP_H[4] = {1.0, 0.978197, 0.579829, -0.202354};
static const double TP_L[4] = {1.0, 6.978267, 2.599096, -0.998254};
static const double TP_HH[6][7] = {
    { 0.5132047, 0.2151778,-0.2818107, 0.1778064,-0.0417661, 0.0,        0.0},
    { 0.3205656, 0.7317883,-1.070786,  0.460504,  0.0,      -0.01578386, 0.0},
    { 0.0,       1.241044, -1.263184,  0.2340379, 0.0,       0.0,        0.0},
    { 0.0,       1.476783,  0.0,      -0.4924179, 0.1600435, 0.0,       -0.003629481},
    {-0.7782567, 0.0,       0.0,       0.0,       0.0,       0.0,        0.0},
    { 0.1885447, 0.0,       0.0,       0.0,       0.0,       0.0,        0.0}};
static const double TP_LL[5][6] = {
    { 1.3293046,-0.40452437, 0.24409490, 0.018660751,-0.12961068, 0.044809953},
    { 1.7018363,-2.2156845,  1.6511057, -0.76736002,  0.37283344,-0.11203160},
    { 5.2246158,-10.124111,  4.9874687, -0.27297694, -0.43083393, 0.13333849},
    { 8.7127675,-9.5000611,  4.3786606, -0.91783782,  0.0,        0.0},
    {-1.8525999, 0.93404690, 0.0,        0.0,         0.0,        0.0}};

/* src/eos_module.f90:2266-2279 — IF97 region-4 saturation pressure. */
double orc_h2o_pww(double t)
{
    double tt = fmin_f(fmax_f(t, 273.15), 647.096);
    double theta = tt - 0.23855557567849e0 / (tt - 0.65017534844798e3);
    double a = theta * theta + 0.11670521452767e4 * theta - 0.72421316703206e6;
    double b = -0.17073846940092e2 * (theta * theta) + 0.12020824702470e5 * theta - 0.32325550322333e7;
    double c = 0.14915108613530e2 * (theta * theta) - 0.48232657361591e4 * theta + 0.40511340542057e6;
    double x = 2.0 * c / (-b + sqrt(b * b - 4.0 * a * c));
    double x2 = x * x;
    return 1.0e6 * (x2 * x2);
}

/* src/eos_module.f90:2340-2354 — IF97 region-4 saturation temperature. */
double orc_h2o_tww(double p)
{
    double beta = pow(p / 1.0e6, 0.25);
    double e = beta * beta - 0.17073846940092e2 * beta + 0.14915108613530e2;
    double f = 0.11670521452767e4 * (beta * beta) + 0.12020824702470e5 * beta - 0.48232657361591e4;
    double g = -0.72421316703206e6 * (beta * beta) - 0.32325550322333e7 * beta + 0.40511340542057e6;
    double d = 2.0 * g / (-f - sqrt(f * f - 4.0 * e * g));
    double s = 0.65017534844798e3 + d;
    return 0.5 * (0.65017534844798e3 + d - sqrt(s * s - 4.0 * (-0.23855557567849e0 + 0.65017534844798e3 * d)));
}

/* src/eos_module.f90:1331-1374 (and the identical EOS half of 1463-1506). */
void orc_iapws97_l(double p, double t, orc_eos2 *e)
{
    double pi = p * 0.00000006049607;
    double tau = 1386.0 / t;
    double pi1 = 1.0 / (7.1 - pi);
    double tau1 = 1.0 / (tau - 1.222);
    double gt = 0.0, gp = 0.0, gpp = 0.0, gtt = 0.0, gpt = 0.0;
    for (int k = 0; k < 34; k++) {
        double g = R1_N[k] * orc_powi(7.1 - pi, R1_I[k]) * orc_powi(tau - 1.222, R1_J[k]);
        double g1 = (double)R1_J[k] * g * tau1;
        double g2 = -((double)R1_I[k] * g * pi1);
        double g3 = -((double)(R1_I[k] - 1) * g2 * pi1);
        double g4 = (double)(R1_J[k] - 1) * g1 * tau1;
        double g5 = (double)R1_J[k] * g2 * tau1;
        gt = gt + g1;
        gp = gp + g2;
        gpp = gpp + g3;
        gtt = gtt + g4;
        gpt = gpt + g5;
    }
    double gp1 = 1.0 / gp;
    double tau3 = tau * tau * tau;
    e->rho = (double)25.841246054191803727f * tau * gp1;
    e->h = 639675.036 * gt;
    e->drdp = -(gpp * 1.5632937721834122e-6 * tau * (gp1 * gp1));
    e->drdt = -(e->rho * 7.215007215007215e-4 * tau) + (double)0.01864447767257706f * gpt * tau3 * (gp1 * gp1);
    e->dhdp = (double)0.03869782431941923775f * gpt;
    e->dhdt = -(461.526 * (tau * tau) * gtt);
}

/* src/eos_module.f90:1376-1461 (and the EOS half of 1541-1626). */
void orc_iapws97_v(double p, double t, orc_eos2 *e)
{
    double pww = orc_h2o_pww(t);
    double pi = p * 1.0e-6;
    double tau = 540.0 / t;
    double pi1 = 1.0e6 / p;
    double tau1 = 1.0 / tau;
    double tau2 = 1.0 / (tau - 0.5);
    double g0t = 0.0, g0p = pi1, g0pp = -(pi1 * pi1), g0tt = 0.0, g0pt = 0.0;
    double grt = 0.0, grp = 0.0, grpp = 0.0, grtt = 0.0, grpt = 0.0;
    const double *n0;
    const int *ir, *jr;
    const double *nr;
    int nterm;
    if (p > pww) {            /* metastable vapour */
        n0 = R2M_N0; ir = R2M_IR; jr = R2M_JR; nr = R2M_NR; nterm = 13;
    } else {
        n0 = R2_N0; ir = R2_IR; jr = R2_JR; nr = R2_NR; nterm = 43;
    }
    for (int k = 0; k < nterm; k++) {
        if (k < 9) {
            double g0 = n0[k] * orc_powi(tau, R2_J0[k]);
            double g01 = (double)R2_J0[k] * g0 * tau1;
            double g02 = (double)(R2_J0[k] - 1) * g01 * tau1;
            g0t = g0t + g01;
            g0tt = g0tt + g02;
        }
        double gr = nr[k] * orc_powi(pi, ir[k]) * orc_powi(tau - 0.5, jr[k]);
        double gr1 = (double)jr[k] * gr * tau2;
        double gr2 = (double)ir[k] * gr * pi1;
        double gr3 = (double)(ir[k] - 1) * gr2 * pi1;
        double gr4 = (double)(jr[k] - 1) * gr1 * tau2;
        double gr5 = (double)jr[k] * gr2 * tau2;
        grt = grt + gr1;
        grp = grp + gr2;
        grpp = grpp + gr3;
        grtt = grtt + gr4;
        grpt = grpt + gr5;
    }
    double gp1 = 1.0 / (g0p + grp);
    double tau3 = tau * tau * tau;
    e->rho = 4.01245401527075799 * tau * gp1;
    e->h = 249224.04 * (g0t + grt);
    e->drdp = -((g0pp + grpp) * 4.01245401527075799e-6 * tau * (gp1 * gp1));
    e->drdt = -(e->rho * (double)0.00185185185185185185f * tau)
              + (double)0.0074304703986495518f * (g0pt + grpt) * tau3 * (gp1 * gp1);
    e->dhdp = 0.24922404 * (g0pt + grpt);
    e->dhdt = -(461.526 * (tau * tau) * (g0tt + grtt));
}

/* IAPS-85 viscosity / conductivity shared by both IF97 *_prop routines
 * (src/eos_module.f90:1508-1537, 1628-1657); no critical enhancement. */
static void iaps85(double t, double rho, orc_prop2 *pr)
{
    double tt = t * 0.00154495032985;
    double rr = fmax_f(1.0e-3, fmin_f(1332.4, rho)) * 0.00314699949333;
    double tt1 = 647.27 / t;
    double mu0 = sqrt(tt) / (TP_H[0] + TP_H[1] * tt1 + TP_H[2] * (tt1 * tt1) + TP_H[3] * (tt1 * tt1 * tt1));
    double mu1 = 0.0;
    for (int i = 0; i <= 5; i++)
        for (int j = 0; j <= 6; j++)
            mu1 = mu1 + TP_HH[i][j] * orc_powi(tt1 - 1.0, i) * orc_powi(rr - 1.0, j);
    mu1 = exp(rr * mu1);
    double mu2 = 1.0;
    pr->vis = 55.071e-6 * mu0 * mu1 * mu2;
    double lam0 = sqrt(tt) / (TP_L[0] + TP_L[1] * tt1 + TP_L[2] * (tt1 * tt1) + TP_L[3] * (tt1 * tt1 * tt1));
    double lam1 = 0.0;
    for (int i = 0; i <= 4; i++)
        for (int j = 0; j <= 5; j++)
            lam1 = lam1 + TP_LL[i][j] * orc_powi(tt1 - 1.0, i) * orc_powi(rr - 1.0, j);
    lam1 = exp(rr * lam1);
    double lam2 = 0.0;
    pr->cond = 0.4945 * (lam0 * lam1 + lam2);
}

void orc_iapws97_l_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr)
{
    orc_iapws97_l(p, t, e);
    iaps85(t, e->rho, pr);
}

void orc_iapws97_v_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr)
{
    orc_iapws97_v(p, t, e);
    iaps85(t, e->rho, pr);
}

/* src/eos_module.f90:1236-1255 */
void orc_ideal(double p, double t, orc_eos2 *e)
{
    const double r1 = 0.00348383500557;
    double tau = 1.0 / t;
    e->rho = p * r1 * tau;
    e->h = 1004.64 * t;
    e->drdp = r1 * tau;
    e->drdt = -(p * r1 * (tau * tau));
    e->dhdp = 0.0;
    e->dhdt = 1004.64;
}

/* src/eos_module.f90:1257-1282 (Sutherland, Pr = 0.72) */
void orc_ideal_prop(double p, double t, orc_eos2 *e, orc_prop2 *pr)
{
    orc_ideal(p, t, e);
    pr->vis = 1.716e-5 * pow(t / 273.11, 1.5) * 383.67 / (t + 110.56);
    pr->cond = 1395.333333 * pr->vis;
}

/* Mixture rules shared by deteos_simple / deteos_prop
 * (src/eos_module.f90:1077-1116, 1191-1229). */
static void mixture(const orc_eos2 *l, const orc_eos2 *v, const orc_eos2 *g,
                    double y1, double y2, double *dv, int *alt_snd2)
{
    dv[2] = l->rho;
    dv[3] = v->rho;
    dv[4] = g->rho;
    double irhol = 1.0 / dv[2];
    double irhov = 1.0 / dv[3];
    double irhog = 1.0 / dv[4];
    double yl = 1.0 - y1 - y2;
    double irho = yl * irhol + y1 * irhov + y2 * irhog;
    dv[0] = 1.0 / irho;
    dv[1] = yl * l->h + y1 * v->h + y2 * g->h;
    double r2 = dv[0] * dv[0];
    dv[6] = r2 * (yl * (irhol * irhol) * l->drdp + y1 * (irhov * irhov) * v->drdp + y2 * (irhog * irhog) * g->drdp);
    dv[7] = r2 * (yl * (irhol * irhol) * l->drdt + y1 * (irhov * irhov) * v->drdt + y2 * (irhog * irhog) * g->drdt);
    dv[8] = r2 * (irhol - irhov);
    dv[9] = r2 * (irhol - irhog);
    dv[10] = yl * l->dhdp + y1 * v->dhdp + y2 * g->dhdp;
    dv[11] = yl * l->dhdt + y1 * v->dhdt + y2 * g->dhdt;
    dv[12] = v->h - l->h;
    dv[13] = g->h - l->h;
    dv[5] = dv[0] * dv[11] / (dv[0] * dv[6] * dv[11] + dv[7] * (1.0 - dv[0] * dv[10]));
    if (dv[5] <= 0.0) {
        /* src/eos_module.f90:1104-1111 */
        if (alt_snd2) (*alt_snd2)++;
        double ic2l = (dv[2] * l->drdp * l->dhdt + l->drdt * (1.0 - dv[2] * l->dhdp)) / (dv[2] * l->dhdt);
        double ic2v = (dv[3] * v->drdp * v->dhdt + v->drdt * (1.0 - dv[3] * v->dhdp)) / (dv[3] * v->dhdt);
        double ic2g = (dv[4] * g->drdp * g->dhdt + g->drdt * (1.0 - dv[4] * g->dhdp)) / (dv[4] * g->dhdt);
        dv[5] = 1.0 / (r2 * (yl * ic2l * (irhol * irhol) + y1 * ic2v * (irhov * irhov) + y2 * ic2g * (irhog * irhog)));
    }
    dv[14] = v->drdp;
    dv[15] = v->drdt;
    dv[16] = l->drdp;
    dv[17] = l->drdt;
}

int orc_alt_snd2_count = 0;

/* src/eos_module.f90:1064-1118 (fluid = IF97, ngas = ideal). */
void orc_deteos_simple(double p, double t, double y1, double y2, double *dv)
{
    orc_eos2 l, v, g;
    orc_iapws97_l(p, t, &l);
    orc_iapws97_v(p, t, &v);
    orc_ideal(p, t, &g);
    mixture(&l, &v, &g, y1, y2, dv, &orc_alt_snd2_count);
}

/* src/eos_module.f90:1176-1234: dv(18) plus tv(1:2); tv(3) untouched. */
void orc_deteos_prop(double p, double t, double y1, double y2, double *dv, double *tv)
{
    orc_eos2 l, v, g;
    orc_prop2 lp, vp, gp;
    orc_iapws97_l_prop(p, t, &l, &lp);
    orc_iapws97_v_prop(p, t, &v, &vp);
    orc_ideal_prop(p, t, &g, &gp);
    mixture(&l, &v, &g, y1, y2, dv, &orc_alt_snd2_count);
    double irhol = 1.0 / dv[2], irhov = 1.0 / dv[3], irhog = 1.0 / dv[4];
    double yl = 1.0 - y1 - y2;
    tv[0] = dv[0] * (lp.vis * yl * irhol + vp.vis * y1 * irhov + gp.vis * y2 * irhog);
    tv[1] = dv[0] * (lp.cond * yl * irhol + vp.cond * y1 * irhov + gp.cond * y2 * irhog);
}

/* src/eos_module.f90:1120-1174: dv only (Euler runs). */
void orc_deteos_only(double p, double t, double y1, double y2, double *dv)
{
    orc_deteos_simple(p, t, y1, y2, dv);
}

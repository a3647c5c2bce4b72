// Pointwise pieces of the discrete adjoint (cav3d_rhs_vjp): the gradient of
// one face's inviscid flux, one face's viscous flux and one cell's
// properties with respect to their inputs, by dual-number passes (Dn<W>) of
// the templated physics in phys.hpp.  Host and device: the device kernels
// call them per face / cell, and the host build lets them be checked against
// finite differences of the same physics compiled for the CPU.
#pragma once
#include "phys.hpp"

namespace cav {

constexpr int NSLOT = 38;   // the reference's face stencil, Fortran numbering 1..38

// Gradient of a scalar function of n inputs by dual passes of W lanes:
// x(q) is the value of input q; f(in, base) evaluates phi reading input q as
// in(q), which carries lane q - base when that lies in 0..W-1; gx[q] =
// d phi / d x[q].  The inputs are seeded where they are read (no array of
// duals is kept), which keeps the private memory per lane small.
template <int W, class X, class F>
HD void dn_grad(int n, X x, F f, double *gx)
{
    for (int base = 0; base < n; base += W) {
        auto in = [&](int q) -> Dn<W> { return Dn<W>::seed(x(q), q - base); };
        const Dn<W> phi = f(in, base);
        for (int l = 0; l < W; l++) if (base + l < n) gx[base + l] = phi.d[l];
    }
}

// muscl_clamps with the origin of every face value: 0 the reconstruction,
// 1 the cell, 2 the pressure floor (a constant in absolute pressure)
__host__ __device__ inline void clamps_src(int npv, double pref, const double *c9, const double *c10, double *xl, double *xr,
                           int *ol, int *orr)
{
    for (int n = 0; n < NPV; n++) ol[n] = orr[n] = 0;
    const double fl = -pref + 10.0;
    if (!(xl[0] > fl)) { xl[0] = fl; ol[0] = 2; }
    if (!(xr[0] > fl)) { xr[0] = fl; orr[0] = 2; }
    if ((xl[5] < 0.0) || (xl[5] > 1.0)) { xl[5] = c9[5]; ol[5] = 1; }
    if ((xr[5] < 0.0) || (xr[5] > 1.0)) { xr[5] = c10[5]; orr[5] = 1; }
    if ((xl[6] < 0.0) || (xl[6] > 1.0)) { xl[6] = c9[6]; ol[6] = 1; }
    if ((xr[6] < 0.0) || (xr[6] > 1.0)) { xr[6] = c10[6]; orr[6] = 1; }
    if ((xl[5] + xl[6]) > 1.0) { xl[5] = c9[5]; xl[6] = c9[6]; ol[5] = ol[6] = 1; }
    if ((xr[5] + xr[6]) > 1.0) { xr[5] = c10[5]; xr[6] = c10[6]; orr[5] = orr[6] = 1; }
    if (npv > 7) {
        if ((xl[7] < 0.0) && (c9[7] > 0.0)) { xl[7] = c9[7]; ol[7] = 1; }
        if ((xr[7] < 0.0) && (c10[7] > 0.0)) { xr[7] = c10[7]; orr[7] = 1; }
        if ((xl[8] < 0.0) && (c9[8] > 0.0)) { xl[8] = c9[8]; ol[8] = 1; }
        if ((xr[8] < 0.0) && (c10[8] > 0.0)) { xr[8] = c10[8]; orr[8] = 1; }
    }
}

// the Roe-averaged c^2 of inviscid_face from its ten inputs z(0..9) =
// {xl0, xr0, xl4, xr4, xl5, xr5, xl6, xr6, rho_L, rho_R}
template <class R, class Z> HD R c2_average(const Par &P, Z z)
{
    R sl = sqrt(z(8)), sr = sqrt(z(9)), rd = R(1) / (sl + sr);
    R ya1 = (sl * z(4) + sr * z(5)) * rd, ya2 = (sl * z(6) + sr * z(7)) * rd;
    return face_c2<R>(R(0.5) * (z(0) + z(1)) + R(P.pref), R(0.5) * (z(2) + z(3)), ya1, ya2);
}

// The inviscid adjoint of one face up to its reconstruction, independent of
// its direction: f0 points at the left cell's pressure, off[ll-1] is stencil
// point ll's offset and ng the component stride; cellp = {rho, h, c^2} of the
// left and right cells (read by first-order faces); g the seed.  Out: the
// seeds of the reconstruction (xlb, xrb: the reconstructed values; c9b, c10b:
// the cells' own values where a clamp chose them), senb the sensor
// pressures' bars (sensor schemes), dcbo[(side*ND + m)*dstr] the two cells'
// property bars, ub the u_ref derivative.  Inlined into its kernel: as a called function its locals
// lived in 5.3 KB of private memory per lane, inlined in 1.4 KB.
template <bool SENSOR>
__host__ __device__ __attribute__((always_inline)) inline void face_adj_core(const Par &P, int npv, const double *f0, long ng, const long *off,
                                      const double *n3, const double *cellp, const double *g, double *xlb, double *xrb,
                                      double *c9b, double *c10b, double *senb, double *dcbo, long dstr, double &ub)
{
    auto xs = [&](int ll, int n) -> double { return f0[(long)n * ng + off[ll - 1]]; };
    // forward recompute: reconstruction, clamps, face EOS, averaged c^2
    double c9[NPV], c10[NPV], xl[NPV], xr[NPV], g3[6][3];
    Par Pi = P;
    Pi.iturb = -3;
    face_recon<double>(Pi, npv, xs, c9, c10, xl, xr, g3);
    double ql[NPV], qr[NPV];
    int ol[NPV], orr[NPV];
    for (int n = 0; n < NPV; n++) { ql[n] = xl[n]; qr[n] = xr[n]; ol[n] = orr[n] = 0; }
    FaceState<double> L, Rt;
    if (P.nmuscl > 0) {
        clamps_src(npv, P.pref, c9, c10, ql, qr, ol, orr);
        face_rho_h<double>(ql[0] + P.pref, ql[4], ql[5], ql[6], L.rho, L.h);
        face_rho_h<double>(qr[0] + P.pref, qr[4], qr[5], qr[6], Rt.rho, Rt.h);
        if (SENSOR) {
            L.c2 = face_c2<double>(ql[0] + P.pref, ql[4], ql[5], ql[6]);
            Rt.c2 = face_c2<double>(qr[0] + P.pref, qr[4], qr[5], qr[6]);
        }
    } else {
        L.rho = cellp[0]; L.h = cellp[1]; L.c2 = cellp[2];
        Rt.rho = cellp[3]; Rt.h = cellp[4]; Rt.c2 = cellp[5];
    }
    for (int n = 0; n < NPV; n++) { L.p[n] = ql[n]; Rt.p[n] = qr[n]; }
    const double za[10] = {ql[0], qr[0], ql[4], qr[4], ql[5], qr[5], ql[6], qr[6], L.rho, Rt.rho};
    auto zav = [&](int q) -> double { return za[q]; };
    const double c2a = c2_average<double>(P, zav);
    double sd[18];
    if (SENSOR) for (int ll = 1; ll <= 18; ll++) sd[ll - 1] = f0[off[ll - 1]];

    // flux core: z = {pL[npv], pR[npv], rhoL, hL, rhoR, hR, c2a, [c2L, c2R, sd[18]], u_ref}
    constexpr int NZ = 2 * NPV + 5 + (SENSOR ? 20 : 0) + 1;
    const int o = 2 * npv;
    const int nz = o + 5 + (SENSOR ? 20 : 0) + 1;
    double zx[NZ], zb[NZ];
    for (int n = 0; n < npv; n++) { zx[n] = L.p[n]; zx[npv + n] = Rt.p[n]; }
    zx[o] = L.rho; zx[o + 1] = L.h; zx[o + 2] = Rt.rho; zx[o + 3] = Rt.h; zx[o + 4] = c2a;
    if (SENSOR) {
        zx[o + 5] = L.c2; zx[o + 6] = Rt.c2;
        for (int q = 0; q < 18; q++) zx[o + 7 + q] = sd[q];
    }
    zx[nz - 1] = P.uref;
    {
        constexpr int W = 2;
        using T = Dn<W>;
        dn_grad<W>(nz, [&](int q) -> double { return zx[q]; }, [&](auto in, int base) -> T {
            FaceState<T> l, r;
            for (int n = 0; n < NPV; n++) {
                l.p[n] = n < npv ? in(n) : T(0.0);
                r.p[n] = n < npv ? in(npv + n) : T(0.0);
            }
            l.rho = in(o); l.h = in(o + 1); r.rho = in(o + 2); r.h = in(o + 3);
            l.c2 = SENSOR ? in(o + 5) : T(0.0);
            r.c2 = SENSOR ? in(o + 6) : T(0.0);
            Par Pu = P;
            const int lu = nz - 1 - base;
            Pu.uref_seed = (lu >= 0 && lu < W) ? lu + 1 : 0;
            const T nt[3] = {T(n3[0]), T(n3[1]), T(n3[2])};
            T fx[NPV];
            if constexpr (SENSOR) {
                T sdt[18];
                for (int q = 0; q < 18; q++) sdt[q] = in(o + 7 + q);
                if (P.nscheme == 2) roem_flux<T>(Pu, npv, nt, l, r, in(o + 4), sdt, fx);
                else ausmpwp_flux<T>(Pu, npv, nt, l, r, in(o + 4), sdt, fx);
            } else {
                if (P.nscheme == 1) roe_flux<T>(Pu, npv, nt, l, r, in(o + 4), fx);
                else ausm_flux<T>(Pu, npv, nt, l, r, in(o + 4), fx);
            }
            T phi(0.0);
            for (int n = 0; n < npv; n++) phi = phi + g[n] * fx[n];
            return phi;
        }, zb);
    }
    ub = zb[nz - 1];
    // bars of the (clamped) face states and of the two sides' rho, h, c^2
    double qlb[NPV], qrb[NPV];
    for (int n = 0; n < NPV; n++) { qlb[n] = n < npv ? zb[n] : 0.0; qrb[n] = n < npv ? zb[npv + n] : 0.0; }
    double rhoLb = zb[o], hLb = zb[o + 1], rhoRb = zb[o + 2], hRb = zb[o + 3];
    double c2Lb = SENSOR ? zb[o + 5] : 0.0, c2Rb = SENSOR ? zb[o + 6] : 0.0;
    {   // averaged c^2 = c^2(pa, ta, ya1, ya2): one 4-lane pass of the EOS, then
        // the closed forms of pa, ta and the Roe-averaged fractions
        const double c2ab = zb[o + 4];
        const double sl = sqrt(L.rho), sr = sqrt(Rt.rho), S = sl + sr, rd = 1.0 / S;
        const double ya1 = (sl * ql[5] + sr * qr[5]) * rd, ya2 = (sl * ql[6] + sr * qr[6]) * rd;
        const double x4[4] = {0.5 * (ql[0] + qr[0]) + P.pref, 0.5 * (ql[4] + qr[4]), ya1, ya2};
        double g4[4];
        dn_grad<4>(4, [&](int k) -> double { return x4[k]; }, [&](auto in, int) -> Dn<4> {
            return c2ab * face_c2<Dn<4>>(in(0), in(1), in(2), in(3));
        }, g4);
        qlb[0] += 0.5 * g4[0]; qrb[0] += 0.5 * g4[0];
        qlb[4] += 0.5 * g4[1]; qrb[4] += 0.5 * g4[1];
        qlb[5] += g4[2] * sl * rd; qrb[5] += g4[2] * sr * rd;
        qlb[6] += g4[3] * sl * rd; qrb[6] += g4[3] * sr * rd;
        // d ya / d sl = (x_L - ya) / S, d sl / d rho_L = 0.5 / sl
        rhoLb += (g4[2] * (ql[5] - ya1) + g4[3] * (ql[6] - ya2)) * rd * (0.5 / sl);
        rhoRb += (g4[2] * (qr[5] - ya1) + g4[3] * (qr[6] - ya2)) * rd * (0.5 / sr);
    }
    double dcb[2 * ND];
    for (int m = 0; m < 2 * ND; m++) dcb[m] = 0.0;
    if (P.nmuscl > 0) {   // face EOS of each side
        for (int side = 0; side < 2; side++) {
            const double *q = side ? qr : ql;
            double *qb = side ? qrb : qlb;
            const double rb_ = side ? rhoRb : rhoLb, hb_ = side ? hRb : hLb, cb_ = side ? c2Rb : c2Lb;
            // rho and h of the face state: the mixture EOS's own derivatives
            // (d rho/dp, dT, dy1, dy2; dh/dp, ...) are their exact gradients
            double d[ND];
            mixture_eos<double>(q[0] + P.pref, q[4], q[5], q[6], d, false);
            qb[0] += rb_ * d[D_RP] + hb_ * d[D_HP];
            qb[4] += rb_ * d[D_RT] + hb_ * d[D_HT];
            qb[5] += rb_ * d[D_RY1] + hb_ * d[D_HY1];
            qb[6] += rb_ * d[D_RY2] + hb_ * d[D_HY2];
            if (SENSOR && cb_ != 0.0) {     // c^2 of the face state (sensor schemes)
                const double x4[4] = {q[0] + P.pref, q[4], q[5], q[6]};
                double gx[4];
                dn_grad<4>(4, [&](int k) -> double { return x4[k]; }, [&](auto in, int) -> Dn<4> {
                    return cb_ * face_c2<Dn<4>>(in(0), in(1), in(2), in(3));
                }, gx);
                qb[0] += gx[0]; qb[4] += gx[1]; qb[5] += gx[2]; qb[6] += gx[3];
            }
        }
    } else {
        dcb[D_RHO] = rhoLb; dcb[D_H] = hLb; dcb[ND + D_RHO] = rhoRb; dcb[ND + D_H] = hRb;
        if (SENSOR) { dcb[D_C2] = c2Lb; dcb[ND + D_C2] = c2Rb; }
    }
    // through the clamps to the raw reconstruction or the cells
    for (int n = 0; n < NPV; n++) {
        xlb[n] = ol[n] == 0 ? qlb[n] : 0.0;
        xrb[n] = orr[n] == 0 ? qrb[n] : 0.0;
        c9b[n] = ol[n] == 1 ? qlb[n] : 0.0;
        c10b[n] = orr[n] == 1 ? qrb[n] : 0.0;
    }
    if (SENSOR) for (int q = 0; q < 18; q++) senb[q] = zb[o + 7 + q];
    for (int m = 0; m < 2 * ND; m++) dcbo[(long)m * dstr] = dcb[m];
}

// The reconstruction part of one face's inviscid adjoint, for variable n
// (xs(ll): its value at stencil point ll): the seeds wl, wr of its left and
// right reconstructed values and c9bn, c10bn of the two cells' own values
// (where a clamp chose them) go back to the stencil points; senb (variable
// 0 of the sensor schemes, else null) adds the sensor pressures' bars.
// Out: sbo[(ll-1)*ss] for every ll in mask.
template <class XS>
__host__ __device__ __attribute__((always_inline)) inline void face_adj_recon(const Par &P, int n, XS xs1, double wl, double wr,
                                      double c9bn, double c10bn, const double *senb, unsigned long long mask,
                                      double *sbo, long ss)
{
    constexpr int TVD6[6] = {9, 10, 23, 32, 37, 38};
    auto xs = [&](int ll, int) -> double { return xs1(ll); };
    double sb[NSLOT];
    for (int q = 0; q < NSLOT; q++) sb[q] = 0.0;
    if (P.nmuscl == 0) {
        sb[8] += wl;
        sb[9] += wr;
    } else if (wl != 0.0 || wr != 0.0) {
        if (P.nmuscl == 2) {
            // the points the MLP value can depend on: the fourteen it reads
            // directly and the extremum each min / max group selects (the
            // first one, as mn / mx keep it); the others enter as constants
            int sup[NSLOT], ns = 0;
            unsigned long long seen = 0;
            auto add = [&](int ll) { if (!((seen >> ll) & 1ull)) { seen |= 1ull << ll; sup[ns++] = ll; } };
            constexpr int DIRECT[14] = {9, 10, 23, 32, 37, 38, 3, 4, 7, 8, 11, 12, 15, 16};
            for (int q = 0; q < 14; q++) add(DIRECT[q]);
            auto ext = [&](int lo, int hi, int step) {
                int amin = lo, amax = lo;
                double vmin = xs(lo, n), vmax = vmin;
                for (int q = lo + step; q <= hi; q += step) {
                    const double v = xs(q, n);
                    if (v < vmin) { vmin = v; amin = q; }
                    if (v > vmax) { vmax = v; amax = q; }
                }
                add(amin);
                add(amax);
            };
            ext(1, 18, 1); ext(19, 27, 1); ext(28, 36, 1); ext(1, 17, 2); ext(2, 18, 2);
            // four support points per pass, each read compared against them
            for (int q = 0; q < ns; q += 4) {
                int sl4[4];
                for (int l = 0; l < 4; l++) sl4[l] = q + l < ns ? sup[q + l] : -1;
                Dn<4> a, b;
                muscl_var<Dn<4>>(P.nlim, true, [&](int ll) -> Dn<4> {
                    Dn<4> r(xs(ll, n));
                    for (int l = 0; l < 4; l++)
                        if (ll == sl4[l]) r.d[l] = 1.0;
                    return r;
                }, a, b);
                const Dn<4> phi = wl * a + wr * b;
                for (int l = 0; l < 4; l++)
                    if (sl4[l] > 0) sb[sl4[l] - 1] += phi.d[l];
            }
        } else {
            double gx[6];
            dn_grad<6>(6, [&](int q) -> double { return xs(TVD6[q], n); }, [&](auto in, int) -> Dn<6> {
                Dn<6> a, b;
                muscl_var<Dn<6>>(P.nlim, false, [&](int ll) -> Dn<6> {
                    return in(ll == 9 ? 0 : ll == 10 ? 1 : ll == 23 ? 2 : ll == 32 ? 3 : ll == 37 ? 4 : 5);
                }, a, b);
                return wl * a + wr * b;
            }, gx);
            for (int q = 0; q < 6; q++) sb[TVD6[q] - 1] += gx[q];
        }
    }
    sb[8] += c9bn;
    sb[9] += c10bn;
    if (senb)
        for (int q = 0; q < 18; q++) sb[q] += senb[q];
    for (int ll = 1; ll <= NSLOT; ll++)
        if ((mask >> (ll - 1)) & 1ull) sbo[(long)(ll - 1) * ss] = sb[ll - 1];
}

// The inviscid adjoint of one face: face_adj_core, then face_adj_recon for
// every variable (the kernels run the two parts as separate passes).  Out
// as face_adj_core, and sbo[(ll-1)*ss + n*sn] the bar of stencil point ll of
// variable n for every ll in mask (bit ll-1).
template <bool SENSOR>
__host__ __device__ __attribute__((always_inline)) inline void face_adj(const Par &P, int npv, const double *f0, long ng, const long *off, const double *n3,
                                      const double *cellp, const double *g, unsigned long long mask, double *sbo,
                                      long ss, long sn, double *dcbo, long dstr, double &ub)
{
    double xlb[NPV], xrb[NPV], c9b[NPV], c10b[NPV], senb[18];
    face_adj_core<SENSOR>(P, npv, f0, ng, off, n3, cellp, g, xlb, xrb, c9b, c10b, senb, dcbo, dstr, ub);
    for (int n = 0; n < npv; n++)
        face_adj_recon(P, n, [&](int ll) -> double { return f0[(long)n * ng + off[ll - 1]]; }, xlb[n], xrb[n], c9b[n],
                       c10b[n], SENSOR && n == 0 ? senb : nullptr, mask, sbo + (long)n * sn, ss);
}

// dual passes of one face's viscous flux over its inputs zx (the layout of
// k_vface_vjp); inlined into each direction's kernel (as a called function
// its locals took 4.9 KB of private memory per lane, inlined 0.9 KB).  omt:
// the index of the two cells' eddy viscosities in zx when they are inputs,
// or -1 (held at ml, mr)
__host__ __device__ __attribute__((always_inline)) inline void visc_adj(const Par &P, int npv, int nv, int ncv, int nz, const double *zx, const double *n3,
                                      const double e4[4][3], const double t4[4][3], const double *gl, const double *gr,
                                      const double *dl, const double *dr, double ml, double mr, int omt, const double *g,
                                      double *zb)
{
    const bool turb = P.iturb == 0;
    const int og = 3 * nv, oc = og + ncv;
    {
        constexpr int W = 2;
        using Tn = Dn<W>;
        dn_grad<W>(nz, [&](int q) -> double { return zx[q]; }, [&](auto in, int) -> Tn {
            Tn gt[6][3], cv[5], dlt[ND], drt[ND], nt[3], et[4][3], tt[4][3], glt[2], grt[2];
            for (int q = 0; q < 6; q++) for (int c = 0; c < 3; c++) gt[q][c] = q < nv ? in(3 * q + c) : Tn(0.0);
            for (int q = 0; q < 5; q++) cv[q] = q < ncv ? in(og + q) : Tn(0.0);
            for (int m = 0; m < ND; m++) { dlt[m] = Tn(dl[m]); drt[m] = Tn(dr[m]); }
            dlt[D_MU] = in(oc); dlt[D_KAP] = in(oc + 1); drt[D_MU] = in(oc + 2); drt[D_KAP] = in(oc + 3);
            if (turb) { dlt[D_HT] = in(oc + 4); drt[D_HT] = in(oc + 5); dlt[D_RHO] = in(oc + 6); drt[D_RHO] = in(oc + 7); }
            for (int c = 0; c < 3; c++) nt[c] = Tn(n3[c]);
            for (int a = 0; a < 4; a++) for (int c = 0; c < 3; c++) { et[a][c] = Tn(e4[a][c]); tt[a][c] = Tn(t4[a][c]); }
            glt[0] = Tn(gl[0]); glt[1] = Tn(gl[1]); grt[0] = Tn(gr[0]); grt[1] = Tn(gr[1]);
            Tn fv[NPV];
            const Tn mlt = omt >= 0 ? in(omt) : Tn(ml), mrt = omt >= 0 ? in(omt + 1) : Tn(mr);
            visc_flux<Tn>(P, nt, et, tt, glt, grt, gt, cv, dlt, drt, mlt, mrt, fv);
            Tn phi(0.0);
            for (int n = 0; n < npv; n++) phi = phi + g[n] * fv[n];
            return phi;
        }, zb);
    }
}

// property bars of one cell back to (p, T, y1, y2) through its EOS
__host__ __device__ __attribute__((always_inline)) inline void eos_adj(bool prop, const double *x4, const double *dcb, double *gx)
{
    dn_grad<4>(4, [&](int q) -> double { return x4[q]; }, [&](auto in, int) -> Dn<4> {
        Dn<4> d[ND];
        mixture_eos<Dn<4>>(in(0), in(1), in(2), in(3), d, prop);
        Dn<4> phi(0.0);
        for (int m = 0; m < ND; m++) if (dcb[m] != 0.0) phi = phi + dcb[m] * d[m];
        return phi;
    }, gx);
}

}  // namespace cav

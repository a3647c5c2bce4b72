// cav3d: the pseudo-time LU-SGS step on MI355X (gfx950).
//
// Layout: every block's ghosted cell arrays (-1:imax+3 etc.) are concatenated
// in one allocation per field and stored structure-of-arrays, i fastest, so a
// wavefront reads 64 consecutive cells of one variable.  Persistent state is
// pv (npv fields) and the eddy viscosity; the 13 cell properties the path
// reads (rho, h, c^2, the eight EOS derivatives, mu, kappa) are kept per cell
// as the reference keeps dv/tv, and are re-derived from pv wherever the
// reference calls the EOS.
//
// One step (t_lusgs%timeinteg, src/update_module.f90:455-733):
//   k_timestep            local dt + global min          (timestep_module)
//   k_bc / k_halo         faces, interfaces, edges, corners (bc_module setbc)
//   k_face<D>             reconstruction + EOS + inviscid/viscous face flux
//   k_combine             flux differences in the reference's i,j,k order
//                         + Merkle / SST / BDF sources   (rhs_module calrhs)
//   k_lhs_prep            sweep-independent terms of D^-1 and of the Jacobians
//   k_sweep_tile<fwd>     the L (U) sweep: tiles of i-pencils, hyperplane order,
//                         LDS hand-over inside a tile, coherent polls across tiles
//   k_update, k_eddy      pv += dq, clip, EOS, norms; SST eddy viscosity
#include "phys.hpp"
#include "../../include/cav3d.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace cav;

#define MAXB 16

namespace {

thread_local std::string g_err;
int fail(const std::string &m, int code = -1) { g_err = m; return code; }

#define HIPCHK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_), -3); \
    } while (0)

struct BlkD {
    int I, J, K, nx, ny;
    long g0, c0, w0, f0[3], cx0, ex0, tx0, p0;   // p0: first (j,k) slot of the block's sweep plane
    long ng, nc, nw, nf[3], ncx, nex, ntx;
};

struct Seg {          // one boundary range (face BC, edge or corner)
    int b, fn, face, pad;
    int is[3], ie[3], org[3], dir[3], nb[6][3];
    long start;
};

struct Con {          // one interface connection
    int src, dst;
    int is[3], ie[3], isd[3], tm[9];
    long start;
};

struct DS {           // device state, passed by value
    int nb, npv, nbcseg[3], ncon;
    long NG, NC, NW, NF[3], NCX, NEX, NTX, NBC[3], NHALO;
    double *pv, *dc, *mut, *vol, *ydn, *cx, *ex, *tx;
    double *res, *dt, *icav, *itt, *ocut, *rcav, *qq;
    double *dcv;
    double *xqf, *xqb; // dq of the sweeps [NW][NPV], a cell's values adjacent: forward
                       // the cross-tile values only, backward every cell's
    double *gx;       // diagonal-inverse terms [NGX][NC]
    double *zero;     // zeros, one more than the longest pencil: the sweeps' source for absent items
    double *jc;       // Jacobian cell terms [2][NC]: beta, h
    double *jf;       // Jacobian face terms [18][NC]: faces -i,+i,-j,+j,-k,+k x (u, lam, vis)
    long NPLANE;      // sum over blocks of (J-1)(K-1)
    double *ff[3];    // face fluxes [2*npv][NF]
    double *red;      // norm partials
    double *scal;     // [0] dt min bits, [1] accumulated time, [2..] norms
    Seg *seg[3];
    Con *con;
    BlkD b[MAXB];
    Par P;
};

enum { FN_NONE = 0, FN_WALLINV, FN_WALLVIS, FN_WALLVISKW, FN_INSUB, FN_OUTSUB, FN_FARFIELD, FN_SYM,
       FN_INSUPER, FN_COPY, FN_EMPTY, FN_EDGEWW, FN_EDGEWWKW, FN_EDGENOWALL, FN_CORNERWW,
       FN_CORNERWWKW, FN_CORNERNOWALL };
enum { F_NONE = -1, F_IMIN = 0, F_IMAX, F_JMIN, F_JMAX, F_KMIN, F_KMAX };

// ------------------------------------------------------------ index helpers
__device__ __forceinline__ long gidx(const BlkD &B, int i, int j, int k)
{
    return B.g0 + (long)(i + 1) + (long)B.nx * ((long)(j + 1) + (long)B.ny * (long)(k + 1));
}
__device__ __forceinline__ long cidx(const BlkD &B, int i, int j, int k)
{
    return B.c0 + (long)(i - 2) + (long)(B.I - 1) * ((long)(j - 2) + (long)(B.J - 1) * (long)(k - 2));
}
__device__ __forceinline__ long widx(const BlkD &B, int i, int j, int k)
{
    return B.w0 + (long)(i - 1) + (long)(B.I + 1) * ((long)(j - 1) + (long)(B.J + 1) * (long)(k - 1));
}
// face-vector arrays by axis: 0 -> cx(1:I,1:J+1,1:K+1), 1 -> ex(1:I+1,1:J,1:K+1), 2 -> tx(1:I+1,1:J+1,1:K)
__device__ __forceinline__ void face_vec(const DS &S, const BlkD &B, int ax, int i, int j, int k, double *v)
{
    const double *a;
    long n, o;
    if (ax == 0) { a = S.cx; n = S.NCX; o = B.cx0 + (i - 1) + (long)B.I * ((j - 1) + (long)(B.J + 1) * (k - 1)); }
    else if (ax == 1) { a = S.ex; n = S.NEX; o = B.ex0 + (i - 1) + (long)(B.I + 1) * ((j - 1) + (long)B.J * (k - 1)); }
    else { a = S.tx; n = S.NTX; o = B.tx0 + (i - 1) + (long)(B.I + 1) * ((j - 1) + (long)(B.J + 1) * (k - 1)); }
    v[0] = a[o]; v[1] = a[n + o]; v[2] = a[2 * n + o];
}

__device__ __forceinline__ int find_blk(const DS &S, long t, int which)
{
    int b = 0;
    for (int q = 1; q < S.nb; q++) {
        long o = which == 0 ? S.b[q].c0 : (which == 1 ? S.b[q].g0 : S.b[q].f0[which - 2]);
        if (t >= o) b = q;
    }
    return b;
}

__device__ __forceinline__ void load_cell(const DS &S, const BlkD &B, int i, int j, int k, CellJ &c)
{
    long g = gidx(B, i, j, k);
    for (int n = 0; n < NPV; n++) if (n < S.npv) c.pv[n] = S.pv[n * S.NG + g];
    for (int m = 0; m < ND; m++) c.d[m] = S.dc[m * S.NG + g];
    c.mut = S.mut ? S.mut[g] : 0.0;
    c.vol = S.vol[widx(B, i, j, k)];
}

__device__ __forceinline__ void cell_faces(const DS &S, const BlkD &B, int i, int j, int k, double f6[6][3])
{
    face_vec(S, B, 0, i - 1, j, k, f6[0]);
    face_vec(S, B, 0, i, j, k, f6[1]);
    face_vec(S, B, 1, i, j - 1, k, f6[2]);
    face_vec(S, B, 1, i, j, k, f6[3]);
    face_vec(S, B, 2, i, j, k - 1, f6[4]);
    face_vec(S, B, 2, i, j, k, f6[5]);
}

__device__ __forceinline__ void cell_of(const BlkD &B, long r, int &i, int &j, int &k)
{
    int ni = B.I - 1, nj = B.J - 1;
    i = 2 + (int)(r % ni);
    j = 2 + (int)((r / ni) % nj);
    k = 2 + (int)(r / ((long)ni * nj));
}

// ------------------------------------------------------------ EOS kernels
// dc (and mut = emutref) of interior cells from pv
__global__ __launch_bounds__(256) void k_eos_cells(DS S, int interior_only, int set_mut)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = interior_only ? S.NC : S.NG;
    if (t >= n) return;
    int b = find_blk(S, t, interior_only ? 0 : 1);
    const BlkD &B = S.b[b];
    long g;
    if (interior_only) {
        int i, j, k;
        cell_of(B, t - B.c0, i, j, k);
        g = gidx(B, i, j, k);
    } else g = t;
    double d[ND];
    mixture_eos<double>(S.pv[g] + S.P.pref, S.pv[4 * S.NG + g], S.pv[5 * S.NG + g], S.pv[6 * S.NG + g], d, S.P.iturb != -3);
    for (int m = 0; m < ND; m++) S.dc[m * S.NG + g] = d[m];
    if (set_mut && S.mut) S.mut[g] = S.P.emutref;
}

// full dv(18), tv(1:2) of every ghosted cell, AoS for the host
__global__ __launch_bounds__(256) void k_eos_full(DS S, long g0, long n, double *dv, double *tv2)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    long g = g0 + q;
    double d[ND], full[18];
    mixture_eos<double>(S.pv[g] + S.P.pref, S.pv[4 * S.NG + g], S.pv[5 * S.NG + g], S.pv[6 * S.NG + g], d,
                        S.P.iturb != -3, full);
    for (int m = 0; m < 18; m++) dv[18 * q + m] = full[m];
    tv2[2 * q] = d[D_MU];
    tv2[2 * q + 1] = d[D_KAP];
}

__global__ __launch_bounds__(256) void k_eos_points(Par P, int n, const double *p, const double *t, const double *y1, const double *y2,
                             double *dv, double *tv)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double d[ND], full[18];
    mixture_eos<double>(p[q], t[q], y1[q], y2[q], d, true, full);
    for (int m = 0; m < 18; m++) dv[18 * q + m] = full[m];
    tv[2 * q] = d[D_MU];
    tv[2 * q + 1] = d[D_KAP];
}

// ------------------------------------------------------------ time step
__global__ __launch_bounds__(256) void k_reset_scal(DS S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long big = 0x7ff0000000000000ull;   // +inf
        reinterpret_cast<unsigned long long *>(S.scal)[0] = big;
    }
}

__global__ __launch_bounds__(256) void k_timestep(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 1.0e300;
    if (t < S.NC) {
        const BlkD &B = S.b[find_blk(S, t, 0)];
        int i, j, k;
        cell_of(B, t - B.c0, i, j, k);
        CellJ c;
        load_cell(S, B, i, j, k, c);
        double f6[6][3];
        cell_faces(S, B, i, j, k, f6);
        dt = (S.P.local == -1) ? S.P.cfl : local_dt(S.P, c, f6);
        S.dt[t] = dt;
    }
    // wavefront min, workgroup min, then one atomic per workgroup
    // (positive doubles order as their bit patterns)
    __shared__ double wmin[4];
    for (int o = 32; o > 0; o >>= 1) dt = fmin(dt, __shfl_xor(dt, o));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmin(m, wmin[q]);
        if (m < 1.0e299)
            atomicMin(reinterpret_cast<unsigned long long *>(S.scal), (unsigned long long)__double_as_longlong(m));
    }
}

__global__ __launch_bounds__(256) void k_time_acc(DS S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double m = __longlong_as_double((long long)reinterpret_cast<unsigned long long *>(S.scal)[0]);
        if (S.P.local == -1) m = S.P.cfl;
        S.scal[1] += m;
        S.scal[2] = m;
    }
}

__global__ __launch_bounds__(256) void k_dt_fill(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < S.NC) S.dt[t] = __longlong_as_double((long long)reinterpret_cast<unsigned long long *>(S.scal)[0]);
}

// ------------------------------------------------------------ boundary conditions
__device__ __forceinline__ void put_cell(const DS &S, long g, const double *pv, const double *d, double mut)
{
    for (int n = 0; n < NPV; n++) if (n < S.npv) S.pv[n * S.NG + g] = pv[n];
    for (int m = 0; m < ND; m++) S.dc[m * S.NG + g] = d[m];
    if (S.mut) S.mut[g] = mut;
}

__device__ void surface_of(const Seg &s, int i, int j, int k, int *o)
{
    o[0] = s.org[0] + s.dir[0] * i;
    o[1] = s.org[1] + s.dir[1] * j;
    o[2] = s.org[2] + s.dir[2] * k;
    switch (s.face) {
    case F_IMIN: o[0] = s.org[0] + s.dir[0] * s.ie[0]; break;
    case F_IMAX: o[0] = s.org[0] + s.dir[0] * s.is[0]; break;
    case F_JMIN: o[1] = s.org[1] + s.dir[1] * s.ie[1]; break;
    case F_JMAX: o[1] = s.org[1] + s.dir[1] * s.is[1]; break;
    case F_KMIN: o[2] = s.org[2] + s.dir[2] * s.ie[2]; break;
    case F_KMAX: o[2] = s.org[2] + s.dir[2] * s.is[2]; break;
    default: break;
    }
}

// outward area vector of a boundary face at its surface cell
__device__ void outward_of(const DS &S, const BlkD &B, int face, const int *s, double *n)
{
    double sg = 1.0;
    switch (face) {
    case F_IMIN: face_vec(S, B, 0, s[0] - 1, s[1], s[2], n); sg = -1.0; break;
    case F_IMAX: face_vec(S, B, 0, s[0], s[1], s[2], n); break;
    case F_JMIN: face_vec(S, B, 1, s[0], s[1] - 1, s[2], n); sg = -1.0; break;
    case F_JMAX: face_vec(S, B, 1, s[0], s[1], s[2], n); break;
    case F_KMIN: face_vec(S, B, 2, s[0], s[1], s[2] - 1, n); sg = -1.0; break;
    default: face_vec(S, B, 2, s[0], s[1], s[2], n); break;
    }
    n[0] *= sg; n[1] *= sg; n[2] *= sg;
}

__device__ void apply_bc_cell(const DS &S, const Seg &s, int i, int j, int k)
{
    const BlkD &B = S.b[s.b];
    const Par &P = S.P;
    const int npv = S.npv;
    const double *rp = P.bc_pv;
    int m[3] = {s.org[0] + s.dir[0] * i, s.org[1] + s.dir[1] * j, s.org[2] + s.dir[2] * k};
    long gm = gidx(B, m[0], m[1], m[2]);
    long gg = gidx(B, i, j, k);
    double pv[NPV], d[ND], mut = 0.0;
    for (int n = 0; n < NPV; n++) if (n < npv) pv[n] = S.pv[n * S.NG + gm];
    for (int q = 0; q < ND; q++) d[q] = S.dc[q * S.NG + gm];
    if (S.mut) mut = S.mut[gm];
    switch (s.fn) {
    case FN_WALLVIS:
        pv[1] = -pv[1]; pv[2] = -pv[2]; pv[3] = -pv[3];
        put_cell(S, gg, pv, d, mut);
        break;
    case FN_WALLVISKW: {
        int o[3];
        surface_of(s, i, j, k, o);
        long gs = gidx(B, o[0], o[1], o[2]);
        double yd = S.ydn[widx(B, o[0], o[1], o[2])];
        double var = 1600.0 * S.dc[D_MU * S.NG + gs] / S.dc[D_RHO * S.NG + gs] / (yd * yd) - pv[8];
        pv[1] = -pv[1]; pv[2] = -pv[2]; pv[3] = -pv[3]; pv[7] = -pv[7]; pv[8] = var;
        put_cell(S, gg, pv, d, -mut);
        break;
    }
    case FN_WALLINV: case FN_SYM: {
        int o[3];
        double n[3];
        surface_of(s, i, j, k, o);
        outward_of(S, B, s.face, o, n);
        n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
        long gs = gidx(B, o[0], o[1], o[2]);
        double ps1 = S.pv[1 * S.NG + gs], ps2 = S.pv[2 * S.NG + gs], ps3 = S.pv[3 * S.NG + gs];
        double dn = ps1 * n[0] + ps2 * n[1] + ps3 * n[2];
        double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        double v1 = 2.0 * (ps1 - n[0] * dn / nn) - pv[1];
        double v2 = 2.0 * (ps2 - n[1] * dn / nn) - pv[2];
        double v3 = 2.0 * (ps3 - n[2] * dn / nn) - pv[3];
        pv[1] = v1; pv[2] = v2; pv[3] = v3;
        put_cell(S, gg, pv, d, mut);
        break;
    }
    case FN_INSUB: case FN_OUTSUB: case FN_FARFIELD: case FN_INSUPER: {
        double press = P.bc_pressure;
        if (s.fn == FN_INSUPER) {
            pv[0] = 2.0 * (press - rp[0]) - pv[0];
            for (int q = 1; q < NPV; q++) if (q < npv) pv[q] = 2.0 * rp[q] - pv[q];
            pv[5] = mn(mx(pv[5], 0.0), 1.0);
            pv[6] = mn(mx(pv[6], 0.0), 1.0);
        } else {
            int o[3];
            double n[3], ps[NPV];
            surface_of(s, i, j, k, o);
            outward_of(S, B, s.face, o, n);
            long gs = gidx(B, o[0], o[1], o[2]);
            for (int q = 0; q < NPV; q++) if (q < npv) ps[q] = S.pv[q * S.NG + gs];
            double rs = S.dc[D_RHO * S.NG + gs], cs = S.dc[D_C2 * S.NG + gs];
            double dl = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double vel = (ps[1] * n[0] + ps[2] * n[1] + ps[3] * n[2]) * dl;
            double a = sqrt(sndp2(P, cs, vel * vel));
            int mode = (s.fn == FN_INSUB) ? 0 : 1;
            if (s.fn == FN_FARFIELD) {
                double mach = vel / sqrt(cs);
                if (mach >= 0.0) mode = (mach >= 1.0) ? 2 : 1;
                else mode = (mach <= -1.0) ? 3 : 0;
            }
            if (mode == 0) {
                double var = press - rp[0] + ps[0]
                           - rs * a * (n[0] * (rp[1] - ps[1]) + n[1] * (rp[2] - ps[2]) + n[2] * (rp[3] - ps[3])) * dl;
                pv[0] = var - pv[0];
                pv[1] = 2.0 * (rp[1] + n[0] * 0.5 * var / rs / a * dl) - pv[1];
                pv[2] = 2.0 * (rp[2] + n[1] * 0.5 * var / rs / a * dl) - pv[2];
                pv[3] = 2.0 * (rp[3] + n[2] * 0.5 * var / rs / a * dl) - pv[3];
                for (int q = 4; q < NPV; q++) if (q < npv) pv[q] = 2.0 * rp[q] - pv[q];
                pv[5] = mn(mx(pv[5], 0.0), 1.0);
                pv[6] = mn(mx(pv[6], 0.0), 1.0);
            } else if (mode == 1) {
                double dp = ps[0] + rp[0] - press;
                pv[0] = 2.0 * (press - rp[0]) - pv[0];
                pv[1] = 2.0 * (ps[1] + n[0] * dp / rs / a * dl) - pv[1];
                pv[2] = 2.0 * (ps[2] + n[1] * dp / rs / a * dl) - pv[2];
                pv[3] = 2.0 * (ps[3] + n[2] * dp / rs / a * dl) - pv[3];
            } else if (mode == 3) {
                pv[0] = 2.0 * (press - rp[0]) - pv[0];
                for (int q = 1; q < NPV; q++) if (q < npv) pv[q] = 2.0 * rp[q] - pv[q];
                pv[5] = mn(mx(pv[5], 0.0), 1.0);
                pv[6] = mn(mx(pv[6], 0.0), 1.0);
            }
        }
        double mu_keep[2] = {d[D_MU], d[D_KAP]};
        mixture_eos<double>(pv[0] + rp[0], pv[4], pv[5], pv[6], d, P.iturb != -3);
        if (P.iturb == -3) { d[D_MU] = mu_keep[0]; d[D_KAP] = mu_keep[1]; }
        put_cell(S, gg, pv, d, mut);
        break;
    }
    case FN_COPY:
        put_cell(S, gg, pv, d, mut);
        break;
    case FN_EDGEWW: case FN_EDGEWWKW: case FN_CORNERWW: case FN_CORNERWWKW: {
        const bool corner = (s.fn == FN_CORNERWW || s.fn == FN_CORNERWWKW);
        const int nn = corner ? 6 : 3;
        double pp[NPV];
        for (int q = 0; q < NPV; q++) if (q < npv) pp[q] = -pv[q];
        for (int t = 0; t < nn; t++) {
            const int *nb = s.nb[t];
            if (!corner && nb[0] == 0 && nb[1] == 0 && nb[2] == 0) continue;
            long gn = gidx(B, nb[0] + s.dir[0] * i, nb[1] + s.dir[1] * j, nb[2] + s.dir[2] * k);
            for (int q = 0; q < NPV; q++) if (q < npv) pp[q] = pp[q] - S.pv[q * S.NG + gn];
        }
        pv[1] = pp[1]; pv[2] = pp[2]; pv[3] = pp[3];
        if (s.fn == FN_EDGEWW || s.fn == FN_CORNERWW) {
            if (npv > 7) { pv[7] = pp[7]; pv[8] = pp[8]; }
        } else {
            double yd = S.ydn[widx(B, m[0], m[1], m[2])];
            double var = 800.0 * d[D_MU] / d[D_RHO] / (yd * yd);
            pv[7] = pp[7];
            pv[8] = (corner ? 8.0 : 4.0) * var + pp[8];
        }
        put_cell(S, gg, pv, d, mut);
        break;
    }
    case FN_EDGENOWALL: case FN_CORNERNOWALL: {
        const bool corner = (s.fn == FN_CORNERNOWALL);
        const int nn = corner ? 6 : 3;
        double mt = mut;
        for (int t = 0; t < nn; t++) {
            const int *nb = s.nb[t];
            if (!corner && nb[0] == 0 && nb[1] == 0 && nb[2] == 0) continue;
            long gn = gidx(B, nb[0] + s.dir[0] * i, nb[1] + s.dir[1] * j, nb[2] + s.dir[2] * k);
            for (int q = 0; q < NPV; q++) if (q < npv) pv[q] = pv[q] + S.pv[q * S.NG + gn];
            if (S.mut) mt = mt + S.mut[gn];
        }
        double f = corner ? 1.0 / 7.0 : 1.0 / 3.0;
        for (int q = 0; q < NPV; q++) if (q < npv) pv[q] = f * pv[q];
        mt = f * mt;
        mixture_eos<double>(pv[0] + rp[0], pv[4], pv[5], pv[6], d, P.iturb != -3);
        put_cell(S, gg, pv, d, mt);
        break;
    }
    default: break;
    }
}

__global__ __launch_bounds__(256) void k_bc(DS S, int which)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NBC[which]) return;
    const Seg *sg = S.seg[which];
    int lo = 0, hi = S.nbcseg[which] - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sg[mid].start <= t) lo = mid; else hi = mid - 1;
    }
    const Seg &s = sg[lo];
    long r = t - s.start;
    int ni = s.ie[0] - s.is[0] + 1, nj = s.ie[1] - s.is[1] + 1;
    int i = s.is[0] + (int)(r % ni);
    int j = s.is[1] + (int)((r / ni) % nj);
    int k = s.is[2] + (int)(r / ((long)ni * nj));
    apply_bc_cell(S, s, i, j, k);
}

// interface copy: the sending block's cells to the donor block's ghost cells
__global__ __launch_bounds__(256) void k_halo(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NHALO) return;
    int lo = 0, hi = S.ncon - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (S.con[mid].start <= t) lo = mid; else hi = mid - 1;
    }
    const Con &c = S.con[lo];
    long r = t - c.start;
    int ni = c.ie[0] - c.is[0] + 1, nj = c.ie[1] - c.is[1] + 1;
    int di = (int)(r % ni), dj = (int)((r / ni) % nj), dk = (int)(r / ((long)ni * nj));
    int ii = c.tm[0] * di + c.tm[1] * dj + c.tm[2] * dk + c.isd[0];
    int jj = c.tm[3] * di + c.tm[4] * dj + c.tm[5] * dk + c.isd[1];
    int kk = c.tm[6] * di + c.tm[7] * dj + c.tm[8] * dk + c.isd[2];
    long gs = gidx(S.b[c.src], c.is[0] + di, c.is[1] + dj, c.is[2] + dk);
    long gd = gidx(S.b[c.dst], ii, jj, kk);
    for (int n = 0; n < NPV; n++) if (n < S.npv) S.pv[n * S.NG + gd] = S.pv[n * S.NG + gs];
    for (int m = 0; m < ND; m++) S.dc[m * S.NG + gd] = S.dc[m * S.NG + gs];
    if (S.mut) S.mut[gd] = S.mut[gs];
}

// ------------------------------------------------------------ face fluxes
// Direction D: normal axis D, first transverse axis E=(D+1)%3, second T=(D+2)%3,
// matching the reference's stencil orderings (src/rhs_module.f90:288-312,
// 374-398, 460-484): ll = 1 + dn + 2(de+1) + 6(dt+1) for dn in {0,1}.
// face of direction D handled by thread t: its (i,j,k) (the left cell) and block
template <int D>
__device__ __forceinline__ const BlkD &face_of(const DS &S, long t, int &i, int &j, int &k)
{
    int b = 0;
    for (int q = 1; q < S.nb; q++) if (t >= S.b[q].f0[D]) b = q;
    const BlkD &B = S.b[b];
    long r = t - B.f0[D];
    if (D == 0) { i = 1 + (int)(r % B.I); j = 2 + (int)((r / B.I) % (B.J - 1)); k = 2 + (int)(r / ((long)B.I * (B.J - 1))); }
    else if (D == 1) { i = 2 + (int)(r % (B.I - 1)); j = 1 + (int)((r / (B.I - 1)) % B.J); k = 2 + (int)(r / ((long)(B.I - 1) * B.J)); }
    else { i = 2 + (int)(r % (B.I - 1)); j = 2 + (int)((r / (B.I - 1)) % (B.J - 1)); k = 1 + (int)(r / ((long)(B.I - 1) * (B.J - 1))); }
    return B;
}

// offset of stencil point ll (Fortran numbering) from the face's left cell
template <int D>
__device__ __forceinline__ long stencil_off(const BlkD &B, int ll)
{
    const long st[3] = {1, (long)B.nx, (long)B.nx * B.ny};
    const long sN = st[D], sE = st[(D + 1) % 3], sT = st[(D + 2) % 3];
    int q = ll - 1;
    if (q < 18) return (long)(q & 1) * sN + (long)((q >> 1) % 3 - 1) * sE + (long)(q / 6 - 1) * sT;
    if (q < 27) { q -= 18; return -sN + (long)(q % 3 - 1) * sE + (long)(q / 3 - 1) * sT; }
    if (q < 36) { q -= 27; return 2 * sN + (long)(q % 3 - 1) * sE + (long)(q / 3 - 1) * sT; }
    return q == 36 ? -2 * sN : 3 * sN;
}

// Inviscid flux of every face of direction D (src/rhs_module.f90:282-337 and
// the j/k analogues): reconstruction, clamps, face EOS, Roe / AUSM+-up.
// SENSOR: the Roe-M / AUSMPW+ instantiation (18 stencil pressures, face
// sound speeds); the other schemes run the lighter one.  R: the arithmetic
// of reconstruction, face EOS and flux -- double, or float for the mixed
// precision mode (the state is read in FP64 and rounded once; gauge
// pressure keeps the low-Mach pressure differences; the metastable-vapour
// switch compares against the FP64 saturation pressure).
template <int D, bool SENSOR, class R>
__global__ __launch_bounds__(256) void k_face(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NF[D]) return;
    int i, j, k;
    const BlkD &B = face_of<D>(S, t, i, j, k);
    const Par &P = S.P;
    const int npv = S.npv;
    const long g9 = gidx(B, i, j, k);
    const long sN = D == 0 ? 1 : (D == 1 ? (long)B.nx : (long)B.nx * B.ny);
    R c9[NPV], c10[NPV], xl[NPV], xr[NPV];
    R g3[6][3];
    const double *f0 = S.pv + g9;
    Par Pi = P;
    Pi.iturb = -3;          // no viscous differences here
    face_recon<R>(Pi, npv, [&](int ll, int n) -> R { return R(f0[(long)n * S.NG + stencil_off<D>(B, ll)]); },
                  c9, c10, xl, xr, g3);
    const long gq = g9 + sN;
    double n3d[3];
    face_vec(S, B, D, i, j, k, n3d);
    const R n3[3] = {R(n3d[0]), R(n3d[1]), R(n3d[2])};
    R rh9[3] = {R(S.dc[D_RHO * S.NG + g9]), R(S.dc[D_H * S.NG + g9]), R(0)};
    R rh10[3] = {R(S.dc[D_RHO * S.NG + gq]), R(S.dc[D_H * S.NG + gq]), R(0)};
    R fx[NPV];
    if constexpr (SENSOR) {
        // the pressure sensor of Roe-M / AUSMPW+: x(1:18,1) of the face stencil
        R sd[18];
        for (int ll = 1; ll <= 18; ll++) sd[ll - 1] = R(f0[stencil_off<D>(B, ll)]);
        rh9[2] = R(S.dc[D_C2 * S.NG + g9]);
        rh10[2] = R(S.dc[D_C2 * S.NG + gq]);
        inviscid_face<R, true>(P, npv, n3, c9, c10, xl, xr, rh9, rh10, fx, sd);
    } else {
        inviscid_face<R, false>(P, npv, n3, c9, c10, xl, xr, rh9, rh10, fx);
    }
    double *out = S.ff[D];
    const long NF = S.NF[D];
    for (int n = 0; n < NPV; n++) if (n < npv) out[n * NF + t] = double(fx[n]);
}

// Viscous flux of every face of direction D (src/vsflux_module.f90:182-398),
// with the metric averaging of src/rhs_module.f90:340-357.  R as k_face.
template <int D, class R>
__global__ __launch_bounds__(256) void k_vface(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NF[D]) return;
    int i, j, k;
    const BlkD &B = face_of<D>(S, t, i, j, k);
    const Par &P = S.P;
    const int npv = S.npv;
    const long g9 = gidx(B, i, j, k);
    const long sN = D == 0 ? 1 : (D == 1 ? (long)B.nx : (long)B.nx * B.ny);
    R g3[6][3] = {};
    R c9[NPV], c10[NPV];
#pragma unroll
    for (int n = 0; n < NPV; n++) {
        c9[n] = c10[n] = R(0);
        if (n >= npv) continue;
        const double *f = S.pv + (long)n * S.NG + g9;
        auto xs = [&](int ll) -> R { return R(f[stencil_off<D>(B, ll)]); };
        c9[n] = R(f[0]);
        c10[n] = R(f[sN]);
        if ((n >= 1 && n <= 4) || n >= 7) {
            const int q = (n <= 4) ? n - 1 : n - 3;     // u, v, w, T, k, omega
            g3[q][0] = xs(10) - xs(9);
            g3[q][1] = R(0.0625) * (-xs(1) - xs(2) + xs(5) + xs(6) + R(2) * (-xs(7) - xs(8) + xs(11) + xs(12))
                                    - xs(13) - xs(14) + xs(17) + xs(18));
            g3[q][2] = R(0.0625) * (-xs(1) - xs(2) + xs(13) + xs(14) + R(2) * (-xs(3) - xs(4) + xs(15) + xs(16))
                                    - xs(5) - xs(6) + xs(17) + xs(18));
        }
    }
    int ip = i, jp = j, kp = k;
    if (D == 0) ip++; else if (D == 1) jp++; else kp++;
    const long gq = g9 + sN;
    double n3d[3];
    face_vec(S, B, D, i, j, k, n3d);
    double e4d[4][3], t4d[4][3];
    const int E = (D + 1) % 3, T = (D + 2) % 3;
    int p0[3] = {i, j, k}, p1[3] = {ip, jp, kp};
    int a0[3], a1[3];
    // e-family: (P,e-1), (P+1,e-1), (P,e), (P+1,e)
    for (int q = 0; q < 3; q++) { a0[q] = p0[q]; a1[q] = p1[q]; }
    a0[E]--; a1[E]--;
    face_vec(S, B, E, a0[0], a0[1], a0[2], e4d[0]);
    face_vec(S, B, E, a1[0], a1[1], a1[2], e4d[1]);
    face_vec(S, B, E, p0[0], p0[1], p0[2], e4d[2]);
    face_vec(S, B, E, p1[0], p1[1], p1[2], e4d[3]);
    // t-family: i faces (P,t-1),(P+1,t-1),(P,t),(P+1,t); j and k faces (P,t-1),(P,t),(P+1,t-1),(P+1,t)
    for (int q = 0; q < 3; q++) { a0[q] = p0[q]; a1[q] = p1[q]; }
    a0[T]--; a1[T]--;
    face_vec(S, B, T, a0[0], a0[1], a0[2], t4d[0]);
    if (D == 0) {
        face_vec(S, B, T, a1[0], a1[1], a1[2], t4d[1]);
        face_vec(S, B, T, p0[0], p0[1], p0[2], t4d[2]);
    } else {
        face_vec(S, B, T, p0[0], p0[1], p0[2], t4d[1]);
        face_vec(S, B, T, a1[0], a1[1], a1[2], t4d[2]);
    }
    face_vec(S, B, T, p1[0], p1[1], p1[2], t4d[3]);
    R n3[3], e4[4][3], t4[4][3];
    for (int q = 0; q < 3; q++) n3[q] = R(n3d[q]);
    for (int a = 0; a < 4; a++)
        for (int q = 0; q < 3; q++) { e4[a][q] = R(e4d[a][q]); t4[a][q] = R(t4d[a][q]); }
    R gl[2], gr[2], dl[ND], dr[ND];
    long wl = widx(B, i, j, k), wr = widx(B, ip, jp, kp);
    gl[0] = R(S.vol[wl]); gr[0] = R(S.vol[wr]);
    gl[1] = S.ydn ? R(S.ydn[wl]) : R(1); gr[1] = S.ydn ? R(S.ydn[wr]) : R(1);
    for (int m = 0; m < ND; m++) { dl[m] = R(S.dc[m * S.NG + g9]); dr[m] = R(S.dc[m * S.NG + gq]); }
    R ml = S.mut ? R(S.mut[g9]) : R(0), mr = S.mut ? R(S.mut[gq]) : R(0);
    R cav[5] = {R(0.5) * (c9[1] + c10[1]), R(0.5) * (c9[2] + c10[2]), R(0.5) * (c9[3] + c10[3]), R(0), R(0)};
    if (npv > 7) { cav[3] = R(0.5) * (c9[7] + c10[7]); cav[4] = R(0.5) * (c9[8] + c10[8]); }
    R fv[NPV];
    visc_flux<R>(P, n3, e4, t4, gl, gr, g3, cav, dl, dr, ml, mr, fv);
    double *out = S.ff[D];
    const long NF = S.NF[D];
    for (int n = 0; n < NPV; n++) if (n < npv) out[(npv + n) * NF + t] = double(fv[n]);
}

__device__ __forceinline__ long fidx(const BlkD &B, int D, int i, int j, int k)
{
    if (D == 0) return B.f0[0] + (i - 1) + (long)B.I * ((j - 2) + (long)(B.J - 1) * (k - 2));
    if (D == 1) return B.f0[1] + (i - 2) + (long)(B.I - 1) * ((j - 1) + (long)B.J * (k - 2));
    return B.f0[2] + (i - 2) + (long)(B.I - 1) * ((j - 2) + (long)(B.J - 1) * (k - 1));
}

// flux differences in the reference's order, then the cell sources
// MODELS: the instantiation with the Kunz / Singhal / Schnerr-Sauer / Lee
// sources (a full EOS per cell) and the DES length scales; Merkle, no
// cavitation and RANS SST run the lighter one.
template <bool MODELS>
__global__ __launch_bounds__(256) void k_combine(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    const Par &P = S.P;
    const int npv = S.npv;
    const bool visc = (P.iturb >= -2);
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    double r[NPV];
    for (int D = 0; D < 3; D++) {
        long f1 = fidx(B, D, i, j, k);
        long f0 = fidx(B, D, i - (D == 0), j - (D == 1), k - (D == 2));
        const double *F = S.ff[D];
        long NF = S.NF[D];
        for (int n = 0; n < NPV; n++) if (n < npv) {
            double a = -(F[n * NF + f1] - F[n * NF + f0]);
            r[n] = (D == 0) ? a : r[n] + a;
            if (visc) r[n] = r[n] + (F[(npv + n) * NF + f1] - F[(npv + n) * NF + f0]);
        }
    }
    long g = gidx(B, i, j, k);
    double d[ND], pc[NPV];
    for (int m = 0; m < ND; m++) d[m] = S.dc[m * S.NG + g];
    for (int n = 0; n < NPV; n++) if (n < npv) pc[n] = S.pv[n * S.NG + g];
    double vol = S.vol[widx(B, i, j, k)];
    if (P.ncav) {
        double rc, rv, ic[4];
        if constexpr (MODELS) {
            if (P.ncav >= 2) {
                double dd[ND], full[18];
                mixture_eos<double>(pc[0] + P.pref, pc[4], pc[5], pc[6], dd, false, full);
                cav_model(P, pc, full, vol, rc, rv, ic);
            } else {
                merkle(P, pc, d, vol, rc, rv, ic);
            }
        } else {
            merkle(P, pc, d, vol, rc, rv, ic);
        }
        r[5] = r[5] + rv - rc;
        for (int q = 0; q < 4; q++) S.icav[q * S.NC + t] = ic[q];
        S.rcav[t] = rc;
        S.rcav[S.NC + t] = rv;
    }
    if (P.iturb == 0) {
        double x7[7][NPV], f6[6][3], src[2], it[4], oc;
        const int di[7] = {-1, 0, 1, 0, 0, 0, 0}, dj[7] = {0, 0, 0, -1, 1, 0, 0}, dk[7] = {0, 0, 0, 0, 0, -1, 1};
        for (int q = 0; q < 7; q++) {
            long gn = gidx(B, i + di[q], j + dj[q], k + dk[q]);
            for (int n = 0; n < NPV; n++) if (n < npv) x7[q][n] = S.pv[n * S.NG + gn];
        }
        cell_faces(S, B, i, j, k, f6);
        sst_source<MODELS>(P, x7, f6, vol, S.ydn[widx(B, i, j, k)], d, S.mut[g], src, it, oc);
        S.ocut[t] = oc;
        r[7] = r[7] + src[0];
        r[8] = r[8] + src[1];
        for (int q = 0; q < 4; q++) S.itt[q * S.NC + t] = it[q];
    }
    if (P.nsteady > 0) {
        double fx[NPV];
        fx[0] = d[D_RHO];
        fx[1] = d[D_RHO] * pc[1];
        fx[2] = d[D_RHO] * pc[2];
        fx[3] = d[D_RHO] * pc[3];
        fx[4] = d[D_RHO] * (d[D_H] + 0.5 * (pc[1] * pc[1] + pc[2] * pc[2] + pc[3] * pc[3])) - (pc[0] + P.pref);
        for (int n = 5; n < NPV; n++) if (n < npv) fx[n] = d[D_RHO] * pc[n];
        for (int n = 0; n < NPV; n++) if (n < npv) {
            double q1 = S.qq[(long)n * S.NC + t], q2 = S.qq[(long)(npv + n) * S.NC + t];
            double v = (double)P.nsteady * (P.bdf_b * fx[n] - q1) + 0.5 * ((double)P.nsteady - 1.0) * q2;
            r[n] = r[n] - v * vol / P.dt_phy;
        }
    }
    for (int n = 0; n < NPV; n++) if (n < npv) S.res[n * S.NC + t] = r[n];
}

// ------------------------------------------------------------ LU-SGS sweeps
// One hyperplane i+j+k = s of block b.  Thread -> (j,k); i = s - j - k.
// Per-cell terms of the LU-SGS operator that do not depend on the sweep:
// the diagonal-inverse coefficients and the Jacobian's cell and face terms
// (src/lhs_module.f90:346-457, src/jacobian_module.f90:182-403).
__global__ __launch_bounds__(128) void k_lhs_prep(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    const Par &P = S.P;
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    CellJ cc;
    load_cell(S, B, i, j, k, cc);
    double f6[6][3];
    cell_faces(S, B, i, j, k, f6);
    double ic[4] = {0, 0, 0, 0}, it[4] = {0, 0, 0, 0};
    if (P.ncav) for (int m = 0; m < 4; m++) ic[m] = S.icav[m * S.NC + t];
    if (P.iturb == 0) for (int m = 0; m < 4; m++) it[m] = S.itt[m * S.NC + t];
    double g[NGX];
    getx_prep(P, S.npv, cc, f6, S.dt[t], ic, it, g);
    for (int q = 0; q < NGX; q++) S.gx[q * S.NC + t] = g[q];
    JacCell jc = jac_cell(P, cc);
    S.jc[t] = jc.beta;
    S.jc[S.NC + t] = jc.h;
    for (int q = 0; q < 6; q++) {
        JacFace f = jac_face(P, cc, f6[q]);
        S.jf[(3 * q) * S.NC + t] = f.u;
        S.jf[(3 * q + 1) * S.NC + t] = f.lam;
        S.jf[(3 * q + 2) * S.NC + t] = f.vis;
    }
}

// ---------------------------------------------------------------- LU-SGS sweeps
// The forward (backward) sweep, src/update_module.f90:486-670:
//   dcv = res + A+(i-1) dq(i-1) + A+(j-1) dq(j-1) + A+(k-1) dq(k-1),  dq = D^-1 dcv
//   (dcv - A-(i+1) dq(i+1) - ... backward), ghost dq = 0.
// Work is split into pencils along i: a 16-lane group owns pencil (j,k) and
// walks i in sweep order.  Lane n accumulates row n of each Jacobian product
// in the reference's order, then the rows are exchanged and lane n evaluates
// component n of D^-1.  dq of a neighbour in another workgroup is published
// with a device-coherent store and read with device-coherent loads; an entry
// still holding the "not yet written" bit pattern is polled again.  Every
// poll is bounded: past the limit the kernel records an error and stops
// waiting, so each wavefront finishes.
constexpr unsigned long long DQ_UNSET = 0x7FF4A5A5A5A5A5A5ull;   // a signalling NaN no arithmetic produces
constexpr long POLL_LIMIT = 1L << 19;

__device__ __forceinline__ double ld_dev(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tiled sweep: a workgroup owns a tile of TJ x TK pencils of one block and
// walks its hyperplanes s = i + j + k in sweep order with one barrier per
// hyperplane.  A pencil's neighbour inside the tile hands over its dq through
// LDS (double-buffered by the parity of s); a neighbour in another tile
// through a device-coherent exchange array that holds a cell's values side by
// side (one write of 72 bytes, not nine 8-byte writes to nine lines), waited
// for with a bounded poll.
// Everything a cell reads that does not depend on the sweep (its residual,
// state, properties, diagonal-inverse terms, and the Jacobian operands of its
// j and k neighbours) is moved ahead of use into an LDS stage, so a step's
// critical path is the dq hand-over and the arithmetic.
// Tiles wait only on tiles upstream in j and k, every tile is resident
// (cooperative launch) and a workgroup takes its tiles in topological order,
// so the chain always advances.
constexpr int TJ = 4, TK = 4;
// LDS stage of one pencil-step (doubles)
enum {
    ST_RES = 0, ST_PV = 9, ST_DC = 18, ST_N3 = 31, ST_JC = 34, ST_JF = 36, ST_GX = 39,
    ST_J = 68, ST_K = 98,                  // neighbour blocks: pv 9, dc 13, n3 3, jc 2, jf 3
    NB_PV = 0, NB_DC = 9, NB_N3 = 22, NB_JC = 25, NB_JF = 27, ST_N = 128
};
static_assert(ST_GX + NGX == ST_J, "stage layout");

__device__ __forceinline__ long face_off(const BlkD &B, int ax, const int *p)
{
    if (ax == 0) return B.cx0 + (p[0] - 1) + (long)B.I * ((p[1] - 1) + (long)(B.J + 1) * (p[2] - 1));
    if (ax == 1) return B.ex0 + (p[0] - 1) + (long)(B.I + 1) * ((p[1] - 1) + (long)B.J * (p[2] - 1));
    return B.tx0 + (p[0] - 1) + (long)(B.I + 1) * ((p[1] - 1) + (long)(B.J + 1) * (p[2] - 1));
}
__device__ __forceinline__ const double *face_base(const DS &S, const BlkD &B, int ax, const int *p, int comp)
{
    const long o = face_off(B, ax, p);
    if (ax == 0) return S.cx + comp * S.NCX + o;
    if (ax == 1) return S.ex + comp * S.NEX + o;
    return S.tx + comp * S.NTX + o;
}

// Stage item e of pencil (j,k) of a sweep: every field is stored i fastest,
// so the item of cell i is base[i]; nullptr where the item is not read.
template <bool FWD>
__device__ __forceinline__ const double *stage_base(const DS &S, const BlkD &B, int j, int k, int e)
{
    const int sg = FWD ? 1 : -1, side = FWD ? 1 : 0, npv = S.npv;
    const long c = cidx(B, 0, j, k), g = gidx(B, 0, j, k);
    if (e < ST_PV) return e < npv ? (FWD ? S.res + e * S.NC + c : S.dcv + e * S.NW + widx(B, 0, j, k)) : nullptr;
    if (e < ST_DC) return (e - ST_PV) < npv ? S.pv + (e - ST_PV) * S.NG + g : nullptr;
    if (e < ST_N3) return S.dc + (e - ST_DC) * S.NG + g;
    if (e < ST_JC) { const int p[3] = {FWD ? 0 : -1, j, k}; return face_base(S, B, 0, p, e - ST_N3); }
    if (e < ST_JF) return S.jc + (e - ST_JC) * S.NC + c;
    if (e < ST_GX) return S.jf + (3 * side + e - ST_JF) * S.NC + c;
    if (e < ST_J) return S.gx + (e - ST_GX) * S.NC + c;
    const int ax = e < ST_K ? 1 : 2;
    const int r = e - (ax == 1 ? ST_J : ST_K);
    int m[3] = {0, j, k};
    m[ax] -= sg;
    const int lim = ax == 1 ? B.J : B.K;
    if (m[ax] < 2 || m[ax] > lim) return nullptr;
    const long gm = gidx(B, 0, m[1], m[2]), cm = cidx(B, 0, m[1], m[2]);
    if (r < NB_DC) return r < npv ? S.pv + r * S.NG + gm : nullptr;
    if (r < NB_N3) return S.dc + (r - NB_DC) * S.NG + gm;
    if (r < NB_JC) { const int self[3] = {0, j, k}; return face_base(S, B, ax, FWD ? m : self, r - NB_N3); }
    if (r < NB_JF) return S.jc + (r - NB_JC) * S.NC + cm;
    return S.jf + (3 * (2 * ax + side) + r - NB_JF) * S.NC + cm;
}

// row `lane` Jacobian operands of a staged block (pv at b + NB_PV, ...)
template <class A>
__device__ __forceinline__ JacNb stage_jnb(const A &st, int b, int lane, int npv)
{
    JacNb q;
    q.pvn = lane < npv ? st[b + NB_PV + lane] : 0.0;
    q.pv1 = st[b + NB_PV + 1]; q.pv2 = st[b + NB_PV + 2]; q.pv3 = st[b + NB_PV + 3];
    const int d = b + NB_DC;
    q.rho = st[d + D_RHO]; q.rp = st[d + D_RP]; q.rt = st[d + D_RT]; q.ry1 = st[d + D_RY1]; q.ry2 = st[d + D_RY2];
    q.hp = st[d + D_HP]; q.ht = st[d + D_HT]; q.hy1 = st[d + D_HY1]; q.hy2 = st[d + D_HY2];
    q.n3[0] = st[b + NB_N3]; q.n3[1] = st[b + NB_N3 + 1]; q.n3[2] = st[b + NB_N3 + 2];
    q.n3n = (lane >= 1 && lane <= 3) ? st[b + NB_N3 + lane - 1] : 0.0;
    q.jc = JacCell{st[b + NB_JC], st[b + NB_JC + 1]};
    q.jf = JacFace{st[b + NB_JF], st[b + NB_JF + 1], st[b + NB_JF + 2]};
    return q;
}
static_assert(ST_PV + 9 == ST_DC && ST_DC + ND == ST_N3 && ST_N3 + 3 == ST_JC && ST_JC + 2 == ST_JF,
              "the own-cell part of the stage has the neighbour-block layout");

// lane q of the lane's 16-lane row, by a DPP row broadcast (no LDS traffic)
template <int Q>
__device__ __forceinline__ double row_lane(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + Q, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + Q, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// yv[q] = v of lane q of the row, q < 9 (every lane of the wavefront active)
__device__ __forceinline__ void bcast_rows(double v, double *yv)
{
    static_assert(NPV == 9, "one broadcast per row of the system");
    yv[0] = row_lane<0>(v); yv[1] = row_lane<1>(v); yv[2] = row_lane<2>(v);
    yv[3] = row_lane<3>(v); yv[4] = row_lane<4>(v); yv[5] = row_lane<5>(v);
    yv[6] = row_lane<6>(v); yv[7] = row_lane<7>(v); yv[8] = row_lane<8>(v);
}

// Workgroup barrier for LDS hand-over: this wavefront's LDS accesses are
// complete (lgkmcnt), nothing else is waited for, so global loads and the
// stage LDS-DMA stay in flight across it (a release fence on LDS would also
// wait for every LDS-DMA).  The asm clobbers keep the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}
// global address space: a pointer read back from LDS would otherwise be
// generic, and generic loads also count against the LDS wait counter
#define GLOBAL __attribute__((address_space(1)))
constexpr int WIN = 4;               // cells of a stage window
// one cell of a stage window: item e at stride WIN
struct WinRow {
    const double *p;
    __device__ double operator[](int e) const { return p[e * WIN]; }
};
constexpr int SWEEP_THREADS = 256 + 64 + 64 + 64;   // 16 pencil groups of 16 lanes, a store, a poll and a loader wavefront
constexpr int NEDGE = 2 * 4 * NPV;   // cross-tile dq values a tile reads per step: two edges of four pencils
// Stage loads, by the loader wavefront, as LDS-DMA (global memory to LDS,
// 16 bytes a lane, no registers).  A pencil at sweep distance d reads window
// d >> 2 (its cell) and window (d + 1) >> 2 (the lookahead); while d & 3 is
// 0, 1 or 2 both are window W = d >> 2 and the other buffer, window W + 1, is
// free.  The loader issues quarters 0 and 1 of window W + 1 at d & 3 = 0 and
// quarters 2 and 3 at d & 3 = 1, and waits for each step's transfers before
// the barrier that ends the next step, so every transfer lands while its
// buffer is free and window W + 1 is complete when it first becomes the
// lookahead (d & 3 = 3).  A window's four cells lie in ascending i in LDS in
// both sweeps, as the transfers copy them.
static_assert(TJ == 4 && TK == 4, "the poll wavefront maps an edge of four pencils");

// Division of labour inside a tile's workgroup:
//  - four pencil wavefronts (16 groups of 16 lanes, lane = row) run the
//    hyperplane steps; they read only LDS (their one global access is the
//    store of a downstream-edge pencil's dq), so nothing they wait for is
//    slower than LDS;
//  - a poll wavefront reads, with device-coherent loads, the dq of the
//    upstream cells in other tiles during the step before its use and hands
//    it over through LDS (it is the only wavefront that waits on other tiles);
//  - a store wavefront writes each hyperplane's dcv (forward) or dq
//    (backward) one step later; the dq of the pencils on the tile's
//    downstream edges the pencil wavefronts store as soon as they form it;
//  - a loader wavefront fills the stage windows ahead of the pencils.
// A step's Jacobian coefficients are formed during the step before from its
// stage without branching on the row, so the dependent chain of a step is
// three row products, the exchange of y and D^-1.
template <bool FWD>
__global__ __launch_bounds__(SWEEP_THREADS) void k_sweep_tile(DS S, const int *tiles, int ntiles, int *err)
{
    // dq hand-over, one LDS object (so that the compiler can tell its reads
    // from the stage LDS-DMA): [0, HX) the dq of the hyperplane by the parity
    // of s, [HX, HZ) the dq of the j / k neighbour in another tile by the
    // parity of s, [HZ, HZ + NPV) zeros for a ghost neighbour
    constexpr int HX = 2 * TJ * TK * NPV, HZ = HX + 2 * 2 * TJ * TK * NPV;
    __shared__ double hand[HZ + NPV];
    auto lds_at = [](int par, int g) { return (par * TJ * TK + g) * NPV; };
    auto xch_at = [](int par, int a, int g) { return HX + ((par * 2 + a) * TJ * TK + g) * NPV; };
    auto lds = [&](int par, int g) { return hand + lds_at(par, g); };
    auto xch = [&](int par, int a, int g) { return hand + xch_at(par, a, g); };
    double *const zrow = hand + HZ;
    // dcv of the last YR hyperplanes (forward sweep), by s mod YR: the store
    // wavefront writes a pencil's dcv four cells at a time
    constexpr int YR = 5;
    __shared__ double ylds[FWD ? YR : 1][TJ * TK][NPV];
    __shared__ double win[2][TJ * TK][ST_N][WIN];   // stage windows (see the pencil wavefronts)
    __shared__ const double *src[TJ * TK][ST_N];
    __shared__ int gave_up;
    const int npv = S.npv;
    const int sg = FWD ? 1 : -1;
    const double sgn = FWD ? 1.0 : -1.0;
    double *const xq = FWD ? S.xqf : S.xqb;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 15, grp = (threadIdx.x >> 4) & 15;
    constexpr int PER = ST_N / 16;                                  // stage items per lane
    if (threadIdx.x == 0) gave_up = 0;
    if (threadIdx.x < NPV) zrow[threadIdx.x] = 0.0;
    for (int tq = blockIdx.x; tq < ntiles; tq += gridDim.x) {
        const int *td = tiles + 4 * (FWD ? tq : ntiles - 1 - tq);
        const BlkD B = S.b[td[0]];           // a copy: the loops read the block's extents from registers
        const long NW = S.NW;
        const int j0 = td[1], k0 = td[2];
        const int jend = min(j0 + TJ - 1, B.J), kend = min(k0 + TK - 1, B.K);
        // the tile's hyperplane range in sweep order
        const int sa = FWD ? 2 + j0 + k0 : B.I + jend + kend;
        const int sb = FWD ? B.I + jend + kend : 2 + j0 + k0;
        // the pencils whose dq the downstream tiles read (j + 1 and k + 1 edges forward)
        auto down_edge = [](int g) {
            return FWD ? (g % TJ == TJ - 1 || g / TJ == TK - 1) : (g % TJ == 0 || g / TJ == 0);
        };
        auto active = [&](int g, int ii) {
            const int jj = j0 + g % TJ, kk = k0 + g / TJ;
            return jj <= jend && kk <= kend && ii - jj - kk >= 2 && ii - jj - kk <= B.I;
        };
        if (wave == 4) {
            // store wavefront: hyperplane s is written out during step s+1,
            // while the pencil wavefronts compute.  The dq that other tiles
            // wait for (the pencils on the tile's downstream edges) the
            // pencil wavefronts store themselves into the exchange array, as
            // soon as it is formed.  Forward, dq is read by nothing else (the
            // backward sweep starts from dcv), so only dcv is written, a
            // pencil's four cells of a stage window (or its last few)
            // together and non-temporal (the backward sweep reads them after
            // this launch); backward, the exchange array also takes the other
            // pencils' dq, the update's input.
            auto publish = [&](int s) {
                const int pb = s & 1;
                if constexpr (FWD) {
                    // the four pencils whose window closes at hyperplane s
                    // (j + k = s - 5 mod 4) write its four cells, formed at
                    // hyperplanes s - 3 .. s; lane -> (pencil, row, cell), so
                    // that four adjacent lanes write one row's 32 bytes
                    const int r0 = (s - j0 - k0 - 5) & 3;
                    for (int v = threadIdx.x - 256; v < 4 * NPV * 4; v += 64) {
                        const int p = v / (4 * NPV), n = (v >> 2) % NPV, c = v & 3;
                        // the class's pencil with gj = p: gj + gk = r0, or r0 + 4 past the tile's anti-diagonal
                        const int gj = p, gk = r0 + 4 * (p > r0) - p;
                        const int jj = j0 + gj, kk = k0 + gk, i = s - jj - kk;
                        if (n < npv && jj <= jend && kk <= kend && i >= 5 && i <= B.I)
                            __builtin_nontemporal_store(ylds[(s - 3 + c) % YR][gj + TJ * gk][n], S.dcv + n * NW + widx(B, i - 3 + c, jj, kk));
                    }
                    // a pencil that ends at hyperplane s in the middle of a window
                    if (s >= B.I + j0 + k0) {
                        for (int e = threadIdx.x - 256; e < TJ * TK * NPV; e += 64) {
                            const int g = e / NPV, n = e - g * NPV;
                            const int jj = j0 + g % TJ, kk = k0 + g / TJ, i = s - jj - kk;
                            const int m = (i - 2) & 3;
                            if (n < npv && i == B.I && m != 3 && jj <= jend && kk <= kend) {
                                double *o = S.dcv + n * NW + widx(B, i - m, jj, kk);
                                for (int c = 0; c <= m; c++) __builtin_nontemporal_store(ylds[(s - m + c) % YR][g][n], o + c);
                            }
                        }
                    }
                } else {
                    for (int e = threadIdx.x - 256; e < TJ * TK * NPV; e += 64) {
                        const int g = e / NPV, n = e - g * NPV;
                        if (n < npv && active(g, s) && !down_edge(g)) {
                            const int jj = j0 + g % TJ, kk = k0 + g / TJ, i = s - jj - kk;
                            xq[widx(B, i, jj, kk) * NPV + n] = lds(pb, g)[n];
                        }
                    }
                }
            };
            __syncthreads();
            for (int s = sa;; s += sg) {
                if (s != sa) publish(s - sg);
                lds_barrier();
                if (s == sb) break;
            }
            publish(sb);
            if (gave_up && threadIdx.x == 256) atomicOr(err, 1);
            __syncthreads();
            continue;
        }
        if (wave == 6) {
            // loader wavefront: in a step, the eight pencils with d & 3 = 0
            // or 1 each get two quarters (32 items x 4 cells, 1 KB each) of
            // their next window, one 16-byte transfer a lane per quarter:
            // lane L moves item L / 2, cells 2 (L & 1) and 2 (L & 1) + 1.
            // The pencils' prologue fills window 0 and the source table is
            // theirs, so the loader starts after the first barrier.
            const int l = threadIdx.x - 384, li = l >> 1, lc = 2 * (l & 1);
            const int BI = B.I;
            __syncthreads();
            int d0 = FWD ? sa - j0 - k0 - 2 : BI - (sa - j0 - k0);      // distance of pencil 0 (it grows by one a step)
            for (int s = sa;; s += sg) {
                // the transfers are issued at once (issuing sixteen of them
                // takes longer than the poll wavefront's two reads; holding
                // them back behind those reads cost 10 us a sweep)
                if (s != sb) {
                    // the eight pencils with d & 3 = e in {0, 1}: four per
                    // class, pencil p of a class at gj = p on the class's
                    // anti-diagonals t = c, c + 4 (no branch between them, so
                    // the pointer reads can be issued ahead of the transfers
                    // that use them, not each waited for in turn)
                    const double *pp[2][4][2];
                    int gg[2][4], ww[2][4];
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        const int c = (FWD ? d0 - e : e - d0) & 3;
#pragma unroll
                        for (int p = 0; p < 4; p++) {
                            const int t = c + 4 * (p > c), g = p + TJ * (t - p);
                            const int d = FWD ? d0 - t : d0 + t;
                            gg[e][p] = g;
                            ww[e][p] = (d >> 2) + 1;
#pragma unroll
                            for (int h = 0; h < 2; h++) pp[e][p][h] = src[g][32 * (2 * e + h) + li];
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 2; e++) {
#pragma unroll
                        for (int p = 0; p < 4; p++) {
                            const int g = gg[e][p], w1 = ww[e][p];
                            // first cell of the lane's pair; a pair that lies
                            // beyond the pencil is clamped to cells 1..I+1 (its
                            // slots hold no cell of the pencil)
                            const int c0 = min(max((FWD ? 2 + 4 * w1 : BI - 4 * w1 - 3) + lc, 1), BI);
#pragma unroll
                            for (int h = 0; h < 2; h++) {
                                const int q = 2 * e + h;
                                __builtin_amdgcn_global_load_lds((const GLOBAL void *)(pp[e][p][h] + c0),
                                                                 (__attribute__((address_space(3))) void *)&win[w1 & 1][g][32 * q][0],
                                                                 16, 0, 0);
                            }
                        }
                    }
                }
                d0++;
                // the transfers of the previous step have landed (16 issued since)
                asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
                lds_barrier();
                if (s == sb) break;
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            continue;
        }
        if (wave == 5) {
            // poll wavefront: item q = lane, lane + 64 of the 72 edge values
            // (edge a = q / 36: the j (a = 0) or k (a = 1) neighbours; pencil
            // e = q / 9 % 4 along the edge; row n = q % 9)
            const int l64 = threadIdx.x - 320;
            const double *dummy = xq + widx(B, min(max(sa - j0 - k0, 2), B.I), j0, k0) * NPV;
            auto item = [&](int q, int s, bool &need) -> const double * {
                const int a = q / (4 * NPV), e = (q / NPV) % 4, n = q % NPV;
                const int g = a == 0 ? e * TJ + (FWD ? 0 : TJ - 1) : (FWD ? 0 : TK - 1) * TJ + e;
                const int jj = j0 + g % TJ, kk = k0 + g / TJ;
                const int jn = a == 0 ? jj - sg : jj, kn = a == 0 ? kk : kk - sg;
                need = q < NEDGE && n < npv && active(g, s) && jn >= 2 && jn <= B.J && kn >= 2 && kn <= B.K;
                return need ? xq + widx(B, s - jj - kk, jn, kn) * NPV + n : dummy;
            };
            bool n0, n1;
            const double *p0, *p1;
            double v0, v1;
            auto issue = [&](int s) {            // loads for step s (unconditional: the count of loads never varies)
                p0 = item(l64, s, n0);
                p1 = item(l64 + 64, s, n1);
                v0 = ld_dev(p0);
                v1 = ld_dev(p1);
            };
            auto deliver = [&](int s) {          // wait (bounded) for step s's values, then hand them over
                auto unset = [](bool nd, double v) { return nd && (unsigned long long)__double_as_longlong(v) == DQ_UNSET; };
                if (__ballot(unset(n0, v0) || unset(n1, v1))) {
                    long polls = 0;
                    do {
                        // another tile's give-up is read now and then, not on every poll
                        if (++polls > POLL_LIMIT ||
                            ((polls & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            gave_up = 1;
                            v0 = v1 = 0.0;
                            break;
                        }
                        if (unset(n0, v0)) v0 = ld_dev(p0);
                        if (unset(n1, v1)) v1 = ld_dev(p1);
                    } while (__ballot(unset(n0, v0) || unset(n1, v1)));
                }
                const int par = s & 1;
                for (int t = 0; t < 2; t++) {
                    const int q = l64 + 64 * t;
                    if (q < NEDGE) {
                        const int a = q / (4 * NPV), e = (q / NPV) % 4, n = q % NPV;
                        const int g = a == 0 ? e * TJ + (FWD ? 0 : TJ - 1) : (FWD ? 0 : TK - 1) * TJ + e;
                        xch(par, a, g)[n] = (t == 0 ? n0 : n1) ? (t == 0 ? v0 : v1) : 0.0;
                    }
                }
            };
            issue(sa);
            deliver(sa);
            __syncthreads();
            for (int s = sa;; s += sg) {
                // the next step's values are read at the start of this step,
                // while the pencils compute, as late as they can be
                if (s != sb) issue(s + sg);
                if (s != sb) deliver(s + sg);
                lds_barrier();
                if (s == sb) break;
            }
            __syncthreads();
            continue;
        }
        // pencil wavefronts
        const int gj = grp % TJ, gk = grp / TJ;
        const bool mine = lane < npv;
        const int ln = lane < NPV ? lane : 0;
        const int j = j0 + gj, k = k0 + gk;
        const bool live = (j <= jend && k <= kend);
        const bool edge = down_edge(grp);
        const int BI = B.I;
        const int jm = j - sg, km = k - sg;
        const bool lj = jm >= j0 && jm <= jend, lk = km >= k0 && km <= kend;
        const bool nj = jm >= 2 && jm <= B.J, nk = km >= 2 && km <= B.K;   // neighbour inside the block
        // where the stage items come from: item e of cell i is src[grp][e][i]
        // (the zero array for items a pencil does not read)
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const double *q = live ? stage_base<FWD>(S, B, j, k, lane + 16 * r) : nullptr;
            src[grp][lane + 16 * r] = q ? q : S.zero;
        }
        // Stage windows: window m holds cells at sweep distance 4m..4m+3 of
        // the pencil (distance d = i - 2 forward, I - i backward), item-major
        // with the four cells adjacent, in LDS buffer m & 1.  Four adjacent
        // lanes load one item's four cells (one cache line between them).
        // The pencils fill window 0 here; the loader wavefront fills the
        // rest during the steps.
        const int l4 = lane >> 2, c4 = lane & 3;
        auto wcell = [&](int m, int c) { return FWD ? 2 + 4 * m + c : BI - 4 * m - 3 + c; };
        auto slot = [](int dd) { return FWD ? (dd & 3) : 3 - (dd & 3); };     // the cell at distance dd in its window
        auto fetchq = [&](int m, int qq, double *v) {
            const int ic = min(max(wcell(m, c4), 2), BI);
#pragma unroll
            for (int r = 0; r < 8; r++) v[r] = ((const GLOBAL double *)src[grp][32 * qq + 4 * r + l4])[ic];
        };
        auto putq = [&](int m, int qq, const double *v) {
#pragma unroll
            for (int r = 0; r < 8; r++) win[m & 1][grp][32 * qq + 4 * r + l4][c4] = v[r];
        };
        // prologue: window 0 complete, window -1's buffer zeroed (steps before
        // the pencil's first cell read it)
        for (int qq = 0; qq < 4; qq++) {
            double v[8];
            fetchq(0, qq, v);
            putq(0, qq, v);
        }
        for (int e = lane; e < ST_N * WIN; e += 16) (&win[1][grp][0][0])[e] = 0.0;
        const int d0 = FWD ? sa - j - k - 2 : BI - (sa - j - k);
        if (lane < NPV) lds((sa - sg) & 1, grp)[lane] = 0.0;       // the first step's upstream reads zeros
        double ci[8], cj[8], ck[8];                                  // coefficients of the coming step
#pragma unroll
        for (int q = 0; q < 8; q++) ci[q] = 0.0;
        {
            const WinRow s0{&win[(d0 >> 2) & 1][grp][0][slot(d0)]};
            jac_row_coef(npv, stage_jnb(s0, ST_J, lane, npv), sg, lane, sgn, cj);
            jac_row_coef(npv, stage_jnb(s0, ST_K, lane, npv), sg, lane, sgn, ck);
        }
        __syncthreads();
        for (int s = sa;; s += sg) {
            const int i = s - j - k;
            const int d = FWD ? i - 2 : BI - i;
            const bool act = live && i >= 2 && i <= BI;
            // dq of the three upstream neighbours: in the tile from the
            // previous hyperplane, from another tile through the poll wavefront
            const WinRow st{&win[(d >> 2) & 1][grp][0][slot(d)]};
            const int pb = (s - sg) & 1;
            // (indices into the hand-over object, not pointers chosen among
            // several, so the compiler sees these reads cannot touch the stage)
            const int oi = lds_at(pb, grp);
            const int oj = lj ? lds_at(pb, grp - sg) : (nj ? xch_at(s & 1, 0, grp) : HZ);
            const int ok = lk ? lds_at(pb, grp - sg * TJ) : (nk ? xch_at(s & 1, 1, grp) : HZ);
            double xi[NPV], xj[NPV], xk[NPV];
#pragma unroll
            for (int q = 0; q < NPV; q++) { xi[q] = hand[oi + q]; xj[q] = hand[oj + q]; xk[q] = hand[ok + q]; }
            // dcv = res + the three upstream products, in the reference's i, j, k order
            double y = mine ? st[ST_RES + lane] : 0.0;
            jac_row_apply(ci, xi, hand[oi + ln], y);
            jac_row_apply(cj, xj, hand[oj + ln], y);
            jac_row_apply(ck, xk, hand[ok + ln], y);
            __builtin_amdgcn_sched_barrier(0);
            // dq = D^-1 dcv
            double yv[NPV], pc[NPV];
            bcast_rows(y, yv);
#pragma unroll
            for (int q = 0; q < NPV; q++) pc[q] = st[ST_PV + q];
            const double xn = getx_apply_sel([&](int q) { return st[ST_GX + q]; }, pc, st[ST_DC + D_RY1],
                                             st[ST_DC + D_RT], yv, lane, y, st[ST_PV + ln]);
            if (lane < NPV) {
                lds(s & 1, grp)[lane] = act ? xn : 0.0;
                if constexpr (FWD) ylds[s % YR][grp][lane] = y;
                // the pencil wavefronts have no loads in flight, so this store
                // costs only its issue
                if (edge && act && mine) st_dev(xq + widx(B, i, j, k) * NPV + lane, xn);
            }
            __builtin_amdgcn_sched_barrier(0);
            // coefficients of the next step (independent of dq): this cell's
            // row through its face to the next cell, and the j / k neighbours
            // of the next cell
            jac_row_coef(npv, stage_jnb(st, ST_PV, lane, npv), sg, lane, sgn, ci);
            const WinRow s1{&win[((d + 1) >> 2) & 1][grp][0][slot(d + 1)]};
            jac_row_coef(npv, stage_jnb(s1, ST_J, lane, npv), sg, lane, sgn, cj);
            jac_row_coef(npv, stage_jnb(s1, ST_K, lane, npv), sg, lane, sgn, ck);
            // formed before the barrier (while the poll wavefront may still be
            // waiting), not after it on the next step's path
#pragma unroll
            for (int q = 0; q < 8; q++) asm volatile("" : "+v"(ci[q]), "+v"(cj[q]), "+v"(ck[q]));
            lds_barrier();
            if (s == sb) break;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_fill_bits(double *p, long n, unsigned long long bits)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = __longlong_as_double((long long)bits);
}

// ------------------------------------------------------------ update + norms
// Two wavefronts per SIMD: left to itself the compiler gives the cell's EOS
// 268 registers, one wavefront per SIMD, and the loads go unhidden (159 us
// against 99 us with a small spill, 524288 cells).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_update(DS S)
{
    __shared__ double sred[4][NPV + 1];
    __shared__ long sidx[4];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const Par &P = S.P;
    const int npv = S.npv;
    double dsq[NPV + 1];
    for (int n = 0; n <= NPV; n++) dsq[n] = 0.0;
    double amax = -1.0;
    long aidx = 0x7fffffffffffffffL;
    if (t < S.NC) {
        int bb = find_blk(S, t, 0);
        const BlkD &B = S.b[bb];
        int i, j, k;
        cell_of(B, t - B.c0, i, j, k);
        long g = gidx(B, i, j, k), w = widx(B, i, j, k);
        double p[NPV], old[NPV];
        for (int n = 0; n < NPV; n++) if (n < npv) { old[n] = S.pv[n * S.NG + g]; p[n] = old[n] + S.xqb[w * NPV + n]; }
        p[0] = mx(-P.pref + 1.0e1, p[0]);
        p[5] = mn(1.0, mx(0.0, p[5]));
        p[6] = mn(1.0, mx(0.0, p[6]));
        if (P.iturb == 0) {
            p[7] = mx(p[7], 1.0e-16);
            p[8] = mx(mx(p[8], 1.0e-16), S.ocut[t]);
        }
        for (int n = 0; n < NPV; n++) if (n < npv) S.pv[n * S.NG + g] = p[n];
        double d[ND];
        mixture_eos<double>(p[0] + P.pref, p[4], p[5], p[6], d, P.iturb != -3);
        for (int m = 0; m < ND; m++) S.dc[m * S.NG + g] = d[m];
        for (int n = 0; n < NPV; n++) if (n < npv) { double q = fabs(old[n] - p[n]); dsq[n] = q * q; }
        double q0 = fabs(old[0] - p[0]);
        if (q0 > 1.0e-12) {
            amax = q0;
            // linear order of the reference's scan: block, then k, j, i
            aidx = ((long)bb << 48) | ((long)k << 32) | ((long)j << 16) | (long)i;
        }
    }
    // wavefront reduction
    for (int o = 32; o > 0; o >>= 1) {
        for (int n = 0; n < NPV; n++) if (n < npv) dsq[n] += __shfl_xor(dsq[n], o);
        double om = __shfl_xor(amax, o);
        long oi = __shfl_xor(aidx, o);
        if (om > amax || (om == amax && oi < aidx)) { amax = om; aidx = oi; }
    }
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int n = 0; n < NPV; n++) if (n < npv) sred[wv][n] = dsq[n];
        sred[wv][NPV] = amax;
        sidx[wv] = aidx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc[NPV + 1];
        double bm = sred[0][NPV];
        long bi = sidx[0];
        for (int n = 0; n < NPV; n++) if (n < npv) acc[n] = sred[0][n];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            for (int n = 0; n < NPV; n++) if (n < npv) acc[n] += sred[q][n];
            if (sred[q][NPV] > bm || (sred[q][NPV] == bm && sidx[q] < bi)) { bm = sred[q][NPV]; bi = sidx[q]; }
        }
        double *o = S.red + (long)blockIdx.x * (NPV + 2);
        for (int n = 0; n < NPV; n++) if (n < npv) o[n] = acc[n];
        o[NPV] = bm;
        reinterpret_cast<long *>(o)[NPV + 1] = bi;
    }
}

// fixed-order final reduction of the per-workgroup partials (deterministic)
__global__ __launch_bounds__(256) void k_norm_final(DS S, int nwg)
{
    __shared__ double part[256][NPV];
    __shared__ double pm[256];
    __shared__ long pi[256];
    int tid = threadIdx.x;
    double acc[NPV];
    for (int n = 0; n < NPV; n++) acc[n] = 0.0;
    double bm = -1.0;
    long bi = 0x7fffffffffffffffL;
    for (int q = tid; q < nwg; q += 256) {
        const double *o = S.red + (long)q * (NPV + 2);
        for (int n = 0; n < NPV; n++) if (n < S.npv) acc[n] += o[n];
        long oi = reinterpret_cast<const long *>(o)[NPV + 1];
        if (o[NPV] > bm || (o[NPV] == bm && oi < bi)) { bm = o[NPV]; bi = oi; }
    }
    for (int n = 0; n < NPV; n++) part[tid][n] = acc[n];
    pm[tid] = bm;
    pi[tid] = bi;
    __syncthreads();
    // pairwise tree in a fixed pairing: the same inputs always give the same bits
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            for (int n = 0; n < NPV; n++) part[tid][n] += part[tid + h][n];
            if (pm[tid + h] > pm[tid] || (pm[tid + h] == pm[tid] && pi[tid + h] < pi[tid])) {
                pm[tid] = pm[tid + h];
                pi[tid] = pi[tid + h];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int n = 0; n < NPV; n++) if (n < S.npv) S.scal[4 + n] = sqrt(part[0][n]);
        S.scal[4 + NPV] = pm[0];
        reinterpret_cast<long *>(S.scal)[5 + NPV] = pi[0];
    }
}

__global__ __launch_bounds__(256) void k_eddy(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    double x7[7][NPV], f6[6][3], d[ND];
    const int di[7] = {-1, 0, 1, 0, 0, 0, 0}, dj[7] = {0, 0, 0, -1, 1, 0, 0}, dk[7] = {0, 0, 0, 0, 0, -1, 1};
    for (int q = 0; q < 7; q++) {
        long gn = gidx(B, i + di[q], j + dj[q], k + dk[q]);
        for (int n = 0; n < NPV; n++) if (n < S.npv) x7[q][n] = S.pv[n * S.NG + gn];
    }
    long g = gidx(B, i, j, k);
    for (int m = 0; m < ND; m++) d[m] = S.dc[m * S.NG + g];
    cell_faces(S, B, i, j, k, f6);
    long w = widx(B, i, j, k);
    S.mut[g] = sst_eddy(x7, f6, S.vol[w], S.ydn[w], d);
}

// dual time: qq(2) = qq(1), qq(1) = Q(pv, dv)
__global__ __launch_bounds__(256) void k_unsteady_shift(DS S)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    long g = gidx(B, i, j, k);
    const int npv = S.npv;
    double pc[NPV];
    for (int n = 0; n < NPV; n++) if (n < npv) pc[n] = S.pv[n * S.NG + g];
    double rho = S.dc[D_RHO * S.NG + g], h = S.dc[D_H * S.NG + g];
    for (int n = 0; n < NPV; n++) if (n < npv) S.qq[(long)(npv + n) * S.NC + t] = S.qq[(long)n * S.NC + t];
    double q[NPV];
    q[0] = rho;
    q[1] = rho * pc[1];
    q[2] = rho * pc[2];
    q[3] = rho * pc[3];
    q[4] = rho * (h + 0.5 * (pc[1] * pc[1] + pc[2] * pc[2] + pc[3] * pc[3])) - pc[0] - S.P.pref;
    for (int n = 5; n < NPV; n++) if (n < npv) q[n] = rho * pc[n];
    for (int n = 0; n < NPV; n++) if (n < npv) S.qq[(long)n * S.NC + t] = q[n];
}

// ------------------------------------------------------------ layout transposes
// host AoS (m components fastest) <-> device SoA for one block's ghosted cells
__global__ __launch_bounds__(256) void k_aos_to_soa(const double *src, int m, int msrc, int moff, double *dst, long nstride, long g0, long n)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    for (int q = 0; q < m; q++) dst[q * nstride + g0 + g] = src[g * msrc + moff + q];
}
__global__ __launch_bounds__(256) void k_soa_to_aos(const double *src, int m, long nstride, long g0, double *dst, long n)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    for (int q = 0; q < m; q++) dst[g * m + q] = src[q * nstride + g0 + g];
}
// dv(18) AoS -> compact dc, tv(ntv) AoS -> mu, kappa, mut
__global__ __launch_bounds__(256) void k_dv_to_dc(const double *dv, const double *tv, int ntv, double *dc, double *mut, long nstride, long g0, long n)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int map[11] = {0, 1, 5, 6, 7, 8, 9, 10, 11, 12, 13};
    for (int q = 0; q < 11; q++) dc[q * nstride + g0 + g] = dv[18 * g + map[q]];
    if (ntv >= 2) {
        dc[D_MU * nstride + g0 + g] = tv[ntv * g];
        dc[D_KAP * nstride + g0 + g] = tv[ntv * g + 1];
    }
    if (ntv == 3 && mut) mut[g0 + g] = tv[ntv * g + 2];
}
// interior-cell arrays (c-indexed SoA, m components) -> reference AoS (m,2:imax,...)
__global__ __launch_bounds__(256) void k_cell_to_aos(const double *src, int m, long nstride, long c0, double *dst, long n)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int a = 0; a < m; a++) dst[q * m + a] = src[a * nstride + c0 + q];
}
__global__ __launch_bounds__(256) void k_cell_from_aos(const double *src, int m, long nstride, long c0, double *dst, long n)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int a = 0; a < m; a++) dst[a * nstride + c0 + q] = src[q * m + a];
}
__global__ __launch_bounds__(256) void k_fill(double *p, double v, long n)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ============================================================ host context
struct cav3d_ctx {
    cav3d_params prm;
    Par P;
    DS S;
    int nb = 0, npv = 0, ntv = 0, ngrd = 0;
    std::vector<BlkD> blk;
    std::vector<std::vector<Seg>> segs = std::vector<std::vector<Seg>>(3);
    std::vector<Con> cons;
    std::vector<std::vector<int>> bcdesc;   // per block: nbc*7 raw
    hipStream_t st = nullptr;
    // the residual's j and k directions run on two more streams, forked from
    // and joined back to st by events (default priority, like st)
    hipStream_t side[2] = {};
    hipEvent_t fork = nullptr, join[2] = {};
    int dev = 0;              // device current at creation (the context's device)
    double *stage = nullptr;
    size_t stage_n = 0;
    int nwg_update = 0;
    int coop_blocks = 0;      // grid of the cooperative sweep kernels
    int *tiles = nullptr;     // tile table of the sweeps (device): block, j0, k0, diagonal
    int ntiles = 0;
    int *sweep_err = nullptr; // set when a sweep wait gave up
    bool profile = false;
    hipEvent_t ev[14] = {};
    double prof_ms[10] = {};  // [8] last cav3d_rhs_vjp, [9] last cav3d_calrhs (device ms)
    std::vector<void *> allocs;
    // grid from vertex coordinates (cav3d_set_grid / cav3d_build_metrics)
    std::vector<std::vector<int>> condesc;   // per block: ncon*22 raw
    double *xv = nullptr;     // vertices, SoA [3][NV], block b from v0[b]
    long NV = 0;
    std::vector<long> v0;
    std::vector<char> has_x;
    double *xc = nullptr;     // cell centroids grd(2:4), SoA [3][NW]
    double *xbuf = nullptr;   // interface-exchange staging
    long xbuf_n = 0;
    // adjoint work arrays (allocated by the first cav3d_rhs_vjp)
    bool adj_ready = false;
    double *a_gdb = nullptr, *a_rsb[3] = {}, *a_snb[3] = {};
    double *a_rb = nullptr, *a_fb[3] = {}, *a_db[3] = {}, *a_ub[3] = {}, *a_mtb[3] = {}, *a_cb = nullptr, *a_cdb = nullptr,
           *a_csb = nullptr,
           *a_pb = nullptr, *a_sum = nullptr, *a_save = nullptr, *a_obj = nullptr;
    long a_obj_n = 0;
};

namespace {

// Every device array carries DPAD zeroed bytes before and after it: the
// sweeps' stage transfers move cells in pairs and may read one element past
// either end of an array.
constexpr size_t DPAD = 256;
template <class T> int dalloc(cav3d_ctx *c, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    char *raw = nullptr;
    HIPCHK(hipMalloc((void **)&raw, n * sizeof(T) + 2 * DPAD));
    HIPCHK(hipMemsetAsync(raw, 0, n * sizeof(T) + 2 * DPAD, c->st));
    c->allocs.push_back(raw);
    *p = (T *)(raw + DPAD);
    return 0;
}

int ensure_stage(cav3d_ctx *c, size_t n)
{
    if (c->stage_n >= n) return 0;
    if (c->stage) { HIPCHK(hipStreamSynchronize(c->st)); HIPCHK(hipFree(c->stage)); c->stage = nullptr; }
    HIPCHK(hipMalloc((void **)&c->stage, n * sizeof(double)));
    c->stage_n = n;
    return 0;
}

int wall_fn(int iturb) { return iturb == 0 ? FN_WALLVISKW : (iturb == -2 ? FN_WALLVIS : FN_WALLINV); }
void set3(int *a, int x, int y, int z) { a[0] = x; a[1] = y; a[2] = z; }

// BC construction (src/bc_module.f90:157-336 faces, :502-866 edges, :870-1185 corners)
int build_bcs(cav3d_ctx *c)
{
    const int iturb = c->prm.iturb;
    for (int w = 0; w < 3; w++) c->segs[w].clear();
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        const std::vector<int> &dsc = c->bcdesc[b];
        int nbc = (int)dsc.size() / 7;
        std::vector<Seg> faces;
        for (int n = 0; n < nbc; n++) {
            Seg s;
            memset(&s, 0, sizeof(s));
            s.b = b;
            int kind = dsc[7 * n] & CAV3D_BC_KIND_MASK;
            for (int d = 0; d < 3; d++) { s.is[d] = dsc[7 * n + 1 + d]; s.ie[d] = dsc[7 * n + 4 + d]; }
            if (s.is[0] == s.ie[0]) {
                set3(s.dir, -1, 1, 1);
                if (s.is[0] == 1) { s.face = F_IMIN; s.is[0] -= 2; s.org[0] = 2 * s.ie[0] + 1; }
                else { s.face = F_IMAX; s.ie[0] += 2; s.org[0] = 2 * s.is[0] - 1; }
            } else if (s.is[1] == s.ie[1]) {
                set3(s.dir, 1, -1, 1);
                if (s.is[1] == 1) { s.face = F_JMIN; s.is[1] -= 2; s.org[1] = 2 * s.ie[1] + 1; }
                else { s.face = F_JMAX; s.ie[1] += 2; s.org[1] = 2 * s.is[1] - 1; }
            } else if (s.is[2] == s.ie[2]) {
                set3(s.dir, 1, 1, -1);
                if (s.is[2] == 1) { s.face = F_KMIN; s.is[2] -= 2; s.org[2] = 2 * s.ie[2] + 1; }
                else { s.face = F_KMAX; s.ie[2] += 2; s.org[2] = 2 * s.is[2] - 1; }
            } else return fail("boundary range of block " + std::to_string(b) + " is not a face", -2);
            switch (kind) {
            case CAV3D_BC_WALL: s.fn = wall_fn(iturb); break;
            case CAV3D_BC_INFLOW_SUBSONIC: s.fn = FN_INSUB; break;
            case CAV3D_BC_OUTFLOW_SUBSONIC: s.fn = FN_OUTSUB; break;
            case CAV3D_BC_FARFIELD: s.fn = FN_FARFIELD; break;
            case CAV3D_BC_SYMMETRY: s.fn = FN_SYM; break;
            case CAV3D_BC_INFLOW_SUPERSONIC: s.fn = FN_INSUPER; break;
            case CAV3D_BC_OUTFLOW_SUPERSONIC: case CAV3D_BC_DEGENERATE: s.fn = FN_COPY; break;
            case CAV3D_BC_INFLOW: case CAV3D_BC_OUTFLOW: s.fn = FN_EMPTY; break;
            default: return fail("unknown boundary kind " + std::to_string(kind), -2);
            }
            s.pad = kind;
            {
                // the face's own axis spans its three ghost layers; the others stay on interior cells
                const int ext[3] = {B.I, B.J, B.K};
                const int ax = (s.face == F_IMIN || s.face == F_IMAX) ? 0 : ((s.face == F_JMIN || s.face == F_JMAX) ? 1 : 2);
                for (int d = 0; d < 3; d++) {
                    const int lo = (d == ax) ? -1 : 2, hi = (d == ax) ? ext[d] + 3 : ext[d];
                    if (s.is[d] < lo || s.ie[d] > hi || s.is[d] > s.ie[d])
                        return fail("boundary range of block " + std::to_string(b) + " lies outside its face", -2);
                }
            }
            faces.push_back(s);
        }
        auto covers = [&](const int *p, int d0, int d1, int d2) {
            for (const Seg &s : faces) {
                if (s.pad != CAV3D_BC_WALL && s.pad != CAV3D_BC_SYMMETRY) continue;
                bool ok = s.is[d0] <= p[d0] && s.ie[d0] >= p[d0] && s.is[d1] <= p[d1] && s.ie[d1] >= p[d1];
                if (d2 >= 0) ok = ok && s.is[d2] <= p[d2] && s.ie[d2] >= p[d2];
                if (ok) return true;
            }
            return false;
        };
        int I = B.I, J = B.J, K = B.K;
        Seg e[12];
        memset(e, 0, sizeof(e));
        set3(e[0].is, -1, -1, 2); set3(e[0].ie, 1, 1, K); set3(e[0].nb[0], 1, 2, 0); set3(e[0].nb[1], 2, 1, 0); set3(e[0].org, 2, 2, 0);
        set3(e[1].is, I + 1, -1, 2); set3(e[1].ie, I + 3, 1, K); set3(e[1].nb[0], I + 1, 2, 0); set3(e[1].nb[1], I, 1, 0); set3(e[1].org, I, 2, 0);
        set3(e[2].is, -1, J + 1, 2); set3(e[2].ie, 1, J + 3, K); set3(e[2].nb[0], 1, J, 0); set3(e[2].nb[1], 2, J + 1, 0); set3(e[2].org, 2, J, 0);
        set3(e[3].is, I + 1, J + 1, 2); set3(e[3].ie, I + 3, J + 3, K); set3(e[3].nb[0], I + 1, J, 0); set3(e[3].nb[1], I, J + 1, 0); set3(e[3].org, I, J, 0);
        set3(e[4].is, -1, 2, -1); set3(e[4].ie, 1, J, 1); set3(e[4].nb[0], 1, 0, 2); set3(e[4].nb[2], 2, 0, 1); set3(e[4].org, 2, 0, 2);
        set3(e[5].is, I + 1, 2, -1); set3(e[5].ie, I + 3, J, 1); set3(e[5].nb[0], I + 1, 0, 2); set3(e[5].nb[2], I, 0, 1); set3(e[5].org, I, 0, 2);
        set3(e[6].is, -1, 2, K + 1); set3(e[6].ie, 1, J, K + 3); set3(e[6].nb[0], 1, 0, K); set3(e[6].nb[2], 2, 0, K + 1); set3(e[6].org, 2, 0, K);
        set3(e[7].is, I + 1, 2, K + 1); set3(e[7].ie, I + 3, J, K + 3); set3(e[7].nb[0], I + 1, 0, K); set3(e[7].nb[2], I, 0, K + 1); set3(e[7].org, I, 0, K);
        set3(e[8].is, 2, -1, -1); set3(e[8].ie, I, 1, 1); set3(e[8].nb[1], 0, 1, 2); set3(e[8].nb[2], 0, 2, 1); set3(e[8].org, 0, 2, 2);
        set3(e[9].is, 2, J + 1, -1); set3(e[9].ie, I, J + 3, 1); set3(e[9].nb[1], 0, J + 1, 2); set3(e[9].nb[2], 0, J, 1); set3(e[9].org, 0, J, 2);
        set3(e[10].is, 2, -1, K + 1); set3(e[10].ie, I, 1, K + 3); set3(e[10].nb[1], 0, 1, K); set3(e[10].nb[2], 0, 2, K + 1); set3(e[10].org, 0, 2, K);
        set3(e[11].is, 2, J + 1, K + 1); set3(e[11].ie, I, J + 3, K + 3); set3(e[11].nb[1], 0, J + 1, K); set3(e[11].nb[2], 0, J, K + 1); set3(e[11].org, 0, J, K);
        const int ww = iturb == 0 ? FN_EDGEWWKW : FN_EDGEWW;
        for (int m = 0; m < 12; m++) { e[m].b = b; e[m].face = F_NONE; }
        for (int m = 0; m < 4; m++) {
            bool is = covers(e[m].nb[0], 0, 1, -1), js = covers(e[m].nb[1], 0, 1, -1);
            if (is && js) e[m].fn = ww;
            else if (is) { e[m].face = (m == 0 || m == 2) ? F_IMIN : F_IMAX; e[m].org[0] = e[m].nb[1][0]; e[m].org[1] = e[m].nb[1][1]; e[m].fn = wall_fn(iturb); }
            else if (js) { e[m].face = (m == 0 || m == 1) ? F_JMIN : F_JMAX; e[m].org[0] = e[m].nb[0][0]; e[m].org[1] = e[m].nb[0][1]; e[m].fn = wall_fn(iturb); }
            else e[m].fn = FN_EDGENOWALL;
            set3(e[m].dir, 0, 0, 1);
        }
        for (int m = 4; m < 8; m++) {
            bool is = covers(e[m].nb[0], 0, 2, -1), ks = covers(e[m].nb[2], 0, 2, -1);
            if (is && ks) e[m].fn = ww;
            else if (is) { e[m].face = (m == 4 || m == 6) ? F_IMIN : F_IMAX; e[m].org[0] = e[m].nb[2][0]; e[m].org[2] = e[m].nb[2][2]; e[m].fn = wall_fn(iturb); }
            else if (ks) { e[m].face = (m == 4 || m == 5) ? F_KMIN : F_KMAX; e[m].org[0] = e[m].nb[0][0]; e[m].org[2] = e[m].nb[0][2]; e[m].fn = wall_fn(iturb); }
            else e[m].fn = FN_EDGENOWALL;
            set3(e[m].dir, 0, 1, 0);
        }
        for (int m = 8; m < 12; m++) {
            bool js = covers(e[m].nb[1], 1, 2, -1), ks = covers(e[m].nb[2], 1, 2, -1);
            if (js && ks) e[m].fn = ww;
            else if (js) { e[m].face = (m == 8 || m == 10) ? F_JMIN : F_JMAX; e[m].org[1] = e[m].nb[2][1]; e[m].org[2] = e[m].nb[2][2]; e[m].fn = wall_fn(iturb); }
            else if (ks) { e[m].face = (m == 8 || m == 9) ? F_KMIN : F_KMAX; e[m].org[1] = e[m].nb[1][1]; e[m].org[2] = e[m].nb[1][2]; e[m].fn = wall_fn(iturb); }
            else e[m].fn = FN_EDGENOWALL;
            set3(e[m].dir, 1, 0, 0);
        }
        Seg q[8];
        memset(q, 0, sizeof(q));
        for (int m = 0; m < 8; m++) {
            int ib = m & 1, jb = (m >> 1) & 1, kb = (m >> 2) & 1;
            int ig = ib ? I + 1 : 1, jg = jb ? J + 1 : 1, kg = kb ? K + 1 : 1;
            int ic = ib ? I : 2, jc = jb ? J : 2, kc = kb ? K : 2;
            int hi[3] = {I + 1, J + 1, K + 1}, bit[3] = {ib, jb, kb};
            for (int d = 0; d < 3; d++) { q[m].is[d] = bit[d] ? hi[d] : -1; q[m].ie[d] = bit[d] ? hi[d] + 2 : 1; }
            set3(q[m].nb[0], ig, jc, kc); set3(q[m].nb[1], ic, jg, kc); set3(q[m].nb[2], ic, jc, kg);
            set3(q[m].org, ic, jc, kc);
            set3(q[m].nb[3], ic, jg, kg); set3(q[m].nb[4], ig, jc, kg); set3(q[m].nb[5], ig, jg, kc);
            q[m].face = F_NONE;
            q[m].b = b;
        }
        set3(q[5].nb[2], I, 1, K + 1);   // neighbor3 of corner 6 as tabulated at src/bc_module.f90:958-960
        for (int m = 0; m < 8; m++) {
            Seg &cm = q[m];
            bool is = covers(cm.nb[0], 0, 1, 2), js = covers(cm.nb[1], 0, 1, 2), ks = covers(cm.nb[2], 0, 1, 2);
            if (is && js && ks) cm.fn = iturb == 0 ? FN_CORNERWWKW : FN_CORNERWW;
            else if (!is && js && ks) { memcpy(cm.org, cm.nb[0], 12); set3(cm.nb[0], 0, 0, 0); memcpy(cm.nb[1], cm.nb[4], 12); memcpy(cm.nb[2], cm.nb[5], 12); cm.fn = ww; }
            else if (is && !js && ks) { memcpy(cm.org, cm.nb[1], 12); memcpy(cm.nb[0], cm.nb[3], 12); set3(cm.nb[1], 0, 0, 0); memcpy(cm.nb[2], cm.nb[5], 12); cm.fn = ww; }
            else if (is && js && !ks) { memcpy(cm.org, cm.nb[2], 12); memcpy(cm.nb[0], cm.nb[3], 12); memcpy(cm.nb[1], cm.nb[4], 12); set3(cm.nb[2], 0, 0, 0); cm.fn = ww; }
            else if (!is && !js && ks) { cm.face = m < 4 ? F_KMIN : F_KMAX; memcpy(cm.org, cm.nb[5], 12); cm.fn = wall_fn(iturb); }
            else if (is && !js && !ks) { cm.face = (m % 2 == 0) ? F_IMIN : F_IMAX; memcpy(cm.org, cm.nb[3], 12); cm.fn = wall_fn(iturb); }
            else if (!is && js && !ks) { cm.face = (m == 0 || m == 1 || m == 4 || m == 5) ? F_JMIN : F_JMAX; memcpy(cm.org, cm.nb[4], 12); cm.fn = wall_fn(iturb); }
            else cm.fn = FN_CORNERNOWALL;
        }
        for (Seg &s : faces) if (s.fn != FN_EMPTY) c->segs[0].push_back(s);
        for (int m = 0; m < 12; m++) c->segs[1].push_back(e[m]);
        for (int m = 0; m < 8; m++) c->segs[2].push_back(q[m]);
    }
    for (int w = 0; w < 3; w++) {
        long start = 0;
        for (Seg &s : c->segs[w]) {
            s.start = start;
            start += (long)(s.ie[0] - s.is[0] + 1) * (s.ie[1] - s.is[1] + 1) * (s.ie[2] - s.is[2] + 1);
        }
        c->S.NBC[w] = start;
        c->S.nbcseg[w] = (int)c->segs[w].size();
    }
    return 0;
}

int launch_setbc(cav3d_ctx *c)
{
    DS &S = c->S;
    if (S.NBC[0]) hipLaunchKernelGGL(k_bc, dim3(nblk(S.NBC[0], 128)), dim3(128), 0, c->st, S, 0);
    if (S.NHALO) hipLaunchKernelGGL(k_halo, dim3(nblk(S.NHALO, 128)), dim3(128), 0, c->st, S);
    if (S.NBC[1]) hipLaunchKernelGGL(k_bc, dim3(nblk(S.NBC[1], 128)), dim3(128), 0, c->st, S, 1);
    if (S.NBC[2]) hipLaunchKernelGGL(k_bc, dim3(nblk(S.NBC[2], 128)), dim3(128), 0, c->st, S, 2);
    return 0;
}

int launch_rhs(cav3d_ctx *c)
{
    DS &S = c->S;
    const bool sensor = S.P.nscheme == 2 || S.P.nscheme == 3;
    // AUSMPW+ stays FP64 in the mixed mode: its pressure weight
    // (p + p_ref + rho c^2)/(rho c^2) - 1 is ~1e-5 in water and cancels in FP32
    const bool mixed = c->prm.precision == CAV3D_MIXED && S.P.nscheme != 3;
    // the three directions are independent (each writes its own face
    // fluxes): i on st, j and k on the two side streams, joined before the
    // combine, so that one direction's tail overlaps the others
    hipStream_t sd[3] = {c->st, c->side[0], c->side[1]};
    hipEventRecord(c->fork, c->st);
    hipStreamWaitEvent(c->side[0], c->fork, 0);
    hipStreamWaitEvent(c->side[1], c->fork, 0);
#define CAV3D_FACES(SEN, R)                                                                             \
    do {                                                                                                \
        hipLaunchKernelGGL((k_face<0, SEN, R>), dim3(nblk(S.NF[0], 256)), dim3(256), 0, sd[0], S);      \
        hipLaunchKernelGGL((k_face<1, SEN, R>), dim3(nblk(S.NF[1], 256)), dim3(256), 0, sd[1], S);      \
        hipLaunchKernelGGL((k_face<2, SEN, R>), dim3(nblk(S.NF[2], 256)), dim3(256), 0, sd[2], S);      \
    } while (0)
    if (sensor && mixed) CAV3D_FACES(true, float);
    else if (sensor) CAV3D_FACES(true, double);
    else if (mixed) CAV3D_FACES(false, float);
    else CAV3D_FACES(false, double);
#undef CAV3D_FACES
    if (c->prm.iturb >= -2) {
        if (c->prm.precision == CAV3D_MIXED) {
            hipLaunchKernelGGL((k_vface<0, float>), dim3(nblk(S.NF[0], 256)), dim3(256), 0, sd[0], S);
            hipLaunchKernelGGL((k_vface<1, float>), dim3(nblk(S.NF[1], 256)), dim3(256), 0, sd[1], S);
            hipLaunchKernelGGL((k_vface<2, float>), dim3(nblk(S.NF[2], 256)), dim3(256), 0, sd[2], S);
        } else {
            hipLaunchKernelGGL((k_vface<0, double>), dim3(nblk(S.NF[0], 256)), dim3(256), 0, sd[0], S);
            hipLaunchKernelGGL((k_vface<1, double>), dim3(nblk(S.NF[1], 256)), dim3(256), 0, sd[1], S);
            hipLaunchKernelGGL((k_vface<2, double>), dim3(nblk(S.NF[2], 256)), dim3(256), 0, sd[2], S);
        }
    }
    hipEventRecord(c->join[0], c->side[0]);
    hipEventRecord(c->join[1], c->side[1]);
    hipStreamWaitEvent(c->st, c->join[0], 0);
    hipStreamWaitEvent(c->st, c->join[1], 0);
    if (S.P.ncav >= 2 || S.P.ides > 0) hipLaunchKernelGGL(k_combine<true>, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S);
    else hipLaunchKernelGGL(k_combine<false>, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S);
    return 0;
}

int launch_timestep(cav3d_ctx *c, hipStream_t st)
{
    DS &S = c->S;
    hipLaunchKernelGGL(k_reset_scal, dim3(1), dim3(64), 0, st, S);
    hipLaunchKernelGGL(k_timestep, dim3(nblk(S.NC, 256)), dim3(256), 0, st, S);
    if (c->prm.local == 0) hipLaunchKernelGGL(k_dt_fill, dim3(nblk(S.NC, 256)), dim3(256), 0, st, S);
    hipLaunchKernelGGL(k_time_acc, dim3(1), dim3(64), 0, st, S);
    return 0;
}

// the sweeps' exchange arrays back to "not yet written"
int launch_sweep_fills(cav3d_ctx *c, hipStream_t st)
{
    DS &S = c->S;
    const long nq = (long)NPV * S.NW;
    hipLaunchKernelGGL(k_fill_bits, dim3(nblk(nq, 256)), dim3(256), 0, st, S.xqf, nq, DQ_UNSET);
    hipLaunchKernelGGL(k_fill_bits, dim3(nblk(nq, 256)), dim3(256), 0, st, S.xqb, nq, DQ_UNSET);
    return 0;
}

int launch_sweeps(cav3d_ctx *c)
{
    DS &S = c->S;
    hipLaunchKernelGGL(k_lhs_prep, dim3(nblk(S.NC, 128)), dim3(128), 0, c->st, S);
    const int *tiles = c->tiles;
    int ntiles = c->ntiles;
    int *err = c->sweep_err;
    void *args[4] = {&S, &tiles, &ntiles, &err};
    dim3 grid(c->coop_blocks), block(SWEEP_THREADS);
    if (c->profile) hipEventRecord(c->ev[6], c->st);
    hipError_t e = hipLaunchCooperativeKernel((const void *)k_sweep_tile<true>, grid, block, args, 0, c->st);
    if (e != hipSuccess) return fail(std::string("forward sweep launch: ") + hipGetErrorString(e), -3);
    if (c->profile) hipEventRecord(c->ev[7], c->st);
    e = hipLaunchCooperativeKernel((const void *)k_sweep_tile<false>, grid, block, args, 0, c->st);
    if (e != hipSuccess) return fail(std::string("backward sweep launch: ") + hipGetErrorString(e), -3);
    if (c->profile) hipEventRecord(c->ev[8], c->st);
    return 0;
}

int launch_update(cav3d_ctx *c)
{
    DS &S = c->S;
    hipLaunchKernelGGL(k_update, dim3(c->nwg_update), dim3(256), 0, c->st, S);
    // the final norm reduction reads only the update's partials: it runs on
    // a side stream beside the eddy viscosity, and reaches st again at the
    // next residual's join (before the next update rewrites the partials);
    // cav3d_residual waits for it
    hipEventRecord(c->fork, c->st);
    hipStreamWaitEvent(c->side[0], c->fork, 0);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, c->side[0], S, c->nwg_update);
    if (c->prm.iturb == 0) hipLaunchKernelGGL(k_eddy, dim3(nblk(S.NC, 128)), dim3(128), 0, c->st, S);
    return 0;
}

// a sweep whose bounded wait gave up leaves a flag; report it (stream synchronised by the caller)
int sweep_status(cav3d_ctx *c)
{
    int e = 0;
    HIPCHK(hipMemcpy(&e, c->sweep_err, sizeof(int), hipMemcpyDeviceToHost));
    if (e) return fail("LU-SGS sweep: a dependency wait reached its limit; results of that iteration are invalid", -5);
    return 0;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e), -3);
    return 0;
}

}  // namespace

// ============================================================ C ABI
extern "C" {

const char *cav3d_last_error(void) { return g_err.c_str(); }

int cav3d_create(const cav3d_params *prm, int nblocks, const cav3d_block *blocks, cav3d_ctx **out)
{
    if (!prm || !blocks || !out) return fail("null argument", -1);
    *out = nullptr;
    if (nblocks < 1 || nblocks > MAXB) return fail("nblocks must be 1.." + std::to_string(MAXB), -1);
    if (prm->iturb != 0 && prm->iturb != -2 && prm->iturb != -3)
        return fail("iturb must be 0 (k-omega SST), -2 (laminar) or -3 (Euler)", -1);
    if (prm->nscheme < 1 || prm->nscheme > 4)
        return fail("nscheme must be 1 (Roe), 2 (Roe-M), 3 (AUSMPW+) or 4 (AUSM+-up)", -1);
    if (prm->nmuscl < 0 || prm->nmuscl > 2 || prm->nlim < 0 || prm->nlim > 10) return fail("bad nmuscl/nlim", -1);
    if (prm->ides < 0 || prm->ides > 4) return fail("ides must be 0 (RANS), 1 (DES), 2 (DDES), 3 (IDDES) or 4 (S-IDDES)", -1);
    if (prm->ncav < 0 || prm->ncav > 5)
        return fail("ncav must be 0 (none), 1 (Merkle), 2 (Kunz), 3 (Singhal), 4 (Schnerr-Sauer) or 5 (Lee)", -1);
    if (prm->cav_clip != 0 && prm->cav_clip != 1) return fail("cav_clip must be 0 (as written) or 1 (species)", -1);
    if (prm->nsteady < 0 || prm->nsteady > 2) return fail("nsteady must be 0, 1 or 2", -1);
    if (!(prm->dt_phy > 0.0)) return fail("dt_phy must be > 0 (it divides the BDF term of getx)", -1);
    if (prm->precision != CAV3D_FP64 && prm->precision != CAV3D_MIXED)
        return fail("precision must be CAV3D_FP64 (0) or CAV3D_MIXED (1)", -1);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device", -4);
    cav3d_ctx *c = new cav3d_ctx();
    hipGetDevice(&c->dev);
    c->prm = *prm;
    c->nb = nblocks;
    c->npv = prm->iturb == 0 ? 9 : 7;
    c->ntv = prm->iturb == 0 ? 3 : (prm->iturb == -2 ? 2 : 0);
    c->ngrd = prm->iturb == 0 ? 5 : 4;
    Par &P = c->P;
    memset(&P, 0, sizeof(P));
    P.npv = c->npv; P.ntv = c->ntv; P.iturb = prm->iturb; P.nscheme = prm->nscheme; P.nmuscl = prm->nmuscl;
    P.nlim = prm->nlim; P.local = prm->local; P.prec = prm->prec; P.ncav = prm->ncav; P.nsteady = prm->nsteady;
    P.cav_clip = prm->cav_clip;
    P.ides = prm->iturb == 0 ? prm->ides : 0;
    P.cfl = prm->cfl; P.pref = prm->pref; P.uref = prm->uref; P.str = prm->str; P.dt_phy = prm->dt_phy;
    P.c_v = prm->c_v; P.c_c = prm->c_c; P.dp_ref = prm->dp_ref; P.t_ref = prm->t_ref; P.t_crit = prm->t_crit;
    for (int q = 0; q < 9; q++) P.bc_pv[q] = prm->bc_pv[q];
    P.bc_pressure = prm->bc_pressure;
    P.emutref = prm->emutref;
    P.bdf_b = prm->nsteady == 1 ? 1.0 : (prm->nsteady == 2 ? 1.5 * 0.5 : 0.0);
    if (hipStreamCreate(&c->st) != hipSuccess) { delete c; return fail("hipStreamCreate failed", -3); }
    for (int q = 0; q < 2; q++)
        if (hipStreamCreateWithFlags(&c->side[q], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->join[q], hipEventDisableTiming) != hipSuccess) {
            cav3d_destroy(c);
            return fail("stream set-up failed", -3);
        }
    if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess) { cav3d_destroy(c); return fail("stream set-up failed", -3); }
    DS &S = c->S;
    memset(&S, 0, sizeof(S));
    S.nb = nblocks;
    S.npv = c->npv;
    S.P = P;
    long g0 = 0, c0 = 0, w0 = 0, f0[3] = {0, 0, 0}, x0 = 0, e0 = 0, t0 = 0, p0 = 0;
    c->blk.resize(nblocks);
    c->bcdesc.resize(nblocks);
    for (int b = 0; b < nblocks; b++) {
        const cav3d_block &cb = blocks[b];
        if (cb.imax < 4 || cb.jmax < 4 || cb.kmax < 4) { delete c; return fail("each block needs at least 3 cells per direction", -1); }
        BlkD &B = c->blk[b];
        B.I = cb.imax; B.J = cb.jmax; B.K = cb.kmax;
        B.nx = B.I + 5; B.ny = B.J + 5;
        B.ng = (long)B.nx * B.ny * (B.K + 5);
        B.nc = (long)(B.I - 1) * (B.J - 1) * (B.K - 1);
        B.nw = (long)(B.I + 1) * (B.J + 1) * (B.K + 1);
        B.nf[0] = (long)B.I * (B.J - 1) * (B.K - 1);
        B.nf[1] = (long)(B.I - 1) * B.J * (B.K - 1);
        B.nf[2] = (long)(B.I - 1) * (B.J - 1) * B.K;
        B.ncx = (long)B.I * (B.J + 1) * (B.K + 1);
        B.nex = (long)(B.I + 1) * B.J * (B.K + 1);
        B.ntx = (long)(B.I + 1) * (B.J + 1) * B.K;
        B.g0 = g0; B.c0 = c0; B.w0 = w0; B.cx0 = x0; B.ex0 = e0; B.tx0 = t0;
        for (int d = 0; d < 3; d++) { B.f0[d] = f0[d]; f0[d] += B.nf[d]; }
        B.p0 = p0;
        p0 += (long)(B.J - 1) * (B.K - 1);
        g0 += B.ng; c0 += B.nc; w0 += B.nw; x0 += B.ncx; e0 += B.nex; t0 += B.ntx;
        S.b[b] = B;
        if (cb.nbc < 0 || (cb.nbc > 0 && !cb.bc)) { delete c; return fail("bad bc list", -1); }
        c->bcdesc[b].assign(cb.bc, cb.bc + 7 * cb.nbc);
    }
    S.NG = g0; S.NC = c0; S.NW = w0; S.NCX = x0; S.NEX = e0; S.NTX = t0; S.NPLANE = p0;
    for (int d = 0; d < 3; d++) S.NF[d] = f0[d];
    // interfaces
    long hs = 0;
    c->condesc.resize(nblocks);
    for (int b = 0; b < nblocks; b++) {
        const cav3d_block &cb = blocks[b];
        if (cb.ncon < 0 || (cb.ncon > 0 && !cb.con)) { delete c; return fail("bad interface list", -1); }
        c->condesc[b].assign(cb.con, cb.con + 22 * cb.ncon);
        for (int n = 0; n < cb.ncon; n++) {
            const int *d = cb.con + 22 * n;
            Con cn;
            memset(&cn, 0, sizeof(cn));
            cn.src = b; cn.dst = d[0];
            if (cn.dst < 0 || cn.dst >= nblocks) { delete c; return fail("interface donor out of range", -1); }
            for (int q = 0; q < 3; q++) { cn.is[q] = d[1 + q]; cn.ie[q] = d[4 + q]; cn.isd[q] = d[7 + q]; }
            for (int q = 0; q < 9; q++) cn.tm[q] = d[13 + q];
            long cnt = 1;
            for (int q = 0; q < 3; q++) cnt *= (cn.ie[q] >= cn.is[q]) ? (cn.ie[q] - cn.is[q] + 1) : 0;
            if (cnt == 0) continue;
            // both the source box and its image must lie inside the ghosted arrays
            const int ext[2][3] = {{blocks[b].imax, blocks[b].jmax, blocks[b].kmax},
                                   {blocks[cn.dst].imax, blocks[cn.dst].jmax, blocks[cn.dst].kmax}};
            bool ok = true;
            for (int q = 0; q < 3; q++) {
                if (cn.is[q] < -1 || cn.ie[q] > ext[0][q] + 3) ok = false;
                for (int r = 0; r < 3; r++)
                    if (cn.tm[3 * q + r] < -1 || cn.tm[3 * q + r] > 1) ok = false;
            }
            for (int corner = 0; ok && corner < 8; corner++) {
                int del[3];
                for (int q = 0; q < 3; q++) del[q] = ((corner >> q) & 1) ? cn.ie[q] - cn.is[q] : 0;
                for (int q = 0; q < 3; q++) {
                    int v = cn.tm[3 * q] * del[0] + cn.tm[3 * q + 1] * del[1] + cn.tm[3 * q + 2] * del[2] + cn.isd[q];
                    if (v < -1 || v > ext[1][q] + 3) ok = false;
                }
            }
            if (!ok) { delete c; return fail("interface " + std::to_string(n) + " of block " + std::to_string(b) +
                                             " maps outside the ghosted arrays", -1); }
            cn.start = hs;
            hs += cnt;
            c->cons.push_back(cn);
        }
    }
    S.NHALO = hs;
    S.ncon = (int)c->cons.size();
    int rc = build_bcs(c);
    if (rc) { cav3d_destroy(c); return rc; }
    const int npv = c->npv;
    if ((rc = dalloc(c, &S.pv, npv * S.NG)) || (rc = dalloc(c, &S.dc, (long)ND * S.NG)) ||
        (rc = dalloc(c, &S.vol, S.NW)) || (rc = dalloc(c, &c->xc, 3 * S.NW)) || (rc = dalloc(c, &S.cx, 3 * S.NCX)) || (rc = dalloc(c, &S.ex, 3 * S.NEX)) ||
        (rc = dalloc(c, &S.tx, 3 * S.NTX)) || (rc = dalloc(c, &S.res, npv * S.NC)) || (rc = dalloc(c, &S.dt, S.NC)) ||
        (rc = dalloc(c, &S.dcv, npv * S.NW)) ||
        (rc = dalloc(c, &S.xqf, (long)NPV * S.NW)) || (rc = dalloc(c, &S.xqb, (long)NPV * S.NW)) ||
        (rc = dalloc(c, &S.scal, 32))) { cav3d_destroy(c); return rc; }
    for (int d = 0; d < 3; d++)
        if ((rc = dalloc(c, &S.ff[d], 2 * npv * S.NF[d]))) { cav3d_destroy(c); return rc; }
    if (prm->iturb == 0) {
        if ((rc = dalloc(c, &S.mut, S.NG)) || (rc = dalloc(c, &S.ydn, S.NW)) || (rc = dalloc(c, &S.itt, 4 * S.NC)) ||
            (rc = dalloc(c, &S.ocut, S.NC))) { cav3d_destroy(c); return rc; }
    }
    if (prm->ncav) {
        if ((rc = dalloc(c, &S.icav, 4 * S.NC)) || (rc = dalloc(c, &S.rcav, 2 * S.NC))) { cav3d_destroy(c); return rc; }
    }
    if (prm->nsteady > 0) {
        if ((rc = dalloc(c, &S.qq, 2 * npv * S.NC))) { cav3d_destroy(c); return rc; }
    }
    int maxi = 0;
    for (int b = 0; b < nblocks; b++) maxi = std::max(maxi, c->blk[b].I);
    if ((rc = dalloc(c, &S.gx, (long)NGX * S.NC)) || (rc = dalloc(c, &S.jc, 2 * S.NC)) ||
        (rc = dalloc(c, &S.jf, 18 * S.NC)) || (rc = dalloc(c, &S.zero, maxi + 8))) { cav3d_destroy(c); return rc; }
    {
        // tile table: TJ x TK pencil tiles of every block, ordered by the tile
        // diagonal tj + tk (the forward sweep takes them in this order, the
        // backward sweep in reverse)
        std::vector<std::array<int, 4>> tl;
        for (int b = 0; b < nblocks; b++) {
            const int ntj = (c->blk[b].J - 1 + TJ - 1) / TJ, ntk = (c->blk[b].K - 1 + TK - 1) / TK;
            for (int tk = 0; tk < ntk; tk++)
                for (int tj = 0; tj < ntj; tj++) tl.push_back({b, 2 + TJ * tj, 2 + TK * tk, tj + tk});
        }
        std::stable_sort(tl.begin(), tl.end(), [](const std::array<int, 4> &x, const std::array<int, 4> &y) {
            return x[3] < y[3];
        });
        std::vector<int> tab;
        for (auto &t : tl) tab.insert(tab.end(), t.begin(), t.end());
        c->ntiles = (int)tl.size();
        if ((rc = dalloc(c, &c->tiles, tab.size())) || (rc = dalloc(c, &c->sweep_err, 1))) {
            cav3d_destroy(c); return rc;
        }
        if (hipMemcpyAsync(c->tiles, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, c->st) != hipSuccess) {
            cav3d_destroy(c); return fail("tile table upload failed", -3);
        }
        int dev = 0, ncu = 0, per = 0, per2 = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_sweep_tile<true>, SWEEP_THREADS, 0);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, (const void *)k_sweep_tile<false>, SWEEP_THREADS, 0);
        per = std::min(per, per2);
        long cap = (long)ncu * per;
        if (cap < 1) { cav3d_destroy(c); return fail("sweep kernel cannot be made resident", -3); }
        c->coop_blocks = (int)std::max(1L, std::min((long)c->ntiles, cap));
    }
    c->nwg_update = (int)nblk(S.NC, 256);
    if ((rc = dalloc(c, &S.red, (long)c->nwg_update * (NPV + 2)))) { cav3d_destroy(c); return rc; }
    for (int w = 0; w < 3; w++) {
        if ((rc = dalloc(c, &S.seg[w], c->segs[w].size()))) { cav3d_destroy(c); return rc; }
        if (!c->segs[w].empty() &&
            hipMemcpyAsync(S.seg[w], c->segs[w].data(), c->segs[w].size() * sizeof(Seg), hipMemcpyHostToDevice, c->st) != hipSuccess) {
            cav3d_destroy(c); return fail("segment upload failed", -3);
        }
    }
    if ((rc = dalloc(c, &S.con, c->cons.size()))) { cav3d_destroy(c); return rc; }
    if (!c->cons.empty() &&
        hipMemcpyAsync(S.con, c->cons.data(), c->cons.size() * sizeof(Con), hipMemcpyHostToDevice, c->st) != hipSuccess) {
        cav3d_destroy(c); return fail("interface upload failed", -3);
    }
    for (int q = 0; q < 14; q++) hipEventCreate(&c->ev[q]);
    if (hipStreamSynchronize(c->st) != hipSuccess) { cav3d_destroy(c); return fail("device set-up failed", -3); }
    *out = c;
    return 0;
}

void cav3d_destroy(cav3d_ctx *c)
{
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    for (void *p : c->allocs) hipFree(p);
    if (c->stage) hipFree(c->stage);
    if (c->xbuf) hipFree(c->xbuf);
    if (c->a_obj) hipFree(c->a_obj);
    for (int q = 0; q < 14; q++) if (c->ev[q]) hipEventDestroy(c->ev[q]);
    for (int q = 0; q < 2; q++) {
        if (c->side[q]) { hipStreamSynchronize(c->side[q]); hipStreamDestroy(c->side[q]); }
        if (c->join[q]) hipEventDestroy(c->join[q]);
    }
    if (c->fork) hipEventDestroy(c->fork);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

int cav3d_set_metrics(cav3d_ctx *c, int b, const double *cx, const double *ex, const double *tx, const double *grd)
{
    if (!c || b < 0 || b >= c->nb || !cx || !ex || !tx || !grd) return fail("bad argument to cav3d_set_metrics", -1);
    const BlkD &B = c->blk[b];
    DS &S = c->S;
    int rc;
    long nmax = std::max(std::max(3 * B.ncx, 3 * B.nex), std::max(3 * B.ntx, (long)c->ngrd * B.nw));
    if ((rc = ensure_stage(c, nmax))) return rc;
    struct { const double *h; long n; double *d; long nst; long o; } parts[3] = {
        {cx, B.ncx, S.cx, S.NCX, B.cx0}, {ex, B.nex, S.ex, S.NEX, B.ex0}, {tx, B.ntx, S.tx, S.NTX, B.tx0}};
    for (auto &p : parts) {
        HIPCHK(hipMemcpyAsync(c->stage, p.h, 3 * p.n * sizeof(double), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(p.n, 256)), dim3(256), 0, c->st, (const double *)c->stage, 3, 3, 0, p.d, p.nst, p.o, p.n);
        HIPCHK(hipStreamSynchronize(c->st));
    }
    HIPCHK(hipMemcpyAsync(c->stage, grd, (long)c->ngrd * B.nw * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(B.nw, 256)), dim3(256), 0, c->st, (const double *)c->stage, 1, c->ngrd, 0, S.vol, S.NW, B.w0, B.nw);
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(B.nw, 256)), dim3(256), 0, c->st, (const double *)c->stage, 3, c->ngrd, 1, c->xc, S.NW, B.w0, B.nw);
    if (S.ydn)
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(B.nw, 256)), dim3(256), 0, c->st, (const double *)c->stage, 1, c->ngrd, 4, S.ydn, S.NW, B.w0, B.nw);
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_set_metrics");
}

int cav3d_put_state(cav3d_ctx *c, int b, const double *pv, const double *dv, const double *tv, const double *qq)
{
    if (!c || b < 0 || b >= c->nb || !pv) return fail("bad argument to cav3d_put_state", -1);
    const BlkD &B = c->blk[b];
    DS &S = c->S;
    int rc;
    if ((rc = ensure_stage(c, (long)(18 + 3) * B.ng + 2L * c->npv * B.nc))) return rc;
    HIPCHK(hipMemcpyAsync(c->stage, pv, (long)c->npv * B.ng * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, (const double *)c->stage, c->npv, c->npv, 0, S.pv, S.NG, B.g0, B.ng);
    HIPCHK(hipStreamSynchronize(c->st));
    if (dv) {
        HIPCHK(hipMemcpyAsync(c->stage, dv, 18L * B.ng * sizeof(double), hipMemcpyHostToDevice, c->st));
        double *tvd = c->stage + 18L * B.ng;
        if (tv && c->ntv) HIPCHK(hipMemcpyAsync(tvd, tv, (long)c->ntv * B.ng * sizeof(double), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_dv_to_dc, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, (const double *)c->stage,
                           (const double *)tvd, (tv ? c->ntv : 0), S.dc, S.mut, S.NG, B.g0, B.ng);
        HIPCHK(hipStreamSynchronize(c->st));
    } else {
        DS S1 = S;   // evaluate this block's interior only
        S1.nb = 1; S1.b[0] = B; S1.b[0].c0 = 0; S1.NC = B.nc;
        S1.b[0].g0 = B.g0;
        hipLaunchKernelGGL(k_eos_cells, dim3(nblk(B.nc, 128)), dim3(128), 0, c->st, S1, 1, 1);
        HIPCHK(hipStreamSynchronize(c->st));
    }
    if (c->prm.nsteady > 0) {
        if (qq) {
            HIPCHK(hipMemcpyAsync(c->stage, qq, 2L * c->npv * B.nc * sizeof(double), hipMemcpyHostToDevice, c->st));
            hipLaunchKernelGGL(k_cell_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage,
                               2 * c->npv, S.NC, B.c0, S.qq, B.nc);
        } else {
            DS S1 = S;
            S1.nb = 1; S1.b[0] = B; S1.NC = S.NC;
            // both levels from the current state (initial_module set_others)
            hipLaunchKernelGGL(k_unsteady_shift, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S);
            hipLaunchKernelGGL(k_unsteady_shift, dim3(nblk(S.NC, 256)), dim3(256), 0, c->st, S);
        }
        HIPCHK(hipStreamSynchronize(c->st));
    }
    return check_launch("cav3d_put_state");
}

int cav3d_get_state(cav3d_ctx *c, int b, double *pv, double *dv, double *tv, double *qq)
{
    if (!c || b < 0 || b >= c->nb) return fail("bad argument to cav3d_get_state", -1);
    const BlkD &B = c->blk[b];
    DS &S = c->S;
    int rc;
    HIPCHK(hipStreamSynchronize(c->st));
    if ((rc = ensure_stage(c, (long)(18 + 3) * B.ng + 2L * c->npv * B.nc))) return rc;
    if (pv) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, (const double *)S.pv, c->npv, S.NG, B.g0, c->stage, B.ng);
        HIPCHK(hipMemcpyAsync(pv, c->stage, (long)c->npv * B.ng * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    if (dv || tv) {
        // evaluate the EOS over this block's ghosted cells into AoS staging
        double *dvd = c->stage, *tvd = c->stage + 18L * B.ng;
        hipLaunchKernelGGL(k_eos_full, dim3(nblk(B.ng, 128)), dim3(128), 0, c->st, S, B.g0, B.ng, dvd, tvd);
        HIPCHK(hipStreamSynchronize(c->st));
        if (dv) HIPCHK(hipMemcpy(dv, dvd, 18L * B.ng * sizeof(double), hipMemcpyDeviceToHost));
        if (tv && c->ntv) {
            std::vector<double> t2(2 * B.ng), mt;
            HIPCHK(hipMemcpy(t2.data(), tvd, 2L * B.ng * sizeof(double), hipMemcpyDeviceToHost));
            if (c->ntv == 3) {
                mt.resize(B.ng);
                HIPCHK(hipMemcpy(mt.data(), S.mut + B.g0, B.ng * sizeof(double), hipMemcpyDeviceToHost));
            }
            for (long g = 0; g < B.ng; g++) {
                tv[c->ntv * g] = t2[2 * g];
                tv[c->ntv * g + 1] = t2[2 * g + 1];
                if (c->ntv == 3) tv[3 * g + 2] = mt[g];
            }
        }
    }
    if (qq && c->prm.nsteady > 0) {
        hipLaunchKernelGGL(k_cell_to_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)S.qq, 2 * c->npv, S.NC, B.c0, c->stage, B.nc);
        HIPCHK(hipMemcpyAsync(qq, c->stage, 2L * c->npv * B.nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    return check_launch("cav3d_get_state");
}

int cav3d_setbc(cav3d_ctx *c)
{
    if (!c) return fail("null context", -1);
    launch_setbc(c);
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_setbc");
}

int cav3d_timestep(cav3d_ctx *c, int b, double *dt)
{
    if (!c || b < 0 || b >= c->nb) return fail("bad argument to cav3d_timestep", -1);
    launch_timestep(c, c->st);
    const BlkD &B = c->blk[b];
    if (dt) HIPCHK(hipMemcpyAsync(dt, c->S.dt + B.c0, B.nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_timestep");
}

int cav3d_calrhs(cav3d_ctx *c, double time, int b, double *res, double *icav, double *itt, double *omega_cut)
{
    (void)time;
    if (!c || b < 0 || b >= c->nb) return fail("bad argument to cav3d_calrhs", -1);
    DS &S = c->S;
    const BlkD &B = c->blk[b];
    int rc;
    if ((rc = ensure_stage(c, (long)c->npv * B.nc))) return rc;
    if (c->profile) hipEventRecord(c->ev[12], c->st);
    launch_rhs(c);
    if (c->profile) {
        hipEventRecord(c->ev[13], c->st);
        hipEventSynchronize(c->ev[13]);
        float ms;
        hipEventElapsedTime(&ms, c->ev[12], c->ev[13]);
        c->prof_ms[9] = ms;
    }
    struct { double *h; const double *d; int m; } outs[4] = {{res, S.res, c->npv}, {icav, S.icav, 4}, {itt, S.itt, 4}, {omega_cut, S.ocut, 1}};
    for (auto &o : outs) {
        if (!o.h) continue;
        if (!o.d) { memset(o.h, 0, sizeof(double) * o.m * B.nc); continue; }
        hipLaunchKernelGGL(k_cell_to_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, o.d, o.m, S.NC, B.c0, c->stage, B.nc);
        HIPCHK(hipMemcpyAsync(o.h, c->stage, (long)o.m * B.nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_calrhs");
}

int cav3d_timeinteg(cav3d_ctx *c, int nt_phy, int nt, double timeprev, double *time)
{
    (void)nt;
    if (!c) return fail("null context", -1);
    if (c->profile) hipEventRecord(c->ev[0], c->st);
    // the time step and the sweeps' exchange-array fills read nothing the
    // boundary fill writes (interior cells and metrics only): they run on the
    // side streams meanwhile, joined again before the combine (launch_rhs),
    // which precedes every reader of dt and of the exchange arrays
    hipEventRecord(c->fork, c->st);
    hipStreamWaitEvent(c->side[0], c->fork, 0);
    hipStreamWaitEvent(c->side[1], c->fork, 0);
    launch_sweep_fills(c, c->side[0]);
    launch_timestep(c, c->side[1]);
    if (c->profile) hipEventRecord(c->ev[1], c->st);
    launch_setbc(c);
    if (c->profile) hipEventRecord(c->ev[2], c->st);
    launch_rhs(c);
    if (c->profile) hipEventRecord(c->ev[3], c->st);
    int lrc = launch_sweeps(c);
    if (lrc) return lrc;
    if (c->profile) hipEventRecord(c->ev[4], c->st);
    launch_update(c);
    if (c->profile) {
        hipEventRecord(c->ev[5], c->st);
        hipEventSynchronize(c->ev[5]);
        float ms;
        for (int q = 0; q < 5; q++) { hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]); c->prof_ms[q] = ms; }
        hipEventElapsedTime(&ms, c->ev[0], c->ev[5]);
        c->prof_ms[5] = ms;
        hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
        c->prof_ms[6] = ms;
        hipEventElapsedTime(&ms, c->ev[7], c->ev[8]);
        c->prof_ms[7] = ms;
    }
    int rc = check_launch("cav3d_timeinteg");
    if (rc) return rc;
    if (time) {
        double tacc;
        HIPCHK(hipMemcpyAsync(&tacc, c->S.scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if ((rc = sweep_status(c))) return rc;
        *time = (c->prm.nsteady == 0) ? tacc + timeprev : c->prm.dt_phy * (double)nt_phy;
    }
    return 0;
}

int cav3d_residual(cav3d_ctx *c, double *l2, int *ijk, double *resmax)
{
    if (!c) return fail("null context", -1);
    double h[32];
    HIPCHK(hipStreamSynchronize(c->side[0]));     // the norms' final reduction
    HIPCHK(hipMemcpyAsync(h, c->S.scal, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    int rc = sweep_status(c);
    if (rc) return rc;
    if (l2) for (int n = 0; n < c->npv; n++) l2[n] = h[4 + n];
    long bi;
    memcpy(&bi, &h[5 + NPV], sizeof(long));
    double bm = h[4 + NPV];
    if (ijk) {
        if (bm > 1.0e-12) {
            ijk[0] = (int)(bi >> 48); ijk[1] = (int)(bi & 0xffff); ijk[2] = (int)((bi >> 16) & 0xffff); ijk[3] = (int)((bi >> 32) & 0xffff);
        } else { ijk[0] = -1; ijk[1] = ijk[2] = ijk[3] = 0; }
    }
    if (resmax) *resmax = bm > 1.0e-12 ? bm : 1.0e-12;
    return 0;
}

int cav3d_unsteady_shift(cav3d_ctx *c)
{
    if (!c) return fail("null context", -1);
    if (c->prm.nsteady == 0) return 0;
    hipLaunchKernelGGL(k_unsteady_shift, dim3(nblk(c->S.NC, 256)), dim3(256), 0, c->st, c->S);
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_unsteady_shift");
}

int cav3d_eos(cav3d_ctx *c, int n, const double *p, const double *t, const double *y1, const double *y2, double *dv, double *tv)
{
    if (!c || n < 0 || !p || !t || !y1 || !y2 || !dv || !tv) return fail("bad argument to cav3d_eos", -1);
    if (n == 0) return 0;
    int rc;
    if ((rc = ensure_stage(c, 24L * n))) return rc;
    double *d = c->stage;
    HIPCHK(hipMemcpyAsync(d, p, n * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d + n, t, n * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d + 2L * n, y1, n * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d + 3L * n, y2, n * 8, hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_eos_points, dim3(nblk(n, 128)), dim3(128), 0, c->st, c->P, n, (const double *)d, (const double *)(d + n),
                       (const double *)(d + 2L * n), (const double *)(d + 3L * n), d + 4L * n, d + 22L * n);
    HIPCHK(hipMemcpyAsync(dv, d + 4L * n, 18L * n * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(tv, d + 22L * n, 2L * n * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_eos");
}

int cav3d_eos_host(int n, const double *p, const double *t, const double *y1, const double *y2, double *dv, double *tv)
{
    if (n < 0 || !p || !t || !y1 || !y2 || !dv || !tv) return fail("bad argument to cav3d_eos_host", -1);
    auto work = [&](int lo, int hi) {
        for (int q = lo; q < hi; q++) {
            double d[ND];
            mixture_eos<double>(p[q], t[q], y1[q], y2[q], d, true, dv + 18L * q);
            tv[2L * q] = d[D_MU];
            tv[2L * q + 1] = d[D_KAP];
        }
    };
    int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (int w = 0; w < nt; w++) th.emplace_back(work, (int)((long)n * w / nt), (int)((long)n * (w + 1) / nt));
    for (auto &x : th) x.join();
    return 0;
}

int cav3d_profile(cav3d_ctx *c, int enable, double *out)
{
    if (!c) return fail("null context", -1);
    c->profile = enable != 0;
    if (out) for (int q = 0; q < 10; q++) out[q] = c->prof_ms[q];
    return 0;
}

int cav3d_synchronize(cav3d_ctx *c)
{
    if (!c) return fail("null context", -1);
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_synchronize");
}

}  // extern "C"

#include "metrics.hpp"
#include "post.hpp"
#include "adjoint.hpp"

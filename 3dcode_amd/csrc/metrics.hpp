// Grid metrics on the device from the vertex coordinates: face vectors,
// cell volumes and centroids, wall distance, with the ghost fills and the
// block-interface exchanges of t_grid's construct sequence
// (src/grid_module.f90:328-330):
//   calnormal  src/grid_module.f90:609-940   cx/ex/tx, degenerate faces, ghost
//              copies, exchange of the tangential face vectors
//   calvolume  src/grid_module.f90:942-1119  centroids, six-pyramid volumes
//              (pyramid :1121-1145), ghost fill, exchange of the volume
//   calydns    src/grid_module.f90:350-607   nearest BCWall face centroid over
//              every block, ghost copies, exchange
// One block stands for one MPI rank of the reference.  An exchange packs
// every sender's values before any receiver is written (the MPI branches);
// a block connected to itself is written at once, in connection order.
//
// The wall distance is the only heavy part: every interior cell against every
// wall-face centroid, FP64, exact minimum.  Wall points are staged through
// LDS in tiles and read as wavefront-uniform broadcasts; each thread keeps two
// cells so one LDS read feeds two distance evaluations.
//
// Included once, from cav3d.hip, after the context definition.

namespace {

struct XB {                 // one block's vertices: SoA, component stride nv
    const double *x;
    long nv;
    int I, J, K;
};
// reference vertex index (2..imax+1 holds CGNS vertices 1..imax)
__device__ __forceinline__ double xv(const XB &X, int c, int i, int j, int k)
{
    return X.x[c * X.nv + (long)(i - 2) + (long)X.I * ((long)(j - 2) + (long)X.J * (long)(k - 2))];
}

struct Arr3 {               // a 1-based (n0,n1,n2) array of ncomp SoA components
    double *a;
    long cs;
    int ncomp, n0, n1, n2;
};
__device__ __host__ __forceinline__ long a3(const Arr3 &A, int i, int j, int k)
{
    return (long)(i - 1) + (long)A.n0 * ((long)(j - 1) + (long)A.n1 * (long)(k - 1));
}

// face vectors of axis AX over the cells/faces the reference loops cover
template <int AX>
__global__ __launch_bounds__(256) void k_face_vectors(XB X, Arr3 A)
{
    const int n0 = AX == 0 ? X.I : X.I - 1, n1 = AX == 1 ? X.J : X.J - 1, n2 = AX == 2 ? X.K : X.K - 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n0 * n1 * n2) return;
    const int i = (AX == 0 ? 1 : 2) + (int)(t % n0);
    const int j = (AX == 1 ? 1 : 2) + (int)((t / n0) % n1);
    const int k = (AX == 2 ? 1 : 2) + (int)(t / ((long)n0 * n1));
    double a[3], b[3];
    for (int c = 0; c < 3; c++) {
        if (AX == 0) {        // e = x(i+1,j+1,k) - x(i+1,j,k+1), s = x(i+1,j+1,k+1) - x(i+1,j,k)
            a[c] = xv(X, c, i + 1, j + 1, k) - xv(X, c, i + 1, j, k + 1);
            b[c] = xv(X, c, i + 1, j + 1, k + 1) - xv(X, c, i + 1, j, k);
        } else if (AX == 1) { // c = x(i,j+1,k) - x(i+1,j+1,k+1), s = x(i,j+1,k+1) - x(i+1,j+1,k)
            a[c] = xv(X, c, i, j + 1, k) - xv(X, c, i + 1, j + 1, k + 1);
            b[c] = xv(X, c, i, j + 1, k + 1) - xv(X, c, i + 1, j + 1, k);
        } else {              // c = x(i+1,j,k+1) - x(i,j+1,k+1), e = x(i+1,j+1,k+1) - x(i,j,k+1)
            a[c] = xv(X, c, i + 1, j, k + 1) - xv(X, c, i, j + 1, k + 1);
            b[c] = xv(X, c, i + 1, j + 1, k + 1) - xv(X, c, i, j, k + 1);
        }
    }
    const long o = a3(A, i, j, k);
    A.a[o] = 0.5 * (a[1] * b[2] - b[1] * a[2]);
    A.a[A.cs + o] = 0.5 * (b[0] * a[2] - a[0] * b[2]);
    A.a[2 * A.cs + o] = 0.5 * (a[0] * b[1] - b[0] * a[1]);
}

// one plane of an array: mode 0 copy from plane src, 1 2*src - src2 (src2 may
// pin the first in-plane index to fixu), 2 zero.  u/v are the two other axes
// in i, j, k order.
__global__ __launch_bounds__(256) void k_plane(Arr3 A, int ax, int dst, int src, int src2, int fixu, int mode,
                                               int ulo, int uhi, int vlo, int vhi)
{
    const int nu = uhi - ulo + 1, nvv = vhi - vlo + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nu <= 0 || nvv <= 0 || t >= (long)nu * nvv) return;
    const int u = ulo + (int)(t % nu), v = vlo + (int)(t / nu);
    const int ua = ax == 0 ? 1 : 0, va = ax == 2 ? 1 : 2;
    int pd[3], ps[3], p2[3];
    pd[ax] = dst; ps[ax] = src; p2[ax] = src2;
    pd[ua] = ps[ua] = u; p2[ua] = fixu > 0 ? fixu : u;
    pd[va] = ps[va] = p2[va] = v;
    const long od = a3(A, pd[0], pd[1], pd[2]), os = a3(A, ps[0], ps[1], ps[2]), o2 = a3(A, p2[0], p2[1], p2[2]);
    for (int c = 0; c < A.ncomp; c++) {
        double *a = A.a + c * A.cs;
        a[od] = mode == 2 ? 0.0 : (mode == 0 ? a[os] : 2.0 * a[os] - a[o2]);
    }
}

// centroid grd(2:4) and volume grd(1) of the interior cells
__device__ __forceinline__ double pyramid(const double (&p)[4][3], double xc, double yc, double zc)
{
    double cs[3], va[3], vb[3];
    for (int c = 0; c < 3; c++) {
        cs[c] = 0.25 * (p[0][c] + p[1][c] + p[2][c] + p[3][c]);
        va[c] = p[2][c] - p[0][c];
        vb[c] = p[3][c] - p[1][c];
    }
    double px = (xc - cs[0]) * (va[1] * vb[2] - va[2] * vb[1]);
    double py = (yc - cs[1]) * (va[2] * vb[0] - va[0] * vb[2]);
    double pz = (zc - cs[2]) * (va[0] * vb[1] - va[1] * vb[0]);
    return fabs(px + py + pz);
}

__global__ __launch_bounds__(256) void k_volumes(XB X, Arr3 vol, Arr3 xc)
{
    const int ni = X.I - 1, nj = X.J - 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ni * nj * (X.K - 1)) return;
    const int i = 2 + (int)(t % ni), j = 2 + (int)((t / ni) % nj), k = 2 + (int)(t / ((long)ni * nj));
    double g[3];
    for (int c = 0; c < 3; c++)
        g[c] = 0.125 * (xv(X, c, i, j, k) + xv(X, c, i + 1, j, k) + xv(X, c, i, j + 1, k) + xv(X, c, i, j, k + 1) +
                        xv(X, c, i + 1, j + 1, k) + xv(X, c, i, j + 1, k + 1) + xv(X, c, i + 1, j, k + 1) +
                        xv(X, c, i + 1, j + 1, k + 1));
    // the six faces in the reference's order and vertex order
    const unsigned char fc[6][4] = {{0, 1, 3, 2}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 3, 7, 5}};
    double vp[6];
    for (int f = 0; f < 6; f++) {
        double p[4][3];
        for (int v = 0; v < 4; v++) {
            const int b = fc[f][v];   // bit 0: +i, bit 1: +j, bit 2: +k
            for (int c = 0; c < 3; c++) p[v][c] = xv(X, c, i + (b & 1), j + ((b >> 1) & 1), k + (b >> 2));
        }
        vp[f] = pyramid(p, g[0], g[1], g[2]);
    }
    const long o = a3(vol, i, j, k);
    vol.a[o] = (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5]) / 6.0;
    for (int c = 0; c < 3; c++) xc.a[c * xc.cs + o] = g[c];
}

// wall-face centroids of one BC range, SoA [3][nw]
struct WSeg {
    int ax, fv, t1, t2;
    int lo1, hi1, lo2, hi2;
    long start;
};
__global__ __launch_bounds__(256) void k_wall_points(XB X, WSeg s, double *w, long nw)
{
    const int n1 = s.hi1 - s.lo1 + 1, n2 = s.hi2 - s.lo2 + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n1 * n2) return;
    const int u1 = s.lo1 + (int)(t % n1), u2 = s.lo2 + (int)(t / n1);
    for (int c = 0; c < 3; c++) {
        double acc = 0.0;
        for (int v = 0; v < 4; v++) {
            int p[3];
            p[s.ax] = s.fv;
            p[s.t1] = u1 + (v & 1);
            p[s.t2] = u2 + (v >> 1);
            const double q = xv(X, c, p[0], p[1], p[2]);
            acc = v == 0 ? q : acc + q;
        }
        w[c * nw + s.start + t] = 0.25 * acc;
    }
}

constexpr int WTILE = 1024;   // wall points per LDS tile (24 KB)

// distance of the interior cells of one block to the nearest wall point:
// min over points of (dx*dx + dy*dy) + dz*dz, then one square root (the
// square root is monotone, so this equals the minimum of the square roots),
// capped at 1e6 as the reference's initial value
__global__ __launch_bounds__(256) void k_wall_distance(Arr3 xc, int I, int J, int K, const double *w, long nw, Arr3 yd)
{
    __shared__ double sx[WTILE], sy[WTILE], sz[WTILE];
    const int ni = I - 1, nj = J - 1;
    const long ncell = (long)ni * nj * (K - 1);
    long cid[2] = {(long)blockIdx.x * 512 + threadIdx.x, (long)blockIdx.x * 512 + 256 + threadIdx.x};
    double p[2][3], m[2];
    long o[2];
    for (int h = 0; h < 2; h++) {
        const long t = cid[h] < ncell ? cid[h] : 0;
        const int i = 2 + (int)(t % ni), j = 2 + (int)((t / ni) % nj), k = 2 + (int)(t / ((long)ni * nj));
        o[h] = a3(xc, i, j, k);
        for (int c = 0; c < 3; c++) p[h][c] = xc.a[c * xc.cs + o[h]];
        m[h] = __builtin_inf();
    }
    for (long t0 = 0; t0 < nw; t0 += WTILE) {
        const int n = (int)((nw - t0) < WTILE ? (nw - t0) : WTILE);
        __syncthreads();
        for (int q = threadIdx.x; q < n; q += 256) {
            sx[q] = w[t0 + q];
            sy[q] = w[nw + t0 + q];
            sz[q] = w[2 * nw + t0 + q];
        }
        __syncthreads();
#pragma unroll 4
        for (int q = 0; q < n; q++) {
            const double wx = sx[q], wy = sy[q], wz = sz[q];
            for (int h = 0; h < 2; h++) {
                const double dx = p[h][0] - wx, dy = p[h][1] - wy, dz = p[h][2] - wz;
                const double d = dx * dx + dy * dy + dz * dz;
                m[h] = fmin(m[h], d);
            }
        }
    }
    for (int h = 0; h < 2; h++) {
        if (cid[h] >= ncell) continue;
        const double s = sqrt(m[h]);
        yd.a[o[h]] = s < 1.0e6 ? s : 1.0e6;
    }
}

// one interface exchange: what 0 tangential face vectors, 1 volume, 2 wall distance
struct XOp {
    int src, dst, what, m, q;
    int is[3], ie[3], isd[3];
    int tm[9];
    double s1, s2;
    long n;
};

__device__ __forceinline__ double *fv_at(const DS &S, const BlkD &B, int ax, int i, int j, int k, long &cs)
{
    if (ax == 0) { cs = S.NCX; return S.cx + B.cx0 + (i - 1) + (long)B.I * ((j - 1) + (long)(B.J + 1) * (k - 1)); }
    if (ax == 1) { cs = S.NEX; return S.ex + B.ex0 + (i - 1) + (long)(B.I + 1) * ((j - 1) + (long)B.J * (k - 1)); }
    cs = S.NTX;
    return S.tx + B.tx0 + (i - 1) + (long)(B.I + 1) * ((j - 1) + (long)(B.J + 1) * (k - 1));
}

__global__ __launch_bounds__(256) void k_xpack(DS S, XOp op, double *buf)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= op.n) return;
    const int ni = op.ie[0] - op.is[0] + 1, nj = op.ie[1] - op.is[1] + 1;
    const int i = op.is[0] + (int)(t % ni), j = op.is[1] + (int)((t / ni) % nj), k = op.is[2] + (int)(t / ((long)ni * nj));
    const BlkD &B = S.b[op.src];
    if (op.what == 0) {
        long c1, c2;
        const double *a = fv_at(S, B, (op.m + 1) % 3, i, j, k, c1), *b = fv_at(S, B, (op.m + 2) % 3, i, j, k, c2);
        for (int m = 0; m < 3; m++) {
            buf[6 * t + 2 * m] = a[m * c1];
            buf[6 * t + 2 * m + 1] = b[m * c2];
        }
    } else {
        const long w = B.w0 + (i - 1) + (long)(B.I + 1) * ((j - 1) + (long)(B.J + 1) * (k - 1));
        buf[t] = op.what == 1 ? S.vol[w] : S.ydn[w];
    }
}

__global__ __launch_bounds__(256) void k_xunpack(DS S, XOp op, const double *buf)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= op.n) return;
    const int ni = op.ie[0] - op.is[0] + 1, nj = op.ie[1] - op.is[1] + 1;
    const int di = (int)(t % ni), dj = (int)((t / ni) % nj), dk = (int)(t / ((long)ni * nj));
    const int ii = op.tm[0] * di + op.tm[1] * dj + op.tm[2] * dk + op.isd[0];
    const int jj = op.tm[3] * di + op.tm[4] * dj + op.tm[5] * dk + op.isd[1];
    const int kk = op.tm[6] * di + op.tm[7] * dj + op.tm[8] * dk + op.isd[2];
    const BlkD &B = S.b[op.dst];
    if (op.what == 0) {
        long c1, c2;
        double *a = fv_at(S, B, (op.q + 1) % 3, ii, jj, kk, c1), *b = fv_at(S, B, (op.q + 2) % 3, ii, jj, kk, c2);
        for (int m = 0; m < 3; m++) {
            a[m * c1] = op.s1 * buf[6 * t + 2 * m];
            b[m * c2] = op.s2 * buf[6 * t + 2 * m + 1];
        }
    } else {
        const long w = B.w0 + (ii - 1) + (long)(B.I + 1) * ((jj - 1) + (long)(B.J + 1) * (kk - 1));
        (op.what == 1 ? S.vol : S.ydn)[w] = buf[t];
    }
}

__global__ __launch_bounds__(256) void k_grd_to_aos(const double *vol, const double *xc, const double *yd, int ngrd,
                                                    long cs, long w0, double *dst, long n)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    dst[g * ngrd] = vol[w0 + g];
    for (int c = 0; c < 3; c++) dst[g * ngrd + 1 + c] = xc[c * cs + w0 + g];
    if (ngrd > 4) dst[g * ngrd + 4] = yd ? yd[w0 + g] : 0.0;
}

// ------------------------------------------------------------ host side
XB xblock(const cav3d_ctx *c, int b)
{
    const BlkD &B = c->blk[b];
    return XB{c->xv + c->v0[b], c->NV, B.I, B.J, B.K};
}
Arr3 arr_cx(const cav3d_ctx *c, int b) { const BlkD &B = c->blk[b]; return Arr3{c->S.cx + B.cx0, c->S.NCX, 3, B.I, B.J + 1, B.K + 1}; }
Arr3 arr_ex(const cav3d_ctx *c, int b) { const BlkD &B = c->blk[b]; return Arr3{c->S.ex + B.ex0, c->S.NEX, 3, B.I + 1, B.J, B.K + 1}; }
Arr3 arr_tx(const cav3d_ctx *c, int b) { const BlkD &B = c->blk[b]; return Arr3{c->S.tx + B.tx0, c->S.NTX, 3, B.I + 1, B.J + 1, B.K}; }
Arr3 arr_w(const cav3d_ctx *c, double *base, int ncomp, int b)
{
    const BlkD &B = c->blk[b];
    return Arr3{base + B.w0, c->S.NW, ncomp, B.I + 1, B.J + 1, B.K + 1};
}

void plane(cav3d_ctx *c, const Arr3 &A, int ax, int dst, int src, int src2, int fixu, int mode, int ulo, int uhi,
           int vlo, int vhi)
{
    long n = (long)std::max(0, uhi - ulo + 1) * std::max(0, vhi - vlo + 1);
    if (n == 0) return;
    hipLaunchKernelGGL(k_plane, dim3(nblk(n, 256)), dim3(256), 0, c->st, A, ax, dst, src, src2, fixu, mode, ulo, uhi, vlo, vhi);
}

// The one-layer ranges of a connection from the extended 22-int form
// (src/grid_module.f90:228-300 conversion, :924-938 re-ordering, the
// src/bc_module.f90:355-488 extension undone); what 0 gets calnormal's
// vertex-index ranges (before the re-ordering).
int xop_from(const cav3d_ctx *c, int b, const int *d, int what, XOp &op)
{
    memset(&op, 0, sizeof(op));
    op.src = b;
    op.dst = d[0];
    op.what = what;
    int is[3], ie[3], isd[3], ied[3];
    for (int r = 0; r < 3; r++) { is[r] = d[1 + r]; ie[r] = d[4 + r]; isd[r] = d[7 + r]; ied[r] = d[10 + r]; }
    for (int r = 0; r < 9; r++) op.tm[r] = d[13 + r];
    const BlkD &A = c->blk[b], &D = c->blk[op.dst];
    const int da[3] = {A.I, A.J, A.K}, dd[3] = {D.I, D.J, D.K};
    int q = -1, m = -1;
    for (int r = 0; r < 3 && q < 0; r++)
        if (isd[r] <= 1 || isd[r] >= dd[r] + 1 || ied[r] <= 1 || ied[r] >= dd[r] + 1) q = r;
    if (q < 0) return fail("interface of block " + std::to_string(b) + ": no ghost layer on the donor side", -2);
    for (int l = 0; l < 3; l++) if (op.tm[3 * q + l] != 0) m = l;
    if (m < 0) return fail("interface of block " + std::to_string(b) + ": singular transform", -2);
    const bool lead = isd[q] == 1 || isd[q] == dd[q] + 1;
    const int orig = lead ? isd[q] : ied[q], face = lead ? is[m] : ie[m];
    if (lead ? is[m] != 2 : ie[m] != da[m])
        return fail("interface of block " + std::to_string(b) + " is not three layers at a block face", -2);
    isd[q] = ied[q] = orig;
    is[m] = ie[m] = face;
    op.m = m;
    op.q = q;
    op.s1 = op.s2 = 1.0;
    for (int r = 0; r < 3; r++) {
        op.is[r] = is[r];
        op.ie[r] = ie[r];
        op.isd[r] = isd[r];
        if (what == 0 && r != m) op.is[r] = is[r] - 1;
        if (what == 0 && r != q && isd[r] < ied[r]) op.isd[r] = isd[r] - 1;
    }
    if (what == 0) {
        // calnormal's signs follow the donor's tangential ranges (src/grid_module.f90:782-807)
        const int t1 = (q + 1) % 3, t2 = (q + 2) % 3;
        auto desc = [&](int r) {
            int s = isd[r], e = ied[r];
            if (s < e) s -= 1; else if (s > e) e -= 1;
            return s > e;
        };
        if (desc(t1)) op.s1 = -1.0;
        if (desc(t2)) op.s2 = -1.0;
    }
    op.n = 1;
    for (int r = 0; r < 3; r++) op.n *= op.ie[r] >= op.is[r] ? op.ie[r] - op.is[r] + 1 : 0;
    // every read and write must land inside the two blocks' metric arrays:
    // face-vector array ax of a block with vertex counts n spans
    // (n0 + (ax != 0), n1 + (ax != 1), n2 + (ax != 2)), cell arrays n + 1
    auto hi = [&](const int *n, int ax, int r) { return ax < 0 ? n[r] + 1 : n[r] + (ax != r ? 1 : 0); };
    const int rd1 = what == 0 ? (m + 1) % 3 : -1, rd2 = what == 0 ? (m + 2) % 3 : -1;
    const int wr1 = what == 0 ? (q + 1) % 3 : -1, wr2 = what == 0 ? (q + 2) % 3 : -1;
    for (int corner = 0; op.n > 0 && corner < 8; corner++) {
        int del[3], o[3], s[3];
        for (int r = 0; r < 3; r++) del[r] = ((corner >> r) & 1) ? op.ie[r] - op.is[r] : 0;
        for (int r = 0; r < 3; r++) {
            s[r] = op.is[r] + del[r];
            o[r] = op.tm[3 * r] * del[0] + op.tm[3 * r + 1] * del[1] + op.tm[3 * r + 2] * del[2] + op.isd[r];
        }
        for (int r = 0; r < 3; r++) {
            if (o[r] < 1 || o[r] > std::min(hi(dd, wr1, r), hi(dd, wr2, r)))
                return fail("interface of block " + std::to_string(b) + " maps outside the donor's metric arrays", -2);
            if (s[r] < 1 || s[r] > std::min(hi(da, rd1, r), hi(da, rd2, r)))
                return fail("interface of block " + std::to_string(b) + " reads outside its metric arrays", -2);
        }
    }
    return 0;
}

int exchange_metrics(cav3d_ctx *c, int what, double *buf)
{
    std::vector<XOp> ops;
    std::vector<long> at;
    long tot = 0;
    for (int b = 0; b < c->nb; b++) {
        const std::vector<int> &cd = c->condesc[b];
        for (size_t n = 0; n + 22 <= cd.size(); n += 22) {
            XOp op;
            int rc = xop_from(c, b, cd.data() + n, what, op);
            if (rc) return rc;
            if (op.n == 0) continue;
            ops.push_back(op);
            at.push_back(tot);
            tot += op.n * (what == 0 ? 6 : 1);
        }
    }
    if (tot > c->xbuf_n) return fail("exchange buffer too small", -3);
    for (size_t n = 0; n < ops.size(); n++) {
        hipLaunchKernelGGL(k_xpack, dim3(nblk(ops[n].n, 256)), dim3(256), 0, c->st, c->S, ops[n], buf + at[n]);
        if (ops[n].src == ops[n].dst)
            hipLaunchKernelGGL(k_xunpack, dim3(nblk(ops[n].n, 256)), dim3(256), 0, c->st, c->S, ops[n], (const double *)(buf + at[n]));
    }
    for (size_t n = 0; n < ops.size(); n++)
        if (ops[n].src != ops[n].dst)
            hipLaunchKernelGGL(k_xunpack, dim3(nblk(ops[n].n, 256)), dim3(256), 0, c->st, c->S, ops[n], (const double *)(buf + at[n]));
    return 0;
}

long exchange_size(const cav3d_ctx *c)
{
    long tot = 0;
    for (int b = 0; b < c->nb; b++)
        for (size_t n = 0; n + 22 <= c->condesc[b].size(); n += 22) {
            const int *d = c->condesc[b].data() + n;
            long cnt = 1;
            for (int r = 0; r < 3; r++) cnt *= std::abs(d[4 + r] - d[1 + r]) + 2;
            tot += 6 * cnt;
        }
    return tot;
}

std::vector<WSeg> wall_segments(const cav3d_ctx *c, int b, long &start, bool &ok)
{
    std::vector<WSeg> out;
    const BlkD &B = c->blk[b];
    const int dim[3] = {B.I, B.J, B.K};
    const std::vector<int> &dsc = c->bcdesc[b];
    for (size_t n = 0; n + 7 <= dsc.size(); n += 7) {
        if ((dsc[n] & CAV3D_BC_KIND_MASK) != CAV3D_BC_WALL) continue;
        const int *s = &dsc[n + 1], *e = &dsc[n + 4];
        for (int ax = 0; ax < 3; ax++) {
            if (s[ax] != e[ax]) continue;
            WSeg w;
            w.ax = ax;
            w.fv = s[ax] == 1 ? 2 : dim[ax] + 1;
            w.t1 = ax == 0 ? 1 : 0;
            w.t2 = ax == 2 ? 1 : 2;
            w.lo1 = s[w.t1]; w.hi1 = e[w.t1]; w.lo2 = s[w.t2]; w.hi2 = e[w.t2];
            if (w.lo1 < 2 || w.hi1 > dim[w.t1] || w.lo2 < 2 || w.hi2 > dim[w.t2]) { ok = false; continue; }
            w.start = start;
            long cnt = (long)std::max(0, w.hi1 - w.lo1 + 1) * std::max(0, w.hi2 - w.lo2 + 1);
            start += cnt;
            out.push_back(w);
        }
    }
    return out;
}

}  // namespace

extern "C" {

int cav3d_set_grid(cav3d_ctx *c, int b, const double *x)
{
    if (!c || b < 0 || b >= c->nb || !x) return fail("bad argument to cav3d_set_grid", -1);
    int rc;
    if (!c->xv) {
        long nv = 0;
        c->v0.assign(c->nb, 0);
        for (int q = 0; q < c->nb; q++) { c->v0[q] = nv; nv += (long)c->blk[q].I * c->blk[q].J * c->blk[q].K; }
        c->NV = nv;
        if ((rc = dalloc(c, &c->xv, 3 * nv))) return rc;
        c->has_x.assign(c->nb, 0);
    }
    const BlkD &B = c->blk[b];
    const long n = (long)B.I * B.J * B.K;
    if ((rc = ensure_stage(c, 3 * n))) return rc;
    HIPCHK(hipMemcpyAsync(c->stage, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(n, 256)), dim3(256), 0, c->st, (const double *)c->stage, 3, 3, 0, c->xv, c->NV, c->v0[b], n);
    HIPCHK(hipStreamSynchronize(c->st));
    c->has_x[b] = 1;
    return check_launch("cav3d_set_grid");
}

int cav3d_build_metrics(cav3d_ctx *c)
{
    if (!c) return fail("null context", -1);
    for (int b = 0; b < c->nb; b++)
        if (!c->xv || !c->has_x[b]) return fail("cav3d_build_metrics: block " + std::to_string(b) + " has no coordinates", -1);
    int rc;
    long need = exchange_size(c);
    if (need > c->xbuf_n) {
        if (c->xbuf) { HIPCHK(hipStreamSynchronize(c->st)); HIPCHK(hipFree(c->xbuf)); c->xbuf = nullptr; }
        HIPCHK(hipMalloc((void **)&c->xbuf, need * sizeof(double)));
        c->xbuf_n = need;
    }
    // calnormal
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        const int I = B.I, J = B.J, K = B.K;
        XB X = xblock(c, b);
        Arr3 cx = arr_cx(c, b), ex = arr_ex(c, b), tx = arr_tx(c, b);
        hipLaunchKernelGGL(k_face_vectors<0>, dim3(nblk((long)I * (J - 1) * (K - 1), 256)), dim3(256), 0, c->st, X, cx);
        hipLaunchKernelGGL(k_face_vectors<1>, dim3(nblk((long)(I - 1) * J * (K - 1), 256)), dim3(256), 0, c->st, X, ex);
        hipLaunchKernelGGL(k_face_vectors<2>, dim3(nblk((long)(I - 1) * (J - 1) * K, 256)), dim3(256), 0, c->st, X, tx);
        const std::vector<int> &dsc = c->bcdesc[b];
        for (size_t n = 0; n + 7 <= dsc.size(); n += 7) {
            if ((dsc[n] & CAV3D_BC_KIND_MASK) != CAV3D_BC_DEGENERATE) continue;
            const int *s = &dsc[n + 1], *e = &dsc[n + 4];
            if (s[0] == e[0]) plane(c, cx, 0, s[0] == 1 ? 1 : I, 0, 0, 0, 2, 1, J + 1, 1, K + 1);
            if (s[1] == e[1]) plane(c, ex, 1, s[1] == 1 ? 1 : J, 0, 0, 0, 2, 1, I + 1, 1, K + 1);
            if (s[2] == e[2]) plane(c, tx, 2, s[2] == 1 ? 1 : K, 0, 0, 0, 2, 1, I + 1, 1, J + 1);
        }
        // ghost copies, statement by statement (src/grid_module.f90:710-723)
        plane(c, cx, 1, 1, 2, 0, 0, 0, 1, I, 1, K + 1);
        plane(c, cx, 1, J + 1, J, 0, 0, 0, 1, I, 1, K + 1);
        plane(c, cx, 2, 1, 2, 0, 0, 0, 1, I, 1, J + 1);
        plane(c, cx, 2, K + 1, K, 0, 0, 0, 1, I, 1, J + 1);
        plane(c, ex, 0, 1, 2, 0, 0, 0, 1, J, 1, K + 1);
        plane(c, ex, 0, I + 1, I, 0, 0, 0, 1, J, 1, K + 1);
        plane(c, ex, 2, 1, 2, 0, 0, 0, 1, I + 1, 1, J);
        plane(c, ex, 2, K + 1, K, 0, 0, 0, 1, I + 1, 1, J);
        plane(c, tx, 0, 1, 2, 0, 0, 0, 1, J + 1, 1, K);
        plane(c, tx, 0, I + 1, I, 0, 0, 0, 1, J + 1, 1, K);
        plane(c, tx, 1, 1, 2, 0, 0, 0, 1, I + 1, 1, K);
        plane(c, tx, 1, J + 1, J, 0, 0, 0, 1, I + 1, 1, K);
    }
    if ((rc = exchange_metrics(c, 0, c->xbuf))) return rc;
    // calvolume
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        const int I = B.I, J = B.J, K = B.K;
        Arr3 vol = arr_w(c, c->S.vol, 1, b), xc = arr_w(c, c->xc, 3, b);
        hipLaunchKernelGGL(k_volumes, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, xblock(c, b), vol, xc);
        // ghost fill (src/grid_module.f90:1073-1113): volumes copied, centroids
        // extrapolated, kmin against grd(:,3,j,3) as written
        plane(c, vol, 1, 1, 2, 0, 0, 0, 2, I, 2, K);
        plane(c, xc, 1, 1, 2, 3, 0, 1, 2, I, 2, K);
        plane(c, vol, 1, J + 1, J, 0, 0, 0, 2, I, 2, K);
        plane(c, xc, 1, J + 1, J, J - 1, 0, 1, 2, I, 2, K);
        plane(c, vol, 0, 1, 2, 0, 0, 0, 1, J + 1, 2, K);
        plane(c, xc, 0, 1, 2, 3, 0, 1, 1, J + 1, 2, K);
        plane(c, vol, 0, I + 1, I, 0, 0, 0, 1, J + 1, 2, K);
        plane(c, xc, 0, I + 1, I, I - 1, 0, 1, 1, J + 1, 2, K);
        plane(c, vol, 2, 1, 2, 0, 0, 0, 1, I + 1, 1, J + 1);
        plane(c, xc, 2, 1, 2, 3, 3, 1, 1, I + 1, 1, J + 1);
        plane(c, vol, 2, K + 1, K, 0, 0, 0, 1, I + 1, 1, J + 1);
        plane(c, xc, 2, K + 1, K, K - 1, 0, 1, 1, I + 1, 1, J + 1);
    }
    if ((rc = exchange_metrics(c, 1, c->xbuf))) return rc;
    // calydns: turbulent runs only (ngrd = 5)
    if (c->S.ydn) {
        long nw = 0;
        std::vector<std::vector<WSeg>> segs(c->nb);
        bool ok = true;
        for (int b = 0; b < c->nb; b++) segs[b] = wall_segments(c, b, nw, ok);
        if (!ok) return fail("cav3d_build_metrics: a BCWall range lies outside its block's cells", -2);
        double *w = nullptr;
        HIPCHK(hipMalloc((void **)&w, 3 * std::max(nw, 1L) * sizeof(double)));
        for (int b = 0; b < c->nb; b++)
            for (const WSeg &s : segs[b]) {
                long n = (long)(s.hi1 - s.lo1 + 1) * (s.hi2 - s.lo2 + 1);
                if (n > 0) hipLaunchKernelGGL(k_wall_points, dim3(nblk(n, 256)), dim3(256), 0, c->st, xblock(c, b), s, w, nw);
            }
        for (int b = 0; b < c->nb; b++) {
            const BlkD &B = c->blk[b];
            const int I = B.I, J = B.J, K = B.K;
            Arr3 yd = arr_w(c, c->S.ydn, 1, b);
            hipLaunchKernelGGL(k_wall_distance, dim3(nblk(B.nc, 512)), dim3(256), 0, c->st, arr_w(c, c->xc, 3, b), I, J, K,
                               (const double *)w, nw, yd);
            plane(c, yd, 1, 1, 2, 0, 0, 0, 2, I, 2, K);
            plane(c, yd, 1, J + 1, J, 0, 0, 0, 2, I, 2, K);
            plane(c, yd, 0, 1, 2, 0, 0, 0, 1, J + 1, 2, K);
            plane(c, yd, 0, I + 1, I, 0, 0, 0, 1, J + 1, 2, K);
            plane(c, yd, 2, 1, 2, 0, 0, 0, 1, I + 1, 1, J + 1);
            plane(c, yd, 2, K + 1, K, 0, 0, 0, 1, I + 1, 1, J + 1);
        }
        hipError_t e = hipStreamSynchronize(c->st);
        hipFree(w);
        if (e != hipSuccess) return fail(std::string("cav3d_build_metrics: ") + hipGetErrorString(e), -3);
        if ((rc = exchange_metrics(c, 2, c->xbuf))) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    return check_launch("cav3d_build_metrics");
}

int cav3d_get_metrics(cav3d_ctx *c, int b, double *cx, double *ex, double *tx, double *grd)
{
    if (!c || b < 0 || b >= c->nb) return fail("bad argument to cav3d_get_metrics", -1);
    const BlkD &B = c->blk[b];
    const DS &S = c->S;
    int rc;
    long nmax = std::max(std::max(3 * B.ncx, 3 * B.nex), std::max(3 * B.ntx, (long)c->ngrd * B.nw));
    if ((rc = ensure_stage(c, nmax))) return rc;
    struct { double *h; long n; const double *d; long nst; long o; } parts[3] = {
        {cx, B.ncx, S.cx, S.NCX, B.cx0}, {ex, B.nex, S.ex, S.NEX, B.ex0}, {tx, B.ntx, S.tx, S.NTX, B.tx0}};
    for (auto &p : parts) {
        if (!p.h) continue;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(p.n, 256)), dim3(256), 0, c->st, p.d, 3, p.nst, p.o, c->stage, p.n);
        HIPCHK(hipMemcpyAsync(p.h, c->stage, 3 * p.n * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    if (grd) {
        hipLaunchKernelGGL(k_grd_to_aos, dim3(nblk(B.nw, 256)), dim3(256), 0, c->st, (const double *)S.vol,
                           (const double *)c->xc, (const double *)S.ydn, c->ngrd, S.NW, B.w0, c->stage, B.nw);
        HIPCHK(hipMemcpyAsync(grd, c->stage, (long)c->ngrd * B.nw * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    return check_launch("cav3d_get_metrics");
}

}  // extern "C"

// Post integrals and checkpoint files of the device-resident state.
//
//   cav3d_integrals  vapour and gas volume, the sum over interior cells of the
//                    vof fields rho*y1/rho_v and rho*y2/rho_g
//                    (src/postvariable_module.f90:282-295) times the cell
//                    volume; pressure and viscous force over the boundary
//                    ranges flagged CAV3D_BC_SURFACE, as clcd_writing sums them
//                    (src/postvariable_module.f90:507-633), with cl and cd
//   cav3d_export /   the out_<iter>.dat restart file: export_variable
//   cav3d_import     (src/variable_module.f90:108-182) and the restart reader
//                    (src/initial_module.f90:245-389)
//
// Both integrals are reductions: per-workgroup partial sums in a fixed tree
// order, added on the host in workgroup order, so a result is reproducible
// run to run (it differs from a sequential sum by rounding only).
//
// Included once, from cav3d.hip, after metrics.hpp.

namespace {

constexpr int POST_WG = 256;

// block-wide sum of NV values per thread into part[blockIdx][NV], fixed order
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double *part)
{
    __shared__ double sh[NV][POST_WG];
    for (int q = 0; q < NV; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = POST_WG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < NV; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; q++) part[(long)blockIdx.x * NV + q] = sh[q][0];
}

// vapour and gas volume of the interior cells of every block
__global__ __launch_bounds__(POST_WG) void k_post_cells(DS S, const double *vol, double *part)
{
    long t = (long)blockIdx.x * POST_WG + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < S.NC) {
        const int b = find_blk(S, t, 0);
        const BlkD &B = S.b[b];
        int i, j, k;
        cell_of(B, t - B.c0, i, j, k);
        const long g = gidx(B, i, j, k);
        double d[ND], full[18];
        mixture_eos<double>(S.pv[g] + S.P.pref, S.pv[4 * S.NG + g], S.pv[5 * S.NG + g], S.pv[6 * S.NG + g], d,
                            false, full);
        const double w = vol[widx(B, i, j, k)];
        v[0] = full[0] * S.pv[5 * S.NG + g] / full[3] * w;
        v[1] = full[0] * S.pv[6 * S.NG + g] / full[4] * w;
    }
    wg_sum<2>(v, part);
}

struct FSeg {        // one surface boundary range
    int b, ax, side; // side 0: min face, 1: max face
    int lo[3], hi[3];
    long start, n;
};

// pressure and viscous force of every surface face
__global__ __launch_bounds__(POST_WG) void k_post_faces(DS S, const FSeg *seg, int nseg, long nface, const double *xvert,
                                                        long nv, const long *v0, const double *xc, double *part)
{
    long t = (long)blockIdx.x * POST_WG + threadIdx.x;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < nface) {
        int q = 0;
        for (int r = 1; r < nseg; r++) if (t >= seg[r].start) q = r;
        const FSeg &s = seg[q];
        const BlkD &B = S.b[s.b];
        const long r = t - s.start;
        const int n0 = s.hi[0] - s.lo[0] + 1, n1 = s.hi[1] - s.lo[1] + 1;
        const int p[3] = {s.lo[0] + (int)(r % n0), s.lo[1] + (int)((r / n0) % n1), s.lo[2] + (int)(r / ((long)n0 * n1))};
        const int dim[3] = {B.I, B.J, B.K};
        const int ax = s.ax;
        // the face vector (outward from the boundary into the cell for a min
        // face, the negated max-face vector otherwise), the interior cell and
        // the vertex plane of the face
        int fi[3] = {p[0], p[1], p[2]}, ci[3] = {p[0], p[1], p[2]};
        fi[ax] = s.side ? dim[ax] : 1;
        ci[ax] = s.side ? dim[ax] : 2;
        const int vplane = s.side ? dim[ax] + 1 : 2;
        double cxv[3];
        face_vec(S, B, ax, fi[0], fi[1], fi[2], cxv);
        if (s.side) for (int c = 0; c < 3; c++) cxv[c] = -cxv[c];
        const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
        const XB X{xvert + v0[s.b], nv, B.I, B.J, B.K};
        double x[3];
        for (int c = 0; c < 3; c++) {
            double acc = 0.0;
            for (int m = 0; m < 4; m++) {
                int u[3];
                u[ax] = vplane;
                u[t1] = p[t1] + (m & 1);
                u[t2] = p[t2] + (m >> 1);
                const double a = xv(X, c, u[0], u[1], u[2]);
                acc = m == 0 ? a : acc + a;
            }
            x[c] = 0.25 * acc;
        }
        const long g = gidx(B, ci[0], ci[1], ci[2]);
        const long w = widx(B, ci[0], ci[1], ci[2]);
        const double pg = S.pv[g], u = S.pv[S.NG + g], vv = S.pv[2 * S.NG + g], ww = S.pv[3 * S.NG + g];
        v[0] = -pg * cxv[0];
        v[1] = -pg * cxv[1];
        v[2] = -pg * cxv[2];
        const double dl = cxv[0] * cxv[0] + cxv[1] * cxv[1] + cxv[2] * cxv[2];
        const double vel = cxv[0] * u + cxv[1] * vv + cxv[2] * ww;
        double vis = 0.0;
        if (S.P.iturb == 0) vis = S.dc[D_MU * S.NG + g] + S.mut[g];
        else if (S.P.iturb == -2) vis = S.dc[D_MU * S.NG + g];
        const double gx = xc[w], gy = xc[S.NW + w], gz = xc[2 * S.NW + w];
        const double dist = sqrt((gx - x[0]) * (gx - x[0]) + (gy - x[1]) * (gy - x[1]) + (gz - x[2]) * (gz - x[2]));
        const double sdl = sqrt(dl);
        const double uu[3] = {u, vv, ww};
        for (int c = 0; c < 3; c++) {
            const double ut = uu[c] - cxv[c] * vel / dl;
            v[3 + c] = vis * ut / dist * sdl;
        }
    }
    wg_sum<6>(v, part);
}

__global__ __launch_bounds__(256) void k_interior_from_aos(const double *src, int m, double *dst, long nstride, BlkD B)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B.nc) return;
    int i, j, k;
    cell_of(B, q, i, j, k);
    const long g = gidx(B, i, j, k);
    for (int a = 0; a < m; a++) dst[a * nstride + g] = src[(g - B.g0) * m + a];
}

// the CAV3D_BC_SURFACE ranges of every block, in block and list order, handed
// to push(FSeg) (empty ranges skipped); nface counts their faces
template <class Push>
void scan_surface(const cav3d_ctx *c, long &nface, bool &ok, Push push)
{
    nface = 0;
    ok = true;
    for (int b = 0; b < c->nb; b++) {
        const BlkD &B = c->blk[b];
        const int dim[3] = {B.I, B.J, B.K};
        const std::vector<int> &dsc = c->bcdesc[b];
        for (size_t n = 0; n + 7 <= dsc.size(); n += 7) {
            if (!(dsc[n] & CAV3D_BC_SURFACE)) continue;
            FSeg s;
            memset(&s, 0, sizeof(s));
            s.b = b;
            s.ax = -1;
            for (int a = 0; a < 3; a++) { s.lo[a] = dsc[n + 1 + a]; s.hi[a] = dsc[n + 4 + a]; }
            for (int a = 0; a < 3 && s.ax < 0; a++) if (s.lo[a] == s.hi[a]) s.ax = a;
            if (s.ax < 0) { ok = false; continue; }
            s.side = s.lo[s.ax] == 1 ? 0 : 1;
            if (s.side && s.lo[s.ax] != dim[s.ax] + 1) { ok = false; continue; }
            for (int a = 0; a < 3; a++)
                if (a != s.ax && (s.lo[a] < 2 || s.hi[a] > dim[a])) ok = false;
            s.n = 1;
            for (int a = 0; a < 3; a++) s.n *= std::max(0, s.hi[a] - s.lo[a] + 1);
            s.start = nface;
            nface += s.n;
            if (s.n > 0) push(s);
        }
    }
}

std::vector<FSeg> surface_segments(const cav3d_ctx *c, long &nface, bool &ok)
{
    std::vector<FSeg> out;
    scan_surface(c, nface, ok, [&](const FSeg &s) { out.push_back(s); });
    return out;
}

// the integrals' scaling of the eight sums: vapour and gas volume, pressure
// and viscous force coefficients, cl and cd (clcd_writing)
void integrals_out(const double *acc, double rhoref, double uref, double area, double aoa, double *out)
{
    const double ref = 2.0 / rhoref / (uref * uref) / area;
    out[0] = acc[0];
    out[1] = acc[1];
    for (int a = 0; a < 6; a++) out[2 + a] = acc[2 + a] * ref;
    const double fx = out[2] + out[5], fy = out[3] + out[6];
    out[8] = fy * cos(aoa) - fx * sin(aoa);
    out[9] = fy * sin(aoa) + fx * cos(aoa);
}

// ---- the allocation-free integrals (cav3d_integrals_noalloc)
// Its device memory is one file-scope array per GPU, sized once here:
//   [0, 4096)       the surface-segment table (at most NOALLOC_MAXSEG ranges)
//   [4096, 4352)    the per-block vertex offsets
//   [4352, 4608)    the eight final sums
//   [4608, end)     the per-workgroup partial sums (2 per cell workgroup,
//                   6 per surface-face workgroup)
// Every call on a GPU uses that GPU's array and leaves its contents behind;
// each call writes every entry it reads before reading it, so nothing needs
// clearing, and calls that overlap on one GPU race.
constexpr long NOALLOC_SCRATCH_BYTES = 2L << 20;
constexpr int NOALLOC_MAXSEG = 64;
constexpr long NOALLOC_SEG_OFF = 0, NOALLOC_V0_OFF = 4096, NOALLOC_SUM_OFF = 4352, NOALLOC_PART_OFF = 4608;
constexpr long NOALLOC_PART_CAP = (NOALLOC_SCRATCH_BYTES - NOALLOC_PART_OFF) / (long)sizeof(double);
static_assert(sizeof(FSeg) * NOALLOC_MAXSEG <= NOALLOC_V0_OFF - NOALLOC_SEG_OFF, "segment table fits");
static_assert(sizeof(long) * MAXB <= NOALLOC_SUM_OFF - NOALLOC_V0_OFF, "vertex offsets fit");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// device address of this GPU's scratch, looked up once per device with the
// device current (the array's host-side name is not a device address)
constexpr int NOALLOC_MAXDEV = 64;
unsigned char *g_scratch_ptr[NOALLOC_MAXDEV] = {};

int scratch_of(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return fail("cav3d_integrals_noalloc: device ordinal out of range", -1);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(std::string("cav3d_integrals_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), -3);
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// rows of W partial sums added in row order into acc (thread 0), staged
// through LDS in chunks of 256 rows by the whole workgroup
template <int W>
__device__ __forceinline__ void seq_rows(const double *p, long rows, double *acc, double *buf)
{
    for (long r0 = 0; r0 < rows; r0 += 256) {
        const int nr = (int)((rows - r0) < 256 ? (rows - r0) : 256);
        __syncthreads();
        for (int q = threadIdx.x; q < nr * W; q += blockDim.x) buf[q] = p[r0 * W + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int r = 0; r < nr; r++)
#pragma unroll
                for (int a = 0; a < W; a++) acc[a] += buf[r * W + a];
    }
}

// the eight sums of the partials in the order cav3d_integrals adds them on the
// host (workgroup order per quantity), so both entry points agree bit for bit
__global__ __launch_bounds__(256) void k_post_final(const double *part, long nwc, long nwf, double *sum8)
{
    __shared__ double buf[256 * 6];
    double ac[2] = {0.0, 0.0}, af[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    seq_rows<2>(part, nwc, ac, buf);
    seq_rows<6>(part + 2 * nwc, nwf, af, buf);
    if (threadIdx.x == 0) {
        sum8[0] = ac[0];
        sum8[1] = ac[1];
        for (int a = 0; a < 6; a++) sum8[2 + a] = af[a];
    }
}

// ---- the config-4 objective J = w_vap (vapour volume) + w_fx (x-force on the
// surface ranges, N) and its gradient over pv.  Each interior cell's vapour
// volume rho y1 / rho_v vol and each surface face's x-force (the integrand of
// k_post_faces) are differentiated through the EOS by one 4-lane dual pass
// over (p, T, y1, y2); the velocity enters the wall shear linearly.
__global__ __launch_bounds__(POST_WG) void k_obj_cells(DS S, const double *vol, double w_vap, double *pb, double *val,
                                                       double *dpref)
{
    long t = (long)blockIdx.x * POST_WG + threadIdx.x;
    if (t >= S.NC) return;
    const BlkD &B = S.b[find_blk(S, t, 0)];
    int i, j, k;
    cell_of(B, t - B.c0, i, j, k);
    const long g = gidx(B, i, j, k);
    using T = Dn<4>;
    T d[ND], full[18];
    mixture_eos<T>(T::seed(S.pv[g] + S.P.pref, 0), T::seed(S.pv[4 * S.NG + g], 1), T::seed(S.pv[5 * S.NG + g], 2),
                   T::seed(S.pv[6 * S.NG + g], 3), d, false, full);
    const double w = vol[widx(B, i, j, k)];
    const T a = full[0] * T::seed(S.pv[5 * S.NG + g], 2) / full[3] * w;
    val[t] = w_vap * a.v;
    dpref[t] = w_vap * a.d[0];
    for (int n = 0; n < S.npv; n++) pb[(long)n * S.NG + g] = 0.0;
    pb[g] = w_vap * a.d[0];
    pb[4 * S.NG + g] = w_vap * a.d[1];
    pb[5 * S.NG + g] = w_vap * a.d[2];
    pb[6 * S.NG + g] = w_vap * a.d[3];
}

// the x-force of the faces of one surface range; adds its gradient to the
// adjacent interior cell (a range holds each cell once, so no two threads of
// a launch touch the same cell)
__global__ __launch_bounds__(POST_WG) void k_obj_faces(DS S, FSeg s, const double *xvert, long nv, long v0b,
                                                       const double *xc, double w_fx, double *pb, double *val,
                                                       double *dpref)
{
    long r = (long)blockIdx.x * POST_WG + threadIdx.x;
    if (r >= s.n) return;
    const BlkD &B = S.b[s.b];
    const int n0 = s.hi[0] - s.lo[0] + 1, n1 = s.hi[1] - s.lo[1] + 1;
    const int p[3] = {s.lo[0] + (int)(r % n0), s.lo[1] + (int)((r / n0) % n1), s.lo[2] + (int)(r / ((long)n0 * n1))};
    const int dim[3] = {B.I, B.J, B.K};
    const int ax = s.ax;
    int fi[3] = {p[0], p[1], p[2]}, ci[3] = {p[0], p[1], p[2]};
    fi[ax] = s.side ? dim[ax] : 1;
    ci[ax] = s.side ? dim[ax] : 2;
    const int vplane = s.side ? dim[ax] + 1 : 2;
    double cxv[3];
    face_vec(S, B, ax, fi[0], fi[1], fi[2], cxv);
    if (s.side) for (int c = 0; c < 3; c++) cxv[c] = -cxv[c];
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    const XB X{xvert + v0b, nv, B.I, B.J, B.K};
    double x[3];
    for (int c = 0; c < 3; c++) {
        double acc = 0.0;
        for (int m = 0; m < 4; m++) {
            int u[3];
            u[ax] = vplane;
            u[t1] = p[t1] + (m & 1);
            u[t2] = p[t2] + (m >> 1);
            const double a = xv(X, c, u[0], u[1], u[2]);
            acc = m == 0 ? a : acc + a;
        }
        x[c] = 0.25 * acc;
    }
    const long g = gidx(B, ci[0], ci[1], ci[2]);
    const long w = widx(B, ci[0], ci[1], ci[2]);
    const double pg = S.pv[g], u = S.pv[S.NG + g], vv = S.pv[2 * S.NG + g], ww = S.pv[3 * S.NG + g];
    const double dl = cxv[0] * cxv[0] + cxv[1] * cxv[1] + cxv[2] * cxv[2];
    const double vel = cxv[0] * u + cxv[1] * vv + cxv[2] * ww;
    const double gx = xc[w], gy = xc[S.NW + w], gz = xc[2 * S.NW + w];
    const double dist = sqrt((gx - x[0]) * (gx - x[0]) + (gy - x[1]) * (gy - x[1]) + (gz - x[2]) * (gz - x[2]));
    const double sdl = sqrt(dl);
    const double ut = u - cxv[0] * vel / dl;
    using T = Dn<4>;
    T vis(0.0);
    if (S.P.iturb >= -2) {
        T d[ND];
        mixture_eos<T>(T::seed(pg + S.P.pref, 0), T::seed(S.pv[4 * S.NG + g], 1), T::seed(S.pv[5 * S.NG + g], 2),
                       T::seed(S.pv[6 * S.NG + g], 3), d, true);
        vis = d[D_MU];
        if (S.P.iturb == 0) vis = vis + S.mut[g];
    }
    const T fx = vis * ut / dist * sdl;
    val[r] = w_fx * (-pg * cxv[0] + fx.v);
    dpref[r] = w_fx * fx.d[0];
    const double c = vis.v / dist * sdl;       // d(shear)/d(u, v, w) = c d(ut)/d(u, v, w)
    pb[g] += w_fx * (-cxv[0] + fx.d[0]);
    pb[S.NG + g] += w_fx * c * (1.0 - cxv[0] * cxv[0] / dl);
    pb[2 * S.NG + g] += w_fx * c * (-cxv[0] * cxv[1] / dl);
    pb[3 * S.NG + g] += w_fx * c * (-cxv[0] * cxv[2] / dl);
    pb[4 * S.NG + g] += w_fx * fx.d[1];
    pb[5 * S.NG + g] += w_fx * fx.d[2];
    pb[6 * S.NG + g] += w_fx * fx.d[3];
}

// ---- restart file records of one block (src/variable_module.f90:126-178)
struct Rec {
    long pv, emut, qq, rcav;
};
Rec records(const cav3d_ctx *c, const BlkD &B)
{
    Rec r;
    r.pv = (long)c->npv * B.ng;
    r.emut = c->prm.iturb >= -1 ? B.ng : 0;
    r.qq = (c->prm.nsteady >= 1 ? 2L : 0L) * c->npv * B.nc;
    r.rcav = c->prm.ncav ? 2 * B.nc : 0;
    return r;
}

}  // namespace

extern "C" {

int cav3d_integrals(cav3d_ctx *c, double rhoref, double uref, double area, double aoa, double *out)
{
    if (!c || !out) return fail("bad argument to cav3d_integrals", -1);
    DS &S = c->S;
    long nface;
    bool ok;
    std::vector<FSeg> segs = surface_segments(c, nface, ok);
    if (!ok) return fail("cav3d_integrals: a surface range is not a block face inside the cells", -2);
    if (nface > 0 && !c->xv) return fail("cav3d_integrals: surface forces need the vertex coordinates (cav3d_set_grid)", -1);
    const long nwc = (long)nblk(S.NC, POST_WG), nwf = nface > 0 ? (long)nblk(nface, POST_WG) : 0;
    double *dpart = nullptr, *dv0 = nullptr;
    FSeg *dseg = nullptr;
    std::vector<double> hp(2 * nwc + 6 * nwf);
    std::vector<long> v0(c->v0.begin(), c->v0.end());
    hipError_t e = hipMalloc((void **)&dpart, hp.size() * sizeof(double));
    if (e == hipSuccess && nface > 0) e = hipMalloc((void **)&dseg, segs.size() * sizeof(FSeg));
    if (e == hipSuccess && nface > 0) e = hipMalloc((void **)&dv0, v0.size() * sizeof(long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_post_cells, dim3(nwc), dim3(POST_WG), 0, c->st, S, (const double *)S.vol, dpart);
        if (nface > 0) {
            e = hipMemcpyAsync(dseg, segs.data(), segs.size() * sizeof(FSeg), hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess) e = hipMemcpyAsync(dv0, v0.data(), v0.size() * sizeof(long), hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess)
                hipLaunchKernelGGL(k_post_faces, dim3(nwf), dim3(POST_WG), 0, c->st, S, (const FSeg *)dseg, (int)segs.size(),
                                   nface, (const double *)c->xv, c->NV, (const long *)dv0, (const double *)c->xc,
                                   dpart + 2 * nwc);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), dpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    if (dpart) hipFree(dpart);
    if (dseg) hipFree(dseg);
    if (dv0) hipFree((void *)dv0);
    if (e != hipSuccess) return fail(std::string("cav3d_integrals: ") + hipGetErrorString(e), -3);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long q = 0; q < nwc; q++) for (int a = 0; a < 2; a++) acc[a] += hp[2 * q + a];
    for (long q = 0; q < nwf; q++) for (int a = 0; a < 6; a++) acc[2 + a] += hp[2 * nwc + 6 * q + a];
    integrals_out(acc, rhoref, uref, area, aoa, out);
    return check_launch("cav3d_integrals");
}

int cav3d_integrals_noalloc(cav3d_ctx *c, double rhoref, double uref, double area, double aoa, double *out)
{
    if (!c || !out) return fail("bad argument to cav3d_integrals_noalloc", -1);
    DS &S = c->S;
    FSeg segs[NOALLOC_MAXSEG];
    int nseg = 0;
    long nface;
    bool ok;
    scan_surface(c, nface, ok, [&](const FSeg &s) {
        if (nseg < NOALLOC_MAXSEG) segs[nseg] = s;
        nseg++;
    });
    if (!ok) return fail("cav3d_integrals_noalloc: a surface range is not a block face inside the cells", -2);
    if (nseg > NOALLOC_MAXSEG)
        return fail("cav3d_integrals_noalloc: " + std::to_string(nseg) + " surface ranges, at most " +
                    std::to_string(NOALLOC_MAXSEG), -2);
    if (nface > 0 && !c->xv) return fail("cav3d_integrals_noalloc: surface forces need the vertex coordinates (cav3d_set_grid)", -1);
    const long nwc = (long)nblk(S.NC, POST_WG), nwf = nface > 0 ? (long)nblk(nface, POST_WG) : 0;
    if (2 * nwc + 6 * nwf > NOALLOC_PART_CAP)
        return fail("cav3d_integrals_noalloc: " + std::to_string(2 * nwc + 6 * nwf) + " partial sums exceed the " +
                    std::to_string(NOALLOC_PART_CAP) + " of its fixed device array (" +
                    std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes); use cav3d_integrals", -2);
    unsigned char *ptr = nullptr;
    int rc = scratch_of(c->dev, &ptr);
    if (rc) return rc;
    FSeg *dseg = (FSeg *)(ptr + NOALLOC_SEG_OFF);
    long *dv0 = (long *)(ptr + NOALLOC_V0_OFF);
    double *sum8 = (double *)(ptr + NOALLOC_SUM_OFF);
    double *dpart = (double *)(ptr + NOALLOC_PART_OFF);
    long v0[MAXB] = {};
    for (int b = 0; b < c->nb && b < (int)c->v0.size(); b++) v0[b] = c->v0[b];
    hipLaunchKernelGGL(k_post_cells, dim3(nwc), dim3(POST_WG), 0, c->st, S, (const double *)S.vol, dpart);
    if (nface > 0) {
        HIPCHK(hipMemcpyAsync(dseg, segs, nseg * sizeof(FSeg), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(dv0, v0, sizeof(v0), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_post_faces, dim3(nwf), dim3(POST_WG), 0, c->st, S, (const FSeg *)dseg, nseg, nface,
                           (const double *)c->xv, c->NV, (const long *)dv0, (const double *)c->xc, dpart + 2 * nwc);
    }
    hipLaunchKernelGGL(k_post_final, dim3(1), dim3(256), 0, c->st, (const double *)dpart, nwc, nwf, sum8);
    double acc[8];
    HIPCHK(hipMemcpyAsync(acc, sum8, sizeof(acc), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if ((rc = check_launch("cav3d_integrals_noalloc"))) return rc;
    integrals_out(acc, rhoref, uref, area, aoa, out);
    return 0;
}

int cav3d_export(cav3d_ctx *c, const char *path, int nt_phy, int nt, double time)
{
    if (!c || !path) return fail("bad argument to cav3d_export", -1);
    DS &S = c->S;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(std::string("cav3d_export: cannot open ") + path, -4);
    int rc = 0;
    std::vector<double> h;
    for (int b = 0; b < c->nb && rc == 0; b++) {
        const BlkD &B = c->blk[b];
        const Rec r = records(c, B);
        if ((rc = ensure_stage(c, std::max(r.pv, std::max(r.qq, r.rcav))))) break;
        const int hdr[2] = {nt_phy, nt};
        bool w = fwrite(hdr, sizeof(int), 2, f) == 2 && fwrite(&time, sizeof(double), 1, f) == 1;
        struct { long n; int m; int kind; const double *src; long stride, off; } parts[4] = {
            {r.pv, c->npv, 0, S.pv, S.NG, B.g0}, {r.emut, 1, 0, S.mut, S.NG, B.g0},
            {r.qq, 2 * c->npv, 1, S.qq, S.NC, B.c0}, {r.rcav, 2, 1, S.rcav, S.NC, B.c0}};
        for (auto &p : parts) {
            if (!w || p.n == 0) continue;
            if (p.kind == 0)
                hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(B.ng, 256)), dim3(256), 0, c->st, p.src, p.m, p.stride, p.off, c->stage, B.ng);
            else
                hipLaunchKernelGGL(k_cell_to_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, p.src, p.m, p.stride, p.off, c->stage, B.nc);
            h.resize(p.n);
            if (hipMemcpyAsync(h.data(), c->stage, p.n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipStreamSynchronize(c->st) != hipSuccess) { rc = fail("cav3d_export: device copy failed", -3); break; }
            w = fwrite(h.data(), sizeof(double), p.n, f) == (size_t)p.n;
        }
        if (rc == 0 && !w) rc = fail(std::string("cav3d_export: write failed on ") + path, -4);
    }
    if (fclose(f) != 0 && rc == 0) rc = fail(std::string("cav3d_export: close failed on ") + path, -4);
    return rc ? rc : check_launch("cav3d_export");
}

int cav3d_import(cav3d_ctx *c, const char *path, int *nps, int *nts, double *timeprev)
{
    if (!c || !path) return fail("bad argument to cav3d_import", -1);
    DS &S = c->S;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(std::string("cav3d_import: cannot open ") + path, -4);
    int rc = 0, hdr[2] = {0, 0};
    double tp = 0.0;
    std::vector<double> h;
    for (int b = 0; b < c->nb && rc == 0; b++) {
        const BlkD &B = c->blk[b];
        const Rec r = records(c, B);
        if ((rc = ensure_stage(c, std::max(r.pv, std::max(r.qq, r.rcav))))) break;
        int hb[2];
        double tb;
        if (fread(hb, sizeof(int), 2, f) != 2 || fread(&tb, sizeof(double), 1, f) != 1) {
            rc = fail(std::string("cav3d_import: short file ") + path, -4);
            break;
        }
        if (b == 0) { hdr[0] = hb[0]; hdr[1] = hb[1]; tp = tb; }
        long ns[4] = {r.pv, r.emut, r.qq, r.rcav};
        for (int q = 0; q < 4 && rc == 0; q++) {
            if (ns[q] == 0) continue;
            h.resize(ns[q]);
            if (fread(h.data(), sizeof(double), ns[q], f) != (size_t)ns[q]) { rc = fail(std::string("cav3d_import: short file ") + path, -4); break; }
            if (hipMemcpyAsync(c->stage, h.data(), ns[q] * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess) {
                rc = fail("cav3d_import: device copy failed", -3);
                break;
            }
            // the reader keeps the interior cells only (src/initial_module.f90:334-371)
            if (q == 0)
                hipLaunchKernelGGL(k_interior_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage, c->npv, S.pv, S.NG, B);
            else if (q == 1 && S.mut)
                hipLaunchKernelGGL(k_interior_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage, 1, S.mut, S.NG, B);
            else if (q == 2 && S.qq)
                hipLaunchKernelGGL(k_cell_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage, 2 * c->npv, S.NC, B.c0, S.qq, B.nc);
            else if (q == 3 && S.rcav)
                hipLaunchKernelGGL(k_cell_from_aos, dim3(nblk(B.nc, 256)), dim3(256), 0, c->st, (const double *)c->stage, 2, S.NC, B.c0, S.rcav, B.nc);
            if (hipStreamSynchronize(c->st) != hipSuccess) rc = fail("cav3d_import: device copy failed", -3);
        }
    }
    fclose(f);
    if (rc) return rc;
    // dv and tv(1:2) of the interior from the EOS
    hipLaunchKernelGGL(k_eos_cells, dim3(nblk(S.NC, 128)), dim3(128), 0, c->st, S, 1, 0);
    HIPCHK(hipStreamSynchronize(c->st));
    // counters as the reader returns them
    int ps = hdr[0] + 1, ts = hdr[1] + 1;
    if (c->prm.nsteady >= 1) ts = 1; else ps = 1;
    if (nps) *nps = ps;
    if (nts) *nts = ts;
    if (timeprev) *timeprev = tp;
    return check_launch("cav3d_import");
}


}  // extern "C"

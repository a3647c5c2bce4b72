/*
 * ORACLE — test infrastructure only (see orc_eos.c header).
 *
 * Grid metrics of a multi-block grid from its vertex coordinates: the plain-C
 * FP64 restatement of t_grid's construct sequence calnormal -> calvolume ->
 * calydns (src/grid_module.f90:328-330), with every block resident in one
 * process and one block standing for one MPI rank of the reference:
 *   calnormal  src/grid_module.f90:609-940   face vectors cx/ex/tx, degenerate
 *              faces, ghost copies, interface exchange of the tangential face
 *              vectors, closing re-ordering of the connection ranges
 *   calvolume  src/grid_module.f90:942-1119  centroid grd(2:4), six-pyramid
 *              volume grd(1), ghost fill, interface exchange of grd(1)
 *   pyramid    src/grid_module.f90:1121-1145
 *   calydns    src/grid_module.f90:350-607   wall distance grd(5): nearest
 *              BCWall face centroid over all blocks, ghost copies, exchange
 * Interface exchanges keep the reference's MPI semantics: every block packs
 * what it sends from its arrays as they stand before any receive is applied;
 * a connection whose donor is the block itself is applied at once, in
 * connection order (the "no mpi boundary" branches).
 *
 * Connections arrive as the host passes them to the solver: 22 ints after the
 * three-layer extension of src/bc_module.f90:355-488 applied to the
 * re-ordered ranges.  grid_conn() recovers the one-layer ranges calvolume /
 * calydns loop over and the vertex-index ranges calnormal loops over.
 *
 * Reference defect not reproduced: for iturb = -2 the reference allocates
 * grd(4,...) (src/grid_module.f90:319-323) but still calls calydns
 * (:330), which stores grd(5,i,j,k), past the first extent.  With ngrd = 4
 * no wall distance is computed here.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define XV(b, c, i, j, k) ((b)->x[(c) + 3 * ((long)((i) - 2) + (long)(b)->imax * ((long)((j) - 2) + (long)(b)->jmax * ((k) - 2)))])
#define CXI(b, i, j, k) (3 * ((long)((i) - 1) + (long)(b)->imax * ((long)((j) - 1) + (long)((b)->jmax + 1) * ((k) - 1))))
#define EXI(b, i, j, k) (3 * ((long)((i) - 1) + (long)((b)->imax + 1) * ((long)((j) - 1) + (long)(b)->jmax * ((k) - 1))))
#define TXI(b, i, j, k) (3 * ((long)((i) - 1) + (long)((b)->imax + 1) * ((long)((j) - 1) + (long)((b)->jmax + 1) * ((k) - 1))))
#define GI(b, ng, q, i, j, k) ((q) + (long)(ng) * ((long)((i) - 1) + (long)((b)->imax + 1) * ((long)((j) - 1) + (long)((b)->jmax + 1) * ((k) - 1))))

typedef struct {
    int src, dst;
    int m, q;              /* normal axis in the sending block / in the donor */
    int is[3], ie[3];      /* sending block's one-layer range (cells) */
    int isd[3], ied[3];    /* donor range, as the re-ordered connectinfo holds it */
    int nis[3], nisd[3];   /* calnormal's (vertex-index) starts */
    int nied[3];
    int tm[3][3];
} gconn;

/* The one-layer grid_module ranges of a connection from its extended form
 * (src/grid_module.f90:228-300 conversion, :924-938 re-ordering,
 * src/bc_module.f90:355-488 extension undone). */
static int grid_conn(const orc_gblock *blk, int nb, int b, const int *d, gconn *g)
{
    memset(g, 0, sizeof(*g));
    g->src = b;
    g->dst = d[0];
    if (g->dst < 0 || g->dst >= nb) return -1;
    for (int r = 0; r < 3; r++) {
        g->is[r] = d[1 + r]; g->ie[r] = d[4 + r]; g->isd[r] = d[7 + r]; g->ied[r] = d[10 + r];
        for (int l = 0; l < 3; l++) g->tm[r][l] = d[13 + 3 * r + l];
    }
    const orc_gblock *A = &blk[b], *D = &blk[g->dst];
    const int da[3] = {A->imax, A->jmax, A->kmax}, dd[3] = {D->imax, D->jmax, D->kmax};
    g->q = -1;
    for (int r = 0; r < 3 && g->q < 0; r++)
        if (g->isd[r] <= 1 || g->isd[r] >= dd[r] + 1 || g->ied[r] <= 1 || g->ied[r] >= dd[r] + 1) g->q = r;
    if (g->q < 0) return -1;
    g->m = -1;
    for (int l = 0; l < 3; l++) if (g->tm[g->q][l] != 0) g->m = l;
    if (g->m < 0) return -1;
    const int q = g->q, m = g->m;
    /* the donor's own ghost layer (1 or dim+1) is the unextended end */
    int lead = (g->isd[q] == 1 || g->isd[q] == dd[q] + 1);
    int orig = lead ? g->isd[q] : g->ied[q];
    int face = lead ? g->is[m] : g->ie[m];
    if (lead ? (g->is[m] != 2) : (g->ie[m] != da[m])) return -1;
    g->isd[q] = g->ied[q] = orig;
    g->is[m] = g->ie[m] = face;
    for (int r = 0; r < 3; r++) {
        g->nis[r] = r == m ? g->is[r] : g->is[r] - 1;
        g->nisd[r] = g->isd[r];
        g->nied[r] = g->ied[r];
        if (r != q) {
            if (g->isd[r] < g->ied[r]) g->nisd[r] = g->isd[r] - 1;
            else if (g->isd[r] > g->ied[r]) g->nied[r] = g->ied[r] - 1;
        }
    }
    return 0;
}

static void xmap(const gconn *g, const int *is, const int *isd, int i, int j, int k, int *o)
{
    const int di[3] = {i - is[0], j - is[1], k - is[2]};
    for (int r = 0; r < 3; r++) o[r] = g->tm[r][0] * di[0] + g->tm[r][1] * di[1] + g->tm[r][2] * di[2] + isd[r];
}

/* ---------------------------------------------------------------- calnormal */
static void face_vectors(orc_gblock *b)
{
    const int I = b->imax, J = b->jmax, K = b->kmax;
    for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
            for (int i = 1; i <= I; i++) {
                double xe = XV(b, 0, i + 1, j + 1, k) - XV(b, 0, i + 1, j, k + 1);
                double ye = XV(b, 1, i + 1, j + 1, k) - XV(b, 1, i + 1, j, k + 1);
                double ze = XV(b, 2, i + 1, j + 1, k) - XV(b, 2, i + 1, j, k + 1);
                double xs = XV(b, 0, i + 1, j + 1, k + 1) - XV(b, 0, i + 1, j, k);
                double ys = XV(b, 1, i + 1, j + 1, k + 1) - XV(b, 1, i + 1, j, k);
                double zs = XV(b, 2, i + 1, j + 1, k + 1) - XV(b, 2, i + 1, j, k);
                double *v = b->cx + CXI(b, i, j, k);
                v[0] = 0.5 * (ye * zs - ys * ze);
                v[1] = 0.5 * (xs * ze - xe * zs);
                v[2] = 0.5 * (xe * ys - xs * ye);
            }
    for (int i = 2; i <= I; i++)
        for (int k = 2; k <= K; k++)
            for (int j = 1; j <= J; j++) {
                double xc = XV(b, 0, i, j + 1, k) - XV(b, 0, i + 1, j + 1, k + 1);
                double yc = XV(b, 1, i, j + 1, k) - XV(b, 1, i + 1, j + 1, k + 1);
                double zc = XV(b, 2, i, j + 1, k) - XV(b, 2, i + 1, j + 1, k + 1);
                double xs = XV(b, 0, i, j + 1, k + 1) - XV(b, 0, i + 1, j + 1, k);
                double ys = XV(b, 1, i, j + 1, k + 1) - XV(b, 1, i + 1, j + 1, k);
                double zs = XV(b, 2, i, j + 1, k + 1) - XV(b, 2, i + 1, j + 1, k);
                double *v = b->ex + EXI(b, i, j, k);
                v[0] = 0.5 * (yc * zs - ys * zc);
                v[1] = 0.5 * (xs * zc - xc * zs);
                v[2] = 0.5 * (xc * ys - xs * yc);
            }
    for (int j = 2; j <= J; j++)
        for (int i = 2; i <= I; i++)
            for (int k = 1; k <= K; k++) {
                double xc = XV(b, 0, i + 1, j, k + 1) - XV(b, 0, i, j + 1, k + 1);
                double yc = XV(b, 1, i + 1, j, k + 1) - XV(b, 1, i, j + 1, k + 1);
                double zc = XV(b, 2, i + 1, j, k + 1) - XV(b, 2, i, j + 1, k + 1);
                double xe = XV(b, 0, i + 1, j + 1, k + 1) - XV(b, 0, i, j, k + 1);
                double ye = XV(b, 1, i + 1, j + 1, k + 1) - XV(b, 1, i, j, k + 1);
                double ze = XV(b, 2, i + 1, j + 1, k + 1) - XV(b, 2, i, j, k + 1);
                double *v = b->tx + TXI(b, i, j, k);
                v[0] = 0.5 * (yc * ze - ye * zc);
                v[1] = 0.5 * (xe * zc - xc * ze);
                v[2] = 0.5 * (xc * ye - xe * yc);
            }
    /* degenerate faces carry no area (src/grid_module.f90:674-708) */
    for (int n = 0; n < b->nbc; n++) {
        const int *d = b->bc + 7 * n;
        if ((d[0] & 0xff) != ORC_BC_DEGENERATE) continue;
        if (d[1] == d[4]) {
            int i0 = d[1] == 1 ? 1 : I;
            for (int k = 1; k <= K + 1; k++) for (int j = 1; j <= J + 1; j++)
                memset(b->cx + CXI(b, i0, j, k), 0, 3 * sizeof(double));
        }
        if (d[2] == d[5]) {
            int j0 = d[2] == 1 ? 1 : J;
            for (int k = 1; k <= K + 1; k++) for (int i = 1; i <= I + 1; i++)
                memset(b->ex + EXI(b, i, j0, k), 0, 3 * sizeof(double));
        }
        if (d[3] == d[6]) {
            int k0 = d[3] == 1 ? 1 : K;
            for (int j = 1; j <= J + 1; j++) for (int i = 1; i <= I + 1; i++)
                memset(b->tx + TXI(b, i, j, k0), 0, 3 * sizeof(double));
        }
    }
    /* ghost copies, statement by statement (src/grid_module.f90:710-723) */
#define CP3(dst, src) memcpy((dst), (src), 3 * sizeof(double))
    for (int k = 1; k <= K + 1; k++) for (int i = 1; i <= I; i++) CP3(b->cx + CXI(b, i, 1, k), b->cx + CXI(b, i, 2, k));
    for (int k = 1; k <= K + 1; k++) for (int i = 1; i <= I; i++) CP3(b->cx + CXI(b, i, J + 1, k), b->cx + CXI(b, i, J, k));
    for (int j = 1; j <= J + 1; j++) for (int i = 1; i <= I; i++) CP3(b->cx + CXI(b, i, j, 1), b->cx + CXI(b, i, j, 2));
    for (int j = 1; j <= J + 1; j++) for (int i = 1; i <= I; i++) CP3(b->cx + CXI(b, i, j, K + 1), b->cx + CXI(b, i, j, K));
    for (int k = 1; k <= K + 1; k++) for (int j = 1; j <= J; j++) CP3(b->ex + EXI(b, 1, j, k), b->ex + EXI(b, 2, j, k));
    for (int k = 1; k <= K + 1; k++) for (int j = 1; j <= J; j++) CP3(b->ex + EXI(b, I + 1, j, k), b->ex + EXI(b, I, j, k));
    for (int j = 1; j <= J; j++) for (int i = 1; i <= I + 1; i++) CP3(b->ex + EXI(b, i, j, 1), b->ex + EXI(b, i, j, 2));
    for (int j = 1; j <= J; j++) for (int i = 1; i <= I + 1; i++) CP3(b->ex + EXI(b, i, j, K + 1), b->ex + EXI(b, i, j, K));
    for (int k = 1; k <= K; k++) for (int j = 1; j <= J + 1; j++) CP3(b->tx + TXI(b, 1, j, k), b->tx + TXI(b, 2, j, k));
    for (int k = 1; k <= K; k++) for (int j = 1; j <= J + 1; j++) CP3(b->tx + TXI(b, I + 1, j, k), b->tx + TXI(b, I, j, k));
    for (int k = 1; k <= K; k++) for (int i = 1; i <= I + 1; i++) CP3(b->tx + TXI(b, i, 1, k), b->tx + TXI(b, i, 2, k));
    for (int k = 1; k <= K; k++) for (int i = 1; i <= I + 1; i++) CP3(b->tx + TXI(b, i, J + 1, k), b->tx + TXI(b, i, J, k));
#undef CP3
}

/* face-vector array and offset by axis: 0 cx, 1 ex, 2 tx */
static double *fv(orc_gblock *b, int ax, int i, int j, int k)
{
    if (ax == 0) return b->cx + CXI(b, i, j, k);
    if (ax == 1) return b->ex + EXI(b, i, j, k);
    return b->tx + TXI(b, i, j, k);
}

static long conn_cells(const int *is, const int *ie)
{
    long n = 1;
    for (int r = 0; r < 3; r++) n *= ie[r] >= is[r] ? ie[r] - is[r] + 1 : 0;
    return n;
}

/* The exchange steps.  what: 0 tangential face vectors (calnormal ranges),
 * 1 grd(1) (calvolume), 2 grd(5) (calydns). */
static void pack(orc_gblock *blk, int ngrd, const gconn *g, int what, double *buf)
{
    orc_gblock *A = &blk[g->src];
    const int *is = what == 0 ? g->nis : g->is;
    int ie[3];
    for (int r = 0; r < 3; r++) ie[r] = g->ie[r];
    long l = 0;
    /* calnormal sends (ex,tx), (tx,cx) or (cx,ex) by the sender's normal */
    const int f1 = (g->m + 1) % 3, f2 = (g->m + 2) % 3;
    for (int k = is[2]; k <= ie[2]; k++)
        for (int j = is[1]; j <= ie[1]; j++)
            for (int i = is[0]; i <= ie[0]; i++) {
                if (what == 0) {
                    const double *a = fv(A, f1, i, j, k), *c = fv(A, f2, i, j, k);
                    for (int m = 0; m < 3; m++) { buf[l++] = a[m]; buf[l++] = c[m]; }
                } else {
                    buf[l++] = A->grd[GI(A, ngrd, what == 1 ? 0 : 4, i, j, k)];
                }
            }
}

static void unpack(orc_gblock *blk, int ngrd, const gconn *g, int what, const double *buf)
{
    orc_gblock *B = &blk[g->dst];
    const int *is = what == 0 ? g->nis : g->is;
    const int *isd = what == 0 ? g->nisd : g->isd;
    const int q = g->q;
    int sign1 = 1, sign2 = 1;
    if (what == 0) {
        /* the donor's tangential ranges set the signs (src/grid_module.f90:782-807) */
        const int t1 = (q + 1) % 3, t2 = (q + 2) % 3;
        if (g->nisd[t1] > g->nied[t1]) sign1 = -sign1;
        if (g->nisd[t2] > g->nied[t2]) sign2 = -sign2;
    }
    const int w1 = (q + 1) % 3, w2 = (q + 2) % 3;
    long l = 0;
    for (int k = is[2]; k <= g->ie[2]; k++)
        for (int j = is[1]; j <= g->ie[1]; j++)
            for (int i = is[0]; i <= g->ie[0]; i++) {
                int o[3];
                xmap(g, is, isd, i, j, k, o);
                if (what == 0) {
                    double *a = fv(B, w1, o[0], o[1], o[2]), *c = fv(B, w2, o[0], o[1], o[2]);
                    for (int m = 0; m < 3; m++) {
                        a[m] = sign1 * buf[l++];
                        c[m] = sign2 * buf[l++];
                    }
                } else {
                    B->grd[GI(B, ngrd, what == 1 ? 0 : 4, o[0], o[1], o[2])] = buf[l++];
                }
            }
}

static int exchange(orc_gblock *blk, int nb, int ngrd, const gconn *gc, int ng, int what)
{
    double **bufs = calloc(ng > 0 ? ng : 1, sizeof(double *));
    if (!bufs) return -1;
    for (int n = 0; n < ng; n++) {
        const gconn *g = &gc[n];
        const int *is = what == 0 ? g->nis : g->is;
        long cnt = conn_cells(is, g->ie) * (what == 0 ? 6 : 1);
        bufs[n] = malloc((cnt > 0 ? cnt : 1) * sizeof(double));
        pack(blk, ngrd, g, what, bufs[n]);
        if (g->dst == g->src) unpack(blk, ngrd, g, what, bufs[n]);
    }
    for (int n = 0; n < ng; n++)
        if (gc[n].dst != gc[n].src) unpack(blk, ngrd, &gc[n], what, bufs[n]);
    for (int n = 0; n < ng; n++) free(bufs[n]);
    free(bufs);
    (void)nb;
    return 0;
}

/* ---------------------------------------------------------------- calvolume */
static double pyramid(const double *pa, const double *pb, const double *pc, const double *pd,
                      double xc, double yc, double zc)
{
    double csx = 0.25 * (pa[0] + pb[0] + pc[0] + pd[0]);
    double csy = 0.25 * (pa[1] + pb[1] + pc[1] + pd[1]);
    double csz = 0.25 * (pa[2] + pb[2] + pc[2] + pd[2]);
    double vax = pc[0] - pa[0], vay = pc[1] - pa[1], vaz = pc[2] - pa[2];
    double vbx = pd[0] - pb[0], vby = pd[1] - pb[1], vbz = pd[2] - pb[2];
    double volpx = (xc - csx) * (vay * vbz - vaz * vby);
    double volpy = (yc - csy) * (vaz * vbx - vax * vbz);
    double volpz = (zc - csz) * (vax * vby - vay * vbx);
    return fabs(volpx + volpy + volpz);
}

static void volumes(orc_gblock *b, int ngrd)
{
    const int I = b->imax, J = b->jmax, K = b->kmax;
    for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
            for (int i = 2; i <= I; i++) {
                double *g = b->grd + GI(b, ngrd, 0, i, j, k);
                for (int c = 0; c < 3; c++)
                    g[1 + c] = 0.125 * (XV(b, c, i, j, k) + XV(b, c, i + 1, j, k) + XV(b, c, i, j + 1, k) +
                                        XV(b, c, i, j, k + 1) + XV(b, c, i + 1, j + 1, k) + XV(b, c, i, j + 1, k + 1) +
                                        XV(b, c, i + 1, j, k + 1) + XV(b, c, i + 1, j + 1, k + 1));
                /* the six faces in the reference's order and vertex order */
                static const int fc[6][4][3] = {
                    {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}, {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},
                    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}}, {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},
                    {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}}, {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}}};
                double vp[6];
                for (int f = 0; f < 6; f++) {
                    double p[4][3];
                    for (int v = 0; v < 4; v++)
                        for (int c = 0; c < 3; c++)
                            p[v][c] = XV(b, c, i + fc[f][v][0], j + fc[f][v][1], k + fc[f][v][2]);
                    vp[f] = pyramid(p[0], p[1], p[2], p[3], g[1], g[2], g[3]);
                }
                g[0] = (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5]) / 6.0;
            }
    /* ghost fill (src/grid_module.f90:1073-1113): volumes copied, centroids
       extrapolated; kmin reads grd(:,3,j,3) as written */
#define G(q, i, j, k) b->grd[GI(b, ngrd, q, i, j, k)]
    for (int k = 2; k <= K; k++)
        for (int i = 2; i <= I; i++) {
            G(0, i, 1, k) = G(0, i, 2, k);
            for (int q = 1; q <= 3; q++) G(q, i, 1, k) = 2.0 * G(q, i, 2, k) - G(q, i, 3, k);
            G(0, i, J + 1, k) = G(0, i, J, k);
            for (int q = 1; q <= 3; q++) G(q, i, J + 1, k) = 2.0 * G(q, i, J, k) - G(q, i, J - 1, k);
        }
    for (int k = 2; k <= K; k++)
        for (int j = 1; j <= J + 1; j++) {
            G(0, 1, j, k) = G(0, 2, j, k);
            for (int q = 1; q <= 3; q++) G(q, 1, j, k) = 2.0 * G(q, 2, j, k) - G(q, 3, j, k);
            G(0, I + 1, j, k) = G(0, I, j, k);
            for (int q = 1; q <= 3; q++) G(q, I + 1, j, k) = 2.0 * G(q, I, j, k) - G(q, I - 1, j, k);
        }
    for (int j = 1; j <= J + 1; j++)
        for (int i = 1; i <= I + 1; i++) {
            G(0, i, j, 1) = G(0, i, j, 2);
            for (int q = 1; q <= 3; q++) G(q, i, j, 1) = 2.0 * G(q, i, j, 2) - G(q, 3, j, 3);
            G(0, i, j, K + 1) = G(0, i, j, K);
            for (int q = 1; q <= 3; q++) G(q, i, j, K + 1) = 2.0 * G(q, i, j, K) - G(q, i, j, K - 1);
        }
}

/* ---------------------------------------------------------------- calydns */
static long wall_points(const orc_gblock *blk, int nb, double *out)
{
    long m = 0;
    for (int bb = 0; bb < nb; bb++) {
        const orc_gblock *b = &blk[bb];
        for (int n = 0; n < b->nbc; n++) {
            const int *d = b->bc + 7 * n;
            if ((d[0] & 0xff) != ORC_BC_WALL) continue;
            const int *s = d + 1, *e = d + 4;
            for (int ax = 0; ax < 3; ax++) {
                if (s[ax] != e[ax]) continue;
                int fv0 = s[ax] == 1 ? 2 : (ax == 0 ? b->imax : ax == 1 ? b->jmax : b->kmax) + 1;
                int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                for (int u2 = s[t2]; u2 <= e[t2]; u2++)
                    for (int u1 = s[t1]; u1 <= e[t1]; u1++) {
                        if (out) {
                            for (int c = 0; c < 3; c++) {
                                int p[4][3];
                                for (int v = 0; v < 4; v++) {
                                    p[v][ax] = fv0;
                                    p[v][t1] = u1 + (v & 1);
                                    p[v][t2] = u2 + (v >> 1);
                                }
                                out[3 * m + c] = 0.25 * (XV(b, c, p[0][0], p[0][1], p[0][2]) + XV(b, c, p[1][0], p[1][1], p[1][2]) +
                                                         XV(b, c, p[2][0], p[2][1], p[2][2]) + XV(b, c, p[3][0], p[3][1], p[3][2]));
                            }
                        }
                        m++;
                    }
            }
        }
    }
    return m;
}

static void wall_distance(orc_gblock *b, int ngrd, const double *w, long nw)
{
    const int I = b->imax, J = b->jmax, K = b->kmax;
#define G5(i, j, k) b->grd[GI(b, ngrd, 4, i, j, k)]
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
            for (int i = 2; i <= I; i++) {
                const double *g = b->grd + GI(b, ngrd, 0, i, j, k);
                double xc = g[1], yc = g[2], zc = g[3];
                double ymin = 1.0e6;
                for (long m = 0; m < nw; m++) {
                    double dx = xc - w[3 * m], dy = yc - w[3 * m + 1], dz = zc - w[3 * m + 2];
                    double yt = sqrt(dx * dx + dy * dy + dz * dz);
                    if (yt < ymin) ymin = yt;
                }
                G5(i, j, k) = ymin;
            }
    for (int k = 2; k <= K; k++)
        for (int i = 2; i <= I; i++) { G5(i, 1, k) = G5(i, 2, k); G5(i, J + 1, k) = G5(i, J, k); }
    for (int k = 2; k <= K; k++)
        for (int j = 1; j <= J + 1; j++) { G5(1, j, k) = G5(2, j, k); G5(I + 1, j, k) = G5(I, j, k); }
    for (int j = 1; j <= J + 1; j++)
        for (int i = 1; i <= I + 1; i++) { G5(i, j, 1) = G5(i, j, 2); G5(i, j, K + 1) = G5(i, j, K); }
#undef G5
#undef G
}

int orc_grid_metrics(int nb, orc_gblock *blk, int ngrd)
{
    int ng = 0;
    for (int b = 0; b < nb; b++) ng += blk[b].ncon;
    gconn *gc = calloc(ng > 0 ? ng : 1, sizeof(gconn));
    if (!gc) return -1;
    int n = 0;
    for (int b = 0; b < nb; b++)
        for (int c = 0; c < blk[b].ncon; c++, n++)
            if (grid_conn(blk, nb, b, blk[b].con + 22 * c, &gc[n])) { free(gc); return -2; }
    for (int b = 0; b < nb; b++) {
        orc_gblock *B = &blk[b];
        memset(B->cx, 0, sizeof(double) * 3 * (long)B->imax * (B->jmax + 1) * (B->kmax + 1));
        memset(B->ex, 0, sizeof(double) * 3 * (long)(B->imax + 1) * B->jmax * (B->kmax + 1));
        memset(B->tx, 0, sizeof(double) * 3 * (long)(B->imax + 1) * (B->jmax + 1) * B->kmax);
        memset(B->grd, 0, sizeof(double) * ngrd * (long)(B->imax + 1) * (B->jmax + 1) * (B->kmax + 1));
        face_vectors(B);
    }
    exchange(blk, nb, ngrd, gc, ng, 0);
    for (int b = 0; b < nb; b++) volumes(&blk[b], ngrd);
    exchange(blk, nb, ngrd, gc, ng, 1);
    if (ngrd >= 5) {
        long nw = wall_points(blk, nb, NULL);
        double *w = malloc(3 * (nw > 0 ? nw : 1) * sizeof(double));
        wall_points(blk, nb, w);
        for (int b = 0; b < nb; b++) wall_distance(&blk[b], ngrd, w, nw);
        free(w);
        exchange(blk, nb, ngrd, gc, ng, 2);
    }
    free(gc);
    return 0;
}

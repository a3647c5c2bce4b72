/*
 * ORACLE — test infrastructure only (see orc_eos.c header).
 *
 * One pseudo-time step of the reference, restated in C:
 *   t_lusgs%timeinteg   src/update_module.f90:455-733
 *   t_rhs%calrhs        src/rhs_module.f90:260-643
 *   localtime/mintime   src/timestep_module.f90:127-316
 *   t_bc%construct      src/bc_module.f90:71-1188 (faces, 12 edges, 8 corners)
 *   t_bc%setbc          src/bc_module.f90:1235-1396 (same-process halo copy)
 *   BC routines         src/bc_module.f90:1806-3132
 *   unsteadysource      src/unsteady_module.f90:96-114
 *   setqres/residual    src/residual_module.f90:94-186
 *   unsteadyupdate      src/solve_module.f90:155-183
 * All blocks live in one process; the MPI exchange of setbc becomes a copy
 * from the sending block's cells into the donor block's ghost cells.
 * OpenMP is used only where the reference's per-cell results do not depend
 * on the loop order: face loops per line, cell loops, and the LU-SGS sweeps
 * by hyperplanes i+j+k = s (each cell reads only the previous plane).
 */
#include "orc_int.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { FN_NONE = 0, FN_WALLINV, FN_WALLVIS, FN_WALLVISKW, FN_INSUB, FN_OUTSUB, FN_FARFIELD, FN_SYM,
       FN_INSUPER, FN_COPY, FN_EMPTY, FN_EDGEWW, FN_EDGEWWKW, FN_EDGENOWALL, FN_CORNERWW,
       FN_CORNERWWKW, FN_CORNERNOWALL };
enum { F_NONE = -1, F_IMIN = 0, F_IMAX, F_JMIN, F_JMAX, F_KMIN, F_KMAX };

typedef struct {
    int kind, fn, face, surface;
    int istart[3], iend[3], origin[3], dir[3], nb[6][3];
    int ostart[3], oend[3];   /* the range as grid_module gives it */
} bcinfo;

typedef struct {
    int donor;
    int is[3], ie[3], isd[3], ied[3], tm[3][3];
} conn;

typedef struct {
    int imax, jmax, kmax, nx, ny, nz;
    double *cx, *ex, *tx, *grd;
    double *pv, *dv, *tv, *qq;
    double *res, *icav, *itt, *ocut, *dt, *dqs, *dcv, *qres, *rcav;
    int nbc, ncon;
    bcinfo *bc;
    bcinfo edge[12], corner[8];
    conn *con;
} blk;

struct orc_ctx {
    orc_cfg c;
    int nb;
    blk *b;
    double time_acc;
};

/* ---------------- index helpers (Fortran layouts) ---------------- */
static inline long cell(const blk *b, int i, int j, int k)
{
    return (long)(i + 1) + (long)b->nx * ((long)(j + 1) + (long)b->ny * (long)(k + 1));
}
static inline double *PV(const orc_ctx *c, const blk *b, int i, int j, int k) { return b->pv + c->c.npv * cell(b, i, j, k); }
static inline double *DV(const blk *b, int i, int j, int k) { return b->dv + 18 * cell(b, i, j, k); }
static inline double *TV(const orc_ctx *c, const blk *b, int i, int j, int k) { return b->tv + c->c.ntv * cell(b, i, j, k); }
static inline double *CX(const blk *b, int i, int j, int k)
{ return b->cx + 3 * ((long)(i - 1) + (long)b->imax * ((long)(j - 1) + (long)(b->jmax + 1) * (k - 1))); }
static inline double *EX(const blk *b, int i, int j, int k)
{ return b->ex + 3 * ((long)(i - 1) + (long)(b->imax + 1) * ((long)(j - 1) + (long)b->jmax * (k - 1))); }
static inline double *TX(const blk *b, int i, int j, int k)
{ return b->tx + 3 * ((long)(i - 1) + (long)(b->imax + 1) * ((long)(j - 1) + (long)(b->jmax + 1) * (k - 1))); }
static inline double *GRD(const orc_ctx *c, const blk *b, int i, int j, int k)
{ return b->grd + c->c.ngrd * ((long)(i - 1) + (long)(b->imax + 1) * ((long)(j - 1) + (long)(b->jmax + 1) * (k - 1))); }
static inline long ii_(const blk *b, int i, int j, int k)
{ return (long)(i - 2) + (long)(b->imax - 1) * ((long)(j - 2) + (long)(b->jmax - 1) * (k - 2)); }
static inline long iw_(const blk *b, int i, int j, int k)
{ return (long)(i - 1) + (long)(b->imax + 1) * ((long)(j - 1) + (long)(b->jmax + 1) * (k - 1)); }

static int nthreads = 1;
void orc_set_threads(int n) { nthreads = n > 0 ? n : 1; }

/* ---------------- EOS at a cell (deteos of the path's EOS object) ---------------- */
static void cell_eos(const orc_ctx *c, double pabs, double t, double y1, double y2, double *dv, double *tv)
{
    if (c->c.iturb == -3) orc_deteos_simple(pabs, t, y1, y2, dv);
    else orc_deteos_prop(pabs, t, y1, y2, dv, tv);
}

/* ---------------- BC construction, src/bc_module.f90:124-218, 502-1185 ---------------- */
static int wall_fn(int iturb)
{
    switch (iturb) {
    case 0: return FN_WALLVISKW;
    case -2: return FN_WALLVIS;
    default: return FN_WALLINV;
    }
}

static void set3(int *a, int x, int y, int z) { a[0] = x; a[1] = y; a[2] = z; }

static int covers(const blk *b, const int *p, int d0, int d1, int d2)
{
    /* any wall/symmetry face BC whose adjusted range holds p in dims d0,d1(,d2) */
    for (int n = 0; n < b->nbc; n++) {
        const bcinfo *bi = &b->bc[n];
        if (bi->kind != ORC_BC_WALL && bi->kind != ORC_BC_SYMMETRY) continue;
        int ok = (bi->istart[d0] <= p[d0]) && (bi->iend[d0] >= p[d0]) &&
                 (bi->istart[d1] <= p[d1]) && (bi->iend[d1] >= p[d1]);
        if (d2 >= 0) ok = ok && (bi->istart[d2] <= p[d2]) && (bi->iend[d2] >= p[d2]);
        if (ok) return 1;
    }
    return 0;
}

static void build_edges_corners(const orc_ctx *c, blk *b)
{
    int I = b->imax, J = b->jmax, K = b->kmax, iturb = c->c.iturb;
    bcinfo *e = b->edge;
    memset(e, 0, sizeof(b->edge));
    /* edge table, src/bc_module.f90:505-623 */
    set3(e[0].istart, -1, -1, 2);          set3(e[0].iend, 1, 1, K);
    set3(e[0].nb[0], 1, 2, 0);             set3(e[0].nb[1], 2, 1, 0);             set3(e[0].origin, 2, 2, 0);
    set3(e[1].istart, I + 1, -1, 2);       set3(e[1].iend, I + 3, 1, K);
    set3(e[1].nb[0], I + 1, 2, 0);         set3(e[1].nb[1], I, 1, 0);             set3(e[1].origin, I, 2, 0);
    set3(e[2].istart, -1, J + 1, 2);       set3(e[2].iend, 1, J + 3, K);
    set3(e[2].nb[0], 1, J, 0);             set3(e[2].nb[1], 2, J + 1, 0);         set3(e[2].origin, 2, J, 0);
    set3(e[3].istart, I + 1, J + 1, 2);    set3(e[3].iend, I + 3, J + 3, K);
    set3(e[3].nb[0], I + 1, J, 0);         set3(e[3].nb[1], I, J + 1, 0);         set3(e[3].origin, I, J, 0);
    set3(e[4].istart, -1, 2, -1);          set3(e[4].iend, 1, J, 1);
    set3(e[4].nb[0], 1, 0, 2);             set3(e[4].nb[2], 2, 0, 1);             set3(e[4].origin, 2, 0, 2);
    set3(e[5].istart, I + 1, 2, -1);       set3(e[5].iend, I + 3, J, 1);
    set3(e[5].nb[0], I + 1, 0, 2);         set3(e[5].nb[2], I, 0, 1);             set3(e[5].origin, I, 0, 2);
    set3(e[6].istart, -1, 2, K + 1);       set3(e[6].iend, 1, J, K + 3);
    set3(e[6].nb[0], 1, 0, K);             set3(e[6].nb[2], 2, 0, K + 1);         set3(e[6].origin, 2, 0, K);
    set3(e[7].istart, I + 1, 2, K + 1);    set3(e[7].iend, I + 3, J, K + 3);
    set3(e[7].nb[0], I + 1, 0, K);         set3(e[7].nb[2], I, 0, K + 1);         set3(e[7].origin, I, 0, K);
    set3(e[8].istart, 2, -1, -1);          set3(e[8].iend, I, 1, 1);
    set3(e[8].nb[1], 0, 1, 2);             set3(e[8].nb[2], 0, 2, 1);             set3(e[8].origin, 0, 2, 2);
    set3(e[9].istart, 2, J + 1, -1);       set3(e[9].iend, I, J + 3, 1);
    set3(e[9].nb[1], 0, J + 1, 2);         set3(e[9].nb[2], 0, J, 1);             set3(e[9].origin, 0, J, 2);
    set3(e[10].istart, 2, -1, K + 1);      set3(e[10].iend, I, 1, K + 3);
    set3(e[10].nb[1], 0, 1, K);            set3(e[10].nb[2], 0, 2, K + 1);        set3(e[10].origin, 0, 2, K);
    set3(e[11].istart, 2, J + 1, K + 1);   set3(e[11].iend, I, J + 3, K + 3);
    set3(e[11].nb[1], 0, J + 1, K);        set3(e[11].nb[2], 0, J, K + 1);        set3(e[11].origin, 0, J, K);
    for (int m = 0; m < 12; m++) { e[m].face = F_NONE; e[m].kind = 0; }

    /* edge type selection, src/bc_module.f90:652-866 */
    for (int m = 0; m < 4; m++) {
        int is = covers(b, e[m].nb[0], 0, 1, -1), js = covers(b, e[m].nb[1], 0, 1, -1);
        if (is && js) e[m].fn = (iturb == 0) ? FN_EDGEWWKW : FN_EDGEWW;
        else if (is) {
            e[m].face = (m == 0 || m == 2) ? F_IMIN : F_IMAX;
            e[m].origin[0] = e[m].nb[1][0]; e[m].origin[1] = e[m].nb[1][1];
            e[m].fn = wall_fn(iturb);
        } else if (js) {
            e[m].face = (m == 0 || m == 1) ? F_JMIN : F_JMAX;
            e[m].origin[0] = e[m].nb[0][0]; e[m].origin[1] = e[m].nb[0][1];
            e[m].fn = wall_fn(iturb);
        } else e[m].fn = FN_EDGENOWALL;
        set3(e[m].dir, 0, 0, 1);
    }
    for (int m = 4; m < 8; m++) {
        int is = covers(b, e[m].nb[0], 0, 2, -1), ks = covers(b, e[m].nb[2], 0, 2, -1);
        if (is && ks) e[m].fn = (iturb == 0) ? FN_EDGEWWKW : FN_EDGEWW;
        else if (is) {
            e[m].face = (m == 4 || m == 6) ? F_IMIN : F_IMAX;
            e[m].origin[0] = e[m].nb[2][0]; e[m].origin[2] = e[m].nb[2][2];
            e[m].fn = wall_fn(iturb);
        } else if (ks) {
            e[m].face = (m == 4 || m == 5) ? F_KMIN : F_KMAX;
            e[m].origin[0] = e[m].nb[0][0]; e[m].origin[2] = e[m].nb[0][2];
            e[m].fn = wall_fn(iturb);
        } else e[m].fn = FN_EDGENOWALL;
        set3(e[m].dir, 0, 1, 0);
    }
    for (int m = 8; m < 12; m++) {
        int js = covers(b, e[m].nb[1], 1, 2, -1), ks = covers(b, e[m].nb[2], 1, 2, -1);
        if (js && ks) e[m].fn = (iturb == 0) ? FN_EDGEWWKW : FN_EDGEWW;
        else if (js) {
            e[m].face = (m == 8 || m == 10) ? F_JMIN : F_JMAX;
            e[m].origin[1] = e[m].nb[2][1]; e[m].origin[2] = e[m].nb[2][2];
            e[m].fn = wall_fn(iturb);
        } else if (ks) {
            e[m].face = (m == 8 || m == 9) ? F_KMIN : F_KMAX;
            e[m].origin[1] = e[m].nb[1][1]; e[m].origin[2] = e[m].nb[1][2];
            e[m].fn = wall_fn(iturb);
        } else e[m].fn = FN_EDGENOWALL;
        set3(e[m].dir, 1, 0, 0);
    }

    /* corner table, src/bc_module.f90:872-998 */
    bcinfo *q = b->corner;
    memset(q, 0, sizeof(b->corner));
    const int lo[3] = {-1, -1, -1};
    int hi[3] = {I + 1, J + 1, K + 1};
    for (int m = 0; m < 8; m++) {
        int ib = m & 1, jb = (m >> 1) & 1, kb = (m >> 2) & 1;
        int ig = ib ? I + 1 : 1, jg = jb ? J + 1 : 1, kg = kb ? K + 1 : 1;   /* first ghost */
        int ic = ib ? I : 2, jc = jb ? J : 2, kc = kb ? K : 2;               /* interior */
        for (int d = 0; d < 3; d++) {
            int bit = (d == 0) ? ib : (d == 1 ? jb : kb);
            q[m].istart[d] = bit ? hi[d] : lo[d];
            q[m].iend[d] = bit ? hi[d] + 2 : 1;
        }
        set3(q[m].nb[0], ig, jc, kc);
        set3(q[m].nb[1], ic, jg, kc);
        set3(q[m].nb[2], ic, jc, kg);
        set3(q[m].origin, ic, jc, kc);
        set3(q[m].nb[3], ic, jg, kg);
        set3(q[m].nb[4], ig, jc, kg);
        set3(q[m].nb[5], ig, jg, kc);
        q[m].face = F_NONE;
    }
    /* corner 6 of the reference table has neighbor3 = (imax,1,kmax+1) (:958-960) */
    set3(q[5].nb[2], I, 1, K + 1);
    for (int m = 0; m < 8; m++) {
        bcinfo *cm = &q[m];
        int is = covers(b, cm->nb[0], 0, 1, 2), js = covers(b, cm->nb[1], 0, 1, 2), ks = covers(b, cm->nb[2], 0, 1, 2);
        int ww = (iturb == 0) ? FN_EDGEWWKW : FN_EDGEWW;
        if (is && js && ks) cm->fn = (iturb == 0) ? FN_CORNERWWKW : FN_CORNERWW;
        else if (!is && js && ks) {
            memcpy(cm->origin, cm->nb[0], sizeof(cm->origin));
            set3(cm->nb[0], 0, 0, 0);
            memcpy(cm->nb[1], cm->nb[4], sizeof(cm->origin));
            memcpy(cm->nb[2], cm->nb[5], sizeof(cm->origin));
            cm->fn = ww;
        } else if (is && !js && ks) {
            memcpy(cm->origin, cm->nb[1], sizeof(cm->origin));
            memcpy(cm->nb[0], cm->nb[3], sizeof(cm->origin));
            set3(cm->nb[1], 0, 0, 0);
            memcpy(cm->nb[2], cm->nb[5], sizeof(cm->origin));
            cm->fn = ww;
        } else if (is && js && !ks) {
            memcpy(cm->origin, cm->nb[2], sizeof(cm->origin));
            memcpy(cm->nb[0], cm->nb[3], sizeof(cm->origin));
            memcpy(cm->nb[1], cm->nb[4], sizeof(cm->origin));
            set3(cm->nb[2], 0, 0, 0);
            cm->fn = ww;
        } else if (!is && !js && ks) {
            cm->face = (m < 4) ? F_KMIN : F_KMAX;
            memcpy(cm->origin, cm->nb[5], sizeof(cm->origin));
            cm->fn = wall_fn(iturb);
        } else if (is && !js && !ks) {
            cm->face = (m % 2 == 0) ? F_IMIN : F_IMAX;
            memcpy(cm->origin, cm->nb[3], sizeof(cm->origin));
            cm->fn = wall_fn(iturb);
        } else if (!is && js && !ks) {
            cm->face = (m == 0 || m == 1 || m == 4 || m == 5) ? F_JMIN : F_JMAX;
            memcpy(cm->origin, cm->nb[4], sizeof(cm->origin));
            cm->fn = wall_fn(iturb);
        } else cm->fn = FN_CORNERNOWALL;
    }
}

/* face BC adjustment and function selection, src/bc_module.f90:157-336 */
static int build_face(const orc_ctx *c, bcinfo *bi)
{
    int *s = bi->istart, *e = bi->iend;
    memset(bi->origin, 0, sizeof(bi->origin));
    if (s[0] == e[0]) {
        set3(bi->dir, -1, 1, 1);
        if (s[0] == 1) { bi->face = F_IMIN; s[0] -= 2; bi->origin[0] = 2 * e[0] + 1; }
        else { bi->face = F_IMAX; e[0] += 2; bi->origin[0] = 2 * s[0] - 1; }
    } else if (s[1] == e[1]) {
        set3(bi->dir, 1, -1, 1);
        if (s[1] == 1) { bi->face = F_JMIN; s[1] -= 2; bi->origin[1] = 2 * e[1] + 1; }
        else { bi->face = F_JMAX; e[1] += 2; bi->origin[1] = 2 * s[1] - 1; }
    } else if (s[2] == e[2]) {
        set3(bi->dir, 1, 1, -1);
        if (s[2] == 1) { bi->face = F_KMIN; s[2] -= 2; bi->origin[2] = 2 * e[2] + 1; }
        else { bi->face = F_KMAX; e[2] += 2; bi->origin[2] = 2 * s[2] - 1; }
    } else return -1;
    switch (bi->kind) {
    case ORC_BC_WALL: bi->fn = wall_fn(c->c.iturb); break;
    case ORC_BC_INFLOW_SUBSONIC: bi->fn = FN_INSUB; break;
    case ORC_BC_OUTFLOW_SUBSONIC: bi->fn = FN_OUTSUB; break;
    case ORC_BC_FARFIELD: bi->fn = FN_FARFIELD; break;
    case ORC_BC_SYMMETRY: bi->fn = FN_SYM; break;
    case ORC_BC_INFLOW_SUPERSONIC: bi->fn = FN_INSUPER; break;
    case ORC_BC_OUTFLOW_SUPERSONIC: case ORC_BC_DEGENERATE: bi->fn = FN_COPY; break;
    case ORC_BC_INFLOW: case ORC_BC_OUTFLOW: bi->fn = FN_EMPTY; break;
    default: return -2;
    }
    return 0;
}

/* ---------------- BC functions ---------------- */
/* surface cell of a face BC (select case(face) blocks) */
static void surface(const bcinfo *bi, int i, int j, int k, int *s)
{
    s[0] = bi->origin[0] + bi->dir[0] * i;
    s[1] = bi->origin[1] + bi->dir[1] * j;
    s[2] = bi->origin[2] + bi->dir[2] * k;
    switch (bi->face) {
    case F_IMIN: s[0] = bi->origin[0] + bi->dir[0] * bi->iend[0]; break;
    case F_IMAX: s[0] = bi->origin[0] + bi->dir[0] * bi->istart[0]; break;
    case F_JMIN: s[1] = bi->origin[1] + bi->dir[1] * bi->iend[1]; break;
    case F_JMAX: s[1] = bi->origin[1] + bi->dir[1] * bi->istart[1]; break;
    case F_KMIN: s[2] = bi->origin[2] + bi->dir[2] * bi->iend[2]; break;
    case F_KMAX: s[2] = bi->origin[2] + bi->dir[2] * bi->istart[2]; break;
    default: break;
    }
}

/* outward area vector at the surface cell (in/outflow and farfield sign) */
static void outward(const blk *b, int face, const int *s, double *n)
{
    const double *p;
    double sg = 1.0;
    switch (face) {
    case F_IMIN: p = CX(b, s[0] - 1, s[1], s[2]); sg = -1.0; break;
    case F_IMAX: p = CX(b, s[0], s[1], s[2]); break;
    case F_JMIN: p = EX(b, s[0], s[1] - 1, s[2]); sg = -1.0; break;
    case F_JMAX: p = EX(b, s[0], s[1], s[2]); break;
    case F_KMIN: p = TX(b, s[0], s[1], s[2] - 1); sg = -1.0; break;
    default: p = TX(b, s[0], s[1], s[2]); break;
    }
    n[0] = sg * p[0]; n[1] = sg * p[1]; n[2] = sg * p[2];
}

static void copy_cell(const orc_ctx *c, blk *b, const int *dst, const double *pv, const double *dv, const double *tv)
{
    memcpy(PV(c, b, dst[0], dst[1], dst[2]), pv, sizeof(double) * c->c.npv);
    memcpy(DV(b, dst[0], dst[1], dst[2]), dv, sizeof(double) * 18);
    if (c->c.ntv) memcpy(TV(c, b, dst[0], dst[1], dst[2]), tv, sizeof(double) * c->c.ntv);
}

static void apply_bc(const orc_ctx *c, blk *b, const bcinfo *bi)
{
    const orc_cfg *g = &c->c;
    int npv = g->npv, ntv = g->ntv;
    const double *rp = g->bc_pv;
    if (bi->fn == FN_EMPTY || bi->fn == FN_NONE) return;
    for (int k = bi->istart[2]; k <= bi->iend[2]; k++)
    for (int j = bi->istart[1]; j <= bi->iend[1]; j++)
    for (int i = bi->istart[0]; i <= bi->iend[0]; i++) {
        int g3[3] = {i, j, k};
        int m[3] = {bi->origin[0] + bi->dir[0] * i, bi->origin[1] + bi->dir[1] * j, bi->origin[2] + bi->dir[2] * k};
        double pv[ORC_NPV], dv[18], tv[3], ppv[ORC_NPV];
        memcpy(pv, PV(c, b, m[0], m[1], m[2]), sizeof(double) * npv);
        memcpy(dv, DV(b, m[0], m[1], m[2]), sizeof(double) * 18);
        if (ntv) memcpy(tv, TV(c, b, m[0], m[1], m[2]), sizeof(double) * ntv);
        switch (bi->fn) {
        case FN_WALLVIS:                      /* :2285-2322 */
            pv[1] = -pv[1]; pv[2] = -pv[2]; pv[3] = -pv[3];
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        case FN_WALLVISKW: {                  /* :2368-2443 */
            int s[3];
            surface(bi, i, j, k, s);
            const double *dvb = DV(b, s[0], s[1], s[2]);
            const double *tvb = TV(c, b, s[0], s[1], s[2]);
            const double *grd = GRD(c, b, s[0], s[1], s[2]);
            double var = 1600.0 * tvb[0] / dvb[0] / (grd[4] * grd[4]) - pv[8];
            pv[1] = -pv[1]; pv[2] = -pv[2]; pv[3] = -pv[3]; pv[7] = -pv[7]; pv[8] = var;
            tv[2] = -tv[2];
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        }
        case FN_WALLINV: case FN_SYM: {       /* :2199-2283, :3049-3132 */
            int s[3];
            double n[3];
            surface(bi, i, j, k, s);
            outward(b, bi->face, s, n);
            n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
            const double *ps = PV(c, b, s[0], s[1], s[2]);
            double dn = ps[1] * n[0] + ps[2] * n[1] + ps[3] * n[2];
            double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            double v2 = 2.0 * (ps[1] - n[0] * dn / nn) - pv[1];
            double v3 = 2.0 * (ps[2] - n[1] * dn / nn) - pv[2];
            double v4 = 2.0 * (ps[3] - n[2] * dn / nn) - pv[3];
            pv[1] = v2; pv[2] = v3; pv[3] = v4;
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        }
        case FN_INSUB: case FN_OUTSUB: case FN_FARFIELD: case FN_INSUPER: {
            int s[3];
            double n[3];
            surface(bi, i, j, k, s);
            double ps[ORC_NPV], dvs[18];
            if (bi->fn != FN_INSUPER) {
                outward(b, bi->face, s, n);
                memcpy(ps, PV(c, b, s[0], s[1], s[2]), sizeof(double) * npv);
                memcpy(dvs, DV(b, s[0], s[1], s[2]), sizeof(double) * 18);
            }
            double press = g->bc_pressure;
            if (bi->fn == FN_INSUPER) {       /* :2642-2678 */
                pv[0] = 2.0 * (press - rp[0]) - pv[0];
                for (int q = 1; q < npv; q++) pv[q] = 2.0 * rp[q] - pv[q];
                pv[5] = fmin_f(fmax_f(pv[5], 0.0), 1.0);
                pv[6] = fmin_f(fmax_f(pv[6], 0.0), 1.0);
            } else {
                double dl = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                double vel = (ps[1] * n[0] + ps[2] * n[1] + ps[3] * n[2]) * dl;
                double a = sqrt(orc_sndp2(g, dvs[5], vel * vel));
                int mode;                     /* 0: in-sub, 1: out-sub, 2: copy, 3: in-super */
                if (bi->fn == FN_INSUB) mode = 0;
                else if (bi->fn == FN_OUTSUB) mode = 1;
                else {                        /* farfield :2969-2998 */
                    double mach = vel / sqrt(dvs[5]);
                    if (mach >= 0.0) mode = (mach >= 1.0) ? 2 : 1;
                    else mode = (mach <= -1.0) ? 3 : 0;
                }
                if (mode == 0) {              /* :2605-2614, :2988-2996 */
                    double var = press - rp[0] + ps[0]
                               - dvs[0] * a * (n[0] * (rp[1] - ps[1]) + n[1] * (rp[2] - ps[2]) + n[2] * (rp[3] - ps[3])) * dl;
                    pv[0] = var - pv[0];
                    pv[1] = 2.0 * (rp[1] + n[0] * 0.5 * var / dvs[0] / a * dl) - pv[1];
                    pv[2] = 2.0 * (rp[2] + n[1] * 0.5 * var / dvs[0] / a * dl) - pv[2];
                    pv[3] = 2.0 * (rp[3] + n[2] * 0.5 * var / dvs[0] / a * dl) - pv[3];
                    for (int q = 4; q < npv; q++) pv[q] = 2.0 * rp[q] - pv[q];
                    pv[5] = fmin_f(fmax_f(pv[5], 0.0), 1.0);
                    pv[6] = fmin_f(fmax_f(pv[6], 0.0), 1.0);
                } else if (mode == 1) {       /* :2840-2843, :2976-2979 */
                    pv[0] = 2.0 * (press - rp[0]) - pv[0];
                    pv[1] = 2.0 * (ps[1] + n[0] * (ps[0] + rp[0] - press) / dvs[0] / a * dl) - pv[1];
                    pv[2] = 2.0 * (ps[2] + n[1] * (ps[0] + rp[0] - press) / dvs[0] / a * dl) - pv[2];
                    pv[3] = 2.0 * (ps[3] + n[2] * (ps[0] + rp[0] - press) / dvs[0] / a * dl) - pv[3];
                } else if (mode == 3) {       /* :2983-2986 */
                    pv[0] = 2.0 * (press - rp[0]) - pv[0];
                    for (int q = 1; q < npv; q++) pv[q] = 2.0 * rp[q] - pv[q];
                    pv[5] = fmin_f(fmax_f(pv[5], 0.0), 1.0);
                    pv[6] = fmin_f(fmax_f(pv[6], 0.0), 1.0);
                }
            }
            cell_eos(c, pv[0] + rp[0], pv[4], pv[5], pv[6], dv, tv);
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        }
        case FN_COPY:
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        case FN_EDGEWW: case FN_EDGEWWKW: case FN_CORNERWW: case FN_CORNERWWKW: {   /* :1806-2074 */
            int corner = (bi->fn == FN_CORNERWW || bi->fn == FN_CORNERWWKW);
            int nn = corner ? 6 : 3;
            for (int q = 0; q < npv; q++) ppv[q] = -pv[q];
            for (int t = 0; t < nn; t++) {
                const int *nb = bi->nb[t];
                if (!corner && nb[0] == 0 && nb[1] == 0 && nb[2] == 0) continue;
                const double *pn = PV(c, b, nb[0] + bi->dir[0] * i, nb[1] + bi->dir[1] * j, nb[2] + bi->dir[2] * k);
                for (int q = 0; q < npv; q++) ppv[q] = ppv[q] - pn[q];
            }
            if (bi->fn == FN_EDGEWW || bi->fn == FN_CORNERWW) {
                pv[1] = ppv[1]; pv[2] = ppv[2]; pv[3] = ppv[3];
                if (npv > 7) { pv[7] = ppv[7]; pv[8] = ppv[8]; }
            } else {
                const double *grd = GRD(c, b, m[0], m[1], m[2]);
                double var = 800.0 * tv[0] / dv[0] / (grd[4] * grd[4]);
                double opv = (corner ? 8.0 : 4.0) * var + ppv[8];
                pv[1] = ppv[1]; pv[2] = ppv[2]; pv[3] = ppv[3]; pv[7] = ppv[7]; pv[8] = opv;
            }
            copy_cell(c, b, g3, pv, dv, tv);
            break;
        }
        case FN_EDGENOWALL: case FN_CORNERNOWALL: {   /* :2076-2197 */
            int corner = (bi->fn == FN_CORNERNOWALL);
            int nn = corner ? 6 : 3;
            double tvs[3] = {0.0, 0.0, 0.0};
            if (ntv) memcpy(tvs, tv, sizeof(double) * ntv);
            for (int t = 0; t < nn; t++) {
                const int *nb = bi->nb[t];
                if (!corner && nb[0] == 0 && nb[1] == 0 && nb[2] == 0) continue;
                int p3[3] = {nb[0] + bi->dir[0] * i, nb[1] + bi->dir[1] * j, nb[2] + bi->dir[2] * k};
                const double *pn = PV(c, b, p3[0], p3[1], p3[2]);
                for (int q = 0; q < npv; q++) pv[q] = pv[q] + pn[q];
                if (ntv) {
                    const double *tn = TV(c, b, p3[0], p3[1], p3[2]);
                    for (int q = 0; q < ntv; q++) tvs[q] = tvs[q] + tn[q];
                }
            }
            double f = corner ? 1.0 / 7.0 : 1.0 / 3.0;
            for (int q = 0; q < npv; q++) pv[q] = f * pv[q];
            for (int q = 0; q < ntv; q++) tvs[q] = f * tvs[q];
            memcpy(PV(c, b, i, j, k), pv, sizeof(double) * npv);
            cell_eos(c, pv[0] + rp[0], pv[4], pv[5], pv[6], dv, tvs);
            memcpy(DV(b, i, j, k), dv, sizeof(double) * 18);
            if (ntv) memcpy(TV(c, b, i, j, k), tvs, sizeof(double) * ntv);
            break;
        }
        default: break;
        }
    }
}

/* interface copy, src/bc_module.f90:1261-1386 (same-process branch :1309-1341) */
static void halo(orc_ctx *c)
{
    int npv = c->c.npv, ntv = c->c.ntv;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        for (int n = 0; n < b->ncon; n++) {
            const conn *cn = &b->con[n];
            blk *d = &c->b[cn->donor];
            for (int k = cn->is[2]; k <= cn->ie[2]; k++)
            for (int j = cn->is[1]; j <= cn->ie[1]; j++)
            for (int i = cn->is[0]; i <= cn->ie[0]; i++) {
                int di = i - cn->is[0], dj = j - cn->is[1], dk = k - cn->is[2];
                int ii = cn->tm[0][0] * di + cn->tm[0][1] * dj + cn->tm[0][2] * dk + cn->isd[0];
                int jj = cn->tm[1][0] * di + cn->tm[1][1] * dj + cn->tm[1][2] * dk + cn->isd[1];
                int kk = cn->tm[2][0] * di + cn->tm[2][1] * dj + cn->tm[2][2] * dk + cn->isd[2];
                memcpy(PV(c, d, ii, jj, kk), PV(c, b, i, j, k), sizeof(double) * npv);
                if (ntv) memcpy(TV(c, d, ii, jj, kk), TV(c, b, i, j, k), sizeof(double) * ntv);
                memcpy(DV(d, ii, jj, kk), DV(b, i, j, k), sizeof(double) * 18);
            }
        }
    }
}

int orc_setbc(orc_ctx *c)
{
    for (int bb = 0; bb < c->nb; bb++)
        for (int n = 0; n < c->b[bb].nbc; n++) apply_bc(c, &c->b[bb], &c->b[bb].bc[n]);
    halo(c);
    for (int bb = 0; bb < c->nb; bb++) {
        for (int n = 0; n < 12; n++) apply_bc(c, &c->b[bb], &c->b[bb].edge[n]);
        for (int n = 0; n < 8; n++) apply_bc(c, &c->b[bb], &c->b[bb].corner[n]);
    }
    return 0;
}

/* ---------------- RHS, src/rhs_module.f90:260-643 ---------------- */
static void face_flux(const orc_ctx *c, const double (*x)[ORC_NPV], const double *n3,
                      const double *dvc_l, const double *dvc_r, const double *grdl, const double *grdr,
                      double *fx)
{
    const orc_cfg *g = &c->c;
    double pvl[ORC_NPV], pvr[ORC_NPV], dvl[18], dvr[18];
    if (g->nmuscl > 0) {
        if (g->nmuscl == 1) orc_tvd(g, x, pvl, pvr);
        else orc_mlp(g, x, pvl, pvr, NULL);
        orc_deteos_simple(pvl[0] + g->pref, pvl[4], pvl[5], pvl[6], dvl);
        orc_deteos_simple(pvr[0] + g->pref, pvr[4], pvr[5], pvr[6], dvr);
    } else {
        memcpy(pvl, x[8], sizeof(double) * g->npv);
        memcpy(pvr, x[9], sizeof(double) * g->npv);
        memcpy(dvl, dvc_l, sizeof(dvl));
        memcpy(dvr, dvc_r, sizeof(dvr));
    }
    (void)grdl; (void)grdr;
    double sd[18];
    for (int q = 0; q < 18; q++) sd[q] = x[q][0];      /* x(1:18,1), src/rhs_module.f90:336 */
    if (g->nscheme == 1) orc_roe(g, n3, pvl, pvr, dvl, dvr, fx);
    else if (g->nscheme == 2) orc_roem(g, n3, pvl, pvr, dvl, dvr, sd, fx);
    else if (g->nscheme == 3) orc_ausmpwp(g, n3, pvl, pvr, dvl, dvr, sd, fx);
    else orc_ausmpup(g, n3, pvl, pvr, dvl, dvr, fx, NULL);
}

/* unsteady source of one cell, src/unsteady_module.f90:96-114 */
static void unsteady_source(const orc_cfg *g, const double *pv, const double *dv, const double *grd,
                            const double *q1, const double *q2, double *fx)
{
    int npv = g->npv;
    double b = (g->nsteady == 1) ? 1.0 : 1.5 * 0.5;
    fx[0] = dv[0];
    fx[1] = dv[0] * pv[1];
    fx[2] = dv[0] * pv[2];
    fx[3] = dv[0] * pv[3];
    fx[4] = dv[0] * (dv[1] + 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3])) - (pv[0] + g->pref);
    for (int n = 5; n < npv; n++) fx[n] = dv[0] * pv[n];
    for (int n = 0; n < npv; n++) {
        fx[n] = (double)g->nsteady * (b * fx[n] - q1[n]) + 0.5 * ((double)g->nsteady - 1.0) * q2[n];
        fx[n] = fx[n] * grd[0] / g->dt_phy;
    }
}

static void calrhs_block(orc_ctx *c, blk *b)
{
    const orc_cfg *g = &c->c;
    int npv = g->npv, I = b->imax, J = b->jmax, K = b->kmax;
    int vis = (g->iturb >= -2);
    int nmax = (I > J ? I : J);
    if (K > nmax) nmax = K;
    /* ---- i faces ---- */
#pragma omp parallel num_threads(nthreads)
    {
        double (*ea)[ORC_NPV] = malloc(sizeof(double) * ORC_NPV * (nmax + 2));
        double (*eva)[ORC_NPV] = malloc(sizeof(double) * ORC_NPV * (nmax + 2));
        double x[38][ORC_NPV];
#pragma omp for collapse(2) schedule(static)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++) {
            for (int i = 1; i <= I; i++) {
                const double *nx = CX(b, i, j, k);
                int ll = 0;
                for (int kk = -1; kk <= 1; kk++) for (int jj = -1; jj <= 1; jj++) for (int ii = 0; ii <= 1; ii++)
                    memcpy(x[ll++], PV(c, b, i + ii, j + jj, k + kk), sizeof(double) * npv);
                if (g->nmuscl > 0) {
                    for (int kk = -1; kk <= 1; kk++) for (int jj = -1; jj <= 1; jj++)
                        memcpy(x[ll++], PV(c, b, i - 1, j + jj, k + kk), sizeof(double) * npv);
                    for (int kk = -1; kk <= 1; kk++) for (int jj = -1; jj <= 1; jj++)
                        memcpy(x[ll++], PV(c, b, i + 2, j + jj, k + kk), sizeof(double) * npv);
                    memcpy(x[36], PV(c, b, i - 2, j, k), sizeof(double) * npv);
                    memcpy(x[37], PV(c, b, i + 3, j, k), sizeof(double) * npv);
                }
                const double *grdl = GRD(c, b, i, j, k), *grdr = GRD(c, b, i + 1, j, k);
                face_flux(c, (const double (*)[ORC_NPV])x, nx, DV(b, i, j, k), DV(b, i + 1, j, k), grdl, grdr, ea[i]);
                if (vis) {
                    const double *e4[4] = {EX(b, i, j - 1, k), EX(b, i + 1, j - 1, k), EX(b, i, j, k), EX(b, i + 1, j, k)};
                    const double *t4[4] = {TX(b, i, j, k - 1), TX(b, i + 1, j, k - 1), TX(b, i, j, k), TX(b, i + 1, j, k)};
                    orc_vsflux(g, nx, e4, t4, grdl, grdr, (const double (*)[ORC_NPV])x, DV(b, i, j, k), DV(b, i + 1, j, k),
                               TV(c, b, i, j, k), TV(c, b, i + 1, j, k), eva[i]);
                }
            }
            for (int i = 2; i <= I; i++) {
                double *r = b->res + npv * ii_(b, i, j, k);
                for (int n = 0; n < npv; n++) {
                    r[n] = -(ea[i][n] - ea[i - 1][n]);
                    if (vis) r[n] = r[n] + (eva[i][n] - eva[i - 1][n]);
                }
            }
        }
        /* ---- j faces ---- */
#pragma omp for collapse(2) schedule(static)
        for (int i = 2; i <= I; i++)
        for (int k = 2; k <= K; k++) {
            for (int j = 1; j <= J; j++) {
                const double *nx = EX(b, i, j, k);
                int ll = 0;
                for (int ii = -1; ii <= 1; ii++) for (int kk = -1; kk <= 1; kk++) for (int jj = 0; jj <= 1; jj++)
                    memcpy(x[ll++], PV(c, b, i + ii, j + jj, k + kk), sizeof(double) * npv);
                if (g->nmuscl > 0) {
                    for (int ii = -1; ii <= 1; ii++) for (int kk = -1; kk <= 1; kk++)
                        memcpy(x[ll++], PV(c, b, i + ii, j - 1, k + kk), sizeof(double) * npv);
                    for (int ii = -1; ii <= 1; ii++) for (int kk = -1; kk <= 1; kk++)
                        memcpy(x[ll++], PV(c, b, i + ii, j + 2, k + kk), sizeof(double) * npv);
                    memcpy(x[36], PV(c, b, i, j - 2, k), sizeof(double) * npv);
                    memcpy(x[37], PV(c, b, i, j + 3, k), sizeof(double) * npv);
                }
                const double *grdl = GRD(c, b, i, j, k), *grdr = GRD(c, b, i, j + 1, k);
                face_flux(c, (const double (*)[ORC_NPV])x, nx, DV(b, i, j, k), DV(b, i, j + 1, k), grdl, grdr, ea[j]);
                if (vis) {
                    const double *e4[4] = {TX(b, i, j, k - 1), TX(b, i, j + 1, k - 1), TX(b, i, j, k), TX(b, i, j + 1, k)};
                    const double *t4[4] = {CX(b, i - 1, j, k), CX(b, i, j, k), CX(b, i - 1, j + 1, k), CX(b, i, j + 1, k)};
                    orc_vsflux(g, nx, e4, t4, grdl, grdr, (const double (*)[ORC_NPV])x, DV(b, i, j, k), DV(b, i, j + 1, k),
                               TV(c, b, i, j, k), TV(c, b, i, j + 1, k), eva[j]);
                }
            }
            for (int j = 2; j <= J; j++) {
                double *r = b->res + npv * ii_(b, i, j, k);
                for (int n = 0; n < npv; n++) {
                    r[n] = r[n] - (ea[j][n] - ea[j - 1][n]);
                    if (vis) r[n] = r[n] + (eva[j][n] - eva[j - 1][n]);
                }
            }
        }
        /* ---- k faces ---- */
#pragma omp for collapse(2) schedule(static)
        for (int j = 2; j <= J; j++)
        for (int i = 2; i <= I; i++) {
            for (int k = 1; k <= K; k++) {
                const double *nx = TX(b, i, j, k);
                int ll = 0;
                for (int jj = -1; jj <= 1; jj++) for (int ii = -1; ii <= 1; ii++) for (int kk = 0; kk <= 1; kk++)
                    memcpy(x[ll++], PV(c, b, i + ii, j + jj, k + kk), sizeof(double) * npv);
                if (g->nmuscl > 0) {
                    for (int jj = -1; jj <= 1; jj++) for (int ii = -1; ii <= 1; ii++)
                        memcpy(x[ll++], PV(c, b, i + ii, j + jj, k - 1), sizeof(double) * npv);
                    for (int jj = -1; jj <= 1; jj++) for (int ii = -1; ii <= 1; ii++)
                        memcpy(x[ll++], PV(c, b, i + ii, j + jj, k + 2), sizeof(double) * npv);
                    memcpy(x[36], PV(c, b, i, j, k - 2), sizeof(double) * npv);
                    memcpy(x[37], PV(c, b, i, j, k + 3), sizeof(double) * npv);
                }
                const double *grdl = GRD(c, b, i, j, k), *grdr = GRD(c, b, i, j, k + 1);
                face_flux(c, (const double (*)[ORC_NPV])x, nx, DV(b, i, j, k), DV(b, i, j, k + 1), grdl, grdr, ea[k]);
                if (vis) {
                    const double *e4[4] = {CX(b, i - 1, j, k), CX(b, i - 1, j, k + 1), CX(b, i, j, k), CX(b, i, j, k + 1)};
                    const double *t4[4] = {EX(b, i, j - 1, k), EX(b, i, j, k), EX(b, i, j - 1, k + 1), EX(b, i, j, k + 1)};
                    orc_vsflux(g, nx, e4, t4, grdl, grdr, (const double (*)[ORC_NPV])x, DV(b, i, j, k), DV(b, i, j, k + 1),
                               TV(c, b, i, j, k), TV(c, b, i, j, k + 1), eva[k]);
                }
            }
            for (int k = 2; k <= K; k++) {
                double *r = b->res + npv * ii_(b, i, j, k);
                for (int n = 0; n < npv; n++) {
                    r[n] = r[n] - (ea[k][n] - ea[k - 1][n]);
                    if (vis) r[n] = r[n] + (eva[k][n] - eva[k - 1][n]);
                }
            }
        }
        free(ea);
        free(eva);
        /* ---- cell sources ---- */
#pragma omp for collapse(2) schedule(static)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
        for (int i = 2; i <= I; i++) {
            const double *f6[6] = {CX(b, i - 1, j, k), CX(b, i, j, k), EX(b, i, j - 1, k), EX(b, i, j, k),
                                   TX(b, i, j, k - 1), TX(b, i, j, k)};
            const double *grd = GRD(c, b, i, j, k);
            memcpy(x[0], PV(c, b, i - 1, j, k), sizeof(double) * npv);
            memcpy(x[1], PV(c, b, i, j, k), sizeof(double) * npv);
            memcpy(x[2], PV(c, b, i + 1, j, k), sizeof(double) * npv);
            memcpy(x[3], PV(c, b, i, j - 1, k), sizeof(double) * npv);
            memcpy(x[4], PV(c, b, i, j + 1, k), sizeof(double) * npv);
            memcpy(x[5], PV(c, b, i, j, k - 1), sizeof(double) * npv);
            memcpy(x[6], PV(c, b, i, j, k + 1), sizeof(double) * npv);
            const double *dv = DV(b, i, j, k);
            const double *tv = TV(c, b, i, j, k);
            long q = ii_(b, i, j, k);
            double *r = b->res + npv * q;
            if (g->ncav) {
                double rc, rv, ic[4];
                orc_cavsource(g, (const double (*)[ORC_NPV])x, dv, grd, &rc, &rv, ic);
                r[5] = r[5] + rv - rc;
                memcpy(b->icav + 4 * q, ic, sizeof(ic));
                b->rcav[2 * q] = rc;
                b->rcav[2 * q + 1] = rv;
            }
            if (g->iturb == 0) {
                double src[2], it[4], oc;
                orc_kwsst_source_des(g->ides, f6, grd, (const double (*)[ORC_NPV])x, dv, tv, src, it, &oc);
                b->ocut[q] = oc;
                r[7] = r[7] + src[0];
                r[8] = r[8] + src[1];
                memcpy(b->itt + 4 * q, it, sizeof(it));
            }
            if (g->nsteady > 0) {
                double fx[ORC_NPV];
                const double *q1 = b->qq + npv * (2 * q), *q2 = b->qq + npv * (2 * q + 1);
                unsteady_source(g, x[1], dv, grd, q1, q2, fx);
                for (int n = 0; n < npv; n++) r[n] = r[n] - fx[n];
            }
        }
    }
}

/* ---------------- timestep, src/timestep_module.f90:127-316 ---------------- */
static void timestep(orc_ctx *c, double timeprev, double *time_out)
{
    const orc_cfg *g = &c->c;
    double dtmin = 1.0e10;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        int I = b->imax, J = b->jmax, K = b->kmax;
        if (g->local == -1) {
            long n = (long)(I - 1) * (J - 1) * (K - 1);
            for (long q = 0; q < n; q++) b->dt[q] = g->cfl;
            continue;
        }
        double bmin = 1.0e10;
#pragma omp parallel for collapse(2) reduction(min:bmin) num_threads(nthreads)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
        for (int i = 2; i <= I; i++) {
            const double *f6[6] = {CX(b, i - 1, j, k), CX(b, i, j, k), EX(b, i, j - 1, k), EX(b, i, j, k),
                                   TX(b, i, j, k - 1), TX(b, i, j, k)};
            double dt = orc_localdt(g, f6, GRD(c, b, i, j, k), PV(c, b, i, j, k), DV(b, i, j, k), TV(c, b, i, j, k));
            b->dt[ii_(b, i, j, k)] = dt;
            if (bmin >= dt) bmin = dt;
        }
        if (bmin < dtmin) dtmin = bmin;
    }
    if (g->local == -1) {
        c->time_acc = c->time_acc + g->cfl;
    } else {
        if (g->local == 0)
            for (int bb = 0; bb < c->nb; bb++) {
                blk *b = &c->b[bb];
                long n = (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1);
                for (long q = 0; q < n; q++) b->dt[q] = dtmin;
            }
        c->time_acc = c->time_acc + dtmin;
    }
    *time_out = c->time_acc + timeprev;
}

/* ---------------- LU-SGS, src/update_module.f90:486-732 ---------------- */
static void cellfaces(const blk *b, int i, int j, int k, const double *f6[6])
{
    f6[0] = CX(b, i - 1, j, k); f6[1] = CX(b, i, j, k);
    f6[2] = EX(b, i, j - 1, k); f6[3] = EX(b, i, j, k);
    f6[4] = TX(b, i, j, k - 1); f6[5] = TX(b, i, j, k);
}

static void sweep_cell(orc_ctx *c, blk *b, int i, int j, int k, int fwd)
{
    const orc_cfg *g = &c->c;
    int npv = g->npv;
    static const double zero4[4] = {0, 0, 0, 0};
    double a[ORC_NPV * ORC_NPV];
    long w = iw_(b, i, j, k);
    double *dcv = b->dcv + npv * w;
    int s = fwd ? 1 : -1;
    int nb[3][3] = {{i - s, j, k}, {i, j - s, k}, {i, j, k - s}};
    const double *nrm[3];
    if (fwd) { nrm[0] = CX(b, i - 1, j, k); nrm[1] = EX(b, i, j - 1, k); nrm[2] = TX(b, i, j, k - 1); }
    else { nrm[0] = CX(b, i, j, k); nrm[1] = EX(b, i, j, k); nrm[2] = TX(b, i, j, k); }
    if (fwd) memcpy(dcv, b->res + npv * ii_(b, i, j, k), sizeof(double) * npv);
    for (int d = 0; d < 3; d++) {
        int *p = nb[d];
        orc_caljac(g, nrm[d], GRD(c, b, p[0], p[1], p[2]), PV(c, b, p[0], p[1], p[2]), DV(b, p[0], p[1], p[2]),
                   TV(c, b, p[0], p[1], p[2]), s, a);
        const double *dq = b->dqs + npv * iw_(b, p[0], p[1], p[2]);
        for (int n = 0; n < npv; n++)
            for (int l = 0; l < npv; l++) {
                if (fwd) dcv[n] = dcv[n] + a[l * npv + n] * dq[l];
                else dcv[n] = dcv[n] - a[l * npv + n] * dq[l];
            }
    }
    const double *f6[6];
    cellfaces(b, i, j, k, f6);
    long q = ii_(b, i, j, k);
    orc_getx(g, f6, GRD(c, b, i, j, k), PV(c, b, i, j, k), DV(b, i, j, k), TV(c, b, i, j, k), b->dt[q],
             g->ncav ? b->icav + 4 * q : zero4, g->iturb == 0 ? b->itt + 4 * q : zero4, dcv, b->dqs + npv * w);
}

static void sweeps(orc_ctx *c, blk *b)
{
    int npv = c->c.npv, I = b->imax, J = b->jmax, K = b->kmax;
    long nw = (long)(I + 1) * (J + 1) * (K + 1) * npv;
    memset(b->dcv, 0, sizeof(double) * nw);
    memset(b->dqs, 0, sizeof(double) * nw);
    for (int s = 6; s <= I + J + K; s++) {
#pragma omp parallel for collapse(2) schedule(dynamic, 16) num_threads(nthreads)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++) {
            int i = s - j - k;
            if (i >= 2 && i <= I) sweep_cell(c, b, i, j, k, 1);
        }
    }
    memset(b->dqs, 0, sizeof(double) * nw);
    for (int s = I + J + K; s >= 6; s--) {
#pragma omp parallel for collapse(2) schedule(dynamic, 16) num_threads(nthreads)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++) {
            int i = s - j - k;
            if (i >= 2 && i <= I) sweep_cell(c, b, i, j, k, 0);
        }
    }
}

static void update_block(orc_ctx *c, blk *b)
{
    const orc_cfg *g = &c->c;
    int npv = g->npv, ntv = g->ntv, I = b->imax, J = b->jmax, K = b->kmax;
    double tv3_keep = (ntv == 3) ? TV(c, b, 2, 2, 2)[2] : 0.0;
#pragma omp parallel for collapse(2) num_threads(nthreads)
    for (int k = 2; k <= K; k++)
    for (int j = 2; j <= J; j++)
    for (int i = 2; i <= I; i++) {
        double *pv = PV(c, b, i, j, k);
        const double *dq = b->dqs + npv * iw_(b, i, j, k);
        double p[ORC_NPV], tv[3];
        for (int n = 0; n < npv; n++) p[n] = pv[n] + dq[n];
        p[0] = fmax_f(-g->pref + 1.0e1, p[0]);
        p[5] = fmin_f(1.0, fmax_f(0.0, p[5]));
        p[6] = fmin_f(1.0, fmax_f(0.0, p[6]));
        if (g->iturb == 0) {
            p[7] = fmax_f(p[7], 1.0e-16);
            p[8] = fmax_f(fmax_f(p[8], 1.0e-16), b->ocut[ii_(b, i, j, k)]);
        }
        memcpy(pv, p, sizeof(double) * npv);
        tv[2] = tv3_keep;
        cell_eos(c, p[0] + g->pref, p[4], p[5], p[6], DV(b, i, j, k), tv);
        if (ntv) memcpy(TV(c, b, i, j, k), tv, sizeof(double) * ntv);
    }
    if (g->iturb == 0) {
#pragma omp parallel for collapse(2) num_threads(nthreads)
        for (int k = 2; k <= K; k++)
        for (int j = 2; j <= J; j++)
        for (int i = 2; i <= I; i++) {
            double x[7][ORC_NPV];
            const double *f6[6];
            cellfaces(b, i, j, k, f6);
            memcpy(x[0], PV(c, b, i - 1, j, k), sizeof(double) * npv);
            memcpy(x[1], PV(c, b, i, j, k), sizeof(double) * npv);
            memcpy(x[2], PV(c, b, i + 1, j, k), sizeof(double) * npv);
            memcpy(x[3], PV(c, b, i, j - 1, k), sizeof(double) * npv);
            memcpy(x[4], PV(c, b, i, j + 1, k), sizeof(double) * npv);
            memcpy(x[5], PV(c, b, i, j, k - 1), sizeof(double) * npv);
            memcpy(x[6], PV(c, b, i, j, k + 1), sizeof(double) * npv);
            TV(c, b, i, j, k)[2] = orc_kwsst_eddy(f6, GRD(c, b, i, j, k), (const double (*)[ORC_NPV])x,
                                                  DV(b, i, j, k), TV(c, b, i, j, k));
        }
    }
}

int orc_timestep(orc_ctx *c, int nt_phy, int nt, double timeprev, double *time)
{
    (void)nt;
    double t;
    timestep(c, timeprev, &t);
    if (c->c.nsteady == 0) *time = t;
    else *time = c->c.dt_phy * (double)nt_phy;
    return 0;
}

int orc_timeinteg(orc_ctx *c, int nt_phy, int nt, double timeprev, double *time)
{
    orc_timestep(c, nt_phy, nt, timeprev, time);
    orc_setbc(c);
    for (int bb = 0; bb < c->nb; bb++) calrhs_block(c, &c->b[bb]);
    for (int bb = 0; bb < c->nb; bb++) {
        sweeps(c, &c->b[bb]);
        update_block(c, &c->b[bb]);
    }
    return 0;
}

/* ---------------- residual, src/residual_module.f90:94-186 ---------------- */
int orc_setqres(orc_ctx *c)
{
    int npv = c->c.npv;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        for (int k = 2; k <= b->kmax; k++)
        for (int j = 2; j <= b->jmax; j++)
        for (int i = 2; i <= b->imax; i++)
            memcpy(b->qres + npv * ii_(b, i, j, k), PV(c, b, i, j, k), sizeof(double) * npv);
    }
    return 0;
}

/* l2[npv]; ijkmax[4] = block, i, j, k of the first strict max of |dp|;
 * blocks are visited in order, as ranks would be concatenated */
int orc_residual(orc_ctx *c, double *l2, int *ijkmax, double *resmax)
{
    int npv = c->c.npv;
    double res[ORC_NPV] = {0};
    double rmax = 1.0e-12;
    int im = 0, jm = 0, km = 0, bm = -1;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        for (int k = 2; k <= b->kmax; k++)
        for (int j = 2; j <= b->jmax; j++)
        for (int i = 2; i <= b->imax; i++) {
            double *q = b->qres + npv * ii_(b, i, j, k);
            const double *pv = PV(c, b, i, j, k);
            for (int n = 0; n < npv; n++) q[n] = fabs(q[n] - pv[n]);
            for (int n = 0; n < npv; n++) res[n] = res[n] + q[n] * q[n];
            if (q[0] > rmax) { rmax = q[0]; im = i; jm = j; km = k; bm = bb; }
        }
    }
    for (int n = 0; n < npv; n++) l2[n] = sqrt(res[n]);
    ijkmax[0] = bm; ijkmax[1] = im; ijkmax[2] = jm; ijkmax[3] = km;
    *resmax = rmax;
    return 0;
}

/* src/solve_module.f90:155-183 */
int orc_unsteady_shift(orc_ctx *c)
{
    const orc_cfg *g = &c->c;
    int npv = g->npv;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        for (int k = 2; k <= b->kmax; k++)
        for (int j = 2; j <= b->jmax; j++)
        for (int i = 2; i <= b->imax; i++) {
            long q = ii_(b, i, j, k);
            double *q1 = b->qq + npv * (2 * q), *q2 = b->qq + npv * (2 * q + 1);
            memcpy(q2, q1, sizeof(double) * npv);
            const double *dv = DV(b, i, j, k), *pv = PV(c, b, i, j, k);
            q1[0] = dv[0];
            q1[1] = dv[0] * pv[1];
            q1[2] = dv[0] * pv[2];
            q1[3] = dv[0] * pv[3];
            q1[4] = dv[0] * (dv[1] + 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3])) - pv[0] - g->pref;
            for (int n = 5; n < npv; n++) q1[n] = dv[0] * pv[n];
        }
    }
    return 0;
}

/* ---------------- context / API ---------------- */
orc_ctx *orc_new(const orc_cfg *cfg, int nblocks, const int *dims)
{
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->c = *cfg;
    c->nb = nblocks;
    c->b = calloc(nblocks, sizeof(blk));
    int npv = cfg->npv, ntv = cfg->ntv;
    for (int bb = 0; bb < nblocks; bb++) {
        blk *b = &c->b[bb];
        b->imax = dims[3 * bb]; b->jmax = dims[3 * bb + 1]; b->kmax = dims[3 * bb + 2];
        b->nx = b->imax + 5; b->ny = b->jmax + 5; b->nz = b->kmax + 5;
        long ng = (long)b->nx * b->ny * b->nz;
        long ni = (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1);
        long nw = (long)(b->imax + 1) * (b->jmax + 1) * (b->kmax + 1);
        b->cx = calloc(3 * (long)b->imax * (b->jmax + 1) * (b->kmax + 1), sizeof(double));
        b->ex = calloc(3 * (long)(b->imax + 1) * b->jmax * (b->kmax + 1), sizeof(double));
        b->tx = calloc(3 * (long)(b->imax + 1) * (b->jmax + 1) * b->kmax, sizeof(double));
        b->grd = calloc(cfg->ngrd * nw, sizeof(double));
        b->pv = calloc(npv * ng, sizeof(double));
        b->dv = calloc(18 * ng, sizeof(double));
        b->tv = calloc((ntv ? ntv : 1) * ng, sizeof(double));
        b->qq = calloc(2 * npv * ni, sizeof(double));
        b->res = calloc(npv * ni, sizeof(double));
        b->icav = calloc(4 * ni, sizeof(double));
        b->itt = calloc(4 * ni, sizeof(double));
        b->ocut = calloc(ni, sizeof(double));
        b->dt = calloc(ni, sizeof(double));
        b->rcav = calloc(2 * ni, sizeof(double));
        b->qres = calloc(npv * ni, sizeof(double));
        b->dqs = calloc(npv * nw, sizeof(double));
        b->dcv = calloc(npv * nw, sizeof(double));
    }
    return c;
}

void orc_free(orc_ctx *c)
{
    if (!c) return;
    for (int bb = 0; bb < c->nb; bb++) {
        blk *b = &c->b[bb];
        double *p[] = {b->cx, b->ex, b->tx, b->grd, b->pv, b->dv, b->tv, b->qq, b->res, b->icav, b->itt,
                       b->ocut, b->dt, b->rcav, b->qres, b->dqs, b->dcv};
        for (unsigned q = 0; q < sizeof(p) / sizeof(p[0]); q++) free(p[q]);
        free(b->bc);
        free(b->con);
    }
    free(c->b);
    free(c);
}

int orc_set_grid(orc_ctx *c, int bb, const double *cx, const double *ex, const double *tx, const double *grd)
{
    blk *b = &c->b[bb];
    memcpy(b->cx, cx, sizeof(double) * 3 * (long)b->imax * (b->jmax + 1) * (b->kmax + 1));
    memcpy(b->ex, ex, sizeof(double) * 3 * (long)(b->imax + 1) * b->jmax * (b->kmax + 1));
    memcpy(b->tx, tx, sizeof(double) * 3 * (long)(b->imax + 1) * (b->jmax + 1) * b->kmax);
    memcpy(b->grd, grd, sizeof(double) * c->c.ngrd * (long)(b->imax + 1) * (b->jmax + 1) * (b->kmax + 1));
    return 0;
}

/* desc: nbc x 7 ints {kind, istart[3], iend[3]} in grid_module convention */
int orc_set_bcs(orc_ctx *c, int bb, int nbc, const int *desc)
{
    blk *b = &c->b[bb];
    free(b->bc);
    b->bc = calloc(nbc > 0 ? nbc : 1, sizeof(bcinfo));
    b->nbc = nbc;
    for (int n = 0; n < nbc; n++) {
        bcinfo *bi = &b->bc[n];
        bi->kind = desc[7 * n] & 0xff;
        bi->surface = (desc[7 * n] & ORC_BC_SURFACE) != 0;
        for (int d = 0; d < 3; d++) { bi->istart[d] = desc[7 * n + 1 + d]; bi->iend[d] = desc[7 * n + 4 + d]; }
        for (int d = 0; d < 3; d++) { bi->ostart[d] = bi->istart[d]; bi->oend[d] = bi->iend[d]; }
        if (build_face(c, bi)) return -1;
    }
    return 0;
}

/* desc: ncon x 22 ints {donor, istart[3], iend[3], istart_donor[3], iend_donor[3], transmat row-major[9]} */
int orc_set_conns(orc_ctx *c, int bb, int ncon, const int *desc)
{
    blk *b = &c->b[bb];
    free(b->con);
    b->con = calloc(ncon > 0 ? ncon : 1, sizeof(conn));
    b->ncon = ncon;
    for (int n = 0; n < ncon; n++) {
        const int *d = desc + 22 * n;
        conn *cn = &b->con[n];
        cn->donor = d[0];
        for (int q = 0; q < 3; q++) {
            cn->is[q] = d[1 + q]; cn->ie[q] = d[4 + q]; cn->isd[q] = d[7 + q]; cn->ied[q] = d[10 + q];
        }
        for (int m = 0; m < 3; m++) for (int l = 0; l < 3; l++) cn->tm[m][l] = d[13 + 3 * m + l];
    }
    return 0;
}

int orc_finalize(orc_ctx *c)
{
    if (c->c.iturb == -1) return -1;
    for (int bb = 0; bb < c->nb; bb++) build_edges_corners(c, &c->b[bb]);
    return 0;
}

int orc_put_state(orc_ctx *c, int bb, const double *pv, const double *dv, const double *tv, const double *qq)
{
    blk *b = &c->b[bb];
    long ng = (long)b->nx * b->ny * b->nz;
    long ni = (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1);
    if (pv) memcpy(b->pv, pv, sizeof(double) * c->c.npv * ng);
    if (dv) memcpy(b->dv, dv, sizeof(double) * 18 * ng);
    if (tv && c->c.ntv) memcpy(b->tv, tv, sizeof(double) * c->c.ntv * ng);
    if (qq) memcpy(b->qq, qq, sizeof(double) * 2 * c->c.npv * ni);
    return 0;
}

int orc_get_state(orc_ctx *c, int bb, double *pv, double *dv, double *tv, double *qq)
{
    blk *b = &c->b[bb];
    long ng = (long)b->nx * b->ny * b->nz;
    long ni = (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1);
    if (pv) memcpy(pv, b->pv, sizeof(double) * c->c.npv * ng);
    if (dv) memcpy(dv, b->dv, sizeof(double) * 18 * ng);
    if (tv && c->c.ntv) memcpy(tv, b->tv, sizeof(double) * c->c.ntv * ng);
    if (qq) memcpy(qq, b->qq, sizeof(double) * 2 * c->c.npv * ni);
    return 0;
}

int orc_calrhs(orc_ctx *c, int bb, double *res, double *icav, double *itt, double *omega_cut)
{
    blk *b = &c->b[bb];
    long ni = (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1);
    calrhs_block(c, b);
    if (res) memcpy(res, b->res, sizeof(double) * c->c.npv * ni);
    if (icav) memcpy(icav, b->icav, sizeof(double) * 4 * ni);
    if (itt) memcpy(itt, b->itt, sizeof(double) * 4 * ni);
    if (omega_cut) memcpy(omega_cut, b->ocut, sizeof(double) * ni);
    return 0;
}

int orc_get_dt(orc_ctx *c, int bb, double *dt)
{
    blk *b = &c->b[bb];
    memcpy(dt, b->dt, sizeof(double) * (long)(b->imax - 1) * (b->jmax - 1) * (b->kmax - 1));
    return 0;
}

/* edge/corner descriptors for tests: 20 x 26 ints {fn, face, istart[3], iend[3], origin[3], dir[3], nb[6][3]-first 12} */
int orc_get_bcinfo(orc_ctx *c, int bb, int *out)
{
    blk *b = &c->b[bb];
    for (int m = 0; m < 20; m++) {
        const bcinfo *bi = m < 12 ? &b->edge[m] : &b->corner[m - 12];
        int *o = out + 32 * m;
        o[0] = bi->fn; o[1] = bi->face;
        for (int d = 0; d < 3; d++) {
            o[2 + d] = bi->istart[d]; o[5 + d] = bi->iend[d]; o[8 + d] = bi->origin[d]; o[11 + d] = bi->dir[d];
        }
        for (int t = 0; t < 6; t++) for (int d = 0; d < 3; d++) o[14 + 3 * t + d] = bi->nb[t][d];
    }
    return 0;
}

/* Integrals of the current state (restated for the checker):
 *   out[0] vapour volume: sum over interior cells of dv(1)*pv(6)/dv(4)*grd(1)
 *   out[1] gas volume:    sum of dv(1)*pv(7)/dv(5)*grd(1)
 *          (the vof_y1 / vof_y2 fields, src/postvariable_module.f90:282-295)
 *   out[2..7] pressure and viscous force over the ranges flagged as the
 *          'Surface' family, scaled by 2/(rhoref*uref^2)/area; out[8..9] cl, cd
 *          (clcd_writing, src/postvariable_module.f90:507-633)
 * x[b]: vertex coordinates x(3,2:imax+1,2:jmax+1,2:kmax+1) of block b. */
int orc_integrals(orc_ctx *c, const double *const *x, double rhoref, double uref, double area, double aoa,
                  double *out)
{
    const int npv = c->c.npv;
    double vv = 0.0, vg = 0.0;
    double fp[3] = {0.0, 0.0, 0.0}, fv[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < c->nb; l++) {
        const blk *b = &c->b[l];
        for (int k = 2; k <= b->kmax; k++)
            for (int j = 2; j <= b->jmax; j++)
                for (int i = 2; i <= b->imax; i++) {
                    const double *pv = PV(c, b, i, j, k), *dv = DV(b, i, j, k);
                    const double vol = GRD(c, b, i, j, k)[0];
                    vv += dv[0] * pv[5] / dv[3] * vol;
                    vg += dv[0] * pv[6] / dv[4] * vol;
                }
    }
    for (int l = 0; l < c->nb; l++) {
        const blk *b = &c->b[l];
        const int I = b->imax, J = b->jmax;
#define XX(cc, i, j, k) (x[l][(cc) + 3 * ((long)((i) - 2) + (long)I * ((long)((j) - 2) + (long)J * ((k) - 2)))])
        for (int m = 0; m < b->nbc; m++) {
            const bcinfo *bi = &b->bc[m];
            if (!bi->surface) continue;
            const int *s = bi->ostart, *e = bi->oend;
            for (int k = s[2]; k <= e[2]; k++)
                for (int j = s[1]; j <= e[1]; j++)
                    for (int i = s[0]; i <= e[0]; i++) {
                        double cx[3], xf[3];
                        const double *grd, *pv, *tv;
                        int ci, cj, ck;
                        if (s[0] == e[0]) {
                            if (s[0] == 1) {
                                memcpy(cx, CX(b, i, j, k), sizeof(cx));
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i + 1, j, k) + XX(q, i + 1, j + 1, k) + XX(q, i + 1, j, k + 1) + XX(q, i + 1, j + 1, k + 1));
                                ci = i + 1; cj = j; ck = k;
                            } else {
                                for (int q = 0; q < 3; q++) cx[q] = -CX(b, i - 1, j, k)[q];
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i, j, k) + XX(q, i, j + 1, k) + XX(q, i, j, k + 1) + XX(q, i, j + 1, k + 1));
                                ci = i - 1; cj = j; ck = k;
                            }
                        } else if (s[1] == e[1]) {
                            if (s[1] == 1) {
                                memcpy(cx, EX(b, i, j, k), sizeof(cx));
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i, j + 1, k) + XX(q, i + 1, j + 1, k) + XX(q, i, j + 1, k + 1) + XX(q, i + 1, j + 1, k + 1));
                                ci = i; cj = j + 1; ck = k;
                            } else {
                                for (int q = 0; q < 3; q++) cx[q] = -EX(b, i, j - 1, k)[q];
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i, j, k) + XX(q, i + 1, j, k) + XX(q, i, j, k + 1) + XX(q, i + 1, j, k + 1));
                                ci = i; cj = j - 1; ck = k;
                            }
                        } else {
                            if (s[2] == 1) {
                                memcpy(cx, TX(b, i, j, k), sizeof(cx));
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i, j, k + 1) + XX(q, i + 1, j, k + 1) + XX(q, i, j + 1, k + 1) + XX(q, i + 1, j + 1, k + 1));
                                ci = i; cj = j; ck = k + 1;
                            } else {
                                for (int q = 0; q < 3; q++) cx[q] = -TX(b, i, j, k - 1)[q];
                                for (int q = 0; q < 3; q++)
                                    xf[q] = 0.25 * (XX(q, i, j, k) + XX(q, i + 1, j, k) + XX(q, i, j + 1, k) + XX(q, i + 1, j + 1, k));
                                ci = i; cj = j; ck = k - 1;
                            }
                        }
                        grd = GRD(c, b, ci, cj, ck);
                        pv = PV(c, b, ci, cj, ck);
                        tv = c->c.ntv ? TV(c, b, ci, cj, ck) : NULL;
                        double dfp[3], dfv[3];
                        for (int q = 0; q < 3; q++) dfp[q] = -pv[0] * cx[q];
                        double dl = cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2];
                        double vel = cx[0] * pv[1] + cx[1] * pv[2] + cx[2] * pv[3];
                        double vis;
                        if (c->c.iturb == -1 || c->c.iturb == 0) vis = tv[0] + tv[2];
                        else if (c->c.iturb == -2) vis = tv[0];
                        else vis = 0.0;
                        for (int q = 0; q < 3; q++) {
                            double ut = pv[1 + q] - cx[q] * vel / dl;
                            dfv[q] = vis * ut / sqrt((grd[1] - xf[0]) * (grd[1] - xf[0]) + (grd[2] - xf[1]) * (grd[2] - xf[1]) +
                                                     (grd[3] - xf[2]) * (grd[3] - xf[2]));
                        }
                        for (int q = 0; q < 3; q++) dfv[q] = dfv[q] * sqrt(dl);
                        for (int q = 0; q < 3; q++) { fp[q] += dfp[q]; fv[q] += dfv[q]; }
                    }
        }
#undef XX
    }
    (void)npv;
    const double ref = 2.0 / rhoref / (uref * uref) / area;
    out[0] = vv;
    out[1] = vg;
    for (int q = 0; q < 3; q++) { out[2 + q] = fp[q] * ref; out[5 + q] = fv[q] * ref; }
    out[8] = (out[3] + out[6]) * cos(aoa) - (out[2] + out[5]) * sin(aoa);
    out[9] = (out[3] + out[6]) * sin(aoa) + (out[2] + out[5]) * cos(aoa);
    return 0;
}

/*
 * oracle/oracle.c — CPU oracle for the closest-point / signed-distance path (TEST INFRASTRUCTURE).
 *
 * Restates zulianp/ssdf's algorithms in plain C for checking the HIP library; see oracle.h.  Built
 * with -O2 -fopenmp -ffp-contract=off so that every floating-point expression rounds as written.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min / std::max semantics (return the first argument unless the second is strictly
 * smaller / larger), and std::min({a,b,c}) as a left fold */
#define ORC_MIN(a, b) ((b) < (a) ? (b) : (a))
#define ORC_MAX(a, b) ((a) < (b) ? (b) : (a))
#define ORC_MIN3(a, b, c) ORC_MIN(ORC_MIN(a, b), c)
#define ORC_MAX3(a, b, c) ORC_MAX(ORC_MAX(a, b), c)

#define REAL float
#define SUF f
#define ORC_ABS fabsf
#define ORC_SQRT sqrtf
#define ORC_CEIL ceilf
#define ORC_REAL_MAX FLT_MAX
#include "orc_real.h"
#undef REAL
#undef SUF
#undef ORC_ABS
#undef ORC_SQRT
#undef ORC_CEIL
#undef ORC_REAL_MAX

#define REAL double
#define SUF d
#define ORC_ABS fabs
#define ORC_SQRT sqrt
#define ORC_CEIL ceil
#define ORC_REAL_MAX DBL_MAX
#include "orc_real.h"
#undef REAL
#undef SUF
#undef ORC_ABS
#undef ORC_SQRT
#undef ORC_CEIL
#undef ORC_REAL_MAX

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ================================================================================================
 * Brute-force FP64 argmin (the ground truth for distance and closest-triangle index).
 * Distances are pt_tri_d2_d (serial/ssdf.hpp:50-121 in double); the scan runs in increasing triangle
 * order with a strict '<', so the lowest index wins exact ties — the canonical tie-break.
 * ============================================================================================== */
void orc_brute_argmin(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel,
                      const int *s0, const int *s1, const int *s2, const double *sx, const double *sy,
                      const double *sz, const double *init_d2, double *out_d2, int *out_tri, int *out_ntie,
                      double *out_d2second) {
#pragma omp parallel for schedule(dynamic, 64)
    for (ptrdiff_t p = 0; p < npoints; p++) {
        const double px = x[p], py = y[p], pz = z[p];
        double best = init_d2 ? init_d2[p] : INFINITY, second = INFINITY;
        int bt = -1, ntie = 0;
        for (ptrdiff_t t = 0; t < nsel; t++) {
            const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
            const double d2 = pt_tri_d2_d(px, py, pz, sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1], sx[i2],
                                          sy[i2], sz[i2]);
            if (d2 < best) {
                if (bt >= 0) second = best;
                best = d2, bt = (int)t, ntie = 1;
            } else if (d2 == best && bt >= 0) {
                ntie++;
                second = d2;
            } else if (d2 < second) {
                second = d2;
            }
        }
        out_d2[p] = best;
        out_tri[p] = bt;
        if (out_ntie) out_ntie[p] = ntie;
        if (out_d2second) out_d2second[p] = second;
    }
}

void orc_pair_d2(ptrdiff_t npoints, const double *x, const double *y, const double *z, const int *tri,
                 const int *s0, const int *s1, const int *s2, const double *sx, const double *sy, const double *sz,
                 double *out_d2) {
#pragma omp parallel for
    for (ptrdiff_t p = 0; p < npoints; p++) {
        const int t = tri[p];
        if (t < 0) {
            out_d2[p] = INFINITY;
            continue;
        }
        const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
        out_d2[p] = pt_tri_d2_d(x[p], y[p], z[p], sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1], sx[i2], sy[i2],
                                sz[i2]);
    }
}

/* ================================================================================================
 * LBVH topology (Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees and k-d
 * trees"): shared by the exact FP64 BVH oracle and the canonical FP32 work-count replay.
 * ============================================================================================== */
typedef struct {
    ptrdiff_t n;      /* leaves (= triangles) */
    int *order;       /* sorted position -> triangle index */
    int *left, *right; /* per internal node: >= 0 internal node, < 0 leaf ~k */
    int *parent;      /* per internal node (root: -1) */
    int *leaf_parent; /* per leaf */
} orc_lbvh;

static uint64_t orc_expand21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

typedef struct {
    uint64_t code;
    int idx;
} orc_coderec;

static int orc_coderec_cmp(const void *pa, const void *pb) {
    const orc_coderec *a = (const orc_coderec *)pa, *b = (const orc_coderec *)pb;
    if (a->code != b->code) return a->code < b->code ? -1 : 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

static int orc_delta(const uint64_t *codes, ptrdiff_t n, ptrdiff_t i, ptrdiff_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = codes[i], b = codes[j];
    if (a != b) return __builtin_clzll(a ^ b);
    return 64 + __builtin_clz((unsigned)(i ^ j));
}

/* Morton keys from float triangle-box centres, normalised per axis by the centre bounds:
 * u = (c - lo) / (hi - lo), q = min(floor(u * 2^21), 2^21 - 1). */
static void orc_lbvh_build(orc_lbvh *b, ptrdiff_t n, const float *bmin, const float *bmax) {
    b->n = n;
    b->order = (int *)malloc(sizeof(int) * n);
    const ptrdiff_t ni = n > 1 ? n - 1 : 1;
    b->left = (int *)malloc(sizeof(int) * ni);
    b->right = (int *)malloc(sizeof(int) * ni);
    b->parent = (int *)malloc(sizeof(int) * ni);
    b->leaf_parent = (int *)malloc(sizeof(int) * n);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (ptrdiff_t t = 0; t < n; t++)
        for (int d = 0; d < 3; d++) {
            const float c = (bmin[3 * t + d] + bmax[3 * t + d]) * 0.5f;
            lo[d] = c < lo[d] ? c : lo[d];
            hi[d] = c > hi[d] ? c : hi[d];
        }
    float ext[3];
    for (int d = 0; d < 3; d++) ext[d] = hi[d] - lo[d];
    orc_coderec *recs = (orc_coderec *)malloc(sizeof(orc_coderec) * n);
    for (ptrdiff_t t = 0; t < n; t++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) {
            const float c = (bmin[3 * t + d] + bmax[3 * t + d]) * 0.5f;
            const float u = ext[d] > 0.0f ? (c - lo[d]) / ext[d] : 0.0f;
            float s = u * 2097152.0f;
            uint32_t qi = s > 0.0f ? (uint32_t)s : 0u;
            q[d] = qi > 2097151u ? 2097151u : qi;
        }
        recs[t].code = (orc_expand21(q[0]) << 2) | (orc_expand21(q[1]) << 1) | orc_expand21(q[2]);
        recs[t].idx = (int)t;
    }
    qsort(recs, (size_t)n, sizeof(orc_coderec), orc_coderec_cmp);
    uint64_t *codes = (uint64_t *)malloc(sizeof(uint64_t) * n);
    for (ptrdiff_t k = 0; k < n; k++) codes[k] = recs[k].code, b->order[k] = recs[k].idx;
    free(recs);
    if (n == 1) {
        b->leaf_parent[0] = -1;
        free(codes);
        return;
    }
    b->parent[0] = -1;
    for (ptrdiff_t i = 0; i < n - 1; i++) {
        const int dir = (orc_delta(codes, n, i, i + 1) - orc_delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = orc_delta(codes, n, i, i - dir);
        ptrdiff_t lmax = 2;
        while (orc_delta(codes, n, i, i + lmax * dir) > dmin) lmax <<= 1;
        ptrdiff_t l = 0;
        for (ptrdiff_t t = lmax >> 1; t >= 1; t >>= 1)
            if (orc_delta(codes, n, i, i + (l + t) * dir) > dmin) l += t;
        const ptrdiff_t j = i + l * dir;
        const int dnode = orc_delta(codes, n, i, j);
        ptrdiff_t s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (orc_delta(codes, n, i, i + (s + t) * dir) > dnode) s += t;
        } while (t > 1);
        const ptrdiff_t gamma = i + s * dir + (dir < 0 ? dir : 0);
        const ptrdiff_t lo_ = i < j ? i : j, hi_ = i < j ? j : i;
        if (lo_ == gamma) {
            b->left[i] = ~(int)gamma;
            b->leaf_parent[gamma] = (int)i;
        } else {
            b->left[i] = (int)gamma;
            b->parent[gamma] = (int)i;
        }
        if (hi_ == gamma + 1) {
            b->right[i] = ~(int)(gamma + 1);
            b->leaf_parent[gamma + 1] = (int)i;
        } else {
            b->right[i] = (int)(gamma + 1);
            b->parent[gamma + 1] = (int)i;
        }
    }
    free(codes);
}

static void orc_lbvh_free(orc_lbvh *b) {
    free(b->order), free(b->left), free(b->right), free(b->parent), free(b->leaf_parent);
}

/* Bottom-up union of boxes (6 values per box: min xyz, max xyz), type-generic via macro. */
#define ORC_REFIT(TYPE, NAME)                                                                    \
    static void NAME(const orc_lbvh *b, const TYPE *leaf_box, TYPE *node_box) {                 \
        const ptrdiff_t n = b->n;                                                                \
        if (n < 2) return;                                                                       \
        int *cnt = (int *)calloc((size_t)(n - 1), sizeof(int));                                  \
        for (ptrdiff_t k = 0; k < n; k++) {                                                      \
            int node = b->leaf_parent[k];                                                        \
            while (node >= 0) {                                                                  \
                if (++cnt[node] < 2) break;                                                      \
                const int c[2] = {b->left[node], b->right[node]};                                \
                const TYPE *cb[2];                                                               \
                for (int h = 0; h < 2; h++) cb[h] = c[h] < 0 ? leaf_box + 6 * (~c[h]) : node_box + 6 * c[h]; \
                for (int d = 0; d < 3; d++) {                                                    \
                    node_box[6 * node + d] = ORC_MIN(cb[0][d], cb[1][d]);                        \
                    node_box[6 * node + 3 + d] = ORC_MAX(cb[0][3 + d], cb[1][3 + d]);            \
                }                                                                                \
                node = b->parent[node];                                                          \
            }                                                                                    \
        }                                                                                        \
        free(cnt);                                                                               \
    }
ORC_REFIT(float, orc_refit_f)
ORC_REFIT(double, orc_refit_d)

static int orc_lbvh_maxdepth(const orc_lbvh *b) {
    if (b->n < 2) return 0;
    int *stack = (int *)malloc(sizeof(int) * 2 * (b->n + 1));
    int sp = 0, maxd = 0;
    stack[sp++] = 0, stack[sp++] = 0; /* depth = edges from the root */
    while (sp) {
        const int depth = stack[--sp], node = stack[--sp];
        const int c[2] = {b->left[node], b->right[node]};
        for (int h = 0; h < 2; h++) {
            if (c[h] < 0)
                maxd = depth + 1 > maxd ? depth + 1 : maxd;
            else
                stack[sp++] = c[h], stack[sp++] = depth + 1;
        }
    }
    free(stack);
    return maxd;
}

/* ================================================================================================
 * Exact FP64 argmin through the LBVH.  A box is skipped only when its FP64 distance exceeds the best
 * distance by a margin delta = 1e-13 * L (L >= 1 bounds every coordinate): FP64 rounding moves a
 * computed point-box or point-triangle distance by a few ulps of L, far below delta, so no triangle
 * whose computed d2 could equal or beat the current best is ever skipped and the result is the
 * brute-force argmin bit for bit (checked in tests/test_oracle.py).
 * ============================================================================================== */
void orc_bvh_argmin(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel,
                    const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const double *sx, const double *sy,
                    const double *sz, double *out_d2, int *out_tri) {
    if (nsel == 0) {
        for (ptrdiff_t p = 0; p < npoints; p++) out_d2[p] = INFINITY, out_tri[p] = -1;
        return;
    }
    double L = 1.0;
    for (ptrdiff_t v = 0; v < nsp; v++) {
        L = fmax(L, fabs(sx[v]));
        L = fmax(L, fabs(sy[v]));
        L = fmax(L, fabs(sz[v]));
    }
    float *fmin_ = (float *)calloc(3 * nsel + 1, sizeof(float)), *fmax_ = (float *)calloc(3 * nsel + 1, sizeof(float));
    double *lbox = (double *)malloc(sizeof(double) * 6 * nsel);
    for (ptrdiff_t t = 0; t < nsel; t++) {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const double *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const double a = co[d][iv[0]], b_ = co[d][iv[1]], c = co[d][iv[2]];
            const double lo = ORC_MIN3(a, b_, c), hi = ORC_MAX3(a, b_, c);
            fmin_[3 * t + d] = (float)lo, fmax_[3 * t + d] = (float)hi;
        }
    }
    orc_lbvh b;
    orc_lbvh_build(&b, nsel, fmin_, fmax_);
    free(fmin_), free(fmax_);
    for (ptrdiff_t k = 0; k < nsel; k++) {
        const int t = b.order[k];
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const double *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const double a = co[d][iv[0]], b_ = co[d][iv[1]], c = co[d][iv[2]];
            lbox[6 * k + d] = ORC_MIN3(a, b_, c);
            lbox[6 * k + 3 + d] = ORC_MAX3(a, b_, c);
        }
    }
    double *nbox = (double *)malloc(sizeof(double) * 6 * (nsel > 1 ? nsel - 1 : 1));
    orc_refit_d(&b, lbox, nbox);
    const int maxdepth = orc_lbvh_maxdepth(&b);

#pragma omp parallel
    {
        int *st_node = (int *)malloc(sizeof(int) * (maxdepth + 2));
        double *st_d = (double *)malloc(sizeof(double) * (maxdepth + 2));
#pragma omp for schedule(dynamic, 256)
        for (ptrdiff_t p = 0; p < npoints; p++) {
            const double px = x[p], py = y[p], pz = z[p];
            const double Lq = fmax(L, fmax(fabs(px), fmax(fabs(py), fabs(pz))));
            const double delta = 1e-13 * Lq;
            double best = INFINITY, bound = INFINITY;
            int bt = -1, sp = 0;
            int node = nsel > 1 ? 0 : ~0;
            for (;;) {
                if (node < 0) {
                    const int k = ~node, t = b.order[k];
                    const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                    const double d2 = pt_tri_d2_d(px, py, pz, sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1],
                                                  sx[i2], sy[i2], sz[i2]);
                    if (d2 < best || (d2 == best && t < bt)) {
                        best = d2, bt = t;
                        const double r = sqrt(best) + delta;
                        bound = r * r;
                    }
                    node = INT32_MAX; /* pop */
                } else {
                    const int c[2] = {b.left[node], b.right[node]};
                    double dd[2];
                    for (int h = 0; h < 2; h++) {
                        const double *bx = c[h] < 0 ? lbox + 6 * (~c[h]) : nbox + 6 * c[h];
                        dd[h] = pt_box_d2_d(px, py, pz, bx[0], bx[3], bx[1], bx[4], bx[2], bx[5]);
                    }
                    const int n0 = dd[0] <= bound, n1 = dd[1] <= bound;
                    if (n0 && n1) {
                        const int nearc = dd[0] <= dd[1] ? 0 : 1;
                        st_node[sp] = c[1 - nearc], st_d[sp] = dd[1 - nearc], sp++;
                        node = c[nearc];
                    } else if (n0) {
                        node = c[0];
                    } else if (n1) {
                        node = c[1];
                    } else {
                        node = INT32_MAX;
                    }
                }
                if (node == INT32_MAX) {
                    while (sp > 0) {
                        sp--;
                        if (st_d[sp] <= bound) {
                            node = st_node[sp];
                            break;
                        }
                    }
                    if (node == INT32_MAX) break;
                }
            }
            out_d2[p] = best;
            out_tri[p] = bt;
        }
        free(st_node), free(st_d);
    }
    free(lbox), free(nbox);
    orc_lbvh_free(&b);
}

/* ================================================================================================
 * Canonical FP32 work counts (see oracle.h).
 * ============================================================================================== */
void orc_lbvh_counts(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nsel,
                     const int *s0, const int *s1, const int *s2, const float *sx, const float *sy, const float *sz,
                     int64_t *out_nbox, int64_t *out_ntri, int *out_maxdepth, float *out_d2) {
    float *fmin_ = (float *)calloc(3 * nsel + 1, sizeof(float)), *fmax_ = (float *)calloc(3 * nsel + 1, sizeof(float));
    for (ptrdiff_t t = 0; t < nsel; t++) {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const float *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const float a = co[d][iv[0]], b_ = co[d][iv[1]], c = co[d][iv[2]];
            fmin_[3 * t + d] = ORC_MIN3(a, b_, c), fmax_[3 * t + d] = ORC_MAX3(a, b_, c);
        }
    }
    orc_lbvh b;
    orc_lbvh_build(&b, nsel, fmin_, fmax_);
    float *lbox = (float *)malloc(sizeof(float) * 6 * nsel);
    for (ptrdiff_t k = 0; k < nsel; k++) {
        const int t = b.order[k];
        for (int d = 0; d < 3; d++) lbox[6 * k + d] = fmin_[3 * t + d], lbox[6 * k + 3 + d] = fmax_[3 * t + d];
    }
    free(fmin_), free(fmax_);
    float *nbox = (float *)malloc(sizeof(float) * 6 * (nsel > 1 ? nsel - 1 : 1));
    orc_refit_f(&b, lbox, nbox);
    const int maxdepth = orc_lbvh_maxdepth(&b);
    *out_maxdepth = maxdepth;

#pragma omp parallel
    {
        int *st_node = (int *)malloc(sizeof(int) * (maxdepth + 2));
        float *st_d = (float *)malloc(sizeof(float) * (maxdepth + 2));
#pragma omp for schedule(dynamic, 256)
        for (ptrdiff_t p = 0; p < npoints; p++) {
            const float px = x[p], py = y[p], pz = z[p];
            float best = INFINITY;
            int64_t nb = 0, nt = 0;
            int sp = 0;
            int node = nsel > 1 ? 0 : ~0;
            for (;;) {
                if (node < 0) {
                    const int t = b.order[~node];
                    const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                    const float d2 = pt_tri_d2_f(px, py, pz, sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1], sx[i2],
                                                 sy[i2], sz[i2]);
                    nt++;
                    if (d2 < best) best = d2;
                    node = INT32_MAX;
                } else {
                    const int c[2] = {b.left[node], b.right[node]};
                    float dd[2];
                    for (int h = 0; h < 2; h++) {
                        const float *bx = c[h] < 0 ? lbox + 6 * (~c[h]) : nbox + 6 * c[h];
                        dd[h] = pt_box_d2_f(px, py, pz, bx[0], bx[3], bx[1], bx[4], bx[2], bx[5]);
                    }
                    nb += 2;
                    const int n0 = dd[0] < best, n1 = dd[1] < best;
                    if (n0 && n1) {
                        const int nearc = dd[0] <= dd[1] ? 0 : 1;
                        st_node[sp] = c[1 - nearc], st_d[sp] = dd[1 - nearc], sp++;
                        node = c[nearc];
                    } else if (n0) {
                        node = c[0];
                    } else if (n1) {
                        node = c[1];
                    } else {
                        node = INT32_MAX;
                    }
                }
                if (node == INT32_MAX) {
                    while (sp > 0) {
                        sp--;
                        if (st_d[sp] < best) {
                            node = st_node[sp];
                            break;
                        }
                    }
                    if (node == INT32_MAX) break;
                }
            }
            out_nbox[p] = nb;
            out_ntri[p] = nt;
            if (out_d2) out_d2[p] = best;
        }
        free(st_node), free(st_d);
    }
    free(lbox), free(nbox);
    orc_lbvh_free(&b);
}

/* Leaf order of the canonical LBVH: the triangle index of each leaf, i.e. the triangles stably
 * sorted by the Morton code of their FP32 box centre (the order a correct GPU sort must produce). */
/* Replay of the GPU's wave-coherent walk (plain FP32) over the 4 x 4 x 4 bricks of a grid: per
 * brick, 64 lanes share one walk of the same LBVH, whose references to subtrees of at most 3 leaves
 * name runs of leaves that every lane evaluates.  At an internal node every lane tests both child
 * boxes; the walk takes the child that the majority of the lanes wanting a child find nearer, and
 * stacks the other one if some lane wants it; a popped entry is re-tested by every lane.  Lanes
 * outside the grid take part with a bound of -inf.  Sums over the sampled bricks (stride `every` in
 * brick order): node visits, pops, leaf references and triangle evaluations (each by 64 lanes).
 * This is the packet walk's executed work, against which orc_lbvh_counts' canonical per-query work
 * gives the fraction that is useful. */
static int orc_subtree_leaves(const orc_lbvh *b, int *nl, int node) {
    int c[2] = {b->left[node], b->right[node]}, n = 0;
    for (int h = 0; h < 2; h++) n += c[h] < 0 ? 1 : orc_subtree_leaves(b, nl, c[h]);
    return nl[node] = n;
}

void orc_packet_counts(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, const float *sx, const float *sy,
                       const float *sz, int nx, int ny, int nz, const float *lo, const float *hi, int every,
                       int64_t *out) {
    float *fmin_ = (float *)calloc(3 * nsel + 1, sizeof(float)), *fmax_ = (float *)calloc(3 * nsel + 1, sizeof(float));
    for (ptrdiff_t t = 0; t < nsel; t++) {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const float *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const float a = co[d][iv[0]], b_ = co[d][iv[1]], c = co[d][iv[2]];
            fmin_[3 * t + d] = ORC_MIN3(a, b_, c), fmax_[3 * t + d] = ORC_MAX3(a, b_, c);
        }
    }
    orc_lbvh b;
    orc_lbvh_build(&b, nsel, fmin_, fmax_);
    float *lbox = (float *)malloc(sizeof(float) * 6 * nsel);
    for (ptrdiff_t k = 0; k < nsel; k++) {
        const int t = b.order[k];
        for (int d = 0; d < 3; d++) lbox[6 * k + d] = fmin_[3 * t + d], lbox[6 * k + 3 + d] = fmax_[3 * t + d];
    }
    free(fmin_), free(fmax_);
    const ptrdiff_t ni = nsel > 1 ? nsel - 1 : 1;
    float *nbox = (float *)malloc(sizeof(float) * 6 * ni);
    orc_refit_f(&b, lbox, nbox);
    int *nl = (int *)calloc((size_t)ni, sizeof(int));
    int *first = (int *)calloc((size_t)ni, sizeof(int)); /* first leaf of each subtree (leaves in order) */
    if (nsel > 1) {
        orc_subtree_leaves(&b, nl, 0);
        /* first leaf position: the leftmost leaf below */
        for (ptrdiff_t i = 0; i < ni; i++) {
            int c = (int)i;
            while (c >= 0) c = b.left[c];
            first[i] = ~c;
        }
    }
    float h[3];
    const int n3[3] = {nx, ny, nz};
    for (int d = 0; d < 3; d++) h[d] = n3[d] > 1 ? (hi[d] - lo[d]) / (float)(n3[d] - 1) : 0.0f;
    const int nbx = (nx + 3) / 4, nby = (ny + 3) / 4, nbz = (nz + 3) / 4;
    const ptrdiff_t nbricks = (ptrdiff_t)nbx * nby * nbz;
    int64_t visits = 0, pops = 0, refs = 0, tris = 0, bricks = 0;
#pragma omp parallel reduction(+ : visits, pops, refs, tris, bricks)
    {
        int st[256];
#pragma omp for schedule(dynamic, 16)
        for (ptrdiff_t bi = 0; bi < nbricks; bi += every) {
            const int bx = (int)(bi % nbx), by = (int)((bi / nbx) % nby), bz = (int)(bi / ((ptrdiff_t)nbx * nby));
            float px[64], py[64], pz[64], best[64];
            for (int l = 0; l < 64; l++) {
                const int i = bx * 4 + (l & 3), j = by * 4 + ((l >> 2) & 3), k = bz * 4 + (l >> 4);
                const int ok = i < nx && j < ny && k < nz;
                px[l] = lo[0] + (float)i * h[0], py[l] = lo[1] + (float)j * h[1], pz[l] = lo[2] + (float)k * h[2];
                best[l] = ok ? INFINITY : -INFINITY;
            }
            bricks++;
            int sp = 0;
            /* references: >= 0 internal node to visit; < 0 a run ~(first leaf) with count cnt */
            int node = nsel > 1 ? 0 : ~0, cnt = 1;
#define ORC_REF(c, ref_, cnt_)                                \
    do {                                                      \
        if ((c) < 0) (ref_) = (c), (cnt_) = 1;                \
        else if (nl[(c)] <= 3) (ref_) = ~first[(c)], (cnt_) = nl[(c)]; \
        else (ref_) = (c), (cnt_) = 0;                        \
    } while (0)
            if (node == 0 && nl[0] <= 3) node = ~first[0], cnt = nl[0];
            int stcnt[256];
            for (;;) {
                while (node >= 0) {
                    visits++;
                    const int c[2] = {b.left[node], b.right[node]};
                    const float *cb[2];
                    for (int hh = 0; hh < 2; hh++) cb[hh] = c[hh] < 0 ? lbox + 6 * (~c[hh]) : nbox + 6 * c[hh];
                    int nlw = 0, nrw = 0, nlt = 0, nany = 0;
                    for (int l = 0; l < 64; l++) {
                        const float d0 = pt_box_d2_f(px[l], py[l], pz[l], cb[0][0], cb[0][3], cb[0][1], cb[0][4],
                                                     cb[0][2], cb[0][5]);
                        const float d1 = pt_box_d2_f(px[l], py[l], pz[l], cb[1][0], cb[1][3], cb[1][1], cb[1][4],
                                                     cb[1][2], cb[1][5]);
                        const int wl = d0 < best[l], wr = d1 < best[l];
                        nlw += wl, nrw += wr, nlt += wl && (!wr || d0 <= d1), nany += wl || wr;
                    }
                    if (nany == 0) {
                        node = INT32_MIN;
                        break;
                    }
                    const int lf = 2 * nlt >= nany;
                    const int take = lf ? c[0] : c[1], other = lf ? c[1] : c[0];
                    if (lf ? nrw > 0 : nlw > 0) {
                        int r, n;
                        ORC_REF(other, r, n);
                        st[sp] = r, stcnt[sp] = n, sp++;
                    }
                    ORC_REF(take, node, cnt);
                }
                if (node != INT32_MIN) { /* a run of cnt leaves from ~node */
                    refs++;
                    for (int j = 0; j < cnt; j++) {
                        const int t = b.order[~node + j];
                        const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                        tris++;
                        for (int l = 0; l < 64; l++) {
                            const float d2 = pt_tri_d2_f(px[l], py[l], pz[l], sx[i0], sy[i0], sz[i0], sx[i1], sy[i1],
                                                         sz[i1], sx[i2], sy[i2], sz[i2]);
                            if (d2 < best[l]) best[l] = d2;
                        }
                    }
                }
                node = INT32_MIN;
                while (sp > 0) {
                    --sp;
                    pops++;
                    const int r = st[sp], n = stcnt[sp];
                    /* the box of the pending reference: a node's, a leaf's, or a run's subtree's */
                    const float *bx;
                    if (r >= 0)
                        bx = nbox + 6 * r;
                    else if (n == 1)
                        bx = lbox + 6 * (~r);
                    else
                        bx = NULL;
                    float rb[6];
                    if (!bx) { /* run: union of its leaves' boxes */
                        for (int d = 0; d < 3; d++) rb[d] = INFINITY, rb[3 + d] = -INFINITY;
                        for (int j = 0; j < n; j++)
                            for (int d = 0; d < 3; d++) {
                                rb[d] = ORC_MIN(rb[d], lbox[6 * (~r + j) + d]);
                                rb[3 + d] = ORC_MAX(rb[3 + d], lbox[6 * (~r + j) + 3 + d]);
                            }
                        bx = rb;
                    }
                    int any = 0;
                    for (int l = 0; l < 64 && !any; l++)
                        any = pt_box_d2_f(px[l], py[l], pz[l], bx[0], bx[3], bx[1], bx[4], bx[2], bx[5]) < best[l];
                    if (any) {
                        node = r, cnt = n;
                        break;
                    }
                }
                if (node == INT32_MIN) break;
            }
#undef ORC_REF
        }
    }
    out[0] = bricks, out[1] = visits, out[2] = pops, out[3] = refs, out[4] = tris;
    free(lbox), free(nbox), free(nl), free(first);
    orc_lbvh_free(&b);
}

void orc_lbvh_order(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, const float *sx, const float *sy,
                    const float *sz, int *out_order) {
    float *fmin_ = (float *)calloc(3 * nsel + 1, sizeof(float)), *fmax_ = (float *)calloc(3 * nsel + 1, sizeof(float));
    for (ptrdiff_t t = 0; t < nsel; t++) {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const float *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const float a = co[d][iv[0]], b_ = co[d][iv[1]], c = co[d][iv[2]];
            fmin_[3 * t + d] = ORC_MIN3(a, b_, c), fmax_[3 * t + d] = ORC_MAX3(a, b_, c);
        }
    }
    orc_lbvh b;
    orc_lbvh_build(&b, nsel, fmin_, fmax_);
    for (ptrdiff_t k = 0; k < nsel; k++) out_order[k] = b.order[k];
    free(fmin_), free(fmax_);
    orc_lbvh_free(&b);
}

/* ================================================================================================
 * Pseudonormals; follows serial/bvh/bvh.cpp:1669-1721.  The reference pairs a face edge with the
 * first face seen on the same undirected edge through a hash map keyed min*V + max (1707-1719); a
 * sort by (key, 3t + e) visits the same pairs in the same order: the first face f0 of a key sums
 * with every later face fk, edge_n[fk] = n[f0] + n[fk], and edge_n[f0] keeps the last such sum.
 * ============================================================================================== */
typedef struct {
    uint64_t key;
    ptrdiff_t slot; /* 3 t + e */
} orc_edgerec;

static int orc_edgerec_cmp(const void *pa, const void *pb) {
    const orc_edgerec *a = (const orc_edgerec *)pa, *b = (const orc_edgerec *)pb;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return (a->slot > b->slot) - (a->slot < b->slot);
}

ptrdiff_t orc_pseudonormals(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp,
                            const double *sx, const double *sy, const double *sz, double *face_n, double *edge_n,
                            double *vertex_n) {
    const int *s[3] = {s0, s1, s2};
    memset(vertex_n, 0, sizeof(double) * 3 * nsp);
    for (ptrdiff_t t = 0; t < nsel; t++) {
        double p[3][3];
        for (int v = 0; v < 3; v++) p[v][0] = sx[s[v][t]], p[v][1] = sy[s[v][t]], p[v][2] = sz[s[v][t]];
        const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double w[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len > 0)
            for (int d = 0; d < 3; d++) n[d] /= len;
        for (int d = 0; d < 3; d++) face_n[3 * t + d] = n[d];
        /* angle-weighted vertex normals (1694-1703) */
        for (int v = 0; v < 3; v++) {
            const double *a = p[v], *b = p[(v + 1) % 3], *c = p[(v + 2) % 3];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            const double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
            const double cosine = (l1 > 0 && l2 > 0) ? (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]) / (l1 * l2) : 1.0;
            const double angle = acos(fmax(-1.0, fmin(1.0, cosine)));
            for (int d = 0; d < 3; d++) vertex_n[3 * (ptrdiff_t)s[v][t] + d] += angle * n[d];
        }
        for (int e = 0; e < 3; e++)
            for (int d = 0; d < 3; d++) edge_n[9 * t + 3 * e + d] = n[d];
    }
    orc_edgerec *recs = (orc_edgerec *)malloc(sizeof(orc_edgerec) * 3 * (nsel ? nsel : 1));
    for (ptrdiff_t t = 0; t < nsel; t++)
        for (int e = 0; e < 3; e++) {
            const uint64_t a = (uint64_t)s[e][t], b = (uint64_t)s[(e + 1) % 3][t];
            recs[3 * t + e].key = (a < b ? a : b) * (uint64_t)nsp + (a < b ? b : a);
            recs[3 * t + e].slot = 3 * t + e;
        }
    qsort(recs, (size_t)(3 * nsel), sizeof(orc_edgerec), orc_edgerec_cmp);
    ptrdiff_t open_edges = 0;
    for (ptrdiff_t i = 0; i < 3 * nsel;) {
        ptrdiff_t j = i + 1;
        while (j < 3 * nsel && recs[j].key == recs[i].key) j++;
        if (j - i != 2) open_edges++;
        const ptrdiff_t f0 = recs[i].slot;
        for (ptrdiff_t k = i + 1; k < j; k++) {
            const ptrdiff_t fk = recs[k].slot;
            for (int d = 0; d < 3; d++) {
                const double sum = face_n[3 * (f0 / 3) + d] + face_n[3 * (fk / 3) + d];
                edge_n[9 * (fk / 3) + 3 * (fk % 3) + d] = sum;
                edge_n[9 * (f0 / 3) + 3 * (f0 % 3) + d] = sum;
            }
        }
        i = j;
    }
    free(recs);
    return open_edges;
}

/* ================================================================================================
 * Sign at q from its closest triangle; follows serial/bvh/bvh.cpp:1734-1792.  The reference's closest
 * point is cuBQL's single-precision P; here it is pt_tri_cp in FP64 (serial/ssdf.hpp:143-239).
 * ============================================================================================== */
void orc_sign(ptrdiff_t npoints, const double *x, const double *y, const double *z, const int *tri, const int *s0,
              const int *s1, const int *s2, const double *sx, const double *sy, const double *sz,
              const double *face_n, const double *edge_n, const double *vertex_n, int *out_sign, int *out_region,
              double *out_dot) {
    const int *s[3] = {s0, s1, s2};
    const double tol = 1e-4; /* 1738 */
#pragma omp parallel for schedule(static)
    for (ptrdiff_t p = 0; p < npoints; p++) {
        const ptrdiff_t t = tri[p];
        if (t < 0) {
            out_sign[p] = 1, out_region[p] = -1, out_dot[p] = 0;
            continue;
        }
        const double q[3] = {x[p], y[p], z[p]};
        double a[3][3];
        for (int v = 0; v < 3; v++) a[v][0] = sx[s[v][t]], a[v][1] = sy[s[v][t]], a[v][2] = sz[s[v][t]];
        double c[3];
        pt_tri_cp_d(q[0], q[1], q[2], a[0][0], a[0][1], a[0][2], a[1][0], a[1][1], a[1][2], a[2][0], a[2][1],
                    a[2][2], c, c + 1, c + 2);
        /* barycentric coordinates of c (1758-1769) */
        const double e1[3] = {a[1][0] - a[0][0], a[1][1] - a[0][1], a[1][2] - a[0][2]};
        const double e2[3] = {a[2][0] - a[0][0], a[2][1] - a[0][1], a[2][2] - a[0][2]};
        const double r[3] = {c[0] - a[0][0], c[1] - a[0][1], c[2] - a[0][2]};
        const double d11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
        const double d12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
        const double d22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
        const double r1 = r[0] * e1[0] + r[1] * e1[1] + r[2] * e1[2];
        const double r2 = r[0] * e2[0] + r[1] * e2[1] + r[2] * e2[2];
        const double det = d11 * d22 - d12 * d12;
        const double l1 = det > 0 ? (d22 * r1 - d12 * r2) / det : 0;
        const double l2 = det > 0 ? (d11 * r2 - d12 * r1) / det : 0;
        const double l[3] = {1 - l1 - l2, l1, l2};
        /* region (1771-1788) */
        int small = 0, last_small = -1, last_large = -1;
        for (int v = 0; v < 3; v++) {
            if (l[v] < tol)
                small++, last_small = v;
            else
                last_large = v;
        }
        const double *n;
        if (small == 0)
            n = &face_n[3 * t];
        else if (small == 1)
            n = &edge_n[9 * t + 3 * ((last_small + 1) % 3)];
        else
            n = &vertex_n[3 * (ptrdiff_t)s[last_large < 0 ? 0 : last_large][t]];
        const double dot = (q[0] - c[0]) * n[0] + (q[1] - c[1]) * n[1] + (q[2] - c[2]) * n[2];
        out_sign[p] = dot < 0 ? -1 : 1;
        out_region[p] = small == 0 ? 0 : (small == 1 ? 1 : 2);
        out_dot[p] = dot;
    }
}

/* ================================================================================================
 * Generalized winding number, brute force FP64: w(q) = sum_t Omega_t(q) / (4 pi), with
 * tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|), a, b, c the vertices
 * relative to q (Van Oosterom and Strackee, IEEE TBME 30(2), 1983).
 * ============================================================================================== */
void orc_gwn(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel, const int *s0,
             const int *s1, const int *s2, const double *sx, const double *sy, const double *sz, double *out_w) {
#pragma omp parallel for schedule(dynamic, 16)
    for (ptrdiff_t p = 0; p < npoints; p++) {
        double acc = 0;
        for (ptrdiff_t t = 0; t < nsel; t++) {
            const int iv[3] = {s0[t], s1[t], s2[t]};
            double v[3][3];
            for (int k = 0; k < 3; k++) v[k][0] = sx[iv[k]] - x[p], v[k][1] = sy[iv[k]] - y[p], v[k][2] = sz[iv[k]] - z[p];
            const double la = sqrt(v[0][0] * v[0][0] + v[0][1] * v[0][1] + v[0][2] * v[0][2]);
            const double lb = sqrt(v[1][0] * v[1][0] + v[1][1] * v[1][1] + v[1][2] * v[1][2]);
            const double lc = sqrt(v[2][0] * v[2][0] + v[2][1] * v[2][1] + v[2][2] * v[2][2]);
            const double bxc[3] = {v[1][1] * v[2][2] - v[1][2] * v[2][1], v[1][2] * v[2][0] - v[1][0] * v[2][2],
                                   v[1][0] * v[2][1] - v[1][1] * v[2][0]};
            const double num = v[0][0] * bxc[0] + v[0][1] * bxc[1] + v[0][2] * bxc[2];
            const double ab = v[0][0] * v[1][0] + v[0][1] * v[1][1] + v[0][2] * v[1][2];
            const double ac = v[0][0] * v[2][0] + v[0][1] * v[2][1] + v[0][2] * v[2][2];
            const double bc = v[1][0] * v[2][0] + v[1][1] * v[2][1] + v[1][2] * v[2][2];
            const double den = la * lb * lc + ab * lc + ac * lb + bc * la;
            acc += 2.0 * atan2(num, den);
        }
        out_w[p] = acc / (4.0 * M_PI);
    }
}

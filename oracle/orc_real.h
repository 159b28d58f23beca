/*
 * oracle/orc_real.h — precision-generic body of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Included twice by oracle.c, once with REAL=float / SUF=f and once with REAL=double / SUF=d, so
 * that every routine exists in both precisions with identical operation order.  Compiled with
 * -ffp-contract=off: each expression below rounds exactly as written, which is what makes the FP64
 * variants bit-reproducible by the GPU's FP64 fixup path.
 *
 * Nothing in the product links this file; only tests/, the smoke check and bench.py's CPU-baseline
 * leg load liboracle.so.
 */

#ifndef REAL
#error "define REAL and SUF before including orc_real.h"
#endif

#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define FN(name) ORC_CAT(name, SUF)

/* Squared distance point -> triangle by Voronoi regions (Ericson, Real-Time Collision Detection
 * 5.1.5).  Follows serial/ssdf.hpp:50-121 term for term: vertex A (70), vertex B (75), edge AB
 * (77-85), vertex C (90), edge AC (92-100), edge BC (102-110), face (112-120). */
static inline REAL FN(pt_tri_d2)(const REAL px, const REAL py, const REAL pz, const REAL ax, const REAL ay,
                                 const REAL az, const REAL bx, const REAL by, const REAL bz, const REAL cx,
                                 const REAL cy, const REAL cz) {
    const REAL abx = bx - ax, aby = by - ay, abz = bz - az;
    const REAL acx = cx - ax, acy = cy - ay, acz = cz - az;
    const REAL apx = px - ax, apy = py - ay, apz = pz - az;
    const REAL d1 = abx * apx + aby * apy + abz * apz;
    const REAL d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= (REAL)0 && d2 <= (REAL)0) return apx * apx + apy * apy + apz * apz;

    const REAL bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const REAL d3 = abx * bpx + aby * bpy + abz * bpz;
    const REAL d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= (REAL)0 && d4 <= d3) return bpx * bpx + bpy * bpy + bpz * bpz;

    const REAL vc = d1 * d4 - d3 * d2;
    if (vc <= (REAL)0 && d1 >= (REAL)0 && d3 <= (REAL)0) {
        const REAL v = d1 / (d1 - d3);
        const REAL qx = ax + v * abx, qy = ay + v * aby, qz = az + v * abz;
        const REAL dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }

    const REAL cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const REAL d5 = abx * cpx + aby * cpy + abz * cpz;
    const REAL d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= (REAL)0 && d5 <= d6) return cpx * cpx + cpy * cpy + cpz * cpz;

    const REAL vb = d5 * d2 - d1 * d6;
    if (vb <= (REAL)0 && d2 >= (REAL)0 && d6 <= (REAL)0) {
        const REAL w = d2 / (d2 - d6);
        const REAL qx = ax + w * acx, qy = ay + w * acy, qz = az + w * acz;
        const REAL dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }

    const REAL va = d3 * d6 - d5 * d4;
    if (va <= (REAL)0 && (d4 - d3) >= (REAL)0 && (d5 - d6) >= (REAL)0) {
        const REAL w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const REAL qx = bx + w * (cx - bx), qy = by + w * (cy - by), qz = bz + w * (cz - bz);
        const REAL dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }

    const REAL denom = (REAL)1 / (va + vb + vc);
    const REAL v = vb * denom;
    const REAL w = vc * denom;
    const REAL qx = ax + abx * v + acx * w;
    const REAL qy = ay + aby * v + acy * w;
    const REAL qz = az + abz * v + acz * w;
    const REAL dx = px - qx, dy = py - qy, dz = pz - qz;
    return dx * dx + dy * dy + dz * dz;
}

/* Closest point on the triangle; follows serial/ssdf.hpp:143-239 (same regions as pt_tri_d2; the
 * edge-BC branch forms (c - b) once, as ssdf.hpp:220 does). */
static inline void FN(pt_tri_cp)(const REAL px, const REAL py, const REAL pz, const REAL ax, const REAL ay,
                                 const REAL az, const REAL bx, const REAL by, const REAL bz, const REAL cx,
                                 const REAL cy, const REAL cz, REAL *ox, REAL *oy, REAL *oz) {
    const REAL abx = bx - ax, aby = by - ay, abz = bz - az;
    const REAL acx = cx - ax, acy = cy - ay, acz = cz - az;
    const REAL apx = px - ax, apy = py - ay, apz = pz - az;
    const REAL d1 = abx * apx + aby * apy + abz * apz;
    const REAL d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= (REAL)0 && d2 <= (REAL)0) {
        *ox = ax, *oy = ay, *oz = az;
        return;
    }
    const REAL bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const REAL d3 = abx * bpx + aby * bpy + abz * bpz;
    const REAL d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= (REAL)0 && d4 <= d3) {
        *ox = bx, *oy = by, *oz = bz;
        return;
    }
    const REAL vc = d1 * d4 - d3 * d2;
    if (vc <= (REAL)0 && d1 >= (REAL)0 && d3 <= (REAL)0) {
        const REAL v = d1 / (d1 - d3);
        *ox = ax + v * abx, *oy = ay + v * aby, *oz = az + v * abz;
        return;
    }
    const REAL cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const REAL d5 = abx * cpx + aby * cpy + abz * cpz;
    const REAL d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= (REAL)0 && d5 <= d6) {
        *ox = cx, *oy = cy, *oz = cz;
        return;
    }
    const REAL vb = d5 * d2 - d1 * d6;
    if (vb <= (REAL)0 && d2 >= (REAL)0 && d6 <= (REAL)0) {
        const REAL w = d2 / (d2 - d6);
        *ox = ax + w * acx, *oy = ay + w * acy, *oz = az + w * acz;
        return;
    }
    const REAL va = d3 * d6 - d5 * d4;
    if (va <= (REAL)0 && (d4 - d3) >= (REAL)0 && (d5 - d6) >= (REAL)0) {
        const REAL w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const REAL bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
        *ox = bx + w * bcx, *oy = by + w * bcy, *oz = bz + w * bcz;
        return;
    }
    const REAL denom = (REAL)1 / (va + vb + vc);
    const REAL v = vb * denom;
    const REAL w = vc * denom;
    *ox = ax + abx * v + acx * w;
    *oy = ay + aby * v + acy * w;
    *oz = az + abz * v + acz * w;
}

/* Squared point -> AABB distance; follows serial/edf/edf.hpp:55-69. */
static inline REAL FN(pt_box_d2)(const REAL px, const REAL py, const REAL pz, const REAL minx, const REAL maxx,
                                 const REAL miny, const REAL maxy, const REAL minz, const REAL maxz) {
    const REAL dx = (px < minx) ? (minx - px) : (px > maxx ? px - maxx : (REAL)0);
    const REAL dy = (py < miny) ? (miny - py) : (py > maxy ? py - maxy : (REAL)0);
    const REAL dz = (pz < minz) ? (minz - pz) : (pz > maxz ? pz - maxz : (REAL)0);
    return dx * dx + dy * dy + dz * dz;
}

/* serial/ssdf.hpp:412-425: strict '<' */
static inline int FN(box_can_improve)(const REAL px, const REAL py, const REAL pz, const REAL best_sq,
                                      const REAL minx, const REAL maxx, const REAL miny, const REAL maxy,
                                      const REAL minz, const REAL maxz) {
    return FN(pt_box_d2)(px, py, pz, minx, maxx, miny, maxy, minz, maxz) < best_sq;
}

/* Signed distance to an AABB (negative inside); follows serial/edf/edf.hpp:193-221. */
REAL FN(orc_pt_box_sdist)(const REAL px, const REAL py, const REAL pz, const REAL minx, const REAL maxx,
                          const REAL miny, const REAL maxy, const REAL minz, const REAL maxz) {
    const REAL hx = (maxx - minx) * (REAL)0.5, hy = (maxy - miny) * (REAL)0.5, hz = (maxz - minz) * (REAL)0.5;
    const REAL cx = (maxx + minx) * (REAL)0.5, cy = (maxy + miny) * (REAL)0.5, cz = (maxz + minz) * (REAL)0.5;
    const REAL qx = ORC_ABS(px - cx) - hx, qy = ORC_ABS(py - cy) - hy, qz = ORC_ABS(pz - cz) - hz;
    const REAL ox = qx > (REAL)0 ? qx : (REAL)0, oy = qy > (REAL)0 ? qy : (REAL)0, oz = qz > (REAL)0 ? qz : (REAL)0;
    const REAL outside = ORC_SQRT(ox * ox + oy * oy + oz * oz);
    const REAL mq = ORC_MAX(qx, ORC_MAX(qy, qz));
    const REAL inside = mq < (REAL)0 ? mq : (REAL)0;
    return outside + inside;
}

/* Bounds of a point cloud (npoints >= 1), seeded with point 0 and reduced by
 * min(a, b) = b < a ? b : a, max(a, b) = a < b ? b : a; follows serial/edf/edf.hpp:500-539.
 * box6 = xmin, xmax, ymin, ymax, zmin, zmax. */
void FN(orc_compute_aabb)(const ptrdiff_t n, const REAL *x, const REAL *y, const REAL *z, REAL *box6) {
    REAL b[6] = {x[0], x[0], y[0], y[0], z[0], z[0]};
    for (ptrdiff_t i = 0; i < n; i++) {
        b[0] = ORC_MIN(b[0], x[i]), b[1] = ORC_MAX(b[1], x[i]);
        b[2] = ORC_MIN(b[2], y[i]), b[3] = ORC_MAX(b[3], y[i]);
        b[4] = ORC_MIN(b[4], z[i]), b[5] = ORC_MAX(b[5], z[i]);
    }
    for (int k = 0; k < 6; k++) box6[k] = b[k];
}

/* Signed distance of every point to a box; follows serial/edf/edf.hpp:541-557. */
void FN(orc_all_points_aabb_sdist)(const ptrdiff_t n, const REAL *x, const REAL *y, const REAL *z, const REAL *box6,
                                   REAL *out) {
    for (ptrdiff_t i = 0; i < n; i++)
        out[i] = FN(orc_pt_box_sdist)(x[i], y[i], z[i], box6[0], box6[1], box6[2], box6[3], box6[4], box6[5]);
}

REAL FN(orc_pt_box_d2)(const REAL px, const REAL py, const REAL pz, const REAL minx, const REAL maxx,
                       const REAL miny, const REAL maxy, const REAL minz, const REAL maxz) {
    return FN(pt_box_d2)(px, py, pz, minx, maxx, miny, maxy, minz, maxz);
}

REAL FN(orc_pt_tri_d2)(const REAL *p, const REAL *a, const REAL *b, const REAL *c) {
    return FN(pt_tri_d2)(p[0], p[1], p[2], a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]);
}

void FN(orc_pt_tri_cp)(const REAL *p, const REAL *a, const REAL *b, const REAL *c, REAL *out) {
    FN(pt_tri_cp)(p[0], p[1], p[2], a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], out, out + 1, out + 2);
}

/* ------------------------------------------------------------------------------------------------
 * edf: exact unsigned distance by three sorted sweeps; follows serial/ssdf.hpp:447-619.
 * out[] is in/out (current best, min-accumulated); an empty surface leaves it untouched (462).
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    const REAL *key;
    ptrdiff_t idx;
} FN(orc_sortrec);

static int FN(orc_sortrec_cmp)(const void *pa, const void *pb) {
    const FN(orc_sortrec) *a = (const FN(orc_sortrec) *)pa, *b = (const FN(orc_sortrec) *)pb;
    const REAL ka = a->key[a->idx], kb = b->key[b->idx];
    if (ka < kb) return -1;
    if (kb < ka) return 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

static ptrdiff_t FN(orc_lower_bound)(const REAL *arr, ptrdiff_t lo, ptrdiff_t hi, const REAL v) {
    /* first index in [lo, hi) with arr[i] >= v (std::lower_bound semantics) */
    while (lo < hi) {
        const ptrdiff_t mid = lo + (hi - lo) / 2;
        if (arr[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

int FN(orc_edf)(const ptrdiff_t npoints, const REAL *x, const REAL *y, const REAL *z, const ptrdiff_t nsel,
                const int *s0, const int *s1, const int *s2, const ptrdiff_t nsp, const REAL *sx, const REAL *sy,
                const REAL *sz, REAL *out) {
    if (nsel == 0 || nsp == 0) return 0;
    REAL *smin[3], *smax[3];
    for (int d = 0; d < 3; d++) {
        smin[d] = (REAL *)malloc(sizeof(REAL) * nsel);
        smax[d] = (REAL *)malloc(sizeof(REAL) * nsel);
    }
    ptrdiff_t *sort_idx = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * nsel);
    FN(orc_sortrec) *recs = (FN(orc_sortrec) *)malloc(sizeof(FN(orc_sortrec)) * nsel);
    REAL *scratch = (REAL *)malloc(sizeof(REAL) * nsel);
    REAL *cum_max = (REAL *)malloc(sizeof(REAL) * nsel);

    /* 1) triangle AABBs (ssdf.hpp:476-492) */
#pragma omp parallel for
    for (ptrdiff_t i = 0; i < nsel; i++) {
        const int i0 = s0[i], i1 = s1[i], i2 = s2[i];
        const REAL *co[3] = {sx, sy, sz};
        for (int d = 0; d < 3; d++) {
            const REAL v0 = co[d][i0], v1 = co[d][i1], v2 = co[d][i2];
            smin[d][i] = ORC_MIN3(v0, v1, v2);
            smax[d][i] = ORC_MAX3(v0, v1, v2);
        }
    }

    const REAL *pc[3] = {x, y, z};
    for (int dim = 0; dim < 3; dim++) {
        REAL *surf_min = smin[dim], *surf_max = smax[dim];
        /* sort by min coordinate (510); qsort with an index tie-break in place of std::sort: the
         * min over the visited set does not depend on the order of equal keys */
        for (ptrdiff_t i = 0; i < nsel; i++) recs[i].key = surf_min, recs[i].idx = i;
        qsort(recs, (size_t)nsel, sizeof(recs[0]), FN(orc_sortrec_cmp));
        for (ptrdiff_t i = 0; i < nsel; i++) sort_idx[i] = recs[i].idx;
        memcpy(scratch, surf_min, sizeof(REAL) * nsel);
        for (ptrdiff_t i = 0; i < nsel; i++) surf_min[i] = scratch[sort_idx[i]];
        memcpy(scratch, surf_max, sizeof(REAL) * nsel);
        for (ptrdiff_t i = 0; i < nsel; i++) surf_max[i] = scratch[sort_idx[i]];
        /* prefix maximum of the permuted maxima (523-532) */
        REAL acc = surf_max[0];
        for (ptrdiff_t i = 0; i < nsel; i++) {
            acc = ORC_MAX(acc, surf_max[i]);
            cum_max[i] = acc;
        }

#pragma omp parallel for schedule(dynamic, 256)
        for (ptrdiff_t p = 0; p < npoints; p++) {
            const REAL px = x[p], py = y[p], pz = z[p];
            const REAL pcoord = pc[dim][p];
            REAL best_sq = out[p] * out[p]; /* re-seeded from out each pass (540) */
            const ptrdiff_t left = FN(orc_lower_bound)(surf_min, 0, nsel, pcoord);
            /* left scan (546-573): starts at left-1, or at 0 when left == 0 */
            for (ptrdiff_t i = (left > 0 ? left - 1 : 0); i >= 0; i--) {
                const REAL margin = pcoord - cum_max[i];
                if (margin >= (REAL)0 && margin * margin >= best_sq) break;
                const ptrdiff_t t = sort_idx[i];
                const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                const REAL tx0 = sx[i0], tx1 = sx[i1], tx2 = sx[i2];
                const REAL ty0 = sy[i0], ty1 = sy[i1], ty2 = sy[i2];
                const REAL tz0 = sz[i0], tz1 = sz[i1], tz2 = sz[i2];
                if (!FN(box_can_improve)(px, py, pz, best_sq, ORC_MIN3(tx0, tx1, tx2), ORC_MAX3(tx0, tx1, tx2),
                                         ORC_MIN3(ty0, ty1, ty2), ORC_MAX3(ty0, ty1, ty2), ORC_MIN3(tz0, tz1, tz2),
                                         ORC_MAX3(tz0, tz1, tz2)))
                    continue;
                const REAL d2 = FN(pt_tri_d2)(px, py, pz, tx0, ty0, tz0, tx1, ty1, tz1, tx2, ty2, tz2);
                best_sq = ORC_MIN(best_sq, d2);
            }
            /* right scan (576-602) */
            for (ptrdiff_t i = left; i < nsel; i++) {
                const REAL margin = surf_min[i] - pcoord;
                if (margin >= (REAL)0 && margin * margin >= best_sq) break;
                const ptrdiff_t t = sort_idx[i];
                const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                const REAL tx0 = sx[i0], tx1 = sx[i1], tx2 = sx[i2];
                const REAL ty0 = sy[i0], ty1 = sy[i1], ty2 = sy[i2];
                const REAL tz0 = sz[i0], tz1 = sz[i1], tz2 = sz[i2];
                if (!FN(box_can_improve)(px, py, pz, best_sq, ORC_MIN3(tx0, tx1, tx2), ORC_MAX3(tx0, tx1, tx2),
                                         ORC_MIN3(ty0, ty1, ty2), ORC_MAX3(ty0, ty1, ty2), ORC_MIN3(tz0, tz1, tz2),
                                         ORC_MAX3(tz0, tz1, tz2)))
                    continue;
                const REAL d2 = FN(pt_tri_d2)(px, py, pz, tx0, ty0, tz0, tx1, ty1, tz1, tx2, ty2, tz2);
                best_sq = ORC_MIN(best_sq, d2);
            }
            out[p] = ORC_SQRT(best_sq); /* 604 */
        }
    }

    for (int d = 0; d < 3; d++) free(smin[d]), free(smax[d]);
    free(sort_idx), free(recs), free(scratch), free(cum_max);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * edf_celllist: 2-D cell grid on the two widest axes + a sorted third axis; follows
 * serial/ssdf.hpp:642-963.  This is the OpenMP CPU baseline that bench.py times.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    const REAL *key;
    int idx;
} FN(orc_cellrec);

static int FN(orc_cellrec_cmp)(const void *pa, const void *pb) {
    const FN(orc_cellrec) *a = (const FN(orc_cellrec) *)pa, *b = (const FN(orc_cellrec) *)pb;
    const REAL ka = a->key[a->idx], kb = b->key[b->idx];
    if (ka < kb) return -1;
    if (kb < ka) return 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

int FN(orc_celllist)(const ptrdiff_t npoints, const REAL *x, const REAL *y, const REAL *z, const ptrdiff_t nsel,
                     const int *s0, const int *s1, const int *s2, const ptrdiff_t nsp, const REAL *sx, const REAL *sy,
                     const REAL *sz, REAL *out) {
    if (nsel == 0 || nsp == 0) return 0; /* 657 */
    REAL *tmin[3], *tmax[3];
    for (int d = 0; d < 3; d++) {
        tmin[d] = (REAL *)malloc(sizeof(REAL) * nsel);
        tmax[d] = (REAL *)malloc(sizeof(REAL) * nsel);
    }
    REAL gmin[3] = {ORC_REAL_MAX, ORC_REAL_MAX, ORC_REAL_MAX};
    REAL gmax[3] = {-ORC_REAL_MAX, -ORC_REAL_MAX, -ORC_REAL_MAX};
    REAL max_extent = 0;
    /* triangle boxes, global bounds and the largest triangle extent (673-723) */
    for (ptrdiff_t i = 0; i < nsel; ++i) {
        const int i0 = s0[i], i1 = s1[i], i2 = s2[i];
        const REAL *co[3] = {sx, sy, sz};
        REAL ext = 0;
        for (int d = 0; d < 3; d++) {
            const REAL v0 = co[d][i0], v1 = co[d][i1], v2 = co[d][i2];
            const REAL lo = ORC_MIN3(v0, v1, v2), hi = ORC_MAX3(v0, v1, v2);
            tmin[d][i] = lo, tmax[d][i] = hi;
            gmin[d] = ORC_MIN(gmin[d], lo);
            gmax[d] = ORC_MAX(gmax[d], hi);
            ext = ORC_MAX(ext, hi - lo);
        }
        max_extent = ORC_MAX(max_extent, ext);
    }

    /* grid axes: the two widest spans, the third is sorted (725-759) */
    const REAL span[3] = {gmax[0] - gmin[0], gmax[1] - gmin[1], gmax[2] - gmin[2]};
    int axis0, axis1, sort_axis;
    if (span[0] >= span[1] && span[0] >= span[2]) {
        axis0 = 0;
        if (span[1] >= span[2]) axis1 = 1, sort_axis = 2;
        else axis1 = 2, sort_axis = 1;
    } else if (span[1] >= span[0] && span[1] >= span[2]) {
        axis0 = 1;
        if (span[0] >= span[2]) axis1 = 0, sort_axis = 2;
        else axis1 = 2, sort_axis = 0;
    } else {
        axis0 = 2;
        if (span[0] >= span[1]) axis1 = 0, sort_axis = 1;
        else axis1 = 1, sort_axis = 0;
    }

    const REAL cell_size = ORC_MAX(max_extent, (REAL)1e-8); /* 761 */
    const REAL amin0 = gmin[axis0], amin1 = gmin[axis1];
    int nc0 = (int)ORC_CEIL(span[axis0] / cell_size);
    int nc1 = (int)ORC_CEIL(span[axis1] / cell_size);
    if (nc0 < 1) nc0 = 1;
    if (nc1 < 1) nc1 = 1;
    const ptrdiff_t ncells = (ptrdiff_t)nc0 * (ptrdiff_t)nc1;

    int *cell_counts = (int *)calloc((size_t)ncells, sizeof(int));
    REAL *cmin[3], *cmax[3];
    for (int d = 0; d < 3; d++) {
        cmin[d] = (REAL *)malloc(sizeof(REAL) * ncells);
        cmax[d] = (REAL *)malloc(sizeof(REAL) * ncells);
        for (ptrdiff_t c = 0; c < ncells; c++) cmin[d][c] = ORC_REAL_MAX, cmax[d][c] = -ORC_REAL_MAX;
    }

    /* cell of a pair of axis coordinates (784-790): truncation toward zero, then clamped */
#define ORC_CELL_ID(a0, a1, out_cid)                                 \
    do {                                                             \
        int ix_ = (int)(((a0) - amin0) / cell_size);                 \
        int iy_ = (int)(((a1) - amin1) / cell_size);                 \
        ix_ = ix_ < nc0 - 1 ? ix_ : nc0 - 1;                         \
        ix_ = ix_ > 0 ? ix_ : 0;                                     \
        iy_ = iy_ < nc1 - 1 ? iy_ : nc1 - 1;                         \
        iy_ = iy_ > 0 ? iy_ : 0;                                     \
        (out_cid) = (ptrdiff_t)iy_ * nc0 + ix_;                      \
    } while (0)

    /* count and aggregate cell boxes by triangle box centre (793-802) */
    ptrdiff_t *tri_cell = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * nsel);
    for (ptrdiff_t i = 0; i < nsel; ++i) {
        ptrdiff_t cid;
        ORC_CELL_ID((tmin[axis0][i] + tmax[axis0][i]) / 2, (tmin[axis1][i] + tmax[axis1][i]) / 2, cid);
        tri_cell[i] = cid;
        cell_counts[cid] += 1;
        for (int d = 0; d < 3; d++) {
            cmin[d][cid] = ORC_MIN(cmin[d][cid], tmin[d][i]);
            cmax[d][cid] = ORC_MAX(cmax[d][cid], tmax[d][i]);
        }
    }
    ptrdiff_t *cell_ptr = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * (ncells + 1));
    cell_ptr[0] = 0;
    for (ptrdiff_t c = 0; c < ncells; ++c) cell_ptr[c + 1] = cell_ptr[c] + cell_counts[c];
    int *cell_idx = (int *)malloc(sizeof(int) * nsel);
    ptrdiff_t *fill = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * ncells);
    memcpy(fill, cell_ptr, sizeof(ptrdiff_t) * ncells);
    for (ptrdiff_t i = 0; i < nsel; ++i) cell_idx[fill[tri_cell[i]]++] = (int)i;

    /* per-cell sort on the sort axis (819-828) */
    const REAL *sorted_min_orig = tmin[sort_axis];
#pragma omp parallel
    {
        FN(orc_cellrec) *tmp = NULL;
        ptrdiff_t cap = 0;
#pragma omp for /* the reference's default (static) schedule, ssdf.hpp:819 */
        for (ptrdiff_t c = 0; c < ncells; ++c) {
            const ptrdiff_t b = cell_ptr[c], e = cell_ptr[c + 1];
            if (e - b < 2) continue;
            if (e - b > cap) {
                cap = e - b;
                tmp = (FN(orc_cellrec) *)realloc(tmp, sizeof(*tmp) * cap);
            }
            for (ptrdiff_t i = b; i < e; i++) tmp[i - b].key = sorted_min_orig, tmp[i - b].idx = cell_idx[i];
            qsort(tmp, (size_t)(e - b), sizeof(*tmp), FN(orc_cellrec_cmp));
            for (ptrdiff_t i = b; i < e; i++) cell_idx[i] = tmp[i - b].idx;
        }
        free(tmp);
    }

    /* permute the boxes into cell order (830-840) */
    {
        REAL *temp = (REAL *)malloc(sizeof(REAL) * nsel);
        REAL *boxes[6] = {tmin[0], tmin[1], tmin[2], tmax[0], tmax[1], tmax[2]};
        for (int d = 0; d < 6; d++) {
            memcpy(temp, boxes[d], sizeof(REAL) * nsel);
            for (ptrdiff_t i = 0; i < nsel; i++) boxes[d][i] = temp[cell_idx[i]];
        }
        free(temp);
    }
    const REAL *sorted_min = tmin[sort_axis], *sorted_max = tmax[sort_axis];

    /* per-cell cumulative maxima (843-854) */
    REAL *cum_max = (REAL *)calloc((size_t)nsel, sizeof(REAL));
    for (ptrdiff_t c = 0; c < ncells; ++c) {
        const ptrdiff_t b = cell_ptr[c], e = cell_ptr[c + 1];
        if (b == e) continue;
        REAL acc = sorted_max[b];
        for (ptrdiff_t i = b; i < e; ++i) {
            acc = ORC_MAX(acc, sorted_max[i]);
            cum_max[i] = acc;
        }
    }

    const REAL *xyz[3] = {x, y, z};
    const REAL *tminx = tmin[0], *tminy = tmin[1], *tminz = tmin[2];
    const REAL *tmaxx = tmax[0], *tmaxy = tmax[1], *tmaxz = tmax[2];

    /* queries (912-959): own cell first, then every other non-empty cell that can improve; the
     * reference's default (static) schedule of its `omp parallel for` (912) */
#pragma omp parallel for
    for (ptrdiff_t p = 0; p < npoints; ++p) {
        const REAL px = x[p], py = y[p], pz = z[p];
        const REAL pcoord = xyz[sort_axis][p];
        REAL best_sq = out[p] * out[p];
        ptrdiff_t first_cid;
        ORC_CELL_ID((xyz[axis0][p] + xyz[axis0][p]) / 2, (xyz[axis1][p] + xyz[axis1][p]) / 2, first_cid);
        for (ptrdiff_t it = -1; it < ncells; ++it) {
            const ptrdiff_t cid = it < 0 ? first_cid : it;
            if (it >= 0 && cid == first_cid) continue;
            if (cell_counts[cid] == 0 ||
                !FN(box_can_improve)(px, py, pz, best_sq, cmin[0][cid], cmax[0][cid], cmin[1][cid], cmax[1][cid],
                                     cmin[2][cid], cmax[2][cid]))
                continue;
            /* process_cell (859-898) */
            const ptrdiff_t b = cell_ptr[cid], e = cell_ptr[cid + 1];
            const ptrdiff_t left = FN(orc_lower_bound)(sorted_min, b, e, pcoord);
            for (ptrdiff_t i = (left > b) ? left - 1 : b; i >= b; --i) {
                const REAL margin = pcoord - cum_max[i];
                if (margin >= (REAL)0 && margin * margin >= best_sq) break;
                if (FN(box_can_improve)(px, py, pz, best_sq, tminx[i], tmaxx[i], tminy[i], tmaxy[i], tminz[i],
                                        tmaxz[i])) {
                    const int t = cell_idx[i];
                    const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                    const REAL d2 = FN(pt_tri_d2)(px, py, pz, sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1],
                                                  sx[i2], sy[i2], sz[i2]);
                    if (d2 < best_sq) best_sq = d2;
                }
            }
            for (ptrdiff_t i = left; i < e; ++i) {
                const REAL margin = sorted_min[i] - pcoord;
                if (margin >= (REAL)0 && margin * margin >= best_sq) break;
                if (!FN(box_can_improve)(px, py, pz, best_sq, tminx[i], tmaxx[i], tminy[i], tmaxy[i], tminz[i],
                                         tmaxz[i]))
                    continue;
                const int t = cell_idx[i];
                const int i0 = s0[t], i1 = s1[t], i2 = s2[t];
                const REAL d2 = FN(pt_tri_d2)(px, py, pz, sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1], sx[i2],
                                              sy[i2], sz[i2]);
                if (d2 < best_sq) best_sq = d2;
            }
        }
        out[p] = ORC_SQRT(best_sq);
    }
#undef ORC_CELL_ID

    for (int d = 0; d < 3; d++) free(tmin[d]), free(tmax[d]), free(cmin[d]), free(cmax[d]);
    free(cell_counts), free(tri_cell), free(cell_ptr), free(cell_idx), free(fill), free(cum_max);
    return 0;
}

#undef FN
#undef ORC_CAT
#undef ORC_CAT2

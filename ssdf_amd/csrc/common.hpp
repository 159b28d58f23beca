// common.hpp — shared host/device definitions of the MI355X distance-field engine.
//
// Layout in HBM (all per mesh, built once and kept resident):
//   Node  (64 B, one per internal LBVH node): both children's AABBs + child references, read by a
//         whole wave with one scalar 64-byte load during wave-coherent traversal.
//   TriP  (144 B, one per triangle, in Morton/leaf order): FP32 vertices a, b, c + original index
//         (the 48-byte TriF header) + the precomputed constants of the FP32 distance routine.
// Child references: r >= 0 is internal node r, r < 0 is leaf ~r (a position in the TriF array).
#pragma once

#include <hip/hip_runtime.h>

#include "ssdf_c.h"

#include <cstddef>
#include <cstdint>

#pragma clang fp contract(off)

namespace ssdf_amd {

// Internal LBVH node: both children's FP32 boxes, the child references and the node's range of
// leaves [first, last] in Morton order.  A child reference is >= 0 for an internal node, and
// ~(k | (n - 1) << 29) for the run of n consecutive leaves k .. k + n - 1 (n = 1 a single leaf): a
// subtree of at most kMaxRun leaves is not visited, its parent's box test admits all its triangles
// (leaf positions are below 2^29, the build's triangle limit).
struct alignas(16) Node {
    float lmin[3], lmax[3];
    float rmin[3], rmax[3];
    int left, right;
    int first, last;
};
static_assert(sizeof(Node) == 64, "Node must be 64 bytes");

constexpr int kLeafBits = 29;
// Measured (FP32 + fixup; 200 k torus 256^3 query / 64-mesh batch step / 512^3 beams): runs of at most 2
// leaves 6.95 / 79.4 / 109.1 ms, 3 6.97 / 78.5 / 103.7 ms, 4 7.18 / 78.2 / 100.6 ms.
constexpr int kMaxRun = 3;
static_assert(kMaxRun >= 1 && kMaxRun <= 4, "the run length takes 2 bits");

// Leaf reference ref (< 0): its first leaf position and its number of leaves.
__host__ __device__ __forceinline__ int leaf_first(const int ref) { return ~ref & ((1 << kLeafBits) - 1); }
__host__ __device__ __forceinline__ int leaf_count(const int ref) { return 1 + ((~ref >> kLeafBits) & 3); }
__host__ __device__ __forceinline__ int leaf_run(const int first, const int n) {
    return ~(first | ((n - 1) << kLeafBits));
}

// Far-field data of an internal node's two children for winding numbers: the child subtree's
// summed area vector sum(a_i n_i), its area-weighted centroid and the squared radius of its box
// around that centroid; plus the child references (as in Node).
struct alignas(16) GwnNode {
    float ln[3], lc[3], lr2;
    float rn[3], rc[3], rr2;
    int left, right;
};
static_assert(sizeof(GwnNode) == 64, "GwnNode must be 64 bytes");

struct alignas(16) TriF {
    float a[3], b[3], c[3];
    int orig;
    int flags;  // bit 0: degenerate (no usable normal)
    int pad;
};
static_assert(sizeof(TriF) == 48, "TriF must be 48 bytes");

// Triangle record of the FP32 traversal: the TriF header plus the triangle in a plane frame of its
// own, so that the branch-free distance (tri_d2_fast) works in two dimensions and a wave reading
// the record with scalar loads spends no VALU on setting it up.  The frame's origin o is the first
// corner of the longest edge, U points along that edge, V lies in the plane towards the third
// corner, and N = U x V is the unit normal (the orientation of (b - a) x (c - a)).  In (U, V)
// coordinates the corners are (0, 0), (bx, 0) and (cx, cy), cy >= 0.  144 bytes, in leaf (Morton)
// order.
struct alignas(16) TriP {
    TriF f;
    float o[3];
    float nh[3];               // N
    float ux[3], vx[3];        // U, V
    float bx, ibx;             // second corner (bx, 0); 1 / bx (0 when bx = 0)
    float e1x, e1y, f1x, f1y;  // edge e1 from (bx, 0) to (cx, cy); e1 / |e1|^2 (0 when e1 = 0)
    float cx, cy, g2x, g2y;    // third corner c; c / |c|^2 (0 when c = 0)
    float in0;                 // inside threshold of the side tests: 0, or +inf on a degenerate triangle
    float pad;
};
static_assert(sizeof(TriP) == 144, "TriP must be 144 bytes");

// Error codes: SSDF_OK / SSDF_ERR_* from the public header.

#define SSDF_INF __builtin_inff()

// ---------------------------------------------------------------------------------------------
// Geometry, FP32 and FP64.  Operation order follows serial/ssdf.hpp:50-121 (distance) and
// 143-239 (closest point) exactly and contraction is off, so the FP64 instantiation rounds like
// the oracle's restatement bit for bit.
// ---------------------------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T pt_tri_d2(const T px, const T py, const T pz, const T ax, const T ay,
                                                const T az, const T bx, const T by, const T bz, const T cx,
                                                const T cy, const T cz) {
    const T abx = bx - ax, aby = by - ay, abz = bz - az;
    const T acx = cx - ax, acy = cy - ay, acz = cz - az;
    const T apx = px - ax, apy = py - ay, apz = pz - az;
    const T d1 = abx * apx + aby * apy + abz * apz;
    const T d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= T(0) && d2 <= T(0)) return apx * apx + apy * apy + apz * apz;
    const T bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const T d3 = abx * bpx + aby * bpy + abz * bpz;
    const T d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= T(0) && d4 <= d3) return bpx * bpx + bpy * bpy + bpz * bpz;
    const T vc = d1 * d4 - d3 * d2;
    if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
        const T v = d1 / (d1 - d3);
        const T qx = ax + v * abx, qy = ay + v * aby, qz = az + v * abz;
        const T dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }
    const T cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const T d5 = abx * cpx + aby * cpy + abz * cpz;
    const T d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= T(0) && d5 <= d6) return cpx * cpx + cpy * cpy + cpz * cpz;
    const T vb = d5 * d2 - d1 * d6;
    if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
        const T w = d2 / (d2 - d6);
        const T qx = ax + w * acx, qy = ay + w * acy, qz = az + w * acz;
        const T dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }
    const T va = d3 * d6 - d5 * d4;
    if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
        const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const T qx = bx + w * (cx - bx), qy = by + w * (cy - by), qz = bz + w * (cz - bz);
        const T dx = px - qx, dy = py - qy, dz = pz - qz;
        return dx * dx + dy * dy + dz * dz;
    }
    const T denom = T(1) / (va + vb + vc);
    const T v = vb * denom;
    const T w = vc * denom;
    const T qx = ax + abx * v + acx * w;
    const T qy = ay + aby * v + acy * w;
    const T qz = az + abz * v + acz * w;
    const T dx = px - qx, dy = py - qy, dz = pz - qz;
    return dx * dx + dy * dy + dz * dz;
}

template <typename T>
__host__ __device__ __forceinline__ void pt_tri_cp(const T px, const T py, const T pz, const T ax, const T ay,
                                                   const T az, const T bx, const T by, const T bz, const T cx,
                                                   const T cy, const T cz, T &ox, T &oy, T &oz) {
    const T abx = bx - ax, aby = by - ay, abz = bz - az;
    const T acx = cx - ax, acy = cy - ay, acz = cz - az;
    const T apx = px - ax, apy = py - ay, apz = pz - az;
    const T d1 = abx * apx + aby * apy + abz * apz;
    const T d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= T(0) && d2 <= T(0)) {
        ox = ax, oy = ay, oz = az;
        return;
    }
    const T bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const T d3 = abx * bpx + aby * bpy + abz * bpz;
    const T d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= T(0) && d4 <= d3) {
        ox = bx, oy = by, oz = bz;
        return;
    }
    const T vc = d1 * d4 - d3 * d2;
    if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
        const T v = d1 / (d1 - d3);
        ox = ax + v * abx, oy = ay + v * aby, oz = az + v * abz;
        return;
    }
    const T cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const T d5 = abx * cpx + aby * cpy + abz * cpz;
    const T d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= T(0) && d5 <= d6) {
        ox = cx, oy = cy, oz = cz;
        return;
    }
    const T vb = d5 * d2 - d1 * d6;
    if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
        const T w = d2 / (d2 - d6);
        ox = ax + w * acx, oy = ay + w * acy, oz = az + w * acz;
        return;
    }
    const T va = d3 * d6 - d5 * d4;
    if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
        const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const T bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
        ox = bx + w * bcx, oy = by + w * bcy, oz = bz + w * bcz;
        return;
    }
    const T denom = T(1) / (va + vb + vc);
    const T v = vb * denom;
    const T w = vc * denom;
    ox = ax + abx * v + acx * w;
    oy = ay + aby * v + acy * w;
    oz = az + abz * v + acz * w;
}

// Squared point-AABB distance (serial/edf/edf.hpp:55-69).  max(lo - p, p - hi, 0) per axis picks
// the same operand the reference's nested selects pick, so the value is identical.
template <typename T>
__host__ __device__ __forceinline__ T pt_box_d2(const T px, const T py, const T pz, const T mnx, const T mny,
                                                const T mnz, const T mxx, const T mxy, const T mxz) {
    const T dx = (px < mnx) ? (mnx - px) : (px > mxx ? px - mxx : T(0));
    const T dy = (py < mny) ? (mny - py) : (py > mxy ? py - mxy : T(0));
    const T dz = (pz < mnz) ? (mnz - pz) : (pz > mxz ? pz - mxz : T(0));
    return dx * dx + dy * dy + dz * dz;
}

// FP32 point-AABB squared distance for pruning: per axis max(lo - p, p - hi, 0) (one v_max3),
// sum with fused multiply-adds.  Same value as pt_box_d2 up to the rounding of the final sum.
// (p - med3(p, lo, hi) would be one instruction shorter, but a VOP3 reads one SGPR only, and both
// bounds arrive in SGPRs from the scalar node load: the compiler adds a v_mov per axis.)
__host__ __device__ __forceinline__ float box_d2_fast(const float px, const float py, const float pz, const float *mn,
                                             const float *mx) {
    const float dx = fmaxf(fmaxf(mn[0] - px, px - mx[0]), 0.0f);
    const float dy = fmaxf(fmaxf(mn[1] - py, py - mx[1]), 0.0f);
    const float dz = fmaxf(fmaxf(mn[2] - pz, pz - mx[2]), 0.0f);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

// Branch-free FP32 squared point-triangle distance in the triangle's plane frame: h^2 plus the
// squared 2D distance from the projection (u, v) of p to the triangle, which is 0 when (u, v) falls
// inside (all three side tests non-negative) and otherwise the nearest of the three edge segments.
// Same quantity as the Voronoi-region routine pt_tri_d2 (serial/ssdf.hpp:50-121) to a few ulps of
// the coordinates, with no per-lane branches or divisions (every lane of a wave runs the same ~40
// instructions); the exact FP64 routine re-decides near-ties.  Degenerate triangles have an
// infinite inside threshold, so they take the segment branch.
//
// It is split in two so that a wave can stop after the first part: tri_plane gives p - o and the
// signed plane distance h = N.(p - o), and |h| bounds the distance from below (plane_bound below);
// tri_d2_rest finishes from there.
struct TriPlane {
    float pa0, pa1, pa2, h;
};
__host__ __device__ __forceinline__ TriPlane tri_plane(const float px, const float py, const float pz, const TriP &t) {
    TriPlane r;
    r.pa0 = px - t.o[0], r.pa1 = py - t.o[1], r.pa2 = pz - t.o[2];
    r.h = __builtin_fmaf(t.nh[0], r.pa0, __builtin_fmaf(t.nh[1], r.pa1, t.nh[2] * r.pa2));
    return r;
}

__host__ __device__ __forceinline__ float clamp01(const float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__host__ __device__ __forceinline__ float tri_d2_rest(const float px, const float py, const float pz, const TriP &t,
                                                      const TriPlane &pl) {
    const float pa0 = pl.pa0, pa1 = pl.pa1, pa2 = pl.pa2, h = pl.h;
    const float u = __builtin_fmaf(t.ux[0], pa0, __builtin_fmaf(t.ux[1], pa1, t.ux[2] * pa2));
    const float v = __builtin_fmaf(t.vx[0], pa0, __builtin_fmaf(t.vx[1], pa1, t.vx[2] * pa2));
    // segments s(t) = start + t e, t = clamp((p - start).e / |e|^2, 0, 1); r = p - s(t)
    // from (0, 0) to (bx, 0)
    const float t0 = clamp01(u * t.ibx);
    const float r0 = __builtin_fmaf(-t0, t.bx, u);
    const float q0 = __builtin_fmaf(r0, r0, v * v);
    // from (bx, 0) to (cx, cy); side test cross(e1, p - b)
    const float d1 = u - t.bx;
    const float s1 = __builtin_fmaf(t.e1x, v, -(t.e1y * d1));
    const float t1 = clamp01(__builtin_fmaf(t.f1x, d1, t.f1y * v));
    const float r1x = __builtin_fmaf(-t1, t.e1x, d1), r1y = __builtin_fmaf(-t1, t.e1y, v);
    const float q1 = __builtin_fmaf(r1x, r1x, r1y * r1y);
    // from (0, 0) to (cx, cy); side test cross(p, c) <= 0
    const float s2 = __builtin_fmaf(t.cy, u, -(t.cx * v));
    const float t2 = clamp01(__builtin_fmaf(t.g2x, u, t.g2y * v));
    const float r2x = __builtin_fmaf(-t2, t.cx, u), r2y = __builtin_fmaf(-t2, t.cy, v);
    const float q2 = __builtin_fmaf(r2x, r2x, r2y * r2y);
    // side test of the first edge: v >= 0
    const bool inside = fminf(v, fminf(s1, s2)) >= t.in0;
    return __builtin_fmaf(h, h, inside ? 0.0f : fminf(q0, fminf(q1, q2)));
}

__host__ __device__ __forceinline__ float tri_d2_fast(const float px, const float py, const float pz, const TriP &t) {
    return tri_d2_rest(px, py, pz, t, tri_plane(px, py, pz, t));
}

// Plane bound of a pruning bound d2b (a squared distance, strict or inclusive): every triangle whose
// tri_d2_fast value is at most d2b has |h| <= plane_bound(r, babs), where r is sqrt(d2b) to 1 ulp
// or above and babs >= 0.  Why: tri_d2_fast rounds h^2 + q with q >= 0 once, so
// |h| <= sqrt(d_fast) (1 + 2^-24) <= r (1 + 2^-24), and the factor on r covers that and its own
// rounding; the 2 babs (babs = 2^-18 max(L, |p|) in the kernels) is margin on top.  r = NaN or
// -inf (no query, or one handed over) gives a bound that no |h| passes.
__host__ __device__ __forceinline__ float plane_bound(const float r, const float babs) {
    return __builtin_fmaf(r, 1.0f + 0x1p-22f, 2.0f * babs);
}

// Fills a TriP from FP32 vertices: the frame and the 2D corners are evaluated in FP64 from the FP32
// values and rounded once.  The frame is Gram-Schmidt on the longest edge: V is the part of the
// third corner's offset perpendicular to U, so the three corners lie in the (U, V) plane to FP64
// rounding even on slivers, whose cross product cancels.  A sliver whose smallest height is at most
// 1e-6 of its longest edge (|n| <= 1e-6 |e|^2, n = 2 x area) is flagged degenerate: its infinite
// inside threshold sends every query to the segment branch of tri_d2_fast, whose error is then at
// most that height (a projection inside the triangle lies within it of an edge).  Thicker triangles
// meet the FP32 error bound of the fixup band (2^-18 of the coordinate scale); thinner unflagged
// ones do not: their FP32 side tests cannot resolve the strip they cover.
__host__ __device__ inline void make_trip(TriP &tr, const float *va, const float *vb, const float *vc, int orig) {
    double v[3][3];
    for (int d = 0; d < 3; d++) {
        tr.f.a[d] = va[d], tr.f.b[d] = vb[d], tr.f.c[d] = vc[d];
        v[0][d] = va[d], v[1][d] = vb[d], v[2][d] = vc[d];
    }
    double e[3][3], n[3], l2[3];
    for (int q = 0; q < 3; q++) {
        for (int d = 0; d < 3; d++) e[q][d] = v[(q + 1) % 3][d] - v[q][d];
        l2[q] = e[q][0] * e[q][0] + e[q][1] * e[q][1] + e[q][2] * e[q][2];
    }
    // n = (b - a) x (c - a) = e0 x (-e2)
    n[0] = e[0][2] * e[2][1] - e[0][1] * e[2][2];
    n[1] = e[0][0] * e[2][2] - e[0][2] * e[2][0];
    n[2] = e[0][1] * e[2][0] - e[0][0] * e[2][1];
    const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    int o = 0;  // the longest edge runs from corner o to corner o + 1 (a cyclic relabelling keeps n)
    for (int q = 1; q < 3; q++)
        if (l2[q] > l2[o]) o = q;
    const double emax = l2[o];
    const bool degenerate = !(emax > 0) || !(nn > 1e-12 * emax * emax);
    const double *p0 = v[o], *p2 = v[(o + 2) % 3];
    const double el = sqrt(emax);
    double U[3] = {1.0, 0.0, 0.0}, V[3], w[3];
    if (el > 0)
        for (int d = 0; d < 3; d++) U[d] = e[o][d] / el;
    for (int d = 0; d < 3; d++) w[d] = p2[d] - p0[d];
    const double wu = w[0] * U[0] + w[1] * U[1] + w[2] * U[2];
    for (int d = 0; d < 3; d++) V[d] = w[d] - wu * U[d];
    double vl = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    if (!(vl > 0)) {
        // collinear corners: any unit vector perpendicular to U (U x the axis U is least along)
        const int k = (fabs(U[0]) <= fabs(U[1]) && fabs(U[0]) <= fabs(U[2])) ? 0 : (fabs(U[1]) <= fabs(U[2]) ? 1 : 2);
        const double ax[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        V[0] = U[1] * ax[2] - U[2] * ax[1], V[1] = U[2] * ax[0] - U[0] * ax[2], V[2] = U[0] * ax[1] - U[1] * ax[0];
        vl = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    }
    for (int d = 0; d < 3; d++) V[d] /= vl;
    const double N[3] = {U[1] * V[2] - U[2] * V[1], U[2] * V[0] - U[0] * V[2], U[0] * V[1] - U[1] * V[0]};
    const double cx = wu, cy = w[0] * V[0] + w[1] * V[1] + w[2] * V[2];
    const double e1x = cx - el, e1y = cy, l1 = e1x * e1x + e1y * e1y, lc = cx * cx + cy * cy;
    for (int d = 0; d < 3; d++) {
        tr.o[d] = (float)p0[d];
        tr.nh[d] = (float)N[d];
        tr.ux[d] = (float)U[d];
        tr.vx[d] = (float)V[d];
    }
    tr.bx = (float)el;
    tr.ibx = el > 0 ? (float)(1.0 / el) : 0.0f;
    tr.e1x = (float)e1x, tr.e1y = (float)e1y;
    tr.f1x = l1 > 0 ? (float)(e1x / l1) : 0.0f, tr.f1y = l1 > 0 ? (float)(e1y / l1) : 0.0f;
    tr.cx = (float)cx, tr.cy = (float)cy;
    tr.g2x = lc > 0 ? (float)(cx / lc) : 0.0f, tr.g2y = lc > 0 ? (float)(cy / lc) : 0.0f;
    tr.in0 = degenerate ? SSDF_INF : 0.0f;
    tr.pad = 0.0f;
    tr.f.orig = orig;
    tr.f.flags = degenerate ? 1 : 0;
    tr.f.pad = 0;
}

// Monotone float <-> int mapping for atomicMin / atomicMax on floats.
__device__ __forceinline__ int float_to_ordered(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ __forceinline__ float ordered_to_float(int i) {
    const int b = i >= 0 ? i : i ^ 0x7fffffff;
    union {
        int i;
        float f;
    } u;
    u.i = b;
    return u.f;
}

// Blocks of one launch are handed to the 8 XCDs round-robin (block b -> XCD b mod 8), each XCD
// taking its blocks in order.  This bijection gives XCD x the logical chunks x, x + 8, x + 16, ...
// of kXcdChunk consecutive tiles, so its private L2 sees one compact region at a time while the 8
// XCDs sweep the grid side by side (one contiguous eighth per XCD left the XCDs with cheap
// eighths idle for the last third of the launch).  Blocks past the last whole stripe of 8 chunks
// keep their own index.
constexpr unsigned kXcdChunk = 128;
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
    const unsigned span = 8u * kXcdChunk;
    if (b >= nblocks / span * span) return b;
    const unsigned x = b & 7u, k = b >> 3;
    return ((k / kXcdChunk) * 8u + x) * kXcdChunk + k % kXcdChunk;
}

}  // namespace ssdf_amd

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

// Triangle record of the FP32 traversal: the TriF header plus per-triangle constants of the
// branch-free distance (tri_d2_fast), so that a wave reading it with scalar loads spends no VALU on
// them: edges e0 = b-a, e1 = c-b, e2 = a-c; in-plane inward edge normals m_i = n x e_i; inverse
// squared edge lengths; unit normal.  144 bytes, in leaf (Morton) order.
struct alignas(16) TriP {
    TriF f;
    float e[9];
    float m[9];
    float inv[3];
    float nh[3];
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

// Branch-free FP32 squared point-triangle distance: the plane distance when the projection of p
// falls inside (all three in-plane side tests non-negative), else the nearest of the three edge
// segments.  Same quantity as the Voronoi-region routine pt_tri_d2 (serial/ssdf.hpp:50-121) to a
// few ulps of the coordinates, with no per-lane branches or divisions (every lane of a wave runs the
// same ~65 instructions); the exact FP64 routine re-decides near-ties.  Degenerate triangles carry
// NaN side normals, so they take the segment branch.
//
// It is split in two so that a wave can stop after the first part: tri_plane gives p - a and the
// signed plane distance h = n.(p - a) (unit normal n; 0 on triangles flagged degenerate), and |h|
// bounds the distance from below (plane_bound below); tri_d2_rest finishes from there.
struct TriPlane {
    float pa0, pa1, pa2, h;
};
__host__ __device__ __forceinline__ TriPlane tri_plane(const float px, const float py, const float pz, const TriP &t) {
    TriPlane r;
    r.pa0 = px - t.f.a[0], r.pa1 = py - t.f.a[1], r.pa2 = pz - t.f.a[2];
    r.h = __builtin_fmaf(t.nh[0], r.pa0, __builtin_fmaf(t.nh[1], r.pa1, t.nh[2] * r.pa2));
    return r;
}

__host__ __device__ __forceinline__ float tri_d2_rest(const float px, const float py, const float pz, const TriP &t,
                                                      const TriPlane &pl) {
    const float pa0 = pl.pa0, pa1 = pl.pa1, pa2 = pl.pa2, h = pl.h;
    const float pb0 = px - t.f.b[0], pb1 = py - t.f.b[1], pb2 = pz - t.f.b[2];
    const float pc0 = px - t.f.c[0], pc1 = py - t.f.c[1], pc2 = pz - t.f.c[2];
    const float s0 = __builtin_fmaf(t.m[0], pa0, __builtin_fmaf(t.m[1], pa1, t.m[2] * pa2));
    const float s1 = __builtin_fmaf(t.m[3], pb0, __builtin_fmaf(t.m[4], pb1, t.m[5] * pb2));
    const float s2 = __builtin_fmaf(t.m[6], pc0, __builtin_fmaf(t.m[7], pc1, t.m[8] * pc2));
    // edge segments: t = clamp(e.p / |e|^2, 0, 1), r = p - t e
    float u0 = __builtin_fmaf(t.e[0], pa0, __builtin_fmaf(t.e[1], pa1, t.e[2] * pa2)) * t.inv[0];
    float u1 = __builtin_fmaf(t.e[3], pb0, __builtin_fmaf(t.e[4], pb1, t.e[5] * pb2)) * t.inv[1];
    float u2 = __builtin_fmaf(t.e[6], pc0, __builtin_fmaf(t.e[7], pc1, t.e[8] * pc2)) * t.inv[2];
    u0 = fminf(fmaxf(u0, 0.0f), 1.0f);
    u1 = fminf(fmaxf(u1, 0.0f), 1.0f);
    u2 = fminf(fmaxf(u2, 0.0f), 1.0f);
    const float r00 = __builtin_fmaf(-u0, t.e[0], pa0), r01 = __builtin_fmaf(-u0, t.e[1], pa1),
                r02 = __builtin_fmaf(-u0, t.e[2], pa2);
    const float r10 = __builtin_fmaf(-u1, t.e[3], pb0), r11 = __builtin_fmaf(-u1, t.e[4], pb1),
                r12 = __builtin_fmaf(-u1, t.e[5], pb2);
    const float r20 = __builtin_fmaf(-u2, t.e[6], pc0), r21 = __builtin_fmaf(-u2, t.e[7], pc1),
                r22 = __builtin_fmaf(-u2, t.e[8], pc2);
    const float q0 = __builtin_fmaf(r00, r00, __builtin_fmaf(r01, r01, r02 * r02));
    const float q1 = __builtin_fmaf(r10, r10, __builtin_fmaf(r11, r11, r12 * r12));
    const float q2 = __builtin_fmaf(r20, r20, __builtin_fmaf(r21, r21, r22 * r22));
    const bool inside = (s0 >= 0.0f) & (s1 >= 0.0f) & (s2 >= 0.0f);
    // bit select, so that both branches stay straight-line code for the whole wave
    const int msk = -(int)inside;
    const int dp = __builtin_bit_cast(int, h * h), ds = __builtin_bit_cast(int, fminf(q0, fminf(q1, q2)));
    return __builtin_bit_cast(float, (dp & msk) | (ds & ~msk));
}

__host__ __device__ __forceinline__ float tri_d2_fast(const float px, const float py, const float pz, const TriP &t) {
    return tri_d2_rest(px, py, pz, t, tri_plane(px, py, pz, t));
}

// Plane bound of a pruning bound d2b (a squared distance, strict or inclusive): every triangle whose
// tri_d2_fast value is at most d2b has |h| <= plane_bound(r, babs), where r is sqrt(d2b) to 1 ulp
// or above and babs >= 2^-18 max(L, |p|) (L: the largest coordinate magnitude of the mesh).  Why:
// sqrt(d_fast) is within 2^-19 max of the exact distance d, |h_exact| <= d, and the FP32 h is
// within 18 * 2^-24 max of h_exact (the roundings of p - a and of the unit normal, three in the
// FMA chain, over |p - a| <= 2 sqrt(3) max); 2 babs covers those terms, and the factor on r its
// rounding.  r = NaN or -inf (no query, or one handed over) gives a bound that no |h| passes.
__host__ __device__ __forceinline__ float plane_bound(const float r, const float babs) {
    return __builtin_fmaf(r, 1.0f + 0x1p-22f, 2.0f * babs);
}

// Fills a TriP from FP32 vertices: the constants are evaluated in FP64 from the FP32 values and
// rounded once.  A sliver whose smallest height is at most 1e-6 of its longest edge
// (|n| <= 1e-6 |e|^2, n = 2 x area) is flagged degenerate: NaN side normals send every query to the
// segment branch of tri_d2_fast, whose error is then at most that height (a projection inside the
// triangle lies within it of an edge).  Thicker triangles meet the FP32 error bound of the fixup
// band (2^-18 of the coordinate scale); thinner unflagged ones do not: their FP32 side tests cannot
// resolve the strip they cover, and points along their axis were taken as inside.
__host__ __device__ inline void make_trip(TriP &tr, const float *va, const float *vb, const float *vc, int orig) {
    double v[3][3];
    for (int d = 0; d < 3; d++) {
        tr.f.a[d] = va[d], tr.f.b[d] = vb[d], tr.f.c[d] = vc[d];
        v[0][d] = va[d], v[1][d] = vb[d], v[2][d] = vc[d];
    }
    double e[3][3], n[3];
    for (int q = 0; q < 3; q++)
        for (int d = 0; d < 3; d++) e[q][d] = v[(q + 1) % 3][d] - v[q][d];
    // n = (b - a) x (c - a) = e0 x (-e2)
    n[0] = e[0][2] * e[2][1] - e[0][1] * e[2][2];
    n[1] = e[0][0] * e[2][2] - e[0][2] * e[2][0];
    n[2] = e[0][1] * e[2][0] - e[0][0] * e[2][1];
    const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double emax = 0;
    for (int q = 0; q < 3; q++) {
        const double l2 = e[q][0] * e[q][0] + e[q][1] * e[q][1] + e[q][2] * e[q][2];
        emax = l2 > emax ? l2 : emax;
    }
    const bool degenerate = !(emax > 0) || !(nn > 1e-12 * emax * emax);
    const double inv_nl = degenerate ? 0.0 : 1.0 / sqrt(nn);
    for (int q = 0; q < 3; q++) {
        const double l2 = e[q][0] * e[q][0] + e[q][1] * e[q][1] + e[q][2] * e[q][2];
        tr.inv[q] = l2 > 0 ? (float)(1.0 / l2) : 0.0f;
        for (int d = 0; d < 3; d++) tr.e[3 * q + d] = (float)e[q][d];
        const double m[3] = {n[1] * e[q][2] - n[2] * e[q][1], n[2] * e[q][0] - n[0] * e[q][2],
                             n[0] * e[q][1] - n[1] * e[q][0]};
        for (int d = 0; d < 3; d++) tr.m[3 * q + d] = degenerate ? __builtin_nanf("") : (float)(m[d] * inv_nl);
    }
    for (int d = 0; d < 3; d++) tr.nh[d] = (float)(n[d] * inv_nl);
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

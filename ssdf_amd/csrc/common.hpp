// common.hpp — shared host/device definitions of the MI355X distance-field engine.
//
// Layout in HBM (all per mesh, built once and kept resident):
//   Node  (64 B, one per internal LBVH node): both children's AABBs + child references, read by a
//         whole wave with one scalar 64-byte load during wave-coherent traversal.
//   TriF  (48 B, one per triangle, in Morton/leaf order): FP32 vertices a, b, c + original index.
// Child references: r >= 0 is internal node r, r < 0 is leaf ~r (a position in the TriF array).
#pragma once

#include <hip/hip_runtime.h>

#include "ssdf_c.h"

#include <cstddef>
#include <cstdint>

#pragma clang fp contract(off)

namespace ssdf_amd {

struct alignas(16) Node {
    float lmin[3], lmax[3];
    float rmin[3], rmax[3];
    int left, right;
    int pad0, pad1;
};
static_assert(sizeof(Node) == 64, "Node must be 64 bytes");

struct alignas(16) TriF {
    float a[3], b[3], c[3];
    int orig;
    int pad0, pad1;
};
static_assert(sizeof(TriF) == 48, "TriF must be 48 bytes");

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

// Blocks of one launch are handed to the 8 XCDs round-robin (block b -> XCD b mod 8).  This
// bijection gives each XCD a contiguous range of logical tiles so its private L2 sees one
// spatially compact region of the mesh.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
    const unsigned q = nblocks >> 3, r = nblocks & 7u;
    const unsigned x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace ssdf_amd

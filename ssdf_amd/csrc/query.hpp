// query.hpp — launch arguments shared by the query kernels and their host wrappers.
#pragma once

#include "common.hpp"
#include "mesh.hpp"

namespace ssdf_amd {

enum QuerySource : int { QSRC_POINTS = 0, QSRC_GRID = 1 };
// EDF: unsigned distance, min-accumulated into out_d (serial/ssdf.hpp:447-460 contract).
// SDF: signed distance + closest triangle, overwriting (serial/bvh/bvh.cpp:1641-1798 contract).
// CLOSEST: distance, closest point and triangle, updated only on strict improvement of out_d
//          (serial/bvh/bvh.cpp:125-197 contract).
enum QueryOut : int { OUT_EDF = 0, OUT_SDF = 1, OUT_CLOSEST = 2 };

// flags of the C ABI (ssdf_c.h)
enum : unsigned {
    QF_FIXUP = 1u,        // FP32 traversal + FP64 re-evaluation of near-tied candidates
    QF_SIGN_FP64 = 2u,    // sign always evaluated in FP64
    QF_PERLANE = 4u,      // one-thread-per-query traversal instead of wave-coherent
};

struct QueryArgs {
    // mesh
    const Node *nodes;
    const TriP *tris;
    const MeshMeta *meta;
    const int *s0, *s1, *s2;
    const void *sx, *sy, *sz;  // caller-precision vertices (float or double)
    const double *face_n, *edge_n, *vertex_n;
    // queries: points (SoA, caller precision) ...
    const void *qx, *qy, *qz;
    long long npoints;
    // ... or grid nodes x_i = gmin + i * gh (x fastest)
    int nx, ny, nz;
    int nbx, nby, nbz;  // 4x4x4 bricks per axis
    double gmin[3], gh[3];
    // outputs (caller precision T)
    void *out_d;
    int *out_tri;
    void *out_cx, *out_cy, *out_cz;
    // precision control
    float tau;      // relative near-tie band on distances
    float band_abs; // absolute near-tie band (FP32 error bound of a distance)
    float sign_eps; // FP64 sign below this distance
    double delta64; // FP64 pruning margin of the exact traversal
    // fixup queue: query ids whose candidate list overflowed
    unsigned *fix_count;
    unsigned *fix_list;
    unsigned fix_cap;
    int *err_flag;  // bit 1: a traversal stack would overflow; bit 2: iteration cap hit
    long long max_iters;  // traversal steps bound (2 x nodes + slack): a malformed tree cannot hang
    unsigned nblocks;  // launched blocks (XCD remap)
    unsigned flags;    // QF_* of the call
};

int launch_query(const DevMesh &m, QueryArgs &a, int qsrc, int out, unsigned flags, int out_is_double,
                 int q_is_double, hipStream_t stream);

}  // namespace ssdf_amd

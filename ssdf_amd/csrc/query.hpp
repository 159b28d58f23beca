// query.hpp — launch arguments shared by the query kernels and their host wrappers.
#pragma once

#include "common.hpp"
#include "mesh.hpp"
#include "radix.hpp"

#include <algorithm>
#include <cmath>

namespace ssdf_amd {

enum QuerySource : int { QSRC_POINTS = 0, QSRC_GRID = 1 };
// EDF: unsigned distance, min-accumulated into out_d (serial/ssdf.hpp:447-460 contract).
// SDF: signed distance + closest triangle, overwriting (serial/bvh/bvh.cpp:1641-1798 contract).
// CLOSEST: distance, closest point and triangle, updated only on strict improvement of out_d
//          (serial/bvh/bvh.cpp:125-197 contract).
// SQEDF: squared unsigned distance, min-accumulated into out_d (serial/bvh/bvh.cpp:1575-1639 contract).
enum QueryOut : int { OUT_EDF = 0, OUT_SDF = 1, OUT_CLOSEST = 2, OUT_SQEDF = 3 };

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
    const unsigned *perm;  // optional: wave slot j answers point perm[j] (Morton order)
    // ... or grid nodes x_i = gmin + i * gh (x fastest)
    int nx, ny, nz;
    int nbx, nby, nbz;  // bricks per axis
    int brick;          // brick shape of a wave's 64 grid nodes (brick_code; 0: 4 x 4 x 4)
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
    unsigned *fix_count;  // [0] queue entries written, [2] next entry for the per-query kernel
    unsigned *fix_list;   // 4 words per entry: query id (low 32 bits), FP32 bound (bits), tag (mesh of
                          // a batch), query id (high 32 bits)
    unsigned fix_cap;     // entries
    unsigned heavy_cap;   // entries packet waves may hand over (the rest is kept for FIX near-ties)
    unsigned fix_tag;
    int *err_flag;  // bit 1: a traversal stack would overflow; bit 2: iteration cap hit; bit 4: fixup queue full
    long long max_iters;  // traversal steps bound (2 x nodes + slack): a malformed tree cannot hang
    int heavy_iters;      // leaves a packet wave walks before handing its queries over (0: never)
    unsigned nblocks;  // launched blocks (XCD remap)
    // Persistent launch (persistent != 0): the waves of total_waves go out in 8 contiguous ranges,
    // one per XCD, through the counters wqueue[32 x] (128 bytes apart, zero at the launch); a wave
    // whose range is exhausted moves on to the next ones.
    unsigned *wqueue;
    unsigned total_waves;
    unsigned persistent;
    unsigned flags;    // QF_* of the call
    // winding numbers (gwn.hip)
    const GwnNode *gwn;
    float beta2;          // far-field acceptance: |c - q|^2 > beta2 * r^2
    float *out_w;         // GW_WINDING
    unsigned char *out_in;  // GW_INSIDE
};

// Compute units of the current device (cached per device).
unsigned device_cus();

int launch_query(const DevMesh &m, QueryArgs &a, int qsrc, int out, unsigned flags, int out_is_double,
                 int q_is_double, hipStream_t stream);

// Brick shapes: a wave answers 2^bx x 2^by x 2^bz grid nodes (bx + by + bz = 6), coded
// bx | by << 4 | bz << 8; code 0 is the default 4 x 4 x 4.
__host__ __device__ constexpr int brick_code(int bx, int by, int bz) { return bx | by << 4 | bz << 8; }
__host__ __device__ __forceinline__ void brick_log2(int code, int &bx, int &by, int &bz) {
    if (code == 0) code = brick_code(2, 2, 2);
    bx = code & 15, by = (code >> 4) & 15, bz = (code >> 8) & 15;
}
inline bool brick_valid(int code) {
    if (code == 0) return true;
    int bx, by, bz;
    brick_log2(code, bx, by, bz);
    return bx + by + bz == 6 && code == brick_code(bx, by, bz);
}
// Bricks per axis and waves of a grid under a brick shape: bricks are grouped 4 x 4 x 4 into
// super-bricks (64 waves, Morton order inside), so the waves cover whole super-bricks.
inline void grid_bricks(int code, int nx, int ny, int nz, int &nbx, int &nby, int &nbz, unsigned long long &waves) {
    int bx, by, bz;
    brick_log2(code, bx, by, bz);
    nbx = (nx + (1 << bx) - 1) >> bx, nby = (ny + (1 << by) - 1) >> by, nbz = (nz + (1 << bz) - 1) >> bz;
    waves = (unsigned long long)((nbx + 3) / 4) * ((nby + 3) / 4) * ((nbz + 3) / 4) * 64;
}

// Precision constants of the queries (single meshes and batches alike).
constexpr float kBandRel = 0x1p-18f;    // FP32 error bound of a distance, relative to the coordinate scale
constexpr float kTau = 1e-5f;           // relative near-tie band on distances
constexpr double kDelta64Rel = 1e-13;   // FP64 pruning margin relative to the coordinate scale
constexpr double kSignEpsRel = 1e-6;    // FP64 sign within this fraction of the bbox diagonal
// Leaves a packet wave walks before its queries go to the per-query kernel.  A wave of grid nodes
// outside or near a surface needs tens to a few thousand; deep inside a bumpy near-sphere, where
// the nearest regions differ from lane to lane, the shared walk grows to 10^5 leaves and more and
// one wave would run for hundreds of milliseconds.  Measured best of 1 Ki - 16 Ki on grids and
// point clouds around 1.4-2 M-triangle near-spheres and on 64 near-spheres of 1 k-500 k
// triangles; smaller budgets fill the hand-over share of the queue with moderately incoherent
// waves before the heavy ones arrive.
constexpr int kHeavyLeaves = 8192;
// The current budget (ssdf_set_heavy_leaves; initially kHeavyLeaves or SSDF_HEAVY_LEAVES).
int heavy_leaves();

// Coordinate scale max(1, |vertex bounds|) and bounding-box diagonal of a built mesh.
inline float meta_scale(const MeshMeta &hm) {
    float L = 1.0f;
    for (int d = 0; d < 6; d++) L = std::max(L, std::fabs(ordered_to_float(hm.bbox_ord[d])));
    return L;
}

inline float meta_diag(const MeshMeta &hm) {
    double s = 0;
    for (int d = 0; d < 3; d++) {
        const double e = (double)ordered_to_float(hm.bbox_ord[3 + d]) - ordered_to_float(hm.bbox_ord[d]);
        s += e * e;
    }
    return (float)std::sqrt(s);
}

// ------------------------------------------------------------------------------------------------
// Batches: many FP32 meshes in shared arenas, their queries answered by one persistent launch.
// Work unit = one wave's query set (a 4x4x4 brick of a mesh's grid, or 64 of its points); units
// are numbered mesh after mesh in the order of the descriptor array.
// ------------------------------------------------------------------------------------------------
struct BatchMesh {
    long long node_off, tri_off, vert_off;  // arena offsets (elements) of this mesh's records
    long long out_off;                      // offset of its first query in the query / output arrays
    long long npoints;                      // points mode
    int nx, ny, nz, nbx, nby, nbz;          // grid mode
    double gmin[3], gh[3];
    float band_abs, sign_eps;
    double delta64;
    unsigned unit_start;  // first work unit
    int mesh;             // index of the mesh in the batch (meta record, fixup tag)
    int brick;            // grid mode: brick shape (brick_code) of this mesh's waves
    int heavy;            // leaf budget of this mesh's packet waves (< 0: the process-wide one)
};

struct BatchArgs {
    QueryArgs base;  // fields shared by every mesh: tolerances, flags, fixup queue, error flag
    const Node *nodes;
    const TriP *tris;
    const MeshMeta *metas;
    const int *s0, *s1, *s2;
    const float *sx, *sy, *sz;
    const double *face_n, *edge_n, *vertex_n;
    const float *qx, *qy, *qz;  // points mode, concatenated per mesh at out_off
    float *out_d;
    int *out_tri;
    float *out_cx, *out_cy, *out_cz;
    int nmeshes;
    unsigned total_units;
    unsigned *queue;  // 8 counters (one per XCD range), 64 words apart
    unsigned chunk;   // units per grab
    int nlev;         // LDS levels of the per-query kernel: the deepest tree of the batch
    unsigned long long *cost;  // optional, per mesh: shader-clock cycles its waves spent (autotuning)
    const unsigned *order;     // optional: position -> work unit, each XCD range's units inside it
    unsigned *unit_cost;       // optional, per unit: shader-clock cycles of its packet wave (saturated)
};

// Waves per block of the packet and winding-number kernels.
constexpr unsigned kWavesPerBlock = 4;
// Internal nodes below which a tree's node records are addressed by 32-bit byte offsets (the
// STACK == 64 traversal); larger trees take the STACK == 128 one, which indexes in 64 bits.
constexpr long long kNarrowNodes = 1ll << 26;

// The waves a block of a packet or winding-number launch answers: its own 4 (one block per 4
// waves, block ids remapped over the XCDs), or, in a persistent launch, one at a time from the
// XCD's range of the wave order and then from the others' (QueryArgs::wqueue).
struct WaveSource {
    unsigned k = 0;      // persistent: ranges exhausted so far
    bool done = false;   // one block per kWavesPerBlock waves: its wave handed out
    __device__ explicit WaveSource(const QueryArgs &) {}
    __device__ bool next(const QueryArgs &a, unsigned &w) {
        if (!a.persistent) {
            if (done) return false;
            done = true;
            w = xcd_remap(blockIdx.x, a.nblocks) * kWavesPerBlock + (threadIdx.x >> 6);
            return true;
        }
        const unsigned U = a.total_waves, x0 = blockIdx.x & 7u;
        for (; k < 8; k++) {
            const unsigned x = (x0 + k) & 7u;
            const unsigned lo = (unsigned)(((unsigned long long)U * x) >> 3);
            const unsigned hi = (unsigned)(((unsigned long long)U * (x + 1)) >> 3);
            unsigned grab = 0;
            if ((threadIdx.x & 63) == 0) grab = atomicAdd(&a.wqueue[32 * x], 1u);
            grab = (unsigned)__builtin_amdgcn_readfirstlane(__shfl((int)grab, 0));
            if (grab < hi - lo) {
                w = lo + grab;
                return true;
            }
        }
        return false;
    }
};

// Launches the reset, the persistent packet kernel and the queue kernel of a batch.
// meshes: device descriptor array of b.nmeshes entries; stack: 64 or 128.
int launch_batch(const BatchArgs &b, const BatchMesh *meshes, int qsrc, int out, bool fix, int stack,
                 unsigned nblocks, hipStream_t stream);

// Scattered points: a wave of 64 consecutive points that lie all over the mesh walks the union of
// their 64 paths.  For FP32 point sets of at least kSortMin points the queries are ordered by a
// 30-bit Morton code of their position in the mesh's root box (4 radix passes) and the waves take
// them through the resulting permutation; results still land at the points' own positions.
constexpr long long kSortMin = 4096;
// Workspace of sort_points for n points: keys and values twice (ping-pong), the radix tile table
// and digit totals, and a margin.
constexpr size_t sort_ws_bytes(long long n) {
    return 16 * (size_t)(n > 1 ? n : 1) + 4 * radix_scratch_words((size_t)(n > 1 ? n : 1)) + 256;
}
// The device memory one query call works in, wherever it lives (a handle's workspace, a serving
// request's buffer, a static array): status words ([0] queue entries written, [1] kernel error
// flag, [2] queue read position), the queue of 4-word entries, and the Morton sort of the points
// (FP32 point sets of kSortMin points or more).
struct QueryWs {
    unsigned *status = nullptr;
    unsigned *list = nullptr;
    size_t entries = 0;
    void *sort = nullptr;
    size_t sort_bytes = 0;
};

constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Status area of a query workspace: words 0-3 (queue count, kernel error flag, queue read position),
// then the 8 wave-queue counters of a persistent launch at words 32 (x + 1), each on a 128-byte line
// of its own (cleared with the status words before every query).
constexpr size_t kStatusBytes = 1280;
constexpr unsigned kWaveQueueWord = 32;

// Queue entries reserved for nq queries (FIX near-ties that overflow the candidate list, and the
// queries of packet waves past their leaf budget): an eighth of the queries, at least 64 Ki.  Past
// that a near-tie keeps its FP32 result and sets bit 4 of the kernel error flag, and a heavy query
// stays in its packet wave.
constexpr size_t fixup_entries(size_t nq) { return std::min(nq, std::max<size_t>(1u << 16, nq / 8)); }

// Bytes of a contiguous QueryWs for nq queries (sorted: with room for the point sort).
constexpr size_t query_ws_bytes(size_t nq, bool sorted) {
    return kStatusBytes + round256(16 * std::max<size_t>(fixup_entries(nq), 1)) +
           (sorted ? round256(sort_ws_bytes((long long)nq)) : 0);
}

// Splits a 256-byte aligned buffer of query_ws_bytes(nq, sorted) bytes into a QueryWs.
inline QueryWs carve_ws(void *base, size_t nq, bool sorted) {
    QueryWs w;
    char *p = (char *)base;
    w.status = (unsigned *)p;
    w.list = (unsigned *)(p + kStatusBytes);
    w.entries = std::max<size_t>(fixup_entries(nq), 1);
    if (sorted) {
        w.sort = p + kStatusBytes + round256(16 * w.entries);
        w.sort_bytes = round256(sort_ws_bytes((long long)nq));
    }
    return w;
}

// A point query on a resident mesh in the caller's workspace `ws` (capi.hip).  handle_call: a call
// of the handle's owner (the allocation-free entry point), ordered after the handle's builds and
// its last call and recorded as its last call; it allocates nothing, so the winding-number tree a
// signed query may need must exist.  Otherwise a serving request: read-only, ordered by the server,
// and it may build that tree (synchronously, on the mesh's stream).
int ws_query_points(ssdf_mesh_t h, void *ws, bool handle_call, int mode, ptrdiff_t n, const void *x,
                    const void *y, const void *z, void *out_d, int *out_tri, void *cx, void *cy, void *cz,
                    unsigned flags, hipStream_t stream);

// perm (in ws) = the Morton order of the n points; ws holds sort_ws_bytes(n) bytes.
int sort_points(const MeshMeta *meta, long long n, const float *x, const float *y, const float *z, void *ws,
                const unsigned **perm, hipStream_t stream);

// Generalized winding number of the queries (gwn.hip): GW_WINDING writes w (float) to out_w,
// GW_INSIDE writes w >= 1/2 (uint8) to out_in, GW_SIGN gives out_d (mesh precision) the sign
// -|d| where w >= 1/2, +|d| elsewhere.  Builds the mesh's dipole tree on first use.
enum GwnMode : int { GW_WINDING = 0, GW_INSIDE = 1, GW_SIGN = 2 };
int launch_gwn(DevMesh &m, QueryArgs &a, int qsrc, int mode, int q_is_double, hipStream_t stream);

template <typename T>
__device__ __forceinline__ T load_as(const void *p, long long i);
template <>
__device__ __forceinline__ float load_as<float>(const void *p, long long i) {
    return ((const float *)p)[i];
}
template <>
__device__ __forceinline__ double load_as<double>(const void *p, long long i) {
    return ((const double *)p)[i];
}

template <typename T>
__device__ __forceinline__ void store_as(void *p, long long i, T v) {
    ((T *)p)[i] = v;
}

// Query of a lane.  Grid: waves map to bricks of 2^bx x 2^by x 2^bz nodes (4 x 4 x 4 by default),
// bricks to 4x4x4 super-bricks in Morton order (spatially compact waves and blocks), lanes x fastest
// inside the brick.  Points: 64 consecutive points per wave.
template <int QSRC, typename Q>
__device__ __forceinline__ bool lane_query(const QueryArgs &a, unsigned wave_id, int lane, long long &qid, Q &x,
                                           Q &y, Q &z) {
    if (QSRC == QSRC_POINTS) {
        const long long j = (long long)wave_id * 64 + lane;
        if (j >= a.npoints) return false;
        qid = a.perm ? (long long)a.perm[j] : j;  // spatially sorted order of scattered points
        x = load_as<Q>(a.qx, qid), y = load_as<Q>(a.qy, qid), z = load_as<Q>(a.qz, qid);
        return true;
    } else {
        const unsigned sb = wave_id >> 6, l = wave_id & 63u;
        const unsigned nsx = (a.nbx + 3) >> 2, nsy = (a.nby + 3) >> 2;
        const unsigned sbx = sb % nsx, sby = (sb / nsx) % nsy, sbz = sb / (nsx * nsy);
        const unsigned lx = (l & 1u) | ((l >> 2) & 2u), ly = ((l >> 1) & 1u) | ((l >> 3) & 2u),
                       lz = ((l >> 2) & 1u) | ((l >> 4) & 2u);
        int bx, by, bz;
        brick_log2(a.brick, bx, by, bz);
        const int i = (int)((sbx * 4 + lx) << bx) + (lane & ((1 << bx) - 1));
        const int j = (int)((sby * 4 + ly) << by) + ((lane >> bx) & ((1 << by) - 1));
        const int k = (int)((sbz * 4 + lz) << bz) + (lane >> (bx + by));
        if (i >= a.nx || j >= a.ny || k >= a.nz) return false;
        qid = (long long)i + (long long)a.nx * ((long long)j + (long long)a.ny * k);
        const Q hx = (Q)a.gh[0], hy = (Q)a.gh[1], hz = (Q)a.gh[2];
        x = (Q)a.gmin[0] + (Q)i * hx;
        y = (Q)a.gmin[1] + (Q)j * hy;
        z = (Q)a.gmin[2] + (Q)k * hz;
        return true;
    }
}


}  // namespace ssdf_amd

// batch.hip — many FP32 meshes resident in shared arenas, queried by one persistent launch.
//
// Every mesh keeps the single-mesh layout (LBVH nodes, Morton-ordered TriP records, pseudonormals,
// meta record), placed at its offset in one arena per kind.  The kernels take the arena bases as
// __restrict__ arguments and a wave adds its mesh's (uniform) offset, so node records still arrive
// by scalar loads.  A query call is three launches on the caller's stream — reset, the persistent
// packet kernel over all meshes' work units, the fixup-queue kernel — so a build + query sequence
// can be captured once into a hipGraph and replayed.
//
// Work order: units are numbered mesh after mesh, meshes sorted by decreasing tree depth (the
// costlier queries first), and split into 8 ranges, one per XCD, with stealing (k_batch_packet).
#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "query.hpp"

using namespace ssdf_amd;

// A mesh's share of the batch's connectivity arena: 3 corners of kConnBytesPerCorner bytes per
// triangle, rounded up to 64 so that every mesh's 8-byte edge keys stay aligned.
constexpr size_t kConnStride = 64;
static_assert(kConnStride >= 3 * kConnBytesPerCorner, "a triangle's corners fit its stride");

// Work-unit order: the packet kernel takes its units through `order` and records each unit's
// cycles in `unit_cost`; after it, the costliest units of each XCD range move to its front for the
// next query with the same plan (`key`).  ctr: per range, counts of the 32 cost classes (zero
// between queries), then per range two cursors and the lowest heavy class (4 words).
struct UnitOrder {
    unsigned *order = nullptr, *unit_cost = nullptr, *ctr = nullptr;
    size_t cap = 0;
    unsigned long long key = 0;
};

struct ssdf_batch {
    int nm = 0;
    bool with_normals = false;
    hipStream_t stream = nullptr;  // allocations are ordered on the creation stream
    std::vector<DevMesh> meshes;   // views into the arenas
    std::vector<long long> tri_off, vert_off, node_off;
    int *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;
    float *sx = nullptr, *sy = nullptr, *sz = nullptr;
    Node *nodes = nullptr;
    TriP *tris = nullptr;
    MeshMeta *metas = nullptr;
    double *face_n = nullptr, *edge_n = nullptr, *vertex_n = nullptr;
    char *conn = nullptr;  // per mesh, the sorted connectivity of its normals (kConnStride bytes a triangle)
    char *ws = nullptr;  // build workspace of lane 0
    size_t ws_cap = 0;
    // Build lanes: meshes are built concurrently on up to kLanes streams (lane 0: the caller's),
    // each lane with a workspace of its own; events fork the lanes off the caller's stream and
    // join them back (in a captured graph too).
    static constexpr int kLanes = 4;
    hipStream_t lane_st[kLanes] = {};
    hipEvent_t lane_ev[kLanes] = {};  // [0]: fork; [l]: join of lane l
    char *lane_ws[kLanes] = {};       // lanes 1..
    size_t lane_ws_cap[kLanes] = {};
    size_t device_bytes = 0;
    // query state
    BatchMesh *d_desc = nullptr, *h_desc = nullptr;  // device / pinned host descriptors
    int desc_cap = 0;
    unsigned *queue = nullptr;   // 8 x 64 words
    unsigned *status = nullptr;  // fixup count, error flag
    unsigned *fix_list = nullptr;
    size_t fix_cap = 0;  // entries
    int maxdepth = 0;
    int num_cus = 256;
    int blocks_per_cu = 0, chunk = 0;  // schedule (0: defaults)
    std::vector<int> mesh_brick;       // per mesh: brick shape of its grid waves (brick_code; 0: 4x4x4)
    std::vector<int> mesh_heavy;       // per mesh: leaf budget of its packet waves (-1: process-wide)
    unsigned long long *cost = nullptr;  // per-mesh cycle counters of the autotuner
    // Work-unit orders of eager queries and of the captured graph (each its own, so that neither
    // sees an order made for the other's plan).
    UnitOrder eager_order, graph_order;
    // The captured graph's own descriptors and fixup queue: eager queries of another plan rewrite
    // (and may reallocate) the shared ones between replays.
    BatchMesh *graph_desc = nullptr;
    unsigned *graph_fix = nullptr;
    size_t graph_fix_cap = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
};

namespace {

// The caller may have destroyed the batch's creation stream already: the device is synchronized
// and the frees go to the null stream.
void destroy_batch(ssdf_batch *b) {
    if (!b) return;
    (void)hipDeviceSynchronize();
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
    const hipStream_t st = nullptr;
    for (void *p : {(void *)b->s0, (void *)b->s1, (void *)b->s2, (void *)b->sx, (void *)b->sy, (void *)b->sz,
                    (void *)b->nodes, (void *)b->tris, (void *)b->metas, (void *)b->face_n, (void *)b->edge_n,
                    (void *)b->vertex_n, (void *)b->conn, (void *)b->ws, (void *)b->d_desc, (void *)b->queue, (void *)b->status,
                    (void *)b->fix_list, (void *)b->cost, (void *)b->eager_order.order, (void *)b->eager_order.unit_cost,
                    (void *)b->eager_order.ctr, (void *)b->graph_order.order, (void *)b->graph_order.unit_cost,
                    (void *)b->graph_order.ctr, (void *)b->graph_desc, (void *)b->graph_fix})
        dev_free(p, st);
    for (int l = 1; l < ssdf_batch::kLanes; l++) dev_free(b->lane_ws[l], st);
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    (void)hipStreamSynchronize(st);
    for (int l = 0; l < ssdf_batch::kLanes; l++) {
        if (b->lane_st[l]) (void)hipStreamDestroy(b->lane_st[l]);
        if (b->lane_ev[l]) (void)hipEventDestroy(b->lane_ev[l]);
    }
    delete b;
}

template <typename T>
int arena(T **p, size_t n, ssdf_batch *b) {
    SSDF_HIP_TRY(dev_alloc((void **)p, sizeof(T) * std::max<size_t>(n, 1), b->stream));
    b->device_bytes += sizeof(T) * n;
    return SSDF_OK;
}

// Builds mesh i (LBVH, and normals) on stream st with lane l's workspace.
int build_one(ssdf_batch *b, int i, int l, bool normals, hipStream_t st) {
    char *&ws = l ? b->lane_ws[l] : b->ws;
    size_t &cap = l ? b->lane_ws_cap[l] : b->ws_cap;
    DevMesh &m = b->meshes[i];
    m.ws = ws, m.ws_cap = cap;
    int rc = build_lbvh(m, st);
    if (!rc && normals) rc = build_normals(m, st);
    ws = m.ws, cap = m.ws_cap;  // grown if it had to (only before any capture: the same meshes
                                // land on the same lanes every time)
    m.ws = nullptr, m.ws_cap = 0;
    return rc;
}

// Builds every mesh, the lanes running concurrently: largest mesh first onto the least loaded
// lane.  Lane l > 0 runs on the batch's own stream, forked off st and joined back into it.
int build_lanes(ssdf_batch *b, bool normals, hipStream_t st) {
    constexpr int L = ssdf_batch::kLanes;
    for (int l = 0; l < L; l++) {
        if (!b->lane_ev[l]) SSDF_HIP_TRY(hipEventCreateWithFlags(&b->lane_ev[l], hipEventDisableTiming));
        if (l && !b->lane_st[l]) SSDF_HIP_TRY(hipStreamCreateWithFlags(&b->lane_st[l], hipStreamNonBlocking));
    }
    std::vector<int> order(b->nm);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->meshes[x].nsel > b->meshes[y].nsel; });
    std::vector<int> lane_of(b->nm, 0);
    long long load[L] = {};
    for (int i : order) {
        int l = 0;
        for (int k = 1; k < L; k++)
            if (load[k] < load[l]) l = k;
        lane_of[i] = l;
        load[l] += std::max<long long>(b->meshes[i].nsel, 0);
    }
    SSDF_HIP_TRY(hipEventRecord(b->lane_ev[0], st));
    for (int l = 1; l < L; l++) SSDF_HIP_TRY(hipStreamWaitEvent(b->lane_st[l], b->lane_ev[0], 0));
    int rc = SSDF_OK;
    for (int i : order)
        if (!rc && b->meshes[i].nsel > 0) {
            const int l = lane_of[i];
            rc = build_one(b, i, l, normals, l ? b->lane_st[l] : st);
        }
    // join every lane, also after an error (a capture must end with its streams joined)
    for (int l = 1; l < L; l++) {
        SSDF_HIP_TRY(hipEventRecord(b->lane_ev[l], b->lane_st[l]));
        SSDF_HIP_TRY(hipStreamWaitEvent(st, b->lane_ev[l], 0));
    }
    return rc;
}

// Copies every meta record back (one synchronisation) and records the deepest tree.
int sync_metas(ssdf_batch *b, hipStream_t st) {
    std::vector<MeshMeta> h(b->nm);
    SSDF_HIP_TRY(hipMemcpyAsync(h.data(), b->metas, sizeof(MeshMeta) * b->nm, hipMemcpyDeviceToHost, st));
    SSDF_HIP_TRY(hipStreamSynchronize(st));
    b->maxdepth = 0;
    for (int i = 0; i < b->nm; i++) {
        if (b->meshes[i].nsel <= 0) {  // never built: no record on the device
            b->meshes[i].hmeta = MeshMeta{};
            continue;
        }
        b->meshes[i].hmeta = h[i];
        b->maxdepth = std::max(b->maxdepth, h[i].maxdepth);
    }
    return SSDF_OK;
}

int rebuild_all(ssdf_batch *b, bool normals, hipStream_t st) {
    if (int rc = build_lanes(b, normals, st)) return rc;
    for (int i = 0; i < b->nm; i++) b->meshes[i].has_normals = normals && b->meshes[i].nsel > 0;
    return SSDF_OK;
}

int ensure_desc(ssdf_batch *b, int n) {
    if (b->desc_cap >= n) return SSDF_OK;
    dev_free(b->d_desc, b->stream);
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    b->d_desc = nullptr, b->h_desc = nullptr, b->desc_cap = 0;
    SSDF_HIP_TRY(dev_alloc((void **)&b->d_desc, sizeof(BatchMesh) * n, b->stream));
    SSDF_HIP_TRY(hipHostMalloc((void **)&b->h_desc, sizeof(BatchMesh) * n, hipHostMallocDefault));
    b->desc_cap = n;
    return SSDF_OK;
}

int ensure_fix(ssdf_batch *b, size_t entries) {
    if (b->fix_cap >= entries) return SSDF_OK;
    dev_free(b->fix_list, b->stream);
    b->fix_list = nullptr, b->fix_cap = 0;
    SSDF_HIP_TRY(dev_alloc((void **)&b->fix_list, sizeof(unsigned) * 4 * entries, b->stream));
    b->fix_cap = entries;
    return SSDF_OK;
}

// Per-mesh descriptors of a query (grid: dims/boxes, points: offsets) in work-unit order.
struct Plan {
    std::vector<BatchMesh> desc;
    unsigned total_units = 0;
    size_t total_queries = 0;
};

int plan_queries(ssdf_batch *b, int qsrc, const ptrdiff_t *dims, const double *boxes, const ptrdiff_t *offsets,
                 Plan &p) {
    std::vector<int> order(b->nm);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return b->meshes[x].hmeta.maxdepth > b->meshes[y].hmeta.maxdepth;
    });
    unsigned long long units = 0;
    for (int i : order) {
        const DevMesh &m = b->meshes[i];
        BatchMesh d{};
        d.node_off = b->node_off[i], d.tri_off = b->tri_off[i], d.vert_off = b->vert_off[i];
        d.mesh = i;
        d.brick = b->mesh_brick[i];
        d.heavy = b->mesh_heavy[i];
        const float L = meta_scale(m.hmeta);
        d.band_abs = kBandRel * L;
        d.sign_eps = (float)(kSignEpsRel * meta_diag(m.hmeta));
        d.delta64 = kDelta64Rel * L;
        unsigned long long u = 0;
        if (qsrc == QSRC_GRID) {
            const ptrdiff_t n[3] = {dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]};
            if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] > 0x7fffffff || n[1] > 0x7fffffff || n[2] > 0x7fffffff) {
                set_error("batch grid: invalid dimensions");
                return SSDF_ERR_INVALID;
            }
            d.nx = (int)n[0], d.ny = (int)n[1], d.nz = (int)n[2];
            d.out_off = offsets[i];
            const double *lo = boxes + 6 * i, *hi = boxes + 6 * i + 3;
            for (int k = 0; k < 3; k++) {
                const float l = (float)lo[k], h = (float)hi[k];
                d.gmin[k] = l;
                d.gh[k] = n[k] > 1 ? (double)((h - l) / (float)(n[k] - 1)) : 0.0;
            }
            unsigned long long w = 0;
            grid_bricks(d.brick, d.nx, d.ny, d.nz, d.nbx, d.nby, d.nbz, w);
            if (m.nsel > 0 && d.nx > 0 && d.ny > 0 && d.nz > 0) u = w;
            p.total_queries += (size_t)n[0] * n[1] * n[2];
        } else {
            d.out_off = offsets[i];
            d.npoints = offsets[i + 1] - offsets[i];
            if (d.npoints < 0) {
                set_error("batch points: offsets must not decrease");
                return SSDF_ERR_INVALID;
            }
            if (m.nsel > 0) u = (unsigned long long)(d.npoints + 63) / 64;
            p.total_queries += (size_t)d.npoints;
        }
        d.unit_start = (unsigned)units;
        units += u;
        if (units > 0xffffffffull) {
            set_error("batch: more than 2^32 work units");
            return SSDF_ERR_INVALID;
        }
        p.desc.push_back(d);
    }
    p.total_units = (unsigned)units;
    return SSDF_OK;
}

// Uploads the plan (pinned staging, stream-ordered) and fills the launch arguments.
int stage(ssdf_batch *b, const Plan &p, int qsrc, int mode, unsigned flags, const float *x, const float *y,
          const float *z, float *out_d, int *out_tri, float *cx, float *cy, float *cz, hipStream_t st,
          BatchArgs &a) {
    if (int rc = ensure_desc(b, b->nm)) return rc;
    if (!b->queue) SSDF_HIP_TRY(dev_alloc((void **)&b->queue, sizeof(unsigned) * 8 * 64, b->stream));
    if (!b->status) SSDF_HIP_TRY(dev_alloc((void **)&b->status, 64, b->stream));
    const bool fix = (flags & SSDF_FLAG_FIXUP) != 0;
    const size_t entries = std::min(p.total_queries, std::max<size_t>(1u << 16, p.total_queries / 8));
    if (int rc = ensure_fix(b, std::max<size_t>(entries, 1))) return rc;
    if (st != b->stream) {
        // the buffers above are ordered on the creation stream
        hipEvent_t ev;
        SSDF_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        SSDF_HIP_TRY(hipEventRecord(ev, b->stream));
        SSDF_HIP_TRY(hipStreamWaitEvent(st, ev, 0));
        (void)hipEventDestroy(ev);
    }
    // the pinned staging buffer may still feed an earlier call's copy
    SSDF_HIP_TRY(hipStreamSynchronize(st));
    std::memcpy(b->h_desc, p.desc.data(), sizeof(BatchMesh) * p.desc.size());
    SSDF_HIP_TRY(hipMemcpyAsync(b->d_desc, b->h_desc, sizeof(BatchMesh) * p.desc.size(), hipMemcpyHostToDevice, st));
    std::memset(&a, 0, sizeof(a));
    a.base.tau = kTau;
    a.base.flags = flags;
    a.base.fix_count = b->status;
    a.base.err_flag = (int *)(b->status + 1);
    a.base.fix_list = b->fix_list;
    a.base.fix_cap = (unsigned)std::min<size_t>(b->fix_cap, 0xffffffffu);
    long long maxsel = 1;
    for (auto &m : b->meshes) maxsel = std::max<long long>(maxsel, m.nsel);
    a.base.max_iters = 4 * maxsel + 64;
    a.base.heavy_iters = heavy_leaves();
    a.base.heavy_cap = a.base.fix_cap / 2;
    a.nlev = std::max(b->maxdepth, 1);
    a.nodes = b->nodes, a.tris = b->tris, a.metas = b->metas;
    a.s0 = b->s0, a.s1 = b->s1, a.s2 = b->s2;
    a.sx = b->sx, a.sy = b->sy, a.sz = b->sz;
    a.face_n = b->face_n, a.edge_n = b->edge_n, a.vertex_n = b->vertex_n;
    a.qx = x, a.qy = y, a.qz = z;
    a.out_d = out_d, a.out_tri = out_tri, a.out_cx = cx, a.out_cy = cy, a.out_cz = cz;
    a.nmeshes = (int)p.desc.size();
    a.total_units = p.total_units;
    a.queue = b->queue;
    a.chunk = (unsigned)(b->chunk > 0 ? b->chunk : 1);
    (void)qsrc;
    (void)mode;
    return SSDF_OK;
}

// ---- cost-ordered work units ----
// Range x of U units is [U x / 8, U (x + 1) / 8) (k_batch_packet's split); a unit's cost class is
// the bit length of its cycle count (32 classes).  The units of a range's kHeavyClasses costliest
// non-empty classes (within 2^kHeavyClasses of its costliest unit) move to the front of the range,
// the others keep their places' order behind them: a heavy unit that starts late would run on alone
// after every other wave has finished, while sorting all units by cost scatters the meshes over
// time and loses the locality of the natural order (mesh after mesh, bricks in Morton order).
constexpr int kOrderClasses = 32;
constexpr int kHeavyClasses = 4;

__device__ __forceinline__ unsigned unit_range(unsigned u, unsigned U) {
    return (unsigned)((8ull * (u + 1ull) - 1ull) / U);
}
__device__ __forceinline__ unsigned cost_class(unsigned c) { return 31u - (unsigned)__clz(c | 1u); }

__global__ __launch_bounds__(256) void k_order_identity(unsigned *order, unsigned U, unsigned *ctr) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < U) order[i] = i;
    if (blockIdx.x == 0 && threadIdx.x < 8 * kOrderClasses) ctr[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(1024) void k_order_count(const unsigned *__restrict__ cost, unsigned U, unsigned *ctr) {
    __shared__ unsigned h[8 * kOrderClasses];
    if (threadIdx.x < 8 * kOrderClasses) h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < U) atomicAdd(&h[unit_range(u, U) * kOrderClasses + cost_class(cost[u])], 1u);
    __syncthreads();
    if (threadIdx.x < 8 * kOrderClasses && h[threadIdx.x]) atomicAdd(&ctr[threadIdx.x], h[threadIdx.x]);
}

// One block of 8 threads, one per range: the lowest heavy class and the two cursors (heavy units
// from the range's start, the others after them), then the counts cleared.
__global__ __launch_bounds__(64) void k_order_offsets(unsigned U, unsigned *ctr) {
    const unsigned x = threadIdx.x;
    if (x >= 8) return;
    unsigned *cnt = ctr + x * kOrderClasses;
    int top = kOrderClasses - 1;
    while (top > 0 && cnt[top] == 0u) top--;
    const int thr = top - kHeavyClasses + 1;
    unsigned heavy = 0;
    for (int c = kOrderClasses - 1; c >= 0 && c >= thr; c--) heavy += cnt[c];
    const unsigned lo = (unsigned)(((unsigned long long)U * x) >> 3);
    unsigned *cur = ctr + 8 * kOrderClasses + 4 * x;
    cur[0] = lo;          // heavy
    cur[1] = lo + heavy;  // the rest
    cur[2] = (unsigned)max(thr, 0);
    for (int c = 0; c < kOrderClasses; c++) cnt[c] = 0u;
}

// A block of 1024 units (at most two ranges) counts its units per cursor in LDS and takes each
// cursor's places with one atomic: one atomic per unit, all on the 16 cursors, cost 14 ms; one per
// wave still 0.57 ms (device-scope atomics on one address from every XCD serialise).
__global__ __launch_bounds__(1024) void k_order_scatter(const unsigned *__restrict__ cost, unsigned U, unsigned *ctr,
                                                        unsigned *order) {
    __shared__ unsigned cnt[16], base[16];
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned u = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned key = 0, rank = 0;
    if (u < U) {
        const unsigned x = unit_range(u, U);
        key = 2 * x + (cost_class(cost[u]) < ctr[8 * kOrderClasses + 4 * x + 2] ? 1u : 0u);
        rank = atomicAdd(&cnt[key], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16 && cnt[threadIdx.x])
        base[threadIdx.x] = atomicAdd(&ctr[8 * kOrderClasses + 4 * (threadIdx.x >> 1) + (threadIdx.x & 1)], cnt[threadIdx.x]);
    __syncthreads();
    if (u < U) order[base[key] + rank] = u;
}

// Key of a plan's work units: the order of one query serves the next only if their units agree.
unsigned long long plan_key(const Plan &p, int qsrc) {
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ull; };
    mix((unsigned long long)qsrc);
    mix(p.total_units);
    for (const BatchMesh &d : p.desc) {
        mix((unsigned long long)d.mesh), mix(d.unit_start), mix((unsigned long long)d.brick);
        mix((unsigned long long)d.nx << 32 | (unsigned)d.ny), mix((unsigned long long)d.nz);
        mix((unsigned long long)d.npoints);
    }
    return h | 1ull;  // never 0 (no order yet)
}

// Order buffers for p's units; an identity order when the last order was made for another plan.
int prepare_order(ssdf_batch *b, UnitOrder &o, const Plan &p, int qsrc, hipStream_t st, BatchArgs &a) {
    const unsigned U = p.total_units;
    if (o.cap < U) {
        dev_free(o.order, b->stream), dev_free(o.unit_cost, b->stream);
        o.order = o.unit_cost = nullptr, o.cap = 0, o.key = 0;
        SSDF_HIP_TRY(dev_alloc((void **)&o.order, sizeof(unsigned) * U, b->stream));
        SSDF_HIP_TRY(dev_alloc((void **)&o.unit_cost, sizeof(unsigned) * U, b->stream));
        o.cap = U;
    }
    if (!o.ctr) SSDF_HIP_TRY(dev_alloc((void **)&o.ctr, sizeof(unsigned) * 2 * 8 * kOrderClasses, b->stream));
    if (st != b->stream) {
        hipEvent_t ev;
        SSDF_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        SSDF_HIP_TRY(hipEventRecord(ev, b->stream));
        SSDF_HIP_TRY(hipStreamWaitEvent(st, ev, 0));
        (void)hipEventDestroy(ev);
    }
    const unsigned long long key = plan_key(p, qsrc);
    if (key != o.key) {
        hipLaunchKernelGGL(k_order_identity, dim3((U + 255) / 256), dim3(256), 0, st, o.order, U, o.ctr);
        SSDF_HIP_TRY(hipGetLastError());
        o.key = key;
    }
    a.order = o.order;
    a.unit_cost = o.unit_cost;
    return SSDF_OK;
}

// After a query: the next order from the cycles its units just recorded.
int reorder_units(const UnitOrder &o, unsigned U, hipStream_t st) {
    hipLaunchKernelGGL(k_order_count, dim3((U + 1023) / 1024), dim3(1024), 0, st, o.unit_cost, U, o.ctr);
    hipLaunchKernelGGL(k_order_offsets, dim3(1), dim3(64), 0, st, U, o.ctr);
    hipLaunchKernelGGL(k_order_scatter, dim3((U + 1023) / 1024), dim3(1024), 0, st, o.unit_cost, U, o.ctr, o.order);
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

// Stack width of a batch's traversal: 64 entries (32-bit node offsets) while every tree is at most
// 64 deep and has fewer than kNarrowNodes internal nodes, else 128 (64-bit offsets).
int stack_of(const ssdf_batch *b) {
    long long maxsel = 0;
    for (const auto &m : b->meshes) maxsel = std::max<long long>(maxsel, m.nsel);
    if (b->maxdepth <= 64 && maxsel - 1 < kNarrowNodes) return 64;
    if (b->maxdepth <= 128) return 128;
    return 0;
}

// Default: 8 blocks per CU (8 waves per SIMD, the packet kernel's occupancy), what the autotuner
// picks on the 64-mesh batch in both modes.
constexpr int kDefaultBlocksPerCu = 8;
unsigned grid_blocks(const ssdf_batch *b, unsigned total_units, bool fix) {
    (void)fix;
    const int bpc = b->blocks_per_cu > 0 ? b->blocks_per_cu : kDefaultBlocksPerCu;
    const unsigned want = (unsigned)(b->num_cus * bpc);
    return std::max(1u, std::min(want, (total_units + 3) / 4));
}

int run_query(ssdf_batch *b, int qsrc, int mode, const ptrdiff_t *dims, const double *boxes,
              const ptrdiff_t *offsets, const float *x, const float *y, const float *z, float *out_d, int *out_tri,
              float *cx, float *cy, float *cz, unsigned flags, hipStream_t st, unsigned long long *cost = nullptr) {
    if (mode == SSDF_MODE_SDF && !b->with_normals) {
        set_error("batch: signed queries need the meshes' pseudonormals");
        return SSDF_ERR_INVALID;
    }
    if (!out_d || (mode == SSDF_MODE_CLOSEST && (!cx || !cy || !cz))) {
        set_error("batch query: null output");
        return SSDF_ERR_INVALID;
    }
    Plan p;
    if (int rc = plan_queries(b, qsrc, dims, boxes, offsets, p)) return rc;
    if (p.total_units == 0) return SSDF_OK;
    BatchArgs a;
    if (int rc = stage(b, p, qsrc, mode, flags, x, y, z, out_d, out_tri, cx, cy, cz, st, a)) return rc;
    const int stack = stack_of(b);
    if (!stack) {
        set_error("batch: LBVH deeper than 128 levels");
        return SSDF_ERR_DEPTH;
    }
    const bool fix = (flags & SSDF_FLAG_FIXUP) != 0;
    a.cost = cost;
    if (int rc = prepare_order(b, b->eager_order, p, qsrc, st, a)) return rc;
    if (int rc = launch_batch(a, b->d_desc, qsrc, mode, fix, stack, grid_blocks(b, p.total_units, fix), st)) return rc;
    return reorder_units(b->eager_order, p.total_units, st);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// One grid query with every mesh's waves timed (shader-clock cycles summed per mesh, packet and
// per-query kernels alike) into cost[nm] on the host.
int timed_costs(ssdf_batch *b, int mode, const ptrdiff_t *dims, const double *boxes, const ptrdiff_t *out_offsets,
                float *out_d, int *out_tri, unsigned flags, hipStream_t st, std::vector<unsigned long long> &cost) {
    if (!b->cost) SSDF_HIP_TRY(dev_alloc((void **)&b->cost, sizeof(unsigned long long) * b->nm, b->stream));
    if (st != b->stream) SSDF_HIP_TRY(hipStreamSynchronize(b->stream));
    hipLaunchKernelGGL(k_zero_u64, dim3((b->nm + 255) / 256), dim3(256), 0, st, b->cost, b->nm);
    if (int rc = run_query(b, QSRC_GRID, mode, dims, boxes, out_offsets, nullptr, nullptr, nullptr, out_d, out_tri,
                           nullptr, nullptr, nullptr, flags, st, b->cost))
        return rc;
    cost.assign(b->nm, 0ull);
    SSDF_HIP_TRY(hipMemcpyAsync(cost.data(), b->cost, sizeof(unsigned long long) * b->nm, hipMemcpyDeviceToHost, st));
    SSDF_HIP_TRY(hipStreamSynchronize(st));
    return SSDF_OK;
}

}  // namespace

extern "C" {

int ssdf_batch_create(int nmeshes, const ptrdiff_t *tri_offsets, const ptrdiff_t *vert_offsets, const int *s0,
                      const int *s1, const int *s2, const float *sx, const float *sy, const float *sz, int device_ptrs,
                      int with_normals, void *stream, ssdf_batch_t *out) {
    if (!out || nmeshes <= 0 || !tri_offsets || !vert_offsets) {
        set_error("ssdf_batch_create: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    *out = nullptr;
    const ptrdiff_t T = tri_offsets[nmeshes], V = vert_offsets[nmeshes];
    if (tri_offsets[0] != 0 || vert_offsets[0] != 0 || T < 0 || V < 0 ||
        (T > 0 && (!s0 || !s1 || !s2)) || (V > 0 && (!sx || !sy || !sz))) {
        set_error("ssdf_batch_create: offsets must start at 0; arrays must be non-null");
        return SSDF_ERR_INVALID;
    }
    for (int i = 0; i < nmeshes; i++) {
        const ptrdiff_t t = tri_offsets[i + 1] - tri_offsets[i], v = vert_offsets[i + 1] - vert_offsets[i];
        if (t < 0 || v < 0 || t >= ((ptrdiff_t)1 << kLeafBits) || (t > 0 && v == 0)) {
            set_error("ssdf_batch_create: offsets must not decrease; a mesh with triangles needs vertices");
            return SSDF_ERR_INVALID;
        }
    }
    ssdf_batch *b = new (std::nothrow) ssdf_batch();
    if (!b) return SSDF_ERR_OOM;
    struct Guard {
        ssdf_batch *b;
        ~Guard() {
            if (b) destroy_batch(b);
        }
    } guard{b};
    b->nm = nmeshes;
    b->stream = (hipStream_t)stream;
    b->with_normals = with_normals != 0;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            b->num_cus = cus;
    }
    b->tri_off.resize(nmeshes);
    b->vert_off.resize(nmeshes);
    b->node_off.resize(nmeshes);
    long long nodes_total = 0;
    for (int i = 0; i < nmeshes; i++) {
        b->tri_off[i] = tri_offsets[i];
        b->vert_off[i] = vert_offsets[i];
        b->node_off[i] = nodes_total;
        const long long t = tri_offsets[i + 1] - tri_offsets[i];
        nodes_total += t > 1 ? t - 1 : 1;
    }
    if (int rc = arena(&b->s0, T, b)) return rc;
    if (int rc = arena(&b->s1, T, b)) return rc;
    if (int rc = arena(&b->s2, T, b)) return rc;
    if (int rc = arena(&b->sx, V, b)) return rc;
    if (int rc = arena(&b->sy, V, b)) return rc;
    if (int rc = arena(&b->sz, V, b)) return rc;
    if (int rc = arena(&b->nodes, nodes_total, b)) return rc;
    if (int rc = arena(&b->tris, T, b)) return rc;
    if (int rc = arena(&b->metas, nmeshes, b)) return rc;
    if (b->with_normals) {
        if (int rc = arena(&b->face_n, 3 * (size_t)T, b)) return rc;
        if (int rc = arena(&b->edge_n, 9 * (size_t)T, b)) return rc;
        if (int rc = arena(&b->vertex_n, 3 * (size_t)V, b)) return rc;
        if (int rc = arena(&b->conn, kConnStride * (size_t)T, b)) return rc;
    }
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipStream_t st = b->stream;
    if (T > 0) {
        SSDF_HIP_TRY(hipMemcpyAsync(b->s0, s0, sizeof(int) * T, kind, st));
        SSDF_HIP_TRY(hipMemcpyAsync(b->s1, s1, sizeof(int) * T, kind, st));
        SSDF_HIP_TRY(hipMemcpyAsync(b->s2, s2, sizeof(int) * T, kind, st));
    }
    if (V > 0) {
        SSDF_HIP_TRY(hipMemcpyAsync(b->sx, sx, sizeof(float) * V, kind, st));
        SSDF_HIP_TRY(hipMemcpyAsync(b->sy, sy, sizeof(float) * V, kind, st));
        SSDF_HIP_TRY(hipMemcpyAsync(b->sz, sz, sizeof(float) * V, kind, st));
    }
    b->meshes.resize(nmeshes);
    b->mesh_brick.assign(nmeshes, 0);
    b->mesh_heavy.assign(nmeshes, -1);
    for (int i = 0; i < nmeshes; i++) {
        DevMesh &m = b->meshes[i];
        m.dtype = 0;
        m.nsel = tri_offsets[i + 1] - tri_offsets[i];
        m.nsp = vert_offsets[i + 1] - vert_offsets[i];
        m.s0 = b->s0 + b->tri_off[i], m.s1 = b->s1 + b->tri_off[i], m.s2 = b->s2 + b->tri_off[i];
        m.sx = b->sx + b->vert_off[i], m.sy = b->sy + b->vert_off[i], m.sz = b->sz + b->vert_off[i];
        m.owns_input = false;
        m.nodes = b->nodes + b->node_off[i];
        m.tris = b->tris + b->tri_off[i];
        m.meta = b->metas + i;
        if (b->with_normals) {
            m.face_n = b->face_n + 3 * b->tri_off[i];
            m.edge_n = b->edge_n + 9 * b->tri_off[i];
            m.vertex_n = b->vertex_n + 3 * b->vert_off[i];
            m.conn = b->conn + kConnStride * b->tri_off[i];
        }
        m.stream = st;
    }
    if (int rc = rebuild_all(b, b->with_normals, st)) return rc;
    if (int rc = sync_metas(b, st)) return rc;
    b->device_bytes += b->ws_cap;
    for (int l = 1; l < ssdf_batch::kLanes; l++) b->device_bytes += b->lane_ws_cap[l];
    *out = b;
    guard.b = nullptr;
    return SSDF_OK;
}

int ssdf_batch_destroy(ssdf_batch_t batch) {
    destroy_batch(batch);
    return SSDF_OK;
}

int ssdf_batch_rebuild(ssdf_batch_t b, int with_normals, void *stream) {
    if (!b) {
        set_error("ssdf_batch_rebuild: null batch");
        return SSDF_ERR_INVALID;
    }
    if (with_normals && !b->with_normals) {
        set_error("ssdf_batch_rebuild: the batch was created without normals");
        return SSDF_ERR_INVALID;
    }
    const hipStream_t st = (hipStream_t)stream;
    if (st != b->stream) SSDF_HIP_TRY(hipStreamSynchronize(b->stream));
    if (int rc = rebuild_all(b, with_normals != 0, st)) return rc;
    return sync_metas(b, st);
}

int ssdf_batch_query_grids(ssdf_batch_t b, int mode, const ptrdiff_t *dims, const double *boxes,
                           const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream) {
    if (!b || !dims || !boxes || !out_offsets || mode == SSDF_MODE_CLOSEST) {
        set_error("ssdf_batch_query_grids: invalid arguments (closest points need point queries)");
        return SSDF_ERR_INVALID;
    }
    return run_query(b, QSRC_GRID, mode, dims, boxes, out_offsets, nullptr, nullptr, nullptr, out_d, out_tri,
                     nullptr, nullptr, nullptr, flags, (hipStream_t)stream);
}

int ssdf_batch_query_points(ssdf_batch_t b, int mode, const ptrdiff_t *point_offsets, const float *x,
                            const float *y, const float *z, float *out_d, int *out_tri, float *out_cx, float *out_cy,
                            float *out_cz, unsigned flags, void *stream) {
    if (!b || !point_offsets || !x || !y || !z || point_offsets[0] != 0) {
        set_error("ssdf_batch_query_points: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    return run_query(b, QSRC_POINTS, mode, nullptr, nullptr, point_offsets, x, y, z, out_d, out_tri, out_cx, out_cy,
                     out_cz, flags, (hipStream_t)stream);
}

int ssdf_batch_set_schedule(ssdf_batch_t b, int blocks_per_cu, int units_per_grab) {
    if (!b || blocks_per_cu < 0 || blocks_per_cu > 32 || units_per_grab < 0 || units_per_grab > 256) {
        set_error("ssdf_batch_set_schedule: blocks_per_cu in [0, 32], units_per_grab in [0, 256]");
        return SSDF_ERR_INVALID;
    }
    b->blocks_per_cu = blocks_per_cu;
    b->chunk = units_per_grab;
    return SSDF_OK;
}

int ssdf_batch_set_mesh_schedule(ssdf_batch_t b, int mesh, int brick_x, int brick_y, int brick_z, int heavy_leaves) {
    if (!b || mesh < 0 || mesh >= b->nm || heavy_leaves < -1) {
        set_error("ssdf_batch_set_mesh_schedule: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    int lg[3] = {-1, -1, -1};
    const int ext[3] = {brick_x, brick_y, brick_z};
    for (int d = 0; d < 3; d++)
        for (int k = 0; k <= 6; k++)
            if (ext[d] == (1 << k)) lg[d] = k;
    if (lg[0] < 0 || lg[1] < 0 || lg[2] < 0 || lg[0] + lg[1] + lg[2] != 6) {
        set_error("ssdf_batch_set_mesh_schedule: the brick is 2^a x 2^b x 2^c = 64 nodes");
        return SSDF_ERR_INVALID;
    }
    b->mesh_brick[mesh] = brick_code(lg[0], lg[1], lg[2]);
    b->mesh_heavy[mesh] = heavy_leaves;
    return SSDF_OK;
}

int ssdf_batch_get_mesh_schedule(ssdf_batch_t b, int mesh, int *brick_x, int *brick_y, int *brick_z,
                                 int *heavy_leaves) {
    if (!b || mesh < 0 || mesh >= b->nm) {
        set_error("ssdf_batch_get_mesh_schedule: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    int bx, by, bz;
    brick_log2(b->mesh_brick[mesh], bx, by, bz);
    if (brick_x) *brick_x = 1 << bx;
    if (brick_y) *brick_y = 1 << by;
    if (brick_z) *brick_z = 1 << bz;
    if (heavy_leaves) *heavy_leaves = b->mesh_heavy[mesh];
    return SSDF_OK;
}

int ssdf_batch_autotune(ssdf_batch_t b, int mode, const ptrdiff_t *dims, const double *boxes,
                        const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream,
                        int *best_blocks_per_cu, int *best_units_per_grab, float *best_ms) {
    if (!b) {
        set_error("ssdf_batch_autotune: null batch");
        return SSDF_ERR_INVALID;
    }
    const hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    SSDF_HIP_TRY(hipEventCreate(&e0));
    SSDF_HIP_TRY(hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } events{e0, e1};
    // one untimed run, then a timed one (events on the stream)
    auto timed = [&](float &ms) -> int {
        for (int rep = 0; rep < 2; rep++) {
            if (rep == 1) (void)hipEventRecord(e0, st);
            if (int rc = ssdf_batch_query_grids(b, mode, dims, boxes, out_offsets, out_d, out_tri, flags, stream))
                return rc;
            if (rep == 1) (void)hipEventRecord(e1, st);
        }
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        return SSDF_OK;
    };
    // 1. per mesh, the brick shape of its grid waves: one query per candidate with every mesh at
    //    that candidate, each mesh's cost read from its waves' clocks (packet and per-query
    //    kernels); every mesh keeps its cheapest.  FP32 + fixup results do not depend on it.
    static const int shapes[] = {brick_code(2, 2, 2), brick_code(3, 2, 1), brick_code(2, 3, 1),
                                 brick_code(1, 2, 3), brick_code(3, 3, 0)};
    std::vector<unsigned long long> cost, best_cost(b->nm, ~0ull);
    std::vector<int> best_shape(b->nm, 0);
    std::fill(b->mesh_heavy.begin(), b->mesh_heavy.end(), -1);
    int rc = run_query(b, QSRC_GRID, mode, dims, boxes, out_offsets, nullptr, nullptr, nullptr, out_d, out_tri, nullptr,
                       nullptr, nullptr, flags, st);  // untimed first run (code objects, caches, workspace)
    for (int sh : shapes) {
        if (rc) break;
        std::fill(b->mesh_brick.begin(), b->mesh_brick.end(), sh);
        rc = timed_costs(b, mode, dims, boxes, out_offsets, out_d, out_tri, flags, st, cost);
        for (int i = 0; i < b->nm && !rc; i++)
            if (cost[i] < best_cost[i]) best_cost[i] = cost[i], best_shape[i] = sh;
    }
    if (rc) return rc;
    b->mesh_brick = best_shape;
    // 2. the leaf budget, by the launch's time: a mesh's summed wave cycles do not show the tail of
    //    its longest waves, which the budget trades against the per-query kernel's work
    static const int budgets[] = {4096, 8192, 16384};
    float best_b = 1e30f;
    int bh = -1;
    for (int hb : budgets) {
        std::fill(b->mesh_heavy.begin(), b->mesh_heavy.end(), hb);
        float ms = 0.0f;
        if ((rc = timed(ms))) return rc;
        if (ms < best_b) best_b = ms, bh = hb;
    }
    std::fill(b->mesh_heavy.begin(), b->mesh_heavy.end(), bh);
    // 3. the launch: blocks per CU x units per grab, timed by events with the per-mesh schedule
    const int bpcs[] = {3, 4, 5, 6, 8};
    const int chunks[] = {1, 2, 4, 8};
    float best = 1e30f;
    int bb = 0, bc = 0;
    for (int bpc : bpcs) {
        for (int ch : chunks) {
            b->blocks_per_cu = bpc, b->chunk = ch;
            float ms = 0.0f;
            if ((rc = timed(ms))) return rc;
            if (ms < best) best = ms, bb = bpc, bc = ch;
        }
    }
    // 4. each of the stages above picked on one timed run per candidate; the schedule is kept only
    //    if it beats the default (4x4x4 bricks, the process-wide leaf budget, 8 blocks per CU, one
    //    unit per grab) over three alternating runs each, so that run-to-run noise cannot leave the
    //    batch slower than it started
    const std::vector<int> tuned_brick = b->mesh_brick, tuned_heavy = b->mesh_heavy;
    auto set_tuned = [&]() {
        b->mesh_brick = tuned_brick, b->mesh_heavy = tuned_heavy;
        b->blocks_per_cu = bb, b->chunk = bc;
    };
    auto set_default = [&]() {
        std::fill(b->mesh_brick.begin(), b->mesh_brick.end(), brick_code(2, 2, 2));
        std::fill(b->mesh_heavy.begin(), b->mesh_heavy.end(), -1);
        b->blocks_per_cu = kDefaultBlocksPerCu, b->chunk = 1;
    };
    float t_tuned = 1e30f, t_default = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        float ms = 0.0f;
        set_default();
        if ((rc = timed(ms))) return rc;
        t_default = std::min(t_default, ms);
        set_tuned();
        if ((rc = timed(ms))) return rc;
        t_tuned = std::min(t_tuned, ms);
    }
    if (t_default <= t_tuned) {
        set_default();
        bb = kDefaultBlocksPerCu, bc = 1, best = t_default;
    } else {
        best = t_tuned;
    }
    if (best_blocks_per_cu) *best_blocks_per_cu = bb;
    if (best_units_per_grab) *best_units_per_grab = bc;
    if (best_ms) *best_ms = best;
    return SSDF_OK;
}

int ssdf_batch_capture(ssdf_batch_t b, int rebuild, int mode, const ptrdiff_t *dims, const double *boxes,
                       const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream) {
    if (!b || !dims || !boxes || !out_offsets || mode == SSDF_MODE_CLOSEST) {
        set_error("ssdf_batch_capture: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    const hipStream_t st = (hipStream_t)stream;
    if (!st) {
        set_error("ssdf_batch_capture: needs a created stream (the null stream cannot be captured)");
        return SSDF_ERR_INVALID;
    }
    if (mode == SSDF_MODE_SDF && !b->with_normals) {
        set_error("batch: signed queries need the meshes' pseudonormals");
        return SSDF_ERR_INVALID;
    }
    Plan p;
    if (int rc = plan_queries(b, QSRC_GRID, dims, boxes, out_offsets, p)) return rc;
    BatchArgs a;
    if (int rc = stage(b, p, QSRC_GRID, mode, flags, nullptr, nullptr, nullptr, out_d, out_tri, nullptr, nullptr,
                       nullptr, st, a))
        return rc;
    const int stack = stack_of(b);
    if (!stack) {
        set_error("batch: LBVH deeper than 128 levels");
        return SSDF_ERR_DEPTH;
    }
    // the previous graph (and any replay of it still running) goes before its buffers are reused
    SSDF_HIP_TRY(hipDeviceSynchronize());
    if (b->graph) (void)hipGraphExecDestroy(b->graph), b->graph = nullptr;
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def), b->graph_def = nullptr;
    // the graph's own descriptors (a copy of the ones stage just uploaded) and fixup queue
    if (!b->graph_desc) SSDF_HIP_TRY(dev_alloc((void **)&b->graph_desc, sizeof(BatchMesh) * b->nm, b->stream));
    if (b->graph_fix_cap < b->fix_cap) {
        dev_free(b->graph_fix, b->stream);
        b->graph_fix = nullptr, b->graph_fix_cap = 0;
        SSDF_HIP_TRY(dev_alloc((void **)&b->graph_fix, sizeof(unsigned) * 4 * b->fix_cap, b->stream));
        b->graph_fix_cap = b->fix_cap;
    }
    SSDF_HIP_TRY(hipStreamSynchronize(b->stream));
    SSDF_HIP_TRY(hipMemcpyAsync(b->graph_desc, b->d_desc, sizeof(BatchMesh) * p.desc.size(), hipMemcpyDeviceToDevice,
                                st));
    a.base.fix_list = b->graph_fix;
    a.base.fix_cap = (unsigned)std::min<size_t>(b->graph_fix_cap, 0xffffffffu);
    a.base.heavy_cap = a.base.fix_cap / 2;
    // the graph's own unit order, starting from the identity (replays reorder it for the next one)
    b->graph_order.key = 0;
    if (p.total_units > 0)
        if (int rc = prepare_order(b, b->graph_order, p, QSRC_GRID, st, a)) return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(st));  // descriptors and buffers in place before capture
    const bool fix = (flags & SSDF_FLAG_FIXUP) != 0;
    SSDF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = SSDF_OK;
    if (rebuild) rc = build_lanes(b, b->with_normals, st);  // workspaces sized since creation
    if (!rc && p.total_units > 0) rc = launch_batch(a, b->graph_desc, QSRC_GRID, mode, fix, stack,
                                                    grid_blocks(b, p.total_units, fix), st);
    if (!rc && p.total_units > 0) rc = reorder_units(b->graph_order, p.total_units, st);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SSDF_HIP_TRY(ce);
    b->graph_def = g;
    SSDF_HIP_TRY(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    return SSDF_OK;
}

int ssdf_batch_replay(ssdf_batch_t b, void *stream) {
    if (!b || !b->graph) {
        set_error("ssdf_batch_replay: nothing captured");
        return SSDF_ERR_INVALID;
    }
    SSDF_HIP_TRY(hipGraphLaunch(b->graph, (hipStream_t)stream));
    return SSDF_OK;
}

int ssdf_batch_last_stats(ssdf_batch_t b, unsigned *fixup_queued, int *kernel_error) {
    if (!b) {
        set_error("ssdf_batch_last_stats: null batch");
        return SSDF_ERR_INVALID;
    }
    unsigned h[2] = {0u, 0u};
    if (b->status) SSDF_HIP_TRY(hipMemcpy(h, b->status, sizeof(h), hipMemcpyDeviceToHost));
    if (fixup_queued) *fixup_queued = h[0];
    if (kernel_error) *kernel_error = (int)h[1];
    return SSDF_OK;
}

int ssdf_batch_info(ssdf_batch_t b, int mesh, ssdf_mesh_info_t *info) {
    if (!b || !info || mesh < -1 || mesh >= b->nm) {
        set_error("ssdf_batch_info: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    std::memset(info, 0, sizeof(*info));
    info->dtype = 0;
    if (mesh == -1) {  // the whole batch
        for (auto &m : b->meshes) info->nselements += m.nsel, info->nspoints += m.nsp;
        info->maxdepth = b->maxdepth;
        info->open_edges = -1;
        info->device_bytes = b->device_bytes;
        return SSDF_OK;
    }
    const DevMesh &m = b->meshes[mesh];
    info->nselements = m.nsel;
    info->nspoints = m.nsp;
    info->maxdepth = m.hmeta.maxdepth;
    info->open_edges = m.has_normals ? m.hmeta.open_edges : -1;
    info->flipped_edges = m.hmeta.flipped_edges;
    for (int d = 0; d < 6; d++) info->bbox[d] = ordered_to_float(m.hmeta.bbox_ord[d]);
    info->device_bytes = 0;
    return SSDF_OK;
}

}  // extern "C"

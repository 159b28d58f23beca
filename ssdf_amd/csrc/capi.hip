// capi.hip — the C ABI (include/ssdf_c.h): reference symbols, resident meshes, device plumbing.
#include "query.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>
#include <string>

using namespace ssdf_amd;

namespace {

// FP32 error bound of a computed distance, relative to the largest coordinate magnitude involved:
// 32 ulps of 1.0 (a few roundings of the projection and of p - q, with ample margin).

// Query workspace a handle keeps between calls (grown on demand, stream-ordered).
struct Workspace {
    unsigned *status = nullptr;  // kStatusBytes: status words and the wave-queue counters (query.hpp)
    unsigned *list = nullptr;
    size_t cap = 0;  // words of list (4 per queue entry)
    void *sort = nullptr;  // Morton sort of scattered query points
    size_t sort_cap = 0;
};

__global__ void k_reset_status(unsigned *status) {
    if (threadIdx.x < 4) status[threadIdx.x] = 0u;
    if (threadIdx.x >= 4 && threadIdx.x < 12) status[kWaveQueueWord * (threadIdx.x - 3)] = 0u;
}

int ws_reserve(Workspace &w, size_t n, hipStream_t stream) {
    if (!w.status) SSDF_HIP_TRY(dev_alloc((void **)&w.status, kStatusBytes, stream));
    if (w.cap < n) {
        dev_free(w.list, stream);
        w.list = nullptr;
        w.cap = 0;
        SSDF_HIP_TRY(dev_alloc((void **)&w.list, sizeof(unsigned) * std::max<size_t>(n, 1), stream));
        w.cap = n;
    }
    return SSDF_OK;
}

void ws_free(Workspace &w, hipStream_t stream) {
    dev_free(w.status, stream);
    dev_free(w.list, stream);
    dev_free(w.sort, stream);
    w = Workspace{};
}

int ws_sort_reserve(Workspace &w, size_t bytes, hipStream_t stream) {
    if (w.sort_cap >= bytes) return SSDF_OK;
    dev_free(w.sort, stream);
    w.sort = nullptr;
    w.sort_cap = 0;
    SSDF_HIP_TRY(dev_alloc(&w.sort, bytes, stream));
    w.sort_cap = bytes;
    return SSDF_OK;
}

// Exact FP64 argmin over every triangle (tiles of 256 TriF headers in LDS).
__global__ __launch_bounds__(256) void k_brute_f64(const TriP *__restrict__ tris, int ntri,
                                                   const float *__restrict__ x, const float *__restrict__ y,
                                                   const float *__restrict__ z, long long npoints,
                                                   double *__restrict__ out_d2, int *__restrict__ out_tri) {
    __shared__ TriF tile[256];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = p < npoints;
    const double px = ok ? (double)x[p] : 0.0, py = ok ? (double)y[p] : 0.0, pz = ok ? (double)z[p] : 0.0;
    double best = __builtin_inf();
    int bt = 0x7fffffff;
    for (int base = 0; base < ntri; base += 256) {
        __syncthreads();
        if (base + (int)threadIdx.x < ntri) tile[threadIdx.x] = tris[base + threadIdx.x].f;
        __syncthreads();
        const int m = std::min(256, ntri - base);
        for (int i = 0; i < m; i++) {
            const TriF &tr = tile[i];
            const double d = pt_tri_d2<double>(px, py, pz, tr.a[0], tr.a[1], tr.a[2], tr.b[0], tr.b[1], tr.b[2],
                                               tr.c[0], tr.c[1], tr.c[2]);
            if (d < best || (d == best && tr.orig < bt)) best = d, bt = tr.orig;
        }
    }
    if (ok) {
        out_d2[p] = best;
        out_tri[p] = bt == 0x7fffffff ? -1 : bt;
    }
}

// Original triangle index of each LBVH leaf, in leaf (Morton) order.
__global__ __launch_bounds__(256) void k_leaf_order(const TriP *__restrict__ tris, int ntri, int *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ntri) out[k] = tris[k].f.orig;
}

// Plain streaming copy (16 B per lane per step): the reference byte count for calibrating the
// HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename G>
int upload_mesh(DevMesh &m, ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp,
                const G *sx, const G *sy, const G *sz, bool device_ptrs, hipStream_t stream) {
    m.dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    m.stream = stream;
    m.nsel = nsel;
    m.nsp = nsp;
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const size_t ib = sizeof(int) * (size_t)nsel, vb = sizeof(G) * (size_t)nsp;
    int **idx[3] = {&m.s0, &m.s1, &m.s2};
    const int *isrc[3] = {s0, s1, s2};
    for (int k = 0; k < 3; k++) {
        SSDF_HIP_TRY(dev_alloc((void **)idx[k], std::max<size_t>(ib, 4), stream));
        SSDF_HIP_TRY(hipMemcpyAsync(*idx[k], isrc[k], ib, kind, stream));
    }
    void **vtx[3] = {&m.sx, &m.sy, &m.sz};
    const G *vsrc[3] = {sx, sy, sz};
    for (int k = 0; k < 3; k++) {
        SSDF_HIP_TRY(dev_alloc(vtx[k], std::max<size_t>(vb, 8), stream));
        SSDF_HIP_TRY(hipMemcpyAsync(*vtx[k], vsrc[k], vb, kind, stream));
    }
    m.owns_input = true;
    m.device_bytes += 3 * (ib + vb);
    return SSDF_OK;
}

}  // namespace

struct ssdf_mesh {
    DevMesh m;
    Workspace ws;
    hipStream_t last_stream = nullptr;  // stream of the last query (its workspace use)
    bool queried = false;
    hipEvent_t order_ev = nullptr;  // orders one stream after another (stream_after)
    int brick = 0;                  // brick shape of grid waves (brick_code; 0: 4 x 4 x 4)
    int heavy = -1;                 // leaf budget of packet waves (-1: the process-wide one)
};

namespace {
std::atomic<int> &heavy_setting() {
    static std::atomic<int> v{[] {
        const char *e = std::getenv("SSDF_HEAVY_LEAVES");
        if (!e || !*e) return kHeavyLeaves;
        const long x = std::strtol(e, nullptr, 10);
        return (int)std::max(0L, std::min(x, 1L << 30));
    }()};
    return v;
}
}  // namespace

int ssdf_amd::heavy_leaves() { return heavy_setting().load(std::memory_order_relaxed); }

namespace {

// Makes `st` wait for the work queued so far on `other` (nothing when they are one stream).
int stream_after(ssdf_mesh *h, hipStream_t st, hipStream_t other) {
    if (st == other) return SSDF_OK;
    if (!h->order_ev) SSDF_HIP_TRY(hipEventCreateWithFlags(&h->order_ev, hipEventDisableTiming));
    SSDF_HIP_TRY(hipEventRecord(h->order_ev, other));
    SSDF_HIP_TRY(hipStreamWaitEvent(st, h->order_ev, 0));
    return SSDF_OK;
}

// Orders `st` after every use of the handle's device buffers so far: its builds (the handle's
// stream) and its last query (whose kernels read the tree and the triangles).
int order_after_handle(ssdf_mesh *h, hipStream_t st) {
    if (int rc = stream_after(h, st, h->m.stream)) return rc;
    if (h->queried)
        if (int rc = stream_after(h, st, h->last_stream)) return rc;
    return SSDF_OK;
}

// Fills the mesh-independent launch arguments of a query over nq queries working in `w`, and clears
// w's status words (a kernel, so that the sequence stays capturable).
void fill_args(const DevMesh &m, QueryArgs &a, const QueryWs &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, stream, w.status);
    a.fix_count = w.status;
    a.err_flag = (int *)(w.status + 1);
    a.wqueue = w.status + kWaveQueueWord;
    a.fix_list = w.list;
    a.fix_cap = (unsigned)std::min<size_t>(w.entries, 0xffffffffu);
    a.heavy_iters = heavy_leaves();
    a.heavy_cap = a.fix_cap / 2;
    const float L = meta_scale(m.hmeta);
    a.tau = kTau;
    a.band_abs = kBandRel * L;
    a.sign_eps = (float)(kSignEpsRel * meta_diag(m.hmeta));
    a.delta64 = kDelta64Rel * L;
}

// The handle's own workspace for a query of nq queries, ordered after its last use on another
// stream (calls on one handle share it).
int handle_ws(ssdf_mesh *h, size_t nq, bool sorted, hipStream_t stream, QueryWs &w) {
    // after the handle's last build (it may have run on another stream) ...
    if (int rc = stream_after(h, stream, h->m.stream)) return rc;
    // ... and after the last query that used this workspace
    if (h->queried && h->last_stream != stream && h->ws.list)
        if (int rc = stream_after(h, stream, h->last_stream)) return rc;
    h->queried = true;
    h->last_stream = stream;
    if (int rc = ws_reserve(h->ws, 4 * fixup_entries(nq), stream)) return rc;
    if (sorted)
        if (int rc = ws_sort_reserve(h->ws, sort_ws_bytes((long long)nq), stream)) return rc;
    w.status = h->ws.status;
    w.list = h->ws.list;
    w.entries = h->ws.cap / 4;
    w.sort = sorted ? h->ws.sort : nullptr;
    w.sort_bytes = sorted ? h->ws.sort_cap : 0;
    return SSDF_OK;
}

int create_mesh(int dtype, ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp,
                const void *sx, const void *sy, const void *sz, bool device_ptrs, bool with_normals,
                hipStream_t stream, ssdf_mesh **out) {
    *out = nullptr;
    if (nsel < 0 || nsp < 0 || (nsel > 0 && (!s0 || !s1 || !s2 || !sx || !sy || !sz))) {
        set_error("ssdf_mesh_create: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    ssdf_mesh *h = new (std::nothrow) ssdf_mesh();
    if (!h) return SSDF_ERR_OOM;
    // the ordering event exists from the start, so that ordering a call (stream_after) never
    // allocates: the allocation-free query orders itself too
    if (hipEventCreateWithFlags(&h->order_ev, hipEventDisableTiming) != hipSuccess) {
        delete h;
        set_error("ssdf_mesh_create: hipEventCreateWithFlags failed");
        return SSDF_ERR_HIP;
    }
    int rc = dtype == 0 ? upload_mesh<float>(h->m, nsel, s0, s1, s2, nsp, (const float *)sx, (const float *)sy,
                                             (const float *)sz, device_ptrs, stream)
                        : upload_mesh<double>(h->m, nsel, s0, s1, s2, nsp, (const double *)sx, (const double *)sy,
                                              (const double *)sz, device_ptrs, stream);
    if (!rc && nsel > 0) rc = build_lbvh(h->m, stream);
    if (!rc && nsel > 0 && with_normals) rc = build_normals(h->m, stream);
    if (!rc && nsel > 0) rc = sync_meta(h->m, stream);
    if (rc) {
        free_mesh(h->m, stream);
        (void)hipEventDestroy(h->order_ev);
        delete h;
        return rc;
    }
    *out = h;
    return SSDF_OK;
}

// streams_alive: the creation and last-query streams still exist (the library's own callers), so
// the frees are ordered behind the handle's work on them.  Otherwise (the public destroy: the
// caller may have destroyed its streams already) the device is synchronized and the frees go to
// the null stream.
void destroy_mesh(ssdf_mesh *h, bool streams_alive) {
    if (!h) return;
    hipStream_t st = nullptr;
    if (streams_alive) {
        // frees are ordered on the build stream: make it wait for the last query's stream
        st = h->m.stream;
        if (h->queried) (void)stream_after(h, st, h->last_stream);
    } else {
        (void)hipDeviceSynchronize();
    }
    ws_free(h->ws, st);
    free_mesh(h->m, st);
    if (h->order_ev) (void)hipEventDestroy(h->order_ev);
    delete h;
}

// Grid nodes x_i = min + i h with h = (max - min) / (n - 1) in the mesh precision
// (serial/bvh/bvh.cpp:1665-1667, 1744).
void setup_grid(const DevMesh &m, QueryArgs &a, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, const double *lo,
                const double *hi) {
    a.nx = (int)nx, a.ny = (int)ny, a.nz = (int)nz;
    const ptrdiff_t ns[3] = {nx, ny, nz};
    for (int d = 0; d < 3; d++) {
        if (m.dtype == 0) {
            const float l = (float)lo[d], u = (float)hi[d];
            a.gmin[d] = l;
            a.gh[d] = ns[d] > 1 ? (double)((u - l) / (float)(ns[d] - 1)) : 0.0;
        } else {
            a.gmin[d] = lo[d];
            a.gh[d] = ns[d] > 1 ? (hi[d] - lo[d]) / (double)(ns[d] - 1) : 0.0;
        }
    }
}

// Signs from winding numbers: always with SSDF_FLAG_SIGN_GWN; with SSDF_FLAG_SIGN_AUTO when the
// edge pass found open or non-manifold (same-direction) edges, where pseudonormals lose meaning.
bool gwn_sign_wanted(const DevMesh &m, unsigned flags) {
    if (flags & SSDF_FLAG_SIGN_GWN) return true;
    return (flags & SSDF_FLAG_SIGN_AUTO) && (m.hmeta.open_edges > 0 || m.hmeta.flipped_edges > 0);
}

__global__ void k_fill_empty(size_t n, float *w, unsigned char *in) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (w) w[i] = 0.0f;
        if (in) in[i] = 0;
    }
}

// An empty surface encloses nothing: w = 0, outside everywhere.
int fill_empty(int mode, size_t n, float *out_w, unsigned char *out_in, hipStream_t stream) {
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_empty, dim3(nb), dim3(256), 0, stream, n, mode == GW_WINDING ? out_w : nullptr,
                       mode == GW_INSIDE ? out_in : nullptr);
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

bool points_sorted(const DevMesh &m, ptrdiff_t n) { return m.dtype == 0 && n >= kSortMin; }

// A point query of mode `mode` on mesh m, working in w (which needs the sort part when
// points_sorted(m, n)).
int query_points_in(DevMesh &m, const QueryWs &w, int mode, ptrdiff_t n, const void *x, const void *y,
                    const void *z, void *out_d, int *out_tri, void *cx, void *cy, void *cz, unsigned flags,
                    hipStream_t stream) {
    QueryArgs a{};
    fill_args(m, a, w, stream);
    a.qx = x, a.qy = y, a.qz = z;
    a.npoints = n;
    a.out_d = out_d;
    a.out_tri = out_tri;
    a.out_cx = cx, a.out_cy = cy, a.out_cz = cz;
    const int dbl = m.dtype;
    if (points_sorted(m, n))
        if (int rc = sort_points(m.meta, n, (const float *)x, (const float *)y, (const float *)z, w.sort, &a.perm,
                                 stream))
            return rc;
    if (int rc = launch_query(m, a, QSRC_POINTS, mode, flags, dbl, dbl, stream)) return rc;
    if (mode == SSDF_MODE_SDF && gwn_sign_wanted(m, flags)) return launch_gwn(m, a, QSRC_POINTS, GW_SIGN, dbl, stream);
    return SSDF_OK;
}

int check_points_args(int mode, const void *x, const void *y, const void *z, const void *out_d, const void *cx,
                      const void *cy, const void *cz) {
    if (!x || !y || !z || !out_d || (mode == SSDF_MODE_CLOSEST && (!cx || !cy || !cz))) {
        set_error("query_points: null argument");
        return SSDF_ERR_INVALID;
    }
    return SSDF_OK;
}

int query_points(ssdf_mesh *h, int mode, ptrdiff_t n, const void *x, const void *y, const void *z, void *out_d,
                 int *out_tri, void *cx, void *cy, void *cz, unsigned flags, hipStream_t stream) {
    if (n <= 0 || h->m.nsel <= 0) return SSDF_OK;
    if (int rc = check_points_args(mode, x, y, z, out_d, cx, cy, cz)) return rc;
    QueryWs w;
    if (int rc = handle_ws(h, (size_t)n, points_sorted(h->m, n), stream, w)) return rc;
    return query_points_in(h->m, w, mode, n, x, y, z, out_d, out_tri, cx, cy, cz, flags, stream);
}

int query_grid(ssdf_mesh *h, int mode, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, const double *lo, const double *hi,
               void *out_d, int *out_tri, unsigned flags, hipStream_t stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || h->m.nsel <= 0) return SSDF_OK;
    if (nx > 0x7fffffff || ny > 0x7fffffff || nz > 0x7fffffff || !out_d) {
        set_error("query_grid: invalid grid");
        return SSDF_ERR_INVALID;
    }
    const size_t nq = (size_t)nx * (size_t)ny * (size_t)nz;
    QueryWs w;
    if (int rc = handle_ws(h, nq, false, stream, w)) return rc;
    QueryArgs a{};
    fill_args(h->m, a, w, stream);
    setup_grid(h->m, a, nx, ny, nz, lo, hi);
    a.brick = h->brick;
    if (h->heavy >= 0) a.heavy_iters = h->heavy;
    a.out_d = out_d;
    a.out_tri = out_tri;
    const int dbl = h->m.dtype;
    if (int rc = launch_query(h->m, a, QSRC_GRID, mode, flags, dbl, dbl, stream)) return rc;
    if (mode == SSDF_MODE_SDF && gwn_sign_wanted(h->m, flags))
        return launch_gwn(h->m, a, QSRC_GRID, GW_SIGN, dbl, stream);
    return SSDF_OK;
}

// Winding numbers (GW_WINDING into out_w) or inside flags (GW_INSIDE into out_in) at points.
int gwn_points(ssdf_mesh *h, int mode, ptrdiff_t n, const void *x, const void *y, const void *z, float *out_w,
               unsigned char *out_in, hipStream_t stream) {
    if (n <= 0) return SSDF_OK;
    if (!x || !y || !z || (mode == GW_WINDING ? !out_w : !out_in)) {
        set_error("winding-number query: null argument");
        return SSDF_ERR_INVALID;
    }
    if (h->m.nsel <= 0) return fill_empty(mode, (size_t)n, out_w, out_in, stream);
    QueryWs w;
    if (int rc = handle_ws(h, (size_t)n, points_sorted(h->m, n), stream, w)) return rc;
    QueryArgs a{};
    fill_args(h->m, a, w, stream);
    a.qx = x, a.qy = y, a.qz = z;
    a.npoints = n;
    a.out_w = out_w;
    a.out_in = out_in;
    if (points_sorted(h->m, n))
        if (int rc = sort_points(h->m.meta, n, (const float *)x, (const float *)y, (const float *)z, w.sort, &a.perm,
                                 stream))
            return rc;
    return launch_gwn(h->m, a, QSRC_POINTS, mode, h->m.dtype, stream);
}

int gwn_grid(ssdf_mesh *h, int mode, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, const double *lo, const double *hi,
             float *out_w, unsigned char *out_in, hipStream_t stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return SSDF_OK;
    if (nx > 0x7fffffff || ny > 0x7fffffff || nz > 0x7fffffff || (mode == GW_WINDING ? !out_w : !out_in)) {
        set_error("winding-number grid: invalid grid or null output");
        return SSDF_ERR_INVALID;
    }
    const size_t nq = (size_t)nx * (size_t)ny * (size_t)nz;
    if (h->m.nsel <= 0) return fill_empty(mode, nq, out_w, out_in, stream);
    QueryWs w;
    if (int rc = handle_ws(h, nq, false, stream, w)) return rc;
    QueryArgs a{};
    fill_args(h->m, a, w, stream);
    setup_grid(h->m, a, nx, ny, nz, lo, hi);
    a.out_w = out_w;
    a.out_in = out_in;
    return launch_gwn(h->m, a, QSRC_GRID, mode, h->m.dtype, stream);
}

// RAII holder of device buffers for the synchronous host-pointer entry points.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

template <typename G>
int h2d(DevBuf &b, const G *src, ptrdiff_t n) {
    SSDF_HIP_TRY(hipMalloc(&b.p, sizeof(G) * (size_t)std::max<ptrdiff_t>(n, 1)));
    if (src && n > 0) SSDF_HIP_TRY(hipMemcpy(b.p, src, sizeof(G) * (size_t)n, hipMemcpyHostToDevice));
    return SSDF_OK;
}

template <typename G>
int d2h(G *dst, const DevBuf &b, ptrdiff_t n) {
    if (dst && n > 0) SSDF_HIP_TRY(hipMemcpy(dst, b.p, sizeof(G) * (size_t)n, hipMemcpyDeviceToHost));
    return SSDF_OK;
}

// The reference's per-call GPU structure (serial/bvh/bvh.cu:80-158): allocate, upload, build,
// query, download, free.  Results are in host memory on return.
template <typename G>
int host_points_call(int mode, ptrdiff_t npoints, const G *x, const G *y, const G *z, ptrdiff_t nsel, const int *s0,
                     const int *s1, const int *s2, ptrdiff_t nsp, const G *sx, const G *sy, const G *sz, G *outd,
                     G *outx, G *outy, G *outz, int *outtri, unsigned flags) {
    if (nsel == 0 || nsp == 0 || npoints <= 0) return SSDF_OK;  // serial/ssdf.hpp:462
    if (!x || !y || !z || !outd) {
        set_error("null point or output array");
        return SSDF_ERR_INVALID;
    }
    hipStream_t stream = nullptr;
    ssdf_mesh *h = nullptr;
    const int dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    if (int rc = create_mesh(dtype, nsel, s0, s1, s2, nsp, sx, sy, sz, false, mode == SSDF_MODE_SDF, stream, &h))
        return rc;
    struct Guard {
        ssdf_mesh *h;
        ~Guard() { destroy_mesh(h, true); }
    } guard{h};
    DevBuf dx, dy, dz, dd, dt, dcx, dcy, dcz;
    if (int rc = h2d(dx, x, npoints)) return rc;
    if (int rc = h2d(dy, y, npoints)) return rc;
    if (int rc = h2d(dz, z, npoints)) return rc;
    if (int rc = h2d(dd, mode == SSDF_MODE_SDF ? (const G *)nullptr : outd, npoints)) return rc;
    if (outtri) {
        if (int rc = h2d(dt, (const int *)nullptr, npoints)) return rc;
        if (mode == SSDF_MODE_CLOSEST) SSDF_HIP_TRY(hipMemcpy(dt.p, outtri, sizeof(int) * npoints, hipMemcpyHostToDevice));
    }
    if (mode == SSDF_MODE_CLOSEST) {
        if (int rc = h2d(dcx, outx, npoints)) return rc;
        if (int rc = h2d(dcy, outy, npoints)) return rc;
        if (int rc = h2d(dcz, outz, npoints)) return rc;
    }
    if (int rc = query_points(h, mode, npoints, dx.p, dy.p, dz.p, dd.p, (int *)dt.p, dcx.p, dcy.p, dcz.p, flags,
                              stream))
        return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    if (int rc = d2h(outd, dd, npoints)) return rc;
    if (outtri)
        if (int rc = d2h(outtri, dt, npoints)) return rc;
    if (mode == SSDF_MODE_CLOSEST) {
        if (int rc = d2h(outx, dcx, npoints)) return rc;
        if (int rc = d2h(outy, dcy, npoints)) return rc;
        if (int rc = d2h(outz, dcz, npoints)) return rc;
    }
    return SSDF_OK;
}

template <typename G>
int host_grid_call(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const G *sx,
                   const G *sy, const G *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, G xmin, G ymin, G zmin, G xmax,
                   G ymax, G zmax, G *out, int *out_tri, unsigned flags) {
    if (nsel == 0 || nsp == 0 || nx <= 0 || ny <= 0 || nz <= 0) return SSDF_OK;
    if (!out) {
        set_error("null output array");
        return SSDF_ERR_INVALID;
    }
    hipStream_t stream = nullptr;
    ssdf_mesh *h = nullptr;
    const int dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    if (int rc = create_mesh(dtype, nsel, s0, s1, s2, nsp, sx, sy, sz, false, true, stream, &h)) return rc;
    struct Guard {
        ssdf_mesh *h;
        ~Guard() { destroy_mesh(h, true); }
    } guard{h};
    const ptrdiff_t nq = nx * ny * nz;
    DevBuf dd, dt;
    if (int rc = h2d(dd, (const G *)nullptr, nq)) return rc;
    if (out_tri)
        if (int rc = h2d(dt, (const int *)nullptr, nq)) return rc;
    const double lo[3] = {(double)xmin, (double)ymin, (double)zmin}, hi[3] = {(double)xmax, (double)ymax, (double)zmax};
    if (int rc = query_grid(h, SSDF_MODE_SDF, nx, ny, nz, lo, hi, dd.p, (int *)dt.p, flags, stream)) return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    if (int rc = d2h(out, dd, nq)) return rc;
    if (out_tri)
        if (int rc = d2h(out_tri, dt, nq)) return rc;
    return SSDF_OK;
}

// sample_sqedf_bvh (serial/bvh/bvh.cpp:1575-1639): out (in/out, squared distances) = min(out, d^2)
// on the grid nodes; synchronous.
template <typename G>
int host_sqedf_call(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const G *sx,
                    const G *sy, const G *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, G xmin, G ymin, G zmin,
                    G xmax, G ymax, G zmax, G *out, unsigned flags) {
    if (nsel == 0 || nsp == 0 || nx <= 0 || ny <= 0 || nz <= 0) return SSDF_OK;
    if (!out) {
        set_error("sample_sqedf: null output array");
        return SSDF_ERR_INVALID;
    }
    const ptrdiff_t nq = nx * ny * nz;
    hipStream_t stream = nullptr;
    ssdf_mesh *h = nullptr;
    const int dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    if (int rc = create_mesh(dtype, nsel, s0, s1, s2, nsp, sx, sy, sz, false, false, stream, &h)) return rc;
    struct Guard {
        ssdf_mesh *h;
        ~Guard() { destroy_mesh(h, true); }
    } guard{h};
    DevBuf dd;
    if (int rc = h2d(dd, out, nq)) return rc;
    const double lo[3] = {(double)xmin, (double)ymin, (double)zmin}, hi[3] = {(double)xmax, (double)ymax, (double)zmax};
    if (int rc = query_grid(h, SSDF_MODE_SQEDF, nx, ny, nz, lo, hi, dd.p, nullptr, flags, stream)) return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    return d2h(out, dd, nq);
}

// points_inside_bvh (serial/bvh/bvh.cpp:318-365): out[i] = 1 inside, 0 outside; synchronous.
template <typename G>
int host_inside_points(ptrdiff_t npoints, const G *x, const G *y, const G *z, ptrdiff_t nsel, const int *s0,
                       const int *s1, const int *s2, ptrdiff_t nsp, const G *sx, const G *sy, const G *sz,
                       uint8_t *out) {
    if (npoints <= 0) return SSDF_OK;
    if (!x || !y || !z || !out) {
        set_error("points_inside: null point or output array");
        return SSDF_ERR_INVALID;
    }
    if (nsel == 0 || nsp == 0) {
        std::memset(out, 0, (size_t)npoints);
        return SSDF_OK;
    }
    hipStream_t stream = nullptr;
    ssdf_mesh *h = nullptr;
    const int dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    if (int rc = create_mesh(dtype, nsel, s0, s1, s2, nsp, sx, sy, sz, false, false, stream, &h)) return rc;
    struct Guard {
        ssdf_mesh *h;
        ~Guard() { destroy_mesh(h, true); }
    } guard{h};
    DevBuf dx, dy, dz, dout;
    if (int rc = h2d(dx, x, npoints)) return rc;
    if (int rc = h2d(dy, y, npoints)) return rc;
    if (int rc = h2d(dz, z, npoints)) return rc;
    if (int rc = h2d(dout, (const uint8_t *)nullptr, npoints)) return rc;
    if (int rc = gwn_points(h, GW_INSIDE, npoints, dx.p, dy.p, dz.p, nullptr, (unsigned char *)dout.p, stream))
        return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    return d2h(out, dout, npoints);
}

// sample_inside_bvh (serial/bvh/bvh.cpp:1800-1866): grid nodes x_i = min + i h, x fastest.
template <typename G>
int host_inside_grid(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const G *sx,
                     const G *sy, const G *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, G xmin, G ymin, G zmin,
                     G xmax, G ymax, G zmax, uint8_t *out) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return SSDF_OK;
    if (!out) {
        set_error("sample_inside: null output array");
        return SSDF_ERR_INVALID;
    }
    const size_t nq = (size_t)nx * (size_t)ny * (size_t)nz;
    if (nsel == 0 || nsp == 0) {
        std::memset(out, 0, nq);
        return SSDF_OK;
    }
    hipStream_t stream = nullptr;
    ssdf_mesh *h = nullptr;
    const int dtype = sizeof(G) == sizeof(double) ? 1 : 0;
    if (int rc = create_mesh(dtype, nsel, s0, s1, s2, nsp, sx, sy, sz, false, false, stream, &h)) return rc;
    struct Guard {
        ssdf_mesh *h;
        ~Guard() { destroy_mesh(h, true); }
    } guard{h};
    DevBuf dout;
    if (int rc = h2d(dout, (const uint8_t *)nullptr, (ptrdiff_t)nq)) return rc;
    const double lo[3] = {(double)xmin, (double)ymin, (double)zmin}, hi[3] = {(double)xmax, (double)ymax, (double)zmax};
    if (int rc = gwn_grid(h, GW_INSIDE, nx, ny, nz, lo, hi, nullptr, (unsigned char *)dout.p, stream)) return rc;
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    return d2h(out, dout, (ptrdiff_t)nq);
}

// SSDF_CELL_VALIDATE (serial/ssdf.hpp:979-1010), read on every call as the reference reads it: the
// engine's result is checked against the exact FP64 argmin of the same inputs, and the maximum and
// the norm of the relative differences are printed to stderr when the maximum exceeds 1e-5; the
// engine's result is returned.
int cell_validate() {
    const char *e = std::getenv("SSDF_CELL_VALIDATE");
    return e ? std::atoi(e) : 0;
}

template <typename T>
int edf_validated(ptrdiff_t npoints, const T *x, const T *y, const T *z, ptrdiff_t nsel, const int *s0,
                  const int *s1, const int *s2, ptrdiff_t nsp, const T *sx, const T *sy, const T *sz, T *out) {
    if (npoints <= 0 || !out) return host_points_call<T>(SSDF_MODE_EDF, npoints, x, y, z, nsel, s0, s1, s2, nsp, sx, sy,
                                                         sz, out, nullptr, nullptr, nullptr, nullptr, 0u);
    std::vector<T> test(out, out + npoints);
    if (int rc = host_points_call<T>(SSDF_MODE_EDF, npoints, x, y, z, nsel, s0, s1, s2, nsp, sx, sy, sz, test.data(),
                                     nullptr, nullptr, nullptr, nullptr, 0u))
        return rc;
    // the exact FP64 search on the same values widened to FP64, seeded with the caller's out
    std::vector<double> X(x, x + npoints), Y(y, y + npoints), Z(z, z + npoints);
    std::vector<double> SX(sx, sx + (nsp > 0 ? nsp : 0)), SY(sy, sy + (nsp > 0 ? nsp : 0)),
        SZ(sz, sz + (nsp > 0 ? nsp : 0));
    std::vector<double> ref(out, out + npoints);
    if (int rc = host_points_call<double>(SSDF_MODE_EDF, npoints, X.data(), Y.data(), Z.data(), nsel, s0, s1, s2, nsp,
                                          SX.data(), SY.data(), SZ.data(), ref.data(), nullptr, nullptr, nullptr,
                                          nullptr, 0u))
        return rc;
    T max_diff = 0, diff_norm = 0;
    for (ptrdiff_t i = 0; i < npoints; ++i) {
        const T r = (T)ref[i];
        const T rel = std::abs(r - test[i]) / (std::abs(r) + (T)1e-16);
        max_diff = std::max<T>(max_diff, rel);
        diff_norm += rel * rel;
    }
    if (max_diff > (T)1e-5) {
        std::fprintf(stderr, "EDF cell-list validation max rel diff: %g\n", (double)max_diff);
        std::fprintf(stderr, "EDF cell-list validation rel diff norm: %g\n", (double)std::sqrt(diff_norm));
    }
    std::copy(test.begin(), test.end(), out);
    return SSDF_OK;
}

}  // namespace

namespace ssdf_amd {
// Destroys a mesh of server_mesh_create, its frees ordered on the server's own (live) streams.
void server_mesh_destroy(ssdf_mesh_t h) { destroy_mesh(h, true); }

// Resident FP32 mesh with normals from host arrays, for the serving loop (server.hip).
int server_mesh_create(const int *s0, const int *s1, const int *s2, ptrdiff_t nsel, const float *sx,
                       const float *sy, const float *sz, ptrdiff_t nsp, hipStream_t stream, ssdf_mesh_t *out,
                       size_t *bytes) {
    ssdf_mesh *h = nullptr;
    if (int rc = create_mesh(0, nsel, s0, s1, s2, nsp, sx, sy, sz, false, true, stream, &h)) return rc;
    *out = h;
    *bytes = h->m.device_bytes + h->m.ws_cap;
    return SSDF_OK;
}

// A point query in a caller-provided contiguous workspace `ws` (256-byte aligned,
// query_ws_bytes(n, n >= kSortMin) bytes) instead of the handle's: calls on one mesh then share
// only its read-only tree and may run concurrently on different streams.  The winding-number tree
// that SSDF_FLAG_SIGN_GWN / AUTO need is built here (allocating, synchronous) when may_build_gwn,
// else its absence is an error.
int ws_query_points(ssdf_mesh_t h, void *ws, bool handle_call, int mode, ptrdiff_t n, const void *x,
                    const void *y, const void *z, void *out_d, int *out_tri, void *cx, void *cy, void *cz,
                    unsigned flags, hipStream_t stream) {
    if (n <= 0 || h->m.nsel <= 0) return SSDF_OK;
    if (int rc = check_points_args(mode, x, y, z, out_d, cx, cy, cz)) return rc;
    if (handle_call) {
        // ordered like the handle's other calls: after its builds and its last call, and recorded
        // as its last call, which its rebuilds and its destroy wait for
        if (int rc = order_after_handle(h, stream)) return rc;
        h->queried = true;
        h->last_stream = stream;
    }
    if (mode == SSDF_MODE_SDF && gwn_sign_wanted(h->m, flags) && !h->m.has_gwn) {
        if (handle_call) {
            set_error("the mesh's winding-number tree is not built (run one allocating winding-number or "
                      "signed query on it first)");
            return SSDF_ERR_INVALID;
        }
        // built once, on the mesh's own stream, and completed before any call (on any stream) uses it
        if (int rc = build_gwn(h->m, h->m.stream)) return rc;
        SSDF_HIP_TRY(hipStreamSynchronize(h->m.stream));
    }
    const QueryWs w = carve_ws(ws, (size_t)n, points_sorted(h->m, n));
    return query_points_in(h->m, w, mode, n, x, y, z, out_d, out_tri, cx, cy, cz, flags, stream);
}
}  // namespace ssdf_amd

extern "C" {

int ssdf_sample_sqedf_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                        const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                        float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, float *out,
                        unsigned flags) {
    return host_sqedf_call<float>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                  ymax, zmax, out, flags);
}

int ssdf_sample_sqedf_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                        const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                        ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                        double *out) {
    return host_sqedf_call<double>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                   ymax, zmax, out, 0u);
}

int ssdf_points_inside_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx,
                         const float *sy, const float *sz, uint8_t *out) {
    return host_inside_points<float>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

int ssdf_points_inside_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const double *sx,
                         const double *sy, const double *sz, uint8_t *out) {
    return host_inside_points<double>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

int ssdf_sample_inside_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                         float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, uint8_t *out) {
    return host_inside_grid<float>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                   ymax, zmax, out);
}

int ssdf_sample_inside_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                         ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                         uint8_t *out) {
    return host_inside_grid<double>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                    ymax, zmax, out);
}

int ssdf_mesh_winding_number(ssdf_mesh_t mesh, ptrdiff_t npoints, const void *x, const void *y, const void *z,
                             float *out_w, void *stream) {
    if (!mesh) {
        set_error("ssdf_mesh_winding_number: null mesh");
        return SSDF_ERR_INVALID;
    }
    return gwn_points(mesh, GW_WINDING, npoints, x, y, z, out_w, nullptr, (hipStream_t)stream);
}

int ssdf_mesh_inside_points(ssdf_mesh_t mesh, ptrdiff_t npoints, const void *x, const void *y, const void *z,
                            uint8_t *out_inside, void *stream) {
    if (!mesh) {
        set_error("ssdf_mesh_inside_points: null mesh");
        return SSDF_ERR_INVALID;
    }
    return gwn_points(mesh, GW_INSIDE, npoints, x, y, z, nullptr, out_inside, (hipStream_t)stream);
}

int ssdf_mesh_inside_grid(ssdf_mesh_t mesh, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, double xmin, double ymin,
                          double zmin, double xmax, double ymax, double zmax, uint8_t *out_inside, void *stream) {
    if (!mesh) {
        set_error("ssdf_mesh_inside_grid: null mesh");
        return SSDF_ERR_INVALID;
    }
    const double lo[3] = {xmin, ymin, zmin}, hi[3] = {xmax, ymax, zmax};
    return gwn_grid(mesh, GW_INSIDE, nx, ny, nz, lo, hi, nullptr, out_inside, (hipStream_t)stream);
}

int ssdf_edf_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy,
               const float *sz, float *out) {
    if (cell_validate())
        return edf_validated<float>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
    return host_points_call<float>(SSDF_MODE_EDF, npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz,
                                   out, nullptr, nullptr, nullptr, nullptr, 0u);
}

int ssdf_edf_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const double *sx, const double *sy,
               const double *sz, double *out) {
    if (cell_validate())
        return edf_validated<double>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
    return host_points_call<double>(SSDF_MODE_EDF, npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz,
                                    out, nullptr, nullptr, nullptr, nullptr, 0u);
}

int ssdf_closest_point_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx,
                         const float *sy, const float *sz, float *outd, float *outx, float *outy, float *outz,
                         int *outtri, unsigned flags) {
    if (!outx || !outy || !outz) {
        set_error("ssdf_closest_point_f: null closest-point output");
        return SSDF_ERR_INVALID;
    }
    return host_points_call<float>(SSDF_MODE_CLOSEST, npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy,
                                   sz, outd, outx, outy, outz, outtri, flags);
}

int ssdf_closest_point_d(ptrdiff_t npoints, const double *x, const double *y, const double *z,
                         ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const double *sx, const double *sy, const double *sz, double *outd, double *outx,
                         double *outy, double *outz, int *outtri) {
    if (!outx || !outy || !outz) {
        set_error("ssdf_closest_point_d: null closest-point output");
        return SSDF_ERR_INVALID;
    }
    return host_points_call<double>(SSDF_MODE_CLOSEST, npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy,
                                    sz, outd, outx, outy, outz, outtri, 0u);
}

int ssdf_sample_sdf_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                      const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                      float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, float *out,
                      int *out_tri, unsigned flags) {
    return host_grid_call<float>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                 ymax, zmax, out, out_tri, flags);
}

int ssdf_sample_sdf_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                      const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                      ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                      double *out, int *out_tri) {
    return host_grid_call<double>(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                  ymax, zmax, out, out_tri, 0u);
}

int ssdf_sdf_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy,
               const float *sz, float *out_signed, int *out_tri, unsigned flags) {
    return host_points_call<float>(SSDF_MODE_SDF, npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz,
                                   out_signed, nullptr, nullptr, nullptr, out_tri, flags);
}

int ssdf_mesh_create(int dtype, ptrdiff_t nselements, const int *s0, const int *s1, const int *s2,
                     ptrdiff_t nspoints, const void *sx, const void *sy, const void *sz, int device_ptrs,
                     int with_normals, void *stream, ssdf_mesh_t *out) {
    if (!out || (dtype != 0 && dtype != 1)) {
        set_error("ssdf_mesh_create: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    return create_mesh(dtype, nselements, s0, s1, s2, nspoints, sx, sy, sz, device_ptrs != 0, with_normals != 0,
                       (hipStream_t)stream, out);
}

int ssdf_mesh_rebuild(ssdf_mesh_t mesh, int with_normals, void *stream) {
    if (!mesh) {
        set_error("ssdf_mesh_rebuild: null mesh");
        return SSDF_ERR_INVALID;
    }
    if (mesh->m.nsel <= 0) return SSDF_OK;
    const hipStream_t st = (hipStream_t)stream;
    // the rebuild overwrites the tree and the triangle records: after every earlier build and query
    if (int rc = order_after_handle(mesh, st)) return rc;
    mesh->m.stream = st;
    mesh->m.has_normals = false;
    mesh->m.has_gwn = false;
    if (int rc = build_lbvh(mesh->m, st)) return rc;
    if (with_normals)
        if (int rc = build_normals(mesh->m, st)) return rc;
    return sync_meta(mesh->m, st);
}

int ssdf_mesh_update_vertices(ssdf_mesh_t mesh, const void *sx, const void *sy, const void *sz, int device_ptrs,
                              int with_normals, void *stream) {
    if (!mesh || !sx || !sy || !sz) {
        set_error("ssdf_mesh_update_vertices: null argument");
        return SSDF_ERR_INVALID;
    }
    const size_t vb = (mesh->m.dtype ? sizeof(double) : sizeof(float)) * (size_t)mesh->m.nsp;
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipStream_t st = (hipStream_t)stream;
    // queries still running on other streams read the vertex arrays (FP64 re-evaluation, normals):
    // the copies wait for them
    if (int rc = order_after_handle(mesh, st)) return rc;
    SSDF_HIP_TRY(hipMemcpyAsync(mesh->m.sx, sx, vb, kind, st));
    SSDF_HIP_TRY(hipMemcpyAsync(mesh->m.sy, sy, vb, kind, st));
    SSDF_HIP_TRY(hipMemcpyAsync(mesh->m.sz, sz, vb, kind, st));
    return ssdf_mesh_rebuild(mesh, with_normals, stream);
}

int ssdf_mesh_destroy(ssdf_mesh_t mesh) {
    destroy_mesh(mesh, false);
    return SSDF_OK;
}

int ssdf_mesh_info(ssdf_mesh_t mesh, ssdf_mesh_info_t *info) {
    if (!mesh || !info) return SSDF_ERR_INVALID;
    if (mesh->m.nsel > 0)
        if (int rc = sync_meta(mesh->m, nullptr)) return rc;
    info->nselements = mesh->m.nsel;
    info->nspoints = mesh->m.nsp;
    info->dtype = mesh->m.dtype;
    info->maxdepth = mesh->m.hmeta.maxdepth;
    info->open_edges = mesh->m.has_normals ? mesh->m.hmeta.open_edges : -1;
    info->flipped_edges = mesh->m.has_normals ? mesh->m.hmeta.flipped_edges : -1;
    for (int d = 0; d < 6; d++) info->bbox[d] = ordered_to_float(mesh->m.hmeta.bbox_ord[d]);
    info->device_bytes = mesh->m.device_bytes;
    return SSDF_OK;
}

int ssdf_mesh_query_points(ssdf_mesh_t mesh, int mode, ptrdiff_t npoints, const void *x, const void *y,
                           const void *z, void *out_d, int *out_tri, void *out_cx, void *out_cy, void *out_cz,
                           unsigned flags, void *stream) {
    if (!mesh || mode < 0 || mode > 3) {
        set_error("ssdf_mesh_query_points: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    return query_points(mesh, mode, npoints, x, y, z, out_d, out_tri, out_cx, out_cy, out_cz, flags,
                        (hipStream_t)stream);
}

int ssdf_mesh_query_grid(ssdf_mesh_t mesh, int mode, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, double xmin,
                         double ymin, double zmin, double xmax, double ymax, double zmax, void *out_d, int *out_tri,
                         unsigned flags, void *stream) {
    if (!mesh || (mode != SSDF_MODE_SDF && mode != SSDF_MODE_EDF && mode != SSDF_MODE_SQEDF)) {
        set_error("ssdf_mesh_query_grid: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    const double lo[3] = {xmin, ymin, zmin}, hi[3] = {xmax, ymax, zmax};
    return query_grid(mesh, mode, nx, ny, nz, lo, hi, out_d, out_tri, flags, (hipStream_t)stream);
}

int ssdf_mesh_set_schedule(ssdf_mesh_t mesh, int brick_x, int brick_y, int brick_z, int heavy_leaves) {
    int lg[3] = {-1, -1, -1};
    const int ext[3] = {brick_x, brick_y, brick_z};
    for (int d = 0; d < 3; d++)
        for (int k = 0; k <= 6; k++)
            if (ext[d] == (1 << k)) lg[d] = k;
    if (!mesh || heavy_leaves < -1 || lg[0] < 0 || lg[1] < 0 || lg[2] < 0 || lg[0] + lg[1] + lg[2] != 6) {
        set_error("ssdf_mesh_set_schedule: the brick is 2^a x 2^b x 2^c = 64 nodes, heavy_leaves >= -1");
        return SSDF_ERR_INVALID;
    }
    mesh->brick = brick_code(lg[0], lg[1], lg[2]);
    mesh->heavy = heavy_leaves;
    return SSDF_OK;
}

int ssdf_mesh_last_stats(ssdf_mesh_t mesh, unsigned *fixup_queued, int *kernel_error) {
    if (!mesh) return SSDF_ERR_INVALID;
    unsigned st[2] = {0, 0};
    if (mesh->ws.status) SSDF_HIP_TRY(hipMemcpy(st, mesh->ws.status, sizeof(st), hipMemcpyDeviceToHost));
    if (fixup_queued) *fixup_queued = st[0];
    if (kernel_error) *kernel_error = (int)st[1];
    return SSDF_OK;
}

int ssdf_mesh_brute_points(ssdf_mesh_t mesh, ptrdiff_t npoints, const float *x, const float *y, const float *z,
                           double *out_d2, int *out_tri, void *stream) {
    if (!mesh || mesh->m.dtype != 0) {
        set_error("ssdf_mesh_brute_points: needs an FP32 mesh");
        return SSDF_ERR_INVALID;
    }
    if (npoints <= 0) return SSDF_OK;
    const unsigned nb = (unsigned)((npoints + 255) / 256);
    hipLaunchKernelGGL(k_brute_f64, dim3(nb), dim3(256), 0, (hipStream_t)stream, mesh->m.tris, (int)mesh->m.nsel,
                       x, y, z, (long long)npoints, out_d2, out_tri);
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

int ssdf_mesh_leaf_order(ssdf_mesh_t mesh, int *out_tri, void *stream) {
    if (!mesh || !out_tri) {
        set_error("ssdf_mesh_leaf_order: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    if (mesh->m.nsel <= 0) return SSDF_OK;
    const unsigned nb = (unsigned)((mesh->m.nsel + 255) / 256);
    hipLaunchKernelGGL(k_leaf_order, dim3(nb), dim3(256), 0, (hipStream_t)stream, mesh->m.tris, (int)mesh->m.nsel,
                       out_tri);
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

int ssdf_device_count(int *count) {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return SSDF_OK;
}
int ssdf_device_malloc(void **ptr, size_t bytes) {
    SSDF_HIP_TRY(hipMalloc(ptr, std::max<size_t>(bytes, 4)));
    return SSDF_OK;
}
int ssdf_device_free(void *ptr) {
    SSDF_HIP_TRY(hipFree(ptr));
    return SSDF_OK;
}
int ssdf_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    SSDF_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SSDF_OK;
}
int ssdf_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    SSDF_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SSDF_OK;
}
int ssdf_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) {
        set_error("ssdf_device_copy: 16-byte multiples and alignment");
        return SSDF_ERR_INVALID;
    }
    const size_t n = bytes / 16;
    if (n == 0) return SSDF_OK;
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, (hipStream_t)stream, (uint4 *)dst, (const uint4 *)src, n);
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}
int ssdf_stream_create(void **stream) {
    SSDF_HIP_TRY(hipStreamCreateWithFlags((hipStream_t *)stream, hipStreamNonBlocking));
    return SSDF_OK;
}
int ssdf_stream_destroy(void *stream) {
    SSDF_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return SSDF_OK;
}
int ssdf_stream_synchronize(void *stream) {
    SSDF_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return SSDF_OK;
}
int ssdf_event_create(void **event) {
    SSDF_HIP_TRY(hipEventCreate((hipEvent_t *)event));
    return SSDF_OK;
}
int ssdf_event_destroy(void *event) {
    SSDF_HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return SSDF_OK;
}
int ssdf_event_record(void *event, void *stream) {
    SSDF_HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return SSDF_OK;
}
int ssdf_event_elapsed_ms(void *start, void *stop, float *ms) {
    SSDF_HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    SSDF_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SSDF_OK;
}

int ssdf_set_heavy_leaves(int leaves) {
    if (leaves < 0) return heavy_setting().load();
    return heavy_setting().exchange(leaves);
}

const char *ssdf_last_error(void) { return get_error(); }
const char *ssdf_version(void) { return "ssdf_amd 0.1.0 (gfx950)"; }

}  // extern "C"

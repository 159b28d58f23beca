// lbvh.hip — LBVH construction and pseudonormals for a mesh resident in HBM.
//
// Build (one stream, no host round trip, graph-capturable):
//   k_mesh_init   meta to neutral values
//   k_tri_boxes   per-triangle FP32 AABB (outward-rounded for FP64 vertices) + per-block partial
//                 vertex/centre bounds, folded by k_reduce_bounds
//   k_morton      63-bit Morton key of each box centre, normalised per axis by the centre bounds
//   radix sort    (key, triangle) pairs, 8 passes of 8 bits, stable -> equal keys keep index order
//   k_karras      internal node i's children and parent links (Karras, HPG 2012), ties in the keys
//                 split by position, so the tree is a pure function of the mesh
//   k_refit_pack  one thread per leaf: packs the Morton-ordered 144-byte TriP record and refits the
//                 boxes and subtree heights bottom-up, in LDS within the block's leaves and with
//                 atomic min/max above; k_finalize_nodes turns the slots into the Nodes' float
//                 boxes and the root's height into the tree depth (sizes the traversal stack)
// The same construction, run serially, is the oracle's canonical replay (oracle/oracle.c), so the
// device tree and the replay's tree are identical.
#include "mesh.hpp"
#include "radix.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace ssdf_amd {

static thread_local std::string g_last_error;

hipError_t dev_alloc(void **ptr, size_t bytes, hipStream_t stream) {
    static std::once_flag once[64];
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
        std::call_once(once[dev], [dev] {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = ~0ull;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
        });
    }
    return hipMallocAsync(ptr, bytes < 16 ? 16 : bytes, stream);
}

void dev_free(void *ptr, hipStream_t stream) {
    if (ptr) (void)hipFreeAsync(ptr, stream);
}

int ensure_ws(DevMesh &m, size_t bytes, hipStream_t stream) {
    if (m.ws_cap >= bytes) return SSDF_OK;
    dev_free(m.ws, stream);
    m.ws = nullptr;
    m.ws_cap = 0;
    SSDF_HIP_TRY(dev_alloc((void **)&m.ws, bytes, stream));
    m.ws_cap = bytes;
    return SSDF_OK;
}
void set_error(const std::string &msg) { g_last_error = msg; }
const char *get_error() { return g_last_error.c_str(); }

__global__ void k_mesh_init(MeshMeta *meta) {
    if (threadIdx.x == 0) {
        meta->root = 0;
        meta->maxdepth = 0;
        for (int d = 0; d < 3; d++) {
            meta->bbox_ord[d] = 0x7fffffff;
            meta->bbox_ord[3 + d] = (int)0x80000000;
            meta->cbox_ord[d] = 0x7fffffff;
            meta->cbox_ord[3 + d] = (int)0x80000000;
            meta->root_box[d] = 0.f;
            meta->root_box[3 + d] = 0.f;
        }
        meta->open_edges = -1;
        meta->flipped_edges = -1;
    }
}

__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename G>
__device__ __forceinline__ float to_f_down(G v);
template <>
__device__ __forceinline__ float to_f_down<float>(float v) {
    return v;
}
template <>
__device__ __forceinline__ float to_f_down<double>(double v) {
    return __double2float_rd(v);
}
template <typename G>
__device__ __forceinline__ float to_f_up(G v);
template <>
__device__ __forceinline__ float to_f_up<float>(float v) {
    return v;
}
template <>
__device__ __forceinline__ float to_f_up<double>(double v) {
    return __double2float_ru(v);
}

// Triangle box (6 floats: min xyz, max xyz).  Each block also reduces its triangles' vertex
// bounds and centre bounds to 12 partials (min xyz, max xyz of each); k_reduce_bounds folds the
// partials into the mesh meta (two stages instead of contended atomics).
template <typename G>
__global__ __launch_bounds__(256) void k_tri_boxes(const int *__restrict__ s0, const int *__restrict__ s1,
                                                   const int *__restrict__ s2, const G *__restrict__ sx,
                                                   const G *__restrict__ sy, const G *__restrict__ sz, int n,
                                                   float *__restrict__ boxes, float *__restrict__ partials) {
    __shared__ float red[4][12];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float v[12];
#pragma unroll
    for (int d = 0; d < 3; d++) v[d] = SSDF_INF, v[3 + d] = -SSDF_INF, v[6 + d] = SSDF_INF, v[9 + d] = -SSDF_INF;
    if (t < n) {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        const G *co[3] = {sx, sy, sz};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const G a = co[d][iv[0]], b = co[d][iv[1]], c = co[d][iv[2]];
            G lo = b < a ? b : a;
            lo = c < lo ? c : lo;
            G hi = a < b ? b : a;
            hi = hi < c ? c : hi;
            const float flo = to_f_down<G>(lo), fhi = to_f_up<G>(hi);
            boxes[(size_t)6 * t + d] = flo;
            boxes[(size_t)6 * t + 3 + d] = fhi;
            const float cc = (flo + fhi) * 0.5f;
            v[d] = flo, v[3 + d] = fhi, v[6 + d] = cc, v[9 + d] = cc;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v[d] = wave_min_f(v[d]);
        v[3 + d] = wave_max_f(v[3 + d]);
        v[6 + d] = wave_min_f(v[6 + d]);
        v[9 + d] = wave_max_f(v[9 + d]);
    }
    if (lane == 0)
        for (int i = 0; i < 12; i++) red[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int i = threadIdx.x;
        const bool is_min = (i % 6) < 3;
        float r = red[0][i];
        for (int ww = 1; ww < 4; ww++) r = is_min ? fminf(r, red[ww][i]) : fmaxf(r, red[ww][i]);
        partials[(size_t)blockIdx.x * 12 + i] = r;
    }
}

__global__ __launch_bounds__(256) void k_reduce_bounds(const float *__restrict__ partials, int nparts,
                                                       MeshMeta *__restrict__ meta) {
    __shared__ float red[4][12];
    float v[12];
#pragma unroll
    for (int i = 0; i < 12; i++) v[i] = (i % 6) < 3 ? SSDF_INF : -SSDF_INF;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const float p = partials[(size_t)b * 12 + i];
            v[i] = (i % 6) < 3 ? fminf(v[i], p) : fmaxf(v[i], p);
        }
#pragma unroll
    for (int i = 0; i < 12; i++) v[i] = (i % 6) < 3 ? wave_min_f(v[i]) : wave_max_f(v[i]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int i = 0; i < 12; i++) red[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int i = threadIdx.x;
        const bool is_min = (i % 6) < 3;
        float r = red[0][i];
        for (int ww = 1; ww < 4; ww++) r = is_min ? fminf(r, red[ww][i]) : fmaxf(r, red[ww][i]);
        if (i < 6)
            meta->bbox_ord[i] = float_to_ordered(r);
        else
            meta->cbox_ord[i - 6] = float_to_ordered(r);
    }
}

__host__ __device__ __forceinline__ uint64_t expand21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(256) void k_morton(const float *__restrict__ boxes, int n,
                                                const MeshMeta *__restrict__ meta, uint64_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float lo = ordered_to_float(meta->cbox_ord[d]);
        const float hi = ordered_to_float(meta->cbox_ord[3 + d]);
        const float ext = hi - lo;
        const float c = (boxes[(size_t)6 * t + d] + boxes[(size_t)6 * t + 3 + d]) * 0.5f;
        const float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
        const float s = u * 2097152.0f;
        const uint32_t qi = s > 0.0f ? (uint32_t)s : 0u;
        q[d] = qi > 2097151u ? 2097151u : qi;
    }
    keys[t] = (expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]);
    vals[t] = (uint32_t)t;
}

__device__ __forceinline__ int kdelta(const uint64_t *__restrict__ codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = codes[i], b = codes[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz(i ^ j);
}

// parent[] holds 2n-1 entries: internal node i at i, leaf k at n-1+k; value (p << 1) | side.
__global__ __launch_bounds__(256) void k_karras(const uint64_t *__restrict__ codes, int n, Node *__restrict__ nodes,
                                                int *__restrict__ parent, int *__restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    (void)flags;
    const int dir = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(codes, n, i, i - dir);
    int lmax = 2;
    while (kdelta(codes, n, i, i + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(codes, n, i, i + (l + t) * dir) > dmin) l += t;
    const int j = i + l * dir;
    const int dnode = kdelta(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(codes, n, i, i + (s + t) * dir) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * dir + (dir < 0 ? dir : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    int left, right;
    if (lo == gamma) {
        left = ~gamma;
        parent[n - 1 + gamma] = (i << 1) | 0;
    } else {
        left = gamma;
        parent[gamma] = (i << 1) | 0;
    }
    if (hi == gamma + 1) {
        right = ~(gamma + 1);
        parent[n - 1 + gamma + 1] = (i << 1) | 1;
    } else {
        right = gamma + 1;
        parent[gamma + 1] = (i << 1) | 1;
    }
    nodes[i].left = left;
    nodes[i].right = right;
    nodes[i].first = lo;
    nodes[i].last = hi;
}

template <typename G>
__device__ __forceinline__ float to_f_near(G v) {
    return (float)v;
}

// Child-slot boxes during the build, as ordered ints (12 per internal node: left min xyz, max xyz,
// right min xyz, max xyz), so that min/max merge with integer atomics.
__global__ __launch_bounds__(256) void k_slots_init(int *__restrict__ slots, size_t nwords, int *__restrict__ hts,
                                                    size_t nh) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) slots[i] = ((i % 6) < 3) ? 0x7fffffff : (int)0x80000000;
    if (i < nh) hts[i] = 0;
}

// One thread per leaf (Morton position k): packs the TriP record, then climbs towards the root.
// While the node lies inside the block's 256 leaves, the boxes meet in LDS: each arriving thread
// writes its side, and the second of the two merges both and climbs on (workgroup-scope ordering
// only).  From the first node that reaches outside the block, the thread merges its box into each
// ancestor's child slot with atomic min/max and stops at the first slot that already contains it.
// Stopping there is exact: each bound of that slot came from a thread of the same subtree, and that
// thread carries the bound (or a tighter one) upward itself.  No global fences and no cross-XCD
// cache writebacks are needed; the next kernel sees every store's and atomic's result.  The LDS
// stage takes the ~6 device-scope atomics per level off the lower levels, where most threads are.
template <typename G>
__global__ __launch_bounds__(256) void k_refit_pack(const uint32_t *__restrict__ order, int n,
                                                    const int *__restrict__ s0, const int *__restrict__ s1,
                                                    const int *__restrict__ s2, const G *__restrict__ sx,
                                                    const G *__restrict__ sy, const G *__restrict__ sz,
                                                    const float *__restrict__ boxes, const int *__restrict__ parent,
                                                    const Node *__restrict__ nodes, int *__restrict__ slots,
                                                    int *__restrict__ hts, TriP *__restrict__ tris,
                                                    MeshMeta *__restrict__ meta) {
    // child boxes and arrival counts of the internal nodes whose leaf range lies inside this
    // block's 256 leaves (such a node's index is one end of its range, so it is inside too)
    __shared__ int lbox[256][12];
    __shared__ int lht[256][2];
    __shared__ int lcnt[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b0 = blockIdx.x * blockDim.x, b1 = min(n - 1, b0 + (int)blockDim.x - 1);
    lcnt[threadIdx.x] = 0;
    __syncthreads();
    if (k >= n) return;
    const int t = (int)order[k];
    {
        const int iv[3] = {s0[t], s1[t], s2[t]};
        float va[3], vb[3], vc[3];
        va[0] = to_f_near<G>(sx[iv[0]]), va[1] = to_f_near<G>(sy[iv[0]]), va[2] = to_f_near<G>(sz[iv[0]]);
        vb[0] = to_f_near<G>(sx[iv[1]]), vb[1] = to_f_near<G>(sy[iv[1]]), vb[2] = to_f_near<G>(sz[iv[1]]);
        vc[0] = to_f_near<G>(sx[iv[2]]), vc[1] = to_f_near<G>(sy[iv[2]]), vc[2] = to_f_near<G>(sz[iv[2]]);
        TriP tr;
        make_trip(tr, va, vb, vc, t);
        tris[k] = tr;
    }
    if (n == 1) {
        for (int d = 0; d < 6; d++) meta->root_box[d] = boxes[(size_t)6 * t + d];
        meta->root = ~0;
        return;
    }
    int b[6];
#pragma unroll
    for (int d = 0; d < 6; d++) b[d] = float_to_ordered(boxes[(size_t)6 * t + d]);
    // h: height of the subtree this thread carries (a leaf: 0; an internal node: 1 + its
    // children's larger height); in the atomic part a lower bound, merged with atomic max
    int p = parent[n - 1 + k];
    int h = 0;
    bool local = true;
    for (int guard = 0; guard < 4096; guard++) {
        const int pi = p >> 1, side = p & 1;
        if (local) local = nodes[pi].first >= b0 && nodes[pi].last <= b1;  // ranges nest: once out, out
        if (local) {
            // inside the block: LDS, and the second of the two threads to arrive merges and climbs
            int *ls = lbox[pi - b0];
#pragma unroll
            for (int d = 0; d < 6; d++) ls[6 * side + d] = b[d];
            lht[pi - b0][side] = h;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (atomicAdd(&lcnt[pi - b0], 1) == 0) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int *os = ls + 6 * (side ^ 1);
            int *gs = slots + (size_t)12 * pi;  // the node's final child boxes, for k_finalize_nodes
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const int o = os[d];
                gs[6 * side + d] = b[d];
                gs[6 * (side ^ 1) + d] = o;
                b[d] = d < 3 ? min(b[d], o) : max(b[d], o);
            }
            const int ho = lht[pi - b0][side ^ 1];
            hts[2 * pi + side] = h;
            hts[2 * pi + (side ^ 1)] = ho;
            h = 1 + max(h, ho);
        } else {
            // across blocks: atomic min/max into the child slot, climbing while the slot grew
            int *sl = slots + (size_t)12 * pi + 6 * side;
            bool grew = atomicMax(hts + 2 * pi + side, h) < h;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                grew |= atomicMin(sl + d, b[d]) > b[d];
                grew |= atomicMax(sl + 3 + d, b[3 + d]) < b[3 + d];
            }
            if (!grew) break;
            h += 1;
        }
        if (pi == 0) break;
        p = parent[pi];
    }
}

// Ordered-int slots -> the float child boxes of each Node; the root's box into the meta.
__global__ __launch_bounds__(256) void k_finalize_nodes(const int *__restrict__ slots, const int *__restrict__ hts,
                                                        int ninternal, Node *__restrict__ nodes,
                                                        MeshMeta *__restrict__ meta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninternal) return;
    const int *sl = slots + (size_t)12 * i;
    Node *nd = nodes + i;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        nd->lmin[d] = ordered_to_float(sl[d]);
        nd->lmax[d] = ordered_to_float(sl[3 + d]);
        nd->rmin[d] = ordered_to_float(sl[6 + d]);
        nd->rmax[d] = ordered_to_float(sl[9 + d]);
    }
    // a child subtree of at most kMaxRun leaves becomes a run reference (its range is contiguous in
    // Morton order and fixed by k_karras, so no other thread of this kernel changes it)
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const int c = side ? nd->right : nd->left;
        if (c < 0) continue;
        const int f = nodes[c].first, n = nodes[c].last - f + 1;
        if (n <= kMaxRun) {
            if (side)
                nd->right = leaf_run(f, n);
            else
                nd->left = leaf_run(f, n);
        }
    }
    if (i == 0) {
        meta->maxdepth = 1 + max(hts[0], hts[1]);  // leaves below the deepest path: internal nodes on it
        for (int d = 0; d < 3; d++) {
            meta->root_box[d] = fminf(ordered_to_float(sl[d]), ordered_to_float(sl[6 + d]));
            meta->root_box[3 + d] = fmaxf(ordered_to_float(sl[3 + d]), ordered_to_float(sl[9 + d]));
        }
    }
}

static unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// Build workspace of build_lbvh and build_normals for n triangles (the layouts below).
static size_t lbvh_ws_bytes(size_t nn) {
    const size_t ni = nn > 1 ? nn - 1 : 1;
    return 16 * nn + 8 * nn + 24 * nn + 8 * nn + 48 * ni + 8 * ni + 48 * (size_t)nblk(nn, 256) +
           4 * radix_scratch_words(nn);
}
static size_t normals_ws_bytes(size_t nn) { return 48 * 3 * nn + 4 * radix_scratch_words(3 * nn); }

size_t mesh_bytes_bound(ptrdiff_t nsel, ptrdiff_t nsp, int dtype, bool with_normals) {
    const size_t n = (size_t)std::max<ptrdiff_t>(nsel, 1), V = (size_t)std::max<ptrdiff_t>(nsp, 1);
    const size_t ni = n > 1 ? n - 1 : 1;
    const size_t vb = dtype ? sizeof(double) : sizeof(float);
    size_t b = 3 * sizeof(int) * n + 3 * vb * V;                     // caller-layout copies
    b += sizeof(Node) * ni + sizeof(TriP) * n + sizeof(MeshMeta);   // LBVH
    size_t ws = lbvh_ws_bytes(n);
    if (with_normals) {
        b += sizeof(double) * (12 * n + 3 * V) + kConnBytesPerCorner * 3 * n;
        ws = std::max(ws, normals_ws_bytes(n));
    }
    return b + ws;
}

int build_lbvh(DevMesh &m, hipStream_t stream) {
    const int n = (int)m.nsel;
    if (m.nsel <= 0 || m.nsel >= ((ptrdiff_t)1 << kLeafBits)) {
        set_error("build_lbvh: triangle count out of range");
        return SSDF_ERR_INVALID;
    }
    const size_t nn = (size_t)n, ni = nn > 1 ? nn - 1 : 1;
    if (!m.nodes) {
        SSDF_HIP_TRY(dev_alloc((void **)&m.meta, sizeof(MeshMeta), stream));
        SSDF_HIP_TRY(dev_alloc((void **)&m.nodes, sizeof(Node) * ni, stream));
        SSDF_HIP_TRY(dev_alloc((void **)&m.tris, sizeof(TriP) * nn, stream));
        m.device_bytes += sizeof(Node) * ni + sizeof(TriP) * nn + sizeof(MeshMeta);
    }
    // one workspace allocation: boxes 24 B, keys 2 x 8 B, vals 2 x 4 B, parent 8 B, slots 48 B and
    // child heights 8 B per triangle, block partials, radix scratch
    const unsigned nb = nblk(nn, 256);
    const size_t off_keys = 0, off_vals = off_keys + 16 * nn, off_boxes = off_vals + 8 * nn,
                 off_parent = off_boxes + 24 * nn, off_slots = off_parent + 8 * nn, off_hts = off_slots + 48 * ni,
                 off_part = off_hts + 8 * ni,
                 off_scr = off_part + 48 * (size_t)nb, total = off_scr + 4 * radix_scratch_words(nn);
    if (int rc = ensure_ws(m, total, stream)) return rc;
    char *ws = m.ws;
    uint64_t *keys = (uint64_t *)(ws + off_keys);
    uint32_t *vals = (uint32_t *)(ws + off_vals);
    float *boxes = (float *)(ws + off_boxes);
    int *parent = (int *)(ws + off_parent);
    int *slots = (int *)(ws + off_slots);
    int *hts = (int *)(ws + off_hts);
    float *partials = (float *)(ws + off_part);
    uint32_t *scratch = (uint32_t *)(ws + off_scr);

    hipLaunchKernelGGL(k_mesh_init, dim3(1), dim3(64), 0, stream, m.meta);
    if (m.dtype == 0)
        hipLaunchKernelGGL(k_tri_boxes<float>, dim3(nb), dim3(256), 0, stream, m.s0, m.s1, m.s2, (const float *)m.sx,
                           (const float *)m.sy, (const float *)m.sz, n, boxes, partials);
    else
        hipLaunchKernelGGL(k_tri_boxes<double>, dim3(nb), dim3(256), 0, stream, m.s0, m.s1, m.s2,
                           (const double *)m.sx, (const double *)m.sy, (const double *)m.sz, n, boxes, partials);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(256), 0, stream, partials, (int)nb, m.meta);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(256), 0, stream, boxes, n, m.meta, keys, vals);
    uint64_t *skeys = nullptr;
    uint32_t *svals = nullptr;
    SSDF_HIP_TRY(radix_sort_pairs<uint64_t>(keys, keys + nn, vals, vals + nn, nn, 64, scratch, stream, &skeys, &svals));
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(nblk(nn - 1, 256)), dim3(256), 0, stream, skeys, n, m.nodes, parent,
                           (int *)nullptr);
        hipLaunchKernelGGL(k_slots_init, dim3(nblk(12 * ni, 256)), dim3(256), 0, stream, slots, 12 * ni, hts, 2 * ni);
    }
    if (m.dtype == 0)
        hipLaunchKernelGGL(k_refit_pack<float>, dim3(nb), dim3(256), 0, stream, svals, n, m.s0, m.s1, m.s2,
                           (const float *)m.sx, (const float *)m.sy, (const float *)m.sz, boxes, parent, m.nodes,
                           slots, hts, m.tris, m.meta);
    else
        hipLaunchKernelGGL(k_refit_pack<double>, dim3(nb), dim3(256), 0, stream, svals, n, m.s0, m.s1, m.s2,
                           (const double *)m.sx, (const double *)m.sy, (const double *)m.sz, boxes, parent,
                           m.nodes, slots, hts, m.tris, m.meta);
    if (n > 1) {
        hipLaunchKernelGGL(k_finalize_nodes, dim3(nblk(ni, 256)), dim3(256), 0, stream, slots, hts, (int)ni, m.nodes,
                           m.meta);
    }
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

int sync_meta(DevMesh &m, hipStream_t stream) {
    SSDF_HIP_TRY(hipMemcpyAsync(&m.hmeta, m.meta, sizeof(MeshMeta), hipMemcpyDeviceToHost, stream));
    SSDF_HIP_TRY(hipStreamSynchronize(stream));
    return SSDF_OK;
}

// ------------------------------------------------------------------------------------------------
// Pseudonormals, FP64; follows serial/bvh/bvh.cpp:1669-1721.  The reference accumulates vertex
// normals in triangle order and pairs face edges through a hash map; here both become stable
// sorts, which visit the same contributions in the same order:
//   vertex: entries 3t+v keyed by vertex id, summed per vertex in increasing t;
//   edge:   entries 3t+e keyed min*V+max; the first face f0 of a key pairs with every later fk,
//           edge_n[fk] = n[f0] + n[fk], and edge_n[f0] keeps the last sum.
// ------------------------------------------------------------------------------------------------
template <typename G>
__global__ __launch_bounds__(256) void k_face_normals(const int *__restrict__ s0, const int *__restrict__ s1,
                                                      const int *__restrict__ s2, const G *__restrict__ sx,
                                                      const G *__restrict__ sy, const G *__restrict__ sz, int n,
                                                      long long nsp, double *__restrict__ face_n,
                                                      double *__restrict__ edge_n, double *__restrict__ angle,
                                                      uint32_t *__restrict__ vkeys, uint32_t *__restrict__ vvals,
                                                      uint64_t *__restrict__ ekeys, uint32_t *__restrict__ evals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int iv[3] = {s0[t], s1[t], s2[t]};
    double p[3][3];
#pragma unroll
    for (int v = 0; v < 3; v++) p[v][0] = (double)sx[iv[v]], p[v][1] = (double)sy[iv[v]], p[v][2] = (double)sz[iv[v]];
    const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    const double w[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    double nn[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    const double len = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    if (len > 0) {
        for (int d = 0; d < 3; d++) nn[d] /= len;
    }
    for (int d = 0; d < 3; d++) face_n[(size_t)3 * t + d] = nn[d];
    for (int e = 0; e < 3; e++)
        for (int d = 0; d < 3; d++) edge_n[(size_t)9 * t + 3 * e + d] = nn[d];
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const double *a = p[v], *b = p[(v + 1) % 3], *c = p[(v + 2) % 3];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        const double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        const double cosine = (l1 > 0 && l2 > 0) ? (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]) / (l1 * l2) : 1.0;
        angle[(size_t)3 * t + v] = acos(fmax(-1.0, fmin(1.0, cosine)));
        if (!vkeys) continue;  // the connectivity's sorted corners are already known
        vkeys[(size_t)3 * t + v] = (uint32_t)iv[v];
        vvals[(size_t)3 * t + v] = (uint32_t)(3 * t + v);
        const uint64_t ea = (uint64_t)iv[v], eb = (uint64_t)iv[(v + 1) % 3];
        ekeys[(size_t)3 * t + v] = (ea < eb ? ea : eb) * (uint64_t)nsp + (ea < eb ? eb : ea);
        evals[(size_t)3 * t + v] = (uint32_t)(3 * t + v);
    }
}

__global__ __launch_bounds__(256) void k_fill_f64(double *__restrict__ p, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_vertex_normals(const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ vals, size_t m,
                                                        const double *__restrict__ face_n,
                                                        const double *__restrict__ angle,
                                                        double *__restrict__ vertex_n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t key = keys[i];
    if (i > 0 && keys[i - 1] == key) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t j = i; j < m && keys[j] == key; j++) {
        const uint32_t e = vals[j];
        const double a = angle[e];
        const size_t t = e / 3;
        for (int d = 0; d < 3; d++) acc[d] += a * face_n[3 * t + d];
    }
    for (int d = 0; d < 3; d++) vertex_n[(size_t)3 * key + d] = acc[d];
}

__global__ __launch_bounds__(256) void k_edge_normals(const uint64_t *__restrict__ keys,
                                                      const uint32_t *__restrict__ vals, size_t m,
                                                      const int *__restrict__ s0, const int *__restrict__ s1,
                                                      const int *__restrict__ s2, const double *__restrict__ face_n,
                                                      double *__restrict__ edge_n, MeshMeta *__restrict__ meta) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t key = keys[i];
    if (i > 0 && keys[i - 1] == key) return;
    size_t j = i + 1;
    while (j < m && keys[j] == key) j++;
    const size_t g = j - i;
    const uint32_t f0 = vals[i];
    const size_t t0 = f0 / 3;
    for (size_t k = i + 1; k < j; k++) {
        const uint32_t fk = vals[k];
        const size_t tk = fk / 3;
        for (int d = 0; d < 3; d++) {
            const double sum = face_n[3 * t0 + d] + face_n[3 * tk + d];
            edge_n[9 * tk + 3 * (fk % 3) + d] = sum;
            edge_n[9 * t0 + 3 * (f0 % 3) + d] = sum;
        }
    }
    if (g != 2) {
        atomicAdd(&meta->open_edges, 1);
    } else {
        const int *s[3] = {s0, s1, s2};
        const uint32_t f1 = vals[i + 1];
        const int a0 = s[f0 % 3][t0], b0 = s[(f0 % 3 + 1) % 3][t0];
        const int a1 = s[f1 % 3][f1 / 3];
        if (a0 == a1 && b0 != a0) atomicAdd(&meta->flipped_edges, 1);
    }
}

__global__ void k_meta_zero_edges(MeshMeta *meta) {
    meta->open_edges = 0;
    meta->flipped_edges = 0;
}

int build_normals(DevMesh &m, hipStream_t stream) {
    const size_t n = (size_t)m.nsel, V = (size_t)m.nsp;
    const size_t e = 3 * n;
    if (!m.face_n) {
        SSDF_HIP_TRY(dev_alloc((void **)&m.face_n, sizeof(double) * 3 * n, stream));
        SSDF_HIP_TRY(dev_alloc((void **)&m.edge_n, sizeof(double) * 9 * n, stream));
        SSDF_HIP_TRY(dev_alloc((void **)&m.vertex_n, sizeof(double) * 3 * V, stream));
        m.device_bytes += sizeof(double) * (12 * n + 3 * V);
    }
    if (!m.conn) {
        SSDF_HIP_TRY(dev_alloc((void **)&m.conn, kConnBytesPerCorner * e, stream));
        m.device_bytes += kConnBytesPerCorner * e;
    }
    // the corners sorted by vertex id and by edge key: they depend on the connectivity alone, which
    // a handle never changes, so only the first build sorts them
    uint32_t *cvk = (uint32_t *)m.conn, *cvv = cvk + e;
    uint64_t *cek = (uint64_t *)(m.conn + 8 * e);
    uint32_t *cev = (uint32_t *)(m.conn + 16 * e);
    const bool sort = !m.conn_ready;
    // workspace: angle 8 B, vertex keys/vals 2 x 2 x 4 B, edge keys 2 x 8 B, edge vals 2 x 4 B per
    // corner, radix scratch
    const size_t off_angle = 0, off_vk = off_angle + 8 * e, off_vv = off_vk + 8 * e, off_ek = off_vv + 8 * e,
                 off_ev = off_ek + 16 * e, off_scr = off_ev + 8 * e, total = off_scr + 4 * radix_scratch_words(e);
    if (int rc = ensure_ws(m, total, stream)) return rc;
    double *angle = (double *)(m.ws + off_angle);
    uint32_t *vkeys = (uint32_t *)(m.ws + off_vk), *vvals = (uint32_t *)(m.ws + off_vv);
    uint64_t *ekeys = (uint64_t *)(m.ws + off_ek);
    uint32_t *evals = (uint32_t *)(m.ws + off_ev), *scratch = (uint32_t *)(m.ws + off_scr);
    hipLaunchKernelGGL(k_meta_zero_edges, dim3(1), dim3(1), 0, stream, m.meta);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * V, 256)), dim3(256), 0, stream, m.vertex_n, 3 * V, 0.0);
    uint32_t *kv = sort ? vkeys : nullptr, *kvv = sort ? vvals : nullptr, *kev = sort ? evals : nullptr;
    uint64_t *kek = sort ? ekeys : nullptr;
    if (m.dtype == 0)
        hipLaunchKernelGGL(k_face_normals<float>, dim3(nblk(n, 256)), dim3(256), 0, stream, m.s0, m.s1, m.s2,
                           (const float *)m.sx, (const float *)m.sy, (const float *)m.sz, (int)n, (long long)V,
                           m.face_n, m.edge_n, angle, kv, kvv, kek, kev);
    else
        hipLaunchKernelGGL(k_face_normals<double>, dim3(nblk(n, 256)), dim3(256), 0, stream, m.s0, m.s1, m.s2,
                           (const double *)m.sx, (const double *)m.sy, (const double *)m.sz, (int)n, (long long)V,
                           m.face_n, m.edge_n, angle, kv, kvv, kek, kev);
    if (sort) {
        int vbits = 1;
        while (vbits < 32 && ((size_t)1 << vbits) < V) vbits++;
        int ebits = 1;
        const unsigned long long maxkey = (unsigned long long)V * (unsigned long long)V;
        while (ebits < 64 && (1ull << ebits) < maxkey) ebits++;
        uint32_t *svk = nullptr, *svv = nullptr, *sev = nullptr;
        uint64_t *sek = nullptr;
        SSDF_HIP_TRY(
            radix_sort_pairs<uint32_t>(vkeys, vkeys + e, vvals, vvals + e, e, vbits, scratch, stream, &svk, &svv));
        SSDF_HIP_TRY(
            radix_sort_pairs<uint64_t>(ekeys, ekeys + e, evals, evals + e, e, ebits, scratch, stream, &sek, &sev));
        SSDF_HIP_TRY(hipMemcpyAsync(cvk, svk, sizeof(uint32_t) * e, hipMemcpyDeviceToDevice, stream));
        SSDF_HIP_TRY(hipMemcpyAsync(cvv, svv, sizeof(uint32_t) * e, hipMemcpyDeviceToDevice, stream));
        SSDF_HIP_TRY(hipMemcpyAsync(cek, sek, sizeof(uint64_t) * e, hipMemcpyDeviceToDevice, stream));
        SSDF_HIP_TRY(hipMemcpyAsync(cev, sev, sizeof(uint32_t) * e, hipMemcpyDeviceToDevice, stream));
        m.conn_ready = true;
    }
    hipLaunchKernelGGL(k_vertex_normals, dim3(nblk(e, 256)), dim3(256), 0, stream, cvk, cvv, e, m.face_n, angle,
                       m.vertex_n);
    hipLaunchKernelGGL(k_edge_normals, dim3(nblk(e, 256)), dim3(256), 0, stream, cek, cev, e, m.s0, m.s1, m.s2,
                       m.face_n, m.edge_n, m.meta);
    SSDF_HIP_TRY(hipGetLastError());
    m.has_normals = true;
    return SSDF_OK;
}

void free_mesh(DevMesh &m, hipStream_t st) {
    if (m.owns_input) {
        dev_free(m.s0, st);
        dev_free(m.s1, st);
        dev_free(m.s2, st);
        dev_free(m.sx, st);
        dev_free(m.sy, st);
        dev_free(m.sz, st);
    }
    dev_free(m.nodes, st);
    dev_free(m.tris, st);
    dev_free(m.meta, st);
    dev_free(m.face_n, st);
    dev_free(m.edge_n, st);
    dev_free(m.vertex_n, st);
    dev_free(m.conn, st);
    dev_free(m.gwn, st);
    dev_free(m.ws, st);
    m = DevMesh{};
}

}  // namespace ssdf_amd

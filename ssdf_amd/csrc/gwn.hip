// gwn.hip — generalized winding number w(q) = (1 / 4 pi) sum of the solid angles of the triangles
// seen from q (1 inside a closed outward-oriented surface, 0 outside, fractional near holes).
// It gives signs and inside/outside flags that stay meaningful on open or non-manifold meshes, where
// pseudonormal signs (serial/bvh/bvh.cpp:1734-1792) are undefined away from the surface, and it is
// the engine behind points_inside_bvh / sample_inside_bvh (serial/bvh/bvh.cpp:318-365, 1800-1866),
// whose reference answer on a closed surface is the same inside/outside decision.
//
// Far field (Barill et al., "Fast winding numbers for soups and clouds", SIGGRAPH 2018, order-1
// terms): a subtree with summed area vector N = sum(a_i n_i) and area-weighted centroid c
// contributes N . (c - q) / |c - q|^3 to the solid angle when |c - q| > beta r, r the radius of its
// box around c; otherwise it is opened, and a leaf triangle contributes its exact solid angle
// (Van Oosterom & Strackee 1983: 2 atan2(a . (b x c), |a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|)).
//
// Build: per-leaf FP64 (area vector, area, area * centroid) in Morton order, an inclusive FP64
// prefix scan, and per internal node the difference of the prefix at its leaf range [first, last]
// for each child: deterministic, no atomics.
//
// Query: wave-coherent like the distance traversal.  The wave visits a node and, for each child,
// adds the far-field term in every lane if every lane accepts it, evaluates the triangle exactly
// if it is a leaf, and otherwise pushes it: every lane sums one shared cut of the tree.
#include <algorithm>
#include <climits>

#include "query.hpp"

namespace ssdf_amd {
namespace {

constexpr int kGwnScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kGwnScanTile = kGwnScanBlock * kScanItems;
constexpr int kCh = 8;  // area vector (3), area, area * centroid (3), unused
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr int kGwnWaves = kWavesPerBlock;
// far-field acceptance |c - q| > beta r; order-1 terms at beta = 3 keep |w error| well below the
// 1/2 threshold away from the surface (tests/test_gpu_gwn.py bounds it against the exact sum)
constexpr float kGwnBeta = 3.0f;

__global__ __launch_bounds__(256) void k_gwn_leaf(const TriP *__restrict__ tris, int n, double *__restrict__ v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const TriF t = tris[k].f;
    const double a[3] = {t.a[0], t.a[1], t.a[2]}, b[3] = {t.b[0], t.b[1], t.b[2]}, c[3] = {t.c[0], t.c[1], t.c[2]};
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double an[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                          0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    const double A = sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
    double *o = v + (size_t)kCh * k;
    o[0] = an[0], o[1] = an[1], o[2] = an[2], o[3] = A;
    o[4] = A * ((a[0] + b[0] + c[0]) / 3.0);
    o[5] = A * ((a[1] + b[1] + c[1]) / 3.0);
    o[6] = A * ((a[2] + b[2] + c[2]) / 3.0);
    o[7] = 0.0;
}

// Inclusive scan of the kCh channels across the block's threads (one value set per thread).
__device__ void block_scan8(double (&run)[kCh], double (&excl)[kCh], double (&total)[kCh]) {
    __shared__ double s[kCh][kGwnScanBlock];
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < kCh; c++) s[c][t] = run[c];
    __syncthreads();
    for (int off = 1; off < kGwnScanBlock; off <<= 1) {
        double tmp[kCh];
#pragma unroll
        for (int c = 0; c < kCh; c++) tmp[c] = t >= off ? s[c][t - off] : 0.0;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kCh; c++) s[c][t] += tmp[c];
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < kCh; c++) {
        excl[c] = t > 0 ? s[c][t - 1] : 0.0;
        total[c] = s[c][kGwnScanBlock - 1];
    }
    __syncthreads();
}

// Local inclusive scan of one tile of kGwnScanTile leaves; the tile's total goes to bsum.
__global__ __launch_bounds__(kGwnScanBlock) void k_scan8_local(double *__restrict__ v, int n, double *__restrict__ bsum) {
    const size_t base = (size_t)blockIdx.x * kGwnScanTile + (size_t)threadIdx.x * kScanItems;
    double run[kCh] = {};
    for (int i = 0; i < kScanItems; i++) {
        const size_t idx = base + i;
        if (idx < (size_t)n) {
            double *o = v + kCh * idx;
#pragma unroll
            for (int c = 0; c < kCh; c++) {
                run[c] += o[c];
                o[c] = run[c];
            }
        }
    }
    double excl[kCh], total[kCh];
    block_scan8(run, excl, total);
    for (int i = 0; i < kScanItems; i++) {
        const size_t idx = base + i;
        if (idx < (size_t)n) {
            double *o = v + kCh * idx;
#pragma unroll
            for (int c = 0; c < kCh; c++) o[c] += excl[c];
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < kCh; c++) bsum[(size_t)kCh * blockIdx.x + c] = total[c];
    }
}

// Exclusive scan of the tile totals in place (one block).
__global__ __launch_bounds__(kGwnScanBlock) void k_scan8_tiles(double *__restrict__ bsum, int nt) {
    double carry[kCh] = {};
    for (int b0 = 0; b0 < nt; b0 += kGwnScanBlock) {
        const int i = b0 + (int)threadIdx.x;
        double run[kCh];
#pragma unroll
        for (int c = 0; c < kCh; c++) run[c] = i < nt ? bsum[(size_t)kCh * i + c] : 0.0;
        double excl[kCh], total[kCh];
        block_scan8(run, excl, total);
        if (i < nt) {
#pragma unroll
            for (int c = 0; c < kCh; c++) bsum[(size_t)kCh * i + c] = carry[c] + excl[c];
        }
#pragma unroll
        for (int c = 0; c < kCh; c++) carry[c] += total[c];
    }
}

__global__ __launch_bounds__(kGwnScanBlock) void k_scan8_add(double *__restrict__ v, int n,
                                                          const double *__restrict__ boff) {
    const size_t base = (size_t)blockIdx.x * kGwnScanTile + (size_t)threadIdx.x * kScanItems;
    const double *off = boff + (size_t)kCh * blockIdx.x;
    for (int i = 0; i < kScanItems; i++) {
        const size_t idx = base + i;
        if (idx < (size_t)n) {
#pragma unroll
            for (int c = 0; c < kCh; c++) v[kCh * idx + c] += off[c];
        }
    }
}

// Far-field records of node i's two children from the inclusive prefix I over the leaves.
__global__ __launch_bounds__(256) void k_gwn_nodes(const Node *__restrict__ nodes, int ni,
                                                   const double *__restrict__ I, GwnNode *__restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const Node nd = nodes[i];
    GwnNode o;
    o.left = nd.left;
    o.right = nd.right;
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const int child = side ? nd.right : nd.left;
        int f, l;
        if (child >= 0) {
            f = nodes[child].first, l = nodes[child].last;
        } else {
            f = leaf_first(child);  // a run of leaf_count(child) leaves
            l = f + leaf_count(child) - 1;
        }
        double S[7];
#pragma unroll
        for (int c = 0; c < 7; c++) S[c] = I[(size_t)kCh * l + c] - (f > 0 ? I[(size_t)kCh * (f - 1) + c] : 0.0);
        const float *bmin = side ? nd.rmin : nd.lmin, *bmax = side ? nd.rmax : nd.lmax;
        double cen[3], r2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double lo = bmin[d], hi = bmax[d];
            double cd = S[3] > 0.0 ? S[4 + d] / S[3] : 0.5 * (lo + hi);
            cd = fmin(fmax(cd, lo), hi);
            cen[d] = cd;
            const double ext = fmax(cd - lo, hi - cd);
            r2 += ext * ext;
        }
        // widened so that the FP32-rounded centre still bounds the box
        const float r2f = (float)(r2 * (1.0 + 1e-5)) + 1e-30f;
        float *pn = side ? o.rn : o.ln, *pc = side ? o.rc : o.lc;
#pragma unroll
        for (int d = 0; d < 3; d++) pn[d] = (float)S[d], pc[d] = (float)cen[d];
        if (side)
            o.rr2 = r2f;
        else
            o.lr2 = r2f;
    }
    g[i] = o;
}

__device__ __forceinline__ float solid_angle(float qx, float qy, float qz, const float *A, const float *B,
                                             const float *C) {
    const float ax = A[0] - qx, ay = A[1] - qy, az = A[2] - qz;
    const float bx = B[0] - qx, by = B[1] - qy, bz = B[2] - qz;
    const float cx = C[0] - qx, cy = C[1] - qy, cz = C[2] - qz;
    // the hardware square root (1 ulp), not the correctly rounded sequence: the winding number is an
    // approximation to 0.03 anyway
    const float la = __builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az);
    const float lb = __builtin_amdgcn_sqrtf(bx * bx + by * by + bz * bz);
    const float lc = __builtin_amdgcn_sqrtf(cx * cx + cy * cy + cz * cz);
    const float det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    const float den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                      (bx * cx + by * cy + bz * cz) * la;
    return 2.0f * atan2f(det, den);
}

template <int STACK, int QSRC, typename QT, int MODE>
__device__ __forceinline__ void gwn_wave(const GwnNode *__restrict__ g, const TriP *__restrict__ tris,
                                         const QueryArgs &a, const unsigned wave_id) {
    const int lane = threadIdx.x & 63;
    long long qid = 0;
    QT qx0 = QT(0), qy0 = QT(0), qz0 = QT(0);
    const bool valid = lane_query<QSRC, QT>(a, wave_id, lane, qid, qx0, qy0, qz0);
    if (__ballot(valid) == 0) return;
    const float qx = (float)qx0, qy = (float)qy0, qz = (float)qz0;
    const float beta2 = a.beta2;
    float acc = 0.0f;
    int node = __builtin_amdgcn_readfirstlane(a.meta->root);
    if (node < 0) {
        const TriF t = tris[~node].f;
        acc = solid_angle(qx, qy, qz, t.a, t.b, t.c);
    } else {
        int stackv = 0, stackv_hi = 0, sp = 0;
        bool overflow = false;
        long long iters = 0;
        auto visit = [&](int child, const float *n, const float *c, float r2) {
            const float dx = c[0] - qx, dy = c[1] - qy, dz = c[2] - qz;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const bool near = valid && !(d2 > beta2 * r2);
            if (__ballot(near) == 0) {
                // |c - q|^-3 from the hardware reciprocal square root (1 ulp; d2 > 0 here)
                const float rs = __builtin_amdgcn_rsqf(d2);
                const float inv = rs * rs * rs;
                acc += (n[0] * dx + n[1] * dy + n[2] * dz) * inv;
            } else if (child < 0) {
                const int k0 = leaf_first(child), nl = leaf_count(child);
                for (int j = 0; j < nl; j++) {
                    const TriF t = tris[k0 + j].f;
                    acc += solid_angle(qx, qy, qz, t.a, t.b, t.c);
                }
            } else if (sp < STACK) {
                if (STACK == 64) {
                    stackv = lane == sp ? child : stackv;
                } else {
                    stackv = lane == sp ? child : stackv;
                    stackv_hi = lane == sp - 64 ? child : stackv_hi;
                }
                sp++;
            } else {
                overflow = true;
            }
        };
        for (;;) {
            const GwnNode gn = g[node];
            visit(gn.left, gn.ln, gn.lc, gn.lr2);
            visit(gn.right, gn.rn, gn.rc, gn.rr2);
            if (sp == 0 || overflow) break;
            --sp;
            node = (STACK == 64 || sp < 64) ? __builtin_amdgcn_readlane(stackv, sp & 63)
                                            : __builtin_amdgcn_readlane(stackv_hi, sp & 63);
            if (++iters > a.max_iters) break;
        }
        if (lane == 0 && (overflow || iters > a.max_iters)) atomicOr(a.err_flag, overflow ? 1 : 2);
    }
    if (!valid) return;
    const float w = acc * kInv4Pi;
    if (MODE == GW_WINDING) {
        a.out_w[qid] = w;
    } else if (MODE == GW_INSIDE) {
        a.out_in[qid] = w >= 0.5f ? 1 : 0;
    } else {
        const QT d = load_as<QT>(a.out_d, qid);
        const QT ad = d < QT(0) ? -d : d;
        store_as<QT>(a.out_d, qid, w >= 0.5f ? -ad : ad);
    }
}

// One block per kGwnWaves waves, or (a.persistent) one resident set taking waves from 8 per-XCD
// ranges with stealing, as k_query_packet does.
template <int STACK, int QSRC, typename QT, int MODE>
__global__ __launch_bounds__(256) void k_gwn(const GwnNode *__restrict__ g, const TriP *__restrict__ tris,
                                             QueryArgs a) {
    WaveSource src(a);
    for (unsigned w; src.next(a, w);) gwn_wave<STACK, QSRC, QT, MODE>(g, tris, a, w);
}

__global__ void k_reset_wqueue(unsigned *wq) {
    if (threadIdx.x < 8) wq[32 * threadIdx.x] = 0u;
}

// Blocks of kernel k resident on one compute unit (cached per instance).
template <int STACK, int QSRC, typename QT, int MODE>
unsigned gwn_blocks_per_cu() {
    static const unsigned n = [] {
        int b = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, reinterpret_cast<const void *>(&k_gwn<STACK, QSRC, QT, MODE>), 64 * kGwnWaves, 0);
        return e == hipSuccess && b > 0 ? (unsigned)b : 1u;
    }();
    return n;
}

template <int STACK, int QSRC, typename QT, int MODE>
void launch_mode(const DevMesh &m, QueryArgs a, unsigned nblocks, hipStream_t stream) {
    // more waves than the GPU holds at once: a persistent launch (the wave-queue counters of the
    // query's workspace, which a distance query before may have used, cleared first)
    const unsigned resident = device_cus() * gwn_blocks_per_cu<STACK, QSRC, QT, MODE>();
    a.persistent = nblocks > resident && a.wqueue ? 1u : 0u;
    if (a.persistent) {
        hipLaunchKernelGGL(k_reset_wqueue, dim3(1), dim3(64), 0, stream, a.wqueue);
        nblocks = resident;
    }
    hipLaunchKernelGGL((k_gwn<STACK, QSRC, QT, MODE>), dim3(nblocks), dim3(64 * kGwnWaves), 0, stream, m.gwn, m.tris,
                       a);
}

template <int STACK, int QSRC, typename QT>
void dispatch_mode(const DevMesh &m, const QueryArgs &a, int mode, unsigned nblocks, hipStream_t stream) {
    if (mode == GW_WINDING)
        launch_mode<STACK, QSRC, QT, GW_WINDING>(m, a, nblocks, stream);
    else if (mode == GW_INSIDE)
        launch_mode<STACK, QSRC, QT, GW_INSIDE>(m, a, nblocks, stream);
    else
        launch_mode<STACK, QSRC, QT, GW_SIGN>(m, a, nblocks, stream);
}

template <int QSRC, typename QT>
int dispatch_stack(const DevMesh &m, const QueryArgs &a, int mode, unsigned nblocks, hipStream_t stream) {
    // pending entries never exceed the depth of the deepest internal node + 2
    const int need = m.hmeta.maxdepth + 1;
    if (need <= 64)
        dispatch_mode<64, QSRC, QT>(m, a, mode, nblocks, stream);
    else if (need <= 128)
        dispatch_mode<128, QSRC, QT>(m, a, mode, nblocks, stream);
    else {
        set_error("LBVH deeper than 127 levels");
        return SSDF_ERR_DEPTH;
    }
    return SSDF_OK;
}

}  // namespace

size_t gwn_bytes_bound(ptrdiff_t nsel) {
    const size_t n = (size_t)std::max<ptrdiff_t>(nsel, 1);
    const size_t nt = (n + kGwnScanTile - 1) / kGwnScanTile;
    return sizeof(GwnNode) * n + sizeof(double) * kCh * (n + nt);
}

// Far-field winding-number tree of the mesh (see the head of this file).
int build_gwn(DevMesh &m, hipStream_t stream) {
    const int n = (int)m.nsel;
    if (n < 2) {
        m.has_gwn = true;  // the root is a leaf, evaluated exactly
        return SSDF_OK;
    }
    const int ni = n - 1;
    if (!m.gwn) {
        SSDF_HIP_TRY(dev_alloc((void **)&m.gwn, sizeof(GwnNode) * (size_t)ni, stream));
        m.device_bytes += sizeof(GwnNode) * (size_t)ni;
    }
    const int nt = (n + kGwnScanTile - 1) / kGwnScanTile;
    if (int rc = ensure_ws(m, sizeof(double) * kCh * ((size_t)n + (size_t)nt), stream)) return rc;
    double *v = (double *)m.ws, *bsum = v + (size_t)kCh * n;
    hipLaunchKernelGGL(k_gwn_leaf, dim3((n + 255) / 256), dim3(256), 0, stream, m.tris, n, v);
    hipLaunchKernelGGL(k_scan8_local, dim3(nt), dim3(kGwnScanBlock), 0, stream, v, n, bsum);
    hipLaunchKernelGGL(k_scan8_tiles, dim3(1), dim3(kGwnScanBlock), 0, stream, bsum, nt);
    hipLaunchKernelGGL(k_scan8_add, dim3(nt), dim3(kGwnScanBlock), 0, stream, v, n, bsum);
    hipLaunchKernelGGL(k_gwn_nodes, dim3((ni + 255) / 256), dim3(256), 0, stream, m.nodes, ni, v, m.gwn);
    SSDF_HIP_TRY(hipGetLastError());
    m.has_gwn = true;
    return SSDF_OK;
}

int launch_gwn(DevMesh &m, QueryArgs &a, int qsrc, int mode, int q_is_double, hipStream_t stream) {
    if (m.nsel <= 0) return SSDF_OK;
    if (q_is_double != m.dtype) {
        set_error("winding numbers take queries in the mesh precision");
        return SSDF_ERR_UNSUPPORTED;
    }
    if (!m.has_gwn)
        if (int rc = build_gwn(m, stream)) return rc;
    a.nodes = m.nodes;
    a.tris = m.tris;
    a.meta = m.meta;
    a.gwn = m.gwn;
    if (a.beta2 <= 0.0f) a.beta2 = kGwnBeta * kGwnBeta;
    a.max_iters = 4 * (long long)m.nsel + 64;
    const long long nq = qsrc == QSRC_POINTS ? a.npoints : (long long)a.nx * a.ny * a.nz;
    if (nq <= 0) return SSDF_OK;
    unsigned nwaves;
    if (qsrc == QSRC_POINTS) {
        nwaves = (unsigned)((nq + 63) / 64);
    } else {
        unsigned long long w = 0;
        grid_bricks(a.brick, a.nx, a.ny, a.nz, a.nbx, a.nby, a.nbz, w);
        if (w > 0xffffffffull - kGwnWaves) {
            set_error("winding-number grid: too many nodes for one launch");
            return SSDF_ERR_INVALID;
        }
        nwaves = (unsigned)w;
    }
    const unsigned nblocks = (nwaves + kGwnWaves - 1) / kGwnWaves;
    a.nblocks = nblocks;
    a.total_waves = nwaves;
    int rc;
    if (m.dtype == 0)
        rc = qsrc == QSRC_POINTS ? dispatch_stack<QSRC_POINTS, float>(m, a, mode, nblocks, stream)
                                 : dispatch_stack<QSRC_GRID, float>(m, a, mode, nblocks, stream);
    else
        rc = qsrc == QSRC_POINTS ? dispatch_stack<QSRC_POINTS, double>(m, a, mode, nblocks, stream)
                                 : dispatch_stack<QSRC_GRID, double>(m, a, mode, nblocks, stream);
    if (rc) return rc;
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

}  // namespace ssdf_amd

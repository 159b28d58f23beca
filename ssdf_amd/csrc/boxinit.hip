// boxinit.hip — the distance driver's seeding helpers on the GPU: the bounds of a point cloud
// (compute_aabb, serial/edf/edf.hpp:500-539), the signed distance of every point to a box
// (all_points_aabb_signed_distance, edf.hpp:541-557, over point_aabb_signed_distance, 193-221), and
// their fusion into the SSDF_INIT_WITH_VOLUME_BOX seed |sdist(p, bounds of all points)|
// (drivers/sdf.exe.cpp:317-327), which callers use as the initial `out` of the edf entry points.
//
// Bounds: each thread folds a grid-strided slice starting from point 0 with the reference's
// min(a, b) = b < a ? b : a and max(a, b) = a < b ? b : a, blocks fold their threads in LDS, one
// block folds the per-block partials.  For NaN-free input this is the exact min / max (as the
// reference's OpenMP reduction, the choice between +0 and -0 depends on the order).  HBM-bound:
// 3 sizeof(T) bytes per point read once.
#include <algorithm>

#include "mesh.hpp"

namespace ssdf_amd {
namespace {

constexpr int kBlock = 256;
constexpr unsigned kMaxParts = 1024;

template <typename T>
__device__ __forceinline__ T ref_min(T a, T b) {
    return b < a ? b : a;
}
template <typename T>
__device__ __forceinline__ T ref_max(T a, T b) {
    return a < b ? b : a;
}

// Folds v[6] (xmin, xmax, ymin, ymax, zmin, zmax) across the block; thread 0 returns the result.
template <typename T>
__device__ void block_fold6(T (&v)[6], T *out6) {
    __shared__ T s[6][kBlock];
#pragma unroll
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < 6; k += 2) {
                s[k][threadIdx.x] = ref_min(s[k][threadIdx.x], s[k][threadIdx.x + off]);
                s[k + 1][threadIdx.x] = ref_max(s[k + 1][threadIdx.x], s[k + 1][threadIdx.x + off]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[threadIdx.x] = s[threadIdx.x][0];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bounds_partial(ptrdiff_t n, const T *__restrict__ x,
                                                           const T *__restrict__ y, const T *__restrict__ z,
                                                           T *__restrict__ part) {
    T v[6] = {x[0], x[0], y[0], y[0], z[0], z[0]};
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * kBlock) {
        const T xi = x[i], yi = y[i], zi = z[i];
        v[0] = ref_min(v[0], xi), v[1] = ref_max(v[1], xi);
        v[2] = ref_min(v[2], yi), v[3] = ref_max(v[3], yi);
        v[4] = ref_min(v[4], zi), v[5] = ref_max(v[5], zi);
    }
    block_fold6(v, part + 6 * (size_t)blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bounds_final(unsigned nparts, const T *__restrict__ part,
                                                         T *__restrict__ box6) {
    T v[6] = {part[0], part[1], part[2], part[3], part[4], part[5]};
    for (unsigned p = threadIdx.x; p < nparts; p += kBlock) {
        v[0] = ref_min(v[0], part[6 * p + 0]), v[1] = ref_max(v[1], part[6 * p + 1]);
        v[2] = ref_min(v[2], part[6 * p + 2]), v[3] = ref_max(v[3], part[6 * p + 3]);
        v[4] = ref_min(v[4], part[6 * p + 4]), v[5] = ref_max(v[5], part[6 * p + 5]);
    }
    block_fold6(v, box6);
}

// point_aabb_signed_distance (serial/edf/edf.hpp:193-221), operation for operation; negative inside.
template <typename T>
__device__ __forceinline__ T box_sdist(T px, T py, T pz, const T *b) {
    const T minx = b[0], maxx = b[1], miny = b[2], maxy = b[3], minz = b[4], maxz = b[5];
    const T hx = (maxx - minx) * T(0.5), hy = (maxy - miny) * T(0.5), hz = (maxz - minz) * T(0.5);
    const T cx = (maxx + minx) * T(0.5), cy = (maxy + miny) * T(0.5), cz = (maxz + minz) * T(0.5);
    const T qx = fabs(px - cx) - hx, qy = fabs(py - cy) - hy, qz = fabs(pz - cz) - hz;
    const T ox = qx > T(0) ? qx : T(0), oy = qy > T(0) ? qy : T(0), oz = qz > T(0) ? qz : T(0);
    const T outside = sqrt(ox * ox + oy * oy + oz * oz);
    const T mq = ref_max(qx, ref_max(qy, qz));
    const T inside = mq < T(0) ? mq : T(0);
    return outside + inside;
}

// box6 is read from device memory, so the fused seed needs no host round trip.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_box_sdist(ptrdiff_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                      const T *__restrict__ z, const T *__restrict__ box6,
                                                      T *__restrict__ out, int absval) {
    const T b[6] = {box6[0], box6[1], box6[2], box6[3], box6[4], box6[5]};
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * kBlock) {
        const T d = box_sdist(x[i], y[i], z[i], b);
        out[i] = absval ? fabs(d) : d;
    }
}

unsigned parts_for(ptrdiff_t n) {
    return (unsigned)std::min<ptrdiff_t>((n + kBlock - 1) / kBlock, (ptrdiff_t)kMaxParts);
}

unsigned blocks_for(ptrdiff_t n) {
    return (unsigned)std::min<ptrdiff_t>((n + kBlock - 1) / kBlock, (ptrdiff_t)65536);
}

// Bounds of n >= 1 device points into box6 (device); part: 6 * kMaxParts T of scratch.
template <typename T>
void launch_bounds(ptrdiff_t n, const T *x, const T *y, const T *z, T *part, T *box6, hipStream_t stream) {
    const unsigned np = parts_for(n);
    hipLaunchKernelGGL(k_bounds_partial<T>, dim3(np), dim3(kBlock), 0, stream, n, x, y, z, part);
    hipLaunchKernelGGL(k_bounds_final<T>, dim3(1), dim3(kBlock), 0, stream, np, part, box6);
}

// RAII device buffer of the synchronous host-pointer entry points.
struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

template <typename T>
int upload3(Buf &b, ptrdiff_t n, const T *x, const T *y, const T *z, size_t extra) {
    SSDF_HIP_TRY(hipMalloc(&b.p, sizeof(T) * (3 * (size_t)n + extra)));
    T *d = (T *)b.p;
    SSDF_HIP_TRY(hipMemcpy(d, x, sizeof(T) * n, hipMemcpyHostToDevice));
    SSDF_HIP_TRY(hipMemcpy(d + n, y, sizeof(T) * n, hipMemcpyHostToDevice));
    SSDF_HIP_TRY(hipMemcpy(d + 2 * n, z, sizeof(T) * n, hipMemcpyHostToDevice));
    return SSDF_OK;
}

template <typename T>
int compute_aabb_host(ptrdiff_t n, const T *x, const T *y, const T *z, T *xmin, T *xmax, T *ymin, T *ymax, T *zmin,
                      T *zmax) {
    if (n <= 0 || !x || !y || !z || !xmin || !xmax || !ymin || !ymax || !zmin || !zmax) {
        set_error("compute_aabb: needs at least one point and non-null arrays");
        return SSDF_ERR_INVALID;
    }
    Buf b;
    if (int rc = upload3(b, n, x, y, z, 6 * (size_t)kMaxParts + 6)) return rc;
    T *d = (T *)b.p, *part = d + 3 * n, *box = part + 6 * kMaxParts;
    launch_bounds<T>(n, d, d + n, d + 2 * n, part, box, nullptr);
    SSDF_HIP_TRY(hipGetLastError());
    T h[6];
    SSDF_HIP_TRY(hipMemcpy(h, box, sizeof(h), hipMemcpyDeviceToHost));
    *xmin = h[0], *xmax = h[1], *ymin = h[2], *ymax = h[3], *zmin = h[4], *zmax = h[5];
    return SSDF_OK;
}

template <typename T>
int all_points_host(ptrdiff_t n, const T *x, const T *y, const T *z, T minx, T maxx, T miny, T maxy, T minz, T maxz,
                    T *out) {
    if (n <= 0) return SSDF_OK;
    if (!x || !y || !z || !out) {
        set_error("all_points_aabb_signed_distance: null array");
        return SSDF_ERR_INVALID;
    }
    Buf b;
    if (int rc = upload3(b, n, x, y, z, (size_t)n + 6)) return rc;
    T *d = (T *)b.p, *o = d + 3 * n, *box = o + n;
    const T h[6] = {minx, maxx, miny, maxy, minz, maxz};
    SSDF_HIP_TRY(hipMemcpy(box, h, sizeof(h), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_box_sdist<T>, dim3(blocks_for(n)), dim3(kBlock), 0, nullptr, n, d, d + n, d + 2 * n,
                       box, o, 0);
    SSDF_HIP_TRY(hipGetLastError());
    SSDF_HIP_TRY(hipMemcpy(out, o, sizeof(T) * n, hipMemcpyDeviceToHost));
    return SSDF_OK;
}

template <typename T>
int init_volume_box(ptrdiff_t n, const T *x, const T *y, const T *z, T *out, hipStream_t stream) {
    void *ws = nullptr;
    SSDF_HIP_TRY(dev_alloc(&ws, sizeof(T) * (6 * (size_t)kMaxParts + 6), stream));
    T *part = (T *)ws, *box = part + 6 * kMaxParts;
    launch_bounds<T>(n, x, y, z, part, box, stream);
    hipLaunchKernelGGL(k_box_sdist<T>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, x, y, z, box, out, 1);
    const hipError_t e = hipGetLastError();
    dev_free(ws, stream);
    SSDF_HIP_TRY(e);
    return SSDF_OK;
}

}  // namespace
}  // namespace ssdf_amd

using namespace ssdf_amd;

extern "C" {

int ssdf_compute_aabb_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, float *xmin, float *xmax,
                        float *ymin, float *ymax, float *zmin, float *zmax) {
    return compute_aabb_host<float>(npoints, x, y, z, xmin, xmax, ymin, ymax, zmin, zmax);
}

int ssdf_compute_aabb_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, double *xmin,
                        double *xmax, double *ymin, double *ymax, double *zmin, double *zmax) {
    return compute_aabb_host<double>(npoints, x, y, z, xmin, xmax, ymin, ymax, zmin, zmax);
}

int ssdf_all_points_aabb_signed_distance_f(ptrdiff_t npoints, const float *x, const float *y, const float *z,
                                           float minx, float maxx, float miny, float maxy, float minz, float maxz,
                                           float *out) {
    return all_points_host<float>(npoints, x, y, z, minx, maxx, miny, maxy, minz, maxz, out);
}

int ssdf_all_points_aabb_signed_distance_d(ptrdiff_t npoints, const double *x, const double *y, const double *z,
                                           double minx, double maxx, double miny, double maxy, double minz,
                                           double maxz, double *out) {
    return all_points_host<double>(npoints, x, y, z, minx, maxx, miny, maxy, minz, maxz, out);
}

int ssdf_init_volume_box(int dtype, ptrdiff_t npoints, const void *x, const void *y, const void *z, void *out,
                         void *stream) {
    if (npoints <= 0) return SSDF_OK;
    if (!x || !y || !z || !out || (dtype != 0 && dtype != 1)) {
        set_error("ssdf_init_volume_box: null array or dtype not 0/1");
        return SSDF_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    return dtype == 0 ? init_volume_box<float>(npoints, (const float *)x, (const float *)y, (const float *)z,
                                                (float *)out, s)
                      : init_volume_box<double>(npoints, (const double *)x, (const double *)y, (const double *)z,
                                                 (double *)out, s);
}

}  // extern "C"

// noalloc.hip — the allocation-free point query (ssdf_mesh_query_points_noalloc).
//
// The workspace of a call (status words, queue, point sort) is `scratch`, a static device array
// sized for SSDF_NOALLOC_MAX_POINTS points, instead of the handle's grown-on-demand buffers.  Each
// GPU holds its own copy from the library's load until the process ends; every call on that GPU
// shares it, so calls that overlap on one GPU race.
#include "query.hpp"

#include <atomic>
#include <mutex>

using namespace ssdf_amd;

namespace {
constexpr long long kNoAllocMaxPoints = SSDF_NOALLOC_MAX_POINTS;
constexpr size_t kNoAllocBytes = query_ws_bytes((size_t)kNoAllocMaxPoints, true);
static_assert(kNoAllocBytes < ((size_t)24 << 20), "the static workspace stays modest");
}  // namespace

// On the host this name is a stand-in, not device memory: host code uses the address that
// noalloc_scratch() looks up for the current device, never the name itself.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::atomic<void *> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per device (with that device
// current) and kept.
int noalloc_scratch(void **out) {
    int dev = 0;
    SSDF_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("ssdf_mesh_query_points_noalloc: device ordinal out of range");
        return SSDF_ERR_UNSUPPORTED;
    }
    void *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
        if (!p) {
            SSDF_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) {
                set_error("ssdf_mesh_query_points_noalloc: no device address for the static workspace");
                return SSDF_ERR_HIP;
            }
            g_scratch_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return SSDF_OK;
}
}  // namespace

extern "C" int ssdf_mesh_query_points_noalloc(ssdf_mesh_t mesh, int mode, ptrdiff_t npoints, const void *x,
                                              const void *y, const void *z, void *out_d, int *out_tri,
                                              void *out_cx, void *out_cy, void *out_cz, unsigned flags,
                                              void *stream) {
    if (!mesh || mode < 0 || mode > 3) {
        set_error("ssdf_mesh_query_points_noalloc: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    if (npoints > kNoAllocMaxPoints) {
        set_error("ssdf_mesh_query_points_noalloc: at most SSDF_NOALLOC_MAX_POINTS (1048576) points per call");
        return SSDF_ERR_INVALID;
    }
    if (npoints <= 0) return SSDF_OK;
    void *ws = nullptr;
    if (int rc = noalloc_scratch(&ws)) return rc;
    return ws_query_points(mesh, ws, true, mode, npoints, x, y, z, out_d, out_tri, out_cx, out_cy, out_cz, flags,
                           (hipStream_t)stream);
}

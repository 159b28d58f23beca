// mesh.hpp — a triangle mesh resident in HBM with its LBVH and (optionally) its pseudonormals.
#pragma once

#include "common.hpp"

#include <string>

namespace ssdf_amd {

// Per-mesh scalars the kernels read from device memory (so that a captured build + query never
// needs a host round trip).
struct MeshMeta {
    int root;       // child reference of the root (internal 0, or leaf ~0 for one triangle)
    int maxdepth;   // deepest leaf, edges from the root
    int bbox_ord[6];  // vertex bounds as ordered ints (min xyz, max xyz)
    int cbox_ord[6];  // triangle-centre bounds as ordered ints
    float root_box[6];
    int open_edges;  // edges not shared by exactly two faces (-1 until normals are built)
    int flipped_edges;  // manifold edges whose two faces traverse it in the same direction
};

constexpr size_t kConnBytesPerCorner = 4 + 4 + 8 + 4;

struct DevMesh {
    int dtype = 0;  // 0: float vertices, 1: double vertices
    ptrdiff_t nsel = 0, nsp = 0;
    // caller-layout copies (SoA), needed for FP64 re-evaluation and normals
    int *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;
    void *sx = nullptr, *sy = nullptr, *sz = nullptr;
    bool owns_input = true;
    // LBVH
    Node *nodes = nullptr;  // nsel - 1 (at least 1)
    TriP *tris = nullptr;   // nsel, leaf order
    MeshMeta *meta = nullptr;  // device
    MeshMeta hmeta{};          // host copy after sync_meta()
    // pseudonormals (FP64): face 3T, edge 9T, vertex 3V
    double *face_n = nullptr, *edge_n = nullptr, *vertex_n = nullptr;
    bool has_normals = false;
    // the normals' view of the connectivity, sorted once by the first normals build (a handle's
    // triangles never change, only its vertex positions): the 3T corners by vertex id (keys,
    // corner ids) and by edge key (keys, corner ids), kConnBytesPerCorner bytes per corner
    char *conn = nullptr;
    bool conn_ready = false;
    // winding-number far field (nsel - 1 records), built on first use, stale after a rebuild
    GwnNode *gwn = nullptr;
    bool has_gwn = false;
    size_t device_bytes = 0;
    hipStream_t stream = nullptr;  // allocations and frees are ordered on this stream
    // build workspace, kept across rebuilds (the LBVH build and the normals build run one after the
    // other on `stream` and share it)
    char *ws = nullptr;
    size_t ws_cap = 0;
};

// Grows m.ws to at least `bytes` (stream-ordered); no-op when it is large enough.
int ensure_ws(DevMesh &m, size_t bytes, hipStream_t stream);

// Stream-ordered device memory from the device's default pool, which keeps freed blocks for reuse
// (release threshold raised to the maximum at first use): repeated builds and queries allocate
// without a device-wide synchronisation.
hipError_t dev_alloc(void **ptr, size_t bytes, hipStream_t stream);
void dev_free(void *ptr, hipStream_t stream);

// Uploads nothing: builds the LBVH of a mesh whose SoA arrays are already on the device.
int build_lbvh(DevMesh &m, hipStream_t stream);
// FP64 pseudonormals on the device (face, edge, vertex), following serial/bvh/bvh.cpp:1669-1721.
int build_normals(DevMesh &m, hipStream_t stream);
// Far-field tree of the generalized winding number (gwn.hip): per internal node, each child's
// area vector, centroid and radius, from FP64 prefix sums over the Morton-ordered leaves.
int build_gwn(DevMesh &m, hipStream_t stream);
// Upper bounds of the device bytes a mesh can hold: its arrays, LBVH, normals (with_normals) and
// build workspace (lbvh.hip); the winding-number tree and its build workspace (gwn.hip).
size_t mesh_bytes_bound(ptrdiff_t nsel, ptrdiff_t nsp, int dtype, bool with_normals);
size_t gwn_bytes_bound(ptrdiff_t nsel);
// Copies meta back (synchronises the stream).
int sync_meta(DevMesh &m, hipStream_t stream);
// Frees the mesh's device arrays, ordered on `st` (a stream that is still alive; null: the null
// stream, for callers that have synchronized the device first).
void free_mesh(DevMesh &m, hipStream_t st);

// thread-local last error message of the C ABI
void set_error(const std::string &msg);
const char *get_error();

#define SSDF_HIP_TRY(expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::ssdf_amd::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
            return e_ == hipErrorOutOfMemory ? SSDF_ERR_OOM : SSDF_ERR_HIP; \
        }                                                                                     \
    } while (0)

}  // namespace ssdf_amd

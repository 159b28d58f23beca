/*
 * ssdf_c.h — C ABI of the MI355X distance-field engine (libssdf_amd.so).
 *
 * Drop-in for zulianp/ssdf's exported symbols plus the entry points its closest-point path needs
 * on a GPU.  Every function returns 0 on success and a non-zero SSDF_ERR_* code otherwise (the
 * reference returns int 0 on its hot path and never fails; this engine reports device errors
 * instead of aborting).  ssdf_last_error() gives the message of the calling thread's last error.
 *
 * Layout conventions are the reference's (serial/ssdf.hpp:427-446): SoA coordinates, 0-based int
 * triangle indices s0, s1, s2, ptrdiff_t counts, x-fastest grids (serial/bvh/bvh.cpp:1792).
 */
#ifndef SSDF_C_H
#define SSDF_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SSDF_OK = 0,
    SSDF_ERR_INVALID = 1,
    SSDF_ERR_HIP = 2,
    SSDF_ERR_OOM = 3,
    SSDF_ERR_DEPTH = 4,
    SSDF_ERR_UNSUPPORTED = 5,
    SSDF_ERR_BUSY = 6
};

/* query flags */
#define SSDF_FLAG_FIXUP 1u     /* FP32 traversal + FP64 re-evaluation of near-tied candidates */
#define SSDF_FLAG_SIGN_FP64 2u /* evaluate every sign in FP64 */
/* signs from the generalized winding number (negative where w >= 1/2) instead of pseudonormals:
 * always (GWN), or only for meshes with open or non-manifold edges (AUTO) */
#define SSDF_FLAG_SIGN_GWN 8u
#define SSDF_FLAG_SIGN_AUTO 16u

/* query modes of the resident-mesh API */
#define SSDF_MODE_EDF 0     /* unsigned distance, min-accumulated into out_d */
#define SSDF_MODE_SDF 1     /* signed distance (pseudonormal sign) + closest triangle, overwriting */
#define SSDF_MODE_CLOSEST 2 /* distance + closest point + triangle, on strict improvement only */
#define SSDF_MODE_SQEDF 3   /* squared unsigned distance, min-accumulated into out_d */

/* ---------------------------------------------------------------------------------------------
 * Reference symbols (host pointers, synchronous; results in host memory on return).
 * --------------------------------------------------------------------------------------------- */

/* Unsigned point-to-surface distance, FP32; out is in/out (min-accumulated), an empty surface
 * leaves it untouched.  Replaces ssdf_edf_f, serial/ssdf.cpp:36-50 (-> edf_select<float,float,int>,
 * serial/ssdf.hpp:965-1016).  One call: upload, LBVH build, wave-coherent FP32 query, download. */
int ssdf_edf_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy,
               const float *sz, float *out);

/* FP64 twin; exact FP64 argmin (the reference's CPU edf<double> semantics).  Replaces ssdf_edf_d,
 * serial/ssdf.cpp:52-66. */
int ssdf_edf_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const double *sx, const double *sy,
               const double *sz, double *out);

/* ---------------------------------------------------------------------------------------------
 * Entry points behind the reference's C++ templates (no C symbol in the reference; the C++
 * header ssdf/ssdf.hpp wraps these with the reference's template signatures).
 * --------------------------------------------------------------------------------------------- */

/* closest_point_bvh (serial/bvh/bvh.hpp:34-51, bvh.cpp:125-197): updates outd, outx/y/z and outtri
 * only where the new distance is strictly below outd. */
int ssdf_closest_point_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx,
                         const float *sy, const float *sz, float *outd, float *outx, float *outy, float *outz,
                         int *outtri, unsigned flags);
int ssdf_closest_point_d(ptrdiff_t npoints, const double *x, const double *y, const double *z,
                         ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const double *sx, const double *sy, const double *sz, double *outd, double *outx,
                         double *outy, double *outz, int *outtri);

/* sample_sdf_bvh (serial/bvh/bvh.hpp:203-227, bvh.cpp:1641-1798): signed distance on the nodes of
 * an nx x ny x nz grid over [min, max], x fastest, positive where the right-handed normals point.
 * out_tri (may be NULL) receives the closest triangle.  Overwrites. */
int ssdf_sample_sdf_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                      const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                      float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, float *out,
                      int *out_tri, unsigned flags);
int ssdf_sample_sdf_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                      const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                      ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                      double *out, int *out_tri);

/* Signed distance at arbitrary points (the per-point form of sample_sdf_bvh). */
int ssdf_sdf_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
               const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy,
               const float *sz, float *out_signed, int *out_tri, unsigned flags);

/* sample_sqedf_bvh (serial/bvh/bvh.hpp:183-201, bvh.cpp:1575-1639): on the nodes of an nx x ny x nz
 * grid over [min, max] (x fastest), out holds squared distances and becomes min(out, d^2) (callers
 * seed it, e.g. with FLT_MAX).  The float variant takes query flags (SSDF_FLAG_FIXUP); the double
 * variant is exact FP64. */
int ssdf_sample_sqedf_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                        const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                        float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, float *out,
                        unsigned flags);
int ssdf_sample_sqedf_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                        const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                        ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                        double *out);

/* points_inside_bvh (serial/bvh/bvh.hpp:53-66, bvh.cpp:318-365): out[i] = 1 if point i is inside
 * the surface, else 0.  The reference decides by cuBQL's ray crossing parity; this engine by the
 * generalized winding number w >= 1/2, the same decision on closed surfaces and a stable one on open
 * ones.  An empty surface gives 0 everywhere. */
int ssdf_points_inside_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const float *sx,
                         const float *sy, const float *sz, uint8_t *out);
int ssdf_points_inside_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nselements,
                         const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints, const double *sx,
                         const double *sy, const double *sz, uint8_t *out);

/* sample_inside_bvh (serial/bvh/bvh.hpp:229-247, bvh.cpp:1800-1866): the same on the nodes of an
 * nx x ny x nz grid over [min, max], x fastest. */
int ssdf_sample_inside_f(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz,
                         float xmin, float ymin, float zmin, float xmax, float ymax, float zmax, uint8_t *out);
int ssdf_sample_inside_d(ptrdiff_t nselements, const int *s0, const int *s1, const int *s2, ptrdiff_t nspoints,
                         const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                         ptrdiff_t nz, double xmin, double ymin, double zmin, double xmax, double ymax, double zmax,
                         uint8_t *out);

/* ---------------------------------------------------------------------------------------------
 * Seeding helpers of the distance driver (host pointers, synchronous).
 * --------------------------------------------------------------------------------------------- */

/* compute_aabb (serial/edf/edf.hpp:500-539): bounds of npoints >= 1 points, seeded with point 0 and
 * reduced with the reference's min/max.  SSDF_ERR_INVALID for npoints <= 0 (the reference reads
 * x[0]). */
int ssdf_compute_aabb_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, float *xmin, float *xmax,
                        float *ymin, float *ymax, float *zmin, float *zmax);
int ssdf_compute_aabb_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, double *xmin,
                        double *xmax, double *ymin, double *ymax, double *zmin, double *zmax);

/* all_points_aabb_signed_distance (serial/edf/edf.hpp:541-557 over point_aabb_signed_distance,
 * 193-221): signed distance of every point to the box, negative inside; overwrites out. */
int ssdf_all_points_aabb_signed_distance_f(ptrdiff_t npoints, const float *x, const float *y, const float *z,
                                           float minx, float maxx, float miny, float maxy, float minz, float maxz,
                                           float *out);
int ssdf_all_points_aabb_signed_distance_d(ptrdiff_t npoints, const double *x, const double *y, const double *z,
                                           double minx, double maxx, double miny, double maxy, double minz,
                                           double maxz, double *out);

/* The SSDF_INIT_WITH_VOLUME_BOX seed of drivers/sdf.exe.cpp:317-327 on DEVICE arrays, asynchronous
 * on `stream`: out[i] = |signed distance of point i to the bounds of all points|, the initial value
 * of `out` for the edf entry points.  dtype 0 float, 1 double; points and out in that precision. */
int ssdf_init_volume_box(int dtype, ptrdiff_t npoints, const void *x, const void *y, const void *z, void *out,
                         void *stream);

/* ---------------------------------------------------------------------------------------------
 * Resident meshes (no reference counterpart: the reference rebuilds its BVH on every call,
 * serial/bvh/bvh.cu:80-158).  A mesh handle keeps its LBVH (and pseudonormals) in HBM; queries take
 * DEVICE pointers and run asynchronously on `stream` (a hipStream_t, NULL = default stream).
 * Calls on one handle must not overlap (they share its fixup workspace).
 * --------------------------------------------------------------------------------------------- */
typedef struct ssdf_mesh *ssdf_mesh_t;

typedef struct {
    ptrdiff_t nselements, nspoints;
    int dtype;           /* 0 float, 1 double */
    int maxdepth;        /* LBVH depth */
    int open_edges;      /* edges not shared by exactly two faces (-1: normals not built) */
    int flipped_edges;   /* shared edges traversed the same way by both faces */
    float bbox[6];       /* min xyz, max xyz */
    size_t device_bytes; /* HBM held by the handle */
} ssdf_mesh_info_t;

/* dtype 0: float vertices, 1: double.  device_ptrs != 0: the arrays are device memory and are
 * copied on the device.  with_normals != 0: also build the FP64 pseudonormals (needed by SDF).
 * A mesh holds at most 2^29 - 1 triangles (SSDF_ERR_INVALID beyond): the BVH's leaf references
 * carry a leaf position and a run length in one int.  The same limit holds for every entry point
 * that builds a BVH, and per mesh in batches. */
int ssdf_mesh_create(int dtype, ptrdiff_t nselements, const int *s0, const int *s1, const int *s2,
                     ptrdiff_t nspoints, const void *sx, const void *sy, const void *sz, int device_ptrs,
                     int with_normals, void *stream, ssdf_mesh_t *out);
/* Waits for the device to finish and frees the handle; it does not use the streams the handle was
 * created or queried on, so it is safe after the caller has destroyed them. */
int ssdf_mesh_destroy(ssdf_mesh_t mesh);
/* Rebuilds the LBVH (and, with_normals != 0, the pseudonormals) from the handle's current arrays,
 * reusing its device memory: no allocation after the first build.  For meshes whose vertices move
 * between queries (the reference's warm-start / FSI reuse workflow keeps the field, not the tree). */
int ssdf_mesh_rebuild(ssdf_mesh_t mesh, int with_normals, void *stream);
/* Copies new vertex coordinates (same count, host or device memory) into the handle and rebuilds. */
int ssdf_mesh_update_vertices(ssdf_mesh_t mesh, const void *sx, const void *sy, const void *sz, int device_ptrs,
                              int with_normals, void *stream);
int ssdf_mesh_info(ssdf_mesh_t mesh, ssdf_mesh_info_t *info);

/* Points (device SoA, mesh precision).  out_tri, out_cx/cy/cz may be NULL where the mode does not
 * write them (EDF, SDF).  EDF/CLOSEST read out_d as the current best. */
int ssdf_mesh_query_points(ssdf_mesh_t mesh, int mode, ptrdiff_t npoints, const void *x, const void *y,
                           const void *z, void *out_d, int *out_tri, void *out_cx, void *out_cy, void *out_cz,
                           unsigned flags, void *stream);

/* Allocation-free variant of ssdf_mesh_query_points.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as ssdf_mesh_query_points, but the
 * call's workspace (status words, queue, point sort) is a static device array of the library, about
 * 18 MiB on each GPU that uses it, held until the process ends; the call itself allocates nothing.
 * npoints <= SSDF_NOALLOC_MAX_POINTS, else SSDF_ERR_INVALID.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (any mesh, any stream) race, so the caller serializes them.
 * It is ordered only by `stream` (no ordering against the handle's other streams).  Signed queries
 * signed by winding numbers (SSDF_FLAG_SIGN_GWN / AUTO) need the mesh's winding-number tree built by
 * an earlier allocating query; ssdf_mesh_last_stats does not report these calls. */
#define SSDF_NOALLOC_MAX_POINTS 1048576
int ssdf_mesh_query_points_noalloc(ssdf_mesh_t mesh, int mode, ptrdiff_t npoints, const void *x, const void *y,
                                   const void *z, void *out_d, int *out_tri, void *out_cx, void *out_cy,
                                   void *out_cz, unsigned flags, void *stream);

/* Grid nodes x_i = min + i (max - min) / (n - 1), computed in the mesh precision.  Modes EDF, SDF
 * or SQEDF. */
int ssdf_mesh_query_grid(ssdf_mesh_t mesh, int mode, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, double xmin,
                         double ymin, double zmin, double xmax, double ymax, double zmax, void *out_d, int *out_tri,
                         unsigned flags, void *stream);

/* Generalized winding number w of each point (device arrays, mesh precision in, float out): 1 inside
 * a closed outward-oriented surface, 0 outside, fractional near holes.  Exact solid angles near the
 * point, order-1 far-field terms for subtrees beyond 3 radii. */
int ssdf_mesh_winding_number(ssdf_mesh_t mesh, ptrdiff_t npoints, const void *x, const void *y, const void *z,
                             float *out_w, void *stream);
/* Inside flags (w >= 1/2) at points or on grid nodes (device arrays). */
int ssdf_mesh_inside_points(ssdf_mesh_t mesh, ptrdiff_t npoints, const void *x, const void *y, const void *z,
                            uint8_t *out_inside, void *stream);
int ssdf_mesh_inside_grid(ssdf_mesh_t mesh, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, double xmin, double ymin,
                          double zmin, double xmax, double ymax, double zmax, uint8_t *out_inside, void *stream);

/* Schedule of this handle's grid queries (no counterpart in the reference; a tuning knob): the
 * brick of grid nodes one wave answers (brick_x x brick_y x brick_z = 64, powers of two; default 4 x
 * 4 x 4) and the leaf budget of its packet waves (-1: the process-wide one, ssdf_set_heavy_leaves).
 * FP32 + fixup results do not depend on either. */
int ssdf_mesh_set_schedule(ssdf_mesh_t mesh, int brick_x, int brick_y, int brick_z, int heavy_leaves);

/* Statistics of the last query on this handle (after its stream has completed): number of
 * queries re-traversed in FP64 because their candidate list overflowed, and the kernel error flag:
 * bit 1 a traversal stack overflowed, bit 2 the iteration bound was hit (malformed tree), bit 4
 * the fixup queue (max(65536, queries / 8) entries) was full and some near-tied queries kept their
 * FP32 result. */
int ssdf_mesh_last_stats(ssdf_mesh_t mesh, unsigned *fixup_queued, int *kernel_error);

/* ---------------------------------------------------------------------------------------------
 * Batches (no reference counterpart: the reference answers one mesh per call, bvh.cu:80-158).
 * Many FP32 meshes live in shared device arenas; one call answers every mesh's queries with a
 * single persistent launch (work queue over (mesh, 64-query unit) items, one range per XCD with
 * stealing).  Meshes are given concatenated: mesh i owns triangles [tri_offsets[i],
 * tri_offsets[i+1]) and vertices [vert_offsets[i], vert_offsets[i+1]), its s0/s1/s2 holding LOCAL
 * vertex indices.  Queries and outputs are device arrays.  Calls on one batch must not overlap.
 * --------------------------------------------------------------------------------------------- */
typedef struct ssdf_batch *ssdf_batch_t;

int ssdf_batch_create(int nmeshes, const ptrdiff_t *tri_offsets, const ptrdiff_t *vert_offsets, const int *s0,
                      const int *s1, const int *s2, const float *sx, const float *sy, const float *sz, int device_ptrs,
                      int with_normals, void *stream, ssdf_batch_t *out);
/* The creation stream orders the batch's own allocations and must outlive its use; destroy waits
 * for the device instead and is safe after the caller has destroyed its streams. */
int ssdf_batch_destroy(ssdf_batch_t batch);
/* Rebuilds every mesh's LBVH (and normals) from the batch's vertex arena, in place. */
int ssdf_batch_rebuild(ssdf_batch_t batch, int with_normals, void *stream);
/* Mesh i's grid: dims[3i..3i+2] nodes over boxes[6i..6i+5] (min xyz, max xyz), x fastest, its
 * results at out_d/out_tri + out_offsets[i].  mode EDF (min-accumulate) or SDF (overwrite).
 * The persistent launch hands out the work units (one wave's 64 queries of one mesh) costliest
 * first, as the batch's previous query of the same plan measured them (the first query of a plan:
 * natural order); results do not depend on the order.  Calls on one batch must not overlap on
 * different streams (they share its work queues). */
int ssdf_batch_query_grids(ssdf_batch_t batch, int mode, const ptrdiff_t *dims, const double *boxes,
                           const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream);
/* Mesh i's points: [point_offsets[i], point_offsets[i+1]) of x/y/z, results at the same positions. */
int ssdf_batch_query_points(ssdf_batch_t batch, int mode, const ptrdiff_t *point_offsets, const float *x,
                            const float *y, const float *z, float *out_d, int *out_tri, float *out_cx, float *out_cy,
                            float *out_cz, unsigned flags, void *stream);
/* Persistent-launch schedule: blocks per CU (grid size) and work units per queue grab; 0 = default. */
int ssdf_batch_set_schedule(ssdf_batch_t batch, int blocks_per_cu, int units_per_grab);
/* Per-mesh schedule of grid queries: the brick of grid nodes one wave answers (brick_x x brick_y x
 * brick_z = 64, powers of two; default 4 x 4 x 4) and the leaf budget of the mesh's packet waves
 * before they hand their queries to the per-query kernel (-1: the process-wide budget,
 * ssdf_set_heavy_leaves; 0: never).  FP32 + fixup results do not depend on either. */
int ssdf_batch_set_mesh_schedule(ssdf_batch_t batch, int mesh, int brick_x, int brick_y, int brick_z,
                                 int heavy_leaves);
int ssdf_batch_get_mesh_schedule(ssdf_batch_t batch, int mesh, int *brick_x, int *brick_y, int *brick_z,
                                 int *heavy_leaves);
/* Autotuner of grid queries.  Per mesh, the brick (4x4x4, 8x4x2, 4x8x2, 2x4x8, 8x8x1), chosen from
 * one query per candidate in which every mesh's waves clock their own cycles (packet and per-query
 * kernels), so each mesh keeps its cheapest.  Then the leaf budget (4096, 8192, 16384) and the launch
 * (blocks per CU x units per grab), timed with events under the per-mesh bricks.  The schedule found
 * is kept only if it beats the default one (4x4x4 bricks, process-wide budget, 8 blocks per CU, one
 * unit per grab) over three alternating timed runs of each; otherwise the default is restored.  The
 * kept schedule stays in the batch (ssdf_batch_get_mesh_schedule reads it); reports the launch
 * schedule and its time.  Runs the query 64 times. */
int ssdf_batch_autotune(ssdf_batch_t batch, int mode, const ptrdiff_t *dims, const double *boxes,
                        const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream,
                        int *best_blocks_per_cu, int *best_units_per_grab, float *best_ms);
/* Captures [rebuild of every mesh +] the grid query into a hipGraph (stream must be a created
 * stream); ssdf_batch_replay launches it.  The graph holds kernel launches only, and owns its
 * descriptors, fixup queue and unit order, so eager queries of other plans may run between
 * replays. */
int ssdf_batch_capture(ssdf_batch_t batch, int rebuild, int mode, const ptrdiff_t *dims, const double *boxes,
                       const ptrdiff_t *out_offsets, float *out_d, int *out_tri, unsigned flags, void *stream);
int ssdf_batch_replay(ssdf_batch_t batch, void *stream);
/* As ssdf_mesh_last_stats, for the batch's last query (synchronous read). */
int ssdf_batch_last_stats(ssdf_batch_t batch, unsigned *fixup_queued, int *kernel_error);
/* mesh = -1: totals of the batch (triangles, vertices, deepest tree, device bytes). */
int ssdf_batch_info(ssdf_batch_t batch, int mesh, ssdf_mesh_info_t *info);

/* ---------------------------------------------------------------------------------------------
 * Serving (no reference counterpart).  A worker thread keeps registered meshes resident (built on
 * first use on a stream of their own, least recently used idle meshes evicted when the HBM budget
 * needs room), admits requests in arrival order when their device buffer (20 B per query of
 * inputs and outputs plus the request's own query workspace) fits the budget, and runs them as soon
 * as they are admitted on a pool of streams (upload, query, download), requests on the same mesh
 * included.  Requests take HOST arrays that must stay valid until the request completes; pinned
 * memory (ssdf_host_alloc) lets the copies run asynchronously.
 * --------------------------------------------------------------------------------------------- */
typedef struct ssdf_server *ssdf_server_t;

typedef struct {
    unsigned long long served, queries, errors, mesh_loads, evictions;
    size_t pending, inflight, bytes_in_use, peak_bytes;
    double p50_s, p99_s, max_s; /* latency: arrival to results in host memory */
    double throughput_qps;      /* queries / (last completion - first arrival) */
} ssdf_server_stats_t;

/* max_inflight: requests on the device at once (0 = 16). */
int ssdf_server_create(size_t hbm_budget_bytes, int max_inflight, ssdf_server_t *out);
/* Registers an FP32 mesh (host arrays kept by the caller for the server's lifetime). */
int ssdf_server_add_mesh(ssdf_server_t server, int mesh_id, ptrdiff_t nselements, const int *s0, const int *s1,
                         const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy, const float *sz);
/* The HBM the server charges against its budget for mesh mesh_id while it is resident (an upper
 * bound of what its tree, normals, winding-number tree and build workspace can allocate) and for a
 * request of npoints points (its inputs, outputs and query workspace); either pointer may be NULL. */
int ssdf_server_charge(ssdf_server_t server, int mesh_id, ptrdiff_t npoints, size_t *mesh_bytes,
                       size_t *request_bytes);
/* Seconds on the server's clock (its creation is 0). */
int ssdf_server_clock(ssdf_server_t server, double *now);
/* mode EDF (out_d and out_tri read as the current best, kept where not improved) or SDF (both
 * overwritten); arrival: timestamp on the server clock for the
 * latency (<= 0: now).  Returns a ticket for ssdf_server_wait / ssdf_server_poll. */
int ssdf_server_submit(ssdf_server_t server, int mesh_id, int mode, ptrdiff_t npoints, const float *x,
                       const float *y, const float *z, float *out_d, int *out_tri, unsigned flags, double arrival,
                       uint64_t *ticket);
/* Blocks until the request completed; releases the ticket. */
int ssdf_server_wait(ssdf_server_t server, uint64_t ticket, double *latency_s);
int ssdf_server_poll(ssdf_server_t server, uint64_t ticket, int *done);
int ssdf_server_stats(ssdf_server_t server, ssdf_server_stats_t *stats);
/* Drains the queue, then stops the worker and frees the resident meshes. */
int ssdf_server_destroy(ssdf_server_t server);
/* Pinned host memory for request buffers. */
int ssdf_host_alloc(void **ptr, size_t bytes);
int ssdf_host_free(void *ptr);

/* ---------------------------------------------------------------------------------------------
 * Device plumbing for callers without a GPU runtime of their own (tests, benchmarks, FFI).
 * --------------------------------------------------------------------------------------------- */
int ssdf_device_count(int *count);
int ssdf_device_malloc(void **ptr, size_t bytes);
int ssdf_device_free(void *ptr);
int ssdf_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ssdf_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Streaming device-to-device copy by a plain kernel (16-byte aligned, multiple of 16 bytes). */
int ssdf_device_copy(void *dst, const void *src, size_t bytes, void *stream);
int ssdf_stream_create(void **stream);
int ssdf_stream_destroy(void *stream);
int ssdf_stream_synchronize(void *stream);
/* Device-side timing on a stream: record start, run work, record stop, read milliseconds. */
int ssdf_event_create(void **event);
int ssdf_event_destroy(void *event);
int ssdf_event_record(void *event, void *stream);
int ssdf_event_elapsed_ms(void *start, void *stop, float *ms);

/* GPU brute force (test reference): exact FP64 argmin over every triangle of a resident FP32
 * mesh, lowest index on ties; out_d2 and out_tri are device arrays. */
int ssdf_mesh_brute_points(ssdf_mesh_t mesh, ptrdiff_t npoints, const float *x, const float *y, const float *z,
                           double *out_d2, int *out_tri, void *stream);

/* LBVH leaf order (test reference): the original triangle index of each leaf, in leaf order, into
 * the device array out_tri[nselements].  Leaves are the triangles stably sorted by the 63-bit
 * Morton code of their box centre, so this exposes the build's radix sort to tests. */
int ssdf_mesh_leaf_order(ssdf_mesh_t mesh, int *out_tri, void *stream);

/* Leaf budget of a packet wave (no counterpart in the reference; a tuning knob of this library).
 * A wave of 64 queries that has evaluated this many leaves hands its unfinished queries to a
 * per-query kernel (one wave per query) with the bound found so far; results are unchanged (FP32 +
 * fixup: bit-identical; plain FP32: the same distances, ties may name another triangle).
 * leaves >= 0 sets it process-wide (0: never hand over) and returns the previous value; leaves < 0
 * only returns the current one.  Initial value 8192, or SSDF_HEAVY_LEAVES from the environment. */
int ssdf_set_heavy_leaves(int leaves);

const char *ssdf_last_error(void);
const char *ssdf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SSDF_C_H */

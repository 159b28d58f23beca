/*
 * oracle/oracle.h — CPU oracle for the closest-point / signed-distance path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of zulianp/ssdf's algorithms (citations per function in orc_real.h and
 * oracle.c).  It is the checker for the HIP library, the CPU baseline that bench.py times, and the
 * source of the canonical work counts for the roofline.  It is never linked into the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so.
 *
 * Parity pinning (what fixes each function independently of this restatement) is listed in
 * tests/test_oracle.py; functions without such a check are marked "parity unpinned" below.
 */
#ifndef SSDF_ORACLE_H
#define SSDF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitives (orc_real.h), float and double ---- */
float orc_pt_tri_d2_f(const float *p, const float *a, const float *b, const float *c);
double orc_pt_tri_d2_d(const double *p, const double *a, const double *b, const double *c);
void orc_pt_tri_cp_f(const float *p, const float *a, const float *b, const float *c, float *out);
void orc_pt_tri_cp_d(const double *p, const double *a, const double *b, const double *c, double *out);
float orc_pt_box_d2_f(float px, float py, float pz, float minx, float maxx, float miny, float maxy, float minz,
                      float maxz);
double orc_pt_box_d2_d(double px, double py, double pz, double minx, double maxx, double miny, double maxy,
                       double minz, double maxz);
float orc_pt_box_sdist_f(float px, float py, float pz, float minx, float maxx, float miny, float maxy, float minz,
                         float maxz);
double orc_pt_box_sdist_d(double px, double py, double pz, double minx, double maxx, double miny, double maxy,
                          double minz, double maxz);
/* compute_aabb / all_points_aabb_signed_distance (serial/edf/edf.hpp:500-557); box6 = xmin, xmax,
 * ymin, ymax, zmin, zmax */
void orc_compute_aabb_f(ptrdiff_t n, const float *x, const float *y, const float *z, float *box6);
void orc_compute_aabb_d(ptrdiff_t n, const double *x, const double *y, const double *z, double *box6);
void orc_all_points_aabb_sdist_f(ptrdiff_t n, const float *x, const float *y, const float *z, const float *box6,
                                 float *out);
void orc_all_points_aabb_sdist_d(ptrdiff_t n, const double *x, const double *y, const double *z, const double *box6,
                                 double *out);

/* ---- the reference's two CPU engines, restated (orc_real.h) ---- */
int orc_edf_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nsel, const int *s0,
              const int *s1, const int *s2, ptrdiff_t nsp, const float *sx, const float *sy, const float *sz,
              float *out);
int orc_edf_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel, const int *s0,
              const int *s1, const int *s2, ptrdiff_t nsp, const double *sx, const double *sy, const double *sz,
              double *out);
int orc_celllist_f(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nsel, const int *s0,
                   const int *s1, const int *s2, ptrdiff_t nsp, const float *sx, const float *sy, const float *sz,
                   float *out);
int orc_celllist_d(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel,
                   const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const double *sx, const double *sy,
                   const double *sz, double *out);

/* ---- FP64 ground truth (oracle.c) ----
 * Exact argmin of the FP64 squared distance (pt_tri_d2 in double) over all triangles, the lowest
 * triangle index winning exact ties.  init_d2 (may be NULL = +inf) seeds the best value and the
 * index stays -1 unless a triangle is strictly below it.  out_ntie counts the triangles whose d2
 * equals the minimum bit for bit; out_d2second is the smallest d2 of any other triangle. */
void orc_brute_argmin(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel,
                      const int *s0, const int *s1, const int *s2, const double *sx, const double *sy,
                      const double *sz, const double *init_d2, double *out_d2, int *out_tri, int *out_ntie,
                      double *out_d2second);

/* The same argmin (identical results) through a CPU BVH with a conservative pruning margin, for
 * inputs where brute force is too slow. */
/* FP64 squared distance of point i to triangle tri[i] (pt_tri_d2, serial/ssdf.hpp:50-121), for
 * judging a reported triangle against the argmin: equal d2 means an exact tie. */
void orc_pair_d2(ptrdiff_t npoints, const double *x, const double *y, const double *z, const int *tri,
                 const int *s0, const int *s1, const int *s2, const double *sx, const double *sy, const double *sz,
                 double *out_d2);
void orc_bvh_argmin(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel,
                    const int *s0, const int *s1, const int *s2, ptrdiff_t nsp, const double *sx, const double *sy,
                    const double *sz, double *out_d2, int *out_tri);

/* Pseudonormals: face (3T), edge (9T, local edge e = vertex e -> (e+1)%3), vertex (3V); follows
 * serial/bvh/bvh.cpp:1669-1721.  Returns the number of edges that are not shared by exactly two
 * faces (boundary or non-manifold). */
ptrdiff_t orc_pseudonormals(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, ptrdiff_t nsp,
                            const double *sx, const double *sy, const double *sz, double *face_n, double *edge_n,
                            double *vertex_n);

/* Sign of the distance at q given the closest triangle t: closest point in FP64, barycentric region
 * with tol 1e-4, pseudonormal, sign of (q - c).n; follows serial/bvh/bvh.cpp:1734-1792.  Writes
 * +1/-1 to out_sign, the region code (0 face, 1 edge, 2 vertex) and the dot product. */
void orc_sign(ptrdiff_t npoints, const double *x, const double *y, const double *z, const int *tri, const int *s0,
              const int *s1, const int *s2, const double *sx, const double *sy, const double *sz,
              const double *face_n, const double *edge_n, const double *vertex_n, int *out_sign, int *out_region,
              double *out_dot);

/* Generalized winding number (sum of signed solid angles / 4 pi; Van Oosterom and Strackee 1983),
 * brute force in FP64.  Not in the reference: it is this engine's sign fallback for open meshes. */
void orc_gwn(ptrdiff_t npoints, const double *x, const double *y, const double *z, ptrdiff_t nsel, const int *s0,
             const int *s1, const int *s2, const double *sx, const double *sy, const double *sz, double *out_w);

/* Canonical LBVH replay (FP32): 63-bit Morton codes of triangle box centres, stable sort by code,
 * Karras hierarchy, leaf size 1, nearest-child-first traversal with strict '<' pruning from +inf.
 * Counts the child-box tests and point-triangle tests per query: the algorithmic work that the
 * roofline's flop count is built from. */
void orc_lbvh_counts(ptrdiff_t npoints, const float *x, const float *y, const float *z, ptrdiff_t nsel,
                     const int *s0, const int *s1, const int *s2, const float *sx, const float *sy, const float *sz,
                     int64_t *out_nbox, int64_t *out_ntri, int *out_maxdepth, float *out_d2);

/* Leaf order of the canonical LBVH (triangle index per leaf: stable sort by Morton code). */
/* Replay of the GPU's wave-coherent packet walk over the 4x4x4 bricks of an nx x ny x nz grid on
 * [lo, hi] (every `every`-th brick): out[0..4] = bricks, node visits, pops, leaf references, triangle
 * evaluations (each performed by all 64 lanes).  Parity unpinned: a model of this build's kernel for
 * the roofline's useful-work fraction, not a reference algorithm. */
void orc_packet_counts(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, const float *sx, const float *sy,
                       const float *sz, int nx, int ny, int nz, const float *lo, const float *hi, int every,
                       int64_t *out);
void orc_lbvh_order(ptrdiff_t nsel, const int *s0, const int *s1, const int *s2, const float *sx, const float *sy,
                    const float *sz, int *out_order);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

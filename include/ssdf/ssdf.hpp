// ssdf/ssdf.hpp — drop-in C++ interface of the MI355X engine, mirroring zulianp/ssdf's templates.
//
// Same names, template parameters <G geometry, T output, I index>, argument order and contracts as
// serial/ssdf.hpp:447-460 (edf), 642-655 (edf_celllist), 965-978 (edf_select) and
// serial/bvh/bvh.hpp:19-51 (edf_bvh, closest_point_bvh), 209-227 (sample_sdf_bvh).  Every engine
// runs on the GPU through libssdf_amd.so's C ABI (ssdf_c.h); link with -lssdf_amd.
//
// Supported instantiations (the reference's: serial/ssdf.cpp:4-30, bvh.cpp:1162-1256, 1906-1942):
//   G=float,T=float,I=int   FP32 path          G=double,T=double,I=int  exact FP64 path
//   G=float,T=double,I=int  FP32 path, output widened (edf_bvh / closest_point_bvh <float,double>)
// Return 0 on success (as the reference), a non-zero SSDF_ERR_* code on a device error.
#pragma once

#include <stddef.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "../ssdf_c.h"

#ifndef SSDF_RESTRICT
#define SSDF_RESTRICT __restrict__
#endif

namespace ssdf {

namespace detail {
template <typename G, typename T, typename I>
struct check_types {
    static_assert(std::is_same<I, int>::value, "ssdf_amd: triangle indices are int");
    static_assert((std::is_same<G, float>::value && (std::is_same<T, float>::value || std::is_same<T, double>::value)) ||
                      (std::is_same<G, double>::value && std::is_same<T, double>::value),
                  "ssdf_amd: supported <G,T> are <float,float>, <float,double>, <double,double>");
};

inline int edf_dispatch(ptrdiff_t n, const float *x, const float *y, const float *z, ptrdiff_t ns, const int *s0,
                        const int *s1, const int *s2, ptrdiff_t nv, const float *sx, const float *sy, const float *sz,
                        float *out) {
    return ssdf_edf_f(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, out);
}
inline int edf_dispatch(ptrdiff_t n, const double *x, const double *y, const double *z, ptrdiff_t ns, const int *s0,
                        const int *s1, const int *s2, ptrdiff_t nv, const double *sx, const double *sy,
                        const double *sz, double *out) {
    return ssdf_edf_d(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, out);
}
// <float, double>: FP32 computation, min-accumulated into the double output
inline int edf_dispatch(ptrdiff_t n, const float *x, const float *y, const float *z, ptrdiff_t ns, const int *s0,
                        const int *s1, const int *s2, ptrdiff_t nv, const float *sx, const float *sy, const float *sz,
                        double *out) {
    std::vector<float> tmp(n > 0 ? (size_t)n : 0);
    for (ptrdiff_t i = 0; i < n; i++) tmp[i] = out[i] < 3.4028234663852886e38 ? (float)out[i] : 3.4028234663852886e38f;
    const int rc = ssdf_edf_f(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, tmp.data());
    if (rc) return rc;
    for (ptrdiff_t i = 0; i < n; i++) out[i] = std::min(out[i], (double)tmp[i]);
    return 0;
}
}  // namespace detail

/// Unsigned point-to-surface distance (serial/ssdf.hpp:427-460): out[] is in/out, min-accumulated;
/// an empty surface leaves it untouched.
template <typename G, typename T, typename I>
int edf(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
        const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
        const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
        const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz,
        T *const SSDF_RESTRICT out) {
    (void)sizeof(detail::check_types<G, T, I>);
    return detail::edf_dispatch(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

/// serial/ssdf.hpp:642-655: same contract as edf (the cell list is a CPU acceleration structure;
/// here the LBVH serves every engine).
template <typename G, typename T, typename I>
int edf_celllist(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
                 const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
                 const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
                 const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz,
                 T *const SSDF_RESTRICT out) {
    return edf<G, T, I>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

/// serial/ssdf.hpp:965-1016 (the C ABI's ssdf_edf_f/d call this).
template <typename G, typename T, typename I>
int edf_select(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
               const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
               const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
               const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz,
               T *const SSDF_RESTRICT out) {
    return edf<G, T, I>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

/// serial/bvh/bvh.hpp:19-32.
template <typename G, typename T, typename I>
int edf_bvh(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
            const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
            const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
            const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz,
            T *const SSDF_RESTRICT out) {
    return edf<G, T, I>(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

/// serial/bvh/bvh.hpp:34-51, bvh.cpp:125-197: distance, closest point and closest triangle, written
/// only where the new distance is strictly below outd.  FP32 uses the FP64 fixup, so the triangle
/// index is the FP64 argmin (lowest index on exact ties).
inline int closest_point_bvh_impl(ptrdiff_t n, const float *x, const float *y, const float *z, ptrdiff_t ns,
                                  const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const float *sx,
                                  const float *sy, const float *sz, float *outd, float *ox, float *oy, float *oz,
                                  int *ot) {
    return ssdf_closest_point_f(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, outd, ox, oy, oz, ot, SSDF_FLAG_FIXUP);
}
inline int closest_point_bvh_impl(ptrdiff_t n, const double *x, const double *y, const double *z, ptrdiff_t ns,
                                  const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const double *sx,
                                  const double *sy, const double *sz, double *outd, double *ox, double *oy,
                                  double *oz, int *ot) {
    return ssdf_closest_point_d(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, outd, ox, oy, oz, ot);
}
inline int closest_point_bvh_impl(ptrdiff_t n, const float *x, const float *y, const float *z, ptrdiff_t ns,
                                  const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const float *sx,
                                  const float *sy, const float *sz, double *outd, float *ox, float *oy, float *oz,
                                  int *ot) {
    std::vector<float> d(n > 0 ? (size_t)n : 0);
    for (ptrdiff_t i = 0; i < n; i++) d[i] = outd[i] < 3.4028234663852886e38 ? (float)outd[i] : 3.4028234663852886e38f;
    const int rc = closest_point_bvh_impl(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, d.data(), ox, oy, oz, ot);
    if (rc) return rc;
    for (ptrdiff_t i = 0; i < n; i++)
        if ((double)d[i] < outd[i]) outd[i] = d[i];
    return 0;
}

template <typename G, typename T, typename I>
int closest_point_bvh(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
                      const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
                      const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
                      const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy,
                      const G *const SSDF_RESTRICT sz, T *const SSDF_RESTRICT outd, G *const SSDF_RESTRICT outx,
                      G *const SSDF_RESTRICT outy, G *const SSDF_RESTRICT outz, I *const SSDF_RESTRICT outtri) {
    (void)sizeof(detail::check_types<G, T, I>);
    return closest_point_bvh_impl(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, outd, outx, outy,
                                  outz, outtri);
}

/// serial/bvh/bvh.hpp:203-227: signed distance on the nodes of a grid, x fastest, positive where
/// the right-handed normals point (angle-weighted pseudonormal sign).  Overwrites out.
inline int sample_sdf_bvh_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv,
                               const float *sx, const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny,
                               ptrdiff_t nz, float x0, float y0, float z0, float x1, float y1, float z1, float *out) {
    return ssdf_sample_sdf_f(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out, nullptr,
                             SSDF_FLAG_FIXUP);
}
inline int sample_sdf_bvh_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv,
                               const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                               ptrdiff_t nz, double x0, double y0, double z0, double x1, double y1, double z1,
                               double *out) {
    return ssdf_sample_sdf_d(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out, nullptr);
}

template <typename G, typename T, typename I>
int sample_sdf_bvh(const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0, const I *const SSDF_RESTRICT s1,
                   const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints, const G *const SSDF_RESTRICT sx,
                   const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz, const ptrdiff_t nx,
                   const ptrdiff_t ny, const ptrdiff_t nz, const G xmin, const G ymin, const G zmin, const G xmax,
                   const G ymax, const G zmax, T *const SSDF_RESTRICT out) {
    static_assert(std::is_same<G, T>::value, "ssdf_amd: sample_sdf_bvh is instantiated for <float,float> and "
                                             "<double,double> (serial/bvh/bvh.cpp:1906-1942)");
    (void)sizeof(detail::check_types<G, T, I>);
    return sample_sdf_bvh_impl(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                               ymax, zmax, out);
}

namespace detail {
inline int sqedf_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const float *sx,
                      const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, float x0, float y0,
                      float z0, float x1, float y1, float z1, float *out) {
    return ssdf_sample_sqedf_f(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out,
                               SSDF_FLAG_FIXUP);
}
inline int sqedf_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const double *sx,
                      const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, double x0,
                      double y0, double z0, double x1, double y1, double z1, double *out) {
    return ssdf_sample_sqedf_d(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out);
}
}  // namespace detail

/// serial/bvh/bvh.hpp:183-201: squared unsigned distance on grid nodes, out in/out (min-accumulated).
template <typename G, typename T, typename I>
int sample_sqedf_bvh(const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0, const I *const SSDF_RESTRICT s1,
                     const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints, const G *const SSDF_RESTRICT sx,
                     const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz, const ptrdiff_t nx,
                     const ptrdiff_t ny, const ptrdiff_t nz, const G xmin, const G ymin, const G zmin, const G xmax,
                     const G ymax, const G zmax, T *const SSDF_RESTRICT out) {
    static_assert(std::is_same<G, T>::value, "ssdf_amd: sample_sqedf_bvh is instantiated for <float,float> and "
                                             "<double,double>");
    (void)sizeof(detail::check_types<G, T, I>);
    return detail::sqedf_impl(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax, ymax,
                              zmax, out);
}

namespace detail {
inline int inside_points_impl(ptrdiff_t n, const float *x, const float *y, const float *z, ptrdiff_t ns,
                              const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const float *sx,
                              const float *sy, const float *sz, uint8_t *out) {
    return ssdf_points_inside_f(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, out);
}
inline int inside_points_impl(ptrdiff_t n, const double *x, const double *y, const double *z, ptrdiff_t ns,
                              const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const double *sx,
                              const double *sy, const double *sz, uint8_t *out) {
    return ssdf_points_inside_d(n, x, y, z, ns, s0, s1, s2, nv, sx, sy, sz, out);
}
inline int inside_grid_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv, const float *sx,
                            const float *sy, const float *sz, ptrdiff_t nx, ptrdiff_t ny, ptrdiff_t nz, float x0,
                            float y0, float z0, float x1, float y1, float z1, uint8_t *out) {
    return ssdf_sample_inside_f(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out);
}
inline int inside_grid_impl(ptrdiff_t ns, const int *s0, const int *s1, const int *s2, ptrdiff_t nv,
                            const double *sx, const double *sy, const double *sz, ptrdiff_t nx, ptrdiff_t ny,
                            ptrdiff_t nz, double x0, double y0, double z0, double x1, double y1, double z1,
                            uint8_t *out) {
    return ssdf_sample_inside_d(ns, s0, s1, s2, nv, sx, sy, sz, nx, ny, nz, x0, y0, z0, x1, y1, z1, out);
}
}  // namespace detail

/// serial/bvh/bvh.hpp:53-66: out[i] = 1 inside the surface, else 0 (generalized winding number >= 1/2,
/// the ray-parity decision of the reference on closed surfaces).
template <typename G, typename I>
int points_inside_bvh(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
                      const G *const SSDF_RESTRICT z, const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0,
                      const I *const SSDF_RESTRICT s1, const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints,
                      const G *const SSDF_RESTRICT sx, const G *const SSDF_RESTRICT sy,
                      const G *const SSDF_RESTRICT sz, uint8_t *const SSDF_RESTRICT out) {
    (void)sizeof(detail::check_types<G, G, I>);
    return detail::inside_points_impl(npoints, x, y, z, nselements, s0, s1, s2, nspoints, sx, sy, sz, out);
}

/// serial/bvh/bvh.hpp:229-247: inside flags on the nodes of a grid, x fastest.
template <typename G, typename I>
int sample_inside_bvh(const ptrdiff_t nselements, const I *const SSDF_RESTRICT s0, const I *const SSDF_RESTRICT s1,
                      const I *const SSDF_RESTRICT s2, const ptrdiff_t nspoints, const G *const SSDF_RESTRICT sx,
                      const G *const SSDF_RESTRICT sy, const G *const SSDF_RESTRICT sz, const ptrdiff_t nx,
                      const ptrdiff_t ny, const ptrdiff_t nz, const G xmin, const G ymin, const G zmin, const G xmax,
                      const G ymax, const G zmax, uint8_t *const SSDF_RESTRICT out) {
    (void)sizeof(detail::check_types<G, G, I>);
    return detail::inside_grid_impl(nselements, s0, s1, s2, nspoints, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax,
                                    ymax, zmax, out);
}

namespace detail {
inline int compute_aabb_impl(ptrdiff_t n, const float *x, const float *y, const float *z, float *b) {
    return ssdf_compute_aabb_f(n, x, y, z, &b[0], &b[1], &b[2], &b[3], &b[4], &b[5]);
}
inline int compute_aabb_impl(ptrdiff_t n, const double *x, const double *y, const double *z, double *b) {
    return ssdf_compute_aabb_d(n, x, y, z, &b[0], &b[1], &b[2], &b[3], &b[4], &b[5]);
}
inline int box_sdist_impl(ptrdiff_t n, const float *x, const float *y, const float *z, const float *b, float *out) {
    return ssdf_all_points_aabb_signed_distance_f(n, x, y, z, b[0], b[1], b[2], b[3], b[4], b[5], out);
}
inline int box_sdist_impl(ptrdiff_t n, const double *x, const double *y, const double *z, const double *b,
                          double *out) {
    return ssdf_all_points_aabb_signed_distance_d(n, x, y, z, b[0], b[1], b[2], b[3], b[4], b[5], out);
}
// T(x[i]) for every point, as the reference converts before computing (serial/edf/edf.hpp:515-517)
template <typename T, typename G>
std::vector<T> widen(ptrdiff_t n, const G *v) {
    return std::vector<T>(v, v + (n > 0 ? n : 0));
}
}  // namespace detail

/// serial/edf/edf.hpp:500-539: bounds of npoints >= 1 points, computed in T on the GPU.
template <typename G, typename T>
int compute_aabb(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x, const G *const SSDF_RESTRICT y,
                 const G *const SSDF_RESTRICT z, T *const SSDF_RESTRICT xmin, T *const SSDF_RESTRICT xmax,
                 T *const SSDF_RESTRICT ymin, T *const SSDF_RESTRICT ymax, T *const SSDF_RESTRICT zmin,
                 T *const SSDF_RESTRICT zmax) {
    T b[6];
    int rc;
    if constexpr (std::is_same<G, T>::value) {
        rc = detail::compute_aabb_impl(npoints, x, y, z, b);
    } else {
        const auto wx = detail::widen<T>(npoints, x), wy = detail::widen<T>(npoints, y),
                   wz = detail::widen<T>(npoints, z);
        rc = detail::compute_aabb_impl(npoints, wx.data(), wy.data(), wz.data(), b);
    }
    if (rc) return rc;
    *xmin = b[0], *xmax = b[1], *ymin = b[2], *ymax = b[3], *zmin = b[4], *zmax = b[5];
    return 0;
}

/// serial/edf/edf.hpp:541-557: signed distance of every point to the box (negative inside), in T.
template <typename G, typename T>
int all_points_aabb_signed_distance(const ptrdiff_t npoints, const G *const SSDF_RESTRICT x,
                                    const G *const SSDF_RESTRICT y, const G *const SSDF_RESTRICT z, const T minx,
                                    const T maxx, const T miny, const T maxy, const T minz, const T maxz,
                                    T *const SSDF_RESTRICT out) {
    const T b[6] = {minx, maxx, miny, maxy, minz, maxz};
    if constexpr (std::is_same<G, T>::value) {
        return detail::box_sdist_impl(npoints, x, y, z, b, out);
    } else {
        const auto wx = detail::widen<T>(npoints, x), wy = detail::widen<T>(npoints, y),
                   wz = detail::widen<T>(npoints, z);
        return detail::box_sdist_impl(npoints, wx.data(), wy.data(), wz.data(), b, out);
    }
}

}  // namespace ssdf

// oracle/ref_edf.cpp — C-ABI shim over the REAL reference header serial/edf/edf.hpp (test
// infrastructure).  Built by oracle/Makefile straight from /root/reference into oracle/_ref/; the
// header is self-contained (standard headers only), unlike serial/ssdf.hpp, which needs the
// cmake-generated ssdf_config.hpp and is therefore not built here.  Used only to pin the oracle's
// box-distance primitives against the reference's own code.
#include "edf/edf.hpp"

extern "C" {
double ref_point_aabb_distance_squared_d(double px, double py, double pz, double minx, double maxx, double miny,
                                         double maxy, double minz, double maxz) {
    return ssdf::point_aabb_distance_squared(px, py, pz, minx, maxx, miny, maxy, minz, maxz);
}
float ref_point_aabb_distance_squared_f(float px, float py, float pz, float minx, float maxx, float miny,
                                        float maxy, float minz, float maxz) {
    return ssdf::point_aabb_distance_squared(px, py, pz, minx, maxx, miny, maxy, minz, maxz);
}
double ref_point_aabb_signed_distance_d(double px, double py, double pz, double minx, double maxx, double miny,
                                        double maxy, double minz, double maxz) {
    return ssdf::point_aabb_signed_distance(px, py, pz, minx, maxx, miny, maxy, minz, maxz);
}
float ref_point_aabb_signed_distance_f(float px, float py, float pz, float minx, float maxx, float miny, float maxy,
                                       float minz, float maxz) {
    return ssdf::point_aabb_signed_distance(px, py, pz, minx, maxx, miny, maxy, minz, maxz);
}
double ref_point_triangle_barycentric_d(const double *q, const double *a, const double *b, const double *c,
                                        double *w) {
    return ssdf::point_triangle_barycentric(q[0], q[1], q[2], a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                                            &w[0], &w[1], &w[2]);
}
float ref_point_triangle_barycentric_f(const float *q, const float *a, const float *b, const float *c, float *w) {
    return ssdf::point_triangle_barycentric(q[0], q[1], q[2], a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                                            &w[0], &w[1], &w[2]);
}
void ref_all_points_aabb_signed_distance_d(ptrdiff_t n, const double *x, const double *y, const double *z,
                                           double *box6, double *out) {
    ssdf::compute_aabb(n, x, y, z, &box6[0], &box6[1], &box6[2], &box6[3], &box6[4], &box6[5]);
    ssdf::all_points_aabb_signed_distance(n, x, y, z, box6[0], box6[1], box6[2], box6[3], box6[4], box6[5], out);
}
}

// A reference caller's view of the engine: the single-triangle case of the reference's
// tests/ssdf_test.cpp and the seeded 50-triangle case of tests/ssdf_brute_force_test.cpp, written
// against include/ssdf/ssdf.hpp exactly as they are against serial/ssdf.hpp.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ssdf/ssdf.hpp"

#define CHECK(...)                                                         \
    do {                                                                   \
        if (!(__VA_ARGS__)) {                                                      \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #__VA_ARGS__); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

template <typename T>
static T brute(const T px, const T py, const T pz, const std::vector<T> &sx, const std::vector<T> &sy,
               const std::vector<T> &sz, const std::vector<int> &s0, const std::vector<int> &s1,
               const std::vector<int> &s2) {
    // squared distance by plane projection / edge clamping, independent of the engine's formula
    T best = 1e300 > 1e30 ? T(1e30) : T(1e30);
    for (size_t t = 0; t < s0.size(); t++) {
        const T a[3] = {sx[s0[t]], sy[s0[t]], sz[s0[t]]}, b[3] = {sx[s1[t]], sy[s1[t]], sz[s1[t]]},
                c[3] = {sx[s2[t]], sy[s2[t]], sz[s2[t]]};
        auto seg = [&](const T *u, const T *v) {
            T d[3] = {v[0] - u[0], v[1] - u[1], v[2] - u[2]}, w[3] = {px - u[0], py - u[1], pz - u[2]};
            T l = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            T s = l > 0 ? (w[0] * d[0] + w[1] * d[1] + w[2] * d[2]) / l : 0;
            s = s < 0 ? 0 : (s > 1 ? 1 : s);
            T r[3] = {w[0] - s * d[0], w[1] - s * d[1], w[2] - s * d[2]};
            return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        };
        T e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        T n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        T nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        T d2 = std::min(seg(a, b), std::min(seg(b, c), seg(c, a)));
        if (nn > 0) {
            T h = ((px - a[0]) * n[0] + (py - a[1]) * n[1] + (pz - a[2]) * n[2]) / nn;
            T q[3] = {px - h * n[0], py - h * n[1], pz - h * n[2]};
            auto side = [&](const T *u, const T *v) {
                T p1[3] = {u[0] - q[0], u[1] - q[1], u[2] - q[2]}, p2[3] = {v[0] - q[0], v[1] - q[1], v[2] - q[2]};
                return (p1[1] * p2[2] - p1[2] * p2[1]) * n[0] + (p1[2] * p2[0] - p1[0] * p2[2]) * n[1] +
                       (p1[0] * p2[1] - p1[1] * p2[0]) * n[2];
            };
            if (side(b, c) >= 0 && side(c, a) >= 0 && side(a, b) >= 0) d2 = h * h * nn;
        }
        best = std::min(best, d2);
    }
    return std::sqrt(best);
}

int main() {
    {  // one triangle, five points, out = 1111111
        std::vector<float> sx = {0.0f, 1.0f, 0.5f}, sy = {0.0f, 0.0f, 1.0f}, sz = {0.0f, 0.0f, 0.0f};
        std::vector<int> s0 = {0}, s1 = {1}, s2 = {2};
        std::vector<float> x = {0.5f, 0.0f, 1.0f, 0.5f, 0.25f}, y = {0.5f, 0.0f, 0.0f, 0.5f, 0.25f},
                           z = {0.0f, 0.0f, 0.0f, 1.0f, 0.0f}, out(5, 1111111.0f);
        CHECK(ssdf::edf_select<float, float, int>(5, x.data(), y.data(), z.data(), 1, s0.data(), s1.data(), s2.data(),
                                                  3, sx.data(), sy.data(), sz.data(), out.data()) == 0);
        for (int i = 0; i < 5; i++) {
            CHECK(std::isfinite(out[i]) && out[i] >= 0.0f);
            CHECK(std::abs(out[i] - brute<float>(x[i], y[i], z[i], sx, sy, sz, s0, s1, s2)) < 1e-5f);
        }
    }
    {  // std::mt19937(1), U(-1,1), 50 triangles, 20 points, FP64, tolerance 1e-7
        std::mt19937 rng(1);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        const int ntri = 50, npt = 20;
        std::vector<double> sx(3 * ntri), sy(3 * ntri), sz(3 * ntri), x(npt), y(npt), z(npt);
        for (int i = 0; i < 3 * ntri; i++) sx[i] = U(rng), sy[i] = U(rng), sz[i] = U(rng);
        std::vector<int> s0(ntri), s1(ntri), s2(ntri);
        for (int t = 0; t < ntri; t++) s0[t] = 3 * t, s1[t] = 3 * t + 1, s2[t] = 3 * t + 2;
        for (int i = 0; i < npt; i++) x[i] = U(rng), y[i] = U(rng), z[i] = U(rng);
        std::vector<double> a(npt, 1e9), b(npt, 1e9), c(npt, 1e9);
        CHECK(ssdf::edf<double, double, int>(npt, x.data(), y.data(), z.data(), ntri, s0.data(), s1.data(), s2.data(),
                                             3 * ntri, sx.data(), sy.data(), sz.data(), a.data()) == 0);
        CHECK(ssdf::edf_select<double, double, int>(npt, x.data(), y.data(), z.data(), ntri, s0.data(), s1.data(),
                                                    s2.data(), 3 * ntri, sx.data(), sy.data(), sz.data(), b.data()) == 0);
        CHECK(ssdf::edf_celllist<double, double, int>(npt, x.data(), y.data(), z.data(), ntri, s0.data(), s1.data(),
                                                      s2.data(), 3 * ntri, sx.data(), sy.data(), sz.data(), c.data()) == 0);
        for (int i = 0; i < npt; i++) {
            const double e = brute<double>(x[i], y[i], z[i], sx, sy, sz, s0, s1, s2);
            CHECK(std::abs(a[i] - e) <= 1e-7 && std::abs(b[i] - e) <= 1e-7 && std::abs(c[i] - e) <= 1e-7);
        }
    }
    std::printf("drop-in edf tests passed\n");
    return 0;
}

// Host build of the engine's FP32 pruning routines (box_d2_fast, tri_d2_fast, tri_plane, make_trip
// from ssdf_amd/csrc/common.hpp): reads (point, triangle) cases on stdin, writes the FP32 squared
// distances, the degenerate flag and the plane distance on stdout; tests/test_fast_routines.py
// checks them against the FP64 oracle.
#include <cstdio>

#include "common.hpp"

int main() {
    int n = 0;
    if (std::scanf("%d", &n) != 1) return 1;
    for (int i = 0; i < n; i++) {
        float p[3], a[3], b[3], c[3];
        for (float *v : {p, a, b, c})
            if (std::scanf("%a %a %a", &v[0], &v[1], &v[2]) != 3) return 1;
        ssdf_amd::TriP t;
        ssdf_amd::make_trip(t, a, b, c, i);
        const float d = ssdf_amd::tri_d2_fast(p[0], p[1], p[2], t);
        float mn[3], mx[3];
        for (int k = 0; k < 3; k++) {
            mn[k] = std::fmin(a[k], std::fmin(b[k], c[k]));
            mx[k] = std::fmax(a[k], std::fmax(b[k], c[k]));
        }
        const float bx = ssdf_amd::box_d2_fast(p[0], p[1], p[2], mn, mx);
        const float h = ssdf_amd::tri_plane(p[0], p[1], p[2], t).h;
        std::printf("%a %a %d %a\n", d, bx, t.f.flags, h);
    }
    return 0;
}

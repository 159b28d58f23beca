// points_inside_main.cpp — the `points_inside` command of zulianp/ssdf (drivers/points_inside.exe.cpp:
// 74-298) on the GPU engine.
//
//   ssdf-points_inside <surface_folder> [surface_folder ...] <points_folder> <output_file>
//
// The surfaces are concatenated into one (each folder's indices offset by the vertices before it,
// 220-255) and every point is classified against it: output uint8[npoints], 1 inside, 0 outside.
// Folders hold x, y, z (float32) and i0, i1, i2 (int32) raw arrays named <c>.raw, as the driver reads
// them, or <c>.float32 / <c>.int32 as the sdf command reads them.  The driver decides by cuBQL's ray
// crossing parity (points_inside_bvh); this engine by the generalized winding number w >= 1/2
// (ssdf_points_inside_f), the same decision on closed surfaces.  Exit status 1 on any error.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cli_io.hpp"
#include "ssdf_c.h"

using namespace ssdf_cli;

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr,
                     "Usage: %s <surf_folder1> [surf_folder2] ... <points_folder> <output_file>\n\n"
                     "Surface folder(s): x, y, z (float32) and i0, i1, i2 (int32) raw arrays (<c>.raw or\n"
                     "  <c>.float32 / <c>.int32).\n"
                     "Points folder: x, y, z (float32).\n"
                     "Output file format: uint8 inside[npoints] (0=outside, 1=inside)\n",
                     argv[0]);
        return 1;
    }
    const char *points_folder = argv[argc - 2];
    const char *output_file = argv[argc - 1];

    std::vector<float> sx, sy, sz;
    std::vector<int> s0, s1, s2;
    for (int i = 1; i < argc - 2; i++) {
        const char *folder = argv[i];
        const std::string fx = find(folder, "x", "float32"), fy = find(folder, "y", "float32"),
                          fz = find(folder, "z", "float32"), f0 = find(folder, "i0", "int32"),
                          f1 = find(folder, "i1", "int32"), f2 = find(folder, "i2", "int32");
        const long nv = count_of(fx, sizeof(float)), nt = count_of(f0, sizeof(int));
        if (nv < 0 || nt < 0) {
            std::fprintf(stderr, "Error: Cannot open or invalid file in '%s'\n", folder);
            return 1;
        }
        if (count_of(fy, sizeof(float)) != nv || count_of(fz, sizeof(float)) != nv) {
            std::fprintf(stderr, "Error: Surface coordinate files have mismatched sizes in '%s'\n", folder);
            return 1;
        }
        if (count_of(f1, sizeof(int)) != nt || count_of(f2, sizeof(int)) != nt) {
            std::fprintf(stderr, "Error: Surface index files have mismatched sizes in '%s'\n", folder);
            return 1;
        }
        std::vector<float> x, y, z;
        std::vector<int> a, b, c;
        if (!read_all(fx, x, nv) || !read_all(fy, y, nv) || !read_all(fz, z, nv) || !read_all(f0, a, nt) ||
            !read_all(f1, b, nt) || !read_all(f2, c, nt)) {
            std::fprintf(stderr, "Error: Failed to read surface '%s'\n", folder);
            return 1;
        }
        const int off = (int)sx.size();  // the folder's local vertex indices follow the vertices before it
        for (long t = 0; t < nt; t++) s0.push_back(a[t] + off), s1.push_back(b[t] + off), s2.push_back(c[t] + off);
        sx.insert(sx.end(), x.begin(), x.end());
        sy.insert(sy.end(), y.begin(), y.end());
        sz.insert(sz.end(), z.begin(), z.end());
        std::printf("  Surface %d (%s): %ld points, %ld elements\n", i, folder, nv, nt);
    }
    const std::string px = find(points_folder, "x", "float32"), py = find(points_folder, "y", "float32"),
                      pz = find(points_folder, "z", "float32");
    const long n = count_of(px, sizeof(float));
    if (n < 0) {
        std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", join(points_folder, "x.raw").c_str());
        return 1;
    }
    if (count_of(py, sizeof(float)) != n || count_of(pz, sizeof(float)) != n) {
        std::fprintf(stderr, "Error: Point coordinate files have mismatched sizes\n");
        return 1;
    }
    std::vector<float> x, y, z;
    if (!read_all(px, x, n) || !read_all(py, y, n) || !read_all(pz, z, n)) {
        std::fprintf(stderr, "Error: Failed to read point coordinates\n");
        return 1;
    }
    std::printf("Running points_inside for %ld points (%zu vertices, %zu triangles)...\n", n, sx.size(), s0.size());
    std::vector<uint8_t> inside((size_t)n, 0);
    const double t0 = now_ms();
    if (n > 0) {
        const int rc = ssdf_points_inside_f(n, x.data(), y.data(), z.data(), (ptrdiff_t)s0.size(), s0.data(),
                                            s1.data(), s2.data(), (ptrdiff_t)sx.size(), sx.data(), sy.data(),
                                            sz.data(), inside.data());
        if (rc) return fail("ssdf_points_inside_f", rc);
    }
    std::printf("points_inside: %.3f ms\n", now_ms() - t0);
    if (!write_all(output_file, inside.data(), inside.size())) {
        std::fprintf(stderr, "Error: Cannot write output file '%s'\n", output_file);
        return 1;
    }
    std::printf("Done. Wrote %ld inside/outside flags to '%s'\n", n, output_file);
    return 0;
}

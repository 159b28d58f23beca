// distance_driver.hpp — the body of the sdf and sqedf commands (drivers/sdf.exe.cpp:128-483,
// drivers/sqedf.exe.cpp:129-469), which share their contract:
//
//   <command> <surface_folder> [surface_folder ...] <points_folder> <output_file>
//
// Each surface folder holds x, y, z (float32) and i0, i1, i2 (int32) raw arrays, the points folder
// x, y, z; the output file is float32[npoints], the distance (sqedf: squared distance) to the union of
// the surfaces, each surface min-accumulated into the running result as the drivers call edf_select /
// sqedf per surface (sdf.exe.cpp:348-459, sqedf.exe.cpp:345-447).  Files are named <c>.float32 /
// <c>.int32 as sdf reads them (sdf.exe.cpp:206-211), or <c>.raw, the names sqedf and the usage texts
// give.  Environment, read as the drivers read it:
//   SSDF_INPUT_SDF=<file>           start from this float32[npoints] field (sdf.exe.cpp:281-311,
//                                   sqedf.exe.cpp:284-314; squared values for sqedf)
//   SSDF_INIT_WITH_VOLUME_BOX=1     sdf only: start from |signed distance to the points' bounding box|
//                                   (sdf.exe.cpp:317-327)
//   SSDF_DOUBLE_PRECISION=1         per surface, exact FP64 results rounded to float
//                                   (sdf.exe.cpp:370-417, sqedf.exe.cpp:378-420)
// otherwise the start value is FLT_MAX.  FP32 runs keep the points and the field in device memory
// across surfaces and download once.  Exit status 1 on any error, 0 on success.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cli_io.hpp"
#include "ssdf_c.h"

namespace ssdf_cli {

struct Surface {
    const char *folder;
    std::string x, y, z, i0, i1, i2;
    long nv, nt;
};

// sqedf<double> of one surface (sqedf.exe.cpp:383-420): out = min(out, exact FP64 squared distance),
// through an FP64 resident mesh (the exact per-lane search) on device copies.
inline int sqedf_d(long n, const std::vector<double> &X, const std::vector<double> &Y, const std::vector<double> &Z,
                   const Surface &s, const std::vector<int> &s0, const std::vector<int> &s1,
                   const std::vector<int> &s2, const std::vector<double> &SX, const std::vector<double> &SY,
                   const std::vector<double> &SZ, std::vector<double> &O) {
    const size_t b = sizeof(double) * (size_t)n;
    void *d[4] = {nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    ssdf_mesh_t m = nullptr;
    for (void *&p : d)
        if (!rc) rc = ssdf_device_malloc(&p, b);
    if (!rc) rc = ssdf_memcpy_h2d(d[0], X.data(), b);
    if (!rc) rc = ssdf_memcpy_h2d(d[1], Y.data(), b);
    if (!rc) rc = ssdf_memcpy_h2d(d[2], Z.data(), b);
    if (!rc) rc = ssdf_memcpy_h2d(d[3], O.data(), b);
    if (!rc)
        rc = ssdf_mesh_create(1, s.nt, s0.data(), s1.data(), s2.data(), s.nv, SX.data(), SY.data(), SZ.data(), 0, 0,
                              nullptr, &m);
    if (!rc)
        rc = ssdf_mesh_query_points(m, SSDF_MODE_SQEDF, n, d[0], d[1], d[2], d[3], nullptr, nullptr, nullptr, nullptr,
                                    0, nullptr);
    if (!rc) rc = ssdf_stream_synchronize(nullptr);
    if (!rc) rc = ssdf_memcpy_d2h(O.data(), d[3], b);
    const int code = rc;
    if (m) ssdf_mesh_destroy(m);
    for (void *p : d)
        if (p) ssdf_device_free(p);
    return code ? fail("FP64 sqedf", code) : 0;
}

// The sdf and sqedf drivers: `squared` selects sqedf (squared distances, min-accumulated; no volume-
// box seed, which the sqedf driver does not read).
inline int distance_driver(int argc, char **argv, bool squared) {
    if (argc < 4) {
        std::fprintf(stderr,
                     "Usage: %s <surface_folder> [surface_folder ...] <points_folder> <output_file>\n\n"
                     "Surface folders: x, y, z (float32) and i0, i1, i2 (int32) raw arrays, named\n"
                     "  <c>.float32 / <c>.int32 (or <c>.raw).\n"
                     "Points folder: x, y, z (float32).\n"
                     "Output: float32[npoints], the %s to the union of the surfaces.\n\n"
                     "Environment: SSDF_INPUT_SDF=<file>, %sSSDF_DOUBLE_PRECISION=1\n",
                     argv[0], squared ? "squared distance" : "distance",
                     squared ? "" : "SSDF_INIT_WITH_VOLUME_BOX=1, ");
        return 1;
    }
    const char *points_folder = argv[argc - 2];
    const char *output_file = argv[argc - 1];

    std::printf("Checking files and determining sizes...\n");
    std::vector<Surface> surfaces;
    for (int i = 1; i < argc - 2; i++) {
        Surface s{argv[i], find(argv[i], "x", "float32"), find(argv[i], "y", "float32"),
                  find(argv[i], "z", "float32"), find(argv[i], "i0", "int32"), find(argv[i], "i1", "int32"),
                  find(argv[i], "i2", "int32"), 0, 0};
        s.nv = count_of(s.x, sizeof(float));
        if (s.nv < 0) {
            std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", join(argv[i], "x.float32").c_str());
            return 1;
        }
        if (count_of(s.y, sizeof(float)) != s.nv || count_of(s.z, sizeof(float)) != s.nv) {
            std::fprintf(stderr, "Error: Surface coordinate files have mismatched sizes in '%s'\n", argv[i]);
            return 1;
        }
        s.nt = count_of(s.i0, sizeof(int));
        if (s.nt < 0) {
            std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", join(argv[i], "i0.int32").c_str());
            return 1;
        }
        if (count_of(s.i1, sizeof(int)) != s.nt || count_of(s.i2, sizeof(int)) != s.nt) {
            std::fprintf(stderr, "Error: Surface index files have mismatched sizes in '%s'\n", argv[i]);
            return 1;
        }
        std::printf("  Surface %zu (%s): %ld points, %ld elements\n", surfaces.size() + 1, argv[i], s.nv, s.nt);
        surfaces.push_back(s);
    }
    const std::string px = find(points_folder, "x", "float32"), py = find(points_folder, "y", "float32"),
                      pz = find(points_folder, "z", "float32");
    const long n = count_of(px, sizeof(float));
    if (n < 0) {
        std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", join(points_folder, "x.float32").c_str());
        return 1;
    }
    if (count_of(py, sizeof(float)) != n || count_of(pz, sizeof(float)) != n) {
        std::fprintf(stderr, "Error: Point coordinate files have mismatched sizes\n");
        return 1;
    }
    std::printf("  Points: %ld points\n", n);

    std::printf("Reading points...\n");
    std::vector<float> x, y, z;
    if (!read_all(px, x, n) || !read_all(py, y, n) || !read_all(pz, z, n)) {
        std::fprintf(stderr, "Error: Failed to read point coordinates\n");
        return 1;
    }

    // initial field
    std::vector<float> out((size_t)n, FLT_MAX);
    const char *input_sdf = std::getenv("SSDF_INPUT_SDF");
    const bool fp64 = env_int("SSDF_DOUBLE_PRECISION") == 1;
    bool seeded_on_device = false;
    if (input_sdf) {
        std::printf("Reading input SDF from '%s'...\n", input_sdf);
        const long m = count_of(input_sdf, sizeof(float));
        if (m < 0) {
            std::fprintf(stderr, "Error: Cannot open input SDF file '%s' or invalid size\n", input_sdf);
            return 1;
        }
        if (m != n) {
            std::fprintf(stderr, "Error: Input SDF file has %ld points, expected %ld\n", m, n);
            return 1;
        }
        if (!read_all(input_sdf, out, n)) {
            std::fprintf(stderr, "Error: Failed to read SDF values from '%s'\n", input_sdf);
            return 1;
        }
    } else if (!squared && env_int("SSDF_INIT_WITH_VOLUME_BOX") && n > 0) {
        seeded_on_device = true;  // below, once the points are on the device
    }

    // device copies of the points and the field (FP32 runs)
    void *dx = nullptr, *dy = nullptr, *dz = nullptr, *dout = nullptr;
    const size_t fb = sizeof(float) * (size_t)(n > 0 ? n : 1);
    int rc;
    if (n > 0) {
        for (void **p : {&dx, &dy, &dz, &dout})
            if ((rc = ssdf_device_malloc(p, fb))) return fail("ssdf_device_malloc", rc);
        if ((rc = ssdf_memcpy_h2d(dx, x.data(), fb)) || (rc = ssdf_memcpy_h2d(dy, y.data(), fb)) ||
            (rc = ssdf_memcpy_h2d(dz, z.data(), fb)))
            return fail("upload of the points", rc);
        if (seeded_on_device) {
            if ((rc = ssdf_init_volume_box(0, n, dx, dy, dz, dout, nullptr))) return fail("ssdf_init_volume_box", rc);
            if ((rc = ssdf_stream_synchronize(nullptr))) return fail("ssdf_stream_synchronize", rc);
        } else if ((rc = ssdf_memcpy_h2d(dout, out.data(), fb))) {
            return fail("upload of the initial field", rc);
        }
    }

    std::printf("Processing surfaces...\n");
    std::vector<float> sx, sy, sz;
    std::vector<int> s0, s1, s2;
    bool host_current = !seeded_on_device;  // `out` on the host holds the field
    for (size_t k = 0; k < surfaces.size(); k++) {
        const Surface &s = surfaces[k];
        std::printf("  Processing surface %zu/%zu: %s (%ld points, %ld elements)\n", k + 1, surfaces.size(),
                    s.folder, s.nv, s.nt);
        if (!read_all(s.x, sx, s.nv) || !read_all(s.y, sy, s.nv) || !read_all(s.z, sz, s.nv)) {
            std::fprintf(stderr, "Error: Failed to read surface coordinates from '%s'\n", s.folder);
            return 1;
        }
        if (!read_all(s.i0, s0, s.nt) || !read_all(s.i1, s1, s.nt) || !read_all(s.i2, s2, s.nt)) {
            std::fprintf(stderr, "Error: Failed to read surface indices from '%s'\n", s.folder);
            return 1;
        }
        const double t0 = now_ms();
        if (n > 0 && s.nt > 0 && s.nv > 0) {
            if (fp64) {
                // edf_select<double>: widen, exact FP64, round back to float per surface
                if (!host_current) {
                    if ((rc = ssdf_memcpy_d2h(out.data(), dout, fb))) return fail("download of the field", rc);
                    host_current = true;
                }
                std::vector<double> X(x.begin(), x.end()), Y(y.begin(), y.end()), Z(z.begin(), z.end());
                std::vector<double> O(out.begin(), out.end());
                std::vector<double> SX(sx.begin(), sx.end()), SY(sy.begin(), sy.end()), SZ(sz.begin(), sz.end());
                if (!squared) {
                    if ((rc = ssdf_edf_d(n, X.data(), Y.data(), Z.data(), s.nt, s0.data(), s1.data(), s2.data(),
                                         s.nv, SX.data(), SY.data(), SZ.data(), O.data())))
                        return fail("ssdf_edf_d", rc);
                } else if ((rc = sqedf_d(n, X, Y, Z, s, s0, s1, s2, SX, SY, SZ, O))) {
                    return rc;
                }
                for (long i = 0; i < n; i++) out[(size_t)i] = (float)O[(size_t)i];
                if ((rc = ssdf_memcpy_h2d(dout, out.data(), fb))) return fail("upload of the field", rc);
            } else {
                ssdf_mesh_t m = nullptr;
                if ((rc = ssdf_mesh_create(0, s.nt, s0.data(), s1.data(), s2.data(), s.nv, sx.data(), sy.data(),
                                           sz.data(), 0, 0, nullptr, &m)))
                    return fail("ssdf_mesh_create", rc);
                rc = ssdf_mesh_query_points(m, squared ? SSDF_MODE_SQEDF : SSDF_MODE_EDF, n, dx, dy, dz, dout, nullptr,
                                            nullptr, nullptr, nullptr, 0, nullptr);
                const int rs = ssdf_stream_synchronize(nullptr);
                ssdf_mesh_destroy(m);
                if (rc) return fail("ssdf_mesh_query_points", rc);
                if (rs) return fail("ssdf_stream_synchronize", rs);
                host_current = false;
            }
        }
        std::printf("Time taken: %f ms\n", now_ms() - t0);
    }
    if (n > 0 && !host_current)
        if ((rc = ssdf_memcpy_d2h(out.data(), dout, fb))) return fail("download of the field", rc);
    for (void *p : {dx, dy, dz, dout})
        if (p) ssdf_device_free(p);

    std::printf("Writing output to '%s'...\n", output_file);
    FILE *f = std::fopen(output_file, "wb");
    if (!f) {
        std::fprintf(stderr, "Error: Cannot open output file '%s'\n", output_file);
        return 1;
    }
    if (n > 0 && std::fwrite(out.data(), sizeof(float), (size_t)n, f) != (size_t)n) {
        std::fprintf(stderr, "Error: Failed to write SDF results\n");
        std::fclose(f);
        return 1;
    }
    std::fclose(f);
    std::printf("Successfully computed %s for %ld points across %zu surfaces, written to '%s'\n",
                squared ? "squared distances" : "SDF", n, surfaces.size(), output_file);
    return 0;
}

}  // namespace ssdf_cli

// mesh_to_sdf_main.cpp — the `mesh_to_sdf` demo of zulianp/ssdf (demo/mesh_to_sdf.exe.cpp:73-208) on
// the GPU engine.
//
//   ssdf-mesh_to_sdf <surface_folder> <nx> <ny> <nz> <output_folder>
//
// The signed distance to the triangles of <surface_folder> on the nodes of an nx x ny x nz grid over
// the surface's bounding box (sample_sdf_bvh: pseudonormal sign, positive where the right-handed
// normals point).  The surface folder holds x, y, z (float32) and i0, i1, i2 (int32) raw arrays, as
// the sdf command reads them.  Environment, read as the demo reads it (73-92, 137-176, 200-201):
//   SSDF_MARGIN, SSDF_MARGIN_X/_Y/_Z  grow the box by this much per side (default 0; the per-axis
//                                     values default to SSDF_MARGIN)
//   SSDF_SCALE_BOX                    then scale it about its centre (default 1)
//   SSDF_BOX_MESH=<folder>            take the box from the x, y, z arrays of this folder instead
//   SSDF_HMAX                         grid spacing at most this: n = ceil(range / hmax) + 1 per axis,
//                                     replacing nx, ny, nz (pass 0 0 0)
//   SSDF_SCALE                        multiply the distances (e.g. -1 flips the sign)
// SSDF_BOX_BLOCKS selects blocks of an smesh mesh; a raw folder has none, so it is an error here.
// Output: <output_folder>/sdf.float32 (float32[nx*ny*nz], x fastest) and <output_folder>/grid.txt
// (dims and box).  The demo writes the grid through smesh, whose file layout is not reproduced.
// Exit status 1 on any error, 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "cli_io.hpp"
#include "ssdf_c.h"

using namespace ssdf_cli;

namespace {

ptrdiff_t parse_extent(const char *s) {
    char *end = nullptr;
    const long long v = std::strtoll(s, &end, 10);
    return (end != s && *end == '\0') ? (ptrdiff_t)v : 0;
}

// The box grown by the margins, then scaled about its centre (mesh_to_sdf.exe.cpp:30-54), in float.
void expand_box(float &xmin, float &xmax, float &ymin, float &ymax, float &zmin, float &zmax, float scale, float mx,
                float my, float mz) {
    const float cx = (xmin + xmax) * 0.5f, cy = (ymin + ymax) * 0.5f, cz = (zmin + zmax) * 0.5f;
    const float hx = ((xmax - xmin) * 0.5f + mx) * scale;
    const float hy = ((ymax - ymin) * 0.5f + my) * scale;
    const float hz = ((zmax - zmin) * 0.5f + mz) * scale;
    xmin = cx - hx, xmax = cx + hx;
    ymin = cy - hy, ymax = cy + hy;
    zmin = cz - hz, zmax = cz + hz;
}

ptrdiff_t points_for_spacing(double range, double hmax) { return (ptrdiff_t)std::ceil(range / hmax) + 1; }

bool read_points(const char *folder, std::vector<float> &x, std::vector<float> &y, std::vector<float> &z) {
    const std::string px = find(folder, "x", "float32"), py = find(folder, "y", "float32"),
                      pz = find(folder, "z", "float32");
    const long n = count_of(px, sizeof(float));
    if (n <= 0 || count_of(py, sizeof(float)) != n || count_of(pz, sizeof(float)) != n) return false;
    return read_all(px, x, n) && read_all(py, y, n) && read_all(pz, z, n);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr,
                     "Usage: %s <surface_folder> <nx> <ny> <nz> <output_folder>\n"
                     "  The signed distance to the triangles of <surface_folder> on a grid.\n"
                     "  SSDF_MARGIN[_X|_Y|_Z]  grow the box by this much (default 0)\n"
                     "  SSDF_SCALE_BOX         then scale it about its centre (default 1)\n"
                     "  SSDF_BOX_MESH          take the box from the points of this folder, not the surface\n"
                     "  SSDF_HMAX              grid spacing at most this; then pass 0 0 0\n"
                     "  SSDF_SCALE             multiply the distances (e.g. -1 flips the sign)\n"
                     "  Output: <output_folder>/sdf.float32 (x fastest) and <output_folder>/grid.txt\n",
                     argv[0]);
        return 1;
    }
    ptrdiff_t nx = parse_extent(argv[2]), ny = parse_extent(argv[3]), nz = parse_extent(argv[4]);
    const double hmax = env_double("SSDF_HMAX", 0.0);
    if (hmax <= 0 && (nx < 2 || ny < 2 || nz < 2)) {
        std::fprintf(stderr, "Error: nx, ny, nz must be integers >= 2 (or set SSDF_HMAX)\n");
        return 1;
    }
    const char *folder = argv[1];
    std::vector<float> sx, sy, sz;
    if (!read_points(folder, sx, sy, sz)) {
        std::fprintf(stderr, "Error: cannot read the x, y, z arrays of '%s'\n", folder);
        return 1;
    }
    const std::string fi0 = find(folder, "i0", "int32"), fi1 = find(folder, "i1", "int32"),
                      fi2 = find(folder, "i2", "int32");
    const long nt = count_of(fi0, sizeof(int));
    std::vector<int> s0, s1, s2;
    if (nt < 0 || count_of(fi1, sizeof(int)) != nt || count_of(fi2, sizeof(int)) != nt || !read_all(fi0, s0, nt) ||
        !read_all(fi1, s1, nt) || !read_all(fi2, s2, nt)) {
        std::fprintf(stderr, "Error: cannot read the i0, i1, i2 arrays of '%s'\n", folder);
        return 1;
    }
    const long nv = (long)sx.size();

    // the box: of the surface, or of another folder's points
    float xmin, xmax, ymin, ymax, zmin, zmax;
    int rc;
    const char *blocks = std::getenv("SSDF_BOX_BLOCKS");
    if (blocks && blocks[0]) {
        std::fprintf(stderr, "Error: SSDF_BOX_BLOCKS selects blocks of an smesh mesh; a raw folder has none\n");
        return 1;
    }
    const char *box_mesh = std::getenv("SSDF_BOX_MESH");
    if (box_mesh && box_mesh[0]) {
        std::vector<float> bx, by, bz;
        if (!read_points(box_mesh, bx, by, bz)) {
            std::fprintf(stderr, "Error: unable to read %s\n", box_mesh);
            return 1;
        }
        rc = ssdf_compute_aabb_f((ptrdiff_t)bx.size(), bx.data(), by.data(), bz.data(), &xmin, &xmax, &ymin, &ymax,
                                 &zmin, &zmax);
    } else {
        rc = ssdf_compute_aabb_f(nv, sx.data(), sy.data(), sz.data(), &xmin, &xmax, &ymin, &ymax, &zmin, &zmax);
    }
    if (rc) return fail("ssdf_compute_aabb_f", rc);
    const float margin = (float)env_double("SSDF_MARGIN", 0.0);
    expand_box(xmin, xmax, ymin, ymax, zmin, zmax, (float)env_double("SSDF_SCALE_BOX", 1.0),
               (float)env_double("SSDF_MARGIN_X", margin), (float)env_double("SSDF_MARGIN_Y", margin),
               (float)env_double("SSDF_MARGIN_Z", margin));
    if (hmax > 0) {
        nx = points_for_spacing(xmax - xmin, hmax);
        ny = points_for_spacing(ymax - ymin, hmax);
        nz = points_for_spacing(zmax - zmin, hmax);
    }
    std::printf("mesh_to_sdf: %ld triangles, grid %td x %td x %td on [%g, %g] x [%g, %g] x [%g, %g]\n", nt, nx, ny,
                nz, xmin, xmax, ymin, ymax, zmin, zmax);

    const size_t nq = (size_t)nx * (size_t)ny * (size_t)nz;
    std::vector<float> out(nq, 0.0f);
    const double t0 = now_ms();
    if ((rc = ssdf_sample_sdf_f(nt, s0.data(), s1.data(), s2.data(), nv, sx.data(), sy.data(), sz.data(), nx, ny, nz,
                                xmin, ymin, zmin, xmax, ymax, zmax, out.data(), nullptr, 0)))
        return fail("ssdf_sample_sdf_f", rc);
    std::printf("sample_sdf_bvh: %.3f ms\n", now_ms() - t0);
    const float scale = (float)env_double("SSDF_SCALE", 1.0);
    if (scale != 1.0f)
        for (float &v : out) v *= scale;

    mkdir(argv[5], 0755);  // may exist already; the writes below report a real failure
    const std::string grid = join(argv[5], "sdf.float32"), meta = join(argv[5], "grid.txt");
    if (!write_all(grid, out.data(), nq)) {
        std::fprintf(stderr, "Error: failed to write grid to %s\n", argv[5]);
        return 1;
    }
    FILE *f = std::fopen(meta.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Error: failed to write %s\n", meta.c_str());
        return 1;
    }
    std::fprintf(f, "nx %td\nny %td\nnz %td\nxmin %.9g\nymin %.9g\nzmin %.9g\nxmax %.9g\nymax %.9g\nzmax %.9g\n", nx,
                 ny, nz, xmin, ymin, zmin, xmax, ymax, zmax);
    std::fclose(f);
    return 0;
}

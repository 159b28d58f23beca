// sdf_sign_main.cpp — the `sdf_sign` command of zulianp/ssdf (drivers/sdf_sign.exe.cpp:74-141): signs
// an unsigned field with inside flags, the last step of the sdf -> points_inside -> sdf_sign chain.
//
//   ssdf-sdf_sign <udf_file> <inside_u8_file> <output_sdf_file>
//
// udf_file float32[n], inside_u8_file uint8[n] (0 outside, 1 inside); output float32[n], the
// distance negated where the flag is set.  A file transform on the host, as in the reference: it
// reads and writes n values once and calls nothing in the engine.  Exit status 1 on any error.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cli_io.hpp"

using namespace ssdf_cli;

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr,
                     "Usage: %s <udf_file> <inside_u8_file> <output_sdf_file>\n\n"
                     "Input formats:\n"
                     "  - udf_file: float udf[npoints] (binary)\n"
                     "  - inside_u8_file: uint8 inside[npoints] (0=outside, 1=inside)\n\n"
                     "Output format:\n"
                     "  - float sdf[npoints] (negative inside)\n",
                     argv[0]);
        return 1;
    }
    const long n_udf = count_of(argv[1], sizeof(float));
    if (n_udf < 0) {
        std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", argv[1]);
        return 1;
    }
    const long n_inside = count_of(argv[2], sizeof(uint8_t));
    if (n_inside < 0) {
        std::fprintf(stderr, "Error: Cannot open or invalid file '%s'\n", argv[2]);
        return 1;
    }
    if (n_udf != n_inside) {
        std::fprintf(stderr, "Error: Size mismatch: udf has %ld entries, inside has %ld entries\n", n_udf, n_inside);
        return 1;
    }
    std::printf("Reading %ld values...\n", n_udf);
    std::vector<float> sdf;
    std::vector<uint8_t> inside;
    if (!read_all(argv[1], sdf, n_udf) || !read_all(argv[2], inside, n_inside)) {
        std::fprintf(stderr, "Error: Failed to read the inputs\n");
        return 1;
    }
    for (long i = 0; i < n_udf; i++)
        if (inside[(size_t)i]) sdf[(size_t)i] = -sdf[(size_t)i];  // negative inside
    std::printf("Writing output to '%s'...\n", argv[3]);
    if (!write_all(argv[3], sdf.data(), sdf.size())) {
        std::fprintf(stderr, "Error: Failed to write output\n");
        return 1;
    }
    return 0;
}

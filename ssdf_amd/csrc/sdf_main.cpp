// sdf_main.cpp — the `sdf` command of zulianp/ssdf (drivers/sdf.exe.cpp:128-483) on the GPU engine:
//
//   ssdf-sdf <surface_folder> [surface_folder ...] <points_folder> <output_file>
//
// The distance to the union of the surfaces; the contract and environment are described in
// distance_driver.hpp.
#include "distance_driver.hpp"

int main(int argc, char **argv) { return ssdf_cli::distance_driver(argc, argv, false); }

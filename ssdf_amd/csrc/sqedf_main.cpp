// sqedf_main.cpp — the `sqedf` command of zulianp/ssdf (drivers/sqedf.exe.cpp:129-469) on the GPU
// engine:
//
//   ssdf-sqedf <surface_folder> [surface_folder ...] <points_folder> <output_file>
//
// The squared distance to the union of the surfaces, min-accumulated per surface as ssdf::sqedf does
// (serial/sqedf.hpp:37-52: out holds the current best squared distances); the contract and
// environment are described in distance_driver.hpp.
#include "distance_driver.hpp"

int main(int argc, char **argv) { return ssdf_cli::distance_driver(argc, argv, true); }

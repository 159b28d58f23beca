// cli_io.hpp — raw-array file helpers shared by the command-line programs (sdf_main.cpp,
// mesh_to_sdf_main.cpp).  A folder holds one file per array: <name>.<ext> (float32 / int32), or
// <name>.raw.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ssdf_c.h"

namespace ssdf_cli {

inline std::string join(const char *folder, const std::string &name) {
    std::string p = folder;
    if (!p.empty() && p.back() != '/') p += '/';
    return p + name;
}

inline long file_bytes(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return -1;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fclose(f);
    return n;
}

// <folder>/<stem>.<ext> if present, else <folder>/<stem>.raw; "" if neither.
inline std::string find(const char *folder, const char *stem, const char *ext) {
    const std::string a = join(folder, std::string(stem) + "." + ext);
    if (file_bytes(a) >= 0) return a;
    const std::string b = join(folder, std::string(stem) + ".raw");
    if (file_bytes(b) >= 0) return b;
    return "";
}

// Element count of a raw array of `esize`-byte elements, -1 if missing or not a whole count.
inline long count_of(const std::string &path, size_t esize) {
    if (path.empty()) return -1;
    const long n = file_bytes(path);
    if (n < 0 || n % (long)esize) return -1;
    return n / (long)esize;
}

template <typename T>
bool read_all(const std::string &path, std::vector<T> &v, long n) {
    v.resize((size_t)n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t got = n ? std::fread(v.data(), sizeof(T), (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

template <typename T>
bool write_all(const std::string &path, const T *v, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t put = n ? std::fwrite(v, sizeof(T), n, f) : 0;
    return std::fclose(f) == 0 && put == n;
}

inline int env_int(const char *name) {
    const char *s = std::getenv(name);
    return s ? std::atoi(s) : 0;
}

inline double env_double(const char *name, double dflt) {
    const char *s = std::getenv(name);
    return s && s[0] ? std::atof(s) : dflt;
}

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int fail(const char *what, int rc) {
    std::fprintf(stderr, "Error: %s failed (code %d): %s\n", what, rc, ssdf_last_error());
    return 1;
}

}  // namespace ssdf_cli

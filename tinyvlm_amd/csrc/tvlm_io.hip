// tvlm_io.hip -- plot3d mesh input for the device paths (host code; reference: MeshIO,
// vlm/src/vlm_mesh.cpp:15-124).
//
// Single-block 2-D plot3d: "1", "ni nj 1", then every x, every y, every z with the chordwise index i
// fastest.  Values are parsed in FP32 (the steady and unsteady drivers' Simulation buffers are
// FP32, vlm/include/vlm.hpp:44-49) or in FP64 (the harmonic-balance driver reads into FP64 buffers,
// lib/libhbvlm3.cpp:158-162), then optionally shifted to the quarter chord in that precision: every
// chordwise row moves a quarter panel aft (0.75f v(j) + 0.25f v(j+1)) and the trailing-edge row is
// extrapolated (4/3 v - 1/3 v_prev) (vlm/src/vlm_mesh.cpp:15-32); the result is FP64 with w = 1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "tvlm_internal.hpp"

#pragma clang fp contract(off)

namespace {

int read_header(std::ifstream& f, const char* path, int64_t& ni, int64_t& nj) {
    int64_t blocks = 0, nk = 0;
    if (!(f >> blocks) || blocks != 1) {
        tvlm::set_error("%s: only single-block plot3d meshes are supported", path);
        return TVLM_ERR_INVALID;
    }
    if (!(f >> ni >> nj >> nk) || nk != 1 || ni < 2 || nj < 2) {
        tvlm::set_error("%s: expected a 2-D plot3d header 'ni nj 1' with ni, nj >= 2", path);
        return TVLM_ERR_INVALID;
    }
    return TVLM_OK;
}

}  // namespace

extern "C" {

int tvlm_plot3d_dims(const char* path, int64_t* ns, int64_t* nc) {
    TVLM_REQUIRE(path && ns && nc, "null argument");
    std::ifstream f(path);
    TVLM_REQUIRE(f.is_open(), "failed to open mesh file %s", path);
    int64_t ni = 0, nj = 0;
    TVLM_TRY(read_header(f, path, ni, nj));
    *nc = ni - 1;
    *ns = nj - 1;
    return TVLM_OK;
}

}  // extern "C"

namespace {

template <typename R>
int read_verts(const char* path, int32_t quarter_chord, double* verts) {
    std::ifstream f(path);
    TVLM_REQUIRE(f.is_open(), "failed to open mesh file %s", path);
    int64_t ni = 0, nj = 0;
    TVLM_TRY(read_header(f, path, ni, nj));
    const int64_t n = ni * nj;  // vertices (j spanwise, i chordwise) -> verts(j, i, k) at j + i nj + k n
    std::vector<R> v((size_t)n * 3);
    std::string tok;
    for (int64_t k = 0; k < 3; k++)
        for (int64_t j = 0; j < nj; j++)
            for (int64_t i = 0; i < ni; i++) {
                TVLM_REQUIRE((bool)(f >> tok), "%s: truncated coordinate block", path);
                char* end = nullptr;
                const R x = sizeof(R) == 4 ? (R)std::strtof(tok.c_str(), &end) : (R)std::strtod(tok.c_str(), &end);
                TVLM_REQUIRE(end && *end == '\0', "%s: bad number '%s'", path, tok.c_str());
                v[(size_t)(j + i * nj + k * n)] = x;
            }
    for (int64_t i = 1; i < ni; i++)
        TVLM_REQUIRE(v[(size_t)(i * nj + n)] == v[(size_t)n], "%s: mesh vertices should be ordered in chordwise direction",
                     path);
    if (quarter_chord) {
        for (int64_t i = 0; i + 1 < ni; i++)
            for (int64_t j = 0; j < nj; j++)
                for (int64_t k = 0; k < 3; k++) {
                    R& a = v[(size_t)(j + i * nj + k * n)];
                    const R b = v[(size_t)(j + (i + 1) * nj + k * n)];
                    a = 0.75f * a + 0.25f * b;
                }
        for (int64_t j = 0; j < nj; j++)
            for (int64_t k = 0; k < 3; k++) {
                R& a = v[(size_t)(j + (ni - 1) * nj + k * n)];
                const R b = v[(size_t)(j + (ni - 2) * nj + k * n)];
                a = (4.f / 3.f) * a - (1.f / 3.f) * b;
            }
    }
    for (int64_t e = 0; e < 3 * n; e++) verts[e] = (double)v[(size_t)e];
    for (int64_t e = 0; e < n; e++) verts[3 * n + e] = 1.0;
    return TVLM_OK;
}

}  // namespace

extern "C" {

int tvlm_plot3d_read(const char* path, int32_t quarter_chord, int32_t parse_f64, double* verts) {
    TVLM_REQUIRE(path && verts, "null argument");
    return parse_f64 ? read_verts<double>(path, quarter_chord, verts) : read_verts<float>(path, quarter_chord, verts);
}

}  // extern "C"

// vlm_coeffs -- the reference's steady known-answer programs (tests/vlm_elliptic_coeffs.cpp and
// tests/vlm_square_coeffs.cpp) over the C++ driver and the HIP backend.
//
//   vlm_coeffs elliptic <mesh_dir> [f32|f64]   elliptic_45x45: CL within 3% of lifting-line theory
//   vlm_coeffs square   <mesh_dir> [f32|f64]   infinite rectangular wings: CL within 1% of 2 pi a
//
// Alphas 0,1,2,3,4,5,10,15 degrees, beta 0, u_inf 1, rho 1.  Besides the table, each solve prints
// one machine-readable line "RESULT <case> <alpha_deg> <cl> <cmx> <cmy> <cmz>" in the working
// precision.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vlm.hpp"

using namespace vlm;

namespace {

// elliptic wing, root chord a, half span b: S = pi a b / 2, AR = (2b)^2 / S
f32 s_ref(f32 a, f32 b) { return 0.5f * PI_f * a * b; }
f32 aspect_ratio(f32 a, f32 b) { return (2 * b) * (2 * b) / s_ref(a, b); }
f32 cl_elliptic(f32 alpha, f32 a, f32 b) { return 2.0f * PI_f * alpha / (1.0f + 2.0f / aspect_ratio(a, b)); }
f32 cl_2d(f32 alpha) { return 2.0f * PI_f * alpha; }

template <typename T>
int run_case(const std::string& label, const std::vector<std::string>& meshes, bool elliptic) {
    const std::vector<f32> degrees = {0, 1, 2, 3, 4, 5, 10, 15};
    int failures = 0;
    for (const auto& backend_name : get_available_backends()) {
        VLMT<T> sim{backend_name, meshes};
        std::printf("\n%s (%s, %s)\n", label.c_str(), backend_name.c_str(), sizeof(T) == 8 ? "f64" : "f32");
        std::printf("|    Alpha   |     CL     |    CMx     |    CMy     |    CMz     |  CL Error   |\n");
        std::printf("|------------|------------|------------|------------|------------|-------------|\n");
        for (const f32 deg : degrees) {
            const FlowData flow{to_radians(deg), 0.0f, 1.0f, 1.0f};
            const auto c = sim.run(flow);
            const f32 exact = elliptic ? cl_elliptic(flow.alpha, 1.0f, 5.0f) : cl_2d(flow.alpha);
            const f32 rerr = exact < EPS_f ? 0.f : std::abs(c.cl - exact) / exact;
            std::printf("| %10.1f | %10.6f | %10.6f | %10.6f | %10.6f | %10.3f%% |\n", deg, c.cl, c.cm.x, c.cm.y, c.cm.z,
                        rerr * 100.0f);
            std::printf("RESULT %s %.1f %.17g %.17g %.17g %.17g\n", label.c_str(), deg, (double)sim.last_cl,
                        (double)sim.last_cm.x, (double)sim.last_cm.y, (double)sim.last_cm.z);
            if (rerr > (elliptic ? 0.03f : 0.01f)) {
                std::printf("Large CL error\n");
                failures++;
            }
        }
    }
    return failures;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: %s elliptic|square <mesh_dir> [f32|f64]\n", argv[0]);
        return 2;
    }
    const std::string which = argv[1], dir = std::string(argv[2]) + "/";
    const bool f64 = argc > 3 && std::strcmp(argv[3], "f64") == 0;
    int failures = 0;
    try {
        if (which == "elliptic") {
            const std::vector<std::string> m = {dir + "elliptic_45x45.x"};
            failures += f64 ? run_case<double>("elliptic_45x45", m, true) : run_case<float>("elliptic_45x45", m, true);
        } else if (which == "square") {
            const std::vector<std::pair<std::string, std::vector<std::string>>> sets = {
                {"infinite_rectangular_2x2_parts",
                 {dir + "infinite_rectangular_2x2_part0.x", dir + "infinite_rectangular_2x2_part1.x"}},
                {"infinite_rectangular_2x4", {dir + "infinite_rectangular_2x4.x"}},
                {"infinite_rectangular_10x5", {dir + "infinite_rectangular_10x5.x"}},
                {"infinite_rectangular_20x5", {dir + "infinite_rectangular_20x5.x"}},
            };
            for (const auto& [label, m] : sets)
                failures += f64 ? run_case<double>(label, m, false) : run_case<float>(label, m, false);
        } else {
            std::printf("unknown case %s\n", which.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 3;
    }
    return failures ? 1 : 0;
}

// vlm_solvers.hpp -- fixed-point acceleration used by the harmonic-balance driver
// (reference: vlm/include/vlm_solvers.hpp:9-25, vlm/src/vlm_solvers.cpp:6-150).
#pragma once

#include <functional>

#include "vlm_backend.hpp"

namespace vlm {

// Type-II Anderson mixing for x = f(x) on device vectors: keeps the last m differences of iterates
// (X) and of residuals g = f(x) - x (G) in circular buffers, and each iteration takes
// x+ = x + g - (X + G) gamma with gamma = argmin ||G gamma - g|| (LSQ::solve).  Stops when
// ||g|| <= tol_res or after max_iter iterations; x0 receives the last iterate.  Returns the
// iteration count (the reference returns nothing and prints on non-convergence).
i32 anderson_acceleration(Backend* backend, const TensorView1dD& x0,
                          const std::function<void(const TensorView1dD& x, const TensorView1dD& y)>& f,
                          i32 max_iter = 100, f64 tol_res = 1e-6, i32 m = 3);
i32 anderson_acceleration(Backend* backend, const TensorView1fD& x0,
                          const std::function<void(const TensorView1fD& x, const TensorView1fD& y)>& f,
                          i32 max_iter = 100, f32 tol_res = 1e-6, i32 m = 3);

}  // namespace vlm

// vlm_types.hpp -- scalar aliases and the constants the reference backends use
// (vlm/include/vlm_types.hpp:8-28 of the reference: same names and values, they feed parity).
#pragma once

#include <cstdint>

namespace vlm {

using i32 = std::int32_t;
using i64 = std::int64_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using f32 = float;
using f64 = double;

static constexpr f64 PI_d = 3.14159265358979;
static constexpr f32 PI_f = 3.141593f;
static constexpr f32 EPS_f = 1.19209e-07f;

}  // namespace vlm

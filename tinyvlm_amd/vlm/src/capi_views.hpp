// capi_views.hpp -- conversions from the C++ data model (TensorView, MultiTensorView) to the
// C ABI's tvlm_view, and C-ABI status to exceptions.  Internal to the C++ layer.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tinyvlm_hip.h"
#include "vlm_memory.hpp"

namespace vlm::capi {

inline void check(int status, const char* what) {
    if (status != TVLM_OK) throw std::runtime_error(std::string(what) + ": " + tvlm_last_error());
}

// trailing dimensions beyond D get extent 1 and the stride that continues the last one
template <typename T, i64 D, Location L>
tvlm_view view(const TensorView<T, D, L>& v) {
    tvlm_view r{};
    r.ptr = v.ptr();
    for (int k = 0; k < 3; k++) {
        r.shape[k] = k < D ? v.shape(k) : 1;
        r.stride[k] = k < D ? v.stride(k) : (k > 0 ? r.stride[k - 1] * r.shape[k - 1] : 1);
    }
    return r;
}

template <typename V>
std::vector<tvlm_view> views(const std::vector<V>& m) {
    std::vector<tvlm_view> r;
    r.reserve(m.size());
    for (const auto& v : m) r.push_back(view(v));
    return r;
}

inline int32_t location(Location l) { return l == Location::Device ? 0 : 1; }

}  // namespace vlm::capi

// vlm_backend.cpp -- the non-virtual Backend helpers (reference: vlm/src/vlm_backend.cpp:40-282).
// wake_shed and displace_wing go to the C ABI as one launch for all surfaces; the multibody
// force helpers loop over surfaces through the virtuals; the multibody coefficients are the
// area-weighted (CL) and area*MAC-weighted (CM) averages of the per-surface coefficients.
#include "capi_views.hpp"
#include "vlm_backend.hpp"

namespace vlm {

Backend::Backend(std::unique_ptr<Memory> memory_, std::unique_ptr<BLAS> blas_)
    : memory(std::move(memory_)), blas(std::move(blas_)) {}

Backend::~Backend() = default;

namespace {

// wake_shed: trailing-edge row of every surface copied into wake row -1-iteration
// (vlm/src/vlm_backend.cpp:42-47)
template <typename T>
void wake_shed_impl(MultiTensorView<T, 3, Location::Device>& wing, MultiTensorView<T, 3, Location::Device>& wake,
                    i32 iteration) {
    if (wing.size() != wake.size()) throw std::invalid_argument("wake_shed: surface count mismatch");
    const auto w = capi::views(wing), k = capi::views(wake);
    if constexpr (sizeof(T) == 8)
        capi::check(tvlm_wake_shed_f64((int32_t)w.size(), w.data(), k.data(), iteration, nullptr), "wake_shed");
    else
        capi::check(tvlm_wake_shed_f32((int32_t)w.size(), w.data(), k.data(), iteration, nullptr), "wake_shed");
}

// displace_wing: verts(ns*nc, 4) = verts_init(ns*nc, 4) * transform^T per surface
// (vlm/src/vlm_backend.cpp:49-69)
template <typename T>
void displace_impl(MultiTensorView<T, 2, Location::Device>& transforms, MultiTensorView<T, 3, Location::Device>& wing,
                   MultiTensorView<T, 3, Location::Device>& init) {
    if (wing.size() != init.size() || wing.size() != transforms.size())
        throw std::invalid_argument("displace_wing: surface count mismatch");
    const auto t = capi::views(transforms), w = capi::views(wing), i = capi::views(init);
    if constexpr (sizeof(T) == 8)
        capi::check(tvlm_displace_wing_f64((int32_t)w.size(), t.data(), w.data(), i.data(), nullptr), "displace_wing");
    else
        capi::check(tvlm_displace_wing_f32((int32_t)w.size(), t.data(), w.data(), i.data(), nullptr), "displace_wing");
}

// area-weighted CL over surfaces, lift axis normalize(freestream x y) (vlm/src/vlm_backend.cpp:71-99)
template <typename T>
T cl_multibody(Backend& b, MultiTensorView<T, 3, Location::Device>& forces, MultiTensorView<T, 2, Location::Device>& areas,
               const linalg::vec3<T>& freestream, T rho) {
    T weighted = 0, total = 0;
    for (size_t m = 0; m < forces.size(); m++) {
        const T area = b.sum(areas[m]);
        auto axis = linalg::normalize(linalg::cross(freestream, linalg::vec3<T>{0, 1, 0}));
        auto fs = freestream;
        weighted += b.coeff_cl(forces[m], axis, fs, rho, area) * area;
        total += area;
    }
    return weighted / total;
}

// CM over surfaces weighted by area*MAC, normalised by (sum area)(sum MAC)
// (vlm/src/vlm_backend.cpp:101-135)
template <typename T>
linalg::vec3<T> cm_multibody(Backend& b, MultiTensorView<T, 3, Location::Device>& forces,
                             MultiTensorView<T, 3, Location::Device>& verts, MultiTensorView<T, 2, Location::Device>& areas,
                             const linalg::vec3<T>& ref_pt, const linalg::vec3<T>& freestream, T rho) {
    linalg::vec3<T> acc{0, 0, 0};
    T total_area = 0, total_mac = 0;
    for (size_t m = 0; m < forces.size(); m++) {
        const T area = b.sum(areas[m]);
        const T mac = b.mesh_mac(verts[m], areas[m]);
        auto rp = ref_pt;
        auto fs = freestream;
        acc = acc + (area * mac) * b.coeff_cm(forces[m], verts[m], rp, fs, rho, area, mac);
        total_area += area;
        total_mac += mac;
    }
    return acc / (total_area * total_mac);
}

}  // namespace

void Backend::wake_shed(MTV3f& verts_wing, MTV3f& verts_wake, i32 iteration) { wake_shed_impl(verts_wing, verts_wake, iteration); }
void Backend::wake_shed(MTV3d& verts_wing, MTV3d& verts_wake, i32 iteration) { wake_shed_impl(verts_wing, verts_wake, iteration); }
void Backend::displace_wing(MTV2f& transforms, MTV3f& verts_wing, MTV3f& verts_wing_init) { displace_impl(transforms, verts_wing, verts_wing_init); }
void Backend::displace_wing(MTV2d& transforms, MTV3d& verts_wing, MTV3d& verts_wing_init) { displace_impl(transforms, verts_wing, verts_wing_init); }

// per-surface loops (vlm/src/vlm_backend.cpp:201-282)
void Backend::forces_steady_multibody(MTV3f& verts_wing, MTV2f& gamma_delta, MTV3f& velocities, MTV3f& forces) {
    for (size_t m = 0; m < verts_wing.size(); m++) forces_steady(verts_wing[m], gamma_delta[m], velocities[m], forces[m]);
}
void Backend::forces_steady_multibody(MTV3d& verts_wing, MTV2d& gamma_delta, MTV3d& velocities, MTV3d& forces) {
    for (size_t m = 0; m < verts_wing.size(); m++) forces_steady(verts_wing[m], gamma_delta[m], velocities[m], forces[m]);
}
void Backend::forces_unsteady_multibody(MTV3f& verts_wing, MTV2f& gamma_delta, MTV2f& dgamma_dt, MTV3f& velocities,
                                        MTV2f& areas, MTV3f& normals, MTV3f& forces) {
    for (size_t m = 0; m < verts_wing.size(); m++)
        forces_unsteady(verts_wing[m], gamma_delta[m], dgamma_dt[m], velocities[m], areas[m], normals[m], forces[m]);
}
void Backend::forces_unsteady_multibody(MTV3d& verts_wing, MTV2d& gamma_delta, MTV2d& dgamma_dt, MTV3d& velocities,
                                        MTV2d& areas, MTV3d& normals, MTV3d& forces) {
    for (size_t m = 0; m < verts_wing.size(); m++)
        forces_unsteady(verts_wing[m], gamma_delta[m], dgamma_dt[m], velocities[m], areas[m], normals[m], forces[m]);
}

f32 Backend::coeff_cl_multibody(MTV3f& aero_forces, MTV2f& areas, f32x3& freestream, f32 rho) {
    return cl_multibody(*this, aero_forces, areas, freestream, rho);
}
f64 Backend::coeff_cl_multibody(MTV3d& aero_forces, MTV2d& areas, f64x3& freestream, f64 rho) {
    return cl_multibody(*this, aero_forces, areas, freestream, rho);
}
Backend::f32x3 Backend::coeff_cm_multibody(MTV3f& aero_forces, MTV3f& verts_wing, MTV2f& areas, f32x3& ref_pt,
                                           f32x3& freestream, f32 rho) {
    return cm_multibody(*this, aero_forces, verts_wing, areas, ref_pt, freestream, rho);
}
Backend::f64x3 Backend::coeff_cm_multibody(MTV3d& aero_forces, MTV3d& verts_wing, MTV2d& areas, f64x3& ref_pt,
                                           f64x3& freestream, f64 rho) {
    return cm_multibody(*this, aero_forces, verts_wing, areas, ref_pt, freestream, rho);
}

}  // namespace vlm

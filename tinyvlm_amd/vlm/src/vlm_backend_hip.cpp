// vlm_backend_hip.cpp -- BackendHIP: the vlm::Backend interface implemented over the C ABI of
// libtinyvlm_hip.so (include/tinyvlm_hip.h).  Each virtual is one C-ABI call on the default
// stream; scalar results come back through a small device buffer and a synchronous copy, which is
// the by-value contract of the reference methods.  C-ABI errors become std::runtime_error (the
// reference prints and exits, vlm_backend_cpu_lapack.cpp:5-13).
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "capi_views.hpp"
#include "vlm_backend.hpp"

namespace vlm {
namespace {

using capi::check;
using capi::location;
inline constexpr auto tv = [](const auto& v) { return capi::view(v); };
template <typename V>
std::vector<tvlm_view> tvs(const std::vector<V>& m) {
    return capi::views(m);
}
int32_t loc(Location l) { return location(l); }

class HIPMemory final : public Memory {
   public:
    HIPMemory() : Memory(false) {}
    void* alloc(Location location, i64 size_bytes) const override {
        if (location == Location::Host) return std::malloc((size_t)size_bytes);
        void* p = nullptr;
        check(tvlm_malloc(&p, size_bytes, 0), "tvlm_malloc");
        return p;
    }
    void free(Location location, void* ptr) const override {
        if (location == Location::Host) return std::free(ptr);
        check(tvlm_free(ptr, 0), "tvlm_free");
    }
    void copy(Location dst_loc, void* dst, i64 dst_stride, Location src_loc, const void* src, i64 src_stride,
              i64 elem_size, i64 size) const override {
        if (dst_loc == Location::Host && src_loc == Location::Host) {
            for (i64 i = 0; i < size; i++)
                std::memcpy((char*)dst + i * dst_stride * elem_size, (const char*)src + i * src_stride * elem_size,
                            (size_t)elem_size);
            return;
        }
        check(tvlm_copy(loc(dst_loc), dst, dst_stride, loc(src_loc), src, src_stride, elem_size, size, nullptr),
              "tvlm_copy");
        check(tvlm_stream_synchronize(nullptr), "tvlm_stream_synchronize");
    }
    // host memory here is plain malloc memory, filled in place
    void fill(Location l, float* ptr, i64 stride, float value, i64 size) const override {
        if (l == Location::Host) return fill_host(ptr, stride, value, size);
        check(tvlm_fill_f32(loc(l), ptr, stride, value, size, nullptr), "tvlm_fill_f32");
    }
    void fill(Location l, double* ptr, i64 stride, double value, i64 size) const override {
        if (l == Location::Host) return fill_host(ptr, stride, value, size);
        check(tvlm_fill_f64(loc(l), ptr, stride, value, size, nullptr), "tvlm_fill_f64");
    }

   private:
    template <typename T>
    static void fill_host(T* ptr, i64 stride, T value, i64 size) {
        for (i64 i = 0; i < size; i++) ptr[i * stride] = value;
    }
};

// one device scalar buffer per object; results are copied back synchronously
class DeviceScalars {
   public:
    DeviceScalars() { check(tvlm_malloc(&m_ptr, 8 * sizeof(double), 0), "tvlm_malloc"); }
    ~DeviceScalars() { tvlm_free(m_ptr, 0); }
    template <typename T>
    T* dev() const {
        return static_cast<T*>(m_ptr);
    }
    template <typename T>
    void get(T* host, int n) const {
        check(tvlm_copy(1, host, 1, 0, m_ptr, 1, sizeof(T), n, nullptr), "tvlm_copy");
        check(tvlm_stream_synchronize(nullptr), "tvlm_stream_synchronize");
    }

   private:
    void* m_ptr = nullptr;
};

class HIPBLAS final : public BLAS {
    DeviceScalars m_s;

    template <typename T>
    void gemv_(T alpha, const TensorView<T, 2, Location::Device>& A, const TensorView<T, 1, Location::Device>& x, T beta,
               const TensorView<T, 1, Location::Device>& y, Trans trans) {
        const tvlm_view a = tv(A), xv = tv(x), yv = tv(y);
        if constexpr (sizeof(T) == 8)
            check(tvlm_gemv_f64(alpha, &a, &xv, beta, &yv, trans == Trans::Yes, nullptr), "gemv");
        else
            check(tvlm_gemv_f32(alpha, &a, &xv, beta, &yv, trans == Trans::Yes, nullptr), "gemv");
    }
    template <typename T>
    void gemm_(T alpha, const TensorView<T, 2, Location::Device>& A, const TensorView<T, 2, Location::Device>& B, T beta,
               const TensorView<T, 2, Location::Device>& C, Trans ta, Trans tb) {
        const tvlm_view a = tv(A), b = tv(B), c = tv(C);
        if constexpr (sizeof(T) == 8)
            check(tvlm_gemm_f64(alpha, &a, &b, beta, &c, ta == Trans::Yes, tb == Trans::Yes, nullptr), "gemm");
        else
            check(tvlm_gemm_f32(alpha, &a, &b, beta, &c, ta == Trans::Yes, tb == Trans::Yes, nullptr), "gemm");
    }
    template <typename T, i64 D>
    void axpy_(T alpha, const TensorView<T, D, Location::Device>& x, const TensorView<T, D, Location::Device>& y) {
        const tvlm_view xv = tv(x), yv = tv(y);
        if constexpr (sizeof(T) == 8)
            check(tvlm_axpy_f64(alpha, &xv, &yv, (int32_t)D, nullptr), "axpy");
        else
            check(tvlm_axpy_f32(alpha, &xv, &yv, (int32_t)D, nullptr), "axpy");
    }
    template <typename T>
    void scal_(T alpha, const TensorView<T, 1, Location::Device>& x) {
        const tvlm_view xv = tv(x);
        if constexpr (sizeof(T) == 8)
            check(tvlm_scal_f64(alpha, &xv, nullptr), "scal");
        else
            check(tvlm_scal_f32(alpha, &xv, nullptr), "scal");
    }
    template <typename T>
    T norm_(const TensorView<T, 1, Location::Device>& x) {
        const tvlm_view xv = tv(x);
        if constexpr (sizeof(T) == 8)
            check(tvlm_nrm2_f64(&xv, m_s.dev<double>(), nullptr), "nrm2");
        else
            check(tvlm_nrm2_f32(&xv, m_s.dev<float>(), nullptr), "nrm2");
        T r;
        m_s.get(&r, 1);
        return r;
    }

   public:
    void gemv(const f32 a, const TensorView2fD& A, const TensorView1fD& x, const f32 b, const TensorView1fD& y, Trans t) override { gemv_(a, A, x, b, y, t); }
    void gemm(const f32 a, const TensorView2fD& A, const TensorView2fD& B, const f32 b, const TensorView2fD& C, Trans ta, Trans tb) override { gemm_(a, A, B, b, C, ta, tb); }
    void axpy(const f32 a, const TensorView1fD& x, const TensorView1fD& y) override { axpy_(a, x, y); }
    void axpy(const f32 a, const TensorView2fD& x, const TensorView2fD& y) override { axpy_(a, x, y); }
    void scal(const f32 a, const TensorView1fD& x) override { scal_(a, x); }
    f32 norm(const TensorView1fD& x) override { return norm_(x); }
    void gemv(const f64 a, const TensorView2dD& A, const TensorView1dD& x, const f64 b, const TensorView1dD& y, Trans t) override { gemv_(a, A, x, b, y, t); }
    void gemm(const f64 a, const TensorView2dD& A, const TensorView2dD& B, const f64 b, const TensorView2dD& C, Trans ta, Trans tb) override { gemm_(a, A, B, b, C, ta, tb); }
    void axpy(const f64 a, const TensorView1dD& x, const TensorView1dD& y) override { axpy_(a, x, y); }
    void axpy(const f64 a, const TensorView2dD& x, const TensorView2dD& y) override { axpy_(a, x, y); }
    void scal(const f64 a, const TensorView1dD& x) override { scal_(a, x); }
    f64 norm(const TensorView1dD& x) override { return norm_(x); }
};

class HIPLU final : public LU {
    Tensor<i32, 1, Location::Device> m_ipiv{m_memory.get()};
    DeviceScalars m_s;

    template <typename T>
    void factorize_(const TensorView<T, 2, Location::Device>& A) {
        const tvlm_view a = tv(A);
        if constexpr (sizeof(T) == 8)
            check(tvlm_getrf_f64(&a, m_ipiv.ptr(), m_s.dev<int32_t>(), nullptr), "getrf");
        else
            check(tvlm_getrf_f32(&a, m_ipiv.ptr(), m_s.dev<int32_t>(), nullptr), "getrf");
        int32_t info = 0;
        m_s.get(&info, 1);
        if (info != 0) throw std::runtime_error("LU: exactly singular matrix (getrf info " + std::to_string(info) + ")");
    }
    template <typename T>
    void solve_(const TensorView<T, 2, Location::Device>& A, const TensorView<T, 2, Location::Device>& x) {
        const tvlm_view a = tv(A), xv = tv(x);
        if constexpr (sizeof(T) == 8)
            check(tvlm_getrs_f64(&a, m_ipiv.ptr(), &xv, nullptr), "getrs");
        else
            check(tvlm_getrs_f32(&a, m_ipiv.ptr(), &xv, nullptr), "getrs");
    }

   public:
    explicit HIPLU(std::unique_ptr<Memory> memory) : LU(std::move(memory)) {}
    using LU::solve;
    void init(const TensorView2fD& A) override { m_ipiv.init({A.shape(0)}); }
    void factorize(const TensorView2fD& A) override { factorize_(A); }
    void solve(const TensorView2fD& A, const TensorView2fD& x) override { solve_(A, x); }
    void init(const TensorView2dD& A) override { m_ipiv.init({A.shape(0)}); }
    void factorize(const TensorView2dD& A) override { factorize_(A); }
    void solve(const TensorView2dD& A, const TensorView2dD& x) override { solve_(A, x); }
};

// LAPACK ?gels 'N' semantics on the device (vlm_backend_cpu_lapack.cpp:96-205); no workspace to
// size, and a rank-deficient A throws like the reference's CHECK_LAPACK exit
class HIPLSQ final : public LSQ {
    DeviceScalars m_s;

    template <typename T>
    void solve_(const TensorView<T, 2, Location::Device>& A, const TensorView<T, 2, Location::Device>& B) {
        const tvlm_view a = tv(A), b = tv(B);
        if constexpr (sizeof(T) == 8)
            check(tvlm_gels_f64(&a, &b, m_s.dev<int32_t>(), nullptr), "gels");
        else
            check(tvlm_gels_f32(&a, &b, m_s.dev<int32_t>(), nullptr), "gels");
        int32_t info = 0;
        m_s.get(&info, 1);
        if (info != 0) throw std::runtime_error("LSQ: rank-deficient matrix (gels info " + std::to_string(info) + ")");
    }

   public:
    explicit HIPLSQ(std::unique_ptr<Memory> memory) : LSQ(std::move(memory)) {}
    void init(const TensorView2fD&, const TensorView2fD&) override {}
    void solve(const TensorView2fD& A, const TensorView2fD& B) override { solve_(A, B); }
    void init(const TensorView2dD&, const TensorView2dD&) override {}
    void solve(const TensorView2dD& A, const TensorView2dD& B) override { solve_(A, B); }
};

class BackendHIP final : public Backend {
    DeviceScalars m_s;

    template <typename T>
    struct P;  // C-ABI dispatch by precision

   public:
    BackendHIP() : Backend(std::make_unique<HIPMemory>(), std::make_unique<HIPBLAS>()) {
        name = "HIP";
        check(tvlm_init(0), "tvlm_init");
    }

#define TVLM_BACKEND_BODY(T, S, TV1, TV2, TV3, MTV2, MTV3, V3)                                                           \
    void lhs_assemble(TV2& lhs, MTV3& colloc, MTV3& normals, MTV3& verts_wing, MTV3& verts_wake,                          \
                      std::vector<i32>& condition, i32 iteration) override {                                             \
        const tvlm_view l = tv(lhs);                                                                                      \
        auto c = tvs(colloc), n = tvs(normals), vw = tvs(verts_wing), vk = tvs(verts_wake);                               \
        check(tvlm_lhs_assemble_##S(&l, (int32_t)c.size(), c.data(), n.data(), vw.data(), vk.data(), iteration,           \
                                    (T)sigma_vatistas, nullptr),                                                          \
              "lhs_assemble");                                                                                            \
        std::fill(condition.begin(), condition.end(), iteration); /* the reference's loop counters end at iteration */  \
    }                                                                                                                     \
    void rhs_assemble_velocities(TV1& rhs, MTV3& normals, MTV3& velocities) override {                                    \
        const tvlm_view r = tv(rhs);                                                                                      \
        auto n = tvs(normals), v = tvs(velocities);                                                                       \
        check(tvlm_rhs_assemble_velocities_##S(&r, (int32_t)n.size(), n.data(), v.data(), nullptr),                       \
              "rhs_assemble_velocities");                                                                                 \
    }                                                                                                                     \
    void rhs_assemble_wake_influence(TV1& rhs, MTV2& gamma_wake, MTV3& colloc, MTV3& normals, MTV3& verts_wake,           \
                                     std::vector<bool>& lifting, i32 iteration) override {                               \
        const tvlm_view r = tv(rhs);                                                                                      \
        auto g = tvs(gamma_wake), c = tvs(colloc), n = tvs(normals), w = tvs(verts_wake);                                 \
        std::vector<uint8_t> lift(lifting.begin(), lifting.end());                                                        \
        check(tvlm_rhs_assemble_wake_influence_##S(&r, (int32_t)c.size(), g.data(), c.data(), n.data(), w.data(),        \
                                                   lift.data(), iteration, (T)sigma_vatistas, nullptr),                   \
              "rhs_assemble_wake_influence");                                                                             \
    }                                                                                                                     \
    void forces_steady(TV3& verts_wing, TV2& gamma_delta, TV3& velocities, TV3& forces) override {                        \
        const tvlm_view v = tv(verts_wing), g = tv(gamma_delta), u = tv(velocities), f = tv(forces);                      \
        check(tvlm_forces_steady_##S(&v, &g, &u, &f, nullptr), "forces_steady");                                          \
    }                                                                                                                     \
    void forces_unsteady(TV3& verts_wing, TV2& gamma_delta, TV2& dgamma_dt, TV3& velocities, TV2& areas, TV3& normals,    \
                         TV3& forces) override {                                                                          \
        const tvlm_view v = tv(verts_wing), g = tv(gamma_delta), gt = tv(dgamma_dt), u = tv(velocities), a = tv(areas),  \
                        n = tv(normals), f = tv(forces);                                                                 \
        check(tvlm_forces_unsteady_##S(&v, &g, &gt, &u, &a, &n, &f, nullptr), "forces_unsteady");                         \
    }                                                                                                                     \
    T coeff_cl(TV3& forces, V3& lift_axis, V3& freestream, T rho, T area) override {                                     \
        const tvlm_view f = tv(forces);                                                                                   \
        const T la[3] = {lift_axis.x, lift_axis.y, lift_axis.z}, fs[3] = {freestream.x, freestream.y, freestream.z};     \
        check(tvlm_coeff_cl_##S(&f, la, fs, rho, area, m_s.dev<T>(), nullptr), "coeff_cl");                              \
        T r;                                                                                                              \
        m_s.get(&r, 1);                                                                                                   \
        return r;                                                                                                         \
    }                                                                                                                     \
    V3 coeff_cm(TV3& forces, TV3& verts_wing, V3& ref_pt, V3& freestream, T rho, T area, T mac) override {               \
        const tvlm_view f = tv(forces), v = tv(verts_wing);                                                               \
        const T rp[3] = {ref_pt.x, ref_pt.y, ref_pt.z}, fs[3] = {freestream.x, freestream.y, freestream.z};              \
        check(tvlm_coeff_cm_##S(&f, &v, rp, fs, rho, area, mac, m_s.dev<T>(), nullptr), "coeff_cm");                     \
        T r[3];                                                                                                           \
        m_s.get(r, 3);                                                                                                    \
        return V3{r[0], r[1], r[2]};                                                                                      \
    }                                                                                                                     \
    void mesh_metrics(T alpha_rad, MTV3& verts_wing, MTV3& colloc, MTV3& normals, MTV2& areas) override {               \
        auto v = tvs(verts_wing), c = tvs(colloc), n = tvs(normals), a = tvs(areas);                                      \
        check(tvlm_mesh_metrics_##S(alpha_rad, (int32_t)v.size(), v.data(), c.data(), n.data(), a.data(), nullptr),      \
              "mesh_metrics");                                                                                            \
    }                                                                                                                     \
    T mesh_mac(TV3& verts_wing, TV2& areas) override {                                                                   \
        const tvlm_view v = tv(verts_wing), a = tv(areas);                                                                \
        check(tvlm_mesh_mac_##S(&v, &a, m_s.dev<T>(), nullptr), "mesh_mac");                                             \
        T r;                                                                                                              \
        m_s.get(&r, 1);                                                                                                   \
        return r;                                                                                                         \
    }                                                                                                                     \
    void gamma_wake_from_coeffs(TV2& gamma_wake, TV2& gamma_coeffs, i32 harmonics, T tn, T omega, T dt,                   \
                                i64 iteration) override {                                                                 \
        const tvlm_view g = tv(gamma_wake), c = tv(gamma_coeffs);                                                         \
        check(tvlm_gamma_wake_from_coeffs_##S(&g, &c, harmonics, tn, omega, dt, iteration, nullptr),                      \
              "gamma_wake_from_coeffs");                                                                                  \
    }                                                                                                                     \
    T sum(TV1& x) override {                                                                                              \
        const tvlm_view v = tv(x);                                                                                        \
        check(tvlm_sum_##S(&v, 1, m_s.dev<T>(), nullptr), "sum");                                                         \
        T r;                                                                                                              \
        m_s.get(&r, 1);                                                                                                   \
        return r;                                                                                                         \
    }                                                                                                                     \
    T sum(TV2& x) override {                                                                                              \
        const tvlm_view v = tv(x);                                                                                        \
        check(tvlm_sum_##S(&v, 2, m_s.dev<T>(), nullptr), "sum");                                                         \
        T r;                                                                                                              \
        m_s.get(&r, 1);                                                                                                   \
        return r;                                                                                                         \
    }

    TVLM_BACKEND_BODY(f32, f32, TV1f, TV2f, TV3f, MTV2f, MTV3f, f32x3)
    TVLM_BACKEND_BODY(f64, f64, TV1d, TV2d, TV3d, MTV2d, MTV3d, f64x3)
#undef TVLM_BACKEND_BODY

    std::unique_ptr<Memory> create_memory_manager() override { return std::make_unique<HIPMemory>(); }
    std::unique_ptr<LU> create_lu_solver() override { return std::make_unique<HIPLU>(create_memory_manager()); }
    std::unique_ptr<BLAS> create_blas() override { return std::make_unique<HIPBLAS>(); }
    std::unique_ptr<LSQ> create_lsq_solver() override { return std::make_unique<HIPLSQ>(create_memory_manager()); }
};

}  // namespace

std::unique_ptr<Backend> create_backend(const std::string& backend_name) {
    if (backend_name == "hip") return std::make_unique<BackendHIP>();
    throw std::runtime_error("Unsupported backend: " + backend_name);
}

std::vector<std::string> get_available_backends() { return {"hip"}; }

}  // namespace vlm

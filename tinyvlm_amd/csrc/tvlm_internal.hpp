// tvlm_internal.hpp -- shared device/host helpers of the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/tinyvlm_hip.h"

namespace tvlm {

// ---- error reporting --------------------------------------------------------------------------
void set_error(const char* fmt, ...);

#define TVLM_HIP(expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            ::tvlm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,      \
                              __LINE__);                                                            \
            return e_ == hipErrorOutOfMemory ? TVLM_ERR_NOMEM : TVLM_ERR_HIP;                       \
        }                                                                                           \
    } while (0)

#define TVLM_REQUIRE(cond, ...)                                                                     \
    do {                                                                                            \
        if (!(cond)) {                                                                              \
            ::tvlm::set_error(__VA_ARGS__);                                                         \
            return TVLM_ERR_INVALID;                                                                \
        }                                                                                           \
    } while (0)

#define TVLM_TRY(expr)                                                                              \
    do {                                                                                            \
        int s_ = (expr);                                                                            \
        if (s_ != TVLM_OK) return s_;                                                               \
    } while (0)

#define TVLM_LAUNCHED() TVLM_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---- view validation --------------------------------------------------------------------------
// 3-D per-surface tensors {a, b, k} must be "plane contiguous": stride[0] == 1 and
// stride[1] == shape[0] (what the reference kernels assume, vlm_backend_cpu.cpp:111, :202).
inline bool plane_contiguous(const tvlm_view& v) {
    return v.ptr != nullptr && v.stride[0] == 1 && v.stride[1] == v.shape[0];
}

// Scratch buffers owned by the library (grow-only), one per device, slot and stream: wake-induction
// tile partials, LU row-interchange lists, triangular-solve block inverses, the exchange buffers of
// the cooperative LU panels.  Work queued on different streams never shares a buffer, so concurrent
// solves on separate streams are safe; work on one stream is ordered by the stream.
enum WorkspaceSlot { kWsWakePartials = 0, kWsLuPerm = 1, kWsSolveInv = 2, kWsLuCoop = 3, kWsLsq = 4, kWorkspaceSlots = 5 };
int device_workspace(size_t bytes, void** out, int slot, hipStream_t stream);
// Frees what the library keeps for one stream on the current device (its scratch buffers, its
// rocBLAS handle, its look-ahead side stream); the stream's work must have finished.
void release_workspaces(hipStream_t stream);
void release_lu_stream(hipStream_t stream);

// A stream with a hardware queue of its own: created with an all-CUs mask (streams created
// without a mask share a small pool of queues, and two streams on one queue run one after the
// other).  Default priority.
int create_queue_stream(hipStream_t* out);

// Looks up the device address of the allocation-free wake induction's fixed scratch array on the
// current device (once per device; tvlm_init calls it so that later calls find it set up).
int noalloc_scratch_setup();

// getrf with every panel in one work-group (no cooperative panels, hence no waits between
// work-groups): the recovery path after a cooperative panel's bounded wait gave up
int getrf_single_group_f64(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream);
int getrf_single_group_f32(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream);
constexpr int32_t kGetrfStalled = -999;  // info of a matrix whose cooperative panel gave up waiting

// Work-group targets of the wake-induction launches, library-wide (tvlm_set_tile_targets): the
// ring-by-ring form's tile shape (wake_tile_shape) and new UVLM plans' lattice launches.
constexpr int kTargetGroups = 2048;     // default of wake_target_groups()
constexpr int kLatTargetGroups = 1024;  // default of lattice_target_groups()
int wake_target_groups();
int lattice_target_groups();

// lhs_assemble in the FP64 seeded form (the served steady solve uses it; exported as
// tvlm_lhs_assemble_seeded_f64)
int lhs_assemble_seeded_f64(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                            const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, void* stream);

// ---- small vector type ------------------------------------------------------------------------
template <typename T>
struct v3 {
    T x, y, z;
};
template <typename T>
__host__ __device__ inline v3<T> operator-(const v3<T>& a, const v3<T>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T>
__host__ __device__ inline v3<T> operator+(const v3<T>& a, const v3<T>& b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T>
__host__ __device__ inline v3<T> operator*(T s, const v3<T>& a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <typename T>
__host__ __device__ inline T dot(const v3<T>& a, const v3<T>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
__host__ __device__ inline v3<T> cross(const v3<T>& a, const v3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---- Biot-Savart constants (vlm_backend_cpu_kernels.ispc:9-12, :80, :87, :393, :400) -----------
template <typename T>
struct BS;
template <>
struct BS<double> {
    static constexpr double fourpi = (double)(4.0f * 3.141593f);  // the float product 4.0f*PI_f
    static constexpr double cut_cross = (double)1e-10f;            // RCUT
    static constexpr double cut_norm = (double)1e-10f;             // FP64 uses RCUT for the norms too
};
template <>
struct BS<float> {
    static constexpr float fourpi = 4.0f * 3.141593f;
    static constexpr float cut_cross = 1e-10f;  // RCUT on |r1 x r2|^2
    static constexpr float cut_norm = 1e-5f;    // RCUT2 on |r1|, |r2|
};

// Exact form of one unit straight segment a->b at c (Katz-Plotkin eq. 10.115), operation for
// operation as kernel_biosavart / dkernel_biosavart (kernels.ispc:69-89 / :382-402) with IEEE
// sqrt and division.  With sigma == 0 the core term sqrt(sq*sq + 0) equals sq exactly in IEEE
// arithmetic, so the SIGMA=false instantiation skips that square root.
template <typename T, bool SIGMA>
__device__ __forceinline__ v3<T> segment_exact(const v3<T>& c, const v3<T>& a, const v3<T>& b, T sigma) {
    const v3<T> r0 = b - a;
    const v3<T> r1 = c - a;
    const v3<T> r2 = c - b;
    const v3<T> cr = cross(r1, r2);
    const T r1n = sqrt(dot(r1, r1));
    const T r2n = sqrt(dot(r2, r2));
    const T sq = dot(cr, cr);
    if ((sq < BS<T>::cut_cross) || (r1n < BS<T>::cut_norm) || (r2n < BS<T>::cut_norm)) return {T(0), T(0), T(0)};
    T core;
    if constexpr (SIGMA) {
        const T sm = sigma * sigma * dot(r0, r0);
        core = sqrt(sq * sq + sm * sm);
    } else {
        core = sq;
    }
    const T coeff = (dot(r0, r1) * r2n - dot(r0, r2) * r1n) / (BS<T>::fourpi * core * r1n * r2n);
    return coeff * cr;
}

// Segment + its image across y = 0 (kernel_symmetry, kernels.ispc:91-102).
template <typename T, bool SIGMA>
__device__ __forceinline__ void symmetry_exact(v3<T>& inf, const v3<T>& c, const v3<T>& a, const v3<T>& b, T sigma) {
    const v3<T> s = segment_exact<T, SIGMA>(c, a, b, sigma);
    inf.x += s.x;
    inf.y += s.y;
    inf.z += s.z;
    const v3<T> cm = {c.x, -c.y, c.z};
    const v3<T> t = segment_exact<T, SIGMA>(cm, a, b, sigma);
    inf.x += t.x;
    inf.y -= t.y;
    inf.z += t.z;
}

// Unit vortex ring v0->v1->v2->v3->v0 with images (kernel_influence body, kernels.ispc:127-133).
template <typename T, bool SIGMA>
__device__ __forceinline__ v3<T> ring_exact(const v3<T>& c, const v3<T>& v0, const v3<T>& v1, const v3<T>& v2,
                                            const v3<T>& v3_, T sigma) {
    v3<T> inf = {T(0), T(0), T(0)};
    symmetry_exact<T, SIGMA>(inf, c, v0, v1, sigma);
    symmetry_exact<T, SIGMA>(inf, c, v1, v2, sigma);
    symmetry_exact<T, SIGMA>(inf, c, v2, v3_, sigma);
    symmetry_exact<T, SIGMA>(inf, c, v3_, v0, sigma);
    return inf;
}

// ---- panel metrics (vlm_backend_cpu.cpp:391-433, vlm_backend_cpud.cpp:258-300) -----------------
// Normal normalize((v3-v1) x (v2-v0)); area 0.5(|f x b| + |b x e|); collocation point
// 0.5(v0+v1) + k * 0.5(v2+v3-v0-v1), k = 1/2 below EPS_f else alpha / (2 (sin alpha + EPS_f)).
template <typename T>
__device__ __forceinline__ void panel_metrics(T alpha, const v3<T>& v0, const v3<T>& v1, const v3<T>& v2,
                                              const v3<T>& v3_, v3<T>& colloc, v3<T>& normal, T& area) {
    const v3<T> nv = cross(v3_ - v1, v2 - v0);
    const T l = sqrt(dot(nv, nv));
    normal = {nv.x / l, nv.y / l, nv.z / l};
    const v3<T> f = v3_ - v0, b = v2 - v0, e = v1 - v0;
    const v3<T> fb = cross(f, b), be = cross(b, e);
    area = T(0.5) * (sqrt(dot(fb, fb)) + sqrt(dot(be, be)));
    const T eps = (T)1.19209e-07f;  // EPS_f (vlm_types.hpp:24) in both precisions
    const T factor = (alpha < eps) ? T(0.5) : T(0.5) * (alpha / (sin(alpha) + eps));
    const v3<T> chord = T(0.5) * (((v2 + v3_) - v0) - v1);
    colloc = T(0.5) * (v0 + v1) + factor * chord;
}

// One homogeneous point times a column-major 4x4 (row r of verts {P,4} * T^T, vlm_backend.cpp:59-67).
template <typename T>
__device__ __forceinline__ T displace_row(const T* M, int64_t m0, int64_t m1, int r, const T* x) {
    T s = x[0] * M[r * m0 + 0 * m1];
    s += x[1] * M[r * m0 + 1 * m1];
    s += x[2] * M[r * m0 + 2 * m1];
    s += x[3] * M[r * m0 + 3 * m1];
    return s;
}

// ---- FP32 fast form ---------------------------------------------------------------------------
// Same mathematics with the hardware reciprocal square root: coeff = (d1/|r1| - d2/|r2|) / (4 pi sq),
// which equals the reference's (d1 |r2| - d2 |r1|) / (4 pi sq |r1| |r2|).  Cutoffs are tested on
// squared norms: |r| < 1e-5  <=>  |r|^2 < 1e-10, and sq < 1e-10 as in the reference.  Each ring
// vertex difference and its rsqrt is computed once and shared by the two segments that meet there.
struct FastVertex {
    float x, y, z, r2, rinv;  // c - v, |c - v|^2, 1/|c - v|
};

__device__ __forceinline__ FastVertex fast_vertex(float cx, float cy, float cz, float vx, float vy, float vz) {
    FastVertex f;
    f.x = cx - vx;
    f.y = cy - vy;
    f.z = cz - vz;
    f.r2 = f.x * f.x + f.y * f.y + f.z * f.z;
    f.rinv = __builtin_amdgcn_rsqf(f.r2);
    return f;
}

// Accumulates s * (velocity of segment a->b) into acc; s = +1 / -1 mixes the y image sign.
__device__ __forceinline__ void fast_segment(float& ax, float& ay, float& az, const FastVertex& r1, const FastVertex& r2,
                                             float sy) {
    const float crx = r1.y * r2.z - r1.z * r2.y;
    const float cry = r1.z * r2.x - r1.x * r2.z;
    const float crz = r1.x * r2.y - r1.y * r2.x;
    const float sq = crx * crx + cry * cry + crz * crz;
    // r0 = b - a = r1 - r2 (as points: (c-a) - (c-b))
    const float r0x = r1.x - r2.x, r0y = r1.y - r2.y, r0z = r1.z - r2.z;
    const float d1 = r0x * r1.x + r0y * r1.y + r0z * r1.z;
    const float d2 = r0x * r2.x + r0y * r2.y + r0z * r2.z;
    const bool ok = (sq >= 1e-10f) & (r1.r2 >= 1e-10f) & (r2.r2 >= 1e-10f);
    const float t = (d1 * r1.rinv - d2 * r2.rinv) * __builtin_amdgcn_rcpf(BS<float>::fourpi * sq);
    const float coeff = ok ? t : 0.0f;
    ax += coeff * crx;
    ay += sy * (coeff * cry);
    az += coeff * crz;
}

// ---- FP64 seeded form ---------------------------------------------------------------------------
// FP64 arithmetic throughout, but 1/sqrt and 1/x from the hardware estimates (v_rsq_f64,
// v_rcp_f64) refined by one Newton step in FP64 instead of the multi-instruction correctly rounded
// FP64 sqrt and division (the Newton step squares the estimate's relative error).  MI355X runs
// FP64 add/mul/fma at the rate of non-packed FP32, so this costs about what the FP32 form costs
// while keeping FP64 accuracy.  Cutoffs as the FP64 reference: |r| < RCUT <=> |r|^2 < RCUT^2.
__device__ __forceinline__ double rsqrt_seeded(double x) {
    const double y = __builtin_amdgcn_rsq(x);  // v_rsq_f64 estimate
    return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ double rcp_seeded(double x) {
    const double y = __builtin_amdgcn_rcp(x);  // v_rcp_f64 estimate
    return y * (2.0 - x * y);
}

struct SeedVertex {
    double x, y, z, r2, rinv;
};

__device__ __forceinline__ SeedVertex seed_vertex(double cx, double cy, double cz, double vx, double vy, double vz) {
    SeedVertex f;
    f.x = cx - vx;
    f.y = cy - vy;
    f.z = cz - vz;
    f.r2 = f.x * f.x + f.y * f.y + f.z * f.z;
    f.rinv = rsqrt_seeded(f.r2);
    return f;
}

__device__ __forceinline__ void seed_segment(double& ax, double& ay, double& az, const SeedVertex& r1,
                                             const SeedVertex& r2, double sy) {
    const double crx = r1.y * r2.z - r1.z * r2.y;
    const double cry = r1.z * r2.x - r1.x * r2.z;
    const double crz = r1.x * r2.y - r1.y * r2.x;
    const double sq = crx * crx + cry * cry + crz * crz;
    const double r0x = r1.x - r2.x, r0y = r1.y - r2.y, r0z = r1.z - r2.z;
    const double d1 = r0x * r1.x + r0y * r1.y + r0z * r1.z;
    const double d2 = r0x * r2.x + r0y * r2.y + r0z * r2.z;
    constexpr double cut2 = BS<double>::cut_norm * BS<double>::cut_norm;
    const bool ok = (sq >= BS<double>::cut_cross) & (r1.r2 >= cut2) & (r2.r2 >= cut2);
    const double t = (d1 * r1.rinv - d2 * r2.rinv) * rcp_seeded(BS<double>::fourpi * sq);
    const double coeff = ok ? t : 0.0;
    ax += coeff * crx;
    ay += sy * (coeff * cry);
    az += coeff * crz;
}

// Segment a->b of strength g (r1 = c - a, r2 = c - b), accumulated with image sign sy on y.
__device__ __forceinline__ void seed_segment_w(double& ax, double& ay, double& az, const SeedVertex& r1,
                                               const SeedVertex& r2, double sy, double g) {
    const double crx = r1.y * r2.z - r1.z * r2.y;
    const double cry = r1.z * r2.x - r1.x * r2.z;
    const double crz = r1.x * r2.y - r1.y * r2.x;
    const double sq = crx * crx + cry * cry + crz * crz;
    const double r0x = r1.x - r2.x, r0y = r1.y - r2.y, r0z = r1.z - r2.z;
    const double d1 = r0x * r1.x + r0y * r1.y + r0z * r1.z;
    const double d2 = r0x * r2.x + r0y * r2.y + r0z * r2.z;
    constexpr double cut2 = BS<double>::cut_norm * BS<double>::cut_norm;
    const bool ok = (sq >= BS<double>::cut_cross) & (r1.r2 >= cut2) & (r2.r2 >= cut2);
    const double t = (d1 * r1.rinv - d2 * r2.rinv) * rcp_seeded(BS<double>::fourpi * sq) * g;
    const double coeff = ok ? t : 0.0;
    ax += coeff * crx;
    ay += sy * (coeff * cry);
    az += coeff * crz;
}

__device__ __forceinline__ void seed_ring(double& ax, double& ay, double& az, double cx, double cy, double cz,
                                          const double* v0, const double* v1, const double* v2, const double* v3) {
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const double y = m == 0 ? cy : -cy;
        const double sy = m == 0 ? 1.0 : -1.0;
        const SeedVertex a = seed_vertex(cx, y, cz, v0[0], v0[1], v0[2]);
        const SeedVertex b = seed_vertex(cx, y, cz, v1[0], v1[1], v1[2]);
        const SeedVertex c = seed_vertex(cx, y, cz, v2[0], v2[1], v2[2]);
        const SeedVertex d = seed_vertex(cx, y, cz, v3[0], v3[1], v3[2]);
        seed_segment(ax, ay, az, a, b, sy);
        seed_segment(ax, ay, az, b, c, sy);
        seed_segment(ax, ay, az, c, d, sy);
        seed_segment(ax, ay, az, d, a, sy);
    }
}

// Ring with image; vertices given as coordinates.
__device__ __forceinline__ void fast_ring(float& ax, float& ay, float& az, float cx, float cy, float cz, const float* v0,
                                          const float* v1, const float* v2, const float* v3) {
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const float y = m == 0 ? cy : -cy;
        const float sy = m == 0 ? 1.0f : -1.0f;
        const FastVertex a = fast_vertex(cx, y, cz, v0[0], v0[1], v0[2]);
        const FastVertex b = fast_vertex(cx, y, cz, v1[0], v1[1], v1[2]);
        const FastVertex c = fast_vertex(cx, y, cz, v2[0], v2[1], v2[2]);
        const FastVertex d = fast_vertex(cx, y, cz, v3[0], v3[1], v3[2]);
        fast_segment(ax, ay, az, a, b, sy);
        fast_segment(ax, ay, az, b, c, sy);
        fast_segment(ax, ay, az, c, d, sy);
        fast_segment(ax, ay, az, d, a, sy);
    }
}

}  // namespace tvlm

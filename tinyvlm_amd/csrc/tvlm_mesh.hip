// tvlm_mesh.hip -- per-panel operations, deterministic reductions, small BLAS and memory.
//
// Everything here is O(N) (or O(N) x 4x4) and HBM/launch bound: one thread per panel or vertex,
// column-major planes read with unit stride.  Reductions (CL, CM, MAC, sums, norms) run in one
// work-group with a fixed summation tree so they are bitwise reproducible run to run; the CUDA
// reference uses float atomics there (vlm_backend_cuda_kernels.cuh:89, :451, :501-503).
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include <map>
#include <tuple>

#include <hip/hip_ext.h>

#include "tvlm_internal.hpp"

namespace tvlm {

// ---- error text -------------------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ---- grow-only workspaces, one per (device, slot, stream) ----------------------------------------
namespace {
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::map<std::tuple<int, int, void*>, Workspace> g_ws;

}  // namespace

void release_workspaces(hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    for (int slot = 0; slot < kWorkspaceSlots; slot++) {
        auto it = g_ws.find(std::make_tuple(dev, slot, static_cast<void*>(stream)));
        if (it == g_ws.end()) continue;
        if (it->second.ptr) (void)hipFree(it->second.ptr);
        g_ws.erase(it);
    }
}

int create_queue_stream(hipStream_t* out) {
    int dev = 0, cus = 0;
    TVLM_HIP(hipGetDevice(&dev));
    TVLM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<uint32_t> mask((cus + 31) / 32, 0u);
    for (int c = 0; c < cus; c++) mask[c / 32] |= 1u << (c % 32);
    TVLM_HIP(hipExtStreamCreateWithCUMask(out, (uint32_t)mask.size(), mask.data()));
    return TVLM_OK;
}

int device_workspace(size_t bytes, void** out, int slot, hipStream_t stream) {
    TVLM_REQUIRE(slot >= 0 && slot < kWorkspaceSlots, "workspace slot %d out of range", slot);
    int dev = 0;
    TVLM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace& w = g_ws[std::make_tuple(dev, slot, static_cast<void*>(stream))];
    if (w.bytes < bytes) {
        if (w.ptr) {
            // a previous launch may still read the old buffer
            TVLM_HIP(hipDeviceSynchronize());
            TVLM_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        size_t b = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
        TVLM_HIP(hipMalloc(&w.ptr, b));
        w.bytes = b;
    }
    *out = w.ptr;
    return TVLM_OK;
}

namespace {

constexpr int kRedThreads = 1024;

// Fixed-order block sum: each thread already holds its strided partial; tree in LDS.
template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v, T* lds) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) lds[t] = lds[t] + lds[t + s];
        __syncthreads();
    }
    const T r = lds[0];
    __syncthreads();
    return r;
}

// ---- mesh metrics (vlm_backend_cpu.cpp:391-433, vlm_backend_cpud.cpp:258-300) -------------------
template <typename T>
__global__ void k_mesh_metrics(T alpha, const T* __restrict__ V, int64_t vs2, int ns, int nc, T* __restrict__ colloc,
                               int64_t cs2, T* __restrict__ normals, int64_t ns2, T* __restrict__ areas,
                               int64_t as1) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ns * nc) return;
    const int i = p % ns, j = p / ns;
    const int64_t vs1 = ns + 1;
    auto vtx = [&](int ii, int jj) -> v3<T> {
        const int64_t o = ii + jj * vs1;
        return {V[o], V[o + vs2], V[o + 2 * vs2]};
    };
    const v3<T> v0 = vtx(i, j), v1 = vtx(i + 1, j), v2 = vtx(i + 1, j + 1), v3_ = vtx(i, j + 1);
    v3<T> cp, nrm;
    T area;
    panel_metrics(alpha, v0, v1, v2, v3_, cp, nrm, area);
    normals[p] = nrm.x;
    normals[p + ns2] = nrm.y;
    normals[p + 2 * ns2] = nrm.z;
    areas[i + (int64_t)j * as1] = area;
    colloc[p] = cp.x;
    colloc[p + cs2] = cp.y;
    colloc[p + 2 * cs2] = cp.z;
}

// ---- rhs velocities (vlm_backend_cpu.cpp:156-179) -------------------------------------------------
template <typename T>
__global__ void k_rhs_velocities(T* __restrict__ rhs, int64_t rs, const T* __restrict__ n, int64_t n2,
                                 const T* __restrict__ v, int64_t v2, int np) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    rhs[p * rs] += -(v[p] * n[p] + v[p + v2] * n[p + n2] + v[p + 2 * v2] * n[p + 2 * n2]);
}

// ---- forces (vlm_backend_cpu.cpp:276-321) ---------------------------------------------------------
template <typename T, bool UNSTEADY>
__global__ void k_forces(const T* __restrict__ V, int64_t vs2, int ns, int nc, const T* __restrict__ gd, int64_t gd1,
                         const T* __restrict__ gdt, int64_t gdt1, const T* __restrict__ vel, int64_t vel2,
                         const T* __restrict__ areas, int64_t a1, const T* __restrict__ nrm, int64_t n2,
                         T* __restrict__ F, int64_t f2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ns * nc) return;
    const int i = p % ns, j = p / ns;
    const T rho = T(1);  // hard-coded in the reference (vlm_backend_cpu.cpp:282, :310)
    const int64_t a = i + (int64_t)j * (ns + 1), b = a + 1;
    const v3<T> U = {vel[p], vel[p + vel2], vel[p + 2 * vel2]};
    const v3<T> d = {V[b] - V[a], V[b + vs2] - V[a + vs2], V[b + 2 * vs2] - V[a + 2 * vs2]};
    const v3<T> cr = cross(U, d);
    const T s = rho * gd[i + (int64_t)j * gd1];
    T fx = s * cr.x, fy = s * cr.y, fz = s * cr.z;
    if constexpr (UNSTEADY) {
        const T u = rho * gdt[i + (int64_t)j * gdt1] * areas[i + (int64_t)j * a1];
        fx += u * nrm[p];
        fy += u * nrm[p + n2];
        fz += u * nrm[p + 2 * n2];
    }
    F[p] = fx;
    F[p + f2] = fy;
    F[p + 2 * f2] = fz;
}

// ---- coefficient reductions (vlm_backend_cpu.cpp:323-362, :443-468, :506-523) ---------------------
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_coeff_cl(const T* __restrict__ F, int64_t f2, int np, T lx, T ly, T lz,
                                                          T den, T* __restrict__ out) {
    __shared__ T lds[kRedThreads];
    T s = T(0);
    for (int p = threadIdx.x; p < np; p += kRedThreads) s += F[p] * lx + F[p + f2] * ly + F[p + 2 * f2] * lz;
    s = block_sum<T, kRedThreads>(s, lds);
    if (threadIdx.x == 0) out[0] = s / den;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_coeff_cm(const T* __restrict__ F, int64_t f2, const T* __restrict__ V,
                                                          int64_t vs2, int ns, int nc, T rx, T ry, T rz, T den,
                                                          T* __restrict__ out) {
    __shared__ T lds[kRedThreads];
    T sx = T(0), sy = T(0), sz = T(0);
    for (int p = threadIdx.x; p < ns * nc; p += kRedThreads) {
        const int i = p % ns, j = p / ns;
        const int64_t a = i + (int64_t)j * (ns + 1), b = a + 1;
        const v3<T> lever = {T(0.5) * (V[a] + V[b]) - rx, T(0.5) * (V[a + vs2] + V[b + vs2]) - ry,
                             T(0.5) * (V[a + 2 * vs2] + V[b + 2 * vs2]) - rz};
        const v3<T> f = {F[p], F[p + f2], F[p + 2 * f2]};
        const v3<T> c = cross(lever, f);
        sx += c.x;
        sy += c.y;
        sz += c.z;
    }
    sx = block_sum<T, kRedThreads>(sx, lds);
    sy = block_sum<T, kRedThreads>(sy, lds);
    sz = block_sum<T, kRedThreads>(sz, lds);
    if (threadIdx.x == 0) {
        out[0] = sx / den;
        out[1] = sy / den;
        out[2] = sz / den;
    }
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_mesh_mac(const T* __restrict__ V, int64_t vs2, int ns, int nc,
                                                          const T* __restrict__ areas, int64_t a1,
                                                          T* __restrict__ out) {
    __shared__ T lds[kRedThreads];
    T mac = T(0);
    const int64_t te = (int64_t)nc * (ns + 1);
    for (int i = threadIdx.x; i < ns; i += kRedThreads) {
        v3<T> d0, d1, d3;
        d0 = {V[i] - V[i + te], V[i + vs2] - V[i + te + vs2], V[i + 2 * vs2] - V[i + te + 2 * vs2]};
        d1 = {V[i + 1] - V[i + 1 + te], V[i + 1 + vs2] - V[i + 1 + te + vs2], V[i + 1 + 2 * vs2] - V[i + 1 + te + 2 * vs2]};
        d3 = {V[i + 1] - V[i], V[i + 1 + vs2] - V[i + vs2], V[i + 1 + 2 * vs2] - V[i + 2 * vs2]};
        const T c0 = sqrt(dot(d0, d0)), c1 = sqrt(dot(d1, d1)), w = sqrt(dot(d3, d3));
        mac += T(0.5) * (c0 * c0 + c1 * c1) * w;
    }
    T area = T(0);
    for (int p = threadIdx.x; p < ns * nc; p += kRedThreads) area += areas[(p % ns) + (int64_t)(p / ns) * a1];
    mac = block_sum<T, kRedThreads>(mac, lds);
    area = block_sum<T, kRedThreads>(area, lds);
    if (threadIdx.x == 0) out[0] = mac / area;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_sum2d(const T* __restrict__ x, int64_t s0, int64_t s1, int64_t n0,
                                                       int64_t n1, T* __restrict__ out, int sqr) {
    __shared__ T lds[kRedThreads];
    T s = T(0);
    const int64_t n = n0 * n1;
    for (int64_t q = threadIdx.x; q < n; q += kRedThreads) {
        const T v = x[(q % n0) * s0 + (q / n0) * s1];
        s += sqr ? v * v : v;
    }
    s = block_sum<T, kRedThreads>(s, lds);
    if (threadIdx.x == 0) out[0] = sqr ? sqrt(s) : s;
}

// ---- wake shed / displace ---------------------------------------------------------------------
template <typename T>
__global__ void k_wake_shed(const T* __restrict__ V, int64_t vs2, int ns, int nc, T* __restrict__ W, int64_t ws2,
                            int row) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * (ns + 1)) return;
    const int i = e % (ns + 1), k = e / (ns + 1);
    W[i + (int64_t)row * (ns + 1) + k * ws2] = V[i + (int64_t)nc * (ns + 1) + k * vs2];
}

template <typename T>
__global__ void k_displace(const T* __restrict__ M, int64_t m0, int64_t m1, const T* __restrict__ X, int64_t xs2,
                           T* __restrict__ Y, int64_t ys2, int64_t nverts) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nverts) return;
    const T x[4] = {X[p], X[p + xs2], X[p + 2 * xs2], X[p + 3 * xs2]};
#pragma unroll
    for (int r = 0; r < 4; r++) Y[p + r * ys2] = displace_row(M, m0, m1, r, x);
}

// Wake vertex rows of a whole prescribed motion: element e = (span vertex j, step i) of one surface
// writes wake row nwv-1-i with that surface's trailing-edge vertex j of X displaced by the step's
// transform M + (i n_surf + m) m2 -- the values k_displace then k_wake_shed leave there.
template <typename T>
__global__ void k_wake_geometry(const T* __restrict__ M, int64_t m0, int64_t m1, int64_t m2, int n_surf, int m,
                                const T* __restrict__ X, int64_t xs2, int ns, int nc, T* __restrict__ W, int64_t ws2,
                                int nwv, int steps) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)steps * (ns + 1)) return;
    const int j = (int)(e % (ns + 1)), i = (int)(e / (ns + 1));
    const T* Mi = M + ((int64_t)i * n_surf + m) * m2;
    const int64_t p = j + (int64_t)nc * (ns + 1);
    const T x[4] = {X[p], X[p + xs2], X[p + 2 * xs2], X[p + 3 * xs2]};
    const int64_t w = j + (int64_t)(nwv - 1 - i) * (ns + 1);
#pragma unroll
    for (int r = 0; r < 4; r++) W[w + r * ws2] = displace_row(Mi, m0, m1, r, x);
}

// ---- gamma_wake_from_coeffs (vlm_backend_cpu.cpp:470-503) -----------------------------------------
template <typename T>
__global__ void k_gamma_wake_coeffs(T* __restrict__ gw, int64_t g1, const T* __restrict__ gc, int64_t c1, int ns,
                                    int64_t start, int64_t nw, int harmonics, T tn, T omega, T dt, T s1, T s2) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t cnt = (nw - start) * ns;
    if (e >= cnt) return;
    const int i = (int)(e % ns);
    const int64_t j = start + e / ns;
    T g = gc[i] * s1;
    for (int h = 0; h < harmonics; h++) {
        const T wk = omega * (T)(h + 1);
        const T arg = wk * (tn - (T)(j - start + 1) * dt);
        g += gc[i + (2 * h + 1) * c1] * cos(arg) * s2;
        g += gc[i + (2 * h + 2) * c1] * sin(arg) * s2;
    }
    gw[i + j * g1] = g;
}

// ---- kinematic velocities (device form of vlm/src/vlm.cpp:419-436) ------------------------------
template <typename T>
struct Mat4 {
    T m[16];  // column-major
};

template <typename T>
__global__ void k_kin_vel(Mat4<T> d, const T* __restrict__ c, int64_t c2, T* __restrict__ v, int64_t v2, int np) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const T x = c[p], y = c[p + c2], z = c[p + 2 * c2];
#pragma unroll
    for (int r = 0; r < 3; r++) v[p + r * v2] = -(d.m[r] * x + d.m[r + 4] * y + d.m[r + 8] * z + d.m[r + 12]);
}

// ---- BLAS-lite ------------------------------------------------------------------------------------
template <typename T>
__global__ void k_axpy2d(T alpha, const T* __restrict__ x, int64_t x0, int64_t x1, T* __restrict__ y, int64_t y0,
                         int64_t y1, int64_t n0, int64_t n1) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n0 * n1) return;
    const int64_t i = q % n0, j = q / n0;
    y[i * y0 + j * y1] += alpha * x[i * x0 + j * x1];
}

template <typename T>
__global__ void k_scal(T alpha, T* __restrict__ x, int64_t s, int64_t n) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) x[q * s] *= alpha;
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, int64_t s, T v, int64_t n) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) x[q * s] = v;
}

// C = alpha op(A) op(B) + beta C, 16x16 tiles through LDS (small operands only: displacement,
// harmonic-balance DFT products).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm(int m, int n, int k, T alpha, const T* __restrict__ A, int64_t lda, int ta,
                                              const T* __restrict__ B, int64_t ldb, int tb, T beta, T* __restrict__ C,
                                              int64_t ldc) {
    __shared__ T sa[16][17], sb[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.x * 16 + tx, col = blockIdx.y * 16 + ty;
    T acc = T(0);
    for (int k0 = 0; k0 < k; k0 += 16) {
        {   // A tile: rows blockIdx.x*16 + tx, k index k0 + ty
            const int r = blockIdx.x * 16 + tx, kk = k0 + ty;
            sa[ty][tx] = (r < m && kk < k) ? (ta ? A[kk + (int64_t)r * lda] : A[r + (int64_t)kk * lda]) : T(0);
        }
        {   // B tile: k index k0 + tx, cols blockIdx.y*16 + ty
            const int kk = k0 + tx, c = blockIdx.y * 16 + ty;
            sb[ty][tx] = (kk < k && c < n) ? (tb ? B[c + (int64_t)kk * ldb] : B[kk + (int64_t)c * ldb]) : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) acc += sa[q][tx] * sb[ty][q];
        __syncthreads();
    }
    if (row < m && col < n) {
        T* c = C + row + (int64_t)col * ldc;
        *c = beta == T(0) ? alpha * acc : alpha * acc + beta * *c;
    }
}

template <typename T>
__global__ void k_gemv(int m, int n, T alpha, const T* __restrict__ A, int64_t lda, int trans, const T* __restrict__ x,
                       int64_t xs, T beta, T* __restrict__ y, int64_t ys) {
    // y(i) over the output length; trans=0: y = alpha A x + beta y (A m x n), else y = alpha A^T x + beta y
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int out_len = trans ? n : m, in_len = trans ? m : n;
    if (i >= out_len) return;
    T s = T(0);
    if (!trans) {
        for (int j = 0; j < in_len; j++) s += A[i + (int64_t)j * lda] * x[(int64_t)j * xs];
    } else {
        const T* col = A + (int64_t)i * lda;
        for (int j = 0; j < in_len; j++) s += col[j] * x[(int64_t)j * xs];
    }
    T* yi = y + (int64_t)i * ys;
    *yi = beta == T(0) ? alpha * s : alpha * s + beta * *yi;
}

inline dim3 blocks_for(int64_t n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

template <typename T>
int mesh_metrics(T alpha, int32_t n_surf, const tvlm_view* verts, const tvlm_view* colloc, const tvlm_view* normals,
                 const tvlm_view* areas, void* stream) {
    TVLM_REQUIRE(verts && colloc && normals && areas, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    for (int m = 0; m < n_surf; m++) {
        const int ns = (int)colloc[m].shape[0], nc = (int)colloc[m].shape[1];
        TVLM_REQUIRE(plane_contiguous(verts[m]) && verts[m].shape[0] == ns + 1 && verts[m].shape[1] == nc + 1,
                     "verts_wing[%d] shape/stride", m);
        TVLM_REQUIRE(plane_contiguous(colloc[m]) && plane_contiguous(normals[m]) && normals[m].shape[0] == ns &&
                         normals[m].shape[1] == nc,
                     "colloc/normals[%d] shape/stride", m);
        TVLM_REQUIRE(areas[m].stride[0] == 1 && areas[m].shape[0] == ns && areas[m].shape[1] == nc, "areas[%d]", m);
        hipLaunchKernelGGL(k_mesh_metrics<T>, blocks_for((int64_t)ns * nc), dim3(256), 0, as_stream(stream), alpha,
                           (const T*)verts[m].ptr, verts[m].stride[2], ns, nc, (T*)colloc[m].ptr, colloc[m].stride[2],
                           (T*)normals[m].ptr, normals[m].stride[2], (T*)areas[m].ptr, areas[m].stride[1]);
        TVLM_LAUNCHED();
    }
    return TVLM_OK;
}

template <typename T>
int rhs_velocities(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* normals, const tvlm_view* vel, void* stream) {
    TVLM_REQUIRE(rhs && normals && vel, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    int64_t off = 0;
    for (int m = 0; m < n_surf; m++) {
        const int64_t np = normals[m].shape[0] * normals[m].shape[1];
        TVLM_REQUIRE(plane_contiguous(normals[m]) && plane_contiguous(vel[m]) &&
                         vel[m].shape[0] == normals[m].shape[0] && vel[m].shape[1] == normals[m].shape[1],
                     "normals/velocities[%d] shape/stride", m);
        TVLM_REQUIRE(off + np <= rhs->shape[0], "rhs too short");
        hipLaunchKernelGGL(k_rhs_velocities<T>, blocks_for(np), dim3(256), 0, as_stream(stream),
                           (T*)rhs->ptr + off * rhs->stride[0], rhs->stride[0], (const T*)normals[m].ptr,
                           normals[m].stride[2], (const T*)vel[m].ptr, vel[m].stride[2], (int)np);
        TVLM_LAUNCHED();
        off += np;
    }
    return TVLM_OK;
}

template <typename T, bool UNSTEADY>
int forces(const tvlm_view* verts, const tvlm_view* gd, const tvlm_view* gdt, const tvlm_view* vel, const tvlm_view* areas,
           const tvlm_view* nrm, const tvlm_view* F, void* stream) {
    TVLM_REQUIRE(verts && gd && vel && F, "null argument");
    const int ns = (int)gd->shape[0], nc = (int)gd->shape[1];
    TVLM_REQUIRE(gd->stride[0] == 1, "gamma_delta stride[0] must be 1");
    TVLM_REQUIRE(plane_contiguous(*verts) && verts->shape[0] == ns + 1 && verts->shape[1] == nc + 1, "verts_wing");
    TVLM_REQUIRE(plane_contiguous(*vel) && vel->shape[0] == ns && vel->shape[1] == nc, "velocities");
    TVLM_REQUIRE(plane_contiguous(*F) && F->shape[0] == ns && F->shape[1] == nc, "forces");
    const T* gdtp = nullptr;
    const T* ap = nullptr;
    const T* np_ = nullptr;
    int64_t gdt1 = 0, a1 = 0, n2 = 0;
    if (UNSTEADY) {
        TVLM_REQUIRE(gdt && areas && nrm, "null argument");
        TVLM_REQUIRE(gdt->stride[0] == 1 && areas->stride[0] == 1 && plane_contiguous(*nrm), "dgamma_dt/areas/normals");
        gdtp = (const T*)gdt->ptr;
        gdt1 = gdt->stride[1];
        ap = (const T*)areas->ptr;
        a1 = areas->stride[1];
        np_ = (const T*)nrm->ptr;
        n2 = nrm->stride[2];
    }
    hipLaunchKernelGGL((k_forces<T, UNSTEADY>), blocks_for((int64_t)ns * nc), dim3(256), 0, as_stream(stream),
                       (const T*)verts->ptr, verts->stride[2], ns, nc, (const T*)gd->ptr, gd->stride[1], gdtp, gdt1,
                       (const T*)vel->ptr, vel->stride[2], ap, a1, np_, n2, (T*)F->ptr, F->stride[2]);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

}  // namespace

}  // namespace tvlm

namespace tvlm {
namespace {
std::atomic<int> g_wake_target{kTargetGroups}, g_lattice_target{kLatTargetGroups};
}  // namespace
int wake_target_groups() { return g_wake_target.load(std::memory_order_relaxed); }
int lattice_target_groups() { return g_lattice_target.load(std::memory_order_relaxed); }
}  // namespace tvlm

using namespace tvlm;

extern "C" {

int tvlm_set_tile_targets(int32_t wake_groups, int32_t lattice_groups) {
    TVLM_REQUIRE((wake_groups == 0 || (wake_groups >= 64 && wake_groups <= 65536)) &&
                     (lattice_groups == 0 || (lattice_groups >= 64 && lattice_groups <= 65536)),
                 "targets %d / %d out of range [64, 65536] (0: default)", wake_groups, lattice_groups);
    g_wake_target.store(wake_groups ? wake_groups : kTargetGroups);
    g_lattice_target.store(lattice_groups ? lattice_groups : kLatTargetGroups);
    return TVLM_OK;
}

int tvlm_get_tile_targets(int32_t* wake_groups, int32_t* lattice_groups) {
    if (wake_groups) *wake_groups = wake_target_groups();
    if (lattice_groups) *lattice_groups = lattice_target_groups();
    return TVLM_OK;
}

const char* tvlm_last_error(void) { return g_last_error.c_str(); }
const char* tvlm_backend_name(void) { return "hip"; }

int tvlm_init(int32_t device) {
    int n = 0;
    TVLM_HIP(hipGetDeviceCount(&n));
    TVLM_REQUIRE(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
    TVLM_HIP(hipSetDevice(device));
    TVLM_TRY(noalloc_scratch_setup());
    return TVLM_OK;
}

int tvlm_stream_create(void** stream) {
    TVLM_REQUIRE(stream, "null argument");
    hipStream_t s = nullptr;
    TVLM_TRY(create_queue_stream(&s));
    *stream = s;
    return TVLM_OK;
}

int tvlm_stream_destroy(void* stream) {
    if (!stream) return TVLM_OK;
    TVLM_HIP(hipStreamSynchronize(as_stream(stream)));
    release_lu_stream(as_stream(stream));
    release_workspaces(as_stream(stream));
    TVLM_HIP(hipStreamDestroy(as_stream(stream)));
    return TVLM_OK;
}

int tvlm_stream_query(void* stream, int32_t* done) {
    TVLM_REQUIRE(done, "null argument");
    const hipError_t e = hipStreamQuery(as_stream(stream));
    *done = e == hipSuccess ? 1 : 0;
    if (e != hipErrorNotReady) TVLM_HIP(e);
    return TVLM_OK;
}

int tvlm_stream_synchronize(void* stream) {
    TVLM_HIP(hipStreamSynchronize(as_stream(stream)));
    return TVLM_OK;
}

int tvlm_malloc(void** ptr, int64_t bytes, int32_t location) {
    TVLM_REQUIRE(ptr && bytes >= 0, "bad argument");
    if (bytes == 0) {
        *ptr = nullptr;
        return TVLM_OK;
    }
    if (location == 0) {
        TVLM_HIP(hipMalloc(ptr, (size_t)bytes));
    } else {
        TVLM_HIP(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
    }
    return TVLM_OK;
}

int tvlm_free(void* ptr, int32_t location) {
    if (!ptr) return TVLM_OK;
    if (location == 0) {
        TVLM_HIP(hipFree(ptr));
    } else {
        TVLM_HIP(hipHostFree(ptr));
    }
    return TVLM_OK;
}

int tvlm_copy(int32_t dst_loc, void* dst, int64_t dst_stride, int32_t src_loc, const void* src, int64_t src_stride,
              int64_t elem_size, int64_t n, void* stream) {
    TVLM_REQUIRE(dst && src && elem_size > 0 && n >= 0, "bad argument");
    if (n == 0) return TVLM_OK;
    hipStream_t st = as_stream(stream);
    const hipMemcpyKind kind = hipMemcpyDefault;
    (void)dst_loc;
    (void)src_loc;
    if (dst_stride == 1 && src_stride == 1) {
        TVLM_HIP(hipMemcpyAsync(dst, src, (size_t)(n * elem_size), kind, st));
    } else {
        TVLM_HIP(hipMemcpy2DAsync(dst, (size_t)(dst_stride * elem_size), src, (size_t)(src_stride * elem_size),
                                  (size_t)elem_size, (size_t)n, kind, st));
    }
    return TVLM_OK;
}

int tvlm_fill_f32(int32_t location, float* ptr, int64_t stride, float value, int64_t n, void* stream) {
    TVLM_REQUIRE(ptr || n == 0, "null pointer");
    if (n == 0) return TVLM_OK;
    if (location != 0) {
        TVLM_HIP(hipStreamSynchronize(as_stream(stream)));
        for (int64_t i = 0; i < n; i++) ptr[i * stride] = value;
        return TVLM_OK;
    }
    hipLaunchKernelGGL(k_fill<float>, blocks_for(n), dim3(256), 0, as_stream(stream), ptr, stride, value, n);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

int tvlm_fill_f64(int32_t location, double* ptr, int64_t stride, double value, int64_t n, void* stream) {
    TVLM_REQUIRE(ptr || n == 0, "null pointer");
    if (n == 0) return TVLM_OK;
    if (location != 0) {
        TVLM_HIP(hipStreamSynchronize(as_stream(stream)));
        for (int64_t i = 0; i < n; i++) ptr[i * stride] = value;
        return TVLM_OK;
    }
    hipLaunchKernelGGL(k_fill<double>, blocks_for(n), dim3(256), 0, as_stream(stream), ptr, stride, value, n);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

#define TVLM_DEFINE_MESH(T, S)                                                                               \
    int tvlm_mesh_metrics_##S(T alpha, int32_t n_surf, const tvlm_view* v, const tvlm_view* c, const tvlm_view* n, \
                              const tvlm_view* a, void* st) {                                                 \
        return mesh_metrics<T>(alpha, n_surf, v, c, n, a, st);                                               \
    }                                                                                                         \
    int tvlm_rhs_assemble_velocities_##S(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* n,            \
                                         const tvlm_view* v, void* st) {                                      \
        return rhs_velocities<T>(rhs, n_surf, n, v, st);                                                      \
    }                                                                                                         \
    int tvlm_forces_steady_##S(const tvlm_view* v, const tvlm_view* gd, const tvlm_view* vel, const tvlm_view* f, \
                               void* st) {                                                                    \
        return forces<T, false>(v, gd, nullptr, vel, nullptr, nullptr, f, st);                                \
    }                                                                                                         \
    int tvlm_forces_unsteady_##S(const tvlm_view* v, const tvlm_view* gd, const tvlm_view* gdt, const tvlm_view* vel, \
                                 const tvlm_view* a, const tvlm_view* n, const tvlm_view* f, void* st) {      \
        return forces<T, true>(v, gd, gdt, vel, a, n, f, st);                                                 \
    }                                                                                                         \
    int tvlm_coeff_cl_##S(const tvlm_view* f, const T* la, const T* fs, T rho, T area, T* out, void* st) {    \
        TVLM_REQUIRE(f && la && fs && out, "null argument");                                                  \
        TVLM_REQUIRE(plane_contiguous(*f), "forces stride");                                                  \
        const T den = (T)0.5 * rho * (fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]) * area;                  \
        hipLaunchKernelGGL(k_coeff_cl<T>, dim3(1), dim3(kRedThreads), 0, as_stream(st), (const T*)f->ptr,     \
                           f->stride[2], (int)(f->shape[0] * f->shape[1]), la[0], la[1], la[2], den, out);    \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_coeff_cm_##S(const tvlm_view* f, const tvlm_view* v, const T* ref, const T* fs, T rho, T area,   \
                          T mac, T* out, void* st) {                                                          \
        TVLM_REQUIRE(f && v && ref && fs && out, "null argument");                                            \
        TVLM_REQUIRE(plane_contiguous(*f) && plane_contiguous(*v) && v->shape[0] == f->shape[0] + 1,          \
                     "forces/verts shape");                                                                   \
        const T den = (T)0.5 * rho * (fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]) * area * mac;            \
        hipLaunchKernelGGL(k_coeff_cm<T>, dim3(1), dim3(kRedThreads), 0, as_stream(st), (const T*)f->ptr,     \
                           f->stride[2], (const T*)v->ptr, v->stride[2], (int)f->shape[0], (int)f->shape[1], \
                           ref[0], ref[1], ref[2], den, out);                                                 \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_mesh_mac_##S(const tvlm_view* v, const tvlm_view* a, T* out, void* st) {                         \
        TVLM_REQUIRE(v && a && out, "null argument");                                                         \
        TVLM_REQUIRE(plane_contiguous(*v) && a->stride[0] == 1 && v->shape[0] == a->shape[0] + 1 &&           \
                         v->shape[1] == a->shape[1] + 1,                                                      \
                     "verts/areas shape");                                                                    \
        hipLaunchKernelGGL(k_mesh_mac<T>, dim3(1), dim3(kRedThreads), 0, as_stream(st), (const T*)v->ptr,     \
                           v->stride[2], (int)a->shape[0], (int)a->shape[1], (const T*)a->ptr, a->stride[1], out); \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_sum_##S(const tvlm_view* x, int32_t ndim, T* out, void* st) {                                    \
        TVLM_REQUIRE(x && out && (ndim == 1 || ndim == 2), "bad argument");                                  \
        hipLaunchKernelGGL(k_sum2d<T>, dim3(1), dim3(kRedThreads), 0, as_stream(st), (const T*)x->ptr,        \
                           x->stride[0], ndim == 2 ? x->stride[1] : 0, x->shape[0], ndim == 2 ? x->shape[1] : 1, \
                           out, 0);                                                                           \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_nrm2_##S(const tvlm_view* x, T* out, void* st) {                                                 \
        TVLM_REQUIRE(x && out, "bad argument");                                                               \
        hipLaunchKernelGGL(k_sum2d<T>, dim3(1), dim3(kRedThreads), 0, as_stream(st), (const T*)x->ptr,        \
                           x->stride[0], (int64_t)0, x->shape[0], (int64_t)1, out, 1);                        \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_wake_shed_##S(int32_t n_surf, const tvlm_view* vw, const tvlm_view* wk, int32_t it, void* st) {  \
        TVLM_REQUIRE(vw && wk && n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "bad argument");                \
        for (int m = 0; m < n_surf; m++) {                                                                    \
            const int ns = (int)vw[m].shape[0] - 1, nc = (int)vw[m].shape[1] - 1;                             \
            const int row = (int)wk[m].shape[1] - 1 - it;                                                     \
            TVLM_REQUIRE(plane_contiguous(vw[m]) && plane_contiguous(wk[m]) && wk[m].shape[0] == ns + 1,      \
                         "verts_wing/verts_wake[%d] shape", m);                                              \
            TVLM_REQUIRE(row >= 0 && row < wk[m].shape[1], "iteration %d outside the wake (%lld rows)", it,   \
                         (long long)wk[m].shape[1]);                                                          \
            hipLaunchKernelGGL(k_wake_shed<T>, blocks_for(4 * (ns + 1)), dim3(256), 0, as_stream(st),         \
                               (const T*)vw[m].ptr, vw[m].stride[2], ns, nc, (T*)wk[m].ptr, wk[m].stride[2], row); \
            TVLM_LAUNCHED();                                                                                  \
        }                                                                                                     \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_displace_wing_##S(int32_t n_surf, const tvlm_view* tr, const tvlm_view* vw, const tvlm_view* vi, \
                               void* st) {                                                                    \
        TVLM_REQUIRE(tr && vw && vi && n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "bad argument");          \
        for (int m = 0; m < n_surf; m++) {                                                                    \
            TVLM_REQUIRE(plane_contiguous(vw[m]) && plane_contiguous(vi[m]) && vw[m].shape[0] == vi[m].shape[0] && \
                             vw[m].shape[1] == vi[m].shape[1] && vw[m].shape[2] == 4 && vi[m].shape[2] == 4,  \
                         "verts[%d] shape", m);                                                               \
            TVLM_REQUIRE(tr[m].shape[0] == 4 && tr[m].shape[1] == 4, "transform[%d] must be 4x4", m);         \
            const int64_t nv = vw[m].shape[0] * vw[m].shape[1];                                               \
            hipLaunchKernelGGL(k_displace<T>, blocks_for(nv), dim3(256), 0, as_stream(st), (const T*)tr[m].ptr, \
                               tr[m].stride[0], tr[m].stride[1], (const T*)vi[m].ptr, vi[m].stride[2],        \
                               (T*)vw[m].ptr, vw[m].stride[2], nv);                                           \
            TVLM_LAUNCHED();                                                                                  \
        }                                                                                                     \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_wake_geometry_##S(int32_t n_surf, const tvlm_view* tr, const tvlm_view* vi, const tvlm_view* wk,  \
                               int32_t steps, void* st) {                                                     \
        TVLM_REQUIRE(tr && vi && wk && n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES && steps >= 0, "bad argument"); \
        TVLM_REQUIRE(tr->ptr && tr->shape[0] == 4 && tr->shape[1] == 4 &&                                     \
                         tr->shape[2] >= (int64_t)steps * n_surf,                                             \
                     "transforms must be {4, 4, >= steps * n_surf}");                                         \
        for (int m = 0; m < n_surf; m++) {                                                                    \
            TVLM_REQUIRE(plane_contiguous(vi[m]) && plane_contiguous(wk[m]) && vi[m].shape[2] == 4 &&         \
                             wk[m].shape[2] == 4 && wk[m].shape[0] == vi[m].shape[0] && vi[m].shape[1] >= 2,  \
                         "verts_wing_init/verts_wake[%d] shape", m);                                          \
            TVLM_REQUIRE(steps <= wk[m].shape[1], "%d steps outside the wake (%lld rows)", steps,             \
                         (long long)wk[m].shape[1]);                                                          \
        }                                                                                                     \
        for (int m = 0; m < n_surf; m++) {                                                                    \
            const int ns = (int)vi[m].shape[0] - 1, nc = (int)vi[m].shape[1] - 1;                             \
            const int64_t cnt = (int64_t)steps * (ns + 1);                                                    \
            if (cnt == 0) continue;                                                                           \
            hipLaunchKernelGGL(k_wake_geometry<T>, blocks_for(cnt), dim3(256), 0, as_stream(st),              \
                               (const T*)tr->ptr, tr->stride[0], tr->stride[1], tr->stride[2], n_surf, m,     \
                               (const T*)vi[m].ptr, vi[m].stride[2], ns, nc, (T*)wk[m].ptr, wk[m].stride[2],  \
                               (int)wk[m].shape[1], steps);                                                   \
            TVLM_LAUNCHED();                                                                                  \
        }                                                                                                     \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_gamma_wake_from_coeffs_##S(const tvlm_view* gw, const tvlm_view* gc, int32_t harmonics, T tn,    \
                                        T omega, T dt, int64_t iteration, void* st) {                        \
        TVLM_REQUIRE(gw && gc && harmonics >= 0, "bad argument");                                            \
        TVLM_REQUIRE(gw->stride[0] == 1 && gc->stride[0] == 1 && gc->shape[0] == gw->shape[0] &&             \
                         gc->shape[1] >= 2 * harmonics + 1,                                                   \
                     "gamma_wake / gamma_coeffs shape");                                                      \
        TVLM_REQUIRE(iteration >= 0 && iteration <= gw->shape[1], "iteration out of range");                 \
        const int32_t K = 2 * harmonics + 1;                                                                  \
        const T s1 = (T)1 / sqrt((T)K), s2 = sqrt((T)2 / (T)K);                                              \
        const int64_t start = gw->shape[1] - iteration;                                                       \
        const int64_t cnt = iteration * gw->shape[0];                                                         \
        if (cnt == 0) return TVLM_OK;                                                                         \
        hipLaunchKernelGGL(k_gamma_wake_coeffs<T>, blocks_for(cnt), dim3(256), 0, as_stream(st), (T*)gw->ptr, \
                           gw->stride[1], (const T*)gc->ptr, gc->stride[1], (int)gw->shape[0], start,         \
                           gw->shape[1], harmonics, tn, omega, dt, s1, s2);                                   \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_axpy_##S(T alpha, const tvlm_view* x, const tvlm_view* y, int32_t ndim, void* st) {              \
        TVLM_REQUIRE(x && y && (ndim == 1 || ndim == 2), "bad argument");                                    \
        const int64_t n0 = x->shape[0], n1 = ndim == 2 ? x->shape[1] : 1;                                     \
        TVLM_REQUIRE(y->shape[0] == n0 && (ndim == 1 || y->shape[1] == n1), "axpy shape mismatch");           \
        if (n0 * n1 == 0) return TVLM_OK;                                                                     \
        hipLaunchKernelGGL(k_axpy2d<T>, blocks_for(n0 * n1), dim3(256), 0, as_stream(st), alpha,              \
                           (const T*)x->ptr, x->stride[0], ndim == 2 ? x->stride[1] : 0, (T*)y->ptr,          \
                           y->stride[0], ndim == 2 ? y->stride[1] : 0, n0, n1);                               \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_scal_##S(T alpha, const tvlm_view* x, void* st) {                                                \
        TVLM_REQUIRE(x, "bad argument");                                                                      \
        if (x->shape[0] == 0) return TVLM_OK;                                                                 \
        hipLaunchKernelGGL(k_scal<T>, blocks_for(x->shape[0]), dim3(256), 0, as_stream(st), alpha, (T*)x->ptr, \
                           x->stride[0], x->shape[0]);                                                        \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_gemm_##S(T alpha, const tvlm_view* a, const tvlm_view* b, T beta, const tvlm_view* c, int32_t ta, \
                      int32_t tb, void* st) {                                                                 \
        TVLM_REQUIRE(a && b && c, "bad argument");                                                            \
        TVLM_REQUIRE(a->stride[0] == 1 && b->stride[0] == 1 && c->stride[0] == 1, "gemm needs unit stride[0]"); \
        const int m = (int)(ta ? a->shape[1] : a->shape[0]);                                                  \
        const int n = (int)(tb ? b->shape[0] : b->shape[1]);                                                  \
        const int k = (int)(ta ? a->shape[0] : a->shape[1]);                                                  \
        TVLM_REQUIRE((int)(tb ? b->shape[1] : b->shape[0]) == k && c->shape[0] == m && c->shape[1] == n,      \
                     "gemm shape mismatch");                                                                  \
        if (m == 0 || n == 0) return TVLM_OK;                                                                 \
        hipLaunchKernelGGL(k_gemm<T>, dim3((m + 15) / 16, (n + 15) / 16), dim3(256), 0, as_stream(st), m, n, k, \
                           alpha, (const T*)a->ptr, a->stride[1], ta, (const T*)b->ptr, b->stride[1], tb, beta, \
                           (T*)c->ptr, c->stride[1]);                                                         \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_kinematic_velocities_##S(int32_t n_surf, const T* dT, const tvlm_view* c, const tvlm_view* v,    \
                                      void* st) {                                                             \
        TVLM_REQUIRE(dT && c && v && n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "bad argument");            \
        for (int m = 0; m < n_surf; m++) {                                                                    \
            TVLM_REQUIRE(plane_contiguous(c[m]) && plane_contiguous(v[m]) && c[m].shape[0] == v[m].shape[0] &&  \
                             c[m].shape[1] == v[m].shape[1],                                                  \
                         "colloc/velocities[%d] shape", m);                                                   \
            Mat4<T> d;                                                                                        \
            for (int k = 0; k < 16; k++) d.m[k] = dT[16 * m + k];                                             \
            const int np = (int)(c[m].shape[0] * c[m].shape[1]);                                              \
            hipLaunchKernelGGL(k_kin_vel<T>, blocks_for(np), dim3(256), 0, as_stream(st), d, (const T*)c[m].ptr, \
                               c[m].stride[2], (T*)v[m].ptr, v[m].stride[2], np);                             \
            TVLM_LAUNCHED();                                                                                  \
        }                                                                                                     \
        return TVLM_OK;                                                                                       \
    }                                                                                                         \
    int tvlm_gemv_##S(T alpha, const tvlm_view* a, const tvlm_view* x, T beta, const tvlm_view* y, int32_t tr, \
                      void* st) {                                                                             \
        TVLM_REQUIRE(a && x && y && a->stride[0] == 1, "bad argument");                                       \
        const int m = (int)a->shape[0], n = (int)a->shape[1];                                                 \
        TVLM_REQUIRE(x->shape[0] == (tr ? m : n) && y->shape[0] == (tr ? n : m), "gemv shape mismatch");      \
        const int out_len = tr ? n : m;                                                                       \
        if (out_len == 0) return TVLM_OK;                                                                     \
        hipLaunchKernelGGL(k_gemv<T>, blocks_for(out_len), dim3(256), 0, as_stream(st), m, n, alpha,          \
                           (const T*)a->ptr, a->stride[1], tr, (const T*)x->ptr, x->stride[0], beta, (T*)y->ptr, \
                           y->stride[0]);                                                                     \
        TVLM_LAUNCHED();                                                                                      \
        return TVLM_OK;                                                                                       \
    }

TVLM_DEFINE_MESH(double, f64)
TVLM_DEFINE_MESH(float, f32)

}  // extern "C"

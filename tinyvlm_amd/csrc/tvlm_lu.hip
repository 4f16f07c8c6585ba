// tvlm_lu.hip -- dense LU with partial pivoting (getrf) and solve (getrs) on gfx950.
//
// Semantics follow LAPACK ?getrf/?getrs as called by the reference (vlm_backend_cpu_lapack.cpp:40-94):
// column-major in place L\U, unit lower L, 1-based ipiv, pivot = first row of max |a| in the
// column, info = first exactly-zero pivot (1-based), factorisation continues past it.
//
// getrf is blocked right-looking with panel width NB:
//   1. k_panel    : one work-group factors the n-k0 x kb panel (pivot search by LDS tree,
//                   row swap inside the panel, reciprocal scaling, rank-1 update);
//   2. k_laswp    : the panel's interchanges applied to the columns left and right of it;
//   3. k_trsm     : U12 = L11^{-1} A12 with L11 staged in LDS, one thread per column;
//   4. k_update   : A22 -= L21 U12, 64x64 output tiles through LDS, 4x4 register blocking.
// getrs applies the interchanges, then forward/backward substitution: one work-group per
// right-hand side for n <= kSmallSolve, otherwise blocked (diagonal block in one work-group,
// off-diagonal update by all work-groups).
#include "tvlm_internal.hpp"

namespace tvlm {
namespace {

constexpr int kPanelThreads = 1024;
constexpr int NB = 64;

template <typename T>
__global__ void k_zero_info(int32_t* info) {
    if (threadIdx.x == 0 && blockIdx.x == 0) info[0] = 0;
}

template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_panel(T* __restrict__ a, int64_t lda, int n, int k0, int kb,
                                                         int32_t* __restrict__ ipiv, int32_t* __restrict__ info) {
    __shared__ T sval[kPanelThreads];
    __shared__ int sidx[kPanelThreads];
    __shared__ T srow[NB];
    const int t = threadIdx.x;
    for (int k = k0; k < k0 + kb; k++) {
        T* colk = a + (int64_t)k * lda;
        // pivot search: max |a(i,k)|, i >= k, first index on ties
        T best = T(-1);
        int bi = k;
        for (int i = k + t; i < n; i += kPanelThreads) {
            const T v = fabs(colk[i]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        sval[t] = best;
        sidx[t] = bi;
        __syncthreads();
        for (int s = kPanelThreads / 2; s > 0; s >>= 1) {
            if (t < s) {
                const T v2 = sval[t + s];
                const int i2 = sidx[t + s];
                if (v2 > sval[t] || (v2 == sval[t] && i2 < sidx[t])) {
                    sval[t] = v2;
                    sidx[t] = i2;
                }
            }
            __syncthreads();
        }
        const int p = sidx[0];
        if (t == 0) ipiv[k] = p + 1;
        // swap rows k and p inside the panel columns, and stage the pivot row
        if (t < kb) {
            const int j = k0 + t;
            T* cj = a + (int64_t)j * lda;
            const T vk = cj[k], vp = cj[p];
            if (p != k) {
                cj[k] = vp;
                cj[p] = vk;
            }
            srow[t] = (p != k) ? vp : vk;
        }
        __syncthreads();
        const T piv = srow[k - k0];
        if (t == 0 && piv == T(0) && info[0] == 0) info[0] = k + 1;
        if (piv != T(0)) {
            const T r = T(1) / piv;
            for (int i = k + 1 + t; i < n; i += kPanelThreads) {
                const T l = colk[i] * r;
                colk[i] = l;
                for (int j = k + 1; j < k0 + kb; j++) {
                    T* cj = a + (int64_t)j * lda;
                    cj[i] -= l * srow[j - k0];
                }
            }
        }
        __syncthreads();
    }
}

// Apply interchanges ipiv[k0 .. k0+kb) to columns [c0, c1) (one thread per column).
template <typename T>
__global__ void k_laswp(T* __restrict__ a, int64_t lda, int c0, int c1, int k0, int kb, const int32_t* __restrict__ ipiv) {
    const int j = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= c1) return;
    T* cj = a + (int64_t)j * lda;
    for (int k = k0; k < k0 + kb; k++) {
        const int p = ipiv[k] - 1;
        if (p != k) {
            const T t = cj[k];
            cj[k] = cj[p];
            cj[p] = t;
        }
    }
}

// U12 = L11^{-1} A12 for columns [j0, n), L11 = unit lower of a(k0:k0+kb, k0:k0+kb).
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(T* __restrict__ a, int64_t lda, int n, int k0, int kb, int j0) {
    __shared__ T L[NB][NB + 1];
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
        const int r = e % kb, c = e / kb;
        L[r][c] = a[(k0 + r) + (int64_t)(k0 + c) * lda];
    }
    __syncthreads();
    const int j = j0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    T* cj = a + (int64_t)j * lda + k0;
    T x[NB];
#pragma unroll
    for (int r = 0; r < NB; r++)
        if (r < kb) x[r] = cj[r];
#pragma unroll
    for (int c = 0; c < NB; c++) {
        if (c < kb) {
            const T xc = x[c];
#pragma unroll
            for (int r = c + 1; r < NB; r++)
                if (r < kb) x[r] -= L[r][c] * xc;
        }
    }
#pragma unroll
    for (int r = 0; r < NB; r++)
        if (r < kb) cj[r] = x[r];
}

// C(m x n) -= A(m x k) B(k x n), all column-major; 64x64 tiles, 256 threads, 4x4 per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_update(T* __restrict__ C, int64_t ldc, const T* __restrict__ A, int64_t lda,
                                                const T* __restrict__ B, int64_t ldb, int m, int n, int k) {
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ T sa[BK][BM + 1];
    __shared__ T sb[BK][BN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.x * BM, col0 = blockIdx.y * BN;
    T acc[4][4] = {};
    for (int kk = 0; kk < k; kk += BK) {
        for (int e = threadIdx.x; e < BM * BK; e += 256) {
            const int r = e % BM, q = e / BM;
            const int gr = row0 + r, gq = kk + q;
            sa[q][r] = (gr < m && gq < k) ? A[gr + (int64_t)gq * lda] : T(0);
        }
        for (int e = threadIdx.x; e < BN * BK; e += 256) {
            const int q = e % BK, c = e / BK;
            const int gq = kk + q, gc = col0 + c;
            sb[q][c] = (gq < k && gc < n) ? B[gq + (int64_t)gc * ldb] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; q++) {
            T ra[4], rb[4];
#pragma unroll
            for (int u = 0; u < 4; u++) ra[u] = sa[q][tx + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; v++) rb[v] = sb[q][ty + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += ra[u] * rb[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int gr = row0 + tx + 16 * u;
        if (gr >= m) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int gc = col0 + ty + 16 * v;
            if (gc < n) C[gr + (int64_t)gc * ldc] -= acc[u][v];
        }
    }
}

// ---- getrs ------------------------------------------------------------------------------------
constexpr int kSmallSolve = 2048;
constexpr int kSolveThreads = 1024;

// Whole solve of one right-hand side column per work-group (n <= kSmallSolve).
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_solve_small(const T* __restrict__ a, int64_t lda, int n,
                                                               const int32_t* __restrict__ ipiv, T* __restrict__ x,
                                                               int64_t ldx) {
    __shared__ T xs[kSmallSolve];
    T* xc = x + (int64_t)blockIdx.x * ldx;
    const int t = threadIdx.x;
    for (int i = t; i < n; i += kSolveThreads) xs[i] = xc[i];
    __syncthreads();
    if (t == 0) {
        for (int k = 0; k < n; k++) {
            const int p = ipiv[k] - 1;
            if (p != k) {
                const T tmp = xs[k];
                xs[k] = xs[p];
                xs[p] = tmp;
            }
        }
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {  // L y = P b (unit diagonal)
        const T xk = xs[k];
        const T* col = a + (int64_t)k * lda;
        for (int i = k + 1 + t; i < n; i += kSolveThreads) xs[i] -= col[i] * xk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {  // U x = y
        const T* col = a + (int64_t)k * lda;
        if (t == 0) xs[k] = xs[k] / col[k];
        __syncthreads();
        const T xk = xs[k];
        for (int i = t; i < k; i += kSolveThreads) xs[i] -= col[i] * xk;
        __syncthreads();
    }
    for (int i = t; i < n; i += kSolveThreads) xc[i] = xs[i];
}

template <typename T>
__global__ void k_apply_pivots(T* __restrict__ x, int64_t ldx, int n, const int32_t* __restrict__ ipiv) {
    T* xc = x + (int64_t)blockIdx.x * ldx;
    if (threadIdx.x != 0) return;
    for (int k = 0; k < n; k++) {
        const int p = ipiv[k] - 1;
        if (p != k) {
            const T tmp = xc[k];
            xc[k] = xc[p];
            xc[p] = tmp;
        }
    }
}

constexpr int SB = 256;  // diagonal block of the blocked solve

// Diagonal block solve: lower (unit) if LOWER, else upper; rows [r0, r0+nb); one WG per rhs.
template <typename T, bool LOWER>
__global__ __launch_bounds__(SB) void k_diag_solve(const T* __restrict__ a, int64_t lda, int r0, int nb,
                                                   T* __restrict__ x, int64_t ldx) {
    __shared__ T xs[SB];
    T* xc = x + (int64_t)blockIdx.x * ldx + r0;
    const int t = threadIdx.x;
    if (t < nb) xs[t] = xc[t];
    __syncthreads();
    if (LOWER) {
        for (int k = 0; k < nb; k++) {
            const T xk = xs[k];
            if (t > k && t < nb) xs[t] -= a[(r0 + t) + (int64_t)(r0 + k) * lda] * xk;
            __syncthreads();
        }
    } else {
        for (int k = nb - 1; k >= 0; k--) {
            if (t == 0) xs[k] = xs[k] / a[(r0 + k) + (int64_t)(r0 + k) * lda];
            __syncthreads();
            const T xk = xs[k];
            if (t < k) xs[t] -= a[(r0 + t) + (int64_t)(r0 + k) * lda] * xk;
            __syncthreads();
        }
    }
    if (t < nb) xc[t] = xs[t];
}

// x(rows [q0, q1)) -= A(rows [q0,q1), cols [c0, c0+nb)) x(c0 : c0+nb); one thread per row, per rhs.
template <typename T>
__global__ void k_off_update(const T* __restrict__ a, int64_t lda, int q0, int q1, int c0, int nb, T* __restrict__ x,
                             int64_t ldx) {
    __shared__ T xb[SB];
    T* xc = x + (int64_t)blockIdx.y * ldx;
    for (int e = threadIdx.x; e < nb; e += blockDim.x) xb[e] = xc[c0 + e];
    __syncthreads();
    const int i = q0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q1) return;
    T s = T(0);
    for (int c = 0; c < nb; c++) s += a[i + (int64_t)(c0 + c) * lda] * xb[c];
    xc[i] -= s;
}

template <typename T>
int getrf(const tvlm_view* av, int32_t* ipiv, int32_t* info, void* stream) {
    TVLM_REQUIRE(av && av->ptr && ipiv && info, "null argument");
    TVLM_REQUIRE(av->stride[0] == 1, "A must be column-major");
    TVLM_REQUIRE(av->shape[0] == av->shape[1], "A must be square (got %lld x %lld)", (long long)av->shape[0],
                 (long long)av->shape[1]);
    const int n = (int)av->shape[0];
    const int64_t lda = av->stride[1];
    TVLM_REQUIRE(lda >= n || n == 0, "lda < n");
    hipStream_t st = as_stream(stream);
    T* a = static_cast<T*>(av->ptr);
    hipLaunchKernelGGL(k_zero_info<T>, dim3(1), dim3(64), 0, st, info);
    TVLM_LAUNCHED();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = (n - k0 < NB) ? n - k0 : NB;
        hipLaunchKernelGGL(k_panel<T>, dim3(1), dim3(kPanelThreads), 0, st, a, lda, n, k0, kb, ipiv, info);
        TVLM_LAUNCHED();
        if (k0 > 0) {
            hipLaunchKernelGGL(k_laswp<T>, dim3((k0 + 255) / 256), dim3(256), 0, st, a, lda, 0, k0, k0, kb, ipiv);
            TVLM_LAUNCHED();
        }
        const int j1 = k0 + kb;
        if (j1 >= n) continue;
        hipLaunchKernelGGL(k_laswp<T>, dim3((n - j1 + 255) / 256), dim3(256), 0, st, a, lda, j1, n, k0, kb, ipiv);
        TVLM_LAUNCHED();
        hipLaunchKernelGGL(k_trsm<T>, dim3((n - j1 + 255) / 256), dim3(256), 0, st, a, lda, n, k0, kb, j1);
        TVLM_LAUNCHED();
        const int m = n - j1;
        hipLaunchKernelGGL(k_update<T>, dim3((m + 63) / 64, (m + 63) / 64), dim3(256), 0, st,
                           a + j1 + (int64_t)j1 * lda, lda, a + j1 + (int64_t)k0 * lda, lda,
                           a + k0 + (int64_t)j1 * lda, lda, m, m, kb);
        TVLM_LAUNCHED();
    }
    return TVLM_OK;
}

template <typename T>
int getrs(const tvlm_view* av, const int32_t* ipiv, const tvlm_view* xv, void* stream) {
    TVLM_REQUIRE(av && xv && av->ptr && xv->ptr && ipiv, "null argument");
    TVLM_REQUIRE(av->stride[0] == 1 && xv->stride[0] == 1, "column-major operands required");
    const int n = (int)av->shape[0];
    TVLM_REQUIRE(av->shape[1] == n && xv->shape[0] == n, "shape mismatch");
    const int nrhs = (int)xv->shape[1];
    if (n == 0 || nrhs == 0) return TVLM_OK;
    const int64_t lda = av->stride[1], ldx = nrhs > 1 ? xv->stride[1] : n;
    hipStream_t st = as_stream(stream);
    const T* a = static_cast<const T*>(av->ptr);
    T* x = static_cast<T*>(xv->ptr);
    if (n <= kSmallSolve) {
        hipLaunchKernelGGL(k_solve_small<T>, dim3(nrhs), dim3(kSolveThreads), 0, st, a, lda, n, ipiv, x, ldx);
        TVLM_LAUNCHED();
        return TVLM_OK;
    }
    hipLaunchKernelGGL(k_apply_pivots<T>, dim3(nrhs), dim3(64), 0, st, x, ldx, n, ipiv);
    TVLM_LAUNCHED();
    for (int r0 = 0; r0 < n; r0 += SB) {  // forward: L
        const int nb = (n - r0 < SB) ? n - r0 : SB;
        hipLaunchKernelGGL((k_diag_solve<T, true>), dim3(nrhs), dim3(SB), 0, st, a, lda, r0, nb, x, ldx);
        TVLM_LAUNCHED();
        if (r0 + nb < n) {
            hipLaunchKernelGGL(k_off_update<T>, dim3((n - r0 - nb + 255) / 256, nrhs), dim3(256), 0, st, a, lda,
                               r0 + nb, n, r0, nb, x, ldx);
            TVLM_LAUNCHED();
        }
    }
    const int nblk = (n + SB - 1) / SB;
    for (int b = nblk - 1; b >= 0; b--) {  // backward: U
        const int r0 = b * SB;
        const int nb = (n - r0 < SB) ? n - r0 : SB;
        hipLaunchKernelGGL((k_diag_solve<T, false>), dim3(nrhs), dim3(SB), 0, st, a, lda, r0, nb, x, ldx);
        TVLM_LAUNCHED();
        if (r0 > 0) {
            hipLaunchKernelGGL(k_off_update<T>, dim3((r0 + 255) / 256, nrhs), dim3(256), 0, st, a, lda, 0, r0, r0, nb,
                               x, ldx);
            TVLM_LAUNCHED();
        }
    }
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

extern "C" {
int tvlm_getrf_f64(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream) {
    return tvlm::getrf<double>(a, ipiv, info, stream);
}
int tvlm_getrf_f32(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream) {
    return tvlm::getrf<float>(a, ipiv, info, stream);
}
int tvlm_getrs_f64(const tvlm_view* a, const int32_t* ipiv, const tvlm_view* x, void* stream) {
    return tvlm::getrs<double>(a, ipiv, x, stream);
}
int tvlm_getrs_f32(const tvlm_view* a, const int32_t* ipiv, const tvlm_view* x, void* stream) {
    return tvlm::getrs<float>(a, ipiv, x, stream);
}
}

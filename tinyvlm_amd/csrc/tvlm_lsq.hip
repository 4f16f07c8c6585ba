// tvlm_lsq.hip -- linear least squares min ||A x - B|| (LSQ::solve) on gfx950.
//
// Semantics of LAPACK ?gels with trans 'N' and M >= N as the reference calls it
// (vlm_backend_cpu_lapack.cpp:141-205): A {M,N} is overwritten by its Householder QR factors in the
// ?geqrf layout (R on and above the diagonal, the reflectors' essential parts below it), B {M,NRHS}
// by Q^T B whose first N rows are then overwritten by the solution of R X = (Q^T B)(0:N).
// info = i (1-based) if R(i,i) is exactly zero (A rank deficient), in which case the back
// substitution is skipped as in ?gels; an all-zero A gives B = 0.  LAPACK's rescaling of A and B
// with norms outside [smlnum, bignum] is not performed.
//
// Reflector j (?larfg): alpha = A(j,j), s = sum_{i>j} A(i,j)^2; s = 0 -> tau = 0 (H = I); otherwise
// beta = -sign(alpha) sqrt(alpha^2 + s), tau = (beta - alpha) / beta, v = (1, A(j+1:,j) / (alpha - beta)),
// A(j,j) = beta.  H = I - tau v v^T is applied to A(j:, j+1:) and to B(j:, :) (?larf, then the
// unblocked ?orm2r order H(1) first), then X solves R X = B(0:N) by column-oriented back substitution
// (?trsm Left/Upper/NoTrans).
//
// One work-group of 1024 threads does the whole factorisation: the problems this serves are
// tall and thin (Anderson histories of <= 10 columns over N x K rows).  Each column step is a
// strided pass over the remaining columns with per-thread partial dot products held in registers
// (statically indexed, N + NRHS <= 32), one fixed-order block reduction, and one rank-1 update.
#include "tvlm_internal.hpp"

namespace tvlm {
namespace {

constexpr int kLsqThreads = 1024;
constexpr int kLsqWaves = kLsqThreads / 64;
constexpr int kLsqMaxCols = 32;  // N + NRHS

template <typename T>
__device__ __forceinline__ T wave_sum_xor(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max_xor(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// target column k of the step at column j: the A columns right of j, then the B columns
template <typename T>
__device__ __forceinline__ T* target(T* a, int64_t lda, T* b, int64_t ldb, int j, int na, int k) {
    return k < na ? a + (int64_t)(j + 1 + k) * lda : b + (int64_t)(k - na) * ldb;
}

template <typename T>
__global__ __launch_bounds__(kLsqThreads) void k_gels(T* __restrict__ a, int64_t lda, int m, int n, T* __restrict__ b,
                                                      int64_t ldb, int nrhs, int32_t* __restrict__ info) {
    __shared__ T red[kLsqWaves][kLsqMaxCols];
    __shared__ T wsh[kLsqMaxCols];
    __shared__ T s_tau, s_beta, s_scal, s_amax;
    __shared__ int s_info;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;

    // all-zero A: B = 0
    T amax = 0;
    for (int j = 0; j < n; j++)
        for (int i = t; i < m; i += kLsqThreads) amax = fmax(amax, fabs(a[i + (int64_t)j * lda]));
    amax = wave_max_xor(amax);
    if (lane == 0) red[wv][0] = amax;
    __syncthreads();
    if (t == 0) {
        T x = 0;
        for (int w = 0; w < kLsqWaves; w++) x = fmax(x, red[w][0]);
        s_amax = x;
    }
    __syncthreads();
    if (s_amax == T(0)) {
        for (int r = 0; r < nrhs; r++)
            for (int i = t; i < m; i += kLsqThreads) b[i + (int64_t)r * ldb] = 0;
        if (t == 0) info[0] = 0;
        return;
    }

    for (int j = 0; j < n; j++) {
        T* cj = a + (int64_t)j * lda;
        // reflector: sum of squares below the diagonal
        T s = 0;
        for (int i = j + 1 + t; i < m; i += kLsqThreads) s += cj[i] * cj[i];
        s = wave_sum_xor(s);
        __syncthreads();  // red[] of the previous step is consumed
        if (lane == 0) red[wv][0] = s;
        __syncthreads();
        if (t == 0) {
            T tot = 0;
            for (int w = 0; w < kLsqWaves; w++) tot += red[w][0];
            const T alpha = cj[j];
            if (tot == T(0)) {
                s_tau = 0;
                s_beta = alpha;
                s_scal = 1;
            } else {
                const T mag = sqrt(alpha * alpha + tot);
                const T beta = alpha >= T(0) ? -mag : mag;
                s_tau = (beta - alpha) / beta;
                s_beta = beta;
                s_scal = T(1) / (alpha - beta);
            }
        }
        __syncthreads();
        const T tau = s_tau;
        if (tau != T(0)) {
            const T scal = s_scal;
            for (int i = j + 1 + t; i < m; i += kLsqThreads) cj[i] *= scal;
            const int na = n - j - 1;
            const int K = na + nrhs;
            // w_k = C(j,k) + sum_{i>j} v_i C(i,k) for every target column k
            T acc[kLsqMaxCols];
#pragma unroll
            for (int k = 0; k < kLsqMaxCols; k++) acc[k] = 0;
            for (int i = j + 1 + t; i < m; i += kLsqThreads) {
                const T vi = cj[i];
#pragma unroll
                for (int k = 0; k < kLsqMaxCols; k++)
                    if (k < K) acc[k] += vi * target(a, lda, b, ldb, j, na, k)[i];
            }
#pragma unroll
            for (int k = 0; k < kLsqMaxCols; k++) {
                if (k < K) {
                    const T v = wave_sum_xor(acc[k]);
                    if (lane == 0) red[wv][k] = v;
                }
            }
            __syncthreads();
            if (t < K) {
                T w = target(a, lda, b, ldb, j, na, t)[j];
                for (int q = 0; q < kLsqWaves; q++) w += red[q][t];
                wsh[t] = tau * w;
            }
            __syncthreads();
            for (int i = j + t; i < m; i += kLsqThreads) {
                const T vi = (i == j) ? T(1) : cj[i];
#pragma unroll
                for (int k = 0; k < kLsqMaxCols; k++)
                    if (k < K) target(a, lda, b, ldb, j, na, k)[i] -= wsh[k] * vi;
            }
        }
        __syncthreads();
        if (t == 0) cj[j] = s_beta;
        __syncthreads();
    }

    // R X = (Q^T B)(0:N), skipped when R has an exactly zero diagonal entry
    if (t == 0) {
        int inf = 0;
        for (int i = 0; i < n; i++)
            if (a[i + (int64_t)i * lda] == T(0)) {
                inf = i + 1;
                break;
            }
        s_info = inf;
        info[0] = inf;
    }
    __syncthreads();
    if (s_info == 0 && t < nrhs) {
        T* x = b + (int64_t)t * ldb;
        for (int jj = n - 1; jj >= 0; jj--) {
            const T* col = a + (int64_t)jj * lda;
            const T xj = x[jj] / col[jj];
            x[jj] = xj;
            for (int i = 0; i < jj; i++) x[i] -= xj * col[i];
        }
    }
}

template <typename T>
int gels(const tvlm_view* av, const tvlm_view* bv, int32_t* info, void* stream) {
    TVLM_REQUIRE(av && bv && av->ptr && bv->ptr && info, "gels: null argument");
    TVLM_REQUIRE(av->stride[0] == 1 && bv->stride[0] == 1, "gels: column-major operands required");
    TVLM_REQUIRE(av->shape[2] == 1 && bv->shape[2] == 1, "gels: 2-D operands required");
    const int64_t m = av->shape[0], n = av->shape[1], nrhs = bv->shape[1];
    TVLM_REQUIRE(bv->shape[0] == m, "gels: B has %lld rows, A has %lld", (long long)bv->shape[0], (long long)m);
    TVLM_REQUIRE(m >= 0 && n >= 0 && nrhs >= 0 && m < (int64_t)1 << 31, "gels: bad sizes");
    if (m < n) {
        set_error("gels: underdetermined systems (M < N) are not supported");
        return TVLM_ERR_UNSUPPORTED;
    }
    if (n + nrhs > kLsqMaxCols) {
        set_error("gels: N + NRHS = %lld exceeds %d", (long long)(n + nrhs), kLsqMaxCols);
        return TVLM_ERR_UNSUPPORTED;
    }
    const int64_t lda = n > 1 ? av->stride[1] : m, ldb = nrhs > 1 ? bv->stride[1] : m;
    TVLM_REQUIRE(lda >= m && ldb >= m, "gels: leading dimensions must be >= M");
    hipLaunchKernelGGL(k_gels<T>, dim3(1), dim3(kLsqThreads), 0, as_stream(stream), static_cast<T*>(av->ptr), lda,
                       (int)m, (int)n, static_cast<T*>(bv->ptr), ldb, (int)nrhs, info);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

extern "C" {
int tvlm_gels_f64(const tvlm_view* a, const tvlm_view* b, int32_t* info, void* stream) {
    return tvlm::gels<double>(a, b, info, stream);
}
int tvlm_gels_f32(const tvlm_view* a, const tvlm_view* b, int32_t* info, void* stream) {
    return tvlm::gels<float>(a, b, info, stream);
}
}

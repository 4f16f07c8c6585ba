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
// The rows are spread over up to 256 work-groups of 256 threads (row i in group (i / 256) % groups),
// with one launch per column: launch j finishes reflector j - 1 and prepares reflector j.
//  - From the previous launch's per-group partial sums (summed in a fixed order, the same in every
//    group), every group forms reflector j - 1's tau, beta and the scale of its vector, and
//    w_k = tau (C(j-1,k) + scal sum_{i>j-1} A(i,j-1) C(i,k)) for each target column k (the columns
//    right of j - 1, then the B columns); the row j - 1 values come from a head record that the
//    group owning that row wrote in the previous launch.  It then applies H = I - tau v v^T to its
//    own rows: v_i = A(i,j-1) scal stored below the diagonal, C(i,k) -= w_k v_i, A(j-1,j-1) = beta.
//  - In the same pass over its rows, each group sums A(i,j)^2 and A(i,j) C(i,k) over its rows i > j
//    (fixed order: rows, then a butterfly in each wave, then the waves in order) into its partial
//    record; the group owning row j writes that row's values into the head record.
// Partial and head records alternate between two buffers, so a launch reads one while it writes the
// other.  Launch 0 also records each group's max |A|; the last launch (j = n) then zeroes B when A
// is all zero (as ?gels does) or, in the group that owns row n - 1 (whose last update it has just
// made), sets info and solves R X = (Q^T B)(0:N).
#include "tvlm_internal.hpp"

namespace tvlm {
namespace {

constexpr int kLsqThreads = 256;
constexpr int kLsqWaves = kLsqThreads / 64;
constexpr int kLsqMaxCols = 32;        // N + NRHS
constexpr int kLsqMaxGroups = 256;
constexpr int kLsqRec = kLsqMaxCols + 1;  // partial / head record: one value, then one per target

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

// target column k of the reflector of column c: the A columns right of c, then the B columns
template <typename T>
__device__ __forceinline__ T* target(T* a, int64_t lda, T* b, int64_t ldb, int c, int na, int k) {
    return k < na ? a + (int64_t)(c + 1 + k) * lda : b + (int64_t)(k - na) * ldb;
}

// ws: part[2][groups][kLsqRec], head[2][kLsqRec], amax[groups]
template <typename T>
__global__ __launch_bounds__(kLsqThreads) void k_qr_step(T* __restrict__ a, int64_t lda, int m, int n,
                                                         T* __restrict__ b, int64_t ldb, int nrhs, T* __restrict__ ws,
                                                         int groups, int j, int32_t* __restrict__ info) {
    __shared__ T sum_prev[kLsqRec];
    __shared__ T wsh[kLsqMaxCols];
    __shared__ T red[kLsqWaves][kLsqRec];
    __shared__ T s_tau, s_beta, s_scal;
    T* part = ws;
    T* head = ws + (int64_t)2 * groups * kLsqRec;
    T* amaxg = head + 2 * kLsqRec;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = blockIdx.x;
    const int64_t step = (int64_t)groups * kLsqThreads;

    // reflector of column c = j - 1
    const int c = j - 1;
    const int nac = n - c - 1, ntc = nac + nrhs;  // A targets, all targets of column c
    T tau = 0, scal = 1, beta = 0;
    if (j > 0) {
        const T* pp = part + (int64_t)(c & 1) * groups * kLsqRec;
        const T* hp = head + (c & 1) * kLsqRec;
        // column sums of the groups' records: wave w takes columns w, w + 4, ...; lane l adds groups
        // l, l + 64, l + 128, l + 192 in that order, then a butterfly (the same order in every group;
        // the loads of a wave's columns are all issued before any is added)
#pragma unroll
        for (int k = 0; k < kLsqRec / kLsqWaves + 1; k++) {
            const int col = wv + k * kLsqWaves;
            if (col <= ntc) {
                T v = 0;
#pragma unroll
                for (int r = 0; r < kLsqMaxGroups / 64; r++) {
                    const int q = lane + 64 * r;
                    if (q < groups) v += pp[(int64_t)q * kLsqRec + col];
                }
                v = wave_sum_xor(v);
                if (lane == 0) sum_prev[col] = v;
            }
        }
        __syncthreads();
        if (t == 0) {
            const T tot = sum_prev[0], alpha = hp[0];
            if (tot == T(0)) {
                s_tau = 0;
                s_beta = alpha;
                s_scal = 1;
            } else {
                const T mag = sqrt(alpha * alpha + tot);
                const T bt = alpha >= T(0) ? -mag : mag;
                s_tau = (bt - alpha) / bt;
                s_beta = bt;
                s_scal = T(1) / (alpha - bt);
            }
        }
        __syncthreads();
        tau = s_tau;
        scal = s_scal;
        beta = s_beta;
        if (t < ntc) wsh[t] = tau * (hp[1 + t] + scal * sum_prev[1 + t]);
        __syncthreads();
    }

    // this group's rows: finish reflector c, then the partial sums of column j
    const bool next = j < n;
    const int naj = n - j - 1, ntj = next ? naj + nrhs : 0;
    T acc[kLsqMaxCols];
#pragma unroll
    for (int k = 0; k < kLsqMaxCols; k++) acc[k] = 0;
    T ss = 0, amax = 0;
    for (int64_t i = (int64_t)g * kLsqThreads + t; i < m; i += step) {
        if (j == 0)
            for (int q = 0; q < n; q++) amax = fmax(amax, fabs(a[i + (int64_t)q * lda]));
        if (j > 0 && i >= c) {
            T* ac = a + (int64_t)c * lda;
            if (tau != T(0)) {
                T v = T(1);
                if (i > c) {
                    v = ac[i] * scal;
                    ac[i] = v;
                }
#pragma unroll
                for (int k = 0; k < kLsqMaxCols; k++)
                    if (k < ntc) target(a, lda, b, ldb, c, nac, k)[i] -= wsh[k] * v;
            }
            if (i == c) ac[i] = beta;
        }
        if (next && i > j) {
            const T aij = a[i + (int64_t)j * lda];
            ss += aij * aij;
#pragma unroll
            for (int k = 0; k < kLsqMaxCols; k++)
                if (k < ntj) acc[k] += aij * target(a, lda, b, ldb, j, naj, k)[i];
        }
        if (next && i == j) {
            T* hw = head + (j & 1) * kLsqRec;
            hw[0] = a[i + (int64_t)j * lda];
            for (int k = 0; k < ntj; k++) hw[1 + k] = target(a, lda, b, ldb, j, naj, k)[i];
        }
    }
    if (j == 0) {
        amax = wave_max_xor(amax);
        if (lane == 0) red[wv][0] = amax;
        __syncthreads();
        if (t == 0) {
            T x = 0;
            for (int w = 0; w < kLsqWaves; w++) x = fmax(x, red[w][0]);
            amaxg[g] = x;
        }
        __syncthreads();
    }
    if (next) {
        ss = wave_sum_xor(ss);
        if (lane == 0) red[wv][0] = ss;
#pragma unroll
        for (int k = 0; k < kLsqMaxCols; k++) {
            if (k < ntj) {
                const T v = wave_sum_xor(acc[k]);
                if (lane == 0) red[wv][1 + k] = v;
            }
        }
        __syncthreads();
        if (t <= ntj) {
            T v = 0;
            for (int w = 0; w < kLsqWaves; w++) v += red[w][t];
            part[((int64_t)(j & 1) * groups + g) * kLsqRec + t] = v;
        }
        return;
    }

    // last launch (j = n): all-zero A -> B = 0, info = 0; else back substitution
    bool zero = n == 0;
    if (!zero) {  // every wave forms the same maximum over the groups' records
        T x = 0;
#pragma unroll
        for (int r = 0; r < kLsqMaxGroups / 64; r++) {
            const int q = lane + 64 * r;
            if (q < groups) x = fmax(x, amaxg[q]);
        }
        zero = wave_max_xor(x) == T(0);
    }
    if (zero) {
        for (int64_t i = (int64_t)g * kLsqThreads + t; i < m; i += step)
            for (int r = 0; r < nrhs; r++) b[i + (int64_t)r * ldb] = 0;
        if (g == 0 && t == 0) info[0] = 0;
        return;
    }
    if (g != ((n - 1) / kLsqThreads) % groups) return;
    __syncthreads();  // row n - 1's last update (one thread of this group) is seen by the others
    __shared__ int s_info;
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
    int64_t groups = (m + kLsqThreads - 1) / kLsqThreads;
    groups = groups < 1 ? 1 : groups > kLsqMaxGroups ? kLsqMaxGroups : groups;
    hipStream_t st = as_stream(stream);
    void* ws = nullptr;
    TVLM_TRY(device_workspace(sizeof(T) * (size_t)(2 * groups * kLsqRec + 2 * kLsqRec + groups), &ws, kWsLsq, st));
    for (int j = 0; j <= n; j++) {
        hipLaunchKernelGGL(k_qr_step<T>, dim3((unsigned)groups), dim3(kLsqThreads), 0, st, static_cast<T*>(av->ptr),
                           lda, (int)m, (int)n, static_cast<T*>(bv->ptr), ldb, (int)nrhs, static_cast<T*>(ws),
                           (int)groups, j, info);
        TVLM_LAUNCHED();
    }
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

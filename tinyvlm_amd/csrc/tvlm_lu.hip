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

// ---- fast getrf: register-resident panel, fused interchange + TRSM, MFMA trailing update ------
//
// k_panel_reg: one work-group of 1024 threads factors the panel rows [k0, n) x columns [k0, k0+kb),
// kb <= NB.  Thread t owns rows k0 + t + 1024 q (q < Q).  The panel is processed left-looking in
// sub-panels of IB columns held in registers (Q x IB values per thread): a sub-panel first receives
// the updates of the panel's earlier sub-panels (blocked forward substitution with the inverses of
// their unit-lower IB x IB diagonal blocks, kept in LDS), then is factored column by column with
// partial pivoting (argmax by wave shuffles + one LDS exchange, first index on ties; the pivot row
// and row k exchange through LDS; reciprocal scaling and rank-1 update in registers).  Each
// interchange is applied at once to the panel's other columns in global memory.  At the end thread
// 0 composes the panel's interchanges into a list of (destination, source) rows for the rest of the
// matrix.
constexpr int kPT = 1024;
constexpr int kPW = kPT / 64;

template <typename T>
__device__ __forceinline__ void argmax_merge(T& best, int& bi, T v, int i) {
    if (v > best || (v == best && i < bi)) {
        best = v;
        bi = i;
    }
}

template <typename T, int Q, int IB>
__global__ __launch_bounds__(kPT) void k_panel_reg(T* __restrict__ a, int64_t lda, int n, int k0, int kb,
                                                   int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                   int32_t* __restrict__ perm) {
    __shared__ T s_val[kPW];
    __shared__ int s_idx[kPW];
    __shared__ T s_rowp[IB], s_rowc[IB];
    __shared__ T s_inv[NB / IB][IB][IB];
    __shared__ T s_blk[IB][IB];
    __shared__ T s_u[IB][IB];
    __shared__ int s_piv[NB];
    __shared__ int s_src[NB], s_okey[NB], s_oval[NB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    T x[Q][IB];

    for (int c0 = k0, s = 0; c0 < k0 + kb; c0 += IB, s++) {
        const int ib = min(IB, k0 + kb - c0);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int r = k0 + t + kPT * q;
#pragma unroll
            for (int j = 0; j < IB; j++) x[q][j] = (r < n && j < ib) ? a[r + (int64_t)(c0 + j) * lda] : T(0);
        }
        // updates from the earlier sub-panels sp < s (rows b0 .. b0+IB of block sp hold U after this)
        for (int sp = 0; sp < s; sp++) {
            const int b0 = k0 + sp * IB;
            const int jj = t - (b0 - k0);  // owners of the block's rows: q = 0
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) s_blk[jj][j] = x[0][j];
            }
            __syncthreads();
            if (t < IB * IB) {
                const int i = t / IB, j = t % IB;
                T u = T(0);
                for (int p = 0; p <= i; p++) u += s_inv[sp][i][p] * s_blk[p][j];
                s_u[i][j] = u;
            }
            __syncthreads();
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) x[0][j] = s_u[jj][j];
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int r = k0 + t + kPT * q;
                if (r >= b0 + IB && r < n) {
                    T l[IB];
#pragma unroll
                    for (int p = 0; p < IB; p++) l[p] = a[r + (int64_t)(b0 + p) * lda];
#pragma unroll
                    for (int j = 0; j < IB; j++) {
                        T acc = x[q][j];
#pragma unroll
                        for (int p = 0; p < IB; p++) acc -= l[p] * s_u[p][j];
                        x[q][j] = acc;
                    }
                }
            }
        }
        // factor the sub-panel
#pragma unroll
        for (int jj = 0; jj < IB; jj++) {
            if (jj < ib) {
                const int c = c0 + jj;
                T best = T(-1);
                int bi = n;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const int r = k0 + t + kPT * q;
                    if (r >= c && r < n) argmax_merge(best, bi, (T)fabs(x[q][jj]), r);
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    const T ob = __shfl_xor(best, o);
                    const int oi = __shfl_xor(bi, o);
                    argmax_merge(best, bi, ob, oi);
                }
                if (lane == 0) {
                    s_val[w] = best;
                    s_idx[w] = bi;
                }
                __syncthreads();
                best = s_val[0];
                bi = s_idx[0];
#pragma unroll
                for (int ww = 1; ww < kPW; ww++) argmax_merge(best, bi, s_val[ww], s_idx[ww]);
                const int p = bi;
                const int tp = (p - k0) % kPT, qp = (p - k0) / kPT;
                const int tc = (c - k0) % kPT, qc = (c - k0) / kPT;
                if (t == tp) {
#pragma unroll
                    for (int q = 0; q < Q; q++)
                        if (q == qp) {
#pragma unroll
                            for (int j = 0; j < IB; j++) s_rowp[j] = x[q][j];
                        }
                }
                if (t == tc) {
#pragma unroll
                    for (int q = 0; q < Q; q++)
                        if (q == qc) {
#pragma unroll
                            for (int j = 0; j < IB; j++) s_rowc[j] = x[q][j];
                        }
                }
                __syncthreads();
                const T piv = s_rowp[jj];
                const bool nz = piv != T(0);
                if (t == 0) {
                    ipiv[c] = p + 1;
                    s_piv[c - k0] = nz ? p : c;
                    if (!nz && info[0] == 0) info[0] = c + 1;
                }
                if (nz) {
                    if (p != c) {
                        if (t == tc) {
#pragma unroll
                            for (int q = 0; q < Q; q++)
                                if (q == qc) {
#pragma unroll
                                    for (int j = 0; j < IB; j++) x[q][j] = s_rowp[j];
                                }
                        }
                        if (t == tp) {
#pragma unroll
                            for (int q = 0; q < Q; q++)
                                if (q == qp) {
#pragma unroll
                                    for (int j = 0; j < IB; j++) x[q][j] = s_rowc[j];
                                }
                        }
                        if (t < kb) {
                            const int col = k0 + t;
                            if (col < c0 || col >= c0 + ib) {
                                T* cc = a + (int64_t)col * lda;
                                const T vc = cc[c], vp = cc[p];
                                cc[c] = vp;
                                cc[p] = vc;
                            }
                        }
                    }
                    const T rinv = T(1) / piv;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const int r = k0 + t + kPT * q;
                        if (r > c && r < n) {
                            const T l = x[q][jj] * rinv;
                            x[q][jj] = l;
#pragma unroll
                            for (int j = jj + 1; j < IB; j++) x[q][j] -= l * s_rowp[j];
                        }
                    }
                }
            }
        }
        // sub-panel back to memory
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int r = k0 + t + kPT * q;
            if (r < n) {
#pragma unroll
                for (int j = 0; j < IB; j++)
                    if (j < ib) a[r + (int64_t)(c0 + j) * lda] = x[q][j];
            }
        }
        // inverse of the unit-lower diagonal block, for the later sub-panels
        if (c0 + ib < k0 + kb) {
            const int jj = t - (c0 - k0);
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) s_blk[jj][j] = x[0][j];
            }
            __syncthreads();
            if (t < IB) {
                const int j = t;
                T col[IB];
#pragma unroll
                for (int i = 0; i < IB; i++) col[i] = i == j ? T(1) : T(0);
#pragma unroll
                for (int i = 0; i < IB; i++) {
                    if (i > j) {
                        T v = T(0);
#pragma unroll
                        for (int p = 0; p < IB; p++)
                            if (p >= j && p < i) v -= s_blk[i][p] * col[p];
                        col[i] = v;
                    }
                }
#pragma unroll
                for (int i = 0; i < IB; i++) s_inv[s][i][j] = col[i];
            }
        }
        __syncthreads();
    }
    // (destination, source) rows of the composed interchanges, for the columns outside the panel
    if (t == 0) {
        int nout = 0;
        for (int i = 0; i < kb; i++) s_src[i] = k0 + i;
        for (int i = 0; i < kb; i++) {
            const int p = s_piv[i];
            if (p == k0 + i) continue;
            int* other;
            if (p < k0 + kb) {
                other = &s_src[p - k0];
            } else {
                int e = 0;
                while (e < nout && s_okey[e] != p) e++;
                if (e == nout) {
                    s_okey[e] = p;
                    s_oval[e] = p;
                    nout++;
                }
                other = &s_oval[e];
            }
            const int tmp = s_src[i];
            s_src[i] = *other;
            *other = tmp;
        }
        int cnt = 0;
        for (int i = 0; i < kb; i++)
            if (s_src[i] != k0 + i) {
                perm[1 + 2 * cnt] = k0 + i;
                perm[2 + 2 * cnt] = s_src[i];
                cnt++;
            }
        for (int e = 0; e < nout; e++)
            if (s_oval[e] != s_okey[e]) {
                perm[1 + 2 * cnt] = s_okey[e];
                perm[2 + 2 * cnt] = s_oval[e];
                cnt++;
            }
        perm[0] = cnt;
    }
}

// Row interchanges of one panel on the columns outside it (one wave, one column per lane), then
// for the columns right of the panel U12 = L11^{-1} A12 with L11 (unit lower) staged in LDS.
template <typename T>
__global__ __launch_bounds__(64) void k_swap_trsm(T* __restrict__ a, int64_t lda, int n, int k0, int kb, int nleft,
                                                  const int32_t* __restrict__ perm) {
    __shared__ T s_v[2 * NB][64];
    __shared__ T s_l[NB][NB + 1];
    const int lane = threadIdx.x;
    const bool left = (int)blockIdx.x < nleft;
    const int col = left ? blockIdx.x * 64 + lane : k0 + kb + (blockIdx.x - nleft) * 64 + lane;
    const bool ok = left ? col < k0 : col < n;
    const int cnt = perm[0];
    T* cc = a + (int64_t)col * lda;
    if (ok) {
        for (int e = 0; e < cnt; e++) s_v[e][lane] = cc[perm[2 + 2 * e]];
        for (int e = 0; e < cnt; e++) cc[perm[1 + 2 * e]] = s_v[e][lane];
    }
    if (left) return;
    for (int e = lane; e < kb * kb; e += 64) {
        const int r = e % kb, c = e / kb;
        s_l[r][c] = a[(k0 + r) + (int64_t)(k0 + c) * lda];
    }
    __syncthreads();
    if (!ok) return;
    T xr[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) xr[r] = r < kb ? cc[k0 + r] : T(0);
#pragma unroll
    for (int c = 0; c < NB; c++) {
        if (c < kb) {
            const T xc = xr[c];
#pragma unroll
            for (int r = c + 1; r < NB; r++)
                if (r < kb) xr[r] -= s_l[r][c] * xc;
        }
    }
#pragma unroll
    for (int r = 0; r < NB; r++)
        if (r < kb) cc[k0 + r] = xr[r];
}

// C(m x n) -= A(m x kb) B(kb x n) on the matrix cores: 64 x 64 tile per work-group of 4 waves, each
// wave a 32 x 32 quarter as 2 x 2 blocks of the 16x16x4 MFMA (A operand: lane l holds A(l%16, l/16);
// B operand: B(l/16, l%16); accumulator register r of lane l holds C(row(l/16, r), l%16) with
// row = 4 (l/16) + r for the FP32 instruction and l/16 + 4 r for the FP64 one, as measured on
// gfx950).  A and B are staged k-major in LDS; the accumulators start from C and A enters negated.
template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(double x, double y, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <>
struct MfmaVec<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(float x, float y, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_update_mfma(T* __restrict__ C, int64_t ldc, const T* __restrict__ A,
                                                     int64_t lda, const T* __restrict__ B, int64_t ldb, int m, int n,
                                                     int kb) {
    using V = typename MfmaVec<T>::v4;
    __shared__ T sa[NB][64 + 1];
    __shared__ T sb[NB][64 + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    for (int e = t; e < 64 * NB; e += 256) {
        const int i = e % 64, k = e / 64;
        sa[k][i] = (i0 + i < m && k < kb) ? -A[(i0 + i) + (int64_t)k * lda] : T(0);
    }
    for (int e = t; e < 64 * NB; e += 256) {
        const int k = e % NB, j = e / NB;
        sb[k][j] = (j0 + j < n && k < kb) ? B[k + (int64_t)(j0 + j) * ldb] : T(0);
    }
    const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
    const int lr = lane & 15, lg = lane >> 4;
    V acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
            const int col = j0 + wj + 16 * bj + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i0 + wi + 16 * bi + MfmaVec<T>::row(lg, r);
                acc[bi][bj][r] = (row < m && col < n) ? C[row + (int64_t)col * ldc] : T(0);
            }
        }
    __syncthreads();
    const int ksteps = (kb + 3) / 4;
    for (int ks = 0; ks < ksteps; ks++) {
        const int k = 4 * ks + lg;
        T af[2], bf[2];
#pragma unroll
        for (int bi = 0; bi < 2; bi++) af[bi] = sa[k][wi + 16 * bi + lr];
#pragma unroll
        for (int bj = 0; bj < 2; bj++) bf[bj] = sb[k][wj + 16 * bj + lr];
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < 2; bj++) acc[bi][bj] = MfmaVec<T>::mfma(af[bi], bf[bj], acc[bi][bj]);
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
            const int col = j0 + wj + 16 * bj + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i0 + wi + 16 * bi + MfmaVec<T>::row(lg, r);
                if (row < m && col < n) C[row + (int64_t)col * ldc] = acc[bi][bj][r];
            }
        }
}

template <typename T, int Q, int IB>
int launch_panel_reg(T* a, int64_t lda, int n, int k0, int kb, int32_t* ipiv, int32_t* info, int32_t* perm,
                     hipStream_t st) {
    hipLaunchKernelGGL((k_panel_reg<T, Q, IB>), dim3(1), dim3(kPT), 0, st, a, lda, n, k0, kb, ipiv, info, perm);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// sub-panel width per rows-per-thread Q: Q x IB values per thread stay in registers without spills
template <typename T>
int panel_fast(T* a, int64_t lda, int n, int k0, int kb, int32_t* ipiv, int32_t* info, int32_t* perm, hipStream_t st) {
    const int m = n - k0;
    constexpr bool D = sizeof(T) == 8;
    if (m <= kPT) return launch_panel_reg<T, 1, 8>(a, lda, n, k0, kb, ipiv, info, perm, st);
    if (m <= 2 * kPT) return launch_panel_reg<T, 2, D ? 4 : 8>(a, lda, n, k0, kb, ipiv, info, perm, st);
    if (m <= 4 * kPT) return launch_panel_reg<T, 4, D ? 4 : 8>(a, lda, n, k0, kb, ipiv, info, perm, st);
    if (m <= 8 * kPT) return launch_panel_reg<T, 8, D ? 2 : 4>(a, lda, n, k0, kb, ipiv, info, perm, st);
    return launch_panel_reg<T, 16, D ? 1 : 2>(a, lda, n, k0, kb, ipiv, info, perm, st);
}

constexpr int kFastPanelRows = 16 * kPT;  // larger matrices take the column-by-column panel

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
    if (n <= kFastPanelRows) {
        void* ws = nullptr;
        TVLM_TRY(device_workspace(sizeof(int32_t) * (2 * 2 * NB + 1), &ws, kWsLuPerm));
        int32_t* perm = static_cast<int32_t*>(ws);
        for (int k0 = 0; k0 < n; k0 += NB) {
            const int kb = (n - k0 < NB) ? n - k0 : NB;
            TVLM_TRY(panel_fast<T>(a, lda, n, k0, kb, ipiv, info, perm, st));
            const int j1 = k0 + kb;
            const int nleft = (k0 + 63) / 64, nright = (n - j1 + 63) / 64;
            if (nleft + nright > 0) {
                hipLaunchKernelGGL(k_swap_trsm<T>, dim3(nleft + nright), dim3(64), 0, st, a, lda, n, k0, kb, nleft,
                                   perm);
                TVLM_LAUNCHED();
            }
            if (j1 >= n) continue;
            const int m = n - j1;
            hipLaunchKernelGGL(k_update_mfma<T>, dim3((m + 63) / 64, (m + 63) / 64), dim3(256), 0, st,
                               a + j1 + (int64_t)j1 * lda, lda, a + j1 + (int64_t)k0 * lda, lda,
                               a + k0 + (int64_t)j1 * lda, lda, m, m, kb);
            TVLM_LAUNCHED();
        }
        return TVLM_OK;
    }
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

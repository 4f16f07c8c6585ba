// tvlm_lu.hip -- dense LU with partial pivoting (getrf) and solve (getrs) on gfx950.
//
// Semantics follow LAPACK ?getrf/?getrs as called by the reference (vlm_backend_cpu_lapack.cpp:40-94):
// column-major in place L\U, unit lower L, 1-based ipiv, pivot = first row of max |a| in the
// column, info = first exactly-zero pivot (1-based), factorisation continues past it.
//
// getrf is blocked right-looking; each panel (64 columns, 32 while more than 2048 rows remain):
//   1. k_panel_reg  : one work-group factors the panel (register-resident sub-panels, DPP argmax,
//                     LDS-only barriers) and lists the panel's composed row interchanges; panels
//                     of more than 1024 rows of one matrix (or of a few) are factored by several
//                     work-groups together instead (k_panel_coop, 64 columns, one exchange through
//                     global memory per column);
//   2. k_swap_trsm  : the interchanges on the columns left and right of the panel, and
//                     U12 = L11^{-1} A12 for the columns right of it;
//   3. k_update     : A22 -= L21 U12 on the matrix cores (double-buffered 16-deep k-slices).
// Panels taller than the cooperative groups hold (and taller than 8192 rows) use the
// column-by-column k_panel + k_laswp + k_trsm instead.
// getrs (n <= 16384): the interchanges are composed per block of 64 pivots and the 64 x 64
// diagonal blocks of L and U inverted (k_solve_prep), then one work-group per right-hand side
// applies the interchanges and both blocked substitutions with the column in LDS
// (k_solve_blocked); from n = 1024 with few right-hand sides, the interchanges are applied alone
// (k_solve_perm) and each 64-row block step runs on all trailing row blocks at once
// (k_trsv_step); larger n uses block-row kernels (k_diag_solve / k_off_update).
#include <atomic>
#include <map>
#include <utility>
#include <mutex>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "tvlm_internal.hpp"

namespace tvlm {
namespace {

constexpr int kPanelThreads = 1024;
constexpr int NB = 64;


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

// ---- fast getrf: register-resident panel, fused interchange + TRSM, MFMA trailing update ------
//
// k_panel_reg: one work-group of NT threads factors the panel rows [k0, n) x columns [k0, k0+kb),
// kb <= NB.  Thread t owns rows k0 + t + NT q (q < Q).  The panel is processed left-looking in
// sub-panels of IB columns held in registers (Q x IB values per thread): a sub-panel first receives
// the updates of the panel's earlier sub-panels (blocked forward substitution with the inverses of
// their unit-lower IB x IB diagonal blocks, kept in LDS), then is factored column by column with
// partial pivoting (argmax by wave shuffles + one LDS exchange, first index on ties; the pivot row
// and row k exchange through LDS; reciprocal scaling and rank-1 update in registers).  Inside the
// column loop only LDS is touched and the work-group synchronises with an LDS-only barrier, so no
// step waits for global memory; pivots stay in LDS until the panel ends.  Each sub-panel's
// interchanges reach the panel's other columns as one gather / scatter, and at the end thread 0
// composes the panel's interchanges into (destination, source) row pairs for the rest of the matrix.

constexpr int kPermInts = 1 + 2 * 2 * NB;  // composed interchanges of one panel

// barrier for LDS traffic only: waits for this wave's LDS operations, not for global memory
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <typename T>
__device__ __forceinline__ void argmax_merge(T& best, int& bi, T v, int i) {
    const bool take = (v > best) | ((v == best) & (i < bi));
    best = take ? v : best;
    bi = take ? i : bi;
}

// One DPP step of a wave argmax: the value and index of the lane CTRL selects (row_shr / row_bcast
// within the row and bank masks) are merged in; lanes without a source keep their own pair.
template <int CTRL, int ROWM>
__device__ __forceinline__ void dpp_argmax_step(double& v, int& i) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWM, 0xf, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWM, 0xf, false);
    const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, ROWM, 0xf, false);
    argmax_merge(v, i, __hiloint2double(ohi, olo), oi);
}
template <int CTRL, int ROWM>
__device__ __forceinline__ void dpp_argmax_step(float& v, int& i) {
    const int ov = __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWM, 0xf, false);
    const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, ROWM, 0xf, false);
    argmax_merge(v, i, __int_as_float(ov), oi);
}
__device__ __forceinline__ double readlane63(double v) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}
__device__ __forceinline__ float readlane63(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// wave-wide argmax (larger value, then smaller index), result in every lane: row_shr 1, 2, 4, 8
// gathers each row's maximum into its lane 15, row_bcast 15 / 31 carry it across rows to lane 63
template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& i) {
    dpp_argmax_step<0x111, 0xf>(v, i);
    dpp_argmax_step<0x112, 0xf>(v, i);
    dpp_argmax_step<0x114, 0xf>(v, i);
    dpp_argmax_step<0x118, 0xf>(v, i);
    dpp_argmax_step<0x142, 0xa>(v, i);
    dpp_argmax_step<0x143, 0xc>(v, i);
    v = readlane63(v);
    i = __builtin_amdgcn_readlane(i, 63);
}

// Composes the interchanges pos[i] <-> piv[i] (i < cnt <= 64, applied in order, piv[i] >= pos[i],
// positions [first, first + cnt)) into (dst, src) row pairs, using one wave: lane l holds the source
// row of position first + l, and the rows outside the block that the interchanges touch are kept as
// (position, source) entries, one per lane.  Every lane calls it; returns the pair count (<= 2 cnt).
__device__ int wave_compose_swaps(int first, int cnt, const int* piv, int* pd, int* ps) {
    const int lane = threadIdx.x & 63;
    int src = first + lane;  // position first + lane
    int okey = -1, oval = -1;
    int nout = 0;
    for (int i = 0; i < cnt; i++) {
        const int p = __builtin_amdgcn_readfirstlane(piv[i]);
        if (p == first + i) continue;
        const int vi = __builtin_amdgcn_readlane(src, i);
        if (p < first + cnt) {
            const int j = p - first;
            const int vj = __builtin_amdgcn_readlane(src, j);
            src = lane == i ? vj : (lane == j ? vi : src);
        } else {
            const unsigned long long m = __ballot(okey == p);
            int e;
            if (m != 0ull) {
                e = __ffsll((long long)m) - 1;
            } else {
                e = nout++;
                okey = lane == e ? p : okey;
                oval = lane == e ? p : oval;
            }
            const int ve = __builtin_amdgcn_readlane(oval, e);
            src = lane == i ? ve : src;
            oval = lane == e ? vi : oval;
        }
    }
    const bool in_pair = lane < cnt && src != first + lane;
    const bool out_pair = lane < nout && oval != okey;
    const unsigned long long mi = __ballot(in_pair), mo = __ballot(out_pair);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int ni = __popcll(mi);
    if (in_pair) {
        const int k = __popcll(mi & below);
        pd[k] = first + lane;
        ps[k] = src;
    }
    if (out_pair) {
        const int k = ni + __popcll(mo & below);
        pd[k] = okey;
        ps[k] = oval;
    }
    return ni + __popcll(mo);
}

template <typename T, int NT, int Q, int IB>
__global__ __launch_bounds__(NT) void k_panel_reg(T* __restrict__ a, int64_t lda, int n, int k0, int kb,
                                                  int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                  int32_t* __restrict__ perm, int64_t sb) {
    constexpr int NW = NT / 64;
    a += (int64_t)blockIdx.z * sb;  // batch: one matrix per blockIdx.z
    ipiv += (int64_t)blockIdx.z * n;
    info += blockIdx.z;
    perm += (int64_t)blockIdx.z * kPermInts;
    __shared__ T s_val[NW];
    __shared__ int s_idx[NW];
    __shared__ T s_rowp[IB], s_rowc[IB];
    __shared__ T s_inv[NB / IB][IB][IB];
    __shared__ T s_blk[IB][IB];
    __shared__ T s_u[IB][IB];
    __shared__ int s_arg[NB], s_piv[NB];
    __shared__ int s_pd[2 * NB], s_ps[2 * NB], s_pcnt, s_zero;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_zero = 0;
    T x[Q][IB];

    for (int c0 = k0, s = 0; c0 < k0 + kb; c0 += IB, s++) {
        const int ib = min(IB, k0 + kb - c0);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int r = k0 + t + NT * q;
#pragma unroll
            for (int j = 0; j < IB; j++) x[q][j] = (r < n && j < ib) ? a[r + (int64_t)(c0 + j) * lda] : T(0);
        }
        // updates from the earlier sub-panels sp < s (rows b0 .. b0+IB of block sp hold U after this);
        // when registers allow, the L block of round sp + 1 is loaded while round sp computes (two
        // register buffers, rounds taken in pairs so that both stay statically indexed)
        T l0[Q][IB], l1[Q][IB];
        auto load_l = [&](T (&l)[Q][IB], int b) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int r = k0 + t + NT * q;
#pragma unroll
                for (int p = 0; p < IB; p++) l[q][p] = (r >= b + IB && r < n) ? a[r + (int64_t)(b + p) * lda] : T(0);
            }
        };
        auto round = [&](int sp, T (&lc)[Q][IB], T (&ln)[Q][IB], bool prefetch) {
            const int b0 = k0 + sp * IB;
            if (prefetch && sp + 1 < s) load_l(ln, b0 + IB);
            const int jj = t - (b0 - k0);  // owners of the block's rows: q = 0
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) s_blk[jj][j] = x[0][j];
            }
            lds_sync();
            if (t < IB * IB) {
                const int i = t / IB, j = t % IB;
                T u = T(0);
                for (int p = 0; p <= i; p++) u += s_inv[sp][i][p] * s_blk[p][j];
                s_u[i][j] = u;
            }
            lds_sync();
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) x[0][j] = s_u[jj][j];
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int r = k0 + t + NT * q;
                if (r >= b0 + IB && r < n) {
#pragma unroll
                    for (int j = 0; j < IB; j++) {
                        T acc = x[q][j];
#pragma unroll
                        for (int p = 0; p < IB; p++) acc -= lc[q][p] * s_u[p][j];
                        x[q][j] = acc;
                    }
                }
            }
            lds_sync();  // s_u, s_blk reused by the next block
        };
        if constexpr (Q * IB * sizeof(T) <= 256) {
            if (s > 0) load_l(l0, k0);
            for (int sp = 0; sp < s; sp += 2) {
                round(sp, l0, l1, true);
                if (sp + 1 < s) round(sp + 1, l1, l0, true);
            }
        } else {  // registers for one buffer only: each round loads its own block first
            for (int sp = 0; sp < s; sp++) {
                load_l(l0, k0 + sp * IB);
                round(sp, l0, l1, false);
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
                    const int r = k0 + t + NT * q;
                    if (r >= c && r < n) argmax_merge(best, bi, (T)fabs(x[q][jj]), r);
                }
                wave_argmax(best, bi);
                if (lane == 0) {
                    s_val[w] = best;
                    s_idx[w] = bi;
                }
                lds_sync();
                best = s_val[0];
                bi = s_idx[0];
#pragma unroll
                for (int ww = 1; ww < NW; ww++) argmax_merge(best, bi, s_val[ww], s_idx[ww]);
                // a column of NaNs wins no comparison and leaves bi = n: take row c then (LAPACK's
                // idamax also stays on the first entry), so every row index below stays in [c, n)
                const int p = (bi >= c && bi < n) ? bi : c;
                const int tp = (p - k0) % NT, qp = (p - k0) / NT;
                const int tc = (c - k0) % NT, qc = (c - k0) / NT;
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
                lds_sync();
                const T piv = s_rowp[jj];
                const bool nz = piv != T(0);
                if (t == 0) {
                    s_arg[c - k0] = p;
                    s_piv[c - k0] = nz ? p : c;
                    if (!nz && s_zero == 0) s_zero = c + 1;
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
                    }
                    const T rinv = T(1) / piv;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const int r = k0 + t + NT * q;
                        if (r > c && r < n) {
                            const T l = x[q][jj] * rinv;
                            x[q][jj] = l;
#pragma unroll
                            for (int j = jj + 1; j < IB; j++) x[q][j] -= l * s_rowp[j];
                        }
                    }
                }
                lds_sync();  // s_rowp / s_rowc / s_val reused by the next column
            }
        }
        // sub-panel back to memory
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int r = k0 + t + NT * q;
            if (r < n) {
#pragma unroll
                for (int j = 0; j < IB; j++)
                    if (j < ib) a[r + (int64_t)(c0 + j) * lda] = x[q][j];
            }
        }
        // the sub-panel's interchanges on the panel's other columns, as one gather / scatter
        if (t < 64) {
            const int np = wave_compose_swaps(c0, ib, s_piv + (c0 - k0), s_pd, s_ps);
            if (t == 0) s_pcnt = np;
        }
        __syncthreads();
        {
            // whole columns per pass: every source of a column is read before any of its rows is written
            const int pairs = s_pcnt;
            if (pairs > 0) {
                const int cpi = NT / pairs, cl = t / pairs, pr = t % pairs;
                for (int cb = 0; cb < kb; cb += cpi) {
                    const int col = k0 + cb + cl;
                    const bool act = cl < cpi && cb + cl < kb && (col < c0 || col >= c0 + ib);
                    T v = T(0);
                    if (act) v = a[s_ps[pr] + (int64_t)col * lda];
                    __syncthreads();
                    if (act) a[s_pd[pr] + (int64_t)col * lda] = v;
                    __syncthreads();
                }
            }
        }
        // inverse of the unit-lower diagonal block, for the later sub-panels
        if (c0 + ib < k0 + kb) {
            const int jj = t - (c0 - k0);
            if (jj >= 0 && jj < IB) {
#pragma unroll
                for (int j = 0; j < IB; j++) s_blk[jj][j] = x[0][j];
            }
            lds_sync();
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
    // pivots, first zero pivot, and the panel's composed interchanges for the rest of the matrix
    if (t < kb) ipiv[k0 + t] = s_arg[t] + 1;
    if (t == 0 && s_zero != 0 && info[0] == 0) info[0] = s_zero;
    if (t < 64) {
        const int np = wave_compose_swaps(k0, kb, s_piv, s_pd, s_ps);
        if (t == 0) s_pcnt = np;
    }
    __syncthreads();
    const int np = s_pcnt;
    for (int e = t; e < np; e += NT) {
        perm[1 + 2 * e] = s_pd[e];
        perm[2 + 2 * e] = s_ps[e];
    }
    if (t == 0) perm[0] = np;
}

// ---- cooperative tall panel ---------------------------------------------------------------------
// k_panel_coop: G work-groups of NT threads factor the panel rows [k0, n) x columns [k0, k0 + kb)
// together (kb <= NB).  Thread t of group g owns row k0 + g NT + t; the group keeps its NT rows of
// the panel in LDS, column-major (consecutive threads, consecutive banks), from the first column to
// the last.  Per column, one exchange through global memory: the first wave of every group
// publishes the group's local pivot candidate (|value|, row) and that row's values (group 0 also
// the diagonal row) from LDS, releases once, and counts the group in at a per-column counter (one
// atomic add each, then a bounded spin).  Every group then
// picks the same pivot (largest |value|, first row on ties: LAPACK's choice), the owners of the
// pivot and diagonal rows exchange them, and each group scales and updates its own rows.  Exchange
// buffers (per matrix): cval / crow [NB][kCoopMaxGroups], xrow [NB][kCoopMaxGroups + 1][NB] (slot
// kCoopMaxGroups: the diagonal row), counters ctr [NB], zeroed before the launch.  A spin that runs
// out (the groups were not all resident), or a pivot row outside [c, n) read back from the exchange,
// leaves info = kCoopStalled instead of waiting for ever or indexing out of bounds.
// Group 0 writes ipiv and the panel's composed interchanges for k_swap_trsm, as k_panel_reg does.
//
// Before the column-by-column exchange, an attempt with two exchanges in all: group 0 factors its
// own NT rows with partial pivoting among them and publishes U11 (its first kb rows); every other
// group eliminates its rows against U11 (the same operations, in the same order, as the exchange
// loop with those pivots) and checks |a(r, c)| <= |U(c, c)| for every row and column before
// scaling.  When that holds everywhere, partial pivoting over all rows picks exactly group 0's
// pivots (on a tie the first row, which is group 0's), so the result is the exchange loop's; the
// groups vote, and on any failure every group reloads its rows and runs the exchange loop.
// Influence matrices of rectangular and swept planforms never pivot outside the diagonal block and
// pass throughout; after kCoopOptFails failed attempts in a row the matrix (ostate, per matrix,
// zeroed with the counters) skips them.  Counters per panel and matrix: ctr[0, NB) the columns,
// ctr[NB] U11 published, ctr[NB + 1] the votes, ctr[NB + 2] a failed vote.
constexpr int kCoopMaxGroups = 64;   // work-groups of one launch; all of them must be resident at once
constexpr int kCoopMinRows = 1024;   // shorter panels stay in one work-group (k_panel_reg)
constexpr unsigned kCoopSpinLimit = 1u << 22;  // default bound of every wait (tvlm_lu_set_coop_spin_limit)
constexpr int kCoopStalled = -999;
constexpr int kCoopCtr = NB + 4;                                                      // counters per panel and matrix
constexpr int64_t kCoopXrow = (int64_t)NB * (kCoopMaxGroups + 1) * NB + NB * NB;  // xrow per matrix (+ U11)
constexpr int kCoopOptFails = 2;

template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_agent(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// threads (= rows) per cooperative work-group: NT rows x NB columns of T = 128 KB of LDS
template <typename T>
constexpr int coop_threads() {
    return sizeof(T) == 4 ? 512 : 256;
}

// Column J of the first group's local factorisation (k_panel_coop's single-exchange attempt); cols()
// runs all NB of them in order, one instantiation per column, so the row stays in registers.
template <typename T, int NT>
struct CoopLocal {
    T* r;
    int k0, n, i;
    bool mine;
    int t, lane, w, wu;
    T* s_val;
    int *s_idx;
    T *s_prow, *s_crow;
    int *s_arg, *s_piv, *s_zero;

    template <int J>
    __device__ __forceinline__ void col() {
        constexpr int NW = NT / 64;
        const int c = k0 + J;
        T best = T(-1);
        int bi = n;
        if (mine && i >= J) {
            best = (T)fabs(r[J]);
            bi = k0 + i;
        }
        wave_argmax(best, bi);
        if (lane == 0) {
            s_val[w] = best;
            s_idx[w] = bi;
        }
        __syncthreads();
        best = s_val[0];
        bi = s_idx[0];
#pragma unroll
        for (int ww = 1; ww < NW; ww++) argmax_merge(best, bi, s_val[ww], s_idx[ww]);
        const int pl = (bi >= k0 + J && bi < k0 + NT && bi < n) ? bi - k0 : J;
        if (wu == (pl >> 6) && lane == (pl & 63)) {
#pragma unroll
            for (int jj = 0; jj < NB; jj++) s_prow[jj] = r[jj];
        }
        if (pl != J && wu == 0 && lane == J) {
#pragma unroll
            for (int jj = 0; jj < NB; jj++) s_crow[jj] = r[jj];
        }
        __syncthreads();
        const T piv = s_prow[J];
        const bool nz = piv != T(0);
        if (t == 0) {
            s_arg[J] = k0 + pl;
            s_piv[J] = nz ? k0 + pl : c;
            if (!nz && *s_zero == 0) *s_zero = c + 1;
        }
        if (nz && pl != J) {
            if (wu == 0 && lane == J) {
#pragma unroll
                for (int jj = 0; jj < NB; jj++) r[jj] = s_prow[jj];
            }
            if (wu == (pl >> 6) && lane == (pl & 63)) {
#pragma unroll
                for (int jj = 0; jj < NB; jj++) r[jj] = s_crow[jj];
            }
        }
        if (nz && mine && i > J) {
            const T l = r[J] * (T(1) / piv);
            r[J] = l;
#pragma unroll
            for (int jj = J + 1; jj < NB; jj++) r[jj] -= l * s_prow[jj];
        }
        __syncthreads();  // s_val, s_prow and s_crow are rewritten for the next column
    }

    template <int... J>
    __device__ __forceinline__ void cols(std::integer_sequence<int, J...>) {
        (col<J>(), ...);
    }
};

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_panel_coop(T* __restrict__ a, int64_t lda, int n, int k0, int kb, int G,
                                                   int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                   int32_t* __restrict__ perm, int64_t sb, int32_t* ctr, T* cval,
                                                   int32_t* crow, T* xrow, int32_t* ostate, unsigned spin_limit,
                                                   int32_t* fail) {
    constexpr int NW = NT / 64, MG = kCoopMaxGroups;
    const int z = blockIdx.z, g = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    a += (int64_t)z * sb;
    ipiv += (int64_t)z * n;
    info += z;
    perm += (int64_t)z * kPermInts;
    ctr += (int64_t)z * kCoopCtr;
    cval += (int64_t)z * NB * MG;
    crow += (int64_t)z * NB * MG;
    xrow += (int64_t)z * kCoopXrow;
    const int m = n - k0, i = g * NT + t;  // this thread's row: k0 + i
    const bool mine = i < m;
    extern __shared__ unsigned char coop_lds[];
    T* xs = reinterpret_cast<T*>(coop_lds);  // xs[jj * NT + t]: column jj of this group's row t
    __shared__ T s_val[NW];
    __shared__ int s_idx[NW];
    __shared__ T s_prow[NB], s_crow[NB];
    __shared__ int s_p, s_fail, s_zero, s_arg[NB], s_piv[NB];
    __shared__ int s_pd[2 * NB], s_ps[2 * NB], s_pcnt, s_opt;
    __shared__ T s_u[NB * (NB + 1) / 2];  // U11 of the attempt, row r from column r on, rows packed
    if (t == 0) {
        s_fail = 0;
        s_zero = 0;
    }
    for (int jj = 0; jj < kb; jj++) xs[jj * NT + t] = mine ? a[k0 + i + (int64_t)(k0 + jj) * lda] : T(0);
    __syncthreads();
    bool settled = false;
    const int fails = ld_agent(&ostate[z]);
    if (kb == NB && fails < kCoopOptFails) {
        int32_t* octr = ctr + NB;
        T* xu = xrow + (int64_t)NB * (MG + 1) * NB;
        bool ok = true;
        if (g == 0) {
            // partial pivoting among this group's rows (all NT of them hold rows: m > kCoopMinRows),
            // each row in its thread's registers (one instantiation per column: every index is a
            // constant); the pivot row is broadcast through s_prow, the row it displaces through s_crow
            T r[NB];
#pragma unroll
            for (int jj = 0; jj < NB; jj++) r[jj] = xs[jj * NT + t];
            CoopLocal<T, NT> L{r, k0, n, i, mine, t, lane, w, __builtin_amdgcn_readfirstlane(w), s_val, s_idx, s_prow,
                               s_crow, s_arg, s_piv, &s_zero};
            L.cols(std::make_integer_sequence<int, NB>{});
#pragma unroll
            for (int jj = 0; jj < NB; jj++) xs[jj * NT + t] = r[jj];
            __syncthreads();
            for (int e = t; e < kb * kb; e += NT) {
                const int r = e / kb, cc = e % kb;
                if (cc >= r) st_agent(&xu[r * NB + cc], xs[cc * NT + r]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (t == 0) __hip_atomic_fetch_add(&octr[0], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (t == 0) {
                unsigned spins = 0;
                while (__hip_atomic_load(&octr[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < 1) {
                    if (++spins > spin_limit) {
                        s_fail = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __syncthreads();
            for (int e = t; e < kb * kb; e += NT) {
                const int r = e / kb, cc = e % kb;
                if (cc >= r) s_u[r * NB - r * (r - 1) / 2 + (cc - r)] = ld_agent(&xu[r * NB + cc]);
            }
            __syncthreads();
            bool okr = s_fail == 0;
            if (mine) {  // the row in registers (kb == NB here: every index below is a constant)
                T r[NB];
#pragma unroll
                for (int jj = 0; jj < NB; jj++) r[jj] = xs[jj * NT + t];
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const int b = j * NB - j * (j - 1) / 2;
                    const T piv = s_u[b];
                    okr = okr && ((T)fabs(r[j]) <= (T)fabs(piv));
                    if (piv != T(0)) {
                        const T l = r[j] * (T(1) / piv);
                        r[j] = l;
#pragma unroll
                        for (int jj = j + 1; jj < NB; jj++) r[jj] -= l * s_u[b + (jj - j)];
                    }
                }
#pragma unroll
                for (int jj = 0; jj < NB; jj++) xs[jj * NT + t] = r[jj];
            }
            ok = __syncthreads_and(okr) != 0;
        }
        // the vote: every group counts in, a failing one first raises the flag
        if (t == 0) {
            if (!ok) st_agent(&octr[2], 1);
            __hip_atomic_fetch_add(&octr[1], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            bool stalled = false;
            while (__hip_atomic_load(&octr[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < G) {
                if (++spins > spin_limit) {
                    stalled = true;
                    s_fail = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            s_opt = (!stalled && ld_agent(&octr[2]) == 0) ? 1 : 0;
        }
        __syncthreads();
        settled = s_opt != 0;
        if (g == 0 && t == 0) st_agent(&ostate[z], settled ? 0 : fails + 1);
        if (!settled) {  // back to the rows as loaded
            if (t == 0) s_zero = 0;
            for (int jj = 0; jj < kb; jj++) xs[jj * NT + t] = mine ? a[k0 + i + (int64_t)(k0 + jj) * lda] : T(0);
            __syncthreads();
        }
    }
    for (int j = 0; j < kb && !settled; j++) {
        const int c = k0 + j;
        // this group's candidate: largest |a(r, c)| over its rows r >= c, first row on ties
        T best = T(-1);
        int bi = n;
        if (mine && i >= j) {
            best = (T)fabs(xs[j * NT + t]);
            bi = k0 + i;
        }
        wave_argmax(best, bi);
        if (lane == 0) {
            s_val[w] = best;
            s_idx[w] = bi;
        }
        __syncthreads();
        best = s_val[0];
        bi = s_idx[0];
#pragma unroll
        for (int ww = 1; ww < NW; ww++) argmax_merge(best, bi, s_val[ww], s_idx[ww]);
        // publish and meet: wave 0 writes this group's candidate row and (group 0) the diagonal row
        // straight from LDS, then the candidate, releases once, and its lane 0 counts the group in
        // and waits for the others; the other waves go straight to the barrier
        T* xr = xrow + (int64_t)j * (MG + 1) * NB;
        if (w == 0) {
            if (bi < n && lane < kb) st_agent(&xr[g * NB + lane], xs[lane * NT + (bi - k0 - g * NT)]);
            if (g == 0 && lane < kb) st_agent(&xr[MG * NB + lane], xs[lane * NT + j]);  // row c: local row j
            if (lane == 0) {
                st_agent(&cval[j * MG + g], best);
                st_agent(&crow[j * MG + g], bi);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0) {
                __hip_atomic_fetch_add(&ctr[j], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                unsigned spins = 0;
                while (__hip_atomic_load(&ctr[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < G) {
                    if (++spins > spin_limit) {
                        s_fail = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();
        // the pivot: the same reduction in every group
        if (w == 0) {
            T v = T(-1);
            int ri = n;
            if (lane < G) {
                v = ld_agent(&cval[j * MG + lane]);
                ri = ld_agent(&crow[j * MG + lane]);
            }
            wave_argmax(v, ri);
            if (lane == 0) {
                // ri == n: no group had a candidate (a column of NaNs): row c, as LAPACK's idamax;
                // any other pivot row outside [c, n) can only come from a failed exchange: keep every
                // index below in bounds and report it
                const bool ok = ri >= c && ri < n;
                s_p = ok ? ri : c;
                if (!ok && ri != n) s_fail = 1;
            }
        }
        __syncthreads();
        const int p = s_p;
        const int gw = (p - k0) / NT;
        if (t < kb) s_prow[t] = ld_agent(&xr[gw * NB + t]);
        if (t >= 64 && t - 64 < kb) s_crow[t - 64] = ld_agent(&xr[MG * NB + (t - 64)]);
        __syncthreads();
        const T piv = s_prow[j];
        const bool nz = piv != T(0);
        if (g == 0 && t == 0) {
            s_arg[j] = p;
            s_piv[j] = nz ? p : c;
            if (!nz && s_zero == 0) s_zero = c + 1;
        }
        if (nz) {
            if (p != c) {
                if (g == 0 && t < kb) xs[t * NT + j] = s_prow[t];
                if (gw == g && t < kb) xs[t * NT + (p - k0 - g * NT)] = s_crow[t];
                __syncthreads();
            }
            if (mine && i > j) {
                const T l = xs[j * NT + t] * (T(1) / piv);
                xs[j * NT + t] = l;
#pragma unroll 8
                for (int jj = j + 1; jj < kb; jj++) xs[jj * NT + t] -= l * s_prow[jj];
            }
        }
        __syncthreads();  // s_val / s_prow / s_crow / s_p reused by the next column
    }
    if (mine)
        for (int jj = 0; jj < kb; jj++) a[k0 + i + (int64_t)(k0 + jj) * lda] = xs[jj * NT + t];
    // any group whose wait ran out marks the matrix (k_coop_fail_info turns the mark into info)
    if (t == 0 && s_fail) __hip_atomic_fetch_or(&fail[z], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g != 0) return;
    if (t < kb) ipiv[k0 + t] = s_arg[t] + 1;
    if (t == 0) {
        if (s_fail)
            info[0] = kCoopStalled;
        else if (s_zero != 0 && info[0] == 0)
            info[0] = s_zero;
    }
    if (t < 64) {
        const int np = wave_compose_swaps(k0, kb, s_piv, s_pd, s_ps);
        if (t == 0) s_pcnt = np;
    }
    __syncthreads();
    const int np = s_pcnt;
    for (int e = t; e < np; e += NT) {
        perm[1 + 2 * e] = s_pd[e];
        perm[2 + 2 * e] = s_ps[e];
    }
    if (t == 0) perm[0] = np;
}

// Row interchanges of one panel on the columns outside it, then for the columns right of the panel
// U12 = L11^{-1} A12 (L11 unit lower).  A work-group of 256 threads takes 64 columns: the swap pairs
// are gathered for all columns at once, then scattered; U12's 64 x 64 block and L11 sit in LDS and
// the forward substitution runs one row per step over all 64 columns.
// mode: 0 = interchanges on both sides and U12 on the right; 1 = interchanges only; 2 = U12 only
// (the right blocks, no interchanges).  With gtab, the left blocks before column left0 take the
// interchanges into their gather lists instead of their columns (deferred interchanges, below).  The
// left blocks start at column lbase, the right ones at rbase (default: right after the panel).
template <typename T>
__global__ __launch_bounds__(256) void k_swap_trsm(T* __restrict__ a, int64_t lda, int n, int k0, int kb, int nleft,
                                                   const int32_t* __restrict__ perm, int64_t sb, int mode = 0,
                                                   int left0 = 0, int32_t* __restrict__ gtab = nullptr, int lbase = 0,
                                                   int rbase = -1) {
    a += (int64_t)blockIdx.z * sb;
    perm += (int64_t)blockIdx.z * kPermInts;
    constexpr int kPer = 2 * NB * 64 / 256;  // gathered elements per thread
    __shared__ int s_pd[2 * NB], s_ps[2 * NB];
    __shared__ T s_x[NB][64 + 1];
    __shared__ T s_l[NB][NB + 1];
    const int t = threadIdx.x;
    const bool left = (int)blockIdx.x < nleft;
    const int col0 = left ? lbase + blockIdx.x * 64 : (rbase < 0 ? k0 + kb : rbase) + (blockIdx.x - nleft) * 64;
    const int colend = left ? k0 : n;
    const int cnt = mode == 2 ? 0 : perm[0];
    if (left && gtab && col0 < left0) {
        // g'[dst] = g[src] for every pair of the panel's composed interchanges (cnt <= 2 NB)
        int32_t* g = gtab + ((int64_t)blockIdx.z * ((n + NB - 1) / NB) + col0 / NB) * n;
        const int v = t < cnt ? g[perm[2 + 2 * t]] : 0;
        __syncthreads();
        if (t < cnt) g[perm[1 + 2 * t]] = v;
        return;
    }
    for (int e = t; e < cnt; e += 256) {
        s_pd[e] = perm[1 + 2 * e];
        s_ps[e] = perm[2 + 2 * e];
    }
    __syncthreads();
    T v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int e = t + 256 * u, pr = e >> 6, col = col0 + (e & 63);
        v[u] = (pr < cnt && col < colend) ? a[s_ps[pr] + (int64_t)col * lda] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int e = t + 256 * u, pr = e >> 6, col = col0 + (e & 63);
        if (pr < cnt && col < colend) a[s_pd[pr] + (int64_t)col * lda] = v[u];
    }
    if (left || mode == 1) return;
    __syncthreads();
    for (int e = t; e < NB * 64; e += 256) {
        const int r = e % NB, c = e / NB;
        s_x[r][c] = (r < kb && col0 + c < colend) ? a[(k0 + r) + (int64_t)(col0 + c) * lda] : T(0);
    }
    for (int e = t; e < kb * kb; e += 256) {
        const int r = e % kb, c = e / kb;
        s_l[r][c] = a[(k0 + r) + (int64_t)(k0 + c) * lda];
    }
    __syncthreads();
    const int c = t & 63, g = t >> 6;
    for (int k = 0; k < kb - 1; k++) {
        const T xk = s_x[k][c];
#pragma unroll
        for (int m = 0; m < NB / 4; m++) {
            const int r = g + 4 * m;
            if (r > k && r < kb) s_x[r][c] -= s_l[r][k] * xk;
        }
        lds_sync();
    }
    for (int e = t; e < NB * 64; e += 256) {
        const int r = e % NB, cc = e / NB;
        if (r < kb && col0 + cc < colend) a[(k0 + r) + (int64_t)(col0 + cc) * lda] = s_x[r][cc];
    }
}

// C(m x n) -= A(m x kb) B(kb x n), the trailing update, on the matrix cores (any kb; batched over
// blockIdx.z with one stride for A, B and C).  One work-group of 4 waves per TI x TJ tile of C,
// waves arranged WAVES_I x (4 / WAVES_I) over it; each wave holds its block as NBI x NBJ MFMA
// blocks (FP32: v_mfma_f32_32x32x2f32, FP64: v_mfma_f64_16x16x4f64).  Operand layouts on gfx950:
// first operand lane l = P(l % BM, l / BM), second Q(l / BM, l % BM) (m, k and k, n); output
// register r of lane l = D(drow(l, r), l % BM).  The kernel computes D = (A B)^T block by block
// (P from B, Q from A), so that each lane's outputs lie along a column of C and the 64 lanes of one
// register read or write whole 128-byte row runs of C.
// A and B pass through LDS in 16-deep k-slices, double-buffered: the global loads of slice s + 1
// are in flight while the MFMAs consume slice s.  The accumulators start from zero and C is read
// once, in the epilogue (C - A B, the order of a BLAS gemm with alpha = -1, beta = 1).
// Tiles: FP32 128 x 64 (4 waves along the rows), 64 x 64 (2 x 2 waves) when the 128-row grid
// would hold fewer than 1024 work-groups; FP64 64 x 64 (4 waves along the rows).  Measured against
// rocBLAS ?gemm_strided_batched at k = 64: 2.79 vs 2.14 ms (m = n = 4032, batch 64, FP32),
// 0.73 vs 0.56 ms (m = n = 16320, FP32), 1.24 vs 1.04 ms (m = n = 16320, FP64).
template <typename T>
struct UpdCfg;
template <>
struct UpdCfg<float> {
    static constexpr int BM = 32, KS = 2, NACC = 16;
    typedef float vacc __attribute__((ext_vector_type(16)));
    static __device__ __forceinline__ vacc mfma(float p, float q, vacc c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(p, q, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
};
template <>
struct UpdCfg<double> {
    static constexpr int BM = 16, KS = 4, NACC = 4;
    typedef double vacc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ vacc mfma(double p, double q, vacc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(p, q, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int l, int r) { return (l >> 4) + 4 * r; }
};

constexpr int kUpdDK = 16;  // k-slice depth

// The FP32 128 x 64 tile is held to 6 waves per SIMD (its accumulators then sit in 80 vector registers
// beside the operands instead of 32 accumulation registers: 5 -> 6 waves; measured ~1 % on getrf).
template <typename T, int TI, int TJ, int WAVES_I>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && TI * TJ == 8192 ? 6 : 1)))
void k_update(T* __restrict__ C, int64_t ldc, const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
              int64_t ldb, int m, int n, int kb, int64_t sbatch) {
    using G = UpdCfg<T>;
    constexpr int BM = G::BM, KS = G::KS, DK = kUpdDK;
    constexpr int WI = TI / WAVES_I, WJ = TJ / (4 / WAVES_I), NBI = WI / BM, NBJ = WJ / BM;
    constexpr int NA = TI * DK / 256, NBq = TJ * DK / 256;  // staged elements per thread and slice
    static_assert(NBI >= 1 && NBJ >= 1 && NA >= 1 && NBq >= 1 && 256 % TI == 0, "tile shape");
    __shared__ T sa[2][DK][TI];
    __shared__ T sb[2][TJ][DK + 1];
    const int i0 = blockIdx.x * TI, j0 = blockIdx.y * TJ;
    // uniform tile bases: every per-lane offset below fits 32 bits (< 4 ld + TI)
    C += (int64_t)blockIdx.z * sbatch + i0 + (int64_t)j0 * ldc;
    A += (int64_t)blockIdx.z * sbatch + i0;
    B += (int64_t)blockIdx.z * sbatch + (int64_t)j0 * ldb;
    m -= i0;
    n -= j0;
    const int t = threadIdx.x, l = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wi = (w % WAVES_I) * WI, wj = (w / WAVES_I) * WJ;
    const int li = l % BM, lg = l / BM;
    const int ta_i = t % TI, ta_k = t / TI;  // A staging: rows along the threads
    const int tb_k = t % DK, tb_j = t / DK;  // B staging: k along the threads
    const int offA = ta_i + ta_k * (int)lda, offB = tb_k + tb_j * (int)ldb;
    const int nsl = (kb + DK - 1) / DK;
    T ra[NA], rb[NBq];
    auto load = [&](int s) {
        const int k0 = s * DK;
        const T* Ak = A + (int64_t)k0 * lda;
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int k = k0 + ta_k + q * (256 / TI);
            ra[q] = (ta_i < m && k < kb) ? Ak[(int64_t)(q * (256 / TI)) * lda + offA] : T(0);
        }
        const T* Bk = B + k0;
#pragma unroll
        for (int q = 0; q < NBq; q++) {
            const int j = tb_j + q * (256 / DK);
            rb[q] = (j < n && k0 + tb_k < kb) ? Bk[(int64_t)(q * (256 / DK)) * ldb + offB] : T(0);
        }
    };
    load(0);
    typename G::vacc acc[NBI][NBJ];
#pragma unroll
    for (int bi = 0; bi < NBI; bi++)
#pragma unroll
        for (int bj = 0; bj < NBJ; bj++)
#pragma unroll
            for (int r = 0; r < G::NACC; r++) acc[bi][bj][r] = T(0);
    for (int s = 0; s < nsl; s++) {
        const int buf = s & 1;
        // buffer buf was last read in slice s - 2, which every wave finished before the barrier of s - 1
#pragma unroll
        for (int q = 0; q < NA; q++) sa[buf][ta_k + q * (256 / TI)][ta_i] = ra[q];
#pragma unroll
        for (int q = 0; q < NBq; q++) sb[buf][tb_j + q * (256 / DK)][tb_k] = rb[q];
        __syncthreads();
        if (s + 1 < nsl) load(s + 1);
#pragma unroll
        for (int kk = 0; kk < DK; kk += KS) {
            T a[NBI], b[NBJ];
#pragma unroll
            for (int bi = 0; bi < NBI; bi++) a[bi] = sa[buf][kk + lg][wi + bi * BM + li];
#pragma unroll
            for (int bj = 0; bj < NBJ; bj++) b[bj] = sb[buf][wj + bj * BM + li][kk + lg];
#pragma unroll
            for (int bi = 0; bi < NBI; bi++)
#pragma unroll
                for (int bj = 0; bj < NBJ; bj++) acc[bi][bj] = G::mfma(b[bj], a[bi], acc[bi][bj]);
        }
    }
    const int offC = li + G::drow(l, 0) * (int)ldc;
#pragma unroll
    for (int bi = 0; bi < NBI; bi++) {
        const int i = wi + bi * BM + li;
#pragma unroll
        for (int bj = 0; bj < NBJ; bj++) {
            T cv[G::NACC];
#pragma unroll
            for (int r = 0; r < G::NACC; r++) {
                const int j = wj + bj * BM + G::drow(l, r);
                const T* Cp = C + (wi + bi * BM) + (int64_t)(wj + bj * BM + G::drow(0, r)) * ldc;
                cv[r] = (i < m && j < n) ? __builtin_nontemporal_load(Cp + offC) : T(0);
            }
#pragma unroll
            for (int r = 0; r < G::NACC; r++) {
                const int j = wj + bj * BM + G::drow(l, r);
                T* Cp = C + (wi + bi * BM) + (int64_t)(wj + bj * BM + G::drow(0, r)) * ldc;
                if (i < m && j < n) __builtin_nontemporal_store(cv[r] - acc[bi][bj][r], Cp + offC);
            }
        }
    }
}

template <typename T>
void launch_update(T* C, int64_t ldc, const T* A, int64_t lda, const T* B, int64_t ldb, int m, int n, int kb,
                   int64_t sb, int batch, hipStream_t st) {
    if (m <= 0 || n <= 0 || kb <= 0) return;
    if constexpr (sizeof(T) == 4) {
        const int64_t tall = (int64_t)((m + 127) / 128) * ((n + 63) / 64) * batch;
        if (tall >= 1024)
            hipLaunchKernelGGL((k_update<T, 128, 64, 4>), dim3((m + 127) / 128, (n + 63) / 64, batch), dim3(256), 0, st,
                               C, ldc, A, lda, B, ldb, m, n, kb, sb);
        else
            hipLaunchKernelGGL((k_update<T, 64, 64, 2>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, st, C,
                               ldc, A, lda, B, ldb, m, n, kb, sb);
    } else {
        hipLaunchKernelGGL((k_update<T, 64, 64, 4>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, st, C, ldc,
                           A, lda, B, ldb, m, n, kb, sb);
    }
}

template <typename T, int NT, int Q, int IB>
int launch_panel_reg(T* a, int64_t lda, int n, int k0, int kb, int32_t* ipiv, int32_t* info, int32_t* perm,
                     hipStream_t st, int batch, int64_t sb) {
    hipLaunchKernelGGL((k_panel_reg<T, NT, Q, IB>), dim3(1, 1, batch), dim3(NT), 0, st, a, lda, n, k0, kb, ipiv, info,
                       perm, sb);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// 256 threads (one wave per SIMD: no wave waits for issue slots, up to 512 registers per lane), rows
// per thread Q and sub-panel width IB per panel height, chosen so that the sub-panel (Q x IB values
// per thread) and its L-block buffers stay in registers.
template <typename T>
int panel_fast(T* a, int64_t lda, int n, int k0, int kb, int32_t* ipiv, int32_t* info, int32_t* perm, hipStream_t st,
               int batch, int64_t sb) {
    const int m = n - k0;
    if (m <= 256) return launch_panel_reg<T, 256, 1, 8>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    if (m <= 512) return launch_panel_reg<T, 256, 2, 8>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    if (m <= 1024) return launch_panel_reg<T, 256, 4, 8>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    if (m <= 2048) return launch_panel_reg<T, 256, 8, 8>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    if constexpr (sizeof(T) == 8) {
        if (m <= 4096) return launch_panel_reg<T, 256, 16, 4>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
        return launch_panel_reg<T, 256, 32, 2>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    } else {
        if (m <= 4096) return launch_panel_reg<T, 256, 16, 4>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
        return launch_panel_reg<T, 256, 32, 4>(a, lda, n, k0, kb, ipiv, info, perm, st, batch, sb);
    }
}

constexpr int kFastPanelRows = 8192;  // taller panels take the column-by-column panel
constexpr int kWidePanelBatch = 8;     // batches of at least this many matrices use 64-column panels

// ---- getrs ------------------------------------------------------------------------------------
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
__global__ void k_zero_info_batch(int32_t* info, int batch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < batch) info[i] = 0;
}

// Work-groups for a cooperative panel of m rows, or 0 when the panel stays in one work-group (short
// panels, batches whose groups would not all fit in kCoopMaxGroups, or panels taller than the
// groups can hold).
template <typename T>
int coop_groups(int m, int batch) {
    if (m <= kCoopMinRows) return 0;
    const int per_group = coop_threads<T>();
    const int G = (m + per_group - 1) / per_group;
    if (G < 2 || G * batch > kCoopMaxGroups) return 0;
    return G;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

constexpr size_t kCoopLdsBytes = 128 * 1024;  // NT rows x NB columns (coop_threads)

std::atomic<unsigned> g_coop_spin_limit{kCoopSpinLimit};

// info[z] = kCoopStalled for every matrix a cooperative panel's wait gave up on (by any group)
__global__ void k_coop_fail_info(const int32_t* __restrict__ fail, int32_t* __restrict__ info, int batch) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z < batch && fail[z] != 0) info[z] = kCoopStalled;
}

template <typename T>
int panel_coop(T* a, int64_t lda, int n, int k0, int kb, int G, int32_t* ipiv, int32_t* info, int32_t* perm,
               hipStream_t st, int batch, int64_t sb, int32_t* ctr, T* cval, int32_t* crow, T* xrow, int32_t* ostate,
               int32_t* fail) {
    constexpr int NT = coop_threads<T>();
    static const hipError_t lds_attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_panel_coop<T, NT>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           (int)kCoopLdsBytes);
    TVLM_HIP(lds_attr);
    TVLM_REQUIRE(G >= 2 && G * batch <= kCoopMaxGroups && kb <= NB && (int64_t)G * NT >= (int64_t)(n - k0),
                 "cooperative panel: %d groups for %d rows", G, n - k0);
    TVLM_REQUIRE(ctr && cval && crow && xrow && ostate && fail, "cooperative panel without exchange buffers");
    hipLaunchKernelGGL((k_panel_coop<T, NT>), dim3(G, 1, batch), dim3(NT), kCoopLdsBytes, st, a, lda, n, k0, kb, G,
                       ipiv, info, perm, sb, ctr, cval, crow, xrow, ostate, g_coop_spin_limit.load(), fail);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// shape[2] > 1 factors a batch of matrices, matrix b at ptr + b * stride[2]; ipiv holds n entries
// per matrix (matrix b at ipiv + b n), info one per matrix.
// Side stream and events of the batched look-ahead, one set per (device, caller stream), created
// on first use (default priority).  Per caller stream, so that factorisations running
// concurrently on different streams each have their own side chain.  The side stream is created
// with an all-CUs mask: such a stream gets a hardware queue of its own, so its updates can overlap
// the caller's panels whatever queue the caller's stream shares (streams without a mask share a
// small pool of queues, and two streams on one queue run one after the other).
struct LookaheadStreams {
    hipStream_t side = nullptr;
    hipEvent_t narrow_done = nullptr, rest_done = nullptr;
};

std::mutex g_lookahead_mu;
std::map<std::pair<int, hipStream_t>, LookaheadStreams> g_lookahead;

int lookahead_streams(hipStream_t caller, LookaheadStreams** out) {
    int dev = 0;
    TVLM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_lookahead_mu);
    LookaheadStreams& L = g_lookahead[std::make_pair(dev, caller)];
    if (!L.side) {
        TVLM_TRY(create_queue_stream(&L.side));
        TVLM_HIP(hipEventCreateWithFlags(&L.narrow_done, hipEventDisableTiming));
        TVLM_HIP(hipEventCreateWithFlags(&L.rest_done, hipEventDisableTiming));
    }
    *out = &L;
    return TVLM_OK;
}

// Batched getrf with a one-panel look-ahead (batches of kWidePanelBatch or more, n <= kFastPanelRows,
// 64-column panels).  After panel k's interchanges and U12 (k_swap_trsm), the trailing update is
// split: the next panel's 64 columns are updated first on the caller's stream, which then factors
// panel k+1 at once, while the rest of the trailing update runs on a side stream; the next
// k_swap_trsm waits for both.  The panels (one work-group per matrix) thus run beside the
// bandwidth-bound updates instead of between them.  Same arithmetic as the plain loop.
// Batched look-ahead factorisation.  Each step starts from a factored block [k0, j1) (one or two
// panels) whose U12 is known on every column right of it, and the trailing matrix not yet updated by
// that block.  The caller's stream updates the step's own columns first and factors them at once,
// while the rest of the trailing update runs on a side stream; what touches the rest (interchanges,
// U12) waits for it.  While the batch's trailing matrices exceed the last-level cache
// (kLookaheadPairBytes) a step takes two panels A and B: A, A's interchanges and U12 on B's columns
// only, B's columns updated by A, B, B's interchanges on A's columns -- all inside the step's 128
// columns, beside the side stream -- then, after it, A's and B's interchanges on the columns left
// and right of the step (in that order), A's U12 on the right, B's rows right of the step updated by
// A, B's U12.  The next trailing update is then one of rank 2 NB: the trailing matrices are read and
// written once per two panels.  perm holds two panels' composed interchanges (A's, then B's).
constexpr double kLookaheadPairBytes = 256.0 * 1024 * 1024;

template <typename T>
int getrf_lookahead(T* a, int64_t lda, int n, int batch, int64_t sb, int32_t* ipiv, int32_t* info, int32_t* perm,
                    hipStream_t st) {
    LookaheadStreams* L = nullptr;
    TVLM_TRY(lookahead_streams(st, &L));
    int32_t* perm2 = perm + (size_t)batch * kPermInts;
    // interchanges (mode 0/1) and U12 (mode 0/2) of the panel at kp on nleft blocks of 64 columns
    // from lbase and nright blocks from rbase
    auto swaps = [&](int kp, int kb, const int32_t* pm, int mode, int lbase, int nleft, int rbase, int nright) -> int {
        if (nleft + nright > 0) {
            hipLaunchKernelGGL(k_swap_trsm<T>, dim3(nleft + nright, 1, batch), dim3(256), 0, st, a, lda, n, kp, kb,
                               nleft, pm, sb, mode, 0, nullptr, lbase, rbase);
            TVLM_LAUNCHED();
        }
        return TVLM_OK;
    };
    auto blocks = [](int cols) { return cols > 0 ? (cols + 63) / 64 : 0; };
    auto at = [&](int r, int c) { return a + r + (int64_t)c * lda; };
    int k0 = 0, j1 = n < NB ? n : NB;
    TVLM_TRY(panel_fast<T>(a, lda, n, k0, j1, ipiv, info, perm, st, batch, sb));
    TVLM_TRY(swaps(k0, j1, perm, 0, 0, 0, j1, blocks(n - j1)));
    while (j1 < n) {
        const int kr = j1 - k0, m = n - j1;  // rank of the pending update; trailing size
        const double tail = (double)batch * (double)m * (double)m * sizeof(T);
        const bool pair = m > 2 * NB && tail > kLookaheadPairBytes;
        const int w = pair ? 2 * NB : (m < NB ? m : NB), jw = j1 + w;  // the step's columns [j1, jw)
        launch_update<T>(at(j1, j1), lda, at(j1, k0), lda, at(k0, j1), lda, m, w, kr, sb, batch, st);
        TVLM_LAUNCHED();
        const bool rest = m > w;
        if (rest) {
            TVLM_HIP(hipEventRecord(L->narrow_done, st));
            TVLM_HIP(hipStreamWaitEvent(L->side, L->narrow_done, 0));
            launch_update<T>(at(j1, jw), lda, at(j1, k0), lda, at(k0, jw), lda, m, m - w, kr, sb, batch, L->side);
            TVLM_LAUNCHED();
            TVLM_HIP(hipEventRecord(L->rest_done, L->side));
        }
        if (!pair) {
            TVLM_TRY(panel_fast<T>(a, lda, n, j1, w, ipiv, info, perm, st, batch, sb));
            if (rest) TVLM_HIP(hipStreamWaitEvent(st, L->rest_done, 0));
            TVLM_TRY(swaps(j1, w, perm, 0, 0, blocks(j1), jw, blocks(n - jw)));
            k0 = j1;
            j1 = jw;
            continue;
        }
        const int jA = j1 + NB;  // panel A = [j1, jA), panel B = [jA, jw)
        TVLM_TRY(panel_fast<T>(a, lda, n, j1, NB, ipiv, info, perm, st, batch, sb));
        TVLM_TRY(swaps(j1, NB, perm, 0, 0, 0, jA, 1));
        launch_update<T>(at(jA, jA), lda, at(jA, j1), lda, at(j1, jA), lda, n - jA, NB, NB, sb, batch, st);
        TVLM_LAUNCHED();
        TVLM_TRY(panel_fast<T>(a, lda, n, jA, NB, ipiv, info, perm2, st, batch, sb));
        TVLM_TRY(swaps(jA, NB, perm2, 1, j1, 1, 0, 0));
        if (rest) TVLM_HIP(hipStreamWaitEvent(st, L->rest_done, 0));
        TVLM_TRY(swaps(j1, NB, perm, 1, 0, blocks(j1), jw, blocks(n - jw)));
        TVLM_TRY(swaps(jA, NB, perm2, 1, 0, blocks(j1), jw, blocks(n - jw)));
        if (jw < n) {
            TVLM_TRY(swaps(j1, NB, perm, 2, 0, 0, jw, blocks(n - jw)));
            launch_update<T>(at(jA, jw), lda, at(jA, j1), lda, at(j1, jw), lda, NB, n - jw, NB, sb, batch, st);
            TVLM_LAUNCHED();
            TVLM_TRY(swaps(jA, NB, perm2, 2, 0, 0, jw, blocks(n - jw)));
        }
        k0 = j1;
        j1 = jw;
    }
    return TVLM_OK;
}

// ---- deferred interchanges of the L columns (cooperative factorisations) ------------------------
//
// LAPACK's getrf applies every panel's row interchanges to all columns left of the panel.  The
// cooperative factorisation applies them only to the columns that a later step still reads (the
// first panel of a pair takes the second's).  Every other column block P (columns [NB P, NB P + NB))
// keeps a gather list g_P of n rows instead, identity at the start: each later panel's composed
// interchanges (dst <- src pairs) update the list, g[dst] = g[src] (a few hundred integers, in
// k_swap_trsm), instead of two scattered rows per pair in each of the block's columns.  At the end
// k_defer_gather (one work-group per column) stages rows [NB (P + 1), n) of its column in LDS and
// writes row r back from row g_P[r]: one contiguous pass over L.  The lists only move rows at or
// below NB (P + 1), so g_P[r] stays in [NB (P + 1), n) there.
constexpr size_t kDeferLdsBytes = 128 * 1024;  // rows [NB (P + 1), n) of one column

__global__ void k_iota_lists(int32_t* __restrict__ g, int n, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) g[i] = (int32_t)(i % n);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_defer_gather(T* __restrict__ a, int64_t lda, int n,
                                                       const int32_t* __restrict__ gtab, int nblocks, int64_t sb) {
    extern __shared__ unsigned char s_raw[];
    T* s_col = reinterpret_cast<T*>(s_raw);
    const int c = blockIdx.x, P = c / NB, t = threadIdx.x;
    const int from = NB * (P + 1);
    if (from >= n) return;
    a += (int64_t)blockIdx.z * sb + (int64_t)c * lda;
    gtab += ((int64_t)blockIdx.z * nblocks + P) * n;
    const int m = n - from;
#pragma unroll 4
    for (int r = t; r < m; r += 1024) s_col[r] = a[from + r];
    __syncthreads();
#pragma unroll 4
    for (int r = t; r < m; r += 1024) {
        const int src = gtab[from + r] - from;
        a[from + r] = s_col[(unsigned)src < (unsigned)m ? src : r];
    }
}

template <typename T>
int defer_gather(T* a, int64_t lda, int n, int batch, int64_t sb, const int32_t* gtab, hipStream_t st) {
    const int nblocks = (n + NB - 1) / NB;
    if (nblocks < 2) return TVLM_OK;
    TVLM_REQUIRE((size_t)n * sizeof(T) <= kDeferLdsBytes, "deferred interchanges: n = %d too large", n);
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_defer_gather<T>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)kDeferLdsBytes);
    TVLM_HIP(attr);
    hipLaunchKernelGGL(k_defer_gather<T>, dim3(n, 1, batch), dim3(1024), sizeof(T) * (size_t)n, st, a, lda, n, gtab,
                       nblocks, sb);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// Factorisations whose tall panels are cooperative (all panels NB wide), taken two panels at a
// time so that the trailing matrix is read and written once per pair: panel A at kp, its
// interchanges and U12 (k_swap_trsm, mode 0); panel B's columns updated by panel A, panel B; panel
// B's interchanges on both sides (mode 1); the rows of panel B right of it updated by panel A (one
// NB-row GEMM); panel B's U12 (mode 2); one rank-2NB update of the trailing matrix.  The same
// products as one panel at a time, summed in a different order.
constexpr int kPairMinRows = 6144;  // trailing matrices at least this tall are updated once per pair

// ---- the factorisation plan -------------------------------------------------------------------
// Which path a getrf call takes and, for the cooperative path, every panel's form and exchange
// buffers; computed on the host before anything is queued (getrf_plan), executed by getrf, and
// exported as tvlm_getrf_plan so that the schedule and the buffer sizing can be checked without a GPU.
//   kPathLookahead : batches of >= kWidePanelBatch matrices of <= kFastPanelRows rows (one-group panels
//                    beside the trailing updates on a side stream);
//   kPathCoopPairs : the first panel is cooperative; then every panel of more than kCoopMinRows rows is
//                    (coop_groups falls with the rows) and takes its own counter set;
//   kPathPanels    : one work-group per panel throughout (k_panel_reg, or the column-by-column k_panel
//                    above kFastPanelRows rows) -- also every factorisation with cooperative panels
//                    switched off.
enum GetrfPath { kPathLookahead = 0, kPathCoopPairs = 1, kPathPanels = 2 };

struct GetrfStep {
    int kp, kb;   // first column and width of the panel
    int groups;   // work-groups of a cooperative panel, 0 for a one-group panel
    int ctr_set;  // counter set of a cooperative panel, -1 otherwise
};

struct GetrfPlan {
    int path = kPathPanels;
    std::vector<GetrfStep> steps;  // kPathCoopPairs only
    int ctr_sets = 0;              // counter sets allocated (one per cooperative panel)
    // exchange buffers (kPathCoopPairs): ints (counters, per-matrix attempt state and failure word,
    // candidate rows, gather lists) and reals (candidate values, published rows and U11)
    size_t nctr = 0, ncand = 0, nrow = 0, ngat = 0;
    size_t coop_bytes(size_t real_bytes) const {
        return nctr || ncand ? sizeof(int32_t) * (nctr + ncand + ngat) + real_bytes * (ncand + nrow) + 512 : 0;
    }
};

template <typename T>
GetrfPlan getrf_plan(int n, int batch, bool allow_coop) {
    GetrfPlan P;
    if (batch >= kWidePanelBatch && n <= kFastPanelRows) {
        P.path = kPathLookahead;
        return P;
    }
    if (!allow_coop || coop_groups<T>(n, batch) == 0) {
        P.path = kPathPanels;
        return P;
    }
    P.path = kPathCoopPairs;
    for (int kp = 0; kp < n; kp += NB) {
        GetrfStep st{kp, n - kp < NB ? n - kp : NB, coop_groups<T>(n - kp, batch), -1};
        if (st.groups > 0) st.ctr_set = P.ctr_sets++;
        P.steps.push_back(st);
    }
    const int panels = (n + NB - 1) / NB;
    P.nctr = (size_t)P.ctr_sets * batch * kCoopCtr + 2 * (size_t)batch;
    P.ncand = (size_t)batch * NB * kCoopMaxGroups;
    P.nrow = (size_t)batch * kCoopXrow;
    P.ngat = (size_t)batch * panels * n;
    return P;
}

template <typename T>
int getrf_coop_pairs(T* a, int64_t lda, int n, int batch, int64_t sb, int32_t* ipiv, int32_t* info, int32_t* perm,
                     hipStream_t st, const GetrfPlan& plan, int32_t* ctr, T* cval, int32_t* crow, T* xrow,
                     int32_t* gtab, int32_t* ostate, int32_t* fail) {
    if (n > NB) {
        const int64_t total = (int64_t)batch * ((n + NB - 1) / NB) * n;
        hipLaunchKernelGGL(k_iota_lists, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, gtab, n, total);
        TVLM_LAUNCHED();
    }
    auto panel = [&](int kp, int kb) -> int {
        // the planned panel at kp: its width, group count and exchange counter set
        const int s = kp / NB;
        TVLM_REQUIRE(kp % NB == 0 && s < (int)plan.steps.size() && plan.steps[s].kp == kp && plan.steps[s].kb == kb,
                     "panel at column %d (width %d) is not in the factorisation plan", kp, kb);
        const GetrfStep& p = plan.steps[s];
        if (p.groups == 0) return panel_fast<T>(a, lda, n, kp, kb, ipiv, info, perm, st, batch, sb);
        TVLM_REQUIRE(p.ctr_set >= 0 && p.ctr_set < plan.ctr_sets, "cooperative panel %d: counter set %d of %d", s,
                     p.ctr_set, plan.ctr_sets);
        return panel_coop<T>(a, lda, n, kp, kb, p.groups, ipiv, info, perm, st, batch, sb,
                             ctr + (size_t)p.ctr_set * batch * kCoopCtr, cval, crow, xrow, ostate, fail);
    };
    // the left columns before left0 take the interchanges into their gather lists
    auto swap_trsm = [&](int kp, int kb, int mode, int left0) -> int {
        const int j1 = kp + kb;
        const int nleft = mode == 2 ? 0 : (kp + 63) / 64, nright = (n - j1 + 63) / 64;
        if (nleft + nright > 0) {
            hipLaunchKernelGGL(k_swap_trsm<T>, dim3(nleft + nright, 1, batch), dim3(256), 0, st, a, lda, n, kp, kb,
                               nleft, perm, sb, mode, left0, gtab);
            TVLM_LAUNCHED();
        }
        return TVLM_OK;
    };
    auto at = [&](int r, int c) { return a + r + (int64_t)c * lda; };
    for (int kp = 0; kp < n;) {
        const int kbA = n - kp < NB ? n - kp : NB, jA = kp + kbA;
        TVLM_TRY(panel(kp, kbA));
        TVLM_TRY(swap_trsm(kp, kbA, 0, kp));
        if (jA >= n) break;
        if (n - jA < kPairMinRows) {  // the trailing matrix stays in the last-level cache: one panel at a time
            launch_update<T>(at(jA, jA), lda, at(jA, kp), lda, at(kp, jA), lda, n - jA, n - jA, kbA, sb, batch, st);
            TVLM_LAUNCHED();
            kp = jA;
            continue;
        }
        const int kbB = n - jA < NB ? n - jA : NB, jB = jA + kbB;
        launch_update<T>(at(jA, jA), lda, at(jA, kp), lda, at(kp, jA), lda, n - jA, kbB, kbA, sb, batch, st);
        TVLM_LAUNCHED();
        TVLM_TRY(panel(jA, kbB));
        TVLM_TRY(swap_trsm(jA, kbB, 1, kp));
        if (jB >= n) break;
        launch_update<T>(at(jA, jB), lda, at(jA, kp), lda, at(kp, jB), lda, kbB, n - jB, kbA, sb, batch, st);
        TVLM_LAUNCHED();
        TVLM_TRY(swap_trsm(jA, kbB, 2, jA));
        launch_update<T>(at(jB, jB), lda, at(jB, kp), lda, at(kp, jB), lda, n - jB, n - jB, kbA + kbB, sb, batch, st);
        TVLM_LAUNCHED();
        kp = jB;
    }
    TVLM_TRY(defer_gather<T>(a, lda, n, batch, sb, gtab, st));
    hipLaunchKernelGGL(k_coop_fail_info, dim3((batch + 255) / 256), dim3(256), 0, st, fail, info, batch);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

template <typename T>
int getrf(const tvlm_view* av, int32_t* ipiv, int32_t* info, void* stream, bool allow_coop = true) {
    TVLM_REQUIRE(av && av->ptr && ipiv && info, "null argument");
    TVLM_REQUIRE(av->stride[0] == 1, "A must be column-major");
    TVLM_REQUIRE(av->shape[0] == av->shape[1], "A must be square (got %lld x %lld)", (long long)av->shape[0],
                 (long long)av->shape[1]);
    const int n = (int)av->shape[0];
    const int batch = (int)av->shape[2];
    const int64_t lda = av->stride[1], sb = av->stride[2];
    TVLM_REQUIRE(lda >= n || n == 0, "lda < n");
    TVLM_REQUIRE(batch >= 1 && batch <= 65535, "batch %d out of range", batch);
    TVLM_REQUIRE(batch == 1 || sb >= lda * n, "matrices of the batch overlap (stride[2] < stride[1] * n)");
    hipStream_t st = as_stream(stream);
    T* a = static_cast<T*>(av->ptr);
    hipLaunchKernelGGL(k_zero_info_batch<T>, dim3((batch + 255) / 256), dim3(256), 0, st, info, batch);
    TVLM_LAUNCHED();
    void* ws = nullptr;
    TVLM_TRY(device_workspace(sizeof(int32_t) * kPermInts * 2 * (size_t)batch, &ws, kWsLuPerm, st));
    int32_t* perm = static_cast<int32_t*>(ws);
    const GetrfPlan plan = getrf_plan<T>(n, batch, allow_coop);
    if (plan.path == kPathLookahead) return getrf_lookahead<T>(a, lda, n, batch, sb, ipiv, info, perm, st);
    if (plan.path == kPathCoopPairs) {
        // Buffers (sized by the plan): per cooperative panel and matrix kCoopCtr counters, per matrix
        // the attempt counter and the failure word (all zeroed here, once for the whole
        // factorisation), the candidates, the published rows and U11, and the deferred
        // interchanges' gather lists.
        void* cws = nullptr;
        TVLM_TRY(device_workspace(plan.coop_bytes(sizeof(T)), &cws, kWsLuCoop, st));
        int32_t* ctr = static_cast<int32_t*>(cws);
        int32_t* crow = ctr + plan.nctr;
        T* cval = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(crow + plan.ncand) + 255) & ~(uintptr_t)255);
        T* xrow = cval + plan.ncand;
        int32_t* gtab =
            reinterpret_cast<int32_t*>((reinterpret_cast<uintptr_t>(xrow + plan.nrow) + 255) & ~(uintptr_t)255);
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((plan.nctr + 255) / 256)), dim3(256), 0, st, ctr,
                           (int64_t)plan.nctr);
        TVLM_LAUNCHED();
        int32_t* ostate = ctr + (size_t)plan.ctr_sets * batch * kCoopCtr;  // per matrix: failed attempts in a row
        int32_t* fail = ostate + batch;                                     // per matrix: a wait gave up
        return getrf_coop_pairs<T>(a, lda, n, batch, sb, ipiv, info, perm, st, plan, ctr, cval, crow, xrow, gtab,
                                   ostate, fail);
    }
    for (int k0 = 0; k0 < n;) {
        // tall panels of one matrix are narrower: the in-panel left-looking traffic grows with width
        // squared.  A batch is bound by the trailing updates instead, whose traffic (the whole
        // trailing matrix read and written per panel) halves with the full width.
        const int width = (n - k0 > 2048 && batch < kWidePanelBatch) ? NB / 2 : NB;
        const int kb = (n - k0 < width) ? n - k0 : width;
        const int j1 = k0 + kb;
        const int kp = k0;  // this panel's first column
        k0 = j1;
        if (n - kp <= kFastPanelRows) {
            TVLM_TRY(panel_fast<T>(a, lda, n, kp, kb, ipiv, info, perm, st, batch, sb));
            const int nleft = (kp + 63) / 64, nright = (n - j1 + 63) / 64;
            if (nleft + nright > 0) {
                hipLaunchKernelGGL(k_swap_trsm<T>, dim3(nleft + nright, 1, batch), dim3(256), 0, st, a, lda, n, kp, kb,
                                   nleft, perm, sb);
                TVLM_LAUNCHED();
            }
        } else {
            for (int z = 0; z < batch; z++) {
                T* az = a + (int64_t)z * sb;
                int32_t* pz = ipiv + (int64_t)z * n;
                hipLaunchKernelGGL(k_panel<T>, dim3(1), dim3(kPanelThreads), 0, st, az, lda, n, kp, kb, pz, info + z);
                TVLM_LAUNCHED();
                if (kp > 0) {
                    hipLaunchKernelGGL(k_laswp<T>, dim3((kp + 255) / 256), dim3(256), 0, st, az, lda, 0, kp, kp, kb, pz);
                    TVLM_LAUNCHED();
                }
                if (j1 < n) {
                    hipLaunchKernelGGL(k_laswp<T>, dim3((n - j1 + 255) / 256), dim3(256), 0, st, az, lda, j1, n, kp, kb,
                                       pz);
                    TVLM_LAUNCHED();
                    hipLaunchKernelGGL(k_trsm<T>, dim3((n - j1 + 255) / 256), dim3(256), 0, st, az, lda, n, kp, kb, j1);
                    TVLM_LAUNCHED();
                }
            }
        }
        if (j1 >= n) continue;
        const int m = n - j1;
        launch_update<T>(a + j1 + (int64_t)j1 * lda, lda, a + j1 + (int64_t)kp * lda, lda, a + kp + (int64_t)j1 * lda,
                         lda, m, m, kb, sb, batch, st);
        TVLM_LAUNCHED();
    }
    return TVLM_OK;
}

// ---- blocked getrs: interchange lists + diagonal-block inverses + one work-group per rhs -------
//
// k_solve_prep, y = 0: composes the interchanges of pivot block b (64 positions) into (dst, src)
//   pairs (one wave); y = 1 / 2: inverts the 64 x 64 diagonal block b of L (unit lower) / U (upper),
//   one column per lane, each row's dot product split over 4 waves.
// k_solve_blocked: the rhs column lives in LDS; the interchange lists are applied block after
//   block; then x_b = Linv_b (x_b - L_{b,<b} x_{<b}) for b ascending and x_b = Uinv_b (x_b -
//   U_{b,>b} x_{>b}) for b descending, each off-diagonal product split over 16 thread groups
//   reading 64 consecutive rows per column, partial sums reduced in LDS.
constexpr int kSB = 64;            // solve block
constexpr int kSolveMaxN = 16384;  // rhs column in LDS
constexpr int kSolvePairs = 1 + 2 * 2 * kSB;

template <typename T>
__global__ __launch_bounds__(256) void k_solve_prep(const T* __restrict__ a, int64_t lda, int n,
                                                    const int32_t* __restrict__ ipiv, int32_t* __restrict__ pairs,
                                                    T* __restrict__ inv, int64_t sa) {
    __shared__ int s_p[kSB];
    __shared__ int s_pd[2 * kSB], s_ps[2 * kSB];
    __shared__ T s_a[kSB][kSB + 1];
    __shared__ T s_x[kSB][kSB + 1];  // the inverse, row by row
    __shared__ T s_red[4][kSB];
    const int nb_ = (n + kSB - 1) / kSB;  // batch: one matrix per blockIdx.z
    a += (int64_t)blockIdx.z * sa;
    ipiv += (int64_t)blockIdx.z * n;
    pairs += (int64_t)blockIdx.z * nb_ * kSolvePairs;
    inv += (int64_t)blockIdx.z * 2 * nb_ * kSB * kSB;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, g = t >> 6;
    const int r0 = b * kSB, rb = min(kSB, n - r0);
    if (blockIdx.y == 0) {
        if (t < kSB && lane < rb) s_p[lane] = ipiv[r0 + lane] - 1;
        __syncthreads();
        if (t < kSB) {
            const int np = wave_compose_swaps(r0, rb, s_p, s_pd, s_ps);
            int32_t* out = pairs + (int64_t)b * kSolvePairs;
            for (int e = lane; e < np; e += 64) {
                out[1 + 2 * e] = s_pd[e];
                out[2 + 2 * e] = s_ps[e];
            }
            if (lane == 0) out[0] = np;
        }
        return;
    }
    const bool lower = blockIdx.y == 1;
    for (int e = t; e < kSB * kSB; e += 256) {
        const int r = e % kSB, c = e / kSB;
        s_a[r][c] = (r < rb && c < rb) ? a[(r0 + r) + (int64_t)(r0 + c) * lda] : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    // Column j of the inverse belongs to lane j; row i of it is formed from the rows already done,
    // the dot product over them split over the 4 waves (wave g takes p = g, g + 4, ...) and summed
    // in a fixed order.  Unit lower: X_ij = delta_ij - sum_{j <= p < i} L_ip X_pj, rows ascending;
    // upper: X_ij = (delta_ij - sum_{i < p <= j} U_ip X_pj) / U_ii, rows descending.
    const int j = lane;
    for (int k = 0; k < kSB; k++) {
        const int i = lower ? k : kSB - 1 - k;
        T part = T(0);
        if (lower) {
            for (int p = j + ((g - j) % 4 + 4) % 4; p < i; p += 4) part += s_a[i][p] * s_x[p][j];
        } else {
            for (int p = i + 1 + g; p <= j; p += 4) part += s_a[i][p] * s_x[p][j];
        }
        s_red[g][j] = part;
        __syncthreads();
        if (g == 0) {
            const T sum = (s_red[0][j] + s_red[1][j]) + (s_red[2][j] + s_red[3][j]);
            T v;
            if (lower)
                v = i < j ? T(0) : (i == j ? T(1) : -sum);
            else
                v = i > j ? T(0) : ((i == j ? T(1) : T(0)) - sum) / s_a[i][i];
            s_x[i][j] = v;
        }
        __syncthreads();
    }
    T* out = inv + ((int64_t)(lower ? 0 : 1) * ((n + kSB - 1) / kSB) + b) * kSB * kSB;
    for (int e = t; e < kSB * kSB; e += 256) {
        const int r = e % kSB, c = e / kSB;
        out[r + (int64_t)c * kSB] = s_x[r][c];
    }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_solve_blocked(const T* __restrict__ a, int64_t lda, int n,
                                                        const int32_t* __restrict__ pairs, const T* __restrict__ inv,
                                                        T* __restrict__ x, int64_t ldx, int64_t sa, int64_t sx) {
    __shared__ T xs[kSolveMaxN];
    __shared__ T red[16][kSB];
    __shared__ T yb[kSB];
    const int t = threadIdx.x, r = t & 63, g = t >> 6;
    const int nblk = (n + kSB - 1) / kSB;
    a += (int64_t)blockIdx.z * sa;  // batch: one matrix per blockIdx.z
    pairs += (int64_t)blockIdx.z * nblk * kSolvePairs;
    inv += (int64_t)blockIdx.z * 2 * nblk * kSB * kSB;
    x += (int64_t)blockIdx.z * sx;
    T* xc = x + (int64_t)blockIdx.x * ldx;
    for (int i = t; i < n; i += 1024) xs[i] = xc[i];
    __syncthreads();
    for (int b = 0; b < nblk; b++) {  // P b: interchange lists in pivot order
        const int32_t* pl = pairs + (int64_t)b * kSolvePairs;
        const int np = pl[0];
        T v = T(0);
        if (t < np) v = xs[pl[2 + 2 * t]];
        __syncthreads();
        if (t < np) xs[pl[1 + 2 * t]] = v;
        __syncthreads();
    }
    const T* linv = inv;
    const T* uinv = inv + (int64_t)nblk * kSB * kSB;
    // L y = P b
    for (int b = 0; b < nblk; b++) {
        const int r0 = b * kSB, rb = min(kSB, n - r0);
        T acc = T(0);
        if (r < rb)
            for (int j = g; j < r0; j += 16) acc += a[(r0 + r) + (int64_t)j * lda] * xs[j];
        red[g][r] = acc;
        __syncthreads();
        if (t < kSB) {
            T y = t < rb ? xs[r0 + t] : T(0);
#pragma unroll
            for (int q = 0; q < 16; q++) y -= red[q][t];
            yb[t] = y;
        }
        __syncthreads();
        T s = T(0);
        const T* lb = linv + (int64_t)b * kSB * kSB;
#pragma unroll
        for (int p = 4 * g; p < 4 * g + 4; p++) s += lb[r + p * kSB] * yb[p];
        red[g][r] = s;
        __syncthreads();
        if (t < rb) {
            T v = T(0);
#pragma unroll
            for (int q = 0; q < 16; q++) v += red[q][t];
            xs[r0 + t] = v;
        }
        __syncthreads();
    }
    // U x = y
    for (int b = nblk - 1; b >= 0; b--) {
        const int r0 = b * kSB, rb = min(kSB, n - r0);
        T acc = T(0);
        if (r < rb)
            for (int j = r0 + kSB + g; j < n; j += 16) acc += a[(r0 + r) + (int64_t)j * lda] * xs[j];
        red[g][r] = acc;
        __syncthreads();
        if (t < kSB) {
            T y = t < rb ? xs[r0 + t] : T(0);
#pragma unroll
            for (int q = 0; q < 16; q++) y -= red[q][t];
            yb[t] = y;
        }
        __syncthreads();
        T s = T(0);
        const T* ub = uinv + (int64_t)b * kSB * kSB;
#pragma unroll
        for (int p = 4 * g; p < 4 * g + 4; p++) s += ub[r + p * kSB] * yb[p];
        red[g][r] = s;
        __syncthreads();
        if (t < rb) {
            T v = T(0);
#pragma unroll
            for (int q = 0; q < 16; q++) v += red[q][t];
            xs[r0 + t] = v;
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += 1024) xc[i] = xs[i];
}

// Multi-work-group form for large n and few right-hand sides (one work-group streaming the whole
// factor per right-hand side is bound by what one CU can read).  k_solve_perm applies the
// interchange lists with the column in LDS (one work-group per rhs); then one k_trsv_step launch per
// 64-row block and triangle.  Forward step b: work-group j takes row block r = b + 1 + j and
// subtracts L_{r,b} x_b (x_b final from the previous launch); the group of r = b + 1 then finishes
// x_r = Linv_r (its updated value).  Backward step b: row block r = b - 1 - j, U_{r,b} x_b, and the
// group of r = b - 1 finishes x_r = Uinv_r (...).  The first launch of each triangle (b = -1 forward,
// b = nblk backward) only finishes its first block.  Each off-diagonal block is read once per rhs:
// 64 consecutive rows per column by each wave, 4 waves splitting the 64 columns.
constexpr int kTrsvMinN = 1024;     // below this, one work-group per rhs is as fast
constexpr int kTrsvMaxRhs = 256;    // rhs x matrices; more right-hand sides fill the device anyway

template <typename T>
__global__ __launch_bounds__(1024) void k_solve_perm(int n, const int32_t* __restrict__ pairs, T* __restrict__ x,
                                                     int64_t ldx, int64_t sx) {
    __shared__ T xs[kSolveMaxN];
    const int t = threadIdx.x;
    const int nblk = (n + kSB - 1) / kSB;
    pairs += (int64_t)blockIdx.z * nblk * kSolvePairs;
    T* xc = x + (int64_t)blockIdx.z * sx + (int64_t)blockIdx.x * ldx;
    for (int i = t; i < n; i += 1024) xs[i] = xc[i];
    __syncthreads();
    for (int b = 0; b < nblk; b++) {
        const int32_t* pl = pairs + (int64_t)b * kSolvePairs;
        const int np = pl[0];
        T v = T(0);
        if (t < np) v = xs[pl[2 + 2 * t]];
        __syncthreads();
        if (t < np) xs[pl[1 + 2 * t]] = v;
        __syncthreads();
    }
    for (int i = t; i < n; i += 1024) xc[i] = xs[i];
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_step(const T* __restrict__ a, int64_t lda, int n,
                                                   const T* __restrict__ inv, T* __restrict__ x, int64_t ldx,
                                                   int64_t sa, int64_t sx, int b) {
    __shared__ T xb[kSB];
    __shared__ T red[4][kSB];
    __shared__ T yb[kSB];
    const int t = threadIdx.x, rl = t & 63, g = t >> 6;
    const int nblk = (n + kSB - 1) / kSB;
    a += (int64_t)blockIdx.z * sa;
    inv += (int64_t)blockIdx.z * 2 * nblk * kSB * kSB + (LOWER ? 0 : (int64_t)nblk * kSB * kSB);
    T* xc = x + (int64_t)blockIdx.z * sx + (int64_t)blockIdx.y * ldx;
    const int j = blockIdx.x;
    const int r = LOWER ? b + 1 + j : b - 1 - j;  // this group's row block
    const int r0 = r * kSB, c0 = b * kSB;
    const bool has_b = LOWER ? b >= 0 : b < nblk;
    T y = T(0);
    if (t < kSB) y = r0 + t < n ? xc[r0 + t] : T(0);
    if (has_b) {
        if (t < kSB) xb[t] = c0 + t < n ? xc[c0 + t] : T(0);
        __syncthreads();
        T acc = T(0);
        if (r0 + rl < n) {
            const T* ab = a + (r0 + rl) + (int64_t)c0 * lda;
#pragma unroll 4
            for (int c = 16 * g; c < 16 * g + 16; c++)
                if (c0 + c < n) acc += ab[(int64_t)c * lda] * xb[c];
        }
        red[g][rl] = acc;
        __syncthreads();
        if (t < kSB) y -= red[0][t] + red[1][t] + red[2][t] + red[3][t];
    }
    if (j != 0) {  // a later block: store the partial update
        if (t < kSB && r0 + t < n) xc[r0 + t] = y;
        return;
    }
    // the next block in order: finish it with its diagonal-block inverse
    if (t < kSB) yb[t] = y;
    __syncthreads();
    const T* ib = inv + (int64_t)r * kSB * kSB;
    T s = T(0);
#pragma unroll
    for (int p = 16 * g; p < 16 * g + 16; p++) s += ib[rl + p * kSB] * yb[p];
    red[g][rl] = s;
    __syncthreads();
    if (t < kSB && r0 + t < n) xc[r0 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// shape[2] > 1 solves a batch: matrix b at A.ptr + b A.stride[2] with pivots ipiv + b n, its
// right-hand sides at X.ptr + b X.stride[2].
template <typename T>
int getrs(const tvlm_view* av, const int32_t* ipiv, const tvlm_view* xv, void* stream) {
    TVLM_REQUIRE(av && xv && av->ptr && xv->ptr && ipiv, "null argument");
    TVLM_REQUIRE(av->stride[0] == 1 && xv->stride[0] == 1, "column-major operands required");
    const int n = (int)av->shape[0];
    TVLM_REQUIRE(av->shape[1] == n && xv->shape[0] == n, "shape mismatch");
    const int batch = (int)av->shape[2];
    TVLM_REQUIRE(xv->shape[2] == batch, "A has %d matrices, X %lld right-hand-side sets", batch,
                 (long long)xv->shape[2]);
    TVLM_REQUIRE(batch >= 1 && batch <= 65535, "batch %d out of range", batch);
    const int nrhs = (int)xv->shape[1];
    if (n == 0 || nrhs == 0) return TVLM_OK;
    const int64_t lda = av->stride[1], ldx = nrhs > 1 ? xv->stride[1] : n;
    const int64_t sa = av->stride[2], sx = xv->stride[2];
    hipStream_t st = as_stream(stream);
    const T* a = static_cast<const T*>(av->ptr);
    T* x = static_cast<T*>(xv->ptr);
    if (n <= kSolveMaxN && (int64_t)n * sizeof(T) <= 128 * 1024) {
        const int nblk = (n + kSB - 1) / kSB;
        void* ws = nullptr;
        const size_t inv_bytes = sizeof(T) * 2 * (size_t)nblk * kSB * kSB * batch;
        TVLM_TRY(device_workspace(inv_bytes + sizeof(int32_t) * (size_t)nblk * kSolvePairs * batch, &ws, kWsSolveInv, st));
        T* inv = static_cast<T*>(ws);
        int32_t* pairs = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + inv_bytes);
        hipLaunchKernelGGL(k_solve_prep<T>, dim3(nblk, 3, batch), dim3(256), 0, st, a, lda, n, ipiv, pairs, inv, sa);
        TVLM_LAUNCHED();
        if (n < kTrsvMinN || (int64_t)nrhs * batch > kTrsvMaxRhs) {
            hipLaunchKernelGGL(k_solve_blocked<T>, dim3(nrhs, 1, batch), dim3(1024), 0, st, a, lda, n, pairs, inv, x,
                               ldx, sa, sx);
            TVLM_LAUNCHED();
            return TVLM_OK;
        }
        hipLaunchKernelGGL(k_solve_perm<T>, dim3(nrhs, 1, batch), dim3(1024), 0, st, n, pairs, x, ldx, sx);
        TVLM_LAUNCHED();
        for (int b = -1; b < nblk - 1; b++) {
            const int groups = b < 0 ? 1 : nblk - 1 - b;
            hipLaunchKernelGGL((k_trsv_step<T, true>), dim3(groups, nrhs, batch), dim3(256), 0, st, a, lda, n, inv, x,
                               ldx, sa, sx, b);
            TVLM_LAUNCHED();
        }
        for (int b = nblk; b > 0; b--) {
            const int groups = b == nblk ? 1 : b;
            hipLaunchKernelGGL((k_trsv_step<T, false>), dim3(groups, nrhs, batch), dim3(256), 0, st, a, lda, n, inv, x,
                               ldx, sa, sx, b);
            TVLM_LAUNCHED();
        }
        return TVLM_OK;
    }
    for (int z = 0; z < batch; z++) {
        const T* az = a + (int64_t)z * sa;
        const int32_t* pz = ipiv + (int64_t)z * n;
        T* xz = x + (int64_t)z * sx;
        hipLaunchKernelGGL(k_apply_pivots<T>, dim3(nrhs), dim3(64), 0, st, xz, ldx, n, pz);
        TVLM_LAUNCHED();
        for (int r0 = 0; r0 < n; r0 += SB) {  // forward: L
            const int nb = (n - r0 < SB) ? n - r0 : SB;
            hipLaunchKernelGGL((k_diag_solve<T, true>), dim3(nrhs), dim3(SB), 0, st, az, lda, r0, nb, xz, ldx);
            TVLM_LAUNCHED();
            if (r0 + nb < n) {
                hipLaunchKernelGGL(k_off_update<T>, dim3((n - r0 - nb + 255) / 256, nrhs), dim3(256), 0, st, az, lda,
                                   r0 + nb, n, r0, nb, xz, ldx);
                TVLM_LAUNCHED();
            }
        }
        const int nblk = (n + SB - 1) / SB;
        for (int b = nblk - 1; b >= 0; b--) {  // backward: U
            const int r0 = b * SB;
            const int nb = (n - r0 < SB) ? n - r0 : SB;
            hipLaunchKernelGGL((k_diag_solve<T, false>), dim3(nrhs), dim3(SB), 0, st, az, lda, r0, nb, xz, ldx);
            TVLM_LAUNCHED();
            if (r0 > 0) {
                hipLaunchKernelGGL(k_off_update<T>, dim3((r0 + 255) / 256, nrhs), dim3(256), 0, st, az, lda, 0, r0, r0,
                                   nb, xz, ldx);
                TVLM_LAUNCHED();
            }
        }
    }
    return TVLM_OK;
}

}  // namespace

void release_lu_stream(hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    LookaheadStreams L;
    {
        std::lock_guard<std::mutex> lock(g_lookahead_mu);
        auto it = g_lookahead.find(std::make_pair(dev, stream));
        if (it == g_lookahead.end()) return;
        L = it->second;
        g_lookahead.erase(it);
    }
    if (L.side) {
        (void)hipStreamSynchronize(L.side);
        (void)hipStreamDestroy(L.side);
    }
    if (L.narrow_done) (void)hipEventDestroy(L.narrow_done);
    if (L.rest_done) (void)hipEventDestroy(L.rest_done);
}

int getrf_single_group_f64(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream) {
    return getrf<double>(a, ipiv, info, stream, false);
}
int getrf_single_group_f32(const tvlm_view* a, int32_t* ipiv, int32_t* info, void* stream) {
    return getrf<float>(a, ipiv, info, stream, false);
}

}  // namespace tvlm

extern "C" {
int tvlm_getrf_plan(int64_t n, int32_t batch, int32_t elem_bytes, int32_t* path, int32_t* n_steps, int32_t* groups,
                    int32_t* ctr_set, int32_t max_steps, int64_t* coop_ws_bytes) {
    TVLM_REQUIRE(path && n_steps, "null argument");
    TVLM_REQUIRE(n >= 0 && n <= (int64_t)1 << 30 && batch >= 1 && batch <= 65535, "n %lld / batch %d out of range",
                 (long long)n, batch);
    TVLM_REQUIRE(elem_bytes == 4 || elem_bytes == 8, "elem_bytes must be 4 or 8");
    const tvlm::GetrfPlan P = elem_bytes == 8 ? tvlm::getrf_plan<double>((int)n, batch, true)
                                              : tvlm::getrf_plan<float>((int)n, batch, true);
    *path = P.path;
    *n_steps = (int32_t)P.steps.size();
    if (coop_ws_bytes) *coop_ws_bytes = (int64_t)P.coop_bytes((size_t)elem_bytes);
    if (groups || ctr_set) {
        TVLM_REQUIRE((int64_t)P.steps.size() <= (int64_t)max_steps, "%zu panels, room for %d", P.steps.size(), max_steps);
        for (size_t k = 0; k < P.steps.size(); k++) {
            if (groups) groups[k] = P.steps[k].groups;
            if (ctr_set) ctr_set[k] = P.steps[k].ctr_set;
        }
    }
    return TVLM_OK;
}
int tvlm_lu_update(int32_t elem_bytes, void* c, int64_t ldc, const void* a, int64_t lda, const void* b, int64_t ldb,
                   int64_t m, int64_t n, int64_t k, int32_t batch, int64_t stride, void* stream) {
    TVLM_REQUIRE(elem_bytes == 4 || elem_bytes == 8, "elem_bytes must be 4 or 8");
    TVLM_REQUIRE(m >= 0 && n >= 0 && k >= 0 && m < (1 << 30) && n < (1 << 30) && k < (1 << 30), "m/n/k out of range");
    TVLM_REQUIRE(batch >= 1 && batch <= 65535 && (batch == 1 || stride >= 0), "batch %d / stride out of range", batch);
    TVLM_REQUIRE(ldc >= (m > 1 ? m : 1) && lda >= (m > 1 ? m : 1) && ldb >= (k > 1 ? k : 1) && ldc < (1 << 28) &&
                     lda < (1 << 28) && ldb < (1 << 28),
                 "leading dimensions out of range");
    if (m == 0 || n == 0 || k == 0) return TVLM_OK;
    TVLM_REQUIRE(c && a && b, "null matrix");
    const hipStream_t st = (hipStream_t)stream;
    if (elem_bytes == 8)
        tvlm::launch_update<double>((double*)c, ldc, (const double*)a, lda, (const double*)b, ldb, (int)m, (int)n, (int)k,
                                    stride, batch, st);
    else
        tvlm::launch_update<float>((float*)c, ldc, (const float*)a, lda, (const float*)b, ldb, (int)m, (int)n, (int)k,
                                   stride, batch, st);
    TVLM_LAUNCHED();
    return TVLM_OK;
}
int tvlm_lu_set_coop_spin_limit(uint32_t limit) {
    tvlm::g_coop_spin_limit.store(limit == 0 ? tvlm::kCoopSpinLimit : limit);
    return TVLM_OK;
}
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

// tvlm_steady.hip -- steady VLM solves of a batch of requests (the caller of the hot path,
// VLM::run, vlm/src/vlm.cpp:97-151, for many angles of attack at once).
//
// A plan owns the device buffers of `batch` requests on one mesh (any number of surfaces).  For
// request b, solve() repeats VLM::run's sequence: wing reset to the input mesh, wake row shed,
// displacement 100 chords along -freestream, second shed, panel metrics at alpha_b, influence matrix
// (iteration 1: the trailing-edge columns also see the one wake row), rhs = -(V_inf . n), solve,
// circulation differences, steady forces and per-surface CL/CM.
//   mode 0 (plain FP64): exact FP64 assembly, FP64 LU and substitution, one request after another;
//                        this is the reference's arithmetic, operation for operation.
//   mode 1 (served)    : every request's FP64 influence matrix in the seeded form, copied to FP32,
//                        one batched FP32 LU for all requests, then FP64 iterative refinement
//                        x += A32^{-1} (b - A64 x) (kRefine + 1 corrections, residual in FP64 against
//                        the FP64 matrix), so the answer carries the FP64 matrix's accuracy.  The
//                        returned x of every request is checked: ||b - A64 x||_inf <= sqrt(N) eps
//                        ||x||_inf ||A64||_inf (the stopping test of LAPACK's dsgesv).  A request that
//                        fails it gets more corrections (up to kMaxCorrections in all), then, if it
//                        still fails, the FP64 LU of its FP64 matrix (the plain path's solve); a request
//                        whose FP32 LU had a cooperative panel give up waiting (info -999) is first
//                        re-factored with one-group panels.  tvlm_vlm_refinement reports what each
//                        request took.
// Per-surface coefficients are combined on the host (area-weighted CL, area*MAC-weighted CM,
// vlm/src/vlm_backend.cpp:71-139).
#include <cmath>
#include <cstring>
#include <vector>

#include "tvlm_internal.hpp"

namespace {
constexpr int kMaxS = TVLM_MAX_SURFACES;
constexpr int kRefine = 3;
constexpr int kMaxCorrections = 8;  // corrections of one request before the FP64 fallback
// tvlm_vlm_refinement status bits
constexpr int32_t kStMoreCorrections = 1, kStFallback64 = 2, kStRefactored = 4, kStUnresolved = 8;
constexpr int kResidRows = 64;     // rows per work-group of the residual
constexpr int kResidWaves = 8;     // waves splitting its columns
}  // namespace

struct tvlm_vlm_plan {
    int32_t mode = 0, n_surf = 0, N = 0, batch = 0, active = 0;
    hipStream_t stream = nullptr;
    int64_t ns[kMaxS] = {}, nc[kMaxS] = {}, off[kMaxS] = {};  // off: first panel of surface m
    int64_t vsz[kMaxS] = {}, wsz[kMaxS] = {}, psz[kMaxS] = {};  // elements per request: verts, wake, panels
    std::vector<void*> allocs;
    int64_t bytes = 0;                                           // device bytes held
    hipEvent_t done = nullptr;  // recorded after the plan's own latest work (creation, solve, results)
    double* init[kMaxS] = {};                                    // {ns+1, nc+1, 4}
    double *verts[kMaxS] = {}, *wake[kMaxS] = {};                // per request b: + b * vsz / wsz
    double *colloc[kMaxS] = {}, *normals[kMaxS] = {}, *areas[kMaxS] = {};
    double *vel[kMaxS] = {}, *forces[kMaxS] = {}, *gdelta[kMaxS] = {};
    double *lhs = nullptr, *rhs = nullptr, *x = nullptr, *res = nullptr, *tr = nullptr, *coef = nullptr;
    float *lhs32 = nullptr, *r32 = nullptr;
    int32_t *ipiv = nullptr, *info = nullptr;
    double *rowabs = nullptr, *berr = nullptr;  // served mode: row sums of |A64|, backward error per request
    int32_t* conv = nullptr;                     // served mode: 1 = the request's x passed the test
    std::vector<int32_t> st_status, st_corr;     // per request of the last solve (tvlm_vlm_refinement)
    std::vector<double> st_berr;
    bool checked = false;                        // the last solve's requests were checked (and rescued)
    double* tr_host = nullptr;  // pinned: transforms of all requests and surfaces
    std::vector<double> alpha, beta, ref;  // per request; ref: CM reference point in the moved frame
    double u_inf = 1.0, rho = 1.0;
    bool solved = false;
    tvlm_lhs_group* group = nullptr;  // served mode: every request's influence matrix, one grouped launch

    int alloc(void** p, size_t bytes) {
        TVLM_HIP(hipMalloc(p, bytes));
        allocs.push_back(*p);
        this->bytes += (int64_t)bytes;
        return TVLM_OK;
    }
};

namespace tvlm {
namespace {

// r = rhs - A x per request (FP64).  A work-group takes 64 rows (one per lane) and its 8 waves
// split the columns (wave w: columns w, w + 8, ...), so each wave reads 64 consecutive rows of a
// column (512 contiguous bytes) and N / 8 work-groups' worth of waves keep enough loads in flight
// to stream the FP64 matrix at HBM rate; x[j] is the same for the whole wave; the 8 partial sums
// are added in LDS in a fixed order.
// With ROWABS the same pass also stores the row sums of |A| (for ||A||_inf of the convergence test).
template <bool ROWABS>
__global__ __launch_bounds__(kResidRows * kResidWaves) void k_residual(const double* __restrict__ A, int n,
                                                                      const double* __restrict__ rhs,
                                                                      const double* __restrict__ x,
                                                                      double* __restrict__ r,
                                                                      double* __restrict__ rowabs) {
    __shared__ double part[kResidWaves][kResidRows];
    __shared__ double apart[ROWABS ? kResidWaves : 1][kResidRows];
    const int64_t b = blockIdx.y;
    A += b * (int64_t)n * n;
    rhs += b * n;
    x += b * n;
    r += b * n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * kResidRows + lane;
    double acc = 0.0, aacc = 0.0;
    if (i < n) {
        const double* Ai = A + i;
#pragma unroll 4
        for (int j = w; j < n; j += kResidWaves) {
            const double aij = Ai[(int64_t)j * n];
            acc += aij * x[j];
            if constexpr (ROWABS) aacc += fabs(aij);
        }
    }
    part[w][lane] = acc;
    if constexpr (ROWABS) apart[w][lane] = aacc;
    __syncthreads();
    if (w == 0 && i < n) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kResidWaves; q++) s += part[q][lane];
        r[i] = rhs[i] - s;
        if constexpr (ROWABS) {
            double sa = 0.0;
#pragma unroll
            for (int q = 0; q < kResidWaves; q++) sa += apart[q][lane];
            rowabs[b * n + i] = sa;
        }
    }
}

// Convergence test of request blockIdx.x (LAPACK dsgesv's stopping test): with rnrm = ||r||_inf,
// xnrm = ||x||_inf, anrm = ||A||_inf (max row sum), converged iff the FP32 LU reported no problem
// (info == 0), every entry of r and x is finite and rnrm <= sqrt(n) eps xnrm anrm.  berr = rnrm /
// (xnrm anrm), the normwise backward error.
__global__ __launch_bounds__(256) void k_refine_check(const double* __restrict__ r, const double* __restrict__ x,
                                                      const double* __restrict__ rowabs,
                                                      const int32_t* __restrict__ info, int n,
                                                      double* __restrict__ berr, int32_t* __restrict__ conv) {
    __shared__ double sr[256], sx[256], sa[256];
    __shared__ int sbad[256];
    const int64_t b = blockIdx.x;
    r += b * n;
    x += b * n;
    rowabs += b * n;
    double mr = 0.0, mx = 0.0, ma = 0.0;
    int bad = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double ri = r[i], xi = x[i];
        bad |= !isfinite(ri) || !isfinite(xi);
        mr = fmax(mr, fabs(ri));
        mx = fmax(mx, fabs(xi));
        ma = fmax(ma, rowabs[i]);
    }
    sr[threadIdx.x] = mr;
    sx[threadIdx.x] = mx;
    sa[threadIdx.x] = ma;
    sbad[threadIdx.x] = bad;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + h]);
            sx[threadIdx.x] = fmax(sx[threadIdx.x], sx[threadIdx.x + h]);
            sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + h]);
            sbad[threadIdx.x] |= sbad[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rn = sr[0], xa = sx[0] * sa[0];
        const bool ok = info[b] == 0 && sbad[0] == 0 && rn <= sqrt((double)n) * 2.220446049250313e-16 * xa;
        berr[b] = sbad[0] ? (double)NAN : (xa > 0.0 ? rn / xa : (rn == 0.0 ? 0.0 : (double)INFINITY));
        conv[b] = ok ? 1 : 0;
    }
}

__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

__global__ void k_add_f32(double* __restrict__ x, const float* __restrict__ d, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += (double)d[i];
}

// gamma_delta(i, j) = gamma(i, j) - gamma(i, j-1), gamma_delta(i, 0) = gamma(i, 0) (vlm.cpp:126-139)
__global__ void k_gamma_delta(const double* __restrict__ g, double* __restrict__ gd, int ns, int nc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ns * nc) return;
    gd[p] = p >= ns ? g[p] - g[p - ns] : g[p];
}

int grid_for(int64_t n) { return (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096); }

void views(tvlm_vlm_plan* P, int b, int m, tvlm_view& v, tvlm_view& w, tvlm_view& c, tvlm_view& nn, tvlm_view& a,
           tvlm_view& vel, tvlm_view& f, tvlm_view& gd) {
    const int64_t ns = P->ns[m], nc = P->nc[m];
    v = {P->verts[m] + b * P->vsz[m], {ns + 1, nc + 1, 4}, {1, ns + 1, (ns + 1) * (nc + 1)}};
    w = {P->wake[m] + b * P->wsz[m], {ns + 1, 2, 4}, {1, ns + 1, (ns + 1) * 2}};
    c = {P->colloc[m] + b * 3 * P->psz[m], {ns, nc, 3}, {1, ns, ns * nc}};
    nn = {P->normals[m] + b * 3 * P->psz[m], {ns, nc, 3}, {1, ns, ns * nc}};
    a = {P->areas[m] + b * P->psz[m], {ns, nc, 1}, {1, ns, ns * nc}};
    vel = {P->vel[m] + b * 3 * P->psz[m], {ns, nc, 3}, {1, ns, ns * nc}};
    f = {P->forces[m] + b * 3 * P->psz[m], {ns, nc, 3}, {1, ns, ns * nc}};
    gd = {P->gdelta[m] + b * P->psz[m], {ns, nc, 1}, {1, ns, ns * nc}};
}

// geometry, influence matrix (matrix = false: zeroed only; the caller assembles it) and rhs of
// request b (vlm.cpp:97-124)
int assemble(tvlm_vlm_plan* P, int b, bool seeded, bool matrix = true) {
    hipStream_t st = P->stream;
    const int n = P->n_surf;
    std::vector<tvlm_view> V(n), W(n), C(n), Nn(n), A(n), VEL(n), F(n), GD(n), TR(n), INIT(n);
    for (int m = 0; m < n; m++) {
        views(P, b, m, V[m], W[m], C[m], Nn[m], A[m], VEL[m], F[m], GD[m]);
        const int64_t ns = P->ns[m], nc = P->nc[m];
        INIT[m] = {P->init[m], {ns + 1, nc + 1, 4}, {1, ns + 1, (ns + 1) * (nc + 1)}};
        TR[m] = {P->tr + ((int64_t)b * n + m) * 16, {4, 4, 1}, {1, 4, 16}};
        TVLM_HIP(hipMemcpyAsync(V[m].ptr, P->init[m], sizeof(double) * P->vsz[m], hipMemcpyDeviceToDevice, st));
    }
    TVLM_TRY(tvlm_wake_shed_f64(n, V.data(), W.data(), 0, st));
    TVLM_TRY(tvlm_displace_wing_f64(n, TR.data(), V.data(), INIT.data(), st));
    TVLM_TRY(tvlm_wake_shed_f64(n, V.data(), W.data(), 1, st));
    TVLM_TRY(tvlm_mesh_metrics_f64(P->alpha[b], n, V.data(), C.data(), Nn.data(), A.data(), st));
    const int64_t NN = (int64_t)P->N * P->N;
    double* lhs = P->lhs + (P->mode == 1 ? b * NN : 0);
    double* rhs = P->rhs + (int64_t)b * P->N;
    TVLM_TRY(tvlm_fill_f64(0, lhs, 1, 0.0, NN, st));
    TVLM_TRY(tvlm_fill_f64(0, rhs, 1, 0.0, P->N, st));
    tvlm_view L = {lhs, {P->N, P->N, 1}, {1, P->N, NN}};
    if (matrix && seeded)
        TVLM_TRY(lhs_assemble_seeded_f64(&L, n, C.data(), Nn.data(), V.data(), W.data(), 1, st));
    else if (matrix)
        TVLM_TRY(tvlm_lhs_assemble_f64(&L, n, C.data(), Nn.data(), V.data(), W.data(), 1, 0.0, st));
    const double a = P->alpha[b], be = P->beta[b], u = P->u_inf;
    const double fs[3] = {u * std::cos(a) * std::cos(be), -u * std::cos(a) * std::sin(be), u * std::sin(a)};
    for (int m = 0; m < n; m++)
        for (int k = 0; k < 3; k++)
            TVLM_TRY(tvlm_fill_f64(0, static_cast<double*>(VEL[m].ptr) + k * P->psz[m], 1, fs[k], P->psz[m], st));
    tvlm_view R = {rhs, {P->N, 1, 1}, {1, P->N, P->N}};
    TVLM_TRY(tvlm_rhs_assemble_velocities_f64(&R, n, Nn.data(), VEL.data(), st));
    return TVLM_OK;
}

// circulation differences, forces and per-surface coefficients of request b (vlm.cpp:126-151);
// coef[(b n_surf + m) * 8 + {0: area, 1: mac, 2: cl, 3..5: cm}]
int post(tvlm_vlm_plan* P, int b) {
    hipStream_t st = P->stream;
    const int n = P->n_surf;
    const double a = P->alpha[b], be = P->beta[b], u = P->u_inf;
    const double fs[3] = {u * std::cos(a) * std::cos(be), -u * std::cos(a) * std::sin(be), u * std::sin(a)};
    // lift axis normalize(fs x y) (vlm/src/vlm_utils.cpp:13-15)
    double la[3] = {-fs[2], 0.0, fs[0]};
    const double ln = std::sqrt(la[0] * la[0] + la[2] * la[2]);
    la[0] /= ln;
    la[2] /= ln;
    const double* ref = P->ref.data() + 3 * b;
    for (int m = 0; m < n; m++) {
        tvlm_view V, W, C, Nn, A, VEL, F, GD;
        views(P, b, m, V, W, C, Nn, A, VEL, F, GD);
        const double* g = P->x + (int64_t)b * P->N + P->off[m];
        hipLaunchKernelGGL(k_gamma_delta, dim3(grid_for(P->psz[m])), dim3(256), 0, st, g,
                           static_cast<double*>(GD.ptr), (int)P->ns[m], (int)P->nc[m]);
        TVLM_LAUNCHED();
        TVLM_TRY(tvlm_forces_steady_f64(&V, &GD, &VEL, &F, st));
        double* cf = P->coef + ((int64_t)b * n + m) * 8;
        tvlm_view A2 = {A.ptr, {P->ns[m], P->nc[m], 1}, {1, P->ns[m], P->psz[m]}};
        TVLM_TRY(tvlm_sum_f64(&A2, 2, cf + 0, st));
        TVLM_TRY(tvlm_mesh_mac_f64(&V, &A2, cf + 1, st));
        // area and MAC enter only the normalisation: CL and CM are formed on the host from these sums
        TVLM_TRY(tvlm_coeff_cl_f64(&F, la, fs, P->rho, 1.0, cf + 2, st));
        TVLM_TRY(tvlm_coeff_cm_f64(&F, &V, ref, fs, P->rho, 1.0, 1.0, cf + 3, st));
    }
    return TVLM_OK;
}

// Served mode, after the solve: reads the per-request test results; for every request that failed,
// (1) if its FP32 LU stalled (info -999), re-factors its FP32 copy with one-group panels and restarts
// its refinement from x = 0; (2) applies more corrections, one at a time with the test after each,
// up to kMaxCorrections in all; (3) if it still fails, solves it with the FP64 LU of its FP64 matrix
// (factored in place; the next solve assembles it again); then redoes its post-processing.  Runs
// synchronously on the plan's stream; requests that passed are not touched.
int rescue(tvlm_vlm_plan* P) {
    const int B = P->active;
    const int64_t N = P->N, NN = N * N;
    hipStream_t st = P->stream;
    std::vector<int32_t> conv(B), info(B);
    std::vector<double> berr(B);
    TVLM_HIP(hipMemcpyAsync(conv.data(), P->conv, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipMemcpyAsync(info.data(), P->info, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipMemcpyAsync(berr.data(), P->berr, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < B; b++) {
        P->st_berr[b] = berr[b];
        if (conv[b]) continue;
        double* A64 = P->lhs + b * NN;
        float* A32 = P->lhs32 + b * NN;
        int32_t* ipiv = P->ipiv + b * N;
        int32_t* inf = P->info + b;
        double* x = P->x + b * N;
        double* rhs = P->rhs + b * N;
        double* res = P->res + b * N;
        float* r32 = P->r32 + b * N;
        tvlm_view L32 = {A32, {N, N, 1}, {1, N, NN}};
        tvlm_view R32 = {r32, {N, 1, 1}, {1, N, N}};
        const dim3 rg((unsigned)((N + kResidRows - 1) / kResidRows), 1);
        int32_t status = 0, corr = P->st_corr[b];
        if (info[b] == kGetrfStalled) {
            hipLaunchKernelGGL(k_to_f32, dim3(grid_for(NN)), dim3(256), 0, st, A64, A32, NN);
            TVLM_LAUNCHED();
            TVLM_TRY(getrf_single_group_f32(&L32, ipiv, inf, st));
            TVLM_TRY(tvlm_fill_f64(0, x, 1, 0.0, N, st));
            TVLM_HIP(hipMemcpyAsync(res, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            status |= kStRefactored;
            corr = 0;
        }
        bool ok = false;
        while (corr < kMaxCorrections && !ok) {
            hipLaunchKernelGGL(k_to_f32, dim3(grid_for(N)), dim3(256), 0, st, res, r32, N);
            TVLM_LAUNCHED();
            TVLM_TRY(tvlm_getrs_f32(&L32, ipiv, &R32, st));
            hipLaunchKernelGGL(k_add_f32, dim3(grid_for(N)), dim3(256), 0, st, x, r32, N);
            TVLM_LAUNCHED();
            corr++;
            hipLaunchKernelGGL(k_residual<false>, rg, dim3(kResidRows * kResidWaves), 0, st, A64, (int)N, rhs, x, res,
                               P->rowabs + b * N);
            TVLM_LAUNCHED();
            hipLaunchKernelGGL(k_refine_check, dim3(1), dim3(256), 0, st, res, x, P->rowabs + b * N, inf, (int)N,
                               P->berr + b, P->conv + b);
            TVLM_LAUNCHED();
            int32_t c = 0;
            TVLM_HIP(hipMemcpyAsync(&c, P->conv + b, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            TVLM_HIP(hipMemcpyAsync(&P->st_berr[b], P->berr + b, sizeof(double), hipMemcpyDeviceToHost, st));
            TVLM_HIP(hipStreamSynchronize(st));
            ok = c != 0;
        }
        if (ok) {
            if (corr > kRefine + 1) status |= kStMoreCorrections;
        } else {
            // the plain path's solve of this request: FP64 LU of its FP64 matrix, in place
            tvlm_view L64 = {A64, {N, N, 1}, {1, N, NN}};
            TVLM_TRY(tvlm_getrf_f64(&L64, ipiv, inf, st));
            int32_t i64 = 0;
            TVLM_HIP(hipMemcpyAsync(&i64, inf, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            TVLM_HIP(hipStreamSynchronize(st));
            if (i64 == kGetrfStalled) {
                // its factors are invalid and the matrix is gone: assemble it again, one-group panels
                TVLM_TRY(assemble(P, b, true));
                TVLM_TRY(getrf_single_group_f64(&L64, ipiv, inf, st));
                TVLM_HIP(hipMemcpyAsync(&i64, inf, sizeof(int32_t), hipMemcpyDeviceToHost, st));
                TVLM_HIP(hipStreamSynchronize(st));
                status |= kStRefactored;
            }
            TVLM_HIP(hipMemcpyAsync(x, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            tvlm_view X = {x, {N, 1, 1}, {1, N, N}};
            TVLM_TRY(tvlm_getrs_f64(&L64, ipiv, &X, st));
            status |= kStFallback64;
            if (i64 != 0) status |= kStUnresolved;
            P->st_berr[b] = NAN;  // the FP64 matrix is factored in place: no residual against it
        }
        P->st_status[b] = status;
        P->st_corr[b] = corr;
        TVLM_TRY(post(P, b));
    }
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

using namespace tvlm;

extern "C" {

int tvlm_vlm_create(tvlm_vlm_plan** plan, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                    const double* const* verts_init, int32_t batch, int32_t mode, void* stream) {
    TVLM_REQUIRE(plan && ns && nc && verts_init, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= kMaxS, "n_surf %d out of range", n_surf);
    TVLM_REQUIRE(batch >= 1 && batch <= 4096, "batch %d out of range", batch);
    TVLM_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (plain FP64) or 1 (served)");
    auto* P = new tvlm_vlm_plan();
    P->mode = mode;
    P->n_surf = n_surf;
    P->batch = batch;
    P->stream = as_stream(stream);
    int64_t N = 0;
    for (int m = 0; m < n_surf; m++) {
        if (ns[m] < 1 || nc[m] < 1) {
            delete P;
            set_error("surface %d: empty mesh", m);
            return TVLM_ERR_INVALID;
        }
        P->ns[m] = ns[m];
        P->nc[m] = nc[m];
        P->off[m] = N;
        P->vsz[m] = (ns[m] + 1) * (nc[m] + 1) * 4;
        P->wsz[m] = (ns[m] + 1) * 2 * 4;
        P->psz[m] = ns[m] * nc[m];
        N += ns[m] * nc[m];
    }
    P->N = (int32_t)N;
    const int64_t NN = N * N;
    const int64_t nb = mode == 1 ? batch : 1;  // plain mode reuses one matrix for every request
    auto fail = [&](int rc) {
        (void)tvlm_lhs_group_destroy(P->group);
        for (void* p : P->allocs) (void)hipFree(p);
        if (P->tr_host) (void)hipHostFree(P->tr_host);
        if (P->done) (void)hipEventDestroy(P->done);
        delete P;
        return rc;
    };
    int rc = TVLM_OK;
    for (int m = 0; m < n_surf && rc == TVLM_OK; m++) {
        const size_t v = sizeof(double) * P->vsz[m], p3 = sizeof(double) * 3 * P->psz[m], p1 = sizeof(double) * P->psz[m];
        rc = P->alloc((void**)&P->init[m], v);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->verts[m], v * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->wake[m], sizeof(double) * P->wsz[m] * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->colloc[m], p3 * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->normals[m], p3 * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->areas[m], p1 * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->vel[m], p3 * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->forces[m], p3 * batch);
        if (rc == TVLM_OK) rc = P->alloc((void**)&P->gdelta[m], p1 * batch);
        if (rc == TVLM_OK) rc = tvlm_copy(0, P->init[m], 1, 1, verts_init[m], 1, sizeof(double), P->vsz[m], stream);
    }
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->lhs, sizeof(double) * NN * nb);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->lhs32, sizeof(float) * NN * nb);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->r32, sizeof(float) * N * batch);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->res, sizeof(double) * N * batch);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->rowabs, sizeof(double) * N * batch);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->berr, sizeof(double) * batch);
    if (rc == TVLM_OK && mode == 1) rc = P->alloc((void**)&P->conv, sizeof(int32_t) * batch);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->rhs, sizeof(double) * N * batch);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->x, sizeof(double) * N * batch);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->tr, sizeof(double) * 16 * n_surf * batch);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->coef, sizeof(double) * 8 * n_surf * batch);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->ipiv, sizeof(int32_t) * N * nb);
    if (rc == TVLM_OK) rc = P->alloc((void**)&P->info, sizeof(int32_t) * batch);
    if (rc == TVLM_OK) {
        hipError_t e = hipHostMalloc((void**)&P->tr_host, sizeof(double) * 16 * n_surf * batch, hipHostMallocDefault);
        if (e != hipSuccess) {
            set_error("hipHostMalloc failed: %s", hipGetErrorString(e));
            rc = TVLM_ERR_NOMEM;
        }
    }
    if (rc == TVLM_OK) {
        // the wake rows' w coordinate is 1 and stays so (vlm.cpp:94)
        for (int m = 0; m < n_surf && rc == TVLM_OK; m++)
            for (int b = 0; b < batch && rc == TVLM_OK; b++)
                rc = tvlm_fill_f64(0, P->wake[m] + b * P->wsz[m] + 3 * (P->ns[m] + 1) * 2, 1, 1.0, (P->ns[m] + 1) * 2,
                                   stream);
    }
    if (rc == TVLM_OK && hipEventCreateWithFlags(&P->done, hipEventDisableTiming) != hipSuccess) {
        set_error("hipEventCreateWithFlags failed");
        rc = TVLM_ERR_HIP;
    }
    if (rc == TVLM_OK && hipEventRecord(P->done, P->stream) != hipSuccess) {
        set_error("hipEventRecord failed");
        rc = TVLM_ERR_HIP;
    }
    if (rc == TVLM_OK && mode == 1) {
        // the grouped assembly's work list: request b's matrix and surfaces (iteration 1: one wake row)
        std::vector<tvlm_lhs_item> items(batch);
        std::vector<tvlm_view> vv((size_t)batch * n_surf * 8);
        for (int b = 0; b < batch; b++) {
            tvlm_view *V = &vv[((size_t)b * 8 + 0) * n_surf], *W = &vv[((size_t)b * 8 + 1) * n_surf],
                      *C = &vv[((size_t)b * 8 + 2) * n_surf], *Nn = &vv[((size_t)b * 8 + 3) * n_surf];
            for (int m = 0; m < n_surf; m++) {
                tvlm_view A, VEL, F, GD;
                views(P, b, m, V[m], W[m], C[m], Nn[m], A, VEL, F, GD);
            }
            items[b] = {{P->lhs + b * NN, {N, N, 1}, {1, N, NN}}, n_surf, C, Nn, V, W, 1};
        }
        rc = tvlm_lhs_group_create(&P->group, batch, items.data(), stream);
    }
    if (rc != TVLM_OK) return fail(rc);
    P->alpha.assign(batch, 0.0);
    P->beta.assign(batch, 0.0);
    P->ref.assign(3 * batch, 0.0);
    *plan = P;
    return TVLM_OK;
}

int tvlm_vlm_solve(tvlm_vlm_plan* P, const double* alpha, const double* beta, double u_inf, double rho,
                   const double* cm_ref) {
    TVLM_REQUIRE(P, "null plan");
    return tvlm_vlm_solve_n(P, P->batch, alpha, beta, u_inf, rho, cm_ref);
}

int tvlm_vlm_solve_n(tvlm_vlm_plan* P, int32_t n_active, const double* alpha, const double* beta, double u_inf,
                     double rho, const double* cm_ref) {
    TVLM_REQUIRE(P && alpha, "null argument");
    TVLM_REQUIRE(n_active >= 1 && n_active <= P->batch, "n_active %d out of range 1..%d", n_active, P->batch);
    const int n = P->n_surf, B = n_active;
    P->active = B;
    P->u_inf = u_inf;
    P->rho = rho;
    const double c0[3] = {cm_ref ? cm_ref[0] : 0.25, cm_ref ? cm_ref[1] : 0.0, cm_ref ? cm_ref[2] : 0.0};
    TVLM_HIP(hipEventSynchronize(P->done));  // the pinned transform buffer may still be read
    for (int b = 0; b < B; b++) {
        P->alpha[b] = alpha[b];
        P->beta[b] = beta ? beta[b] : 0.0;
        // translation by -100 u (cos a, 0, sin a): 100-chord wake row (vlm.cpp:104-108)
        const double tx = -100.0 * u_inf * std::cos(alpha[b]), tz = -100.0 * u_inf * std::sin(alpha[b]);
        for (int m = 0; m < n; m++) {
            double* T = P->tr_host + ((int64_t)b * n + m) * 16;
            for (int k = 0; k < 16; k++) T[k] = (k % 5 == 0) ? 1.0 : 0.0;
            T[12] = tx;
            T[14] = tz;
        }
        P->ref[3 * b + 0] = c0[0] + tx;
        P->ref[3 * b + 1] = c0[1];
        P->ref[3 * b + 2] = c0[2] + tz;
    }
    hipStream_t st = P->stream;
    TVLM_HIP(hipMemcpyAsync(P->tr, P->tr_host, sizeof(double) * 16 * n * B, hipMemcpyHostToDevice, st));
    const int64_t N = P->N, NN = N * N;
    if (P->mode == 0) {
        tvlm_view L = {P->lhs, {N, N, 1}, {1, N, NN}};
        for (int b = 0; b < B; b++) {
            TVLM_TRY(assemble(P, b, false));
            TVLM_TRY(tvlm_getrf_f64(&L, P->ipiv, P->info + b, st));
            TVLM_HIP(hipMemcpyAsync(P->x + b * N, P->rhs + b * N, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            tvlm_view X = {P->x + b * N, {N, 1, 1}, {1, N, N}};
            TVLM_TRY(tvlm_getrs_f64(&L, P->ipiv, &X, st));
            TVLM_TRY(post(P, b));
        }
    } else {
        // geometry and rhs request by request, then every request's matrix in one grouped launch
        for (int b = 0; b < B; b++) TVLM_TRY(assemble(P, b, true, false));
        TVLM_TRY(tvlm_lhs_group_assemble_f64(P->group, B, 1, 0.0, st));
        hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, st, P->lhs, P->lhs32, NN * B);
        TVLM_LAUNCHED();
        tvlm_view L32 = {P->lhs32, {N, N, B}, {1, N, NN}};
        TVLM_TRY(tvlm_getrf_f32(&L32, P->ipiv, P->info, st));
        tvlm_view R32 = {P->r32, {N, 1, B}, {1, N, N}};
        // x = 0, r = b; then kRefine + 1 corrections x += A32^{-1} r, r = b - A64 x
        TVLM_TRY(tvlm_fill_f64(0, P->x, 1, 0.0, N * B, st));
        const dim3 rg((unsigned)((N + kResidRows - 1) / kResidRows), B);
        for (int it = 0; it <= kRefine; it++) {
            if (it == 0) {
                hipLaunchKernelGGL(k_to_f32, dim3(grid_for(N * B)), dim3(256), 0, st, P->rhs, P->r32, N * B);
            } else {
                if (it == 1)
                    hipLaunchKernelGGL(k_residual<true>, rg, dim3(kResidRows * kResidWaves), 0, st, P->lhs, (int)N,
                                       P->rhs, P->x, P->res, P->rowabs);
                else
                    hipLaunchKernelGGL(k_residual<false>, rg, dim3(kResidRows * kResidWaves), 0, st, P->lhs, (int)N,
                                       P->rhs, P->x, P->res, P->rowabs);
                TVLM_LAUNCHED();
                hipLaunchKernelGGL(k_to_f32, dim3(grid_for(N * B)), dim3(256), 0, st, P->res, P->r32, N * B);
            }
            TVLM_LAUNCHED();
            TVLM_TRY(tvlm_getrs_f32(&L32, P->ipiv, &R32, st));
            hipLaunchKernelGGL(k_add_f32, dim3(grid_for(N * B)), dim3(256), 0, st, P->x, P->r32, N * B);
            TVLM_LAUNCHED();
        }
        // the returned x of every request: its residual and the convergence test (the rescue of the
        // requests that fail it runs in tvlm_vlm_results)
        hipLaunchKernelGGL(k_residual<false>, rg, dim3(kResidRows * kResidWaves), 0, st, P->lhs, (int)N, P->rhs, P->x,
                           P->res, P->rowabs);
        TVLM_LAUNCHED();
        hipLaunchKernelGGL(k_refine_check, dim3(B), dim3(256), 0, st, P->res, P->x, P->rowabs, P->info, (int)N,
                           P->berr, P->conv);
        TVLM_LAUNCHED();
        for (int b = 0; b < B; b++) TVLM_TRY(post(P, b));
    }
    TVLM_HIP(hipEventRecord(P->done, st));
    P->solved = true;
    P->checked = false;
    P->st_status.assign(B, 0);
    P->st_corr.assign(B, P->mode == 1 ? kRefine + 1 : 0);
    P->st_berr.assign(B, 0.0);
    return TVLM_OK;
}

// Waits for the plan's last solve.  cl[n], cm[n][3], gamma[n][N], info[n] for the n requests of
// the last solve; null pointers are skipped.  The copies are queued on the plan's stream (not the
// null stream), so reading one plan's results does not wait for work on other streams.
int tvlm_vlm_results(tvlm_vlm_plan* P, double* cl, double* cm, double* gamma, int32_t* info) {
    TVLM_REQUIRE(P && P->solved, "no solve has run on this plan");
    const int n = P->n_surf, B = P->active;
    hipStream_t st = P->stream;
    if (P->mode == 1 && !P->checked) {
        TVLM_TRY(rescue(P));
        P->checked = true;
    }
    std::vector<double> coef((size_t)8 * n * B);
    TVLM_HIP(hipMemcpyAsync(coef.data(), P->coef, sizeof(double) * coef.size(), hipMemcpyDeviceToHost, st));
    if (gamma) TVLM_HIP(hipMemcpyAsync(gamma, P->x, sizeof(double) * P->N * B, hipMemcpyDeviceToHost, st));
    if (info) TVLM_HIP(hipMemcpyAsync(info, P->info, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipEventRecord(P->done, st));
    TVLM_HIP(hipEventSynchronize(P->done));
    for (int b = 0; b < B; b++) {
        // per surface: cf[2] = sum(F . l) / q and cf[3..5] = sum(r x F) / q (q = rho |V|^2 / 2,
        // vlm_backend_cpu.cpp:323-362); CL area-weighted, CM area*MAC-weighted over the surfaces
        double cl_sum = 0.0, area_sum = 0.0, mac_sum = 0.0, cm_sum[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < n; m++) {
            const double* cf = coef.data() + ((size_t)b * n + m) * 8;
            cl_sum += cf[2];
            for (int k = 0; k < 3; k++) cm_sum[k] += cf[3 + k];
            area_sum += cf[0];
            mac_sum += cf[1];
        }
        if (cl) cl[b] = cl_sum / area_sum;
        if (cm)
            for (int k = 0; k < 3; k++) cm[3 * b + k] = cm_sum[k] / (area_sum * mac_sum);
    }
    return TVLM_OK;
}

// What the served refinement took for each request of the last solve (after tvlm_vlm_results).
int tvlm_vlm_refinement(tvlm_vlm_plan* P, int32_t* status, int32_t* corrections, double* backward_error) {
    TVLM_REQUIRE(P && P->solved, "no solve has run on this plan");
    if (P->mode == 1 && !P->checked) {
        TVLM_TRY(rescue(P));
        P->checked = true;
    }
    const int B = P->active;
    for (int b = 0; b < B; b++) {
        if (status) status[b] = P->st_status[b];
        if (corrections) corrections[b] = P->st_corr[b];
        if (backward_error) backward_error[b] = P->st_berr[b];
    }
    return TVLM_OK;
}

// Moves the plan to another stream: later solves queue on `stream`, after the plan's own latest
// work on its previous stream (a device-side wait on the plan's event; the host does not block, and
// whatever else runs on the previous stream is not waited for).
int tvlm_vlm_set_stream(tvlm_vlm_plan* P, void* stream) {
    TVLM_REQUIRE(P, "null plan");
    if (as_stream(stream) == P->stream) return TVLM_OK;
    TVLM_HIP(hipStreamWaitEvent(as_stream(stream), P->done, 0));
    P->stream = as_stream(stream);
    return TVLM_OK;
}

// *done = 1 once the plan's latest work has finished (does not block)
int tvlm_vlm_query(const tvlm_vlm_plan* P, int32_t* done) {
    TVLM_REQUIRE(P && done, "null argument");
    const hipError_t e = hipEventQuery(P->done);
    *done = e == hipSuccess ? 1 : 0;
    if (e != hipErrorNotReady) TVLM_HIP(e);
    return TVLM_OK;
}

int tvlm_vlm_footprint(const tvlm_vlm_plan* P, int64_t* device_bytes) {
    TVLM_REQUIRE(P && device_bytes, "null argument");
    *device_bytes = P->bytes;
    return TVLM_OK;
}

int tvlm_vlm_destroy(tvlm_vlm_plan* P) {
    if (!P) return TVLM_OK;
    (void)hipEventSynchronize(P->done);
    (void)tvlm_lhs_group_destroy(P->group);
    for (void* p : P->allocs) (void)hipFree(p);
    (void)hipEventDestroy(P->done);
    if (P->tr_host) (void)hipHostFree(P->tr_host);
    delete P;
    return TVLM_OK;
}

}  // extern "C"

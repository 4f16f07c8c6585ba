// tvlm_uvlm.hip -- device-resident unsteady VLM time loop (the caller of the hot path, UVLM::run,
// vlm/src/vlm.cpp:358-509, kept on the GPU end to end).
//
// The prescribed kinematics are a function of time only, so the host evaluates the dual 4x4 of
// every surface at every step once (value T and derivative dT) and uploads them; the loop then
// needs no host interaction: per step
//   k_pre    : wake shed of the trailing-edge vertex row, kinematic velocities -(dT [x;1]) at the
//              t = 0 collocation points, rhs = -(V . n)                        (vlm.cpp:412-438)
//   wake induction + k_wake_reduce : rhs -= wake induction                    (vlm.cpp:439)
//              plain mode: k_wake_partials, exact FP64 Biot-Savart ring by ring;
//              fast mode : k_wake_lattice, every wake segment once with the difference of the
//              circulations of the two rings it borders
//   solve    : plain mode: LU forward/backward substitution (getrs);
//              fast mode : gamma = A^{-1} rhs with the FP64 inverse formed once from the LU
//   k_post_advance (one work-group): gamma post-processing, shed of Gamma_TE into the wake,
//              unsteady forces, CL and CM of all surfaces with fixed-order reductions
//              (vlm.cpp:441-487); then displacement by T(t+dt), panel metrics (vlm.cpp:489-507)
//              and the next step's wake shed, velocities and rhs
// The fast mode evaluates the wake induction in the FP64 seeded form (FP64 arithmetic, rsqrt and
// reciprocal from the FP64 hardware estimates plus one Newton step) and replays the whole loop as
// one hipGraph.  Its geometry does not depend on the solution (prescribed motion), so k_geom_all
// computes every step's pose, panel metrics, kinematic velocities, velocity rhs and wake vertex row
// once at setup; a step is then the wake lattice, its reduction into rhs = rhs_vel - wake, the GEMV
// partials of A^{-1} rhs and k_post (gamma post-processing, wake Gamma, forces, CL/CM).
#include <algorithm>
#include <cstring>
#include <vector>

#include "tvlm_kernels.hpp"

struct tvlm_uvlm_plan;

namespace tvlm {
namespace {

constexpr int kMaxS = TVLM_MAX_SURFACES;

struct USurf {
    double *verts_pos, *verts, *colloc, *colloc_h, *normals, *areas, *vel, *forces;
    double *gprev, *gdelta, *gdt, *wake, *gamma_wake;
    // fast mode: geometry of every step (step s at + s * {nv*3, np*3, np*3, np, np*3})
    double *g_verts, *g_colloc, *g_normals, *g_areas, *g_vel;
    int32_t ns, nc, off, v_off, te_off;
};

struct UTab {
    int32_t n, N, NV, NTE;
    int32_t t_steps;
    USurf s[kMaxS];
};

// Simulation z of a plan holding several simulations (fast mode): every per-simulation buffer of
// simulation z lies z * zs doubles after simulation 0's (one contiguous block per simulation).
__device__ __forceinline__ USurf sim_surf(const USurf& a, int64_t o) {
    USurf b = a;
    b.verts_pos += o;
    b.verts += o;
    b.colloc += o;
    b.colloc_h += o;
    b.normals += o;
    b.areas += o;
    b.vel += o;
    b.forces += o;
    b.gprev += o;
    b.gdelta += o;
    b.gdt += o;
    b.wake += o;
    b.gamma_wake += o;
    b.g_verts += o;
    b.g_colloc += o;
    b.g_normals += o;
    b.g_areas += o;
    b.g_vel += o;
    return b;
}

__device__ __forceinline__ int surf_of(const UTab& t, int idx, int which) {
    // which: 0 panel offsets, 1 vertex offsets, 2 trailing-edge copy offsets
    int m = 0;
    while (m + 1 < t.n) {
        const int next = which == 0 ? t.s[m + 1].off : which == 1 ? t.s[m + 1].v_off : t.s[m + 1].te_off;
        if (idx < next) break;
        m++;
    }
    return m;
}

// wake shed + kinematic velocities + rhs velocities
__global__ void k_pre(UTab tab, const double* __restrict__ dT_all, int step, double* __restrict__ rhs) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < tab.N) {
        const int m = surf_of(tab, e, 0);
        const USurf& S = tab.s[m];
        const int p = e - S.off;
        const int np = S.ns * S.nc;
        const double* d = dT_all + ((int64_t)step * tab.n + m) * 16;
        const double x = S.colloc_h[p], y = S.colloc_h[p + np], z = S.colloc_h[p + 2 * np];
        double V[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            V[r] = -(d[r] * x + d[r + 4] * y + d[r + 8] * z + d[r + 12]);
            S.vel[p + r * np] = V[r];
        }
        rhs[e] = -(V[0] * S.normals[p] + V[1] * S.normals[p + np] + V[2] * S.normals[p + 2 * np]);
        return;
    }
    const int q = e - tab.N;
    if (q >= tab.NTE) return;
    const int m = surf_of(tab, q, 2);
    const USurf& S = tab.s[m];
    const int l = q - S.te_off;
    const int vs1 = S.ns + 1;
    const int i = l % vs1, k = l / vs1;
    const int64_t vs2 = (int64_t)vs1 * (S.nc + 1), ws2 = (int64_t)vs1 * tab.t_steps;
    const int row = tab.t_steps - 1 - step;
    S.wake[i + (int64_t)row * vs1 + k * ws2] = S.verts[i + (int64_t)S.nc * vs1 + k * vs2];
}

// y(partial cb) = A(:, cols of block cb) x(cols); 64 rows per work-group, 4 waves split the columns
__global__ __launch_bounds__(256) void k_gemv_part(const double* __restrict__ A, int64_t lda, int n, int cbs,
                                                   const double* __restrict__ x, double* __restrict__ part,
                                                   int64_t zs = 0) {
    __shared__ double red[4][64];
    A += (int64_t)blockIdx.z * zs;  // simulation blockIdx.z
    x += (int64_t)blockIdx.z * zs;
    part += (int64_t)blockIdx.z * zs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = blockIdx.x * 64 + lane;
    const int c0 = blockIdx.y * cbs, c1 = min(n, c0 + cbs);
    double s = 0.0;
    if (row < n)
        for (int cb0 = c0 + wave; cb0 < c1; cb0 += 4 * 8) {  // eight loads in flight, summed in order
            double a[8], xv[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int c = cb0 + 4 * k;
                a[k] = c < c1 ? A[row + (int64_t)c * lda] : 0.0;
                xv[k] = c < c1 ? x[c] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (cb0 + 4 * k < c1) s += a[k] * xv[k];
        }
    red[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && row < n) part[(int64_t)blockIdx.y * n + row] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

constexpr int kPostThreads = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums K per-thread values over the work-group: butterfly inside each wave, then thread 0 adds
// the waves in index order (fixed order: bitwise reproducible).  Result valid in thread 0 only.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double (*lds)[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = lds[0][k];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) s += lds[w][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// FROM_PARTIALS: a panel's circulation and its upstream chordwise neighbour's are summed from the
// solve's partials inside the panel loop (same order, so the same values) instead of a first pass
// through gamma[] and a barrier; the panel's partial and geometry loads are then issued together.
template <int NT, bool FROM_PARTIALS = false>
__device__ __forceinline__ void post_step(const UTab& tab, const double* __restrict__ gsrc, int cb, int step,
                                          double inv_dt, const double* __restrict__ fs_all,
                                          const double* __restrict__ ref_all, double rho, double* __restrict__ gamma,
                                          double* __restrict__ cl_out, double* __restrict__ cm_out,
                                          double (*lds)[6], int64_t soff = 0) {
    const int t = threadIdx.x;
    const int N = tab.N;
    auto gsum = [&](int P) { return strided_sum(gsrc, (int64_t)N, P, 0, cb, 1); };
    if constexpr (!FROM_PARTIALS) {
        for (int P = t; P < N; P += NT) gamma[P] = gsum(P);
        __syncthreads();
    }
    const v3<double> fs = {fs_all[3 * step], fs_all[3 * step + 1], fs_all[3 * step + 2]};
    const v3<double> ref = {ref_all[3 * step], ref_all[3 * step + 1], ref_all[3 * step + 2]};
    // lift axis = normalize(freestream x (0,1,0))  (vlm/src/vlm_backend.cpp:86)
    v3<double> la = cross(fs, v3<double>{0.0, 1.0, 0.0});
    const double ll = sqrt(dot(la, la));
    la = {la.x / ll, la.y / ll, la.z / ll};
    const double q = 0.5 * rho * dot(fs, fs);
    double cl_num = 0.0, tot_a = 0.0, tot_mac = 0.0;
    double cmn[3] = {0.0, 0.0, 0.0};
    const bool forces = step > 0;  // the reference skips the coefficients of step 0 (vlm.cpp:461)
    for (int m = 0; m < tab.n; m++) {
        const USurf S = sim_surf(tab.s[m], soff);
        const int ns = S.ns, nc = S.nc, np = ns * nc;
        const int64_t vs1 = ns + 1, vs2 = vs1 * (nc + 1);
        double a_acc = 0.0, c_acc = 0.0, mx = 0.0, my = 0.0, mz = 0.0, mac_acc = 0.0;
        for (int p = t; p < np; p += NT) {
            const int i = p % ns, j = p / ns;
            // every operand of the panel is loaded before the first store (the stores below go
            // through pointers the compiler must assume may alias them), so the loads overlap
            const double g = FROM_PARTIALS ? gsum(S.off + p) : gamma[S.off + p];
            const double gup = j > 0 ? (FROM_PARTIALS ? gsum(S.off + p - ns) : gamma[S.off + p - ns]) : 0.0;
            const double gprev = S.gprev[p];
            const int64_t a = i + (int64_t)j * vs1, b = a + 1;
            v3<double> U{}, va{}, vb{}, nrm{};
            double area = 0.0;
            if (forces) {
                U = {S.vel[p], S.vel[p + np], S.vel[p + 2 * np]};
                va = {S.verts[a], S.verts[a + vs2], S.verts[a + 2 * vs2]};
                vb = {S.verts[b], S.verts[b + vs2], S.verts[b + 2 * vs2]};
                nrm = {S.normals[p], S.normals[p + np], S.normals[p + 2 * np]};
                area = S.areas[p];
            }
            const double gd = j > 0 ? g + (-1.0) * gup : g;
            if constexpr (FROM_PARTIALS) gamma[S.off + p] = g;
            const double gt = (g + (-1.0) * gprev) * inv_dt;
            S.gdelta[p] = gd;
            S.gdt[p] = gt;
            S.gprev[p] = g;
            if (j == nc - 1) S.gamma_wake[i + (int64_t)(tab.t_steps - 2 - step) * ns] = g;
            if (forces) {
                const v3<double> cr = cross(U, vb - va);
                const double s1 = rho * gd;
                const double u = rho * gt * area;
                const v3<double> F = {s1 * cr.x + u * nrm.x, s1 * cr.y + u * nrm.y, s1 * cr.z + u * nrm.z};
                S.forces[p] = F.x;
                S.forces[p + np] = F.y;
                S.forces[p + 2 * np] = F.z;
                c_acc += dot(F, la);
                const v3<double> lever = {0.5 * (va.x + vb.x) - ref.x, 0.5 * (va.y + vb.y) - ref.y,
                                          0.5 * (va.z + vb.z) - ref.z};
                const v3<double> mo = cross(lever, F);
                mx += mo.x;
                my += mo.y;
                mz += mo.z;
                a_acc += area;
            }
        }
        if (forces) {
            for (int i = t; i < ns; i += NT) {
                const int64_t te = (int64_t)nc * vs1;
                const double* V = S.verts;
                const v3<double> d0 = {V[i] - V[i + te], V[i + vs2] - V[i + te + vs2], V[i + 2 * vs2] - V[i + te + 2 * vs2]};
                const v3<double> d1 = {V[i + 1] - V[i + 1 + te], V[i + 1 + vs2] - V[i + 1 + te + vs2],
                                       V[i + 1 + 2 * vs2] - V[i + 1 + te + 2 * vs2]};
                const v3<double> d3 = {V[i + 1] - V[i], V[i + 1 + vs2] - V[i + vs2], V[i + 1 + 2 * vs2] - V[i + 2 * vs2]};
                const double c0 = sqrt(dot(d0, d0)), c1 = sqrt(dot(d1, d1)), w = sqrt(dot(d3, d3));
                mac_acc += 0.5 * (c0 * c0 + c1 * c1) * w;
            }
            double v[6] = {a_acc, c_acc, mx, my, mz, mac_acc};
            block_sums<6>(v, lds);
            const double area = v[0], cs = v[1], sx = v[2], sy = v[3], sz = v[4], macs = v[5];
            // coeff_cl / coeff_cm / mesh_mac / *_multibody (vlm_backend_cpu.cpp:323-362, vlm_backend.cpp:71-139)
            const double mac = macs / area;
            const double cl_m = cs / (q * area);
            const double den = q * area * mac;
            cl_num += cl_m * area;
            cmn[0] += (sx / den) * area * mac;
            cmn[1] += (sy / den) * area * mac;
            cmn[2] += (sz / den) * area * mac;
            tot_a += area;
            tot_mac += mac;
        }
    }
    if (t == 0) {
        cl_out[step] = forces ? cl_num / tot_a : 0.0;
        for (int k = 0; k < 3; k++) cm_out[3 * step + k] = forces ? cmn[k] / (tot_a * tot_mac) : 0.0;
    }
}

// Displacement by T(step) (verts = T verts_pos) and panel metrics at alpha = 0 (vlm.cpp:489-507),
// fused with the start of the next time step when do_pre is set (vlm.cpp:412-438): the thread of a
// trailing-edge vertex also sheds it into wake row t_steps-1-step, the thread of a panel also
// forms the kinematic velocity at its t = 0 collocation point and rhs = -(V . n).
__device__ __forceinline__ void advance_item(const UTab& tab, const double* __restrict__ T_all,
                                             const double* __restrict__ dT_all, int step, int do_pre,
                                             double* __restrict__ rhs, int e) {
    if (e < tab.NV) {
        const int m = surf_of(tab, e, 1);
        const USurf& S = tab.s[m];
        const int p = e - S.v_off;
        const int64_t nv = (int64_t)(S.ns + 1) * (S.nc + 1);
        const double* M = T_all + ((int64_t)step * tab.n + m) * 16;
        const double x[4] = {S.verts_pos[p], S.verts_pos[p + nv], S.verts_pos[p + 2 * nv], S.verts_pos[p + 3 * nv]};
        double y[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            y[r] = displace_row(M, 1, 4, r, x);
            S.verts[p + r * nv] = y[r];
        }
        const int vs1 = S.ns + 1;
        if (do_pre && p / vs1 == S.nc) {
            const int i = p % vs1;
            const int64_t ws2 = (int64_t)vs1 * tab.t_steps;
            const int row = tab.t_steps - 1 - step;
#pragma unroll
            for (int k = 0; k < 4; k++) S.wake[i + (int64_t)row * vs1 + k * ws2] = y[k];
        }
        return;
    }
    const int P = e - tab.NV;
    if (P >= tab.N) return;
    const int m = surf_of(tab, P, 0);
    const USurf& S = tab.s[m];
    const int p = P - S.off;
    const int ns = S.ns, nc = S.nc, np = ns * nc;
    const int i = p % ns, j = p / ns;
    const int64_t vs1 = ns + 1, nv = vs1 * (nc + 1);
    const double* M = T_all + ((int64_t)step * tab.n + m) * 16;
    auto vtx = [&](int ii, int jj) -> v3<double> {
        const int64_t o = ii + jj * vs1;
        const double x[4] = {S.verts_pos[o], S.verts_pos[o + nv], S.verts_pos[o + 2 * nv], S.verts_pos[o + 3 * nv]};
        return {displace_row(M, 1, 4, 0, x), displace_row(M, 1, 4, 1, x), displace_row(M, 1, 4, 2, x)};
    };
    v3<double> c, n;
    double area;
    panel_metrics(0.0, vtx(i, j), vtx(i + 1, j), vtx(i + 1, j + 1), vtx(i, j + 1), c, n, area);
    S.colloc[p] = c.x;
    S.colloc[p + np] = c.y;
    S.colloc[p + 2 * np] = c.z;
    S.normals[p] = n.x;
    S.normals[p + np] = n.y;
    S.normals[p + 2 * np] = n.z;
    S.areas[p] = area;
    if (do_pre) {
        const double* d = dT_all + ((int64_t)step * tab.n + m) * 16;
        const double x = S.colloc_h[p], y = S.colloc_h[p + np], z = S.colloc_h[p + 2 * np];
        double V[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            V[r] = -(d[r] * x + d[r + 4] * y + d[r + 8] * z + d[r + 12]);
            S.vel[p + r * np] = V[r];
        }
        rhs[P] = -(V[0] * n.x + V[1] * n.y + V[2] * n.z);
    }
}

// One work-group per step for everything O(N) after the solve: gamma post-processing, wake Gamma,
// forces, CL/CM (post_step), then, after a barrier, displacement + metrics + the next step's
// shed / velocities / rhs (advance_item) -- two launches and their gaps fewer per step.
__global__ __launch_bounds__(kPostThreads) void k_post_advance(UTab tab, const double* __restrict__ gsrc, int cb,
                                                               int step, double inv_dt,
                                                               const double* __restrict__ fs_all,
                                                               const double* __restrict__ ref_all, double rho,
                                                               double* __restrict__ gamma, double* __restrict__ cl_out,
                                                               double* __restrict__ cm_out,
                                                               const double* __restrict__ T_all,
                                                               const double* __restrict__ dT_all, int do_pre,
                                                               double* __restrict__ rhs) {
    __shared__ double lds[kPostThreads / 64][6];
    post_step<kPostThreads>(tab, gsrc, cb, step, inv_dt, fs_all, ref_all, rho, gamma, cl_out, cm_out, lds);
    __syncthreads();
    for (int e = threadIdx.x; e < tab.NV + tab.N; e += kPostThreads)
        advance_item(tab, T_all, dT_all, step + 1, do_pre, rhs, e);
}

// Fast mode: the geometry of steps 0 .. nsteps-1 at once.  Step s poses the mesh at T_all[s]
// applied to the t = 0 pose (step 0: the t = 0 pose itself, vlm.cpp:366-373, 489-507), stores the
// vertex coordinates, the panel metrics at alpha = 0, the kinematic velocity -(dT_all[s] [x_h; 1])
// at the t = 0 collocation point and rhs_vel = -(V . n) (vlm.cpp:419-438), and sheds the
// trailing-edge vertex row into wake row t_steps-1-s (vlm.cpp:412-418).  Same arithmetic as
// k_pre / advance_item.
__global__ void k_geom_all(UTab tab, const double* __restrict__ T_all, const double* __restrict__ dT_all,
                           int nsteps, double* __restrict__ g_rhs, int64_t zs = 0) {
    const int64_t so = (int64_t)blockIdx.y * zs;  // simulation blockIdx.y
    T_all += so;
    dT_all += so;
    g_rhs += so;
    const int per = tab.NV + tab.N;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)per * nsteps) return;
    const int s = (int)(e / per), r = (int)(e % per);
    if (r < tab.NV) {
        const int m = surf_of(tab, r, 1);
        const USurf S = sim_surf(tab.s[m], so);
        const int p = r - S.v_off;
        const int64_t nv = (int64_t)(S.ns + 1) * (S.nc + 1);
        const double* M = T_all + ((int64_t)s * tab.n + m) * 16;
        const double x[4] = {S.verts_pos[p], S.verts_pos[p + nv], S.verts_pos[p + 2 * nv], S.verts_pos[p + 3 * nv]};
        double y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) y[k] = s == 0 ? x[k] : displace_row(M, 1, 4, k, x);
        double* gv = S.g_verts + (int64_t)s * nv * 3;
#pragma unroll
        for (int k = 0; k < 3; k++) gv[p + k * nv] = y[k];
        const int vs1 = S.ns + 1;
        if (p / vs1 == S.nc) {
            const int i = p % vs1;
            const int64_t ws2 = (int64_t)vs1 * tab.t_steps;
            const int row = tab.t_steps - 1 - s;
#pragma unroll
            for (int k = 0; k < 4; k++) S.wake[i + (int64_t)row * vs1 + k * ws2] = y[k];
        }
        return;
    }
    const int P = r - tab.NV;
    const int m = surf_of(tab, P, 0);
    const USurf S = sim_surf(tab.s[m], so);
    const int p = P - S.off;
    const int ns = S.ns, nc = S.nc, np = ns * nc;
    const int i = p % ns, j = p / ns;
    const int64_t vs1 = ns + 1, nv = vs1 * (nc + 1);
    const double* M = T_all + ((int64_t)s * tab.n + m) * 16;
    auto vtx = [&](int ii, int jj) -> v3<double> {
        const int64_t o = ii + jj * vs1;
        const double x[4] = {S.verts_pos[o], S.verts_pos[o + nv], S.verts_pos[o + 2 * nv], S.verts_pos[o + 3 * nv]};
        if (s == 0) return {x[0], x[1], x[2]};
        return {displace_row(M, 1, 4, 0, x), displace_row(M, 1, 4, 1, x), displace_row(M, 1, 4, 2, x)};
    };
    v3<double> c, n;
    double area;
    panel_metrics(0.0, vtx(i, j), vtx(i + 1, j), vtx(i + 1, j + 1), vtx(i, j + 1), c, n, area);
    const int64_t o3 = (int64_t)s * np * 3;
    S.g_colloc[o3 + p] = c.x;
    S.g_colloc[o3 + p + np] = c.y;
    S.g_colloc[o3 + p + 2 * np] = c.z;
    S.g_normals[o3 + p] = n.x;
    S.g_normals[o3 + p + np] = n.y;
    S.g_normals[o3 + p + 2 * np] = n.z;
    S.g_areas[(int64_t)s * np + p] = area;
    const double* d = dT_all + ((int64_t)s * tab.n + m) * 16;
    const double x = S.colloc_h[p], y = S.colloc_h[p + np], z = S.colloc_h[p + 2 * np];
    double V[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        V[k] = -(d[k] * x + d[k + 4] * y + d[k + 8] * z + d[k + 12]);
        S.g_vel[o3 + p + k * np] = V[k];
    }
    g_rhs[(int64_t)s * tab.N + P] = -(V[0] * n.x + V[1] * n.y + V[2] * n.z);
}

__global__ __launch_bounds__(kPostThreads) void k_post(UTab tab, const double* __restrict__ gsrc, int cb, int step,
                                                       double inv_dt, const double* __restrict__ fs_all,
                                                       const double* __restrict__ ref_all, double rho,
                                                       double* __restrict__ gamma, double* __restrict__ cl_out,
                                                       double* __restrict__ cm_out, int64_t zs = 0) {
    __shared__ double lds[kPostThreads / 64][6];
    const int64_t o = (int64_t)blockIdx.x * zs;  // simulation blockIdx.x
    post_step<kPostThreads, true>(tab, gsrc + o, cb, step, inv_dt, fs_all + o, ref_all + o, rho, gamma + o, cl_out + o,
                                  cm_out + o, lds, o);
}

// rhs[p] = base[p] - sum over tiles of partial[tile][p] (fixed order, as k_wake_reduce)
__global__ __launch_bounds__(64 * kRedWaves) void k_wake_reduce_base(double* __restrict__ rhs,
                                                                     const double* __restrict__ base,
                                                                     const double* __restrict__ partial, int tiles,
                                                                     int total, int64_t zs = 0, int64_t wzs = 0) {
    __shared__ double red[kRedWaves][64];
    rhs += (int64_t)blockIdx.z * zs;  // simulation blockIdx.z
    base += (int64_t)blockIdx.z * zs;
    partial += (int64_t)blockIdx.z * wzs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (p < total) s = strided_sum(partial, (int64_t)total, p, wave, tiles, kRedWaves);
    red[wave][lane] = s;
    __syncthreads();
#pragma unroll
    for (int h = kRedWaves / 2; h > 0; h >>= 1) {
        if (wave < h) red[wave][lane] += red[wave + h][lane];
        __syncthreads();
    }
    if (wave == 0 && p < total) rhs[p] = base[p] - red[0][lane];
}

__global__ void k_fill64(double* x, int64_t n, double v) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) x[q] = v;
}

__global__ void k_eye(double* a, int n) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < (int64_t)n * n) a[q] = (q % n) == (q / n) ? 1.0 : 0.0;
}

inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

inline int64_t off_total(const int64_t* ns, const int64_t* nc, int n_surf) {
    int64_t n = 0;
    for (int m = 0; m < n_surf; m++) n += ns[m] * nc[m];
    return n;
}

}  // namespace
}  // namespace tvlm

using namespace tvlm;

struct tvlm_uvlm_plan {
    int32_t mode = 0;
    int32_t n_surf = 0;
    int32_t N = 0;
    int64_t t_steps = 0;
    double dt = 0.0;
    hipStream_t stream = nullptr;
    UTab tab{};
    std::vector<void*> allocs;
    double *init[kMaxS] = {};
    double *lhs = nullptr, *ainv = nullptr, *rhs = nullptr, *gamma = nullptr, *gpart = nullptr, *wpart = nullptr;
    double *cl = nullptr, *cm = nullptr, *T_all = nullptr, *dT_all = nullptr, *fs_all = nullptr, *ref_all = nullptr;
    int32_t *ipiv = nullptr, *info = nullptr;
    double* g_rhs = nullptr;  // fast mode: rhs_vel of every step, [t_steps-1][N]
    int cb = 1, cbs = 0;
    size_t wpart_elems = 0;
    bool motion_set = false;
    // simulations held by the plan (fast mode may hold several of one mesh, each with its own
    // motion): simulation z's per-simulation buffers lie z * zs doubles after simulation 0's, its
    // wake-induction partials z * wzs after
    int32_t sims = 1;
    int64_t zs = 0, wzs = 0;
    std::vector<uint8_t> sim_motion;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> wake_events;  // start/stop around the replayed wake launches
    // far-wake aggregation (fast mode): levels of age threshold far_A, row group far_G, span group far_S
    int32_t far_n = 0;
    int32_t far_A[8] = {}, far_G[8] = {}, far_S[8] = {};
    int32_t lat_target = lattice_target_groups();  // work-group target of the fast mode's wake launches

    int alloc(void** p, size_t bytes) {
        TVLM_HIP(hipMalloc(p, bytes));
        allocs.push_back(*p);
        return TVLM_OK;
    }
};

namespace tvlm {
namespace {

int plan_views(tvlm_uvlm_plan* P, int m, tvlm_view& verts, tvlm_view& colloc, tvlm_view& normals, tvlm_view& areas,
               tvlm_view& wake, int sim = 0) {
    const USurf& S = P->tab.s[m];
    const int64_t ns = S.ns, nc = S.nc, o = (int64_t)sim * P->zs;
    verts = {S.verts + o, {ns + 1, nc + 1, 4}, {1, ns + 1, (ns + 1) * (nc + 1)}};
    colloc = {S.colloc + o, {ns, nc, 3}, {1, ns, ns * nc}};
    normals = {S.normals + o, {ns, nc, 3}, {1, ns, ns * nc}};
    areas = {S.areas + o, {ns, nc, 1}, {1, ns, ns * nc}};
    wake = {S.wake + o, {ns + 1, P->t_steps, 4}, {1, ns + 1, (ns + 1) * P->t_steps}};
    return TVLM_OK;
}

// Far-wake partition of iteration `it` (ring rows [end - it, end) exist, end = t_steps - 1; the
// newest row is end - it): from the oldest row on, level l (coarsest first) takes nb[l] blocks of
// far_G[l] rows, rows [lo[l], lo[l] + nb[l] far_G[l]); a level's blocks are counted from the oldest
// row, the C = floor((it - far_A[l]) / far_G[l]) oldest of them are older than far_A[l] steps, and
// the coarser levels' blocks are taken out first.  Rows [end - it, exact_end) stay exact.
// Returns exact_end.
int far_partition(const tvlm_uvlm_plan* P, int it, int* lo, int* nb) {
    const int end = (int)P->t_steps - 1;
    int hi = end;
    for (int l = P->far_n - 1; l >= 0; l--) {
        const int A = P->far_A[l], G = P->far_G[l];
        const int C = it >= A ? (it - A) / G : 0;
        const int cov = (end - hi) / G;
        nb[l] = C > cov ? C - cov : 0;
        lo[l] = hi - nb[l] * G;
        hi = lo[l];
    }
    return hi;
}

// (receiver, wake ring) pairs of iteration `it`: exact rings plus the far levels' coarse rings.
double wake_pairs(const tvlm_uvlm_plan* P, int it) {
    int lo[8], nb[8];
    const int exact_end = far_partition(P, it, lo, nb);
    const int end = (int)P->t_steps - 1;
    double rings = 0.0;
    for (int m = 0; m < P->n_surf; m++) {
        const int ns = P->tab.s[m].ns;
        rings += (double)ns * (double)(exact_end - (end - it));
        for (int l = 0; l < P->far_n; l++) rings += (double)((ns + P->far_S[l] - 1) / P->far_S[l]) * nb[l];
    }
    return rings * (double)P->N;
}

// The wake tile table for iteration `it` (rows [t_steps-1-it, t_steps-1) of every surface; in fast
// mode with far-wake levels, the exact rows plus one coarse lattice per level that holds blocks).
void wake_tables(const tvlm_uvlm_plan* P, int it, RecvTable<double>& rt, WakeTable<double>& wt, int& tiles) {
    rt = {};
    wt = {};
    rt.n = P->n_surf;
    for (int m = 0; m < P->n_surf; m++) {
        const USurf& S = P->tab.s[m];
        const int np = S.ns * S.nc;
        rt.offset[m] = S.off;
        if (P->mode == 1)  // the step's precomputed geometry
            rt.r[m] = {S.g_colloc + (int64_t)it * np * 3, np, S.g_normals + (int64_t)it * np * 3, np, np};
        else
            rt.r[m] = {S.colloc, np, S.normals, np, np};
    }
    rt.offset[P->n_surf] = P->N;
    rt.total = P->N;
    tiles = 0;
    int lo[8], nb[8];
    const bool far = P->mode == 1 && P->far_n > 0;
    const int exact_end = far ? far_partition(P, it, lo, nb) : (int)P->t_steps - 1;
    const int recv_blocks = (P->N + kRecv - 1) / kRecv;
    for (int m = 0; m < P->n_surf; m++) {
        const USurf& S = P->tab.s[m];
        WakeSurface<double> base{};
        base.verts = S.wake;
        base.ws2 = (int64_t)(S.ns + 1) * P->t_steps;
        base.gamma = S.gamma_wake;
        base.ns = S.ns;
        base.nwv = (int32_t)P->t_steps;
        auto add = [&](const WakeSurface<double>& w0) {
            WakeSurface<double>& w = wt.s[wt.n];
            w = w0;
            if (P->mode == 1)
                lattice_tile_shape(w.ns, w.row_end - w.row_begin, recv_blocks, w.sw, w.rh);
            else
                wake_tile_shape(w.ns, w.row_end - w.row_begin, recv_blocks, w.sw, w.rh);
            w.tiles_span = (w.ns + w.sw - 1) / w.sw;
            w.tile_begin = tiles;
            tiles += w.tiles_span * ((w.row_end - w.row_begin + w.rh - 1) / w.rh);
            wt.n++;
        };
        WakeSurface<double> w = base;
        w.row_begin = (int32_t)P->t_steps - 1 - it;
        w.row_end = exact_end;
        if (w.row_end > w.row_begin || !far) add(w);
        if (!far) continue;
        for (int l = 0; l < P->far_n; l++) {
            if (nb[l] == 0) continue;
            WakeSurface<double> c = base;
            c.ns = (S.ns + P->far_S[l] - 1) / P->far_S[l];
            c.nsf = S.ns;
            c.vrs = P->far_G[l];
            c.vcs = P->far_S[l];
            c.vrow0 = lo[l];
            c.row_begin = 0;
            c.row_end = nb[l];
            add(c);
        }
    }
    wt.total_tiles = tiles;
}

int setup(tvlm_uvlm_plan* P) {
    hipStream_t st = P->stream;
    const int n = P->n_surf;
    const int64_t NN = (int64_t)P->N * P->N;
    for (int z = 0; z < P->sims; z++) {
        const int64_t o = (int64_t)z * P->zs;
        // pose at t = 0 (vlm.cpp:366-373): verts_pos = T(0) verts_init; verts = verts_pos; metrics; colloc_h
        for (int m = 0; m < n; m++) {
            USurf& S = P->tab.s[m];
            const int64_t nv = (int64_t)(S.ns + 1) * (S.nc + 1);
            tvlm_view tr = {P->T_all + o + (int64_t)m * 16, {4, 4, 1}, {1, 4, 16}};
            tvlm_view vi = {P->init[m], {S.ns + 1, S.nc + 1, 4}, {1, S.ns + 1, nv}};
            tvlm_view vp = {S.verts_pos + o, {S.ns + 1, S.nc + 1, 4}, {1, S.ns + 1, nv}};
            TVLM_TRY(tvlm_displace_wing_f64(1, &tr, &vp, &vi, st));
            TVLM_HIP(hipMemcpyAsync(S.verts + o, S.verts_pos + o, sizeof(double) * nv * 4, hipMemcpyDeviceToDevice,
                                    st));
            hipLaunchKernelGGL(k_fill64, dim3(nblk(S.ns * S.nc)), dim3(256), 0, st, S.gprev + o,
                               (int64_t)S.ns * S.nc, 0.0);
            TVLM_LAUNCHED();
        }
        std::vector<tvlm_view> V(n), C(n), Nn(n), A(n), W(n);
        for (int m = 0; m < n; m++) plan_views(P, m, V[m], C[m], Nn[m], A[m], W[m], z);
        TVLM_TRY(tvlm_mesh_metrics_f64(0.0, n, V.data(), C.data(), Nn.data(), A.data(), st));
        for (int m = 0; m < n; m++) {
            USurf& S = P->tab.s[m];
            TVLM_HIP(hipMemcpyAsync(S.colloc_h + o, S.colloc + o, sizeof(double) * S.ns * S.nc * 3,
                                    hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(k_fill64, dim3(nblk(NN)), dim3(256), 0, st, P->lhs + o, NN, 0.0);
        TVLM_LAUNCHED();
        hipLaunchKernelGGL(k_fill64, dim3(nblk(P->t_steps)), dim3(256), 0, st, P->cl + o, (int64_t)P->t_steps, 0.0);
        TVLM_LAUNCHED();
        // LHS once, iteration 0 (vlm.cpp:402)
        tvlm_view lhs = {P->lhs + o, {P->N, P->N, 1}, {1, P->N, NN}};
        TVLM_TRY(tvlm_lhs_assemble_f64(&lhs, n, C.data(), Nn.data(), V.data(), W.data(), 0, 0.0, st));
        if (P->mode == 1) {
            hipLaunchKernelGGL(k_eye, dim3(nblk(NN)), dim3(256), 0, st, P->ainv + o, P->N);
            TVLM_LAUNCHED();
        }
    }
    // the LU of every simulation's LHS at once (vlm.cpp:403), in fast mode its inverse
    tvlm_view lhs = {P->lhs, {P->N, P->N, P->sims}, {1, P->N, P->sims > 1 ? P->zs : NN}};
    TVLM_TRY(tvlm_getrf_f64(&lhs, P->ipiv, P->info, st));
    if (P->mode == 1) {
        tvlm_view x = {P->ainv, {P->N, P->N, P->sims}, {1, P->N, P->sims > 1 ? P->zs : NN}};
        TVLM_TRY(tvlm_getrs_f64(&lhs, P->ipiv, &x, st));
        const int nsteps = (int)P->t_steps - 1;
        hipLaunchKernelGGL(k_geom_all, dim3(nblk((int64_t)(P->tab.NV + P->N) * nsteps), P->sims), dim3(256), 0, st,
                           P->tab, P->T_all, P->dT_all, nsteps, P->g_rhs, P->zs);
        TVLM_LAUNCHED();
    }
    return TVLM_OK;
}

// wake-induction partials of step i (> 0) into P->wpart; tiles = partial rows written
int launch_wake(tvlm_uvlm_plan* P, int i, int& tiles) {
    hipStream_t st = P->stream;
    RecvTable<double> rt;
    WakeTable<double> wt;
    tiles = 0;
    wake_tables(P, i, rt, wt, tiles);
    TVLM_REQUIRE((size_t)tiles * P->N <= P->wpart_elems, "wake partial buffer too small");
    dim3 grid((P->N + kRecv - 1) / kRecv, tiles, P->sims);
    if (P->mode == 1) {
        int rh = 1, split = 1;
        // the launch's work-groups count every simulation's receiver blocks
        lattice_launch_shape(wt, (P->N + kRecv - 1) / kRecv * P->sims, P->lat_target, rh, split);
        int t2 = 0;
        for (int m = 0; m < wt.n; m++) {
            WakeSurface<double>& w = wt.s[m];
            w.rh = rh;
            w.sw = lattice_span_tile(w.ns, split);
            w.tiles_span = (w.ns + w.sw - 1) / w.sw;
            w.tile_begin = t2;
            t2 += w.tiles_span * ((w.row_end - w.row_begin + rh - 1) / rh);
        }
        wt.total_tiles = t2;
        tiles = t2;
        grid.y = tiles;
        TVLM_REQUIRE((size_t)tiles * P->N <= P->wpart_elems, "wake partial buffer too small");
        if (rh == 1)
            hipLaunchKernelGGL((k_wake_lattice<double, 1>), grid, dim3(kThreads), 0, st, P->wpart, rt, wt, P->zs,
                               P->wzs);
        else if (rh == 2)
            hipLaunchKernelGGL((k_wake_lattice<double, 2>), grid, dim3(kThreads), 0, st, P->wpart, rt, wt, P->zs,
                               P->wzs);
        else
            hipLaunchKernelGGL((k_wake_lattice<double, 4>), grid, dim3(kThreads), 0, st, P->wpart, rt, wt, P->zs,
                               P->wzs);
    } else
        hipLaunchKernelGGL((k_wake_partials<double, double, 0>), grid, dim3(kThreads), 0, st, P->wpart, rt, wt, 0.0);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// Fast mode step i: wake lattice + reduction into rhs = rhs_vel - wake (i > 0), gamma partials of
// A^{-1} rhs, post-processing on the step's precomputed geometry.
int enqueue_step_fast(tvlm_uvlm_plan* P, int i) {
    hipStream_t st = P->stream;
    const double* x = P->g_rhs;  // step 0 has no wake: rhs = rhs_vel
    if (i > 0) {
        int tiles = 0;
        TVLM_TRY(launch_wake(P, i, tiles));
        hipLaunchKernelGGL(k_wake_reduce_base, dim3((P->N + 63) / 64, 1, P->sims), dim3(64 * kRedWaves), 0, st,
                           P->rhs, P->g_rhs + (int64_t)i * P->N, P->wpart, tiles, P->N, P->zs, P->wzs);
        TVLM_LAUNCHED();
        x = P->rhs;
    }
    UTab t = P->tab;
    for (int m = 0; m < t.n; m++) {
        USurf& S = t.s[m];
        const int64_t np = (int64_t)S.ns * S.nc, nv = (int64_t)(S.ns + 1) * (S.nc + 1);
        S.verts = S.g_verts + i * nv * 3;
        S.vel = S.g_vel + i * np * 3;
        S.normals = S.g_normals + i * np * 3;
        S.areas = S.g_areas + i * np;
    }
    hipLaunchKernelGGL(k_gemv_part, dim3((P->N + 63) / 64, P->cb, P->sims), dim3(256), 0, st, P->ainv,
                       (int64_t)P->N, P->N, P->cbs, x, P->gpart, P->zs);
    TVLM_LAUNCHED();
    hipLaunchKernelGGL(k_post, dim3(P->sims), dim3(kPostThreads), 0, st, t, P->gpart, P->cb, i, 1.0 / P->dt,
                       P->fs_all, P->ref_all, 1.0, P->gamma, P->cl, P->cm, P->zs);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

int enqueue_step(tvlm_uvlm_plan* P, int i) {
    if (P->mode == 1) return enqueue_step_fast(P, i);
    hipStream_t st = P->stream;
    UTab& tab = P->tab;
    if (i == 0) {  // later steps start inside the previous step's k_advance
        hipLaunchKernelGGL(k_pre, dim3(nblk(tab.N + tab.NTE)), dim3(256), 0, st, tab, P->dT_all, i, P->rhs);
        TVLM_LAUNCHED();
    }
    if (i > 0) {
        int tiles = 0;
        TVLM_TRY(launch_wake(P, i, tiles));
        launch_wake_reduce<double>(P->rhs, 1, P->wpart, tiles, P->N, st);
        TVLM_LAUNCHED();
    }
    const double* gsrc;
    int cb;
    if (P->mode == 1) {
        hipLaunchKernelGGL(k_gemv_part, dim3((P->N + 63) / 64, P->cb), dim3(256), 0, st, P->ainv, (int64_t)P->N, P->N,
                           P->cbs, P->rhs, P->gpart);
        TVLM_LAUNCHED();
        gsrc = P->gpart;
        cb = P->cb;
    } else {
        tvlm_view lhs = {P->lhs, {P->N, P->N, 1}, {1, P->N, (int64_t)P->N * P->N}};
        tvlm_view x = {P->rhs, {P->N, 1, 1}, {1, P->N, P->N}};
        TVLM_TRY(tvlm_getrs_f64(&lhs, P->ipiv, &x, st));
        gsrc = P->rhs;
        cb = 1;
    }
    const int do_pre = (i + 1 < (int)P->t_steps - 1) ? 1 : 0;
    hipLaunchKernelGGL(k_post_advance, dim3(1), dim3(kPostThreads), 0, st, tab, gsrc, cb, i, 1.0 / P->dt, P->fs_all,
                       P->ref_all, 1.0, P->gamma, P->cl, P->cm, P->T_all, P->dT_all, do_pre, P->rhs);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

extern "C" {

int tvlm_uvlm_create_batched(tvlm_uvlm_plan** out, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                             const double* const* verts_init, int64_t t_steps, double dt, int32_t mode,
                             int32_t sims, void* stream) {
    TVLM_REQUIRE(out && ns && nc && verts_init, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    TVLM_REQUIRE(t_steps >= 2 && t_steps < (1 << 24), "t_steps %lld out of range", (long long)t_steps);
    TVLM_REQUIRE(dt > 0.0, "dt must be positive");
    TVLM_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (plain FP64) or 1 (FP32 wake, inverse solve, graph)");
    TVLM_REQUIRE(sims >= 1 && sims <= 65535, "sims %d out of range", sims);
    TVLM_REQUIRE(sims == 1 || mode == 1, "several simulations per plan need mode 1");
    auto* P = new tvlm_uvlm_plan();
    P->mode = mode;
    P->n_surf = n_surf;
    P->t_steps = t_steps;
    P->dt = dt;
    P->sims = sims;
    P->sim_motion.assign((size_t)sims, 0);
    P->stream = as_stream(stream);
    UTab& tab = P->tab;
    tab.n = n_surf;
    tab.t_steps = (int32_t)t_steps;
    int off = 0, voff = 0, teoff = 0;
    int rc = TVLM_OK;
    auto fail = [&](int code) {
        for (void* p : P->allocs) (void)hipFree(p);
        delete P;
        return code;
    };
    size_t max_tiles = 0;
    for (int m = 0; m < n_surf; m++) {
        if (ns[m] < 1 || nc[m] < 1 || !verts_init[m]) {
            set_error("surface %d: bad dimensions", m);
            return fail(TVLM_ERR_INVALID);
        }
        USurf& S = tab.s[m];
        S.ns = (int32_t)ns[m];
        S.nc = (int32_t)nc[m];
        S.off = off;
        S.v_off = voff;
        S.te_off = teoff;
        off += (int)(ns[m] * nc[m]);
        voff += (int)((ns[m] + 1) * (nc[m] + 1));
        teoff += 4 * (int)(ns[m] + 1);
        const int sw = ns[m] < kMaxSpanTile ? (int)ns[m] : kMaxSpanTile;
        max_tiles += (size_t)((ns[m] + sw - 1) / sw) * (size_t)t_steps;  // worst case: one row per tile
    }
    {
        // fast mode: a launch whose spans are split holds fewer than 2 kLatTargetGroups work-groups
        const size_t recv_blocks = (size_t)((off_total(ns, nc, n_surf) + kRecv - 1) / kRecv);
        const size_t split_tiles = 2 * (size_t)lattice_target_groups() / recv_blocks + (size_t)kMaxWakes;
        if (mode == 1 && split_tiles > max_tiles) max_tiles = split_tiles;
    }
    tab.N = off;
    tab.NV = voff;
    tab.NTE = teoff;
    P->N = off;
    const size_t N = (size_t)off;
    P->cbs = 64;
    P->cb = (int)((N + P->cbs - 1) / P->cbs);
    P->wpart_elems = max_tiles * N;
    // every per-simulation buffer: one block per simulation (256-byte aligned parts)
    const size_t st1 = (size_t)(t_steps - 1), ts = (size_t)t_steps;
    std::vector<std::pair<double**, size_t>> lay;
    for (int m = 0; m < n_surf; m++) {
        USurf& S = tab.s[m];
        const size_t np = (size_t)ns[m] * nc[m], nv = (size_t)(ns[m] + 1) * (nc[m] + 1);
        for (double** b3 : {&S.colloc, &S.colloc_h, &S.normals, &S.vel, &S.forces}) lay.push_back({b3, np * 3});
        for (double** b1 : {&S.areas, &S.gprev, &S.gdelta, &S.gdt}) lay.push_back({b1, np});
        lay.push_back({&S.verts_pos, nv * 4});
        lay.push_back({&S.verts, nv * 4});
        lay.push_back({&S.wake, (size_t)(ns[m] + 1) * ts * 4});
        lay.push_back({&S.gamma_wake, (size_t)ns[m] * st1});
        if (mode == 1) {
            lay.push_back({&S.g_verts, nv * 3 * st1});
            for (double** g3 : {&S.g_colloc, &S.g_normals, &S.g_vel}) lay.push_back({g3, np * 3 * st1});
            lay.push_back({&S.g_areas, np * st1});
        }
    }
    lay.push_back({&P->lhs, N * N});
    if (mode == 1) lay.push_back({&P->ainv, N * N});
    lay.push_back({&P->rhs, N});
    lay.push_back({&P->gamma, N});
    lay.push_back({&P->gpart, N * (size_t)P->cb});
    lay.push_back({&P->cl, ts});
    lay.push_back({&P->cm, 3 * ts});
    lay.push_back({&P->T_all, 16 * (size_t)n_surf * ts});
    lay.push_back({&P->dT_all, 16 * (size_t)n_surf * ts});
    lay.push_back({&P->fs_all, 3 * ts});
    lay.push_back({&P->ref_all, 3 * ts});
    if (mode == 1) lay.push_back({&P->g_rhs, N * st1});
    std::vector<size_t> where(lay.size());
    size_t blk = 0;
    for (size_t k = 0; k < lay.size(); k++) {
        where[k] = blk;
        blk += (lay[k].second + 31) / 32 * 32;
    }
    P->zs = (int64_t)blk;
    double* base = nullptr;
    if ((rc = P->alloc((void**)&base, sizeof(double) * blk * (size_t)sims)) != TVLM_OK) return fail(rc);
    for (size_t k = 0; k < lay.size(); k++) *lay[k].first = base + where[k];
    P->wzs = (int64_t)((P->wpart_elems + 31) / 32 * 32);  // every simulation's partials 256-byte aligned
    if ((rc = P->alloc((void**)&P->wpart, sizeof(double) * (size_t)P->wzs * (size_t)sims)) != TVLM_OK)
        return fail(rc);
    if ((rc = P->alloc((void**)&P->ipiv, sizeof(int32_t) * N * (size_t)sims)) != TVLM_OK) return fail(rc);
    if ((rc = P->alloc((void**)&P->info, sizeof(int32_t) * (size_t)sims)) != TVLM_OK) return fail(rc);
    for (int m = 0; m < n_surf; m++) {
        const size_t nv = (size_t)(ns[m] + 1) * (nc[m] + 1);
        if ((rc = P->alloc((void**)&P->init[m], sizeof(double) * nv * 4)) != TVLM_OK) return fail(rc);
        if (hipMemcpyAsync(P->init[m], verts_init[m], sizeof(double) * nv * 4, hipMemcpyHostToDevice, P->stream) !=
            hipSuccess) {
            set_error("upload of surface %d failed", m);
            return fail(TVLM_ERR_HIP);
        }
    }
    if (hipStreamSynchronize(P->stream) != hipSuccess) {
        set_error("stream synchronize failed");
        return fail(TVLM_ERR_HIP);
    }
    *out = P;
    return TVLM_OK;
}

int tvlm_uvlm_create(tvlm_uvlm_plan** out, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                     const double* const* verts_init, int64_t t_steps, double dt, int32_t mode, void* stream) {
    return tvlm_uvlm_create_batched(out, n_surf, ns, nc, verts_init, t_steps, dt, mode, 1, stream);
}

// T, dT: [t_steps][n_surf][16] column-major 4x4 value and time derivative of each surface transform
// at t_i = i dt; freestream, ref_pt: [t_steps][3] (the CM reference point moves with surface 0).
int tvlm_uvlm_set_motion_sim(tvlm_uvlm_plan* P, int32_t sim, const double* T, const double* dT,
                             const double* freestream, const double* ref_pt) {
    TVLM_REQUIRE(P && T && dT && freestream && ref_pt, "null argument");
    TVLM_REQUIRE(sim >= 0 && sim < P->sims, "simulation %d of %d", sim, P->sims);
    const size_t n16 = sizeof(double) * 16 * P->n_surf * P->t_steps, n3 = sizeof(double) * 3 * P->t_steps;
    const int64_t o = (int64_t)sim * P->zs;
    TVLM_HIP(hipMemcpyAsync(P->T_all + o, T, n16, hipMemcpyHostToDevice, P->stream));
    TVLM_HIP(hipMemcpyAsync(P->dT_all + o, dT, n16, hipMemcpyHostToDevice, P->stream));
    TVLM_HIP(hipMemcpyAsync(P->fs_all + o, freestream, n3, hipMemcpyHostToDevice, P->stream));
    TVLM_HIP(hipMemcpyAsync(P->ref_all + o, ref_pt, n3, hipMemcpyHostToDevice, P->stream));
    TVLM_HIP(hipStreamSynchronize(P->stream));
    P->sim_motion[(size_t)sim] = 1;
    P->motion_set = std::all_of(P->sim_motion.begin(), P->sim_motion.end(), [](uint8_t v) { return v != 0; });
    return TVLM_OK;
}

int tvlm_uvlm_set_motion(tvlm_uvlm_plan* P, const double* T, const double* dT, const double* freestream,
                         const double* ref_pt) {
    TVLM_REQUIRE(P, "null plan");
    for (int32_t z = 0; z < P->sims; z++) TVLM_TRY(tvlm_uvlm_set_motion_sim(P, z, T, dT, freestream, ref_pt));
    return TVLM_OK;
}

// One full simulation: setup (pose, LHS, LU[, inverse]) then steps 0 .. t_steps-2.  use_graph
// replays the time loop as one captured hipGraph (captured on the first call).  Asynchronous.
int tvlm_uvlm_run(tvlm_uvlm_plan* P, int32_t use_graph) {
    TVLM_REQUIRE(P && P->motion_set, "plan has no motion");
    TVLM_TRY(setup(P));
    const int last = (int)P->t_steps - 1;
    if (!use_graph) {
        for (int i = 0; i < last; i++) TVLM_TRY(enqueue_step(P, i));
        return TVLM_OK;
    }
    TVLM_REQUIRE(P->stream != nullptr, "hipGraph capture needs a non-default stream");
    // the plain loop's per-step getrs uses the library's per-stream scratch, which a later, larger
    // solve on the same stream may reallocate under a captured graph: plain mode is never captured
    TVLM_REQUIRE(P->mode == 1, "use_graph needs mode 1 (plain mode runs call by call)");
    if (!P->exec) {
        TVLM_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
        int rc = TVLM_OK;
        for (int i = 0; i < last && rc == TVLM_OK; i++) rc = enqueue_step(P, i);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(P->stream, &g);
        if (rc != TVLM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        TVLM_HIP(e);
        P->graph = g;
        TVLM_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    }
    TVLM_HIP(hipGraphLaunch(P->exec, P->stream));
    return TVLM_OK;
}

// One simulation without graph capture, with HIP events recorded on the plan's stream around every
// wake-induction launch; returns the summed kernel time, the number of launches and the number of
// (receiver, wake ring) pairs they evaluated.  Synchronous.
int tvlm_uvlm_time_wake(tvlm_uvlm_plan* P, double* total_ms, int32_t* launches, double* pairs) {
    TVLM_REQUIRE(P && P->motion_set && total_ms && launches && pairs, "bad argument");
    const int last = (int)P->t_steps - 1;
    if (P->wake_events.empty()) {
        P->wake_events.resize(2, nullptr);
        for (auto& e : P->wake_events) TVLM_HIP(hipEventCreate(&e));
    }
    // one whole simulation fills the wake; then every step's wake launch (same grid and operand
    // sizes as in the loop) is replayed back to back between two events
    TVLM_TRY(setup(P));
    for (int i = 0; i < last; i++) TVLM_TRY(enqueue_step(P, i));
    TVLM_HIP(hipEventRecord(P->wake_events[0], P->stream));
    for (int i = 1; i < last; i++) {
        int tiles = 0;
        TVLM_TRY(launch_wake(P, i, tiles));
    }
    TVLM_HIP(hipEventRecord(P->wake_events[1], P->stream));
    TVLM_HIP(hipStreamSynchronize(P->stream));
    double pr = 0.0;
    for (int i = 1; i < last; i++) pr += wake_pairs(P, i) * (double)P->sims;
    float ms = 0.f;
    TVLM_HIP(hipEventElapsedTime(&ms, P->wake_events[0], P->wake_events[1]));
    *total_ms = ms;
    *launches = last - 1;
    *pairs = pr;
    return TVLM_OK;
}

// Far-wake aggregation of the fast mode's wake induction (see far_partition): n_levels (0: exact
// wake) triples {age threshold, row group, span group}, ages increasing, each row group a multiple
// of the one before.  Drops a captured graph (its launches change).
int tvlm_uvlm_set_far_wake(tvlm_uvlm_plan* P, int32_t n_levels, const int32_t* levels) {
    TVLM_REQUIRE(P, "null plan");
    TVLM_REQUIRE(P->mode == 1, "far-wake aggregation needs mode 1 (the plain mode is the exact reference)");
    TVLM_REQUIRE(n_levels >= 0 && n_levels <= 8, "n_levels %d out of range [0, 8]", n_levels);
    TVLM_REQUIRE(n_levels == 0 || levels, "null levels");
    TVLM_REQUIRE(P->n_surf * (1 + n_levels) <= tvlm::kMaxWakes, "%d surfaces x (1 + %d levels) exceed %d wake tables",
                 P->n_surf, n_levels, tvlm::kMaxWakes);
    for (int l = 0; l < n_levels; l++) {
        const int32_t A = levels[3 * l], G = levels[3 * l + 1], S = levels[3 * l + 2];
        TVLM_REQUIRE(A >= 0 && G >= 1 && G <= (1 << 20) && S >= 1, "level %d: bad {age %d, rows %d, spans %d}", l, A,
                     G, S);
        if (l > 0)
            TVLM_REQUIRE(A > levels[3 * (l - 1)] && G % levels[3 * (l - 1) + 1] == 0,
                         "level %d: ages must increase and row groups nest", l);
    }
    if (P->exec) {
        TVLM_HIP(hipStreamSynchronize(P->stream));
        TVLM_HIP(hipGraphExecDestroy(P->exec));
        P->exec = nullptr;
    }
    if (P->graph) {
        TVLM_HIP(hipGraphDestroy(P->graph));
        P->graph = nullptr;
    }
    P->far_n = n_levels;
    for (int l = 0; l < n_levels; l++) {
        P->far_A[l] = levels[3 * l];
        P->far_G[l] = levels[3 * l + 1];
        P->far_S[l] = levels[3 * l + 2];
    }
    return TVLM_OK;
}

// Work-group target of the fast mode's wake launches (lattice_launch_shape); grows the tile
// partial buffer when a larger target needs it.  Drops a captured graph.
int tvlm_uvlm_set_lattice_target(tvlm_uvlm_plan* P, int32_t groups) {
    TVLM_REQUIRE(P, "null plan");
    TVLM_REQUIRE(P->mode == 1, "the lattice target applies to mode 1");
    TVLM_REQUIRE(groups >= 64 && groups <= 65536, "groups %d out of range [64, 65536]", groups);
    const size_t recv_blocks = (size_t)((P->N + kRecv - 1) / kRecv);
    const size_t need = (2 * (size_t)groups / recv_blocks + (size_t)kMaxWakes) * (size_t)P->N;
    if (P->exec) {
        TVLM_HIP(hipStreamSynchronize(P->stream));
        TVLM_HIP(hipGraphExecDestroy(P->exec));
        P->exec = nullptr;
    }
    if (P->graph) {
        TVLM_HIP(hipGraphDestroy(P->graph));
        P->graph = nullptr;
    }
    if (need > P->wpart_elems) {
        TVLM_HIP(hipStreamSynchronize(P->stream));
        double* w = nullptr;
        const size_t wzs = (need + 31) / 32 * 32;
        TVLM_TRY(P->alloc((void**)&w, sizeof(double) * wzs * (size_t)P->sims));
        P->wpart = w;  // the smaller buffer stays in the plan's allocations until destroy
        P->wpart_elems = need;
        P->wzs = (int64_t)wzs;
    }
    P->lat_target = groups;
    return TVLM_OK;
}

// Copies (synchronously) CL[t_steps-1], CM[t_steps-1][3] (step 0 entries are 0), the last
// circulation gamma[N] and the LU info; any pointer may be null.  The copies are queued on the
// plan's stream (not the null stream), so they do not wait for work on other streams.
int tvlm_uvlm_results_sim(tvlm_uvlm_plan* P, int32_t sim, double* cl, double* cm, double* gamma, int32_t* info) {
    TVLM_REQUIRE(P, "null plan");
    TVLM_REQUIRE(sim >= 0 && sim < P->sims, "simulation %d of %d", sim, P->sims);
    hipStream_t st = P->stream;
    const size_t n = (size_t)(P->t_steps - 1);
    const int64_t o = (int64_t)sim * P->zs;
    if (cl) TVLM_HIP(hipMemcpyAsync(cl, P->cl + o, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (cm) TVLM_HIP(hipMemcpyAsync(cm, P->cm + o, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    if (gamma) TVLM_HIP(hipMemcpyAsync(gamma, P->gamma + o, sizeof(double) * P->N, hipMemcpyDeviceToHost, st));
    if (info) TVLM_HIP(hipMemcpyAsync(info, P->info + sim, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipStreamSynchronize(st));
    return TVLM_OK;
}

int tvlm_uvlm_results(tvlm_uvlm_plan* P, double* cl, double* cm, double* gamma, int32_t* info) {
    return tvlm_uvlm_results_sim(P, 0, cl, cm, gamma, info);
}

// Prescribed freestream + heave + pitch chain shared by all surfaces, evaluated with dual numbers
// on the host for every step (vlm/include/vlm_kinematics.hpp:149-239, headeronly/tinyad.hpp):
//   fs    = translation(-u t, 0, 0)
//   heave = translation(0, 0, h_amp sin(omega t + h_phase))        after fs
//   pitch = rotation((x_pivot,0,0), (0,1,0), a_amp sin(omega t + a_phase))  after heave
// T_all[0] = T(0), T_all[k] = T((k-1) dt + dt) (the t + dt of vlm.cpp:500); dT_all[i] = dT(i dt);
// freestream = (u, 0, 0); CM reference = T(t) (0.5, 0, 0) (vlm.cpp:464-466).
int tvlm_motion_pitch_heave(double u, double h_amp, double h_phase, double a_amp, double a_phase, double omega,
                            double x_pivot, int32_t n_surf, int64_t t_steps, double dt, double* T_out, double* dT_out,
                            double* fs_out, double* ref_out) {
    TVLM_REQUIRE(T_out && dT_out && fs_out && ref_out && n_surf >= 1 && t_steps >= 1, "bad argument");
    struct D {
        double v, d;
    };
    auto mul = [](D a, D b) { return D{a.v * b.v, b.v * a.d + a.v * b.d}; };
    auto add = [](D a, D b) { return D{a.v + b.v, a.d + b.d}; };
    struct M4 {
        D m[16];  // column-major
    };
    auto ident = []() {
        M4 r;
        for (int k = 0; k < 16; k++) r.m[k] = D{(k % 5) == 0 ? 1.0 : 0.0, 0.0};
        return r;
    };
    auto mm = [&](const M4& a, const M4& b) {
        M4 r;
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) {
                D acc = mul(a.m[0 * 4 + i], b.m[j * 4 + 0]);
                for (int k = 1; k < 4; k++) acc = add(acc, mul(a.m[k * 4 + i], b.m[j * 4 + k]));
                r.m[j * 4 + i] = acc;
            }
        return r;
    };
    auto eval = [&](double t, double* T, double* dT) {
        M4 fs = ident(), hv = ident(), pt = ident();
        fs.m[12] = D{-u * t, -u};
        hv.m[14] = D{h_amp * sin(omega * t + h_phase), h_amp * omega * cos(omega * t + h_phase)};
        const D a = {a_amp * sin(omega * t + a_phase), a_amp * omega * cos(omega * t + a_phase)};
        const D sn = {sin(a.v), cos(a.v) * a.d}, cs = {cos(a.v), -sin(a.v) * a.d};
        pt.m[0] = cs;
        pt.m[2] = D{-sn.v, -sn.d};
        pt.m[8] = sn;
        pt.m[10] = cs;
        pt.m[12] = D{(1.0 - cs.v) * x_pivot, -cs.d * x_pivot};
        pt.m[14] = D{sn.v * x_pivot, sn.d * x_pivot};
        const M4 r = mm(fs, mm(hv, pt));
        for (int k = 0; k < 16; k++) {
            T[k] = r.m[k].v;
            dT[k] = r.m[k].d;
        }
    };
    const int64_t ts = t_steps;
    const int n = n_surf;
    double *T = T_out, *dT = dT_out, *fs = fs_out, *ref = ref_out;
    double t16[16], d16[16];
    for (int64_t k = 0; k < ts; k++) {
        const double ti = (double)k * dt;
        // value for displacement: T(0), then T(t_{k-1} + dt)
        eval(k == 0 ? 0.0 : (double)(k - 1) * dt + dt, t16, d16);
        for (int m = 0; m < n; m++) std::memcpy(&T[(k * n + m) * 16], t16, sizeof t16);
        eval(ti, t16, d16);
        for (int m = 0; m < n; m++) std::memcpy(&dT[(k * n + m) * 16], d16, sizeof d16);
        fs[3 * k] = u;
        fs[3 * k + 1] = 0.0;
        fs[3 * k + 2] = 0.0;
        for (int r = 0; r < 3; r++) ref[3 * k + r] = t16[r] * 0.5 + t16[r + 12];
    }
    return TVLM_OK;
}

int tvlm_uvlm_set_pitch_heave_sim(tvlm_uvlm_plan* P, int32_t sim, double u, double h_amp, double h_phase,
                                  double a_amp, double a_phase, double omega, double x_pivot) {
    TVLM_REQUIRE(P, "null plan");
    TVLM_REQUIRE(sim >= -1 && sim < P->sims, "simulation %d of %d", sim, P->sims);
    const int64_t ts = P->t_steps;
    std::vector<double> T(16 * P->n_surf * ts), dT(16 * P->n_surf * ts), fs(3 * ts), ref(3 * ts);
    TVLM_TRY(tvlm_motion_pitch_heave(u, h_amp, h_phase, a_amp, a_phase, omega, x_pivot, P->n_surf, ts, P->dt, T.data(),
                                     dT.data(), fs.data(), ref.data()));
    if (sim < 0) return tvlm_uvlm_set_motion(P, T.data(), dT.data(), fs.data(), ref.data());
    return tvlm_uvlm_set_motion_sim(P, sim, T.data(), dT.data(), fs.data(), ref.data());
}

int tvlm_uvlm_set_pitch_heave(tvlm_uvlm_plan* P, double u, double h_amp, double h_phase, double a_amp,
                              double a_phase, double omega, double x_pivot) {
    return tvlm_uvlm_set_pitch_heave_sim(P, -1, u, h_amp, h_phase, a_amp, a_phase, omega, x_pivot);
}

int tvlm_uvlm_destroy(tvlm_uvlm_plan* P) {
    if (!P) return TVLM_OK;
    (void)hipStreamSynchronize(P->stream);
    for (auto e : P->wake_events)
        if (e) (void)hipEventDestroy(e);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    for (void* p : P->allocs) (void)hipFree(p);
    delete P;
    return TVLM_OK;
}

}  // extern "C"

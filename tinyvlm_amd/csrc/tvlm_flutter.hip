// tvlm_flutter.hip -- the coupled pitch/heave aeroelastic driver (the reference's UVLM_2DOF,
// tests/uvlm_2dof.cpp:428-865, with its Newmark-beta integrator, vlm/src/vlm_integrator.cpp:26-68)
// for many airspeeds of one wing at once, kept on the device: a flutter sweep.
//
// The wing moves rigidly, so its influence matrix is assembled once at the initial pose, factored
// and inverted once (shared by every airspeed).  Each airspeed b carries its own structural state,
// pose, wake and circulations; the airspeeds advance in lockstep, one time step at a time, and each
// step runs the reference's fixed-point (Newton) iterations on the structural increment:
//   k_fl_struct : Newmark step from the load increment (2 x 2 algebra), trial state, the section's
//                 pose T and its time derivative in closed form (freestream, heave, pitch about the
//                 elastic axis), the moment reference point
//   k_fl_geom   : wing vertices at the pose, panel metrics, trailing edge shed into the wake ring,
//                 kinematic velocities at the initial collocation points, rhs = -(V . n)
//   k_fl_wake   : wake induction on the vortex lattice in the FP64 seeded form (the served UVLM
//                 loop's arithmetic), over the newest R ring rows only (the wake is a ring buffer of
//                 R + 1 vertex rows and R + 1 circulation rows: older rings are dropped)
//   k_fl_reduce : rhs -= the tile partials, fixed order
//   k_fl_gemm   : gamma = A^{-1} rhs for all airspeeds at once (each A^{-1} block read once per
//                 8 airspeeds)
//   k_fl_post   : circulation differences, dGamma/dt, the wake's newest circulation row, unsteady
//                 forces, CL and CM about the elastic axis, the structural loads and the next load
//                 increment, the convergence test |du_prev - du| <= EPS_f (at most 51 iterations)
// An airspeed that has converged skips the rest of the step's iterations (every kernel returns
// early for it); the host reads the per-airspeed flags after each iteration and commits the step
// (k_fl_commit) once all have converged.
// All arithmetic is FP64; Biot-Savart in the seeded form (FP64 with the hardware 1/sqrt and 1/x
// estimates refined by one Newton step).
#include <cmath>
#include <cstring>
#include <vector>

#include "tvlm_kernels.hpp"

namespace tvlm {
namespace {

constexpr int kFlMaxIter = 50;                                 // iterations stop once it > 50 (uvlm_2dof.cpp:637, :795)
constexpr double kFlEps = (double)1.19209e-07f;                // EPS_f (vlm_types.hpp:24)
constexpr double kPiF = (double)3.141593f;                     // PI_f (vlm_types.hpp:21)
constexpr int kFlChunk = 8;                                    // airspeeds per k_fl_gemm work-group
constexpr int kFlGemmCols = 512;                               // columns per k_fl_gemm work-group
constexpr int kFlPostThreads = 1024;

struct FlConst {
    double dt, inv_dt, a_h, rho, mu, r_a;
    double x0, x1, xd0, xd1, xdd0, xdd1;  // Newmark-beta coefficients (vlm_integrator.cpp:26-44)
    double M[4];                          // mass matrix, column-major
    double beta[4];                       // torsional spring coefficients
    int32_t N, ns, nc, slots;             // panels, spanwise / chordwise panels, ring slots R + 1
    int32_t B;
};

struct FlAir {
    double C[4];     // damping matrix
    double Kinv[4];  // inverse of K_eff = K + x0 M + x1 C (K = 0: the stiffness is on the load side)
    double kh, ka;   // (omega / U_a)^2, 1 / U_a^2
};

struct FlState {
    double u[2], v[2], a[2];     // committed state (step i)
    double u1[2], v1[2], a1[2];  // trial state (step i + 1)
    double du[2], du_k[2], dF[2];
    double F[2], F1[2];
    double T[16], dT[16], ref[3];
    int32_t it, done;
};

// per-airspeed device buffers: base + b * stride (strides in elements)
struct FlBufs {
    const double* vinit;   // {ns+1, nc+1, 4} dimensionless wing, shared
    const double* colloc_h;  // {ns, nc, 3} collocation points of the initial pose, shared
    double *verts, *colloc, *normals, *areas, *vel, *rhs_vel, *rhs, *gamma, *gprev;
    double *wverts, *wgam;  // wake rings: vertex rows [slots][ns+1][3 planes], circulations [slots][ns]
    int64_t sv, sp3, sp, swv, swg;  // strides: verts (3 planes), panels x 3, panels, wake vertex ring, wake gamma ring
};

__device__ __forceinline__ double torsional(const FlConst& c, double al) {
    return c.beta[0] + c.beta[1] * al + c.beta[2] * al * al + c.beta[3] * al * al * al;
}

// Newmark step and pose of every airspeed still iterating (one thread each)
__global__ void k_fl_struct(FlConst c, const FlAir* __restrict__ air, FlState* __restrict__ st, int step) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= c.B) return;
    FlState& s = st[b];
    if (s.done) return;
    const FlAir& A = air[b];
    s.du_k[0] = s.du[0];
    s.du_k[1] = s.du[1];
    // K_eff du = dF + M (xd0 v + xdd0 a) + C (xd1 v + xdd1 a)  (vlm_integrator.cpp:46-68)
    const double p0 = c.xd0 * s.v[0] + c.xdd0 * s.a[0], p1 = c.xd0 * s.v[1] + c.xdd0 * s.a[1];
    const double q0 = c.xd1 * s.v[0] + c.xdd1 * s.a[0], q1 = c.xd1 * s.v[1] + c.xdd1 * s.a[1];
    const double r0 = s.dF[0] + (c.M[0] * p0 + c.M[2] * p1) + (A.C[0] * q0 + A.C[2] * q1);
    const double r1 = s.dF[1] + (c.M[1] * p0 + c.M[3] * p1) + (A.C[1] * q0 + A.C[3] * q1);
    double du[2], dv[2], da[2];
    du[0] = A.Kinv[0] * r0 + A.Kinv[2] * r1;
    du[1] = A.Kinv[1] * r0 + A.Kinv[3] * r1;
    for (int k = 0; k < 2; k++) {
        dv[k] = c.x1 * du[k] - c.xd1 * s.v[k] - c.xdd1 * s.a[k];
        da[k] = c.x0 * du[k] - c.xd0 * s.v[k] - c.xdd0 * s.a[k];
        s.du[k] = du[k];
        s.u1[k] = s.u[k] + du[k];
        s.v1[k] = s.v[k] + dv[k];
        s.a1[k] = s.a[k] + da[k];
    }
    // pose F(t) H P at t = (i + 1) dt: freestream translation (-t, 0, 0), heave (0, 0, -h - h' tau),
    // pitch about (1 + a_h, 0, 0), axis +y, by alpha + alpha' tau, and its tau-derivative at tau = 0
    // (tests/uvlm_2dof.cpp:645-680); column-major
    const double t = (double)step * c.dt + c.dt;
    const double h = s.u1[0], hd = s.v1[0], al = s.u1[1], ald = s.v1[1];
    const double px = 1.0 + c.a_h;
    const double cs = cos(al), sn = sin(al);
    double* T = s.T;
    double* D = s.dT;
    for (int k = 0; k < 16; k++) {
        T[k] = 0.0;
        D[k] = 0.0;
    }
    T[0] = cs;
    T[2] = -sn;
    T[5] = 1.0;
    T[8] = sn;
    T[10] = cs;
    T[12] = px - cs * px - t;
    T[14] = sn * px - h;
    T[15] = 1.0;
    D[0] = -sn * ald;
    D[2] = -cs * ald;
    D[8] = cs * ald;
    D[10] = -sn * ald;
    D[12] = sn * px * ald - 1.0;
    D[14] = cs * px * ald - hd;
    s.ref[0] = T[0] * px + T[12];
    s.ref[1] = 0.0;
    s.ref[2] = T[2] * px + T[14];
}

// wing vertices, metrics, trailing-edge shed, velocities and rhs of airspeed blockIdx.y
__global__ void k_fl_geom(FlConst c, FlBufs d, const FlState* __restrict__ st, int step) {
    const int b = blockIdx.y;
    const FlState& s = st[b];
    if (s.done) return;
    const int ns = c.ns, nc = c.nc, np = c.N;
    const int64_t vs1 = ns + 1, nv = vs1 * (nc + 1);
    const int NV = (int)nv;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const double* M = s.T;
    if (e < NV) {
        const double x[4] = {d.vinit[e], d.vinit[e + nv], d.vinit[e + 2 * nv], d.vinit[e + 3 * nv]};
        double* V = d.verts + b * d.sv;
        double y[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            y[r] = displace_row(M, 1, 4, r, x);
            V[e + r * nv] = y[r];
        }
        if (e / vs1 == nc) {  // trailing edge -> the wake ring's vertex row of step i + 1 (wake_shed)
            const int i = e % vs1;
            double* W = d.wverts + b * d.swv + (int64_t)((step + 1) % c.slots) * vs1 * 3;
#pragma unroll
            for (int r = 0; r < 3; r++) W[i + r * vs1] = y[r];
        }
        return;
    }
    const int p = e - NV;
    if (p >= np) return;
    const int i = p % ns, j = p / ns;
    auto vtx = [&](int ii, int jj) -> v3<double> {
        const int64_t o = ii + jj * vs1;
        const double x[4] = {d.vinit[o], d.vinit[o + nv], d.vinit[o + 2 * nv], d.vinit[o + 3 * nv]};
        return {displace_row(M, 1, 4, 0, x), displace_row(M, 1, 4, 1, x), displace_row(M, 1, 4, 2, x)};
    };
    v3<double> cl, n;
    double area;
    panel_metrics(0.0, vtx(i, j), vtx(i + 1, j), vtx(i + 1, j + 1), vtx(i, j + 1), cl, n, area);
    double* C = d.colloc + b * d.sp3;
    double* Nn = d.normals + b * d.sp3;
    C[p] = cl.x;
    C[p + np] = cl.y;
    C[p + 2 * np] = cl.z;
    Nn[p] = n.x;
    Nn[p + np] = n.y;
    Nn[p + 2 * np] = n.z;
    d.areas[b * d.sp + p] = area;
    // velocity -(dT [x_h; 1]) at the initial pose's collocation point (tests/uvlm_2dof.cpp:316-330, 684-686)
    const double* D = s.dT;
    const double x = d.colloc_h[p], y = d.colloc_h[p + np], z = d.colloc_h[p + 2 * np];
    double U[3];
    double* Vel = d.vel + b * d.sp3;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        U[r] = -(D[r] * x + D[r + 4] * y + D[r + 8] * z + D[r + 12]);
        Vel[p + r * np] = U[r];
    }
    d.rhs_vel[b * d.sp + p] = -(U[0] * n.x + U[1] * n.y + U[2] * n.z);
}

// Wake tile shape: spans per tile (<= 4 x kLatChunkMax) and rows per tile (1, 2 or 4): the largest
// rows per tile that still gives one airspeed at least kFlTargetGroups work-groups.  It does not
// depend on how many airspeeds run together, so an airspeed's result is the same alone and in a
// batch, bit for bit.
constexpr int kFlTargetGroups = 1024;

inline void fl_tile_shape(int ns, int rows, int recv_blocks, int& sw, int& rh, int& tiles) {
    sw = ns < kWaves * kLatChunkMax ? ns : kWaves * kLatChunkMax;
    const int tiles_span = (ns + sw - 1) / sw;
    rh = 4;
    while (rh > 1 && (int64_t)recv_blocks * tiles_span * ((rows + rh - 1) / rh) < kFlTargetGroups) rh /= 2;
    tiles = tiles_span * ((rows + rh - 1) / rh);
}

// Wake induction of airspeed blockIdx.z over the lattice of its newest `rows` ring rows.  Logical
// vertex row l (0 = the trailing edge shed at step i + 1, l = the row shed l steps earlier) lives in
// ring slot (i + 1 - l) mod slots; logical ring row q (between vertex rows q and q + 1) carries the
// circulation shed at step i - q, in slot (i - q) mod slots.  Same tiling and arithmetic as
// k_wake_lattice.
template <int RH>
__global__ __launch_bounds__(kThreads) void k_fl_wake(FlConst c, FlBufs d, const FlState* __restrict__ st,
                                                      double* __restrict__ partial, int64_t spart, int step, int rows,
                                                      int sw, int tiles_span) {
    constexpr int SWG = kWaves * kLatChunkMax;
    __shared__ double sverts[(RH + 1) * (SWG + 1) * 3];
    __shared__ double sgam[RH * SWG];
    __shared__ double sred[kWaves][kRecv];
    const int b = blockIdx.z;
    if (st[b].done) return;
    const int ns = c.ns, N = c.N;
    const int64_t vs1 = ns + 1;
    const int tile = blockIdx.y;
    const int ts = tile % tiles_span, tr = tile / tiles_span;
    const int s0 = ts * sw, r0 = tr * RH;
    const int swe = min(sw, ns - s0);
    const int rhe = min(RH, rows - r0);
    const int vw = swe + 1;
    const double* WV = d.wverts + b * d.swv;
    const double* WG = d.wgam + b * d.swg;
    for (int e = threadIdx.x; e < (rhe + 1) * vw; e += kThreads) {
        const int q = e / vw, sp = e % vw;
        const int slot = (step + 1 - (r0 + q)) % c.slots;
        const double* row = WV + (int64_t)slot * vs1 * 3 + (s0 + sp);
        sverts[3 * e + 0] = row[0];
        sverts[3 * e + 1] = row[vs1];
        sverts[3 * e + 2] = row[2 * vs1];
    }
    for (int e = threadIdx.x; e < rhe * swe; e += kThreads) {
        const int q = e / swe, sp = e % swe;
        const int slot = (step - (r0 + q)) % c.slots;
        sgam[e] = WG[(int64_t)slot * ns + s0 + sp];
    }
    __syncthreads();
    const int lane = threadIdx.x & (kRecv - 1);
    const int wave = threadIdx.x / kRecv;
    const int p = blockIdx.x * kRecv + lane;
    double part = 0.0;
    const int chunk = (swe + kWaves - 1) / kWaves;
    const int sa = wave * chunk, sb = min(swe, sa + chunk);
    if (p < N && sa < sb) {
        const double* C = d.colloc + b * d.sp3;
        const double* Nn = d.normals + b * d.sp3;
        double ax = 0.0, ay = 0.0, az = 0.0;
        lattice_chunk<RH>(sverts, sgam, swe, rhe, vw, sa, sb, C[p], C[p + N], C[p + 2 * N], ax, ay, az);
        part = ax * Nn[p] + ay * Nn[p + N] + az * Nn[p + 2 * N];
    }
    sred[wave][lane] = part;
    __syncthreads();
    if (wave == 0 && p < N)
        partial[b * spart + (int64_t)tile * N + p] = (sred[0][lane] + sred[1][lane]) + (sred[2][lane] + sred[3][lane]);
}

// rhs = rhs_vel - sum of the tile partials (fixed order), airspeed blockIdx.y
__global__ __launch_bounds__(64 * kRedWaves) void k_fl_reduce(FlConst c, FlBufs d, const FlState* __restrict__ st,
                                                              const double* __restrict__ partial, int64_t spart,
                                                              int tiles) {
    __shared__ double red[kRedWaves][64];
    const int b = blockIdx.y;
    if (st[b].done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const double* P = partial + b * spart;
    double s = 0.0;
    if (p < c.N)
        for (int t = wave; t < tiles; t += kRedWaves) s += P[(int64_t)t * c.N + p];
    red[wave][lane] = s;
    __syncthreads();
#pragma unroll
    for (int h = kRedWaves / 2; h > 0; h >>= 1) {
        if (wave < h) red[wave][lane] += red[wave + h][lane];
        __syncthreads();
    }
    if (wave == 0 && p < c.N) d.rhs[b * d.sp + p] = d.rhs_vel[b * d.sp + p] - red[0][lane];
}

// gpart[b][cb][row] = A^{-1}(row, columns of block cb) rhs_b: 64 rows per work-group (one per lane),
// its 4 waves split the block's columns, kFlChunk airspeeds (blockIdx.z) per work-group so that each
// element of A^{-1} is read once per chunk.
__global__ __launch_bounds__(256) void k_fl_gemm(FlConst c, FlBufs d, const double* __restrict__ Ainv,
                                                 double* __restrict__ gpart, int cbs) {
    __shared__ double red[4][kFlChunk][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int N = c.N;
    const int row = blockIdx.x * 64 + lane;
    const int c0 = blockIdx.y * cbs, c1 = min(N, c0 + cbs);
    const int b0 = blockIdx.z * kFlChunk;
    const int nb = min(kFlChunk, c.B - b0);
    double acc[kFlChunk];
#pragma unroll
    for (int k = 0; k < kFlChunk; k++) acc[k] = 0.0;
    if (row < N) {
        for (int col = c0 + wave; col < c1; col += 4) {
            const double a = Ainv[row + (int64_t)col * N];
#pragma unroll
            for (int k = 0; k < kFlChunk; k++)
                if (k < nb) acc[k] += a * d.rhs[(b0 + k) * d.sp + col];
        }
    }
#pragma unroll
    for (int k = 0; k < kFlChunk; k++) red[wave][k][lane] = acc[k];
    __syncthreads();
    if (wave == 0 && row < N) {
        const int ncb = gridDim.y;
        for (int k = 0; k < nb; k++)
            gpart[((int64_t)(b0 + k) * ncb + blockIdx.y) * N + row] =
                (red[0][k][lane] + red[1][k][lane]) + (red[2][k][lane] + red[3][k][lane]);
    }
}

__device__ __forceinline__ double fl_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Circulation, forces, coefficients, loads and the convergence test of airspeed blockIdx.x
// (tests/uvlm_2dof.cpp:712-807; coefficients as coeff_cl/cm_multibody, vlm/src/vlm_backend.cpp:71-139)
__global__ __launch_bounds__(kFlPostThreads) void k_fl_post(FlConst c, FlBufs d, const FlAir* __restrict__ air,
                                                            FlState* __restrict__ st, const double* __restrict__ gpart,
                                                            int ncb, int step) {
    __shared__ double lds[kFlPostThreads / 64][6];
    const int b = blockIdx.x;
    FlState& s = st[b];
    if (s.done) return;  // uniform over the work-group
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int ns = c.ns, nc = c.nc, np = c.N;
    const int64_t vs1 = ns + 1, vs2 = vs1 * (nc + 1);
    const double* gp = gpart + (int64_t)b * ncb * np;
    auto gsum = [&](int P) {
        double g = 0.0;
        for (int k = 0; k < ncb; k++) g += gp[(int64_t)k * np + P];
        return g;
    };
    const double* V = d.verts + b * d.sv;
    const double* Nn = d.normals + b * d.sp3;
    const double* Vel = d.vel + b * d.sp3;
    const double* Ar = d.areas + b * d.sp;
    double* gam = d.gamma + b * d.sp;
    const double* gpr = d.gprev + b * d.sp;
    double* WG = d.wgam + b * d.swg + (int64_t)((step + 1) % c.slots) * ns;
    // freestream (1, 0, 0), lift axis (0, 0, 1): the section flies along -x at unit speed
    const double q = 0.5 * c.rho;
    double a_acc = 0.0, c_acc = 0.0, mx = 0.0, my = 0.0, mz = 0.0, mac_acc = 0.0;
    for (int p = t; p < np; p += kFlPostThreads) {
        const int i = p % ns, j = p / ns;
        const double g = gsum(p);
        const double gd = j > 0 ? g + (-1.0) * gsum(p - ns) : g;
        gam[p] = g;
        const double gt = (g + (-1.0) * gpr[p]) * c.inv_dt;
        if (j == nc - 1) WG[i] = g;
        const int64_t ia = i + (int64_t)j * vs1, ib = ia + 1;
        const v3<double> U = {Vel[p], Vel[p + np], Vel[p + 2 * np]};
        const v3<double> va = {V[ia], V[ia + vs2], V[ia + 2 * vs2]};
        const v3<double> vb = {V[ib], V[ib + vs2], V[ib + 2 * vs2]};
        const v3<double> cr = cross(U, vb - va);
        const double s1 = c.rho * gd;
        const double u = c.rho * gt * Ar[p];
        const v3<double> F = {s1 * cr.x + u * Nn[p], s1 * cr.y + u * Nn[p + np], s1 * cr.z + u * Nn[p + 2 * np]};
        c_acc += F.z;
        const v3<double> lever = {0.5 * (va.x + vb.x) - s.ref[0], 0.5 * (va.y + vb.y) - s.ref[1],
                                  0.5 * (va.z + vb.z) - s.ref[2]};
        const v3<double> mo = cross(lever, F);
        mx += mo.x;
        my += mo.y;
        mz += mo.z;
        a_acc += Ar[p];
    }
    for (int i = t; i < ns; i += kFlPostThreads) {
        const int64_t te = (int64_t)nc * vs1;
        const v3<double> d0 = {V[i] - V[i + te], V[i + vs2] - V[i + te + vs2], V[i + 2 * vs2] - V[i + te + 2 * vs2]};
        const v3<double> d1 = {V[i + 1] - V[i + 1 + te], V[i + 1 + vs2] - V[i + 1 + te + vs2],
                               V[i + 1 + 2 * vs2] - V[i + 1 + te + 2 * vs2]};
        const v3<double> d3 = {V[i + 1] - V[i], V[i + 1 + vs2] - V[i + vs2], V[i + 1 + 2 * vs2] - V[i + 2 * vs2]};
        const double c0 = sqrt(dot(d0, d0)), c1 = sqrt(dot(d1, d1)), w = sqrt(dot(d3, d3));
        mac_acc += 0.5 * (c0 * c0 + c1 * c1) * w;
    }
    double v[6] = {a_acc, c_acc, mx, my, mz, mac_acc};
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = fl_wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; k++) lds[wave][k] = v[k];
    __syncthreads();
    if (t != 0) return;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double sum = lds[0][k];
        for (int w = 1; w < kFlPostThreads / 64; w++) sum += lds[w][k];
        v[k] = sum;
    }
    const double area = v[0], mac = v[5] / area;
    const double cl = v[1] / (q * area);
    const double cmy = v[3] / (q * area * mac);
    const FlAir& A = air[b];
    s.F1[0] = -cl / (kPiF * c.mu);
    s.F1[1] = 2.0 * cmy / (kPiF * c.mu * (c.r_a * c.r_a));
    s.dF[0] = s.F1[0] - s.F[0] - A.kh * s.du[0];
    s.dF[1] = s.F1[1] - s.F[1] - A.ka * (torsional(c, s.u1[1]) - torsional(c, s.u[1]));
    s.du_k[0] = s.du_k[0] - s.du[0];
    s.du_k[1] = s.du_k[1] - s.du[1];
    s.it += 1;
    const double nrm = sqrt(s.du_k[0] * s.du_k[0] + s.du_k[1] * s.du_k[1]);
    s.done = (!(nrm > kFlEps) || s.it > kFlMaxIter) ? 1 : 0;
}

// The step is over for every airspeed: commit the trial state, record the history, gamma_prev <-
// gamma, reset the iteration state.
__global__ void k_fl_commit(FlConst c, FlBufs d, FlState* __restrict__ st, double* __restrict__ hist,
                            int32_t* __restrict__ iters, int step) {
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < c.N) d.gprev[b * d.sp + e] = d.gamma[b * d.sp + e];
    if (e != 0) return;
    FlState& s = st[b];
    for (int k = 0; k < 2; k++) {
        s.u[k] = s.u1[k];
        s.v[k] = s.v1[k];
        s.a[k] = s.a1[k];
        s.F[k] = s.F1[k];
        s.du[k] = 0.0;
        s.du_k[k] = 1.0;
        s.dF[k] = 0.0;
    }
    double* h = hist + ((int64_t)(step + 1) * c.B + b) * 8;
    h[0] = s.u[0];
    h[1] = s.u[1];
    h[2] = s.v[0];
    h[3] = s.v[1];
    h[4] = s.a[0];
    h[5] = s.a[1];
    h[6] = s.F[0];
    h[7] = s.F[1];
    iters[(int64_t)step * c.B + b] = s.it;
    s.it = 0;
    s.done = 0;
}

// initial wake rows and circulations of every airspeed: the t = 0 trailing edge in ring slot 0, the
// t = 0 trailing-edge circulation in slot 0, gamma_prev = gamma(0) (tests/uvlm_2dof.cpp:547-559)
__global__ void k_fl_init(FlConst c, FlBufs d, const double* __restrict__ verts0, const double* __restrict__ gamma0) {
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = c.ns, nc = c.nc;
    const int64_t vs1 = ns + 1, nv = vs1 * (nc + 1);
    if (e < c.N) {
        d.gprev[b * d.sp + e] = gamma0[e];
        d.gamma[b * d.sp + e] = gamma0[e];
        if (e / ns == nc - 1) d.wgam[b * d.swg + e % ns] = gamma0[e];
    }
    if (e < vs1) {
        double* W = d.wverts + b * d.swv;
        for (int r = 0; r < 3; r++) W[e + r * vs1] = verts0[e + nc * vs1 + r * nv];
    }
}

__global__ void k_fl_eye(double* a, int n) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < (int64_t)n * n) a[q] = (q % n) == (q / n) ? 1.0 : 0.0;
}

__global__ void k_fl_fill(double* x, int64_t n, double v) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) x[q] = v;
}

inline unsigned fl_blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace tvlm

using namespace tvlm;

struct tvlm_flutter_plan {
    FlConst c{};
    FlBufs d{};
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    FlAir* air = nullptr;
    FlState* st = nullptr;
    double *vinit = nullptr, *colloc_h = nullptr, *verts0 = nullptr, *lhs = nullptr, *ainv = nullptr;
    double *partial = nullptr, *gpart = nullptr, *hist = nullptr, *gamma0 = nullptr, *tr0 = nullptr;
    int32_t *ipiv = nullptr, *info = nullptr, *iters = nullptr;
    int32_t* done_host = nullptr;  // pinned: the per-airspeed flags after an iteration
    int64_t spart = 0;             // partial elements per airspeed
    int R = 0, ncb = 1, cbs = 0;
    int64_t max_steps = 0, step = 0;
    double init_angle = 0.0;
    int32_t lu_info = 0;
    std::vector<hipEvent_t> events;

    int alloc(void** p, size_t bytes) {
        TVLM_HIP(hipMalloc(p, bytes));
        allocs.push_back(*p);
        return TVLM_OK;
    }
};

namespace tvlm {
namespace {

int fl_wake_launch(tvlm_flutter_plan* P, int step, int& tiles) {
    const int rows = (int)(step + 1 < P->R ? step + 1 : P->R);
    const int recv = (P->c.N + kRecv - 1) / kRecv;
    int sw = 0, rh = 1;
    fl_tile_shape(P->c.ns, rows, recv, sw, rh, tiles);
    TVLM_REQUIRE((int64_t)tiles * P->c.N <= P->spart, "wake partial buffer too small (%d tiles)", tiles);
    const int tiles_span = (P->c.ns + sw - 1) / sw;
    const dim3 grid(recv, tiles, P->c.B);
    hipStream_t st = P->stream;
    if (rh == 4)
        hipLaunchKernelGGL((k_fl_wake<4>), grid, dim3(kThreads), 0, st, P->c, P->d, P->st, P->partial, P->spart, step,
                           rows, sw, tiles_span);
    else if (rh == 2)
        hipLaunchKernelGGL((k_fl_wake<2>), grid, dim3(kThreads), 0, st, P->c, P->d, P->st, P->partial, P->spart, step,
                           rows, sw, tiles_span);
    else
        hipLaunchKernelGGL((k_fl_wake<1>), grid, dim3(kThreads), 0, st, P->c, P->d, P->st, P->partial, P->spart, step,
                           rows, sw, tiles_span);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// one fixed-point iteration of every airspeed still iterating at step `step`
int fl_iteration(tvlm_flutter_plan* P, int step) {
    hipStream_t st = P->stream;
    const FlConst& c = P->c;
    hipLaunchKernelGGL(k_fl_struct, dim3(fl_blocks(c.B, 64)), dim3(64), 0, st, c, P->air, P->st, step);
    TVLM_LAUNCHED();
    const int64_t nv = (int64_t)(c.ns + 1) * (c.nc + 1);
    hipLaunchKernelGGL(k_fl_geom, dim3(fl_blocks(nv + c.N), c.B), dim3(256), 0, st, c, P->d, P->st, step);
    TVLM_LAUNCHED();
    int tiles = 0;
    TVLM_TRY(fl_wake_launch(P, step, tiles));
    hipLaunchKernelGGL(k_fl_reduce, dim3((c.N + 63) / 64, c.B), dim3(64 * kRedWaves), 0, st, c, P->d, P->st,
                       P->partial, P->spart, tiles);
    TVLM_LAUNCHED();
    hipLaunchKernelGGL(k_fl_gemm, dim3((c.N + 63) / 64, P->ncb, (c.B + kFlChunk - 1) / kFlChunk), dim3(256), 0, st, c,
                       P->d, P->ainv, P->gpart, P->cbs);
    TVLM_LAUNCHED();
    hipLaunchKernelGGL(k_fl_post, dim3(c.B), dim3(kFlPostThreads), 0, st, c, P->d, P->air, P->st, P->gpart, P->ncb,
                       step);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

extern "C" {

int tvlm_flutter_create(tvlm_flutter_plan** out, int64_t ns, int64_t nc, const double* verts_init,
                        const tvlm_flutter_params* prm, int32_t n_airspeeds, const double* U_a, int32_t wake_rows,
                        int64_t max_steps, void* stream) {
    TVLM_REQUIRE(out && verts_init && prm && U_a, "null argument");
    TVLM_REQUIRE(ns >= 1 && nc >= 1 && ns * nc <= (1 << 20), "mesh %lld x %lld out of range", (long long)ns,
                 (long long)nc);
    TVLM_REQUIRE(n_airspeeds >= 1 && n_airspeeds <= 1024, "n_airspeeds %d out of range", n_airspeeds);
    TVLM_REQUIRE(wake_rows >= 1 && wake_rows <= 100000, "wake_rows %d out of range", wake_rows);
    TVLM_REQUIRE(max_steps >= 1 && max_steps <= 10000000, "max_steps out of range");
    TVLM_REQUIRE(prm->b > 0.0 && prm->mu > 0.0 && prm->r_a > 0.0, "b, mu and r_a must be positive");
    for (int b = 0; b < n_airspeeds; b++) TVLM_REQUIRE(U_a[b] > 0.0, "U_a[%d] must be positive", b);
    auto* P = new tvlm_flutter_plan();
    auto fail = [&](int rc) {
        for (void* p : P->allocs) (void)hipFree(p);
        if (P->done_host) (void)hipHostFree(P->done_host);
        delete P;
        return rc;
    };
    const int N = (int)(ns * nc);
    const int B = n_airspeeds;
    const int64_t vs1 = ns + 1, nv = vs1 * (nc + 1);
    P->stream = as_stream(stream);
    P->R = wake_rows;
    P->max_steps = max_steps;
    hipStream_t st = P->stream;
    // dimensionless wing (nondimensionalize_verts, tests/uvlm_2dof.cpp:417-425) and the time step:
    // the last chordwise panel of span station 0 (tests/uvlm_2dof.cpp:474-481)
    std::vector<double> vin(verts_init, verts_init + nv * 4);
    for (int64_t k = 0; k < nv * 3; k++) vin[k] = vin[k] / prm->b;
    const int64_t l1 = 0 + nc * vs1, l0 = 0 + (nc - 1) * vs1;
    const double ex = vin[l1] - vin[l0], ey = vin[l1 + nv] - vin[l0 + nv], ez = vin[l1 + 2 * nv] - vin[l0 + 2 * nv];
    const double dt = std::sqrt(ex * ex + ey * ey + ez * ez);
    if (!(dt > 0.0)) {
        delete P;
        set_error("zero last-panel chord: no time step");
        return TVLM_ERR_INVALID;
    }
    FlConst& c = P->c;
    c.dt = dt;
    c.inv_dt = 1.0 / dt;
    c.a_h = prm->a_h;
    c.rho = prm->rho;
    c.mu = prm->mu;
    c.r_a = prm->r_a;
    const double nb = 0.25, ng = 0.5;  // average acceleration
    c.x0 = 1.0 / (nb * dt * dt);
    c.x1 = ng / (nb * dt);
    c.xd0 = 1.0 / (nb * dt);
    c.xd1 = ng / nb;
    c.xdd0 = 1.0 / (2.0 * nb);
    c.xdd1 = -dt * (1.0 - ng / (2.0 * nb));
    // M = [[1, x_a], [x_a / r_a^2, 1]] (tests/uvlm_2dof.cpp:334-345), column-major
    c.M[0] = 1.0;
    c.M[1] = prm->x_a / (prm->r_a * prm->r_a);
    c.M[2] = prm->x_a;
    c.M[3] = 1.0;
    for (int k = 0; k < 4; k++) c.beta[k] = prm->beta[k];
    c.N = N;
    c.ns = (int32_t)ns;
    c.nc = (int32_t)nc;
    c.slots = wake_rows + 1;
    c.B = B;
    P->init_angle = prm->init_pitch;
    // per-airspeed structural constants and initial state
    std::vector<FlAir> air(B);
    std::vector<FlState> s0(B);
    std::vector<double> hist0((size_t)B * 8);
    for (int b = 0; b < B; b++) {
        FlAir& A = air[b];
        const double Ua = U_a[b];
        A.C[0] = 2.0 * prm->zeta_h * (prm->omega / Ua);
        A.C[1] = 0.0;
        A.C[2] = 0.0;
        A.C[3] = 2.0 * prm->zeta_a / Ua;
        A.kh = (prm->omega / Ua) * (prm->omega / Ua);
        A.ka = 1.0 / (Ua * Ua);
        double K[4];
        for (int k = 0; k < 4; k++) K[k] = c.x0 * c.M[k] + c.x1 * A.C[k];
        const double det = K[0] * K[3] - K[2] * K[1];
        A.Kinv[0] = K[3] / det;
        A.Kinv[1] = -K[1] / det;
        A.Kinv[2] = -K[2] / det;
        A.Kinv[3] = K[0] / det;
        FlState& s = s0[b];
        std::memset(&s, 0, sizeof(s));
        s.u[1] = prm->init_pitch;
        // M a0 = [-kh h0, -ka f(alpha0)] - C v0 (tests/uvlm_2dof.cpp:395-412)
        const double f0 = -A.kh * s.u[0], f1 = -A.ka * (c.beta[0] + c.beta[1] * s.u[1] + c.beta[2] * s.u[1] * s.u[1] +
                                                         c.beta[3] * s.u[1] * s.u[1] * s.u[1]);
        const double mdet = c.M[0] * c.M[3] - c.M[2] * c.M[1];
        s.a[0] = (c.M[3] * f0 - c.M[2] * f1) / mdet;
        s.a[1] = (-c.M[1] * f0 + c.M[0] * f1) / mdet;
        s.du_k[0] = s.du_k[1] = 1.0;
        double* h = hist0.data() + (size_t)b * 8;
        h[0] = s.u[0];
        h[1] = s.u[1];
        h[2] = s.v[0];
        h[3] = s.v[1];
        h[4] = s.a[0];
        h[5] = s.a[1];
    }
    // buffers
    const int recv = (N + kRecv - 1) / kRecv;
    int max_tiles = 0;
    for (int rows = 1; rows <= wake_rows; rows++) {
        int sw, rh, tiles;
        fl_tile_shape((int)ns, rows, recv, sw, rh, tiles);
        max_tiles = tiles > max_tiles ? tiles : max_tiles;
    }
    P->spart = (int64_t)max_tiles * N;
    P->cbs = kFlGemmCols;
    P->ncb = (N + P->cbs - 1) / P->cbs;
    FlBufs& d = P->d;
    d.sv = nv * 3;
    d.sp3 = (int64_t)N * 3;
    d.sp = N;
    d.swv = (int64_t)c.slots * vs1 * 3;
    d.swg = (int64_t)c.slots * ns;
    const size_t D8 = sizeof(double);
    int rc = TVLM_OK;
    auto A = [&](void** p, size_t bytes) {
        if (rc == TVLM_OK) rc = P->alloc(p, bytes);
    };
    A((void**)&P->air, sizeof(FlAir) * B);
    A((void**)&P->st, sizeof(FlState) * B);
    A((void**)&P->vinit, D8 * nv * 4);
    A((void**)&P->colloc_h, D8 * N * 3);
    A((void**)&P->verts0, D8 * nv * 4);
    A((void**)&P->gamma0, D8 * N);
    A((void**)&P->lhs, D8 * (size_t)N * N);
    A((void**)&P->ainv, D8 * (size_t)N * N);
    A((void**)&d.verts, D8 * d.sv * B);
    A((void**)&d.colloc, D8 * d.sp3 * B);
    A((void**)&d.normals, D8 * d.sp3 * B);
    A((void**)&d.areas, D8 * d.sp * B);
    A((void**)&d.vel, D8 * d.sp3 * B);
    A((void**)&d.rhs_vel, D8 * d.sp * B);
    A((void**)&d.rhs, D8 * d.sp * B);
    A((void**)&d.gamma, D8 * d.sp * B);
    A((void**)&d.gprev, D8 * d.sp * B);
    A((void**)&d.wverts, D8 * d.swv * B);
    A((void**)&d.wgam, D8 * d.swg * B);
    A((void**)&P->partial, D8 * (size_t)P->spart * B);
    A((void**)&P->gpart, D8 * (size_t)P->ncb * N * B);
    A((void**)&P->hist, D8 * (size_t)(max_steps + 1) * B * 8);
    A((void**)&P->iters, sizeof(int32_t) * (size_t)max_steps * B);
    A((void**)&P->ipiv, sizeof(int32_t) * N);
    A((void**)&P->info, sizeof(int32_t));
    A((void**)&P->tr0, sizeof(double) * 16);
    if (rc == TVLM_OK && hipHostMalloc((void**)&P->done_host, sizeof(int32_t) * B, hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        rc = TVLM_ERR_NOMEM;
    }
    if (rc != TVLM_OK) return fail(rc);
    d.vinit = P->vinit;
    d.colloc_h = P->colloc_h;
    auto H2D = [&](void* dst, const void* src, size_t bytes) {
        if (rc == TVLM_OK && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("hipMemcpyAsync failed");
            rc = TVLM_ERR_HIP;
        }
    };
    H2D(P->air, air.data(), sizeof(FlAir) * B);
    H2D(P->st, s0.data(), sizeof(FlState) * B);
    H2D(P->vinit, vin.data(), D8 * nv * 4);
    H2D(P->hist, hist0.data(), D8 * B * 8);
    if (rc != TVLM_OK) return fail(rc);
    // wake rings start zeroed (only slots already shed are ever read; zeros keep the rest defined)
    hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(d.swv * B)), dim3(256), 0, st, d.wverts, d.swv * B, 0.0);
    hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(d.swg * B)), dim3(256), 0, st, d.wgam, d.swg * B, 0.0);
    if (hipGetLastError() != hipSuccess) return fail(TVLM_ERR_HIP);
    // initial pose: rotation by the initial pitch about the elastic axis (tests/uvlm_2dof.cpp:505-520)
    {
        const double px = 1.0 + prm->a_h, cs = std::cos(prm->init_pitch), sn = std::sin(prm->init_pitch);
        double T0[16] = {cs, 0, -sn, 0, 0, 1, 0, 0, sn, 0, cs, 0, px - cs * px, 0, sn * px, 1};
        double* T0d = P->tr0;
        H2D(T0d, T0, sizeof(T0));
        tvlm_view tr = {T0d, {4, 4, 1}, {1, 4, 16}};
        tvlm_view vi = {P->vinit, {ns + 1, nc + 1, 4}, {1, ns + 1, nv}};
        tvlm_view v0 = {P->verts0, {ns + 1, nc + 1, 4}, {1, ns + 1, nv}};
        if (rc == TVLM_OK) rc = tvlm_displace_wing_f64(1, &tr, &v0, &vi, st);
        // metrics of the initial pose into the first airspeed's buffers, then shared copies
        tvlm_view C0 = {d.colloc, {ns, nc, 3}, {1, ns, (int64_t)N}};
        tvlm_view N0 = {d.normals, {ns, nc, 3}, {1, ns, (int64_t)N}};
        tvlm_view A0 = {d.areas, {ns, nc, 1}, {1, ns, (int64_t)N}};
        if (rc == TVLM_OK) rc = tvlm_mesh_metrics_f64(0.0, 1, &v0, &C0, &N0, &A0, st);
        if (rc == TVLM_OK && hipMemcpyAsync(P->colloc_h, d.colloc, D8 * N * 3, hipMemcpyDeviceToDevice, st) != hipSuccess)
            rc = TVLM_ERR_HIP;
        // rigid influence matrix at the initial pose (iteration 0), its LU and inverse
        const int64_t NN = (int64_t)N * N;
        if (rc == TVLM_OK) {
            hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(NN)), dim3(256), 0, st, P->lhs, NN, 0.0);
            if (hipGetLastError() != hipSuccess) rc = TVLM_ERR_HIP;
        }
        tvlm_view L = {P->lhs, {N, N, 1}, {1, N, NN}};
        tvlm_view W = {d.wverts, {ns + 1, 2, 4}, {1, ns + 1, (ns + 1) * 2}};  // unused with iteration 0
        if (rc == TVLM_OK) rc = tvlm_lhs_assemble_f64(&L, 1, &C0, &N0, &v0, &W, 0, 0.0, st);
        if (rc == TVLM_OK) rc = tvlm_getrf_f64(&L, P->ipiv, P->info, st);
        if (rc == TVLM_OK) {
            hipLaunchKernelGGL(k_fl_eye, dim3(fl_blocks(NN)), dim3(256), 0, st, P->ainv, N);
            if (hipGetLastError() != hipSuccess) rc = TVLM_ERR_HIP;
        }
        tvlm_view X = {P->ainv, {N, N, 1}, {1, N, NN}};
        if (rc == TVLM_OK) rc = tvlm_getrs_f64(&L, P->ipiv, &X, st);
        // t = 0: freestream only, V = (1, 0, 0); gamma(0) by the LU (tests/uvlm_2dof.cpp:547-559)
        if (rc == TVLM_OK) {
            hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(N)), dim3(256), 0, st, d.vel, (int64_t)N, 1.0);
            hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(2 * (int64_t)N)), dim3(256), 0, st, d.vel + N, 2 * (int64_t)N,
                               0.0);
            hipLaunchKernelGGL(k_fl_fill, dim3(fl_blocks(N)), dim3(256), 0, st, P->gamma0, (int64_t)N, 0.0);
            if (hipGetLastError() != hipSuccess) rc = TVLM_ERR_HIP;
        }
        tvlm_view G0 = {P->gamma0, {N, 1, 1}, {1, N, N}};
        tvlm_view Vl = {d.vel, {ns, nc, 3}, {1, ns, (int64_t)N}};
        if (rc == TVLM_OK) rc = tvlm_rhs_assemble_velocities_f64(&G0, 1, &N0, &Vl, st);
        if (rc == TVLM_OK) rc = tvlm_getrs_f64(&L, P->ipiv, &G0, st);
        if (rc == TVLM_OK) {
            hipLaunchKernelGGL(k_fl_init, dim3(fl_blocks(N > vs1 ? N : vs1), B), dim3(256), 0, st, P->c, P->d, P->verts0,
                               P->gamma0);
            if (hipGetLastError() != hipSuccess) rc = TVLM_ERR_HIP;
        }
        if (rc == TVLM_OK && hipMemcpyAsync(&P->lu_info, P->info, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = TVLM_ERR_HIP;
        if (rc == TVLM_OK && hipStreamSynchronize(st) != hipSuccess) {
            set_error("flutter plan setup failed on the device");
            rc = TVLM_ERR_HIP;
        }
        if (rc == TVLM_OK && P->lu_info != 0) {
            set_error("singular influence matrix (getrf info %d)", P->lu_info);
            rc = TVLM_ERR_SINGULAR;
        }
    }
    if (rc != TVLM_OK) return fail(rc);
    *out = P;
    return TVLM_OK;
}

int tvlm_flutter_run(tvlm_flutter_plan* P, int64_t steps) {
    TVLM_REQUIRE(P, "null plan");
    TVLM_REQUIRE(steps >= 0 && P->step + steps <= P->max_steps, "%lld steps past the plan's %lld",
                 (long long)(P->step + steps), (long long)P->max_steps);
    hipStream_t st = P->stream;
    const int B = P->c.B;
    for (int64_t k = 0; k < steps; k++) {
        const int step = (int)P->step;
        for (int it = 0; it <= kFlMaxIter; it++) {
            TVLM_TRY(fl_iteration(P, step));
            // the flags are strided inside the state records: gather them with one 2-D copy
            TVLM_HIP(hipMemcpy2DAsync(P->done_host, sizeof(int32_t), &P->st[0].done, sizeof(FlState), sizeof(int32_t), B,
                                      hipMemcpyDeviceToHost, st));
            TVLM_HIP(hipStreamSynchronize(st));
            bool all = true;
            for (int b = 0; b < B && all; b++) all = P->done_host[b] != 0;
            if (all) break;
        }
        hipLaunchKernelGGL(k_fl_commit, dim3(fl_blocks(P->c.N), B), dim3(256), 0, st, P->c, P->d, P->st, P->hist,
                           P->iters, step);
        TVLM_LAUNCHED();
        P->step++;
    }
    TVLM_HIP(hipStreamSynchronize(st));
    return TVLM_OK;
}

int tvlm_flutter_results(tvlm_flutter_plan* P, int64_t* steps_done, double* dt, double* states, int32_t* iters) {
    TVLM_REQUIRE(P, "null plan");
    hipStream_t st = P->stream;
    const int B = P->c.B;
    if (steps_done) *steps_done = P->step;
    if (dt) *dt = P->c.dt;
    if (states)
        TVLM_HIP(hipMemcpyAsync(states, P->hist, sizeof(double) * (size_t)(P->step + 1) * B * 8, hipMemcpyDeviceToHost,
                                st));
    if (iters && P->step > 0)
        TVLM_HIP(hipMemcpyAsync(iters, P->iters, sizeof(int32_t) * (size_t)P->step * B, hipMemcpyDeviceToHost, st));
    TVLM_HIP(hipStreamSynchronize(st));
    return TVLM_OK;
}

int tvlm_flutter_time_wake(tvlm_flutter_plan* P, int32_t launches, double* total_ms, double* pairs) {
    TVLM_REQUIRE(P && total_ms && pairs && launches >= 1, "bad argument");
    TVLM_REQUIRE(P->step >= 1, "run at least one step first");
    hipStream_t st = P->stream;
    if (P->events.empty()) {
        P->events.resize(2, nullptr);
        for (auto& e : P->events) TVLM_HIP(hipEventCreate(&e));
    }
    // the wake launch of the next step's first iteration, on the current state (every airspeed active)
    const int step = (int)P->step;
    int tiles = 0;
    TVLM_TRY(fl_wake_launch(P, step, tiles));  // warm-up
    TVLM_HIP(hipEventRecord(P->events[0], st));
    for (int k = 0; k < launches; k++) TVLM_TRY(fl_wake_launch(P, step, tiles));
    TVLM_HIP(hipEventRecord(P->events[1], st));
    TVLM_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    TVLM_HIP(hipEventElapsedTime(&ms, P->events[0], P->events[1]));
    const int rows = step + 1 < P->R ? step + 1 : P->R;
    *total_ms = ms;
    *pairs = (double)P->c.B * P->c.N * (double)rows * P->c.ns;
    return TVLM_OK;
}

int tvlm_flutter_destroy(tvlm_flutter_plan* P) {
    if (!P) return TVLM_OK;
    (void)hipStreamSynchronize(P->stream);
    for (void* p : P->allocs) (void)hipFree(p);
    for (auto e : P->events) (void)hipEventDestroy(e);
    if (P->done_host) (void)hipHostFree(P->done_host);
    delete P;
    return TVLM_OK;
}

}  // extern "C"

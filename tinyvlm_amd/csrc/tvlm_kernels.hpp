// tvlm_kernels.hpp -- Biot-Savart assembly kernels shared by the Backend entry points and the
// device-resident solver loops.
//
// lhs block: a work-group owns 64 receivers (one per lane, colloc + normal in registers) and TC
// rings whose 12 vertex coordinates are staged in LDS; its 4 waves take every 4th ring, so each
// wave reads LDS as a broadcast and writes one coalesced 64-element column segment per ring.
//
// wake partials: the wake rings are cut into tiles of RH rows x SW spans; a work-group stages one
// tile's vertex rows and circulations in LDS, each wave takes every 4th ring of the tile for the
// work-group's 64 receivers, and the 4 waves reduce in LDS to one partial per (tile, receiver).
// k_wake_reduce then sums the partials of all tiles in a fixed order (bitwise reproducible).
#pragma once

#include "tvlm_internal.hpp"

namespace tvlm {
namespace {

constexpr int kRecv = 64;      // receivers per work-group (one per lane)
constexpr int kWaves = 4;      // waves per work-group
constexpr int kThreads = kRecv * kWaves;

template <typename T>
struct RingSource {
    const T* verts;   // {ns+1, nc+1, 4} plane contiguous
    int64_t vs2;      // plane stride of verts
    int32_t ns, nc;   // panels spanwise / chordwise of the ring surface
    const T* wake;    // {ns+1, nwv, 4} plane contiguous (TE columns only)
    int64_t ws2;
    int32_t nwv;
    int32_t iteration;
};

template <typename T>
struct Receivers {
    const T* colloc;  // {ns, nc, 3}: element p + k*cs2
    int64_t cs2;
    const T* normals;
    int64_t ns2;
    int32_t m;        // number of receivers (panels)
};

// Velocity of the ring (v0, v1, v2, v3) with image at c in the form selected by MODE:
// 0 exact (sigma = 0), 1 exact with core, 2 FP32 rsqrt form, 3 FP64 seeded form.
template <typename T, int MODE>
__device__ __forceinline__ v3<T> ring_vel(const v3<T>& c, const T* v0, const T* v1, const T* v2, const T* v3_,
                                          T sigma) {
    if constexpr (MODE == 2) {
        float ax = 0.f, ay = 0.f, az = 0.f;
        fast_ring(ax, ay, az, c.x, c.y, c.z, v0, v1, v2, v3_);
        return {ax, ay, az};
    } else if constexpr (MODE == 3) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        seed_ring(ax, ay, az, c.x, c.y, c.z, v0, v1, v2, v3_);
        return {ax, ay, az};
    } else {
        return ring_exact<T, MODE == 1>(c, {v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]},
                                        {v3_[0], v3_[1], v3_[2]}, sigma);
    }
}

// Loads ring r's vertices (v0, v1 = +1 span, v2 = +1 span +1 chord, v3 = +1 chord) from a vertex
// grid whose row (chordwise) stride is ns+1.
template <typename T>
__device__ __forceinline__ void load_ring(T* dst, const T* grid, int64_t s2, int64_t base, int32_t ns) {
    const int64_t off[4] = {base, base + 1, base + ns + 2, base + ns + 1};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        dst[3 * q + 0] = grid[off[q]];
        dst[3 * q + 1] = grid[off[q] + s2];
        dst[3 * q + 2] = grid[off[q] + 2 * s2];
    }
}

// One tile of the lhs block of a (receiver surface, ring surface) pair: receivers [bx kRecv, +kRecv)
// x rings [by TC, +TC).  lhs points at (offset_i, offset_j).  Every thread of the work-group calls it
// (it holds a barrier); sv is the caller's LDS staging of TC rings.
template <typename T, int TC, int MODE>  // MODE as ring_vel
__device__ __forceinline__ void lhs_tile(T (*sv)[12], T* __restrict__ lhs, int64_t ld, const Receivers<T>& rc,
                                         const RingSource<T>& src, T sigma, int bx, int by) {
    const int n_rings = src.ns * src.nc;
    const int col0 = by * TC;
    for (int e = threadIdx.x; e < TC * 4; e += kThreads) {
        const int r = e >> 2, q = e & 3;
        const int lidx = col0 + r;
        if (lidx < n_rings) {
            const int64_t base = lidx + lidx / src.ns;  // vertex 0 of ring lidx (vlm_backend_cpu.cpp:111)
            const int64_t off = q == 0 ? base : q == 1 ? base + 1 : q == 2 ? base + src.ns + 2 : base + src.ns + 1;
            sv[r][3 * q + 0] = src.verts[off];
            sv[r][3 * q + 1] = src.verts[off + src.vs2];
            sv[r][3 * q + 2] = src.verts[off + 2 * src.vs2];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & (kRecv - 1);
    const int wave = threadIdx.x / kRecv;
    const int i = bx * kRecv + lane;
    if (i >= rc.m) return;
    const v3<T> c = {rc.colloc[i], rc.colloc[i + rc.cs2], rc.colloc[i + 2 * rc.cs2]};
    const v3<T> n = {rc.normals[i], rc.normals[i + rc.ns2], rc.normals[i + 2 * rc.ns2]};
    const int ncols = min(TC, n_rings - col0);
    const int te0 = (src.nc - 1) * src.ns;
    for (int r = wave; r < ncols; r += kWaves) {
        const int lidx = col0 + r;
        T* dst = lhs + i + (int64_t)lidx * ld;
        T acc = *dst;
        acc += dot(ring_vel<T, MODE>(c, sv[r], sv[r] + 3, sv[r] + 6, sv[r] + 9, sigma), n);
        if (src.iteration > 0 && lidx >= te0) {
            // trailing-edge column: wake rings of rows nwv-2-k, k < iteration (vlm_backend_cpu.cpp:121-139)
            const int j = lidx - te0;
            for (int k = 0; k < src.iteration; k++) {
                T w[12];
                load_ring(w, src.wake, src.ws2, (int64_t)j + (int64_t)(src.nwv - 2 - k) * (src.ns + 1), src.ns);
                acc += dot(ring_vel<T, MODE>(c, w, w + 3, w + 6, w + 9, sigma), n);
            }
        }
        *dst = acc;
    }
}

// lhs block of one (receiver surface, ring surface) pair, one tile per work-group.
template <typename T, int TC, int MODE>  // MODE as ring_vel
__global__ __launch_bounds__(kThreads) void k_lhs_block(T* __restrict__ lhs, int64_t ld, Receivers<T> rc,
                                                        RingSource<T> src, T sigma) {
    __shared__ T sv[TC][12];
    lhs_tile<T, TC, MODE>(sv, lhs, ld, rc, src, sigma, blockIdx.x, blockIdx.y);
}

// One (receiver surface, ring surface) block of one influence matrix in a grouped launch: the tiles
// of all sections of all matrices are numbered in section order, section s owning tiles
// [tile0, tile0 + tiles_x tiles_y); tile t of it is receiver block t % tiles_x, ring block t / tiles_x.
template <typename T>
struct LhsSection {
    T* lhs;  // the section's origin in its matrix
    int64_t ld;
    Receivers<T> rc;
    RingSource<T> src;
    int32_t tiles_x, tile0;
};

// A value every lane of the wave holds alike, moved to a scalar register (the section fields the
// grouped kernel reads from LDS would otherwise sit in vector registers and cost it occupancy).
__device__ __forceinline__ int32_t wave_uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename P>
__device__ __forceinline__ P* wave_uniform(P* p) {
    return reinterpret_cast<P*>(wave_uniform((int64_t)reinterpret_cast<uintptr_t>(p)));
}
template <typename T>
__device__ __forceinline__ LhsSection<T> wave_uniform(const LhsSection<T>& a) {
    LhsSection<T> b;
    b.lhs = wave_uniform(a.lhs);
    b.ld = wave_uniform(a.ld);
    b.rc.colloc = wave_uniform(a.rc.colloc);
    b.rc.cs2 = wave_uniform(a.rc.cs2);
    b.rc.normals = wave_uniform(a.rc.normals);
    b.rc.ns2 = wave_uniform(a.rc.ns2);
    b.rc.m = wave_uniform(a.rc.m);
    b.src.verts = wave_uniform(a.src.verts);
    b.src.vs2 = wave_uniform(a.src.vs2);
    b.src.ns = wave_uniform(a.src.ns);
    b.src.nc = wave_uniform(a.src.nc);
    b.src.wake = wave_uniform(a.src.wake);
    b.src.ws2 = wave_uniform(a.src.ws2);
    b.src.nwv = wave_uniform(a.src.nwv);
    b.src.iteration = wave_uniform(a.src.iteration);
    b.tiles_x = wave_uniform(a.tiles_x);
    b.tile0 = wave_uniform(a.tile0);
    return b;
}

// Grouped assembly: many influence matrices of any sizes (different meshes, requests) in one
// launch, one tile per work-group over the work list of all their tiles (the first `tiles`: a
// prefix of whole matrices), numbered section by section; the hardware hands the work-groups out in
// order, so the ones running together work on neighbouring tiles, as in one per-section launch.
// tile0[s] is the first tile of section s (tile0[nsec] = tiles); chunk[c] is the section holding tile
// c * kGroupChunk, so a tile's section is chunk[t / kGroupChunk] or one of the few after it.  Every
// tile runs lhs_tile, so each matrix is bitwise the per-section launches' result.
constexpr int kGroupChunk = 64;

template <typename T, int TC, int MODE>
__global__ __launch_bounds__(kThreads) void k_lhs_grouped(const LhsSection<T>* __restrict__ secs,
                                                          const int32_t* __restrict__ tile0,
                                                          const int32_t* __restrict__ chunk, int tiles, T sigma) {
    __shared__ T sv[TC][12];
    __shared__ LhsSection<T> s_sec;
    const int t = blockIdx.x;
    if (t >= tiles) return;
    if (threadIdx.x == 0) {
        int s = chunk[t / kGroupChunk];
        while (tile0[s + 1] <= t) s++;  // sections are non-empty and tile0 ends with `tiles`
        s_sec = secs[s];
    }
    __syncthreads();
    const LhsSection<T> S = wave_uniform(s_sec);
    const int local = t - S.tile0;
    lhs_tile<T, TC, MODE>(sv, S.lhs, S.ld, S.rc, S.src, sigma, local % S.tiles_x, local / S.tiles_x);
}

// ---- wake influence ---------------------------------------------------------------------------
constexpr int kMaxWakes = TVLM_MAX_SURFACES;

template <typename T>
struct WakeSurface {
    const T* verts;   // {ns+1, nwv, 4} plane contiguous
    int64_t ws2;
    const T* gamma;   // {ns, nwv-1} contiguous
    int32_t ns, nwv;
    int32_t row_begin, row_end;  // ring rows [row_begin, row_end)
    int32_t sw, rh;              // tile extent: spans, rows
    int32_t tiles_span, tile_begin;
    // k_wake_lattice only: a coarse lattice over a finer wake of nsf spans (0: this is the wake
    // itself).  Coarse vertex (column c, row r) is the fine vertex (min(c vcs, nsf), vrow0 + r vrs);
    // coarse ring (r, c) carries the mean circulation of the fine rings of rows
    // [vrow0 + r vrs, vrow0 + (r+1) vrs) and spans [c vcs, min((c+1) vcs, nsf)).
    int32_t vrs = 1, vcs = 1, vrow0 = 0, nsf = 0;
};

template <typename T>
struct WakeTable {
    int32_t n;
    int32_t total_tiles;
    WakeSurface<T> s[kMaxWakes];
};

template <typename T>
struct RecvTable {
    int32_t n;
    int32_t total;
    int32_t offset[kMaxWakes + 1];
    Receivers<T> r[kMaxWakes];
};

constexpr int kTileRings = 256;  // target rings per tile
constexpr int kMaxSpanTile = 64;

template <typename T>
__device__ __forceinline__ int find_surface(const RecvTable<T>& t, int p) {
    int s = 0;
    while (s + 1 < t.n && p >= t.offset[s + 1]) s++;
    return s;
}

// TS: storage type of the inputs and partials, T: compute type (float for MODE 2, double for MODE 3)
template <typename TS, typename T, int MODE>
__global__ __launch_bounds__(kThreads) void k_wake_partials(TS* __restrict__ partial, RecvTable<TS> rt,
                                                            WakeTable<TS> wt, T sigma) {
    // LDS: vertex rows of the tile (rh+1) x (sw+1) x 3 and circulations rh x sw.  With
    // rh = floor(kTileRings / sw) and 1 <= sw <= kMaxSpanTile, (rh+1)(sw+1) <= 2 (kTileRings+1).
    __shared__ T sverts[2 * (kTileRings + 1) * 3];
    __shared__ T sgam[kTileRings];
    __shared__ T sred[kWaves][kRecv];

    const int tile = blockIdx.y;
    int w = 0;
    while (w + 1 < wt.n && tile >= wt.s[w + 1].tile_begin) w++;
    const WakeSurface<TS>& ws = wt.s[w];
    const int local = tile - ws.tile_begin;
    const int ts = local % ws.tiles_span, tr = local / ws.tiles_span;
    const int s0 = ts * ws.sw, r0 = ws.row_begin + tr * ws.rh;
    const int swe = min(ws.sw, ws.ns - s0);
    const int rhe = min(ws.rh, ws.row_end - r0);
    const int vw = swe + 1;
    // stage vertex rows r0 .. r0+rhe and circulations
    for (int e = threadIdx.x; e < (rhe + 1) * vw; e += kThreads) {
        const int q = e / vw, sp = e % vw;
        const int64_t g = (int64_t)(s0 + sp) + (int64_t)(r0 + q) * (ws.ns + 1);
        sverts[3 * e + 0] = (T)ws.verts[g];
        sverts[3 * e + 1] = (T)ws.verts[g + ws.ws2];
        sverts[3 * e + 2] = (T)ws.verts[g + 2 * ws.ws2];
    }
    for (int e = threadIdx.x; e < rhe * swe; e += kThreads) {
        const int q = e / swe, sp = e % swe;
        sgam[e] = (T)ws.gamma[(int64_t)(r0 + q) * ws.ns + s0 + sp];
    }
    __syncthreads();

    const int lane = threadIdx.x & (kRecv - 1);
    const int wave = threadIdx.x / kRecv;
    const int p = blockIdx.x * kRecv + lane;
    T part = T(0);
    if (p < rt.total) {
        const int si = find_surface(rt, p);
        const Receivers<TS>& rc = rt.r[si];
        const int li = p - rt.offset[si];
        const v3<T> c = {(T)rc.colloc[li], (T)rc.colloc[li + rc.cs2], (T)rc.colloc[li + 2 * rc.cs2]};
        const v3<T> n = {(T)rc.normals[li], (T)rc.normals[li + rc.ns2], (T)rc.normals[li + 2 * rc.ns2]};
        v3<T> acc = {T(0), T(0), T(0)};
        for (int q = wave; q < rhe * swe; q += kWaves) {
            const int row = q / swe, sp = q % swe;
            const T* a = sverts + 3 * (row * vw + sp);
            const T* d = sverts + 3 * ((row + 1) * vw + sp);
            const T g = sgam[q];
            const v3<T> ind = ring_vel<T, MODE>(c, a, a + 3, d + 3, d, sigma);
            acc.x += ind.x * g;
            acc.y += ind.y * g;
            acc.z += ind.z * g;
        }
        part = dot(acc, n);
    }
    sred[wave][lane] = part;
    __syncthreads();
    if (wave == 0 && p < rt.total) {
        const T s = (sred[0][lane] + sred[1][lane]) + (sred[2][lane] + sred[3][lane]);
        partial[(int64_t)tile * rt.total + p] = (TS)s;
    }
}

// rhs[p] -= sum over tiles of partial[tile][p]: 64 receivers per work-group, 16 waves each sum
// every 16th tile, then a fixed-order tree over the waves (bitwise reproducible).
constexpr int kRedWaves = 16;

// sum over t = first, first + step, ... < end of x[t * ld + p], in that order, with the loads of
// eight terms issued before any is added (a plain loop waits for each load in turn).
template <typename T>
__device__ __forceinline__ T strided_sum(const T* __restrict__ x, int64_t ld, int p, int first, int end, int step) {
    T s = T(0);
    for (int t0 = first; t0 < end; t0 += 8 * step) {
        T v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int t = t0 + k * step;
            v[k] = t < end ? x[(int64_t)t * ld + p] : T(0);
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (t0 + k * step < end) s += v[k];
    }
    return s;
}

template <typename T>
__global__ __launch_bounds__(64 * kRedWaves) void k_wake_reduce(T* __restrict__ rhs, int64_t rhs_stride,
                                                                const T* __restrict__ partial, int tiles, int total) {
    __shared__ T red[kRedWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    T s = T(0);
    if (p < total) s = strided_sum(partial, (int64_t)total, p, wave, tiles, kRedWaves);
    red[wave][lane] = s;
    __syncthreads();
#pragma unroll
    for (int h = kRedWaves / 2; h > 0; h >>= 1) {
        if (wave < h) red[wave][lane] += red[wave + h][lane];
        __syncthreads();
    }
    if (wave == 0 && p < total) rhs[(int64_t)p * rhs_stride] -= red[0][lane];
}

template <typename T>
inline void launch_wake_reduce(T* rhs, int64_t rhs_stride, const T* partial, int tiles, int total, hipStream_t st) {
    hipLaunchKernelGGL(k_wake_reduce<T>, dim3((total + 63) / 64), dim3(64 * kRedWaves), 0, st, rhs, rhs_stride, partial,
                       tiles, total);
}

// Tile shape of one wake surface: spans per tile sw <= kMaxSpanTile, rows per tile rh with
// rh * sw <= kTileRings, and small enough that the launch has about wake_target_groups()
// work-groups (receiver blocks x tiles; default kTargetGroups) so every SIMD holds several waves.

inline void wake_tile_shape(int ns, int rows, int recv_blocks, int& sw, int& rh) {
    sw = ns < kMaxSpanTile ? ns : kMaxSpanTile;
    const int tiles_span = (ns + sw - 1) / sw;
    int rh_max = kTileRings / sw;
    if (rh_max < 1) rh_max = 1;
    int want_row_tiles = wake_target_groups() / (recv_blocks * tiles_span);
    if (want_row_tiles < 1) want_row_tiles = 1;
    rh = (rows + want_row_tiles - 1) / want_row_tiles;
    if (rh < 1) rh = 1;
    if (rh > rh_max) rh = rh_max;
}

// Wake induction of the seeded FP64 form evaluated on the vortex lattice.  Every segment of the
// tile is evaluated once with strength = difference of the circulations of the two rings that share
// it (spanwise segment of vertex row q, span s: G(q,s) - G(q-1,s); chordwise segment of vertex
// column s: G(q,s-1) - G(q,s); rings outside the work-group tile count 0 -- their own tile adds
// them).  The work-group tile is RH rows x up to 4*CH spans; wave w walks the ring columns of its
// span chunk left to right keeping the vertex terms of the previous vertex column in registers, so
// a ring costs (RH+1)/RH vertex terms and (2 RH + 1)/RH segments (both images share x and z).
struct LatVertex {
    double x, y, z, ri;  // c - v, 1/|c - v|
    double ym, rim;      // image receiver (y -> -y): y component and 1/|.|
};

__device__ __forceinline__ LatVertex lat_vertex(double cx, double cy, double cz, const double* v) {
    LatVertex f;
    f.x = cx - v[0];
    f.y = cy - v[1];
    f.z = cz - v[2];
    f.ym = -cy - v[1];
    const double xz = f.x * f.x + f.z * f.z;
    f.ri = rsqrt_seeded(xz + f.y * f.y);
    f.rim = rsqrt_seeded(xz + f.ym * f.ym);
    return f;
}

// Segment a->b of strength g for the receiver and for its image in y = 0, without the 1/(4 pi) factor
// (lattice_chunk applies it once to its sum).  The two evaluations share every term without a y
// component: r0 = b - a (taken as (c - a) - (c - b), whose y part is the same for the image), the y
// component of r1 x r2 and its square, and the x and z parts of r0.r1 and r0.r2.
__device__ __forceinline__ void lat_pair(double& ax, double& ay, double& az, const LatVertex& a, const LatVertex& b,
                                         double g) {
    const double r0x = a.x - b.x, r0y = a.y - b.y, r0z = a.z - b.z;
    const double cry = a.z * b.x - a.x * b.z;
    const double cry2 = cry * cry;
    const double d1xz = r0x * a.x + r0z * a.z;
    const double d2xz = r0x * b.x + r0z * b.z;
    constexpr double cut = BS<double>::cut_cross;
    // |r| < RCUT  <=>  1/|r| > 1/RCUT
    constexpr double rimax = 1.0 / BS<double>::cut_norm;
    {
        const double crx = a.y * b.z - a.z * b.y;
        const double crz = a.x * b.y - a.y * b.x;
        const double sq = crx * crx + crz * crz + cry2;
        const double d1 = d1xz + r0y * a.y, d2 = d2xz + r0y * b.y;
        const bool ok = (sq >= cut) & (a.ri <= rimax) & (b.ri <= rimax);
        const double t = (d1 * a.ri - d2 * b.ri) * rcp_seeded(sq) * g;
        const double coeff = ok ? t : 0.0;
        ax += coeff * crx;
        ay += coeff * cry;
        az += coeff * crz;
    }
    {
        const double crx = a.ym * b.z - a.z * b.ym;
        const double crz = a.x * b.ym - a.ym * b.x;
        const double sq = crx * crx + crz * crz + cry2;
        const double d1 = d1xz + r0y * a.ym, d2 = d2xz + r0y * b.ym;
        const bool ok = (sq >= cut) & (a.rim <= rimax) & (b.rim <= rimax);
        const double t = (d1 * a.rim - d2 * b.rim) * rcp_seeded(sq) * g;
        const double coeff = ok ? t : 0.0;
        ax += coeff * crx;
        ay -= coeff * cry;
        az += coeff * crz;
    }
}

constexpr int kLatChunkMax = 16;  // spans per wave

// The lattice walk of one wave over its span chunk [sa, sb) of a staged tile (vertex rows 0..rhe,
// vw = swe + 1 vertex columns, 3 coordinates each in sverts; circulations rhe x swe in sgam) for
// the receiver (cx, cy, cz): chordwise segments of each vertex column with the circulation
// difference of the two ring columns beside it, spanwise segments of each vertex row with the
// difference of the two ring rows, the previous vertex column's terms kept in registers; the tile's
// right edge only when sb == swe.  Adds the induced velocity to (vx, vy, vz).
template <int RH>
__device__ __forceinline__ void lattice_chunk(const double* sverts, const double* sgam, int swe, int rhe, int vw,
                                              int sa, int sb, double cx, double cy, double cz, double& vx, double& vy,
                                              double& vz) {
    double ax = 0.0, ay = 0.0, az = 0.0;  // 4 pi x the induced velocity (lat_pair leaves the factor out)
    LatVertex prev[RH + 1], cur[RH + 1];
#pragma unroll
    for (int q = 0; q <= RH; q++)
        if (q <= rhe) prev[q] = lat_vertex(cx, cy, cz, sverts + 3 * (q * vw + sa));
    for (int s = sa; s < sb; s++) {
        // chordwise segments of vertex column s (left edge of ring column s)
#pragma unroll
        for (int q = 0; q < RH; q++) {
            if (q < rhe) {
                double g = -sgam[q * swe + s];
                if (s > 0) g += sgam[q * swe + s - 1];
                lat_pair(ax, ay, az, prev[q], prev[q + 1], g);
            }
        }
        // vertex column s+1 and the spanwise segments of ring column s
#pragma unroll
        for (int q = 0; q <= RH; q++) {
            if (q <= rhe) {
                cur[q] = lat_vertex(cx, cy, cz, sverts + 3 * (q * vw + s + 1));
                double g = 0.0;
                if (q < rhe) g += sgam[q * swe + s];
                if (q > 0) g -= sgam[(q - 1) * swe + s];
                lat_pair(ax, ay, az, prev[q], cur[q], g);
            }
        }
#pragma unroll
        for (int q = 0; q <= RH; q++) prev[q] = cur[q];
    }
    if (sb == swe) {  // right edge of the work-group tile: chordwise segments of vertex column swe
#pragma unroll
        for (int q = 0; q < RH; q++)
            if (q < rhe) lat_pair(ax, ay, az, prev[q], prev[q + 1], sgam[q * swe + swe - 1]);
    }
    constexpr double inv_fourpi = 1.0 / BS<double>::fourpi;
    vx += ax * inv_fourpi;
    vy += ay * inv_fourpi;
    vz += az * inv_fourpi;
}

// zs: simulations (blockIdx.z) of one launch find their receivers, wake and circulations z * zs
// elements after simulation 0's (the tables' pointers) and write their partials z * wzs after.
// At least 4 waves per SIMD: the hint once held the 4-row form at 128 registers when it needed 130
// unhinted (3 waves); since a segment and its image share their y-free terms it takes 126 either way,
// and the hint keeps a later change from dropping a wave unnoticed.
template <typename TS, int RH>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_wake_lattice(TS* __restrict__ partial, RecvTable<TS> rt,
                                                           WakeTable<TS> wt, int64_t zs = 0, int64_t wzs = 0) {
    const int64_t zo = (int64_t)blockIdx.z * zs;
    partial += (int64_t)blockIdx.z * wzs;
    constexpr int SWG = kWaves * kLatChunkMax;
    __shared__ double sverts[(RH + 1) * (SWG + 1) * 3];
    __shared__ double sgam[RH * SWG];
    __shared__ double sred[kWaves][kRecv];
    const int tile = blockIdx.y;
    int w = 0;
    while (w + 1 < wt.n && tile >= wt.s[w + 1].tile_begin) w++;
    const WakeSurface<TS>& ws = wt.s[w];
    const int local = tile - ws.tile_begin;
    const int ts = local % ws.tiles_span, tr = local / ws.tiles_span;
    const int s0 = ts * ws.sw, r0 = ws.row_begin + tr * ws.rh;
    const int swe = min(ws.sw, ws.ns - s0);
    const int rhe = min(ws.rh, ws.row_end - r0);
    const int vw = swe + 1;
    const int nsf = ws.nsf ? ws.nsf : ws.ns;
    for (int e = threadIdx.x; e < (rhe + 1) * vw; e += kThreads) {
        const int q = e / vw, sp = e % vw;
        const int col = min((s0 + sp) * ws.vcs, nsf);
        const int64_t g = zo + (int64_t)col + (int64_t)(ws.vrow0 + (r0 + q) * ws.vrs) * (nsf + 1);
        sverts[3 * e + 0] = (double)ws.verts[g];
        sverts[3 * e + 1] = (double)ws.verts[g + ws.ws2];
        sverts[3 * e + 2] = (double)ws.verts[g + 2 * ws.ws2];
    }
    if (ws.vrs == 1 && ws.vcs == 1) {
        for (int e = threadIdx.x; e < rhe * swe; e += kThreads) {
            const int q = e / swe, sp = e % swe;
            sgam[e] = (double)ws.gamma[zo + (int64_t)(ws.vrow0 + r0 + q) * nsf + s0 + sp];
        }
    } else {  // coarse ring: mean of the fine rings it replaces, rows then spans in order
        for (int e = threadIdx.x; e < rhe * swe; e += kThreads) {
            const int q = e / swe, sp = e % swe;
            const int fr0 = ws.vrow0 + (r0 + q) * ws.vrs, fc0 = (s0 + sp) * ws.vcs;
            const int fc1 = min(fc0 + ws.vcs, nsf);
            double sum = 0.0;
            for (int fr = fr0; fr < fr0 + ws.vrs; fr++)
                for (int fc = fc0; fc < fc1; fc++) sum += (double)ws.gamma[zo + (int64_t)fr * nsf + fc];
            sgam[e] = sum / (double)(ws.vrs * (fc1 - fc0));
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & (kRecv - 1);
    const int wave = threadIdx.x / kRecv;
    const int p = blockIdx.x * kRecv + lane;
    double part = 0.0;
    const int chunk = (swe + kWaves - 1) / kWaves;
    const int sa = wave * chunk, sb = min(swe, sa + chunk);
    if (p < rt.total && sa < sb) {
        const int si = find_surface(rt, p);
        const Receivers<TS>& rc = rt.r[si];
        const int li = p - rt.offset[si];
        const TS* co = rc.colloc + zo;
        const TS* no = rc.normals + zo;
        const double cx = co[li], cy = co[li + rc.cs2], cz = co[li + 2 * rc.cs2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        lattice_chunk<RH>(sverts, sgam, swe, rhe, vw, sa, sb, cx, cy, cz, ax, ay, az);
        part = ax * (double)no[li] + ay * (double)no[li + rc.ns2] + az * (double)no[li + 2 * rc.ns2];
    }
    sred[wave][lane] = part;
    __syncthreads();
    if (wave == 0 && p < rt.total)
        partial[(int64_t)tile * rt.total + p] = (TS)((sred[0][lane] + sred[1][lane]) + (sred[2][lane] + sred[3][lane]));
}

// Tile shape for k_wake_lattice: sw <= 4 * kLatChunkMax spans, rh in {1, 2, 4} rows, chosen so
// that the launch has about kLatTargetGroups work-groups.

inline void lattice_tile_shape(int ns, int rows, int recv_blocks, int& sw, int& rh) {
    sw = ns < kWaves * kLatChunkMax ? ns : kWaves * kLatChunkMax;
    const int tiles_span = (ns + sw - 1) / sw;
    int want_row_tiles = kLatTargetGroups / (recv_blocks * tiles_span);
    if (want_row_tiles < 1) want_row_tiles = 1;
    const int need = (rows + want_row_tiles - 1) / want_row_tiles;
    rh = need <= 1 ? 1 : need <= 2 ? 2 : 4;
}

// Span extent of a lattice tile when the launch's spans are split `split` ways: ceil(ns / split)
// spans, at least 2 per wave (or all ns), at most kWaves * kLatChunkMax.
inline int lattice_span_tile(int ns, int split) {
    int sw = (ns + split - 1) / split;
    const int lo = ns < 2 * kWaves ? ns : 2 * kWaves;
    if (sw < lo) sw = lo;
    if (sw > kWaves * kLatChunkMax) sw = kWaves * kLatChunkMax;
    return sw;
}

// One shape for every surface of a k_wake_lattice launch (rows per tile rh, the kernel's template
// value, and a span split): a launch of more than `target` work-groups with one row per tile takes
// 2 or 4 rows per tile (the smallest that gets within the target, else 4); a smaller launch splits
// its spans (2, 3, ... ways, at least 2 spans per wave) until it reaches the target or cannot split
// further, so that a short wake still spreads over the CUs.  A split launch holds fewer than
// 2 target work-groups.
template <typename T>
inline void lattice_launch_shape(const WakeTable<T>& wt, int recv_blocks, int target, int& rh, int& split) {
    auto groups = [&](int rh_, int k) {
        int64_t g = 0;
        for (int m = 0; m < wt.n; m++) {
            const WakeSurface<T>& w = wt.s[m];
            const int sw = lattice_span_tile(w.ns, k);
            g += (int64_t)recv_blocks * ((w.ns + sw - 1) / sw) * ((w.row_end - w.row_begin + rh_ - 1) / rh_);
        }
        return g;
    };
    rh = 1;
    split = 1;
    if (groups(1, 1) > target) {
        rh = groups(2, 1) <= target ? 2 : 4;
        return;
    }
    for (int k = 2; k <= 64 && groups(1, split) < target; k++) {
        bool finer = false;  // does k split any surface finer than the current split?
        for (int m = 0; m < wt.n; m++)
            if (lattice_span_tile(wt.s[m].ns, k) < lattice_span_tile(wt.s[m].ns, split)) finer = true;
        if (!finer) continue;
        if (groups(1, k) > 2 * (int64_t)target) break;
        split = k;
    }
}

}  // namespace
}  // namespace tvlm

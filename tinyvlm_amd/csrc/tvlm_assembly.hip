// tvlm_assembly.hip -- influence-matrix assembly and wake induction on gfx950.
//
// lhs_assemble: every (receiver, ring) pair is independent.  A work-group owns 64 receivers
// (one per lane, colloc + normal in registers) and TC rings whose 12 vertex coordinates are staged
// in LDS; its 4 waves take every 4th ring, so each wave reads LDS as a broadcast and writes one
// coalesced 64-element column segment per ring.  The work is VALU-bound (16 sqrt + 8 div per pair
// in the exact forms, 16 rsq + 8 rcp in the fast FP32 form).
//
// rhs_assemble_wake_influence: receivers x wake rings with a Gamma-weighted sum per receiver.
// The wake rings are cut into tiles of RH rows x SW spans; a work-group stages one tile's vertex
// rows and circulations in LDS and reduces its 4 waves in LDS to one partial per receiver.  A
// second kernel sums the partials of all tiles in a fixed order, so results are bitwise
// reproducible (the CUDA reference uses float atomics, vlm_backend_cuda_kernels.cuh:370).
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

// Fast FP32 ring evaluation at one receiver (with image) -> dot with normal.
__device__ __forceinline__ float ring_dot_fast(float cx, float cy, float cz, float nx, float ny, float nz, const float* v) {
    float ax = 0.f, ay = 0.f, az = 0.f;
    fast_ring(ax, ay, az, cx, cy, cz, v, v + 3, v + 6, v + 9);
    return ax * nx + ay * ny + az * nz;
}

template <typename T, bool SIGMA>
__device__ __forceinline__ T ring_dot_exact(const v3<T>& c, const v3<T>& n, const T* v, T sigma) {
    const v3<T> inf = ring_exact<T, SIGMA>(c, {v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]},
                                           {v[9], v[10], v[11]}, sigma);
    return dot(inf, n);
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

// lhs block of one (receiver surface, ring surface) pair.  lhs points at (offset_i, offset_j).
template <typename T, int TC, int MODE>  // MODE 0: exact sigma=0, 1: exact with sigma, 2: fast fp32
__global__ __launch_bounds__(kThreads) void k_lhs_block(T* __restrict__ lhs, int64_t ld, Receivers<T> rc,
                                                        RingSource<T> src, T sigma) {
    __shared__ T sv[TC][12];
    const int n_rings = src.ns * src.nc;
    const int col0 = blockIdx.y * TC;
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
    const int i = blockIdx.x * kRecv + lane;
    if (i >= rc.m) return;
    const v3<T> c = {rc.colloc[i], rc.colloc[i + rc.cs2], rc.colloc[i + 2 * rc.cs2]};
    const v3<T> n = {rc.normals[i], rc.normals[i + rc.ns2], rc.normals[i + 2 * rc.ns2]};
    const int ncols = min(TC, n_rings - col0);
    const int te0 = (src.nc - 1) * src.ns;
    for (int r = wave; r < ncols; r += kWaves) {
        const int lidx = col0 + r;
        T* dst = lhs + i + (int64_t)lidx * ld;
        T acc = *dst;
        if constexpr (MODE == 2) {
            acc += ring_dot_fast(c.x, c.y, c.z, n.x, n.y, n.z, sv[r]);
        } else {
            acc += ring_dot_exact<T, MODE == 1>(c, n, sv[r], sigma);
        }
        if (src.iteration > 0 && lidx >= te0) {
            // trailing-edge column: wake rings of rows nwv-2-k, k < iteration (vlm_backend_cpu.cpp:121-139)
            const int j = lidx - te0;
            for (int k = 0; k < src.iteration; k++) {
                T w[12];
                load_ring(w, src.wake, src.ws2, (int64_t)j + (int64_t)(src.nwv - 2 - k) * (src.ns + 1), src.ns);
                if constexpr (MODE == 2) {
                    acc += ring_dot_fast(c.x, c.y, c.z, n.x, n.y, n.z, w);
                } else {
                    acc += ring_dot_exact<T, MODE == 1>(c, n, w, sigma);
                }
            }
        }
        *dst = acc;
    }
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

template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void k_wake_partials(T* __restrict__ partial, RecvTable<T> rt, WakeTable<T> wt,
                                                            T sigma) {
    // LDS: vertex rows of the tile (rh+1) x (sw+1) x 3 and circulations rh x sw.  With
    // rh = floor(kTileRings / sw) and 1 <= sw <= kMaxSpanTile, (rh+1)(sw+1) <= 2 (kTileRings+1).
    __shared__ T sverts[2 * (kTileRings + 1) * 3];
    __shared__ T sgam[kTileRings];
    __shared__ T sred[kWaves][kRecv];

    const int tile = blockIdx.y;
    int w = 0;
    while (w + 1 < wt.n && tile >= wt.s[w + 1].tile_begin) w++;
    const WakeSurface<T>& ws = wt.s[w];
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
        sverts[3 * e + 0] = ws.verts[g];
        sverts[3 * e + 1] = ws.verts[g + ws.ws2];
        sverts[3 * e + 2] = ws.verts[g + 2 * ws.ws2];
    }
    for (int e = threadIdx.x; e < rhe * swe; e += kThreads) {
        const int q = e / swe, sp = e % swe;
        sgam[e] = ws.gamma[(int64_t)(r0 + q) * ws.ns + s0 + sp];
    }
    __syncthreads();

    const int lane = threadIdx.x & (kRecv - 1);
    const int wave = threadIdx.x / kRecv;
    const int p = blockIdx.x * kRecv + lane;
    T part = T(0);
    if (p < rt.total) {
        const int si = find_surface(rt, p);
        const Receivers<T>& rc = rt.r[si];
        const int li = p - rt.offset[si];
        const v3<T> c = {rc.colloc[li], rc.colloc[li + rc.cs2], rc.colloc[li + 2 * rc.cs2]};
        const v3<T> n = {rc.normals[li], rc.normals[li + rc.ns2], rc.normals[li + 2 * rc.ns2]};
        v3<T> acc = {T(0), T(0), T(0)};
        for (int q = wave; q < rhe * swe; q += kWaves) {
            const int row = q / swe, sp = q % swe;
            const T* a = sverts + 3 * (row * vw + sp);
            const T* d = sverts + 3 * ((row + 1) * vw + sp);
            const T g = sgam[q];
            if constexpr (MODE == 2) {
                float ax = 0.f, ay = 0.f, az = 0.f;
                fast_ring(ax, ay, az, c.x, c.y, c.z, a, a + 3, d + 3, d);
                acc.x += ax * g;
                acc.y += ay * g;
                acc.z += az * g;
            } else {
                const v3<T> ind = ring_exact<T, MODE == 1>(c, {a[0], a[1], a[2]}, {a[3], a[4], a[5]},
                                                          {d[3], d[4], d[5]}, {d[0], d[1], d[2]}, sigma);
                acc.x += ind.x * g;
                acc.y += ind.y * g;
                acc.z += ind.z * g;
            }
        }
        part = dot(acc, n);
    }
    sred[wave][lane] = part;
    __syncthreads();
    if (wave == 0 && p < rt.total) {
        const T s = (sred[0][lane] + sred[1][lane]) + (sred[2][lane] + sred[3][lane]);
        partial[(int64_t)tile * rt.total + p] = s;
    }
}

template <typename T>
__global__ void k_wake_reduce(T* __restrict__ rhs, int64_t rhs_stride, const T* __restrict__ partial, int tiles,
                              int total) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    T s = T(0);
    for (int t = 0; t < tiles; t++) s += partial[(int64_t)t * total + p];
    rhs[(int64_t)p * rhs_stride] -= s;
}

template <typename T>
int validate_surface3(const tvlm_view& v, int64_t a, int64_t b, const char* what, int32_t m) {
    TVLM_REQUIRE(plane_contiguous(v), "%s[%d]: expected stride[0]==1 and stride[1]==shape[0]", what, m);
    TVLM_REQUIRE(v.shape[0] == a && v.shape[1] == b, "%s[%d]: shape (%lld,%lld) != (%lld,%lld)", what, m,
                 (long long)v.shape[0], (long long)v.shape[1], (long long)a, (long long)b);
    return TVLM_OK;
}

template <typename T>
int lhs_assemble(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                 const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, T sigma, void* stream) {
    TVLM_REQUIRE(lhs && colloc && normals && verts_wing && verts_wake, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    TVLM_REQUIRE(iteration >= 0, "negative iteration");
    TVLM_REQUIRE(lhs->stride[0] == 1, "lhs must be column-major with unit stride[0]");
    int64_t total = 0;
    for (int m = 0; m < n_surf; m++) {
        const int64_t ns = colloc[m].shape[0], nc = colloc[m].shape[1];
        TVLM_TRY(validate_surface3<T>(colloc[m], ns, nc, "colloc", m));
        TVLM_TRY(validate_surface3<T>(normals[m], ns, nc, "normals", m));
        TVLM_TRY(validate_surface3<T>(verts_wing[m], ns + 1, nc + 1, "verts_wing", m));
        if (iteration > 0) {
            TVLM_TRY(validate_surface3<T>(verts_wake[m], ns + 1, verts_wake[m].shape[1], "verts_wake", m));
            TVLM_REQUIRE(verts_wake[m].shape[1] >= iteration + 1, "verts_wake[%d] has %lld rows < iteration+1", m,
                         (long long)verts_wake[m].shape[1]);
        }
        total += ns * nc;
    }
    TVLM_REQUIRE(lhs->shape[0] >= total && lhs->shape[1] >= total, "lhs smaller than the total panel count %lld",
                 (long long)total);
    const int64_t ld = lhs->stride[1];
    hipStream_t st = as_stream(stream);
    int64_t offset_j = 0;
    for (int mj = 0; mj < n_surf; mj++) {
        RingSource<T> src;
        src.verts = static_cast<const T*>(verts_wing[mj].ptr);
        src.vs2 = verts_wing[mj].stride[2];
        src.ns = (int32_t)colloc[mj].shape[0];
        src.nc = (int32_t)colloc[mj].shape[1];
        src.wake = static_cast<const T*>(verts_wake[mj].ptr);
        src.ws2 = verts_wake[mj].stride[2];
        src.nwv = (int32_t)verts_wake[mj].shape[1];
        src.iteration = iteration;
        const int n_rings = src.ns * src.nc;
        int64_t offset_i = 0;
        for (int mi = 0; mi < n_surf; mi++) {
            Receivers<T> rc;
            rc.colloc = static_cast<const T*>(colloc[mi].ptr);
            rc.cs2 = colloc[mi].stride[2];
            rc.normals = static_cast<const T*>(normals[mi].ptr);
            rc.ns2 = normals[mi].stride[2];
            rc.m = (int32_t)(colloc[mi].shape[0] * colloc[mi].shape[1]);
            T* section = static_cast<T*>(lhs->ptr) + offset_i + offset_j * ld;
            constexpr int TC = sizeof(T) == 8 ? 32 : 64;
            dim3 grid((rc.m + kRecv - 1) / kRecv, (n_rings + TC - 1) / TC);
            if constexpr (sizeof(T) == 4) {
                if (sigma != T(0)) {
                    // the fast form has no core model; use the exact float form for sigma != 0
                    hipLaunchKernelGGL((k_lhs_block<T, TC, 1>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
                } else {
                    hipLaunchKernelGGL((k_lhs_block<T, TC, 2>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
                }
            } else {
                if (sigma != T(0)) {
                    hipLaunchKernelGGL((k_lhs_block<T, TC, 1>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
                } else {
                    hipLaunchKernelGGL((k_lhs_block<T, TC, 0>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
                }
            }
            TVLM_LAUNCHED();
            offset_i += rc.m;
        }
        offset_j += n_rings;
    }
    return TVLM_OK;
}

template <typename T>
int wake_influence(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake, const tvlm_view* colloc,
                   const tvlm_view* normals, const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                   T sigma, void* stream) {
    TVLM_REQUIRE(rhs && gamma_wake && colloc && normals && verts_wake && lifting, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    TVLM_REQUIRE(iteration >= 0, "negative iteration");
    RecvTable<T> rt{};
    rt.n = n_surf;
    int total = 0;
    for (int m = 0; m < n_surf; m++) {
        const int64_t ns = colloc[m].shape[0], nc = colloc[m].shape[1];
        TVLM_TRY(validate_surface3<T>(colloc[m], ns, nc, "colloc", m));
        TVLM_TRY(validate_surface3<T>(normals[m], ns, nc, "normals", m));
        rt.offset[m] = total;
        rt.r[m] = {static_cast<const T*>(colloc[m].ptr), colloc[m].stride[2], static_cast<const T*>(normals[m].ptr),
                   normals[m].stride[2], (int32_t)(ns * nc)};
        total += (int)(ns * nc);
    }
    rt.offset[n_surf] = total;
    rt.total = total;
    TVLM_REQUIRE(rhs->shape[0] >= total, "rhs shorter than the panel count");
    WakeTable<T> wt{};
    int tiles = 0;
    for (int m = 0; m < n_surf; m++) {
        if (!lifting[m] || iteration == 0) continue;
        const int64_t ns = colloc[m].shape[0];
        const tvlm_view& wv = verts_wake[m];
        TVLM_TRY(validate_surface3<T>(wv, ns + 1, wv.shape[1], "verts_wake", m));
        const int32_t nwv = (int32_t)wv.shape[1];
        TVLM_REQUIRE(nwv >= iteration + 1, "verts_wake[%d]: %d vertex rows < iteration+1", m, nwv);
        TVLM_REQUIRE(gamma_wake[m].stride[0] == 1 && gamma_wake[m].stride[1] == ns && gamma_wake[m].shape[0] == ns &&
                         gamma_wake[m].shape[1] >= nwv - 1,
                     "gamma_wake[%d]: expected contiguous {%lld, >=%d}", m, (long long)ns, nwv - 1);
        WakeSurface<T>& s = wt.s[wt.n];
        s.verts = static_cast<const T*>(wv.ptr);
        s.ws2 = wv.stride[2];
        s.gamma = static_cast<const T*>(gamma_wake[m].ptr);
        s.ns = (int32_t)ns;
        s.nwv = nwv;
        s.row_begin = nwv - 1 - iteration;
        s.row_end = nwv - 1;
        if (sizeof(T) == 4 && iteration > 5000) s.row_end = nwv - 1 - iteration + 5000;  // kernels.ispc:157-162
        s.sw = (int32_t)(ns < kMaxSpanTile ? ns : kMaxSpanTile);
        s.rh = kTileRings / s.sw;
        if (s.rh < 1) s.rh = 1;
        s.tiles_span = (s.ns + s.sw - 1) / s.sw;
        const int tiles_row = (s.row_end - s.row_begin + s.rh - 1) / s.rh;
        s.tile_begin = tiles;
        tiles += s.tiles_span * tiles_row;
        wt.n++;
    }
    wt.total_tiles = tiles;
    if (tiles == 0 || total == 0) return TVLM_OK;
    void* ws = nullptr;
    TVLM_TRY(device_workspace(sizeof(T) * (size_t)tiles * (size_t)total, &ws));
    hipStream_t st = as_stream(stream);
    dim3 grid((total + kRecv - 1) / kRecv, tiles);
    if constexpr (sizeof(T) == 4) {
        if (sigma != T(0))
            hipLaunchKernelGGL((k_wake_partials<T, 1>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
        else
            hipLaunchKernelGGL((k_wake_partials<T, 2>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
    } else {
        if (sigma != T(0))
            hipLaunchKernelGGL((k_wake_partials<T, 1>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
        else
            hipLaunchKernelGGL((k_wake_partials<T, 0>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
    }
    TVLM_LAUNCHED();
    hipLaunchKernelGGL(k_wake_reduce<T>, dim3((total + 255) / 256), dim3(256), 0, st, static_cast<T*>(rhs->ptr),
                       rhs->stride[0], (const T*)ws, tiles, total);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

}  // namespace
}  // namespace tvlm

extern "C" {

int tvlm_lhs_assemble_f64(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                          const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, double sigma,
                          void* stream) {
    return tvlm::lhs_assemble<double>(lhs, n_surf, colloc, normals, verts_wing, verts_wake, iteration, sigma, stream);
}
int tvlm_lhs_assemble_f32(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                          const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, float sigma,
                          void* stream) {
    return tvlm::lhs_assemble<float>(lhs, n_surf, colloc, normals, verts_wing, verts_wake, iteration, sigma, stream);
}
int tvlm_rhs_assemble_wake_influence_f64(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                         const tvlm_view* colloc, const tvlm_view* normals, const tvlm_view* verts_wake,
                                         const uint8_t* lifting, int32_t iteration, double sigma, void* stream) {
    return tvlm::wake_influence<double>(rhs, n_surf, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                                        stream);
}
int tvlm_rhs_assemble_wake_influence_f32(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                         const tvlm_view* colloc, const tvlm_view* normals, const tvlm_view* verts_wake,
                                         const uint8_t* lifting, int32_t iteration, float sigma, void* stream) {
    return tvlm::wake_influence<float>(rhs, n_surf, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                                       stream);
}

}  // extern "C"

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
#include <mutex>

#include "tvlm_kernels.hpp"

namespace tvlm {
namespace {

template <typename T>
int validate_surface3(const tvlm_view& v, int64_t a, int64_t b, const char* what, int32_t m) {
    TVLM_REQUIRE(plane_contiguous(v), "%s[%d]: expected stride[0]==1 and stride[1]==shape[0]", what, m);
    TVLM_REQUIRE(v.shape[0] == a && v.shape[1] == b, "%s[%d]: shape (%lld,%lld) != (%lld,%lld)", what, m,
                 (long long)v.shape[0], (long long)v.shape[1], (long long)a, (long long)b);
    return TVLM_OK;
}

// The (receiver surface, ring surface) sections of one influence matrix, checked against the
// reference's shapes, appended to `out` with their tiles numbered from `tiles` on (advanced).
template <typename T>
int lhs_sections(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                 const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration,
                 std::vector<LhsSection<T>>& out, int64_t& tiles) {
    TVLM_REQUIRE(lhs && colloc && normals && verts_wing && verts_wake, "null argument");
    TVLM_REQUIRE(n_surf >= 1 && n_surf <= TVLM_MAX_SURFACES, "n_surf %d out of range", n_surf);
    TVLM_REQUIRE(iteration >= 0, "negative iteration");
    TVLM_REQUIRE(lhs->ptr != nullptr && lhs->stride[0] == 1, "lhs must be column-major with unit stride[0]");
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
    constexpr int TC = sizeof(T) == 8 ? 32 : 64;
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
            LhsSection<T> sec;
            sec.rc.colloc = static_cast<const T*>(colloc[mi].ptr);
            sec.rc.cs2 = colloc[mi].stride[2];
            sec.rc.normals = static_cast<const T*>(normals[mi].ptr);
            sec.rc.ns2 = normals[mi].stride[2];
            sec.rc.m = (int32_t)(colloc[mi].shape[0] * colloc[mi].shape[1]);
            sec.src = src;
            sec.lhs = static_cast<T*>(lhs->ptr) + offset_i + offset_j * ld;
            sec.ld = ld;
            sec.tiles_x = (sec.rc.m + kRecv - 1) / kRecv;
            const int64_t tiles_y = (n_rings + TC - 1) / TC;
            TVLM_REQUIRE(tiles + (int64_t)sec.tiles_x * tiles_y < (int64_t)1 << 31, "more than 2^31 tiles");
            sec.tile0 = (int32_t)tiles;
            tiles += (int64_t)sec.tiles_x * tiles_y;
            out.push_back(sec);
            offset_i += sec.rc.m;
        }
        offset_j += n_rings;
    }
    return TVLM_OK;
}

// seeded = true (FP64 only): the FP64 seeded form (hardware 1/sqrt and 1/x estimates + one Newton
// step) instead of the exact form; used by the served steady solve, which refines against the exact
// FP64 residual anyway.
template <typename T>
int lhs_assemble(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                 const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, T sigma, void* stream,
                 bool seeded = false) {
    std::vector<LhsSection<T>> secs;
    int64_t tiles = 0;
    TVLM_TRY(lhs_sections<T>(lhs, n_surf, colloc, normals, verts_wing, verts_wake, iteration, secs, tiles));
    hipStream_t st = as_stream(stream);
    constexpr int TC = sizeof(T) == 8 ? 32 : 64;
    for (const LhsSection<T>& sc : secs) {
        T* section = sc.lhs;
        const int64_t ld = sc.ld;
        const Receivers<T>& rc = sc.rc;
        const RingSource<T>& src = sc.src;
        dim3 grid(sc.tiles_x, (src.ns * src.nc + TC - 1) / TC);
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
            } else if (seeded) {
                hipLaunchKernelGGL((k_lhs_block<T, TC, 3>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
            } else {
                hipLaunchKernelGGL((k_lhs_block<T, TC, 0>), grid, dim3(kThreads), 0, st, section, ld, rc, src, sigma);
            }
        }
        TVLM_LAUNCHED();
    }
    return TVLM_OK;
}

// fixed_ws: when non-null, the tile partials go to this caller-provided device buffer of
// fixed_bytes (nothing is allocated; a launch that needs more is rejected before any work is
// queued); when null, to the library's per-stream workspace (grown on demand).
// lattice (FP64 only): the wake's vortex lattice in the FP64 seeded form (k_wake_lattice, the served
// UVLM loop's kernel: every lattice segment once, with the difference of the circulations of the
// two rings beside it) instead of ring by ring.
template <typename T>
int wake_influence(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake, const tvlm_view* colloc,
                   const tvlm_view* normals, const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                   T sigma, void* stream, void* fixed_ws = nullptr, size_t fixed_bytes = 0, bool lattice = false) {
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
        wake_tile_shape(s.ns, s.row_end - s.row_begin, (total + kRecv - 1) / kRecv, s.sw, s.rh);
        s.tiles_span = (s.ns + s.sw - 1) / s.sw;
        const int tiles_row = (s.row_end - s.row_begin + s.rh - 1) / s.rh;
        s.tile_begin = tiles;
        tiles += s.tiles_span * tiles_row;
        wt.n++;
    }
    wt.total_tiles = tiles;
    if (tiles == 0 || total == 0) return TVLM_OK;
    const int recv_blocks = (total + kRecv - 1) / kRecv;
    int lat_rh = 1;
    if (lattice) {  // one rows-per-tile for the launch, spans split while it holds few work-groups
        TVLM_REQUIRE(sizeof(T) == 8 && sigma == T(0), "the lattice form is FP64 without a core model");
        int split = 1;
        lattice_launch_shape(wt, recv_blocks, lattice_target_groups(), lat_rh, split);
        tiles = 0;
        for (int m = 0; m < wt.n; m++) {
            WakeSurface<T>& w = wt.s[m];
            w.rh = lat_rh;
            w.sw = lattice_span_tile(w.ns, split);
            w.tiles_span = (w.ns + w.sw - 1) / w.sw;
            w.tile_begin = tiles;
            tiles += w.tiles_span * ((w.row_end - w.row_begin + lat_rh - 1) / lat_rh);
        }
        wt.total_tiles = tiles;
    }
    void* ws = nullptr;
    hipStream_t st = as_stream(stream);
    const size_t need = sizeof(T) * (size_t)tiles * (size_t)total;
    if (fixed_ws) {
        TVLM_REQUIRE(need <= fixed_bytes,
                     "wake induction needs %zu bytes of tile partials (%d tiles x %d receivers), above the %zu-byte "
                     "fixed scratch of the allocation-free entry point",
                     need, tiles, total, fixed_bytes);
        ws = fixed_ws;
    } else {
        TVLM_TRY(device_workspace(need, &ws, kWsWakePartials, st));
    }
    dim3 grid((total + kRecv - 1) / kRecv, tiles);
    if constexpr (sizeof(T) == 8) {
        if (lattice) {
            if (lat_rh == 1)
                hipLaunchKernelGGL((k_wake_lattice<T, 1>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, 0, 0);
            else if (lat_rh == 2)
                hipLaunchKernelGGL((k_wake_lattice<T, 2>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, 0, 0);
            else
                hipLaunchKernelGGL((k_wake_lattice<T, 4>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, 0, 0);
            TVLM_LAUNCHED();
            launch_wake_reduce<T>(static_cast<T*>(rhs->ptr), rhs->stride[0], (const T*)ws, tiles, total, st);
            TVLM_LAUNCHED();
            return TVLM_OK;
        }
    }
    if constexpr (sizeof(T) == 4) {
        if (sigma != T(0))
            hipLaunchKernelGGL((k_wake_partials<T, T, 1>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
        else
            hipLaunchKernelGGL((k_wake_partials<T, T, 2>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
    } else {
        if (sigma != T(0))
            hipLaunchKernelGGL((k_wake_partials<T, T, 1>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
        else
            hipLaunchKernelGGL((k_wake_partials<T, T, 0>), grid, dim3(kThreads), 0, st, (T*)ws, rt, wt, sigma);
    }
    TVLM_LAUNCHED();
    launch_wake_reduce<T>(static_cast<T*>(rhs->ptr), rhs->stride[0], (const T*)ws, tiles, total, st);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

// ---- allocation-free wake induction ------------------------------------------------------------
// The tile partials of tvlm_rhs_assemble_wake_influence_noalloc_f64 live in this fixed device array
// (each device that runs the library holds its own copy until the process ends).  Host code never
// passes the array's name to a kernel: its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_dev.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up on the first call per device
int scratch_address(void** out) {
    int dev = 0;
    TVLM_HIP(hipGetDevice(&dev));
    TVLM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d devices the library tracks", dev,
                 kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        TVLM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        TVLM_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned a null address for the scratch array");
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return TVLM_OK;
}

}  // namespace

int noalloc_scratch_setup() {
    void* ptr = nullptr;
    return scratch_address(&ptr);
}

int lhs_assemble_seeded_f64(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc, const tvlm_view* normals,
                            const tvlm_view* verts_wing, const tvlm_view* verts_wake, int32_t iteration, void* stream) {
    return lhs_assemble<double>(lhs, n_surf, colloc, normals, verts_wing, verts_wake, iteration, 0.0, stream, true);
}

}  // namespace tvlm

// Sections of several FP64 influence matrices, item by item, on the device for grouped launches.
struct tvlm_lhs_group {
    tvlm::LhsSection<double>* secs = nullptr;  // device
    int32_t* tile0 = nullptr;                  // device: first tile of every section, then the total
    int32_t* chunk = nullptr;                  // device: section of tile c * kGroupChunk
    int32_t n_items = 0;
    std::vector<int32_t> item_secs;   // sections of items [0, k): item_secs[k]
    std::vector<int64_t> item_tiles;  // tiles of items [0, k): item_tiles[k]
};

extern "C" {

int tvlm_lhs_group_create(tvlm_lhs_group** out, int32_t n_items, const tvlm_lhs_item* items, void* stream) {
    TVLM_REQUIRE(out && items, "null argument");
    TVLM_REQUIRE(n_items >= 1, "n_items %d < 1", n_items);
    std::vector<tvlm::LhsSection<double>> secs;
    std::vector<int32_t> isec(1, 0);
    std::vector<int64_t> itile(1, 0);
    int64_t tiles = 0;
    for (int32_t k = 0; k < n_items; k++) {
        const tvlm_lhs_item& it = items[k];
        TVLM_TRY(tvlm::lhs_sections<double>(&it.lhs, it.n_surf, it.colloc, it.normals, it.verts_wing, it.verts_wake,
                                            it.iteration, secs, tiles));
        isec.push_back((int32_t)secs.size());
        itile.push_back(tiles);
    }
    auto* G = new tvlm_lhs_group();
    const size_t bytes = sizeof(tvlm::LhsSection<double>) * secs.size();
    std::vector<int32_t> t0(secs.size() + 1);
    for (size_t k = 0; k < secs.size(); k++) t0[k] = secs[k].tile0;
    t0[secs.size()] = (int32_t)tiles;
    std::vector<int32_t> ch((size_t)((tiles + tvlm::kGroupChunk - 1) / tvlm::kGroupChunk));
    for (size_t c = 0, k = 0; c < ch.size(); c++) {
        const int64_t t = (int64_t)c * tvlm::kGroupChunk;
        while (t0[k + 1] <= t) k++;
        ch[c] = (int32_t)k;
    }
    // uploaded on the caller's stream (not the null stream, which would wait for every other stream)
    hipStream_t st = tvlm::as_stream(stream);
    if (hipMalloc((void**)&G->secs, bytes) != hipSuccess ||
        hipMalloc((void**)&G->tile0, sizeof(int32_t) * t0.size()) != hipSuccess ||
        hipMalloc((void**)&G->chunk, sizeof(int32_t) * (ch.empty() ? 1 : ch.size())) != hipSuccess ||
        hipMemcpyAsync(G->chunk, ch.data(), sizeof(int32_t) * ch.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(G->secs, secs.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(G->tile0, t0.data(), sizeof(int32_t) * t0.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        if (G->secs) (void)hipFree(G->secs);
        if (G->tile0) (void)hipFree(G->tile0);
        if (G->chunk) (void)hipFree(G->chunk);
        delete G;
        tvlm::set_error("group table of %zu sections: device allocation or upload failed", secs.size());
        return TVLM_ERR_HIP;
    }
    G->n_items = n_items;
    G->item_secs = std::move(isec);
    G->item_tiles = std::move(itile);
    *out = G;
    return TVLM_OK;
}

int tvlm_lhs_group_assemble_f64(tvlm_lhs_group* G, int32_t n_items, int32_t seeded, double sigma, void* stream) {
    TVLM_REQUIRE(G, "null group");
    TVLM_REQUIRE(n_items >= 0 && n_items <= G->n_items, "n_items %d of %d", n_items, G->n_items);
    TVLM_REQUIRE(!(seeded && sigma != 0.0), "the seeded form has no core model (sigma must be 0)");
    const int64_t tiles = G->item_tiles[n_items];
    if (tiles == 0) return TVLM_OK;
    hipStream_t st = tvlm::as_stream(stream);
    using tvlm::k_lhs_grouped;
    const dim3 grid((unsigned)tiles);
    if (sigma != 0.0)
        hipLaunchKernelGGL((k_lhs_grouped<double, 32, 1>), grid, dim3(tvlm::kThreads), 0, st, G->secs, G->tile0,
                           G->chunk, (int)tiles, sigma);
    else if (seeded)
        hipLaunchKernelGGL((k_lhs_grouped<double, 32, 3>), grid, dim3(tvlm::kThreads), 0, st, G->secs, G->tile0,
                           G->chunk, (int)tiles, sigma);
    else
        hipLaunchKernelGGL((k_lhs_grouped<double, 32, 0>), grid, dim3(tvlm::kThreads), 0, st, G->secs, G->tile0,
                           G->chunk, (int)tiles, sigma);
    TVLM_LAUNCHED();
    return TVLM_OK;
}

int tvlm_lhs_group_destroy(tvlm_lhs_group* G) {
    if (!G) return TVLM_OK;
    if (G->secs) (void)hipFree(G->secs);
    if (G->tile0) (void)hipFree(G->tile0);
    if (G->chunk) (void)hipFree(G->chunk);
    delete G;
    return TVLM_OK;
}

int tvlm_lhs_assemble_seeded_f64(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc,
                                 const tvlm_view* normals, const tvlm_view* verts_wing, const tvlm_view* verts_wake,
                                 int32_t iteration, void* stream) {
    return tvlm::lhs_assemble_seeded_f64(lhs, n_surf, colloc, normals, verts_wing, verts_wake, iteration, stream);
}

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
int tvlm_rhs_assemble_wake_influence_noalloc_f64(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                                 const tvlm_view* colloc, const tvlm_view* normals,
                                                 const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                                                 double sigma, void* stream) {
    void* ws = nullptr;
    TVLM_TRY(tvlm::scratch_address(&ws));
    return tvlm::wake_influence<double>(rhs, n_surf, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                                        stream, ws, tvlm::kNoAllocScratchBytes);
}
int tvlm_rhs_assemble_wake_influence_lattice_f64(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                                 const tvlm_view* colloc, const tvlm_view* normals,
                                                 const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                                                 void* stream) {
    return tvlm::wake_influence<double>(rhs, n_surf, gamma_wake, colloc, normals, verts_wake, lifting, iteration, 0.0,
                                        stream, nullptr, 0, true);
}
int tvlm_rhs_assemble_wake_influence_f32(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                         const tvlm_view* colloc, const tvlm_view* normals, const tvlm_view* verts_wake,
                                         const uint8_t* lifting, int32_t iteration, float sigma, void* stream) {
    return tvlm::wake_influence<float>(rhs, n_surf, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                                       stream);
}

}  // extern "C"

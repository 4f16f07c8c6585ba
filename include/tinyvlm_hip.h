/*
 * tinyvlm_hip.h -- C ABI of the MI355X (gfx950) HIP backend for tinyvlm.
 *
 * Drop-in boundary: every entry point below is one method of the reference's plugin interface
 * (class vlm::Backend, vlm::LU, vlm::BLAS, vlm::Memory in vlm/include/vlm_backend.hpp and
 * vlm/include/vlm_memory.hpp), flattened to plain C: device pointers, shapes and strides
 * (tvlm_view), an explicit HIP stream, and an int status instead of exit()/exceptions.
 * The citation beside each declaration names the reference interface it replaces.
 *
 * Conventions
 *  - Every tensor is column-major like the reference TensorView (vlm/include/vlm_memory.hpp:306-319):
 *    element (i,j,k) lives at ptr + i*stride[0] + j*stride[1] + k*stride[2]; stride[0] must be 1
 *    for 2-D/3-D operands.  Unused trailing dimensions have shape 1.
 *  - Multi-surface arguments (the reference's MultiTensorView = std::vector<TensorView>) are
 *    host arrays of n_surf tvlm_view, one per surface, 1 <= n_surf <= TVLM_MAX_SURFACES.
 *  - Suffix _f64 computes in double, _f32 in float.  `stream` is a hipStream_t (NULL = default).
 *  - All compute entry points are asynchronous on `stream`.  Scalar results (CL, CM, sums,
 *    norms) are written to DEVICE memory (`*_dev` arguments) so that callers can keep whole
 *    solver loops on the device; the C++ adapter (tinyvlm_amd/vlm) turns them back into the
 *    reference's by-value returns.
 *  - Outputs the reference accumulates into (lhs, rhs) are accumulated here too;
 *    the caller zero-fills them, as the reference drivers do (vlm/src/vlm.cpp:99-100, :437).
 *  - Return value: TVLM_OK or an error code; tvlm_last_error() gives the message.
 */
#ifndef TINYVLM_HIP_H
#define TINYVLM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TVLM_MAX_SURFACES 16

typedef enum tvlm_status {
    TVLM_OK = 0,
    TVLM_ERR_INVALID = 1,     /* bad shape / stride / argument */
    TVLM_ERR_HIP = 2,         /* HIP runtime error */
    TVLM_ERR_SINGULAR = 3,    /* getrf: exactly zero pivot (info > 0) */
    TVLM_ERR_UNSUPPORTED = 4, /* request outside what this build implements */
    TVLM_ERR_NOMEM = 5        /* device allocation failed */
} tvlm_status;

typedef struct tvlm_view {
    void* ptr;         /* address of element (0,0,0) */
    int64_t shape[3];  /* extents; 1 for unused trailing dimensions */
    int64_t stride[3]; /* in elements */
} tvlm_view;

/* ---- library / device ---------------------------------------------------------------------- */
/* BackendCUDA::BackendCUDA device selection and banner (vlm/backends/cuda/src/vlm_backend_cuda.cu:132-142) */
int tvlm_init(int32_t device);
/* error text of the last failing call on this thread (replaces the reference's printf + exit,
 * vlm/backends/cuda/src/vlm_backend_cuda.cu:21-29) */
const char* tvlm_last_error(void);
/* create_backend name registry (vlm/src/vlm_backend.cpp:15-38): "hip" */
const char* tvlm_backend_name(void);
int tvlm_stream_synchronize(void* stream);
/* Streams for callers that run several plans at once (no counterpart in the reference, whose
 * backends use one implicit stream): tvlm_stream_create makes a default-priority stream with a
 * hardware queue of its own (all-CUs mask; unmasked streams share a small pool of queues and two
 * streams on one queue run one after the other).  tvlm_stream_query sets *done = 1 once all
 * work queued on the stream has finished, 0 while some is still running; it does not block.
 * tvlm_stream_destroy waits for the stream and also frees what the library kept for it (scratch
 * buffers, its rocBLAS handle, its LU look-ahead side stream). */
int tvlm_stream_create(void** stream);
int tvlm_stream_destroy(void* stream);
int tvlm_stream_query(void* stream, int32_t* done);

/* ---- Memory (vlm/include/vlm_memory.hpp:22-35) ---------------------------------------------- */
/* location: 0 = Device, 1 = Host (pinned) -- enum class Location { Device, Host } */
int tvlm_malloc(void** ptr, int64_t bytes, int32_t location);
int tvlm_free(void* ptr, int32_t location);
/* Memory::copy: n elements of elem_size bytes, strided in elements */
int tvlm_copy(int32_t dst_loc, void* dst, int64_t dst_stride, int32_t src_loc, const void* src,
              int64_t src_stride, int64_t elem_size, int64_t n, void* stream);
/* Memory::fill (float / double overloads) */
int tvlm_fill_f32(int32_t location, float* ptr, int64_t stride, float value, int64_t n, void* stream);
int tvlm_fill_f64(int32_t location, double* ptr, int64_t stride, double value, int64_t n, void* stream);

/* ---- Backend virtuals ----------------------------------------------------------------------- */
#define TVLM_DECLARE_BACKEND(T, S)                                                                  \
    /* Backend::lhs_assemble (vlm_backend.hpp:31/:44): lhs{N,N} += ring influence; TE columns also \
     * get wake rows -2-k, k < iteration */                                                         \
    int tvlm_lhs_assemble_##S(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc,      \
                              const tvlm_view* normals, const tvlm_view* verts_wing,                \
                              const tvlm_view* verts_wake, int32_t iteration, T sigma,             \
                              void* stream);                                                        \
    /* Backend::rhs_assemble_velocities (vlm_backend.hpp:32/:45): rhs += -(V . n) */                \
    int tvlm_rhs_assemble_velocities_##S(const tvlm_view* rhs, int32_t n_surf,                     \
                                         const tvlm_view* normals, const tvlm_view* velocities,     \
                                         void* stream);                                             \
    /* Backend::rhs_assemble_wake_influence (vlm_backend.hpp:33/:46) */                             \
    int tvlm_rhs_assemble_wake_influence_##S(const tvlm_view* rhs, int32_t n_surf,                 \
                                             const tvlm_view* gamma_wake, const tvlm_view* colloc,  \
                                             const tvlm_view* normals,                              \
                                             const tvlm_view* verts_wake, const uint8_t* lifting,   \
                                             int32_t iteration, T sigma, void* stream);             \
    /* Backend::forces_steady (vlm_backend.hpp:34/:47) */                                           \
    int tvlm_forces_steady_##S(const tvlm_view* verts_wing, const tvlm_view* gamma_delta,          \
                               const tvlm_view* velocities, const tvlm_view* forces, void* stream); \
    /* Backend::forces_unsteady (vlm_backend.hpp:35/:48) */                                         \
    int tvlm_forces_unsteady_##S(const tvlm_view* verts_wing, const tvlm_view* gamma_delta,        \
                                 const tvlm_view* dgamma_dt, const tvlm_view* velocities,           \
                                 const tvlm_view* areas, const tvlm_view* normals,                  \
                                 const tvlm_view* forces, void* stream);                            \
    /* Backend::coeff_cl (vlm_backend.hpp:36/:49); result -> cl_dev[0] */                           \
    int tvlm_coeff_cl_##S(const tvlm_view* forces, const T* lift_axis, const T* freestream, T rho, \
                          T area, T* cl_dev, void* stream);                                         \
    /* Backend::coeff_cm (vlm_backend.hpp:37/:50); result -> cm_dev[0..2] */                        \
    int tvlm_coeff_cm_##S(const tvlm_view* forces, const tvlm_view* verts_wing, const T* ref_pt,   \
                          const T* freestream, T rho, T area, T mac, T* cm_dev, void* stream);      \
    /* Backend::mesh_metrics (vlm_backend.hpp:38/:51) */                                            \
    int tvlm_mesh_metrics_##S(T alpha_rad, int32_t n_surf, const tvlm_view* verts_wing,            \
                              const tvlm_view* colloc, const tvlm_view* normals,                    \
                              const tvlm_view* areas, void* stream);                                \
    /* Backend::mesh_mac (vlm_backend.hpp:39/:52); result -> mac_dev[0] */                          \
    int tvlm_mesh_mac_##S(const tvlm_view* verts_wing, const tvlm_view* areas, T* mac_dev,         \
                          void* stream);                                                            \
    /* Backend::gamma_wake_from_coeffs (vlm_backend.hpp:40/:53) */                                  \
    int tvlm_gamma_wake_from_coeffs_##S(const tvlm_view* gamma_wake,                               \
                                        const tvlm_view* gamma_coeffs, int32_t harmonics, T tn,     \
                                        T omega, T dt, int64_t iteration, void* stream);            \
    /* Backend::sum (TV1 / TV2 overloads, vlm_backend.hpp:41-42/:54-55); ndim 1 or 2 */             \
    int tvlm_sum_##S(const tvlm_view* x, int32_t ndim, T* sum_dev, void* stream);                  \
    /* Backend::wake_shed (vlm_backend.hpp:59/:66, impl vlm/src/vlm_backend.cpp:42-47) */           \
    int tvlm_wake_shed_##S(int32_t n_surf, const tvlm_view* verts_wing,                            \
                           const tvlm_view* verts_wake, int32_t iteration, void* stream);           \
    /* Backend::displace_wing (vlm_backend.hpp:58/:65, impl vlm/src/vlm_backend.cpp:49-69) */       \
    int tvlm_displace_wing_##S(int32_t n_surf, const tvlm_view* transforms,                        \
                               const tvlm_view* verts_wing, const tvlm_view* verts_wing_init,       \
                               void* stream);                                                       \
    /* The wake geometry of a prescribed motion (no reference counterpart: one launch per surface  \
     * for the per-step displace_wing + wake_shed loop of lib/libhbvlm3.cpp:306-319).  For every    \
     * step i < steps, surface m's wake vertex row shape[1]-1-i becomes its wing trailing-edge row   \
     * of verts_wing_init displaced by transforms[:, :, i*n_surf + m] ({4,4,>= steps*n_surf}): the  \
     * values steps alternating displace_wing / wake_shed(i) calls leave there.  The wing vertices   \
     * are not touched; other wake rows are left as they are. */                                    \
    int tvlm_wake_geometry_##S(int32_t n_surf, const tvlm_view* transforms,                        \
                               const tvlm_view* verts_wing_init, const tvlm_view* verts_wake,       \
                               int32_t steps, void* stream);                                        \
    /* LU::factorize (vlm_backend.hpp:105/:112): in-place L\U with row pivots; ipiv_dev{N} is      \
     * 1-based like LAPACK; info_dev[0] = 0 or k (1-based) of the first exactly zero pivot.         \
     * a->shape[2] = B > 1 factors B independent matrices in one pass (matrix b at                  \
     * ptr + b*stride[2], pivots at ipiv_dev + b*N, status info_dev[b]). */                         \
    int tvlm_getrf_##S(const tvlm_view* a, int32_t* ipiv_dev, int32_t* info_dev, void* stream);    \
    /* LU::solve (vlm_backend.hpp:106/:113): x{N,nrhs} <- A^{-1} x ('N'); with a batch, x is       \
     * {N, nrhs, B} and set b is solved with matrix b */                                            \
    int tvlm_getrs_##S(const tvlm_view* a, const int32_t* ipiv_dev, const tvlm_view* x,            \
                       void* stream);                                                               \
    /* LSQ::init/solve (vlm_backend.hpp:121-129; LAPACKE_?gels 'N' in vlm_backend_cpu_lapack.cpp:   \
     * 124-205): min ||A X - B|| for A{M,N}, M >= N, N + NRHS <= 32; A <- Householder QR (?geqrf   \
     * layout), B{M,NRHS} <- Q^T B with X in rows 0..N-1; info_dev[0] = i if R(i,i) == 0 (X not  \
     * computed), else 0.  No workspace (LSQ::init has nothing to size). */                        \
    int tvlm_gels_##S(const tvlm_view* a, const tvlm_view* b, int32_t* info_dev, void* stream);    \
    /* BLAS::gemv (vlm_backend.hpp:84/:91), trans: 0 = No, 1 = Yes */                               \
    int tvlm_gemv_##S(T alpha, const tvlm_view* a, const tvlm_view* x, T beta, const tvlm_view* y, \
                      int32_t trans, void* stream);                                                 \
    /* BLAS::gemm (vlm_backend.hpp:85/:92) */                                                       \
    int tvlm_gemm_##S(T alpha, const tvlm_view* a, const tvlm_view* b, T beta, const tvlm_view* c, \
                      int32_t trans_a, int32_t trans_b, void* stream);                              \
    /* BLAS::axpy 1-D and column-wise 2-D (vlm_backend.hpp:86-87/:93-94); ndim 1 or 2 */            \
    int tvlm_axpy_##S(T alpha, const tvlm_view* x, const tvlm_view* y, int32_t ndim, void* stream);\
    /* BLAS::scal (vlm_backend.hpp:88/:95) */                                                       \
    int tvlm_scal_##S(T alpha, const tvlm_view* x, void* stream);                                  \
    /* BLAS::norm (vlm_backend.hpp:89/:96) -> nrm_dev[0] */                                         \
    int tvlm_nrm2_##S(const tvlm_view* x, T* nrm_dev, void* stream);                              \
    /* Device form of UVLM::run's host velocity loop (vlm/src/vlm.cpp:419-436): for surface m,     \
     * velocities(p) = -(dT_m [x_p; 1]) with x_p = colloc(p); dT holds n_surf column-major 4x4      \
     * time derivatives of the surface transforms (host memory, 16*n_surf values) */                \
    int tvlm_kinematic_velocities_##S(int32_t n_surf, const T* dT, const tvlm_view* colloc,       \
                                      const tvlm_view* velocities, void* stream);

TVLM_DECLARE_BACKEND(double, f64)
TVLM_DECLARE_BACKEND(float, f32)

/* LU factorisation plan (no counterpart in the reference; host only, no device needed): the path
 * tvlm_getrf_{f32,f64} takes for n x n matrices in batches of `batch` (elem_bytes 4 or 8):
 * *path = 0 batched look-ahead (one-group panels), 1 cooperative panels (several work-groups per
 * panel, waiting on each other through bounded spins), 2 one-group panels throughout.  For path 1,
 * *n_steps panels of 64 columns: groups[k] work-groups of panel k (0 = one-group panel) and the
 * exchange counter set ctr_set[k] it uses (-1 for one-group panels); *coop_ws_bytes the exchange
 * buffers getrf allocates.  groups/ctr_set may be null; otherwise they hold max_steps entries. */
int tvlm_getrf_plan(int64_t n, int32_t batch, int32_t elem_bytes, int32_t* path, int32_t* n_steps, int32_t* groups,
                    int32_t* ctr_set, int32_t max_steps, int64_t* coop_ws_bytes);
/* C -= A B for `batch` matrices (matrix z at c/a/b + z * stride elements; elem_bytes 4 or 8,
 * column-major, C m x n, A m x k, B k x n): the trailing update inside tvlm_getrf_{f32,f64}, on the
 * matrix cores.  No counterpart in the reference, where the update runs inside LAPACKE_?getrf
 * (vlm/backends/cpu/src/vlm_backend_cpu_lapack.cpp:40-50) and cuSOLVER ?getrf
 * (vlm/backends/cuda/src/vlm_backend_cuda_solvers.cu:71-102); exported so that tests can check
 * it against a vendor GEMM at the shapes getrf uses. */
int tvlm_lu_update(int32_t elem_bytes, void* c, int64_t ldc, const void* a, int64_t lda, const void* b, int64_t ldb,
                   int64_t m, int64_t n, int64_t k, int32_t batch, int64_t stride, void* stream);
/* Bound of every wait of the cooperative LU panels, in spins (default 4194304; 0 restores it).  No
 * counterpart in the reference: a diagnostic knob with which tests make the bounded waits give up,
 * so that the recovery of a stalled factorisation (info = -999, see tvlm_vlm_refinement) runs.  A
 * wait that gives up never hangs or faults; it leaves that matrix's factors invalid and info -999. */
int tvlm_lu_set_coop_spin_limit(uint32_t limit);

/* Library-wide work-group targets of the wake-induction launches (no counterpart in the reference;
 * tuning knobs that tinyvlm_amd.autotune sets from its cached measurements): wake_groups for the
 * ring-by-ring form (Backend::rhs_assemble_wake_influence and the plain UVLM loop, default 2048),
 * lattice_groups for UVLM plans created afterwards (fast mode, default 1024; per plan:
 * tvlm_uvlm_set_lattice_target).  0 restores a default; range [64, 65536]. */
int tvlm_set_tile_targets(int32_t wake_groups, int32_t lattice_groups);
int tvlm_get_tile_targets(int32_t* wake_groups, int32_t* lattice_groups);

/* Grouped assembly: many FP64 influence matrices of any sizes -- different meshes, different
 * requests -- in ONE launch over a work list of (matrix, section, tile).  Each item is one
 * Backend::lhs_assemble call (vlm_backend.hpp:44) with its own arguments; the result of every matrix
 * is bitwise that of tvlm_lhs_assemble_f64 (seeded = 0) or tvlm_lhs_assemble_seeded_f64 (seeded = 1)
 * on the same item.  Like those, it accumulates into lhs (zero it first).  No counterpart in the
 * reference, which launches per surface pair (vlm_backend_cuda.cu:193-252).  The group keeps the
 * items' device pointers (views and the surfaces' data must stay where they are); create uploads
 * the work list once on `stream` and synchronizes it; assemble launches the first n_items items. */
typedef struct tvlm_lhs_item {
    tvlm_view lhs;               /* {N, N} column-major (stride[0] = 1), N = the item's panel count */
    int32_t n_surf;              /* surfaces of the item; the views below have n_surf entries */
    const tvlm_view* colloc;     /* host arrays of device views, as tvlm_lhs_assemble_f64 takes them */
    const tvlm_view* normals;
    const tvlm_view* verts_wing;
    const tvlm_view* verts_wake;
    int32_t iteration;
} tvlm_lhs_item;
typedef struct tvlm_lhs_group tvlm_lhs_group;
int tvlm_lhs_group_create(tvlm_lhs_group** group, int32_t n_items, const tvlm_lhs_item* items, void* stream);
int tvlm_lhs_group_assemble_f64(tvlm_lhs_group* group, int32_t n_items, int32_t seeded, double sigma,
                                void* stream);
int tvlm_lhs_group_destroy(tvlm_lhs_group* group);

/* Backend::lhs_assemble (vlm_backend.hpp:44) in the FP64 seeded form the served steady path uses
 * (FP64 arithmetic; 1/sqrt and 1/x from the FP64 hardware estimates plus one Newton step each; no
 * core model, sigma = 0): same arguments and result as tvlm_lhs_assemble_f64 to about 1e-15
 * relative.  No counterpart in the reference, which assembles in FP32 or exact FP64. */
int tvlm_lhs_assemble_seeded_f64(const tvlm_view* lhs, int32_t n_surf, const tvlm_view* colloc,
                                 const tvlm_view* normals, const tvlm_view* verts_wing, const tvlm_view* verts_wake,
                                 int32_t iteration, void* stream);
/* Allocation-free form of tvlm_rhs_assemble_wake_influence_f64 (same arguments, same kernels, same
 * result bit for bit).  No counterpart in the reference: it is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It calls no hipMalloc/hipFree and no allocating library routine: the tile partials go
 * to a fixed 32 MiB device array of the library (one per device, held until the process ends).
 * A call whose partials would exceed it (tiles x receivers x 8 bytes > 32 MiB) is rejected with
 * TVLM_ERR_INVALID before any work is queued.  Every call on a device shares that device's array:
 * calls that overlap on one device (different streams, or host threads) race and must be ordered
 * by the caller.  The array's address is looked up by tvlm_init, or else by the first call on a
 * device (outside any capture). */
int tvlm_rhs_assemble_wake_influence_noalloc_f64(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                                 const tvlm_view* colloc, const tvlm_view* normals,
                                                 const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                                                 double sigma, void* stream);
/* Backend::rhs_assemble_wake_influence (vlm_backend.hpp:46) evaluated on the wake's vortex lattice
 * in the FP64 seeded form (1/sqrt and 1/x from the FP64 hardware estimates plus one Newton step;
 * every lattice segment once, with the difference of the circulations of the rings beside it): the
 * same arguments (no core model: sigma = 0) and the same rhs as tvlm_rhs_assemble_wake_influence_f64
 * to rounding (within 2e-12 of the largest entry: the estimates' few ulp are amplified where a
 * receiver lies near the extension of a segment), in fewer operations.  No counterpart in the
 * reference; the HB solver takes it with wake_form "lattice". */
int tvlm_rhs_assemble_wake_influence_lattice_f64(const tvlm_view* rhs, int32_t n_surf, const tvlm_view* gamma_wake,
                                                 const tvlm_view* colloc, const tvlm_view* normals,
                                                 const tvlm_view* verts_wake, const uint8_t* lifting, int32_t iteration,
                                                 void* stream);

/* ---- device-resident UVLM time loop ---------------------------------------------------------
 * The caller of the hot path, UVLM::run (vlm/src/vlm.cpp:358-509), kept on the device: the host
 * evaluates the prescribed kinematics once for all steps (tvlm_uvlm_set_motion) and the whole
 * loop -- wake shed, kinematic velocities, rhs, wake induction, solve, gamma post-processing,
 * unsteady forces, CL/CM, displacement, panel metrics -- runs without host round trips.
 * mode 0: plain FP64 (exact FP64 Biot-Savart, LU forward/backward substitution every step).
 * mode 1: wake induction in the FP64 seeded form (FP64 arithmetic, 1/sqrt and 1/x from the FP64
 *         hardware estimates plus one Newton step), gamma = A^{-1} rhs with the FP64 inverse formed
 *         once from the LU; use_graph replays the loop as a hipGraph (needs a non-default stream;
 *         mode 1 only: plain mode runs call by call, use_graph != 0 is TVLM_ERR_INVALID).
 * All surfaces are lifting; LHS assembly uses iteration 0 as the reference does (vlm.cpp:402). */
typedef struct tvlm_uvlm_plan tvlm_uvlm_plan;
/* verts_init[m]: host {ns+1, nc+1, 4} column-major double (w = 1); t_steps as UVLM::run's
 * (the loop runs t_steps-1 steps); dt the fixed time step (vlm.cpp:380-386) */
int tvlm_uvlm_create(tvlm_uvlm_plan** plan, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                     const double* const* verts_init, int64_t t_steps, double dt, int32_t mode,
                     void* stream);
/* T, dT: host [t_steps][n_surf][16] column-major value / time derivative of each surface's
 * transform at t_i = i*dt; freestream, ref_pt: host [t_steps][3] */
int tvlm_uvlm_set_motion(tvlm_uvlm_plan* plan, const double* T, const double* dT,
                         const double* freestream, const double* ref_pt);
/* prescribed motion evaluated natively for all steps: freestream translation(-u t), heave
 * h_amp sin(omega t + h_phase) after it, pitch a_amp sin(omega t + a_phase) about (x_pivot,0,0)
 * and the y axis after that (tests/uvlm_theodorsen.cpp:184-208 style chain); CM about
 * T(t) (0.5,0,0) (vlm/src/vlm.cpp:464-466) */
int tvlm_uvlm_set_pitch_heave(tvlm_uvlm_plan* plan, double u, double h_amp, double h_phase,
                              double a_amp, double a_phase, double omega, double x_pivot);
/* host-only evaluation of the same chain into caller arrays laid out as tvlm_uvlm_set_motion
 * takes them (no device work) */
int tvlm_motion_pitch_heave(double u, double h_amp, double h_phase, double a_amp, double a_phase,
                            double omega, double x_pivot, int32_t n_surf, int64_t t_steps, double dt,
                            double* T, double* dT, double* freestream, double* ref_pt);
/* one whole simulation (setup + all steps), asynchronous on the plan's stream */
int tvlm_uvlm_run(tvlm_uvlm_plan* plan, int32_t use_graph);
/* synchronises; CL[t_steps-1], CM[t_steps-1][3] (step 0 = 0 as the reference skips it),
 * last gamma[N], getrf info; null pointers are skipped */
int tvlm_uvlm_results(tvlm_uvlm_plan* plan, double* cl, double* cm, double* gamma, int32_t* info);
int tvlm_uvlm_destroy(tvlm_uvlm_plan* plan);
/* measurement: one simulation without capture, then every step's wake-induction launch replayed
 * back to back between two HIP events on the plan's stream; returns their total ms, the launch
 * count and the (receiver, wake ring) pairs they evaluate.  Leaves the results of that simulation
 * in the plan. */
int tvlm_uvlm_time_wake(tvlm_uvlm_plan* plan, double* total_ms, int32_t* launches, double* pairs);
/* Far-wake aggregation of the fast mode (mode 1) wake induction; no counterpart in the reference,
 * whose UVLM evaluates every wake ring (vlm/src/vlm.cpp:439).  n_levels (0: exact wake, the
 * default) triples {age A, row group G, span group S}, A increasing, each G a multiple of the one
 * before.  Counted from the oldest ring row, level l's blocks are G consecutive rows; the
 * floor((it - A) / G) oldest blocks are older than A steps; the coarsest level takes its blocks
 * first, each finer level those of its own count that remain, the newest rows stay exact.  Each
 * (block, group of S spans) becomes one vortex ring through the wake vertices at its corners with
 * the mean circulation of the rings it replaces.  An approximation: accept it only where |dCL|
 * against the exact wake stays within the caller's bar.  Drops a captured graph. */
int tvlm_uvlm_set_far_wake(tvlm_uvlm_plan* plan, int32_t n_levels, const int32_t* levels);
/* Work-group target of the fast mode's wake-lattice launches (default 1024): a launch above it takes
 * 2 or 4 wake rows per tile, a launch below it splits its spans (at least 2 per wave) to reach it.
 * No counterpart in the reference (a tuning knob; tinyvlm_amd.autotune picks it per workload).
 * Drops a captured graph. */
int tvlm_uvlm_set_lattice_target(tvlm_uvlm_plan* plan, int32_t groups);
/* Several simulations of one mesh in one plan (mode 1 only; sims in [1, 65535]), each with its own
 * prescribed motion: every launch of the time loop covers all of them, so a batch of short
 * simulations fills the device that one alone leaves idle.  Each simulation's results equal those
 * of a one-simulation plan with its motion.  No counterpart in the reference (one UVLM per
 * Backend, vlm/src/vlm.cpp:358); for callers that run many motions of one mesh (a sweep, a
 * served queue).  tvlm_uvlm_create is this with sims = 1; tvlm_uvlm_set_motion and
 * tvlm_uvlm_set_pitch_heave set every simulation's motion, tvlm_uvlm_results reads simulation 0.
 * tvlm_uvlm_time_wake counts the pairs of all simulations. */
int tvlm_uvlm_create_batched(tvlm_uvlm_plan** plan, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                             const double* const* verts_init, int64_t t_steps, double dt, int32_t mode,
                             int32_t sims, void* stream);
/* motion of simulation sim, laid out as tvlm_uvlm_set_motion takes it */
int tvlm_uvlm_set_motion_sim(tvlm_uvlm_plan* plan, int32_t sim, const double* T, const double* dT,
                             const double* freestream, const double* ref_pt);
/* tvlm_uvlm_set_pitch_heave for simulation sim (-1: every simulation) */
int tvlm_uvlm_set_pitch_heave_sim(tvlm_uvlm_plan* plan, int32_t sim, double u, double h_amp,
                                  double h_phase, double a_amp, double a_phase, double omega,
                                  double x_pivot);
/* tvlm_uvlm_results for simulation sim */
int tvlm_uvlm_results_sim(tvlm_uvlm_plan* plan, int32_t sim, double* cl, double* cm, double* gamma,
                          int32_t* info);

/* ---- coupled pitch/heave flutter sweep ------------------------------------------------------
 * The coupled aeroelastic driver UVLM_2DOF (tests/uvlm_2dof.cpp:428-865, COUPLED path, with the
 * Newmark-beta integrator of vlm/src/vlm_integrator.cpp:26-68) for n_airspeeds reduced velocities of
 * one lifting wing at once, resident on the device.  Per airspeed and time step: the reference's
 * fixed-point iterations on the structural increment (Newmark step, pose, panel metrics, wake shed,
 * kinematic velocities, rhs, wake induction, solve, circulation post-processing, unsteady forces,
 * CL/CM about the elastic axis, loads) until |du_k - du| <= EPS_f or 51 iterations.  The influence
 * matrix is assembled and factored once at the initial pose (rigid motion) and inverted once; the
 * wake is a ring buffer of the newest wake_rows ring rows (older rings are dropped, as if their
 * circulation were zero).  FP64 throughout; wake induction in the FP64 seeded form (hardware
 * 1/sqrt and 1/x estimates refined by one Newton step).  No counterpart in the reference beyond the
 * single-airspeed driver it batches. */
typedef struct tvlm_flutter_params {
    double b;           /* half chord (the mesh is divided by it: nondimensionalize_verts) */
    double rho, a_h, omega, zeta_a, zeta_h, x_a, mu, r_a; /* UVLM_2DOF_Vars (tests/uvlm_2dof.cpp:23-39) */
    double beta[4];     /* torsional spring f(alpha) = beta0 + beta1 a + beta2 a^2 + beta3 a^3 (:285) */
    double init_pitch;  /* initial pitch, radians (the reference: to_radians(3)) */
} tvlm_flutter_params;
typedef struct tvlm_flutter_plan tvlm_flutter_plan;
/* verts_init: host {ns+1, nc+1, 4} column-major double (one surface, as Simulation reads it: FP32
 * parse + quarter chord); U_a: host [n_airspeeds] reduced velocities; the time step is the
 * reference's (last chordwise panel of the dimensionless mesh).  Sets up everything (state at t = 0,
 * LU and inverse of the influence matrix) and synchronises. */
int tvlm_flutter_create(tvlm_flutter_plan** plan, int64_t ns, int64_t nc, const double* verts_init,
                        const tvlm_flutter_params* params, int32_t n_airspeeds, const double* U_a,
                        int32_t wake_rows, int64_t max_steps, void* stream);
/* advances every airspeed by `steps` coupled time steps (synchronous: the host reads the per-airspeed
 * convergence flags after each iteration) */
int tvlm_flutter_run(tvlm_flutter_plan* plan, int64_t steps);
/* steps done so far, the time step, states [steps_done + 1][n_airspeeds][8] (h, alpha, h', alpha',
 * h'', alpha'', F_h, F_alpha) and iterations [steps_done][n_airspeeds]; null pointers are skipped */
int tvlm_flutter_results(tvlm_flutter_plan* plan, int64_t* steps_done, double* dt, double* states, int32_t* iters);
/* measurement: the wake-induction launch of the next step (every airspeed), replayed `launches`
 * times between two HIP events on the plan's stream; total ms and (receiver, wake ring) pairs per
 * launch */
int tvlm_flutter_time_wake(tvlm_flutter_plan* plan, int32_t launches, double* total_ms, double* pairs);
int tvlm_flutter_destroy(tvlm_flutter_plan* plan);

/* ---- plot3d meshes (MeshIO, vlm/src/vlm_mesh.cpp:34-124); host only, no device needed ----------
 * dims: ns spanwise and nc chordwise panels.  read: verts {ns+1, nc+1, 4} column-major double, w = 1,
 * parsed in FP32 (parse_f64 = 0: the VLM/UVLM drivers' buffers) or FP64 (the HB driver's) and, if
 * quarter_chord != 0, quarter-chord shifted in that precision */
int tvlm_plot3d_dims(const char* path, int64_t* ns, int64_t* nc);
int tvlm_plot3d_read(const char* path, int32_t quarter_chord, int32_t parse_f64, double* verts);

/* ---- steady solves of a batch of requests ---------------------------------------------------
 * The caller of the hot path VLM::run (vlm/src/vlm.cpp:97-151) for `batch` angles of attack on one
 * mesh, every step on the device.  mode 0: plain FP64, requests one after another (exact FP64
 * assembly, FP64 LU + substitution).  mode 1: served -- the FP64 influence matrices of all requests
 * in the seeded form, one batched FP32 LU, FP64 iterative refinement against the FP64 matrix.
 * CM is taken about cm_ref moved with the wing (the reference returns 0).  Plans on different
 * streams may run concurrently (the library's scratch buffers are per stream). */
typedef struct tvlm_vlm_plan tvlm_vlm_plan;
/* verts_init[m]: host {ns+1, nc+1, 4} column-major double (w = 1), e.g. a plot3d mesh after the
 * quarter-chord shift */
int tvlm_vlm_create(tvlm_vlm_plan** plan, int32_t n_surf, const int64_t* ns, const int64_t* nc,
                    const double* const* verts_init, int32_t batch, int32_t mode, void* stream);
/* alpha, beta: host [batch] (radians; beta may be null); cm_ref: host [3] or null = (0.25, 0, 0);
 * asynchronous on the plan's stream */
int tvlm_vlm_solve(tvlm_vlm_plan* plan, const double* alpha, const double* beta, double u_inf,
                   double rho, const double* cm_ref);
/* the first n_active (1..batch) requests only: alpha, beta host [n_active]; a serving loop fills a
 * plan with however many requests of its mesh are waiting */
int tvlm_vlm_solve_n(tvlm_vlm_plan* plan, int32_t n_active, const double* alpha, const double* beta, double u_inf,
                     double rho, const double* cm_ref);
/* waits for the plan's last solve; cl[n], cm[n][3], gamma[n][N], info[n] (getrf status per request)
 * for the n requests of the last solve; null pointers are skipped */
int tvlm_vlm_results(tvlm_vlm_plan* plan, double* cl, double* cm, double* gamma, int32_t* info);
/* What the served refinement (mode 1) took for each of the n requests of the last solve; null
 * pointers are skipped.  The returned x of every request is tested (||b - A x||_inf <= sqrt(N) eps
 * ||x||_inf ||A||_inf against the FP64 matrix, LAPACK dsgesv's stopping test) before results are
 * returned; status bits: 0 = passed after the fixed 4 corrections, 1 = needed more corrections (at
 * most 8 in all), 2 = failed them and was solved by the FP64 LU of its FP64 matrix (the plain
 * path's solve), 4 = a cooperative panel of its LU gave up waiting (info -999) and the matrix was
 * re-factored with one-group panels, 8 = unresolved (the FP64 matrix is singular: info > 0).
 * corrections[b]: FP32 corrections applied; backward_error[b]: ||b - A x||_inf / (||x||_inf
 * ||A||_inf) of the returned x (NaN after the FP64 fallback).  Mode 0: all zero.  The test, and the
 * extra work of requests that fail it, run inside tvlm_vlm_results (or here, whichever comes first). */
int tvlm_vlm_refinement(tvlm_vlm_plan* plan, int32_t* status, int32_t* corrections, double* backward_error);
/* moves the plan to another stream: its later solves queue there after its own latest work (a
 * device-side wait; the host does not block).  A serving loop keeps plans per mesh and runs each
 * batch on whichever stream is free.  Not while a solve of the plan is being queued elsewhere. */
int tvlm_vlm_set_stream(tvlm_vlm_plan* plan, void* stream);
/* *done = 1 once the plan's latest solve (or creation) has finished on the device, else 0; does
 * not block */
int tvlm_vlm_query(const tvlm_vlm_plan* plan, int32_t* done);
/* device bytes the plan holds (its admission cost in a serving loop) */
int tvlm_vlm_footprint(const tvlm_vlm_plan* plan, int64_t* device_bytes);
int tvlm_vlm_destroy(tvlm_vlm_plan* plan);

#ifdef __cplusplus
}
#endif
#endif /* TINYVLM_HIP_H */

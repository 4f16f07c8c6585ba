/*
 * vlm_oracle.h -- CPU restatement of the tinyvlm reference backend (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity checker for the HIP backend.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it.  Nothing on the product path links or calls it.
 *
 * It restates, in plain C with OpenMP loops, the arithmetic of the reference CPU backend
 * (ISPC kernels + Taskflow scheduling + LAPACK) and of the drivers that call it.  Every function
 * cites the reference file:line it follows.  Two precisions are provided:
 *   orc_*_d  : restatement of the FP64 backend (vlm/backends/cpu/src/vlm_backend_cpud.cpp and the
 *              d* ISPC kernels), including the float-typed constants the reference uses inside
 *              FP64 code (4.0f*PI_f, RCUT=1e-10f for all three cutoffs, EPS_f).
 *   orc_*_f  : restatement of the FP32 backend (vlm/backends/cpu/src/vlm_backend_cpu.cpp and the
 *              FP32 ISPC kernels: RCUT=1e-10f on |r1 x r2|^2, RCUT2=1e-5f on |r1|,|r2|, 5000-row
 *              wake truncation).
 *
 * Data layout (identical to the reference's column-major TensorView, vlm/include/vlm_memory.hpp:306-319):
 *   verts      {ns+1, nc+1, 4}   element (i,j,k) at  i + j*(ns+1) + k*(ns+1)*(nc+1)
 *   verts_wake {ns+1, nwv, 4}    element (i,j,k) at  i + j*(ns+1) + k*(ns+1)*nwv
 *   colloc, normals, velocities, forces {ns, nc, 3}   (i,j,k) at i + j*ns + k*ns*nc
 *   areas, gamma {ns, nc};  gamma_wake {ns, nwv-1}
 *   lhs {N, N} with leading dimension ld;  panel index p = i + j*ns, surfaces concatenated.
 *
 * Parity pins (each function is checked by tests/test_oracle_pins.py against closed forms,
 * brute force in long double, numpy/OpenBLAS, or the reference's analytic known-answer tests).
 */
#ifndef VLM_ORACLE_H
#define VLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DECLARE(REAL, S)                                                                          \
    void orc_biosavart_##S(const REAL* c, const REAL* v1, const REAL* v2, REAL sigma, REAL* out);     \
    void orc_lhs_assemble_##S(REAL* lhs, int64_t ld, int32_t nsurf, const int64_t* ns,                \
                              const int64_t* nc, const int64_t* nwv, REAL* const* colloc,             \
                              REAL* const* normals, REAL* const* verts_wing,                          \
                              REAL* const* verts_wake, int32_t iteration, REAL sigma);                \
    void orc_rhs_assemble_velocities_##S(REAL* rhs, int32_t nsurf, const int64_t* ns,                 \
                                         const int64_t* nc, REAL* const* normals,                     \
                                         REAL* const* velocities);                                    \
    void orc_rhs_assemble_wake_influence_##S(REAL* rhs, int32_t nsurf, const int64_t* ns,             \
                                             const int64_t* nc, const int64_t* nwv,                   \
                                             REAL* const* gamma_wake, REAL* const* colloc,            \
                                             REAL* const* normals, REAL* const* verts_wake,           \
                                             const int32_t* lifting, int32_t iteration, REAL sigma);  \
    void orc_rhs_assemble_wake_influence_far_##S(REAL* rhs, int32_t nsurf, const int64_t* ns,         \
                                                 const int64_t* nc, const int64_t* nwv,               \
                                                 REAL* const* gamma_wake, REAL* const* colloc,        \
                                                 REAL* const* normals, REAL* const* verts_wake,       \
                                                 const int32_t* lifting, int32_t iteration,           \
                                                 REAL sigma, const int32_t* far);                     \
    void orc_mesh_metrics_##S(REAL alpha_rad, int32_t nsurf, const int64_t* ns, const int64_t* nc,    \
                              REAL* const* verts_wing, REAL* const* colloc, REAL* const* normals,     \
                              REAL* const* areas);                                                    \
    void orc_forces_steady_##S(int64_t ns, int64_t nc, const REAL* verts_wing,                        \
                               const REAL* gamma_delta, const REAL* velocities, REAL* forces);        \
    void orc_forces_unsteady_##S(int64_t ns, int64_t nc, const REAL* verts_wing,                      \
                                 const REAL* gamma_delta, const REAL* dgamma_dt,                      \
                                 const REAL* velocities, const REAL* areas, const REAL* normals,      \
                                 REAL* forces);                                                       \
    REAL orc_coeff_cl_##S(int64_t ns, int64_t nc, const REAL* forces, const REAL* lift_axis,          \
                          const REAL* freestream, REAL rho, REAL area);                               \
    void orc_coeff_cm_##S(int64_t ns, int64_t nc, const REAL* forces, const REAL* verts_wing,         \
                          const REAL* ref_pt, const REAL* freestream, REAL rho, REAL area, REAL mac,  \
                          REAL* out);                                                                 \
    REAL orc_sum_##S(int64_t n, const REAL* x);                                                       \
    REAL orc_mesh_mac_##S(int64_t ns, int64_t nc, const REAL* verts_wing, const REAL* areas);         \
    REAL orc_coeff_cl_multibody_##S(int32_t nsurf, const int64_t* ns, const int64_t* nc,              \
                                    REAL* const* forces, REAL* const* areas,                          \
                                    const REAL* freestream, REAL rho);                                \
    void orc_coeff_cm_multibody_##S(int32_t nsurf, const int64_t* ns, const int64_t* nc,              \
                                    REAL* const* forces, REAL* const* verts_wing,                     \
                                    REAL* const* areas, const REAL* ref_pt,                           \
                                    const REAL* freestream, REAL rho, REAL* out);                     \
    void orc_wake_shed_##S(int64_t ns, int64_t nc, int64_t nwv, const REAL* verts_wing,               \
                           REAL* verts_wake, int32_t iteration);                                      \
    void orc_displace_wing_##S(int64_t nverts, const REAL* transform, const REAL* verts_init,          \
                               REAL* verts);                                                          \
    void orc_gamma_wake_from_coeffs_##S(int64_t ns, int64_t nw, REAL* gamma_wake,                     \
                                        const REAL* gamma_coeffs, int32_t harmonics, REAL tn,         \
                                        REAL omega, REAL dt, int64_t iteration);                      \
    int orc_getrf_##S(int64_t n, REAL* a, int64_t lda, int32_t* ipiv);                                \
    void orc_getrs_##S(int64_t n, int64_t nrhs, const REAL* a, int64_t lda, const int32_t* ipiv,      \
                       REAL* b, int64_t ldb);                                                         \
    int orc_gels_##S(int64_t m, int64_t n, int64_t nrhs, REAL* a, int64_t lda, REAL* b, int64_t ldb); \
    void orc_quarterchord_##S(int64_t ns, int64_t nc, REAL* verts);                                   \
    int orc_vlm_run_##S(int32_t nsurf, const int64_t* ns, const int64_t* nc,                          \
                        REAL* const* verts_init, REAL alpha, REAL beta, REAL u_inf, REAL rho,         \
                        const REAL* cm_ref_pt, REAL* gamma_out, REAL* coeffs_out);                    \
    int orc_uvlm_run_##S(int32_t nsurf, const int64_t* ns, const int64_t* nc,                         \
                         REAL* const* verts_init, const REAL* motion, int64_t t_steps_override,       \
                         REAL t_final, REAL* cl_out, REAL* cm_out, REAL* gamma_out,                   \
                         int64_t* t_steps_out);                                                       \
    int orc_uvlm_run_far_##S(int32_t nsurf, const int64_t* ns, const int64_t* nc,                     \
                             REAL* const* verts_init, const REAL* motion, int64_t t_steps_override,   \
                             REAL t_final, REAL* cl_out, REAL* cm_out, REAL* gamma_out,               \
                             int64_t* t_steps_out, const int32_t* far);

ORC_DECLARE(double, d)
ORC_DECLARE(float, f)

/* Dual-number kinematics of the prescribed pitch/heave motion used by the unsteady drivers:
 *   T(t) = Tfs(t) * Theave(t) * Tpitch(t)
 *   Tfs    = translation(-u*t, 0, 0)                                  (tests/uvlm_theodorsen.cpp:198-200)
 *   Theave = translation(0, 0, h_amp*sin(omega*t + h_phase))
 *   Tpitch = rotation(point (x_pivot,0,0), axis (0,1,0), a_amp*sin(omega*t + a_phase))
 *                                                                     (vlm/include/vlm_kinematics.hpp:184-202)
 * motion[] = {u, h_amp, h_phase, a_amp, a_phase, omega, x_pivot}.
 * Writes the 4x4 column-major value T and its time derivative dT. */
void orc_motion_transform_d(const double* motion, double t, double* T, double* dT);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

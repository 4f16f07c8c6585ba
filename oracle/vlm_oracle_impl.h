/*
 * vlm_oracle_impl.h -- precision-generic body of the CPU restatement (TEST INFRASTRUCTURE ONLY).
 * Included twice by vlm_oracle.c: once with REAL=double (FP64 backend, vlm_backend_cpud.cpp and
 * the d* ISPC kernels) and once with REAL=float (FP32 backend, vlm_backend_cpu.cpp and the FP32
 * ISPC kernels).  Expects: REAL, FN(name), IS_F64, SQRT, SIN, COS.
 */

/* ---- constants ------------------------------------------------------------------------------ */
/* vlm/backends/cpu/src/vlm_backend_cpu_kernels.ispc:9-12 : RCUT 1e-10f, RCUT2 1e-5f, PI_f 3.141593f.
 * Both precisions use the float product 4.0f*PI_f (kernels.ispc:87 and :400).                     */
#define ORC_FOURPI ((REAL)(4.0f * 3.141593f))
#if IS_F64
/* dkernel_biosavart uses RCUT for all three tests (kernels.ispc:393) */
#define ORC_CUT_CROSS ((REAL)1e-10f)
#define ORC_CUT_NORM ((REAL)1e-10f)
#else
/* kernel_biosavart: RCUT on |r1 x r2|^2, RCUT2 on the norms (kernels.ispc:80) */
#define ORC_CUT_CROSS ((REAL)1e-10f)
#define ORC_CUT_NORM ((REAL)1e-5f)
#endif
/* EPS_f (vlm/include/vlm_types.hpp:24), used by mesh_metrics in both precisions */
#define ORC_EPS_F ((REAL)1.19209e-07f)

/* ---- small vector helpers (same operation order as the reference macros, kernels.ispc:26-35) -- */
static inline REAL FN(dot3)(const REAL* a, const REAL* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void FN(cross3)(const REAL* a, const REAL* b, REAL* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* Katz-Plotkin eq. 10.115 straight segment, kernels.ispc:69-89 (FP32) / :382-402 (FP64). */
void FN(orc_biosavart)(const REAL* c, const REAL* v1, const REAL* v2, REAL sigma, REAL* out) {
    REAL r0[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
    REAL r1[3] = {c[0] - v1[0], c[1] - v1[1], c[2] - v1[2]};
    REAL r2[3] = {c[0] - v2[0], c[1] - v2[1], c[2] - v2[2]};
    REAL cr[3];
    FN(cross3)(r1, r2, cr);
    const REAL r1n = SQRT(FN(dot3)(r1, r1));
    const REAL r2n = SQRT(FN(dot3)(r2, r2));
    const REAL sq = FN(dot3)(cr, cr);
    if ((sq < ORC_CUT_CROSS) || (r1n < ORC_CUT_NORM) || (r2n < ORC_CUT_NORM)) {
        out[0] = out[1] = out[2] = (REAL)0;
        return;
    }
    const REAL smoother = sigma * sigma * FN(dot3)(r0, r0);
    const REAL coeff = (FN(dot3)(r0, r1) * r2n - FN(dot3)(r0, r2) * r1n) /
                       (ORC_FOURPI * SQRT(sq * sq + smoother * smoother) * r1n * r2n);
    out[0] = cr[0] * coeff;
    out[1] = cr[1] * coeff;
    out[2] = cr[2] * coeff;
}

/* Segment plus its image in the y=0 symmetry plane, kernels.ispc:91-102 / :404-415. */
static inline void FN(symmetry)(REAL* inf, const REAL* c, const REAL* a, const REAL* b, REAL sigma) {
    REAL s[3];
    FN(orc_biosavart)(c, a, b, sigma, s);
    inf[0] += s[0];
    inf[1] += s[1];
    inf[2] += s[2];
    const REAL cm[3] = {c[0], -c[1], c[2]};
    FN(orc_biosavart)(cm, a, b, sigma, s);
    inf[0] += s[0];
    inf[1] -= s[1];
    inf[2] += s[2];
}

/* Induced velocity of one unit vortex ring (v0,v1,v2,v3) at c, segment order 01,12,23,30. */
static inline void FN(ring)(REAL* inf, const REAL* c, const REAL* v0, const REAL* v1, const REAL* v2,
                            const REAL* v3, REAL sigma) {
    inf[0] = inf[1] = inf[2] = (REAL)0;
    FN(symmetry)(inf, c, v0, v1, sigma);
    FN(symmetry)(inf, c, v1, v2, sigma);
    FN(symmetry)(inf, c, v2, v3, sigma);
    FN(symmetry)(inf, c, v3, v0, sigma);
}

/* kernel_influence / dkernel_influence (kernels.ispc:106-135 / :419-448): one ring whose vertex 0
 * is at v (vertex1 = v+1, vertex3 = v+v_rld, vertex2 = v+v_rld+1; planes v_ld apart) acting on m
 * receivers; lhs_col[i] += (ring velocity at colloc i) . normal i. */
static void FN(influence)(int64_t m, REAL* lhs_col, const REAL* colloc, int64_t cld, const REAL* v,
                          int64_t v_ld, int64_t v_rld, const REAL* normals, int64_t nld, REAL sigma) {
    const REAL v0[3] = {v[0], v[v_ld], v[2 * v_ld]};
    const REAL v1[3] = {v[1], v[v_ld + 1], v[2 * v_ld + 1]};
    const REAL v2[3] = {v[v_rld + 1], v[v_ld + v_rld + 1], v[2 * v_ld + v_rld + 1]};
    const REAL v3[3] = {v[v_rld], v[v_ld + v_rld], v[2 * v_ld + v_rld]};
    for (int64_t i = 0; i < m; i++) {
        const REAL c[3] = {colloc[i], colloc[cld + i], colloc[2 * cld + i]};
        const REAL n[3] = {normals[i], normals[nld + i], normals[2 * nld + i]};
        REAL inf[3];
        FN(ring)(inf, c, v0, v1, v2, v3, sigma);
        lhs_col[i] += FN(dot3)(inf, n);
    }
}

/* Backend::lhs_assemble (vlm_backend_cpu.cpp:84-148, vlm_backend_cpud.cpp:27-91).
 * Surface blocks (m_i receivers, m_j rings) at lhs offset (offset_i, offset_j).  Every ring of
 * surface m_j is applied to every receiver of m_i (wing pass + last_row pass); then the trailing
 * edge columns additionally receive the wake rings of wake rows -2-k for k < iteration
 * (the Taskflow condition loop, vlm_backend_cpu.cpp:121-139).  Accumulates into lhs. */
void FN(orc_lhs_assemble)(REAL* lhs, int64_t ld, int32_t nsurf, const int64_t* ns, const int64_t* nc,
                          const int64_t* nwv, REAL* const* colloc, REAL* const* normals,
                          REAL* const* verts_wing, REAL* const* verts_wake, int32_t iteration,
                          REAL sigma) {
    int64_t offset_j = 0;
    for (int32_t mj = 0; mj < nsurf; mj++) {
        int64_t offset_i = 0;
        const int64_t npan_j = ns[mj] * nc[mj];
        for (int32_t mi = 0; mi < nsurf; mi++) {
            const int64_t m = ns[mi] * nc[mi];
            REAL* section = lhs + offset_i + offset_j * ld;
            const REAL* ci = colloc[mi];
            const REAL* ni = normals[mi];
            const REAL* vw = verts_wing[mj];
            const int64_t vw_ld = (ns[mj] + 1) * (nc[mj] + 1);
#pragma omp parallel for schedule(static)
            for (int64_t lidx = 0; lidx < npan_j; lidx++) {
                FN(influence)(m, section + lidx * ld, ci, m, vw + lidx + lidx / ns[mj], vw_ld, ns[mj] + 1, ni, m, sigma);
            }
            const int64_t end_wing = (nc[mj] - 1) * ns[mj];
            const int64_t vk_ld = (ns[mj] + 1) * nwv[mj];
            for (int32_t k = 0; k < iteration; k++) {
                const int64_t row = nwv[mj] - 2 - k;
#pragma omp parallel for schedule(static)
                for (int64_t j = 0; j < ns[mj]; j++) {
                    FN(influence)(m, section + (j + end_wing) * ld, ci, m, verts_wake[mj] + j + row * (ns[mj] + 1), vk_ld,
                                  ns[mj] + 1, ni, m, sigma);
                }
            }
            offset_i += m;
        }
        offset_j += npan_j;
    }
}

/* Backend::rhs_assemble_velocities (vlm_backend_cpu.cpp:156-179): rhs(p) += -(V_p . n_p). */
void FN(orc_rhs_assemble_velocities)(REAL* rhs, int32_t nsurf, const int64_t* ns, const int64_t* nc,
                                     REAL* const* normals, REAL* const* velocities) {
    int64_t offset = 0;
    for (int32_t m = 0; m < nsurf; m++) {
        const int64_t np = ns[m] * nc[m];
        const REAL* n = normals[m];
        const REAL* v = velocities[m];
        for (int64_t p = 0; p < np; p++) {
            rhs[offset + p] += -(v[p] * n[p] + v[np + p] * n[np + p] + v[2 * np + p] * n[2 * np + p]);
        }
        offset += np;
    }
}

/* kernel_wake_influence / dkernel_wake_influence (kernels.ispc:150-196 / :463-502) for one receiver:
 * rings of wake rows [v_m-1-iteration, end) with end = v_m-1 (FP32 keeps only the newest 5000
 * rows, kernels.ispc:157-162); ring (i,j) carries gamma[i*(v_n-1)+j]. */
static void FN(wake_influence_one)(const REAL* c, const REAL* n, const REAL* v, int64_t v_ld, int64_t v_m,
                                   int64_t v_n, const REAL* gamma, REAL* rhs, REAL sigma, int32_t iteration) {
    REAL ind_tot[3] = {0, 0, 0};
    int64_t end = v_m - 1;
#if !IS_F64
    if (iteration > 5000) end = v_m - 1 - iteration + 5000;
#endif
    for (int64_t i = v_m - iteration - 1; i < end; i++) {
        for (int64_t j = 0; j < v_n - 1; j++) {
            const int64_t i0 = i * v_n + j, i1 = i0 + 1, i2 = (i + 1) * v_n + j + 1, i3 = (i + 1) * v_n + j;
            const REAL v0[3] = {v[i0], v[v_ld + i0], v[2 * v_ld + i0]};
            const REAL v1[3] = {v[i1], v[v_ld + i1], v[2 * v_ld + i1]};
            const REAL v2[3] = {v[i2], v[v_ld + i2], v[2 * v_ld + i2]};
            const REAL v3[3] = {v[i3], v[v_ld + i3], v[2 * v_ld + i3]};
            REAL ind[3];
            FN(ring)(ind, c, v0, v1, v2, v3, sigma);
            const REAL g = gamma[i * (v_n - 1) + j];
            ind_tot[0] += ind[0] * g;
            ind_tot[1] += ind[1] * g;
            ind_tot[2] += ind[2] * g;
        }
    }
    *rhs -= FN(dot3)(ind_tot, n);
}

/* Far-wake aggregation (no counterpart in the reference; the served UVLM loop's approximation, accepted
 * only at |dCL| <= 1e-6 against the exact wake).  far = {L, A_1, G_1, S_1, ..., A_L, G_L, S_L}: L levels
 * with age thresholds A_1 < ... < A_L, row groups G_l (each a multiple of the one before) and span
 * groups S_l.  At iteration it, level l's blocks are G_l consecutive ring rows counted from the oldest
 * row (block b: rows [end-(b+1)G_l, end-bG_l), end = v_m-1); the C_l = floor((it - A_l) / G_l) oldest
 * blocks are older than A_l.  From the oldest end, level L takes its C_L blocks, each finer level the
 * blocks of its own count that the coarser levels left, and the remaining (newest) rows stay exact.
 * Within a level, each block is cut into groups of S_l spans (group j: spans [jS, min((j+1)S, ns))) and
 * each (block, group) is replaced by ONE vortex ring through the lattice vertices at its corners,
 * carrying the mean circulation of the G_l x S rings it replaces (summed rows-then-spans in order). */
static void FN(far_levels)(int32_t iteration, int64_t end, const int32_t* far, int64_t* lo, int64_t* nb) {
    const int32_t L = far[0];
    int64_t hi = end;
    for (int32_t l = L - 1; l >= 0; l--) {
        const int64_t A = far[1 + 3 * l], G = far[2 + 3 * l];
        const int64_t C = iteration >= A ? (iteration - A) / G : 0;
        const int64_t cov = (end - hi) / G;
        nb[l] = C > cov ? C - cov : 0;
        lo[l] = hi - nb[l] * G;
        hi = lo[l];
    }
}

static void FN(wake_influence_far_one)(const REAL* c, const REAL* n, const REAL* v, int64_t v_ld, int64_t v_m,
                                       int64_t v_n, const REAL* gamma, REAL* rhs, REAL sigma, int32_t iteration,
                                       const int32_t* far) {
    const int64_t ns = v_n - 1, end = v_m - 1;
    const int32_t L = far[0];
    int64_t lo[16], nb[16];
    FN(far_levels)(iteration, end, far, lo, nb);
    const int64_t near_end = L > 0 ? lo[0] : end;
    REAL ind_tot[3] = {0, 0, 0};
    for (int64_t i = v_m - iteration - 1; i < near_end; i++) {
        for (int64_t j = 0; j < ns; j++) {
            const int64_t i0 = i * v_n + j, i1 = i0 + 1, i2 = (i + 1) * v_n + j + 1, i3 = (i + 1) * v_n + j;
            const REAL v0[3] = {v[i0], v[v_ld + i0], v[2 * v_ld + i0]};
            const REAL v1[3] = {v[i1], v[v_ld + i1], v[2 * v_ld + i1]};
            const REAL v2[3] = {v[i2], v[v_ld + i2], v[2 * v_ld + i2]};
            const REAL v3[3] = {v[i3], v[v_ld + i3], v[2 * v_ld + i3]};
            REAL ind[3];
            FN(ring)(ind, c, v0, v1, v2, v3, sigma);
            const REAL g = gamma[i * ns + j];
            ind_tot[0] += ind[0] * g;
            ind_tot[1] += ind[1] * g;
            ind_tot[2] += ind[2] * g;
        }
    }
    for (int32_t l = 0; l < L; l++) {
        const int64_t G = far[2 + 3 * l], S = far[3 + 3 * l];
        for (int64_t q = 0; q < nb[l]; q++) {
            const int64_t rt = lo[l] + q * G, rb = rt + G; /* vertex rows of the block's edges */
            for (int64_t j0 = 0; j0 < ns; j0 += S) {
                const int64_t j1 = j0 + S < ns ? j0 + S : ns;
                REAL gs = 0;
                for (int64_t i = rt; i < rb; i++)
                    for (int64_t j = j0; j < j1; j++) gs += gamma[i * ns + j];
                const REAL g = gs / (REAL)((rb - rt) * (j1 - j0));
                const int64_t i0 = rt * v_n + j0, i1 = rt * v_n + j1, i2 = rb * v_n + j1, i3 = rb * v_n + j0;
                const REAL v0[3] = {v[i0], v[v_ld + i0], v[2 * v_ld + i0]};
                const REAL v1[3] = {v[i1], v[v_ld + i1], v[2 * v_ld + i1]};
                const REAL v2[3] = {v[i2], v[v_ld + i2], v[2 * v_ld + i2]};
                const REAL v3[3] = {v[i3], v[v_ld + i3], v[2 * v_ld + i3]};
                REAL ind[3];
                FN(ring)(ind, c, v0, v1, v2, v3, sigma);
                ind_tot[0] += ind[0] * g;
                ind_tot[1] += ind[1] * g;
                ind_tot[2] += ind[2] * g;
            }
        }
    }
    *rhs -= FN(dot3)(ind_tot, n);
}

/* Backend::rhs_assemble_wake_influence (vlm_backend_cpu.cpp:181-224, vlm_backend_cpud.cpp:124-167):
 * every receiver of every surface sees the wake of every lifting surface. */
static void FN(wake_influence_all)(REAL* rhs, int32_t nsurf, const int64_t* ns, const int64_t* nc,
                                   const int64_t* nwv, REAL* const* gamma_wake, REAL* const* colloc,
                                   REAL* const* normals, REAL* const* verts_wake, const int32_t* lifting,
                                   int32_t iteration, REAL sigma, const int32_t* far) {
    int64_t offset = 0;
    for (int32_t mi = 0; mi < nsurf; mi++) {
        const int64_t np = ns[mi] * nc[mi];
        const REAL* ci = colloc[mi];
        const REAL* ni = normals[mi];
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t idx = 0; idx < np; idx++) {
            const REAL c[3] = {ci[idx], ci[np + idx], ci[2 * np + idx]};
            const REAL n[3] = {ni[idx], ni[np + idx], ni[2 * np + idx]};
            for (int32_t mj = 0; mj < nsurf; mj++) {
                if (!lifting[mj]) continue;
                if (far && far[0] > 0)
                    FN(wake_influence_far_one)(c, n, verts_wake[mj], (ns[mj] + 1) * nwv[mj], nwv[mj], ns[mj] + 1,
                                               gamma_wake[mj], rhs + offset + idx, sigma, iteration, far);
                else
                    FN(wake_influence_one)(c, n, verts_wake[mj], (ns[mj] + 1) * nwv[mj], nwv[mj], ns[mj] + 1,
                                           gamma_wake[mj], rhs + offset + idx, sigma, iteration);
            }
        }
        offset += np;
    }
}
void FN(orc_rhs_assemble_wake_influence)(REAL* rhs, int32_t nsurf, const int64_t* ns, const int64_t* nc,
                                         const int64_t* nwv, REAL* const* gamma_wake, REAL* const* colloc,
                                         REAL* const* normals, REAL* const* verts_wake, const int32_t* lifting,
                                         int32_t iteration, REAL sigma) {
    FN(wake_influence_all)(rhs, nsurf, ns, nc, nwv, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                           NULL);
}

/* The same with far-wake aggregation (far as in orc_uvlm_run_far; NULL or L = 0: exact). */
void FN(orc_rhs_assemble_wake_influence_far)(REAL* rhs, int32_t nsurf, const int64_t* ns, const int64_t* nc,
                                             const int64_t* nwv, REAL* const* gamma_wake, REAL* const* colloc,
                                             REAL* const* normals, REAL* const* verts_wake, const int32_t* lifting,
                                             int32_t iteration, REAL sigma, const int32_t* far) {
    FN(wake_influence_all)(rhs, nsurf, ns, nc, nwv, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma,
                           far);
}

/* Backend::mesh_metrics (vlm_backend_cpu.cpp:391-433, vlm_backend_cpud.cpp:258-300). */
void FN(orc_mesh_metrics)(REAL alpha_rad, int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* verts_wing,
                          REAL* const* colloc, REAL* const* normals, REAL* const* areas) {
    for (int32_t m = 0; m < nsurf; m++) {
        const int64_t vs1 = ns[m] + 1, vs2 = (ns[m] + 1) * (nc[m] + 1), ps2 = ns[m] * nc[m];
        const REAL* V = verts_wing[m];
        for (int64_t j = 0; j < nc[m]; j++) {
            for (int64_t i = 0; i < ns[m]; i++) {
#define ORC_VTX(dst, ii, jj) const REAL dst[3] = {V[(ii) + (jj)*vs1], V[(ii) + (jj)*vs1 + vs2], V[(ii) + (jj)*vs1 + 2 * vs2]}
                ORC_VTX(v0, i, j);
                ORC_VTX(v1, i + 1, j);
                ORC_VTX(v2, i + 1, j + 1);
                ORC_VTX(v3, i, j + 1);
#undef ORC_VTX
                const int64_t p = i + j * ns[m];
                const REAL a[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
                const REAL b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
                REAL nv[3];
                FN(cross3)(a, b, nv);
                const REAL l = SQRT(FN(dot3)(nv, nv));
                normals[m][p] = nv[0] / l;
                normals[m][p + ps2] = nv[1] / l;
                normals[m][p + 2 * ps2] = nv[2] / l;

                const REAL f[3] = {v3[0] - v0[0], v3[1] - v0[1], v3[2] - v0[2]};
                const REAL bb[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
                const REAL e[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
                REAL fb[3], be[3];
                FN(cross3)(f, bb, fb);
                FN(cross3)(bb, e, be);
                areas[m][p] = (REAL)0.5 * (SQRT(FN(dot3)(fb, fb)) + SQRT(FN(dot3)(be, be)));

                const REAL factor =
                    (alpha_rad < ORC_EPS_F) ? (REAL)0.5 : (REAL)0.5 * (alpha_rad / (SIN(alpha_rad) + ORC_EPS_F));
                for (int k = 0; k < 3; k++) {
                    const REAL chord = (REAL)0.5 * (((v2[k] + v3[k]) - v0[k]) - v1[k]);
                    colloc[m][p + k * ps2] = (REAL)0.5 * (v0[k] + v1[k]) + factor * chord;
                }
            }
        }
    }
}

/* Backend::forces_steady (vlm_backend_cpu.cpp:276-297): F = rho*dGamma * (V x (v(i+1,j) - v(i,j))), rho = 1. */
void FN(orc_forces_steady)(int64_t ns, int64_t nc, const REAL* verts_wing, const REAL* gamma_delta, const REAL* velocities,
                           REAL* forces) {
    const REAL rho = (REAL)1;
    const int64_t vs1 = ns + 1, vs2 = (ns + 1) * (nc + 1), ps2 = ns * nc;
    for (int64_t j = 0; j < nc; j++) {
        for (int64_t i = 0; i < ns; i++) {
            const int64_t p = i + j * ns;
            const REAL V[3] = {velocities[p], velocities[p + ps2], velocities[p + 2 * ps2]};
            const int64_t a = i + j * vs1, b = i + 1 + j * vs1;
            const REAL d[3] = {verts_wing[b] - verts_wing[a], verts_wing[b + vs2] - verts_wing[a + vs2],
                               verts_wing[b + 2 * vs2] - verts_wing[a + 2 * vs2]};
            REAL c[3];
            FN(cross3)(V, d, c);
            const REAL s = rho * gamma_delta[p];
            forces[p] = s * c[0];
            forces[p + ps2] = s * c[1];
            forces[p + 2 * ps2] = s * c[2];
        }
    }
}

/* Backend::forces_unsteady (vlm_backend_cpu.cpp:299-321): steady part + rho*dGamma/dt*A*n. */
void FN(orc_forces_unsteady)(int64_t ns, int64_t nc, const REAL* verts_wing, const REAL* gamma_delta, const REAL* dgamma_dt,
                             const REAL* velocities, const REAL* areas, const REAL* normals, REAL* forces) {
    FN(orc_forces_steady)(ns, nc, verts_wing, gamma_delta, velocities, forces);
    const REAL rho = (REAL)1;
    const int64_t ps2 = ns * nc;
    for (int64_t p = 0; p < ps2; p++) {
        const REAL s = rho * dgamma_dt[p] * areas[p];
        forces[p] += s * normals[p];
        forces[p + ps2] += s * normals[p + ps2];
        forces[p + 2 * ps2] += s * normals[p + 2 * ps2];
    }
}

/* Backend::coeff_cl (vlm_backend_cpu.cpp:323-338). */
REAL FN(orc_coeff_cl)(int64_t ns, int64_t nc, const REAL* forces, const REAL* lift_axis, const REAL* freestream, REAL rho,
                      REAL area) {
    const int64_t ps2 = ns * nc;
    REAL cl = 0;
    for (int64_t p = 0; p < ps2; p++) {
        const REAL f[3] = {forces[p], forces[p + ps2], forces[p + 2 * ps2]};
        cl += FN(dot3)(f, lift_axis);
    }
    return cl / ((REAL)0.5 * rho * FN(dot3)(freestream, freestream) * area);
}

/* Backend::coeff_cm (vlm_backend_cpu.cpp:340-362): sum of (mid leading vortex line - ref) x F. */
void FN(orc_coeff_cm)(int64_t ns, int64_t nc, const REAL* forces, const REAL* verts_wing, const REAL* ref_pt,
                      const REAL* freestream, REAL rho, REAL area, REAL mac, REAL* out) {
    const int64_t vs1 = ns + 1, vs2 = (ns + 1) * (nc + 1), ps2 = ns * nc;
    REAL cm[3] = {0, 0, 0};
    for (int64_t j = 0; j < nc; j++) {
        for (int64_t i = 0; i < ns; i++) {
            const int64_t p = i + j * ns, a = i + j * vs1, b = i + 1 + j * vs1;
            REAL lever[3], f[3], c[3];
            for (int k = 0; k < 3; k++) {
                lever[k] = (REAL)0.5 * (verts_wing[a + k * vs2] + verts_wing[b + k * vs2]) - ref_pt[k];
                f[k] = forces[p + k * ps2];
            }
            FN(cross3)(lever, f, c);
            cm[0] += c[0];
            cm[1] += c[1];
            cm[2] += c[2];
        }
    }
    const REAL den = (REAL)0.5 * rho * FN(dot3)(freestream, freestream) * area * mac;
    out[0] = cm[0] / den;
    out[1] = cm[1] / den;
    out[2] = cm[2] / den;
}

/* Backend::sum (vlm_backend_cpu.cpp:506-523): sequential sum in storage order. */
REAL FN(orc_sum)(int64_t n, const REAL* x) {
    REAL s = 0;
    for (int64_t i = 0; i < n; i++) s += x[i];
    return s;
}

/* Backend::mesh_mac (vlm_backend_cpu.cpp:443-468): trapezoidal integral of c(y)^2 over S. */
REAL FN(orc_mesh_mac)(int64_t ns, int64_t nc, const REAL* verts_wing, const REAL* areas) {
    const int64_t vs1 = ns + 1, vs2 = (ns + 1) * (nc + 1);
    REAL mac = 0;
    for (int64_t i = 0; i < ns; i++) {
        REAL d0[3], d1[3], d3[3];
        for (int k = 0; k < 3; k++) {
            d0[k] = verts_wing[i + k * vs2] - verts_wing[i + nc * vs1 + k * vs2];
            d1[k] = verts_wing[i + 1 + k * vs2] - verts_wing[i + 1 + nc * vs1 + k * vs2];
            d3[k] = verts_wing[i + 1 + k * vs2] - verts_wing[i + k * vs2];
        }
        const REAL c0 = SQRT(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
        const REAL c1 = SQRT(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
        const REAL w = SQRT(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
        mac += (REAL)0.5 * (c0 * c0 + c1 * c1) * w;
    }
    return mac / FN(orc_sum)(ns * nc, areas);
}

/* coeff_cl_multibody_impl (vlm/src/vlm_backend.cpp:71-100): area-weighted CL over surfaces,
 * lift axis = normalize(freestream x (0,1,0)). */
REAL FN(orc_coeff_cl_multibody)(int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* forces, REAL* const* areas,
                                const REAL* freestream, REAL rho) {
    REAL cl = 0, total = 0;
    const REAL yax[3] = {0, 1, 0};
    REAL la[3];
    FN(cross3)(freestream, yax, la);
    const REAL l = SQRT(FN(dot3)(la, la));
    la[0] /= l;
    la[1] /= l;
    la[2] /= l;
    for (int32_t m = 0; m < nsurf; m++) {
        const REAL a = FN(orc_sum)(ns[m] * nc[m], areas[m]);
        cl += FN(orc_coeff_cl)(ns[m], nc[m], forces[m], la, freestream, rho, a) * a;
        total += a;
    }
    return cl / total;
}

/* coeff_cm_multibody_impl (vlm/src/vlm_backend.cpp:102-139). */
void FN(orc_coeff_cm_multibody)(int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* forces,
                                REAL* const* verts_wing, REAL* const* areas, const REAL* ref_pt, const REAL* freestream,
                                REAL rho, REAL* out) {
    REAL cm[3] = {0, 0, 0}, total_area = 0, total_mac = 0;
    for (int32_t m = 0; m < nsurf; m++) {
        const REAL a = FN(orc_sum)(ns[m] * nc[m], areas[m]);
        const REAL mac = FN(orc_mesh_mac)(ns[m], nc[m], verts_wing[m], areas[m]);
        REAL lc[3];
        FN(orc_coeff_cm)(ns[m], nc[m], forces[m], verts_wing[m], ref_pt, freestream, rho, a, mac, lc);
        for (int k = 0; k < 3; k++) cm[k] += lc[k] * a * mac;
        total_area += a;
        total_mac += mac;
    }
    for (int k = 0; k < 3; k++) out[k] = cm[k] / (total_area * total_mac);
}

/* wake_shed_impl (vlm/src/vlm_backend.cpp:42-47): wing TE vertex row -> wake vertex row -1-iteration. */
void FN(orc_wake_shed)(int64_t ns, int64_t nc, int64_t nwv, const REAL* verts_wing, REAL* verts_wake, int32_t iteration) {
    const int64_t vs1 = ns + 1, vs2 = (ns + 1) * (nc + 1), ws2 = (ns + 1) * nwv;
    const int64_t row = nwv - 1 - iteration;
    for (int k = 0; k < 4; k++)
        for (int64_t i = 0; i <= ns; i++) verts_wake[i + row * vs1 + k * ws2] = verts_wing[i + nc * vs1 + k * vs2];
}

/* displace_wing_impl (vlm/src/vlm_backend.cpp:49-69): verts {P,4} = verts_init {P,4} * T^T, where the
 * {4,4} column-major transform holds the homogeneous matrix (headeronly/linalg.h store()). */
void FN(orc_displace_wing)(int64_t nverts, const REAL* transform, const REAL* verts_init, REAL* verts) {
    for (int64_t p = 0; p < nverts; p++) {
        for (int r = 0; r < 4; r++) {
            REAL s = 0;
            for (int k = 0; k < 4; k++) s += verts_init[p + k * nverts] * transform[r + 4 * k];
            verts[p + r * nverts] = s;
        }
    }
}

/* Backend::gamma_wake_from_coeffs (vlm_backend_cpu.cpp:470-503, vlm_backend_cpud.cpp:337-370). */
void FN(orc_gamma_wake_from_coeffs)(int64_t ns, int64_t nw, REAL* gamma_wake, const REAL* gamma_coeffs, int32_t harmonics,
                                    REAL tn, REAL omega, REAL dt, int64_t iteration) {
    const int32_t K = 2 * harmonics + 1;
    const REAL s1 = (REAL)1 / SQRT((REAL)K);
    const REAL s2 = SQRT((REAL)2 / (REAL)K);
    const int64_t start = nw - iteration;
    for (int64_t j = start; j < nw; j++) {
        for (int64_t i = 0; i < ns; i++) {
            REAL g = gamma_coeffs[i] * s1;
            for (int64_t h = 0; h < harmonics; h++) {
                const REAL wk = omega * (REAL)(h + 1);
                const REAL arg = wk * (tn - (REAL)(j - start + 1) * dt);
                g += gamma_coeffs[i + (2 * h + 1) * ns] * COS(arg) * s2;
                g += gamma_coeffs[i + (2 * h + 2) * ns] * SIN(arg) * s2;
            }
            gamma_wake[i + j * ns] = g;
        }
    }
}

/* LU::factorize -> LAPACKE_?getrf (vlm_backend_cpu_lapack.cpp:40-50, :70-80): textbook partial
 * pivoting (pivot = first index of max |a|, as i?amax), column-major, 1-based ipiv.  Blocked
 * right-looking form (panel + row swaps + triangular solve + Schur update), same mathematics as
 * LAPACK's dgetrf.  Returns info (0, or k>0 if U(k,k) == 0). */
int FN(orc_getrf)(int64_t n, REAL* a, int64_t lda, int32_t* ipiv) {
    const int64_t nb = 64;
    int info = 0;
#define A_(i, j) a[(i) + (j)*lda]
    for (int64_t k0 = 0; k0 < n; k0 += nb) {
        const int64_t kb = (n - k0 < nb) ? n - k0 : nb;
        /* unblocked panel on columns [k0, k0+kb) over rows [k0, n), swaps applied to the full rows */
        for (int64_t k = k0; k < k0 + kb; k++) {
            int64_t p = k;
            REAL amax = A_(k, k) < 0 ? -A_(k, k) : A_(k, k);
            for (int64_t i = k + 1; i < n; i++) {
                const REAL v = A_(i, k) < 0 ? -A_(i, k) : A_(i, k);
                if (v > amax) {
                    amax = v;
                    p = i;
                }
            }
            ipiv[k] = (int32_t)(p + 1);
            if (p != k) {
                for (int64_t j = 0; j < n; j++) {
                    const REAL t = A_(k, j);
                    A_(k, j) = A_(p, j);
                    A_(p, j) = t;
                }
            }
            if (A_(k, k) != (REAL)0) {
                const REAL r = (REAL)1 / A_(k, k);
                for (int64_t i = k + 1; i < n; i++) A_(i, k) *= r;
            } else if (info == 0) {
                info = (int)(k + 1);
            }
            for (int64_t j = k + 1; j < k0 + kb; j++) {
                const REAL ukj = A_(k, j);
                for (int64_t i = k + 1; i < n; i++) A_(i, j) -= A_(i, k) * ukj;
            }
        }
        const int64_t j1 = k0 + kb;
        if (j1 >= n) continue;
        /* U12 = L11^{-1} A12 */
#pragma omp parallel for schedule(static)
        for (int64_t j = j1; j < n; j++) {
            for (int64_t k = k0; k < j1; k++) {
                const REAL ukj = A_(k, j);
                for (int64_t i = k + 1; i < j1; i++) A_(i, j) -= A_(i, k) * ukj;
            }
        }
        /* A22 -= L21 * U12 */
#pragma omp parallel for schedule(static)
        for (int64_t j = j1; j < n; j++) {
            for (int64_t k = k0; k < j1; k++) {
                const REAL ukj = A_(k, j);
                for (int64_t i = j1; i < n; i++) A_(i, j) -= A_(i, k) * ukj;
            }
        }
    }
#undef A_
    return info;
}

/* LU::solve -> LAPACKE_?getrs('N') (vlm_backend_cpu_lapack.cpp:52-64): apply row swaps, then
 * L (unit) forward and U backward substitution on every right-hand side. */
void FN(orc_getrs)(int64_t n, int64_t nrhs, const REAL* a, int64_t lda, const int32_t* ipiv, REAL* b, int64_t ldb) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nrhs; c++) {
        REAL* x = b + c * ldb;
        for (int64_t k = 0; k < n; k++) {
            const int64_t p = ipiv[k] - 1;
            if (p != k) {
                const REAL t = x[k];
                x[k] = x[p];
                x[p] = t;
            }
        }
        for (int64_t k = 0; k < n; k++) {
            const REAL xk = x[k];
            for (int64_t i = k + 1; i < n; i++) x[i] -= a[i + k * lda] * xk;
        }
        for (int64_t k = n - 1; k >= 0; k--) {
            x[k] /= a[k + k * lda];
            const REAL xk = x[k];
            for (int64_t i = 0; i < k; i++) x[i] -= a[i + k * lda] * xk;
        }
    }
}

/* LSQ::solve = LAPACKE_?gels 'N' (vlm/backends/cpu/src/vlm_backend_cpu_lapack.cpp:141-205), M >= N:
 * ?geqr2 Householder QR (reflector ?larfg: beta = -sign(alpha) sqrt(alpha^2 + |x|^2),
 * tau = (beta - alpha) / beta, v = x / (alpha - beta)), each H(j) applied to the trailing columns
 * (?larf) and to B in the order H(1), H(2), ... (?orm2r), then R X = (Q^T B)(0:N) by
 * column-oriented back substitution (?trsm).  Returns info = i if R(i,i) == 0 (then X is not
 * computed); an all-zero A sets B to zero.  LAPACK's rescaling of extreme norms is not restated. */
int FN(orc_gels)(int64_t m, int64_t n, int64_t nrhs, REAL* a, int64_t lda, REAL* b, int64_t ldb) {
    if (m < n) return -1;
    REAL amax = 0;
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) {
            const REAL v = a[i + j * lda] < 0 ? -a[i + j * lda] : a[i + j * lda];
            if (v > amax) amax = v;
        }
    if (amax == 0) {
        for (int64_t r = 0; r < nrhs; r++)
            for (int64_t i = 0; i < m; i++) b[i + r * ldb] = 0;
        return 0;
    }
    for (int64_t j = 0; j < n; j++) {
        REAL* cj = a + j * lda;
        REAL s = 0;
        for (int64_t i = j + 1; i < m; i++) s += cj[i] * cj[i];
        const REAL alpha = cj[j];
        REAL tau = 0, beta = alpha;
        if (s != 0) {
            const REAL mag = SQRT(alpha * alpha + s);
            beta = alpha >= 0 ? -mag : mag;
            tau = (beta - alpha) / beta;
            const REAL scal = (REAL)1 / (alpha - beta);
            for (int64_t i = j + 1; i < m; i++) cj[i] *= scal;
        }
        if (tau != 0) {
            const int64_t na = n - j - 1;
            for (int64_t k = 0; k < na + nrhs; k++) {
                REAL* c = k < na ? a + (j + 1 + k) * lda : b + (k - na) * ldb;
                REAL w = c[j];
                for (int64_t i = j + 1; i < m; i++) w += cj[i] * c[i];
                w *= tau;
                c[j] -= w;
                for (int64_t i = j + 1; i < m; i++) c[i] -= w * cj[i];
            }
        }
        cj[j] = beta;
    }
    for (int64_t i = 0; i < n; i++)
        if (a[i + i * lda] == 0) return (int)(i + 1);
    for (int64_t r = 0; r < nrhs; r++) {
        REAL* x = b + r * ldb;
        for (int64_t jj = n - 1; jj >= 0; jj--) {
            const REAL* col = a + jj * lda;
            const REAL xj = x[jj] / col[jj];
            x[jj] = xj;
            for (int64_t i = 0; i < jj; i++) x[i] -= xj * col[i];
        }
    }
    return 0;
}

/* mesh_quarterchord (vlm/src/vlm_mesh.cpp:15-32): move each chordwise vertex row to its quarter
 * chord, then extrapolate the trailing edge with the float constants 4.f/3.f and 1.f/3.f. */
void FN(orc_quarterchord)(int64_t ns, int64_t nc, REAL* v) {
    const int64_t vs1 = ns + 1, vs2 = (ns + 1) * (nc + 1);
    for (int64_t j = 0; j < nc; j++)
        for (int64_t i = 0; i <= ns; i++)
            for (int k = 0; k < 3; k++)
                v[i + j * vs1 + k * vs2] = (REAL)0.75f * v[i + j * vs1 + k * vs2] + (REAL)0.25f * v[i + (j + 1) * vs1 + k * vs2];
    const int64_t j = nc - 1;
    for (int64_t i = 0; i <= ns; i++)
        for (int k = 0; k < 3; k++)
            v[i + (j + 1) * vs1 + k * vs2] =
                (REAL)(4.f / 3.f) * v[i + (j + 1) * vs1 + k * vs2] - (REAL)(1.f / 3.f) * v[i + j * vs1 + k * vs2];
}

/* ---------------------------------------------------------------------------------------------- */
/* Drivers                                                                                         */
/* ---------------------------------------------------------------------------------------------- */

/* gamma post-processing of the drivers (vlm/src/vlm.cpp:129-142 and :442-458): copy the solution
 * slice, chordwise delta (gamma(i,j) - gamma(i,j-1) for j>0). */
static void FN(gamma_delta)(int64_t ns, int64_t nc, const REAL* gamma, REAL* delta) {
    for (int64_t p = 0; p < ns * nc; p++) delta[p] = gamma[p];
    for (int64_t j = 1; j < nc; j++)
        for (int64_t i = 0; i < ns; i++) delta[i + j * ns] += -(REAL)1 * gamma[i + (j - 1) * ns];
}

/* VLM::run (vlm/src/vlm.cpp:97-151) in precision REAL: 100*u_inf chord wake, iteration=1 LHS, LU,
 * gamma post-processing, steady forces, area-weighted CL.  Additionally returns CM about
 * cm_ref_pt transformed by the same wake translation as the wing (the reference driver returns
 * CM = 0; coeff_cm_multibody is the reference routine).  coeffs_out = {CL, CMx, CMy, CMz}.
 * Returns the getrf info. */
int FN(orc_vlm_run)(int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* verts_init, REAL alpha, REAL beta,
                    REAL u_inf, REAL rho, const REAL* cm_ref_pt, REAL* gamma_out, REAL* coeffs_out) {
    if (nsurf < 1 || nsurf > 64) return -1;
    int64_t N = 0;
    for (int32_t m = 0; m < nsurf; m++) N += ns[m] * nc[m];
    REAL* lhs = (REAL*)calloc((size_t)(N * N), sizeof(REAL));
    REAL* rhs = (REAL*)calloc((size_t)N, sizeof(REAL));
    int32_t* ipiv = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    REAL *vw[64], *vk[64], *co[64], *no[64], *ar[64], *ve[64], *fo[64], *gd[64];
    int64_t nwv[64];
    const REAL fs[3] = {u_inf * COS(alpha) * COS(beta), -u_inf * COS(alpha) * SIN(beta), u_inf * SIN(alpha)};
    REAL T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, (REAL)-100 * u_inf * COS(alpha), 0, (REAL)-100 * u_inf * SIN(alpha), 1};
    for (int32_t m = 0; m < nsurf; m++) {
        const int64_t nv = (ns[m] + 1) * (nc[m] + 1), np = ns[m] * nc[m];
        nwv[m] = 2;
        vw[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(nv * 4));
        vk[m] = (REAL*)calloc((size_t)((ns[m] + 1) * 2 * 4), sizeof(REAL));
        co[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        no[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        ve[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        fo[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        ar[m] = (REAL*)malloc(sizeof(REAL) * (size_t)np);
        gd[m] = (REAL*)malloc(sizeof(REAL) * (size_t)np);
        memcpy(vw[m], verts_init[m], sizeof(REAL) * (size_t)(nv * 4));
        FN(orc_wake_shed)(ns[m], nc[m], 2, vw[m], vk[m], 0);
        FN(orc_displace_wing)(nv, T, verts_init[m], vw[m]);
        FN(orc_wake_shed)(ns[m], nc[m], 2, vw[m], vk[m], 1);
        for (int64_t p = 0; p < np; p++) {
            ve[m][p] = fs[0];
            ve[m][p + np] = fs[1];
            ve[m][p + 2 * np] = fs[2];
        }
    }
    FN(orc_mesh_metrics)(alpha, nsurf, ns, nc, vw, co, no, ar);
    FN(orc_lhs_assemble)(lhs, N, nsurf, ns, nc, nwv, co, no, vw, vk, 1, (REAL)0);
    FN(orc_rhs_assemble_velocities)(rhs, nsurf, ns, nc, no, ve);
    const int info = FN(orc_getrf)(N, lhs, N, ipiv);
    FN(orc_getrs)(N, 1, lhs, N, ipiv, rhs, N);
    int64_t off = 0;
    for (int32_t m = 0; m < nsurf; m++) {
        FN(gamma_delta)(ns[m], nc[m], rhs + off, gd[m]);
        FN(orc_forces_steady)(ns[m], nc[m], vw[m], gd[m], ve[m], fo[m]);
        off += ns[m] * nc[m];
    }
    memcpy(gamma_out, rhs, sizeof(REAL) * (size_t)N);
    coeffs_out[0] = FN(orc_coeff_cl_multibody)(nsurf, ns, nc, fo, ar, fs, rho);
    REAL ref[3];
    for (int r = 0; r < 3; r++) ref[r] = T[r] * cm_ref_pt[0] + T[r + 4] * cm_ref_pt[1] + T[r + 8] * cm_ref_pt[2] + T[r + 12];
    FN(orc_coeff_cm_multibody)(nsurf, ns, nc, fo, vw, ar, ref, fs, rho, coeffs_out + 1);
    for (int32_t m = 0; m < nsurf; m++) {
        free(vw[m]); free(vk[m]); free(co[m]); free(no[m]); free(ve[m]); free(fo[m]); free(ar[m]); free(gd[m]);
    }
    free(lhs);
    free(rhs);
    free(ipiv);
    return info;
}

/* UVLM::run (vlm/src/vlm.cpp:358-509) in precision REAL for the prescribed pitch/heave motion of
 * orc_motion_transform_d (all surfaces share it).  dt = last panel chord / |freestream|,
 * t_steps = (int64)(t_final/dt) unless t_steps_override > 0; the loop runs i = 0 .. t_steps-2.
 * Outputs per step i (valid for i >= 1; step 0 writes 0): cl_out[i], cm_out[i] (y component,
 * about T(t)*(0.5,0,0)).  gamma_out receives the last step's wing circulation (N values). */
int FN(orc_uvlm_run_far)(int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* verts_init,
                         const REAL* motion, int64_t t_steps_override, REAL t_final, REAL* cl_out, REAL* cm_out,
                         REAL* gamma_out, int64_t* t_steps_out, const int32_t* far) {
    if (nsurf < 1 || nsurf > 64) return -1;
    double mot[7];
    for (int k = 0; k < 7; k++) mot[k] = (double)motion[k];
    const REAL rho = (REAL)1;
    int64_t N = 0;
    for (int32_t m = 0; m < nsurf; m++) N += ns[m] * nc[m];
    double Td[16], dTd[16];
    REAL T[16];
    orc_motion_transform_d(mot, 0.0, Td, dTd);
    for (int k = 0; k < 16; k++) T[k] = (REAL)Td[k];

    /* dt from surface 0's last panel chord at station 0 (vlm.cpp:380-386) and the body speed |u| */
    const REAL* v0 = verts_init[0];
    const int64_t vs1 = ns[0] + 1, vs2 = (ns[0] + 1) * (nc[0] + 1);
    const REAL dx = v0[nc[0] * vs1] - v0[(nc[0] - 1) * vs1];
    const REAL dy = v0[nc[0] * vs1 + vs2] - v0[(nc[0] - 1) * vs1 + vs2];
    const REAL dz = v0[nc[0] * vs1 + 2 * vs2] - v0[(nc[0] - 1) * vs1 + 2 * vs2];
    const REAL chord = SQRT(dx * dx + dy * dy + dz * dz);
    const REAL dt = chord / (REAL)fabs(mot[0]);
    const int64_t t_steps = t_steps_override > 0 ? t_steps_override : (int64_t)(t_final / dt);
    *t_steps_out = t_steps;

    REAL* lhs = (REAL*)calloc((size_t)(N * N), sizeof(REAL));
    REAL* rhs = (REAL*)calloc((size_t)N, sizeof(REAL));
    int32_t* ipiv = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    REAL *vp[64], *vw[64], *vk[64], *co[64], *ch[64], *no[64], *ar[64], *ve[64], *fo[64], *gd[64], *gp[64], *gt[64], *gk[64];
    int64_t nwv[64];
    int32_t lifting[64];
    for (int32_t m = 0; m < nsurf; m++) {
        const int64_t nv = (ns[m] + 1) * (nc[m] + 1), np = ns[m] * nc[m];
        nwv[m] = t_steps;
        lifting[m] = 1;
        vp[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(nv * 4));
        vw[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(nv * 4));
        vk[m] = (REAL*)calloc((size_t)((ns[m] + 1) * t_steps * 4), sizeof(REAL));
        gk[m] = (REAL*)calloc((size_t)(ns[m] * (t_steps - 1)), sizeof(REAL));
        co[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        ch[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        no[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        ve[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        fo[m] = (REAL*)malloc(sizeof(REAL) * (size_t)(np * 3));
        ar[m] = (REAL*)malloc(sizeof(REAL) * (size_t)np);
        gd[m] = (REAL*)malloc(sizeof(REAL) * (size_t)np);
        gp[m] = (REAL*)calloc((size_t)np, sizeof(REAL));
        gt[m] = (REAL*)malloc(sizeof(REAL) * (size_t)np);
        FN(orc_displace_wing)(nv, T, verts_init[m], vp[m]);
        memcpy(vw[m], vp[m], sizeof(REAL) * (size_t)(nv * 4));
    }
    FN(orc_mesh_metrics)((REAL)0, nsurf, ns, nc, vw, co, no, ar);
    for (int32_t m = 0; m < nsurf; m++) memcpy(ch[m], co[m], sizeof(REAL) * (size_t)(ns[m] * nc[m] * 3));
    FN(orc_lhs_assemble)(lhs, N, nsurf, ns, nc, nwv, co, no, vw, vk, 0, (REAL)0);
    const int info = FN(orc_getrf)(N, lhs, N, ipiv);

    for (int64_t i = 0; i < t_steps - 1; i++) {
        const REAL t = (REAL)i * dt;
        orc_motion_transform_d(mot, (double)t, Td, dTd);
        /* freestream = -(velocity of the origin under the freestream node translation(-u t,0,0)) (vlm.cpp:414) */
        const REAL fsv[3] = {(REAL)mot[0], (REAL)0, (REAL)0};
        for (int32_t m = 0; m < nsurf; m++) {
            FN(orc_wake_shed)(ns[m], nc[m], nwv[m], vw[m], vk[m], (int32_t)i);
            const int64_t np = ns[m] * nc[m];
            for (int64_t p = 0; p < np; p++) {
                const double x = ch[m][p], y = ch[m][p + np], z = ch[m][p + 2 * np];
                for (int r = 0; r < 3; r++)
                    ve[m][p + r * np] = (REAL)(-(dTd[r] * x + dTd[r + 4] * y + dTd[r + 8] * z + dTd[r + 12]));
            }
        }
        memset(rhs, 0, sizeof(REAL) * (size_t)N);
        FN(orc_rhs_assemble_velocities)(rhs, nsurf, ns, nc, no, ve);
        FN(wake_influence_all)(rhs, nsurf, ns, nc, nwv, gk, co, no, vk, lifting, (int32_t)i, (REAL)0, far);
        FN(orc_getrs)(N, 1, lhs, N, ipiv, rhs, N);
        int64_t off = 0;
        const REAL inv_dt = (REAL)1 / dt;
        for (int32_t m = 0; m < nsurf; m++) {
            const int64_t np = ns[m] * nc[m];
            const REAL* g = rhs + off;
            FN(gamma_delta)(ns[m], nc[m], g, gd[m]);
            for (int64_t s = 0; s < ns[m]; s++) gk[m][s + (nwv[m] - 2 - i) * ns[m]] = g[s + (nc[m] - 1) * ns[m]];
            for (int64_t p = 0; p < np; p++) gt[m][p] = (g[p] + -(REAL)1 * gp[m][p]) * inv_dt;
            off += np;
        }
        if (i > 0) {
            for (int32_t m = 0; m < nsurf; m++)
                FN(orc_forces_unsteady)(ns[m], nc[m], vw[m], gd[m], gt[m], ve[m], ar[m], no[m], fo[m]);
            cl_out[i] = FN(orc_coeff_cl_multibody)(nsurf, ns, nc, fo, ar, fsv, rho);
            REAL ref[3], cm[3];
            for (int r = 0; r < 3; r++) ref[r] = (REAL)(Td[r] * 0.5 + Td[r + 12]);
            FN(orc_coeff_cm_multibody)(nsurf, ns, nc, fo, vw, ar, ref, fsv, rho, cm);
            cm_out[i] = cm[1];
        } else {
            cl_out[i] = 0;
            cm_out[i] = 0;
        }
        off = 0;
        for (int32_t m = 0; m < nsurf; m++) {
            memcpy(gp[m], rhs + off, sizeof(REAL) * (size_t)(ns[m] * nc[m]));
            off += ns[m] * nc[m];
        }
        orc_motion_transform_d(mot, (double)(t + dt), Td, dTd);
        for (int k = 0; k < 16; k++) T[k] = (REAL)Td[k];
        for (int32_t m = 0; m < nsurf; m++) FN(orc_displace_wing)((ns[m] + 1) * (nc[m] + 1), T, vp[m], vw[m]);
        FN(orc_mesh_metrics)((REAL)0, nsurf, ns, nc, vw, co, no, ar);
    }
    memcpy(gamma_out, rhs, sizeof(REAL) * (size_t)N);
    for (int32_t m = 0; m < nsurf; m++) {
        free(vp[m]); free(vw[m]); free(vk[m]); free(gk[m]); free(co[m]); free(ch[m]); free(no[m]);
        free(ve[m]); free(fo[m]); free(ar[m]); free(gd[m]); free(gp[m]); free(gt[m]);
    }
    free(lhs);
    free(rhs);
    free(ipiv);
    return info;
}

/* UVLM::run with the exact wake (the reference), or with far-wake aggregation: far = {L, A_1, G_1, S_1,
 * ...} (far-wake aggregation above), NULL or L = 0 for the exact wake. */
int FN(orc_uvlm_run)(int32_t nsurf, const int64_t* ns, const int64_t* nc, REAL* const* verts_init, const REAL* motion,
                     int64_t t_steps_override, REAL t_final, REAL* cl_out, REAL* cm_out, REAL* gamma_out,
                     int64_t* t_steps_out) {
    return FN(orc_uvlm_run_far)(nsurf, ns, nc, verts_init, motion, t_steps_override, t_final, cl_out, cm_out,
                                gamma_out, t_steps_out, NULL);
}

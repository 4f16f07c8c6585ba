/*
 * vlm_oracle.c -- CPU restatement of the tinyvlm reference backend (TEST INFRASTRUCTURE ONLY).
 * Only tests/, the smoke check and bench.py's CPU-baseline leg load the library built from this
 * file; the HIP product path never does.  Built with -ffp-contract=off so that every operation is
 * rounded as written in the reference sources.
 */
#include "vlm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- dual numbers (headeronly/tinyad.hpp:7-232 restated): value + d/dt ---------------------- */
typedef struct {
    double v, d;
} dual;

static dual dmul(dual a, dual b) { return (dual){a.v * b.v, b.v * a.d + a.v * b.d}; }
static dual dadd(dual a, dual b) { return (dual){a.v + b.v, a.d + b.d}; }

/* 4x4 dual matrix, column-major m[col*4+row] as linalg::mat stores columns x,y,z,w. */
typedef struct {
    dual m[16];
} dmat4;

static dmat4 dmat_identity(void) {
    dmat4 r;
    for (int k = 0; k < 16; k++) r.m[k] = (dual){(k % 5) == 0 ? 1.0 : 0.0, 0.0};
    return r;
}

/* linalg::mul(mat,mat): column j of the product is sum_k a.col(k) * b(k,j), k = 0..3 in order. */
static dmat4 dmat_mul(const dmat4* a, const dmat4* b) {
    dmat4 r;
    for (int j = 0; j < 4; j++) {
        for (int i = 0; i < 4; i++) {
            dual s = dmul(a->m[0 * 4 + i], b->m[j * 4 + 0]);
            for (int k = 1; k < 4; k++) s = dadd(s, dmul(a->m[k * 4 + i], b->m[j * 4 + k]));
            r.m[j * 4 + i] = s;
        }
    }
    return r;
}

/* translation_matrix (vlm/include/vlm_kinematics.hpp:149-157) */
static dmat4 dtranslation(dual x, dual y, dual z) {
    dmat4 r = dmat_identity();
    r.m[12] = x;
    r.m[13] = y;
    r.m[14] = z;
    return r;
}

/* rotation_matrix about the y axis through (px,0,0) by angle a (Rodrigues, vlm_kinematics.hpp:184-202):
 * R = [[cos,0,sin],[0,1,0],[-sin,0,cos]], translation part (I-R)p. */
static dmat4 drotation_y(double px, dual a) {
    const dual s = {sin(a.v), cos(a.v) * a.d};
    const dual c = {cos(a.v), -sin(a.v) * a.d};
    dmat4 r = dmat_identity();
    r.m[0] = c;                      /* (0,0) */
    r.m[2] = (dual){-s.v, -s.d};     /* (2,0) */
    r.m[8] = s;                      /* (0,2) */
    r.m[10] = c;                     /* (2,2) */
    r.m[12] = (dual){(1.0 - c.v) * px, -c.d * px}; /* ((I-R)p).x */
    r.m[13] = (dual){0.0, 0.0};
    r.m[14] = (dual){s.v * px, s.d * px};          /* ((I-R)p).z = -(-sin)*px */
    return r;
}

void orc_motion_transform_d(const double* motion, double t, double* T, double* dT) {
    const double u = motion[0], h_amp = motion[1], h_ph = motion[2], a_amp = motion[3], a_ph = motion[4];
    const double om = motion[5], xp = motion[6];
    const dmat4 fs = dtranslation((dual){-u * t, -u}, (dual){0, 0}, (dual){0, 0});
    const dmat4 hv = dtranslation((dual){0, 0}, (dual){0, 0},
                                  (dual){h_amp * sin(om * t + h_ph), h_amp * om * cos(om * t + h_ph)});
    const dmat4 pt = drotation_y(xp, (dual){a_amp * sin(om * t + a_ph), a_amp * om * cos(om * t + a_ph)});
    /* node chain pitch.after(heave).after(fs): transform_dual = fs * (heave * pitch) (vlm_kinematics.hpp:214-222) */
    const dmat4 hp = dmat_mul(&hv, &pt);
    const dmat4 all = dmat_mul(&fs, &hp);
    for (int k = 0; k < 16; k++) {
        T[k] = all.m[k].v;
        dT[k] = all.m[k].d;
    }
}

#ifdef _OPENMP
#include <omp.h>
#endif
/* OpenMP threads the parallel loops use (the CPU baseline reports it as its core count). */
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- FP64 instantiation ---------------------------------------------------------------------- */
#define REAL double
#define FN(name) name##_d
#define IS_F64 1
#define SQRT sqrt
#define SIN sin
#define COS cos
#include "vlm_oracle_impl.h"
#undef REAL
#undef FN
#undef IS_F64
#undef SQRT
#undef SIN
#undef COS
#undef ORC_FOURPI
#undef ORC_CUT_CROSS
#undef ORC_CUT_NORM
#undef ORC_EPS_F

/* ---- FP32 instantiation ---------------------------------------------------------------------- */
#define REAL float
#define FN(name) name##_f
#define IS_F64 0
#define SQRT sqrtf
#define SIN sinf
#define COS cosf
#include "vlm_oracle_impl.h"

/*
 * Oracle translation unit: TEST INFRASTRUCTURE ONLY (never linked into the
 * product library).  Instantiates oracle_impl.h twice:
 *   *_f64  FP64 oracle (libm double functions),
 *   *_f32  FP32 port of the reference algorithm (libm float functions),
 *          also the timed CPU baseline.
 * Built by oracle/Makefile into oracle/build/liboracle.so.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---- FP64 oracle ---- */
#define REAL double
#define SFX(name) name##_f64
#define EXP exp
#define SIN sin
#define COS cos
#define SQRT sqrt
#define FABS fabs
#define HYPOT hypot
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef EXP
#undef SIN
#undef COS
#undef SQRT
#undef FABS
#undef HYPOT

/* ---- FP32 port ---- */
#define REAL float
#define SFX(name) name##_f32
#define EXP expf
#define SIN sinf
#define COS cosf
#define SQRT sqrtf
#define FABS fabsf
#define HYPOT hypotf
#include "oracle_impl.h"

int oracle_version(void) {
    return 1;
}

/*
 * Parity oracle for the bml ELLPACK-R / ELLBLOCK thresholded multiply path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C (+OpenMP) restatement of the reference
 * algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product library never links it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no -march, so
 * that real-typed sums round exactly as the reference's do on x86-64).
 *
 * The typed body lives in orc_typed.inc and is instantiated once per
 * precision, as the reference compiles its *_typed.c files once per
 * precision (reference cmake/bmlAddTypedLibrary.cmake).
 */
#include "orc.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CAT3_(a, b, c) a##_##b##_##c
#define CAT3(a, b, c) CAT3_(a, b, c)

/* single real */
#define T float
#define SFX f32
#define ORC(name) CAT3(orc, name, SFX)
#define MAG(x) (fabsf(x))
#define ABOVE(x, t) (fabsf(x) > (float) (t))
#define REALP(x) (x)
#include "orc_typed.inc"
#undef T
#undef SFX
#undef ORC
#undef MAG
#undef ABOVE
#undef REALP

/* double real */
#define T double
#define SFX f64
#define ORC(name) CAT3(orc, name, SFX)
#define MAG(x) (fabs(x))
#define ABOVE(x, t) (fabs(x) > (double) (t))
#define REALP(x) (x)
#include "orc_typed.inc"
#undef T
#undef SFX
#undef ORC
#undef MAG
#undef ABOVE
#undef REALP

/* single complex */
#define T float _Complex
#define SFX c64
#define ORC(name) CAT3(orc, name, SFX)
#define MAG(x) (cabsf(x))
#define ABOVE(x, t) (cabsf(x) > (float) (t))
#define REALP(x) (crealf(x))
#include "orc_typed.inc"
#undef T
#undef SFX
#undef ORC
#undef MAG
#undef ABOVE
#undef REALP

/* double complex */
#define T double _Complex
#define SFX c128
#define ORC(name) CAT3(orc, name, SFX)
#define MAG(x) (cabs(x))
#define ABOVE(x, t) (cabs(x) > (double) (t))
#define REALP(x) (creal(x))
#include "orc_typed.inc"
#undef T
#undef SFX
#undef ORC
#undef MAG
#undef ABOVE
#undef REALP

int orc_version(void)
{
    return 1;
}

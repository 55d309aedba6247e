/*
 * Parity oracle (TEST INFRASTRUCTURE ONLY) -- see orc.c.
 *
 * Every function exists in four precisions, suffix _f32, _f64, _c64, _c128
 * (float, double, float complex, double complex); values are passed as
 * void pointers of that element type.  Return value 0 = success,
 * ORC_ERR_M_OVERFLOW = a row needed more than its M slots (the reference
 * exits the process there), ORC_ERR_ALLOC = out of host memory.
 */
#ifndef BML_ORACLE_H
#define BML_ORACLE_H

#define ORC_ERR_M_OVERFLOW 1
#define ORC_ERR_ALLOC 2

int orc_version(void);

#endif

/*
 * Host-side internals shared by the C files of the HIP backend: error
 * policy, synchronisation policy, per-matrix bookkeeping.
 */
#ifndef BML_HIP_HOST_H
#define BML_HIP_HOST_H

#include "../../include/bml_hip.h"
#include "bmlhip_abi.h"

#include <stddef.h>

/* Per-matrix bookkeeping kept on the host. */
struct bml_hip_aux {
    int stats_valid;       /* max_nnz/total_nnz describe the device state */
    int max_nnz;
    long long total_nnz;
    int touched_hint;      /* distinct columns per row seen the last time this
                              matrix was squared (0 = unknown) */
    int hint_max_nnz;      /* its longest row at that time */
};

/* Record an error.  In BML_HIP_ERRORS_EXIT mode prints it the way bml's
 * LOG_ERROR does and exits the process with EXIT_FAILURE. */
void bmlh_error(int code, const char *fmt, ...);
/* 1 if an error is pending (only possible in BML_HIP_ERRORS_RECORD mode). */
int bmlh_failed(void);
/* Make sure the device and the library stream exist; 0 on success. */
int bmlh_ready(void);
/* Check a bmlhip_* return code; reports the device-side message. */
int bmlh_ok(int rc, const char *what);
/* End of an operation: in BML_HIP_SYNC_EACH_CALL mode wait, read the status
 * block, raise M overflow as bml does, and return 1 with *st filled;
 * otherwise return 0 (nothing read). */
int bmlh_finish(const char *what, bmlhip_status *st);
/* Like bmlh_finish but always synchronises (for calls that return values). */
int bmlh_finish_sync(const char *what, bmlhip_status *st);
int bmlh_sync_mode(void);

size_t bmlh_elem_size(int precision);
/* block-product engine for products of A (BML_HIP_BLOCK_ENGINE_*, resolved) */
int bmlh_block_engine_for(const bml_matrix_ellblock_hip_t *A);
/* An ELLBLOCK work matrix with undefined tile values (indices and counts
 * zero), for drivers that write every tile before reading it. */
bml_matrix_ellblock_hip_t *bmlh_block_matrix_ellblock_work(int precision, int NB, int MB, int M,
                                                           const int *bsize, int distrib_mode);
/* the engine set by bml_hip_set_block_engine (BML_HIP_BLOCK_ENGINE_*) */
int bmlh_block_engine_setting(void);
/* the real scalar s as one element of precision prec, into buf (16 bytes) */
void bmlh_put_scalar(int prec, double s, void *buf);
struct bml_hip_aux *bmlh_aux_new(void);
void bmlh_set_stats(struct bml_hip_aux *aux, const bmlhip_status *st);
/* The row-major dense N x N host array (malloc'd) of an ellpack_hip or
 * ellblock_hip matrix A, its elements cast to precision; N into *N_out. */
void *bmlh_dense_as(const void *A, int precision, int *N_out, const char *what);

#endif

/*
 * ellpack_hip: bml's ELLPACK-R matrix type with its three arrays in HBM.
 *
 * C host side of the backend.  Each entry point mirrors the bml ELLPACK
 * backend function named in include/bml_hip.h (argument meaning, in-place
 * semantics, fatal errors) and queues gfx950 kernels through the thin C ABI
 * of bmlhip_abi.h.
 */
#include "bml_hip_host.h"
#include "bml_hip_tune.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- helpers ------------------------------------------------------------ */

static int valid_ell(const bml_matrix_ellpack_hip_t *A, const char *what)
{
    if (!A) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: matrix is NULL", what);
        return 0;
    }
    if (A->matrix_type != BML_HIP_TYPE_ELLPACK) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: not an ellpack_hip matrix (type %d)", what,
                   A->matrix_type);
        return 0;
    }
    if (bmlh_elem_size(A->matrix_precision) == 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown precision %d", what,
                   A->matrix_precision);
        return 0;
    }
    return 1;
}

static int same_shape(const bml_matrix_ellpack_hip_t *A, const bml_matrix_ellpack_hip_t *B,
                      const char *what)
{
    if (A->matrix_precision != B->matrix_precision) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: precisions differ (%d vs %d)", what,
                   A->matrix_precision, B->matrix_precision);
        return 0;
    }
    if (A->N != B->N) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: dimensions differ (%d vs %d)", what, A->N, B->N);
        return 0;
    }
    return 1;
}

static int max_nnz_of(const bml_matrix_ellpack_hip_t *A)
{
    return (A->aux && A->aux->stats_valid) ? A->aux->max_nnz : A->M;
}

static bml_matrix_ellpack_hip_t *alloc_ell(int prec, int N, int M, int mode, int zero,
                                           const char *what)
{
    if (bmlh_ready())
        return NULL;
    size_t es = bmlh_elem_size(prec);
    if (!es) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown precision %d", what, prec);
        return NULL;
    }
    if (N < 0 || M <= 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bad shape N=%d M=%d", what, N, M);
        return NULL;
    }
    /* the row kernels address slot i*M + jp in 32-bit unsigned arithmetic */
    if ((unsigned long long) N * (unsigned long long) M > (unsigned long long) UINT32_MAX) {
        bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: N*M = %llu slots exceeds 2^32 - 1", what,
                   (unsigned long long) N * (unsigned long long) M);
        return NULL;
    }
    bml_matrix_ellpack_hip_t *A = calloc(1, sizeof *A);
    if (!A) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    A->matrix_type = BML_HIP_TYPE_ELLPACK;
    A->matrix_precision = prec;
    A->distribution_mode = mode;
    A->N = N;
    A->M = M;
    A->aux = bmlh_aux_new();
    size_t nm = (size_t) N * (size_t) M;
    if (!bmlh_ok(bmlhip_malloc(&A->value, nm * es), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &A->index, nm * sizeof(int)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &A->nnz, (size_t) N * sizeof(int)), what)) {
        bml_deallocate_ellpack_hip(A);
        return NULL;
    }
    if (zero) {
        bmlhip_zero(A->value, nm * es);
        bmlhip_zero(A->index, nm * sizeof(int));
    }
    bmlhip_zero(A->nnz, (size_t) N * sizeof(int));
    A->aux->stats_valid = 1;
    return A;
}

static void finish_into(const char *what, bml_matrix_ellpack_hip_t *out)
{
    bmlhip_status st;
    if (bmlh_finish(what, &st))
        bmlh_set_stats(out->aux, &st);
    else
        bmlh_set_stats(out->aux, NULL);
}

/* ---- allocation and conversion -------------------------------------------- */

bml_matrix_ellpack_hip_t *bml_zero_matrix_ellpack_hip(int precision, int N, int M,
                                                      int distrib_mode)
{
    return alloc_ell(precision, N, M, distrib_mode, 1, "bml_zero_matrix_ellpack_hip");
}

bml_matrix_ellpack_hip_t *bml_noinit_matrix_ellpack_hip(int precision, int N, int M,
                                                        int distrib_mode)
{
    return alloc_ell(precision, N, M, distrib_mode, 0, "bml_noinit_matrix_ellpack_hip");
}

bml_matrix_ellpack_hip_t *bml_identity_matrix_ellpack_hip(int precision, int N, int M,
                                                          int distrib_mode)
{
    bml_matrix_ellpack_hip_t *A =
        alloc_ell(precision, N, M, distrib_mode, 1, "bml_identity_matrix_ellpack_hip");
    if (!A)
        return NULL;
    if (!bmlh_ok(bmlhip_identity(precision, N, M, A->value, A->index, A->nnz),
                 "bml_identity_matrix_ellpack_hip"))
        return A;
    A->aux->stats_valid = 1;
    A->aux->max_nnz = N > 0 ? 1 : 0;
    A->aux->total_nnz = N;
    return A;
}

void bml_deallocate_ellpack_hip(bml_matrix_ellpack_hip_t *A)
{
    if (!A)
        return;
    bmlhip_free(A->value);
    bmlhip_free(A->index);
    bmlhip_free(A->nnz);
    free(A->aux);
    free(A);
}

void bml_clear_ellpack_hip(bml_matrix_ellpack_hip_t *A)
{
    if (!valid_ell(A, "bml_clear_ellpack_hip"))
        return;
    bmlhip_zero(A->nnz, (size_t) A->N * sizeof(int));
    A->aux->stats_valid = 1;
    A->aux->max_nnz = 0;
    A->aux->total_nnz = 0;
}

void bml_copy_ellpack_hip(const bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B)
{
    const char *what = "bml_copy_ellpack_hip";
    if (!valid_ell(A, what) || !valid_ell(B, what) || !same_shape(A, B, what))
        return;
    if (A == B)
        return;
    size_t es = bmlh_elem_size(A->matrix_precision);
    if (A->M == B->M) {
        size_t nm = (size_t) A->N * A->M;
        if (!bmlh_ok(bmlhip_d2d(B->value, A->value, nm * es), what) ||
            !bmlh_ok(bmlhip_d2d(B->index, A->index, nm * sizeof(int)), what) ||
            !bmlh_ok(bmlhip_d2d(B->nnz, A->nnz, (size_t) A->N * sizeof(int)), what))
            return;
        *B->aux = *A->aux;
        bmlhip_status st;
        bmlh_finish(what, &st);
        return;
    }
    if (!bmlh_ok(bmlhip_copy_rows(A->matrix_precision, A->N, A->M, A->value, A->index, A->nnz,
                                  B->M, B->value, B->index, B->nnz), what))
        return;
    *B->aux = *A->aux;
    bmlhip_status st;
    bmlh_finish(what, &st);
}

bml_matrix_ellpack_hip_t *bml_copy_ellpack_hip_new(const bml_matrix_ellpack_hip_t *A)
{
    if (!valid_ell(A, "bml_copy_ellpack_hip_new"))
        return NULL;
    bml_matrix_ellpack_hip_t *B = bml_noinit_matrix_ellpack_hip(A->matrix_precision, A->N, A->M,
                                                                A->distribution_mode);
    if (B)
        bml_copy_ellpack_hip(A, B);
    return B;
}

void bml_ellpack_hip_set_arrays(bml_matrix_ellpack_hip_t *A, const void *value,
                                const int *index, const int *nnz, int src_M)
{
    const char *what = "bml_ellpack_hip_set_arrays";
    if (!valid_ell(A, what))
        return;
    const int N = A->N, M = A->M;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    if (src_M <= 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bad source stride %d", what, src_M);
        return;
    }
    char *v = malloc((size_t) N * M * es + 1);
    int *ix = calloc((size_t) N * M + 1, sizeof(int));
    unsigned char *seen = calloc((size_t) N + 1, 1);
    if (!v || !ix || !seen) {
        free(v);
        free(ix);
        free(seen);
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return;
    }
    memset(v, 0, (size_t) N * M * es + 1);
    int maxn = 0;
    long long tot = 0;
    for (int i = 0; i < N; i++) {
        const int n = nnz[i];
        if (n < 0 || n > M || n > src_M) {
            bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                       "%s: row %d has %d entries, M=%d: Number of non-zeroes per row > M",
                       what, i, n, M);
            goto out;
        }
        for (int k = 0; k < n; k++) {
            const int c = index[(size_t) i * src_M + k];
            if (c < 0 || c >= N) {
                bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: row %d column %d out of range", what, i, c);
                goto out;
            }
            if (seen[c]) {
                bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: row %d repeats column %d", what, i, c);
                goto out;
            }
            seen[c] = 1;
        }
        for (int k = 0; k < n; k++)
            seen[index[(size_t) i * src_M + k]] = 0;
        memcpy(v + (size_t) i * M * es, (const char *) value + (size_t) i * src_M * es,
               (size_t) n * es);
        memcpy(ix + (size_t) i * M, index + (size_t) i * src_M, (size_t) n * sizeof(int));
        if (n > maxn)
            maxn = n;
        tot += n;
    }
    if (bmlh_ok(bmlhip_h2d(A->value, v, (size_t) N * M * es), what) &&
        bmlh_ok(bmlhip_h2d(A->index, ix, (size_t) N * M * sizeof(int)), what) &&
        bmlh_ok(bmlhip_h2d(A->nnz, nnz, (size_t) N * sizeof(int)), what)) {
        A->aux->stats_valid = 1;
        A->aux->max_nnz = maxn;
        A->aux->total_nnz = tot;
        A->aux->touched_hint = 0;
    }
out:
    free(v);
    free(ix);
    free(seen);
}

void bml_ellpack_hip_get_arrays(const bml_matrix_ellpack_hip_t *A, void *value, int *index,
                                int *nnz)
{
    const char *what = "bml_ellpack_hip_get_arrays";
    if (!valid_ell(A, what))
        return;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return;
    const size_t nm = (size_t) A->N * A->M;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    if (value && !bmlh_ok(bmlhip_d2h(value, A->value, nm * es), what))
        return;
    if (index && !bmlh_ok(bmlhip_d2h(index, A->index, nm * sizeof(int)), what))
        return;
    if (nnz)
        bmlh_ok(bmlhip_d2h(nnz, A->nnz, (size_t) A->N * sizeof(int)), what);
}

bml_matrix_ellpack_hip_t *bml_import_from_dense_ellpack_hip(int precision, int order, int N,
                                                            const void *A, double threshold,
                                                            int M, int distrib_mode)
{
    const char *what = "bml_import_from_dense_ellpack_hip";
    size_t es = bmlh_elem_size(precision);
    if (!es) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown precision %d", what, precision);
        return NULL;
    }
    bml_matrix_ellpack_hip_t *B = bml_zero_matrix_ellpack_hip(precision, N, M, distrib_mode);
    if (!B)
        return NULL;
    char *v = calloc((size_t) N * M + 1, es);
    int *ix = calloc((size_t) N * M + 1, sizeof(int));
    int *nn = calloc((size_t) N + 1, sizeof(int));
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    for (int i = 0; i < N; i++) {
        int n = 0;
        for (int j = 0; j < N; j++) {
            size_t src = order == BML_HIP_DENSE_COLUMN_MAJOR ? (size_t) j * N + i
                                                             : (size_t) i * N + j;
            const char *e = (const char *) A + src * es;
            double mag;
            switch (precision) {
            case BML_HIP_SINGLE_REAL: mag = fabsf(*(const float *) e); break;
            case BML_HIP_DOUBLE_REAL: mag = fabs(*(const double *) e); break;
            case BML_HIP_SINGLE_COMPLEX:
                mag = hypotf(((const float *) e)[0], ((const float *) e)[1]);
                break;
            default: mag = hypot(((const double *) e)[0], ((const double *) e)[1]); break;
            }
            int keep = precision == BML_HIP_SINGLE_REAL || precision == BML_HIP_SINGLE_COMPLEX
                           ? (float) mag > (float) threshold
                           : mag > threshold;
            if (!keep)
                continue;
            if (n >= M) {
                bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                           "%s: row %d: Number of non-zeroes per row > M", what, i);
                goto out;
            }
            memcpy(v + ((size_t) i * M + n) * es, e, es);
            ix[(size_t) i * M + n] = j;
            n++;
        }
        nn[i] = n;
    }
    bml_ellpack_hip_set_arrays(B, v, ix, nn, M);
out:
    free(v);
    free(ix);
    free(nn);
    return B;
}

void *bml_export_to_dense_ellpack_hip(const bml_matrix_ellpack_hip_t *A, int order)
{
    const char *what = "bml_export_to_dense_ellpack_hip";
    if (!valid_ell(A, what))
        return NULL;
    const int N = A->N, M = A->M;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    char *v = malloc((size_t) N * M * es + 1);
    int *ix = malloc((size_t) N * M * sizeof(int) + 1);
    int *nn = malloc((size_t) N * sizeof(int) + 1);
    char *D = calloc((size_t) N * N + 1, es);
    if (!v || !ix || !nn || !D) {
        free(v);
        free(ix);
        free(nn);
        free(D);
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    bml_ellpack_hip_get_arrays(A, v, ix, nn);
    if (!bmlh_failed())
        for (int i = 0; i < N; i++)
            for (int k = 0; k < nn[i]; k++) {
                const int j = ix[(size_t) i * M + k];
                size_t dst = order == BML_HIP_DENSE_COLUMN_MAJOR ? (size_t) j * N + i
                                                                 : (size_t) i * N + j;
                memcpy(D + dst * es, v + ((size_t) i * M + k) * es, es);
            }
    free(v);
    free(ix);
    free(nn);
    return D;
}

void bml_ellpack_hip_row_stats(const bml_matrix_ellpack_hip_t *A, int *max_nnz,
                               long long *total_nnz)
{
    const char *what = "bml_ellpack_hip_row_stats";
    if (!valid_ell(A, what))
        return;
    if (!A->aux->stats_valid) {
        bmlhip_status st;
        if (!bmlh_ok(bmlhip_row_stats(A->N, A->nnz), what) || !bmlh_finish_sync(what, &st))
            return;
        bmlh_set_stats(A->aux, &st);
    }
    if (max_nnz)
        *max_nnz = A->aux->max_nnz;
    if (total_nnz)
        *total_nnz = A->aux->total_nnz;
}

/* ---- the multiply path ---------------------------------------------------- */

/* C = A*B into C's storage.  kind 0 = x2 (traces, kept-count check), 1 = AB
 * (touched-count check), 2 = AB with the kept-count check (adjust).  With
 * ab (alpha, beta): C <- beta*C + alpha*A*B on the exact engine, C merged in
 * the product's epilogue. */
static void apply_shape(bmlhip_spgemm_args *a, const bmlh_spgemm_shape *sh)
{
    a->group = sh->group;
    a->table = sh->table;
    a->bucket = sh->bucket;
    a->remap = sh->remap;
    a->waves = sh->waves;
    a->engine = sh->engine;
    a->wbits = sh->wbits;
    a->wrows = sh->wrows;
}

/* the autotuner's measurement: the product with shape sh, timed */
static int measure_product(const bmlh_spgemm_shape *sh, void *ctx, double *ms)
{
    bmlhip_spgemm_args a = *(const bmlhip_spgemm_args *) ctx;
    apply_shape(&a, sh);
    if (bmlhip_time_begin())
        return 1;
    const int rc = bmlhip_spgemm(&a);
    const int rt = bmlhip_time_end(ms);
    return rc || rt;
}

static int run_product_ws(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                          bml_matrix_ellpack_hip_t *C, double threshold, int kind,
                          struct bml_hip_aux *hint_owner, void *ws, size_t ws_bytes,
                          const double *ab, const char *what)
{
    bmlhip_spgemm_args a;
    memset(&a, 0, sizeof a);
    a.prec = A->matrix_precision;
    a.kind = kind;
    a.N = A->N;
    a.MA = A->M;
    a.MB = B->M;
    a.MC = C->M;
    a.d_Av = A->value;
    a.d_Ai = A->index;
    a.d_An = A->nnz;
    a.d_Bv = B->value;
    a.d_Bi = B->index;
    a.d_Bn = B->nnz;
    a.d_Cv = C->value;
    a.d_Ci = C->index;
    a.d_Cn = C->nnz;
    a.threshold = threshold;
    /* BML_HIP_ASSUME_SYMMETRIC=1: the caller asserts X is symmetric
     * (Hermitian) and x2 forms the upper triangle and mirrors it, as the SP2
     * drivers do for a symmetric H (a test hook for that path) */
    const char *e = getenv("BML_HIP_ASSUME_SYMMETRIC");
    a.sym = kind == 0 && !ws && !ab && e && atoi(e) != 0;
    a.d_ws = ws;
    a.ws_bytes = ws_bytes;
    if (ab) {
        a.fuse = 1;
        a.alpha = ab[0];
        a.beta = ab[1];
    }
    bmlh_spgemm_shape sh;
    /* on a workspace the window engine (squares only) is not offered; the
     * autotuner measures x2 and AB, whose runs overwrite C (not the fused
     * multiply, which reads C, nor the workspace and adjust forms) */
    const int tune = !ab && !ws && kind != 2;
    bmlh_tuned_spgemm(a.prec, A->N, max_nnz_of(A), max_nnz_of(B),
                      hint_owner ? hint_owner->touched_hint : 0,
                      hint_owner ? hint_owner->hint_max_nnz : 0,
                      (A == B && !ws ? BMLH_PICK_SQUARE : 0) | (ab ? BMLH_PICK_EXACT : 0), &sh,
                      tune ? measure_product : NULL, &a);
    apply_shape(&a, &sh);
    return bmlh_ok(bmlhip_spgemm(&a), what);
}

static int run_product(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                       bml_matrix_ellpack_hip_t *C, double threshold, int kind,
                       struct bml_hip_aux *hint_owner, const char *what)
{
    return run_product_ws(A, B, C, threshold, kind, hint_owner, NULL, 0, NULL, what);
}

void *bml_multiply_x2_ellpack_hip(bml_matrix_ellpack_hip_t *X, bml_matrix_ellpack_hip_t *X2,
                                  double threshold)
{
    const char *what = "bml_multiply_x2_ellpack_hip";
    if (!valid_ell(X, what) || !valid_ell(X2, what) || !same_shape(X, X2, what))
        return NULL;
    if (X == X2) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: X and X2 must be distinct matrices", what);
        return NULL;
    }
    if (!run_product(X, X, X2, threshold, 0, X->aux, what))
        return NULL;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return NULL;
    bmlh_set_stats(X2->aux, &st);
    X->aux->touched_hint = st.max_touched;
    X->aux->hint_max_nnz = max_nnz_of(X);
    double *trace = malloc(2 * sizeof(double));
    if (!trace) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    trace[0] = st.scalars[0];
    trace[1] = st.scalars[1];
    return trace;
}

void bml_multiply_x2_ellpack_hip_noalloc(bml_matrix_ellpack_hip_t *X,
                                         bml_matrix_ellpack_hip_t *X2, double threshold,
                                         double *traces)
{
    const char *what = "bml_multiply_x2_ellpack_hip_noalloc";
    if (!valid_ell(X, what) || !valid_ell(X2, what) || !same_shape(X, X2, what))
        return;
    if (X == X2 || !traces) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: X and X2 must be distinct and traces not NULL",
                   what);
        return;
    }
    size_t ws_bytes = 0;
    void *ws = bmlhip_noalloc_workspace(&ws_bytes);
    if (!ws) {
        bmlh_error(BML_HIP_ERR_DEVICE, "%s: the library is not initialised on this GPU", what);
        return;
    }
    if (!run_product_ws(X, X, X2, threshold, 0, X->aux, ws, ws_bytes, NULL, what))
        return;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return;
    bmlh_set_stats(X2->aux, &st);
    X->aux->touched_hint = st.max_touched;
    X->aux->hint_max_nnz = max_nnz_of(X);
    traces[0] = st.scalars[0];
    traces[1] = st.scalars[1];
}

void bml_multiply_AB_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                                 bml_matrix_ellpack_hip_t *C, double threshold)
{
    const char *what = "bml_multiply_AB_ellpack_hip";
    if (!valid_ell(A, what) || !valid_ell(B, what) || !valid_ell(C, what) ||
        !same_shape(A, B, what) || !same_shape(A, C, what))
        return;
    if (C == A || C == B) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: C must be distinct from A and B", what);
        return;
    }
    if (!run_product(A, B, C, threshold, 1, A == B ? A->aux : NULL, what))
        return;
    finish_into(what, C);
}

void bml_multiply_adjust_AB_ellpack_hip(bml_matrix_ellpack_hip_t *A,
                                        bml_matrix_ellpack_hip_t *B,
                                        bml_matrix_ellpack_hip_t *C, double threshold)
{
    const char *what = "bml_multiply_adjust_AB_ellpack_hip";
    if (!valid_ell(A, what) || !valid_ell(B, what) || !valid_ell(C, what) ||
        !same_shape(A, B, what) || !same_shape(A, C, what))
        return;
    if (C == A || C == B) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: C must be distinct from A and B", what);
        return;
    }
    /* the threshold is held in the matrix's element type, as the
     * reference's REAL_T adjust_threshold */
    const int single = A->matrix_precision == BML_HIP_SINGLE_REAL ||
                       A->matrix_precision == BML_HIP_SINGLE_COMPLEX;
    double t = single ? (double) (float) threshold : threshold;
    for (int pass = 0; pass < 2048; pass++) {
        if (!run_product(A, B, C, t, 2, NULL, what))
            return;
        bmlhip_status st;
        if (!bmlh_ok(bmlhip_status_fetch(&st), what))
            return;
        if (st.m_overflow_rows == 0) {
            bmlh_set_stats(C->aux, &st);
            return;
        }
        bmlhip_status_clear_errors();
        t = single ? (double) (float) (t * 2.0) : t * 2.0;
    }
    bmlh_error(BML_HIP_ERR_M_OVERFLOW, "%s: Number of non-zeroes per row > M", what);
}

static void add_impl(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B, double alpha,
                     double beta, double threshold, int mode, double *norm, const char *what)
{
    if (!valid_ell(A, what) || !valid_ell(B, what) || !same_shape(A, B, what))
        return;
    if (!bmlh_ok(bmlhip_add(A->matrix_precision, A->N, A->M, A->value, A->index, A->nnz, B->M,
                            B->value, B->index, B->nnz, alpha, beta, threshold, mode),
                 what))
        return;
    if (norm) {
        bmlhip_status st;
        if (!bmlh_finish_sync(what, &st))
            return;
        bmlh_set_stats(A->aux, &st);
        *norm = st.scalars[2];
    } else {
        finish_into(what, A);
    }
    A->aux->touched_hint = 0;
}

void bml_add_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B, double alpha,
                         double beta, double threshold)
{
    add_impl(A, B, alpha, beta, threshold, 0, NULL, "bml_add_ellpack_hip");
}

double bml_add_norm_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                                double alpha, double beta, double threshold)
{
    double n = 0.0;
    add_impl(A, B, alpha, beta, threshold, 1, &n, "bml_add_norm_ellpack_hip");
    return n;
}

/* add_norm whose norm is sum |A - B|^2 (the SP2 idempotency measure) */
double bmlh_add_frob_ellpack(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                             double alpha, double beta, double threshold)
{
    double n = 0.0;
    add_impl(A, B, alpha, beta, threshold, 2, &n, "bml_add_norm_ellpack_hip");
    return n;
}

void bml_multiply_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                              bml_matrix_ellpack_hip_t *C, double alpha, double beta,
                              double threshold)
{
    const char *what = "bml_multiply_ellpack_hip";
    if (!valid_ell(A, what) || !valid_ell(B, what) || !valid_ell(C, what) ||
        !same_shape(A, B, what) || !same_shape(A, C, what))
        return;
    if (A == B && alpha == 1.0 && beta == 0.0) {
        free(bml_multiply_x2_ellpack_hip(A, C, threshold));
        return;
    }
    if (C != A && C != B) {
        /* C <- beta*C + alpha*A*B in one pass: the product's epilogue merges
         * C's row as bml_add does, no temporary */
        const double ab[2] = {alpha, beta};
        if (run_product_ws(A, B, C, threshold, A == B ? 0 : 1, A == B ? A->aux : NULL, NULL, 0,
                           ab, what))
            finish_into(what, C);
        else
            bmlh_set_stats(C->aux, NULL);
        C->aux->touched_hint = 0;
        return;
    }
    /* C is A or B: the product into a temporary of C's shape, then
     * C <- beta*C + alpha*T */
    bml_matrix_ellpack_hip_t *T =
        bml_noinit_matrix_ellpack_hip(A->matrix_precision, A->N, C->M, A->distribution_mode);
    if (!T)
        return;
    if (A == B)
        free(bml_multiply_x2_ellpack_hip(A, T, threshold));
    else
        bml_multiply_AB_ellpack_hip(A, B, T, threshold);
    if (!bmlh_failed())
        bml_add_ellpack_hip(C, T, beta, alpha, threshold);
    bml_deallocate_ellpack_hip(T);
}

void bml_add_identity_ellpack_hip(bml_matrix_ellpack_hip_t *A, double beta, double threshold)
{
    const char *what = "bml_add_identity_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    if (!bmlh_ok(bmlhip_scale_add_identity(A->matrix_precision, A->N, A->M, A->value, A->index,
                                           A->nnz, NULL, beta, threshold, 1), what))
        return;
    finish_into(what, A);
    A->aux->stats_valid = 0;
}

void bml_scale_add_identity_ellpack_hip(bml_matrix_ellpack_hip_t *A, double alpha, double beta,
                                        double threshold)
{
    const char *what = "bml_scale_add_identity_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    unsigned char s[16];
    bmlh_put_scalar(A->matrix_precision, alpha, s);
    if (!bmlh_ok(bmlhip_scale_add_identity(A->matrix_precision, A->N, A->M, A->value, A->index,
                                           A->nnz, s, beta, threshold, 1), what))
        return;
    finish_into(what, A);
    A->aux->stats_valid = 0;
}

void bml_scale_inplace_ellpack_hip(const void *scale_factor, bml_matrix_ellpack_hip_t *A)
{
    const char *what = "bml_scale_inplace_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    if (!bmlh_ok(bmlhip_scale_add_identity(A->matrix_precision, A->N, A->M, A->value, A->index,
                                           A->nnz, scale_factor, 0.0, 0.0, 0), what))
        return;
    bmlhip_status st;
    bmlh_finish(what, &st);
}

void bml_threshold_ellpack_hip(bml_matrix_ellpack_hip_t *A, double threshold)
{
    const char *what = "bml_threshold_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    if (!bmlh_ok(bmlhip_threshold(A->matrix_precision, A->N, A->M, A->value, A->index, A->nnz,
                                  threshold), what))
        return;
    bmlhip_status st;
    bmlh_finish(what, &st);
    A->aux->stats_valid = 0;
    A->aux->touched_hint = 0;
}

bml_matrix_ellpack_hip_t *bml_threshold_new_ellpack_hip(const bml_matrix_ellpack_hip_t *A,
                                                        double threshold)
{
    bml_matrix_ellpack_hip_t *B = bml_copy_ellpack_hip_new(A);
    if (B)
        bml_threshold_ellpack_hip(B, threshold);
    return B;
}

double bml_trace_ellpack_hip(const bml_matrix_ellpack_hip_t *A)
{
    const char *what = "bml_trace_ellpack_hip";
    if (!valid_ell(A, what))
        return 0.0;
    if (!bmlh_ok(bmlhip_trace(A->matrix_precision, A->N, A->M, A->value, A->index, A->nnz),
                 what))
        return 0.0;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return 0.0;
    return st.scalars[0];
}

void *bml_gershgorin_ellpack_hip(const bml_matrix_ellpack_hip_t *A)
{
    const char *what = "bml_gershgorin_ellpack_hip";
    if (!valid_ell(A, what))
        return NULL;
    if (!bmlh_ok(bmlhip_gershgorin(A->matrix_precision, A->N, A->M, A->value, A->index, A->nnz),
                 what))
        return NULL;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return NULL;
    double *e = malloc(2 * sizeof(double));
    if (!e) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    e[0] = st.scalars[0];
    e[1] = st.scalars[1];
    return e;
}

void bml_normalize_ellpack_hip(bml_matrix_ellpack_hip_t *A, double mineval, double maxeval)
{
    const char *what = "bml_normalize_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    const double span = maxeval - mineval;
    unsigned char s[16];
    bmlh_put_scalar(A->matrix_precision, -1.0 / span, s);
    if (!bmlh_ok(bmlhip_scale_add_identity(A->matrix_precision, A->N, A->M, A->value, A->index,
                                           A->nnz, s, maxeval / span, 0.0, 1), what))
        return;
    finish_into(what, A);
    A->aux->stats_valid = 0;
}

/* ---- Matrix Market files -------------------------------------------------- */

/* bml_write_bml_matrix_ellpack (ellpack/bml_utilities_ellpack_typed.c:128-190):
 * a coordinate file, "%%%MatrixMarket matrix coordinate real general" (three
 * '%', as the reference prints it; "complex" for complex precisions), then
 * "N N nnz", then one line per stored entry in row order and each row's
 * storage order: 1-based row and column and the value as %20.15g (complex:
 * real and imaginary parts). */
void bml_write_bml_matrix_ellpack_hip(const bml_matrix_ellpack_hip_t *A, const char *filename)
{
    const char *what = "bml_write_bml_matrix_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    const int N = A->N, M = A->M, prec = A->matrix_precision;
    const size_t es = bmlh_elem_size(prec);
    char *v = malloc((size_t) N * M * es + 1);
    int *ix = malloc(((size_t) N * M + 1) * sizeof(int));
    int *nn = malloc(((size_t) N + 1) * sizeof(int));
    FILE *f = NULL;
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    bml_ellpack_hip_get_arrays(A, v, ix, nn);
    if (bmlh_failed())
        goto out;
    if (!(f = fopen(filename, "w"))) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: cannot open %s", what, filename);
        goto out;
    }
    const int cplx = prec == BML_HIP_SINGLE_COMPLEX || prec == BML_HIP_DOUBLE_COMPLEX;
    fprintf(f, "%%%%%%MatrixMarket matrix coordinate %s general\n", cplx ? "complex" : "real");
    long long tot = 0;
    for (int i = 0; i < N; i++)
        tot += nn[i];
    fprintf(f, "%d %d %lld\n", N, N, tot);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < nn[i]; k++) {
            const size_t at = (size_t) i * M + k;
            const int c = ix[at] + 1;
            switch (prec) {
            case BML_HIP_SINGLE_REAL:
                fprintf(f, "%d %d %20.15g\n", i + 1, c, (double) ((const float *) v)[at]);
                break;
            case BML_HIP_DOUBLE_REAL:
                fprintf(f, "%d %d %20.15g\n", i + 1, c, ((const double *) v)[at]);
                break;
            case BML_HIP_SINGLE_COMPLEX:
                fprintf(f, "%d %d %20.15g %20.15g\n", i + 1, c,
                        (double) ((const float *) v)[2 * at], (double) ((const float *) v)[2 * at + 1]);
                break;
            default:
                fprintf(f, "%d %d %20.15g %20.15g\n", i + 1, c, ((const double *) v)[2 * at],
                        ((const double *) v)[2 * at + 1]);
                break;
            }
        }
    if (fclose(f) != 0)
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: write to %s failed", what, filename);
    f = NULL;
out:
    if (f)
        fclose(f);
    free(v);
    free(ix);
    free(nn);
}

/* bml_read_bml_matrix_ellpack (ellpack/bml_utilities_ellpack_typed.c:22-120):
 * appends the file's entries to A's rows in file order (A is normally a zero
 * matrix); a "symmetric" file also stores each off-diagonal entry at its
 * mirror.  Values are read as %f (single real) or %lf and rounded to A's
 * precision.  Where the reference would overrun a row, this reports the
 * M overflow. */
void bml_read_bml_matrix_ellpack_hip(bml_matrix_ellpack_hip_t *A, const char *filename)
{
    const char *what = "bml_read_bml_matrix_ellpack_hip";
    if (!valid_ell(A, what))
        return;
    const int N = A->N, M = A->M, prec = A->matrix_precision;
    const size_t es = bmlh_elem_size(prec);
    const int cplx = prec == BML_HIP_SINGLE_COMPLEX || prec == BML_HIP_DOUBLE_COMPLEX;
    char *v = malloc((size_t) N * M * es + 1);
    int *ix = malloc(((size_t) N * M + 1) * sizeof(int));
    int *nn = malloc(((size_t) N + 1) * sizeof(int));
    FILE *f = NULL;
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    bml_ellpack_hip_get_arrays(A, v, ix, nn);
    if (bmlh_failed())
        goto out;
    if (!(f = fopen(filename, "r"))) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: cannot open %s", what, filename);
        goto out;
    }
    char h[5][64];
    int n1, n2, nz;
    if (fscanf(f, "%63s %63s %63s %63s %63s", h[0], h[1], h[2], h[3], h[4]) != 5 ||
        fscanf(f, "%d %d %d", &n1, &n2, &nz) != 3) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: %s: read error on the header", what, filename);
        goto out;
    }
    if (n1 > N || n2 > N || nz < 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: %s is %d x %d, the matrix %d x %d", what,
                   filename, n1, n2, N, N);
        goto out;
    }
    const int sym = strcmp(h[4], "symmetric") == 0;
    for (int e = 0; e < nz; e++) {
        int r, c, got;
        double re = 0.0, im = 0.0;
        float fv = 0.0f;
        if (prec == BML_HIP_SINGLE_REAL)
            got = fscanf(f, "%d %d %f", &r, &c, &fv) == 3;
        else if (!cplx)
            got = fscanf(f, "%d %d %lf", &r, &c, &re) == 3;
        else
            got = fscanf(f, "%d %d %lf %lf", &r, &c, &re, &im) == 4;
        if (!got) {
            bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: %s: read error at entry %d", what, filename, e);
            goto out;
        }
        r--;
        c--;
        if (r < 0 || r >= N || c < 0 || c >= N) {
            bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: %s: entry %d (%d, %d) out of range", what,
                       filename, e, r + 1, c + 1);
            goto out;
        }
        for (int pass = 0; pass < (sym && c != r ? 2 : 1); pass++) {
            const int row = pass ? c : r, col = pass ? r : c;
            if (nn[row] >= M) {
                bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                           "%s: row %d: Number of non-zeroes per row > M", what, row);
                goto out;
            }
            const size_t at = (size_t) row * M + nn[row]++;
            ix[at] = col;
            switch (prec) {
            case BML_HIP_SINGLE_REAL: ((float *) v)[at] = fv; break;
            case BML_HIP_DOUBLE_REAL: ((double *) v)[at] = re; break;
            case BML_HIP_SINGLE_COMPLEX:
                ((float *) v)[2 * at] = (float) re;
                ((float *) v)[2 * at + 1] = (float) im;
                break;
            default:
                ((double *) v)[2 * at] = re;
                ((double *) v)[2 * at + 1] = im;
                break;
            }
        }
    }
    bml_ellpack_hip_set_arrays(A, v, ix, nn, M);
out:
    if (f)
        fclose(f);
    free(v);
    free(ix);
    free(nn);
}

/* ---- bml_convert ------------------------------------------------------------ */

/* One element of precision pin at src, as an element of precision pout at dst:
 * float <-> double as C converts them, a real value gets a zero imaginary
 * part, and a complex value keeps its real part (C's complex-to-real
 * conversion). */
static void cast_elem(int pin, const void *src, int pout, void *dst)
{
    double re, im = 0.0;
    switch (pin) {
    case BML_HIP_SINGLE_REAL: re = *(const float *) src; break;
    case BML_HIP_DOUBLE_REAL: re = *(const double *) src; break;
    case BML_HIP_SINGLE_COMPLEX:
        re = ((const float *) src)[0];
        im = ((const float *) src)[1];
        break;
    default:
        re = ((const double *) src)[0];
        im = ((const double *) src)[1];
        break;
    }
    switch (pout) {
    case BML_HIP_SINGLE_REAL: *(float *) dst = (float) re; break;
    case BML_HIP_DOUBLE_REAL: *(double *) dst = re; break;
    case BML_HIP_SINGLE_COMPLEX:
        ((float *) dst)[0] = (float) re;
        ((float *) dst)[1] = (float) im;
        break;
    default:
        ((double *) dst)[0] = re;
        ((double *) dst)[1] = im;
        break;
    }
}

void *bmlh_dense_as(const void *A, int precision, int *N_out, const char *what)
{
    if (!A) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: matrix is NULL", what);
        return NULL;
    }
    const size_t es = bmlh_elem_size(precision);
    if (!es) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown precision %d", what, precision);
        return NULL;
    }
    const int type = *(const int *) A;
    int N, pin;
    void *D;
    if (type == BML_HIP_TYPE_ELLPACK) {
        const bml_matrix_ellpack_hip_t *E = A;
        N = E->N;
        pin = E->matrix_precision;
        D = bml_export_to_dense_ellpack_hip(E, BML_HIP_DENSE_ROW_MAJOR);
    } else if (type == BML_HIP_TYPE_ELLBLOCK) {
        const bml_matrix_ellblock_hip_t *E = A;
        N = E->N;
        pin = E->matrix_precision;
        D = bml_export_to_dense_ellblock_hip(E, BML_HIP_DENSE_ROW_MAJOR);
    } else {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown matrix type %d", what, type);
        return NULL;
    }
    if (!D || bmlh_failed()) {
        free(D);
        return NULL;
    }
    if (pin != precision) {
        const size_t nn = (size_t) N * N, ein = bmlh_elem_size(pin);
        char *C = malloc(nn * es + 1);
        if (!C) {
            free(D);
            bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
            return NULL;
        }
        for (size_t k = 0; k < nn; k++)
            cast_elem(pin, (const char *) D + k * ein, precision, C + k * es);
        free(D);
        D = C;
    }
    *N_out = N;
    return D;
}

bml_matrix_ellpack_hip_t *bml_convert_ellpack_hip(const void *A, int precision, int M,
                                                  int distrib_mode)
{
    const char *what = "bml_convert_ellpack_hip";
    int N = 0;
    char *D = bmlh_dense_as(A, precision, &N, what);
    if (!D)
        return NULL;
    /* every row holds all N columns, zeros included (bml_set_element per
     * element), so M must hold N */
    if (N > M) {
        free(D);
        bmlh_error(BML_HIP_ERR_M_OVERFLOW, "%s: row 0: Number of non-zeroes per row > M", what);
        return NULL;
    }
    const size_t es = bmlh_elem_size(precision);
    bml_matrix_ellpack_hip_t *B = bml_zero_matrix_ellpack_hip(precision, N, M, distrib_mode);
    char *v = B ? calloc((size_t) N * M + 1, es) : NULL;
    int *ix = B ? calloc((size_t) N * M + 1, sizeof(int)) : NULL;
    int *nn = B ? calloc((size_t) N + 1, sizeof(int)) : NULL;
    if (B && (!v || !ix || !nn)) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        bml_deallocate_ellpack_hip(B);
        B = NULL;
    }
    if (B) {
        for (int i = 0; i < N; i++) {
            memcpy(v + (size_t) i * M * es, D + (size_t) i * N * es, (size_t) N * es);
            for (int j = 0; j < N; j++)
                ix[(size_t) i * M + j] = j;
            nn[i] = N;
        }
        bml_ellpack_hip_set_arrays(B, v, ix, nn, M);
    }
    free(v);
    free(ix);
    free(nn);
    free(D);
    return B;
}

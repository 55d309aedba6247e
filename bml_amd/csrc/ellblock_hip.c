/*
 * ellblock_hip: bml's ELLBLOCK matrix type with its blocks in HBM.
 *
 * C host side of the backend.  Each entry point mirrors the bml ELLBLOCK
 * backend function named in include/bml_hip.h (argument meaning, in-place
 * semantics, fatal errors) and queues gfx950 kernels through the C ABI of
 * bmlhip_abi.h.  Blocks are stored as bt x bt tiles (bt = 8, 16 or 32),
 * zero-padded, one per block slot.
 */
#include "bml_hip_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- block-size registry (ellblock/bml_allocate_ellblock.c:14-91) -------- */

static int *s_bsize = NULL;
static int s_nb = 0, s_mb = 0;
static const int s_default_block_dim = 4;
static int s_engine = BML_HIP_BLOCK_ENGINE_AUTO;

void bml_hip_set_block_sizes(const int *bsize, int nb, int mb)
{
    const char *what = "bml_hip_set_block_sizes";
    if (nb <= 0 || mb <= 0 || !bsize) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: nb=%d mb=%d", what, nb, mb);
        return;
    }
    if (s_nb != 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: block sizes are set once only", what);
        return;
    }
    int *b = malloc(sizeof(int) * (size_t) nb);
    if (!b) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return;
    }
    memcpy(b, bsize, sizeof(int) * (size_t) nb);
    s_bsize = b;
    s_nb = nb;
    s_mb = mb;
}

const int *bml_hip_get_block_sizes(int N, int M)
{
    if (!s_bsize) {
        if (N <= 0 || M <= 0) {
            bmlh_error(BML_HIP_ERR_ARGUMENT, "bml_hip_get_block_sizes: N=%d M=%d", N, M);
            return NULL;
        }
        int nb = (N + s_default_block_dim - 1) / s_default_block_dim;
        int mb = (M + s_default_block_dim - 1) / s_default_block_dim;
        if (mb < 4)
            mb = 4;
        if (mb > nb)
            mb = nb;
        int *b = malloc(sizeof(int) * (size_t) nb);
        if (!b) {
            bmlh_error(BML_HIP_ERR_ALLOC, "bml_hip_get_block_sizes: host allocation failed");
            return NULL;
        }
        for (int ib = 0; ib < nb - 1; ib++)
            b[ib] = s_default_block_dim;
        b[nb - 1] = N - s_default_block_dim * (nb - 1);
        s_bsize = b;
        s_nb = nb;
        s_mb = mb;
    }
    return s_bsize;
}

int bml_hip_get_nb(void)
{
    return s_nb;
}

int bml_hip_get_mb(void)
{
    return s_mb;
}

void bml_hip_set_block_engine(int engine)
{
    if (engine < BML_HIP_BLOCK_ENGINE_AUTO || engine > BML_HIP_BLOCK_ENGINE_MFMA_BF16) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "bml_hip_set_block_engine: unknown engine %d", engine);
        return;
    }
    s_engine = engine;
}

int bmlh_block_engine_setting(void)
{
    return s_engine;
}

/* ---- helpers ------------------------------------------------------------ */

static int valid_eb(const bml_matrix_ellblock_hip_t *A, const char *what)
{
    if (!A) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: matrix is NULL", what);
        return 0;
    }
    if (A->matrix_type != BML_HIP_TYPE_ELLBLOCK) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: not an ellblock_hip matrix (type %d)", what,
                   A->matrix_type);
        return 0;
    }
    return 1;
}

static int same_blocks(const bml_matrix_ellblock_hip_t *A, const bml_matrix_ellblock_hip_t *B,
                       const char *what)
{
    if (A->matrix_precision != B->matrix_precision) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: precisions differ (%d vs %d)", what,
                   A->matrix_precision, B->matrix_precision);
        return 0;
    }
    if (A->NB != B->NB || A->bt != B->bt ||
        memcmp(A->bsize_host, B->bsize_host, sizeof(int) * (size_t) A->NB) != 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: block structures differ", what);
        return 0;
    }
    return 1;
}

static size_t tile_bytes(const bml_matrix_ellblock_hip_t *A)
{
    return (size_t) A->bt * A->bt * bmlh_elem_size(A->matrix_precision);
}

/* engine for a product: AUTO is MFMA on 32x32 tiles of real precision */
int bmlh_block_engine_for(const bml_matrix_ellblock_hip_t *A)
{
    const int real = A->matrix_precision == BML_HIP_SINGLE_REAL ||
                     A->matrix_precision == BML_HIP_DOUBLE_REAL;
    int e = s_engine;
    const char *env = getenv("BML_HIP_BLOCK_ENGINE");
    if (env)
        e = atoi(env);
    if (e == BML_HIP_BLOCK_ENGINE_AUTO)
        e = (real && A->bt == 32) ? BML_HIP_BLOCK_ENGINE_MFMA : BML_HIP_BLOCK_ENGINE_EXACT;
    if (e == BML_HIP_BLOCK_ENGINE_MFMA && !(real && A->bt == 32))
        e = BML_HIP_BLOCK_ENGINE_EXACT;
    /* bf16 inputs apply to single-precision products; double-precision
     * products (the refinement of a mixed solve) stay on FP64 MFMA */
    if (e == BML_HIP_BLOCK_ENGINE_MFMA_BF16 && A->matrix_precision != BML_HIP_SINGLE_REAL)
        e = (real && A->bt == 32) ? BML_HIP_BLOCK_ENGINE_MFMA : BML_HIP_BLOCK_ENGINE_EXACT;
    if (e == BML_HIP_BLOCK_ENGINE_MFMA_BF16 && A->bt != 32)
        e = BML_HIP_BLOCK_ENGINE_EXACT;
    return e;
}

static void eb_finish_into(const char *what, bml_matrix_ellblock_hip_t *out)
{
    bmlhip_status st;
    if (bmlh_finish(what, &st))
        bmlh_set_stats(out->aux, &st);
    else
        bmlh_set_stats(out->aux, NULL);
}

/* ---- allocation ------------------------------------------------------------ */

static bml_matrix_ellblock_hip_t *eb_alloc(int precision, int NB, int MB, int M,
                                           const int *bsize, int distrib_mode, int zero_values)
{
    const char *what = "bml_block_matrix_ellblock_hip";
    if (bmlh_ready())
        return NULL;
    const size_t es = bmlh_elem_size(precision);
    if (!es) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: unknown precision %d", what, precision);
        return NULL;
    }
    if (NB <= 0 || MB <= 0 || !bsize) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bad shape NB=%d MB=%d", what, NB, MB);
        return NULL;
    }
    int N = 0, bmax = 0;
    for (int ib = 0; ib < NB; ib++) {
        if (bsize[ib] <= 0 || bsize[ib] > 32) {
            bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: block size %d (block row %d) not in 1..32",
                       what, bsize[ib], ib);
            return NULL;
        }
        N += bsize[ib];
        if (bsize[ib] > bmax)
            bmax = bsize[ib];
    }
    bml_matrix_ellblock_hip_t *A = calloc(1, sizeof *A);
    if (!A) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    A->matrix_type = BML_HIP_TYPE_ELLBLOCK;
    A->matrix_precision = precision;
    A->distribution_mode = distrib_mode;
    A->N = N;
    A->M = M;
    A->NB = NB;
    A->MB = MB;
    A->bt = bmax <= 8 ? 8 : bmax <= 16 ? 16 : 32;
    A->aux = bmlh_aux_new();
    A->bsize_host = malloc(sizeof(int) * (size_t) NB);
    if (!A->aux || !A->bsize_host) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        bml_deallocate_ellblock_hip(A);
        return NULL;
    }
    memcpy(A->bsize_host, bsize, sizeof(int) * (size_t) NB);
    const size_t slots = (size_t) NB * MB;
    if (!bmlh_ok(bmlhip_malloc(&A->value, slots * tile_bytes(A)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &A->indexb, slots * sizeof(int)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &A->nnzb, (size_t) NB * sizeof(int)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &A->bsize, (size_t) NB * sizeof(int)), what) ||
        !bmlh_ok(bmlhip_h2d(A->bsize, bsize, (size_t) NB * sizeof(int)), what)) {
        bml_deallocate_ellblock_hip(A);
        return NULL;
    }
    if (zero_values)
        bmlhip_zero(A->value, slots * tile_bytes(A));
    bmlhip_zero(A->indexb, slots * sizeof(int));
    bmlhip_zero(A->nnzb, (size_t) NB * sizeof(int));
    A->aux->stats_valid = 1;
    if (s_mb == 0)
        bml_hip_set_block_sizes(bsize, NB, MB);
    return A;
}

bml_matrix_ellblock_hip_t *bml_block_matrix_ellblock_hip(int precision, int NB, int MB, int M,
                                                         const int *bsize, int distrib_mode)
{
    return eb_alloc(precision, NB, MB, M, bsize, distrib_mode, 1);
}

/* A work matrix whose tile values start undefined (block counts zero): for
 * drivers that write every tile they later read, so the NB*MB tiles of a
 * wide matrix are not cleared on every call */
bml_matrix_ellblock_hip_t *bmlh_block_matrix_ellblock_work(int precision, int NB, int MB, int M,
                                                           const int *bsize, int distrib_mode)
{
    return eb_alloc(precision, NB, MB, M, bsize, distrib_mode, 0);
}

bml_matrix_ellblock_hip_t *bml_zero_matrix_ellblock_hip(int precision, int N, int M,
                                                        int distrib_mode)
{
    const int *bsize = bml_hip_get_block_sizes(N, M);
    if (!bsize)
        return NULL;
    return bml_block_matrix_ellblock_hip(precision, s_nb, s_mb, M, bsize, distrib_mode);
}

void bml_deallocate_ellblock_hip(bml_matrix_ellblock_hip_t *A)
{
    if (!A)
        return;
    bmlhip_free(A->value);
    bmlhip_free(A->indexb);
    bmlhip_free(A->nnzb);
    bmlhip_free(A->bsize);
    free(A->bsize_host);
    free(A->aux);
    free(A);
}

void bml_clear_ellblock_hip(bml_matrix_ellblock_hip_t *A)
{
    if (!valid_eb(A, "bml_clear_ellblock_hip"))
        return;
    bmlhip_zero(A->nnzb, (size_t) A->NB * sizeof(int));
    A->aux->stats_valid = 1;
    A->aux->max_nnz = 0;
    A->aux->total_nnz = 0;
}

void bml_copy_ellblock_hip(const bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B)
{
    const char *what = "bml_copy_ellblock_hip";
    if (!valid_eb(A, what) || !valid_eb(B, what) || !same_blocks(A, B, what) || A == B)
        return;
    if (!bmlh_ok(bmlhip_eb_copy(A->matrix_precision, A->NB, A->bt, A->MB, A->value, A->indexb,
                                A->nnzb, B->MB, B->value, B->indexb, B->nnzb), what))
        return;
    *B->aux = *A->aux;
    bmlhip_status st;
    bmlh_finish(what, &st);
}

bml_matrix_ellblock_hip_t *bml_copy_ellblock_hip_new(const bml_matrix_ellblock_hip_t *A)
{
    if (!valid_eb(A, "bml_copy_ellblock_hip_new"))
        return NULL;
    bml_matrix_ellblock_hip_t *B = bml_block_matrix_ellblock_hip(
        A->matrix_precision, A->NB, A->MB, A->M, A->bsize_host, A->distribution_mode);
    if (B)
        bml_copy_ellblock_hip(A, B);
    return B;
}

void bml_ellblock_hip_set_blocks(bml_matrix_ellblock_hip_t *A, const void *value,
                                 const int *indexb, const int *nnzb, int src_MB, int bmax)
{
    const char *what = "bml_ellblock_hip_set_blocks";
    if (!valid_eb(A, what))
        return;
    const int NB = A->NB, MB = A->MB, bt = A->bt;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    int maxb = 0;
    for (int ib = 0; ib < NB; ib++)
        if (A->bsize_host[ib] > maxb)
            maxb = A->bsize_host[ib];
    if (src_MB <= 0 || bmax < maxb) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bad source layout MB=%d bmax=%d", what, src_MB,
                   bmax);
        return;
    }
    const size_t tb = (size_t) bt * bt * es;
    char *v = calloc((size_t) NB * MB, tb);
    int *ix = calloc((size_t) NB * MB + 1, sizeof(int));
    unsigned char *seen = calloc((size_t) NB + 1, 1);
    int maxn = 0;
    long long tot = 0;
    if (!v || !ix || !seen) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    for (int ib = 0; ib < NB; ib++) {
        const int n = nnzb[ib];
        if (n < 0 || n > MB || n > src_MB) {
            bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                       "%s: block row %d has %d blocks, MB=%d: Number of non-zeroes per row > M",
                       what, ib, n, MB);
            goto out;
        }
        for (int jp = 0; jp < n; jp++) {
            const int jb = indexb[(size_t) ib * src_MB + jp];
            if (jb < 0 || jb >= NB || seen[jb]) {
                bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: block row %d: block column %d %s", what,
                           ib, jb, (jb < 0 || jb >= NB) ? "out of range" : "repeated");
                goto out;
            }
            seen[jb] = 1;
            ix[(size_t) ib * MB + jp] = jb;
            const char *src = (const char *) value + ((size_t) ib * src_MB + jp) * bmax * bmax * es;
            char *dst = v + ((size_t) ib * MB + jp) * tb;
            for (int ii = 0; ii < A->bsize_host[ib]; ii++)
                memcpy(dst + (size_t) ii * bt * es, src + (size_t) ii * bmax * es,
                       (size_t) A->bsize_host[jb] * es);
        }
        for (int jp = 0; jp < n; jp++)
            seen[indexb[(size_t) ib * src_MB + jp]] = 0;
        if (n > maxn)
            maxn = n;
        tot += n;
    }
    if (bmlh_ok(bmlhip_h2d(A->value, v, (size_t) NB * MB * tb), what) &&
        bmlh_ok(bmlhip_h2d(A->indexb, ix, (size_t) NB * MB * sizeof(int)), what) &&
        bmlh_ok(bmlhip_h2d(A->nnzb, nnzb, (size_t) NB * sizeof(int)), what)) {
        A->aux->stats_valid = 1;
        A->aux->max_nnz = maxn;
        A->aux->total_nnz = tot;
    }
out:
    free(v);
    free(ix);
    free(seen);
}

void bml_ellblock_hip_get_blocks(const bml_matrix_ellblock_hip_t *A, void *value, int *indexb,
                                 int *nnzb, int bmax)
{
    const char *what = "bml_ellblock_hip_get_blocks";
    if (!valid_eb(A, what))
        return;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return;
    const int NB = A->NB, MB = A->MB, bt = A->bt;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    const size_t tb = (size_t) bt * bt * es;
    if (bmax < bt) {
        for (int ib = 0; ib < NB; ib++)
            if (A->bsize_host[ib] > bmax) {
                bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bmax %d below block size %d", what, bmax,
                           A->bsize_host[ib]);
                return;
            }
    }
    char *v = malloc((size_t) NB * MB * tb + 1);
    int *ix = malloc(sizeof(int) * ((size_t) NB * MB + 1));
    int *nn = malloc(sizeof(int) * ((size_t) NB + 1));
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    if (!bmlh_ok(bmlhip_d2h(v, A->value, (size_t) NB * MB * tb), what) ||
        !bmlh_ok(bmlhip_d2h(ix, A->indexb, (size_t) NB * MB * sizeof(int)), what) ||
        !bmlh_ok(bmlhip_d2h(nn, A->nnzb, (size_t) NB * sizeof(int)), what))
        goto out;
    if (nnzb)
        memcpy(nnzb, nn, sizeof(int) * (size_t) NB);
    if (indexb)
        memcpy(indexb, ix, sizeof(int) * (size_t) NB * MB);
    if (value) {
        memset(value, 0, (size_t) NB * MB * bmax * bmax * es);
        for (int ib = 0; ib < NB; ib++)
            for (int jp = 0; jp < nn[ib]; jp++) {
                const int jb = ix[(size_t) ib * MB + jp];
                const char *src = v + ((size_t) ib * MB + jp) * tb;
                char *dst = (char *) value + ((size_t) ib * MB + jp) * bmax * bmax * es;
                for (int ii = 0; ii < A->bsize_host[ib]; ii++)
                    memcpy(dst + (size_t) ii * bmax * es, src + (size_t) ii * bt * es,
                           (size_t) A->bsize_host[jb] * es);
            }
    }
out:
    free(v);
    free(ix);
    free(nn);
}

void *bml_export_to_dense_ellblock_hip(const bml_matrix_ellblock_hip_t *A, int order)
{
    const char *what = "bml_export_to_dense_ellblock_hip";
    if (!valid_eb(A, what))
        return NULL;
    const int NB = A->NB, MB = A->MB, bt = A->bt, N = A->N;
    const size_t es = bmlh_elem_size(A->matrix_precision);
    char *v = malloc((size_t) NB * MB * bt * bt * es + 1);
    int *ix = malloc(sizeof(int) * ((size_t) NB * MB + 1));
    int *nn = malloc(sizeof(int) * ((size_t) NB + 1));
    int *offs = malloc(sizeof(int) * ((size_t) NB + 1));
    char *D = calloc((size_t) N * N + 1, es);
    if (!v || !ix || !nn || !offs || !D) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        free(D);
        D = NULL;
        goto out;
    }
    bml_ellblock_hip_get_blocks(A, v, ix, nn, bt);
    if (bmlh_failed()) {
        free(D);
        D = NULL;
        goto out;
    }
    offs[0] = 0;
    for (int ib = 0; ib < NB; ib++)
        offs[ib + 1] = offs[ib] + A->bsize_host[ib];
    for (int ib = 0; ib < NB; ib++)
        for (int jp = 0; jp < nn[ib]; jp++) {
            const int jb = ix[(size_t) ib * MB + jp];
            const char *t = v + ((size_t) ib * MB + jp) * bt * bt * es;
            for (int ii = 0; ii < A->bsize_host[ib]; ii++)
                for (int jj = 0; jj < A->bsize_host[jb]; jj++) {
                    const size_t r = (size_t) offs[ib] + ii, c = (size_t) offs[jb] + jj;
                    const size_t at = order == BML_HIP_DENSE_COLUMN_MAJOR ? c * N + r : r * N + c;
                    memcpy(D + at * es, t + ((size_t) ii * bt + jj) * es, es);
                }
        }
out:
    free(v);
    free(ix);
    free(nn);
    free(offs);
    return D;
}

/* ---- the multiply path ---------------------------------------------------- */

/* C = A*B blockwise; kind 0 = x2 (traces), 1 = AB */
static int run_eb_product(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                          bml_matrix_ellblock_hip_t *C, double threshold, int kind,
                          const char *what)
{
    bmlhip_eb_args a;
    memset(&a, 0, sizeof a);
    a.prec = A->matrix_precision;
    a.kind = kind;
    a.NB = A->NB;
    a.bt = A->bt;
    a.d_bsize = A->bsize;
    a.MA = A->MB;
    a.MB = B->MB;
    a.MC = C->MB;
    a.d_Av = A->value;
    a.d_Ai = A->indexb;
    a.d_An = A->nnzb;
    a.d_Bv = B->value;
    a.d_Bi = B->indexb;
    a.d_Bn = B->nnzb;
    a.d_Cv = C->value;
    a.d_Ci = C->indexb;
    a.d_Cn = C->nnzb;
    a.threshold = threshold;
    a.engine = bmlh_block_engine_for(A);
    /* BML_HIP_ASSUME_SYMMETRIC=1: the caller asserts X is symmetric
     * (Hermitian) and x2 on an MFMA engine forms the blocks kb >= ib and
     * mirrors them, as the SP2 driver does for a symmetric H (a test hook for
     * that path) */
    const char *e = getenv("BML_HIP_ASSUME_SYMMETRIC");
    a.sym = kind == 0 && e && atoi(e) != 0;
    return bmlh_ok(bmlhip_eb_product(&a), what);
}

void *bml_multiply_x2_ellblock_hip(bml_matrix_ellblock_hip_t *X, bml_matrix_ellblock_hip_t *X2,
                                   double threshold)
{
    const char *what = "bml_multiply_x2_ellblock_hip";
    if (!valid_eb(X, what) || !valid_eb(X2, what) || !same_blocks(X, X2, what))
        return NULL;
    if (X == X2) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: X and X2 must be distinct matrices", what);
        return NULL;
    }
    if (!run_eb_product(X, X, X2, threshold, 0, what))
        return NULL;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return NULL;
    bmlh_set_stats(X2->aux, &st);
    double *trace = malloc(2 * sizeof(double));
    if (!trace) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        return NULL;
    }
    trace[0] = st.scalars[0];
    trace[1] = st.scalars[1];
    return trace;
}

void bml_multiply_AB_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                                  bml_matrix_ellblock_hip_t *C, double threshold)
{
    const char *what = "bml_multiply_AB_ellblock_hip";
    if (!valid_eb(A, what) || !valid_eb(B, what) || !valid_eb(C, what) ||
        !same_blocks(A, B, what) || !same_blocks(A, C, what))
        return;
    if (C == A || C == B) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: C must be distinct from A and B", what);
        return;
    }
    if (!run_eb_product(A, B, C, threshold, 1, what))
        return;
    eb_finish_into(what, C);
}

static void eb_add_impl(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                        double alpha, double beta, double threshold, int mode, double *norm,
                        const char *what)
{
    if (!valid_eb(A, what) || !valid_eb(B, what) || !same_blocks(A, B, what))
        return;
    if (!bmlh_ok(bmlhip_eb_add(A->matrix_precision, A->NB, A->bt, A->bsize, A->MB, A->value,
                               A->indexb, A->nnzb, B->MB, B->value, B->indexb, B->nnzb, alpha,
                               beta, threshold, mode, NULL, NULL),
                 what))
        return;
    if (norm) {
        bmlhip_status st;
        if (!bmlh_finish_sync(what, &st))
            return;
        bmlh_set_stats(A->aux, &st);
        *norm = st.scalars[2];
    } else {
        eb_finish_into(what, A);
    }
}

void bml_add_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                          double alpha, double beta, double threshold)
{
    eb_add_impl(A, B, alpha, beta, threshold, 0, NULL, "bml_add_ellblock_hip");
}

/* add whose value is sum |A - B|^2 over the union (the SP2 idempotency
 * measure) */
double bmlh_add_frob_ellblock(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                              double alpha, double beta, double threshold)
{
    double n = 0.0;
    eb_add_impl(A, B, alpha, beta, threshold, 2, &n, "bml_add_norm_ellblock_hip");
    return n;
}

void bml_multiply_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                               bml_matrix_ellblock_hip_t *C, double alpha, double beta,
                               double threshold)
{
    const char *what = "bml_multiply_ellblock_hip";
    if (!valid_eb(A, what) || !valid_eb(B, what) || !valid_eb(C, what) ||
        !same_blocks(A, B, what) || !same_blocks(A, C, what))
        return;
    if (A == B && alpha == 1.0 && beta == 0.0) {
        free(bml_multiply_x2_ellblock_hip(A, C, threshold));
        return;
    }
    /* product into a temporary with C's block slots, then C <- beta*C + alpha*T */
    bml_matrix_ellblock_hip_t *T = bml_block_matrix_ellblock_hip(
        A->matrix_precision, A->NB, C->MB, C->M, A->bsize_host, A->distribution_mode);
    if (!T)
        return;
    if (A == B)
        free(bml_multiply_x2_ellblock_hip(A, T, threshold));
    else
        bml_multiply_AB_ellblock_hip(A, B, T, threshold);
    if (!bmlh_failed())
        bml_add_ellblock_hip(C, T, beta, alpha, threshold);
    bml_deallocate_ellblock_hip(T);
}

void bml_threshold_ellblock_hip(bml_matrix_ellblock_hip_t *A, double threshold)
{
    const char *what = "bml_threshold_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    if (!bmlh_ok(bmlhip_eb_threshold(A->matrix_precision, A->NB, A->bt, A->bsize, A->MB,
                                     A->value, A->indexb, A->nnzb, threshold), what))
        return;
    bmlhip_status st;
    bmlh_finish(what, &st);
    A->aux->stats_valid = 0;
}

double bml_trace_ellblock_hip(const bml_matrix_ellblock_hip_t *A)
{
    const char *what = "bml_trace_ellblock_hip";
    if (!valid_eb(A, what))
        return 0.0;
    if (!bmlh_ok(bmlhip_eb_trace(A->matrix_precision, A->NB, A->bt, A->bsize, A->MB, A->value,
                                 A->indexb, A->nnzb), what))
        return 0.0;
    bmlhip_status st;
    if (!bmlh_finish_sync(what, &st))
        return 0.0;
    return st.scalars[0];
}

/* ---- SP2 initialisation ---------------------------------------------------- */

/* bml_gershgorin_ellblock (ellblock/bml_normalize_ellblock_typed.c:98-170) */
void *bml_gershgorin_ellblock_hip(const bml_matrix_ellblock_hip_t *A)
{
    const char *what = "bml_gershgorin_ellblock_hip";
    if (!valid_eb(A, what))
        return NULL;
    if (!bmlh_ok(bmlhip_eb_gershgorin(A->matrix_precision, A->NB, A->bt, A->bsize, A->MB,
                                      A->value, A->indexb, A->nnzb), what))
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

/* bml_scale_inplace_ellblock (ellblock/bml_scale_ellblock_typed.c:81-87) */
void bml_scale_inplace_ellblock_hip(const void *scale_factor, bml_matrix_ellblock_hip_t *A)
{
    const char *what = "bml_scale_inplace_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    if (!bmlh_ok(bmlhip_eb_scale(A->matrix_precision, A->NB, A->bt, A->MB, A->value, A->nnzb,
                                 scale_factor), what))
        return;
    bmlhip_status st;
    bmlh_finish(what, &st);
}

/* bml_add_identity_ellblock (ellblock/bml_add_ellblock_typed.c:354-370):
 * A <- A + beta*I through the blockwise add, identity blocks of A's sizes */
void bml_add_identity_ellblock_hip(bml_matrix_ellblock_hip_t *A, double beta, double threshold)
{
    const char *what = "bml_add_identity_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    bml_matrix_ellblock_hip_t *I = bml_block_matrix_ellblock_hip(
        A->matrix_precision, A->NB, 1, A->N, A->bsize_host, A->distribution_mode);
    if (!I)
        return;
    if (bmlh_ok(bmlhip_eb_identity(I->matrix_precision, I->NB, I->bt, I->bsize, I->MB, I->value,
                                   I->indexb, I->nnzb), what))
        bml_add_ellblock_hip(A, I, 1.0, beta, threshold);
    bml_deallocate_ellblock_hip(I);
}

/* bml_normalize_ellblock (ellblock/bml_normalize_ellblock_typed.c:33-46):
 * A <- (maxeval*I - A)/(maxeval - mineval) */
void bml_normalize_ellblock_hip(bml_matrix_ellblock_hip_t *A, double mineval, double maxeval)
{
    const char *what = "bml_normalize_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    const double d = maxeval - mineval;
    unsigned char s[16];
    bmlh_put_scalar(A->matrix_precision, -1.0 / d, s);
    bml_scale_inplace_ellblock_hip(s, A);
    if (!bmlh_failed())
        bml_add_identity_ellblock_hip(A, maxeval / d, 0.0);
}

/* ---- Matrix Market files -------------------------------------------------- */

/* bml_write_bml_matrix_ellblock (ellblock/bml_utilities_ellblock_typed.c:174-260):
 * "%%%MatrixMarket matrix coordinate real general" ("complex" for complex
 * precisions), "N N n" with n the elements of the stored blocks, then every
 * element of every stored block (block rows in order, blocks in storage
 * order, each block row-major, zeros included) as 1-based global row and
 * column and %20.15e (complex: real and imaginary parts). */
void bml_write_bml_matrix_ellblock_hip(const bml_matrix_ellblock_hip_t *A, const char *filename)
{
    const char *what = "bml_write_bml_matrix_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    const int NB = A->NB, MB = A->MB, bt = A->bt, prec = A->matrix_precision;
    const int *bs = A->bsize_host;
    const size_t es = bmlh_elem_size(prec), tile = (size_t) bt * bt;
    char *v = malloc((size_t) NB * MB * tile * es + 1);
    int *ix = malloc(((size_t) NB * MB + 1) * sizeof(int));
    int *nn = malloc(((size_t) NB + 1) * sizeof(int));
    int *off = malloc(((size_t) NB + 1) * sizeof(int));
    FILE *f = NULL;
    if (!v || !ix || !nn || !off) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    bml_ellblock_hip_get_blocks(A, v, ix, nn, bt);
    if (bmlh_failed())
        goto out;
    if (!(f = fopen(filename, "w"))) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: cannot open %s", what, filename);
        goto out;
    }
    const int cplx = prec == BML_HIP_SINGLE_COMPLEX || prec == BML_HIP_DOUBLE_COMPLEX;
    fprintf(f, "%%%%%%MatrixMarket matrix coordinate %s general\n", cplx ? "complex" : "real");
    off[0] = 0;
    for (int ib = 1; ib <= NB; ib++)
        off[ib] = off[ib - 1] + bs[ib - 1];
    long long tot = 0;
    for (int ib = 0; ib < NB; ib++)
        for (int jp = 0; jp < nn[ib]; jp++)
            tot += (long long) bs[ib] * bs[ix[(size_t) ib * MB + jp]];
    fprintf(f, "%d %d %lld\n", A->N, A->N, tot);
    for (int ib = 0; ib < NB; ib++)
        for (int jp = 0; jp < nn[ib]; jp++) {
            const size_t slot = (size_t) ib * MB + jp;
            const int jb = ix[slot];
            for (int ii = 0; ii < bs[ib]; ii++)
                for (int jj = 0; jj < bs[jb]; jj++) {
                    const size_t at = slot * tile + (size_t) ii * bt + jj;
                    const int r = off[ib] + 1 + ii, c = off[jb] + 1 + jj;
                    switch (prec) {
                    case BML_HIP_SINGLE_REAL:
                        fprintf(f, "%d %d %20.15e\n", r, c, (double) ((const float *) v)[at]);
                        break;
                    case BML_HIP_DOUBLE_REAL:
                        fprintf(f, "%d %d %20.15e\n", r, c, ((const double *) v)[at]);
                        break;
                    case BML_HIP_SINGLE_COMPLEX:
                        fprintf(f, "%d %d %20.15e %20.15e\n", r, c,
                                (double) ((const float *) v)[2 * at],
                                (double) ((const float *) v)[2 * at + 1]);
                        break;
                    default:
                        fprintf(f, "%d %d %20.15e %20.15e\n", r, c, ((const double *) v)[2 * at],
                                ((const double *) v)[2 * at + 1]);
                        break;
                    }
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
    free(off);
}

/* bml_read_bml_matrix_ellblock (ellblock/bml_utilities_ellblock_typed.c:24-165):
 * each entry of the file goes to its block, which is appended to the block
 * row (zeroed) the first time one of its elements appears.  A "symmetric"
 * file also stores each off-diagonal element at its mirror.  The reference's
 * own symmetric branch writes the new block's column at the wrong slot and
 * never allocates the block, and it tests the in-block indices for the
 * diagonal; this follows the general branch for the mirror instead.  A
 * block row past MB is the M overflow (the reference asserts). */
void bml_read_bml_matrix_ellblock_hip(bml_matrix_ellblock_hip_t *A, const char *filename)
{
    const char *what = "bml_read_bml_matrix_ellblock_hip";
    if (!valid_eb(A, what))
        return;
    const int NB = A->NB, MB = A->MB, bt = A->bt, prec = A->matrix_precision, N = A->N;
    const int *bs = A->bsize_host;
    const int cplx = prec == BML_HIP_SINGLE_COMPLEX || prec == BML_HIP_DOUBLE_COMPLEX;
    const size_t es = bmlh_elem_size(prec), tile = (size_t) bt * bt;
    char *v = malloc((size_t) NB * MB * tile * es + 1);
    int *ix = malloc(((size_t) NB * MB + 1) * sizeof(int));
    int *nn = malloc(((size_t) NB + 1) * sizeof(int));
    int *blk = malloc(((size_t) N + 1) * sizeof(int)), *in = malloc(((size_t) N + 1) * sizeof(int));
    FILE *f = NULL;
    if (!v || !ix || !nn || !blk || !in) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    bml_ellblock_hip_get_blocks(A, v, ix, nn, bt);
    if (bmlh_failed())
        goto out;
    for (int ib = 0, r = 0; ib < NB; ib++)
        for (int ii = 0; ii < bs[ib]; ii++, r++) {
            blk[r] = ib;
            in[r] = ii;
        }
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
            got = fscanf(f, "%d %d %e", &r, &c, &fv) == 3;
        else if (!cplx)
            got = fscanf(f, "%d %d %le", &r, &c, &re) == 3;
        else
            got = fscanf(f, "%d %d %le %le", &r, &c, &re, &im) == 4;
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
            const int ib = blk[row], jb = blk[col];
            int jp = 0;
            while (jp < nn[ib] && ix[(size_t) ib * MB + jp] != jb)
                jp++;
            if (jp == nn[ib]) {
                if (nn[ib] >= MB) {
                    bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                               "%s: block row %d: Number of non-zeroes per row > M", what, ib);
                    goto out;
                }
                ix[(size_t) ib * MB + jp] = jb;
                memset(v + ((size_t) ib * MB + jp) * tile * es, 0, tile * es);
                nn[ib]++;
            }
            const size_t at = ((size_t) ib * MB + jp) * tile + (size_t) in[row] * bt + in[col];
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
    bml_ellblock_hip_set_blocks(A, v, ix, nn, MB, bt);
out:
    if (f)
        fclose(f);
    free(v);
    free(ix);
    free(nn);
    free(blk);
    free(in);
}

/* bml_import_from_dense_ellblock (ellblock/bml_import_ellblock_typed.c:20-96):
 * a zero matrix with the registered block sizes, then every block (ib, jb)
 * of the dense N x N array (row- or column-major), block columns in order,
 * kept iff its max |element| (bml_norm_inf, bml_utilities_ellblock_typed.c:
 * 260-277: fabs/cabs, as fabs/hypot, in the element's precision, compared
 * in double) is
 * above the threshold; the diagonal block is not forced.  A block row past
 * MB is the M overflow (the reference would overrun it). */
bml_matrix_ellblock_hip_t *bml_import_from_dense_ellblock_hip(int precision, int order, int N,
                                                             const void *A, double threshold,
                                                             int M, int distrib_mode)
{
    const char *what = "bml_import_from_dense_ellblock_hip";
    const size_t es = bmlh_elem_size(precision);
    if (!es || !A || (order != 0 && order != 1)) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: bad precision %d, order %d or array", what,
                   precision, order);
        return NULL;
    }
    bml_matrix_ellblock_hip_t *B = bml_zero_matrix_ellblock_hip(precision, N, M, distrib_mode);
    if (!B)
        return NULL;
    const int NB = B->NB, MB = B->MB, bt = B->bt;
    const int *bs = B->bsize_host;
    int tot = 0;
    for (int ib = 0; ib < NB; ib++)
        tot += bs[ib];
    if (tot != N) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: registered block sizes cover %d rows, not %d",
                   what, tot, N);
        bml_deallocate_ellblock_hip(B);
        return NULL;
    }
    const size_t tile = (size_t) bt * bt;
    char *v = calloc((size_t) NB * MB * tile + 1, es);
    int *ix = calloc((size_t) NB * MB + 1, sizeof(int));
    int *nn = calloc((size_t) NB + 1, sizeof(int));
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto fail;
    }
    for (int ib = 0, r0 = 0; ib < NB; r0 += bs[ib], ib++)
        for (int jb = 0, c0 = 0; jb < NB; c0 += bs[jb], jb++) {
            double norm = 0.0;
            for (int ii = 0; ii < bs[ib]; ii++)
                for (int jj = 0; jj < bs[jb]; jj++) {
                    const size_t at = order == 0 ? (size_t) (r0 + ii) * N + (c0 + jj)
                                                 : (size_t) (c0 + jj) * N + (r0 + ii);
                    double a;
                    switch (precision) {
                    case BML_HIP_SINGLE_REAL: a = (double) fabsf(((const float *) A)[at]); break;
                    case BML_HIP_DOUBLE_REAL: a = fabs(((const double *) A)[at]); break;
                    case BML_HIP_SINGLE_COMPLEX: /* cabsf: glibc's is hypotf */
                        a = (double) hypotf(((const float *) A)[2 * at], ((const float *) A)[2 * at + 1]);
                        break;
                    default: /* cabs: hypot */
                        a = hypot(((const double *) A)[2 * at], ((const double *) A)[2 * at + 1]);
                        break;
                    }
                    if (a > norm)
                        norm = a;
                }
            if (!(norm > threshold))
                continue;
            if (nn[ib] >= MB) {
                bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                           "%s: block row %d: Number of non-zeroes per row > M", what, ib);
                goto fail;
            }
            const size_t slot = (size_t) ib * MB + nn[ib]++;
            ix[slot] = jb;
            for (int ii = 0; ii < bs[ib]; ii++)
                for (int jj = 0; jj < bs[jb]; jj++) {
                    const size_t at = order == 0 ? (size_t) (r0 + ii) * N + (c0 + jj)
                                                 : (size_t) (c0 + jj) * N + (r0 + ii);
                    memcpy(v + (slot * tile + (size_t) ii * bt + jj) * es,
                           (const char *) A + at * es, es);
                }
        }
    bml_ellblock_hip_set_blocks(B, v, ix, nn, MB, bt);
    if (bmlh_failed())
        goto fail;
    free(v);
    free(ix);
    free(nn);
    return B;
fail:
    free(v);
    free(ix);
    free(nn);
    bml_deallocate_ellblock_hip(B);
    return NULL;
}

/* ---- bml_convert ------------------------------------------------------------ */

bml_matrix_ellblock_hip_t *bml_convert_ellblock_hip(const void *A, int precision, int M,
                                                    int distrib_mode)
{
    const char *what = "bml_convert_ellblock_hip";
    int N = 0;
    char *D = bmlh_dense_as(A, precision, &N, what);
    if (!D)
        return NULL;
    /* the registered block sizes (set from N if none are); every block of
     * the NB x NB block grid is stored, zeros included, in block-column order */
    const int *bsize = bml_hip_get_block_sizes(N, N);
    const int NB = bml_hip_get_nb();
    int tot = 0;
    for (int ib = 0; bsize && ib < NB; ib++)
        tot += bsize[ib];
    if (!bsize || tot != N) {
        free(D);
        if (bsize)
            bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: registered block sizes cover %d rows, not %d",
                       what, tot, N);
        return NULL;
    }
    bml_matrix_ellblock_hip_t *B =
        bml_block_matrix_ellblock_hip(precision, NB, NB, M, bsize, distrib_mode);
    if (!B) {
        free(D);
        return NULL;
    }
    const int bt = B->bt;
    const size_t es = bmlh_elem_size(precision), tile = (size_t) bt * bt;
    char *v = calloc((size_t) NB * NB * tile + 1, es);
    int *ix = calloc((size_t) NB * NB + 1, sizeof(int));
    int *nn = calloc((size_t) NB + 1, sizeof(int));
    if (!v || !ix || !nn) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        bml_deallocate_ellblock_hip(B);
        B = NULL;
        goto out;
    }
    for (int ib = 0, r0 = 0; ib < NB; r0 += bsize[ib], ib++) {
        nn[ib] = NB;
        for (int jb = 0, c0 = 0; jb < NB; c0 += bsize[jb], jb++) {
            const size_t slot = (size_t) ib * NB + jb;
            ix[slot] = jb;
            for (int ii = 0; ii < bsize[ib]; ii++)
                memcpy(v + (slot * tile + (size_t) ii * bt) * es,
                       D + ((size_t) (r0 + ii) * N + c0) * es, (size_t) bsize[jb] * es);
        }
    }
    bml_ellblock_hip_set_blocks(B, v, ix, nn, NB, bt);
    if (bmlh_failed()) {
        bml_deallocate_ellblock_hip(B);
        B = NULL;
    }
out:
    free(v);
    free(ix);
    free(nn);
    free(D);
    return B;
}

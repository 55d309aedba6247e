/*
 * Driver around the real reference (bml 2.4.0 C sources, built by ref.mk
 * into oracle/_ref/libbmlref.so).  TEST INFRASTRUCTURE ONLY: it writes the
 * golden vectors under tests/golden/ (through tests/golden/make_golden.py)
 * and times the reference CPU path for bench.py's cpu_baseline leg.
 *
 * Every operation goes through bml's public C API (bml_multiply_x2,
 * bml_multiply, bml_add, ...), so the results are what a bml client gets.
 *
 * File formats (little endian, written by bml_amd.io):
 *   ELL1: "ELL1", int32 precision (bml_matrix_precision_t), N, M,
 *         value[N*M] (element type of the precision), index int32[N*M],
 *         nnz int32[N]                          (row-major, as the reference)
 *   BLK1: "BLK1", int32 precision, NB, MB, bmax, bsize int32[NB],
 *         value[NB*MB*bmax*bmax] (bmax x bmax tile per block slot, block in
 *         the top-left corner, row-major), indexb int32[NB*MB], nnzb int32[NB]
 * Scalars on the command line accept C99 hex floats (exact).
 */
#include "bml.h"
#include "ellpack/bml_types_ellpack.h"
#include "ellblock/bml_types_ellblock.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Defined in the reference's bml_allocate.c:323 but declared in none of its
 * headers. */
bml_matrix_t *bml_block_matrix(bml_matrix_type_t matrix_type,
                               bml_matrix_precision_t matrix_precision,
                               int NB, int MB, int M, int *bsizes,
                               bml_distribution_mode_t distrib_mode);

static size_t elem_size(int prec)
{
    switch (prec) {
    case single_real: return 4;
    case double_real: return 8;
    case single_complex: return 8;
    case double_complex: return 16;
    }
    fprintf(stderr, "bad precision %d\n", prec);
    exit(3);
}

static void die(const char *m)
{
    fprintf(stderr, "ref_driver: %s\n", m);
    exit(3);
}

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (n && fread(p, sz, n, f) != n)
        die("short read");
}

static void wr(const void *p, size_t sz, size_t n, FILE *f)
{
    if (n && fwrite(p, sz, n, f) != n)
        die("short write");
}

/* ---- ELLPACK ---------------------------------------------------------- */

static bml_matrix_ellpack_t *read_ell(const char *path, int Mover)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        die("cannot open input");
    char magic[4];
    int hdr[3];
    rd(magic, 1, 4, f);
    if (memcmp(magic, "ELL1", 4))
        die("not an ELL1 file");
    rd(hdr, 4, 3, f);
    int prec = hdr[0], N = hdr[1], M = hdr[2];
    int Mm = Mover > 0 ? Mover : M;
    size_t es = elem_size(prec);
    bml_matrix_ellpack_t *A =
        bml_zero_matrix(ellpack, (bml_matrix_precision_t) prec, N, Mm, sequential);
    char *v = malloc(es * (size_t) N * M + 1);
    int *ix = malloc(sizeof(int) * (size_t) N * M + 1);
    rd(v, es, (size_t) N * M, f);
    rd(ix, 4, (size_t) N * M, f);
    rd(A->nnz, 4, (size_t) N, f);
    fclose(f);
    for (int i = 0; i < N; i++) {
        if (A->nnz[i] > Mm)
            die("row does not fit the requested M");
        memcpy((char *) A->value + es * (size_t) i * Mm, v + es * (size_t) i * M,
               es * (size_t) A->nnz[i]);
        memcpy(A->index + (size_t) i * Mm, ix + (size_t) i * M,
               4 * (size_t) A->nnz[i]);
    }
    free(v);
    free(ix);
    return A;
}

static void write_ell(const char *path, bml_matrix_ellpack_t *A)
{
    FILE *f = fopen(path, "wb");
    if (!f)
        die("cannot open output");
    int hdr[3] = { A->matrix_precision, A->N, A->M };
    size_t es = elem_size(A->matrix_precision);
    wr("ELL1", 1, 4, f);
    wr(hdr, 4, 3, f);
    wr(A->value, es, (size_t) A->N * A->M, f);
    wr(A->index, 4, (size_t) A->N * A->M, f);
    wr(A->nnz, 4, (size_t) A->N, f);
    fclose(f);
}

/* ---- ELLBLOCK --------------------------------------------------------- */

static int g_bmax = 0;

static bml_matrix_ellblock_t *read_blk(const char *path, int MBover)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        die("cannot open input");
    char magic[4];
    int hdr[4];
    rd(magic, 1, 4, f);
    if (memcmp(magic, "BLK1", 4))
        die("not a BLK1 file");
    rd(hdr, 4, 4, f);
    int prec = hdr[0], NB = hdr[1], MB = hdr[2], bmax = hdr[3];
    int MBm = MBover > 0 ? MBover : MB;
    size_t es = elem_size(prec), b2 = (size_t) bmax * bmax;
    int *bsize = malloc(sizeof(int) * NB);
    rd(bsize, 4, NB, f);
    int N = 0;
    for (int ib = 0; ib < NB; ib++)
        N += bsize[ib];
    bml_matrix_ellblock_t *A = bml_block_matrix(ellblock,
        (bml_matrix_precision_t) prec, NB, MBm, N, bsize, sequential);
    char *v = malloc(es * (size_t) NB * MB * b2 + 1);
    int *ix = malloc(sizeof(int) * (size_t) NB * MB + 1);
    rd(v, es, (size_t) NB * MB * b2, f);
    rd(ix, 4, (size_t) NB * MB, f);
    rd(A->nnzb, 4, NB, f);
    fclose(f);
    for (int ib = 0; ib < NB; ib++) {
        if (A->nnzb[ib] > MBm)
            die("block row does not fit the requested MB");
        for (int jp = 0; jp < A->nnzb[ib]; jp++) {
            int jb = ix[(size_t) ib * MB + jp];
            int bi = bsize[ib], bj = bsize[jb];
            char *blk = bml_allocate_memory(es * (size_t) bi * bj);
            const char *t = v + es * ((size_t) ib * MB + jp) * b2;
            for (int ii = 0; ii < bi; ii++)
                memcpy(blk + es * (size_t) ii * bj, t + es * (size_t) ii * bmax,
                       es * (size_t) bj);
            A->ptr_value[(size_t) ib * MBm + jp] = blk;
            A->indexb[(size_t) ib * MBm + jp] = jb;
        }
    }
    free(v);
    free(ix);
    free(bsize);
    g_bmax = bmax;
    return A;
}

static void write_blk(const char *path, bml_matrix_ellblock_t *A)
{
    FILE *f = fopen(path, "wb");
    if (!f)
        die("cannot open output");
    int NB = A->NB, MB = A->MB, bmax = g_bmax;
    size_t es = elem_size(A->matrix_precision), b2 = (size_t) bmax * bmax;
    int hdr[4] = { A->matrix_precision, NB, MB, bmax };
    wr("BLK1", 1, 4, f);
    wr(hdr, 4, 4, f);
    wr(A->bsize, 4, NB, f);
    char *v = calloc(es * (size_t) NB * MB * b2 + 1, 1);
    int *ix = malloc(sizeof(int) * (size_t) NB * MB + 1);
    for (size_t s = 0; s < (size_t) NB * MB; s++)
        ix[s] = -1;
    for (int ib = 0; ib < NB; ib++)
        for (int jp = 0; jp < A->nnzb[ib]; jp++) {
            size_t s = (size_t) ib * MB + jp;
            int jb = A->indexb[s];
            int bi = A->bsize[ib], bj = A->bsize[jb];
            const char *blk = A->ptr_value[s];
            for (int ii = 0; ii < bi; ii++)
                memcpy(v + es * (s * b2 + (size_t) ii * bmax),
                       blk + es * (size_t) ii * bj, es * (size_t) bj);
            ix[s] = jb;
        }
    wr(v, es, (size_t) NB * MB * b2, f);
    wr(ix, 4, (size_t) NB * MB, f);
    wr(A->nnzb, 4, NB, f);
    fclose(f);
    free(v);
    free(ix);
}

/* ---- helpers ---------------------------------------------------------- */

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

static int cmp_d(const void *a, const void *b)
{
    double x = *(const double *) a, y = *(const double *) b;
    return (x > y) - (x < y);
}

static double num(const char *s)
{
    return strtod(s, NULL);
}

/* Scale every stored value of an ELLPACK matrix by (T) s, as BLAS ?scal
 * would inside bml_scale_inplace (absent here: the reference is built
 * without BLAS). */
static void scale_values(bml_matrix_ellpack_t *A, double s)
{
    size_t n = (size_t) A->N * A->M;
    switch (A->matrix_precision) {
    case single_real: {
        float *v = A->value, f = (float) s;
        for (size_t k = 0; k < n; k++) v[k] = f * v[k];
        break;
    }
    case double_real: {
        double *v = A->value;
        for (size_t k = 0; k < n; k++) v[k] = s * v[k];
        break;
    }
    case single_complex: {
        float _Complex *v = A->value, f = (float _Complex) s;
        for (size_t k = 0; k < n; k++) v[k] = f * v[k];
        break;
    }
    case double_complex: {
        double _Complex *v = A->value, f = (double _Complex) s;
        for (size_t k = 0; k < n; k++) v[k] = f * v[k];
        break;
    }
    default:
        die("bad precision");
    }
}

/*
 * SP2 purification (Niklasson, Phys. Rev. B 66, 155115 (2002)) driven
 * through bml's API the way client codes drive it: Gershgorin bounds,
 * X0 = (emax*I - H)/(emax - emin), then per iteration X2 = X*X (thresholded,
 * returns tr X and tr X^2), pick X <- X^2 or X <- 2X - X^2 by whichever trace
 * lands closer to nocc, and apply it with bml_add_norm, whose return value
 * is ||X - X^2||_F^2.  Stops at ||X - X^2||_F <= tol or maxit iterations.
 */
static int sp2(bml_matrix_ellpack_t *H, int M, double nocc, double thr,
               double tol, int maxit, bml_matrix_ellpack_t **Xout, int verbose)
{
    bml_matrix_ellpack_t *X = bml_zero_matrix(ellpack, H->matrix_precision,
                                              H->N, M, sequential);
    size_t es = elem_size(H->matrix_precision);
    for (int i = 0; i < H->N; i++) {
        memcpy((char *) X->value + es * (size_t) i * M,
               (char *) H->value + es * (size_t) i * H->M, es * H->nnz[i]);
        memcpy(X->index + (size_t) i * M, H->index + (size_t) i * H->M,
               4 * (size_t) H->nnz[i]);
        X->nnz[i] = H->nnz[i];
    }
    double *g = bml_gershgorin(H);
    double emin = g[0], emax = g[1];
    bml_free_memory(g);
    scale_values(X, -1.0 / (emax - emin));
    bml_add_identity(X, emax / (emax - emin), 0.0);
    bml_matrix_ellpack_t *X2 = bml_zero_matrix(ellpack, H->matrix_precision,
                                               H->N, M, sequential);
    int it;
    for (it = 0; it < maxit; it++) {
        double *tr = bml_multiply_x2(X, X2, thr);
        double trx = tr[0], trx2 = tr[1];
        bml_free_memory(tr);
        int branch = fabs(trx2 - nocc) < fabs(2.0 * trx - trx2 - nocc) ? 0 : 1;
        double n2 = branch ? bml_add_norm(X, X2, 2.0, -1.0, thr)
                           : bml_add_norm(X, X2, 0.0, 1.0, thr);
        double nrm = sqrt(n2);
        if (verbose)
            printf("iter %d %a %a %a %d\n", it, trx, trx2, nrm, branch);
        if (nrm <= tol) {
            it++;
            break;
        }
    }
    bml_deallocate((bml_matrix_t **) &X2);
    *Xout = X;
    return it;
}

int main(int argc, char **argv)
{
    if (argc < 2)
        die("usage: ref_driver <op> ...");
    const char *op = argv[1];
#define NEED(n) if (argc < (n)) die("missing arguments")
    if (!strcmp(op, "x2")) {
        NEED(6);
        bml_matrix_ellpack_t *X = read_ell(argv[2], 0);
        bml_matrix_ellpack_t *X2 = bml_zero_matrix(ellpack, X->matrix_precision,
                                                   X->N, atoi(argv[4]), sequential);
        double *tr = bml_multiply_x2(X, X2, num(argv[5]));
        printf("trace %a %a\n", tr[0], tr[1]);
        write_ell(argv[3], X2);
    } else if (!strcmp(op, "ab")) {
        NEED(7);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0), *B = read_ell(argv[3], 0);
        bml_matrix_ellpack_t *C = bml_zero_matrix(ellpack, A->matrix_precision,
                                                  A->N, atoi(argv[5]), sequential);
        bml_multiply_AB(A, B, C, num(argv[6]));
        write_ell(argv[4], C);
    } else if (!strcmp(op, "multiply")) {
        NEED(9);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0);
        bml_matrix_ellpack_t *B = strcmp(argv[3], "same") ? read_ell(argv[3], 0) : A;
        bml_matrix_ellpack_t *C = read_ell(argv[4], 0);
        bml_multiply(A, B, C, num(argv[6]), num(argv[7]), num(argv[8]));
        write_ell(argv[5], C);
    } else if (!strcmp(op, "add") || !strcmp(op, "add_norm")) {
        NEED(8);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0), *B = read_ell(argv[3], 0);
        if (!strcmp(op, "add"))
            bml_add(A, B, num(argv[5]), num(argv[6]), num(argv[7]));
        else
            printf("norm %a\n", bml_add_norm(A, B, num(argv[5]), num(argv[6]),
                                             num(argv[7])));
        write_ell(argv[4], A);
    } else if (!strcmp(op, "add_identity")) {
        NEED(6);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0);
        bml_add_identity(A, num(argv[4]), num(argv[5]));
        write_ell(argv[3], A);
    } else if (!strcmp(op, "threshold")) {
        NEED(5);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0);
        bml_threshold(A, num(argv[4]));
        write_ell(argv[3], A);
    } else if (!strcmp(op, "trace")) {
        NEED(3);
        printf("trace %a\n", bml_trace(read_ell(argv[2], 0)));
    } else if (!strcmp(op, "gershgorin")) {
        NEED(3);
        double *g = bml_gershgorin(read_ell(argv[2], 0));
        printf("gersh %a %a\n", g[0], g[1]);
    } else if (!strcmp(op, "mm_write")) {
        /* mm_write <in.ell> <out.mtx>: bml_write_bml_matrix */
        NEED(4);
        bml_write_bml_matrix(read_ell(argv[2], 0), argv[3]);
    } else if (!strcmp(op, "mm_read")) {
        /* mm_read <in.mtx> <precision> <N> <M> <out.ell>: bml_read_bml_matrix
         * into a zero matrix */
        NEED(7);
        bml_matrix_ellpack_t *A = bml_zero_matrix(ellpack, (bml_matrix_precision_t) atoi(argv[3]),
                                                  atoi(argv[4]), atoi(argv[5]), sequential);
        bml_read_bml_matrix(A, argv[2]);
        write_ell(argv[6], A);
    } else if (!strcmp(op, "eb_mm_write")) {
        /* eb_mm_write <in.blk> <out.mtx>: bml_write_bml_matrix */
        NEED(4);
        bml_write_bml_matrix(read_blk(argv[2], 0), argv[3]);
    } else if (!strcmp(op, "eb_mm_read")) {
        /* eb_mm_read <in.mtx> <shape.blk> <out.blk>: bml_read_bml_matrix into
         * a zero matrix of shape.blk's precision, block sizes and MB */
        NEED(5);
        bml_matrix_ellblock_t *S = read_blk(argv[3], 0);
        bml_matrix_ellblock_t *A = bml_block_matrix(ellblock, S->matrix_precision, S->NB,
                                                    S->MB, S->N, S->bsize, sequential);
        bml_read_bml_matrix(A, argv[2]);
        write_blk(argv[4], A);
    } else if (!strcmp(op, "convert")) {
        /* convert <in.ell> <M> <out.ell>: bml_convert to ellpack */
        NEED(5);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0);
        bml_matrix_ellpack_t *B = bml_convert(A, ellpack, A->matrix_precision, atoi(argv[3]),
                                              sequential);
        write_ell(argv[4], B);
    } else if (!strcmp(op, "convert_blk")) {
        /* convert_blk <in.blk> <M> <out.ell>: bml_convert of an ellblock
         * matrix to ellpack */
        NEED(5);
        bml_matrix_ellblock_t *A = read_blk(argv[2], 0);
        bml_matrix_ellpack_t *B = bml_convert(A, ellpack, A->matrix_precision, atoi(argv[3]),
                                              sequential);
        write_ell(argv[4], B);
    } else if (!strcmp(op, "eb_convert")) {
        /* eb_convert <in.ell> <M> <out.blk>: bml_convert of an ellpack matrix
         * to ellblock (the default block sizes of the registry) */
        NEED(5);
        bml_matrix_ellpack_t *A = read_ell(argv[2], 0);
        bml_matrix_ellblock_t *B = bml_convert(A, ellblock, A->matrix_precision, atoi(argv[3]),
                                               sequential);
        g_bmax = 0;
        for (int ib = 0; ib < B->NB; ib++)
            if (B->bsize[ib] > g_bmax)
                g_bmax = B->bsize[ib];
        write_blk(argv[4], B);
    } else if (!strcmp(op, "eb_x2")) {
        NEED(6);
        bml_matrix_ellblock_t *X = read_blk(argv[2], 0);
        bml_matrix_ellblock_t *X2 = bml_block_matrix(ellblock, X->matrix_precision,
            X->NB, atoi(argv[4]), X->N, X->bsize, sequential);
        double *tr = bml_multiply_x2(X, X2, num(argv[5]));
        printf("trace %a %a\n", tr[0], tr[1]);
        write_blk(argv[3], X2);
    } else if (!strcmp(op, "eb_ab")) {
        NEED(7);
        bml_matrix_ellblock_t *A = read_blk(argv[2], 0), *B = read_blk(argv[3], 0);
        bml_matrix_ellblock_t *C = bml_block_matrix(ellblock, A->matrix_precision,
            A->NB, atoi(argv[5]), A->N, A->bsize, sequential);
        bml_multiply_AB(A, B, C, num(argv[6]));
        write_blk(argv[4], C);
    } else if (!strcmp(op, "eb_multiply")) {
        NEED(9);
        bml_matrix_ellblock_t *A = read_blk(argv[2], 0);
        bml_matrix_ellblock_t *B = strcmp(argv[3], "same") ? read_blk(argv[3], 0) : A;
        bml_matrix_ellblock_t *C = read_blk(argv[4], 0);
        bml_multiply(A, B, C, num(argv[6]), num(argv[7]), num(argv[8]));
        write_blk(argv[5], C);
    } else if (!strcmp(op, "eb_add")) {
        NEED(8);
        bml_matrix_ellblock_t *A = read_blk(argv[2], 0), *B = read_blk(argv[3], 0);
        bml_add(A, B, num(argv[5]), num(argv[6]), num(argv[7]));
        write_blk(argv[4], A);
    } else if (!strcmp(op, "eb_threshold")) {
        NEED(5);
        bml_matrix_ellblock_t *A = read_blk(argv[2], 0);
        bml_threshold(A, num(argv[4]));
        write_blk(argv[3], A);
    } else if (!strcmp(op, "eb_trace")) {
        NEED(3);
        printf("trace %a\n", bml_trace(read_blk(argv[2], 0)));
    } else if (!strcmp(op, "sp2")) {
        NEED(9);
        bml_matrix_ellpack_t *H = read_ell(argv[2], 0), *X = NULL;
        int it = sp2(H, atoi(argv[4]), num(argv[5]), num(argv[6]), num(argv[7]),
                     atoi(argv[8]), &X, 1);
        printf("iters %d\n", it);
        write_ell(argv[3], X);
    } else if (!strcmp(op, "sp2_iters")) {
        /* time K SP2 iterations (after Gershgorin + normalize) through the
         * reference: sp2_iters <H> <M> <nocc> <thr> <K> */
        NEED(7);
        bml_matrix_ellpack_t *H = read_ell(argv[2], 0);
        int M = atoi(argv[3]), K = atoi(argv[6]);
        double nocc = num(argv[4]), thr = num(argv[5]);
        bml_matrix_ellpack_t *X = bml_zero_matrix(ellpack, H->matrix_precision, H->N, M,
                                                  sequential);
        size_t es = elem_size(H->matrix_precision);
        for (int i = 0; i < H->N; i++) {
            memcpy((char *) X->value + es * (size_t) i * M,
                   (char *) H->value + es * (size_t) i * H->M, es * H->nnz[i]);
            memcpy(X->index + (size_t) i * M, H->index + (size_t) i * H->M,
                   4 * (size_t) H->nnz[i]);
            X->nnz[i] = H->nnz[i];
        }
        double *g = bml_gershgorin(H);
        scale_values(X, -1.0 / (g[1] - g[0]));
        bml_add_identity(X, g[1] / (g[1] - g[0]), 0.0);
        bml_matrix_ellpack_t *X2 = bml_zero_matrix(ellpack, H->matrix_precision, H->N, M,
                                                   sequential);
        double *t = malloc(sizeof(double) * (K + 1));
        for (int it = 0; it < K; it++) {
            double t0 = now_ms();
            double *tr = bml_multiply_x2(X, X2, thr);
            int branch = fabs(tr[1] - nocc) < fabs(2.0 * tr[0] - tr[1] - nocc) ? 0 : 1;
            bml_free_memory(tr);
            if (branch)
                bml_add_norm(X, X2, 2.0, -1.0, thr);
            else
                bml_add_norm(X, X2, 0.0, 1.0, thr);
            t[it] = now_ms() - t0;
            printf("iter_ms %d %.6f\n", it, t[it]);
        }
        qsort(t, K, sizeof(double), cmp_d);
        printf("median_iter_ms %.6f\n", t[K / 2]);
    } else if (!strcmp(op, "bench_x2")) {
        NEED(6);
        bml_matrix_ellpack_t *X = read_ell(argv[2], 0);
        bml_matrix_ellpack_t *X2 = bml_zero_matrix(ellpack, X->matrix_precision,
                                                   X->N, atoi(argv[3]), sequential);
        int reps = atoi(argv[5]);
        double *t = malloc(sizeof(double) * (reps + 1));
        bml_free_memory(bml_multiply_x2(X, X2, num(argv[4])));
        for (int r = 0; r < reps; r++) {
            double t0 = now_ms();
            bml_free_memory(bml_multiply_x2(X, X2, num(argv[4])));
            t[r] = now_ms() - t0;
        }
        qsort(t, reps, sizeof(double), cmp_d);
        printf("median_ms %.6f\n", t[reps / 2]);
    } else if (!strcmp(op, "bench_x2_rows")) {
        /* bench_x2_rows <in.ell> <M2> <thr> <rows> <reps>: bml_multiply_x2 on
         * the row block [0, rows) (the matrix's domain, which the reference's
         * loop runs over: bml_multiply_ellpack_typed.c:155-157); prints the
         * median time and the block's partial products */
        NEED(7);
        bml_matrix_ellpack_t *X = read_ell(argv[2], 0);
        bml_matrix_ellpack_t *X2 = bml_zero_matrix(ellpack, X->matrix_precision,
                                                   X->N, atoi(argv[3]), sequential);
        int rows = atoi(argv[5]), reps = atoi(argv[6]);
        if (rows < 1 || rows > X->N)
            die("rows out of range");
        X->domain->localRowMin[0] = 0;
        X->domain->localRowMax[0] = rows;
        long long prods = 0;
        for (int i = 0; i < rows; i++)
            for (int jp = 0; jp < X->nnz[i]; jp++)
                prods += X->nnz[X->index[(size_t) i * X->M + jp]];
        double *t = malloc(sizeof(double) * (reps + 1));
        for (int r = 0; r < reps; r++) {
            double t0 = now_ms();
            bml_free_memory(bml_multiply_x2(X, X2, num(argv[4])));
            t[r] = now_ms() - t0;
        }
        qsort(t, reps, sizeof(double), cmp_d);
        printf("median_ms %.6f products %lld\n", t[reps / 2], prods);
    } else if (!strcmp(op, "bench_sp2")) {
        NEED(9);
        bml_matrix_ellpack_t *H = read_ell(argv[2], 0);
        int reps = atoi(argv[8]);
        double *t = malloc(sizeof(double) * (reps + 1));
        int it = 0;
        for (int r = 0; r < reps; r++) {
            bml_matrix_ellpack_t *X = NULL;
            double t0 = now_ms();
            it = sp2(H, atoi(argv[3]), num(argv[4]), num(argv[5]), num(argv[6]),
                     atoi(argv[7]), &X, 0);
            t[r] = now_ms() - t0;
            bml_deallocate((bml_matrix_t **) &X);
        }
        qsort(t, reps, sizeof(double), cmp_d);
        printf("iters %d\nmedian_ms %.6f\n", it, t[reps / 2]);
    } else {
        die("unknown op");
    }
    return 0;
}

/*
 * A C client of libbml_hip.so that calls the C ABI through include/bml_hip.h
 * directly (no ctypes), in the argument order of the bml backend functions
 * each entry point replaces, and checks the results against plain dense
 * loops written here.
 *
 * Built with gcc by tests/test_c_abi.py:
 *   - always; run with the argument "link" it only proves that it linked
 *     against the library and exits 0 (no GPU needed);
 *   - where bml's sources are present: with -DBML_REF_INCLUDE, which includes
 *     the reference's struct definitions (ellpack/bml_types_ellpack.h:11-35,
 *     ellblock/bml_types_ellblock.h:11-54) and static_asserts that the
 *     leading members the dispatchers read sit at the same offsets;
 *   - on a GPU ("run" argument): the checks below.
 */
#include "../../include/bml_hip.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef BML_REF_INCLUDE
#include "ellblock/bml_types_ellblock.h"
#include "ellpack/bml_types_ellpack.h"
#define SAME_OFFSET(A, B, f) \
    _Static_assert(offsetof(A, f) == offsetof(B, f), "offset of " #f " differs")
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, matrix_type);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, matrix_precision);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, distribution_mode);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, N);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, M);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, value);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, index);
SAME_OFFSET(bml_matrix_ellpack_t, bml_matrix_ellpack_hip_t, nnz);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, matrix_type);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, matrix_precision);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, distribution_mode);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, N);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, M);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, NB);
SAME_OFFSET(bml_matrix_ellblock_t, bml_matrix_ellblock_hip_t, MB);
/* the type ids: appended after distributed2d */
_Static_assert(BML_HIP_TYPE_ELLPACK == distributed2d + 1, "ellpack_hip id");
_Static_assert(BML_HIP_TYPE_ELLBLOCK == distributed2d + 2, "ellblock_hip id");
_Static_assert(BML_HIP_DOUBLE_REAL == double_real && BML_HIP_SINGLE_COMPLEX == single_complex,
               "precision ids");
_Static_assert(BML_HIP_DENSE_COLUMN_MAJOR == dense_column_major, "dense order ids");
_Static_assert(sizeof(bml_matrix_type_t) == sizeof(int) &&
                   sizeof(bml_matrix_precision_t) == sizeof(int) &&
                   sizeof(bml_distribution_mode_t) == sizeof(int),
               "enum tags are ints");
#endif

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            fprintf(stderr, "abi_client: FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");                         \
            failures++;                                    \
        }                                                  \
    } while (0)

/* symmetric banded test matrix: a_ij = exp(-|i-j|/3) cos(i+2j) for |i-j| <= 6 */
static void banded(int N, double *A)
{
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const int d = abs(i - j);
            A[(size_t) i * N + j] = d <= 6 ? exp(-d / 3.0) * cos((double) (i + j) * 0.37) : 0.0;
        }
}

/* X2 = X*X kept as bml_multiply_x2 keeps it: the diagonal, and |v| > thr */
static void dense_x2(int N, const double *X, double thr, double *X2, double *tr)
{
    tr[0] = tr[1] = 0.0;
    for (int i = 0; i < N; i++) {
        tr[0] += X[(size_t) i * N + i];
        for (int k = 0; k < N; k++) {
            double s = 0.0;
            for (int j = 0; j < N; j++) s += X[(size_t) i * N + j] * X[(size_t) j * N + k];
            X2[(size_t) i * N + k] = (i == k || fabs(s) > thr) ? s : 0.0;
        }
        tr[1] += X2[(size_t) i * N + i];
    }
}

static double maxdiff(int n, const double *a, const double *b)
{
    double m = 0.0;
    for (int k = 0; k < n; k++) {
        const double d = fabs(a[k] - b[k]);
        m = d > m ? d : m;
    }
    return m;
}

static void run_ellpack(void)
{
    const int N = 160, M = 64;
    const double thr = 1e-4;
    double *A = malloc(sizeof(double) * N * N), *R = malloc(sizeof(double) * N * N);
    double rtr[2];
    banded(N, A);
    dense_x2(N, A, thr, R, rtr);
    /* bml_import_from_dense_ellpack(prec, order, N, A, threshold, M, mode) */
    bml_matrix_ellpack_hip_t *X =
        bml_import_from_dense_ellpack_hip(BML_HIP_DOUBLE_REAL, BML_HIP_DENSE_ROW_MAJOR, N, A, 0.0,
                                          M, 0);
    bml_matrix_ellpack_hip_t *X2 = bml_zero_matrix_ellpack_hip(BML_HIP_DOUBLE_REAL, N, M, 0);
    CHECK(X && X2, "allocation");
    CHECK(X->matrix_type == BML_HIP_TYPE_ELLPACK && X->N == N && X->M == M, "struct fields");
    for (int engine = 0; engine <= 2; engine++) {
        bml_hip_set_row_engine(engine);
        double *tr = bml_multiply_x2_ellpack_hip(X, X2, thr);
        CHECK(tr && fabs(tr[0] - rtr[0]) <= 1e-12 * fabs(rtr[0]) &&
                  fabs(tr[1] - rtr[1]) <= 1e-12 * fabs(rtr[1]),
              "x2 traces, engine %d", engine);
        free(tr);
        double *D = bml_export_to_dense_ellpack_hip(X2, BML_HIP_DENSE_ROW_MAJOR);
        CHECK(D && maxdiff(N * N, D, R) <= 1e-12, "x2 values, engine %d (%g)", engine,
              D ? maxdiff(N * N, D, R) : -1.0);
        free(D);
    }
    bml_hip_set_row_engine(BML_HIP_ROW_ENGINE_EXACT);
    /* bml_add(A, B, alpha, beta, threshold): X2 <- 0.5 X2 - 2 X */
    bml_add_ellpack_hip(X2, X, 0.5, -2.0, thr);
    double *D = bml_export_to_dense_ellpack_hip(X2, BML_HIP_DENSE_ROW_MAJOR);
    double worst = 0.0;
    for (int k = 0; k < N * N; k++) {
        double v = 0.5 * R[k] - 2.0 * A[k];
        v = fabs(v) > thr ? v : 0.0;
        worst = fabs(D[k] - v) > worst ? fabs(D[k] - v) : worst;
    }
    CHECK(worst <= 1e-12, "add (%g)", worst);
    free(D);
    /* bml_trace, bml_threshold */
    double t = bml_trace_ellpack_hip(X);
    CHECK(fabs(t - rtr[0]) <= 1e-12 * fabs(rtr[0]), "trace");
    bml_threshold_ellpack_hip(X, 0.1);
    D = bml_export_to_dense_ellpack_hip(X, BML_HIP_DENSE_ROW_MAJOR);
    worst = 0.0;
    for (int k = 0; k < N * N; k++) {
        const double v = fabs(A[k]) > 0.1 ? A[k] : 0.0;
        worst = fabs(D[k] - v) > worst ? fabs(D[k] - v) : worst;
    }
    CHECK(worst == 0.0, "threshold");
    free(D);
    /* M overflow, recorded instead of exiting */
    bml_hip_set_error_mode(BML_HIP_ERRORS_RECORD);
    bml_matrix_ellpack_hip_t *S = bml_zero_matrix_ellpack_hip(BML_HIP_DOUBLE_REAL, N, 4, 0);
    double *tr2 = bml_multiply_x2_ellpack_hip(X2, S, 0.0);
    char msg[256];
    CHECK(tr2 == NULL && bml_hip_last_error(msg, sizeof msg) == BML_HIP_ERR_M_OVERFLOW,
          "M overflow recorded");
    bml_hip_clear_error();
    bml_hip_set_error_mode(BML_HIP_ERRORS_EXIT);
    bml_deallocate_ellpack_hip(S);
    bml_deallocate_ellpack_hip(X);
    bml_deallocate_ellpack_hip(X2);
    free(A);
    free(R);
}

static void run_ellblock(void)
{
    const int NB = 12, b = 8, N = NB * b, MB = NB;
    int bs[12];
    for (int k = 0; k < NB; k++) bs[k] = b;
    bml_hip_set_block_sizes(bs, NB, MB);
    double *A = malloc(sizeof(double) * N * N), *R = malloc(sizeof(double) * N * N);
    double rtr[2];
    banded(N, A);
    dense_x2(N, A, 0.0, R, rtr);
    /* bml_import_from_dense_ellblock(prec, order, N, A, threshold, M, mode) */
    bml_matrix_ellblock_hip_t *X = bml_import_from_dense_ellblock_hip(
        BML_HIP_DOUBLE_REAL, BML_HIP_DENSE_ROW_MAJOR, N, A, 0.0, N, 0);
    bml_matrix_ellblock_hip_t *X2 = bml_zero_matrix_ellblock_hip(BML_HIP_DOUBLE_REAL, N, N, 0);
    CHECK(X && X2 && X->NB == NB && X->MB == MB, "ellblock allocation");
    double *tr = bml_multiply_x2_ellblock_hip(X, X2, 0.0);
    CHECK(tr && fabs(tr[0] - rtr[0]) <= 1e-12 * fabs(rtr[0]) &&
              fabs(tr[1] - rtr[1]) <= 1e-12 * fabs(rtr[1]),
          "ellblock x2 traces");
    free(tr);
    double *D = bml_export_to_dense_ellblock_hip(X2, BML_HIP_DENSE_ROW_MAJOR);
    CHECK(D && maxdiff(N * N, D, R) <= 1e-12, "ellblock x2 values");
    free(D);
    bml_deallocate_ellblock_hip(X);
    bml_deallocate_ellblock_hip(X2);
    free(A);
    free(R);
}

static void run_sp2(void)
{
    /* a gapped chain: onsite +-1 alternating, hopping -0.2, periodic */
    const int N = 64;
    double *H = calloc((size_t) N * N, sizeof(double));
    for (int i = 0; i < N; i++) {
        H[(size_t) i * N + i] = (i % 2) ? -1.0 : 1.0;
        H[(size_t) i * N + (i + 1) % N] = H[(size_t) ((i + 1) % N) * N + i] = -0.2;
    }
    bml_matrix_ellpack_hip_t *Hm =
        bml_import_from_dense_ellpack_hip(BML_HIP_DOUBLE_REAL, 0, N, H, 0.0, N, 0);
    bml_matrix_ellpack_hip_t *X = bml_zero_matrix_ellpack_hip(BML_HIP_DOUBLE_REAL, N, N, 0);
    bml_hip_sp2_options o;
    bml_hip_sp2_defaults(&o);
    o.threshold = 0.0;
    o.tol = 1e-10;
    bml_hip_sp2_result r;
    bml_hip_sp2_ellpack(Hm, X, N / 2, &o, &r, NULL);
    CHECK(r.converged && fabs(r.trace_x - N / 2) < 1e-8, "sp2 converged to tr X = N/2");
    bml_deallocate_ellpack_hip(Hm);
    bml_deallocate_ellpack_hip(X);
    free(H);
}

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "link") == 0) {
        printf("abi_client: linked, %s\n", bml_hip_version());
        return 0;
    }
    if (bml_hip_init(0) != 0) {
        fprintf(stderr, "abi_client: no device\n");
        return 2;
    }
    run_ellpack();
    run_ellblock();
    run_sp2();
    bml_hip_shutdown();
    if (failures) {
        fprintf(stderr, "abi_client: %d check(s) failed\n", failures);
        return 1;
    }
    printf("abi_client: all checks passed\n");
    return 0;
}

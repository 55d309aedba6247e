/*
 * bml_hip.h -- C ABI of the MI355X (gfx950) ELLPACK-R / ELLBLOCK matrix types
 * `ellpack_hip` and `ellblock_hip`, drop-ins behind bml's C API.
 *
 * Every entry point below replaces the bml backend function cited beside it
 * (paths relative to bml's src/C-interface/).  bml's public dispatchers
 * (bml_multiply_x2, bml_multiply, bml_add, bml_threshold, bml_trace, ...)
 * switch on the type tag read from the first int of the matrix
 * (bml_introspection.c:22-35); the two structs below start with the same
 * three tag fields as bml_matrix_ellpack_t (ellpack/bml_types_ellpack.h:13-18)
 * so a `case ellpack_hip:` / `case ellblock_hip:` line in each dispatcher
 * is all the integration needs.
 *
 * Conventions shared with bml:
 *   - precision: bml_matrix_precision_t values (1 single_real, 2 double_real,
 *     3 single_complex, 4 double_complex), bml_types.h:24-36;
 *   - ELLPACK-R storage: value[N*M], index[N*M] row-major
 *     (ROWMAJOR(i,jp,N,M) = i*M+jp, macros.h:13), nnz[N]; here the three
 *     arrays live in device memory (HBM);
 *   - thresholds compare |v| > threshold strictly (typed.h:36,73,112,150);
 *   - errors are fatal: by default the library prints the message and exits
 *     with EXIT_FAILURE, as LOG_ERROR does (bml_logger.c:111-115).  A caller
 *     may select BML_HIP_ERRORS_RECORD to get a return to the caller with the
 *     message kept for bml_hip_last_error() instead.
 *   - results are returned the way bml returns them: bml_multiply_x2 and
 *     bml_gershgorin hand back a malloc'd double[2] the caller frees;
 *     exports hand back malloc'd host arrays.
 *
 * All device work runs on one HIP stream per process (default priority).
 * With the default BML_HIP_SYNC_EACH_CALL every entry point returns with its
 * work complete and its overflow checks done; BML_HIP_SYNC_DEFERRED lets
 * void entry points return early, checks deferred to the next call that
 * synchronises (an export, a trace, bml_hip_synchronize()).
 */
#ifndef BML_HIP_H
#define BML_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* bml_matrix_type_t values for the new types, appended after distributed2d
 * (= 5, bml_types.h:7-21) so the Fortran ids 0-4 stay unchanged. */
#define BML_HIP_TYPE_ELLPACK 6
#define BML_HIP_TYPE_ELLBLOCK 7

/* bml_matrix_precision_t (bml_types.h:24-36) */
#define BML_HIP_SINGLE_REAL 1
#define BML_HIP_DOUBLE_REAL 2
#define BML_HIP_SINGLE_COMPLEX 3
#define BML_HIP_DOUBLE_COMPLEX 4

/* bml_dense_order_t (bml_types.h:38-44) */
#define BML_HIP_DENSE_ROW_MAJOR 0
#define BML_HIP_DENSE_COLUMN_MAJOR 1

/* error modes */
#define BML_HIP_ERRORS_EXIT 0   /* print and exit(EXIT_FAILURE), as LOG_ERROR */
#define BML_HIP_ERRORS_RECORD 1 /* keep the message, return to the caller */

/* synchronisation modes */
#define BML_HIP_SYNC_EACH_CALL 0
#define BML_HIP_SYNC_DEFERRED 1

/* error codes reported by bml_hip_last_error() */
#define BML_HIP_OK 0
#define BML_HIP_ERR_M_OVERFLOW 1   /* "Number of non-zeroes per row > M" */
#define BML_HIP_ERR_ARGUMENT 2     /* bad argument, mismatched shapes/types */
#define BML_HIP_ERR_DEVICE 3       /* HIP runtime failure */
#define BML_HIP_ERR_ALLOC 4        /* device or host allocation failed */
#define BML_HIP_ERR_UNSUPPORTED 5  /* precision/shape the backend lacks */

struct bml_hip_aux;

/* ELLPACK-R matrix resident in HBM.  Field order of the first five members
 * is that of bml_matrix_ellpack_t (ellpack/bml_types_ellpack.h:11-24). */
typedef struct bml_matrix_ellpack_hip_t {
    int matrix_type;       /* BML_HIP_TYPE_ELLPACK */
    int matrix_precision;  /* bml_matrix_precision_t */
    int distribution_mode; /* bml_distribution_mode_t; sequential only */
    int N;                 /* rows (and columns) */
    int M;                 /* slots per row */
    void *value;           /* device: N*M elements, row-major */
    int *index;            /* device: N*M column indices */
    int *nnz;              /* device: N live-slot counts */
    struct bml_hip_aux *aux; /* library-private bookkeeping */
} bml_matrix_ellpack_hip_t;

/* ELLBLOCK matrix resident in HBM.  Block slot (ib, jp) owns one bt x bt
 * tile, bt = 8, 16 or 32 (the smallest that holds every block size), with
 * the bsize[ib] x bsize[jb] block in its top-left corner and zeros
 * elsewhere: value[((ib*MB + jp)*bt + ii)*bt + jj].  The tile replaces the
 * per-block pointers into a memory pool of bml_matrix_ellblock_t
 * (ellblock/bml_types_ellblock.h:11-56); the first seven members keep that
 * struct's order. */
typedef struct bml_matrix_ellblock_hip_t {
    int matrix_type;       /* BML_HIP_TYPE_ELLBLOCK */
    int matrix_precision;  /* bml_matrix_precision_t */
    int distribution_mode; /* sequential only */
    int N;                 /* sum of bsize */
    int M;                 /* bml's nominal non-zeroes per row */
    int NB;                /* block rows */
    int MB;                /* block slots per block row */
    int bt;                /* tile edge */
    void *value;           /* device: NB*MB*bt*bt elements */
    int *indexb;           /* device: NB*MB block columns */
    int *nnzb;             /* device: NB live blocks per block row */
    int *bsize;            /* device: NB block sizes */
    int *bsize_host;       /* host copy of bsize */
    struct bml_hip_aux *aux; /* library-private bookkeeping */
} bml_matrix_ellblock_hip_t;

/* ---- runtime ------------------------------------------------------------ */

/* Select the device and create the library stream; called implicitly by the
 * first allocation (device 0).  No bml counterpart (bml_init.c initialises
 * MPI only). */
int bml_hip_init(int device);
void bml_hip_shutdown(void);
void bml_hip_set_error_mode(int mode);
void bml_hip_set_sync_mode(int mode);
/* Last error code (BML_HIP_OK if none); copies the message into buf. */
int bml_hip_last_error(char *buf, int len);
void bml_hip_clear_error(void);
/* Wait for all queued work and run the deferred overflow checks. */
void bml_hip_synchronize(void);
/* The hipStream_t every kernel of the library runs on. */
void *bml_hip_stream(void);
/* Library build string. */
const char *bml_hip_version(void);

/* Engine of the ELLPACK row products (x2, AB, multiply, and the SP2 drivers):
 *   EXACT      one wavefront per row sums each column in the reference's
 *              order (jp ascending, bml_multiply_ellpack_typed.c:280-293)
 *              and writes the kept columns in first-touch order: results
 *              bit-identical to the reference (the default);
 *   UNORDERED  rows of 64+ lanes: four or eight wavefronts share each
 *              row's table and add products with LDS atomics into double
 *              accumulators, so a column's sum is formed in arrival order
 *              (agrees with EXACT to rounding and may vary in the last bits
 *              between runs) and the kept columns are written in table
 *              order.  Rows of shorter groups use EXACT.  Squares (x2 and
 *              the SP2 iterations) whose rows touch more distinct columns
 *              than an LDS hash table holds (estimated from the previous
 *              square of the same matrix) use WINDOW;
 *   WINDOW     squares only, whatever their fill: the column range is cut
 *              into windows of 2^11..2^13 columns and each output row adds
 *              its products into a dense LDS window per window it touches
 *              (no hashing), summed in arrival order like UNORDERED, kept
 *              columns written sorted; other products run UNORDERED.  The
 *              windows of one product number at most 64 (N <= 2^19).
 * No bml counterpart. */
#define BML_HIP_ROW_ENGINE_EXACT 0
#define BML_HIP_ROW_ENGINE_UNORDERED 1
#define BML_HIP_ROW_ENGINE_WINDOW 2
void bml_hip_set_row_engine(int engine);
int bml_hip_get_row_engine(void);

/* Measured launch shapes.  With autotuning on (also BML_HIP_AUTOTUNE=1),
 * the first ELLPACK x2 or AB product (and SP2 iteration) of each kind of
 * operand -- precision, engine and the sizes in steps of 2^(1/4) -- times
 * candidate launch shapes on its own operands (lanes per row, table size,
 * bucket width and wavefronts; window width and rows per block on the
 * window engine) and keeps the fastest.  cache_path (also
 * BML_HIP_TUNE_CACHE) names a text file the shapes are read from and
 * appended to, so later processes reuse them; NULL keeps the current one.
 * Shapes set through the environment (BML_HIP_GROUP and the like) are not
 * tuned.  entries counts the shapes held; clear forgets them (the file is
 * left alone).  No bml counterpart. */
void bml_hip_set_autotune(int on, const char *cache_path);
int bml_hip_get_autotune(void);
int bml_hip_autotune_entries(void);
void bml_hip_autotune_clear(void);

/* ---- ELLPACK-R: allocation and conversion ------------------------------- */

/* bml_zero_matrix_ellpack (ellpack/bml_allocate_ellpack.c, typed
 * bml_allocate_ellpack_typed.c:190-262) */
bml_matrix_ellpack_hip_t *bml_zero_matrix_ellpack_hip(
    int precision, int N, int M, int distrib_mode);
/* bml_noinit_matrix_ellpack (bml_allocate_ellpack_typed.c:146-188) */
bml_matrix_ellpack_hip_t *bml_noinit_matrix_ellpack_hip(
    int precision, int N, int M, int distrib_mode);
/* bml_identity_matrix_ellpack (bml_allocate_ellpack_typed.c) */
bml_matrix_ellpack_hip_t *bml_identity_matrix_ellpack_hip(
    int precision, int N, int M, int distrib_mode);
/* bml_deallocate_ellpack (ellpack/bml_allocate_ellpack.c) */
void bml_deallocate_ellpack_hip(bml_matrix_ellpack_hip_t *A);
/* bml_clear_ellpack: every row emptied */
void bml_clear_ellpack_hip(bml_matrix_ellpack_hip_t *A);
/* bml_copy_ellpack_new / bml_copy_ellpack (ellpack/bml_copy_ellpack_typed.c);
 * B may have a different M as long as every row fits. */
bml_matrix_ellpack_hip_t *bml_copy_ellpack_hip_new(const bml_matrix_ellpack_hip_t *A);
void bml_copy_ellpack_hip(const bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B);
/* bml_import_from_dense_ellpack (ellpack/bml_import_ellpack.h:6-13,
 * typed ellpack/bml_import_ellpack_typed.c:28-76), same argument order:
 * entries with |a| > threshold are kept, in column order. */
bml_matrix_ellpack_hip_t *bml_import_from_dense_ellpack_hip(
    int precision, int order, int N, const void *A, double threshold, int M,
    int distrib_mode);
/* bml_export_to_dense_ellpack (ellpack/bml_export_ellpack_typed.c:18-73):
 * returns a malloc'd N*N host array. */
void *bml_export_to_dense_ellpack_hip(const bml_matrix_ellpack_hip_t *A, int order);
/* bml_convert_ellpack (ellpack/bml_convert_ellpack_typed.c:13-40, reached
 * through bml_convert.c:21-51 with matrix_type ellpack_hip): a new N x M
 * matrix whose row i holds every column j = 0..N-1 in order with A(i, j),
 * zeros included (bml_set_element for each element).  A is an ellpack_hip or
 * ellblock_hip matrix; elements are cast to precision as C converts them
 * (the reference reinterprets A's element bytes as the new precision).
 * N > M is the M-overflow error (the reference writes past the row). */
bml_matrix_ellpack_hip_t *bml_convert_ellpack_hip(const void *A, int precision, int M,
                                                  int distrib_mode);
/* Upload/download the ELLPACK-R arrays of a host bml ellpack matrix
 * (value/index/nnz of bml_matrix_ellpack_t, bml_types_ellpack.h:19-26).
 * src_M is the row stride of the host arrays; rows must fit A->M. */
void bml_ellpack_hip_set_arrays(bml_matrix_ellpack_hip_t *A, const void *value,
                                const int *index, const int *nnz, int src_M);
void bml_ellpack_hip_get_arrays(const bml_matrix_ellpack_hip_t *A, void *value,
                                int *index, int *nnz);
/* bml_write_bml_matrix / bml_read_bml_matrix for ellpack (bml_utilities.h:38-45,
 * ellpack/bml_utilities_ellpack_typed.c:22-190): Matrix Market coordinate
 * files in the reference's text format; read appends the entries to A's
 * rows (a "symmetric" file also stores the mirrors). */
void bml_write_bml_matrix_ellpack_hip(const bml_matrix_ellpack_hip_t *A, const char *filename);
void bml_read_bml_matrix_ellpack_hip(bml_matrix_ellpack_hip_t *A, const char *filename);

/* ---- ELLPACK-R: the thresholded multiply path --------------------------- */

/* bml_multiply_x2_ellpack (ellpack/bml_multiply_ellpack_typed.c:130-349):
 * X2 <- X*X; diagonal always stored, off-diagonals iff |v| > threshold;
 * returns malloc'd double[2] = {tr X, tr X^2}. */
void *bml_multiply_x2_ellpack_hip(bml_matrix_ellpack_hip_t *X,
                                  bml_matrix_ellpack_hip_t *X2, double threshold);
/* No counterpart in bml: bml_multiply_x2_ellpack_hip for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same product on the exact or unordered row
 * engine (never the window engine, never the symmetric shortcut), its
 * scratch taken from a fixed 64 MiB device array that the library holds on
 * each GPU until the process ends, tr X and tr X^2 written to traces[0..1]
 * (no host allocation either).  Inputs whose per-row records (36 bytes per
 * row) and one global table of the second pass (2 * min(M*M, N) + 64 slots
 * of value + 8 bytes) do not fit in it are refused with
 * BML_HIP_ERR_UNSUPPORTED.  Calls on one GPU share the array: calls that
 * overlap (several host threads or processes on one GPU) race. */
void bml_multiply_x2_ellpack_hip_noalloc(bml_matrix_ellpack_hip_t *X,
                                         bml_matrix_ellpack_hip_t *X2, double threshold,
                                         double *traces);
/* bml_multiply_ellpack (bml_multiply_ellpack_typed.c:55-118):
 * C <- alpha*A*B + beta*C. */
void bml_multiply_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                              bml_matrix_ellpack_hip_t *C, double alpha, double beta,
                              double threshold);
/* bml_multiply_AB_ellpack (bml_multiply_ellpack_typed.c:362-553): C <- A*B. */
void bml_multiply_AB_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                                 bml_matrix_ellpack_hip_t *C, double threshold);
/* bml_multiply_adjust_AB_ellpack (bml_multiply_ellpack_typed.c:566-707):
 * C <- A*B, threshold doubled until every row fits C->M. */
void bml_multiply_adjust_AB_ellpack_hip(bml_matrix_ellpack_hip_t *A,
                                        bml_matrix_ellpack_hip_t *B,
                                        bml_matrix_ellpack_hip_t *C, double threshold);
/* bml_add_ellpack (ellpack/bml_add_ellpack_typed.c:50-216):
 * A <- alpha*A + beta*B, every entry kept iff |v| > threshold. */
void bml_add_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                         double alpha, double beta, double threshold);
/* bml_add_norm_ellpack (bml_add_ellpack_typed.c:230-408): as add; returns
 * sum over the union of (A - B)^2 before the update. */
double bml_add_norm_ellpack_hip(bml_matrix_ellpack_hip_t *A, bml_matrix_ellpack_hip_t *B,
                                double alpha, double beta, double threshold);
/* bml_add_identity_ellpack (bml_add_ellpack_typed.c:420-549) */
void bml_add_identity_ellpack_hip(bml_matrix_ellpack_hip_t *A, double beta,
                                  double threshold);
/* bml_scale_add_identity_ellpack (bml_add_ellpack_typed.c:562-573) */
void bml_scale_add_identity_ellpack_hip(bml_matrix_ellpack_hip_t *A, double alpha,
                                        double beta, double threshold);
/* bml_scale_inplace_ellpack (ellpack/bml_scale_ellpack_typed.c:93-130):
 * scale_factor points to one element of A's precision. */
void bml_scale_inplace_ellpack_hip(const void *scale_factor, bml_matrix_ellpack_hip_t *A);
/* bml_threshold_ellpack (ellpack/bml_threshold_ellpack_typed.c:88-133) */
void bml_threshold_ellpack_hip(bml_matrix_ellpack_hip_t *A, double threshold);
/* bml_threshold_new_ellpack (bml_threshold_ellpack_typed.c:28-78) */
bml_matrix_ellpack_hip_t *bml_threshold_new_ellpack_hip(const bml_matrix_ellpack_hip_t *A,
                                                        double threshold);
/* bml_trace_ellpack (ellpack/bml_trace_ellpack_typed.c:29-92) */
double bml_trace_ellpack_hip(const bml_matrix_ellpack_hip_t *A);
/* bml_gershgorin_ellpack (ellpack/bml_normalize_ellpack_typed.c:83-175):
 * returns malloc'd double[2] = {min eigenvalue bound, max bound}. */
void *bml_gershgorin_ellpack_hip(const bml_matrix_ellpack_hip_t *A);
/* bml_normalize_ellpack (bml_normalize_ellpack_typed.c:33-45):
 * A <- (maxeval*I - A)/(maxeval - mineval). */
void bml_normalize_ellpack_hip(bml_matrix_ellpack_hip_t *A, double mineval,
                               double maxeval);
/* Row statistics of the last synchronised state: max and total nnz.
 * (no bml counterpart; bml_get_sparsity is the nearest,
 * ellpack/bml_getters_ellpack_typed.c) */
void bml_ellpack_hip_row_stats(const bml_matrix_ellpack_hip_t *A, int *max_nnz,
                               long long *total_nnz);


/* ---- ELLBLOCK ----------------------------------------------------------- */

/* Block products: the engine that multiplies bt x bt tiles.
 *   EXACT   each element's length-bsize dot product left to right, then
 *           added, in the reference's order (bml_multiply_block1..6,
 *           ellblock/bml_multiply_ellblock_typed.c:28-168): bit-identical
 *           to the reference's block kernels; any precision and tile;
 *   MFMA    matrix cores (v_mfma_f32_32x32x2f32 for single real,
 *           v_mfma_f64_16x16x4f64 for double real) on 32 x 32 tiles;
 *   MFMA_BF16  single real with inputs rounded to bfloat16 and FP32
 *           accumulation (v_mfma_f32_32x32x16_bf16); double-precision
 *           products under this setting use MFMA;
 *   AUTO    MFMA for real precisions with 32 x 32 tiles, else EXACT.
 * No bml counterpart: the reference picks its kernel by block size
 * (libxsmm above 6, :380). */
#define BML_HIP_BLOCK_ENGINE_AUTO -1
#define BML_HIP_BLOCK_ENGINE_EXACT 0
#define BML_HIP_BLOCK_ENGINE_MFMA 1
#define BML_HIP_BLOCK_ENGINE_MFMA_BF16 2
void bml_hip_set_block_engine(int engine);

/* Block-size registry: bml_set_block_sizes / bml_get_block_sizes /
 * bml_get_nb / bml_get_mb (ellblock/bml_allocate_ellblock.c:24-91).  Set at
 * most once (the reference asserts); unset, the default is blocks of 4 rows,
 * nb = ceil(N/4), mb = max(4, ceil(M/4)) capped at nb, the last block taking
 * the remainder. */
void bml_hip_set_block_sizes(const int *bsize, int nb, int mb);
const int *bml_hip_get_block_sizes(int N, int M);
int bml_hip_get_nb(void);
int bml_hip_get_mb(void);

/* bml_block_matrix_ellblock (ellblock/bml_allocate_ellblock_typed.c:176-243):
 * zero matrix with NB block rows of the given sizes (each <= 32) and MB
 * slots per block row; registers the sizes when none are set. */
bml_matrix_ellblock_hip_t *bml_block_matrix_ellblock_hip(
    int precision, int NB, int MB, int M, const int *bsize, int distrib_mode);
/* bml_zero_matrix_ellblock (bml_allocate_ellblock_typed.c:245-262): block
 * sizes from the registry */
bml_matrix_ellblock_hip_t *bml_zero_matrix_ellblock_hip(int precision, int N, int M,
                                                        int distrib_mode);
/* bml_deallocate_ellblock (ellblock/bml_allocate_ellblock.c) */
void bml_deallocate_ellblock_hip(bml_matrix_ellblock_hip_t *A);
/* bml_clear_ellblock (bml_allocate_ellblock_typed.c:127-145) */
void bml_clear_ellblock_hip(bml_matrix_ellblock_hip_t *A);
/* bml_copy_ellblock_new / bml_copy_ellblock (ellblock/bml_copy_ellblock_typed.c);
 * B may have a different MB as long as every block row fits. */
bml_matrix_ellblock_hip_t *bml_copy_ellblock_hip_new(const bml_matrix_ellblock_hip_t *A);
void bml_copy_ellblock_hip(const bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B);
/* Upload/download blocks as bmax x bmax host tiles per slot
 * (value[((ib*src_MB + jp)*bmax + ii)*bmax + jj]), block columns and
 * counts; bmax >= every block size. */
void bml_ellblock_hip_set_blocks(bml_matrix_ellblock_hip_t *A, const void *value,
                                 const int *indexb, const int *nnzb, int src_MB, int bmax);
void bml_ellblock_hip_get_blocks(const bml_matrix_ellblock_hip_t *A, void *value,
                                 int *indexb, int *nnzb, int bmax);
/* bml_write_bml_matrix / bml_read_bml_matrix for ellblock (bml_utilities.h:38-45,
 * ellblock/bml_utilities_ellblock_typed.c:24-260): Matrix Market coordinate
 * files in the reference's text format (every element of every stored
 * block); read adds each entry to its block, appending blocks as needed. */
void bml_write_bml_matrix_ellblock_hip(const bml_matrix_ellblock_hip_t *A, const char *filename);
void bml_read_bml_matrix_ellblock_hip(bml_matrix_ellblock_hip_t *A, const char *filename);
/* bml_multiply_x2_ellblock (ellblock/bml_multiply_ellblock_typed.c:246-442):
 * X2 <- X*X blockwise; a product block is kept iff it is the diagonal block
 * or its max |element| > threshold (no elementwise threshold); fatal if a
 * block row touches more than X2->MB blocks; returns malloc'd
 * double[2] = {tr X, tr X^2}. */
void *bml_multiply_x2_ellblock_hip(bml_matrix_ellblock_hip_t *X,
                                   bml_matrix_ellblock_hip_t *X2, double threshold);
/* bml_multiply_ellblock (bml_multiply_ellblock_typed.c:183-231):
 * C <- alpha*A*B + beta*C. */
void bml_multiply_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                               bml_matrix_ellblock_hip_t *C, double alpha, double beta,
                               double threshold);
/* bml_multiply_AB_ellblock (bml_multiply_ellblock_typed.c:455-621): C <- A*B */
void bml_multiply_AB_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                                  bml_matrix_ellblock_hip_t *C, double threshold);
/* bml_add_ellblock (ellblock/bml_add_ellblock_typed.c:35-197): A <- alpha*A +
 * beta*B blockwise, a block kept iff its max |element| > threshold. */
void bml_add_ellblock_hip(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                          double alpha, double beta, double threshold);
/* bml_threshold_ellblock (ellblock/bml_threshold_ellblock_typed.c:95-140):
 * drop blocks whose max |element| <= threshold, zero the small elements of
 * the others. */
void bml_threshold_ellblock_hip(bml_matrix_ellblock_hip_t *A, double threshold);
/* bml_trace_ellblock (ellblock/bml_trace_ellblock_typed.c:27-58) */
double bml_trace_ellblock_hip(const bml_matrix_ellblock_hip_t *A);
/* bml_gershgorin_ellblock (ellblock/bml_normalize_ellblock_typed.c:98-170):
 * returns malloc'd double[2] = {min eigenvalue bound, max bound}. */
void *bml_gershgorin_ellblock_hip(const bml_matrix_ellblock_hip_t *A);
/* bml_normalize_ellblock (bml_normalize_ellblock_typed.c:33-46):
 * A <- (maxeval*I - A)/(maxeval - mineval). */
void bml_normalize_ellblock_hip(bml_matrix_ellblock_hip_t *A, double mineval, double maxeval);
/* bml_scale_inplace_ellblock (ellblock/bml_scale_ellblock_typed.c:81-87):
 * scale_factor points to one element of A's precision. */
void bml_scale_inplace_ellblock_hip(const void *scale_factor, bml_matrix_ellblock_hip_t *A);
/* bml_add_identity_ellblock (ellblock/bml_add_ellblock_typed.c:354-370):
 * A <- A + beta*I through the blockwise add. */
void bml_add_identity_ellblock_hip(bml_matrix_ellblock_hip_t *A, double beta, double threshold);
/* bml_import_from_dense_ellblock (ellblock/bml_import_ellblock_typed.c:20-96):
 * a matrix with the registered block sizes (which must cover N rows) holding
 * the blocks of the dense N x N array whose max |element| exceeds threshold,
 * in block-column order; order 0 row-major, 1 column-major. */
bml_matrix_ellblock_hip_t *bml_import_from_dense_ellblock_hip(int precision, int order, int N,
                                                             const void *A, double threshold,
                                                             int M, int distrib_mode);
/* bml_export_to_dense_ellblock (ellblock/bml_export_ellblock_typed.c):
 * returns a malloc'd N*N host array. */
void *bml_export_to_dense_ellblock_hip(const bml_matrix_ellblock_hip_t *A, int order);
/* bml_convert_ellblock (ellblock/bml_convert_ellblock_typed.c:16-72): a new
 * matrix with the registered block sizes (bml_get_block_sizes(N, N)) and
 * MB = NB, every block of the block grid stored, zeros included, in
 * block-column order.  A and the element cast as for bml_convert_ellpack_hip. */
bml_matrix_ellblock_hip_t *bml_convert_ellblock_hip(const void *A, int precision, int M,
                                                    int distrib_mode);

/* ---- SP2 purification ----------------------------------------------------
 * No bml counterpart: client codes (PROGRESS, LATTE) build SP2 from
 * bml_gershgorin, bml_normalize, bml_multiply_x2 and bml_add_norm; this runs
 * the same steps resident on the GPU, the branch decision and convergence
 * test included (Niklasson, Phys. Rev. B 66, 155115 (2002)). */
typedef struct bml_hip_sp2_options {
    double threshold;  /* drop |x| <= threshold in every product and update */
    double tol;        /* stop when ||X - X^2||_F <= tol */
    int maxit;         /* iteration cap */
    int mixed;         /* 1: iterate in single precision, the last
                          refine_iters iterations in double */
    int refine_iters;  /* double-precision iterations at the end (mixed) */
    int poll;          /* read the convergence flag every poll iterations */
    int graph;         /* replay each iteration's kernels as a captured
                          hipGraph (1) or launch them one by one (0): batched
                          solves, and single ELLPACK solves with poll > 1 */
    int mixed_switch;  /* mixed: also move to double as soon as the single-
                          precision iterate stops improving (1, needs poll 1),
                          or only for the last refine_iters iterations (0) */
    int resume;        /* single solves: 1 continues the iteration from the
                          iterate X holds (no Gershgorin bounds, no
                          normalization; H is only tested for symmetry) */
} bml_hip_sp2_options;

typedef struct bml_hip_sp2_result {
    int iterations;        /* iterations run */
    int single_iterations; /* of which in single precision */
    int converged;         /* ||X - X^2||_F <= tol was reached */
    double trace_x;        /* tr X of the last iteration's input */
    double trace_x2;       /* tr X^2 of the last iteration */
    double idempotency;    /* ||X - X^2||_F of the last iteration */
    double emin, emax;     /* Gershgorin bounds of H */
} bml_hip_sp2_result;

void bml_hip_sp2_defaults(bml_hip_sp2_options *opt);
/* X <- density matrix of H for nocc electrons.  H and X are ellpack_hip
 * matrices of equal N and precision (double_real or double_complex when
 * opt->mixed); X->M bounds the fill.  history, if not NULL, receives
 * opt->maxit * 4 doubles: tr X, tr X^2, ||X - X^2||_F, branch (0: X <- X^2,
 * 1: X <- 2X - X^2) per iteration. */
void bml_hip_sp2_ellpack(const bml_matrix_ellpack_hip_t *H, bml_matrix_ellpack_hip_t *X,
                         double nocc, const bml_hip_sp2_options *opt,
                         bml_hip_sp2_result *res, double *history);
/* nsys independent SP2 solves at once (H[s], X[s] of any N; every X[s] with
 * the same M and precision): the systems are packed into one block-diagonal
 * ELLPACK-R matrix, every iteration runs one launch per phase over all rows
 * (rows of finished systems skip their work), and each system keeps its own
 * branch decisions, traces and convergence test.  res receives nsys
 * results.  No bml counterpart. */
void bml_hip_sp2_ellpack_batch(int nsys, bml_matrix_ellpack_hip_t *const *H,
                               bml_matrix_ellpack_hip_t *const *X, const double *nocc,
                               const bml_hip_sp2_options *opt, bml_hip_sp2_result *res);
/* The same on ELLBLOCK matrices (block products on the engine chosen by
 * bml_hip_set_block_engine; in mixed precision the single-precision
 * iterations use MFMA or, with BML_HIP_BLOCK_ENGINE_MFMA_BF16, bfloat16
 * inputs).  H and X share their block sizes; X->MB bounds the fill. */
void bml_hip_sp2_ellblock(const bml_matrix_ellblock_hip_t *H, bml_matrix_ellblock_hip_t *X,
                          double nocc, const bml_hip_sp2_options *opt,
                          bml_hip_sp2_result *res, double *history);

/* ---- profiling window -----------------------------------------------------
 * No bml counterpart (bml has no timers).  Between begin and end every
 * ELLPACK row-product kernel launch is timed with HIP events on the library
 * stream and its work is counted. */
typedef struct bml_hip_profile {
    /* per precision 1..4 (bml_matrix_precision_t; index 0 unused) */
    int launches[5];         /* row-product launches timed */
    double kernel_ms[5];     /* their summed durations */
    long long products[5];   /* partial products (sum over rows of A of nnz of the rows of B) */
    long long rows[5];       /* output rows produced */
    long long in_nnz[5];     /* entries of A's rows read */
    long long out_nnz[5];    /* entries written */
    long long ovf_rows[5];   /* rows whose LDS table overflowed (second pass) */
    long long touched[5];    /* distinct columns before thresholding */
    long long streamed[5];   /* entries of B's rows read: the products, or fewer
                                where several output rows share each read */
} bml_hip_profile;
void bml_hip_profile_begin(void);
void bml_hip_profile_end(bml_hip_profile *out);

#ifdef __cplusplus
}
#endif

#endif /* BML_HIP_H */

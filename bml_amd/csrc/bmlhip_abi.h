/*
 * Thin C ABI between the C host layer (ellpack_hip.c, ellblock_hip.c,
 * sp2_hip.c) and the HIP sources under hip/.  Plain pointers and sizes only;
 * device pointers are marked d_.  Every function returns 0 on success or a
 * BML_HIP_ERR_* code, with a message available from bmlhip_errmsg().
 */
#ifndef BMLHIP_ABI_H
#define BMLHIP_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status block resident in device memory, written by kernels, read back by
 * bmlhip_status_fetch() (one per process). */
typedef struct bmlhip_status {
    int m_overflow_rows;     /* rows that needed more slots than M */
    int first_m_overflow;    /* lowest such row (or INT_MAX) */
    int table_overflow_rows; /* rows handed to the global-table pass */
    int max_nnz;             /* largest row written by the last op */
    int max_touched;         /* largest pre-threshold row of the last product */
    int cursor;              /* work-queue cursor of the global-table pass */
    long long total_nnz;     /* entries written by the last op */
    double scalars[4];       /* trace X, trace X^2 / norm / gershgorin */
    /* accumulated over a profiling window (not reset per operation), per
     * precision 1..4 (index 0 unused) */
    long long prof_products[5]; /* partial products of row-product kernels */
    long long prof_rows[5];     /* rows those kernels produced */
    long long prof_in_nnz[5];   /* entries of A's rows they read */
    long long prof_out_nnz[5];  /* entries they wrote */
    long long prof_ovf_rows[5]; /* rows finished by the global-table pass */
    long long prof_touched[5];  /* distinct columns before thresholding */
    long long prof_streamed[5]; /* entries of B's rows read (the products, unless
                                   several output rows share each read) */
} bmlhip_status;

/* Per-row records: kernels store (kept, distinct, products, nnz of A's row)
 * per row as int4 into scratch slot BMLHIP_ROWSTAT_SLOT (kept = -1: row
 * skipped) and bmlhip_rowstat_reduce folds them into the status block, with
 * a few atomics per block instead of several per row; `what` is a mask of: */
#define BMLHIP_ROWSTAT_SLOT 7
#define BMLHIP_SYMFILL_SLOT 8 /* symmetric X^2: per-row mirror counters, kept zero */
#define BMLHIP_RS_OVF_KEPT 1 /* rows with kept > MC overflow */
#define BMLHIP_RS_OVF_DIST 2 /* rows with distinct > MC overflow */
#define BMLHIP_RS_STORED 4   /* max_nnz, total_nnz from min(kept, MC) */
#define BMLHIP_RS_PROF 8     /* profiling sums of precision `prof` */
#define BMLHIP_RS_OWN_STREAM 16 /* the kernel added its own streamed count */

/* ---- runtime (hip/runtime.hip) ----------------------------------------- */
int bmlhip_init(int device);
/* compute units of the current GPU (0 before init) */
int bmlhip_device_cus(void);
/* a timed region on the library stream: end waits for it and gives ms */
int bmlhip_time_begin(void);
int bmlhip_time_end(double *ms);
void bmlhip_shutdown(void);
const char *bmlhip_errmsg(void);
/* The fixed workspace of the entry points that allocate nothing: a
 * file-scope device array of BMLHIP_NOALLOC_BYTES, one per GPU, held until
 * the process ends; its address on the current device (taken once at
 * bmlhip_init), NULL before.  Calls that overlap on one GPU share it. */
#define BMLHIP_NOALLOC_BYTES ((size_t) 64 << 20)
void *bmlhip_noalloc_workspace(size_t *bytes);
void bmlhip_set_errmsg(const char *msg);
int bmlhip_malloc(void **d_ptr, size_t bytes);
void bmlhip_free(void *d_ptr);
int bmlhip_h2d(void *d_dst, const void *src, size_t bytes);
int bmlhip_d2h(void *dst, const void *d_src, size_t bytes);
int bmlhip_d2d(void *d_dst, const void *d_src, size_t bytes);
int bmlhip_zero(void *d_ptr, size_t bytes);
int bmlhip_sync(void);
void *bmlhip_stream(void);
/* Library-owned scratch, grown on demand; slot selects one of a few
 * independent buffers.  The pointer stays valid until the next call with
 * the same slot and a larger size. */
#define BMLHIP_SCRATCH_SLOTS 24
void *bmlhip_scratch(int slot, size_t bytes);
/* as bmlhip_scratch, zero-filled whenever the buffer is (re)allocated */
void *bmlhip_scratch_zeroed(int slot, size_t bytes);
bmlhip_status *bmlhip_status_dev(void);
/* reset the per-operation fields (not the overflow fields) */
int bmlhip_status_reset(void);
/* reset everything, overflow fields included */
int bmlhip_status_clear_errors(void);
int bmlhip_status_fetch(bmlhip_status *host);

/* ---- ELLPACK kernels (hip/ellpack.hip) ---------------------------------- */

/* C = A*B, thresholded (diagonal kept).  kind 0 = x2: per-row traces are
 * reduced into status->scalars[0..1] and the kept count is checked against
 * MC; kind 1 = AB: no traces, the touched count is checked against MC (as
 * the reference's multiply_AB does).  group = lanes per row (8..64),
 * table = hash slots per row in LDS (power of two, >= 64). */
/* Batched systems: one ELLPACK-R matrix holds nsys independent systems as
 * consecutive row ranges (their column indices offset likewise, so the
 * matrix is block diagonal and its products are the systems' products). */
typedef struct bmlhip_batch {
    int nsys;
    const int *d_row_sys;  /* [N] system of each row */
    const int *d_row_off;  /* [nsys + 1] first row of each system */
    int *d_done;           /* [nsys] converged: its rows are skipped */
    double *d_ab;          /* [2 nsys] alpha, beta of each system's pending add */
    double *d_trx, *d_trx2, *d_norm2; /* [nsys] tr X, tr X^2, ||X - X^2||_F^2 */
    double *d_rowtr;       /* [2 N] per-row trace partials */
    double *d_rownorm;     /* [N] per-row norm partials */
    /* work queue (all three set, or NULL: one block per row, rows of
     * finished systems return at once): the rows of the unfinished systems
     * in order, rebuilt on the device before each product; the row kernels
     * then run persistent, taking rows from it with an atomic cursor */
    int *d_list;           /* [N] */
    int *d_qinfo;          /* [4]: rows listed, product cursor, add cursor, systems left */
    int *d_act_off;        /* [nsys + 1] first list position of each system */
} bmlhip_batch;

typedef struct bmlhip_spgemm_args {
    int prec, kind, N;
    int MA, MB, MC;
    const void *d_Av;
    const int *d_Ai, *d_An;
    const void *d_Bv;
    const int *d_Bi, *d_Bn;
    void *d_Cv;
    int *d_Ci, *d_Cn;
    double threshold;
    int group;   /* lanes per output row: 8, 16, 32 or 64 */
    int table;   /* LDS hash-table slots per row (rounded up to a multiple of 8) */
    int bucket;  /* slots per probe: 2, 4 or 8 */
    int remap;   /* 1: XCD-aware block-to-row order */
    int waves;   /* group 64: wavefronts sharing one row's table (1 = one) */
    int engine;  /* BMLHIP_ROW_EXACT: sums in the reference's jp order, first-touch
                    output order; BMLHIP_ROW_UNORDERED (group 64): waves
                    wavefronts per row with LDS atomic adds, table order */
    int sym;     /* x2 of a symmetric / Hermitian X on the unordered or window
                    engine: rows form columns k >= i, the rest is mirrored */
    int wbits;   /* BMLHIP_ROW_WINDOW: log2 of the column window (10..13) */
    int wrows;   /* BMLHIP_ROW_WINDOW: output rows per block (1, 2 or 4) */
    const int *d_skip; /* if not NULL and *d_skip != 0 the kernels do nothing */
    const bmlhip_batch *batch; /* not NULL: per-system traces into batch->d_trx/d_trx2 */
    void *d_ws;       /* not NULL: the product takes its scratch from these
                         ws_bytes of device memory and allocates nothing
                         (exact and unordered engines, no batch, no sym) */
    size_t ws_bytes;
    int fuse;         /* 1: C <- beta*C + alpha*(A*B), C merged into the
                         product's epilogue (exact engine, kind 0 or 1, one
                         system, C distinct from A and B) */
    double alpha, beta;
} bmlhip_spgemm_args;
#define BMLHIP_ROW_EXACT 0
#define BMLHIP_ROW_UNORDERED 1
#define BMLHIP_ROW_WINDOW 2 /* x2 only (A = B): waves wavefronts per row add
                               into a dense LDS window of 2^wbits columns at a
                               time; rows come out sorted by column */
/* scratch slots of the window engine: packed rows, window offsets, window
 * masks, per-row norms */
#define BMLHIP_WIN_PACK_SLOT 12
#define BMLHIP_WIN_BOFF_SLOT 13
#define BMLHIP_WIN_MASK_SLOT 14
#define BMLHIP_WIN_NORM_SLOT 15
/* row blocks: union columns, coefficients, union sizes */
#define BMLHIP_WIN_UJ_SLOT 16
#define BMLHIP_WIN_UC_SLOT 17
#define BMLHIP_WIN_UN_SLOT 18
/* ELLBLOCK bf16 engine: bfloat16 copies of A's tiles and of B's, transposed */
#define BMLHIP_EB_A16_SLOT 19
#define BMLHIP_EB_B16_SLOT 20
/* window engine limits: at most 64 windows per row range */
#define BMLHIP_WIN_MAX_WINDOWS 64
int bmlhip_spgemm(const bmlhip_spgemm_args *a);
/* rebuild b's work queue (d_list, d_qinfo) from the done flags: queued
 * before each batched iteration's product (nothing when b has no queue) */
int bmlhip_batch_queue(const bmlhip_batch *b, int N);

/* A <- alpha*A + beta*B (mode 0), or that plus the norm over the union of
 * (A - B): mode 1 Re((A-B)^2) as bml_add_norm, mode 2 |A - B|^2; the norm
 * goes to status->scalars[2].  Entries kept iff |v| > threshold. */
int bmlhip_add(int prec, int N, int MA, void *d_Av, int *d_Ai, int *d_An,
               int MB, const void *d_Bv, const int *d_Bi, const int *d_Bn,
               double alpha, double beta, double threshold, int mode);
int bmlhip_rowstat_reduce(int N, int MC, int what, int prof, const int *d_skip);
/* *symmetric = 1 when every stored (i, j, v) has (j, i, conj(v)) stored,
 * equal by value, +0.0 == -0.0 (synchronous; not for use under graph capture) */
int bmlhip_is_symmetric(int prec, int N, int M, const void *d_Av, const int *d_Ai,
                        const int *d_An, int *symmetric);
/* as bmlhip_add with alpha = d_ab[0], beta = d_ab[1] read on the device and
 * nothing done when *d_skip != 0 */
int bmlhip_add_dev(int prec, int N, int MA, void *d_Av, int *d_Ai, int *d_An,
                   int MB, const void *d_Bv, const int *d_Bi, const int *d_Bn,
                   const double *d_ab, double threshold, int mode, const int *d_skip);

/* ---- profiling window ----------------------------------------------------- */
/* begin: zero the prof_* counters and time every row-product launch with HIP
 * events on the library stream; end: wait, return launches and total kernel
 * milliseconds (the prof_* counters come with bmlhip_status_fetch). */
int bmlhip_prof_begin(void);
int bmlhip_prof_end(int launches[5], double kernel_ms[5]);
/* event pair around one launch of a precision-prec kernel (launchers) */
void bmlhip_prof_mark_begin(int prec);
void bmlhip_prof_mark_end(void);

/* ---- SP2 (hip/sp2.hip) ---------------------------------------------------- */
/* Device state of one SP2 solve. */
typedef struct bmlhip_sp2_state {
    int it;          /* iterations done */
    int done;        /* converged: later iterations are no-ops */
    int branch;      /* last branch: 0 X <- X^2, 1 X <- 2X - X^2 */
    int max_nnz;     /* X's longest row after the last update */
    int max_touched; /* longest pre-threshold row of the last product */
    int max_nnz_x2;  /* longest row of the last X^2 */
    double ab[2];    /* alpha, beta of the pending add */
    double trx, trx2, norm; /* last tr X, tr X^2, ||X - X^2||_F */
} bmlhip_sp2_state;
/* After x2 (status scalars 0,1 hold tr X, tr X^2): choose the branch that
 * lands tr closer to nocc, store alpha/beta, record history[it*4..]. */
int bmlhip_sp2_decide(bmlhip_sp2_state *d_state, double nocc, double *d_hist, int maxit);
/* After add (status scalars[2] holds ||X - X^2||^2): record the norm, set
 * done when it is <= tol, advance it; touched = 1 also records the status
 * block's longest pre-threshold row (when the product ran after the
 * decision, as in the fused window step). */
int bmlhip_sp2_check(bmlhip_sp2_state *d_state, double tol, double *d_hist, int maxit,
                     int touched);
/* One SP2 iteration of a symmetric (Hermitian) X on the window engine, X
 * updated in place: tr X and tr X^2 = sum |X_ik|^2 reduced from X, the
 * branch chosen (bmlhip_sp2_decide), then one fused pass forms the upper
 * triangle of X^2 (kept as bml_multiply_x2 keeps it), writes alpha X + beta
 * X^2 thresholded as bml_add_norm does, mirrors it and reduces ||X - X^2||^2
 * into scalars[2]; then bmlhip_sp2_check.  Nothing runs once d_state->done. */
int bmlhip_sp2_window_step(int prec, int N, int M, void *d_Xv, int *d_Xi, int *d_Xn,
                           double threshold, int wbits, int waves, int wrows,
                           bmlhip_sp2_state *d_state, double nocc, double tol, double *d_hist,
                           int maxit);
/* Capture the kernels queued on the library stream between begin and end
 * into an executable graph (no allocation, memset or host sync may happen
 * in between); launch replays them on the library stream. */
int bmlhip_graph_begin(void);
int bmlhip_graph_end(void **exec);
int bmlhip_graph_launch(void *exec);
void bmlhip_graph_destroy(void *exec);

/* batched add with norm mode (1 or 2): alpha, beta per system from
 * b->d_ab, per-system norms into b->d_norm2 */
int bmlhip_add_batch(int prec, int N, int MA, void *d_Av, int *d_Ai, int *d_An,
                     int MB, const void *d_Bv, const int *d_Bi, const int *d_Bn,
                     double threshold, int mode, const bmlhip_batch *b);
/* per-system Gershgorin bounds of a batched matrix -> d_emin, d_emax [nsys] */
int bmlhip_gershgorin_batch(int prec, int N, int M, const void *d_Av, const int *d_Ai,
                            const int *d_An, const bmlhip_batch *b, double *d_emin,
                            double *d_emax);
/* per-system X0 = (emax I - X)/(emax - emin) in place (normalize) */
int bmlhip_normalize_batch(int prec, int N, int M, void *d_Av, int *d_Ai, int *d_An,
                           const bmlhip_batch *b, const double *d_emin, const double *d_emax);
/* per-system branch decision / convergence test (after x2 / after add);
 * d_alldone counts the finished systems; d_stats[0] receives the product's
 * longest pre-threshold row, d_stats[1] the longest row of X after the add */
int bmlhip_sp2_decide_batch(const bmlhip_batch *b, const double *d_nocc, int *d_branch,
                            int *d_stats);
int bmlhip_sp2_check_batch(const bmlhip_batch *b, double tol, int maxit, int *d_it,
                           double *d_norm, int *d_alldone, int *d_stats);
/* copy rows from (N, MA) to (N, MB) adding col_shift to every column index */
int bmlhip_copy_rows_shift(int prec, int N, int MA, const void *d_Av, const int *d_Ai,
                           const int *d_An, int MB, void *d_Bv, int *d_Bi, int *d_Bn,
                           int col_shift);
/* values of an ELLPACK-R matrix from one precision to another (same N, M) */
int bmlhip_convert(int prec_in, int prec_out, size_t count, const void *d_in, void *d_out);
/* the first d_cnt[r] * unit elements of each of nrows rows, stride elements
 * apart (the live slots of a matrix) */
int bmlhip_convert_rows(int prec_in, int prec_out, int nrows, size_t stride, size_t unit,
                        const int *d_cnt, const void *d_in, void *d_out);
/* A <- A + beta*I then threshold (add_identity); scale: values *= s first
 * (s_bits points to one element of the precision; NULL = no scaling). */
int bmlhip_scale_add_identity(int prec, int N, int M, void *d_Av, int *d_Ai,
                              int *d_An, const void *s_bits, double beta,
                              double threshold, int do_identity);
int bmlhip_threshold(int prec, int N, int M, void *d_Av, int *d_Ai, int *d_An,
                     double threshold);
/* trace -> status->scalars[0] */
int bmlhip_trace(int prec, int N, int M, const void *d_Av, const int *d_Ai,
                 const int *d_An);
/* gershgorin -> status->scalars[0] (min), [1] (max) */
int bmlhip_gershgorin(int prec, int N, int M, const void *d_Av, const int *d_Ai,
                      const int *d_An);
/* copy rows from an (N, MA) ELLPACK-R matrix to an (N, MB) one */
int bmlhip_copy_rows(int prec, int N, int MA, const void *d_Av, const int *d_Ai,
                     const int *d_An, int MB, void *d_Bv, int *d_Bi, int *d_Bn);
/* fill the diagonal: value 1, nnz 1 (identity) */
int bmlhip_identity(int prec, int N, int M, void *d_Av, int *d_Ai, int *d_An);
/* nnz statistics of a matrix -> status max_nnz / total_nnz */
int bmlhip_row_stats(int N, const int *d_An);

/* ---- ELLBLOCK: one bt x bt tile per block slot ------------------------------ */

typedef struct bmlhip_eb_args {
    int prec;
    int kind;           /* 0: x2 (traces -> scalars[0], [1]), 1: AB */
    int NB, bt;
    const int *d_bsize;
    int MA, MB, MC;     /* block slots per block row of A, B, C */
    const void *d_Av;
    const int *d_Ai, *d_An;
    const void *d_Bv;
    const int *d_Bi, *d_Bn;
    void *d_Cv;
    int *d_Ci, *d_Cn;
    double threshold;
    int engine;         /* 0 exact, 1 MFMA, 2 MFMA with bf16 inputs */
    const int *d_skip;  /* if not NULL and *d_skip != 0 the kernels do nothing */
    int sym;            /* x2 on an MFMA engine of an exactly symmetric (Hermitian)
                         * X: form the blocks kb >= ib, mirror the others */
    /* fused SP2 update (with sym): C <- alpha X + beta (X^2 thresholded),
     * alpha, beta from d_ab; blocks kept iff max |element| > threshold (the
     * diagonal block always); |X - X^2|^2 -> scalars[2]; C's trace and
     * sum |C_ij|^2 (the next iteration's tr X and tr X^2) -> scalars[0], [1] */
    int fuse;
    const double *d_ab;
} bmlhip_eb_args;
int bmlhip_eb_product(const bmlhip_eb_args *a);
/* tr X -> scalars[0] and the sum of |X_ij|^2 (tr X^2 of a symmetric or
 * Hermitian X) -> scalars[1] */
int bmlhip_eb_trace_frob(int prec, int NB, int bt, int MB, const void *d_Av, const int *d_Ai,
                         const int *d_An, const int *d_skip);
/* *symmetric = 1 iff every stored block (ib, jb) has a stored (jb, ib) whose
 * tile is its conjugate transpose, bit for bit */
int bmlhip_eb_is_symmetric(int prec, int NB, int bt, int MB, const void *d_Av, const int *d_Ai,
                           const int *d_An, int *symmetric);
/* A <- alpha*A + beta*B blockwise; mode 2 adds |A - B|^2 over the union
 * into scalars[2] (alpha, beta from d_ab when not NULL) */
int bmlhip_eb_add(int prec, int NB, int bt, const int *d_bsize, int MA, void *d_Av,
                  int *d_Ai, int *d_An, int MB, const void *d_Bv, const int *d_Bi,
                  const int *d_Bn, double alpha, double beta, double threshold, int mode,
                  const double *d_ab, const int *d_skip);
int bmlhip_eb_threshold(int prec, int NB, int bt, const int *d_bsize, int MB, void *d_Av,
                        int *d_Ai, int *d_An, double threshold);
/* trace -> scalars[0] */
int bmlhip_eb_trace(int prec, int NB, int bt, const int *d_bsize, int MB, const void *d_Av,
                    const int *d_Ai, const int *d_An);
/* copy block rows between matrices of different MB */
int bmlhip_eb_copy(int prec, int NB, int bt, int MA, const void *d_Av, const int *d_Ai,
                   const int *d_An, int MB, void *d_Bv, int *d_Bi, int *d_Bn);
/* deterministic sum of n doubles (fixed tree) -> status->scalars[which] */
int bmlhip_sum_to_status(const double *d_x, int n, int which);
/* min of lo[0..n) -> scalars[0], max of hi[0..n) -> scalars[1] (the max
 * starts at DBL_MIN, as bml's Gershgorin) */
int bmlhip_minmax_to_status(const double *d_lo, const double *d_hi, int n);
/* Gershgorin bounds of an ELLBLOCK matrix -> scalars[0] (min), [1] (max) */
int bmlhip_eb_gershgorin(int prec, int NB, int bt, const int *d_bsize, int MB,
                         const void *d_Av, const int *d_Ai, const int *d_An);
/* values of the live blocks *= s (s_bits: one element of the precision) */
int bmlhip_eb_scale(int prec, int NB, int bt, int MB, void *d_Av, const int *d_An,
                    const void *s_bits);
/* identity: one diagonal block per block row */
int bmlhip_eb_identity(int prec, int NB, int bt, const int *d_bsize, int MB, void *d_Av,
                       int *d_Ai, int *d_An);
/* block-row statistics -> status max_nnz / total_nnz */
int bmlhip_eb_row_stats(int NB, const int *d_An);

#ifdef __cplusplus
}
#endif

#endif /* BMLHIP_ABI_H */

/*
 * SP2 density-matrix purification resident on the GPU.
 *
 * The steps are those a bml client runs (Gershgorin bounds, normalize, then
 * per iteration multiply_x2 and add_norm) with the iteration's branch
 * decision and convergence test done by one-thread kernels on the device,
 * so the host only polls a flag.  Mixed precision: the iterations run in
 * single precision except the last opt->refine_iters, which continue in
 * double from the converted single-precision iterate; traces and norms are
 * always reduced in double.
 */
#include "bml_hip_host.h"
#include "bml_hip_tune.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void bml_hip_sp2_defaults(bml_hip_sp2_options *o)
{
    o->threshold = 1e-5;
    o->tol = 1e-8;
    o->maxit = 60;
    o->mixed = 0;
    o->refine_iters = 4;
    o->poll = 1;
    o->graph = 1;
    o->mixed_switch = 1;
    o->resume = 0;
}

static int single_of(int prec)
{
    return prec == BML_HIP_DOUBLE_REAL ? BML_HIP_SINGLE_REAL
           : prec == BML_HIP_DOUBLE_COMPLEX ? BML_HIP_SINGLE_COMPLEX : 0;
}

/* Stall test of the single-precision phase: the idempotency measure read at
 * every iteration no longer falls below its value two iterations earlier
 * (SP2 alternates branches, and a converging iterate shrinks it at least
 * every other step; once rounding dominates, it grows), or, once it is below
 * 1e-5 sqrt(nocc) (the quadratic regime, where every step shrinks it), it
 * grows at all.  Needs poll == 1. */
typedef struct {
    int seen;
    double prev1, prev2;
} stall_track;

static int stalled(const bml_hip_sp2_options *o, stall_track *t, double norm, double nocc)
{
    if (!o->mixed_switch || o->poll != 1)
        return 0;
    const int stall = (t->seen >= 2 && norm >= t->prev2) ||
                      (t->seen >= 1 && norm >= t->prev1 && t->prev1 < 1e-5 * sqrt(nocc));
    t->prev2 = t->prev1;
    t->prev1 = norm;
    t->seen++;
    return stall;
}

/* one SP2 iteration on (X, *X2p), all queued on the library stream; *X2p
 * (the product's storage, of X's shape) is allocated on first use: the
 * fused window step updates X in place and needs none */
typedef struct {
    bml_matrix_ellpack_hip_t *X, **X2p;
    double threshold;
    int sym;
} sp2_tune_ctx;

/* the autotuner's measurement for an SP2 iteration: X^2 (the symmetric
 * form when the solve uses it) into X2, timed; X is left as it is */
static int measure_sp2(const bmlh_spgemm_shape *sh, void *ctx, double *ms)
{
    sp2_tune_ctx *c = ctx;
    bml_matrix_ellpack_hip_t *X = c->X;
    if (!*c->X2p && !(*c->X2p = bml_noinit_matrix_ellpack_hip(X->matrix_precision, X->N, X->M,
                                                              X->distribution_mode)))
        return 1;
    bml_matrix_ellpack_hip_t *X2 = *c->X2p;
    bmlhip_spgemm_args a;
    memset(&a, 0, sizeof a);
    a.prec = X->matrix_precision;
    a.N = X->N;
    a.MA = a.MB = X->M;
    a.MC = X2->M;
    a.d_Av = a.d_Bv = X->value;
    a.d_Ai = a.d_Bi = X->index;
    a.d_An = a.d_Bn = X->nnz;
    a.d_Cv = X2->value;
    a.d_Ci = X2->index;
    a.d_Cn = X2->nnz;
    a.threshold = c->threshold;
    a.group = sh->group;
    a.table = sh->table;
    a.bucket = sh->bucket;
    a.remap = sh->remap;
    a.waves = sh->waves;
    a.engine = sh->engine;
    a.wbits = sh->wbits;
    a.wrows = sh->wrows;
    a.sym = c->sym;
    if (bmlhip_time_begin())
        return 1;
    const int rc = bmlhip_spgemm(&a);
    const int rt = bmlhip_time_end(ms);
    return rc || rt;
}

/* launch shape of an SP2 iteration on X (the tuner reads the row counts
 * the last poll recorded; with autotuning on, measured) */
static void sp2_shape(bml_matrix_ellpack_hip_t *X, bml_matrix_ellpack_hip_t **X2p,
                      double threshold, int sym, bmlh_spgemm_shape *sh)
{
    const int mx = X->aux->stats_valid ? X->aux->max_nnz : X->M;
    sp2_tune_ctx c = { X, X2p, threshold, sym };
    bmlh_tuned_spgemm(X->matrix_precision, X->N, mx, mx, X->aux->touched_hint,
                      X->aux->hint_max_nnz, BMLH_PICK_SQUARE, sh, measure_sp2, &c);
}

static int same_shape(const bmlh_spgemm_shape *a, const bmlh_spgemm_shape *b)
{
    return a->group == b->group && a->table == b->table && a->bucket == b->bucket &&
           a->remap == b->remap && a->waves == b->waves && a->engine == b->engine &&
           a->wbits == b->wbits && a->wrows == b->wrows;
}

static int sp2_iteration(bml_matrix_ellpack_hip_t *X, bml_matrix_ellpack_hip_t **X2p,
                         bmlhip_sp2_state *d_state, double *d_hist,
                         const bml_hip_sp2_options *o, double nocc, int sym,
                         const bmlh_spgemm_shape *shp, const char *what)
{
    const bmlh_spgemm_shape sh = *shp;
    bmlhip_spgemm_args a;
    memset(&a, 0, sizeof a);
    a.prec = X->matrix_precision;
    a.kind = 0;
    a.N = X->N;
    a.MA = a.MB = X->M;
    a.d_Av = a.d_Bv = X->value;
    a.d_Ai = a.d_Bi = X->index;
    a.d_An = a.d_Bn = X->nnz;
    a.threshold = o->threshold;
    a.d_skip = &d_state->done;
    a.group = sh.group;
    a.table = sh.table;
    a.bucket = sh.bucket;
    a.remap = sh.remap;
    a.waves = sh.waves;
    a.engine = sh.engine;
    a.wbits = sh.wbits;
    a.wrows = sh.wrows;
    a.sym = sym;
    if (sh.engine == BMLHIP_ROW_WINDOW && sym)
        /* one fused pass: X^2's upper triangle, the update and the norm,
         * X rewritten in place (X2 is not used) */
        return bmlh_ok(bmlhip_sp2_window_step(X->matrix_precision, X->N, X->M, X->value,
                                              X->index, X->nnz, o->threshold, sh.wbits,
                                              sh.waves, sh.wrows, d_state, nocc, o->tol, d_hist,
                                              o->maxit),
                       what);
    if (!*X2p && !(*X2p = bml_noinit_matrix_ellpack_hip(X->matrix_precision, X->N, X->M,
                                                        X->distribution_mode)))
        return 0;
    bml_matrix_ellpack_hip_t *X2 = *X2p;
    a.MC = X2->M;
    a.d_Cv = X2->value;
    a.d_Ci = X2->index;
    a.d_Cn = X2->nnz;
    return bmlh_ok(bmlhip_spgemm(&a), what) &&
           bmlh_ok(bmlhip_sp2_decide(d_state, nocc, d_hist, o->maxit), what) &&
           bmlh_ok(bmlhip_add_dev(X->matrix_precision, X->N, X->M, X->value, X->index, X->nnz,
                                  X2->M, X2->value, X2->index, X2->nnz, d_state->ab,
                                  o->threshold, 2, &d_state->done), what) &&
           bmlh_ok(bmlhip_sp2_check(d_state, o->tol, d_hist, o->maxit, 0), what);
}

void bml_hip_sp2_ellpack(const bml_matrix_ellpack_hip_t *H, bml_matrix_ellpack_hip_t *X,
                         double nocc, const bml_hip_sp2_options *opt,
                         bml_hip_sp2_result *res, double *history)
{
    const char *what = "bml_hip_sp2_ellpack";
    bml_hip_sp2_options o;
    if (opt)
        o = *opt;
    else
        bml_hip_sp2_defaults(&o);
    if (o.poll < 1)
        o.poll = 1;
    if (o.maxit < 1)
        o.maxit = 1;
    if (!H || !X || H->matrix_type != BML_HIP_TYPE_ELLPACK ||
        X->matrix_type != BML_HIP_TYPE_ELLPACK || H->N != X->N ||
        H->matrix_precision != X->matrix_precision || H == X) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: H and X must be distinct ellpack_hip matrices "
                   "of equal size and precision", what);
        return;
    }
    const int prec = X->matrix_precision, N = X->N, M = X->M;
    const int sprec = single_of(prec);
    if (o.mixed && !sprec) {
        bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: mixed precision needs a double-precision X",
                   what);
        return;
    }
    bml_hip_sp2_result r;
    memset(&r, 0, sizeof r);

    /* X0 = (emax I - H) / (emax - emin), unless X holds the iterate to
     * continue from */
    if (!o.resume) {
        double *g = bml_gershgorin_ellpack_hip(H);
        if (!g)
            return;
        r.emin = g[0];
        r.emax = g[1];
        free(g);
    }
    /* A symmetric (Hermitian) H gives symmetric iterates: on the unordered
     * engine, X^2 then forms its upper triangle and mirrors it */
    int sym = 0;
    if (bml_hip_get_row_engine() != BML_HIP_ROW_ENGINE_EXACT &&
        !bmlh_ok(bmlhip_is_symmetric(H->matrix_precision, H->N, H->M, H->value, H->index,
                                     H->nnz, &sym), what))
        return;
    if (!o.resume) {
        bml_copy_ellpack_hip(H, X);
        if (bmlh_failed())
            return;
        bml_normalize_ellpack_hip(X, r.emin, r.emax);
        if (bmlh_failed())
            return;
    }

    bml_matrix_ellpack_hip_t *X2 = NULL, *Xs = NULL, *X2s = NULL;
    bmlhip_sp2_state *d_state = NULL;
    double *d_hist = NULL;
    void *graph = NULL;
    const bml_matrix_ellpack_hip_t *gcur = NULL;
    bmlh_spgemm_shape gshape;
    memset(&gshape, 0, sizeof gshape);
    if (!bmlh_ok(bmlhip_malloc((void **) &d_state, sizeof(bmlhip_sp2_state)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &d_hist, sizeof(double) * 4 * (size_t) o.maxit), what))
        goto out;
    bmlhip_zero(d_state, sizeof(bmlhip_sp2_state));
    bmlhip_zero(d_hist, sizeof(double) * 4 * (size_t) o.maxit);

    int single = 0;
    if (o.mixed && o.maxit > o.refine_iters) {
        Xs = bml_noinit_matrix_ellpack_hip(sprec, N, M, X->distribution_mode);
        if (!Xs)
            goto out;
        if (!bmlh_ok(bmlhip_convert(prec, sprec, (size_t) N * M, X->value, Xs->value), what) ||
            !bmlh_ok(bmlhip_d2d(Xs->index, X->index, (size_t) N * M * sizeof(int)), what) ||
            !bmlh_ok(bmlhip_d2d(Xs->nnz, X->nnz, (size_t) N * sizeof(int)), what))
            goto out;
        *Xs->aux = *X->aux;
        single = 1;
    }

    bmlhip_sp2_state st;
    memset(&st, 0, sizeof st);
    stall_track stall = { 0, 0.0, 0.0 };
    for (int it = 0; it < o.maxit; it++) {
        bml_matrix_ellpack_hip_t *cur = single ? Xs : X;
        bmlh_spgemm_shape sh;
        sp2_shape(cur, single ? &X2s : &X2, o.threshold, sym, &sh);
        if (o.graph && o.poll > 1) {
            /* between two polls every iteration is the same launches: replay
             * them as a graph, captured anew when the iterate or its launch
             * shape changes (the first iteration of each runs eagerly, which
             * sizes every scratch buffer and allocates X2; the captured
             * kernels are not run by the capture) */
            if (!graph || gcur != cur || !same_shape(&sh, &gshape)) {
                bmlhip_graph_destroy(graph);
                graph = NULL;
                if (!sp2_iteration(cur, single ? &X2s : &X2, d_state, d_hist, &o, nocc, sym, &sh,
                                   what))
                    goto out;
                if (!bmlh_ok(bmlhip_graph_begin(), what))
                    goto out;
                const int okc = sp2_iteration(cur, single ? &X2s : &X2, d_state, d_hist, &o, nocc,
                                              sym, &sh, what);
                void *ge = NULL;
                const int oke = bmlh_ok(bmlhip_graph_end(&ge), what);
                if (!okc || !oke)
                    goto out;
                graph = ge;
                gshape = sh;
                gcur = cur;
            } else if (!bmlh_ok(bmlhip_graph_launch(graph), what)) {
                goto out;
            }
        } else if (!sp2_iteration(cur, single ? &X2s : &X2, d_state, d_hist, &o, nocc, sym, &sh,
                                  what)) {
            goto out;
        }
        int last_single = single && it + 1 >= o.maxit - o.refine_iters;
        if ((it + 1) % o.poll == 0 || last_single || it + 1 == o.maxit) {
            if (!bmlh_ok(bmlhip_d2h(&st, d_state, sizeof st), what))
                goto out;
            /* the tuner reads max_nnz; the totals are refreshed at the end */
            cur->aux->touched_hint = st.max_touched;
            cur->aux->hint_max_nnz = cur->aux->stats_valid ? cur->aux->max_nnz : cur->M;
            cur->aux->stats_valid = 1;
            cur->aux->max_nnz = st.max_nnz;
            if (single)
                r.single_iterations = st.it;
            if (st.done)
                break;
            if (single && stalled(&o, &stall, st.norm, nocc))
                last_single = 1;
        }
        if (last_single) {
            /* continue in double from the single-precision iterate */
            if (!bmlh_ok(bmlhip_convert(sprec, prec, (size_t) N * M, Xs->value, X->value), what) ||
                !bmlh_ok(bmlhip_d2d(X->index, Xs->index, (size_t) N * M * sizeof(int)), what) ||
                !bmlh_ok(bmlhip_d2d(X->nnz, Xs->nnz, (size_t) N * sizeof(int)), what))
                goto out;
            *X->aux = *Xs->aux;
            single = 0;
        }
    }
    if (!bmlh_ok(bmlhip_d2h(&st, d_state, sizeof st), what))
        goto out;
    if (single) {
        /* converged while still in single precision: hand back in X */
        if (!bmlh_ok(bmlhip_convert(sprec, prec, (size_t) N * M, Xs->value, X->value), what) ||
            !bmlh_ok(bmlhip_d2d(X->index, Xs->index, (size_t) N * M * sizeof(int)), what) ||
            !bmlh_ok(bmlhip_d2d(X->nnz, Xs->nnz, (size_t) N * sizeof(int)), what))
            goto out;
        *X->aux = *Xs->aux;
    }
    {
        bmlhip_status s;
        if (!bmlh_finish_sync(what, &s))
            goto out;
    }
    X->aux->stats_valid = 0; /* row totals recomputed on demand */
    r.iterations = st.it;
    r.converged = st.done;
    r.trace_x = st.trx;
    r.trace_x2 = st.trx2;
    r.idempotency = st.norm;
    if (history)
        bmlh_ok(bmlhip_d2h(history, d_hist, sizeof(double) * 4 * (size_t) o.maxit), what);
    if (res)
        *res = r;
out:
    bmlhip_graph_destroy(graph);
    bmlhip_free(d_state);
    bmlhip_free(d_hist);
    bml_deallocate_ellpack_hip(X2);
    bml_deallocate_ellpack_hip(Xs);
    bml_deallocate_ellpack_hip(X2s);
}

/* ---- batched solves --------------------------------------------------------- */

typedef struct {
    bmlhip_batch b;
    int *row_sys, *row_off, *done, *branch, *it, *alldone, *stats;
    double *ab, *trx, *trx2, *norm2, *norm, *rowtr, *rownorm, *nocc, *emin, *emax;
    int *list, *qinfo, *act_off; /* the row work queue */
    int sym; /* every H symmetric (Hermitian): symmetric x2 on the unordered engine */
} batch_dev;

static void batch_free(batch_dev *d)
{
    void *p[] = { d->row_sys, d->row_off, d->done, d->branch, d->it, d->alldone, d->stats,
                  d->ab, d->trx, d->trx2, d->norm2, d->norm, d->rowtr, d->rownorm, d->nocc,
                  d->emin, d->emax, d->list, d->qinfo, d->act_off };
    for (size_t k = 0; k < sizeof p / sizeof p[0]; k++)
        bmlhip_free(p[k]);
}

static int batch_alloc(batch_dev *d, int nsys, int N, const char *what)
{
    memset(d, 0, sizeof *d);
    const size_t S = (size_t) nsys, R = (size_t) N;
    int ok = bmlh_ok(bmlhip_malloc((void **) &d->row_sys, R * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->row_off, (S + 1) * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->done, S * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->branch, S * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->it, S * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->alldone, 2 * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->stats, 2 * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->ab, 2 * S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->trx, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->trx2, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->norm2, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->norm, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->rowtr, 2 * R * sizeof(double) + 8), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->rownorm, R * sizeof(double) + 8), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->nocc, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->emin, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->emax, S * sizeof(double)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->list, R * sizeof(int) + 8), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->qinfo, 4 * sizeof(int)), what) &&
             bmlh_ok(bmlhip_malloc((void **) &d->act_off, (S + 1) * sizeof(int)), what);
    if (!ok)
        return 0;
    bmlhip_zero(d->done, S * sizeof(int));
    bmlhip_zero(d->branch, S * sizeof(int));
    bmlhip_zero(d->it, S * sizeof(int));
    bmlhip_zero(d->alldone, 2 * sizeof(int));
    bmlhip_zero(d->stats, 2 * sizeof(int));
    bmlhip_zero(d->ab, 2 * S * sizeof(double));
    bmlhip_zero(d->trx, S * sizeof(double));
    bmlhip_zero(d->trx2, S * sizeof(double));
    bmlhip_zero(d->norm2, S * sizeof(double));
    bmlhip_zero(d->norm, S * sizeof(double));
    d->b.nsys = nsys;
    d->b.d_row_sys = d->row_sys;
    d->b.d_row_off = d->row_off;
    d->b.d_done = d->done;
    d->b.d_ab = d->ab;
    d->b.d_trx = d->trx;
    d->b.d_trx2 = d->trx2;
    d->b.d_norm2 = d->norm2;
    d->b.d_rowtr = d->rowtr;
    d->b.d_rownorm = d->rownorm;
    /* The work queue is on for the exact engine and off for the others
     * (their batched iterations ran faster with one block per row: 147
     * against 129 solves/s on the 512-system batch); BML_HIP_BATCH_QUEUE=1
     * or 0 sets it either way. */
    const char *e = getenv("BML_HIP_BATCH_QUEUE");
    const int queue = e ? atoi(e) != 0 : bml_hip_get_row_engine() == BML_HIP_ROW_ENGINE_EXACT;
    if (queue) {
        d->b.d_list = d->list;
        d->b.d_qinfo = d->qinfo;
        d->b.d_act_off = d->act_off;
    }
    return 1;
}

/* one batched iteration on (X, X2), all queued on the library stream */
static int batch_iteration(bml_matrix_ellpack_hip_t *X, bml_matrix_ellpack_hip_t *X2,
                           batch_dev *d, const bml_hip_sp2_options *o,
                           const bmlh_spgemm_shape *sh, const char *what)
{
    bmlhip_spgemm_args a;
    memset(&a, 0, sizeof a);
    a.prec = X->matrix_precision;
    a.kind = 0;
    a.N = X->N;
    a.MA = a.MB = X->M;
    a.MC = X2->M;
    a.d_Av = a.d_Bv = X->value;
    a.d_Ai = a.d_Bi = X->index;
    a.d_An = a.d_Bn = X->nnz;
    a.d_Cv = X2->value;
    a.d_Ci = X2->index;
    a.d_Cn = X2->nnz;
    a.threshold = o->threshold;
    a.group = sh->group;
    a.table = sh->table;
    a.bucket = sh->bucket;
    a.remap = sh->remap;
    a.waves = sh->waves;
    a.engine = sh->engine;
    a.wbits = sh->wbits;
    a.wrows = sh->wrows;
    a.sym = d->sym;
    a.batch = &d->b;
    return bmlh_ok(bmlhip_batch_queue(&d->b, X->N), what) && bmlh_ok(bmlhip_spgemm(&a), what) &&
           bmlh_ok(bmlhip_sp2_decide_batch(&d->b, d->nocc, d->branch, d->stats), what) &&
           bmlh_ok(bmlhip_add_batch(X->matrix_precision, X->N, X->M, X->value, X->index, X->nnz,
                                    X2->M, X2->value, X2->index, X2->nnz, o->threshold, 2,
                                    &d->b), what) &&
           bmlh_ok(bmlhip_sp2_check_batch(&d->b, o->tol, o->maxit, d->it, d->norm, d->alldone,
                                          d->stats), what);
}

void bml_hip_sp2_ellpack_batch(int nsys, bml_matrix_ellpack_hip_t *const *H,
                               bml_matrix_ellpack_hip_t *const *X, const double *nocc,
                               const bml_hip_sp2_options *opt, bml_hip_sp2_result *res)
{
    const char *what = "bml_hip_sp2_ellpack_batch";
    bml_hip_sp2_options o;
    if (opt)
        o = *opt;
    else
        bml_hip_sp2_defaults(&o);
    if (o.poll < 1)
        o.poll = 1;
    if (o.maxit < 1)
        o.maxit = 1;
    if (nsys < 1 || !H || !X || !nocc) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: nsys=%d", what, nsys);
        return;
    }
    const int prec = X[0] ? X[0]->matrix_precision : 0, M = X[0] ? X[0]->M : 0;
    long long Ntot = 0;
    for (int s = 0; s < nsys; s++) {
        if (!H[s] || !X[s] || H[s]->matrix_type != BML_HIP_TYPE_ELLPACK ||
            X[s]->matrix_type != BML_HIP_TYPE_ELLPACK || H[s]->N != X[s]->N ||
            H[s]->matrix_precision != prec || X[s]->matrix_precision != prec || X[s]->M != M ||
            H[s] == X[s]) {
            bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: system %d: H and X must be distinct "
                       "ellpack_hip matrices of equal N, one precision and one M for every X",
                       what, s);
            return;
        }
        Ntot += H[s]->N;
    }
    if (Ntot >= (1ll << 24)) {
        bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: %lld rows in all exceed 2^24", what, Ntot);
        return;
    }
    const int N = (int) Ntot;
    const int sprec = single_of(prec);
    if (o.mixed && !sprec) {
        bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: mixed precision needs double precision", what);
        return;
    }
    int *row_off = malloc(sizeof(int) * ((size_t) nsys + 1));
    int *row_sys = malloc(sizeof(int) * ((size_t) N + 1));
    int *h_it = malloc(sizeof(int) * (size_t) nsys);
    int *h_done = malloc(sizeof(int) * (size_t) nsys);
    double *h_d = malloc(sizeof(double) * 5 * (size_t) nsys);
    batch_dev d;
    memset(&d, 0, sizeof d);
    bml_matrix_ellpack_hip_t *Xc = NULL, *X2c = NULL, *Xs = NULL, *X2s = NULL;
    void *graph = NULL;
    bmlh_spgemm_shape gshape;
    memset(&gshape, 0, sizeof gshape);
    if (!row_off || !row_sys || !h_it || !h_done || !h_d) {
        bmlh_error(BML_HIP_ERR_ALLOC, "%s: host allocation failed", what);
        goto out;
    }
    row_off[0] = 0;
    for (int s = 0; s < nsys; s++) {
        row_off[s + 1] = row_off[s] + H[s]->N;
        for (int i = row_off[s]; i < row_off[s + 1]; i++)
            row_sys[i] = s;
    }
    if (!batch_alloc(&d, nsys, N, what))
        goto out;
    if (!bmlh_ok(bmlhip_h2d(d.row_off, row_off, sizeof(int) * ((size_t) nsys + 1)), what) ||
        !bmlh_ok(bmlhip_h2d(d.row_sys, row_sys, sizeof(int) * (size_t) N), what) ||
        !bmlh_ok(bmlhip_h2d(d.nocc, nocc, sizeof(double) * (size_t) nsys), what))
        goto out;

    /* pack H into one block-diagonal matrix with X's M, then per-system
     * Gershgorin bounds and X0 = (emax I - H)/(emax - emin) */
    Xc = bml_zero_matrix_ellpack_hip(prec, N, M, X[0]->distribution_mode);
    X2c = bml_noinit_matrix_ellpack_hip(prec, N, M, X[0]->distribution_mode);
    if (!Xc || !X2c)
        goto out;
    {
        const size_t es = bmlh_elem_size(prec);
        int hmax = 0;
        for (int s = 0; s < nsys; s++) {
            const size_t r0 = (size_t) row_off[s];
            if (!bmlh_ok(bmlhip_copy_rows_shift(prec, H[s]->N, H[s]->M, H[s]->value,
                                                H[s]->index, H[s]->nnz, M,
                                                (char *) Xc->value + r0 * M * es,
                                                Xc->index + r0 * M, Xc->nnz + r0, row_off[s]),
                         what))
                goto out;
            const int hm = H[s]->aux->stats_valid ? H[s]->aux->max_nnz : H[s]->M;
            if (hm > hmax)
                hmax = hm;
        }
        bmlhip_status st;
        if (!bmlh_finish_sync(what, &st))
            goto out;
        Xc->aux->stats_valid = 1;
        Xc->aux->max_nnz = hmax + 1;
    }
    d.sym = 0;
    if (bml_hip_get_row_engine() != BML_HIP_ROW_ENGINE_EXACT &&
        !bmlh_ok(bmlhip_is_symmetric(prec, N, M, Xc->value, Xc->index, Xc->nnz, &d.sym), what))
        goto out;
    if (!bmlh_ok(bmlhip_gershgorin_batch(prec, N, M, Xc->value, Xc->index, Xc->nnz, &d.b,
                                         d.emin, d.emax), what) ||
        !bmlh_ok(bmlhip_normalize_batch(prec, N, M, Xc->value, Xc->index, Xc->nnz, &d.b,
                                        d.emin, d.emax), what))
        goto out;

    int single = 0;
    if (o.mixed && o.maxit > o.refine_iters) {
        Xs = bml_noinit_matrix_ellpack_hip(sprec, N, M, Xc->distribution_mode);
        X2s = bml_noinit_matrix_ellpack_hip(sprec, N, M, Xc->distribution_mode);
        if (!Xs || !X2s)
            goto out;
        if (!bmlh_ok(bmlhip_convert(prec, sprec, (size_t) N * M, Xc->value, Xs->value), what) ||
            !bmlh_ok(bmlhip_d2d(Xs->index, Xc->index, (size_t) N * M * sizeof(int)), what) ||
            !bmlh_ok(bmlhip_d2d(Xs->nnz, Xc->nnz, (size_t) N * sizeof(int)), what))
            goto out;
        *Xs->aux = *Xc->aux;
        single = 1;
    }

    int touched = 0, hint_max = 0, stats[2] = { 0, 0 }, alldone[2] = { 0, 0 };
    for (int it = 0; it < o.maxit; it++) {
        bml_matrix_ellpack_hip_t *cur = single ? Xs : Xc, *cur2 = single ? X2s : X2c;
        const int mx = cur->aux->stats_valid ? cur->aux->max_nnz : cur->M;
        bmlh_spgemm_shape sh;
        bmlh_pick_spgemm(cur->matrix_precision, N, mx, mx, touched, hint_max, BMLH_PICK_SQUARE,
                         &sh);
        if (o.graph) {
            /* replay the iteration as a graph; capture anew when the launch
             * shape or the precision changes (the first pass runs eagerly so
             * every scratch buffer is sized before capture) */
            if (!graph || !same_shape(&sh, &gshape)) {
                bmlhip_graph_destroy(graph);
                graph = NULL;
                if (!batch_iteration(cur, cur2, &d, &o, &sh, what))
                    goto out;
                if (!bmlh_ok(bmlhip_graph_begin(), what))
                    goto out;
                const int okc = batch_iteration(cur, cur2, &d, &o, &sh, what);
                void *ge = NULL;
                const int oke = bmlh_ok(bmlhip_graph_end(&ge), what);
                if (!okc || !oke)
                    goto out;
                graph = ge;
                gshape = sh;
                /* the captured kernels were not run: the eager pass above was
                 * this iteration */
            } else if (!bmlh_ok(bmlhip_graph_launch(graph), what)) {
                goto out;
            }
        } else if (!batch_iteration(cur, cur2, &d, &o, &sh, what)) {
            goto out;
        }
        const int last_single = single && it + 1 >= o.maxit - o.refine_iters;
        if ((it + 1) % o.poll == 0 || last_single || it + 1 == o.maxit) {
            if (!bmlh_ok(bmlhip_d2h(alldone, d.alldone, sizeof alldone), what) ||
                !bmlh_ok(bmlhip_d2h(stats, d.stats, sizeof stats), what))
                goto out;
            touched = stats[0];
            hint_max = mx;
            cur->aux->stats_valid = 1;
            cur->aux->max_nnz = stats[1];
            if (alldone[0] >= nsys)
                break;
        }
        if (last_single) {
            if (!bmlh_ok(bmlhip_convert(sprec, prec, (size_t) N * M, Xs->value, Xc->value), what) ||
                !bmlh_ok(bmlhip_d2d(Xc->index, Xs->index, (size_t) N * M * sizeof(int)), what) ||
                !bmlh_ok(bmlhip_d2d(Xc->nnz, Xs->nnz, (size_t) N * sizeof(int)), what))
                goto out;
            *Xc->aux = *Xs->aux;
            single = 0;
            bmlhip_graph_destroy(graph);
            graph = NULL;
        }
    }
    if (single) {
        if (!bmlh_ok(bmlhip_convert(sprec, prec, (size_t) N * M, Xs->value, Xc->value), what) ||
            !bmlh_ok(bmlhip_d2d(Xc->index, Xs->index, (size_t) N * M * sizeof(int)), what) ||
            !bmlh_ok(bmlhip_d2d(Xc->nnz, Xs->nnz, (size_t) N * sizeof(int)), what))
            goto out;
    }
    /* unpack every system's rows into its X */
    {
        const size_t es = bmlh_elem_size(prec);
        for (int s = 0; s < nsys; s++) {
            const size_t r0 = (size_t) row_off[s];
            if (!bmlh_ok(bmlhip_copy_rows_shift(prec, X[s]->N, M, (char *) Xc->value + r0 * M * es,
                                                Xc->index + r0 * M, Xc->nnz + r0, M, X[s]->value,
                                                X[s]->index, X[s]->nnz, -row_off[s]), what))
                goto out;
            X[s]->aux->stats_valid = 0;
        }
        bmlhip_status st;
        if (!bmlh_finish_sync(what, &st))
            goto out;
    }
    if (res) {
        if (!bmlh_ok(bmlhip_d2h(h_it, d.it, sizeof(int) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_done, d.done, sizeof(int) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_d, d.trx, sizeof(double) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_d + nsys, d.trx2, sizeof(double) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_d + 2 * nsys, d.norm, sizeof(double) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_d + 3 * nsys, d.emin, sizeof(double) * (size_t) nsys), what) ||
            !bmlh_ok(bmlhip_d2h(h_d + 4 * nsys, d.emax, sizeof(double) * (size_t) nsys), what))
            goto out;
        for (int s = 0; s < nsys; s++) {
            bml_hip_sp2_result *r = &res[s];
            memset(r, 0, sizeof *r);
            r->iterations = h_it[s];
            r->single_iterations = o.mixed ? (h_it[s] < o.maxit - o.refine_iters
                                                  ? h_it[s] : o.maxit - o.refine_iters)
                                           : 0;
            r->converged = h_done[s] == 1;
            r->trace_x = h_d[s];
            r->trace_x2 = h_d[nsys + s];
            r->idempotency = h_d[2 * nsys + s];
            r->emin = h_d[3 * nsys + s];
            r->emax = h_d[4 * nsys + s];
        }
    }
out:
    bmlhip_graph_destroy(graph);
    batch_free(&d);
    bml_deallocate_ellpack_hip(Xc);
    bml_deallocate_ellpack_hip(X2c);
    bml_deallocate_ellpack_hip(Xs);
    bml_deallocate_ellpack_hip(X2s);
    free(row_off);
    free(row_sys);
    free(h_it);
    free(h_done);
    free(h_d);
}

/* ---- ELLBLOCK ------------------------------------------------------------- */

static int eb_sp2_iteration(bml_matrix_ellblock_hip_t *X, bml_matrix_ellblock_hip_t *X2,
                            bmlhip_sp2_state *d_state, double *d_hist,
                            const bml_hip_sp2_options *o, double nocc, int sym,
                            const char *what)
{
    bmlhip_eb_args a;
    memset(&a, 0, sizeof a);
    a.prec = X->matrix_precision;
    a.kind = 0;
    a.NB = X->NB;
    a.bt = X->bt;
    a.d_bsize = X->bsize;
    a.MA = a.MB = X->MB;
    a.MC = X2->MB;
    a.d_Av = a.d_Bv = X->value;
    a.d_Ai = a.d_Bi = X->indexb;
    a.d_An = a.d_Bn = X->nnzb;
    a.d_Cv = X2->value;
    a.d_Ci = X2->indexb;
    a.d_Cn = X2->nnzb;
    a.threshold = o->threshold;
    a.engine = bmlh_block_engine_for(X);
    a.d_skip = &d_state->done;
    a.sym = sym;
    return bmlh_ok(bmlhip_eb_product(&a), what) &&
           bmlh_ok(bmlhip_sp2_decide(d_state, nocc, d_hist, o->maxit), what) &&
           bmlh_ok(bmlhip_eb_add(X->matrix_precision, X->NB, X->bt, X->bsize, X->MB, X->value,
                                 X->indexb, X->nnzb, X2->MB, X2->value, X2->indexb, X2->nnzb,
                                 0.0, 0.0, o->threshold, 2, d_state->ab, &d_state->done),
                   what) &&
           bmlh_ok(bmlhip_sp2_check(d_state, o->tol, d_hist, o->maxit, 0), what);
}

/* One iteration with the update fused into the product (MFMA engines, an
 * exactly symmetric X): the branch is decided from tr X and
 * tr X^2 = sum |X_ij|^2 before the product, which then writes
 * alpha X + beta X^2 into X2, reduces |X - X^2|^2 and leaves the new
 * iterate's two traces for the next decision (have_tr: the previous
 * iteration was fused, so they are in place; otherwise a pre-pass reduces
 * them from X); the caller swaps the two matrices' buffers. */
static int eb_sp2_iteration_fused(bml_matrix_ellblock_hip_t *X, bml_matrix_ellblock_hip_t *X2,
                                  bmlhip_sp2_state *d_state, double *d_hist,
                                  const bml_hip_sp2_options *o, double nocc, int have_tr,
                                  const char *what)
{
    bmlhip_eb_args a;
    memset(&a, 0, sizeof a);
    a.prec = X->matrix_precision;
    a.kind = 0;
    a.NB = X->NB;
    a.bt = X->bt;
    a.d_bsize = X->bsize;
    a.MA = a.MB = X->MB;
    a.MC = X2->MB;
    a.d_Av = a.d_Bv = X->value;
    a.d_Ai = a.d_Bi = X->indexb;
    a.d_An = a.d_Bn = X->nnzb;
    a.d_Cv = X2->value;
    a.d_Ci = X2->indexb;
    a.d_Cn = X2->nnzb;
    a.threshold = o->threshold;
    a.engine = bmlh_block_engine_for(X);
    a.d_skip = &d_state->done;
    a.sym = 1;
    a.fuse = 1;
    a.d_ab = d_state->ab;
    return (have_tr ||
            bmlh_ok(bmlhip_eb_trace_frob(X->matrix_precision, X->NB, X->bt, X->MB, X->value,
                                         X->indexb, X->nnzb, &d_state->done),
                    what)) &&
           bmlh_ok(bmlhip_sp2_decide(d_state, nocc, d_hist, o->maxit), what) &&
           bmlh_ok(bmlhip_eb_product(&a), what) &&
           bmlh_ok(bmlhip_sp2_check(d_state, o->tol, d_hist, o->maxit, 1), what);
}

static void eb_swap_buffers(bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B)
{
    void *v = A->value;
    int *i = A->indexb, *n = A->nnzb;
    A->value = B->value;
    A->indexb = B->indexb;
    A->nnzb = B->nnzb;
    B->value = v;
    B->indexb = i;
    B->nnzb = n;
}

/* values, block columns and counts of an ELLBLOCK matrix into one of the
 * same shape in another precision */
static int eb_convert(const bml_matrix_ellblock_hip_t *A, bml_matrix_ellblock_hip_t *B,
                      const char *what)
{
    const size_t slots = (size_t) A->NB * A->MB, tile = (size_t) A->bt * A->bt;
    return bmlh_ok(bmlhip_convert_rows(A->matrix_precision, B->matrix_precision, A->NB,
                                       (size_t) A->MB * tile, tile, A->nnzb, A->value, B->value),
                   what) &&
           bmlh_ok(bmlhip_d2d(B->indexb, A->indexb, slots * sizeof(int)), what) &&
           bmlh_ok(bmlhip_d2d(B->nnzb, A->nnzb, (size_t) A->NB * sizeof(int)), what);
}

void bml_hip_sp2_ellblock(const bml_matrix_ellblock_hip_t *H, bml_matrix_ellblock_hip_t *X,
                          double nocc, const bml_hip_sp2_options *opt,
                          bml_hip_sp2_result *res, double *history)
{
    const char *what = "bml_hip_sp2_ellblock";
    bml_hip_sp2_options o;
    if (opt)
        o = *opt;
    else
        bml_hip_sp2_defaults(&o);
    if (o.poll < 1)
        o.poll = 1;
    if (o.maxit < 1)
        o.maxit = 1;
    if (!H || !X || H->matrix_type != BML_HIP_TYPE_ELLBLOCK ||
        X->matrix_type != BML_HIP_TYPE_ELLBLOCK || H->NB != X->NB || H->bt != X->bt ||
        H->matrix_precision != X->matrix_precision || H == X ||
        memcmp(H->bsize_host, X->bsize_host, sizeof(int) * (size_t) H->NB) != 0) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "%s: H and X must be distinct ellblock_hip matrices "
                   "of equal block structure and precision", what);
        return;
    }
    const int prec = X->matrix_precision, NB = X->NB, MB = X->MB;
    const int sprec = single_of(prec);
    if (o.mixed && !sprec) {
        bmlh_error(BML_HIP_ERR_UNSUPPORTED, "%s: mixed precision needs a double-precision X",
                   what);
        return;
    }
    bml_hip_sp2_result r;
    memset(&r, 0, sizeof r);
    const int engine0 = bmlh_block_engine_setting(); /* restored at the end */
    double *g = bml_gershgorin_ellblock_hip(H);
    if (!g)
        return;
    r.emin = g[0];
    r.emax = g[1];
    free(g);
    /* A symmetric H gives symmetric iterates: on the MFMA engines X^2 then
     * forms the blocks kb >= ib and mirrors them */
    int sym = 0;
    if (bmlh_block_engine_for(H) != BML_HIP_BLOCK_ENGINE_EXACT &&
        !bmlh_ok(bmlhip_eb_is_symmetric(H->matrix_precision, H->NB, H->bt, H->MB, H->value,
                                        H->indexb, H->nnzb, &sym), what))
        return;
    bml_copy_ellblock_hip(H, X);
    if (bmlh_failed())
        return;
    bml_normalize_ellblock_hip(X, r.emin, r.emax);
    if (bmlh_failed())
        return;

    bml_matrix_ellblock_hip_t *X2 = NULL, *Xs = NULL, *X2s = NULL;
    bmlhip_sp2_state *d_state = NULL;
    double *d_hist = NULL;
    /* work matrices: every tile they hold is written before it is read */
    X2 = bmlh_block_matrix_ellblock_work(prec, NB, MB, X->M, X->bsize_host, X->distribution_mode);
    if (!X2)
        goto out;
    if (!bmlh_ok(bmlhip_malloc((void **) &d_state, sizeof(bmlhip_sp2_state)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &d_hist, sizeof(double) * 4 * (size_t) o.maxit), what))
        goto out;
    bmlhip_zero(d_state, sizeof(bmlhip_sp2_state));
    bmlhip_zero(d_hist, sizeof(double) * 4 * (size_t) o.maxit);

    int single = 0;
    if (o.mixed && o.maxit > o.refine_iters) {
        Xs = bmlh_block_matrix_ellblock_work(sprec, NB, MB, X->M, X->bsize_host,
                                             X->distribution_mode);
        X2s = bmlh_block_matrix_ellblock_work(sprec, NB, MB, X->M, X->bsize_host,
                                              X->distribution_mode);
        if (!Xs || !X2s || !eb_convert(X, Xs, what))
            goto out;
        single = 1;
    }

    bmlhip_sp2_state st;
    memset(&st, 0, sizeof st);
    stall_track stall = { 0, 0.0, 0.0 };
    /* bf16 inputs: once the bf16 iterate stalls, or nears the bf16 rounding
     * floor, the single-precision phase goes on with FP32 MFMA until that
     * stalls too, then double (FP32 iterations bring it far closer to
     * idempotency for less than the double-precision ones would cost) */
    int bf16 = single && bmlh_block_engine_for(Xs) == BML_HIP_BLOCK_ENGINE_MFMA_BF16;
    double recover = 0.0; /* the measure when bf16 gave way to FP32 (0: none pending) */
    /* the update fused into the product: MFMA engines on a symmetric H,
     * with the state read after every iteration (so no iteration runs, or
     * swaps buffers, after convergence); BML_HIP_EB_FUSE=0 turns it off */
    const char *fe = getenv("BML_HIP_EB_FUSE");
    const int fuse_ok = sym && o.poll == 1 && !(fe && atoi(fe) == 0);
    void *x_value0 = X->value;
    int have_tr = 0; /* the status holds the iterate's traces (a fused iteration left them) */
    for (int it = 0; it < o.maxit; it++) {
        bml_matrix_ellblock_hip_t *cur = single ? Xs : X, *cur2 = single ? X2s : X2;
        if (fuse_ok && bmlh_block_engine_for(cur) != BML_HIP_BLOCK_ENGINE_EXACT) {
            if (!eb_sp2_iteration_fused(cur, cur2, d_state, d_hist, &o, nocc, have_tr, what))
                goto out;
            eb_swap_buffers(cur, cur2);
            have_tr = 1;
        } else {
            if (!eb_sp2_iteration(cur, cur2, d_state, d_hist, &o, nocc, sym, what))
                goto out;
            have_tr = 0;
        }
        int last_single = single && it + 1 >= o.maxit - o.refine_iters;
        if ((it + 1) % o.poll == 0 || last_single || it + 1 == o.maxit) {
            if (!bmlh_ok(bmlhip_d2h(&st, d_state, sizeof st), what))
                goto out;
            if (single)
                r.single_iterations = st.it;
            if (st.done)
                break;
            /* bf16 inputs round each entry to 2^-8 of itself, which leaves
             * X^2 about 8 * 2^-8 * sqrt(nocc) from idempotent in the
             * Frobenius norm (config 3: the bf16 iterate diverges once the
             * measure is below ~3 = 4 * 2^-8 * sqrt(32768), and FP32 then
             * spends iterations climbing back): below that level the phase
             * goes on in FP32 */
            const int bf16_floor = bf16 && o.mixed_switch && st.norm < 8.0 * 0x1p-8 * sqrt(nocc);
            /* after the switch to FP32 the measure first climbs while FP32
             * corrects the bf16 iterate; the FP32 stall test waits until it
             * is back below its value at the switch */
            const int armed = !(recover > 0.0 && st.norm >= recover);
            if (armed)
                recover = 0.0;
            if (single && (bf16_floor || (armed && stalled(&o, &stall, st.norm, nocc)))) {
                if (bf16) {
                    bf16 = 0;
                    bml_hip_set_block_engine(BML_HIP_BLOCK_ENGINE_MFMA);
                    stall.seen = 0;
                    recover = st.norm;
                } else {
                    last_single = 1;
                }
            }
        }
        if (last_single) {
            if (!eb_convert(Xs, X, what))
                goto out;
            single = 0;
        }
    }
    if (!bmlh_ok(bmlhip_d2h(&st, d_state, sizeof st), what))
        goto out;
    if (single && !eb_convert(Xs, X, what))
        goto out;
    if (X->value != x_value0) {
        /* the iterate lies in the buffers X started with and X2 holds:
         * copy it there and give X its own buffers back */
        bml_copy_ellblock_hip(X, X2);
        if (bmlh_failed())
            goto out;
        eb_swap_buffers(X, X2);
    }
    {
        bmlhip_status s;
        if (!bmlh_finish_sync(what, &s))
            goto out;
    }
    X->aux->stats_valid = 0;
    r.iterations = st.it;
    r.converged = st.done;
    r.trace_x = st.trx;
    r.trace_x2 = st.trx2;
    r.idempotency = st.norm;
    if (history)
        bmlh_ok(bmlhip_d2h(history, d_hist, sizeof(double) * 4 * (size_t) o.maxit), what);
    if (res)
        *res = r;
out:
    if (bmlh_block_engine_setting() != engine0)
        bml_hip_set_block_engine(engine0);
    bmlhip_free(d_state);
    bmlhip_free(d_hist);
    bml_deallocate_ellblock_hip(X2);
    bml_deallocate_ellblock_hip(Xs);
    bml_deallocate_ellblock_hip(X2s);
}

void bml_hip_profile_begin(void)
{
    if (bmlh_ready())
        return;
    bmlh_ok(bmlhip_prof_begin(), "bml_hip_profile_begin");
}

void bml_hip_profile_end(bml_hip_profile *out)
{
    bml_hip_profile p;
    memset(&p, 0, sizeof p);
    bmlhip_status st;
    if (bmlh_ok(bmlhip_prof_end(p.launches, p.kernel_ms), "bml_hip_profile_end") &&
        bmlh_ok(bmlhip_status_fetch(&st), "bml_hip_profile_end")) {
        for (int k = 0; k < 5; k++) {
            p.products[k] = st.prof_products[k];
            p.rows[k] = st.prof_rows[k];
            p.in_nnz[k] = st.prof_in_nnz[k];
            p.out_nnz[k] = st.prof_out_nnz[k];
            p.ovf_rows[k] = st.prof_ovf_rows[k];
            p.touched[k] = st.prof_touched[k];
            p.streamed[k] = st.prof_streamed[k];
        }
    }
    if (out)
        *out = p;
}

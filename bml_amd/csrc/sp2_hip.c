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
}

static int single_of(int prec)
{
    return prec == BML_HIP_DOUBLE_REAL ? BML_HIP_SINGLE_REAL
           : prec == BML_HIP_DOUBLE_COMPLEX ? BML_HIP_SINGLE_COMPLEX : 0;
}

/* one SP2 iteration on (X, X2), all queued on the library stream */
static int sp2_iteration(bml_matrix_ellpack_hip_t *X, bml_matrix_ellpack_hip_t *X2,
                         bmlhip_sp2_state *d_state, double *d_hist,
                         const bml_hip_sp2_options *o, double nocc, const char *what)
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
    a.d_skip = &d_state->done;
    int mx = X->aux->stats_valid ? X->aux->max_nnz : X->M;
    bmlh_spgemm_shape sh;
    bmlh_pick_spgemm(a.prec, X->N, mx, mx, X->aux->touched_hint, X->aux->hint_max_nnz, &sh);
    a.group = sh.group;
    a.table = sh.table;
    a.bucket = sh.bucket;
    a.remap = sh.remap;
    return bmlh_ok(bmlhip_spgemm(&a), what) &&
           bmlh_ok(bmlhip_sp2_decide(d_state, nocc, d_hist, o->maxit), what) &&
           bmlh_ok(bmlhip_add_dev(X->matrix_precision, X->N, X->M, X->value, X->index, X->nnz,
                                  X2->M, X2->value, X2->index, X2->nnz, d_state->ab,
                                  o->threshold, 2, &d_state->done), what) &&
           bmlh_ok(bmlhip_sp2_check(d_state, o->tol, d_hist, o->maxit), what);
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

    /* X0 = (emax I - H) / (emax - emin) */
    double *g = bml_gershgorin_ellpack_hip(H);
    if (!g)
        return;
    r.emin = g[0];
    r.emax = g[1];
    free(g);
    bml_copy_ellpack_hip(H, X);
    if (bmlh_failed())
        return;
    bml_normalize_ellpack_hip(X, r.emin, r.emax);
    if (bmlh_failed())
        return;

    bml_matrix_ellpack_hip_t *X2 = NULL, *Xs = NULL, *X2s = NULL;
    bmlhip_sp2_state *d_state = NULL;
    double *d_hist = NULL;
    X2 = bml_noinit_matrix_ellpack_hip(prec, N, M, X->distribution_mode);
    if (!X2)
        goto out;
    if (!bmlh_ok(bmlhip_malloc((void **) &d_state, sizeof(bmlhip_sp2_state)), what) ||
        !bmlh_ok(bmlhip_malloc((void **) &d_hist, sizeof(double) * 4 * (size_t) o.maxit), what))
        goto out;
    bmlhip_zero(d_state, sizeof(bmlhip_sp2_state));
    bmlhip_zero(d_hist, sizeof(double) * 4 * (size_t) o.maxit);

    int single = 0;
    if (o.mixed && o.maxit > o.refine_iters) {
        Xs = bml_noinit_matrix_ellpack_hip(sprec, N, M, X->distribution_mode);
        X2s = bml_noinit_matrix_ellpack_hip(sprec, N, M, X->distribution_mode);
        if (!Xs || !X2s)
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
    for (int it = 0; it < o.maxit; it++) {
        bml_matrix_ellpack_hip_t *cur = single ? Xs : X, *cur2 = single ? X2s : X2;
        if (!sp2_iteration(cur, cur2, d_state, d_hist, &o, nocc, what))
            goto out;
        const int last_single = single && it + 1 >= o.maxit - o.refine_iters;
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
    bmlhip_free(d_state);
    bmlhip_free(d_hist);
    bml_deallocate_ellpack_hip(X2);
    bml_deallocate_ellpack_hip(Xs);
    bml_deallocate_ellpack_hip(X2s);
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
        }
    }
    if (out)
        *out = p;
}

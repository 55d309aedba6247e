/*
 * Runtime entry points of the HIP backend: initialisation, error and
 * synchronisation policy.
 *
 * Error policy mirrors bml's logger: an error is fatal, printed with an
 * "ERROR" prefix and ends the process with EXIT_FAILURE
 * (src/C-interface/bml_logger.c:57-117).  BML_HIP_ERRORS_RECORD keeps the
 * message instead so that a host language can raise it.
 */
#include "bml_hip_host.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int g_error_mode = BML_HIP_ERRORS_EXIT;
static int g_sync_mode = BML_HIP_SYNC_EACH_CALL;
static int g_last_code = BML_HIP_OK;
static char g_last_msg[1024];
static int g_ready = 0;

void bmlh_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_msg, sizeof g_last_msg, fmt, ap);
    va_end(ap);
    g_last_code = code;
    if (g_error_mode == BML_HIP_ERRORS_EXIT) {
        fprintf(stderr, "[ERROR] bml_hip: %s\n", g_last_msg);
        fflush(stderr);
        exit(EXIT_FAILURE);
    }
}

int bmlh_failed(void)
{
    return g_last_code != BML_HIP_OK;
}

int bmlh_ok(int rc, const char *what)
{
    if (rc == 0)
        return 1;
    bmlh_error(rc, "%s: %s", what, bmlhip_errmsg());
    return 0;
}

int bmlh_ready(void)
{
    if (g_ready)
        return 0;
    const char *dev = getenv("BML_HIP_DEVICE");
    int rc = bmlhip_init(dev ? atoi(dev) : 0);
    if (rc) {
        bmlh_error(rc, "device initialisation failed: %s", bmlhip_errmsg());
        return rc;
    }
    g_ready = 1;
    return 0;
}

int bmlh_sync_mode(void)
{
    return g_sync_mode;
}

static int check_status(const char *what, bmlhip_status *st)
{
    if (!bmlh_ok(bmlhip_status_fetch(st), what))
        return 0;
    if (st->m_overflow_rows > 0) {
        int rows = st->m_overflow_rows, first = st->first_m_overflow;
        bmlhip_status_clear_errors();
        bmlh_error(BML_HIP_ERR_M_OVERFLOW,
                   "%s: Number of non-zeroes per row > M, Increase M "
                   "(%d rows, first row %d)", what, rows, first);
        return 0;
    }
    return 1;
}

int bmlh_finish(const char *what, bmlhip_status *st)
{
    if (g_sync_mode != BML_HIP_SYNC_EACH_CALL)
        return 0;
    return check_status(what, st);
}

int bmlh_finish_sync(const char *what, bmlhip_status *st)
{
    return check_status(what, st);
}

size_t bmlh_elem_size(int precision)
{
    switch (precision) {
    case BML_HIP_SINGLE_REAL: return 4;
    case BML_HIP_DOUBLE_REAL: return 8;
    case BML_HIP_SINGLE_COMPLEX: return 8;
    case BML_HIP_DOUBLE_COMPLEX: return 16;
    }
    return 0;
}

/* T-typed scalar s for precision prec into buf */
void bmlh_put_scalar(int prec, double s, void *buf)
{
    switch (prec) {
    case BML_HIP_SINGLE_REAL: {
        float f = (float) s;
        memcpy(buf, &f, 4);
        break;
    }
    case BML_HIP_DOUBLE_REAL:
        memcpy(buf, &s, 8);
        break;
    case BML_HIP_SINGLE_COMPLEX: {
        float f[2] = { (float) s, 0.0f };
        memcpy(buf, f, 8);
        break;
    }
    case BML_HIP_DOUBLE_COMPLEX: {
        double d[2] = { s, 0.0 };
        memcpy(buf, d, 16);
        break;
    }
    }
}

struct bml_hip_aux *bmlh_aux_new(void)
{
    return (struct bml_hip_aux *) calloc(1, sizeof(struct bml_hip_aux));
}

void bmlh_set_stats(struct bml_hip_aux *aux, const bmlhip_status *st)
{
    if (!aux)
        return;
    if (st) {
        aux->stats_valid = 1;
        aux->max_nnz = st->max_nnz;
        aux->total_nnz = st->total_nnz;
    } else {
        aux->stats_valid = 0;
    }
}

/* ---- public runtime API ------------------------------------------------ */

int bml_hip_init(int device)
{
    if (g_ready)
        return 0;
    int rc = bmlhip_init(device);
    if (rc) {
        bmlh_error(rc, "device initialisation failed: %s", bmlhip_errmsg());
        return rc;
    }
    g_ready = 1;
    return 0;
}

void bml_hip_shutdown(void)
{
    if (!g_ready)
        return;
    bmlhip_shutdown();
    g_ready = 0;
}

void bml_hip_set_error_mode(int mode)
{
    g_error_mode = mode == BML_HIP_ERRORS_RECORD ? BML_HIP_ERRORS_RECORD
                                                 : BML_HIP_ERRORS_EXIT;
}

void bml_hip_set_sync_mode(int mode)
{
    g_sync_mode = mode == BML_HIP_SYNC_DEFERRED ? BML_HIP_SYNC_DEFERRED
                                                : BML_HIP_SYNC_EACH_CALL;
}

int bml_hip_last_error(char *buf, int len)
{
    if (buf && len > 0) {
        strncpy(buf, g_last_code ? g_last_msg : "", (size_t) len - 1);
        buf[len - 1] = '\0';
    }
    return g_last_code;
}

void bml_hip_clear_error(void)
{
    g_last_code = BML_HIP_OK;
    g_last_msg[0] = '\0';
}

void bml_hip_synchronize(void)
{
    if (bmlh_ready())
        return;
    bmlhip_status st;
    check_status("bml_hip_synchronize", &st);
}

void *bml_hip_stream(void)
{
    if (bmlh_ready())
        return NULL;
    return bmlhip_stream();
}

const char *bml_hip_version(void)
{
    return "bml_hip 0.1 (gfx950)";
}

/* ---- row-product engine ------------------------------------------------- */

static int s_row_engine = BML_HIP_ROW_ENGINE_EXACT;

void bml_hip_set_row_engine(int engine)
{
    if (engine != BML_HIP_ROW_ENGINE_EXACT && engine != BML_HIP_ROW_ENGINE_UNORDERED &&
        engine != BML_HIP_ROW_ENGINE_WINDOW) {
        bmlh_error(BML_HIP_ERR_ARGUMENT, "bml_hip_set_row_engine: unknown engine %d", engine);
        return;
    }
    s_row_engine = engine;
}

int bml_hip_get_row_engine(void)
{
    return s_row_engine;
}

/*
 * Launch-shape choice for the ELLPACK row-product kernels.
 *
 * group (lanes per output row) follows the length of the rows of B that a
 * step streams: a step holds three GROUP-wide chunks of a row of B in
 * registers, so GROUP is the smallest power of two with 3*GROUP >= max_b and
 * a 64-lane wavefront carries 64/GROUP output rows.
 *
 * The LDS table holds a row's distinct columns.  est, the largest distinct
 * count per row, comes from the previous product of the same matrix when
 * known (scaled by the growth of its rows since), else from the bound
 * min(N, max_a*max_b).
 * The table gets max(3/2 est, est + max_b) + 64 slots in 4-slot buckets: a
 * row stops (and goes to the global-table pass) once count + (row length of
 * B) could pass the table, and at ~2/3 load the probe rounds per step (the
 * longest probe among the 3*64 lanes of a step) stay low; a smaller table
 * buys occupancy but costs more rounds (on the 32^3 lattice X^2, 1177
 * distinct columns per row, 1792 slots ran 1.4x faster than 1536 and as
 * fast as 2048).  One wavefront's tables are capped at 32 KiB so at least
 * five wavefronts share a CU.  Rows that outgrow the table are finished by the global-table
 * pass, so the cap costs speed, never correctness.
 *
 * Environment overrides: BML_HIP_GROUP, BML_HIP_TABLE, BML_HIP_BUCKET,
 * BML_HIP_LDS_BUDGET (bytes per wavefront), BML_HIP_XCD_REMAP (0/1).
 */
#include "bml_hip_tune.h"
#include "bml_hip_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void bmlh_pick_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                      int hint_max_a, int flags, bmlh_spgemm_shape *out)
{
    const int square = flags & BMLH_PICK_SQUARE;
    int g = 8;
    while (g < 64 && 3 * g < max_b)
        g <<= 1;
    long long bound = (long long) max_a * (long long) max_b;
    if (bound > N)
        bound = N;
    long long est = bound;
    if (touched_hint > 0) {
        /* the rows grew since the hint was taken: a row's distinct count
         * grows about as its length */
        double grow = hint_max_a > 0 && max_a > hint_max_a ? (double) max_a / hint_max_a : 1.0;
        est = (long long) (touched_hint * grow) + 1;
        if (est > bound)
            est = bound;
    }
    long long need = est * 3 / 2;
    if (need < est + max_b)
        need = est + max_b;
    need = (need + 64 + 63) & ~63ll;
    const long long slot = (long long) bmlh_elem_size(precision) + 4 + 2;
    const char *eb = getenv("BML_HIP_LDS_BUDGET");
    const long long budget = eb ? atoll(eb) : 32 * 1024;
    /* more rows per wavefront only while their tables fit */
    while (g < 64 && (64 / g) * (need + 8 + g) * slot > budget)
        g <<= 1;
    /* The kernel is latency bound, so its speed follows the wavefronts
     * resident per CU, which the LDS sets (160 KiB per CU, allocated in
     * 1 KiB granules here).  Take the largest table that still keeps the
     * most wavefronts with at least `need` slots: on the 32^3 lattice X^2
     * (1177 distinct columns per row, need 1856) 1792 slots (6 wavefronts
     * per CU) ran 4.8 ms, 1856 slots (5) 5.4 ms. */
    const int R = 64 / g;
    const long long pad = 8 + g;
    long long t = 0;
    for (int k = 8; k >= 5 && !t; k--) {
        const long long per = (160 * 1024 / k) / 1024 * 1024;
        const long long tk = (per / (R * slot) - pad) / 64 * 64;
        if (tk >= need - 64)
            t = tk;
    }
    if (!t) {
        t = need;
        while (t > 64 && R * (t + pad) * slot > budget)
            t -= 64;
    }
    /* A row whose table does not fit one wavefront's budget: several
     * wavefronts (at most 3 chunks of the rows of B each) share one table
     * of up to 120 KB, so the row still fits in LDS.  When the table fits,
     * one wavefront is faster: on the 32^3 lattice X^2, 6.3 ms against
     * 16 ms with three wavefronts (a step's fixed work and the barrier
     * outweigh the occupancy gain).  On the 64^3 complex X^2 (73 KB tables,
     * 5 chunks) two wavefronts ran in 203 ms, 3 in 211, 5 in 243, one
     * wavefront with the whole table in 323, and the capped table (rows in
     * the global-table pass) in 3200. */
    const int nch = (max_b + 63) / 64;
    int waves = 1;
    if (g == 64 && nch >= 2 && (need + 72) * slot > budget) {
        waves = (nch + 2) / 3;
        if (waves < 2)
            waves = 2;
        if (waves == 7)
            waves = 8;
        if (waves > 8)
            waves = 8;
        long long tm = need;
        const long long mw_budget = 120 * 1024;
        while (tm > 64 && (tm + 72) * slot > mw_budget)
            tm -= 64;
        t = tm;
    }
    const int sel =
        flags & BMLH_PICK_EXACT ? BML_HIP_ROW_ENGINE_EXACT : bml_hip_get_row_engine();
    int engine = (sel == BML_HIP_ROW_ENGINE_UNORDERED || sel == BML_HIP_ROW_ENGINE_WINDOW) &&
                         g == 64
                     ? BMLHIP_ROW_UNORDERED
                     : BMLHIP_ROW_EXACT;
    const char *e;
    if (engine == BMLHIP_ROW_UNORDERED) {
        /* W wavefronts share a row's table; no first-touch list is kept
         * (key + double-component accumulator per slot).  Insertions are
         * counted as they are claimed, so the table needs the row's
         * distinct columns plus slack against growth: max(3/2 est, est +
         * 256) + 64 slots.  For W = 4 and 8, take the largest table that
         * keeps the most blocks per CU (at most 20 wavefronts per CU), and
         * the W with the most wavefronts per CU (ties: the larger table).
         * 32^3 lattice X^2 (FP64, 1177 distinct columns per row): 4 x 2560
         * slots (5 blocks) 2.01 ms, the exact engine 4.76 ms; 64^3 complex
         * X^2 (2139 distinct): 8 wavefronts 101-107 ms, 4 wavefronts 153 ms.
         */
        long long nu = est * 3 / 2;
        if (nu < est + 256)
            nu = est + 256;
        nu = (nu + 64 + 63) & ~63ll;
        /* double-component accumulators (8 or 16 bytes) + key */
        const long long su = (precision == 3 || precision == 4 ? 16 : 8) + 4;
        long long best_waves = -1;
        for (int wv = 4; wv <= 8; wv *= 2) {
            long long tw = 0, kw = 0;
            for (int k = 32 / wv; k >= 1 && !tw; k--) {
                const long long per = (160 * 1024 / k) / 1024 * 1024 - 1024;
                const long long tk = (per / su - 72) / 64 * 64;
                if (tk >= nu - 64) {
                    tw = tk;
                    kw = k;
                }
            }
            if (!tw) {
                tw = nu;
                while (tw > 64 && (tw + 72) * su > 159 * 1024)
                    tw -= 64;
                kw = 1;
            }
            /* beyond ~20 wavefronts per CU the table's load and the
             * epilogue's idle wavefronts cost more than the occupancy buys
             * (lattice X^2: 8 x 3200 slots, 32 per CU, 2.33 ms) */
            while (kw * wv > 20) {
                kw--;
                const long long per = (160 * 1024 / kw) / 1024 * 1024 - 1024;
                tw = (per / su - 72) / 64 * 64;
            }
            if (kw * wv > best_waves || (kw * wv == best_waves && tw > t)) {
                best_waves = kw * wv;
                t = tw;
                waves = wv;
            }
        }
    }
    /* Window engine (squares): forced by the WINDOW setting, or on the
     * unordered setting when a row's distinct columns would pass what an
     * LDS hash table holds at two blocks per CU (beyond that the hash
     * engine's rows fall to the global-table pass).  Windows of 2^11
     * columns (10% faster than 2^12 on the 32^3 lattice at hopping -0.5:
     * smaller windows keep more blocks per CU), eight wavefronts per row;
     * at most 64 windows.  Blocks of 4 consecutive output rows share each
     * streamed row of X (the launcher takes fewer when 4 rows' columns
     * would not fit its LDS union table), with windows of 2^10 columns. */
    out->wbits = 0;
    out->wrows = 1;
    if (square && (sel == BML_HIP_ROW_ENGINE_WINDOW ||
                   (sel == BML_HIP_ROW_ENGINE_UNORDERED && engine == BMLHIP_ROW_UNORDERED))) {
        int wr = (e = getenv("BML_HIP_WIN_ROWS")) ? atoi(e) : 4;
        if (wr != 1 && wr != 2)
            wr = 4;
        int wb = (e = getenv("BML_HIP_WINDOW_BITS")) ? atoi(e) : (wr > 1 ? 10 : 11);
        if (wb < 10)
            wb = 10;
        if (wb > 13)
            wb = 13;
        /* smallest window with at most 64 windows over N columns */
        while (wb < 13 && ((long long) N + (1ll << wb) - 1) >> wb > BMLHIP_WIN_MAX_WINDOWS)
            wb++;
        const int fits = (((long long) N + (1ll << wb) - 1) >> wb) <= BMLHIP_WIN_MAX_WINDOWS;
        const long long wmin = (e = getenv("BML_HIP_WINDOW_MIN")) ? atoll(e) : 3000;
        if (fits && (sel == BML_HIP_ROW_ENGINE_WINDOW || est > wmin)) {
            engine = BMLHIP_ROW_WINDOW;
            out->wbits = wb;
            out->wrows = wr;
            waves = 8;
        }
    }
    out->engine = engine;
    out->waves = (e = getenv("BML_HIP_WAVES")) ? atoi(e) : waves;
    out->group = (e = getenv("BML_HIP_GROUP")) ? atoi(e) : g;
    out->table = (e = getenv("BML_HIP_TABLE")) ? atoi(e) : (int) t;
    /* 2-slot buckets for the unordered engine (its tables sit at ~1/2
     * load): 2.05 against 2.24 ms with 4-slot buckets on the lattice X^2 */
    out->bucket = (e = getenv("BML_HIP_BUCKET")) ? atoi(e)
                                                 : (engine == BMLHIP_ROW_UNORDERED ? 2 : 4);
    /* contiguous rows per XCD cut the HBM fetch of the lattice X^2 5x but
     * ran 1.3x slower than the hardware's round-robin order */
    out->remap = (e = getenv("BML_HIP_XCD_REMAP")) ? atoi(e) : 0;
}

/* ---- measured shapes ------------------------------------------------------
 *
 * With autotuning on, the first product of each kind of operand (a key of
 * precision, flags, the analytic engine and the sizes, the sizes in steps of
 * 2^(1/4)) times candidate shapes around the analytic one on the operands
 * themselves and keeps the fastest, in memory and, when a cache file is
 * named, on disk (one line per key, read back at the next start).  The
 * search is coordinate-wise: for each parameter in turn (exact engine:
 * bucket, lanes per row, table size, wavefronts; unordered: wavefronts,
 * bucket, table; window: window width, rows per block) every value is
 * timed with the others at the best so far.  Each candidate runs twice and
 * the faster time counts; a candidate that fails to launch is skipped.
 * Shapes set through the environment (BML_HIP_GROUP and the like) are left
 * alone. */
typedef struct {
    char key[128];
    bmlh_spgemm_shape sh;
    double ms;
} tune_entry;

static int g_tune = -1; /* -1: not read from the environment yet */
static char g_tune_path[1024];
static tune_entry *g_tab;
static int g_ntab, g_captab, g_loaded;

static const char *const g_override_env[] = {
    "BML_HIP_GROUP", "BML_HIP_TABLE", "BML_HIP_BUCKET", "BML_HIP_WAVES", "BML_HIP_WIN_ROWS",
    "BML_HIP_WINDOW_BITS", "BML_HIP_LDS_BUDGET", "BML_HIP_XCD_REMAP"};

static void tune_env(void)
{
    if (g_tune >= 0)
        return;
    const char *e = getenv("BML_HIP_AUTOTUNE");
    g_tune = e && atoi(e) != 0;
    const char *p = getenv("BML_HIP_TUNE_CACHE");
    if (p)
        snprintf(g_tune_path, sizeof g_tune_path, "%s", p);
}

static int tune_add(const char *key, const bmlh_spgemm_shape *sh, double ms)
{
    for (int k = 0; k < g_ntab; k++)
        if (!strcmp(g_tab[k].key, key)) {
            g_tab[k].sh = *sh;
            g_tab[k].ms = ms;
            return 1;
        }
    if (g_ntab == g_captab) {
        const int cap = g_captab ? 2 * g_captab : 64;
        tune_entry *t = realloc(g_tab, sizeof(tune_entry) * (size_t) cap);
        if (!t)
            return 0;
        g_tab = t;
        g_captab = cap;
    }
    tune_entry *t = &g_tab[g_ntab++];
    snprintf(t->key, sizeof t->key, "%s", key);
    t->sh = *sh;
    t->ms = ms;
    return 1;
}

static void tune_load(void)
{
    if (g_loaded || !g_tune_path[0])
        return;
    g_loaded = 1;
    FILE *f = fopen(g_tune_path, "r");
    if (!f)
        return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char key[128];
        bmlh_spgemm_shape sh;
        double ms;
        if (sscanf(line, "%127s %d %d %d %d %d %d %d %d %lf", key, &sh.group, &sh.table,
                   &sh.bucket, &sh.remap, &sh.waves, &sh.engine, &sh.wbits, &sh.wrows, &ms) == 10)
            tune_add(key, &sh, ms);
    }
    fclose(f);
}

static void tune_save(const char *key, const bmlh_spgemm_shape *sh, double ms)
{
    if (!g_tune_path[0])
        return;
    FILE *f = fopen(g_tune_path, "a");
    if (!f)
        return;
    fprintf(f, "%s %d %d %d %d %d %d %d %d %.6f\n", key, sh->group, sh->table, sh->bucket,
            sh->remap, sh->waves, sh->engine, sh->wbits, sh->wrows, ms);
    fclose(f);
}

void bml_hip_set_autotune(int on, const char *cache_path)
{
    tune_env();
    g_tune = on != 0;
    if (cache_path) {
        snprintf(g_tune_path, sizeof g_tune_path, "%s", cache_path);
        g_loaded = 0;
    }
}

int bml_hip_get_autotune(void)
{
    tune_env();
    return g_tune;
}

int bml_hip_autotune_entries(void)
{
    tune_env();
    tune_load();
    return g_ntab;
}

void bml_hip_autotune_clear(void)
{
    free(g_tab);
    g_tab = NULL;
    g_ntab = g_captab = 0;
    g_loaded = 1; /* the file is not read again until a path is set */
}

static int qbucket(long long x)
{
    return x <= 1 ? 0 : (int) ceil(4.0 * log2((double) x));
}

/* time one shape: the faster of two runs, < 0 if it does not launch */
static double tune_time(const bmlh_spgemm_shape *sh, bmlh_measure_fn fn, void *ctx)
{
    double best = -1.0;
    for (int r = 0; r < 2; r++) {
        double ms = 0.0;
        if (fn(sh, ctx, &ms))
            return -1.0;
        if (best < 0.0 || ms < best)
            best = ms;
    }
    return best;
}

static void try_shape(const bmlh_spgemm_shape *c, bmlh_spgemm_shape *best, double *best_ms,
                      bmlh_measure_fn fn, void *ctx)
{
    const double ms = tune_time(c, fn, ctx);
    if (ms >= 0.0 && (*best_ms < 0.0 || ms < *best_ms)) {
        *best = *c;
        *best_ms = ms;
    }
}

void bmlh_tuned_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                       int hint_max_a, int flags, bmlh_spgemm_shape *out, bmlh_measure_fn fn,
                       void *ctx)
{
    bmlh_pick_spgemm(precision, N, max_a, max_b, touched_hint, hint_max_a, flags, out);
    tune_env();
    if (!g_tune || !fn || N <= 0)
        return;
    for (size_t k = 0; k < sizeof g_override_env / sizeof g_override_env[0]; k++)
        if (getenv(g_override_env[k]))
            return;
    tune_load();
    char key[128];
    snprintf(key, sizeof key, "cu%d:p%d:f%d:e%d:n%d:a%d:b%d:t%d", bmlhip_device_cus(), precision,
             flags, out->engine, qbucket(N), qbucket(max_a), qbucket(max_b),
             qbucket(touched_hint));
    for (int k = 0; k < g_ntab; k++)
        if (!strcmp(g_tab[k].key, key)) {
            *out = g_tab[k].sh;
            return;
        }
    bmlh_spgemm_shape best = *out;
    double best_ms = tune_time(out, fn, ctx);
    if (best_ms < 0.0)
        return; /* the analytic shape itself fails: leave it to report */
    bmlh_spgemm_shape c;
    if (out->engine == BMLHIP_ROW_WINDOW) {
        static const int wb[] = {10, 11};
        static const int wr[] = {1, 2, 4};
        for (int k = 0; k < 2; k++) {
            c = best;
            c.wbits = wb[k];
            if (c.wbits != best.wbits)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
        for (int k = 0; k < 3; k++) {
            c = best;
            c.wrows = wr[k];
            if (c.wrows != best.wrows)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
    } else if (out->engine == BMLHIP_ROW_UNORDERED) {
        static const int wv[] = {4, 8};
        for (int k = 0; k < 2; k++) {
            c = best;
            c.waves = wv[k];
            if (c.waves != best.waves)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
        c = best;
        c.bucket = best.bucket == 2 ? 4 : 2;
        try_shape(&c, &best, &best_ms, fn, ctx);
        const int t0 = best.table;
        for (int k = 0; k < 2; k++) {
            c = best;
            c.table = (k ? t0 * 3 / 2 : t0 * 3 / 4) / 64 * 64;
            if (c.table >= 256)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
    } else {
        static const int bk[] = {2, 4, 8};
        for (int k = 0; k < 3; k++) {
            c = best;
            c.bucket = bk[k];
            if (c.bucket != best.bucket)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
        const int g0 = best.group;
        for (int g = g0 / 2; g <= 2 * g0 && g <= 64; g *= 2) {
            if (g < 8 || g == g0)
                continue;
            c = best;
            c.group = g;
            c.waves = 1;
            try_shape(&c, &best, &best_ms, fn, ctx);
        }
        const int t0 = best.table;
        for (int k = 0; k < 2; k++) {
            c = best;
            c.table = (k ? t0 * 3 / 2 : t0 * 3 / 4) / 64 * 64;
            if (c.table >= 64)
                try_shape(&c, &best, &best_ms, fn, ctx);
        }
        if (best.group == 64) {
            static const int wv[] = {1, 2, 3, 4};
            for (int k = 0; k < 4; k++) {
                c = best;
                c.waves = wv[k];
                if (c.waves != best.waves)
                    try_shape(&c, &best, &best_ms, fn, ctx);
            }
        }
    }
    *out = best;
    if (tune_add(key, &best, best_ms))
        tune_save(key, &best, best_ms);
}

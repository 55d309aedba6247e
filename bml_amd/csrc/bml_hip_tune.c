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

#include <stdlib.h>

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

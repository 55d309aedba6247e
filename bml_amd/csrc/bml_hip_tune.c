/*
 * Launch-shape choice for the ELLPACK row-product kernels.
 *
 * group (lanes per output row) follows the length of the rows of B that a
 * step streams: one GROUP-wide chunk per step when rows are short, so a
 * 64-lane wavefront carries 64/GROUP output rows.  The LDS table holds the
 * row's distinct columns at <= 3/4 load; its size comes from the distinct
 * count observed on the previous product of the same matrix when known,
 * else from the bound min(N, max_a*max_b), and is capped so one wavefront's
 * tables stay within 40 KiB (four wavefronts per CU at least).  Rows that
 * outgrow the table are finished by the global-table pass, so the cap costs
 * speed, never correctness.  Environment overrides: BML_HIP_GROUP,
 * BML_HIP_TABLE.
 */
#include "bml_hip_tune.h"
#include "bml_hip_host.h"

#include <stdlib.h>

static long long next_pow2(long long v)
{
    long long p = 64;
    while (p < v)
        p <<= 1;
    return p;
}

void bmlh_pick_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                      int *group, int *table)
{
    int g = max_b <= 8 ? 8 : max_b <= 16 ? 16 : max_b <= 32 ? 32 : 64;
    long long bound = (long long) max_a * (long long) max_b;
    if (bound > N)
        bound = N;
    long long est = touched_hint > 0 ? touched_hint : bound;
    if (est > bound && bound > 0)
        est = bound;
    long long t = next_pow2(est + est / 3 + g + 16);
    const long long slot = (long long) bmlh_elem_size(precision) + 4 + 2;
    const long long budget = 40 * 1024;
    /* more rows per wavefront only while their tables fit */
    while (g < 64 && (64 / g) * t * slot > budget)
        g <<= 1;
    while (t > 64 && (64 / g) * t * slot > budget)
        t >>= 1;
    const char *eg = getenv("BML_HIP_GROUP");
    const char *et = getenv("BML_HIP_TABLE");
    if (eg)
        g = atoi(eg);
    if (et)
        t = atoll(et);
    *group = g;
    *table = (int) t;
}

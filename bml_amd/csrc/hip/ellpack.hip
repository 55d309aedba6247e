// gfx950 kernels for the ELLPACK-R thresholded multiply path.
//
// Row accumulator.  Every output row is built by a "group" of GROUP lanes
// (8..64, so 64/GROUP rows per 64-lane wavefront) that owns an open-addressed
// hash table in LDS: keys[] (column, -1 = empty), vals[] (the running sum)
// and ord[] (slot of the r-th distinct column, in first-touch order).  The
// group walks row i of A in jp order; for each (a, j) it streams row j of B
// in GROUP-wide chunks (one coalesced load per chunk), probes/inserts the
// columns (linear probing, CAS to claim a slot) and adds a*b into vals[].
// All columns of one row of B are distinct, so within a chunk no two lanes
// touch the same slot, and successive chunks/steps are ordered by the
// wavefront's program order: every output value is summed in exactly the
// reference's order (jp ascending), with no atomics on values.  New columns
// get their rank from a ballot prefix count in lane (= kp) order, so ord[]
// reproduces the reference's first-touch order and the epilogue emits each
// output row in the reference's column order.
//
// Epilogue: the group walks ord[], keeps the diagonal and |v| > threshold,
// compacts with a ballot prefix count into the output row, records per-row
// trace contributions and resets the touched slots.
//
// Rows whose distinct-column count would pass 3/4 of the LDS table stop and
// are listed for a second pass that runs the same code on a large table in
// global memory (one wavefront per row), so no input is ever refused.
#include "../bmlhip_abi.h"
#include "devtypes.h"

#include <limits.h>
#include <type_traits>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

using namespace bmlhip;

namespace {

constexpr int EMPTY = -1;

// bml_matrix_precision_t of an element type
template <typename T> struct PrecOf;
template <> struct PrecOf<float> { static constexpr int value = 1; };
template <> struct PrecOf<double> { static constexpr int value = 2; };
template <> struct PrecOf<cf32> { static constexpr int value = 3; };
template <> struct PrecOf<cf64> { static constexpr int value = 4; };
// global-table pass: at most this many blocks, and at most this many bytes
// of tables in total
constexpr int FALLBACK_BLOCKS = 256;
constexpr size_t FALLBACK_BYTES = (size_t) 1 << 30;

template <typename T>
__device__ __forceinline__ T shfl_any(T v, int src) {
    static_assert(sizeof(T) % 4 == 0, "element must be a multiple of 4 bytes");
    union U {
        T t;
        int w[sizeof(T) / 4];
    } u, r;
    u.t = v;
#pragma unroll
    for (int k = 0; k < (int) (sizeof(T) / 4); k++) r.w[k] = __shfl(u.w[k], src, 64);
    return r.t;
}

__device__ __forceinline__ unsigned hash_slot(int col, int tbits) {
    return ((unsigned) col * 0x9E3779B1u) >> (32 - tbits);
}


struct SpParams {
    int N, MA, MB, MC, kind;
    const void *Av;
    const int *Ai, *An;
    const void *Bv;
    const int *Bi, *Bn;
    void *Cv;
    int *Ci, *Cn;
    double thr;
    double *rowtr;  // [2*N]: per-row tr X and tr X^2 contributions (x2)
    bmlhip_status *st;
    int *ovf;         // rows handed to the global-table pass
    const int *skip;  // when set and nonzero, the launch does nothing
    int4 *rowstat;    // [N]: (kept, distinct columns, products, nnz of A's row),
                      // kept = -1 for a skipped row; reduced by k_row_stats
    int remap;        // XCD-aware block-to-row order
    int sym;          // symmetric x2: rows form only the columns k >= i
    // batched systems: row i belongs to system row_sys[i]; rows of systems
    // with sys_done set are skipped (both NULL: one system)
    const int *row_sys, *sys_done;
    // fuse: C <- beta*C + alpha*(A*B), merged in the epilogue
    int fuse;
    double alpha, beta;
    // batched work queue (NULL: one block per row): the rows to form are
    // qlist[0 .. qinfo[0]), taken with the cursor qinfo[1]
    const int *qlist;
    int *qinfo;
};

// Broadcast lane src's value to the caller's lanes.  With one group per
// wavefront src is wave-uniform and a scalar readlane does it.
template <int GROUP, typename T>
__device__ __forceinline__ T bcast(T v, int src) {
    if constexpr (GROUP == 64) {
        static_assert(sizeof(T) % 4 == 0, "element must be a multiple of 4 bytes");
        union U {
            T t;
            int w[sizeof(T) / 4];
        } u, r;
        u.t = v;
#pragma unroll
        for (int k = 0; k < (int) (sizeof(T) / 4); k++)
            r.w[k] = __builtin_amdgcn_readlane(u.w[k], src);
        return r.t;
    } else {
        return shfl_any(v, src);
    }
}

// A step holds up to NCH GROUP-wide chunks of a row of B in registers.
constexpr int NCH = 3;

template <typename T>
struct BRow {
    int col[NCH];
    T val[NCH];
};

// The table is an array of buckets of BK slots (BK = 2, 4 or 8 keys, read
// with one or two 8/16-byte LDS accesses).  A column's home bucket comes
// from a multiplicative hash scaled to the bucket count (any count: tables
// need not be powers of two).  Tables in LDS (fewer than 2^16 buckets) use
// 24-bit multiplies, full rate on CDNA; tables in global memory may be larger
// and use a 64-bit product.
template <bool BIG>
__device__ __forceinline__ unsigned home_bucket(int col, unsigned nbk) {
    if constexpr (BIG) {
        const unsigned x = (unsigned) col * 0x9E3779B1u;
        return (unsigned) (((unsigned long long) x * nbk) >> 32);
    } else {
        const unsigned x = (unsigned) __umul24(col, 0x9E3779u) >> 16;
        return (unsigned) __umul24(x, nbk) >> 16;
    }
}

template <int BK>
struct Bucket {
    int k[BK];
};

template <int BK>
__device__ __forceinline__ Bucket<BK> load_bucket(const int *keys, unsigned at) {
    Bucket<BK> b;
    if constexpr (BK == 2) {
        const int2 v = *reinterpret_cast<const int2 *>(keys + at);
        b.k[0] = v.x;
        b.k[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < BK / 4; q++) {
            const int4 v = *reinterpret_cast<const int4 *>(keys + at + 4 * q);
            b.k[4 * q] = v.x;
            b.k[4 * q + 1] = v.y;
            b.k[4 * q + 2] = v.z;
            b.k[4 * q + 3] = v.w;
        }
    }
    return b;
}

// Where col sits in a bucket (BK if absent), and how many of its slots are
// filled.  Buckets fill front to back and nothing is deleted while a row is
// built, so a present column is among the filled slots and the first empty
// slot is at index `filled` (BK: the bucket is full).  EMPTY (-1) is the
// only key with the sign bit set.
template <int BK>
__device__ __forceinline__ void bucket_scan(const Bucket<BK> &b, int col, unsigned &hit_at,
                                            unsigned &filled) {
    unsigned e_hit = BK, empties = 0;
#pragma unroll
    for (int e = BK - 1; e >= 0; e--) {
        e_hit = b.k[e] == col ? (unsigned) e : e_hit;
        empties += (unsigned) b.k[e] >> 31;
    }
    hit_at = e_hit;
    filled = BK - empties;
}

// Slot of col in a finished table, or -1 when absent (the probe sequence of
// the build: an empty slot in a bucket proves the column absent).
template <int BK, bool BIG>
__device__ __forceinline__ int find_slot(const int *keys, int col, unsigned nbk, unsigned tsize) {
    unsigned h = home_bucket<BIG>(col, nbk) * BK;
    while (true) {
        const Bucket<BK> kb = load_bucket<BK>(keys, h);
        unsigned eh, filled;
        bucket_scan<BK>(kb, col, eh, filled);
        if (eh < (unsigned) BK) return (int) (h + eh);
        if (filled < (unsigned) BK) return -1;
        h += BK;
        if (h >= tsize) h -= tsize;
    }
}

// A product slot whose column C's row already took (the key keeps its sign
// bit clear, so probes still see the slot filled; columns are below 2^24).
constexpr int TAKEN = 1 << 30;

// Epilogue of C <- beta*C + alpha*P for one row, P the row's product in the
// table (count columns, first-touch order in ord).  It is the reference's
// bml_multiply: the product thresholded as bml_multiply_x2 / _AB store it
// (bml_multiply_ellpack_typed.c:86-101; the diagonal always kept, :306,527), then
// bml_add(C, P, beta, alpha) (bml_add_ellpack_typed.c:162-196): C's columns
// in C's order, each 0 + beta*c (when beta != 0) plus alpha*p where P keeps
// that column, then P's other kept columns in first-touch order as
// 0 + alpha*p (when alpha != 0), every sum kept iff |v| > threshold.  C's
// row is read and compacted in place (an entry is only ever written at or
// before its own position).  Returns the kept count, or MC + 1 when the
// product alone overflows C (x2: kept entries, AB: distinct columns); d2 is
// the product's diagonal.  Cleans the table.
template <typename T, int GROUP, int BK, bool BIG, typename ORD>
__device__ int fused_epilogue(const SpParams &p, int i, bool rowact, int *keys, T *vals,
                              const ORD *ord, unsigned tsize, int count, uint64_t gmask,
                              T &d2) {
    using E = Elem<T>;
    const int lane = lane_id(), gl = lane % GROUP;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const unsigned nbk = tsize / BK;
    const int MC = p.MC;
    const size_t rowC = (size_t) (rowact ? i : 0) * MC;
    T *Cv = (T *) p.Cv;
    const bool usec = p.beta > 0.0 || p.beta < 0.0;
    const bool usep = p.alpha > 0.0 || p.alpha < 0.0;
    auto pkeeps = [&](int col, T pv) { return col == i || E::above(pv, p.thr); };
    const int nC = rowact && usec ? p.Cn[i] : 0;
    int kept = 0;
    for (int r0 = 0; __any(r0 < nC); r0 += GROUP) {
        const int r = r0 + gl;
        const bool act = r < nC;
        int col = 0;
        T v = E::zero();
        if (act) {
            col = p.Ci[rowC + r];
            v = E::axpy(E::zero(), p.beta, Cv[rowC + r]);
            const int s = find_slot<BK, BIG>(keys, col, nbk, tsize);
            if (s >= 0) {
                const T pv = vals[s];
                if (usep && pkeeps(col, pv)) {
                    v = E::axpy(v, p.alpha, pv);
                    keys[s] = col | TAKEN;
                }
            }
        }
        const bool keep = act && E::above(v, p.thr);
        const uint64_t km = ballot64(keep) & gmask;
        const int pos = kept + popc64(km & ltmask);
        if (keep && pos < MC) {
            p.Ci[rowC + pos] = col;
            Cv[rowC + pos] = v;
        }
        kept += popc64(km);
    }
    int keptp = 0;
    for (int r0 = 0; __any(rowact && r0 < count); r0 += GROUP) {
        const int r = r0 + gl;
        const bool act = rowact && r < count;
        int s = 0, key = 0;
        T pv = E::zero();
        if (act) {
            s = (int) ord[r];
            key = keys[s];
            pv = vals[s];
        }
        const int col = key & ~TAKEN;
        const bool pk = act && pkeeps(col, pv);
        keptp += popc64(ballot64(pk) & gmask);
        const uint64_t dm = ballot64(act && col == i) & gmask;
        if (dm) d2 = shfl_any(pv, __ffsll((unsigned long long) dm) - 1);
        const T v = E::axpy(E::zero(), p.alpha, pv);
        const bool keep = usep && pk && !(key & TAKEN) && E::above(v, p.thr);
        const uint64_t km = ballot64(keep) & gmask;
        const int pos = kept + popc64(km & ltmask);
        if (keep && pos < MC) {
            p.Ci[rowC + pos] = col;
            Cv[rowC + pos] = v;
        }
        if (act) {
            keys[s] = EMPTY;
            vals[s] = E::zero();
        }
        kept += popc64(km);
    }
    const bool povf = p.kind == 0 ? keptp > MC : count > MC;
    return povf && kept <= MC ? MC + 1 : kept;
}

// One row per group.  keys/vals/ord point at the group's table (LDS or
// global): tsize slots (nbk = tsize/BK buckets), then a dummy bucket at slot
// tsize that finished lanes read, then (from slot tsize + 8) one private
// slot per lane that idle lanes aim their CAS and value accesses at, so that
// every LDS access of the probe loop is unconditional (no exec-mask branches,
// which would force a wait after each access) and conflict-free for idle
// lanes.
//
// Per step (one entry (a, j) of row i) the chunks of row j of B are in
// registers (loaded one step ahead, branch-free, into two buffers used
// alternately).  Their columns are distinct, so all NCH chunks are probed
// together: one bucket read per chunk per round, NCH reads in flight.
// Insertion takes the first empty slot of a bucket and nothing is deleted
// while a row is built, so the filled slots of every bucket form a prefix:
// the first slot holding the column or EMPTY ends the search (a match, or
// proof that the column is absent, and the lane claims that slot with CAS --
// if another lane took it first the lane reads the bucket again); a bucket
// with neither sends the lane to the next bucket (linear probing over
// buckets).  New columns are ranked in (chunk, lane) = kp order, then the
// products are added into their slots (plain read-modify-write: a step's
// slots are distinct and steps follow program order).
template <typename T, int GROUP, int BK, bool BIG, typename ORD>
__device__ void spgemm_row(const SpParams &p, int i, int *keys, T *vals, ORD *ord, unsigned tsize,
                           int cap) {
    using E = Elem<T>;
    const int lane = lane_id();
    const int g = lane / GROUP, gl = lane % GROUP;
    const uint64_t gmask = GROUP == 64 ? ~0ull : (((1ull << GROUP) - 1ull) << (g * GROUP));
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const T *Av = (const T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    const int *Ai = p.Ai, *Bi = p.Bi, *Bn = p.Bn;
    const bool live = i >= 0;
    const int nA = live ? p.An[i] : 0;
    const unsigned rowA = live ? (unsigned) i * (unsigned) p.MA : 0u;
    const unsigned MB = (unsigned) p.MB;
    const unsigned nbk = tsize / BK, dummy = tsize, priv = tsize + 8u + (unsigned) gl;

    int count = 0;        // distinct columns so far (group-uniform)
    bool tovf = false;    // table would overflow (group-uniform)
    long long prods = 0;  // partial products of the row (group-uniform)
    T trx = E::zero();

    // branch-free loads of row j's first NCH chunks (idle lanes read element 0)
    auto load_b = [&](BRow<T> &r, int j, int nb, bool act) {
        const unsigned base = (unsigned) j * MB;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int kp = c * GROUP + gl;
            const unsigned at = (act & (kp < nb)) ? base + (unsigned) kp : 0u;
            r.col[c] = Bi[at];
            r.val[c] = Bv[at];
        }
    };

    // the NCH chunks of a row of B in r (entries kbase.. of the row) times a
    auto accumulate = [&](const BRow<T> &r, int kbase, int nb, bool gact, T a) {
        // h: bucket being probed, dummy once the lane is done (or idle)
        unsigned h[NCH], slot[NCH];
        bool isnew[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            bool act = gact & (kbase + c * GROUP + gl < nb);
            // the global-table pass of a symmetric x2 forms columns k >= i only
            if constexpr (BIG) act = act & (!p.sym | (r.col[c] >= i));
            h[c] = act ? home_bucket<BIG>(r.col[c], nbk) * BK : dummy;
            slot[c] = priv;
            isnew[c] = false;
        }
        while (true) {
            Bucket<BK> kb[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) kb[c] = load_bucket<BK>(keys, h[c]);
            unsigned at[NCH];
            bool ins[NCH], anyins = false;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const bool pend = h[c] != dummy;
                unsigned eh, nf;
                bucket_scan<BK>(kb[c], r.col[c], eh, nf);
                const bool found = eh < (unsigned) BK;
                at[c] = h[c] + (found ? eh : nf);
                const bool hit = pend & found;
                ins[c] = pend & !found & (nf < (unsigned) BK);
                slot[c] = hit ? at[c] : slot[c];
                unsigned hn = h[c] + BK;
                hn = hn >= tsize ? hn - tsize : hn;
                h[c] = (hit | !pend) ? dummy : (ins[c] ? h[c] : hn);
                anyins = anyins | ins[c];
            }
            if (__any(anyins)) {
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int old = atomicCAS(&keys[ins[c] ? at[c] : priv], EMPTY, r.col[c]);
                    const bool won = ins[c] & (old == EMPTY);
                    isnew[c] = isnew[c] | won;
                    slot[c] = won ? at[c] : slot[c];
                    h[c] = won ? dummy : h[c];
                }
            }
            bool anyp = false;
#pragma unroll
            for (int c = 0; c < NCH; c++) anyp = anyp | (h[c] != dummy);
            if (!__any(anyp)) break;
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const uint64_t nm = ballot64(isnew[c]) & gmask;
            if (isnew[c]) ord[count + popc64(nm & ltmask)] = (ORD) slot[c];
            count += popc64(nm);
        }
        // idle lanes add into their private slot
        T v[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) v[c] = vals[slot[c]];
#pragma unroll
        for (int c = 0; c < NCH; c++) vals[slot[c]] = E::add(v[c], E::mul(a, r.val[c]));
    };

    BRow<T> buf0, buf1;
    for (int c0 = 0; __any(c0 < nA); c0 += GROUP) {
        // GROUP entries of row i, and nnz of the rows they name
        const int jl = c0 + gl;
        const unsigned at = jl < nA ? rowA + (unsigned) jl : 0u;
        const T av = Av[at];
        const int aj = jl < nA ? Ai[at] : 0;
        const int an = Bn[aj];
        const int steps = nA - c0 < GROUP ? (nA - c0 > 0 ? nA - c0 : 0) : GROUP;

        T a = bcast<GROUP>(av, g * GROUP);
        int j = bcast<GROUP>(aj, g * GROUP);
        int nb = bcast<GROUP>(an, g * GROUP);
        load_b(buf0, j, nb, 0 < steps && !tovf);

        // step t: accumulate row j (in cur) times a, prefetch step t+1 into nxt
        auto step = [&](int t, BRow<T> &cur, BRow<T> &nxt) {
            const int t1 = t + 1 < GROUP ? t + 1 : t;
            const T a1 = bcast<GROUP>(av, g * GROUP + t1);
            const int j1 = bcast<GROUP>(aj, g * GROUP + t1);
            const int nb1 = bcast<GROUP>(an, g * GROUP + t1);
            load_b(nxt, j1, nb1, t + 1 < steps && !tovf);
            // a step adds at most nb columns and the table must keep an empty
            // slot: a row that could fill it stops and goes to the
            // global-table pass
            bool gact = t < steps && !tovf;
            if (gact && count + nb > cap) {
                tovf = true;
                gact = false;
            }
            if (gact) {
                prods += nb;
                if (j == i) trx = E::add(trx, a);
            }
            accumulate(cur, 0, nb, gact, a);
            // rows of B longer than NCH*GROUP entries: the rest NCH chunks at a time
            for (int k0 = NCH * GROUP; __any(gact && k0 < nb); k0 += NCH * GROUP) {
                BRow<T> extra;
                const unsigned base = (unsigned) j * MB;
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int kp = k0 + c * GROUP + gl;
                    const unsigned at2 = gact && kp < nb ? base + (unsigned) kp : 0u;
                    extra.col[c] = Bi[at2];
                    extra.val[c] = Bv[at2];
                }
                accumulate(extra, k0, nb, gact, a);
            }
            a = a1;
            j = j1;
            nb = nb1;
        };
        for (int t = 0; __any(t < steps && !tovf); t += 2) {
            step(t, buf0, buf1);
            step(t + 1, buf1, buf0);
        }
    }

    // epilogue: thresholded compaction in first-touch order
    const bool rowact = live && !tovf;
    const int MC = p.MC;
    const size_t rowC = (size_t) (live ? i : 0) * MC;
    T *Cv = (T *) p.Cv;
    int kept = 0;
    T d2 = E::zero();
    if (p.fuse)
        kept = fused_epilogue<T, GROUP, BK, BIG, ORD>(p, i, rowact, keys, vals, ord, tsize, count,
                                                      gmask, d2);
    else for (int r0 = 0; __any(rowact && r0 < count); r0 += GROUP) {
        const int r = r0 + gl;
        const bool act = rowact && r < count;
        int s = 0, col = 0;
        T v = E::zero();
        if (act) {
            s = (int) ord[r];
            col = keys[s];
            v = vals[s];
        }
        const bool isdiag = act && col == i;
        const bool keep = isdiag || (act && E::above(v, p.thr));
        const uint64_t km = ballot64(keep) & gmask;
        const int pos = kept + popc64(km & ltmask);
        if (keep && pos < MC) {
            p.Ci[rowC + pos] = col;
            Cv[rowC + pos] = v;
        }
        const uint64_t dm = ballot64(isdiag) & gmask;
        if (dm) d2 = shfl_any(v, __ffsll((unsigned long long) dm) - 1);
        if (act) {
            keys[s] = EMPTY;
            vals[s] = E::zero();
        }
        kept += popc64(km);
    }
    if (live && tovf) {
        // leave the table clean, then hand the row to the global-table pass
        for (int r = gl; r < count; r += GROUP) {
            const int s = (int) ord[r];
            keys[s] = EMPTY;
            vals[s] = E::zero();
        }
        if (gl == 0) {
            const int q = atomicAdd(&p.st->table_overflow_rows, 1);
            p.ovf[q] = i;
            atomicAdd((unsigned long long *) &p.st->prof_ovf_rows[PrecOf<T>::value], 1ull);
        }
    }
    if (rowact && gl == 0) {
        p.Cn[i] = kept < MC ? kept : MC;
        // plain stores: same-address atomics from every row serialise (eight
        // per row cost 1.3 ms over 32768 rows)
        p.rowstat[i] = make_int4(kept, count, (int) prods, nA);
        if (p.kind == 0) {
            p.rowtr[i] = E::re(trx);
            p.rowtr[p.N + i] = E::re(d2);
        }
    }
}

// slots per table beyond tsize: the dummy bucket (8 slots) and one private
// slot per lane of the group; a multiple of 8 keeps every table 16-byte
// aligned
template <int GROUP>
constexpr unsigned table_pad() { return 8u + (unsigned) GROUP; }

// Several wavefronts per row (long rows of B).  The W wavefronts of a block
// share one row's table; at each step wavefront w takes its own range of the
// chunks of row j of B (q = ceil(chunks / W) consecutive chunks, so the
// wavefronts' ranges follow kp order).  The columns of one row of B are
// distinct, so within a step the wavefronts touch distinct slots; one
// barrier per step orders the steps: every wavefront's probes and inserts
// of step t are done before any wavefront ranks or adds (the new-column
// counts of the step are exchanged through LDS, double-buffered), and the
// adds of step t come before the adds of step t+1.  Sums are therefore
// formed in the reference's order, as in the one-wavefront kernel.
template <typename T, int W, int BK>
__global__ __launch_bounds__(64 * W) void k_spgemm_mw(SpParams p, unsigned tsize) {
    using E = Elem<T>;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int cnt_s[2][W];
    if (p.skip && *p.skip) return;
    const unsigned stride = tsize + table_pad<64>();
    T *vals = (T *) smem;
    int *keys = (int *) (smem + (size_t) stride * sizeof(T));
    uint16_t *ord = (uint16_t *) (smem + (size_t) stride * (sizeof(T) + sizeof(int)));
    for (unsigned s = threadIdx.x; s < stride; s += 64 * W) {
        keys[s] = EMPTY;
        vals[s] = E::zero();
    }
    __syncthreads();
    const int row = blockIdx.x;
    if (row >= p.N) return;
    if (p.row_sys && p.sys_done[p.row_sys[row]]) {  // uniform
        if (threadIdx.x == 0) p.rowstat[row] = make_int4(-1, 0, 0, 0);
        return;
    }
    const int i = row;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const T *Av = (const T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    const int *Ai = p.Ai, *Bi = p.Bi, *Bn = p.Bn;
    const int nA = p.An[i];
    const unsigned rowA = (unsigned) i * (unsigned) p.MA;
    const unsigned MB = (unsigned) p.MB;
    const unsigned nbk = tsize / BK, dummy = tsize, priv = tsize + 8u + (unsigned) lane;
    const int cap = (int) tsize - 8;

    int count = 0;
    bool tovf = false;
    long long prods = 0;
    T trx = E::zero();

    // first kp of this wavefront's chunks for a row of nb entries
    auto first_kp = [&](int nb) {
        const int nch = (nb + 63) / 64;
        const int q = (nch + W - 1) / W;
        return w * q * 64;
    };
    auto last_kp = [&](int nb) {
        const int nch = (nb + 63) / 64;
        const int q = (nch + W - 1) / W;
        const int e = (w + 1) * q * 64;
        return e < nb ? e : nb;
    };
    auto load_b = [&](BRow<T> &r, int j, int nb, bool act) {
        const unsigned base = (unsigned) j * MB;
        const int k0 = first_kp(nb), k1 = last_kp(nb);
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int kp = k0 + c * 64 + lane;
            const unsigned at = (act & (kp < k1)) ? base + (unsigned) kp : 0u;
            r.col[c] = Bi[at];
            r.val[c] = Bv[at];
        }
    };

    BRow<T> buf0, buf1;
    for (int c0 = 0; c0 < nA; c0 += 64) {
        const int jl = c0 + lane;
        const unsigned at = jl < nA ? rowA + (unsigned) jl : 0u;
        const T av = Av[at];
        const int aj = jl < nA ? Ai[at] : 0;
        const int an = Bn[aj];
        const int steps = nA - c0 < 64 ? nA - c0 : 64;
        T a = bcast<64>(av, 0);
        int j = bcast<64>(aj, 0);
        int nb = bcast<64>(an, 0);
        load_b(buf0, j, nb, !tovf);
        // step t: row j (in cur) times a, prefetch step t+1 into nxt; the two
        // buffers are named, not indexed, so they stay in registers
        auto step = [&](int t, BRow<T> &cur, BRow<T> &nxt) {
            const int t1 = t + 1 < 64 ? t + 1 : t;
            const T a1 = bcast<64>(av, t1);
            const int j1 = bcast<64>(aj, t1);
            const int nb1 = bcast<64>(an, t1);
            bool gact = !tovf;
            // the table must keep an empty slot, and a wavefront holds at
            // most NCH chunks of a row: otherwise the row goes to the
            // global-table pass
            if (gact && (count + nb > cap || ((nb + 63) / 64 + W - 1) / W > NCH)) {
                tovf = true;
                gact = false;
            }
            load_b(nxt, j1, nb1, t + 1 < steps && gact);
            if (gact) {
                prods += nb;
                if (j == i) trx = E::add(trx, a);
            }
            const int k0 = first_kp(nb), k1 = last_kp(nb);
            // probe / insert this wavefront's chunks
            unsigned h[NCH], slot[NCH];
            bool act[NCH], isnew[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                act[c] = gact & (k0 + c * 64 + lane < k1);
                h[c] = act[c] ? home_bucket<false>(cur.col[c], nbk) * BK : dummy;
                slot[c] = priv;
                isnew[c] = false;
            }
            while (true) {
                Bucket<BK> kb[NCH];
#pragma unroll
                for (int c = 0; c < NCH; c++) kb[c] = load_bucket<BK>(keys, h[c]);
                unsigned atc[NCH];
                bool ins[NCH], anyins = false;
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const bool pend = h[c] != dummy;
                    unsigned eh, nf;
                    bucket_scan<BK>(kb[c], cur.col[c], eh, nf);
                    const bool found = eh < (unsigned) BK;
                    atc[c] = h[c] + (found ? eh : nf);
                    const bool hit = pend & found;
                    ins[c] = pend & !found & (nf < (unsigned) BK);
                    slot[c] = hit ? atc[c] : slot[c];
                    unsigned hn = h[c] + BK;
                    hn = hn >= tsize ? hn - tsize : hn;
                    h[c] = (hit | !pend) ? dummy : (ins[c] ? h[c] : hn);
                    anyins = anyins | ins[c];
                }
                if (__any(anyins)) {
#pragma unroll
                    for (int c = 0; c < NCH; c++) {
                        const int old = atomicCAS(&keys[ins[c] ? atc[c] : priv], EMPTY, cur.col[c]);
                        const bool won = ins[c] & (old == EMPTY);
                        isnew[c] = isnew[c] | won;
                        slot[c] = won ? atc[c] : slot[c];
                        h[c] = won ? dummy : h[c];
                    }
                }
                bool anyp = false;
#pragma unroll
                for (int c = 0; c < NCH; c++) anyp = anyp | (h[c] != dummy);
                if (!__any(anyp)) break;
            }
            int nw = 0;
#pragma unroll
            for (int c = 0; c < NCH; c++) nw += popc64(ballot64(isnew[c]));
            if (lane == 0) cnt_s[t & 1][w] = nw;
            __syncthreads();
            int base = count, tot = 0;
#pragma unroll
            for (int v = 0; v < W; v++) {
                const int cv = cnt_s[t & 1][v];
                base += v < w ? cv : 0;
                tot += cv;
            }
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const uint64_t nm = ballot64(isnew[c]);
                if (isnew[c]) ord[base + popc64(nm & ltmask)] = (uint16_t) slot[c];
                base += popc64(nm);
            }
            count += tot;
            T v[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) v[c] = vals[slot[c]];
#pragma unroll
            for (int c = 0; c < NCH; c++) vals[slot[c]] = E::add(v[c], E::mul(a, cur.val[c]));
            a = a1;
            j = j1;
            nb = nb1;
        };
        for (int t = 0; t < steps; t += 2) {
            step(t, buf0, buf1);
            if (t + 1 < steps) step(t + 1, buf1, buf0);
        }
    }
    __syncthreads();  // every wavefront's last adds are in
    if (w != 0) return;
    // epilogue (wavefront 0): thresholded compaction in first-touch order
    const int MC = p.MC;
    const size_t rowC = (size_t) i * MC;
    T *Cv = (T *) p.Cv;
    int kept = 0;
    T d2 = E::zero();
    if (!tovf && p.fuse) {
        kept = fused_epilogue<T, 64, BK, false, uint16_t>(p, i, true, keys, vals, ord, tsize, count,
                                                          ~0ull, d2);
    } else if (!tovf) {
        for (int r0 = 0; r0 < count; r0 += 64) {
            const int r = r0 + lane;
            const bool act = r < count;
            int s = 0, col = 0;
            T v = E::zero();
            if (act) {
                s = (int) ord[r];
                col = keys[s];
                v = vals[s];
            }
            const bool isdiag = act && col == i;
            const bool keep = isdiag || (act && E::above(v, p.thr));
            const uint64_t km = ballot64(keep);
            const int pos = kept + popc64(km & ltmask);
            if (keep && pos < MC) {
                p.Ci[rowC + pos] = col;
                Cv[rowC + pos] = v;
            }
            const uint64_t dm = ballot64(isdiag);
            if (dm) d2 = shfl_any(v, __ffsll((unsigned long long) dm) - 1);
            kept += popc64(km);
        }
    } else if (lane == 0) {
        const int q = atomicAdd(&p.st->table_overflow_rows, 1);
        p.ovf[q] = i;
        atomicAdd((unsigned long long *) &p.st->prof_ovf_rows[PrecOf<T>::value], 1ull);
    }
    if (!tovf && lane == 0) {
        p.Cn[i] = kept < MC ? kept : MC;
        p.rowstat[i] = make_int4(kept, count, (int) prods, nA);
        if (p.kind == 0) {
            p.rowtr[i] = E::re(trx);
            p.rowtr[p.N + i] = E::re(d2);
        }
    }
}

// ---------------------------------------------------------------------------
// Unordered engine.  W wavefronts share one row's LDS table and take the
// entries of row i of A round-robin (wavefront w: jp = w, w + W, ...) with no
// barrier between steps; products go into the table with LDS atomic adds.
// A column's sum is therefore formed in arrival order rather than the
// reference's jp order (it agrees with the exact engine to rounding, a few
// ulps of the sum's magnitudes, and may differ between runs in the last
// bits), and the output row lists its kept columns in table order.  Every
// insertion attempt is counted on a block counter before its CAS (a wave's
// attempts of a round with one atomic, which also gives back its earlier
// attempts that filled no slot); the count bounds the filled slots and is
// kept within tsize - 8, so a probe always finds an empty slot.  A row that
// would pass that goes to the global-table pass, which runs the exact
// engine.
//
// Inserts follow the exact engine's bucket protocol (first empty slot of a
// bucket, claimed with CAS; a lane that loses reads the bucket again, unless
// the winner inserted the same column), which keeps every bucket's filled
// slots a prefix and every column in one slot when several wavefronts insert
// at once.  probe_claim returns 0, or -1 when the claims would pass cap.
template <int BK, int NC = NCH>
__device__ __forceinline__ int probe_claim(int *keys, const int (&col)[NC], const bool (&act)[NC],
                                           unsigned nbk, unsigned tsize, unsigned priv,
                                           int *claims, int cap, int &lost,
                                           unsigned (&slot)[NC]) {
    const unsigned dummy = tsize;
    unsigned h[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        // every lane hashes (idle lanes hold column of element 0): no branch
        const unsigned hb = home_bucket<false>(col[c], nbk) * BK;
        h[c] = act[c] ? hb : dummy;
        slot[c] = priv;
    }
    while (true) {
        Bucket<BK> kb[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) kb[c] = load_bucket<BK>(keys, h[c]);
        unsigned at[NC];
        bool ins[NC], anyins = false;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const bool pend = h[c] != dummy;
            unsigned eh, nf;
            bucket_scan<BK>(kb[c], col[c], eh, nf);
            const bool found = eh < (unsigned) BK;
            at[c] = h[c] + (found ? eh : nf);
            const bool hit = pend & found;
            ins[c] = pend & !found & (nf < (unsigned) BK);
            slot[c] = hit ? at[c] : slot[c];
            unsigned hn = h[c] + BK;
            hn = hn >= tsize ? hn - tsize : hn;
            h[c] = (hit | !pend) ? dummy : (ins[c] ? h[c] : hn);
            anyins = anyins | ins[c];
        }
        if (__any(anyins)) {
            // count the claims on the block's counter before making them: the
            // count bounds the filled slots, and keeping it within tsize - 8
            // leaves empty slots, so every probe ends
            // (claims this wavefront lost earlier are given back here)
            int n = 0;
#pragma unroll
            for (int c = 0; c < NC; c++) n += popc64(ballot64(ins[c]));
            int base = 0;
            if (lane_id() == 0) base = atomicAdd(claims, n - lost);
            base = __builtin_amdgcn_readfirstlane(base);
            lost = 0;
            if (base + n > cap) return -1;
            int old[NC];
#pragma unroll
            for (int c = 0; c < NC; c++)
                old[c] = ins[c] ? atomicCAS(&keys[at[c]], EMPTY, col[c]) : EMPTY;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                // lost to another wavefront inserting the same column: a hit
                const bool won = ins[c] & (old[c] == EMPTY);
                const bool met = ins[c] & (old[c] == col[c]);
                slot[c] = (won | met) ? at[c] : slot[c];
                h[c] = (won | met) ? dummy : h[c];
                lost += popc64(ballot64(ins[c] & !won));
            }
        }
        bool anyp = false;
#pragma unroll
        for (int c = 0; c < NC; c++) anyp = anyp | (h[c] != dummy);
        if (!__any(anyp)) break;
    }
    return 0;
}

// The unordered engine accumulates in double components: on gfx950 an LDS
// atomic add of doubles (ds_add_f64) runs about 8x faster than one of floats
// (ds_add_f32): 0.67 against 5.1 ms for the same 10^9 random adds.  Single
// precision products are rounded to float, summed in double and the sum
// rounded to float once.
template <typename T> struct UAcc;
template <> struct UAcc<float> {
    using type = double;
    __device__ static double to(float v) { return (double) v; }
    __device__ static float from(double v) { return (float) v; }
};
template <> struct UAcc<double> {
    using type = double;
    __device__ static double to(double v) { return v; }
    __device__ static double from(double v) { return v; }
};
template <> struct UAcc<cf32> {
    using type = cf64;
    __device__ static cf64 to(cf32 v) { return {(double) v.x, (double) v.y}; }
    __device__ static cf32 from(cf64 v) { return {(float) v.x, (float) v.y}; }
};
template <> struct UAcc<cf64> {
    using type = cf64;
    __device__ static cf64 to(cf64 v) { return v; }
    __device__ static cf64 from(cf64 v) { return v; }
};

// One row i of the unordered engine on the block's table (keys, vals:
// cleared, res_s and ovf_s zero); ends with wavefront 0's epilogue, after
// a barrier every wavefront reaches.
template <typename T, int W, int BK, bool SYM>
__device__ void unord_row(const SpParams &p, unsigned tsize, int i, int *keys,
                          typename UAcc<T>::type *vals, int &res_s, int &ovf_s,
                          long long (&prods_s)[W], T (&trx_s)[W]) {
    using E = Elem<T>;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const T *Av = (const T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    const int *Ai = p.Ai, *Bi = p.Bi, *Bn = p.Bn;
    const int nA = p.An[i];
    const unsigned rowA = (unsigned) i * (unsigned) p.MA;
    const unsigned MB = (unsigned) p.MB;
    const unsigned nbk = tsize / BK, priv = tsize + 8u + (unsigned) lane;
    const int cap = (int) tsize - 8;

    long long prods = 0;
    T trx = E::zero();
    bool stop = false;  // wave-uniform: the row's claims passed the table
    int lost = 0;       // wave-uniform: claims that filled no slot, not yet given back

    auto load_b = [&](BRow<T> &r, int j, int nb, bool act) {
        const unsigned base = (unsigned) j * MB;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int kp = c * 64 + lane;
            const unsigned at = (act & (kp < nb)) ? base + (unsigned) kp : 0u;
            r.col[c] = Bi[at];
            r.val[c] = Bv[at];
        }
    };
    // add a times the chunks of r (entries k0.. of row j of B) into the table
    // (-1: the table is full, the row goes to the global-table pass)
    auto chunks = [&](const BRow<T> &r, int k0, int nb, T a) {
        bool act[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            act[c] = k0 + c * 64 + lane < nb;
            if constexpr (SYM) act[c] = act[c] & (r.col[c] >= i);
        }
        unsigned slot[NCH];
        if (probe_claim<BK>(keys, r.col, act, nbk, tsize, priv, &res_s, cap, lost, slot) < 0)
            return -1;
#pragma unroll
        for (int c = 0; c < NCH; c++)
            if (act[c]) lds_add(&vals[slot[c]], UAcc<T>::to(E::mul(a, r.val[c])));
        return 0;
    };

    for (int c0 = 0; c0 < nA && !stop; c0 += 64) {
        const int jl = c0 + lane;
        const unsigned at = jl < nA ? rowA + (unsigned) jl : 0u;
        const T av = Av[at];
        const int aj = jl < nA ? Ai[at] : 0;
        const int an = Bn[aj];
        const int steps = nA - c0 < 64 ? nA - c0 : 64;
        BRow<T> buf0, buf1;
        int t = w;
        T a = bcast<64>(av, t < steps ? t : 0);
        int j = bcast<64>(aj, t < steps ? t : 0);
        int nb = bcast<64>(an, t < steps ? t : 0);
        load_b(buf0, j, nb, t < steps);
        auto step = [&](BRow<T> &cur, BRow<T> &nxt) {
            const int t1 = t + W;
            const int src = t1 < steps ? t1 : 0;
            const T a1 = bcast<64>(av, src);
            const int j1 = bcast<64>(aj, src);
            const int nb1 = bcast<64>(an, src);
            load_b(nxt, j1, nb1, t1 < steps);
            if (t < steps && !stop) {
                prods += nb;
                if (j == i) trx = E::add(trx, a);
                int rc = chunks(cur, 0, nb, a);
                for (int k0 = NCH * 64; k0 < nb && rc == 0; k0 += NCH * 64) {
                    BRow<T> extra;
                    const unsigned base = (unsigned) j * MB;
#pragma unroll
                    for (int c = 0; c < NCH; c++) {
                        const int kp = k0 + c * 64 + lane;
                        const unsigned at2 = kp < nb ? base + (unsigned) kp : 0u;
                        extra.col[c] = Bi[at2];
                        extra.val[c] = Bv[at2];
                    }
                    rc = chunks(extra, k0, nb, a);
                }
                if (rc < 0) {
                    stop = true;
                    if (lane == 0) ovf_s = 1;
                }
            }
            t = t1;
            a = a1;
            j = j1;
            nb = nb1;
        };
        while (t < steps && !stop) {
            step(buf0, buf1);
            if (t < steps && !stop) step(buf1, buf0);
        }
    }
    if (lane == 0) {
        prods_s[w] = prods;
        trx_s[w] = trx;
    }
    __syncthreads();  // every wavefront's adds are in
    if (w != 0) return;
    if (ovf_s) {
        if (lane == 0) {
            const int q = atomicAdd(&p.st->table_overflow_rows, 1);
            p.ovf[q] = i;
            atomicAdd((unsigned long long *) &p.st->prof_ovf_rows[PrecOf<T>::value], 1ull);
        }
        return;
    }
    // epilogue (wavefront 0): the kept columns in table order
    const int MC = p.MC;
    const size_t rowC = (size_t) i * MC;
    T *Cv = (T *) p.Cv;
    int kept = 0, count = 0;
    T d2 = E::zero();
    for (unsigned s0 = 0; s0 < tsize; s0 += 64) {
        const unsigned s = s0 + (unsigned) lane;
        const int col = s < tsize ? keys[s] : EMPTY;
        const T v = UAcc<T>::from(vals[s]);
        const bool occ = col != EMPTY;
        const bool isdiag = occ && col == i;
        const bool keep = isdiag || (occ && E::above(v, p.thr));
        const uint64_t km = ballot64(keep);
        const int pos = kept + popc64(km & ltmask);
        if (keep && pos < MC) {
            p.Ci[rowC + pos] = col;
            Cv[rowC + pos] = v;
        }
        const uint64_t dm = ballot64(isdiag);
        if (dm) d2 = shfl_any(v, __ffsll((unsigned long long) dm) - 1);
        kept += popc64(km);
        count += popc64(ballot64(occ));
    }
    if (lane == 0) {
        long long pr = 0;
        T tx = E::zero();
        for (int k = 0; k < W; k++) {
            pr += prods_s[k];
            tx = E::add(tx, trx_s[k]);
        }
        p.Cn[i] = kept < MC ? kept : MC;
        p.rowstat[i] = make_int4(kept, count, (int) pr, nA);
        if (p.kind == 0) {
            p.rowtr[i] = E::re(tx);
            p.rowtr[p.N + i] = E::re(d2);
        }
    }
}

// SYM (X symmetric or Hermitian, SP2 on such an H): row i forms only the
// columns k >= i of X^2 (the other lanes of a chunk sit idle: no probe, no
// add), the table holds about half the row's columns, and k_sym_mirror
// copies each kept (i, k), k > i, to (k, i) conjugated.
template <typename T, int W, int BK, bool SYM>
__global__ __launch_bounds__(64 * W) void k_spgemm_unord(SpParams p, unsigned tsize) {
    using A = typename UAcc<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int res_s, ovf_s, item_s;
    __shared__ long long prods_s[W];
    __shared__ T trx_s[W];
    if (p.skip && *p.skip) return;
    const unsigned stride = tsize + table_pad<64>();
    A *vals = (A *) smem;
    int *keys = (int *) (smem + (size_t) stride * sizeof(A));
    auto clear = [&]() {
        for (unsigned s = threadIdx.x; s < stride; s += 64 * W) {
            keys[s] = EMPTY;
            vals[s] = Elem<A>::zero();
        }
        if (threadIdx.x == 0) {
            res_s = 0;
            ovf_s = 0;
        }
    };
    clear();
    __syncthreads();
    if (!p.qlist) {
        const int row = p.remap ? xcd_block(blockIdx.x, gridDim.x) : (int) blockIdx.x;
        if (row >= p.N) return;
        if (p.row_sys && p.sys_done[p.row_sys[row]]) {  // uniform
            if (threadIdx.x == 0) p.rowstat[row] = make_int4(-1, 0, 0, 0);
            return;
        }
        unord_row<T, W, BK, SYM>(p, tsize, row, keys, vals, res_s, ovf_s, prods_s, trx_s);
        return;
    }
    // persistent over the work queue, one listed row per take; the next
    // row is taken while the current one runs (its atomic's latency hides
    // behind the row's loads).  item_s is read by every thread after a
    // barrier and written again only after the next one.
    const int n = p.qinfo[0];
    if (threadIdx.x == 0) item_s = atomicAdd(&p.qinfo[1], 1);
    __syncthreads();
    int q = item_s;
    while (q < n) {
        int nxt = 0;
        if (threadIdx.x == 0) nxt = atomicAdd(&p.qinfo[1], 1);
        unord_row<T, W, BK, SYM>(p, tsize, p.qlist[q], keys, vals, res_s, ovf_s, prods_s, trx_s);
        __syncthreads();  // the epilogue's table reads are done
        clear();
        if (threadIdx.x == 0) item_s = nxt;
        __syncthreads();
        q = item_s;
    }
}

// Mirror of the upper rows: entry (i, k), k > i, of row i's upper part goes
// to row k at position Cn[k] + fill[k]++ (conjugated).  One wavefront per
// row; rows of finished systems are skipped.
template <typename T>
__global__ __launch_bounds__(256) void k_sym_mirror(SpParams p, int *fill) {
    if (p.skip && *p.skip) return;
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    if (i >= p.N) return;
    if (p.row_sys && p.sys_done[p.row_sys[i]]) return;
    const int lane = threadIdx.x & 63;
    const int MC = p.MC;
    T *Cv = (T *) p.Cv;
    const int u = p.Cn[i];
    for (int e = lane; e < u; e += 64) {
        const int k = p.Ci[(size_t) i * MC + e];
        if (k > i) {
            const int pos = p.Cn[k] + atomicAdd(&fill[k], 1);
            if (pos < MC) {
                p.Ci[(size_t) k * MC + pos] = i;
                Cv[(size_t) k * MC + pos] = Elem<T>::conj(Cv[(size_t) i * MC + e]);
            }
        }
    }
}

// Final row lengths after the mirror: upper + mirrored entries (capped at
// MC; the row record keeps the full count for the overflow check), and the
// fill counters cleared for the next product.
__global__ __launch_bounds__(256) void k_sym_finish(SpParams p, int *fill) {
    if (p.skip && *p.skip) return;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= p.N) return;
    if (p.row_sys && p.sys_done[p.row_sys[k]]) return;
    const int f = fill[k];
    if (f == 0) return;
    fill[k] = 0;
    const int tot = p.Cn[k] + f;
    p.Cn[k] = tot < p.MC ? tot : p.MC;
    p.rowstat[k].x += f;
}

// ---------------------------------------------------------------------------
// Window engine, for wide rows (X^2 of a matrix whose product rows touch more
// distinct columns than an LDS hash table holds: the 32^3 lattice density
// matrix at hopping -0.5 keeps ~2600 entries per row and its X^2 rows touch
// ~10^4 of the 32768 columns).  The column range is cut into windows of WC =
// 2^WB columns.  A pack pass (k_win_pack) copies every row of X with its
// entries grouped by window, as (value, column) records in consecutive
// 32-bit words (one load per entry), with per-row window offsets (boff) and
// a mask of the windows the row touches.  Then one block per output row i
// takes the windows its rows j touch, in column order; per window its
// wavefronts take the entries (a, j) of row i round-robin and add a * X[j,k]
// for the segment of row j in the window straight into a dense LDS window
// (ds_add_f64: no hashing, no probing, no capacity limit), and the epilogue
// walks the window in column order: output rows come out sorted by column.
// Sums are formed in arrival order, as on the unordered engine.
//
// SYM (X symmetric / Hermitian): row i forms only its columns k >= i (the
// windows below i's are skipped, lanes with k < i sit idle) and
// k_sym_mirror fills the rest.  FUSE (SP2): the epilogue applies the SP2
// update X <- alpha X + beta X^2 (alpha, beta chosen beforehand from tr X and
// tr X^2 = sum |X_ik|^2, which the pack pass reduces) with the thresholds of
// bml_multiply_x2 and bml_add_norm, accumulates ||X - X^2||_F^2 over the row,
// and writes the new row in place of X's (the products read only the packed
// copy).

// A packed entry: the value's words, then the column.
template <typename T>
struct PkW {
    static constexpr int n = (int) (sizeof(T) / 4) + 1;
};
template <int NW>
struct __attribute__((aligned(4))) Words {
    unsigned w[NW];
};
template <typename T>
__device__ __forceinline__ void pk_get(const unsigned *P, size_t e, T &v, int &c) {
    constexpr int NW = PkW<T>::n;
    const Words<NW> x = *reinterpret_cast<const Words<NW> *>(P + e * NW);
    union {
        T t;
        unsigned u[NW - 1];
    } cv;
#pragma unroll
    for (int q = 0; q < NW - 1; q++) cv.u[q] = x.w[q];
    v = cv.t;
    c = (int) x.w[NW - 1];
}
// Entry k of a packed row starting at Pj: a scalar base and a 32-bit
// per-lane byte offset (one global load with a scalar address; rows hold
// fewer than 2^24 entries).
template <typename T>
__device__ __forceinline__ void pk_get_at(const unsigned *Pj, unsigned k, T &v, int &c) {
    constexpr int NW = PkW<T>::n;
    const unsigned off = __umul24(k, 4u * NW);
    const Words<NW> x = *reinterpret_cast<const Words<NW> *>((const char *) Pj + off);
    union {
        T t;
        unsigned u[NW - 1];
    } cv;
#pragma unroll
    for (int q = 0; q < NW - 1; q++) cv.u[q] = x.w[q];
    v = cv.t;
    c = (int) x.w[NW - 1];
}
template <typename T>
__device__ __forceinline__ int pk_col(const unsigned *P, size_t e) {
    return (int) P[e * PkW<T>::n + PkW<T>::n - 1];
}
template <typename T>
__device__ __forceinline__ void pk_put(unsigned *P, size_t e, T v, int c) {
    constexpr int NW = PkW<T>::n;
    union {
        T t;
        unsigned u[NW - 1];
    } cv;
    cv.t = v;
    Words<NW> x;
#pragma unroll
    for (int q = 0; q < NW - 1; q++) x.w[q] = cv.u[q];
    x.w[NW - 1] = (unsigned) c;
    *reinterpret_cast<Words<NW> *>(P + e * NW) = x;
}

__device__ __forceinline__ unsigned long long or64_wave(unsigned long long m) {
    unsigned lo = (unsigned) m, hi = (unsigned) (m >> 32);
    for (int off = 32; off > 0; off >>= 1) {
        lo |= (unsigned) __shfl_xor((int) lo, off, 64);
        hi |= (unsigned) __shfl_xor((int) hi, off, 64);
    }
    return ((unsigned long long) hi << 32) | lo;
}

struct WinParams {
    int N, MP, nwin, MC;
    const unsigned *P;               // packed rows: N x MP entries
    const int *boff;                 // [N][nwin + 1] window offsets within each packed row
    const unsigned long long *wmask; // [N] windows each row touches
    const double *rowmax;            // [N] largest |X_jk| of each row (row blocks' scale)
    int fx_extra;                    // row blocks: scale exponent lowered by this (wide rows)
    void *Cv;
    int *Ci, *Cn;
    double thr;
    double *rowtr;    // [2N] x2: per-row tr X, tr X^2 partials
    double *rownorm;  // [N] FUSE: per-row ||X - X^2||^2 partials
    const double *ab; // FUSE: (alpha, beta) per system (index row_sys[i], or 0)
    int4 *rowstat;
    const int *skip;
    const int *row_sys, *sys_done;
    int remap;
    // row blocks (k_spgemm_winb): per block of R rows the union of their
    // columns and the R coefficients of each
    const int *Uj;     // [nblk][ustride] union columns j
    const void *Uc;    // [nblk][ustride][R] X[i0 + r, j] (T; zero where row r lacks j)
    const int *Un;     // [nblk] union sizes
    size_t ustride;
    unsigned long long *streamed;  // prof_streamed of the precision
};

// Pack pass: one wavefront per row.  Row j's entries are grouped by window
// (counted, scanned, scattered with LDS counters; the order within a window
// is the arrival order), boff[j][c] is the first entry of window c,
// boff[j][nwin] the row length, wmask[j] the windows with entries.  With
// rowtr the row's tr X (its diagonal) and sum |X_jk|^2 partials are written
// too.
template <typename T>
__global__ __launch_bounds__(256) void k_win_pack(int N, int M, const T *Xv, const int *Xi,
                                                  const int *Xn, int MP, int nwin, int wb,
                                                  unsigned *P, int *boff,
                                                  unsigned long long *wmask, double *rowtr,
                                                  double *rowmax, const int *skip,
                                                  const int *row_sys, const int *sys_done) {
    __shared__ int cnt[4][65];
    if (skip && *skip) return;  // uniform
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + w;
    const bool live = j < N && !(row_sys && sys_done[row_sys[j]]);
    const int n = live ? Xn[j] : 0;
    const size_t rowX = live ? (size_t) j * M : 0;
    cnt[w][lane] = 0;
    __syncthreads();
    double trx = 0.0, sq = 0.0, mx = 0.0;
    for (int e = lane; e < n; e += 64) {
        const int c = Xi[rowX + e];
        atomicAdd(&cnt[w][c >> wb], 1);
        if (rowtr || rowmax) {
            const T v = Xv[rowX + e];
            if (c == j) trx += Elem<T>::re(v);
            sq += Elem<T>::abs2(v);
            mx = fmax(mx, Elem<T>::mag(v));
        }
    }
    if (rowmax && live) {
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        if (lane == 0) rowmax[j] = mx;
    }
    __syncthreads();
    const int x = lane < nwin ? cnt[w][lane] : 0;
    int inc = x;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    const unsigned long long m = ballot64(x > 0);
    __syncthreads();
    cnt[w][lane] = inc - x;  // next free position of window `lane`
    if (live) {
        int *bo = boff + (size_t) j * (nwin + 1);
        if (lane < nwin) bo[lane] = inc - x;
        if (lane == 0) {
            bo[nwin] = n;
            wmask[j] = m;
        }
    }
    __syncthreads();
    const size_t rowP = (size_t) (live ? j : 0) * MP;
    for (int e = lane; e < n; e += 64) {
        const int c = Xi[rowX + e];
        const T v = Xv[rowX + e];
        const int pos = atomicAdd(&cnt[w][c >> wb], 1);
        pk_put<T>(P, rowP + pos, v, c);
    }
    if (rowtr && live) {
        for (int off = 32; off > 0; off >>= 1) {
            trx += __shfl_down(trx, off, 64);
            sq += __shfl_down(sq, off, 64);
        }
        if (lane == 0) {
            rowtr[j] = trx;
            rowtr[N + j] = sq;
        }
    }
}

// One block of W wavefronts per output row; LDS: the window's accumulators
// (double components) and, with FUSE, X's row i in the window.
template <typename T, int W, int WB, bool SYM, bool FUSE>
__global__ __launch_bounds__(64 * W) void k_spgemm_win(WinParams p) {
    using E = Elem<T>;
    using A = typename UAcc<T>::type;
    constexpr int WC = 1 << WB;
    constexpr int NT = 64 * W;
    constexpr int PER = WC / W;  // epilogue columns per wavefront
    static_assert(PER % 64 == 0, "window too small for the block");
    extern __shared__ __align__(16) unsigned char smem[];
    A *win = (A *) smem;  // WC accumulators, then one dummy slot per lane
    T *xwin = (T *) (smem + sizeof(A) * (WC + 64));
    const int dummy = WC + (int) (threadIdx.x & 63);
    __shared__ unsigned long long mask_s;
    __shared__ int cnt_s[W], dflag_s;
    __shared__ long long prods_s[W];
    __shared__ double nrm_s[W];
    __shared__ T trx_s, d2_s;
    if (p.skip && *p.skip) return;  // uniform
    const int i = p.remap ? xcd_block(blockIdx.x, gridDim.x) : (int) blockIdx.x;
    if (i >= p.N) return;
    const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
    if (p.row_sys && p.sys_done[p.row_sys[i]]) {  // uniform
        if (tid == 0) p.rowstat[i] = make_int4(-1, 0, 0, 0);
        return;
    }
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const unsigned *P = p.P;
    const size_t rowP = (size_t) i * p.MP;
    const int nw1 = p.nwin + 1;
    const int *bo_i = p.boff + (size_t) i * nw1;
    const int nA = bo_i[p.nwin];
    if (tid == 0) {
        mask_s = 0ull;
        trx_s = E::zero();
        d2_s = E::zero();
    }
    __syncthreads();
    // the windows row i touches: the union of its rows' masks
    unsigned long long m = 0ull;
    // the entries of B streamed by this wavefront (wave-uniform): the
    // segments of the windows taken, so with SYM about half of the full
    // product's partial products
    long long prods = 0;
    for (int e = tid; e < nA; e += NT) {
        T a;
        int j;
        pk_get<T>(P, rowP + e, a, j);
        m |= p.wmask[j];
        if (!FUSE && j == i) trx_s = a;  // one entry per row holds the diagonal
    }
    m = or64_wave(m);
    if (lane == 0 && m) atomicOr(&mask_s, m);
    __syncthreads();
    m = mask_s;
    // the update alpha X + beta X^2 also keeps X's own entries in windows
    // that no product reaches (a row without its diagonal entry)
    if constexpr (FUSE) m |= p.wmask[i];
    if constexpr (SYM) m &= ~((1ull << (i >> WB)) - 1ull);
    double alpha = 0.0, beta = 0.0;
    if constexpr (FUSE) {
        const int sys = p.row_sys ? p.row_sys[i] : 0;
        alpha = p.ab[2 * sys];
        beta = p.ab[2 * sys + 1];
    }
    const int MC = p.MC;
    const size_t rowC = (size_t) i * MC;
    T *Cv = (T *) p.Cv;
    int kept = 0, distinct = 0;  // block-uniform running counts
    double nrm = 0.0;            // this lane's norm partial
    while (m) {
        const int c = __ffsll((long long) m) - 1;
        m &= m - 1ull;
        const int cbase = c << WB;
        for (int s = tid; s < WC; s += NT) {
            win[s] = Elem<A>::zero();
            if constexpr (FUSE) xwin[s] = E::zero();
        }
        if (tid == 0) dflag_s = SYM ? (nA > 0) : 0;
        bool dtouch = false;  // this lane added into column i (!SYM)
        __syncthreads();
        // products: wavefront w takes groups g = w, w + W, ... of 64 entries
        // of row i; lane q of a group holds entry q's (a, j) and the bounds
        // of row j's segment in window c
        for (int g0 = w * 64; g0 < nA; g0 += NT) {
            const int e = g0 + lane;
            T aq = E::zero();
            int jq = 0, lo = 0, hi = 0;
            if (e < nA) {
                pk_get<T>(P, rowP + e, aq, jq);
                const int *bo = p.boff + (size_t) jq * nw1;
                lo = bo[c];
                hi = bo[c + 1];
            }
            uint64_t todo = ballot64(hi > lo);
            while (todo) {
                const int q = __ffsll((long long) todo) - 1;
                todo &= todo - 1ull;
                const T a = bcast<64>(aq, q);
                const int j = __builtin_amdgcn_readlane(jq, q);
                const int l0 = __builtin_amdgcn_readlane(lo, q);
                const int l1 = __builtin_amdgcn_readlane(hi, q);
                const size_t rb = (size_t) j * p.MP;
                prods += l1 - l0;
                int k0 = l0 + lane;
                // two entries per lane per round, branch-free: both loads
                // issue before the adds, and (SYM) the products of columns
                // below i go to the lane's own dummy slot past the window
                for (; k0 + 64 < l1; k0 += 128) {
                    T b0, b1;
                    int c0, c1;
                    pk_get<T>(P, rb + k0, b0, c0);
                    pk_get<T>(P, rb + k0 + 64, b1, c1);
                    const int s0 = (!SYM || c0 >= i) ? c0 - cbase : dummy;
                    const int s1 = (!SYM || c1 >= i) ? c1 - cbase : dummy;
                    lds_add(&win[s0], UAcc<T>::to(E::mul(a, b0)));
                    lds_add(&win[s1], UAcc<T>::to(E::mul(a, b1)));
                    dtouch = dtouch | (c0 == i) | (c1 == i);
                }
                if (k0 < l1) {
                    T b0;
                    int c0;
                    pk_get<T>(P, rb + k0, b0, c0);
                    const int s0 = (!SYM || c0 >= i) ? c0 - cbase : dummy;
                    lds_add(&win[s0], UAcc<T>::to(E::mul(a, b0)));
                    dtouch = dtouch | (c0 == i);
                }
            }
        }
        if constexpr (!SYM) {
            if (__any(dtouch) && lane == 0) dflag_s = 1;
        }
        if constexpr (FUSE) {
            // X's own row i in this window
            for (int e = bo_i[c] + tid; e < bo_i[c + 1]; e += NT) {
                T v;
                int k;
                pk_get<T>(P, rowP + e, v, k);
                xwin[k - cbase] = v;
            }
        }
        __syncthreads();
        const bool dtouched = dflag_s != 0;
        // epilogue: wavefront w walks columns [w PER, (w+1) PER) of the
        // window twice, counting, then writing at its offset
        auto entry = [&](int s, T &out, bool &keep, bool &nz, double &dn) {
            const int k = cbase + s;
            const bool valid = k < p.N && (!SYM || k >= i);
            const T y = UAcc<T>::from(win[s]);
            nz = valid && (E::mag(y) != 0.0 || k == i);
            const bool ykeep = valid && (k == i ? dtouched : E::above(y, p.thr));
            dn = 0.0;
            if constexpr (FUSE) {
                const T x = xwin[s];
                const T yk = ykeep ? y : E::zero();
                if (valid) dn = E::abs2(E::sub(x, yk)) * ((SYM && k > i) ? 2.0 : 1.0);
                out = E::axpy(E::axpy(E::zero(), alpha, x), beta, yk);
                keep = valid && E::above(out, p.thr);
            } else {
                out = y;
                keep = ykeep;
            }
        };
        int cw = 0, dw = 0;
        for (int s = w * PER + lane; s < (w + 1) * PER; s += 64) {
            T v;
            bool keep, nz;
            double dn;
            entry(s, v, keep, nz, dn);
            nrm += dn;
            cw += popc64(ballot64(keep));
            dw += popc64(ballot64(nz));
            if (!FUSE && cbase + s == i) d2_s = v;
        }
        if (lane == 0) {
            cnt_s[w] = cw;
            prods_s[w] = dw;  // distinct counts, summed below
        }
        __syncthreads();
        int pos0 = kept, tot = 0, dtot = 0;
        for (int q = 0; q < W; q++) {
            pos0 += q < w ? cnt_s[q] : 0;
            tot += cnt_s[q];
            dtot += (int) prods_s[q];
        }
        for (int s = w * PER + lane; s < (w + 1) * PER; s += 64) {
            T v;
            bool keep, nz;
            double dn;
            entry(s, v, keep, nz, dn);
            const uint64_t km = ballot64(keep);
            const int pos = pos0 + popc64(km & ltmask);
            if (keep && pos < MC) {
                p.Ci[rowC + pos] = cbase + s;
                Cv[rowC + pos] = v;
            }
            pos0 += popc64(km);
        }
        kept += tot;
        distinct += dtot;
        __syncthreads();  // the window and cnt_s are reused by the next window
    }
    // row records, traces, norm
    if (lane == 0) prods_s[w] = prods;
    double r = nrm;
    if constexpr (FUSE) {
        for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
        if (lane == 0) nrm_s[w] = r;
    }
    __syncthreads();
    if (tid == 0) {
        long long pr = 0;
        double nr = 0.0;
        for (int q = 0; q < W; q++) {
            pr += prods_s[q];
            if constexpr (FUSE) nr += nrm_s[q];
        }
        p.Cn[i] = kept < MC ? kept : MC;
        p.rowstat[i] = make_int4(kept, distinct, (int) (pr < INT_MAX ? pr : INT_MAX), nA);
        if constexpr (FUSE) {
            p.rownorm[i] = nr;
        } else {
            p.rowtr[i] = E::re(trx_s);
            p.rowtr[p.N + i] = E::re(d2_s);
        }
    }
}

// ---------------------------------------------------------------------------
// Window engine on row blocks.  On a lattice, neighbouring rows of the
// density matrix reference mostly the same rows j (on the 32^3 iterate at
// hopping -0.5, four consecutive rows reference ~1.3x the rows one does), and
// the one-row kernel streams each row j of X once per output row that
// references it.  Here a block takes R consecutive output rows i0 .. i0+R-1:
// k_win_union lists the union U of their columns j with the R coefficients
// X[i0 + r, j] (zero where row r lacks j), and k_spgemm_winb walks U once
// per window, streaming each row j's segment once and adding c_r * X[j, k]
// into the window of every row r with c_r != 0 (a wave-uniform test).  The
// windows, epilogue, symmetric and fused cases are those of k_spgemm_win,
// one row at a time.

// Union of the columns of rows i0 .. i0+R-1 (one block per row block): an
// LDS hash of TS slots (TS > R*M, so a probe always ends) gives each
// distinct column a slot, a block-wide scan ranks the filled slots, and the
// coefficients are written at the ranks.  Rows of finished systems, and
// rows past N, count as empty.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_win_union(int N, int M, const T *Xv, const int *Xi,
                                                   const int *Xn, int tbits, int *Uj, T *Uc,
                                                   int *Un, size_t ustride, const int *skip,
                                                   const int *row_sys, const int *sys_done) {
    extern __shared__ __align__(16) int ukeys[];  // TS keys, then TS 16-bit ranks
    __shared__ int part_s[256];
    if (skip && *skip) return;  // uniform
    const int TS = 1 << tbits;
    uint16_t *rank = (uint16_t *) (ukeys + TS);
    const int b = blockIdx.x, i0 = b * R, tid = threadIdx.x;
    for (int t = tid; t < TS; t += 256) ukeys[t] = -1;
    __syncthreads();
    auto slot_of = [&](int j) -> int {
        unsigned h = ((unsigned) j * 2654435761u) >> (32 - tbits);
        while (true) {
            const int k = ukeys[h];
            if (k == j) return (int) h;
            if (k < 0) {
                const int old = atomicCAS(&ukeys[h], -1, j);
                if (old < 0 || old == j) return (int) h;
            }
            h = (h + 1) & (unsigned) (TS - 1);
        }
    };
    int nr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = i0 + r;
        nr[r] = (i < N && !(row_sys && sys_done[row_sys[i]])) ? Xn[i] : 0;
        for (int e = tid; e < nr[r]; e += 256) (void) slot_of(Xi[(size_t) i * M + e]);
    }
    __syncthreads();
    // ranks of the filled slots in slot order: each thread counts a
    // contiguous chunk, a block scan offsets the chunks
    const int per = TS / 256;
    int cnt = 0;
    for (int t = tid * per; t < (tid + 1) * per; t++) cnt += ukeys[t] >= 0;
    part_s[tid] = cnt;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int v = tid >= d ? part_s[tid - d] : 0;
        __syncthreads();
        part_s[tid] += v;
        __syncthreads();
    }
    int u = part_s[tid] - cnt;
    const int total = part_s[255];
    int *uj = Uj + (size_t) b * ustride;
    T *uc = Uc + (size_t) b * ustride * R;
    for (int t = tid * per; t < (tid + 1) * per; t++) {
        const int k = ukeys[t];
        if (k >= 0) {
            rank[t] = (uint16_t) u;
            uj[u] = k;
#pragma unroll
            for (int r = 0; r < R; r++) uc[(size_t) u * R + r] = Elem<T>::zero();
            u++;
        }
    }
    if (tid == 0) Un[b] = total;
    __syncthreads();  // the zeroed coefficients land before they are set
#pragma unroll
    for (int r = 0; r < R; r++) {
        const size_t row = (size_t) (i0 + r) * M;
        for (int e = tid; e < nr[r]; e += 256) {
            const int h = slot_of(Xi[row + e]);
            uc[(size_t) rank[h] * R + r] = Xv[row + e];
        }
    }
}

// The same union from a bitmap of the N columns (N/8 bytes of LDS, and a
// word prefix as large, against the hash's 6 * 2^14 bytes at M = 3072, which
// leaves one block per CU): each entry sets its column's bit, and a block
// scan of the words' bit counts ranks the columns, so U comes out in column
// order.  For N <= 2^16.  The order U comes in matters to the product: on
// the 32^3 SP2 step (one MI355X) the FP32 step took 55.6 ms with this
// union against 56.4 with the hash's, but the FP64 product 58.8 against
// 54.1 ms (and 64.4 ms with U dealt round-robin over the groups, each group
// sampling the whole column range), so the FP64 and complex products keep
// the hash's order.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_win_union_bm(int N, int M, const T *Xv, const int *Xi,
                                                      const int *Xn, int *Uj, T *Uc, int *Un,
                                                      size_t ustride, const int *skip,
                                                      const int *row_sys, const int *sys_done) {
    extern __shared__ __align__(16) unsigned bmw[];  // NW bit words, then NW word ranks
    __shared__ int part_s[256];
    if (skip && *skip) return;  // uniform
    const int NW = (N + 31) >> 5;
    int *wr = (int *) (bmw + NW);
    const int b = blockIdx.x, i0 = b * R, tid = threadIdx.x;
    for (int t = tid; t < NW; t += 256) bmw[t] = 0u;
    __syncthreads();
    int nr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = i0 + r;
        nr[r] = (i < N && !(row_sys && sys_done[row_sys[i]])) ? Xn[i] : 0;
        for (int e = tid; e < nr[r]; e += 256) {
            const int j = Xi[(size_t) i * M + e];
            atomicOr(&bmw[j >> 5], 1u << (j & 31));
        }
    }
    __syncthreads();
    // each thread ranks a contiguous run of words; a block scan offsets the runs
    const int per = (NW + 255) / 256, w0 = tid * per;
    const int w1 = w0 + per < NW ? w0 + per : NW;
    int cnt = 0;
    for (int t = w0; t < w1; t++) cnt += __popc(bmw[t]);
    part_s[tid] = cnt;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int v = tid >= d ? part_s[tid - d] : 0;
        __syncthreads();
        part_s[tid] += v;
        __syncthreads();
    }
    int u = part_s[tid] - cnt;
    const int total = part_s[255];
    int *uj = Uj + (size_t) b * ustride;
    T *uc = Uc + (size_t) b * ustride * R;
    for (int t = w0; t < w1; t++) {
        wr[t] = u;
        unsigned m = bmw[t];
        while (m) {
            const int bit = __ffs(m) - 1;
            m &= m - 1u;
            uj[u] = (t << 5) + bit;
            u++;
        }
    }
    for (int k = tid; k < total * R; k += 256) uc[k] = Elem<T>::zero();
    if (tid == 0) Un[b] = total;
    __syncthreads();  // ranks in LDS; the zeroed coefficients land before they are set
#pragma unroll
    for (int r = 0; r < R; r++) {
        const size_t row = (size_t) (i0 + r) * M;
        for (int e = tid; e < nr[r]; e += 256) {
            const int j = Xi[row + e];
            const unsigned below = bmw[j >> 5] & ((1u << (j & 31)) - 1u);
            uc[(size_t) (wr[j >> 5] + __popc(below)) * R + r] = Xv[row + e];
        }
    }
}

// Fixed-point window sums (row blocks).  A row block sums its products as
// 64-bit integers in units of 2^-E: integer adds are associative, so the sums
// do not depend on the order the wavefronts' atomics arrive in, and an LDS
// add of 64-bit integers (ds_add_u64) runs about 1.8x as many random-address
// adds per second as one of doubles (ds_add_f64) on gfx950.  E is the
// block's own: with ma the largest |X_ik| of its rows and mb the largest of
// the rows j it streams, 2^(e-1) <= ma * mb < 2^e, E = 50 - e - fx_extra, so
// every product is below 2^50 units and a slot's sum of at most M products
// below 2^62.  A product a * b is rounded to units by one fused multiply-add
// t = (a 2^E) * b + 1.5 * 2^52: t lies in [2^52, 2^53), where consecutive
// doubles are one unit apart, so t's bits minus those of 1.5 * 2^52 are the
// rounded product in two's complement (the constant's low word is zero).
// The rounding error is at most 2^-(E+1) per product, about 2^-52 ma mb:
// below the tolerance of the double-precision unordered sums.  Entries must
// be finite.
template <typename T> struct FxW;
template <> struct FxW<float> {
    static constexpr int NC = 1;  // 64-bit slots per column
    using V = double;             // an entry widened (products of floats are exact)
    __device__ static double widen(float v) { return (double) v; }
    __device__ static float out(const unsigned long long *w, int s, double inv) {
        return (float) ((double) (long long) w[s] * inv);
    }
};
template <> struct FxW<double> {
    static constexpr int NC = 1;
    using V = double;
    __device__ static double widen(double v) { return v; }
    __device__ static double out(const unsigned long long *w, int s, double inv) {
        return (double) (long long) w[s] * inv;
    }
};
template <> struct FxW<cf32> {
    static constexpr int NC = 2;
    using V = cf64;
    __device__ static cf64 widen(cf32 v) { return {(double) v.x, (double) v.y}; }
    __device__ static cf32 out(const unsigned long long *w, int s, double inv) {
        return {(float) ((double) (long long) w[2 * s] * inv),
                (float) ((double) (long long) w[2 * s + 1] * inv)};
    }
};
template <> struct FxW<cf64> {
    static constexpr int NC = 2;
    using V = cf64;
    __device__ static cf64 widen(cf64 v) { return v; }
    __device__ static cf64 out(const unsigned long long *w, int s, double inv) {
        return {(double) (long long) w[2 * s] * inv, (double) (long long) w[2 * s + 1] * inv};
    }
};
__device__ __forceinline__ void lds_add_u64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// (t's bits) - (bits of 1.5 * 2^52): round(x) for t = x + 1.5 * 2^52, |x| < 2^51
__device__ __forceinline__ unsigned long long fx_units(double t) {
    return (unsigned long long) __double_as_longlong(t) - 0x4338000000000000ull;
}
constexpr double FX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
// a (scaled by 2^E) times b, rounded to units, added into slot s
__device__ __forceinline__ void fx_add(unsigned long long *win, int s, double a, double b) {
    lds_add_u64(&win[s], fx_units(__builtin_fma(a, b, FX_MAGIC)));
}
__device__ __forceinline__ void fx_add(unsigned long long *win, int s, cf64 a, cf64 b) {
    // re = ar br - ai bi, im = ar bi + ai br; the inner sum rounds to units
    // first, so each component takes two roundings
    const double re = __builtin_fma(a.x, b.x, __builtin_fma(-a.y, b.y, FX_MAGIC));
    const double im = __builtin_fma(a.x, b.y, __builtin_fma(a.y, b.x, FX_MAGIC));
    lds_add_u64(&win[2 * s], fx_units(re));
    lds_add_u64(&win[2 * s + 1], fx_units(im));
}
__device__ __forceinline__ double fx_scale(double v, double sc) { return v * sc; }
__device__ __forceinline__ cf64 fx_scale(cf64 v, double sc) { return {v.x * sc, v.y * sc}; }

// Waves per SIMD the compiler is asked to fit (registers).  The fused step's
// 4-row blocks of 2^10-column windows need 37 KB (float) and 42 KB (double)
// of LDS, so 3 blocks of 8 wavefronts fit a CU; real types hold their
// registers to 6 waves per SIMD (some spill to scratch), which ran the
// 32^3 SP2 step's kernel 1.2x (float) and 1.15x (double) faster than the
// 4-wave default.  Complex blocks are LDS-bound at 1-2 per CU.
template <typename T> struct WinbOcc { static constexpr int value = 1; };
template <> struct WinbOcc<float> { static constexpr int value = 6; };
template <> struct WinbOcc<double> { static constexpr int value = 6; };
template <typename T, int W, int WB, int R, bool SYM, bool FUSE>
__global__ __launch_bounds__(64 * W, WinbOcc<T>::value) void k_spgemm_winb(WinParams p) {
    using E = Elem<T>;
    using F = FxW<T>;
    using V = typename F::V;
    constexpr int WC = 1 << WB;
    constexpr int NT = 64 * W;
    constexpr int NC = F::NC;
    constexpr int PER = WC / W;  // epilogue columns per wavefront
    static_assert(PER % 64 == 0, "window too small for the block");
    static_assert(WC % R == 0, "a row block must lie in one window");
    extern __shared__ __align__(16) unsigned char smem[];
    // R windows of WC slots, then one dummy slot per lane; a slot is NC
    // 64-bit integers
    unsigned long long *win = (unsigned long long *) smem;
    __shared__ unsigned long long mask_s, mb_s;
    __shared__ int grp_s;
    __shared__ int cnt_s[W], dcnt_s[W], dflag_s[R];
    __shared__ long long prods_s[W][R + 1];
    __shared__ double nrm_s[W][R];
    __shared__ T trx_s[R], d2_s[R];
    if (p.skip && *p.skip) return;  // uniform
    // with p.remap each XCD takes a contiguous range of row blocks
    const int b = p.remap ? xcd_block(blockIdx.x, gridDim.x) : (int) blockIdx.x, i0 = b * R;
    if (i0 >= p.N) return;
    const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
    const int dummy = R * WC + lane;
    // FUSE: each wavefront's PER columns of X's row, staged per row in the
    // epilogue
    T *xbuf = (T *) (smem + sizeof(unsigned long long) * NC * (R * WC + 64)) + w * PER;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const unsigned *P = p.P;
    const int nw1 = p.nwin + 1;
    bool live[R];
    int nA[R];
    double ma = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = i0 + r;
        live[r] = i < p.N && !(p.row_sys && p.sys_done[p.row_sys[i]]);
        nA[r] = live[r] ? p.boff[(size_t) i * nw1 + p.nwin] : 0;
        if (live[r]) ma = fmax(ma, p.rowmax[i]);
    }
    const int nU = p.Un[b];
    const int *Uj = p.Uj + (size_t) b * p.ustride;
    const T *Uc = (const T *) p.Uc + (size_t) b * p.ustride * R;
    if (tid == 0) {
        mask_s = 0ull;
        mb_s = 0ull;
#pragma unroll
        for (int r = 0; r < R; r++) {
            trx_s[r] = E::zero();
            d2_s[r] = E::zero();
        }
    }
    __syncthreads();
    // the windows the block touches (the union of its rows' rows' masks) and
    // the largest magnitude in the rows it streams
    unsigned long long m = 0ull;
    double mb = 0.0;
    for (int e = tid; e < nU; e += NT) {
        const int j = Uj[e];
        m |= p.wmask[j];
        mb = fmax(mb, p.rowmax[j]);
        if constexpr (!FUSE) {
            const int r = j - i0;  // a diagonal coefficient X[i0 + r, i0 + r]
            if (r >= 0 && r < R) trx_s[r] = Uc[(size_t) e * R + r];
        }
    }
    if constexpr (FUSE) {
        // the update alpha X + beta X^2 also keeps X's own entries in windows
        // no product reaches
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r]) m |= p.wmask[i0 + r];
    }
    m = or64_wave(m);
    for (int off = 32; off > 0; off >>= 1) mb = fmax(mb, __shfl_xor(mb, off, 64));
    if (lane == 0) {
        if (m) atomicOr(&mask_s, m);
        // non-negative doubles order as their bits do
        atomicMax(&mb_s, (unsigned long long) __double_as_longlong(mb));
    }
    __syncthreads();
    m = mask_s;
    if constexpr (SYM) m &= ~((1ull << (i0 >> WB)) - 1ull);
    // the block's scale 2^E (block-uniform)
    int ex = 0;
    {
        const double pm = ma * __longlong_as_double((long long) mb_s);
        if (pm > 0.0 && pm <= 1.7976931348623157e308) {
            int e;
            (void) frexp(pm, &e);  // pm < 2^e
            ex = 50 - e - p.fx_extra;
            ex = ex < -1000 ? -1000 : (ex > 1000 ? 1000 : ex);
        }
    }
    const double sc = ldexp(1.0, ex), inv = ldexp(1.0, -ex);
    double alpha = 0.0, beta = 0.0;
    if constexpr (FUSE) {
        alpha = p.ab[0];
        beta = p.ab[1];
    }
    const int MC = p.MC;
    T *Cv = (T *) p.Cv;
    int kept[R], distinct[R];  // block-uniform running counts
    long long prods[R];        // this wavefront's products per row (wave-uniform)
    long long streamed = 0;    // this wavefront's entries of B read (wave-uniform)
#pragma unroll
    for (int r = 0; r < R; r++) {
        kept[r] = 0;
        distinct[r] = 0;
        prods[r] = 0;
        if (lane == 0) nrm_s[w][r] = 0.0;  // FUSE: this wavefront's norm partials
    }
    const int dwin = i0 >> WB;  // the window holding the block's diagonal
    while (m) {
        const int c = __ffsll((long long) m) - 1;
        m &= m - 1ull;
        const int cbase = c << WB;
        for (int s = tid; s < NC * R * WC; s += NT) win[s] = 0ull;
#pragma unroll
        for (int r = 0; r < R; r++)
            if (tid == r) dflag_s[r] = SYM ? (nA[r] > 0) : 0;
        if (tid == 0) grp_s = 0;
        bool dtouch[R];  // this lane added into column i0 + r (!SYM)
#pragma unroll
        for (int r = 0; r < R; r++) dtouch[r] = false;
        __syncthreads();
        // products: wavefront w takes groups g = w, w + W, ... of 64 union
        // entries; lane q of a group holds entry q's column j, coefficients
        // and the bounds of row j's segment in window c.  Only the window
        // holding the diagonal (DIAG) sends SYM's columns below row i0 + r
        // to the lane's dummy slot and notes (!SYM) the diagonal's touch.
        auto products = [&](auto diag) {
            constexpr bool DIAG = decltype(diag)::value;
            // an LDS base with the window's first column at 0 (slot of
            // column k in row r: r * WC + k - cbase)
            unsigned long long *wc = win - (size_t) NC * cbase;
            // groups of 64 union entries are taken from a block counter (the
            // groups' work varies with their segments' lengths)
            auto next_group = [&]() {
                int g = 0;
                if (lane == 0) g = 64 * atomicAdd(&grp_s, 1);
                return __builtin_amdgcn_readfirstlane(g);
            };
            for (int g0 = next_group(); g0 < nU; g0 = next_group()) {
                const int e = g0 + lane;
                int jq = 0, lo = 0, hi = 0;
                unsigned rq = 0;  // rows with a nonzero coefficient
                V cs[R];          // the coefficients, widened and scaled by 2^E
#pragma unroll
                for (int r = 0; r < R; r++) cs[r] = fx_scale(F::widen(E::zero()), 0.0);
                if (e < nU) {
                    jq = Uj[e];
                    const int *bo = p.boff + (size_t) jq * nw1;
                    lo = bo[c];
                    hi = bo[c + 1];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const T t = Uc[(size_t) e * R + r];
                        cs[r] = fx_scale(F::widen(t), sc);
                        if (E::nonzero(t)) rq |= 1u << r;
                    }
                }
                uint64_t todo = ballot64(hi > lo);
                if (!todo) continue;
                // The group's segments as a sequence of rounds of 256
                // entries (four per lane), software-pipelined: the next
                // round's loads are issued before the current round's adds.
                // Segment state is wave-uniform (scalar registers).
                struct Seg {
                    const unsigned *Pj;  // row j's packed entries
                    int l1;              // segment end
                    unsigned rm;
                    V a[R];
                };
                auto fetch = [&](Seg &sg, int &kb) {
                    const int q = __ffsll((long long) todo) - 1;
                    todo &= todo - 1ull;
                    const int j = __builtin_amdgcn_readlane(jq, q);
                    kb = __builtin_amdgcn_readlane(lo, q);
                    sg.l1 = __builtin_amdgcn_readlane(hi, q);
                    sg.rm = (unsigned) __builtin_amdgcn_readlane((int) rq, q);
                    sg.Pj = P + (size_t) j * p.MP * PkW<T>::n;
#pragma unroll
                    for (int r = 0; r < R; r++) sg.a[r] = bcast<64>(cs[r], q);
                    const int len = sg.l1 - kb;
                    streamed += len;
#pragma unroll
                    for (int r = 0; r < R; r++)
                        if (sg.rm >> r & 1u) prods[r] += len;
                };
                // EPL loads per lane, always issued (straight-line code, so
                // that the adds wait only for their own round's loads):
                // lanes past the segment's end reload its last entry and
                // mark it column -1.
                constexpr int EPL = 4;  // 256 entries per round
                auto load = [&](const Seg &sg, int kb, T (&bv)[EPL], int (&cv)[EPL]) {
                    const int k0 = kb + lane, kl = sg.l1 - 1;
#pragma unroll
                    for (int u = 0; u < EPL; u++) {
                        const int k = k0 + 64 * u;
                        pk_get_at<T>(sg.Pj, (unsigned) (k < kl ? k : kl), bv[u], cv[u]);
                    }
#pragma unroll
                    for (int u = 0; u < EPL; u++)
                        if (k0 + 64 * u > kl) cv[u] = -1;
                };
                auto add = [&](const Seg &sg, T bv, int col) {
                    if (col < 0) return;
                    const V wv = F::widen(bv);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        if (!(sg.rm >> r & 1u)) continue;  // uniform
                        int sl = r * WC + col;
                        if constexpr (DIAG && SYM) sl = col >= i0 + r ? sl : dummy + cbase;
                        fx_add(wc, sl, sg.a[r], wv);
                        if constexpr (DIAG && !SYM) dtouch[r] = dtouch[r] | (col == i0 + r);
                    }
                };
                Seg cur, nxt;
                int kc, kn;
                T bv[EPL], nv[EPL];
                int cv[EPL], dv[EPL];
                fetch(cur, kc);
                load(cur, kc, bv, cv);
                while (true) {
                    // the round after (kc): the rest of this segment, or the next
                    const bool same = kc + 64 * EPL < cur.l1;
                    const bool more = same || todo != 0ull;
                    if (same) {
                        nxt = cur;
                        kn = kc + 64 * EPL;
                    } else if (more) {
                        fetch(nxt, kn);
                    } else {
                        nxt = cur;  // a round that is loaded and never added
                        kn = kc;
                    }
                    load(nxt, kn, nv, dv);
#pragma unroll
                    for (int u = 0; u < EPL; u++) add(cur, bv[u], cv[u]);
                    if (!more) break;
                    cur = nxt;
                    kc = kn;
#pragma unroll
                    for (int u = 0; u < EPL; u++) {
                        bv[u] = nv[u];
                        cv[u] = dv[u];
                    }
                }
            }
        };
        if (c == dwin)
            products(std::true_type{});
        else
            products(std::false_type{});
        if constexpr (!SYM) {
#pragma unroll
            for (int r = 0; r < R; r++)
                if (__any(dtouch[r]) && lane == 0) dflag_s[r] = 1;
        }
        __syncthreads();
        // epilogue, one row at a time: wavefront w walks columns
        // [w PER, (w+1) PER) of row r's window twice, counting, then
        // writing at its offset
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!live[r]) continue;  // uniform
            const int i = i0 + r;
            const size_t rowC = (size_t) i * MC;
            const bool dtouched = dflag_s[r] != 0;
            if constexpr (FUSE) {
                // X's row i in this wavefront's columns of the window
                for (int s = lane; s < PER; s += 64) xbuf[s] = E::zero();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int *bo = p.boff + (size_t) i * nw1;
                const int kw = cbase + w * PER;
                for (int e = bo[c] + lane; e < bo[c + 1]; e += 64) {
                    T v;
                    int k;
                    pk_get<T>(P, (size_t) i * p.MP + e, v, k);
                    if (k >= kw && k < kw + PER) xbuf[k - kw] = v;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            auto entry = [&](int s, T &out, bool &keep, bool &nz, double &dn) {
                const int k = cbase + s;
                const bool valid = k < p.N && (!SYM || k >= i);
                const T y = F::out(win, r * WC + s, inv);
                nz = valid && (E::mag(y) != 0.0 || k == i);
                const bool ykeep = valid && (k == i ? dtouched : E::above(y, p.thr));
                dn = 0.0;
                if constexpr (FUSE) {
                    const T x = xbuf[s - w * PER];
                    const T yk = ykeep ? y : E::zero();
                    if (valid) dn = E::abs2(E::sub(x, yk)) * ((SYM && k > i) ? 2.0 : 1.0);
                    out = E::axpy(E::axpy(E::zero(), alpha, x), beta, yk);
                    keep = valid && E::above(out, p.thr);
                } else {
                    out = y;
                    keep = ykeep;
                }
            };
            int cw = 0, dw = 0;
            double nw = 0.0;
            for (int s = w * PER + lane; s < (w + 1) * PER; s += 64) {
                T v;
                bool keep, nz;
                double dn;
                entry(s, v, keep, nz, dn);
                nw += dn;
                cw += popc64(ballot64(keep));
                dw += popc64(ballot64(nz));
                if (!FUSE && cbase + s == i) d2_s[r] = v;
            }
            if constexpr (FUSE) {
                for (int off = 32; off > 0; off >>= 1) nw += __shfl_down(nw, off, 64);
            }
            if (lane == 0) {
                cnt_s[w] = cw;
                dcnt_s[w] = dw;
                if constexpr (FUSE) nrm_s[w][r] += nw;
            }
            __syncthreads();
            int pos0 = kept[r], tot = 0, dtot = 0;
            for (int q = 0; q < W; q++) {
                pos0 += q < w ? cnt_s[q] : 0;
                tot += cnt_s[q];
                dtot += dcnt_s[q];
            }
            for (int s = w * PER + lane; s < (w + 1) * PER; s += 64) {
                T v;
                bool keep, nz;
                double dn;
                entry(s, v, keep, nz, dn);
                const uint64_t km = ballot64(keep);
                const int pos = pos0 + popc64(km & ltmask);
                if (keep && pos < MC) {
                    p.Ci[rowC + pos] = cbase + s;
                    Cv[rowC + pos] = v;
                }
                pos0 += popc64(km);
            }
            kept[r] += tot;
            distinct[r] += dtot;
            __syncthreads();  // cnt_s is reused by the next row
        }
    }
    // row records, traces, norms
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < R; r++) prods_s[w][r] = prods[r];
        prods_s[w][R] = streamed;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = i0 + r;
        if (tid == r && i < p.N) {
            if (!live[r]) {
                p.rowstat[i] = make_int4(-1, 0, 0, 0);
            } else {
                long long pr = 0;
                double nr = 0.0;
                for (int q = 0; q < W; q++) {
                    pr += prods_s[q][r];
                    if constexpr (FUSE) nr += nrm_s[q][r];
                }
                p.Cn[i] = kept[r] < MC ? kept[r] : MC;
                p.rowstat[i] = make_int4(kept[r], distinct[r], (int) (pr < INT_MAX ? pr : INT_MAX),
                                         nA[r]);
                if constexpr (FUSE) {
                    p.rownorm[i] = nr;
                } else {
                    p.rowtr[i] = E::re(trx_s[r]);
                    p.rowtr[p.N + i] = E::re(d2_s[r]);
                }
            }
        }
    }
    if (tid == 0 && p.streamed) {
        long long st = 0;
        for (int q = 0; q < W; q++) st += prods_s[q][R];
        atomicAdd(p.streamed, (unsigned long long) st);
    }
}

// First pass: table in LDS, 64/GROUP rows per 64-thread block.
template <typename T, int GROUP, int BK>
__global__ __launch_bounds__(64) void k_spgemm_lds(SpParams p, unsigned tsize) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (p.skip && *p.skip) return;
    constexpr int R = 64 / GROUP;
    const unsigned stride = tsize + table_pad<GROUP>();
    T *vals = (T *) smem;
    int *keys = (int *) (smem + (size_t) R * stride * sizeof(T));
    uint16_t *ord = (uint16_t *) (smem + (size_t) R * stride * (sizeof(T) + sizeof(int)));
    for (unsigned s = threadIdx.x; s < R * stride; s += 64) {
        keys[s] = EMPTY;
        vals[s] = Elem<T>::zero();
    }
    __syncthreads();
    const int g = threadIdx.x / GROUP;
    const int cap = (int) tsize - 8;
    if (p.qlist) {
        // persistent over the work queue, R listed rows per take (each
        // row's epilogue leaves its table clean)
        // (the next take is issued while the current rows run)
        const int n = p.qinfo[0];
        int q = 0;
        if (threadIdx.x == 0) q = atomicAdd(&p.qinfo[1], 1);
        q = __shfl(q, 0, 64);
        while ((long long) q * R < n) {
            int nxt = 0;
            if (threadIdx.x == 0) nxt = atomicAdd(&p.qinfo[1], 1);
            const int v = q * R + g;
            spgemm_row<T, GROUP, BK, false, uint16_t>(p, v < n ? p.qlist[v] : -1, keys + g * stride,
                                                      vals + (size_t) g * stride, ord + g * stride,
                                                      tsize, cap);
            __syncthreads();
            q = __shfl(nxt, 0, 64);
        }
        return;
    }
    // with p.remap each XCD takes a contiguous range of rows
    const int vb = p.remap ? xcd_block(blockIdx.x, gridDim.x) : (int) blockIdx.x;
    const int row = vb * R + g;
    const int i = (row < p.N && !(p.row_sys && p.sys_done[p.row_sys[row]])) ? row : -1;
    if (row < p.N && i < 0 && threadIdx.x % GROUP == 0) p.rowstat[row] = make_int4(-1, 0, 0, 0);
    spgemm_row<T, GROUP, BK, false, uint16_t>(p, i, keys + g * stride, vals + (size_t) g * stride,
                                              ord + g * stride, tsize, cap);
}

// Second pass: rows listed in p.ovf, one wavefront per row, table in global
// memory (tsize slots per block, at least twice the row's bound).
template <typename T>
__global__ __launch_bounds__(64) void k_spgemm_global(SpParams p, unsigned tsize,
                                                      unsigned char *tabs, int *cursor) {
    if (p.skip && *p.skip) return;
    const int n = p.st->table_overflow_rows;
    if (n == 0) return;
    const size_t stride = (size_t) tsize + table_pad<64>();
    unsigned char *base = tabs + blockIdx.x * stride * (sizeof(T) + 2 * sizeof(int));
    T *vals = (T *) base;
    int *keys = (int *) (base + stride * sizeof(T));
    int *ord = (int *) (base + stride * (sizeof(T) + sizeof(int)));
    for (size_t s = threadIdx.x; s < stride; s += 64) {
        keys[s] = EMPTY;
        vals[s] = Elem<T>::zero();
    }
    __syncthreads();
    while (true) {
        int q = 0;
        if (threadIdx.x == 0) q = atomicAdd(cursor, 1);
        q = __shfl(q, 0, 64);
        if (q >= n) break;
        // the table holds twice the row's bound, so a step never fills it
        spgemm_row<T, 64, 2, true, int>(p, p.ovf[q], keys, vals, ord, tsize, (int) tsize - 2);
        __syncthreads();
    }
}

// Reduces per-row records (kept, distinct, products, nnz of A's row), kept
// = -1 marking a skipped row, into the status block.  `what` selects:
// BMLHIP_RS_OVF_KEPT / BMLHIP_RS_OVF_DIST count the rows whose kept /
// distinct count exceeds MC (and the first of them), BMLHIP_RS_STORED the
// largest and total stored row (kept capped at MC), BMLHIP_RS_PROF the
// profiling sums of precision `prof`.  One set of atomics per block.
__global__ __launch_bounds__(256) void k_reduce_rowstat(const int4 *rs, int N, int MC, int what,
                                                        int prof, const int *skip,
                                                        bmlhip_status *st) {
    if (skip && *skip) return;
    int ovf = 0, first = INT_MAX, mx = 0, mt = 0;
    unsigned long long tot = 0, prods = 0, rows = 0, innz = 0, dist = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const int4 r = rs[i];
        if (r.x < 0) continue;
        const int stored = r.x < MC ? r.x : MC;
        if (((what & BMLHIP_RS_OVF_KEPT) && r.x > MC) || ((what & BMLHIP_RS_OVF_DIST) && r.y > MC)) {
            ovf++;
            first = i < first ? i : first;
        }
        mx = stored > mx ? stored : mx;
        mt = r.y > mt ? r.y : mt;
        tot += (unsigned long long) stored;
        prods += (unsigned long long) (unsigned) r.z;
        rows++;
        innz += (unsigned long long) r.w;
        dist += (unsigned long long) r.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ovf += __shfl_down(ovf, off, 64);
        const int f = __shfl_down(first, off, 64);
        first = f < first ? f : first;
        const int a = __shfl_down(mx, off, 64);
        mx = a > mx ? a : mx;
        const int b = __shfl_down(mt, off, 64);
        mt = b > mt ? b : mt;
        tot += __shfl_down(tot, off, 64);
        prods += __shfl_down(prods, off, 64);
        rows += __shfl_down(rows, off, 64);
        innz += __shfl_down(innz, off, 64);
        dist += __shfl_down(dist, off, 64);
    }
    __shared__ int si[4][4];
    __shared__ unsigned long long sl[4][5];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        si[w][0] = ovf; si[w][1] = first; si[w][2] = mx; si[w][3] = mt;
        sl[w][0] = tot; sl[w][1] = prods; sl[w][2] = rows; sl[w][3] = innz; sl[w][4] = dist;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            ovf += si[k][0];
            first = si[k][1] < first ? si[k][1] : first;
            mx = si[k][2] > mx ? si[k][2] : mx;
            mt = si[k][3] > mt ? si[k][3] : mt;
            tot += sl[k][0]; prods += sl[k][1]; rows += sl[k][2]; innz += sl[k][3]; dist += sl[k][4];
        }
        if (ovf) {
            atomicAdd(&st->m_overflow_rows, ovf);
            atomicMin(&st->first_m_overflow, first);
        }
        if ((what & BMLHIP_RS_STORED) && rows) {
            atomicMax(&st->max_nnz, mx);
            atomicAdd((unsigned long long *) &st->total_nnz, tot);
        }
        if ((what & BMLHIP_RS_PROF) && rows) {
            atomicMax(&st->max_touched, mt);
            atomicAdd((unsigned long long *) &st->prof_products[prof], prods);
            atomicAdd((unsigned long long *) &st->prof_rows[prof], rows);
            atomicAdd((unsigned long long *) &st->prof_in_nnz[prof], innz);
            atomicAdd((unsigned long long *) &st->prof_out_nnz[prof], tot);
            atomicAdd((unsigned long long *) &st->prof_touched[prof], dist);
            if (!(what & BMLHIP_RS_OWN_STREAM))
                atomicAdd((unsigned long long *) &st->prof_streamed[prof], prods);
        }
    }
}

inline int rowstat_blocks(int N) {
    const int b = (N + 2047) / 2048;
    return b < 1 ? 1 : (b > 240 ? 240 : b);
}

// Deterministic sum of n doubles (fixed partition, fixed tree).
__global__ __launch_bounds__(1024) void k_sum(const double *x, int n, double *out) {
    __shared__ double part[1024];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 1024) s += x[k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// min of x[0..n) -> out[0], max of y[0..n) -> out[1]
__global__ __launch_bounds__(1024) void k_minmax(const double *x, const double *y, int n,
                                                 double *out, double init_max) {
    __shared__ double lo[1024], hi[1024];
    double a = 1.7976931348623157e308, b = init_max;
    for (int k = threadIdx.x; k < n; k += 1024) {
        a = x[k] < a ? x[k] : a;
        b = y[k] > b ? y[k] : b;
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) {
            lo[threadIdx.x] = lo[threadIdx.x + w] < lo[threadIdx.x] ? lo[threadIdx.x + w]
                                                                    : lo[threadIdx.x];
            hi[threadIdx.x] = hi[threadIdx.x + w] > hi[threadIdx.x] ? hi[threadIdx.x + w]
                                                                    : hi[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = lo[0];
        out[1] = hi[0];
    }
}

// Per-system sums of x over row ranges [off[s], off[s+1]) (fixed partition
// and tree: deterministic); systems with done[s] set are left alone.
// Work queue of a batched iteration.  k_batch_scan (one block): act_off[s]
// = rows of the unfinished systems before s, qinfo = (rows listed, 0, 0,
// systems left).  k_batch_fill (one block per system): an unfinished
// system's rows go to list[act_off[s] ..] in order; a finished system's
// row records are marked skipped (kept = -1), as the row kernels would
// have marked them.
__global__ __launch_bounds__(1024) void k_batch_scan(int nsys, const int *row_off, const int *done,
                                                     int *act_off, int *qinfo) {
    __shared__ int sums[1024];
    __shared__ int carry_s, left_s;
    const int t = threadIdx.x;
    if (t == 0) {
        carry_s = 0;
        left_s = 0;
    }
    __syncthreads();
    for (int base = 0; base < nsys; base += 1024) {
        const int s = base + t;
        const bool act = s < nsys && !done[s];
        const int n = act ? row_off[s + 1] - row_off[s] : 0;
        sums[t] = n;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int v = t >= off ? sums[t - off] : 0;
            __syncthreads();
            sums[t] += v;
            __syncthreads();
        }
        const int carry = carry_s;
        if (s < nsys) act_off[s] = carry + sums[t] - n;
        if (act) atomicAdd(&left_s, 1);
        __syncthreads();
        if (t == 1023) carry_s = carry + sums[1023];
        __syncthreads();
    }
    if (t == 0) {
        act_off[nsys] = carry_s;
        qinfo[0] = carry_s;
        qinfo[1] = 0;
        qinfo[2] = 0;
        qinfo[3] = left_s;
    }
}

__global__ __launch_bounds__(256) void k_batch_fill(const int *row_off, const int *done,
                                                    const int *act_off, int *list, int4 *rowstat) {
    const int s = blockIdx.x;
    const int r0 = row_off[s], n = row_off[s + 1] - r0;
    if (done[s]) {
        for (int k = threadIdx.x; k < n; k += 256) rowstat[r0 + k] = make_int4(-1, 0, 0, 0);
    } else {
        const int o = act_off[s];
        for (int k = threadIdx.x; k < n; k += 256) list[o + k] = r0 + k;
    }
}

__global__ __launch_bounds__(256) void k_seg_sum(const double *x, const int *off, const int *done,
                                                 double *out) {
    __shared__ double part[256];
    const int s = blockIdx.x;
    if (done && done[s]) return;
    double acc = 0.0;
    for (int k = off[s] + (int) threadIdx.x; k < off[s + 1]; k += 256) acc += x[k];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = part[0];
}

// Per-system min of lo and max of hi (the max starting at DBL_MIN, as bml's
// Gershgorin).
__global__ __launch_bounds__(256) void k_seg_minmax(const double *lo, const double *hi,
                                                    const int *off, double *omin, double *omax) {
    __shared__ double a_s[256], b_s[256];
    const int s = blockIdx.x;
    double a = 1.7976931348623157e308, b = 2.2250738585072014e-308;
    for (int k = off[s] + (int) threadIdx.x; k < off[s + 1]; k += 256) {
        a = lo[k] < a ? lo[k] : a;
        b = hi[k] > b ? hi[k] : b;
    }
    a_s[threadIdx.x] = a;
    b_s[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) {
            a_s[threadIdx.x] = a_s[threadIdx.x + w] < a_s[threadIdx.x] ? a_s[threadIdx.x + w] : a_s[threadIdx.x];
            b_s[threadIdx.x] = b_s[threadIdx.x + w] > b_s[threadIdx.x] ? b_s[threadIdx.x + w] : b_s[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        omin[s] = a_s[0];
        omax[s] = b_s[0];
    }
}

// ---------------------------------------------------------------------------
// add / add_norm: A <- alpha*A + beta*B per row, union in first-touch order
// (A's columns, then B's new ones), every entry kept iff |v| > threshold.
// mode 0: add (A skipped if alpha == 0, B if beta == 0); mode 1: add_norm
// with Re((A-B)^2); mode 2: add_norm with |A-B|^2.  One wavefront per row;
// y (A - B) lives in a second value table.

struct AddParams {
    int N, MA, MB, mode;
    void *Av;
    int *Ai, *An;
    const void *Bv;
    const int *Bi, *Bn;
    double alpha, beta, thr;
    double *rownorm;
    bmlhip_status *st;
    const double *ab;  // when set: alpha = ab[0], beta = ab[1] (read on the device)
    const int *skip;   // when set and nonzero, the launch does nothing
    // batched systems: alpha, beta of row i at ab_sys[2*row_sys[i]]; rows of
    // systems with sys_done set are skipped
    const int *row_sys, *sys_done;
    const double *ab_sys;
    int4 *rowstat;  // [N]: (kept, 0, 0, 0), kept = -1 for a skipped row
    // batched work queue (NULL: one block per row): rows qlist[0 ..
    // qinfo[0]), taken with the cursor qinfo[2]
    const int *qlist;
    int *qinfo;
};

// tsize: any table size; a column's home slot is its multiplicative hash
// scaled to the table, probing is linear with wrap-around
template <typename T, typename ORD>
__device__ void add_row(const AddParams &p, int i, int *keys, T *xs, T *ys, ORD *ord,
                        unsigned tsize) {
    using E = Elem<T>;
    const int lane = lane_id();
    const uint64_t ltmask = (1ull << lane) - 1ull;
    T *Av = (T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    const size_t rA = (size_t) i * p.MA, rB = (size_t) i * p.MB;
    int count = 0;
    const double *ab = p.ab_sys ? p.ab_sys + 2 * p.row_sys[i] : p.ab;
    const double alpha = ab ? ab[0] : p.alpha;
    const double beta = ab ? ab[1] : p.beta;
    for (int part = 0; part < 2; part++) {
        const double s = part ? beta : alpha;
        if (p.mode == 0 && !(s > 0.0 || s < 0.0)) continue;
        const int n = part ? p.Bn[i] : p.An[i];
        for (int k0 = 0; k0 < n; k0 += 64) {
            const int kp = k0 + lane;
            const bool act = kp < n;
            int col = 0, slot = 0;
            bool isnew = false;
            T v = E::zero();
            if (act) {
                col = part ? p.Bi[rB + kp] : p.Ai[rA + kp];
                v = part ? Bv[rB + kp] : Av[rA + kp];
                unsigned h = __umulhi((unsigned) col * 0x9E3779B1u, tsize);
                while (true) {
                    const int kk = keys[h];
                    if (kk == col) break;
                    if (kk == EMPTY) {
                        const int old = atomicCAS(&keys[h], EMPTY, col);
                        if (old == EMPTY) {
                            isnew = true;
                            break;
                        }
                        if (old == col) break;
                    }
                    h = h + 1u == tsize ? 0u : h + 1u;
                }
                slot = (int) h;
            }
            const uint64_t nm = ballot64(isnew);
            if (isnew) ord[count + popc64(nm & ltmask)] = (ORD) slot;
            count += popc64(nm);
            if (act) {
                xs[slot] = E::axpy(xs[slot], s, v);
                if (p.mode) ys[slot] = part ? E::sub(ys[slot], v) : E::add(ys[slot], v);
            }
        }
    }
    int kept = 0;
    double nrm = 0.0;
    for (int r0 = 0; r0 < count; r0 += 64) {
        const int r = r0 + lane;
        const bool act = r < count;
        int s = 0, col = 0;
        T x = E::zero();
        if (act) {
            s = ord[r];
            col = keys[s];
            x = xs[s];
            if (p.mode == 1) nrm += E::re_sq(ys[s]);
            if (p.mode == 2) nrm += E::abs2(ys[s]);
        }
        const bool keep = act && E::above(x, p.thr);
        const uint64_t km = ballot64(keep);
        const int pos = kept + popc64(km & ltmask);
        if (keep && pos < p.MA) {
            p.Ai[rA + pos] = col;
            Av[rA + pos] = x;
        }
        if (act) {
            keys[s] = EMPTY;
            xs[s] = E::zero();
            ys[s] = E::zero();
        }
        kept += popc64(km);
    }
    // fixed-order wavefront reduction of the row's norm
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, 64);
    if (lane == 0) {
        p.An[i] = kept < p.MA ? kept : p.MA;
        p.rowstat[i] = make_int4(kept, 0, 0, 0);
        if (p.mode) p.rownorm[i] = nrm;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_add_lds(AddParams p, unsigned tsize) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (p.skip && *p.skip) return;
    T *xs = (T *) smem;
    T *ys = xs + tsize;
    int *keys = (int *) (ys + tsize);
    uint16_t *ord = (uint16_t *) (keys + tsize);
    for (unsigned s = threadIdx.x; s < tsize; s += 64) {
        keys[s] = EMPTY;
        xs[s] = Elem<T>::zero();
        ys[s] = Elem<T>::zero();
    }
    __syncthreads();
    if (p.qlist) {
        // persistent over the work queue (each row leaves the table clean;
        // the next take is issued while the current row runs)
        const int n = p.qinfo[0];
        int q = 0;
        if (threadIdx.x == 0) q = atomicAdd(&p.qinfo[2], 1);
        q = __shfl(q, 0, 64);
        while (q < n) {
            int nxt = 0;
            if (threadIdx.x == 0) nxt = atomicAdd(&p.qinfo[2], 1);
            add_row<T, uint16_t>(p, p.qlist[q], keys, xs, ys, ord, tsize);
            __syncthreads();
            q = __shfl(nxt, 0, 64);
        }
        return;
    }
    const int i = blockIdx.x;
    if (p.row_sys && p.sys_done[p.row_sys[i]]) {
        if (threadIdx.x == 0) p.rowstat[i] = make_int4(-1, 0, 0, 0);
        return;
    }
    add_row<T, uint16_t>(p, i, keys, xs, ys, ord, tsize);
}

template <typename T>
__global__ __launch_bounds__(64) void k_add_global(AddParams p, int tbits, unsigned char *tabs) {
    if (p.skip && *p.skip) return;
    const size_t tsize = (size_t) 1 << tbits;
    unsigned char *base = tabs + blockIdx.x * tsize * (2 * sizeof(T) + 2 * sizeof(int));
    T *xs = (T *) base;
    T *ys = xs + tsize;
    int *keys = (int *) (ys + tsize);
    int *ord = keys + tsize;
    for (size_t s = threadIdx.x; s < tsize; s += 64) {
        keys[s] = EMPTY;
        xs[s] = Elem<T>::zero();
        ys[s] = Elem<T>::zero();
    }
    __syncthreads();
    for (int i = blockIdx.x; i < p.N; i += gridDim.x) {
        if (p.row_sys && p.sys_done[p.row_sys[i]]) {
            if (threadIdx.x == 0) p.rowstat[i] = make_int4(-1, 0, 0, 0);
            continue;
        }
        add_row<T, int>(p, i, keys, xs, ys, ord, (unsigned) tsize);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// threshold: in-place stable compaction, one wavefront per row.
template <typename T>
__global__ __launch_bounds__(256) void k_threshold(int N, int M, T *Av, int *Ai, int *An,
                                                   double thr) {
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    if (i >= N) return;
    const int lane = threadIdx.x & 63;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const size_t r = (size_t) i * M;
    const int n = An[i];
    int kept = 0;
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int kp = k0 + lane;
        const bool act = kp < n;
        T v = Elem<T>::zero();
        int c = 0;
        if (act) {
            v = Av[r + kp];
            c = Ai[r + kp];
        }
        const bool keep = act && Elem<T>::above(v, thr);
        const uint64_t km = ballot64(keep);
        const int pos = kept + popc64(km & ltmask);
        if (keep) {
            Av[r + pos] = v;
            Ai[r + pos] = c;
        }
        kept += popc64(km);
    }
    if (lane == 0) An[i] = kept;
}

// trace: first diagonal entry of each row -> rowv[i]
template <typename T>
__global__ __launch_bounds__(256) void k_trace_rows(int N, int M, const T *Av, const int *Ai,
                                                    const int *An, double *rowv) {
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    if (i >= N) return;
    const int lane = threadIdx.x & 63;
    const size_t r = (size_t) i * M;
    const int n = An[i];
    double v = 0.0;
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int kp = k0 + lane;
        const uint64_t dm = ballot64(kp < n && Ai[r + kp] == i);
        if (dm) {
            const int first = k0 + __ffsll((unsigned long long) dm) - 1;
            v = Elem<T>::re(Av[r + first]);
            break;
        }
    }
    if (lane == 0) rowv[i] = v;
}

// gershgorin: per row, d = last diagonal entry, radius = sum |offdiag| held
// in T (sequential, the reference's order); rowlo = Re(d - r), rowhi = Re(d + r)
template <typename T>
__global__ __launch_bounds__(256) void k_gersh_rows(int N, int M, const T *Av, const int *Ai,
                                                    const int *An, double *rowlo,
                                                    double *rowhi) {
    using E = Elem<T>;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const size_t r = (size_t) i * M;
    T d = E::zero(), rad = E::zero();
    for (int k = 0; k < An[i]; k++) {
        const T v = Av[r + k];
        if (Ai[r + k] == i)
            d = v;
        else
            rad = E::acc_mag(rad, E::mag(v));
    }
    rowlo[i] = E::re_minus(d, rad);
    rowhi[i] = E::re_plus(d, rad);
}

// scale (optional) then add_identity with threshold, row i, sequential like
// the reference (bml_add_ellpack_typed.c:420-549)
template <typename T>
__device__ void scale_add_identity_row(int i, int M, T *Av, int *Ai, int *An, T s, int do_scale,
                                       double beta, double thr, int do_identity,
                                       bmlhip_status *st) {
    using E = Elem<T>;
    const size_t r = (size_t) i * M;
    const int n = An[i];
    if (do_scale)
        for (int k = 0; k < n; k++) Av[r + k] = E::mul(s, Av[r + k]);
    if (!do_identity) return;
    int diag = -1;
    for (int k = 0; k < n; k++)
        if (Ai[r + k] == i) diag = k;
    bool extra = false;
    T ev = E::zero();
    if (beta > 0.0 || beta < 0.0) {
        if (diag < 0) {
            extra = true;
            ev = E::from_real(beta);
        } else {
            Av[r + diag] = E::add_real(Av[r + diag], beta);
        }
    }
    int kept = 0;
    for (int k = 0; k < n; k++) {
        const T v = Av[r + k];
        if (E::above(v, thr)) {
            Av[r + kept] = v;
            Ai[r + kept] = Ai[r + k];
            kept++;
        }
    }
    if (extra && E::above(ev, thr)) {
        if (kept < M) {
            Av[r + kept] = ev;
            Ai[r + kept] = i;
            kept++;
        } else {
            atomicAdd(&st->m_overflow_rows, 1);
            atomicMin(&st->first_m_overflow, i);
        }
    }
    An[i] = kept;
}

// one thread per row
template <typename T>
__global__ __launch_bounds__(256) void k_scale_add_identity(int N, int M, T *Av, int *Ai,
                                                            int *An, T s, int do_scale,
                                                            double beta, double thr,
                                                            int do_identity,
                                                            bmlhip_status *st) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    scale_add_identity_row<T>(i, M, Av, Ai, An, s, do_scale, beta, thr, do_identity, st);
}

// batched normalize: row i of system s = row_sys[i] gets
// X0 = (emax I - X)/(emax - emin) with that system's bounds, as
// bml_normalize_ellpack_typed.c:33-45 (scale by (T)(-1/(max-min)), then
// add_identity(max/(max-min)) at threshold 0)
template <typename T>
__global__ __launch_bounds__(256) void k_normalize_batch(int N, int M, T *Av, int *Ai, int *An,
                                                         const int *row_sys, const double *emin,
                                                         const double *emax, bmlhip_status *st) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int sy = row_sys[i];
    const double d = emax[sy] - emin[sy];
    scale_add_identity_row<T>(i, M, Av, Ai, An, Elem<T>::from_real(-1.0 / d), 1, emax[sy] / d,
                              0.0, 1, st);
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_rows(int N, int MA, const T *Av, const int *Ai,
                                                   const int *An, int MB, T *Bv, int *Bi,
                                                   int *Bn, bmlhip_status *st, int shift) {
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    if (i >= N) return;
    const int lane = threadIdx.x & 63;
    const int n = An[i];
    const int m = n < MB ? n : MB;
    for (int k = lane; k < m; k += 64) {
        Bv[(size_t) i * MB + k] = Av[(size_t) i * MA + k];
        Bi[(size_t) i * MB + k] = Ai[(size_t) i * MA + k] + shift;
    }
    if (lane == 0) {
        Bn[i] = m;
        if (n > MB) {
            atomicAdd(&st->m_overflow_rows, 1);
            atomicMin(&st->first_m_overflow, i);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_identity(int N, int M, T *Av, int *Ai, int *An) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    Av[(size_t) i * M] = Elem<T>::from_real(1.0);
    Ai[(size_t) i * M] = i;
    An[i] = 1;
}

__global__ __launch_bounds__(256) void k_row_stats(int N, const int *An, bmlhip_status *st) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    int n = i < N ? An[i] : 0;
    // wavefront reductions, then one atomic per wavefront
    int mx = n;
    unsigned long long sm = (unsigned long long) n;
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_down(mx, off, 64);
        mx = o > mx ? o : mx;
        sm += __shfl_down(sm, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&st->max_nnz, mx);
        atomicAdd((unsigned long long *) &st->total_nnz, sm);
    }
}

inline int ceil_log2(long long v) {
    int b = 0;
    while ((1ll << b) < v) b++;
    return b;
}

int fail_msg(int code, const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    bmlhip_set_errmsg(buf);
    return code;
}

int report(const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    return fail_msg(3, "%s: %s", what, hipGetErrorString(e));
}

// Grid of a persistent kernel: as many blocks as the GPU holds at once
// (blocks per CU for this kernel, block size and LDS, times the CUs), at
// most `items`.
int persistent_blocks(const void *fn, int threads, size_t lds, long long items) {
    static int cus[64];
    int dev = 0;
    (void) hipGetDevice(&dev);
    const int d = dev >= 0 && dev < 64 ? dev : 0;
    if (!cus[d]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cus[d] = n;
    }
    // the occupancy of a (kernel, block, LDS) triple is looked up once, so a
    // graph capture (whose eager pass came first) makes no query
    struct Occ {
        const void *fn;
        int threads;
        size_t lds;
        int per;
    };
    static Occ occ[64];
    static int nocc = 0;
    int per = 0;
    for (int k = 0; k < nocc && !per; k++)
        if (occ[k].fn == fn && occ[k].threads == threads && occ[k].lds == lds) per = occ[k].per;
    if (!per) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) != hipSuccess ||
            per < 1)
            per = 1;
        if (nocc < 64) occ[nocc++] = {fn, threads, lds, per};
    }
    long long b = (long long) per * cus[d];
    if (b > items) b = items;
    return b < 1 ? 1 : (int) b;
}

template <typename T, int GROUP, int BK>
int launch_spgemm_lds(const SpParams &p, unsigned tsize, hipStream_t s) {
    constexpr int R = 64 / GROUP;
    const size_t lds = (size_t) R * (tsize + table_pad<GROUP>()) * (sizeof(T) + sizeof(int) + 2);
    if (lds > 160 * 1024) {
        return fail_msg(2, "LDS table of %zu bytes exceeds 160 KiB", lds);
    }
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_spgemm_lds<T, GROUP, BK>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    const int blocks = !p.qlist || p.N == 0 ? (p.N + R - 1) / R
                       : persistent_blocks((const void *) k_spgemm_lds<T, GROUP, BK>, 64, lds,
                                           (p.N + R - 1) / R);
    if (blocks > 0) {
        bmlhip_prof_mark_begin(PrecOf<T>::value);
        hipLaunchKernelGGL((k_spgemm_lds<T, GROUP, BK>), dim3(blocks), dim3(64), lds, s, p, tsize);
        bmlhip_prof_mark_end();
    }
    return report("k_spgemm_lds");
}

template <typename T, int W>
int launch_spgemm_mw(const SpParams &p, unsigned tsize, hipStream_t s) {
    const size_t lds = (size_t) (tsize + table_pad<64>()) * (sizeof(T) + sizeof(int) + 2);
    if (lds > 160 * 1024 - 64) {
        return fail_msg(2, "LDS table of %zu bytes exceeds 160 KiB", lds);
    }
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_spgemm_mw<T, W, 4>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    if (p.N > 0) {
        bmlhip_prof_mark_begin(PrecOf<T>::value);
        hipLaunchKernelGGL((k_spgemm_mw<T, W, 4>), dim3(p.N), dim3(64 * W), lds, s, p, tsize);
        bmlhip_prof_mark_end();
    }
    return report("k_spgemm_mw");
}

// (the symmetric x2's marks are set by the caller around the row kernel
// and the mirror together)
template <typename T, int W, int BK, bool SYM>
int launch_spgemm_unord(const SpParams &p, unsigned tsize, hipStream_t s) {
    const size_t lds = (size_t) (tsize + table_pad<64>()) *
                       (sizeof(typename UAcc<T>::type) + sizeof(int));
    if (lds > 160 * 1024 - 256) {
        return fail_msg(2, "LDS table of %zu bytes exceeds 160 KiB", lds);
    }
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_spgemm_unord<T, W, BK, SYM>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    if (p.N > 0) {
        const int blocks = !p.qlist ? p.N
                           : persistent_blocks((const void *) k_spgemm_unord<T, W, BK, SYM>,
                                               64 * W, lds, p.N);
        if (!SYM) bmlhip_prof_mark_begin(PrecOf<T>::value);
        hipLaunchKernelGGL((k_spgemm_unord<T, W, BK, SYM>), dim3(blocks), dim3(64 * W), lds, s, p,
                           tsize);
        if (!SYM) bmlhip_prof_mark_end();
    }
    return report("k_spgemm_unord");
}

template <typename T, int BK, bool SYM>
int launch_spgemm_unord_w(const SpParams &p, unsigned tsize, int waves, hipStream_t s) {
    switch (waves) {
    case 1: return launch_spgemm_unord<T, 1, BK, SYM>(p, tsize, s);
    case 2: return launch_spgemm_unord<T, 2, BK, SYM>(p, tsize, s);
    case 4: return launch_spgemm_unord<T, 4, BK, SYM>(p, tsize, s);
    case 8: return launch_spgemm_unord<T, 8, BK, SYM>(p, tsize, s);
    default: return fail_msg(2, "unordered engine: waves %d not one of 1/2/4/8", waves);
    }
}

template <typename T, int GROUP>
int launch_spgemm_bk(const SpParams &p, unsigned tsize, int bucket, hipStream_t s) {
    switch (bucket) {
    case 2: return launch_spgemm_lds<T, GROUP, 2>(p, tsize, s);
    case 4: return launch_spgemm_lds<T, GROUP, 4>(p, tsize, s);
    case 8: return launch_spgemm_lds<T, GROUP, 8>(p, tsize, s);
    default: return fail_msg(2, "bucket %d not one of 2/4/8", bucket);
    }
}

// ---- window engine launchers ----------------------------------------------

template <typename T, int W, int WB, bool SYM, bool FUSE>
int launch_win(const WinParams &wp, hipStream_t s) {
    using A = typename UAcc<T>::type;
    const size_t lds = sizeof(A) * (((size_t) 1 << WB) + 64) + (FUSE ? sizeof(T) * ((size_t) 1 << WB) : 0);
    if (lds > 160 * 1024 - 1024)
        return fail_msg(5, "window engine: a window of 2^%d columns needs %zu bytes of LDS", WB, lds);
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_spgemm_win<T, W, WB, SYM, FUSE>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    if (wp.N > 0)
        hipLaunchKernelGGL((k_spgemm_win<T, W, WB, SYM, FUSE>), dim3(wp.N), dim3(64 * W), lds, s,
                           wp);
    return report("k_spgemm_win");
}

template <typename T, bool SYM, bool FUSE>
int launch_win_shape(const WinParams &wp, int wbits, int waves, hipStream_t s) {
    if (waves == 8) {
        switch (wbits) {
        case 10: return launch_win<T, 8, 10, SYM, FUSE>(wp, s);
        case 11: return launch_win<T, 8, 11, SYM, FUSE>(wp, s);
        case 12: return launch_win<T, 8, 12, SYM, FUSE>(wp, s);
        case 13: return launch_win<T, 8, 13, SYM, FUSE>(wp, s);
        default: break;
        }
    } else if (waves == 4) {
        switch (wbits) {
        case 10: return launch_win<T, 4, 10, SYM, FUSE>(wp, s);
        case 11: return launch_win<T, 4, 11, SYM, FUSE>(wp, s);
        case 12: return launch_win<T, 4, 12, SYM, FUSE>(wp, s);
        case 13: return launch_win<T, 4, 13, SYM, FUSE>(wp, s);
        default: break;
        }
    }
    return fail_msg(2, "window engine: window 2^%d with %d wavefronts not built (2^10..2^13, 4 or 8)",
                    wbits, waves);
}

template <typename T, int W, int WB, int R, bool SYM, bool FUSE>
int launch_winb(const WinParams &wp, hipStream_t s) {
    const size_t WC = (size_t) 1 << WB;
    const size_t lds = sizeof(unsigned long long) * FxW<T>::NC * (R * WC + 64) +
                       (FUSE ? sizeof(T) * WC : 0);
    if (lds > 160 * 1024 - 2048)
        return fail_msg(5, "window engine: %d rows of 2^%d columns need %zu bytes of LDS", R, WB,
                        lds);
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_spgemm_winb<T, W, WB, R, SYM, FUSE>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    if (wp.N > 0)
        hipLaunchKernelGGL((k_spgemm_winb<T, W, WB, R, SYM, FUSE>), dim3((wp.N + R - 1) / R),
                           dim3(64 * W), lds, s, wp);
    return report("k_spgemm_winb");
}

// The product on blocks of `rows` output rows (2 or 4, after win_union), or
// one row per block (rows 1).
template <typename T, bool SYM, bool FUSE>
int launch_win_rows(const WinParams &wp, int rows, int wbits, int waves, hipStream_t s) {
    if (rows == 1) return launch_win_shape<T, SYM, FUSE>(wp, wbits, waves, s);
    if (waves == 8 && rows == 4) {
        switch (wbits) {
        case 10: return launch_winb<T, 8, 10, 4, SYM, FUSE>(wp, s);
        case 11: return launch_winb<T, 8, 11, 4, SYM, FUSE>(wp, s);
        default: break;
        }
    } else if (waves == 8 && rows == 2) {
        switch (wbits) {
        case 10: return launch_winb<T, 8, 10, 2, SYM, FUSE>(wp, s);
        case 11: return launch_winb<T, 8, 11, 2, SYM, FUSE>(wp, s);
        default: break;
        }
    }
    return fail_msg(2, "window engine: blocks of %d rows with windows of 2^%d columns and %d "
                       "wavefronts not built (2 or 4 rows, 2^10 or 2^11, 8 wavefronts)",
                    rows, wbits, waves);
}

// The union lists of the row blocks (k_win_union) into the window scratch.
// rows (4 or 2) is lowered until rows * M + 1 slots fit the union table's
// 2^14 (96 KB of LDS); at 1 nothing is built and the one-row kernel runs,
// as it does for the shapes the row blocks are not built for (windows
// wider than 2^11 columns, other than 8 wavefronts).
template <typename T>
int win_union(int &rows, int N, int M, int wbits, int waves, const void *Xv, const int *Xi,
              const int *Xn, const int *skip, const int *row_sys, const int *sys_done,
              WinParams &wp, hipStream_t s) {
    if (wbits > 11 || waves != 8) rows = 1;
    while (rows > 1 && (long long) rows * M + 1 > (1ll << 14)) rows >>= 1;
    if (rows != 2 && rows != 4) {
        rows = 1;
        return 0;
    }
    int tbits = 8;
    while ((1ll << tbits) < (long long) rows * M + 1) tbits++;
    const size_t lds = (size_t) 6 << tbits;  // 32-bit keys and 16-bit ranks
    const int nblk = (N + rows - 1) / rows;
    const size_t ustride = (size_t) rows * M < (size_t) N ? (size_t) rows * M : (size_t) N;
    int *Uj = (int *) bmlhip_scratch(BMLHIP_WIN_UJ_SLOT, sizeof(int) * (size_t) nblk * ustride + 256);
    T *Uc = (T *) bmlhip_scratch(BMLHIP_WIN_UC_SLOT,
                                 sizeof(T) * (size_t) nblk * ustride * rows + 256);
    int *Un = (int *) bmlhip_scratch(BMLHIP_WIN_UN_SLOT, sizeof(int) * ((size_t) nblk + 1));
    if (!Uj || !Uc || !Un) return fail_msg(4, "scratch allocation failed");
    wp.Uj = Uj;
    wp.Uc = Uc;
    wp.Un = Un;
    wp.ustride = ustride;
    if (N == 0) return 0;
    // BML_HIP_WIN_UNION=0 or 1 (read per call) picks the hash or the bitmap
    // union for every element type; by default 4-byte elements take the
    // bitmap's
    const char *ue = getenv("BML_HIP_WIN_UNION");
    const int umode = ue ? atoi(ue) : -1;
    const bool bitmap = umode < 0 ? sizeof(T) == 4 : umode > 0;
    if (N <= (1 << 16) && bitmap) {
        // bitmap union: N/8 bytes of bits and N/8 of word ranks
        const size_t lb = (size_t) 8 * ((N + 31) / 32);
        if (rows == 4)
            hipLaunchKernelGGL((k_win_union_bm<T, 4>), dim3(nblk), dim3(256), lb, s, N, M,
                               (const T *) Xv, Xi, Xn, Uj, Uc, Un, ustride, skip, row_sys,
                               sys_done);
        else
            hipLaunchKernelGGL((k_win_union_bm<T, 2>), dim3(nblk), dim3(256), lb, s, N, M,
                               (const T *) Xv, Xi, Xn, Uj, Uc, Un, ustride, skip, row_sys,
                               sys_done);
        return report("k_win_union_bm");
    }
    if (rows == 4) {
        if (lds > 64 * 1024)
            (void) hipFuncSetAttribute((const void *) k_win_union<T, 4>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        hipLaunchKernelGGL((k_win_union<T, 4>), dim3(nblk), dim3(256), lds, s, N, M,
                           (const T *) Xv, Xi, Xn, tbits, Uj, Uc, Un, ustride, skip, row_sys,
                           sys_done);
    } else {
        if (lds > 64 * 1024)
            (void) hipFuncSetAttribute((const void *) k_win_union<T, 2>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        hipLaunchKernelGGL((k_win_union<T, 2>), dim3(nblk), dim3(256), lds, s, N, M,
                           (const T *) Xv, Xi, Xn, tbits, Uj, Uc, Un, ustride, skip, row_sys,
                           sys_done);
    }
    return report("k_win_union");
}

// Window count of an N-column product, or -1 when it passes the limit.
inline int win_count(int N, int wbits) {
    const long long n = ((long long) N + (1ll << wbits) - 1) >> wbits;
    return n <= BMLHIP_WIN_MAX_WINDOWS ? (int) (n < 1 ? 1 : n) : -1;
}

// Pack X (N rows, stride M) into the window scratch; traces: the row
// partials of tr X and sum |X_ik|^2 into rowtr.  Fills wp's P, boff, wmask.
template <typename T>
int win_pack(int N, int M, const void *Xv, const int *Xi, const int *Xn, int wbits,
             double *rowtr, const int *skip, const int *row_sys, const int *sys_done,
             WinParams &wp, hipStream_t s) {
    const int nwin = win_count(N, wbits);
    if (nwin < 0)
        return fail_msg(5, "window engine: N=%d needs more than %d windows of 2^%d columns", N,
                        BMLHIP_WIN_MAX_WINDOWS, wbits);
    if (M >= (1 << 24))  // packed rows are addressed with 24-bit entry offsets
        return fail_msg(5, "window engine: rows of M=%d entries (at most 2^24 - 1)", M);
    unsigned *P = (unsigned *) bmlhip_scratch(BMLHIP_WIN_PACK_SLOT,
                                              sizeof(unsigned) * PkW<T>::n * ((size_t) N * M + 1));
    int *boff = (int *) bmlhip_scratch(BMLHIP_WIN_BOFF_SLOT,
                                       sizeof(int) * ((size_t) N * (nwin + 1) + 1));
    // the row masks, then the row maxima
    unsigned long long *wmask = (unsigned long long *) bmlhip_scratch(
        BMLHIP_WIN_MASK_SLOT, sizeof(unsigned long long) * 2 * ((size_t) N + 1));
    if (!P || !boff || !wmask) return fail_msg(4, "scratch allocation failed");
    double *rowmax = (double *) (wmask + N + 1);
    wp.N = N;
    wp.MP = M;
    wp.nwin = nwin;
    wp.P = P;
    wp.boff = boff;
    wp.wmask = wmask;
    wp.rowmax = rowmax;
    // a window slot sums at most M products; each is below 2^(50 - fx_extra)
    // units, so the sum stays below 2^62 while M <= 2^(12 + fx_extra)
    wp.fx_extra = 0;
    while ((long long) M > (1ll << (12 + wp.fx_extra))) wp.fx_extra++;
    if (N > 0)
        hipLaunchKernelGGL(k_win_pack<T>, dim3((N + 3) / 4), dim3(256), 0, s, N, M, (const T *) Xv,
                           Xi, Xn, M, nwin, wbits, P, boff, wmask, rowtr, rowmax, skip, row_sys,
                           sys_done);
    return report("k_win_pack");
}

// X^2 on the window engine (a->d_Av == a->d_Bv): pack, rows, mirror (sym),
// row records and traces, as spgemm_typed's other engines.
template <typename T>
int spgemm_win_typed(const bmlhip_spgemm_args *a) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    WinParams wp;
    memset(&wp, 0, sizeof wp);
    const int *row_sys = a->batch ? a->batch->d_row_sys : nullptr;
    const int *sys_done = a->batch ? a->batch->d_done : nullptr;
    if ((rc = win_pack<T>(a->N, a->MB, a->d_Bv, a->d_Bi, a->d_Bn, a->wbits, nullptr, a->d_skip,
                          row_sys, sys_done, wp, s)))
        return rc;
    int rows = a->wrows;
    if ((rc = win_union<T>(rows, a->N, a->MB, a->wbits, a->waves, a->d_Bv, a->d_Bi, a->d_Bn,
                           a->d_skip, row_sys, sys_done, wp, s)))
        return rc;
    wp.streamed = rows > 1 ? (unsigned long long *) &bmlhip_status_dev()->prof_streamed[PrecOf<T>::value]
                           : nullptr;
    // the profiling window spans the product and the mirror
    if (a->N > 0) bmlhip_prof_mark_begin(PrecOf<T>::value);
    wp.MC = a->MC;
    wp.Cv = a->d_Cv;
    wp.Ci = a->d_Ci;
    wp.Cn = a->d_Cn;
    wp.thr = a->threshold;
    wp.rowtr = a->batch ? a->batch->d_rowtr
                        : (double *) bmlhip_scratch(0, sizeof(double) * 2 * (size_t) (a->N + 1));
    wp.rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (a->N + 1));
    wp.skip = a->d_skip;
    wp.row_sys = row_sys;
    wp.sys_done = sys_done;
    wp.remap = a->remap;
    if (!wp.rowtr || !wp.rowstat) return fail_msg(4, "scratch allocation failed");
    const bool sym = a->sym != 0;
    int *fill = nullptr;
    if (sym) {
        fill = (int *) bmlhip_scratch_zeroed(BMLHIP_SYMFILL_SLOT, sizeof(int) * (size_t) (a->N + 1));
        if (!fill) return fail_msg(4, "scratch allocation failed");
        rc = launch_win_rows<T, true, false>(wp, rows, a->wbits, a->waves, s);
    } else {
        rc = launch_win_rows<T, false, false>(wp, rows, a->wbits, a->waves, s);
    }
    if (rc) return rc;
    SpParams p{};
    memset(&p, 0, sizeof p);
    p.N = a->N;
    p.MC = a->MC;
    p.Cv = a->d_Cv;
    p.Ci = a->d_Ci;
    p.Cn = a->d_Cn;
    p.skip = a->d_skip;
    p.rowstat = wp.rowstat;
    p.row_sys = row_sys;
    p.sys_done = sys_done;
    if (sym && a->N > 0) {
        hipLaunchKernelGGL(k_sym_mirror<T>, dim3((a->N + 3) / 4), dim3(256), 0, s, p, fill);
        hipLaunchKernelGGL(k_sym_finish, dim3((a->N + 255) / 256), dim3(256), 0, s, p, fill);
        if ((rc = report("k_sym_mirror"))) return rc;
    }
    if (a->N > 0) bmlhip_prof_mark_end();
    hipLaunchKernelGGL(k_reduce_rowstat, dim3(rowstat_blocks(a->N)), dim3(256), 0, s,
                       (const int4 *) wp.rowstat, a->N, a->MC,
                       BMLHIP_RS_OVF_KEPT | BMLHIP_RS_STORED | BMLHIP_RS_PROF |
                           (rows > 1 ? BMLHIP_RS_OWN_STREAM : 0),
                       (int) PrecOf<T>::value, a->d_skip, bmlhip_status_dev());
    if ((rc = report("k_reduce_rowstat"))) return rc;
    bmlhip_status *st = bmlhip_status_dev();
    if (a->batch) {
        const bmlhip_batch *b = a->batch;
        hipLaunchKernelGGL(k_seg_sum, dim3(b->nsys), dim3(256), 0, s, wp.rowtr, b->d_row_off,
                           b->d_done, b->d_trx);
        hipLaunchKernelGGL(k_seg_sum, dim3(b->nsys), dim3(256), 0, s, wp.rowtr + a->N,
                           b->d_row_off, b->d_done, b->d_trx2);
        return report("k_seg_sum");
    }
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, wp.rowtr, a->N, &st->scalars[0]);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, wp.rowtr + a->N, a->N, &st->scalars[1]);
    return report("k_sum");
}

// One fused SP2 iteration on the window engine (symmetric X, in place).
template <typename T>
int sp2_window_typed(int N, int M, void *Xv, int *Xi, int *Xn, double thr, int wbits, int waves,
                     int wrows, bmlhip_sp2_state *d_state, double nocc, double tol,
                     double *d_hist, int maxit) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    bmlhip_status *st = bmlhip_status_dev();
    const int *skip = &d_state->done;
    double *rowtr = (double *) bmlhip_scratch(0, sizeof(double) * 2 * (size_t) (N + 1));
    double *rownorm = (double *) bmlhip_scratch(BMLHIP_WIN_NORM_SLOT, sizeof(double) * (size_t) (N + 1));
    int4 *rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (N + 1));
    int *fill = (int *) bmlhip_scratch_zeroed(BMLHIP_SYMFILL_SLOT, sizeof(int) * (size_t) (N + 1));
    if (!rowtr || !rownorm || !rowstat || !fill) return fail_msg(4, "scratch allocation failed");
    WinParams wp;
    memset(&wp, 0, sizeof wp);
    if ((rc = win_pack<T>(N, M, Xv, Xi, Xn, wbits, rowtr, skip, nullptr, nullptr, wp, s))) return rc;
    int rows = wrows;
    if ((rc = win_union<T>(rows, N, M, wbits, waves, Xv, Xi, Xn, skip, nullptr, nullptr, wp, s)))
        return rc;
    wp.streamed = rows > 1 ? (unsigned long long *) &st->prof_streamed[PrecOf<T>::value] : nullptr;
    // the branch needs tr X and tr X^2 before the product
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, rowtr, N, &st->scalars[0]);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, rowtr + N, N, &st->scalars[1]);
    if ((rc = report("k_sum"))) return rc;
    if ((rc = bmlhip_sp2_decide(d_state, nocc, d_hist, maxit))) return rc;
    wp.MC = M;
    wp.Cv = Xv;
    wp.Ci = Xi;
    wp.Cn = Xn;
    wp.thr = thr;
    wp.rownorm = rownorm;
    wp.ab = d_state->ab;
    wp.rowstat = rowstat;
    wp.skip = skip;
    // the profiling window spans the product and the mirror
    if (N > 0) bmlhip_prof_mark_begin(PrecOf<T>::value);
    if ((rc = launch_win_rows<T, true, true>(wp, rows, wbits, waves, s))) return rc;
    SpParams p{};
    memset(&p, 0, sizeof p);
    p.N = N;
    p.MC = M;
    p.Cv = Xv;
    p.Ci = Xi;
    p.Cn = Xn;
    p.skip = skip;
    p.rowstat = rowstat;
    if (N > 0) {
        hipLaunchKernelGGL(k_sym_mirror<T>, dim3((N + 3) / 4), dim3(256), 0, s, p, fill);
        hipLaunchKernelGGL(k_sym_finish, dim3((N + 255) / 256), dim3(256), 0, s, p, fill);
        if ((rc = report("k_sym_mirror"))) return rc;
        bmlhip_prof_mark_end();
    }
    hipLaunchKernelGGL(k_reduce_rowstat, dim3(rowstat_blocks(N)), dim3(256), 0, s,
                       (const int4 *) rowstat, N, M,
                       BMLHIP_RS_OVF_KEPT | BMLHIP_RS_STORED | BMLHIP_RS_PROF |
                           (rows > 1 ? BMLHIP_RS_OWN_STREAM : 0),
                       (int) PrecOf<T>::value, skip, st);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, rownorm, N, &st->scalars[2]);
    if ((rc = report("k_sum"))) return rc;
    return bmlhip_sp2_check(d_state, tol, d_hist, maxit, 1);
}

template <typename T>
int spgemm_typed(const bmlhip_spgemm_args *a) {
    if (a->fuse && (a->engine != BMLHIP_ROW_EXACT || a->batch || a->sym || a->d_ws ||
                    a->d_Cv == a->d_Av || a->d_Cv == a->d_Bv || a->kind == 2))
        return fail_msg(2, "fused multiply: exact engine, one system, C distinct from A and B");
    if (!a->d_ws && !a->fuse && a->engine == BMLHIP_ROW_WINDOW && a->kind == 0 && a->d_Av == a->d_Bv &&
        a->d_Ai == a->d_Bi && a->d_An == a->d_Bn && a->MA == a->MB)
        return spgemm_win_typed<T>(a);
    hipStream_t s = (hipStream_t) bmlhip_stream();
    // the row kernels hash columns with a 24-bit multiply and address rows
    // with 32-bit element offsets
    if (a->N >= (1 << 24) || (size_t) a->N * (size_t) (a->MA > a->MB ? a->MA : a->MB) >= (1ull << 32) ||
        (size_t) a->N * (size_t) a->MC >= (1ull << 32))
        return fail_msg(5, "row product: N=%d with M up to %d exceeds the 2^24 columns / 2^32 "
                           "entries the kernels address", a->N, a->MA > a->MB ? a->MA : a->MB);
    // a row's product count is kept in 32 bits
    if ((size_t) a->MA * (size_t) a->MB >= (1ull << 32))
        return fail_msg(5, "row product: MA*MB = %d*%d exceeds 2^32", a->MA, a->MB);
    SpParams p{};
    p.N = a->N;
    p.MA = a->MA;
    p.MB = a->MB;
    p.MC = a->MC;
    p.kind = a->kind;
    p.Av = a->d_Av;
    p.Ai = a->d_Ai;
    p.An = a->d_An;
    p.Bv = a->d_Bv;
    p.Bi = a->d_Bi;
    p.Bn = a->d_Bn;
    p.Cv = a->d_Cv;
    p.Ci = a->d_Ci;
    p.Cn = a->d_Cn;
    p.thr = a->threshold;
    p.skip = a->d_skip;
    p.st = bmlhip_status_dev();
    p.row_sys = a->batch ? a->batch->d_row_sys : nullptr;
    p.sys_done = a->batch ? a->batch->d_done : nullptr;
    p.remap = a->remap;
    p.sym = 0;
    p.fuse = a->fuse;
    p.alpha = a->alpha;
    p.beta = a->beta;
    // global-table pass: one table per block, >= 2x the row bound
    long long bound = (long long) a->MA * a->MB;
    if (bound > a->N) bound = a->N;
    long long gsize = 2 * bound + 64;
    if (gsize < 1024) gsize = 1024;
    gsize = (gsize + 7) & ~7ll;
    const size_t per_block = ((size_t) gsize + table_pad<64>()) * (sizeof(T) + 2 * sizeof(int));
    int gblocks = (int) (FALLBACK_BYTES / per_block);
    if (gblocks > FALLBACK_BLOCKS) gblocks = FALLBACK_BLOCKS;
    if (gblocks < 1) gblocks = 1;
    unsigned char *gtab;
    if (a->d_ws) {
        // the caller's workspace: the per-row arrays first (256-byte
        // aligned), the global-pass tables in what is left; nothing is
        // allocated and nothing in it needs to start at zero (each kernel
        // writes what a later one reads, the global tables are cleared by
        // their own kernel)
        if (a->batch || a->sym || a->engine == BMLHIP_ROW_WINDOW)
            return fail_msg(2, "row product on a workspace: no batch, symmetric or window path");
        unsigned char *w = (unsigned char *) a->d_ws;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            void *q = w + off;
            off += (bytes + 255) & ~(size_t) 255;
            return q;
        };
        p.rowtr = (double *) take(sizeof(double) * 2 * (size_t) (a->N + 1));
        p.ovf = (int *) take(sizeof(int) * (size_t) (a->N + 64));
        p.rowstat = (int4 *) take(sizeof(int4) * (size_t) (a->N + 1));
        if (off + per_block > a->ws_bytes)
            return fail_msg(5, "row product: N=%d (%zu bytes of row records) and one %zu-byte "
                               "global table exceed the %zu-byte workspace", a->N, off,
                            per_block, a->ws_bytes);
        const size_t fit = (a->ws_bytes - off) / per_block;
        if ((size_t) gblocks > fit) gblocks = (int) fit;
        gtab = w + off;
    } else {
        p.rowtr = a->batch ? a->batch->d_rowtr
                           : (double *) bmlhip_scratch(0, sizeof(double) * 2 * (size_t) (a->N + 1));
        p.ovf = (int *) bmlhip_scratch(1, sizeof(int) * (size_t) (a->N + 64));
        p.rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (a->N + 1));
        gtab = (unsigned char *) bmlhip_scratch(2, (size_t) gblocks * per_block + 256);
    }
    int *cursor = &p.st->cursor;
    if (!p.rowtr || !p.ovf || !p.rowstat || !gtab) {
        return fail_msg(4, "scratch allocation failed");
    }
    // LDS table: a multiple of 8 slots, at least 64, below 2^16 buckets
    long long tsize = ((long long) a->table + 7) & ~7ll;
    if (tsize < 64) tsize = 64;
    if (tsize > 65528) tsize = 65528;
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    const unsigned ts = (unsigned) tsize;
    const bool sym = a->sym && a->kind == 0 && a->engine == BMLHIP_ROW_UNORDERED && a->group == 64;
    // batched: the one-wavefront kernels run persistent over the work queue
    // (bmlhip_batch_queue built it for this iteration).  The unordered
    // engine's multi-wavefront blocks keep one block per row (rows of
    // finished systems return at once): on the 512-system batch at hopping
    // -0.1 they ran 147 solves/s so, against 129 persistent (each take waits
    // for the listed row's index before the row's loads can start).
    if (a->batch && a->batch->d_list && !(a->group == 64 && a->waves > 1)) {
        p.qlist = a->batch->d_list;
        p.qinfo = a->batch->d_qinfo;
    }
    int *fill = nullptr;
    if (sym) {
        fill = (int *) bmlhip_scratch_zeroed(BMLHIP_SYMFILL_SLOT, sizeof(int) * (size_t) (a->N + 1));
        if (!fill) return fail_msg(4, "scratch allocation failed");
        p.sym = 1;
        if (a->N > 0) bmlhip_prof_mark_begin(PrecOf<T>::value);  // ended after the mirror
        switch (a->bucket) {
        case 2: rc = launch_spgemm_unord_w<T, 2, true>(p, ts, a->waves, s); break;
        case 4: rc = launch_spgemm_unord_w<T, 4, true>(p, ts, a->waves, s); break;
        default: return fail_msg(2, "symmetric x2: bucket %d not one of 2/4", a->bucket);
        }
    } else if (a->engine == BMLHIP_ROW_UNORDERED && a->group == 64) {
        switch (a->bucket) {
        case 2: rc = launch_spgemm_unord_w<T, 2, false>(p, ts, a->waves, s); break;
        case 4: rc = launch_spgemm_unord_w<T, 4, false>(p, ts, a->waves, s); break;
        default: return fail_msg(2, "unordered engine: bucket %d not one of 2/4", a->bucket);
        }
    } else if (a->group == 64 && a->waves > 1) {
        switch (a->waves) {
        case 2: rc = launch_spgemm_mw<T, 2>(p, ts, s); break;
        case 3: rc = launch_spgemm_mw<T, 3>(p, ts, s); break;
        case 4: rc = launch_spgemm_mw<T, 4>(p, ts, s); break;
        case 5: rc = launch_spgemm_mw<T, 5>(p, ts, s); break;
        case 6: rc = launch_spgemm_mw<T, 6>(p, ts, s); break;
        case 8: rc = launch_spgemm_mw<T, 8>(p, ts, s); break;
        default: return fail_msg(2, "waves %d not one of 1..6 or 8", a->waves);
        }
    } else switch (a->group) {
    case 8: rc = launch_spgemm_bk<T, 8>(p, ts, a->bucket, s); break;
    case 16: rc = launch_spgemm_bk<T, 16>(p, ts, a->bucket, s); break;
    case 32: rc = launch_spgemm_bk<T, 32>(p, ts, a->bucket, s); break;
    case 64: rc = launch_spgemm_bk<T, 64>(p, ts, a->bucket, s); break;
    default:
        return fail_msg(2, "group %d not one of 8/16/32/64", a->group);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_spgemm_global<T>, dim3(gblocks), dim3(64), 0, s, p, (unsigned) gsize, gtab,
                       cursor);
    if ((rc = report("k_spgemm_global"))) return rc;
    if (sym && a->N > 0) {
        hipLaunchKernelGGL(k_sym_mirror<T>, dim3((a->N + 3) / 4), dim3(256), 0, s, p, fill);
        hipLaunchKernelGGL(k_sym_finish, dim3((a->N + 255) / 256), dim3(256), 0, s, p, fill);
        bmlhip_prof_mark_end();
        if ((rc = report("k_sym_mirror"))) return rc;
    }
    // kind 1 (AB) checks the distinct count against MC, as the reference does
    // (the fused epilogue folds that check into the kept count it records)
    hipLaunchKernelGGL(k_reduce_rowstat, dim3(rowstat_blocks(a->N)), dim3(256), 0, s,
                       (const int4 *) p.rowstat, a->N, a->MC,
                       (a->kind == 1 && !a->fuse ? BMLHIP_RS_OVF_DIST : BMLHIP_RS_OVF_KEPT) |
                           BMLHIP_RS_STORED | BMLHIP_RS_PROF,
                       (int) PrecOf<T>::value, p.skip, p.st);
    if ((rc = report("k_reduce_rowstat"))) return rc;
    if (a->kind == 0 && a->batch) {
        const bmlhip_batch *b = a->batch;
        hipLaunchKernelGGL(k_seg_sum, dim3(b->nsys), dim3(256), 0, s, p.rowtr, b->d_row_off,
                           b->d_done, b->d_trx);
        hipLaunchKernelGGL(k_seg_sum, dim3(b->nsys), dim3(256), 0, s, p.rowtr + a->N,
                           b->d_row_off, b->d_done, b->d_trx2);
        if ((rc = report("k_seg_sum"))) return rc;
    } else if (a->kind == 0) {
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, p.rowtr, a->N, &p.st->scalars[0]);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, p.rowtr + a->N, a->N,
                           &p.st->scalars[1]);
        if ((rc = report("k_sum"))) return rc;
    }
    return 0;
}

template <typename T>
int add_typed(int N, int MA, void *Av, int *Ai, int *An, int MB, const void *Bv, const int *Bi,
              const int *Bn, double alpha, double beta, double thr, int mode,
              const double *abdev = nullptr, const int *skip = nullptr,
              const bmlhip_batch *batch = nullptr, double *rownorm = nullptr) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    AddParams p{};
    p.N = N;
    p.MA = MA;
    p.MB = MB;
    p.mode = mode;
    p.Av = Av;
    p.Ai = Ai;
    p.An = An;
    p.Bv = Bv;
    p.Bi = Bi;
    p.Bn = Bn;
    p.alpha = alpha;
    p.beta = beta;
    p.thr = thr;
    p.ab = abdev;
    p.skip = skip;
    p.row_sys = batch ? batch->d_row_sys : nullptr;
    p.sys_done = batch ? batch->d_done : nullptr;
    p.ab_sys = batch ? batch->d_ab : nullptr;
    p.qlist = batch ? batch->d_list : nullptr;
    p.qinfo = batch ? batch->d_qinfo : nullptr;
    p.st = bmlhip_status_dev();
    p.rownorm = rownorm ? rownorm : (double *) bmlhip_scratch(0, sizeof(double) * (size_t) (N + 1));
    p.rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (N + 1));
    if (!p.rownorm || !p.rowstat) return 4;
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    // a row's union has at most MA + MB columns: the LDS table keeps it at
    // most 4/5 full (multiple of 64 slots; first-touch list in 16 bits)
    const int tbits = ceil_log2((long long) (MA + MB) * 4 / 3 + 64);
    const size_t lds = ((size_t) 1 << tbits) * (2 * sizeof(T) + 2 * sizeof(int));
    const unsigned tl = (unsigned) (((long long) (MA + MB) * 5 / 4 + 64 + 63) & ~63ll);
    const size_t lds_l = (size_t) tl * (2 * sizeof(T) + sizeof(int) + sizeof(uint16_t));
    if (N > 0) {
        if (lds_l <= 64 * 1024 && tl < 65536) {
            const int blocks = !p.qlist ? N
                               : persistent_blocks((const void *) k_add_lds<T>, 64, lds_l, N);
            hipLaunchKernelGGL(k_add_lds<T>, dim3(blocks), dim3(64), lds_l, s, p, tl);
            if ((rc = report("k_add_lds"))) return rc;
        } else {
            const int blocks = 256;
            unsigned char *tabs = (unsigned char *) bmlhip_scratch(2, blocks * lds + 256);
            if (!tabs) return 4;
            hipLaunchKernelGGL(k_add_global<T>, dim3(blocks), dim3(64), 0, s, p, tbits, tabs);
            if ((rc = report("k_add_global"))) return rc;
        }
        hipLaunchKernelGGL(k_reduce_rowstat, dim3(rowstat_blocks(N)), dim3(256), 0, s,
                           (const int4 *) p.rowstat, N, MA,
                           BMLHIP_RS_OVF_KEPT | BMLHIP_RS_STORED, 0, skip, p.st);
        if ((rc = report("k_reduce_rowstat"))) return rc;
    }
    if (mode && batch) {
        hipLaunchKernelGGL(k_seg_sum, dim3(batch->nsys), dim3(256), 0, s, p.rownorm,
                           batch->d_row_off, batch->d_done, batch->d_norm2);
        if ((rc = report("k_seg_sum"))) return rc;
    } else if (mode) {
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, p.rownorm, N, &p.st->scalars[2]);
        if ((rc = report("k_sum"))) return rc;
    }
    return 0;
}

template <typename T>
int scale_add_identity_typed(int N, int M, void *Av, int *Ai, int *An, const void *s_bits,
                             double beta, double thr, int do_identity) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    T sv;
    memset(&sv, 0, sizeof sv);
    if (s_bits) memcpy(&sv, s_bits, sizeof(T));
    if (N > 0)
        hipLaunchKernelGGL(k_scale_add_identity<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, M,
                           (T *) Av, Ai, An, sv, s_bits ? 1 : 0, beta, thr, do_identity,
                           bmlhip_status_dev());
    return report("k_scale_add_identity");
}

template <typename T>
int threshold_typed(int N, int M, void *Av, int *Ai, int *An, double thr) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (N > 0)
        hipLaunchKernelGGL(k_threshold<T>, dim3((N + 3) / 4), dim3(256), 0, s, N, M, (T *) Av, Ai,
                           An, thr);
    return report("k_threshold");
}

template <typename T>
int trace_typed(int N, int M, const void *Av, const int *Ai, const int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    double *rowv = (double *) bmlhip_scratch(0, sizeof(double) * (size_t) (N + 1));
    if (!rowv) return 4;
    if (N > 0)
        hipLaunchKernelGGL(k_trace_rows<T>, dim3((N + 3) / 4), dim3(256), 0, s, N, M,
                           (const T *) Av, Ai, An, rowv);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, rowv, N, &bmlhip_status_dev()->scalars[0]);
    return report("k_trace");
}

template <typename T>
int gershgorin_typed(int N, int M, const void *Av, const int *Ai, const int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    double *rows = (double *) bmlhip_scratch(0, sizeof(double) * 2 * (size_t) (N + 1));
    if (!rows) return 4;
    if (N > 0)
        hipLaunchKernelGGL(k_gersh_rows<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, M,
                           (const T *) Av, Ai, An, rows, rows + N);
    // emax starts at DBL_MIN as in the reference
    hipLaunchKernelGGL(k_minmax, dim3(1), dim3(1024), 0, s, rows, rows + N, N,
                       &bmlhip_status_dev()->scalars[0], 2.2250738585072014e-308);
    return report("k_gershgorin");
}

template <typename T>
int copy_rows_typed(int N, int MA, const void *Av, const int *Ai, const int *An, int MB, void *Bv,
                    int *Bi, int *Bn, int shift = 0) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (N > 0)
        hipLaunchKernelGGL(k_copy_rows<T>, dim3((N + 3) / 4), dim3(256), 0, s, N, MA,
                           (const T *) Av, Ai, An, MB, (T *) Bv, Bi, Bn, bmlhip_status_dev(),
                           shift);
    return report("k_copy_rows");
}

template <typename T>
int gershgorin_batch_typed(int N, int M, const void *Av, const int *Ai, const int *An,
                           const bmlhip_batch *b, double *emin, double *emax) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    double *rows = (double *) bmlhip_scratch(0, sizeof(double) * 2 * (size_t) (N + 1));
    if (!rows) return 4;
    if (N > 0)
        hipLaunchKernelGGL(k_gersh_rows<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, M,
                           (const T *) Av, Ai, An, rows, rows + N);
    hipLaunchKernelGGL(k_seg_minmax, dim3(b->nsys), dim3(256), 0, s, rows, rows + N, b->d_row_off,
                       emin, emax);
    return report("k_gershgorin_batch");
}

template <typename T>
int normalize_batch_typed(int N, int M, void *Av, int *Ai, int *An, const bmlhip_batch *b,
                          const double *emin, const double *emax) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (N > 0)
        hipLaunchKernelGGL(k_normalize_batch<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, M,
                           (T *) Av, Ai, An, b->d_row_sys, emin, emax, bmlhip_status_dev());
    return report("k_normalize_batch");
}

template <typename T>
int identity_typed(int N, int M, void *Av, int *Ai, int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (N > 0)
        hipLaunchKernelGGL(k_identity<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, M, (T *) Av,
                           Ai, An);
    return report("k_identity");
}

// Symmetry test: entry (i, j, v) needs (j, i, conj(v)) in row j, equal by
// value (same_value).  One thread per stored entry; mismatches are counted.
template <typename T>
__global__ __launch_bounds__(256) void k_check_sym(int N, int M, const T *Av, const int *Ai,
                                                   const int *An, int *bad) {
    const size_t t = (size_t) blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t) N * M) return;
    const int i = (int) (t / M), e = (int) (t % M);
    if (e >= An[i]) return;
    const int j = Ai[t];
    if (j < 0 || j >= N) {
        atomicAdd(bad, 1);
        return;
    }
    const T v = Elem<T>::conj(Av[t]);
    const int n = An[j];
    bool ok = false;
    for (int k = 0; k < n && !ok; k++)
        if (Ai[(size_t) j * M + k] == i) ok = same_value(Av[(size_t) j * M + k], v);
    if (!ok) atomicAdd(bad, 1);
}

template <typename T>
int check_sym_typed(int N, int M, const void *Av, const int *Ai, const int *An, int *symmetric) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int *bad = (int *) bmlhip_scratch(BMLHIP_SYMFILL_SLOT + 1, sizeof(int) * 64);
    if (!bad) return fail_msg(4, "scratch allocation failed");
    if (hipMemsetAsync(bad, 0, sizeof(int), s) != hipSuccess) return fail_msg(3, "memset failed");
    const size_t n = (size_t) N * M;
    if (n > 0)
        hipLaunchKernelGGL(k_check_sym<T>, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, N, M,
                           (const T *) Av, Ai, An, bad);
    int rc = report("k_check_sym");
    if (rc) return rc;
    int h = 0;
    if (hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail_msg(3, "symmetry test: copy failed");
    *symmetric = h == 0;
    return 0;
}

}  // namespace

#define DISPATCH(prec, call_tmpl)                                                  \
    switch (prec) {                                                                \
    case 1: return call_tmpl(float);                                               \
    case 2: return call_tmpl(double);                                              \
    case 3: return call_tmpl(cf32);                                                \
    case 4: return call_tmpl(cf64);                                                \
    default: return fail_msg(2, "unknown precision %d", prec);                      \
    }

extern "C" {

int bmlhip_spgemm(const bmlhip_spgemm_args *a) {
#define CALL(T) spgemm_typed<T>(a)
    DISPATCH(a->prec, CALL)
#undef CALL
}

int bmlhip_add(int prec, int N, int MA, void *Av, int *Ai, int *An, int MB, const void *Bv,
               const int *Bi, const int *Bn, double alpha, double beta, double thr, int mode) {
#define CALL(T) add_typed<T>(N, MA, Av, Ai, An, MB, Bv, Bi, Bn, alpha, beta, thr, mode)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_rowstat_reduce(int N, int MC, int what, int prof, const int *skip) {
    void *rs = bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (N + 1));
    if (!rs) return fail_msg(4, "scratch allocation failed");
    hipLaunchKernelGGL(k_reduce_rowstat, dim3(rowstat_blocks(N)), dim3(256), 0,
                       (hipStream_t) bmlhip_stream(), (const int4 *) rs, N, MC, what, prof, skip,
                       bmlhip_status_dev());
    return report("k_reduce_rowstat");
}

int bmlhip_sp2_window_step(int prec, int N, int M, void *Xv, int *Xi, int *Xn, double thr,
                           int wbits, int waves, int wrows, bmlhip_sp2_state *d_state,
                           double nocc, double tol, double *d_hist, int maxit) {
#define CALL(T) sp2_window_typed<T>(N, M, Xv, Xi, Xn, thr, wbits, waves, wrows, d_state, nocc, \
                                    tol, d_hist, maxit)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_is_symmetric(int prec, int N, int M, const void *Av, const int *Ai, const int *An,
                        int *symmetric) {
#define CALL(T) check_sym_typed<T>(N, M, Av, Ai, An, symmetric)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_add_dev(int prec, int N, int MA, void *Av, int *Ai, int *An, int MB, const void *Bv,
                   const int *Bi, const int *Bn, const double *ab, double thr, int mode,
                   const int *skip) {
#define CALL(T) add_typed<T>(N, MA, Av, Ai, An, MB, Bv, Bi, Bn, 0.0, 0.0, thr, mode, ab, skip)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_scale_add_identity(int prec, int N, int M, void *Av, int *Ai, int *An,
                              const void *s_bits, double beta, double thr, int do_identity) {
#define CALL(T) scale_add_identity_typed<T>(N, M, Av, Ai, An, s_bits, beta, thr, do_identity)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_threshold(int prec, int N, int M, void *Av, int *Ai, int *An, double thr) {
#define CALL(T) threshold_typed<T>(N, M, Av, Ai, An, thr)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_trace(int prec, int N, int M, const void *Av, const int *Ai, const int *An) {
#define CALL(T) trace_typed<T>(N, M, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_gershgorin(int prec, int N, int M, const void *Av, const int *Ai, const int *An) {
#define CALL(T) gershgorin_typed<T>(N, M, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_copy_rows(int prec, int N, int MA, const void *Av, const int *Ai, const int *An,
                     int MB, void *Bv, int *Bi, int *Bn) {
#define CALL(T) copy_rows_typed<T>(N, MA, Av, Ai, An, MB, Bv, Bi, Bn)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_identity(int prec, int N, int M, void *Av, int *Ai, int *An) {
#define CALL(T) identity_typed<T>(N, M, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_batch_queue(const bmlhip_batch *b, int N) {
    if (!b->d_list || !b->d_qinfo || !b->d_act_off) return 0;
    int4 *rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (size_t) (N + 1));
    if (!rowstat) return fail_msg(4, "scratch allocation failed");
    hipStream_t s = (hipStream_t) bmlhip_stream();
    hipLaunchKernelGGL(k_batch_scan, dim3(1), dim3(1024), 0, s, b->nsys, b->d_row_off, b->d_done,
                       b->d_act_off, b->d_qinfo);
    if (b->nsys > 0)
        hipLaunchKernelGGL(k_batch_fill, dim3(b->nsys), dim3(256), 0, s, b->d_row_off, b->d_done,
                           b->d_act_off, b->d_list, rowstat);
    return report("k_batch_fill");
}

int bmlhip_add_batch(int prec, int N, int MA, void *Av, int *Ai, int *An, int MB, const void *Bv,
                     const int *Bi, const int *Bn, double thr, int mode, const bmlhip_batch *b) {
#define CALL(T) add_typed<T>(N, MA, Av, Ai, An, MB, Bv, Bi, Bn, 0.0, 0.0, thr, mode, nullptr, nullptr, b, b->d_rownorm)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_gershgorin_batch(int prec, int N, int M, const void *Av, const int *Ai, const int *An,
                            const bmlhip_batch *b, double *emin, double *emax) {
#define CALL(T) gershgorin_batch_typed<T>(N, M, Av, Ai, An, b, emin, emax)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_normalize_batch(int prec, int N, int M, void *Av, int *Ai, int *An,
                           const bmlhip_batch *b, const double *emin, const double *emax) {
#define CALL(T) normalize_batch_typed<T>(N, M, Av, Ai, An, b, emin, emax)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_copy_rows_shift(int prec, int N, int MA, const void *Av, const int *Ai, const int *An,
                           int MB, void *Bv, int *Bi, int *Bn, int shift) {
#define CALL(T) copy_rows_typed<T>(N, MA, Av, Ai, An, MB, Bv, Bi, Bn, shift)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_sum_to_status(const double *x, int n, int which) {
    if (which < 0 || which > 3) return fail_msg(2, "status scalar %d out of range", which);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, (hipStream_t) bmlhip_stream(), x, n,
                       &bmlhip_status_dev()->scalars[which]);
    return report("k_sum");
}

int bmlhip_minmax_to_status(const double *lo, const double *hi, int n) {
    // emax starts at DBL_MIN as in the reference
    hipLaunchKernelGGL(k_minmax, dim3(1), dim3(1024), 0, (hipStream_t) bmlhip_stream(), lo, hi,
                       n, &bmlhip_status_dev()->scalars[0], 2.2250738585072014e-308);
    return report("k_minmax");
}

int bmlhip_row_stats(int N, const int *An) {
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (N > 0)
        hipLaunchKernelGGL(k_row_stats, dim3((N + 255) / 256), dim3(256), 0,
                           (hipStream_t) bmlhip_stream(), N, An, bmlhip_status_dev());
    return report("k_row_stats");
}

}  // extern "C"

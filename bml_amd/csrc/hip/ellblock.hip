// gfx950 kernels for the ELLBLOCK thresholded multiply path.
//
// Storage: block slot (ib, jp) owns one BT x BT tile (BT = 8, 16 or 32),
// the bsize[ib] x bsize[jb] block in its top-left corner, zeros elsewhere.
// Zero padding makes every tile product a full BT x BT product: padded
// columns of the left tile meet padded rows of the right one.
//
// Product C = A*B (x2 when A = B), one block row ib at a time, in two passes:
//
//  k_eb_symbolic  one wavefront per block row: walks the (jp, kp) block
//                 pairs in the reference's order (jp, then kp), gives every
//                 touched block column kb a slot in first-touch order (LDS
//                 hash; a ballot prefix count ranks the new ones), counts the
//                 pairs per slot, and writes the pairs grouped by slot (jp
//                 order kept within a slot: a stable counting sort).
//  k_eb_numeric   four wavefronts per block row claim the row's slots (MFMA
//                 engines: batches of whole slots) from an LDS counter,
//                 accumulate each slot's tile products in registers (exact
//                 VALU dot products in the reference's order, or MFMA),
//                 take the block's max |element| and store a kept tile at
//                 its slot's position; the kept tiles are then compacted
//                 into slot order (the reference's output order).  For the
//                 x2 of an exactly symmetric X each block row forms only
//                 the blocks it owns and mirrors them (k_eb_compact then
//                 compacts).
//
// Keep rule (bml_multiply_ellblock_typed.c:406-408): the diagonal block, or
// max |element| > threshold, compared in double.  A block row touching more
// block columns than C holds is the reference's fatal overflow (:393-398).
#include "../bmlhip_abi.h"
#include "devtypes.h"

#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <type_traits>

using namespace bmlhip;

namespace {

constexpr int EMPTY = -1;
#ifndef BMLHIP_EB_WAVES
#define BMLHIP_EB_WAVES 4
#endif
#ifndef BMLHIP_EB_OCC_BF16
#define BMLHIP_EB_OCC_BF16 3
#endif
#ifndef BMLHIP_EB_BATCH
#define BMLHIP_EB_BATCH 16
#endif
constexpr int NUM_WAVES = BMLHIP_EB_WAVES;  // wavefronts per block row in k_eb_numeric
// pairs per batch, the unit of work a wavefront of k_eb_numeric claims: batch
// k holds the slots whose first pair lies in [EB_BATCH k, EB_BATCH (k+1)).
// 16 against 32: 1.44 against 1.47 ms (FP64), 0.85 against 0.90 ms (FP32)
// for the symmetric X^2 of the 16x16x8-site, 32-orbital lattice; 2 or 8
// wavefronts per block row ran 10-20% slower than 4.
constexpr int EB_BATCH = BMLHIP_EB_BATCH;

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct EbParams {
    int prec, kind, NB, bt;
    const int *bsize;
    int MA, MB, MC;
    const void *Av;
    const int *Ai, *An;
    const void *Bv;
    const int *Bi, *Bn;
    void *Cv;
    int *Ci, *Cn;
    double thr;
    int maxs, maxp;  // slots and pairs per block row held in the scratch lists
    int *cnt;        // [NB] touched block columns
    int *skb;        // [NB*maxs] block column of each slot
    int *soff;       // [NB*(maxs+1)] start of each slot's pair list
    int *pairs;      // [NB*maxp] jp << 16 | kp, grouped by slot
    int *pslot;      // [NB*maxp] slot of each pair
    int *bfirst;     // [NB*(maxp/EB_BATCH+2)] batch k: slots [bfirst[k], bfirst[k+1])
    double *rowtr;   // [2*NB] per block row tr X, tr X^2 (x2)
    int4 *rowstat;   // [NB] (kept, touched, block pairs, nnzb of A's row)
    bmlhip_status *st;
    const int *skip;
    int remap;  // numeric pass: XCD-aware block-row order (xcd_block)
    int sym;    // x2 of an exactly symmetric X: each row forms the slots it owns (sym_owner)
    int *kflag; // [NB*maxs] sym: kept flag of every slot, set by its row or its mirror
    const __bf16 *A16, *B16;  // bf16 engine: A's tiles row-major, B's transposed (k_eb_bf16)
    int ka;  // MFMA engines: the first ka A tiles of a block row are staged in LDS
    // fused SP2 update (sym x2 on an MFMA engine): C <- ab[0] X + ab[1] X^2,
    // A's block columns in LDS at byte acol_off of the dynamic area, the
    // rows' |X - X^2|^2 into rownorm
    int fuse, acol_off;
    const double *ab;
    double *rownorm;
};

// The symmetric X^2 forms each off-diagonal pair of blocks (i, k), (k, i)
// once, in the block row this picks: a hash of the unordered pair, so every
// block row forms about half of its blocks whatever the ordering (on a
// periodic lattice "k >= i" would leave the first planes' rows with most of
// the work and the last planes' with little, one XCD's share against
// another's).  Symmetric in (i, k); the diagonal block is its row's.
__device__ __forceinline__ bool sym_owner(int ib, int kb) {
    const unsigned lo = (unsigned) (ib < kb ? ib : kb), hi = (unsigned) (ib < kb ? kb : ib);
    const unsigned h = lo * 0x9E3779B1u ^ hi * 0x85EBCA77u;
    return ib == kb || ((h >> 31) != 0) == (ib == (int) lo);
}

__device__ __forceinline__ unsigned eb_hash(int kb, unsigned mask) {
    return ((unsigned) kb * 0x9E3779B1u >> 7) & mask;
}

// ---------------------------------------------------------------------------
// symbolic pass

// Find kb in the LDS table (keys/sval, tsize slots) or insert it; returns the
// table position, *isnew set when this lane inserted it.  The columns of one
// chunk are distinct (a row of B holds each block column once), so no two
// lanes of a chunk insert the same key.
// With insert false a missing key is not added: the result is NOSLOT.
constexpr unsigned NOSLOT = ~0u;
__device__ __forceinline__ unsigned find_or_insert(int *keys, int kb, unsigned mask, bool *isnew,
                                                   bool insert) {
    unsigned h = eb_hash(kb, mask);
    while (true) {
        const int k = keys[h];
        if (k == kb) return h;
        if (k == EMPTY) {
            if (!insert) return NOSLOT;
            const int old = atomicCAS(&keys[h], EMPTY, kb);
            if (old == EMPTY) {
                *isnew = true;
                return h;
            }
            if (old == kb) return h;
        }
        h = (h + 1u) & mask;
    }
}

// cap: slots given out (a row touching more than MC block columns is an
// overflow whatever the rest; with count >= cap no slot is added, and the
// pairs of columns without a slot are left out of both passes alike), so
// the table and counters are sized by MC, not by the NB columns a row
// could touch
__global__ __launch_bounds__(64) void k_eb_symbolic(EbParams p, int tbits, int cap) {
    extern __shared__ int sm[];
    if (p.skip && *p.skip) return;
    const int ib = blockIdx.x;
    const int lane = threadIdx.x;
    const unsigned tsize = 1u << tbits, mask = tsize - 1u;
    int *keys = sm;
    int *sval = sm + tsize;
    int *fill = sm + 2 * tsize;  // cap + 63 counters
    // the row's block columns jb and the lengths of rows jb of B, loaded
    // lane-parallel once (not two dependent loads per jp)
    int *jbs = fill + cap + 64, *nbs = jbs + p.MA;
    const int nA = p.An[ib];
    for (unsigned s = lane; s < tsize; s += 64) keys[s] = EMPTY;
    for (int s = lane; s < cap + 64; s += 64) fill[s] = 0;
    for (int q = lane; q < nA; q += 64) {
        const int jb = p.Ai[(size_t) ib * p.MA + q];
        jbs[q] = jb;
        nbs[q] = p.Bn[jb];
    }
    __syncthreads();
    const uint64_t ltmask = (1ull << lane) - 1ull;
    int count = 0;
    // the first chunk of row jbs[jp] of B, lane-wise; the next one is loaded
    // while the current is hashed
    auto first_chunk = [&](int jp) {
        return lane < nbs[jp] ? p.Bi[(size_t) jbs[jp] * p.MB + lane] : -1;
    };
    // fused SP2 update: the row's own block columns take the first slots,
    // in A's order (every one is reached through its row's diagonal block,
    // which the fused update keeps); near convergence the kept blocks are
    // about X's own, so the compaction moves few tiles
    if (p.fuse)
        for (int q0 = 0; q0 < nA; q0 += 64) {
            const int q = q0 + lane;
            bool isnew = false;
            unsigned h = NOSLOT;
            if (q < nA) h = find_or_insert(keys, jbs[q], mask, &isnew, count < cap);
            const uint64_t nm = ballot64(isnew);
            if (isnew) {
                const int r = count + popc64(nm & ltmask);
                sval[h] = r;
                p.skb[(size_t) ib * p.maxs + r] = jbs[q];
            }
            count += popc64(nm);
            __syncthreads();
        }
    // pass 1: slots in first-touch order and pairs per slot
    int kbn = nA > 0 ? first_chunk(0) : -1;
    for (int jp = 0; jp < nA; jp++) {
        const int jb = jbs[jp];
        const int nb = nbs[jp];
        const int kbc = kbn;
        if (jp + 1 < nA) kbn = first_chunk(jp + 1);
        for (int k0 = 0; k0 < nb; k0 += 64) {
            const int kp = k0 + lane;
            const bool act = kp < nb;
            bool isnew = false;
            unsigned h = 0;
            int kb = 0;
            if (act) {
                kb = k0 == 0 ? kbc : p.Bi[(size_t) jb * p.MB + kp];
                h = find_or_insert(keys, kb, mask, &isnew, count < cap);
            }
            const uint64_t nm = ballot64(isnew);
            if (isnew) {
                const int r = count + popc64(nm & ltmask);
                sval[h] = r;
                p.skb[(size_t) ib * p.maxs + r] = kb;
            }
            count += popc64(nm);
            __syncthreads();  // new slot numbers visible to every lane
            // the chunk's slots are distinct; with p.sym the slots row ib
            // does not own get no pairs (their tiles are mirrors)
            if (act && h != NOSLOT && (!p.sym || sym_owner(ib, kb))) fill[sval[h]] += 1;
            __syncthreads();
        }
    }
    // exclusive scan of the per-slot counts -> pair-list offsets
    int running = 0;
    int *off = p.soff + (size_t) ib * (p.maxs + 1);
    for (int s0 = 0; s0 < count; s0 += 64) {
        const int s = s0 + lane;
        const int v = s < count ? fill[s] : 0;
        int inc = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (s < count) {
            off[s] = running + inc - v;
            fill[s] = running + inc - v;  // next write position of the slot
        }
        running += __shfl(inc, 63, 64);
    }
    if (lane == 0) off[count] = running;
    // the loop below reads fill[s - 1], written by lane s - 1 above
    __syncthreads();
    // batch starts: bfirst[k] = the first slot s with off[s] >= EB_BATCH k
    // (count for the end), k <= running / EB_BATCH + 1
    int *bf = p.bfirst + (size_t) ib * (p.maxp / EB_BATCH + 2);
    for (int s0 = 0; s0 <= count; s0 += 64) {
        const int s = s0 + lane;
        if (s <= count) {
            const int cur = s < count ? fill[s] : running;
            const int klo = s == 0 ? 0 : fill[s - 1] / EB_BATCH + 1;
            for (int k = klo; k <= cur / EB_BATCH; k++) bf[k] = s;
        }
    }
    if (lane == 0) bf[running / EB_BATCH + 1] = count;
    __syncthreads();
    // pass 2: pairs grouped by slot, jp order kept within a slot
    int *pr = p.pairs + (size_t) ib * p.maxp;
    kbn = nA > 0 ? first_chunk(0) : -1;
    for (int jp = 0; jp < nA; jp++) {
        const int jb = jbs[jp];
        const int nb = nbs[jp];
        const int kbc = kbn;
        if (jp + 1 < nA) kbn = first_chunk(jp + 1);
        for (int k0 = 0; k0 < nb; k0 += 64) {
            const int kp = k0 + lane;
            const int kb = kp < nb ? (k0 == 0 ? kbc : p.Bi[(size_t) jb * p.MB + kp]) : -1;
            unsigned h = eb_hash(kb, mask);
            if (kp < nb)
                while (keys[h] != kb && keys[h] != EMPTY) h = (h + 1u) & mask;
            if (kp < nb && keys[h] == kb && (!p.sym || sym_owner(ib, kb))) {
                const int s = sval[h];
                pr[fill[s]] = (int) (((unsigned) jp << 16) | (unsigned) kp);
                p.pslot[(size_t) ib * p.maxp + fill[s]] = s;
                fill[s] += 1;
            }
            __syncthreads();
        }
    }
    if (p.sym)
        for (int q = lane; q < count; q += 64) p.kflag[(size_t) ib * p.maxs + q] = 0;
    // the numeric pass records the row's counts (touched > MC is an overflow)
    if (lane == 0) p.cnt[ib] = count;
}

// ---------------------------------------------------------------------------
// tile accumulators

// Exact engine: lane l holds elements e = l + 64q of the tile.  Each element
// takes its length-bk dot product left to right, then is added to the
// accumulator: the order of the reference's block kernels.
template <typename T, int BT>
struct ExactAcc {
    static constexpr bool kPipelined = false;
    static constexpr bool kStage = false;
    static constexpr int Q = (BT * BT + 63) / 64;
    T a[Q];
    __device__ void zero() {
#pragma unroll
        for (int q = 0; q < Q; q++) a[q] = Elem<T>::zero();
    }
    __device__ void mac(const T *L, const T *R, int bi, int bk, int bj) {
        using E = Elem<T>;
        const int lane = lane_id();
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = lane + 64 * q;
            const int ii = e / BT, jj = e % BT;
            if (e < BT * BT && ii < bi && jj < bj) {
                T d = E::mul(L[ii * BT], R[jj]);
                for (int k = 1; k < bk; k++) d = E::add(d, E::mul(L[ii * BT + k], R[k * BT + jj]));
                a[q] = E::add(a[q], d);
            }
        }
    }
    __device__ double maxmag() const {
        double m = 0.0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const double v = Elem<T>::mag(a[q]);
            m = v > m ? v : m;
        }
        return m;
    }
    // the tile into C, rows ld elements apart
    __device__ void store(T *C, int ld = BT) const {
        const int lane = lane_id();
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = lane + 64 * q;
            if (e < BT * BT) C[(e / BT) * ld + e % BT] = a[q];
        }
    }
    // diagonal element kk held by this lane (in diag[kk]) for kk < BT
    __device__ void diag(T *diag) const {
        const int lane = lane_id();
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = lane + 64 * q;
            if (e < BT * BT && e / BT == e % BT) diag[e / BT] = a[q];
        }
    }
};

// loads through an explicit address space, global (1) or LDS (3), so that
// code choosing between a staged and an unstaged tile keeps both kinds of
// load instead of falling back to flat loads
template <typename X> __device__ __forceinline__ X gld(const X *a) {
    return *(const __attribute__((address_space(1))) X *) a;
}
template <typename X> using lds_ptr = const __attribute__((address_space(3))) X *;

// MFMA engine, single real, 32 x 32 tiles: v_mfma_f32_32x32x2f32 takes a
// 32x2 slice of L (lane l: row l%32) and a 2x32 slice of R (lane l: column
// l%32); the two k of an instruction come from the two lane halves.  Step s
// pairs k = s (lanes 0-31) with k = 16 + s (lanes 32-63), so a lane reads 16
// contiguous elements of its row of L and the lanes of a half read one
// contiguous row of R per step.  Output: lane l, register v holds row
// 8*(v/4) + 4*(l/32) + v%4, column l%32.
struct MfmaAccF32 {
    static constexpr bool kPipelined = true;
    floatx16 a;
    // one pair's operands in registers
    struct Frag {
        float lv[16], rv[16];
    };
    __device__ void zero() {
#pragma unroll
        for (int v = 0; v < 16; v++) a[v] = 0.0f;
    }
    __device__ static void load(const float *L, const float *R, Frag &f) {
        const int lane = lane_id();
        const int i = lane & 31, h = lane >> 5;
        const floatx4 *l4 = reinterpret_cast<const floatx4 *>(L + i * 32 + h * 16);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const floatx4 t = gld(l4 + q);
            f.lv[4 * q] = t[0];
            f.lv[4 * q + 1] = t[1];
            f.lv[4 * q + 2] = t[2];
            f.lv[4 * q + 3] = t[3];
        }
#pragma unroll
        for (int s = 0; s < 16; s++) f.rv[s] = gld(R + (h * 16 + s) * 32 + i);
    }
    // A tile staged in LDS: row r's 16-byte chunk c (8 per row) lives at
    // chunk c ^ (r & 7), so the 8 lanes that read rows r .. r + 7 at one
    // chunk hit 8 different 16-byte bank groups
    using TA = float;
    static constexpr bool kStage = true;  // left tiles staged in LDS
    static constexpr int kChunks = 8;     // 16-byte chunks per tile row
    __device__ static int swz(int r, int c) { return c ^ (r & 7); }
    __device__ static void load_s(lds_ptr<float> Ls, const float *R, Frag &f) {
        const int lane = lane_id();
        const int i = lane & 31, h = lane >> 5;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const floatx4 t = *(lds_ptr<floatx4>) (Ls + i * 32 + 4 * swz(i, 4 * h + q));
            f.lv[4 * q] = t[0];
            f.lv[4 * q + 1] = t[1];
            f.lv[4 * q + 2] = t[2];
            f.lv[4 * q + 3] = t[3];
        }
#pragma unroll
        for (int s = 0; s < 16; s++) f.rv[s] = gld(R + (h * 16 + s) * 32 + i);
    }
    __device__ void mma(const Frag &f) {
#pragma unroll
        for (int s = 0; s < 16; s++) a = __builtin_amdgcn_mfma_f32_32x32x2f32(f.lv[s], f.rv[s], a, 0, 0, 0);
    }
    __device__ void mac(const float *L, const float *R, int, int, int) {
        Frag f;
        load(L, R, f);
        mma(f);
    }
    __device__ double maxmag() const {
        float m = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; v++) m = fmaxf(m, fabsf(a[v]));
        return (double) m;
    }
    __device__ static int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
    __device__ void store(float *C, int ld = 32) const {
        const int lane = lane_id();
#pragma unroll
        for (int v = 0; v < 16; v++) C[row(v, lane) * ld + (lane & 31)] = a[v];
    }
    __device__ void diag(float *dg) const {
        const int lane = lane_id();
#pragma unroll
        for (int v = 0; v < 16; v++)
            if (row(v, lane) == (lane & 31)) dg[lane & 31] = a[v];
    }
};

// MFMA engine, single real inputs rounded to bfloat16, FP32 accumulation:
// v_mfma_f32_32x32x16_bf16 takes 8 consecutive k per lane (lanes 0-31 one
// group of 8, lanes 32-63 the next); instruction q covers k = 16q .. 16q+15.
// Same output layout as the FP32 engine.  The operands come from bfloat16
// copies of the tiles made before the pass (k_eb_bf16): A's tile row-major,
// B's transposed, so a lane's 8 consecutive k of its row of A, or of its
// column of B, are one 16-byte load (2 KB per tile against the 4 KB of the
// single-precision tile).
struct MfmaAccBF16 : MfmaAccF32 {
    struct Frag {
        bf16x8 av[2], bv[2];
    };
    __device__ static void load(const __bf16 *L, const __bf16 *Rt, Frag &f) {
        const int lane = lane_id();
        const int i = lane & 31, h = lane >> 5;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int k0 = 16 * q + 8 * h;
            f.av[q] = gld(reinterpret_cast<const bf16x8 *>(L + i * 32 + k0));
            f.bv[q] = gld(reinterpret_cast<const bf16x8 *>(Rt + i * 32 + k0));
        }
    }
    // A tile staged in LDS: row r (64 bytes, 4 chunks) keeps chunk c at
    // c ^ ((r >> 1) & 3), so rows r .. r + 7 at one chunk fill one 128-byte
    // span of banks
    using TA = __bf16;
    static constexpr int kChunks = 4;
    __device__ static int swz(int r, int c) { return c ^ ((r >> 1) & 3); }
    __device__ static void load_s(lds_ptr<__bf16> Ls, const __bf16 *Rt, Frag &f) {
        const int lane = lane_id();
        const int i = lane & 31, h = lane >> 5;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            f.av[q] = *(lds_ptr<bf16x8>) (Ls + i * 32 + 8 * swz(i, 2 * q + h));
            f.bv[q] = gld(reinterpret_cast<const bf16x8 *>(Rt + i * 32 + 16 * q + 8 * h));
        }
    }
    __device__ void mma(const Frag &f) {
#pragma unroll
        for (int q = 0; q < 2; q++) a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.av[q], f.bv[q], a, 0, 0, 0);
    }
};

// bfloat16 copies of the stored tiles of a single-precision matrix with
// 32 x 32 tiles, for the bf16 engine: `rm` (if not null) row-major as
// stored, `tr` (if not null) transposed.  One block per block row, its
// stored slots in turn; each element rounded to nearest even, as a
// conversion inside the product would round it.  Slots past nnzb are not
// written (no pair reads them).
__global__ __launch_bounds__(256) void k_eb_bf16(int M, const float *V, const int *nnzb,
                                                 __bf16 *rm, __bf16 *tr, const int *skip) {
    // each wavefront converts tiles w, w + 4, ... of the row on its own: a
    // lane's 16 loads of a tile issue together, and the transpose goes
    // through the wavefront's own LDS tile (no block barrier per tile)
    __shared__ float t[4][32][33];
    if (skip && *skip) return;  // uniform
    const int ib = blockIdx.x;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int n = nnzb[ib];
    for (int jp = w; jp < n; jp += 4) {
        const size_t base = ((size_t) ib * M + jp) * 1024;
        float v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = V[base + lane + 64 * q];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = lane + 64 * q;
            if (rm) rm[base + e] = (__bf16) v[q];
            t[w][e >> 5][e & 31] = v[q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (tr)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int e = lane + 64 * q;
                tr[base + e] = (__bf16) t[w][e & 31][e >> 5];
            }
        // the tile is read before the wavefront's next one overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// MFMA engine, double real, 32 x 32 tiles as 2 x 2 tiles of 16 x 16:
// v_mfma_f64_16x16x4f64 takes a 16x4 slice of L (lane l: row l%16, k group
// l/16) and a 4x16 slice of R.  Step s pairs k = 8g + s for the four lane
// groups g, so a lane reads 8 contiguous elements of its row of L.  Output:
// lane l, register v holds row l/16 + 4v, column l%16 of its 16x16 tile.
struct MfmaAccF64 {
    static constexpr bool kPipelined = true;
    doublex4 a[2][2];
    struct Frag {
        double lv[2][8], rv[2][8];
    };
    __device__ void zero() {
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 4; v++) a[ti][tj][v] = 0.0;
    }
    __device__ static void load(const double *L, const double *R, Frag &f) {
        const int lane = lane_id();
        const int c = lane & 15, g = lane >> 4;
#pragma unroll
        for (int ti = 0; ti < 2; ti++) {
            const doublex2 *l2 = reinterpret_cast<const doublex2 *>(L + (ti * 16 + c) * 32 + g * 8);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const doublex2 t = gld(l2 + q);
                f.lv[ti][2 * q] = t[0];
                f.lv[ti][2 * q + 1] = t[1];
            }
        }
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int s = 0; s < 8; s++) f.rv[tj][s] = gld(R + (g * 8 + s) * 32 + tj * 16 + c);
    }
    // A tile staged in LDS: row r (256 bytes, 16 chunks) keeps chunk c at
    // c ^ (r & 7)
    using TA = double;
    // not staged: two blocks per CU leave room for 5 FP64 tiles, and the
    // second load path costs the pipelined loop its last registers
    static constexpr bool kStage = false;
    static constexpr int kChunks = 16;
    __device__ static int swz(int r, int c) { return c ^ (r & 7); }
    __device__ static void load_s(lds_ptr<double> Ls, const double *R, Frag &f) {
        const int lane = lane_id();
        const int c = lane & 15, g = lane >> 4;
#pragma unroll
        for (int ti = 0; ti < 2; ti++) {
            const int r = ti * 16 + c;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const doublex2 t = *(lds_ptr<doublex2>) (Ls + r * 32 + 2 * swz(r, 4 * g + q));
                f.lv[ti][2 * q] = t[0];
                f.lv[ti][2 * q + 1] = t[1];
            }
        }
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int s = 0; s < 8; s++) f.rv[tj][s] = gld(R + (g * 8 + s) * 32 + tj * 16 + c);
    }
    __device__ void mma(const Frag &f) {
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int ti = 0; ti < 2; ti++)
#pragma unroll
                for (int tj = 0; tj < 2; tj++)
                    a[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.lv[ti][s], f.rv[tj][s],
                                                                     a[ti][tj], 0, 0, 0);
    }
    __device__ void mac(const double *L, const double *R, int, int, int) {
        Frag f;
        load(L, R, f);
        mma(f);
    }
    __device__ double maxmag() const {
        double m = 0.0;
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 4; v++) m = fmax(m, fabs(a[ti][tj][v]));
        return m;
    }
    __device__ void store(double *C, int ld = 32) const {
        const int lane = lane_id();
        const int c = lane & 15, g = lane >> 4;
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 4; v++) C[(ti * 16 + g + 4 * v) * ld + tj * 16 + c] = a[ti][tj][v];
    }
    __device__ void diag(double *dg) const {
        const int lane = lane_id();
        const int c = lane & 15, g = lane >> 4;
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (g + 4 * v == c) dg[t * 16 + c] = a[t][t][v];
    }
};

template <typename T, int ENGINE, int BT> struct AccOf { using type = ExactAcc<T, BT>; };
template <> struct AccOf<float, 1, 32> { using type = MfmaAccF32; };
template <> struct AccOf<float, 2, 32> { using type = MfmaAccBF16; };
template <> struct AccOf<double, 1, 32> { using type = MfmaAccF64; };

// lane src's value (src wave-uniform): two scalar reads of the halves
__device__ __forceinline__ size_t lane_get(size_t v, int src) {
    const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) v, src);
    const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (v >> 32), src);
    return ((size_t) hi << 32) | lo;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// ---------------------------------------------------------------------------
// numeric pass
//
// Phase 1: the wavefronts of a block row claim slots one at a time from an
// LDS counter (no barrier between slots: a slot's pair count ranges from 1 to
// the row's nnzb, and a round of equal shares left a third of the waves idle
// on a lattice).  Each forms its slot's tile, takes the max |element| and
// stores a kept tile at row position s, which is free as long as the row
// does not overflow (touched blocks <= MC).  Phase 2 (compact_kept) moves
// the kept tiles to their ranks in slot order.  Every tile is formed by one
// wavefront in a fixed order, so the result does not depend on which
// wavefront claims which slot.
//
// With p.sym (x2 of an exactly symmetric X, so X^2 is symmetric): row ib
// forms only the slots it owns (sym_owner).  A tile is staged in LDS and stored twice:
// at (ib, s) and, transposed, at (kb, s') where s' is the slot of ib in row
// kb; the kept flag goes to both slots.  The diagonal tile takes its upper
// triangle for both halves, so X^2 is exactly symmetric.  The compaction
// then runs in k_eb_compact, once every row's mirrors have landed.

// Move each kept slot's tile (flags in kbits, slots [0, nbits)) of row ib to
// its rank among the kept slots.  rank(s) <= s, so moving the tiles in
// increasing s, NUM_WAVES at a time with a barrier between a group's reads
// and its writes, never overwrites a tile not yet moved.  Called by all
// NUM_WAVES wavefronts of the block.
template <typename T, int BT>
__device__ void compact_kept(const EbParams &p, int ib, int nbits, const unsigned *kbits) {
    constexpr size_t B2 = (size_t) BT * BT;
    constexpr int Q = (int) ((B2 + 63) / 64);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    T *Cv = (T *) p.Cv;
    const int *skb = p.skb + (size_t) ib * p.maxs;
    int running = 0;
    for (int c0 = 0; c0 < nbits; c0 += 64) {
        const int sl = c0 + lane;
        const bool kp = sl < nbits && ((kbits[sl >> 5] >> (sl & 31)) & 1u);
        const uint64_t km = ballot64(kp);
        const int pos = running + popc64(km & ltmask);
        running += popc64(km);
        const bool mv = kp && pos != sl;
        const uint64_t mm = ballot64(mv);
        const int rank = popc64(mm & ltmask);
        const int nmv = popc64(mm);
        for (int g0 = 0; g0 < nmv; g0 += NUM_WAVES) {
            const uint64_t mine = ballot64(mv && rank == g0 + w);
            T buf[Q];
            int src = -1, dst = -1;
            if (mine) {
                const int l = __ffsll((unsigned long long) mine) - 1;
                src = c0 + l;
                dst = __shfl(pos, l, 64);
                const T *S = Cv + ((size_t) ib * p.MC + src) * B2;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const size_t e = (size_t) lane + 64 * q;
                    if (e < B2) buf[q] = S[e];
                }
            }
            __syncthreads();
            if (mine) {
                T *D = Cv + ((size_t) ib * p.MC + dst) * B2;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const size_t e = (size_t) lane + 64 * q;
                    if (e < B2) D[e] = buf[q];
                }
                if (lane == 0) p.Ci[(size_t) ib * p.MC + dst] = skb[src];
            }
        }
    }
}

// Slot of block column col among the slots of block row r (-1: none).  The
// first 3 chunks of r's slot list are loaded by mirror_probe when a slot
// starts, so their latency hides behind the slot's products; mirror_slot
// then matches them, and scans any further chunks itself.
struct MirrorProbe {
    int n;
    int v[3];
};
__device__ __forceinline__ MirrorProbe mirror_probe(const EbParams &p, int r) {
    MirrorProbe m;
    const int *skb = p.skb + (size_t) r * p.maxs;
    m.n = p.cnt[r];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int q = 64 * c + lane_id();
        m.v[c] = q < p.maxs ? skb[q] : -1;
    }
    return m;
}
__device__ __forceinline__ int mirror_slot(const EbParams &p, const MirrorProbe &m, int r, int col) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const uint64_t b = ballot64(64 * c + lane_id() < m.n && m.v[c] == col);
        if (b) return 64 * c + __ffsll((unsigned long long) b) - 1;
    }
    const int *skb = p.skb + (size_t) r * p.maxs;
    for (int c0 = 192; c0 < m.n; c0 += 64) {
        const int q = c0 + lane_id();
        const uint64_t b = ballot64(q < m.n && skb[q] == col);
        if (b) return c0 + __ffsll((unsigned long long) b) - 1;
    }
    return -1;
}

// blocks per CU the registers are held to: 3 for the bf16 engine (its
// 48 KB LDS share leaves room for 3; FP32 would spill 94 registers), else 2
template <typename T, int ENGINE> struct EbOcc { static constexpr int value = 2; };
template <> struct EbOcc<float, 2> { static constexpr int value = BMLHIP_EB_OCC_BF16; };
template <typename T, int ENGINE, int BT>
__global__ __launch_bounds__(64 * NUM_WAVES, (EbOcc<T, ENGINE>::value)) void k_eb_numeric(EbParams p) {
    using E = Elem<T>;
    using Acc = typename AccOf<T, ENGINE, BT>::type;
    // dynamic LDS: the kept flag of each slot s < min(maxs, MC), then (sym)
    // one staging tile per wavefront, rows BT + 1 elements apart
    extern __shared__ __align__(16) unsigned kbits[];
    __shared__ int next_s, kept_s;
    __shared__ T diag_s[NUM_WAVES][BT];
    if (p.skip && *p.skip) return;
    const int ib = p.remap ? xcd_block(blockIdx.x, gridDim.x) : (int) blockIdx.x;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int count = p.cnt[ib];
    const int bi = p.bsize[ib];
    const T *Av = (const T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    T *Cv = (T *) p.Cv;
    constexpr size_t B2 = (size_t) BT * BT;
    constexpr int LD = BT + 1;
    const int *off = p.soff + (size_t) ib * (p.maxs + 1);
    const int *pr = p.pairs + (size_t) ib * p.maxp;
    const int *skb = p.skb + (size_t) ib * p.maxs;
    const int nbits = count < p.MC ? count : p.MC;  // slots with a row position
    const int mbits = p.maxs < p.MC ? p.maxs : p.MC;
    T *stg0 = (T *) (kbits + ((mbits + 31) / 32 + 4) / 4 * 4);
    T *stg = stg0 + (size_t) w * BT * LD;
    for (int q = threadIdx.x; q < (nbits + 31) / 32; q += 64 * NUM_WAVES) kbits[q] = 0u;
    // MFMA engines: the row's first nst A tiles, copied once into LDS (after
    // the staging tiles), each row's 16-byte chunks permuted by Acc::swz;
    // every pair with jp < nst then reads its left operand from LDS (a tile
    // is used by about as many pairs as its row jb holds blocks)
    int nst = 0;
    typedef __attribute__((address_space(3))) unsigned char lds_u8;
    lds_u8 *ast = (lds_u8 *) (stg0 + (p.sym ? (size_t) NUM_WAVES * BT * LD : 0));
    // fused update: A's block columns of this row in LDS, alpha and beta
    typedef __attribute__((address_space(3))) int lds_int;
    lds_int *acol = (lds_int *) ((unsigned char *) kbits + p.acol_off);
    const int nAr = p.An[ib];
    // fused: this lane's |X - X^2|^2, and the next iterate's sum |U_ij|^2 and
    // trace over the kept blocks (the next iteration's tr X^2 and tr X)
    double alpha = 0.0, beta = 0.0, nrm_l = 0.0, frob_l = 0.0, trn_l = 0.0;
    if (p.fuse) {
        alpha = p.ab[0];
        beta = p.ab[1];
        for (int q = threadIdx.x; q < nAr; q += 64 * NUM_WAVES) acol[q] = p.Ai[(size_t) ib * p.MA + q];
    }
    if constexpr (Acc::kPipelined && Acc::kStage) {
        using TA = typename Acc::TA;
        constexpr int CH = Acc::kChunks, TE = 16 / (int) sizeof(TA);
        const TA *src = (const TA *) (ENGINE == 2 ? (const void *) p.A16 : p.Av);
        nst = p.An[ib] < p.ka ? p.An[ib] : p.ka;
        for (int e = threadIdx.x; e < nst * BT * CH; e += 64 * NUM_WAVES) {
            const int t = e / (BT * CH), r = e / CH % BT, c = e % CH;
            const floatx4 v = *reinterpret_cast<const floatx4 *>(src + ((size_t) ib * p.MA + t) * B2 + r * BT + c * TE);
            *(__attribute__((address_space(3))) floatx4 *) (ast + sizeof(TA) * (t * (int) B2 + r * BT + Acc::swz(r, c) * TE)) = v;
        }
    }
    if (threadIdx.x == 0) {
        next_s = 0;
        kept_s = 0;
    }
    T tx = E::zero(), tx2 = E::zero();
    if (p.kind == 0 && !p.fuse && threadIdx.x == 0) {
        // tr X: the diagonal of every diagonal block met in the expand phase
        for (int jp = 0; jp < p.An[ib]; jp++) {
            const size_t sa = (size_t) ib * p.MA + jp;
            if (p.Ai[sa] == ib)
                for (int kk = 0; kk < bi; kk++) tx = E::add(tx, Av[sa * B2 + kk * BT + kk]);
        }
    }
    __syncthreads();
    // the next slot this wavefront forms (-1: none left); with p.sym the
    // slots row ib does not own are skipped (mirrors).  Every formed slot has
    // at least one pair: a slot exists because a pair touched it.
    auto claim = [&]() -> int {
        while (true) {
            int c = 0;
            if (lane == 0) c = atomicAdd(&next_s, 1);
            c = __shfl(c, 0, 64);
            if (c >= count) return -1;
            if (!(p.sym && !sym_owner(ib, skb[c]))) return c;
        }
    };
    // keep rule, tr X^2 and the stores of slot s's finished tile
    auto finish_slot = [&](const Acc &acc, int s, int kb, const MirrorProbe &mp) __attribute__((always_inline)) {
        const double nrm = wave_max(acc.maxmag());
        const bool keep = kb == ib || nrm > p.thr;
        if (p.kind == 0 && !p.fuse && kb == ib) {
            // tr X^2: the diagonal block's diagonal, summed in order by lane 0
            acc.diag(diag_s[w]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0)
                for (int kk = 0; kk < bi; kk++) tx2 = E::add(tx2, diag_s[w][kk]);
        }
        if (!p.sym) {
            if (keep) {
                if (s < p.MC) {
                    const size_t sc = (size_t) ib * p.MC + s;
                    acc.store(Cv + sc * B2);
                    if (lane == 0) {
                        p.Ci[sc] = kb;
                        atomicOr(&kbits[s >> 5], 1u << (s & 31));
                    }
                }
                if (lane == 0) atomicAdd(&kept_s, 1);
            }
            return;
        }
        // sym: the tile at (ib, s) and its transpose at (kb, s')
        const int sp = kb != ib ? mirror_slot(p, mp, kb, ib) : -1;
        auto wave_sync = [&]() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        bool kept = keep;
        if (p.fuse) {
            // U = alpha X[ib, kb] + beta Y, Y the product's block when the x2
            // rule keeps it (else 0), formed in the staging tile; U kept iff
            // it is the diagonal block or max |U| > thr (the add's rule;
            // every X block is reached through its row's diagonal block, so
            // none is left out); |X - Y|^2 counted twice off the diagonal
            // (the mirror block)
            acc.store(stg, LD);
            wave_sync();
            if (kb == ib) {  // the diagonal block takes its upper triangle
                for (int e = lane; e < (int) B2; e += 64) {
                    const int r = e / BT, c = e % BT;
                    if (r > c) stg[r * LD + c] = E::conj(stg[c * LD + r]);
                }
                wave_sync();
            }
            int jx = -1;
            for (int c0 = 0; c0 < nAr; c0 += 64) {
                const uint64_t b = ballot64(c0 + lane < nAr && acol[c0 + lane] == kb);
                if (b) {
                    jx = c0 + __ffsll((unsigned long long) b) - 1;
                    break;
                }
            }
            // X's block in chunks of a lane's elements (all of them for 4-byte
            // elements, two at a time for 8-byte ones, whose kernel has no
            // registers to spare), each chunk's loads issued before any is
            // used (from the LDS stage when it holds the exact tile)
            constexpr int QE = (int) (B2 / 64);
            static_assert(B2 % 64 == 0, "a tile is whole rows of 64 lanes");
            constexpr int CH = sizeof(T) <= 4 ? QE : (QE >= 2 ? 2 : QE);
            bool staged = false;
            if constexpr (Acc::kStage) {
                using TA = typename Acc::TA;
                if constexpr (std::is_same<TA, T>::value) staged = jx >= 0 && jx < nst;
            }
            const T *xg = Av + ((size_t) ib * p.MA + (jx >= 0 ? jx : 0)) * B2;
            double mu = 0.0, dn = 0.0, fu = 0.0, tu = 0.0;
#pragma unroll 1
            for (int q0 = 0; q0 < QE; q0 += CH) {
                T xv[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    const int e = lane + 64 * (q0 + j);
                    xv[j] = E::zero();
                    if constexpr (Acc::kStage) {
                        using TA = typename Acc::TA;
                        if constexpr (std::is_same<TA, T>::value) {
                            constexpr int TE = 16 / (int) sizeof(TA);
                            const int r = e / BT, c = e % BT;
                            if (staged)
                                xv[j] = *((lds_ptr<TA>) ast + jx * (int) B2 + r * BT +
                                          Acc::swz(r, c / TE) * TE + c % TE);
                        }
                    }
                    if (!staged && jx >= 0) xv[j] = xg[e];
                }
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    const int e = lane + 64 * (q0 + j), r = e / BT, c = e % BT;
                    const T y = keep ? stg[r * LD + c] : E::zero();
                    const T u = E::axpy(E::axpy(E::zero(), alpha, xv[j]), beta, y);
                    stg[r * LD + c] = u;
                    mu = fmax(mu, E::mag(u));
                    dn += E::abs2(E::sub(xv[j], y));
                    fu += E::abs2(u);
                    if (r == c) tu += E::re(u);
                }
            }
            nrm_l += kb == ib ? dn : 2.0 * dn;
            kept = kb == ib || wave_max(mu) > p.thr;
            if (kept && (s < p.MC || (sp >= 0 && sp < p.MC))) {
                frob_l += kb == ib ? fu : 2.0 * fu;
                if (kb == ib) trn_l += tu;
            }
            wave_sync();
        }
        if (lane == 0) {
            p.kflag[(size_t) ib * p.maxs + s] = kept;
            if (sp >= 0) p.kflag[(size_t) kb * p.maxs + sp] = kept;
        }
        if (kept && (s < p.MC || (sp >= 0 && sp < p.MC))) {
            if (!p.fuse) {
                acc.store(stg, LD);
                wave_sync();
            }
            if (s < p.MC) {
                const size_t sc = (size_t) ib * p.MC + s;
                T *D = Cv + sc * B2;
                for (int e = lane; e < (int) B2; e += 64) {
                    const int r = e / BT, c = e % BT;
                    D[e] = (kb == ib && r > c) ? E::conj(stg[c * LD + r]) : stg[r * LD + c];
                }
                if (lane == 0) p.Ci[sc] = kb;
            }
            if (sp >= 0 && sp < p.MC) {
                const size_t sc = (size_t) kb * p.MC + sp;
                T *D = Cv + sc * B2;
                for (int e = lane; e < (int) B2; e += 64) {
                    const int r = e / BT, c = e % BT;
                    D[e] = E::conj(stg[c * LD + r]);
                }
                if (lane == 0) p.Ci[sc] = ib;
            }
            // the staging tile is read before the next slot overwrites it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    };
    if constexpr (Acc::kPipelined) {
        // One stream of pairs per wavefront.  It claims batches (about
        // EB_BATCH pairs of whole slots) from an LDS counter; a unit is up to
        // 64 pairs of a batch, and lane q holds pair t0 + q's A tile, B tile,
        // slot and block column.  The next unit is decoded as soon as a unit
        // starts, and the next pair's tiles load while the current pair's
        // MFMAs run, across slots and units, so a slot's keep test and stores
        // and the index loads behind the next tiles overlap the MFMAs.
        using Frag = typename Acc::Frag;
        const int *ps = p.pslot + (size_t) ib * p.maxp;
        const int *bf = p.bfirst + (size_t) ib * (p.maxp / EB_BATCH + 2);
        const int nbatch = off[count] / EB_BATCH + 1;
        struct Unit {
            int t0, te, np;
            int sa, bt, sid, kb;  // per lane
        };
        auto decode = [&](Unit &u) {
            u.np = u.te - u.t0 < 64 ? u.te - u.t0 : 64;
            u.sa = 0;
            u.bt = 0;
            u.sid = -1;
            u.kb = -1;
            if (lane < u.np) {
                const int t = u.t0 + lane;
                const unsigned pv = (unsigned) pr[t];
                u.sid = ps[t];
                u.sa = ib * p.MA + (int) (pv >> 16);
                u.bt = p.Ai[u.sa] * p.MB + (int) (pv & 0xffffu);
                u.kb = skb[u.sid];
            }
        };
        // the unit after u (u.np == 0: the first); false when no batch is left
        auto next_unit = [&](Unit &n, const Unit &u) -> bool {
            if (u.np > 0 && u.t0 + u.np < u.te) {
                n.t0 = u.t0 + u.np;
                n.te = u.te;
            } else {
                while (true) {
                    int k = 0;
                    if (lane == 0) k = atomicAdd(&next_s, 1);
                    k = __shfl(k, 0, 64);
                    if (k >= nbatch) return false;
                    n.t0 = off[bf[k]];
                    n.te = off[bf[k + 1]];
                    if (n.te > n.t0) break;
                }
            }
            decode(n);
            return true;
        };
        auto tile_a = [&](const Unit &u, int q) {
            const size_t t = (size_t) __builtin_amdgcn_readlane(u.sa, q) * B2;
            if constexpr (ENGINE == 2) return p.A16 + t;
            else return Av + t;
        };
        auto tile_b = [&](const Unit &u, int q) {
            const size_t t = (size_t) __builtin_amdgcn_readlane(u.bt, q) * B2;
            if constexpr (ENGINE == 2) return p.B16 + t;
            else return Bv + t;
        };
        // pair q of unit u into f: the left tile from LDS when staged
        auto ld = [&](const Unit &u, int q, Frag &f) {
            if constexpr (Acc::kStage) {
                const int jp = __builtin_amdgcn_readlane(u.sa, q) - ib * p.MA;
                if (jp < nst) {
                    Acc::load_s((lds_ptr<typename Acc::TA>) ast + jp * (int) B2, tile_b(u, q), f);
                    return;
                }
            }
            Acc::load(tile_a(u, q), tile_b(u, q), f);
        };
        Unit cu, nu;
        cu.np = 0;
        if (next_unit(cu, cu)) {
            bool have_nu = next_unit(nu, cu);
            Acc acc;
            acc.zero();
            Frag f0, f1;
            int q = 0, slot = -1, kb = -1;
            MirrorProbe mp;
            ld(cu, 0, f0);
            // run pair q of cu (in cur) while the pair after it loads into nxt
            auto step = [&](Frag &cur, Frag &nxt) -> bool {
                if (q + 1 < cu.np)
                    ld(cu, q + 1, nxt);
                else if (have_nu)
                    ld(nu, 0, nxt);
                const int sq = __builtin_amdgcn_readlane(cu.sid, q);
                if (sq != slot) {
                    if (slot >= 0) {
                        finish_slot(acc, slot, kb, mp);
                        acc.zero();
                    }
                    slot = sq;
                    kb = __builtin_amdgcn_readlane(cu.kb, q);
                    if (p.sym && kb != ib) mp = mirror_probe(p, kb);
                }
                acc.mma(cur);
                if (q + 1 < cu.np) {
                    q++;
                    return true;
                }
                if (!have_nu) {
                    finish_slot(acc, slot, kb, mp);
                    return false;
                }
                cu = nu;
                q = 0;
                have_nu = next_unit(nu, cu);
                return true;
            };
            while (step(f0, f1) && step(f1, f0)) {
            }
        }
    } else {
        while (true) {
            const int s = claim();
            if (s < 0) break;
            const int kb = skb[s];
            const int bj = p.bsize[kb];
            Acc acc;
            acc.zero();
            for (int t = off[s]; t < off[s + 1]; t++) {
                const int jp = (int) ((unsigned) pr[t] >> 16), kp = pr[t] & 0xffff;
                const size_t sa = (size_t) ib * p.MA + jp;
                const int jb = p.Ai[sa];
                acc.mac(Av + sa * B2, Bv + ((size_t) jb * p.MB + kp) * B2, bi, p.bsize[jb], bj);
            }
            MirrorProbe mp;
            if (p.sym && kb != ib) mp = mirror_probe(p, kb);
            finish_slot(acc, s, kb, mp);
        }
    }
    __syncthreads();
    if (!p.sym) compact_kept<T, BT>(p, ib, nbits, kbits);
    // tx2 lives in lane 0 of the wavefront that formed the diagonal block
    __shared__ double tr2_s[NUM_WAVES];
    if (p.fuse) {
        __shared__ double fz_s[3][NUM_WAVES];
        for (int off = 32; off > 0; off >>= 1) {
            nrm_l += __shfl_xor(nrm_l, off, 64);
            frob_l += __shfl_xor(frob_l, off, 64);
            trn_l += __shfl_xor(trn_l, off, 64);
        }
        if (lane == 0) {
            fz_s[0][w] = trn_l;
            fz_s[1][w] = frob_l;
            fz_s[2][w] = nrm_l;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            double t = 0.0;
            for (int v = 0; v < NUM_WAVES; v++) t += fz_s[threadIdx.x][v];
            p.rownorm[(size_t) threadIdx.x * p.NB + ib] = t;
        }
        return;
    }
    if (lane == 0) tr2_s[w] = E::re(tx2);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!p.sym) {
            const int kept = kept_s;
            p.Cn[ib] = kept < p.MC ? kept : p.MC;
            p.rowstat[ib] = make_int4(kept, count, off[count], p.An[ib]);
        }
        if (p.kind == 0) {
            double t2 = 0.0;
            for (int v = 0; v < NUM_WAVES; v++) t2 += tr2_s[v];
            p.rowtr[ib] = E::re(tx);
            p.rowtr[p.NB + ib] = t2;
        }
    }
}

// sym: after every row's tiles and mirrors are stored, each block row
// gathers its kept flags (its own slots and the mirrored ones), compacts the
// kept tiles into slot order and records its counts.
template <typename T, int BT>
__global__ __launch_bounds__(64) void k_eb_compact(EbParams p) {
    // one wavefront per block row.  The kept slots below MC are listed with
    // their ranks (src -> dst, dst < src) in LDS, then the tiles that move
    // are copied in increasing src order, each tile's loads issued before
    // the previous tile's stores: a destination lies below its own source
    // and every later one, so no load reads a slot an earlier store wrote,
    // and a store waits for its own tile's loads, which return in order
    // after those of every earlier tile (whose sources it may overwrite).
    extern __shared__ __align__(16) int mv[];  // pairs (src, dst), at most MC
    if (p.skip && *p.skip) return;
    const int ib = blockIdx.x, lane = threadIdx.x;
    constexpr size_t B2 = (size_t) BT * BT;
    constexpr int Q = (int) ((B2 + 63) / 64);
    const int count = p.cnt[ib];
    const int nbits = count < p.MC ? count : p.MC;
    const int *fl = p.kflag + (size_t) ib * p.maxs;
    const int *skb = p.skb + (size_t) ib * p.maxs;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    int kept = 0, nmv = 0;
    for (int c0 = 0; c0 < count; c0 += 64) {
        const int sl = c0 + lane;
        const bool kp = sl < count && fl[sl] != 0;
        const uint64_t km = ballot64(kp);
        const int pos = kept + popc64(km & ltmask);
        const bool m = kp && sl < nbits && pos != sl;
        const uint64_t mm = ballot64(m);
        if (m) {
            const int r = nmv + popc64(mm & ltmask);
            mv[2 * r] = sl;
            mv[2 * r + 1] = pos;
        }
        kept += popc64(km);
        nmv += popc64(mm);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    T *Cv = (T *) p.Cv;
    T *row = Cv + (size_t) ib * p.MC * B2;
    auto load = [&](int k, T (&b)[Q]) {
        const T *S = row + (size_t) mv[2 * k] * B2;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const size_t e = (size_t) lane + 64 * q;
            if (e < B2) b[q] = S[e];
        }
    };
    auto store = [&](int k, const T (&b)[Q]) {
        const int src = mv[2 * k], dst = mv[2 * k + 1];
        T *D = row + (size_t) dst * B2;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const size_t e = (size_t) lane + 64 * q;
            if (e < B2) D[e] = b[q];
        }
        if (lane == 0) p.Ci[(size_t) ib * p.MC + dst] = skb[src];
    };
    T b0[Q], b1[Q];
    if (nmv > 0) load(0, b0);
    for (int k = 0; k < nmv; k += 2) {
        if (k + 1 < nmv) load(k + 1, b1);
        store(k, b0);
        if (k + 1 >= nmv) break;
        if (k + 2 < nmv) load(k + 2, b0);
        store(k + 1, b1);
    }
    if (lane == 0) {
        const int *off = p.soff + (size_t) ib * (p.maxs + 1);
        p.Cn[ib] = kept < p.MC ? kept : p.MC;
        p.rowstat[ib] = make_int4(kept, count, off[count], p.An[ib]);
    }
}

// ---------------------------------------------------------------------------
// add / add with norm: A <- alpha*A + beta*B, one wavefront per block row.
// Union order: A's block columns, then B's new ones
// (ellblock/bml_add_ellblock_typed.c:35-197); in mode 0 A is skipped when
// alpha == 0 and B when beta == 0.  Block kept iff its max |element|, held in
// T, is above the threshold.  Mode 2 also sums |A - B|^2 over the union into
// rownorm.
struct EbAddParams {
    int NB, bt, MA, MB, mode;
    const int *bsize;
    void *Av;
    int *Ai, *An;
    const void *Bv;
    const int *Bi, *Bn;
    double alpha, beta, thr;
    const double *ab;
    const int *skip;
    double *rownorm;
    int4 *rowstat;  // [NB] (kept, 0, 0, 0)
    bmlhip_status *st;
};

// One block row per block: the block lists the union (A's blocks, then B's
// new ones, as bml_add_ellblock_typed.c:35-197), then EB_ADD_WAVES
// wavefronts take consecutive union entries, one each per round: every
// wavefront reads its two tiles and forms alpha*A + beta*B with the block's
// max norm, the round's keep flags are ranked in union order, and the kept
// tiles are written in place into A.  A tile is read before any write of
// its round, and writes only reach slots whose A tiles were read in earlier
// rounds or this one.
#ifndef BMLHIP_EB_ADD_WAVES
#define BMLHIP_EB_ADD_WAVES 4
#endif
constexpr int EB_ADD_WAVES = BMLHIP_EB_ADD_WAVES;
// union entries per wavefront per round: 2 (more loads in flight, 178
// VGPRs in FP64) ran the FP64 SP2 solve of the 16x16x8-site, 32-orbital
// lattice at 22.3 solves/s against 25.3 with 1
#ifndef BMLHIP_EB_ADD_PER
#define BMLHIP_EB_ADD_PER 1
#endif
constexpr int EB_ADD_PER = BMLHIP_EB_ADD_PER;

template <typename T, int BT>
__global__ __launch_bounds__(64 * EB_ADD_WAVES) void k_eb_add(EbAddParams p) {
    using E = Elem<T>;
    extern __shared__ int usm[];
    __shared__ int keep_s[2][EB_ADD_WAVES * EB_ADD_PER];  // by round parity
    __shared__ double nrm_s[EB_ADD_WAVES];
    if (p.skip && *p.skip) return;
    const int ib = blockIdx.x;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int cap = p.MA + p.MB;
    int *ucol = usm, *ua = usm + cap, *ub = usm + 2 * cap;
    const double alpha = p.ab ? p.ab[0] : p.alpha;
    const double beta = p.ab ? p.ab[1] : p.beta;
    const bool useA = p.mode != 0 || alpha > 0.0 || alpha < 0.0;
    const bool useB = p.mode != 0 || beta > 0.0 || beta < 0.0;
    const int nA = useA ? p.An[ib] : 0, nB = useB ? p.Bn[ib] : 0;
    T *Av = (T *) p.Av;
    const T *Bv = (const T *) p.Bv;
    constexpr size_t B2 = (size_t) BT * BT;
    // union (rows hold at most MA + MB block columns): A's list copied,
    // then each thread takes one of B's entries, looks for its column in A's
    // list (an LDS scan, every lane reading the same word) and the columns A
    // lacks are appended in B's order, ranked by a block-wide prefix count
    __shared__ int wcnt_s[EB_ADD_WAVES];
    const uint64_t ltmask = (1ull << lane) - 1ull;
    for (int jp = threadIdx.x; jp < nA; jp += 64 * EB_ADD_WAVES) {
        ucol[jp] = p.Ai[(size_t) ib * p.MA + jp];
        ua[jp] = jp;
        ub[jp] = -1;
    }
    __syncthreads();
    int n = nA;  // block-uniform
    for (int k0 = 0; k0 < nB; k0 += 64 * EB_ADD_WAVES) {
        const int kp = k0 + (int) threadIdx.x;
        int kb = -1, u = -1;
        if (kp < nB) {
            kb = p.Bi[(size_t) ib * p.MB + kp];
            for (int v = 0; v < nA; v++)
                if (ucol[v] == kb) {
                    u = v;
                    break;
                }
        }
        const bool isnew = kp < nB && u < 0;
        const uint64_t m = ballot64(isnew);
        if (lane == 0) wcnt_s[w] = popc64(m);
        if (u >= 0) ub[u] = kp;
        __syncthreads();
        int pos = n, tot = 0;
        for (int q = 0; q < EB_ADD_WAVES; q++) {
            pos += q < w ? wcnt_s[q] : 0;
            tot += wcnt_s[q];
        }
        if (isnew) {
            pos += popc64(m & ltmask);
            ucol[pos] = kb;
            ua[pos] = -1;
            ub[pos] = kp;
        }
        n += tot;
        __syncthreads();  // wcnt_s is reused; the list is complete after the last chunk
    }
    const int bi = p.bsize[ib];
    static_assert(BT * BT % 64 == 0, "a tile is whole rows of 64 lanes");
    constexpr int Q = BT * BT / 64;
    int ll = 0;
    double nrm2 = 0.0;
    // each round: EB_ADD_PER union entries per wavefront, entry u0 + r with
    // r = k EB_ADD_WAVES + w; all their tiles are read (and loads of the
    // round's entries overlap) before any kept tile is written.  A kept
    // tile of entry u goes to position <= u, below every A slot a later
    // round reads, so a wavefront starts its next round's reads as soon as
    // its own writes are issued (the keep flags alternate between two
    // buffers, so the next round's flags do not overwrite unread ones)
    int par = 0;
    for (int u0 = 0; u0 < n; u0 += EB_ADD_WAVES * EB_ADD_PER, par ^= 1) {
        T x[EB_ADD_PER][Q];
        bool keep[EB_ADD_PER];
        int jbk[EB_ADD_PER];
#pragma unroll
        for (int k = 0; k < EB_ADD_PER; k++) {
            const int u = u0 + k * EB_ADD_WAVES + w;
            const bool live = u < n;
            const int jb = live ? ucol[u] : 0, sa = live ? ua[u] : -1, sb = live ? ub[u] : -1;
            const int bj = p.bsize[jb];
            jbk[k] = jb;
            double m = 0.0;
            // both tiles' loads issued together (sa, sb are wave-uniform; BT^2
            // is a multiple of 64), the block's bsize[ib] x bsize[jb] corner
            // selected afterwards
            T av[Q], bv[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                av[q] = E::zero();
                bv[q] = E::zero();
            }
            if (sa >= 0) {
                const T *ta = Av + ((size_t) ib * p.MA + sa) * B2;
#pragma unroll
                for (int q = 0; q < Q; q++) av[q] = ta[lane + 64 * q];
            }
            if (sb >= 0) {
                const T *tb = Bv + ((size_t) ib * p.MB + sb) * B2;
#pragma unroll
                for (int q = 0; q < Q; q++) bv[q] = tb[lane + 64 * q];
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int e = lane + 64 * q;
                x[k][q] = E::zero();
                T y = E::zero();
                if (e / BT < bi && e % BT < bj) {
                    if (sa >= 0) {
                        x[k][q] = E::axpy(x[k][q], alpha, av[q]);
                        y = av[q];
                    }
                    if (sb >= 0) {
                        x[k][q] = E::axpy(x[k][q], beta, bv[q]);
                        y = E::sub(y, bv[q]);
                    }
                }
                const double mg = E::mag(x[k][q]);
                m = mg > m ? mg : m;
                if (p.mode == 2) nrm2 += E::abs2(y);
            }
            m = wave_max(m);
            keep[k] = live && E::above(E::from_real(m), p.thr);
            if (lane == 0) keep_s[par][k * EB_ADD_WAVES + w] = keep[k] ? 1 : 0;
        }
        __syncthreads();  // every tile of the round read; keep flags in
        int tot = 0;
        for (int r = 0; r < EB_ADD_WAVES * EB_ADD_PER; r++) tot += keep_s[par][r];
#pragma unroll
        for (int k = 0; k < EB_ADD_PER; k++) {
            if (!keep[k]) continue;
            int pos = ll;
            for (int r = 0; r < k * EB_ADD_WAVES + w; r++) pos += keep_s[par][r];
            if (pos < p.MA) {
                const size_t d = ((size_t) ib * p.MA + pos) * B2;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const int e = lane + 64 * q;
                    if (e < BT * BT) Av[d + e] = x[k][q];
                }
                if (lane == 0) p.Ai[(size_t) ib * p.MA + pos] = jbk[k];
            } else if (lane == 0 && pos == p.MA) {
                atomicAdd(&p.st->m_overflow_rows, 1);
                atomicMin(&p.st->first_m_overflow, ib);
            }
        }
        ll += tot;
    }
    for (int off = 32; off > 0; off >>= 1) nrm2 += __shfl_down(nrm2, off, 64);
    if (lane == 0) nrm_s[w] = nrm2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int v = 0; v < EB_ADD_WAVES; v++) t += nrm_s[v];
        p.An[ib] = ll < p.MA ? ll : p.MA;
        p.rowstat[ib] = make_int4(ll, 0, 0, 0);
        if (p.mode == 2) p.rownorm[ib] = t;
    }
}

// threshold (ellblock/bml_threshold_ellblock_typed.c:95-140): drop a block
// whose max |element| (held in T) is not above thr, zero the elements of
// the others that are not above thr, keep the order.  One wavefront per
// block row.
template <typename T, int BT>
__global__ __launch_bounds__(64) void k_eb_threshold(int NB, int MB, const int *bsize, T *Av,
                                                     int *Ai, int *An, double thr) {
    using E = Elem<T>;
    const int ib = blockIdx.x, lane = threadIdx.x;
    constexpr size_t B2 = (size_t) BT * BT;
    constexpr int Q = (BT * BT + 63) / 64;
    const int bi = bsize[ib];
    const int n = An[ib];
    int ll = 0;
    for (int jp = 0; jp < n; jp++) {
        const size_t s = (size_t) ib * MB + jp;
        const int jb = Ai[s];
        const int bj = bsize[jb];
        T x[Q];
        double m = 0.0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = lane + 64 * q;
            x[q] = e < BT * BT ? Av[s * B2 + e] : E::zero();
            if (e < BT * BT && e / BT < bi && e % BT < bj) {
                const double mg = E::mag(x[q]);
                m = mg > m ? mg : m;
            }
        }
        m = wave_max(m);
        if (!E::above(E::from_real(m), thr)) continue;
        const size_t d = (size_t) ib * MB + ll;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = lane + 64 * q;
            if (e < BT * BT) {
                const T v = E::above(x[q], thr) ? x[q] : E::zero();
                Av[d * B2 + e] = v;
            }
        }
        if (lane == 0) Ai[d] = jb;
        ll++;
        __syncthreads();
    }
    if (lane == 0) An[ib] = ll;
}

// trace (ellblock/bml_trace_ellblock_typed.c:27-58): the diagonal of the
// first diagonal block of every block row, per block row in double.
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_eb_trace_rows(int NB, int MB, const int *bsize,
                                                       const T *Av, const int *Ai, const int *An,
                                                       double *rowtr) {
    const int ib = blockIdx.x * 256 + threadIdx.x;
    if (ib >= NB) return;
    constexpr size_t B2 = (size_t) BT * BT;
    double s = 0.0;
    for (int jp = 0; jp < An[ib]; jp++) {
        const size_t sl = (size_t) ib * MB + jp;
        if (Ai[sl] == ib) {
            for (int kk = 0; kk < bsize[ib]; kk++) s += Elem<T>::re(Av[sl * B2 + kk * BT + kk]);
            break;
        }
    }
    rowtr[ib] = s;
}

// SP2 with the update fused into the product: per block row, tr X (the
// diagonal of its diagonal block) and the sum of |X_ij|^2 over its blocks,
// which is the row's share of tr X^2 when X is symmetric (Hermitian); in
// double, four elements per thread per tile with a fixed reduction tree
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_eb_trace_frob(int NB, int MB, const T *Av, const int *Ai,
                                                       const int *An, double *rowtr,
                                                       const int *skip) {
    __shared__ double part[2][256];
    if (skip && *skip) return;  // uniform
    const int ib = blockIdx.x;
    constexpr int B2 = BT * BT;
    const int n = An[ib];
    double tr = 0.0, fr = 0.0;
    for (int jp = 0; jp < n; jp++) {
        const size_t sl = (size_t) ib * MB + jp;
        const T *t = Av + sl * B2;
        const bool dg = Ai[sl] == ib;
        T v[(B2 + 255) / 256];
#pragma unroll
        for (int q = 0; q < (B2 + 255) / 256; q++) {
            const int e = threadIdx.x + 256 * q;
            v[q] = e < B2 ? t[e] : Elem<T>::zero();
        }
#pragma unroll
        for (int q = 0; q < (B2 + 255) / 256; q++) {
            const int e = threadIdx.x + 256 * q;
            fr += Elem<T>::abs2(v[q]);
            if (dg && e < B2 && e / BT == e % BT) tr += Elem<T>::re(v[q]);
        }
    }
    part[0][threadIdx.x] = tr;
    part[1][threadIdx.x] = fr;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) {
            part[0][threadIdx.x] += part[0][threadIdx.x + w];
            part[1][threadIdx.x] += part[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rowtr[ib] = part[0][0];
        rowtr[NB + ib] = part[1][0];
    }
}

// the fused SP2 product's three per-row sums (next tr X, next tr X^2, and
// |X - X^2|^2) into status scalars 0, 1, 2: one block of 1024 threads, each
// sum over a fixed partition and a fixed tree (one launch, not three)
__global__ __launch_bounds__(1024) void k_eb_sum3(const double *x, int n, bmlhip_status *st,
                                                  const int *skip) {
    __shared__ double part[1024];
    if (skip && *skip) return;  // uniform
    for (int a = 0; a < 3; a++) {
        const double *v = x + (size_t) a * n;
        double s = 0.0;
        for (int k = threadIdx.x; k < n; k += 1024) s += v[k];
        part[threadIdx.x] = s;
        __syncthreads();
        for (int w = 512; w > 0; w >>= 1) {
            if ((int) threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) st->scalars[a] = part[0];
        __syncthreads();
    }
}

// copy block rows between matrices of different MB
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_eb_copy(int NB, int MA, const T *Av, const int *Ai,
                                                 const int *An, int MB, T *Bv, int *Bi, int *Bn,
                                                 bmlhip_status *st) {
    const int ib = blockIdx.x;
    constexpr size_t B2 = (size_t) BT * BT;
    const int n = An[ib];
    const int m = n < MB ? n : MB;
    for (size_t e = threadIdx.x; e < (size_t) m * B2; e += 256)
        Bv[(size_t) ib * MB * B2 + e] = Av[(size_t) ib * MA * B2 + e];
    for (int jp = threadIdx.x; jp < m; jp += 256) Bi[(size_t) ib * MB + jp] = Ai[(size_t) ib * MA + jp];
    if (threadIdx.x == 0) {
        Bn[ib] = m;
        if (n > MB) {
            atomicAdd(&st->m_overflow_rows, 1);
            atomicMin(&st->first_m_overflow, ib);
        }
    }
}

// Gershgorin rows (ellblock/bml_normalize_ellblock_typed.c:98-170): element
// row i = (ib, ii): d = A(i,i), r = sum of |A(i,j)|, j != i, in block then
// column order, held in T.  Written at ib*BT + ii; padding rows get values
// that never win the min/max.
template <typename T, int BT>
__global__ __launch_bounds__(64) void k_eb_gersh_rows(int NB, int MB, const int *bsize,
                                                      const T *Av, const int *Ai,
                                                      const int *An, double *rowlo,
                                                      double *rowhi) {
    using E = Elem<T>;
    const int ib = blockIdx.x, ii = threadIdx.x;
    if (ii >= BT) return;
    const size_t at = (size_t) ib * BT + ii;
    if (ii >= bsize[ib]) {
        rowlo[at] = 1.7976931348623157e308;
        rowhi[at] = -1.7976931348623157e308;
        return;
    }
    constexpr size_t B2 = (size_t) BT * BT;
    T d = E::zero(), rad = E::zero();
    for (int jp = 0; jp < An[ib]; jp++) {
        const size_t sl = (size_t) ib * MB + jp;
        const int jb = Ai[sl];
        for (int jj = 0; jj < bsize[jb]; jj++) {
            const T v = Av[sl * B2 + ii * BT + jj];
            if (jb == ib && jj == ii)
                d = v;
            else
                rad = E::acc_mag(rad, E::mag(v));
        }
    }
    rowlo[at] = E::re_minus(d, rad);
    rowhi[at] = E::re_plus(d, rad);
}

// x <- s*x over the live blocks (bml_scale_ellblock)
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_eb_scale(int MB, T *Av, const int *An, T sv) {
    const int ib = blockIdx.x;
    constexpr size_t B2 = (size_t) BT * BT;
    const size_t n = (size_t) An[ib] * B2;
    T *row = Av + (size_t) ib * MB * B2;
    for (size_t e = threadIdx.x; e < n; e += 256) row[e] = Elem<T>::mul(sv, row[e]);
}

// identity: one diagonal block per block row
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_eb_identity(int MB, const int *bsize, T *Av, int *Ai,
                                                     int *An) {
    const int ib = blockIdx.x;
    constexpr int B2 = BT * BT;
    T *t = Av + (size_t) ib * MB * B2;
    for (int e = threadIdx.x; e < B2; e += 256) {
        const int ii = e / BT, jj = e % BT;
        t[e] = (ii == jj && ii < bsize[ib]) ? Elem<T>::from_real(1.0) : Elem<T>::zero();
    }
    if (threadIdx.x == 0) {
        Ai[(size_t) ib * MB] = ib;
        An[ib] = 1;
    }
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

template <typename T, int ENGINE, int BT>
int launch_numeric(const EbParams &p, hipStream_t s) {
    if constexpr (ENGINE != 0 && !(BT == 32 && (std::is_same<T, float>::value ||
                                                (ENGINE == 1 && std::is_same<T, double>::value)))) {
        return fail_msg(2, "no MFMA engine for this element type and tile");
    } else {
        // kept flags, then (sym) a staging tile per wavefront, as the kernel
        // lays them out
        const int mbits = p.maxs < p.MC ? p.maxs : p.MC;
        size_t lds = sizeof(unsigned) * (size_t) ((mbits + 31) / 32 + 4) / 4 * 4;
        if (p.sym) lds += sizeof(T) * (size_t) NUM_WAVES * BT * (BT + 1);
        if (lds > 160 * 1024) return fail_msg(5, "ellblock product: %d block slots per row exceed LDS", p.MC);
        EbParams q = p;
        q.ka = 0;
        // fused update: A's block columns after the A tiles
        const size_t acol_bytes = p.fuse ? ((size_t) p.MA * sizeof(int) + 15) / 16 * 16 : 0;
        if (AccOf<T, ENGINE, BT>::type::kStage) {
            // A tiles staged in LDS: as many as fit in the block's share of a
            // CU's LDS, at most the row width.  Default share: 78 KB (two
            // blocks per CU) for FP32 tiles, 48 KB (three) for bf16 ones:
            // on the 16x16x8-site, 32-orbital X^2, FP32 0.551 ms at 78 KB
            // against 0.580 at 48, bf16 0.476 against 0.457
            const size_t tb = (size_t) BT * BT * (ENGINE == 2 ? sizeof(__bf16) : sizeof(T));
            static long budget = -1, cap = -1;
            if (budget < 0) {
                const char *e = getenv("BML_HIP_EB_ALDS_KB");
                budget = e ? 1024L * atol(e) : 0;
                e = getenv("BML_HIP_EB_ASTAGE");
                cap = e ? atol(e) : 1L << 30;
            }
            const long share = budget > 0 ? budget : 1024L * (ENGINE == 2 ? 48 : 78);
            const long b = share < 156L * 1024 ? share : 156L * 1024;
            const long used = (long) (lds + acol_bytes);
            long ka = b > used ? (b - used) / (long) tb : 0;
            if (ka > cap) ka = cap;
            if (ka > p.MA) ka = p.MA;
            q.ka = (int) ka;
            lds += (size_t) ka * tb;
        }
        q.acol_off = (int) lds;
        lds += acol_bytes;
        if (lds > 160 * 1024) return fail_msg(5, "ellblock product: %d block slots per row exceed LDS", p.MA);
        if (lds > 64 * 1024)
            (void) hipFuncSetAttribute((const void *) k_eb_numeric<T, ENGINE, BT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        hipLaunchKernelGGL((k_eb_numeric<T, ENGINE, BT>), dim3(p.NB), dim3(64 * NUM_WAVES), lds, s,
                           q);
        int rc = report("k_eb_numeric");
        if (rc || !p.sym) return rc;
        const int nbits = p.maxs < p.MC ? p.maxs : p.MC;
        const size_t lc = 2 * sizeof(int) * (size_t) (nbits + 1);  // the move list
        if (lc > 64 * 1024) {
            if (lc > 160 * 1024) return fail_msg(5, "ellblock product: %d block slots per row exceed LDS", p.MC);
            (void) hipFuncSetAttribute((const void *) k_eb_compact<T, BT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) lc);
        }
        hipLaunchKernelGGL((k_eb_compact<T, BT>), dim3(p.NB), dim3(64), lc, s, p);
        return report("k_eb_compact");
    }
}

template <typename T>
int eb_product_typed(const bmlhip_eb_args *a) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (a->MA >= 65536 || a->MB >= 65536)
        return fail_msg(5, "ellblock product: more than 65535 block slots per row");
    EbParams p;
    memset(&p, 0, sizeof p);
    p.prec = a->prec;
    p.kind = a->kind;
    p.NB = a->NB;
    p.bt = a->bt;
    p.bsize = a->d_bsize;
    p.MA = a->MA;
    p.MB = a->MB;
    p.MC = a->MC;
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
    {
        /* contiguous block rows per XCD: 1.68 -> 1.60 ms for the FP64 MFMA
         * X^2 of the 16x16x8-site, 32-orbital lattice */
        const char *e = getenv("BML_HIP_EB_REMAP");
        p.remap = e ? atoi(e) : 1;
    }
    const long long maxp = (long long) a->MA * a->MB;
    // distinct block columns a row's slots can name: those its pairs touch,
    // plus (fused update) A's own columns, which take the first slots
    const int fuse_req = a->kind == 0 && a->sym && a->engine != 0 && a->fuse;
    const long long named = maxp + (fuse_req ? a->MA : 0);
    const long long maxs = named < a->NB ? named : a->NB;
    // slots the symbolic pass gives out: MC + 1 show an overflow
    const int cap = (int) (maxs < (long long) a->MC + 1 ? maxs : (long long) a->MC + 1);
    int tbits = 6;
    while ((1ll << tbits) < 2 * ((long long) cap + 64)) tbits++;
    const size_t lds = (size_t) (2 * (1ll << tbits) + cap + 64 + 2 * (long long) a->MA) * sizeof(int);
    if (lds > 160 * 1024)
        return fail_msg(5, "ellblock product: %lld block pairs per row exceed the LDS slot map",
                        maxp);
    p.maxs = (int) maxs;
    p.maxp = (int) maxp;
    const size_t NB = (size_t) a->NB;
    p.rowtr = (double *) bmlhip_scratch(0, sizeof(double) * 3 * (NB + 1));
    // fused update: the next iterate's traces, then the rows' norms
    p.rownorm = p.rowtr;
    p.cnt = (int *) bmlhip_scratch(3, sizeof(int) * (NB + 1));
    p.skb = (int *) bmlhip_scratch(4, sizeof(int) * (NB * maxs + 1));
    p.soff = (int *) bmlhip_scratch(5, sizeof(int) * (NB * (maxs + 1) + 1));
    p.pairs = (int *) bmlhip_scratch(6, sizeof(int) * (NB * maxp + 1));
    p.pslot = (int *) bmlhip_scratch(10, sizeof(int) * (NB * maxp + 1));
    p.bfirst = (int *) bmlhip_scratch(11, sizeof(int) * (NB * (maxp / EB_BATCH + 2) + 1));
    if ((long long) a->NB * a->MA >= (1ll << 31) || (long long) a->NB * a->MB >= (1ll << 31))
        return fail_msg(5, "ellblock product: more than 2^31 block slots");
    p.sym = a->kind == 0 && a->sym && a->engine != 0;
    p.fuse = p.sym && a->fuse;
    p.ab = a->d_ab;
    if (p.fuse && !p.ab) return fail_msg(2, "ellblock product: the fused update needs alpha, beta");
    if (p.sym && !(p.kflag = (int *) bmlhip_scratch(1, sizeof(int) * (NB * maxs + 1))))
        return fail_msg(4, "ellblock product: scratch allocation failed");
    p.rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * (NB + 1));
    if (!p.rowtr || !p.cnt || !p.skb || !p.soff || !p.pairs || !p.pslot || !p.bfirst || !p.rowstat)
        return fail_msg(4, "ellblock product: scratch allocation failed");
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (a->NB == 0) return 0;
    if (lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) k_eb_symbolic,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
    hipLaunchKernelGGL(k_eb_symbolic, dim3(a->NB), dim3(64), lds, s, p, tbits, cap);
    if ((rc = report("k_eb_symbolic"))) return rc;
    if (std::is_same<T, float>::value && a->bt == 32 && a->engine == 2) {
        // bf16 engine: the tiles in bfloat16 once per product (A row-major,
        // B transposed; one pass when A is B)
        __bf16 *a16 = (__bf16 *) bmlhip_scratch(BMLHIP_EB_A16_SLOT,
                                                sizeof(__bf16) * 1024 * (NB * a->MA) + 256);
        __bf16 *b16 = (__bf16 *) bmlhip_scratch(BMLHIP_EB_B16_SLOT,
                                                sizeof(__bf16) * 1024 * (NB * a->MB) + 256);
        if (!a16 || !b16) return fail_msg(4, "ellblock product: scratch allocation failed");
        if (a->d_Av == a->d_Bv && a->MA == a->MB) {
            hipLaunchKernelGGL(k_eb_bf16, dim3(a->NB), dim3(256), 0, s, a->MA,
                               (const float *) a->d_Av, a->d_An, a16, b16, p.skip);
        } else {
            hipLaunchKernelGGL(k_eb_bf16, dim3(a->NB), dim3(256), 0, s, a->MA,
                               (const float *) a->d_Av, a->d_An, a16, (__bf16 *) nullptr, p.skip);
            hipLaunchKernelGGL(k_eb_bf16, dim3(a->NB), dim3(256), 0, s, a->MB,
                               (const float *) a->d_Bv, a->d_Bn, (__bf16 *) nullptr, b16, p.skip);
        }
        if ((rc = report("k_eb_bf16"))) return rc;
        p.A16 = a16;
        p.B16 = b16;
    }
    const int prec_mark = a->prec;
    bmlhip_prof_mark_begin(prec_mark);
    if (std::is_same<T, float>::value && a->bt == 32 && a->engine == 1) {
        rc = launch_numeric<T, 1, 32>(p, s);
    } else if (std::is_same<T, float>::value && a->bt == 32 && a->engine == 2) {
        rc = launch_numeric<T, 2, 32>(p, s);
    } else if (std::is_same<T, double>::value && a->bt == 32 && a->engine == 1) {
        rc = launch_numeric<T, 1, 32>(p, s);
    } else {
        switch (a->bt) {
        case 8: rc = launch_numeric<T, 0, 8>(p, s); break;
        case 16: rc = launch_numeric<T, 0, 16>(p, s); break;
        case 32: rc = launch_numeric<T, 0, 32>(p, s); break;
        default: rc = fail_msg(2, "tile edge %d not one of 8/16/32", a->bt);
        }
    }
    bmlhip_prof_mark_end();
    if (rc) return rc;
    if ((rc = bmlhip_rowstat_reduce(a->NB, a->MC,
                                    BMLHIP_RS_OVF_DIST | BMLHIP_RS_STORED | BMLHIP_RS_PROF,
                                    a->prec, p.skip)))
        return rc;
    if (p.fuse) {
        // the rows' |X - X^2|^2, and the next iterate's traces for the next
        // branch decision
        hipLaunchKernelGGL(k_eb_sum3, dim3(1), dim3(1024), 0, s, p.rownorm, a->NB, p.st, p.skip);
        return report("k_eb_sum3");
    }
    if (a->kind == 0) {
        if ((rc = bmlhip_sum_to_status(p.rowtr, a->NB, 0))) return rc;
        if ((rc = bmlhip_sum_to_status(p.rowtr + a->NB, a->NB, 1))) return rc;
    }
    return 0;
}

template <typename T>
int eb_add_typed(int NB, int bt, const int *bsize, int MA, void *Av, int *Ai, int *An, int MB,
                 const void *Bv, const int *Bi, const int *Bn, double alpha, double beta,
                 double thr, int mode, const double *ab, const int *skip) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    EbAddParams p;
    p.NB = NB;
    p.bt = bt;
    p.MA = MA;
    p.MB = MB;
    p.mode = mode;
    p.bsize = bsize;
    p.Av = Av;
    p.Ai = Ai;
    p.An = An;
    p.Bv = Bv;
    p.Bi = Bi;
    p.Bn = Bn;
    p.alpha = alpha;
    p.beta = beta;
    p.thr = thr;
    p.ab = ab;
    p.skip = skip;
    p.st = bmlhip_status_dev();
    p.rownorm = (double *) bmlhip_scratch(0, sizeof(double) * ((size_t) NB + 1));
    p.rowstat = (int4 *) bmlhip_scratch(BMLHIP_ROWSTAT_SLOT, sizeof(int4) * ((size_t) NB + 1));
    if (!p.rownorm || !p.rowstat) return fail_msg(4, "ellblock add: scratch allocation failed");
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (NB == 0) return 0;
    const size_t lds = (size_t) (3 * (MA + MB) + 1) * sizeof(int);
    if (lds > 64 * 1024)
        return fail_msg(5, "ellblock add: %d + %d block slots per row exceed the LDS union list",
                        MA, MB);
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_add<T, 8>), dim3(NB), dim3(64 * EB_ADD_WAVES), lds, s, p); break;
    case 16: hipLaunchKernelGGL((k_eb_add<T, 16>), dim3(NB), dim3(64 * EB_ADD_WAVES), lds, s, p); break;
    case 32: hipLaunchKernelGGL((k_eb_add<T, 32>), dim3(NB), dim3(64 * EB_ADD_WAVES), lds, s, p); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    if ((rc = report("k_eb_add"))) return rc;
    // overflowing rows were counted in the kernel (the first overflowing entry)
    if ((rc = bmlhip_rowstat_reduce(NB, MA, BMLHIP_RS_STORED, 0, skip))) return rc;
    if (mode == 2) return bmlhip_sum_to_status(p.rownorm, NB, 2);
    return 0;
}

template <typename T>
int eb_threshold_typed(int NB, int bt, const int *bsize, int MB, void *Av, int *Ai, int *An,
                       double thr) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (NB == 0) return 0;
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_threshold<T, 8>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (T *) Av, Ai, An, thr); break;
    case 16: hipLaunchKernelGGL((k_eb_threshold<T, 16>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (T *) Av, Ai, An, thr); break;
    case 32: hipLaunchKernelGGL((k_eb_threshold<T, 32>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (T *) Av, Ai, An, thr); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    return report("k_eb_threshold");
}

template <typename T>
int eb_trace_typed(int NB, int bt, const int *bsize, int MB, const void *Av, const int *Ai,
                   const int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    double *rowtr = (double *) bmlhip_scratch(0, sizeof(double) * ((size_t) NB + 1));
    if (!rowtr) return fail_msg(4, "ellblock trace: scratch allocation failed");
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (NB == 0) return 0;
    const dim3 g((NB + 255) / 256);
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_trace_rows<T, 8>), g, dim3(256), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rowtr); break;
    case 16: hipLaunchKernelGGL((k_eb_trace_rows<T, 16>), g, dim3(256), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rowtr); break;
    case 32: hipLaunchKernelGGL((k_eb_trace_rows<T, 32>), g, dim3(256), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rowtr); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    if ((rc = report("k_eb_trace_rows"))) return rc;
    return bmlhip_sum_to_status(rowtr, NB, 0);
}

template <typename T>
int eb_trace_frob_typed(int NB, int bt, int MB, const void *Av, const int *Ai, const int *An,
                        const int *skip) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    double *rowtr = (double *) bmlhip_scratch(0, sizeof(double) * (2 * (size_t) NB + 1));
    if (!rowtr) return fail_msg(4, "ellblock trace: scratch allocation failed");
    if (NB == 0) return 0;
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_trace_frob<T, 8>), dim3(NB), dim3(256), 0, s, NB, MB, (const T *) Av, Ai, An, rowtr, skip); break;
    case 16: hipLaunchKernelGGL((k_eb_trace_frob<T, 16>), dim3(NB), dim3(256), 0, s, NB, MB, (const T *) Av, Ai, An, rowtr, skip); break;
    case 32: hipLaunchKernelGGL((k_eb_trace_frob<T, 32>), dim3(NB), dim3(256), 0, s, NB, MB, (const T *) Av, Ai, An, rowtr, skip); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    int rc = report("k_eb_trace_frob");
    if (rc) return rc;
    if ((rc = bmlhip_sum_to_status(rowtr, NB, 0))) return rc;
    return bmlhip_sum_to_status(rowtr + NB, NB, 1);
}

template <typename T>
int eb_copy_typed(int NB, int bt, int MA, const void *Av, const int *Ai, const int *An, int MB,
                  void *Bv, int *Bi, int *Bn) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    if (NB == 0) return 0;
    bmlhip_status *st = bmlhip_status_dev();
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_copy<T, 8>), dim3(NB), dim3(256), 0, s, NB, MA, (const T *) Av, Ai, An, MB, (T *) Bv, Bi, Bn, st); break;
    case 16: hipLaunchKernelGGL((k_eb_copy<T, 16>), dim3(NB), dim3(256), 0, s, NB, MA, (const T *) Av, Ai, An, MB, (T *) Bv, Bi, Bn, st); break;
    case 32: hipLaunchKernelGGL((k_eb_copy<T, 32>), dim3(NB), dim3(256), 0, s, NB, MA, (const T *) Av, Ai, An, MB, (T *) Bv, Bi, Bn, st); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    return report("k_eb_copy");
}

template <typename T>
int eb_gershgorin_typed(int NB, int bt, const int *bsize, int MB, const void *Av, const int *Ai,
                        const int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int rc = bmlhip_status_reset();
    if (rc) return rc;
    const size_t n = (size_t) NB * bt;
    double *rows = (double *) bmlhip_scratch(0, sizeof(double) * 2 * (n + 1));
    if (!rows) return fail_msg(4, "ellblock gershgorin: scratch allocation failed");
    if (NB > 0) {
        switch (bt) {
        case 8: hipLaunchKernelGGL((k_eb_gersh_rows<T, 8>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rows, rows + n); break;
        case 16: hipLaunchKernelGGL((k_eb_gersh_rows<T, 16>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rows, rows + n); break;
        case 32: hipLaunchKernelGGL((k_eb_gersh_rows<T, 32>), dim3(NB), dim3(64), 0, s, NB, MB, bsize, (const T *) Av, Ai, An, rows, rows + n); break;
        default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
        }
        if ((rc = report("k_eb_gersh_rows"))) return rc;
    }
    return bmlhip_minmax_to_status(rows, rows + n, (int) n);
}

template <typename T>
int eb_scale_typed(int NB, int bt, int MB, void *Av, const int *An, const void *s_bits) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    T sv;
    memcpy(&sv, s_bits, sizeof(T));
    if (NB == 0) return 0;
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_scale<T, 8>), dim3(NB), dim3(256), 0, s, MB, (T *) Av, An, sv); break;
    case 16: hipLaunchKernelGGL((k_eb_scale<T, 16>), dim3(NB), dim3(256), 0, s, MB, (T *) Av, An, sv); break;
    case 32: hipLaunchKernelGGL((k_eb_scale<T, 32>), dim3(NB), dim3(256), 0, s, MB, (T *) Av, An, sv); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    return report("k_eb_scale");
}

template <typename T>
int eb_identity_typed(int NB, int bt, const int *bsize, int MB, void *Av, int *Ai, int *An) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (NB == 0) return 0;
    switch (bt) {
    case 8: hipLaunchKernelGGL((k_eb_identity<T, 8>), dim3(NB), dim3(256), 0, s, MB, bsize, (T *) Av, Ai, An); break;
    case 16: hipLaunchKernelGGL((k_eb_identity<T, 16>), dim3(NB), dim3(256), 0, s, MB, bsize, (T *) Av, Ai, An); break;
    case 32: hipLaunchKernelGGL((k_eb_identity<T, 32>), dim3(NB), dim3(256), 0, s, MB, bsize, (T *) Av, Ai, An); break;
    default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
    }
    return report("k_eb_identity");
}

// ---------------------------------------------------------------------------
// symmetry test: one wavefront per block row; every stored block (ib, jb)
// needs a stored (jb, ib) holding its conjugate transpose (values compared
// with ==, so +0 and -0 match and a NaN never does)

template <typename T, int BT>
__global__ __launch_bounds__(64) void k_eb_check_sym(int NB, int MB, const T *Av, const int *Ai,
                                                     const int *An, int *bad) {
    const int ib = blockIdx.x, lane = threadIdx.x;
    constexpr int B2 = BT * BT;
    bool ok = true;
    const int n = An[ib];
    for (int jp = 0; jp < n && ok; jp++) {
        const size_t sa = (size_t) ib * MB + jp;
        const int jb = Ai[sa];
        if (jb < 0 || jb >= NB) {
            ok = false;
            break;
        }
        const int nj = An[jb];
        int sp = -1;
        for (int c0 = 0; c0 < nj && sp < 0; c0 += 64) {
            const int q = c0 + lane;
            const uint64_t m = ballot64(q < nj && Ai[(size_t) jb * MB + q] == ib);
            if (m) sp = c0 + __ffsll((unsigned long long) m) - 1;
        }
        if (sp < 0) {
            ok = false;
            break;
        }
        const T *X = Av + sa * B2, *Y = Av + ((size_t) jb * MB + sp) * B2;
        bool eq = true;
        for (int e = lane; e < B2; e += 64)
            eq = eq && same_value(X[e], Elem<T>::conj(Y[(e % BT) * BT + e / BT]));
        ok = ballot64(!eq) == 0;
    }
    if (!ok && lane == 0) atomicAdd(bad, 1);
}

template <typename T>
int eb_check_sym_typed(int NB, int bt, int MB, const void *Av, const int *Ai, const int *An,
                       int *symmetric) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    int *bad = (int *) bmlhip_scratch(BMLHIP_SYMFILL_SLOT + 1, sizeof(int) * 64);
    if (!bad) return fail_msg(4, "scratch allocation failed");
    if (hipMemsetAsync(bad, 0, sizeof(int), s) != hipSuccess) return fail_msg(3, "memset failed");
    if (NB > 0) {
        switch (bt) {
        case 8:
            hipLaunchKernelGGL((k_eb_check_sym<T, 8>), dim3(NB), dim3(64), 0, s, NB, MB,
                               (const T *) Av, Ai, An, bad);
            break;
        case 16:
            hipLaunchKernelGGL((k_eb_check_sym<T, 16>), dim3(NB), dim3(64), 0, s, NB, MB,
                               (const T *) Av, Ai, An, bad);
            break;
        case 32:
            hipLaunchKernelGGL((k_eb_check_sym<T, 32>), dim3(NB), dim3(64), 0, s, NB, MB,
                               (const T *) Av, Ai, An, bad);
            break;
        default: return fail_msg(2, "tile edge %d not one of 8/16/32", bt);
        }
    }
    int rc = report("k_eb_check_sym");
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

int bmlhip_eb_product(const bmlhip_eb_args *a) {
#define CALL(T) eb_product_typed<T>(a)
    DISPATCH(a->prec, CALL)
#undef CALL
}

int bmlhip_eb_is_symmetric(int prec, int NB, int bt, int MB, const void *Av, const int *Ai,
                           const int *An, int *symmetric) {
#define CALL(T) eb_check_sym_typed<T>(NB, bt, MB, Av, Ai, An, symmetric)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_add(int prec, int NB, int bt, const int *bsize, int MA, void *Av, int *Ai,
                  int *An, int MB, const void *Bv, const int *Bi, const int *Bn, double alpha,
                  double beta, double thr, int mode, const double *ab, const int *skip) {
#define CALL(T) eb_add_typed<T>(NB, bt, bsize, MA, Av, Ai, An, MB, Bv, Bi, Bn, alpha, beta, thr, mode, ab, skip)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_threshold(int prec, int NB, int bt, const int *bsize, int MB, void *Av, int *Ai,
                        int *An, double thr) {
#define CALL(T) eb_threshold_typed<T>(NB, bt, bsize, MB, Av, Ai, An, thr)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_trace(int prec, int NB, int bt, const int *bsize, int MB, const void *Av,
                    const int *Ai, const int *An) {
#define CALL(T) eb_trace_typed<T>(NB, bt, bsize, MB, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_trace_frob(int prec, int NB, int bt, int MB, const void *Av, const int *Ai,
                         const int *An, const int *skip) {
#define CALL(T) eb_trace_frob_typed<T>(NB, bt, MB, Av, Ai, An, skip)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_copy(int prec, int NB, int bt, int MA, const void *Av, const int *Ai,
                   const int *An, int MB, void *Bv, int *Bi, int *Bn) {
#define CALL(T) eb_copy_typed<T>(NB, bt, MA, Av, Ai, An, MB, Bv, Bi, Bn)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_gershgorin(int prec, int NB, int bt, const int *bsize, int MB, const void *Av,
                         const int *Ai, const int *An) {
#define CALL(T) eb_gershgorin_typed<T>(NB, bt, bsize, MB, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_scale(int prec, int NB, int bt, int MB, void *Av, const int *An,
                    const void *s_bits) {
#define CALL(T) eb_scale_typed<T>(NB, bt, MB, Av, An, s_bits)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_identity(int prec, int NB, int bt, const int *bsize, int MB, void *Av, int *Ai,
                       int *An) {
#define CALL(T) eb_identity_typed<T>(NB, bt, bsize, MB, Av, Ai, An)
    DISPATCH(prec, CALL)
#undef CALL
}

int bmlhip_eb_row_stats(int NB, const int *An) { return bmlhip_row_stats(NB, An); }

}  // extern "C"

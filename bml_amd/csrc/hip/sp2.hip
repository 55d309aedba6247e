// Device-side pieces of SP2 purification: the branch decision and the
// convergence test run as one-thread kernels on the library stream, so an
// iteration needs no host round trip; and value conversion between
// precisions for the mixed-precision schedule.
#include "../bmlhip_abi.h"
#include "devtypes.h"

#include <math.h>
#include <stdio.h>

using namespace bmlhip;

namespace {

// Niklasson's trace-correcting choice: X^2 lowers the trace, 2X - X^2 raises
// it; take the one whose trace lands closer to nocc.
__global__ void k_sp2_decide(bmlhip_sp2_state *s, const bmlhip_status *st, double nocc,
                             double *hist, int maxit) {
    if (threadIdx.x != 0 || s->done) return;
    const double trx = st->scalars[0], trx2 = st->scalars[1];
    const int branch = fabs(trx2 - nocc) < fabs(2.0 * trx - trx2 - nocc) ? 0 : 1;
    s->branch = branch;
    s->ab[0] = branch ? 2.0 : 0.0;
    s->ab[1] = branch ? -1.0 : 1.0;
    s->trx = trx;
    s->trx2 = trx2;
    s->max_touched = st->max_touched;
    s->max_nnz_x2 = st->max_nnz;
    if (hist && s->it < maxit) {
        hist[4 * s->it + 0] = trx;
        hist[4 * s->it + 1] = trx2;
        hist[4 * s->it + 3] = (double) branch;
    }
}

__global__ void k_sp2_check(bmlhip_sp2_state *s, const bmlhip_status *st, double tol,
                            double *hist, int maxit, int touched) {
    if (threadIdx.x != 0 || s->done) return;
    const double nrm = sqrt(st->scalars[2]);
    s->norm = nrm;
    s->max_nnz = st->max_nnz;
    if (touched) s->max_touched = st->max_touched;
    if (hist && s->it < maxit) hist[4 * s->it + 2] = nrm;
    s->it++;
    if (nrm <= tol) s->done = 1;
}

// batched: one thread per system
__global__ void k_sp2_decide_batch(int nsys, const int *done, const double *trx,
                                   const double *trx2, const double *nocc, double *ab,
                                   int *branch, const bmlhip_status *st, int *stats) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) stats[0] = st->max_touched;
    if (s >= nsys || done[s]) return;
    const int br = fabs(trx2[s] - nocc[s]) < fabs(2.0 * trx[s] - trx2[s] - nocc[s]) ? 0 : 1;
    branch[s] = br;
    ab[2 * s] = br ? 2.0 : 0.0;
    ab[2 * s + 1] = br ? -1.0 : 1.0;
}

__global__ void k_sp2_check_batch(int nsys, int *done, const double *norm2, double tol,
                                  int maxit, int *it, double *norm, int *alldone,
                                  const bmlhip_status *st, int *stats) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) stats[1] = st->max_nnz;
    if (s >= nsys || done[s]) return;
    const double nrm = sqrt(norm2[s]);
    norm[s] = nrm;
    it[s] += 1;
    if (nrm <= tol || it[s] >= maxit) {
        done[s] = nrm <= tol ? 1 : 2;  // 1 converged, 2 iteration cap
        atomicAdd(alldone, 1);
    }
}

template <typename A, typename B>
__device__ B cvt(A a);
template <> __device__ double cvt<float, double>(float a) { return (double) a; }
template <> __device__ float cvt<double, float>(double a) { return (float) a; }
template <> __device__ cf64 cvt<cf32, cf64>(cf32 a) { return {(double) a.x, (double) a.y}; }
template <> __device__ cf32 cvt<cf64, cf32>(cf64 a) { return {(float) a.x, (float) a.y}; }

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_convert(size_t n, const A *in, B *out) {
    for (size_t k = (size_t) blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t) gridDim.x * 256)
        out[k] = cvt<A, B>(in[k]);
}

// the first cnt[r] * unit elements of each row r (rows `stride` elements
// apart): the live slots of an ELLPACK or ELLBLOCK matrix; one block per row
template <typename A, typename B>
__global__ __launch_bounds__(256) void k_convert_rows(size_t stride, size_t unit, const int *cnt,
                                                      const A *in, B *out) {
    const size_t r = blockIdx.x, n = (size_t) cnt[r] * unit;
    const A *x = in + r * stride;
    B *y = out + r * stride;
    size_t k = threadIdx.x;
    for (; k + 768 < n; k += 1024) {  // four loads in flight per thread
        A v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = x[k + 256 * q];
#pragma unroll
        for (int q = 0; q < 4; q++) y[k + 256 * q] = cvt<A, B>(v[q]);
    }
    for (; k < n; k += 256) y[k] = cvt<A, B>(x[k]);
}

int launch_error(const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    bmlhip_set_errmsg(buf);
    return 3;
}

}  // namespace

extern "C" {

int bmlhip_sp2_decide(bmlhip_sp2_state *d_state, double nocc, double *d_hist, int maxit) {
    hipLaunchKernelGGL(k_sp2_decide, dim3(1), dim3(64), 0, (hipStream_t) bmlhip_stream(), d_state,
                       bmlhip_status_dev(), nocc, d_hist, maxit);
    return launch_error("k_sp2_decide");
}

int bmlhip_sp2_check(bmlhip_sp2_state *d_state, double tol, double *d_hist, int maxit,
                     int touched) {
    hipLaunchKernelGGL(k_sp2_check, dim3(1), dim3(64), 0, (hipStream_t) bmlhip_stream(), d_state,
                       bmlhip_status_dev(), tol, d_hist, maxit, touched);
    return launch_error("k_sp2_check");
}

int bmlhip_sp2_decide_batch(const bmlhip_batch *b, const double *nocc, int *branch,
                            int *stats) {
    hipLaunchKernelGGL(k_sp2_decide_batch, dim3((b->nsys + 255) / 256), dim3(256), 0,
                       (hipStream_t) bmlhip_stream(), b->nsys, b->d_done, b->d_trx, b->d_trx2,
                       nocc, b->d_ab, branch, bmlhip_status_dev(), stats);
    return launch_error("k_sp2_decide_batch");
}

int bmlhip_sp2_check_batch(const bmlhip_batch *b, double tol, int maxit, int *it, double *norm,
                           int *alldone, int *stats) {
    hipLaunchKernelGGL(k_sp2_check_batch, dim3((b->nsys + 255) / 256), dim3(256), 0,
                       (hipStream_t) bmlhip_stream(), b->nsys, b->d_done, b->d_norm2, tol, maxit,
                       it, norm, alldone, bmlhip_status_dev(), stats);
    return launch_error("k_sp2_check_batch");
}

int bmlhip_convert(int prec_in, int prec_out, size_t count, const void *in, void *out) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    const int blocks = (int) ((count + 255) / 256 < 65536 ? (count + 255) / 256 : 65536);
    if (count == 0) return 0;
    if (prec_in == 1 && prec_out == 2)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks), dim3(256), 0, s, count,
                           (const float *) in, (double *) out);
    else if (prec_in == 2 && prec_out == 1)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks), dim3(256), 0, s, count,
                           (const double *) in, (float *) out);
    else if (prec_in == 3 && prec_out == 4)
        hipLaunchKernelGGL((k_convert<cf32, cf64>), dim3(blocks), dim3(256), 0, s, count,
                           (const cf32 *) in, (cf64 *) out);
    else if (prec_in == 4 && prec_out == 3)
        hipLaunchKernelGGL((k_convert<cf64, cf32>), dim3(blocks), dim3(256), 0, s, count,
                           (const cf64 *) in, (cf32 *) out);
    else {
        bmlhip_set_errmsg("unsupported precision conversion");
        return 2;
    }
    return launch_error("k_convert");
}

int bmlhip_convert_rows(int prec_in, int prec_out, int nrows, size_t stride, size_t unit,
                        const int *d_cnt, const void *in, void *out) {
    hipStream_t s = (hipStream_t) bmlhip_stream();
    if (nrows <= 0) return 0;
#define CVR(A, B)                                                                            \
    hipLaunchKernelGGL((k_convert_rows<A, B>), dim3(nrows), dim3(256), 0, s, stride, unit, d_cnt, \
                       (const A *) in, (B *) out)
    if (prec_in == 1 && prec_out == 2) CVR(float, double);
    else if (prec_in == 2 && prec_out == 1) CVR(double, float);
    else if (prec_in == 3 && prec_out == 4) CVR(cf32, cf64);
    else if (prec_in == 4 && prec_out == 3) CVR(cf64, cf32);
    else {
        bmlhip_set_errmsg("unsupported precision conversion");
        return 2;
    }
#undef CVR
    return launch_error("k_convert_rows");
}

}  // extern "C"

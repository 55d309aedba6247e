// Element types and arithmetic for the gfx950 kernels.
//
// The ELLPACK kernels are compiled with -ffp-contract=off: every a*b+c below
// is a rounded multiply followed by a rounded add, the way the reference's C
// loops evaluate `x = x + a*b` on x86-64, so real-typed results match the
// reference bit for bit when the summation order matches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bmlhip {

struct cf32 {
    float x, y;
};
struct cf64 {
    double x, y;
};

template <typename T> struct Elem;

template <> struct Elem<float> {
    using real = float;
    static constexpr bool complex = false;
    __device__ static float zero() { return 0.0f; }
    __device__ static float conj(float v) { return v; }
    __device__ static float add(float a, float b) { return a + b; }
    __device__ static float sub(float a, float b) { return a - b; }
    __device__ static float mul(float a, float b) { return a * b; }
    // |v| > t exactly as the reference's is_above_threshold: fabsf(x) > (float) t
    __device__ static bool above(float v, double t) { return fabsf(v) > (float) t; }
    __device__ static bool nonzero(float v) { return v != 0.0f; }
    __device__ static double re(float v) { return (double) v; }
    __device__ static double mag(float v) { return (double) fabsf(v); }
    // |v|^2 as the oracle forms it: the magnitude squared in the element's
    // real type, then widened
    __device__ static double abs2(float v) {
        float m = fabsf(v);
        return (double) (m * m);
    }
    // x + s*a with s a double: C promotes to double, rounds once on store
    __device__ static float axpy(float x, double s, float a) {
        return (float) ((double) x + s * (double) a);
    }
    // x + s with s a double (add_identity's diagonal update)
    __device__ static float add_real(float x, double s) { return (float) ((double) x + s); }
    __device__ static float from_real(double s) { return (float) s; }
    // radius += (double) |a| with radius held in T
    __device__ static float acc_mag(float r, double m) { return (float) ((double) r + m); }
    __device__ static double re_plus(float d, float r) { return (double) (d + r); }
    __device__ static double re_minus(float d, float r) { return (double) (d - r); }
    // Re(y*y) accumulated into a double (the reference's add_norm)
    __device__ static double re_sq(float y) { return (double) (y * y); }
};

template <> struct Elem<double> {
    using real = double;
    static constexpr bool complex = false;
    __device__ static double zero() { return 0.0; }
    __device__ static double conj(double v) { return v; }
    __device__ static double add(double a, double b) { return a + b; }
    __device__ static double sub(double a, double b) { return a - b; }
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static bool above(double v, double t) { return fabs(v) > t; }
    __device__ static bool nonzero(double v) { return v != 0.0; }
    __device__ static double re(double v) { return v; }
    __device__ static double mag(double v) { return fabs(v); }
    __device__ static double abs2(double v) { return v * v; }
    __device__ static double axpy(double x, double s, double a) { return x + s * a; }
    __device__ static double add_real(double x, double s) { return x + s; }
    __device__ static double from_real(double s) { return s; }
    __device__ static double acc_mag(double r, double m) { return r + m; }
    __device__ static double re_plus(double d, double r) { return d + r; }
    __device__ static double re_minus(double d, double r) { return d - r; }
    __device__ static double re_sq(double y) { return y * y; }
};

template <> struct Elem<cf32> {
    using real = float;
    static constexpr bool complex = true;
    __device__ static cf32 zero() { return {0.0f, 0.0f}; }
    __device__ static cf32 conj(cf32 v) { return {v.x, -v.y}; }
    __device__ static cf32 add(cf32 a, cf32 b) { return {a.x + b.x, a.y + b.y}; }
    __device__ static cf32 sub(cf32 a, cf32 b) { return {a.x - b.x, a.y - b.y}; }
    __device__ static cf32 mul(cf32 a, cf32 b) {
        return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
    }
    __device__ static bool above(cf32 v, double t) { return hypotf(v.x, v.y) > (float) t; }
    __device__ static bool nonzero(cf32 v) { return v.x != 0.0f || v.y != 0.0f; }
    __device__ static double re(cf32 v) { return (double) v.x; }
    __device__ static double mag(cf32 v) { return (double) hypotf(v.x, v.y); }
    __device__ static double abs2(cf32 v) {
        float m = hypotf(v.x, v.y);
        return (double) (m * m);
    }
    // float complex + double*(float complex): evaluated in double complex
    __device__ static cf32 axpy(cf32 x, double s, cf32 a) {
        return {(float) ((double) x.x + s * (double) a.x),
                (float) ((double) x.y + s * (double) a.y)};
    }
    __device__ static cf32 add_real(cf32 x, double s) {
        return {(float) ((double) x.x + s), x.y};
    }
    __device__ static cf32 from_real(double s) { return {(float) s, 0.0f}; }
    __device__ static cf32 acc_mag(cf32 r, double m) {
        return {(float) ((double) r.x + m), r.y};
    }
    __device__ static double re_plus(cf32 d, cf32 r) { return (double) (d.x + r.x); }
    __device__ static double re_minus(cf32 d, cf32 r) { return (double) (d.x - r.x); }
    __device__ static double re_sq(cf32 y) { return (double) (y.x * y.x - y.y * y.y); }
};

template <> struct Elem<cf64> {
    using real = double;
    static constexpr bool complex = true;
    __device__ static cf64 zero() { return {0.0, 0.0}; }
    __device__ static cf64 conj(cf64 v) { return {v.x, -v.y}; }
    __device__ static cf64 add(cf64 a, cf64 b) { return {a.x + b.x, a.y + b.y}; }
    __device__ static cf64 sub(cf64 a, cf64 b) { return {a.x - b.x, a.y - b.y}; }
    __device__ static cf64 mul(cf64 a, cf64 b) {
        return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
    }
    __device__ static bool above(cf64 v, double t) { return hypot(v.x, v.y) > t; }
    __device__ static bool nonzero(cf64 v) { return v.x != 0.0 || v.y != 0.0; }
    __device__ static double re(cf64 v) { return v.x; }
    __device__ static double mag(cf64 v) { return hypot(v.x, v.y); }
    __device__ static double abs2(cf64 v) {
        double m = hypot(v.x, v.y);
        return m * m;
    }
    __device__ static cf64 axpy(cf64 x, double s, cf64 a) {
        return {x.x + s * a.x, x.y + s * a.y};
    }
    __device__ static cf64 add_real(cf64 x, double s) { return {x.x + s, x.y}; }
    __device__ static cf64 from_real(double s) { return {s, 0.0}; }
    __device__ static cf64 acc_mag(cf64 r, double m) { return {r.x + m, r.y}; }
    __device__ static double re_plus(cf64 d, cf64 r) { return d.x + r.x; }
    __device__ static double re_minus(cf64 d, cf64 r) { return d.x - r.x; }
    __device__ static double re_sq(cf64 y) { return y.x * y.x - y.y * y.y; }
};

// Atomic add into an LDS slot (ds_add_f32 / ds_add_f64; a complex value is
// two component adds).  Concurrent adds to one slot land in arrival order.
__device__ __forceinline__ void lds_add(float *p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(cf32 *p, cf32 v) {
    lds_add(&p->x, v.x);
    lds_add(&p->y, v.y);
}
__device__ __forceinline__ void lds_add(cf64 *p, cf64 v) {
    lds_add(&p->x, v.x);
    lds_add(&p->y, v.y);
}

// XCD-aware block order.  The dispatcher deals the blocks of a grid to the 8
// XCDs round-robin (block b runs on XCD b % 8) and each XCD runs its share
// in order.  xcd_block(b, n) renumbers the n blocks so that XCD x takes the
// x-th contiguous range of about n/8 blocks: blocks resident on one XCD at a
// time are neighbours, and the operand rows they share stay in its L2.  A
// bijection on [0, n).
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int per = (n + 7) / 8;
    const int xcd = b % 8, k = b / 8;
    const int full = n % 8 == 0 ? 8 : n % 8;  // XCDs that take `per` blocks
    return xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
}

// Value equality (==, so +0.0 equals -0.0; complex by component): the
// symmetry tests compare a stored entry with the conjugate of its mirror,
// and conj turns a +0.0 imaginary part into -0.0.
__device__ __forceinline__ bool same_value(float a, float b) { return a == b; }
__device__ __forceinline__ bool same_value(double a, double b) { return a == b; }
__device__ __forceinline__ bool same_value(cf32 a, cf32 b) { return a.x == b.x && a.y == b.y; }
__device__ __forceinline__ bool same_value(cf64 a, cf64 b) { return a.x == b.x && a.y == b.y; }

// 64-lane wavefront helpers
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }
__device__ __forceinline__ int lane_id() { return (int) __lane_id(); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

}  // namespace bmlhip

// lkl_gsomath.hpp — host/device math of the GSO movement step.
//
// The glowworm move (glowworm.rs:128-150) slerps quaternions (qt.rs:80-104)
// with f64 acos and sin.  The reference takes them from the host libm, whose
// results for these arguments are the correctly rounded ones; the device's own
// acos/sin are accurate to an ulp or two but not always correctly rounded.  So
// the device move evaluates them in double-double arithmetic (about 100 bits)
// and rounds once, which gives the correctly rounded value, and with it the
// host's bits.  Everything else in the move is plain f64 arithmetic in the
// reference's operation order with contraction off.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace lkl {
namespace gm {

#define LKL_HD __host__ __device__ __forceinline__

struct dd {
    double hi, lo;
};

// Knuth's two-sum and the fma-based two-product: exact error terms.
LKL_HD dd two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
LKL_HD dd quick_two_sum(double a, double b) {  // |a| >= |b|
#pragma clang fp contract(off)
    const double s = a + b;
    return {s, b - (s - a)};
}
LKL_HD dd two_prod(double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}
LKL_HD dd add(dd x, dd y) {
#pragma clang fp contract(off)
    dd s = two_sum(x.hi, y.hi);
    const dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
LKL_HD dd mul(dd x, dd y) {
#pragma clang fp contract(off)
    dd p = two_prod(x.hi, y.hi);
    p.lo += x.hi * y.lo + x.lo * y.hi;
    return quick_two_sum(p.hi, p.lo);
}
// x / d for a small exact integer d, given inv = 1 / d rounded: q1 = x.hi * inv
// is within two ulps of the quotient, the remainder x - q1 d is formed exactly
// (two_prod) and its quotient q2 corrects q1, which leaves ~2^-103 relative
// error, far below what the final rounding needs.
LKL_HD dd div_d(dd x, double d, double inv) {
#pragma clang fp contract(off)
    const double q1 = x.hi * inv;
    const dd p = two_prod(q1, d);
    const dd r = add(x, {-p.hi, -p.lo});
    const double q2 = r.hi * inv;
    return quick_two_sum(q1, q2);
}

// The series divisors (2k)(2k+1) of sin and (2k-1)(2k) of cos, k = 1..29,
// and their rounded reciprocals.
struct SeriesDiv {
    double sin_d[30], sin_inv[30], cos_d[30], cos_inv[30];
    constexpr SeriesDiv() : sin_d(), sin_inv(), cos_d(), cos_inv() {
        for (int k = 1; k < 30; ++k) {
            sin_d[k] = (double)((2 * k) * (2 * k + 1));
            sin_inv[k] = 1.0 / sin_d[k];
            cos_d[k] = (double)((2 * k - 1) * (2 * k));
            cos_inv[k] = 1.0 / cos_d[k];
        }
    }
};
#ifdef __HIP_DEVICE_COMPILE__
__constant__ constexpr SeriesDiv kSeries{};
#else
constexpr SeriesDiv kSeries{};
#endif
LKL_HD double round_dd(dd x) { return x.hi + x.lo; }

// sin and cos of a double-double argument |x| <= 1.6 by their Taylor series
// (terms until they fall below 1e-34 of the result).
LKL_HD dd sin_dd(dd x) {
    const dd x2 = mul(x, x);
    dd term = x, sum = x;
    for (int k = 1; k < 30; ++k) {
        term = div_d(mul(term, x2), kSeries.sin_d[k], kSeries.sin_inv[k]);
        term = {-term.hi, -term.lo};
        sum = add(sum, term);
        if (std::fabs(term.hi) < 1e-34 * std::fabs(sum.hi)) break;
    }
    return sum;
}
LKL_HD dd cos_dd(dd x) {
    const dd x2 = mul(x, x);
    dd term = {1.0, 0.0}, sum = {1.0, 0.0};
    for (int k = 1; k < 30; ++k) {
        term = div_d(mul(term, x2), kSeries.cos_d[k], kSeries.cos_inv[k]);
        term = {-term.hi, -term.lo};
        sum = add(sum, term);
        if (std::fabs(term.hi) < 1e-34 * std::fabs(sum.hi)) break;
    }
    return sum;
}

// Correctly rounded sin(x) for 0 <= x <= 1.6 (the range the slerp needs).
LKL_HD double sin_cr(double x) { return round_dd(sin_dd({x, 0.0})); }

// Correctly rounded acos(d) for 0 <= d <= 1: one Newton step on cos(w) = d
// from the library value, carried in double-double.  *w0_out and *sin_w0
// (may be NULL) receive the library value and sin_cr of it, computed on the way.
LKL_HD double acos_cr(double d, double* w0_out = nullptr, double* sin_w0 = nullptr) {
#pragma clang fp contract(off)
    const double w0 = std::acos(d);
    if (w0_out) *w0_out = w0;
    if (w0 == 0.0) {
        if (sin_w0) *sin_w0 = 0.0;
        return 0.0;
    }
    const dd c = cos_dd({w0, 0.0});
    const dd s = sin_dd({w0, 0.0});
    if (sin_w0) *sin_w0 = round_dd(s);
    const dd diff = add(c, {-d, 0.0});  // cos(w0) - d
    const double corr = round_dd(diff) / round_dd(s);
    return round_dd(add({w0, 0.0}, {corr, 0.0}));
}

// qt.rs:41-47
LKL_HD void normalize4(double& w, double& x, double& y, double& z) {
#pragma clang fp contract(off)
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    w /= n;
    x /= n;
    y /= n;
    z /= n;
}

// qt.rs:80-104 (linear threshold 0.9995, constants.rs:11) with the correctly
// rounded acos/sin; a and b are (w, x, y, z).
LKL_HD void slerp_cr(const double* a_in, const double* b_in, double t, double* out) {
#pragma clang fp contract(off)
    double aw = a_in[0], ax = a_in[1], ay = a_in[2], az = a_in[3];
    double bw = b_in[0], bx = b_in[1], by = b_in[2], bz = b_in[3];
    normalize4(aw, ax, ay, az);
    normalize4(bw, bx, by, bz);
    double d = aw * bw + ax * bx + ay * by + az * bz;
    if (d < 0.0) {
        aw = -aw; ax = -ax; ay = -ay; az = -az;
        d *= -1.0;
    }
    if (d > 0.9995) {
        double rw = aw + (bw - aw) * t, rx = ax + (bx - ax) * t, ry = ay + (by - ay) * t,
               rz = az + (bz - az) * t;
        normalize4(rw, rx, ry, rz);
        out[0] = rw; out[1] = rx; out[2] = ry; out[3] = rz;
        return;
    }
    d = std::fmax(std::fmin(d, 1.0), -1.0);
    // sin_cr of an argument already evaluated is reused: sin(om) when om is
    // the library acos (its series ran inside acos_cr), and sin((1 - t) om)
    // when it equals sin(t om)'s argument (t = 1/2)
    double w0, sw0;
    const double om = acos_cr(d, &w0, &sw0);
    const double so = om == w0 ? sw0 : sin_cr(om);
    const double a1 = (1.0 - t) * om, a2 = t * om;
    const double sa1 = sin_cr(a1), sa2 = a2 == a1 ? sa1 : sin_cr(a2);
    const double s1 = sa1 / so, s2 = sa2 / so;
    out[0] = aw * s1 + bw * s2;
    out[1] = ax * s1 + bx * s2;
    out[2] = ay * s1 + by * s2;
    out[3] = az * s1 + bz * s2;
}

#undef LKL_HD

}  // namespace gm
}  // namespace lkl

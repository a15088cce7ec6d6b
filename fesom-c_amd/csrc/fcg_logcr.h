// Correctly rounded natural logarithm in FP64, for the one logarithm of the
// FP64 step: d3_end's bottom boundary value bott_R = 16.6 kappa^2 /
// log(delta_b/z0b_min)^2 (src/fv_mixing.F90:235-238).  The reference gets
// log from the host's libm (glibc), which rounds correctly for all but a
// few in 10^4 arguments; the device library's log is within 1 ulp but not
// correctly rounded.  Where d3_end's iteration does not converge in its
// 200 sweeps (a column a few centimetres deep at a drying shoreline) one ulp
// in bott_R grows into a percent of bt, so the FP64 step evaluates this
// log correctly rounded instead.
//
// Method: x = 2^k m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
// s = (m-1)/(m+1), |s| <= 0.1716, summed as s * sum_n s^(2n)/(2n+1) for
// n = 0..22 in double-double arithmetic (Dekker/Knuth error-free products
// and sums through fma), plus k ln 2 with ln 2 split into two doubles.  The
// double-double result is accurate to about 2^-104 relative, so rounding it
// to double gives the correctly rounded log except for arguments within
// that distance of a rounding tie.  Arguments that are not positive, finite
// and normal go to the library log.
//
// The header compiles for the device and, for its CPU check, for the host.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define FCG_HD __host__ __device__ __forceinline__
#else
#define FCG_HD inline
#endif

namespace fcglog {

struct dd {
    double hi, lo;
};

FCG_HD dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
FCG_HD dd fast_two_sum(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}
FCG_HD dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}
FCG_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
FCG_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
FCG_HD dd dd_div(dd a, dd b) {  // a/b, two correction steps of the double quotient
    const double q1 = a.hi / b.hi;
    dd r = dd_add(a, dd_mul(dd{-q1, 0.0}, b));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul(dd{-q2, 0.0}, b));
    const double q3 = r.hi / b.hi;
    return dd_add(fast_two_sum(q1, q2), dd{q3, 0.0});
}

FCG_HD double log_cr(double x) {
    if (!(x > 0.0) || !(x < INFINITY) || x < 2.2250738585072014e-308) return std::log(x);
    int k = 0;
    double m = std::frexp(x, &k);  // x = m 2^k, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        k -= 1;
    }
    const dd num = {m - 1.0, 0.0};            // exact: m within a factor 2 of 1
    const dd den = two_sum(m, 1.0);
    const dd s = dd_div(num, den);
    const dd s2 = dd_mul(s, s);
    // sum_{n=0}^{22} s^(2n)/(2n+1) by Horner from the highest term; each
    // coefficient 1/(2n+1) as a double-double (its residual through fma)
    dd p = {0.0, 0.0};
    for (int n = 22; n >= 0; --n) {
        const double d = (double)(2 * n + 1);
        const double c = 1.0 / d;
        const dd cn = {c, std::fma(-c, d, 1.0) / d};
        p = dd_add(dd_mul(p, s2), cn);
    }
    dd lm = dd_mul(s, p);
    lm.hi *= 2.0;
    lm.lo *= 2.0;
    // ln 2 = 0x1.62e42fefa39efp-1 + 0x1.abc9e3b39803fp-56
    const dd l2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
    const dd kl = dd_add(two_prod((double)k, l2.hi), dd{(double)k * l2.lo, 0.0});
    const dd r = dd_add(kl, lm);
    return r.hi + r.lo;
}

}  // namespace fcglog

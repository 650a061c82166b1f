// Fast FP64 evaluation of the integrated Green's function's corner term for
// the mixed-precision solve.
//
// The corner term of FFTPoissonSolver::integratedGreensFunction
// (src/Solvers/FFTPoissonSolver.cpp:479-489) is
//   F = -z^2/2 atan(xy/(zr)) - y^2/2 atan(xz/(yr)) - x^2/2 atan(yz/(xr))
//       + yz log(x+r) + xz log(y+r) + xy log(z+r),   r = |(x, y, z)|.
// Its 8-corner difference cancels ~(r/h)^3 of the magnitude, so F needs close
// to full FP64 accuracy even when the transform that follows is FP32 (a
// relative error e in F moves far-field G by ~e (r/h)^3: at n=128, e=1e-15
// gives ~1e-8, far below FP32 rounding).  The FP64 path uses the correctly
// rounded functions of crmath.hpp, which reproduce the reference's bits; this
// path trades the last ulp for about a third of the instructions:
//   - atan of a quotient p/q with one division: the quotient is reduced to
//     |t| <= tan(pi/8) by the three cases t = |p|/|q|, (|p|-|q|)/(|p|+|q|)
//     (offset pi/4), -|q|/|p| (offset pi/2), then atan(t) = t sum_k (-t^2)^k/(2k+1);
//   - the third atan from the identity
//       atan(xy/(zr)) + atan(xz/(yr)) + atan(yz/(xr)) = sgn(x) sgn(y) sgn(z) pi/2,
//     which holds for all nonzero x, y, z (each term is odd in each
//     coordinate; at x = y = z = 1 the terms are pi/6 each) and, with the
//     signs of signed zeros, gives the quotients' own limits at a zero z;
//   - log(a) = e ln2 + 2 atanh(u), a = m 2^e with m in [sqrt(1/2), sqrt(2)),
//     u = (m-1)/(m+1), atanh(u) = u sum_k u^(2k)/(2k+1).
// Both series use their exact coefficients 1/(2k+1) and are truncated where
// the first omitted term is below 1e-17 relative.  Every operation is an IEEE
// operation (fma, division, sqrt, frexp) that rounds the same way on the host
// and on the device, so a host build returns the device's bits.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "crmath.hpp"

namespace opalsc {
namespace igf {

constexpr double kPi2 = 1.5707963267948966;     // pi/2 rounded
constexpr double kPi4 = 0.78539816339744828;    // pi/4 rounded
constexpr double kTanPi8 = 0.41421356237309503; // tan(pi/8) = sqrt(2)-1 rounded
constexpr double kSqrtHalf = 0.70710678118654757;
// ln2 split so that e*kLn2Hi is exact for |e| < 2^11 (the hi part has 32 bits)
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;

// The reduced argument of atan(p/q) for finite p, q not both zero: atan(p/q) =
// sgn * (off + atan(num/den)) with |num/den| <= tan(pi/8) (rounded).  q = +-0
// gives the limit +-pi/2 with the quotient's sign, p = +-0 gives +-0.
struct AtanArg {
    double num, den, off;
    bool neg;
};
__host__ __device__ inline AtanArg atan_arg(double p, double q) {
    const double ap = fabs(p), aq = fabs(q);
    AtanArg a{ap, aq, 0.0, signbit(p) != signbit(q)};
    if (ap > kTanPi8 * aq) {
        if (aq <= kTanPi8 * ap) {  // |p/q| >= tan(3pi/8): pi/2 - atan(|q|/|p|)
            a.num = -aq;
            a.den = ap;
            a.off = kPi2;
        } else {  // atan(|p|/|q|) = pi/4 + atan((|p|-|q|)/(|p|+|q|))
            a.num = ap - aq;
            a.den = ap + aq;
            a.off = kPi4;
        }
    }
    return a;
}

// atan of the reduced argument, t = num/den.
__host__ __device__ inline double atan_finish(const AtanArg &a, double t) {
    const double s = t * t;
    // |t| <= tan(pi/8): s <= 0.1716, s^21/43 < 2e-18
    double c = 1.0 / 41;
    c = fma(c, -s, 1.0 / 39);
    c = fma(c, -s, 1.0 / 37);
    c = fma(c, -s, 1.0 / 35);
    c = fma(c, -s, 1.0 / 33);
    c = fma(c, -s, 1.0 / 31);
    c = fma(c, -s, 1.0 / 29);
    c = fma(c, -s, 1.0 / 27);
    c = fma(c, -s, 1.0 / 25);
    c = fma(c, -s, 1.0 / 23);
    c = fma(c, -s, 1.0 / 21);
    c = fma(c, -s, 1.0 / 19);
    c = fma(c, -s, 1.0 / 17);
    c = fma(c, -s, 1.0 / 15);
    c = fma(c, -s, 1.0 / 13);
    c = fma(c, -s, 1.0 / 11);
    c = fma(c, -s, 1.0 / 9);
    c = fma(c, -s, 1.0 / 7);
    c = fma(c, -s, 1.0 / 5);
    c = fma(c, -s, 1.0 / 3);
    c = fma(c, -s, 1.0);
    const double v = fma(t, c, a.off);
    return a.neg ? -v : v;
}

// The reduced argument of log(a) for a positive normal a: a = m 2^e with m in
// [sqrt(1/2), sqrt(2)), log(a) = e ln2 + 2 atanh((m-1)/(m+1)).
struct LogArg {
    double num, den;
    int e;
};
__host__ __device__ inline LogArg log_arg(double a) {
    int e;
    double m = frexp(a, &e);  // m in [1/2, 1)
    if (m < kSqrtHalf) {
        m = m * 2.0;
        e = e - 1;
    }
    return LogArg{m - 1.0, m + 1.0, e};
}

// log of the reduced argument, u = num/den.
__host__ __device__ inline double log_finish(const LogArg &a, double u) {
    const double s = u * u;  // |u| <= 0.1716, s <= 0.0295, s^11/23 < 1e-18
    double c = 1.0 / 21;
    c = fma(c, s, 1.0 / 19);
    c = fma(c, s, 1.0 / 17);
    c = fma(c, s, 1.0 / 15);
    c = fma(c, s, 1.0 / 13);
    c = fma(c, s, 1.0 / 11);
    c = fma(c, s, 1.0 / 9);
    c = fma(c, s, 1.0 / 7);
    c = fma(c, s, 1.0 / 5);
    c = fma(c, s, 1.0 / 3);
    c = fma(c, s, 1.0);
    const double de = (double)a.e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, 2.0 * u * c));
}

// The corner term F(x, y, z) / cellVolume of FFTPoissonSolver.cpp:479-489
// with correctly rounded atan and log, in the reference's order of operations
// (the FP64 solve; with shifted z also FFTPoissonSolver.cpp:563-573, 589-599).
__host__ __device__ inline double term_cr(double v0, double v1, double v2, double cellVolume) {
    const double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    const double a0 = crm::atan_cr(v0 * v1 / (v2 * r));
    const double a1 = crm::atan_cr(v0 * v2 / (v1 * r));
    const double a2 = crm::atan_cr(v1 * v2 / (v0 * r));
    const double l0 = crm::log_cr(v0 + r);
    const double l1 = crm::log_cr(v1 + r);
    const double l2 = crm::log_cr(v2 + r);
    double t = -v2 * v2 * a0 / 2;
    t += -v1 * v1 * a1 / 2;
    t += -v0 * v0 * a2 / 2;
    t += v1 * v2 * l0;
    t += v0 * v2 * l1;
    t += v0 * v1 * l2;
    return t / cellVolume;
}

// The corner term F(x, y, z) times inv_volume = 1/(h0 h1 h2), in the
// reference's term order.  The five quotients of the reduced arguments share
// one division (Montgomery's simultaneous inversion: the prefix products of
// the denominators, one reciprocal, then one multiplication per quotient
// back), each quotient within a few ulps.
__host__ __device__ inline double term_fast(double x, double y, double z, double inv_volume) {
    const double r = sqrt(x * x + y * y + z * z);
    const AtanArg A0 = atan_arg(x * y, z * r);
    const AtanArg A1 = atan_arg(x * z, y * r);
    const LogArg L0 = log_arg(x + r), L1 = log_arg(y + r), L2 = log_arg(z + r);
    const double c1 = A0.den * A1.den, c2 = c1 * L0.den, c3 = c2 * L1.den, c4 = c3 * L2.den;
    double inv = 1.0 / c4;
    const double i4 = inv * c3;
    inv = inv * L2.den;
    const double i3 = inv * c2;
    inv = inv * L1.den;
    const double i2 = inv * c1;
    inv = inv * L0.den;
    const double i1 = inv * A0.den, i0 = inv * A1.den;
    const double a0 = atan_finish(A0, A0.num * i0);
    const double a1 = atan_finish(A1, A1.num * i1);
    const bool neg = (signbit(x) != signbit(y)) != signbit(z);
    const double a2 = ((neg ? -kPi2 : kPi2) - a0) - a1;
    double t = -z * z * a0 / 2;
    t += -y * y * a1 / 2;
    t += -x * x * a2 / 2;
    t += y * z * log_finish(L0, L0.num * i2);
    t += x * z * log_finish(L1, L1.num * i3);
    t += x * y * log_finish(L2, L2.num * i4);
    return t * inv_volume;
}

}  // namespace igf
}  // namespace opalsc

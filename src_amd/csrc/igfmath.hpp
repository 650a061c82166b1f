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

// ---------------------------------------------------------------------------
// The Green's function value itself, for the mixed-precision solve
// ---------------------------------------------------------------------------
// G at a grid point is the 8-corner difference of F / cellVolume, i.e. the
// average of 1/|r| over the cell of size h0 x h1 x h2 centred at c.  Far from
// the origin the average has the Taylor expansion (odd moments vanish,
// <s_d^2> = h_d^2/12, <s_d^4> = h_d^4/80)
//   <1/r> = 1/r + sum_d h_d^2/24 d_dd(1/r) + sum_d h_d^4/1920 d_dddd(1/r)
//           + sum_{d<e} h_d^2 h_e^2/576 d_ddee(1/r) + O((h/r)^6 / r),
//   d_dd(1/r)   = (3 x_d^2 - r^2) / r^5,
//   d_dddd(1/r) = 3 (35 x_d^4 - 30 x_d^2 r^2 + 3 r^4) / r^9,
//   d_ddee(1/r) = (105 x_d^2 x_e^2 - 15 r^2 (x_d^2 + x_e^2) + 3 r^4) / r^9.
// At |c| >= kGreenFarCells * max(h) its relative error is below 2e-10 (and
// falls as (h/r)^6), smaller than the rounding the 8-corner difference of
// FP64 corner terms keeps there (~2e-15 (r/h)^3: 8e-12 at 16 cells, 3e-8 at
// 256); nearer cells take the exact difference.
constexpr double kGreenFarCells = 16.0;

struct GreenMoments {
    double a[3];  // h_d^2/24
    double b[3];  // 3 h_d^4/1920
    double c[3];  // h_0^2 h_1^2/576, h_0^2 h_2^2/576, h_1^2 h_2^2/576
    double far2;  // (kGreenFarCells * max h)^2
};
__host__ __device__ inline GreenMoments green_moments(double h0, double h1, double h2) {
    GreenMoments m;
    const double hs[3] = {h0 * h0, h1 * h1, h2 * h2};
    for (int d = 0; d < 3; ++d) {
        m.a[d] = hs[d] / 24;
        m.b[d] = 3 * hs[d] * hs[d] / 1920;
    }
    m.c[0] = hs[0] * hs[1] / 576;
    m.c[1] = hs[0] * hs[2] / 576;
    m.c[2] = hs[1] * hs[2] / 576;
    const double hm = fmax(h0, fmax(h1, h2)) * kGreenFarCells;
    m.far2 = hm * hm;
    return m;
}

// The order-4 cell average of 1/r for a cell centred at (x, y, z).
__host__ __device__ inline double green_far(const GreenMoments &m, double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
    const double ir = 1.0 / sqrt(r2), ir2 = ir * ir;
    const double ir5 = ir * ir2 * ir2, ir9 = ir5 * ir2 * ir2;
    const double r4 = r2 * r2;
    const double s2 = m.a[0] * (3 * x2 - r2) + m.a[1] * (3 * y2 - r2) + m.a[2] * (3 * z2 - r2);
    double s4 = m.b[0] * (35 * x2 * x2 - 30 * x2 * r2 + 3 * r4);
    s4 += m.b[1] * (35 * y2 * y2 - 30 * y2 * r2 + 3 * r4);
    s4 += m.b[2] * (35 * z2 * z2 - 30 * z2 * r2 + 3 * r4);
    s4 += m.c[0] * (105 * x2 * y2 - 15 * r2 * (x2 + y2) + 3 * r4);
    s4 += m.c[1] * (105 * x2 * z2 - 15 * r2 * (x2 + z2) + 3 * r4);
    s4 += m.c[2] * (105 * y2 * z2 - 15 * r2 * (y2 + z2) + 3 * r4);
    return ir + (s2 * ir5 + s4 * ir9);
}

// The same expansion in FP32, for a caller that transforms G in FP32 (the
// mixed solve's x DCT): coordinates in units of the largest cell side hm, so
// that r^-9 stays well inside FP32's range for any cell size (r >= 16 and
// coordinates below ~500 in those units), and the result scaled back by 1/hm.
// The terms add no cancellation beyond 3 x_d^2 - r^2 (whose rounding enters
// at (h/r)^2 relative), so its error is a few FP32 ulps.  Unpacked FP32 and
// FP64 vector instructions issue at the same rate on MI355X; the saving is
// the square root and the division, short sequences in FP32 and Newton
// iterations in FP64 (the mixed x DCT with the octant: 0.237 -> 0.226 ms at
// 256^3).
struct GreenMomentsF {
    float a[3], b[3], c[3];  // as GreenMoments, in units of hm
    float ihm;               // 1/hm
    double rhm;              // 1/hm, for the FP64 coordinates
};
__host__ __device__ inline GreenMomentsF green_moments_f(double h0, double h1, double h2) {
    const double hm = fmax(h0, fmax(h1, h2));
    const double u[3] = {h0 / hm, h1 / hm, h2 / hm};
    const double us[3] = {u[0] * u[0], u[1] * u[1], u[2] * u[2]};
    GreenMomentsF m;
    for (int d = 0; d < 3; ++d) {
        m.a[d] = (float)(us[d] / 24);
        m.b[d] = (float)(3 * us[d] * us[d] / 1920);
    }
    m.c[0] = (float)(us[0] * us[1] / 576);
    m.c[1] = (float)(us[0] * us[2] / 576);
    m.c[2] = (float)(us[1] * us[2] / 576);
    m.ihm = (float)(1.0 / hm);
    m.rhm = 1.0 / hm;
    return m;
}
__host__ __device__ inline float green_far_f(const GreenMomentsF &m, float x, float y, float z) {
    const float x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
    const float ir = 1.0f / sqrtf(r2), ir2 = ir * ir;
    const float ir5 = ir * ir2 * ir2, ir9 = ir5 * ir2 * ir2;
    const float r4 = r2 * r2;
    const float s2 = m.a[0] * (3 * x2 - r2) + m.a[1] * (3 * y2 - r2) + m.a[2] * (3 * z2 - r2);
    float s4 = m.b[0] * (35 * x2 * x2 - 30 * x2 * r2 + 3 * r4);
    s4 += m.b[1] * (35 * y2 * y2 - 30 * y2 * r2 + 3 * r4);
    s4 += m.b[2] * (35 * z2 * z2 - 30 * z2 * r2 + 3 * r4);
    s4 += m.c[0] * (105 * x2 * y2 - 15 * r2 * (x2 + y2) + 3 * r4);
    s4 += m.c[1] * (105 * x2 * z2 - 15 * r2 * (x2 + z2) + 3 * r4);
    s4 += m.c[2] * (105 * y2 * z2 - 15 * r2 * (y2 + z2) + 3 * r4);
    return (ir + (s2 * ir5 + s4 * ir9)) * m.ihm;
}

// G at octant point (i, j, k) of the mixed-precision solve: zshift = 0 and
// second = false for FFTPoissonSolver::integratedGreensFunction (the caller
// applies the G[0,0,0] := G[0,0,1] quirk), otherwise the table of
// shiftedIntGreensFunction (FFTPoissonSolver.cpp:552-603) with z corners
// (k h2 - h2/2) + zshift [- n2 h2].  Near cells: the 8-corner difference of
// the corner terms in the order of FFTPoissonSolver.cpp:510-517.
__host__ __device__ inline double green_point_fast(const GreenMoments &m, int i, int j, int k, double h0, double h1,
                                                   double h2, double zshift, bool shifted, bool second, int n2,
                                                   double inv_volume) {
    const double cx = i * h0, cy = j * h1;
    double cz = k * h2;
    if (shifted) {
        cz = cz + zshift;
        if (second) cz = cz - n2 * h2;
    }
    if (cx * cx + cy * cy + cz * cz >= m.far2) return green_far(m, cx, cy, cz);
    // the corners in the reference's sum order +(1,1,1) +(1,0,0) +(0,1,0) +(0,0,1)
    // -(1,1,0) -(1,0,1) -(0,1,1) -(0,0,0), one nibble (dx dy dz) each; a rolled
    // loop, so this rare path does not size the registers of the common one
    constexpr unsigned kOrder = 0x03561247u;
    double g = 0.0;
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        const unsigned c = (kOrder >> (4 * q)) & 7u;
        const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
        const double v0 = (i + dx) * h0 - h0 / 2, v1 = (j + dy) * h1 - h1 / 2;
        double v2 = (k + dz) * h2 - h2 / 2;
        if (shifted) {
            v2 = v2 + zshift;
            if (second) v2 = v2 - n2 * h2;
        }
        const double t = term_fast(v0, v1, v2, inv_volume);
        g = q < 4 ? g + t : g - t;
    }
    return g;
}

// green_point_fast rounded to FP32 with the far field evaluated in FP32: the
// near/far choice is green_point_fast's own (FP64 coordinates), near points
// take its exact FP64 difference.
__host__ __device__ inline float green_point_mixed(const GreenMoments &m, const GreenMomentsF &mf, int i, int j,
                                                   int k, double h0, double h1, double h2, double zshift,
                                                   bool shifted, bool second, int n2, double inv_volume) {
    const double cx = i * h0, cy = j * h1;
    double cz = k * h2;
    if (shifted) {
        cz = cz + zshift;
        if (second) cz = cz - n2 * h2;
    }
    if (cx * cx + cy * cy + cz * cz >= m.far2)
        return green_far_f(mf, (float)(cx * mf.rhm), (float)(cy * mf.rhm), (float)(cz * mf.rhm));
    return (float)green_point_fast(m, i, j, k, h0, h1, h2, zshift, shifted, second, n2, inv_volume);
}

}  // namespace igf
}  // namespace opalsc

/*
 * opal_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of OPAL 2.0's space-charge PIC step
 * (cloud-in-cell scatter, Hockney open-boundary FFT Poisson solve with the
 * integrated Green's function, CIC gather, Boris kick/push, bunch moments,
 * beam-frame rotation and reference-frame update).  It is the checker that the
 * HIP library is compared against; nothing in the product links or calls it.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Every function follows the operation order of the reference function it
 * cites (file:line relative to the OPAL source tree), because bit-exact
 * agreement of cell indices and interpolation weights depends on that order.
 * Build with -ffp-contract=off (see oracle/Makefile): no FMA contraction.
 *
 * Threads: sequential by default (what the tests use).  orc_set_threads(n)
 * runs the independent loops (FFT lines, Green's function corners, CIC gather,
 * Grad, push, kick, coordinate transformations) on n OpenMP threads with
 * results identical to the sequential ones; the scatter and the moment sums,
 * which accumulate, then split the particles into n contiguous ranges added in
 * range order, so they round differently (deterministically) for n > 1.  This
 * serves the CPU baseline timing on all host cores.
 *
 * Layouts (same as the reference): particle vectors are AoS double[N][3];
 * fields are x-fastest, interior cells only (no guard layer); the vector
 * field E is AoS double[n0*n1*n2][3].
 *
 * Reference-side differences that are deliberate (and documented at the
 * function): the antiparallel branch of getQuaternion draws a random axis in
 * the reference; here it is the fixed axis (1,0,0).  Particles whose cell
 * lookup is non-finite or out of int range are dropped from the deposit and
 * gather zero field (the reference would abort or index out of bounds).
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_C 299792458.0              /* src/Classic/Physics/Physics.h:52 */
#define ORC_EPSILON_0 8.854187817e-12  /* src/Classic/Physics/Physics.h:58 */
#define ORC_PI 3.14159265358979323846  /* src/Classic/Physics/Physics.h:31 */

typedef double complex cplx;

int orc_version(void) { return 1; }

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }
/* an independent loop on g_threads threads (sequential when g_threads is 1) */
#define ORC_PAR _Pragma("omp parallel for num_threads(g_threads) if(g_threads > 1) schedule(static)")

/* ------------------------------------------------------------------------ */
/* Bounds and mesh                                                          */
/* ------------------------------------------------------------------------ */

/* PartBunchBase<T,Dim>::getLocalBounds, src/Classic/Algorithms/PartBunchBase.hpp:1203-1218.
 * Note the `else if`: a value smaller than rmin is never compared with rmax. */
void orc_get_bounds(int64_t N, const double *R, double rmin[3], double rmax[3]) {
    if (N == 0) {
        for (int d = 0; d < 3; ++d) rmin[d] = rmax[d] = 0.0;
        return;
    }
    for (int d = 0; d < 3; ++d) rmin[d] = rmax[d] = R[d];
    for (int64_t i = 1; i < N; ++i) {
        for (int d = 0; d < 3; ++d) {
            double v = R[3 * i + d];
            if (rmin[d] > v) rmin[d] = v;
            else if (rmax[d] < v) rmax[d] = v;
        }
    }
}

/* PartBunchBase<T,Dim>::boundp, src/Classic/Algorithms/PartBunchBase.hpp:802-883
 * (not emitting).  dh = BBOXINCR/100 (src/Structure/FieldSolver.cpp:487).  dc:
 * DC beam, z not padded but given the period length (PartBunchBase.hpp:814,
 * 840-843).  Returns 0, or -1 when h0*h1*h2 <= 0 ("h<0, can not build a mesh"). */
static int boundp_impl(int64_t N, const double *R, const int n[3], double dh, int dc, double period,
                       double hr[3], double origin[3]) {
    double rmin[3], rmax[3];
    orc_get_bounds(N, R, rmin, rmax);
    const int dims = dc ? 2 : 3;
    for (int i = 0; i < dims; ++i) {
        double length = fabs(rmax[i] - rmin[i]);
        if (length < 1e-10) {
            rmax[i] += 1e-10;
            rmin[i] -= 1e-10;
        } else {
            rmax[i] += dh * length;
            rmin[i] -= dh * length;
        }
        hr[i] = (rmax[i] - rmin[i]) / (n[i] - 1);
    }
    if (dc) {
        rmax[2] = rmin[2] + period;
        hr[2] = period / (n[2] - 1);
    }
    if (hr[0] * hr[1] * hr[2] <= 0) return -1;
    for (int i = 0; i < 3; ++i) origin[i] = rmin[i] - hr[i] / 2.0;
    return 0;
}

int orc_boundp(int64_t N, const double *R, const int n[3], double dh,
               double hr[3], double origin[3]) {
    return boundp_impl(N, R, n, dh, 0, 0.0, hr, origin);
}

int orc_boundp_dc(int64_t N, const double *R, const int n[3], double dh, double period,
                  double hr[3], double origin[3]) {
    return boundp_impl(N, R, n, dh, 1, period, hr, origin);
}

/* boundp's update() for a DC beam (ippl/src/Particle/ParticleSpatialLayout.hpp:296-297):
 * ParticlePeriodicBCond (ippl/src/Particle/ParticleBConds.h:38-45, applied by
 * ParticleBConds::apply, ParticleBConds.h:113-120) on z over the region of the
 * mesh's cells, min = 0*h + origin, max = n*h + origin (RegionLayout::convert_index,
 * ippl/src/Region/RegionLayout.hpp:526-556, getVertexPosition,
 * ippl/src/Meshes/UniformCartesian.hpp:1054-1055).  x and y: ParticleNoBCond. */
void orc_wrap_periodic_z(int64_t N, double *R, const double origin[3], const double hr[3], const int n[3]) {
    const double m1 = 0 * hr[2] + origin[2], m2 = n[2] * hr[2] + origin[2];
    for (int64_t i = 0; i < N; ++i) {
        const double t = R[3 * i + 2];
        if (t < m1)
            R[3 * i + 2] = m2 - (m1 - t);
        else if (t >= m2)
            R[3 * i + 2] = m1 + (t - m2);
    }
}

/* ------------------------------------------------------------------------ */
/* Cloud-in-cell                                                            */
/* ------------------------------------------------------------------------ */

/* One dimension of IntCICImpl<3>::scatter/gather's cell lookup:
 *   FindNGP -> UniformCartesian::getVertexBelow (ippl/src/Meshes/UniformCartesian.hpp:1025-1039)
 *   FindPos -> getCellPosition (UniformCartesian.hpp:1084-1097)
 *   FindDelta -> getDeltaCell (UniformCartesian.hpp:1151-1158) = h*1
 *   lgp offset and dpos (ippl/src/Particle/IntCIC.h:837-851).
 * Returns 0 when the lookup is out of int range or non-finite. */
static int cic_dim(double x, double o, double h, int *lgp, double *dpos) {
    double t = (x - o) / h;
    if (!(t > -1073741824.0 && t < 1073741824.0)) return 0;
    int i = (int)t;
    int ngp = (x >= o) ? i : i - 1;
    double g = ((double)ngp + 0.5) * h + o;
    double delta = h * 1;
    int off = 0;
    if (g > x) {
        off = -1;
        g = g - delta;
    }
    *lgp = ngp + off;
    *dpos = (x - g) / delta;
    return 1;
}

/* Per-particle lower grid point (lgp, the cell the 8-point stencil starts at)
 * and normalised offset dpos; valid[i] = 0 for dropped particles. */
void orc_cic_index(int64_t N, const double *R, const double origin[3], const double hr[3],
                   int32_t *lgp, double *dpos, uint8_t *valid) {
    ORC_PAR
    for (int64_t p = 0; p < N; ++p) {
        int ok = 1;
        for (int d = 0; d < 3; ++d) {
            int l = 0;
            double dp = 0.0;
            ok &= cic_dim(R[3 * p + d], origin[d], hr[d], &l, &dp);
            lgp[3 * p + d] = l;
            dpos[3 * p + d] = dp;
        }
        if (valid) valid[p] = (uint8_t)ok;
    }
}

static inline int in_grid(int i, int j, int k, const int n[3]) {
    return i >= 0 && j >= 0 && k >= 0 && i < n[0] && j < n[1] && k < n[2];
}

/* IntCICImpl<3>::scatter, ippl/src/Particle/IntCIC.h:816-860, applied
 * sequentially over particles as ParticleAttrib::scatter does
 * (ippl/src/Particle/ParticleAttrib.h:201-225).  rho is accumulated, not
 * cleared.  Deposits that land outside the interior are dropped: on one rank
 * accumGuardCells has no neighbour to add guard cells into
 * (ippl/src/Field/BareField.hpp:764+).  Returns the number of dropped
 * (particle, corner) deposits. */
static int64_t scatter_range(int64_t p0, int64_t p1, const double *R, const double *pdata,
                             const double origin[3], const double hr[3], const int n[3], double *rho) {
    int64_t dropped = 0;
    for (int64_t p = p0; p < p1; ++p) {
        int l[3];
        double dp[3];
        int ok = 1;
        for (int d = 0; d < 3; ++d) ok &= cic_dim(R[3 * p + d], origin[d], hr[d], &l[d], &dp[d]);
        if (!ok) {
            dropped += 8;
            continue;
        }
        const double q = pdata[p];
        const double a0 = 1 - dp[0], a1 = 1 - dp[1], a2 = 1 - dp[2];
        const double b0 = dp[0], b1 = dp[1], b2 = dp[2];
        double w[8];
        w[0] = a0 * a1 * a2 * q;
        w[1] = b0 * a1 * a2 * q;
        w[2] = a0 * b1 * a2 * q;
        w[3] = b0 * b1 * a2 * q;
        w[4] = a0 * a1 * b2 * q;
        w[5] = b0 * a1 * b2 * q;
        w[6] = a0 * b1 * b2 * q;
        w[7] = b0 * b1 * b2 * q;
        for (int c = 0; c < 8; ++c) {
            int i = l[0] + (c & 1), j = l[1] + ((c >> 1) & 1), k = l[2] + ((c >> 2) & 1);
            if (in_grid(i, j, k, n)) rho[((int64_t)k * n[1] + j) * n[0] + i] += w[c];
            else ++dropped;
        }
    }
    return dropped;
}

int64_t orc_scatter(int64_t N, const double *R, const double *pdata, const double origin[3],
                    const double hr[3], const int n[3], double *rho) {
    const int T = g_threads;
    if (T <= 1 || N < 4096) return scatter_range(0, N, R, pdata, origin, hr, n, rho);
    /* thread t deposits its contiguous range into a grid of its own; the grids
     * are then added to rho in thread order */
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    double *priv = (double *)calloc((size_t)T * ncell, sizeof(double));
    int64_t dropped = 0;
#pragma omp parallel for num_threads(T) schedule(static, 1) reduction(+ : dropped)
    for (int t = 0; t < T; ++t)
        dropped += scatter_range(N * t / T, N * (t + 1) / T, R, pdata, origin, hr, n, priv + (size_t)t * ncell);
#pragma omp parallel for num_threads(T) schedule(static)
    for (size_t c = 0; c < ncell; ++c)
        for (int t = 0; t < T; ++t) rho[c] += priv[(size_t)t * ncell + c];
    free(priv);
    return dropped;
}

/* IntCICImpl<3>::gather, ippl/src/Particle/IntCIC.h:933-977 for a Vektor field:
 * Ef = sum over the 8 corners, in order 000,100,...,111, of ((a*b)*c)*E.
 * Corners outside the interior read the guard value: 0 for ZeroFace; with
 * periodic z faces (pz, a DC beam's PeriodicFace, filled by Field::fillGuardCells
 * before the gather, ippl/src/Field/Field.hpp:366-381) the z guard layers hold
 * the opposite cells: k = n2 reads k = 0, k = -1 reads k = n2-1. */
static void gather_impl(int64_t N, const double *R, const double *E, const double origin[3],
                        const double hr[3], const int n[3], int pz, double *Ef) {
    ORC_PAR
    for (int64_t p = 0; p < N; ++p) {
        int l[3];
        double dp[3];
        int ok = 1;
        for (int d = 0; d < 3; ++d) ok &= cic_dim(R[3 * p + d], origin[d], hr[d], &l[d], &dp[d]);
        if (!ok) {
            Ef[3 * p] = Ef[3 * p + 1] = Ef[3 * p + 2] = 0.0;
            continue;
        }
        const double a0 = 1 - dp[0], a1 = 1 - dp[1], a2 = 1 - dp[2];
        const double b0 = dp[0], b1 = dp[1], b2 = dp[2];
        double w[8];
        w[0] = a0 * a1 * a2;
        w[1] = b0 * a1 * a2;
        w[2] = a0 * b1 * a2;
        w[3] = b0 * b1 * a2;
        w[4] = a0 * a1 * b2;
        w[5] = b0 * a1 * b2;
        w[6] = a0 * b1 * b2;
        w[7] = b0 * b1 * b2;
        for (int d = 0; d < 3; ++d) {
            double acc = 0.0;
            for (int c = 0; c < 8; ++c) {
                int i = l[0] + (c & 1), j = l[1] + ((c >> 1) & 1), k = l[2] + ((c >> 2) & 1);
                if (pz && k == n[2]) k = 0;
                else if (pz && k == -1) k = n[2] - 1;
                double e = in_grid(i, j, k, n) ? E[3 * (((int64_t)k * n[1] + j) * n[0] + i) + d] : 0.0;
                double term = w[c] * e;
                acc = (c == 0) ? term : acc + term;
            }
            Ef[3 * p + d] = acc;
        }
    }
}

void orc_gather(int64_t N, const double *R, const double *E, const double origin[3],
                const double hr[3], const int n[3], double *Ef) {
    gather_impl(N, R, E, origin, hr, n, 0, Ef);
}

void orc_gather_periodic_z(int64_t N, const double *R, const double *E, const double origin[3],
                           const double hr[3], const int n[3], double *Ef) {
    gather_impl(N, R, E, origin, hr, n, 1, Ef);
}

/* ------------------------------------------------------------------------ */
/* FFT (test-only mixed-radix complex FFT; any sign convention gives the    */
/* same circular convolution, which is all the reference's RC/CR pair       */
/* computes: ippl/src/FFT/FFT.hpp:768-775,1214,1308,1571,1673)              */
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

/* Decimation-in-time recursion: X[0..n) = DFT of x[0], x[xs], ... with
 * tw[k*tws] = exp(sign*2*pi*i*k/n). */
static void fft_rec(const cplx *x, size_t xs, cplx *X, int n, const cplx *tw, size_t tws) {
    if (n == 1) {
        X[0] = x[0];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(x + r * xs, xs * p, X + (size_t)r * m, m, tw, tws * p);
    cplx t[64];
    cplx *tt = (p <= 64) ? t : (cplx *)malloc(sizeof(cplx) * p);
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) tt[r] = X[(size_t)r * m + k] * tw[((size_t)r * k % n) * tws];
        for (int q = 0; q < p; ++q) {
            cplx s = 0;
            for (int r = 0; r < p; ++r) s += tt[r] * tw[((size_t)r * q * m % n) * tws];
            X[k + (size_t)q * m] = s;
        }
    }
    if (tt != t) free(tt);
}

/* In-place 1D complex DFT of a contiguous line; sign -1 forward, +1 backward, unnormalised. */
void orc_fft1d(cplx *line, int n, int sign) {
    cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
    cplx *in = (cplx *)malloc(sizeof(cplx) * n);
    for (int k = 0; k < n; ++k) {
        double a = 2.0 * ORC_PI * (double)k / (double)n;
        tw[k] = cos(a) + sign * sin(a) * I;
    }
    memcpy(in, line, sizeof(cplx) * n);
    fft_rec(in, 1, line, n, tw, 1);
    free(in);
    free(tw);
}

/* In-place 3D complex DFT, x-fastest dims m[0..2]. */
void orc_fft3d(cplx *a, const int m[3], int sign) {
    int maxm = m[0] > m[1] ? m[0] : m[1];
    if (m[2] > maxm) maxm = m[2];
    cplx *tw = (cplx *)malloc(sizeof(cplx) * maxm);
    const size_t stride[3] = {1, (size_t)m[0], (size_t)m[0] * m[1]};
    for (int ax = 0; ax < 3; ++ax) {
        int L = m[ax];
        for (int k = 0; k < L; ++k) {
            double ang = 2.0 * ORC_PI * (double)k / (double)L;
            tw[k] = cos(ang) + sign * sin(ang) * I;
        }
        int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        const int64_t nlines = (int64_t)m[o2] * m[o1];
#pragma omp parallel num_threads(g_threads) if (g_threads > 1)
        {
            cplx *line = (cplx *)malloc(sizeof(cplx) * maxm);
            cplx *in = (cplx *)malloc(sizeof(cplx) * maxm);
#pragma omp for schedule(static)
            for (int64_t uv = 0; uv < nlines; ++uv) {
                const int u = (int)(uv / m[o1]), v = (int)(uv % m[o1]);
                size_t base = (size_t)u * stride[o2] + (size_t)v * stride[o1];
                for (int k = 0; k < L; ++k) in[k] = a[base + (size_t)k * stride[ax]];
                fft_rec(in, 1, line, L, tw, 1);
                for (int k = 0; k < L; ++k) a[base + (size_t)k * stride[ax]] = line[k];
            }
            free(line);
            free(in);
        }
    }
    free(tw);
}

/* ------------------------------------------------------------------------ */
/* Green's functions                                                        */
/* ------------------------------------------------------------------------ */

/* atan and log inside the IGF corner term.  The 8-corner difference cancels
 * ~(r/h)^3 of the corner magnitude, so one-ulp libm differences reach ~1e-9 of
 * far-field G and ~1e-11 of E at n=128: two implementations only agree to
 * the 1e-12 parity bar if their atan/log agree bit for bit.  Mode 1 (default)
 * evaluates them in binary128 and rounds once to double, i.e. correctly
 * rounded values, which no libm choice can dispute; mode 0 uses the host libm
 * as the reference binary does (glibc 2.35 here: not correctly rounded for
 * ~0.2% of atan and ~0.01% of log arguments). */
static int g_igf_libm_mode = 1;
void orc_set_igf_libm(int mode) { g_igf_libm_mode = mode; }
static double igf_atan(double x) {
    return g_igf_libm_mode ? (double)atanq((__float128)x) : atan(x);
}
static double igf_log(double x) {
    return g_igf_libm_mode ? (double)logq((__float128)x) : log(x);
}
double orc_atan_cr(double x) { return (double)atanq((__float128)x); }
double orc_log_cr(double x) { return (double)logq((__float128)x); }

/* Antiderivative term of FFTPoissonSolver::integratedGreensFunction,
 * src/Solvers/FFTPoissonSolver.cpp:472-489, at corner (i,j,k). */
static double igf_corner(int i, int j, int k, const double hr[3], double cellVolume) {
    double v0 = i * hr[0] - hr[0] / 2;
    double v1 = j * hr[1] - hr[1] / 2;
    double v2 = k * hr[2] - hr[2] / 2;
    double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    double t = -v2 * v2 * igf_atan(v0 * v1 / (v2 * r)) / 2;
    t += -v1 * v1 * igf_atan(v0 * v2 / (v1 * r)) / 2;
    t += -v0 * v0 * igf_atan(v1 * v2 / (v0 * r)) / 2;
    t += v1 * v2 * igf_log(v0 + r);
    t += v0 * v2 * igf_log(v1 + r);
    t += v0 * v1 * igf_log(v2 + r);
    return t / cellVolume;
}

/* Integrated Green's function octant G[i,j,k], i in [0,n0], j in [0,n1],
 * k in [0,n2] (x-fastest, (n0+1)(n1+1)(n2+1) values):
 *   corners T (src/Solvers/FFTPoissonSolver.cpp:469-495),
 *   8-corner difference in the reference's order (FFTPoissonSolver.cpp:510-517),
 *   and, if apply_quirk, G[0,0,0] := G[0,0,1] (FFTPoissonSolver.cpp:657).
 * Index n needs T at n+1, which the reference reads from outside its
 * tmpgreen field; any value there is harmless because the physical potential
 * never reads G at index n (x-y spans (-n,n) on the 2n-periodic grid).  Here T
 * at n+1 is the same formula. */
void orc_igf_octant(const int n[3], const double hr[3], double *G, int apply_quirk) {
    const int t0 = n[0] + 2, t1 = n[1] + 2, t2 = n[2] + 2;
    double cellVolume = hr[0] * hr[1] * hr[2];
    double *T = (double *)malloc(sizeof(double) * (size_t)t0 * t1 * t2);
    ORC_PAR
    for (int k = 0; k < t2; ++k)
        for (int j = 0; j < t1; ++j)
            for (int i = 0; i < t0; ++i)
                T[((size_t)k * t1 + j) * t0 + i] = igf_corner(i, j, k, hr, cellVolume);
#define TT(a, b, c) T[((size_t)(c) * t1 + (b)) * t0 + (a)]
    const int g0 = n[0] + 1, g1 = n[1] + 1, g2 = n[2] + 1;
    ORC_PAR
    for (int k = 0; k < g2; ++k)
        for (int j = 0; j < g1; ++j)
            for (int i = 0; i < g0; ++i) {
                double g = TT(i + 1, j + 1, k + 1);
                g += TT(i + 1, j, k);
                g += TT(i, j + 1, k);
                g += TT(i, j, k + 1);
                g -= TT(i + 1, j + 1, k);
                g -= TT(i + 1, j, k + 1);
                g -= TT(i, j + 1, k + 1);
                g -= TT(i, j, k);
                G[((size_t)k * g1 + j) * g0 + i] = g;
            }
#undef TT
    free(T);
    if (apply_quirk) G[0] = G[(size_t)1 * g1 * g0];
}

/* STANDARD Green's function octant, SpecializedGreensFunction<3>::calculate
 * (src/Solvers/FFTPoissonSolver.cpp:32-43) with grnIField (FFTPoissonSolver.cpp:228-234):
 * G = 1/sqrt(i^2 h0^2 + j^2 h1^2 + k^2 h2^2), G[0,0,0] := G[0,0,1]. */
void orc_green_standard_octant(const int n[3], const double hr[3], double *G) {
    const int g0 = n[0] + 1, g1 = n[1] + 1, g2 = n[2] + 1;
    double hsq0 = hr[0] * hr[0], hsq1 = hr[1] * hr[1], hsq2 = hr[2] * hr[2];
    for (int k = 0; k < g2; ++k)
        for (int j = 0; j < g1; ++j)
            for (int i = 0; i < g0; ++i) {
                double g = (double)(i * i) * hsq0 + (double)(j * j) * hsq1 + (double)(k * k) * hsq2;
                G[((size_t)k * g1 + j) * g0 + i] = 1.0 / sqrt(g);
            }
    G[0] = G[(size_t)1 * g1 * g0];
}

/* Even mirror of the octant onto the doubled grid (2n0,2n1,2n2):
 * FFTPoissonSolver::mirrorRhoField, src/Solvers/FFTPoissonSolver.cpp:641-664. */
void orc_mirror_green(const int n[3], const double *Goct, double *G2) {
    const int m0 = 2 * n[0], m1 = 2 * n[1], m2 = 2 * n[2];
    const int g0 = n[0] + 1, g1 = n[1] + 1;
    ORC_PAR
    for (int z = 0; z < m2; ++z) {
        int fz = z <= n[2] ? z : m2 - z;
        for (int y = 0; y < m1; ++y) {
            int fy = y <= n[1] ? y : m1 - y;
            for (int x = 0; x < m0; ++x) {
                int fx = x <= n[0] ? x : m0 - x;
                G2[((size_t)z * m1 + y) * m0 + x] = Goct[((size_t)fz * g1 + fy) * g0 + fx];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Poisson solve                                                            */
/* ------------------------------------------------------------------------ */

/* Circular convolution on the doubled grid of the zero-padded rho with the
 * doubled Green's function; rho_inout holds rho on entry and phi on exit
 * (x-fastest n0*n1*n2).  FFTPoissonSolver::computePotential,
 * src/Solvers/FFTPoissonSolver.cpp:337-382 (forward transform of rho2 and of
 * G unnormalised, pointwise product, backward transform scaled by 1/M). */
void orc_convolve_doubled(const int n[3], double *rho_inout, const double *G2) {
    const int m[3] = {2 * n[0], 2 * n[1], 2 * n[2]};
    const size_t M = (size_t)m[0] * m[1] * m[2];
    cplx *a = (cplx *)calloc(M, sizeof(cplx));
    cplx *g = (cplx *)malloc(M * sizeof(cplx));
    ORC_PAR
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                a[((size_t)k * m[1] + j) * m[0] + i] = rho_inout[((size_t)k * n[1] + j) * n[0] + i];
    ORC_PAR
    for (size_t q = 0; q < M; ++q) g[q] = G2[q];
    orc_fft3d(a, m, -1);
    orc_fft3d(g, m, -1);
    ORC_PAR
    for (size_t q = 0; q < M; ++q) a[q] *= g[q];
    orc_fft3d(a, m, +1);
    double inv = 1.0 / (double)M;
    ORC_PAR
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                rho_inout[((size_t)k * n[1] + j) * n[0] + i] = creal(a[((size_t)k * m[1] + j) * m[0] + i]) * inv;
    free(a);
    free(g);
}

/* FFTPoissonSolver::computePotential(rho, hr), FFTPoissonSolver.cpp:337-382,
 * with the Green's function chosen by `greens` (0 INTEGRATED, 1 STANDARD).
 * In place: rho in, phi out; no h^3 or 1/(4 pi eps0) factors. */
void orc_compute_potential(const int n[3], double *rho_inout, const double hr[3], int greens) {
    const size_t go = (size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    const size_t M = (size_t)8 * n[0] * n[1] * n[2];
    double *Goct = (double *)malloc(go * sizeof(double));
    double *G2 = (double *)malloc(M * sizeof(double));
    if (greens == 1) orc_green_standard_octant(n, hr, Goct);
    else orc_igf_octant(n, hr, Goct, 1);
    orc_mirror_green(n, Goct, G2);
    orc_convolve_doubled(n, rho_inout, G2);
    free(Goct);
    free(G2);
}

/* ------------------------------------------------------------------------ */
/* Periodic-z solve (DC beams, FFTPoissonSolver with bcz_m)                 */
/* ------------------------------------------------------------------------ */

/* The Green's function on the (2n0, 2n1, n2) grid of the periodic-z solver:
 * the octant (with the [0,0,0] quirk) mirrored in x and y only, z read one-sided
 * for k in [0, n2) (FFTPoissonSolver::initializeFields, FFTPoissonSolver.cpp:130-155:
 * z not doubled; mirrorRhoField, FFTPoissonSolver.cpp:661-662: no z mirror). */
void orc_mirror_green_periodic_z(const int n[3], const double *Goct, double *G2) {
    const int m0 = 2 * n[0], m1 = 2 * n[1], m2 = n[2];
    const int g0 = n[0] + 1, g1 = n[1] + 1;
    for (int z = 0; z < m2; ++z)
        for (int y = 0; y < m1; ++y) {
            int fy = y <= n[1] ? y : m1 - y;
            for (int x = 0; x < m0; ++x) {
                int fx = x <= n[0] ? x : m0 - x;
                G2[((size_t)z * m1 + y) * m0 + x] = Goct[((size_t)z * g1 + fy) * g0 + fx];
            }
        }
}

/* FFTPoissonSolver::computePotential(rho, hr) with bcz_m
 * (FFTPoissonSolver.cpp:337-382 over the grid of FFTPoissonSolver.cpp:130-155):
 * circular convolution on (2n0, 2n1, n2) of rho zero-padded in x and y with the
 * Green's function above, scaled by 1/(2n0 2n1 n2).  In place: rho in, phi out. */
void orc_compute_potential_periodic_z(const int n[3], double *rho_inout, const double hr[3], int greens) {
    const int m[3] = {2 * n[0], 2 * n[1], n[2]};
    const size_t go = (size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    const size_t M = (size_t)m[0] * m[1] * m[2];
    double *Goct = (double *)malloc(go * sizeof(double));
    double *G2 = (double *)malloc(M * sizeof(double));
    if (greens == 1) orc_green_standard_octant(n, hr, Goct);
    else orc_igf_octant(n, hr, Goct, 1);
    orc_mirror_green_periodic_z(n, Goct, G2);
    cplx *a = (cplx *)calloc(M, sizeof(cplx));
    cplx *g = (cplx *)malloc(M * sizeof(cplx));
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                a[((size_t)k * m[1] + j) * m[0] + i] = rho_inout[((size_t)k * n[1] + j) * n[0] + i];
    for (size_t q = 0; q < M; ++q) g[q] = G2[q];
    orc_fft3d(a, m, -1);
    orc_fft3d(g, m, -1);
    for (size_t q = 0; q < M; ++q) a[q] *= g[q];
    orc_fft3d(a, m, +1);
    double inv = 1.0 / (double)M;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                rho_inout[((size_t)k * n[1] + j) * n[0] + i] = creal(a[((size_t)k * m[1] + j) * m[0] + i]) * inv;
    free(a);
    free(g);
    free(Goct);
    free(G2);
}

/* ------------------------------------------------------------------------ */
/* Image-charge (shifted) solve                                             */
/* ------------------------------------------------------------------------ */

/* Corner antiderivative of FFTPoissonSolver::shiftedIntGreensFunction,
 * src/Solvers/FFTPoissonSolver.cpp:552-603: the IGF corner term with the z
 * coordinate (k*h2 - h2/2) + zshift, and for the second table additionally
 * - n2*h2 (left-to-right evaluation as in the reference). */
static double igf_corner_shifted(int i, int j, int k, const double hr[3], double cellVolume,
                                 double zshift, int second, int n2) {
    double v0 = i * hr[0] - hr[0] / 2;
    double v1 = j * hr[1] - hr[1] / 2;
    double v2 = k * hr[2] - hr[2] / 2 + zshift;
    if (second) v2 = v2 - n2 * hr[2];
    double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    double t = -v2 * v2 * igf_atan(v0 * v1 / (v2 * r)) / 2;
    t += -v1 * v1 * igf_atan(v0 * v2 / (v1 * r)) / 2;
    t += -v0 * v0 * igf_atan(v1 * v2 / (v0 * r)) / 2;
    t += v1 * v2 * igf_log(v0 + r);
    t += v0 * v2 * igf_log(v1 + r);
    t += v0 * v1 * igf_log(v2 + r);
    return t / cellVolume;
}

/* The shifted Green's function on the doubled grid (2n0,2n1,2n2), x-fastest:
 * two corner tables (FFTPoissonSolver.cpp:552-603), their 8-corner differences
 * on [0,n]^3 (FFTPoissonSolver.cpp:621-639, no [0,0,0] quirk), the first for
 * z in [0,n2], the second shifted to z in [n2+1, 2n2) (no z mirror), both
 * mirrored evenly in x and y (mirrorRhoField(Field_t&), FFTPoissonSolver.cpp:666-691). */
void orc_shifted_green_doubled(const int n[3], const double hr[3], double zshift, double *G2) {
    const int t0 = n[0] + 2, t1 = n[1] + 2, t2 = n[2] + 2;
    const int g0 = n[0] + 1, g1 = n[1] + 1, g2 = n[2] + 1;
    const double cellVolume = hr[0] * hr[1] * hr[2];
    const size_t nt = (size_t)t0 * t1 * t2, ng = (size_t)g0 * g1 * g2;
    double *T = (double *)malloc(sizeof(double) * nt);
    double *Ga = (double *)malloc(sizeof(double) * ng);
    double *Gb = (double *)malloc(sizeof(double) * ng);
    for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < t2; ++k)
            for (int j = 0; j < t1; ++j)
                for (int i = 0; i < t0; ++i)
                    T[((size_t)k * t1 + j) * t0 + i] = igf_corner_shifted(i, j, k, hr, cellVolume, zshift, pass, n[2]);
        double *G = pass ? Gb : Ga;
#define TT(a, b, c) T[((size_t)(c) * t1 + (b)) * t0 + (a)]
        for (int k = 0; k < g2; ++k)
            for (int j = 0; j < g1; ++j)
                for (int i = 0; i < g0; ++i) {
                    double g = TT(i + 1, j + 1, k + 1);
                    g += TT(i + 1, j, k);
                    g += TT(i, j + 1, k);
                    g += TT(i, j, k + 1);
                    g -= TT(i + 1, j + 1, k);
                    g -= TT(i + 1, j, k + 1);
                    g -= TT(i, j + 1, k + 1);
                    g -= TT(i, j, k);
                    G[((size_t)k * g1 + j) * g0 + i] = g;
                }
#undef TT
    }
    const int m0 = 2 * n[0], m1 = 2 * n[1], m2 = 2 * n[2];
    for (int z = 0; z < m2; ++z) {
        const double *G = z <= n[2] ? Ga : Gb;
        const int fz = z <= n[2] ? z : z - n[2];
        for (int y = 0; y < m1; ++y) {
            const int fy = y <= n[1] ? y : m1 - y;
            for (int x = 0; x < m0; ++x) {
                const int fx = x <= n[0] ? x : m0 - x;
                G2[((size_t)z * m1 + y) * m0 + x] = G[((size_t)fz * g1 + fy) * g0 + fx];
            }
        }
    }
    free(T);
    free(Ga);
    free(Gb);
}

/* FFTPoissonSolver::computePotential(rho, hr, zshift), FFTPoissonSolver.cpp:289-330:
 * circular convolution with the shifted Green's function, negated (image
 * charge), and the result flipped in z: phi[i][j][k] = conv[i][j][n2-1-k].
 * In place: rho in, image potential out. */
void orc_compute_potential_shifted(const int n[3], double *rho_inout, const double hr[3], double zshift) {
    const size_t M = (size_t)8 * n[0] * n[1] * n[2];
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    double *G2 = (double *)malloc(M * sizeof(double));
    orc_shifted_green_doubled(n, hr, zshift, G2);
    for (size_t q = 0; q < M; ++q) G2[q] = -G2[q];  /* -(rho^ * G^) == rho^ * (-G)^ exactly */
    double *conv = (double *)malloc(ncell * sizeof(double));
    memcpy(conv, rho_inout, ncell * sizeof(double));
    orc_convolve_doubled(n, conv, G2);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                rho_inout[((size_t)k * n[1] + j) * n[0] + i] = conv[((size_t)(n[2] - k - 1) * n[1] + j) * n[0] + i];
    free(G2);
    free(conv);
}

/* ------------------------------------------------------------------------ */
/* Gradient                                                                 */
/* ------------------------------------------------------------------------ */

/* Grad Cell->Cell, ippl/src/Meshes/UniformCartesian.hpp:2127-2200: central
 * differences (guard value 0 from ZeroFace, ippl/src/Field/BCond.h:594-605),
 * then the border faces of component d replaced by one-sided second-order
 * differences.  Output r = +Grad (the caller negates). */
void orc_grad(const int n[3], const double *x, const double hr[3], double *r) {
    const int n0 = n[0], n1 = n[1], n2 = n[2];
    const double id0 = 0.5 / hr[0], id1 = 0.5 / hr[1], id2 = 0.5 / hr[2];
#define XV(i, j, k) (((i) < 0 || (j) < 0 || (k) < 0 || (i) >= n0 || (j) >= n1 || (k) >= n2) ? 0.0 : x[((size_t)(k) * n1 + (j)) * n0 + (i)])
#define RR(i, j, k, d) r[3 * (((size_t)(k) * n1 + (j)) * n0 + (i)) + (d)]
    ORC_PAR
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                double sx = XV(i + 1, j, k) - XV(i - 1, j, k);
                double sy = XV(i, j + 1, k) - XV(i, j - 1, k);
                double sz = XV(i, j, k + 1) - XV(i, j, k - 1);
                RR(i, j, k, 0) = (id0 * sx + 0.0 * sy) + 0.0 * sz;
                RR(i, j, k, 1) = (0.0 * sx + id1 * sy) + 0.0 * sz;
                RR(i, j, k, 2) = (0.0 * sx + 0.0 * sy) + id2 * sz;
            }
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j) {
            double s = (-1.0 * XV(2, j, k) + 4.0 * XV(1, j, k)) - 3.0 * XV(0, j, k);
            RR(0, j, k, 0) = id0 * s;
            s = (1.0 * XV(n0 - 3, j, k) - 4.0 * XV(n0 - 2, j, k)) + 3.0 * XV(n0 - 1, j, k);
            RR(n0 - 1, j, k, 0) = id0 * s;
        }
    for (int k = 0; k < n2; ++k)
        for (int i = 0; i < n0; ++i) {
            double s = (-1.0 * XV(i, 2, k) + 4.0 * XV(i, 1, k)) - 3.0 * XV(i, 0, k);
            RR(i, 0, k, 1) = id1 * s;
            s = (1.0 * XV(i, n1 - 3, k) - 4.0 * XV(i, n1 - 2, k)) + 3.0 * XV(i, n1 - 1, k);
            RR(i, n1 - 1, k, 1) = id1 * s;
        }
    for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
            double s = (-1.0 * XV(i, j, 2) + 4.0 * XV(i, j, 1)) - 3.0 * XV(i, j, 0);
            RR(i, j, 0, 2) = id2 * s;
            s = (1.0 * XV(i, j, n2 - 3) - 4.0 * XV(i, j, n2 - 2)) + 3.0 * XV(i, j, n2 - 1);
            RR(i, j, n2 - 1, 2) = id2 * s;
        }
#undef XV
#undef RR
}

/* ------------------------------------------------------------------------ */
/* Self fields                                                              */
/* ------------------------------------------------------------------------ */

/* PartBunch::computeSelfFields(), src/Classic/Algorithms/PartBunch.cpp:482-648.
 * Q is modified in place exactly as the reference does (Q *= dt; scatter;
 * Q /= dt).  Ef is overwritten, Bf accumulated.  Optional outputs rho_out
 * (scaled charge density that enters the solver), phi_out (potential in V)
 * and E_out (AoS grid field after the Lorentz scaling) may be NULL.
 * Returns the number of dropped deposits. */
static int64_t self_fields_impl(int64_t N, const double *R, const double *P, double *Q,
                                const double *dt, double dT, const double origin[3],
                                const double hr[3], const int n[3], double coupling, int greens, int pz,
                                double *Ef, double *Bf, double *rho_out, double *phi_out,
                                double *E_out) {
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    double *rho = (double *)calloc(ncell, sizeof(double));
    double *E = (double *)malloc(3 * ncell * sizeof(double));
    for (int64_t i = 0; i < N; ++i) Q[i] *= dt[i];
    int64_t dropped = orc_scatter(N, R, Q, origin, hr, n, rho);
    for (int64_t i = 0; i < N; ++i) Q[i] /= dt[i];
    for (size_t c = 0; c < ncell; ++c) rho[c] /= dT;

    double gammaz = 0.0;
    for (int64_t i = 0; i < N; ++i) gammaz += P[3 * i + 2];
    gammaz /= (double)N;
    gammaz *= gammaz;
    gammaz = sqrt(gammaz + 1.0);
    double hs[3] = {hr[0] * 1.0, hr[1] * 1.0, hr[2] * 1.0};
    hs[2] *= gammaz;
    double tmp2 = 1 / hs[0] * 1 / hs[1] * 1 / hs[2];
    for (size_t c = 0; c < ncell; ++c) rho[c] *= tmp2;
    if (rho_out) memcpy(rho_out, rho, ncell * sizeof(double));

    if (pz)
        orc_compute_potential_periodic_z(n, rho, hs, greens);
    else
        orc_compute_potential(n, rho, hs, greens);

    double vol = hs[0] * hs[1] * hs[2];
    for (size_t c = 0; c < ncell; ++c) rho[c] *= vol;
    for (size_t c = 0; c < ncell; ++c) rho[c] *= coupling;
    if (phi_out) memcpy(phi_out, rho, ncell * sizeof(double));

    orc_grad(n, rho, hr, E);
    const double sc[3] = {gammaz / 1.0, gammaz / 1.0, 1.0 / (1.0 * gammaz)};
    for (size_t c = 0; c < ncell; ++c)
        for (int d = 0; d < 3; ++d) E[3 * c + d] = -E[3 * c + d] * sc[d];
    if (E_out) memcpy(E_out, E, 3 * ncell * sizeof(double));

    gather_impl(N, R, E, origin, hr, n, pz, Ef);

    double betaC = sqrt(gammaz * gammaz - 1.0) / gammaz / ORC_C;
    for (int64_t i = 0; i < N; ++i) {
        Bf[3 * i + 0] = Bf[3 * i + 0] - betaC * Ef[3 * i + 1];
        Bf[3 * i + 1] = Bf[3 * i + 1] + betaC * Ef[3 * i + 0];
    }
    free(rho);
    free(E);
    return dropped;
}

int64_t orc_compute_self_fields(int64_t N, const double *R, const double *P, double *Q,
                                const double *dt, double dT, const double origin[3],
                                const double hr[3], const int n[3], double coupling, int greens,
                                double *Ef, double *Bf, double *rho_out, double *phi_out,
                                double *E_out) {
    return self_fields_impl(N, R, P, Q, dt, dT, origin, hr, n, coupling, greens, 0, Ef, Bf, rho_out,
                            phi_out, E_out);
}

/* The same for a DC beam (setBCForDCBeam, src/Classic/Algorithms/PartBunch.cpp:1079-1098):
 * periodic-z solve (FFTPoissonSolver.cpp:130-155) and periodic z guards for
 * the gather; deposits beyond the z faces are dropped (accumGuardCells has no
 * neighbour field on one rank, ippl/src/Field/BareField.hpp:764-960). */
int64_t orc_compute_self_fields_periodic_z(int64_t N, const double *R, const double *P, double *Q,
                                           const double *dt, double dT, const double origin[3],
                                           const double hr[3], const int n[3], double coupling, int greens,
                                           double *Ef, double *Bf, double *rho_out, double *phi_out,
                                           double *E_out) {
    return self_fields_impl(N, R, P, Q, dt, dT, origin, hr, n, coupling, greens, 1, Ef, Bf, rho_out,
                            phi_out, E_out);
}

/* PartBunchBase::calcGammas, src/Classic/Algorithms/PartBunchBase.hpp:639-682:
 * mean of sqrt(1 + P.P) over each bin's particles, summed sequentially in
 * particle order; an empty bin gets 1.  Bins outside [0, nbins) are ignored. */
void orc_calc_gammas(int64_t N, const double *P, const int *bin, int nbins, double *gamma) {
    int64_t *cnt = (int64_t *)calloc((size_t)nbins, sizeof(int64_t));
    for (int b = 0; b < nbins; ++b) gamma[b] = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const int b = bin[i];
        if (b < 0 || b >= nbins) continue;
        gamma[b] += sqrt(1.0 + (P[3 * i] * P[3 * i] + P[3 * i + 1] * P[3 * i + 1] + P[3 * i + 2] * P[3 * i + 2]));
        ++cnt[b];
    }
    for (int b = 0; b < nbins; ++b) gamma[b] = cnt[b] != 0 ? gamma[b] / (double)cnt[b] : 1.0;
    free(cnt);
}

/* Field of one bin (and of its image charge) on the grid, AoS E, after the
 * Lorentz scaling: the body of PartBunch::computeSelfFields(int),
 * src/Classic/Algorithms/PartBunch.cpp:138-430, up to each gather.
 * rho: the bin's charge density exactly as scattered ((qi*dt) deposits, then
 * / dT); E_self and E_img receive -Grad(phi)*(g, g, 1/g). */
static void binned_fields(const int n[3], const double *rho_scattered, const double origin[3],
                          const double hr[3], double gammaz, double coupling, int greens, int image,
                          double *E_self, double *E_img) {
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    const double scaleFactor = 1.0;
    double *rho = (double *)malloc(ncell * sizeof(double));
    double *img = (double *)malloc(ncell * sizeof(double));
    double *E = (double *)malloc(3 * ncell * sizeof(double));
    memcpy(rho, rho_scattered, ncell * sizeof(double));
    /* PartBunch.cpp:181-182 */
    double tmp2 = 1 / hr[0] * 1 / hr[1] * 1 / hr[2] / (scaleFactor * scaleFactor * scaleFactor) / gammaz;
    for (size_t c = 0; c < ncell; ++c) rho[c] *= tmp2;
    double hs[3] = {hr[0] * scaleFactor, hr[1] * scaleFactor, hr[2] * scaleFactor};
    hs[2] *= gammaz;
    memcpy(img, rho, ncell * sizeof(double));
    orc_compute_potential(n, rho, hs, greens);
    const double vol = hs[0] * hs[1] * hs[2];
    const double sc[3] = {gammaz / scaleFactor, gammaz / scaleFactor, 1.0 / (scaleFactor * gammaz)};
    for (size_t c = 0; c < ncell; ++c) rho[c] *= vol;
    for (size_t c = 0; c < ncell; ++c) rho[c] *= coupling;
    orc_grad(n, rho, hr, E);
    for (size_t c = 0; c < ncell; ++c)
        for (int d = 0; d < 3; ++d) E_self[3 * c + d] = -E[3 * c + d] * sc[d];
    if (image) {
        /* PartBunch.cpp:261-264: zshift from the mesh origin and the z extent */
        const double hz = hr[2];
        const double zshift = -(2 * origin[2] + (0 + (n[2] - 1) + 1) * hz) * gammaz * scaleFactor;
        orc_compute_potential_shifted(n, img, hs, zshift);
        for (size_t c = 0; c < ncell; ++c) img[c] *= vol;
        for (size_t c = 0; c < ncell; ++c) img[c] *= coupling;
        orc_grad(n, img, hr, E);
        for (size_t c = 0; c < ncell; ++c)
            for (int d = 0; d < 3; ++d) E_img[3 * c + d] = -E[3 * c + d] * sc[d];
    }
    free(rho);
    free(img);
    free(E);
}

/* The energy-binned self fields: the bin loop of
 * ParallelTTracker::computeSpaceChargeFields (src/Algorithms/ParallelTTracker.cpp:567-579)
 * over PartBunch::computeSelfFields(int) (PartBunch.cpp:138-430).  Bin b is
 * given charge qi per particle (setBinCharge, PartBunchBase.hpp:716-724), its
 * field and, if `image`, its image charge's field are gathered at every
 * particle and added to Ef; Bf gets -/+ betaC*E_y, +/- betaC*E_x for the bunch
 * and the image (opposite signs).  Ef is overwritten (the tracker zeroes it
 * first), Bf accumulated.  gamma_in: per-bin Lorentz factors, or NULL to use
 * calcGammas.  Returns the number of dropped deposits. */
int64_t orc_compute_self_fields_binned(int64_t N, const double *R, const double *P, const int *bin,
                                       int nbins, double qi, const double *dt, double dT,
                                       const double origin[3], const double hr[3], const int n[3],
                                       double coupling, int greens, int image,
                                       const double *gamma_in, double *Ef, double *Bf) {
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    double *gam = (double *)malloc(sizeof(double) * (nbins > 0 ? nbins : 1));
    if (gamma_in) memcpy(gam, gamma_in, sizeof(double) * nbins);
    else orc_calc_gammas(N, P, bin, nbins, gam);
    double *Qb = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
    double *rho = (double *)malloc(ncell * sizeof(double));
    double *Es = (double *)calloc(3 * ncell, sizeof(double));
    double *Ei = (double *)calloc(3 * ncell, sizeof(double));
    double *tmp = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
    int64_t dropped = 0;
    memset(Ef, 0, sizeof(double) * 3 * N);
    double *Rb = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
    for (int b = 0; b < nbins; ++b) {
        /* the reference scatters every particle with Q = qi*dt in the bin and 0
         * elsewhere; adding the zeros changes no value, so only the bin's
         * particles are scattered (in particle order), and only their
         * out-of-grid deposits are counted */
        int64_t nb = 0;
        for (int64_t i = 0; i < N; ++i)
            if (bin[i] == b) {
                Qb[nb] = qi;
                Qb[nb] *= dt[i];
                memcpy(Rb + 3 * nb, R + 3 * i, 3 * sizeof(double));
                ++nb;
            }
        memset(rho, 0, ncell * sizeof(double));
        dropped += orc_scatter(nb, Rb, Qb, origin, hr, n, rho);
        for (size_t c = 0; c < ncell; ++c) rho[c] /= dT;
        binned_fields(n, rho, origin, hr, gam[b], coupling, greens, image, Es, Ei);
        const double betaC = sqrt(gam[b] * gam[b] - 1.0) / gam[b] / ORC_C;
        orc_gather(N, R, Es, origin, hr, n, tmp);
        for (int64_t i = 0; i < N; ++i) {
            Bf[3 * i + 0] = Bf[3 * i + 0] - betaC * tmp[3 * i + 1];
            Bf[3 * i + 1] = Bf[3 * i + 1] + betaC * tmp[3 * i + 0];
        }
        for (int64_t i = 0; i < 3 * N; ++i) Ef[i] += tmp[i];
        if (image) {
            orc_gather(N, R, Ei, origin, hr, n, tmp);
            for (int64_t i = 0; i < N; ++i) {
                Bf[3 * i + 0] = Bf[3 * i + 0] + betaC * tmp[3 * i + 1];
                Bf[3 * i + 1] = Bf[3 * i + 1] - betaC * tmp[3 * i + 0];
            }
            for (int64_t i = 0; i < 3 * N; ++i) Ef[i] += tmp[i];
        }
    }
    free(gam);
    free(Qb);
    free(Rb);
    free(rho);
    free(Es);
    free(Ei);
    free(tmp);
    return dropped;
}

/* ------------------------------------------------------------------------ */
/* Boris pusher                                                             */
/* ------------------------------------------------------------------------ */

/* ParallelTTracker::pushParticles (src/Algorithms/ParallelTTracker.h:448-455):
 * switchToUnitlessPositions(true) (PartBunchBase.hpp:468-488),
 * BorisPusher::push (src/Steppers/BorisPusher.h:116-124),
 * switchOffUnitlessPositions(true) (PartBunchBase.hpp:493-513). */
void orc_push(int64_t N, double *R, const double *P, const double *dt) {
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) {
        double s = ORC_C * dt[i];
        for (int d = 0; d < 3; ++d) R[3 * i + d] /= s;
    }
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) {
        const double *p = P + 3 * i;
        double g = sqrt(1.0 + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
        for (int d = 0; d < 3; ++d) R[3 * i + d] += 0.5 * p[d] / g;
    }
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) {
        double s = ORC_C * dt[i];
        for (int d = 0; d < 3; ++d) R[3 * i + d] *= s;
    }
}

static void boris_kick1(double P[3], const double Ef[3], const double Bf[3], double dt,
                        double mass, double charge) {
    /* BorisPusher::kick, src/Steppers/BorisPusher.h:56-113 */
    double e = 0.5 * dt * charge * ORC_C / mass;
    for (int d = 0; d < 3; ++d) P[d] += e * Ef[d];
    double gamma = sqrt(1.0 + (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]));
    double tf = 0.5 * dt * charge * ORC_C * ORC_C / (gamma * mass);
    double t[3] = {tf * Bf[0], tf * Bf[1], tf * Bf[2]};
    double w[3] = {P[0] + (P[1] * t[2] - P[2] * t[1]), P[1] + (P[2] * t[0] - P[0] * t[2]),
                   P[2] + (P[0] * t[1] - P[1] * t[0])};
    double sf = 2.0 / (1.0 + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
    double s[3] = {sf * t[0], sf * t[1], sf * t[2]};
    double c[3] = {w[1] * s[2] - w[2] * s[1], w[2] * s[0] - w[0] * s[2], w[0] * s[1] - w[1] * s[0]};
    for (int d = 0; d < 3; ++d) P[d] += c[d];
    double e2 = 0.5 * dt * charge * ORC_C / mass;
    for (int d = 0; d < 3; ++d) P[d] += e2 * Ef[d];
}

/* ParallelTTracker::kickParticles, src/Algorithms/ParallelTTracker.h:442-446. */
void orc_kick(int64_t N, double *P, const double *Ef, const double *Bf, const double *dt,
              double charge, double mass) {
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) boris_kick1(P + 3 * i, Ef + 3 * i, Bf + 3 * i, dt[i], mass, charge);
}

/* ------------------------------------------------------------------------ */
/* Moments and beam parameters                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    double rmean[3], pmean[3], rrms[3], prms[3], eps_norm[3], eps[3], rprms[3];
    double gamma_mean, ekin_MeV;
    double moments[27];
    int64_t n;
} orc_stats;

/* PartBunchBase::calcMoments, src/Classic/Algorithms/PartBunchBase.hpp:2185-2258
 * (sequential accumulation; out[0..5] first moments, out[6..26] lower triangle
 * row by row). */
static void moments_range(int64_t k0, int64_t k1, const double *R, const double *P, double out[27]) {
    for (int q = 0; q < 27; ++q) out[q] = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        double part[6] = {R[3 * k], P[3 * k], R[3 * k + 1], P[3 * k + 1], R[3 * k + 2], P[3 * k + 2]};
        int l = 6;
        for (int i = 0; i < 6; ++i) {
            out[i] += part[i];
            for (int j = 0; j <= i; ++j) out[l++] += part[i] * part[j];
        }
    }
}

void orc_calc_moments(int64_t N, const double *R, const double *P, double out[27]) {
    const int T = g_threads;
    if (T <= 1 || N < 4096) {
        moments_range(0, N, R, P, out);
        return;
    }
    double *part = (double *)malloc(sizeof(double) * 27 * T);
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) moments_range(N * t / T, N * (t + 1) / T, R, P, part + 27 * t);
    for (int q = 0; q < 27; ++q) {
        out[q] = 0.0;
        for (int t = 0; t < T; ++t) out[q] += part[27 * t + q];
    }
    free(part);
}

static double mom(const double m[27], int i, int j) {
    if (j > i) {
        int t = i;
        i = j;
        j = t;
    }
    return m[6 + i * (i + 1) / 2 + j];
}

/* PartBunchBase::calcBeamParameters, PartBunchBase.hpp:1529-1613, with
 * calcEMean (PartBunchBase.hpp:2099-2116).  mass_eV is the species rest energy. */
void orc_beam_parameters_from_moments(int64_t N, const double m[27], double gamma_sum,
                                      double mass_eV, orc_stats *s) {
    memset(s, 0, sizeof(*s));
    s->n = N;
    memcpy(s->moments, m, sizeof(double) * 27);
    if (N == 0) return;
    const double Nd = (double)N;
    double rsq[3], psq[3], rp[3], eps2[3];
    for (int i = 0; i < 3; ++i) {
        s->rmean[i] = m[2 * i] / Nd;
        s->pmean[i] = m[2 * i + 1] / Nd;
        rsq[i] = mom(m, 2 * i, 2 * i) - Nd * s->rmean[i] * s->rmean[i];
        psq[i] = mom(m, 2 * i + 1, 2 * i + 1) - Nd * s->pmean[i] * s->pmean[i];
        if (psq[i] < 0) psq[i] = 0;
        rp[i] = mom(m, 2 * i, 2 * i + 1) - Nd * s->rmean[i] * s->pmean[i];
    }
    for (int i = 0; i < 3; ++i) eps2[i] = (rsq[i] * psq[i] - rp[i] * rp[i]) / (Nd * Nd);
    for (int i = 0; i < 3; ++i) rp[i] /= Nd;
    for (int i = 0; i < 3; ++i) {
        s->rrms[i] = sqrt(rsq[i] / Nd);
        s->prms[i] = sqrt(psq[i] / Nd);
        s->eps_norm[i] = sqrt(eps2[i] > 0.0 ? eps2[i] : 0.0);
        double tmp = s->rrms[i] * s->prms[i];
        double fac = (tmp == 0) ? 0.0 : 1.0 / tmp;
        s->rprms[i] = rp[i] * fac;
    }
    double gamma = gamma_sum / Nd;
    s->gamma_mean = gamma;
    for (int i = 0; i < 3; ++i) s->eps[i] = s->eps_norm[i] / (gamma * sqrt(1.0 - 1.0 / (gamma * gamma)));
    (void)mass_eV;
}

void orc_calc_beam_parameters(int64_t N, const double *R, const double *P, double mass_eV,
                              orc_stats *s) {
    double m[27];
    orc_calc_moments(N, R, P, m);
    double gsum = 0.0, esum = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const double *p = P + 3 * i;
        double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        gsum += sqrt(1.0 + pp);
        esum += sqrt(pp + 1.0);
    }
    orc_beam_parameters_from_moments(N, m, gsum, mass_eV, s);
    if (N > 0) {
        esum -= (double)N;
        esum *= mass_eV * 1.0e-6;
        s->ekin_MeV = esum / (double)N;
    }
}

/* ------------------------------------------------------------------------ */
/* Quaternions and coordinate-system transformations                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    double origin[3];
    double q[4];
    double M[9]; /* row-major */
} orc_trafo;

static void quat_mul(const double a[4], const double b[4], double out[4]) {
    /* Quaternion::operator*=, src/Classic/Algorithms/Quaternion.cpp:84-93 */
    double d = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    double c0 = a[2] * b[3] - a[3] * b[2];
    double c1 = a[3] * b[1] - a[1] * b[3];
    double c2 = a[1] * b[2] - a[2] * b[1];
    double r0 = a[0] * b[0] - d;
    double r1 = a[0] * b[1] + b[0] * a[1] + c0;
    double r2 = a[0] * b[2] + b[0] * a[2] + c1;
    double r3 = a[0] * b[3] + b[0] * a[3] + c2;
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
    out[3] = r3;
}

static void quat_rotate(const double q[4], const double v[3], double out[3]) {
    /* Quaternion::rotate, Quaternion.cpp:122-133: (q * (v * conj(q))).imag() */
    double qv[4] = {0.0, v[0], v[1], v[2]};
    double qc[4] = {q[0], -q[1], -q[2], -q[3]};
    double t[4], r[4];
    quat_mul(qv, qc, t);
    quat_mul(q, t, r);
    out[0] = r[1];
    out[1] = r[2];
    out[2] = r[3];
}

static void quat_rotation_matrix(const double qin[4], double M[9]) {
    /* Quaternion::getRotationMatrix, Quaternion.cpp:135-150 (normalises first) */
    double len = sqrt(qin[0] * qin[0] + qin[1] * qin[1] + qin[2] * qin[2] + qin[3] * qin[3]);
    double r[4] = {qin[0] / len, qin[1] / len, qin[2] / len, qin[3] / len};
    M[0] = 1 - 2 * (r[2] * r[2] + r[3] * r[3]);
    M[1] = 2 * (-r[0] * r[3] + r[1] * r[2]);
    M[2] = 2 * (r[0] * r[2] + r[1] * r[3]);
    M[3] = 2 * (r[0] * r[3] + r[1] * r[2]);
    M[4] = 1 - 2 * (r[1] * r[1] + r[3] * r[3]);
    M[5] = 2 * (-r[0] * r[1] + r[2] * r[3]);
    M[6] = 2 * (-r[0] * r[2] + r[1] * r[3]);
    M[7] = 2 * (r[0] * r[1] + r[2] * r[3]);
    M[8] = 1 - 2 * (r[1] * r[1] + r[2] * r[2]);
}

/* getQuaternion(u, ref=(0,0,1)), Quaternion.cpp:34-69.  The antiparallel
 * branch draws a random axis in the reference; here it is (1,0,0). */
void orc_get_quaternion_to_z(const double uin[3], double q[4]) {
    double len = sqrt(uin[0] * uin[0] + uin[1] * uin[1] + uin[2] * uin[2]);
    if (len < 1e-12) { /* the reference throws ("length of vector less than 1e-12") */
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
        return;
    }
    double u[3] = {uin[0] / len, uin[1] / len, uin[2] / len};
    const double ref[3] = {0.0 / 1.0, 0.0 / 1.0, 1.0 / 1.0};
    double axis[3] = {u[1] * ref[2] - u[2] * ref[1], u[2] * ref[0] - u[0] * ref[2],
                      u[0] * ref[1] - u[1] * ref[0]};
    double normAxis = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    double dur = u[0] * ref[0] + u[1] * ref[1] + u[2] * ref[2];
    if (normAxis < 1e-12) {
        if (fabs(dur - 1.0) < 1e-12) {
            q[0] = 1.0;
            q[1] = q[2] = q[3] = 0.0;
            return;
        }
        q[0] = 0.0;
        q[1] = 1.0;
        q[2] = 0.0;
        q[3] = 0.0;
        return;
    }
    for (int d = 0; d < 3; ++d) axis[d] /= normAxis;
    double cosAngle = sqrt(0.5 * (1 + dur));
    double sinAngle = sqrt(1 - cosAngle * cosAngle);
    q[0] = cosAngle;
    q[1] = sinAngle * axis[0];
    q[2] = sinAngle * axis[1];
    q[3] = sinAngle * axis[2];
}

/* CoordinateSystemTrafo(origin, orientation), CoordinateSystemTrafo.cpp:19-23 */
void orc_trafo_make(const double origin[3], const double q[4], orc_trafo *t) {
    memcpy(t->origin, origin, sizeof(double) * 3);
    memcpy(t->q, q, sizeof(double) * 4);
    quat_rotation_matrix(q, t->M);
}

/* CoordinateSystemTrafo::invert, CoordinateSystemTrafo.cpp:25-29 */
void orc_trafo_inverted(const orc_trafo *a, orc_trafo *t) {
    double o[3];
    quat_rotate(a->q, a->origin, o);
    for (int d = 0; d < 3; ++d) t->origin[d] = -o[d];
    t->q[0] = a->q[0];
    for (int d = 1; d < 4; ++d) t->q[d] = -a->q[d];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t->M[3 * r + c] = a->M[3 * c + r];
}

/* transformTo / rotateTo, CoordinateSystemTrafo.h:77-94 (dot(Tenzor,Vektor),
 * ippl/src/AppTypes/TSVMetaDot.h:255-263). */
static void mat_vec(const double M[9], const double v[3], double out[3]) {
    double a = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    double b = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    double c = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
    out[0] = a;
    out[1] = b;
    out[2] = c;
}

void orc_transform_to(const orc_trafo *t, int64_t N, double *R) {
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) {
        double v[3] = {R[3 * i] - t->origin[0], R[3 * i + 1] - t->origin[1], R[3 * i + 2] - t->origin[2]};
        mat_vec(t->M, v, R + 3 * i);
    }
}

void orc_rotate_to(const orc_trafo *t, int64_t N, double *V) {
    ORC_PAR
    for (int64_t i = 0; i < N; ++i) {
        double v[3] = {V[3 * i], V[3 * i + 1], V[3 * i + 2]};
        mat_vec(t->M, v, V + 3 * i);
    }
}

/* ------------------------------------------------------------------------ */
/* Tracker                                                                  */
/* ------------------------------------------------------------------------ */

/* A Solenoid (kind 0, FM2DMagnetoStatic map: a = Bz, b = Br) or an RFCavity
 * (kind 1, FM2DDynamic map: a = Ez, b = Er, c = Bt), grids nz*nr indexed
 * i + j*nz, in SI units after the map's normalisation. */
typedef struct orc_fieldmap_s {
    int kind, nz, nr;
    double hz, hr, zstart, length, scale, omega, phase;
    const double *a, *b, *c;
} orc_fieldmap;

/* FM2DMagnetoStatic::getFieldstrength (src/Classic/Fields/FM2DMagnetoStatic.cpp:161-190),
 * FM2DDynamic::getFieldstrength (src/Classic/Fields/FM2DDynamic.cpp:169-205). */
static int fieldmap_2d(const orc_fieldmap *m, double x, double y, double z, double e[3], double b[3]) {
    const double RR = sqrt(x * x + y * y);
    int indexr, indexz;
    if (m->kind == 0) {
        indexr = abs((int)floor(RR / m->hr));
        indexz = abs((int)floor(z / m->hz));
    } else {
        indexr = (int)floor(RR / m->hr);
        indexz = (int)floor(z / m->hz);
    }
    const double leverr = RR / m->hr - indexr;
    const double leverz = z / m->hz - indexz;
    if ((indexz < 0) || (indexz + 2 > m->nz)) return 0;
    if (indexr + 2 > m->nr) return 1;
    const int index1 = indexz + indexr * m->nz;
    const int index2 = index1 + m->nz;
#define ORC_LERP(f) ((1.0 - leverz) * (1.0 - leverr) * (f)[index1] + leverz * (1.0 - leverr) * (f)[index1 + 1] + \
                     (1.0 - leverz) * leverr * (f)[index2] + leverz * leverr * (f)[index2 + 1])
    if (m->kind == 0) {
        const double BfieldR = ORC_LERP(m->b), BfieldZ = ORC_LERP(m->a);
        if (RR != 0) {
            b[0] += BfieldR * x / RR;
            b[1] += BfieldR * y / RR;
        }
        b[2] += BfieldZ;
    } else {
        const double EfieldR = ORC_LERP(m->b), EfieldZ = ORC_LERP(m->a), BfieldT = ORC_LERP(m->c);
        if (RR > 1e-10) {
            e[0] += EfieldR * x / RR;
            e[1] += EfieldR * y / RR;
            b[0] -= BfieldT * y / RR;
            b[1] += BfieldT * x / RR;
        }
        e[2] += EfieldZ;
    }
#undef ORC_LERP
    return 0;
}

/* Solenoid::apply (src/Classic/AbsBeamline/Solenoid.cpp:167-179) and
 * RFCavity::apply (src/Classic/AbsBeamline/RFCavity.cpp:232-250) at the lab
 * position (x, y, zlab): 1 when the particle is beyond the map's radius. */
int orc_fieldmap_apply(const orc_fieldmap *m, double x, double y, double zlab, double t, double E[3], double B[3]) {
    const double z = zlab - m->zstart;
    if (!(z >= 0.0 && z < m->length)) return 0;
    double te[3] = {0.0, 0.0, 0.0}, tb[3] = {0.0, 0.0, 0.0};
    if (fieldmap_2d(m, x, y, z, te, tb)) return 1;
    if (m->kind == 0) {
        for (int d = 0; d < 3; ++d) B[d] += m->scale * tb[d];
    } else {
        const double ce = m->scale * cos(m->omega * t + m->phase), sb = m->scale * sin(m->omega * t + m->phase);
        for (int d = 0; d < 3; ++d) E[d] += ce * te[d];
        for (int d = 0; d < 3; ++d) B[d] -= sb * tb[d];
    }
    return 0;
}

/* Every map at lab positions R (AoS) at time t into E, B (overwritten), maps
 * in order, a particle beyond one map's radius skipping the later ones
 * (computeExternalFields, src/Algorithms/ParallelTTracker.cpp:614-640).
 * Returns the number of such particles. */
int64_t orc_fieldmap_eval(int nmaps, const orc_fieldmap *maps, int64_t N, const double *R, double t, double *E,
                          double *B) {
    int64_t lost = 0;
    for (int64_t i = 0; i < N; ++i) {
        double e[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
        int out = 0;
        for (int k = 0; k < nmaps && !out; ++k) out = orc_fieldmap_apply(&maps[k], R[3 * i], R[3 * i + 1], R[3 * i + 2], t, e, b);
        for (int d = 0; d < 3; ++d) {
            E[3 * i + d] = e[d];
            B[3 * i + d] = b[d];
        }
        lost += out;
    }
    return lost;
}

size_t orc_sizeof_fieldmap(void) { return sizeof(orc_fieldmap); }

typedef struct {
    int n[3];
    double dh;        /* BBOXINCR/100 */
    int greens;       /* 0 INTEGRATED, 1 STANDARD */
    int space_charge; /* 0: FSTYPE=NONE */
    double dT;
    double charge;  /* species charge in units of e */
    double mass_eV; /* species rest energy */
    double coupling;
    /* analytic stand-ins for element fields, in the reference frame:
     * B = (0,0,Bz), E = (0,0,Ez0*sin(2*pi*rf_freq*t + rf_phase)) */
    double Bz, Ez0, rf_freq, rf_phase;
    /* tracker state */
    double t, pathLength;
    double RefPartR[3], RefPartP[3];
    /* DC beam (BCFFTZ=PERIODIC): periodic z and its period length */
    int bcz;
    double period;
    /* field maps on the beamline (see orc_fieldmap); the lab frame is the
     * reference frame moved by (0, 0, pathLength) */
    int nmaps;
    const struct orc_fieldmap_s *maps;
} orc_tracker;

static void ext_fields(const orc_tracker *tr, double t, double E[3], double B[3]) {
    E[0] = 0.0;
    E[1] = 0.0;
    E[2] = tr->Ez0 == 0.0 ? 0.0 : tr->Ez0 * sin(2.0 * ORC_PI * tr->rf_freq * t + tr->rf_phase);
    B[0] = 0.0;
    B[1] = 0.0;
    B[2] = tr->Bz;
}

/* ParallelTTracker::updateReferenceParticle (src/Algorithms/ParallelTTracker.cpp:1123-1187)
 * and updateRefToLabCSTrafo (ParallelTTracker.cpp:1199-1213).  Returns the
 * update trafo applied to the bunch by transformBunch. */
static void update_ref(orc_tracker *tr, orc_trafo *upd) {
    double dt = tr->t < tr->dT ? tr->t : tr->dT;
    double scale = ORC_C * dt;
    double *R = tr->RefPartR, *P = tr->RefPartP;
    for (int d = 0; d < 3; ++d) R[d] /= scale;
    {
        double g = sqrt(1.0 + (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]));
        for (int d = 0; d < 3; ++d) R[d] += 0.5 * P[d] / g;
    }
    for (int d = 0; d < 3; ++d) R[d] *= scale;
    double E[3], B[3];
    ext_fields(tr, tr->t - 0.5 * dt, E, B);
    double Ef[3] = {0.0 + E[0], 0.0 + E[1], 0.0 + E[2]};
    double Bf[3] = {0.0 + B[0], 0.0 + B[1], 0.0 + B[2]};
    /* elements at the reference particle (ParallelTTracker.cpp:1134-1155) */
    for (int k = 0; k < tr->nmaps; ++k)
        if (orc_fieldmap_apply(&tr->maps[k], R[0], R[1], R[2] + tr->pathLength, tr->t - 0.5 * dt, Ef, Bf)) break;
    boris_kick1(P, Ef, Bf, dt, tr->mass_eV, tr->charge);
    for (int d = 0; d < 3; ++d) R[d] /= scale;
    {
        double g = sqrt(1.0 + (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]));
        for (int d = 0; d < 3; ++d) R[d] += 0.5 * P[d] / g;
    }
    for (int d = 0; d < 3; ++d) R[d] *= scale;

    tr->pathLength += sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    double q[4];
    orc_get_quaternion_to_z(P, q);
    orc_trafo_make(R, q, upd);
    double nr[3], np[3];
    double v[3] = {R[0] - upd->origin[0], R[1] - upd->origin[1], R[2] - upd->origin[2]};
    mat_vec(upd->M, v, nr);
    mat_vec(upd->M, P, np);
    memcpy(R, nr, sizeof(nr));
    memcpy(P, np, sizeof(np));
}

/* The space-charge part of one step, ParallelTTracker::computeSpaceChargeFields
 * (src/Algorithms/ParallelTTracker.cpp:546-592) for a bunch without energy bins:
 * moments -> quaternion -> rotate R into the beam frame -> boundp ->
 * computeSelfFields -> rotate R, Ef, Bf back.  mesh_out (hr, origin) optional. */
typedef struct {
    const int *bin; /* per-particle energy bin, NULL: no bins */
    int nbins;
    double qi;      /* macro charge given to the particles of the bin being solved */
    int image;
} orc_bins;

static int64_t space_charge_impl(orc_tracker *tr, int64_t N, double *R, const double *P, double *Q,
                                 const double *dt, double *Ef, double *Bf, double mesh_out[6],
                                 double *rho_out, double *phi_out, double *E_out, const orc_bins *bins);

int64_t orc_space_charge(orc_tracker *tr, int64_t N, double *R, const double *P, double *Q,
                         const double *dt, double *Ef, double *Bf, double mesh_out[6],
                         double *rho_out, double *phi_out, double *E_out) {
    return space_charge_impl(tr, N, R, P, Q, dt, Ef, Bf, mesh_out, rho_out, phi_out, E_out, NULL);
}

/* With energy bins: calcGammas, then the bin loop (ParallelTTracker.cpp:567-579). */
static int64_t space_charge_impl(orc_tracker *tr, int64_t N, double *R, const double *P, double *Q,
                                 const double *dt, double *Ef, double *Bf, double mesh_out[6],
                                 double *rho_out, double *phi_out, double *E_out, const orc_bins *bins) {
    double m[27];
    orc_calc_moments(N, R, P, m);
    double pmean[3] = {m[1] / (double)N, m[3] / (double)N, m[5] / (double)N};
    double qa[4], qb[4];
    orc_get_quaternion_to_z(pmean, qa);
    qb[0] = qa[0];
    qb[1] = -qa[1];
    qb[2] = -qa[2];
    qb[3] = -qa[3];
    double o[3] = {0.0, 0.0, tr->pathLength};
    orc_trafo b2r, r2b;
    orc_trafo_make(o, qb, &b2r);
    orc_trafo_inverted(&b2r, &r2b);
    orc_transform_to(&r2b, N, R);
    double hr[3], origin[3];
    if (tr->bcz && bins && bins->bin) return -1; /* no energy bins for a DC beam */
    if (boundp_impl(N, R, tr->n, tr->dh, tr->bcz, tr->period, hr, origin) != 0) return -1;
    if (tr->bcz) orc_wrap_periodic_z(N, R, origin, hr, tr->n);
    if (mesh_out) {
        memcpy(mesh_out, hr, sizeof(hr));
        memcpy(mesh_out + 3, origin, sizeof(origin));
    }
    int64_t dropped;
    if (bins && bins->bin)
        dropped = orc_compute_self_fields_binned(N, R, P, bins->bin, bins->nbins, bins->qi, dt, tr->dT,
                                                 origin, hr, tr->n, tr->coupling, tr->greens, bins->image,
                                                 NULL, Ef, Bf);
    else
        dropped = self_fields_impl(N, R, P, Q, dt, tr->dT, origin, hr, tr->n, tr->coupling, tr->greens,
                                   tr->bcz, Ef, Bf, rho_out, phi_out, E_out);
    orc_transform_to(&b2r, N, R);
    orc_rotate_to(&b2r, N, Ef);
    orc_rotate_to(&b2r, N, Bf);
    return dropped;
}

/* ParallelTTracker::execute loop body, src/Algorithms/ParallelTTracker.cpp:343-388,
 * for nsteps steps with per-step statistics (DataSink stat quantities after
 * each step).  Returns 0 or a negative error. */
static int track_impl(orc_tracker *tr, int64_t N, double *R, double *P, double *Q, double *dt, int nsteps,
                      orc_stats *stats, const orc_bins *bins);

int orc_track(orc_tracker *tr, int64_t N, double *R, double *P, double *Q, double *dt, int nsteps,
              orc_stats *stats) {
    return track_impl(tr, N, R, P, Q, dt, nsteps, stats, NULL);
}

/* The same loop with energy bins (space charge by bin + image charge). */
int orc_track_binned(orc_tracker *tr, int64_t N, double *R, double *P, double *Q, double *dt,
                     const int *bin, int nbins, double qi, int image, int nsteps, orc_stats *stats) {
    orc_bins b = {bin, nbins, qi, image};
    return track_impl(tr, N, R, P, Q, dt, nsteps, stats, &b);
}

static int track_impl(orc_tracker *tr, int64_t N, double *R, double *P, double *Q, double *dt, int nsteps,
                      orc_stats *stats, const orc_bins *bins) {
    double *Ef = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
    double *Bf = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
    for (int s = 0; s < nsteps; ++s) {
        orc_push(N, R, P, dt);
        memset(Ef, 0, sizeof(double) * 3 * N);
        memset(Bf, 0, sizeof(double) * 3 * N);
        if (tr->space_charge && N > 0) {
            if (space_charge_impl(tr, N, R, P, Q, dt, Ef, Bf, NULL, NULL, NULL, NULL, bins) < 0) {
                free(Ef);
                free(Bf);
                return -1;
            }
        }
        for (int64_t i = 0; i < N; ++i) {
            double E[3], B[3];
            ext_fields(tr, tr->t + 0.5 * dt[i], E, B);
            for (int d = 0; d < 3; ++d) {
                Ef[3 * i + d] += E[d];
                Bf[3 * i + d] += B[d];
            }
            /* computeExternalFields (ParallelTTracker.cpp:595-640) */
            for (int k = 0; k < tr->nmaps; ++k)
                if (orc_fieldmap_apply(&tr->maps[k], R[3 * i], R[3 * i + 1], R[3 * i + 2] + tr->pathLength,
                                       tr->t + 0.5 * dt[i], &Ef[3 * i], &Bf[3 * i]))
                    break;
        }
        orc_kick(N, P, Ef, Bf, dt, tr->charge, tr->mass_eV);
        orc_push(N, R, P, dt);
        for (int64_t i = 0; i < N; ++i) dt[i] = tr->dT;
        tr->t += tr->dT;
        if (tr->t > 0.0) {
            orc_trafo upd;
            update_ref(tr, &upd);
            orc_transform_to(&upd, N, R);
            orc_rotate_to(&upd, N, P);
        }
        if (stats) orc_calc_beam_parameters(N, R, P, tr->mass_eV, &stats[s]);
    }
    free(Ef);
    free(Bf);
    return 0;
}

size_t orc_sizeof_stats(void) { return sizeof(orc_stats); }
size_t orc_sizeof_tracker(void) { return sizeof(orc_tracker); }

// Hockney open-boundary Poisson solve on the doubled grid, gfx950.
//
// Reference: FFTPoissonSolver::computePotential and integratedGreensFunction
// (src/Solvers/FFTPoissonSolver.cpp:337-382,459-522,641-664).  The reference
// zero-pads rho to (2n)^3, r2c-transforms rho and the mirrored Green's
// function, multiplies, and c2r-transforms back: an exact circular
// convolution.  Here the same convolution is computed with pruned passes that
// never touch the zero half of the padded input or the unused half of the
// output (complex buffers hold K0 = n0+1 frequencies in x, stored with a row
// pitch kp >= K0 that is a whole number of 128-byte tiles; padding stays 0):
//
//   X-fwd  rho (n0*n1*n2)          -> A[z<n2][y<n1][kx<K0]     r2c along x
//   Y-fwd  A                       -> B[z<n2][ky<2n1][kx]      along y, input y<n1
//   Z-conv B (in place)            fwd along z (input z<n2), * Ghat, inverse, keep z<n2
//   Y-inv  B                       -> A                        keep y<n1
//   X-inv  A                       -> phi (n0*n1*n2)           c2r, keep x<n0
//
// The Green's function is real and even in every axis on the doubled grid, so
// its transform Ghat is real and even as well: it is the 3D DCT-I of the
// (n+1)^3 octant, computed here with three DCT passes instead of a (2n)^3 r2c.
#include "../../include/opalsc.h"
#include "common.hpp"
#include "fft.hpp"
#include "tuning.hpp"  // generated from src_amd/tuning.json by src_amd/build.py
#include "igfmath.hpp"
#include "kernels.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace opalsc {

// ---------------------------------------------------------------------------
// Green's function octant (FP64 in both precisions)
// ---------------------------------------------------------------------------

// FFTPoissonSolver::integratedGreensFunction, FFTPoissonSolver.cpp:472-491: the
// corner term (igf::term_cr, the reference's operations with correctly rounded
// atan and log) at corner (i, j, k)
__device__ inline double igf_corner(int i, int j, int k, double h0, double h1, double h2, double cellVolume) {
    double v0 = i * h0 - h0 / 2;
    double v1 = j * h1 - h1 / 2;
    double v2 = k * h2 - h2 / 2;
    return igf::term_cr(v0, v1, v2, cellVolume);
}

// Corner antiderivatives T[k][j][i], i,j,k in [0, n+1], each evaluated once
// (FFTPoissonSolver.cpp:472-491; the reference's tmpgreen field) for the FP64
// solve.  Grid: x over the (n0+2)*(n1+2) corners of a plane, y over k.
__global__ __launch_bounds__(256) void k_igf_corners(double *T, int n0, int n1, int n2,
                                                     const Scalars *sc, size_t tB) {
    sc += blockIdx.z;  // batch: one table per grid, tB doubles apart
    T += blockIdx.z * tB;
    const double h0 = sc->hs[0], h1 = sc->hs[1], h2 = sc->hs[2];
    const double vol = h0 * h1 * h2;
    const int m0 = n0 + 2, plane = m0 * (n1 + 2);
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= plane) return;
    const int j = div_small(c, m0, 1.0f / (float)m0), i = c - j * m0, k = blockIdx.y;
    T[(size_t)k * plane + c] = igf_corner(i, j, k, h0, h1, h2, vol);
}

// The two corner tables of FFTPoissonSolver::shiftedIntGreensFunction
// (FFTPoissonSolver.cpp:552-603): z coordinate (k*h2 - h2/2) + zshift, and for
// the second table also - n2*h2, evaluated left to right as there (FP64 solve).
__global__ __launch_bounds__(256) void k_igf_corners_shifted(double *T, int n0, int n1, int n2,
                                                             const Scalars *sc, int second, size_t tB) {
    sc += blockIdx.z;
    T += blockIdx.z * tB;
    const double h0 = sc->hs[0], h1 = sc->hs[1], h2 = sc->hs[2], zshift = sc->zshift;
    const double vol = h0 * h1 * h2;
    const int m0 = n0 + 2, plane = m0 * (n1 + 2);
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= plane) return;
    const int j = div_small(c, m0, 1.0f / (float)m0), i = c - j * m0, k = blockIdx.y;
    double v0 = i * h0 - h0 / 2;
    double v1 = j * h1 - h1 / 2;
    double v2 = k * h2 - h2 / 2 + zshift;
    if (second) v2 = v2 - n2 * h2;
    T[(size_t)k * plane + c] = igf::term_cr(v0, v1, v2, vol);
}

// G octant [0,n0]x[0,n1]x[0,n2] from T: the 8-corner difference in the
// reference's order (FFTPoissonSolver.cpp:510-517) with G[0,0,0] := G[0,0,1]
// (FFTPoissonSolver.cpp:657).  Index n uses T at n+1 from the same formula;
// the physical potential never reads G at index n.  Grid: x over the
// (n0+1)*(n1+1) points of a plane, y over runs of kIgfDiffK planes; each
// thread walks its column up the run and keeps the four corners of the plane
// below in registers, so every T value is loaded once per run instead of
// twice.  The shifted (image-charge) tables use the same difference without
// the quirk (FFTPoissonSolver.cpp:621-639).
constexpr int kIgfDiffK = 8;
// The quirk G[0,0,0] := G[0,0,1] is written by the thread of column 0 in the
// run that holds planes 0 and 1 (k = 1 is computed after k = 0 in the same
// thread, so no other block writes G[0] afterwards): the first run must hold
// both planes, and n2 >= 4 (PoissonGrid::init) guarantees plane 1 exists.
static_assert(kIgfDiffK >= 2, "the run of planes 0..kIgfDiffK-1 must include plane 1 for the G[0,0,0] quirk");
__global__ __launch_bounds__(256) void k_igf_diff(const double *T, double *G, int n0, int n1, int n2, int quirk,
                                                  size_t tB, size_t gB) {
    T += blockIdx.z * tB;
    G += blockIdx.z * gB;
    const int g0 = n0 + 1, gplane = g0 * (n1 + 1);
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= gplane) return;
    const int j = div_small(c, g0, 1.0f / (float)g0), i = c - j * g0;
    const int kbeg = blockIdx.y * kIgfDiffK;
    const int kend = kbeg + kIgfDiffK < n2 + 1 ? kbeg + kIgfDiffK : n2 + 1;
    const size_t m0 = (size_t)n0 + 2, plane = m0 * (n1 + 2);
    const double *t = T + (size_t)kbeg * plane + (size_t)j * m0 + i;
    double a00 = t[0], a01 = t[1], a10 = t[m0], a11 = t[m0 + 1];
    for (int k = kbeg; k < kend; ++k) {
        t += plane;
        const double b00 = t[0], b01 = t[1], b10 = t[m0], b11 = t[m0 + 1];
        double g = b11;
        g += a01;
        g += a10;
        g += b00;
        g -= a11;
        g -= b01;
        g -= b10;
        g -= a00;
        G[(size_t)k * gplane + c] = g;
        if (quirk && c == 0 && k == 1) G[0] = g;
        a00 = b00;
        a01 = b01;
        a10 = b10;
        a11 = b11;
    }
}

// STANDARD Green's function, SpecializedGreensFunction<3>::calculate
// (FFTPoissonSolver.cpp:32-43): G = 1/sqrt(i^2 h0^2 + j^2 h1^2 + k^2 h2^2), G[0] := G[0,0,1].
__global__ __launch_bounds__(256) void k_green_standard(double *G, int n0, int n1, int n2,
                                                        const Scalars *sc, size_t gB) {
    sc += blockIdx.z;
    G += blockIdx.z * gB;
    const size_t total = (size_t)(n0 + 1) * (n1 + 1) * (n2 + 1);
    const double hsq0 = sc->hs[0] * sc->hs[0], hsq1 = sc->hs[1] * sc->hs[1],
                 hsq2 = sc->hs[2] * sc->hs[2];
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
         q += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(q % (n0 + 1));
        int j = (int)((q / (n0 + 1)) % (n1 + 1));
        int k = (int)(q / ((size_t)(n0 + 1) * (n1 + 1)));
        if (i == 0 && j == 0 && k == 0) k = 1;
        double g = (double)(i * i) * hsq0 + (double)(j * j) * hsq1 + (double)(k * k) * hsq2;
        G[q] = 1.0 / sqrt(g);
    }
}

// The mixed-precision solve's octant in one pass, without a corner table:
// igf::green_point_fast at every point (the exact 8-corner difference near
// the origin, the cell-average expansion elsewhere).  shifted: the image
// charge's tables (second: the one shifted by a further -n2*h2), which have no
// G[0,0,0] quirk.  Grid: x over the (n0+1)*(n1+1) points of a plane, y over
// runs of kGreenPlanes planes, z over the batch's grids (gB doubles apart).
constexpr int kGreenPlanes = 4;
__global__ __launch_bounds__(256) void k_green_fast(double *G, int n0, int n1, int n2, const Scalars *sc,
                                                   int shifted, int second, size_t gB) {
    sc += blockIdx.z;
    G += blockIdx.z * gB;
    const double h0 = sc->hs[0], h1 = sc->hs[1], h2 = sc->hs[2], zshift = sc->zshift;
    const double ivol = 1.0 / (h0 * h1 * h2);
    const igf::GreenMoments m = igf::green_moments(h0, h1, h2);
    const int g0 = n0 + 1, gplane = g0 * (n1 + 1);
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= gplane) return;
    const int j = div_small(c, g0, 1.0f / (float)g0), i = c - j * g0;
    // kGreenPlanes planes per thread: the cell moments and 1/volume (FP64
    // divisions, the same for every point) once per thread, not per point
#pragma unroll
    for (int t = 0; t < kGreenPlanes; ++t) {
        const int k = blockIdx.y * kGreenPlanes + t;
        if (k > n2) break;
        const int kq = (!shifted && c == 0 && k == 0) ? 1 : k;  // G[0,0,0] := G[0,0,1] (FFTPoissonSolver.cpp:657)
        G[(size_t)k * gplane + c] =
            igf::green_point_fast(m, i, j, kq, h0, h1, h2, zshift, shifted != 0, second != 0, n2, ivol);
    }
}
void launch_green_fast(hipStream_t s, double *G, const int n[3], const Scalars *sc, int shifted, int second,
                       int nb, size_t gB) {
    dim3 grid(((n[0] + 1) * (n[1] + 1) + 255) / 256, (n[2] + kGreenPlanes) / kGreenPlanes, nb);
    k_green_fast<<<grid, 256, 0, s>>>(G, n[0], n[1], n[2], sc, shifted, second, gB);
}

// ---------------------------------------------------------------------------
// DCT-I passes (even extension of each octant line, length-2n FFT, real part)
// ---------------------------------------------------------------------------

// Tile loads: every kernel first issues all of its thread's global loads into
// registers (the loop over q is unrolled; at most QH = E/2+1 elements per
// thread for half-length and n+1-length lines, E for full lines), then writes
// them to LDS, so the loads of a thread are in flight together.
template <int E>
constexpr int half_q() { return E / 2 + 1; }

// The DCT-I passes transform two real lines in one complex line: the even
// extension of a real line has a real DFT, so FFT(a + i b) = A + i B with A and
// B real, and the real and imaginary parts of the result are the two lines'
// transforms.  The pair is two neighbouring lines (of similar magnitude, so
// the rounding each adds to the other is that of the pair's larger value).

// Octant lines (k, j) of G (pitch n0+1, FP64) -> rows of Gh (pitch gp); complex
// line p of the tile holds real lines 2p and 2p+1 of the block's 2W.
// GREEN (the mixed-precision grid's IGF tables): the octant values are
// evaluated here instead of read from G, so the octant is never stored (table
// 0: the G[0,0,0] quirk; 1, 2: the image charge's shifted tables; scalars
// sc + blockIdx.z): on an FP32 grid k_green_fast's near-field values rounded
// to FP32 and its far-field expansion evaluated in FP32 (green_point_mixed).
template <typename T, int E, bool GREEN = false>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_dct_x(
    const double *G, T *Gh, FftPlan pl, int n0, int nlines, int gp, size_t gB, size_t ghB,
    const Scalars *sc = nullptr, int table = 0, int n1 = 0, int n2 = 0) {
    using C = typename cplx_of<T>::type;
    constexpr int Q = half_q<E>();
    G += blockIdx.z * gB;
    Gh += blockIdx.z * ghB;
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    const int W = pl.W, P = pl.P, L = pl.L, m = n0 + 1, total = W * m;
    const float rm = 1.0f / (float)m;
    const int line0 = blockIdx.x * 2 * W;
    const C *tw = fft_twiddles(buf, pl);
    if constexpr (GREEN) {
        // no loads to keep in flight: each pair of values goes to LDS as it is formed
        sc += blockIdx.z;
        const double h0 = sc->hs[0], h1 = sc->hs[1], h2 = sc->hs[2], zshift = sc->zshift;
        const double ivol = 1.0 / (h0 * h1 * h2);
        const igf::GreenMoments gm = igf::green_moments(h0, h1, h2);
        const int m1 = n1 + 1;
        const float rm1 = 1.0f / (float)m1;
        const bool shifted = table != 0, second = table == 2;
        // octant point (i, line): k, j of the line, with G[0,0,0] := G[0,0,1]
        auto coords = [&](int line, int i, int &j, int &kq) {
            const int k = div_small(line, m1, rm1);
            j = line - k * m1;
            kq = (table == 0 && i == 0 && j == 0 && k == 0) ? 1 : k;
        };
        // on an FP32 grid the far field in FP32 (green_point_mixed)
        const igf::GreenMomentsF gmf = igf::green_moments_f(h0, h1, h2);
        auto point = [&](int line, int i) -> T {
            int j, kq;
            coords(line, i, j, kq);
            if constexpr (sizeof(T) == 4)
                return igf::green_point_mixed(gm, gmf, i, j, kq, h0, h1, h2, zshift, shifted, second, n2, ivol);
            else
                return (T)igf::green_point_fast(gm, i, j, kq, h0, h1, h2, zshift, shifted, second, n2, ivol);
        };
        for (int e = threadIdx.x; e < total; e += kFftThreads) {
            const int ln = div_small(e, m, rm), i = e - ln * m;
            const C v = C{point(min(line0 + 2 * ln, nlines - 1), i), point(min(line0 + 2 * ln + 1, nlines - 1), i)};
            buf[__umul24(i, P) + ln] = v;
            if (i >= 1 && i <= n0 - 1) buf[__umul24(L - i, P) + ln] = v;
        }
    } else {
        T r[Q], u[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int e = min(threadIdx.x + q * kFftThreads, total - 1);  // clamped: loads stay unconditional
            const int ln = div_small(e, m, rm), i = e - ln * m;
            r[q] = (T)G[(size_t)min(line0 + 2 * ln, nlines - 1) * m + i];
            u[q] = (T)G[(size_t)min(line0 + 2 * ln + 1, nlines - 1) * m + i];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int e = threadIdx.x + q * kFftThreads;
            if (e < total) {
                const int ln = div_small(e, m, rm), i = e - ln * m;
                buf[__umul24(i, P) + ln] = C{r[q], u[q]};
                if (i >= 1 && i <= n0 - 1) buf[__umul24(L - i, P) + ln] = C{r[q], u[q]};
            }
        }
    }
    __syncthreads();
    fft_lds<C, E>(buf, tw, pl, false);
    for (int e = threadIdx.x; e < total; e += kFftThreads) {
        const int ln = div_small(e, m, rm), k = e - ln * m;
        const C v = buf[__umul24(k, P) + ln];
        if (line0 + 2 * ln < nlines) Gh[(size_t)(line0 + 2 * ln) * gp + k] = v.x;
        if (line0 + 2 * ln + 1 < nlines) Gh[(size_t)(line0 + 2 * ln + 1) * gp + k] = v.y;
    }
}

// DCT-I along a strided axis of Gh in place: element i of line (x, o) at
// o*ostride + x + i*S, x in [x0, x0+2W) (whole tiles: the row pitch is a
// multiple of 2W), i in [0, npts); complex column p of the tile holds the real
// columns x0+2p and x0+2p+1 (one two-element load and store each).  Columns
// x >= ncols are padding: they hold zeros and are stored as zeros, since the
// transform of a padding column paired with a real one is that column's
// rounding, which later calls would transform again and feed back.
template <typename T, int E>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_dct_strided(T *Gh, FftPlan pl, int npts, size_t S,
                                                             size_t ostride, size_t ghB, int ncols) {
    using C = typename cplx_of<T>::type;
    constexpr int Q = half_q<E>();
    int bx, by, bz;
    xcd_block(bx, by, bz);
    Gh += bz * ghB;
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    const int W = pl.W, lw = pl.lw, P = pl.P, L = pl.L, total = W * npts;
    const int x0 = bx * 2 * W;
    const size_t base = (size_t)by * ostride + x0;
    const int nn = npts - 1;
    const C *tw = fft_twiddles(buf, pl);
    C r[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = min(threadIdx.x + q * kFftThreads, total - 1);
        r[q] = *(const C *)(Gh + base + 2 * (e & (W - 1)) + (size_t)(e >> lw) * S);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kFftThreads;
        if (e < total) {
            const int ln = e & (W - 1), i = e >> lw;
            buf[__umul24(i, P) + ln] = r[q];
            if (i >= 1 && i <= nn - 1) buf[__umul24(L - i, P) + ln] = r[q];
        }
    }
    __syncthreads();
    fft_lds<C, E>(buf, tw, pl, false);
    for (int e = threadIdx.x; e < total; e += kFftThreads) {
        const int ln = e & (W - 1), k = e >> lw, x = x0 + 2 * ln;
        C v = buf[__umul24(k, P) + ln];
        if (x >= ncols) v.x = T(0);
        if (x + 1 >= ncols) v.y = T(0);
        *(C *)(Gh + base + 2 * ln + (size_t)k * S) = v;
    }
}

// z transform of the image-charge Green's function, which is even in x and y
// but not in z (mirrorRhoField(Field_t&), FFTPoissonSolver.cpp:666-691): the
// doubled z line of column (kx, ky) is Ha[z] for z <= n2 and Hb[z - n2] for
// n2 < z < 2n2 (Ha, Hb: the two octant tables after the x and y DCTs, pitch
// gp), transformed forward and stored complete (L2 = 2n2 complex values) in
// Gi[kz][ky][kx] (pitch gp).
template <typename T, int E>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_green_zfft(
    const T *Ha, const T *Hb, typename cplx_of<T>::type *Gi, FftPlan pl, int n1, int n2, int gp, size_t hB,
    size_t giB) {
    using C = typename cplx_of<T>::type;
    Ha += blockIdx.z * hB;
    Hb += blockIdx.z * hB;
    Gi += blockIdx.z * giB;
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    const int W = pl.W, lw = pl.lw, P = pl.P, L = pl.L, total = W * L;
    const int x0 = blockIdx.x * W, ky = blockIdx.y;
    const size_t S = (size_t)gp * (n1 + 1), base = (size_t)ky * gp + x0;
    const C *tw = fft_twiddles(buf, pl);
    T r[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = threadIdx.x + q * kFftThreads;  // total = E * kFftThreads exactly
        const int z = e >> lw;
        r[q] = z <= n2 ? Ha[base + (e & (W - 1)) + (size_t)z * S] : Hb[base + (e & (W - 1)) + (size_t)(z - n2) * S];
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = threadIdx.x + q * kFftThreads;
        buf[__umul24(e >> lw, P) + (e & (W - 1))] = C{r[q], T(0)};
    }
    __syncthreads();
    fft_lds<C, E>(buf, tw, pl, false);
    const size_t Sg = (size_t)gp * (n1 + 1);
    for (int e = threadIdx.x; e < total; e += kFftThreads) {
        const int ln = e & (W - 1), kz = e >> lw;
        Gi[(size_t)kz * Sg + (size_t)ky * gp + x0 + ln] = buf[__umul24(kz, P) + ln];
    }
}

// ---------------------------------------------------------------------------
// Convolution passes
// ---------------------------------------------------------------------------

// SRC 0: int64 fixed-point charge (value = i*q2e, then /dT, *tmp2:
//        PartBunch.cpp:494-510); SRC 1: double rho as given.
// The x passes transform two real lines in one complex line, as the DCT passes
// do: line 2p of a block's 2W lines is the real part of complex line p and line
// 2p + 1 its imaginary part.  With Z = FFT(a + i b), A[k] = (Z[k] + conj Z[L-k])/2
// and B[k] = -i (Z[k] - conj Z[L-k])/2 (k < L/2 + 1, Z[L] = Z[0]); the inverse
// pass transforms Z = Y_a + i Y_b with Y's Hermitian extension and takes the real
// and imaginary parts.  The imaginary parts of Y[0] and Y[L/2], which a c2r
// transform ignores, are dropped before the merge.
template <typename T, int E, int SRC>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_xfwd(const void *src, typename cplx_of<T>::type *A,
                                                      FftPlan pl, int n0, int nlines, int kp,
                                                      const Scalars *sc, size_t sB, size_t aB) {
    using C = typename cplx_of<T>::type;
    constexpr int Q = E;  // 2W lines of n0 = L/2: E elements per thread
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    T *bufr = (T *)smem;
    sc += blockIdx.z;  // batch: grid blockIdx.z, its source sB cells and its spectrum aB elements apart
    A += blockIdx.z * aB;
    const int W = pl.W, lw = pl.lw, P = pl.P, L = pl.L, K0 = n0 + 1, total = 2 * W * n0;
    const float rn = 1.0f / (float)n0, rk = 1.0f / (float)K0;
    const int line0 = blockIdx.x * 2 * W;
    double q2e = 1.0, dT = 1.0, tmp2 = 1.0;
    if constexpr (SRC == 0) {
        q2e = sc->q2e;
        dT = sc->dT;
        tmp2 = sc->tmp2;
    }
    const C *tw = fft_twiddles(buf, pl);
    // SRC 0: the fixed-point charge grid, two words per cell (hi, lo) on the
    // FP64 path and one on the mixed path
    constexpr bool kTwoWords = SRC == 0 && std::is_same<T, double>::value;
    using S = typename std::conditional<SRC == 0, typename std::conditional<kTwoWords, longlong2, long long>::type,
                                        double>::type;
    double q2e_lo = 0.0;
    int lsh = 0;
    long long lmask = 0;
    if constexpr (kTwoWords) {
        q2e_lo = sc->q2e_lo;
        lsh = sc->lo_shift;
        lmask = ((long long)1 << lsh) - 1;
    }
    const S *s = (const S *)src + blockIdx.z * sB;
    S r[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = min(threadIdx.x + q * kFftThreads, total - 1);
        const int ln = div_small(e, n0, rn), x = e - ln * n0;
        r[q] = s[(size_t)min(line0 + ln, nlines - 1) * n0 + x];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kFftThreads;
        if (e < total) {
            const int ln = div_small(e, n0, rn), x = e - ln * n0;
            double d;
            if constexpr (kTwoWords) {
                // the canonical words (lo in [0, 2^lo_shift)) of the cell's exact
                // sum, so the value does not depend on how deposits were grouped
                const long long hi = r[q].x + (r[q].y >> lsh), lo = r[q].y & lmask;
                d = (double)hi * q2e + (double)lo * q2e_lo;
                d = d / dT;
                d = d * tmp2;
            } else if constexpr (SRC == 0) {
                d = (double)r[q] * q2e;
                d = d / dT;
                d = d * tmp2;
            } else {
                d = r[q];
            }
            bufr[2 * (__umul24(x, P) + (ln >> 1)) + (ln & 1)] = (T)d;
        }
    }
    for (int e = threadIdx.x; e < W * (L - n0); e += kFftThreads) {
        const int ln = e & (W - 1), x = n0 + (e >> lw);
        buf[__umul24(x, P) + ln] = C{T(0), T(0)};
    }
    __syncthreads();
    fft_lds<C, E>(buf, tw, pl, false);
    for (int e = threadIdx.x; e < 2 * W * K0; e += kFftThreads) {
        const int ln = div_small(e, K0, rk), k = e - ln * K0;
        if (line0 + ln < nlines) {
            const C a = buf[__umul24(k, P) + (ln >> 1)], b = buf[__umul24(k == 0 ? 0 : L - k, P) + (ln >> 1)];
            const C v = (ln & 1) ? C{(a.y + b.y) * T(0.5), (b.x - a.x) * T(0.5)}
                                 : C{(a.x + b.x) * T(0.5), (a.y - b.y) * T(0.5)};
            A[(size_t)(line0 + ln) * kp + k] = v;
        }
    }
}

// A strided pass over one tile with the plan's first and last radix equal (R):
// every Stockham stage reads, for the thread's butterflies, rows j + r L/R of
// its line, and the last stage of a transform writes exactly those rows.  So
// the first stage reads its inputs from HBM (rows at and beyond nin are zero)
// and the last stage writes HBM (rows below nout); in the convolution passes
// (MODE 2: Ghat; 3: image, -Gi, output flipped in z; 4: periodic z, Gper) the
// last forward stage, the product and the first inverse stage stay in
// registers.  Modes as k_strided's (0 and 1 compile, but k_strided takes the
// LDS path for them).
template <typename T, int E, int R, int MODE>
__device__ inline void strided_fused(const typename cplx_of<T>::type *in, typename cplx_of<T>::type *out,
                                     typename cplx_of<T>::type *buf, const typename cplx_of<T>::type *tw,
                                     const FftPlan &pl, size_t ib, size_t ob, size_t inS, size_t outS, int nin,
                                     int nout, const T *Gh, int gp, int n1, int n2, int fy, int kx0) {
    using C = typename cplx_of<T>::type;
    constexpr bool CONV = MODE >= 2, INV1 = MODE == 1;
    using G = typename std::conditional<(MODE >= 3), C, T>::type;  // complex factors of the image / periodic z
    constexpr int NB = E / R;
    constexpr int lr = R == 8 ? 3 : (R == 4 ? 2 : 1);
    const int W = pl.W, lw = pl.lw, P = pl.P, L = pl.L, LR = L / R, nst = pl.nst;
    C v[NB][R];
    G g[CONV ? NB : 1][CONV ? R : 1];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int e = threadIdx.x + b * kFftThreads;
        const int line = e & (W - 1), j = e >> lw;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = j + r * LR;
            v[b][r] = row < nin ? in[ib + (size_t)row * inS + line] : C{T(0), T(0)};
            if constexpr (MODE == 2) {
                const int fz = row <= n2 ? row : L - row;  // Ghat folded into the octant
                g[b][r] = Gh[((size_t)fz * (n1 + 1) + fy) * gp + kx0 + line];
            } else if constexpr (CONV) {
                g[b][r] = ((const C *)Gh)[((size_t)row * (n1 + 1) + fy) * gp + kx0 + line];
            }
        }
        dftR<C, R>(v[b], INV1);  // first stage: its twiddles are all 1
    }
    stage_write<C, E, R>(buf, L, lw, P, 0, v);
    __syncthreads();
    int ls = fft_lds_range<C, E>(buf, tw, pl, 1, nst - 1, lr, INV1);
    stage_read<C, E, R>(buf, tw, L, lw, P, ls, INV1, v);  // last stage: rows j + r L/R
    if constexpr (CONV) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (MODE == 2) {
                    v[b][r] = C{v[b][r].x * g[b][r], v[b][r].y * g[b][r]};
                } else {
                    const C a = v[b][r], q = g[b][r];
                    const C p = C{a.x * q.x - a.y * q.y, a.x * q.y + a.y * q.x};
                    v[b][r] = MODE == 3 ? C{-p.x, -p.y} : p;
                }
            }
            dftR<C, R>(v[b], true);  // first inverse stage, same rows, twiddles 1
        }
        __syncthreads();  // every thread has read the forward transform's last stage
        stage_write<C, E, R>(buf, L, lw, P, 0, v);
        __syncthreads();
        ls = fft_lds_range<C, E>(buf, tw, pl, 1, nst - 1, lr, true);
        stage_read<C, E, R>(buf, tw, L, lw, P, ls, true, v);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int e = threadIdx.x + b * kFftThreads;
        const int line = e & (W - 1), j = e >> lw;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = j + r * LR;
            // the image potential is read back flipped in z (FFTPoissonSolver.cpp:326-329)
            if (row < nout) out[ob + (size_t)(MODE == 3 ? nout - 1 - row : row) * outS + line] = v[b][r];
        }
    }
}

// Strided pass over one tile of W consecutive kx at fixed outer index o:
// element i of line kx at in[o*inO + i*inS + kx], i < nin (zero beyond), the
// transformed line written to out[o*outO + i*outS + kx] for i < nout.  Tiles
// are whole (the row pitch is a multiple of W, padding columns hold zeros and
// stay zero), so every row of a tile is one aligned 128-byte line.
// MODE 0: forward; MODE 1: inverse; MODE 2: forward, *Ghat, inverse (z-conv);
// MODE 3: image charge; MODE 4: periodic z (whole lines, complex Gper, no flip).
template <typename T, int E, int MODE>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_strided(const typename cplx_of<T>::type *in,
                                                         typename cplx_of<T>::type *out, FftPlan pl,
                                                         int nin, int nout, size_t inS, size_t inO,
                                                         size_t outS, size_t outO, const T *Gh,
                                                         int gp, int n1, int n2, int L1, size_t inB, size_t outB,
                                                         size_t ghB) {
    using C = typename cplx_of<T>::type;
    int bx, o, bz;
    xcd_block(bx, o, bz);
    in += bz * inB;
    out += bz * outB;
    if (Gh) Gh += bz * ghB;
    constexpr int Q = (MODE == 1 || MODE == 4) ? E : E / 2;  // whole-line passes read E per thread
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    const int W = pl.W, lw = pl.lw, P = pl.P, L = pl.L, total = W * nin;
    const int kx0 = bx * W;
    const size_t ib = (size_t)o * inO + kx0, ob = (size_t)o * outO + kx0;
    const C *tw = fft_twiddles(buf, pl);
    // the convolution passes only: the fused single transforms of the y passes
    // measured slower than the LDS path (0.103 against 0.097 ms at 256^3)
    if (MODE >= 2 && pl.nst >= 2 && pl.radix[0] == pl.radix[pl.nst - 1]) {
        const int fy = o <= n1 ? o : L1 - o;  // the y row of the factors
        const int R0 = pl.radix[0];
        if constexpr (E % 8 == 0)
            if (R0 == 8) {
                strided_fused<T, E, 8, MODE>(in, out, buf, tw, pl, ib, ob, inS, outS, nin, nout, Gh, gp, n1, n2, fy,
                                             kx0);
                return;
            }
        if constexpr (E % 4 == 0)
            if (R0 == 4) {
                strided_fused<T, E, 4, MODE>(in, out, buf, tw, pl, ib, ob, inS, outS, nin, nout, Gh, gp, n1, n2, fy,
                                             kx0);
                return;
            }
        if constexpr (E % 2 == 0)
            if (R0 == 2) {
                strided_fused<T, E, 2, MODE>(in, out, buf, tw, pl, ib, ob, inS, outS, nin, nout, Gh, gp, n1, n2, fy,
                                             kx0);
                return;
            }
    }
    C r[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = min(threadIdx.x + q * kFftThreads, total - 1);
        r[q] = in[ib + (size_t)(e >> lw) * inS + (e & (W - 1))];
    }
    // MODE 2 (o is ky; Ghat folded into the octant, Ghat[2n-k] = Ghat[k]):
    // the thread's Ghat factors are loaded with the input, so that their
    // latency overlaps the input's instead of following the forward transform,
    // when E of them fit in 16 registers; larger tiles load them after the
    // transform, in halves
    constexpr bool GPRE = MODE == 2 && E * sizeof(T) <= 64;
    T g[GPRE ? E : 1];
    const int fy = o <= n1 ? o : L1 - o;
    auto gload = [&](int q) -> T {
        const int e = threadIdx.x + q * kFftThreads;
        const int kz = e >> lw;
        const int fz = kz <= n2 ? kz : L - kz;
        return Gh[((size_t)fz * (n1 + 1) + fy) * gp + kx0 + (e & (W - 1))];
    };
    if constexpr (GPRE)
#pragma unroll
        for (int q = 0; q < E; ++q) g[q] = gload(q);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kFftThreads;
        if (e < total) buf[__umul24(e >> lw, P) + (e & (W - 1))] = r[q];
    }
    for (int e = threadIdx.x; e < W * (L - nin); e += kFftThreads) {
        const int ln = e & (W - 1), i = nin + (e >> lw);
        buf[__umul24(i, P) + ln] = C{T(0), T(0)};
    }
    __syncthreads();
    if constexpr (MODE == 2) {
        fft_lds<C, E>(buf, tw, pl, false);
        if constexpr (GPRE) {
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int e = threadIdx.x + q * kFftThreads;
                const int idx = __umul24(e >> lw, P) + (e & (W - 1));
                const C v = buf[idx];
                buf[idx] = C{v.x * g[q], v.y * g[q]};
            }
        } else {
            constexpr int H = E / 2;
#pragma unroll
            for (int h = 0; h < E; h += H) {
                T gh[H];
#pragma unroll
                for (int q = 0; q < H; ++q) gh[q] = gload(h + q);
#pragma unroll
                for (int q = 0; q < H; ++q) {
                    const int e = threadIdx.x + (h + q) * kFftThreads;
                    const int idx = __umul24(e >> lw, P) + (e & (W - 1));
                    const C v = buf[idx];
                    buf[idx] = C{v.x * gh[q], v.y * gh[q]};
                }
            }
        }
        __syncthreads();
        fft_lds<C, E>(buf, tw, pl, true);
    } else if constexpr (MODE == 3 || MODE == 4) {
        // image charge: times -Gi (complex, not even in z; y folded), FFTPoissonSolver.cpp:316;
        // periodic z: times Gper (complex, one-sided in z; y folded)
        fft_lds<C, E>(buf, tw, pl, false);
        const C *Gi = (const C *)Gh;
        const int fy = o <= n1 ? o : L1 - o;
        C g[E];
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = threadIdx.x + q * kFftThreads;
            g[q] = Gi[((size_t)(e >> lw) * (n1 + 1) + fy) * gp + kx0 + (e & (W - 1))];
        }
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = threadIdx.x + q * kFftThreads;
            const int idx = __umul24(e >> lw, P) + (e & (W - 1));
            const C v = buf[idx];
            const C p = C{v.x * g[q].x - v.y * g[q].y, v.x * g[q].y + v.y * g[q].x};
            buf[idx] = MODE == 3 ? C{-p.x, -p.y} : p;
        }
        __syncthreads();
        fft_lds<C, E>(buf, tw, pl, true);
    } else {
        fft_lds<C, E>(buf, tw, pl, MODE == 1);
    }
    for (int e = threadIdx.x; e < W * nout; e += kFftThreads) {
        const int ln = e & (W - 1), i = e >> lw;
        // the image potential is read back flipped in z (FFTPoissonSolver.cpp:326-329)
        const int io = MODE == 3 ? nout - 1 - i : i;
        out[ob + (size_t)io * outS + ln] = buf[__umul24(i, P) + ln];
    }
}

// c2r along x, keep x < n0, phi = ((raw*invM)*s1)*s2.  OUTSCALE 1: s1 = vol,
// s2 = coupling (PartBunch.cpp:542-554); OUTSCALE 0: s1 = s2 = 1.
template <typename T, int E, int OUTSCALE>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_xinv(const typename cplx_of<T>::type *A, T *phi,
                                                      FftPlan pl, int n0, int nlines, int kp,
                                                      double invM, const Scalars *sc, size_t aB, size_t phiB) {
    using C = typename cplx_of<T>::type;
    constexpr int Q = half_q<E>();  // W line pairs of K0 = L/2 + 1
    sc += blockIdx.z;
    A += blockIdx.z * aB;
    phi += blockIdx.z * phiB;
    extern __shared__ char smem[];
    C *buf = (C *)smem;
    const int W = pl.W, P = pl.P, L = pl.L, K0 = n0 + 1, total = W * K0;
    const float rn = 1.0f / (float)n0, rk = 1.0f / (float)K0;
    const int line0 = blockIdx.x * 2 * W;
    const C *tw = fft_twiddles(buf, pl);
    C ra[Q], rb[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = min(threadIdx.x + q * kFftThreads, total - 1);
        const int p = div_small(e, K0, rk), k = e - p * K0;
        ra[q] = A[(size_t)min(line0 + 2 * p, nlines - 1) * kp + k];
        rb[q] = A[(size_t)min(line0 + 2 * p + 1, nlines - 1) * kp + k];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kFftThreads;
        if (e < total) {
            const int p = div_small(e, K0, rk), k = e - p * K0;
            C a = ra[q], b = rb[q];
            if (k == 0 || k == n0) a.y = b.y = T(0);
            buf[__umul24(k, P) + p] = C{a.x - b.y, a.y + b.x};
            if (k >= 1 && k <= n0 - 1) buf[__umul24(L - k, P) + p] = C{a.x + b.y, b.x - a.y};
        }
    }
    __syncthreads();
    fft_lds<C, E>(buf, tw, pl, true);
    double s1 = 1.0, s2 = 1.0;
    if constexpr (OUTSCALE == 1) {
        s1 = sc->vol;
        s2 = sc->coupling;
    }
    const T *bufr = (const T *)smem;
    for (int e = threadIdx.x; e < 2 * W * n0; e += kFftThreads) {
        const int ln = div_small(e, n0, rn), x = e - ln * n0;
        const int line = line0 + ln;
        if (line < nlines) {
            const T v = bufr[2 * (__umul24(x, P) + (ln >> 1)) + (ln & 1)];
            if constexpr (sizeof(T) == 8) {
                double d = v * invM;
                d = d * s1;
                d = d * s2;
                phi[(size_t)line * n0 + x] = d;
            } else {
                float d = v * (float)invM;
                d = d * (float)s1;
                d = d * (float)s2;
                phi[(size_t)line * n0 + x] = d;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// E = -Grad(phi) * (gz, gz, 1/gz)   (PartBunch.cpp:573-575; Grad Cell->Cell,
// ippl/src/Meshes/UniformCartesian.hpp:2127-2200; ZeroFace guards)
// ---------------------------------------------------------------------------
// Grad at cell (i, j, k) of phi (p points at it): central differences, the
// one-sided second-order stencils on the border faces of each component.
template <typename T>
__device__ inline void grad_stencil(const T *p, int i, int j, int k, int n0, int n1, int n2, size_t sy, size_t sz,
                                    T id0, T id1, T id2, T &r0, T &r1, T &r2) {
    if (i == 0) r0 = id0 * ((T(-1.0) * p[2] + T(4.0) * p[1]) - T(3.0) * p[0]);
    else if (i == n0 - 1) r0 = id0 * ((T(1.0) * p[-2] - T(4.0) * p[-1]) + T(3.0) * p[0]);
    else r0 = id0 * (p[1] - p[-1]);
    if (j == 0) r1 = id1 * ((T(-1.0) * p[2 * sy] + T(4.0) * p[sy]) - T(3.0) * p[0]);
    else if (j == n1 - 1) r1 = id1 * ((T(1.0) * p[-2 * (ptrdiff_t)sy] - T(4.0) * p[-(ptrdiff_t)sy]) + T(3.0) * p[0]);
    else r1 = id1 * (p[sy] - p[-(ptrdiff_t)sy]);
    if (k == 0) r2 = id2 * ((T(-1.0) * p[2 * sz] + T(4.0) * p[sz]) - T(3.0) * p[0]);
    else if (k == n2 - 1) r2 = id2 * ((T(1.0) * p[-2 * (ptrdiff_t)sz] - T(4.0) * p[-(ptrdiff_t)sz]) + T(3.0) * p[0]);
    else r2 = id2 * (p[sz] - p[-(ptrdiff_t)sz]);
}

// Cell tiles of the Grad kernels: a block of 256 threads covers tw consecutive
// i (a power of two, lane-fastest) by 256/tw rows j of plane k = blockIdx.y;
// blockIdx.x = bj * nbi + bi.  No per-cell integer division.
struct CellTile {
    int lw, nbi;  // log2 of the tile width; tiles along i
};
static CellTile cell_tile(int n0) {
    int lw = 6;  // 64 cells, narrower for narrow grids
    while (lw > 3 && (1 << (lw - 1)) >= n0) --lw;
    return CellTile{lw, (n0 + (1 << lw) - 1) >> lw};
}
__device__ inline bool tile_cell(const CellTile &t, int n0, int n1, int &i, int &j) {
    const int bj = (int)blockIdx.x / t.nbi, bi = (int)blockIdx.x - bj * t.nbi;
    i = (bi << t.lw) + ((int)threadIdx.x & ((1 << t.lw) - 1));
    j = bj * (256 >> t.lw) + ((int)threadIdx.x >> t.lw);
    return i < n0 && j < n1;
}

// grid b = blockIdx.z of a batch: phi + b*total, Eg + b*total, scalars sc + b
template <typename T>
__global__ __launch_bounds__(256) void k_grad(const T *phi, typename vec4_of<T>::type *Eg, int n0,
                                              int n1, int n2, const Scalars *sc, CellTile t) {
    const size_t total = (size_t)n0 * n1 * n2;
    phi += blockIdx.z * total;
    Eg += blockIdx.z * total;
    sc += blockIdx.z;
    const T id0 = (T)(0.5 / sc->hr[0]), id1 = (T)(0.5 / sc->hr[1]), id2 = (T)(0.5 / sc->hr[2]);
    const T g0 = (T)sc->gscale[0], g1 = (T)sc->gscale[1], g2 = (T)sc->gscale[2];
    const size_t sy = n0, sz = (size_t)n0 * n1;
    int i, j;
    if (!tile_cell(t, n0, n1, i, j)) return;
    const int k = blockIdx.y;
    const size_t c = (size_t)k * sz + (size_t)j * sy + i;
    T r0, r1, r2;
    grad_stencil(phi + c, i, j, k, n0, n1, n2, sy, sz, id0, id1, id2, r0, r1, r2);
    typename vec4_of<T>::type e;
    e.x = -r0 * g0;
    e.y = -r1 * g1;
    e.z = -r2 * g2;
    e.w = T(0);
    Eg[c] = e;
}

// The energy bins' fields summed on the grid in one pass: for each of the nb
// bins (scalars scb[b], potential phi + b*phiB, and its image's phii + b*phiB
// when phii is given), E_b = -Grad(phi_b) * (g_b, g_b, 1/g_b) and the magnetic
// field betaC_b * (-E_y, E_x) of the bin (betaC_b * (E_y, -E_x) of its image),
// added in bin order (bin, then its image) as the per-bin accumulation adds
// them; Eg and Bg are written once.  The bins' factors are formed once per
// block, in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_field_sums(const T *phi, const T *phii, size_t phiB, int nb,
                                                    typename vec4_of<T>::type *Eg, typename vec4_of<T>::type *Bg,
                                                    int n0, int n1, int n2, const Scalars *scb, CellTile t) {
    using V = typename vec4_of<T>::type;
    __shared__ T fac[kMaxBins][7];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const Scalars *sc = scb + b;
        fac[b][0] = (T)(0.5 / sc->hr[0]);
        fac[b][1] = (T)(0.5 / sc->hr[1]);
        fac[b][2] = (T)(0.5 / sc->hr[2]);
        fac[b][3] = (T)sc->gscale[0];
        fac[b][4] = (T)sc->gscale[1];
        fac[b][5] = (T)sc->gscale[2];
        fac[b][6] = (T)sc->betaC;
    }
    __syncthreads();
    const size_t sy = n0, sz = (size_t)n0 * n1;
    int i, j;
    if (!tile_cell(t, n0, n1, i, j)) return;
    const int k = blockIdx.y;
    const size_t c = (size_t)k * sz + (size_t)j * sy + i;
    V E = {T(0), T(0), T(0), T(0)}, B = {T(0), T(0), T(0), T(0)};
    for (int b = 0; b < nb; ++b) {
        const T id0 = fac[b][0], id1 = fac[b][1], id2 = fac[b][2];
        const T g0 = fac[b][3], g1 = fac[b][4], g2 = fac[b][5];
        for (int img = 0; img < (phii ? 2 : 1); ++img) {
            const T *p = (img ? phii : phi) + b * phiB + c;
            const T bs = img ? fac[b][6] : -fac[b][6];
            T r0, r1, r2;
            grad_stencil(p, i, j, k, n0, n1, n2, sy, sz, id0, id1, id2, r0, r1, r2);
            const T ex = -r0 * g0, ey = -r1 * g1, ez = -r2 * g2;
            E.x += ex;
            E.y += ey;
            E.z += ez;
            B.x += bs * ey;
            B.y -= bs * ex;
        }
    }
    Eg[c] = E;
    Bg[c] = B;
}

// rho entering the solver, for diagnostics
__global__ void k_rho_to_double(const long long *rho, double *out, size_t n, const Scalars *sc, int two_words) {
    const double q2e = sc->q2e, q2e_lo = sc->q2e_lo, dT = sc->dT, tmp2 = sc->tmp2;
    const int lsh = sc->lo_shift;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        double d;
        if (two_words) {  // canonical words, as the x pass reads them
            const long long hi = rho[2 * c] + (rho[2 * c + 1] >> lsh);
            const long long lo = rho[2 * c + 1] & (((long long)1 << lsh) - 1);
            d = (double)hi * q2e + (double)lo * q2e_lo;
        } else {
            d = (double)rho[c] * q2e;
        }
        d = d / dT;
        d = d * tmp2;
        out[c] = d;
    }
}

template <typename T>
__global__ void k_widen(const T *in, double *out, size_t n, int comps, int stride) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x)
        for (int d = 0; d < comps; ++d) out[c * comps + d] = (double)in[c * stride + d];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// ---- tile shapes: the autotuned table, overrides, the default rule ----------
// Tiles are chosen per transform length L and precision by E, the elements per
// thread (W = 512 E / L lines per workgroup).  The table the autotuner wrote
// (src_amd/tuning.json, compiled in as tuning.hpp) comes first, then
// overrides set through opalsc_tune_fft_set, then the default rule: the
// smallest E whose tile rows are whole 128-byte lines.
struct TileEntry {
    int L, fp64, E;
};
static const TileEntry kTunedTiles[] = {OPALSC_TUNED_FFT_TILES{0, 0, 0}};
static std::vector<TileEntry> g_tile_override;

// E values the kernels can take for length L (radices 8/4/2 then 3 dividing
// E, W a power of two, the tile within the register budget and within the
// 160 KiB of LDS a gfx950 workgroup can hold: the tile of pitch W + 1, and the
// FP32 twiddles)
// the factorisation L = 2^a * odd, odd in {1, 3, 5}; false otherwise
static bool split_length(int L, int &a, int &odd) {
    a = 0;
    int m = L;
    while (m > 0 && m % 2 == 0) { m /= 2; ++a; }
    odd = m;
    return a >= 2 && (m == 1 || m == 3 || m == 5);
}

static int tile_candidates(int L, bool fp64, int *out, int cap) {
    int a = 0, odd = 1;
    if (!split_length(L, a, odd)) return -1;
    // E a multiple of every radix: the odd one and 2 (E even)
    static const int c2[] = {4, 8, 16, 32}, c3[] = {6, 12, 24, 48}, c5[] = {10, 20, 40, 0};
    const int *cand = odd == 3 ? c3 : odd == 5 ? c5 : c2;
    const int emax = fp64 ? (odd == 1 ? 16 : odd == 3 ? 24 : 20) : (odd == 1 ? 32 : odd == 3 ? 48 : 40);
    int k = 0;
    for (int c = 0; c < 4; ++c) {
        const int e = cand[c];
        if (e == 0 || e > emax || (e * kFftThreads) % L != 0) continue;
        const int W = e * kFftThreads / L;
        if (W < 1 || (W & (W - 1)) != 0) continue;
        const size_t cs = fp64 ? 16 : 8;
        if ((size_t)L * (W + 1) * cs + (fp64 ? 0 : (size_t)L * cs) > 160 * 1024) continue;
        if (k < cap) out[k] = e;
        ++k;
    }
    return k;
}

static int tile_lookup(int L, bool fp64) {
    for (const TileEntry &t : g_tile_override)
        if (t.L == L && t.fp64 == (fp64 ? 1 : 0)) return t.E;
    for (const TileEntry &t : kTunedTiles)
        if (t.L == L && t.fp64 == (fp64 ? 1 : 0)) return t.E;
    return 0;
}

// Tile shape for length L: the tuned or overridden E if the kernels take it,
// else the smallest E whose tile has W >= 128 bytes / complex size lines,
// within the register budget.  LDS pitch: the x passes fill and drain their
// tiles along the transform (lanes of a wave P elements apart), which an odd
// pitch W + 1 spreads over the banks; the strided (y, z) passes fill, drain
// and run every Stockham stage across the tile's W lines (lanes 1 element
// apart, rows of W), where the pitch W keeps two rows of a W = 8 FP32 tile
// on disjoint banks and W + 1 would overlap them.
static bool make_plan(int L, bool fp64, bool strided, FftPlan &pl) {
    int a = 0, odd = 1;
    if (!split_length(L, a, odd)) return false;
    pl.L = L;
    const int wmin = fp64 ? 8 : 16;
    int cand[8];
    const int nc = tile_candidates(L, fp64, cand, 8);
    pl.E = 0;
    const int tuned = tile_lookup(L, fp64);
    for (int c = 0; c < nc && tuned; ++c)
        if (cand[c] == tuned) pl.E = tuned;
    for (int c = 0; c < nc && pl.E == 0; ++c) {
        if (cand[c] * kFftThreads / L >= wmin || c == nc - 1) pl.E = cand[c];
    }
    if (pl.E == 0) return false;
    pl.W = pl.E * kFftThreads / L;
    if (pl.W < 1) return false;
    pl.lw = 0;
    while ((1 << pl.lw) < pl.W) ++pl.lw;
    if ((1 << pl.lw) != pl.W) return false;
    pl.P = strided ? pl.W : pl.W + 1;
    pl.nst = 0;
    // every radix must divide E (each thread owns E/R whole butterflies)
    const bool use8 = (pl.E % 8) == 0, use4 = (pl.E % 4) == 0;
    int rem = a;
    while (rem >= 3 && use8) { pl.radix[pl.nst++] = 8; rem -= 3; }
    while (rem >= 2 && use4) { pl.radix[pl.nst++] = 4; rem -= 2; }
    while (rem >= 1) { pl.radix[pl.nst++] = 2; rem -= 1; }
    // the powers of two in any order: the first radix also last where it occurs
    // twice, so that a convolution pass can keep the last forward and the first
    // inverse stage in registers (strided_fused)
    if (pl.nst >= 3 && pl.radix[1] == pl.radix[0]) {
        const int r0 = pl.radix[0];
        for (int k = 1; k < pl.nst - 1; ++k) pl.radix[k] = pl.radix[k + 1];
        pl.radix[pl.nst - 1] = r0;
    }
    if (odd > 1) pl.radix[pl.nst++] = odd;  // last: every earlier span is a power of two
    if (pl.nst > 16) return false;
    for (int s = 0; s < pl.nst; ++s)
        if (pl.E % pl.radix[s] != 0) return false;
    return true;
}

template <typename T>
static void make_twiddles(int L, std::vector<char> &out) {
    using C = typename cplx_of<T>::type;
    out.resize(sizeof(C) * L);
    C *p = (C *)out.data();
    for (int m = 0; m < L; ++m) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)L;
        p[m].x = (T)cosl(ang);
        p[m].y = (T)sinl(ang);
    }
}

// tile + (FP32) twiddle table
static size_t lds_bytes(const FftPlan &pl, size_t csize) {
    return (size_t)pl.L * pl.P * csize + (csize == 8 ? (size_t)pl.L * csize : 0);
}

#define OPALSC_DISPATCH_E(E_, ...)                             \
    switch (E_) {                                              \
    case 4: { constexpr int EV = 4; __VA_ARGS__; } break;      \
    case 6: { constexpr int EV = 6; __VA_ARGS__; } break;      \
    case 8: { constexpr int EV = 8; __VA_ARGS__; } break;      \
    case 10: { constexpr int EV = 10; __VA_ARGS__; } break;    \
    case 12: { constexpr int EV = 12; __VA_ARGS__; } break;    \
    case 16: { constexpr int EV = 16; __VA_ARGS__; } break;    \
    case 20: { constexpr int EV = 20; __VA_ARGS__; } break;    \
    case 24: { constexpr int EV = 24; __VA_ARGS__; } break;    \
    case 32: if constexpr (sizeof(T) == 4) { constexpr int EV = 32; __VA_ARGS__; } break; \
    case 40: if constexpr (sizeof(T) == 4) { constexpr int EV = 40; __VA_ARGS__; } break; \
    case 48: if constexpr (sizeof(T) == 4) { constexpr int EV = 48; __VA_ARGS__; } break; \
    default: break;                                            \
    }

int PoissonGrid::init(const int nn[3], bool fp64_, int greens_, bool periodic_z, hipStream_t s,
                      std::string &err) {
    stream = s;
    fp64 = fp64_;
    greens = greens_;
    pz = periodic_z;
    for (int d = 0; d < 3; ++d) {
        n[d] = nn[d];
        // periodic z: the z transform has the grid's own length (FFTPoissonSolver.cpp:130-155)
        L[d] = (pz && d == 2) ? nn[d] : 2 * nn[d];
        if (nn[d] < 4) { err = "n must be >= 4 in every dimension"; return -1; }
        if (!make_plan(L[d], fp64, d != 0, plan[d])) {
            err = "unsupported grid size n=" + std::to_string(nn[d]) +
                  ((pz && d == 2) ? " for periodic z (need n = 2^a, 3*2^a or 5*2^a, a>=2)"
                                  : " (need n = 2^a, 3*2^a or 5*2^a, a>=1)");
            return -1;
        }
    }
    K0 = n[0] + 1;
    const size_t cs = fp64 ? 16 : 8, rs = fp64 ? 8 : 4;
    int lds_max = 0;
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            lds_max = 65536;
    }
    for (int d = 0; d < 3; ++d)
        if (lds_bytes(plan[d], cs) > (size_t)lds_max) {
            err = "grid size n=" + std::to_string(n[d]) + " needs more LDS per block than the device has";
            return -1;
        }
    // tile-aligned pitches: whole 128-byte rows in every strided pass
    const int tilew = std::max(std::max(plan[1].W, plan[2].W), (int)(128 / cs));
    kp = (K0 + tilew - 1) / tilew * tilew;
    // Ghat rows: whole 128-byte lines and whole tiles of the DCT passes (2W real columns)
    const int gtile = std::max((int)(128 / rs), 2 * std::max(plan[1].W, plan[2].W));
    gp = (kp + gtile - 1) / gtile * gtile;
    ncell = (size_t)n[0] * n[1] * n[2];
    noct = (size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    const size_t szA = (size_t)kp * n[1] * n[2] * cs;
    const size_t szB = (size_t)kp * L[1] * n[2] * cs;
    const size_t szG = (size_t)gp * (n[1] + 1) * (n[2] + 1) * rs;
    auto alloc = [&](void **p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) { err = "hipMalloc failed (" + std::to_string(bytes) + " bytes)"; return false; }
        return true;
    };
    if (!alloc(&rho_i64, rho_words() * 8) || !alloc(&rho_f64, ncell * 8) || !alloc(&A, szA) ||
        !alloc(&B, szB) || !alloc(&Goct, noct * 8) ||
        (fp64 && !alloc(&Tcorner, (size_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2) * 8)) || !alloc(&Ghat, szG) ||
        !alloc(&phi, ncell * rs) || !alloc(&Eg, ncell * 4 * rs))
        return -1;
    // padding columns are read as whole tiles and must hold (and keep) zeros
    if (hipMemset(A, 0, szA) != hipSuccess || hipMemset(B, 0, szB) != hipSuccess ||
        hipMemset(Ghat, 0, szG) != hipSuccess) {
        err = "hipMemset failed";
        return -1;
    }
    if (pz) {
        const size_t szP = (size_t)L[2] * (n[1] + 1) * gp * cs;
        if (!alloc(&Gper, szP)) return -1;
        if (hipMemset(Gper, 0, szP) != hipSuccess) {
            err = "hipMemset failed";
            return -1;
        }
    }
    for (int d = 0; d < 3; ++d) {
        std::vector<char> tw;
        if (fp64) make_twiddles<double>(L[d], tw);
        else make_twiddles<float>(L[d], tw);
        if (!alloc(&twd[d], tw.size())) return -1;
        if (hipMemcpy(twd[d], tw.data(), tw.size(), hipMemcpyHostToDevice) != hipSuccess) {
            err = "twiddle upload failed";
            return -1;
        }
        plan[d].tw = twd[d];
    }
    return 0;
}

void PoissonGrid::destroy() {
    void *ps[] = {rho_i64, rho_f64, A, B, Goct, Tcorner, Ghat, phi, Eg, twd[0], twd[1], twd[2],
                  HimgA, HimgB, Gimg, Esum, Bsum, Gper, bb_rho, bb_A, bb_B, bb_Bi, bb_phi, bb_phii, bb_T, bb_G,
                  bb_Gh, bb_HA, bb_HB, bb_Gi, bb_E};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    bb_rho = bb_A = bb_B = bb_Bi = bb_phi = bb_phii = bb_T = bb_G = bb_Gh = bb_HA = bb_HB = bb_Gi = bb_E = nullptr;
    bin_cap = 0;
    bin_img_cap = false;
    rho_i64 = rho_f64 = A = B = Goct = Tcorner = Ghat = phi = Eg = nullptr;
    HimgA = HimgB = Gimg = Esum = Bsum = Gper = nullptr;
    twd[0] = twd[1] = twd[2] = nullptr;
}

// The buffers a Green's function or a solve works on: nb grids of this
// shape, grid b's array starting b*stride elements after grid 0's (nb = 1,
// strides 0: the context's own grid).  Every kernel takes grid b as
// blockIdx.z and its scalars as d_sc + b.
struct GridSet {
    int nb = 1;
    const void *rho = nullptr;  // x-pass source (fixed-point words or doubles), rhoB cells apart
    size_t rhoB = 0;
    void *A = nullptr, *B = nullptr, *Bi = nullptr;  // spectra (complex), AB / BB elements apart
    size_t AB = 0, BB = 0;
    void *phi = nullptr, *phii = nullptr;  // potential of the charge and of its image, phiB apart
    size_t phiB = 0;
    double *T = nullptr, *Goct = nullptr;  // corner table, octant (FP64)
    size_t TB = 0, GB = 0;
    void *Ghat = nullptr, *HA = nullptr, *HB = nullptr;  // real transforms (grid precision)
    size_t GhB = 0;
    void *Gimg = nullptr;  // image: complex transform over the doubled z
    size_t GiB = 0;
};

static GridSet own_set(PoissonGrid &g, int src_kind) {
    GridSet b;
    b.rho = src_kind == 0 ? g.rho_i64 : g.rho_f64;
    b.A = g.A;
    b.B = g.B;
    b.phi = g.phi;
    b.T = (double *)g.Tcorner;
    b.Goct = (double *)g.Goct;
    b.Ghat = g.Ghat;
    b.HA = g.HimgA;
    b.HB = g.HimgB;
    b.Gimg = g.Gimg;
    return b;
}

// DCT-I along x (octant G, FP64, pitch n0+1 -> H, pitch gp) and along y, in place in H.
// table >= 0: the octant of IGF table `table` evaluated in the x pass (G unused).
template <typename T>
static void dct_xy(PoissonGrid &g, const GridSet &gs, const double *G, T *H, hipStream_t st, int table = -1,
                   const Scalars *d_sc = nullptr) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C), gp = (size_t)g.gp;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    {
        const FftPlan &pl = g.plan[0];
        int nlines = (n1 + 1) * (n2 + 1);
        dim3 grid((nlines + 2 * pl.W - 1) / (2 * pl.W), 1, gs.nb);
        if (table >= 0) {
            OPALSC_DISPATCH_E(pl.E, k_dct_x<T, EV, true><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                        nullptr, H, pl, n0, nlines, g.gp, 0, gs.GhB, d_sc, table, n1, n2));
        } else {
            OPALSC_DISPATCH_E(pl.E, k_dct_x<T, EV><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                        G, H, pl, n0, nlines, g.gp, gs.GB, gs.GhB));
        }
    }
    {
        const FftPlan &pl = g.plan[1];
        dim3 grid((n0 + 1 + 2 * pl.W - 1) / (2 * pl.W), n2 + 1, gs.nb);
        OPALSC_DISPATCH_E(pl.E, k_dct_strided<T, EV><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                    H, pl, n1 + 1, gp, gp * (n1 + 1), gs.GhB, n0 + 1));
    }
}

// The IGF octant of table 0 (integratedGreensFunction, with the G[0,0,0] quirk)
// or of the image charge's table 1 / 2 (shiftedIntGreensFunction's first and
// second corner tables) into gs.Goct: on the FP64 grid the corner table with
// correctly rounded atan/log and its 8-corner difference, on the mixed grid
// the one-pass octant.
static void igf_octant(PoissonGrid &g, const GridSet &gs, const Scalars *d_sc, int table, hipStream_t st) {
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    if (!g.fp64) {
        launch_green_fast(st, gs.Goct, g.n, d_sc, table != 0, table == 2, gs.nb, gs.GB);
        return;
    }
    dim3 grid(((n0 + 2) * (n1 + 2) + 255) / 256, n2 + 2, gs.nb);
    if (table == 0) k_igf_corners<<<grid, 256, 0, st>>>(gs.T, n0, n1, n2, d_sc, gs.TB);
    else k_igf_corners_shifted<<<grid, 256, 0, st>>>(gs.T, n0, n1, n2, d_sc, table == 2, gs.TB);
    dim3 grid2(((n0 + 1) * (n1 + 1) + 255) / 256, (n2 + kIgfDiffK) / kIgfDiffK, gs.nb);
    k_igf_diff<<<grid2, 256, 0, st>>>(gs.T, gs.Goct, n0, n1, n2, table == 0, gs.TB, gs.GB);
}

template <typename T>
static void green_impl(PoissonGrid &g, const GridSet &gs, const Scalars *d_sc, hipStream_t st) {
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const bool igf = g.greens != 1;
    // the mixed grid's IGF octant is evaluated inside the x pass of the DCT
    const bool fused = igf && !g.fp64;
    if (!igf) {
        k_green_standard<<<dim3(1024, 1, gs.nb), 256, 0, st>>>(gs.Goct, n0, n1, n2, d_sc, gs.GB);
    } else if (!fused) {
        igf_octant(g, gs, d_sc, 0, st);
    }
    g.profile_mark(PROF_GREEN);
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C);
    const size_t gp = (size_t)g.gp;
    dct_xy<T>(g, gs, gs.Goct, (T *)gs.Ghat, st, fused ? 0 : -1, d_sc);
    if (g.pz) {
        // periodic z: the z line is G[k] for k in [0, n2), one-sided (mirrorRhoField,
        // FFTPoissonSolver.cpp:661-662), so its transform is complex (one grid only)
        const FftPlan &pl = g.plan[2];
        dim3 grid((n0 + 1 + pl.W - 1) / pl.W, n1 + 1);
        OPALSC_DISPATCH_E(pl.E, k_green_zfft<T, EV><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                    (const T *)g.Ghat, (const T *)g.Ghat, (C *)g.Gper, pl, n1, n2, g.gp, 0, 0));
    } else {
        const FftPlan &pl = g.plan[2];
        dim3 grid((n0 + 1 + 2 * pl.W - 1) / (2 * pl.W), n1 + 1, gs.nb);
        OPALSC_DISPATCH_E(pl.E, k_dct_strided<T, EV><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                    (T *)gs.Ghat, pl, n2 + 1, gp * (n1 + 1), gp, gs.GhB, n0 + 1));
    }
    g.profile_mark(PROF_DCT);
}

// Transform of the image-charge Green's function (FFTPoissonSolver::shiftedIntGreensFunction,
// FFTPoissonSolver.cpp:546-639) for the shift in the scalars: both corner
// tables, their differences, DCT-I in x and y, complex FFT along the doubled z.
template <typename T>
static void green_image_impl(PoissonGrid &g, const GridSet &gs, const Scalars *d_sc, hipStream_t st) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C);
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    T *H[2] = {(T *)gs.HA, (T *)gs.HB};
    for (int second = 0; second < 2; ++second) {
        if (g.fp64) {
            igf_octant(g, gs, d_sc, 1 + second, st);
            dct_xy<T>(g, gs, gs.Goct, H[second], st);
        } else {
            dct_xy<T>(g, gs, nullptr, H[second], st, 1 + second, d_sc);  // octant evaluated in the x pass
        }
    }
    g.profile_mark(PROF_GREEN);
    {
        const FftPlan &pl = g.plan[2];
        dim3 grid((n0 + 1 + pl.W - 1) / pl.W, n1 + 1, gs.nb);
        OPALSC_DISPATCH_E(pl.E, k_green_zfft<T, EV><<<grid, kFftThreads, lds_bytes(pl, cs), st>>>(
                                    H[0], H[1], (C *)gs.Gimg, pl, n1, n2, g.gp, gs.GhB, gs.GiB));
    }
    g.profile_mark(PROF_DCT);
}

void PoissonGrid::octant(const Scalars *d_sc, int table, hipStream_t s) {
    if (!s) s = stream;
    igf_octant(*this, own_set(*this, 0), d_sc, table, s);
}

void PoissonGrid::green(const Scalars *d_sc, hipStream_t s) {
    if (!s) s = stream;
    const GridSet gs = own_set(*this, 0);
    if (fp64) green_impl<double>(*this, gs, d_sc, s);
    else green_impl<float>(*this, gs, d_sc, s);
}

int PoissonGrid::ensure_image(std::string &err) {
    if (Gimg) return 0;
    const size_t rs = fp64 ? 8 : 4, cs = 2 * rs;
    const size_t szH = (size_t)gp * (n[1] + 1) * (n[2] + 1) * rs;
    const size_t szG = (size_t)L[2] * (n[1] + 1) * gp * cs;
    if (hipMalloc(&HimgA, szH) != hipSuccess || hipMalloc(&HimgB, szH) != hipSuccess ||
        hipMalloc(&Gimg, szG) != hipSuccess) {
        err = "hipMalloc of the image-charge Green's function failed";
        return -1;
    }
    // padding columns are read as whole tiles and must hold (and keep) zeros
    if (hipMemset(HimgA, 0, szH) != hipSuccess || hipMemset(HimgB, 0, szH) != hipSuccess ||
        hipMemset(Gimg, 0, szG) != hipSuccess) {
        err = "hipMemset failed";
        return -1;
    }
    return 0;
}

void PoissonGrid::green_image(const Scalars *d_sc, hipStream_t s) {
    if (!s) s = stream;
    const GridSet gs = own_set(*this, 0);
    if (fp64) green_image_impl<double>(*this, gs, d_sc, s);
    else green_image_impl<float>(*this, gs, d_sc, s);
}

// The convolution passes (see the top of this file) on a GridSet.
template <typename T>
static void pass_x_fwd(PoissonGrid &g, const GridSet &gs, const Scalars *d_sc, int src_kind) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C);
    const int n0 = g.n[0], nlines = g.n[1] * g.n[2];
    const FftPlan &pl = g.plan[0];
    dim3 grid((nlines + 2 * pl.W - 1) / (2 * pl.W), 1, gs.nb);  // 2W real lines per block
    if (src_kind == 0)
        OPALSC_DISPATCH_E(pl.E, k_xfwd<T, EV, 0><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    gs.rho, (C *)gs.A, pl, n0, nlines, g.kp, d_sc, gs.rhoB, gs.AB))
    else
        OPALSC_DISPATCH_E(pl.E, k_xfwd<T, EV, 1><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    gs.rho, (C *)gs.A, pl, n0, nlines, g.kp, d_sc, gs.rhoB, gs.AB))
    g.profile_mark(PROF_FFT_X);
}

template <typename T>
static void pass_y(PoissonGrid &g, const GridSet &gs, bool inverse, const void *in, void *out) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C), kp = (size_t)g.kp;
    const int n1 = g.n[1], n2 = g.n[2], L1 = g.L[1];
    const FftPlan &pl = g.plan[1];
    dim3 grid((g.K0 + pl.W - 1) / pl.W, n2, gs.nb);
    if (!inverse)
        OPALSC_DISPATCH_E(pl.E, k_strided<T, EV, 0><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    (const C *)in, (C *)out, pl, n1, L1, kp, kp * n1, kp, kp * L1, nullptr, 0, n1, n2,
                                    L1, gs.AB, gs.BB, 0))
    else
        OPALSC_DISPATCH_E(pl.E, k_strided<T, EV, 1><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    (const C *)in, (C *)out, pl, L1, n1, kp, kp * L1, kp, kp * n1, nullptr, 0, n1, n2,
                                    L1, gs.BB, gs.AB, 0))
    g.profile_mark(PROF_FFT_Y);
}

// MODE 2: times Ghat (in place allowed); 3: image (times -Gimg, z flipped); 4: periodic z
template <typename T>
static void pass_z(PoissonGrid &g, const GridSet &gs, int mode, const void *in, void *out) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C), kp = (size_t)g.kp;
    const int n1 = g.n[1], n2 = g.n[2], L1 = g.L[1];
    const FftPlan &pl = g.plan[2];
    dim3 grid((g.K0 + pl.W - 1) / pl.W, L1, gs.nb);
    const C *ci = (const C *)in;
    C *co = (C *)out;
    if (mode == 4)
        OPALSC_DISPATCH_E(pl.E, k_strided<T, EV, 4><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    ci, co, pl, n2, n2, kp * L1, kp, kp * L1, kp, (const T *)g.Gper, g.gp, n1, n2, L1,
                                    gs.BB, gs.BB, 0))
    else if (mode == 3)
        OPALSC_DISPATCH_E(pl.E, k_strided<T, EV, 3><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    ci, co, pl, n2, n2, kp * L1, kp, kp * L1, kp, (const T *)gs.Gimg, g.gp, n1, n2, L1,
                                    gs.BB, gs.BB, 2 * gs.GiB))
    else
        OPALSC_DISPATCH_E(pl.E, k_strided<T, EV, 2><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    ci, co, pl, n2, n2, kp * L1, kp, kp * L1, kp, (const T *)gs.Ghat, g.gp, n1, n2, L1,
                                    gs.BB, gs.BB, gs.GhB))
    g.profile_mark(PROF_FFT_Z);
}

template <typename T>
static void pass_x_inv(PoissonGrid &g, const GridSet &gs, const Scalars *d_sc, int outscale, void *phi) {
    using C = typename cplx_of<T>::type;
    const size_t cs = sizeof(C);
    const int n0 = g.n[0], nlines = g.n[1] * g.n[2];
    const FftPlan &pl = g.plan[0];
    dim3 grid((nlines + 2 * pl.W - 1) / (2 * pl.W), 1, gs.nb);  // 2W real lines per block
    const double invM = 1.0 / ((double)g.L[0] * (double)g.L[1] * (double)g.L[2]);
    if (outscale)
        OPALSC_DISPATCH_E(pl.E, k_xinv<T, EV, 1><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    (const C *)gs.A, (T *)phi, pl, n0, nlines, g.kp, invM, d_sc, gs.AB, gs.phiB))
    else
        OPALSC_DISPATCH_E(pl.E, k_xinv<T, EV, 0><<<grid, kFftThreads, lds_bytes(pl, cs), g.stream>>>(
                                    (const C *)gs.A, (T *)phi, pl, n0, nlines, g.kp, invM, d_sc, gs.AB, gs.phiB))
    g.profile_mark(PROF_FFT_X);
}

template <typename T>
static void solve_impl(PoissonGrid &g, const Scalars *d_sc, int src_kind, int outscale, bool image,
                       hipEvent_t green_ready) {
    const GridSet gs = own_set(g, src_kind);
    pass_x_fwd<T>(g, gs, d_sc, src_kind);
    pass_y<T>(g, gs, false, gs.A, gs.B);
    // the Green's function, computed on another stream, is read from here on
    if (green_ready) (void)hipStreamWaitEvent(g.stream, green_ready, 0);
    pass_z<T>(g, gs, g.pz ? 4 : image ? 3 : 2, gs.B, gs.B);
    pass_y<T>(g, gs, true, gs.B, gs.A);
    pass_x_inv<T>(g, gs, d_sc, outscale, gs.phi);
}

void PoissonGrid::solve(const Scalars *d_sc, int src_kind, int outscale, bool image, hipEvent_t green_ready) {
    if (fp64) solve_impl<double>(*this, d_sc, src_kind, outscale, image, green_ready);
    else solve_impl<float>(*this, d_sc, src_kind, outscale, image, green_ready);
}

// Per-bin buffers of the batched energy-bin solve (allocated on first use,
// grown when more bins or the image charge are asked for).
int PoissonGrid::ensure_bins(int nb, bool image, std::string &err) {
    if (nb <= bin_cap && (!image || bin_img_cap)) return 0;
    void **ps[] = {&bb_rho, &bb_A, &bb_B, &bb_Bi, &bb_phi, &bb_phii, &bb_T, &bb_G, &bb_Gh, &bb_HA, &bb_HB, &bb_Gi, &bb_E};
    for (void **q : ps) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    bin_cap = 0;
    bin_img_cap = false;
    const size_t rs = fp64 ? 8 : 4, cs = 2 * rs, m = (size_t)nb;
    const size_t szA = (size_t)kp * n[1] * n[2] * cs, szB = (size_t)kp * L[1] * n[2] * cs;
    // the corner table only for the FP64 solve (the mixed one evaluates G in one pass)
    const size_t szT = fp64 ? (size_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2) * 8 : 0, szGh = (size_t)gp * (n[1] + 1) * (n[2] + 1) * rs;
    const size_t szGi = (size_t)L[2] * (n[1] + 1) * gp * cs;
    struct {
        void **p;
        size_t bytes;
        bool zero;
    } want[] = {{&bb_rho, rho_words() * 8, false}, {&bb_A, szA, true},      {&bb_B, szB, true},
                {&bb_phi, ncell * rs, false},      {&bb_T, szT, false},     {&bb_G, noct * 8, false},
                {&bb_Gh, szGh, true},              {&bb_Bi, image ? szB : 0, true},
                {&bb_phii, image ? ncell * rs : 0, false}, {&bb_HA, image ? szGh : 0, true},
                {&bb_HB, image ? szGh : 0, true},  {&bb_Gi, image ? szGi : 0, true}};
    for (auto &w : want) {
        if (w.bytes == 0) continue;
        if (hipMalloc(w.p, m * w.bytes) != hipSuccess) {
            err = "hipMalloc of the energy-bin grids failed (" + std::to_string(m * w.bytes) + " bytes)";
            return -1;
        }
        // padding columns are read as whole tiles and must hold (and keep) zeros
        if (w.zero && hipMemset(*w.p, 0, m * w.bytes) != hipSuccess) {
            err = "hipMemset failed";
            return -1;
        }
    }
    bin_cap = nb;
    bin_img_cap = image;
    return 0;
}

static GridSet bin_set(PoissonGrid &g, int nb) {
    GridSet b;
    b.nb = nb;
    b.rho = g.bb_rho;
    b.rhoB = g.ncell;
    b.A = g.bb_A;
    b.AB = (size_t)g.kp * g.n[1] * g.n[2];
    b.B = g.bb_B;
    b.Bi = g.bb_Bi;
    b.BB = (size_t)g.kp * g.L[1] * g.n[2];
    b.phi = g.bb_phi;
    b.phii = g.bb_phii;
    b.phiB = g.ncell;
    b.T = (double *)g.bb_T;
    b.TB = (size_t)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
    b.Goct = (double *)g.bb_G;
    b.GB = g.noct;
    b.Ghat = g.bb_Gh;
    b.HA = g.bb_HA;
    b.HB = g.bb_HB;
    b.GhB = (size_t)g.gp * (g.n[1] + 1) * (g.n[2] + 1);
    b.Gimg = g.bb_Gi;
    b.GiB = (size_t)g.L[2] * (g.n[1] + 1) * g.gp;
    return b;
}

void PoissonGrid::zero_bin_rho(int nb, hipStream_t s) { launch_fill_i64(s, (size_t)nb * rho_words(), (long long *)bb_rho); }

void PoissonGrid::green_bins(const Scalars *scb, int nb, bool image, hipStream_t s) {
    if (!s) s = stream;
    const GridSet gs = bin_set(*this, nb);
    if (fp64) {
        green_impl<double>(*this, gs, scb, s);
        if (image) green_image_impl<double>(*this, gs, scb, s);
    } else {
        green_impl<float>(*this, gs, scb, s);
        if (image) green_image_impl<float>(*this, gs, scb, s);
    }
}

// Every bin's solve (and its image's) as one batch: the forward x and y
// passes once for both (they transform the same charge), the image z pass into
// its own spectrum before the bin's in-place z pass, then the inverse passes.
template <typename T>
static void solve_bins_impl(PoissonGrid &g, const GridSet &gs, const Scalars *scb, bool image, hipEvent_t green_ready) {
    pass_x_fwd<T>(g, gs, scb, 0);
    pass_y<T>(g, gs, false, gs.A, gs.B);
    if (green_ready) (void)hipStreamWaitEvent(g.stream, green_ready, 0);
    if (image) pass_z<T>(g, gs, 3, gs.B, gs.Bi);
    pass_z<T>(g, gs, 2, gs.B, gs.B);
    pass_y<T>(g, gs, true, gs.B, gs.A);
    pass_x_inv<T>(g, gs, scb, 1, gs.phi);
    if (image) {
        pass_y<T>(g, gs, true, gs.Bi, gs.A);
        pass_x_inv<T>(g, gs, scb, 1, gs.phii);
    }
}

void PoissonGrid::solve_bins(const Scalars *scb, int nb, bool image, hipEvent_t green_ready) {
    const GridSet gs = bin_set(*this, nb);
    if (fp64) solve_bins_impl<double>(*this, gs, scb, image, green_ready);
    else solve_bins_impl<float>(*this, gs, scb, image, green_ready);
}

// grid of the Grad kernels: cell tiles over (i, j), planes k, then the batch
static dim3 grad_grid(const int n[3], const CellTile &t, int nb) {
    return dim3((unsigned)(t.nbi * ((n[1] + (256 >> t.lw) - 1) / (256 >> t.lw))), (unsigned)n[2], (unsigned)nb);
}

void PoissonGrid::field_sums(const Scalars *scb, int nb, bool image) {
    const CellTile t = cell_tile(n[0]);
    const dim3 grid = grad_grid(n, t, 1);
    if (fp64)
        k_field_sums<double><<<grid, 256, 0, stream>>>((const double *)bb_phi, image ? (const double *)bb_phii : nullptr,
                                                       ncell, nb, (double4 *)Esum, (double4 *)Bsum, n[0], n[1], n[2],
                                                       scb, t);
    else
        k_field_sums<float><<<grid, 256, 0, stream>>>((const float *)bb_phi, image ? (const float *)bb_phii : nullptr,
                                                      ncell, nb, (float4 *)Esum, (float4 *)Bsum, n[0], n[1], n[2],
                                                      scb, t);
    profile_mark(PROF_GRAD);
}

int PoissonGrid::ensure_bin_fields(std::string &err) {
    if (bb_E) return 0;
    const size_t bytes = (size_t)bin_cap * ncell * 4 * (fp64 ? 8 : 4);
    if (bytes == 0 || hipMalloc(&bb_E, bytes) != hipSuccess) {
        bb_E = nullptr;
        err = "hipMalloc of the batched field grids failed";
        return -1;
    }
    return 0;
}

void PoissonGrid::grad_bins(const Scalars *scb, int nb) {
    const CellTile t = cell_tile(n[0]);
    const dim3 grid = grad_grid(n, t, nb);
    if (fp64) k_grad<double><<<grid, 256, 0, stream>>>((const double *)bb_phi, (double4 *)bb_E, n[0], n[1], n[2], scb, t);
    else k_grad<float><<<grid, 256, 0, stream>>>((const float *)bb_phi, (float4 *)bb_E, n[0], n[1], n[2], scb, t);
    profile_mark(PROF_GRAD);
}

void PoissonGrid::grad(const Scalars *d_sc) {
    const CellTile t = cell_tile(n[0]);
    const dim3 grid = grad_grid(n, t, 1);
    if (fp64) k_grad<double><<<grid, 256, 0, stream>>>((const double *)phi, (double4 *)Eg, n[0], n[1], n[2], d_sc, t);
    else k_grad<float><<<grid, 256, 0, stream>>>((const float *)phi, (float4 *)Eg, n[0], n[1], n[2], d_sc, t);
    profile_mark(PROF_GRAD);
}

int PoissonGrid::ensure_sums(std::string &err) {
    if (Esum) return 0;
    const size_t bytes = ncell * 4 * (fp64 ? 8 : 4);
    if (hipMalloc(&Esum, bytes) != hipSuccess || hipMalloc(&Bsum, bytes) != hipSuccess) {
        err = "hipMalloc of the field sums failed";
        return -1;
    }
    return 0;
}

void PoissonGrid::clear_sums() {
    const size_t words = ncell * (fp64 ? 4 : 2);  // one 4-vector per cell, in 8-byte words
    launch_fill_i64(stream, words, (long long *)Esum);
    launch_fill_i64(stream, words, (long long *)Bsum);
}

void PoissonGrid::zero_rho(hipStream_t s) { launch_fill_i64(s, rho_words(), (long long *)rho_i64); }

void PoissonGrid::rho_to_double(const Scalars *d_sc, double *out) {
    k_rho_to_double<<<1024, 256, 0, stream>>>((const long long *)rho_i64, out, ncell, d_sc, fp64 ? 1 : 0);
}

void PoissonGrid::widen(int which, double *out, const void *efield) {
    if (which == 1) {
        if (fp64) (void)hipMemcpyAsync(out, phi, ncell * 8, hipMemcpyDeviceToDevice, stream);
        else k_widen<float><<<1024, 256, 0, stream>>>((const float *)phi, out, ncell, 1, 1);
    } else {
        const void *E = efield ? efield : Eg;
        if (fp64) k_widen<double><<<1024, 256, 0, stream>>>((const double *)E, out, ncell, 3, 4);
        else k_widen<float><<<1024, 256, 0, stream>>>((const float *)E, out, ncell, 3, 4);
    }
}

}  // namespace opalsc

// ---- autotuning entry points (include/opalsc.h) ----
extern "C" int opalsc_tune_fft_candidates(int L, int precision, int *E, int cap) {
    if (cap < 0 || (cap > 0 && !E)) return OPALSC_EINVAL;
    const int k = opalsc::tile_candidates(L, precision == OPALSC_FP64, E, cap);
    return k < 0 ? OPALSC_EINVAL : k;
}

extern "C" int opalsc_tune_fft_set(int L, int precision, int E) {
    using opalsc::g_tile_override;
    const int fp64 = precision == OPALSC_FP64 ? 1 : 0;
    bool ok = E == 0;
    int cand[8];
    const int nc = opalsc::tile_candidates(L, fp64 != 0, cand, 8);
    for (int c = 0; c < nc; ++c) ok |= cand[c] == E;
    if (!ok) return OPALSC_EINVAL;  // the previous setting stays
    for (size_t i = 0; i < g_tile_override.size(); ++i)
        if (g_tile_override[i].L == L && g_tile_override[i].fp64 == fp64) {
            g_tile_override.erase(g_tile_override.begin() + (long)i);
            break;
        }
    if (E != 0) g_tile_override.push_back({L, fp64, E});
    return OPALSC_OK;
}

extern "C" int opalsc_tune_fft_get(int L, int precision) {
    opalsc::FftPlan pl;
    if (!opalsc::make_plan(L, precision == OPALSC_FP64, false, pl)) return OPALSC_EINVAL;
    return pl.E;
}


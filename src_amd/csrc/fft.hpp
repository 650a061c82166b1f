// LDS-resident batched 1D FFT for gfx950 (no vendor FFT library).
//
// A workgroup of kFftThreads threads transforms W lines of length L held in LDS
// with an interleaved layout: element (line, idx) lives at idx*P + line,
// P = W + 1.  Consecutive threads work on consecutive lines, so every LDS
// access of a stage touches consecutive addresses and the odd pitch keeps
// the transposing accesses of the x passes spread over the banks.  W is a
// power of two chosen so that one row of a tile is a whole 128-byte line of
// HBM in the strided passes.
//
// Each stage is a Stockham autosort radix-R step: thread-owned butterflies
// are read into registers, twiddled, transformed, and written back after a
// barrier.  E = W*L/kFftThreads elements per thread is a compile-time constant
// divisible by every radix used, so each thread owns exactly E/R butterflies
// per stage.  Radices 8, 4 and 2 come first and radix 3 (L = 3*2^a) or 5
// (L = 5*2^a) last, so the span Ns of every stage is a power of two and all
// index arithmetic is shifts and masks.  Twiddles come from LDS (FP32) or through the cache (FP64,
// whose tiles leave no room for them).
#pragma once
#include "common.hpp"

namespace opalsc {

constexpr int kFftThreads = 512;  // threads per FFT workgroup

// a / d for 0 <= a < 2^22, d > 0: float estimate, then the one correction it can need
__device__ inline int div_small(int a, int d, float rd) {
    int q = (int)((float)a * rd);
    if (q * d > a) --q;
    else if ((q + 1) * d <= a) ++q;
    return q;
}

struct FftPlan {
    int L;           // transform length
    int W;           // lines per workgroup (power of two)
    int lw;          // log2(W)
    int P;           // LDS pitch (W + 1)
    int E;           // elements per thread (template parameter of the kernels)
    int nst;         // number of stages
    int radix[16];   // radix of each stage
    const void *tw;  // exp(-2 pi i m / L), m in [0, L), complex T
};

template <typename C>
__device__ inline C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ inline C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
// Complex product with explicit fused multiply-adds (4 instructions instead of
// 6; the library is built without contraction, so the FFTs' twiddle products
// are the one place that asks for them)
__device__ inline float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename C>
__device__ inline C cmul(C a, C b) { return C{fmadd(a.x, b.x, -(a.y * b.y)), fmadd(a.x, b.y, a.y * b.x)}; }
// multiply by -i (forward) or +i (inverse)
template <typename C>
__device__ inline C cmul_mi(C a, bool inv) { return inv ? C{-a.y, a.x} : C{a.y, -a.x}; }

template <typename C>
__device__ inline void dft2(C *v, bool) {
    C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <typename C>
__device__ inline void dft3(C *v, bool inv) {
    using T = decltype(v[0].x);
    const T h = T(0.866025403784438646763723170752936183);  // sqrt(3)/2
    C t1 = cadd(v[1], v[2]);
    C t2 = C{v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    C d = csub(v[1], v[2]);
    C s = C{h * d.x, h * d.y};
    C r = cmul_mi(s, inv);  // forward: -i*s
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, r);
    v[2] = csub(t2, r);
}

// X_k = sum_n x_n w^(nk), w = exp(-+2 pi i / 5): with t1 = x1 + x4, t2 = x2 + x3,
// t3 = x1 - x4, t4 = x2 - x3, X1,4 = (x0 + c1 t1 + c2 t2) -+ i (s1 t3 + s2 t4) and
// X2,3 = (x0 + c2 t1 + c1 t2) -+ i (s2 t3 - s1 t4), c_m = cos(2 pi m / 5), s_m = sin(2 pi m / 5)
template <typename C>
__device__ inline void dft5(C *v, bool inv) {
    using T = decltype(v[0].x);
    const T c1 = T(0.309016994374947424102293417182819059), c2 = T(-0.809016994374947424102293417182819059);
    const T s1 = T(0.951056516295153572116439333379382143), s2 = T(0.587785252292473129168705954639072769);
    const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const C a1 = C{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    const C a2 = C{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    const C b1 = cmul_mi(C{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, inv);
    const C b2 = cmul_mi(C{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, inv);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <typename C>
__device__ inline void dft4(C *v, bool inv) {
    C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C a2 = cadd(v[1], v[3]), a3 = cmul_mi(csub(v[1], v[3]), inv);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <typename C>
__device__ inline void dft8(C *v, bool inv) {
    using T = decltype(v[0].x);
    const T r = T(0.707106781186547524400844362104849039);
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, inv);
    dft4(o, inv);
    // twiddles W8^k, forward W8 = (1-i)/sqrt2
    C o1 = inv ? C{r * (o[1].x - o[1].y), r * (o[1].x + o[1].y)}
               : C{r * (o[1].x + o[1].y), r * (o[1].y - o[1].x)};
    C o2 = cmul_mi(o[2], inv);
    C o3 = inv ? C{-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y)}
               : C{r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

template <typename C, int R>
__device__ inline void dftR(C *v, bool inv) {
    if constexpr (R == 2) dft2(v, inv);
    else if constexpr (R == 3) dft3(v, inv);
    else if constexpr (R == 4) dft4(v, inv);
    else if constexpr (R == 5) dft5(v, inv);
    else dft8(v, inv);
}

// One Stockham stage on LDS data; Ns = 2^ls = product of the radices already
// applied.  Every stage reads, for the thread's butterfly (line, j), the rows
// j + r L/R of its line; it writes rows idxD + r Ns.  stage_read loads,
// twiddles and transforms the thread's butterflies into registers;
// stage_write stores them.
// LDS indices and twiddle indices stay below 2^24: 24-bit multiplies (full
// rate; 32-bit ones are quarter rate), once per butterfly, and the R elements
// of a butterfly at uniform strides.
template <typename C, int E, int R>
__device__ inline void stage_read(const C *buf, const C *tw, int L, int lw, int P, int ls, bool inv,
                                  C (&v)[E / R][R]) {
    constexpr int NB = E / R;
    const int LR = L / R, Ns = 1 << ls, W = 1 << lw;
    const int twstep = L / (Ns * R);
    const int LRP = LR * P;
    // the lane's element index, opaque to loop-invariant code motion: hoisted
    // out of the stage loop, the E load and store addresses stay live across
    // every stage (and both transforms of the z pass) and spill at E = 16;
    // recomputed here they cost a few integer operations per stage
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int g = t + b * kFftThreads;
        const int line = g & (W - 1), j = g >> lw;
        const int k = j & (Ns - 1);
        const int rb = __umul24(j, P) + line;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = buf[rb + r * LRP];
        // the first stage's twiddles are all 1 (k = 0)
        if (ls != 0) {
            const int kt = __umul24(k, twstep);
            int ti = 0;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                ti += kt;  // r * k * twstep
                C w = tw[ti];
                if (inv) w.y = -w.y;
                v[b][r] = cmul(v[b][r], w);
            }
        }
        dftR<C, R>(v[b], inv);
    }
}
template <typename C, int E, int R>
__device__ inline void stage_write(C *buf, int L, int lw, int P, int ls, const C (&v)[E / R][R]) {
    constexpr int NB = E / R;
    const int Ns = 1 << ls, W = 1 << lw;
    const int NsP = Ns * P;
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int g = t + b * kFftThreads;
        const int line = g & (W - 1), j = g >> lw;
        const int k = j & (Ns - 1);
        const int idxD = (((j >> ls) * R) << ls) + k;
        const int wb = __umul24(idxD, P) + line;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[wb + r * NsP] = v[b][r];
    }
}
template <typename C, int E, int R>
__device__ inline void fft_stage(C *buf, const C *tw, int L, int lw, int P, int ls, bool inv) {
    C v[E / R][R];
    stage_read<C, E, R>(buf, tw, L, lw, P, ls, inv, v);
    __syncthreads();
    stage_write<C, E, R>(buf, L, lw, P, ls, v);
    __syncthreads();
}

// Stages [s0, s1) of the plan on LDS data, Ns = 2^ls before stage s0 (every
// earlier radix a power of two); returns ls after them (the odd radix, last,
// leaves it unchanged).
template <typename C, int E>
__device__ inline int fft_lds_range(C *buf, const C *tw, const FftPlan &pl, int s0, int s1, int ls, bool inv) {
    for (int s = s0; s < s1; ++s) {
        const int R = pl.radix[s];
        if constexpr (E % 8 == 0) {
            if (R == 8) { fft_stage<C, E, 8>(buf, tw, pl.L, pl.lw, pl.P, ls, inv); ls += 3; continue; }
        }
        if constexpr (E % 4 == 0) {
            if (R == 4) { fft_stage<C, E, 4>(buf, tw, pl.L, pl.lw, pl.P, ls, inv); ls += 2; continue; }
        }
        if constexpr (E % 2 == 0) {
            if (R == 2) { fft_stage<C, E, 2>(buf, tw, pl.L, pl.lw, pl.P, ls, inv); ls += 1; continue; }
        }
        if constexpr (E % 3 == 0) {
            if (R == 3) { fft_stage<C, E, 3>(buf, tw, pl.L, pl.lw, pl.P, ls, inv); continue; }
        }
        if constexpr (E % 5 == 0) {
            if (R == 5) { fft_stage<C, E, 5>(buf, tw, pl.L, pl.lw, pl.P, ls, inv); continue; }
        }
    }
    return ls;
}

// Transform the W lines in buf (caller has synchronised after filling it and tw).
template <typename C, int E>
__device__ inline void fft_lds(C *buf, const C *tw, const FftPlan &pl, bool inv) {
    (void)fft_lds_range<C, E>(buf, tw, pl, 0, pl.nst, 0, inv);
}

// Twiddle source of a kernel: FP32 tiles copy the table into LDS after the
// data (call before the barrier that follows the tile load); FP64 reads it
// through the cache.
template <typename C>
__device__ inline const C *fft_twiddles(C *buf, const FftPlan &pl) {
    if constexpr (sizeof(C) == 8) {
        C *t = buf + pl.L * pl.P;
        const C *src = (const C *)pl.tw;
        for (int m = threadIdx.x; m < pl.L; m += kFftThreads) t[m] = src[m];
        return t;
    } else {
        return (const C *)pl.tw;
    }
}

}  // namespace opalsc

// Deterministic block reductions shared by the particle kernels: per-thread
// values are combined across a wave with butterfly shuffles and across waves
// in a fixed order, sums as double-double, so results do not depend on timing.
#pragma once
#include "common.hpp"
#include "crmath.hpp"
#include "kernels.hpp"

namespace opalsc {

using crm::dd;
using crm::dd_add;

constexpr int kReduceThreads = 512;   // single-block final reductions of the partials

inline int stream_blocks(int64_t N) {
    int64_t b = (N + kThreads - 1) / kThreads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

// ---- deterministic block reductions ---------------------------------------
// One step of a recursive-halving wave reduction of 2H double-doubles: the
// lane keeps the half selected by its lane bit d and adds the other half of
// lane ^ d (the same operands as a butterfly step; dd_add is commutative bit
// for bit, since TwoSum's error term is exact).
template <int H>
__device__ inline void dd_halve(const double (&h)[2 * H], const double (&l)[2 * H], double (&ho)[H],
                                double (&lo)[H], int lane, int d) {
    const bool up = lane & d;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double kh = up ? h[H + i] : h[i], kl = up ? l[H + i] : l[i];
        const double sh = up ? h[i] : h[H + i], sl = up ? l[i] : l[H + i];
        const dd r = dd_add(dd{kh, kl}, dd{__shfl_xor(sh, d, 64), __shfl_xor(sl, d, 64)});
        ho[i] = r.hi;
        lo[i] = r.lo;
    }
}

// Block sum of K double-doubles per thread; thread 0 receives the result.
// Across a wave: butterflies (lane ^ 32, ^ 16, ..., ^ 1); across waves: wave
// 0's sum, then waves 1, 2, ... added in order.  K = 16 computes the same sums
// with recursive halving within the wave (17 exchanged values instead of 96)
// and one thread per value across the waves.
template <int K>
__device__ inline void block_sum_dd(double (&h)[K], double (&l)[K], double *lds) {
    if constexpr (K == 16) {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
        double h8[8], l8[8], h4[4], l4[4], h2[2], l2[2], h1[1], l1[1];
        dd_halve<8>(h, l, h8, l8, lane, 32);
        dd_halve<4>(h8, l8, h4, l4, lane, 16);
        dd_halve<2>(h4, l4, h2, l2, lane, 8);
        dd_halve<1>(h2, l2, h1, l1, lane, 4);
        dd t = {h1[0], l1[0]};
        for (int off = 2; off > 0; off >>= 1) t = dd_add(t, dd{__shfl_xor(t.hi, off, 64), __shfl_xor(t.lo, off, 64)});
        const int k = lane >> 2;  // the value t sums
        if ((lane & 3) == 0) {
            lds[(wave * K + k) * 2] = t.hi;
            lds[(wave * K + k) * 2 + 1] = t.lo;
        }
        __syncthreads();
        if (threadIdx.x < K) {
            const int q = threadIdx.x;
            dd a = {lds[2 * q], lds[2 * q + 1]};
            for (int w = 1; w < nw; ++w) a = dd_add(a, dd{lds[(w * K + q) * 2], lds[(w * K + q) * 2 + 1]});
            lds[2 * q] = a.hi;  // only thread q reads or writes value q's slots here
            lds[2 * q + 1] = a.lo;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < K; ++q) {
                h[q] = lds[2 * q];
                l[q] = lds[2 * q + 1];
            }
        __syncthreads();
        return;
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double oh = __shfl_xor(h[k], off, 64), ol = __shfl_xor(l[k], off, 64);
            dd r = dd_add(dd{h[k], l[k]}, dd{oh, ol});
            h[k] = r.hi;
            l[k] = r.lo;
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (lane == 0)
        for (int k = 0; k < K; ++k) {
            lds[(wave * K + k) * 2] = h[k];
            lds[(wave * K + k) * 2 + 1] = l[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; ++k) {
            dd a = {lds[2 * k], lds[2 * k + 1]};
            for (int w = 1; w < nw; ++w) a = dd_add(a, dd{lds[(w * K + k) * 2], lds[(w * K + k) * 2 + 1]});
            h[k] = a.hi;
            l[k] = a.lo;
        }
    }
    __syncthreads();
}

__device__ inline void block_minmax(double (&mn)[3], double (&mx)[3], double *lds) {
    for (int d = 0; d < 3; ++d) {
        mn[d] = wave_min(mn[d]);
        mx[d] = wave_max(mx[d]);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (lane == 0)
        for (int d = 0; d < 3; ++d) {
            lds[wave * 6 + d] = mn[d];
            lds[wave * 6 + 3 + d] = mx[d];
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < nw; ++w)
            for (int d = 0; d < 3; ++d) {
                mn[d] = fmin(mn[d], lds[w * 6 + d]);
                mx[d] = fmax(mx[d], lds[w * 6 + 3 + d]);
            }
    __syncthreads();
}

// block_minmax for the gather passes' block boxes: the maxima negated into
// minima (exact), the six minima reduced in a wave by recursive halving (10
// exchanged values instead of 36; a minimum does not depend on the order),
// thread 0 gets the block's result.
__device__ inline void block_minmax_split(double (&mn)[3], double (&mx)[3], double *lds) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const double v[8] = {mn[0], mn[1], mn[2], -mx[0], -mx[1], -mx[2], __builtin_inf(), __builtin_inf()};
    double a[4], b[2];
    {
        const bool up = lane & 32;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double keep = up ? v[4 + i] : v[i], send = up ? v[i] : v[4 + i];
            a[i] = fmin(keep, __shfl_xor(send, 32, 64));
        }
    }
    {
        const bool up = lane & 16;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double keep = up ? a[2 + i] : a[i], send = up ? a[i] : a[2 + i];
            b[i] = fmin(keep, __shfl_xor(send, 16, 64));
        }
    }
    const bool up = lane & 8;
    double t = fmin(up ? b[1] : b[0], __shfl_xor(up ? b[0] : b[1], 8, 64));
    t = fmin(t, __shfl_xor(t, 4, 64));
    t = fmin(t, __shfl_xor(t, 2, 64));
    t = fmin(t, __shfl_xor(t, 1, 64));
    const int idx = lane >> 3;  // the value this lane's t reduces
    if ((lane & 7) == 0 && idx < 6) lds[wave * 6 + idx] = idx < 3 ? t : -t;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            mn[d] = lds[d];
            mx[d] = lds[3 + d];
            for (int w = 1; w < nw; ++w) {
                mn[d] = fmin(mn[d], lds[w * 6 + d]);
                mx[d] = fmax(mx[d], lds[w * 6 + 3 + d]);
            }
        }
    __syncthreads();
}

// Per-thread chunk statistics of positions: min[3], max[3], sum[3], sum of squares[3].
__device__ inline void box_stats_init(double (&st)[kBoxStats]) {
    for (int d = 0; d < 3; ++d) {
        st[d] = __builtin_inf();
        st[3 + d] = -__builtin_inf();
        st[6 + d] = st[9 + d] = 0.0;
    }
}
__device__ inline void box_stats_add(double (&st)[kBoxStats], const double r[3]) {
    for (int d = 0; d < 3; ++d) {
        st[d] = fmin(st[d], r[d]);
        st[3 + d] = fmax(st[3 + d], r[d]);
        st[6 + d] += r[d];
        st[9 + d] += r[d] * r[d];
    }
}

// Block min/max and sums of positions and squared positions (the chunk
// statistics the deposit sizes its LDS box from); result valid in thread 0.
__device__ inline void block_chunk_stats(double (&st)[kBoxStats], double *lds) {
    for (int d = 0; d < 3; ++d) {
        st[d] = wave_min(st[d]);
        st[3 + d] = wave_max(st[3 + d]);
    }
    for (int k = 6; k < kBoxStats; ++k) st[k] = wave_sum(st[k]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (lane == 0)
        for (int k = 0; k < kBoxStats; ++k) lds[wave * kBoxStats + k] = st[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < nw; ++w) {
            for (int d = 0; d < 3; ++d) {
                st[d] = fmin(st[d], lds[w * kBoxStats + d]);
                st[3 + d] = fmax(st[3 + d], lds[w * kBoxStats + 3 + d]);
            }
            for (int k = 6; k < kBoxStats; ++k) st[k] += lds[w * kBoxStats + k];
        }
    __syncthreads();
}

// Sum K dd values over nb partial blocks (slots 2k, 2k+1) with one block of
// 256 threads; result valid in thread 0.
template <int K>
__device__ inline void reduce_partials_dd(const double *partials, int nb, double (&out)[K], double *lds) {
    double h[K], l[K];
    for (int k = 0; k < K; ++k) h[k] = l[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < K; ++k) {
            dd r = dd_add(dd{h[k], l[k]}, dd{partials[b * kSlots + 2 * k], partials[b * kSlots + 2 * k + 1]});
            h[k] = r.hi;
            l[k] = r.lo;
        }
    block_sum_dd<K>(h, l, lds);
    for (int k = 0; k < K; ++k) out[k] = h[k] + l[k];
}

__device__ inline void reduce_partials_minmax(const double *partials, int nb, double (&mn)[3],
                                              double (&mx)[3], double *lds) {
    for (int d = 0; d < 3; ++d) {
        mn[d] = __builtin_inf();
        mx[d] = -__builtin_inf();
    }
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], partials[b * kSlots + d]);
            mx[d] = fmax(mx[d], partials[b * kSlots + 3 + d]);
        }
    block_minmax_split(mn, mx, lds);
}

}  // namespace opalsc

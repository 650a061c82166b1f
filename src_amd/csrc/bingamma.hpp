// Per-chunk sums of the Lorentz factor for the energy bins' mean gammas
// (PartBunchBase::calcGammas, PartBunchBase.hpp:639-682: gamma_b = mean of
// sqrt(1 + P.P) over the particles of bin b).
//
// Particles are kept bin-major, so bin b is the index range [start[b],
// start[b+1]).  A chunk of C consecutive particles (the one-shot gather's
// block: C = kOsChunk mixed, kOsChunk64 FP64) holds sum g0 over its particles
// below `split`, the end of the bin its first particle lies in, and g1 over
// the rest.  When every non-empty bin holds at least C particles, the rest all
// lie in the next non-empty bin (or past the last bin), so the chunks that meet
// bin b give its sum: g0 of those starting inside b, g1 of the one starting
// before it.  Both the gather pass (which produces the next step's momenta)
// and the standalone pass of a call's first step compute these sums with the
// same thread mapping (particle chunk*C + u*kThreads + t, u ascending) and the
// same reduction (wave butterflies, then the block's waves in order), so a
// multi-step call and single-step calls get the same bits.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

namespace opalsc {

// Every non-empty bin holds at least C particles (then a chunk meets at most
// two of them).
inline bool bins_chunked(const BinRanges &r, int64_t C) {
    if (r.nbins <= 0) return false;
    for (int k = 0; k < r.nbins; ++k) {
        const int64_t n = r.start[k + 1] - r.start[k];
        if (n > 0 && n < C) return false;
    }
    return true;
}

// The end of the bin that holds particle `first`; N past the last bin.
__device__ inline int64_t bin_chunk_split(const BinRanges &r, int64_t first, int64_t N) {
    if (r.nbins <= 0 || first >= r.start[r.nbins]) return N;
    int lo = 0, hi = r.nbins - 1;  // smallest k with start[k+1] > first
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (r.start[mid + 1] > first) hi = mid;
        else lo = mid + 1;
    }
    return r.start[lo + 1];
}

__device__ inline double lorentz_gamma(double px, double py, double pz) {
    return sqrt(1.0 + (px * px + py * py + pz * pz));
}

// The block's two sums into binp[2*chunk], binp[2*chunk+1] (every thread of
// the block calls this; lds holds 2 doubles per wave).
__device__ inline void bin_chunk_store(double g0, double g1, double *lds, double *binp, int64_t chunk,
                                       bool write) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    if (lane == 0) {
        lds[2 * wave] = g0;
        lds[2 * wave + 1] = g1;
    }
    __syncthreads();
    if (write && threadIdx.x < 2) {
        double t = lds[threadIdx.x];
        for (int w = 1; w < kThreads / 64; ++w) t += lds[2 * w + threadIdx.x];
        binp[2 * chunk + threadIdx.x] = t;
    }
}

}  // namespace opalsc

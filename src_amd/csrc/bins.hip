// Energy-binned self fields, gfx950: ParallelTTracker::computeSpaceChargeFields
// with energy bins (src/Algorithms/ParallelTTracker.cpp:567-579) over
// PartBunch::computeSelfFields(int) (src/Classic/Algorithms/PartBunch.cpp:138-430).
//
// Bin b gives its particles the macro charge qi (setBinCharge,
// PartBunchBase.hpp:716-724) and is solved with its own Lorentz factor
// (calcGammas, PartBunchBase.hpp:639-682), once for the bunch and once for its
// image charge.  The particles are ordered by bin, so each bin's deposit is
// one contiguous range.  The fields of all bins are summed on the grid and
// gathered once (CIC gathering is linear).
#include "bingamma.hpp"
#include "common.hpp"
#include "gather.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace opalsc {

// keys for ordering by bin: bins outside [0, nbins) sort last and are ignored
__global__ void k_bin_keys(int64_t N, const int32_t *bin, int nbins, uint32_t *keys, uint32_t *vals) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = bin[i];
        keys[i] = (b >= 0 && b < nbins) ? (uint32_t)b : (uint32_t)nbins;
        vals[i] = (uint32_t)i;
    }
}
void launch_bin_keys(hipStream_t s, int64_t N, const int32_t *bin, int nbins, uint32_t *keys, uint32_t *vals) {
    if (N > 0) k_bin_keys<<<stream_blocks(N), kThreads, 0, s>>>(N, bin, nbins, keys, vals);
}

// particles per bin (LDS histogram per block)
__global__ __launch_bounds__(256) void k_bin_count(int64_t N, const int32_t *bin, int nbins,
                                                   unsigned long long *counts) {
    __shared__ unsigned int h[kMaxBins];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = bin[i];
        if (b >= 0 && b < nbins) atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (h[b]) atomicAdd(&counts[b], (unsigned long long)h[b]);
}
void launch_bin_count(hipStream_t s, int64_t N, const int32_t *bin, int nbins, unsigned long long *counts) {
    if (N > 0) k_bin_count<<<std::min(stream_blocks(N), 4096), kThreads, 0, s>>>(N, bin, nbins, counts);
}

// the bin-ordered copies the deposit reads: positions and qd = qi*dt
__global__ void k_bin_order(int64_t N, const uint32_t *perm, const double *x, const double *y,
                            const double *z, const double *dt, double qi, double *ox, double *oy,
                            double *oz, double *oqd) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t j = perm[i];
        ox[i] = x[j];
        oy[i] = y[j];
        oz[i] = z[j];
        double qd = qi;  // Q = qi (setBinCharge), Q *= dt (PartBunch.cpp:154)
        qd *= dt[j];
        oqd[i] = qd;
    }
}
void launch_bin_order(hipStream_t s, int64_t N, const uint32_t *perm, const double *x, const double *y,
                      const double *z, const double *dt, double qi, double *ox, double *oy, double *oz,
                      double *oqd) {
    if (N > 0) k_bin_order<<<stream_blocks(N), kThreads, 0, s>>>(N, perm, x, y, z, dt, qi, ox, oy, oz, oqd);
}

// calcGammas: per bin, the sum of sqrt(1 + P.P) over its particles (double-double
// across threads and blocks), then the mean; an empty bin gets 1.
__global__ __launch_bounds__(256) void k_bin_gamma_partials(int64_t N, const uint32_t *perm, const double *P,
                                                            double *partials) {
    __shared__ double lds[4 * 2];
    double h[1] = {0.0}, l[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double *p = P + 3 * (size_t)perm[i];
        const dd r = dd_add(dd{h[0], l[0]}, dd{sqrt(1.0 + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2])), 0.0});
        h[0] = r.hi;
        l[0] = r.lo;
    }
    block_sum_dd<1>(h, l, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * kSlots] = h[0];
        partials[blockIdx.x * kSlots + 1] = l[0];
    }
}
__global__ __launch_bounds__(kReduceThreads) void k_bin_gamma_final(const double *partials, int nb, int64_t count,
                                                                    int b, double *gammas) {
    __shared__ double lds[16 * 2];
    double out[1];
    reduce_partials_dd<1>(partials, nb, out, lds);
    if (threadIdx.x == 0) gammas[b] = count != 0 ? out[0] / (double)count : 1.0;
}
void launch_bin_gamma(hipStream_t s, int64_t count, const uint32_t *perm, const double *P, double *partials,
                      int b, double *gammas) {
    const int nb = count > 0 ? red_blocks(count) : 0;
    if (nb > 0) k_bin_gamma_partials<<<nb, kThreads, 0, s>>>(count, perm, P, partials);
    k_bin_gamma_final<<<1, kReduceThreads, 0, s>>>(partials, nb, count, b, gammas);
}

// The scalars of one bin (PartBunch.cpp:172-191,215,236,261-264): Lorentz
// factor, density scale 1/h0*1/h1*1/h2/gamma, solver spacing (h0, h1, h2*gamma),
// field scale (gamma, gamma, 1/gamma), betaC and the image shift; fixed-point
// scale for count deposits of at most qdmax.  Mesh, n, dT and coupling are
// already set in sc.
__device__ inline void bin_scalars(Scalars *sc, double gammaz, int64_t count, double qdmax, bool exact) {
    const double scaleFactor = 1.0;
    const double *hr = sc->hr;
    sc->gammaz = gammaz;
    sc->tmp2 = 1 / hr[0] * 1 / hr[1] * 1 / hr[2] / (scaleFactor * scaleFactor * scaleFactor) / gammaz;
    double hs[3] = {hr[0] * scaleFactor, hr[1] * scaleFactor, hr[2] * scaleFactor};
    hs[2] *= gammaz;
    for (int d = 0; d < 3; ++d) sc->hs[d] = hs[d];
    sc->vol = hs[0] * hs[1] * hs[2];
    sc->gscale[0] = gammaz / scaleFactor;
    sc->gscale[1] = gammaz / scaleFactor;
    sc->gscale[2] = 1.0 / (scaleFactor * gammaz);
    sc->betaC = sqrt(gammaz * gammaz - 1.0) / gammaz / kC;
    const double hz = hr[2];
    sc->zshift = -(2 * sc->origin[2] + (0 + (sc->n[2] - 1) + 1) * hz) * gammaz * scaleFactor;
    // fixed point: count deposits of |qd| <= qdmax
    fixed_point_scale(sc, count, qdmax, exact);
}

__global__ void k_bin_scalars(Scalars *sc, const double *gammas, int b, int64_t count,
                              const unsigned long long *qdmax_bits, int exact) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bin_scalars(sc, gammas[b], count, __longlong_as_double((long long)*qdmax_bits), exact != 0);
}
void launch_bin_scalars(hipStream_t s, Scalars *sc, const double *gammas, int b, int64_t count,
                        const unsigned long long *qdmax_bits, bool exact) {
    k_bin_scalars<<<1, 64, 0, s>>>(sc, gammas, b, count, qdmax_bits, exact ? 1 : 0);
}

// ---- tracker with energy bins: the bunch is ordered by bin, bin b holding
// particles start[b] <= i < start[b+1]

// calcGammas over the bin ranges: block (x, b) sums sqrt(1 + P.P) over a
// grid-stride share of bin b (double-double); gridDim.x is sized for the most
// populated bin.  Then one block per bin reduces its gridDim.x partials to the
// mean (1 for an empty bin).
__global__ __launch_bounds__(256) void k_track_bin_gamma_partials(const double *px, const double *py,
                                                                  const double *pz, BinRanges r,
                                                                  double *partials) {
    __shared__ double lds[4 * 2];
    const int b = blockIdx.y;
    double h[1] = {0.0}, l[1] = {0.0};
    for (int64_t i = r.start[b] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < r.start[b + 1];
         i += (int64_t)gridDim.x * blockDim.x) {
        const double p0 = px[i], p1 = py[i], p2 = pz[i];
        const dd v = dd_add(dd{h[0], l[0]}, dd{sqrt(1.0 + (p0 * p0 + p1 * p1 + p2 * p2)), 0.0});
        h[0] = v.hi;
        l[0] = v.lo;
    }
    block_sum_dd<1>(h, l, lds);
    if (threadIdx.x == 0) {
        double *o = partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
        o[0] = h[0];
        o[1] = l[0];
    }
}
__global__ __launch_bounds__(kReduceThreads) void k_track_bin_gamma_final(const double *partials, int nx,
                                                                          BinRanges r, double *gammas) {
    __shared__ double lds[16 * 2];
    const int b = blockIdx.x;
    const double *pb = partials + (size_t)b * nx * 2;
    double h[1] = {0.0}, l[1] = {0.0};
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        const dd v = dd_add(dd{h[0], l[0]}, dd{pb[2 * x], pb[2 * x + 1]});
        h[0] = v.hi;
        l[0] = v.lo;
    }
    block_sum_dd<1>(h, l, lds);
    const int64_t count = r.start[b + 1] - r.start[b];
    if (threadIdx.x == 0) gammas[b] = count != 0 ? (h[0] + l[0]) / (double)count : 1.0;
}

// Per-bin scalars of the step: copies of the step's scalars (mesh, frames)
// with bin b's Lorentz-frame values and charge qi.
// Also writes the scalars of the non-empty bins, in bin order, to scbc (the
// batched solve's bins: an empty bin has no charge and no field).
__global__ void k_track_bin_scalars(const Scalars *sc, const double *gammas, BinRanges r, double qi,
                                    int exact, Scalars *scb, Scalars *scbc) {
    const int b = threadIdx.x;
    if (b >= r.nbins) return;
    Scalars s = *sc;
    s.qi = qi;
    const int64_t count = r.start[b + 1] - r.start[b];
    // every deposit is |qi*dT| <= |qi|*dT*(1+2^-52)
    bin_scalars(&s, gammas[b], count, fabs(qi) * s.dT * (1.0 + 0x1p-50), exact != 0);
    scb[b] = s;
    if (count > 0) {
        int z = 0;
        for (int k = 0; k < b; ++k) z += r.start[k + 1] > r.start[k] ? 1 : 0;
        scbc[z] = s;
    }
}

// calcGammas from the chunk sums (the layout bingamma.hpp describes): block b
// adds, in chunk order, g0 of the chunks that start inside bin b and g1 of the
// chunk that starts before it (double-double), then divides by the count.
__global__ __launch_bounds__(kReduceThreads) void k_bin_gamma_from_chunks(const double *binp, int64_t C,
                                                                         BinRanges r, double *gammas) {
    __shared__ double lds[16 * 2];
    const int b = blockIdx.x;
    const int64_t s0 = r.start[b], s1 = r.start[b + 1];
    double h[1] = {0.0}, l[1] = {0.0};
    if (s1 > s0) {
        const int64_t c0 = s0 / C, c1 = (s1 - 1) / C;
        for (int64_t c = c0 + threadIdx.x; c <= c1; c += blockDim.x) {
            const double g = c * C >= s0 ? binp[2 * c] : binp[2 * c + 1];
            const dd v = dd_add(dd{h[0], l[0]}, dd{g, 0.0});
            h[0] = v.hi;
            l[0] = v.lo;
        }
    }
    block_sum_dd<1>(h, l, lds);
    if (threadIdx.x == 0) gammas[b] = s1 > s0 ? (h[0] + l[0]) / (double)(s1 - s0) : 1.0;
}

// The chunk sums of the resident momenta, for a step whose previous gather
// pass did not leave them: U particles per thread, as in the one-shot gather
// pass of the same chunk size C = U*kThreads.
template <int U>
__global__ __launch_bounds__(256) void k_bin_gamma_chunks(int64_t N, const double *__restrict__ px,
                                                          const double *__restrict__ py,
                                                          const double *__restrict__ pz, BinRanges r,
                                                          double *binp) {
    __shared__ double lds[2 * (kThreads / 64)];
    constexpr int64_t C = (int64_t)U * kThreads;
    const int64_t chunk = blockIdx.x, first = chunk * C, base = first + threadIdx.x;
    const int64_t split = bin_chunk_split(r, first, N);
    double g0 = 0.0, g1 = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = base + u * kThreads;
        if (i >= N) break;
        const double g = lorentz_gamma(px[i], py[i], pz[i]);
        if (i < split) g0 += g;
        else g1 += g;
    }
    bin_chunk_store(g0, g1, lds, binp, chunk, true);
}

void launch_bin_gamma_chunks(hipStream_t s, bool fp64, const Bunch &bunch, const BinRanges &r, double *binp) {
    if (bunch.N <= 0 || r.nbins <= 0) return;
    const int64_t C = fp64 ? kOsChunk64 : kOsChunk;
    const unsigned nch = (unsigned)((bunch.N + C - 1) / C);
    if (fp64) k_bin_gamma_chunks<kOsU64><<<nch, kThreads, 0, s>>>(bunch.N, bunch.px, bunch.py, bunch.pz, r, binp);
    else k_bin_gamma_chunks<kOsU><<<nch, kThreads, 0, s>>>(bunch.N, bunch.px, bunch.py, bunch.pz, r, binp);
}

void launch_track_bin_scalars(hipStream_t s, const Bunch &bunch, const BinRanges &r, double qi, bool exact,
                              const Scalars *sc, double *partials, const double *binp, int64_t chunk,
                              double *gammas, Scalars *scb, Scalars *scbc) {
    if (r.nbins <= 0) return;
    static_assert(kMaxBins <= 64, "one thread per bin");
    if (binp) {
        k_bin_gamma_from_chunks<<<r.nbins, kReduceThreads, 0, s>>>(binp, chunk, r, gammas);
    } else {
        // blocks per bin: about 8 particles per thread in the most populated bin
        int64_t most = 0;
        for (int k = 0; k < r.nbins; ++k) most = std::max(most, r.start[k + 1] - r.start[k]);
        const int nx = (int)std::min<int64_t>(kBinGammaBlocks, std::max<int64_t>(1, (most + 2047) / 2048));
        k_track_bin_gamma_partials<<<dim3(nx, r.nbins), kThreads, 0, s>>>(bunch.px, bunch.py, bunch.pz, r,
                                                                           partials);
        k_track_bin_gamma_final<<<r.nbins, kReduceThreads, 0, s>>>(partials, nx, r, gammas);
    }
    k_track_bin_scalars<<<1, 64, 0, s>>>(sc, gammas, r, qi, exact ? 1 : 0, scb, scbc);
}

// bin keys of the resident bunch for its first bin-major order: the bin of
// particle id[i] (bins outside [0, nbins) become nbins and sort last)
__global__ void k_bin_keys_by_id(int64_t N, const int32_t *bin, const uint32_t *id, int nbins, uint32_t *keys,
                                 uint32_t *vals) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = bin[id[i]];
        keys[i] = (b >= 0 && b < nbins) ? (uint32_t)b : (uint32_t)nbins;
        vals[i] = (uint32_t)i;
    }
}
void launch_bin_keys_by_id(hipStream_t s, int64_t N, const int32_t *bin, const uint32_t *id, int nbins,
                           uint32_t *keys, uint32_t *vals) {
    if (N > 0) k_bin_keys_by_id<<<stream_blocks(N), kThreads, 0, s>>>(N, bin, id, nbins, keys, vals);
}

// Ef = gather(sum over bins of E), Bf += gather(sum over bins of the implied B)
// (PartBunch.cpp:238-248,419-426 for every bin), AoS outputs.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_binned(int64_t N, const double *x, const double *y, const double *z,
                                                       const typename vec4_of<T>::type *Es,
                                                       const typename vec4_of<T>::type *Bs, const Scalars *sc,
                                                       double *Ef, double *Bf) {
    const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
    const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        int l[3];
        double d[3], e[3] = {0.0, 0.0, 0.0}, bb[3] = {0.0, 0.0, 0.0};
        bool ok = cic_dim(x[p], o0, h0, l[0], d[0]);
        ok &= cic_dim(y[p], o1, h1, l[1], d[1]);
        ok &= cic_dim(z[p], o2, h2, l[2], d[2]);
        if (ok) {
            gather_E<T>(Es, l, d, n0, n1, n2, e);
            gather_E<T>(Bs, l, d, n0, n1, n2, bb);
        }
        Ef[3 * p] = e[0];
        Ef[3 * p + 1] = e[1];
        Ef[3 * p + 2] = e[2];
        Bf[3 * p] = Bf[3 * p] + bb[0];
        Bf[3 * p + 1] = Bf[3 * p + 1] + bb[1];
    }
}
void launch_gather_binned(hipStream_t s, bool fp64, int64_t N, const double *x, const double *y, const double *z,
                          const void *Es, const void *Bs, const Scalars *sc, double *Ef, double *Bf) {
    if (N <= 0) return;
    if (fp64)
        k_gather_binned<double><<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, (const double4 *)Es,
                                                                      (const double4 *)Bs, sc, Ef, Bf);
    else
        k_gather_binned<float><<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, (const float4 *)Es,
                                                                     (const float4 *)Bs, sc, Ef, Bf);
}

}  // namespace opalsc

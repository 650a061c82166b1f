// Particle kernels of the space-charge step, gfx950.
//
// Device layout of the bunch is structure-of-arrays (x,y,z,px,py,pz,q,id):
// every particle kernel streams 8-byte words with unit stride across a wave.
// Reductions are deterministic: each thread accumulates its grid-stride
// particles in order, threads and blocks are combined in a fixed butterfly /
// block order, and sums are carried as double-double so that the moment sums
// are the correctly rounded sums of the per-particle terms.
#include "common.hpp"
#include "bingamma.hpp"
#include "crmath.hpp"
#include "frame.hpp"
#include "igfmath.hpp"
#include "kernels.hpp"
#include "gather.hpp"
#include "reduce.hpp"
#include "tuning.hpp"  // generated from src_amd/tuning.json by src_amd/build.py

namespace opalsc {


int red_blocks(int64_t N) {
    int64_t b = (N + kThreads - 1) / kThreads;
    if (b > kRedBlocksMax) b = kRedBlocksMax;
    return (int)b;
}

// ---- layout conversion ------------------------------------------------------
__global__ void k_soa_from_aos(int64_t N, const double *R, double *x, double *y, double *z) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = R[3 * i];
        y[i] = R[3 * i + 1];
        z[i] = R[3 * i + 2];
    }
}
__global__ void k_aos_from_soa(int64_t N, const double *x, const double *y, const double *z, double *R) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        R[3 * i] = x[i];
        R[3 * i + 1] = y[i];
        R[3 * i + 2] = z[i];
    }
}
void launch_soa_from_aos(hipStream_t s, int64_t N, const double *R, double *x, double *y, double *z) {
    if (N > 0) k_soa_from_aos<<<stream_blocks(N), kThreads, 0, s>>>(N, R, x, y, z);
}
void launch_aos_from_soa(hipStream_t s, int64_t N, const double *x, const double *y, const double *z, double *R) {
    if (N > 0) k_aos_from_soa<<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, R);
}

// ---- computeSelfFields entry point -----------------------------------------
// Q *= dt; scatter; Q /= dt (PartBunch.cpp:493-495) and sum(P)[2] (PartBunch.cpp:499).
__global__ __launch_bounds__(256) void k_selffield_prep(int64_t N, const double *P, double *Q,
                                                        const double *dt, double *qd,
                                                        double *partials,
                                                        unsigned long long *qdmax_bits) {
    __shared__ double lds[4 * 2 * 2];
    double h[1] = {0.0}, l[1] = {0.0};
    double amax = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = Q[i] * dt[i];
        qd[i] = v;
        Q[i] = v / dt[i];
        amax = fmax(amax, fabs(v));
        dd r = crm::two_sum(h[0], P[3 * i + 2]);
        h[0] = r.hi;
        l[0] += r.lo;
    }
    block_sum_dd<1>(h, l, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * kSlots] = h[0];
        partials[blockIdx.x * kSlots + 1] = l[0];
    }
    for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(qdmax_bits, (unsigned long long)__double_as_longlong(amax));
}

void launch_selffield_prep(hipStream_t s, int64_t N, const double *P, double *Q, const double *dt,
                           double *qd, double *partials, unsigned long long *qdmax_bits) {
    if (N <= 0) return;
    k_selffield_prep<<<red_blocks(N), kThreads, 0, s>>>(N, P, Q, dt, qd, partials, qdmax_bits);
}

// Fixed-point scale of a deposit of all N particles (fixed_point_scale).
__device__ inline void set_fixed_point(Scalars *sc, int64_t N, double qdmax, bool exact) {
    fixed_point_scale(sc, N, qdmax, exact);
}

// The computeSelfFields scalars (PartBunch.cpp:497-511,542-554,573-575,641).
__device__ inline void selffield_scalars(Scalars *sc, double psum_z, int64_t N, double dT,
                                         double coupling) {
    double gammaz = psum_z / (double)N;
    gammaz *= gammaz;
    gammaz = sqrt(gammaz + 1.0);
    sc->gammaz = gammaz;
    sc->dT = dT;
    sc->coupling = coupling;
    double hs[3] = {sc->hr[0] * 1.0, sc->hr[1] * 1.0, sc->hr[2] * 1.0};
    hs[2] *= gammaz;
    for (int d = 0; d < 3; ++d) sc->hs[d] = hs[d];
    sc->tmp2 = 1 / hs[0] * 1 / hs[1] * 1 / hs[2];
    sc->vol = hs[0] * hs[1] * hs[2];
    sc->gscale[0] = gammaz / 1.0;
    sc->gscale[1] = gammaz / 1.0;
    sc->gscale[2] = 1.0 / (1.0 * gammaz);
    sc->betaC = sqrt(gammaz * gammaz - 1.0) / gammaz / kC;
}

__global__ __launch_bounds__(512) void k_selffield_scalars(const double *partials, int nb, int64_t N,
                                                           double dT, double o0, double o1, double o2,
                                                           double h0, double h1, double h2, int n0,
                                                           int n1, int n2, double coupling,
                                                           const unsigned long long *qdmax_bits,
                                                           int exact, int pz, Scalars *sc) {
    __shared__ double lds[16 * 2 * 2];
    double out[1];
    reduce_partials_dd<1>(partials, nb, out, lds);
    if (threadIdx.x != 0) return;
    sc->n[0] = n0;
    sc->n[1] = n1;
    sc->n[2] = n2;
    sc->origin[0] = o0;
    sc->origin[1] = o1;
    sc->origin[2] = o2;
    sc->hr[0] = h0;
    sc->hr[1] = h1;
    sc->hr[2] = h2;
    sc->N = N;
    sc->mesh_ok = 1;
    sc->pz = pz;
    sc->psum[2] = out[0];
    selffield_scalars(sc, out[0], N, dT, coupling);
    set_fixed_point(sc, N, __longlong_as_double((long long)*qdmax_bits), exact != 0);
}

void launch_selffield_scalars(hipStream_t s, const double *partials, int nb, int64_t N, double dT,
                              const double origin[3], const double hr[3], const int n[3],
                              double coupling, const unsigned long long *qdmax_bits, bool exact,
                              bool pz, Scalars *sc) {
    k_selffield_scalars<<<1, kReduceThreads, 0, s>>>(partials, nb, N, dT, origin[0], origin[1], origin[2],
                                               hr[0], hr[1], hr[2], n[0], n[1], n[2], coupling,
                                               qdmax_bits, exact ? 1 : 0, pz ? 1 : 0, sc);
}

// ---- CIC scatter into int64 fixed point -------------------------------------
// IntCICImpl<3>::scatter (ippl/src/Particle/IntCIC.h:816-860): the 8 weights
// ((a*b)*c)*qd are rounded to the fixed-point grid and added with 64-bit
// integer atomics, so the charge grid does not depend on the order in which
// particles are processed.  MODE 0: qd given; MODE 1: qd = Q*dT with Q
// updated to (Q*dT)/dT as the reference leaves it (PartBunch.cpp:493-495).

// LDS-tiled deposit.  A block owns a contiguous chunk of kChunk particles;
// each thread loads its kPPT consecutive particles into registers at once and
// the block bounds the chunk's cells with a min/max of the positions (or takes
// the box the bounds pass recorded).  If the target box fits kBoxCap cells,
// contributions are accumulated in an LDS copy of the box (64-bit LDS
// atomics) and the non-zero cells are flushed to HBM with one global atomic
// each; otherwise they go straight to HBM.  Before either, a thread sums its
// consecutive particles of one cell in registers, so a bunch sorted by cell
// issues one atomic per run and corner instead of one per particle, and lanes
// of a wave rarely hit the same LDS address.  Integer addition makes the
// result independent of all of this.
constexpr int kPPT = 8;
constexpr int kChunk = kThreads * kPPT;
static_assert(kChunk == kDepositChunk, "the bounds pass and the deposit share chunks");

// a / b for 0 <= a < 2^24, 0 < b, from a float estimate corrected to the exact quotient
__device__ inline int small_div(int a, int b, float rb) {
    int qt = (int)((float)a * rb);
    if (qt * b > a) --qt;
    else if ((qt + 1) * b <= a) ++qt;
    return qt;
}
// LDS box capacities (int64 cells): 38 KiB lets four blocks share a CU's
// 160 KiB of LDS, so a large bunch's deposit runs at four waves per SIMD
// (0.98 -> 0.91 ms at 1e8 particles against 44 KiB at three; 31 KiB at five
// waves spills and slows); a bunch of a few million particles or fewer fills
// the device with fewer blocks than that, and there the larger box (fewer
// chunks falling back to global atomics) wins (64 small bunches: 6.2 -> 5.9 ms).
constexpr int kBoxCapSmall = 4864, kBoxCapLarge = 5632;
// Bunches of at most this many particles deposit through the larger LDS box
// (one block fewer per CU); set by the autotuner's table or opalsc_tune_deposit_set.
static int64_t g_large_box_max = OPALSC_TUNED_LARGE_BOX_MAX;

// Adds a cell sum v (deposit scale) to the charge grid: one word per cell on
// the mixed path; on the FP64 path the exact split of fixed_point_scale,
// hi = floor(v / 2^s) to word 2c and lo = v - hi * 2^s in [0, 2^s) to word 2c+1.
template <bool EXACT>
__device__ inline void rho_add(long long *rho, size_t c, long long v, int s) {
    if constexpr (EXACT) {
        const long long hi = v >> s;
        const long long lo = v - hi * ((long long)1 << s);
        if (hi) atomicAdd((unsigned long long *)&rho[2 * c], (unsigned long long)hi);
        if (lo) atomicAdd((unsigned long long *)&rho[2 * c + 1], (unsigned long long)lo);
    } else {
        atomicAdd((unsigned long long *)&rho[c], (unsigned long long)v);
    }
}

// MODE 0: qd given; 1 (tracker): qd = Q*dT with Q left as (Q*dT)/dT (FP64 path), positions
// given in the reference frame and rotated into the beam frame (M_rb, o_rb)
// exactly as the bounds pass rotated them; 2 (tracker, energy bin): as 1 with
// qd = qi*dT for every particle (setBinCharge) and Q untouched; 3 (tracker, a
// bunch whose particles all carry the same Q): as 1 with qd = Scalars::qd_u,
// the charge kept as one number (Scalars::qu).  EXACT: the
// FP64 path (reference-order CIC, 64-bit fixed point per deposit); otherwise
// the mixed path (CIC with 1/h, FP32 weights, |deposit| < 2^24 in 32-bit
// integers, summed in 64 bits).  BOX: chunk statistics supplied by the bounds
// pass.  Only particles lo <= p < N deposit; block b owns chunk chunk0 + b of
// the Nall-particle bunch.
template <int MODE, bool EXACT, bool BOX, int kBoxCap, bool BATCH = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_deposit_lds(int64_t lo, int64_t N, int64_t Nall, int64_t chunk0,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z, double *q,
                                                     const double *qdv, long long *rho,
                                                     const Scalars *sc, unsigned long long *dropped,
                                                     const double *chunkbox, BatchView bv = BatchView{}) {
    __shared__ long long box[kBoxCap];
    __shared__ double red[4 * kBoxStats];
    __shared__ int boxinfo[7];
    int bid = (int)blockIdx.x;
    if constexpr (BATCH) {
        int nblk = 0;
        const BatchItem &it = bv.items[batch_find(bv, bid, nblk)];
        lo = 0;
        N = Nall = it.N;
        chunk0 = 0;
        x = it.x;
        y = it.y;
        z = it.z;
        q = it.q;
        rho = it.rho;
        sc = it.sc;
        dropped = it.dropped;
        chunkbox = it.chunkbox;
    }
    const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
    const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
    const double ih0 = 1.0 / h0, ih1 = 1.0 / h1, ih2 = 1.0 / h2;
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    const double e2q = sc->e2q, dT = sc->dT;
    const int lsh = sc->lo_shift;
    const int64_t chunk = chunk0 + bid;
    const int64_t base = chunk * kChunk;

    // 1. the chunk's particles, all loads issued before any is used; thread t
    // owns the kPPT consecutive particles p0 = base + t*kPPT ... p0 + kPPT-1
    const int64_t p0 = base + (int64_t)threadIdx.x * kPPT;
    const double *qsrc = MODE == 0 ? qdv : q;
    double px[kPPT], py[kPPT], pz[kPPT], pq[kPPT];
    if (MODE >= 2) {
        double qb = sc->qd_u;
        if (MODE == 2) {
            qb = sc->qi;  // Q = qi (setBinCharge), Q *= dt (PartBunch.cpp:154)
            qb *= dT;
        }
#pragma unroll
        for (int k = 0; k < kPPT; ++k) pq[k] = qb;
    }
    if (p0 >= lo && p0 + kPPT <= N) {
#pragma unroll
        for (int m = 0; m < kPPT / 2; ++m) {
            const double2 tx = ((const double2 *)(x + p0))[m], ty = ((const double2 *)(y + p0))[m];
            const double2 tz = ((const double2 *)(z + p0))[m];
            px[2 * m] = tx.x;
            px[2 * m + 1] = tx.y;
            py[2 * m] = ty.x;
            py[2 * m + 1] = ty.y;
            pz[2 * m] = tz.x;
            pz[2 * m + 1] = tz.y;
            if constexpr (MODE < 2) {
                const double2 tq = ((const double2 *)(qsrc + p0))[m];
                pq[2 * m] = tq.x;
                pq[2 * m + 1] = tq.y;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kPPT; ++k) {
            const int64_t p = p0 + k;
            const bool in = p >= lo && p < N;
            px[k] = in ? x[p] : 0.0;
            py[k] = in ? y[p] : 0.0;
            pz[k] = in ? z[p] : 0.0;
            if constexpr (MODE < 2) pq[k] = in ? qsrc[p] : 0.0;
        }
    }
    if constexpr (MODE != 0) {
        // the tracker keeps positions in the reference frame; the mesh is in the beam frame
        double M[9], o[3];
        for (int k = 0; k < 9; ++k) M[k] = sc->M_rb[k];
        for (int d = 0; d < 3; ++d) o[d] = sc->o_rb[d];
        const bool dcz = sc->pz != 0;
        const double wlo = sc->wrap_lo, whi = sc->wrap_hi;
#pragma unroll
        for (int k = 0; k < kPPT; ++k) {
            double r[3] = {px[k], py[k], pz[k]};
            if constexpr (EXACT) transform_to(M, o, r);
            else transform_to_fast(M, o, r);
            px[k] = r[0];
            py[k] = r[1];
            pz[k] = dcz ? wrap_periodic(r[2], wlo, whi) : r[2];  // boundp's update() of a DC beam
        }
    }

    // 2. the chunk's LDS box: its bounding box if that fits kBoxCap cells, else
    // the largest box about the mean (mean +- a*rms per axis) that fits, with
    // the particles outside it going to HBM directly
    const double ob[3] = {o0, o1, o2}, hb[3] = {h0, h1, h2};
    const int nb[3] = {n0, n1, n2};
    // cells [lo, hi] whose corners particles in [a, b] can reach: (x-o)/h is
    // monotonic in x; lgp is in [floor(t)-1, floor(t)] for t >= 0 and
    // >= floor(t)-2 below the origin; corners reach lgp+1
    auto span = [&](int d, double a, double b, int &lo, int &hi) -> bool {
        const double tl = (a - ob[d]) / hb[d], th = (b - ob[d]) / hb[d];
        if (!(tl > -1e9 && th < 1e9)) return false;
        lo = (int)floor(tl) - (tl < 0.0 ? 2 : 1);
        hi = (int)floor(th) + 1;
        lo = lo < 0 ? 0 : lo;
        hi = hi > nb[d] - 1 ? nb[d] - 1 : hi;
        if (hi < lo) hi = lo;
        return true;
    };
    double st[kBoxStats];
    // own: the block forms its chunk's statistics itself (no statistics supplied,
    // or only a bounding box, NaN sums, that overflows the LDS box)
    bool own = !BOX;
    if constexpr (BOX) own = chunkbox == nullptr;  // a batched bunch re-sorted in this step
    if (!own) {
        for (int k = 0; k < kBoxStats; ++k) st[k] = chunkbox[chunk * kBoxStats + k];
        if (!(st[6] == st[6])) {
            int l0 = 0, h0_ = 0;
            long long v = 1;
            bool fin = true;
            for (int d = 0; d < 3 && fin; ++d) {
                fin = span(d, st[d], st[3 + d], l0, h0_);
                v *= (h0_ - l0 + 1);
            }
            own = !fin || v > kBoxCap;  // the same for every thread of the block
        }
    }
    if (own) {
        box_stats_init(st);
#pragma unroll
        for (int k = 0; k < kPPT; ++k) {
            if (p0 + k >= lo && p0 + k < N) {
                const double r[3] = {px[k], py[k], pz[k]};
                box_stats_add(st, r);
            }
        }
        block_chunk_stats(st, red);
    }
    if (threadIdx.x == 0) {
        // particles the statistics cover: the whole chunk (bounds pass) or its part in [lo, N)
        const int64_t c0 = !own ? base : (lo > base ? lo : base);
        const int64_t c1 = !own ? (Nall < base + kChunk ? Nall : base + kChunk) : (N < base + kChunk ? N : base + kChunk);
        const double cnt = (double)(c1 - c0);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        long long vol = 1;
        bool ok = true;
        for (int d = 0; d < 3 && ok; ++d) {
            ok = span(d, st[d], st[3 + d], lo[d], hi[d]);
            vol *= (hi[d] - lo[d] + 1);
        }
        if (ok && vol > kBoxCap) {
            static constexpr double kA[5] = {3.0, 2.0, 1.5, 1.0, 0.5};
            ok = false;
            for (int t = 0; t < 5 && !ok; ++t) {
                int l2[3], h2[3];
                long long v2 = 1;
                bool fin = true;
                for (int d = 0; d < 3; ++d) {
                    const double mu = st[6 + d] / cnt;
                    const double var = st[9 + d] / cnt - mu * mu;
                    const double w = kA[t] * sqrt(var > 0.0 ? var : 0.0);
                    fin &= span(d, fmax(mu - w, st[d]), fmin(mu + w, st[3 + d]), l2[d], h2[d]);
                    v2 *= (h2[d] - l2[d] + 1);
                }
                if (fin && v2 <= kBoxCap) {
                    ok = true;
                    vol = v2;
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = l2[d];
                        hi[d] = h2[d];
                    }
                }
            }
        }
        for (int d = 0; d < 3; ++d) {
            boxinfo[d] = lo[d];
            boxinfo[3 + d] = hi[d] - lo[d] + 1;
        }
        boxinfo[6] = (ok && vol <= kBoxCap) ? (int)vol : -1;
    }
    __syncthreads();
    const int vol = boxinfo[6];
    const int bx0 = boxinfo[0], by0 = boxinfo[1], bz0 = boxinfo[2];
    const int bx = boxinfo[3], by = boxinfo[4], bz = boxinfo[5];
    const bool use_lds = vol > 0;
    if (use_lds) {
        for (int e = threadIdx.x; e < vol; e += kThreads) box[e] = 0;
        __syncthreads();
    }

    // 3. CIC; a thread sums its consecutive particles of one cell in registers
    // and adds each run to the box (or the grid) when the cell changes
    using V = typename std::conditional<EXACT, long long, int>::type;
    const int sxy = bx * by;
    unsigned long long nd = 0;
    V acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cur_key = -100, cl[3] = {0, 0, 0};  // no particle key is -100
    bool cur_live = false;
    auto flush = [&]() {
        const int li = cl[0] - bx0, lj = cl[1] - by0, lk = cl[2] - bz0;
        // the whole stencil inside the LDS box (hence inside the grid): the common case
        if (use_lds && li >= 0 && lj >= 0 && lk >= 0 && li < bx - 1 && lj < by - 1 && lk < bz - 1) {
            long long *b = &box[(lk * by + lj) * bx + li];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int off = (c & 1) + ((c >> 1) & 1) * bx + ((c >> 2) & 1) * sxy;
                atomicAdd((unsigned long long *)(b + off), (unsigned long long)(long long)acc[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int i = cl[0] + (c & 1), j = cl[1] + ((c >> 1) & 1), kk = cl[2] + ((c >> 2) & 1);
                if (i >= 0 && j >= 0 && kk >= 0 && i < n0 && j < n1 && kk < n2) {
                    if (acc[c] == 0) continue;
                    const int ci = i - bx0, cj = j - by0, ck = kk - bz0;
                    const unsigned long long add = (unsigned long long)(long long)acc[c];
                    if (use_lds && ci >= 0 && cj >= 0 && ck >= 0 && ci < bx && cj < by && ck < bz)
                        atomicAdd((unsigned long long *)&box[(ck * by + cj) * bx + ci], add);
                    else
                        rho_add<EXACT>(rho, ((size_t)kk * n1 + j) * n0 + i, (long long)acc[c], lsh);
                }
            }
        }
    };
#pragma unroll
    for (int k = 0; k < kPPT; ++k) {
        const int64_t p = p0 + k;
        V v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int key = -1 - k;  // unique within the thread: never merges
        int l[3] = {0, 0, 0};
        bool valid = false;
        if (p >= lo && p < N) {
            double qd;
            if constexpr (MODE != 1) {
                qd = pq[k];
            } else {
                const double qv = pq[k];
                qd = qv * dT;
                if constexpr (EXACT) {
                    // the FP64 path leaves Q as (Q*dT)/dT like the reference; the
                    // mixed path keeps Q (they differ by at most an ulp)
                    const double qn = qd / dT;
                    if (qn != qv) q[p] = qn;
                }
            }
            double d[3];
            if constexpr (EXACT) {
                valid = cic_dim(px[k], o0, h0, l[0], d[0]);
                valid &= cic_dim(py[k], o1, h1, l[1], d[1]);
                valid &= cic_dim(pz[k], o2, h2, l[2], d[2]);
            } else {
                valid = cic_dim_fast(px[k], o0, ih0, l[0], d[0]);
                valid &= cic_dim_fast(py[k], o1, ih1, l[1], d[1]);
                valid &= cic_dim_fast(pz[k], o2, ih2, l[2], d[2]);
            }
            if (valid) {
                // a stencil starting outside [-1, n-1] has no corner in the grid
                const bool near = l[0] >= -1 && l[1] >= -1 && l[2] >= -1 && l[0] < n0 && l[1] < n1 && l[2] < n2;
                if (near) {
                    if constexpr (EXACT) {
                        double w[8];
                        cic_weights(d, w);
#pragma unroll
                        for (int c = 0; c < 8; ++c) v[c] = (long long)rint(w[c] * qd * e2q);
                    } else {
                        const float a0 = 1.0f - (float)d[0], a1 = 1.0f - (float)d[1], a2 = 1.0f - (float)d[2];
                        const float b0 = (float)d[0], b1 = (float)d[1], b2 = (float)d[2];
                        const float qs = (float)(qd * e2q);
                        const float a01 = a0 * a1 * qs, b01 = b0 * a1 * qs, a0b1 = a0 * b1 * qs, b0b1 = b0 * b1 * qs;
                        v[0] = (int)rintf(a01 * a2);
                        v[1] = (int)rintf(b01 * a2);
                        v[2] = (int)rintf(a0b1 * a2);
                        v[3] = (int)rintf(b0b1 * a2);
                        v[4] = (int)rintf(a01 * b2);
                        v[5] = (int)rintf(b01 * b2);
                        v[6] = (int)rintf(a0b1 * b2);
                        v[7] = (int)rintf(b0b1 * b2);
                    }
                    key = ((l[2] + 1) * (n1 + 1) + (l[1] + 1)) * (n0 + 1) + (l[0] + 1);
                }
            } else {
                nd += 8;
            }
        }
        if (key == cur_key) {
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] += v[c];
        } else {
            if (cur_live) flush();
            cur_key = key;
            cur_live = key >= 0;  // valid and near the grid
            cl[0] = l[0];
            cl[1] = l[1];
            cl[2] = l[2];
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = v[c];
        }
        // out-of-grid corners of valid particles are dropped, counted per particle
        if (valid && !(l[0] >= 0 && l[1] >= 0 && l[2] >= 0 && l[0] < n0 - 1 && l[1] < n1 - 1 && l[2] < n2 - 1)) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int i = l[0] + (c & 1), j = l[1] + ((c >> 1) & 1), kk = l[2] + ((c >> 2) & 1);
                if (!(i >= 0 && j >= 0 && kk >= 0 && i < n0 && j < n1 && kk < n2)) ++nd;
            }
        }
    }
    if (cur_live) flush();
    if (use_lds) {
        __syncthreads();
        const float rbx = 1.0f / (float)bx, rsxy = 1.0f / (float)sxy;
        for (int e = threadIdx.x; e < vol; e += kThreads) {
            const long long v = box[e];
            if (v != 0) {
                const int lk = small_div(e, sxy, rsxy);
                const int r = e - lk * sxy;
                const int lj = small_div(r, bx, rbx);
                const int li = r - lj * bx;
                rho_add<EXACT>(rho, ((size_t)(lk + bz0) * n1 + (lj + by0)) * n0 + (li + bx0), v, lsh);
            }
        }
    }
    if (nd) atomicAdd(dropped, nd);
}

bool deposit_large_box(int64_t n) { return n <= g_large_box_max; }
void set_deposit_large_box_max(int64_t n) { g_large_box_max = n; }
int64_t deposit_large_box_max() { return g_large_box_max; }

void launch_deposit_range(hipStream_t s, int mode, bool exact, int64_t lo, int64_t hi, int64_t Nall,
                          const double *x, const double *y, const double *z, double *q, const double *qd,
                          long long *rho, const Scalars *sc, unsigned long long *dropped,
                          const double *chunkbox) {
    if (hi <= lo) return;
    const int64_t chunk0 = lo / kChunk;
    const unsigned blocks = (unsigned)((hi - 1) / kChunk - chunk0 + 1);
    const bool box = chunkbox != nullptr;
    const bool large = deposit_large_box(hi - lo);
#define OPALSC_DEP(M_, E_, B_)                                                                              \
    do {                                                                                                    \
        if (large)                                                                                          \
            k_deposit_lds<M_, E_, B_, kBoxCapLarge><<<blocks, kThreads, 0, s>>>(                            \
                lo, hi, Nall, chunk0, x, y, z, q, qd, rho, sc, dropped, chunkbox);                          \
        else                                                                                                \
            k_deposit_lds<M_, E_, B_, kBoxCapSmall><<<blocks, kThreads, 0, s>>>(                            \
                lo, hi, Nall, chunk0, x, y, z, q, qd, rho, sc, dropped, chunkbox);                          \
    } while (0)
    if (mode == 0) {
        if (exact) OPALSC_DEP(0, true, false);
        else OPALSC_DEP(0, false, false);
    } else if (mode == 1) {
        if (box) {
            if (exact) OPALSC_DEP(1, true, true);
            else OPALSC_DEP(1, false, true);
        } else {
            if (exact) OPALSC_DEP(1, true, false);
            else OPALSC_DEP(1, false, false);
        }
    } else if (mode == 2) {
        if (box) {
            if (exact) OPALSC_DEP(2, true, true);
            else OPALSC_DEP(2, false, true);
        } else {
            if (exact) OPALSC_DEP(2, true, false);
            else OPALSC_DEP(2, false, false);
        }
    } else {
        if (box) {
            if (exact) OPALSC_DEP(3, true, true);
            else OPALSC_DEP(3, false, true);
        } else {
            if (exact) OPALSC_DEP(3, true, false);
            else OPALSC_DEP(3, false, false);
        }
    }
#undef OPALSC_DEP
}
void launch_deposit(hipStream_t s, int mode, bool exact, int64_t N, const double *x, const double *y,
                    const double *z, double *q, const double *qd, long long *rho, const Scalars *sc,
                    unsigned long long *dropped, const double *chunkbox) {
    launch_deposit_range(s, mode, exact, 0, N, N, x, y, z, q, qd, rho, sc, dropped, chunkbox);
}

// ---- CIC gather ---------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_gather_selffield(int64_t N, const double *x, const double *y,
                                                          const double *z,
                                                          const typename vec4_of<T>::type *Eg,
                                                          const Scalars *sc, double *Ef, double *Bf) {
    const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
    const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    const double betaC = sc->betaC;
    const bool pz = sc->pz != 0;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        int l[3];
        double d[3], e[3] = {0.0, 0.0, 0.0};
        bool ok = cic_dim(x[p], o0, h0, l[0], d[0]);
        ok &= cic_dim(y[p], o1, h1, l[1], d[1]);
        ok &= cic_dim(z[p], o2, h2, l[2], d[2]);
        if (ok) gather_E<T>(Eg, l, d, n0, n1, n2, e, pz);
        Ef[3 * p] = e[0];
        Ef[3 * p + 1] = e[1];
        Ef[3 * p + 2] = e[2];
        Bf[3 * p] = Bf[3 * p] - betaC * e[1];
        Bf[3 * p + 1] = Bf[3 * p + 1] + betaC * e[0];
    }
}

void launch_gather_selffield(hipStream_t s, bool fp64, int64_t N, const double *x, const double *y,
                             const double *z, const void *Eg, const Scalars *sc, double *Ef,
                             double *Bf) {
    if (N <= 0) return;
    if (fp64) k_gather_selffield<double><<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, (const double4 *)Eg, sc, Ef, Bf);
    else k_gather_selffield<float><<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, (const float4 *)Eg, sc, Ef, Bf);
}

__global__ void k_cic_index(int64_t N, const double *R, double o0, double o1, double o2, double h0,
                            double h1, double h2, int32_t *lgp, double *dpos) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        const double o[3] = {o0, o1, o2}, h[3] = {h0, h1, h2};
        for (int d = 0; d < 3; ++d) {
            int l = 0;
            double dp = 0.0;
            cic_dim(R[3 * p + d], o[d], h[d], l, dp);
            lgp[3 * p + d] = l;
            dpos[3 * p + d] = dp;
        }
    }
}

void launch_cic_index(hipStream_t s, int64_t N, const double *R, const double origin[3],
                      const double hr[3], int32_t *lgp, double *dpos) {
    if (N <= 0) return;
    k_cic_index<<<stream_blocks(N), kThreads, 0, s>>>(N, R, origin[0], origin[1], origin[2], hr[0],
                                                      hr[1], hr[2], lgp, dpos);
}

// ---- bounds, push, kick, moments (AoS entry points) --------------------------
__global__ __launch_bounds__(256) void k_bounds(int64_t N, const double *R, double *partials) {
    __shared__ double lds[4 * 6];
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], R[3 * i + d]);
            mx[d] = fmax(mx[d], R[3 * i + d]);
        }
    block_minmax(mn, mx, lds);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            partials[blockIdx.x * kSlots + d] = mn[d];
            partials[blockIdx.x * kSlots + 3 + d] = mx[d];
        }
}

void launch_bounds(hipStream_t s, int64_t N, const double *R, double *partials) {
    if (N > 0) k_bounds<<<red_blocks(N), kThreads, 0, s>>>(N, R, partials);
}

__global__ __launch_bounds__(512) void k_boundp_scalars(const double *partials, int nb, int n0, int n1,
                                                        int n2, double dh, int pz, double period, Scalars *sc) {
    __shared__ double lds[16 * 6];
    double mn[3], mx[3];
    reduce_partials_minmax(partials, nb, mn, mx, lds);
    if (threadIdx.x != 0) return;
    if (nb == 0)
        for (int d = 0; d < 3; ++d) mn[d] = mx[d] = 0.0;
    const int n[3] = {n0, n1, n2};
    for (int d = 0; d < 3; ++d) {
        sc->n[d] = n[d];
        sc->rmin[d] = mn[d];
        sc->rmax[d] = mx[d];
    }
    sc->mesh_ok = boundp_mesh(mn, mx, n, dh, sc->hr, sc->origin, pz != 0, period) ? 1 : 0;
}

void launch_boundp_scalars(hipStream_t s, const double *partials, int nb, const int n[3], double dh,
                           bool pz, double period, Scalars *sc) {
    k_boundp_scalars<<<1, kReduceThreads, 0, s>>>(partials, nb, n[0], n[1], n[2], dh, pz ? 1 : 0, period, sc);
}

__global__ void k_wrap_periodic_z(int64_t N, double *R, double lo, double hi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        R[3 * i + 2] = wrap_periodic(R[3 * i + 2], lo, hi);
}
void launch_wrap_periodic_z(hipStream_t s, int64_t N, double *R, double lo, double hi) {
    if (N > 0) k_wrap_periodic_z<<<stream_blocks(N), kThreads, 0, s>>>(N, R, lo, hi);
}

__global__ void k_fieldmap_eval(int64_t N, const double *R, double t, const FieldMapDesc *maps, int nmaps,
                                const double *data, double *E, double *B, unsigned long long *lost) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        double e[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
        bool out = false;  // a particle beyond one map's radius sees no later map (computeExternalFields)
        for (int k = 0; k < nmaps && !out; ++k)
            out = fieldmap_apply(maps[k], data, R[3 * i], R[3 * i + 1], R[3 * i + 2], t, e, b);
        for (int d = 0; d < 3; ++d) {
            E[3 * i + d] = e[d];
            B[3 * i + d] = b[d];
        }
        if (out) atomicAdd(lost, 1ull);
    }
}
void launch_fieldmap_eval(hipStream_t s, int64_t N, const double *R, double t, const FieldMapDesc *maps, int nmaps,
                          const double *data, double *E, double *B, unsigned long long *lost) {
    if (N > 0) k_fieldmap_eval<<<stream_blocks(N), kThreads, 0, s>>>(N, R, t, maps, nmaps, data, E, B, lost);
}

__global__ void k_push(int64_t N, double *R, const double *P, const double *dt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        double x = R[3 * i], y = R[3 * i + 1], z = R[3 * i + 2];
        push_half(x, y, z, P[3 * i], P[3 * i + 1], P[3 * i + 2], dt[i]);
        R[3 * i] = x;
        R[3 * i + 1] = y;
        R[3 * i + 2] = z;
    }
}
void launch_push(hipStream_t s, int64_t N, double *R, const double *P, const double *dt) {
    if (N > 0) k_push<<<stream_blocks(N), kThreads, 0, s>>>(N, R, P, dt);
}

__global__ void k_kick(int64_t N, double *P, const double *Ef, const double *Bf, const double *dt,
                       double charge, double mass) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        double p[3] = {P[3 * i], P[3 * i + 1], P[3 * i + 2]};
        const double e[3] = {Ef[3 * i], Ef[3 * i + 1], Ef[3 * i + 2]};
        const double b[3] = {Bf[3 * i], Bf[3 * i + 1], Bf[3 * i + 2]};
        boris_kick(p, e, b, dt[i], mass, charge);
        P[3 * i] = p[0];
        P[3 * i + 1] = p[1];
        P[3 * i + 2] = p[2];
    }
}
void launch_kick(hipStream_t s, int64_t N, double *P, const double *Ef, const double *Bf,
                 const double *dt, double charge, double mass) {
    if (N > 0) k_kick<<<stream_blocks(N), kThreads, 0, s>>>(N, P, Ef, Bf, dt, charge, mass);
}

// part = (x, px, y, py, z, pz): 6 sums, 21 products (calcMoments order), gamma sum.
__device__ inline void accumulate_moments(double (&m)[kMom], double x, double y, double z, double px,
                                          double py, double pz) {
    const double part[6] = {x, px, y, py, z, pz};
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] += part[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double r = part[2 * d], p = part[2 * d + 1];
        m[6 + 3 * d] += r * r;
        m[7 + 3 * d] += p * p;
        m[8 + 3 * d] += p * r;
    }
    m[15] += sqrt(1.0 + (px * px + py * py + pz * pz));
}
// The same, returning the particle's sqrt(1 + P.P) (the value added to m[15]).
__device__ inline double accumulate_moments_g(double (&m)[kMom], double x, double y, double z, double px,
                                              double py, double pz) {
    const double part[6] = {x, px, y, py, z, pz};
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] += part[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double r = part[2 * d], p = part[2 * d + 1];
        m[6 + 3 * d] += r * r;
        m[7 + 3 * d] += p * p;
        m[8 + 3 * d] += p * r;
    }
    const double g = lorentz_gamma(px, py, pz);
    m[15] += g;
    return g;
}

__device__ inline void store_moment_partials(double (&m)[kMom], double *partials, double *lds) {
    double l[kMom];
    for (int k = 0; k < kMom; ++k) l[k] = 0.0;
    block_sum_dd<kMom>(m, l, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < kMom; ++k) {
            partials[blockIdx.x * kSlots + 2 * k] = m[k];
            partials[blockIdx.x * kSlots + 2 * k + 1] = l[k];
        }
}

__global__ __launch_bounds__(256) void k_moments(int64_t N, const double *R, const double *P, double *partials) {
    __shared__ double lds[4 * kMom * 2];
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        accumulate_moments(m, R[3 * i], R[3 * i + 1], R[3 * i + 2], P[3 * i], P[3 * i + 1], P[3 * i + 2]);
    store_moment_partials(m, partials, lds);
}
void launch_moments(hipStream_t s, int64_t N, const double *R, const double *P, double *partials) {
    if (N > 0) k_moments<<<red_blocks(N), kThreads, 0, s>>>(N, R, P, partials);
}

__global__ __launch_bounds__(256) void k_moments_soa(int64_t N, const double *x, const double *y,
                                                     const double *z, const double *px,
                                                     const double *py, const double *pz, double *partials) {
    __shared__ double lds[4 * kMom * 2];
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        accumulate_moments(m, x[i], y[i], z[i], px[i], py[i], pz[i]);
    store_moment_partials(m, partials, lds);
}
void launch_moments_soa(hipStream_t s, const Bunch &b, double *partials) {
    if (b.N > 0) k_moments_soa<<<red_blocks(b.N), kThreads, 0, s>>>(b.N, b.x, b.y, b.z, b.px, b.py, b.pz, partials);
}

// Statistics file extras (calcBeamParameters' dispersion moments and the
// charge sum, get_bounds' maximum, PartBunchBase.hpp:1529-1613,1178-1218,1727-1729):
// per block double-double sums of x*pz, px*pz, y*pz, py*pz and Q (slots 0-9)
// and the maxima of x, y, z (slots 10-12); the particle whose id is 0 (the
// first one set) copied to p0[0..5].
__global__ __launch_bounds__(256) void k_stat_extra(int64_t N, const double *x, const double *y, const double *z,
                                                    const double *px, const double *py, const double *pz,
                                                    const double *q, double q_uniform, const uint32_t *id,
                                                    double *partials, double *p0) {
    __shared__ double lds[4 * 5 * 2];
    __shared__ double lmm[4 * 6];
    double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, l[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = __builtin_inf();
        mx[d] = -__builtin_inf();
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double v[5] = {x[i] * pz[i], px[i] * pz[i], y[i] * pz[i], py[i] * pz[i], q ? q[i] : q_uniform};
        for (int k = 0; k < 5; ++k) {
            dd r = dd_add(dd{h[k], l[k]}, dd{v[k], 0.0});
            h[k] = r.hi;
            l[k] = r.lo;
        }
        const double r3[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) mx[d] = fmax(mx[d], r3[d]);
        if (id[i] == 0) {
            p0[0] = x[i];
            p0[1] = y[i];
            p0[2] = z[i];
            p0[3] = px[i];
            p0[4] = py[i];
            p0[5] = pz[i];
        }
    }
    block_sum_dd<5>(h, l, lds);
    block_minmax(mn, mx, lmm);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) {
            partials[blockIdx.x * kSlots + 2 * k] = h[k];
            partials[blockIdx.x * kSlots + 2 * k + 1] = l[k];
        }
        for (int d = 0; d < 3; ++d) partials[blockIdx.x * kSlots + 10 + d] = mx[d];
    }
}

__global__ __launch_bounds__(512) void k_stat_extra_final(const double *partials, int nb, double *out) {
    __shared__ double lds[16 * 5 * 2];
    double s[5];
    reduce_partials_dd<5>(partials, nb, s, lds);
    __shared__ double lmm[16 * 6];
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = __builtin_inf();
        mx[d] = -__builtin_inf();
    }
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int d = 0; d < 3; ++d) mx[d] = fmax(mx[d], partials[b * kSlots + 10 + d]);
    block_minmax(mn, mx, lmm);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) out[k] = s[k];
        for (int d = 0; d < 3; ++d) out[5 + d] = mx[d];
    }
}

void launch_stat_extra(hipStream_t s, const Bunch &b, bool q_uniform, double q_value, double *partials,
                       double *out) {
    if (b.N <= 0) return;
    const int nb = red_blocks(b.N);
    k_stat_extra<<<nb, kThreads, 0, s>>>(b.N, b.x, b.y, b.z, b.px, b.py, b.pz, q_uniform ? nullptr : b.q,
                                         q_value, b.id, partials, out + 8);
    k_stat_extra_final<<<1, kReduceThreads, 0, s>>>(partials, nb, out);
}

__global__ __launch_bounds__(512) void k_stats_from_partials(const double *partials, int nb, int64_t N,
                                                             double mass_eV, StatsOut *out) {
    __shared__ double lds[16 * kMom * 2];
    double m[kMom];
    reduce_partials_dd<kMom>(partials, nb, m, lds);
    if (threadIdx.x != 0) return;
    StatsOut s;
    beam_parameters(N, m, mass_eV, s);
    s.t = 0.0;
    s.pathLength = 0.0;
    s.dropped = 0;
    *out = s;
}
void launch_stats_from_partials(hipStream_t s, const double *partials, int nb, int64_t N,
                                double mass_eV, void *stats_out) {
    k_stats_from_partials<<<1, kReduceThreads, 0, s>>>(partials, nb, N, mass_eV, (StatsOut *)stats_out);
}

// ---- sorting support ------------------------------------------------------------
// rotate: positions are in the reference frame; the cell is that of the
// beam-frame position M_rb (r - o_rb) the deposit uses.
// With energy bins (bin != nullptr) the key is bin-major: (bin << bshift) | (cell key >> drop).
__global__ void k_cell_keys(int64_t N, const double *x, const double *y, const double *z,
                            const Scalars *sc, int rotate, const int32_t *bin, int bshift, int drop,
                            uint32_t *keys, uint32_t *vals) {
    const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
    const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    const uint32_t out = cell_key_count(n0, n1, n2);
    const int hb = (31 - __builtin_clz(out)) / 3;  // bits per axis of the key cube
    double M[9], o[3];
    for (int k = 0; k < 9; ++k) M[k] = sc->M_rb[k];
    for (int d = 0; d < 3; ++d) o[d] = sc->o_rb[d];
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        int l[3];
        double d[3];
        double r[3] = {x[p], y[p], z[p]};
        if (rotate) transform_to_fast(M, o, r);
        bool ok = cic_dim(r[0], o0, h0, l[0], d[0]);
        ok &= cic_dim(r[1], o1, h1, l[1], d[1]);
        ok &= cic_dim(r[2], o2, h2, l[2], d[2]);
        // the tracker's re-sort (rotate: positions in the reference frame) orders
        // for locality only: Hilbert order, and a particle outside the grid files
        // with the last cell, so the keys need no extra bit for it; the exact
        // sort by cell orders by the Morton key, outside particles last
        uint32_t key = rotate ? out - 1 : out;
        if (ok && l[0] >= 0 && l[1] >= 0 && l[2] >= 0 && l[0] < n0 && l[1] < n1 && l[2] < n2)
            key = rotate ? cell_key_hilbert(l[0], l[1], l[2], hb) : cell_key(l[0], l[1], l[2], n0, n1);
        if (bin) key = ((uint32_t)bin[p] << bshift) | (key >> drop);
        else key >>= drop;
        keys[p] = key;
        vals[p] = (uint32_t)p;
    }
}
void launch_cell_keys(hipStream_t s, int64_t N, const double *x, const double *y, const double *z,
                      const Scalars *sc, bool rotate, uint32_t *keys, uint32_t *vals, const int32_t *bin,
                      int bshift, int drop) {
    if (N > 0)
        k_cell_keys<<<stream_blocks(N), kThreads, 0, s>>>(N, x, y, z, sc, rotate ? 1 : 0, bin, bshift, drop, keys,
                                                          vals);
}

// dst[i] = src[perm[i]]: each thread gathers kPermU elements with all of its
// loads in flight before the first store.
constexpr int kPermU = 8;
template <typename T>
__global__ __launch_bounds__(256) void k_permute(int64_t N, const uint32_t *__restrict__ perm,
                                                 const T *__restrict__ src, T *__restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * (kThreads * kPermU) + threadIdx.x;
    uint32_t pi[kPermU];
#pragma unroll
    for (int u = 0; u < kPermU; ++u) {
        const int64_t i = base + u * kThreads;
        pi[u] = perm[i < N ? i : N - 1];
    }
    T v[kPermU];
#pragma unroll
    for (int u = 0; u < kPermU; ++u) v[u] = src[pi[u]];
#pragma unroll
    for (int u = 0; u < kPermU; ++u) {
        const int64_t i = base + u * kThreads;
        if (i < N) dst[i] = v[u];
    }
}
static unsigned permute_blocks(int64_t N) { return (unsigned)((N + kThreads * kPermU - 1) / (kThreads * kPermU)); }
void launch_permute(hipStream_t s, int64_t N, const uint32_t *perm, const double *src, double *dst) {
    if (N > 0) k_permute<double><<<permute_blocks(N), kThreads, 0, s>>>(N, perm, src, dst);
}
void launch_permute_u32(hipStream_t s, int64_t N, const uint32_t *perm, const uint32_t *src, uint32_t *dst) {
    if (N > 0) k_permute<uint32_t><<<permute_blocks(N), kThreads, 0, s>>>(N, perm, src, dst);
}
__global__ void k_iota(int64_t N, uint32_t *v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}
void launch_iota(hipStream_t s, int64_t N, uint32_t *v) {
    if (N > 0) k_iota<<<stream_blocks(N), kThreads, 0, s>>>(N, v);
}

// Plain copy, one 16-byte element per thread over a full grid: the fastest
// copy form measured on MI355X (a persistent grid-stride loop reaches ~25 %
// less), used as the achievable-bandwidth reference and to calibrate the
// FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy_calibration(size_t n16, const uint4 *__restrict__ src,
                                                          uint4 *__restrict__ dst) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}
void launch_copy_calibration(hipStream_t s, size_t bytes, const void *src, void *dst) {
    const size_t n16 = bytes / 16;
    const size_t blocks = (n16 + 255) / 256;
    if (blocks == 0 || blocks > 0x7fffffff) return;
    k_copy_calibration<<<(unsigned)blocks, kThreads, 0, s>>>(n16, (const uint4 *)src, (uint4 *)dst);
}

__global__ void k_fill_i64(size_t n, long long *p) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}
void launch_fill_i64(hipStream_t s, size_t n, long long *p) {
    int blocks = (int)std::min<size_t>((n + 1023) / 1024, 16384);
    k_fill_i64<<<blocks, kThreads, 0, s>>>(n, p);
}

__global__ void k_absmax(int64_t N, const double *v, unsigned long long *out_bits) {
    double a = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        a = fmax(a, fabs(v[i]));
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, (unsigned long long)__double_as_longlong(a));
}
void launch_absmax(hipStream_t s, int64_t N, const double *v, unsigned long long *out_bits) {
    if (N > 0) k_absmax<<<red_blocks(N), kThreads, 0, s>>>(N, v, out_bits);
}

// ---- synthetic bunch ---------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double uniform01(uint64_t seed, uint64_t i, int s) {
    uint64_t h = splitmix64(seed ^ splitmix64(i * 16 + (uint64_t)s));
    return ((double)(h >> 11) + 0.5) * 0x1p-53;
}

__global__ void k_generate(int64_t N, uint64_t seed, double sx, double sy, double sz, double spx,
                           double spy, double spz, double p0, double qeach, double halo_frac,
                           double halo_radius, double *x, double *y, double *z, double *px,
                           double *py, double *pz, double *q, uint32_t *id) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        double nrm[6];
        for (int k = 0; k < 3; ++k) {
            double u1 = uniform01(seed, i, 2 * k), u2 = uniform01(seed, i, 2 * k + 1);
            double r = sqrt(-2.0 * log(u1)), th = 2.0 * kPi * u2;
            nrm[2 * k] = r * cos(th);
            nrm[2 * k + 1] = r * sin(th);
        }
        double X = sx * nrm[0], Y = sy * nrm[1], Z = sz * nrm[2];
        if (halo_frac > 0.0 && uniform01(seed, i, 6) < halo_frac) {
            double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            double rad = halo_radius * cbrt(uniform01(seed, i, 7)) / len;
            X = sx * nrm[0] * rad;
            Y = sy * nrm[1] * rad;
            Z = sz * nrm[2] * rad;
        }
        x[i] = X;
        y[i] = Y;
        z[i] = Z;
        px[i] = spx * nrm[3];
        py[i] = spy * nrm[4];
        pz[i] = p0 + spz * nrm[5];
        q[i] = qeach;
        id[i] = (uint32_t)i;
    }
}
void launch_generate(hipStream_t s, int64_t N, uint64_t seed, const double sr[3], const double sp[3],
                     double p0, double qeach, double halo_frac, double halo_radius, Bunch &b) {
    if (N > 0)
        k_generate<<<stream_blocks(N), kThreads, 0, s>>>(N, seed, sr[0], sr[1], sr[2], sp[0], sp[1], sp[2], p0,
                                                         qeach, halo_frac, halo_radius, b.x, b.y, b.z, b.px,
                                                         b.py, b.pz, b.q, b.id);
}

__global__ void k_cr_test(int which, int64_t n, const double *x, double *y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = which == 0 ? crm::atan_cr(x[i]) : crm::log_cr(x[i]);
}
void launch_cr_test(hipStream_t s, int which, int64_t n, const double *x, double *y) {
    if (n > 0) k_cr_test<<<stream_blocks(n), kThreads, 0, s>>>(which, n, x, y);
}
__global__ void k_igf_term_test(int fast, int64_t n, const double *v, double *y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fast ? igf::term_fast(v[3 * i], v[3 * i + 1], v[3 * i + 2], 1.0)
                    : igf::term_cr(v[3 * i], v[3 * i + 1], v[3 * i + 2], 1.0);
}
void launch_igf_term_test(hipStream_t s, int fast, int64_t n, const double *v, double *y) {
    if (n > 0) k_igf_term_test<<<stream_blocks(n), kThreads, 0, s>>>(fast, n, v, y);
}

// ---- tracker ----------------------------------------------------------------------
// Step-start scalars: beam-frame trafos from the bunch's mean momentum
// (ParallelTTracker::computeSpaceChargeFields, ParallelTTracker.cpp:549-559),
// external fields at t + dT/2 (ParallelTTracker.cpp:630), then t += dT and the
// reference-particle update whose trafo transformBunch applies at step end
// (ParallelTTracker.cpp:362-367,1199-1213).
template <bool BATCH = false>
__global__ void k_track_pre(TrackState *ts, Scalars *sc, int64_t N, unsigned long long *dropped,
                            BatchView bv = BatchView{}) {
    if constexpr (BATCH) {
        if (threadIdx.x != 0) return;
        const BatchItem &it = bv.items[blockIdx.x];
        ts = it.ts;
        sc = it.sc;
        N = it.N;
        dropped = it.dropped;
    } else if (threadIdx.x != 0 || blockIdx.x != 0) {
        return;
    }
    *dropped = 0;  // the step's count of deposits outside the grid
    // the maps are formed in registers and stored once (stores through sc
    // could alias what is read back, which would reload every operand)
    TrackState s = *ts;
    sc->dT = s.dT;
    sc->coupling = s.coupling;
    sc->N = N;
    double Mb[9], Mr[9], ob[3], orb[3];
    if (s.space_charge && N > 0) {
        const double pmean[3] = {sc->psum[0] / (double)N, sc->psum[1] / (double)N, sc->psum[2] / (double)N};
        double qa[4];
        quat_to_z(pmean, qa);
        const double qb[4] = {qa[0], -qa[1], -qa[2], -qa[3]};
        const double o[3] = {0.0, 0.0, s.pathLength};
        Trafo b2r = trafo_make(o, qb);
        Trafo r2b = trafo_inverted(b2r);
        for (int k = 0; k < 9; ++k) {
            Mb[k] = b2r.M[k];
            Mr[k] = r2b.M[k];
        }
        for (int d = 0; d < 3; ++d) {
            ob[d] = b2r.origin[d];
            orb[d] = r2b.origin[d];
        }
    } else {
        // no beam frame: the identity (only the mixed-precision gather pass reads it then)
        for (int k = 0; k < 9; ++k) Mb[k] = Mr[k] = (k % 4 == 0) ? 1.0 : 0.0;
        for (int d = 0; d < 3; ++d) ob[d] = orb[d] = 0.0;
    }
    double Eext[3], Bext[3];
    ext_fields(s, s.t + 0.5 * s.dT, Eext, Bext);
    sc->fm_s0 = s.pathLength;  // the lab frame of the field maps during this step
    sc->fm_t = s.t + 0.5 * s.dT;
    s.t += s.dT;
    Trafo upd;
    if (s.t > 0.0) {
        upd = update_reference(s);
    } else {
        const double o[3] = {0.0, 0.0, 0.0}, q[4] = {1.0, 0.0, 0.0, 0.0};
        upd = trafo_make(o, q);
    }
    // composed maps of the mixed-precision gather pass: r_ref = M_br (r_b - o_br)
    // and r_new = M_u (r_ref - o_u) give r_new = M_c (r_b - o_c) with M_c = M_u M_br,
    // o_c = o_br + M_br^T o_u (M_br orthogonal)
    const double *Mu = upd.M;
    double Mc[9], oc[3], Eb[3], Bb[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Mc[3 * i + j] = Mu[3 * i] * Mb[j] + Mu[3 * i + 1] * Mb[3 + j] + Mu[3 * i + 2] * Mb[6 + j];
    for (int d = 0; d < 3; ++d)
        oc[d] = ob[d] + (Mb[d] * upd.origin[0] + Mb[3 + d] * upd.origin[1] + Mb[6 + d] * upd.origin[2]);
    mat_vec(Mr, Eext, Eb);
    mat_vec(Mr, Bext, Bb);
    for (int k = 0; k < 9; ++k) {
        sc->M_br[k] = Mb[k];
        sc->M_rb[k] = Mr[k];
        sc->M_u[k] = upd.M[k];
        sc->M_c[k] = Mc[k];
    }
    for (int d = 0; d < 3; ++d) {
        sc->o_br[d] = ob[d];
        sc->o_rb[d] = orb[d];
        sc->o_u[d] = upd.origin[d];
        sc->o_c[d] = oc[d];
        sc->o_c2[d] = oc[d];
        sc->o_rb2[d] = orb[d];
        sc->ih[d] = 0.0;
        sc->Eext[d] = Eext[d];
        sc->Bext[d] = Bext[d];
        sc->Eext_b[d] = Eb[d];
        sc->Bext_b[d] = Bb[d];
    }
    s.step += 1;
    *ts = s;
}
void launch_track_pre(hipStream_t s, TrackState *ts, Scalars *sc, int64_t N, unsigned long long *dropped) {
    k_track_pre<<<1, 64, 0, s>>>(ts, sc, N, dropped);
}

// PUSH: the step's first half push (timeIntegration1), written back; otherwise
// the positions already hold it (the previous step's gather pass applied it).
// Then the rotation into the beam frame (not stored: the deposit and the gather
// pass repeat it bit for bit) and its bounds (boundp's get_bounds).  Each block
// walks contiguous chunks of kChunk particles (the deposit's chunks), writes
// every chunk's beam-frame position statistics (bounds, sums, sums of squares)
// to chunkbox (so the deposit reads positions once), and leaves its running
// bounds in partials.
template <bool EXACT, bool PUSH, bool BATCH = false>
__global__ __launch_bounds__(256) void k_track_alpha(int64_t N, double *__restrict__ x,
                                                     double *__restrict__ y, double *__restrict__ z,
                                                     const double *__restrict__ px,
                                                     const double *__restrict__ py,
                                                     const double *__restrict__ pz, const Scalars *sc,
                                                     const TrackState *ts, double *partials,
                                                     double *chunkbox, BatchView bv = BatchView{}) {
    __shared__ double lds[4 * kBoxStats];
    int bid = (int)blockIdx.x, nblk = (int)gridDim.x;
    if constexpr (BATCH) {
        const BatchItem &it = bv.items[batch_find(bv, bid, nblk)];
        N = it.N;
        x = it.x;
        y = it.y;
        z = it.z;
        px = it.px;
        py = it.py;
        pz = it.pz;
        sc = it.sc;
        ts = it.ts;
        partials = it.partials;
        chunkbox = it.chunkbox;
    }
    const double dT = ts->dT;
    const double half_cdt = 0.5 * (kC * dT);
    const bool sc_on = ts->space_charge != 0;
    double M[9], o[3];
    for (int k = 0; k < 9; ++k) M[k] = sc->M_rb[k];
    for (int d = 0; d < 3; ++d) o[d] = sc->o_rb[d];
    double bmn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double bmx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    for (int64_t c = bid; c < nchunks; c += nblk) {
        double st[kBoxStats];
        for (int d = 0; d < 3; ++d) {
            st[d] = __builtin_inf();
            st[3 + d] = -__builtin_inf();
            st[6 + d] = st[9 + d] = 0.0;
        }
        const int64_t base = c * kChunk + threadIdx.x;
#pragma unroll 2
        for (int k = 0; k < kPPT; ++k) {
            const int64_t i = base + k * kThreads;
            if (i >= N) break;
            double r[3] = {x[i], y[i], z[i]};
            if constexpr (PUSH) {
                if constexpr (EXACT) push_half(r[0], r[1], r[2], px[i], py[i], pz[i], dT);
                else push_half_fast(r[0], r[1], r[2], px[i], py[i], pz[i], half_cdt);
                x[i] = r[0];
                y[i] = r[1];
                z[i] = r[2];
            }
            if (sc_on) {
                if constexpr (EXACT) transform_to(M, o, r);
                else transform_to_fast(M, o, r);
            }
            for (int d = 0; d < 3; ++d) {
                st[d] = fmin(st[d], r[d]);
                st[3 + d] = fmax(st[3 + d], r[d]);
                st[6 + d] += r[d];
                st[9 + d] += r[d] * r[d];
            }
        }
        block_chunk_stats(st, lds);
        if (threadIdx.x == 0) {
            for (int k = 0; k < kBoxStats; ++k) chunkbox[c * kBoxStats + k] = st[k];
            for (int d = 0; d < 3; ++d) {
                bmn[d] = fmin(bmn[d], st[d]);
                bmx[d] = fmax(bmx[d], st[3 + d]);
            }
        }
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            partials[bid * kSlots + d] = bmn[d];
            partials[bid * kSlots + 3 + d] = bmx[d];
        }
}
__host__ __device__ inline int alpha_blocks_of(int64_t N) {
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    return (int)(nchunks < 1 ? 1 : nchunks > kRedBlocksMax ? kRedBlocksMax : nchunks);
}
int alpha_blocks(int64_t N) { return alpha_blocks_of(N); }
void launch_track_alpha(hipStream_t s, bool exact, bool push, Bunch &b, const Scalars *sc,
                        const TrackState *ts, double *partials) {
    if (b.N <= 0) return;
#define OPALSC_ALPHA(E_, P_)                                                                              \
    k_track_alpha<E_, P_><<<alpha_blocks(b.N), kThreads, 0, s>>>(b.N, b.x, b.y, b.z, b.px, b.py, b.pz, sc, \
                                                                 ts, partials, b.chunkbox)
    if (exact) {
        if (push) OPALSC_ALPHA(true, true);
        else OPALSC_ALPHA(true, false);
    } else {
        if (push) OPALSC_ALPHA(false, true);
        else OPALSC_ALPHA(false, false);
    }
#undef OPALSC_ALPHA
}

// ---- bounds of the beam-frame positions from the previous gather pass's boxes
// The one-shot gather pass leaves, per block of consecutive particles, the
// statistics (min, max, sum, sum of squares per axis) of the positions it
// wrote, which are the positions the next step's bounds pass reads.  Rotated
// into this step's beam frame, each 2048-particle chunk's box gives an interval
// [L_d, U_d] that holds every particle's beam-frame coordinate d, and a particle
// on the box's face bounds the chunk's minimum from above (its maximum from
// below).  Only chunks whose interval reaches past the best such bound over all
// chunks can hold a global extreme; k_bounds_scan rotates their particles
// exactly as the full bounds pass does, so the bounds (hence the mesh) are the
// same bit for bit while a few percent of the positions are read.  The
// intervals are widened by 16 ulps of sum_j |M_dj| max|r_j - o_j|, more than the
// rounding of either the interval or a particle's rotation can move them.
constexpr double kBoxMarginUlps = 16.0 * 0x1p-53;

// OSPC: one-shot gather blocks per 2048-particle chunk (2 mixed, 4 FP64).  Also
// writes each chunk's widened beam-frame interval as the deposit's bounding
// box, with NaN sums: a deposit block whose box overflows its LDS box forms
// the chunk's statistics itself.
// The block's bounds go to slots kBoxRowSlot.. of its row of the partial-sums
// buffer, beside the slots 0-5 that k_bounds_scan writes: per-block rows
// instead of six device-scope atomics per block, which execute at the memory
// side one after another on the same six words.
constexpr int kBoxRowSlot = 8;
template <int OSPC>
__global__ __launch_bounds__(256) void k_bounds_from_boxes(int64_t N, const double *__restrict__ osbox,
                                                           const Scalars *sc, double *chunkbox,
                                                           double *rows) {
    __shared__ double lds[4 * 6];
    constexpr int kOs = kChunk / OSPC;
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    const int64_t nos = (N + kOs - 1) / kOs;
    double M[9], o[3];
    for (int k = 0; k < 9; ++k) M[k] = sc->M_rb[k];
    for (int d = 0; d < 3; ++d) o[d] = sc->o_rb[d];
    double umin[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double lmax[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nchunks; c += (int64_t)gridDim.x * blockDim.x) {
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = __builtin_inf();
            hi[d] = -__builtin_inf();
        }
        for (int q = 0; q < OSPC; ++q) {
            const int64_t k = c * OSPC + q;
            if (k >= nos) break;
            const double *b = osbox + k * kOsBoxStats;
            for (int d = 0; d < 3; ++d) {
                lo[d] = fmin(lo[d], b[d]);
                hi[d] = fmax(hi[d], b[3 + d]);
            }
        }
        double wl[3], wh[3];
        for (int j = 0; j < 3; ++j) {
            wl[j] = lo[j] - o[j];
            wh[j] = hi[j] - o[j];
        }
        double *cb = chunkbox + c * kBoxStats;
        for (int d = 0; d < 3; ++d) {
            double a[3], b[3], S = 0.0;
            for (int j = 0; j < 3; ++j) {
                a[j] = M[3 * d + j] * wl[j];
                b[j] = M[3 * d + j] * wh[j];
                S += fabs(M[3 * d + j]) * fmax(fabs(wl[j]), fabs(wh[j]));
            }
            const double mg = kBoxMarginUlps * S + 1e-300;
            double L = 0.0, U = 0.0, Lx = 0.0, Ux = 0.0;
            for (int j = 0; j < 3; ++j) {
                L += fmin(a[j], b[j]);
                U += fmax(a[j], b[j]);
                if (j != d) {
                    Lx += fmin(a[j], b[j]);
                    Ux += fmax(a[j], b[j]);
                }
            }
            const bool pos = M[4 * d] >= 0.0;
            umin[d] = fmin(umin[d], (pos ? a[d] : b[d]) + Ux + mg);
            lmax[d] = fmax(lmax[d], (pos ? b[d] : a[d]) + Lx - mg);
            cb[d] = L - mg;
            cb[3 + d] = U + mg;
            cb[6 + d] = cb[9 + d] = __builtin_nan("");
        }
    }
    block_minmax(umin, lmax, lds);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            rows[blockIdx.x * kSlots + kBoxRowSlot + d] = umin[d];
            rows[blockIdx.x * kSlots + kBoxRowSlot + 3 + d] = lmax[d];
        }
}

// The exact bounds over the chunks that can hold an extreme (see above), in
// the full bounds pass's partials layout (min in slots 0-2, max in 3-5).
template <bool EXACT>
__global__ __launch_bounds__(256) void k_bounds_scan(int64_t N, const double *__restrict__ x,
                                                     const double *__restrict__ y, const double *__restrict__ z,
                                                     const Scalars *sc, const double *chunkbox, int nbox,
                                                     double *partials) {
    __shared__ double lds[4 * 6];
    __shared__ double best[6];
    double M[9], o[3], umin[3], lmax[3];
    for (int k = 0; k < 9; ++k) M[k] = sc->M_rb[k];
    for (int d = 0; d < 3; ++d) {
        o[d] = sc->o_rb[d];
        umin[d] = __builtin_inf();
        lmax[d] = -__builtin_inf();
    }
    // the best bounds over the nbox rows k_bounds_from_boxes left (slots the
    // blocks' own results, slots 0-5, do not touch)
    for (int r = threadIdx.x; r < nbox; r += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            umin[d] = fmin(umin[d], partials[r * kSlots + kBoxRowSlot + d]);
            lmax[d] = fmax(lmax[d], partials[r * kSlots + kBoxRowSlot + 3 + d]);
        }
    block_minmax(umin, lmax, lds);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            best[d] = umin[d];
            best[3 + d] = lmax[d];
        }
    __syncthreads();
    for (int d = 0; d < 3; ++d) {
        umin[d] = best[d];
        lmax[d] = best[3 + d];
    }
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    // the block's chunks are blockIdx.x + k gridDim.x (the candidates of a
    // sorted bunch lie close together in memory; striding spreads them over the
    // blocks).  Each thread tests one chunk's box, kThreads chunks at a time, so
    // the tests' loads overlap; the block then scans the candidates in order.
    __shared__ unsigned long long cmask[kThreads / 64];
    for (int64_t k0 = 0; blockIdx.x + k0 * gridDim.x < nchunks; k0 += kThreads) {
        const int64_t c = blockIdx.x + (k0 + threadIdx.x) * gridDim.x;
        bool cand = false;
        if (c < nchunks) {
            const double *cb = chunkbox + c * kBoxStats;
            for (int d = 0; d < 3; ++d) cand |= (cb[d] <= umin[d]) || (cb[3 + d] >= lmax[d]);
        }
        const unsigned long long m = __ballot(cand);
        if ((threadIdx.x & 63) == 0) cmask[threadIdx.x >> 6] = m;
        __syncthreads();
        for (int w = 0; w < kThreads / 64; ++w) {
            unsigned long long mw = cmask[w];
            while (mw) {
                const int b = __builtin_ctzll(mw);
                mw &= mw - 1;
                const int64_t base = (blockIdx.x + (k0 + w * 64 + b) * gridDim.x) * kChunk + threadIdx.x;
#pragma unroll 2
                for (int k = 0; k < kPPT; ++k) {
                    const int64_t i = base + k * kThreads;
                    if (i >= N) break;
                    double r[3] = {x[i], y[i], z[i]};
                    if constexpr (EXACT) transform_to(M, o, r);
                    else transform_to_fast(M, o, r);
                    for (int d = 0; d < 3; ++d) {
                        mn[d] = fmin(mn[d], r[d]);
                        mx[d] = fmax(mx[d], r[d]);
                    }
                }
            }
        }
        __syncthreads();  // every wave has read cmask before the next round writes it
    }
    block_minmax(mn, mx, lds);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            partials[blockIdx.x * kSlots + d] = mn[d];
            partials[blockIdx.x * kSlots + 3 + d] = mx[d];
        }
}

int64_t os_box_count(bool fp64, int64_t N) {
    const int64_t k = fp64 ? kOsChunk64 : kOsChunk;
    return (N + k - 1) / k;
}
static int box_bound_blocks(int64_t N) {
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    return (int)std::min<int64_t>((nchunks + kThreads - 1) / kThreads, 1024);
}
void launch_bounds_from_boxes(hipStream_t s, bool fp64, const Bunch &b, const double *osbox, const Scalars *sc,
                              double *partials) {
    if (b.N <= 0) return;
    const int blocks = box_bound_blocks(b.N);
    static_assert(kChunk == 2 * kOsChunk && kChunk == 4 * kOsChunk64, "one-shot blocks tile the chunks");
    static_assert(kBoxRowSlot >= 6 && kBoxRowSlot + 6 <= kSlots, "box rows beside the scan's slots 0-5");
    if (fp64) k_bounds_from_boxes<4><<<blocks, kThreads, 0, s>>>(b.N, osbox, sc, b.chunkbox, partials);
    else k_bounds_from_boxes<2><<<blocks, kThreads, 0, s>>>(b.N, osbox, sc, b.chunkbox, partials);
}
// A few thousand chunk boxes are checked per block (one broadcast read each)
// and only the candidates scanned, so 1024 blocks suffice; fewer rows for
// launch_track_mesh to fold.
constexpr int kScanBlocksMax = 1024;
int bounds_scan_blocks(int64_t N) { return std::min(alpha_blocks_of(N), kScanBlocksMax); }
void launch_bounds_scan(hipStream_t s, bool exact, const Bunch &b, const Scalars *sc, double *partials) {
    if (b.N <= 0) return;
    const int blocks = bounds_scan_blocks(b.N);  // the partial rows launch_track_mesh reduces
    const int nbox = box_bound_blocks(b.N);
    if (exact) k_bounds_scan<true><<<blocks, kThreads, 0, s>>>(b.N, b.x, b.y, b.z, sc, b.chunkbox, nbox, partials);
    else k_bounds_scan<false><<<blocks, kThreads, 0, s>>>(b.N, b.x, b.y, b.z, sc, b.chunkbox, nbox, partials);
}

template <bool BATCH = false>
__global__ __launch_bounds__(512) void k_track_mesh(const double *partials, int nb, int n0, int n1,
                                                    int n2, double dh, const TrackState *ts, int exact,
                                                    int q_roundtrip, Scalars *sc, BatchView bv = BatchView{}) {
    __shared__ double lds[16 * 6];
    if constexpr (BATCH) {
        const BatchItem &it = bv.items[blockIdx.x];
        partials = it.partials;
        nb = alpha_blocks_of(it.N);
        n0 = it.n[0];
        n1 = it.n[1];
        n2 = it.n[2];
        dh = it.dh;
        ts = it.ts;
        sc = it.sc;
    }
    double mn[3], mx[3];
    reduce_partials_minmax(partials, nb, mn, mx, lds);
    if (threadIdx.x != 0) return;
    const int n[3] = {n0, n1, n2};
    for (int d = 0; d < 3; ++d) {
        sc->n[d] = n[d];
        sc->rmin[d] = mn[d];
        sc->rmax[d] = mx[d];
    }
    sc->mesh_ok = boundp_mesh(mn, mx, n, dh, sc->hr, sc->origin, ts->pz != 0, ts->period) ? 1 : 0;
    sc->pz = ts->pz;
    sc->period = ts->period;
    sc->wrap_lo = 0 * sc->hr[2] + sc->origin[2];
    sc->wrap_hi = n2 * sc->hr[2] + sc->origin[2];
    // mixed-precision gather pass: beam-frame positions relative to the mesh origin
    const double *Mr = sc->M_rb, *og = sc->origin;
    for (int d = 0; d < 3; ++d) {
        sc->o_rb2[d] = sc->o_rb[d] + (Mr[d] * og[0] + Mr[3 + d] * og[1] + Mr[6 + d] * og[2]);
        sc->o_c2[d] = sc->o_c[d] - og[d];
        sc->ih[d] = 1.0 / sc->hr[d];
    }
    const int64_t N = sc->N;
    // a bunch of equal charges: Q *= dt for the scatter, Q /= dt after it
    // (PartBunch.cpp:493-495), once for all particles; the mixed path keeps Q,
    // and so does the energy-binned step (its loop restores Q)
    if (q_roundtrip) {
        sc->qd_u = sc->qu * ts->dT;
        if (exact) sc->qu = sc->qd_u / ts->dT;
    }
    selffield_scalars(sc, sc->psum[2], N, ts->dT, ts->coupling);
    // every particle deposits |Q*dT| <= max|Q|*dT*(1+2^-52)
    set_fixed_point(sc, N, sc->qd_absmax, exact != 0);
}
void launch_track_mesh(hipStream_t s, const double *partials, int nb, const int n[3], double dh,
                       const TrackState *ts, bool exact, bool q_roundtrip, Scalars *sc) {
    k_track_mesh<<<1, kReduceThreads, 0, s>>>(partials, nb, n[0], n[1], n[2], dh, ts, exact ? 1 : 0,
                                              q_roundtrip ? 1 : 0, sc);
}

// Rotation into the beam frame (the bounds pass's, repeated bit for bit),
// gather + B from E (PartBunch.cpp:632-644), rotation back to the reference
// frame (ParallelTTracker.cpp:586-591), external fields, kick + push
// (timeIntegration2), transformBunch with the reference-frame update, and the
// moment sums of the step's statistics.  NEXT: also the next step's first half
// push (timeIntegration1 with the updated P), so that pass reads positions only.
// BINNED (energy bins): Eg and Bg hold the sums over the bins of E and of the
// magnetic field each bin and its image imply (ParallelTTracker.cpp:567-579),
// both gathered.  Four waves per SIMD (at most 128 VGPRs) keep enough loads in
// flight.  The FP64 case without a DC beam or field maps runs in
// k_track_gamma_os instead.
template <bool NEXT, bool BINNED, bool MAPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_track_gamma(
    int64_t N, double *x, double *y, double *z, double *px, double *py, double *pz, const double4 *Eg,
    const double4 *Bg, const Scalars *sc, const TrackState *ts, double *partials) {
    __shared__ double lds[4 * kMom * 2];
    const bool sc_on = ts->space_charge != 0;
    const double dT = ts->dT, charge = ts->charge, mass = ts->mass_eV;
    const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
    const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    const double betaC = sc->betaC;
    const bool dcz = sc->pz != 0;
    const double wlo = sc->wrap_lo, whi = sc->wrap_hi;
    double Mrb[9], orb[3], Mb[9], ob[3], Mu[9], ou[3], Ee[3], Be[3];
    for (int k = 0; k < 9; ++k) {
        Mrb[k] = sc->M_rb[k];
        Mb[k] = sc->M_br[k];
        Mu[k] = sc->M_u[k];
    }
    for (int d = 0; d < 3; ++d) {
        orb[d] = sc->o_rb[d];
        ob[d] = sc->o_br[d];
        ou[d] = sc->o_u[d];
        Ee[d] = sc->Eext[d];
        Be[d] = sc->Bext[d];
    }
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    // XCD-aware grid-stride: workgroups are dispatched to the XCDs round-robin
    // (block b runs on XCD b % kXcds), so the blocks of one XCD stream one
    // contiguous eighth of the bunch together.  A sorted bunch then gathers
    // each field-grid line into one XCD's L2 instead of into several.  (One
    // block per 2048-particle chunk in the same XCD order measured slower:
    // 2.81 against 2.45 ms at 1e8 particles; the per-block moment reduction
    // and set-up outweigh the streaming gain.)
    const int xcds = (gridDim.x % kXcds == 0) ? kXcds : 1;
    const int64_t xper = ((N + xcds - 1) / xcds + kThreads - 1) / kThreads * kThreads;
    const int64_t xbeg = (int64_t)(blockIdx.x % xcds) * xper;
    const int64_t xend = xbeg + xper < N ? xbeg + xper : N;
    const int64_t xstride = (int64_t)(gridDim.x / xcds) * kThreads;
    for (int64_t i = xbeg + (int64_t)(blockIdx.x / xcds) * kThreads + threadIdx.x; i < xend; i += xstride) {
        double r[3] = {x[i], y[i], z[i]};
        double p[3] = {px[i], py[i], pz[i]};
        double Ef[3] = {0.0, 0.0, 0.0}, Bf[3] = {0.0, 0.0, 0.0};
        if (sc_on) {
            transform_to(Mrb, orb, r);
            if (dcz) r[2] = wrap_periodic(r[2], wlo, whi);  // boundp's update() of a DC beam
            int l[3];
            double d[3];
            bool ok = cic_dim(r[0], o0, h0, l[0], d[0]);
            ok &= cic_dim(r[1], o1, h1, l[1], d[1]);
            ok &= cic_dim(r[2], o2, h2, l[2], d[2]);
            if (ok) {
                gather_E<double>(Eg, l, d, n0, n1, n2, Ef, dcz);
                if constexpr (BINNED) gather_E<double>(Bg, l, d, n0, n1, n2, Bf);
            }
            if constexpr (!BINNED) {
                Bf[0] = Bf[0] - betaC * Ef[1];
                Bf[1] = Bf[1] + betaC * Ef[0];
            }
            double t3[3];
            transform_to(Mb, ob, r);
            mat_vec(Mb, Ef, t3);
            Ef[0] = t3[0];
            Ef[1] = t3[1];
            Ef[2] = t3[2];
            mat_vec(Mb, Bf, t3);
            Bf[0] = t3[0];
            Bf[1] = t3[1];
            Bf[2] = t3[2];
        }
        for (int k = 0; k < 3; ++k) {
            Ef[k] += Ee[k];
            Bf[k] += Be[k];
        }
        if constexpr (MAPS) {
            // computeExternalFields (ParallelTTracker.cpp:595-640) at the
            // position the space-charge rotation left, in the step's lab frame
            const double zlab = r[2] + sc->fm_s0, tf = sc->fm_t;
            for (int k = 0; k < ts->nmaps; ++k)
                if (fieldmap_apply(ts->maps[k], ts->mapdata, r[0], r[1], zlab, tf, Ef, Bf)) break;
        }
        double pn[3];
        boris_kick(p, Ef, Bf, dT, mass, charge);
        push_half(r[0], r[1], r[2], p[0], p[1], p[2], dT);
        transform_to(Mu, ou, r);
        mat_vec(Mu, p, pn);
        accumulate_moments(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
        if constexpr (NEXT) push_half(r[0], r[1], r[2], pn[0], pn[1], pn[2], dT);
        x[i] = r[0];
        y[i] = r[1];
        z[i] = r[2];
        px[i] = pn[0];
        py[i] = pn[1];
        pz[i] = pn[2];
    }
    store_moment_partials(m, partials, lds);
}

// The mixed-precision gather pass: the same step with FP32 field grids, the
// kick and push done in the beam frame (Boris kick and drift commute with
// rotations) and one composed map M_c to the updated reference frame, so the
// loop needs 33 frame constants instead of 42 and no round trip of the
// position through the reference frame.  Positions, momenta and moments stay
// in FP64.  EXT: 0 no external field (the loop's 27 frame constants then fit
// the scalar registers), 1 the analytic fields, 2 those and field maps.  DC: a
// DC beam (periodic z), a separate instance so the common case carries none of it.
template <bool NEXT, bool BINNED, int EXT, bool DC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_track_gamma_mixed(
    int64_t N, double *x, double *y, double *z, double *px, double *py, double *pz, const float4 *Eg,
    const float4 *Bg, const Scalars *sc, const TrackState *ts, double *partials) {
    __shared__ double lds[4 * kMom * 2];
    const bool sc_on = ts->space_charge != 0;
    const double dT = ts->dT, charge = ts->charge, mass = ts->mass_eV;
    const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
    const double betaC = sc->betaC;
    const double zper = DC ? n2 * sc->hr[2] : 0.0;  // DC beam: period of the mesh-relative z
    double Mr[9], Mc[9], orb[3], oc[3], ih[3], Ee[3], Be[3];
    for (int k = 0; k < 9; ++k) {
        Mr[k] = sc->M_rb[k];
        Mc[k] = sc->M_c[k];
    }
    for (int d = 0; d < 3; ++d) {
        orb[d] = sc->o_rb2[d];
        oc[d] = sc->o_c2[d];
        ih[d] = sc->ih[d];
        Ee[d] = sc->Eext_b[d];
        Be[d] = sc->Bext_b[d];
    }
    const double e_kick = 0.5 * dT * charge * kC / mass, b_kick = 0.5 * dT * charge * kC * kC / mass;
    const double half_cdt = 0.5 * (kC * dT);
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    // XCD-aware grid-stride, as in k_track_gamma
    const int xcds = (gridDim.x % kXcds == 0) ? kXcds : 1;
    const int64_t xper = ((N + xcds - 1) / xcds + kThreads - 1) / kThreads * kThreads;
    const int64_t xbeg = (int64_t)(blockIdx.x % xcds) * xper;
    const int64_t xend = xbeg + xper < N ? xbeg + xper : N;
    const int64_t xstride = (int64_t)(gridDim.x / xcds) * kThreads;
    // Software pipelined: the next particle's position and momentum are loaded
    // before this particle's field gather, so the two dependent round trips
    // (particle, then its field cells) overlap across iterations.  Three waves
    // per SIMD leave room for the extra registers (measured 2.41 -> 2.31 ms at
    // 1e8 particles; at four waves the loop spills and slows to 3.8 ms).
    int64_t i = xbeg + (int64_t)(blockIdx.x / xcds) * kThreads + threadIdx.x;
    double nr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < xend) {
        nr[0] = x[i];
        nr[1] = y[i];
        nr[2] = z[i];
        nr[3] = px[i];
        nr[4] = py[i];
        nr[5] = pz[i];
    }
    for (; i < xend; i += xstride) {
        double r[3] = {nr[0], nr[1], nr[2]};
        double p[3] = {nr[3], nr[4], nr[5]};
        double em[3] = {0.0, 0.0, 0.0}, bm[3] = {0.0, 0.0, 0.0};
        if constexpr (EXT == 2) {
            // field maps at the reference-frame position in the step's lab frame
            const double zlab = r[2] + sc->fm_s0, tf = sc->fm_t;
            for (int k = 0; k < ts->nmaps; ++k)
                if (fieldmap_apply(ts->maps[k], ts->mapdata, r[0], r[1], zlab, tf, em, bm)) break;
        }
        const int64_t j = i + xstride;
        if (j < xend) {
            nr[0] = x[j];
            nr[1] = y[j];
            nr[2] = z[j];
            nr[3] = px[j];
            nr[4] = py[j];
            nr[5] = pz[j];
        }
        transform_to_fast(Mr, orb, r);  // beam frame, relative to the mesh origin
        if constexpr (DC)
            if (sc_on) r[2] = wrap_periodic(r[2], 0.0, zper);  // boundp's update() of a DC beam
        double pb[3];
        mat_vec_fast(Mr, p, pb);
        double Ef[3] = {0.0, 0.0, 0.0}, Bf[3] = {0.0, 0.0, 0.0};
        if (sc_on) {
            int l[3];
            double d[3];
            bool ok = cic_dim_fast(r[0], 0.0, ih[0], l[0], d[0]);
            ok &= cic_dim_fast(r[1], 0.0, ih[1], l[1], d[1]);
            ok &= cic_dim_fast(r[2], 0.0, ih[2], l[2], d[2]);
            if (ok) {
                gather_E_fast<DC>(Eg, l, d, n0, n1, n2, Ef);
                if constexpr (BINNED) gather_E_fast(Bg, l, d, n0, n1, n2, Bf);
            }
            if constexpr (!BINNED) {
                Bf[0] = -betaC * Ef[1];
                Bf[1] = betaC * Ef[0];
            }
        }
        if constexpr (EXT != 0)
            for (int k = 0; k < 3; ++k) {
                Ef[k] += Ee[k];
                Bf[k] += Be[k];
            }
        if constexpr (EXT == 2) {
            double t3[3];
            mat_vec_fast(Mr, em, t3);
            for (int k = 0; k < 3; ++k) Ef[k] += t3[k];
            mat_vec_fast(Mr, bm, t3);
            for (int k = 0; k < 3; ++k) Bf[k] += t3[k];
        }
        boris_kick_fast(pb, Ef, Bf, e_kick, b_kick);
        push_half_fast(r[0], r[1], r[2], pb[0], pb[1], pb[2], half_cdt);
        transform_to_fast(Mc, oc, r);
        double pn[3];
        mat_vec_fast(Mc, pb, pn);
        accumulate_moments(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
        if constexpr (NEXT) push_half_fast(r[0], r[1], r[2], pn[0], pn[1], pn[2], half_cdt);
        x[i] = r[0];
        y[i] = r[1];
        z[i] = r[2];
        px[i] = pn[0];
        py[i] = pn[1];
        pz[i] = pn[2];
    }
    store_moment_partials(m, partials, lds);
}
// The common case of the mixed-precision gather pass (no DC beam, no field
// maps; with BINNED the energy bins' summed E and B grids) in blocks that run once: block b owns kOsU*kThreads
// consecutive particles (its XCD's share of the bunch in order, as in the
// persistent loop), loads all of them before the first field gather, and
// leaves the block's plain-double moment sums in dense[b*kMom + k] (wave
// butterflies, then the four waves in order).  A stream of short-lived blocks
// keeps more particle loads in flight than a persistent grid-stride loop
// (2.30 -> 2.20 ms at 1e8 particles; 2 or 8 particles per thread: 2.27, 2.31 ms).

template <bool NEXT, int EXT, bool BINNED, bool BATCH = false, bool PERM = false>
__global__ __launch_bounds__(256) void k_track_gamma_mixed_os(int64_t N, double *x, double *y, double *z, double *px,
                                                              double *py, double *pz, const float4 *Eg,
                                                              const float4 *Bg, const Scalars *sc,
                                                              const TrackState *ts, double *dense, double *osbox,
                                                              BinRanges br, double *binp,
                                                              BatchView bv = BatchView{}, PermSrc ps = PermSrc{}) {
    __shared__ double wsum[kThreads / 64][kMom];
    __shared__ double lbox[4 * 6];
    __shared__ double lgam[2 * (kThreads / 64)];
    int bid = (int)blockIdx.x;
    if constexpr (BATCH) {
        int nblk = 0;
        const BatchItem &it = bv.items[batch_find(bv, bid, nblk)];
        N = it.N;
        x = it.x;
        y = it.y;
        z = it.z;
        px = it.px;
        py = it.py;
        pz = it.pz;
        Eg = (const float4 *)it.E;
        sc = it.sc;
        ts = it.ts;
        dense = it.dense;
    }
    const int64_t nchunks = (N + kOsChunk - 1) / kOsChunk;
    const int64_t per = (nchunks + kXcds - 1) / kXcds;
    const int64_t chunk = (int64_t)(bid % kXcds) * per + bid / kXcds;
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    double bmn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double bmx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    double g0 = 0.0, g1 = 0.0;  // BINNED with binp: the chunk's sums of bingamma.hpp
    if (chunk < nchunks) {
        const bool sc_on = ts->space_charge != 0;
        const double dT = ts->dT, charge = ts->charge, mass = ts->mass_eV;
        const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
        const double betaC = sc->betaC;
        double Mr[9], Mc[9], orb[3], oc[3], ih[3], Ee[3], Be[3];
        for (int k = 0; k < 9; ++k) {
            Mr[k] = sc->M_rb[k];
            Mc[k] = sc->M_c[k];
        }
        for (int d = 0; d < 3; ++d) {
            orb[d] = sc->o_rb2[d];
            oc[d] = sc->o_c2[d];
            ih[d] = sc->ih[d];
            Ee[d] = sc->Eext_b[d];
            Be[d] = sc->Bext_b[d];
        }
        const double e_kick = 0.5 * dT * charge * kC / mass, b_kick = 0.5 * dT * charge * kC * kC / mass;
        const double half_cdt = 0.5 * (kC * dT);
        const int64_t base = chunk * kOsChunk + threadIdx.x;
        double rr[kOsU][6];
        uint32_t pj[kOsU];
#pragma unroll
        for (int u = 0; u < kOsU; ++u) {
            const int64_t i = base + u * kThreads;
            const int64_t j = i < N ? i : N - 1;
            if constexpr (PERM) {
                pj[u] = ps.perm[j];
                rr[u][0] = ps.x[pj[u]];
                rr[u][1] = ps.y[pj[u]];
                rr[u][2] = ps.z[pj[u]];
                rr[u][3] = ps.px[pj[u]];
                rr[u][4] = ps.py[pj[u]];
                rr[u][5] = ps.pz[pj[u]];
            } else {
                // read once per step: non-temporal, so the field grid keeps the L2
                rr[u][0] = __builtin_nontemporal_load(x + j);
                rr[u][1] = __builtin_nontemporal_load(y + j);
                rr[u][2] = __builtin_nontemporal_load(z + j);
                rr[u][3] = __builtin_nontemporal_load(px + j);
                rr[u][4] = __builtin_nontemporal_load(py + j);
                rr[u][5] = __builtin_nontemporal_load(pz + j);
            }
        }
        if constexpr (PERM)
#pragma unroll
            for (int u = 0; u < kOsU; ++u) {
                const int64_t i = base + u * kThreads;
                if (i < N) {
                    ps.id_out[i] = ps.id[pj[u]];
                    if (ps.q) ps.q_out[i] = ps.q[pj[u]];
                }
            }
        int64_t split = 0;
        if constexpr (BINNED && NEXT)
            if (binp) split = bin_chunk_split(br, chunk * kOsChunk, N);
#pragma unroll
        for (int u = 0; u < kOsU; ++u) {
            const int64_t i = base + u * kThreads;
            if (i >= N) break;
            double r[3] = {rr[u][0], rr[u][1], rr[u][2]};
            double p[3] = {rr[u][3], rr[u][4], rr[u][5]};
            transform_to_fast(Mr, orb, r);
            double pb[3];
            mat_vec_fast(Mr, p, pb);
            double Ef[3] = {0.0, 0.0, 0.0}, Bf[3] = {0.0, 0.0, 0.0};
            if (sc_on) {
                int l[3];
                double d[3];
                bool ok = cic_dim_fast(r[0], 0.0, ih[0], l[0], d[0]);
                ok &= cic_dim_fast(r[1], 0.0, ih[1], l[1], d[1]);
                ok &= cic_dim_fast(r[2], 0.0, ih[2], l[2], d[2]);
                if (ok) {
                    gather_E_fast(Eg, l, d, n0, n1, n2, Ef);
                    if constexpr (BINNED) gather_E_fast(Bg, l, d, n0, n1, n2, Bf);
                }
                if constexpr (!BINNED) {
                    Bf[0] = -betaC * Ef[1];
                    Bf[1] = betaC * Ef[0];
                }
            }
            if constexpr (EXT != 0)
                for (int k = 0; k < 3; ++k) {
                    Ef[k] += Ee[k];
                    Bf[k] += Be[k];
                }
            boris_kick_fast(pb, Ef, Bf, e_kick, b_kick);
            push_half_fast(r[0], r[1], r[2], pb[0], pb[1], pb[2], half_cdt);
            transform_to_fast(Mc, oc, r);
            double pn[3];
            mat_vec_fast(Mc, pb, pn);
            if constexpr (BINNED && NEXT) {
                const double g = accumulate_moments_g(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
                if (i < split) g0 += g;
                else g1 += g;
            } else {
                accumulate_moments(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
            }
            if constexpr (NEXT) {
                push_half_fast(r[0], r[1], r[2], pn[0], pn[1], pn[2], half_cdt);
                for (int d = 0; d < 3; ++d) {
                    bmn[d] = fmin(bmn[d], r[d]);
                    bmx[d] = fmax(bmx[d], r[d]);
                }
            }
            __builtin_nontemporal_store(r[0], x + i);
            __builtin_nontemporal_store(r[1], y + i);
            __builtin_nontemporal_store(r[2], z + i);
            __builtin_nontemporal_store(pn[0], px + i);
            __builtin_nontemporal_store(pn[1], py + i);
            __builtin_nontemporal_store(pn[2], pz + i);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    static_assert(kMom == 16, "wave_sum16_split reduces 16 values");
    {
        const double t = wave_sum16_split(m, lane);  // the wave's sum of moment lane >> 2
        if ((lane & 3) == 0) wsum[wave][lane >> 2] = t;
    }
    __syncthreads();
    if (threadIdx.x < kMom) {
        double t = wsum[0][threadIdx.x];
        for (int w = 1; w < kThreads / 64; ++w) t += wsum[w][threadIdx.x];
        dense[(int64_t)bid * kMom + threadIdx.x] = t;
    }
    if constexpr (NEXT) {
        if (osbox) {
            block_minmax_split(bmn, bmx, lbox);
            if (threadIdx.x == 0 && chunk < nchunks)
                for (int d = 0; d < 3; ++d) {
                    osbox[chunk * kOsBoxStats + d] = bmn[d];
                    osbox[chunk * kOsBoxStats + 3 + d] = bmx[d];
                }
        }
    }
    if constexpr (BINNED && NEXT)
        if (binp) bin_chunk_store(g0, g1, lgam, binp, chunk, chunk < nchunks);
}

// The FP64 gather pass (no DC beam or field maps; with BINNED the energy
// bins' summed E and B grids) in blocks that run once, kOsU64 particles per thread, the per-particle sequence of
// k_track_gamma unchanged; moment sums as in k_track_gamma_mixed_os.

template <bool NEXT, bool BINNED, bool BATCH = false, bool PERM = false>
__global__ __launch_bounds__(256) void k_track_gamma_os(int64_t N, double *x, double *y, double *z, double *px,
                                                        double *py, double *pz, const double4 *Eg, const double4 *Bg,
                                                        const Scalars *sc, const TrackState *ts, double *dense,
                                                        double *osbox, BinRanges br, double *binp,
                                                        BatchView bv = BatchView{}, PermSrc ps = PermSrc{}) {
    __shared__ double wsum[kThreads / 64][kMom];
    __shared__ double lbox[4 * 6];
    __shared__ double lgam[2 * (kThreads / 64)];
    int bid = (int)blockIdx.x;
    if constexpr (BATCH) {
        int nblk = 0;
        const BatchItem &it = bv.items[batch_find(bv, bid, nblk)];
        N = it.N;
        x = it.x;
        y = it.y;
        z = it.z;
        px = it.px;
        py = it.py;
        pz = it.pz;
        Eg = (const double4 *)it.E;
        sc = it.sc;
        ts = it.ts;
        dense = it.dense;
    }
    const int64_t nchunks = (N + kOsChunk64 - 1) / kOsChunk64;
    const int64_t per = (nchunks + kXcds - 1) / kXcds;
    const int64_t chunk = (int64_t)(bid % kXcds) * per + bid / kXcds;
    double m[kMom];
    for (int k = 0; k < kMom; ++k) m[k] = 0.0;
    double bmn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double bmx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    double g0 = 0.0, g1 = 0.0;  // BINNED with binp: the chunk's sums of bingamma.hpp
    if (chunk < nchunks) {
        const bool sc_on = ts->space_charge != 0;
        const double dT = ts->dT, charge = ts->charge, mass = ts->mass_eV;
        const double o0 = sc->origin[0], o1 = sc->origin[1], o2 = sc->origin[2];
        const double h0 = sc->hr[0], h1 = sc->hr[1], h2 = sc->hr[2];
        const int n0 = sc->n[0], n1 = sc->n[1], n2 = sc->n[2];
        const double betaC = sc->betaC;
        const int64_t base = chunk * kOsChunk64 + threadIdx.x;
        double rr[kOsU64][6];
        uint32_t pj[kOsU64];
#pragma unroll
        for (int u = 0; u < kOsU64; ++u) {
            const int64_t i = base + u * kThreads;
            const int64_t j = i < N ? i : N - 1;
            if constexpr (PERM) {
                pj[u] = ps.perm[j];
                rr[u][0] = ps.x[pj[u]];
                rr[u][1] = ps.y[pj[u]];
                rr[u][2] = ps.z[pj[u]];
                rr[u][3] = ps.px[pj[u]];
                rr[u][4] = ps.py[pj[u]];
                rr[u][5] = ps.pz[pj[u]];
            } else {
                rr[u][0] = __builtin_nontemporal_load(x + j);
                rr[u][1] = __builtin_nontemporal_load(y + j);
                rr[u][2] = __builtin_nontemporal_load(z + j);
                rr[u][3] = __builtin_nontemporal_load(px + j);
                rr[u][4] = __builtin_nontemporal_load(py + j);
                rr[u][5] = __builtin_nontemporal_load(pz + j);
            }
        }
        if constexpr (PERM)
#pragma unroll
            for (int u = 0; u < kOsU64; ++u) {
                const int64_t i = base + u * kThreads;
                if (i < N) {
                    ps.id_out[i] = ps.id[pj[u]];
                    if (ps.q) ps.q_out[i] = ps.q[pj[u]];
                }
            }
        int64_t split = 0;
        if constexpr (BINNED && NEXT)
            if (binp) split = bin_chunk_split(br, chunk * kOsChunk64, N);
#pragma unroll
        for (int u = 0; u < kOsU64; ++u) {
            const int64_t i = base + u * kThreads;
            if (i >= N) break;
            double r[3] = {rr[u][0], rr[u][1], rr[u][2]};
            double p[3] = {rr[u][3], rr[u][4], rr[u][5]};
            double Ef[3] = {0.0, 0.0, 0.0}, Bf[3] = {0.0, 0.0, 0.0};
            if (sc_on) {
                transform_to(sc->M_rb, sc->o_rb, r);
                int l[3];
                double d[3];
                bool ok = cic_dim(r[0], o0, h0, l[0], d[0]);
                ok &= cic_dim(r[1], o1, h1, l[1], d[1]);
                ok &= cic_dim(r[2], o2, h2, l[2], d[2]);
                if (ok) {
                    gather_E<double>(Eg, l, d, n0, n1, n2, Ef);
                    if constexpr (BINNED) gather_E<double>(Bg, l, d, n0, n1, n2, Bf);
                }
                if constexpr (!BINNED) {
                    Bf[0] = Bf[0] - betaC * Ef[1];
                    Bf[1] = Bf[1] + betaC * Ef[0];
                }
                double t3[3];
                transform_to(sc->M_br, sc->o_br, r);
                mat_vec(sc->M_br, Ef, t3);
                Ef[0] = t3[0];
                Ef[1] = t3[1];
                Ef[2] = t3[2];
                mat_vec(sc->M_br, Bf, t3);
                Bf[0] = t3[0];
                Bf[1] = t3[1];
                Bf[2] = t3[2];
            }
            for (int k = 0; k < 3; ++k) {
                Ef[k] += sc->Eext[k];
                Bf[k] += sc->Bext[k];
            }
            double pn[3];
            boris_kick(p, Ef, Bf, dT, mass, charge);
            push_half(r[0], r[1], r[2], p[0], p[1], p[2], dT);
            transform_to(sc->M_u, sc->o_u, r);
            mat_vec(sc->M_u, p, pn);
            if constexpr (BINNED && NEXT) {
                const double g = accumulate_moments_g(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
                if (i < split) g0 += g;
                else g1 += g;
            } else {
                accumulate_moments(m, r[0], r[1], r[2], pn[0], pn[1], pn[2]);
            }
            if constexpr (NEXT) {
                push_half(r[0], r[1], r[2], pn[0], pn[1], pn[2], dT);
                for (int d = 0; d < 3; ++d) {
                    bmn[d] = fmin(bmn[d], r[d]);
                    bmx[d] = fmax(bmx[d], r[d]);
                }
            }
            __builtin_nontemporal_store(r[0], x + i);
            __builtin_nontemporal_store(r[1], y + i);
            __builtin_nontemporal_store(r[2], z + i);
            __builtin_nontemporal_store(pn[0], px + i);
            __builtin_nontemporal_store(pn[1], py + i);
            __builtin_nontemporal_store(pn[2], pz + i);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    static_assert(kMom == 16, "wave_sum16_split reduces 16 values");
    {
        const double t = wave_sum16_split(m, lane);  // the wave's sum of moment lane >> 2
        if ((lane & 3) == 0) wsum[wave][lane >> 2] = t;
    }
    __syncthreads();
    if (threadIdx.x < kMom) {
        double t = wsum[0][threadIdx.x];
        for (int w = 1; w < kThreads / 64; ++w) t += wsum[w][threadIdx.x];
        dense[(int64_t)bid * kMom + threadIdx.x] = t;
    }
    if constexpr (NEXT) {
        if (osbox) {
            block_minmax_split(bmn, bmx, lbox);
            if (threadIdx.x == 0 && chunk < nchunks)
                for (int d = 0; d < 3; ++d) {
                    osbox[chunk * kOsBoxStats + d] = bmn[d];
                    osbox[chunk * kOsBoxStats + 3 + d] = bmx[d];
                }
        }
    }
    if constexpr (BINNED && NEXT)
        if (binp) bin_chunk_store(g0, g1, lgam, binp, chunk, chunk < nchunks);
}

// Dense per-block moment sums -> min(rows, kOsReduceBlocks) double-double
// partials (rows in a fixed order per block), the layout k_track_post reads.
// With at most kOsReduceBlocks rows each partial is one row, {v + 0, 0}: what
// the block sum of one value and zeros gives (adding an exact zero to a
// double-double is the identity, up to the sign of a zero), without the block
// reduction.
constexpr int kOsReduceBlocks = 256;
__host__ __device__ inline int os_reduce_blocks(int64_t rows) {
    return (int)(rows < kOsReduceBlocks ? (rows < 1 ? 1 : rows) : kOsReduceBlocks);
}
template <bool BATCH = false, bool FP64 = false>
__global__ __launch_bounds__(256) void k_reduce_dense(const double *dense, int64_t rows, double *partials,
                                                      BatchView bv = BatchView{}) {
    __shared__ double lds[4 * kMom * 2];
    double h[kMom], l[kMom];
    for (int k = 0; k < kMom; ++k) h[k] = l[k] = 0.0;
    int bid = (int)blockIdx.x, nblk = (int)gridDim.x;
    if constexpr (BATCH) {
        const BatchItem &it = bv.items[batch_find(bv, bid, nblk)];
        dense = it.dense;
        partials = it.partials;
        rows = FP64 ? os_blocks64(it.N) : os_blocks(it.N);
    }
    const int64_t per = (rows + nblk - 1) / nblk;
    const int64_t r0 = bid * per, r1 = r0 + per < rows ? r0 + per : rows;
    if (per == 1) {
        if (threadIdx.x < kMom) {
            const double v = r0 < rows ? dense[r0 * kMom + threadIdx.x] : 0.0;
            partials[bid * kSlots + 2 * threadIdx.x] = v + 0.0;
            partials[bid * kSlots + 2 * threadIdx.x + 1] = 0.0;
        }
        return;
    }
    if (per <= 64) {
        // every row in the first wave: the other waves' sums are exact zeros, so
        // the block sum is the first wave's butterfly followed by three
        // additions of a zero (kept for the sign of a zero sum)
        static_assert(kThreads == 256, "four waves per block");
        if (threadIdx.x >= 64) return;
        if (r0 + threadIdx.x < r1)
            for (int k = 0; k < kMom; ++k) h[k] = dense[(r0 + threadIdx.x) * kMom + k] + 0.0;
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int k = 0; k < kMom; ++k) {
                const double oh = __shfl_xor(h[k], off, 64), ol = __shfl_xor(l[k], off, 64);
                const dd r = dd_add(dd{h[k], l[k]}, dd{oh, ol});
                h[k] = r.hi;
                l[k] = r.lo;
            }
        if (threadIdx.x == 0)
            for (int k = 0; k < kMom; ++k) {
                dd a = {h[k], l[k]};
                for (int w = 1; w < kThreads / 64; ++w) a = dd_add(a, dd{0.0, 0.0});
                partials[bid * kSlots + 2 * k] = a.hi;
                partials[bid * kSlots + 2 * k + 1] = a.lo;
            }
        return;
    }
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x)
        for (int k = 0; k < kMom; ++k) {
            dd s = dd_add(dd{h[k], l[k]}, dd{dense[r * kMom + k], 0.0});
            h[k] = s.hi;
            l[k] = s.lo;
        }
    block_sum_dd<kMom>(h, l, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < kMom; ++k) {
            partials[bid * kSlots + 2 * k] = h[k];
            partials[bid * kSlots + 2 * k + 1] = l[k];
        }
}

namespace {
struct OsArgs {
    hipStream_t s;
    Bunch *b;
    const void *Eg, *Bg;
    const Scalars *sc;
    const TrackState *ts;
    double *dense, *osbox;
    const BinRanges *br;
    double *binp;
    const PermSrc *perm;
};
template <bool NEXT, bool BINNED, bool PERM>
void launch_os64_p(const OsArgs &a, int blocks) {
    k_track_gamma_os<NEXT, BINNED, false, PERM><<<blocks, kThreads, 0, a.s>>>(
        a.b->N, a.b->x, a.b->y, a.b->z, a.b->px, a.b->py, a.b->pz, (const double4 *)a.Eg, (const double4 *)a.Bg, a.sc,
        a.ts, a.dense, NEXT ? a.osbox : nullptr, *a.br, BINNED && NEXT ? a.binp : nullptr, BatchView{},
        PERM ? *a.perm : PermSrc{});
}
template <bool NEXT, bool BINNED>
void launch_os64(const OsArgs &a, int blocks) {
    if (a.perm) launch_os64_p<NEXT, BINNED, true>(a, blocks);
    else launch_os64_p<NEXT, BINNED, false>(a, blocks);
}
template <bool NEXT, int EXT, bool BINNED, bool PERM>
void launch_os_mixed_p(const OsArgs &a, int blocks) {
    k_track_gamma_mixed_os<NEXT, EXT, BINNED, false, PERM><<<blocks, kThreads, 0, a.s>>>(
        a.b->N, a.b->x, a.b->y, a.b->z, a.b->px, a.b->py, a.b->pz, (const float4 *)a.Eg, (const float4 *)a.Bg, a.sc,
        a.ts, a.dense, NEXT ? a.osbox : nullptr, *a.br, BINNED && NEXT ? a.binp : nullptr, BatchView{},
        PERM ? *a.perm : PermSrc{});
}
template <bool NEXT, int EXT, bool BINNED>
void launch_os_mixed(const OsArgs &a, int blocks) {
    if (a.perm) launch_os_mixed_p<NEXT, EXT, BINNED, true>(a, blocks);
    else launch_os_mixed_p<NEXT, EXT, BINNED, false>(a, blocks);
}
template <bool NEXT, int EXT>
void launch_os_mixed_b(const OsArgs &a, int blocks) {
    if (a.Bg) launch_os_mixed<NEXT, EXT, true>(a, blocks);
    else launch_os_mixed<NEXT, EXT, false>(a, blocks);
}
}  // namespace

int launch_track_gamma_os(hipStream_t s, bool fp64, bool next, Bunch &b, const void *Eg, const Scalars *sc,
                          const TrackState *ts, double *dense, double *partials, int ext, double *osbox,
                          const void *Bg, const BinRanges *bins, double *binp, const PermSrc *perm) {
    static const BinRanges none;
    const OsArgs a{s, &b, Eg, Bg, sc, ts, dense, osbox, bins ? bins : &none, bins ? binp : nullptr, perm};
    if (fp64) {
        const int blocks64 = os_blocks64(b.N);
        if (next) {
            if (Bg) launch_os64<true, true>(a, blocks64);
            else launch_os64<true, false>(a, blocks64);
        } else {
            if (Bg) launch_os64<false, true>(a, blocks64);
            else launch_os64<false, false>(a, blocks64);
        }
        const int nr = os_reduce_blocks(blocks64);
        k_reduce_dense<<<nr, kThreads, 0, s>>>(dense, blocks64, partials);
        return nr;
    }
    const int blocks = os_blocks(b.N);
    if (next) {
        if (ext) launch_os_mixed_b<true, 1>(a, blocks);
        else launch_os_mixed_b<true, 0>(a, blocks);
    } else {
        if (ext) launch_os_mixed_b<false, 1>(a, blocks);
        else launch_os_mixed_b<false, 0>(a, blocks);
    }
    const int nr = os_reduce_blocks(blocks);
    k_reduce_dense<<<nr, kThreads, 0, s>>>(dense, blocks, partials);
    return nr;
}

// Blocks of `kernel` (kThreads each) that the device holds at once, per device.
template <typename K>
static int resident_blocks(K kernel) {
    static int cache[64] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (cache[dev] == 0) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0) != hipSuccess || per <= 0)
            per = 1;
        cache[dev] = cus * per;
    }
    return cache[dev];
}

// One wave of resident blocks (fewer partial sums for the step statistics).
int gamma_blocks(bool fp64, int64_t N) {
    const int r = fp64 ? resident_blocks(k_track_gamma<true, false, false>)
                       : resident_blocks(k_track_gamma_mixed<true, false, 1, false>);
    const int b = red_blocks(N);
    return b < r ? b : r;
}

namespace {
struct GammaArgs {
    unsigned blocks;
    hipStream_t s;
    Bunch *b;
    const void *Eg, *Bg;
    const Scalars *sc;
    const TrackState *ts;
    double *partials;
};
template <bool NEXT, bool BINNED, int EXT, bool DC>
void launch_mixed(const GammaArgs &a) {
    k_track_gamma_mixed<NEXT, BINNED, EXT, DC><<<a.blocks, kThreads, 0, a.s>>>(
        a.b->N, a.b->x, a.b->y, a.b->z, a.b->px, a.b->py, a.b->pz, (const float4 *)a.Eg, (const float4 *)a.Bg, a.sc,
        a.ts, a.partials);
}
template <bool NEXT, int EXT>
void launch_mixed_nb(const GammaArgs &a, bool dc) {
    if (a.Bg) launch_mixed<NEXT, true, EXT, false>(a);  // energy bins never come with a DC beam
    else if (dc) launch_mixed<NEXT, false, EXT, true>(a);
    else launch_mixed<NEXT, false, EXT, false>(a);
}
template <bool NEXT, bool MAPS>
void launch_fp64(const GammaArgs &a) {
    if (a.Bg)
        k_track_gamma<NEXT, true, MAPS><<<a.blocks, kThreads, 0, a.s>>>(
            a.b->N, a.b->x, a.b->y, a.b->z, a.b->px, a.b->py, a.b->pz, (const double4 *)a.Eg, (const double4 *)a.Bg,
            a.sc, a.ts, a.partials);
    else
        k_track_gamma<NEXT, false, MAPS><<<a.blocks, kThreads, 0, a.s>>>(
            a.b->N, a.b->x, a.b->y, a.b->z, a.b->px, a.b->py, a.b->pz, (const double4 *)a.Eg, nullptr, a.sc, a.ts,
            a.partials);
}
}  // namespace

// ext: 0 no external field, 1 the analytic fields, 2 those and field maps;
// dc: a DC beam (periodic z)
void launch_track_gamma(hipStream_t s, bool fp64, bool next, Bunch &b, const void *Eg, const Scalars *sc,
                        const TrackState *ts, double *partials, const void *Bg, int ext, bool dc) {
    if (b.N <= 0) return;
    const GammaArgs a{(unsigned)gamma_blocks(fp64, b.N), s, &b, Eg, Bg, sc, ts, partials};
    if (fp64) {
        if (ext == 2) next ? launch_fp64<true, true>(a) : launch_fp64<false, true>(a);
        else next ? launch_fp64<true, false>(a) : launch_fp64<false, false>(a);
    } else if (ext == 2) {
        next ? launch_mixed_nb<true, 2>(a, dc) : launch_mixed_nb<false, 2>(a, dc);
    } else if (ext == 1) {
        next ? launch_mixed_nb<true, 1>(a, dc) : launch_mixed_nb<false, 1>(a, dc);
    } else {
        next ? launch_mixed_nb<true, 0>(a, dc) : launch_mixed_nb<false, 0>(a, dc);
    }
}

// Step statistics (DataSink / calcBeamParameters) and the momentum sums the
// next step's rotation and Lorentz factor need.
template <bool BATCH = false>
__global__ __launch_bounds__(512) void k_track_post(const double *partials, int nb, TrackState *ts,
                                                    Scalars *sc, StatsOut *out,
                                                    const unsigned long long *dropped, BatchView bv = BatchView{}) {
    __shared__ double lds[16 * kMom * 2];
    if constexpr (BATCH) {
        // nb: 1 for the FP64 gather pass's rows, 0 for the mixed one's
        const BatchItem &it = bv.items[blockIdx.x];
        partials = it.partials;
        nb = os_reduce_blocks(nb ? os_blocks64(it.N) : os_blocks(it.N));
        ts = it.ts;
        sc = it.sc;
        out = (StatsOut *)it.stats;
        dropped = it.dropped;
    }
    double m[kMom];
    reduce_partials_dd<kMom>(partials, nb, m, lds);
    if (threadIdx.x != 0) return;
    sc->psum[0] = m[1];
    sc->psum[1] = m[3];
    sc->psum[2] = m[5];
    if (out) {
        StatsOut s;
        beam_parameters(sc->N, m, ts->mass_eV, s);
        s.t = ts->t;
        s.pathLength = ts->pathLength;
        s.dropped = dropped ? (int64_t)*dropped : 0;
        out[ts->step - ts->call_step0 - 1] = s;  // the step's slot, so that a captured step replays unchanged
    }
}
__global__ void k_track_begin_call(TrackState *ts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ts->call_step0 = ts->step;
}
void launch_track_begin_call(hipStream_t s, TrackState *ts) { k_track_begin_call<<<1, 64, 0, s>>>(ts); }
void launch_track_post(hipStream_t s, const double *partials, int nb, TrackState *ts, Scalars *sc,
                       void *stats_out, const unsigned long long *dropped) {
    k_track_post<<<1, kReduceThreads, 0, s>>>(partials, nb, ts, sc, (StatsOut *)stats_out, dropped);
}
void launch_track_init_moments(hipStream_t s, const double *partials, int nb, TrackState *ts,
                               Scalars *sc) {
    k_track_post<<<1, kReduceThreads, 0, s>>>(partials, nb, ts, sc, nullptr, nullptr);
}


// ---- batched steps of many bunches (one launch per pass and kernel class) -----
// Each kernel is the single-bunch kernel with its arguments taken from the
// bunch's BatchItem and its block index relative to the bunch's first block,
// so a batched step computes what single steps of each bunch compute, bit for
// bit.

// DIR 0: the call starts (call_step0 := step, the context's scalars into the
// batch's array); 1: it ends (the batch's scalars back to the context)
template <int DIR>
__global__ void k_batch_scalars(BatchView bv) {
    const BatchItem &it = bv.items[blockIdx.x];
    constexpr int W = (int)(sizeof(Scalars) / sizeof(double));
    static_assert(sizeof(Scalars) % sizeof(double) == 0, "Scalars copies as doubles");
    const double *src = (const double *)(DIR == 0 ? it.sc_ctx : it.sc);
    double *dst = (double *)(DIR == 0 ? it.sc : it.sc_ctx);
    for (int k = threadIdx.x; k < W; k += blockDim.x) dst[k] = src[k];
    if (DIR == 0 && threadIdx.x == 0) it.ts->call_step0 = it.ts->step;
}
void launch_batch_begin(hipStream_t s, const BatchView &v) {
    if (v.n > 0) k_batch_scalars<0><<<v.n, 64, 0, s>>>(v);
}
void launch_batch_end(hipStream_t s, const BatchView &v) {
    if (v.n > 0) k_batch_scalars<1><<<v.n, 64, 0, s>>>(v);
}
void launch_batch_pre(hipStream_t s, const BatchView &v) {
    if (v.n > 0) k_track_pre<true><<<v.n, 64, 0, s>>>(nullptr, nullptr, 0, nullptr, v);
}
static unsigned batch_blocks(const BatchView &v) { return v.n > 0 ? (unsigned)v.blocks : 0u; }
void launch_batch_alpha(hipStream_t s, const BatchView &v, bool exact, bool push) {
    const unsigned nb = batch_blocks(v);
    if (nb == 0) return;
#define OPALSC_BALPHA(E_, P_)                                                                                     \
    k_track_alpha<E_, P_, true><<<nb, kThreads, 0, s>>>(0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, \
                                                        nullptr, nullptr, nullptr, nullptr, v)
    if (exact) {
        if (push) OPALSC_BALPHA(true, true);
        else OPALSC_BALPHA(true, false);
    } else {
        if (push) OPALSC_BALPHA(false, true);
        else OPALSC_BALPHA(false, false);
    }
#undef OPALSC_BALPHA
}
void launch_batch_mesh(hipStream_t s, const BatchView &v, bool exact) {
    if (v.n > 0)
        k_track_mesh<true><<<v.n, kReduceThreads, 0, s>>>(nullptr, 0, 0, 0, 0, 0.0, nullptr, exact ? 1 : 0, 1,
                                                          nullptr, v);
}
void launch_batch_deposit(hipStream_t s, const BatchView &v, bool exact, bool quni, bool large) {
    const unsigned nb = batch_blocks(v);
    if (nb == 0) return;
#define OPALSC_BDEP(M_, E_, B_)                                                                            \
    do {                                                                                                   \
        if (large)                                                                                         \
            k_deposit_lds<M_, E_, B_, kBoxCapLarge, true><<<nb, kThreads, 0, s>>>(                          \
                0, 0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, v); \
        else                                                                                               \
            k_deposit_lds<M_, E_, B_, kBoxCapSmall, true><<<nb, kThreads, 0, s>>>(                          \
                0, 0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, v); \
    } while (0)
    // BOX instances: a bunch whose chunk boxes are stale (re-sorted in this
    // step) has a null chunkbox in its item and forms its own
    if (quni) {
        if (exact) OPALSC_BDEP(3, true, true);
        else OPALSC_BDEP(3, false, true);
    } else {
        if (exact) OPALSC_BDEP(1, true, true);
        else OPALSC_BDEP(1, false, true);
    }
#undef OPALSC_BDEP
}
template <bool NEXT>
static void batch_gather(hipStream_t s, const BatchView &v, unsigned nb, bool exact, bool ext) {
    static const BinRanges none;
    if (exact)
        k_track_gamma_os<NEXT, false, true><<<nb, kThreads, 0, s>>>(0, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                     nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                     nullptr, nullptr, none, nullptr, v);
    else if (ext)
        k_track_gamma_mixed_os<NEXT, 1, false, true><<<nb, kThreads, 0, s>>>(
            0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, none, nullptr, v);
    else
        k_track_gamma_mixed_os<NEXT, 0, false, true><<<nb, kThreads, 0, s>>>(
            0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, none, nullptr, v);
}
void launch_batch_gather(hipStream_t s, const BatchView &v, bool exact, bool ext, bool next) {
    const unsigned nb = batch_blocks(v);
    if (nb == 0) return;
    if (next) batch_gather<true>(s, v, nb, exact, ext);
    else batch_gather<false>(s, v, nb, exact, ext);
}
int batch_reduce_blocks(bool exact, int64_t N) { return os_reduce_blocks(exact ? os_blocks64(N) : os_blocks(N)); }
void launch_batch_reduce(hipStream_t s, const BatchView &v, bool exact) {
    const unsigned nb = batch_blocks(v);
    if (nb == 0) return;
    if (exact) k_reduce_dense<true, true><<<nb, kThreads, 0, s>>>(nullptr, 0, nullptr, v);
    else k_reduce_dense<true, false><<<nb, kThreads, 0, s>>>(nullptr, 0, nullptr, v);
}
void launch_batch_post(hipStream_t s, const BatchView &v, bool exact) {
    if (v.n > 0)
        k_track_post<true><<<v.n, kReduceThreads, 0, s>>>(nullptr, exact ? 1 : 0, nullptr, nullptr, nullptr, nullptr,
                                                           v);
}

}  // namespace opalsc

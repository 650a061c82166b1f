// Stable key/value radix sort of (cell key, particle index) pairs.  The
// deposit and gather kernels are correct in any particle order; sorting the
// bunch by cell gives them cache-resident grid neighbourhoods.  An LSD radix
// sort is stable, so equal keys keep their particle order and the
// permutation is unique.
//
// Each pass sorts by one digit of at most 9 bits (512 buckets):
//  1. k_rs_hist: per block of 2048 keys, the digit histogram (LDS), stored
//     digit-major, hist[digit * nblocks + block].
//  2. k_scan_*: the exclusive prefix sum of hist in that order, i.e. for every
//     (digit, block) the output position of the block's first key with that
//     digit (all smaller digits first, then the same digit in earlier blocks).
//  3. k_rs_scatter: the block ranks its keys stably: a wave takes its keys 64
//     at a time in input order, finds the lanes that share its digit with one
//     ballot per digit bit, and counts the earlier lanes and the earlier keys of
//     the wave with that digit (a per-wave counter in LDS updated by the group's
//     first lane); waves then follow each other in input order.  The block
//     stages its keys in LDS in (digit, rank) order and writes them out in runs
//     that are contiguous in the output.
// Pass p reads its input from the previous pass's output; the caller's input
// arrays are only read, and the last pass writes the caller's output arrays.
#include "kernels.hpp"

namespace opalsc {
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsIPT = 8;                     // keys per thread
constexpr int kRsTile = kRsThreads * kRsIPT;  // keys per block
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsMaxBits = 9;
constexpr int kRsBuckets = 1 << kRsMaxBits;
constexpr int kScanIPT = 8;
constexpr int kScanSeg = kRsThreads * kScanIPT;  // entries per block of the prefix sum
static_assert(kRsIPT == 8, "the histogram loads eight keys per thread as two 16-byte words");

int64_t rs_blocks(int64_t N) { return (N + kRsTile - 1) / kRsTile; }

// Exclusive prefix sum over the block of v (one value per thread); returns the
// block total in every thread.  lds: kRsWaves words.
__device__ inline uint32_t block_exclusive_scan(uint32_t &v, uint32_t *lds) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < kRsWaves; ++w) {
        const uint32_t t = lds[w];
        before += w < wave ? t : 0u;
        total += t;
    }
    __syncthreads();
    v = before + inc - v;
    return total;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t *__restrict__ keys, int64_t N, int shift,
                                                        int bits, uint32_t *__restrict__ hist, int64_t nblocks) {
    __shared__ uint32_t h[kRsBuckets];
    const int nb = 1 << bits;
    const uint32_t mask = (uint32_t)nb - 1u;
    for (int d = threadIdx.x; d < nb; d += kRsThreads) h[d] = 0;
    __syncthreads();
    // thread t counts the kRsIPT consecutive keys from base: a run of equal
    // digits (the common case in the high digits of a bunch sorted a few steps
    // ago) is one LDS atomic instead of one per key
    const int64_t base = (int64_t)blockIdx.x * kRsTile + (int64_t)threadIdx.x * kRsIPT;
    uint32_t kk[kRsIPT];
    if (base + kRsIPT <= N) {
        const uint4 a = ((const uint4 *)(keys + base))[0], b = ((const uint4 *)(keys + base))[1];
        kk[0] = a.x; kk[1] = a.y; kk[2] = a.z; kk[3] = a.w;
        kk[4] = b.x; kk[5] = b.y; kk[6] = b.z; kk[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kRsIPT; ++k) kk[k] = base + k < N ? keys[base + k] : 0u;
    }
    const int cnt = base >= N ? 0 : (N - base < kRsIPT ? (int)(N - base) : kRsIPT);
    uint32_t run_d = 0, run_n = 0;
#pragma unroll
    for (int k = 0; k < kRsIPT; ++k) {
        if (k < cnt) {
            const uint32_t d = (kk[k] >> shift) & mask;
            if (run_n && d != run_d) {
                atomicAdd(&h[run_d], run_n);
                run_n = 0;
            }
            run_d = d;
            ++run_n;
        }
    }
    if (run_n) atomicAdd(&h[run_d], run_n);
    __syncthreads();
    for (int d = threadIdx.x; d < nb; d += kRsThreads) hist[(int64_t)d * nblocks + blockIdx.x] = h[d];
}

// Exclusive prefix sum of a[0, L) in place: segment sums, their prefix sum (one
// block), then each segment's own prefix sum plus its offset.
__global__ __launch_bounds__(kRsThreads) void k_scan_seg_sums(const uint32_t *__restrict__ a, int64_t L,
                                                              uint32_t *__restrict__ seg) {
    __shared__ uint32_t lds[kRsWaves];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * kScanIPT;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanIPT; ++k)
        if (base + k < L) s += a[base + k];
    const uint32_t total = block_exclusive_scan(s, lds);
    if (threadIdx.x == 0) seg[blockIdx.x] = total;
}

__global__ __launch_bounds__(kRsThreads) void k_scan_top(uint32_t *seg, int64_t nseg) {
    __shared__ uint32_t lds[kRsWaves];
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nseg; b0 += kScanSeg) {
        const int64_t base = b0 + (int64_t)threadIdx.x * kScanIPT;
        uint32_t v[kScanIPT], s = 0;
#pragma unroll
        for (int k = 0; k < kScanIPT; ++k) {
            v[k] = base + k < nseg ? seg[base + k] : 0u;
            s += v[k];
        }
        const uint32_t total = block_exclusive_scan(s, lds);
        uint32_t run = carry + s;
#pragma unroll
        for (int k = 0; k < kScanIPT; ++k)
            if (base + k < nseg) {
                seg[base + k] = run;
                run += v[k];
            }
        carry += total;
    }
}

__global__ __launch_bounds__(kRsThreads) void k_scan_apply(uint32_t *__restrict__ a, int64_t L,
                                                           const uint32_t *__restrict__ seg) {
    __shared__ uint32_t lds[kRsWaves];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * kScanIPT;
    uint32_t v[kScanIPT], s = 0;
#pragma unroll
    for (int k = 0; k < kScanIPT; ++k) {
        v[k] = base + k < L ? a[base + k] : 0u;
        s += v[k];
    }
    (void)block_exclusive_scan(s, lds);
    uint32_t run = seg[blockIdx.x] + s;
#pragma unroll
    for (int k = 0; k < kScanIPT; ++k)
        if (base + k < L) {
            a[base + k] = run;
            run += v[k];
        }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                           int64_t N, int shift, int bits,
                                                           const uint32_t *__restrict__ offs, int64_t nblocks) {
    __shared__ uint32_t whist[kRsWaves][kRsBuckets];
    __shared__ uint32_t bstart[kRsBuckets];
    __shared__ uint32_t sk[kRsTile], sv[kRsTile];
    __shared__ uint32_t lds[kRsWaves];
    const int nb = 1 << bits;
    const uint32_t mask = (uint32_t)nb - 1u;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int e = threadIdx.x; e < kRsWaves * kRsBuckets; e += kRsThreads) (&whist[0][0])[e] = 0;
    __syncthreads();
    const int64_t tile0 = (int64_t)blockIdx.x * kRsTile;
    // wave w takes keys tile0 + w*64*IPT + k*64 + lane: (w, k, lane) is input order
    const int64_t wbase = tile0 + (int64_t)wave * (64 * kRsIPT) + lane;
    uint32_t key[kRsIPT], val[kRsIPT], rank[kRsIPT];
#pragma unroll
    for (int k = 0; k < kRsIPT; ++k) {
        const int64_t i = wbase + k * 64;
        const bool ok = i < N;
        key[k] = ok ? kin[i] : 0u;
        val[k] = ok ? vin[i] : 0u;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kRsIPT; ++k) {
        const bool ok = wbase + k * 64 < N;
        const uint32_t d = (key[k] >> shift) & mask;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < kRsMaxBits; ++b) {
            if (b < bits) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bb = __ballot(ok && bit);
                peers &= bit ? bb : ~bb;
            }
        }
        // the keys of this wave with digit d taken before (earlier k), then the
        // earlier lanes of this k; the group's first lane advances the counter
        const uint32_t before = ok ? whist[wave][d] : 0u;
        rank[k] = before + (uint32_t)__popcll(peers & below);
        if (ok && lane == __ffsll((long long)peers) - 1) whist[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit: the waves' offsets in order, and the block's count
    for (int d = threadIdx.x; d < kRsBuckets; d += kRsThreads) {
        uint32_t run = 0;
        for (int w = 0; w < kRsWaves; ++w) {
            const uint32_t c = whist[w][d];
            whist[w][d] = run;
            run += c;
        }
        bstart[d] = run;
    }
    __syncthreads();
    // the block's first position of every digit (exclusive sum over digits, two per thread)
    {
        static_assert(kRsBuckets == 2 * kRsThreads, "two buckets per thread");
        const uint32_t a = bstart[2 * threadIdx.x], b = bstart[2 * threadIdx.x + 1];
        uint32_t s = a + b;
        (void)block_exclusive_scan(s, lds);
        bstart[2 * threadIdx.x] = s;
        bstart[2 * threadIdx.x + 1] = s + a;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsIPT; ++k) {
        if (wbase + k * 64 < N) {
            const uint32_t d = (key[k] >> shift) & mask;
            const uint32_t r = bstart[d] + whist[wave][d] + rank[k];
            if (r < (uint32_t)kRsTile) {  // always: the ranks are a permutation of the tile
                sk[r] = key[k];
                sv[r] = val[k];
            }
        }
    }
    __syncthreads();
    const int cnt = (int)(N - tile0 < kRsTile ? N - tile0 : kRsTile);
    for (int r = threadIdx.x; r < cnt; r += kRsThreads) {
        const uint32_t kk = sk[r];
        const uint32_t d = (kk >> shift) & mask;
        const uint32_t dest = offs[(int64_t)d * nblocks + blockIdx.x] + ((uint32_t)r - bstart[d]);
        if (dest < (uint64_t)N) {  // always: the offsets are an exclusive prefix sum of the counts
            kout[dest] = kk;
            vout[dest] = sv[r];
        }
    }
}

}  // namespace

// Scratch: a key/value buffer pair of N words each for the passes between
// input and output, the (digit, block) offsets and the segment sums.
size_t sort_temp_bytes(int64_t N) {
    if (N <= 0) return 0;
    const int64_t L = (int64_t)kRsBuckets * rs_blocks(N);
    const int64_t nseg = (L + kScanSeg - 1) / kScanSeg;
    const size_t n4 = (size_t)(N + 3) / 4 * 4;  // whole 16-byte words per key/value buffer
    return n4 * 8 + (size_t)L * 4 + (size_t)(nseg + 1) * 4 + 256;
}

int sort_pairs(hipStream_t s, void *tmp, size_t tmp_bytes, const uint32_t *kin, uint32_t *kout,
               const uint32_t *vin, uint32_t *vout, int64_t N, int end_bit) {
    if (N <= 0) return 0;
    if (N > (int64_t)0xFFFFFFF0 || end_bit > 32 || tmp_bytes < sort_temp_bytes(N)) return -1;
    if (end_bit <= 0) {
        if (hipMemcpyAsync(kout, kin, (size_t)N * 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(vout, vin, (size_t)N * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return -1;
        return 0;
    }
    const int64_t nblocks = rs_blocks(N);
    const int passes = (end_bit + kRsMaxBits - 1) / kRsMaxBits;
    const int width = (end_bit + passes - 1) / passes;
    const int64_t n4 = (N + 3) / 4 * 4;  // keeps every buffer 16-byte aligned
    uint32_t *tk = (uint32_t *)tmp, *tv = tk + n4;
    uint32_t *hist = tv + n4;
    // (input, output) of each pass: A -> B, B -> T, T -> B, ... for an odd
    // number of passes, A -> T, T -> B, B -> T, ... for an even one
    const uint32_t *sk = kin, *sv = vin;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * width;
        const int bits = end_bit - shift < width ? end_bit - shift : width;
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t *dk = to_out ? kout : tk, *dv = to_out ? vout : tv;
        const int64_t L = (int64_t)(1 << bits) * nblocks;
        const int64_t nseg = (L + kScanSeg - 1) / kScanSeg;
        uint32_t *seg = hist + L;
        k_rs_hist<<<(unsigned)nblocks, kRsThreads, 0, s>>>(sk, N, shift, bits, hist, nblocks);
        k_scan_seg_sums<<<(unsigned)nseg, kRsThreads, 0, s>>>(hist, L, seg);
        k_scan_top<<<1, kRsThreads, 0, s>>>(seg, nseg);
        k_scan_apply<<<(unsigned)nseg, kRsThreads, 0, s>>>(hist, L, seg);
        k_rs_scatter<<<(unsigned)nblocks, kRsThreads, 0, s>>>(sk, sv, dk, dv, N, shift, bits, hist, nblocks);
        sk = dk;
        sv = dv;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace opalsc

// opalsc_beam_parameters_noalloc: calcBeamParameters of device arrays with
// no device allocation, for callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The moment reduction's only scratch is its per-block partial sums (kSlots
// doubles per block, at most kRedBlocksMax blocks: the streaming kernels
// grid-stride over any N).  Here they live in a file-scope device array of
// kSlots * kRedBlocksMax doubles (4 MiB) instead of the context's buffer.
// Each GPU holds its own copy of the array from the first call on it until
// the process ends; its address on a device is looked up once per device with
// hipGetSymbolAddress and only that address reaches kernels.  Every call on a
// GPU shares that GPU's array: each reduction block writes its own row before
// the final block reads the rows, so nothing needs clearing, and calls that
// overlap on one GPU (two streams, two threads) race on it.
#include "../../include/opalsc.h"
#include "kernels.hpp"

#include <mutex>
#include <string>

using namespace opalsc;

namespace {

constexpr size_t kScratchDoubles = (size_t)kSlots * kRedBlocksMax;  // 4 MiB
static_assert(kScratchDoubles * sizeof(double) == (size_t)4 << 20, "stated scratch size");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {nullptr};

// The device address of this device's copy of `scratch` (looked up once per device).
int scratch_of_current_device(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_global_error("opalsc_beam_parameters_noalloc: no current HIP device (or ordinal >= 64)");
        return OPALSC_EHIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_global_error(std::string("opalsc_beam_parameters_noalloc: hipGetSymbolAddress(scratch): ") +
                             hipGetErrorString(e));
            return OPALSC_EHIP;
        }
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return OPALSC_OK;
}

}  // namespace

extern "C" int opalsc_beam_parameters_noalloc(int64_t N, const double *R_dev, const double *P_dev,
                                              double mass_eV, opalsc_stats *out_dev, void *stream) {
    if (N < 0 || N > (int64_t)0xFFFFFFF0) {
        set_global_error("opalsc_beam_parameters_noalloc: N out of range (0 .. 2^32-16)");
        return OPALSC_EINVAL;
    }
    // the partial sums of red_blocks(N) <= kRedBlocksMax blocks, one kSlots row each
    const size_t rows = N > 0 ? (size_t)red_blocks(N) : 0;
    if (rows * kSlots > kScratchDoubles) {
        set_global_error("opalsc_beam_parameters_noalloc: " + std::to_string(rows) +
                         " partial-sum rows exceed the 4 MiB scratch array");
        return OPALSC_EINVAL;
    }
    if (!device_range_ok(out_dev, sizeof(opalsc_stats))) {
        set_global_error("opalsc_beam_parameters_noalloc: out_dev is not device memory of one opalsc_stats");
        return OPALSC_EINVAL;
    }
    const size_t v = (size_t)N * 3 * sizeof(double);
    if (N > 0 && (!device_range_ok(R_dev, v) || !device_range_ok(P_dev, v))) {
        set_global_error("opalsc_beam_parameters_noalloc: R_dev and P_dev must be device memory of N x 3 doubles");
        return OPALSC_EINVAL;
    }
    double *partials = nullptr;
    if (int e = scratch_of_current_device(&partials)) return e;
    hipStream_t s = (hipStream_t)stream;
    static_assert(sizeof(opalsc_stats) % 8 == 0, "stats cleared as 64-bit words");
    if (N == 0) {
        // an empty bunch has all-zero statistics (a fill kernel, not a memset node)
        launch_fill_i64(s, sizeof(opalsc_stats) / 8, (long long *)out_dev);
    } else {
        launch_moments(s, N, R_dev, P_dev, partials);
        launch_stats_from_partials(s, partials, (int)rows, N, mass_eV, out_dev);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_global_error(std::string("opalsc_beam_parameters_noalloc: launch failed: ") + hipGetErrorString(e));
        return OPALSC_EHIP;
    }
    return OPALSC_OK;
}

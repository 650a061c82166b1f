// C ABI of the space-charge step (include/opalsc.h).  Host code only: owns
// the context's device memory and stream and sequences the kernels.  No C++
// exception crosses the boundary; every entry point returns a status code.
#include "../../include/opalsc.h"
#include "bingamma.hpp"
#include "common.hpp"
#include "crmath.hpp"
#include "frame.hpp"
#include "igfmath.hpp"
#include "kernels.hpp"
#include "tuning.hpp"  // generated from src_amd/tuning.json by src_amd/build.py

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace opalsc;

static_assert(sizeof(opalsc_stats) == sizeof(StatsOut), "stats layout");

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t b) {
        if (b <= bytes) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (b == 0) return true;
        if (hipMalloc(&p, b) != hipSuccess) return false;
        bytes = b;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return (T *)p; }
};

std::string g_create_error;

}  // namespace

// Errors of the entry points without a context (creation, file I/O), read by
// opalsc_last_error(NULL).
void opalsc::set_global_error(const std::string &msg) { g_create_error = msg; }

struct opalsc_ctx {
    opalsc_config cfg;
    std::string err;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // the tracker's Green's function, concurrent with the deposit
    hipEvent_t ev_mesh = nullptr, ev_green = nullptr, ev_gimg = nullptr, ev_sort = nullptr;
    hipEvent_t ev_start = nullptr, ev_zero = nullptr;  // a step's start; its charge grid cleared
    PoissonGrid grid;
    Bunch bunch;
    Scalars *d_sc = nullptr;
    TrackState *d_ts = nullptr;
    double *partials = nullptr;
    unsigned long long *d_counters = nullptr;  // [0] qd max bits, [1] dropped, [8..13] bound keys
    DevBuf aosR, aosP, aosEf, aosBf, dtb, statsb, tmpd;
    std::vector<FieldMapHost> fmaps;  // field maps on the beamline, in the order they were added
    DevBuf fm_desc, fm_data;
    DevBuf os_dense;  // per-block moment sums of the one-shot gather pass
    DevBuf osbox;     // per one-shot block: statistics of the positions it wrote (the next bounds)
    DevBuf binb, gamb, cntb;  // energy bins: per-particle bin, per-bin gamma, per-bin count
    // tracker energy bins: the resident bunch is ordered by bin
    BinRanges bins;
    double bin_qi = 0.0;
    bool bin_image = true;
    DevBuf binsorted;         // int32 bin of each resident particle (bin order)
    DevBuf binpart;           // calcGammas partial sums
    DevBuf binchunks;         // calcGammas chunk sums (bingamma.hpp), 2 per one-shot gather chunk
    DevBuf scb;               // Scalars per bin
    DevBuf scbc;              // Scalars of the non-empty bins, in bin order (the batched solve's)
    double qmax = 0.0;        // max |Q| of the resident bunch
    bool q_uniform = false;   // every particle of the resident bunch carries the charge q_value:
    double q_value = 0.0;     // the tracker then keeps Q as that one number (no per-particle array)
    bool have_bunch = false, have_track = false;
    // the resident bunch was re-sorted by cell since it was set: a re-sort can
    // then be applied by the gather pass (an arbitrary order is permuted first)
    bool cell_ordered = false;
    bool track_sc = false;    // the tracker computes space-charge fields
    bool track_ext = false;   // the tracker has a nonzero external field
    Profiler prof;
    double dh = 0.02;
    double period = kC / 1e9;  // DC beam period length (PartBunchBase.hpp:132: c / 1 GHz)
    int steps_done = 0;
    uint64_t gen = 0;  // bumped whenever buffers or sizes a captured step refers to change
};

#define FAIL(ctx, code, msg)   \
    do {                       \
        (ctx)->err = (msg);    \
        return (code);         \
    } while (0)

#define HIPCHK(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) FAIL(ctx, OPALSC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK_LAUNCH(ctx) HIPCHK(ctx, hipGetLastError())

// Every entry point works on the device the context was created on.
static int use_device(opalsc_ctx *c) {
    if (hipSetDevice(c->cfg.device) != hipSuccess) FAIL(c, OPALSC_EHIP, "hipSetDevice failed");
    return 0;
}

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
namespace opalsc {
const char *prof_name(int id) {
    static const char *names[PROF_COUNT] = {"push_rotate_bounds", "scalar", "sort", "zero_rho",
                                            "deposit", "green", "green_dct", "fft_x", "fft_y",
                                            "fft_z_conv", "grad", "gather_kick_push", "rotate_bounds",
                                            "gather_kick_push_resort"};
    return (id >= 0 && id < PROF_COUNT) ? names[id] : "?";
}
void Profiler::start() {
    if (!enabled) return;
    mark(-1);
}
void Profiler::mark(int id) {
    if (!enabled) return;
    if (used == (int)pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        pool.push_back(e);
        ids.push_back(0);
    }
    ids[used] = id;
    (void)hipEventRecord(pool[used], stream);
    ++used;
}
void Profiler::flush() {
    if (!enabled || used == 0) return;
    (void)hipEventSynchronize(pool[used - 1]);
    for (int i = 1; i < used; ++i) {
        if (ids[i] < 0) continue;
        float t = 0.f;
        if (hipEventElapsedTime(&t, pool[i - 1], pool[i]) == hipSuccess) {
            ms[ids[i]] += t;
            launches[ids[i]] += 1;
        }
    }
    used = 0;
}
void Profiler::destroy() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
    ids.clear();
    used = 0;
}
}  // namespace opalsc

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int opalsc_abi_version(void) { return OPALSC_ABI_VERSION; }

extern "C" const char *opalsc_last_error(const opalsc_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

static int alloc_bunch(opalsc_ctx *c, int64_t cap) {
    Bunch &b = c->bunch;
    b.cap = cap;
    double **arrs[] = {&b.x, &b.y, &b.z, &b.px, &b.py, &b.pz, &b.q, &b.tmp, &b.qd};
    for (double **a : arrs)
        if (hipMalloc((void **)a, (size_t)std::max<int64_t>(cap, 1) * 8) != hipSuccess)
            FAIL(c, OPALSC_ENOMEM, "hipMalloc of particle arrays failed");
    uint32_t **u[] = {&b.id, &b.keys, &b.vals, &b.keys2, &b.vals2};
    for (uint32_t **a : u)
        if (hipMalloc((void **)a, (size_t)std::max<int64_t>(cap, 1) * 4) != hipSuccess)
            FAIL(c, OPALSC_ENOMEM, "hipMalloc of particle index arrays failed");
    b.sort_tmp_bytes = sort_temp_bytes(std::max<int64_t>(cap, 1));
    if (hipMalloc(&b.sort_tmp, std::max<size_t>(b.sort_tmp_bytes, 16)) != hipSuccess)
        FAIL(c, OPALSC_ENOMEM, "hipMalloc of sort scratch failed");
    const size_t nchunks = (size_t)(std::max<int64_t>(cap, 1) + kDepositChunk - 1) / kDepositChunk;
    if (hipMalloc((void **)&b.chunkbox, nchunks * kBoxStats * sizeof(double)) != hipSuccess)
        FAIL(c, OPALSC_ENOMEM, "hipMalloc of chunk boxes failed");
    // a tracker that re-sorts: the gather pass writes the re-sorted arrays here
    if (c->cfg.sort_interval > 0)
        for (double *&a : b.alt)
            if (hipMalloc((void **)&a, (size_t)std::max<int64_t>(cap, 1) * 8) != hipSuccess)
                FAIL(c, OPALSC_ENOMEM, "hipMalloc of re-sort arrays failed");
    return 0;
}

static void free_bunch(Bunch &b) {
    for (double *a : b.alt)
        if (a) (void)hipFree(a);
    void *ps[] = {b.x, b.y, b.z, b.px, b.py, b.pz, b.q, b.tmp, b.qd, b.id, b.keys, b.vals, b.keys2, b.vals2, b.sort_tmp,
                  b.chunkbox};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    b = Bunch();
}

extern "C" int opalsc_create(opalsc_ctx **out, const opalsc_config *cfg) {
    if (!out || !cfg) {
        g_create_error = "null argument";
        return OPALSC_EINVAL;
    }
    *out = nullptr;
    if (cfg->precision != OPALSC_FP64 && cfg->precision != OPALSC_MIXED) {
        g_create_error = "precision must be OPALSC_FP64 or OPALSC_MIXED";
        return OPALSC_EINVAL;
    }
    if (cfg->greens != OPALSC_GREENS_INTEGRATED && cfg->greens != OPALSC_GREENS_STANDARD) {
        g_create_error = "greens must be INTEGRATED or STANDARD";
        return OPALSC_EINVAL;
    }
    if (cfg->max_particles < 0 || cfg->max_particles > (int64_t)0xFFFFFFF0) {
        g_create_error = "max_particles out of range (0 .. 2^32-16)";
        return OPALSC_EINVAL;
    }
    if (cfg->bcz != OPALSC_BC_OPEN && cfg->bcz != OPALSC_BC_PERIODIC) {
        g_create_error = "bcz must be OPALSC_BC_OPEN or OPALSC_BC_PERIODIC";
        return OPALSC_EINVAL;
    }
    opalsc_ctx *c = new (std::nothrow) opalsc_ctx();
    if (!c) {
        g_create_error = "out of host memory";
        return OPALSC_ENOMEM;
    }
    c->cfg = *cfg;
    c->dh = cfg->bboxincr / 100.0;
    auto bail = [&](int code) {
        g_create_error = c->err;
        opalsc_destroy(c);
        return code;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) {
        c->err = "hipSetDevice failed";
        return bail(OPALSC_EHIP);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(OPALSC_EHIP);
    }
    c->prof.stream = c->stream;
    if (c->grid.init(cfg->n, cfg->precision == OPALSC_FP64, cfg->greens, cfg->bcz == OPALSC_BC_PERIODIC, c->stream,
                     c->err) != 0)
        return bail(OPALSC_EINVAL);
    c->grid.prof = &c->prof;
    if (alloc_bunch(c, cfg->max_particles) != 0) return bail(OPALSC_ENOMEM);
    if (hipMalloc((void **)&c->d_sc, sizeof(Scalars)) != hipSuccess ||
        hipMalloc((void **)&c->d_ts, sizeof(TrackState)) != hipSuccess ||
        hipMalloc((void **)&c->partials, sizeof(double) * kSlots * kRedBlocksMax) != hipSuccess ||
        hipMalloc((void **)&c->d_counters, sizeof(unsigned long long) * 16) != hipSuccess) {
        c->err = "hipMalloc of scalar buffers failed";
        return bail(OPALSC_ENOMEM);
    }
    Scalars z;
    memset(&z, 0, sizeof(z));
    for (int d = 0; d < 3; ++d) z.n[d] = cfg->n[d];
    if (hipMemcpy(c->d_sc, &z, sizeof(z), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->d_counters, 0, sizeof(unsigned long long) * 16) != hipSuccess) {
        c->err = "initialisation copy failed";
        return bail(OPALSC_EHIP);
    }
    *out = c;
    return OPALSC_OK;
}

extern "C" int opalsc_destroy(opalsc_ctx *c) {
    if (!c) return OPALSC_OK;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->grid.destroy();
    free_bunch(c->bunch);
    DevBuf *bs[] = {&c->aosR, &c->aosP, &c->aosEf, &c->aosBf, &c->dtb,      &c->statsb,  &c->tmpd,
                    &c->binb, &c->gamb, &c->cntb,  &c->binsorted, &c->binpart, &c->scb, &c->scbc, &c->fm_desc, &c->fm_data,
                    &c->os_dense, &c->osbox, &c->binchunks};
    for (DevBuf *b : bs) b->release();
    void *ps[] = {c->d_sc, c->d_ts, c->partials, c->d_counters};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    c->prof.destroy();
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->ev_mesh) (void)hipEventDestroy(c->ev_mesh);
    if (c->ev_green) (void)hipEventDestroy(c->ev_green);
    if (c->ev_sort) (void)hipEventDestroy(c->ev_sort);
    if (c->ev_gimg) (void)hipEventDestroy(c->ev_gimg);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_zero) (void)hipEventDestroy(c->ev_zero);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OPALSC_OK;
}

extern "C" int opalsc_synchronize(opalsc_ctx *c) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

static int check_n(opalsc_ctx *c, int64_t N) {
    if (N < 0) FAIL(c, OPALSC_EINVAL, "negative particle count");
    if (N > c->bunch.cap) FAIL(c, OPALSC_ECAP, "more particles than max_particles");
    return 0;
}

// A periodic-z (DC beam) context has no energy bins or image charge (the
// shifted Green's function is an open-boundary construction).
static int reject_periodic_z(opalsc_ctx *c) {
    if (c->grid.pz) FAIL(c, OPALSC_EINVAL, "energy bins and image charge need an open-boundary context");
    return 0;
}

static int upload(opalsc_ctx *c, DevBuf &b, const void *src, size_t bytes) {
    if (!b.ensure(bytes)) FAIL(c, OPALSC_ENOMEM, "hipMalloc of staging buffer failed");
    if (bytes) HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
}

static void upload_scalars(opalsc_ctx *c, const Scalars &s) {
    (void)hipMemcpyAsync(c->d_sc, &s, sizeof(s), hipMemcpyHostToDevice, c->stream);
}

static int download_scalars(opalsc_ctx *c, Scalars &s) {
    HIPCHK(c, hipMemcpyAsync(&s, c->d_sc, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Poisson solver
// ---------------------------------------------------------------------------
// Every byte of [p, p + bytes) is device memory the library may read and write.
static bool is_device_range(const void *p, size_t bytes) { return device_range_ok(p, bytes); }

static void potential_scalars(opalsc_ctx *c, const double hr[3]) {
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) {
        s.n[d] = c->cfg.n[d];
        s.hr[d] = hr[d];
        s.hs[d] = hr[d];
    }
    s.dT = 1.0;
    s.tmp2 = 1.0;
    s.q2e = 1.0;
    upload_scalars(c, s);
}

extern "C" int opalsc_compute_potential_dev(opalsc_ctx *c, double *rho_dev, const double hr[3]) {
    if (!c || !rho_dev || !hr) return c ? (c->err = "null argument", OPALSC_EINVAL) : OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    const size_t nc = c->grid.ncell;
    if (!is_device_range(rho_dev, nc * 8)) FAIL(c, OPALSC_EINVAL, "rho_dev is not device memory");
    potential_scalars(c, hr);
    HIPCHK(c, hipMemcpyAsync(c->grid.rho_f64, rho_dev, nc * 8, hipMemcpyDeviceToDevice, c->stream));
    c->grid.green(c->d_sc);
    c->grid.solve(c->d_sc, 1, 0);
    CHECK_LAUNCH(c);
    c->grid.widen(1, rho_dev);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_compute_potential(opalsc_ctx *c, double *rho, const double hr[3]) {
    if (!c || !rho || !hr) return c ? (c->err = "null argument", OPALSC_EINVAL) : OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) {
        s.n[d] = c->cfg.n[d];
        s.hr[d] = hr[d];
        s.hs[d] = hr[d];
    }
    s.dT = 1.0;
    s.tmp2 = 1.0;
    s.q2e = 1.0;
    upload_scalars(c, s);
    const size_t nc = c->grid.ncell;
    HIPCHK(c, hipMemcpyAsync(c->grid.rho_f64, rho, nc * 8, hipMemcpyHostToDevice, c->stream));
    c->grid.green(c->d_sc);
    c->grid.solve(c->d_sc, 1, 0);
    CHECK_LAUNCH(c);
    c->grid.widen(1, (double *)c->grid.rho_f64);
    HIPCHK(c, hipMemcpyAsync(rho, c->grid.rho_f64, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_compute_potential_shifted(opalsc_ctx *c, double *rho, const double hr[3], double zshift) {
    if (!c || !rho || !hr) return c ? (c->err = "null argument", OPALSC_EINVAL) : OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = reject_periodic_z(c)) return e;
    if (c->grid.ensure_image(c->err)) return OPALSC_ENOMEM;
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) {
        s.n[d] = c->cfg.n[d];
        s.hr[d] = hr[d];
        s.hs[d] = hr[d];
    }
    s.dT = 1.0;
    s.tmp2 = 1.0;
    s.q2e = 1.0;
    s.zshift = zshift;
    upload_scalars(c, s);
    const size_t nc = c->grid.ncell;
    HIPCHK(c, hipMemcpyAsync(c->grid.rho_f64, rho, nc * 8, hipMemcpyHostToDevice, c->stream));
    c->grid.green_image(c->d_sc);
    c->grid.solve(c->d_sc, 1, 0, true);
    CHECK_LAUNCH(c);
    c->grid.widen(1, (double *)c->grid.rho_f64);
    HIPCHK(c, hipMemcpyAsync(rho, c->grid.rho_f64, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// mesh, CIC, sort
// ---------------------------------------------------------------------------
extern "C" int opalsc_boundp(opalsc_ctx *c, int64_t N, const double *R, double hr[3], double origin[3]) {
    if (!c || (!R && N > 0) || !hr || !origin) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    launch_bounds(c->stream, N, c->aosR.as<double>(), c->partials);
    launch_boundp_scalars(c->stream, c->partials, N > 0 ? red_blocks(N) : 0, c->cfg.n, c->dh, c->grid.pz, c->period,
                          c->d_sc);
    CHECK_LAUNCH(c);
    Scalars s;
    if (int e = download_scalars(c, s)) return e;
    if (!s.mesh_ok) FAIL(c, OPALSC_EMESH, "boundp: h<0, can not build a mesh");
    for (int d = 0; d < 3; ++d) {
        hr[d] = s.hr[d];
        origin[d] = s.origin[d];
    }
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// field maps on the beamline
// ---------------------------------------------------------------------------
// Uploads every map (descriptors with offsets into one data array) and points
// the tracker state at them.
static int upload_fieldmaps(opalsc_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<FieldMapDesc> desc;
    std::vector<double> data;
    for (const FieldMapHost &h : c->fmaps) {
        FieldMapDesc d = h.d;
        const int64_t base = (int64_t)data.size();
        d.a = base;
        data.insert(data.end(), h.a.begin(), h.a.end());
        d.b = (int64_t)data.size();
        data.insert(data.end(), h.b.begin(), h.b.end());
        d.c = (int64_t)data.size();
        data.insert(data.end(), h.c.begin(), h.c.end());
        desc.push_back(d);
    }
    if (!c->fm_desc.ensure(sizeof(FieldMapDesc) * kMaxFieldMaps) || !c->fm_data.ensure(data.size() * 8))
        FAIL(c, OPALSC_ENOMEM, "field map allocation failed");
    if (!desc.empty())
        HIPCHK(c, hipMemcpyAsync(c->fm_desc.p, desc.data(), desc.size() * sizeof(FieldMapDesc),
                                 hipMemcpyHostToDevice, c->stream));
    if (!data.empty())
        HIPCHK(c, hipMemcpyAsync(c->fm_data.p, data.data(), data.size() * 8, hipMemcpyHostToDevice, c->stream));
    if (c->have_track) {
        struct {
            int nmaps;
            const FieldMapDesc *maps;
            const double *mapdata;
        } f = {(int)desc.size(), c->fm_desc.as<FieldMapDesc>(), c->fm_data.as<double>()};
        static_assert(offsetof(TrackState, maps) - offsetof(TrackState, nmaps) == offsetof(decltype(f), maps),
                      "field map members of TrackState");
        static_assert(offsetof(TrackState, mapdata) - offsetof(TrackState, nmaps) == offsetof(decltype(f), mapdata),
                      "field map members of TrackState");
        HIPCHK(c, hipMemcpyAsync((char *)c->d_ts + offsetof(TrackState, nmaps), &f, sizeof(f), hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ++c->gen;
    return OPALSC_OK;
}

// A Solenoid (2D magnetostatic map) or RFCavity (2D dynamic map) whose map
// starts at elemedge on the beamline (Solenoid::initialise, Solenoid.cpp:195-214;
// RFCavity::initialise, RFCavity.cpp:285-310): scale is KS or VOLT, freq_hz
// the cavity's frequency (the map's when 0 or more than 1 % away from it, as
// there), phase its LAG.  Returns the element's index or a negative error.
extern "C" int opalsc_fieldmap_add(opalsc_ctx *c, const char *path, double elemedge, double scale,
                                   double freq_hz, double phase) {
    if (!c || !path) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if ((int)c->fmaps.size() >= kMaxFieldMaps) FAIL(c, OPALSC_EINVAL, "at most 8 field maps");
    FieldMapHost h;
    std::string err;
    if (parse_fieldmap_2d(path, h, err) != 0) FAIL(c, OPALSC_EINVAL, err);
    h.d.zstart = elemedge + h.d.zstart;
    h.d.scale = scale;
    if (h.d.kind == 1) {
        const double two_pi = 2.0 * kPi;
        double omega = freq_hz * two_pi;
        if (!(omega > 0.0) || std::abs((omega - h.frequency) / omega) > 0.01) omega = h.frequency;
        h.d.omega = omega;
        h.d.phase = phase;
    }
    c->fmaps.push_back(std::move(h));
    if (int e = upload_fieldmaps(c)) {
        c->fmaps.pop_back();
        return e;
    }
    return (int)c->fmaps.size() - 1;
}

extern "C" int opalsc_fieldmap_clear(opalsc_ctx *c) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    c->fmaps.clear();
    return upload_fieldmaps(c);
}

// The field of every map at lab positions R (AoS) and time t, summed in the
// order the maps were added into E and B (overwritten); lost (may be NULL)
// counts the particles beyond a map's radius.
extern "C" int opalsc_fieldmap_eval(opalsc_ctx *c, int64_t N, const double *R, double t, double *E, double *B,
                                    int64_t *lost) {
    if (!c || (N > 0 && (!R || !E || !B))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (!c->fm_desc.p && upload_fieldmaps(c)) return OPALSC_ENOMEM;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (!c->aosEf.ensure((size_t)N * 24 + 8) || !c->aosBf.ensure((size_t)N * 24 + 8))
        FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    HIPCHK(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * 2, c->stream));
    launch_fieldmap_eval(c->stream, N, c->aosR.as<double>(), t, c->fm_desc.as<FieldMapDesc>(),
                         (int)c->fmaps.size(), c->fm_data.as<double>(), c->aosEf.as<double>(),
                         c->aosBf.as<double>(), &c->d_counters[1]);
    CHECK_LAUNCH(c);
    unsigned long long nl = 0;
    if (N > 0) {
        HIPCHK(c, hipMemcpyAsync(E, c->aosEf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(B, c->aosBf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(&nl, &c->d_counters[1], sizeof(nl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (lost) *lost = (int64_t)nl;
    return OPALSC_OK;
}

// PartBunchBase::setBeamFrequency (src/Classic/Algorithms/PartBunchBase.hpp:2534-2537):
// the DC beam's period length becomes c / f.
extern "C" int opalsc_set_beam_frequency(opalsc_ctx *c, double f) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!(f > 0.0) || !std::isfinite(f)) FAIL(c, OPALSC_EINVAL, "beam frequency must be positive and finite");
    c->period = kC / f;
    if (c->have_track)
        HIPCHK(c, hipMemcpyAsync((char *)c->d_ts + offsetof(TrackState, period), &c->period, sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

// The particle boundary condition boundp's update() applies to a DC beam
// (ParticleSpatialLayout.hpp:296-297): z wrapped into [origin, n*h + origin).
extern "C" int opalsc_wrap_periodic_z(opalsc_ctx *c, int64_t N, double *R, const double origin[3],
                                      const double hr[3]) {
    if (!c || (!R && N > 0) || !origin || !hr) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    const double lo = 0 * hr[2] + origin[2], hi = c->cfg.n[2] * hr[2] + origin[2];
    if (!(hi > lo)) FAIL(c, OPALSC_EINVAL, "the mesh has no extent in z");
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    launch_wrap_periodic_z(c->stream, N, c->aosR.as<double>(), lo, hi);
    CHECK_LAUNCH(c);
    if (N > 0) HIPCHK(c, hipMemcpyAsync(R, c->aosR.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_cic_index(opalsc_ctx *c, int64_t N, const double *R, const double origin[3],
                                const double hr[3], int32_t *lgp, double *dpos) {
    if (!c || (N > 0 && (!R || !lgp || !dpos)) || !origin || !hr) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (N == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (!c->aosP.ensure((size_t)N * 24) || !c->aosEf.ensure((size_t)N * 24))
        FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    launch_cic_index(c->stream, N, c->aosR.as<double>(), origin, hr, c->aosP.as<int32_t>(), c->aosEf.as<double>());
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(lgp, c->aosP.p, (size_t)N * 12, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(dpos, c->aosEf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

static int key_bits(const int n[3]) {
    uint64_t m = (uint64_t)cell_key_count(n[0], n[1], n[2]) + 1;
    int b = 1;
    while (b < 32 && (1ull << b) < m) ++b;
    return b;
}
// the bits of the cell keys alone (cell_key_count is a power of two)
static int key_bits_in_grid(const int n[3]) {
    const uint64_t m = (uint64_t)cell_key_count(n[0], n[1], n[2]);
    int b = 0;
    while (b < 32 && (1ull << b) < m) ++b;
    return b;
}

extern "C" int opalsc_sort_by_cell(opalsc_ctx *c, int64_t N, const double *R, const double origin[3],
                                   const double hr[3], uint32_t *perm) {
    if (!c || (N > 0 && (!R || !perm)) || !origin || !hr) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (N == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    Bunch &b = c->bunch;
    launch_soa_from_aos(c->stream, N, c->aosR.as<double>(), b.x, b.y, b.z);
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) {
        s.n[d] = c->cfg.n[d];
        s.hr[d] = hr[d];
        s.origin[d] = origin[d];
    }
    upload_scalars(c, s);
    launch_cell_keys(c->stream, N, b.x, b.y, b.z, c->d_sc, false, b.keys, b.vals);
    if (sort_pairs(c->stream, b.sort_tmp, b.sort_tmp_bytes, b.keys, b.keys2, b.vals, b.vals2, N, key_bits(c->cfg.n)))
        FAIL(c, OPALSC_EHIP, "radix sort failed");
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(perm, b.vals2, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_bunch = c->have_track = false;  // SoA buffers were reused
    ++c->gen;
    c->bins.nbins = 0;
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// self fields
// ---------------------------------------------------------------------------
static int self_fields_checks(opalsc_ctx *c, int64_t N, const void *R, const void *P, const void *Q, const void *dt,
                              const double origin[3], const double hr[3], const void *Ef, const void *Bf) {
    if (!c || !origin || !hr) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (N > 0 && (!R || !P || !Q || !dt || !Ef || !Bf)) FAIL(c, OPALSC_EINVAL, "null particle array");
    if (int e = check_n(c, N)) return e;
    if (N == 0) FAIL(c, OPALSC_EINVAL, "computeSelfFields needs particles (the Lorentz factor divides by N)");
    return OPALSC_OK;
}

// PartBunch::computeSelfFields on device arrays (AoS R, P, Ef, Bf; Q updated in place).
static int self_fields_core(opalsc_ctx *c, int64_t N, const double *R, const double *P, double *Q, const double *dt,
                            double dT, const double origin[3], const double hr[3], double coupling, double *Ef,
                            double *Bf) {
    Bunch &b = c->bunch;
    HIPCHK(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * 2, c->stream));
    launch_soa_from_aos(c->stream, N, R, b.x, b.y, b.z);
    launch_selffield_prep(c->stream, N, P, Q, dt, b.qd, c->partials, &c->d_counters[0]);
    launch_selffield_scalars(c->stream, c->partials, red_blocks(N), N, dT, origin, hr, c->cfg.n, coupling,
                             &c->d_counters[0], c->grid.fp64, c->grid.pz, c->d_sc);
    c->grid.zero_rho(c->stream);
    launch_deposit(c->stream, 0, c->grid.fp64, N, b.x, b.y, b.z, nullptr, b.qd, (long long *)c->grid.rho_i64,
                   c->d_sc, &c->d_counters[1], nullptr);
    c->grid.green(c->d_sc);
    c->grid.solve(c->d_sc, 0, 1);
    c->grid.grad(c->d_sc);
    launch_gather_selffield(c->stream, c->grid.fp64, N, b.x, b.y, b.z, c->grid.Eg, c->d_sc, Ef, Bf);
    CHECK_LAUNCH(c);
    c->have_bunch = c->have_track = false;  // SoA buffers were reused
    ++c->gen;
    c->bins.nbins = 0;
    return OPALSC_OK;
}

extern "C" int opalsc_compute_self_fields(opalsc_ctx *c, int64_t N, const double *R, const double *P,
                                          double *Q, const double *dt, double dT,
                                          const double origin[3], const double hr[3],
                                          double coupling, double *Ef, double *Bf) {
    if (int e = self_fields_checks(c, N, R, P, Q, dt, origin, hr, Ef, Bf)) return e;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosBf, Bf, (size_t)N * 24)) return e;
    if (int e = upload(c, c->dtb, dt, (size_t)N * 8)) return e;
    if (!c->aosEf.ensure((size_t)N * 24)) FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    double *q = c->bunch.q;
    HIPCHK(c, hipMemcpyAsync(q, Q, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    if (int e = self_fields_core(c, N, c->aosR.as<double>(), c->aosP.as<double>(), q, c->dtb.as<double>(), dT,
                                 origin, hr, coupling, c->aosEf.as<double>(), c->aosBf.as<double>()))
        return e;
    HIPCHK(c, hipMemcpyAsync(Ef, c->aosEf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Bf, c->aosBf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Q, q, (size_t)N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_compute_self_fields_dev(opalsc_ctx *c, int64_t N, const double *R_dev, const double *P_dev,
                                              double *Q_dev, const double *dt_dev, double dT,
                                              const double origin[3], const double hr[3], double coupling,
                                              double *Ef_dev, double *Bf_dev) {
    if (int e = self_fields_checks(c, N, R_dev, P_dev, Q_dev, dt_dev, origin, hr, Ef_dev, Bf_dev)) return e;
    const size_t v = (size_t)N * 24, sz = (size_t)N * 8;
    if (!is_device_range(R_dev, v) || !is_device_range(P_dev, v) || !is_device_range(Q_dev, sz) ||
        !is_device_range(dt_dev, sz) || !is_device_range(Ef_dev, v) || !is_device_range(Bf_dev, v))
        FAIL(c, OPALSC_EINVAL, "a _dev argument is not device memory");
    if (int e = self_fields_core(c, N, R_dev, P_dev, Q_dev, dt_dev, dT, origin, hr, coupling, Ef_dev, Bf_dev))
        return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_compute_self_fields_binned(opalsc_ctx *c, int64_t N, const double *R, const double *P,
                                                 const int32_t *bin, int nbins, double qi, const double *dt,
                                                 double dT, const double origin[3], const double hr[3],
                                                 double coupling, const double *gamma_bin, int image,
                                                 double *Ef, double *Bf) {
    if (!c || !origin || !hr) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = reject_periodic_z(c)) return e;
    if (N > 0 && (!R || !P || !bin || !dt || !Ef || !Bf)) FAIL(c, OPALSC_EINVAL, "null particle array");
    if (nbins < 1 || nbins > kMaxBins) FAIL(c, OPALSC_EINVAL, "nbins must be in [1, 64]");
    if (int e = check_n(c, N)) return e;
    if (N == 0) return OPALSC_OK;
    if (c->grid.ensure_sums(c->err)) return OPALSC_ENOMEM;
    if (!c->scbc.ensure(sizeof(Scalars) * kMaxBins)) FAIL(c, OPALSC_ENOMEM, "bin scalars");
    Bunch &b = c->bunch;
    const bool exact = c->grid.fp64;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosBf, Bf, (size_t)N * 24)) return e;
    if (int e = upload(c, c->dtb, dt, (size_t)N * 8)) return e;
    if (int e = upload(c, c->binb, bin, (size_t)N * 4)) return e;
    if (!c->aosEf.ensure((size_t)N * 24) || !c->gamb.ensure(kMaxBins * 8) || !c->cntb.ensure(kMaxBins * 8))
        FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    if (gamma_bin)
        HIPCHK(c, hipMemcpyAsync(c->gamb.p, gamma_bin, (size_t)nbins * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * 2, c->stream));
    HIPCHK(c, hipMemsetAsync(c->cntb.p, 0, kMaxBins * 8, c->stream));
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) {
        s.n[d] = c->cfg.n[d];
        s.hr[d] = hr[d];
        s.origin[d] = origin[d];
    }
    s.dT = dT;
    s.coupling = coupling;
    s.N = N;
    upload_scalars(c, s);
    // order by bin: each bin's particles become one range of the deposit
    launch_soa_from_aos(c->stream, N, c->aosR.as<double>(), b.x, b.y, b.z);
    launch_bin_count(c->stream, N, c->binb.as<int32_t>(), nbins, c->cntb.as<unsigned long long>());
    launch_bin_keys(c->stream, N, c->binb.as<int32_t>(), nbins, b.keys, b.vals);
    int bits = 1;
    while ((1 << bits) <= nbins) ++bits;
    if (sort_pairs(c->stream, b.sort_tmp, b.sort_tmp_bytes, b.keys, b.keys2, b.vals, b.vals2, N, bits))
        FAIL(c, OPALSC_EHIP, "radix sort failed");
    launch_bin_order(c->stream, N, b.vals2, b.x, b.y, b.z, c->dtb.as<double>(), qi, b.px, b.py, b.pz, b.qd);
    launch_absmax(c->stream, N, b.qd, &c->d_counters[0]);
    unsigned long long cnt[kMaxBins];
    HIPCHK(c, hipMemcpyAsync(cnt, c->cntb.p, (size_t)nbins * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t start[kMaxBins];
    int64_t acc = 0;
    for (int k = 0; k < nbins; ++k) {
        start[k] = acc;
        acc += (int64_t)cnt[k];
    }
    if (!gamma_bin)
        for (int k = 0; k < nbins; ++k)
            launch_bin_gamma(c->stream, (int64_t)cnt[k], b.vals2 + start[k], c->aosP.as<double>(), c->partials, k,
                             c->gamb.as<double>());
    // the non-empty bins (an empty bin has no field) solved as one batch, each
    // with the mesh scalars and its own Lorentz frame, charge scale and shift
    int nz = 0;
    for (int k = 0; k < nbins; ++k) nz += cnt[k] > 0 ? 1 : 0;
    if (nz == 0) {
        c->grid.clear_sums();
    } else {
        if (c->grid.ensure_bins(nz, image != 0, c->err)) return OPALSC_ENOMEM;
        Scalars *scz = c->scbc.as<Scalars>();
        c->grid.zero_bin_rho(nz, c->stream);
        for (int k = 0, z = 0; k < nbins; ++k) {
            if (cnt[k] == 0) continue;
            const int64_t s0 = start[k], nk = (int64_t)cnt[k];
            HIPCHK(c, hipMemcpyAsync(scz + z, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToDevice, c->stream));
            launch_bin_scalars(c->stream, scz + z, c->gamb.as<double>(), k, nk, &c->d_counters[0], exact);
            launch_deposit(c->stream, 0, exact, nk, b.px + s0, b.py + s0, b.pz + s0, nullptr, b.qd + s0,
                           c->grid.bin_rho(z), scz + z, &c->d_counters[1], nullptr);
            ++z;
        }
        c->grid.green_bins(scz, nz, image != 0, c->stream);
        c->grid.solve_bins(scz, nz, image != 0, nullptr);
        c->grid.field_sums(scz, nz, image != 0);
        CHECK_LAUNCH(c);
    }
    launch_gather_binned(c->stream, c->grid.fp64, N, b.x, b.y, b.z, c->grid.Esum, c->grid.Bsum, c->d_sc,
                         c->aosEf.as<double>(), c->aosBf.as<double>());
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(Ef, c->aosEf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Bf, c->aosBf.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_bunch = c->have_track = false;  // SoA buffers were reused
    ++c->gen;
    c->bins.nbins = 0;
    return OPALSC_OK;
}

extern "C" int opalsc_get_grid(opalsc_ctx *c, int which, double *out) {
    if (!c || !out) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    const size_t nc = c->grid.ncell;
    if (which == OPALSC_GRID_RHO) {
        c->grid.rho_to_double(c->d_sc, (double *)c->grid.rho_f64);
        HIPCHK(c, hipMemcpyAsync(out, c->grid.rho_f64, nc * 8, hipMemcpyDeviceToHost, c->stream));
    } else if (which == OPALSC_GRID_PHI) {
        c->grid.widen(1, (double *)c->grid.rho_f64);
        HIPCHK(c, hipMemcpyAsync(out, c->grid.rho_f64, nc * 8, hipMemcpyDeviceToHost, c->stream));
    } else if (which == OPALSC_GRID_EFIELD) {
        if (!c->tmpd.ensure(nc * 24)) FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
        // a tracker step with energy bins leaves the bins' summed field
        const bool sums = c->bins.nbins > 0 && c->track_sc && c->grid.Esum && c->steps_done > 0;
        c->grid.widen(2, c->tmpd.as<double>(), sums ? c->grid.Esum : nullptr);
        HIPCHK(c, hipMemcpyAsync(out, c->tmpd.p, nc * 24, hipMemcpyDeviceToHost, c->stream));
    } else {
        FAIL(c, OPALSC_EINVAL, "unknown grid");
    }
    CHECK_LAUNCH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// pusher, moments
// ---------------------------------------------------------------------------
extern "C" int opalsc_push_half(opalsc_ctx *c, int64_t N, double *R, const double *P, const double *dt) {
    if (!c || (N > 0 && (!R || !P || !dt))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (N == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
    if (int e = upload(c, c->dtb, dt, (size_t)N * 8)) return e;
    launch_push(c->stream, N, c->aosR.as<double>(), c->aosP.as<double>(), c->dtb.as<double>());
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(R, c->aosR.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_kick(opalsc_ctx *c, int64_t N, double *P, const double *Ef, const double *Bf,
                           const double *dt, double charge, double mass_eV) {
    if (!c || (N > 0 && (!P || !Ef || !Bf || !dt))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (N == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosEf, Ef, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosBf, Bf, (size_t)N * 24)) return e;
    if (int e = upload(c, c->dtb, dt, (size_t)N * 8)) return e;
    launch_kick(c->stream, N, c->aosP.as<double>(), c->aosEf.as<double>(), c->aosBf.as<double>(),
                c->dtb.as<double>(), charge, mass_eV);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(P, c->aosP.p, (size_t)N * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_beam_parameters(opalsc_ctx *c, int64_t N, const double *R, const double *P,
                                      double mass_eV, opalsc_stats *out) {
    if (!c || !out || (N > 0 && (!R || !P))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    if (N == 0) {
        memset(out, 0, sizeof(*out));
        return OPALSC_OK;
    }
    if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
    if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
    if (!c->statsb.ensure(sizeof(opalsc_stats))) FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    launch_moments(c->stream, N, c->aosR.as<double>(), c->aosP.as<double>(), c->partials);
    launch_stats_from_partials(c->stream, c->partials, red_blocks(N), N, mass_eV, c->statsb.p);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(out, c->statsb.p, sizeof(opalsc_stats), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// device-resident tracker
// ---------------------------------------------------------------------------
extern "C" int opalsc_bunch_set(opalsc_ctx *c, int64_t N, const double *R, const double *P, const double *Q) {
    if (!c || (N > 0 && (!R || !P || !Q))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, N)) return e;
    Bunch &b = c->bunch;
    b.N = N;
    if (N > 0) {
        if (int e = upload(c, c->aosR, R, (size_t)N * 24)) return e;
        launch_soa_from_aos(c->stream, N, c->aosR.as<double>(), b.x, b.y, b.z);
        if (int e = upload(c, c->aosP, P, (size_t)N * 24)) return e;
        launch_soa_from_aos(c->stream, N, c->aosP.as<double>(), b.px, b.py, b.pz);
        HIPCHK(c, hipMemcpyAsync(b.q, Q, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
        launch_iota(c->stream, N, b.id);
        CHECK_LAUNCH(c);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // the staging copies are not needed by the tracker
        c->aosR.release();
        c->aosP.release();
    }
    double qm = 0.0;
    bool uni = true;
    for (int64_t i = 0; i < N; ++i) {
        qm = std::max(qm, std::fabs(Q[i]));
        uni = uni && Q[i] == Q[0];
    }
    c->qmax = qm;
    c->q_uniform = uni && N > 0;
    c->q_value = N > 0 ? Q[0] : 0.0;
    c->have_bunch = true;
    c->have_track = false;
    c->cell_ordered = false;
    ++c->gen;
    c->bins.nbins = 0;
    return OPALSC_OK;
}

extern "C" int opalsc_bunch_generate(opalsc_ctx *c, const opalsc_gauss_params *g) {
    if (!c || !g) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (int e = check_n(c, g->n)) return e;
    Bunch &b = c->bunch;
    b.N = g->n;
    const double qeach = g->n > 0 ? g->q_total / (double)g->n : 0.0;
    launch_generate(c->stream, g->n, g->seed, g->sigma_r, g->sigma_p, g->p0, qeach, g->halo_frac,
                    g->halo_radius, b);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->qmax = std::fabs(qeach);
    c->q_uniform = true;
    c->q_value = qeach;
    c->have_bunch = true;
    c->have_track = false;
    c->cell_ordered = false;
    ++c->gen;
    c->bins.nbins = 0;
    return OPALSC_OK;
}

extern "C" int opalsc_bunch_get(opalsc_ctx *c, double *R, double *P, double *Q) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!c->have_bunch) FAIL(c, OPALSC_ESTATE, "no resident bunch");
    Bunch &b = c->bunch;
    const int64_t N = b.N;
    if (N == 0) return OPALSC_OK;
    std::vector<uint32_t> id(N);
    std::vector<double> tmp(N);
    HIPCHK(c, hipMemcpyAsync(id.data(), b.id, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto fetch = [&](const double *src, double *dst, int stride, int off) -> int {
        HIPCHK(c, hipMemcpy(tmp.data(), src, (size_t)N * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) dst[(size_t)id[i] * stride + off] = tmp[i];
        return 0;
    };
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (R) {
        if (int e = fetch(b.x, R, 3, 0)) return e;
        if (int e = fetch(b.y, R, 3, 1)) return e;
        if (int e = fetch(b.z, R, 3, 2)) return e;
    }
    if (P) {
        if (int e = fetch(b.px, P, 3, 0)) return e;
        if (int e = fetch(b.py, P, 3, 1)) return e;
        if (int e = fetch(b.pz, P, 3, 2)) return e;
    }
    if (Q && c->q_uniform) {
        double qv = c->q_value;
        if (c->have_track) {
            Scalars s;
            if (int e = download_scalars(c, s)) return e;
            qv = s.qu;
        }
        for (int64_t i = 0; i < N; ++i) Q[i] = qv;
    } else if (Q) {
        if (int e = fetch(b.q, Q, 1, 0)) return e;
    }
    return OPALSC_OK;
}

extern "C" int opalsc_track_init(opalsc_ctx *c, const opalsc_track_params *p) {
    if (!c || !p) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (c->bins.nbins > 0)
        if (int e = reject_periodic_z(c)) return e;
    if (!c->have_bunch) FAIL(c, OPALSC_ESTATE, "track_init before bunch_set/bunch_generate");
    if (!(p->dT > 0.0)) FAIL(c, OPALSC_EINVAL, "dT must be positive");
    TrackState ts;
    memset(&ts, 0, sizeof(ts));
    ts.dT = p->dT;
    ts.charge = p->charge;
    ts.mass_eV = p->mass_eV;
    ts.coupling = p->coupling;
    ts.Bz = p->Bz;
    ts.Ez0 = p->Ez0;
    ts.rf_freq = p->rf_freq;
    ts.rf_phase = p->rf_phase;
    ts.t = p->t;
    ts.pathLength = p->pathLength;
    for (int d = 0; d < 3; ++d) {
        ts.RefPartR[d] = p->RefPartR[d];
        ts.RefPartP[d] = p->RefPartP[d];
    }
    ts.space_charge = p->space_charge;
    ts.step = 0;
    ts.pz = c->grid.pz ? 1 : 0;
    ts.nmaps = (int)c->fmaps.size();
    ts.maps = c->fm_desc.as<FieldMapDesc>();
    ts.mapdata = c->fm_data.as<double>();
    ts.period = c->period;
    HIPCHK(c, hipMemcpyAsync(c->d_ts, &ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    Scalars s;
    memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) s.n[d] = c->cfg.n[d];
    s.N = c->bunch.N;
    s.qd_absmax = c->qmax * p->dT * (1.0 + 0x1p-50);
    s.qu = c->q_value;
    s.dT = p->dT;
    s.coupling = p->coupling;
    upload_scalars(c, s);
    launch_moments_soa(c->stream, c->bunch, c->partials);
    launch_track_init_moments(c->stream, c->partials, red_blocks(c->bunch.N), c->d_ts, c->d_sc);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t cap1 = std::max<int64_t>(c->bunch.cap, 1);
    if (!c->os_dense.ensure((size_t)std::max(os_blocks(cap1), os_blocks64(cap1)) * kMom * 8))
        c->os_dense.release();  // the persistent gather pass then serves
    if (!c->osbox.ensure((size_t)std::max(os_box_count(true, cap1), os_box_count(false, cap1)) * kOsBoxStats * 8))
        c->osbox.release();  // every step then runs the full bounds pass
    const int64_t nch = std::max((cap1 + kOsChunk64 - 1) / kOsChunk64, (cap1 + kOsChunk - 1) / kOsChunk);
    if (!c->binchunks.ensure((size_t)nch * 2 * 8))
        c->binchunks.release();  // energy bins then sum over the bins' momenta every step
    c->have_track = true;
    ++c->gen;
    c->track_sc = p->space_charge != 0;
    c->track_ext = p->Bz != 0.0 || p->Ez0 != 0.0;
    c->steps_done = 0;
    return OPALSC_OK;
}

static int bin_bits(int nbins) {  // bits of the values 0 .. nbins
    int bits = 1;
    while ((1 << bits) <= nbins) ++bits;
    return bits;
}

// Reorders every particle array of the resident bunch by perm (b.vals2), on
// stream st (default: the context's).
static void permute_bunch(opalsc_ctx *c, hipStream_t st = nullptr) {
    if (!st) st = c->stream;
    ++c->gen;  // the arrays trade places with the permutation buffer
    Bunch &b = c->bunch;
    const int64_t N = b.N;
    // one gather pass per array (measured faster than gathering all arrays in one pass)
    double **arrs[] = {&b.x, &b.y, &b.z, &b.px, &b.py, &b.pz, &b.q};
    for (double **a : arrs) {
        if (*a == b.q && c->q_uniform) continue;  // the charge is one number
        launch_permute(st, N, b.vals2, *a, b.tmp);
        std::swap(*a, b.tmp);
    }
    launch_permute_u32(st, N, b.vals2, b.id, b.keys);  // the sort's input keys are no longer needed
    std::swap(b.id, b.keys);
}

// The permutation (b.vals2) that orders the resident bunch by cell along a
// Hilbert curve; with energy bins, by cell within each bin, so the bins keep
// their ranges.  sc: the step's mesh (default: the context's scalars); st: the
// stream (default: the context's).  The key drops its g_sort_drop lowest bits
// (blocks of 2^(drop/3) cells per axis; particles keep their relative order
// within a block), set by the tuning table or opalsc_tune_sort_set.
static int g_sort_drop = OPALSC_TUNED_SORT_DROP_BITS;

static void sort_keys(opalsc_ctx *c, const Scalars *sc = nullptr, hipStream_t st = nullptr) {
    if (!sc) sc = c->d_sc;
    if (!st) st = c->stream;
    Bunch &b = c->bunch;
    const int64_t N = b.N;
    int bits = key_bits(c->cfg.n);
    const int coarse = std::min(g_sort_drop, bits - 1);
    if (c->bins.nbins > 0) {
        const int bb = bin_bits(c->bins.nbins);
        const int drop = std::max(coarse, bits + bb > 32 ? bits + bb - 32 : 0);
        launch_cell_keys(st, N, b.x, b.y, b.z, sc, true, b.keys, b.vals, c->binsorted.as<int32_t>(), bits - drop,
                         drop);
        bits = bits - drop + bb;
    } else if (coarse > 0) {
        launch_cell_keys(st, N, b.x, b.y, b.z, sc, true, b.keys, b.vals, nullptr, 0, coarse);
        bits = std::max(1, key_bits_in_grid(c->cfg.n) - coarse);
    } else {
        launch_cell_keys(st, N, b.x, b.y, b.z, sc, true, b.keys, b.vals);
        bits = key_bits_in_grid(c->cfg.n);
    }
    sort_pairs(st, b.sort_tmp, b.sort_tmp_bytes, b.keys, b.keys2, b.vals, b.vals2, N, bits);
}

// Cell order of the resident bunch, every array permuted.
static void sort_bunch(opalsc_ctx *c, const Scalars *sc = nullptr, hipStream_t st = nullptr) {
    sort_keys(c, sc, st);
    permute_bunch(c, st);
}

extern "C" int opalsc_bunch_set_bins(opalsc_ctx *c, const int32_t *bin, int nbins, double qi, int image) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!c->have_bunch) FAIL(c, OPALSC_ESTATE, "set_bins before bunch_set/bunch_generate");
    if (nbins > 0)
        if (int e = reject_periodic_z(c)) return e;
    if (nbins == 0) {
        c->bins.nbins = 0;
        ++c->gen;
        return OPALSC_OK;
    }
    if (nbins < 0 || nbins > kMaxBins) FAIL(c, OPALSC_EINVAL, "nbins must be in [0, 64]");
    c->cell_ordered = false;  // ordered by bin from here on
    Bunch &b = c->bunch;
    const int64_t N = b.N;
    if (N > 0 && !bin) FAIL(c, OPALSC_EINVAL, "null bin array");
    if (!c->binsorted.ensure((size_t)(N > 0 ? N : 1) * 4) || !c->cntb.ensure(kMaxBins * 8) ||
        !c->gamb.ensure(kMaxBins * 8) || !c->scb.ensure(sizeof(Scalars) * kMaxBins) ||
        !c->binpart.ensure(sizeof(double) * kMaxBins * kBinGammaBlocks * 2))
        FAIL(c, OPALSC_ENOMEM, "bin buffers");
    unsigned long long cnt[kMaxBins] = {0};
    if (N > 0) {
        if (int e = upload(c, c->binb, bin, (size_t)N * 4)) return e;
        // bin order, stable: particles keep their relative order within a bin
        launch_bin_keys_by_id(c->stream, N, c->binb.as<int32_t>(), b.id, nbins, b.keys, b.vals);
        if (sort_pairs(c->stream, b.sort_tmp, b.sort_tmp_bytes, b.keys, b.keys2, b.vals, b.vals2, N,
                       bin_bits(nbins)))
            FAIL(c, OPALSC_EHIP, "radix sort failed");
        HIPCHK(c, hipMemcpyAsync(c->binsorted.p, b.keys2, (size_t)N * 4, hipMemcpyDeviceToDevice, c->stream));
        permute_bunch(c);
        HIPCHK(c, hipMemsetAsync(c->cntb.p, 0, kMaxBins * 8, c->stream));
        launch_bin_count(c->stream, N, c->binsorted.as<int32_t>(), nbins, c->cntb.as<unsigned long long>());
        CHECK_LAUNCH(c);
        HIPCHK(c, hipMemcpyAsync(cnt, c->cntb.p, (size_t)nbins * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    ++c->gen;
    c->bins.nbins = nbins;
    c->bins.start[0] = 0;
    for (int k = 0; k < nbins; ++k) c->bins.start[k + 1] = c->bins.start[k] + (int64_t)cnt[k];
    c->bin_qi = qi;
    c->bin_image = image != 0;
    return OPALSC_OK;
}

extern "C" int opalsc_bunch_get_bins(opalsc_ctx *c, int *nbins, int64_t *counts, double *gammas) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    const int nb = c->bins.nbins;
    if (nbins) *nbins = nb;
    if (counts)
        for (int k = 0; k < nb; ++k) counts[k] = c->bins.start[k + 1] - c->bins.start[k];
    if (gammas && nb > 0) {
        HIPCHK(c, hipMemcpyAsync(gammas, c->gamb.p, (size_t)nb * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return OPALSC_OK;
}

// Buffers the space-charge step of c needs beyond those of create.
static int nonempty_bins(const BinRanges &r) {
    int z = 0;
    for (int k = 0; k < r.nbins; ++k) z += r.start[k + 1] > r.start[k] ? 1 : 0;
    return z;
}

static int prepare_track(opalsc_ctx *c) {
    if (c->bins.nbins > 0 && c->track_sc) {
        if (c->grid.ensure_sums(c->err)) return OPALSC_ENOMEM;
        if (c->grid.ensure_bins(std::max(nonempty_bins(c->bins), 1), c->bin_image, c->err)) return OPALSC_ENOMEM;
        if (!c->scbc.ensure(sizeof(Scalars) * kMaxBins)) FAIL(c, OPALSC_ENOMEM, "bin scalars");
    }
    (void)gamma_blocks(c->grid.fp64, c->bunch.N);  // occupancy query outside any stream capture
    return OPALSC_OK;
}

// One ParallelTTracker step of c on its stream (kernel launches only, so a
// step can be captured into a graph).  push: the step's first half push
// (otherwise the previous step's gather pass applied it); next: the gather
// pass also applies the next step's first half push; sort_now: re-sort the
// bunch by cell after this step's mesh.  Statistics go to the call's slot
// ts->step - ts->call_step0 - 1 of statsb.  overlap: the Green's function may
// run on the context's second stream (a single bunch; a batch already runs
// bunches side by side).
static bool enqueue_step(opalsc_ctx *c, bool push, bool next, bool sort_now, bool overlap, bool boxes) {
    Bunch &b = c->bunch;
    const bool exact = c->grid.fp64;  // reference-order particle arithmetic on the FP64 path
    const bool binned = c->bins.nbins > 0;
    // energy bins whose calcGammas comes from chunk sums (bingamma.hpp): the
    // previous gather pass of this call left them, or they are summed here,
    // before the bounds pass and the sort move particles
    const int64_t gchunk = c->grid.fp64 ? kOsChunk64 : kOsChunk;
    const bool chunked = binned && c->track_sc && c->binchunks.p && bins_chunked(c->bins, gchunk);
    const int ext = c->fmaps.empty() ? (c->track_ext ? 1 : 0) : 2;
    // the gather pass in blocks that run once unless a DC beam or field maps need
    // the persistent loop (2.30 -> 2.20 ms at 1e8 particles, mixed)
    const bool oneshot = !c->grid.pz && ext != 2 && b.N > 0 && c->os_dense.p;
    // a re-sort is applied by the one-shot gather pass (it reads the old
    // arrays through the permutation and writes the new ones), so the deposit
    // reads the old order with its chunk boxes and no array is permuted on
    // its own
    const bool deferred = sort_now && oneshot && b.alt[0] && c->cell_ordered;
    // the second stream takes the Green's function beside the deposit and the
    // x/y passes on grids of 96^3 cells and more (on smaller ones the extra
    // synchronisation costs more than it hides), and then a deferred re-sort
    // (keys and radix passes), which only the gather pass waits for
    const bool side = overlap && !c->prof.enabled && c->stream2 && c->ev_sort && c->grid.ncell >= (size_t)96 * 96 * 96;
    const bool sort_side = deferred && side;
    bool sort_launched = false;
    // the single grid's charge words are cleared on the second stream at the
    // step's start, beside the statistics, bounds and mesh kernels (one block
    // each), instead of between the mesh and the deposit; the previous step's
    // x pass, which read them, is done by then
    const bool zero_side = side && c->track_sc && !binned && c->ev_zero;
    if (zero_side) {
        (void)hipEventRecord(c->ev_start, c->stream);
        (void)hipStreamWaitEvent(c->stream2, c->ev_start, 0);
        c->grid.zero_rho(c->stream2);
        (void)hipEventRecord(c->ev_zero, c->stream2);
    }
    launch_track_pre(c->stream, c->d_ts, c->d_sc, b.N, &c->d_counters[1]);
    if (chunked && !boxes) launch_bin_gamma_chunks(c->stream, c->grid.fp64, b, c->bins, c->binchunks.as<double>());
    c->prof.mark(PROF_SCALAR);
    int bound_rows = alpha_blocks(b.N);  // partial rows of the bounds for launch_track_mesh
    if (c->track_sc && !push && boxes) {
        // the previous gather pass left the statistics of these positions
        launch_bounds_from_boxes(c->stream, c->grid.fp64, b, c->osbox.as<double>(), c->d_sc, c->partials);
        launch_bounds_scan(c->stream, exact, b, c->d_sc, c->partials);
        bound_rows = bounds_scan_blocks(b.N);
        c->prof.mark(PROF_ROTATE_BOUNDS);
    } else if (c->track_sc || push) {
        launch_track_alpha(c->stream, exact, push, b, c->d_sc, c->d_ts, c->partials);
        c->prof.mark(push ? PROF_PUSH_BOUNDS : PROF_ROTATE_BOUNDS);
    }
    if (c->track_sc) {
        launch_track_mesh(c->stream, c->partials, bound_rows, c->cfg.n, c->dh, c->d_ts, exact, !binned, c->d_sc);
        c->prof.mark(PROF_SCALAR);
        if (sort_now && !sort_side) {
            if (deferred) sort_keys(c);
            else sort_bunch(c);  // the bounds pass's chunk boxes no longer match the particle order
            c->prof.mark(PROF_SORT);
        }
        if (sort_now) c->cell_ordered = true;
        const double *boxes_in = sort_now && !deferred ? nullptr : b.chunkbox;
        if (binned) {
            // ParallelTTracker.cpp:567-579: every bin in its own Lorentz frame, and its
            // image; the non-empty bins are solved as one batch (one launch per pass)
            Scalars *scz = c->scbc.as<Scalars>();
            launch_track_bin_scalars(c->stream, b, c->bins, c->bin_qi, exact, c->d_sc, c->binpart.as<double>(),
                                     chunked ? c->binchunks.as<double>() : nullptr, gchunk, c->gamb.as<double>(),
                                     c->scb.as<Scalars>(), scz);
            const int nz = nonempty_bins(c->bins);
            c->prof.mark(PROF_SCALAR);
            // the bins' Green's functions need only their scalars: they run on the
            // second stream beside the deposits and the forward x/y passes
            const bool ov = nz > 0 && side;
            if (ov || sort_side) {
                (void)hipEventRecord(c->ev_mesh, c->stream);
                (void)hipStreamWaitEvent(c->stream2, c->ev_mesh, 0);
            }
            if (ov) {
                c->grid.green_bins(scz, nz, c->bin_image, c->stream2);
                (void)hipEventRecord(c->ev_green, c->stream2);
            }
            if (sort_side) {
                sort_keys(c, nullptr, c->stream2);
                (void)hipEventRecord(c->ev_sort, c->stream2);
                sort_launched = true;
            }
            if (nz > 0) c->grid.zero_bin_rho(nz, c->stream);
            c->prof.mark(PROF_ZERO);
            for (int k = 0, z = 0; k < c->bins.nbins; ++k) {
                const int64_t lo = c->bins.start[k], hi = c->bins.start[k + 1];
                if (hi == lo) continue;  // an empty bin has no charge
                launch_deposit_range(c->stream, 2, exact, lo, hi, b.N, b.x, b.y, b.z, nullptr, nullptr,
                                     c->grid.bin_rho(z), scz + z, &c->d_counters[1], boxes_in);
                ++z;
            }
            c->prof.mark(PROF_DEPOSIT);
            if (nz > 0) {
                if (!ov) c->grid.green_bins(scz, nz, c->bin_image, c->stream);
                c->grid.solve_bins(scz, nz, c->bin_image, ov ? c->ev_green : nullptr);
                c->grid.field_sums(scz, nz, c->bin_image);
            } else {
                c->grid.clear_sums();
            }
        } else {
            // The Green's function needs only the mesh: without profiling it runs
            // on the second stream, beside the deposit and the x and y passes,
            // and the z pass waits for it.
            // (on small grids the extra synchronisation costs more than it hides)
            overlap = side;
            if (overlap) {
                (void)hipEventRecord(c->ev_mesh, c->stream);
                (void)hipStreamWaitEvent(c->stream2, c->ev_mesh, 0);
                c->grid.green(c->d_sc, c->stream2);
                (void)hipEventRecord(c->ev_green, c->stream2);
                if (sort_side) {
                    sort_keys(c, nullptr, c->stream2);
                    (void)hipEventRecord(c->ev_sort, c->stream2);
                    sort_launched = true;
                }
            }
            if (zero_side) (void)hipStreamWaitEvent(c->stream, c->ev_zero, 0);
            else c->grid.zero_rho(c->stream);
            c->prof.mark(PROF_ZERO);
            launch_deposit(c->stream, c->q_uniform ? 3 : 1, exact, b.N, b.x, b.y, b.z, b.q, nullptr,
                           (long long *)c->grid.rho_i64, c->d_sc, &c->d_counters[1], boxes_in);
            c->prof.mark(PROF_DEPOSIT);
            if (!overlap) c->grid.green(c->d_sc);
            c->grid.solve(c->d_sc, 0, 1, false, overlap ? c->ev_green : nullptr);
            c->grid.grad(c->d_sc);
        }
    }
    int nparts = gamma_blocks(c->grid.fp64, b.N);
    // the next step of this call takes its bounds from the statistics this pass leaves
    const bool boxes_out = oneshot && next && c->track_sc && c->osbox.p;
    const bool bsum = binned && c->track_sc;  // the bins' summed fields
    if (oneshot) {
        PermSrc ps{b.vals2, b.x, b.y, b.z, b.px, b.py, b.pz, b.id, b.keys, c->q_uniform ? nullptr : b.q, b.tmp};
        Bunch out = b;  // the arrays the pass writes
        if (deferred)
            for (int k = 0; k < 6; ++k) {
                double **dst[] = {&out.x, &out.y, &out.z, &out.px, &out.py, &out.pz};
                *dst[k] = b.alt[k];
            }
        if (sort_launched) (void)hipStreamWaitEvent(c->stream, c->ev_sort, 0);  // the permutation
        nparts = launch_track_gamma_os(c->stream, c->grid.fp64, next, out, bsum ? c->grid.Esum : c->grid.Eg, c->d_sc,
                                       c->d_ts, c->os_dense.as<double>(), c->partials, ext,
                                       boxes_out ? c->osbox.as<double>() : nullptr,
                                       bsum ? c->grid.Bsum : nullptr, &c->bins,
                                       boxes_out && chunked ? c->binchunks.as<double>() : nullptr,
                                       deferred ? &ps : nullptr);
        if (deferred) {
            // the re-sorted arrays are the bunch's from here on
            ++c->gen;
            double **cur[] = {&b.x, &b.y, &b.z, &b.px, &b.py, &b.pz};
            for (int k = 0; k < 6; ++k) std::swap(*cur[k], b.alt[k]);
            std::swap(b.id, b.keys);
            if (!c->q_uniform) std::swap(b.q, b.tmp);
        }
    } else if (bsum)
        launch_track_gamma(c->stream, c->grid.fp64, next, b, c->grid.Esum, c->d_sc, c->d_ts, c->partials,
                           c->grid.Bsum, ext, false);
    else
        launch_track_gamma(c->stream, c->grid.fp64, next, b, c->grid.Eg, c->d_sc, c->d_ts, c->partials, nullptr,
                           ext, c->grid.pz);
    c->prof.mark(deferred ? PROF_GATHER_RESORT : PROF_GATHER_PUSH);
    launch_track_post(c->stream, c->partials, nparts, c->d_ts, c->d_sc, c->statsb.p, &c->d_counters[1]);
    c->prof.mark(PROF_SCALAR);
    return boxes_out;
}

extern "C" int opalsc_track(opalsc_ctx *c, int nsteps, opalsc_stats *per_step) {
    if (!c || nsteps < 0) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!c->have_track) FAIL(c, OPALSC_ESTATE, "track before track_init");
    if (nsteps == 0) return OPALSC_OK;
    if (!c->statsb.ensure(sizeof(opalsc_stats) * (size_t)nsteps)) FAIL(c, OPALSC_ENOMEM, "stats buffer");
    if (int e = prepare_track(c)) return e;
    // the second stream is made on first use: streams are spread over the
    // device's hardware queues in creation order, and contexts that are only
    // ever stepped in a batch keep one stream each
    if (!c->stream2) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_mesh, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_green, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_sort, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_gimg, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_start, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_zero, hipEventDisableTiming));
    }
    c->prof.start();
    launch_track_begin_call(c->stream, c->d_ts);
    // Within one call the gather pass of step s also applies step s+1's first
    // half push, so the bounds pass of step s+1 only reads positions (or, after
    // a one-shot gather pass, only the chunks that can hold an extreme).  The
    // last step of a call leaves the bunch at its end-of-step state.
    bool boxes = false;
    for (int s = 0; s < nsteps; ++s) {
        const int interval = c->cfg.sort_interval;
        const bool sort_now = c->track_sc && interval > 0 && (c->steps_done % interval) == 0;
        boxes = enqueue_step(c, s == 0, s + 1 < nsteps, sort_now, true, boxes);
        ++c->steps_done;
        CHECK_LAUNCH(c);
    }
    if (per_step)
        HIPCHK(c, hipMemcpyAsync(per_step, c->statsb.p, sizeof(opalsc_stats) * (size_t)nsteps,
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.flush();
    return OPALSC_OK;
}

extern "C" int opalsc_step(opalsc_ctx *c, int nsteps, opalsc_stats *per_step) {
    return opalsc_track(c, nsteps, per_step);
}

// ---------------------------------------------------------------------------
// many independent bunches: one graph of one step of every bunch
// ---------------------------------------------------------------------------
// Fused form (the default): the bunches fall into kernel classes (precision,
// one charge for every particle, external field, space charge, deposit box
// size) and grid groups (grid shape, precision, Green's function).  Every
// particle pass is one launch per class over a device table of BatchItems;
// every Poisson pass is one launch per group over that group's grids (the
// strided per-grid buffers of PoissonGrid's batched solve, owned by the
// batch), each group's Green's function and solve on its own stream.  A step
// of every bunch is then a fixed sequence of a few dozen launches, captured
// once as a HIP graph and replayed per step.  Bunches with energy bins, field
// maps or a DC beam, or a profiled context, take the per-bunch form: one step
// of every bunch as a parallel branch of a graph.
namespace {
struct BatchClass {
    bool exact = false, quni = false, ext = false, sc_on = false, large = false;
    int i0 = 0, n = 0;            // rows [i0, i0 + n) of the item table
    int off[BP_COUNT] = {0};      // each pass's block offsets: first int in the offset table
    int blocks[BP_COUNT] = {0};   // each pass's blocks
};
struct BatchGroup {
    int n[3] = {0, 0, 0};
    bool fp64 = false;
    int greens = 0;
    int nb = 0, sc0 = 0;  // grids of the group; its first Scalars in the batch's array
    PoissonGrid pg;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
};
constexpr int kStageSlots = 8;  // item-table uploads in flight before the host waits
}  // namespace

struct opalsc_batch {
    std::vector<opalsc_ctx *> ctx;
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::string err;
    // per-bunch form
    std::vector<uint64_t> gen;      // generations the graph was captured at
    std::vector<void *> stats_ptr;  // statistics buffers it writes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // fused form
    bool fused_on = true;
    bool plan_ok = false;
    std::vector<long long> sig;       // the bunches' shapes and classes the plan was made for
    std::vector<int> row_ctx;         // context of each item-table row
    std::vector<int> ctx_group, ctx_slot, ctx_scal;  // per context: grid group (-1: none), slot, Scalars index
    std::vector<BatchClass> classes;
    std::vector<BatchGroup *> groups;
    BatchItem *items_d = nullptr;
    int *offs_d = nullptr;
    Scalars *scal_d = nullptr;
    BatchItem *stage = nullptr;  // pinned staging of the item table, kStageSlots copies
    int stage_slot = 0;
    hipEvent_t ev_mesh = nullptr, ev_dep = nullptr;
    // captured steps by (first half push, next step's half push fused): index push + 2 * next
    hipGraph_t fgraph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t fexec[4] = {nullptr, nullptr, nullptr, nullptr};
};

#define BFAIL(b, code, msg) \
    do {                    \
        (b)->err = (msg);   \
        return (code);      \
    } while (0)
#define BHIP(b, call)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) BFAIL(b, OPALSC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

static void batch_drop_graph(opalsc_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void fused_drop_graph(opalsc_batch *b) {
    for (int k = 0; k < 4; ++k) {
        if (b->fexec[k]) (void)hipGraphExecDestroy(b->fexec[k]);
        if (b->fgraph[k]) (void)hipGraphDestroy(b->fgraph[k]);
        b->fexec[k] = nullptr;
        b->fgraph[k] = nullptr;
    }
}

static void fused_release(opalsc_batch *b) {
    fused_drop_graph(b);
    for (BatchGroup *g : b->groups) {
        g->pg.destroy();
        if (g->done) (void)hipEventDestroy(g->done);
        if (g->s) (void)hipStreamDestroy(g->s);
        delete g;
    }
    b->groups.clear();
    b->classes.clear();
    b->row_ctx.clear();
    b->ctx_group.clear();
    b->ctx_slot.clear();
    b->ctx_scal.clear();
    if (b->items_d) (void)hipFree(b->items_d);
    if (b->offs_d) (void)hipFree(b->offs_d);
    if (b->scal_d) (void)hipFree(b->scal_d);
    if (b->stage) (void)hipHostFree(b->stage);
    b->items_d = nullptr;
    b->offs_d = nullptr;
    b->scal_d = nullptr;
    b->stage = nullptr;
    b->stage_slot = 0;
    b->plan_ok = false;
    b->sig.clear();
}

extern "C" int opalsc_batch_create(opalsc_batch **out, opalsc_ctx *const *ctxs, int n) {
    if (!out || !ctxs || n < 1) return OPALSC_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return OPALSC_EINVAL;
        if (ctxs[i]->cfg.device != ctxs[0]->cfg.device) return OPALSC_EINVAL;  // one device per batch
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return OPALSC_EINVAL;  // each bunch once
    }
    if (hipSetDevice(ctxs[0]->cfg.device) != hipSuccess) return OPALSC_EHIP;
    opalsc_batch *b = new (std::nothrow) opalsc_batch();
    if (!b) return OPALSC_ENOMEM;
    b->ctx.assign(ctxs, ctxs + n);
    b->gen.assign(n, ~(uint64_t)0);
    b->stats_ptr.assign(n, nullptr);
    b->join.assign(n, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->master, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_mesh, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_dep, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < n; ++i) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        if (b->master) (void)hipStreamDestroy(b->master);
        hipEvent_t evs[] = {b->fork, b->ev_mesh, b->ev_dep};
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : b->join)
            if (e) (void)hipEventDestroy(e);
        delete b;
        return OPALSC_EHIP;
    }
    *out = b;
    return OPALSC_OK;
}

extern "C" int opalsc_batch_destroy(opalsc_batch *b) {
    if (!b) return OPALSC_OK;
    (void)hipSetDevice(b->ctx[0]->cfg.device);
    if (b->master) (void)hipStreamSynchronize(b->master);
    batch_drop_graph(b);
    fused_release(b);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    hipEvent_t evs[] = {b->fork, b->ev_mesh, b->ev_dep};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (b->master) (void)hipStreamDestroy(b->master);
    delete b;
    return OPALSC_OK;
}

extern "C" const char *opalsc_batch_last_error(const opalsc_batch *b) { return b ? b->err.c_str() : ""; }

extern "C" int opalsc_tune_deposit_set(int64_t max_particles) {
    if (max_particles < 0) return OPALSC_EINVAL;
    set_deposit_large_box_max(max_particles);
    return OPALSC_OK;
}

extern "C" int64_t opalsc_tune_deposit_get(void) { return deposit_large_box_max(); }

extern "C" int opalsc_tune_sort_set(int drop_bits) {
    if (drop_bits < 0 || drop_bits > 24) return OPALSC_EINVAL;
    g_sort_drop = drop_bits;
    return OPALSC_OK;
}

extern "C" int opalsc_tune_sort_get(void) { return g_sort_drop; }

extern "C" int opalsc_batch_set_fused(opalsc_batch *b, int enable) {
    if (!b) return OPALSC_EINVAL;
    b->fused_on = enable != 0;
    return OPALSC_OK;
}

// Captures one step of every bunch (push, no fused next push, no sort) as the
// parallel branches of one graph: each bunch's kernels on its own stream,
// forked from and joined back into the batch's stream.
static int batch_capture(opalsc_batch *b) {
    batch_drop_graph(b);
    const int n = (int)b->ctx.size();
    BHIP(b, hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(b->fork, b->master);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        opalsc_ctx *c = b->ctx[i];
        e = hipStreamWaitEvent(c->stream, b->fork, 0);
        if (e != hipSuccess) break;
        (void)enqueue_step(c, true, false, false, false, false);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(b->join[i], c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->master, b->join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->master, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        BFAIL(b, OPALSC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    for (int i = 0; i < n; ++i) {
        b->gen[i] = b->ctx[i]->gen;
        b->stats_ptr[i] = b->ctx[i]->statsb.p;
    }
    return OPALSC_OK;
}

// ---- fused form ----

// A context the fused step covers: tracking set up, a non-empty bunch, no
// energy bins, field maps or DC beam, the one-shot gather pass available, no
// profiling.
static bool fused_ok(const opalsc_ctx *c) {
    return c->have_track && c->bunch.N > 0 && c->bins.nbins == 0 && c->fmaps.empty() && !c->grid.pz &&
           c->os_dense.p && !c->prof.enabled;
}

// What the plan depends on, per context (a change makes a new plan).
static std::vector<long long> fused_signature(const opalsc_batch *b) {
    std::vector<long long> s;
    for (opalsc_ctx *c : b->ctx) {
        if (!fused_ok(c)) return {};
        s.push_back(c->bunch.N);
        for (int d = 0; d < 3; ++d) s.push_back(c->cfg.n[d]);
        s.push_back((c->grid.fp64 ? 1 : 0) | (c->q_uniform ? 2 : 0) | (c->track_ext ? 4 : 0) |
                    (c->track_sc ? 8 : 0));
        s.push_back(c->grid.greens);
    }
    return s;
}

static BatchView fused_view(const opalsc_batch *b, const BatchClass &k, int pass) {
    BatchView v;
    v.items = b->items_d + k.i0;
    v.n = k.n;
    v.off = pass == BP_ITEM ? nullptr : b->offs_d + k.off[pass];
    v.blocks = pass == BP_ITEM ? k.n : k.blocks[pass];
    return v;
}

static BatchView fused_view_all(const opalsc_batch *b) {
    BatchView v;
    v.items = b->items_d;
    v.n = (int)b->row_ctx.size();
    v.off = nullptr;
    v.blocks = v.n;
    return v;
}

static int fused_plan(opalsc_batch *b, std::vector<long long> &&sig) {
    fused_release(b);
    const int n = (int)b->ctx.size();
    b->ctx_group.assign(n, -1);
    b->ctx_slot.assign(n, 0);
    b->ctx_scal.assign(n, 0);
    // grid groups: the space-charge bunches of one grid shape, precision and Green's function
    for (int i = 0; i < n; ++i) {
        opalsc_ctx *c = b->ctx[i];
        if (!c->track_sc) continue;
        int g = 0;
        for (; g < (int)b->groups.size(); ++g) {
            const BatchGroup *q = b->groups[g];
            if (q->n[0] == c->cfg.n[0] && q->n[1] == c->cfg.n[1] && q->n[2] == c->cfg.n[2] &&
                q->fp64 == c->grid.fp64 && q->greens == c->grid.greens)
                break;
        }
        if (g == (int)b->groups.size()) {
            BatchGroup *q = new (std::nothrow) BatchGroup();
            if (!q) BFAIL(b, OPALSC_ENOMEM, "batch group");
            for (int d = 0; d < 3; ++d) q->n[d] = c->cfg.n[d];
            q->fp64 = c->grid.fp64;
            q->greens = c->grid.greens;
            b->groups.push_back(q);
        }
        b->ctx_group[i] = g;
        b->ctx_slot[i] = b->groups[g]->nb++;
    }
    // the batch's Scalars: each group's grids in slot order, then the bunches without space charge
    int nscal = 0;
    for (BatchGroup *q : b->groups) {
        q->sc0 = nscal;
        nscal += q->nb;
    }
    for (int i = 0; i < n; ++i)
        b->ctx_scal[i] = b->ctx_group[i] >= 0 ? b->groups[b->ctx_group[i]]->sc0 + b->ctx_slot[i] : nscal++;
    // kernel classes: item-table rows ordered by class, then by context
    auto key = [&](int i) {
        const opalsc_ctx *c = b->ctx[i];
        return (c->grid.fp64 ? 1 : 0) | (c->q_uniform ? 2 : 0) | (c->track_ext ? 4 : 0) | (c->track_sc ? 8 : 0) |
               (deposit_large_box(c->bunch.N) ? 16 : 0);
    };
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
    b->row_ctx = order;
    std::vector<int> offs;
    for (int r = 0; r < n;) {
        const int k0 = key(order[r]);
        BatchClass k;
        k.exact = k0 & 1;
        k.quni = (k0 & 2) != 0;
        k.ext = (k0 & 4) != 0;
        k.sc_on = (k0 & 8) != 0;
        k.large = (k0 & 16) != 0;
        k.i0 = r;
        while (r < n && key(order[r]) == k0) ++r;
        k.n = r - k.i0;
        for (int pass = 0; pass < BP_ITEM; ++pass) {
            k.off[pass] = (int)offs.size();
            int acc = 0;
            for (int j = k.i0; j < k.i0 + k.n; ++j) {
                offs.push_back(acc);
                const int64_t N = b->ctx[order[j]]->bunch.N;
                if (pass == BP_ALPHA) acc += alpha_blocks(N);
                else if (pass == BP_DEPOSIT) acc += k.sc_on ? (int)((N + kDepositChunk - 1) / kDepositChunk) : 0;
                else if (pass == BP_GATHER) acc += k.exact ? os_blocks64(N) : os_blocks(N);
                else acc += batch_reduce_blocks(k.exact, N);
            }
            offs.push_back(acc);
            k.blocks[pass] = acc;
        }
        b->classes.push_back(k);
    }
    BHIP(b, hipMalloc((void **)&b->items_d, sizeof(BatchItem) * n));
    BHIP(b, hipMalloc((void **)&b->offs_d, sizeof(int) * offs.size()));
    BHIP(b, hipMalloc((void **)&b->scal_d, sizeof(Scalars) * std::max(nscal, 1)));
    BHIP(b, hipHostMalloc((void **)&b->stage, sizeof(BatchItem) * n * kStageSlots, hipHostMallocDefault));
    BHIP(b, hipMemcpy(b->offs_d, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice));
    for (BatchGroup *q : b->groups) {
        BHIP(b, hipStreamCreateWithFlags(&q->s, hipStreamNonBlocking));
        BHIP(b, hipEventCreateWithFlags(&q->done, hipEventDisableTiming));
        std::string e;
        if (q->pg.init(q->n, q->fp64, q->greens, false, q->s, e) || q->pg.ensure_bins(q->nb, false, e) ||
            q->pg.ensure_bin_fields(e))
            BFAIL(b, OPALSC_ENOMEM, "batch grid group: " + e);
    }
    b->sig = std::move(sig);
    b->plan_ok = true;
    return OPALSC_OK;
}

// The item table of the current bunch arrays, uploaded in stream order on the
// batch's stream.  nobox: contexts whose deposit forms its own chunk boxes
// (re-sorted in this step).
static int fused_upload(opalsc_batch *b, const std::vector<char> *nobox = nullptr) {
    const int n = (int)b->row_ctx.size();
    if (b->stage_slot == kStageSlots) {
        BHIP(b, hipStreamSynchronize(b->master));  // every earlier upload has been read
        b->stage_slot = 0;
    }
    BatchItem *st = b->stage + (size_t)b->stage_slot * n;
    ++b->stage_slot;
    for (int r = 0; r < n; ++r) {
        const int i = b->row_ctx[r];
        opalsc_ctx *c = b->ctx[i];
        const Bunch &u = c->bunch;
        BatchItem it;
        std::memset(&it, 0, sizeof(it));
        it.N = u.N;
        it.x = u.x;
        it.y = u.y;
        it.z = u.z;
        it.px = u.px;
        it.py = u.py;
        it.pz = u.pz;
        it.q = u.q;
        it.chunkbox = (nobox && (*nobox)[i]) ? nullptr : u.chunkbox;
        it.sc = b->scal_d + b->ctx_scal[i];
        it.sc_ctx = c->d_sc;
        it.ts = c->d_ts;
        it.partials = c->partials;
        it.dense = c->os_dense.as<double>();
        it.dropped = &c->d_counters[1];
        it.stats = c->statsb.p;
        for (int d = 0; d < 3; ++d) it.n[d] = c->cfg.n[d];
        it.dh = c->dh;
        const int g = b->ctx_group[i];
        if (g >= 0) {
            it.rho = b->groups[g]->pg.bin_rho(b->ctx_slot[i]);
            it.E = b->groups[g]->pg.bin_E(b->ctx_slot[i]);
        }
        st[r] = it;
    }
    BHIP(b, hipMemcpyAsync(b->items_d, st, sizeof(BatchItem) * n, hipMemcpyHostToDevice, b->master));
    return OPALSC_OK;
}

// One step of every bunch on the batch's stream, the step opalsc_track takes
// inside a call: push, the step's first half push (the call's first step);
// next, the gather pass also applies the next step's (all but the call's last
// step), so a later step's bounds pass reads positions only.  sorted: contexts
// re-sorted by cell after their mesh (then this step is launched directly, not
// captured: the sort moves the bunch arrays).
static int fused_step(opalsc_batch *b, const std::vector<int> &sorted, bool push, bool next) {
    hipStream_t m = b->master;
    for (const BatchClass &k : b->classes) launch_batch_pre(m, fused_view(b, k, BP_ITEM));
    for (const BatchClass &k : b->classes)
        if (push || k.sc_on) launch_batch_alpha(m, fused_view(b, k, BP_ALPHA), k.exact, push);
    for (const BatchClass &k : b->classes)
        if (k.sc_on) launch_batch_mesh(m, fused_view(b, k, BP_ITEM), k.exact);
    if (!sorted.empty()) {
        BHIP(b, hipEventRecord(b->fork, m));
        std::vector<char> nobox(b->ctx.size(), 0);
        for (int i : sorted) {
            opalsc_ctx *c = b->ctx[i];
            BHIP(b, hipStreamWaitEvent(c->stream, b->fork, 0));
            sort_bunch(c, b->scal_d + b->ctx_scal[i], c->stream);
            c->cell_ordered = true;
            BHIP(b, hipEventRecord(b->join[i], c->stream));
            BHIP(b, hipStreamWaitEvent(m, b->join[i], 0));
            nobox[i] = 1;
        }
        if (int e = fused_upload(b, &nobox)) return e;  // the sorted bunches' arrays changed places
    }
    // the Green's functions need only the mesh: each group's on its own stream,
    // beside the deposit (enqueued first, so that it starts first)
    BHIP(b, hipEventRecord(b->ev_mesh, m));
    for (BatchGroup *g : b->groups) g->pg.zero_bin_rho(g->nb, m);
    for (const BatchClass &k : b->classes)
        if (k.sc_on) launch_batch_deposit(m, fused_view(b, k, BP_DEPOSIT), k.exact, k.quni, k.large);
    BHIP(b, hipEventRecord(b->ev_dep, m));
    for (BatchGroup *g : b->groups) {
        BHIP(b, hipStreamWaitEvent(g->s, b->ev_mesh, 0));
        g->pg.green_bins(b->scal_d + g->sc0, g->nb, false, g->s);
    }
    for (BatchGroup *g : b->groups) {
        BHIP(b, hipStreamWaitEvent(g->s, b->ev_dep, 0));
        g->pg.solve_bins(b->scal_d + g->sc0, g->nb, false, nullptr);
        g->pg.grad_bins(b->scal_d + g->sc0, g->nb);
        BHIP(b, hipEventRecord(g->done, g->s));
        BHIP(b, hipStreamWaitEvent(m, g->done, 0));
    }
    for (const BatchClass &k : b->classes)
        launch_batch_gather(m, fused_view(b, k, BP_GATHER), k.exact, k.ext, next);
    for (const BatchClass &k : b->classes) launch_batch_reduce(m, fused_view(b, k, BP_REDUCE), k.exact);
    for (const BatchClass &k : b->classes) launch_batch_post(m, fused_view(b, k, BP_ITEM), k.exact);
    BHIP(b, hipGetLastError());
    if (!sorted.empty())
        if (int e = fused_upload(b)) return e;  // later steps take their boxes from the bounds pass again
    return OPALSC_OK;
}

static int fused_capture(opalsc_batch *b, bool push, bool next) {
    const int v = (push ? 1 : 0) + (next ? 2 : 0);
    BHIP(b, hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal));
    const int e = fused_step(b, {}, push, next);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->master, &g);
    if (e != OPALSC_OK || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (e != OPALSC_OK) return e;
        BFAIL(b, OPALSC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e2));
    }
    b->fgraph[v] = g;
    BHIP(b, hipGraphInstantiate(&b->fexec[v], b->fgraph[v], nullptr, nullptr, 0));
    return OPALSC_OK;
}

static int fused_run(opalsc_batch *b, int nsteps, bool use_graph) {
    hipStream_t m = b->master;
    const int n = (int)b->ctx.size();
    if (int e = fused_upload(b)) return e;
    launch_batch_begin(m, fused_view_all(b));
    for (int s = 0; s < nsteps; ++s) {
        std::vector<int> sorted;
        for (int i = 0; i < n; ++i) {
            const opalsc_ctx *c = b->ctx[i];
            const int interval = c->cfg.sort_interval;
            if (c->track_sc && interval > 0 && (c->steps_done + s) % interval == 0) sorted.push_back(i);
        }
        const bool push = s == 0, next = s + 1 < nsteps;
        const int v = (push ? 1 : 0) + (next ? 2 : 0);
        if (sorted.empty() && use_graph) {
            if (!b->fexec[v])
                if (int e = fused_capture(b, push, next)) return e;
            BHIP(b, hipGraphLaunch(b->fexec[v], m));
        } else if (int e = fused_step(b, sorted, push, next)) {
            return e;
        }
    }
    launch_batch_end(m, fused_view_all(b));
    BHIP(b, hipGetLastError());
    return OPALSC_OK;
}

extern "C" int opalsc_step_batched(opalsc_batch *b, int nsteps, opalsc_stats *per_step, int use_graph) {
    if (!b || nsteps < 0) return OPALSC_EINVAL;
    if (nsteps == 0) return OPALSC_OK;
    const int n = (int)b->ctx.size();
    BHIP(b, hipSetDevice(b->ctx[0]->cfg.device));
    bool recapture = !b->exec;
    for (int i = 0; i < n; ++i) {
        opalsc_ctx *c = b->ctx[i];
        if (!c->have_track) BFAIL(b, OPALSC_ESTATE, "bunch " + std::to_string(i) + ": step before track_init");
        if (c->prof.enabled) BFAIL(b, OPALSC_ESTATE, "bunch " + std::to_string(i) + ": profiling is on");
        if (!c->statsb.ensure(sizeof(opalsc_stats) * (size_t)nsteps))
            BFAIL(b, OPALSC_ENOMEM, "bunch " + std::to_string(i) + ": stats buffer");
        if (prepare_track(c) != OPALSC_OK) BFAIL(b, OPALSC_ENOMEM, "bunch " + std::to_string(i) + ": " + c->err);
        recapture |= b->gen[i] != c->gen || b->stats_ptr[i] != c->statsb.p;
    }
    std::vector<long long> sig = b->fused_on ? fused_signature(b) : std::vector<long long>();
    const bool fused = !sig.empty();
    // every bunch's earlier work on its own stream comes first
    for (int i = 0; i < n; ++i) {
        opalsc_ctx *c = b->ctx[i];
        if (!fused) {
            launch_track_begin_call(c->stream, c->d_ts);
            BHIP(b, hipGetLastError());
        }
        BHIP(b, hipEventRecord(b->join[i], c->stream));
        BHIP(b, hipStreamWaitEvent(b->master, b->join[i], 0));
    }
    if (fused) {
        if (!b->plan_ok || sig != b->sig)
            if (int e = fused_plan(b, std::move(sig))) return e;
        if (int e = fused_run(b, nsteps, use_graph != 0)) return e;
        BHIP(b, hipEventRecord(b->fork, b->master));
        for (int i = 0; i < n; ++i) BHIP(b, hipStreamWaitEvent(b->ctx[i]->stream, b->fork, 0));
    } else if (use_graph) {
        if (recapture)
            if (int e = batch_capture(b)) return e;
        for (int s = 0; s < nsteps; ++s) BHIP(b, hipGraphLaunch(b->exec, b->master));
        // later work on every bunch's stream follows the graph
        BHIP(b, hipEventRecord(b->fork, b->master));
        for (int i = 0; i < n; ++i) BHIP(b, hipStreamWaitEvent(b->ctx[i]->stream, b->fork, 0));
    } else {
        for (int s = 0; s < nsteps; ++s)
            for (int i = 0; i < n; ++i) {
                (void)enqueue_step(b->ctx[i], true, false, false, false, false);
                BHIP(b, hipGetLastError());
            }
    }
    BHIP(b, hipStreamSynchronize(b->master));
    for (int i = 0; i < n; ++i) {
        opalsc_ctx *c = b->ctx[i];
        c->steps_done += nsteps;
        BHIP(b, hipStreamSynchronize(c->stream));
        if (per_step) {
            std::vector<opalsc_stats> st((size_t)nsteps);
            BHIP(b, hipMemcpy(st.data(), c->statsb.p, sizeof(opalsc_stats) * (size_t)nsteps, hipMemcpyDeviceToHost));
            for (int s = 0; s < nsteps; ++s) per_step[(size_t)s * n + i] = st[s];
        }
    }
    return OPALSC_OK;
}

// One line of the DataSink statistics file for the resident bunch after
// opalsc_track: calcBeamParameters (PartBunchBase.hpp:1529-1613) with the
// dispersion moments and the energy spread, get_bounds' maximum, sum(Q), the
// tracker's time, path length and reference particle, the external field at
// the reference particle at t - dT/2 (ParallelTTracker::writePhaseSpace,
// ParallelTTracker.cpp:1030-1040) and the first particle set (R[0], P[0]).
extern "C" int opalsc_stat_row_get(opalsc_ctx *c, opalsc_stat_row *row) {
    if (!c || !row) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!c->have_track) FAIL(c, OPALSC_ESTATE, "stat_row_get needs a tracked bunch (opalsc_track_init)");
    const Bunch &b = c->bunch;
    const int64_t N = b.N;
    memset(row, 0, sizeof(*row));
    TrackState ts;
    HIPCHK(c, hipMemcpyAsync(&ts, c->d_ts, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    row->t_ns = ts.t * 1e9;
    row->s = ts.pathLength;
    row->num_particles = N;
    for (int d = 0; d < 3; ++d) {
        row->ref_r[d] = ts.RefPartR[d];
        row->ref_p[d] = ts.RefPartP[d];
    }
    double E[3], B[3];
    ext_fields(ts, ts.t - 0.5 * ts.dT, E, B);
    for (int d = 0; d < 3; ++d) {
        row->B_ref[d] = B[d];
        row->E_ref[d] = E[d] * 1e-6;
    }
    row->dt_ns = ts.dT * 1e9;
    if (N == 0) return OPALSC_OK;
    if (!c->statsb.ensure(sizeof(StatsOut)) || !c->tmpd.ensure(16 * sizeof(double)))
        FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    launch_moments_soa(c->stream, b, c->partials);
    launch_stats_from_partials(c->stream, c->partials, red_blocks(N), N, ts.mass_eV, c->statsb.p);
    CHECK_LAUNCH(c);
    opalsc_stats st;
    HIPCHK(c, hipMemcpyAsync(&st, c->statsb.p, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    launch_stat_extra(c->stream, b, c->q_uniform, c->q_value, c->partials, c->tmpd.as<double>());
    CHECK_LAUNCH(c);
    double ex[14];
    HIPCHK(c, hipMemcpyAsync(ex, c->tmpd.p, sizeof(ex), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double Nd = (double)N;
    row->charge = ex[4];
    row->energy_MeV = st.ekin_MeV;
    for (int d = 0; d < 3; ++d) {
        row->rms_r[d] = st.rrms[d];
        row->rms_p[d] = st.prms[d];
        row->emit[d] = st.eps_norm[d];
        row->mean_r[d] = st.rmean[d];
        row->max_r[d] = ex[5 + d];
        row->corr[d] = st.rprms[d];
        row->R0[d] = ex[8 + d];
        row->P0[d] = ex[11 + d];
    }
    row->Dx = ex[0] / Nd;
    row->DDx = ex[1] / Nd;
    row->Dy = ex[2] / Nd;
    row->DDy = ex[3] / Nd;
    // energy spread from the longitudinal momentum spread (PartBunchBase.hpp:1595-1604)
    const double m0 = ts.mass_eV * 1.E-6;
    const double tmp = 1.0 / std::pow(st.ekin_MeV / m0 + 1., 2.0);
    row->dE_MeV = st.prms[2] * m0 * std::sqrt(1.0 - tmp);
    row->has_particle0 = 1;
    return OPALSC_OK;
}

extern "C" int opalsc_track_get_params(opalsc_ctx *c, opalsc_track_params *out) {
    if (!c || !out) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (!c->have_track) FAIL(c, OPALSC_ESTATE, "no tracker state");
    TrackState ts;
    HIPCHK(c, hipMemcpyAsync(&ts, c->d_ts, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->dT = ts.dT;
    out->charge = ts.charge;
    out->mass_eV = ts.mass_eV;
    out->coupling = ts.coupling;
    out->Bz = ts.Bz;
    out->Ez0 = ts.Ez0;
    out->rf_freq = ts.rf_freq;
    out->rf_phase = ts.rf_phase;
    out->t = ts.t;
    out->pathLength = ts.pathLength;
    for (int d = 0; d < 3; ++d) {
        out->RefPartR[d] = ts.RefPartR[d];
        out->RefPartP[d] = ts.RefPartP[d];
    }
    out->space_charge = ts.space_charge;
    return OPALSC_OK;
}

extern "C" int opalsc_get_mesh(opalsc_ctx *c, double hr[3], double origin[3]) {
    if (!c || !hr || !origin) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    Scalars s;
    if (int e = download_scalars(c, s)) return e;
    for (int d = 0; d < 3; ++d) {
        hr[d] = s.hr[d];
        origin[d] = s.origin[d];
    }
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// correctly rounded math hooks
// ---------------------------------------------------------------------------
extern "C" int opalsc_host_atan_cr(int64_t n, const double *x, double *y) {
    if (n < 0 || (n > 0 && (!x || !y))) return OPALSC_EINVAL;
    for (int64_t i = 0; i < n; ++i) y[i] = crm::atan_cr(x[i]);
    return OPALSC_OK;
}
extern "C" int opalsc_host_log_cr(int64_t n, const double *x, double *y) {
    if (n < 0 || (n > 0 && (!x || !y))) return OPALSC_EINVAL;
    for (int64_t i = 0; i < n; ++i) y[i] = crm::log_cr(x[i]);
    return OPALSC_OK;
}
static int device_cr(opalsc_ctx *c, int which, int64_t n, const double *x, double *y) {
    if (!c || n < 0 || (n > 0 && (!x || !y))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (n == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosR, x, (size_t)n * 8)) return e;
    if (!c->aosP.ensure((size_t)n * 8)) FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    launch_cr_test(c->stream, which, n, c->aosR.as<double>(), c->aosP.as<double>());
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(y, c->aosP.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}
extern "C" int opalsc_device_atan_cr(opalsc_ctx *c, int64_t n, const double *x, double *y) {
    return device_cr(c, 0, n, x, y);
}
extern "C" int opalsc_device_log_cr(opalsc_ctx *c, int64_t n, const double *x, double *y) {
    return device_cr(c, 1, n, x, y);
}
extern "C" int opalsc_host_igf_term(int64_t n, const double *v, int fast, double *y) {
    if (n < 0 || (n > 0 && (!v || !y))) return OPALSC_EINVAL;
    for (int64_t i = 0; i < n; ++i)
        y[i] = fast ? igf::term_fast(v[3 * i], v[3 * i + 1], v[3 * i + 2], 1.0)
                    : igf::term_cr(v[3 * i], v[3 * i + 1], v[3 * i + 2], 1.0);
    return OPALSC_OK;
}
extern "C" int opalsc_host_green_octant(const int n[3], const double h[3], double zshift, int table, int fast,
                                        double *out) {
    if (!n || !h || !out || table < 0 || table > 2 || n[0] < 1 || n[1] < 1 || n[2] < 1) return OPALSC_EINVAL;
    const int g0 = n[0] + 1, g1 = n[1] + 1, g2 = n[2] + 1;
    const double vol = h[0] * h[1] * h[2], ivol = 1.0 / vol;
    const bool shifted = table != 0, second = table == 2;
    if (fast) {
        const igf::GreenMoments m = igf::green_moments(h[0], h[1], h[2]);
        for (int k = 0; k < g2; ++k)
            for (int j = 0; j < g1; ++j)
                for (int i = 0; i < g0; ++i) {
                    const int kq = (!shifted && i == 0 && j == 0 && k == 0) ? 1 : k;
                    out[((size_t)k * g1 + j) * g0 + i] =
                        igf::green_point_fast(m, i, j, kq, h[0], h[1], h[2], zshift, shifted, second, n[2], ivol);
                }
        return OPALSC_OK;
    }
    // the corner table and its 8-corner difference, as the FP64 kernels evaluate them
    const int m0 = n[0] + 2, m1 = n[1] + 2, m2 = n[2] + 2;
    std::vector<double> T((size_t)m0 * m1 * m2);
    for (int k = 0; k < m2; ++k)
        for (int j = 0; j < m1; ++j)
            for (int i = 0; i < m0; ++i) {
                const double v0 = i * h[0] - h[0] / 2, v1 = j * h[1] - h[1] / 2;
                double v2 = k * h[2] - h[2] / 2;
                if (shifted) {
                    v2 = v2 + zshift;
                    if (second) v2 = v2 - n[2] * h[2];
                }
                T[((size_t)k * m1 + j) * m0 + i] = igf::term_cr(v0, v1, v2, vol);
            }
    auto t = [&](int i, int j, int k) { return T[((size_t)k * m1 + j) * m0 + i]; };
    for (int k = 0; k < g2; ++k)
        for (int j = 0; j < g1; ++j)
            for (int i = 0; i < g0; ++i) {
                double g = t(i + 1, j + 1, k + 1);
                g += t(i + 1, j, k);
                g += t(i, j + 1, k);
                g += t(i, j, k + 1);
                g -= t(i + 1, j + 1, k);
                g -= t(i + 1, j, k + 1);
                g -= t(i, j + 1, k + 1);
                g -= t(i, j, k);
                out[((size_t)k * g1 + j) * g0 + i] = g;
            }
    if (!shifted) out[0] = out[(size_t)g1 * g0];  // G[0,0,0] := G[0,0,1]
    return OPALSC_OK;
}

extern "C" int opalsc_device_green_octant(opalsc_ctx *c, const double h[3], double zshift, int table, double *out) {
    if (!c || !h || !out || table < 0 || table > 2) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (c->grid.greens == 1) FAIL(c, OPALSC_EINVAL, "the context's Green's function is not the integrated one");
    if (table != 0 && c->grid.pz) FAIL(c, OPALSC_EINVAL, "a periodic-z context has no image-charge tables");
    Scalars s;
    std::memset(&s, 0, sizeof(s));
    for (int d = 0; d < 3; ++d) s.hs[d] = h[d];
    s.zshift = zshift;
    if (int e = upload(c, c->aosR, &s, sizeof(s))) return e;
    c->grid.octant(c->aosR.as<Scalars>(), table, c->stream);
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(out, c->grid.Goct, c->grid.noct * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

extern "C" int opalsc_device_igf_term(opalsc_ctx *c, int64_t n, const double *v, int fast, double *y) {
    if (!c || n < 0 || (n > 0 && (!v || !y))) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    if (n == 0) return OPALSC_OK;
    if (int e = upload(c, c->aosR, v, (size_t)n * 24)) return e;
    if (!c->aosP.ensure((size_t)n * 8)) FAIL(c, OPALSC_ENOMEM, "staging allocation failed");
    launch_igf_term_test(c->stream, fast, n, c->aosR.as<double>(), c->aosP.as<double>());
    CHECK_LAUNCH(c);
    HIPCHK(c, hipMemcpyAsync(y, c->aosP.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OPALSC_OK;
}

// ---------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------
extern "C" int opalsc_profile_enable(opalsc_ctx *c, int enable) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    c->prof.flush();
    c->prof.enabled = enable != 0;
    for (int i = 0; i < PROF_COUNT; ++i) {
        c->prof.ms[i] = 0.0;
        c->prof.launches[i] = 0;
    }
    return OPALSC_OK;
}

extern "C" int opalsc_copy_bandwidth(opalsc_ctx *c, size_t bytes, int reps, double *gbs) {
    if (!c || !gbs || reps <= 0 || bytes < 16) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) FAIL(c, OPALSC_ENOMEM, "copy buffer allocation failed");
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        FAIL(c, OPALSC_ENOMEM, "copy buffer allocation failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemsetAsync(a, 1, bytes, c->stream);
    launch_copy_calibration(c->stream, bytes, a, b);  // warm-up
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch_copy_calibration(c->stream, bytes, a, b);
    (void)hipEventRecord(e1, c->stream);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) FAIL(c, OPALSC_EHIP, "copy calibration failed");
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return OPALSC_OK;
}

extern "C" int opalsc_profile_read(opalsc_ctx *c, int cap, const char **names, double *ms, int64_t *launches) {
    if (!c) return OPALSC_EINVAL;
    if (int e_ = use_device(c)) return e_;
    c->prof.flush();
    int n = std::min<int>(cap, PROF_COUNT);
    for (int i = 0; i < n; ++i) {
        if (names) names[i] = prof_name(i);
        if (ms) ms[i] = c->prof.ms[i];
        if (launches) launches[i] = c->prof.launches[i];
    }
    return PROF_COUNT;
}

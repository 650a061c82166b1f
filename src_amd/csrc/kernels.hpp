// Host-side launch interfaces between the C ABI and the kernels.
#pragma once
#include "common.hpp"
#include "fft.hpp"

#include <string>
#include <vector>

namespace opalsc {

void set_global_error(const std::string &msg);

// Every byte of [p, p + bytes) is device (or managed) memory: the pointer's
// attributes, then the extent of the allocation that holds it.  Allocates
// nothing.
inline bool device_range_ok(const void *p, size_t bytes) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // memory from a pool whose extent the runtime does not report
    }
    return (const char *)p + bytes <= (const char *)base + size;
}

// A parsed 2D field map (host), FieldMapDesc offsets relative to its own grids.
struct FieldMapHost {
    FieldMapDesc d;
    double frequency;  // the map file's frequency (rad/s; dynamic maps)
    std::vector<double> a, b, c;
};
// FM2DMagnetoStatic / FM2DDynamic file reading (io.cpp); 0 or -1 with err set.
int parse_fieldmap_2d(const char *path, FieldMapHost &out, std::string &err);


enum ProfId {
    PROF_PUSH_BOUNDS = 0,  // push + beam-frame rotation + bounds
    PROF_SCALAR,           // single-thread scalar kernels
    PROF_SORT,             // cell keys, radix sort, permutation
    PROF_ZERO,             // clear charge grid
    PROF_DEPOSIT,          // CIC scatter (int64 fixed point)
    PROF_GREEN,            // IGF / standard Green's function octant
    PROF_DCT,              // DCT-I of the octant -> Ghat
    PROF_FFT_X,            // r2c / c2r along x
    PROF_FFT_Y,            // along y
    PROF_FFT_Z,            // z: forward * Ghat * inverse
    PROF_GRAD,             // E = -Grad(phi)
    PROF_GATHER_PUSH,      // gather + kick + push + frame update + moments
    PROF_ROTATE_BOUNDS,    // beam-frame rotation + bounds of already pushed positions
    PROF_GATHER_RESORT,    // the gather pass that also applies a re-sort (reads through the permutation)
    PROF_COUNT
};

const char *prof_name(int id);

// Accumulates HIP-event intervals between consecutive marks on one stream.
struct Profiler {
    bool enabled = false;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<int> ids;
    int used = 0;
    double ms[PROF_COUNT] = {0};
    long long launches[PROF_COUNT] = {0};
    void start();
    void mark(int id);
    void flush();  // synchronises and folds the recorded intervals into ms[]
    void destroy();
};

struct PoissonGrid {
    int n[3] = {0, 0, 0}, L[3] = {0, 0, 0}, K0 = 0;
    int kp = 0;   // row pitch (complex elements) of A and B: K0 rounded up to whole tiles
    int gp = 0;   // row pitch (real elements) of Ghat, >= kp
    bool fp64 = true;
    bool pz = false;  // periodic in z (z not doubled): the DC-beam solver
    int greens = 0;
    size_t ncell = 0, noct = 0;
    FftPlan plan[3];
    void *twd[3] = {nullptr, nullptr, nullptr};
    void *rho_i64 = nullptr;  // fixed-point charge: long long[ncell] (mixed), [ncell][2] hi/lo (FP64)
    void *rho_f64 = nullptr;  // double[ncell] (compute_potential input)
    void *A = nullptr, *B = nullptr;
    void *Goct = nullptr;     // double[noct]: G on the octant
    void *Tcorner = nullptr;  // double[(n0+2)(n1+2)(n2+2)]: IGF corner antiderivatives
    void *Ghat = nullptr;     // T[n2+1][n1+1][gp]
    void *phi = nullptr;      // T[ncell]
    void *Eg = nullptr;       // vec4<T>[ncell]
    // image-charge Green's function (allocated on first use): the two shifted
    // octant tables after the x and y DCTs (T[n2+1][n1+1][gp] each) and the
    // transform over the doubled z (complex T[2 n2][n1+1][gp])
    void *HimgA = nullptr, *HimgB = nullptr, *Gimg = nullptr;
    // energy bins (allocated on first use): sums over bins of E and of the
    // magnetic field it implies, vec4<T>[ncell] each
    void *Esum = nullptr, *Bsum = nullptr;
    // periodic z: the Green's function transform, complex T[n2][n1+1][gp] (not even in z)
    void *Gper = nullptr;
    hipStream_t stream = nullptr;
    Profiler *prof = nullptr;

    int init(const int nn[3], bool fp64, int greens, bool periodic_z, hipStream_t s, std::string &err);
    void destroy();
    // the FP64 grid evaluates the reference's corner table with correctly rounded
    // atan/log, the mixed grid the one-pass octant (igfmath.hpp)
    void green(const Scalars *d_sc, hipStream_t s = nullptr);  // s: default stream
    // only the real octant G of table 0 (plain) or 1/2 (the image charge's) into Goct
    void octant(const Scalars *d_sc, int table, hipStream_t s = nullptr);
    int ensure_image(std::string &err);
    void green_image(const Scalars *d_sc, hipStream_t s = nullptr);  // shift: Scalars::zshift
    // src_kind 0: int64 fixed-point rho_i64 (self-field scaling); 1: rho_f64 as is
    // outscale 1: phi *= vol*coupling (self fields); 0: potential only
    // image: convolve with the image-charge Green's function (negated, z flipped)
    // green_ready: the z pass waits for it (Green's function computed on another stream)
    void solve(const Scalars *d_sc, int src_kind, int outscale, bool image = false, hipEvent_t green_ready = nullptr);
    void grad(const Scalars *d_sc);
    int ensure_sums(std::string &err);
    void clear_sums();
    void rho_to_double(const Scalars *d_sc, double *out);
    size_t rho_words() const { return fp64 ? 2 * ncell : ncell; }
    void zero_rho(hipStream_t s);  // clears every word of rho_i64

    // Energy bins solved as one batch: nb grids of this shape, each with its own
    // fixed-point charge, spectra, potentials (the bin's and its image's) and
    // Green's function tables; every pass is one launch for all of them.
    int bin_cap = 0;
    bool bin_img_cap = false;
    void *bb_rho = nullptr, *bb_A = nullptr, *bb_B = nullptr, *bb_Bi = nullptr, *bb_phi = nullptr,
         *bb_phii = nullptr, *bb_T = nullptr, *bb_G = nullptr, *bb_Gh = nullptr, *bb_HA = nullptr,
         *bb_HB = nullptr, *bb_Gi = nullptr;
    int ensure_bins(int nb, bool image, std::string &err);
    long long *bin_rho(int b) const { return (long long *)bb_rho + (size_t)b * rho_words(); }
    void zero_bin_rho(int nb, hipStream_t s);
    // scb: nb Scalars, one per grid of the batch
    void green_bins(const Scalars *scb, int nb, bool image, hipStream_t s);
    void solve_bins(const Scalars *scb, int nb, bool image, hipEvent_t green_ready);
    // Esum, Bsum := the sums over the nb grids (and their images) of E and of
    // the magnetic field it implies (written, not accumulated)
    void field_sums(const Scalars *scb, int nb, bool image);
    // Independent bunches of this grid shape solved as one batch (the same
    // per-grid buffers): E of every grid, vec4<T>[ncell] each, one launch
    void *bb_E = nullptr;
    int ensure_bin_fields(std::string &err);  // bb_E for bin_cap grids
    void *bin_E(int b) const { return (char *)bb_E + (size_t)b * ncell * 4 * (fp64 ? 8 : 4); }
    void grad_bins(const Scalars *scb, int nb);
    void widen(int which, double *out, const void *efield = nullptr);  // efield: another E grid (default Eg)
    void profile_mark(int id) {
        if (prof) prof->mark(id);
    }
};

constexpr int kSlots = 64;           // doubles per block in a partial-sums buffer
constexpr int kBoxStats = 12;        // per chunk: min[3], max[3], sum[3], sum of squares[3]
constexpr int kOsBoxStats = 6;       // per one-shot gather block: min[3], max[3] of the positions written
// particles per thread and per block of the one-shot gather passes (mixed, FP64)
constexpr int kOsU = 4;
constexpr int kOsChunk = kThreads * kOsU;
constexpr int kOsU64 = 2;
constexpr int kOsChunk64 = kThreads * kOsU64;

// Particle buffers in structure-of-arrays form.
struct Bunch {
    int64_t cap = 0, N = 0;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *q = nullptr;
    uint32_t *id = nullptr;
    double *tmp = nullptr;       // permutation scratch
    double *qd = nullptr;        // Q*dt (self-field entry point)
    double *chunkbox = nullptr;  // per-chunk beam-frame position statistics (kBoxStats doubles)
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr;
    // tracker re-sorts applied by the gather pass: the arrays it writes x, y, z, px, py, pz into
    double *alt[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
};

// ---- Batched steps of many independent bunches --------------------------------
// One bunch of a batched step: its particle arrays, per-bunch state and the
// grid slot of its shape group.  api.hip keeps a device table of these; each
// particle pass is one launch over the bunches of a kernel class, a flat block
// index finding its bunch through the pass's block offsets.
struct BatchItem {
    int64_t N;
    double *x, *y, *z, *px, *py, *pz, *q;
    double *chunkbox;
    Scalars *sc;       // the bunch's scalars during the call (the batch's per-group array)
    Scalars *sc_ctx;   // the context's own scalars (copied in and out around a call)
    TrackState *ts;
    double *partials;  // per-block partial sums
    double *dense;     // per-block moment sums of the one-shot gather pass
    long long *rho;    // the bunch's fixed-point charge grid
    const void *E;     // the bunch's field grid (vec4 of the grid precision)
    unsigned long long *dropped;
    void *stats;       // the call's statistics array
    int n[3];
    double dh;
};
enum BatchPass { BP_ALPHA = 0, BP_DEPOSIT, BP_GATHER, BP_REDUCE, BP_ITEM, BP_COUNT };
// The bunches of one kernel class in a pass (passed by value): items[0..n),
// off[j] = first flat block of item j (device), off[n] = blocks = the blocks of
// the launch (host copy).
struct BatchView {
    const BatchItem *items;
    const int *off;
    int n;
    int blocks;
};
// flat block -> (item, block within the item, the item's block count)
__device__ inline int batch_find(const BatchView &bv, int &bid, int &nblk) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = bv.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bv.off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    bid = b - bv.off[lo];
    nblk = bv.off[lo + 1] - bv.off[lo];
    return lo;
}

// Launchers of the batched particle passes (particles.hip).  Class flags:
// exact (FP64 arithmetic), quni (one charge for every particle), ext (an
// external field), large (the deposit's larger LDS box).
void launch_batch_begin(hipStream_t s, const BatchView &v);          // call_step0, scalars in
void launch_batch_end(hipStream_t s, const BatchView &v);            // scalars out
void launch_batch_pre(hipStream_t s, const BatchView &v);
// push: the step's first half push (else the previous gather pass applied it)
void launch_batch_alpha(hipStream_t s, const BatchView &v, bool exact, bool push);
void launch_batch_mesh(hipStream_t s, const BatchView &v, bool exact);
void launch_batch_deposit(hipStream_t s, const BatchView &v, bool exact, bool quni, bool large);
// next: also the next step's first half push
void launch_batch_gather(hipStream_t s, const BatchView &v, bool exact, bool ext, bool next);
void launch_batch_reduce(hipStream_t s, const BatchView &v, bool exact);
void launch_batch_post(hipStream_t s, const BatchView &v, bool exact);
// partial rows the batched gather pass of an N-particle bunch leaves for the post pass
int batch_reduce_blocks(bool exact, int64_t N);
bool deposit_large_box(int64_t n);  // the deposit of n particles uses the larger LDS box
void set_deposit_large_box_max(int64_t n);
int64_t deposit_large_box_max();

// A re-sort applied by the one-shot gather pass itself: particle i of the
// sorted order is read from the old arrays at perm[i] and written at i of the
// pass's (new) arrays; id (and Q when it is not one number) move along.
struct PermSrc {
    const uint32_t *perm;
    const double *x, *y, *z, *px, *py, *pz;
    const uint32_t *id;
    uint32_t *id_out;
    const double *q;  // null: the bunch keeps Q as one number
    double *q_out;
};

// Launchers (particles.hip)
int red_blocks(int64_t N);
void launch_soa_from_aos(hipStream_t s, int64_t N, const double *R, double *x, double *y, double *z);
void launch_aos_from_soa(hipStream_t s, int64_t N, const double *x, const double *y, const double *z, double *R);
void launch_selffield_prep(hipStream_t s, int64_t N, const double *P, double *Q, const double *dt,
                           double *qd, double *partials, unsigned long long *qdmax_bits);
void launch_selffield_scalars(hipStream_t s, const double *partials, int nb, int64_t N, double dT,
                              const double origin[3], const double hr[3], const int n[3],
                              double coupling, const unsigned long long *qdmax_bits, bool exact,
                              bool pz, Scalars *sc);
void launch_deposit(hipStream_t s, int mode, bool exact, int64_t N, const double *x, const double *y,
                    const double *z, double *q, const double *qd, long long *rho, const Scalars *sc,
                    unsigned long long *dropped, const double *chunkbox);
// deposit of particles lo <= p < hi of an Nall-particle bunch (chunkbox indexed over the bunch)
void launch_deposit_range(hipStream_t s, int mode, bool exact, int64_t lo, int64_t hi, int64_t Nall,
                          const double *x, const double *y, const double *z, double *q, const double *qd,
                          long long *rho, const Scalars *sc, unsigned long long *dropped,
                          const double *chunkbox);
void launch_gather_selffield(hipStream_t s, bool fp64, int64_t N, const double *x, const double *y,
                             const double *z, const void *Eg, const Scalars *sc, double *Ef,
                             double *Bf);
void launch_cic_index(hipStream_t s, int64_t N, const double *R, const double origin[3],
                      const double hr[3], int32_t *lgp, double *dpos);
void launch_bounds(hipStream_t s, int64_t N, const double *R, double *partials);
// pz: DC beam (z left to the period length, PartBunchBase.hpp:840-843)
void launch_boundp_scalars(hipStream_t s, const double *partials, int nb, const int n[3], double dh,
                           bool pz, double period, Scalars *sc);
// ParticlePeriodicBCond in z over [origin, n*h + origin) (AoS R, in place)
void launch_wrap_periodic_z(hipStream_t s, int64_t N, double *R, double lo, double hi);
// every field map at lab positions R (AoS) and time t into E, B (overwritten)
void launch_fieldmap_eval(hipStream_t s, int64_t N, const double *R, double t, const FieldMapDesc *maps, int nmaps,
                          const double *data, double *E, double *B, unsigned long long *lost);
void launch_push(hipStream_t s, int64_t N, double *R, const double *P, const double *dt);
void launch_kick(hipStream_t s, int64_t N, double *P, const double *Ef, const double *Bf,
                 const double *dt, double charge, double mass);
void launch_moments(hipStream_t s, int64_t N, const double *R, const double *P, double *partials);
// statistics-file extras of the resident bunch: out[0..4] = sums of x*pz,
// px*pz, y*pz, py*pz, Q; out[5..7] = max of x, y, z; out[8..13] = R, P of the
// particle with id 0
void launch_stat_extra(hipStream_t s, const Bunch &b, bool q_uniform, double q_value, double *partials,
                       double *out);
void launch_stats_from_partials(hipStream_t s, const double *partials, int nb, int64_t N,
                                double mass_eV, void *stats_out);
void launch_cell_keys(hipStream_t s, int64_t N, const double *x, const double *y, const double *z,
                      const Scalars *sc, bool rotate, uint32_t *keys, uint32_t *vals,
                      const int32_t *bin = nullptr, int bshift = 0, int drop = 0);
void launch_permute(hipStream_t s, int64_t N, const uint32_t *perm, const double *src, double *dst);
void launch_permute_u32(hipStream_t s, int64_t N, const uint32_t *perm, const uint32_t *src, uint32_t *dst);
void launch_generate(hipStream_t s, int64_t N, uint64_t seed, const double sr[3], const double sp[3],
                     double p0, double qeach, double halo_frac, double halo_radius, Bunch &b);
void launch_absmax(hipStream_t s, int64_t N, const double *v, unsigned long long *out_bits);
void launch_iota(hipStream_t s, int64_t N, uint32_t *v);
void launch_fill_i64(hipStream_t s, size_t n, long long *p);
void launch_copy_calibration(hipStream_t s, size_t bytes, const void *src, void *dst);
void launch_cr_test(hipStream_t s, int which, int64_t n, const double *x, double *y);
void launch_igf_term_test(hipStream_t s, int fast, int64_t n, const double *v, double *y);

// Tracker kernels
void launch_track_pre(hipStream_t s, TrackState *ts, Scalars *sc, int64_t N, unsigned long long *dropped);
int alpha_blocks(int64_t N);
void launch_track_alpha(hipStream_t s, bool exact, bool push, Bunch &b, const Scalars *sc,
                        const TrackState *ts, double *partials);
// q_roundtrip: the step scatters the bunch's own charge (plain, not binned): Q*dT, then /dT
void launch_track_mesh(hipStream_t s, const double *partials, int nb, const int n[3], double dh,
                       const TrackState *ts, bool exact, bool q_roundtrip, Scalars *sc);
int gamma_blocks(bool fp64, int64_t N);  // partial-sum blocks launch_track_gamma writes
// the gather pass in blocks that run once (no DC beam or field maps):
// os_blocks(N) rows of kMom doubles in dense; returns the partial blocks for k_track_post.
// Block b takes chunk (b mod 8) * per + b / 8: the blocks of one XCD stream one
// contiguous eighth of the bunch.
constexpr int kXcds = 8;  // XCDs of an MI355X; workgroups are dealt to them round-robin
__host__ __device__ inline int os_blocks(int64_t N) {
    const int64_t nchunks = (N + kOsChunk - 1) / kOsChunk;
    const int64_t per = (nchunks + kXcds - 1) / kXcds;
    return (int)(per * kXcds);
}
// The block's (x, y, z) index with the grid's blocks reordered so that each
// XCD takes one contiguous range of the x-fastest linear order: the workgroup
// of linear index g (in dispatch order) runs on XCD g mod 8, and neighbouring
// indices of the returned order (tiles that share 128-byte lines) run on the
// same XCD, one after the other, so the second fetch of a shared line hits
// that XCD's L2.  A bijection of the grid onto itself for any size.
__device__ inline void xcd_block(int &bx, int &by, int &bz) {
    const int gx = (int)gridDim.x, gxy = gx * (int)gridDim.y, nt = gxy * (int)gridDim.z;
    const int g = (int)blockIdx.x + gx * ((int)blockIdx.y + (int)gridDim.y * (int)blockIdx.z);
    const int q = nt / kXcds, r = nt % kXcds, xcd = g % kXcds, slot = g / kXcds;
    const int l = xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
    bz = l / gxy;
    by = (l - bz * gxy) / gx;
    bx = l - bz * gxy - by * gx;
}
__host__ __device__ inline int os_blocks64(int64_t N) {
    const int64_t nchunks = (N + kOsChunk64 - 1) / kOsChunk64;
    const int64_t per = (nchunks + kXcds - 1) / kXcds;
    return (int)(per * kXcds);
}
// osbox (used with next): per one-shot block the bounds (kOsBoxStats doubles) of the
// positions written, which the next step's bounds come from.  Bg: the energy
// bins' summed magnetic field grid (then B is gathered, not formed from E);
// with it and next, binp (not null) receives the chunk sums of bingamma.hpp
// for the bin ranges *bins.
struct BinRanges;
// perm (may be null): the pass also applies a re-sort (PermSrc); b holds the new arrays.
int launch_track_gamma_os(hipStream_t s, bool fp64, bool next, Bunch &b, const void *Eg, const Scalars *sc,
                          const TrackState *ts, double *dense, double *partials, int ext, double *osbox,
                          const void *Bg = nullptr, const BinRanges *bins = nullptr, double *binp = nullptr,
                          const PermSrc *perm = nullptr);
// blocks of statistics launch_track_gamma_os leaves (one per one-shot block)
int64_t os_box_count(bool fp64, int64_t N);
// The bounds pass from the previous gather pass's statistics: the beam-frame
// chunk statistics into b.chunkbox and per-block bounds on the extremes into
// spare slots of partials' rows, then the exact bounds of the chunks that can
// hold an extreme into partials (bounds_scan_blocks(N) rows, slots 0-5, in
// launch_track_alpha's layout).
int bounds_scan_blocks(int64_t N);
void launch_bounds_from_boxes(hipStream_t s, bool fp64, const Bunch &b, const double *osbox, const Scalars *sc,
                              double *partials);
void launch_bounds_scan(hipStream_t s, bool exact, const Bunch &b, const Scalars *sc, double *partials);
// Bg != nullptr: energy bins, Eg/Bg the summed bin fields (grid E and B both gathered);
// ext: the tracker has a nonzero external field
void launch_track_gamma(hipStream_t s, bool fp64, bool next, Bunch &b, const void *Eg, const Scalars *sc,
                        const TrackState *ts, double *partials, const void *Bg, int ext, bool dc);
// stats_out: the call's statistics array, written at slot ts->step - ts->call_step0 - 1
void launch_track_post(hipStream_t s, const double *partials, int nb, TrackState *ts, Scalars *sc,
                       void *stats_out, const unsigned long long *dropped);
void launch_track_begin_call(hipStream_t s, TrackState *ts);  // call_step0 := step
void launch_track_init_moments(hipStream_t s, const double *partials, int nb, TrackState *ts,
                               Scalars *sc);
void launch_moments_soa(hipStream_t s, const Bunch &b, double *partials);

// Energy bins (bins.hip)
constexpr int kMaxBins = 64;
void launch_bin_keys(hipStream_t s, int64_t N, const int32_t *bin, int nbins, uint32_t *keys, uint32_t *vals);
void launch_bin_count(hipStream_t s, int64_t N, const int32_t *bin, int nbins, unsigned long long *counts);
void launch_bin_order(hipStream_t s, int64_t N, const uint32_t *perm, const double *x, const double *y,
                      const double *z, const double *dt, double qi, double *ox, double *oy, double *oz,
                      double *oqd);
void launch_bin_gamma(hipStream_t s, int64_t count, const uint32_t *perm, const double *P, double *partials,
                      int b, double *gammas);
void launch_bin_scalars(hipStream_t s, Scalars *sc, const double *gammas, int b, int64_t count,
                        const unsigned long long *qdmax_bits, bool exact);
void launch_gather_binned(hipStream_t s, bool fp64, int64_t N, const double *x, const double *y, const double *z,
                          const void *Es, const void *Bs, const Scalars *sc, double *Ef, double *Bf);
// tracker: bin b holds the particles start[b] <= i < start[b+1] of the bin-ordered bunch
struct BinRanges {
    int nbins = 0;
    int64_t start[kMaxBins + 1] = {0};
};
constexpr int kBinGammaBlocks = 4096;  // at most this many partial-sum blocks per bin for calcGammas
// per-bin calcGammas and the per-bin scalars scb[b], and those of the non-empty
// bins in bin order in scbc.  binp: the chunk sums (chunk particles per chunk)
// of bingamma.hpp; null: a pass over the bins' momenta (partials:
// kMaxBins*kBinGammaBlocks*2 doubles)
void launch_track_bin_scalars(hipStream_t s, const Bunch &bunch, const BinRanges &r, double qi, bool exact,
                              const Scalars *sc, double *partials, const double *binp, int64_t chunk,
                              double *gammas, Scalars *scb, Scalars *scbc);
// The chunk sums of bingamma.hpp from the resident momenta (2 doubles per
// one-shot chunk of the precision's gather pass).
void launch_bin_gamma_chunks(hipStream_t s, bool fp64, const Bunch &bunch, const BinRanges &r, double *binp);
void launch_bin_keys_by_id(hipStream_t s, int64_t N, const int32_t *bin, const uint32_t *id, int nbins,
                           uint32_t *keys, uint32_t *vals);

// Sorting (sort.hip)
size_t sort_temp_bytes(int64_t N);
int sort_pairs(hipStream_t s, void *tmp, size_t tmp_bytes, const uint32_t *kin, uint32_t *kout,
               const uint32_t *vin, uint32_t *vout, int64_t N, int end_bit);

}  // namespace opalsc

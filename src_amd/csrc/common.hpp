// Shared definitions for the gfx950 space-charge kernels.
//
// All translation units are compiled with -ffp-contract=off: every multiply
// and add below is rounded separately, in the operation order of the OPAL
// function each helper cites, so that cell indices and CIC weights come out
// bit-identical to the reference's scalar code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace opalsc {

constexpr double kC = 299792458.0;               // src/Classic/Physics/Physics.h:52
constexpr double kPi = 3.14159265358979323846;   // src/Classic/Physics/Physics.h:31
constexpr int kThreads = 256;                    // threads per block for particle kernels
constexpr int kRedBlocksMax = 8192;              // reduction partial slots (blocks of the streaming kernels)
constexpr int kMomentSlots = 32;                 // 27 moments + gamma sum + spare
constexpr int kDepositChunk = 2048;              // particles per deposit / bounds chunk

template <typename T> struct vec4_of;
template <> struct vec4_of<float> { using type = float4; };
template <> struct vec4_of<double> { using type = double4; };
template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };

// Mesh, step scalars and trafos, produced on the device by the scalar kernels
// and read by every grid and particle kernel of a step.
struct Scalars {
    int n[3];
    double hr[3], origin[3];   // mesh of the beam frame (boundp)
    double rmin[3], rmax[3];   // particle bounds
    // deposit fixed point (fixed_point_scale): each deposit is rounded to an
    // integer at the scale e2q.  Mixed path: one word per cell, value = int * q2e.
    // FP64 path: two words per cell, value = hi * q2e + lo * q2e_lo, where a
    // flushed sum v is split into hi = v >> lo_shift and lo = v - (hi << lo_shift).
    double q2e, e2q, q2e_lo;
    int lo_shift;
    double qd_absmax;
    // computeSelfFields scalars (src/Classic/Algorithms/PartBunch.cpp:496-644)
    double dT, gammaz, betaC, hs[3], tmp2, vol, coupling;
    double gscale[3];          // (gz, gz, 1/gz)
    double psum[3];            // sum of P over the bunch
    // coordinate transformations for the step (row-major 3x3)
    double M_rb[9], o_rb[3];   // reference -> beam
    double M_br[9], o_br[3];   // beam -> reference
    double M_u[9], o_u[3];     // reference-frame update (transformBunch)
    double Eext[3], Bext[3];   // external field for the bunch at t + dT/2
    // mixed-precision gather pass: kick and push in the beam frame, then one map
    // to the updated reference frame.  r_b' = M_rb (r - o_rb2) is the beam-frame
    // position relative to the mesh origin, ih = 1/hr, M_c = M_u M_br maps
    // beam-frame vectors to the updated reference frame, r_new = M_c (r_b' - o_c2),
    // Eext_b/Bext_b the external fields in the beam frame.
    double o_rb2[3], ih[3], M_c[9], o_c[3], o_c2[3], Eext_b[3], Bext_b[3];
    double zshift;             // image-charge Green's function shift (FFTPoissonSolver.cpp:289)
    double qi;                 // energy bins: macro charge of the bin's particles (setBinCharge)
    double qu, qd_u;           // tracker, every particle of charge qu: qu and this step's qu*dT
    int mesh_ok;
    int64_t N;
    // DC beam (periodic z): period length, and the z range the particles are
    // wrapped into after boundp, [wrap_lo, wrap_hi) in the beam frame
    int pz;
    double period, wrap_lo, wrap_hi;
    // field maps: the path length the step's lab frame is offset by, and the
    // time of the bunch's field evaluation (t + dT/2)
    double fm_s0, fm_t;
};

// Fixed-point scale of the charge deposits of `count` particles, each
// depositing at most |qdmax| (IntCIC weights are <= 1).
// Mixed path: one deposit below 2^24 (|qd| * 2^e < 2^24), so 64 deposits sum
// in 32 bits and any cell sum stays far below 2^63.
// FP64 path: the high word's unit 2^-e with count * qdmax * 2^e < 2^62 bounds
// every cell sum of the high words; each deposit is rounded at the finer scale
// 2^(e+s) with 2^s <= count / kDepositChunk, which bounds the sum of one
// deposit chunk (kDepositChunk particles) in an int64 LDS cell.  A chunk's
// flushed cell sum v is split exactly into hi = v >> s (added to the high word)
// and lo = v - (hi << s) in [0, 2^s) (added to the low word, which sums at most
// 8 * count such terms, < 2^56).  The rounding error per deposit is then
// 2^-(e+s) instead of 2^-e: at 1e8 particles 15 bits finer than one global
// scale for the whole bunch.
__host__ __device__ inline void fixed_point_scale(Scalars *sc, int64_t count, double qdmax, bool exact) {
    const double B = (exact ? (double)(count > 0 ? count : 1) : 1.0) * qdmax * (1.0 + 0x1p-40);
    int e = 0, s = 0;
    if (B > 0.0) {
        int E;
        (void)frexp(B, &E);
        e = (exact ? 62 : 24) - E;
    }
    if (exact)
        while (((int64_t)kDepositChunk << (s + 1)) <= count) ++s;
    if (e + s > 1000) {  // tiny charges: keep the scales finite
        e = 1000 - s;
    }
    sc->e2q = ldexp(1.0, e + s);
    sc->q2e = ldexp(1.0, -e);
    sc->q2e_lo = ldexp(1.0, -(e + s));
    sc->lo_shift = s;
    sc->qd_absmax = qdmax;
}

// ---- external field maps ------------------------------------------------------
// A Solenoid with a 2D magnetostatic map (FM2DMagnetoStatic) or an RFCavity
// with a 2D dynamic map (FM2DDynamic) on the beamline's z axis.  Grids are
// nz*nr doubles indexed i + j*nz (z index i, r index j), as the reference
// stores them (FM2DMagnetoStatic.cpp:110-130, FM2DDynamic.cpp:118-140):
// magnetostatic a = Bz, b = Br; dynamic a = Ez, b = Er, c = Bt (SI units).
constexpr int kMaxFieldMaps = 8;
struct FieldMapDesc {
    int kind;          // 0: solenoid (2D magnetostatic), 1: RF cavity (2D dynamic)
    int nz, nr;        // grid points
    double hz, hr;     // grid spacings (m)
    double zstart;     // lab z of the map's first z point: element edge + zbegin
    double length;     // zend - zbegin (m)
    double scale;      // KS (solenoid) or VOLT (cavity)
    double omega, phase;
    int64_t a, b, c;   // offsets of the grids in the map data array
};

// Bilinear field of one map at (x, y, z) in its own frame (z from the map's
// first point): FM2DMagnetoStatic::getFieldstrength (FM2DMagnetoStatic.cpp:161-190)
// and FM2DDynamic::getFieldstrength (FM2DDynamic.cpp:169-205), accumulated into
// e, b.  Returns true when the particle is beyond the map's radius.
__host__ __device__ inline bool fieldmap_2d(const FieldMapDesc &m, const double *data, double x, double y, double z,
                                            double e[3], double bf[3]) {
    const double RR = sqrt(x * x + y * y);
    int indexr, indexz;
    if (m.kind == 0) {
        indexr = abs((int)floor(RR / m.hr));
        indexz = abs((int)floor(z / m.hz));
    } else {
        indexr = (int)floor(RR / m.hr);
        indexz = (int)floor(z / m.hz);
    }
    const double leverr = RR / m.hr - indexr;
    const double leverz = z / m.hz - indexz;
    if ((indexz < 0) || (indexz + 2 > m.nz)) return false;
    if (indexr + 2 > m.nr) return true;
    const int index1 = indexz + indexr * m.nz;
    const int index2 = index1 + m.nz;
    auto lerp = [&](const double *f) {
        return (1.0 - leverz) * (1.0 - leverr) * f[index1] + leverz * (1.0 - leverr) * f[index1 + 1] +
               (1.0 - leverz) * leverr * f[index2] + leverz * leverr * f[index2 + 1];
    };
    if (m.kind == 0) {
        const double BfieldR = lerp(data + m.b), BfieldZ = lerp(data + m.a);
        if (RR != 0) {
            bf[0] += BfieldR * x / RR;
            bf[1] += BfieldR * y / RR;
        }
        bf[2] += BfieldZ;
    } else {
        const double EfieldR = lerp(data + m.b), EfieldZ = lerp(data + m.a), BfieldT = lerp(data + m.c);
        if (RR > 1e-10) {
            e[0] += EfieldR * x / RR;
            e[1] += EfieldR * y / RR;
            bf[0] -= BfieldT * y / RR;
            bf[1] += BfieldT * x / RR;
        }
        e[2] += EfieldZ;
    }
    return false;
}

// Solenoid::apply / RFCavity::apply (Solenoid.cpp:167-179, RFCavity.cpp:232-250)
// at lab position (x, y, zlab) and time t: inside [0, length) of the map the
// scaled field (a cavity's E times cos(omega t + phase), its B times
// -sin(omega t + phase)) is added to E, B.  Returns true for a particle beyond
// the map's radius (the reference removes it; it gets no field here).
__host__ __device__ inline bool fieldmap_apply(const FieldMapDesc &m, const double *data, double x, double y,
                                               double zlab, double t, double E[3], double B[3]) {
    const double z = zlab - m.zstart;
    if (!(z >= 0.0 && z < m.length)) return false;
    double te[3] = {0.0, 0.0, 0.0}, tb[3] = {0.0, 0.0, 0.0};
    if (fieldmap_2d(m, data, x, y, z, te, tb)) return true;
    if (m.kind == 0) {
        for (int d = 0; d < 3; ++d) B[d] += m.scale * tb[d];
    } else {
        const double ce = m.scale * cos(m.omega * t + m.phase), sb = m.scale * sin(m.omega * t + m.phase);
        for (int d = 0; d < 3; ++d) E[d] += ce * te[d];
        for (int d = 0; d < 3; ++d) B[d] -= sb * tb[d];
    }
    return false;
}

// Tracker state carried across steps on the device.
struct TrackState {
    double dT, charge, mass_eV, coupling;
    double Bz, Ez0, rf_freq, rf_phase;
    double t, pathLength;
    double RefPartR[3], RefPartP[3];
    int space_charge;
    int step;
    int call_step0;  // step at the start of the current tracking call (statistics slot = step - call_step0 - 1)
    int pz;          // DC beam: periodic z (mesh rule, particle wrap, periodic guards and solve)
    double period;   // its period length (PartBunchBase::setBeamFrequency: c / f)
    // field maps on the beamline (device arrays); the lab frame is the
    // reference frame moved by (0, 0, pathLength)
    int nmaps;
    const FieldMapDesc *maps;
    const double *mapdata;
};

// The particle boundary condition of a DC beam in z, applied by boundp's
// update() (ParticleSpatialLayout.hpp:296-297): ParticlePeriodicBCond
// (ippl/src/Particle/ParticleBConds.h:38-45) over the region of the mesh's
// cells, [origin, n*h + origin) (RegionLayout::convert_index,
// ippl/src/Region/RegionLayout.hpp:526-556; getVertexPosition,
// ippl/src/Meshes/UniformCartesian.hpp:1054-1055).
__host__ __device__ inline double wrap_periodic(double t, double lo, double hi) {
    if (t < lo) return hi - (lo - t);
    if (t >= hi) return lo + (t - hi);
    return t;
}

// Cell index of a stencil corner along z: with a periodic z face the guard
// layers hold the opposite cells (PeriodicFace, filled before the gather by
// Field::fillGuardCells, ippl/src/Field/Field.hpp:366-381), otherwise zero.
// Returns -1 for a corner that reads 0.
__host__ __device__ inline int corner_z(int k, int n2, bool pz) {
    if (k >= 0 && k < n2) return k;
    if (pz) {
        if (k == n2) return 0;
        if (k == -1) return n2 - 1;
    }
    return -1;
}

// ---- cloud-in-cell lookup -------------------------------------------------
// One dimension of IntCICImpl<3>::scatter/gather (ippl/src/Particle/IntCIC.h:826-851):
// getVertexBelow (ippl/src/Meshes/UniformCartesian.hpp:1025-1039), getCellPosition
// (UniformCartesian.hpp:1084-1097), getDeltaCell = h*1 (UniformCartesian.hpp:1151-1158),
// lgp offset -1 when the cell centre lies above the particle.  Returns false
// (particle dropped) when (x-o)/h is non-finite or beyond +-2^30.
__host__ __device__ inline bool cic_dim(double x, double o, double h, int &lgp, double &dpos) {
    double t = (x - o) / h;
    if (!(t > -1073741824.0 && t < 1073741824.0)) return false;
    int i = (int)t;
    int ngp = (x >= o) ? i : i - 1;
    double g = ((double)ngp + 0.5) * h + o;
    double delta = h * 1;
    int off = 0;
    if (g > x) {
        off = -1;
        g = g - delta;
    }
    lgp = ngp + off;
    dpos = (x - g) / delta;
    return true;
}

// Sort order of cells: the 3D Morton (Z-order) index, bits of i, j, k
// interleaved (i lowest).  Any run of consecutive cells in this order spans a
// box of about its own volume in every dimension, so a chunk of particles
// sorted this way has a compact bounding box at every particle density (the
// deposit's LDS box) and a compact cache footprint (the gather).
// Cells up to 1024 per dimension; keys are < cell_key_count(n) = (2^b)^3 with
// 2^b >= max(n).
__host__ __device__ inline uint32_t morton_spread(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint32_t cell_key(int i, int j, int k, int, int) {
    return morton_spread((uint32_t)i) | (morton_spread((uint32_t)j) << 1) | (morton_spread((uint32_t)k) << 2);
}
// Hilbert-curve index of cell (i, j, k) on a 2^b cube (b >= 1), by Skilling's
// transpose method ("Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004): undo the excess rotations and reflections level by level, Gray-code,
// then interleave the bits (i most significant in each triple).  Consecutive
// indices are face neighbours, and each aligned run of 8 is a 2x2x2 block.
__host__ __device__ inline uint32_t cell_key_hilbert(int i, int j, int k, int b) {
    uint32_t X[3] = {(uint32_t)i, (uint32_t)j, (uint32_t)k};
    const uint32_t M = 1u << (b - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int a = 0; a < 3; ++a) {
            if (X[a] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[a]) & P;
                X[0] ^= t;
                X[a] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    uint32_t key = 0;
    for (int q = b - 1; q >= 0; --q)
        key = (key << 3) | (((X[0] >> q) & 1u) << 2) | (((X[1] >> q) & 1u) << 1) | ((X[2] >> q) & 1u);
    return key;
}
__host__ __device__ inline uint32_t cell_key_count(int n0, int n1, int n2) {
    int m = n0 > n1 ? n0 : n1;
    m = m > n2 ? m : n2;
    uint32_t b = 1;
    while ((1u << b) < (uint32_t)m) ++b;
    return 1u << (3 * b);
}

// CIC corner weights in the reference order 000,100,010,110,001,101,011,111,
// each ((a*b)*c) (ippl/src/Particle/IntCIC.h:853-860).
__host__ __device__ inline void cic_weights(const double d[3], double w[8]) {
    const double a0 = 1 - d[0], a1 = 1 - d[1], a2 = 1 - d[2];
    const double b0 = d[0], b1 = d[1], b2 = d[2];
    w[0] = a0 * a1 * a2;
    w[1] = b0 * a1 * a2;
    w[2] = a0 * b1 * a2;
    w[3] = b0 * b1 * a2;
    w[4] = a0 * a1 * b2;
    w[5] = b0 * a1 * b2;
    w[6] = a0 * b1 * b2;
    w[7] = b0 * b1 * b2;
}

// ---- Boris pusher ----------------------------------------------------------
// Half drift in unitless positions: switchToUnitlessPositions, BorisPusher::push,
// switchOffUnitlessPositions (src/Classic/Algorithms/PartBunchBase.hpp:468-513,
// src/Steppers/BorisPusher.h:116-124).
__host__ __device__ inline void push_half(double &x, double &y, double &z, double px, double py,
                                          double pz, double dt) {
    const double s = kC * dt;
    x /= s;
    y /= s;
    z /= s;
    const double g = sqrt(1.0 + (px * px + py * py + pz * pz));
    x += 0.5 * px / g;
    y += 0.5 * py / g;
    z += 0.5 * pz / g;
    x *= s;
    y *= s;
    z *= s;
}

// BorisPusher::kick (src/Steppers/BorisPusher.h:56-113).
__host__ __device__ inline void boris_kick(double P[3], const double E[3], const double B[3],
                                           double dt, double mass, double charge) {
    const double e = 0.5 * dt * charge * kC / mass;
    P[0] += e * E[0];
    P[1] += e * E[1];
    P[2] += e * E[2];
    const double gamma = sqrt(1.0 + (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]));
    const double tf = 0.5 * dt * charge * kC * kC / (gamma * mass);
    const double t0 = tf * B[0], t1 = tf * B[1], t2 = tf * B[2];
    const double w0 = P[0] + (P[1] * t2 - P[2] * t1);
    const double w1 = P[1] + (P[2] * t0 - P[0] * t2);
    const double w2 = P[2] + (P[0] * t1 - P[1] * t0);
    const double sf = 2.0 / (1.0 + (t0 * t0 + t1 * t1 + t2 * t2));
    const double s0 = sf * t0, s1 = sf * t1, s2 = sf * t2;
    const double c0 = w1 * s2 - w2 * s1;
    const double c1 = w2 * s0 - w0 * s2;
    const double c2 = w0 * s1 - w1 * s0;
    P[0] += c0;
    P[1] += c1;
    P[2] += c2;
    const double e2 = 0.5 * dt * charge * kC / mass;
    P[0] += e2 * E[0];
    P[1] += e2 * E[1];
    P[2] += e2 * E[2];
}

// dot(Tenzor<double,3>, Vektor<double,3>) (ippl/src/AppTypes/TSVMetaDot.h:255-263).
__host__ __device__ inline void mat_vec(const double *M, const double v[3], double out[3]) {
    const double a = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    const double b = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    const double c = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
    out[0] = a;
    out[1] = b;
    out[2] = c;
}

// CoordinateSystemTrafo::transformTo (src/Classic/Algorithms/CoordinateSystemTrafo.h:77-80).
__host__ __device__ inline void transform_to(const double *M, const double *o, double r[3]) {
    const double v[3] = {r[0] - o[0], r[1] - o[1], r[2] - o[2]};
    mat_vec(M, v, r);
}

// ---- particle arithmetic of the mixed-precision path ----------------------
// The FP64 path above repeats the reference's operation order (bit-exact cell
// indices, push and kick).  The mixed path keeps particle state in FP64 but
// evaluates the same formulas with FMAs and one reciprocal square root instead
// of three divisions; results differ from the exact path by a few ulps.

// cic_dim with a precomputed 1/h: t = (x-o)/h - 1/2, lgp = floor(t), dpos = t - lgp
// (the same stencil as cic_dim: the cell centre below x is lgp).
__host__ __device__ inline bool cic_dim_fast(double x, double o, double invh, int &lgp, double &dpos) {
    const double t = (x - o) * invh - 0.5;
    if (!(t > -1073741824.0 && t < 1073741824.0)) return false;
    const double f = floor(t);
    lgp = (int)f;
    dpos = t - f;
    return true;
}

__device__ inline void push_half_fast(double &x, double &y, double &z, double px, double py,
                                      double pz, double half_cdt) {
    const double h = half_cdt * rsqrt(fma(pz, pz, fma(py, py, fma(px, px, 1.0))));
    x = fma(px, h, x);
    y = fma(py, h, y);
    z = fma(pz, h, z);
}

__device__ inline void mat_vec_fast(const double *M, const double v[3], double out[3]) {
    const double a = fma(M[0], v[0], fma(M[1], v[1], M[2] * v[2]));
    const double b = fma(M[3], v[0], fma(M[4], v[1], M[5] * v[2]));
    const double c = fma(M[6], v[0], fma(M[7], v[1], M[8] * v[2]));
    out[0] = a;
    out[1] = b;
    out[2] = c;
}

__device__ inline void transform_to_fast(const double *M, const double *o, double r[3]) {
    const double v[3] = {r[0] - o[0], r[1] - o[1], r[2] - o[2]};
    mat_vec_fast(M, v, r);
}

// BorisPusher::kick with e = dt*q*c/(2m), bfac = dt*q*c^2/(2m) precomputed.
__device__ inline void boris_kick_fast(double P[3], const double E[3], const double B[3], double e,
                                       double bfac) {
    P[0] = fma(e, E[0], P[0]);
    P[1] = fma(e, E[1], P[1]);
    P[2] = fma(e, E[2], P[2]);
    const double tf = bfac * rsqrt(fma(P[2], P[2], fma(P[1], P[1], fma(P[0], P[0], 1.0))));
    const double t0 = tf * B[0], t1 = tf * B[1], t2 = tf * B[2];
    const double w0 = P[0] + fma(P[1], t2, -P[2] * t1);
    const double w1 = P[1] + fma(P[2], t0, -P[0] * t2);
    const double w2 = P[2] + fma(P[0], t1, -P[1] * t0);
    const double sf = 2.0 / fma(t2, t2, fma(t1, t1, fma(t0, t0, 1.0)));
    const double s0 = sf * t0, s1 = sf * t1, s2 = sf * t2;
    P[0] += fma(w1, s2, -w2 * s1);
    P[1] += fma(w2, s0, -w0 * s2);
    P[2] += fma(w0, s1, -w1 * s0);
    P[0] = fma(e, E[0], P[0]);
    P[1] = fma(e, E[1], P[1]);
    P[2] = fma(e, E[2], P[2]);
}

// ---- wave/block reductions ------------------------------------------------
__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Wave sums of 16 values per lane by recursive halving: at distance d = 32,
// 16, 8, 4 a lane keeps the half of its remaining values selected by its lane
// bit d, sends the other half to lane ^ d and adds what that lane sends; then
// two butterfly steps (d = 2, 1).  Every partial sum pairs the same lanes in
// the same order as wave_sum (v + partner at d = 32, 16, ..., 1), so the sums
// equal wave_sum's bit for bit, with 17 exchanged values instead of 96.
// Returns the sum of value (lane >> 2) over the wave.
__device__ inline double wave_sum16_split(const double (&v)[16], int lane) {
    double a[8], b[4], c[2];
    {
        const bool up = lane & 32;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double keep = up ? v[8 + i] : v[i], send = up ? v[i] : v[8 + i];
            a[i] = keep + __shfl_xor(send, 32, 64);
        }
    }
    {
        const bool up = lane & 16;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double keep = up ? a[4 + i] : a[i], send = up ? a[i] : a[4 + i];
            b[i] = keep + __shfl_xor(send, 16, 64);
        }
    }
    {
        const bool up = lane & 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double keep = up ? b[2 + i] : b[i], send = up ? b[i] : b[2 + i];
            c[i] = keep + __shfl_xor(send, 8, 64);
        }
    }
    const bool up = lane & 4;
    const double keep = up ? c[1] : c[0], send = up ? c[0] : c[1];
    double t = keep + __shfl_xor(send, 4, 64);
    t += __shfl_xor(t, 2, 64);
    t += __shfl_xor(t, 1, 64);
    return t;
}
__device__ inline double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace opalsc

// Scalar bookkeeping of one ParallelTTracker step, evaluated on the device by
// a single thread so that a step needs no host round trip: boundp's mesh,
// the quaternion/trafo algebra of the beam-frame rotation and of the
// reference-frame update, and calcBeamParameters' statistics.  Each function
// keeps the operation order of the reference function it cites.
#pragma once
#include "common.hpp"

namespace opalsc {

// PartBunchBase::boundp (src/Classic/Algorithms/PartBunchBase.hpp:826-879),
// non-DC beam, not emitting; returns false for "h<0, can not build a mesh".
__host__ __device__ inline bool boundp_mesh(const double rmin_in[3], const double rmax_in[3],
                                            const int n[3], double dh, double hr[3],
                                            double origin[3], bool dc = false, double period = 0.0) {
    double rmin[3] = {rmin_in[0], rmin_in[1], rmin_in[2]};
    double rmax[3] = {rmax_in[0], rmax_in[1], rmax_in[2]};
    // a DC beam (periodic in z) leaves z to the period (PartBunchBase.hpp:814,840-843)
    const int dims = dc ? 2 : 3;
    for (int i = 0; i < dims; ++i) {
        double length = fabs(rmax[i] - rmin[i]);
        if (length < 1e-10) {
            rmax[i] += 1e-10;
            rmin[i] -= 1e-10;
        } else {
            rmax[i] += dh * length;
            rmin[i] -= dh * length;
        }
        hr[i] = (rmax[i] - rmin[i]) / (n[i] - 1);
    }
    if (dc) {
        rmax[2] = rmin[2] + period;
        hr[2] = period / (n[2] - 1);
    }
    if (hr[0] * hr[1] * hr[2] <= 0) return false;
    for (int i = 0; i < 3; ++i) origin[i] = rmin[i] - hr[i] / 2.0;
    return true;
}

// Quaternion::operator*= (src/Classic/Algorithms/Quaternion.cpp:84-93)
__host__ __device__ inline void quat_mul(const double a[4], const double b[4], double out[4]) {
    const double d = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    const double c0 = a[2] * b[3] - a[3] * b[2];
    const double c1 = a[3] * b[1] - a[1] * b[3];
    const double c2 = a[1] * b[2] - a[2] * b[1];
    const double r0 = a[0] * b[0] - d;
    const double r1 = a[0] * b[1] + b[0] * a[1] + c0;
    const double r2 = a[0] * b[2] + b[0] * a[2] + c1;
    const double r3 = a[0] * b[3] + b[0] * a[3] + c2;
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
    out[3] = r3;
}

// Quaternion::rotate (Quaternion.cpp:122-133)
__host__ __device__ inline void quat_rotate(const double q[4], const double v[3], double out[3]) {
    const double qv[4] = {0.0, v[0], v[1], v[2]};
    const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
    double t[4], r[4];
    quat_mul(qv, qc, t);
    quat_mul(q, t, r);
    out[0] = r[1];
    out[1] = r[2];
    out[2] = r[3];
}

// Quaternion::getRotationMatrix (Quaternion.cpp:135-150), normalising first.
__host__ __device__ inline void quat_matrix(const double qin[4], double M[9]) {
    const double len = sqrt(qin[0] * qin[0] + qin[1] * qin[1] + qin[2] * qin[2] + qin[3] * qin[3]);
    const double r[4] = {qin[0] / len, qin[1] / len, qin[2] / len, qin[3] / len};
    M[0] = 1 - 2 * (r[2] * r[2] + r[3] * r[3]);
    M[1] = 2 * (-r[0] * r[3] + r[1] * r[2]);
    M[2] = 2 * (r[0] * r[2] + r[1] * r[3]);
    M[3] = 2 * (r[0] * r[3] + r[1] * r[2]);
    M[4] = 1 - 2 * (r[1] * r[1] + r[3] * r[3]);
    M[5] = 2 * (-r[0] * r[1] + r[2] * r[3]);
    M[6] = 2 * (-r[0] * r[2] + r[1] * r[3]);
    M[7] = 2 * (r[0] * r[1] + r[2] * r[3]);
    M[8] = 1 - 2 * (r[1] * r[1] + r[2] * r[2]);
}

// getQuaternion(u, (0,0,1)) (Quaternion.cpp:34-69).  For an antiparallel u the
// reference draws a random axis; here the axis is (1,0,0).  A zero-length u
// (the reference throws) gives the identity.
__host__ __device__ inline void quat_to_z(const double uin[3], double q[4]) {
    const double len = sqrt(uin[0] * uin[0] + uin[1] * uin[1] + uin[2] * uin[2]);
    if (len < 1e-12) {
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
        return;
    }
    const double u[3] = {uin[0] / len, uin[1] / len, uin[2] / len};
    const double ref[3] = {0.0 / 1.0, 0.0 / 1.0, 1.0 / 1.0};
    double axis[3] = {u[1] * ref[2] - u[2] * ref[1], u[2] * ref[0] - u[0] * ref[2],
                      u[0] * ref[1] - u[1] * ref[0]};
    const double normAxis = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    const double dur = u[0] * ref[0] + u[1] * ref[1] + u[2] * ref[2];
    if (normAxis < 1e-12) {
        if (fabs(dur - 1.0) < 1e-12) {
            q[0] = 1.0;
            q[1] = q[2] = q[3] = 0.0;
        } else {
            q[0] = 0.0;
            q[1] = 1.0;
            q[2] = 0.0;
            q[3] = 0.0;
        }
        return;
    }
    for (int d = 0; d < 3; ++d) axis[d] /= normAxis;
    const double cosAngle = sqrt(0.5 * (1 + dur));
    const double sinAngle = sqrt(1 - cosAngle * cosAngle);
    q[0] = cosAngle;
    q[1] = sinAngle * axis[0];
    q[2] = sinAngle * axis[1];
    q[3] = sinAngle * axis[2];
}

struct Trafo {
    double origin[3], q[4], M[9];
};

// CoordinateSystemTrafo(origin, orientation) (CoordinateSystemTrafo.cpp:19-23)
__host__ __device__ inline Trafo trafo_make(const double o[3], const double q[4]) {
    Trafo t;
    for (int d = 0; d < 3; ++d) t.origin[d] = o[d];
    for (int d = 0; d < 4; ++d) t.q[d] = q[d];
    quat_matrix(q, t.M);
    return t;
}

// CoordinateSystemTrafo::invert (CoordinateSystemTrafo.cpp:25-29)
__host__ __device__ inline Trafo trafo_inverted(const Trafo &a) {
    Trafo t;
    double o[3];
    quat_rotate(a.q, a.origin, o);
    for (int d = 0; d < 3; ++d) t.origin[d] = -o[d];
    t.q[0] = a.q[0];
    for (int d = 1; d < 4; ++d) t.q[d] = -a.q[d];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t.M[3 * r + c] = a.M[3 * c + r];
    return t;
}

// External field stand-in in the reference frame at time t.
__host__ __device__ inline void ext_fields(const TrackState &ts, double t, double E[3], double B[3]) {
    E[0] = 0.0;
    E[1] = 0.0;
    E[2] = ts.Ez0 == 0.0 ? 0.0 : ts.Ez0 * sin(2.0 * kPi * ts.rf_freq * t + ts.rf_phase);
    B[0] = 0.0;
    B[1] = 0.0;
    B[2] = ts.Bz;
}

// ParallelTTracker::updateReferenceParticle + updateRefToLabCSTrafo
// (src/Algorithms/ParallelTTracker.cpp:1123-1213), called after t += dT.
// Returns the trafo that transformBunch applies to the bunch.
__host__ __device__ inline Trafo update_reference(TrackState &ts) {
    const double dt = ts.t < ts.dT ? ts.t : ts.dT;
    const double scale = kC * dt;
    double *R = ts.RefPartR, *P = ts.RefPartP;
    push_half(R[0], R[1], R[2], P[0], P[1], P[2], dt);
    double E[3], B[3];
    ext_fields(ts, ts.t - 0.5 * dt, E, B);
    double Ef[3] = {0.0 + E[0], 0.0 + E[1], 0.0 + E[2]};
    double Bf[3] = {0.0 + B[0], 0.0 + B[1], 0.0 + B[2]};
    // field maps (applyToReferenceParticle) in the lab frame of the step
    for (int k = 0; k < ts.nmaps; ++k)
        if (fieldmap_apply(ts.maps[k], ts.mapdata, R[0], R[1], R[2] + ts.pathLength, ts.t - 0.5 * dt, Ef, Bf)) break;
    boris_kick(P, Ef, Bf, dt, ts.mass_eV, ts.charge);
    push_half(R[0], R[1], R[2], P[0], P[1], P[2], dt);
    (void)scale;
    ts.pathLength += sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    double q[4];
    quat_to_z(P, q);
    Trafo upd = trafo_make(R, q);
    double nr[3] = {R[0], R[1], R[2]};
    transform_to(upd.M, upd.origin, nr);
    double np[3];
    mat_vec(upd.M, P, np);
    for (int d = 0; d < 3; ++d) {
        R[d] = nr[d];
        P[d] = np[d];
    }
    return upd;
}

// Index of moment (i,j) in the 27-entry layout of PartBunchBase::calcMoments
// (PartBunchBase.hpp:2185-2258): 6 first moments, then the lower triangle row by row.
// Moment sums that calcBeamParameters reads (PartBunchBase.hpp:1557-1611):
// calcMoments (PartBunchBase.hpp:2185-2258) accumulates the full lower
// triangle of part_i*part_j, but only the 9 entries below enter the
// statistics, so only they are summed here (same products, same values):
//   [0..5]   sum of x, px, y, py, z, pz
//   [6+3d]   sum r_d^2,  [7+3d] sum p_d^2,  [8+3d] sum r_d*p_d   (d = 0..2)
//   [15]     sum sqrt(1 + p.p)
constexpr int kMom = 16;

// calcBeamParameters from the moment sums (PartBunchBase.hpp:1557-1611) and
// calcEMean (PartBunchBase.hpp:2099-2116).  Output order matches opalsc_stats.
struct StatsOut {
    double rmean[3], pmean[3], rrms[3], prms[3], eps_norm[3], eps[3], rprms[3];
    double gamma_mean, ekin_MeV;
    double t, pathLength;
    int64_t n;
    int64_t dropped;
};

__host__ __device__ inline void beam_parameters(int64_t N, const double m[kMom], double mass_eV,
                                                StatsOut &s) {
    const double gamma_sum = m[15];
    for (int i = 0; i < 3; ++i)
        s.rmean[i] = s.pmean[i] = s.rrms[i] = s.prms[i] = s.eps_norm[i] = s.eps[i] = s.rprms[i] = 0.0;
    s.gamma_mean = 0.0;
    s.ekin_MeV = 0.0;
    s.n = N;
    if (N == 0) return;
    const double Nd = (double)N;
    double rsq[3], psq[3], rp[3], eps2[3];
    for (int i = 0; i < 3; ++i) {
        s.rmean[i] = m[2 * i] / Nd;
        s.pmean[i] = m[2 * i + 1] / Nd;
        rsq[i] = m[6 + 3 * i] - Nd * s.rmean[i] * s.rmean[i];
        psq[i] = m[7 + 3 * i] - Nd * s.pmean[i] * s.pmean[i];
        if (psq[i] < 0) psq[i] = 0;
        rp[i] = m[8 + 3 * i] - Nd * s.rmean[i] * s.pmean[i];
    }
    for (int i = 0; i < 3; ++i) eps2[i] = (rsq[i] * psq[i] - rp[i] * rp[i]) / (Nd * Nd);
    for (int i = 0; i < 3; ++i) rp[i] /= Nd;
    for (int i = 0; i < 3; ++i) {
        s.rrms[i] = sqrt(rsq[i] / Nd);
        s.prms[i] = sqrt(psq[i] / Nd);
        s.eps_norm[i] = sqrt(eps2[i] > 0.0 ? eps2[i] : 0.0);
        const double tmp = s.rrms[i] * s.prms[i];
        const double fac = (tmp == 0) ? 0.0 : 1.0 / tmp;
        s.rprms[i] = rp[i] * fac;
    }
    const double gamma = gamma_sum / Nd;
    s.gamma_mean = gamma;
    for (int i = 0; i < 3; ++i) s.eps[i] = s.eps_norm[i] / (gamma * sqrt(1.0 - 1.0 / (gamma * gamma)));
    double ek = gamma_sum - Nd;
    ek *= mass_eV * 1.0e-6;
    s.ekin_MeV = ek / Nd;
}

}  // namespace opalsc

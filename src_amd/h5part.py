"""H5Part phase-space files of OPAL-T (include/opalsc_h5.h, libopalsc_h5.so).

H5PartWrapperForPT (src/Structure/H5PartWrapperForPT.cpp) over the HDF5 C
library: create() is writeHeader, write_step() writeStep, read_step()
readStep; write_tracker_step() forms the step attributes of writeStepHeader
(:276-377) from a tracking context's resident bunch, and restart() is what
Distribution::doRestartOpalT (src/Distribution/Distribution.cpp:519-552)
restores from the last step: the particles, the time, the path length and
the reference particle.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopalsc_h5.so")

EXPORTS = ["opalsc_h5_last_error", "opalsc_h5_create", "opalsc_h5_write_step", "opalsc_h5_num_steps",
           "opalsc_h5_num_particles", "opalsc_h5_read_header", "opalsc_h5_read_step"]

EPSILON_0 = 8.854187817e-12   # src/Classic/Physics/Physics.h:58
CLIGHT = 299792458.0          # src/Classic/Physics/Physics.h:52


class H5Step(C.Structure):
    """opalsc_h5_step: the step attributes of H5PartWrapperForPT::writeStepHeader."""
    _fields_ = ([("spos", C.c_double), ("time", C.c_double)] +
                [(n, C.c_double * 3) for n in ("ref_r", "ref_p", "centroid", "rms_x", "mean_p", "rms_p",
                                               "tait_bryant", "eps_norm", "eps_geom", "min_x", "max_x",
                                               "min_p", "max_p")] +
                [("step", C.c_int64), ("local_track_step", C.c_int64), ("global_track_step", C.c_int64),
                 ("sigma", C.c_double), ("energy", C.c_double), ("dE", C.c_double), ("mass", C.c_double),
                 ("charge", C.c_double), ("num_bunch", C.c_int64), ("step_to_last_inj", C.c_int64),
                 ("B_ref", C.c_double * 3), ("E_ref", C.c_double * 3)])


class H5PartError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libopalsc_h5.so not found at %s (built by src_amd/build.py where HDF5 is installed)"
                      % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    i64p, dp = C.POINTER(C.c_int64), C.c_void_p
    sig = {
        "opalsc_h5_last_error": ([], C.c_char_p),
        "opalsc_h5_create": ([C.c_char_p, C.c_char_p, C.c_int64, C.c_double], C.c_int),
        "opalsc_h5_write_step": ([C.c_char_p, C.c_int64, dp, dp, dp, dp, dp, C.POINTER(H5Step), i64p], C.c_int),
        "opalsc_h5_num_steps": ([C.c_char_p, i64p], C.c_int),
        "opalsc_h5_num_particles": ([C.c_char_p, C.c_int64, i64p], C.c_int),
        "opalsc_h5_read_header": ([C.c_char_p, i64p, C.POINTER(C.c_double)], C.c_int),
        "opalsc_h5_read_step": ([C.c_char_p, C.c_int64, C.c_int64, C.c_int64, dp, dp, dp, dp,
                                 C.POINTER(H5Step)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise H5PartError("opalsc_h5 error %d: %s" % (rc, lib().opalsc_h5_last_error().decode()))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def create(path, opal_version=None, dump_frequency=1, dphi_global=0.0):
    """H5PartWrapperForPT::writeHeader: a new file with the file attributes."""
    _check(lib().opalsc_h5_create(str(path).encode(), None if opal_version is None else opal_version.encode(),
                                  int(dump_frequency), float(dphi_global)))


def write_step(path, R, P, Q, attrs, ids=None, ptype=None):
    """H5PartWrapperForPT::writeStep: appends a Step#k group; returns k."""
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
    N = len(R)
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(N, 3)
    Q = np.ascontiguousarray(np.broadcast_to(np.asarray(Q, dtype=np.float64), (N,)))
    ids = np.arange(N, dtype=np.int64) if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
    if ids.shape != (N,):
        raise ValueError("ids must hold one entry per particle")
    if ptype is not None:
        ptype = np.ascontiguousarray(ptype, dtype=np.int32)
        if ptype.shape != (N,):
            raise ValueError("ptype must hold one entry per particle")
    k = C.c_int64(-1)
    _check(lib().opalsc_h5_write_step(str(path).encode(), N, _p(R), _p(P), _p(Q), _p(ids), _p(ptype),
                                      C.byref(attrs), C.byref(k)))
    return k.value


def num_steps(path):
    n = C.c_int64()
    _check(lib().opalsc_h5_num_steps(str(path).encode(), C.byref(n)))
    return n.value


def num_particles(path, step=-1):
    n = C.c_int64()
    _check(lib().opalsc_h5_num_particles(str(path).encode(), int(step), C.byref(n)))
    return n.value


def read_header(path):
    f, d = C.c_int64(0), C.c_double(0.0)
    _check(lib().opalsc_h5_read_header(str(path).encode(), C.byref(f), C.byref(d)))
    return {"dump_frequency": f.value, "dPhiGlobal": d.value}


def read_step(path, step=-1, first=0, last=None):
    """readStep: attributes and particles first..last (inclusive) of a step (-1: the last)."""
    if last is None:
        last = num_particles(path, step) - 1
    n = max(int(last) - int(first) + 1, 0)
    R, P, Q, ids = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.int64)
    a = H5Step()
    _check(lib().opalsc_h5_read_step(str(path).encode(), int(step), int(first), int(last), _p(R), _p(P), _p(Q),
                                     _p(ids), C.byref(a)))
    return {"R": R, "P": P, "Q": Q, "id": ids, "attrs": a}


def tracker_step_attrs(ctx, local_track_step=0, global_track_step=0):
    """The writeStepHeader attributes of a tracking context's resident bunch
    (H5PartWrapperForPT.cpp:276-377), from calcBeamParameters and the
    statistics row the library computes, the bounds of R and P, and the
    tracker's reference particle; the rotation to the lab frame is the
    identity on this straight beamline (TaitBryantAngles 0)."""
    R, P, Q = ctx.get_bunch()
    tp = ctx.track_params()
    st = ctx.calcBeamParameters(R, P, mass_eV=tp.mass_eV)
    row = ctx.stat_row()
    a = H5Step()
    a.spos = row.s
    a.time = tp.t
    a.ref_r[:] = list(tp.RefPartR)
    a.ref_p[:] = list(tp.RefPartP)
    a.centroid[:] = list(st["rmean"])
    a.rms_x[:] = list(st["rrms"])
    a.mean_p[:] = list(st["pmean"])
    a.rms_p[:] = list(st["prms"])
    a.eps_norm[:] = list(st["eps_norm"])
    a.eps_geom[:] = list(st["eps"])
    a.min_x[:] = list(R.min(axis=0))
    a.max_x[:] = list(R.max(axis=0))
    a.min_p[:] = list(P.min(axis=0))
    a.max_p[:] = list(P.max(axis=0))
    a.local_track_step = int(local_track_step)
    a.global_track_step = int(global_track_step)
    # writeStepHeader's sigma: I_0 = 4 pi eps0 c M / Q (M in eV, Q in e), gamma = Ekin / M + 1
    eks, m = st["ekin_MeV"], tp.mass_eV
    gamma = eks / (m * 1e-6) + 1.0
    I0 = 4.0 * math.pi * EPSILON_0 * CLIGHT * m / tp.charge
    ex, ey = st["eps"][0], st["eps"][1]
    den = 2.0 * gamma * I0 * (ex * ex + ey * ey)
    a.sigma = (st["rrms"][0] ** 2 + st["rrms"][1] ** 2) / den if den != 0 else 0.0
    a.energy = eks
    a.dE = row.dE_MeV
    a.mass = 1.0e-9 * m
    a.charge = float(np.sum(Q))
    a.num_bunch, a.step_to_last_inj = 1, 0
    a.B_ref[:] = list(row.B_ref)
    a.E_ref[:] = list(row.E_ref)
    return a, R, P, Q


def write_tracker_step(path, ctx, local_track_step=0, global_track_step=0):
    """ParallelTTracker::writePhaseSpace's H5Part dump of the resident bunch."""
    a, R, P, Q = tracker_step_attrs(ctx, local_track_step, global_track_step)
    return write_step(path, R, P, Q, a)


def restart(ctx, path, dT, step=-1, **track_kw):
    """Distribution::doRestartOpalT over readStep: the particles of a step
    (default the last) as the context's bunch, and the tracker started at the
    step's TIME, SPOS and reference particle."""
    d = read_step(path, step)
    a = d["attrs"]
    ctx.set_bunch(d["R"], d["P"], d["Q"])
    ctx.track_init(dT, t=a.time, path_length=a.spos, ref_r=tuple(a.ref_r), ref_p=tuple(a.ref_p), **track_kw)
    return d

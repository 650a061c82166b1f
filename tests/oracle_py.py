"""ctypes bindings to the CPU oracle (oracle/libopal_oracle.so).

Test infrastructure only: the oracle is the FP64 CPU restatement of OPAL's
space-charge step that the HIP library is checked against.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "libopal_oracle.so")

PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
PI32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
PU8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
PC = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

CLIGHT = 299792458.0
EPSILON_0 = 8.854187817e-12
PI = 3.14159265358979323846
COUPLING = 1.0 / (4 * PI * EPSILON_0)
M_P_EV = 0.93827204e+00 * 1e9


class Stats(C.Structure):
    _fields_ = [(name, C.c_double * 3) for name in
                ("rmean", "pmean", "rrms", "prms", "eps_norm", "eps", "rprms")] + [
        ("gamma_mean", C.c_double), ("ekin_MeV", C.c_double),
        ("moments", C.c_double * 27), ("n", C.c_int64)]


class Tracker(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("dh", C.c_double), ("greens", C.c_int),
                ("space_charge", C.c_int), ("dT", C.c_double), ("charge", C.c_double),
                ("mass_eV", C.c_double), ("coupling", C.c_double), ("Bz", C.c_double),
                ("Ez0", C.c_double), ("rf_freq", C.c_double), ("rf_phase", C.c_double),
                ("t", C.c_double), ("pathLength", C.c_double), ("RefPartR", C.c_double * 3),
                ("RefPartP", C.c_double * 3), ("bcz", C.c_int), ("period", C.c_double),
                ("nmaps", C.c_int), ("maps", C.c_void_p)]


class FieldMap(C.Structure):
    _fields_ = [("kind", C.c_int), ("nz", C.c_int), ("nr", C.c_int), ("hz", C.c_double), ("hr", C.c_double),
                ("zstart", C.c_double), ("length", C.c_double), ("scale", C.c_double), ("omega", C.c_double),
                ("phase", C.c_double), ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(_ROOT, "oracle")])
    L = C.CDLL(_LIB_PATH)
    i64, i32p = C.c_int64, C.POINTER(C.c_int)
    d3 = C.c_double * 3
    L.orc_get_bounds.argtypes = [i64, PD, PD, PD]
    L.orc_boundp.argtypes = [i64, PD, PI32, C.c_double, PD, PD]
    L.orc_boundp.restype = C.c_int
    L.orc_cic_index.argtypes = [i64, PD, PD, PD, PI32, PD, PU8]
    L.orc_scatter.argtypes = [i64, PD, PD, PD, PD, PI32, PD]
    L.orc_scatter.restype = i64
    L.orc_gather.argtypes = [i64, PD, PD, PD, PD, PI32, PD]
    L.orc_gather_periodic_z.argtypes = [i64, PD, PD, PD, PD, PI32, PD]
    L.orc_boundp_dc.argtypes = [i64, PD, PI32, C.c_double, C.c_double, PD, PD]
    L.orc_boundp_dc.restype = C.c_int
    L.orc_wrap_periodic_z.argtypes = [i64, PD, PD, PD, PI32]
    L.orc_fft3d.argtypes = [PC, PI32, C.c_int]
    L.orc_fft1d.argtypes = [PC, C.c_int, C.c_int]
    L.orc_igf_octant.argtypes = [PI32, PD, PD, C.c_int]
    L.orc_green_standard_octant.argtypes = [PI32, PD, PD]
    L.orc_mirror_green.argtypes = [PI32, PD, PD]
    L.orc_convolve_doubled.argtypes = [PI32, PD, PD]
    L.orc_compute_potential.argtypes = [PI32, PD, PD, C.c_int]
    L.orc_grad.argtypes = [PI32, PD, PD, PD]
    L.orc_compute_self_fields.argtypes = [i64, PD, PD, PD, PD, C.c_double, PD, PD, PI32,
                                          C.c_double, C.c_int, PD, PD, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
    L.orc_compute_self_fields.restype = i64
    L.orc_compute_self_fields_periodic_z.argtypes = L.orc_compute_self_fields.argtypes
    L.orc_compute_self_fields_periodic_z.restype = i64
    L.orc_push.argtypes = [i64, PD, PD, PD]
    L.orc_kick.argtypes = [i64, PD, PD, PD, PD, C.c_double, C.c_double]
    L.orc_calc_moments.argtypes = [i64, PD, PD, PD]
    L.orc_calc_beam_parameters.argtypes = [i64, PD, PD, C.c_double, C.POINTER(Stats)]
    L.orc_get_quaternion_to_z.argtypes = [PD, PD]
    L.orc_track.argtypes = [C.POINTER(Tracker), i64, PD, PD, PD, PD, C.c_int, C.POINTER(Stats)]
    L.orc_track.restype = C.c_int
    L.orc_space_charge.argtypes = [C.POINTER(Tracker), i64, PD, PD, PD, PD, PD, PD, PD,
                                   C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_space_charge.restype = i64
    L.orc_shifted_green_doubled.argtypes = [PI32, PD, C.c_double, PD]
    L.orc_compute_potential_shifted.argtypes = [PI32, PD, PD, C.c_double]
    L.orc_compute_potential_periodic_z.argtypes = [PI32, PD, PD, C.c_int]
    L.orc_mirror_green_periodic_z.argtypes = [PI32, PD, PD]
    L.orc_calc_gammas.argtypes = [i64, PD, PI32, C.c_int, PD]
    L.orc_compute_self_fields_binned.argtypes = [i64, PD, PD, PI32, C.c_int, C.c_double, PD, C.c_double,
                                                 PD, PD, PI32, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                                 PD, PD]
    L.orc_compute_self_fields_binned.restype = i64
    L.orc_track_binned.argtypes = [C.POINTER(Tracker), i64, PD, PD, PD, PD, PI32, C.c_int, C.c_double,
                                   C.c_int, C.c_int, C.POINTER(Stats)]
    L.orc_track_binned.restype = C.c_int
    L.orc_fieldmap_eval.argtypes = [C.c_int, C.c_void_p, i64, PD, C.c_double, PD, PD]
    L.orc_fieldmap_eval.restype = i64
    assert L.orc_sizeof_fieldmap() == C.sizeof(FieldMap)
    assert L.orc_sizeof_stats() == C.sizeof(Stats)
    assert L.orc_sizeof_tracker() == C.sizeof(Tracker)
    _lib = L
    return L


def _n(n):
    return np.ascontiguousarray(n, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def get_bounds(R):
    R = _d(R)
    lo, hi = np.zeros(3), np.zeros(3)
    lib().orc_get_bounds(len(R), R, lo, hi)
    return lo, hi


def boundp(R, n, dh=0.02, period=None):
    """period: DC beam (periodic z) with this period length."""
    R = _d(R)
    hr, origin = np.zeros(3), np.zeros(3)
    if period is None:
        rc = lib().orc_boundp(len(R), R, _n(n), dh, hr, origin)
    else:
        rc = lib().orc_boundp_dc(len(R), R, _n(n), dh, float(period), hr, origin)
    if rc != 0:
        raise ValueError("h<0, can not build a mesh")
    return hr, origin


def wrap_periodic_z(R, origin, hr, n):
    R = _d(R).copy()
    lib().orc_wrap_periodic_z(len(R), R, _d(origin), _d(hr), _n(n))
    return R


def cic_index(R, origin, hr):
    R = _d(R)
    N = len(R)
    lgp = np.zeros((N, 3), np.int32)
    dpos = np.zeros((N, 3))
    valid = np.zeros(N, np.uint8)
    lib().orc_cic_index(N, R, _d(origin), _d(hr), lgp, dpos, valid)
    return lgp, dpos, valid


def scatter(R, q, origin, hr, n, rho=None):
    R = _d(R)
    n = _n(n)
    if rho is None:
        rho = np.zeros(int(n[0]) * int(n[1]) * int(n[2]))
    dropped = lib().orc_scatter(len(R), R, _d(q), _d(origin), _d(hr), n, rho)
    return rho, dropped


def gather(R, E, origin, hr, n, periodic_z=False):
    R = _d(R)
    Ef = np.zeros((len(R), 3))
    f = lib().orc_gather_periodic_z if periodic_z else lib().orc_gather
    f(len(R), R, _d(E), _d(origin), _d(hr), _n(n), Ef)
    return Ef


def fft3d(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    m = _n(a.shape[::-1])
    lib().orc_fft3d(a, m, sign)
    return a


def igf_octant(n, hr, quirk=True):
    n = _n(n)
    G = np.zeros(int(np.prod(n + 1)))
    lib().orc_igf_octant(n, _d(hr), G, 1 if quirk else 0)
    return G.reshape(int(n[2]) + 1, int(n[1]) + 1, int(n[0]) + 1)


def green_standard_octant(n, hr):
    n = _n(n)
    G = np.zeros(int(np.prod(n + 1)))
    lib().orc_green_standard_octant(n, _d(hr), G)
    return G.reshape(int(n[2]) + 1, int(n[1]) + 1, int(n[0]) + 1)


def mirror_green(n, Goct):
    n = _n(n)
    G2 = np.zeros(int(np.prod(2 * n)))
    lib().orc_mirror_green(n, _d(Goct).ravel(), G2)
    return G2.reshape(2 * int(n[2]), 2 * int(n[1]), 2 * int(n[0]))


def convolve_doubled(n, rho, G2):
    phi = _d(rho).ravel().copy()
    lib().orc_convolve_doubled(_n(n), phi, _d(G2).ravel())
    return phi


def compute_potential(n, rho, hr, greens=0):
    phi = _d(rho).ravel().copy()
    lib().orc_compute_potential(_n(n), phi, _d(hr), greens)
    return phi


def compute_potential_periodic_z(n, rho, hr, greens=0):
    phi = _d(rho).ravel().copy()
    lib().orc_compute_potential_periodic_z(_n(n), phi, _d(hr), greens)
    return phi


def mirror_green_periodic_z(n, Goct):
    n = _n(n)
    G2 = np.zeros(4 * int(n[0]) * int(n[1]) * int(n[2]))
    lib().orc_mirror_green_periodic_z(n, _d(Goct).ravel(), G2)
    return G2.reshape(int(n[2]), 2 * int(n[1]), 2 * int(n[0]))


def shifted_green_doubled(n, hr, zshift):
    n = _n(n)
    G2 = np.zeros(8 * int(np.prod(n)))
    lib().orc_shifted_green_doubled(n, _d(hr), float(zshift), G2)
    return G2.reshape(2 * int(n[2]), 2 * int(n[1]), 2 * int(n[0]))


def compute_potential_shifted(n, rho, hr, zshift):
    phi = _d(rho).ravel().copy()
    lib().orc_compute_potential_shifted(_n(n), phi, _d(hr), float(zshift))
    return phi


def calc_gammas(P, bins, nbins):
    g = np.zeros(nbins)
    lib().orc_calc_gammas(len(P), _d(P), np.ascontiguousarray(bins, dtype=np.int32), nbins, g)
    return g


def compute_self_fields_binned(R, P, bins, nbins, qi, dt, dT, origin, hr, n, coupling=COUPLING,
                               greens=0, image=True, gammas=None, Bf=None):
    R, P = _d(R), _d(P)
    N = len(R)
    Ef = np.zeros((N, 3))
    Bf = np.zeros((N, 3)) if Bf is None else _d(Bf).copy()
    g = None if gammas is None else _d(gammas)
    dropped = lib().orc_compute_self_fields_binned(
        N, R, P, np.ascontiguousarray(bins, dtype=np.int32), nbins, qi, _d(dt), dT, _d(origin), _d(hr),
        _n(n), coupling, greens, int(image), g.ctypes.data if g is not None else None, Ef, Bf)
    return dict(Ef=Ef, Bf=Bf, dropped=dropped)


def grad(n, phi, hr):
    n = _n(n)
    r = np.zeros(int(np.prod(n)) * 3)
    lib().orc_grad(n, _d(phi).ravel(), _d(hr), r)
    return r.reshape(-1, 3)


def compute_self_fields(R, P, Q, dt, dT, origin, hr, n, coupling=COUPLING, greens=0, Bf=None,
                        want_grids=False, periodic_z=False):
    R, P = _d(R), _d(P)
    Q = _d(Q).copy()
    N = len(R)
    n = _n(n)
    ncell = int(np.prod(n))
    Ef = np.zeros((N, 3))
    Bf = np.zeros((N, 3)) if Bf is None else _d(Bf).copy()
    rho = np.zeros(ncell) if want_grids else None
    phi = np.zeros(ncell) if want_grids else None
    E = np.zeros(3 * ncell) if want_grids else None
    ptr = (lambda a: a.ctypes.data if a is not None else None)
    f = lib().orc_compute_self_fields_periodic_z if periodic_z else lib().orc_compute_self_fields
    dropped = f(N, R, P, Q, _d(dt), dT, _d(origin), _d(hr), n, coupling, greens, Ef, Bf, ptr(rho), ptr(phi),
                ptr(E))
    out = dict(Ef=Ef, Bf=Bf, Q=Q, dropped=dropped)
    if want_grids:
        out.update(rho=rho, phi=phi, E=E.reshape(-1, 3))
    return out


def push(R, P, dt):
    R = _d(R).copy()
    lib().orc_push(len(R), R, _d(P), _d(dt))
    return R


def kick(P, Ef, Bf, dt, charge=1.0, mass=M_P_EV):
    P = _d(P).copy()
    lib().orc_kick(len(P), P, _d(Ef), _d(Bf), _d(dt), charge, mass)
    return P


def calc_moments(R, P):
    m = np.zeros(27)
    lib().orc_calc_moments(len(R), _d(R), _d(P), m)
    return m


def beam_parameters(R, P, mass=M_P_EV):
    s = Stats()
    lib().orc_calc_beam_parameters(len(R), _d(R), _d(P), mass, C.byref(s))
    return stats_dict(s)


def stats_dict(s):
    d = {k: np.array(getattr(s, k)[:]) for k in
         ("rmean", "pmean", "rrms", "prms", "eps_norm", "eps", "rprms")}
    d.update(gamma_mean=s.gamma_mean, ekin_MeV=s.ekin_MeV, moments=np.array(s.moments[:]), n=s.n)
    return d


def quaternion_to_z(u):
    q = np.zeros(4)
    lib().orc_get_quaternion_to_z(_d(u), q)
    return q


def make_tracker(n, dT, charge=1.0, mass=M_P_EV, dh=0.02, greens=0, space_charge=True, Bz=0.0,
                 Ez0=0.0, rf_freq=0.0, rf_phase=0.0, ref_p=(0.0, 0.0, 1.0), coupling=COUPLING,
                 periodic_z=False, period=CLIGHT / 1e9, maps=None):
    t = Tracker()
    t.n[:] = list(n)
    t.dh, t.greens, t.space_charge, t.dT = dh, greens, int(space_charge), dT
    t.charge, t.mass_eV, t.coupling = charge, mass, coupling
    t.Bz, t.Ez0, t.rf_freq, t.rf_phase = Bz, Ez0, rf_freq, rf_phase
    t.t, t.pathLength = 0.0, 0.0
    t.RefPartR[:] = [0.0, 0.0, 0.0]
    t.RefPartP[:] = list(ref_p)
    t.bcz, t.period = int(periodic_z), float(period)
    if maps:
        arr = make_fieldmaps(maps)
        t.nmaps, t.maps = len(maps), C.cast(arr, C.c_void_p)
        t._maps = arr
    return t


def track(tracker, R, P, Q, dt, nsteps):
    R, P, Q, dt = (_d(a).copy() for a in (R, P, Q, dt))
    stats = (Stats * max(nsteps, 1))()
    rc = lib().orc_track(C.byref(tracker), len(R), R, P, Q, dt, nsteps, stats)
    if rc != 0:
        raise RuntimeError("orc_track failed: %d" % rc)
    return R, P, Q, dt, [stats_dict(stats[i]) for i in range(nsteps)]


def track_binned(tracker, R, P, Q, dt, bins, nbins, qi, nsteps, image=True):
    R, P, Q, dt = (_d(a).copy() for a in (R, P, Q, dt))
    stats = (Stats * max(nsteps, 1))()
    rc = lib().orc_track_binned(C.byref(tracker), len(R), R, P, Q, dt,
                                np.ascontiguousarray(bins, dtype=np.int32), nbins, qi, int(image),
                                nsteps, stats)
    if rc != 0:
        raise RuntimeError("orc_track_binned failed: %d" % rc)
    return R, P, Q, dt, [stats_dict(stats[i]) for i in range(nsteps)]


# --- 2D field maps -----------------------------------------------------------------
_ALNUM = set("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789.-+\x89")


def _content_lines(path):
    """Fieldmap::getLine (src/Classic/Fields/Fieldmap.cpp:519-540): lines with an
    alphanumeric character once a '#' comment is cut, trimmed to the first and
    last such character."""
    out = []
    for line in open(path, encoding="latin-1").read().split("\n"):
        line = line.split("#", 1)[0]
        idx = [k for k, ch in enumerate(line) if ch in _ALNUM]
        if idx:
            out.append(line[idx[0]:idx[-1] + 1])
    return out


def parse_fieldmap_2d(path):
    """FM2DMagnetoStatic / FM2DDynamic file reading, restated: returns a dict
    with kind (0 magnetostatic, 1 dynamic), nz, nr (points), hz, hr (m), zbegin,
    length (m), frequency (rad/s) and the grids a, b[, c] (index i + j*nz),
    normalised as the reference does.  Raises ValueError where it throws."""
    lines = _content_lines(path)
    head = lines[0].split()
    kind = {"2DMa": 0, "2DDy": 1}.get(lines[0][:4])
    if kind is None or len(head) != 2 or head[1] not in ("XZ", "ZX"):
        raise ValueError("An error occured when reading the fieldmap")
    swap = head[1] == "ZX"
    k = 1
    def ints(line):
        a, b, n = line.split()
        return float(a), float(b), int(n)
    if swap:
        rb, re_, nr = ints(lines[k]); k += 1
        freq = 0.0
        if kind == 1:
            freq = float(lines[k]); k += 1
        zb, ze, nz = ints(lines[k]); k += 1
    else:
        zb, ze, nz = ints(lines[k]); k += 1
        freq = 0.0
        if kind == 1:
            freq = float(lines[k]); k += 1
        rb, re_, nr = ints(lines[k]); k += 1
    gz, gr = nz + 1, nr + 1
    nv = 2 if kind == 0 else 4
    rows = [list(map(float, l.split())) for l in lines[k:]]
    if len(rows) != gz * gr or any(len(r) != nv for r in rows):
        raise ValueError("An error occured when reading the fieldmap")
    a = np.zeros(gz * gr); b = np.zeros(gz * gr); c = np.zeros(gz * gr)
    it = iter(rows)
    outer, inner = (gz, gr) if swap else (gr, gz)
    for o in range(outer):
        for n in range(inner):
            i, j = (o, n) if swap else (n, o)
            v = next(it)
            q = i + j * gz
            if kind == 0:
                a[q], b[q] = (v[1], v[0]) if swap else (v[0], v[1])
            else:
                if swap:
                    b[q], a[q], c[q] = v[0], v[1], v[2]
                else:
                    a[q], b[q], c[q] = v[0], v[1], v[3]
    vmax = np.max(np.abs(a[:gz]))
    if kind == 0:
        a = a / vmax
        b = b / vmax
    else:
        mu_0 = 1.256637061e-06
        a = a * (1.e6 / vmax)
        b = b * (1.e6 / vmax)
        c = c * (mu_0 / vmax)
    rb, re_, zb, ze = rb / 100., re_ / 100., zb / 100., ze / 100.
    return dict(kind=kind, nz=gz, nr=gr, hr=(re_ - rb) / nr, hz=(ze - zb) / nz, zbegin=zb, length=ze - zb,
                frequency=freq * (2.0 * PI) * 1e6, a=a, b=b, c=c)


def make_fieldmaps(specs):
    """specs: list of (parsed map dict, elemedge, scale, freq_hz, phase).  Returns
    the ctypes array (keep it and the grids alive while it is used)."""
    arr = (FieldMap * max(len(specs), 1))()
    keep = []
    for k, (m, edge, scale, freq_hz, phase) in enumerate(specs):
        f = arr[k]
        f.kind, f.nz, f.nr, f.hz, f.hr = m["kind"], m["nz"], m["nr"], m["hz"], m["hr"]
        f.zstart, f.length, f.scale = edge + m["zbegin"], m["length"], scale
        omega = 0.0
        if m["kind"] == 1:
            omega = freq_hz * (2.0 * PI)
            if not omega > 0.0 or abs((omega - m["frequency"]) / omega) > 0.01:
                omega = m["frequency"]
        f.omega, f.phase = omega, (phase if m["kind"] == 1 else 0.0)
        grids = [np.ascontiguousarray(m[g], dtype=np.float64) for g in ("a", "b", "c")]
        keep += grids
        f.a, f.b, f.c = (g.ctypes.data for g in grids)
    arr._keep = keep
    return arr


def fieldmap_eval(maps, n, R, t):
    R = _d(R)
    E = np.zeros((len(R), 3)); B = np.zeros((len(R), 3))
    lost = lib().orc_fieldmap_eval(n, C.cast(maps, C.c_void_p), len(R), R, float(t), E, B)
    return E, B, lost

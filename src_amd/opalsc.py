"""Python host mirror of the space-charge step's C ABI (include/opalsc.h).

Names follow the reference interfaces they stand in for:
  FFTPoissonSolver.computePotential    src/Solvers/FFTPoissonSolver.cpp:337-382
  SpaceChargeContext.boundp            src/Classic/Algorithms/PartBunchBase.hpp:802-883
  SpaceChargeContext.computeSelfFields src/Classic/Algorithms/PartBunch.cpp:482-648
  SpaceChargeContext.pushParticles     src/Algorithms/ParallelTTracker.h:448-455
  SpaceChargeContext.kickParticles     src/Algorithms/ParallelTTracker.h:442-446
  SpaceChargeContext.calcBeamParameters src/Classic/Algorithms/PartBunchBase.hpp:1529-1613
  Tracker.track                         src/Algorithms/ParallelTTracker.cpp:343-388

All work runs in the HIP library libopalsc.so; importing this module raises
if the library is missing, and there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OPALSC_LIB", os.path.join(_HERE, "libopalsc.so"))

FP64, MIXED = 0, 1
GREENS_INTEGRATED, GREENS_STANDARD = 0, 1
GRID_RHO, GRID_PHI, GRID_EFIELD = 0, 1, 2
BC_OPEN, BC_PERIODIC = 0, 1

CLIGHT = 299792458.0                       # src/Classic/Physics/Physics.h:52
EPSILON_0 = 8.854187817e-12                # src/Classic/Physics/Physics.h:58
PI = 3.14159265358979323846
COUPLING = 1.0 / (4 * PI * EPSILON_0)      # src/Track/TrackRun.cpp:542
M_P_EV = 0.93827204e+00 * 1e9              # src/Classic/Physics/Physics.h:88 (GeV -> eV)

EXPORTS = [
    "opalsc_abi_version", "opalsc_create", "opalsc_destroy", "opalsc_last_error",
    "opalsc_synchronize", "opalsc_compute_potential", "opalsc_compute_potential_shifted",
    "opalsc_compute_potential_dev", "opalsc_compute_self_fields_dev", "opalsc_step",
    "opalsc_boundp", "opalsc_set_beam_frequency", "opalsc_wrap_periodic_z", "opalsc_cic_index",
    "opalsc_sort_by_cell", "opalsc_compute_self_fields", "opalsc_compute_self_fields_binned",
    "opalsc_get_grid", "opalsc_push_half",
    "opalsc_kick", "opalsc_beam_parameters", "opalsc_beam_parameters_noalloc", "opalsc_bunch_set",
    "opalsc_bunch_generate",
    "opalsc_bunch_get", "opalsc_bunch_set_bins", "opalsc_bunch_get_bins", "opalsc_track_init",
    "opalsc_batch_create", "opalsc_batch_destroy", "opalsc_batch_last_error", "opalsc_step_batched", "opalsc_batch_set_fused", "opalsc_track", "opalsc_track_get_params",
    "opalsc_get_mesh", "opalsc_host_atan_cr", "opalsc_host_log_cr", "opalsc_device_atan_cr",
    "opalsc_device_log_cr", "opalsc_host_igf_term", "opalsc_device_igf_term",
    "opalsc_host_green_octant", "opalsc_device_green_octant", "opalsc_profile_enable", "opalsc_profile_read",
    "opalsc_copy_bandwidth", "opalsc_fromfile_count", "opalsc_fromfile_read", "opalsc_stat_row_get",
    "opalsc_stat_write", "opalsc_fieldmap_add", "opalsc_fieldmap_clear", "opalsc_fieldmap_eval",
    "opalsc_tune_fft_candidates", "opalsc_tune_fft_set", "opalsc_tune_fft_get", "opalsc_tune_deposit_set",
    "opalsc_tune_deposit_get", "opalsc_tune_sort_set", "opalsc_tune_sort_get",
]


class Config(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("greens", C.c_int), ("precision", C.c_int),
                ("bboxincr", C.c_double), ("max_particles", C.c_int64), ("device", C.c_int),
                ("sort_interval", C.c_int), ("bcz", C.c_int)]


class Stats(C.Structure):
    _fields_ = [(name, C.c_double * 3) for name in
                ("rmean", "pmean", "rrms", "prms", "eps_norm", "eps", "rprms")] + [
        ("gamma_mean", C.c_double), ("ekin_MeV", C.c_double), ("t", C.c_double),
        ("pathLength", C.c_double), ("n", C.c_int64), ("dropped", C.c_int64)]


class StatRow(C.Structure):
    """opalsc_stat_row: one line of the DataSink statistics file."""
    _fields_ = [("t_ns", C.c_double), ("s", C.c_double), ("num_particles", C.c_int64),
                ("charge", C.c_double), ("energy_MeV", C.c_double)] + [
        (name, C.c_double * 3) for name in
        ("rms_r", "rms_p", "emit", "mean_r", "ref_r", "ref_p", "max_r", "corr")] + [
        ("Dx", C.c_double), ("DDx", C.c_double), ("Dy", C.c_double), ("DDy", C.c_double),
        ("B_ref", C.c_double * 3), ("E_ref", C.c_double * 3), ("dE_MeV", C.c_double),
        ("dt_ns", C.c_double), ("parts_outside", C.c_int64), ("has_particle0", C.c_int),
        ("R0", C.c_double * 3), ("P0", C.c_double * 3)]


class TrackParams(C.Structure):
    _fields_ = [("dT", C.c_double), ("charge", C.c_double), ("mass_eV", C.c_double),
                ("coupling", C.c_double), ("Bz", C.c_double), ("Ez0", C.c_double),
                ("rf_freq", C.c_double), ("rf_phase", C.c_double), ("t", C.c_double),
                ("pathLength", C.c_double), ("RefPartR", C.c_double * 3),
                ("RefPartP", C.c_double * 3), ("space_charge", C.c_int)]


class GaussParams(C.Structure):
    _fields_ = [("n", C.c_int64), ("seed", C.c_uint64), ("sigma_r", C.c_double * 3),
                ("sigma_p", C.c_double * 3), ("p0", C.c_double), ("q_total", C.c_double),
                ("halo_frac", C.c_double), ("halo_radius", C.c_double)]


class OpalscError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libopalsc.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libopalsc.so not found at %s: build it with "
                      "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i64, dp = C.c_void_p, C.c_int64, C.c_void_p
    sig = {
        "opalsc_abi_version": ([], C.c_int),
        "opalsc_create": ([C.POINTER(vp), C.POINTER(Config)], C.c_int),
        "opalsc_destroy": ([vp], C.c_int),
        "opalsc_last_error": ([vp], C.c_char_p),
        "opalsc_synchronize": ([vp], C.c_int),
        "opalsc_compute_potential": ([vp, dp, dp], C.c_int),
        "opalsc_compute_potential_shifted": ([vp, dp, dp, C.c_double], C.c_int),
        "opalsc_compute_potential_dev": ([vp, vp, dp], C.c_int),
        "opalsc_compute_self_fields_dev": ([vp, i64, vp, vp, vp, vp, C.c_double, dp, dp, C.c_double,
                                            vp, vp], C.c_int),
        "opalsc_step": ([vp, C.c_int, dp], C.c_int),
        "opalsc_boundp": ([vp, i64, dp, dp, dp], C.c_int),
        "opalsc_set_beam_frequency": ([vp, C.c_double], C.c_int),
        "opalsc_wrap_periodic_z": ([vp, i64, dp, dp, dp], C.c_int),
        "opalsc_cic_index": ([vp, i64, dp, dp, dp, dp, dp], C.c_int),
        "opalsc_sort_by_cell": ([vp, i64, dp, dp, dp, dp], C.c_int),
        "opalsc_compute_self_fields": ([vp, i64, dp, dp, dp, dp, C.c_double, dp, dp, C.c_double,
                                        dp, dp], C.c_int),
        "opalsc_compute_self_fields_binned": ([vp, i64, dp, dp, dp, C.c_int, C.c_double, dp,
                                               C.c_double, dp, dp, C.c_double, dp, C.c_int, dp, dp],
                                              C.c_int),
        "opalsc_get_grid": ([vp, C.c_int, dp], C.c_int),
        "opalsc_push_half": ([vp, i64, dp, dp, dp], C.c_int),
        "opalsc_kick": ([vp, i64, dp, dp, dp, dp, C.c_double, C.c_double], C.c_int),
        "opalsc_beam_parameters": ([vp, i64, dp, dp, C.c_double, C.POINTER(Stats)], C.c_int),
        "opalsc_beam_parameters_noalloc": ([i64, vp, vp, C.c_double, vp, vp], C.c_int),
        "opalsc_bunch_set": ([vp, i64, dp, dp, dp], C.c_int),
        "opalsc_bunch_generate": ([vp, C.POINTER(GaussParams)], C.c_int),
        "opalsc_bunch_get": ([vp, dp, dp, dp], C.c_int),
        "opalsc_bunch_set_bins": ([vp, dp, C.c_int, C.c_double, C.c_int], C.c_int),
        "opalsc_bunch_get_bins": ([vp, C.POINTER(C.c_int), dp, dp], C.c_int),
        "opalsc_track_init": ([vp, C.POINTER(TrackParams)], C.c_int),
        "opalsc_track": ([vp, C.c_int, dp], C.c_int),
        "opalsc_batch_create": ([C.POINTER(vp), C.POINTER(vp), C.c_int], C.c_int),
        "opalsc_batch_destroy": ([vp], C.c_int),
        "opalsc_batch_last_error": ([vp], C.c_char_p),
        "opalsc_step_batched": ([vp, C.c_int, dp, C.c_int], C.c_int),
        "opalsc_batch_set_fused": ([vp, C.c_int], C.c_int),
        "opalsc_track_get_params": ([vp, C.POINTER(TrackParams)], C.c_int),
        "opalsc_get_mesh": ([vp, dp, dp], C.c_int),
        "opalsc_host_atan_cr": ([i64, dp, dp], C.c_int),
        "opalsc_fromfile_count": ([C.c_char_p, C.POINTER(C.c_int64)], C.c_int),
        "opalsc_fromfile_read": ([C.c_char_p, i64, C.c_int, C.c_double, dp, dp], C.c_int),
        "opalsc_stat_row_get": ([vp, C.POINTER(StatRow)], C.c_int),
        "opalsc_fieldmap_add": ([vp, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
        "opalsc_fieldmap_clear": ([vp], C.c_int),
        "opalsc_fieldmap_eval": ([vp, i64, dp, C.c_double, dp, dp, C.POINTER(C.c_int64)], C.c_int),
        "opalsc_stat_write": ([C.c_char_p, C.c_char_p, C.c_int, C.POINTER(StatRow)], C.c_int),
        "opalsc_host_log_cr": ([i64, dp, dp], C.c_int),
        "opalsc_device_atan_cr": ([vp, i64, dp, dp], C.c_int),
        "opalsc_device_log_cr": ([vp, i64, dp, dp], C.c_int),
        "opalsc_host_igf_term": ([i64, dp, C.c_int, dp], C.c_int),
        "opalsc_device_igf_term": ([vp, i64, dp, C.c_int, dp], C.c_int),
        "opalsc_host_green_octant": ([C.POINTER(C.c_int), dp, C.c_double, C.c_int, C.c_int, dp], C.c_int),
        "opalsc_device_green_octant": ([vp, dp, C.c_double, C.c_int, dp], C.c_int),
        "opalsc_profile_enable": ([vp, C.c_int], C.c_int),
        "opalsc_profile_read": ([vp, C.c_int, C.POINTER(C.c_char_p), dp, dp], C.c_int),
        "opalsc_copy_bandwidth": ([vp, C.c_size_t, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "opalsc_tune_fft_candidates": ([C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int], C.c_int),
        "opalsc_tune_fft_set": ([C.c_int, C.c_int, C.c_int], C.c_int),
        "opalsc_tune_fft_get": ([C.c_int, C.c_int], C.c_int),
        "opalsc_tune_deposit_set": ([C.c_int64], C.c_int),
        "opalsc_tune_deposit_get": ([], C.c_int64),
        "opalsc_tune_sort_set": ([C.c_int], C.c_int),
        "opalsc_tune_sort_get": ([], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _per_particle(a, N, name, vec=False):
    """a as a contiguous float64 array of N scalars or N 3-vectors (AoS); a
    ValueError otherwise, before any pointer reaches the C side."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    want = (N, 3) if vec else (N,)
    if vec and a.ndim == 1 and a.size == 3 * N:
        a = a.reshape(N, 3)
    if a.shape != want:
        raise ValueError("%s must have shape %s, got %s" % (name, want, a.shape))
    return a


def tune_fft_candidates(L, precision):
    """Tile parameters E (elements per thread) the FFT kernels take for length L."""
    buf = (C.c_int * 8)()
    k = lib().opalsc_tune_fft_candidates(int(L), int(precision), buf, 8)
    if k < 0:
        raise ValueError("no FFT tiles for length %d" % L)
    return list(buf[:k])


def tune_fft_set(L, precision, E):
    """Override the tile of length-L transforms for grids created afterwards (E = 0: the table's)."""
    _check(None, lib().opalsc_tune_fft_set(int(L), int(precision), int(E)))


def tune_fft_get(L, precision):
    return lib().opalsc_tune_fft_get(int(L), int(precision))


def tune_deposit_set(max_particles):
    """Bunches of at most max_particles deposit through the larger LDS box."""
    _check(None, lib().opalsc_tune_deposit_set(int(max_particles)))


def tune_deposit_get():
    return int(lib().opalsc_tune_deposit_get())


def tune_sort_set(drop_bits):
    """The tracker's re-sort orders by the cell's Hilbert-curve key without its drop_bits lowest bits."""
    _check(None, lib().opalsc_tune_sort_set(int(drop_bits)))


def tune_sort_get():
    return int(lib().opalsc_tune_sort_get())


def stats_to_dict(s):
    d = {k: np.array(getattr(s, k)[:]) for k in
         ("rmean", "pmean", "rrms", "prms", "eps_norm", "eps", "rprms")}
    d.update(gamma_mean=s.gamma_mean, ekin_MeV=s.ekin_MeV, t=s.t, pathLength=s.pathLength,
             n=s.n, dropped=s.dropped)
    return d


def fromfile_read(path, momenta_in_ev=False, mass_eV=M_P_EV):
    """Distribution TYPE=FROMFILE: (R, P) as OPAL reads them (INPUTMOUNITS=EV
    with momenta_in_ev)."""
    n = C.c_int64()
    p = str(path).encode()
    _check(None, lib().opalsc_fromfile_count(p, C.byref(n)))
    R = np.zeros((n.value, 3))
    P = np.zeros((n.value, 3))
    _check(None, lib().opalsc_fromfile_read(p, n.value, int(bool(momenta_in_ev)), float(mass_eV), _ptr(R), _ptr(P)))
    return R, P


def stat_write(path, row, input_fn="", new_file=False):
    """Append one StatRow to a DataSink .stat file (new_file: header first)."""
    _check(None, lib().opalsc_stat_write(str(path).encode(), str(input_fn).encode(), int(bool(new_file)),
                                         C.byref(row)))


def host_atan_cr(x):
    x = _d(x).ravel()
    y = np.empty_like(x)
    _check(None, lib().opalsc_host_atan_cr(len(x), _ptr(x), _ptr(y)))
    return y


def host_log_cr(x):
    x = _d(x).ravel()
    y = np.empty_like(x)
    _check(None, lib().opalsc_host_log_cr(len(x), _ptr(x), _ptr(y)))
    return y


def host_igf_term(v, fast):
    """The IGF corner term at points v (n x 3), cell volume 1, evaluated on the
    host: fast=False the correctly rounded form, fast=True the mixed-precision
    solve's form."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    y = np.empty(len(v))
    _check(None, lib().opalsc_host_igf_term(len(v), _ptr(v), 1 if fast else 0, _ptr(y)))
    return y


def host_green_octant(n, h, zshift=0.0, table=0, fast=False):
    """The real Green's function octant G[k, j, i] of an n-cell grid with
    spacing h on the host (table 0 plain, 1 and 2 the image charge's)."""
    nn = (C.c_int * 3)(*[int(v) for v in n])
    hh = np.ascontiguousarray(h, dtype=np.float64)
    out = np.empty((int(n[2]) + 1, int(n[1]) + 1, int(n[0]) + 1))
    _check(None, lib().opalsc_host_green_octant(nn, _ptr(hh), float(zshift), int(table), 1 if fast else 0, _ptr(out)))
    return out


def _check(ctx, rc):
    if rc != 0:
        msg = lib().opalsc_last_error(ctx)
        raise OpalscError("opalsc error %d: %s" % (rc, msg.decode() if msg else ""))


class SpaceChargeContext:
    """One opalsc context: a FIELDSOLVER of n cells per dimension on one GPU."""

    def __init__(self, n, precision=FP64, greens=GREENS_INTEGRATED, bboxincr=2.0,
                 max_particles=1 << 20, device=0, sort_interval=0, bcz=BC_OPEN):
        cfg = Config()
        cfg.bcz = bcz
        cfg.n[:] = [int(v) for v in n]
        cfg.greens, cfg.precision, cfg.bboxincr = greens, precision, bboxincr
        cfg.max_particles, cfg.device, cfg.sort_interval = int(max_particles), device, sort_interval
        self.n = tuple(int(v) for v in n)
        self.precision = precision
        h = C.c_void_p()
        rc = lib().opalsc_create(C.byref(h), C.byref(cfg))
        if rc != 0:
            _check(None, rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().opalsc_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(self._h, rc)

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]

    # FFTPoissonSolver::computePotential
    def computePotential(self, rho, hr):
        phi = _d(rho).ravel().copy()
        assert phi.size == self.ncell
        self._c(lib().opalsc_compute_potential(self._h, _ptr(phi), _ptr(_d(hr))))
        return phi

    # FFTPoissonSolver::computePotential(rho, hr, zshift): image-charge potential
    def computePotentialShifted(self, rho, hr, zshift):
        phi = _d(rho).ravel().copy()
        assert phi.size == self.ncell
        self._c(lib().opalsc_compute_potential_shifted(self._h, _ptr(phi), _ptr(_d(hr)), float(zshift)))
        return phi

    def boundp(self, R):
        R = _d(R, (-1, 3))
        hr, origin = np.zeros(3), np.zeros(3)
        self._c(lib().opalsc_boundp(self._h, len(R), _ptr(R), _ptr(hr), _ptr(origin)))
        return hr, origin

    # PartBunchBase::setBeamFrequency: the DC beam's period length is c / f
    def setBeamFrequency(self, f):
        self._c(lib().opalsc_set_beam_frequency(self._h, float(f)))

    # boundp's update() for a DC beam: z wrapped into [origin, n*h + origin)
    def wrap_periodic_z(self, R, origin, hr):
        R = _d(R, (-1, 3)).copy()
        self._c(lib().opalsc_wrap_periodic_z(self._h, len(R), _ptr(R), _ptr(_d(origin)), _ptr(_d(hr))))
        return R

    def cic_index(self, R, origin, hr):
        R = _d(R, (-1, 3))
        lgp = np.zeros((len(R), 3), np.int32)
        dpos = np.zeros((len(R), 3))
        self._c(lib().opalsc_cic_index(self._h, len(R), _ptr(R), _ptr(_d(origin)), _ptr(_d(hr)),
                                       _ptr(lgp), _ptr(dpos)))
        return lgp, dpos

    def sort_by_cell(self, R, origin, hr):
        R = _d(R, (-1, 3))
        perm = np.zeros(len(R), np.uint32)
        self._c(lib().opalsc_sort_by_cell(self._h, len(R), _ptr(R), _ptr(_d(origin)),
                                          _ptr(_d(hr)), _ptr(perm)))
        return perm

    # PartBunch::computeSelfFields
    def computeSelfFields(self, R, P, Q, dt, dT, origin, hr, coupling=COUPLING, Bf=None):
        R = _d(R, (-1, 3))
        N = len(R)
        P = _per_particle(P, N, "P", vec=True)
        Q = _per_particle(Q, N, "Q").copy()
        dt = _per_particle(dt, N, "dt")
        Ef = np.zeros((N, 3))
        Bf = np.zeros((N, 3)) if Bf is None else _per_particle(Bf, N, "Bf", vec=True).copy()
        self._c(lib().opalsc_compute_self_fields(self._h, N, _ptr(R), _ptr(P), _ptr(Q), _ptr(dt),
                                                 dT, _ptr(_d(origin)), _ptr(_d(hr)), coupling,
                                                 _ptr(Ef), _ptr(Bf)))
        return Ef, Bf, Q

    # ParallelTTracker::computeSpaceChargeFields with energy bins over
    # PartBunch::computeSelfFields(int bin)
    def computeSelfFieldsBinned(self, R, P, bin, nbins, qi, dt, dT, origin, hr,
                                coupling=COUPLING, gammas=None, image=False, Bf=None):
        R = _d(R, (-1, 3))
        N = len(R)
        P = _per_particle(P, N, "P", vec=True)
        dt = _per_particle(dt, N, "dt")
        bin = np.ascontiguousarray(bin, dtype=np.int32)
        if bin.shape != (N,):
            raise ValueError("bin must hold one entry per particle")
        Ef = np.zeros((N, 3))
        Bf = np.zeros((N, 3)) if Bf is None else _per_particle(Bf, N, "Bf", vec=True).copy()
        g = None if gammas is None else _d(gammas)
        if g is not None and len(g) < nbins:
            raise ValueError("gammas must hold nbins entries")
        self._c(lib().opalsc_compute_self_fields_binned(
            self._h, N, _ptr(R), _ptr(P), _ptr(bin), int(nbins), float(qi), _ptr(dt), dT,
            _ptr(_d(origin)), _ptr(_d(hr)), coupling, None if g is None else _ptr(g),
            1 if image else 0, _ptr(Ef), _ptr(Bf)))
        return Ef, Bf

    def get_grid(self, which):
        out = np.zeros(self.ncell * (3 if which == GRID_EFIELD else 1))
        self._c(lib().opalsc_get_grid(self._h, which, _ptr(out)))
        return out.reshape(-1, 3) if which == GRID_EFIELD else out

    def pushParticles(self, R, P, dt):
        R = _d(R, (-1, 3)).copy()
        N = len(R)
        P = _per_particle(P, N, "P", vec=True)
        dt = _per_particle(dt, N, "dt")
        self._c(lib().opalsc_push_half(self._h, N, _ptr(R), _ptr(P), _ptr(dt)))
        return R

    def kickParticles(self, P, Ef, Bf, dt, charge=1.0, mass_eV=M_P_EV):
        P = _d(P, (-1, 3)).copy()
        N = len(P)
        Ef = _per_particle(Ef, N, "Ef", vec=True)
        Bf = _per_particle(Bf, N, "Bf", vec=True)
        dt = _per_particle(dt, N, "dt")
        self._c(lib().opalsc_kick(self._h, N, _ptr(P), _ptr(Ef), _ptr(Bf), _ptr(dt), charge, mass_eV))
        return P

    def calcBeamParameters(self, R, P, mass_eV=M_P_EV):
        s = Stats()
        R = _d(R, (-1, 3))
        P = _per_particle(P, len(R), "P", vec=True)
        self._c(lib().opalsc_beam_parameters(self._h, len(R), _ptr(R), _ptr(P), mass_eV,
                                             C.byref(s)))
        return stats_to_dict(s)

    def device_atan_cr(self, x):
        x = _d(x).ravel()
        y = np.empty_like(x)
        self._c(lib().opalsc_device_atan_cr(self._h, len(x), _ptr(x), _ptr(y)))
        return y

    def device_log_cr(self, x):
        x = _d(x).ravel()
        y = np.empty_like(x)
        self._c(lib().opalsc_device_log_cr(self._h, len(x), _ptr(x), _ptr(y)))
        return y

    def device_green_octant(self, h, zshift=0.0, table=0):
        hh = np.ascontiguousarray(h, dtype=np.float64)
        out = np.empty((self.n[2] + 1, self.n[1] + 1, self.n[0] + 1))
        self._c(lib().opalsc_device_green_octant(self._h, _ptr(hh), float(zshift), int(table), _ptr(out)))
        return out

    def device_igf_term(self, v, fast):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        y = np.empty(len(v))
        self._c(lib().opalsc_device_igf_term(self._h, len(v), _ptr(v), 1 if fast else 0, _ptr(y)))
        return y

    # ---- device-resident tracker (ParallelTTracker::execute) ----
    def set_bunch(self, R, P, Q):
        R = _d(R, (-1, 3))
        P = _per_particle(P, len(R), "P", vec=True)
        Q = _per_particle(Q, len(R), "Q")
        self._c(lib().opalsc_bunch_set(self._h, len(R), _ptr(R), _ptr(P), _ptr(Q)))
        self.N = len(R)

    def generate_bunch(self, n, seed, sigma_r, sigma_p, p0, q_total, halo_frac=0.0,
                       halo_radius=5.0):
        g = GaussParams()
        g.n, g.seed = int(n), int(seed)
        g.sigma_r[:] = list(sigma_r)
        g.sigma_p[:] = list(sigma_p)
        g.p0, g.q_total, g.halo_frac, g.halo_radius = p0, q_total, halo_frac, halo_radius
        self._c(lib().opalsc_bunch_generate(self._h, C.byref(g)))
        self.N = int(n)

    def get_bunch(self):
        R = np.zeros((self.N, 3))
        P = np.zeros((self.N, 3))
        Q = np.zeros(self.N)
        self._c(lib().opalsc_bunch_get(self._h, _ptr(R), _ptr(P), _ptr(Q)))
        return R, P, Q

    def set_bins(self, bin, nbins, qi, image=True):
        """Energy bins of the resident bunch (bin[i] for the i-th particle given
        to set_bunch); the tracker then solves each bin in its own frame and,
        with image, its image charge.  nbins=0 removes the bins."""
        bin = np.ascontiguousarray(bin if bin is not None else np.zeros(0), dtype=np.int32)
        if nbins and bin.shape != (self.N,):
            raise ValueError("bin must hold one entry per particle")
        self._c(lib().opalsc_bunch_set_bins(self._h, _ptr(bin) if nbins else None, int(nbins), float(qi),
                                            1 if image else 0))

    def get_bins(self):
        """(counts, gammas) of the energy bins; gammas are those of the most recent step."""
        nb = C.c_int(0)
        counts = np.zeros(64, np.int64)
        gammas = np.zeros(64)
        self._c(lib().opalsc_bunch_get_bins(self._h, C.byref(nb), _ptr(counts), _ptr(gammas)))
        return counts[:nb.value], gammas[:nb.value]

    def track_init(self, dT, charge=1.0, mass_eV=M_P_EV, coupling=COUPLING, Bz=0.0, Ez0=0.0,
                   rf_freq=0.0, rf_phase=0.0, ref_p=(0.0, 0.0, 1.0), space_charge=True, t=0.0,
                   path_length=0.0, ref_r=(0.0, 0.0, 0.0)):
        p = TrackParams()
        p.dT, p.charge, p.mass_eV, p.coupling = dT, charge, mass_eV, coupling
        p.Bz, p.Ez0, p.rf_freq, p.rf_phase = Bz, Ez0, rf_freq, rf_phase
        p.t, p.pathLength = float(t), float(path_length)
        p.RefPartR[:] = list(ref_r)
        p.RefPartP[:] = list(ref_p)
        p.space_charge = int(space_charge)
        self._c(lib().opalsc_track_init(self._h, C.byref(p)))

    def track(self, nsteps, want_stats=True):
        stats = (Stats * max(nsteps, 1))() if want_stats else None
        self._c(lib().opalsc_track(self._h, int(nsteps),
                                   C.cast(stats, C.c_void_p) if want_stats else None))
        return [stats_to_dict(stats[i]) for i in range(nsteps)] if want_stats else None

    # a Solenoid (2D magnetostatic map, scale = KS) or RFCavity (2D dynamic map,
    # scale = VOLT, freq_hz = FREQ, phase = LAG) whose map starts at elemedge
    def add_fieldmap(self, path, elemedge=0.0, scale=1.0, freq_hz=0.0, phase=0.0):
        rc = lib().opalsc_fieldmap_add(self._h, str(path).encode(), float(elemedge), float(scale),
                                       float(freq_hz), float(phase))
        if rc < 0:
            self._c(rc)
        return rc

    def clear_fieldmaps(self):
        self._c(lib().opalsc_fieldmap_clear(self._h))

    def fieldmap_eval(self, R, t):
        R = _d(R, (-1, 3))
        E, B = np.zeros_like(R), np.zeros_like(R)
        lost = C.c_int64()
        self._c(lib().opalsc_fieldmap_eval(self._h, len(R), _ptr(R), float(t), _ptr(E), _ptr(B), C.byref(lost)))
        return E, B, lost.value

    def stat_row(self):
        """The DataSink statistics line of the resident bunch (after track)."""
        row = StatRow()
        self._c(lib().opalsc_stat_row_get(self._h, C.byref(row)))
        return row

    def track_params(self):
        p = TrackParams()
        self._c(lib().opalsc_track_get_params(self._h, C.byref(p)))
        return p

    def mesh(self):
        hr, origin = np.zeros(3), np.zeros(3)
        self._c(lib().opalsc_get_mesh(self._h, _ptr(hr), _ptr(origin)))
        return hr, origin

    def profile_enable(self, on=True):
        self._c(lib().opalsc_profile_enable(self._h, int(on)))

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        g = C.c_double()
        self._c(lib().opalsc_copy_bandwidth(self._h, int(nbytes), int(reps), C.byref(g)))
        return g.value

    def profile_read(self):
        names = (C.c_char_p * 32)()
        ms = np.zeros(32)
        launches = np.zeros(32, np.int64)
        n = lib().opalsc_profile_read(self._h, 32, names, _ptr(ms), _ptr(launches))
        return {names[i].decode(): (float(ms[i]), int(launches[i])) for i in range(n)}


class Batch:
    """Several independent bunches stepped together (opalsc_step_batched).
    The contexts must be set up for tracking and outlive the batch."""

    def __init__(self, contexts, fused=True):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c._h for c in self.contexts])
        h = C.c_void_p()
        rc = lib().opalsc_batch_create(C.byref(h), arr, len(self.contexts))
        if rc != 0:
            raise OpalscError("opalsc_batch_create failed: %d" % rc)
        self._h = h
        lib().opalsc_batch_set_fused(self._h, 1 if fused else 0)

    def step(self, nsteps, graph=True, want_stats=True):
        n = len(self.contexts)
        stats = (Stats * max(nsteps * n, 1))() if want_stats else None
        rc = lib().opalsc_step_batched(self._h, int(nsteps), C.cast(stats, C.c_void_p) if want_stats else None,
                                       1 if graph else 0)
        if rc != 0:
            msg = lib().opalsc_batch_last_error(self._h)
            raise OpalscError("opalsc error %d: %s" % (rc, msg.decode() if msg else ""))
        if not want_stats:
            return None
        return [[stats_to_dict(stats[s * n + i]) for i in range(n)] for s in range(nsteps)]

    def close(self):
        if getattr(self, "_h", None):
            lib().opalsc_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class FFTPoissonSolver:
    """PoissonSolver interface (src/Solvers/PoissonSolver.h:16-70) over the HIP solver."""

    def __init__(self, n, greens="INTEGRATED", precision=FP64, device=0, bcz="OPEN"):
        g = GREENS_INTEGRATED if greens.upper() == "INTEGRATED" else GREENS_STANDARD
        b = BC_PERIODIC if str(bcz).upper() == "PERIODIC" else BC_OPEN
        self._ctx = SpaceChargeContext(n, precision=precision, greens=g, max_particles=0,
                                       device=device, bcz=b)

    def computePotential(self, rho, hr, zshift=None):
        """In place in the reference (rho in, phi out); here rho is returned as phi.
        With zshift: the image-charge overload (FFTPoissonSolver.cpp:289-330)."""
        if zshift is None:
            return self._ctx.computePotential(rho, hr)
        return self._ctx.computePotentialShifted(rho, hr, zshift)

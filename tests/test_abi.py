"""The C-ABI library without a GPU: it loads, exports every entry point that
include/opalsc.h declares, validates arguments, and its correctly rounded
atan/log (the same double-double code the IGF kernel runs) agree bit for bit
with mpmath and with the oracle's binary128 evaluation."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_py as O

sc = pytest.importorskip("src_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "opalsc.h")).read()
    return sorted(set(re.findall(r"\b(opalsc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = sc.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    assert sorted(sc.EXPORTS) == names
    assert L.opalsc_abi_version() == 2


def test_create_validates_config_before_touching_a_device():
    with pytest.raises(sc.OpalscError, match="precision"):
        sc.SpaceChargeContext((32, 32, 32), precision=7)
    with pytest.raises(sc.OpalscError, match="greens"):
        sc.SpaceChargeContext((32, 32, 32), greens=5)
    with pytest.raises(sc.OpalscError, match="max_particles"):
        sc.SpaceChargeContext((32, 32, 32), max_particles=-1)
    with pytest.raises(sc.OpalscError, match="bcz"):
        sc.SpaceChargeContext((32, 32, 32), bcz=3)


def test_config_struct_matches_header():
    """The ctypes mirror of opalsc_config has the header's fields in order."""
    text = open(os.path.join(ROOT, "include", "opalsc.h")).read()
    body = re.search(r"typedef struct \{(.*?)\} opalsc_config;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b([a-z_]+)(?:\[3\])?;", body)
    assert fields == [f[0] for f in sc.opalsc.Config._fields_]


def test_host_cr_math_equals_quad_oracle():
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.uniform(-5, 5, 100000), 10 ** rng.uniform(-30, 30, 100000),
                        rng.uniform(0.999, 1.001, 20000)])
    ya, yl = sc.host_atan_cr(x), sc.host_log_cr(np.abs(x))
    L = O.lib()
    L.orc_atan_cr.restype = L.orc_log_cr.restype = C.c_double
    L.orc_atan_cr.argtypes = L.orc_log_cr.argtypes = [C.c_double]
    qa = np.array([L.orc_atan_cr(float(v)) for v in x])
    ql = np.array([L.orc_log_cr(float(abs(v))) for v in x])
    assert np.array_equal(ya, qa)
    assert np.array_equal(yl, ql)


def test_host_cr_math_equals_mpmath_on_igf_arguments():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 160
    rng = np.random.default_rng(22)
    n, hr = 128, np.array([1.3e-4, 1.1e-4, 2.9e-4])
    idx = rng.integers(0, n + 2, size=(3000, 3))
    v = idx * hr - hr / 2
    r = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    a_args = np.concatenate([v[:, 0] * v[:, 1] / (v[:, 2] * r), v[:, 1] * v[:, 2] / (v[:, 0] * r)])
    l_args = np.concatenate([v[:, 0] + r, v[:, 2] + r])
    ya, yl = sc.host_atan_cr(a_args), sc.host_log_cr(l_args)
    assert all(y == float(mpmath.atan(mpmath.mpf(float(a)))) for a, y in zip(a_args, ya))
    assert all(y == float(mpmath.log(mpmath.mpf(float(a)))) for a, y in zip(l_args, yl))


def test_special_values():
    x = np.array([0.0, -0.0, np.inf, -np.inf, 1e-310, 1.0])
    ya = sc.host_atan_cr(x)
    assert ya[0] == 0.0 and ya[4] == 1e-310
    assert ya[2] == np.arctan(np.inf) and ya[3] == -np.arctan(np.inf)
    yl = sc.host_log_cr(np.array([1.0, 2.0, 1e-310, np.inf]))
    assert yl[0] == 0.0 and yl[1] == np.log(2.0) and yl[3] == np.inf
    assert yl[2] == float(np.log(np.float64(1e-310)))


def test_mirror_rejects_short_per_particle_arrays_before_the_c_call():
    """Every per-particle argument of the Python mirror must hold N scalars or
    N 3-vectors; a scalar Q or a short dt raises ValueError before a pointer
    reaches the library (the method never touches the context's handle)."""
    c = object.__new__(sc.SpaceChargeContext)  # no device context behind it
    R = np.zeros((4, 3))
    good = dict(P=np.zeros((4, 3)), Q=np.zeros(4), dt=np.zeros(4))
    bad = [("Q", 1e-12), ("dt", np.zeros(3)), ("P", np.zeros((3, 3)))]
    for name, val in bad:
        args = dict(good, **{name: val})
        with pytest.raises(ValueError, match=name):
            c.computeSelfFields(R, args["P"], args["Q"], args["dt"], 1e-11, np.zeros(3), np.ones(3))
    with pytest.raises(ValueError, match="Bf"):
        c.computeSelfFields(R, good["P"], good["Q"], good["dt"], 1e-11, np.zeros(3), np.ones(3),
                            Bf=np.zeros((5, 3)))
    with pytest.raises(ValueError, match="dt"):
        c.pushParticles(R, good["P"], np.zeros(2))
    with pytest.raises(ValueError, match="Ef"):
        c.kickParticles(good["P"], np.zeros((2, 3)), np.zeros((4, 3)), good["dt"])
    with pytest.raises(ValueError, match="Q"):
        c.set_bunch(R, good["P"], np.zeros(5))
    with pytest.raises(ValueError, match="dt"):
        c.computeSelfFieldsBinned(R, good["P"], np.zeros(4, np.int32), 1, 1e-12, np.zeros(1), 1e-11,
                                  np.zeros(3), np.ones(3))


def test_fft_tile_candidates_and_overrides():
    """The autotuner's knobs: the tiles the kernels take per transform length
    (E elements per thread, W = 512 E / L lines a power of two, radix 3 only
    for L = 3 * 2^a), overrides that apply and reset, invalid tiles refused."""
    assert sc.tune_fft_candidates(32, sc.MIXED) == [4, 8, 16, 32]
    assert sc.tune_fft_candidates(96, sc.MIXED) == [6, 12, 24]
    assert sc.tune_fft_candidates(96, sc.FP64) == [6, 12]
    assert sc.tune_fft_candidates(512, sc.FP64) == [4, 8, 16]
    for L in (32, 64, 96, 128, 256, 512):
        for p in (sc.MIXED, sc.FP64):
            cs = 16 if p == sc.FP64 else 8
            for E in sc.tune_fft_candidates(L, p):
                W = E * 512 // L
                assert E * 512 % L == 0 and W & (W - 1) == 0
                assert L * (W + 1) * cs + (0 if p == sc.FP64 else L * cs) <= 160 * 1024
    assert sc.tune_fft_candidates(80, sc.MIXED) == [10, 20]
    assert sc.tune_fft_candidates(80, sc.FP64) == [10]
    for L in (56, 200, 36, 6):  # 7 * 2^3, 25 * 2^3, 9 * 2^2, 3 * 2^1: no tiles
        with pytest.raises(ValueError):
            sc.tune_fft_candidates(L, sc.MIXED)
    base = sc.tune_fft_get(128, sc.MIXED)
    assert base in sc.tune_fft_candidates(128, sc.MIXED)
    try:
        sc.tune_fft_set(128, sc.MIXED, 32)
        assert sc.tune_fft_get(128, sc.MIXED) == 32
        assert sc.tune_fft_get(128, sc.FP64) in sc.tune_fft_candidates(128, sc.FP64)
        with pytest.raises(sc.OpalscError):
            sc.tune_fft_set(128, sc.MIXED, 6)  # not a tile of a power-of-two length
        assert sc.tune_fft_get(128, sc.MIXED) == 32  # a refused tile leaves the override
    finally:
        sc.tune_fft_set(128, sc.MIXED, 0)
    assert sc.tune_fft_get(128, sc.MIXED) == base


def test_deposit_tuning_threshold_round_trip():
    old = sc.tune_deposit_get()
    try:
        sc.tune_deposit_set(12345)
        assert sc.tune_deposit_get() == 12345
        with pytest.raises(sc.OpalscError):
            sc.tune_deposit_set(-1)
    finally:
        sc.tune_deposit_set(old)
    assert sc.tune_deposit_get() == old


def test_sort_granularity_round_trip():
    old = sc.tune_sort_get()
    try:
        for d in (0, 3, 6):
            sc.tune_sort_set(d)
            assert sc.tune_sort_get() == d
        with pytest.raises(sc.OpalscError):
            sc.tune_sort_set(-3)
    finally:
        sc.tune_sort_set(old)

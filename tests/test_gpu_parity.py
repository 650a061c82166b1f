"""Parity of the HIP library (through its C ABI) with the CPU oracle.

Bars (from the FP64 drop-in requirement): cell indices, CIC offsets, the
cell sort, push and kick bit-exact; rho/phi/E/Ef/Bf within 1e-12 relative L2;
statistics within 1e-10 relative where the statistic is well conditioned.
"""
import os

import numpy as np
import pytest

import oracle_py as O
from conftest import gaussian_bunch, rel_l2

pytestmark = pytest.mark.gpu

sc = pytest.importorskip("src_amd")


def ctx(n, precision=sc.FP64, greens=sc.GREENS_INTEGRATED, N=1 << 17, sort_interval=0):
    return sc.SpaceChargeContext(n, precision=precision, greens=greens, max_particles=N,
                                 sort_interval=sort_interval)


def test_cr_math_device_equals_host_and_quad():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-4, 4, 50000), 10 ** rng.uniform(-20, 20, 50000)])
    with ctx((8, 8, 8), N=200000) as c:
        da, dl = c.device_atan_cr(x), c.device_log_cr(np.abs(x) + 1e-300)
    ha, hl = sc.host_atan_cr(x), sc.host_log_cr(np.abs(x) + 1e-300)
    assert np.array_equal(da, ha)
    assert np.array_equal(dl, hl)
    L = O.lib()
    L.orc_atan_cr.restype = L.orc_log_cr.restype = O.C.c_double
    L.orc_atan_cr.argtypes = L.orc_log_cr.argtypes = [O.C.c_double]
    idx = rng.integers(0, len(x), 3000)
    assert all(L.orc_atan_cr(float(x[i])) == ha[i] for i in idx)
    assert all(L.orc_log_cr(float(abs(x[i]) + 1e-300)) == hl[i] for i in idx)


@pytest.mark.parametrize("n", [(32, 32, 32), (16, 24, 48)])
def test_boundp_and_cic_index_bit_exact(n):
    R, P, Q = gaussian_bunch(100000, seed=1)
    R[:, 2] *= 3.0
    hr_o, org_o = O.boundp(R, n)
    lgp_o, dpos_o, valid = O.cic_index(R, org_o, hr_o)
    with ctx(n) as c:
        hr, org = c.boundp(R)
        lgp, dpos = c.cic_index(R, org, hr)
    assert np.array_equal(hr, hr_o) and np.array_equal(org, org_o)
    assert valid.all()
    assert np.array_equal(lgp, lgp_o)
    assert np.array_equal(dpos, dpos_o)


def test_sort_by_cell_is_stable_and_bit_exact():
    n = (32, 32, 32)
    R, P, Q = gaussian_bunch(100000, seed=2)
    hr, org = O.boundp(R, n)
    lgp, _, _ = O.cic_index(R, org, hr)
    key = np.zeros(len(lgp), np.int64)
    for b in range(10):  # 3D Morton order, i in the lowest bit of each triple
        for d in range(3):
            key |= ((lgp[:, d].astype(np.int64) >> b) & 1) << (3 * b + d)
    with ctx(n) as c:
        perm = c.sort_by_cell(R, org, hr)
    assert np.array_equal(perm.astype(np.int64), np.argsort(key, kind="stable"))


def _morton_keys(lgp, n):
    m = 1
    while (1 << m) < max(n):
        m += 1
    key = np.zeros(len(lgp), np.int64)
    for b in range(m):
        for d in range(3):
            key |= ((lgp[:, d].astype(np.int64) >> b) & 1) << (3 * b + d)
    inside = np.all((lgp >= 0) & (lgp < np.asarray(n)), axis=1)
    key[~inside] = 1 << (3 * m)  # outside the grid: after every cell
    return key


@pytest.mark.parametrize("n,N", [((256, 256, 256), 1_000_003), ((48, 24, 96), 300_001), ((8, 8, 8), 5000)])
def test_sort_by_cell_large_keys_ragged_and_ties(n, N):
    """The radix sort over the key widths the grids need (25 bits for 256^3:
    three digit passes; 22 bits for 96 cells: three; 10 bits for 8^3: two
    passes), a particle count that is no multiple of the 2048-key blocks,
    a dense clump (thousands of equal keys) and particles outside the mesh:
    the permutation is numpy's stable argsort of the Morton keys."""
    rng = np.random.default_rng(N)
    R = rng.normal(size=(N, 3)) * 1e-3
    R[: N // 4] *= 0.02  # a clump: long runs of equal keys
    hr, org = O.boundp(R[: N // 2], n)  # the mesh of half the bunch: the rest may fall outside
    lgp, _, _ = O.cic_index(R, org, hr)
    key = _morton_keys(lgp, n)
    with ctx(n, N=N) as c:
        perm = c.sort_by_cell(R, org, hr)
    assert np.array_equal(perm.astype(np.int64), np.argsort(key, kind="stable"))


@pytest.mark.parametrize("greens", [0, 1])
@pytest.mark.parametrize("n", [(16, 16, 16), (32, 16, 24), (12, 12, 12), (8, 48, 12), (64, 32, 96), (20, 40, 10),
                               (80, 24, 40), (16, 6, 10), (12, 10, 6)])
def test_compute_potential_fp64(n, greens):
    """Against the oracle's solve; (16, 6, 10) and (12, 10, 6) have 60 x lines,
    fewer than one block's 2W, so the x passes' blocks are partly empty
    (clamped loads, no stores)."""
    rng = np.random.default_rng(4)
    rho = rng.normal(size=n[0] * n[1] * n[2])
    hr = np.array([1.1e-4, 0.9e-4, 2.3e-4])
    ref = O.compute_potential(n, rho, hr, greens)
    with ctx(n, greens=greens) as c:
        phi = c.computePotential(rho, hr)
    assert rel_l2(phi, ref) < 1e-12


@pytest.mark.parametrize("precision,tol", [(sc.FP64, 1e-12), (sc.MIXED, 1e-5)])
@pytest.mark.parametrize("n", [(32, 48, 20), (64, 24, 40)])
def test_every_fft_tile_the_autotuner_can_pick(n, precision, tol):
    """Every tile (elements per thread E) the kernels offer for each transform
    length, set as the autotuner sets it: the solve stays within the bar of
    the oracle (radix sequences 8/4/2 with 3 or 5 last, tiles of 4 to 64 lines)."""
    rng = np.random.default_rng(21)
    rho = rng.normal(size=n[0] * n[1] * n[2])
    hr = np.array([1.1e-4, 0.9e-4, 2.3e-4])
    ref = O.compute_potential(n, rho, hr, 0)
    lengths = sorted({2 * m for m in n})
    cands = {L: sc.tune_fft_candidates(L, precision) for L in lengths}
    try:
        for k in range(max(len(v) for v in cands.values())):
            for L in lengths:
                sc.tune_fft_set(L, precision, cands[L][min(k, len(cands[L]) - 1)])
            with ctx(n, precision=precision) as c:
                phi = c.computePotential(rho, hr)
            err = rel_l2(phi, ref)
            assert err < tol, ({L: sc.tune_fft_get(L, precision) for L in lengths}, err)
    finally:
        for L in lengths:
            sc.tune_fft_set(L, precision, 0)


@pytest.mark.parametrize("n", [(16, 16, 16), (32, 16, 24), (8, 48, 12), (128, 64, 32), (40, 20, 80), (16, 6, 10)])
def test_compute_potential_mixed(n):
    """FP32 convolution (FP64 Green's function): tolerance 1e-5 relative L2."""
    rng = np.random.default_rng(9)
    rho = rng.normal(size=n[0] * n[1] * n[2])
    hr = np.array([1.1e-4, 0.9e-4, 2.3e-4])
    ref = O.compute_potential(n, rho, hr, 0)
    with ctx(n, precision=sc.MIXED) as c:
        phi = c.computePotential(rho, hr)
    assert rel_l2(phi, ref) < 1e-5


def _self_fields_case(n, N, seed, precision=sc.FP64):
    R, P, Q = gaussian_bunch(N, seed=seed, q_total=1e-9)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n)
    ref = O.compute_self_fields(R, P, Q, dt, 1e-11, org, hr, n, want_grids=True)
    with ctx(n, precision=precision, N=N) as c:
        Ef, Bf, Q2 = c.computeSelfFields(R, P, Q, dt, 1e-11, org, hr)
        grids = dict(rho=c.get_grid(sc.GRID_RHO), phi=c.get_grid(sc.GRID_PHI),
                     E=c.get_grid(sc.GRID_EFIELD))
    return ref, Ef, Bf, Q2, grids


@pytest.mark.parametrize("n,N", [((32, 32, 32), 100000), ((16, 32, 24), 20000)])
def test_compute_self_fields_fp64(n, N):
    ref, Ef, Bf, Q2, g = _self_fields_case(n, N, seed=5)
    assert ref["dropped"] == 0
    assert np.array_equal(Q2, ref["Q"])
    assert rel_l2(g["rho"], ref["rho"]) < 1e-12
    assert rel_l2(g["phi"], ref["phi"]) < 1e-12
    assert rel_l2(g["E"], ref["E"]) < 1e-12
    assert rel_l2(Ef, ref["Ef"]) < 1e-12
    assert rel_l2(Bf, ref["Bf"]) < 1e-12


def test_compute_self_fields_mixed_within_1e5():
    ref, Ef, Bf, Q2, g = _self_fields_case((32, 32, 32), 100000, seed=6, precision=sc.MIXED)
    assert rel_l2(g["E"], ref["E"]) < 1e-5
    assert rel_l2(Ef, ref["Ef"]) < 1e-5


def test_push_and_kick_bit_exact():
    R, P, Q = gaussian_bunch(50000, seed=7)
    rng = np.random.default_rng(8)
    dt = rng.uniform(0.5e-11, 2e-11, len(R))
    Ef = rng.normal(size=R.shape) * 1e6
    Bf = rng.normal(size=R.shape) * 0.5
    with ctx((8, 8, 8), N=len(R)) as c:
        R2 = c.pushParticles(R, P, dt)
        P2 = c.kickParticles(P, Ef, Bf, dt, 1.0, O.M_P_EV)
    assert np.array_equal(R2, O.push(R, P, dt))
    assert np.array_equal(P2, O.kick(P, Ef, Bf, dt, 1.0, O.M_P_EV))


def test_beam_parameters():
    """The GPU sums are the correctly rounded sums of the per-particle terms
    (double-double across threads); the oracle sums sequentially like the
    reference.  GPU vs exact: ~1 ulp; GPU vs oracle: the oracle's rounding."""
    import math
    R, P, Q = gaussian_bunch(200000, seed=9)
    ref = O.beam_parameters(R, P)
    with ctx((8, 8, 8), N=len(R)) as c:
        s = c.calcBeamParameters(R, P)
    for k in ("rmean", "pmean", "rrms", "prms", "eps_norm"):
        assert np.allclose(s[k], ref[k], rtol=1e-10, atol=0), k
    g_exact = math.fsum(np.sqrt(1.0 + (P[:, 0] * P[:, 0] + P[:, 1] * P[:, 1] + P[:, 2] * P[:, 2]))) / len(P)
    assert abs(s["gamma_mean"] - g_exact) <= 2e-16 * g_exact
    assert abs(ref["gamma_mean"] - g_exact) <= 1e-12 * g_exact
    for d in range(3):
        assert abs(s["rmean"][d] - math.fsum(R[:, d]) / len(R)) <= 1e-15 * np.abs(R[:, d]).mean()


def test_tracker_matches_oracle_c1_shape():
    n, N, steps = (32, 32, 32), 100000, 5
    R, P, Q = gaussian_bunch(N, seed=1, q_total=1e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181))
    Ro, Po, Qo, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
        sg = c.track(steps)
        Rg, Pg, Qg = c.get_bunch()
    assert rel_l2(Rg, Ro) < 1e-12
    assert rel_l2(Pg - P, Po - P) < 1e-9
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm", "prms"):
            assert np.allclose(a[k], b[k], rtol=1e-10, atol=0), (k, a[k], b[k])


def test_tracker_sorted_and_unsorted_agree():
    """The int64 deposit makes rho independent of particle order and the
    per-particle arithmetic does not see the order at all; only the moment
    sums (double per thread, double-double across threads) can differ in the
    last bit between orders.  Re-sorting every step therefore agrees with
    never sorting to rounding level, and a fixed schedule is bit-reproducible."""
    n, N, steps = (32, 32, 32), 200000, 4
    R, P, Q = gaussian_bunch(N, seed=21, q_total=2e-9)
    out = []
    for interval in (0, 1, 1):
        with ctx(n, N=N, sort_interval=interval) as c:
            c.set_bunch(R, P, Q)
            c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
            st = c.track(steps)
            out.append((c.get_bunch(), st))
    (R0, P0, Q0), s0 = out[0]
    (R1, P1, Q1), s1 = out[1]
    (R2, P2, Q2), s2 = out[2]
    assert np.array_equal(R1, R2) and np.array_equal(P1, P2)
    assert np.array_equal(Q0, Q1)
    assert rel_l2(R0, R1) < 1e-14 and rel_l2(P0 - P, P1 - P) < 1e-12
    for a, b in zip(s0, s1):
        assert np.allclose(a["eps_norm"], b["eps_norm"], rtol=1e-12, atol=0)
        assert np.allclose(a["rrms"], b["rrms"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("precision", [sc.FP64, sc.MIXED])
def test_tracker_one_call_equals_single_steps(precision):
    """Inside one track() call the gather pass of step s also applies the first
    half push of step s+1; a call of one step pushes in the bounds pass.  Both
    schedules run the same per-particle operations, so they agree bit for bit."""
    n, N, steps = (32, 32, 32), 150000, 5
    R, P, Q = gaussian_bunch(N, seed=31, q_total=2e-9)
    runs = []
    for split in (False, True):
        with ctx(n, precision=precision, N=N, sort_interval=2) as c:
            c.set_bunch(R, P, Q)
            c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, Ez0=1e6, rf_freq=5e7)
            st = [c.track(1)[0] for _ in range(steps)] if split else c.track(steps)
            runs.append((c.get_bunch(), st, c.track_params()))
    (Ra, Pa, Qa), sa, pa = runs[0]
    (Rb, Pb, Qb), sb, pb = runs[1]
    assert np.array_equal(Ra, Rb) and np.array_equal(Pa, Pb) and np.array_equal(Qa, Qb)
    for a, b in zip(sa, sb):
        for k in ("rmean", "pmean", "rrms", "prms", "eps_norm"):
            assert np.array_equal(a[k], b[k]), k
    assert pa.t == pb.t and pa.pathLength == pb.pathLength


@pytest.mark.parametrize("precision", [sc.FP64, sc.MIXED])
def test_tracker_tilted_bunch_one_call_equals_single_steps(precision):
    """A bunch whose mean momentum points well away from z (beam frame rotated
    by ~50 degrees) and a halo: inside one call the bounds of steps 2.. come
    from the previous gather pass's chunk statistics and an exact scan of the
    chunks that can hold an extreme; single-step calls run the full bounds
    pass.  Meshes, particles and statistics agree bit for bit."""
    n, N, steps = (24, 32, 16), 300000, 6
    R, P, Q = gaussian_bunch(N, seed=33, q_total=3e-9, halo_frac=0.1, sigma_r=(1e-3, 2e-3, 0.5e-3))
    P += np.array([0.35, -0.2, 0.3])
    runs = []
    for split in (False, True):
        with ctx(n, precision=precision, N=N, sort_interval=3) as c:
            c.set_bunch(R, P, Q)
            c.track_init(1e-11, ref_p=(0.35, -0.2, 0.346181), Bz=0.2)
            st, meshes = [], []
            if split:
                for _ in range(steps):
                    st += c.track(1)
                    meshes.append(c.mesh())
            else:
                st = c.track(steps)
                meshes.append(c.mesh())
            runs.append((c.get_bunch(), st, meshes[-1]))
    (Ra, Pa, Qa), sa, ma = runs[0]
    (Rb, Pb, Qb), sb, mb = runs[1]
    assert np.array_equal(ma[0], mb[0]) and np.array_equal(ma[1], mb[1])
    assert np.array_equal(Ra, Rb) and np.array_equal(Pa, Pb)
    for a, b in zip(sa, sb):
        for k in ("rmean", "pmean", "rrms", "prms", "eps_norm"):
            assert np.array_equal(a[k], b[k]), k


def test_tracker_without_space_charge_matches_oracle():
    n, N, steps = (16, 16, 16), 50000, 4
    R, P, Q = gaussian_bunch(N, seed=5, q_total=1e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, space_charge=False)
    Ro, Po, _, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, space_charge=False)
        sg = c.track(steps)
        Rg, Pg, _ = c.get_bunch()
    assert np.array_equal(Rg, Ro) and np.array_equal(Pg, Po)
    for a, b in zip(sg, so):
        assert np.allclose(a["rrms"], b["rrms"], rtol=1e-12, atol=0)


def test_tracker_matches_oracle_solenoid_1e6_64():
    n, N, steps = (64, 64, 64), 1000000, 3
    R, P, Q = gaussian_bunch(N, seed=2, q_total=10e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
    Ro, Po, _, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, N=N, sort_interval=2) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
        sg = c.track(steps)
        Rg, Pg, _ = c.get_bunch()
    assert rel_l2(Rg, Ro) < 1e-12
    assert rel_l2(Pg - P, Po - P) < 1e-10
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm"):
            for d in range(2):
                assert abs(a[k][d] / b[k][d] - 1) < 1e-10, (k, d, a[k], b[k])
        assert abs(a["rrms"][2] / b["rrms"][2] - 1) < 1e-10
        # longitudinal emittance: psq_z = sum pz^2 - N<pz>^2 cancels (<pz>/sigma_pz)^2 ~ 2e3,
        # so the reference's own sequential sum carries ~1e-9 here
        assert abs(a["eps_norm"][2] / b["eps_norm"][2] - 1) < 1e-8


@pytest.mark.slow
def test_compute_self_fields_fp64_c2_full_size():
    """Plain FP64 drop-in configuration: 1e7 particles, 128^3 grid (256^3 doubled FFT)."""
    n, N = (128, 128, 128), 10_000_000
    R, P, Q = gaussian_bunch(N, seed=2, q_total=10e-9)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n)
    lgp_o, dpos_o, _ = O.cic_index(R, org, hr)
    ref = O.compute_self_fields(R, P, Q, dt, 1e-11, org, hr, n, want_grids=True)
    with ctx(n, N=N) as c:
        hr_g, org_g = c.boundp(R)
        lgp, dpos = c.cic_index(R, org, hr)
        Ef, Bf, Q2 = c.computeSelfFields(R, P, Q, dt, 1e-11, org, hr)
        phi = c.get_grid(sc.GRID_PHI)
        E = c.get_grid(sc.GRID_EFIELD)
        rho = c.get_grid(sc.GRID_RHO)
    assert np.array_equal(hr_g, hr) and np.array_equal(org_g, org)
    assert np.array_equal(lgp, lgp_o) and np.array_equal(dpos, dpos_o)
    assert np.array_equal(Q2, ref["Q"])
    for name, a, b in (("rho", rho, ref["rho"]), ("phi", phi, ref["phi"]), ("E", E, ref["E"]),
                       ("Ef", Ef, ref["Ef"]), ("Bf", Bf, ref["Bf"])):
        err = rel_l2(a, b)
        print("C2 %s rel L2 %.3e" % (name, err))
        assert err < 1e-12, name


@pytest.mark.parametrize("precision,tol", [(sc.FP64, 1e-12), (sc.MIXED, 1e-5)])
@pytest.mark.parametrize("n,zshift", [((16, 16, 16), -0.021), ((32, 16, 24), 0.013), ((12, 24, 8), -0.35)])
def test_compute_potential_shifted(n, zshift, precision, tol):
    """Image-charge solve (shifted Green's function, negated, z flipped).

    The reference's image tables are 8-corner differences of FP64 corner terms
    whose rounding the difference amplifies by ~(r/h)^3: with the image 0.35 m
    (~1500 cells) away its G is off by ~7e-6 relative.  The FP64 solve keeps
    those bits; the mixed solve's cell-average form is accurate to 5e-10
    (tests/test_igf_term.py), so it agrees with the reference to its own
    tolerance plus the reference tables' error, measured here against that form."""
    rng = np.random.default_rng(17)
    rho = rng.normal(size=n[0] * n[1] * n[2])
    hr = np.array([1.1e-4, 0.9e-4, 2.3e-4])
    ref = O.compute_potential_shifted(n, rho, hr, zshift)
    with ctx(n, precision=precision) as c:
        phi = c.computePotentialShifted(rho, hr, zshift)
        phi0 = c.computePotential(rho, hr)  # the plain solve still works after the image one
    ref_table_err = 0.0
    if precision == sc.MIXED:
        for table in (1, 2):
            g_ref = sc.host_green_octant(n, hr, zshift, table, fast=False)
            g_acc = sc.host_green_octant(n, hr, zshift, table, fast=True)
            ref_table_err = max(ref_table_err, float(np.max(np.abs(g_ref - g_acc) / np.abs(g_acc))))
        if abs(zshift) < 0.1:
            assert ref_table_err < 1e-7  # the image within ~200 cells: the tables agree closely
    assert rel_l2(phi, ref) < tol + 2 * ref_table_err
    assert rel_l2(phi0, O.compute_potential(n, rho, hr, 0)) < max(tol, 1e-12)


def _binned_case(N, n, nbins, seed, precision, image, given_gammas):
    R, P, _ = gaussian_bunch(N, seed=seed, q_total=1e-9)
    R = R + np.array([0.0, 0.0, 0.004])  # bunch 4 mm downstream of the cathode plane z = 0
    P[:, 2] += np.linspace(0.0, 0.3, N)  # energy spread across the bins
    rng = np.random.default_rng(seed + 1)
    dt = rng.uniform(0.8e-11, 1.2e-11, N)
    # bins by energy; bin 1 left empty, a few particles in no bin
    bins = np.minimum((np.argsort(np.argsort(P[:, 2])) * nbins) // N, nbins - 1).astype(np.int32)
    bins[bins == 1] = 2
    bins[rng.random(N) < 0.01] = -1
    bins[:3] = nbins  # out of range: no bin
    qi = 1e-9 / N
    hr, org = O.boundp(R, n)
    Bf0 = rng.normal(size=(N, 3)) * 1e-3
    g = None
    if given_gammas:
        g = 1.0 + rng.uniform(0.01, 0.5, nbins)
    ref = O.compute_self_fields_binned(R, P, bins, nbins, qi, dt, 1e-11, org, hr, n, image=image,
                                       gammas=g, Bf=Bf0)
    with ctx(n, precision=precision, N=N) as c:
        Ef, Bf = c.computeSelfFieldsBinned(R, P, bins, nbins, qi, dt, 1e-11, org, hr, gammas=g,
                                           image=image, Bf=Bf0)
        # a second call on the same context gives the same fields
        Ef2, Bf2 = c.computeSelfFieldsBinned(R, P, bins, nbins, qi, dt, 1e-11, org, hr, gammas=g,
                                             image=image, Bf=Bf0)
    assert np.array_equal(Ef, Ef2) and np.array_equal(Bf, Bf2)
    return ref, Ef, Bf, Bf0


@pytest.mark.parametrize("image", [True, False])
@pytest.mark.parametrize("given_gammas", [True, False])
def test_compute_self_fields_binned_fp64(image, given_gammas):
    """Energy-binned self fields with the cathode image charge, FP64 path."""
    ref, Ef, Bf, Bf0 = _binned_case(60000, (16, 24, 32), 5, 21, sc.FP64, image, given_gammas)
    assert ref["dropped"] == 0
    assert rel_l2(Ef, ref["Ef"]) < 1e-12
    assert rel_l2(Bf - Bf0, ref["Bf"] - Bf0) < 1e-12
    assert np.array_equal(Bf[:, 2], Bf0[:, 2])


def test_compute_self_fields_binned_mixed():
    ref, Ef, Bf, Bf0 = _binned_case(60000, (32, 32, 32), 4, 22, sc.MIXED, True, False)
    assert rel_l2(Ef, ref["Ef"]) < 1e-5
    assert rel_l2(Bf - Bf0, ref["Bf"] - Bf0) < 1e-5


def test_compute_self_fields_binned_rejects_bad_nbins():
    R, P, _ = gaussian_bunch(100, seed=3)
    dt = np.full(100, 1e-11)
    hr, org = O.boundp(R, (8, 8, 8))
    with ctx((8, 8, 8), N=100) as c:
        for nb in (0, 65):
            with pytest.raises(sc.OpalscError):
                c.computeSelfFieldsBinned(R, P, np.zeros(100, np.int32), nb, 1e-12, dt, 1e-11, org, hr)


def geometric_energy_bins(P, nbins):
    """Bins by longitudinal momentum rank with populations in the ratio 2:1
    from one bin to the next (bin 0 the largest, holding the lowest pz)."""
    N = len(P)
    w = 0.5 ** np.arange(nbins)
    edges = np.floor(np.cumsum(w) / w.sum() * N).astype(np.int64)
    edges[-1] = N
    rank = np.empty(N, np.int64)
    rank[np.argsort(P[:, 2], kind="stable")] = np.arange(N)
    return np.searchsorted(edges, rank, side="right").astype(np.int32)


def _binned_track_case(n, N, nbins, steps, precision=sc.FP64, sort_interval=0, image=True, nobin=0.0, seed=31):
    R, P, Q = gaussian_bunch(N, seed=seed, q_total=1e-9, sigma_p=(1e-3, 1e-3, 5e-3))
    R[:, 2] += 0.005  # 5 sigma downstream of the cathode plane
    bins = geometric_energy_bins(P, nbins)
    if nobin:
        bins[np.random.default_rng(seed).random(N) < nobin] = -1
    qi = 1e-9 / N
    dt = np.full(N, 1e-11)
    kw = dict(Bz=1.0, Ez0=1e6, rf_freq=50e6, ref_p=(0.0, 0.0, 0.046181))
    tr = O.make_tracker(n, 1e-11, **kw)
    ref = O.track_binned(tr, R, P, Q, dt, bins, nbins, qi, steps, image=image)
    with ctx(n, precision=precision, N=N, sort_interval=sort_interval) as c:
        c.set_bunch(R, P, Q)
        c.set_bins(bins, nbins, qi, image=image)
        counts, _ = c.get_bins()
        c.track_init(1e-11, **kw)
        sg = c.track(steps)
        Rg, Pg, Qg = c.get_bunch()
        _, gammas = c.get_bins()
    assert np.array_equal(counts, np.bincount(bins[bins >= 0], minlength=nbins))
    assert np.array_equal(Qg, Q)  # the bin loop restores Q
    return R, P, ref, (Rg, Pg, sg), gammas


@pytest.mark.parametrize("sort_interval", [0, 2])
def test_tracker_binned_matches_oracle(sort_interval):
    """Energy bins + image charge + solenoid + RF field, FP64 path."""
    R, P, ref, gpu, gammas = _binned_track_case((16, 16, 32), 60000, 6, 3, sort_interval=sort_interval,
                                                nobin=0.02)
    Ro, Po, _, _, so = ref
    Rg, Pg, sg = gpu
    assert np.all(gammas > 1.0)
    assert rel_l2(Rg, Ro) < 1e-12
    assert rel_l2(Pg - P, Po - P) < 1e-9
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm", "prms"):
            assert np.allclose(a[k], b[k], rtol=1e-10, atol=0), (k, a[k], b[k])


@pytest.mark.parametrize("precision", [sc.FP64, sc.MIXED])
@pytest.mark.parametrize("tiny_bin", [False, True])
def test_tracker_binned_one_call_equals_single_steps(precision, tiny_bin):
    """Energy bins with an empty bin and unbinned particles, re-sorted inside
    the call.  Inside one call each bin's mean gamma comes from the chunk sums
    the previous gather pass left; a call of one step sums the same chunks
    itself.  With every non-empty bin larger than a gather chunk both agree bit
    for bit; a bin smaller than a chunk makes every step sum over the bins'
    momenta, which agrees as well."""
    n, N, steps = (16, 16, 32), 120000, 5
    R, P, Q = gaussian_bunch(N, seed=41, q_total=1e-9, sigma_p=(1e-3, 1e-3, 5e-3))
    R[:, 2] += 0.005
    rank = np.empty(N, np.int64)
    rank[np.argsort(P[:, 2], kind="stable")] = np.arange(N)
    # bins 0, 1, 3, 4 by pz rank (bin 2 empty), the highest-pz group unbinned
    edges = np.array([50000, 80000, 100000, 100300 if tiny_bin else 117000, N])
    bins = np.array([0, 1, 3, 4, -1], np.int32)[np.searchsorted(edges, rank, side="right")]
    bins[np.random.default_rng(5).random(N) < 0.02] = -1
    runs = []
    for split in (False, True):
        with ctx(n, precision=precision, N=N, sort_interval=2) as c:
            c.set_bunch(R, P, Q)
            c.set_bins(bins, 5, 1e-9 / N, image=True)
            c.track_init(1e-11, Bz=1.0, Ez0=1e6, rf_freq=50e6, ref_p=(0.0, 0.0, 0.046181))
            st = [c.track(1)[0] for _ in range(steps)] if split else c.track(steps)
            runs.append((c.get_bunch(), st, c.get_bins()[1]))
    (Ra, Pa, _), sa, ga = runs[0]
    (Rb, Pb, _), sb, gb = runs[1]
    counts = np.bincount(bins[bins >= 0], minlength=5)
    assert counts[2] == 0 and (counts[[0, 1, 3, 4]].min() < 1024) == tiny_bin
    assert np.array_equal(ga, gb)
    assert np.array_equal(Ra, Rb) and np.array_equal(Pa, Pb)
    for a, b in zip(sa, sb):
        for k in ("rmean", "pmean", "rrms", "prms", "eps_norm"):
            assert np.array_equal(a[k], b[k]), k


def test_tracker_binned_without_image_and_mixed():
    R, P, ref, gpu, _ = _binned_track_case((32, 32, 32), 80000, 10, 3, precision=sc.MIXED, image=False)
    Ro, Po, _, _, so = ref
    Rg, Pg, sg = gpu
    assert rel_l2(Pg - P, Po - P) < 1e-5
    for a, b in zip(sg, so):
        assert np.allclose(a["eps_norm"], b["eps_norm"], rtol=1e-6, atol=0)


@pytest.mark.parametrize("precision,tol", [(sc.FP64, 1e-12), (sc.MIXED, 1e-5)])
@pytest.mark.parametrize("greens", [0, 1])
@pytest.mark.parametrize("n", [(16, 16, 16), (32, 16, 24), (12, 24, 8), (64, 32, 96), (10, 16, 20)])
def test_compute_potential_periodic_z(n, greens, precision, tol):
    """DC-beam solver (BCFFTZ=PERIODIC): x and y doubled, z periodic with the
    one-sided Green's function."""
    rng = np.random.default_rng(19)
    rho = rng.normal(size=n[0] * n[1] * n[2])
    hr = np.array([1.1e-4, 0.9e-4, 2.3e-4])
    ref = O.compute_potential_periodic_z(n, rho, hr, greens)
    with sc.SpaceChargeContext(n, precision=precision, greens=greens, max_particles=0,
                               bcz=sc.BC_PERIODIC) as c:
        phi = c.computePotential(rho, hr)
        phi2 = c.computePotential(rho, hr)
    assert np.array_equal(phi, phi2)
    assert rel_l2(phi, ref) < tol


def test_periodic_z_context_rejects_bins_and_image_charge():
    R, P, Q = gaussian_bunch(100, seed=3)
    dt = np.full(100, 1e-11)
    with sc.SpaceChargeContext((8, 8, 8), max_particles=100, bcz=sc.BC_PERIODIC) as c:
        hr, org = c.boundp(R)
        with pytest.raises(sc.OpalscError, match="open-boundary"):
            c.computePotentialShifted(np.zeros(512), hr, 0.1)
        with pytest.raises(sc.OpalscError, match="open-boundary"):
            c.computeSelfFieldsBinned(R, P, np.zeros(100, np.int32), 1, 1e-12, dt, 1e-11, org, hr)
        c.set_bunch(R, P, Q)
        with pytest.raises(sc.OpalscError, match="open-boundary"):
            c.set_bins(np.zeros(100, np.int32), 1, 1e-12)
        with pytest.raises(sc.OpalscError, match="beam frequency"):
            c.setBeamFrequency(0.0)
    with pytest.raises(sc.OpalscError, match="periodic z"):
        sc.SpaceChargeContext((8, 8, 6), max_particles=0, bcz=sc.BC_PERIODIC)


def _dc_bunch(N, seed, zlen):
    """Round beam, uniform in z over zlen, <P> exactly along z (momenta in +- pairs)."""
    rng = np.random.default_rng(seed)
    R = np.column_stack([rng.normal(0, 1e-3, N), rng.normal(0, 1e-3, N), rng.uniform(-zlen / 2, zlen / 2, N)])
    P = np.zeros((N, 3))
    for d in (0, 1):
        P[:, d] = rng.normal(0, 1e-3, N // 2).repeat(2) * np.tile([1.0, -1.0], N // 2)
    P[:, 2] = 0.046181 + rng.normal(0, 1e-3, N)
    return R, P, np.full(N, 1e-9 / N)


@pytest.mark.parametrize("freq", [1e9, 2.5e10])
def test_dc_boundp_and_wrap_bit_exact(freq):
    """DC-beam mesh rule (z given the period c/f) and the particle wrap, against
    the oracle bit for bit; the second frequency makes the period shorter than
    the bunch, so particles wrap."""
    n = (16, 16, 32)
    R, P, Q = _dc_bunch(20000, seed=12, zlen=0.02)
    period = O.CLIGHT / freq
    hro, orgo = O.boundp(R, n, period=period)
    Rwo = O.wrap_periodic_z(R, orgo, hro, n)
    with sc.SpaceChargeContext(n, max_particles=len(R), bcz=sc.BC_PERIODIC) as c:
        c.setBeamFrequency(freq)
        hr, org = c.boundp(R)
        Rw = c.wrap_periodic_z(R, org, hr)
    assert np.array_equal(hr, hro) and np.array_equal(org, orgo)
    assert np.array_equal(Rw, Rwo)
    assert (not np.array_equal(Rw, R)) == (period < 0.02)


@pytest.mark.parametrize("precision,tol", [(sc.FP64, 1e-12), (sc.MIXED, 1e-5)])
def test_dc_self_fields_match_oracle(precision, tol):
    """computeSelfFields of a DC beam: periodic-z solve and periodic gather."""
    n, N, freq = (16, 16, 32), 60000, 2.5e10
    R, P, Q = _dc_bunch(N, seed=13, zlen=0.02)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n, period=O.CLIGHT / freq)
    Rw = O.wrap_periodic_z(R, org, hr, n)
    ref = O.compute_self_fields(Rw, P, Q, dt, 1e-11, org, hr, n, want_grids=True, periodic_z=True)
    with sc.SpaceChargeContext(n, precision=precision, max_particles=N, bcz=sc.BC_PERIODIC) as c:
        Ef, Bf, Q2 = c.computeSelfFields(Rw, P, Q, dt, 1e-11, org, hr)
        E = c.get_grid(sc.GRID_EFIELD)
    assert rel_l2(E, ref["E"]) < tol
    assert rel_l2(Ef, ref["Ef"]) < tol
    assert rel_l2(Bf, ref["Bf"]) < tol
    # the top layer's particles read layer 0 through the periodic guard
    top = np.floor((Rw[:, 2] - org[2]) / hr[2] - 0.5) >= n[2] - 1
    assert top.any()


def test_dc_tracker_matches_oracle():
    """The DC-beam tracker (period shorter than the bunch: particles wrap in the
    beam frame every step) against the oracle's ParallelTTracker loop."""
    n, N, steps, freq = (16, 16, 32), 50000, 4, 2.5e10
    R, P, Q = _dc_bunch(N, seed=14, zlen=0.02)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, periodic_z=True,
                        period=O.CLIGHT / freq)
    Ro, Po, Qo, _, so = O.track(tr, R, P, Q, dt, steps)
    with sc.SpaceChargeContext(n, max_particles=N, bcz=sc.BC_PERIODIC, sort_interval=2) as c:
        c.setBeamFrequency(freq)
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
        sg = c.track(steps)
        Rg, Pg, Qg = c.get_bunch()
    assert np.abs(Ro[:, 2] - R[:, 2]).max() > 1e-3  # some particles wrapped
    assert rel_l2(Rg, Ro) < 1e-12
    assert rel_l2(Pg - P, Po - P) < 1e-9
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm", "prms"):
            assert np.allclose(a[k], b[k], rtol=1e-10, atol=0), (k, a[k], b[k])


def test_dc_tracker_mixed_close_to_fp64():
    n, N, steps, freq = (16, 16, 32), 50000, 4, 2.5e10
    R, P, Q = _dc_bunch(N, seed=15, zlen=0.02)
    out = []
    for prec in (sc.FP64, sc.MIXED):
        with sc.SpaceChargeContext(n, precision=prec, max_particles=N, bcz=sc.BC_PERIODIC) as c:
            c.setBeamFrequency(freq)
            c.set_bunch(R, P, Q)
            c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
            out.append(c.track(steps))
    for a, b in zip(out[0], out[1]):
        for k in ("rrms", "eps_norm"):
            assert np.allclose(a[k], b[k], rtol=1e-4, atol=0), (k, a[k], b[k])


def _batch_setups():
    """Four independent bunches: different sizes, grids, precisions, one with energy bins."""
    out = []
    specs = [(3000, (16, 16, 16), sc.FP64, 0.0, 0), (20000, (32, 32, 32), sc.MIXED, 0.5, 0),
             (5000, (12, 24, 16), sc.FP64, 1.0, 3), (60000, (16, 32, 24), sc.MIXED, 0.0, 0)]
    for k, (N, n, prec, Bz, nbins) in enumerate(specs):
        R, P, Q = gaussian_bunch(N, seed=40 + k, q_total=1e-9 * (k + 1))
        bins = None
        if nbins:
            R[:, 2] += 0.005
            bins = geometric_energy_bins(P, nbins)
        out.append(dict(N=N, n=n, prec=prec, Bz=Bz, R=R, P=P, Q=Q, bins=bins, nbins=nbins))
    return out


def _make_batch_ctx(u):
    c = sc.SpaceChargeContext(u["n"], precision=u["prec"], max_particles=u["N"])
    c.set_bunch(u["R"], u["P"], u["Q"])
    if u["nbins"]:
        c.set_bins(u["bins"], u["nbins"], 1e-9 / u["N"], image=True)
    c.track_init(1e-11, Bz=u["Bz"], ref_p=(0.0, 0.0, 0.046181))
    return c


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_steps(graph):
    """A batched step is opalsc_track(ctx, 1): same bunch and statistics bit for bit."""
    setups = _batch_setups()
    steps = 4
    ref = []
    for u in setups:
        c = _make_batch_ctx(u)
        st = [c.track(1)[0] for _ in range(steps)]
        ref.append((c.get_bunch(), st))
        c.close()
    ctxs = [_make_batch_ctx(u) for u in setups]
    with sc.Batch(ctxs) as b:
        st1 = b.step(2, graph=graph)
        st2 = b.step(2, graph=graph)  # the second call replays the captured graph
    sts = st1 + st2
    for i, (c, (bunch_ref, st_ref)) in enumerate(zip(ctxs, ref)):
        R, P, Q = c.get_bunch()
        assert np.array_equal(R, bunch_ref[0]) and np.array_equal(P, bunch_ref[1]), i
        for s in range(steps):
            for k in ("rrms", "prms", "eps_norm", "rmean", "pmean"):
                assert np.array_equal(sts[s][i][k], st_ref[s][k]), (i, s, k)
        c.close()


def _fused_setups():
    """Bunches of every kernel class the fused batch forms: FP64 and mixed,
    equal and unequal charges, with and without an external field, without
    space charge; two share a grid shape (one grid group, two slots), two
    re-sort within the call (every 2 and 3 steps)."""
    specs = [  # N, n, precision, Bz, unequal charges, space charge, sort interval
        (3000, (16, 16, 16), sc.FP64, 0.0, False, True, 0),
        (20000, (32, 32, 32), sc.MIXED, 0.5, False, True, 2),
        (5000, (12, 24, 16), sc.FP64, 1.0, True, True, 3),
        (60000, (16, 32, 24), sc.MIXED, 0.0, True, True, 0),
        (7000, (32, 32, 32), sc.MIXED, 0.0, False, True, 0),
        (4000, (16, 16, 16), sc.MIXED, 0.3, False, False, 0),
    ]
    out = []
    for k, (N, n, prec, Bz, uneq, scon, si) in enumerate(specs):
        R, P, Q = gaussian_bunch(N, seed=70 + k, q_total=1e-9 * (k + 1))
        if uneq:
            Q = Q * np.random.default_rng(80 + k).uniform(0.7, 1.3, N)
        out.append(dict(N=N, n=n, prec=prec, Bz=Bz, R=R, P=P, Q=Q, sc=scon, si=si))
    return out


def _make_fused_ctx(u):
    c = sc.SpaceChargeContext(u["n"], precision=u["prec"], max_particles=u["N"], sort_interval=u["si"])
    c.set_bunch(u["R"], u["P"], u["Q"])
    c.track_init(1e-11, Bz=u["Bz"], ref_p=(0.0, 0.0, 0.046181), space_charge=u["sc"])
    return c


@pytest.mark.parametrize("graph", [True, False])
def test_fused_batch_equals_single_steps(graph):
    """The fused batch (one launch per pass and kernel class, one per pass and
    grid shape) equals opalsc_track(ctx, 1) per bunch bit for bit, with re-sorts
    inside a call (launched directly) between graph replays."""
    setups = _fused_setups()
    steps = 7
    ref = []
    for u in setups:
        c = _make_fused_ctx(u)
        st = [c.track(1)[0] for _ in range(steps)]
        ref.append((c.get_bunch(), st))
        c.close()
    ctxs = [_make_fused_ctx(u) for u in setups]
    try:
        with sc.Batch(ctxs) as b:
            sts = b.step(4, graph=graph) + b.step(3, graph=graph)
        for i, (c, (bunch_ref, st_ref)) in enumerate(zip(ctxs, ref)):
            R, P, Q = c.get_bunch()
            assert np.array_equal(R, bunch_ref[0]) and np.array_equal(P, bunch_ref[1]), i
            assert np.array_equal(Q, bunch_ref[2]), i
            for s in range(steps):
                for k in ("rrms", "prms", "eps_norm", "rmean", "pmean", "t", "dropped"):
                    assert np.array_equal(sts[s][i][k], st_ref[s][k]), (i, s, k)
        # a single step of a context after the batch continues from the batch's state
        u = setups[1]
        c = _make_fused_ctx(u)
        for _ in range(steps + 1):
            c.track(1)
        ctxs[1].track(1)
        assert np.array_equal(ctxs[1].get_bunch()[0], c.get_bunch()[0])
        c.close()
    finally:
        for c in ctxs:
            c.close()


def test_fused_and_per_bunch_batches_agree():
    """The fused and the per-bunch forms of a batch give the same steps."""
    setups = [u for u in _fused_setups() if u["si"] == 0]
    out = []
    for fused in (True, False):
        ctxs = [_make_fused_ctx(u) for u in setups]
        with sc.Batch(ctxs, fused=fused) as b:
            st = b.step(3)
        out.append(([c.get_bunch() for c in ctxs], st))
        for c in ctxs:
            c.close()
    (b1, s1), (b2, s2) = out
    for i in range(len(setups)):
        assert np.array_equal(b1[i][0], b2[i][0]) and np.array_equal(b1[i][1], b2[i][1]), i
        for s in range(3):
            assert np.array_equal(s1[s][i]["eps_norm"], s2[s][i]["eps_norm"]), (i, s)


def test_batch_recaptures_after_a_new_bunch():
    setups = _batch_setups()[:2]
    ctxs = [_make_batch_ctx(u) for u in setups]
    with sc.Batch(ctxs) as b:
        b.step(1)
        # a new bunch in context 0 (other size): the graph must not replay stale buffers
        R, P, Q = gaussian_bunch(1500, seed=99, q_total=1e-9)
        ctxs[0].set_bunch(R, P, Q)
        ctxs[0].track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
        st = b.step(2)
    c = sc.SpaceChargeContext(setups[0]["n"], precision=setups[0]["prec"], max_particles=3000)
    c.set_bunch(R, P, Q)
    c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
    ref = [c.track(1)[0] for _ in range(2)]
    assert np.array_equal(ctxs[0].get_bunch()[0], c.get_bunch()[0])
    assert np.array_equal(st[1][0]["eps_norm"], ref[1]["eps_norm"])
    for x in ctxs + [c]:
        x.close()


@pytest.mark.parametrize("precision,tol_r,tol_s", [(sc.FP64, 1e-12, 1e-10), (sc.MIXED, 1e-9, 1e-6)])
def test_tracker_with_unequal_charges(precision, tol_r, tol_s):
    """Per-particle macro charges (the deposit reads Q and, on the FP64 path,
    leaves it as (Q*dT)/dT); bunches of equal charges keep Q as one number."""
    n, N, steps = (32, 32, 32), 80000, 4
    R, P, Q = gaussian_bunch(N, seed=12, q_total=1e-9)
    Q = Q * np.random.default_rng(13).uniform(0.7, 1.3, N)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
    Ro, Po, Qo, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, precision=precision, N=N, sort_interval=2) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
        sg = c.track(steps)
        Rg, Pg, Qg = c.get_bunch()
    assert rel_l2(Rg, Ro) < tol_r
    if precision == sc.FP64:
        assert np.array_equal(Qg, Qo)
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm"):
            assert np.allclose(a[k], b[k], rtol=tol_s, atol=0), (k, a[k], b[k])


def test_tracker_equal_charges_round_trip_like_the_reference():
    """A bunch of equal charges: Q after tracking equals the oracle's per-particle (Q*dT)/dT."""
    n, N, steps = (16, 16, 16), 20000, 3
    R, P, Q = gaussian_bunch(N, seed=14, q_total=3e-9)
    Q[:] = 8.294255678822374e-13  # (Q*dT)/dT rounds to the next double
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181))
    _, _, Qo, _, _ = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
        c.track(steps)
        _, _, Qg = c.get_bunch()
    assert np.array_equal(Qg, Qo)


@pytest.mark.parametrize("precision", [sc.FP64, sc.MIXED])
@pytest.mark.parametrize("nbins", [0, 4])
def test_green_function_overlap_changes_nothing(precision, nbins):
    """On large grids the tracker computes the Green's functions on a second
    stream beside the deposit; with profiling on it runs them in line.  Both
    schedules give the same bunch and statistics bit for bit."""
    n, N, steps = (96, 96, 96), 100000, 3
    R, P, Q = gaussian_bunch(N, seed=15, q_total=1e-9, sigma_p=(1e-3, 1e-3, 5e-3))
    R[:, 2] += 0.005
    out = []
    for prof in (False, True):
        with ctx(n, precision=precision, N=N) as c:
            c.set_bunch(R, P, Q)
            if nbins:
                c.set_bins(geometric_energy_bins(P, nbins), nbins, 1e-9 / N, image=True)
            c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
            c.profile_enable(prof)
            st = c.track(steps)
            out.append((c.get_bunch(), st))
    (R0, P0, _), s0 = out[0]
    (R1, P1, _), s1 = out[1]
    assert np.array_equal(R0, R1) and np.array_equal(P0, P1)
    for a, b in zip(s0, s1):
        assert np.array_equal(a["eps_norm"], b["eps_norm"])


def test_compute_self_fields_particles_outside_the_mesh():
    """A mesh (from boundp of the core) that leaves a tenth of the bunch outside:
    their deposits are dropped and they gather zero field (ZeroFace)."""
    n, N = (32, 32, 32), 50000
    R, P, Q = gaussian_bunch(N, seed=23, q_total=1e-9)
    dt = np.full(N, 1e-11)
    r2 = (R ** 2).sum(1)
    core = R[r2 < np.quantile(r2, 0.9)]
    hr, org = O.boundp(core, n)
    ref = O.compute_self_fields(R, P, Q, dt, 1e-11, org, hr, n, want_grids=True)
    assert ref["dropped"] > 0
    with ctx(n, N=N) as c:
        Ef, Bf, Q2 = c.computeSelfFields(R, P, Q, dt, 1e-11, org, hr)
        rho = c.get_grid(sc.GRID_RHO)
    assert rel_l2(rho, ref["rho"]) < 1e-12
    assert rel_l2(Ef, ref["Ef"]) < 1e-12
    assert rel_l2(Bf, ref["Bf"]) < 1e-12
    t = (R - org) / hr  # cell coordinates: more than two cells outside the mesh reads no field
    far = ((t < -2) | (t > np.array(n) + 1)).any(axis=1)
    assert far.sum() > 0
    assert np.all(Ef[far] == 0.0) and np.all(ref["Ef"][far] == 0.0)


@pytest.mark.parametrize("N", [1, 7])
def test_compute_self_fields_tiny_bunches(N):
    """One particle (boundp's degenerate-length branch in every dimension) and a few."""
    n = (8, 8, 8)
    R, P, Q = gaussian_bunch(N, seed=24, q_total=1e-12)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n)
    ref = O.compute_self_fields(R, P, Q, dt, 1e-11, org, hr, n, want_grids=True)
    with ctx(n, N=16) as c:
        hr2, org2 = c.boundp(R)
        Ef, Bf, Q2 = c.computeSelfFields(R, P, Q, dt, 1e-11, org, hr)
        rho = c.get_grid(sc.GRID_RHO)
        E = c.get_grid(sc.GRID_EFIELD)
    assert np.array_equal(hr2, hr) and np.array_equal(org2, org)
    assert np.array_equal(Q2, ref["Q"])
    assert rel_l2(rho, ref["rho"]) < 1e-12
    assert rel_l2(E, ref["E"]) < 1e-12
    # a lone particle sits where its own field cancels: compare on the grid's scale
    scale = np.abs(ref["E"]).max()
    assert np.abs(Ef - ref["Ef"]).max() <= 1e-12 * scale


def test_tracker_few_particles():
    n, N, steps = (8, 8, 8), 5, 4
    R, P, Q = gaussian_bunch(N, seed=25, q_total=1e-12)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.3)
    Ro, Po, _, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.3)
        sg = c.track(steps)
        Rg, Pg, _ = c.get_bunch()
    assert rel_l2(Rg, Ro) < 1e-12
    assert rel_l2(Pg - P, Po - P) < 1e-9
    for a, b in zip(sg, so):
        assert np.allclose(a["rrms"], b["rrms"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("precision,tol_r,tol_s", [(sc.FP64, 1e-12, 1e-10), (sc.MIXED, 1e-9, 1e-6)])
def test_tracker_on_a_radix5_grid(precision, tol_r, tol_s):
    """A 40 x 20 x 40 grid (transform lengths 80 and 40: a radix-5 stage) in the
    tracker, five steps against the oracle."""
    n, N, steps = (40, 20, 40), 30000, 5
    R, P, Q = gaussian_bunch(N, seed=17, q_total=1e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.4)
    Ro, Po, _, _, so = O.track(tr, R, P, Q, dt, steps)
    with ctx(n, precision=precision, N=N, sort_interval=2) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.4)
        sg = c.track(steps)
        Rg, Pg, _ = c.get_bunch()
    assert rel_l2(Rg, Ro) < tol_r
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm"):
            assert np.allclose(a[k], b[k], rtol=tol_s, atol=0), (k, a[k], b[k])


@pytest.mark.parametrize("precision,N", [(sc.FP64, 511), (sc.FP64, 513), (sc.FP64, 4097),
                                         (sc.MIXED, 1023), (sc.MIXED, 1025), (sc.MIXED, 8193)])
def test_tracker_ragged_block_edges(precision, N):
    """Bunch sizes one either side of the gather pass's block edges: FP64
    one-shot blocks of 512 particles (4096 per XCD share), mixed one-shot
    blocks of 1024 (8192 per XCD share, with the N-1 clamp of their loads);
    a partial last block and XCD shares of unequal length.  Mixed tolerances:
    positions 1e-9, statistics 1e-6 relative."""
    n, steps = (8, 8, 8), 3
    R, P, Q = gaussian_bunch(N, seed=31, q_total=1e-11)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.3)
    Ro, Po, _, _, so = O.track(tr, R, P, Q, dt, steps)
    fp64 = precision == sc.FP64
    with ctx(n, precision=precision, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.3)
        sg = c.track(steps)
        Rg, Pg, _ = c.get_bunch()
    assert rel_l2(Rg, Ro) < (1e-12 if fp64 else 1e-9)
    assert rel_l2(Pg - P, Po - P) < (1e-9 if fp64 else 1e-6)
    rtol = 1e-10 if fp64 else 1e-6
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm", "prms"):
            assert np.allclose(a[k], b[k], rtol=rtol, atol=0), (k, a[k], b[k])
        tol_mean = 1e-12 if fp64 else 1e-6
        assert np.all(np.abs(np.asarray(a["rmean"]) - b["rmean"]) <= tol_mean * np.asarray(b["rrms"]))


class _HipArrays:
    """Device arrays from the HIP runtime the library itself links (hipMalloc/hipMemcpy via ctypes)."""

    def __init__(self):
        import ctypes
        import ctypes.util
        self.C = ctypes
        path = None
        for line in open("/proc/self/maps"):
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
        self.hip = ctypes.CDLL(path or "libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), self.C.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(self.C.c_void_p), self.C.c_size_t(a.nbytes), 1) == 0
        self.ptrs.append(p)
        return p, a.shape

    def get(self, pa):
        p, shape = pa
        out = np.zeros(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), p, self.C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def test_device_array_entry_points_equal_host_ones():
    """opalsc_compute_self_fields_dev / opalsc_compute_potential_dev on arrays
    already in HBM give what the host-array entry points give, bit for bit;
    a host array passed as a device one is refused."""
    n, N = (32, 32, 32), 40000
    R, P, Q = gaussian_bunch(N, seed=26, q_total=1e-9)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n)
    Bf0 = np.random.default_rng(27).normal(size=(N, 3)) * 1e-4
    rho = np.random.default_rng(28).normal(size=n[0] * n[1] * n[2])
    L = sc.lib()
    ptr = sc.opalsc._ptr
    with ctx(n, N=N) as c:
        Ef_h, Bf_h, Q_h = c.computeSelfFields(R, P, Q, dt, 1e-11, org, hr, Bf=Bf0)
        phi_h = c.computePotential(rho, hr)
        H = _HipArrays()
        try:
            d = {k: H.put(v) for k, v in dict(R=R, P=P, Q=Q, dt=dt, Bf=Bf0, Ef=np.zeros((N, 3)),
                                                 rho=rho).items()}
            c._c(L.opalsc_compute_self_fields_dev(c._h, N, d["R"][0], d["P"][0], d["Q"][0], d["dt"][0], 1e-11,
                                                  ptr(org), ptr(hr), sc.COUPLING, d["Ef"][0], d["Bf"][0]))
            assert np.array_equal(H.get(d["Ef"]), Ef_h)
            assert np.array_equal(H.get(d["Bf"]), Bf_h)
            assert np.array_equal(H.get(d["Q"]), Q_h)
            c._c(L.opalsc_compute_potential_dev(c._h, d["rho"][0], ptr(hr)))
            assert np.array_equal(H.get(d["rho"]), phi_h)
            host = np.zeros(n[0] * n[1] * n[2])
            with pytest.raises(sc.OpalscError, match="device memory"):
                c._c(L.opalsc_compute_potential_dev(c._h, host.ctypes.data, ptr(hr)))
        finally:
            H.free()


def test_h5part_dump_and_restart(tmp_path):
    """A tracker dump (writeStep of the resident bunch) and a restart from it
    (doRestartOpalT over readStep): the restarted tracker continues as the
    original one does, and the step attributes are the tracker's statistics."""
    h5 = pytest.importorskip("src_amd.h5part")
    if not os.path.exists(h5.LIB_PATH):
        pytest.skip("libopalsc_h5.so not built (no HDF5)")
    n, N = (16, 16, 16), 5000
    R, P, Q = gaussian_bunch(N, seed=91, q_total=1e-9)
    path = str(tmp_path / "dump.h5")
    with ctx(n, N=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5)
        st = c.track(3)
        h5.create(path, dump_frequency=1)
        k = h5.write_tracker_step(path, c, local_track_step=3, global_track_step=3)
        tp = c.track_params()
        Rd, Pd, Qd = c.get_bunch()
        ref = c.track(3)
        Rr, Pr, _ = c.get_bunch()
    assert k == 0
    d = h5.read_step(path)
    a = d["attrs"]
    assert np.array_equal(d["R"], Rd) and np.array_equal(d["P"], Pd) and np.array_equal(d["Q"], Qd)
    assert a.time == tp.t and a.spos == tp.pathLength and list(a.ref_p) == list(tp.RefPartP)
    assert np.allclose(list(a.rms_x), st[-1]["rrms"], rtol=1e-12, atol=0)
    assert np.allclose(list(a.eps_norm), st[-1]["eps_norm"], rtol=1e-12, atol=0)
    assert list(a.max_x) == list(Rd.max(axis=0)) and list(a.min_p) == list(Pd.min(axis=0))
    assert a.charge == pytest.approx(1e-9, rel=1e-12) and a.local_track_step == 3
    with ctx(n, N=N) as c2:
        h5.restart(c2, path, 1e-11, Bz=0.5)
        got = c2.track(3)
        Rg, Pg, _ = c2.get_bunch()
    assert rel_l2(Rg, Rr) < 1e-13 and rel_l2(Pg - Pd, Pr - Pd) < 1e-10
    for x, y in zip(got, ref):
        assert x["t"] == y["t"] and x["pathLength"] == pytest.approx(y["pathLength"], rel=1e-14)
        assert np.allclose(x["eps_norm"], y["eps_norm"], rtol=1e-10, atol=0)

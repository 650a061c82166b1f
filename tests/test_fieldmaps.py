"""External field maps on the beamline: a Solenoid with a 2D magnetostatic map
(FM2DMagnetoStatic) and an RFCavity with a 2D dynamic map (FM2DDynamic),
evaluated in the tracker's kick (ParallelTTracker::computeExternalFields).

The oracle's reading of the map files and its evaluation are pinned by maps
whose fields are known in closed form (uniform, linear in r, one hot cell);
the library is compared with the oracle on the same files."""
import math

import numpy as np
import pytest

import oracle_py as O
from conftest import gaussian_bunch, rel_l2

sc = pytest.importorskip("src_amd")

MU_0 = 1.256637061e-06


def write_magnetostatic(path, zb, ze, nz, rb, re_, nr, bz, br, swap=False, comment=True):
    """bz, br: callables of (z index i, r index j) giving the file's values."""
    lines = ["# solenoid map written by the test"] if comment else []
    lines.append("2DMagnetoStatic " + ("ZX" if swap else "XZ"))
    if swap:
        lines += ["%r %r %d" % (rb, re_, nr), "%r %r %d  # z range" % (zb, ze, nz)]
        for i in range(nz + 1):
            for j in range(nr + 1):
                lines.append("%r %r" % (float(br(i, j)), float(bz(i, j))))
    else:
        lines += ["%r %r %d" % (zb, ze, nz), "%r %r %d" % (rb, re_, nr)]
        for j in range(nr + 1):
            for i in range(nz + 1):
                lines.append("%r %r" % (float(bz(i, j)), float(br(i, j))))
    path.write_text("\n".join(lines) + "\n")
    return path


def write_dynamic(path, zb, ze, nz, rb, re_, nr, freq_mhz, ez, er, ht, swap=False):
    lines = ["2DDynamic " + ("ZX" if swap else "XZ")]
    if swap:
        lines += ["%r %r %d" % (rb, re_, nr), "%r" % freq_mhz, "%r %r %d" % (zb, ze, nz)]
        for i in range(nz + 1):
            for j in range(nr + 1):
                lines.append("%r %r %r 0.0" % (float(er(i, j)), float(ez(i, j)), float(ht(i, j))))
    else:
        lines += ["%r %r %d" % (zb, ze, nz), "%r" % freq_mhz, "%r %r %d" % (rb, re_, nr)]
        for j in range(nr + 1):
            for i in range(nz + 1):
                lines.append("%r %r 0.0 %r" % (float(ez(i, j)), float(er(i, j)), float(ht(i, j))))
    path.write_text("\n".join(lines) + "\n")
    return path


def test_map_reading_both_orientations_agree_and_normalise(tmp_path):
    """XZ (z fastest, Bz Br) and ZX (r fastest, Br Bz) files of the same field
    read to the same grids; fields are divided by the largest on-axis |Bz|;
    lengths go from cm to m."""
    bz = lambda i, j: 2.0 + 0.1 * i - 0.05 * j
    br = lambda i, j: 0.01 * i * j
    a = O.parse_fieldmap_2d(write_magnetostatic(tmp_path / "xz.t7", 0.0, 20.0, 10, 0.0, 2.0, 4, bz, br))
    b = O.parse_fieldmap_2d(write_magnetostatic(tmp_path / "zx.t7", 0.0, 20.0, 10, 0.0, 2.0, 4, bz, br, swap=True))
    for k in ("a", "b"):
        assert np.array_equal(a[k], b[k])
    assert (a["nz"], a["nr"], a["hz"], a["hr"], a["length"]) == (11, 5, 0.02, 0.005, 0.2)
    vmax = max(abs(bz(i, 0)) for i in range(11))
    assert a["a"][3 + 2 * 11] == bz(3, 2) / vmax and a["b"][3 + 2 * 11] == br(3, 2) / vmax
    assert np.max(np.abs(a["a"][:11])) == 1.0


def test_uniform_solenoid_and_radial_field_closed_form(tmp_path):
    """Bz uniform and Br proportional to r: B = KS*(Br*x/r, Br*y/r, Bz) inside
    [edge, edge + length), nothing outside; bilinear interpolation is exact for
    fields linear in r and z."""
    m = O.parse_fieldmap_2d(write_magnetostatic(tmp_path / "s.t7", 0.0, 10.0, 5, 0.0, 1.0, 4,
                                                lambda i, j: 1.0, lambda i, j: 0.5 * j))
    maps = O.make_fieldmaps([(m, 0.3, 0.8, 0.0, 0.0)])
    R = np.array([[0.003, 0.004, 0.35], [0.0, 0.0, 0.31], [0.003, 0.004, 0.29], [0.003, 0.004, 0.4]])
    E, B, lost = O.fieldmap_eval(maps, 1, R, 0.0)
    assert lost == 0 and np.all(E == 0)
    r = 0.005
    br = 0.8 * 0.5 * (r / m["hr"])
    assert B[0] == pytest.approx([br * 0.6, br * 0.8, 0.8], rel=1e-14)
    assert np.array_equal(B[1], [0.0, 0.0, 0.8])  # on axis: no radial part
    assert np.all(B[2] == 0) and np.all(B[3] == 0)  # before the map, and at its end
    # beyond the map's radius: lost
    _, _, lost = O.fieldmap_eval(maps, 1, np.array([[0.02, 0.0, 0.35]]), 0.0)
    assert lost == 1


def test_rf_cavity_field_closed_form(tmp_path):
    """A dynamic map with uniform Ez, Er = 0 and H_theta linear in r:
    E = VOLT*1e6*cos(w t + lag) (0, 0, 1), B_theta = -VOLT*sin(w t + lag)*mu_0*H/Ezmax."""
    m = O.parse_fieldmap_2d(write_dynamic(tmp_path / "c.t7", 0.0, 10.0, 5, 0.0, 1.0, 4, 50.0,
                                          lambda i, j: 2.0, lambda i, j: 0.0, lambda i, j: 1000.0 * j))
    assert m["frequency"] == 50.0 * 2 * math.pi * 1e6
    maps = O.make_fieldmaps([(m, 0.0, 3.0, 50.1e6, 0.3)])  # within 1 %: the element's frequency
    t, x, y = 1.3e-9, 0.004, -0.003
    E, B, _ = O.fieldmap_eval(maps, 1, np.array([[x, y, 0.05]]), t)
    ph = 50.1e6 * 2 * math.pi * t + 0.3
    assert E[0] == pytest.approx([0.0, 0.0, 3.0 * math.cos(ph) * 1e6], rel=1e-14, abs=1e-9)
    bt = -3.0 * math.sin(ph) * MU_0 * 1000.0 * (0.005 / m["hr"]) / 2.0
    assert B[0] == pytest.approx([-bt * y / 0.005, bt * x / 0.005, 0.0], rel=1e-12, abs=1e-20)
    far = O.make_fieldmaps([(m, 0.0, 3.0, 80e6, 0.3)])  # more than 1 % away: the map's frequency
    assert far[0].omega == m["frequency"]


def test_map_files_the_reference_rejects(tmp_path):
    bad = tmp_path / "third.t7"
    bad.write_text("2DMagnetoStatic XZ TRUE\n0 10 1\n0 1 1\n1 0\n1 0\n1 0\n1 0\n")
    short = write_magnetostatic(tmp_path / "short.t7", 0.0, 10.0, 2, 0.0, 1.0, 1, lambda i, j: 1.0,
                                lambda i, j: 0.0)
    short.write_text("\n".join(short.read_text().split("\n")[:-3]) + "\n")
    extra = write_magnetostatic(tmp_path / "extra.t7", 0.0, 10.0, 2, 0.0, 1.0, 1, lambda i, j: 1.0,
                                lambda i, j: 0.0)
    extra.write_text(extra.read_text() + "1.0 2.0\n")
    for p in (bad, short, extra):
        with pytest.raises(ValueError):
            O.parse_fieldmap_2d(p)


@pytest.mark.gpu
def test_library_reads_and_evaluates_maps_like_the_oracle(tmp_path):
    rng = np.random.default_rng(2)
    g1 = rng.normal(size=(21, 9))
    g2 = rng.normal(size=(21, 9))
    g3 = rng.normal(size=(21, 9))
    f1 = write_magnetostatic(tmp_path / "sol.t7", -5.0, 15.0, 20, 0.0, 1.6, 8,
                             lambda i, j: 1.0 + 0.2 * g1[i, j], lambda i, j: 0.1 * g2[i, j], swap=True)
    f2 = write_dynamic(tmp_path / "cav.t7", 0.0, 12.0, 20, 0.0, 1.6, 8, 325.0,
                       lambda i, j: 5.0 + g1[i, j], lambda i, j: g2[i, j], lambda i, j: 100 * g3[i, j])
    specs = [(O.parse_fieldmap_2d(f1), 0.02, 0.7, 0.0, 0.0), (O.parse_fieldmap_2d(f2), 0.1, 12.0, 325.5e6, 1.1)]
    maps = O.make_fieldmaps(specs)
    R = np.column_stack([rng.uniform(-0.02, 0.02, 5000), rng.uniform(-0.02, 0.02, 5000),
                         rng.uniform(-0.05, 0.3, 5000)])
    with sc.SpaceChargeContext((8, 8, 8), max_particles=len(R)) as c:
        assert c.add_fieldmap(f1, 0.02, 0.7) == 0
        assert c.add_fieldmap(f2, 0.1, 12.0, 325.5e6, 1.1) == 1
        for t in (0.0, 2.7e-9):
            E, B, lost = c.fieldmap_eval(R, t)
            Eo, Bo, losto = O.fieldmap_eval(maps, 2, R, t)
            assert lost == losto and lost > 0
            assert rel_l2(E, Eo) < 1e-14 and rel_l2(B, Bo) < 1e-14
        with pytest.raises(sc.OpalscError, match="error occured"):
            bad = tmp_path / "bad.t7"
            bad.write_text("2DMagnetoStatic XZ FALSE\n")
            c.add_fieldmap(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol_r,tol_p", [(sc.FP64, 1e-12, 1e-9), (sc.MIXED, 1e-8, 1e-5)])
def test_tracker_with_solenoid_and_cavity_matches_oracle(tmp_path, precision, tol_r, tol_p):
    """The tracker's kick adds the maps' fields (bunch at t + dT/2, reference
    particle at t - dT/2, both in the step's lab frame); against the oracle's
    ParallelTTracker loop with the same elements."""
    n, N, steps = (16, 16, 16), 40000, 6
    f1 = write_magnetostatic(tmp_path / "sol.t7", 0.0, 2.0, 40, 0.0, 1.0, 20,
                             lambda i, j: 1.0 - 0.002 * (i - 20) ** 2, lambda i, j: 0.001 * (i - 20) * j)
    f2 = write_dynamic(tmp_path / "cav.t7", 0.0, 2.0, 40, 0.0, 1.0, 20, 50.0,
                       lambda i, j: 1.0 + 0.01 * i, lambda i, j: -0.002 * j, lambda i, j: 40.0 * j)
    specs = [(O.parse_fieldmap_2d(f1), -0.01, 0.5, 0.0, 0.0), (O.parse_fieldmap_2d(f2), -0.005, 2.0, 0.0, 0.4)]
    R, P, Q = gaussian_bunch(N, seed=23, q_total=1e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), maps=specs)
    Ro, Po, Qo, _, so = O.track(tr, R, P, Q, dt, steps)
    with sc.SpaceChargeContext(n, precision=precision, max_particles=N, sort_interval=3) as c:
        c.add_fieldmap(f1, -0.01, 0.5)
        c.add_fieldmap(f2, -0.005, 2.0, 0.0, 0.4)
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181))
        sg = c.track(steps)
        Rg, Pg, Qg = c.get_bunch()
        par = c.track_params()
    # the maps act: a run without them ends elsewhere
    tr0 = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181))
    R0, P0, _, _, _ = O.track(tr0, R, P, Q, dt, steps)
    assert rel_l2(Po - P, P0 - P) > 1e-3
    assert rel_l2(Rg, Ro) < tol_r
    assert rel_l2(Pg - P, Po - P) < tol_p
    assert np.allclose(list(par.RefPartP), list(tr.RefPartP), rtol=1e-13, atol=0)
    for a, b in zip(sg, so):
        for k in ("rrms", "eps_norm"):
            assert np.allclose(a[k], b[k], rtol=1e-9 if precision == sc.FP64 else 1e-5, atol=0), (k, a[k], b[k])

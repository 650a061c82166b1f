"""Golden fixtures (tests/golden/make_golden.py): the oracle must keep
reproducing them bit for bit; the GPU must match them at the parity bars."""
import os

import numpy as np
import pytest

import oracle_py as O
from conftest import rel_l2

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_oracle_reproduces_self_field_fixture():
    f = load("self_fields_1k_16.npz")
    out = O.compute_self_fields(f["R"], f["P"], f["Q"], f["dt"], 1e-11, f["origin"], f["hr"],
                                f["n"], want_grids=True)
    for k in ("rho", "phi", "E", "Ef", "Bf"):
        assert np.array_equal(out[k], f[k]), k
    assert np.array_equal(out["Q"], f["Qout"])


def test_oracle_reproduces_track_fixture():
    f = load("track_1k_16_3steps.npz")
    tr = O.make_tracker(tuple(f["n"]), float(f["dT"]), ref_p=(0.0, 0.0, float(f["p0"])),
                        Bz=float(f["Bz"]))
    R, P, _, _, st = O.track(tr, f["R"], f["P"], f["Q"], np.full(len(f["R"]), float(f["dT"])), 3)
    assert np.array_equal(R, f["Rout"]) and np.array_equal(P, f["Pout"])
    assert np.array_equal(np.array([s["eps_norm"] for s in st]), f["eps_norm"])


@pytest.mark.gpu
def test_gpu_matches_fixtures():
    import src_amd as sc
    f = load("self_fields_1k_16.npz")
    with sc.SpaceChargeContext(tuple(f["n"]), max_particles=len(f["R"])) as c:
        Ef, Bf, Q = c.computeSelfFields(f["R"], f["P"], f["Q"], f["dt"], 1e-11, f["origin"], f["hr"])
        assert rel_l2(c.get_grid(sc.GRID_PHI), f["phi"]) < 1e-12
        assert rel_l2(c.get_grid(sc.GRID_RHO), f["rho"]) < 1e-12
    assert rel_l2(Ef, f["Ef"]) < 1e-12 and rel_l2(Bf, f["Bf"]) < 1e-12
    assert np.array_equal(Q, f["Qout"])
    t = load("track_1k_16_3steps.npz")
    with sc.SpaceChargeContext(tuple(t["n"]), max_particles=len(t["R"]), sort_interval=1) as c:
        c.set_bunch(t["R"], t["P"], t["Q"])
        c.track_init(float(t["dT"]), ref_p=(0.0, 0.0, float(t["p0"])), Bz=float(t["Bz"]))
        st = c.track(3)
        R, P, _ = c.get_bunch()
    assert rel_l2(R, t["Rout"]) < 1e-13
    for i, s in enumerate(st):
        assert np.allclose(s["rrms"], t["rrms"][i], rtol=1e-12, atol=0)
        assert np.allclose(s["eps_norm"][:2], t["eps_norm"][i][:2], rtol=1e-10, atol=0)

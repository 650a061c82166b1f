"""Writes tests/golden/*.npz: small, seeded cases of the space-charge step
computed by the oracle (1e3 particles on a 16^3 grid, 3 tracker steps;
one self-field solve with its grids).  They pin the oracle against silent
regressions and let GPU tests compare with stored vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_py as O  # noqa: E402
from conftest import gaussian_bunch  # noqa: E402

N, n, steps, dT, p0 = 1000, (16, 16, 16), 3, 1e-11, 0.046181


def make():
    R, P, Q = gaussian_bunch(N, seed=101, q_total=1e-9, p0=p0)
    dt = np.full(N, dT)
    hr, origin = O.boundp(R, n)
    sf = O.compute_self_fields(R, P, Q, dt, dT, origin, hr, n, want_grids=True)
    np.savez_compressed(os.path.join(HERE, "self_fields_1k_16.npz"), R=R, P=P, Q=Q, dt=dt,
                        hr=hr, origin=origin, n=np.array(n), rho=sf["rho"], phi=sf["phi"],
                        E=sf["E"], Ef=sf["Ef"], Bf=sf["Bf"], Qout=sf["Q"])
    tr = O.make_tracker(n, dT, ref_p=(0.0, 0.0, p0), Bz=0.5)
    R2, P2, Q2, _, st = O.track(tr, R, P, Q, dt, steps)
    keys = ("rrms", "prms", "eps_norm", "rmean", "pmean")
    np.savez_compressed(os.path.join(HERE, "track_1k_16_3steps.npz"), R=R, P=P, Q=Q, Rout=R2,
                        Pout=P2, Bz=np.array(0.5), dT=np.array(dT), p0=np.array(p0),
                        n=np.array(n), **{k: np.array([s[k] for s in st]) for k in keys})


if __name__ == "__main__":
    make()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))

"""H5Part phase-space files (include/opalsc_h5.h): the layout of
H5PartWrapperForPT (src/Structure/H5PartWrapperForPT.cpp:193-424) as HDF5's
own h5dump reads it back, and write/read round trips (CPU only: the H5Part
library is host code).  The GPU restart test is in test_gpu_parity.py."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

H5 = pytest.importorskip("src_amd.h5part")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

if not os.path.exists(H5.LIB_PATH):
    pytest.skip("libopalsc_h5.so not built (no HDF5)", allow_module_level=True)


def _attrs(seed):
    rng = np.random.default_rng(seed)
    a = H5.H5Step()
    for name, _ in H5.H5Step._fields_:
        v = getattr(a, name)
        if isinstance(v, int):
            setattr(a, name, int(rng.integers(0, 1000)))
        elif isinstance(v, float):
            setattr(a, name, float(rng.normal()))
        else:
            v[:] = list(rng.normal(size=3))
    return a


def _bunch(N, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(N, 3)) * 1e-3, rng.normal(size=(N, 3)), rng.uniform(1e-15, 2e-15, N)


def test_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "opalsc_h5.h")).read()
    names = sorted(set(re.findall(r"\b(opalsc_h5_[a-z0-9_]+)\s*\(", text)))
    L = H5.lib()
    assert [n for n in names if not hasattr(L, n)] == []
    assert sorted(H5.EXPORTS) == names


def test_round_trip_of_two_steps(tmp_path):
    path = str(tmp_path / "bunch.h5")
    H5.create(path, "OPAL 2.0.0 test", dump_frequency=10, dphi_global=0.25)
    data = [(_bunch(1000, 1), _attrs(2)), (_bunch(777, 3), _attrs(4))]
    for k, ((R, P, Q), a) in enumerate(data):
        assert H5.write_step(path, R, P, Q, a, ids=np.arange(len(R)) + 5) == k
    assert H5.num_steps(path) == 2
    assert H5.num_particles(path, 0) == 1000 and H5.num_particles(path) == 777
    assert H5.read_header(path) == {"dump_frequency": 10, "dPhiGlobal": 0.25}
    for k, ((R, P, Q), a) in enumerate(data):
        d = H5.read_step(path, k)
        assert np.array_equal(d["R"], R) and np.array_equal(d["P"], P) and np.array_equal(d["Q"], Q)
        assert np.array_equal(d["id"], np.arange(len(R)) + 5)
        b = d["attrs"]
        for name, _ in H5.H5Step._fields_:
            x, y = getattr(a, name), getattr(b, name)
            if name == "step":
                assert y == k  # the writer numbers the steps
            elif isinstance(x, (int, float)):
                assert x == y, name
            else:
                assert list(x) == list(y), name
    # the default step is the last (readStep); a slice [first, last]
    d = H5.read_step(path, -1, first=10, last=19)
    assert np.array_equal(d["R"], data[1][0][0][10:20])
    with pytest.raises(H5.H5PartError, match="particle numbers"):
        H5.read_step(path, 1, first=0, last=777)
    with pytest.raises(H5.H5PartError):
        H5.read_step(path, 2)


def test_an_empty_bunch_writes_no_step(tmp_path):
    path = str(tmp_path / "empty.h5")
    H5.create(path)
    assert H5.write_step(path, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), _attrs(5)) == -1
    assert H5.num_steps(path) == 0


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_layout_as_hdf5_reads_it(tmp_path):
    """What H5PartWrapperForPT writes through H5hut, read by HDF5's h5dump
    (not by this library): root attributes, groups Step#0 and Step#1, the
    step attributes with their types and lengths, one 1D dataset per
    quantity (float64 x..q, int64 id, int32 ptype)."""
    path = str(tmp_path / "layout.h5")
    H5.create(path, "OPAL x", dump_frequency=3)
    for k in range(2):
        R, P, Q = _bunch(50, 10 + k)
        H5.write_step(path, R, P, Q, _attrs(20 + k))
    out = subprocess.run([H5DUMP, "-H", "-A", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    blocks = re.split(r'\n\s*(?=GROUP "|DATASET "|ATTRIBUTE ")', out)

    def block(kind, name):
        for b in blocks:
            if b.startswith('%s "%s"' % (kind, name)):
                return b
        raise AssertionError("%s %s missing" % (kind, name))

    assert 'GROUP "Step#0"' in out and 'GROUP "Step#1"' in out and 'GROUP "Step#2"' not in out
    # file attributes (writeHeader)
    b = block("ATTRIBUTE", "dump frequency")
    assert "H5T_STD_I64LE" in b and "SIMPLE { ( 1 ) / ( 1 ) }" in b
    assert "H5T_IEEE_F64LE" in block("ATTRIBUTE", "dPhiGlobal")
    for name, unit in (("xUnit", "m"), ("pxUnit", "#beta#gamma"), ("qUnit", "C"), ("ENERGYUnit", "MeV")):
        b = block("ATTRIBUTE", name)
        assert "STRSIZE %d;" % (len(unit) + 1) in b and "H5T_STR_NULLTERM" in b and "SCALAR" in b, name
    # step attributes (writeStepHeader): 3-vectors, scalars, int64 counters, the flavour string
    for name in ("RefPartR", "centroid", "RMSX", "RefPartP", "MEANP", "RMSP", "TaitBryantAngles", "#varepsilon",
                 "#varepsilon-geom", "minX", "maxX", "minP", "maxP", "B-ref", "E-ref"):
        b = block("ATTRIBUTE", name)
        assert "H5T_IEEE_F64LE" in b and "SIMPLE { ( 3 ) / ( 3 ) }" in b, name
    for name in ("SPOS", "#sigma", "TIME", "ENERGY", "dE", "MASS", "CHARGE"):
        b = block("ATTRIBUTE", name)
        assert "H5T_IEEE_F64LE" in b and "SIMPLE { ( 1 ) / ( 1 ) }" in b, name
    for name in ("Step", "LocalTrackStep", "GlobalTrackStep", "NumBunch", "SteptoLastInj"):
        assert "H5T_STD_I64LE" in block("ATTRIBUTE", name), name
    assert "STRSIZE 7;" in block("ATTRIBUTE", "OPAL_flavour")
    # particle datasets (writeStepData)
    for name in ("x", "y", "z", "px", "py", "pz", "q"):
        b = block("DATASET", name)
        assert "H5T_IEEE_F64LE" in b and "SIMPLE { ( 50 ) / ( 50 ) }" in b, name
    assert "H5T_STD_I64LE" in block("DATASET", "id")
    assert "H5T_STD_I32LE" in block("DATASET", "ptype")

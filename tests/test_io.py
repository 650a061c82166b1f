"""File formats beside the space-charge step: FROMFILE distributions
(Distribution::createDistributionFromFile, src/Distribution/Distribution.cpp:1073-1250)
and the DataSink statistics file (src/Structure/DataSink.cpp:281-430,585-925).

The readers and writers are host code in the library (no GPU needed); the
statistics line of a tracked bunch is computed on the device and checked
against the oracle in the gpu-marked test."""
import math

import numpy as np
import pytest

import oracle_py as O
from conftest import gaussian_bunch

sc = pytest.importorskip("src_amd")

# The reference column order (DataSink::writeSDDSHeader); numbering 1..48.
STAT_COLUMNS = (
    "t s numParticles charge energy rms_x rms_y rms_s rms_px rms_py rms_ps emit_x emit_y emit_s "
    "mean_x mean_y mean_s ref_x ref_y ref_z ref_px ref_py ref_pz max_x max_y max_s xpx ypy zpz "
    "Dx DDx Dy DDy Bx_ref By_ref Bz_ref Ex_ref Ey_ref Ez_ref dE dt partsOutside "
    "R0_x R0_y R0_s P0_x P0_y P0_s").split()


def _write(tmp_path, text, name="dist.dat"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_fromfile_reads_comments_count_and_free_layout(tmp_path):
    """Lines empty after '#' comments are removed are skipped before the count;
    the six values per particle are read as a whitespace stream (line breaks
    anywhere), in the order x px y py z pz."""
    text = ("# distribution written by hand\n"
            "\n"
            "   # indented comment\n"
            "3   # particles\n"
            "1.0 0.1 2.0 0.2 3.0 0.3\n"
            "-1e-3 5e-4\n -2e-3 6e-4 -3e-3\n 7e-4\n"
            "0 0 0 0 0 0.046181\n")
    R, P = sc.fromfile_read(_write(tmp_path, text))
    assert np.array_equal(R, [[1.0, 2.0, 3.0], [-1e-3, -2e-3, -3e-3], [0.0, 0.0, 0.0]])
    assert np.array_equal(P, [[0.1, 0.2, 0.3], [5e-4, 6e-4, 7e-4], [0.0, 0.0, 0.046181]])


def test_fromfile_momenta_in_ev():
    """converteVToBetaGamma: sqrt((|E|/m + 1)^2 - 1) with the sign of E; below
    machine epsilon the small-value form sqrt(2|E|/m)."""
    m = 938.27204e6
    vals = [2.5e5, -1.0e6, 1e-9, -3e-10, 0.0]
    def conv(v):
        b = math.copysign(math.sqrt((abs(v) / m + 1.0) ** 2 - 1.0), v)
        if abs(b) < np.finfo(float).eps:
            b = math.copysign(math.sqrt(2 * abs(v) / m), v)
        return b
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ev.dat")
        with open(p, "w") as f:
            f.write("1\n0 %r 0 %r 0 %r\n" % (vals[0], vals[1], vals[2]))
        R, P = sc.fromfile_read(p, momenta_in_ev=True, mass_eV=m)
        with open(p, "w") as f:
            f.write("1\n0 %r 0 %r 0 %r\n" % (vals[3], vals[4], 1.0))
        R2, P2 = sc.fromfile_read(p, momenta_in_ev=True, mass_eV=m)
    assert P[0, 0] == conv(vals[0]) and P[0, 1] == conv(vals[1]) and P[0, 2] == conv(vals[2])
    assert P2[0, 0] == conv(vals[3]) and P2[0, 1] == 0.0
    assert P[0, 2] == math.sqrt(2e-9 / m)  # the small-value branch


def test_fromfile_errors_and_short_files(tmp_path):
    with pytest.raises(sc.OpalscError, match="really exists"):
        sc.fromfile_read(tmp_path / "missing.dat")
    with pytest.raises(sc.OpalscError, match="does not seem"):
        sc.fromfile_read(_write(tmp_path, "# only\n0\n1 2 3 4 5 6\n", "zero.dat"))
    with pytest.raises(sc.OpalscError, match="does not seem"):
        sc.fromfile_read(_write(tmp_path, "   \n2\n", "blank.dat"))  # a blank line is not skipped
    # as in the reference, a value the stream cannot read leaves zeros behind it
    R, P = sc.fromfile_read(_write(tmp_path, "2\n1 2 3 4 5 6\n7 8 9\n", "short.dat"))
    assert np.array_equal(R, [[1, 3, 5], [7, 9, 0]]) and np.array_equal(P, [[2, 4, 6], [8, 0, 0]])


def _row(seed=0):
    rng = np.random.default_rng(seed)
    row = sc.StatRow()
    for name, _ in sc.StatRow._fields_:
        v = getattr(row, name)
        if hasattr(v, "__len__"):
            v[:] = list(rng.normal(size=len(v)) * 10.0 ** rng.integers(-12, 6, size=len(v)))
        elif name in ("num_particles", "parts_outside"):
            setattr(row, name, int(rng.integers(0, 10 ** 9)))
        elif name == "has_particle0":
            row.has_particle0 = 1
        else:
            setattr(row, name, float(rng.normal() * 10.0 ** rng.integers(-12, 6)))
    return row


def _row_values(row):
    out = [row.t_ns, row.s, row.num_particles, row.charge, row.energy_MeV]
    for k in ("rms_r", "rms_p", "emit", "mean_r", "ref_r", "ref_p", "max_r", "corr"):
        out += list(getattr(row, k))
    out += [row.Dx, row.DDx, row.Dy, row.DDy] + list(row.B_ref) + list(row.E_ref)
    out += [row.dE_MeV, row.dt_ns, row.parts_outside] + list(row.R0) + list(row.P0)
    return out


def _read_stat(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "SDDS1"
    cols = [l.split("=", 1)[1].rstrip(",") for l in lines if l.strip().startswith("name=")]
    i = lines.index("&data")
    assert lines[i + 1].strip() == "mode=ascii," and lines[i + 2].strip() == "no_row_counts=1"
    params = lines[i + 4:i + 7]
    rows = [l for l in lines[i + 7:] if l]
    return cols, params, rows


def test_stat_file_header_and_rows(tmp_path):
    """SDDS header with the reference's parameters and 48 columns in its
    order; each value followed by a tab padded to width 10 (9 spaces and a
    tab); doubles in scientific notation with 15 digits, so the row reads back
    to 1e-15 relative."""
    p = tmp_path / "run.stat"
    rows = [_row(1), _row(2)]
    sc.stat_write(p, rows[0], input_fn="run.in", new_file=True)
    sc.stat_write(p, rows[1])
    cols, params, lines = _read_stat(p)
    assert cols[:3] == ["processors", "revision", "flavor"]
    assert cols[3:] == STAT_COLUMNS
    assert params[0] == "1" and params[2] == "opal-t"
    assert "Statistics data 'run.in'" in open(p).read()
    assert len(lines) == 2
    for line, row in zip(lines, rows):
        fields = line.split("\t")
        assert fields[-1] == ""
        assert all(f.endswith(" " * 9) for f in fields[:-1])
        got = [float(f) for f in fields[:-1]]
        want = _row_values(row)
        assert len(got) == len(want) == 48
        for g, w, name in zip(got, want, STAT_COLUMNS):
            assert g == pytest.approx(w, rel=1e-14, abs=0), name
        mant = fields[0].strip().lstrip("-").split("e")[0]
        assert len(mant) == 17  # d.ddddddddddddddd


def test_stat_file_without_particle0_has_42_columns(tmp_path):
    p = tmp_path / "np.stat"
    row = _row(3)
    row.has_particle0 = 0
    sc.stat_write(p, row, new_file=True)
    cols, _, lines = _read_stat(p)
    assert cols[3:] == STAT_COLUMNS[:42]
    assert len(lines[0].split("\t")) == 43


@pytest.mark.gpu
def test_stat_row_of_a_tracked_bunch_matches_oracle():
    """calcBeamParameters with the dispersion moments, energy spread, maximum,
    charge sum and particle 0 of a tracked bunch (FP64) against the oracle run
    of the same steps."""
    n, N, steps = (16, 16, 16), 30000, 3
    R, P, Q = gaussian_bunch(N, seed=17, q_total=1e-9)
    dt = np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, Ez0=1e6, rf_freq=5e7)
    Ro, Po, Qo, _, so = O.track(tr, R, P, Q, dt, steps)
    with sc.SpaceChargeContext(n, max_particles=N) as c:
        c.set_bunch(R, P, Q)
        c.track_init(1e-11, ref_p=(0.0, 0.0, 0.046181), Bz=0.5, Ez0=1e6, rf_freq=5e7)
        c.track(steps)
        row = c.stat_row()
    st = O.beam_parameters(Ro, Po)
    m = O.calc_moments(Ro, Po)
    Nd = float(N)
    assert row.num_particles == N
    assert row.t_ns == pytest.approx(tr.t * 1e9, rel=1e-15)
    assert row.s == pytest.approx(tr.pathLength, rel=1e-12)
    assert row.charge == pytest.approx(Qo.sum(), rel=1e-13)
    assert row.energy_MeV == pytest.approx(st["ekin_MeV"], rel=1e-10)
    for k, name in (("rms_r", "rrms"), ("rms_p", "prms"), ("emit", "eps_norm"), ("mean_r", "rmean"),
                    ("corr", "rprms")):
        assert np.allclose(list(getattr(row, k)), st[name], rtol=1e-9, atol=0), k
    assert np.array_equal(list(row.max_r), Ro.max(axis=0))
    assert row.Dx == pytest.approx(m[21] / Nd, rel=1e-9)
    assert row.DDx == pytest.approx(m[22] / Nd, rel=1e-9)
    assert row.Dy == pytest.approx(m[23] / Nd, rel=1e-9)
    assert row.DDy == pytest.approx(m[24] / Nd, rel=1e-9)
    m0 = O.M_P_EV * 1e-6
    dE = st["prms"][2] * m0 * math.sqrt(1.0 - 1.0 / (st["ekin_MeV"] / m0 + 1.0) ** 2)
    assert row.dE_MeV == pytest.approx(dE, rel=1e-9)
    assert np.array_equal(list(row.R0), Ro[0]) and np.array_equal(list(row.P0), Po[0])
    assert list(row.B_ref) == [0.0, 0.0, 0.5]
    ez = 1e6 * math.sin(2 * math.pi * 5e7 * (tr.t - 0.5e-11)) * 1e-6
    assert row.E_ref[2] == pytest.approx(ez, rel=1e-12)
    assert row.dt_ns == pytest.approx(1e-2, rel=1e-15)

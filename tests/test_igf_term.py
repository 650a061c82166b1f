"""The integrated Green's function's corner term in its two forms.

The term (src/Solvers/FFTPoissonSolver.cpp:479-489)
    F = -z^2/2 atan(xy/(zr)) - y^2/2 atan(xz/(yr)) - x^2/2 atan(yz/(xr))
        + yz log(x+r) + xz log(y+r) + xy log(z+r)
is evaluated by the FP64 solve with correctly rounded atan/log in the
reference's operation order, and by the mixed-precision solve in a
reduced-instruction form (own series for atan and log, the third atan from the
identity atan(xy/(zr)) + atan(xz/(yr)) + atan(yz/(xr)) = sgn(xyz) pi/2, one
shared division).  Both are checked against the same formula in 80-bit long
double at the corner lattices the solver evaluates (v = (i - 1/2) h, plus the
image charge's shifted z), with the error measured against r^2, the size of the
terms whose rounding the 8-corner difference keeps.
"""
import numpy as np
import pytest

from src_amd import opalsc as sc

LD = np.longdouble


def _F_longdouble(v):
    x, y, z = (v[:, k].astype(LD) for k in range(3))
    r = np.sqrt(x * x + y * y + z * z)
    with np.errstate(divide="ignore"):
        t = (-z * z * np.arctan(x * y / (z * r)) / 2 - y * y * np.arctan(x * z / (y * r)) / 2
             - x * x * np.arctan(y * z / (x * r)) / 2
             + y * z * np.log(x + r) + x * z * np.log(y + r) + x * y * np.log(z + r))
    return t


def _corners(h, n, zshift, count, seed):
    m = n + 2
    rng = np.random.default_rng(seed)
    q = rng.choice(m ** 3, size=min(count, m ** 3), replace=False)
    i, j, k = q % m, (q // m) % m, q // (m * m)
    return np.stack([i * h[0] - h[0] / 2, j * h[1] - h[1] / 2, k * h[2] - h[2] / 2 + zshift], 1)


CASES = [
    ((1e-5, 2e-5, 3e-5), 64, 0.0),
    ((1.3e-4, 1.3e-4, 7e-4), 128, 0.0),   # z stretched by a Lorentz factor
    ((2e-6, 2e-6, 2e-6), 256, 0.0),
    ((1e-5, 1e-5, 2e-5), 64, 3.3e-4),     # image charge: z shifted into the lattice
    ((1e-5, 1e-5, 2e-5), 64, -2.1e-3),    # image charge: all z negative
]


@pytest.mark.parametrize("h,n,zshift", CASES)
@pytest.mark.parametrize("fast", [False, True])
def test_corner_term_matches_long_double(h, n, zshift, fast):
    v = _corners(h, n, zshift, 20000, 7)
    ref = _F_longdouble(v)
    f = sc.host_igf_term(v, fast)
    err = np.abs(f.astype(LD) - ref) / (v ** 2).sum(1)
    # the rounding of six O(r^2) products summed in double: a few 1e-16 r^2
    assert float(err.max()) < 4e-15


def test_fast_term_at_signed_zero_z_takes_the_quotients_limits():
    # at z = +-0, atan(xy/(zr)) is +-pi/2 with the quotient's sign and the other
    # two atans are +-0: the identity-based third atan must give the same value
    xy = np.array([[1e-5, 2e-5], [-1e-5, 2e-5], [1e-5, -3e-5], [-2e-5, -1e-5]])
    for z in (0.0, -0.0):
        v = np.column_stack([xy, np.full(len(xy), z)])
        cr = sc.host_igf_term(v, False)
        fast = sc.host_igf_term(v, True)
        assert np.all(np.isfinite(fast))
        np.testing.assert_allclose(fast, cr, rtol=1e-14, atol=1e-15 * float((v ** 2).sum(1).max()))


def test_fast_term_g_differences_match_correctly_rounded_form():
    # the 8-corner difference of both forms on a whole lattice: far-field G is
    # ~(r/h)^3 smaller than F, so this is where an error in F would show
    h, n = (1e-5, 1.5e-5, 4e-5), 24
    m = n + 2
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    v = np.stack([(i * h[0] - h[0] / 2).ravel(), (j * h[1] - h[1] / 2).ravel(),
                  (k * h[2] - h[2] / 2).ravel()], 1)
    vol = h[0] * h[1] * h[2]

    def g_of(T):
        T = T.reshape(m, m, m)
        return (T[1:, 1:, 1:] + T[1:, :-1, :-1] + T[:-1, 1:, :-1] + T[:-1, :-1, 1:]
                - T[1:, 1:, :-1] - T[1:, :-1, 1:] - T[:-1, 1:, 1:] - T[:-1, :-1, :-1])

    g_ld = g_of(_F_longdouble(v) / vol)
    for fast in (False, True):
        g = g_of(sc.host_igf_term(v, fast) / vol)
        rel = np.abs(g.astype(LD) - g_ld) / np.abs(g_ld)
        assert float(rel.max()) < 1e-9, fast


@pytest.mark.gpu
@pytest.mark.parametrize("fast", [False, True])
def test_device_corner_term_equals_host_bits(fast):
    vs = [_corners(h, n, zs, 4000, 3) for (h, n, zs) in CASES]
    v = np.concatenate(vs + [np.array([[1e-5, 2e-5, 0.0], [1e-5, -2e-5, -0.0]])])
    with sc.SpaceChargeContext((8, 8, 8), max_particles=16) as c:
        np.testing.assert_array_equal(c.device_igf_term(v, fast), sc.host_igf_term(v, fast))


def _G_longdouble(n, h, zshift, table):
    """The octant from the 8-corner difference of the corner term in 80-bit
    long double (the cell average of 1/r to ~1e-19 (r/h)^3)."""
    hx, hy, hz = (LD(v) for v in h)
    k, j, i = np.meshgrid(np.arange(n[2] + 1), np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
    g = np.zeros(i.shape, LD)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                z = (k + dz).astype(LD) * hz - hz / 2
                if table:
                    z = z + LD(zshift)
                    if table == 2:
                        z = z - LD(n[2]) * hz
                v = np.stack([((i + dx).astype(LD) * hx - hx / 2).ravel(), ((j + dy).astype(LD) * hy - hy / 2).ravel(),
                              z.ravel()], 1)
                # + for the corners with an odd count of upper indices (+++, +00, 0+0, 00+)
                g += (-1) ** (3 - dx - dy - dz) * _F_longdouble(v).reshape(i.shape)
    g /= hx * hy * hz
    if table == 0:
        g[0, 0, 0] = g[1, 0, 0]
    return g


OCTANTS = [
    ((40, 36, 44), (1e-5, 1.2e-5, 1.3e-5), 0.0, 0),
    ((40, 40, 40), (1e-5, 1e-5, 3e-5), 0.0, 0),        # z cells 3x longer
    ((32, 32, 48), (1e-5, 1e-5, 2e-5), -20.3 * 2e-5, 1),  # image table with the singularity inside
    ((32, 32, 48), (1e-5, 1e-5, 2e-5), -20.3 * 2e-5, 2),
]


@pytest.mark.parametrize("n,h,zshift,table", OCTANTS)
@pytest.mark.parametrize("fast", [False, True])
def test_green_octant_matches_long_double(n, h, zshift, table, fast):
    """Both forms of the octant against the long-double difference at every
    point.  The FP64 form (the reference's FP64 corner table) keeps the
    rounding its 8-corner difference cancels up, ~2e-15 (r/h)^3 (measured up to
    3.8e-9 on these grids); the mixed form (that difference near the origin,
    the order-4 cell average of 1/r from 16 cells on) stays below 3.1e-10,
    largest just past the 16-cell boundary of the 1:1:3 grid."""
    ref = _G_longdouble(n, h, zshift, table)
    g = sc.host_green_octant(n, h, zshift, table, fast)
    rel = np.abs(g.astype(LD) - ref) / np.abs(ref)
    assert float(rel.max()) < (5e-10 if fast else 1e-8)


def test_green_octant_forms_agree_on_a_large_grid():
    n, h = (64, 64, 64), (1e-5, 1e-5, 1.27e-5)
    a = sc.host_green_octant(n, h, fast=False)
    b = sc.host_green_octant(n, h, fast=True)
    assert np.abs(b - a).max() / np.abs(a).max() < 1e-10
    assert b[0, 0, 0] == b[1, 0, 0]  # G[0,0,0] := G[0,0,1]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [sc.FP64, sc.MIXED])
@pytest.mark.parametrize("table", [0, 1, 2])
def test_device_green_octant_equals_host_bits(precision, table):
    n, h, zshift = (24, 16, 32), (1e-5, 1.1e-5, 2e-5), -12.7 * 2e-5
    with sc.SpaceChargeContext(n, precision=precision, max_particles=16) as c:
        d = c.device_green_octant(h, zshift, table)
    np.testing.assert_array_equal(d, sc.host_green_octant(n, h, zshift, table, precision == sc.MIXED))

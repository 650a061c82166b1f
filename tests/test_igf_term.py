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

"""Pins of the CPU oracle against mathematics that does not depend on it.

The reference ships no golden vectors for the space-charge path (its tests
never run the FFT solver), so every oracle function is checked here against
an independent fact: closed forms, brute force, numerical quadrature,
numpy's pocketfft, mpmath, or a conservation law.  Each check is chosen so a
plausible slip (a dropped IGF term, a wrong sign, a transposed index, an
off-by-one cell) fails it.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from conftest import gaussian_bunch, rel_l2

C = O.CLIGHT


# ---------------------------------------------------------------- bounds
def test_bounds_hand_case_and_padding():
    R = np.array([[0.0, 1.0, -2.0], [3.0, -1.0, 5.0], [1.0, 0.5, 0.0]])
    lo, hi = O.get_bounds(R)
    assert lo.tolist() == [0.0, -1.0, -2.0] and hi.tolist() == [3.0, 1.0, 5.0]
    n = (8, 16, 32)
    hr, origin = O.boundp(R, n, dh=0.02)
    for d in range(3):
        L = hi[d] - lo[d]
        mn, mx = lo[d] - 0.02 * L, hi[d] + 0.02 * L
        assert hr[d] == (mx - mn) / (n[d] - 1)
        assert origin[d] == mn - hr[d] / 2.0


def test_boundp_degenerate_extent_branch():
    R = np.zeros((5, 3))
    R[:, 0] = np.linspace(-1, 1, 5)
    hr, origin = O.boundp(R, (8, 8, 8))
    assert hr[1] == (1e-10 - (-1e-10)) / 7
    assert origin[2] == -1e-10 - hr[2] / 2


# ---------------------------------------------------------------- CIC
def test_cic_particle_on_cell_centre_deposits_to_one_cell():
    n, hr, origin = (8, 8, 8), np.array([0.1, 0.2, 0.3]), np.array([-0.4, -0.8, -1.2])
    i, j, k = 3, 4, 5
    x = np.array([[(i + 0.5) * hr[0] + origin[0], (j + 0.5) * hr[1] + origin[1],
                   (k + 0.5) * hr[2] + origin[2]]])
    lgp, dpos, valid = O.cic_index(x, origin, hr)
    assert lgp.tolist() == [[i, j, k]] and dpos.tolist() == [[0.0, 0.0, 0.0]]
    rho, dropped = O.scatter(x, [2.5], origin, hr, n)
    assert dropped == 0
    assert rho[(k * n[1] + j) * n[0] + i] == 2.5 and np.count_nonzero(rho) == 1


def test_cic_hand_two_particles():
    n, hr, origin = (4, 4, 4), np.ones(3), np.zeros(3)
    # cell centres at 0.5,1.5,...: x=1.25 -> lgp 0, d=0.75 ; x=2.0 -> lgp 1, d=0.5
    R = np.array([[1.25, 0.5, 0.5], [2.0, 2.0, 0.5]])
    lgp, dpos, _ = O.cic_index(R, origin, hr)
    assert lgp.tolist() == [[0, 0, 0], [1, 1, 0]]
    assert dpos.tolist() == [[0.75, 0.0, 0.0], [0.5, 0.5, 0.0]]
    rho, _ = O.scatter(R, [1.0, 2.0], origin, hr, n)
    r = rho.reshape(4, 4, 4)
    assert r[0, 0, 0] == 0.25 and r[0, 0, 1] == 0.75
    assert r[0, 1, 1] == r[0, 1, 2] == r[0, 2, 1] == r[0, 2, 2] == 0.5


def test_cic_conservation_and_partition_of_unity():
    R, P, Q = gaussian_bunch(20000, seed=3)
    n = (16, 16, 16)
    hr, origin = O.boundp(R, n)
    rng = np.random.default_rng(1)
    q = rng.uniform(0.5, 1.5, len(R))
    rho, dropped = O.scatter(R, q, origin, hr, n)
    assert dropped == 0
    assert abs(rho.sum() - math.fsum(q)) < 1e-12 * q.sum()
    lgp, dpos, valid = O.cic_index(R, origin, hr)
    assert valid.all() and (dpos >= 0).all() and (dpos <= 1).all()
    assert (lgp >= 0).all() and (lgp <= np.array(n) - 2).all()


def test_cic_negative_side_uses_vertex_below():
    # getVertexBelow: (int)((x-o)/h), minus one when x < origin
    lgp, dpos, _ = O.cic_index(np.array([[-0.3, -1.0, 0.2]]), np.zeros(3), np.ones(3))
    assert lgp.tolist() == [[-1, -2, -1]]
    rho, dropped = O.scatter(np.array([[-0.3, 0.5, 0.5]]), [1.0], np.zeros(3), np.ones(3), (4, 4, 4))
    assert dropped == 4


def test_gather_linear_field_exact():
    n, hr, origin = (10, 12, 14), np.array([0.1, 0.2, 0.15]), np.array([-0.5, -1.2, -1.05])
    a = np.array([1.5, -0.7, 2.25])
    ii, jj, kk = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    pos = [(g + 0.5) * hr[d] + origin[d] for d, g in enumerate((ii, jj, kk))]
    f = a[0] * pos[0] + a[1] * pos[1] + a[2] * pos[2] + 0.3
    E = np.zeros((n[2], n[1], n[0], 3))
    for d in range(3):
        E[..., d] = (f * (d + 1)).transpose(2, 1, 0)
    rng = np.random.default_rng(2)
    lo = origin + hr * 0.6
    hi = origin + hr * (np.array(n) - 0.6)
    R = lo + rng.random((500, 3)) * (hi - lo)
    Ef = O.gather(R, E.reshape(-1, 3), origin, hr, n)
    expect = (R @ a + 0.3)[:, None] * np.array([1, 2, 3])
    assert np.max(np.abs(Ef - expect)) < 1e-12 * np.max(np.abs(expect))


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(6, 8, 12), (4, 4, 4), (3, 10, 16), (24, 6, 8)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(4)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    assert np.max(np.abs(O.fft3d(a, -1) - np.fft.fftn(a))) < 1e-12 * np.abs(a).sum()
    assert np.max(np.abs(O.fft3d(a, +1) - np.fft.ifftn(a) * a.size)) < 1e-12 * np.abs(a).sum()


# ---------------------------------------------------------------- Green's functions
def _gl_cube_integral(c, h, m=24):
    """Gauss-Legendre integral of 1/r over the cell centred at c (not containing 0)."""
    x, w = np.polynomial.legendre.leggauss(m)
    pts = [c[d] + 0.5 * h[d] * x for d in range(3)]
    X, Y, Z = np.meshgrid(*pts, indexing="ij")
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    return (W / np.sqrt(X * X + Y * Y + Z * Z)).sum() * (0.5 ** 3) * h[0] * h[1] * h[2]


def test_igf_self_cell_closed_form():
    # average of 1/r over the unit cube centred at 0 = 3 ln(2+sqrt3) - pi/2 (before the quirk)
    G = O.igf_octant((4, 4, 4), np.ones(3), quirk=False)
    exact = 3 * math.log(2 + math.sqrt(3)) - math.pi / 2
    assert abs(G[0, 0, 0] - exact) < 1e-14
    # pyramid reduction: int_[0,1]^3 1/r = 3/2 int_[0,1]^2 (1+u^2+v^2)^-1/2 du dv
    from scipy import integrate
    I2, _ = integrate.dblquad(lambda v, u: 1 / math.sqrt(1 + u * u + v * v), 0, 1, 0, 1,
                              epsabs=1e-14, epsrel=1e-14)
    assert abs(2 * 1.5 * I2 - exact) < 1e-12


@pytest.mark.parametrize("cell", [(1, 0, 0), (1, 1, 1), (2, 1, 0), (0, 3, 2)])
def test_igf_matches_quadrature(cell):
    hr = np.array([1.0e-4, 0.7e-4, 1.9e-4])
    n = (6, 6, 6)
    G = O.igf_octant(n, hr, quirk=False)
    c = np.array(cell) * hr
    ref = _gl_cube_integral(c, hr, m=40) / (hr[0] * hr[1] * hr[2])
    assert abs(G[cell[2], cell[1], cell[0]] - ref) < 1e-10 * ref


def test_igf_far_field_is_one_over_r_and_quirk():
    n, hr = (24, 24, 24), np.array([1.0, 1.0, 1.0])
    G = O.igf_octant(n, hr)
    assert G[0, 0, 0] == G[1, 0, 0]  # G[x=0,y=0,z=0] := G[x=0,y=0,z=1]
    for (i, j, k) in [(10, 4, 7), (20, 0, 3), (5, 15, 12)]:
        r = math.sqrt(i * i + j * j + k * k)
        # cube average of the harmonic 1/r differs from 1/r at O(h^4/r^5)
        assert abs(G[k, j, i] * r - 1) < 2.0 / r ** 4
    # homogeneity: scaling h by s scales the cell average of 1/r by 1/s
    G2 = O.igf_octant((6, 6, 6), np.array([2.0, 2.0, 2.0]), quirk=False)
    G1 = O.igf_octant((6, 6, 6), np.array([1.0, 1.0, 1.0]), quirk=False)
    assert np.allclose(G2, G1 / 2, rtol=1e-9, atol=0)


def test_igf_quad_vs_glibc_far_field_sensitivity():
    """glibc's atan/log are not correctly rounded for a few IGF arguments; the
    8-corner difference turns those ulps into ~1e-9 of far-field G.  The
    oracle uses correctly rounded values; the two agree to that level."""
    L = O.lib()
    L.orc_set_igf_libm.argtypes = [O.C.c_int]
    n, hr = (48, 48, 48), np.array([1.0e-4, 1.2e-4, 2.5e-4])
    Gcr = O.igf_octant(n, hr)
    L.orc_set_igf_libm(0)
    try:
        Gg = O.igf_octant(n, hr)
    finally:
        L.orc_set_igf_libm(1)
    d = np.abs(Gcr - Gg) / np.abs(Gcr)
    assert d.max() < 1e-8
    assert (d > 0).mean() < 0.2


def test_standard_green_function():
    n, hr = (5, 6, 7), np.array([0.3, 0.5, 0.7])
    G = O.green_standard_octant(n, hr)
    assert abs(G[2, 3, 4] * math.sqrt(16 * 0.09 + 9 * 0.25 + 4 * 0.49) - 1) < 1e-15
    assert G[0, 0, 0] == G[1, 0, 0]
    assert abs(G[1, 0, 0] * 0.7 - 1) < 1e-15


def test_mirror_is_even_extension():
    n = (3, 4, 5)
    Goct = np.arange(4 * 5 * 6, dtype=float).reshape(6, 5, 4)
    G2 = O.mirror_green(n, Goct)
    for (z, y, x) in [(7, 1, 5), (9, 7, 3), (0, 4, 4), (5, 5, 3)]:
        f = lambda i, m: i if i <= m else 2 * m - i  # noqa: E731
        assert G2[z, y, x] == Goct[f(z, 5), f(y, 4), f(x, 3)]


# ---------------------------------------------------------------- Poisson solve
def test_convolution_equals_brute_force():
    rng = np.random.default_rng(5)
    n = (5, 4, 6)
    rho = rng.normal(size=np.prod(n))
    G2 = O.mirror_green(n, O.igf_octant(n, np.array([1.0, 0.8, 1.3])))
    phi = O.convolve_doubled(n, rho, G2)
    r3 = rho.reshape(n[2], n[1], n[0])
    M = G2.shape
    bf = np.zeros_like(r3)
    zz, yy, xx = np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                bf[z, y, x] = (G2[(z - zz) % M[0], (y - yy) % M[1], (x - xx) % M[2]] * r3).sum()
    assert rel_l2(phi, bf) < 1e-14


def test_point_charge_potential_is_green_function():
    n, hr = (8, 8, 8), np.array([1.0, 1.0, 1.0])
    rho = np.zeros(512)
    c = (2, 5, 3)
    rho[(c[2] * 8 + c[1]) * 8 + c[0]] = 1.0
    phi = O.compute_potential(n, rho, hr).reshape(8, 8, 8)
    G = O.igf_octant(n, hr)
    for (x, y, z) in [(2, 5, 3), (7, 0, 0), (0, 2, 6)]:
        g = G[abs(z - c[2]), abs(y - c[1]), abs(x - c[0])]
        assert abs(phi[z, y, x] - g) < 1e-14 * G.max()


# ------------------------------------------------------- periodic-z (DC beam) solve
def test_periodic_z_green_layout():
    """FFTPoissonSolver with bcz_m: the grid is doubled in x and y only and the
    octant is mirrored in x and y only (FFTPoissonSolver.cpp:130-155,661-662),
    so z reads G one-sided for k in [0, n2)."""
    n, hr = (3, 4, 5), np.array([1.0, 0.8, 1.3])
    Goct = O.igf_octant(n, hr)
    G2 = O.mirror_green_periodic_z(n, Goct)
    assert G2.shape == (5, 8, 6)
    f = lambda i, m: i if i <= m else 2 * m - i  # noqa: E731
    for z in range(5):
        for y in range(8):
            for x in range(6):
                assert G2[z, y, x] == Goct[z, f(y, 4), f(x, 3)]


def test_periodic_z_solve_equals_brute_force():
    """Circular convolution on (2n0, 2n1, n2), Green's function built here from
    the octant, summed directly."""
    rng = np.random.default_rng(7)
    n, hr = (4, 5, 6), np.array([1.0, 0.7, 1.4])
    rho = rng.normal(size=np.prod(n))
    Goct = O.igf_octant(n, hr)
    f = lambda i, m: i if i <= m else 2 * m - i  # noqa: E731
    M = (n[2], 2 * n[1], 2 * n[0])
    G = np.zeros(M)
    for z in range(M[0]):
        for y in range(M[1]):
            for x in range(M[2]):
                G[z, y, x] = Goct[z, f(y, n[1]), f(x, n[0])]
    r3 = rho.reshape(n[2], n[1], n[0])
    zz, yy, xx = np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")
    bf = np.zeros_like(r3)
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                bf[z, y, x] = (G[(z - zz) % M[0], (y - yy) % M[1], (x - xx) % M[2]] * r3).sum()
    phi = O.compute_potential_periodic_z(n, rho, hr)
    assert rel_l2(phi, bf) < 1e-14


@pytest.mark.parametrize("greens", [0, 1])
def test_periodic_z_point_charge(greens):
    """A point charge at c: phi(x, y, z) = G(|x-cx|, |y-cy|, (z-cz) mod n2) --
    z wraps with the one-sided table; x and y see the free-space value."""
    n, hr = (6, 6, 8), np.array([1.0, 1.2, 0.9])
    rho = np.zeros(np.prod(n))
    c = (2, 4, 6)
    rho[(c[2] * n[1] + c[1]) * n[0] + c[0]] = 1.0
    phi = O.compute_potential_periodic_z(n, rho, hr, greens).reshape(n[2], n[1], n[0])
    G = O.igf_octant(n, hr) if greens == 0 else O.green_standard_octant(n, hr)
    for (x, y, z) in [(2, 4, 6), (5, 0, 0), (0, 2, 7), (3, 3, 1)]:
        g = G[(z - c[2]) % n[2], abs(y - c[1]), abs(x - c[0])]
        assert abs(phi[z, y, x] - g) < 1e-13 * np.abs(G).max()


def test_compute_potential_matches_numpy_fft():
    rng = np.random.default_rng(6)
    n, hr = (12, 8, 10), np.array([1e-4, 2e-4, 1.5e-4])
    rho = rng.normal(size=np.prod(n))
    phi = O.compute_potential(n, rho, hr)
    G2 = O.mirror_green(n, O.igf_octant(n, hr))
    a = np.zeros(G2.shape)
    a[:n[2], :n[1], :n[0]] = rho.reshape(n[2], n[1], n[0])
    ref = np.real(np.fft.ifftn(np.fft.fftn(a) * np.fft.fftn(G2)))[:n[2], :n[1], :n[0]]
    assert rel_l2(phi, ref) < 1e-13


# ---------------------------------------------------------------- gradient
def test_grad_exact_on_quadratics_including_borders():
    n, hr = (7, 6, 9), np.array([0.5, 0.3, 0.2])
    z, y, x = np.meshgrid(*[np.arange(n[d]) * hr[d] for d in (2, 1, 0)], indexing="ij")
    phi = 1.3 * x * x - 0.4 * x * y + 2.0 * z * z + 0.7 * y - 1.1 * z * x + 5.0
    g = O.grad(n, phi.ravel(), hr).reshape(n[2], n[1], n[0], 3)
    ex = 2.6 * x - 0.4 * y - 1.1 * z
    ey = -0.4 * x + 0.7 + 0 * y
    ez = 4.0 * z - 1.1 * x
    for d, e in enumerate((ex, ey, ez)):
        assert np.max(np.abs(g[..., d] - e)) < 1e-11


# ---------------------------------------------------------------- self fields
def test_self_field_of_gaussian_bunch_matches_erf_formula():
    """Rest-frame-like test (p0 tiny): the field of a spherical Gaussian of
    charge Q is Q/(4 pi eps0 r^2) [erf(r/(sqrt2 s)) - sqrt(2/pi) r/s exp(-r^2/2s^2)].
    CIC smoothing biases the grid field at O(h^2/s^2), so the check is a
    convergence test: the error must be a few percent and shrink with n (a
    wrong scale factor or a missing h^3 would be off by O(1))."""
    from scipy.special import erf
    N, s, Qt = 400000, 1e-3, 1e-9
    R, P, Q = gaussian_bunch(N, seed=12, q_total=Qt, p0=1e-9, sigma_p=(0, 0, 0))
    r = np.linalg.norm(R, axis=1)
    sel = (r > 0.8 * s) & (r < 2.5 * s)
    rr = r[sel]
    x = rr / (math.sqrt(2) * s)
    Er = O.COUPLING * Qt / rr ** 2 * (erf(x) - np.sqrt(2 / np.pi) * rr / s * np.exp(-x * x))
    errs = []
    for n in ((32, 32, 32), (64, 64, 64)):
        hr, origin = O.boundp(R, n)
        out = O.compute_self_fields(R, P, Q, np.full(N, 1e-11), 1e-11, origin, hr, n)
        Erad = (out["Ef"][sel] * R[sel]).sum(axis=1) / rr
        errs.append(np.median(np.abs(Erad / Er - 1)))
    assert errs[1] < 0.03 and errs[1] < 0.6 * errs[0], errs


def _gaussian_field(X, sig, Qt):
    """E at points X of a Gaussian charge Qt with per-axis sigmas sig, from
    1/r = (2/sqrt(pi)) int_0^inf exp(-r^2 s^2) ds: convolving the heat kernel
    with the Gaussian gives a one-dimensional integral per point,
    E_i = k Qt (2/sqrt(pi)) int_0^inf 2 s^2 x_i / a_i prod_j a_j^(-1/2)
    exp(-s^2 x_j^2 / a_j) ds with a_j = 1 + 2 s^2 sig_j^2, evaluated by
    400-point Gauss-Legendre in s = tan(theta)/min(sig)."""
    th, w = np.polynomial.legendre.leggauss(400)
    th, w = (th + 1) * (np.pi / 4), w * (np.pi / 4)
    s0 = 1.0 / min(sig)
    s, ds = np.tan(th) * s0, s0 / np.cos(th) ** 2
    sig = np.asarray(sig, dtype=np.float64)
    E = np.zeros_like(X)
    for k in range(len(s)):
        a = 1 + 2 * s[k] ** 2 * sig ** 2
        f = np.prod(a ** -0.5) * np.exp(-(s[k] ** 2 * X ** 2 / a).sum(axis=1))
        E += (w[k] * ds[k] * 2 * s[k] ** 2) * f[:, None] * X / a
    return O.COUPLING * Qt * (2 / np.sqrt(np.pi)) * E


def test_gaussian_field_helper_against_erf_and_quad():
    """The 1D-integral field of a Gaussian equals the erf closed form for a
    sphere and scipy's adaptive quadrature for an elongated bunch."""
    from scipy import integrate
    from scipy.special import erf
    s, Qt = 1e-3, 1e-9
    X = np.random.default_rng(0).normal(size=(50, 3)) * s
    r = np.linalg.norm(X, axis=1)
    x = r / (math.sqrt(2) * s)
    Er = O.COUPLING * Qt / r ** 2 * (erf(x) - np.sqrt(2 / np.pi) * r / s * np.exp(-x * x))
    E = _gaussian_field(X, (s, s, s), Qt)
    assert np.max(np.abs((E * X).sum(1) / r / Er - 1)) < 1e-12
    sig = np.array([s, s, 2 * s])
    for p in X[:3]:
        def f(t, i):
            a = 1 + 2 * t * t * sig ** 2
            return 2 * t * t * p[i] / a[i] * np.prod(a ** -0.5) * np.exp(-(t * t * p ** 2 / a).sum())
        q = [integrate.quad(f, 0, np.inf, args=(i,), limit=200, epsabs=0, epsrel=1e-12)[0] for i in range(3)]
        Eq = O.COUPLING * Qt * 2 / np.sqrt(np.pi) * np.array(q)
        assert np.max(np.abs(_gaussian_field(p[None], sig, Qt)[0] / Eq - 1)) < 1e-12


def _boosted_gaussian_fields(R, s, Qt, g):
    """Lab-frame E and B of a spherical Gaussian (sigma s) moving along z
    with Lorentz factor g: the rest-frame field of the bunch stretched to
    (s, s, g s), evaluated at (x, y, g z), boosted as E_perp = g E'_perp,
    E_z = E'_z, B = v x E / c^2."""
    X = R.copy()
    X[:, 2] *= g
    E = _gaussian_field(X, (s, s, g * s), Qt) * np.array([g, g, 1.0])
    beta = math.sqrt(g * g - 1) / g
    B = np.stack([-beta / C * E[:, 1], beta / C * E[:, 0], np.zeros(len(E))], axis=1)
    return E, B


def test_self_fields_of_moving_gaussian_match_the_boosted_rest_frame_field():
    """PartBunch::computeSelfFields at gamma = 2 (PartBunch.cpp:497-511,573-575,
    641-644): the mesh stretch h_z*gamma, the 1/gamma density, the
    (gamma, gamma, 1/gamma) field scaling and B = beta/c x E together must give
    the field of a moving charge.  The oracle converges to the boosted closed
    form at O(h^2) (8 % at 32^3, 2.3 % at 64^3); swapping gamma and 1/gamma
    gives a 70 % error and the unboosted field 33 %."""
    N, s, Qt, g = 400000, 1e-3, 1e-9, 2.0
    R, P, Q = gaussian_bunch(N, seed=21, q_total=Qt, p0=math.sqrt(g * g - 1), sigma_p=(0, 0, 0))
    r = np.linalg.norm(R, axis=1)
    sel = (r > 0.5 * s) & (r < 2.5 * s)
    Ea, Ba = _boosted_gaussian_fields(R[sel], s, Qt, g)
    Xw = R[sel].copy()
    Xw[:, 2] /= g
    Ewrong = _gaussian_field(Xw, (s, s, s / g), Qt) * np.array([1 / g, 1 / g, 1.0])
    errs = []
    for n in ((32, 32, 32), (64, 64, 64)):
        hr, origin = O.boundp(R, n)
        out = O.compute_self_fields(R, P, Q, np.full(N, 1e-11), 1e-11, origin, hr, n)
        e = rel_l2(out["Ef"][sel], Ea)
        errs.append(e)
        assert rel_l2(out["Bf"][sel], Ba) < 1.05 * e + 1e-3
        assert rel_l2(out["Ef"][sel], Ewrong) > 0.5
    assert errs[1] < 0.035 and errs[1] < 0.4 * errs[0], errs


def test_binned_self_fields_of_moving_gaussian_match_the_boosted_field():
    """The binned path's own Lorentz handling (calcGammas, the /gamma density,
    h_z*gamma, PartBunch.cpp:155-182,239-248) at gamma = 2: one bin holding the
    cold bunch gives the boosted closed-form field, as the plain path does."""
    N, s, Qt, g = 400000, 1e-3, 1e-9, 2.0
    R, P, _ = gaussian_bunch(N, seed=22, q_total=Qt, p0=math.sqrt(g * g - 1), sigma_p=(0, 0, 0))
    assert abs(O.calc_gammas(P, np.zeros(N, np.int32), 1)[0] - g) < 1e-14
    r = np.linalg.norm(R, axis=1)
    sel = (r > 0.5 * s) & (r < 2.5 * s)
    Ea, Ba = _boosted_gaussian_fields(R[sel], s, Qt, g)
    n = (64, 64, 64)
    hr, origin = O.boundp(R, n)
    out = O.compute_self_fields_binned(R, P, np.zeros(N, np.int32), 1, Qt / N, np.full(N, 1e-11), 1e-11,
                                       origin, hr, n, image=False)
    e = rel_l2(out["Ef"][sel], Ea)
    assert e < 0.035, e
    assert rel_l2(out["Bf"][sel], Ba) < 1.05 * e + 1e-3


@pytest.mark.parametrize("g", [1.0, 1.7])
def test_image_charge_is_the_mirrored_bunch(g):
    """The image solve (zshift = -(2 o_z + n h_z) gamma, PartBunch.cpp:252-264;
    FFTPoissonSolver.cpp:289-330) gives the field of the bunch mirrored in the
    cathode plane z = 0 with opposite charge.  Independent construction: on a
    mesh with 2 o_z / h_z = 20 the mirrored cells fall on the bunch's own
    lattice, so one open-boundary solve on a z-extended grid that holds the
    negated, mirrored charge below the bunch's cells gives the image field on
    those cells.  Particles whose stencils avoid the border cells (where Grad
    is one-sided) must agree to solver rounding; the image moved by one cell
    is 5 % off."""
    n, M = (16, 16, 16), 20
    h = np.array([0.5e-3, 0.5e-3, 0.5e-3])
    org = np.array([-8 * h[0], -8 * h[1], (M / 2) * h[2]])
    R, P, _ = gaussian_bunch(20000, seed=31, sigma_r=(0.8e-3, 0.8e-3, 0.8e-3), p0=1e-6, sigma_p=(0, 0, 0))
    R[:, 2] += org[2] + 8 * h[2]
    lgp, _, _ = O.cic_index(R, org, h)
    keep = np.all((lgp >= 1) & (lgp <= np.array(n) - 3), axis=1)
    R, P = R[keep], P[keep]
    N = len(R)
    qi, dT = 1e-9 / N, 1e-11
    dt = np.full(N, dT)
    bins = np.zeros(N, np.int32)
    a = O.compute_self_fields_binned(R, P, bins, 1, qi, dt, dT, org, h, n, image=True, gammas=[g])
    b = O.compute_self_fields_binned(R, P, bins, 1, qi, dt, dT, org, h, n, image=False, gammas=[g])
    Eimg = a["Ef"] - b["Ef"]
    assert np.all(Eimg[:, 2] < 0)  # the image attracts the bunch towards the cathode
    rho, _ = O.scatter(R, np.full(N, qi * dT), org, h, n)
    rho = (rho / dT * (1 / h[0] * 1 / h[1] * 1 / h[2] / g)).reshape(n[2], n[1], n[0])
    nzb = M + 2 * n[2]  # image cells m = -(M+n) .. -(M+1) at z = -z_k', bunch cells m = 0 .. n-1
    big = np.zeros((nzb, n[1], n[0]))
    big[:n[2]] = -rho[::-1]  # image of cell k' at m = -M-1-k', i.e. index n-1-k'
    hs = h * np.array([1.0, 1.0, g])
    nb = (n[0], n[1], nzb)
    phi = O.compute_potential(nb, big.ravel(), hs) * np.prod(hs) * O.COUPLING
    E = (-O.grad(nb, phi, h) * np.array([g, g, 1 / g])).reshape(nzb, n[1], n[0], 3)
    Eg = O.gather(R, np.ascontiguousarray(E[M + n[2]:]).reshape(-1, 3), org, h, n)
    assert rel_l2(Eimg, Eg) < 1e-9
    Eoff = O.gather(R, np.ascontiguousarray(E[M + n[2] - 1:M + 2 * n[2] - 1]).reshape(-1, 3), org, h, n)
    assert rel_l2(Eimg, Eoff) > 0.01


# ---------------------------------------------------------------- pusher
def test_push_is_half_drift():
    rng = np.random.default_rng(7)
    R = rng.normal(size=(100, 3))
    P = rng.normal(size=(100, 3))
    dt = rng.uniform(1e-12, 1e-11, 100)
    g = np.sqrt(1 + (P * P).sum(1))
    R2 = O.push(R, P, dt)
    assert np.allclose(R2, R + 0.5 * C * dt[:, None] * P / g[:, None], rtol=1e-14, atol=1e-20)


def test_kick_pure_E_and_pure_B():
    m, q, dt = O.M_P_EV, 1.0, 1e-11
    P = np.array([[0.0, 0.01, 0.5]])
    E = np.array([[1e6, -2e5, 3e5]])
    P2 = O.kick(P, E, np.zeros((1, 3)), [dt], q, m)
    assert np.allclose(P2, P + dt * q * C / m * E, rtol=1e-15, atol=1e-18)
    B = np.array([[0.0, 0.0, 1.5]])
    P3 = O.kick(P, np.zeros((1, 3)), B, [dt], q, m)
    assert abs(np.linalg.norm(P3) - np.linalg.norm(P)) < 1e-15
    gamma = math.sqrt(1 + (P * P).sum())
    t = 0.5 * dt * q * C * C / (gamma * m) * 1.5
    ang = math.atan2(P3[0, 0], P3[0, 1]) - math.atan2(P[0, 0], P[0, 1])
    assert abs(abs(ang) - 2 * math.atan(t)) < 1e-12
    assert P3[0, 2] == P[0, 2]


# ---------------------------------------------------------------- statistics
def test_moments_and_beam_parameters_vs_fsum():
    R, P, Q = gaussian_bunch(5000, seed=13)
    m = O.calc_moments(R, P)
    part = np.stack([R[:, 0], P[:, 0], R[:, 1], P[:, 1], R[:, 2], P[:, 2]], 1)
    l = 6
    for i in range(6):
        assert abs(m[i] - math.fsum(part[:, i])) <= 1e-12 * np.abs(part[:, i]).sum()
        for j in range(i + 1):
            ex = math.fsum(part[:, i] * part[:, j])
            assert abs(m[l] - ex) <= 1e-12 * np.abs(part[:, i] * part[:, j]).sum()
            l += 1
    s = O.beam_parameters(R, P)
    for d in range(3):
        x, p = part[:, 2 * d], part[:, 2 * d + 1]
        assert abs(s["rrms"][d] - x.std()) < 1e-9 * x.std()
        cov = np.cov(x, p, bias=True)
        assert abs(s["eps_norm"][d] - math.sqrt(np.linalg.det(cov))) < 1e-8 * s["eps_norm"][d]
    g = np.sqrt(1 + (P * P).sum(1)).mean()
    assert abs(s["gamma_mean"] - g) < 1e-14


def test_quaternion_aligns_mean_momentum_with_z():
    u = np.array([0.01, -0.02, 0.8])
    q = O.quaternion_to_z(u)
    assert abs(np.dot(q, q) - 1) < 1e-15
    w, x, y, z = q
    M = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (w * y + x * z)],
                  [2 * (w * z + x * y), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (w * x + y * z), 1 - 2 * (x * x + y * y)]])
    v = M @ (u / np.linalg.norm(u))
    assert np.allclose(v, [0, 0, 1], atol=1e-13)
    assert np.allclose(O.quaternion_to_z([0, 0, 2.0]), [1, 0, 0, 0])


# ---------------------------------------------------------------- tracker
def test_track_zero_charge_is_ballistic_in_reference_frame():
    N, steps, dT = 2000, 7, 1e-11
    R, P, Q = gaussian_bunch(N, seed=14, p0=0.3)
    Q[:] = 0.0
    tr = O.make_tracker((16, 16, 16), dT, ref_p=(0, 0, 0.3))
    R2, P2, _, _, stats = O.track(tr, R, P, Q, np.full(N, dT), steps)
    g = np.sqrt(1 + (P * P).sum(1))
    gref = math.sqrt(1 + 0.09)
    expect = R + steps * C * dT * (P / g[:, None] - np.array([0, 0, 0.3 / gref]))
    assert np.allclose(R2, expect, rtol=0, atol=1e-15)
    assert np.array_equal(P2, P)
    assert abs(tr.pathLength - steps * C * dT * 0.3 / gref) < 1e-15


def test_track_single_particle_gyration_in_solenoid():
    """No space charge, uniform Bz: the transverse momentum rotates by
    2*atan(t) per step with |p| conserved (Boris rotation)."""
    N, steps, dT, Bz = 1, 50, 1e-11, 0.5
    R = np.zeros((1, 3))
    P = np.array([[0.01, 0.0, 0.046181]])
    tr = O.make_tracker((8, 8, 8), dT, space_charge=False, Bz=Bz, ref_p=(0, 0, 0.046181))
    R2, P2, _, _, _ = O.track(tr, R, P, np.zeros(1), np.full(1, dT), steps)
    gamma = math.sqrt(1 + (P * P).sum())
    t = 0.5 * dT * 1.0 * C * C / (gamma * O.M_P_EV) * Bz
    ang = steps * 2 * math.atan(t)
    assert abs(np.hypot(*P2[0, :2]) - 0.01) < 1e-15
    got = math.atan2(P2[0, 1], P2[0, 0])
    assert abs((got - (-ang) + math.pi) % (2 * math.pi) - math.pi) < 1e-12


# ---------------------------------------------------------------- image charge / energy bins
def _pad_z(rho, n, m, reverse=False):
    """rho (x-fastest n0*n1*n2) zero-padded to n2+m planes, optionally z-reversed."""
    r = rho.reshape(n[2], n[1], n[0])
    if reverse:
        r = r[::-1]
    out = np.zeros((n[2] + m, n[1], n[0]))
    out[: n[2]] = r
    return out.ravel()


@pytest.mark.parametrize("sign", [+1, -1])
def test_shifted_solve_equals_standard_solve_on_extended_grid(sign):
    """With zshift = +-m*h2 and m >= n2 no charge sits at zero distance from a
    field point, so the shifted (non-mirrored in z), negated, z-flipped
    convolution must equal the ordinary Hockney solve of the zero-padded
    (for a negative shift: z-reversed) charge on a grid taller by m planes,
    read m planes up.  Both solves share only the IGF corner formula; the
    doubled-grid layout, the z wrap of the second table, the sign and the
    flip are all different code."""
    n = (6, 5, 4)
    m = 5
    rng = np.random.default_rng(11)
    rho = rng.normal(size=int(np.prod(n)))
    hr = np.array([1.3e-3, 0.8e-3, 1.1e-3])
    img = O.compute_potential_shifted(n, rho, hr, sign * m * hr[2]).reshape(n[2], n[1], n[0])
    ext = (n[0], n[1], n[2] + m)
    phi = O.compute_potential(ext, _pad_z(rho, n, m, reverse=(sign < 0)), hr, 0).reshape(n[2] + m, n[1], n[0])
    if sign > 0:
        # img(k) = -conv(n2-1-k), conv(j) = phi_ext(j+m)
        expect = -phi[m:][::-1]
    else:
        # img(k) = -phi_ext[reversed rho](k+m)
        expect = -phi[m:]
    assert rel_l2(img, expect) < 1e-10


def test_shifted_solve_far_field_is_minus_total_charge_over_distance():
    n = (8, 8, 8)
    hr = np.array([1e-3, 1e-3, 1e-3])
    rho = np.random.default_rng(12).random(512)
    zs = -2.0  # 250x the bunch size
    img = O.compute_potential_shifted(n, rho, hr, zs)
    # every pair distance is |zs| up to O(size/zs): phi_img ~ -sum(rho)/|zs|
    assert np.allclose(img, -rho.sum() / abs(zs), rtol=2e-2)
    assert abs(img.mean() / (-rho.sum() / abs(zs)) - 1) < 1e-2


def test_shifted_green_layout():
    n = (4, 4, 4)
    hr = np.array([1e-3, 1e-3, 1e-3])
    G = O.shifted_green_doubled(n, hr, -0.01)
    # even in x and y on the doubled grid, not in z
    for x in range(1, 8):
        assert np.array_equal(G[:, :, x], G[:, :, 8 - x])
    for y in range(1, 8):
        assert np.array_equal(G[:, y, :], G[:, 8 - y, :])
    # z: the kernel at offsets d*h - 10 mm for d = z (z <= 4) and d = z - 8 (z > 4)
    d = np.array([z if z <= 4 else z - 8 for z in range(8)])
    dist = np.abs(d * 1e-3 - 0.01)
    order = np.argsort(dist)
    assert np.all(np.diff(G[order, 0, 0]) < 0)  # closer offsets -> larger kernel
    assert np.allclose(G[:, 0, 0], 1.0 / dist, rtol=1e-3)


def test_calc_gammas_vs_fsum():
    rng = np.random.default_rng(13)
    P = rng.normal(size=(1000, 3)) * [0.1, 0.1, 1.0] + [0, 0, 2.0]
    bins = rng.integers(0, 4, 1000).astype(np.int32)
    bins[bins == 2] = 3  # bin 2 empty
    g = O.calc_gammas(P, bins, 4)
    for b in range(4):
        sel = bins == b
        if sel.any():
            exact = math.fsum(np.sqrt(1.0 + (P[sel] ** 2).sum(1))) / sel.sum()
            assert abs(g[b] / exact - 1) < 1e-14
        else:
            assert g[b] == 1.0


def test_binned_single_bin_without_image_equals_plain_self_fields():
    """One bin holding every particle, no image, the plain path's Lorentz
    factor passed in: the binned body reduces to computeSelfFields() up to the
    order of the 1/h products in the density scale."""
    N, n = 5000, (16, 16, 16)
    R, P, _ = gaussian_bunch(N, seed=14, q_total=1e-9)
    qi = 1e-9 / N
    Q = np.full(N, qi)
    dt = np.full(N, 1e-11)
    hr, org = O.boundp(R, n)
    gz = math.sqrt((P[:, 2].sum() / N) ** 2 + 1.0)
    plain = O.compute_self_fields(R, P, Q, dt, 1e-11, org, hr, n)
    binned = O.compute_self_fields_binned(R, P, np.zeros(N, np.int32), 1, qi, dt, 1e-11, org, hr, n,
                                          image=False, gammas=[gz])
    assert rel_l2(binned["Ef"], plain["Ef"]) < 1e-13
    assert rel_l2(binned["Bf"], plain["Bf"]) < 1e-13


def test_binned_fields_are_linear_in_the_bins():
    """With one Lorentz factor for every bin, the sum over bins of the bin
    fields (and of the image fields) equals the field of one bin holding all
    particles."""
    N, n = 6000, (16, 16, 16)
    R, P, _ = gaussian_bunch(N, seed=15, q_total=1e-9)
    qi = 1e-9 / N
    dt = np.full(N, 1e-11)
    R = R + np.array([0.0, 0.0, 0.05])  # bunch 5 cm downstream of the cathode plane z = 0
    hr, org = O.boundp(R, n)
    bins = np.random.default_rng(16).integers(0, 3, N).astype(np.int32)
    g = [1.2, 1.2, 1.2]
    three = O.compute_self_fields_binned(R, P, bins, 3, qi, dt, 1e-11, org, hr, n, gammas=g)
    one = O.compute_self_fields_binned(R, P, np.zeros(N, np.int32), 1, qi, dt, 1e-11, org, hr, n, gammas=g[:1])
    assert rel_l2(three["Ef"], one["Ef"]) < 1e-13
    assert rel_l2(three["Bf"], one["Bf"]) < 1e-13
    # the image field is present and opposes the bunch's own longitudinal push
    noimg = O.compute_self_fields_binned(R, P, np.zeros(N, np.int32), 1, qi, dt, 1e-11, org, hr, n,
                                         image=False, gammas=g[:1])
    dEz = one["Ef"][:, 2] - noimg["Ef"][:, 2]
    assert np.all(dEz < 0) or np.all(dEz > 0)


# --- DC beam (BCFFTZ=PERIODIC): particle side -----------------------------------

def test_boundp_dc_rule_by_hand():
    """boundp of a DC beam (PartBunchBase.hpp:814,840-843): x and y as for an
    open beam, z not padded: hz = period/(n2-1), origin_z = zmin - hz/2."""
    R = np.array([[0.0, 0.0, 0.0], [1e-3, 2e-3, 5e-3], [-1e-3, -2e-3, 1e-3]])
    n, period = (8, 8, 16), 0.01
    hr, org = O.boundp(R, n, period=period)
    hro, orgo = O.boundp(R, n)
    assert np.array_equal(hr[:2], hro[:2]) and np.array_equal(org[:2], orgo[:2])
    assert hr[2] == period / 15
    assert org[2] == 0.0 - hr[2] / 2.0
    assert hro[2] != hr[2]  # the open rule pads the z extent instead


def test_wrap_periodic_z_by_hand():
    """ParticlePeriodicBCond over [origin, n*h + origin): below the region
    t -> max - (min - t), at or above it t -> min + (t - max); x, y untouched."""
    origin, hr, n = np.array([0.0, 0.0, -0.5]), np.array([1.0, 1.0, 0.25]), (4, 4, 8)
    z = np.array([-0.75, 1.5, 1.75, 0.3, -0.5, 1.25])
    R = np.zeros((len(z), 3))
    R[:, 0], R[:, 1], R[:, 2] = 7.0, -3.0, z
    W = O.wrap_periodic_z(R, origin, hr, n)
    assert np.array_equal(W[:, 2], [1.25, -0.5, -0.25, 0.3, -0.5, 1.25])
    assert np.all(W[:, 0] == 7.0) and np.all(W[:, 1] == -3.0)


def test_periodic_gather_reads_the_opposite_layer():
    """With periodic z faces the gather's z guard layers hold the opposite
    cells: a particle in the top cell reads layer 0 with its upper z weight, one
    below the first cell centre reads layer n2-1 with its lower z weight; with
    ZeroFace both read 0 there.  Weights 0.5 (x, y) and 0.25/0.75 (z) are exact."""
    n = (4, 4, 6)
    E = np.zeros((n[2], n[1], n[0], 3))
    E[0] = (1.0, 2.0, 3.0)
    E[n[2] - 1] = (4.0, 5.0, 6.0)
    E = E.reshape(-1, 3)
    origin, hr = np.zeros(3), np.ones(3)
    R = np.array([[2.0, 2.0, 5.75],    # lgp = (1, 1, 5), dz = 0.25: upper corners at k = 6 -> 0
                  [2.0, 2.0, 0.25]])   # lgp = (1, 1, -1), dz = 0.75: lower corners at k = -1 -> 5
    Ep = O.gather(R, E, origin, hr, n, periodic_z=True)
    Eo = O.gather(R, E, origin, hr, n)
    assert np.array_equal(Ep[0], 0.25 * np.array([1.0, 2.0, 3.0]) + 0.75 * np.array([4.0, 5.0, 6.0]))
    assert np.array_equal(Ep[1], 0.25 * np.array([4.0, 5.0, 6.0]) + 0.75 * np.array([1.0, 2.0, 3.0]))
    assert np.array_equal(Eo[0], 0.75 * np.array([4.0, 5.0, 6.0]))
    assert np.array_equal(Eo[1], 0.75 * np.array([1.0, 2.0, 3.0]))


def test_dc_self_fields_of_a_long_beam_have_less_axial_field():
    """A beam uniform in z over the period: with the periodic solve and the
    periodic gather its middle sees a much weaker E_z than the same particles
    on an open mesh, which see the end fields of a finite cylinder.  (The DC
    mesh is not exactly periodic for the particles: its period is
    n2*hz = period*n2/(n2-1), the top layer's upper deposits are dropped and
    Grad is one-sided at the z faces, so E_z does not vanish near the ends.)"""
    rng = np.random.default_rng(4)
    N, n, period = 40000, (16, 16, 32), 0.02
    R = np.column_stack([rng.normal(0, 1e-3, N), rng.normal(0, 1e-3, N), rng.uniform(0, period, N)])
    P = np.zeros((N, 3))
    P[:, 2] = 0.046181
    Q, dt = np.full(N, 1e-9 / N), np.full(N, 1e-11)
    hr, org = O.boundp(R, n, period=period)
    Rw = O.wrap_periodic_z(R, org, hr, n)
    assert np.array_equal(Rw, R)  # every particle lies inside the region already
    dc = O.compute_self_fields(Rw, P, Q, dt, 1e-11, org, hr, n, periodic_z=True)
    hro, orgo = O.boundp(R, n)
    op = O.compute_self_fields(R, P, Q, dt, 1e-11, orgo, hro, n)
    mid = np.abs(R[:, 2] - period / 2) < 0.1 * period
    def rms_z(Ef):
        return np.sqrt(np.mean(Ef[mid, 2] ** 2))
    def rms_r(Ef):
        return np.sqrt(np.mean(Ef[mid, 0] ** 2 + Ef[mid, 1] ** 2))
    assert rms_z(dc["Ef"]) < 0.5 * rms_z(op["Ef"])
    assert rms_r(dc["Ef"]) > 0.5 * rms_r(op["Ef"])  # the transverse field is the same order


def test_dc_tracker_wraps_in_the_beam_frame():
    """In the tracker's space-charge step the beam-frame z of every particle is
    wrapped into the mesh's region after boundp; with <P> along z the beam frame
    is the reference frame shifted by the path length (0 here)."""
    rng = np.random.default_rng(8)
    N, n, period = 4000, (8, 8, 16), 0.01
    R = np.column_stack([rng.normal(0, 1e-3, N), rng.normal(0, 1e-3, N), rng.uniform(-0.008, 0.008, N)])
    P = np.zeros((N, 3))
    P[:, 0] = rng.normal(0, 1e-3, N // 2).repeat(2) * np.tile([1.0, -1.0], N // 2)
    P[:, 1] = rng.normal(0, 1e-3, N // 2).repeat(2) * np.tile([1.0, -1.0], N // 2)
    P[:, 2] = 0.046181
    Q, dt = np.full(N, 1e-9 / N), np.full(N, 1e-11)
    tr = O.make_tracker(n, 1e-11, ref_p=(0.0, 0.0, 0.046181), periodic_z=True, period=period)
    R2, Ef, Bf = R.copy(), np.zeros((N, 3)), np.zeros((N, 3))
    mesh = np.zeros(6)
    import ctypes as C
    O.lib().orc_space_charge(C.byref(tr), N, R2, P, Q.copy(), dt, Ef, Bf, mesh, None, None, None)
    hr, org = mesh[:3], mesh[3:]
    lo, hi = org[2], n[2] * hr[2] + org[2]
    assert np.all(R2[:, 2] >= lo) and np.all(R2[:, 2] < hi + 1e-15)
    moved = R2[:, 2] != R[:, 2]
    assert moved.any() and np.allclose(R[moved, 2] - R2[moved, 2], hi - lo, rtol=0, atol=1e-15)

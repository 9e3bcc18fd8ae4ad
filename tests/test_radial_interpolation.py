"""The radial interpolation behind the multipole path's nodes mode, checked
on the CPU against SciPy's spherical Bessel functions: within a bin of width
h with q h <= 1.04, the interpolant of j_l(q r) at the bin's 8 Chebyshev
nodes is within (q h / 2)^8 / (2^7 8!) = 1.0e-9 of it, for every order, and
the moments it implies reproduce the per-point sum."""
import numpy as np
from scipy.special import spherical_jn

NODES = 8


def chebyshev(n=NODES):
    x = np.cos((2 * np.arange(n) + 1) * np.pi / (2 * n))
    den = np.array([np.prod([x[m] - x[k] for k in range(n) if k != m]) for m in range(n)])
    return x, den


def lagrange(u, x, den):
    """w[i, m]: the Lagrange basis of nodes x at points u (prefix and suffix
    products, as the geometry kernel forms them)."""
    u = np.asarray(u)[:, None]
    d = u - x[None, :]
    pre = np.cumprod(np.hstack([np.ones_like(u), d]), axis=1)[:, :-1]
    suf = np.cumprod(np.hstack([d, np.ones_like(u)])[:, ::-1], axis=1)[:, ::-1][:, 1:]
    return pre * suf / den[None, :]


def test_error_bound_holds_for_every_order():
    x, den = chebyshev()
    bound = 0.52 ** 8 / (2 ** 7 * 40320)
    assert 1.0e-9 < bound < 1.05e-9
    rng = np.random.default_rng(0)
    worst = 0.0
    for h, q in ((2.0, 0.52), (1.0, 1.04), (0.25, 4.16)):
        for b in (0, 3, 49):
            centre = (b + 0.5) * h
            r = centre + 0.5 * h * rng.uniform(-1, 1, 200)
            w = lagrange((r - centre) / (0.5 * h), x, den)
            nodes = centre + 0.5 * h * x
            for l in range(0, 65, 3):
                exact = spherical_jn(l, q * r)
                interp = w @ spherical_jn(l, q * nodes)
                worst = max(worst, np.abs(interp - exact).max())
    assert worst <= bound, worst


def test_moments_reproduce_the_point_sum():
    """sum_i occ_i j_l(q r_i) Y_i = sum_m j_l(q rho_m) M_m with M_m =
    sum_i occ_i w_m(r_i) Y_i, for several q at once (the moments do not
    depend on q)."""
    x, den = chebyshev()
    rng = np.random.default_rng(1)
    h = 2.0
    b = 7
    centre = (b + 0.5) * h
    r = centre + 0.5 * h * rng.uniform(-1, 1, 500)
    occ = rng.uniform(0.5, 1.0, 500)
    y = rng.normal(size=500)  # one harmonic column
    m = (occ * y) @ lagrange((r - centre) / (0.5 * h), x, den)
    nodes = centre + 0.5 * h * x
    for q in (0.0, 0.1, 0.3, 0.52):
        for l in (0, 5, 20, 40):
            direct = np.sum(occ * spherical_jn(l, q * r) * y)
            via = np.sum(spherical_jn(l, q * nodes) * m)
            assert abs(via - direct) <= 1.1e-9 * np.sum(np.abs(occ * y)), (q, l)


def test_chebyshev_moments_give_the_lagrange_weights():
    """The radial Debye sweep accumulates c T_p(u) per pair and maps the
    moments to the nodes by (2 - [p = 0]) / 8 T_p(x_m): at the Chebyshev
    nodes this equals the Lagrange basis L_m(u) (discrete orthogonality), so
    the node sums match the Lagrange form for any u in [-1, 1]."""
    x, den = chebyshev()
    n = len(x)
    u = np.random.default_rng(3).uniform(-1, 1, 400)
    t = np.cos(np.outer(np.arange(n), np.arccos(u)))  # T_p(u), [p][pair]
    rec = np.empty_like(t)  # the kernel's recurrence
    rec[0], rec[1] = 1.0, u
    for p in range(2, n):
        rec[p] = 2 * u * rec[p - 1] - rec[p - 2]
    assert np.abs(rec - t).max() < 1e-13
    mp = np.array([[(1.0 if p == 0 else 2.0) / n * np.cos(p * (2 * m + 1) * np.pi / (2 * n))
                    for p in range(n)] for m in range(n)])
    assert np.allclose(mp @ rec, lagrange(u, x, den).T, atol=1e-13)
    # a wrong normalisation (2 / n for p = 0 too) misses it
    bad = mp.copy()
    bad[:, 0] *= 2
    assert not np.allclose(bad @ rec, lagrange(u, x, den).T, atol=1e-6)

"""Host-side fitting and ensemble arithmetic of the library (no GPU needed),
checked against the FP64 oracle, finite differences and the reference's own
fit tests (fit.rs:489-660) with bases computed by the oracle."""
import math

import numpy as np
import pytest

import crabsaxs_amd as csx
import oracle_py as O


def curve(q, y, s):
    return csx.ExperimentalCurve(np.asarray(q), np.asarray(y), np.asarray(s))


@pytest.mark.parametrize("scale_mode,bg_mode", [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2)])
def test_solve_scale_background_bitwise(scale_mode, bg_mode):
    rng = np.random.default_rng(1)
    x = rng.uniform(1, 100, 40)
    y = 1.3 * x + 2 + rng.normal(size=40)
    s = rng.uniform(0.5, 2, 40)
    sm = csx.ScaleMode.Fixed(1.25) if scale_mode else csx.ScaleMode.Fit
    bm = [csx.BackgroundMode.Fit, csx.BackgroundMode.FixedZero, csx.BackgroundMode.Fixed(1.5)][bg_mode]
    got = csx.solve_scale_background(x, curve(x, y, s), sm, bm)
    want = O.solve_scale_background(x, y, s, scale_mode, 1.25, bg_mode, 1.5)
    assert got == tuple(float(v) for v in want)


def test_hermite_simplex_weights_bitwise():
    rng = np.random.default_rng(2)
    x = np.sort(rng.uniform(0, 1, 12))
    y = rng.normal(size=12)
    t = np.linspace(x[0], x[-1], 57)
    assert np.array_equal(csx.cubic_hermite_interpolate(x, y, t), O.cubic_hermite(x, y, t))
    v = rng.normal(size=9)
    assert np.array_equal(csx.project_simplex(v), O.project_simplex(v))
    q = np.linspace(0.01, 0.5, 60)
    curves = np.stack([np.exp(-(q * r) ** 2 / 3) * 1e4 for r in (10, 15, 20, 30)])
    yv = 0.7 * curves[0] + 0.3 * curves[3] + 3.0
    sg = 0.01 * yv
    for uniform in (False, True):
        w, sc, bg, chi, fitted = csx.optimize_weights(curves, curve(q, yv, sg), uniform=uniform,
                                                      regularization=0.1)
        ow, (osc, obg, ochi) = O.optimize_weights(curves, yv, sg, uniform=uniform,
                                                  regularization=0.1)
        assert np.array_equal(w, ow) and (sc, bg, chi) == (osc, obg, ochi)
        assert np.allclose(fitted, sc * (w @ curves) + bg, rtol=1e-14)


def synthetic_basis(q):
    atoms = O.Atoms([[0, 0, 0], [2, 0, 0]], [1, 3])
    return O.debye_solvent_basis(atoms, q, implicit_h=False)


def test_chi2_gradient_matches_finite_differences():
    q = np.linspace(0.01, 0.5, 50)
    b = synthetic_basis(q)
    y = 1.7 * O.combine_solvent_basis(b, q, 1.03, 0.4) + 0.02
    e = curve(q, y * (1 + 0.01 * np.sin(40 * q)), 0.01 * y + 1e-3)
    p = csx.FitParams(scale=1.6, background=0.05, c1=1.01, c2=0.3)
    chi, g = csx.chi2_grad(b, e, p)
    h = [1e-6, 1e-8, 1e-7, 1e-7]
    for k, name in enumerate(["scale", "background", "c1", "c2"]):
        up, dn = csx.FitParams(**vars(p)), csx.FitParams(**vars(p))
        setattr(up, name, getattr(p, name) + h[k])
        setattr(dn, name, getattr(p, name) - h[k])
        fd = (csx.chi2_grad(b, e, up)[0] - csx.chi2_grad(b, e, dn)[0]) / (2 * h[k])
        assert abs(fd - g[k]) <= 1e-5 * max(abs(fd), 1e-3), (name, fd, g[k])


@pytest.mark.parametrize("optimizer", [csx.Optimizer.NelderMead, csx.Optimizer.Gradient])
def test_exact_recovery_from_basis(optimizer):
    q = np.linspace(0.01, 0.5, 80)
    b = synthetic_basis(q)
    y = 1.7 * O.combine_solvent_basis(b, q, 1.03, 0.4) + 0.02
    e = curve(q, y, np.full(q.size, 0.01))
    opts = csx.FitOptions(optimizer=optimizer, max_iterations=200)
    r = csx.fit_from_basis(b, e, csx.FitParams(), opts)
    assert abs(r.params.c1 - 1.03) < 1e-3 and abs(r.params.c2 - 0.4) < 1e-2, r
    assert abs(r.params.scale - 1.7) < 1e-2 and r.chi2 < 1e-4, r


def lcg_noise(n):
    """The deterministic LCG + Box-Muller noise of fit.rs:555-566."""
    state = 0x5EED
    out = []
    for _ in range(n):
        vals = []
        for _ in range(2):
            state = (state * 6364136223846793005 + 1) & 0xFFFFFFFFFFFFFFFF
            vals.append(max((state >> 11) / float(1 << 53), 2.2250738585072014e-308))
        u1, u2 = vals
        out.append(math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2))
    return np.array(out)


@pytest.mark.parametrize("optimizer", [csx.Optimizer.NelderMead, csx.Optimizer.Gradient])
def test_recovers_solvent_parameters_with_deterministic_noise(optimizer):
    # fit.rs:536-599 with the oracle's multipole basis (two atoms, explicit H)
    atoms = O.Atoms([[0, 0, 0], [2, 0, 0]], [1, 3])
    q = np.arange(1, 81) * 0.006
    raw = O.multipole_solvent(atoms, q, 8, c1=1.02, c2=0.55, implicit_h=False, exact=False)
    y = 1.4e-4 * raw + 0.03 + 0.0001 * lcg_noise(q.size)
    e = curve(q, y, np.full(q.size, 0.002))
    basis = O.multipole_solvent(atoms, q, 9, basis=True, implicit_h=False, exact=False)
    opts = csx.FitOptions(method=csx.ScatteringMethod.Multipole,
                          hydrogen_mode=csx.HydrogenMode.Explicit, max_iterations=100,
                          q_sampling_stride=1, optimizer=optimizer)
    r = csx.fit_from_basis(basis, e, csx.FitParams(1.4e-4, 0.03, 1.02, 0.55), opts)
    if optimizer == csx.Optimizer.Gradient:
        # chi^2 has a flat c1/c2 valley here; the gradient fit must reach a chi^2
        # no worse than Nelder-Mead's (it may stop elsewhere in the valley).
        opts.optimizer = csx.Optimizer.NelderMead
        nm = csx.fit_from_basis(basis, e, csx.FitParams(1.4e-4, 0.03, 1.02, 0.55), opts)
        assert r.chi2 <= nm.chi2 * (1 + 1e-6) and r.chi2 < 0.01, (r, nm)
        return
    assert abs(r.params.c1 - 1.02) < 0.02, r
    assert abs(r.params.c2 - 0.55) < 0.08, r
    assert abs(r.params.scale - 1.4e-4) < 0.4e-4, r
    assert abs(r.params.background - 0.03) < 0.01, r
    assert r.chi2 < 0.01, r


def test_nelder_mead_respects_bounds_and_penalty():
    q = np.linspace(0.01, 0.5, 64)
    b = synthetic_basis(q)
    y = 1.0 * O.combine_solvent_basis(b, q, 1.05, 4.0) + 0.0
    e = curve(q, y, np.full(q.size, 0.01))
    r = csx.fit_from_basis(b, e, csx.FitParams(c1=2.0, c2=9.0), csx.FitOptions())
    assert 0.95 <= r.params.c1 <= 1.05 and -2.0 <= r.params.c2 <= 4.0
    assert r.iterations <= 64


def test_invalid_fit_options_raise():
    q = np.linspace(0.01, 0.5, 10)
    b = synthetic_basis(q)
    e = curve(q, np.ones(10), np.ones(10))
    with pytest.raises(csx.InvalidInput):
        csx.fit_from_basis(b, e, csx.FitParams(), csx.FitOptions(c1_bounds=(1.1, 1.0)))
    with pytest.raises(csx.InvalidInput):
        csx.fit_from_basis(b, e, csx.FitParams(), csx.FitOptions(max_iterations=0))
    with pytest.raises(csx.InvalidInput):
        csx.fit_from_basis(b, curve(q, np.ones(10), np.zeros(10)), csx.FitParams(), csx.FitOptions())


def test_compare_curves_and_quality_presets():
    q = np.linspace(0.01, 0.3, 30)
    y = np.exp(-q * 10)
    e = curve(q, y, np.full(30, 0.01))
    c = csx.compare_curves(e, y * 1.01)
    assert c.r_factor == pytest.approx(0.01, rel=1e-9) and c.pearson_correlation > 0.999
    assert csx.Quality.Fast.fit_options(csx.FitOptions()).max_iterations == 32
    assert csx.Quality.Accurate.fit_options(csx.FitOptions()).q_sampling_stride == 1
    with pytest.raises(csx.InvalidInput):
        csx.compare_curves(e, y[:5])


def _nnls_reference(curves, y, sigma, fit_bg=True):
    """scipy.optimize.nnls (Lawson-Hanson) on the same background-eliminated,
    sigma-weighted system."""
    from scipy.optimize import nnls as scipy_nnls
    u = 1.0 / sigma
    E = (curves / sigma).T.copy()
    f = y / sigma
    if fit_bg:
        P = np.eye(len(y)) - np.outer(u, u) / (u @ u)
        E, f = P @ E, P @ f
    a, _ = scipy_nnls(E, f, maxiter=50 * E.shape[1])
    return a


@pytest.mark.parametrize("fit_bg", [True, False])
def test_nnls_weights_match_scipy_and_recover_mixtures(fit_bg):
    rng = np.random.default_rng(17)
    q = np.linspace(0.01, 0.5, 120)
    m = 40
    # smooth positive curves of varying decay: strongly correlated columns
    rg = rng.uniform(10.0, 40.0, size=m)
    curves = np.stack([np.exp(-(q * r) ** 2 / 3.0) * (1 + 0.1 * np.sin(q * r * 3)) for r in rg])
    truth = np.zeros(m)
    pick = rng.choice(m, size=4, replace=False)
    truth[pick] = rng.dirichlet(np.ones(4))
    bg = 0.02 if fit_bg else 0.0
    y = 3.0 * truth @ curves + bg
    sigma = 0.01 * y
    opts = csx.FitOptions(background_mode=csx.BackgroundMode.Fit if fit_bg
                          else csx.BackgroundMode.FixedZero)
    w, scale, b, chi, fitted = csx.nnls_weights(curves, curve(q, y, sigma), opts)
    assert abs(np.sum(w) - 1) < 1e-12 and np.all(w >= 0)
    # noise-free data: an exact fit
    assert chi < 1e-12, chi
    assert np.allclose(fitted, y, rtol=1e-7)
    # same optimum as scipy's NNLS on the same system
    a = _nnls_reference(curves, y, sigma, fit_bg)
    got = scale * w
    assert np.allclose(curves.T @ got, curves.T @ a, rtol=1e-7)
    # noisy data: chi2 of the exact NNLS optimum
    y2 = y + sigma * rng.normal(size=q.size)
    w2, s2, b2, chi2, _ = csx.nnls_weights(curves, curve(q, y2, sigma), opts)
    a2 = _nnls_reference(curves, y2, sigma, fit_bg)
    u = 1.0 / sigma
    r_ref = (y2 - curves.T @ a2) / sigma
    if fit_bg:
        r_ref = r_ref - (r_ref @ u) / (u @ u) * u
    ref_chi = float(r_ref @ r_ref) / (q.size - 1)
    assert chi2 <= ref_chi * (1 + 1e-9) + 1e-15
    # the reference's projected gradient never beats the exact optimum
    _, _, _, chi_pg, _ = csx.optimize_weights(curves, curve(q, y2, sigma), opts)
    assert chi2 <= chi_pg * (1 + 1e-12)


def test_nnls_weights_reject_fixed_scale_and_regularization():
    q = np.linspace(0.01, 0.3, 20)
    curves = np.stack([np.exp(-q * k) for k in (1.0, 2.0, 3.0)])
    y = curves[0] + 0.5 * curves[2]
    with pytest.raises(csx.InvalidInput):
        csx.nnls_weights(curves, curve(q, y, 0.01 * y),
                         csx.FitOptions(scale_mode=csx.ScaleMode.Fixed(1.0)))


def _chacha(key, counter, stream, rounds):
    import ctypes as C
    from crabsaxs_amd import _lib
    k = (C.c_uint32 * 8)(*key)
    out = (C.c_uint32 * 16)()
    assert _lib.load().csx_chacha_block(k, counter, stream, rounds, out) == 0
    return b"".join(int(w).to_bytes(4, "little") for w in out)


def test_chacha_block_matches_published_vectors():
    """The block function under the bootstrap stream: RFC 7539 section 2.3.2
    (key 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00),
    RFC 7539 A.1 #1 (all-zero key, nonce, counter) and the all-zero-key
    ChaCha8 keystream block."""
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    got = _chacha(key, 0x0900000000000001, 0x4A000000, 20)
    assert got.hex() == ("10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
                         "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert _chacha([0] * 8, 0, 0, 20).hex() == (
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert _chacha([0] * 8, 0, 0, 8).hex() == (
        "3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e"
        "984ce172b9216f419f445367456d5619314a42a3da86b001387bfdb80e0cfe42")


def test_bootstrap_indices_stream():
    a = csx.bootstrap_indices(0xC0FFEE, 150, 40)
    assert a.shape == (40, 150) and a.min() >= 0 and a.max() < 150
    assert np.array_equal(a, csx.bootstrap_indices(0xC0FFEE, 150, 40))
    # replicate-major: fewer replicates are a prefix of the same stream
    assert np.array_equal(csx.bootstrap_indices(0xC0FFEE, 150, 3), a[:3])
    assert not np.array_equal(csx.bootstrap_indices(1, 150, 3), a[:3])
    # uniform: chi-square over 150 cells of 6000 draws
    counts = np.bincount(a.ravel(), minlength=150)
    chi = np.sum((counts - 40.0) ** 2 / 40.0)
    assert chi < 220, chi
    # n = 1 always draws 0
    assert np.all(csx.bootstrap_indices(7, 1, 5) == 0)


def test_pcg32_matches_published_vectors():
    """The PCG32 generator of the bootstrap's seed expansion against the
    reference PCG32 demo (pcg32_srandom(42, 54): increment 2*54+1, state
    advanced by the published seeding steps) and its first six outputs."""
    import ctypes as C
    from crabsaxs_amd import _lib
    mask = (1 << 64) - 1
    mul = 6364136223846793005
    inc = (54 << 1) | 1
    state = (((inc + 42) & mask) * mul + inc) & mask  # step, add 42, step
    out = (C.c_uint32 * 6)()
    assert _lib.load().csx_pcg32(state, inc, 6, 0, out) == 0
    assert [int(v) for v in out] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293,
                                     0xbfa4784b, 0xcbed606e]
    # the seed expansion's convention: outputs of the state after each step
    new = (C.c_uint32 * 5)()
    assert _lib.load().csx_pcg32((state - inc) * pow(mul, -1, 1 << 64) & mask, inc, 5, 1, new) == 0
    assert [int(v) for v in new] == [int(v) for v in out[:5]]

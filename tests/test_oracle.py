"""Pins of the FP64 CPU oracle: golden vectors (40-digit mpmath), closed forms,
library routines (scipy), invariants and the reference's own test assertions.
Each pin is chosen so that a dropped term, a wrong sign or index or a swapped
operand in the oracle fails it."""
import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy.special import spherical_jn

import oracle_py as O

GOLD = json.loads((Path(__file__).parent / "golden" / "debye_golden.json").read_text())
SPECIAL = json.loads((Path(__file__).parent / "golden" / "special_golden.json").read_text())


def atoms(name):
    s = GOLD["structures"][name]
    return O.Atoms(s["xyz"], s["element"])


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize("name", sorted(GOLD["structures"]))
@pytest.mark.parametrize("tag", ["low", "high"])
def test_debye_and_basis_match_mpmath(name, tag):
    q = np.array(GOLD[f"q_{tag}"])
    a = atoms(name)
    want = np.array(GOLD["structures"][name][f"debye_{tag}"])
    assert rel(O.debye_intensity(a, q), want).max() < 1e-12
    basis = np.array(GOLD["structures"][name][f"basis_{tag}"])
    got = O.debye_solvent_basis(a, q)
    scale = np.abs(basis).max(axis=1, keepdims=True)
    assert (np.abs(got - basis) / scale).max() < 1e-12


def test_sinc_closed_form():
    L = O.lib()
    assert L.or_sinc(0.0) == 1.0  # scattering.rs:246-248
    assert L.or_sinc(5e-9) == 1.0
    for x in (1e-7, 0.3, 2.0, 31.4, -4.0):
        assert abs(L.or_sinc(x) - math.sin(x) / x) <= 1e-16


def test_two_atom_closed_form():
    # scattering.rs:250-257
    a = O.Atoms([[0, 0, 0], [2, 0, 0]], [1, 1])
    q = 0.2
    f = O.lib().or_form_factor(1, -1, 1.0, 1, q)
    assert abs(O.debye_intensity(a, [q])[0] - 2 * f * f * (1 + math.sin(2 * q) / (2 * q))) < 1e-10


def test_q_zero_is_square_of_sum_and_invariance():
    rng = np.random.default_rng(1)
    xyz = rng.normal(size=(30, 3)) * 5
    el = rng.integers(1, 5, size=30)
    a = O.Atoms(xyz, el)
    f0 = sum(O.lib().or_form_factor(int(e), -1, 1.0, 1, 0.0) for e in el)
    assert rel(O.debye_intensity(a, [0.0]), np.array([f0 * f0])).max() < 1e-13
    q = np.linspace(0.01, 0.8, 17)
    base = O.debye_intensity(a, q)
    th = 0.7
    rot = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    moved = O.Atoms(xyz @ rot.T + np.array([3.0, -2.0, 7.5]), el)
    assert rel(O.debye_intensity(moved, q), base).max() < 1e-11
    perm = rng.permutation(30)
    assert rel(O.debye_intensity(O.Atoms(xyz[perm], el[perm]), q), base).max() < 1e-11


def test_basis_recombines_to_direct_curve():
    # solvent.rs:651-665, plus the analytic combination identity
    a = O.Atoms([[-1, 0, 0], [1, 0, 0]], [1, 3])
    q = [0.0, 0.1, 0.3, 0.5]
    b = O.debye_solvent_basis(a, q, c1=1.04, c2=0.35)
    comb = O.combine_solvent_basis(b, q, 1.04, 0.35)
    # direct: single amplitude A = v - G e + c2 h in a two-atom Debye sum
    L = O.lib()
    for k, qq in enumerate(q):
        g = L.or_excluded_volume_scale(qq, 1.04)
        amp = []
        for el in (1, 3):
            v = L.or_form_factor(el, -1, 1.0, 1, qq)
            vol = L.or_excluded_volume(el, -1, 1)
            e = 0.334 * vol * math.exp(-vol ** (2 / 3) * qq * qq / (4 * math.pi))
            r = (3 * vol / (4 * math.pi)) ** (1 / 3)
            sv = 4 * math.pi * (r + 1.5) ** 2 * 3
            h = 0.334 * 0.1 * sv * math.exp(-sv ** (2 / 3) * qq * qq / (4 * math.pi))
            amp.append(v - g * e + 0.35 * h)
        s = 1.0 if qq == 0 else math.sin(2 * qq) / (2 * qq)
        direct = amp[0] ** 2 + amp[1] ** 2 + 2 * amp[0] * amp[1] * s
        assert abs(comb[k] - direct) <= max(abs(direct), 1) * 1e-12


def test_excluded_volume_scale_closed_form():
    L = O.lib()
    assert L.or_excluded_volume_scale(0.3, 1.0) == 1.0
    for q, c1 in [(0.1, 1.05), (0.5, 0.95)]:
        want = c1 ** 3 * math.exp(-(c1 * c1 - 1) * 1.62 ** 2 * q * q / 4)
        assert abs(L.or_excluded_volume_scale(q, c1) - want) < 1e-15


def test_water_form_factor_ten_electrons():
    assert abs(O.lib().or_water_form_factor(0.0) - 10.0) < 1e-3  # solvent.rs:774-777


@pytest.mark.parametrize("row", SPECIAL["bessel"], ids=lambda r: f"L{r['l_max']}x{r['x']}")
def test_exact_bessel_matches_mpmath(row):
    got = O.bessel_exact(row["l_max"], row["x"])
    want = np.array(row["j"])
    assert (np.abs(got - want) / np.abs(want).max()).max() < 1e-13
    assert np.allclose(spherical_jn(np.arange(row["l_max"] + 1), row["x"]), want,
                       rtol=1e-10, atol=1e-14 * np.abs(want).max())


def test_reference_bessel_accurate_regime_and_zero_limit():
    assert list(O.bessel_reference(4, 0.0)) == [1.0, 0.0, 0.0, 0.0, 0.0]  # spherical.rs:592-594
    for L, x in [(15, 0.5), (15, 20.0), (64, 3.0), (8, 7.5)]:
        ref = spherical_jn(np.arange(L + 1), x)
        assert np.abs(O.bessel_reference(L, x) - ref).max() / np.abs(ref).max() < 1e-13
    # and it is known to go wrong for x >> L (the regime the exact j_l covers)
    ref = spherical_jn(np.arange(16), 45.0)
    assert np.abs(O.bessel_reference(15, 45.0) - ref).max() / np.abs(ref).max() > 0.5


@pytest.mark.parametrize("row", SPECIAL["harmonics"], ids=lambda r: f"L{r['l_max']}")
def test_real_harmonics_match_mpmath(row):
    got = O.real_harmonics(row["l_max"], row["cos_theta"], row["phi"])
    want = np.array(row["y"])
    assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max())


def test_harmonics_addition_theorem():
    for ct, phi in [(0.2, 0.4), (-0.9, 2.0), (1.0, 0.0)]:
        y = O.real_harmonics(20, ct, phi)
        for l in range(21):
            s = np.sum(y[l * l:(l + 1) ** 2] ** 2)
            assert abs(s - (2 * l + 1) / (4 * math.pi)) < 1e-12


def test_multipole_matches_debye_reference_tests():
    # spherical.rs:596-634 and 637-678: two atoms, L = 32, 1e-6
    for xyz, el, q in [([[0, 0, 0], [2, 0, 0]], [1, 1], [0.1, 0.2, 0.5]),
                       ([[-20, 0, 0], [20, 0, 0]], [1, 3], [0.2, 0.35, 0.5])]:
        a = O.Atoms(xyz, el)
        m = O.multipole_intensity(a, q, 32, exact=False)
        d = O.debye_intensity(a, q)
        assert rel(m, d).max() < 1e-6
    # spherical.rs:681-727: 24-atom helix, L = 24, NRMSE <= 1e-3
    a = atoms("helix24")
    q = np.arange(1, 101) * 0.005
    for exact in (False, True):
        m = O.multipole_intensity(a, q, 24, exact=exact)
        d = O.debye_intensity(a, q)
        assert np.sqrt(((m - d) ** 2).sum() / (d ** 2).sum()) <= 1e-3


def test_adaptive_multipole_matches_debye_spiral():
    # scattering.rs:312-365: 96-atom spiral, adaptive l_max, rel-L2 <= 1e-3
    a = atoms("spiral96")
    q = np.arange(1, 101) * 0.005
    c = a.xyz.mean(axis=0)
    rmax = np.linalg.norm(a.xyz - c, axis=1).max()
    L = int(min(max(math.ceil(q.max() * rmax) + 8, 8), 64))
    m = O.multipole_intensity(a, q, L, exact=False)
    d = O.debye_intensity(a, q)
    assert np.sqrt(((m - d) ** 2).sum() / (d ** 2).sum()) <= 1e-3


def test_multipole_basis_matches_debye_with_adaptive_grid():
    # spherical.rs:730-787: 8 atoms, AdaptiveGrid at 8 A, all six terms <= 1e-3
    a = atoms("eight")
    q = np.arange(1, 51) * 0.01
    sites, vol = O.adaptive_grid(a.xyz, spacing=8.0)
    m = O.multipole_solvent(a, q, 32, sites=sites, voxel_volume=vol, basis=True, exact=False)
    d = O.debye_solvent_basis(a, q, sites=sites, voxel_volume=vol)
    for t in range(6):
        err = math.sqrt(((m[t] - d[t]) ** 2).sum() / max((d[t] ** 2).sum(), 1e-30))
        assert err <= 1e-3, (t, err)


def test_multipole_solvent_intensity_is_combined_basis_without_sites():
    a = atoms("helix24")
    q = np.linspace(0.01, 0.4, 20)
    b = O.multipole_solvent(a, q, 30, c1=1.02, c2=0.5, basis=True)
    i = O.multipole_solvent(a, q, 30, c1=1.02, c2=0.5, basis=False)
    assert rel(O.combine_solvent_basis(b, q, 1.02, 0.5), i).max() < 1e-10


def test_scale_background_exact_recovery():
    # fit.rs:490-518 and 520-533
    x = np.linspace(1, 5, 8) ** 2
    y = 1.7 * x + 0.02
    s = np.full(8, 0.01)
    scale, bg, chi = O.solve_scale_background(x, y, s)
    assert abs(scale - 1.7) < 1e-8 and abs(bg - 0.02) < 1e-8 and chi < 1e-16
    scale, bg, chi = O.solve_scale_background([1, 2, 3], [2, 4, 6], [1, 2, 1], bg_mode=1)
    assert abs(scale - 2.0) < 1e-12 and bg == 0.0 and chi < 1e-24
    scale, bg, chi = O.solve_scale_background(x, y, s, scale_mode=1, scale_value=1.7)
    assert abs(bg - 0.02) < 1e-10


def test_cubic_hermite_and_sampling():
    x = np.linspace(0, 1, 11)
    y = 2 * x + 1
    t = np.linspace(0, 1, 101)
    assert np.allclose(O.cubic_hermite(x, y, t), 2 * t + 1, atol=1e-14)
    assert O.sampled_indices(10, 20) == list(range(10))
    idx = O.sampled_indices(101, 20)
    assert idx == [0, 20, 40, 60, 80, 100]
    assert O.sampled_indices(102, 20)[-1] == 101


def brute_project(v):
    # minimise ||w - v|| on the simplex by exhaustive support search
    best = None
    m = len(v)
    for mask in range(1, 1 << m):
        idx = [i for i in range(m) if mask >> i & 1]
        theta = (sum(v[i] for i in idx) - 1) / len(idx)
        w = np.zeros(m)
        for i in idx:
            w[i] = v[i] - theta
        if np.all(w >= -1e-15):
            d = np.sum((w - v) ** 2)
            if best is None or d < best[0]:
                best = (d, w)
    return best[1]


def test_project_simplex_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(20):
        v = rng.normal(size=5)
        assert np.allclose(O.project_simplex(v), brute_project(v), atol=1e-12)


def test_optimize_weights_properties():
    q = np.linspace(0.01, 0.5, 60)
    curves = np.stack([np.exp(-(q * r) ** 2 / 3) * 1e4 for r in (10, 15, 20, 30)])
    y = 1.3 * curves[2] + 2.0
    sigma = 0.01 * y
    w, (scale, bg, chi) = O.optimize_weights(curves, y, sigma)
    assert abs(w.sum() - 1) < 1e-12 and np.all(w >= 0)
    wu, (_, _, chi_u) = O.optimize_weights(curves, y, sigma, uniform=True)
    assert np.allclose(wu, 0.25) and chi_u > chi
    # one conformer: weight 1 and the two-step linear fit of ensemble.rs:383-434
    # (scale without background first, then the background at that scale)
    w1, (s1, b1, c1) = O.optimize_weights(curves[2:3], y, sigma)
    x, iw = curves[2], 1 / sigma ** 2
    s_want = np.sum(x * y * iw) / np.sum(x * x * iw)
    b_want = np.sum((y - s_want * x) * iw) / np.sum(iw)
    chi_want = np.sum(((s_want * x + b_want - y) / sigma) ** 2) / (len(q) - 1)
    assert w1[0] == 1.0 and abs(s1 - s_want) < 1e-12 * s_want and abs(b1 - b_want) < 1e-9
    assert abs(c1 - chi_want) < 1e-9 * chi_want


@pytest.mark.parametrize("solvent", [None, (1.03, 0.7, 3.0, 0.334)])
def test_oracle_gradient_matches_central_differences(solvent):
    rng = np.random.default_rng(8)
    xyz = rng.normal(size=(12, 3)) * 4
    el = rng.integers(1, 5, size=12)
    q = np.array([0.01, 0.1, 0.35, 0.8])
    jac = O.debye_grad_x(O.Atoms(xyz, el), q, solvent=solvent, jacobian=True)

    def intensity(x):
        a = O.Atoms(x, el)
        if solvent is None:
            return O.debye_intensity(a, q)
        c1, c2, t, rho = solvent
        return O.combine_solvent_basis(O.debye_solvent_basis(a, q, c1, c2, t, rho), q, c1, c2)

    h = 1e-4
    for i in (0, 5, 11):
        for c in range(3):
            up, dn = xyz.copy(), xyz.copy()
            up[i, c] += h
            dn[i, c] -= h
            fd = (intensity(up) - intensity(dn)) / (2 * h)
            scale = np.abs(jac[:, :, :]).max(axis=(1, 2))
            assert np.all(np.abs(fd - jac[:, i, c]) <= 1e-6 * scale + 1e-9), (i, c, fd, jac[:, i, c])
    w = rng.normal(size=q.size)
    g = O.debye_grad_x(O.Atoms(xyz, el), q, weights=w, solvent=solvent)
    assert np.allclose(g, np.einsum("k,kic->ic", w, jac), rtol=1e-12, atol=1e-12)
    # translation invariance: the gradients sum to zero
    assert np.abs(jac.sum(axis=1)).max() <= 1e-10 * np.abs(jac).max()


def _pr_brute(xyz, w, bins):
    """numpy brute force of coordinate_pr: the same FP64 operations per pair
    (so the same bins), summed in a different order."""
    n = len(xyz)
    il, ir = np.triu_indices(n, 1)
    dx = xyz[il, 0] - xyz[ir, 0]
    dy = xyz[il, 1] - xyz[ir, 1]
    dz = xyz[il, 2] - xyz[ir, 2]
    d2 = dx * dx + dy * dy + dz * dz
    dmax = np.sqrt(d2.max())
    step = dmax / bins
    b = np.minimum(np.floor(np.sqrt(d2) / step), bins - 1).astype(int)
    return dmax, np.bincount(b, weights=w[il] * w[ir], minlength=bins), (il, ir, b)


def test_coordinate_pr_reference_fixture():
    # metrics.rs:651-684: two GLY carbons 4 A apart
    atoms = O.Atoms([[0, 0, 0], [4, 0, 0]], [1, 1],
                    group=[O.lib().or_united_atom_group(b"GLY", b"C1")] * 2)
    rg, dmax, hist = O.coordinate_features(atoms, 160)
    assert abs(dmax - 4.0) < 1e-12 and rg > 1.0
    r = (np.arange(160) + 0.5) * dmax / 160
    norm = O.normalize_distribution(r, hist)
    assert norm is not None and len(norm[1]) > 0
    # the single pair lands in the last bin (floor(4 / step) = 160 -> 159)
    assert np.count_nonzero(hist) == 1 and hist[-1] > 0


@pytest.mark.parametrize("n,bins", [(300, 160), (517, 16), (64, 300)])
def test_coordinate_pr_matches_brute_force(n, bins):
    rng = np.random.default_rng(n + bins)
    xyz = rng.normal(scale=9.0, size=(n, 3))
    el = rng.choice([1, 2, 3, 4], size=n)
    atoms = O.Atoms(xyz, el)
    rg, dmax, hist = O.coordinate_features(atoms, bins)
    w = np.array([max(abs(O.lib().or_form_factor(int(e), -1, 1.0, 1, 0.0)), 1e-6) for e in el])
    dmax_b, hist_b, _ = _pr_brute(xyz, w, max(bins, 16))
    assert dmax == dmax_b
    assert np.allclose(hist, hist_b, rtol=1e-12, atol=0)
    assert abs(hist.sum() - (w.sum() ** 2 - (w * w).sum()) / 2) <= 1e-12 * hist.sum()
    c = (xyz * w[:, None]).sum(axis=0) / w.sum()
    assert abs(rg - np.sqrt((w * ((xyz - c) ** 2).sum(axis=1)).sum() / w.sum())) < 1e-12 * rg


def test_coordinate_pr_degenerate():
    rg, dmax, hist = O.coordinate_features(O.Atoms([[1.0, 2.0, 3.0]], [1]), 160)
    # (the weighted centre x w / w rounds, so Rg is ~1e-16, as in the reference)
    assert dmax == 0.0 and hist is None and rg < 1e-12


# ------------------------------------------------------------------ grids --
# The oracle's grids against an independent brute force (tests/grid_cases.py)
# and the reference's grid tests (solvent.rs:688-771).

import grid_cases as G  # noqa: E402


@pytest.mark.parametrize("spacing", [1.0, 3.0, 4.0, 8.0])
def test_adaptive_grid_matches_brute_force(spacing):
    xyz, _ = G.globule(250, 11.0, seed=4)
    ref, vol = O.adaptive_grid(xyz, spacing=spacing)
    assert len(ref) > 0 and np.array_equal(ref, G.adaptive_grid(xyz, spacing=spacing))
    assert vol == (2.0 * (spacing / 2) ** 3 if spacing > 4.0 else spacing ** 3)


@pytest.mark.parametrize("spacing,probe,shell", [(1.0, 1.4, 3.0), (1.5, 1.2, 2.0),
                                                 (0.8, 0.0, 1.0)])
def test_exterior_grid_matches_brute_force(spacing, probe, shell):
    xyz, el = G.globule(250, 11.0, seed=5)
    ref, vol = O.exterior_grid(xyz, el, spacing=spacing, probe=probe, shell=shell)
    assert len(ref) > 0 and vol == spacing ** 3
    assert np.array_equal(ref, G.exterior_grid(xyz, el, spacing, probe, shell))


def test_exterior_grid_excludes_enclosed_void_and_adaptive_keeps_it():
    xyz, el = G.hollow_sphere(9.0)
    ext, _ = O.exterior_grid(xyz, el)
    assert np.array_equal(ext, G.exterior_grid(xyz, el))
    r = np.linalg.norm(ext, axis=1)
    assert r.min() > 9.0 and len(ext) > 0  # no site inside the sealed shell
    ada, _ = O.adaptive_grid(xyz, spacing=1.0)
    assert np.array_equal(ada, G.adaptive_grid(xyz, spacing=1.0))
    assert np.linalg.norm(ada, axis=1).min() < 6.0 + 1e-9  # voids are included


def test_grids_reference_properties():
    one = np.array([[-1.0, 0.0, 0.0]])
    ext, vol = O.exterior_grid(one, [1])  # solvent.rs:688-705
    d = np.linalg.norm(ext - one[0], axis=1)
    assert len(d) > 0 and vol == 1.0
    assert np.all(d > 1.70 + 1.4) and np.all(d <= 1.70 + 1.4 + 3.0 + 1e-12)
    two = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    a, vol = O.adaptive_grid(two)  # solvent.rs:716-726
    assert len(a) > 0 and vol == 64.0
    assert O.adaptive_grid(two, spacing=8.0)[1] == 2.0 * 4.0 ** 3  # solvent.rs:728-738
    shift = np.array([3.17, -8.4, 2.6])  # solvent.rs:740-771
    for build in (lambda x: O.adaptive_grid(x)[0], lambda x: O.exterior_grid(x, [1, 3])[0]):
        g0, g1 = build(two), build(two + shift)
        assert g0.shape == g1.shape and np.abs(g1 - g0 - shift).max() < 1e-12

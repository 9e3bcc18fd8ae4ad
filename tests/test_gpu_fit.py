"""GPU tests of the fitting and ensemble layer through the C ABI (fit.rs,
analysis.rs, ensemble.rs tests and behaviours)."""
import numpy as np
import pytest

import crabsaxs_amd as csx
from test_fit_host import lcg_noise

pytestmark = pytest.mark.gpu


def two_atoms():
    return csx.Structure([csx.Atom(pos=(0, 0, 0), element=csx.Element.C, residue_name="TST",
                                   atom_name="C1"),
                          csx.Atom(pos=(2, 0, 0), element=csx.Element.O, residue_name="TST",
                                   atom_name="O1")])


def test_recovers_scale_and_background(gpu_ctx):
    # fit.rs:489-518
    s = two_atoms()
    q = np.arange(1, 9) * 0.03
    p = csx.SolventParams(c1=1.05, c2=0.4)
    raw = csx.multipole_solvent_intensity(s, q, csx.MultipoleParams(8, csx.HydrogenMode.Explicit), p)
    e = csx.ExperimentalCurve(q, 1.7 * raw + 0.02, np.full(8, 0.01))
    scale, bg, chi2 = csx.solve_scale_background(raw, e)
    assert abs(scale - 1.7) < 1e-8 and abs(bg - 0.02) < 1e-8 and chi2 < 1e-16


@pytest.mark.parametrize("optimizer", [csx.Optimizer.NelderMead, csx.Optimizer.Gradient])
def test_recovers_solvent_parameters_with_deterministic_noise(gpu_ctx, optimizer):
    # fit.rs:536-599, end to end on the GPU multipole basis
    s = two_atoms()
    q = np.arange(1, 81) * 0.006
    expected = csx.SolventParams(c1=1.02, c2=0.55)
    raw = csx.multipole_solvent_intensity(s, q, csx.MultipoleParams(8, csx.HydrogenMode.Explicit),
                                          expected)
    e = csx.ExperimentalCurve(q, 1.4e-4 * raw + 0.03 + 0.0001 * lcg_noise(q.size),
                              np.full(q.size, 0.002))
    opts = csx.FitOptions(method=csx.ScatteringMethod.Multipole,
                          hydrogen_mode=csx.HydrogenMode.Explicit, max_iterations=100,
                          q_sampling_stride=1, optimizer=optimizer)
    r = csx.fit_to_experimental_with_options(s, e, csx.FitParams(1.4e-4, 0.03, 1.02, 0.55), opts)
    assert r.chi2 < 0.01, r
    if optimizer == csx.Optimizer.NelderMead:
        assert abs(r.params.c1 - 1.02) < 0.02, r
        assert abs(r.params.c2 - 0.55) < 0.08, r
        assert abs(r.params.scale - 1.4e-4) < 0.4e-4, r
        assert abs(r.params.background - 0.03) < 0.01, r


def test_sampled_basis_matches_full_basis_fit(gpu_ctx):
    # fit.rs:601-660
    s = two_atoms()
    q = np.arange(1, 102) * 0.005
    raw = csx.multipole_solvent_intensity(s, q, csx.MultipoleParams(8, csx.HydrogenMode.Explicit),
                                          csx.SolventParams(c1=1.03, c2=0.3))
    e = csx.ExperimentalCurve(q, 0.8 * raw + 0.01, np.full(q.size, 0.01))
    common = dict(method=csx.ScatteringMethod.Multipole, hydrogen_mode=csx.HydrogenMode.Explicit,
                  max_iterations=64)
    full = csx.fit_to_experimental_with_options(s, e, csx.FitParams(),
                                                csx.FitOptions(q_sampling_stride=1, **common))
    sampled = csx.fit_to_experimental_with_options(s, e, csx.FitParams(), csx.FitOptions(**common))
    rel = np.sqrt(np.sum((full.fitted_curve - sampled.fitted_curve) ** 2)
                  / np.sum(full.fitted_curve ** 2))
    assert rel < 1e-3, rel


def test_debye_fit_and_scorer_on_globule(gpu_ctx):
    s = csx.generate_globule(1200, 16.0, 1.0, seed=31)
    q = np.linspace(0.01, 0.5, 150)
    truth = csx.SolventParams(c1=1.02, c2=1.0)
    y = 2.3 * csx.solvent_corrected_intensity(s, q, truth) + 1e-3
    e = csx.ExperimentalCurve(q, y, 0.01 * y)
    for opt in (csx.Optimizer.NelderMead, csx.Optimizer.Gradient):
        # 150 points: below the density where stride-20 Hermite sampling of a
        # 16 A globule's basis is accurate, so fit on the full grid.
        r = csx.fit_to_experimental_with_options(
            s, e, csx.FitParams(), csx.FitOptions(method=csx.ScatteringMethod.Debye,
                                                  optimizer=opt, max_iterations=200,
                                                  q_sampling_stride=1))
        assert abs(r.params.c1 - 1.02) < 5e-3 and abs(r.params.c2 - 1.0) < 0.05, r
        assert abs(r.params.scale - 2.3) < 0.05 and r.chi2 < 1e-3, r
    # The scorer's default method (Auto) resolves to the multipole path at
    # 1,200 atoms, whose analytic shell Gaussian exp(-q^2 r_s^2 / 4)
    # (spherical.rs:205-231) differs from the Debye path's Fraser Gaussian
    # (solvent.rs:454-477) by 1-2% of the shell terms: data made with Debye
    # is scored with Debye.
    score = csx.SaxsScorer(e, csx.ScoreOptions(fit=csx.FitOptions(method=csx.ScatteringMethod.Debye),
                                               quality=csx.Quality.Accurate)).score(s)
    assert score.points == 150 and score.reduced_chi2 < 1e-2 and score.q_unit == "angstrom^-1"
    assert abs(score.chi2 - score.reduced_chi2 * 149) < 1e-9 * max(score.chi2, 1)
    fitter = csx.SaxsFitter(csx.FitOptions(method=csx.ScatteringMethod.Debye, q_sampling_stride=1))
    assert fitter.fit(s, e).chi2 < 1e-2


def test_ensemble_recovers_mixture(gpu_ctx):
    conformers = [csx.generate_globule(800 + 100 * k, 13.0 + k, 1.0, seed=40 + k)
                  for k in range(4)]
    q = np.linspace(0.01, 0.4, 80)
    p = csx.SolventParams(c1=1.0, c2=1.0)
    curves = np.stack([csx.solvent_corrected_intensity(c, q, p) for c in conformers])
    y = 1.5 * (0.6 * curves[1] + 0.4 * curves[3]) + 2.0
    e = csx.ExperimentalCurve(q, y, 0.005 * y)
    opts = csx.EnsembleOptions(fit=csx.FitOptions(method=csx.ScatteringMethod.Debye),
                               min_weight=0.05, bootstrap=3)
    fitter = csx.EnsembleFitter(conformers, ["a", "b", "c", "d"], e, opts)
    r = fitter.fit()
    assert abs(sum(r.weights) - 1) < 1e-12 and np.all(np.asarray(r.weights) >= 0)
    assert set(r.conformers) >= {"b", "d"}
    assert len(r.bootstrap_reduced_chi2) == 3 and len(r.weight_intervals) == len(r.conformers)
    # the best candidate fits at least as well as the grid point that made the data
    assert 0.95 <= r.params.c1 <= 1.05 and -2.0 <= r.params.c2 <= 4.0
    b0 = fitter.bases()
    sel = [["a", "b", "c", "d"].index(n) for n in r.conformers]
    truth = np.stack([csx.combine_solvent_basis(b0[i], q, 1.0, 1.0) for i in sel])
    _, _, _, chi_truth, _ = csx.optimize_weights(truth, e, opts.fit)
    assert r.reduced_chi2 <= chi_truth * (1 + 1e-9)
    # re-contracted bases equal directly computed candidate profiles
    b = fitter.bases()
    direct = csx.solvent_corrected_intensity(conformers[2], q, csx.SolventParams(c1=0.975, c2=-0.5))
    recon = csx.combine_solvent_basis(b[2], q, 0.975, -0.5)
    assert np.max(np.abs(recon - direct) / np.abs(direct)) < 1e-12


def _fit_core(b, e, opts, uniform, reg, solver, ctx=None):
    import ctypes as C
    from crabsaxs_amd import _lib
    from crabsaxs_amd._lib import FitParamsC, dptr
    m, n = b.shape[0], len(e)
    w, chi, fitted, p = np.zeros(m), C.c_double(), np.zeros(n), FitParamsC()
    o = opts.to_c()
    args = (dptr(np.ascontiguousarray(b)), m, dptr(e.q), dptr(e.intensity), dptr(e.sigma), n,
            C.byref(o), int(uniform), reg, int(solver), dptr(w), C.byref(p), C.byref(chi),
            dptr(fitted))
    L = _lib.load()
    st = L.csx_ensemble_fit_core_device(ctx.ptr, *args) if ctx else L.csx_ensemble_fit_core(*args)
    assert st == 0
    return w, (p.scale, p.background, p.c1, p.c2), chi.value, fitted


@pytest.mark.parametrize("m,n,uniform,reg,bg", [(37, 90, False, 0.0, "Fit"), (512, 200, False, 0.0, "Fit"),
                                               (5, 64, False, 0.3, "FixedZero"), (8, 50, True, 0.0, "Fit"),
                                               (300, 120, False, 0.0, "Fixed")])
def test_device_fit_core_is_bitwise_the_host_fit_core(gpu_ctx, m, n, uniform, reg, bg):
    """The 25-candidate weight fit on the GPU against the host restatement
    (which is bitwise the oracle's optimize_weights): identical bits."""
    rng = np.random.default_rng(m * 1000 + n)
    q = np.linspace(0.01, 0.5, n)
    base = np.exp(-np.outer(rng.uniform(5, 30, m), q) ** 2 / 3.0)
    b = np.zeros((m, 6, n))
    b[:, 0] = 1e4 * base
    b[:, 1] = 2e3 * base * np.exp(-q ** 2)
    b[:, 2] = 50 * base
    b[:, 3] = 4e3 * base
    b[:, 4] = 300 * base
    b[:, 5] = 90 * base
    y = 1.7 * (b[:m // 2 + 1, 0].mean(axis=0)) + 3.0
    e = csx.ExperimentalCurve(q, y + 0.01 * y * rng.normal(size=n), 0.01 * y + 0.1)
    mode = {"Fit": csx.BackgroundMode.Fit, "FixedZero": csx.BackgroundMode.FixedZero,
            "Fixed": csx.BackgroundMode.Fixed(2.5)}[bg]
    opts = csx.FitOptions(background_mode=mode)
    host = _fit_core(b, e, opts, uniform, reg, csx.WeightSolver.ProjectedGradient)
    dev = _fit_core(b, e, opts, uniform, reg, csx.WeightSolver.ProjectedGradient, gpu_ctx)
    assert np.array_equal(host[0], dev[0])
    assert host[1] == dev[1] and host[2] == dev[2]
    assert np.array_equal(host[3], dev[3])

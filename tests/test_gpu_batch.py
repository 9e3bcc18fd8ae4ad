"""Batched conformer profiles, the persistent launch and the autotuner."""
import numpy as np
import pytest

import crabsaxs_amd as csx
from crabsaxs_amd import autotune

pytestmark = pytest.mark.gpu


def conformers():
    return [csx.generate_globule(n, 10.0 + n / 400, 1.0, seed=60 + n) for n in (300, 1700, 2500)] + \
           [csx.generate_domain_chain(3200, 3, seed=9)]


def test_batch_matches_single_bitwise(gpu_ctx):
    cs = conformers()
    q = np.linspace(0.0, 0.8, 90)
    p = csx.SolventParams(c1=1.01, c2=0.5)
    got = csx.batch_debye_solvent_basis(cs, q, p)
    got_i = csx.batch_debye_intensity(cs, q)
    for k, c in enumerate(cs):
        assert np.array_equal(got[k], csx.debye_solvent_basis(c, q, p))
        assert np.array_equal(got_i[k], csx.debye_intensity(c, q))
    ref = csx.debye_intensity(cs[1], q, precision=csx.Precision.FP64_PLAIN)
    assert np.max(np.abs(got_i[1] - ref) / ref) < 1e-5


def test_persistent_launch_is_bitwise_identical(gpu_ctx):
    cs = conformers()
    q = np.linspace(0.0, 1.0, 130)
    base = csx.batch_debye_intensity(cs, q)
    try:
        for wgs in (0, 1, 4, 8):
            gpu_ctx.set_persistence(wgs)
            assert np.array_equal(csx.batch_debye_intensity(cs, q), base)
    finally:
        gpu_ctx.set_persistence()


def test_every_plan_group_of_a_batch_is_timed(gpu_ctx):
    """A 300-atom set (4.5e6 pair.q, the direct sweep) and a 4,000-atom set
    (the recurrence) have different plans, so the batch launches the pair
    kernel twice and waits after each; the kernel timer counts both."""
    cs = [csx.generate_globule(300, 9.0, 1.0, seed=5), csx.generate_globule(4000, 20.0, 1.0, seed=6)]
    q = np.linspace(0.0, 0.5, 100)
    gpu_ctx.reset_timing()
    gpu_ctx.set_timing(True)
    try:
        got = csx.batch_debye_intensity(cs, q)
    finally:
        gpu_ctx.set_timing(False)
    n, ms = gpu_ctx.kernel_time("debye_pairs_f32")
    assert n == 2 and ms > 0
    for k, c in enumerate(cs):
        assert np.array_equal(got[k], csx.debye_intensity(c, q))


def test_autotuner_picks_a_measured_shape(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", tmp_path / "tuning.json")
    s = csx.generate_globule(4000, 20.0, 1.0, seed=3)
    q = np.linspace(0.0, 1.0, 128)
    want = csx.debye_intensity(s, q, precision=csx.Precision.FP64_PLAIN)
    try:
        best = autotune.autotune(s, q, gpu_ctx, candidates={"q_tile": (32, 64),
                                                            "j_chunk": (256, 512),
                                                            "persist": (0, 4)})
        assert best["tried"] == 8 and best["kernel_ms"] > 0 and not best["cached"]
        got = csx.debye_intensity(s, q)
        assert np.max(np.abs(got - want) / want) < 1e-5
        again = autotune.autotune(s, q, gpu_ctx)
        assert again["cached"] and again["q_tile"] == best["q_tile"]
    finally:
        gpu_ctx.set_tuning()
        gpu_ctx.set_persistence()


def test_ensemble_fitter_uses_batched_bases(gpu_ctx):
    cs = conformers()
    q = np.linspace(0.01, 0.4, 70)
    p = csx.SolventParams(c1=1.0, c2=1.0)
    curves = csx.batch_debye_solvent_basis(cs, q, p)
    mix = 0.5 * csx.combine_solvent_basis(curves[0], q, 1.0, 1.0) + \
        0.5 * csx.combine_solvent_basis(curves[2], q, 1.0, 1.0)
    e = csx.ExperimentalCurve(q, 2.0 * mix + 1.0, 0.01 * mix + 1.0)
    opts = csx.EnsembleOptions(fit=csx.FitOptions(method=csx.ScatteringMethod.Debye))
    fitter = csx.EnsembleFitter(cs, ["a", "b", "c", "d"], e, opts)
    b = fitter.bases()
    for k, c in enumerate(cs):
        assert np.array_equal(b[k], csx.debye_solvent_basis(c, q, csx.SolventParams()))
    r = fitter.fit()
    assert abs(np.sum(r.weights) - 1) < 1e-12


def test_every_q_tile_gives_identical_sums(gpu_ctx):
    """Each q value's sum runs over the same pairs in the same order whatever
    the tile, so every instantiated tile (and the automatic choice) agrees
    bitwise.  The 2,300-atom globule has 41 work items (four element types,
    256-atom i blocks, 512-atom j chunks): its time is set by the longest
    item, so the automatic choice is a narrow tile (10 wide: 200 q in 20
    sweeps, the least modelled time), launched one workgroup per task (820
    tasks, fewer than two per persistent workgroup)."""
    s = csx.generate_globule(2300, 15.0, 1.0, seed=12)
    q = np.linspace(0.005, 0.5, 200)
    try:
        gpu_ctx.set_tuning(64, 2, 512)
        base = csx.debye_intensity(s, q)
        for qt in (0, 8, 10, 16, 20, 24, 25, 32, 40, 48, 50, 56, 63):
            gpu_ctx.set_tuning(qt, 2, 512)
            assert np.array_equal(csx.debye_intensity(s, q), base), qt
        assert gpu_ctx.last_launch()["q_tile"] == 63
        gpu_ctx.set_tuning(0, 2, 512)
        csx.debye_intensity(s, q)
        assert gpu_ctx.last_launch() == {"q_tile": 10, "sin_mode": 2, "persistent_workgroups": 0}
        gpu_ctx.set_tuning(40, 1, 512)  # no 40-wide tile for the reduced sine: automatic
        csx.debye_intensity(s, q)
        assert gpu_ctx.last_launch()["q_tile"] in (8, 16, 32, 64)
    finally:
        gpu_ctx.set_tuning()


def test_automatic_shape_for_a_large_structure(gpu_ctx):
    """20,000 atoms (1,671 work items) are bound by total work.  On a uniform
    grid they take the sine recurrence: for 500 q a 100-wide tile (no
    padding, the per-pair start amortised over 100 steps); for 200 q the
    0.005 A^-1 step turns the largest pair's phase by 0.058 revolutions, so
    tiles are held to 70 q (4 revolutions) and the choice is 50.  The direct
    hardware sine (mode 2) takes the tile with the least padded work (50 for
    200 q, four sweeps without padding; 63 for 500 q).  Both launch
    persistent, 4 workgroups per CU, as there are more than two tasks per
    workgroup."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    cus = gpu_ctx.cu_count()
    gpu_ctx.set_persistence()
    try:
        for mode, tiles in ((3, (50, 100)), (2, (50, 63))):
            gpu_ctx.set_tuning(0, mode, 0)
            for nq, tile in zip((200, 500), tiles):
                csx.debye_intensity(s, np.linspace(0.0, 1.0, nq))
                assert gpu_ctx.last_launch() == {"q_tile": tile, "sin_mode": mode,
                                                 "persistent_workgroups": 4 * cus}, (mode, nq)
    finally:
        gpu_ctx.set_tuning()


def test_ensemble_nnls_recovers_weights(gpu_ctx):
    cs = conformers() + [csx.generate_globule(2100, 13.0, 1.4, seed=77)]
    q = np.linspace(0.01, 0.45, 90)
    curves = csx.batch_debye_solvent_basis(cs, q, csx.SolventParams())
    prof = np.stack([csx.combine_solvent_basis(b, q, 1.0, 1.0) for b in curves])
    w_true = np.array([0.0, 0.3, 0.0, 0.0, 0.7])
    y = 2.0 * w_true @ prof + 5.0
    e = csx.ExperimentalCurve(q, y, 0.01 * y)
    opts = csx.EnsembleOptions(fit=csx.FitOptions(method=csx.ScatteringMethod.Debye),
                               weight_solver=csx.WeightSolver.NNLS)
    r = csx.EnsembleFitter(cs, list("abcde"), e, opts).fit()
    w = dict(zip(r.conformers, r.weights))
    assert abs(w.get("b", 0) - 0.3) < 1e-6 and abs(w.get("e", 0) - 0.7) < 1e-6, w
    assert abs(r.params.scale - 2.0) < 1e-6 and r.reduced_chi2 < 1e-10


def test_batch_autotuner_applies_a_measured_shape(gpu_ctx):
    cs = conformers()
    q = np.linspace(0.005, 0.5, 60)
    try:
        best = autotune.autotune_batch(cs, q, gpu_ctx, sample=3,
                                       candidates={"q_tile": (0, 16), "j_chunk": (256,),
                                                   "persist": (0, 2)})
        assert best["tried"] == 4 and best["sample"] == 3 and best["kernel_ms"] > 0
        tuned = csx.batch_debye_intensity(cs, q)
        gpu_ctx.set_tuning(best["q_tile"], best["sin_mode"], 256)
        gpu_ctx.set_persistence(0)
        # with the same j chunk, sweep and q tile both launch forms sum the
        # same pairs in the same order
        assert np.array_equal(tuned, csx.batch_debye_intensity(cs, q))
    finally:
        gpu_ctx.set_tuning()
        gpu_ctx.set_persistence()


@pytest.mark.parametrize("model", [csx.HydrationModel.AdaptiveGrid, csx.HydrationModel.Grid])
def test_batch_with_grid_hydration_matches_single_bitwise(gpu_ctx, model):
    """Grid hydration sites form an extra point type per structure; the batch
    concatenates them with the atoms and keeps every bin separate."""
    cs = conformers()[:3]
    q = np.linspace(0.0, 0.6, 50)
    p = csx.SolventParams(c1=1.01, c2=0.7, hydration_model=model,
                          hydration_grid_spacing_angstrom=4.0 if model == csx.HydrationModel.AdaptiveGrid
                          else 2.0)
    got = csx.batch_debye_solvent_basis(cs, q, p)
    for k, c in enumerate(cs):
        assert np.array_equal(got[k], csx.debye_solvent_basis(c, q, p))

"""GPU parity of the Debye path (through the C ABI) against the FP64 oracle.

Bars: FP64_PLAIN within 1e-12 relative of the oracle; FP32_COMPENSATED within
1e-5 maximum relative error per q (the north-star gate), both against the
oracle at oracle-feasible sizes and against the FP64 GPU kernel at the full
20,000-atom, 500-q precision-study size.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import crabsaxs_amd as csx
import oracle_py as O

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "debye_golden.json").read_text())
F32 = csx.Precision.FP32_COMPENSATED
F64 = csx.Precision.FP64_PLAIN


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def structure(name):
    s = GOLD["structures"][name]
    return csx.Structure.from_arrays(s["xyz"], s["element"])


def oracle_atoms(s):
    return O.Atoms(s.xyz, s.element, s.group, s.occupancy)


@pytest.mark.parametrize("name", sorted(GOLD["structures"]))
def test_debye_golden_vectors(gpu_ctx, name):
    s = structure(name)
    for tag in ("low", "high"):
        q = np.array(GOLD[f"q_{tag}"])
        want = np.array(GOLD["structures"][name][f"debye_{tag}"])
        got64 = csx.debye_intensity(s, q, precision=F64)
        got32 = csx.debye_intensity(s, q, precision=F32)
        print(name, tag, rel(got64, want).max(), rel(got32, want).max(), flush=True)
        assert rel(got64, want).max() < 1e-12
        assert rel(got32, want).max() < 1e-5
        basis = np.array(GOLD["structures"][name][f"basis_{tag}"])
        b64 = csx.debye_solvent_basis(s, q, csx.SolventParams(), precision=F64)
        b32 = csx.debye_solvent_basis(s, q, csx.SolventParams(), precision=F32)
        scale = np.abs(basis).max(axis=1, keepdims=True)
        assert (np.abs(b64 - basis) / scale).max() < 1e-12
        assert (np.abs(b32 - basis) / scale).max() < 1e-5


def test_two_atom_closed_form(gpu_ctx):
    # scattering.rs:250-257
    s = csx.Structure([csx.Atom(pos=(0, 0, 0), residue_name="TST"),
                       csx.Atom(pos=(2, 0, 0), residue_name="TST", atom_name="C2")])
    f = csx.form_factor_for_atom(csx.Atom(pos=(0, 0, 0), residue_name="TST"), True, 0.2)
    want = 2 * f * f * (1 + np.sin(0.4) / 0.4)
    for p in (F64, F32):
        got = csx.debye_intensity(s, [0.2], precision=p)[0]
        assert abs(got - want) < 1e-10 * want if p == F64 else abs(got - want) < 1e-6 * want


@pytest.mark.parametrize("seed", [11, 12])
def test_globule_1200_matches_oracle(gpu_ctx, seed):
    """1,200-atom globule (R = 16 A), q = 0.005..0.5 (100 points)."""
    s = csx.generate_globule(1200, 16.0, 1.0, seed=seed)
    q = np.arange(1, 101) * 0.005
    want = O.debye_intensity(oracle_atoms(s), q)
    got64 = csx.debye_intensity(s, q, precision=F64)
    got32 = csx.debye_intensity(s, q, precision=F32)
    print("1200 f64", rel(got64, want).max(), "f32", rel(got32, want).max(), flush=True)
    assert rel(got64, want).max() < 1e-12
    assert rel(got32, want).max() < 1e-5


@pytest.mark.parametrize("model", [csx.HydrationModel.Analytic, csx.HydrationModel.AdaptiveGrid,
                                   csx.HydrationModel.Grid])
def test_solvent_basis_matches_oracle(gpu_ctx, model):
    s = csx.generate_globule(400, 12.0, 1.0, seed=4)
    q = np.linspace(0.0, 0.5, 41)
    p = csx.SolventParams(c1=1.03, c2=0.4, hydration_model=model,
                          hydration_grid_spacing_angstrom=4.0)
    sites, vol = None, 0.0
    if model == csx.HydrationModel.AdaptiveGrid:
        sites, vol = O.adaptive_grid(s.xyz, spacing=4.0)
    elif model == csx.HydrationModel.Grid:
        sites, vol = O.exterior_grid(s.xyz, s.element, spacing=1.0, shell=3.0)
    want = O.debye_solvent_basis(oracle_atoms(s), q, c1=1.03, c2=0.4, sites=sites,
                                 voxel_volume=vol)
    scale = np.abs(want).max(axis=1, keepdims=True)
    # FP64 differs from the oracle only by summation order and the last bit of
    # sin: ~1e-16 per term, over 8e4 atom pairs and up to 1e8 site pairs
    # (1 A exterior grid) with cancellation between terms.
    for prec, tol in ((F64, 1e-9), (F32, 1e-5)):
        got = csx.debye_solvent_basis(s, q, p, precision=prec)
        err = (np.abs(got - want) / scale).max()
        print(model.name, prec.name, err, flush=True)
        assert err < tol
        inten = csx.solvent_corrected_intensity(s, q, p, precision=prec)
        ref = O.combine_solvent_basis(want, q, 1.03, 0.4)
        assert rel(inten, ref).max() < tol * 10


def test_pair_tables_recontract(gpu_ctx):
    """S_ab tables re-contract to the vacuum intensity with the type amplitudes."""
    s = csx.generate_globule(600, 12.0, 1.0, seed=9)
    q = np.linspace(0, 0.8, 33)
    S, self_terms = csx.debye_pair_tables(s, q)
    types = sorted(set(int(e) for e in s.element))
    f = np.array([[csx.form_factor_for_atom(csx.Atom(pos=(0, 0, 0), element=csx.Element(t)),
                                            True, qq) for qq in q] for t in types])
    T = len(types)
    inten = (f ** 2 * self_terms[:, None]).sum(axis=0)
    b = 0
    for a in range(T):
        for c in range(a, T):
            inten = inten + 2 * f[a] * f[c] * S[b]
            b += 1
    want = csx.debye_intensity(s, q, precision=F64)
    assert rel(inten, want).max() < 1e-5
    # q = 0 row is exact: sum_{i<j} occ_i occ_j per bin
    counts = np.array([np.sum(s.element == t) for t in types], dtype=float)
    b = 0
    for a in range(T):
        for c in range(a, T):
            exp0 = counts[a] * (counts[a] - 1) / 2 if a == c else counts[a] * counts[c]
            assert S[b, 0] == exp0
            b += 1


def test_edge_cases(gpu_ctx):
    empty = csx.Structure.from_arrays(np.zeros((0, 3)), np.zeros(0, dtype=np.uint8))
    q = np.array([0.0, 0.1, 0.2])
    for p in (F64, F32):
        assert np.array_equal(csx.debye_intensity(empty, q, precision=p), np.zeros(3))
    one = csx.Structure.from_arrays([[1.0, 2.0, 3.0]], [1])
    f = [csx.form_factor_for_atom(csx.Atom(pos=(0, 0, 0)), True, qq) for qq in q]
    for p in (F64, F32):
        assert np.allclose(csx.debye_intensity(one, q, precision=p), np.square(f), rtol=1e-12)
    # coincident atoms: sinc(0) = 1 for the pair
    dup = csx.Structure.from_arrays([[0.5, 0.5, 0.5]] * 3 + [[4.0, 0.0, 0.0]], [1, 1, 3, 2])
    qq = np.array([0.0, 1e-9, 0.05, 0.7])
    want = O.debye_intensity(oracle_atoms(dup), qq)
    for p, tol in ((F64, 1e-12), (F32, 1e-6)):
        assert rel(csx.debye_intensity(dup, qq, precision=p), want).max() < tol
    # fractional occupancy and hydrogens in both modes
    atoms = [csx.Atom(pos=(0, 0, 0), element=csx.Element.C, residue_name="ALA", atom_name="CB",
                      occupancy=0.5),
             csx.Atom(pos=(1.1, 0, 0), element=csx.Element.H, residue_name="ALA",
                      atom_name="HB1"),
             csx.Atom(pos=(0, 1.5, 0.3), element=csx.Element.N, residue_name="ALA",
                      atom_name="N", occupancy=0.8)]
    s = csx.Structure(atoms)
    qv = np.linspace(0, 0.9, 19)
    for imp in (True, False):
        want = O.debye_intensity(oracle_atoms(s), qv, implicit_h=imp)
        for p, tol in ((F64, 1e-12), (F32, 1e-6)):
            got = csx.debye_intensity_with_hydrogens(s, qv, imp, precision=p)
            assert rel(got, want).max() < tol


def test_deterministic(gpu_ctx):
    s = csx.generate_globule(3000, 20.0, 1.0, seed=5)
    q = np.linspace(0, 1, 64)
    a = csx.debye_intensity(s, q)
    b = csx.debye_intensity(s, q)
    assert np.array_equal(a, b)


def expected_auto_chunk(n, cus):
    """The automatic j chunk: the largest of 512, 256, 128, 64 that still
    gives n^2 / (2 * 256 * chunk) >= one item per CU, else 32."""
    for c in (512, 256, 128, 64):
        if n * n / (2.0 * 256 * c) >= cus:
            return c
    return 32


@pytest.mark.parametrize("n", [1200, 2300, 5000, 20000])
def test_automatic_j_chunk(gpu_ctx, n):
    """Small structures take short j sweeps so that their work items spread
    over the CUs; the chunk is a function of the point count and the CU
    count, the results are bitwise reproducible (also on a fresh handle and
    with the same chunk set explicitly) and within 1e-5 of FP64."""
    s = csx.generate_globule(n, 35.0 * (n / 20000) ** (1 / 3), 1.0, seed=9)
    q = np.linspace(0.0, 1.0, 100)
    a = csx.debye_intensity(s, q)
    info = gpu_ctx.last_pair_items()
    assert info["j_chunk"] == expected_auto_chunk(n, gpu_ctx.cu_count()), info
    b = csx.debye_intensity(csx.Structure.from_arrays(s.xyz, s.element), q)
    assert np.array_equal(a, b)
    gpu_ctx.set_tuning(0, 3, info["j_chunk"])
    try:
        c = csx.debye_intensity(csx.Structure.from_arrays(s.xyz, s.element), q)
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(a, c)
    if n <= 5000:
        want = csx.debye_intensity(s, q, precision=F64)
        print("auto chunk", n, info, rel(a, want).max(), flush=True)
        assert rel(a, want).max() < 1e-5


@pytest.mark.parametrize("q_tile,sin_mode,j_chunk", [(8, 0, 256), (32, 0, 4096), (16, 1, 2048),
                                                     (64, 2, 512), (64, 0, 512), (8, 2, 64),
                                                     (24, 2, 128), (10, 2, 32), (20, 2, 64),
                                                     (25, 2, 256)])
def test_tuning_variants_agree(gpu_ctx, q_tile, sin_mode, j_chunk):
    s = csx.generate_globule(1500, 18.0, 1.0, seed=6)
    q = np.linspace(0, 1.0, 77)
    want = csx.debye_intensity(s, q, precision=F64)
    gpu_ctx.set_tuning(q_tile, sin_mode, j_chunk)
    try:
        s2 = csx.Structure.from_arrays(s.xyz, s.element)
        got = csx.debye_intensity(s2, q)
    finally:
        gpu_ctx.set_tuning()
    err = rel(got, want).max()
    print("tuning", q_tile, sin_mode, j_chunk, err, flush=True)
    assert err < 1e-5


def test_precision_study_20k_against_fp64_gpu(gpu_ctx):
    """20,000 atoms, R = 35 A, q = linspace(0, 1, 500): FP32 vs FP64 GPU Debye."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    q = np.linspace(0.0, 1.0, 500)
    ref = csx.debye_intensity(s, q, precision=F64)
    got = csx.debye_intensity(s, q, precision=F32)
    err = rel(got, ref)
    print("20k vacuum max", err.max(), "mean", err.mean(), flush=True)
    assert err.max() < 1e-5
    p = csx.SolventParams()
    ref = csx.solvent_corrected_intensity(s, q, p, precision=F64)
    got = csx.solvent_corrected_intensity(s, q, p, precision=F32)
    err = rel(got, ref)
    print("20k solvent max", err.max(), "mean", err.mean(), flush=True)
    assert err.max() < 1e-5


def test_precision_study_20k_against_cpu_oracle(gpu_ctx):
    """The 20,000-atom precision-study output against the FP64 CPU oracle
    (solvent.rs:490-584 restated) at 8 of its 500 q, q = 1 included, where
    the pair sum's condition number is largest: FP32 per q within 1e-5.

    The oracle sums 2e8 pair terms one after another in FP64, as the
    reference does; where they share a sign (small q) its own error reaches
    about N eps = 2e-8, so the FP64 GPU path is held to 5e-8 against it.  At
    q = 0, where I(0) = (sum_i A_i)^2 is known exactly from the per-atom
    amplitudes A_i = f_i(0) - G(0, c1) e_i(0), the FP32 path (whose q = 0
    tables are exact occupancy products) is held to 1e-12 and the FP64 GPU
    path (blocked FP64 sums of the 2e8 terms) to 1e-10."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    q = np.linspace(0.0, 1.0, 500)
    idx = np.array([0, 1, 50, 125, 250, 375, 450, 499])
    p = csx.SolventParams()
    got = csx.solvent_corrected_intensity(s, q, p, precision=F32)[idx]
    got64 = csx.solvent_corrected_intensity(s, q[idx], p, precision=F64)
    basis = O.debye_solvent_basis(oracle_atoms(s), q[idx])
    want = O.combine_solvent_basis(basis, q[idx], p.c1, p.c2)
    err = rel(got, want)
    print("20k vs CPU oracle per q:", dict(zip(q[idx].round(4).tolist(), err.tolist())), flush=True)
    print("20k FP64 GPU vs CPU oracle per q:", rel(got64, want).tolist(), flush=True)
    assert err.max() < 1e-5
    assert rel(got64, want).max() < 5e-8
    L = O.lib()
    g0 = L.or_excluded_volume_scale(0.0, p.c1)
    amp = {e: L.or_form_factor(e, -1, 1.0, 1, 0.0)
           - g0 * p.bulk_density_e_per_a3 * L.or_excluded_volume(e, -1, 1)
           for e in set(s.element.tolist())}
    exact0 = math.fsum(amp[e] for e in s.element.tolist()) ** 2
    print("I(0): exact", exact0, "FP32", got[0], "FP64 GPU", got64[0], "oracle", want[0], flush=True)
    assert abs(got[0] - exact0) < 1e-12 * exact0
    assert abs(got64[0] - exact0) < 1e-10 * exact0


@pytest.mark.parametrize("n", [1200, 20000])
def test_coordinate_features_match_oracle(gpu_ctx, n):
    """P(r) pair histogram on the GPU against the oracle's restatement of
    coordinate_pr: Dmax and Rg bitwise, the normalised distribution to 1e-12
    (only the FP64 summation order differs)."""
    s = csx.generate_globule(n, 35.0 * (n / 20000) ** (1 / 3), 1.0, seed=17)
    f = csx.coordinate_features(s, 160)
    rg, dmax, hist = O.coordinate_features(O.Atoms(s.xyz, s.element), 160)
    assert f.dmax == dmax and f.rg == rg
    r = (np.arange(160) + 0.5) * (dmax / 160)  # metrics.rs:529-531
    r_o, p_o, rg_o, dmax_o = O.normalize_distribution(r, hist)
    assert np.array_equal(f.p_r.r, r_o)
    assert np.max(np.abs(f.p_r.p - p_o)) <= 1e-12 * np.max(p_o)
    assert abs(f.p_r.rg - rg_o) <= 1e-12 * rg_o and f.p_r.dmax == dmax_o


def test_coordinate_features_reference_fixture_and_edges(gpu_ctx):
    # metrics.rs:651-684
    two = csx.Structure.from_arrays([[0, 0, 0], [4, 0, 0]], [1, 1])
    f = csx.coordinate_features(two)
    assert abs(f.dmax - 4.0) < 1e-12 and f.rg > 1.0 and f.p_r.p.size > 0
    one = csx.Structure.from_arrays([[1.0, 2.0, 3.0]], [1])
    f1 = csx.coordinate_features(one)
    assert f1.dmax == 0.0 and list(f1.p_r.r) == [0.0] and list(f1.p_r.p) == [1.0]
    assert f1.p_r.rg == 0.0 and f1.p_r.dmax == 0.0
    with pytest.raises(csx.InvalidInput):
        csx.coordinate_features(csx.Structure.from_arrays(np.zeros((0, 3)), np.zeros(0, np.uint8)))
    with pytest.raises(csx.InvalidInput):
        csx.coordinate_features(two, bins=5000)


# ------------------------------------ radial interpolation sweep (mode 4) --

def test_radial_sweep_against_fp64_and_oracle(gpu_ctx):
    """Mode 4 (pair distances binned at h = the largest power of two with
    h max|q| <= 1.04, j_0 interpolated at 8 Chebyshev nodes per bin, 64-bit
    fixed-point sums) on a uniform and an irregular grid against FP64 GPU
    Debye, and the solvent intensity against the FP64 CPU oracle."""
    s = csx.generate_globule(3000, 20.0, 1.0, seed=21)
    grids = (np.linspace(0.0, 1.0, 301), np.sort(np.random.default_rng(2).uniform(0.0, 1.2, 97)))
    small = csx.generate_globule(1200, 16.0, 1.0, seed=3)
    qq = 0.003 + 0.0049 * np.arange(100)
    p = csx.SolventParams(c1=1.02, c2=1.5)
    gpu_ctx.set_tuning(0, 4, 0)
    try:
        for q in grids:
            got = csx.debye_intensity(s, q)
            assert gpu_ctx.last_launch() == {"q_tile": 1, "sin_mode": 4, "persistent_workgroups": 0}
            err = rel(got, csx.debye_intensity(s, q, precision=F64)).max()
            print("radial vs fp64", q.size, err, flush=True)
            assert err < 1e-5
        got = csx.solvent_corrected_intensity(small, qq, p)
    finally:
        gpu_ctx.set_tuning()
    want = O.combine_solvent_basis(O.debye_solvent_basis(oracle_atoms(small), qq, c1=1.02, c2=1.5),
                                   qq, 1.02, 1.5)
    assert rel(got, want).max() < 1e-5


def test_radial_sweep_sums_do_not_depend_on_order(gpu_ctx):
    """Integer sums: the tables are the same bits for every j chunk (which
    splits and orders the pairs) and every count of LDS accumulator copies,
    alone or in a batch, call after call."""
    cs = [csx.generate_globule(n, 10.0 + n / 400, 1.0, seed=60 + n) for n in (300, 1700, 2500)]
    q = np.linspace(0.0, 0.8, 90)
    try:
        gpu_ctx.set_tuning(0, 4, 0)
        base = [csx.debye_intensity(c, q) for c in cs]
        batch = csx.batch_debye_intensity(cs, q)
        for k in range(3):
            assert np.array_equal(batch[k], base[k])
        for copies, chunk in ((0, 64), (8, 256), (16, 1024), (32, 256)):
            gpu_ctx.set_tuning(copies, 4, chunk)
            for c, b in zip(cs, base):
                assert np.array_equal(csx.debye_intensity(c, q), b), (copies, chunk)
    finally:
        gpu_ctx.set_tuning()


def test_radial_sweep_copies_are_validated(gpu_ctx):
    with pytest.raises(csx.InvalidInput):
        gpu_ctx.set_tuning(24, 4, 0)
    gpu_ctx.set_tuning()


def test_radial_sweep_falls_back_beyond_its_bins(gpu_ctx):
    """A structure whose pair distances need more bins than a workgroup's
    LDS holds (2,000 A apart at h = 1 A) takes a direct sweep instead."""
    s = csx.Structure.from_arrays([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [2000.0, 0.0, 0.0]], [1, 1, 3])
    q = np.linspace(0.0, 1.0, 20)
    gpu_ctx.set_tuning(0, 4, 0)
    try:
        got = csx.debye_intensity(s, q)
        assert gpu_ctx.last_launch()["sin_mode"] in (1, 2)
    finally:
        gpu_ctx.set_tuning()
    assert rel(got, csx.debye_intensity(s, q, precision=F64)).max() < 1e-5


# ------------------------------------------------ sine recurrence (mode 3) --

@pytest.mark.parametrize("tile", [0, 16, 25, 32, 50, 64, 100, 125, 128])
def test_recurrence_tiles_against_fp64(gpu_ctx, tile):
    """Every recurrence tile (restarts at every tile start) against FP64 GPU
    Debye on a uniform grid, including q = 0 and a last tile with padding."""
    s = csx.generate_globule(3000, 20.0, 1.0, seed=21)
    q = np.linspace(0.0, 1.0, 301)
    want = csx.debye_intensity(s, q, precision=F64)
    gpu_ctx.set_tuning(tile, 3, 0)
    try:
        got = csx.debye_intensity(s, q)
        launch = gpu_ctx.last_launch()
    finally:
        gpu_ctx.set_tuning()
    assert launch["sin_mode"] == 3 and launch["q_tile"] == (tile or launch["q_tile"]), launch
    err = rel(got, want).max()
    print("recurrence tile", tile, err, flush=True)
    assert err < 1e-5


def test_recurrence_against_cpu_oracle(gpu_ctx):
    """The recurrence's solvent intensity against the FP64 CPU oracle on a
    uniform grid that does not start at 0, in 25-wide tiles (every tile
    starts mid-grid)."""
    s = csx.generate_globule(1200, 16.0, 1.0, seed=3)
    q = 0.003 + 0.0049 * np.arange(100)
    p = csx.SolventParams(c1=1.02, c2=1.5)
    gpu_ctx.set_tuning(25, 3, 0)
    try:
        got = csx.solvent_corrected_intensity(s, q, p)
        assert gpu_ctx.last_launch()["sin_mode"] == 3
    finally:
        gpu_ctx.set_tuning()
    want = O.combine_solvent_basis(O.debye_solvent_basis(oracle_atoms(s), q, c1=1.02, c2=1.5),
                                   q, 1.02, 1.5)
    err = rel(got, want).max()
    print("recurrence vs oracle", err, flush=True)
    assert err < 1e-5


def test_recurrence_fallbacks(gpu_ctx):
    """A grid that is not uniform, or whose step turns the largest pair's
    phase by more than 1/4 revolution, takes the direct hardware sine, with
    the bits of an explicit mode 2; a descending uniform grid takes the
    recurrence."""
    s = csx.generate_globule(2000, 17.0, 1.0, seed=8)
    gpu_ctx.set_tuning()
    q = np.linspace(0.0, 0.8, 120)
    q[57] += 1e-6
    a = csx.debye_intensity(s, q)
    assert gpu_ctx.last_launch()["sin_mode"] == 2
    gpu_ctx.set_tuning(0, 2, 0)
    try:
        b = csx.debye_intensity(csx.Structure.from_arrays(s.xyz, s.element), q)
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(a, b)
    coarse = np.linspace(0.0, 3.0, 5)  # 0.75 A^-1 steps over a ~36 A diameter
    got = csx.debye_intensity(s, coarse)
    assert gpu_ctx.last_launch()["sin_mode"] == 2
    assert rel(got, csx.debye_intensity(s, coarse, precision=F64)).max() < 1e-5
    down = np.linspace(0.9, 0.01, 90)
    got = csx.debye_intensity(s, down)
    assert gpu_ctx.last_launch()["sin_mode"] == 3
    assert rel(got, csx.debye_intensity(s, down, precision=F64)).max() < 1e-5


def test_recurrence_mode_is_bitwise_reproducible(gpu_ctx):
    """For a given j chunk and q tile the recurrence's tables are the same
    on every call, for the grid and persistent launch forms alike."""
    s = csx.generate_globule(6000, 24.0, 1.0, seed=13)
    q = np.linspace(0.0, 0.6, 200)
    with gpu_ctx.caching(0):
        base = csx.debye_intensity(s, q)
        try:
            for wgs in (0, 1, 4):
                gpu_ctx.set_persistence(wgs)
                assert np.array_equal(csx.debye_intensity(s, q), base), wgs
        finally:
            gpu_ctx.set_persistence()


def test_recurrence_tiles_are_held_to_the_phase_limit(gpu_ctx):
    """A coarse uniform grid (0.02 A^-1 steps over a ~50 A diameter: about
    0.16 revolutions of the largest pair's phase per step) keeps the
    recurrence's tiles to 4 revolutions: a tuned 128-wide tile falls back to
    one of at most 25 q, and the result stays within 1e-5 of FP64."""
    s = csx.generate_globule(6000, 24.0, 1.0, seed=17)
    q = np.linspace(0.0, 2.0, 100)
    want = csx.debye_intensity(s, q, precision=F64)
    gpu_ctx.set_tuning(128, 3, 0)
    try:
        got = csx.debye_intensity(s, q)
        launch = gpu_ctx.last_launch()
    finally:
        gpu_ctx.set_tuning()
    assert launch["sin_mode"] == 3 and launch["q_tile"] in (16, 25), launch
    assert rel(got, want).max() < 1e-5


def test_tiny_calls_keep_the_direct_sweep(gpu_ctx):
    """Below 2e7 pair.q a call is one short task per CU, where the direct
    sweep's shorter chain is faster: 300 atoms x 100 q (4.5e6 pair.q) take
    the direct hardware sine by default, 1,200 atoms x 100 q (7.2e7) the
    recurrence; a tuned recurrence tile overrides the rule.  All within
    1e-5 of FP64."""
    gpu_ctx.set_tuning()
    for n, mode in ((300, 2), (1200, 3)):
        s = csx.generate_globule(n, 35.0 * (n / 20000) ** (1 / 3), 1.0, seed=5)
        q = np.linspace(0.0, 1.0, 100)
        got = csx.debye_intensity(s, q)
        assert gpu_ctx.last_launch()["sin_mode"] == mode, n
        assert rel(got, csx.debye_intensity(s, q, precision=F64)).max() < 1e-5
    gpu_ctx.set_tuning(16, 3, 0)
    try:
        csx.debye_intensity(csx.generate_globule(300, 9.0, 1.0, seed=5), np.linspace(0.0, 1.0, 100))
        assert gpu_ctx.last_launch()["sin_mode"] == 3
    finally:
        gpu_ctx.set_tuning()

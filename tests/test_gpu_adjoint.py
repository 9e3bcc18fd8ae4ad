"""GPU adjoint (dI/dx) against the FP64 oracle gradient and central finite
differences (h = 1e-4 A) of the FP64 GPU intensity."""
import numpy as np
import pytest

import crabsaxs_amd as csx
import oracle_py as O

pytestmark = pytest.mark.gpu
F64 = csx.Precision.FP64_PLAIN


def oatoms(s):
    return O.Atoms(s.xyz, s.element, s.group, s.occupancy)


@pytest.mark.parametrize("solvent", [None, csx.SolventParams(c1=1.02, c2=2.0)])
def test_jacobian_and_vjp_match_oracle(gpu_ctx, solvent):
    s = csx.generate_globule(1500, 16.0, 1.0, seed=51)
    q = np.linspace(0.0, 1.0, 40)
    osolv = None if solvent is None else (solvent.c1, solvent.c2, 3.0, 0.334)
    want_j = O.debye_grad_x(oatoms(s), q, solvent=osolv, jacobian=True)
    got_j = csx.debye_jacobian(s, q, solvent)
    scale = np.abs(want_j).max(axis=(1, 2), keepdims=True)
    err = np.abs(got_j - want_j) / np.maximum(scale, 1e-300)
    print("jacobian max err (rel to per-q max)", err.max(), flush=True)
    assert err.max() < 1e-4
    w = np.random.default_rng(3).normal(size=q.size)
    want = np.einsum("k,kic->ic", w, want_j)
    got = csx.debye_grad_x(s, q, w, solvent)
    e = np.abs(got - want).max() / np.abs(want).max()
    print("vjp max err", e, flush=True)
    assert e < 1e-4


_TILE_CASE = {}


@pytest.mark.parametrize("tile", [0, 16, 32, 48, 60, 64])
def test_vjp_tiles_match_oracle(gpu_ctx, tile):
    """Every VJP q tile (0: the automatic choice, 16-wide for a structure
    this small) against the oracle gradient; 100 q leave a partial last
    sweep for every tile."""
    if not _TILE_CASE:
        s = csx.generate_globule(1200, 15.0, 1.0, seed=52)
        q = np.linspace(0.01, 0.9, 100)
        solvent = csx.SolventParams(c1=0.98, c2=1.5)
        w = np.random.default_rng(4).normal(size=q.size)
        want = O.debye_grad_x(oatoms(s), q, w, solvent=(solvent.c1, solvent.c2, 3.0, 0.334))
        _TILE_CASE.update(s=s, q=q, solvent=solvent, w=w, want=want)
    c = _TILE_CASE
    try:
        gpu_ctx.set_vjp_tile(tile)
        got = csx.debye_grad_x(c["s"], c["q"], c["w"], c["solvent"])
    finally:
        gpu_ctx.set_vjp_tile()
    e = np.abs(got - c["want"]).max() / np.abs(c["want"]).max()
    print("vjp tile", tile, "max err", e, flush=True)
    assert e < 1e-4
    with pytest.raises(csx.InvalidInput):
        gpu_ctx.set_vjp_tile(40)


def test_vjp_direct_and_recurrence_sweeps_match_oracle(gpu_ctx):
    """A uniform grid takes the sine recurrence, the same grid with one q
    moved by 1e-7 the direct hardware sines; both against the oracle."""
    s = csx.generate_globule(900, 13.0, 1.0, seed=56)
    solvent = csx.SolventParams(c1=1.0, c2=1.2)
    w = np.random.default_rng(6).normal(size=150)
    for shift in (0.0, 1e-7):
        q = np.linspace(0.005, 1.2, 150)
        q[71] += shift
        want = O.debye_grad_x(oatoms(s), q, w, solvent=(1.0, 1.2, 3.0, 0.334))
        got = csx.debye_grad_x(s, q, w, solvent)
        e = np.abs(got - want).max() / np.abs(want).max()
        print("vjp shift", shift, gpu_ctx.last_vjp_launch(), "max err", e, flush=True)
        assert e < 1e-5


def test_directional_derivative_full_size(gpu_ctx):
    """20,000 atoms, 300 q: sum_i grad_i . d against central differences of
    the FP64 GPU intensity along a random direction d (h = 1e-4 A)."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    q = np.linspace(0.01, 1.0, 300)
    p = csx.SolventParams(c1=1.02, c2=2.0)
    rng = np.random.default_rng(5)
    w = rng.normal(size=q.size)
    g = csx.debye_grad_x(s, q, w, p)
    d = rng.normal(size=s.xyz.shape)
    h = 1e-4
    up = csx.Structure.from_arrays(s.xyz + h * d, s.element)
    dn = csx.Structure.from_arrays(s.xyz - h * d, s.element)
    fd = (csx.solvent_corrected_intensity(up, q, p, precision=F64)
          - csx.solvent_corrected_intensity(dn, q, p, precision=F64)) / (2 * h)
    lhs, rhs = float(np.sum(g * d)), float(w @ fd)
    print("directional", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 2e-4 * np.sum(np.abs(w * fd)) + 1e-6 * abs(rhs)


def test_vjp_is_deterministic_and_symmetric_in_pairs(gpu_ctx):
    """Every unordered pair is visited once and its two ends get opposite
    contributions: the gradient sums to zero (translation invariance of I),
    and two calls agree bitwise."""
    s = csx.generate_domain_chain(3000, 3, seed=8)
    q = np.linspace(0.0, 0.6, 70)
    w = np.random.default_rng(9).normal(size=q.size)
    g1 = csx.debye_grad_x(s, q, w)
    g2 = csx.debye_grad_x(s, q, w)
    assert np.array_equal(g1, g2)
    assert np.abs(g1.sum(axis=0)).max() <= 1e-6 * np.abs(g1).sum(axis=0).max()


def test_vjp_over_several_launches(gpu_ctx):
    """A 1 MiB partial-sum workspace splits the 20,000-atom VJP into many
    launches; the result equals the one-launch result to summation order."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    q = np.linspace(0.01, 1.0, 300)
    w = np.random.default_rng(5).normal(size=q.size)
    one = csx.debye_grad_x(s, q, w)
    try:
        gpu_ctx.set_workspace_limit(1 << 20)
        many = csx.debye_grad_x(s, q, w)
    finally:
        gpu_ctx.set_workspace_limit()
    assert np.abs(many - one).max() <= 1e-12 * np.abs(one).max()


def test_vjp_step_split_and_several_launches(gpu_ctx):
    """A 300-atom structure: the automatic shape splits each sweep over 8
    workgroups; a 1 MiB workspace then takes its block pairs in several
    launches.  Every split (1, 2, 4, 8), in one launch or several, matches
    the oracle gradient."""
    s = csx.generate_globule(300, 9.0, 1.0, seed=57)
    q = np.linspace(0.01, 0.9, 120)
    w = np.random.default_rng(8).normal(size=q.size)
    solvent = csx.SolventParams(c1=1.01, c2=0.8)
    want = O.debye_grad_x(oatoms(s), q, w, solvent=(1.01, 0.8, 3.0, 0.334))
    results = {}
    try:
        csx.debye_grad_x(s, q, w, solvent)
        assert gpu_ctx.last_vjp_launch()["step_split"] == 8
        assert gpu_ctx.last_vjp_launch()["launches"] == 1
        gpu_ctx.set_workspace_limit(1 << 20)
        for split in (1, 2, 4, 8):
            gpu_ctx.set_vjp_split(split)
            got = csx.debye_grad_x(s, q, w, solvent)
            shape = gpu_ctx.last_vjp_launch()
            assert shape["step_split"] == split
            if split == 8:
                assert shape["launches"] > 1, shape
            e = np.abs(got - want).max() / np.abs(want).max()
            print("vjp split", split, shape, "max err", e, flush=True)
            assert e < 1e-5
            results[split] = got
    finally:
        gpu_ctx.set_vjp_split()
        gpu_ctx.set_workspace_limit()
    for split in (2, 4, 8):
        assert np.abs(results[split] - results[1]).max() < 1e-5 * np.abs(want).max()


# -------------------------------------------- radial interpolation (mode 4) --

@pytest.fixture
def radial_mode(gpu_ctx):
    gpu_ctx.set_tuning(0, 4, 0)
    yield gpu_ctx
    gpu_ctx.set_tuning()
    gpu_ctx.set_vjp_split()


def test_radial_vjp_matches_oracle(radial_mode):
    """Sine mode 4: F_ab(r) from per-bin Chebyshev coefficients instead of
    the q loop, on the shapes of the q-loop tests (uniform and non-uniform
    grids, step splits 1-8) against the oracle gradient."""
    s = csx.generate_globule(900, 13.0, 1.0, seed=56)
    solvent = csx.SolventParams(c1=1.0, c2=1.2)
    w = np.random.default_rng(6).normal(size=150)
    for shift in (0.0, 1e-7):
        q = np.linspace(0.005, 1.2, 150)
        q[71] += shift
        want = O.debye_grad_x(oatoms(s), q, w, solvent=(1.0, 1.2, 3.0, 0.334))
        for split in (0, 1, 8):
            radial_mode.set_vjp_split(split)
            got = csx.debye_grad_x(s, q, w, solvent)
            assert radial_mode.last_vjp_launch()["q_tile"] == 0
            e = np.abs(got - want).max() / np.abs(want).max()
            print("radial vjp shift", shift, "split", split, radial_mode.last_vjp_launch(),
                  "max err", e, flush=True)
            assert e < 1e-5
    radial_mode.set_vjp_split()
    # deterministic, and the pair contributions cancel in the sum
    q = np.linspace(0.0, 0.6, 70)
    w = np.random.default_rng(9).normal(size=q.size)
    c = csx.generate_domain_chain(3000, 3, seed=8)
    g1 = csx.debye_grad_x(c, q, w)
    assert np.array_equal(g1, csx.debye_grad_x(c, q, w))
    assert np.abs(g1.sum(axis=0)).max() <= 1e-6 * np.abs(g1).sum(axis=0).max()


def test_radial_vjp_directional_derivative_full_size(radial_mode):
    """20,000 atoms, 300 q in mode 4 against central differences of the
    FP64 GPU intensity (the bound of the q-loop test)."""
    s = csx.generate_globule(20000, 35.0, 1.0, seed=1)
    q = np.linspace(0.01, 1.0, 300)
    p = csx.SolventParams(c1=1.02, c2=2.0)
    rng = np.random.default_rng(5)
    w = rng.normal(size=q.size)
    g = csx.debye_grad_x(s, q, w, p)
    assert radial_mode.last_vjp_launch()["q_tile"] == 0
    d = rng.normal(size=s.xyz.shape)
    h = 1e-4
    up = csx.Structure.from_arrays(s.xyz + h * d, s.element)
    dn = csx.Structure.from_arrays(s.xyz - h * d, s.element)
    fd = (csx.solvent_corrected_intensity(up, q, p, precision=F64)
          - csx.solvent_corrected_intensity(dn, q, p, precision=F64)) / (2 * h)
    lhs, rhs = float(np.sum(g * d)), float(w @ fd)
    print("radial directional", lhs, rhs, abs(lhs - rhs) / np.sum(np.abs(w * fd)), flush=True)
    assert abs(lhs - rhs) <= 2e-4 * np.sum(np.abs(w * fd)) + 1e-6 * abs(rhs)


def test_radial_vjp_falls_back_beyond_its_bins(radial_mode):
    """Points 20,000 A apart need more than 2,048 bins of 1 A: the q-loop
    sweep runs instead."""
    s = csx.Structure.from_arrays([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [20000.0, 0.0, 0.0]],
                                  [1, 1, 3])
    q = np.linspace(0.01, 1.0, 20)
    w = np.ones(q.size)
    got = csx.debye_grad_x(s, q, w)
    assert radial_mode.last_vjp_launch()["q_tile"] != 0
    want = O.debye_grad_x(oatoms(s), q, w)
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()

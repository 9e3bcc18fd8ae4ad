"""GPU parity of the multipole path (FP16-split MFMA contraction, exact j_l)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import crabsaxs_amd as csx
import oracle_py as O

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "debye_golden.json").read_text())


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def gold(name):
    s = GOLD["structures"][name]
    return csx.Structure.from_arrays(s["xyz"], s["element"])


def oatoms(s):
    return O.Atoms(s.xyz, s.element, s.group, s.occupancy)


def test_two_atoms_match_debye(gpu_ctx):
    # spherical.rs:596-634 and 637-678 (L = 32, 1e-6)
    for xyz, el, q in [([[0, 0, 0], [2, 0, 0]], [1, 1], [0.1, 0.2, 0.5]),
                       ([[-20, 0, 0], [20, 0, 0]], [1, 3], [0.2, 0.35, 0.5])]:
        s = csx.Structure.from_arrays(xyz, el)
        m = csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=32))
        d = csx.debye_intensity(s, q, precision=csx.Precision.FP64_PLAIN)
        print("two atoms", rel(m, d).max(), flush=True)
        assert rel(m, d).max() < 2e-6


def test_helix_nrmse(gpu_ctx):
    # spherical.rs:681-727 (24 atoms, L = 24, NRMSE <= 1e-3) and parity with
    # the reference's own Miller j_l, accurate at these q r.
    s = gold("helix24")
    q = np.arange(1, 101) * 0.005
    m = csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=24))
    d = csx.debye_intensity(s, q, precision=csx.Precision.FP64_PLAIN)
    assert math.sqrt(((m - d) ** 2).sum() / (d ** 2).sum()) <= 1e-3
    ref = O.multipole_intensity(oatoms(s), q, 24, exact=False)
    print("helix vs reference-miller oracle", rel(m, ref).max(), flush=True)
    assert rel(m, ref).max() < 1e-5


def test_compute_curve_auto_multipole_spiral(gpu_ctx):
    # scattering.rs:312-365: adaptive l_max multipole within 1e-3 (rel-L2) of Debye
    s = gold("spiral96")
    q = np.arange(1, 101) * 0.005
    c = csx.compute_curve(s, q, csx.ComputeOptions(method=csx.ScatteringMethod.Multipole))
    assert c.metadata.method == "Multipole"
    d = csx.debye_intensity(s, q, precision=csx.Precision.FP64_PLAIN)
    assert math.sqrt(((c.intensity - d) ** 2).sum() / (d ** 2).sum()) <= 1e-3


def test_adaptive_grid_basis_matches_debye(gpu_ctx):
    # spherical.rs:730-787
    s = gold("eight")
    q = np.arange(1, 51) * 0.01
    p = csx.SolventParams(hydration_model=csx.HydrationModel.AdaptiveGrid,
                          hydration_grid_spacing_angstrom=8.0)
    m = csx.multipole_solvent_basis(s, q, csx.MultipoleParams(l_max=32), p)
    d = csx.debye_solvent_basis(s, q, p, precision=csx.Precision.FP64_PLAIN)
    for t in range(6):
        err = math.sqrt(((m[t] - d[t]) ** 2).sum() / max((d[t] ** 2).sum(), 1e-30))
        assert err <= 1e-3, (t, err)


@pytest.mark.parametrize("l_max", [15, 25, 40, 64])
def test_globule_matches_exact_oracle(gpu_ctx, l_max):
    s = csx.generate_globule(1200, 16.0, 1.0, seed=21)
    q = np.linspace(0.005, 0.5, 60)
    got = csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=l_max))
    want = O.multipole_intensity(oatoms(s), q, l_max, exact=True)
    print("globule L", l_max, rel(got, want).max(), flush=True)
    assert rel(got, want).max() < 1e-5


def test_occupancies_above_one(gpu_ctx):
    """Occupancies up to 6 scale the FP16 j_l operands by 2^11 instead of
    2^14 (the largest must stay below the FP16 range); the sums match the
    exact oracle as before."""
    s = csx.generate_globule(600, 12.0, 1.0, seed=31)
    occ = np.random.default_rng(31).uniform(0.05, 6.0, len(s))
    s = csx.Structure.from_arrays(s.xyz, s.element, occupancy=occ)
    q = np.linspace(0.005, 0.5, 40)
    for l_max in (15, 25):
        got = csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=l_max))
        want = O.multipole_intensity(oatoms(s), q, l_max, exact=True)
        print("occupancy up to 6, L", l_max, rel(got, want).max(), flush=True)
        assert rel(got, want).max() < 1e-5


def test_centre_point_and_distant_neighbours(gpu_ctx):
    """A point exactly at the centre (q r = 0: j_0 = 1, j_l>0 = 0) and
    neighbours in the radius order that lie far apart (radii 0, 5.2 and 20 A
    in one slab) take the generator's one-column path; the sums still match
    the exact oracle, and Debye where the expansion has converged."""
    cube = [[x, y, z] for x in (-3, 3) for y in (-3, 3) for z in (-3, 3)]
    octa = [[20, 0, 0], [-20, 0, 0], [0, 20, 0], [0, -20, 0], [0, 0, 20], [0, 0, -20]]
    xyz = np.array([[0.0, 0.0, 0.0]] + cube + octa)
    s = csx.Structure.from_arrays(xyz, np.full(len(xyz), 1, dtype=np.uint8))
    q = np.linspace(0.0, 0.5, 41)
    for l_max in (15, 40):
        got = csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=l_max))
        want = O.multipole_intensity(oatoms(s), q, l_max, exact=True)
        print("centre point, L", l_max, rel(got, want).max(), flush=True)
        assert rel(got, want).max() < 1e-5
    d = csx.debye_intensity(s, q, precision=csx.Precision.FP64_PLAIN)
    assert rel(got, d).max() < 1e-5  # L = 40 has converged for q R <= 10


@pytest.fixture(params=[1, 2], ids=["points", "nodes"])
def radial(request, gpu_ctx):
    """Every point in the coefficient kernel (1), or radial interpolation
    nodes (2)."""
    gpu_ctx.set_multipole_radial(request.param)
    yield request.param
    gpu_ctx.set_multipole_radial()


@pytest.mark.parametrize("model", [csx.HydrationModel.Analytic, csx.HydrationModel.AdaptiveGrid])
def test_solvent_paths_match_exact_oracle(gpu_ctx, model, radial):
    s = csx.generate_globule(700, 14.0, 1.0, seed=22)
    q = np.linspace(0.0, 0.5, 41)
    p = csx.SolventParams(c1=1.02, c2=0.6, hydration_model=model,
                          hydration_grid_spacing_angstrom=4.0)
    sites, vol = (None, 0.0)
    if model == csx.HydrationModel.AdaptiveGrid:
        sites, vol = O.adaptive_grid(s.xyz, spacing=4.0)
    mp = csx.MultipoleParams(l_max=30)
    want_i = O.multipole_solvent(oatoms(s), q, 30, c1=1.02, c2=0.6, sites=sites,
                                 voxel_volume=vol, exact=True)
    want_b = O.multipole_solvent(oatoms(s), q, 30, c1=1.02, c2=0.6, sites=sites,
                                 voxel_volume=vol, basis=True, exact=True)
    got_i = csx.multipole_solvent_intensity(s, q, mp, p)
    got_b = csx.multipole_solvent_basis(s, q, mp, p)
    assert (gpu_ctx.last_multipole()["bin_width"] > 0) == (radial == 2)
    print(model.name, radial, rel(got_i, want_i).max(), flush=True)
    assert rel(got_i, want_i).max() < 1e-5
    scale = np.abs(want_b).max(axis=1, keepdims=True)
    assert (np.abs(got_b - want_b) / scale).max() < 1e-5


@pytest.mark.parametrize("l_max", [4, 11])
@pytest.mark.parametrize("model", [csx.HydrationModel.Analytic, csx.HydrationModel.AdaptiveGrid])
def test_low_order_basis_matches_exact_oracle(gpu_ctx, model, l_max):
    # at most 16 column tiles: the three-channel kernel with 4 accumulators
    # per wave (one tile group)
    s = csx.generate_globule(500, 12.0, 1.0, seed=24)
    q = np.linspace(0.0, 0.3, 33)
    p = csx.SolventParams(c1=0.98, c2=1.4, hydration_model=model,
                          hydration_grid_spacing_angstrom=4.0)
    sites, vol = (None, 0.0)
    if model == csx.HydrationModel.AdaptiveGrid:
        sites, vol = O.adaptive_grid(s.xyz, spacing=4.0)
    want = O.multipole_solvent(oatoms(s), q, l_max, c1=0.98, c2=1.4, sites=sites,
                               voxel_volume=vol, basis=True, exact=True)
    got = csx.multipole_solvent_basis(s, q, csx.MultipoleParams(l_max=l_max), p)
    scale = np.abs(want).max(axis=1, keepdims=True)
    err = (np.abs(got - want) / scale).max()
    print(model.name, "L", l_max, err, flush=True)
    assert err < 1e-5


def test_empty_and_single(gpu_ctx):
    empty = csx.Structure.from_arrays(np.zeros((0, 3)), np.zeros(0, dtype=np.uint8))
    q = np.array([0.0, 0.1])
    assert np.array_equal(csx.multipole_intensity(empty, q), np.zeros(2))
    one = csx.Structure.from_arrays([[3.0, -1.0, 2.0]], [3])
    f = np.array([csx.form_factor_for_atom(csx.Atom(pos=(0, 0, 0), element=csx.Element.O),
                                           True, v) for v in q])
    assert rel(csx.multipole_intensity(one, q, csx.MultipoleParams(l_max=8)), f * f).max() < 1e-6


def test_prolate_with_shell_sites_large(gpu_ctx):
    """10,000-atom 2:1 prolate globule with 3 A shell pseudo-atoms, L = 25,
    against the exact-j_l FP64 oracle on a q subset."""
    s = csx.generate_globule(10000, 30.0 / 2 ** (1 / 3), 2.0, seed=23)
    sites = csx.generate_shell_sites(s, 3.0, 3.0, 6.0)
    xyz = np.vstack([s.xyz, sites])
    el = np.concatenate([s.element, np.full(len(sites), 3, dtype=np.uint8)])
    full = csx.Structure.from_arrays(xyz, el)
    q = np.linspace(0.005, 0.5, 24)
    got = csx.multipole_intensity(full, q, csx.MultipoleParams(l_max=25))
    want = O.multipole_intensity(oatoms(full), q, 25, exact=True)
    print("prolate", len(full), rel(got, want).max(), flush=True)
    assert rel(got, want).max() < 1e-5


def test_radial_nodes_match_every_point(gpu_ctx):
    """Radial interpolation (8 Chebyshev nodes per bin of width h, the
    largest power of two with h max|q| <= 1.04) against every point in the
    coefficient kernel, on the 10,000-atom prolate with shell sites: both
    forms agree to FP32 accuracy and meet the exact oracle; the bin width
    follows the q range; the automatic choice takes the nodes here and the
    points for a structure smaller than its nodes."""
    s = csx.generate_globule(10000, 30.0 / 2 ** (1 / 3), 2.0, seed=23)
    sites = csx.generate_shell_sites(s, 3.0, 3.0, 6.0)
    full = csx.Structure.from_arrays(np.vstack([s.xyz, sites]),
                                     np.concatenate([s.element, np.full(len(sites), 3, dtype=np.uint8)]))
    n = len(full)
    q = np.linspace(0.005, 0.5, 24)
    mp = csx.MultipoleParams(l_max=25)
    try:
        got = {}
        for mode in (1, 2):
            gpu_ctx.set_multipole_radial(mode)
            got[mode] = csx.multipole_intensity(full, q, mp)
            info = gpu_ctx.last_multipole()
            if mode == 1:
                assert info == {"bin_width": 0.0, "points": n}
            else:
                assert info["bin_width"] == 2.0 and 0 < info["points"] <= n // 2, info
        print("nodes vs points", rel(got[2], got[1]).max(), flush=True)
        assert rel(got[2], got[1]).max() < 2e-6
        want = O.multipole_intensity(oatoms(full), q, 25, exact=True)
        assert rel(got[2], want).max() < 1e-5
        # the bin width follows max|q|: 1.04 / 1.0 -> 1, 1.04 / 0.25 -> 4
        for qmax, h in ((1.0, 1.0), (0.25, 4.0)):
            csx.multipole_intensity(full, np.linspace(0.0, qmax, 8), mp)
            assert gpu_ctx.last_multipole()["bin_width"] == h, qmax
        gpu_ctx.set_multipole_radial(0)
        csx.multipole_intensity(full, q, mp)
        assert gpu_ctx.last_multipole()["bin_width"] == 2.0
        small = csx.generate_globule(40, 20.0, 1.0, seed=3)
        csx.multipole_intensity(small, q, mp)
        assert gpu_ctx.last_multipole() == {"bin_width": 0.0, "points": 40}
    finally:
        gpu_ctx.set_multipole_radial()


def test_radial_mode_is_validated(gpu_ctx):
    with pytest.raises(csx.InvalidInput):
        gpu_ctx.set_multipole_radial(3)

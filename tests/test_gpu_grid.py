"""Hydration grids on the GPU (solvent.rs:121-223, 259-401) against the
oracle: sites bitwise equal and in the reference's scan order, on the
reference's test structures and its assertions (solvent.rs:688-771), random
globules up to 20,000 atoms, and a sealed hollow shell (an enclosed void
that the exterior flood fill must leave out)."""
import numpy as np
import pytest

import crabsaxs_amd as csx
import grid_cases as G
import oracle_py as O

pytestmark = pytest.mark.gpu


def structure(xyz, el):
    return csx.Structure.from_arrays(np.asarray(xyz, float), np.asarray(el, np.uint8))


def two_atoms():  # solvent.rs:606-631 (carbon and oxygen 2 A apart)
    return structure([[-1, 0, 0], [1, 0, 0]], [1, 3])


def test_reference_grid_tests(gpu_ctx):
    one = structure([[-1, 0, 0]], [1])
    opts = dict(spacing_angstrom=1.0, probe_radius_angstrom=1.4, shell_thickness_angstrom=3.0)
    e1, e2 = csx.build_hydration_grid(one, **opts), csx.build_hydration_grid(one, **opts)
    assert len(e1.sites) > 0 and np.array_equal(e1.sites, e2.sites)  # solvent.rs:688-705
    d = np.linalg.norm(e1.sites - one.xyz[0], axis=1)
    assert np.all(d > 1.70 + 1.4) and np.all(d <= 1.70 + 1.4 + 3.0 + 1e-12)
    ref, vol = O.exterior_grid(one.xyz, one.element)
    assert np.array_equal(e1.sites, ref) and e1.voxel_volume_a3 == vol == 1.0
    s = two_atoms()
    a1, a2 = csx.build_adaptive_hydration_grid(s), csx.build_adaptive_hydration_grid(s)
    assert len(a1.sites) > 0 and np.array_equal(a1.sites, a2.sites)  # solvent.rs:716-726
    assert a1.voxel_volume_a3 == 64.0 and np.array_equal(a1.sites, O.adaptive_grid(s.xyz)[0])
    fast = csx.build_adaptive_hydration_grid(s, spacing_angstrom=8.0)  # solvent.rs:728-738
    assert fast.voxel_volume_a3 == 2.0 * 4.0 ** 3
    assert np.array_equal(fast.sites, O.adaptive_grid(s.xyz, spacing=8.0)[0])
    shift = np.array([3.17, -8.4, 2.6])  # solvent.rs:740-771
    t = structure(s.xyz + shift, s.element)
    for build in (csx.build_adaptive_hydration_grid, csx.build_hydration_grid):
        g0, g1 = build(s).sites, build(t).sites
        assert g0.shape == g1.shape and np.abs(g1 - g0 - shift).max() < 1e-12


@pytest.mark.parametrize("n,radius", [(400, 12.0), (20000, 35.0)])
def test_grids_match_oracle_on_globules(gpu_ctx, n, radius):
    s = csx.generate_globule(n, radius, 1.0, seed=3)
    print(f"grid parity n={n}", flush=True)
    for h in (1.0, 3.0, 4.0, 8.0):
        a = csx.build_adaptive_hydration_grid(s, spacing_angstrom=h)
        ref, vol = O.adaptive_grid(s.xyz, spacing=h)
        assert len(ref) > 0 and np.array_equal(a.sites, ref) and a.voxel_volume_a3 == vol
    for h, probe, shell in ((1.0, 1.4, 3.0), (1.5, 1.2, 2.0), (0.7, 1.4, 3.0)):
        e = csx.build_hydration_grid(s, spacing_angstrom=h, probe_radius_angstrom=probe,
                                     shell_thickness_angstrom=shell)
        ref, vol = O.exterior_grid(s.xyz, s.element, spacing=h, probe=probe, shell=shell)
        assert len(ref) > 0 and np.array_equal(e.sites, ref) and e.voxel_volume_a3 == vol


def test_exterior_grid_leaves_out_enclosed_voids(gpu_ctx):
    # Two nested sealed shells: the gap between them and the core are both voids.
    x1, e1 = G.hollow_sphere(9.0)
    x2, e2 = G.hollow_sphere(17.0)
    s = structure(np.vstack([x1, x2]), np.concatenate([e1, e2]))
    for h in (1.0, 0.6):
        ext = csx.build_hydration_grid(s, spacing_angstrom=h)
        ref, _ = O.exterior_grid(s.xyz, s.element, spacing=h)
        assert np.array_equal(ext.sites, ref)
        assert np.linalg.norm(ext.sites, axis=1).min() > 17.0
    ada = csx.build_adaptive_hydration_grid(s, spacing_angstrom=1.0)
    assert np.array_equal(ada.sites, O.adaptive_grid(s.xyz, spacing=1.0)[0])


def test_grid_edge_cases(gpu_ctx):
    empty = structure(np.zeros((0, 3)), np.zeros(0))
    g = csx.build_hydration_grid(empty, spacing_angstrom=2.0)
    assert g.sites.shape == (0, 3) and g.voxel_volume_a3 == 8.0
    g = csx.build_adaptive_hydration_grid(empty, spacing_angstrom=6.0)
    assert g.sites.shape == (0, 3) and g.voxel_volume_a3 == 2.0 * 27.0
    s = two_atoms()
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(csx.InvalidInput):
            csx.build_hydration_grid(s, spacing_angstrom=bad)
        with pytest.raises(csx.InvalidInput):
            csx.build_adaptive_hydration_grid(s, spacing_angstrom=bad)
    with pytest.raises(csx.InvalidInput):
        csx.build_hydration_grid(s, shell_thickness_angstrom=0.0)
    with pytest.raises(csx.InvalidInput):
        csx.build_adaptive_hydration_grid(s, molecular_clearance_angstrom=0.0)
    with pytest.raises(csx.InvalidInput, match="limit is 250000000"):  # solvent.rs:302-306
        csx.build_hydration_grid(structure([[0, 0, 0], [400, 400, 400]], [1, 1]),
                                 spacing_angstrom=0.5)
    # a single-voxel-thick grid axis and atoms on voxel centres
    line = structure([[0, 0, 0], [2, 0, 0], [4, 0, 0]], [1, 1, 1])
    for h in (0.5, 1.0, 2.0):
        assert np.array_equal(csx.build_hydration_grid(line, spacing_angstrom=h).sites,
                              O.exterior_grid(line.xyz, line.element, spacing=h)[0])
        assert np.array_equal(
            csx.build_adaptive_hydration_grid(line, spacing_angstrom=h, padding_angstrom=0.0).sites,
            O.adaptive_grid(line.xyz, spacing=h, padding=0.0)[0])

"""Seeded synthetic workloads of the benchmark configurations.

  globule_1200      1,200 atoms, R = 16 A, q = 0.005..0.5 step 0.005 (100)
  precision_20k     20,000 atoms, R = 35 A, q = linspace(0, 1, 500)
  prolate_100k      100,000 atoms in a 2:1 prolate spheroid of the volume of a
                    60 A sphere (semi-axes 47.6, 47.6, 95.2 A) plus hydration
                    pseudo-atoms on a 3 A grid 3-6 A from the nearest atom,
                    attached as hydration sites (voxel volume 27 A^3, used
                    with HydrationModel.ExplicitSites); q = linspace(0.005,
                    0.5, 200)
  fit_target_20k    precision_20k structure, 300 q = linspace(0.01, 1, 300),
                    target with scale 2.3, bg 1e-3, c1 1.02, c2 2.0 and
                    Gaussian noise sigma = 1% I + 1e-4 I(0)
  ensemble_512      512 conformers, atom counts uniform in [2000, 30000],
                    alternately a globule (density of the 20k globule) or a
                    2-4 domain chain; q = linspace(0.005, 0.5, 200)
  distinct_shapes   five bodies unlike any globule or domain chain, each of
                    about 6,000 atoms at the 20k globule's density: a rod,
                    a hollow shell, a disk, a dumbbell and a 3:1 prolate
"""
from __future__ import annotations

import numpy as np

from .api import (Element, SolventParams, Structure, generate_domain_chain, generate_globule,
                  generate_shell_sites, solvent_corrected_intensity)

PRECISION_Q = np.linspace(0.0, 1.0, 500)


def globule_1200(seed: int = 0):
    return generate_globule(1200, 16.0, 1.0, seed=seed), np.arange(1, 101) * 0.005


def precision_20k(seed: int = 1):
    return generate_globule(20000, 35.0, 1.0, seed=seed), PRECISION_Q.copy()


SITE_SPACING = 3.0


def prolate_with_sites(n: int, seed: int, radius: float):
    """A 2:1 prolate globule of n atoms (equatorial semi-axis `radius`) with
    its hydration pseudo-atoms (3 A grid, nearest atom 3-6 A away) attached
    as hydration sites of voxel volume 27 A^3."""
    s = generate_globule(n, radius, 2.0, seed=seed)
    sites = generate_shell_sites(s, SITE_SPACING, 3.0, 6.0)
    s.set_hydration_sites(sites, SITE_SPACING ** 3)
    return s, sites


def prolate_100k(seed: int = 2, n: int = 100000):
    s, sites = prolate_with_sites(n, seed, 60.0 / 2 ** (1 / 3))
    return s, np.linspace(0.005, 0.5, 200), len(sites)


def prolate_points_100k(seed: int = 2, n: int = 100000):
    """The prolate_100k points with the pseudo-atoms as oxygen atoms (one
    vacuum point set of atoms and sites together)."""
    s, sites = prolate_with_sites(n, seed, 60.0 / 2 ** (1 / 3))
    xyz = np.vstack([s.xyz, sites])
    el = np.concatenate([s.element, np.full(len(sites), int(Element.O), dtype=np.uint8)])
    return Structure.from_arrays(xyz, el), np.linspace(0.005, 0.5, 200), len(sites)


def fit_target_20k(seed: int = 1, noise_seed: int = 7, ctx=None):
    s, _ = precision_20k(seed)
    q = np.linspace(0.01, 1.0, 300)
    truth = SolventParams(c1=1.02, c2=2.0)
    clean = 2.3 * solvent_corrected_intensity(s, q, truth, ctx=ctx) + 1e-3
    i0 = clean[0]
    sigma = 0.01 * clean + 1e-4 * i0
    rng = np.random.default_rng(noise_seed)
    return s, q, clean + sigma * rng.normal(size=q.size), sigma


def ensemble_512(m: int = 512, seed: int = 4, lo: int = 2000, hi: int = 30000):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=m)
    conformers = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            radius = 35.0 * (n / 20000.0) ** (1 / 3)
            conformers.append(generate_globule(int(n), radius, 1.0, seed=seed * 100003 + k))
        else:
            conformers.append(generate_domain_chain(int(n), int(2 + k % 3), seed=seed * 100003 + k))
    return conformers, np.linspace(0.005, 0.5, 200)


# atoms per cubic angstrom of the 20,000-atom, 35 A globule
GLOBULE_DENSITY = 20000 / (4.0 / 3.0 * np.pi * 35.0 ** 3)


def distinct_shapes(seed: int = 9):
    """Rod (r 8 A, 270 A), hollow shell (30 / 24 A), disk (r 35 A, 12 A
    thick), dumbbell (two 19 A spheres 80 A apart) and prolate (semi-axes
    17, 17, 51 A): points uniform in each body at the globule density, with
    the element mix of a generated globule."""
    rng = np.random.default_rng(seed)

    def fill(inside, box, volume):
        n = int(round(volume * GLOBULE_DENSITY))
        pts = np.empty((0, 3))
        lo, hi = np.asarray(box[0], float), np.asarray(box[1], float)
        while len(pts) < n:
            c = lo + (hi - lo) * rng.random((2 * n, 3))
            pts = np.concatenate([pts, c[inside(c)]])
        pts = pts[:n]
        el = generate_globule(n, 35.0 * (n / 20000.0) ** (1 / 3), 1.0,
                              seed=int(rng.integers(1 << 30))).element
        return Structure.from_arrays(pts, el)

    r2 = lambda c: (c ** 2).sum(axis=1)
    rod = fill(lambda c: c[:, 0] ** 2 + c[:, 1] ** 2 <= 64.0, ([-8, -8, -135], [8, 8, 135]),
               np.pi * 64.0 * 270.0)
    shell = fill(lambda c: (r2(c) <= 900.0) & (r2(c) >= 576.0), ([-30] * 3, [30] * 3),
                 4.0 / 3.0 * np.pi * (30.0 ** 3 - 24.0 ** 3))
    disk = fill(lambda c: c[:, 0] ** 2 + c[:, 1] ** 2 <= 35.0 ** 2, ([-35, -35, -6], [35, 35, 6]),
                np.pi * 35.0 ** 2 * 12.0)
    off = np.array([0.0, 0.0, 40.0])
    bell = fill(lambda c: (r2(c - off) <= 361.0) | (r2(c + off) <= 361.0),
                ([-19, -19, -59], [19, 19, 59]), 2 * 4.0 / 3.0 * np.pi * 19.0 ** 3)
    prolate = fill(lambda c: (c[:, 0] ** 2 + c[:, 1] ** 2) / 289.0 + c[:, 2] ** 2 / 2601.0 <= 1.0,
                   ([-17, -17, -51], [17, 17, 51]), 4.0 / 3.0 * np.pi * 17.0 * 17.0 * 51.0)
    return [rod, shell, disk, bell, prolate]

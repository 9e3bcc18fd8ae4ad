"""The C ABI is called concurrently, as the reference's seam is from rayon
workers (ensemble.rs:244-256): threads sharing structure handles (point sets,
grids, multipole geometry built on first use under the structure's locks,
streams leased from the context) give bitwise the sequential results."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import crabsaxs_amd as csx

pytestmark = pytest.mark.gpu


def structures():
    return [csx.generate_globule(n, 16.0 * (n / 1200) ** (1 / 3), 1.0, seed=40 + k)
            for k, n in enumerate((700, 1500, 2600))]


def work(s):
    q = np.linspace(0.01, 0.5, 48)
    p = csx.SolventParams(c1=1.01, c2=0.5)
    grid = csx.SolventParams(c1=1.0, c2=1.0, hydration_model=csx.HydrationModel.AdaptiveGrid)
    return [csx.debye_intensity(s, q),
            csx.solvent_corrected_intensity(s, q, p),
            csx.solvent_corrected_intensity(s, q, grid),
            csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=20)),
            csx.build_hydration_grid(s, spacing_angstrom=1.5).sites,
            csx.coordinate_features(s, 64).p_r.p]


def test_threads_sharing_structures_match_sequential(gpu_ctx):
    sequential = [work(s) for s in structures()]
    shared = structures()  # fresh: every cache is built while the threads race
    jobs = [k % len(shared) for k in range(12)]
    with ThreadPoolExecutor(max_workers=8) as pool:
        got = list(pool.map(lambda k: (k, work(shared[k])), jobs))
    for k, res in got:
        for a, b in zip(res, sequential[k]):
            assert np.array_equal(a, b)


def test_reused_and_regrown_scratch_give_identical_results():
    """A context keeps each stream's temporaries and grows them on demand: a
    small call, a larger one (the buffers are regrown) and the small call
    again give bitwise the first result, on a fresh context and its destroy."""
    ctx = csx.Context(0)
    try:
        small = csx.generate_globule(600, 12.0, 1.0, seed=50)
        large = csx.generate_globule(5000, 26.0, 1.0, seed=51)
        q_small = np.linspace(0.01, 0.4, 24)
        q_large = np.linspace(0.0, 0.8, 300)

        def run(s, q):
            return [csx.debye_intensity(s, q, ctx=ctx),
                    csx.multipole_intensity(s, q, csx.MultipoleParams(l_max=30), ctx=ctx),
                    csx.coordinate_features(s, 48, ctx=ctx).p_r.p]

        first = run(small, q_small)
        run(large, q_large)
        again = run(small, q_small)
        for a, b in zip(first, again):
            assert np.array_equal(a, b)
    finally:
        ctx.close()

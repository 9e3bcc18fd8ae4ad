"""Brute-force restatement of the two hydration grids (solvent.rs:121-223,
259-401) for pinning the oracle, and the structures the grid tests share.

The brute force tests every voxel centre against every atom in numpy and
finds the exterior as the six-connected components of unknown voxels that
touch the grid boundary (scipy.ndimage.label), an independent route to the
reference's breadth-first flood fill."""
import numpy as np
from scipy import ndimage

# van_der_waals_radius_for (solvent.rs:235-253), in element-code order
# H C N O S P Na Mg Cl Ca Fe Zn K I Other.
VDW = np.array([1.20, 1.70, 1.55, 1.52, 1.80, 1.80, 2.27, 1.73, 1.75, 2.31, 1.72, 1.39, 2.75,
                1.98, 1.70])


def _box(xyz, margin, h):
    lo = xyz.min(axis=0) - margin
    hi = xyz.max(axis=0) + margin
    dims = np.ceil((hi - lo) / h - 1e-10).astype(np.int64) + 1
    return lo, dims


def _centres(lo, dims, h):
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]),
                          indexing="ij")
    return np.stack([lo[0] + x * h, lo[1] + y * h, lo[2] + z * h], axis=-1)


def _within(centres, xyz, radii):
    out = np.zeros(centres.shape[:3], bool)
    for p, r in zip(xyz, radii):
        d = centres - p
        d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
        out |= d2 <= r * r + 1e-10
    return out


def adaptive_grid(xyz, spacing=4.0, clearance=3.0, padding=12.0):
    xyz = np.asarray(xyz, float).reshape(-1, 3)
    n = len(xyz)
    c = [0.0, 0.0, 0.0]
    for p in xyz:
        for a in range(3):
            c[a] += p[a]
    c = [v / n for v in c]
    ss = 0.0
    for p in xyz:
        d = [p[a] - c[a] for a in range(3)]
        ss += d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    rg = (ss / n) ** 0.5
    width = 3.0 if rg <= 15.0 else 5.0 if rg >= 20.0 else 3.0 + 0.4 * (rg - 15.0)
    lo, dims = _box(xyz, padding, spacing)
    cen = _centres(lo, dims, spacing)
    keep = (_within(cen, xyz, np.full(n, clearance + width))
            & ~_within(cen, xyz, np.full(n, clearance)))
    return cen[keep]


def exterior_grid(xyz, element, spacing=1.0, probe=1.4, shell=3.0):
    xyz = np.asarray(xyz, float).reshape(-1, 3)
    r = VDW[np.asarray(element, int)]
    lo, dims = _box(xyz, r.max() + probe + shell + spacing, spacing)
    cen = _centres(lo, dims, spacing)
    unknown = ~_within(cen, xyz, r + probe)
    lab, _ = ndimage.label(unknown)  # six-connected in 3-D
    faces = np.concatenate([lab[0].ravel(), lab[-1].ravel(), lab[:, 0].ravel(),
                            lab[:, -1].ravel(), lab[:, :, 0].ravel(), lab[:, :, -1].ravel()])
    exterior = np.isin(lab, np.unique(faces[faces > 0]))
    keep = exterior & _within(cen, xyz, r + (probe + shell))
    return cen[keep]


def globule(n, radius, seed):
    """Random atoms in a ball (C/N/O/S), for grid tests."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v *= (radius * rng.random(n) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
    return v, rng.choice([1, 2, 3, 4], size=n, p=[0.6, 0.2, 0.18, 0.02]).astype(np.uint8)


def hollow_sphere(radius, step=1.5, centre=(0.0, 0.0, 0.0)):
    """Carbon atoms on a Fibonacci sphere about `step` apart: the probe-expanded
    spheres seal the interior, which is then an enclosed void."""
    m = int(4 * np.pi * radius ** 2 / step ** 2)
    k = np.arange(m) + 0.5
    z = 1 - 2 * k / m
    phi = np.pi * (1 + 5 ** 0.5) * k
    s = np.sqrt(1 - z * z)
    xyz = radius * np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1) + np.asarray(centre)
    return xyz, np.ones(m, np.uint8)

"""ctypes access to the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.  Builds the oracle with `make` on first use
when the shared library is missing.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB = ORACLE_DIR / "liboracle.so"

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.or_sinc.restype = C.c_double
        L.or_sinc.argtypes = [C.c_double]
        L.or_united_atom_group.argtypes = [C.c_char_p, C.c_char_p]
        L.or_form_factor.restype = C.c_double
        L.or_form_factor.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
        L.or_excluded_volume.restype = C.c_double
        L.or_excluded_volume.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_water_form_factor.restype = C.c_double
        L.or_water_form_factor.argtypes = [C.c_double]
        L.or_excluded_volume_scale.restype = C.c_double
        L.or_excluded_volume_scale.argtypes = [C.c_double, C.c_double]
        L.or_debye_intensity.argtypes = [_d, _i, _i, _d, C.c_size_t, _d, C.c_size_t, C.c_int,
                                         C.c_int, _d]
        L.or_debye_solvent_basis.argtypes = [_d, _i, _i, _d, C.c_size_t, _d, C.c_size_t,
                                             C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_int, _d, C.c_size_t, C.c_double, C.c_int,
                                             C.c_int, _d]
        L.or_combine_solvent_basis.argtypes = [_d, _d, C.c_size_t, C.c_double, C.c_double, _d]
        L.or_adaptive_grid.restype = C.c_long
        L.or_adaptive_grid.argtypes = [_d, C.c_size_t, C.c_double, C.c_double, C.c_double, _d,
                                       C.c_size_t, _d]
        L.or_exterior_grid.restype = C.c_long
        L.or_exterior_grid.argtypes = [_d, _i, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                       _d, C.c_size_t, _d]
        L.or_real_harmonics.argtypes = [C.c_int, C.c_double, C.c_double, _d]
        L.or_bessel_reference.argtypes = [C.c_int, C.c_double, _d]
        L.or_bessel_exact.argtypes = [C.c_int, C.c_double, _d]
        L.or_multipole_intensity.argtypes = [_d, _i, _i, _d, C.c_size_t, _d, C.c_size_t, C.c_int,
                                             C.c_int, C.c_int, C.c_int, _d]
        L.or_multipole_solvent.argtypes = [_d, _i, _i, _d, C.c_size_t, _d, C.c_size_t, C.c_int,
                                           C.c_int, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, _d, C.c_size_t, C.c_double,
                                           C.c_int, C.c_int, C.c_int, _d]
        L.or_solve_scale_background.argtypes = [_d, _d, _d, C.c_size_t, C.c_int, C.c_double,
                                                C.c_int, C.c_double, _d]
        L.or_cubic_hermite.argtypes = [_d, _d, C.c_size_t, _d, C.c_size_t, _d]
        L.or_sampled_indices.restype = C.c_size_t
        L.or_sampled_indices.argtypes = [C.c_size_t, C.c_size_t, C.POINTER(C.c_size_t)]
        L.or_optimize_weights.argtypes = [_d, C.c_size_t, _d, _d, C.c_size_t, C.c_int,
                                          C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                          _d, _d]
        L.or_project_simplex.argtypes = [_d, C.c_size_t, _d]
        L.or_debye_grad_x.argtypes = [_d, _i, _i, _d, C.c_size_t, _d, C.c_size_t, _d, C.c_int,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, _d, _d]
        L.or_coordinate_features.argtypes = [_d, _i, _i, _d, C.c_size_t, C.c_size_t, _d, _d, _d]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ii(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _p(a):
    return a.ctypes.data_as(_d)


def _pi(a):
    return a.ctypes.data_as(_i)


class Atoms:
    """SoA atom arrays for the oracle."""

    def __init__(self, xyz, element, group=None, occ=None):
        self.xyz = _f(xyz).reshape(-1, 3)
        n = self.xyz.shape[0]
        self.element = _ii(element).reshape(n)
        self.group = _ii(np.full(n, -1) if group is None else group).reshape(n)
        self.occ = _f(np.ones(n) if occ is None else occ).reshape(n)

    def __len__(self):
        return self.xyz.shape[0]

    def args(self):
        return (_p(self.xyz), _pi(self.element), _pi(self.group), _p(self.occ), len(self))


def debye_intensity(atoms: Atoms, q, implicit_h=True, threads=0):
    q = _f(q)
    out = np.zeros(q.size)
    lib().or_debye_intensity(*atoms.args(), _p(q), q.size, int(implicit_h), threads, _p(out))
    return out


def debye_solvent_basis(atoms: Atoms, q, c1=1.0, c2=0.0, thickness=3.0, rho=0.334,
                        sites=None, voxel_volume=0.0, implicit_h=True, threads=0):
    q = _f(q)
    out = np.zeros(6 * q.size)
    grid = sites is not None
    s = _f(np.zeros((0, 3)) if sites is None else sites)
    lib().or_debye_solvent_basis(*atoms.args(), _p(q), q.size, c1, c2, thickness, rho, int(grid),
                                 _p(s), s.shape[0] if s.ndim == 2 else 0, voxel_volume,
                                 int(implicit_h), threads, _p(out))
    return out.reshape(6, q.size)


def debye_grad_x(atoms: Atoms, q, weights=None, solvent=None, implicit_h=True, threads=0,
                 jacobian=False):
    """solvent: None (vacuum) or (c1, c2, thickness, rho)."""
    q = _f(q)
    w = _f(np.ones(q.size) if weights is None else weights)
    n = len(atoms)
    grad = np.zeros(3 * n)
    jac = np.zeros(q.size * 3 * n) if jacobian else None
    c1, c2, t, rho = solvent if solvent is not None else (1.0, 0.0, 3.0, 0.334)
    lib().or_debye_grad_x(*atoms.args(), _p(q), q.size, _p(w), int(solvent is not None), c1, c2,
                          t, rho, int(implicit_h), threads, _p(grad),
                          _p(jac) if jacobian else None)
    if jacobian:
        return jac.reshape(q.size, n, 3)
    return grad.reshape(n, 3)


def coordinate_features(atoms: Atoms, bins=160):
    """(rg, dmax, unnormalised histogram or None) of coordinate_features /
    coordinate_pr (metrics.rs:236-270, 475-548)."""
    nb = max(int(bins), 16)
    hist = np.zeros(nb)
    rg, dmax = C.c_double(), C.c_double()
    lib().or_coordinate_features(*atoms.args(), nb, C.byref(rg), C.byref(dmax), _p(hist))
    return rg.value, dmax.value, (hist if dmax.value > 0 else None)


def normalize_distribution(r, p):
    """normalize_distribution (metrics.rs:550-584): (r, p, rg, dmax) or None."""
    r, p = list(map(float, r)), list(map(float, p))
    if len(r) != len(p) or not p:
        return None
    area = 0.0
    for k in range(len(r) - 1):
        area += 0.5 * (r[k + 1] - r[k]) * (p[k] + p[k + 1])
    if not np.isfinite(area) or area <= 0.0:
        return None
    p = [v / area for v in p]
    moment = 0.0
    for k in range(len(r) - 1):
        moment += 0.5 * (r[k + 1] - r[k]) * (r[k] * r[k] * p[k] + r[k + 1] * r[k + 1] * p[k + 1])
    rg = float(np.sqrt(max(moment / 2.0, 0.0)))
    maximum = 0.0
    for v in p:
        maximum = max(maximum, v)
    dmax = next((r[k] for k in range(len(p) - 1, -1, -1) if p[k] > maximum * 0.01), None)
    return np.array(r), np.array(p), (rg if np.isfinite(rg) else None), dmax


def combine_solvent_basis(basis, q, c1, c2):
    q = _f(q)
    b = _f(basis).reshape(6, q.size)
    out = np.zeros(q.size)
    lib().or_combine_solvent_basis(_p(b), _p(q), q.size, c1, c2, _p(out))
    return out


def adaptive_grid(xyz, spacing=4.0, clearance=3.0, padding=12.0):
    xyz = _f(xyz).reshape(-1, 3)
    vol = C.c_double()
    n = lib().or_adaptive_grid(_p(xyz), xyz.shape[0], spacing, clearance, padding, None, 0,
                               C.byref(vol))
    if n < 0:
        raise ValueError("invalid grid input")
    sites = np.zeros((n, 3))
    lib().or_adaptive_grid(_p(xyz), xyz.shape[0], spacing, clearance, padding, _p(sites), n,
                           C.byref(vol))
    return sites, vol.value


def exterior_grid(xyz, element, spacing=1.0, probe=1.4, shell=3.0):
    xyz = _f(xyz).reshape(-1, 3)
    el = _ii(element)
    vol = C.c_double()
    n = lib().or_exterior_grid(_p(xyz), _pi(el), xyz.shape[0], spacing, probe, shell, None, 0,
                               C.byref(vol))
    if n < 0:
        raise ValueError("invalid grid input")
    sites = np.zeros((n, 3))
    lib().or_exterior_grid(_p(xyz), _pi(el), xyz.shape[0], spacing, probe, shell, _p(sites), n,
                           C.byref(vol))
    return sites, vol.value


def real_harmonics(l_max, cos_theta, phi):
    out = np.zeros((l_max + 1) ** 2)
    lib().or_real_harmonics(l_max, cos_theta, phi, _p(out))
    return out


def bessel_reference(l_max, x):
    out = np.zeros(l_max + 1)
    lib().or_bessel_reference(l_max, x, _p(out))
    return out


def bessel_exact(l_max, x):
    out = np.zeros(l_max + 1)
    lib().or_bessel_exact(l_max, x, _p(out))
    return out


def multipole_intensity(atoms: Atoms, q, l_max, implicit_h=True, exact=True, threads=0):
    q = _f(q)
    out = np.zeros(q.size)
    lib().or_multipole_intensity(*atoms.args(), _p(q), q.size, l_max, int(implicit_h),
                                 int(exact), threads, _p(out))
    return out


def multipole_solvent(atoms: Atoms, q, l_max, c1=1.0, c2=0.0, thickness=3.0, rho=0.334,
                      sites=None, voxel_volume=0.0, basis=False, implicit_h=True, exact=True,
                      threads=0):
    q = _f(q)
    out = np.zeros((6 if basis else 1) * q.size)
    grid = sites is not None
    s = _f(np.zeros((0, 3)) if sites is None else sites).reshape(-1, 3)
    lib().or_multipole_solvent(*atoms.args(), _p(q), q.size, l_max, int(implicit_h), c1, c2,
                               thickness, rho, int(grid), _p(s), s.shape[0], voxel_volume,
                               int(basis), int(exact), threads, _p(out))
    return out.reshape(6, q.size) if basis else out


def solve_scale_background(theory, intensity, sigma, scale_mode=0, scale_value=0.0, bg_mode=0,
                           bg_value=0.0):
    t, y, s = _f(theory), _f(intensity), _f(sigma)
    out = np.zeros(3)
    lib().or_solve_scale_background(_p(t), _p(y), _p(s), t.size, scale_mode, scale_value,
                                    bg_mode, bg_value, _p(out))
    return tuple(out)


def cubic_hermite(x, y, targets):
    x, y, t = _f(x), _f(y), _f(targets)
    out = np.zeros(t.size)
    lib().or_cubic_hermite(_p(x), _p(y), x.size, _p(t), t.size, _p(out))
    return out


def sampled_indices(length, stride):
    out = (C.c_size_t * (length + 1))()
    n = lib().or_sampled_indices(length, stride, out)
    return [out[i] for i in range(n)]


def optimize_weights(curves, intensity, sigma, uniform=False, regularization=0.0,
                     scale_mode=0, scale_value=0.0, bg_mode=0, bg_value=0.0):
    c = _f(curves)
    m, nq = c.shape
    y, s = _f(intensity), _f(sigma)
    w = np.zeros(m)
    out = np.zeros(3)
    lib().or_optimize_weights(_p(c), m, _p(y), _p(s), nq, scale_mode, scale_value, bg_mode,
                              bg_value, int(uniform), regularization, _p(w), _p(out))
    return w, tuple(out)


def project_simplex(values):
    v = _f(values)
    out = np.zeros(v.size)
    lib().or_project_simplex(_p(v), v.size, _p(out))
    return out

"""Host-side mirror of crabSAXS's profile API over the C ABI.

Names, argument meaning and error behaviour follow the reference's public
Rust API (src/scattering.rs, src/solvent.rs, src/spherical.rs, src/fit.rs,
src/ensemble.rs, src/analysis.rs); all arithmetic of the profile path runs in
libcrabsaxs_hip.so on the GPU.  Errors raise SaxsError subclasses mirroring
`SaxsError` (src/error.rs:4-25).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import enum
import threading
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ComputeOptionsC, SolventParamsC, dptr, f64


# ------------------------------------------------------------------ errors --

class SaxsError(Exception):
    """Base error, `SaxsError` (src/error.rs:4-25)."""


class InvalidInput(SaxsError):
    pass


class FitDidNotConverge(SaxsError):
    pass


class NoDevice(SaxsError):
    pass


class DeviceError(SaxsError):
    pass


class OutOfMemory(SaxsError):
    pass


class ParseError(SaxsError):
    """`SaxsError::StructureParse` / `SaxsError::ExperimentalDataParse`."""


class UnsupportedElement(SaxsError):
    pass


class IoError(SaxsError):
    pass


_STATUS = {1: InvalidInput, 2: FitDidNotConverge, 3: NoDevice, 4: DeviceError, 5: OutOfMemory,
           6: ParseError, 7: UnsupportedElement, 8: IoError}


def _check(status: int) -> None:
    if status != 0:
        msg = _lib.load().csx_last_error().decode()
        raise _STATUS.get(status, SaxsError)(msg)


# ------------------------------------------------------------------- enums --

class Element(enum.IntEnum):
    """`Element` (src/structure.rs:90-106); codes match include/crabsaxs_tables.h."""
    H = 0
    C = 1
    N = 2
    O = 3
    S = 4
    P = 5
    Na = 6
    Mg = 7
    Cl = 8
    Ca = 9
    Fe = 10
    Zn = 11
    K = 12
    I = 13
    Other = 14

    @classmethod
    def from_symbol(cls, symbol: str) -> "Element":
        return cls(_lib.load().csx_element_from_symbol(symbol.encode()))


class HydrogenMode(enum.Enum):
    Implicit = "Implicit"
    Explicit = "Explicit"


class ScatteringMethod(enum.IntEnum):
    Auto = 0
    Debye = 1
    Multipole = 2


class HydrationModel(enum.IntEnum):
    """`HydrationModel` (src/solvent.rs:17-25), plus ExplicitSites (no
    reference counterpart): the sites given with Structure.set_hydration_sites,
    entering like the sites of a Grid model."""
    Analytic = 0
    Grid = 1
    AdaptiveGrid = 2
    ExplicitSites = 3


class Precision(enum.IntEnum):
    """Pair-sum arithmetic (no reference counterpart: the reference is FP64)."""
    FP64_PLAIN = 0
    FP32_COMPENSATED = 1


# --------------------------------------------------------------- structure --

@dataclass
class Atom:
    """`Atom` (src/structure.rs:151-160)."""
    pos: Sequence[float]
    element: Element = Element.C
    occupancy: float = 1.0
    is_hetatm: bool = False
    atom_name: str = "C"
    residue_name: str = "UNK"
    chain_id: str = "A"
    residue_id: int = 1


class AltlocPolicy(enum.IntEnum):
    """`AltlocPolicy` (src/structure.rs:13-21)."""
    Highest = 0
    First = 1
    All = 2


class OccupancyPolicy(enum.IntEnum):
    """`OccupancyPolicy` (src/structure.rs:166-175)."""
    Reported = 0
    SelectedConformerFull = 1


class ResidueKind(enum.IntEnum):
    """`ResidueKind` (src/structure.rs:24-33)."""
    Polymer = 0
    Glycan = 1
    Water = 2
    Other = 3


def residue_kind(name: str) -> ResidueKind:
    """`residue_kind` (src/structure.rs:34-88)."""
    return ResidueKind(_lib.load().csx_residue_kind(name.encode()))


@dataclass
class ParseOptions:
    """`ParseOptions` (src/structure.rs:177-202)."""
    include_hetatm: bool = False
    include_waters: bool = False
    include_glycans: bool = True
    model_index: int = 0
    allow_unknown_elements: bool = False
    occupancy_policy: OccupancyPolicy = OccupancyPolicy.Reported
    altloc_policy: AltlocPolicy = AltlocPolicy.Highest
    chains: Optional[Sequence[str]] = None

    def _c(self) -> "_lib.ParseOptionsC":
        return _lib.ParseOptionsC(
            int(self.include_hetatm), int(self.include_waters), int(self.include_glycans),
            int(self.model_index), int(self.allow_unknown_elements), int(self.occupancy_policy),
            int(self.altloc_policy),
            None if self.chains is None else ",".join(self.chains).encode())


_handle_lock = threading.Lock()  # creation of Structure device handles


def _read_atoms(path, options: Optional[ParseOptions]):
    h = C.c_void_p()
    opts = (options or ParseOptions())._c()
    _check(_lib.load().csx_read_structure_file(str(path).encode(), C.byref(opts), C.byref(h)))
    return h


class Structure:
    """`Structure` (src/structure.rs:163-165), held as SoA arrays.

    Built from `Atom`s (typed by residue/atom names exactly like
    formfactor.rs:357-417) or directly from arrays.  The device copy is made
    on first use and reused; treat a Structure as immutable once computed on.
    """

    def __init__(self, atoms: Sequence[Atom] = ()):
        atoms = list(atoms)
        lib = _lib.load()
        self.xyz = f64([a.pos for a in atoms]).reshape(-1, 3)
        self.element = np.array([int(a.element) for a in atoms], dtype=np.uint8)
        self.group = np.array(
            [lib.csx_united_atom_group(a.residue_name.encode(), a.atom_name.encode())
             for a in atoms], dtype=np.int8)
        self.occupancy = f64([a.occupancy for a in atoms])
        self.atoms = atoms
        self._handles = {}
        self.sites = None

    @classmethod
    def from_pdb_file(cls, path, options: "ParseOptions" = None) -> "Structure":
        """`Structure::from_pdb_file[_with_options]` (src/structure.rs:205-326):
        PDB or mmCIF."""
        return cls._from_atoms_handle(_read_atoms(path, options))

    from_pdb_file_with_options = from_pdb_file

    @classmethod
    def from_pdb_file_models_with_options(cls, path, options: "ParseOptions" = None):
        """Every model of a file (src/structure.rs:330-346)."""
        options = options or ParseOptions()
        models = []
        index = 0
        while True:
            try:
                models.append(cls.from_pdb_file(path, replace(options, model_index=index)))
            except SaxsError:
                if index > 0 and models:
                    break
                raise
            index += 1
        return models

    @classmethod
    def parse(cls, text: str, format: int = 0, options: "ParseOptions" = None) -> "Structure":
        """A PDB (format 1) or mmCIF (2) text; 0 detects the format."""
        h = C.c_void_p()
        data = text.encode()
        opts = (options or ParseOptions())._c()
        _check(_lib.load().csx_parse_structure(data, len(data), format, C.byref(opts),
                                               C.byref(h)))
        return cls._from_atoms_handle(h)

    @classmethod
    def _from_atoms_handle(cls, h) -> "Structure":
        lib = _lib.load()
        try:
            n = lib.csx_atoms_count(h)
            recs = (_lib.AtomRecordC * max(n, 1))()
            _check(lib.csx_atoms_get(h, recs))
        finally:
            lib.csx_atoms_destroy(h)
        atoms = [Atom(pos=(r.x, r.y, r.z), element=Element(r.element), occupancy=r.occupancy,
                      is_hetatm=bool(r.is_hetatm), atom_name=r.atom_name.decode(),
                      residue_name=r.residue_name.decode(), chain_id=r.chain_id.decode(),
                      residue_id=int(r.residue_id)) for r in recs[:n]]
        return cls(atoms)

    @classmethod
    def from_arrays(cls, xyz, element, occupancy=None, group=None) -> "Structure":
        s = cls.__new__(cls)
        s.xyz = f64(xyz).reshape(-1, 3)
        n = s.xyz.shape[0]
        s.element = np.ascontiguousarray(np.asarray(element, dtype=np.uint8).reshape(n))
        s.occupancy = f64(np.ones(n) if occupancy is None else occupancy).reshape(n)
        s.group = np.ascontiguousarray(
            np.full(n, -1, dtype=np.int8) if group is None else np.asarray(group, dtype=np.int8))
        s.atoms = None
        s._handles = {}
        s.sites = None
        return s

    def set_hydration_sites(self, sites, voxel_volume: float) -> None:
        """Hydration sites (n x 3, A) of one voxel volume (A^3) used with
        HydrationModel.ExplicitSites, in place of the sites a Grid or
        AdaptiveGrid model builds (no reference counterpart)."""
        xyz = f64(sites).reshape(-1, 3)
        self.sites = (xyz, float(voxel_volume))
        for _, h in list(self._handles.values()):
            self._attach_sites(h)

    def _attach_sites(self, h) -> None:
        xyz, vol = self.sites
        _check(_lib.load().csx_structure_set_hydration_sites(h, dptr(xyz), xyz.shape[0], vol))

    def __len__(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def atoms_count(self) -> int:
        return len(self)

    def handle(self, ctx: "Context" = None):
        """The device copy on `ctx`, uploaded on first use.  One handle per
        context, created under a lock and kept until the structure is freed,
        so threads may share a structure."""
        ctx = ctx or default_context()
        if not ctx.ptr:
            raise InvalidInput("the context has been closed")
        entry = self._handles.get(id(ctx))
        if entry is not None and entry[0] is ctx:
            return entry[1]
        with _handle_lock:
            entry = self._handles.get(id(ctx))
            if entry is not None and entry[0] is ctx:
                return entry[1]
            h = C.c_void_p()
            _check(_lib.load().csx_structure_create(
                ctx.ptr, len(self), dptr(self.xyz),
                self.element.ctypes.data_as(C.POINTER(C.c_uint8)),
                self.group.ctypes.data_as(C.POINTER(C.c_int8)), dptr(self.occupancy),
                C.byref(h)))
            self._handles[id(ctx)] = (ctx, h)
            if getattr(self, "sites", None) is not None:
                self._attach_sites(h)
            return h

    def _free(self):
        handles = getattr(self, "_handles", {})
        for _, h in handles.values():
            _lib.load().csx_structure_destroy(h)
        handles.clear()

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


# ----------------------------------------------------------------- context --

class Context:
    """One HIP device; owns streams and kernel timing statistics."""

    DEFAULT_CACHE_LIMIT = 256 << 20

    def __init__(self, device: int = 0):
        self.ptr = C.c_void_p()
        _check(_lib.load().csx_ctx_create(device, C.byref(self.ptr)))
        self.device = device
        self.cache_limit = self.DEFAULT_CACHE_LIMIT

    def set_timing(self, enabled: bool) -> None:
        _check(_lib.load().csx_ctx_set_timing(self.ptr, int(enabled)))

    def reset_timing(self) -> None:
        _check(_lib.load().csx_ctx_reset_timing(self.ptr))

    def kernel_time(self, name: str):
        n, ms = C.c_int64(), C.c_double()
        _check(_lib.load().csx_ctx_kernel_time(self.ptr, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def set_tuning(self, q_tile: int = 0, sin_mode: int = 3, j_chunk: int = 0) -> None:
        """Pair-kernel launch shape; q_tile 0 picks the tile per call and
        j_chunk 0 the j chunk per structure; sin_mode 3 (the default) runs the
        sine recurrence on uniform q grids where it models faster, 2 always
        evaluates the hardware sine, 4 interpolates j_0(q r) radially over
        the pair distances (8 Chebyshev nodes per distance bin, order-free
        fixed-point sums)."""
        _check(_lib.load().csx_ctx_set_tuning(self.ptr, q_tile, sin_mode, j_chunk))

    def set_cache_limit(self, nbytes: int = DEFAULT_CACHE_LIMIT) -> None:
        """Bytes of cached pair tables and multipole coefficients kept per
        structure and solvent model (0: no caching)."""
        _check(_lib.load().csx_ctx_set_cache_limit(self.ptr, int(nbytes)))
        self.cache_limit = int(nbytes)

    @contextlib.contextmanager
    def caching(self, nbytes: int):
        """Run a block with the given cache limit (0: every call computes),
        then restore the previous one."""
        prev = self.cache_limit
        self.set_cache_limit(nbytes)
        try:
            yield self
        finally:
            self.set_cache_limit(prev)

    def set_workspace_limit(self, nbytes: int = 2 << 30) -> None:
        """Bytes of adjoint partial sums per launch (larger runs split)."""
        _check(_lib.load().csx_ctx_set_workspace_limit(self.ptr, int(nbytes)))

    def last_launch(self) -> dict:
        """Shape of the most recent pair-kernel launch on this context."""
        v = [C.c_int() for _ in range(3)]
        _check(_lib.load().csx_ctx_last_launch(self.ptr, *(C.byref(x) for x in v)))
        return {"q_tile": v[0].value, "sin_mode": v[1].value, "persistent_workgroups": v[2].value}

    def last_pair_items(self) -> dict:
        """j chunk and work items of the most recent pair-kernel launch."""
        v = [C.c_int() for _ in range(2)]
        _check(_lib.load().csx_ctx_last_pair_items(self.ptr, *(C.byref(x) for x in v)))
        return {"j_chunk": v[0].value, "items": v[1].value}

    def cu_count(self) -> int:
        n = C.c_int()
        _check(_lib.load().csx_ctx_device_info(self.ptr, C.byref(n), None, 0))
        return int(n.value)

    def arch(self) -> str:
        buf = C.create_string_buffer(256)
        _check(_lib.load().csx_ctx_device_info(self.ptr, None, buf, 256))
        return buf.value.decode()

    def set_vjp_tile(self, q_tile: int = 0) -> None:
        """q values per sweep of the coordinate adjoint; 0 picks per call."""
        _check(_lib.load().csx_ctx_set_vjp_tile(self.ptr, q_tile))

    def set_multipole_group(self, tiles: int = 0) -> None:
        """Column tiles per multipole tile group (32, 64, 128); 0 picks per call."""
        _check(_lib.load().csx_ctx_set_multipole_group(self.ptr, tiles))

    def set_multipole_radial(self, mode: int = 0) -> None:
        """Radial interpolation of the multipole coefficients: 0 automatic,
        1 every point, 2 always (8 Chebyshev nodes per radial bin)."""
        _check(_lib.load().csx_ctx_set_multipole_radial(self.ptr, mode))

    def last_multipole(self) -> dict:
        """Radial bin width (0: per point) and points or nodes contracted by
        the last multipole coefficient computation."""
        h, n = C.c_double(), C.c_int()
        _check(_lib.load().csx_ctx_last_multipole(self.ptr, C.byref(h), C.byref(n)))
        return {"bin_width": h.value, "points": n.value}

    def set_vjp_split(self, split: int = 0) -> None:
        """Workgroups per sweep of the coordinate adjoint; 0 picks per call."""
        _check(_lib.load().csx_ctx_set_vjp_split(self.ptr, split))

    def last_vjp_launch(self) -> dict:
        """Shape of the most recent coordinate-adjoint call on this context."""
        v = [C.c_int() for _ in range(4)]
        _check(_lib.load().csx_ctx_last_vjp_launch(self.ptr, *(C.byref(x) for x in v)))
        return {"q_tile": v[0].value, "q_split": v[1].value, "step_split": v[2].value,
                "launches": v[3].value}

    def set_persistence(self, workgroups_per_cu: int = -1) -> None:
        """Pair-kernel launch form: a persistent grid of this many workgroups
        per CU, 0 for one workgroup per task, -1 (default) to choose per call."""
        _check(_lib.load().csx_ctx_set_persistence(self.ptr, workgroups_per_cu))

    def close(self):
        """Free the context and the device temporaries it keeps.  Structures
        used on it may outlive it but are not used on it again."""
        if self.ptr:
            _lib.load().csx_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()


_default_ctx = None
_ctx_lock = threading.Lock()


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def device_count() -> int:
    return int(_lib.load().csx_device_count())


# ---------------------------------------------------------------- options --

@dataclass
class SolventParams:
    """`SolventParams` (src/solvent.rs:27-50)."""
    c1: float = 1.0
    c2: float = 0.0
    shell_thickness_angstrom: float = 3.0
    bulk_density_e_per_a3: float = 0.334
    hydration_model: HydrationModel = HydrationModel.Analytic
    hydration_grid_spacing_angstrom: float = 4.0

    def to_c(self) -> SolventParamsC:
        return SolventParamsC(self.c1, self.c2, self.shell_thickness_angstrom,
                              self.bulk_density_e_per_a3, int(self.hydration_model),
                              self.hydration_grid_spacing_angstrom)


@dataclass
class MultipoleOptions:
    l_max: Optional[int] = None


@dataclass
class ComputeOptions:
    """`ComputeOptions` (src/scattering.rs:28-47) plus the pair-sum precision."""
    method: ScatteringMethod = ScatteringMethod.Auto
    hydrogen_mode: HydrogenMode = HydrogenMode.Implicit
    include_hetatm: bool = False
    multipole: MultipoleOptions = field(default_factory=MultipoleOptions)
    solvent: Optional[SolventParams] = None
    precision: Precision = Precision.FP32_COMPENSATED

    def to_c(self) -> ComputeOptionsC:
        o = ComputeOptionsC()
        o.method = int(self.method)
        o.implicit_hydrogen = int(self.hydrogen_mode == HydrogenMode.Implicit)
        o.has_solvent = int(self.solvent is not None)
        o.solvent = (self.solvent or SolventParams()).to_c()
        o.l_max = 0 if self.multipole.l_max is None else int(self.multipole.l_max)
        o.precision = int(self.precision)
        return o


@dataclass
class CurveMetadata:
    method: str
    atom_count: int


@dataclass
class ScatteringCurve:
    q: np.ndarray
    intensity: np.ndarray
    metadata: CurveMetadata


# --------------------------------------------------------------- functions --

def _out(n: int) -> np.ndarray:
    return np.zeros(n, dtype=np.float64)


def compute_curve(structure: Structure, q_values, options: ComputeOptions = None,
                  ctx: Context = None) -> ScatteringCurve:
    """`compute_curve` (src/scattering.rs:112-212)."""
    options = options or ComputeOptions()
    q = f64(q_values).reshape(-1)
    out = _out(q.size)
    used = C.c_int()
    opts = options.to_c()
    _check(_lib.load().csx_compute_curve(structure.handle(ctx), dptr(q), q.size, C.byref(opts),
                                         dptr(out), C.byref(used)))
    return ScatteringCurve(q.copy(), out, CurveMetadata(ScatteringMethod(used.value).name,
                                                        len(structure)))


def debye_intensity_with_hydrogens(structure: Structure, q_values, implicit_hydrogen: bool,
                                   precision: Precision = Precision.FP32_COMPENSATED,
                                   ctx: Context = None) -> np.ndarray:
    """`debye_intensity_with_hydrogens` (src/scattering.rs:86-110)."""
    q = f64(q_values).reshape(-1)
    out = _out(q.size)
    _check(_lib.load().csx_debye_intensity(structure.handle(ctx), dptr(q), q.size,
                                           int(implicit_hydrogen), int(precision), dptr(out)))
    return out


def debye_intensity(structure: Structure, q_values,
                    precision: Precision = Precision.FP32_COMPENSATED,
                    ctx: Context = None) -> np.ndarray:
    """`debye_intensity` (src/scattering.rs:82-84): implicit hydrogens."""
    return debye_intensity_with_hydrogens(structure, q_values, True, precision, ctx)


def debye_solvent_basis(structure: Structure, q_values, params: SolventParams,
                        implicit_hydrogen: bool = True,
                        precision: Precision = Precision.FP32_COMPENSATED,
                        ctx: Context = None) -> np.ndarray:
    """`debye_solvent_basis` (src/solvent.rs:504-584): array [6, nq]."""
    q = f64(q_values).reshape(-1)
    out = _out(6 * q.size)
    p = params.to_c()
    _check(_lib.load().csx_debye_solvent_basis(structure.handle(ctx), dptr(q), q.size,
                                               C.byref(p), int(implicit_hydrogen),
                                               int(precision), dptr(out)))
    return out.reshape(6, q.size)


def solvent_corrected_intensity(structure: Structure, q_values, params: SolventParams,
                                implicit_hydrogen: bool = True,
                                precision: Precision = Precision.FP32_COMPENSATED,
                                ctx: Context = None) -> np.ndarray:
    """`solvent_corrected_intensity` (src/solvent.rs:490-500)."""
    q = f64(q_values).reshape(-1)
    out = _out(q.size)
    p = params.to_c()
    _check(_lib.load().csx_solvent_corrected_intensity(structure.handle(ctx), dptr(q), q.size,
                                                       C.byref(p), int(implicit_hydrogen),
                                                       int(precision), dptr(out)))
    return out


def debye_pair_tables(structure: Structure, q_values, params: SolventParams = None,
                      ctx: Context = None):
    """Per-type pair tables: (S [nbins, nq], self [T])."""
    q = f64(q_values).reshape(-1)
    lib = _lib.load()
    h = structure.handle(ctx)
    pc = params.to_c() if params is not None else None
    pp = C.byref(pc) if pc is not None else None
    nt = C.c_size_t()
    _check(lib.csx_debye_pair_tables(h, dptr(q), q.size, pp, None, None, C.byref(nt)))
    t = nt.value
    nb = t * (t + 1) // 2
    out = _out(nb * q.size)
    self_terms = _out(t)
    _check(lib.csx_debye_pair_tables(h, dptr(q), q.size, pp, dptr(out), dptr(self_terms),
                                     C.byref(nt)))
    return out.reshape(nb, q.size), self_terms


def _handles(structures: Sequence[Structure], ctx):
    arr = (C.c_void_p * len(structures))()
    for k, s in enumerate(structures):
        arr[k] = s.handle(ctx).value
    return arr


def batch_debye_intensity(structures: Sequence[Structure], q_values,
                          implicit_hydrogen: bool = True, ctx: Context = None) -> np.ndarray:
    """Vacuum Debye profiles of many structures from one kernel launch: [m, nq]."""
    q = f64(q_values).reshape(-1)
    out = _out(len(structures) * q.size)
    _check(_lib.load().csx_batch_debye_intensity(_handles(structures, ctx), len(structures),
                                                 dptr(q), q.size, int(implicit_hydrogen),
                                                 dptr(out)))
    return out.reshape(len(structures), q.size)


def batch_debye_solvent_basis(structures: Sequence[Structure], q_values, params: SolventParams,
                              implicit_hydrogen: bool = True, ctx: Context = None) -> np.ndarray:
    """Six-term Debye bases of many structures from one kernel launch: [m, 6, nq]."""
    q = f64(q_values).reshape(-1)
    out = _out(len(structures) * 6 * q.size)
    p = params.to_c()
    _check(_lib.load().csx_batch_debye_solvent_basis(_handles(structures, ctx), len(structures),
                                                     dptr(q), q.size, C.byref(p),
                                                     int(implicit_hydrogen), dptr(out)))
    return out.reshape(len(structures), 6, q.size)


def debye_grad_x(structure: Structure, q_values, weights, params: SolventParams = None,
                 implicit_hydrogen: bool = True, ctx: Context = None) -> np.ndarray:
    """sum_q weights[q] dI(q)/dx for every atom: array [n, 3] (adjoint kernel)."""
    q = f64(q_values).reshape(-1)
    w = f64(weights).reshape(-1)
    if w.size != q.size:
        raise InvalidInput("weights and q differ in length")
    out = _out(3 * len(structure))
    pc = params.to_c() if params is not None else None
    _check(_lib.load().csx_debye_grad_x(structure.handle(ctx), dptr(q), q.size, dptr(w),
                                        C.byref(pc) if pc is not None else None,
                                        int(implicit_hydrogen), dptr(out)))
    return out.reshape(len(structure), 3)


def debye_jacobian(structure: Structure, q_values, params: SolventParams = None,
                   implicit_hydrogen: bool = True, ctx: Context = None) -> np.ndarray:
    """dI(q)/dx: array [nq, n, 3] (adjoint kernel)."""
    q = f64(q_values).reshape(-1)
    out = _out(q.size * 3 * len(structure))
    pc = params.to_c() if params is not None else None
    _check(_lib.load().csx_debye_jacobian(structure.handle(ctx), dptr(q), q.size,
                                          C.byref(pc) if pc is not None else None,
                                          int(implicit_hydrogen), dptr(out)))
    return out.reshape(q.size, len(structure), 3)


@dataclass
class MultipoleParams:
    """`MultipoleParams` (src/spherical.rs:19-31)."""
    l_max: int = 20
    hydrogen_mode: HydrogenMode = HydrogenMode.Implicit


def multipole_intensity(structure: Structure, q_values, params: MultipoleParams = None,
                        ctx: Context = None) -> np.ndarray:
    """`multipole_intensity` (src/spherical.rs:266-312)."""
    params = params or MultipoleParams()
    q = f64(q_values).reshape(-1)
    out = _out(q.size)
    _check(_lib.load().csx_multipole_intensity(
        structure.handle(ctx), dptr(q), q.size, int(params.l_max),
        int(params.hydrogen_mode == HydrogenMode.Implicit), dptr(out)))
    return out


def multipole_solvent_intensity(structure: Structure, q_values, params: MultipoleParams,
                                solvent: SolventParams, ctx: Context = None) -> np.ndarray:
    """`multipole_solvent_intensity` (src/spherical.rs:408-470)."""
    q = f64(q_values).reshape(-1)
    out = _out(q.size)
    p = solvent.to_c()
    _check(_lib.load().csx_multipole_solvent_intensity(
        structure.handle(ctx), dptr(q), q.size, int(params.l_max),
        int(params.hydrogen_mode == HydrogenMode.Implicit), C.byref(p), dptr(out)))
    return out


def multipole_solvent_basis(structure: Structure, q_values, params: MultipoleParams,
                            solvent: SolventParams, ctx: Context = None) -> np.ndarray:
    """`multipole_solvent_basis` (src/spherical.rs:474-565): array [6, nq]."""
    q = f64(q_values).reshape(-1)
    out = _out(6 * q.size)
    p = solvent.to_c()
    _check(_lib.load().csx_multipole_solvent_basis(
        structure.handle(ctx), dptr(q), q.size, int(params.l_max),
        int(params.hydrogen_mode == HydrogenMode.Implicit), C.byref(p), dptr(out)))
    return out.reshape(6, q.size)


def combine_solvent_basis(basis, q_values, c1: float, c2: float) -> np.ndarray:
    """`combine_solvent_basis` (src/spherical.rs:567-584)."""
    q = f64(q_values).reshape(-1)
    b = f64(basis).reshape(6, q.size)
    out = _out(q.size)
    _check(_lib.load().csx_combine_solvent_basis(dptr(b), dptr(q), q.size, c1, c2, dptr(out)))
    return out


def excluded_volume_scale(q: float, radius_scale: float) -> float:
    """`excluded_volume_scale` (src/solvent.rs:448-452)."""
    return float(_lib.load().csx_excluded_volume_scale(q, radius_scale))


def water_form_factor(q: float) -> float:
    """`water_form_factor` (src/solvent.rs:227-231)."""
    return float(_lib.load().csx_water_form_factor(q))


def form_factor_for_atom(atom: Atom, implicit_hydrogen: bool, q: float) -> float:
    """`form_factor_for_atom` (src/formfactor.rs:425-439)."""
    lib = _lib.load()
    group = lib.csx_united_atom_group(atom.residue_name.encode(), atom.atom_name.encode())
    return float(lib.csx_form_factor(int(atom.element), group, atom.occupancy,
                                     int(implicit_hydrogen), q))


@dataclass
class HydrationGrid:
    sites: np.ndarray
    voxel_volume_a3: float


def _grid(call, *args) -> HydrationGrid:
    count, vol = C.c_size_t(), C.c_double()
    _check(call(*args, None, 0, C.byref(count), C.byref(vol)))
    sites = np.zeros((count.value, 3))
    _check(call(*args, dptr(sites), count.value, C.byref(count), C.byref(vol)))
    return HydrationGrid(sites, float(vol.value))


def build_adaptive_hydration_grid(structure: Structure, spacing_angstrom: float = 4.0,
                                  molecular_clearance_angstrom: float = 3.0,
                                  padding_angstrom: float = 12.0,
                                  ctx: Context = None) -> HydrationGrid:
    """`build_adaptive_hydration_grid` (src/solvent.rs:121-223), rasterised on
    the GPU."""
    lib = _lib.load()
    return _grid(lib.csx_build_adaptive_hydration_grid, structure.handle(ctx),
                 spacing_angstrom, molecular_clearance_angstrom, padding_angstrom)


def build_hydration_grid(structure: Structure, spacing_angstrom: float = 1.0,
                         probe_radius_angstrom: float = 1.4,
                         shell_thickness_angstrom: float = 3.0,
                         ctx: Context = None) -> HydrationGrid:
    """`build_hydration_grid` (src/solvent.rs:259-401): rasterisation, exterior
    flood fill and shell on the GPU."""
    lib = _lib.load()
    return _grid(lib.csx_build_hydration_grid, structure.handle(ctx),
                 spacing_angstrom, probe_radius_angstrom, shell_thickness_angstrom)


# -------------------------------------------------------- synthetic inputs --

def generate_globule(n: int, radius: float, elongation: float = 1.0,
                     seed: int = 0) -> Structure:
    """Seeded synthetic globular protein (see include/crabsaxs.h)."""
    xyz = np.zeros((n, 3))
    el = np.zeros(n, dtype=np.uint8)
    _check(_lib.load().csx_generate_globule(n, radius, elongation, seed, dptr(xyz),
                                            el.ctypes.data_as(C.POINTER(C.c_uint8))))
    return Structure.from_arrays(xyz, el)


def generate_domain_chain(n: int, domains: int, seed: int = 0) -> Structure:
    xyz = np.zeros((n, 3))
    el = np.zeros(n, dtype=np.uint8)
    _check(_lib.load().csx_generate_domain_chain(n, domains, seed, dptr(xyz),
                                                 el.ctypes.data_as(C.POINTER(C.c_uint8))))
    return Structure.from_arrays(xyz, el)


def generate_shell_sites(structure: Structure, spacing: float, inner: float,
                         outer: float) -> np.ndarray:
    lib = _lib.load()
    count = C.c_size_t()
    _check(lib.csx_generate_shell_sites(dptr(structure.xyz), len(structure), spacing, inner,
                                        outer, None, 0, C.byref(count)))
    sites = np.zeros((count.value, 3))
    _check(lib.csx_generate_shell_sites(dptr(structure.xyz), len(structure), spacing, inner,
                                        outer, dptr(sites), count.value, C.byref(count)))
    return sites


# --------------------------------------------------------------- analysis --

@dataclass
class PairDistribution:
    """`PairDistribution` (src/metrics.rs:32-37)."""
    r: np.ndarray
    p: np.ndarray
    rg: Optional[float]
    dmax: Optional[float]


@dataclass
class CoordinateFeatures:
    """`CoordinateFeatures` (src/metrics.rs:61-65)."""
    rg: float
    dmax: float
    p_r: PairDistribution


def coordinate_features(structure: Structure, bins: int = 160,
                        ctx: Context = None) -> CoordinateFeatures:
    """`coordinate_features` (src/metrics.rs:236-270), PrOptions.bins = `bins`;
    the pair histogram runs on the GPU."""
    nb = max(int(bins), 16)
    r, p = np.zeros(nb), np.zeros(nb)
    rg, dmax, prg, pdmax = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    count, flags = C.c_size_t(), C.c_int()
    _check(_lib.load().csx_coordinate_features(structure.handle(ctx), int(bins), C.byref(rg),
                                               C.byref(dmax), dptr(r), dptr(p), C.byref(count),
                                               C.byref(prg), C.byref(pdmax), C.byref(flags)))
    k = count.value
    pr = PairDistribution(r[:k].copy(), p[:k].copy(), prg.value if flags.value & 1 else None,
                          pdmax.value if flags.value & 2 else None)
    return CoordinateFeatures(rg.value, dmax.value, pr)


class ProfileCalculator:
    """`ProfileCalculator` (src/analysis.rs:58-63, 90-105)."""

    def __init__(self, q, options: ComputeOptions = None):
        q = f64(q).reshape(-1)
        if (not np.all(np.isfinite(q))) or np.any(q < 0) or np.any(np.diff(q) < 0):
            raise InvalidInput("q grid must be finite, non-negative, and non-decreasing")
        self.q = q
        self.options = options or ComputeOptions()

    def calculate(self, structure: Structure, ctx: Context = None) -> ScatteringCurve:
        return compute_curve(structure, self.q, self.options, ctx)

"""Fitting, scoring and ensemble API mirroring src/fit.rs, src/analysis.rs and
src/ensemble.rs.  The arithmetic (least squares, Nelder-Mead, gradient fit,
weight optimisation) runs in libcrabsaxs_hip.so; profiles and bases on the GPU.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import FitOptionsC, FitParamsC, FitResultC, dptr, f64
from .api import (Context, HydrationModel, HydrogenMode, InvalidInput, Precision, SaxsError,
                  ScatteringMethod, SolventParams, Structure, _check)


@dataclass
class ExperimentalCurve:
    """`ExperimentalCurve` (src/io.rs:8-16, 113-141)."""
    q: np.ndarray
    intensity: np.ndarray
    sigma: np.ndarray
    has_errors: bool = True

    def __post_init__(self):
        self.q, self.intensity, self.sigma = f64(self.q), f64(self.intensity), f64(self.sigma)

    def __len__(self):
        return int(self.q.size)

    def is_empty(self) -> bool:
        return self.q.size == 0

    def through_qmax(self, qmax: float) -> "ExperimentalCurve":
        keep = self.q <= qmax
        return ExperimentalCurve(self.q[keep], self.intensity[keep], self.sigma[keep],
                                 self.has_errors)

    def q_unit(self) -> str:
        return "angstrom^-1"

    @classmethod
    def from_dat_file(cls, path) -> "ExperimentalCurve":
        """`ExperimentalCurve::from_dat_file` (src/io.rs:19-112)."""
        return cls._read(_lib.load().csx_read_dat_file, str(path).encode())

    @classmethod
    def parse_dat(cls, text: str) -> "ExperimentalCurve":
        """The same rules applied to text in memory."""
        data = text.encode()
        return cls._read(_lib.load().csx_parse_dat, data, len(data))

    @classmethod
    def _read(cls, call, *args) -> "ExperimentalCurve":
        n, has = C.c_size_t(), C.c_int()
        _check(call(*args, None, None, None, 0, C.byref(n), C.byref(has)))
        q, i, s = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
        _check(call(*args, dptr(q), dptr(i), dptr(s), n.value, C.byref(n), C.byref(has)))
        return cls(q, i, s, bool(has.value))


def write_curve_dat(path, q, intensity) -> None:
    """`write_curve_dat` (src/io.rs:143-155): a header and `{:.8e}` columns."""
    _write_curve(path, q, intensity, "# q (1/A) intensity", " ")


def write_curve_csv(path, q, intensity) -> None:
    """`write_curve_csv` (src/io.rs:157-169)."""
    _write_curve(path, q, intensity, "q,intensity", ",")


def _write_curve(path, q, intensity, header, sep) -> None:
    q, intensity = np.asarray(q, float), np.asarray(intensity, float)
    if q.shape != intensity.shape:
        raise InvalidInput("q and intensity lengths differ")
    with open(path, "w") as f:
        f.write(header + "\n")
        for a, b in zip(q, intensity):
            f.write(f"{_rust_exp(a)}{sep}{_rust_exp(b)}\n")


def _rust_exp(v: float) -> str:
    """Rust's `{:.8e}`: mantissa with 8 decimals, exponent without sign
    padding or '+' (1.00000000e-2, 2.50000000e1)."""
    if not np.isfinite(v):
        return "NaN" if np.isnan(v) else ("inf" if v > 0 else "-inf")
    m, e = f"{v:.8e}".split("e")
    return f"{m}e{int(e)}"


class BackgroundMode:
    """`BackgroundMode` (src/fit.rs:41-47): Fit, FixedZero or Fixed(value)."""
    def __init__(self, kind: str = "Fit", value: float = 0.0):
        self.kind, self.value = kind, float(value)

    @classmethod
    def Fixed(cls, value: float) -> "BackgroundMode":
        return cls("Fixed", value)

    def code(self) -> int:
        return {"Fit": 0, "FixedZero": 1, "Fixed": 2}[self.kind]


BackgroundMode.Fit = BackgroundMode("Fit")
BackgroundMode.FixedZero = BackgroundMode("FixedZero")


class ScaleMode:
    """`ScaleMode` (src/fit.rs:29-34): Fit or Fixed(value)."""
    def __init__(self, kind: str = "Fit", value: float = 0.0):
        self.kind, self.value = kind, float(value)

    @classmethod
    def Fixed(cls, value: float) -> "ScaleMode":
        return cls("Fixed", value)

    def code(self) -> int:
        return 0 if self.kind == "Fit" else 1


ScaleMode.Fit = ScaleMode("Fit")


class Optimizer(enum.IntEnum):
    NelderMead = 0
    Gradient = 1


@dataclass
class FitParams:
    """`FitParams` (src/fit.rs:11-17, defaults 125-135)."""
    scale: float = 1.0
    background: float = 0.0
    c1: float = 1.0
    c2: float = 0.0

    def to_c(self) -> FitParamsC:
        return FitParamsC(self.scale, self.background, self.c1, self.c2)


@dataclass
class FitOptions:
    """`FitOptions` (src/fit.rs:49-82) plus the optimiser and pair-sum precision."""
    method: ScatteringMethod = ScatteringMethod.Auto
    hydrogen_mode: HydrogenMode = HydrogenMode.Implicit
    c1_bounds: tuple = (0.95, 1.05)
    c2_bounds: tuple = (-2.0, 4.0)
    max_iterations: int = 64
    q_sampling_stride: int = 20
    solvent: SolventParams = field(default_factory=SolventParams)
    background_mode: BackgroundMode = field(default_factory=lambda: BackgroundMode.Fit)
    scale_mode: ScaleMode = field(default_factory=lambda: ScaleMode.Fit)
    optimizer: Optimizer = Optimizer.NelderMead
    precision: Precision = Precision.FP32_COMPENSATED

    def to_c(self) -> FitOptionsC:
        o = FitOptionsC()
        o.method = int(self.method)
        o.implicit_hydrogen = int(self.hydrogen_mode == HydrogenMode.Implicit)
        o.c1_bounds[0], o.c1_bounds[1] = self.c1_bounds
        o.c2_bounds[0], o.c2_bounds[1] = self.c2_bounds
        o.max_iterations = int(self.max_iterations)
        o.q_sampling_stride = int(self.q_sampling_stride)
        o.solvent = self.solvent.to_c()
        o.background_mode = self.background_mode.code()
        o.background_value = self.background_mode.value
        o.scale_mode = self.scale_mode.code()
        o.scale_value = self.scale_mode.value
        o.optimizer = int(self.optimizer)
        o.precision = int(self.precision)
        return o


@dataclass
class FitResult:
    """`FitResult` (src/fit.rs:19-27)."""
    params: FitParams
    chi2: float
    fitted_curve: np.ndarray
    iterations: int
    converged: bool
    termination: str


def _result(r: FitResultC, fitted: np.ndarray) -> FitResult:
    p = r.params
    return FitResult(FitParams(p.scale, p.background, p.c1, p.c2), float(r.chi2), fitted,
                     int(r.iterations), bool(r.converged), r.termination.decode())


def fit_to_experimental_with_options(structure: Structure, experimental: ExperimentalCurve,
                                     initial: FitParams = None, options: FitOptions = None,
                                     ctx: Context = None) -> FitResult:
    """`fit_to_experimental_with_options` (src/fit.rs:288-454)."""
    initial = initial or FitParams()
    options = options or FitOptions()
    e = experimental
    n = len(e)
    if not (e.q.size == e.intensity.size == e.sigma.size):
        raise InvalidInput("fit requires a non-empty structure and equal-length finite q, "
                           "intensity, and positive-sigma arrays on a non-decreasing q grid")
    res = FitResultC()
    fitted = np.zeros(n)
    o, p = options.to_c(), initial.to_c()
    _check(_lib.load().csx_fit_to_experimental(structure.handle(ctx), dptr(e.q),
                                               dptr(e.intensity), dptr(e.sigma), n, C.byref(p),
                                               C.byref(o), C.byref(res), dptr(fitted)))
    return _result(res, fitted)


def fit_to_experimental(structure: Structure, experimental: ExperimentalCurve,
                        initial_guess: FitParams = None, ctx: Context = None) -> FitResult:
    """`fit_to_experimental` (src/fit.rs:275-286)."""
    return fit_to_experimental_with_options(structure, experimental, initial_guess,
                                            FitOptions(), ctx)


def fit_basis(structure: Structure, q, options: FitOptions, ctx: Context = None) -> np.ndarray:
    """Six-term basis on the q grid as fit.rs:344-384 builds it: [6, n]."""
    q = f64(q)
    out = np.zeros(6 * q.size)
    o = options.to_c()
    _check(_lib.load().csx_fit_basis(structure.handle(ctx), dptr(q), q.size, C.byref(o),
                                     dptr(out)))
    return out.reshape(6, q.size)


def fit_from_basis(basis, experimental: ExperimentalCurve, initial: FitParams = None,
                   options: FitOptions = None) -> FitResult:
    """The fit of fit.rs:385-453 on a precomputed basis (host only)."""
    initial = initial or FitParams()
    options = options or FitOptions()
    e = experimental
    b = f64(basis).reshape(6, len(e))
    res = FitResultC()
    fitted = np.zeros(len(e))
    o, p = options.to_c(), initial.to_c()
    _check(_lib.load().csx_fit_from_basis(dptr(b), dptr(e.q), dptr(e.intensity), dptr(e.sigma),
                                          len(e), C.byref(p), C.byref(o), C.byref(res),
                                          dptr(fitted)))
    return _result(res, fitted)


def solve_scale_background(theory, experimental: ExperimentalCurve,
                           scale_mode: ScaleMode = ScaleMode.Fit,
                           background_mode: BackgroundMode = BackgroundMode.Fit):
    """`solve_scale_background` (src/fit.rs:200-244): (scale, background, chi2)."""
    t = f64(theory)
    e = experimental
    out = np.zeros(3)
    _check(_lib.load().csx_solve_scale_background(dptr(t), dptr(e.intensity), dptr(e.sigma),
                                                  t.size, scale_mode.code(), scale_mode.value,
                                                  background_mode.code(), background_mode.value,
                                                  dptr(out)))
    return float(out[0]), float(out[1]), float(out[2])


def cubic_hermite_interpolate(x, y, targets) -> np.ndarray:
    """`cubic_hermite_interpolate` (src/fit.rs:93-123)."""
    x, y, t = f64(x), f64(y), f64(targets)
    out = np.zeros(t.size)
    _check(_lib.load().csx_cubic_hermite_interpolate(dptr(x), dptr(y), x.size, dptr(t), t.size,
                                                     dptr(out)))
    return out


def chi2_grad(basis, experimental: ExperimentalCurve, params: FitParams):
    """chi^2 and d chi^2 / d(scale, background, c1, c2) from a six-term basis."""
    e = experimental
    b = f64(basis).reshape(6, len(e))
    chi = C.c_double()
    g = np.zeros(4)
    p = params.to_c()
    _check(_lib.load().csx_chi2_grad(dptr(b), dptr(e.q), dptr(e.intensity), dptr(e.sigma),
                                     len(e), C.byref(p), C.byref(chi), dptr(g)))
    return float(chi.value), g


# ----------------------------------------------------------------- analysis --

class Quality(enum.Enum):
    """`Quality` presets (src/analysis.rs:11-56)."""
    Fast = "fast"
    Balanced = "balanced"
    Accurate = "accurate"

    def fit_options(self, options: FitOptions) -> FitOptions:
        import copy
        o = copy.deepcopy(options)
        if self is Quality.Fast:
            o.q_sampling_stride = max(o.q_sampling_stride, 20)
            o.max_iterations = min(o.max_iterations, 32)
            o.solvent.hydration_grid_spacing_angstrom = 8.0
        elif self is Quality.Balanced:
            o.q_sampling_stride = 20
            o.max_iterations = max(o.max_iterations, 64)
            o.solvent.hydration_grid_spacing_angstrom = 4.0
        else:
            o.q_sampling_stride = 1
            o.max_iterations = max(o.max_iterations, 128)
            o.solvent.hydration_grid_spacing_angstrom = 3.0
        return o

    def compute_options(self, options):
        import copy
        o = copy.deepcopy(options)
        if o.solvent is not None:
            if self is Quality.Fast:
                o.solvent.hydration_grid_spacing_angstrom = 8.0
            elif self is Quality.Accurate:
                o.solvent.hydration_grid_spacing_angstrom = 3.0
        return o


class SaxsFitter:
    """`SaxsFitter` (src/analysis.rs:65-88)."""

    def __init__(self, options: FitOptions = None):
        self.options = options or FitOptions()

    def fit(self, structure: Structure, experimental: ExperimentalCurve,
            ctx: Context = None) -> FitResult:
        return fit_to_experimental_with_options(structure, experimental, FitParams(),
                                                self.options, ctx)


@dataclass
class ScoreOptions:
    fit: FitOptions = field(default_factory=FitOptions)
    quality: Quality = Quality.Fast


@dataclass
class ScoreResult:
    chi2: float
    reduced_chi2: float
    scale: float
    background: float
    q_min: float
    q_max: float
    points: int
    has_errors: bool
    q_unit: str
    fit: FitResult


class SaxsScorer:
    """`SaxsScorer` (src/analysis.rs:136-178)."""

    def __init__(self, experimental: ExperimentalCurve, options: ScoreOptions = None):
        if experimental.is_empty():
            raise InvalidInput("experimental curve is empty")
        self.experimental = experimental
        self.options = options or ScoreOptions()

    def score(self, structure: Structure, ctx: Context = None) -> ScoreResult:
        o = self.options.quality.fit_options(self.options.fit)
        fit = fit_to_experimental_with_options(structure, self.experimental, FitParams(), o, ctx)
        e = self.experimental
        dof = max(len(e) - 1, 1)
        return ScoreResult(fit.chi2 * dof, fit.chi2, fit.params.scale, fit.params.background,
                           float(e.q[0]), float(e.q[-1]) if len(e) else 0.0, len(e),
                           e.has_errors, e.q_unit(), fit)


@dataclass
class ComparisonResult:
    chi2: float
    reduced_chi2: float
    pearson_correlation: float
    r_factor: float
    weighted_rmse: float
    residuals: np.ndarray


def compare_curves(experimental: ExperimentalCurve, calculated) -> ComparisonResult:
    """`compare_curves` (src/analysis.rs:190-251); O(n) host arithmetic."""
    calc = f64(calculated)
    e = experimental
    if len(e) != calc.size or e.is_empty():
        raise InvalidInput("experimental and calculated curves must have equal non-zero lengths")
    res = calc - e.intensity
    dof = max(len(e) - 1, 1)
    chi = float(np.sum((res / e.sigma) ** 2))
    mo, mm = e.intensity.mean(), calc.mean()
    cov = float(np.sum((e.intensity - mo) * (calc - mm)))
    vo, vm = float(np.sum((e.intensity - mo) ** 2)), float(np.sum((calc - mm) ** 2))
    pearson = cov / np.sqrt(vo * vm) if vo > 0 and vm > 0 else 0.0
    den = float(np.sum(np.abs(e.intensity)))
    rf = float(np.sum(np.abs(res))) / den if den > 0 else 0.0
    return ComparisonResult(chi, chi / dof, pearson, rf, float(np.sqrt(chi / len(e))), res)


# ----------------------------------------------------------------- ensemble --

def project_simplex(values) -> np.ndarray:
    """`project_simplex` (src/ensemble.rs:449-465)."""
    v = f64(values)
    out = np.zeros(v.size)
    _check(_lib.load().csx_project_simplex(dptr(v), v.size, dptr(out)))
    return out


def optimize_weights(curves, experimental: ExperimentalCurve, options: FitOptions = None,
                     uniform: bool = False, regularization: float = 0.0):
    """`optimize_weights` (src/ensemble.rs:317-381): (weights, scale, bg, chi2, fitted)."""
    options = options or FitOptions()
    cv = f64(curves)
    m, n = cv.shape
    e = experimental
    w, out3, fitted = np.zeros(m), np.zeros(3), np.zeros(n)
    o = options.to_c()
    _check(_lib.load().csx_optimize_weights(dptr(cv), m, dptr(e.intensity), dptr(e.sigma), n,
                                            C.byref(o), int(uniform), regularization, dptr(w),
                                            dptr(out3), dptr(fitted)))
    return w, float(out3[0]), float(out3[1]), float(out3[2]), fitted


def nnls_weights(curves, experimental: ExperimentalCurve, options: FitOptions = None):
    """Exact non-negative least-squares weights (Lawson-Hanson active set):
    (weights, scale, bg, chi2, fitted) like `optimize_weights`.  No reference
    counterpart; needs a fitted scale."""
    options = options or FitOptions()
    cv = f64(curves)
    m, n = cv.shape
    e = experimental
    w, out3, fitted = np.zeros(m), np.zeros(3), np.zeros(n)
    o = options.to_c()
    _check(_lib.load().csx_nnls_weights(dptr(cv), m, dptr(e.intensity), dptr(e.sigma), n,
                                        C.byref(o), dptr(w), dptr(out3), dptr(fitted)))
    return w, float(out3[0]), float(out3[1]), float(out3[2]), fitted


def bootstrap_indices(seed: int, n: int, replicates: int) -> np.ndarray:
    """The resampling draws of EnsembleFitter::fit (ensemble.rs:184-216):
    ChaCha8Rng::seed_from_u64(seed), gen_range(0..n); [replicates, n]."""
    out = np.zeros(replicates * n, dtype=np.uint64)
    _check(_lib.load().csx_bootstrap_indices(int(seed) & (2**64 - 1), n, replicates,
                                             out.ctypes.data_as(C.POINTER(C.c_uint64))))
    return out.reshape(replicates, n).astype(np.int64)


class WeightSolver(enum.IntEnum):
    """Ensemble weight solver: the reference's projected gradient
    (ensemble.rs:317-381) or exact non-negative least squares."""
    ProjectedGradient = 0
    NNLS = 1


@dataclass
class EnsembleOptions:
    """`EnsembleOptions` (src/ensemble.rs:13-38)."""
    fit_weights: bool = True
    uniform: bool = False
    regularization: float = 0.0
    max_conformers: Optional[int] = None
    min_weight: float = 0.0
    bootstrap: int = 0
    seed: int = 0xC0FFEE
    fit: FitOptions = field(default_factory=FitOptions)
    weight_solver: WeightSolver = WeightSolver.ProjectedGradient


@dataclass
class WeightInterval:
    median: float
    lower: float
    upper: float


@dataclass
class EnsembleResult:
    """`EnsembleResult` (src/ensemble.rs:47-59)."""
    conformers: List[str]
    weights: np.ndarray
    weight_intervals: List[WeightInterval]
    params: FitParams
    chi2: float
    reduced_chi2: float
    fitted_curve: np.ndarray
    bootstrap_reduced_chi2: List[float]
    iterations: int
    converged: bool


def _quantile(sorted_values, p):
    if len(sorted_values) == 0:
        return float("nan")
    pos = min(max(p, 0.0), 1.0) * (len(sorted_values) - 1)
    lo, hi = int(np.floor(pos)), int(np.ceil(pos))
    return sorted_values[lo] + (sorted_values[hi] - sorted_values[lo]) * (pos - lo)


class EnsembleFitter:
    """`EnsembleFitter` (src/ensemble.rs:61-300).

    Each conformer's six-term basis on the experimental q grid is computed
    once on the GPU; the 25 (c1, c2) candidates of fit_core, the pruning
    passes and every bootstrap replicate re-contract those bases on the host
    instead of recomputing profiles (the results are the same: every
    candidate profile of the reference equals combine_solvent_basis of the
    conformer's basis).
    """

    def __init__(self, conformers: Sequence[Structure], names: Sequence[str],
                 experimental: ExperimentalCurve, options: EnsembleOptions = None):
        options = options or EnsembleOptions()
        if len(conformers) == 0 or any(len(c) == 0 for c in conformers):
            raise InvalidInput("ensemble must contain non-empty conformers")
        if len(names) != len(conformers):
            raise InvalidInput("ensemble names and conformers differ in length")
        e = experimental
        if e.is_empty() or not (e.q.size == e.intensity.size == e.sigma.size):
            raise InvalidInput("ensemble requires a validated experimental curve")
        if (not np.isfinite(options.regularization) or options.regularization < 0
                or not np.isfinite(options.min_weight) or not 0.0 <= options.min_weight <= 1.0
                or options.max_conformers == 0):
            raise InvalidInput("ensemble regularization and pruning options are invalid")
        self.conformers, self.names = list(conformers), list(names)
        self.experimental, self.options = experimental, options
        self._bases = None

    def bases(self, ctx: Context = None) -> np.ndarray:
        """Per-conformer bases [m, 6, n] (computed once)."""
        if self._bases is None:
            import copy
            from .api import batch_debye_solvent_basis
            o = copy.deepcopy(self.options.fit)
            o.q_sampling_stride = 1
            if (o.method == ScatteringMethod.Debye
                    and o.precision == Precision.FP32_COMPENSATED):
                # every conformer's pair tables from one launch
                self._bases = batch_debye_solvent_basis(
                    self.conformers, self.experimental.q, o.solvent,
                    o.hydrogen_mode == HydrogenMode.Implicit, ctx)
            else:
                self._bases = np.stack([fit_basis(c, self.experimental.q, o, ctx)
                                        for c in self.conformers])
        return self._bases

    def _fit_core(self, selected, data: ExperimentalCurve, ctx):
        """fit_core on the GPU (csx_ensemble_fit_core_device; bitwise the
        host restatement's result)."""
        from .api import default_context
        ctx = ctx or default_context()
        b = np.ascontiguousarray(self.bases(ctx)[selected])
        m, n = len(selected), len(data)
        w, chi, fitted = np.zeros(m), C.c_double(), np.zeros(n)
        p = FitParamsC()
        o = self.options.fit.to_c()
        uniform = self.options.uniform or not self.options.fit_weights
        _check(_lib.load().csx_ensemble_fit_core_device(
            ctx.ptr, dptr(b), m, dptr(data.q), dptr(data.intensity), dptr(data.sigma), n,
            C.byref(o), int(uniform), self.options.regularization,
            int(self.options.weight_solver), dptr(w), C.byref(p), C.byref(chi), dptr(fitted)))
        return w, FitParams(p.scale, p.background, p.c1, p.c2), float(chi.value), fitted

    def fit(self, ctx: Context = None) -> EnsembleResult:
        """`EnsembleFitter::fit` (src/ensemble.rs:127-232)."""
        opts = self.options
        selected = list(range(len(self.conformers)))
        w, params, chi, fitted = self._fit_core(selected, self.experimental, ctx)
        changed = True
        while changed:
            changed = False
            keep = [idx for pos, idx in enumerate(selected) if w[pos] + 1e-12 >= opts.min_weight]
            if not keep:
                keep = [selected[int(np.argmax(w))]]
            if opts.max_conformers is not None and len(keep) > opts.max_conformers:
                ranked = sorted(((w[pos], idx) for pos, idx in enumerate(selected)),
                                key=lambda t: -t[0])
                keep = [idx for _, idx in ranked[:opts.max_conformers]]
            keep.sort()
            if keep != selected:
                selected = keep
                w, params, chi, fitted = self._fit_core(selected, self.experimental, ctx)
                changed = True
        intervals = [WeightInterval(x, x, x) for x in w]
        boot_chi = []
        if opts.bootstrap > 0:
            e = self.experimental
            resid = (e.intensity - fitted) / e.sigma
            samples = [[] for _ in selected]
            picks = bootstrap_indices(opts.seed, len(e), opts.bootstrap)
            for r in range(opts.bootstrap):
                pick = picks[r]
                synth = ExperimentalCurve(e.q, fitted + resid[pick] * e.sigma, e.sigma,
                                          e.has_errors)
                rw, _, rchi, _ = self._fit_core(selected, synth, ctx)
                for pos, val in enumerate(rw):
                    samples[pos].append(val)
                boot_chi.append(rchi)
            for pos, s in enumerate(samples):
                s.sort()
                intervals[pos] = WeightInterval(_quantile(s, 0.5), _quantile(s, 0.025),
                                                _quantile(s, 0.975))
        dof = max(len(self.experimental) - 1, 1)
        return EnsembleResult([self.names[i] for i in selected], w, intervals, params,
                              chi * dof, chi, fitted, boot_chi, 250, True)

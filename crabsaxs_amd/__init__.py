"""crabsaxs_amd: MI355X-native crabSAXS profile path (HIP kernels behind a C ABI).

The public names mirror the reference crate's API (src/lib.rs:16-50).
"""
from .api import (  # noqa: F401
    Atom, ComputeOptions, Context, CoordinateFeatures, CurveMetadata, DeviceError, Element,
    FitDidNotConverge, PairDistribution, coordinate_features,
    AltlocPolicy, IoError, OccupancyPolicy, ParseError, ParseOptions, ResidueKind,
    UnsupportedElement, residue_kind,
    batch_debye_intensity, batch_debye_solvent_basis,
    HydrationGrid, HydrationModel, HydrogenMode, InvalidInput, MultipoleOptions,
    MultipoleParams, NoDevice, multipole_intensity, multipole_solvent_basis,
    multipole_solvent_intensity,
    OutOfMemory, Precision, ProfileCalculator, SaxsError, ScatteringCurve, ScatteringMethod,
    SolventParams, Structure, build_adaptive_hydration_grid, build_hydration_grid,
    combine_solvent_basis, compute_curve, debye_grad_x, debye_intensity,
    debye_intensity_with_hydrogens, debye_jacobian,
    debye_pair_tables, debye_solvent_basis, default_context, device_count,
    excluded_volume_scale, form_factor_for_atom, generate_domain_chain, generate_globule,
    generate_shell_sites, solvent_corrected_intensity, water_form_factor,
)
from .fitting import (  # noqa: F401
    BackgroundMode, ComparisonResult, EnsembleFitter, EnsembleOptions, EnsembleResult,
    ExperimentalCurve, FitOptions, FitParams, FitResult, Optimizer, Quality, SaxsFitter,
    SaxsScorer, ScaleMode, ScoreOptions, ScoreResult, WeightInterval, WeightSolver,
    bootstrap_indices, chi2_grad,
    compare_curves, cubic_hermite_interpolate, fit_basis, fit_from_basis, fit_to_experimental,
    fit_to_experimental_with_options, nnls_weights, optimize_weights, project_simplex,
    solve_scale_background, write_curve_csv, write_curve_dat,
)

// The reference's own unit tests for the profile path (src/formfactor.rs,
// src/scattering.rs, src/solvent.rs, src/spherical.rs, src/fit.rs,
// src/structure.rs, src/io.rs, tests/end_to_end.rs), restated
// over the C++ API of include/crabsaxs.hpp, plus drop-in checks of the
// analysis and ensemble wrappers.  Each test names the reference test it
// restates (file:line).  Tests marked GPU compute profiles and need a device.
//
//   reference_suite --cpu | --gpu | --all | --list | <test name>
//
// Not restated: adaptive_shell_width_matches_published_piecewise_rule and
// bessel_zero_limit test private helpers (covered through the oracle in
// tests/test_oracle.py and tests/test_host.py).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "crabsaxs.hpp"

using namespace crabsaxs;

namespace {

struct Failure {
  std::string what;
};

struct TestCase {
  const char* name;
  bool gpu;
  void (*fn)();
};

std::vector<TestCase>& registry() {
  static std::vector<TestCase> r;
  return r;
}

struct Register {
  Register(const char* name, bool gpu, void (*fn)()) { registry().push_back({name, gpu, fn}); }
};

#define CSX_TEST(name, gpu)                          \
  void name();                                       \
  const Register register_##name(#name, gpu, name);  \
  void name()

#define CHECK(cond)                                                                       \
  do {                                                                                    \
    if (!(cond))                                                                          \
      throw Failure{std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " #cond}; \
  } while (0)

template <class E, class F>
bool throws(F f) {
  try {
    f();
  } catch (const E&) {
    return true;
  }
  return false;
}

Atom atom(double x, double y, double z, Element e, const std::string& name,
          const std::string& residue = "TST", long residue_id = 1) {
  Atom a;
  a.pos = {x, y, z};
  a.element = e;
  a.atom_name = name;
  a.residue_name = residue;
  a.residue_id = residue_id;
  return a;
}

// Fixture files of the reference's tests (examples/data), next to this suite
// under tests/golden/structures.
std::string g_data_dir = "tests/golden/structures/";
std::string data_file(const char* name) { return g_data_dir + name; }

double element_factor(Element e, double q) {  // coefficients_for(e).evaluate(q)
  return csx_form_factor(static_cast<int>(e), -1, 1.0, 0, q);
}

std::vector<double> range_q(int lo, int hi, double step) {
  std::vector<double> q;
  for (int i = lo; i <= hi; ++i) q.push_back(i * step);
  return q;
}

double relative_l2(const std::vector<double>& a, const std::vector<double>& b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / std::max(den, 1e-30));
}

// ------------------------------------------------------ formfactor.rs:440-530

CSX_TEST(carbon_form_factor_at_q_zero_is_atomic_number, false) {
  CHECK(std::fabs(element_factor(Element::C, 0.0) - 6.0) < 0.01);
}

CSX_TEST(form_factors_decay_with_q, false) {
  CHECK(element_factor(Element::C, 0.5) < element_factor(Element::C, 0.0));
}

CSX_TEST(supported_ion_and_metal_factors_match_atomic_numbers_at_zero, false) {
  const std::pair<Element, double> cases[] = {{Element::Na, 11.0}, {Element::Mg, 12.0},
                                              {Element::Cl, 17.0}, {Element::K, 19.0},
                                              {Element::Ca, 20.0}, {Element::Fe, 26.0},
                                              {Element::Zn, 30.0}, {Element::I, 53.0}};
  for (const auto& c : cases) {
    CHECK(std::fabs(element_factor(c.first, 0.0) - c.second) < 0.02);
    CHECK(element_factor(c.first, 0.5) < element_factor(c.first, 0.0));
  }
}

CSX_TEST(united_atom_electron_counts_match_chemistry_at_q_zero, false) {
  CHECK(std::fabs(form_factor_for_atom(atom(0, 0, 0, Element::C, "CB", "ALA"), true, 0.0) - 9.0) <
        0.01);
  CHECK(std::fabs(form_factor_for_atom(atom(0, 0, 0, Element::C, "CA", "GLY"), true, 0.0) - 8.0) <
        0.01);
  CHECK(std::fabs(form_factor_for_atom(atom(0, 0, 0, Element::N, "NZ", "LYS"), true, 0.0) - 9.0) <
        0.01);
  CHECK(std::fabs(form_factor_for_atom(atom(0, 0, 0, Element::O, "OD1", "ASP"), true, 0.0) -
                  8.5) < 0.01);
}

CSX_TEST(nucleotide_united_atoms_include_covalent_hydrogens, false) {
  struct C {
    const char* residue;
    const char* name;
    Element e;
    double electrons;
  };
  const C cases[] = {{"DA", "C2'", Element::C, 8.0},
                     {"A", "C2'", Element::C, 7.0},
                     {"A", "N6", Element::N, 9.0},
                     {"DG", "N1", Element::N, 8.0},
                     {"DT", "C7", Element::C, 9.0}};
  for (const C& c : cases)
    CHECK(std::fabs(form_factor_for_atom(atom(0, 0, 0, c.e, c.name, c.residue), true, 0.0) -
                    c.electrons) < 0.02);
}

CSX_TEST(explicit_hydrogens_are_not_double_counted_in_implicit_mode, false) {
  const Atom h = atom(0, 0, 0, Element::H, "HB1", "ALA");
  CHECK(form_factor_for_atom(h, true, 0.1) == 0.0);
  CHECK(form_factor_for_atom(h, false, 0.1) > 0.0);
}

// ------------------------------------------------------ scattering.rs:215-366

Structure two_carbons() {  // scattering.rs:219-244
  return Structure({atom(0, 0, 0, Element::C, "C"), atom(2, 0, 0, Element::C, "C2")});
}

CSX_TEST(sinc_at_zero_is_one, false) { CHECK(sinc(0.0) == 1.0); }

CSX_TEST(two_atoms_match_closed_form, true) {
  const Structure s = two_carbons();
  const double q = 0.2;
  // the reference's FP64 arithmetic: the FP64 plain GPU path
  const double got = debye_intensity(s, {q}, Precision::FP64Plain)[0];
  const double f = form_factor_for_atom(s.atoms[0], true, q);
  const double expected = 2.0 * f * f * (1.0 + sinc(q * 2.0));
  CHECK(std::fabs(got - expected) < 1e-10);
  // and the FP32-compensated path at its own tolerance
  CHECK(std::fabs(debye_intensity(s, {q})[0] - expected) < 1e-6 * expected);
}

CSX_TEST(rejects_bad_q, true) {
  CHECK(throws<InvalidInput>([] { compute_curve(two_carbons(), {std::nan("")}, ComputeOptions{}); }));
  CHECK(throws<InvalidInput>([] { compute_curve(two_carbons(), {0.2, 0.1}, ComputeOptions{}); }));
}

CSX_TEST(auto_solution_curve_uses_validated_multipole_path, true) {
  ComputeOptions o;
  SolventParams p;
  p.c2 = 0.02;
  p.hydration_model = HydrationModel::AdaptiveGrid;
  p.hydration_grid_spacing_angstrom = 8.0;
  o.solvent = p;
  const ScatteringCurve curve = compute_curve(two_carbons(), {0.05, 0.1}, o);
  CHECK(curve.metadata.method == "Multipole");
  for (double v : curve.intensity) CHECK(std::isfinite(v) && v >= 0.0);
}

CSX_TEST(debye_supports_adaptive_grid_solution_curve, true) {
  ComputeOptions o;
  o.method = ScatteringMethod::Debye;
  SolventParams p;
  p.c2 = 0.02;
  p.hydration_model = HydrationModel::AdaptiveGrid;
  p.hydration_grid_spacing_angstrom = 8.0;
  o.solvent = p;
  const ScatteringCurve curve = compute_curve(two_carbons(), {0.05, 0.1}, o);
  CHECK(curve.metadata.method == "Debye");
  for (double v : curve.intensity) CHECK(std::isfinite(v) && v >= 0.0);
}

CSX_TEST(adaptive_public_multipole_curve_matches_debye_below_one_tenth_percent, true) {
  std::vector<Atom> atoms;
  const Element cycle[] = {Element::C, Element::N, Element::O, Element::S, Element::P};
  for (int i = 0; i < 96; ++i) {
    const double angle = i * 2.399963229728653, radial = 4.0 + (i % 13);
    atoms.push_back(atom(radial * std::cos(angle), radial * std::sin(angle), (i - 47.5) * 0.31,
                         cycle[i % 5], "A" + std::to_string(i), "TST", i));
  }
  const Structure s(atoms);
  const std::vector<double> q = range_q(1, 100, 0.005);
  ComputeOptions o;
  o.method = ScatteringMethod::Multipole;
  const std::vector<double> multipole = compute_curve(s, q, o).intensity;
  CHECK(relative_l2(multipole, debye_intensity(s, q)) <= 0.001);
}

// --------------------------------------------------------- solvent.rs:618-800

Structure carbon_oxygen() {  // solvent.rs:623-648
  return Structure({atom(-1, 0, 0, Element::C, "C"), atom(1, 0, 0, Element::O, "O")});
}

CSX_TEST(solvent_basis_recombines_to_direct_debye_curve, true) {
  const Structure s = carbon_oxygen();
  const std::vector<double> q = {0.0, 0.1, 0.3, 0.5};
  SolventParams p;
  p.c1 = 1.04;
  p.c2 = 0.35;
  for (Precision prec : {Precision::FP64Plain, Precision::FP32Compensated}) {
    const Basis b = debye_solvent_basis(s, q, p, true, prec);
    const std::vector<double> combined = combine_solvent_basis(b, q, p.c1, p.c2);
    const std::vector<double> direct = solvent_corrected_intensity(s, q, p, true, prec);
    for (size_t i = 0; i < q.size(); ++i)
      CHECK(std::fabs(combined[i] - direct[i]) <= std::max(std::fabs(direct[i]), 1.0) * 1e-12);
  }
}

CSX_TEST(vacuum_minus_solvent_is_finite_for_structure_fixture, true) {
  const Structure s = Structure::from_pdb_file(data_file("test_structure.pdb"));
  const std::vector<double> q = range_q(0, 100, 0.005);
  SolventParams p;
  p.c1 = 1.0;
  p.c2 = 0.0;
  const std::vector<double> curve = solvent_corrected_intensity(s, q, p, true);
  for (double v : curve) CHECK(std::isfinite(v) && v >= 0.0);
  CHECK(curve[0] > 0.0);
}

CSX_TEST(hydration_grid_is_deterministic_and_exterior, true) {
  const Structure s({carbon_oxygen().atoms[0]});
  const SolventGridOptions o;
  const HydrationGrid first = build_hydration_grid(s, o), second = build_hydration_grid(s, o);
  CHECK(!first.sites.empty());
  CHECK(first.sites.size() == second.sites.size());
  const double exclusion = CSX_VDW_RADIUS[CSX_EL_C] + o.probe_radius_angstrom;
  const double outer = exclusion + o.shell_thickness_angstrom;
  for (size_t k = 0; k < first.sites.size(); ++k) {
    const Vector3& a = first.sites[k];
    const Vector3& b = second.sites[k];
    CHECK(a.x == b.x && a.y == b.y && a.z == b.z);
    const Vector3& c = s.atoms[0].pos;
    const double d = std::sqrt((a.x - c.x) * (a.x - c.x) + (a.y - c.y) * (a.y - c.y) +
                               (a.z - c.z) * (a.z - c.z));
    CHECK(d > exclusion);
    CHECK(d <= outer + 1e-12);
  }
}

CSX_TEST(adaptive_hydration_grid_is_deterministic, true) {
  const AdaptiveHydrationOptions o;
  const HydrationGrid first = build_adaptive_hydration_grid(carbon_oxygen(), o);
  const HydrationGrid second = build_adaptive_hydration_grid(carbon_oxygen(), o);
  CHECK(!first.sites.empty());
  CHECK(first.voxel_volume_a3 == std::pow(o.spacing_angstrom, 3));
  CHECK(first.sites.size() == second.sites.size());
  for (size_t k = 0; k < first.sites.size(); ++k)
    CHECK(first.sites[k].x == second.sites[k].x && first.sites[k].y == second.sites[k].y &&
          first.sites[k].z == second.sites[k].z);
}

CSX_TEST(adaptive_fast_grid_halves_linear_point_density_weight, true) {
  AdaptiveHydrationOptions o;
  o.spacing_angstrom = 8.0;
  CHECK(build_adaptive_hydration_grid(carbon_oxygen(), o).voxel_volume_a3 == 2.0 * 64.0);
}

CSX_TEST(hydration_grids_are_translation_invariant, true) {
  const Structure s = carbon_oxygen();
  const Vector3 shift{3.17, -8.4, 2.6};
  Structure moved = s;
  for (Atom& a : moved.atoms) a.pos = {a.pos.x + shift.x, a.pos.y + shift.y, a.pos.z + shift.z};
  auto same = [&](const HydrationGrid& a, const HydrationGrid& b) {
    CHECK(a.sites.size() == b.sites.size());
    for (size_t k = 0; k < a.sites.size(); ++k) {
      const double dx = b.sites[k].x - a.sites[k].x - shift.x,
                   dy = b.sites[k].y - a.sites[k].y - shift.y,
                   dz = b.sites[k].z - a.sites[k].z - shift.z;
      CHECK(std::sqrt(dx * dx + dy * dy + dz * dz) < 1e-12);
    }
  };
  same(build_adaptive_hydration_grid(s, {}), build_adaptive_hydration_grid(moved, {}));
  same(build_hydration_grid(s, {}), build_hydration_grid(moved, {}));
}

CSX_TEST(water_form_factor_has_ten_electrons_at_zero, false) {
  CHECK(std::fabs(water_form_factor(0.0) - 10.0) < 1e-3);
}

CSX_TEST(displaced_volume_table_matches_cited_values, false) {
  const std::pair<Element, double> cases[] = {
      {Element::H, 5.15},  {Element::C, 16.44}, {Element::N, 2.49},  {Element::O, 9.13},
      {Element::P, 3.37},  {Element::S, 26.09}, {Element::Na, 4.45}, {Element::Mg, 1.56},
      {Element::Cl, 24.84}, {Element::K, 11.01}, {Element::Ca, 4.19}, {Element::Fe, 7.99},
      {Element::Zn, 9.85}, {Element::I, 44.60}};
  for (const auto& c : cases) CHECK(CSX_EXCLUDED_VOLUME[static_cast<int>(c.first)] == c.second);
}

// ------------------------------------------------------- spherical.rs:586-790

CSX_TEST(multipole_matches_debye_for_two_atoms, true) {
  const std::vector<double> q = {0.1, 0.2, 0.5};
  const Structure s = two_carbons();
  const std::vector<double> a = multipole_intensity(s, q, {32, HydrogenMode::Implicit});
  const std::vector<double> b = debye_intensity(s, q, Precision::FP64Plain);
  for (size_t i = 0; i < q.size(); ++i) CHECK(std::fabs(a[i] - b[i]) / b[i] < 1e-6);
}

CSX_TEST(multipole_matches_debye_at_large_qr, true) {
  const Structure s({atom(-20, 0, 0, Element::C, "C1"), atom(20, 0, 0, Element::O, "O1")});
  const std::vector<double> q = {0.2, 0.35, 0.5};
  const std::vector<double> m = multipole_intensity(s, q, {32, HydrogenMode::Implicit});
  const std::vector<double> d = debye_intensity(s, q, Precision::FP64Plain);
  for (size_t i = 0; i < q.size(); ++i) CHECK(std::fabs(m[i] - d[i]) / d[i] < 1e-6);
}

CSX_TEST(multipole_curve_matches_debye_below_one_tenth_percent, true) {
  std::vector<Atom> atoms;
  const Element cycle[] = {Element::C, Element::N, Element::O, Element::S};
  for (int i = 0; i < 24; ++i) {
    const double angle = i * 0.71;
    atoms.push_back(atom(7.0 * std::cos(angle), 7.0 * std::sin(angle), i * 0.45 - 5.0,
                         cycle[i % 4], "A" + std::to_string(i), "TST", i));
  }
  const Structure s(atoms);
  const std::vector<double> q = range_q(1, 100, 0.005);
  CHECK(relative_l2(multipole_intensity(s, q, {24, HydrogenMode::Implicit}),
                    debye_intensity(s, q)) <= 0.001);
}

CSX_TEST(adaptive_grid_solvent_basis_matches_debye_oracle, true) {
  std::vector<Atom> atoms;
  const Element cycle[] = {Element::C, Element::N, Element::O, Element::S};
  for (int i = 0; i < 8; ++i)
    atoms.push_back(atom(i * 1.4 - 4.9, std::sin(i * 0.9) * 2.0, std::cos(i * 0.7), cycle[i % 4],
                         "A" + std::to_string(i), "TST", i));
  const Structure s(atoms);
  const std::vector<double> q = range_q(1, 50, 0.01);
  SolventParams p;
  p.hydration_model = HydrationModel::AdaptiveGrid;
  p.hydration_grid_spacing_angstrom = 8.0;
  const Basis m = multipole_solvent_basis(s, q, {32, HydrogenMode::Implicit}, p);
  const Basis d = debye_solvent_basis(s, q, p, true);
  for (int c = 0; c < 6; ++c) CHECK(relative_l2(m[c], d[c]) <= 0.001);
}

// ------------------------------------------------------------- fit.rs:455-660

Structure carbon_oxygen_2a() {  // fit.rs:462-487
  return Structure({atom(0, 0, 0, Element::C, "C1"), atom(2, 0, 0, Element::O, "O1")});
}

ExperimentalCurve curve_of(const std::vector<double>& q, const std::vector<double>& y,
                           const std::vector<double>& sigma) {
  ExperimentalCurve e;
  e.q = q;
  e.intensity = y;
  e.sigma = sigma;
  e.has_errors = true;
  return e;
}

CSX_TEST(recovers_scale_and_background_on_synthetic_curve, true) {
  const Structure s = carbon_oxygen_2a();
  const std::vector<double> q = range_q(1, 8, 0.03);
  SolventParams p;
  p.c1 = 1.05;
  p.c2 = 0.4;
  const std::vector<double> raw = multipole_solvent_intensity(s, q, {8, HydrogenMode::Explicit}, p);
  std::vector<double> y;
  for (double v : raw) y.push_back(1.7 * v + 0.02);
  const LinearFit r = solve_scale_background(raw, curve_of(q, y, std::vector<double>(q.size(), 0.01)));
  CHECK(std::fabs(r.scale - 1.7) < 1e-8);
  CHECK(std::fabs(r.background - 0.02) < 1e-8);
  CHECK(r.chi2 < 1e-16);
}

CSX_TEST(fixed_zero_background_uses_weighted_scale_only, false) {
  const LinearFit r = solve_scale_background(
      {1.0, 2.0, 3.0}, curve_of({0.1, 0.2, 0.3}, {2.0, 4.0, 6.0}, {1.0, 2.0, 1.0}),
      ScaleMode::Fit(), BackgroundMode::FixedZero());
  CHECK(std::fabs(r.scale - 2.0) < 1e-12);
  CHECK(r.background == 0.0);
  CHECK(r.chi2 < 1e-24);
}

CSX_TEST(recovers_solvent_parameters_with_deterministic_noise, true) {
  const Structure s = carbon_oxygen_2a();
  const std::vector<double> q = range_q(1, 80, 0.006);
  SolventParams expected;
  expected.c1 = 1.02;
  expected.c2 = 0.55;
  const std::vector<double> raw =
      multipole_solvent_intensity(s, q, {8, HydrogenMode::Explicit}, expected);
  uint64_t state = 0x5eedull;
  auto uniform = [&]() {
    state = state * 6364136223846793005ull + 1ull;
    return std::max(static_cast<double>(state >> 11) / static_cast<double>(1ull << 53),
                    2.2250738585072014e-308);
  };
  auto gaussian = [&]() {
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  };
  std::vector<double> y;
  for (double v : raw) y.push_back(1.4e-4 * v + 0.03 + 0.0001 * gaussian());
  const ExperimentalCurve e = curve_of(q, y, std::vector<double>(q.size(), 0.002));
  FitOptions o;
  o.method = ScatteringMethod::Multipole;
  o.hydrogen_mode = HydrogenMode::Explicit;
  o.max_iterations = 100;
  o.q_sampling_stride = 1;
  const FitParams start{1.4e-4, 0.03, expected.c1, expected.c2};
  const FitResult r = fit_to_experimental_with_options(s, e, start, o);
  CHECK(std::fabs(r.params.c1 - expected.c1) < 0.02);
  CHECK(std::fabs(r.params.c2 - expected.c2) < 0.08);
  CHECK(std::fabs(r.params.scale - 1.4e-4) < 0.4e-4);
  CHECK(std::fabs(r.params.background - 0.03) < 0.01);
  CHECK(r.chi2 < 0.01);
  // The gradient fitter beside it: chi^2 is nearly flat along a (c1, c2)
  // valley here, so it is held to the objective, not to the parameters.
  o.optimizer = Optimizer::Gradient;
  const FitResult g = fit_to_experimental_with_options(s, e, start, o);
  CHECK(g.chi2 < 0.01 && g.chi2 <= r.chi2 * (1.0 + 1e-6));
}

CSX_TEST(sampled_basis_matches_full_basis_fit, true) {
  const Structure s = carbon_oxygen_2a();
  const std::vector<double> q = range_q(1, 101, 0.005);
  SolventParams p;
  p.c1 = 1.03;
  p.c2 = 0.3;
  const std::vector<double> raw = multipole_solvent_intensity(s, q, {8, HydrogenMode::Explicit}, p);
  std::vector<double> y;
  for (double v : raw) y.push_back(0.8 * v + 0.01);
  const ExperimentalCurve e = curve_of(q, y, std::vector<double>(q.size(), 0.01));
  FitOptions common;
  common.method = ScatteringMethod::Multipole;
  common.hydrogen_mode = HydrogenMode::Explicit;
  common.max_iterations = 64;
  FitOptions full_options = common;
  full_options.q_sampling_stride = 1;
  const FitResult full = fit_to_experimental_with_options(s, e, FitParams{}, full_options);
  const FitResult sampled = fit_to_experimental_with_options(s, e, FitParams{}, common);
  CHECK(relative_l2(sampled.fitted_curve, full.fitted_curve) < 1e-3);
}

// ------------------------------------------------- analysis.rs / ensemble.rs

CSX_TEST(profile_calculator_and_scorer_validate_inputs, false) {
  CHECK(throws<InvalidInput>([] { ProfileCalculator({0.2, 0.1}, ComputeOptions{}); }));
  CHECK(throws<InvalidInput>([] { ProfileCalculator({-0.1}, ComputeOptions{}); }));
  CHECK(throws<InvalidInput>([] { SaxsScorer(ExperimentalCurve{}, ScoreOptions{}); }));
  const ExperimentalCurve e = curve_of({0.1, 0.2}, {1.0, 2.0}, {0.1, 0.1});
  CHECK(throws<InvalidInput>([&] { EnsembleFitter({}, {}, e, EnsembleOptions{}); }));
  CHECK(throws<InvalidInput>([&] { EnsembleFitter({two_carbons()}, {}, e, EnsembleOptions{}); }));
  EnsembleOptions bad;
  bad.min_weight = 1.5;
  CHECK(throws<InvalidInput>([&] { EnsembleFitter({two_carbons()}, {"a"}, e, bad); }));
}

CSX_TEST(compare_curves_metrics, false) {
  const ExperimentalCurve e = curve_of({0.1, 0.2, 0.3, 0.4}, {4.0, 3.0, 2.0, 1.0},
                                       {0.5, 0.5, 0.5, 0.5});
  const ComparisonResult same = compare_curves(e, {4.0, 3.0, 2.0, 1.0});
  CHECK(same.chi2 == 0.0 && same.r_factor == 0.0 && std::fabs(same.pearson_correlation - 1.0) < 1e-15);
  const ComparisonResult off = compare_curves(e, {5.0, 3.0, 2.0, 1.0});
  CHECK(std::fabs(off.chi2 - 4.0) < 1e-15 && std::fabs(off.reduced_chi2 - 4.0 / 3.0) < 1e-15);
  CHECK(std::fabs(off.r_factor - 0.1) < 1e-15 && std::fabs(off.weighted_rmse - 1.0) < 1e-15);
  CHECK(throws<InvalidInput>([&] { compare_curves(e, {1.0}); }));
}

CSX_TEST(profile_calculator_scorer_and_fitter_drop_in, true) {
  std::vector<Atom> atoms;
  for (int i = 0; i < 60; ++i)
    atoms.push_back(atom(3.0 * std::cos(i * 0.9), 3.0 * std::sin(i * 0.9), 0.4 * i - 12.0,
                         i % 3 ? Element::C : Element::O, "A" + std::to_string(i)));
  const Structure s(atoms);
  const std::vector<double> q = range_q(1, 120, 0.004);
  ComputeOptions co;
  co.method = ScatteringMethod::Debye;
  SolventParams truth;
  truth.c1 = 1.02;
  truth.c2 = 1.0;
  co.solvent = truth;
  const ScatteringCurve curve = ProfileCalculator(q, co).calculate(s);
  CHECK(curve.metadata.method == "Debye" && curve.metadata.atom_count == 60);
  std::vector<double> y, sigma;
  for (double v : curve.intensity) {
    y.push_back(2.0 * v + 1.0);
    sigma.push_back(0.01 * (2.0 * v + 1.0));
  }
  const ExperimentalCurve e = curve_of(q, y, sigma);
  FitOptions fo;
  fo.method = ScatteringMethod::Debye;
  fo.q_sampling_stride = 1;
  fo.max_iterations = 200;
  const FitResult fit = SaxsFitter(fo).fit(s, e);
  CHECK(fit.chi2 < 1e-3 && std::fabs(fit.params.scale - 2.0) < 0.02);
  ScoreOptions so;
  so.fit = fo;
  so.quality = Quality::Accurate;
  const ScoreResult score = SaxsScorer(e, so).score(s);
  CHECK(score.points == 120 && score.reduced_chi2 < 1e-3 && score.q_unit == "angstrom^-1");
  CHECK(std::fabs(score.chi2 - score.reduced_chi2 * 119.0) <= 1e-9 * std::max(score.chi2, 1.0));
}

CSX_TEST(coordinate_features_report_pair_distance, true) {
  // metrics.rs:678-684
  const Structure s({atom(0, 0, 0, Element::C, "C1"), atom(4, 0, 0, Element::C, "C2")});
  const CoordinateFeatures f = coordinate_features(s);
  CHECK(std::fabs(f.dmax - 4.0) < 1e-12);
  CHECK(f.rg > 1.0);
  CHECK(!f.p_r.p.empty());
  bool threw = false;
  try {
    coordinate_features(Structure(std::vector<Atom>{}));
  } catch (const InvalidInput&) {
    threw = true;
  }
  CHECK(threw);
}

// ------------------------------------------------------ structure.rs:349-441

CSX_TEST(parses_known_atom_count_and_altloc, false) {
  const Structure s = Structure::from_pdb_file(data_file("test_structure.pdb"));
  CHECK(s.atom_count() == 2);
  for (const Atom& a : s.atoms)
    if (a.atom_name == "CA") CHECK(std::fabs(a.occupancy - 0.8) < 1e-12);
}

CSX_TEST(selected_conformer_can_be_treated_as_full_occupancy, false) {
  ParseOptions o;
  o.occupancy_policy = OccupancyPolicy::SelectedConformerFull;
  const Structure s = Structure::from_pdb_file_with_options(data_file("test_structure.pdb"), o);
  for (const Atom& a : s.atoms)
    if (a.atom_name == "CA") CHECK(a.occupancy == 1.0);
}

CSX_TEST(applies_hetatm_water_and_model_policies, false) {
  const std::string path = data_file("test_structure.pdb");
  ParseOptions het;
  het.include_hetatm = true;
  CHECK(Structure::from_pdb_file_with_options(path, het).atom_count() == 2);
  ParseOptions water = het;
  water.include_waters = true;
  CHECK(Structure::from_pdb_file_with_options(path, water).atom_count() == 3);
  ParseOptions second;
  second.model_index = 1;
  const Structure m = Structure::from_pdb_file_with_options(path, second);
  CHECK(m.atom_count() == 1 && m.atoms[0].pos.x == 9.0);
}

CSX_TEST(parses_first_model_from_mmcif, false) {
  const Structure s = Structure::from_pdb_file(data_file("test_structure.cif"));
  CHECK(s.atom_count() == 2);
  CHECK(s.atoms[0].chain_id == "A");
}

CSX_TEST(malformed_structure_returns_an_error, false) {
  CHECK(throws<SaxsError>([] { Structure::from_pdb_file(data_file("malformed.pdb")); }));
}

// ------------------------------------------------------------ io.rs:171-212

CSX_TEST(parses_validated_three_column_curve, false) {
  const ExperimentalCurve c =
      ExperimentalCurve::parse_dat("header\n# comment\n0.01 2.0 0.1\n0.02,1.5,0.2\n0.03 0.0 0.0\n");
  CHECK((c.q == std::vector<double>{0.01, 0.02}));
  CHECK((c.sigma == std::vector<double>{0.1, 0.2}));
}

CSX_TEST(accepts_two_column_data_and_rejects_nonpositive_sigma_and_unsorted_q, false) {
  const ExperimentalCurve c = ExperimentalCurve::parse_dat("0.01 2.0\n0.02 1.0\n");
  CHECK(!c.has_errors);
  CHECK((c.sigma == std::vector<double>{1.0, 1.0}));
  for (const char* bad : {"0.01 2.0 0\n", "0.02 2.0 0.1\n0.01 1.0 0.1\n", "-0.01 2.0 0.1\n"})
    CHECK(throws<ParseError>([&] { ExperimentalCurve::parse_dat(bad); }));
}

CSX_TEST(normalizes_nm_inverse_angstrom_header, false) {
  const ExperimentalCurve c =
      ExperimentalCurve::parse_dat("# q (1/nm) I sigma\n1.0 2.0 0.1\n2.0 1.0 0.1\n");
  CHECK((c.q == std::vector<double>{0.1, 0.2}));
  CHECK(std::string(c.q_unit()) == "angstrom^-1");
  const ExperimentalCurve f = ExperimentalCurve::from_dat_file(data_file("test_experimental.dat"));
  CHECK(f.len() == 5 && f.has_errors);
}

// ------------------------------------------------------ tests/end_to_end.rs:3-10

CSX_TEST(compute_curve_for_example_structure_is_well_formed, true) {
  const Structure s = Structure::from_pdb_file(data_file("small.pdb"));
  const std::vector<double> q = range_q(0, 99, 0.005);
  const std::vector<double> curve = debye_intensity(s, q);
  CHECK(curve.size() == q.size());
  for (double v : curve) CHECK(std::isfinite(v) && v >= 0.0);
}

// The seam is called from rayon workers (ensemble.rs:244-256): threads sharing
// structures get bitwise the sequential profiles.
CSX_TEST(concurrent_calls_match_sequential, true) {
  std::vector<Structure> shared, fresh;
  for (int k = 0; k < 3; ++k) {
    std::vector<Atom> atoms;
    for (int i = 0; i < 300 + 200 * k; ++i)
      atoms.push_back(atom(9.0 * std::cos(0.37 * i), 9.0 * std::sin(0.37 * i), 0.3 * i - 40.0,
                           i % 5 ? Element::C : Element::O, "A" + std::to_string(i)));
    shared.emplace_back(atoms);
    fresh.emplace_back(atoms);
  }
  const std::vector<double> q = range_q(1, 60, 0.008);
  SolventParams p;
  p.c1 = 1.01;
  p.c2 = 0.5;
  std::vector<std::vector<double>> want;
  for (const Structure& s : fresh) want.push_back(solvent_corrected_intensity(s, q, p, true));
  std::vector<std::vector<double>> got(12);
  std::vector<std::thread> threads;
  for (int i = 0; i < 12; ++i)
    threads.emplace_back([&, i] {
      try {
        got[i] = solvent_corrected_intensity(shared[i % 3], q, p, true);
      } catch (const std::exception&) {
        got[i].clear();  // reported by the comparison below
      }
    });
  for (std::thread& t : threads) t.join();
  for (int i = 0; i < 12; ++i) CHECK(got[i] == want[i % 3]);
}

CSX_TEST(ensemble_fitter_recovers_a_mixture, true) {
  std::vector<Structure> conformers;
  for (int k = 0; k < 4; ++k) {
    std::vector<Atom> atoms;
    for (int i = 0; i < 40 + 10 * k; ++i)
      atoms.push_back(atom((4.0 + k) * std::cos(i * 0.7), (4.0 + k) * std::sin(i * 0.7),
                           0.5 * i - 10.0, i % 4 ? Element::C : Element::N,
                           "A" + std::to_string(i)));
    conformers.emplace_back(atoms);
  }
  const std::vector<double> q = range_q(1, 80, 0.005);
  SolventParams p;
  p.c1 = 1.0;
  p.c2 = 1.0;
  std::vector<double> y(q.size(), 0.0), sigma;
  const double w[4] = {0.0, 0.6, 0.0, 0.4};
  for (int k = 0; k < 4; ++k) {
    const std::vector<double> c = solvent_corrected_intensity(conformers[k], q, p, true);
    for (size_t i = 0; i < q.size(); ++i) y[i] += 1.5 * w[k] * c[i];
  }
  for (double& v : y) {
    v += 2.0;
    sigma.push_back(0.005 * v);
  }
  EnsembleOptions o;
  o.fit.method = ScatteringMethod::Debye;
  o.weight_solver = WeightSolver::NNLS;
  o.bootstrap = 3;
  EnsembleFitter fitter(conformers, {"a", "b", "c", "d"}, curve_of(q, y, sigma), o);
  const EnsembleResult r = fitter.fit();
  double wb = 0.0, wd = 0.0, total = 0.0;
  for (size_t i = 0; i < r.conformers.size(); ++i) {
    total += r.weights[i];
    if (r.conformers[i] == "b") wb = r.weights[i];
    if (r.conformers[i] == "d") wd = r.weights[i];
  }
  CHECK(std::fabs(total - 1.0) < 1e-12);
  CHECK(std::fabs(wb - 0.6) < 1e-4 && std::fabs(wd - 0.4) < 1e-4);
  CHECK(r.reduced_chi2 < 1e-6 && r.bootstrap_reduced_chi2.size() == 3);
}

// Without a HIP device every compute call reports NoDevice (SaxsError) and
// nothing else fails or leaks; host-only calls keep working.  With a device
// the same calls succeed.  (No reference counterpart: the reference has no
// device.)
CSX_TEST(no_device_paths_report_cleanly, false) {
  Structure s({atom(0, 0, 0, Element::C, "C1"), atom(2, 0, 0, Element::C, "C2")});
  const std::vector<double> q = {0.0, 0.1, 0.2};
  if (csx_device_count() == 0) {
    CHECK(throws<NoDevice>([] { Context ctx(0); }));
    csx_ctx* raw = reinterpret_cast<csx_ctx*>(0x1);
    CHECK(csx_ctx_create(0, &raw) == CSX_NO_DEVICE && raw == nullptr);
    CHECK(std::string(csx_last_error()).find("no HIP device") != std::string::npos);
  } else {
    Context ctx(0);
    CHECK(debye_intensity(s, q, Precision::FP32Compensated, ctx).size() == q.size());
  }
  // host-only paths
  CHECK(std::fabs(csx_water_form_factor(0.0) - 10.0) < 1e-3);
  std::vector<double> basis(6 * q.size(), 1.0), out(q.size());
  CHECK(csx_combine_solvent_basis(basis.data(), q.data(), q.size(), 1.0, 0.0, out.data()) == CSX_OK);
  uint32_t words[4];
  CHECK(csx_pcg32(1, 3, 4, 1, words) == CSX_OK);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "--all";
  const std::string self = argv[0];
  const size_t slash = self.rfind('/');
  if (slash != std::string::npos) g_data_dir = self.substr(0, slash) + "/../golden/structures/";
  int run = 0, failed = 0;
  for (const TestCase& t : registry()) {
    if (mode == "--list") {
      std::printf("%s%s\n", t.name, t.gpu ? " [gpu]" : "");
      continue;
    }
    const bool pick = mode == "--all" || (mode == "--cpu" && !t.gpu) ||
                      (mode == "--gpu" && t.gpu) || mode == t.name;
    if (!pick) continue;
    ++run;
    std::printf("%-72s ", t.name);
    std::fflush(stdout);
    try {
      t.fn();
      std::printf("ok\n");
    } catch (const Failure& f) {
      ++failed;
      std::printf("FAILED\n    %s\n", f.what.c_str());
    } catch (const std::exception& e) {
      ++failed;
      std::printf("FAILED\n    exception: %s\n", e.what());
    }
    std::fflush(stdout);
  }
  if (mode != "--list") std::printf("%d run, %d failed\n", run, failed);
  return failed == 0 && (run > 0 || mode == "--list") ? 0 : 1;
}

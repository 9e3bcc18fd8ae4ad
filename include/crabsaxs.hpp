// crabsaxs.hpp -- C++17 host API over the C ABI of libcrabsaxs_hip.
//
// The types and functions of crabSAXS's profile path with the reference's
// names, fields, defaults and validation, so that code written against the
// reference's API (src/scattering.rs, src/solvent.rs, src/spherical.rs,
// src/fit.rs, src/analysis.rs, src/ensemble.rs) reads the same.  Functions
// that return `Result<_, SaxsError>` in the reference throw a SaxsError
// subclass here.  All computation runs in the HIP library on the GPU of a
// Context (the default one unless given); there is no CPU fallback.
//
// Header-only; link with -lcrabsaxs_hip.
#ifndef CRABSAXS_HPP
#define CRABSAXS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "crabsaxs.h"
#include "crabsaxs_tables.h"

namespace crabsaxs {

// ------------------------------------------------------------------ errors
// SaxsError (src/error.rs:4-27).

class SaxsError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};
class InvalidInput : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class FitDidNotConverge : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class NoDevice : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class DeviceError : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class OutOfMemory : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
// StructureParse and ExperimentalDataParse.
class ParseError : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class UnsupportedElement : public SaxsError {
 public:
  using SaxsError::SaxsError;
};
class IoError : public SaxsError {
 public:
  using SaxsError::SaxsError;
};

namespace detail {
inline void check(csx_status s) {
  if (s == CSX_OK) return;
  const std::string msg = csx_last_error();
  switch (s) {
    case CSX_INVALID_INPUT: throw InvalidInput(msg);
    case CSX_NOT_CONVERGED: throw FitDidNotConverge(msg);
    case CSX_NO_DEVICE: throw NoDevice(msg);
    case CSX_OUT_OF_MEMORY: throw OutOfMemory(msg);
    case CSX_PARSE_ERROR: throw ParseError(msg);
    case CSX_UNSUPPORTED_ELEMENT: throw UnsupportedElement(msg);
    case CSX_IO_ERROR: throw IoError(msg);
    default: throw DeviceError(msg);
  }
}
}  // namespace detail

// --------------------------------------------------------------- structure
// Element, Atom, Structure (src/structure.rs:90-165).

enum class Element : uint8_t {
  H = CSX_EL_H, C = CSX_EL_C, N = CSX_EL_N, O = CSX_EL_O, S = CSX_EL_S, P = CSX_EL_P,
  Na = CSX_EL_NA, Mg = CSX_EL_MG, Cl = CSX_EL_CL, Ca = CSX_EL_CA, Fe = CSX_EL_FE,
  Zn = CSX_EL_ZN, K = CSX_EL_K, I = CSX_EL_I, Other = CSX_EL_OTHER
};

inline Element element_from_symbol(const std::string& symbol) {
  return static_cast<Element>(csx_element_from_symbol(symbol.c_str()));
}

struct Vector3 {
  double x = 0.0, y = 0.0, z = 0.0;
};

struct Atom {
  Vector3 pos;
  Element element = Element::C;
  double occupancy = 1.0;
  bool is_hetatm = false;
  std::string atom_name;
  std::string residue_name = "UNK";
  std::string chain_id = "A";
  long residue_id = 0;
};

struct ParseOptions;

// A HIP device context (one per GPU; reentrant).
class Context {
 public:
  explicit Context(int device = 0) { detail::check(csx_ctx_create(device, &ptr_)); }
  ~Context() {
    if (ptr_) csx_ctx_destroy(ptr_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  csx_ctx* get() const { return ptr_; }

  // The process-wide context on device 0, created on first use and never
  // destroyed: a static destructor would run after the HIP runtime has been
  // torn down at exit and release streams and memory into it.
  static Context& default_context() {
    static std::mutex mu;
    static Context* ctx = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx) ctx = new Context(0);
    return *ctx;
  }

  void set_timing(bool on) { detail::check(csx_ctx_set_timing(ptr_, on ? 1 : 0)); }
  void reset_timing() { detail::check(csx_ctx_reset_timing(ptr_)); }
  // (launches, total milliseconds) of a kernel while timing was on.
  std::pair<int64_t, double> kernel_time(const std::string& name) const {
    int64_t n = 0;
    double ms = 0.0;
    detail::check(csx_ctx_kernel_time(ptr_, name.c_str(), &n, &ms));
    return {n, ms};
  }
  // Pair-kernel launch shape (csx_ctx_set_tuning); sin_mode 3, the default,
  // takes the sine recurrence on uniform q grids.
  void set_tuning(int q_tile = 0, int sin_mode = 3, int j_chunk = 0) {
    detail::check(csx_ctx_set_tuning(ptr_, q_tile, sin_mode, j_chunk));
  }
  void set_vjp_tile(int q_tile = 0) { detail::check(csx_ctx_set_vjp_tile(ptr_, q_tile)); }
  void set_persistence(int workgroups_per_cu = -1) {
    detail::check(csx_ctx_set_persistence(ptr_, workgroups_per_cu));
  }
  void set_workspace_limit(size_t bytes) {
    detail::check(csx_ctx_set_workspace_limit(ptr_, bytes));
  }
  // Cached pair tables and multipole coefficients per structure (0: off).
  void set_cache_limit(size_t bytes) { detail::check(csx_ctx_set_cache_limit(ptr_, bytes)); }
  int cu_count() const {
    int n = 0;
    detail::check(csx_ctx_device_info(ptr_, &n, nullptr, 0));
    return n;
  }
  std::string arch() const {
    char buf[256] = {0};
    detail::check(csx_ctx_device_info(ptr_, nullptr, buf, sizeof buf));
    return buf;
  }

 private:
  csx_ctx* ptr_ = nullptr;
};

class Structure {
 public:
  std::vector<Atom> atoms;

  Structure() = default;
  explicit Structure(std::vector<Atom> a) : atoms(std::move(a)) {}
  Structure(const Structure& o) : atoms(o.atoms) { copy_sites(o); }
  Structure& operator=(const Structure& o) {
    if (this == &o) return *this;
    atoms = o.atoms;
    copy_sites(o);
    std::lock_guard<std::mutex> lock(state_->mu);
    state_->uploads.clear();
    return *this;
  }
  Structure(Structure&& o) : atoms(std::move(o.atoms)), state_(std::move(o.state_)) {
    o.state_.reset(new State);
  }
  Structure& operator=(Structure&& o) {
    atoms = std::move(o.atoms);
    std::swap(state_, o.state_);
    std::lock_guard<std::mutex> lock(o.state_->mu);
    o.state_->uploads.clear();
    return *this;
  }

  size_t atom_count() const { return atoms.size(); }

  // Hydration sites used with HydrationModel::ExplicitSites (no reference
  // counterpart): 3 coordinates per site and one voxel volume.
  void set_hydration_sites(std::vector<double> xyz, double voxel_volume) {
    if (xyz.size() % 3 != 0) throw InvalidInput("hydration sites need 3 coordinates each");
    std::lock_guard<std::mutex> lock(state_->mu);
    state_->sites = std::move(xyz);
    state_->site_volume = voxel_volume;
    state_->has_sites = true;
    ++state_->sites_version;
  }

  // PDB / mmCIF input (src/structure.rs:205-346); defined with ParseOptions.
  static Structure from_pdb_file(const std::string& path);
  static Structure from_pdb_file_with_options(const std::string& path, const ParseOptions& o);
  static std::vector<Structure> from_pdb_file_models_with_options(const std::string& path,
                                                                  const ParseOptions& o);
  // PDB (CSX_FORMAT_PDB) or mmCIF (CSX_FORMAT_MMCIF) text; CSX_FORMAT_AUTO detects.
  static Structure parse(const std::string& text, int format, const ParseOptions& o);

  // The device copy of this structure's current atoms on `ctx`: uploaded on
  // first use and again whenever the atoms have changed since.  One copy per
  // context, so threads may share a structure across contexts; changing the
  // atoms while another thread computes on the structure is a data race.
  csx_structure* handle(Context& ctx) const {
    const uint64_t fp = fingerprint();
    std::lock_guard<std::mutex> lock(state_->mu);
    std::unique_ptr<Upload>* slot = nullptr;
    for (auto& u : state_->uploads)
      if (u->ctx == ctx.get()) slot = &u;
    if (slot && (*slot)->fingerprint == fp && (*slot)->n == atoms.size()) {
      attach_sites(**slot);
      return (*slot)->h;
    }
    const size_t n = atoms.size();
    std::vector<double> xyz(3 * n), occ(n);
    std::vector<uint8_t> el(n);
    std::vector<int8_t> group(n);
    for (size_t i = 0; i < n; ++i) {
      const Atom& a = atoms[i];
      xyz[3 * i] = a.pos.x;
      xyz[3 * i + 1] = a.pos.y;
      xyz[3 * i + 2] = a.pos.z;
      occ[i] = a.occupancy;
      el[i] = static_cast<uint8_t>(a.element);
      group[i] = static_cast<int8_t>(csx_united_atom_group(a.residue_name.c_str(),
                                                           a.atom_name.c_str()));
    }
    std::unique_ptr<Upload> fresh(new Upload);
    detail::check(csx_structure_create(ctx.get(), n, xyz.data(), el.data(), group.data(),
                                       occ.data(), &fresh->h));
    fresh->ctx = ctx.get();
    fresh->fingerprint = fp;
    fresh->n = n;
    csx_structure* h = fresh->h;
    attach_sites(*fresh);
    if (slot)
      *slot = std::move(fresh);
    else
      state_->uploads.push_back(std::move(fresh));
    return h;
  }

 private:
  struct Upload {
    csx_ctx* ctx = nullptr;
    uint64_t fingerprint = 0;
    size_t n = 0;
    uint64_t sites_version = 0;
    csx_structure* h = nullptr;
    ~Upload() {
      if (h) csx_structure_destroy(h);
    }
  };
  struct State {
    std::mutex mu;
    std::vector<std::unique_ptr<Upload>> uploads;  // one per context
    bool has_sites = false;
    std::vector<double> sites;
    double site_volume = 0.0;
    uint64_t sites_version = 0;
  };
  void copy_sites(const Structure& o) {
    bool has;
    std::vector<double> xyz;
    double vol;
    {
      std::lock_guard<std::mutex> lock(o.state_->mu);
      has = o.state_->has_sites;
      xyz = o.state_->sites;
      vol = o.state_->site_volume;
    }
    std::lock_guard<std::mutex> lock(state_->mu);
    state_->has_sites = has;
    state_->sites = std::move(xyz);
    state_->site_volume = vol;
    ++state_->sites_version;
  }
  // Called with state_->mu held.
  void attach_sites(Upload& u) const {
    if (!state_->has_sites || u.sites_version == state_->sites_version) return;
    detail::check(csx_structure_set_hydration_sites(u.h, state_->sites.data(),
                                                    state_->sites.size() / 3,
                                                    state_->site_volume));
    u.sites_version = state_->sites_version;
  }
  std::unique_ptr<State> state_{new State};

  static uint64_t mix(uint64_t h, uint64_t w) {
    h ^= w + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h * 0xBF58476D1CE4E5B9ull;
  }
  static uint64_t bits(double d) {
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return u;
  }
  uint64_t fingerprint() const {
    uint64_t h = 0x243F6A8885A308D3ull ^ atoms.size();
    for (const Atom& a : atoms) {
      h = mix(h, bits(a.pos.x));
      h = mix(h, bits(a.pos.y));
      h = mix(h, bits(a.pos.z));
      h = mix(h, bits(a.occupancy) ^ static_cast<uint64_t>(a.element));
      for (char c : a.atom_name) h = mix(h, static_cast<unsigned char>(c));
      for (char c : a.residue_name) h = mix(h, static_cast<unsigned char>(c) | 0x100u);
    }
    return h;
  }
};

// ------------------------------------------------------------------ ingest
// AltlocPolicy, ResidueKind, OccupancyPolicy, ParseOptions
// (src/structure.rs:13-202).

enum class AltlocPolicy { Highest = CSX_ALTLOC_HIGHEST, First = CSX_ALTLOC_FIRST, All = CSX_ALTLOC_ALL };
enum class OccupancyPolicy {
  Reported = CSX_OCCUPANCY_REPORTED,
  SelectedConformerFull = CSX_OCCUPANCY_SELECTED_FULL
};
enum class ResidueKind {
  Polymer = CSX_RESIDUE_POLYMER,
  Glycan = CSX_RESIDUE_GLYCAN,
  Water = CSX_RESIDUE_WATER,
  Other = CSX_RESIDUE_OTHER
};

inline ResidueKind residue_kind(const std::string& name) {
  return static_cast<ResidueKind>(csx_residue_kind(name.c_str()));
}

struct ParseOptions {
  bool include_hetatm = false;
  bool include_waters = false;
  bool include_glycans = true;
  size_t model_index = 0;
  bool allow_unknown_elements = false;
  OccupancyPolicy occupancy_policy = OccupancyPolicy::Reported;
  AltlocPolicy altloc_policy = AltlocPolicy::Highest;
  std::optional<std::vector<std::string>> chains;
};

namespace detail {
// csx_parse_options of o; `chains` holds the comma-joined chain list.
inline csx_parse_options parse_options(const ParseOptions& o, std::string& chains) {
  csx_parse_options c;
  csx_parse_options_default(&c);
  c.include_hetatm = o.include_hetatm;
  c.include_waters = o.include_waters;
  c.include_glycans = o.include_glycans;
  c.model_index = o.model_index;
  c.allow_unknown_elements = o.allow_unknown_elements;
  c.occupancy_policy = static_cast<int>(o.occupancy_policy);
  c.altloc_policy = static_cast<int>(o.altloc_policy);
  if (o.chains) {
    for (size_t i = 0; i < o.chains->size(); ++i) chains += (i ? "," : "") + (*o.chains)[i];
    c.chains = chains.c_str();
  }
  return c;
}

// Structure of a parsed atom set (takes ownership of h).
inline Structure structure_of(csx_atoms* h) {
  std::unique_ptr<csx_atoms, void (*)(csx_atoms*)> owner(h, csx_atoms_destroy);
  std::vector<csx_atom_record> recs(csx_atoms_count(h));
  check(csx_atoms_get(h, recs.data()));
  std::vector<Atom> atoms(recs.size());
  for (size_t i = 0; i < recs.size(); ++i) {
    const csx_atom_record& r = recs[i];
    atoms[i].pos = {r.x, r.y, r.z};
    atoms[i].element = static_cast<Element>(r.element);
    atoms[i].occupancy = r.occupancy;
    atoms[i].is_hetatm = r.is_hetatm != 0;
    atoms[i].atom_name = r.atom_name;
    atoms[i].residue_name = r.residue_name;
    atoms[i].chain_id = r.chain_id;
    atoms[i].residue_id = static_cast<long>(r.residue_id);
  }
  return Structure(std::move(atoms));
}
}  // namespace detail

inline Structure Structure::from_pdb_file_with_options(const std::string& path,
                                                       const ParseOptions& o) {
  std::string chains;
  const csx_parse_options c = detail::parse_options(o, chains);
  csx_atoms* h = nullptr;
  detail::check(csx_read_structure_file(path.c_str(), &c, &h));
  return detail::structure_of(h);
}

inline Structure Structure::from_pdb_file(const std::string& path) {
  return from_pdb_file_with_options(path, ParseOptions{});
}

inline std::vector<Structure> Structure::from_pdb_file_models_with_options(
    const std::string& path, const ParseOptions& o) {
  std::vector<Structure> models;
  for (size_t index = 0;; ++index) {
    ParseOptions m = o;
    m.model_index = index;
    try {
      models.push_back(from_pdb_file_with_options(path, m));
    } catch (const SaxsError&) {
      if (index > 0 && !models.empty()) break;
      throw;
    }
  }
  return models;
}

inline Structure Structure::parse(const std::string& text, int format, const ParseOptions& o) {
  std::string chains;
  const csx_parse_options c = detail::parse_options(o, chains);
  csx_atoms* h = nullptr;
  detail::check(csx_parse_structure(text.data(), text.size(), format, &c, &h));
  return detail::structure_of(h);
}

// ------------------------------------------------------------- form factors

enum class HydrogenMode { Implicit, Explicit };

inline bool implicit(HydrogenMode m) { return m == HydrogenMode::Implicit; }

// form_factor_for_atom (src/formfactor.rs:425-439).
inline double form_factor_for_atom(const Atom& a, bool implicit_hydrogen, double q) {
  const int g = csx_united_atom_group(a.residue_name.c_str(), a.atom_name.c_str());
  return csx_form_factor(static_cast<int>(a.element), g, a.occupancy, implicit_hydrogen ? 1 : 0,
                         q);
}

// excluded_volume_for_atom (src/solvent.rs:428-438).
inline double excluded_volume_for_atom(const Atom& a, bool implicit_hydrogen) {
  const int g = csx_united_atom_group(a.residue_name.c_str(), a.atom_name.c_str());
  return csx_excluded_volume(static_cast<int>(a.element), g, implicit_hydrogen ? 1 : 0);
}

// ------------------------------------------------------------------ solvent
// HydrationModel, SolventParams (src/solvent.rs:17-50).

// ExplicitSites (no reference counterpart): the sites given with
// Structure::set_hydration_sites, entering like the sites of a Grid model.
enum class HydrationModel { Analytic = 0, Grid = 1, AdaptiveGrid = 2, ExplicitSites = 3 };

struct SolventParams {
  double c1 = 1.0;
  double c2 = 0.0;
  double shell_thickness_angstrom = 3.0;
  double bulk_density_e_per_a3 = 0.334;
  HydrationModel hydration_model = HydrationModel::Analytic;
  double hydration_grid_spacing_angstrom = 4.0;

  csx_solvent_params to_c() const {
    return {c1, c2, shell_thickness_angstrom, bulk_density_e_per_a3,
            static_cast<int32_t>(hydration_model), hydration_grid_spacing_angstrom};
  }
};

// Pair-sum arithmetic of the Debye path (no reference counterpart).
enum class Precision { FP64Plain = CSX_FP64_PLAIN, FP32Compensated = CSX_FP32_COMPENSATED };

using Basis = std::array<std::vector<double>, 6>;  // vv, ee, hh, ve, vh, eh

namespace detail {
inline Basis split6(const std::vector<double>& flat, size_t nq) {
  Basis b;
  for (int t = 0; t < 6; ++t) b[t].assign(flat.begin() + t * nq, flat.begin() + (t + 1) * nq);
  return b;
}
inline std::vector<double> join6(const Basis& b) {
  const size_t nq = b[0].size();
  std::vector<double> flat(6 * nq);
  for (int t = 0; t < 6; ++t) {
    if (b[t].size() != nq) throw InvalidInput("basis terms differ in length");
    std::copy(b[t].begin(), b[t].end(), flat.begin() + t * nq);
  }
  return flat;
}
}  // namespace detail

// excluded_volume_scale (src/solvent.rs:448-452).
inline double excluded_volume_scale(double q, double radius_scale) {
  return csx_excluded_volume_scale(q, radius_scale);
}

// water_form_factor (src/solvent.rs:227-231).
inline double water_form_factor(double q) { return csx_water_form_factor(q); }

// debye_solvent_basis (src/solvent.rs:504-584).
inline Basis debye_solvent_basis(const Structure& s, const std::vector<double>& q,
                                 const SolventParams& p, bool implicit_hydrogen,
                                 Precision precision = Precision::FP32Compensated,
                                 Context& ctx = Context::default_context()) {
  std::vector<double> out(6 * q.size());
  const csx_solvent_params c = p.to_c();
  detail::check(csx_debye_solvent_basis(s.handle(ctx), q.data(), q.size(), &c,
                                        implicit_hydrogen ? 1 : 0, static_cast<int>(precision),
                                        out.data()));
  return detail::split6(out, q.size());
}

// solvent_corrected_intensity (src/solvent.rs:490-500).
inline std::vector<double> solvent_corrected_intensity(
    const Structure& s, const std::vector<double>& q, const SolventParams& p,
    bool implicit_hydrogen, Precision precision = Precision::FP32Compensated,
    Context& ctx = Context::default_context()) {
  std::vector<double> out(q.size());
  const csx_solvent_params c = p.to_c();
  detail::check(csx_solvent_corrected_intensity(s.handle(ctx), q.data(), q.size(), &c,
                                                implicit_hydrogen ? 1 : 0,
                                                static_cast<int>(precision), out.data()));
  return out;
}

// HydrationGrid and its builders (src/solvent.rs:53-401).
struct HydrationGrid {
  std::vector<Vector3> sites;
  double voxel_volume_a3 = 0.0;
};
struct AdaptiveHydrationOptions {
  double spacing_angstrom = 4.0;
  double molecular_clearance_angstrom = 3.0;
  double padding_angstrom = 12.0;
};
struct SolventGridOptions {
  double spacing_angstrom = 1.0;
  double probe_radius_angstrom = 1.4;
  double shell_thickness_angstrom = 3.0;
};

namespace detail {
template <class Call>
HydrationGrid grid(Call call) {
  size_t count = 0;
  double vol = 0.0;
  check(call(nullptr, 0, &count, &vol));
  std::vector<double> sites(3 * count);
  check(call(sites.data(), count, &count, &vol));
  HydrationGrid g;
  g.voxel_volume_a3 = vol;
  g.sites.resize(count);
  for (size_t k = 0; k < count; ++k) g.sites[k] = {sites[3 * k], sites[3 * k + 1], sites[3 * k + 2]};
  return g;
}
}  // namespace detail

// Both grids are built on the structure's device (bitwise the reference's
// sites, in its scan order).
inline HydrationGrid build_adaptive_hydration_grid(const Structure& s,
                                                   const AdaptiveHydrationOptions& o,
                                                   Context& ctx = Context::default_context()) {
  csx_structure* h = s.handle(ctx);
  return detail::grid([&](double* sites, size_t cap, size_t* count, double* vol) {
    return csx_build_adaptive_hydration_grid(h, o.spacing_angstrom, o.molecular_clearance_angstrom,
                                             o.padding_angstrom, sites, cap, count, vol);
  });
}

inline HydrationGrid build_hydration_grid(const Structure& s, const SolventGridOptions& o,
                                          Context& ctx = Context::default_context()) {
  csx_structure* h = s.handle(ctx);
  return detail::grid([&](double* sites, size_t cap, size_t* count, double* vol) {
    return csx_build_hydration_grid(h, o.spacing_angstrom, o.probe_radius_angstrom,
                                    o.shell_thickness_angstrom, sites, cap, count, vol);
  });
}

// ---------------------------------------------------------------- scattering
// src/scattering.rs.

enum class ScatteringMethod { Auto = CSX_METHOD_AUTO, Debye = CSX_METHOD_DEBYE,
                              Multipole = CSX_METHOD_MULTIPOLE };

// sinc (scattering.rs:62-69).
inline double sinc(double x) { return std::fabs(x) < 1e-8 ? 1.0 : std::sin(x) / x; }

struct MultipoleOptions {
  std::optional<size_t> l_max;
};

struct ComputeOptions {
  ScatteringMethod method = ScatteringMethod::Auto;
  HydrogenMode hydrogen_mode = HydrogenMode::Implicit;
  bool include_hetatm = false;
  MultipoleOptions multipole;
  std::optional<SolventParams> solvent;
  Precision precision = Precision::FP32Compensated;  // Debye pair sums
};

struct CurveMetadata {
  std::string method;
  size_t atom_count = 0;
};

struct ScatteringCurve {
  std::vector<double> q;
  std::vector<double> intensity;
  CurveMetadata metadata;
};

// debye_intensity_with_hydrogens (scattering.rs:86-110).
inline std::vector<double> debye_intensity_with_hydrogens(
    const Structure& s, const std::vector<double>& q, bool implicit_hydrogen,
    Precision precision = Precision::FP32Compensated,
    Context& ctx = Context::default_context()) {
  std::vector<double> out(q.size());
  detail::check(csx_debye_intensity(s.handle(ctx), q.data(), q.size(), implicit_hydrogen ? 1 : 0,
                                    static_cast<int>(precision), out.data()));
  return out;
}

// debye_intensity (scattering.rs:82-84).
inline std::vector<double> debye_intensity(const Structure& s, const std::vector<double>& q,
                                           Precision precision = Precision::FP32Compensated,
                                           Context& ctx = Context::default_context()) {
  return debye_intensity_with_hydrogens(s, q, true, precision, ctx);
}

// compute_curve (scattering.rs:112-212): validation, Auto choice, adaptive
// l_max and dispatch run in the library.
inline ScatteringCurve compute_curve(const Structure& s, const std::vector<double>& q,
                                     const ComputeOptions& o,
                                     Context& ctx = Context::default_context()) {
  csx_compute_options c;
  csx_compute_options_default(&c);
  c.method = static_cast<int32_t>(o.method);
  c.implicit_hydrogen = implicit(o.hydrogen_mode) ? 1 : 0;
  c.has_solvent = o.solvent.has_value() ? 1 : 0;
  if (o.solvent) c.solvent = o.solvent->to_c();
  if (o.multipole.l_max) {
    // the library range-checks l_max (8..=64) with the reference's message
    c.l_max = static_cast<int32_t>(std::min<size_t>(*o.multipole.l_max, 1 << 20));
    if (c.l_max == 0) c.l_max = -1;
  }
  c.precision = static_cast<int32_t>(o.precision);
  ScatteringCurve curve;
  curve.q = q;
  curve.intensity.resize(q.size());
  int used = 0;
  detail::check(csx_compute_curve(s.handle(ctx), q.data(), q.size(), &c, curve.intensity.data(),
                                  &used));
  curve.metadata.method = used == CSX_METHOD_DEBYE ? "Debye" : "Multipole";
  curve.metadata.atom_count = s.atom_count();
  return curve;
}

// ----------------------------------------------------------------- multipole
// MultipoleParams (src/spherical.rs:19-31).
struct MultipoleParams {
  size_t l_max = 20;
  HydrogenMode hydrogen_mode = HydrogenMode::Implicit;
};

// multipole_intensity (spherical.rs:266-312).
inline std::vector<double> multipole_intensity(const Structure& s, const std::vector<double>& q,
                                               const MultipoleParams& p,
                                               Context& ctx = Context::default_context()) {
  std::vector<double> out(q.size());
  detail::check(csx_multipole_intensity(s.handle(ctx), q.data(), q.size(),
                                        static_cast<int>(p.l_max), implicit(p.hydrogen_mode),
                                        out.data()));
  return out;
}

// multipole_solvent_intensity (spherical.rs:408-470).
inline std::vector<double> multipole_solvent_intensity(
    const Structure& s, const std::vector<double>& q, const MultipoleParams& p,
    const SolventParams& solvent, Context& ctx = Context::default_context()) {
  std::vector<double> out(q.size());
  const csx_solvent_params c = solvent.to_c();
  detail::check(csx_multipole_solvent_intensity(s.handle(ctx), q.data(), q.size(),
                                                static_cast<int>(p.l_max),
                                                implicit(p.hydrogen_mode), &c, out.data()));
  return out;
}

// multipole_solvent_basis (spherical.rs:474-565).
inline Basis multipole_solvent_basis(const Structure& s, const std::vector<double>& q,
                                     const MultipoleParams& p, const SolventParams& solvent,
                                     Context& ctx = Context::default_context()) {
  std::vector<double> out(6 * q.size());
  const csx_solvent_params c = solvent.to_c();
  detail::check(csx_multipole_solvent_basis(s.handle(ctx), q.data(), q.size(),
                                            static_cast<int>(p.l_max), implicit(p.hydrogen_mode),
                                            &c, out.data()));
  return detail::split6(out, q.size());
}

// combine_solvent_basis (spherical.rs:567-584).
inline std::vector<double> combine_solvent_basis(const Basis& basis, const std::vector<double>& q,
                                                 double c1, double c2) {
  const std::vector<double> flat = detail::join6(basis);
  if (flat.size() != 6 * q.size()) throw InvalidInput("basis and q differ in length");
  std::vector<double> out(q.size());
  detail::check(csx_combine_solvent_basis(flat.data(), q.data(), q.size(), c1, c2, out.data()));
  return out;
}

// ---------------------------------------------- batches and the adjoint
// (no reference counterparts)

namespace detail {
inline std::vector<csx_structure*> handles(const std::vector<const Structure*>& ss,
                                           Context& ctx) {
  std::vector<csx_structure*> h(ss.size());
  for (size_t k = 0; k < ss.size(); ++k) h[k] = ss[k]->handle(ctx);
  return h;
}
}  // namespace detail

// Six-term bases of many structures from one launch of the pair kernel.
inline std::vector<Basis> batch_debye_solvent_basis(const std::vector<const Structure*>& ss,
                                                    const std::vector<double>& q,
                                                    const SolventParams& p, bool implicit_hydrogen,
                                                    Context& ctx = Context::default_context()) {
  const std::vector<csx_structure*> h = detail::handles(ss, ctx);
  std::vector<double> out(ss.size() * 6 * q.size());
  const csx_solvent_params c = p.to_c();
  detail::check(csx_batch_debye_solvent_basis(h.data(), h.size(), q.data(), q.size(), &c,
                                              implicit_hydrogen ? 1 : 0, out.data()));
  std::vector<Basis> bases(ss.size());
  for (size_t k = 0; k < ss.size(); ++k)
    bases[k] = detail::split6(
        std::vector<double>(out.begin() + k * 6 * q.size(), out.begin() + (k + 1) * 6 * q.size()),
        q.size());
  return bases;
}

// sum_q weights[q] dI(q)/dx_i for every atom (n x 3, row-major); `solvent`
// empty differentiates the vacuum Debye intensity.
inline std::vector<double> debye_grad_x(const Structure& s, const std::vector<double>& q,
                                        const std::vector<double>& weights,
                                        const std::optional<SolventParams>& solvent,
                                        bool implicit_hydrogen,
                                        Context& ctx = Context::default_context()) {
  if (weights.size() != q.size()) throw InvalidInput("weights and q differ in length");
  std::vector<double> out(3 * s.atom_count());
  csx_solvent_params c{};
  if (solvent) c = solvent->to_c();
  detail::check(csx_debye_grad_x(s.handle(ctx), q.data(), q.size(), weights.data(),
                                 solvent ? &c : nullptr, implicit_hydrogen ? 1 : 0, out.data()));
  return out;
}

// ----------------------------------------------------------------- metrics
// PrOptions, PairDistribution, CoordinateFeatures (src/metrics.rs:15-37, 61-65).

struct PrOptions {
  size_t bins = 160;
  std::optional<double> max_r;  // used by the experimental transform only
};

struct PairDistribution {
  std::vector<double> r, p;
  std::optional<double> rg, dmax;
};

struct CoordinateFeatures {
  double rg = 0.0, dmax = 0.0;
  PairDistribution p_r;
};

// coordinate_features (metrics.rs:236-270): weighted Rg and Dmax of the
// coordinates and their pair-distance distribution (pair histogram on the GPU).
inline CoordinateFeatures coordinate_features(const Structure& s, const PrOptions& o = {},
                                              Context& ctx = Context::default_context()) {
  const size_t nb = std::max<size_t>(o.bins, 16);
  std::vector<double> r(nb), p(nb);
  size_t n_out = 0;
  double pr_rg = 0.0, pr_dmax = 0.0;
  int flags = 0;
  CoordinateFeatures f;
  detail::check(csx_coordinate_features(s.handle(ctx), o.bins, &f.rg, &f.dmax, r.data(), p.data(),
                                        &n_out, &pr_rg, &pr_dmax, &flags));
  r.resize(n_out);
  p.resize(n_out);
  f.p_r.r = std::move(r);
  f.p_r.p = std::move(p);
  if (flags & 1) f.p_r.rg = pr_rg;
  if (flags & 2) f.p_r.dmax = pr_dmax;
  return f;
}

// --------------------------------------------------------------------- io
// ExperimentalCurve (src/io.rs:8-16, 113-141).

struct ExperimentalCurve {
  std::vector<double> q, intensity, sigma;
  bool has_errors = false;

  size_t len() const { return q.size(); }
  bool is_empty() const { return q.empty(); }
  const char* q_unit() const { return "angstrom^-1"; }
  // ExperimentalCurve::from_dat_file (src/io.rs:19-112).
  static ExperimentalCurve from_dat_file(const std::string& path) {
    return read([&](double* q, double* i, double* s, size_t cap, size_t* n, int* has) {
      return csx_read_dat_file(path.c_str(), q, i, s, cap, n, has);
    });
  }
  // The same rules applied to text in memory.
  static ExperimentalCurve parse_dat(const std::string& text) {
    return read([&](double* q, double* i, double* s, size_t cap, size_t* n, int* has) {
      return csx_parse_dat(text.data(), text.size(), q, i, s, cap, n, has);
    });
  }
  ExperimentalCurve through_qmax(double qmax) const {
    ExperimentalCurve c;
    c.has_errors = has_errors;
    for (size_t i = 0; i < q.size(); ++i)
      if (q[i] <= qmax) {
        c.q.push_back(q[i]);
        c.intensity.push_back(intensity[i]);
        c.sigma.push_back(sigma[i]);
      }
    return c;
  }

 private:
  template <class Call>
  static ExperimentalCurve read(Call call) {
    size_t n = 0;
    int has = 0;
    detail::check(call(nullptr, nullptr, nullptr, 0, &n, &has));
    ExperimentalCurve c;
    c.q.resize(n);
    c.intensity.resize(n);
    c.sigma.resize(n);
    detail::check(call(c.q.data(), c.intensity.data(), c.sigma.data(), n, &n, &has));
    c.has_errors = has != 0;
    return c;
  }
};

namespace detail {
// Rust's `{:.8e}`: 8 decimals, exponent without '+' or leading zeros.
inline std::string rust_exp(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.8e", v);
  std::string s(buf);
  const size_t e = s.find('e');
  return s.substr(0, e + 1) + std::to_string(std::stoi(s.substr(e + 1)));
}
inline void write_curve(const std::string& path, const std::vector<double>& q,
                        const std::vector<double>& intensity, const char* header, char sep) {
  if (q.size() != intensity.size()) throw InvalidInput("q and intensity lengths differ");
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw IoError("cannot write " + path);
  std::fprintf(f, "%s\n", header);
  for (size_t i = 0; i < q.size(); ++i)
    std::fprintf(f, "%s%c%s\n", rust_exp(q[i]).c_str(), sep, rust_exp(intensity[i]).c_str());
  std::fclose(f);
}
}  // namespace detail

// write_curve_dat / write_curve_csv (src/io.rs:143-169).
inline void write_curve_dat(const std::string& path, const std::vector<double>& q,
                            const std::vector<double>& intensity) {
  detail::write_curve(path, q, intensity, "# q (1/A) intensity", ' ');
}
inline void write_curve_csv(const std::string& path, const std::vector<double>& q,
                            const std::vector<double>& intensity) {
  detail::write_curve(path, q, intensity, "q,intensity", ',');
}

// -------------------------------------------------------------------- fit
// src/fit.rs:11-135.

struct FitParams {
  double scale = 1.0;
  double background = 0.0;
  double c1 = 1.0;  // SolventParams::default().c1
  double c2 = 0.0;  // SolventParams::default().c2
  csx_fit_params to_c() const { return {scale, background, c1, c2}; }
};

struct FitResult {
  FitParams params;
  double chi2 = 0.0;
  std::vector<double> fitted_curve;
  size_t iterations = 0;
  bool converged = false;
  std::string termination;
};

struct ScaleMode {
  enum class Kind { Fit, Fixed } kind = Kind::Fit;
  double value = 0.0;
  static ScaleMode Fit() { return {}; }
  static ScaleMode Fixed(double v) { return {Kind::Fixed, v}; }
};

struct BackgroundMode {
  enum class Kind { Fit, FixedZero, Fixed } kind = Kind::Fit;
  double value = 0.0;
  static BackgroundMode Fit() { return {}; }
  static BackgroundMode FixedZero() { return {Kind::FixedZero, 0.0}; }
  static BackgroundMode Fixed(double v) { return {Kind::Fixed, v}; }
};

// Optimiser over (c1, c2): the reference's Nelder-Mead, or the gradient
// fitter (no reference counterpart).
enum class Optimizer { NelderMead = CSX_OPTIMIZER_NELDER_MEAD, Gradient = CSX_OPTIMIZER_GRADIENT };

struct FitOptions {
  ScatteringMethod method = ScatteringMethod::Auto;
  HydrogenMode hydrogen_mode = HydrogenMode::Implicit;
  std::pair<double, double> c1_bounds{0.95, 1.05};
  std::pair<double, double> c2_bounds{-2.0, 4.0};
  size_t max_iterations = 64;
  size_t q_sampling_stride = 20;
  SolventParams solvent;
  BackgroundMode background_mode;
  ScaleMode scale_mode;
  Optimizer optimizer = Optimizer::NelderMead;
  Precision precision = Precision::FP32Compensated;

  csx_fit_options to_c() const {
    csx_fit_options o;
    csx_fit_options_default(&o);
    o.method = static_cast<int32_t>(method);
    o.implicit_hydrogen = implicit(hydrogen_mode) ? 1 : 0;
    o.c1_bounds[0] = c1_bounds.first;
    o.c1_bounds[1] = c1_bounds.second;
    o.c2_bounds[0] = c2_bounds.first;
    o.c2_bounds[1] = c2_bounds.second;
    o.max_iterations = static_cast<int32_t>(std::min<size_t>(max_iterations, INT32_MAX));
    o.q_sampling_stride = static_cast<int32_t>(std::min<size_t>(q_sampling_stride, INT32_MAX));
    o.solvent = solvent.to_c();
    o.background_mode = background_mode.kind == BackgroundMode::Kind::Fit ? CSX_BACKGROUND_FIT
                        : background_mode.kind == BackgroundMode::Kind::FixedZero
                            ? CSX_BACKGROUND_FIXED_ZERO
                            : CSX_BACKGROUND_FIXED;
    o.background_value = background_mode.value;
    o.scale_mode = scale_mode.kind == ScaleMode::Kind::Fit ? CSX_SCALE_FIT : CSX_SCALE_FIXED;
    o.scale_value = scale_mode.value;
    o.optimizer = static_cast<int32_t>(optimizer);
    o.precision = static_cast<int32_t>(precision);
    return o;
  }
};

namespace detail {
inline FitResult result(const csx_fit_result& r, std::vector<double> fitted) {
  FitResult f;
  f.params = {r.params.scale, r.params.background, r.params.c1, r.params.c2};
  f.chi2 = r.chi2;
  f.fitted_curve = std::move(fitted);
  f.iterations = static_cast<size_t>(r.iterations);
  f.converged = r.converged != 0;
  f.termination = r.termination;
  return f;
}
inline void check_curve(const ExperimentalCurve& e) {
  if (e.q.size() != e.intensity.size() || e.q.size() != e.sigma.size())
    throw InvalidInput(
        "fit requires a non-empty structure and equal-length finite q, intensity, and "
        "positive-sigma arrays on a non-decreasing q grid");
}
}  // namespace detail

// fit_to_experimental_with_options (fit.rs:288-454).
inline FitResult fit_to_experimental_with_options(const Structure& s, const ExperimentalCurve& e,
                                                  const FitParams& initial,
                                                  const FitOptions& options,
                                                  Context& ctx = Context::default_context()) {
  detail::check_curve(e);
  csx_fit_result r;
  std::vector<double> fitted(e.len());
  const csx_fit_params p = initial.to_c();
  const csx_fit_options o = options.to_c();
  detail::check(csx_fit_to_experimental(s.handle(ctx), e.q.data(), e.intensity.data(),
                                        e.sigma.data(), e.len(), &p, &o, &r, fitted.data()));
  return detail::result(r, std::move(fitted));
}

// fit_to_experimental (fit.rs:275-286).
inline FitResult fit_to_experimental(const Structure& s, const ExperimentalCurve& e,
                                     const FitParams& initial_guess,
                                     Context& ctx = Context::default_context()) {
  return fit_to_experimental_with_options(s, e, initial_guess, FitOptions{}, ctx);
}

struct LinearFit {
  double scale, background, chi2;
};

// solve_scale_background (fit.rs:200-244).
inline LinearFit solve_scale_background(const std::vector<double>& theory,
                                        const ExperimentalCurve& e,
                                        ScaleMode scale_mode = ScaleMode::Fit(),
                                        BackgroundMode background_mode = BackgroundMode::Fit()) {
  if (theory.size() != e.len()) throw InvalidInput("theory and experimental curve differ in length");
  double out[3];
  FitOptions modes;
  modes.scale_mode = scale_mode;
  modes.background_mode = background_mode;
  const csx_fit_options o = modes.to_c();
  detail::check(csx_solve_scale_background(theory.data(), e.intensity.data(), e.sigma.data(),
                                           e.len(), o.scale_mode, o.scale_value,
                                           o.background_mode, o.background_value, out));
  return {out[0], out[1], out[2]};
}

// cubic_hermite_interpolate (fit.rs:93-123).
inline std::vector<double> cubic_hermite_interpolate(const std::vector<double>& x,
                                                     const std::vector<double>& y,
                                                     const std::vector<double>& targets) {
  if (x.size() != y.size() || x.empty()) throw InvalidInput("x and y must be equal and non-empty");
  std::vector<double> out(targets.size());
  detail::check(csx_cubic_hermite_interpolate(x.data(), y.data(), x.size(), targets.data(),
                                              targets.size(), out.data()));
  return out;
}

// ---------------------------------------------------------------- analysis
// src/analysis.rs:11-251.

enum class Quality { Fast, Balanced, Accurate };

inline FitOptions fit_options(Quality quality, FitOptions o) {
  switch (quality) {
    case Quality::Fast:
      o.q_sampling_stride = std::max<size_t>(o.q_sampling_stride, 20);
      o.max_iterations = std::min<size_t>(o.max_iterations, 32);
      o.solvent.hydration_grid_spacing_angstrom = 8.0;
      break;
    case Quality::Balanced:
      o.q_sampling_stride = 20;
      o.max_iterations = std::max<size_t>(o.max_iterations, 64);
      o.solvent.hydration_grid_spacing_angstrom = 4.0;
      break;
    case Quality::Accurate:
      o.q_sampling_stride = 1;
      o.max_iterations = std::max<size_t>(o.max_iterations, 128);
      o.solvent.hydration_grid_spacing_angstrom = 3.0;
      break;
  }
  return o;
}

inline ComputeOptions compute_options(Quality quality, ComputeOptions o) {
  if (o.solvent) {
    if (quality == Quality::Fast) o.solvent->hydration_grid_spacing_angstrom = 8.0;
    if (quality == Quality::Accurate) o.solvent->hydration_grid_spacing_angstrom = 3.0;
  }
  return o;
}

class ProfileCalculator {
 public:
  std::vector<double> q;
  ComputeOptions options;

  ProfileCalculator(std::vector<double> q_values, ComputeOptions o)
      : q(std::move(q_values)), options(std::move(o)) {
    for (size_t i = 0; i < q.size(); ++i)
      if (!std::isfinite(q[i]) || q[i] < 0.0 || (i > 0 && q[i] < q[i - 1]))
        throw InvalidInput("q grid must be finite, non-negative, and non-decreasing");
  }
  ScatteringCurve calculate(const Structure& s, Context& ctx = Context::default_context()) const {
    return compute_curve(s, q, options, ctx);
  }
};

class SaxsFitter {
 public:
  FitOptions options;
  explicit SaxsFitter(FitOptions o) : options(std::move(o)) {}
  FitResult fit(const Structure& s, const ExperimentalCurve& e,
                Context& ctx = Context::default_context()) const {
    return fit_to_experimental_with_options(s, e, FitParams{}, options, ctx);
  }
};

struct ScoreOptions {
  FitOptions fit;
  Quality quality = Quality::Fast;
};

struct ScoreResult {
  double chi2 = 0.0, reduced_chi2 = 0.0, scale = 0.0, background = 0.0;
  double q_min = 0.0, q_max = 0.0;
  size_t points = 0;
  bool has_errors = false;
  std::string q_unit;
  FitResult fit;
};

class SaxsScorer {
 public:
  ExperimentalCurve experimental;
  ScoreOptions options;

  SaxsScorer(ExperimentalCurve e, ScoreOptions o) : experimental(std::move(e)), options(o) {
    if (experimental.is_empty()) throw InvalidInput("experimental curve is empty");
  }
  ScoreResult score(const Structure& s, Context& ctx = Context::default_context()) const {
    const FitOptions fo = fit_options(options.quality, options.fit);
    ScoreResult r;
    r.fit = fit_to_experimental_with_options(s, experimental, FitParams{}, fo, ctx);
    const double dof = static_cast<double>(std::max<size_t>(experimental.len(), 2) - 1);
    r.chi2 = r.fit.chi2 * dof;
    r.reduced_chi2 = r.fit.chi2;
    r.scale = r.fit.params.scale;
    r.background = r.fit.params.background;
    r.q_min = experimental.q[0];
    r.q_max = experimental.q.back();
    r.points = experimental.len();
    r.has_errors = experimental.has_errors;
    r.q_unit = experimental.q_unit();
    return r;
  }
};

struct ComparisonResult {
  double chi2, reduced_chi2, pearson_correlation, r_factor, weighted_rmse;
  std::vector<double> residuals;
};

// compare_curves (analysis.rs:190-251): O(n) host arithmetic.
inline ComparisonResult compare_curves(const ExperimentalCurve& e,
                                       const std::vector<double>& calculated) {
  if (e.len() != calculated.size() || e.is_empty())
    throw InvalidInput("experimental and calculated curves must have equal non-zero lengths");
  const size_t n = e.len();
  ComparisonResult r{};
  r.residuals.resize(n);
  double chi = 0.0, mo = 0.0, mm = 0.0;
  for (size_t i = 0; i < n; ++i) {
    r.residuals[i] = calculated[i] - e.intensity[i];
    chi += std::pow(r.residuals[i] / e.sigma[i], 2);
    mo += e.intensity[i];
    mm += calculated[i];
  }
  mo /= static_cast<double>(n);
  mm /= static_cast<double>(n);
  double cov = 0.0, vo = 0.0, vm = 0.0, den = 0.0, num = 0.0;
  for (size_t i = 0; i < n; ++i) {
    cov += (e.intensity[i] - mo) * (calculated[i] - mm);
    vo += std::pow(e.intensity[i] - mo, 2);
    vm += std::pow(calculated[i] - mm, 2);
    den += std::fabs(e.intensity[i]);
    num += std::fabs(r.residuals[i]);
  }
  r.chi2 = chi;
  r.reduced_chi2 = chi / static_cast<double>(std::max<size_t>(n, 2) - 1);
  r.pearson_correlation = vo > 0.0 && vm > 0.0 ? cov / std::sqrt(vo * vm) : 0.0;
  r.r_factor = den > 0.0 ? num / den : 0.0;
  r.weighted_rmse = std::sqrt(chi / static_cast<double>(n));
  return r;
}

// ---------------------------------------------------------------- ensemble
// src/ensemble.rs:13-475.

enum class WeightSolver { ProjectedGradient = CSX_WEIGHTS_PROJECTED_GRADIENT,
                          NNLS = CSX_WEIGHTS_NNLS };

struct EnsembleOptions {
  bool fit_weights = true;
  bool uniform = false;
  double regularization = 0.0;
  std::optional<size_t> max_conformers;
  double min_weight = 0.0;
  size_t bootstrap = 0;
  uint64_t seed = 0xC0FFEE;
  FitOptions fit;
  WeightSolver weight_solver = WeightSolver::ProjectedGradient;  // no reference counterpart
};

struct WeightInterval {
  double median, lower, upper;
};

struct EnsembleResult {
  std::vector<std::string> conformers;
  std::vector<double> weights;
  std::vector<WeightInterval> weight_intervals;
  FitParams params;
  double chi2 = 0.0, reduced_chi2 = 0.0;
  std::vector<double> fitted_curve;
  std::vector<double> bootstrap_reduced_chi2;
  size_t iterations = 250;
  bool converged = true;
};

// Every conformer's six-term basis on the experimental grid is computed once
// (one launch for all conformers on the FP32 Debye path); the 25 (c1, c2)
// candidates of fit_core, the pruning passes and every bootstrap replicate
// re-contract those bases.  Each candidate profile of the reference equals
// combine_solvent_basis of the conformer's basis.
class EnsembleFitter {
 public:
  std::vector<Structure> conformers;
  std::vector<std::string> names;
  ExperimentalCurve experimental;
  EnsembleOptions options;

  EnsembleFitter(std::vector<Structure> c, std::vector<std::string> n, ExperimentalCurve e,
                 EnsembleOptions o)
      : conformers(std::move(c)), names(std::move(n)), experimental(std::move(e)),
        options(std::move(o)) {
    if (conformers.empty() ||
        std::any_of(conformers.begin(), conformers.end(),
                    [](const Structure& s) { return s.atoms.empty(); }))
      throw InvalidInput("ensemble must contain non-empty conformers");
    if (names.size() != conformers.size())
      throw InvalidInput("ensemble names and conformers differ in length");
    if (experimental.is_empty() || experimental.q.size() != experimental.intensity.size() ||
        experimental.q.size() != experimental.sigma.size())
      throw InvalidInput("ensemble requires a validated experimental curve");
    if (!std::isfinite(options.regularization) || options.regularization < 0.0 ||
        !std::isfinite(options.min_weight) || options.min_weight < 0.0 ||
        options.min_weight > 1.0 || (options.max_conformers && *options.max_conformers == 0))
      throw InvalidInput("ensemble regularization and pruning options are invalid");
  }

  // Per-conformer bases, m x 6 x n (computed once).
  const std::vector<double>& bases(Context& ctx = Context::default_context()) {
    if (!bases_.empty()) return bases_;
    const size_t n = experimental.len(), m = conformers.size();
    FitOptions o = options.fit;
    o.q_sampling_stride = 1;
    bases_.assign(m * 6 * n, 0.0);
    if (o.method == ScatteringMethod::Debye && o.precision == Precision::FP32Compensated) {
      std::vector<const Structure*> ss;
      for (const Structure& s : conformers) ss.push_back(&s);
      const std::vector<Basis> b =
          batch_debye_solvent_basis(ss, experimental.q, o.solvent, implicit(o.hydrogen_mode), ctx);
      for (size_t k = 0; k < m; ++k)
        for (int t = 0; t < 6; ++t)
          std::copy(b[k][t].begin(), b[k][t].end(), bases_.begin() + (k * 6 + t) * n);
    } else {
      const csx_fit_options c = o.to_c();
      for (size_t k = 0; k < m; ++k)
        detail::check(csx_fit_basis(conformers[k].handle(ctx), experimental.q.data(), n, &c,
                                    bases_.data() + k * 6 * n));
    }
    return bases_;
  }

  // EnsembleFitter::fit (ensemble.rs:127-232).
  EnsembleResult fit(Context& ctx = Context::default_context()) {
    std::vector<size_t> selected(conformers.size());
    for (size_t i = 0; i < selected.size(); ++i) selected[i] = i;
    Core core = fit_core(selected, experimental, ctx);
    for (bool changed = true; changed;) {
      changed = false;
      std::vector<size_t> keep;
      for (size_t p = 0; p < selected.size(); ++p)
        if (core.weights[p] + 1e-12 >= options.min_weight) keep.push_back(selected[p]);
      if (keep.empty())
        keep.push_back(selected[std::max_element(core.weights.begin(), core.weights.end()) -
                                core.weights.begin()]);
      if (options.max_conformers && keep.size() > *options.max_conformers) {
        std::vector<std::pair<double, size_t>> ranked;
        for (size_t p = 0; p < selected.size(); ++p) ranked.push_back({core.weights[p], selected[p]});
        std::stable_sort(ranked.begin(), ranked.end(),
                         [](const auto& a, const auto& b) { return a.first > b.first; });
        keep.clear();
        for (size_t k = 0; k < *options.max_conformers; ++k) keep.push_back(ranked[k].second);
      }
      std::sort(keep.begin(), keep.end());
      if (keep != selected) {
        selected = keep;
        core = fit_core(selected, experimental, ctx);
        changed = true;
      }
    }
    EnsembleResult r;
    for (size_t i : selected) r.conformers.push_back(names[i]);
    for (double w : core.weights) r.weight_intervals.push_back({w, w, w});
    if (options.bootstrap > 0) {
      const size_t n = experimental.len();
      std::vector<double> resid(n);
      for (size_t i = 0; i < n; ++i)
        resid[i] = (experimental.intensity[i] - core.fitted[i]) / experimental.sigma[i];
      std::vector<uint64_t> pick(options.bootstrap * n);
      detail::check(csx_bootstrap_indices(options.seed, n, options.bootstrap, pick.data()));
      std::vector<std::vector<double>> samples(selected.size());
      for (size_t b = 0; b < options.bootstrap; ++b) {
        ExperimentalCurve synth = experimental;
        for (size_t i = 0; i < n; ++i)
          synth.intensity[i] = core.fitted[i] + resid[pick[b * n + i]] * synth.sigma[i];
        const Core rep = fit_core(selected, synth, ctx);
        for (size_t p = 0; p < rep.weights.size(); ++p) samples[p].push_back(rep.weights[p]);
        r.bootstrap_reduced_chi2.push_back(rep.chi2);
      }
      for (size_t p = 0; p < samples.size(); ++p) {
        std::sort(samples[p].begin(), samples[p].end());
        r.weight_intervals[p] = {quantile(samples[p], 0.5), quantile(samples[p], 0.025),
                                 quantile(samples[p], 0.975)};
      }
    }
    r.weights = core.weights;
    r.params = core.params;
    r.reduced_chi2 = core.chi2;
    r.chi2 = core.chi2 * static_cast<double>(std::max<size_t>(experimental.len(), 2) - 1);
    r.fitted_curve = core.fitted;
    return r;
  }

 private:
  struct Core {
    std::vector<double> weights;
    FitParams params;
    double chi2 = 0.0;
    std::vector<double> fitted;
  };
  std::vector<double> bases_;

  static double quantile(const std::vector<double>& v, double p) {
    if (v.empty()) return std::nan("");
    const double pos = std::min(std::max(p, 0.0), 1.0) * static_cast<double>(v.size() - 1);
    const size_t lo = static_cast<size_t>(std::floor(pos)), hi = static_cast<size_t>(std::ceil(pos));
    return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
  }

  // fit_core (ensemble.rs:234-299) on the cached bases.
  Core fit_core(const std::vector<size_t>& selected, const ExperimentalCurve& data,
                Context& ctx) {
    const std::vector<double>& all = bases(ctx);
    const size_t n = data.len(), m = selected.size();
    std::vector<double> b(m * 6 * n);
    for (size_t k = 0; k < m; ++k)
      std::copy(all.begin() + selected[k] * 6 * n, all.begin() + (selected[k] + 1) * 6 * n,
                b.begin() + k * 6 * n);
    Core c;
    c.weights.resize(m);
    c.fitted.resize(n);
    const csx_fit_options o = options.fit.to_c();
    csx_fit_params p;
    detail::check(csx_ensemble_fit_core_device(
        ctx.get(), b.data(), m, data.q.data(), data.intensity.data(), data.sigma.data(), n, &o,
        (options.uniform || !options.fit_weights) ? 1 : 0, options.regularization,
        static_cast<int>(options.weight_solver), c.weights.data(), &p, &c.chi2,
        c.fitted.data()));
    c.params = {p.scale, p.background, p.c1, p.c2};
    return c;
  }
};

}  // namespace crabsaxs

#endif  // CRABSAXS_HPP

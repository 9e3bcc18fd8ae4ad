// Host-side physics of the profile path (see physics.hpp).
#include "physics.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <unordered_map>

#include "../../include/crabsaxs_tables.h"

namespace csx {

// ---------------------------------------------------------------- typing --

static std::string trim_upper(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  std::string out = s.substr(b, e - b);
  for (auto& c : out) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return out;
}

int element_from_symbol(const std::string& symbol) {
  static const std::pair<const char*, int> kSymbols[] = {
      {"H", CSX_EL_H},   {"D", CSX_EL_H},   {"C", CSX_EL_C},   {"N", CSX_EL_N},
      {"O", CSX_EL_O},   {"S", CSX_EL_S},   {"P", CSX_EL_P},   {"NA", CSX_EL_NA},
      {"MG", CSX_EL_MG}, {"CL", CSX_EL_CL}, {"CA", CSX_EL_CA}, {"FE", CSX_EL_FE},
      {"ZN", CSX_EL_ZN}, {"K", CSX_EL_K},   {"I", CSX_EL_I}};
  const std::string key = trim_upper(symbol);
  for (const auto& kv : kSymbols)
    if (key == kv.first) return kv.second;
  return CSX_EL_OTHER;
}

namespace {

// A '|'-separated alternative list matched against one token.
bool one_of(const char* alternatives, const std::string& token) {
  const char* p = alternatives;
  while (*p) {
    const char* bar = std::strchr(p, '|');
    const size_t len = bar ? static_cast<size_t>(bar - p) : std::strlen(p);
    if (token.size() == len && token.compare(0, len, p, len) == 0) return true;
    if (!bar) break;
    p = bar + 1;
  }
  return false;
}

// (residues, atom names) -> group, first match wins; restates the match arms
// of formfactor.rs:378-416 as data.
struct Rule {
  const char* residues;
  const char* names;
  int group;
};
const Rule kAminoRules[] = {
    {"ALA", "CB", CSX_UA_CSP3H3},
    {"ARG", "CB|CG|CD", CSX_UA_CSP3H2},
    {"ARG", "NE", CSX_UA_GUANIDINENH},
    {"ARG", "NH1|NH2", CSX_UA_GUANIDINENH2},
    {"ASN", "CB", CSX_UA_CSP3H2},
    {"ASP|ASH", "CB", CSX_UA_CSP3H2},
    {"CYS|CYX", "CB", CSX_UA_CSP3H2},
    {"ASN", "ND2", CSX_UA_NH2},
    {"ASP|GLU", "OD1|OD2|OE1|OE2", CSX_UA_ORESONANCE},
    {"ASH", "OD2", CSX_UA_OACIDH},
    {"GLH", "OE2", CSX_UA_OACIDH},
    {"CYS", "SG", CSX_UA_SH},
    {"GLN", "CB|CG", CSX_UA_CSP3H2},
    {"GLU|GLH", "CB|CG", CSX_UA_CSP3H2},
    {"GLN", "NE2", CSX_UA_NH2},
    {"HIS|HID|HIE|HIP", "CB", CSX_UA_CSP3H2},
    {"HIS|HID", "ND1", CSX_UA_NH},
    {"HIE", "NE2", CSX_UA_NH},
    {"HIP", "ND1|NE2", CSX_UA_NHPOSITIVE},
    {"HIS|HID|HIE|HIP", "CD2|CE1", CSX_UA_CAROMATICH},
    {"ILE", "CB", CSX_UA_CSP3H},
    {"LEU", "CG", CSX_UA_CSP3H},
    {"THR", "CB", CSX_UA_CSP3H},
    {"VAL", "CB", CSX_UA_CSP3H},
    {"ILE", "CG1", CSX_UA_CSP3H2},
    {"LEU", "CB", CSX_UA_CSP3H2},
    {"ILE", "CG2|CD1", CSX_UA_CSP3H3},
    {"LEU", "CD1|CD2", CSX_UA_CSP3H3},
    {"LYS", "CB|CG|CD|CE", CSX_UA_CSP3H2},
    {"LYS", "NZ", CSX_UA_NH3POSITIVE},
    {"MET|MSE", "CB|CG", CSX_UA_CSP3H2},
    {"MET|MSE", "CE", CSX_UA_CSP3H3},
    {"PHE|TYR|TRP", "CB", CSX_UA_CSP3H2},
    {"PHE", "CD1|CD2|CE1|CE2|CZ", CSX_UA_CAROMATICH},
    {"TYR", "CD1|CD2|CE1|CE2", CSX_UA_CAROMATICH},
    {"TRP", "CD1|CE3|CZ2|CZ3|CH2", CSX_UA_CAROMATICH},
    {"PRO", "CB|CG|CD", CSX_UA_CSP3H2},
    {"SER", "CB", CSX_UA_CSP3H2},
    {"SER", "OG", CSX_UA_OALCOHOLH},
    {"THR", "OG1", CSX_UA_OALCOHOLH},
    {"TYR", "OH", CSX_UA_OALCOHOLH},
    {"THR", "CG2", CSX_UA_CSP3H3},
    {"VAL", "CG1|CG2", CSX_UA_CSP3H3},
    {"TRP", "NE1", CSX_UA_NH},
    {"PCA", "CB|CG", CSX_UA_CSP3H2},
};

const char* kAminoAcids =
    "ALA|ARG|ASN|ASP|ASH|CYS|CYX|GLN|GLU|GLH|GLY|HIS|HID|HIE|HIP|ILE|LEU|"
    "LYS|MET|MSE|PHE|PRO|SER|THR|TRP|TYR|VAL|PCA";

// nucleotide_base, formfactor.rs:314-323.
char nucleotide_base(const std::string& residue) {
  if (one_of("A|DA|RA|ADE", residue)) return 'A';
  if (one_of("C|DC|RC|CYT", residue)) return 'C';
  if (one_of("G|DG|RG|GUA", residue)) return 'G';
  if (one_of("U|DU|RU|URA", residue)) return 'U';
  if (one_of("T|DT|RT|THY", residue)) return 'T';
  return 0;
}

// nucleotide_group, formfactor.rs:325-355.
int nucleotide_group(const std::string& residue, char base, std::string name) {
  std::replace(name.begin(), name.end(), '*', '\'');
  const bool deoxy = (!residue.empty() && residue[0] == 'D') || base == 'T';
  if (one_of("C1'|C3'|C4'", name)) return CSX_UA_CSP3H;
  if (name == "C2'") return deoxy ? CSX_UA_CSP3H2 : CSX_UA_CSP3H;
  if (name == "C5'") return CSX_UA_CSP3H2;
  if (name == "O2'" && !deoxy) return CSX_UA_OALCOHOLH;
  if (one_of("OP1|OP2|O1P|O2P", name)) return CSX_UA_ORESONANCE;
  if (name == "C8" && (base == 'A' || base == 'G')) return CSX_UA_CAROMATICH;
  struct BaseRule { char base; const char* names; int group; };
  static const BaseRule kBaseRules[] = {
      {'A', "C2", CSX_UA_CAROMATICH}, {'A', "N6", CSX_UA_NH2},
      {'G', "N1", CSX_UA_NH},         {'G', "N2", CSX_UA_NH2},
      {'C', "N4", CSX_UA_NH2},        {'C', "C5|C6", CSX_UA_CAROMATICH},
      {'U', "C5|C6", CSX_UA_CAROMATICH}, {'T', "C6", CSX_UA_CAROMATICH},
      {'U', "N3", CSX_UA_NH},         {'T', "N3", CSX_UA_NH},
      {'T', "C7|C5M", CSX_UA_CSP3H3}};
  for (const auto& r : kBaseRules)
    if (r.base == base && one_of(r.names, name)) return r.group;
  return -1;
}

}  // namespace

int united_atom_group(const std::string& residue_in, const std::string& name_in) {
  const std::string residue = trim_upper(residue_in);
  const std::string name = trim_upper(name_in);
  if (const char base = nucleotide_base(residue)) return nucleotide_group(residue, base, name);
  if (!one_of(kAminoAcids, residue)) return -1;
  if (name == "N") return one_of("PRO|PCA", residue) ? -1 : CSX_UA_NH;
  if (name == "CA") return residue == "GLY" ? CSX_UA_CSP3H2 : CSX_UA_CSP3H;
  for (const auto& r : kAminoRules)
    if (one_of(r.residues, residue) && one_of(r.names, name)) return r.group;
  return -1;
}

// ---------------------------------------------------------- form factors --

static double gaussian_sum(const double* a, const double* b, int terms, double c, double q) {
  const double s = q / (4.0 * kPi);
  const double s2 = s * s;
  double sum = 0.0;
  for (int k = 0; k < terms; ++k) sum += a[k] * std::exp(-b[k] * s2);
  return sum + c;
}

static double element_form_factor(int el, double q) {
  return gaussian_sum(CSX_CM_A[el], CSX_CM_B[el], 4, CSX_CM_C[el], q);
}

double type_form_factor(const TypeKey& t, bool implicit_h, double q) {
  if (implicit_h && t.element == CSX_EL_H) return 0.0;
  if (implicit_h && t.group >= 0)
    return gaussian_sum(CSX_UA_A[t.group], CSX_UA_B[t.group], 5, CSX_UA_C[t.group], q);
  return element_form_factor(t.element, q);
}

double type_excluded_volume(const TypeKey& t, bool implicit_h) {
  if (implicit_h && t.element == CSX_EL_H) return 0.0;
  const double nh = (implicit_h && t.group >= 0) ? CSX_UA_HYDROGENS[t.group] : 0.0;
  return CSX_EXCLUDED_VOLUME[t.element] + (implicit_h ? nh * CSX_EXCLUDED_VOLUME[CSX_EL_H] : 0.0);
}

static double sinc(double x) { return std::fabs(x) < 1e-8 ? 1.0 : std::sin(x) / x; }

double water_form_factor(double q) {
  const double oxygen = element_form_factor(CSX_EL_O, q);
  const double hydrogen = element_form_factor(CSX_EL_H, q);
  return oxygen + 2.0 * hydrogen * sinc(q * CSX_WATER_OH_DISTANCE);
}

double fraser_gaussian(double q, double volume) {
  return volume * std::exp(-std::pow(volume, 2.0 / 3.0) * q * q / (4.0 * kPi));
}

double excluded_volume_scale(double q, double c1) {
  const double r2 = CSX_MEAN_ATOMIC_RADIUS * CSX_MEAN_ATOMIC_RADIUS;
  return c1 * c1 * c1 * std::exp(-(c1 * c1 - 1.0) * r2 * q * q / 4.0);
}

Amplitudes debye_amplitudes(const TypeKey& t, bool implicit_h, double q, double thickness,
                            double rho) {
  const double volume = type_excluded_volume(t, implicit_h);
  const double radius = std::cbrt(3.0 * volume / (4.0 * kPi));
  const double rs = radius + thickness / 2.0;
  const double shell_volume = 4.0 * kPi * (rs * rs) * thickness;
  return {type_form_factor(t, implicit_h, q), rho * fraser_gaussian(q, volume),
          rho * 0.1 * fraser_gaussian(q, shell_volume)};
}

Amplitudes multipole_amplitudes(const TypeKey& t, bool implicit_h, double q, double thickness,
                                double rho) {
  const double volume = type_excluded_volume(t, implicit_h);
  const double radius = std::cbrt(3.0 * volume / (4.0 * kPi));
  const double rs = radius + thickness / 2.0;
  const double shell_volume = 4.0 * kPi * (rs * rs) * thickness;
  const double e = rho * volume * std::exp(-std::pow(volume, 2.0 / 3.0) * q * q / (4.0 * kPi));
  const double h = rho * 0.1 * shell_volume * std::exp(-q * q * rs * rs / 4.0);
  return {type_form_factor(t, implicit_h, q), e, h};
}

// ---------------------------------------------------------------- grids --

bool grid_box(const double* xyz, size_t n, double margin, double spacing, GridBox& box) {
  for (int a = 0; a < 3; ++a) box.lo[a] = box.hi[a] = xyz[a];
  for (size_t i = 1; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = std::min(box.lo[a], xyz[3 * i + a]);
      box.hi[a] = std::max(box.hi[a], xyz[3 * i + a]);
    }
  box.count = 1;
  for (int a = 0; a < 3; ++a) {
    box.lo[a] -= margin;
    box.hi[a] += margin;
    const double cells = std::ceil((box.hi[a] - box.lo[a]) / spacing - 1e-10);
    if (!(cells < 1e15)) return false;
    box.dim[a] = static_cast<size_t>(cells) + 1;
    if (box.count > std::numeric_limits<size_t>::max() / box.dim[a]) return false;
    box.count *= box.dim[a];
  }
  return true;
}

double adaptive_site_volume(double spacing) {
  if (spacing > 4.0) {
    const double h = spacing / 2.0;
    return 2.0 * (h * h * h);
  }
  return spacing * spacing * spacing;
}

double adaptive_shell_width(double rg) {
  if (rg <= 15.0) return 3.0;
  if (rg >= 20.0) return 5.0;
  return 3.0 + 0.4 * (rg - 15.0);
}

namespace {

// Visit every voxel whose centre lies within `radius` of `p`, with the bound
// rounding of solvent.rs:187-199.
template <class F>
void for_voxels_within(const GridBox& b, double spacing, const double* p, double radius, F&& f) {
  size_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    const double lower = (p[a] - radius - b.lo[a]) / spacing;
    const double upper = (p[a] + radius - b.lo[a]) / spacing;
    const double l = std::max(std::floor(lower - 1e-10), 0.0);
    const double u = std::ceil(upper + 1e-10);
    lo[a] = static_cast<size_t>(l);
    hi[a] = u <= 0.0 ? 0 : std::min(static_cast<size_t>(u), b.dim[a] - 1);
  }
  const double r2 = radius * radius;
  for (size_t z = lo[2]; z <= hi[2]; ++z)
    for (size_t y = lo[1]; y <= hi[1]; ++y)
      for (size_t x = lo[0]; x <= hi[0]; ++x) {
        const double dx = (b.lo[0] + static_cast<double>(x) * spacing) - p[0];
        const double dy = (b.lo[1] + static_cast<double>(y) * spacing) - p[1];
        const double dz = (b.lo[2] + static_cast<double>(z) * spacing) - p[2];
        if (dx * dx + dy * dy + dz * dz <= r2 + 1e-10) f((z * b.dim[1] + y) * b.dim[0] + x);
      }
}

}  // namespace

// Synthetic shell sites for workloads: grid points whose nearest atom lies in
// (inner, outer], from the two-mask scheme of solvent.rs:182-218 on the host.
Grid shell_sites(const double* xyz, size_t n, double spacing, double inner, double outer) {
  if (n == 0) return Grid{};
  Grid g;
  if (!std::isfinite(spacing) || spacing <= 0.0 || !(inner >= 0.0) || !(outer > inner)) {
    g.error = "shell-site spacing and radii must be finite with outer > inner >= 0";
    return g;
  }
  g.voxel_volume = spacing * spacing * spacing;
  GridBox b;
  if (!grid_box(xyz, n, outer + spacing, spacing, b) || b.count > (size_t(1) << 31)) {
    g.error = "shell-site grid is too large";
    return g;
  }
  std::vector<uint8_t> shell(b.count, 0), molecular(b.count, 0);
  for (size_t i = 0; i < n; ++i)
    for_voxels_within(b, spacing, xyz + 3 * i, outer, [&](size_t c) { shell[c] = 1; });
  for (size_t i = 0; i < n; ++i)
    for_voxels_within(b, spacing, xyz + 3 * i, inner, [&](size_t c) { molecular[c] = 1; });
  for (size_t z = 0; z < b.dim[2]; ++z)
    for (size_t y = 0; y < b.dim[1]; ++y)
      for (size_t x = 0; x < b.dim[0]; ++x) {
        const size_t c = (z * b.dim[1] + y) * b.dim[0] + x;
        if (shell[c] && !molecular[c]) {
          g.sites.push_back(b.lo[0] + static_cast<double>(x) * spacing);
          g.sites.push_back(b.lo[1] + static_cast<double>(y) * spacing);
          g.sites.push_back(b.lo[2] + static_cast<double>(z) * spacing);
        }
      }
  return g;
}

// ------------------------------------------------------------ generator --

namespace {

struct Rng {
  std::mt19937_64 engine;
  explicit Rng(uint64_t seed) : engine(seed) {}
  double uniform() { return static_cast<double>(engine() >> 11) * (1.0 / 9007199254740992.0); }
  std::array<double, 3> direction() {
    const double z = 2.0 * uniform() - 1.0;
    const double phi = 2.0 * kPi * uniform();
    const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    return {s * std::cos(phi), s * std::sin(phi), z};
  }
};

using Vec3 = std::array<double, 3>;

// Self-avoiding C-alpha walk of `count` residues (3.8 A steps, `exclusion`
// A minimum C-alpha separation) confined to an ellipsoid around `centre`.
void confined_walk(Rng& rng, size_t count, const Vec3& centre, double radius, double elongation,
                   std::vector<Vec3>& out) {
  const double step = 3.8;
  double exclusion = 3.0;
  const double cell = 3.0;
  std::unordered_map<int64_t, std::vector<size_t>> hash;
  auto key = [&](const Vec3& p) {
    const int64_t ix = static_cast<int64_t>(std::floor(p[0] / cell)) + (1 << 20);
    const int64_t iy = static_cast<int64_t>(std::floor(p[1] / cell)) + (1 << 20);
    const int64_t iz = static_cast<int64_t>(std::floor(p[2] / cell)) + (1 << 20);
    return (ix << 42) | (iy << 21) | iz;
  };
  std::vector<Vec3> walk;
  walk.reserve(count);
  auto inside = [&](const Vec3& p) {
    const double x = (p[0] - centre[0]) / radius, y = (p[1] - centre[1]) / radius;
    const double z = (p[2] - centre[2]) / (radius * elongation);
    return x * x + y * y + z * z <= 1.0;
  };
  auto clear = [&](const Vec3& p) {
    const int64_t ix = static_cast<int64_t>(std::floor(p[0] / cell));
    const int64_t iy = static_cast<int64_t>(std::floor(p[1] / cell));
    const int64_t iz = static_cast<int64_t>(std::floor(p[2] / cell));
    for (int64_t dz = -1; dz <= 1; ++dz)
      for (int64_t dy = -1; dy <= 1; ++dy)
        for (int64_t dx = -1; dx <= 1; ++dx) {
          const int64_t k = ((ix + dx + (1 << 20)) << 42) | ((iy + dy + (1 << 20)) << 21) |
                            (iz + dz + (1 << 20));
          auto it = hash.find(k);
          if (it == hash.end()) continue;
          for (size_t idx : it->second) {
            const Vec3& o = walk[idx];
            const double ex = o[0] - p[0], ey = o[1] - p[1], ez = o[2] - p[2];
            if (ex * ex + ey * ey + ez * ez < exclusion * exclusion) return false;
          }
        }
    return true;
  };
  auto push = [&](const Vec3& p) {
    hash[key(p)].push_back(walk.size());
    walk.push_back(p);
  };
  push(centre);
  size_t backtracks = 0;
  while (walk.size() < count) {
    bool placed = false;
    for (int attempt = 0; attempt < 64 && !placed; ++attempt) {
      const Vec3 d = rng.direction();
      const Vec3& last = walk.back();
      const Vec3 p = {last[0] + step * d[0], last[1] + step * d[1], last[2] + step * d[2]};
      if (inside(p) && clear(p)) {
        push(p);
        placed = true;
      }
    }
    if (placed) continue;
    // Dead end: step back one residue (never the first), and relax the
    // exclusion slightly if the walk keeps getting trapped.
    if (walk.size() > 1) {
      hash[key(walk.back())].pop_back();
      walk.pop_back();
    }
    if (++backtracks % 4096 == 0) exclusion *= 0.95;
  }
  out.insert(out.end(), walk.begin(), walk.end());
}

void place_atoms(Rng& rng, const std::vector<Vec3>& ca, size_t n, double* xyz, uint8_t* element) {
  size_t atom = 0;
  for (size_t r = 0; r < ca.size() && atom < n; ++r)
    for (int k = 0; k < 8 && atom < n; ++k, ++atom) {
      const double u = rng.uniform();
      element[atom] = u < 0.63 ? CSX_EL_C : u < 0.80 ? CSX_EL_N : u < 0.99 ? CSX_EL_O : CSX_EL_S;
      Vec3 p = ca[r];
      if (k > 0) {
        const double d = 1.2 + 0.6 * rng.uniform();
        const Vec3 dir = rng.direction();
        for (int a = 0; a < 3; ++a) p[a] += d * dir[a];
      }
      for (int a = 0; a < 3; ++a) xyz[3 * atom + a] = p[a];
    }
  double c[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) c[a] += xyz[3 * i + a];
  for (int a = 0; a < 3; ++a) c[a] /= static_cast<double>(n == 0 ? 1 : n);
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] -= c[a];
}

}  // namespace

void generate_globule(size_t n, double radius, double elongation, uint64_t seed, double* xyz,
                      uint8_t* element) {
  Rng rng(seed);
  std::vector<Vec3> ca;
  confined_walk(rng, (n + 7) / 8, {0.0, 0.0, 0.0}, radius, elongation, ca);
  place_atoms(rng, ca, n, xyz, element);
}

void generate_domain_chain(size_t n, int domains, uint64_t seed, double* xyz, uint8_t* element) {
  Rng rng(seed);
  const size_t residues = (n + 7) / 8;
  const size_t linker = 6;
  const size_t k = static_cast<size_t>(std::max(1, domains));
  const size_t body = residues > linker * (k - 1) + 20 * k ? residues - linker * (k - 1) : residues;
  std::vector<double> share(k);
  double total = 0.0;
  for (auto& s : share) total += (s = 0.5 + rng.uniform());
  std::vector<size_t> size(k);
  size_t used = 0;
  for (size_t d = 0; d < k; ++d) {
    size[d] = d + 1 == k ? body - used
                         : std::max<size_t>(1, static_cast<size_t>(body * share[d] / total));
    used += size[d];
  }
  std::vector<double> rad(k);
  for (size_t d = 0; d < k; ++d) rad[d] = 35.0 * std::cbrt(static_cast<double>(size[d]) / 2500.0);
  std::vector<Vec3> centre(k);
  centre[0] = {0.0, 0.0, 0.0};
  for (size_t d = 1; d < k; ++d) {
    Vec3 best{};
    for (int attempt = 0; attempt < 100; ++attempt) {
      const Vec3 dir = rng.direction();
      const double sep = rad[d - 1] + rad[d] + 8.0;
      const Vec3 p = {centre[d - 1][0] + sep * dir[0], centre[d - 1][1] + sep * dir[1],
                      centre[d - 1][2] + sep * dir[2]};
      best = p;
      bool ok = true;
      for (size_t e = 0; e + 1 < d; ++e) {
        const double dx = p[0] - centre[e][0], dy = p[1] - centre[e][1], dz = p[2] - centre[e][2];
        if (std::sqrt(dx * dx + dy * dy + dz * dz) < rad[d] + rad[e] + 4.0) ok = false;
      }
      if (ok) break;
    }
    centre[d] = best;
  }
  std::vector<Vec3> ca;
  for (size_t d = 0; d < k; ++d) {
    std::vector<Vec3> dom;
    confined_walk(rng, size[d], centre[d], rad[d], 1.0, dom);
    if (d > 0 && body != residues) {
      // Flexible linker: jittered C-alpha path from the previous domain's
      // last residue to this domain's first.
      const Vec3 a = ca.back(), b = dom.front();
      for (size_t l = 1; l <= linker; ++l) {
        const double t = static_cast<double>(l) / static_cast<double>(linker + 1);
        const Vec3 j = rng.direction();
        ca.push_back({a[0] + t * (b[0] - a[0]) + j[0], a[1] + t * (b[1] - a[1]) + j[1],
                      a[2] + t * (b[2] - a[2]) + j[2]});
      }
    }
    ca.insert(ca.end(), dom.begin(), dom.end());
  }
  place_atoms(rng, ca, n, xyz, element);
}

}  // namespace csx

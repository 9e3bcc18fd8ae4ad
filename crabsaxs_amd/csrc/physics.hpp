// Host-side physics of the profile path: atom typing, form factors, solvent
// amplitudes, hydration-grid builders and the synthetic workload generator.
// All FP64 and evaluated in the reference's operation order (this file is
// compiled with -ffp-contract=off so no multiply-add is fused, as in Rust).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace csx {

constexpr double kPi = 3.14159265358979323846;

// One atom type of the pair tables: every per-atom quantity of the profile
// path except occupancy is a function of (element, united-atom group).
struct TypeKey {
  int element;  // CSX_EL_*
  int group;    // CSX_UA_* or -1
  bool operator<(const TypeKey& o) const {
    return element != o.element ? element < o.element : group < o.group;
  }
  bool operator==(const TypeKey& o) const {
    return element == o.element && group == o.group;
  }
};

// Form factor of one type (no occupancy), formfactor.rs:26-50, 425-439.
double type_form_factor(const TypeKey& t, bool implicit_h, double q);
// excluded_volume_for_atom, solvent.rs:428-438.
double type_excluded_volume(const TypeKey& t, bool implicit_h);
// united_atom_group, formfactor.rs:357-417.
int united_atom_group(const std::string& residue, const std::string& name);
// Element::from_symbol, structure.rs:109-127.
int element_from_symbol(const std::string& symbol);
double water_form_factor(double q);                     // solvent.rs:227-231
double excluded_volume_scale(double q, double c1);      // solvent.rs:448-452
double fraser_gaussian(double q, double volume);        // solvent.rs:441-443

// Per-type solvent amplitudes at one q, without occupancy.
//   Debye form (component_amplitudes, solvent.rs:454-488):
//     v = f(q), e = rho * fraser(q, V), h = rho * 0.1 * fraser(q, V_shell)
//   Multipole form (solvent_components_for_atom, spherical.rs:205-231):
//     h = rho * 0.1 * V_shell * exp(-q^2 R_s^2 / 4).
struct Amplitudes { double v, e, h; };
Amplitudes debye_amplitudes(const TypeKey& t, bool implicit_h, double q,
                            double shell_thickness, double rho);
Amplitudes multipole_amplitudes(const TypeKey& t, bool implicit_h, double q,
                                double shell_thickness, double rho);

// Grid sites in (z, y, x) scan order, as the reference emits them; errors are
// reported as a non-empty string.
struct Grid {
  std::vector<double> sites;  // 3 per site
  double voxel_volume = 0.0;
  std::string error;
};
// Voxel box of the hydration grids (solvent.rs:166-181, 286-306): the
// atoms' bounding box grown by `margin`, ceil((hi - lo) / h - 1e-10) + 1
// voxels per axis, count = their product.  False when the count overflows.
struct GridBox {
  double lo[3], hi[3];
  size_t dim[3];
  size_t count;
};
bool grid_box(const double* xyz, size_t n, double margin, double spacing, GridBox& box);
double adaptive_site_volume(double spacing);  // solvent.rs:106-115
double adaptive_shell_width(double rg);       // solvent.rs:96-104
// Grid points whose nearest atom is in (inner, outer] (synthetic workloads;
// host-side).
Grid shell_sites(const double* xyz, size_t n, double spacing, double inner,
                 double outer);

// Synthetic workloads (deterministic on every platform).
void generate_globule(size_t n, double radius, double elongation,
                      uint64_t seed, double* xyz, uint8_t* element);
void generate_domain_chain(size_t n, int domains, uint64_t seed, double* xyz,
                           uint8_t* element);

}  // namespace csx

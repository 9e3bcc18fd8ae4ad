// Hydration grids of a structure, built on its device: the host part (the
// O(N) centre, Rg, bounding box and validation, in the reference's order and
// with its messages) and the launches of grid.hip.
#include <algorithm>
#include <cmath>
#include <sstream>

#include "../../include/crabsaxs_tables.h"
#include "capi_internal.hpp"

namespace csx {

namespace {

// Leaves 2^16 of headroom below INT32_MAX for the 32-bit voxel indices and
// the padded compaction blocks.
constexpr size_t kMaxDeviceVoxels = (size_t(1) << 31) - (size_t(1) << 16);
// solvent.rs:302-306
constexpr size_t kMaxExteriorVoxels = 250000000;

GridGeom geometry(const GridBox& b, double spacing) {
  GridGeom g;
  for (int a = 0; a < 3; ++a) g.lo[a] = b.lo[a];
  g.h = spacing;
  g.nx = static_cast<int>(b.dim[0]);
  g.ny = static_cast<int>(b.dim[1]);
  g.nz = static_cast<int>(b.dim[2]);
  return g;
}

ElementRadii vdw_radii() {
  ElementRadii r{};
  for (int e = 0; e < CSX_NUM_ELEMENTS; ++e) r.r[e] = CSX_VDW_RADIUS[e];
  return r;
}

// Kept voxels of the masks as sites (3 doubles each), in scan order.
csx_status compact(Call& call, int mode, const uint8_t* a, const uint8_t* b, long long padded,
                   const GridGeom& g, Grid& out) {
  const hipStream_t st = call.stream();
  const long long nblocks = padded / 4096;
  DevBuf d_counts, d_offsets, d_sites;
  CSX_HIP(d_counts.alloc(static_cast<size_t>(nblocks) * sizeof(long long), st), "hipMallocAsync");
  CSX_HIP(d_offsets.alloc(static_cast<size_t>(nblocks + 1) * sizeof(long long), st),
          "hipMallocAsync");
  {
    TimerScope scope(call.timer(), "grid_compact", st);
    CSX_HIP(launch_grid_count(mode, a, b, padded, d_counts.as<long long>(),
                              d_offsets.as<long long>(), st), "grid_count");
  }
  long long kept = 0;
  CSX_HIP(hipMemcpyAsync(&kept, d_offsets.as<long long>() + nblocks, sizeof kept,
                         hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  out.sites.assign(static_cast<size_t>(kept) * 3, 0.0);
  if (kept > 0) {
    CSX_HIP(d_sites.alloc(out.sites.size() * sizeof(double), st), "hipMallocAsync");
    {
      TimerScope scope(call.timer(), "grid_compact", st);
      CSX_HIP(launch_grid_emit(mode, a, b, padded, g, d_offsets.as<long long>(),
                               d_sites.as<double>(), st), "grid_emit");
    }
    CSX_HIP(hipMemcpyAsync(out.sites.data(), d_sites.as<void>(), out.sites.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  }
  d_sites.reset();
  d_counts.reset();
  d_offsets.reset();
  return call.finish();
}

csx_status upload_atoms(const csx_structure* s, hipStream_t st, DevBuf& xyz, DevBuf& element) {
  CSX_HIP(xyz.alloc(s->n * 3 * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(element.alloc(s->n, st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(xyz.as<void>(), s->xyz.data(), s->n * 3 * sizeof(double),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(element.as<void>(), s->element.data(), s->n, hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  return CSX_OK;
}

// build_adaptive_hydration_grid (solvent.rs:121-223).
csx_status adaptive_grid(csx_structure* s, double spacing, double clearance, double padding,
                         Grid& g) {
  g.voxel_volume = adaptive_site_volume(spacing);
  const size_t n = s->n;
  if (n == 0) return CSX_OK;
  if (!std::isfinite(spacing) || spacing <= 0.0 || !std::isfinite(clearance) ||
      clearance <= 0.0 || !std::isfinite(padding) || padding < 0.0)
    return fail(CSX_INVALID_INPUT, "adaptive hydration-grid dimensions must be finite and positive");
  const double* xyz = s->xyz.data();
  double c[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) c[a] += xyz[3 * i + a];
  for (int a = 0; a < 3; ++a) c[a] /= static_cast<double>(n);
  double ss = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double dx = xyz[3 * i] - c[0], dy = xyz[3 * i + 1] - c[1], dz = xyz[3 * i + 2] - c[2];
    ss += dx * dx + dy * dy + dz * dz;
  }
  const double rg = std::sqrt(ss / static_cast<double>(n));
  const double outer = clearance + adaptive_shell_width(rg);
  GridBox b;
  if (!grid_box(xyz, n, padding, spacing, b) || b.count > kMaxDeviceVoxels)
    return fail(CSX_INVALID_INPUT, "adaptive hydration grid is too large");
  const GridGeom geo = geometry(b, spacing);
  const long long padded = grid_padded_voxels(static_cast<long long>(b.count));
  Call call(s->ctx);
  CSX_TRY(call.begin());
  const hipStream_t st = call.stream();
  DevBuf d_xyz, d_el, d_shell, d_mol;
  CSX_TRY(upload_atoms(s, st, d_xyz, d_el));
  CSX_HIP(d_shell.alloc(static_cast<size_t>(padded), st), "hipMallocAsync");
  CSX_HIP(d_mol.alloc(static_cast<size_t>(padded), st), "hipMallocAsync");
  CSX_HIP(launch_grid_fill(d_shell.as<uint8_t>(), padded, st), "grid_fill");
  CSX_HIP(launch_grid_fill(d_mol.as<uint8_t>(), padded, st), "grid_fill");
  const ElementRadii none{};
  const int ni = static_cast<int>(n);
  CSX_HIP(launch_grid_raster(d_xyz.as<double>(), d_el.as<uint8_t>(), ni, none, 0, outer, geo,
                             d_shell.as<uint8_t>(), 1, 0, st, call.timer()), "grid_raster");
  CSX_HIP(launch_grid_raster(d_xyz.as<double>(), d_el.as<uint8_t>(), ni, none, 0, clearance, geo,
                             d_mol.as<uint8_t>(), 1, 0, st, call.timer()), "grid_raster");
  d_xyz.reset();
  d_el.reset();
  return compact(call, 1, d_shell.as<uint8_t>(), d_mol.as<uint8_t>(), padded, geo, g);
}

// build_hydration_grid (solvent.rs:259-401).
csx_status exterior_grid(csx_structure* s, double spacing, double probe, double shell, Grid& g) {
  g.voxel_volume = spacing * spacing * spacing;
  const size_t n = s->n;
  if (n == 0) return CSX_OK;
  if (!std::isfinite(spacing) || spacing <= 0.0 || !std::isfinite(probe) || probe < 0.0 ||
      !std::isfinite(shell) || shell <= 0.0)
    return fail(CSX_INVALID_INPUT, "solvent-grid dimensions must be finite and positive");
  double vdw_max = 0.0;
  for (size_t i = 0; i < n; ++i) vdw_max = std::max(vdw_max, CSX_VDW_RADIUS[s->element[i]]);
  const double margin = vdw_max + probe + shell + spacing;
  GridBox b;
  if (!grid_box(s->xyz.data(), n, margin, spacing, b))
    return fail(CSX_INVALID_INPUT, "solvent grid is too large");
  if (b.count > kMaxExteriorVoxels)
    return fail(CSX_INVALID_INPUT, "solvent grid requires " + std::to_string(b.count) +
                                       " voxels; limit is 250000000");
  const GridGeom geo = geometry(b, spacing);
  const long long padded = grid_padded_voxels(static_cast<long long>(b.count));
  Call call(s->ctx);
  CSX_TRY(call.begin());
  const hipStream_t st = call.stream();
  DevBuf d_xyz, d_el, d_state, d_labels, d_ext;
  CSX_TRY(upload_atoms(s, st, d_xyz, d_el));
  CSX_HIP(d_state.alloc(static_cast<size_t>(padded), st), "hipMallocAsync");
  CSX_HIP(d_ext.alloc(static_cast<size_t>(padded), st), "hipMallocAsync");
  CSX_HIP(d_labels.alloc(b.count * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(launch_grid_fill(d_state.as<uint8_t>(), padded, st), "grid_fill");
  CSX_HIP(launch_grid_fill(d_ext.as<uint8_t>(), padded, st), "grid_fill");
  const ElementRadii vdw = vdw_radii();
  const int ni = static_cast<int>(n);
  // 1: molecular (vdW + probe); 2: exterior; 3: exterior within vdW + (probe + shell)
  CSX_HIP(launch_grid_raster(d_xyz.as<double>(), d_el.as<uint8_t>(), ni, vdw, 1, probe, geo,
                             d_state.as<uint8_t>(), 1, 0, st, call.timer()), "grid_raster");
  CSX_HIP(launch_grid_flood(d_state.as<uint8_t>(), geo, d_labels.as<int>(), d_ext.as<uint8_t>(),
                            st, call.timer()), "grid_flood");
  CSX_HIP(launch_grid_raster(d_xyz.as<double>(), d_el.as<uint8_t>(), ni, vdw, 1, probe + shell,
                             geo, d_state.as<uint8_t>(), 3, 1, st, call.timer()), "grid_raster");
  d_labels.reset();
  d_ext.reset();
  d_xyz.reset();
  d_el.reset();
  return compact(call, 0, d_state.as<uint8_t>(), nullptr, padded, geo, g);
}

std::string grid_key(int kind, double a, double b, double c) {
  std::ostringstream k;
  k.precision(17);
  k << (kind == kAdaptiveGrid ? "adaptive:" : "exterior:") << a << ':' << b << ':' << c;
  return k.str();
}

}  // namespace

csx_status structure_grid(csx_structure* s, int kind, double spacing, double a, double b,
                          const Grid** out) {
  const std::string key = grid_key(kind, spacing, a, b);
  auto it = s->grids.find(key);
  if (it == s->grids.end()) {
    Grid g;
    CSX_TRY(kind == kAdaptiveGrid ? adaptive_grid(s, spacing, a, b, g)
                                  : exterior_grid(s, spacing, a, b, g));
    it = s->grids.emplace(key, std::move(g)).first;
  }
  *out = &it->second;
  return CSX_OK;
}

}  // namespace csx

using namespace csx;

namespace {

csx_status grid_call(csx_structure* s, int kind, double spacing, double a, double b,
                     double* sites, size_t capacity, size_t* count, double* volume) {
  if (!s) return fail(CSX_INVALID_INPUT, "structure is required");
  const Grid* g = nullptr;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    CSX_TRY(structure_grid(s, kind, spacing, a, b, &g));
  }
  const size_t ns = g->sites.size() / 3;
  if (count) *count = ns;
  if (volume) *volume = g->voxel_volume;
  if (sites && capacity)
    std::copy(g->sites.begin(), g->sites.begin() + 3 * std::min(capacity, ns), sites);
  return CSX_OK;
}

}  // namespace

extern "C" {

csx_status csx_build_adaptive_hydration_grid(csx_structure* s, double spacing, double clearance,
                                             double padding, double* sites, size_t capacity,
                                             size_t* count, double* voxel_volume) {
  CSX_API_RANGE();
  return grid_call(s, kAdaptiveGrid, spacing, clearance, padding, sites, capacity, count,
                   voxel_volume);
}

csx_status csx_build_hydration_grid(csx_structure* s, double spacing, double probe_radius,
                                    double shell_thickness, double* sites, size_t capacity,
                                    size_t* count, double* voxel_volume) {
  CSX_API_RANGE();
  return grid_call(s, kExteriorGrid, spacing, probe_radius, shell_thickness, sites, capacity,
                   count, voxel_volume);
}

}  // extern "C"

// Coordinate features (metrics.rs:236-270): weighted Rg, Dmax and the
// normalised pair-distance distribution, with the pair histogram on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "capi_internal.hpp"

namespace csx {

size_t pr_hist_lds_bytes(int bins);
hipError_t launch_pr_dmax2(const double4* pts, int n, const int2* items, int nitems,
                           unsigned long long* out, hipStream_t stream);
hipError_t launch_pr_hist(const double4* pts, int n, const int2* items, int nitems, double step,
                          int bins, double* partial, double* out, hipStream_t stream,
                          KernelTimer* timer);

namespace {

constexpr int kBlock = 256;
// four per-wave FP64 histograms and the 8 KiB point tile within 64 KiB of LDS
constexpr size_t kMaxBins = 1536;

struct Distribution {
  std::vector<double> r, p;
  double rg = 0.0, dmax = 0.0;
  bool has_rg = false, has_dmax = false;
};

// normalize_distribution (metrics.rs:550-584); false when the area is not
// finite and positive.
bool normalize_distribution(std::vector<double> r, std::vector<double> p, Distribution& out) {
  if (r.size() != p.size() || p.empty()) return false;
  double area = 0.0;
  for (size_t k = 0; k + 1 < r.size(); ++k) area += 0.5 * (r[k + 1] - r[k]) * (p[k] + p[k + 1]);
  if (!std::isfinite(area) || area <= 0.0) return false;
  for (double& v : p) v /= area;
  double moment = 0.0;
  for (size_t k = 0; k + 1 < r.size(); ++k)
    moment += 0.5 * (r[k + 1] - r[k]) * (r[k] * r[k] * p[k] + r[k + 1] * r[k + 1] * p[k + 1]);
  const double rg = std::sqrt(std::max(moment / 2.0, 0.0));
  double maximum = 0.0;
  for (double v : p) maximum = std::max(maximum, v);
  out.has_dmax = false;
  for (size_t k = p.size(); k-- > 0;)
    if (p[k] > maximum * 0.01) {
      out.dmax = r[k];
      out.has_dmax = true;
      break;
    }
  out.has_rg = std::isfinite(rg);
  out.rg = rg;
  out.r = std::move(r);
  out.p = std::move(p);
  return true;
}

}  // namespace

}  // namespace csx

using namespace csx;

extern "C" {

csx_status csx_coordinate_features(csx_structure* s, size_t bins_in, double* rg,
                                   double* dmax_out, double* r_out, double* p_out,
                                   size_t* n_out, double* pr_rg, double* pr_dmax, int* pr_flags) {
  CSX_API_RANGE();
  if (!s || !rg || !dmax_out || !r_out || !p_out || !n_out || !pr_rg || !pr_dmax || !pr_flags)
    return fail(CSX_INVALID_INPUT, "structure and output buffers are required");
  const size_t n = s->n;
  if (n == 0) return fail(CSX_INVALID_INPUT, "coordinate features require a non-empty structure");
  const size_t bins = std::max<size_t>(bins_in, 16);
  if (bins > kMaxBins) return fail(CSX_INVALID_INPUT, "at most 1536 P(r) bins are supported");
  if (n > static_cast<size_t>(INT32_MAX / 2)) return fail(CSX_INVALID_INPUT, "structure too large");
  // weights |f(0)| (implicit hydrogens) floored at 1e-6, weighted centre and Rg
  std::vector<double> w(n), pts(4 * n);
  double total = 0.0;
  for (size_t i = 0; i < n; ++i) {
    w[i] = std::max(std::fabs(s->occ[i] * type_form_factor(s->types[s->type_of[i]], true, 0.0)),
                    1.0e-6);
    total += w[i];
  }
  double centre[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) centre[a] = centre[a] + s->xyz[3 * i + a] * w[i];
  for (int a = 0; a < 3; ++a) centre[a] = centre[a] / total;
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double dx = s->xyz[3 * i] - centre[0], dy = s->xyz[3 * i + 1] - centre[1],
                 dz = s->xyz[3 * i + 2] - centre[2];
    acc += w[i] * (dx * dx + dy * dy + dz * dz);
  }
  *rg = std::sqrt(acc / total);
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) pts[4 * i + a] = s->xyz[3 * i + a];
    pts[4 * i + 3] = w[i];
  }
  // block pairs I <= J of 256 atoms
  const int nb = static_cast<int>((n + kBlock - 1) / kBlock);
  std::vector<int> items;
  for (int I = 0; I < nb; ++I)
    for (int J = I; J < nb; ++J) items.insert(items.end(), {I, J});
  const int nitems = static_cast<int>(items.size() / 2);
  Call call(s->ctx);
  CSX_TRY(call.begin());
  const hipStream_t st = call.stream();
  Scratch d_pts(call), d_items(call), d_max(call), d_partial(call), d_hist(call);
  CSX_HIP(d_pts.alloc(pts.size() * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_items.alloc(items.size() * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(d_max.alloc(sizeof(unsigned long long), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_pts.as<void>(), pts.data(), pts.size() * sizeof(double),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_items.as<void>(), items.data(), items.size() * sizeof(int),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  const unsigned long long zero = 0;
  CSX_HIP(hipMemcpyAsync(d_max.as<void>(), &zero, sizeof zero, hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(launch_pr_dmax2(d_pts.as<double4>(), static_cast<int>(n), d_items.as<int2>(), nitems,
                          d_max.as<unsigned long long>(), st), "pr_dmax2");
  unsigned long long bitsmax = 0;
  CSX_HIP(hipMemcpyAsync(&bitsmax, d_max.as<void>(), sizeof bitsmax, hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  double dmax2;
  std::memcpy(&dmax2, &bitsmax, sizeof dmax2);
  Distribution d;
  if (!(dmax2 > 0.0)) {
    *dmax_out = 0.0;
    d.r = {0.0};
    d.p = {1.0};
    d.rg = 0.0;
    d.dmax = 0.0;
    d.has_rg = d.has_dmax = true;
  } else {
    const double dmax = std::sqrt(dmax2);
    const double step = dmax / static_cast<double>(bins);
    *dmax_out = dmax;
    CSX_HIP(d_partial.alloc(static_cast<size_t>(nitems) * bins * sizeof(double), st),
            "hipMallocAsync");
    CSX_HIP(d_hist.alloc(bins * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(launch_pr_hist(d_pts.as<double4>(), static_cast<int>(n), d_items.as<int2>(), nitems,
                           step, static_cast<int>(bins), d_partial.as<double>(),
                           d_hist.as<double>(), st, call.timer()),
            "pr_hist");
    std::vector<double> p(bins), r(bins);
    CSX_HIP(hipMemcpyAsync(p.data(), d_hist.as<void>(), bins * sizeof(double),
                           hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    for (size_t b = 0; b < bins; ++b) r[b] = (static_cast<double>(b) + 0.5) * step;
    d_partial.reset();
    d_hist.reset();
    CSX_TRY(call.finish());
    if (!normalize_distribution(r, p, d)) {
      d.r = {step * 0.5};
      d.p = {1.0};
      d.rg = 0.0;
      d.dmax = dmax;
      d.has_rg = d.has_dmax = true;
    }
  }
  d_pts.reset();
  d_items.reset();
  d_max.reset();
  CSX_TRY(call.finish());
  std::copy(d.r.begin(), d.r.end(), r_out);
  std::copy(d.p.begin(), d.p.end(), p_out);
  *n_out = d.r.size();
  *pr_rg = d.rg;
  *pr_dmax = d.dmax;
  *pr_flags = (d.has_rg ? 1 : 0) | (d.has_dmax ? 2 : 0);
  return CSX_OK;
}

}  // extern "C"

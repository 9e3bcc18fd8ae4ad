// Multipole path (spherical.rs:266-565): geometry cache, per-q expansion
// orders, amplitude tables and the C-ABI entry points.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <cmath>

#include "capi_internal.hpp"

namespace csx {

int multipole_chunk();
int multipole_slab();
int multipole_tiles_per_group(int ntiles, int tuned);
size_t multipole_coeff_lds_bytes(int l_range);
int multipole_max_types();
hipError_t launch_multipole_geometry(const double* xyz, int npts, const double* centre, int l_max,
                                     const double* norms, const int* tile_first, int ntiles,
                                     int npad, float* radius, void* Y, hipStream_t stream);
int multipole_nodes();
int multipole_harmonic_exponent();
hipError_t launch_multipole_point_terms(const double* xyz, const double* occ, const int* bin_of,
                                        int p0, int p1, const double* centre, int l_max,
                                        const double* norms, double h, const double* node_x,
                                        const double* node_den, double* terms, double* weights,
                                        hipStream_t stream);
hipError_t launch_multipole_bin_moments(const double* terms, const double* weights,
                                        const int* bins, int nbins, int p0, int nlm,
                                        double* moments, hipStream_t stream);
hipError_t launch_multipole_moment_max(const double* moments, size_t n, unsigned long long* out,
                                       hipStream_t stream);
hipError_t launch_multipole_moment_tiles(const double* moments, int nvirt, int l_max,
                                         const int* tile_first, int ntiles, int npad, int yexp,
                                         void* Y, hipStream_t stream);
hipError_t launch_multipole_coeff(const float* radius, const float* occ, const int* chunks,
                                  int nchunk, int npad, const void* Y, int ntiles,
                                  const int* tile_l, int tpg, int max_l_range, int aexp,
                                  int yexp, const int* type_chunk, int ntypes, const float* qv,
                                  const int* qlmax, int nqp, float* partial, double* coef,
                                  hipStream_t stream, KernelTimer* timer);
hipError_t launch_multipole_contract(const double* coef, const double* amp, int ntypes, int nch,
                                     const int* tile_l, const int* qlmax, int nqp, int ncol,
                                     double* out, hipStream_t stream);

namespace {

// Largest l range of a column-tile group (the LDS rows one workgroup needs).
int max_group_l_range(const std::vector<int>& tile_l, int tiles_per_group) {
  int r = 1;
  for (size_t g0 = 0; g0 < tile_l.size(); g0 += tiles_per_group) {
    const size_t g1 = std::min(tile_l.size(), g0 + tiles_per_group);
    r = std::max(r, tile_l[g1 - 1] - tile_l[g0] + 1);
  }
  return r;
}

// harmonic_norms (spherical.rs:33-51).
std::vector<double> harmonic_norms(int L) {
  std::vector<double> norms((L + 1) * (L + 1), 0.0);
  for (int l = 0; l <= L; ++l)
    for (int m = 0; m <= l; ++m) {
      double ratio = 1.0;
      for (int k = l - m + 1; k <= l + m; ++k) ratio /= static_cast<double>(k);
      norms[l * l + l + m] = std::sqrt((static_cast<double>(2 * l + 1) / (4.0 * kPi)) * ratio);
    }
  return norms;
}

// adaptive_l_max (spherical.rs:201-203).
int adaptive_l_max(double q, double max_radius, int L) {
  const double o = std::ceil(q * max_radius);
  long v = (o > 1e6 ? 1000000L : static_cast<long>(std::max(o, 0.0))) + 8;
  v = std::max<long>(v, std::min(L, 8));
  return static_cast<int>(std::min<long>(v, L));
}

}  // namespace

// Radial-interpolation choice of a point set for one call: 1 = per point,
// 2 = nodes always, 0 = nodes where they are fewer than half the points.
// The bin width is the largest power of two h with q_max h <= 1.04, where
// 8 Chebyshev nodes per bin interpolate j_l(q r) to 1.0e-9 (0 = per point:
// tuned off, or q_max beyond 2^10 h-steps).
static double radial_bin_width(const double* q, size_t nq, int mode) {
  if (mode == 1) return 0.0;
  double qmax = 0.0;
  for (size_t k = 0; k < nq; ++k) qmax = std::max(qmax, std::fabs(q[k]));
  if (!(qmax > 0.0)) return 128.0;
  const double h = std::min(128.0, std::exp2(std::floor(std::log2(1.04 / qmax))));
  return h >= 1.0 / 1024.0 ? h : 0.0;
}

// Chebyshev nodes of the first kind on [-1, 1] and their Lagrange
// denominators prod_{k != m} (x_m - x_k).
static void chebyshev_nodes(int n, std::vector<double>& x, std::vector<double>& den) {
  x.resize(n);
  den.resize(n);
  for (int m = 0; m < n; ++m) x[m] = std::cos((2.0 * m + 1.0) * kPi / (2.0 * n));
  for (int m = 0; m < n; ++m) {
    double d = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != m) d *= x[m] - x[k];
    den[m] = d;
  }
}

// The multipole geometry of a point set for l_max L and bin width h (0: per
// point) under radial mode `mode`, built on first use and cached on the set.
static csx_status get_geometry(Call& call, csx_structure* s, PointSet& ps, int L, double h,
                               int mode, MultipoleGeometry** out) {
  std::lock_guard<std::mutex> lock(ps.geom_mu);
  auto& slot = ps.geoms[std::make_tuple(L, h, mode)];
  if (slot) {
    *out = slot.get();
    return CSX_OK;
  }
  std::unique_ptr<MultipoleGeometry> g(new MultipoleGeometry);
  g->L = L;
  const int n = ps.natoms, ns = ps.nsites;
  g->npts = n + ns;
  std::vector<int> tile_first(L + 2), tile_l;
  for (int l = 0; l <= L; ++l) {
    tile_first[l] = static_cast<int>(tile_l.size());
    for (int t = 0; t < (2 * l + 1 + 15) / 16; ++t) tile_l.push_back(l);
  }
  tile_first[L + 1] = static_cast<int>(tile_l.size());
  g->ntiles = static_cast<int>(tile_l.size());
  g->h_tile_l = tile_l;
  // Points (atoms, then hydration sites as type T) in segments by type, each
  // segment ordered by distance from the centre and padded to whole slabs
  // with zero-occupancy copies of its last point.  Every chunk of the
  // coefficient kernel then holds one type, whose amplitudes factor out of
  // its sum, and the lanes generating j_l(q r) for one slab see nearly equal
  // radii (nearly equal recurrence trip counts).  The sums over points do not
  // depend on the order.
  // The order is by (type, FP32 r^2, index): an LSD radix sort of keys that
  // start in index order, 3 passes of 11 bits over the non-negative FP32 bit
  // patterns (which order like the values), then a stable pass by type.
  const int T = static_cast<int>(s->types.size());
  const int npts = g->npts;
  auto point = [&](int i) { return i < n ? &s->xyz[3 * size_t(i)] : &ps.site_xyz[3 * size_t(i - n)]; };
  auto type_of = [&](int i) { return i < n ? s->type_of[i] : T; };
  std::vector<uint64_t> key(npts), tmp(npts);
  for (int i = 0; i < npts; ++i) {
    const double* p = point(i);
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a) r2 += (p[a] - s->centre[a]) * (p[a] - s->centre[a]);
    const float f = static_cast<float>(r2);
    uint32_t bits;
    std::memcpy(&bits, &f, sizeof bits);
    key[i] = (uint64_t(bits) << 32) | uint32_t(i);
  }
  for (int pass = 0; pass < 3; ++pass) {
    const int shift = 32 + 11 * pass;
    std::vector<int> first(2049, 0);
    for (int i = 0; i < npts; ++i) ++first[((key[i] >> shift) & 2047) + 1];
    for (int b = 0; b < 2048; ++b) first[b + 1] += first[b];
    for (int i = 0; i < npts; ++i) tmp[first[(key[i] >> shift) & 2047]++] = key[i];
    key.swap(tmp);
  }
  std::vector<int> seg(T + 2, 0);  // segment starts in the unpadded order
  for (int i = 0; i < npts; ++i) ++seg[type_of(i) + 1];
  for (int t = 0; t <= T; ++t) seg[t + 1] += seg[t];
  {
    std::vector<int> next(seg.begin(), seg.end() - 1);
    for (int i = 0; i < npts; ++i) {
      const int idx = static_cast<int>(key[i] & 0xffffffffu);
      tmp[next[type_of(idx)]++] = key[i];
    }
    key.swap(tmp);
  }
  for (int i = 0; i < npts; ++i) {
    const double* p = point(i);
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double d = p[a] - s->centre[a];
      r2 += d * d;
    }
    double& rm = i < n ? g->rmax_atoms : g->rmax_sites;
    rm = std::max(rm, std::sqrt(r2));
  }
  // Radial bins: per type, the runs of points whose sort-key radius falls in
  // one [b h, (b + 1) h); each bin becomes kNodes virtual points.
  struct Bin {
    int t, b, p0, p1;
  };
  std::vector<Bin> bins;
  std::vector<int> bin_of;
  const int nodes = multipole_nodes();
  if (h > 0.0) {
    bin_of.resize(npts);
    for (int t = 0; t <= T; ++t)
      for (int k = seg[t]; k < seg[t + 1]; ++k) {
        float r2f;
        const uint32_t bits = static_cast<uint32_t>(key[k] >> 32);
        std::memcpy(&r2f, &bits, sizeof r2f);
        const int b = static_cast<int>(std::floor(std::sqrt(static_cast<double>(r2f)) / h));
        if (bins.empty() || bins.back().t != t || bins.back().b != b) bins.push_back({t, b, k, k});
        bins.back().p1 = k + 1;
        bin_of[k] = b;
      }
    const size_t nvirt = bins.size() * static_cast<size_t>(nodes);
    if (bins.size() > 65535 || (mode != 2 && 2 * nvirt > static_cast<size_t>(npts))) {
      bins.clear();
      h = 0.0;
    }
  }
  g->bin_width = h;
  const std::vector<double> norms = harmonic_norms(L);
  const hipStream_t st = call.stream();
  if (h == 0.0) {
    // Padded layout and the chunk table (first slab, end slab, type).
    std::vector<int> chunks;
    std::vector<double> xyz;
    std::vector<float> occ;
    int slab = 0;
    std::vector<int> type_chunk(T + 2, 0);
    const int slab_points = multipole_slab();
    float occ_max = 0.0f;
    for (int t = 0; t <= T; ++t) {
      const int count = seg[t + 1] - seg[t];
      type_chunk[t] = static_cast<int>(chunks.size() / 4);
      if (count == 0) continue;
      const int slabs = (count + slab_points - 1) / slab_points;
      for (int k = 0; k < slabs * slab_points; ++k) {
        const int idx = static_cast<int>(key[seg[t] + std::min(k, count - 1)] & 0xffffffffu);
        const double* p = point(idx);
        xyz.insert(xyz.end(), p, p + 3);
        const float o = k >= count ? 0.0f : idx < n ? static_cast<float>(s->occ[idx]) : 1.0f;
        occ.push_back(o);
        occ_max = std::max(occ_max, std::fabs(o));
      }
      const int per = multipole_chunk() / slab_points;
      for (int c = 0; c < slabs; c += per)
        chunks.insert(chunks.end(), {slab + c, slab + std::min(slabs, c + per), t, 0});
      slab += slabs;
    }
    g->npad = slab * slab_points;
    // j_l occ enters the FP16 contraction times 2^aexp: the largest occupancy
    // maps below 2^14 (occupancy 1 -> 2^14, j_l <= 1).
    g->aexp = 14;
    if (occ_max > 1.0f) g->aexp -= static_cast<int>(std::ceil(std::log2(static_cast<double>(occ_max))));
    if (!(occ_max < 1e30f)) return fail(CSX_INVALID_INPUT, "occupancy not finite");
    g->nchunk = static_cast<int>(chunks.size() / 4);
    type_chunk[T + 1] = g->nchunk;
    // Types of the point set: the atom types, then the site type when the set
    // has sites; the chunk table keeps the site chunks last.
    if (ps.nsites == 0) type_chunk.pop_back();
    g->h_type_chunk = type_chunk;
    DevBuf d_xyz, d_centre, d_norms, d_first;
    CSX_HIP(d_xyz.alloc(xyz.size() * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_centre.alloc(3 * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_norms.alloc(norms.size() * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_first.alloc(tile_first.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(hipMemcpyAsync(d_xyz.as<void>(), xyz.data(), xyz.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_centre.as<void>(), s->centre, 3 * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_norms.as<void>(), norms.data(), norms.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_first.as<void>(), tile_first.data(), tile_first.size() * sizeof(int),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(g->radius.alloc(g->npad * sizeof(float)), "hipMalloc radii");
    CSX_HIP(g->occ.upload(occ.data(), occ.size() * sizeof(float)), "upload occupancy");
    CSX_HIP(g->chunks.upload(chunks.data(), chunks.size() * sizeof(int)), "upload chunks");
    CSX_HIP(g->tile_l.upload(tile_l.data(), tile_l.size() * sizeof(int)), "upload tiles");
    CSX_HIP(g->type_chunk.upload(type_chunk.data(), type_chunk.size() * sizeof(int)),
            "upload type chunks");
    const size_t ybytes = static_cast<size_t>(g->ntiles) * g->npad * 16 * sizeof(float);
    CSX_HIP(g->Y.alloc(ybytes), "hipMalloc harmonics");
    CSX_HIP(launch_multipole_geometry(d_xyz.as<double>(), g->npad, d_centre.as<double>(), L,
                                      d_norms.as<double>(), d_first.as<int>(), g->ntiles, g->npad,
                                      g->radius.as<float>(), g->Y.as<void>(), st),
            "multipole_geometry");
    d_xyz.reset();
    d_centre.reset();
    d_norms.reset();
    d_first.reset();
    CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    g->yexp = multipole_harmonic_exponent();
  } else {
    // Virtual points per type: kNodes per bin at the bin's Chebyshev nodes,
    // occupancy 1 (the point occupancies are in the moments), padded to
    // whole slabs with zero occupancy; chunks of one slab.
    std::vector<double> nx, nden;
    chebyshev_nodes(nodes, nx, nden);
    const int slab_points = multipole_slab();
    std::vector<int> chunks, hbins, type_chunk(T + 2, 0);
    std::vector<float> vrad, vocc;
    int vslab = 0;
    size_t bi = 0;
    for (int t = 0; t <= T; ++t) {
      type_chunk[t] = static_cast<int>(chunks.size() / 4);
      const size_t b0 = bi;
      while (bi < bins.size() && bins[bi].t == t) ++bi;
      const int count = static_cast<int>(bi - b0) * nodes;
      if (count == 0) continue;
      const int slabs = (count + slab_points - 1) / slab_points;
      for (size_t b = b0; b < bi; ++b) {
        hbins.insert(hbins.end(), {bins[b].p0, bins[b].p1, static_cast<int>(vrad.size()), 0});
        for (int m = 0; m < nodes; ++m) {
          vrad.push_back(static_cast<float>((bins[b].b + 0.5) * h + 0.5 * h * nx[m]));
          vocc.push_back(1.0f);
        }
      }
      for (int k = count; k < slabs * slab_points; ++k) {
        vrad.push_back(vrad.back());
        vocc.push_back(0.0f);
      }
      for (int c = 0; c < slabs; ++c) chunks.insert(chunks.end(), {vslab + c, vslab + c + 1, t, 0});
      vslab += slabs;
    }
    g->npad = vslab * slab_points;
    g->aexp = 14;  // occupancy 1: j_l maps below 2^14
    g->nchunk = static_cast<int>(chunks.size() / 4);
    type_chunk[T + 1] = g->nchunk;
    if (ps.nsites == 0) type_chunk.pop_back();
    g->h_type_chunk = type_chunk;
    // Real points in the sorted order, FP64 coordinates and occupancies.
    std::vector<double> xyz(3 * size_t(npts)), occ(npts);
    for (int k = 0; k < npts; ++k) {
      const int idx = static_cast<int>(key[k] & 0xffffffffu);
      const double* p = point(idx);
      std::copy(p, p + 3, xyz.begin() + 3 * size_t(k));
      occ[k] = idx < n ? s->occ[idx] : 1.0;
      if (!std::isfinite(occ[k])) return fail(CSX_INVALID_INPUT, "occupancy not finite");
    }
    const int nlm = (L + 1) * (L + 1);
    // Point chunks of whole bins, each with at most ~512 MB of FP64 terms
    // (or one bin).
    const size_t cap = std::max<size_t>(1, (size_t(512) << 20) / (size_t(nlm) * sizeof(double)));
    std::vector<std::pair<size_t, size_t>> pchunks;  // bin ranges
    size_t maxpts = 0;
    for (size_t b = 0; b < bins.size();) {
      size_t e = b, pts = 0;
      while (e < bins.size() && (e == b || pts + size_t(bins[e].p1 - bins[e].p0) <= cap))
        pts += size_t(bins[e].p1 - bins[e].p0), ++e;
      pchunks.push_back({b, e});
      maxpts = std::max(maxpts, pts);
      b = e;
    }
    DevBuf d_xyz, d_occ, d_bin, d_centre, d_norms, d_first, d_bins, d_terms, d_w, d_mom, d_max;
    CSX_HIP(d_xyz.alloc(xyz.size() * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_occ.alloc(occ.size() * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_bin.alloc(bin_of.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(d_centre.alloc(3 * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_norms.alloc(norms.size() * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_first.alloc(tile_first.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(d_bins.alloc(hbins.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(d_terms.alloc((maxpts + 15) / 16 * 16 * nlm * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_w.alloc(maxpts * nodes * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_mom.alloc(size_t(g->npad) * nlm * sizeof(double), st), "hipMallocAsync");
    CSX_HIP(d_max.alloc(sizeof(unsigned long long), st), "hipMallocAsync");
    CSX_HIP(hipMemsetAsync(d_mom.as<void>(), 0, size_t(g->npad) * nlm * sizeof(double), st),
            "hipMemsetAsync");
    CSX_HIP(hipMemcpyAsync(d_xyz.as<void>(), xyz.data(), xyz.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_occ.as<void>(), occ.data(), occ.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_bin.as<void>(), bin_of.data(), bin_of.size() * sizeof(int),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_centre.as<void>(), s->centre, 3 * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_norms.as<void>(), norms.data(), norms.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_first.as<void>(), tile_first.data(), tile_first.size() * sizeof(int),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_bins.as<void>(), hbins.data(), hbins.size() * sizeof(int),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    for (const auto& pc : pchunks) {
      const int p0 = bins[pc.first].p0, p1 = bins[pc.second - 1].p1;
      CSX_HIP(launch_multipole_point_terms(d_xyz.as<double>(), d_occ.as<double>(), d_bin.as<int>(),
                                           p0, p1, d_centre.as<double>(), L, d_norms.as<double>(),
                                           h, nx.data(), nden.data(), d_terms.as<double>(),
                                           d_w.as<double>(), st),
              "multipole_point_terms");
      CSX_HIP(launch_multipole_bin_moments(d_terms.as<double>(), d_w.as<double>(),
                                           d_bins.as<int>() + 4 * pc.first,
                                           static_cast<int>(pc.second - pc.first), p0, nlm,
                                           d_mom.as<double>(), st),
              "multipole_bin_moments");
    }
    // The moments enter the FP16 contraction times 2^yexp: the largest maps
    // below 2^14.
    unsigned long long mbits = 0;
    CSX_HIP(launch_multipole_moment_max(d_mom.as<double>(), size_t(g->npad) * nlm,
                                        d_max.as<unsigned long long>(), st),
            "multipole_moment_max");
    CSX_HIP(hipMemcpyAsync(&mbits, d_max.as<void>(), sizeof mbits, hipMemcpyDeviceToHost, st),
            "hipMemcpyAsync");
    CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    double mmax;
    std::memcpy(&mmax, &mbits, sizeof mmax);
    if (!std::isfinite(mmax)) return fail(CSX_INVALID_INPUT, "multipole moments not finite");
    g->yexp = mmax > 0.0 ? std::max(-60, std::min(60, static_cast<int>(std::floor(std::log2(16384.0 / mmax)))))
                         : multipole_harmonic_exponent();
    const size_t ybytes = static_cast<size_t>(g->ntiles) * g->npad * 16 * sizeof(float);
    CSX_HIP(g->Y.alloc(ybytes), "hipMalloc harmonics");
    CSX_HIP(launch_multipole_moment_tiles(d_mom.as<double>(), g->npad, L, d_first.as<int>(),
                                          g->ntiles, g->npad, g->yexp, g->Y.as<void>(), st),
            "multipole_moment_tiles");
    CSX_HIP(g->radius.upload(vrad.data(), vrad.size() * sizeof(float)), "upload node radii");
    CSX_HIP(g->occ.upload(vocc.data(), vocc.size() * sizeof(float)), "upload node occupancy");
    CSX_HIP(g->chunks.upload(chunks.data(), chunks.size() * sizeof(int)), "upload chunks");
    CSX_HIP(g->tile_l.upload(tile_l.data(), tile_l.size() * sizeof(int)), "upload tiles");
    CSX_HIP(g->type_chunk.upload(type_chunk.data(), type_chunk.size() * sizeof(int)),
            "upload type chunks");
    g->nvirtual = static_cast<int>(bins.size()) * nodes;
    CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  }
  *out = g.get();
  slot = std::move(g);
  return CSX_OK;
}

// Per-type coefficient tables S_t[q][lm] of a point set for a q grid and
// l_max, in the point set's cache: computed by the coefficient kernel on a
// miss.  On return the entry is ready and *lock holds it.
static csx_status type_coefficients(Call& call, PointSet& ps, const MultipoleGeometry& g,
                                    const double* q, size_t nq, double rmax,
                                    std::shared_ptr<CacheEntry>* out,
                                    std::unique_lock<std::mutex>* lock) {
  std::string key = "multipole:" + std::to_string(g.L) + ":";
  key.append(reinterpret_cast<const char*>(&g.bin_width), sizeof g.bin_width);
  key.append(reinterpret_cast<const char*>(q), nq * sizeof(double));
  std::shared_ptr<CacheEntry> e = cache_acquire(call.ctx(), ps, key, lock);
  *out = e;
  if (e->ready) return CSX_OK;
  struct Dropper {
    PointSet& ps;
    const std::shared_ptr<CacheEntry>& e;
    ~Dropper() {
      if (!e->ready) cache_drop(ps, e);
    }
  } dropper{ps, e};
  const int TT = static_cast<int>(g.h_type_chunk.size()) - 1;
  if (TT != ps.ntypes) return fail(CSX_DEVICE_ERROR, "multipole type table mismatch");
  const size_t nqp = (nq + 15) / 16 * 16;
  const int ncol = g.ntiles * 16;
  std::vector<float> qv(nqp, 0.0f);
  std::vector<int> qlmax(nqp, 0);
  for (size_t k = 0; k < nq; ++k) {
    qv[k] = static_cast<float>(q[k]);
    qlmax[k] = adaptive_l_max(q[k], rmax, g.L);
  }
  const hipStream_t st = call.stream();
  // A cached entry owns its arrays; a private one (no cache) lives in the
  // call's scratch and ends with the call.
  const size_t coef_bytes = size_t(TT) * nqp * ncol * sizeof(double);
  if (e->cached) {
    CSX_HIP(e->qlmax.alloc(nqp * sizeof(int)), "hipMalloc orders");
    CSX_HIP(e->coef.alloc(coef_bytes), "hipMalloc coefficients");
    e->qlmax_ptr = e->qlmax.as<int>();
    e->coef_ptr = e->coef.as<double>();
  } else {
    void* pq = nullptr;
    void* pc = nullptr;
    CSX_HIP(call.take_scratch(nqp * sizeof(int), &pq), "hipMalloc orders");
    CSX_HIP(call.take_scratch(coef_bytes, &pc), "hipMalloc coefficients");
    e->qlmax_ptr = static_cast<int*>(pq);
    e->coef_ptr = static_cast<double*>(pc);
  }
  Scratch d_q(call), d_partial(call);
  CSX_HIP(d_q.alloc(nqp * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_partial.alloc(static_cast<size_t>(g.nchunk) * nqp * ncol * sizeof(float), st),
          "hipMallocAsync");
  void* h_in = nullptr;
  CSX_HIP(call.take_pinned(2, nqp * (sizeof(float) + sizeof(int)), &h_in), "hipHostMalloc");
  char* in = static_cast<char*>(h_in);
  std::memcpy(in, qv.data(), nqp * sizeof(float));
  std::memcpy(in + nqp * sizeof(float), qlmax.data(), nqp * sizeof(int));
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), in, nqp * sizeof(float), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(e->qlmax_ptr, in + nqp * sizeof(float), nqp * sizeof(int),
                         hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  int tuned;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    tuned = call.ctx()->tuning.mp_group_tiles;
  }
  const int tpg = multipole_tiles_per_group(g.ntiles, tuned);
  const int l_range = max_group_l_range(g.h_tile_l, tpg);
  if (multipole_coeff_lds_bytes(l_range) > call.ctx()->lds_per_workgroup)
    return fail(CSX_INVALID_INPUT, "multipole l_max too large for the device's LDS");
  CSX_HIP(launch_multipole_coeff(g.radius.as<float>(), g.occ.as<float>(), g.chunks.as<int>(),
                                 g.nchunk, g.npad, g.Y.as<void>(), g.ntiles, g.tile_l.as<int>(),
                                 tpg, l_range, g.aexp, g.yexp, g.type_chunk.as<int>(), TT,
                                 d_q.as<float>(),
                                 e->qlmax_ptr, static_cast<int>(nqp), d_partial.as<float>(),
                                 e->coef_ptr, st, call.timer()),
          "multipole_coeff");
  // A cached entry is complete once the stream has run the kernels; every
  // later use of it is stream-ordered after a synchronisation of this call.
  // A private one is used only on this stream.
  if (e->cached) CSX_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  e->nqp = static_cast<int>(nqp);
  e->ncol = ncol;
  e->bytes = size_t(TT) * nqp * ncol * sizeof(double) + nqp * sizeof(int) + key.size();
  e->ready = true;
  cache_commit(call.ctx(), ps, e);
  return CSX_OK;
}

// mode 0: vacuum intensity; 1: solvent intensity; 2: six-term basis.
static csx_status multipole_run(csx_structure* s, const double* q, size_t nq, int L,
                                bool implicit_h, const csx_solvent_params* p, int mode,
                                double* out) {
  const int nterms = mode == 2 ? 6 : 1;
  if (s->n == 0) {
    std::fill(out, out + nterms * nq, 0.0);
    return CSX_OK;
  }
  Call call(s->ctx);
  CSX_TRY(call.begin());
  std::shared_ptr<PointSet> ps;
  CSX_TRY(get_point_set(s, mode == 0 ? nullptr : p, &ps));
  if (ps->ntypes > multipole_max_types())
    return fail(CSX_INVALID_INPUT, "too many form-factor types for the multipole path");
  int radial;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    radial = call.ctx()->tuning.mp_radial;
  }
  MultipoleGeometry* g = nullptr;
  CSX_TRY(get_geometry(call, s, *ps, L, radial_bin_width(q, nq, radial), radial, &g));
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    call.ctx()->last_mp_bin = g->bin_width;
    call.ctx()->last_mp_points = g->bin_width > 0.0 ? g->nvirtual : g->npts;
  }
  const bool grid = mode != 0 && p->hydration_model != CSX_HYDRATION_ANALYTIC;
  const double rmax = grid ? std::max(g->rmax_atoms, g->rmax_sites) : g->rmax_atoms;
  std::shared_ptr<CacheEntry> coef;
  std::unique_lock<std::mutex> coef_lock;
  CSX_TRY(type_coefficients(call, *ps, *g, q, nq, rmax, &coef, &coef_lock));
  // FP64 amplitudes per type and q: the vacuum form factor (mode 0), the
  // combined solvent amplitude v - G(c1) e + c2 h (mode 1), or the three
  // channels v, e, h (mode 2); the site type carries the water amplitude of
  // spherical.rs:449-451.
  const int nch = mode == 2 ? 3 : 1;
  const int nqp = coef->nqp;
  const int T = static_cast<int>(s->types.size());
  const int TT = ps->ntypes;
  std::vector<double> amp(static_cast<size_t>(TT) * nqp * nch, 0.0);
  for (size_t k = 0; k < nq; ++k) {
    const double gscale = mode == 1 ? excluded_volume_scale(q[k], p->c1) : 0.0;
    for (int t = 0; t < T; ++t) {
      double* a = &amp[(static_cast<size_t>(t) * nqp + k) * nch];
      if (mode == 0) {
        a[0] = type_form_factor(s->types[t], implicit_h, q[k]);
        continue;
      }
      const Amplitudes c = multipole_amplitudes(s->types[t], implicit_h, q[k],
                                                p->shell_thickness_angstrom,
                                                p->bulk_density_e_per_a3);
      if (mode == 1) {
        a[0] = c.v - gscale * c.e + (grid ? 0.0 : p->c2 * c.h);
      } else {
        a[0] = c.v;
        a[1] = c.e;
        a[2] = grid ? 0.0 : c.h;
      }
    }
    if (grid) {
      const double site = p->bulk_density_e_per_a3 / 10.0 * ps->site_volume * water_form_factor(q[k]);
      double* a = &amp[(static_cast<size_t>(T) * nqp + k) * nch];
      if (mode == 1) a[0] = p->c2 * site;
      else a[2] = site;
    }
  }
  const hipStream_t st = call.stream();
  Scratch d_amp(call), d_out(call);
  CSX_HIP(d_amp.alloc(amp.size() * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(size_t(nqp) * nterms * sizeof(double), st), "hipMallocAsync");
  void* h_in = nullptr;
  void* h_res = nullptr;
  CSX_HIP(call.take_pinned(0, amp.size() * sizeof(double), &h_in), "hipHostMalloc");
  CSX_HIP(call.take_pinned(1, size_t(nqp) * nterms * sizeof(double), &h_res), "hipHostMalloc");
  std::memcpy(h_in, amp.data(), amp.size() * sizeof(double));
  CSX_HIP(hipMemcpyAsync(d_amp.as<void>(), h_in, amp.size() * sizeof(double),
                         hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(launch_multipole_contract(coef->coef_ptr, d_amp.as<double>(), TT, nch,
                                    g->tile_l.as<int>(), coef->qlmax_ptr, nqp, coef->ncol,
                                    d_out.as<double>(), st),
          "multipole_contract");
  CSX_HIP(hipMemcpyAsync(h_res, d_out.as<void>(), size_t(nqp) * nterms * sizeof(double),
                         hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  d_amp.reset();
  d_out.reset();
  CSX_TRY(call.finish());
  const double* res = static_cast<const double*>(h_res);
  const double scale = 4.0 * kPi;
  for (size_t k = 0; k < nq; ++k)
    for (int t = 0; t < nterms; ++t) {
      const double v = scale * res[k * nterms + t];
      out[t * nq + k] = mode == 2 ? v : std::max(v, 0.0);
    }
  return CSX_OK;
}

static csx_status check(csx_structure* s, const double* q, size_t nq, int L, const void* out) {
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (nq > 0 && (!q || !out)) return fail(CSX_INVALID_INPUT, "q and output buffers are required");
  if (L < 0 || L > 64) return fail(CSX_INVALID_INPUT, "multipole l_max must be between 0 and 64");
  if (nq > (1u << 20)) return fail(CSX_INVALID_INPUT, "too many q values");
  return CSX_OK;
}

csx_status multipole_dispatch(csx_structure* s, const double* q, size_t nq, int l_max,
                              bool implicit_h, const csx_solvent_params* p, int precision,
                              double* out) {
  (void)precision;  // the multipole contraction has one (FP32 MFMA) arithmetic
  CSX_TRY(check(s, q, nq, l_max, out));
  if (nq == 0) return CSX_OK;
  return multipole_run(s, q, nq, l_max, implicit_h, p, p ? 1 : 0, out);
}

}  // namespace csx

using namespace csx;

extern "C" {

csx_status csx_multipole_intensity(csx_structure* s, const double* q, size_t nq, int l_max,
                                   int implicit_h, double* out) {
  CSX_API_RANGE();
  CSX_TRY(check(s, q, nq, l_max, out));
  if (nq == 0) return CSX_OK;
  return multipole_run(s, q, nq, l_max, implicit_h != 0, nullptr, 0, out);
}

csx_status csx_multipole_solvent_intensity(csx_structure* s, const double* q, size_t nq,
                                           int l_max, int implicit_h,
                                           const csx_solvent_params* p, double* out) {
  CSX_API_RANGE();
  CSX_TRY(check(s, q, nq, l_max, out));
  if (!p) return fail(CSX_INVALID_INPUT, "solvent parameters are required");
  if (nq == 0) return CSX_OK;
  return multipole_run(s, q, nq, l_max, implicit_h != 0, p, 1, out);
}

csx_status csx_multipole_solvent_basis(csx_structure* s, const double* q, size_t nq, int l_max,
                                       int implicit_h, const csx_solvent_params* p, double* out) {
  CSX_API_RANGE();
  CSX_TRY(check(s, q, nq, l_max, out));
  if (!p) return fail(CSX_INVALID_INPUT, "solvent parameters are required");
  if (nq == 0) return CSX_OK;
  return multipole_run(s, q, nq, l_max, implicit_h != 0, p, 2, out);
}

}  // extern "C"

// C-ABI entry points: contexts, structures, the Debye path and compute_curve.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <functional>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>

#include "capi_internal.hpp"
#include "../../include/crabsaxs_tables.h"

namespace csx {

static thread_local std::string g_error;

csx_status fail(csx_status st, const std::string& msg) {
  g_error = msg;
  return st;
}

csx_status hip_fail(hipError_t e, const char* what) {
  const csx_status st = e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? CSX_OUT_OF_MEMORY
                        : e == hipErrorNoDevice || e == hipErrorInvalidDevice   ? CSX_NO_DEVICE
                                                                                 : CSX_DEVICE_ERROR;
  return fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------ device memory --

static std::atomic<size_t> g_device_bytes{0};

static size_t device_limit() {
  static const size_t limit = [] {
    const char* v = std::getenv("CSX_DEVICE_MEMORY_LIMIT");
    if (!v || !*v) return SIZE_MAX;
    char* end = nullptr;
    const unsigned long long n = std::strtoull(v, &end, 10);
    return end && *end == '\0' ? static_cast<size_t>(n) : SIZE_MAX;
  }();
  return limit;
}

static bool reserve(size_t bytes) {
  const size_t limit = device_limit();
  size_t held = g_device_bytes.load();
  do {
    if (bytes > limit || held > limit - bytes) return false;
  } while (!g_device_bytes.compare_exchange_weak(held, held + bytes));
  return true;
}

hipError_t device_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (!reserve(bytes)) return hipErrorOutOfMemory;
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) g_device_bytes -= bytes;
  return e;
}

hipError_t device_alloc_async(void** p, size_t bytes, hipStream_t s) {
  *p = nullptr;
  if (!reserve(bytes)) return hipErrorOutOfMemory;
  const hipError_t e = hipMallocAsync(p, bytes, s);
  if (e != hipSuccess) g_device_bytes -= bytes;
  return e;
}

void device_free(void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  g_device_bytes -= bytes;
}

void device_free_async(void* p, size_t bytes, hipStream_t s) {
  if (!p) return;
  (void)hipFreeAsync(p, s);
  g_device_bytes -= bytes;
}

// ------------------------------------------------------------ point sets --

static std::string set_key(const csx_solvent_params* p) {
  if (!p || p->hydration_model == CSX_HYDRATION_ANALYTIC) return "atoms";
  if (p->hydration_model == CSX_HYDRATION_EXPLICIT_SITES) return "sites";
  std::ostringstream k;
  k.precision(17);
  if (p->hydration_model == CSX_HYDRATION_GRID)
    k << "grid:" << p->shell_thickness_angstrom;
  else
    k << "adaptive:" << p->hydration_grid_spacing_angstrom;
  return k.str();
}

static csx_status build_point_set(csx_structure* s, const csx_solvent_params* p, PointSet& ps) {
  const int T = static_cast<int>(s->types.size());
  const size_t n = s->n;
  if (p && p->hydration_model == CSX_HYDRATION_EXPLICIT_SITES) {
    if (!s->has_sites)
      return fail(CSX_INVALID_INPUT,
                  "the explicit-sites hydration model needs sites attached to the structure");
    ps.site_xyz = s->sites;
    ps.site_volume = s->site_volume;
    ps.nsites = static_cast<int>(ps.site_xyz.size() / 3);
  } else if (p && p->hydration_model != CSX_HYDRATION_ANALYTIC) {
    const Grid* g = nullptr;
    CSX_TRY(p->hydration_model == CSX_HYDRATION_GRID
                ? structure_grid(s, kExteriorGrid, 1.0, 1.4, p->shell_thickness_angstrom, &g)
                : structure_grid(s, kAdaptiveGrid, p->hydration_grid_spacing_angstrom, 3.0, 12.0,
                                 &g));
    ps.site_xyz = g->sites;
    ps.site_volume = g->voxel_volume;
    ps.nsites = static_cast<int>(ps.site_xyz.size() / 3);
  }
  const bool with_sites = p && p->hydration_model != CSX_HYDRATION_ANALYTIC;
  ps.natoms = static_cast<int>(n);
  ps.ntypes = T + (with_sites ? 1 : 0);
  const int TT = ps.ntypes;
  const size_t npts = n + ps.nsites;
  if (npts > static_cast<size_t>(INT32_MAX / 2)) return fail(CSX_INVALID_INPUT, "too many points");

  // Type-sorted FP32 layout: a counting sort, stable, so atoms of one type
  // keep their input order.
  std::vector<size_t> order(n), next(T + 1, 0);
  for (size_t i = 0; i < n; ++i) ++next[s->type_of[i] + 1];
  for (int t = 0; t < T; ++t) next[t + 1] += next[t];
  for (size_t i = 0; i < n; ++i) order[next[s->type_of[i]]++] = i;
  std::vector<float> pts32(4 * npts);
  std::vector<int> seg_start(TT + 1, 0);
  ps.self.assign(TT, 0.0);
  ps.wsum.assign(TT, 0.0);
  for (size_t k = 0; k < n; ++k) {
    const size_t i = order[k];
    const int t = s->type_of[i];
    for (int a = 0; a < 3; ++a) pts32[4 * k + a] = static_cast<float>(s->xyz[3 * i + a] - s->centre[a]);
    pts32[4 * k + 3] = static_cast<float>(s->occ[i]);
    ps.occ_max = std::max(ps.occ_max, std::fabs(static_cast<double>(pts32[4 * k + 3])));
    ps.max_radius = std::max(ps.max_radius, std::sqrt(double(pts32[4 * k]) * pts32[4 * k] +
                                                      double(pts32[4 * k + 1]) * pts32[4 * k + 1] +
                                                      double(pts32[4 * k + 2]) * pts32[4 * k + 2]));
    seg_start[t + 1] += 1;
    ps.self[t] += s->occ[i] * s->occ[i];
    ps.wsum[t] += s->occ[i];
  }
  for (int k = 0; k < ps.nsites; ++k) {
    for (int a = 0; a < 3; ++a)
      pts32[4 * (n + k) + a] = static_cast<float>(ps.site_xyz[3 * k + a] - s->centre[a]);
    pts32[4 * (n + k) + 3] = 1.0f;
    ps.occ_max = std::max(ps.occ_max, 1.0);
    const float* p3 = &pts32[4 * (n + k)];
    ps.max_radius = std::max(ps.max_radius, std::sqrt(double(p3[0]) * p3[0] + double(p3[1]) * p3[1] +
                                                      double(p3[2]) * p3[2]));
  }
  if (with_sites) {
    seg_start[T + 1] = ps.nsites;
    ps.self[T] = ps.nsites;
    ps.wsum[T] = ps.nsites;
  }
  for (int t = 0; t < TT; ++t) seg_start[t + 1] += seg_start[t];

  ps.nbins = TT * (TT + 1) / 2;
  // The tables at q = 0 are the occupancy products sum_{i<j} occ_i occ_j.
  std::vector<double> q0;
  for (int a = 0; a < TT; ++a)
    for (int b = a; b < TT; ++b)
      q0.push_back(a == b ? 0.5 * (ps.wsum[a] * ps.wsum[a] - ps.self[a]) : ps.wsum[a] * ps.wsum[b]);

  // FP64 layout in original order.
  std::vector<double> pts64(4 * npts, 0.0), occ64(npts);
  std::vector<int> type64(npts);
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) pts64[4 * i + a] = s->xyz[3 * i + a];
    occ64[i] = s->occ[i];
    type64[i] = s->type_of[i];
  }
  for (int k = 0; k < ps.nsites; ++k) {
    for (int a = 0; a < 3; ++a) pts64[4 * (n + k) + a] = ps.site_xyz[3 * k + a];
    occ64[n + k] = 1.0;
    type64[n + k] = T;
  }
  std::vector<int> orig32(npts), type32(npts);
  for (size_t k = 0; k < n; ++k) {
    orig32[k] = static_cast<int>(order[k]);
    type32[k] = s->type_of[order[k]];
  }
  for (int k = 0; k < ps.nsites; ++k) {
    orig32[n + k] = static_cast<int>(n) + k;
    type32[n + k] = T;
  }
  CSX_HIP(ps.orig32.upload(orig32.data(), orig32.size() * sizeof(int)), "upload order");
  CSX_HIP(ps.type32.upload(type32.data(), type32.size() * sizeof(int)), "upload types");
  CSX_HIP(ps.seg32.upload(seg_start.data(), seg_start.size() * sizeof(int)), "upload segments");
  ps.h_seg = seg_start;
  CSX_HIP(ps.pts32.upload(pts32.data(), pts32.size() * sizeof(float)), "upload points");
  CSX_HIP(ps.q0.upload(q0.data(), q0.size() * sizeof(double)), "upload q = 0 tables");
  CSX_HIP(ps.pts64.upload(pts64.data(), pts64.size() * sizeof(double)), "upload points");
  CSX_HIP(ps.type64.upload(type64.data(), type64.size() * sizeof(int)), "upload types");
  CSX_HIP(ps.occ64.upload(occ64.data(), occ64.size() * sizeof(double)), "upload occupancy");
  return CSX_OK;
}

csx_status get_point_set(csx_structure* s, const csx_solvent_params* p,
                         std::shared_ptr<PointSet>* out) {
  std::lock_guard<std::mutex> lock(s->mu);
  const std::string key = set_key(p);
  auto it = s->sets.find(key);
  if (it != s->sets.end()) {
    *out = it->second;
    return CSX_OK;
  }
  auto ps = std::make_shared<PointSet>();
  CSX_TRY(build_point_set(s, p, *ps));
  s->sets[key] = ps;
  *out = std::move(ps);
  return CSX_OK;
}

// j points per work item: the tuned value, or automatically the largest of
// 512, 256, 128, 64 and 32 that still gives at least one full 256 x chunk
// item per CU (n^2 / 2 pairs over 256 x chunk per item); small sets take
// short j sweeps so that their work spreads over the device (1,200 points on
// 256 CUs: 32).  The chunk orders the FP64 sums, so a set's tables depend on
// it; the automatic value depends only on the set's point count and the CU
// count.
int pair_j_chunk(int tuned, const PointSet& ps, int cus) {
  if (tuned > 0) return tuned;
  const double n = static_cast<double>(ps.natoms + ps.nsites);
  for (int c : {512, 256, 128, 64})
    if (n * n / (2.0 * 256.0 * c) >= std::max(cus, 1)) return c;
  return 32;
}

csx_status get_item_set(PointSet& ps, int chunk, std::shared_ptr<ItemSet>* out) {
  std::lock_guard<std::mutex> lock(ps.items_mu);
  auto it = ps.item_sets.find(chunk);
  if (it != ps.item_sets.end()) {
    *out = it->second;
    return CSX_OK;
  }
  auto is = std::make_shared<ItemSet>();
  const int kI = 256;
  const int TT = ps.ntypes;
  const std::vector<int>& seg = ps.h_seg;
  std::vector<int>& items = is->h_items;
  std::vector<int>& bin_first = is->h_bin_first;
  for (int a = 0; a < TT; ++a)
    for (int b = a; b < TT; ++b) {
      bin_first.push_back(static_cast<int>(items.size() / 4));
      for (int i0 = seg[a]; i0 < seg[a + 1]; i0 += kI) {
        const int ilen = std::min(kI, seg[a + 1] - i0);
        const int jlo = std::max(seg[b], i0), jhi = seg[b + 1];
        for (int j0 = jlo; j0 < jhi; j0 += chunk)
          items.insert(items.end(), {i0, ilen, j0, std::min(chunk, jhi - j0)});
      }
    }
  bin_first.push_back(static_cast<int>(items.size() / 4));
  is->nitems = static_cast<int>(items.size() / 4);
  is->j_chunk = chunk;
  CSX_HIP(is->items.upload(items.data(), items.size() * sizeof(int)), "upload items");
  CSX_HIP(is->bin_first.upload(bin_first.data(), bin_first.size() * sizeof(int)), "upload bins");
  ps.item_sets[chunk] = is;
  *out = std::move(is);
  return CSX_OK;
}

// ---------------------------------------------------------- result cache --

std::shared_ptr<CacheEntry> cache_acquire(csx_ctx* ctx, PointSet& ps, const std::string& key,
                                          std::unique_lock<std::mutex>* lock) {
  size_t limit;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    limit = ctx->tuning.cache_bytes;
  }
  std::shared_ptr<CacheEntry> e;
  if (limit > 0) {
    std::lock_guard<std::mutex> g(ps.cache_mu);
    for (size_t k = 0; k < ps.cache.size(); ++k)
      if (ps.cache[k]->key == key) {
        e = ps.cache[k];
        ps.cache.erase(ps.cache.begin() + k);  // to the front: most recently used
        break;
      }
    if (!e) {
      e = std::make_shared<CacheEntry>();
      e->key = key;
      e->cached = true;
    }
    ps.cache.insert(ps.cache.begin(), e);
  } else {
    e = std::make_shared<CacheEntry>();
    e->key = key;
  }
  *lock = std::unique_lock<std::mutex>(e->mu);
  return e;
}

void cache_commit(csx_ctx* ctx, PointSet& ps, const std::shared_ptr<CacheEntry>& e) {
  size_t limit;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    limit = ctx->tuning.cache_bytes;
  }
  if (limit == 0 || !e->cached) return;
  std::lock_guard<std::mutex> g(ps.cache_mu);
  // The entry goes first (another call may have dropped it meanwhile); then
  // the most recently used ready entries are kept within the limit.  Entries
  // still being computed by other calls are kept; entries dropped here stay
  // alive for the calls that hold them.
  std::vector<std::shared_ptr<CacheEntry>> kept{e};
  size_t total = e->bytes;
  for (auto& c : ps.cache) {
    if (c == e) continue;
    if (!c->ready) {
      kept.push_back(c);
      continue;
    }
    if (total + c->bytes <= limit) {
      total += c->bytes;
      kept.push_back(c);
    }
  }
  ps.cache.swap(kept);
}

std::shared_ptr<CacheEntry> cache_find(csx_ctx* ctx, PointSet& ps, const std::string& key) {
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->tuning.cache_bytes == 0) return nullptr;
  }
  std::lock_guard<std::mutex> g(ps.cache_mu);
  for (size_t k = 0; k < ps.cache.size(); ++k)
    if (ps.cache[k]->key == key && ps.cache[k]->ready) {
      std::shared_ptr<CacheEntry> e = ps.cache[k];
      ps.cache.erase(ps.cache.begin() + k);
      ps.cache.insert(ps.cache.begin(), e);
      return e;
    }
  return nullptr;
}

void cache_store(csx_ctx* ctx, PointSet& ps, const std::string& key,
                 const std::vector<double>& tables) {
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->tuning.cache_bytes == 0) return;
  }
  {
    std::lock_guard<std::mutex> g(ps.cache_mu);
    for (auto& c : ps.cache)
      if (c->key == key) return;  // present, or being computed by another call
  }
  auto e = std::make_shared<CacheEntry>();
  e->key = key;
  e->cached = true;
  e->tables = tables;
  e->bytes = tables.size() * sizeof(double) + key.size();
  e->ready = true;
  {
    std::lock_guard<std::mutex> g(ps.cache_mu);
    ps.cache.insert(ps.cache.begin(), e);
  }
  cache_commit(ctx, ps, e);
}

void cache_drop(PointSet& ps, const std::shared_ptr<CacheEntry>& e) {
  std::lock_guard<std::mutex> g(ps.cache_mu);
  for (size_t k = 0; k < ps.cache.size(); ++k)
    if (ps.cache[k] == e) {
      ps.cache.erase(ps.cache.begin() + k);
      break;
    }
}

// --------------------------------------------------------- pair tables --

// Tiles instantiated for every sine mode, those of sine mode 2 only, and
// those of the recurrence (mode 3).
static const int kTilesAll[] = {8, 16, 32, 64};
static const int kTilesHw[] = {8, 10, 16, 20, 24, 25, 32, 40, 48, 50, 56, 63, 64};
static const int kTilesRec[] = {16, 25, 32, 50, 64, 100, 125, 128};

static bool tile_listed(const int* tiles, int n, int qt) {
  for (int k = 0; k < n; ++k)
    if (tiles[k] == qt) return true;
  return false;
}

// Modelled cost of a pair sweep over QT q values: per pair, the q loop plus
// the per-pair work (distance, refined square root, LDS read; for the
// recurrence also its start: two hardware sines and the step constant's
// series), in units of one q step of the mode.  Fitted on MI355X over
// globules of 300-20,000 atoms at 100, 200 and 500 q (kernel times of every
// tile): the recurrence's per-pair work is about 8 steps and its longest
// item weighs half a CU's share; the direct sweep's 12 evaluations and two
// CU shares.
static double sweep_cost(int sin_mode, int qt) { return sin_mode == 3 ? qt + 8.0 : qt + 12.0; }
static double tail_weight(int sin_mode) { return sin_mode == 3 ? 0.5 : 2.0; }

// The tile of a sweep mode with the least modelled time
//   sweep_cost(QT) (ceil(items ceil(nq / QT) / cus) sum_j / items + t max_j),
// sum_j and max_j the sum and the largest of the work items' j counts: the
// tasks (item, q chunk) of the busiest CU, each an average item long, plus
// the tail of the longest item (one workgroup sweeps an item's j atoms in
// turn, so a structure of a few thousand atoms, with few items, is bound by
// that length and takes narrow tiles; t = tail_weight).  Counting whole
// tasks per CU matters when there are few of them: with the direct sweep,
// 1,200 atoms x 100 q (122 items) take 2 tasks per CU with 25-wide tiles
// and 3 with 20-wide ones, 20.0 against 22.3 us.  Over the fitted sizes the
// recurrence's pick is within 2.4 % of its best tile on average (10 % at
// most).
// The tuned tile wins where it is instantiated for the mode.
static int choose_q_tile(int tuned, int sin_mode, size_t nq, size_t nitems, double sum_j,
                         double max_j, int cus, int max_tile = 1 << 30) {
  const int* tiles = sin_mode == 3 ? kTilesRec : sin_mode == 2 ? kTilesHw : kTilesAll;
  const int ntiles = sin_mode == 3 ? 8 : sin_mode == 2 ? 13 : 4;
  int best = tiles[0];
  double best_cost = 1e300;
  const double items = static_cast<double>(std::max<size_t>(nitems, 1));
  for (int k = 0; k < ntiles; ++k) {
    if (tiles[k] > max_tile && k > 0) continue;
    const double sweeps = static_cast<double>((nq + tiles[k] - 1) / tiles[k]);
    const double per_cu = std::ceil(items * sweeps / std::max(cus, 1));
    const double cost = sweep_cost(sin_mode, tiles[k]) *
                        (per_cu * sum_j / items + tail_weight(sin_mode) * max_j);
    if (tiles[k] == tuned) return tuned;
    if (cost < best_cost) best = tiles[k], best_cost = cost;
  }
  return best;
}

// The step of a uniform q grid in revolutions (q_k = q_0 + k step within 1e-9
// of the grid's largest |q|), as FP32, or 0 where the grid is not uniform,
// has fewer than 3 points, or a step whose phase advance over the largest
// pair distance (2 r_max) exceeds 1/4 revolution (the recurrence's step
// constant comes from a series valid up to there).  The recurrence then
// evaluates q_k = fp32(q_t / 2 pi) + (k - t) step per tile starting at t,
// within about 1e-7 of q_k relative: as close as the FP32 rounding of a
// direct sweep's q.
float uniform_q_step(const double* q, size_t nq, double rmax) {
  if (nq < 3) return 0.f;
  const double step = (q[nq - 1] - q[0]) / static_cast<double>(nq - 1);
  double qmax = 0.0;
  for (size_t k = 0; k < nq; ++k) qmax = std::max(qmax, std::fabs(q[k]));
  if (!(step != 0.0) || !std::isfinite(step)) return 0.f;
  for (size_t k = 0; k < nq; ++k)
    if (!(std::fabs(q[k] - (q[0] + static_cast<double>(k) * step)) <= 1e-9 * qmax)) return 0.f;
  const float qstep = static_cast<float>(step / (2.0 * kPi));
  if (!(std::fabs(static_cast<double>(qstep)) * 2.0 * rmax <= 0.25)) return 0.f;
  return qstep;
}

// Key of a point set's pair tables: they depend on the q grid, the sweep
// mode, the j chunk (which orders the FP64 sums) and, for the recurrence,
// the q tile (where it restarts); a direct sweep gives the same bits for
// every q tile and both launch forms.
static std::string table_key(int sin_mode, int qt, int j_chunk, const double* q, size_t nq) {
  std::string k = "pairs:" + std::to_string(sin_mode) + ":" + std::to_string(j_chunk) + ":";
  if (sin_mode == 3) k += "qt" + std::to_string(qt) + ":";
  k.append(reinterpret_cast<const char*>(q), nq * sizeof(double));
  return k;
}

// The sweep's sine mode for a q grid and the largest point radius, before
// the recurrence is considered: the tuned mode, except that the unreduced
// hardware sine (modes 2 and 3) takes |q r / 2 pi| <= 256 revolutions, so
// larger phases take mode 1.
static int pair_sin_mode(int tuned, const double* q, size_t nq, double rmax) {
  double umax = 0.0;
  for (size_t k = 0; k < nq; ++k)
    umax = std::max(umax, std::fabs(static_cast<double>(static_cast<float>(q[k] / (2.0 * kPi)))) *
                              2.0 * rmax);
  return tuned >= 2 && !(umax <= 250.0) ? 1 : tuned;
}

// The widest recurrence tile for a step (revolutions) and point radius: the
// recurrence's phase error grows with the phase it accumulates within a
// tile, (QT - 1) th for the largest pair (th = |step| 2 r_max): each step
// carries the FP32 rounding of th and of the step constant, where a direct
// sweep rounds each phase once.  Tiles are held to kRecTilePhase
// revolutions of it.  Measured on MI355X against FP64 GPU Debye: 20,000
// atoms x 500 q (2.8 revolutions at QT 125) 1.0e-6 max rel err (direct
// sweep 5.8e-7); 108,237 points x 200 q (r_max 100 A, 0.08 revolutions per
// step) 5.5e-6 at QT 100 (7.9 revolutions), 2.4e-6 at QT 50 (3.9), direct
// sweep 2.8e-6.  The 16-wide tile is always allowed (uniform_q_step keeps
// th <= 1/4, so 3.75 revolutions at most).
constexpr double kRecTilePhase = 4.0;
int rec_max_tile(float step, double rmax) {
  const double th = std::fabs(static_cast<double>(step)) * 2.0 * rmax;
  return th > 0.0 ? static_cast<int>(std::min(1e6, 1.0 + kRecTilePhase / th)) : 1 << 30;
}

// Sweep mode, q tile and recurrence step of a pair-table launch for one
// point set.  Mode 3 (the default) takes the recurrence wherever the grid
// allows it (uniform_q_step) and the call has at least kRecMinPairQ pair.q,
// and the direct hardware sine (mode 2) otherwise or when the tuned tile is
// one that only the direct sweep has.  On MI355X the recurrence's best tile
// beats the direct sweep's from 1,200 atoms x 100 q (7e7 pair.q) up by
// 1.3-1.9x; a call of a few million pair.q is one short task per CU, whose
// length the recurrence's per-pair start and dependent steps set (300 atoms
// x 100 q: 6.8 us direct, 9.2 us recurrence).  The plan follows from the q
// grid, the tuning and the set itself, so a set gets the same plan, and the
// same bits, alone and in a batch.
constexpr double kRecMinPairQ = 2e7;

struct PairPlan {
  int sin_mode = 2;
  int qt = 64;
  float qstep = 0.f;
  double h = 0.0;  // radial interpolation (mode 4): bin width and bins per table
  int nk = 0;
};

// Radial interpolation of the pair tables (sine mode 4): bins of width h,
// the largest power of two with h max|q| <= 1.04 (q the effective FP32 q of
// the table columns), where 8 Chebyshev nodes interpolate j_0(q r) to
// 1.0e-9; pair distances up to 2 r_max need floor(2 r_max / h) + 2 bins,
// whose LDS accumulators (64 B per bin) must fit a workgroup beside the
// 4 KB point tile.  Returns false where they do not.
// Radial bins of width h, the largest power of two (at most 64 A) with
// h max|q| <= 1.04 over the FP32 phase grid's q, covering distances up to
// 2 rmax: false where that takes more than max_bins bins.
bool radial_distance_bins(const double* q, size_t nq, double rmax, double max_bins, double* h,
                          int* nk) {
  double qmax = 0.0;
  for (size_t k = 0; k < nq; ++k)
    qmax = std::max(qmax, 2.0 * kPi * std::fabs(static_cast<double>(static_cast<float>(q[k] / (2.0 * kPi)))));
  *h = qmax > 0.0 ? std::min(64.0, std::exp2(std::floor(std::log2(1.04 / qmax)))) : 64.0;
  const double bins = std::floor(2.0 * rmax / *h) + 2.0;
  if (!(*h >= 1.0 / 4096.0) || !(bins <= max_bins)) return false;
  *nk = static_cast<int>(bins);
  return true;
}

// The radial pair sweep's bins: one accumulator copy of them beside the
// 4 KB point tile in a workgroup's LDS.
static bool radial_bins(const double* q, size_t nq, double rmax, size_t lds, double* h, int* nk) {
  return radial_distance_bins(q, nq, rmax,
                              (static_cast<double>(lds) - 4104.0) /
                                  static_cast<double>(radial_lds_bytes(1)),
                              h, nk);
}

// Fixed-point exponent of a point set's radial accumulators: one
// contribution occ_i occ_j T_p(u) (|T_p| <= 1 on the bin, with a margin of
// 1.3 for u rounded past its ends) times 2^F stays below 2^30, and a table
// bin's sum over all the set's pairs below 2^62.
static int radial_scale_exponent(const PointSet& ps) {
  const double cmax = std::max(ps.occ_max * ps.occ_max * 1.3, 1e-300);
  const double n = static_cast<double>(ps.natoms + ps.nsites);
  const double pairs = std::max(0.5 * n * (n - 1.0), 1.0);
  const int f1 = static_cast<int>(std::floor(30.0 - std::log2(cmax)));
  const int f2 = static_cast<int>(std::floor(62.0 - std::log2(pairs * cmax)));
  return std::max(-60, std::min({f1, f2, 60}));
}

static PairPlan pair_plan(const Tuning& tune, const double* q, size_t nq, double rmax,
                          double npts, size_t nitems, double sum_j, double max_j, int cus,
                          size_t lds) {
  PairPlan plan;
  if (tune.sin_mode == 4) {
    if (radial_bins(q, nq, rmax, lds, &plan.h, &plan.nk)) {
      plan.sin_mode = 4;
      plan.qt = 1;
      return plan;
    }
    Tuning direct = tune;
    direct.sin_mode = 2;
    return pair_plan(direct, q, nq, rmax, npts, nitems, sum_j, max_j, cus, lds);
  }
  plan.sin_mode = pair_sin_mode(tune.sin_mode, q, nq, rmax);
  if (plan.sin_mode == 3) {
    const float step = uniform_q_step(q, nq, rmax);
    const bool direct_only = tune.q_tile != 0 && !tile_listed(kTilesRec, 8, tune.q_tile) &&
                             tile_listed(kTilesHw, 13, tune.q_tile);
    const bool tiny = tune.q_tile == 0 && 0.5 * npts * npts * static_cast<double>(nq) < kRecMinPairQ;
    if (step != 0.f && !direct_only && !tiny) {
      plan.qt = choose_q_tile(tune.q_tile, 3, nq, nitems, sum_j, max_j, cus,
                              rec_max_tile(step, rmax));
      plan.qstep = step;
      return plan;
    }
    plan.sin_mode = 2;
  }
  plan.qt = choose_q_tile(tune.q_tile, plan.sin_mode, nq, nitems, sum_j, max_j, cus);
  return plan;
}

// The q (revolutions) of table column k that the sweep evaluated (the
// device's finish_tables computes the same).
static double plan_q_rev(const PairPlan& plan, const std::vector<float>& qrev, size_t k) {
  if (plan.qstep == 0.f) return static_cast<double>(qrev[k]);
  const size_t kt = k % static_cast<size_t>(plan.qt);
  return static_cast<double>(qrev[k - kt]) +
         static_cast<double>(kt) * static_cast<double>(plan.qstep);
}

// Automatic launch form: the persistent grid pays (4-8 % on large calls)
// once every one of its workgroups takes at least two tasks; with fewer
// tasks one workgroup per task starts sooner.
static int pair_persist_wgs(int tuned, size_t nitems, size_t nq_pad, int qt, int cus) {
  const double tasks = static_cast<double>(nitems) * static_cast<double>(nq_pad / qt);
  return tuned > 0 ? tuned * cus
         : tuned < 0 && tasks >= 2.0 * kAutoPersist * cus ? kAutoPersist * cus
                                                          : 0;
}

// Pair tables of one or more point sets.  Tables already cached on a set
// for the same q grid and arithmetic are reused; the others come from one
// launch of the pair kernel (a batch concatenates the sets' points, work
// items and bins) and are cached:
// P[c][b*nq + k] = sum_{i<j in bin b of set c} occ_i occ_j sinc(q_k r_ij).
// One set waits for a concurrent call computing the same tables; in a batch,
// tables that another call is still computing are computed again.
// Sum and largest of the work items' j counts of some item sets.
static void item_stats(const std::vector<std::shared_ptr<ItemSet>>& isets, size_t* nitems,
                       double* sum_j, double* max_j) {
  *nitems = 0;
  *sum_j = 0.0;
  *max_j = 0.0;
  for (const auto& is : isets) {
    *nitems += static_cast<size_t>(is->nitems);
    for (size_t k = 3; k < is->h_items.size(); k += 4) {
      *sum_j += is->h_items[k];
      *max_j = std::max(*max_j, static_cast<double>(is->h_items[k]));
    }
  }
}

// One launch of the pair kernel over the sets `idx` (their points, work
// items and bins concatenated) with one plan; fills P[c] for c in idx.
// `overlap`, when given, runs on the host while the kernel runs.
static csx_status pair_tables_launch(Call& call, const std::vector<PointSet*>& sets,
                                     const std::vector<size_t>& idx,
                                     const std::vector<std::shared_ptr<ItemSet>>& all_items,
                                     const std::vector<int>& chunk, const PairPlan& plan,
                                     const Tuning& tune, const double* q, size_t nq,
                                     std::vector<std::vector<double>>& P,
                                     const std::function<void()>& overlap) {
  const int cus = call.ctx()->cu_count;
  const int sin_mode = plan.sin_mode, qt = plan.qt;
  std::vector<size_t> todo = idx;
  std::vector<std::shared_ptr<ItemSet>> isets(todo.size());
  size_t npts = 0, nitems = 0, nbins = 0;
  for (size_t t = 0; t < todo.size(); ++t) {
    PointSet* ps = sets[todo[t]];
    isets[t] = all_items[todo[t]];
    npts += static_cast<size_t>(ps->natoms + ps->nsites);
    nitems += static_cast<size_t>(isets[t]->nitems);
    nbins += static_cast<size_t>(ps->nbins);
  }
  if (npts > static_cast<size_t>(INT32_MAX / 2) || nitems > static_cast<size_t>(INT32_MAX / 2) ||
      nbins > static_cast<size_t>(INT32_MAX / 2))
    return fail(CSX_INVALID_INPUT, "batch too large for one launch");
  const size_t nq_pad = (nq + qt - 1) / qt * qt;
  std::vector<float> qrev(nq_pad, 0.0f);
  for (size_t k = 0; k < nq; ++k) qrev[k] = static_cast<float>(q[k] / (2.0 * kPi));
  const hipStream_t st = call.stream();
  Scratch d_q(call), d_partial(call), d_out(call), d_counter(call), d_pts(call), d_items(call),
      d_bins(call);
  CSX_HIP(d_q.alloc(nq_pad * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_partial.alloc(sin_mode == 4 ? size_t(nbins) * plan.nk * radial_node_count() *
                                              sizeof(unsigned long long)
                                        : nitems * nq_pad * sizeof(double),
                          st),
          "hipMallocAsync");
  CSX_HIP(d_out.alloc(nbins * nq_pad * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_counter.alloc(sizeof(int), st), "hipMallocAsync");
  // q and the tables go through the stream's pinned staging, so both copies
  // are asynchronous.
  void* h_q = nullptr;
  void* h_out = nullptr;
  CSX_HIP(call.take_pinned(0, nq_pad * sizeof(float), &h_q), "hipHostMalloc");
  CSX_HIP(call.take_pinned(1, nbins * nq_pad * sizeof(double), &h_out), "hipHostMalloc");
  std::memcpy(h_q, qrev.data(), nq_pad * sizeof(float));
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), h_q, nq_pad * sizeof(float), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  const float4* pts = nullptr;
  const int4* items = nullptr;
  const int* bins = nullptr;
  std::vector<int> h_items, h_bins;
  if (todo.size() == 1) {
    pts = sets[todo[0]]->pts32.as<float4>();
    items = isets[0]->items.as<int4>();
    bins = isets[0]->bin_first.as<int>();
  } else {
    CSX_HIP(d_pts.alloc(npts * sizeof(float4), st), "hipMallocAsync");
    h_items.reserve(4 * nitems);
    h_bins.reserve(nbins + 1);
    size_t off = 0, item_off = 0;
    for (size_t t = 0; t < todo.size(); ++t) {
      const PointSet* ps = sets[todo[t]];
      const ItemSet& is = *isets[t];
      const size_t np = static_cast<size_t>(ps->natoms + ps->nsites);
      CSX_HIP(hipMemcpyAsync(d_pts.as<float4>() + off, ps->pts32.as<void>(), np * sizeof(float4),
                             hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync");
      for (size_t k = 0; k < is.h_items.size(); k += 4) {
        h_items.push_back(is.h_items[k] + static_cast<int>(off));
        h_items.push_back(is.h_items[k + 1]);
        h_items.push_back(is.h_items[k + 2] + static_cast<int>(off));
        h_items.push_back(is.h_items[k + 3]);
      }
      for (int b = 0; b < ps->nbins; ++b)
        h_bins.push_back(is.h_bin_first[b] + static_cast<int>(item_off));
      off += np;
      item_off += static_cast<size_t>(is.nitems);
    }
    h_bins.push_back(static_cast<int>(item_off));
    CSX_HIP(d_items.alloc(h_items.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(d_bins.alloc(h_bins.size() * sizeof(int), st), "hipMallocAsync");
    CSX_HIP(hipMemcpyAsync(d_items.as<void>(), h_items.data(), h_items.size() * sizeof(int),
                           hipMemcpyHostToDevice, st),
            "hipMemcpyAsync");
    CSX_HIP(hipMemcpyAsync(d_bins.as<void>(), h_bins.data(), h_bins.size() * sizeof(int),
                           hipMemcpyHostToDevice, st),
            "hipMemcpyAsync");
    pts = d_pts.as<float4>();
    items = d_items.as<int4>();
    bins = d_bins.as<int>();
  }
  const int persist_wgs = sin_mode == 4 ? 0 : pair_persist_wgs(tune.persist, nitems, nq_pad, qt, cus);
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    call.ctx()->last_launch[0] = qt;
    call.ctx()->last_launch[1] = sin_mode;
    call.ctx()->last_launch[2] = persist_wgs;
    call.ctx()->last_items[0] = chunk[todo[0]];
    call.ctx()->last_items[1] = static_cast<int>(nitems);
  }
  if (sin_mode == 4) {
    // Per table bin, the fixed-point scale of its point set.
    const size_t sbytes = nbins * (sizeof(float) + sizeof(double));
    void* h_s = nullptr;
    CSX_HIP(call.take_pinned(3, sbytes, &h_s), "hipHostMalloc");
    double* unscale = static_cast<double*>(h_s);
    float* scale = reinterpret_cast<float*>(unscale + nbins);
    size_t b = 0;
    for (size_t t = 0; t < todo.size(); ++t) {
      const PointSet& ps = *sets[todo[t]];
      const int f = radial_scale_exponent(ps);
      for (int k = 0; k < ps.nbins; ++k, ++b) {
        scale[b] = std::ldexp(1.0f, f);
        unscale[b] = std::ldexp(1.0, -f);
      }
    }
    Scratch d_scale(call);
    CSX_HIP(d_scale.alloc(sbytes, st), "hipMallocAsync");
    CSX_HIP(hipMemcpyAsync(d_scale.as<void>(), h_s, sbytes, hipMemcpyHostToDevice, st),
            "hipMemcpyAsync");
    const double* d_unscale = d_scale.as<double>();
    const float* d_bscale = reinterpret_cast<const float*>(d_unscale + nbins);
    CSX_HIP(launch_debye_pairs_radial(pts, items, static_cast<int>(nitems), bins,
                                      static_cast<int>(nbins), d_bscale, d_unscale, plan.h,
                                      plan.nk, d_q.as<float>(),
                                      static_cast<int>(nq_pad), tune.q_tile,
                                      d_partial.as<unsigned long long>(), d_out.as<double>(),
                                      call.ctx()->lds_per_workgroup, st, call.timer()),
            "debye_pairs_radial");
  } else {
    CSX_HIP(launch_debye_pairs(pts, items, static_cast<int>(nitems), bins, static_cast<int>(nbins),
                               d_q.as<float>(), plan.qstep, static_cast<int>(nq_pad), qt,
                               sin_mode, persist_wgs, d_counter.as<int>(),
                               d_partial.as<double>(), d_out.as<double>(), st, call.timer()),
            "debye_pairs_f32");
  }
  CSX_HIP(hipMemcpyAsync(h_out, d_out.as<void>(), nbins * nq_pad * sizeof(double),
                         hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  if (overlap) overlap();
  d_q.reset();
  d_partial.reset();
  d_out.reset();
  d_counter.reset();
  d_pts.reset();
  d_items.reset();
  d_bins.reset();
  CSX_TRY(call.finish());
  const double* raw = static_cast<const double*>(h_out);
  size_t bin_off = 0;
  for (size_t t = 0; t < todo.size(); ++t) {
    const PointSet& ps = *sets[todo[t]];
    std::vector<double>& Pc = P[todo[t]];
    Pc.assign(size_t(ps.nbins) * nq, 0.0);
    const int TT = ps.ntypes;
    int b = 0;
    for (int a = 0; a < TT; ++a)
      for (int u = a; u < TT; ++u, ++b)
        for (size_t k = 0; k < nq; ++k) {
          const double qe = 2.0 * kPi * plan_q_rev(plan, qrev, k);
          if (qe == 0.0) {
            Pc[b * nq + k] = a == u ? 0.5 * (ps.wsum[a] * ps.wsum[a] - ps.self[a])
                                    : ps.wsum[a] * ps.wsum[u];
          } else {
            Pc[b * nq + k] = raw[(bin_off + b) * nq_pad + k] / qe;
          }
        }
    bin_off += static_cast<size_t>(ps.nbins);
  }
  return CSX_OK;
}

// Pair tables of one or more point sets.  Each set has its own launch plan
// (sweep mode and q tile, from its own work items), so its tables are the
// same whether it is computed alone or in a batch; tables already cached on
// a set for the same q grid and plan are reused.  The others come from one
// launch of the pair kernel per plan among them (a launch concatenates the
// sets' points, work items and bins; sets on the direct sweep, whose bits
// do not depend on the tile, share one launch) and are cached:
// P[c][b*nq + k] = sum_{i<j in bin b of set c} occ_i occ_j sinc(q_k r_ij).
// One set waits for a concurrent call computing the same tables; in a batch,
// tables that another call is still computing are computed again.
// `overlap`, when given, runs on the host while the (first) pair kernel
// runs, or at once when nothing is launched.
static csx_status pair_tables_batch(Call& call, const std::vector<PointSet*>& sets,
                                    const double* q, size_t nq,
                                    std::vector<std::vector<double>>& P,
                                    const std::function<void()>& overlap = nullptr) {
  Tuning tune;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    tune = call.ctx()->tuning;
  }
  const int cus = call.ctx()->cu_count;
  // Work items of every set (made when the structure is created, or once
  // per j chunk) and the set's plan.
  std::vector<int> chunk(sets.size());
  std::vector<std::shared_ptr<ItemSet>> all_items(sets.size());
  std::vector<PairPlan> plans(sets.size());
  std::vector<std::string> keys(sets.size());
  for (size_t c = 0; c < sets.size(); ++c) {
    chunk[c] = pair_j_chunk(tune.j_chunk, *sets[c], cus);
    CSX_TRY(get_item_set(*sets[c], chunk[c], &all_items[c]));
    size_t ni;
    double sum_j, max_j;
    item_stats({all_items[c]}, &ni, &sum_j, &max_j);
    plans[c] = pair_plan(tune, q, nq, sets[c]->max_radius,
                         static_cast<double>(sets[c]->natoms + sets[c]->nsites), ni, sum_j, max_j,
                         cus, call.ctx()->lds_per_workgroup);
    // the radial sweep's integer sums do not depend on the chunk
    keys[c] = table_key(plans[c].sin_mode, plans[c].qt, plans[c].sin_mode == 4 ? 0 : chunk[c], q,
                        nq);
  }
  // The radial sweep's work items: 256 x 256 by default (on 20,000 atoms
  // x 500 q, 0.61 ms against 0.65 at the pair sweep's 512 and 0.80 at
  // 1024).  Its tables do not depend on the chunk.
  for (size_t c = 0; c < sets.size(); ++c)
    if (plans[c].sin_mode == 4 && tune.j_chunk == 0 && chunk[c] > 256) {
      chunk[c] = 256;
      CSX_TRY(get_item_set(*sets[c], chunk[c], &all_items[c]));
    }
  P.assign(sets.size(), std::vector<double>());
  std::vector<size_t> todo;
  std::shared_ptr<CacheEntry> entry;  // the one set's entry, locked while computed
  std::unique_lock<std::mutex> entry_lock;
  if (sets.size() == 1) {
    entry = cache_acquire(call.ctx(), *sets[0], keys[0], &entry_lock);
    if (entry->ready) {
      P[0] = entry->tables;
      if (overlap) overlap();
      return CSX_OK;
    }
    todo.push_back(0);
  } else {
    for (size_t c = 0; c < sets.size(); ++c) {
      std::shared_ptr<CacheEntry> e = cache_find(call.ctx(), *sets[c], keys[c]);
      if (e) P[c] = e->tables;
      else todo.push_back(c);
    }
    if (todo.empty()) {
      if (overlap) overlap();
      return CSX_OK;
    }
  }
  // A failed computation leaves no half-made entry behind.
  struct Dropper {
    PointSet* ps;
    std::shared_ptr<CacheEntry>* e;
    ~Dropper() {
      if (*e && !(*e)->ready) cache_drop(*ps, *e);
    }
  } dropper{sets[0], &entry};

  // Groups of sets with one plan; direct sweeps of one sine mode share a
  // launch, whose tile follows the group's work items.
  std::vector<std::vector<size_t>> groups;
  std::vector<PairPlan> gplans;
  for (size_t c : todo) {
    size_t g = 0;
    for (; g < groups.size(); ++g) {
      const PairPlan& a = gplans[g];
      const PairPlan& b = plans[c];
      if (a.sin_mode == b.sin_mode && (b.sin_mode != 3 || (a.qt == b.qt && a.qstep == b.qstep)))
        break;
    }
    if (g == groups.size()) {
      groups.emplace_back();
      gplans.push_back(plans[c]);
    }
    groups[g].push_back(c);
  }
  for (size_t g = 0; g < groups.size(); ++g) {
    if (gplans[g].sin_mode == 4) {
      for (size_t c : groups[g]) gplans[g].nk = std::max(gplans[g].nk, plans[c].nk);
    } else if (gplans[g].sin_mode != 3) {
      std::vector<std::shared_ptr<ItemSet>> gi;
      for (size_t c : groups[g]) gi.push_back(all_items[c]);
      size_t ni;
      double sum_j, max_j;
      item_stats(gi, &ni, &sum_j, &max_j);
      gplans[g].qt = choose_q_tile(tune.q_tile, gplans[g].sin_mode, nq, ni, sum_j, max_j, cus);
    }
    CSX_TRY(pair_tables_launch(call, sets, groups[g], all_items, chunk, gplans[g], tune, q, nq, P,
                               g == 0 ? overlap : nullptr));
  }
  if (entry) {
    entry->tables = P[0];
    entry->bytes = P[0].size() * sizeof(double) + keys[0].size();
    entry->ready = true;
    cache_commit(call.ctx(), *sets[0], entry);
  } else {
    for (size_t t = 0; t < todo.size(); ++t)
      cache_store(call.ctx(), *sets[todo[t]], keys[todo[t]], P[todo[t]]);
  }
  return CSX_OK;
}

static csx_status pair_tables_f32(Call& call, PointSet& ps, const double* q, size_t nq,
                                  std::vector<double>& P,
                                  const std::function<void()>& overlap = nullptr) {
  std::vector<std::vector<double>> all;
  CSX_TRY(pair_tables_batch(call, {&ps}, q, nq, all, overlap));
  P = std::move(all[0]);
  return CSX_OK;
}

// out[k] = sum_a X_a Y_a self_a + sum_{a<=b} (X_a Y_b + X_b Y_a) P_ab.
static void contract(const PointSet& ps, const std::vector<double>& P, size_t nq,
                     const std::vector<double>& X, const std::vector<double>& Y, double* out) {
  const int TT = ps.ntypes;
  for (size_t k = 0; k < nq; ++k) {
    double sum = 0.0;
    for (int a = 0; a < TT; ++a) sum += X[a * nq + k] * Y[a * nq + k] * ps.self[a];
    int b = 0;
    for (int a = 0; a < TT; ++a)
      for (int c = a; c < TT; ++c, ++b)
        sum += (X[a * nq + k] * Y[c * nq + k] + X[c * nq + k] * Y[a * nq + k]) * P[b * nq + k];
    out[k] = sum;
  }
}

// Per-type Debye amplitudes [TT][nq] (v, e, h); the site type carries only
// the water amplitude of solvent.rs:560-562 in h.
static void type_amplitudes(const csx_structure* s, const PointSet& ps, const double* q, size_t nq,
                            bool implicit_h, const csx_solvent_params* p, std::vector<double>& V,
                            std::vector<double>& E, std::vector<double>& H) {
  const int T = static_cast<int>(s->types.size());
  const int TT = ps.ntypes;
  V.assign(size_t(TT) * nq, 0.0);
  E.assign(size_t(TT) * nq, 0.0);
  H.assign(size_t(TT) * nq, 0.0);
  const bool grid = p && p->hydration_model != CSX_HYDRATION_ANALYTIC;
  for (int t = 0; t < T; ++t)
    for (size_t k = 0; k < nq; ++k) {
      if (!p) {
        V[t * nq + k] = type_form_factor(s->types[t], implicit_h, q[k]);
        continue;
      }
      const Amplitudes a = debye_amplitudes(s->types[t], implicit_h, q[k],
                                            p->shell_thickness_angstrom, p->bulk_density_e_per_a3);
      V[t * nq + k] = a.v;
      E[t * nq + k] = a.e;
      H[t * nq + k] = grid ? 0.0 : a.h;
    }
  if (grid)
    for (size_t k = 0; k < nq; ++k)
      H[T * nq + k] = p->bulk_density_e_per_a3 / 10.0 * ps.site_volume * water_form_factor(q[k]);
}

// FP64 plain Debye over per-point amplitudes: 1 channel -> I, 3 -> 6 terms.
static csx_status plain_f64(Call& call, const PointSet& ps, const double* q, size_t nq, int nch,
                            const std::vector<double>* tables[3], double* out) {
  const hipStream_t st = call.stream();
  const int npts = nch == 1 ? ps.natoms : ps.natoms + ps.nsites;
  const int TT = ps.ntypes;
  const int nterms = nch == 1 ? 1 : 6;
  if (npts == 0) {
    std::fill(out, out + nterms * nq, 0.0);
    return CSX_OK;
  }
  std::vector<double> table(size_t(nch) * TT * nq);
  for (int c = 0; c < nch; ++c)
    std::copy(tables[c]->begin(), tables[c]->end(), table.begin() + size_t(c) * TT * nq);
  const int nblk = (npts + 255) / 256;
  Scratch d_table(call), d_amp(call), d_q(call), d_partial(call), d_out(call);
  CSX_HIP(d_table.alloc(table.size() * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_amp.alloc(size_t(nch) * nq * npts * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_q.alloc(nq * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_partial.alloc(size_t(nq) * nblk * nterms * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(size_t(nterms) * nq * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_table.as<void>(), table.data(), table.size() * sizeof(double),
                         hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), q, nq * sizeof(double), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(launch_expand_amplitudes(d_table.as<double>(), ps.type64.as<int>(), ps.occ64.as<double>(),
                                   npts, TT, static_cast<int>(nq), nch, d_amp.as<double>(), st),
          "expand_amplitudes");
  CSX_HIP(launch_debye_plain(ps.pts64.as<double4>(), d_amp.as<double>(), d_q.as<double>(), npts,
                             static_cast<int>(nq), nch, nterms, d_partial.as<double>(),
                             d_out.as<double>(), st, call.timer()),
          "debye_plain_f64");
  CSX_HIP(hipMemcpyAsync(out, d_out.as<void>(), size_t(nterms) * nq * sizeof(double),
                         hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  d_table.reset();
  d_amp.reset();
  d_q.reset();
  d_partial.reset();
  d_out.reset();
  return call.finish();
}

static csx_status check_structure(const csx_structure* s, const void* q, size_t nq,
                                  const void* out) {
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (nq > 0 && (!q || !out)) return fail(CSX_INVALID_INPUT, "q and output buffers are required");
  if (nq > static_cast<size_t>(INT32_MAX / 64)) return fail(CSX_INVALID_INPUT, "too many q values");
  return CSX_OK;
}

static csx_status debye_basis(csx_structure* s, const double* q, size_t nq,
                              const csx_solvent_params* p, bool implicit_h, int precision,
                              double* out) {
  Call call(s->ctx);
  CSX_TRY(call.begin());
  std::shared_ptr<PointSet> ps;
  CSX_TRY(get_point_set(s, p, &ps));
  std::vector<double> V, E, H;
  auto amplitudes = [&] { type_amplitudes(s, *ps, q, nq, implicit_h, p, V, E, H); };
  if (precision == CSX_FP64_PLAIN) {
    amplitudes();
    const std::vector<double>* tabs[3] = {&V, &E, &H};
    return plain_f64(call, *ps, q, nq, 3, tabs, out);
  }
  std::vector<double> P;
  CSX_TRY(pair_tables_f32(call, *ps, q, nq, P, amplitudes));
  contract(*ps, P, nq, V, V, out + 0 * nq);
  contract(*ps, P, nq, E, E, out + 1 * nq);
  contract(*ps, P, nq, H, H, out + 2 * nq);
  contract(*ps, P, nq, V, E, out + 3 * nq);
  contract(*ps, P, nq, V, H, out + 4 * nq);
  contract(*ps, P, nq, E, H, out + 5 * nq);
  return CSX_OK;
}

static void combine(const double* basis, const double* q, size_t nq, double c1, double c2,
                    double* out) {
  for (size_t i = 0; i < nq; ++i) {
    const double g = excluded_volume_scale(q[i], c1);
    const double v = basis[i] + g * g * basis[nq + i] + c2 * c2 * basis[2 * nq + i] -
                     2.0 * g * basis[3 * nq + i] + 2.0 * c2 * basis[4 * nq + i] -
                     2.0 * g * c2 * basis[5 * nq + i];
    out[i] = std::max(v, 0.0);
  }
}

static bool precision_ok(int precision) {
  return precision == CSX_FP64_PLAIN || precision == CSX_FP32_COMPENSATED;
}

}  // namespace csx

using namespace csx;

// ================================================================ C ABI ==

extern "C" {

const char* csx_last_error(void) { return g_error.c_str(); }

void csx_solvent_default(csx_solvent_params* out) {
  if (!out) return;
  out->c1 = 1.0;
  out->c2 = 0.0;
  out->shell_thickness_angstrom = 3.0;
  out->bulk_density_e_per_a3 = 0.334;
  out->hydration_model = CSX_HYDRATION_ANALYTIC;
  out->hydration_grid_spacing_angstrom = 4.0;
}

void csx_compute_options_default(csx_compute_options* out) {
  if (!out) return;
  out->method = CSX_METHOD_AUTO;
  out->implicit_hydrogen = 1;
  out->has_solvent = 0;
  csx_solvent_default(&out->solvent);
  out->l_max = 0;
  out->precision = CSX_FP32_COMPENSATED;
}

int csx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

csx_status csx_ctx_create(int device, csx_ctx** out) {
  CSX_API_RANGE();
  if (!out) return fail(CSX_INVALID_INPUT, "output pointer is null");
  *out = nullptr;
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(CSX_NO_DEVICE, "no HIP device is available");
  if (device < 0 || device >= n) return fail(CSX_NO_DEVICE, "device index out of range");
  csx_ctx* ctx = new (std::nothrow) csx_ctx;
  if (!ctx) return fail(CSX_OUT_OF_MEMORY, "context allocation failed");
  ctx->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    ctx->cu_count = cus;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds > 0)
    ctx->lds_per_workgroup = static_cast<size_t>(lds);
  // The no-allocation entry point's device scratch, located once per device
  // with the device current; a failure here disables that entry point only.
  hipError_t e2 = hipSetDevice(device);
  if (e2 == hipSuccess) e2 = noalloc_scratch_setup(device);
  ctx->scratch_status = static_cast<int>(e2);
  *out = ctx;
  return CSX_OK;
}

void csx_ctx_destroy(csx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (hipStream_t s : ctx->free_streams) (void)hipStreamDestroy(s);
  for (auto& entry : ctx->scratch)
    for (auto& slot : entry.second)
      if (slot.first) device_free(slot.first, slot.second);
  for (auto& entry : ctx->pinned)
    for (auto& slot : entry.second)
      if (slot.first) (void)hipHostFree(slot.first);
  delete ctx;
}

csx_status csx_ctx_set_timing(csx_ctx* ctx, int enabled) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  ctx->timing.store(enabled != 0);
  return CSX_OK;
}

csx_status csx_ctx_kernel_time(csx_ctx* ctx, const char* name, int64_t* launches, double* total_ms) {
  if (!ctx || !name) return fail(CSX_INVALID_INPUT, "context and name are required");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->stats.find(name);
  if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.second;
  return CSX_OK;
}

csx_status csx_ctx_reset_timing(csx_ctx* ctx) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->stats.clear();
  return CSX_OK;
}

csx_status csx_ctx_set_tuning(csx_ctx* ctx, int q_tile, int sin_mode, int j_chunk) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  const bool tile_ok = q_tile == 100 || q_tile == 125 || q_tile == 128 ||
                       (q_tile >= 0 && q_tile <= 64 &&
                        (q_tile % 8 == 0 || q_tile == 10 || q_tile == 20 || q_tile == 25 ||
                         q_tile == 50 || q_tile == 63));
  if (!tile_ok)
    return fail(CSX_INVALID_INPUT,
                "q_tile must be 0 (automatic), a multiple of 8 up to 64, 10, 20, 25, 50, 63, "
                "100, 125 or 128");
  if (sin_mode < 0 || sin_mode > 4)
    return fail(CSX_INVALID_INPUT, "sin_mode must be 0, 1, 2, 3 or 4");
  if (sin_mode == 4 && !(q_tile == 0 || q_tile == 8 || q_tile == 16 || q_tile == 32))
    return fail(CSX_INVALID_INPUT,
                "with sin_mode 4, q_tile sets the LDS accumulator copies: 0 (automatic), 8, 16 "
                "or 32");
  if (!(j_chunk == 0 || j_chunk == 32 || j_chunk == 64 || j_chunk == 128 ||
        (j_chunk >= 256 && j_chunk % 256 == 0 && j_chunk <= (1 << 20))))
    return fail(CSX_INVALID_INPUT,
                "j_chunk must be 0 (automatic), 32, 64, 128 or a multiple of 256 in [256, 2^20]");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.q_tile = q_tile;
  ctx->tuning.sin_mode = sin_mode;
  ctx->tuning.j_chunk = j_chunk;
  return CSX_OK;
}

csx_status csx_ctx_set_cache_limit(csx_ctx* ctx, size_t bytes) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.cache_bytes = bytes;
  return CSX_OK;
}

csx_status csx_ctx_set_workspace_limit(csx_ctx* ctx, size_t bytes) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (bytes < (size_t(1) << 20)) return fail(CSX_INVALID_INPUT, "workspace limit must be >= 1 MiB");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.workspace_limit = bytes;
  return CSX_OK;
}

csx_status csx_ctx_set_vjp_tile(csx_ctx* ctx, int q_tile) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (q_tile != 0 && q_tile != 16 && q_tile != 32 && q_tile != 48 && q_tile != 60 && q_tile != 64)
    return fail(CSX_INVALID_INPUT, "vjp q_tile must be 0 (automatic), 16, 32, 48, 60 or 64");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.vjp_q_tile = q_tile;
  return CSX_OK;
}

csx_status csx_ctx_set_multipole_group(csx_ctx* ctx, int tiles) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (tiles != 0 && tiles != 32 && tiles != 48 && tiles != 64 && tiles != 128)
    return fail(CSX_INVALID_INPUT,
                "multipole tile group must be 0 (automatic), 32, 48, 64 or 128");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.mp_group_tiles = tiles;
  return CSX_OK;
}

csx_status csx_ctx_set_multipole_radial(csx_ctx* ctx, int mode) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (mode < 0 || mode > 2)
    return fail(CSX_INVALID_INPUT,
                "multipole radial mode must be 0 (automatic), 1 (per point) or 2 (nodes)");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.mp_radial = mode;
  return CSX_OK;
}

csx_status csx_ctx_last_multipole(csx_ctx* ctx, double* bin_width, int* points) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (bin_width) *bin_width = ctx->last_mp_bin;
  if (points) *points = ctx->last_mp_points;
  return CSX_OK;
}

csx_status csx_ctx_set_vjp_split(csx_ctx* ctx, int split) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (split != 0 && split != 1 && split != 2 && split != 4 && split != 8)
    return fail(CSX_INVALID_INPUT, "vjp split must be 0 (automatic), 1, 2, 4 or 8");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.vjp_split = split;
  return CSX_OK;
}

csx_status csx_ctx_last_vjp_launch(csx_ctx* ctx, int* q_tile, int* q_split, int* step_split,
                                   int* launches) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (q_tile) *q_tile = ctx->last_vjp[0];
  if (q_split) *q_split = ctx->last_vjp[1];
  if (step_split) *step_split = ctx->last_vjp[2];
  if (launches) *launches = ctx->last_vjp[3];
  return CSX_OK;
}

csx_status csx_ctx_last_pair_items(csx_ctx* ctx, int* j_chunk, int* items) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (j_chunk) *j_chunk = ctx->last_items[0];
  if (items) *items = ctx->last_items[1];
  return CSX_OK;
}

csx_status csx_ctx_last_launch(csx_ctx* ctx, int* q_tile, int* sin_mode, int* persistent_wgs) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (q_tile) *q_tile = ctx->last_launch[0];
  if (sin_mode) *sin_mode = ctx->last_launch[1];
  if (persistent_wgs) *persistent_wgs = ctx->last_launch[2];
  return CSX_OK;
}

int csx_element_from_symbol(const char* symbol) {
  return symbol ? element_from_symbol(symbol) : CSX_EL_OTHER;
}

int csx_united_atom_group(const char* residue, const char* atom_name) {
  if (!residue || !atom_name) return -1;
  return united_atom_group(residue, atom_name);
}

double csx_form_factor(int element, int group, double occupancy, int implicit_h, double q) {
  if (element < 0 || element >= CSX_NUM_ELEMENTS || group < -1 || group >= CSX_NUM_UA_GROUPS)
    return 0.0;
  return occupancy * type_form_factor({element, group}, implicit_h != 0, q);
}

double csx_excluded_volume(int element, int group, int implicit_h) {
  if (element < 0 || element >= CSX_NUM_ELEMENTS || group < -1 || group >= CSX_NUM_UA_GROUPS)
    return 0.0;
  return type_excluded_volume({element, group}, implicit_h != 0);
}

double csx_water_form_factor(double q) { return water_form_factor(q); }

double csx_excluded_volume_scale(double q, double radius_scale) {
  return excluded_volume_scale(q, radius_scale);
}

csx_status csx_structure_create(csx_ctx* ctx, size_t n, const double* xyz, const uint8_t* element,
                                const int8_t* group, const double* occupancy,
                                csx_structure** out) {
  CSX_API_RANGE();
  if (!ctx || !out) return fail(CSX_INVALID_INPUT, "context and output pointer are required");
  *out = nullptr;
  if (n > 0 && (!xyz || !element || !occupancy))
    return fail(CSX_INVALID_INPUT, "xyz, element and occupancy arrays are required");
  if (n > static_cast<size_t>(INT32_MAX / 4)) return fail(CSX_INVALID_INPUT, "too many atoms");
  std::unique_ptr<csx_structure> s(new (std::nothrow) csx_structure);
  if (!s) return fail(CSX_OUT_OF_MEMORY, "structure allocation failed");
  s->ctx = ctx;
  s->device = ctx->device;
  s->n = n;
  s->xyz.assign(xyz, xyz + 3 * n);
  s->occ.assign(occupancy, occupancy + n);
  s->element.assign(element, element + n);
  s->group.assign(n, -1);
  if (group) s->group.assign(group, group + n);
  for (size_t i = 0; i < n; ++i) {
    if (s->element[i] >= CSX_NUM_ELEMENTS) return fail(CSX_INVALID_INPUT, "element code out of range");
    if (s->group[i] < -1 || s->group[i] >= CSX_NUM_UA_GROUPS)
      return fail(CSX_INVALID_INPUT, "united-atom group out of range");
  }
  // Type partition: the (element, group) keys present, in sorted order,
  // from a table over the small key range (element, then group from -1).
  constexpr int kGroups = CSX_NUM_UA_GROUPS + 1;
  std::vector<int> slot(static_cast<size_t>(CSX_NUM_ELEMENTS) * kGroups, -1);
  auto key_of = [&](size_t i) { return size_t(s->element[i]) * kGroups + (s->group[i] + 1); };
  for (size_t i = 0; i < n; ++i) slot[key_of(i)] = 0;
  for (int e = 0; e < CSX_NUM_ELEMENTS; ++e)
    for (int g = 0; g < kGroups; ++g)
      if (slot[size_t(e) * kGroups + g] == 0) {
        slot[size_t(e) * kGroups + g] = static_cast<int>(s->types.size()) + 1;
        s->types.push_back({e, g - 1});
      }
  s->type_of.resize(n);
  for (size_t i = 0; i < n; ++i) s->type_of[i] = slot[key_of(i)] - 1;
  // Centroid and maximum radius, as compute_curve does (scattering.rs:157-167).
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) s->centre[a] += s->xyz[3 * i + a];
  for (int a = 0; a < 3; ++a) s->centre[a] /= static_cast<double>(std::max<size_t>(n, 1));
  for (size_t i = 0; i < n; ++i) {
    const double dx = s->xyz[3 * i] - s->centre[0], dy = s->xyz[3 * i + 1] - s->centre[1],
                 dz = s->xyz[3 * i + 2] - s->centre[2];
    s->max_radius = std::max(s->max_radius, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  const hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  std::shared_ptr<PointSet> ps;
  CSX_TRY(get_point_set(s.get(), nullptr, &ps));
  int j_chunk;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    j_chunk = ctx->tuning.j_chunk;
  }
  std::shared_ptr<ItemSet> is;
  CSX_TRY(get_item_set(*ps, pair_j_chunk(j_chunk, *ps, ctx->cu_count), &is));
  *out = s.release();
  return CSX_OK;
}

void csx_structure_destroy(csx_structure* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  delete s;
}

size_t csx_structure_atom_count(const csx_structure* s) { return s ? s->n : 0; }

csx_status csx_structure_set_hydration_sites(csx_structure* s, const double* xyz, size_t count,
                                             double voxel_volume) {
  CSX_API_RANGE();
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (count > 0 && !xyz) return fail(CSX_INVALID_INPUT, "site coordinates are required");
  if (!std::isfinite(voxel_volume) || voxel_volume < 0.0)
    return fail(CSX_INVALID_INPUT, "voxel volume must be finite and non-negative");
  if (count > static_cast<size_t>(INT32_MAX / 4)) return fail(CSX_INVALID_INPUT, "too many sites");
  for (size_t k = 0; k < 3 * count; ++k)
    if (!std::isfinite(xyz[k])) return fail(CSX_INVALID_INPUT, "site coordinates must be finite");
  std::lock_guard<std::mutex> lock(s->mu);
  s->has_sites = true;
  s->sites.assign(xyz, xyz + 3 * count);
  s->site_volume = voxel_volume;
  // calls already holding the previous point set finish with it
  s->sets.erase("sites");
  return CSX_OK;
}

static csx_status copy_grid(const Grid& g, double* sites, size_t capacity, size_t* count,
                            double* volume) {
  if (!g.error.empty()) return fail(CSX_INVALID_INPUT, g.error);
  const size_t ns = g.sites.size() / 3;
  if (count) *count = ns;
  if (volume) *volume = g.voxel_volume;
  if (sites && capacity) std::copy(g.sites.begin(), g.sites.begin() + 3 * std::min(capacity, ns), sites);
  return CSX_OK;
}

csx_status csx_generate_shell_sites(const double* xyz, size_t n, double spacing, double inner,
                                    double outer, double* sites, size_t capacity, size_t* count) {
  CSX_API_RANGE();
  if (n > 0 && !xyz) return fail(CSX_INVALID_INPUT, "xyz is required");
  return copy_grid(shell_sites(xyz, n, spacing, inner, outer), sites, capacity, count, nullptr);
}

csx_status csx_generate_globule(size_t n, double radius, double elongation, uint64_t seed,
                                double* xyz, uint8_t* element) {
  CSX_API_RANGE();
  if (n > 0 && (!xyz || !element)) return fail(CSX_INVALID_INPUT, "output arrays are required");
  if (!(radius > 0.0) || !(elongation > 0.0) || !std::isfinite(radius) || !std::isfinite(elongation))
    return fail(CSX_INVALID_INPUT, "radius and elongation must be finite and positive");
  generate_globule(n, radius, elongation, seed, xyz, element);
  return CSX_OK;
}

csx_status csx_generate_domain_chain(size_t n, int domains, uint64_t seed, double* xyz,
                                     uint8_t* element) {
  CSX_API_RANGE();
  if (n > 0 && (!xyz || !element)) return fail(CSX_INVALID_INPUT, "output arrays are required");
  if (domains < 1 || domains > 16) return fail(CSX_INVALID_INPUT, "domains must be in [1, 16]");
  generate_domain_chain(n, domains, seed, xyz, element);
  return CSX_OK;
}

csx_status csx_debye_intensity(csx_structure* s, const double* q, size_t nq, int implicit_h,
                               int precision, double* out) {
  CSX_API_RANGE();
  CSX_TRY(check_structure(s, q, nq, out));
  if (!precision_ok(precision)) return fail(CSX_INVALID_INPUT, "unknown precision");
  if (nq == 0) return CSX_OK;
  Call call(s->ctx);
  CSX_TRY(call.begin());
  std::shared_ptr<PointSet> ps;
  CSX_TRY(get_point_set(s, nullptr, &ps));
  std::vector<double> V, E, H;
  auto amplitudes = [&] { type_amplitudes(s, *ps, q, nq, implicit_h != 0, nullptr, V, E, H); };
  if (precision == CSX_FP64_PLAIN) {
    amplitudes();
    const std::vector<double>* tabs[3] = {&V, &V, &V};
    CSX_TRY(plain_f64(call, *ps, q, nq, 1, tabs, out));
  } else {
    std::vector<double> P;
    CSX_TRY(pair_tables_f32(call, *ps, q, nq, P, amplitudes));
    contract(*ps, P, nq, V, V, out);
  }
  for (size_t k = 0; k < nq; ++k) out[k] = std::max(out[k], 0.0);
  return CSX_OK;
}

csx_status csx_debye_solvent_basis(csx_structure* s, const double* q, size_t nq,
                                   const csx_solvent_params* p, int implicit_h, int precision,
                                   double* out) {
  CSX_API_RANGE();
  CSX_TRY(check_structure(s, q, nq, out));
  if (!p) return fail(CSX_INVALID_INPUT, "solvent parameters are required");
  if (!precision_ok(precision)) return fail(CSX_INVALID_INPUT, "unknown precision");
  if (nq == 0) return CSX_OK;
  return debye_basis(s, q, nq, p, implicit_h != 0, precision, out);
}

csx_status csx_solvent_corrected_intensity(csx_structure* s, const double* q, size_t nq,
                                           const csx_solvent_params* p, int implicit_h,
                                           int precision, double* out) {
  CSX_API_RANGE();
  CSX_TRY(check_structure(s, q, nq, out));
  if (!p) return fail(CSX_INVALID_INPUT, "solvent parameters are required");
  if (!precision_ok(precision)) return fail(CSX_INVALID_INPUT, "unknown precision");
  if (nq == 0) return CSX_OK;
  std::vector<double> basis(6 * nq);
  CSX_TRY(debye_basis(s, q, nq, p, implicit_h != 0, precision, basis.data()));
  combine(basis.data(), q, nq, p->c1, p->c2, out);
  return CSX_OK;
}

csx_status csx_debye_pair_tables(csx_structure* s, const double* q, size_t nq,
                                 const csx_solvent_params* p, double* out, double* self_terms,
                                 size_t* n_types) {
  CSX_API_RANGE();
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (nq > 0 && !q) return fail(CSX_INVALID_INPUT, "q is required");
  std::shared_ptr<PointSet> ps;
  {
    const hipError_t e = hipSetDevice(s->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  }
  CSX_TRY(get_point_set(s, p, &ps));
  if (n_types) *n_types = static_cast<size_t>(ps->ntypes);
  if (self_terms) std::copy(ps->self.begin(), ps->self.end(), self_terms);
  if (!out || nq == 0) return CSX_OK;
  Call call(s->ctx);
  CSX_TRY(call.begin());
  std::vector<double> P;
  CSX_TRY(pair_tables_f32(call, *ps, q, nq, P));
  std::copy(P.begin(), P.end(), out);
  return CSX_OK;
}

csx_status csx_debye_pair_tables_noalloc(csx_structure* s, const double* q, size_t nq,
                                         double* tables, void* stream) {
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (nq == 0) return CSX_OK;
  if (!q || !tables) return fail(CSX_INVALID_INPUT, "q and the device table buffer are required");
  if (nq > static_cast<size_t>(INT32_MAX / 64)) return fail(CSX_INVALID_INPUT, "too many q values");
  csx_ctx* ctx = s->ctx;
  CSX_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  Tuning tune;
  int scratch_status;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    tune = ctx->tuning;
    scratch_status = ctx->scratch_status;
  }
  // the radial sweep's accumulators are not part of the fixed scratch: the
  // direct sweep computes these tables
  if (tune.sin_mode == 4) tune.sin_mode = 2;
  size_t cap = 0;
  unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(ctx->device, &cap));
  if (!base)
    return fail(CSX_DEVICE_ERROR,
                std::string("the no-allocation scratch is not available on this device: ") +
                    hipGetErrorString(static_cast<hipError_t>(scratch_status)));
  // The atoms' point set and work items are made when the structure is
  // created (and for another j chunk by the first ordinary call with it).
  std::shared_ptr<PointSet> ps;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->sets.find("atoms");
    if (it != s->sets.end()) ps = it->second;
  }
  if (!ps) return fail(CSX_INVALID_INPUT, "the structure has no device point set");
  std::shared_ptr<ItemSet> is;
  {
    std::lock_guard<std::mutex> lock(ps->items_mu);
    auto it = ps->item_sets.find(pair_j_chunk(tune.j_chunk, *ps, ctx->cu_count));
    if (it != ps->item_sets.end()) is = it->second;
  }
  if (!is)
    return fail(CSX_INVALID_INPUT,
                "no work items for the context's j chunk on this structure: call "
                "csx_debye_pair_tables on it once after changing the j chunk");
  double sum_j = 0.0, max_j = 0.0;
  for (size_t k = 3; k < is->h_items.size(); k += 4) {
    sum_j += is->h_items[k];
    max_j = std::max(max_j, static_cast<double>(is->h_items[k]));
  }
  const int cus = ctx->cu_count;
  const PairPlan plan = pair_plan(tune, q, nq, ps->max_radius,
                                  static_cast<double>(ps->natoms + ps->nsites),
                                  static_cast<size_t>(is->nitems), sum_j, max_j, cus,
                                  ctx->lds_per_workgroup);
  const int sin_mode = plan.sin_mode, qt = plan.qt;
  const size_t nq_pad = (nq + qt - 1) / qt * qt;
  const size_t nitems = static_cast<size_t>(is->nitems), nbins = static_cast<size_t>(ps->nbins);
  const int persist_wgs = pair_persist_wgs(tune.persist, nitems, nq_pad, qt, cus);
  // scratch: task counter | q (revolutions) | item partials | raw tables
  auto round256 = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t off_q = 256;
  const size_t off_partial = off_q + round256(nq_pad * sizeof(float));
  const size_t off_raw = off_partial + round256(nitems * nq_pad * sizeof(double));
  const size_t need = off_raw + round256(nbins * nq_pad * sizeof(double));
  if (need > cap) {
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "this call needs %zu bytes of device scratch; the no-allocation entry point "
                  "has %zu (use csx_debye_pair_tables for larger problems)",
                  need, cap);
    return fail(CSX_INVALID_INPUT, msg);
  }
  int* counter = reinterpret_cast<int*>(base);
  float* qrev = reinterpret_cast<float*>(base + off_q);
  double* partial = reinterpret_cast<double*>(base + off_partial);
  double* raw = reinterpret_cast<double*>(base + off_raw);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_launch[0] = qt;
    ctx->last_launch[1] = sin_mode;
    ctx->last_launch[2] = persist_wgs;
  }
  CSX_HIP(launch_write_floats(qrev, q, static_cast<int>(nq), static_cast<int>(nq_pad), 2.0 * kPi,
                              st),
          "write_floats");
  CSX_HIP(launch_debye_pairs(ps->pts32.as<float4>(), is->items.as<int4>(), static_cast<int>(nitems),
                             is->bin_first.as<int>(), static_cast<int>(nbins), qrev, plan.qstep,
                             static_cast<int>(nq_pad), qt, sin_mode, persist_wgs, counter, partial,
                             raw, st, nullptr),
          "debye_pairs_f32");
  CSX_HIP(launch_finish_tables(raw, qrev, qt, plan.qstep, ps->q0.as<double>(),
                               static_cast<int>(nbins),
                               static_cast<int>(nq), static_cast<int>(nq_pad), tables, st),
          "finish_tables");
  return CSX_OK;
}

// Batched Debye profiles: one launch of the pair kernel over every
// conformer's work items.  p == NULL: vacuum intensities (out: m*nq);
// otherwise six-term bases (out: m*6*nq).
static csx_status batch_debye(csx_structure* const* structures, size_t m, const double* q,
                              size_t nq, const csx_solvent_params* p, bool implicit_h,
                              double* out) {
  if (m == 0) return CSX_OK;
  if (!structures || (nq > 0 && (!q || !out)))
    return fail(CSX_INVALID_INPUT, "structures, q and output buffers are required");
  csx_ctx* ctx = structures[0] ? structures[0]->ctx : nullptr;
  for (size_t c = 0; c < m; ++c)
    if (!structures[c] || structures[c]->ctx != ctx)
      return fail(CSX_INVALID_INPUT, "all structures must be non-null and share one context");
  if (nq == 0) return CSX_OK;
  Call call(ctx);
  CSX_TRY(call.begin());
  std::vector<std::shared_ptr<PointSet>> held(m);
  std::vector<PointSet*> sets(m);
  for (size_t c = 0; c < m; ++c) {
    CSX_TRY(get_point_set(structures[c], p, &held[c]));
    sets[c] = held[c].get();
  }
  std::vector<std::vector<double>> P, V(m), E(m), H(m);
  CSX_TRY(pair_tables_batch(call, sets, q, nq, P, [&] {
    for (size_t c = 0; c < m; ++c)
      type_amplitudes(structures[c], *sets[c], q, nq, implicit_h, p, V[c], E[c], H[c]);
  }));
  for (size_t c = 0; c < m; ++c) {
    if (!p) {
      double* o = out + c * nq;
      contract(*sets[c], P[c], nq, V[c], V[c], o);
      for (size_t k = 0; k < nq; ++k) o[k] = std::max(o[k], 0.0);
      continue;
    }
    double* o = out + c * 6 * nq;
    contract(*sets[c], P[c], nq, V[c], V[c], o + 0 * nq);
    contract(*sets[c], P[c], nq, E[c], E[c], o + 1 * nq);
    contract(*sets[c], P[c], nq, H[c], H[c], o + 2 * nq);
    contract(*sets[c], P[c], nq, V[c], E[c], o + 3 * nq);
    contract(*sets[c], P[c], nq, V[c], H[c], o + 4 * nq);
    contract(*sets[c], P[c], nq, E[c], H[c], o + 5 * nq);
  }
  return CSX_OK;
}

csx_status csx_batch_debye_intensity(csx_structure* const* structures, size_t m, const double* q,
                                     size_t nq, int implicit_h, double* out) {
  CSX_API_RANGE();
  return batch_debye(structures, m, q, nq, nullptr, implicit_h != 0, out);
}

csx_status csx_batch_debye_solvent_basis(csx_structure* const* structures, size_t m,
                                         const double* q, size_t nq, const csx_solvent_params* p,
                                         int implicit_h, double* out) {
  CSX_API_RANGE();
  if (!p) return fail(CSX_INVALID_INPUT, "solvent parameters are required");
  return batch_debye(structures, m, q, nq, p, implicit_h != 0, out);
}

csx_status csx_ctx_device_info(csx_ctx* ctx, int* cu_count, char* arch, size_t arch_len) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (cu_count) *cu_count = ctx->cu_count;
  if (arch && arch_len) {
    hipDeviceProp_t prop;
    const hipError_t e = hipGetDeviceProperties(&prop, ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    std::snprintf(arch, arch_len, "%s", prop.gcnArchName);
  }
  return CSX_OK;
}

csx_status csx_ctx_set_persistence(csx_ctx* ctx, int workgroups_per_cu) {
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (workgroups_per_cu < -1 || workgroups_per_cu > 64)
    return fail(CSX_INVALID_INPUT, "workgroups_per_cu must be -1 (automatic) or in [0, 64]");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning.persist = workgroups_per_cu;
  return CSX_OK;
}

csx_status csx_combine_solvent_basis(const double* basis, const double* q, size_t nq, double c1,
                                     double c2, double* out) {
  if (nq > 0 && (!basis || !q || !out)) return fail(CSX_INVALID_INPUT, "buffers are required");
  combine(basis, q, nq, c1, c2, out);
  return CSX_OK;
}

csx_status csx_compute_curve(csx_structure* s, const double* q, size_t nq,
                             const csx_compute_options* o, double* out, int* method_used) {
  CSX_API_RANGE();
  CSX_TRY(check_structure(s, q, nq, out));
  if (!o) return fail(CSX_INVALID_INPUT, "options are required");
  for (size_t k = 0; k < nq; ++k)
    if (!std::isfinite(q[k]) || q[k] < 0.0 || (k > 0 && q[k] < q[k - 1]))
      return fail(CSX_INVALID_INPUT, "q values must be finite, non-negative, and non-decreasing");
  const csx_solvent_params* p = o->has_solvent ? &o->solvent : nullptr;
  if (p) {
    const bool ok = std::isfinite(p->c1) && p->c1 > 0.0 && std::isfinite(p->c2) &&
                    std::isfinite(p->shell_thickness_angstrom) &&
                    p->shell_thickness_angstrom > 0.0 &&
                    std::isfinite(p->bulk_density_e_per_a3) && p->bulk_density_e_per_a3 > 0.0 &&
                    std::isfinite(p->hydration_grid_spacing_angstrom) &&
                    p->hydration_grid_spacing_angstrom > 0.0;
    if (!ok)
      return fail(CSX_INVALID_INPUT,
                  "solvent parameters must be finite with positive dimensions, density, and c1");
    if (p->hydration_model < 0 || p->hydration_model > CSX_HYDRATION_EXPLICIT_SITES)
      return fail(CSX_INVALID_INPUT, "unknown hydration model");
  }
  if (o->l_max != 0 && (o->l_max < 8 || o->l_max > 64))
    return fail(CSX_INVALID_INPUT, "multipole l_max must be between 8 and 64");
  if (!precision_ok(o->precision)) return fail(CSX_INVALID_INPUT, "unknown precision");
  int method = o->method;
  if (method == CSX_METHOD_AUTO) {
    if (p && p->hydration_model != CSX_HYDRATION_ANALYTIC) method = CSX_METHOD_MULTIPOLE;
    else if (s->n < 500) method = CSX_METHOD_DEBYE;
    else method = CSX_METHOD_MULTIPOLE;
  }
  if (method != CSX_METHOD_DEBYE && method != CSX_METHOD_MULTIPOLE)
    return fail(CSX_INVALID_INPUT, "unknown scattering method");
  if (method_used) *method_used = method;
  if (method == CSX_METHOD_DEBYE) {
    if (!p) return csx_debye_intensity(s, q, nq, o->implicit_hydrogen, o->precision, out);
    return csx_solvent_corrected_intensity(s, q, nq, p, o->implicit_hydrogen, o->precision, out);
  }
  double qmax = 0.0;
  for (size_t k = 0; k < nq; ++k) qmax = std::max(qmax, q[k]);
  const double order = std::ceil(qmax * s->max_radius);
  const int l_max = o->l_max != 0 ? o->l_max
                                  : static_cast<int>(std::min(std::max(order, 0.0) + 8.0, 64.0) < 8.0
                                                         ? 8.0
                                                         : std::min(std::max(order, 0.0) + 8.0, 64.0));
  return multipole_dispatch(s, q, nq, l_max, o->implicit_hydrogen != 0, p, o->precision, out);
}

}  // extern "C"

// Internal state behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/crabsaxs.h"
#include "api_range.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace csx {

// Thread-local error message + status helper.
csx_status fail(csx_status st, const std::string& msg);
csx_status hip_fail(hipError_t e, const char* what);

// compute_curve's multipole branch (scattering.rs:183-201).
csx_status multipole_dispatch(csx_structure* s, const double* q, size_t nq, int l_max,
                              bool implicit_h, const csx_solvent_params* p, int precision,
                              double* out);

// Workgroups per CU of the automatic persistent launch.
constexpr int kAutoPersist = 4;

// Launch shape of debye_pairs_f32.  Defaults are the fastest shape measured
// on MI355X for a 20,000-atom x 500-q profile (all shapes meet the same
// accuracy); the autotuner can replace them per workload.
struct Tuning {
  int q_tile = 0;      // q values per sweep of a pair (0: automatic)
  int sin_mode = 3;    // 0 software polynomial, 1 reduced + V_SIN_F32, 2 V_SIN_F32,
                       // 3 sine recurrence on uniform q grids where it models faster, else 2
  int j_chunk = 0;     // j atoms per work item (0: automatic, from the point count)
  int persist = -1;    // persistent launch with this many workgroups per CU (0: grid;
                       // -1: automatic, kAutoPersist per CU when tasks are plenty)
  size_t workspace_limit = size_t(2) << 30;  // bytes of adjoint partials per launch
  int vjp_q_tile = 0;  // q values per sweep of the coordinate adjoint (0: automatic)
  int vjp_split = 0;   // workgroups per sweep of the coordinate adjoint (0: automatic)
  int mp_group_tiles = 0;  // column tiles per multipole tile group (0: automatic)
  int mp_radial = 0;  // multipole radial interpolation: 0 automatic, 1 per point, 2 always
  size_t cache_bytes = size_t(256) << 20;  // cached results kept per point set (0: off)
};

}  // namespace csx

struct csx_ctx {
  int device = 0;
  std::mutex mu;
  std::vector<hipStream_t> free_streams;
  std::atomic<bool> timing{false};
  std::map<std::string, std::pair<int64_t, double>> stats;
  csx::Tuning tuning;
  int cu_count = 256;
  size_t lds_per_workgroup = 64 * 1024;  // the device's per-workgroup LDS limit
  int last_launch[3] = {0, 0, 0};  // q tile, sin mode, persistent workgroups
  int last_items[2] = {0, 0};      // j chunk and work items of the last pair-table launch
  int scratch_status = 0;          // hipError_t of the no-allocation scratch set-up
  int last_vjp[4] = {0, 0, 0, 0};  // q tile, q split, step split, launches
  double last_mp_bin = 0.0;        // radial bin width of the last multipole geometry used
  int last_mp_points = 0;          // points (or nodes) it contracts
  // Grow-only device buffers kept per stream (see csx::Scratch), and
  // grow-only pinned host staging buffers per stream; freed with the context.
  std::map<hipStream_t, std::vector<std::pair<void*, size_t>>> scratch, pinned;
};

namespace csx {

// Every device allocation of the library goes through these, which count
// the bytes it holds.  With CSX_DEVICE_MEMORY_LIMIT=<bytes> in the
// environment (read once), an allocation that would take that count above
// the limit fails with hipErrorOutOfMemory (CSX_OUT_OF_MEMORY to the
// caller) before reaching HIP: out-of-memory paths can then be tested on a
// device with plenty of memory.
hipError_t device_alloc(void** p, size_t bytes);
hipError_t device_alloc_async(void** p, size_t bytes, hipStream_t s);
void device_free(void* p, size_t bytes);
void device_free_async(void* p, size_t bytes, hipStream_t s);

// Device buffer freed on its stream (stream-ordered allocator).
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  hipError_t alloc(size_t bytes, hipStream_t s) {
    reset();
    stream_ = s;
    bytes_ = bytes == 0 ? 16 : bytes;
    const hipError_t e = device_alloc_async(&p_, bytes_, s);
    if (e != hipSuccess) p_ = nullptr;
    return e;
  }
  void reset() {
    if (p_) device_free_async(p_, bytes_, stream_);
    p_ = nullptr;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
  hipStream_t stream_ = nullptr;
};

// Persistent device allocation owned by a structure.
class DevArray {
 public:
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() {
    if (p_) device_free(p_, bytes_);
  }
  hipError_t alloc(size_t bytes) {
    if (p_) device_free(p_, bytes_);
    p_ = nullptr;
    bytes_ = bytes == 0 ? 16 : bytes;
    const hipError_t e = device_alloc(&p_, bytes_);
    if (e != hipSuccess) p_ = nullptr;
    return e;
  }
  hipError_t upload(const void* host, size_t bytes) {
    hipError_t e = alloc(bytes);
    if (e != hipSuccess) return e;
    return bytes ? hipMemcpy(p_, host, bytes, hipMemcpyHostToDevice) : hipSuccess;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// Multipole geometry of one point set at one l_max: the points in segments
// by type (atoms, then sites), each ordered by radius and padded to whole
// slabs, and the FP32 real spherical harmonics in MFMA B-operand tiles; or,
// with radial interpolation, per type the Chebyshev nodes of each radial
// bin as virtual points and their harmonic moments in the same tiles.
struct MultipoleGeometry {
  int L = -1, ntiles = 0, npad = 0, npts = 0, nchunk = 0;
  int aexp = 14;  // exponent of the j_l occ operand scale of the coefficient kernel
  int yexp = 12;  // exponent of the harmonic (or moment) operand scale
  double rmax_atoms = 0.0, rmax_sites = 0.0;
  // Radial interpolation: bin width (0: the points themselves) and the
  // number of virtual points (kNodes per nonempty bin of each type).
  double bin_width = 0.0;
  int nvirtual = 0;
  // Per padded point (or virtual point): radius, occupancy (0 for padding);
  // chunks: (first slab, end slab, type) of each one-type chunk of the
  // coefficient kernel.
  DevArray radius, occ, chunks, Y, tile_l;
  std::vector<int> h_tile_l;      // l of each column tile
  std::vector<int> h_type_chunk;  // per type t: its chunks are [h_type_chunk[t], [t + 1])
  DevArray type_chunk;
};

// Work items of the FP32 pair kernel over one point set for one j chunk:
// type-pure blocks of 256 i points x chunks of j points over the upper
// triangle, bin-major over the T(T+1)/2 type pairs (a <= b).
struct ItemSet {
  DevArray items;       // int4[nitems]: i0, ilen, j0, jlen
  DevArray bin_first;   // int[nbins + 1]
  std::vector<int> h_items, h_bin_first;  // host copies (batch concatenation)
  int nitems = 0, j_chunk = 0;
};

// One cached result of a point set: the pair tables of a q grid (host,
// FP64 [nbins][nq]) or the per-type multipole coefficients of a q grid and
// l_max (device, FP64 [ntypes][nqp][ncol], with the per-q orders).  `key`
// holds the exact inputs the result depends on.  The entry's mutex is held
// while it is computed, so concurrent callers of one key wait for the first.
struct CacheEntry {
  std::string key;
  std::mutex mu;
  std::atomic<bool> ready{false};
  bool cached = false;  // created in the point set's cache (else private to one call)
  size_t bytes = 0;
  std::vector<double> tables;
  // Multipole coefficients and per-q orders: in coef / qlmax for a cached
  // entry, in the computing call's scratch for a private one.
  DevArray coef, qlmax;
  double* coef_ptr = nullptr;
  int* qlmax_ptr = nullptr;
  int nqp = 0, ncol = 0;
};

// One set of scattering points (atoms, optionally followed by hydration
// sites as an extra type) laid out for both Debye kernels.
struct PointSet {
  std::mutex geom_mu;
  // by (l_max, radial bin width (0: per point), radial mode)
  std::map<std::tuple<int, double, int>, std::unique_ptr<MultipoleGeometry>> geoms;
  int ntypes = 0;         // atom types (+1 when sites are present)
  int natoms = 0, nsites = 0;
  int nbins = 0;          // ntypes (ntypes + 1) / 2
  double site_volume = 0.0;
  std::vector<double> self;   // per type: sum occ^2
  std::vector<double> wsum;   // per type: sum occ
  DevArray q0;                // double[nbins]: the pair tables at q = 0
  // FP32 pair-table layout: type-sorted, centred, w = occupancy.
  DevArray pts32;             // float4[natoms + nsites]
  double max_radius = 0.0;    // of the FP32 points about the centroid
  double occ_max = 0.0;       // largest |occupancy| of the FP32 points
  DevArray orig32;            // int[npts]: original index of each sorted point
  DevArray type32;            // int[npts]: type of each sorted point
  DevArray seg32;             // int[ntypes + 1]: type segments of the sorted order
  std::vector<int> h_seg;     // host copy of the type segments
  // FP64 plain layout: original order, atoms then sites.
  DevArray pts64;             // double4[natoms + nsites]
  DevArray type64;            // int[natoms + nsites]
  DevArray occ64;             // double[natoms + nsites]
  std::vector<double> site_xyz;
  // Work items by j chunk, built on first use.
  std::mutex items_mu;
  std::map<int, std::shared_ptr<ItemSet>> item_sets;
  // Cached results, most recently used first.
  std::mutex cache_mu;
  std::vector<std::shared_ptr<CacheEntry>> cache;
};

}  // namespace csx

struct csx_structure {
  csx_ctx* ctx = nullptr;
  int device = 0;  // the context's device; destroy needs no live context
  size_t n = 0;
  std::vector<double> xyz, occ;
  std::vector<uint8_t> element;
  std::vector<int8_t> group;
  std::vector<csx::TypeKey> types;  // sorted unique keys
  std::vector<int> type_of;         // per atom
  double centre[3] = {0, 0, 0};
  double max_radius = 0.0;          // from the centroid
  std::mutex mu;
  // Point sets by solvent model; a caller holds its set for the call, so a
  // set replaced meanwhile (new hydration sites) outlives the calls using it.
  std::map<std::string, std::shared_ptr<csx::PointSet>> sets;
  std::map<std::string, csx::Grid> grids;  // hydration grids by kind and options
  // Hydration sites given by the caller (CSX_HYDRATION_EXPLICIT_SITES).
  bool has_sites = false;
  std::vector<double> sites;
  double site_volume = 0.0;
};

namespace csx {

// Scope of one API call on a device: current device, a leased stream, an
// optional kernel timer merged into the context statistics at the end.
class Call {
 public:
  explicit Call(csx_ctx* ctx) : ctx_(ctx) {}
  ~Call() {
    if (stream_) {
      // A stream goes back to the pool idle (a call that returned early may
      // have left work queued), so the next call may overwrite its staging.
      if (hipStreamQuery(stream_) == hipErrorNotReady) (void)hipStreamSynchronize(stream_);
      std::lock_guard<std::mutex> lock(ctx_->mu);
      ctx_->free_streams.push_back(stream_);
    }
  }
  csx_status begin() {
    hipError_t e = hipSetDevice(ctx_->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    {
      std::lock_guard<std::mutex> lock(ctx_->mu);
      if (!ctx_->free_streams.empty()) {
        stream_ = ctx_->free_streams.back();
        ctx_->free_streams.pop_back();
      }
    }
    if (!stream_) {
      e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
      if (e != hipSuccess) {
        stream_ = nullptr;
        return hip_fail(e, "hipStreamCreate");
      }
    }
    if (ctx_->timing.load()) timer_.reset(new KernelTimer);
    return CSX_OK;
  }
  // Wait for the stream and fold kernel timings into the context.
  csx_status finish() {
    hipError_t e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (timer_) {
      std::lock_guard<std::mutex> lock(ctx_->mu);
      for (auto& p : timer_->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
          auto& st = ctx_->stats[p.name];
          st.first += 1;
          st.second += ms;
        }
      }
      // a call may launch and finish again (one launch per plan group of a
      // batch): later launches are timed too
      timer_.reset(new KernelTimer);
    }
    return CSX_OK;
  }
  hipStream_t stream() const { return stream_; }
  KernelTimer* timer() const { return timer_.get(); }
  csx_ctx* ctx() const { return ctx_; }
  // The call's next grow-only buffer of at least `bytes` (see Scratch).
  hipError_t take_scratch(size_t bytes, void** out) {
    if (!cache_) {
      std::lock_guard<std::mutex> lock(ctx_->mu);
      cache_ = &ctx_->scratch[stream_];  // map nodes are stable
    }
    if (next_ == cache_->size()) cache_->push_back({nullptr, 0});
    auto& slot = (*cache_)[next_++];
    if (slot.second < bytes) {
      if (slot.first) {
        // hipFree waits for the device, so no queued work still uses it
        device_free(slot.first, slot.second);
        slot = {nullptr, 0};
      }
      const size_t rounded = (bytes + (size_t(1) << 20) - 1) >> 20 << 20;
      const hipError_t e = device_alloc(&slot.first, rounded);
      if (e != hipSuccess) {
        slot.first = nullptr;
        return e;
      }
      slot.second = rounded;
    }
    *out = slot.first;
    return hipSuccess;
  }
  // Pinned host staging buffer `index` of the call's stream, at least
  // `bytes` (grow-only).  Copies through it are asynchronous; the stream is
  // idle when a call starts, so the buffer is free to overwrite.
  hipError_t take_pinned(size_t index, size_t bytes, void** out) {
    std::vector<std::pair<void*, size_t>>* pins;
    {
      std::lock_guard<std::mutex> lock(ctx_->mu);
      pins = &ctx_->pinned[stream_];
    }
    if (pins->size() <= index) pins->resize(index + 1, {nullptr, 0});
    auto& slot = (*pins)[index];
    if (slot.second < bytes) {
      if (slot.first) (void)hipHostFree(slot.first);
      slot = {nullptr, 0};
      const size_t rounded = (bytes + 4095) / 4096 * 4096;
      const hipError_t e = hipHostMalloc(&slot.first, rounded, hipHostMallocDefault);
      if (e != hipSuccess) {
        slot.first = nullptr;
        return e;
      }
      slot.second = rounded;
    }
    *out = slot.first;
    return hipSuccess;
  }

 private:
  csx_ctx* ctx_;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<KernelTimer> timer_;
  std::vector<std::pair<void*, size_t>>* cache_ = nullptr;
  size_t next_ = 0;
};

// Device buffer of one call taken from grow-only buffers kept with the
// call's stream, so that a call allocates and frees nothing once its sizes
// have been seen.  A stream serves one call at a time, so the buffers are the
// call's alone, and work that an earlier call queued on the same stream runs
// before this call's.  Requests above kScratchCacheLimit are allocated and
// freed on the stream like DevBuf, so a context does not keep them.
class Scratch {
 public:
  static constexpr size_t kScratchCacheLimit = size_t(1) << 30;
  explicit Scratch(Call& call) : call_(call) {}
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  hipError_t alloc(size_t bytes, hipStream_t s) {
    reset();
    if (bytes > kScratchCacheLimit) return big_.alloc(bytes, s);
    return call_.take_scratch(bytes == 0 ? 16 : bytes, &p_);
  }
  void reset() {
    big_.reset();
    p_ = nullptr;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_ ? p_ : big_.as<void>()); }

 private:
  Call& call_;
  void* p_ = nullptr;
  DevBuf big_;
};

#define CSX_HIP(expr, what)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)
#define CSX_TRY(expr)                \
  do {                               \
    csx_status st_ = (expr);         \
    if (st_ != CSX_OK) return st_;   \
  } while (0)

// Point set of a structure for a solvent model (atoms only when p is NULL or
// analytic), built and uploaded on first use.
csx_status get_point_set(csx_structure* s, const csx_solvent_params* p,
                         std::shared_ptr<PointSet>* out);

// Work items of a point set for a j chunk, built on first use.
csx_status get_item_set(PointSet& ps, int j_chunk, std::shared_ptr<ItemSet>* out);
int pair_j_chunk(int tuned, const PointSet& ps, int cus);
// The step of a uniform q grid in revolutions as FP32, or 0 where the sine
// recurrence cannot take the grid (not uniform, fewer than 3 points, or a
// step turning the phase of the largest pair distance 2 r_max by more than
// 1/4 revolution).
float uniform_q_step(const double* q, size_t nq, double rmax);
bool radial_distance_bins(const double* q, size_t nq, double rmax, double max_bins, double* h,
                          int* nk);
// The widest q tile the recurrence takes for that step and point radius (its
// phase error grows with the phase a tile accumulates).
int rec_max_tile(float step, double rmax);

// The cached result of `key` on a point set.  On return the entry's mutex
// is locked by *lock; entry->ready tells whether it holds the result.  When
// it does not, the caller computes it, sets `bytes` and `ready`, and calls
// cache_commit (or cache_drop on failure) before releasing the lock.  With
// caching off the entry is private to the call.
std::shared_ptr<CacheEntry> cache_acquire(csx_ctx* ctx, PointSet& ps, const std::string& key,
                                          std::unique_lock<std::mutex>* lock);
void cache_commit(csx_ctx* ctx, PointSet& ps, const std::shared_ptr<CacheEntry>& e);
void cache_drop(PointSet& ps, const std::shared_ptr<CacheEntry>& e);
// The ready entry of `key` without waiting (null when absent, still being
// computed, or with caching off), and the store of a result computed without
// an entry (kept unless the key is already present).
std::shared_ptr<CacheEntry> cache_find(csx_ctx* ctx, PointSet& ps, const std::string& key);
void cache_store(csx_ctx* ctx, PointSet& ps, const std::string& key,
                 const std::vector<double>& tables);

// Hydration grid of a structure, built on its device on first use and kept
// with the structure; the caller holds s->mu.  kAdaptiveGrid takes (spacing,
// clearance, padding), kExteriorGrid (spacing, probe radius, shell thickness).
constexpr int kAdaptiveGrid = 0;
constexpr int kExteriorGrid = 1;
csx_status structure_grid(csx_structure* s, int kind, double spacing, double a, double b,
                          const Grid** out);

}  // namespace csx

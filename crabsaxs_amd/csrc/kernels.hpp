// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace csx {

// Per-call collection of kernel durations measured with HIP events recorded
// on the launch stream.  Events are read after the call's stream has been
// synchronised, so timing adds no synchronisation of its own.
struct KernelTimer {
  struct Pending {
    std::string name;
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
  ~KernelTimer() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
  }
};

class TimerScope {
 public:
  TimerScope(KernelTimer* t, const char* name, hipStream_t s) : timer_(t), stream_(s) {
    if (!timer_) return;
    KernelTimer::Pending p{name, nullptr, nullptr};
    if (hipEventCreate(&p.start) != hipSuccess || hipEventCreate(&p.stop) != hipSuccess) {
      timer_ = nullptr;
      return;
    }
    (void)hipEventRecord(p.start, stream_);
    timer_->pending.push_back(p);
    index_ = timer_->pending.size() - 1;
  }
  void stop() {
    if (timer_ && !stopped_) (void)hipEventRecord(timer_->pending[index_].stop, stream_);
    stopped_ = true;
  }
  ~TimerScope() { stop(); }

 private:
  KernelTimer* timer_;
  hipStream_t stream_;
  size_t index_ = 0;
  bool stopped_ = false;
};

// Hydration grids (grid.hip).  A grid of nx * ny * nz voxels (at most
// INT32_MAX) with voxel (x, y, z) centred at lo + (x h, y h, z h); byte masks
// are padded to grid_padded_voxels(count) and zero beyond the grid.
struct GridGeom {
  double lo[3];
  double h;
  int nx, ny, nz;
};
struct ElementRadii {
  double r[16];
};
long long grid_padded_voxels(long long count);
hipError_t launch_grid_fill(uint8_t* p, long long padded, hipStream_t stream);
hipError_t launch_grid_raster(const double* xyz, const uint8_t* element, int n,
                              const ElementRadii& radii, int per_element, double extra,
                              const GridGeom& g, uint8_t* mask, int value, int exterior_only,
                              hipStream_t stream, KernelTimer* timer);
// Unknown voxels (0) connected to the grid boundary become exterior (2).
hipError_t launch_grid_flood(uint8_t* state, const GridGeom& g, int* labels, uint8_t* exterior,
                             hipStream_t stream, KernelTimer* timer);
// mode 0 keeps a[v] == 3, mode 1 keeps a[v] != 0 && b[v] == 0; counts has
// padded / 4096 entries, offsets one more (offsets[last] = kept voxels).
hipError_t launch_grid_count(int mode, const uint8_t* a, const uint8_t* b, long long padded,
                             long long* counts, long long* offsets, hipStream_t stream);
hipError_t launch_grid_emit(int mode, const uint8_t* a, const uint8_t* b, long long padded,
                            const GridGeom& g, const long long* offsets, double* sites,
                            hipStream_t stream);

// EnsembleFitter::fit_core candidates (ensemble.hip): profiles of every
// (c1, c2) candidate and their projected-gradient weight loops.
hipError_t launch_ensemble_candidates(const double* bases, const double* g, const double* c2v,
                                      int m, int n, int ncand, const double* y, const double* s,
                                      int uniform, double reg, int scale_mode, double scale_value,
                                      int bg_mode, double bg_value, double* curves, double* ws,
                                      double* sortbuf, int mpad, double* w_out, double* lin_out,
                                      double* fitted_out, hipStream_t stream);

hipError_t launch_debye_plain(const double4* pts, const double* amp, const double* q, int npts,
                              int nq, int nch, int nterms, double* partial, double* out,
                              hipStream_t stream, KernelTimer* timer);
hipError_t launch_expand_amplitudes(const double* table, const int* type, const double* occ,
                                    int npts, int ntypes, int nq, int nch, double* amp,
                                    hipStream_t stream);
// persist_wgs > 0: persistent launch of that many workgroups over a task
// counter (device int, reset by the launcher); 0: one workgroup per task.
// sin_mode 0-2 evaluate a sine per pair and q (qstep 0); 3 runs the sine
// recurrence over the uniform grid qrev[tile start] + k qstep (revolutions).
hipError_t launch_debye_pairs(const float4* pts, const int4* items, int nitems, const int* bin_first,
                              int nbins, const float* qrev, float qstep, int nq_pad, int qt,
                              int sin_mode, int persist_wgs, int* counter, double* partial,
                              double* out, hipStream_t stream, KernelTimer* timer);

// The no-allocation entry point's device scratch (debye.hip): its address on
// a device (hipGetSymbolAddress, once per device with that device current),
// then the address and size for use by host code.
hipError_t noalloc_scratch_setup(int device);
void* noalloc_scratch(int device, size_t* bytes);
// dst[k] = float(q[k] / divisor) for k < n and 0 for n <= k < n_total, the
// values passed as kernel arguments (no staging allocation).
hipError_t launch_write_floats(float* dst, const double* q, int n, int n_total, double divisor,
                               hipStream_t stream);
// qstep != 0: the recurrence's q values (tile start qrev plus steps of qstep).
// Pair tables by radial interpolation of the pair-distance distribution
// (debye.hip): per table bin b the Chebyshev moment accumulators (int64
// [nbins][nk][8], then FP64 node weights in place,
// scaled by bin_scale[b] = 2^F_b) and out[b][q] = the raw table (sum of
// occ_i occ_j sin(2 pi q r) / r) at the q in revolutions qrev.  copies_wanted:
// LDS accumulator copies per workgroup (0: automatic, else a power of two).
// Distance bins of the radial coordinate adjoint: its per-pair coefficient
// table (32 B per bin) sits in LDS.
constexpr int kVjpRadialMaxBins = 2048;
int radial_node_count();
size_t radial_lds_bytes(int nk);
hipError_t launch_debye_pairs_radial(const float4* pts, const int4* items, int nitems,
                                     const int* bin_first, int nbins, const float* bin_scale,
                                     const double* bin_unscale, double h, int nk,
                                     const float* qrev, int nq_pad, int copies_wanted,
                                     unsigned long long* acc, double* out, size_t lds,
                                     hipStream_t stream, KernelTimer* timer);
hipError_t launch_finish_tables(const double* raw, const float* qrev, int qt, float qstep,
                                const double* q0, int nbins, int nq, int nq_pad, double* tables,
                                hipStream_t stream);

}  // namespace csx

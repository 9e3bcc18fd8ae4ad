// Internal state behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/crabsaxs.h"
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

struct Tuning {
  int q_tile = 16;     // QT of debye_pairs_f32: q values per sweep of a pair
  int sin_mode = 0;    // 0 software polynomial, 1 hardware V_SIN_F32
  int j_chunk = 2048;  // j atoms per work item
};

}  // namespace csx

struct csx_ctx {
  int device = 0;
  std::mutex mu;
  std::vector<hipStream_t> free_streams;
  std::atomic<bool> timing{false};
  std::map<std::string, std::pair<int64_t, double>> stats;
  csx::Tuning tuning;
};

namespace csx {

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
    return hipMallocAsync(&p_, bytes == 0 ? 16 : bytes, s);
  }
  void reset() {
    if (p_) (void)hipFreeAsync(p_, stream_);
    p_ = nullptr;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// Persistent device allocation owned by a structure.
class DevArray {
 public:
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() {
    if (p_) (void)hipFree(p_);
  }
  hipError_t upload(const void* host, size_t bytes) {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    hipError_t e = hipMalloc(&p_, bytes == 0 ? 16 : bytes);
    if (e != hipSuccess) return e;
    return bytes ? hipMemcpy(p_, host, bytes, hipMemcpyHostToDevice) : hipSuccess;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
};

// One set of scattering points (atoms, optionally followed by hydration
// sites as an extra type) laid out for both Debye kernels.
struct PointSet {
  int ntypes = 0;         // atom types (+1 when sites are present)
  int natoms = 0, nsites = 0;
  double site_volume = 0.0;
  std::vector<double> self;   // per type: sum occ^2
  std::vector<double> wsum;   // per type: sum occ
  // FP32 pair-table layout: type-sorted, centred, w = occupancy.
  DevArray pts32;             // float4[natoms + nsites]
  DevArray items;             // int4[nitems]: i0, ilen, j0, jlen
  DevArray bin_first;         // int[nbins + 1]
  int nitems = 0, nbins = 0, j_chunk = 0;
  // FP64 plain layout: original order, atoms then sites.
  DevArray pts64;             // double4[natoms + nsites]
  DevArray type64;            // int[natoms + nsites]
  DevArray occ64;             // double[natoms + nsites]
  std::vector<double> site_xyz;
};

}  // namespace csx

struct csx_structure {
  csx_ctx* ctx = nullptr;
  size_t n = 0;
  std::vector<double> xyz, occ;
  std::vector<uint8_t> element;
  std::vector<int8_t> group;
  std::vector<csx::TypeKey> types;  // sorted unique keys
  std::vector<int> type_of;         // per atom
  double centre[3] = {0, 0, 0};
  double max_radius = 0.0;          // from the centroid
  std::mutex mu;
  std::map<std::string, std::unique_ptr<csx::PointSet>> sets;
};

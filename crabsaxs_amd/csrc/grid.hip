// Hydration grids on the GPU (solvent.rs:121-223 and 259-401).
//
// Rasterisation visits, for every atom, the voxel box of the reference
// (floor/ceil with the 1e-10 slack, clamped to the grid) with one wave per
// atom and sets a byte per voxel whose centre passes the reference's FP64
// test |p - x|^2 <= r^2 + 1e-10.  Setting a mask is an OR over atoms, so the
// result does not depend on the order of the atoms or of the stores.  This
// file is compiled without FMA contraction, so voxel centres, box bounds and
// distances are bitwise the reference's.
//
// The exterior flood fill of solvent.rs:341-375 (breadth-first from the
// boundary through unknown voxels, six-connected) marks exactly the unknown
// voxels whose connected component touches the grid boundary.  Here that set
// comes from a union-find labelling of the unknown voxels: every x-run of
// unknown voxels starts as one tree under its first voxel (one pass per row),
// runs are joined along y and z by lock-free unions (atomicMin on parent
// labels, path halving in every find), a flattening pass points every voxel
// at its root, and the roots of boundary voxels are flagged.
//
// Sites are emitted in linear voxel order (z, y, x), the reference's scan
// order, by a three-kernel compaction: counts per 4096-voxel block, one
// exclusive scan, and an ordered write per block.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;  // consecutive voxels per thread in the compaction

// Grids hold at most INT32_MAX voxels, so voxel indices are 32-bit.
inline __device__ int voxel_index(const GridGeom& g, int x, int y, int z) {
  return (z * g.ny + y) * g.nx + x;
}

// Lower and upper voxel of one axis (solvent.rs:187-193): the floor / ceil of
// (p -/+ r - lo) / h with 1e-10 slack, clamped to [0, n - 1].
inline __device__ void axis_bounds(double p, double r, double lo, double h, int n, int& a,
                                   int& b) {
  const double lower = (p - r - lo) / h;
  const double upper = (p + r - lo) / h;
  const double l = fmin(fmax(floor(lower - 1e-10), 0.0), static_cast<double>(n));
  const double u = ceil(upper + 1e-10);
  a = static_cast<int>(l);
  b = u <= 0.0 ? 0 : static_cast<int>(fmin(u, static_cast<double>(n - 1)));
}

// One wave per atom: mask[v] = value for every voxel v of the atom's box whose
// centre lies within radius (r_elem[element] + extra, or extra alone when
// per_element is 0); with exterior_only, only voxels holding 2 change.
__global__ __launch_bounds__(kThreads) void grid_raster(const double* __restrict__ xyz,
                                                        const uint8_t* __restrict__ element, int n,
                                                        ElementRadii radii, int per_element,
                                                        double extra, GridGeom g,
                                                        uint8_t* __restrict__ mask, int value,
                                                        int exterior_only) {
  const int atom = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (atom >= n) return;
  const double px = xyz[3 * atom], py = xyz[3 * atom + 1], pz = xyz[3 * atom + 2];
  const double radius = per_element ? radii.r[element[atom]] + extra : extra;
  int x0, x1, y0, y1, z0, z1;
  axis_bounds(px, radius, g.lo[0], g.h, g.nx, x0, x1);
  axis_bounds(py, radius, g.lo[1], g.h, g.ny, y0, y1);
  axis_bounds(pz, radius, g.lo[2], g.h, g.nz, z0, z1);
  if (x1 < x0 || y1 < y0 || z1 < z0) return;
  const unsigned wx = x1 - x0 + 1, wy = y1 - y0 + 1;  // a box within the grid
  const double r2 = radius * radius;
  const uint8_t v = static_cast<uint8_t>(value);
  const int plane = g.nx * g.ny;
  // lanes take (x, y) columns of the box and walk z; dx^2 + dy^2 is formed
  // once per column, which is the reference's left-to-right order
  for (unsigned k = lane; k < wx * wy; k += 64) {
    const int x = x0 + static_cast<int>(k % wx);
    const int y = y0 + static_cast<int>(k / wx);
    const double dx = (g.lo[0] + static_cast<double>(x) * g.h) - px;
    const double dy = (g.lo[1] + static_cast<double>(y) * g.h) - py;
    const double dxy = dx * dx + dy * dy;
    int c = voxel_index(g, x, y, z0);
    for (int z = z0; z <= z1; ++z, c += plane) {
      const double dz = (g.lo[2] + static_cast<double>(z) * g.h) - pz;
      if (dxy + dz * dz <= r2 + 1e-10 && (!exterior_only || mask[c] == 2)) mask[c] = v;
    }
  }
}

__global__ void grid_fill(uint8_t* __restrict__ p, long long bytes16) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < bytes16) reinterpret_cast<uint4*>(p)[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ------------------------------------------------------------ union-find --

inline __device__ int uf_load(const int* L, int a) {
  return __hip_atomic_load(L + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of a's tree, halving the path on the way: a is re-pointed at its
// grandparent with atomicMin, which keeps every label pointing into the same
// set and never undoes a smaller link made concurrently.  Labels only
// decrease (L[a] <= a), so every chain ends.
inline __device__ int uf_find(int* L, int a) {
  for (;;) {
    const int p = uf_load(L, a);
    if (p == a) return a;
    const int g = uf_load(L, p);
    if (g == p) return p;
    atomicMin(L + a, g);
    a = g;
  }
}

// Join the trees of a and b: the larger root is linked below the smaller.
// When the larger root was linked elsewhere meanwhile (atomicMin returns a
// smaller parent), the loop continues with that parent, so no link is lost.
inline __device__ void uf_unite(int* L, int a, int b) {
  for (;;) {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(L + a, b);
    if (old == a) return;
    a = old;
  }
}

inline __device__ int thread_voxel() { return blockIdx.x * blockDim.x + threadIdx.x; }

// One thread per x-row: L[c] = first voxel of c's run of unknown voxels, -1
// for the other voxels.
__global__ void cc_rows(const uint8_t* __restrict__ state, GridGeom g, int* __restrict__ L) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.ny * g.nz) return;
  const int c0 = row * g.nx;
  int start = -1;
  for (int x = 0; x < g.nx; ++x) {
    const int c = c0 + x;
    if (state[c] == 0) {
      if (start < 0) start = c;
      L[c] = start;
    } else {
      start = -1;
      L[c] = -1;
    }
  }
}

// Join the runs along y and z.  A voxel whose left neighbour and that
// neighbour's y (z) neighbour are both unknown skips the y (z) union: the two
// runs are joined by the first voxel of that overlap.
__global__ void cc_merge(const uint8_t* __restrict__ state, GridGeom g, int count, int* L) {
  const int c = thread_voxel();
  if (c >= count || state[c] != 0) return;
  const int plane = g.nx * g.ny;
  const int x = c % g.nx, y = (c / g.nx) % g.ny, z = c / plane;
  const bool left = x > 0 && state[c - 1] == 0;
  if (y + 1 < g.ny && state[c + g.nx] == 0 && !(left && state[c - 1 + g.nx] == 0))
    uf_unite(L, c, c + g.nx);
  if (z + 1 < g.nz && state[c + plane] == 0 && !(left && state[c - 1 + plane] == 0))
    uf_unite(L, c, c + plane);
}

// L[c] = root of c; roots of boundary voxels are flagged exterior.
__global__ void cc_flatten(const uint8_t* __restrict__ state, GridGeom g, int count, int* L,
                           uint8_t* __restrict__ exterior) {
  const int c = thread_voxel();
  if (c >= count || state[c] != 0) return;
  const int r = uf_find(L, c);
  __hip_atomic_store(L + c, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int plane = g.nx * g.ny;
  const int x = c % g.nx, y = (c / g.nx) % g.ny, z = c / plane;
  if (x == 0 || x + 1 == g.nx || y == 0 || y + 1 == g.ny || z == 0 || z + 1 == g.nz)
    exterior[r] = 1;
}

__global__ void cc_apply(uint8_t* __restrict__ state, int count, const int* __restrict__ L,
                         const uint8_t* __restrict__ exterior) {
  const int c = thread_voxel();
  if (c < count && state[c] == 0 && exterior[L[c]]) state[c] = 2;
}

// ----------------------------------------------------------- compaction --

// Bit k set when voxel base + k is kept: a[v] == 3 (exterior grid) or
// a[v] != 0 && b[v] == 0 (adaptive shell minus the molecular mask).
inline __device__ unsigned keep16(int mode, const uint8_t* a, const uint8_t* b, long long base) {
  const uint4 va = *reinterpret_cast<const uint4*>(a + base);
  const unsigned wa[4] = {va.x, va.y, va.z, va.w};
  unsigned wb[4] = {0u, 0u, 0u, 0u};
  if (mode == 1) {
    const uint4 vb = *reinterpret_cast<const uint4*>(b + base);
    wb[0] = vb.x;
    wb[1] = vb.y;
    wb[2] = vb.z;
    wb[3] = vb.w;
  }
  unsigned m = 0u;
  for (int k = 0; k < kPerThread; ++k) {
    const unsigned ba = (wa[k >> 2] >> (8 * (k & 3))) & 0xffu;
    const unsigned bb = (wb[k >> 2] >> (8 * (k & 3))) & 0xffu;
    const bool keep = mode == 1 ? (ba != 0u && bb == 0u) : ba == 3u;
    m |= (keep ? 1u : 0u) << k;
  }
  return m;
}

// Exclusive prefix of v over the workgroup (256 threads); *total the sum.
inline __device__ long long block_exclusive(long long v, long long* total) {
  __shared__ long long wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const long long o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  long long before = 0, all = 0;
  for (int w = 0; w < kThreads / 64; ++w) {
    if (w < wave) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(kThreads) void grid_count(int mode, const uint8_t* __restrict__ a,
                                                       const uint8_t* __restrict__ b,
                                                       long long* __restrict__ counts) {
  const long long base =
      static_cast<long long>(blockIdx.x) * (kThreads * kPerThread) + threadIdx.x * kPerThread;
  long long total;
  block_exclusive(__popc(keep16(mode, a, b, base)), &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// offsets[k] = counts[0] + ... + counts[k - 1] for k <= nblocks (one workgroup).
__global__ __launch_bounds__(kThreads) void grid_scan(const long long* __restrict__ counts,
                                                      long long nblocks,
                                                      long long* __restrict__ offsets) {
  long long carry = 0;
  for (long long start = 0; start < nblocks; start += kThreads) {
    const long long i = start + threadIdx.x;
    const long long v = i < nblocks ? counts[i] : 0;
    long long total;
    const long long excl = block_exclusive(v, &total);
    if (i < nblocks) offsets[i] = carry + excl;
    carry += total;
  }
  if (threadIdx.x == 0) offsets[nblocks] = carry;
}

__global__ __launch_bounds__(kThreads) void grid_emit(int mode, const uint8_t* __restrict__ a,
                                                      const uint8_t* __restrict__ b, GridGeom g,
                                                      const long long* __restrict__ offsets,
                                                      double* __restrict__ sites) {
  const long long base =
      static_cast<long long>(blockIdx.x) * (kThreads * kPerThread) + threadIdx.x * kPerThread;
  unsigned m = keep16(mode, a, b, base);
  long long total;
  long long pos = offsets[blockIdx.x] + block_exclusive(__popc(m), &total);
  while (m) {
    const int k = __ffs(m) - 1;
    m &= m - 1;
    const int c = static_cast<int>(base + k);  // a kept voxel lies inside the grid
    const int x = c % g.nx, y = (c / g.nx) % g.ny, z = c / (g.nx * g.ny);
    sites[3 * pos] = g.lo[0] + static_cast<double>(x) * g.h;
    sites[3 * pos + 1] = g.lo[1] + static_cast<double>(y) * g.h;
    sites[3 * pos + 2] = g.lo[2] + static_cast<double>(z) * g.h;
    ++pos;
  }
}

inline unsigned blocks_for(long long n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

long long grid_padded_voxels(long long count) {
  const long long blk = kThreads * kPerThread;
  return (count + blk - 1) / blk * blk;
}

hipError_t launch_grid_fill(uint8_t* p, long long padded, hipStream_t stream) {
  const long long n16 = padded / 16;
  if (n16 > 0) grid_fill<<<blocks_for(n16, kThreads), kThreads, 0, stream>>>(p, n16);
  return hipGetLastError();
}

hipError_t launch_grid_raster(const double* xyz, const uint8_t* element, int n,
                              const ElementRadii& radii, int per_element, double extra,
                              const GridGeom& g, uint8_t* mask, int value, int exterior_only,
                              hipStream_t stream, KernelTimer* timer) {
  if (n <= 0) return hipSuccess;
  TimerScope scope(timer, "grid_raster", stream);
  grid_raster<<<blocks_for(n, kThreads / 64), kThreads, 0, stream>>>(
      xyz, element, n, radii, per_element, extra, g, mask, value, exterior_only);
  return hipGetLastError();
}

hipError_t launch_grid_flood(uint8_t* state, const GridGeom& g, int* labels, uint8_t* exterior,
                             hipStream_t stream, KernelTimer* timer) {
  const long long count = static_cast<long long>(g.nx) * g.ny * g.nz;
  if (count > INT32_MAX) return hipErrorInvalidValue;
  const int n = static_cast<int>(count);
  const unsigned blocks = blocks_for(count, kThreads);
  TimerScope scope(timer, "grid_flood", stream);
  cc_rows<<<blocks_for(static_cast<long long>(g.ny) * g.nz, kThreads), kThreads, 0, stream>>>(
      state, g, labels);
  cc_merge<<<blocks, kThreads, 0, stream>>>(state, g, n, labels);
  cc_flatten<<<blocks, kThreads, 0, stream>>>(state, g, n, labels, exterior);
  cc_apply<<<blocks, kThreads, 0, stream>>>(state, n, labels, exterior);
  return hipGetLastError();
}

hipError_t launch_grid_count(int mode, const uint8_t* a, const uint8_t* b, long long padded,
                             long long* counts, long long* offsets, hipStream_t stream) {
  const long long nblocks = padded / (kThreads * kPerThread);
  if (nblocks <= 0) return hipErrorInvalidValue;
  grid_count<<<static_cast<unsigned>(nblocks), kThreads, 0, stream>>>(mode, a, b, counts);
  grid_scan<<<1, kThreads, 0, stream>>>(counts, nblocks, offsets);
  return hipGetLastError();
}

hipError_t launch_grid_emit(int mode, const uint8_t* a, const uint8_t* b, long long padded,
                            const GridGeom& g, const long long* offsets, double* sites,
                            hipStream_t stream) {
  const long long nblocks = padded / (kThreads * kPerThread);
  if (nblocks <= 0) return hipErrorInvalidValue;
  grid_emit<<<static_cast<unsigned>(nblocks), kThreads, 0, stream>>>(mode, a, b, g, offsets, sites);
  return hipGetLastError();
}

}  // namespace csx

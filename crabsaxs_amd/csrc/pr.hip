// Coordinate pair-distance distribution P(r) (metrics.rs:475-548): the
// weighted histogram of all atom-pair distances, on the GPU.
//
// The pairs are visited as block pairs (I <= J) of 256 atoms, like the
// adjoint.  Distances are FP64 with the reference's operation order
// ((dx dx + dy dy) + dz dz, correctly rounded sqrt and division; this file is
// compiled without FMA contraction), so Dmax and every pair's bin index are
// exactly the reference's.  Only the order of the FP64 weight sums differs
// (per-wave LDS histograms, then a fixed-order reduction over block pairs).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kBlock = 256;

// max over i < j of |x_i - x_j|^2, as the bit pattern of a non-negative
// double (whose unsigned order is the numeric order).
__global__ __launch_bounds__(kBlock) void pr_dmax2(const double4* __restrict__ pts, int n,
                                                   const int2* __restrict__ items,
                                                   unsigned long long* __restrict__ out) {
  __shared__ double4 tile[kBlock];
  __shared__ double wmax[kBlock / 64];
  const int2 it = items[blockIdx.x];
  const int tid = threadIdx.x;
  const int i = it.x * kBlock + tid, j0 = it.y * kBlock;
  const double4 pi = i < n ? pts[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  tile[tid] = j0 + tid < n ? pts[j0 + tid] : make_double4(0.0, 0.0, 0.0, 0.0);
  __syncthreads();
  const int jn = min(kBlock, n - j0);
  double m = 0.0;
  if (i < n)
    for (int t = (it.x == it.y ? tid + 1 : 0); t < jn; ++t) {
      const double dx = pi.x - tile[t].x, dy = pi.y - tile[t].y, dz = pi.z - tile[t].z;
      const double d2 = dx * dx + dy * dy + dz * dz;
      m = d2 > m ? d2 : m;
    }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((tid & 63) == 0) wmax[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    double b = wmax[0];
    for (int w = 1; w < kBlock / 64; ++w) b = wmax[w] > b ? wmax[w] : b;
    atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(b)));
  }
}

// partial[item][b] = sum of w_i w_j over the item's pairs in bin
// b = min(floor(sqrt(d2) / step), bins - 1).
__global__ __launch_bounds__(kBlock) void pr_hist(const double4* __restrict__ pts, int n,
                                                  const int2* __restrict__ items, double step,
                                                  int bins, double* __restrict__ partial) {
  extern __shared__ double hist[];  // [4 waves][bins]
  __shared__ double4 tile[kBlock];
  const int2 it = items[blockIdx.x];
  const int tid = threadIdx.x, wave = tid >> 6;
  for (int b = tid; b < 4 * bins; b += kBlock) hist[b] = 0.0;
  const int i = it.x * kBlock + tid, j0 = it.y * kBlock;
  const double4 pi = i < n ? pts[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  tile[tid] = j0 + tid < n ? pts[j0 + tid] : make_double4(0.0, 0.0, 0.0, 0.0);
  __syncthreads();
  const int jn = min(kBlock, n - j0);
  double* h = hist + wave * bins;
  if (i < n)
    for (int t = (it.x == it.y ? tid + 1 : 0); t < jn; ++t) {
      const double4 pj = tile[t];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double f = floor(d / step);
      const int b = f < double(bins - 1) ? int(f) : bins - 1;
      atomicAdd(h + b, pi.w * pj.w);
    }
  __syncthreads();
  for (int b = tid; b < bins; b += kBlock)
    partial[size_t(blockIdx.x) * bins + b] =
        ((hist[b] + hist[bins + b]) + hist[2 * bins + b]) + hist[3 * bins + b];
}

// out[b] = sum over items, in item order, of partial[item][b].
__global__ void pr_reduce(const double* __restrict__ partial, int nitems, int bins,
                          double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bins) return;
  double s = 0.0;
#pragma unroll 32
  for (int k = 0; k < nitems; ++k) s += partial[size_t(k) * bins + b];
  out[b] = s;
}

}  // namespace

size_t pr_hist_lds_bytes(int bins) { return size_t(4) * bins * sizeof(double); }

hipError_t launch_pr_dmax2(const double4* pts, int n, const int2* items, int nitems,
                           unsigned long long* out, hipStream_t stream) {
  if (nitems <= 0) return hipSuccess;
  pr_dmax2<<<nitems, kBlock, 0, stream>>>(pts, n, items, out);
  return hipGetLastError();
}

hipError_t launch_pr_hist(const double4* pts, int n, const int2* items, int nitems, double step,
                          int bins, double* partial, double* out, hipStream_t stream,
                          KernelTimer* timer) {
  if (nitems <= 0 || bins <= 0) return hipErrorInvalidValue;
  {
    TimerScope scope(timer, "pr_hist", stream);
    pr_hist<<<nitems, kBlock, pr_hist_lds_bytes(bins), stream>>>(pts, n, items, step, bins,
                                                                 partial);
  }
  pr_reduce<<<(bins + 255) / 256, 256, 0, stream>>>(partial, nitems, bins, out);
  return hipGetLastError();
}

}  // namespace csx

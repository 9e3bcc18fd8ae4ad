// EnsembleFitter::fit_core on the GPU (ensemble.rs:234-299 with the
// optimize_weights loop of ensemble.rs:317-381 and project_simplex of
// ensemble.rs:449-465): one workgroup per (c1, c2) candidate runs the 250
// projected-gradient iterations on device-resident profiles.
//
// Every sum runs in the host restatement's order (sequential where the host
// loop is sequential) and this file is compiled without FMA contraction, so
// the weights, scale, background and chi^2 are bitwise those of the host
// fit_core: parallel lanes only take independent elements (one profile
// point, one conformer), and the order-dependent reductions (the linear
// solve, the gradient norm, the simplex threshold, the step size) run in
// one lane.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kThreads = 256;

inline __device__ double clamp0(double v) { return v < 0.0 ? 0.0 : v; }  // std::max(v, 0.0)

// curves[cand][c][p] = max(combine(basis_c, g, c2), 0) with g = G(q_p, c1)
// from the host (same expression and order as the host combine).
__global__ void combine_candidates(const double* __restrict__ bases, const double* __restrict__ g,
                                   const double* __restrict__ c2v, int m, int n, int ncand,
                                   double* __restrict__ curves) {
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = size_t(ncand) * m * n;
  if (idx >= total) return;
  const int p = int(idx % n);
  const int c = int((idx / n) % m);
  const int cand = int(idx / (size_t(n) * m));
  const double* b = bases + size_t(c) * 6 * n;
  const double gg = g[size_t(cand) * n + p], cc = c2v[cand];
  const double v = b[p] + gg * gg * b[n + p] + cc * cc * b[2 * n + p] - 2.0 * gg * b[3 * n + p] +
                   2.0 * cc * b[4 * n + p] - 2.0 * gg * cc * b[5 * n + p];
  curves[idx] = clamp0(v);
}

struct LinFit {
  double scale, bg, chi;
};

// solve_linear_fixed (ensemble.rs:383-434), sequential as the host loop.
__device__ LinFit linear_fixed(const double* x, const double* y, const double* s, int n,
                               int scale_mode, double scale_value, int bg_mode, double bg_value) {
  double scale;
  if (scale_mode == 0) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
      num += x[i] * y[i] / (s[i] * s[i]);
      den += x[i] * x[i] / (s[i] * s[i]);
    }
    scale = den > 0.0 ? num / den : 0.0;
  } else {
    scale = scale_value;
  }
  scale = scale < 0.0 ? 0.0 : scale;  // std::max(scale, 0.0)
  double bg;
  if (bg_mode == 0) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
      num += (y[i] - scale * x[i]) / (s[i] * s[i]);
      den += 1.0 / (s[i] * s[i]);
    }
    bg = den > 0.0 ? num / den : 0.0;
  } else {
    bg = bg_mode == 1 ? 0.0 : bg_value;
  }
  double chi = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (scale * x[i] + bg - y[i]) / s[i];
    chi += r * r;
  }
  return {scale, bg, chi / (n > 1 ? double(n - 1) : 1.0)};
}

// Bitonic sort of buf[0 .. mpad) into descending order (mpad a power of two).
__device__ void sort_descending(double* buf, int mpad) {
  for (int k = 2; k <= mpad; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < mpad; i += blockDim.x) {
        const int partner = i ^ j;
        if (partner > i) {
          const bool desc = (i & k) == 0;
          const double a = buf[i], b = buf[partner];
          if (desc ? a < b : a > b) {
            buf[i] = b;
            buf[partner] = a;
          }
        }
      }
      __syncthreads();
    }
}

// One workgroup per candidate.  ws: per candidate [w m][prop m][g m]
// [mixture n][res n]; sortbuf: per candidate mpad doubles.
__global__ __launch_bounds__(kThreads) void optimize_weights_kernel(
    const double* __restrict__ curves_all, const double* __restrict__ y, const double* __restrict__ s,
    int m, int n, int mpad, int uniform, double reg, int scale_mode, double scale_value,
    int bg_mode, double bg_value, double* __restrict__ ws, double* __restrict__ sortbuf_all,
    double* __restrict__ w_out, double* __restrict__ lin_out, double* __restrict__ fitted_out) {
  const int cand = blockIdx.x;
  const double* curves = curves_all + size_t(cand) * m * n;
  double* w = ws + size_t(cand) * (3 * size_t(m) + 2 * size_t(n));
  double* prop = w + m;
  double* g = prop + m;
  double* mix = g + m;
  double* res = mix + n;
  double* sorted = sortbuf_all + size_t(cand) * mpad;
  __shared__ double sh_scale, sh_bg, sh_norm, sh_theta;
  __shared__ int sh_stop;
  for (int c = threadIdx.x; c < m; c += blockDim.x) w[c] = 1.0 / double(m);
  __syncthreads();
  auto mixture = [&]() {
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
      double v = 0.0;
      for (int c = 0; c < m; ++c) v += curves[size_t(c) * n + p] * w[c];
      mix[p] = v;
    }
    __syncthreads();
  };
  if (!uniform) {
    for (int it = 0; it < 250; ++it) {
      mixture();
      if (threadIdx.x == 0) {
        const LinFit lin = linear_fixed(mix, y, s, n, scale_mode, scale_value, bg_mode, bg_value);
        sh_scale = lin.scale;
        sh_bg = lin.bg;
      }
      __syncthreads();
      const double scale = sh_scale, bg = sh_bg;
      for (int p = threadIdx.x; p < n; p += blockDim.x) res[p] = (scale * mix[p] + bg - y[p]) / s[p];
      __syncthreads();
      for (int c = threadIdx.x; c < m; c += blockDim.x) {
        double v = 0.0;
        for (int p = 0; p < n; ++p) v += res[p] * scale * curves[size_t(c) * n + p] / s[p];
        g[c] = 2.0 * v + 2.0 * reg * (w[c] - 1.0 / double(m));
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double norm = 0.0;
        for (int c = 0; c < m; ++c) norm += g[c] * g[c];
        norm = sqrt(norm);
        sh_norm = norm < 1.0 ? 1.0 : norm;  // std::max(sqrt(norm), 1.0)
      }
      __syncthreads();
      const double norm = sh_norm;
      for (int c = threadIdx.x; c < mpad; c += blockDim.x) {
        if (c < m) {
          prop[c] = w[c] - 0.05 * g[c] / norm;
          sorted[c] = prop[c];
        } else {
          sorted[c] = -INFINITY;
        }
      }
      __syncthreads();
      sort_descending(sorted, mpad);
      if (threadIdx.x == 0) {
        double cum = 0.0;
        int rho = 0;
        for (int i = 0; i < m; ++i) {
          cum += sorted[i];
          if (sorted[i] + (1.0 - cum) / double(i + 1) > 0.0) rho = i + 1;
        }
        double top = 0.0;
        for (int i = 0; i < rho; ++i) top += sorted[i];
        sh_theta = (top - 1.0) / double(rho > 1 ? rho : 1);
      }
      __syncthreads();
      const double theta = sh_theta;
      for (int c = threadIdx.x; c < m; c += blockDim.x) prop[c] = clamp0(prop[c] - theta);
      __syncthreads();
      if (threadIdx.x == 0) {
        double delta = 0.0;
        for (int c = 0; c < m; ++c) delta += fabs(prop[c] - w[c]);
        sh_stop = delta < 1e-8;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < m; c += blockDim.x) w[c] = prop[c];
      __syncthreads();
      if (sh_stop) break;
    }
  }
  mixture();
  if (threadIdx.x == 0) {
    const LinFit lin = linear_fixed(mix, y, s, n, scale_mode, scale_value, bg_mode, bg_value);
    sh_scale = lin.scale;
    sh_bg = lin.bg;
    lin_out[3 * cand] = lin.scale;
    lin_out[3 * cand + 1] = lin.bg;
    lin_out[3 * cand + 2] = lin.chi;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < m; c += blockDim.x) w_out[size_t(cand) * m + c] = w[c];
  for (int p = threadIdx.x; p < n; p += blockDim.x)
    fitted_out[size_t(cand) * n + p] = sh_scale * mix[p] + sh_bg;
}

}  // namespace

hipError_t launch_ensemble_candidates(const double* bases, const double* g, const double* c2v,
                                      int m, int n, int ncand, const double* y, const double* s,
                                      int uniform, double reg, int scale_mode, double scale_value,
                                      int bg_mode, double bg_value, double* curves, double* ws,
                                      double* sortbuf, int mpad, double* w_out, double* lin_out,
                                      double* fitted_out, hipStream_t stream) {
  if (m <= 0 || n <= 0 || ncand <= 0 || mpad < m || (mpad & (mpad - 1)) != 0)
    return hipErrorInvalidValue;
  const size_t total = size_t(ncand) * m * n;
  combine_candidates<<<(total + 255) / 256, 256, 0, stream>>>(bases, g, c2v, m, n, ncand, curves);
  optimize_weights_kernel<<<ncand, kThreads, 0, stream>>>(curves, y, s, m, n, mpad, uniform, reg,
                                                          scale_mode, scale_value, bg_mode,
                                                          bg_value, ws, sortbuf, w_out, lin_out,
                                                          fitted_out);
  return hipGetLastError();
}

}  // namespace csx

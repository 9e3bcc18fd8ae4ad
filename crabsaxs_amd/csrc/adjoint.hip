// Adjoint of the Debye intensity with respect to atom coordinates (no
// reference counterpart; the forward sum is scattering.rs:100-106 and
// solvent.rs:533-558 with the per-type amplitudes A_a(q)).
//
// I(q) = sum_{i,j} A_i A_j sinc(q r_ij), A_i = occ_i A_{type(i)}(q), so
//   dI/dx_i = sum_{j != i} W'_{ab}(q) occ_i occ_j (u cos u - sin u) / r^3 (x_i - x_j)
// with u = q r_ij and W'_{ab}(q) = 2 A_a(q) A_b(q) / q (zero at q = 0).
// debye_adjoint_f32: a workgroup owns 256 consecutive type-sorted atoms (one
// per lane) and a chunk of QT q values and sweeps every other atom j through
// LDS tiles, one type segment at a time (W' of the lane's type and the
// segment's type for the chunk is held in registers).  sin and cos come from
// V_SIN_F32 / V_COS_F32 in revolutions after an exact reduction; FP32 tile
// sums are folded into FP64 per tile.
//  JAC = false: the q weights w_q are folded into W' on the host; the kernel
//               writes out[qchunk][atom][3] (summed over chunks on the host).
//  JAC = true : out[q][atom][3], the full Jacobian.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kAdjBlock = 256;

template <int QT, bool JAC>
__global__ __launch_bounds__(kAdjBlock) void debye_adjoint_f32(
    const float4* __restrict__ pts, const int* __restrict__ orig, const int* __restrict__ type,
    int npts, const int* __restrict__ seg_start, int ntypes, const float* __restrict__ qrev,
    const float* __restrict__ Wp /* [ntypes][ntypes][nq_pad] */, int nq_pad,
    double* __restrict__ out) {
  __shared__ float4 tile[kAdjBlock];
  const int i = blockIdx.x * kAdjBlock + threadIdx.x;
  const int qc = blockIdx.y * QT;
  const bool vi = i < npts;
  const float4 pi = vi ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int ta = vi ? type[i] : 0;
  float qk[QT];
#pragma unroll
  for (int k = 0; k < QT; ++k) qk[k] = qrev[qc + k];
  constexpr int NACC = JAC ? QT : 1;
  double g64[NACC][3];
  float g[NACC][3];
#pragma unroll
  for (int k = 0; k < NACC; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      g64[k][c] = 0.0;
      g[k][c] = 0.f;
    }
  const float two_pi = 6.283185307179586f;
  for (int b = 0; b < ntypes; ++b) {
    float w[QT];
#pragma unroll
    for (int k = 0; k < QT; ++k) w[k] = Wp[(size_t(ta) * ntypes + b) * nq_pad + qc + k];
    const int s0 = seg_start[b], s1 = seg_start[b + 1];
    for (int t0 = s0; t0 < s1; t0 += kAdjBlock) {
      __syncthreads();
      const int j = t0 + threadIdx.x;
      tile[threadIdx.x] = j < s1 ? pts[j] : make_float4(0.f, 0.f, 0.f, 0.f);
      __syncthreads();
      const int jn = min(kAdjBlock, s1 - t0);
      for (int t = 0; t < jn; ++t) {
        const float4 pj = tile[t];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        // coincident atoms (including j == i) have no gradient
        const bool use = d2 > 1e-20f;
        const float r = __builtin_sqrtf(use ? d2 : 1.0f);
        const float ir = __builtin_amdgcn_rcpf(r);
        const float ir3 = use ? pj.w * ir * ir * ir : 0.f;
        float F = 0.f;
#pragma unroll
        for (int k = 0; k < QT; ++k) {
          const float v = qk[k] * r;                 // revolutions
          const float f = v - __builtin_rintf(v);    // exact reduction
          const float s = __builtin_amdgcn_sinf(f);
          const float c = __builtin_amdgcn_cosf(f);
          const float u = two_pi * v;                // radians
          const float term = __builtin_fmaf(u, c, -s);
          if (JAC) {
            const float h = w[k] * term * ir3;
            g[k % NACC][0] = __builtin_fmaf(h, dx, g[k % NACC][0]);
            g[k % NACC][1] = __builtin_fmaf(h, dy, g[k % NACC][1]);
            g[k % NACC][2] = __builtin_fmaf(h, dz, g[k % NACC][2]);
          } else {
            F = __builtin_fmaf(w[k], term, F);
          }
        }
        if (!JAC) {
          const float h = F * ir3;
          g[0][0] = __builtin_fmaf(h, dx, g[0][0]);
          g[0][1] = __builtin_fmaf(h, dy, g[0][1]);
          g[0][2] = __builtin_fmaf(h, dz, g[0][2]);
        }
      }
#pragma unroll
      for (int k = 0; k < NACC; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          g64[k][c] += double(g[k][c]);
          g[k][c] = 0.f;
        }
    }
  }
  if (!vi) return;
  const double oi = double(pi.w);
  const int o = orig[i];
  if (JAC) {
#pragma unroll
    for (int k = 0; k < QT; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) out[(size_t(qc + k) * npts + o) * 3 + c] = oi * g64[k][c];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[(size_t(blockIdx.y) * npts + o) * 3 + c] = oi * g64[0][c];
  }
}

}  // namespace

hipError_t launch_debye_adjoint(const float4* pts, const int* orig, const int* type, int npts,
                                const int* seg_start, int ntypes, const float* qrev,
                                const float* Wp, int nq_pad, bool jacobian, double* out,
                                hipStream_t stream, KernelTimer* timer) {
  if (npts <= 0) return hipSuccess;
  const int qt = jacobian ? 8 : 32;
  if (nq_pad % qt != 0) return hipErrorInvalidValue;
  dim3 grid((npts + kAdjBlock - 1) / kAdjBlock, nq_pad / qt);
  TimerScope scope(timer, "debye_adjoint_f32", stream);
  if (jacobian)
    debye_adjoint_f32<8, true><<<grid, kAdjBlock, 0, stream>>>(pts, orig, type, npts, seg_start,
                                                               ntypes, qrev, Wp, nq_pad, out);
  else
    debye_adjoint_f32<32, false><<<grid, kAdjBlock, 0, stream>>>(pts, orig, type, npts, seg_start,
                                                                 ntypes, qrev, Wp, nq_pad, out);
  return hipGetLastError();
}

int adjoint_q_tile(bool jacobian) { return jacobian ? 8 : 32; }

}  // namespace csx

// Adjoint of the Debye intensity with respect to atom coordinates (no
// reference counterpart; the forward sum is scattering.rs:100-106 and
// solvent.rs:533-558 with the per-type amplitudes A_a(q)).
//
// I(q) = sum_{i,j} A_i A_j sinc(q r_ij), A_i = occ_i A_{type(i)}(q), so
//   dI/dx_i = sum_{j != i} W'_{ab}(q) occ_i occ_j (u cos u - sin u) / r^3 (x_i - x_j)
// with u = q r_ij and W'_{ab}(q) = 2 A_a(q) A_b(q) / q (zero at q = 0).
//
// debye_vjp_f32 (sum_q w_q dI(q)/dx): every unordered pair once.  Work items
// are pairs (I <= J) of type-pure 256-atom blocks; lane l owns atom I_l and
// at step t of a sweep meets atom J_((t + l) mod 256), so the lanes of a wave
// touch 64 distinct j slots and the j-side sums go into per-wave LDS arrays
// without bank conflicts or races (fixed order: deterministic).  With the host
// tables B_ab(q) = w_q W'_ab(q) and A_ab(q) = q B_ab(q) the q loop is
//   v = q r / 2 pi,  C += A cos 2 pi v,  S += B sin 2 pi v,  F = r C - S,
// i.e. two FMAs and one V_SIN_F32 / V_COS_F32 pair per pair.q.  FP32 sums
// per q chunk are folded into FP64 partials per (item, side), which a second
// kernel adds per block in fixed order.
//
// debye_jacobian_f32 (dI(q)/dx for every q): a workgroup owns 256
// type-sorted atoms and QT q values and sweeps all other atoms (each pair is
// visited from both ends; the output is q x n x 3 anyway).  Per pair.q one
// V_SIN_F32 / V_COS_F32 pair (on q r / 2 pi itself where the host has
// checked it is at most 250 revolutions, else after v - rint(v)), the weight
// and three FMAs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kAdjBlock = 256;
constexpr float kTwoPi = 6.283185307179586f;

// sin and cos of 2 pi v, v in revolutions: SM 2 takes V_SIN_F32 / V_COS_F32
// on v itself (|v| <= 256, checked on the host), SM 1 after the exact
// reduction v - rint(v).
template <int SM>
__device__ inline void sincos_rev(float v, float& s, float& c) {
  const float f = SM == 2 ? v : v - __builtin_rintf(v);
  s = __builtin_amdgcn_sinf(f);
  c = __builtin_amdgcn_cosf(f);
}

// 2 sin(pi t) and cos(pi t) for |t| <= 1/4 (t in revolutions): Taylor series
// to t^9 and t^10, truncation below 4e-9 of the values; the sine keeps its
// relative accuracy for tiny t.
__device__ inline float two_sin_pi_small(float t) {
  const float t2 = t * t;
  float p = 0.16429176926612854f;
  p = __builtin_fmaf(p, t2, -1.1985290050506592f);
  p = __builtin_fmaf(p, t2, 5.100327968597412f);
  p = __builtin_fmaf(p, t2, -10.33542537689209f);
  // 2 pi t with 2 pi split in two FP32 parts (the FP32 2 pi alone is 2.8e-8
  // too large, a bias every pair's step would share)
  return __builtin_fmaf(t, 6.2831854820251465f,
                        __builtin_fmaf(t * t2, p, t * -1.7484555314695172e-07f));
}
__device__ inline float cos_pi_small(float t) {
  const float t2 = t * t;
  float p = -0.025806890800595284f;
  p = __builtin_fmaf(p, t2, 0.23533062636852264f);
  p = __builtin_fmaf(p, t2, -1.3352627754211426f);
  p = __builtin_fmaf(p, t2, 4.058712005615234f);
  p = __builtin_fmaf(p, t2, -4.934802055358887f);
  return __builtin_fmaf(p, t2, 1.0f);
}

// The pair's sums over the chunk's QT q values, C = sum_k A_k cos 2 pi v_k
// and S = sum_k B_k sin 2 pi v_k, v_k = q_k r.  SM 1 and 2 evaluate the sine
// and cosine per q.  SM 3 (uniform q grid: q_k = q0 + k qstep) carries
// s_k and c_k by Reinsch's recurrence with the step difference kept as its
// own variable (d_{k+1} = d_k + lam x_k, x_{k+1} = x_k + d_{k+1}, lam =
// -(2 sin(pi th))^2, th = qstep r), four FMAs and two adds per q instead of
// a multiply, two quarter-rate transcendentals and two FMAs.  Its start
// comes from the sine and cosine of the half step before q0,
// (sh, ch) at 2 pi (q0 r - th / 2): s_0 = sh cos(pi th) + ch sin(pi th),
// c_0 = ch cos(pi th) - sh sin(pi th), and the first differences
// s_0 - s_{-1} = 2 sin(pi th) ch, c_0 - c_{-1} = -2 sin(pi th) sh.
template <int QT, int SM>
__device__ __attribute__((always_inline)) inline void pair_q_sums(float r, const float* qk,
                                                                  float qstep, const float* ak,
                                                                  const float* bk, float& C,
                                                                  float& S) {
  C = 0.f;
  S = 0.f;
  if constexpr (SM == 3) {
    const float th = qstep * r;
    const float h2 = two_sin_pi_small(th);
    const float g = cos_pi_small(th);
    const float hh = 0.5f * h2;
    const float v = __builtin_fmaf(-0.5f, th, qk[0] * r);
    const float sh = __builtin_amdgcn_sinf(v), ch = __builtin_amdgcn_cosf(v);
    float sn = __builtin_fmaf(ch, hh, sh * g);
    float cs = __builtin_fmaf(-sh, hh, ch * g);
    float ds = h2 * ch, dc = -(h2 * sh);
    const float lam = -(h2 * h2);
#pragma unroll
    for (int k = 0; k < QT; ++k) {
      C = __builtin_fmaf(ak[k], cs, C);
      S = __builtin_fmaf(bk[k], sn, S);
      if (k + 1 < QT) {
        ds = __builtin_fmaf(lam, sn, ds);
        dc = __builtin_fmaf(lam, cs, dc);
        sn += ds;
        cs += dc;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < QT; ++k) {
      float sn, cs;
      sincos_rev<SM>(qk[k] * r, sn, cs);
      C = __builtin_fmaf(ak[k], cs, C);
      S = __builtin_fmaf(bk[k], sn, S);
    }
  }
}

template <int QT, int SM>
__global__ __launch_bounds__(kAdjBlock) void debye_vjp_f32(
    const float4* __restrict__ pts, const int4* __restrict__ blocks /* start, len, type, 0 */,
    const int2* __restrict__ items /* (I, J), I <= J */, int ntypes,
    const float* __restrict__ qrev, float qstep,
    const float* __restrict__ AB /* [T][T][2][nq_pad] */, int nq_pad, int qsplit, int tsplit,
    double* __restrict__ part /* [items][qsplit * tsplit][2][3][256] */) {
  __shared__ float4 tile[kAdjBlock];
  __shared__ float accj[4][3][kAdjBlock];
  __shared__ float qab[3][QT];
  const int tid = threadIdx.x, wave = tid >> 6;
  const int2 it = items[blockIdx.x];
  const int4 bi = blocks[it.x], bj = blocks[it.y];
  const bool diag = it.x == it.y;
  const float4 pi = tid < bi.y ? pts[bi.x + tid] : make_float4(0.f, 0.f, 0.f, 0.f);
  tile[tid] = tid < bj.y ? pts[bj.x + tid] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int c = 0; c < 3; ++c) accj[w][c][tid] = 0.f;
  const float* A = AB + (size_t(bi.z) * ntypes + bj.z) * 2 * nq_pad;
  const float* B = A + nq_pad;
  double gi64[3] = {0.0, 0.0, 0.0}, gj64[3] = {0.0, 0.0, 0.0};
  // this workgroup's share of the q chunks and of the 256 steps of a sweep
  // (blockIdx.y = q part * tsplit + step part; tsplit divides 256)
  const int nqc = nq_pad / QT;
  const int qpart = int(blockIdx.y) / tsplit, tpart = int(blockIdx.y) % tsplit;
  const int c0 = qpart * nqc / qsplit, c1 = (qpart + 1) * nqc / qsplit;
  const int ta = tpart * (kAdjBlock / tsplit), tb = ta + kAdjBlock / tsplit;
  __syncthreads();
  for (int qc = c0 * QT; qc < c1 * QT; qc += QT) {
    // The chunk's q, A and B pass through LDS so that they sit in VGPRs:
    // as uniform global loads they would take 3 QT SGPRs and spill at QT 32.
    if (tid < QT) {
      qab[0][tid] = qrev[qc + tid];
      qab[1][tid] = A[qc + tid];
      qab[2][tid] = B[qc + tid];
    }
    __syncthreads();
    constexpr int kQk = SM == 3 ? 1 : QT;  // the recurrence needs the chunk's first q only
    float qk[kQk], ak[QT], bk[QT];
#pragma unroll
    for (int k = 0; k < QT; ++k) {
      if (k < kQk) qk[k] = qab[0][k];
      ak[k] = qab[1][k];
      bk[k] = qab[2][k];
    }
    float gi[3] = {0.f, 0.f, 0.f};
    float4 pnext = tile[(ta + tid) & (kAdjBlock - 1)];  // the next step's j, read one step ahead
    for (int t = ta; t < tb; ++t) {
      const int jj = (t + tid) & (kAdjBlock - 1);
      const float4 pj = pnext;
      pnext = tile[(t + 1 + tid) & (kAdjBlock - 1)];
      const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      // coincident points have no gradient; a diagonal block counts j > i only
      const bool use = d2 > 1e-20f && (!diag || jj > tid);
      // r = d2 rsq(d2) with one Newton step (unbiased; see the pair kernel)
      const float dd = use ? d2 : 1.0f;
      const float ir = __builtin_amdgcn_rsqf(dd);
      const float r0 = dd * ir;
      const float r = __builtin_fmaf(__builtin_fmaf(-r0, r0, dd), 0.5f * ir, r0);
      const float wgt = use ? pi.w * pj.w * ir * ir * ir : 0.f;
      float C, S;
      pair_q_sums<QT, SM>(r, qk, qstep, ak, bk, C, S);
      const float h = __builtin_fmaf(r, C, -S) * wgt;
      gi[0] = __builtin_fmaf(h, dx, gi[0]);
      gi[1] = __builtin_fmaf(h, dy, gi[1]);
      gi[2] = __builtin_fmaf(h, dz, gi[2]);
      // Plain read-modify-write: slot jj belongs to this lane at this step
      // and LDS operations of one wave complete in order (LDS atomics here
      // measured 2.2x slower than the whole kernel without them).
      accj[wave][0][jj] -= h * dx;
      accj[wave][1][jj] -= h * dy;
      accj[wave][2][jj] -= h * dz;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gi64[c] += double(gi[c]);
      gj64[c] += ((double(accj[0][c][tid]) + double(accj[1][c][tid])) + double(accj[2][c][tid])) +
                 double(accj[3][c][tid]);
#pragma unroll
      for (int w = 0; w < 4; ++w) accj[w][c][tid] = 0.f;
    }
    __syncthreads();
  }
  double* out = part + (size_t(blockIdx.x) * gridDim.y + blockIdx.y) * 2 * 3 * kAdjBlock;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    out[c * kAdjBlock + tid] = gi64[c];
    out[(3 + c) * kAdjBlock + tid] = gj64[c];
  }
}

// ---------------------------------------------- radial interpolation --
// Sine mode 4: for one type pair the q sum of a pair,
//   F_ab(r) = sum_q B_ab(q) (q r cos q r - sin q r),  B_ab = 2 w_q A_a A_b / q,
// is a function of r alone, band-limited by q_max.  On each radial bin of
// width h (h q_max <= 1.04) it is replaced by its interpolant at 8 Chebyshev
// nodes, kept as the Chebyshev coefficients c_p of that bin, so a pair reads
// 8 coefficients and runs Clenshaw's recurrence instead of a q loop.  The
// error is at most (q_max h / 2)^8 / (2^7 8!) sum_q |B q^8| (q r + 9) per pair
// (the 8th derivative of u cos u - sin u is bounded by u + 8).
constexpr int kVjpNodes = 8;
struct VjpNodes {
  double x[kVjpNodes];             // Chebyshev nodes on [-1, 1]
  double m[kVjpNodes][kVjpNodes];  // coefficient p's weight of node m: (2 - [p = 0]) / 8 T_p(x_m)
};

// coef[ab][k][p] (FP32) from FP64 node values.  One 64-lane workgroup per
// (type pair, bin): lane m + 8 s sums node m over the q values q = s mod 8.
__global__ __launch_bounds__(64) void vjp_radial_coefficients(
    const double* __restrict__ qv, const double* __restrict__ B /* [T*T][nq] */, int nq, int nk,
    double h, VjpNodes nodes, float* __restrict__ coef) {
  __shared__ double part[8][kVjpNodes];
  __shared__ double fnode[kVjpNodes];
  const int m = threadIdx.x & 7, sl = threadIdx.x >> 3;
  const int ab = blockIdx.x / nk, k = blockIdx.x % nk;
  const double rho = (double(k) + 0.5 + 0.5 * nodes.x[m]) * h;
  const double* Bq = B + size_t(ab) * nq;
  double f = 0.0;
  for (int i = sl; i < nq; i += 8) {
    const double u = qv[i] * rho;
    double sn, cs;
    sincos(u, &sn, &cs);
    f = __builtin_fma(Bq[i], __builtin_fma(u, cs, -sn), f);
  }
  part[sl][m] = f;
  __syncthreads();
  if (threadIdx.x < kVjpNodes) {
    double v = 0.0;
    for (int r = 0; r < 8; ++r) v += part[r][threadIdx.x];
    fnode[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x < kVjpNodes) {
    double c = 0.0;
    for (int n = 0; n < kVjpNodes; ++n) c = __builtin_fma(nodes.m[threadIdx.x][n], fnode[n], c);
    coef[size_t(blockIdx.x) * kVjpNodes + threadIdx.x] = float(c);
  }
}

// The pair sweep of debye_vjp_f32 with F(r) from the bin's coefficients
// (the item's type pair's table in LDS: 2 float4 per bin).  No q chunks: the
// FP32 sums of one sweep fold into the FP64 partials once.
__global__ __launch_bounds__(kAdjBlock) void debye_vjp_radial(
    const float4* __restrict__ pts, const int4* __restrict__ blocks /* start, len, type, 0 */,
    const int2* __restrict__ items /* (I, J), I <= J */, int ntypes,
    const float* __restrict__ coef /* [T][T][nk][8] */, int nk, float inv_h, int tsplit,
    double* __restrict__ part /* [items][tsplit][2][3][256] */) {
  extern __shared__ float4 ctab[];
  __shared__ float4 tile[kAdjBlock];
  __shared__ float accj[4][3][kAdjBlock];
  const int tid = threadIdx.x, wave = tid >> 6;
  const int2 it = items[blockIdx.x];
  const int4 bi = blocks[it.x], bj = blocks[it.y];
  const bool diag = it.x == it.y;
  const float4 pi = tid < bi.y ? pts[bi.x + tid] : make_float4(0.f, 0.f, 0.f, 0.f);
  tile[tid] = tid < bj.y ? pts[bj.x + tid] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4* cg =
      reinterpret_cast<const float4*>(coef + (size_t(bi.z) * ntypes + bj.z) * nk * kVjpNodes);
  for (int e = tid; e < 2 * nk; e += kAdjBlock) ctab[e] = cg[e];
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int c = 0; c < 3; ++c) accj[w][c][tid] = 0.f;
  const int ta = int(blockIdx.y) * (kAdjBlock / tsplit), tb = ta + kAdjBlock / tsplit;
  const float two_inv_h = 2.f * inv_h;
  __syncthreads();
  float gi[3] = {0.f, 0.f, 0.f};
  float4 pnext = tile[(ta + tid) & (kAdjBlock - 1)];
  for (int t = ta; t < tb; ++t) {
    const int jj = (t + tid) & (kAdjBlock - 1);
    const float4 pj = pnext;
    pnext = tile[(t + 1 + tid) & (kAdjBlock - 1)];
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const bool use = d2 > 1e-20f && (!diag || jj > tid);
    const float dd = use ? d2 : 1.0f;
    const float ir = __builtin_amdgcn_rsqf(dd);
    const float r0 = dd * ir;
    const float r = __builtin_fmaf(__builtin_fmaf(-r0, r0, dd), 0.5f * ir, r0);
    const float wgt = use ? pi.w * pj.w * ir * ir * ir : 0.f;
    const int k = min(int(r * inv_h), nk - 1);
    const float u = __builtin_fmaf(r, two_inv_h, -float(2 * k + 1));
    const float4 lo = ctab[2 * k], hi = ctab[2 * k + 1];  // c_0..c_3, c_4..c_7
    // Clenshaw: b_p = c_p + 2u b_{p+1} - b_{p+2}; F = c_0 + u b_1 - b_2
    const float u2 = u + u;
    const float b7 = hi.w;
    const float b6 = __builtin_fmaf(u2, b7, hi.z);
    const float b5 = __builtin_fmaf(u2, b6, hi.y - b7);
    const float b4 = __builtin_fmaf(u2, b5, hi.x - b6);
    const float b3 = __builtin_fmaf(u2, b4, lo.w - b5);
    const float b2 = __builtin_fmaf(u2, b3, lo.z - b4);
    const float b1 = __builtin_fmaf(u2, b2, lo.y - b3);
    const float F = __builtin_fmaf(u, b1, lo.x - b2);
    const float hh = F * wgt;
    gi[0] = __builtin_fmaf(hh, dx, gi[0]);
    gi[1] = __builtin_fmaf(hh, dy, gi[1]);
    gi[2] = __builtin_fmaf(hh, dz, gi[2]);
    // slot jj belongs to this lane at this step (see debye_vjp_f32)
    accj[wave][0][jj] -= hh * dx;
    accj[wave][1][jj] -= hh * dy;
    accj[wave][2][jj] -= hh * dz;
  }
  __syncthreads();
  double* out = part + (size_t(blockIdx.x) * gridDim.y + blockIdx.y) * 2 * 3 * kAdjBlock;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    out[c * kAdjBlock + tid] = double(gi[c]);
    out[(3 + c) * kAdjBlock + tid] =
        ((double(accj[0][c][tid]) + double(accj[1][c][tid])) + double(accj[2][c][tid])) +
        double(accj[3][c][tid]);
  }
}

// Position of block pair (I, J), I <= J, in the row-major upper triangle.
__device__ inline size_t item_index(int I, int J, int nb) {
  const size_t i = size_t(I);
  return i * size_t(nb) - (i * (i > 0 ? i - 1 : 0)) / 2 + size_t(J - I);
}

// gacc[B][c][l] += the partials of items [k0, k1) that hold block B: the
// i side of (B, J), J ascending, then the j side of (I, B), I ascending.
__global__ void reduce_vjp(const double* __restrict__ part, size_t k0, size_t k1, int qsplit,
                           int nb, double* __restrict__ gacc) {
  const int Bk = blockIdx.x, tid = threadIdx.x;
  double g[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) g[c] = gacc[(size_t(Bk) * 3 + c) * kAdjBlock + tid];
  for (int side = 0; side < 2; ++side) {
    for (int o = side == 0 ? Bk : 0; o < (side == 0 ? nb : Bk + 1); ++o) {
      const size_t k = side == 0 ? item_index(Bk, o, nb) : item_index(o, Bk, nb);
      if (k < k0 || k >= k1) continue;
      for (int s = 0; s < qsplit; ++s) {
        const double* p = part + (((k - k0) * qsplit + s) * 2 + side) * 3 * kAdjBlock;
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] += p[c * kAdjBlock + tid];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) gacc[(size_t(Bk) * 3 + c) * kAdjBlock + tid] = g[c];
}

__global__ void zero_f64(double* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// out[orig atom][c] = gacc of its block slot.
__global__ void scatter_vjp(const double* __restrict__ gacc, const int4* __restrict__ blocks,
                            const int* __restrict__ orig, double* __restrict__ out) {
  const int Bk = blockIdx.x, tid = threadIdx.x;
  const int4 b = blocks[Bk];
  if (tid >= b.y) return;
  const int o = orig[b.x + tid];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[size_t(o) * 3 + c] = gacc[(size_t(Bk) * 3 + c) * kAdjBlock + tid];
}

template <int QT, int SM>
__global__ __launch_bounds__(kAdjBlock) void debye_jacobian_f32(
    const float4* __restrict__ pts, const int* __restrict__ orig, const int* __restrict__ type,
    int npts, const int* __restrict__ seg_start, int ntypes, const float* __restrict__ qrev,
    const float* __restrict__ Wp /* [ntypes][ntypes][nq_pad] */, int nq_pad,
    double* __restrict__ out /* [nq_pad][npts][3] */) {
  __shared__ float4 tile[kAdjBlock];
  const int i = blockIdx.x * kAdjBlock + threadIdx.x;
  const int qc = blockIdx.y * QT;
  const bool vi = i < npts;
  const float4 pi = vi ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int ta = vi ? type[i] : 0;
  float qk[QT];
#pragma unroll
  for (int k = 0; k < QT; ++k) qk[k] = qrev[qc + k];
  double g64[QT][3];
  float g[QT][3];
#pragma unroll
  for (int k = 0; k < QT; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      g64[k][c] = 0.0;
      g[k][c] = 0.f;
    }
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
        // occ_j (x_i - x_j) / r^3 once per pair; per q the weight, the sine
        // and cosine, and three FMAs
        const float ex = ir3 * dx, ey = ir3 * dy, ez = ir3 * dz;
#pragma unroll
        for (int k = 0; k < QT; ++k) {
          const float v = qk[k] * r;  // revolutions
          float sn, cs;
          sincos_rev<SM>(v, sn, cs);
          const float h = w[k] * __builtin_fmaf(kTwoPi * v, cs, -sn);
          g[k][0] = __builtin_fmaf(h, ex, g[k][0]);
          g[k][1] = __builtin_fmaf(h, ey, g[k][1]);
          g[k][2] = __builtin_fmaf(h, ez, g[k][2]);
        }
      }
#pragma unroll
      for (int k = 0; k < QT; ++k)
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
#pragma unroll
  for (int k = 0; k < QT; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[(size_t(qc + k) * npts + o) * 3 + c] = oi * g64[k][c];
}

constexpr int kJacQT = 8;

}  // namespace

int jacobian_q_tile() { return kJacQT; }

// Workgroups per block pair: a block pair's q chunks are split over up to
// this many workgroups, so that small problems still launch about 16384.
int vjp_qsplit(size_t nqc, size_t nitems) {
  nitems = std::max<size_t>(nitems, 1);
  return static_cast<int>(std::min<size_t>(nqc, std::max<size_t>(1, (16384 + nitems - 1) / nitems)));
}

// q values per pair sweep.  A sweep of a block pair costs about 7 VALU
// cycles per pair.q (one multiply, V_SIN and V_COS at two cycles each, two
// FMAs) plus about 45 per pair (distance, correctly rounded sqrt, weights,
// the three sums), times 1.055 for the tiles of 48 and up (two waves per
// SIMD instead of three; measured at 20k atoms x 300 q: QT 32 / 48 / 60 / 64
// = 13.3 / 14.0 / 12.1 / 12.9 ms).  The modelled time is that cost times
// the sweeps of one workgroup times the rounds of one workgroup per CU:
//   ceil(nqc / qsplit) ceil(nitems qsplit / cus),  nqc = ceil(nq / QT).
// Large structures reduce to the least padded work; small ones, with few
// block pairs, to one sweep per workgroup, which narrow tiles shorten.
// Measured on MI355X over globules of 300 to 20,000 atoms and 50 to 300 q:
// the chosen tile is within 3 % of the fastest.  Under the recurrence (sine
// mode 3) a pair.q costs four FMAs and two adds and a pair about 75 more
// (the recurrence's start adds two series and the sine and cosine of the
// half step), and its tiles are held to max_tile (the pair kernel's limit
// on the phase a tile accumulates).  Sine mode 1 has the tiles
// 16 and 32 only; a tuned tile instantiated for the sine mode is taken as it
// is.
int vjp_q_tile(int tuned, size_t nq, int sin_mode, size_t nitems, int cus, int max_tile) {
  static const int kTiles[] = {16, 32, 48, 60, 64};
  const int ntiles = sin_mode >= 2 ? 5 : 2;
  for (int k = 0; k < ntiles; ++k)
    if (kTiles[k] == tuned && (k == 0 || tuned <= max_tile)) return tuned;
  int best = 16;
  double best_cost = 1e300;
  for (int k = 0; k < ntiles; ++k) {
    const int qt = kTiles[k];
    if (k > 0 && qt > max_tile) continue;
    const size_t nqc = (nq + qt - 1) / qt;
    const size_t qsplit = static_cast<size_t>(vjp_qsplit(nqc, nitems));
    const size_t c = static_cast<size_t>(std::max(cus, 1));
    const double sweeps = double((nqc + qsplit - 1) / qsplit);
    const double rounds = double((nitems * qsplit + c - 1) / c);
    const double cost = (sin_mode == 3 ? 6.0 * qt + 75.0 : (7.0 * qt + 45.0) * (qt >= 48 ? 1.055 : 1.0)) *
                        sweeps * rounds;
    if (cost < best_cost) best = qt, best_cost = cost;
  }
  return best;
}

hipError_t launch_debye_jacobian(const float4* pts, const int* orig, const int* type, int npts,
                                 const int* seg_start, int ntypes, const float* qrev,
                                 const float* Wp, int nq_pad, int sin_mode, double* out,
                                 hipStream_t stream, KernelTimer* timer) {
  if (npts <= 0) return hipSuccess;
  if (nq_pad % kJacQT != 0 || (sin_mode != 1 && sin_mode != 2)) return hipErrorInvalidValue;
  dim3 grid((npts + kAdjBlock - 1) / kAdjBlock, nq_pad / kJacQT);
  TimerScope scope(timer, "debye_jacobian_f32", stream);
  if (sin_mode == 2)
    debye_jacobian_f32<kJacQT, 2><<<grid, kAdjBlock, 0, stream>>>(pts, orig, type, npts, seg_start,
                                                                  ntypes, qrev, Wp, nq_pad, out);
  else
    debye_jacobian_f32<kJacQT, 1><<<grid, kAdjBlock, 0, stream>>>(pts, orig, type, npts, seg_start,
                                                                  ntypes, qrev, Wp, nq_pad, out);
  return hipGetLastError();
}

hipError_t launch_debye_vjp(const float4* pts, const int4* blocks, int nb, const int2* items,
                            size_t nitems, const int* orig, int ntypes, const float* qrev,
                            float qstep, const float* AB, int nq_pad, int qt, int sin_mode,
                            int qsplit, int tsplit, size_t items_per_launch, double* part,
                            double* gacc, double* out, hipStream_t stream, KernelTimer* timer) {
  if (nb <= 0) return hipSuccess;
  const bool hw_only = qt == 48 || qt == 60 || qt == 64;  // sine modes 2 and 3 only
  if ((sin_mode == 3) != (qstep != 0.f) || sin_mode < 1 || sin_mode > 3) return hipErrorInvalidValue;
  if ((qt != 16 && qt != 32 && !(hw_only && sin_mode >= 2)) || nq_pad % qt != 0 || qsplit < 1 ||
      qsplit > nq_pad / qt || items_per_launch == 0 ||
      (tsplit != 1 && tsplit != 2 && tsplit != 4 && tsplit != 8))
    return hipErrorInvalidValue;
  const size_t nacc = size_t(nb) * 3 * kAdjBlock;
  zero_f64<<<(nacc + 255) / 256, 256, 0, stream>>>(gacc, nacc);
  for (size_t k0 = 0; k0 < nitems; k0 += items_per_launch) {
    const size_t k1 = std::min(nitems, k0 + items_per_launch);
    dim3 grid(static_cast<unsigned>(k1 - k0), qsplit * tsplit);
    {
      TimerScope scope(timer, "debye_vjp_f32", stream);
#define CSX_VJP(QT_, SM_)                                                               \
  debye_vjp_f32<QT_, SM_><<<grid, kAdjBlock, 0, stream>>>(pts, blocks, items + k0, ntypes, qrev, \
                                                          qstep, AB, nq_pad, qsplit, tsplit, part)
#define CSX_VJP_MODES(QT_)  \
  if (sin_mode == 3)        \
    CSX_VJP(QT_, 3);        \
  else if (sin_mode == 2)   \
    CSX_VJP(QT_, 2);
      if (qt == 64) {
        CSX_VJP_MODES(64)
      } else if (qt == 60) {
        CSX_VJP_MODES(60)
      } else if (qt == 48) {
        CSX_VJP_MODES(48)
      } else if (qt == 32) {
        CSX_VJP_MODES(32)
        else CSX_VJP(32, 1);
      } else {
        CSX_VJP_MODES(16)
        else CSX_VJP(16, 1);
      }
#undef CSX_VJP_MODES
#undef CSX_VJP
    }
    reduce_vjp<<<nb, kAdjBlock, 0, stream>>>(part, k0, k1, qsplit * tsplit, nb, gacc);
  }
  scatter_vjp<<<nb, kAdjBlock, 0, stream>>>(gacc, blocks, orig, out);
  return hipGetLastError();
}

// Sine mode 4 (see vjp_radial_coefficients): the coefficient tables into
// coef (T * T * nk * 8 floats), then the sweeps and the same reduction.
hipError_t launch_debye_vjp_radial(const float4* pts, const int4* blocks, int nb, const int2* items,
                                   size_t nitems, const int* orig, int ntypes, const double* qv,
                                   const double* B, int nq, double h, int nk, int tsplit,
                                   size_t items_per_launch, float* coef, double* part, double* gacc,
                                   double* out, hipStream_t stream, KernelTimer* timer) {
  if (nb <= 0) return hipSuccess;
  if (!(h > 0.0) || nk <= 0 || nk > kVjpRadialMaxBins || nq <= 0 || ntypes <= 0 ||
      items_per_launch == 0 || (tsplit != 1 && tsplit != 2 && tsplit != 4 && tsplit != 8))
    return hipErrorInvalidValue;
  VjpNodes nodes;
  const double pi = 3.141592653589793;
  for (int m = 0; m < kVjpNodes; ++m) nodes.x[m] = std::cos((2 * m + 1) * pi / (2 * kVjpNodes));
  for (int p = 0; p < kVjpNodes; ++p)
    for (int m = 0; m < kVjpNodes; ++m)
      nodes.m[p][m] = (p == 0 ? 1.0 : 2.0) / kVjpNodes *
                      std::cos(p * (2 * m + 1) * pi / (2 * kVjpNodes));
  vjp_radial_coefficients<<<unsigned(ntypes * ntypes * nk), 64, 0, stream>>>(qv, B, nq, nk, h,
                                                                             nodes, coef);
  const size_t nacc = size_t(nb) * 3 * kAdjBlock;
  zero_f64<<<(nacc + 255) / 256, 256, 0, stream>>>(gacc, nacc);
  const size_t lds = size_t(nk) * kVjpNodes * sizeof(float);
  for (size_t k0 = 0; k0 < nitems; k0 += items_per_launch) {
    const size_t k1 = std::min(nitems, k0 + items_per_launch);
    dim3 grid(static_cast<unsigned>(k1 - k0), tsplit);
    {
      TimerScope scope(timer, "debye_vjp_radial", stream);
      debye_vjp_radial<<<grid, kAdjBlock, lds, stream>>>(pts, blocks, items + k0, ntypes, coef, nk,
                                                         float(1.0 / h), tsplit, part);
    }
    reduce_vjp<<<nb, kAdjBlock, 0, stream>>>(part, k0, k1, tsplit, nb, gacc);
  }
  scatter_vjp<<<nb, kAdjBlock, 0, stream>>>(gacc, blocks, orig, out);
  return hipGetLastError();
}

}  // namespace csx

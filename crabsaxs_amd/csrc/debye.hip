// Debye pair-sum kernels for gfx950 (CDNA4).
//
//  debye_plain_f64   FP64 restatement of the reference pair loop
//                    (scattering.rs:100-106, solvent.rs:533-558): per-atom
//                    amplitudes, r_ij, sin and division for every pair and q.
//                    One workgroup per (q, 256-row block); rows i of the block
//                    sweep all j > i through LDS tiles.  This is the FP64
//                    baseline the fast path is measured against.
//  debye_pairs_f32   Per-type pair tables.  A work item pairs a type-pure block
//                    of 256 "i" atoms (one per lane) with a chunk of "j" atoms
//                    of one type; r_ij is computed once per pair and swept over
//                    QT q values.  The term occ_j sin(q r)/r is accumulated in
//                    FP32 per LDS tile (<= 256 pairs per lane), folded into
//                    FP64 per tile, and the workgroup writes one FP64 partial
//                    per (item, q); a second kernel sums partials per type bin
//                    in a fixed order, so results are bitwise reproducible run
//                    to run.  Sweep modes 0-2 evaluate a sine per pair and q
//                    (q in scalar registers); mode 3, for a uniform q grid,
//                    runs Reinsch's form of the sine recurrence along the
//                    tile's q values (three FMA-class operations per pair and
//                    q instead of a multiply, a two-cycle V_SIN_F32 and an
//                    FMA), restarted from hardware sines at every tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>

#include "kernels.hpp"

namespace csx {
namespace {

constexpr int kBlock = 256;
constexpr int kFloatsPerLaunch = 512;

// Fixed-order FP64 sum over a 256-thread workgroup; valid in thread 0.
template <int NV>
__device__ inline void block_sum_f64(double (&v)[NV], double* scratch /* [4][NV] */) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) scratch[wave * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      v[k] = ((scratch[k] + scratch[NV + k]) + scratch[2 * NV + k]) + scratch[3 * NV + k];
  }
}

__device__ inline double sinc64(double x) { return fabs(x) < 1e-8 ? 1.0 : sin(x) / x; }

// ---------------------------------------------------------------- FP64 --
// NCH amplitude channels per point (1: vacuum f; 3: v, e, h), NT output
// terms (1: I; 6: vv, ee, hh, ve, vh, eh).  amp is [NCH][nq][npts].
template <int NCH, int NT>
__global__ __launch_bounds__(kBlock) void debye_plain_f64(const double4* __restrict__ pts,
                                                          const double* __restrict__ amp,
                                                          const double* __restrict__ qv, int npts,
                                                          int nq, double* __restrict__ partial) {
  __shared__ double4 tile_p[kBlock];
  __shared__ double tile_a[NCH][kBlock];
  __shared__ double scratch[4 * NT];
  const int blk = blockIdx.x, qi = blockIdx.y;
  const int i = blk * kBlock + threadIdx.x;
  const double q = qv[qi];
  const bool vi = i < npts;
  const double4 pi = vi ? pts[i] : make_double4(0, 0, 0, 0);
  double ai[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) ai[c] = vi ? amp[(size_t(c) * nq + qi) * npts + i] : 0.0;
  double acc[NT];
  if (NT == 1) {
    acc[0] = ai[0] * ai[0];
  } else {
    acc[0] = ai[0] * ai[0];
    acc[1] = ai[1] * ai[1];
    acc[2] = ai[2] * ai[2];
    acc[3] = ai[0] * ai[1];
    acc[4] = ai[0] * ai[2];
    acc[5] = ai[1] * ai[2];
  }
  for (int t0 = blk * kBlock; t0 < npts; t0 += kBlock) {
    __syncthreads();
    const int j = t0 + threadIdx.x;
    tile_p[threadIdx.x] = j < npts ? pts[j] : make_double4(0, 0, 0, 0);
#pragma unroll
    for (int c = 0; c < NCH; ++c)
      tile_a[c][threadIdx.x] = j < npts ? amp[(size_t(c) * nq + qi) * npts + j] : 0.0;
    __syncthreads();
    const int jn = min(kBlock, npts - t0);
    for (int t = 0; t < jn; ++t) {
      if (!vi || t0 + t <= i) continue;
      const double4 pj = tile_p[t];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double s = sinc64(q * r);
      if (NT == 1) {
        acc[0] += 2.0 * ai[0] * tile_a[0][t] * s;
      } else {
        const double vj = tile_a[0][t], ej = tile_a[1][t], hj = tile_a[NCH - 1][t];
        acc[0] += 2.0 * ai[0] * vj * s;
        acc[1] += 2.0 * ai[1] * ej * s;
        acc[2] += 2.0 * ai[2] * hj * s;
        acc[3] += (ai[0] * ej + vj * ai[1]) * s;
        acc[4] += (ai[0] * hj + vj * ai[2]) * s;
        acc[5] += (ai[1] * hj + ej * ai[2]) * s;
      }
    }
  }
  block_sum_f64<NT>(acc, scratch);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NT; ++k) partial[(size_t(qi) * gridDim.x + blk) * NT + k] = acc[k];
  }
}

// out[k*nq + qi] = sum over blocks of partial[qi][blk][k], fixed order.
__global__ void reduce_plain_partials(const double* __restrict__ partial, int nblk, int nq, int nt,
                                      double* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nq * nt) return;
  const int qi = idx / nt, k = idx % nt;
  double s = 0.0;
#pragma unroll 32
  for (int b = 0; b < nblk; ++b) s += partial[(size_t(qi) * nblk + b) * nt + k];
  out[size_t(k) * nq + qi] = s;
}

// amp[c][qi][i] = occ_i * table[c][type_i][qi].
__global__ void expand_amplitudes(const double* __restrict__ table, const int* __restrict__ type,
                                  const double* __restrict__ occ, int npts, int ntypes, int nq,
                                  int nch, double* __restrict__ amp) {
  const size_t total = size_t(nch) * nq * npts;
  for (size_t idx = blockIdx.x * size_t(blockDim.x) + threadIdx.x; idx < total;
       idx += size_t(gridDim.x) * blockDim.x) {
    const int i = int(idx % npts);
    const size_t cq = idx / npts;
    const int qi = int(cq % nq), c = int(cq / nq);
    amp[idx] = occ[i] * table[(size_t(c) * ntypes + type[i]) * nq + qi];
  }
}

// ---------------------------------------------------------------- FP32 --

// sin(2 pi u) with an exact reduction in revolutions: g = u - rint(u) lies in
// [-1/2, 1/2] and is exact (|u| < 2^22), so only the product q*r before it is
// rounded; then an odd degree-13 polynomial (max error 1.2e-9 in exact
// arithmetic, fitted to sin(2 pi g) on [-1/2, 1/2]).
__device__ inline float sin2pi_soft(float u) {
  const float g = u - __builtin_rintf(u);
  const float g2 = g * g;
  float p = 3.1996936798095703f;
  p = __builtin_fmaf(p, g2, -14.868611335754395f);
  p = __builtin_fmaf(p, g2, 42.01616668701172f);
  p = __builtin_fmaf(p, g2, -76.70155334472656f);
  p = __builtin_fmaf(p, g2, 81.60502624511719f);
  p = __builtin_fmaf(p, g2, -41.341697692871094f);
  p = __builtin_fmaf(p, g2, 6.2831854820251465f);
  return g * p;
}

// Hardware V_SIN_F32 (argument in revolutions) after the same exact
// reduction.
__device__ inline float sin2pi_hw(float u) {
  const float g = u - __builtin_rintf(u);
  return __builtin_amdgcn_sinf(g);
}

// Hardware V_SIN_F32 on the unreduced argument; its own reduction takes the
// fraction exactly, valid for |u| <= 256 revolutions (checked on the host).
__device__ inline float sin2pi_hw_direct(float u) { return __builtin_amdgcn_sinf(u); }

template <int SINMODE>
__device__ inline float sin2pi(float u) {
  if (SINMODE == 0) return sin2pi_soft(u);
  if (SINMODE == 1) return sin2pi_hw(u);
  return sin2pi_hw_direct(u);
}

// 2 sin(pi t) for |t| <= 1/4 (t in revolutions): the Taylor series to t^9,
// truncation below 3.4e-9 of the value, relative accuracy kept for tiny t
// (the recurrence's step constants need it for short pairs) and no bias:
// the step sets the frequency of every pair's recurrence, and a relative
// bias e in it evaluates each tile at q (1 + e) after k steps, which moves
// I(q) by e dlnI/dlnq where it is steep.
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

// The pairs of lane atom pi with tile[0 .. jn): acc[k] += occ_j / r *
// sin(2 pi qk[k] r).  r = d2 * rsq(d2) refined by one Newton step (the raw
// square-root approximations are biased, which shifts every phase q r the
// same way and measurably moves I(q) where it is steep; the step removes the
// bias at 4 VALU cycles instead of the ~16 of the correctly rounded square
// root); 1/r is the rsq itself.  MASK keeps only j > i (the tile overlapping
// the lane's own block in a diagonal work item).
template <int QT, int SINMODE, bool MASK>
__device__ __attribute__((always_inline)) inline void tile_sweep(const float4* tile, int jn,
                                                                 const float4 pi, int jt_minus_i,
                                                                 const float* qk, float* acc) {
  for (int t = 0; t < jn; ++t) {
    const float4 pj = tile[t];
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float d2 = fmaxf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), 1e-30f);
    const float y = __builtin_amdgcn_rsqf(d2);
    const float r0 = d2 * y;
    const float r = __builtin_fmaf(__builtin_fmaf(-r0, r0, d2), 0.5f * y, r0);
    const float w = (!MASK || t + jt_minus_i > 0) ? pj.w * y : 0.f;
#pragma unroll
    for (int k = 0; k < QT; ++k) {
      const float u = qk[k] * r;
      const float s = sin2pi<SINMODE>(u);
      acc[k] = __builtin_fmaf(w, s, acc[k]);
    }
  }
}

// The same pairs over the uniform q grid q0 + k qstep (revolutions), k < QT.
// Each pair runs Reinsch's form of the sine recurrence on its weighted sine
// s_k = w sin(2 pi (q0 + k qstep) r), w = occ_j / r:
//   d_{k+1} = d_k + lam s_k,  s_{k+1} = s_k + d_{k+1},
//   lam = 2 cos(2 pi th) - 2 = -(2 sin(pi th))^2,  th = qstep r,
// so a q costs one FMA and two adds (acc[k] += s_k), against a multiply, a
// V_SIN_F32 (a quarter-rate transcendental) and an FMA for a direct sweep.
// The form keeps the step difference d_k = s_k - s_{k-1} as its own
// variable: a rounding of s moves the sinusoid's amplitude by that rounding,
// and one of d by d's relative rounding, so the error does not grow as 1/th
// for short pairs the way the three-term form s_{k+1} = 2 cos(th) s_k -
// s_{k-1} does.  Start: s_0 = w sin(2 pi q0 r) and d_0 = s_0 - s_{-1} =
// w 2 sin(pi th) cos(2 pi (q0 r - th / 2)), from V_SIN_F32 / V_COS_F32 on
// the unreduced arguments (the host keeps |q0 r| <= 250 revolutions) and
// the step constant from its own series (the host keeps |th| <= 1/4).
// Masked pairs (j <= i) get w = 0 and add exact zeros.
template <int QT, bool MASK>
__device__ __attribute__((always_inline)) inline void tile_sweep_rec(const float4* tile, int jn,
                                                                     const float4 pi,
                                                                     int jt_minus_i, float q0,
                                                                     float qstep, float* acc) {
  for (int t = 0; t < jn; ++t) {
    const float4 pj = tile[t];
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float d2 = fmaxf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), 1e-30f);
    const float y = __builtin_amdgcn_rsqf(d2);
    const float r0 = d2 * y;
    const float r = __builtin_fmaf(__builtin_fmaf(-r0, r0, d2), 0.5f * y, r0);
    const float w = (!MASK || t + jt_minus_i > 0) ? pj.w * y : 0.f;
    const float th = qstep * r;
    const float h2 = two_sin_pi_small(th);
    const float lam = -(h2 * h2);
    const float u0 = q0 * r;
    float s = w * __builtin_amdgcn_sinf(u0);
    float d = (w * h2) * __builtin_amdgcn_cosf(__builtin_fmaf(-0.5f, th, u0));
#pragma unroll
    for (int k = 0; k < QT; ++k) {
      acc[k] += s;
      if (k + 1 < QT) {
        d = __builtin_fmaf(lam, s, d);
        s += d;
      }
    }
  }
}

template <int QT, int SINMODE, bool MASK>
__device__ __attribute__((always_inline)) inline void sweep(const float4* tile, int jn,
                                                            const float4 pi, int jt_minus_i,
                                                            const float* qk, float qstep,
                                                            float* acc) {
  if constexpr (SINMODE == 3)
    tile_sweep_rec<QT, MASK>(tile, jn, pi, jt_minus_i, qk[0], qstep, acc);
  else
    tile_sweep<QT, SINMODE, MASK>(tile, jn, pi, jt_minus_i, qk, acc);
}

constexpr int pow2_tile(int n) { return n <= 8 ? 8 : n <= 16 ? 16 : n <= 32 ? 32 : 64; }

// Wave sums of N <= 64 lane values acc[0 .. N), each weighted by wi first,
// and acc cleared: a transposing butterfly.  Each step exchanges half of a
// lane's values with its partner at xor offset o and keeps the half that
// the lane's bit o selects, so every value meets the same partners, in the
// same order (offsets 32 .. 1), as a full xor butterfly per value, at about
// one shuffle per value instead of six.  Lane (64 / pow2_tile(N)) * k ends
// with value k's sum.
template <int N>
__device__ __attribute__((always_inline)) inline float wave_tile_sums(float* acc, float wi,
                                                                      int lane) {
  constexpr int kPow = pow2_tile(N);
  constexpr int kStride = 64 / kPow;
  float red[kPow];
#pragma unroll
  for (int k = 0; k < kPow; ++k) {
    red[k] = k < N ? acc[k] * wi : 0.f;
    if (k < N) acc[k] = 0.f;
  }
#pragma unroll
  for (int o = 32, h = kPow / 2; o >= kStride; o >>= 1, h >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int k = 0; k < h; ++k) {
      // The empty asm hands the selects register values, not array reads:
      // otherwise the compiler folds them into one read at a lane-dependent
      // index, and the array leaves registers for compare-select chains.
      float lo = red[k], hi = red[h + k];
      asm("" : "+v"(lo), "+v"(hi));
      const float keep = up ? hi : lo;
      const float send = up ? lo : hi;
      red[k] = keep + __shfl_xor(send, o, 64);
    }
  }
  float v = red[0];
#pragma unroll
  for (int o = kStride / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One work item (type-pure block of 256 i atoms x chunk of j atoms) for the
// QT q values starting at qc; writes partial_row[qc .. qc + QT).  QT <= 128:
// the tile sums go through the wave reduction in one or two groups of <= 64.
template <int QT, int SINMODE>
__device__ __attribute__((always_inline)) inline void pair_item(const float4* __restrict__ pts, const int4 item, const int qc,
                                 const float* __restrict__ qrev, float qstep,
                                 double* __restrict__ partial_row) {
  static_assert(QT >= 1 && QT <= 128, "q tile");
  static_assert(SINMODE == 3 || QT <= 64, "direct sweeps hold q in registers");
  constexpr int kN0 = QT < 64 ? QT : 64;
  constexpr int kN1 = QT - kN0;
  constexpr int kStride0 = 64 / pow2_tile(kN0);
  constexpr int kStride1 = kN1 > 0 ? 64 / pow2_tile(kN1) : 1;
  constexpr int kQk = SINMODE == 3 ? 1 : QT;
  __shared__ float4 tile[kBlock];
  __shared__ double scratch[4 * QT];
  const int i = item.x + threadIdx.x;
  const bool vi = threadIdx.x < item.y;
  const float4 pi = vi ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  float qk[kQk];
#pragma unroll
  for (int k = 0; k < kQk; ++k) qk[k] = qrev[qc + k];
  // Per lane: FP32 tile sums acc[k] (<= 256 pairs).  After each tile they are
  // weighted by occ_i, summed over the wave, and the lane holding q_k's wave
  // sum adds it to its FP64 running sum (run0 for the first 64 q, run1 for
  // the rest).
  float acc[QT];
#pragma unroll
  for (int k = 0; k < QT; ++k) acc[k] = 0.f;
  double run0 = 0.0, run1 = 0.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int jend = item.z + item.w;
  // This wave's first i; a wave with no atoms of the block, or whose atoms
  // all lie at or past a diagonal tile's last j (every pair masked), skips
  // the tile's sweep: its pairs would add exact zeros.
  const int wave_i0 = item.x + 64 * wave;
  const bool wave_live = 64 * wave < item.y;
  for (int t0 = item.z; t0 < jend; t0 += kBlock) {
    __syncthreads();
    const int j = t0 + threadIdx.x;
    tile[threadIdx.x] = j < jend ? pts[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int jn = min(kBlock, jend - t0);
    // only a tile overlapping the i block needs the j > i mask
    if (t0 < item.x + item.y && t0 + kBlock > item.x) {
      if (wave_live && wave_i0 < t0 + jn - 1)
        sweep<QT, SINMODE, true>(tile, jn, pi, t0 - i, qk, qstep, acc);
    } else if (wave_live) {
      sweep<QT, SINMODE, false>(tile, jn, pi, t0 - i, qk, qstep, acc);
    }
    const float v0 = wave_tile_sums<kN0>(acc, pi.w, lane);
    if ((lane & (kStride0 - 1)) == 0 && lane / kStride0 < kN0) run0 += double(v0);
    if constexpr (kN1 > 0) {
      const float v1 = wave_tile_sums<kN1>(acc + kN0, pi.w, lane);
      if ((lane & (kStride1 - 1)) == 0 && lane / kStride1 < kN1) run1 += double(v1);
    }
  }
  if ((lane & (kStride0 - 1)) == 0 && lane / kStride0 < kN0)
    scratch[wave * QT + lane / kStride0] = run0;
  if constexpr (kN1 > 0) {
    if ((lane & (kStride1 - 1)) == 0 && lane / kStride1 < kN1)
      scratch[wave * QT + kN0 + lane / kStride1] = run1;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < QT; k += kBlock)
    partial_row[qc + k] =
        ((scratch[k] + scratch[QT + k]) + scratch[2 * QT + k]) + scratch[3 * QT + k];
}

// Waves per SIMD the kernel is compiled for: the recurrence's wide tiles
// (up to 128 accumulators) are held to three (168 VGPRs), so that other
// waves cover each wave's dependent FMA-add steps.
template <int SINMODE>
constexpr int pair_min_waves() {
  return SINMODE == 3 ? 3 : 1;
}

// One workgroup per (work item, q chunk).
template <int QT, int SINMODE>
__global__ __launch_bounds__(kBlock, pair_min_waves<SINMODE>()) void debye_pairs_f32(const float4* __restrict__ pts,
                                                          const int4* __restrict__ items,
                                                          const float* __restrict__ qrev,
                                                          float qstep, int nq_pad,
                                                          double* __restrict__ partial) {
  pair_item<QT, SINMODE>(pts, items[blockIdx.x], blockIdx.y * QT, qrev, qstep,
                         partial + size_t(blockIdx.x) * nq_pad);
}

// Persistent form: a fixed grid of workgroups takes (work item, q chunk)
// tasks from a global counter until all `total` tasks are done; every
// workgroup exits once the counter passes `total`.
template <int QT, int SINMODE>
__global__ __launch_bounds__(kBlock, pair_min_waves<SINMODE>()) void debye_pairs_f32_persistent(
    const float4* __restrict__ pts, const int4* __restrict__ items, const float* __restrict__ qrev,
    float qstep, int nq_pad, int total, int* __restrict__ counter, double* __restrict__ partial) {
  __shared__ int task;
  const int nqc = nq_pad / QT;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) task = atomicAdd(counter, 1);
    __syncthreads();
    const int w = task;
    if (w >= total) break;
    const int it = w / nqc;
    pair_item<QT, SINMODE>(pts, items[it], (w - it * nqc) * QT, qrev, qstep,
                           partial + size_t(it) * nq_pad);
  }
}

__global__ void reset_counter(int* counter) { *counter = 0; }

// dst[0 .. n) = v[0 .. n): values handed over as a kernel argument, so that
// host data reaches the device without a staging allocation.
struct FloatBlock {
  float v[kFloatsPerLaunch];
};
__global__ void write_floats(float* __restrict__ dst, int n, FloatBlock b) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = b.v[k];
}

// tables[b * nq + k] = raw[b * nq_pad + k] / (2 pi q_k), or q0[b] where
// that is 0 (the occupancy products of the q = 0 row), q_k the q (in
// revolutions) the sweep evaluated: qrev[k] for a direct sweep, and for the
// recurrence (qstep != 0) its tile's first value plus (k mod qt) steps.  The
// host's final step of the pair tables, in the same operations.
__global__ void finish_tables(const double* __restrict__ raw, const float* __restrict__ qrev,
                              int qt, float qstep, const double* __restrict__ q0, int nbins,
                              int nq, int nq_pad, double* __restrict__ tables) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= nq || b >= nbins) return;
  const int kt = k % qt;
  const double qk =
      qstep != 0.f ? double(qrev[k - kt]) + double(kt) * double(qstep) : double(qrev[k]);
  const double qe = 2.0 * 3.14159265358979323846 * qk;
  tables[size_t(b) * nq + k] = qe == 0.0 ? q0[b] : raw[size_t(b) * nq_pad + k] / qe;
}

// out[b][q] = sum_{items of bin b, in order} partial[item][q].
__global__ void reduce_bins(const double* __restrict__ partial, const int* __restrict__ bin_first,
                            int nq_pad, double* __restrict__ out) {
  const int b = blockIdx.x;
  const int qi = blockIdx.y * blockDim.x + threadIdx.x;
  if (qi >= nq_pad) return;
  double s = 0.0;
  // unrolled so that 32 independent loads issue ahead of the in-order sum
#pragma unroll 32
  for (int it = bin_first[b]; it < bin_first[b + 1]; ++it) s += partial[size_t(it) * nq_pad + qi];
  out[size_t(b) * nq_pad + qi] = s;
}

}  // namespace

// ------------------------------------------------------------- launchers --

// ------------------------------------------------ no-allocation scratch --

// Device memory of the no-allocation pair-table entry point: the launch's
// task counter, q in revolutions, item partials and raw tables.  Every GPU
// that loads the library holds one copy until the process ends; calls on
// one GPU share it, so calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;  // device -> address of `scratch`

hipError_t noalloc_scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return hipSuccess;
  int current = -1;
  hipError_t e = hipGetDevice(&current);
  if (e != hipSuccess) return e;
  if (current != device) return hipErrorInvalidDevice;  // the caller makes it current
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}

void* noalloc_scratch(int device, size_t* bytes) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (bytes) *bytes = kNoAllocScratchBytes;
  return it == g_scratch_ptr.end() ? nullptr : it->second;
}

hipError_t launch_write_floats(float* dst, const double* q, int n, int n_total, double divisor,
                               hipStream_t stream) {
  for (int k0 = 0; k0 < n_total; k0 += kFloatsPerLaunch) {
    FloatBlock b;
    const int m = n_total - k0 < kFloatsPerLaunch ? n_total - k0 : kFloatsPerLaunch;
    for (int k = 0; k < kFloatsPerLaunch; ++k)
      b.v[k] = k < m && k0 + k < n ? static_cast<float>(q[k0 + k] / divisor) : 0.f;
    write_floats<<<(m + 255) / 256, 256, 0, stream>>>(dst + k0, m, b);
  }
  return hipGetLastError();
}

hipError_t launch_finish_tables(const double* raw, const float* qrev, int qt, float qstep,
                                const double* q0, int nbins, int nq, int nq_pad, double* tables,
                                hipStream_t stream) {
  if (nbins <= 0 || nq <= 0) return hipSuccess;
  if (qt <= 0) return hipErrorInvalidValue;
  finish_tables<<<dim3((nq + 255) / 256, nbins), 256, 0, stream>>>(raw, qrev, qt, qstep, q0,
                                                                  nbins, nq, nq_pad, tables);
  return hipGetLastError();
}

hipError_t launch_debye_plain(const double4* pts, const double* amp, const double* q, int npts,
                              int nq, int nch, int nterms, double* partial, double* out,
                              hipStream_t stream, KernelTimer* timer) {
  const int nblk = (npts + kBlock - 1) / kBlock;
  dim3 grid(nblk, nq);
  TimerScope scope(timer, "debye_plain_f64", stream);
  if (nch == 1 && nterms == 1)
    debye_plain_f64<1, 1><<<grid, kBlock, 0, stream>>>(pts, amp, q, npts, nq, partial);
  else if (nch == 3 && nterms == 6)
    debye_plain_f64<3, 6><<<grid, kBlock, 0, stream>>>(pts, amp, q, npts, nq, partial);
  else
    return hipErrorInvalidValue;
  scope.stop();
  const int total = nq * nterms;
  reduce_plain_partials<<<(total + 255) / 256, 256, 0, stream>>>(partial, nblk, nq, nterms, out);
  return hipGetLastError();
}

hipError_t launch_expand_amplitudes(const double* table, const int* type, const double* occ,
                                    int npts, int ntypes, int nq, int nch, double* amp,
                                    hipStream_t stream) {
  const size_t total = size_t(nch) * nq * npts;
  const int blocks = int(std::min<size_t>((total + 255) / 256, 65536));
  expand_amplitudes<<<blocks, 256, 0, stream>>>(table, type, occ, npts, ntypes, nq, nch, amp);
  return hipGetLastError();
}

template <int QT, int SM>
static void launch_pairs_t(int nitems, int persist_wgs, int* counter, const float4* pts,
                           const int4* items, const float* qrev, float qstep, int nq_pad,
                           double* partial, hipStream_t stream) {
  const int nqc = nq_pad / QT;
  if (persist_wgs > 0) {
    const int total = nitems * nqc;
    reset_counter<<<1, 1, 0, stream>>>(counter);
    debye_pairs_f32_persistent<QT, SM><<<std::min(persist_wgs, total), kBlock, 0, stream>>>(
        pts, items, qrev, qstep, nq_pad, total, counter, partial);
  } else {
    debye_pairs_f32<QT, SM><<<dim3(nitems, nqc), kBlock, 0, stream>>>(pts, items, qrev, qstep,
                                                                       nq_pad, partial);
  }
}

namespace {

// ------------------------------------------ radial interpolation sweep --
// Pair tables from the pair-distance distribution: for a pair at distance r
// in radial bin k = floor(r / h), j_0(q r) is replaced by its interpolant at
// the bin's 8 Chebyshev nodes rho_km = (k + 1/2 + x_m / 2) h,
// x_m = cos((2m + 1) pi / 16), whose error is at most (q h / 2)^8 / (2^7 8!)
// (1.0e-9 at q h = 1.04; every derivative of j_0 is bounded by 1).  So
//   sum_pairs occ_i occ_j sin(q r) / r = sum_km A_km sin(q rho_km) / rho_km,
//   A_km = sum over the bin's pairs of occ_i occ_j L_m(u), u = 2 r / h - 2k - 1,
// with L_m the Lagrange basis of the nodes.  By the discrete orthogonality
// of the Chebyshev polynomials at these nodes,
//   L_m(u) = sum_p (2 - [p = 0]) / 8 T_p(x_m) T_p(u),
// so a pair adds its Chebyshev moments c T_p(u), p = 0..7 (|T_p| <= 1, from
// the three-term recurrence: 8 VALU operations for the 8), and the 8 x 8 map
// to the nodes runs once per bin after the sums.  Each moment is rounded to
// a fixed-point integer (times 2^F, F per point set) and added to 64-bit
// integer accumulators, first per work item in LDS and then per table bin in
// global memory: integer addition does not depend on the order, so the
// tables are the same bits in every run, alone or in a batch.
constexpr int kRadNodes = 8;
struct RadialMap {
  double x[kRadNodes];             // Chebyshev nodes on [-1, 1]
  double m[kRadNodes][kRadNodes];  // node m's weight of moment p: (2 - [p = 0]) / 8 T_p(x_m)
};

__global__ __launch_bounds__(kBlock) void debye_pairs_radial(
    const float4* __restrict__ pts, const int4* __restrict__ items,
    const int* __restrict__ bin_first, int nbins, const float* __restrict__ bin_scale, float inv_h,
    int nk, int copies, unsigned long long* __restrict__ acc) {
  // [kRadNodes][nk][copies], lane l adding to copy l mod copies: the copies
  // of one (moment, bin) are neighbouring words, so lanes of different copies
  // never meet in a bank whatever their bins, and lanes of one copy meet only
  // where their bins agree modulo (banks / copies)
  extern __shared__ unsigned long long lacc[];
  __shared__ float4 tile[kBlock];
  __shared__ int sbin;
  const int4 item = items[blockIdx.x];
  if (threadIdx.x == 0) {  // the item's table bin: the last b with bin_first[b] <= item
    int lo = 0, hi = nbins - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bin_first[mid] <= int(blockIdx.x)) lo = mid; else hi = mid - 1;
    }
    sbin = lo;
  }
  for (int e = threadIdx.x; e < nk * kRadNodes * copies; e += kBlock) lacc[e] = 0ull;
  __syncthreads();
  unsigned long long* const mine = lacc + threadIdx.x % copies;
  const int node_stride = nk * copies;
  const int bin = sbin;
  const float fscale = bin_scale[bin];
  const int i = item.x + threadIdx.x;
  const bool vi = threadIdx.x < item.y;
  const float4 pi = vi ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float ci = pi.w * fscale;
  const int jend = item.z + item.w;
  const float two_inv_h = 2.f * inv_h;
  for (int t0 = item.z; t0 < jend; t0 += kBlock) {
    __syncthreads();
    const int j = t0 + threadIdx.x;
    tile[threadIdx.x] = j < jend ? pts[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    if (!vi) continue;
    const int jn = min(kBlock, jend - t0);
    for (int t = 0; t < jn; ++t) {
      if (t0 + t <= i) continue;  // each unordered pair once (j > i)
      const float4 pj = tile[t];
      const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const float d2 = fmaxf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), 1e-30f);
      const float y = __builtin_amdgcn_rsqf(d2);
      const float r0 = d2 * y;
      const float r = __builtin_fmaf(__builtin_fmaf(-r0, r0, d2), 0.5f * y, r0);
      const int k = min(int(r * inv_h), nk - 1);
      const float u = __builtin_fmaf(r, two_inv_h, -float(2 * k + 1));
      // c T_p(u): T_0 = 1, T_1 = u, T_{p+1} = 2u T_p - T_{p-1}
      const float u2 = u + u;
      float s0 = ci * pj.w;
      float s1 = s0 * u;
      unsigned long long* const slot = mine + k * copies;
      atomicAdd(slot, static_cast<unsigned long long>(static_cast<long long>(__float2int_rn(s0))));
      atomicAdd(slot + node_stride,
                static_cast<unsigned long long>(static_cast<long long>(__float2int_rn(s1))));
#pragma unroll
      for (int p = 2; p < kRadNodes; ++p) {
        const float s2 = __builtin_fmaf(u2, s1, -s0);
        atomicAdd(slot + p * node_stride,
                  static_cast<unsigned long long>(static_cast<long long>(__float2int_rn(s2))));
        s0 = s1;
        s1 = s2;
      }
    }
  }
  __syncthreads();
  unsigned long long* out = acc + size_t(bin) * nk * kRadNodes;  // [nk][kRadNodes]
  for (int e = threadIdx.x; e < nk * kRadNodes; e += kBlock) {
    const unsigned long long* c = lacc + ((e % kRadNodes) * nk + e / kRadNodes) * copies;
    unsigned long long v = 0ull;
    for (int r = 0; r < copies; ++r) v += c[r];
    if (v != 0ull) atomicAdd(out + e, v);
  }
}

// a_km = 2^-F_b A_km / rho_km in FP64, A_km = sum_p map[m][p] (moment p), in
// place of the integer moment sums (one thread per (bin, k): [bin][k][8]).
__global__ void radial_node_weights(unsigned long long* __restrict__ acc, size_t nbk, int nk,
                                    double h, RadialMap map,
                                    const double* __restrict__ bin_unscale) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nbk) return;
  const int k = int(e % size_t(nk));
  const double unscale = bin_unscale[e / size_t(nk)];
  double t[kRadNodes];
#pragma unroll
  for (int p = 0; p < kRadNodes; ++p)
    t[p] = double(static_cast<long long>(acc[e * kRadNodes + p]));
  double* out = reinterpret_cast<double*>(acc) + e * kRadNodes;
#pragma unroll
  for (int m = 0; m < kRadNodes; ++m) {
    double a = 0.0;
#pragma unroll
    for (int p = 0; p < kRadNodes; ++p) a = __builtin_fma(map.m[m][p], t[p], a);
    out[m] = a * unscale / ((double(k) + 0.5 + 0.5 * map.x[m]) * h);
  }
}

// out[b][q] = sum_km a_km sin(2 pi q rho_km) in FP64, q in revolutions (the
// pair tables' raw form: the host divides by 2 pi q).  A workgroup owns 64 q
// of one bin.  The bins k come in ranges of kNodeRange; each range's sum
// rotates the 8 nodes' phasors from one sincos at its first bin by the step
// angle 2 pi q h (an FP64 rotation: after n steps the phasor is off by ~n
// ulp).  Wave w adds ranges w, w + kNodeWaves, ... in order and the waves'
// sums are added in wave order: the bits depend on the bins' accumulators
// only, not on how many empty bins follow them (a batch gives its sets the
// largest set's bin count).
constexpr int kNodeWaves = 4;
constexpr int kNodeRange = 16;
__global__ __launch_bounds__(64 * kNodeWaves) void radial_node_sums(
    const double* __restrict__ a, int nk, double h, RadialMap map,
    const float* __restrict__ qrev, int nq_pad, double* __restrict__ out) {
  __shared__ double part[kNodeWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const int b = blockIdx.y;
  const double w = q < nq_pad ? 2.0 * 3.141592653589793 * double(qrev[q]) : 0.0;
  const double* ab = a + size_t(b) * nk * kRadNodes;
  double s = 0.0;
  if (w != 0.0) {
    double sd, cd;
    sincos(w * h, &sd, &cd);
    for (int k0 = wave * kNodeRange; k0 < nk; k0 += kNodeWaves * kNodeRange) {
      const int k1 = min(nk, k0 + kNodeRange);
      double sn[kRadNodes], cs[kRadNodes];
#pragma unroll
      for (int m = 0; m < kRadNodes; ++m)
        sincos(w * ((double(k0) + 0.5 + 0.5 * map.x[m]) * h), &sn[m], &cs[m]);
      double r = 0.0;
      for (int k = k0; k < k1; ++k) {
#pragma unroll
        for (int m = 0; m < kRadNodes; ++m) {
          r = __builtin_fma(ab[k * kRadNodes + m], sn[m], r);
          const double t = __builtin_fma(sn[m], cd, cs[m] * sd);
          cs[m] = __builtin_fma(cs[m], cd, -(sn[m] * sd));
          sn[m] = t;
        }
      }
      s += r;
    }
  }
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && q < nq_pad) {
    double v = part[0][lane];
    for (int r = 1; r < kNodeWaves; ++r) v += part[r][lane];
    out[size_t(b) * nq_pad + q] = v;
  }
}

__global__ void zero_u64_array(unsigned long long* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

int radial_node_count() { return kRadNodes; }

size_t radial_lds_bytes(int nk) { return size_t(nk) * kRadNodes * sizeof(unsigned long long); }

// Accumulator copies per workgroup: the requested power of two (1-32), or by
// default 8, fewer where they do not fit beside the point tile.  (20,000
// atoms x 500 q: 0.61 ms at 8 copies, 1.22 ms at 16 or 32, whose LDS leaves
// room for fewer workgroups per CU.)
static int radial_copies(int nk, size_t lds, int want) {
  int c = want > 0 ? std::min(want, 32) : 8;
  while (c > 1 && c * radial_lds_bytes(nk) + 4104 > lds) c >>= 1;
  return c;
}

hipError_t launch_debye_pairs_radial(const float4* pts, const int4* items, int nitems,
                                     const int* bin_first, int nbins, const float* bin_scale,
                                     const double* bin_unscale, double h, int nk,
                                     const float* qrev, int nq_pad, int copies_wanted,
                                     unsigned long long* acc, double* out, size_t lds,
                                     hipStream_t stream, KernelTimer* timer) {
  if (!(h > 0.0) || nk <= 0 || nbins <= 0 || nbins > 65535 || radial_lds_bytes(nk) + 4104 > lds ||
      copies_wanted < 0 || (copies_wanted & (copies_wanted - 1)) != 0)
    return hipErrorInvalidValue;
  const int copies = radial_copies(nk, lds, copies_wanted);
  RadialMap map;
  const double pi = 3.141592653589793;
  for (int m = 0; m < kRadNodes; ++m) {
    map.x[m] = std::cos((2 * m + 1) * pi / (2 * kRadNodes));
    for (int p = 0; p < kRadNodes; ++p)
      map.m[m][p] = (p == 0 ? 1.0 : 2.0) / kRadNodes * std::cos(p * (2 * m + 1) * pi / (2 * kRadNodes));
  }
  const size_t nbk = size_t(nbins) * nk;
  const size_t nacc = nbk * kRadNodes;
  zero_u64_array<<<unsigned((nacc + 255) / 256), 256, 0, stream>>>(acc, nacc);
  {
    TimerScope scope(timer, "debye_pairs_radial", stream);
    if (nitems > 0)
      debye_pairs_radial<<<nitems, kBlock, copies * radial_lds_bytes(nk), stream>>>(
          pts, items, bin_first, nbins, bin_scale, float(1.0 / h), nk, copies, acc);
  }
  radial_node_weights<<<unsigned((nbk + 255) / 256), 256, 0, stream>>>(acc, nbk, nk, h, map,
                                                                       bin_unscale);
  radial_node_sums<<<dim3((nq_pad + 63) / 64, nbins), 64 * kNodeWaves, 0, stream>>>(
      reinterpret_cast<const double*>(acc), nk, h, map, qrev, nq_pad, out);
  return hipGetLastError();
}

hipError_t launch_debye_pairs(const float4* pts, const int4* items, int nitems, const int* bin_first,
                              int nbins, const float* qrev, float qstep, int nq_pad, int qt,
                              int sin_mode, int persist_wgs, int* counter, double* partial,
                              double* out, hipStream_t stream, KernelTimer* timer) {
  if (qt <= 0 || nq_pad % qt != 0) return hipErrorInvalidValue;
  if ((sin_mode == 3) != (qstep != 0.f)) return hipErrorInvalidValue;
  if (persist_wgs > 0 && (!counter || static_cast<long long>(nitems) * (nq_pad / qt) > INT32_MAX))
    persist_wgs = 0;
  {
    TimerScope scope(timer, "debye_pairs_f32", stream);
    if (nitems > 0) {
#define CSX_PAIRS_CASE(QT_, SM_)                                                              \
  else if (qt == QT_ && sin_mode == SM_) launch_pairs_t<QT_, SM_>(nitems, persist_wgs, counter, \
                                                                  pts, items, qrev, qstep,     \
                                                                  nq_pad, partial, stream);
      if (false) {
      }
      CSX_PAIRS_CASE(8, 0) CSX_PAIRS_CASE(8, 1) CSX_PAIRS_CASE(8, 2)
      CSX_PAIRS_CASE(16, 0) CSX_PAIRS_CASE(16, 1) CSX_PAIRS_CASE(16, 2)
      CSX_PAIRS_CASE(32, 0) CSX_PAIRS_CASE(32, 1) CSX_PAIRS_CASE(32, 2)
      CSX_PAIRS_CASE(64, 0) CSX_PAIRS_CASE(64, 1) CSX_PAIRS_CASE(64, 2)
      CSX_PAIRS_CASE(24, 2) CSX_PAIRS_CASE(40, 2) CSX_PAIRS_CASE(48, 2) CSX_PAIRS_CASE(50, 2)
      CSX_PAIRS_CASE(56, 2) CSX_PAIRS_CASE(63, 2)
      CSX_PAIRS_CASE(10, 2) CSX_PAIRS_CASE(20, 2) CSX_PAIRS_CASE(25, 2)
      CSX_PAIRS_CASE(16, 3) CSX_PAIRS_CASE(25, 3) CSX_PAIRS_CASE(32, 3) CSX_PAIRS_CASE(50, 3)
      CSX_PAIRS_CASE(64, 3) CSX_PAIRS_CASE(100, 3) CSX_PAIRS_CASE(125, 3) CSX_PAIRS_CASE(128, 3)
      else return hipErrorInvalidValue;
#undef CSX_PAIRS_CASE
    }
    scope.stop();
  }
  if (nbins > 0) {
    dim3 rgrid(nbins, (nq_pad + 255) / 256);
    reduce_bins<<<rgrid, 256, 0, stream>>>(partial, bin_first, nq_pad, out);
  }
  return hipGetLastError();
}

}  // namespace csx

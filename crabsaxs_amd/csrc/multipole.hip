// Spherical-multipole kernels for gfx950 (spherical.rs:41-565 restated for
// the GPU).
//
//  multipole_geometry  once per (point set, l_max): FP64 real spherical
//                      harmonics by the associated-Legendre recurrence of
//                      spherical.rs:54-93, stored FP32 as MFMA B operands.
//                      Each l owns ceil((2l+1)/16) column tiles of 16 (column
//                      c of l is m = c - l; padding is 0).  A tile is stored
//                      in slabs of 16 points, each slab 256 floats laid out so
//                      that lane 16k + n of a wave reads its four k-step values
//                      (points 4 ks + k, column n) with one 16-byte load.
//  multipole_coeff     S_t[q][lm] = sum_{i of type t} occ_i j_l(q r_i) Y_lm,i
//                      as FP32 MFMA (v_mfma_f32_16x16x4f32: 16 q rows x 16
//                      columns, k = 4 points).  A workgroup owns a chunk of up
//                      to 512 points of one type, a block of 16 q and a group
//                      of column tiles (4 waves x MAXT accumulators).  The
//                      type's amplitudes factor out of its sum, so they are
//                      applied after it (multipole_contract).  Per slab of 16 points,
//                      each thread generates occ_i j_l(q r_i) for one
//                      (point, q) and every l of the group into LDS (FP32
//                      sin/cos; FP32 upward recurrence for l <= q r; FP64
//                      Miller downward recurrence from l + 14 above it,
//                      normalised on the upward value, two passes so no step
//                      divides), then the waves run the MFMAs of their tiles.  Tiles above the
//                      block's largest per-q order are skipped.  Partial
//                      coefficients per chunk are written in FP32.
//  multipole_sum_types FP64 sum of the chunk partials of each type, in chunk
//                      order: the per-type coefficient tables S_t[q][lm],
//                      which depend on the points, q grid and l_max only.
//  multipole_contract  per q, the channel coefficients C_c = sum_t A_c,t(q)
//                      S_t (FP64 amplitudes) and the quadratic forms
//                      sum_lm C_a C_b (1 or 6 per q) over l <= that q's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"

namespace csx {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kSlab = 16;    // points per slab (4 MFMA k-steps of 4)
constexpr int kQB = 16;      // q values per MFMA row block
constexpr int kThreads = 256;
constexpr int kChunk = 512;  // points per workgroup

// Offset of (point p of a slab, column n) inside a 256-float tile slab: lane
// 16 k + n holds the values of points 4 ks + k, ks = 0..3, contiguously.
__device__ inline int slab_offset(int p, int n) { return ((p & 3) * 16 + n) * 4 + (p >> 2); }

// j_l operands of one slab in LDS: 256 floats per l, q-major.  The value for
// (q, point 4 ks + k) sits at q * 16 + 4 xslot(q, k) + ks, so the MFMA lane
// (k, n = q) reads its four k-step values with one 16-byte load.  The slot
// is k XOR two bits of q: the 16-byte reads of 16 consecutive lanes (64
// banks) or of 8 (32 banks) then fall on distinct banks, and a generating
// wave (4 q x 16 points) stores a permutation of 64 consecutive words.
constexpr int kXL = 256;  // floats per l
__device__ inline int xslot(int q, int k) { return k ^ (((q >> 1) ^ (q >> 3)) & 3); }

// ------------------------------------------------------------ geometry --
__global__ void multipole_geometry(const double* __restrict__ xyz, int npts,
                                   const double* __restrict__ centre3, int l_max,
                                   const double* __restrict__ norms, const int* __restrict__ tile_first,
                                   int npad, float* __restrict__ radius, float* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const double px = xyz[3 * i] - centre3[0], py = xyz[3 * i + 1] - centre3[1],
               pz = xyz[3 * i + 2] - centre3[2];
  const double r = sqrt(px * px + py * py + pz * pz);
  radius[i] = float(r);
  double x = 1.0, phi = 0.0;
  if (!(r < 1e-12)) {
    x = fmin(fmax(pz / r, -1.0), 1.0);
    phi = atan2(py, px);
  }
  const double st = sqrt(fmax(1.0 - x * x, 0.0));
  const double sp = sin(phi), cp = cos(phi);
  double sm = 0.0, cm = 1.0, pmm = 1.0;
  const double sqrt2 = 1.4142135623730951;
  // Y(l, column) with column = l + m (cos part) or l - m (sin part).
  const int nslab = npad / kSlab;
  auto store = [&](int l, int col, double v) {
    const int t = tile_first[l] + col / 16;
    Y[(size_t(t) * nslab + i / kSlab) * 256 + slab_offset(i % kSlab, col & 15)] = float(v);
  };
  for (int m = 0; m <= l_max; ++m) {
    if (m > 0) {
      pmm *= -double(2 * m - 1) * st;
      const double nc = cm * cp - sm * sp;
      sm = sm * cp + cm * sp;
      cm = nc;
    }
    auto emit = [&](int l, double p) {
      const double nrm = norms[l * l + l + m];
      if (m == 0) {
        store(l, l, nrm * p);
      } else {
        const double v = sqrt2 * nrm * p;
        store(l, l + m, v * cm);
        store(l, l - m, v * sm);
      }
    };
    emit(m, pmm);
    if (m == l_max) continue;
    const double pm1 = x * double(2 * m + 1) * pmm;
    emit(m + 1, pm1);
    double prev = pmm, cur = pm1;
    for (int l = m + 2; l <= l_max; ++l) {
      const double next = (double(2 * l - 1) * x * cur - double(l + m - 1) * prev) / double(l - m);
      emit(l, next);
      prev = cur;
      cur = next;
    }
  }
}


__global__ void fill_zero(float* __restrict__ p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// 1 / a for a positive normal double whose reciprocal is normal: the
// hardware estimate refined by two Newton steps (within an ulp; the correctly
// rounded division sequence costs twice the instructions).
__device__ inline double recip(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  return fma(r, fma(-a, r, 1.0), r);
}

// One (point, q) column of spherical Bessel values: X[(l - l_lo) * kXL] =
// j_l(x) for l in [l_lo, min(lq, l_hi)] and 0 for the rest of [l_lo, l_hi].
// l <= floor(x): FP32 upward recurrence from j_0, j_1 (formed in FP64 from
// the FP32 sine and cosine; the recurrence is stable there).  Above:
// Miller's downward recurrence f_{l-1} = (2l+1)/x f_l - f_{l+1} started at
// S = top + 14 (f = 1e-250, 0) and normalised by f at floor(x) against the
// upward value.  Pass 1 runs it down to floor(x) for the normalisation and
// keeps the state at `top`; pass 2 resumes from that state and stores.
// Where the growth bound ((2S+1)/x)^(S - floor(x)) stays below 2^1800 the
// recurrence cannot leave the double range from 1e-250 and runs without
// overflow tests.  Otherwise f is rescaled by 1e-250 on overflow and values
// computed before the last rescale (below 1e-250 of the normalising value)
// are stored as 0.  Every stored value is the FP32 j_l times `scale` (one
// FP32 multiply; 1 leaves it unchanged).
__device__ inline void bessel_column(float xf, int lq, int l_lo, int l_hi, float scale, float* X) {
  const int top = min(lq, l_hi);
  int l_zero = l_lo;  // first l to clear
  if (top >= l_lo) {
    if (!(xf >= 1e-8f)) {
      if (l_lo == 0) X[0] = scale;
      l_zero = max(l_lo, 1);
    } else {
      const double x = double(xf);
      float sf, cf;
      sincosf(xf, &sf, &cf);
      const double ix = recip(x);
      const double j0 = double(sf) * ix;
      const double j1 = (double(sf) * ix - double(cf)) * ix;
      const int lu = min(top, int(floorf(xf)));
      if (l_lo == 0) X[0] = float(j0) * scale;
      double jlu = j0;
      if (lu >= 1) {
        if (l_lo <= 1) X[(1 - l_lo) * kXL] = float(j1) * scale;
        // FP32 upward recurrence; (2l + 1) / x is formed from the exact odd
        // integer every step, so the coefficient carries one rounding.
        const float ixf = float(ix);
        float jm = float(j0), jc = float(j1), odd = 3.f;
        for (int l = 1; l < lu; ++l) {
          const float jn = (odd * ixf) * jc - jm;
          jm = jc;
          jc = jn;
          odd += 2.f;
          if (l + 1 >= l_lo) X[(l + 1 - l_lo) * kXL] = jn * scale;
        }
        jlu = double(jc);
      }
      if (top > lu) {
        const int S = top + 14;
        const int lo = max(lu + 1, l_lo);
        const bool safe = float(S - lu) * (__log2f(float(2 * S + 1)) - __log2f(xf)) < 1800.f;
        if (safe) {
          // state at index l: f = f_l, fp = f_{l+1}, c = (2l + 1) / x.  The
          // loops take two steps per trip, f and fp trading roles, so the
          // state needs no register moves.
          const double d = 2.0 * ix;
          double fp = 0.0, f = 1e-250, c = double(2 * S + 1) * ix;
#pragma unroll
          for (int k = 0; k < 14; k += 2) {  // S down to top
            fp = c * f - fp;
            c -= d;
            f = c * fp - f;
            c -= d;
          }
          const double f_top = f, fp_top = fp, c_top = c;
          int l = top;
          for (; l - 1 > lu; l -= 2) {
            fp = c * f - fp;
            c -= d;
            f = c * fp - f;
            c -= d;
          }
          if (l > lu) f = c * f - fp;
          // f_lu > 0 (j_l(x) > 0 for l >= floor(x)), and on this path it is at
          // most 2^1800 times 1e-250, so its reciprocal is a normal double.
          const double ns = jlu * recip(f) * double(scale);
          f = f_top;
          fp = fp_top;
          c = c_top;
          for (l = top; l - 1 >= lo; l -= 2) {
            X[(l - l_lo) * kXL] = float(ns * f);
            fp = c * f - fp;
            c -= d;
            X[(l - 1 - l_lo) * kXL] = float(ns * fp);
            f = c * fp - f;
            c -= d;
          }
          if (l >= lo) X[(l - l_lo) * kXL] = float(ns * f);
        } else {
          // pass 1: normalisation and the last rescale
          double fp = 0.0, f = 1e-250, c = double(2 * S + 1) * ix;
          int last = S + 1;
          for (int l = S; l > lu; --l) {
            const double fn = c * f - fp;
            fp = f;
            f = fn;
            c -= 2.0 * ix;
            if (fabs(f) > 1e250) {
              f *= 1e-250;
              fp *= 1e-250;
              last = l - 1;
            }
          }
          const double norm = jlu / f;
          // pass 2: the same recurrence, storing l in [lo, top]
          fp = 0.0;
          f = 1e-250;
          c = double(2 * S + 1) * ix;
          for (int l = S; l >= lo; --l) {
            if (l <= top) X[(l - l_lo) * kXL] = l <= last ? float(norm * f) * scale : 0.f;
            const double fn = c * f - fp;
            fp = f;
            f = fn;
            c -= 2.0 * ix;
            if (fabs(f) > 1e250) {
              f *= 1e-250;
              fp *= 1e-250;
            }
          }
        }
      }
      l_zero = top + 1;
    }
  }
  for (int l = max(l_zero, l_lo); l <= l_hi; ++l) X[(l - l_lo) * kXL] = 0.f;
}

// partial[(chunk * nqp + q) * ncol + col]: the sum over the chunk's points
// (all of one type) of occ j_l Y.
// At least four waves per SIMD (128 registers).
template <int MAXT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void multipole_coeff(
    const float* __restrict__ radius, const float* __restrict__ occ,
    const int4* __restrict__ chunks /* (first slab, end slab, type) */, int nchunk, int nslab,
    const float* __restrict__ Y, int ntiles, const int* __restrict__ tile_l,
    const float* __restrict__ qv, const int* __restrict__ qlmax, int nqp, int ngroups,
    float* __restrict__ partial) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // XCD-aware order: workgroups go to the 8 XCDs round robin, so give each
  // XCD a contiguous run of logical ids; the q blocks of one point chunk
  // (which read the same harmonic slabs) then share an L2.
  const int nqb = nqp / kQB;
  const int total = nqb * nchunk * ngroups;
  const int b = blockIdx.x;
  const int per = total / 8, extra = total % 8, xcd = b % 8;
  const int logical = xcd * per + min(xcd, extra) + b / 8;
  const int qb = logical % nqb;
  const int chunk = (logical / nqb) % nchunk;
  const int grp = logical / (nqb * nchunk);
  const int q0 = qb * kQB;
  int Lb = 0;
  for (int k = 0; k < kQB; ++k) Lb = max(Lb, qlmax[q0 + k]);
  const int g0 = grp * 4 * MAXT;
  const int g1 = min(g0 + 4 * MAXT, ntiles);
  const int l_lo = tile_l[g0];
  if (l_lo > Lb) return;  // every tile of the group is above this block's orders
  const int l_hi = min(tile_l[g1 - 1], Lb);
  const int lr = l_hi - l_lo + 1;
  const int4 ck = chunks[chunk];
  const int s0 = ck.x, s1 = ck.y;
  float* X = smem;                   // [lr][16 q][4 slots][4 ks]
  float* PR = X + lr * kXL;          // [kChunk] radius and occupancy of the
  float* PO = PR + kChunk;           //   chunk's points
  for (int k = threadIdx.x; k < (s1 - s0) * kSlab; k += kThreads) {
    PR[k] = radius[s0 * kSlab + k];
    PO[k] = occ[s0 * kSlab + k];
  }

  v4f acc[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};

  // generation role: wave w takes q = 4 w + lane / 16 and point lane % 16 of
  // the slab.  Points are sorted by radius, so the lanes of a wave see
  // nearly equal q r and run the recurrences with little divergence.
  const int gq = 4 * wave + (lane >> 4), gp = lane & 15;
  const float qg = qv[q0 + gq];
  const int lg = qlmax[q0 + gq];
  // MFMA role: lane = 16 k + n
  const int mk = lane >> 4, mn = lane & 15;

  // This wave's tiles are g0 + wave + 4 t.  tile_l is non-decreasing, so the
  // live ones (inside the group and at most l_hi) are t < nt; xoff[t] is the
  // LDS offset of tile t's order.  Both are wave-uniform.
  int nt = 0;
  int xoff[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    const int tile = g0 + wave + 4 * t;
    const int l = tile < g1 ? tile_l[tile] : l_hi + 1;
    if (l <= l_hi) nt = t + 1;
    xoff[t] = (min(l, l_hi) - l_lo) * kXL + mn * 16 + 4 * xslot(mn, mk);
  }
  // Harmonic tile rows read by this wave; tiles past the live ones read the
  // last live (or, with none, the last existing) tile, so every load in the
  // loop is unconditional and in bounds.
  size_t yrow[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t)
    yrow[t] = size_t(min(g0 + wave + 4 * max(min(t, nt - 1), 0), ntiles - 1)) * nslab * 256;

  // Harmonic tiles are read in batches of kB: the loads of batch b + 1 are in
  // flight while batch b runs its MFMAs.
  constexpr int kB = 2, kNB = MAXT / kB;
  static_assert(MAXT % kB == 0, "whole batches of tiles");
  for (int s = s0; s < s1; ++s) {
    __syncthreads();  // the previous slab's operands have been read
    // j_l times the point's occupancy (padding points have occupancy 0)
    const int k = (s - s0) * kSlab + gp;
    bessel_column(qg * PR[k], lg, l_lo, l_hi, PO[k],
                  X + gq * 16 + 4 * xslot(gq, gp & 3) + (gp >> 2));
    __syncthreads();
    const float* Ys = Y + size_t(s) * 256 + lane * 4;
    v4f yb[2][kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) yb[0][u] = *reinterpret_cast<const v4f*>(Ys + yrow[u]);
#pragma unroll
    for (int b = 0; b < kNB; ++b) {
      if (b + 1 < kNB) {
#pragma unroll
        for (int u = 0; u < kB; ++u)
          yb[(b + 1) & 1][u] = *reinterpret_cast<const v4f*>(Ys + yrow[(b + 1) * kB + u]);
      }
      // both tiles' A operands are read before either tile's MFMAs (offsets
      // clamped to the group's orders, so the reads are in bounds)
      v4f xa[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) xa[u] = *reinterpret_cast<const v4f*>(X + xoff[b * kB + u]);
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int t = b * kB + u;
        if (t >= nt) continue;
        const v4f a = xa[u];
        const v4f y = yb[b & 1][u];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], y[ks], acc[t], 0, 0, 0);
      }
    }
  }
  const int ncol = ntiles * 16;
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    if (t >= nt) continue;
    const int tile = g0 + wave + 4 * t;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = q0 + 4 * mk + r;
      partial[(size_t(chunk) * nqp + q) * ncol + tile * 16 + mn] = acc[t][r];
    }
  }
}

// coef[(t * nqp + q) * ncol + col] = FP64 sum, in chunk order, of the
// partial coefficients of type t's chunks; 0 for the columns with
// l > qlmax[q] (not written for this q's block).  One thread per (q, column).
__global__ __launch_bounds__(256) void multipole_sum_types(const float* __restrict__ partial,
                                                           const int* __restrict__ type_chunk,
                                                           const int* __restrict__ tile_l,
                                                           const int* __restrict__ qlmax, int nqp,
                                                           int ncol, double* __restrict__ coef) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y, t = blockIdx.z;
  if (col >= ncol) return;
  double v = 0.0;
  if (tile_l[col / 16] <= qlmax[q]) {
    const int c1 = type_chunk[t + 1];
#pragma unroll 16
    for (int ch = type_chunk[t]; ch < c1; ++ch)
      v += double(partial[(size_t(ch) * nqp + q) * ncol + col]);
  }
  coef[(size_t(t) * nqp + q) * ncol + col] = v;
}

// out[q * NT + k], NCH = 1 -> sum C^2; NCH = 3 -> vv, ee, hh, ve, vh, eh,
// over the columns with l <= qlmax[q], where C_c = sum_t amp[(t nqp + q) NCH
// + c] coef_t (the types in order).  One workgroup per q.
template <int NCH>
__global__ __launch_bounds__(256) void multipole_contract(const double* __restrict__ coef,
                                                          const double* __restrict__ amp,
                                                          int ntypes,
                                                          const int* __restrict__ tile_l,
                                                          const int* __restrict__ qlmax,
                                                          int nqp, int ncol,
                                                          double* __restrict__ out) {
  constexpr int NT = NCH == 1 ? 1 : 6;
  constexpr int kMaxTypes = 256;
  __shared__ double A[kMaxTypes * NCH];
  __shared__ double scratch[4 * NT];
  const int q = blockIdx.x;
  const int lq = qlmax[q];
  for (int k = threadIdx.x; k < ntypes * NCH; k += blockDim.x)
    A[k] = amp[(size_t(k / NCH) * nqp + q) * NCH + k % NCH];
  __syncthreads();
  double s[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) s[k] = 0.0;
  for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
    if (tile_l[col / 16] > lq) continue;
    double cc[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) cc[c] = 0.0;
    for (int t = 0; t < ntypes; ++t) {
      const double v = coef[(size_t(t) * nqp + q) * ncol + col];
#pragma unroll
      for (int c = 0; c < NCH; ++c) cc[c] += A[t * NCH + c] * v;
    }
    if (NCH == 1) {
      s[0] += cc[0] * cc[0];
    } else {
      s[0] += cc[0] * cc[0];
      s[1] += cc[1 % NCH] * cc[1 % NCH];
      s[2] += cc[2 % NCH] * cc[2 % NCH];
      s[3] += cc[0] * cc[1 % NCH];
      s[4] += cc[0] * cc[2 % NCH];
      s[5] += cc[1 % NCH] * cc[2 % NCH];
    }
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    double x = s[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    s[k] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NT; ++k) scratch[wave * NT + k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NT; ++k)
      out[size_t(q) * NT + k] =
          ((scratch[k] + scratch[NT + k]) + scratch[2 * NT + k]) + scratch[3 * NT + k];
}

// Accumulator tiles per wave (4 x MAXT tiles per tile group): 16, or 8
// where 8 still covers the tiles in as many groups.  Fewer accumulators
// raise occupancy; more groups generate j_l once more each.  At 108k points
// x 200 q, MAXT 8 vs 16: L = 15 (one group either way) 0.46 vs 0.52 ms;
// L = 25 (two groups vs one) 0.99 vs 0.81 ms.
constexpr int kMaxT = 16;

}  // namespace

int multipole_chunk() { return kChunk; }

int multipole_tiles_per_group(int ntiles) {
  const int big = 4 * kMaxT, small = big / 2;
  return (ntiles + small - 1) / small == (ntiles + big - 1) / big ? small : big;
}

size_t multipole_coeff_lds_bytes(int l_range) {
  return (size_t(l_range) * kXL + 2 * kChunk) * sizeof(float);
}

int multipole_max_types() { return 256; }

hipError_t launch_multipole_geometry(const double* xyz, int npts, const double* centre,
                                     int l_max, const double* norms, const int* tile_first,
                                     int ntiles, int npad, float* radius, float* Y,
                                     hipStream_t stream) {
  if (npts <= 0) return hipSuccess;
  if (npad % kSlab != 0) return hipErrorInvalidValue;
  const size_t ny = size_t(ntiles) * npad * 16;
  fill_zero<<<(ny + 255) / 256, 256, 0, stream>>>(Y, ny);
  multipole_geometry<<<(npts + 127) / 128, 128, 0, stream>>>(xyz, npts, centre, l_max, norms,
                                                             tile_first, npad, radius, Y);
  return hipGetLastError();
}

hipError_t launch_multipole_coeff(const float* radius, const float* occ, const int* chunks,
                                  int nchunk, int npad, const float* Y, int ntiles,
                                  const int* tile_l, int max_l_range, const int* type_chunk,
                                  int ntypes, const float* qv, const int* qlmax, int nqp,
                                  float* partial, double* coef, hipStream_t stream,
                                  KernelTimer* timer) {
  if (nqp % kQB != 0 || npad % kSlab != 0 || ntiles <= 0 || nchunk <= 0 || ntypes <= 0)
    return hipErrorInvalidValue;
  const int tpg = multipole_tiles_per_group(ntiles);
  const int ngroups = (ntiles + tpg - 1) / tpg;
  const long long total = (long long)(nqp / kQB) * nchunk * ngroups;
  if (total > INT32_MAX) return hipErrorInvalidValue;
  const size_t lds = multipole_coeff_lds_bytes(max_l_range);
  const int4* ck = reinterpret_cast<const int4*>(chunks);
  const int nslab = npad / kSlab;
  {
    TimerScope scope(timer, "multipole_coeff", stream);
    if (tpg == 4 * kMaxT)
      multipole_coeff<kMaxT><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, partial);
    else
      multipole_coeff<kMaxT / 2><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, partial);
  }
  const int ncol = ntiles * 16;
  multipole_sum_types<<<dim3((ncol + 255) / 256, nqp, ntypes), 256, 0, stream>>>(
      partial, type_chunk, tile_l, qlmax, nqp, ncol, coef);
  return hipGetLastError();
}

hipError_t launch_multipole_contract(const double* coef, const double* amp, int ntypes, int nch,
                                     const int* tile_l, const int* qlmax, int nqp, int ncol,
                                     double* out, hipStream_t stream) {
  if (ntypes <= 0 || ntypes > multipole_max_types() || (nch != 1 && nch != 3))
    return hipErrorInvalidValue;
  if (nch == 1)
    multipole_contract<1><<<nqp, 256, 0, stream>>>(coef, amp, ntypes, tile_l, qlmax, nqp, ncol, out);
  else
    multipole_contract<3><<<nqp, 256, 0, stream>>>(coef, amp, ntypes, tile_l, qlmax, nqp, ncol, out);
  return hipGetLastError();
}

}  // namespace csx

// Spherical-multipole kernels for gfx950 (spherical.rs:41-565 restated for
// the GPU).
//
//  multipole_geometry  once per (point set, l_max): FP64 real spherical
//                      harmonics by the associated-Legendre recurrence of
//                      spherical.rs:54-93, stored as FP16 (hi, lo) pairs in
//                      MFMA B-operand fragments.  Each l owns ceil((2l+1)/16)
//                      column tiles of 16 (column c of l is m = c - l; padding
//                      is 0).  A tile is stored in slabs of 32 points.
//  multipole_point_terms, multipole_bin_moments, multipole_moment_tiles
//                      the radial-interpolation geometry, once per (point
//                      set, l_max, bin width h): occ_i Y_lm,i per point and
//                      its Lagrange weights at the 8 Chebyshev nodes of its
//                      radial bin; per (bin, lm) the node moments (FP64, in
//                      point order); the moments as FP16 (hi, lo) tiles.  The
//                      coefficient kernel then takes the nodes as its points.
//  multipole_coeff     S_t[q][lm] = sum_{i of type t} occ_i j_l(q r_i) Y_lm,i
//                      on v_mfma_f32_16x16x32_f16 (16 q rows x 16 columns,
//                      k = 32 points), each FP32 product formed from FP16
//                      halves as hi*hi + hi*lo + lo*hi.  A workgroup owns a
//                      chunk of up to 512 points of one type, a block of 16 q
//                      and a group of column tiles (4 waves x MAXT
//                      accumulators).  The type's amplitudes factor out of its
//                      sum, so they are applied after it (multipole_contract).
//                      Per slab of 32 points, each thread generates occ_i
//                      j_l(q r_i) for two neighbouring points of one q and
//                      every l of the group into LDS (FP32 sin/cos; FP32
//                      upward recurrence for l <= q r; FP64 Miller downward
//                      recurrence from l + 14 above it, normalised on the
//                      upward value, two passes so no step divides), then the
//                      waves run the MFMAs of their tiles.  Tiles above the
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
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int kSlab = 32;    // points per slab: the k = 32 of one MFMA
constexpr int kQB = 16;      // q values per MFMA row block
constexpr int kThreads = 256;
constexpr int kChunk = 512;  // points per workgroup (16 slabs)
constexpr int kFrag = 512;   // FP16 values in one operand fragment (64 lanes x 8)
constexpr int kNodes = 8;    // Chebyshev nodes per radial bin

// FP32 operands as FP16 pairs x = hi + lo (22 significant bits together).
// The harmonics are stored times 2^12 (|Y_lm| <= sqrt((2l + 1) / 4 pi) stays
// below the FP16 maximum for l < 1600) and the j_l occ operands times
// 2^kAExp, chosen per launch from the largest occupancy, so that all but
// negligible values are normal FP16 numbers.  The dropped lo*lo term is
// ~2^-22 of a product; the sums are scaled back exactly when written.
constexpr int kYExp = 12;

// Operand fragment of one slab: lane 16 g + c holds row (A: q) or column
// (B: l, m) c at the points 8 g .. 8 g + 7 of the slab, the MFMA's k =
// 8 g + j.  In memory, point p of row/column c is FP16 number frag_half.
__device__ inline int frag_half(int p, int c) { return (16 * (p >> 3) + c) * 8 + (p & 7); }

// In LDS the 16-byte fragment of lane 16 g + c sits at slot 16 g + (c ^ 4g):
// the MFMA reads of 16 lanes (one g) cover 16 consecutive slots, and a
// generating wave (4 rows x 16 point pairs) writes 64 distinct banks.
__device__ inline int lds_slot(int g, int c) { return 16 * g + (c ^ (4 * g)); }

// Two FP32 values (points 2k and 2k + 1 of one row) as packed FP16 pairs:
// hi keeps the leading 11 significant bits (exact in FP16 from 2^-14 up), lo
// is the FP16 (toward zero) of the remainder.
__device__ inline void split_pair(float v0, float v1, uint32_t& hi, uint32_t& lo) {
  const float h0 = __uint_as_float(__float_as_uint(v0) & 0xffffe000u);
  const float h1 = __uint_as_float(__float_as_uint(v1) & 0xffffe000u);
  hi = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(h0, h1));
  lo = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(v0 - h0, v1 - h1));
}

// ------------------------------------------------------------ geometry --
// The real harmonics of one point (offset p from the centre) by the
// associated-Legendre recurrence of spherical.rs:54-93, in FP64:
// store(l, column, Y) with column = l + m for the cos part and l - m for the
// sin part (m > 0), column l for m = 0.  Returns the radius.
template <class Store>
__device__ inline double point_harmonics(double px, double py, double pz, int l_max,
                                         const double* __restrict__ norms, Store store) {
  const double r = sqrt(px * px + py * py + pz * pz);
  double x = 1.0, phi = 0.0;
  if (!(r < 1e-12)) {
    x = fmin(fmax(pz / r, -1.0), 1.0);
    phi = atan2(py, px);
  }
  const double st = sqrt(fmax(1.0 - x * x, 0.0));
  const double sp = sin(phi), cp = cos(phi);
  double sm = 0.0, cm = 1.0, pmm = 1.0;
  const double sqrt2 = 1.4142135623730951;
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
  return r;
}

// Per point: its radius and its harmonics as FP16 (hi, lo) pairs times
// 2^kYExp in the column tiles.
__global__ void multipole_geometry(const double* __restrict__ xyz, int npts,
                                   const double* __restrict__ centre3, int l_max,
                                   const double* __restrict__ norms, const int* __restrict__ tile_first,
                                   int npad, float* __restrict__ radius, _Float16* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const int nslab = npad / kSlab;
  const double r = point_harmonics(
      xyz[3 * i] - centre3[0], xyz[3 * i + 1] - centre3[1], xyz[3 * i + 2] - centre3[2], l_max,
      norms, [&](int l, int col, double v) {
        const int t = tile_first[l] + col / 16;
        const double y = v * double(1 << kYExp);
        const _Float16 hi = static_cast<_Float16>(static_cast<float>(y));
        const _Float16 lo = static_cast<_Float16>(static_cast<float>(y - double(hi)));
        const size_t h = (size_t(t) * nslab + i / kSlab) * (2 * kFrag) + frag_half(i % kSlab, col & 15);
        Y[h] = hi;
        Y[h + kFrag] = lo;
      });
  radius[i] = float(r);
}

// ------------------------------------------- radial interpolation nodes --
// Radial bins of width h (from the centre) hold consecutive points of one
// type.  Within a bin, j_l(q r) is replaced by its interpolant at the
// kNodes Chebyshev nodes of the bin: sum_i occ_i j_l(q r_i) Y_lm,i =
// sum_m j_l(q rho_m) M_m,lm with the moments M_m,lm = sum_i occ_i w_m(r_i)
// Y_lm,i (w_m: the Lagrange basis of the nodes).  |d^n j_l(x) / dx^n| <= 1,
// so the error is at most (q h / 2)^n / (2^(n-1) n!) per point (1.0e-9 at
// n = 8, q h = 1.04).

// Chebyshev nodes on [-1, 1] and the Lagrange denominators prod_{k != m}
// (x_m - x_k), from the host (multipole_chebyshev_nodes).
struct NodeTable {
  double x[kNodes], den[kNodes];
};

// The terms of the radial-interpolation geometry: occ_i Y_lm,i in FP64 in
// tiles of 16 points, terms[((i - p0) / 16 * nlm + lm) * 16 + (i - p0) % 16],
// so that a wave's stores of one (l, m) over 64 consecutive points and a
// wave's loads of 64 consecutive lm for one tile are both contiguous.
__device__ inline size_t term_index(int k, int lm, int nlm) {
  return (size_t(k >> 4) * nlm + lm) * 16 + (k & 15);
}

// One thread per (point, m) of the sorted real points [p0, p1): the m chain
// of the associated-Legendre recurrence (spherical.rs:54-93, as
// point_harmonics), times occ_i; the m = 0 threads also write the point's
// kNodes Lagrange weights at u = (r - centre of its bin) / (h / 2).
__global__ void multipole_point_terms(const double* __restrict__ xyz, const double* __restrict__ occ,
                                      const int* __restrict__ bin_of, int p0, int p1,
                                      const double* __restrict__ centre3, int l_max,
                                      const double* __restrict__ norms, double h, NodeTable nodes,
                                      double* __restrict__ terms, double* __restrict__ weights) {
  const int i = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (i >= p1) return;
  const int nlm = (l_max + 1) * (l_max + 1);
  const int k = i - p0;
  const double px = xyz[3 * i] - centre3[0], py = xyz[3 * i + 1] - centre3[1],
               pz = xyz[3 * i + 2] - centre3[2];
  const double r = sqrt(px * px + py * py + pz * pz);
  double x = 1.0, phi = 0.0;
  if (!(r < 1e-12)) {
    x = fmin(fmax(pz / r, -1.0), 1.0);
    phi = atan2(py, px);
  }
  const double st = sqrt(fmax(1.0 - x * x, 0.0));
  double pmm = 1.0;
  for (int a = 1; a <= m; ++a) pmm *= -double(2 * a - 1) * st;
  double sm = 0.0, cm = 1.0;
  if (m > 0) sincos(double(m) * phi, &sm, &cm);
  const double o = occ[i];
  const double sqrt2 = 1.4142135623730951;
  auto emit = [&](int l, double p) {
    const double nrm = norms[l * l + l + m];
    if (m == 0) {
      terms[term_index(k, l * l + l, nlm)] = o * nrm * p;
    } else {
      const double v = o * sqrt2 * nrm * p;
      terms[term_index(k, l * l + l + m, nlm)] = v * cm;
      terms[term_index(k, l * l + l - m, nlm)] = v * sm;
    }
  };
  emit(m, pmm);
  if (m < l_max) {
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
  if (m != 0) return;
  const double u = (r - (double(bin_of[i]) + 0.5) * h) / (0.5 * h);
  double pre[kNodes + 1], suf[kNodes + 1];
  pre[0] = 1.0;
  suf[kNodes] = 1.0;
#pragma unroll
  for (int a = 0; a < kNodes; ++a) pre[a + 1] = pre[a] * (u - nodes.x[a]);
#pragma unroll
  for (int a = kNodes - 1; a >= 0; --a) suf[a] = suf[a + 1] * (u - nodes.x[a]);
#pragma unroll
  for (int a = 0; a < kNodes; ++a) weights[size_t(k) * kNodes + a] = pre[a] * suf[a + 1] / nodes.den[a];
}

// moments[(v0 + m) * nlm + lm] = sum over the bin's points, in order, of
// weight_m * term_lm.  One thread per (bin, lm): per 16-point tile of the
// bin, one contiguous 128-byte row of terms (the wave's 64 lm together read
// 8 KB); points of a tile outside the bin are skipped.  bins: (first point,
// end point, first virtual point, 0), all inside [p0, ...).
__global__ void multipole_bin_moments(const double* __restrict__ terms,
                                      const double* __restrict__ weights,
                                      const int4* __restrict__ bins, int p0, int nlm,
                                      double* __restrict__ moments) {
  __shared__ double wt[16 * kNodes];  // the tile's weights
  const int lm = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = lm < nlm;
  const int4 b = bins[blockIdx.y];
  double acc[kNodes];
#pragma unroll
  for (int m = 0; m < kNodes; ++m) acc[m] = 0.0;
  const int k0 = b.x - p0, k1 = b.y - p0;
  for (int t = k0 >> 4; t <= (k1 - 1) >> 4; ++t) {
    __syncthreads();  // the previous tile's weights have been read
    for (int e = threadIdx.x; e < 16 * kNodes; e += blockDim.x) {
      const int kk = t * 16 + e / kNodes;
      wt[e] = kk >= k0 && kk < k1 ? weights[size_t(kk) * kNodes + e % kNodes] : 0.0;
    }
    __syncthreads();
    if (!live) continue;
    const double2* row = reinterpret_cast<const double2*>(terms + (size_t(t) * nlm + lm) * 16);
    double v[16];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double2 d = row[u];
      v[2 * u] = d.x;
      v[2 * u + 1] = d.y;
    }
    // the tile's points in order; the bounds are the same for every lane
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int kk = t * 16 + u;
      if (kk < k0 || kk >= k1) continue;
#pragma unroll
      for (int m = 0; m < kNodes; ++m) acc[m] = fma(wt[u * kNodes + m], v[u], acc[m]);
    }
  }
  if (!live) return;
#pragma unroll
  for (int m = 0; m < kNodes; ++m) moments[size_t(b.z + m) * nlm + lm] = acc[m];
}

// *out = max |v| over v[0, n) as the bit pattern of a non-negative double
// (such patterns order like the values); *out starts at 0.
__global__ void abs_max_f64(const double* __restrict__ v, size_t n,
                            unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    m = fmax(m, fabs(v[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __double_as_longlong(m));
}

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }

// The moments of virtual point v as FP16 (hi, lo) pairs times 2^yexp in the
// column tiles (the B operand layout of multipole_geometry).
__global__ void multipole_moment_tiles(const double* __restrict__ moments, int nvirt, int l_max,
                                       const int* __restrict__ tile_first, int npad, int yexp,
                                       _Float16* __restrict__ Y) {
  const int nlm = (l_max + 1) * (l_max + 1);
  const int lm = blockIdx.x * blockDim.x + threadIdx.x;
  if (lm >= nlm) return;
  int l = int(sqrtf(float(lm)));
  while (l * l > lm) --l;
  while ((l + 1) * (l + 1) <= lm) ++l;
  const int col = lm - l * l;
  const int nslab = npad / kSlab;
  const int t = tile_first[l] + col / 16;
  for (int v = blockIdx.y; v < nvirt; v += gridDim.y) {
    const double y = ldexp(moments[size_t(v) * nlm + lm], yexp);
    const _Float16 hi = static_cast<_Float16>(static_cast<float>(y));
    const _Float16 lo = static_cast<_Float16>(static_cast<float>(y - double(hi)));
    const size_t hh = (size_t(t) * nslab + v / kSlab) * (2 * kFrag) + frag_half(v % kSlab, col & 15);
    Y[hh] = hi;
    Y[hh + kFrag] = lo;
  }
}

__global__ void fill_zero(_Float16* __restrict__ p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = static_cast<_Float16>(0.f);
}

// 1 / a for a positive normal double whose reciprocal is normal: the
// hardware estimate refined by two Newton steps (within an ulp; the correctly
// rounded division sequence costs twice the instructions).
__device__ inline double recip(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  return fma(r, fma(-a, r, 1.0), r);
}

// Miller start above the highest stored order.
constexpr int kMillerExtra = 14;

// Whether the Miller recurrence from 1e-250 at S down to lu stays inside the
// double range without tests: its growth bound ((2S+1)/x)^(S - lu) is below
// 2^1800.
__device__ inline bool miller_safe(int S, int lu, float x) {
  return float(S - lu) * (__log2f(float(2 * S + 1)) - __log2f(x)) < 1800.f;
}

// One (point, q) column of spherical Bessel values: put(l, j_l(x) scale)
// for l in [l_lo, min(lq, l_hi)] and put(l, 0) for the rest of [l_lo, l_hi].
// l <= floor(x): FP32 upward recurrence from j_0, j_1 (formed in FP64 from
// the FP32 sine and cosine; the recurrence is stable there).  Above:
// Miller's downward recurrence f_{l-1} = (2l+1)/x f_l - f_{l+1} started at
// S = top + 14 (f = 1e-250, 0) and normalised by f at floor(x) against the
// upward value.  Pass 1 runs it down to floor(x) for the normalisation and
// keeps the state at `top`; pass 2 resumes from that state and stores.
// Where miller_safe holds the recurrence runs without overflow tests.
// Otherwise f is rescaled by 1e-250 on overflow and values computed before
// the last rescale (below 1e-250 of the normalising value) are stored as 0.
// Every stored value is the FP32 j_l times `scale` (one FP32 multiply).
template <class Put>
__device__ inline void bessel_column(float xf, int lq, int l_lo, int l_hi, float scale, Put put) {
  const int top = min(lq, l_hi);
  int l_zero = l_lo;  // first l to clear
  if (top >= l_lo) {
    if (!(xf >= 1e-8f)) {
      if (l_lo == 0) put(0, scale);
      l_zero = max(l_lo, 1);
    } else {
      const double x = double(xf);
      float sf, cf;
      sincosf(xf, &sf, &cf);
      const double ix = recip(x);
      const double j0 = double(sf) * ix;
      const double j1 = (double(sf) * ix - double(cf)) * ix;
      const int lu = min(top, int(floorf(xf)));
      if (l_lo == 0) put(0, float(j0) * scale);
      double jlu = j0;
      if (lu >= 1) {
        if (l_lo <= 1) put(1, float(j1) * scale);
        // FP32 upward recurrence; (2l + 1) / x is formed from the exact odd
        // integer every step, so the coefficient carries one rounding.
        const float ixf = float(ix);
        float jm = float(j0), jc = float(j1), odd = 3.f;
        for (int l = 1; l < lu; ++l) {
          const float jn = (odd * ixf) * jc - jm;
          jm = jc;
          jc = jn;
          odd += 2.f;
          if (l + 1 >= l_lo) put(l + 1, jn * scale);
        }
        jlu = double(jc);
      }
      if (top > lu) {
        const int S = top + kMillerExtra;
        const int lo = max(lu + 1, l_lo);
        if (miller_safe(S, lu, xf)) {
          // state at index l: f = f_l, fp = f_{l+1}, c = (2l + 1) / x.  The
          // loops take two steps per trip, f and fp trading roles, so the
          // state needs no register moves.
          const double d = 2.0 * ix;
          double fp = 0.0, f = 1e-250, c = double(2 * S + 1) * ix;
#pragma unroll
          for (int k = 0; k < kMillerExtra; k += 2) {  // S down to top
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
            put(l, float(ns * f));
            fp = c * f - fp;
            c -= d;
            put(l - 1, float(ns * fp));
            f = c * fp - f;
            c -= d;
          }
          if (l >= lo) put(l, float(ns * f));
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
            if (l <= top) put(l, l <= last ? float(norm * f) * scale : 0.f);
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
  for (int l = max(l_zero, l_lo); l <= l_hi; ++l) put(l, 0.f);
}

// The columns of two points of one q (x0, x1) in lock step: put2(l, v0, v1)
// with j_l(x0) s0 and j_l(x1) s1 for l in [l_lo, min(lq, l_hi)].  Orders
// above lq are left unwritten: the kernel's sums for them are never read
// (multipole_sum_types takes l <= lq only, and an MFMA row is one q).  Both
// columns use the split order lu = floor(min(x0, x1)): for the larger x, lu
// is floor(x) or one less, where j_lu(x) is still positive (the first zero
// of j_l lies beyond l + 2) and the upward recurrence still stable.  The
// recurrences are linear, so the scales ride along from the start values,
// and the Miller storing pass runs on the normalised FP64 state (values
// j_l s, so it stores conversions only); that needs the Miller growth bound
// below 2^900 so that the normalised top value cannot leave the double
// range.  Returns false, having stored nothing, where floor(x) differs by
// more than one, x is below 1e-8, or the growth bound is larger; the caller
// then takes the two columns one at a time (bessel_column).
template <class Put2>
__device__ inline bool bessel_pair(float x0f, float x1f, int lq, int l_lo, int l_hi, float s0,
                                   float s1, Put2 put2) {
  const int top = min(lq, l_hi);
  if (top < l_lo) return true;
  const float xmin = fminf(x0f, x1f);
  if (!(xmin >= 1e-8f)) return false;
  const int fl = int(floorf(xmin));
  if (int(floorf(fmaxf(x0f, x1f))) > fl + 1) return false;
  const int lu = min(top, fl);
  const int S = top + kMillerExtra;
  if (top > lu &&
      !(float(S - lu) * (__log2f(float(2 * S + 1)) - __log2f(xmin)) < 900.f))
    return false;
  float sf0, cf0, sf1, cf1;
  sincosf(x0f, &sf0, &cf0);
  sincosf(x1f, &sf1, &cf1);
  const double ix0 = recip(double(x0f)), ix1 = recip(double(x1f));
  const double a0 = double(sf0) * ix0, a1 = double(sf1) * ix1;
  const double b0 = (a0 - double(cf0)) * ix0, b1 = (a1 - double(cf1)) * ix1;
  // scaled j_{l-1}, j_l of the upward recurrence at l = 1
  float jm0 = float(a0) * s0, jc0 = float(b0) * s0, jm1 = float(a1) * s1, jc1 = float(b1) * s1;
  if (l_lo == 0) put2(0, jm0, jm1);
  float u0 = jm0, u1 = jm1;  // scaled j_lu
  if (lu >= 1) {
    if (l_lo <= 1) put2(1, jc0, jc1);
    const float ixf0 = float(ix0), ixf1 = float(ix1);
    float odd = 3.f;
    int l = 1;
    for (; l < lu && l + 1 < l_lo; ++l) {  // orders below the group: not stored
      const float jn0 = (odd * ixf0) * jc0 - jm0;
      const float jn1 = (odd * ixf1) * jc1 - jm1;
      jm0 = jc0;
      jc0 = jn0;
      jm1 = jc1;
      jc1 = jn1;
      odd += 2.f;
    }
    // two orders per trip, j_{l-1} and j_l trading registers
    for (; l + 1 < lu; l += 2) {
      jm0 = (odd * ixf0) * jc0 - jm0;
      jm1 = (odd * ixf1) * jc1 - jm1;
      put2(l + 1, jm0, jm1);
      odd += 2.f;
      jc0 = (odd * ixf0) * jm0 - jc0;
      jc1 = (odd * ixf1) * jm1 - jc1;
      put2(l + 2, jc0, jc1);
      odd += 2.f;
    }
    if (l < lu) {
      const float jn0 = (odd * ixf0) * jc0 - jm0;
      const float jn1 = (odd * ixf1) * jc1 - jm1;
      put2(l + 1, jn0, jn1);
      jc0 = jn0;
      jc1 = jn1;
    }
    u0 = jc0;
    u1 = jc1;
  }
  if (top > lu) {
    const int lo = max(lu + 1, l_lo);
    const double d0 = 2.0 * ix0, d1 = 2.0 * ix1;
    double p0 = 0.0, f0 = 1e-250, c0 = double(2 * S + 1) * ix0;
    double p1 = 0.0, f1 = 1e-250, c1 = double(2 * S + 1) * ix1;
#pragma unroll
    for (int k = 0; k < kMillerExtra; k += 2) {
      p0 = c0 * f0 - p0;
      p1 = c1 * f1 - p1;
      c0 -= d0;
      c1 -= d1;
      f0 = c0 * p0 - f0;
      f1 = c1 * p1 - f1;
      c0 -= d0;
      c1 -= d1;
    }
    const double F0 = f0, P0 = p0, C0 = c0, F1 = f1, P1 = p1, C1 = c1;
    int l = top;
    for (; l - 1 > lu; l -= 2) {
      p0 = c0 * f0 - p0;
      p1 = c1 * f1 - p1;
      c0 -= d0;
      c1 -= d1;
      f0 = c0 * p0 - f0;
      f1 = c1 * p1 - f1;
      c0 -= d0;
      c1 -= d1;
    }
    if (l > lu) {
      f0 = c0 * f0 - p0;
      f1 = c1 * f1 - p1;
    }
    // f_lu > 0 (j_l(x) > 0 for l >= floor(x)), and on this path it is at
    // most 2^900 times 1e-250, so its reciprocal is a normal double and the
    // normalised state at top (>= 2^-900 j_lu s) stays normal too.
    const double n0 = double(u0) * recip(f0), n1 = double(u1) * recip(f1);
    f0 = F0 * n0;
    p0 = P0 * n0;
    c0 = C0;
    f1 = F1 * n1;
    p1 = P1 * n1;
    c1 = C1;
    for (l = top; l - 1 >= lo; l -= 2) {
      put2(l, float(f0), float(f1));
      p0 = c0 * f0 - p0;
      p1 = c1 * f1 - p1;
      c0 -= d0;
      c1 -= d1;
      put2(l - 1, float(p0), float(p1));
      f0 = c0 * p0 - f0;
      f1 = c1 * p1 - f1;
      c0 -= d0;
      c1 -= d1;
    }
    if (l >= lo) put2(l, float(f0), float(f1));
  }
  return true;
}

// partial[(chunk * nqp + q) * ncol + col]: the sum over the chunk's points
// (all of one type) of occ j_l Y, times 2^-(yexp + aexp) where ascale =
// 2^aexp scales the j_l occ operands and 2^yexp the harmonic operands.
// At least WPE waves per SIMD: 2 (256 registers), 3 (168) or 4 (128).
template <int MAXT, int WPE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void multipole_coeff(
    const float* __restrict__ radius, const float* __restrict__ occ,
    const int4* __restrict__ chunks /* (first slab, end slab, type) */, int nchunk, int nslab,
    const _Float16* __restrict__ Y, int ntiles, const int* __restrict__ tile_l,
    const float* __restrict__ qv, const int* __restrict__ qlmax, int nqp, int ngroups,
    float ascale, float unscale, float* __restrict__ partial) {
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
  const int4 ck = chunks[chunk];
  const int s0 = ck.x, s1 = ck.y;
  _Float16* X = reinterpret_cast<_Float16*>(smem);  // [l][hi, lo fragment]

  v4f acc[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};

  // generation role: wave w takes rows q = 4 w + lane / 16 and the point
  // pair 2 (lane % 16), + 1 of the slab.  Points are sorted by radius, so the
  // lanes of a wave see nearly equal q r and run the recurrences with little
  // divergence.
  const int gq = 4 * wave + (lane >> 4), gpp = lane & 15;
  const float qg = qv[q0 + gq];
  const int lg = qlmax[q0 + gq];
  // this thread's dword in the hi fragment of order l: Xw[(l - l_lo) * kFrag]
  // (lo fragment kFrag / 2 dwords further)
  uint32_t* const Xw = reinterpret_cast<uint32_t*>(X) + 4 * lds_slot(gpp >> 2, gq) + (gpp & 3);
  _Float16* const Xh = X + 8 * lds_slot(gpp >> 2, gq) + 2 * (gpp & 3);
  // MFMA role: lane = 16 g + c, rows/columns c, points 8 g .. 8 g + 7
  const int mk = lane >> 4, mn = lane & 15;
  const _Float16* const Xa = X + 8 * lds_slot(mk, mn);

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
    xoff[t] = (min(l, l_hi) - l_lo) * (2 * kFrag);
  }
  // Harmonic tile rows read by this wave; tiles past the live ones read the
  // last live (or, with none, the last existing) tile, so every load in the
  // loop is in bounds.
  size_t yrow[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t)
    yrow[t] = size_t(min(g0 + wave + 4 * max(min(t, nt - 1), 0), ntiles - 1)) * nslab * (2 * kFrag);

  // Harmonic tiles are read in batches of kB, batch b + 1 while batch b runs
  // its MFMAs; nothing is held in registers across the generation.
  constexpr int kB = 2, kNB = MAXT / kB;
  static_assert(MAXT % kB == 0, "whole batches of tiles");
  for (int s = s0; s < s1; ++s) {
    const _Float16* Ys = Y + size_t(s) * (2 * kFrag) + lane * 8;
    // radii and occupancies of this thread's two points (read while the
    // barrier waits)
    const float2 pr = *reinterpret_cast<const float2*>(radius + s * kSlab + 2 * gpp);
    const float2 po = *reinterpret_cast<const float2*>(occ + s * kSlab + 2 * gpp);
    __syncthreads();  // the previous slab's operands have been read
    // j_l times the points' occupancy (padding points have occupancy 0)
    {
      const float x0 = qg * pr.x, x1 = qg * pr.y;
      const float o0 = po.x * ascale, o1 = po.y * ascale;
      const bool paired = bessel_pair(x0, x1, lg, l_lo, l_hi, o0, o1, [&](int l, float v0, float v1) {
        uint32_t hi, lo;
        split_pair(v0, v1, hi, lo);
        uint32_t* p = Xw + (l - l_lo) * kFrag;
        p[0] = hi;
        p[kFrag / 2] = lo;
      });
      if (!paired) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
          bessel_column(j ? x1 : x0, lg, l_lo, l_hi, j ? o1 : o0, [&](int l, float v) {
            const _Float16 hi = static_cast<_Float16>(v);
            const _Float16 lo = static_cast<_Float16>(v - static_cast<float>(hi));
            _Float16* p = Xh + (l - l_lo) * (2 * kFrag) + j;
            p[0] = hi;
            p[kFrag] = lo;
          });
      }
    }
    __syncthreads();
    v8h yh[2][kB], yl[2][kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      yh[0][u] = *reinterpret_cast<const v8h*>(Ys + yrow[u]);
      yl[0][u] = *reinterpret_cast<const v8h*>(Ys + yrow[u] + kFrag);
    }
#pragma unroll
    for (int b = 0; b < kNB; ++b) {
      if (b * kB >= nt) break;
      if (b + 1 < kNB && (b + 1) * kB < nt) {
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const size_t r = yrow[(b + 1) * kB + u];
          yh[(b + 1) & 1][u] = *reinterpret_cast<const v8h*>(Ys + r);
          yl[(b + 1) & 1][u] = *reinterpret_cast<const v8h*>(Ys + r + kFrag);
        }
      }
      // both tiles' A operands are read before either tile's MFMAs (offsets
      // clamped to the group's orders, so the reads are in bounds)
      v8h ah[kB], al[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        ah[u] = *reinterpret_cast<const v8h*>(Xa + xoff[b * kB + u]);
        al[u] = *reinterpret_cast<const v8h*>(Xa + xoff[b * kB + u] + kFrag);
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int t = b * kB + u;
        if (t >= nt) continue;
        const v8h bh = yh[b & 1][u], bl = yl[b & 1][u];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[u], bh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[u], bl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[u], bh, acc[t], 0, 0, 0);
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
      partial[(size_t(chunk) * nqp + q) * ncol + tile * 16 + mn] = acc[t][r] * unscale;
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

// Accumulator tiles per wave (4 x MAXT tiles per tile group): 8, 12, 16 or
// 32.  Each tile group generates j_l once more, more accumulators lower the
// occupancy, and a group's l range sets its LDS (2 KB per order).  The
// automatic choice is one group of 64 where 33 to 64 tiles cover the
// expansion and groups of 32 otherwise: on the 108k-point prolate with
// sites, 200 q, groups of 32 / 48 / 64 took 0.22 / 0.25 / 0.25 ms at
// l_max 15, 0.50 / 0.72 / 0.42 ms at 25, 0.95 / 1.00 / 1.06 ms at 40 and
// 2.09 / 2.22 / 2.48 ms at 64.
constexpr int kGroupTiles[] = {32, 48, 64, 128};

}  // namespace

int multipole_nodes() { return kNodes; }

int multipole_harmonic_exponent() { return kYExp; }

hipError_t launch_multipole_point_terms(const double* xyz, const double* occ, const int* bin_of,
                                        int p0, int p1, const double* centre, int l_max,
                                        const double* norms, double h, const double* node_x,
                                        const double* node_den, double* terms, double* weights,
                                        hipStream_t stream) {
  if (p1 <= p0) return hipSuccess;
  if (!(h > 0.0) || l_max < 0 || l_max > 65534) return hipErrorInvalidValue;
  NodeTable nodes;
  for (int k = 0; k < kNodes; ++k) {
    nodes.x[k] = node_x[k];
    nodes.den[k] = node_den[k];
  }
  multipole_point_terms<<<dim3((p1 - p0 + 63) / 64, l_max + 1), 64, 0, stream>>>(
      xyz, occ, bin_of, p0, p1, centre, l_max, norms, h, nodes, terms, weights);
  return hipGetLastError();
}

hipError_t launch_multipole_bin_moments(const double* terms, const double* weights,
                                        const int* bins, int nbins, int p0, int nlm,
                                        double* moments, hipStream_t stream) {
  if (nbins <= 0) return hipSuccess;
  if (nbins > 65535) return hipErrorInvalidValue;
  multipole_bin_moments<<<dim3((nlm + 63) / 64, nbins), 64, 0, stream>>>(
      terms, weights, reinterpret_cast<const int4*>(bins), p0, nlm, moments);
  return hipGetLastError();
}

hipError_t launch_multipole_moment_max(const double* moments, size_t n, unsigned long long* out,
                                       hipStream_t stream) {
  zero_u64<<<1, 1, 0, stream>>>(out);
  const size_t blocks = std::min<size_t>((n + 255) / 256, 256);
  if (blocks > 0) abs_max_f64<<<unsigned(blocks), 256, 0, stream>>>(moments, n, out);
  return hipGetLastError();
}

hipError_t launch_multipole_moment_tiles(const double* moments, int nvirt, int l_max,
                                         const int* tile_first, int ntiles, int npad, int yexp,
                                         void* Yraw, hipStream_t stream) {
  _Float16* Y = static_cast<_Float16*>(Yraw);
  if (npad % kSlab != 0 || nvirt > npad) return hipErrorInvalidValue;
  const size_t ny = size_t(ntiles) * npad * 32;  // 16 columns x (hi, lo) per point
  fill_zero<<<(ny + 255) / 256, 256, 0, stream>>>(Y, ny);
  const int nlm = (l_max + 1) * (l_max + 1);
  if (nvirt > 0)
    multipole_moment_tiles<<<dim3((nlm + 255) / 256, std::min(nvirt, 4096)), 256, 0, stream>>>(
        moments, nvirt, l_max, tile_first, npad, yexp, Y);
  return hipGetLastError();
}

int multipole_chunk() { return kChunk; }

int multipole_slab() { return kSlab; }

int multipole_tiles_per_group(int ntiles, int tuned) {
  for (int t : kGroupTiles)
    if (t == tuned) return t;
  return ntiles > 32 && ntiles <= 64 ? 64 : 32;
}

size_t multipole_coeff_lds_bytes(int l_range) {
  return size_t(l_range) * 2 * kFrag * sizeof(_Float16);
}

int multipole_max_types() { return 256; }

hipError_t launch_multipole_geometry(const double* xyz, int npts, const double* centre,
                                     int l_max, const double* norms, const int* tile_first,
                                     int ntiles, int npad, float* radius, void* Yraw,
                                     hipStream_t stream) {
  _Float16* Y = static_cast<_Float16*>(Yraw);
  if (npts <= 0) return hipSuccess;
  if (npad % kSlab != 0) return hipErrorInvalidValue;
  const size_t ny = size_t(ntiles) * npad * 32;  // 16 columns x (hi, lo) per point
  fill_zero<<<(ny + 255) / 256, 256, 0, stream>>>(Y, ny);
  multipole_geometry<<<(npts + 127) / 128, 128, 0, stream>>>(xyz, npts, centre, l_max, norms,
                                                             tile_first, npad, radius, Y);
  return hipGetLastError();
}

hipError_t launch_multipole_coeff(const float* radius, const float* occ, const int* chunks,
                                  int nchunk, int npad, const void* Yraw, int ntiles,
                                  const int* tile_l, int tpg, int max_l_range, int aexp,
                                  int yexp, const int* type_chunk, int ntypes, const float* qv,
                                  const int* qlmax, int nqp, float* partial, double* coef,
                                  hipStream_t stream, KernelTimer* timer) {
  if (nqp % kQB != 0 || npad % kSlab != 0 || ntiles <= 0 || nchunk <= 0 || ntypes <= 0 ||
      (tpg != 32 && tpg != 48 && tpg != 64 && tpg != 128) || aexp < -60 || aexp > 60 ||
      yexp < -60 || yexp > 60)
    return hipErrorInvalidValue;
  const int ngroups = (ntiles + tpg - 1) / tpg;
  const long long total = (long long)(nqp / kQB) * nchunk * ngroups;
  if (total > INT32_MAX) return hipErrorInvalidValue;
  const size_t lds = multipole_coeff_lds_bytes(max_l_range);
  const int4* ck = reinterpret_cast<const int4*>(chunks);
  const int nslab = npad / kSlab;
  const _Float16* Y = static_cast<const _Float16*>(Yraw);
  const float ascale = ldexpf(1.f, aexp), unscale = ldexpf(1.f, -(yexp + aexp));
  {
    TimerScope scope(timer, "multipole_coeff", stream);
    if (tpg == 128)
      multipole_coeff<32, 2><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, ascale,
          unscale, partial);
    else if (tpg == 48)
      multipole_coeff<12, 3><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, ascale,
          unscale, partial);
    else if (tpg == 64)
      multipole_coeff<16, 3><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, ascale,
          unscale, partial);
    else
      multipole_coeff<8, 4><<<unsigned(total), kThreads, lds, stream>>>(
          radius, occ, ck, nchunk, nslab, Y, ntiles, tile_l, qv, qlmax, nqp, ngroups, ascale,
          unscale, partial);
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

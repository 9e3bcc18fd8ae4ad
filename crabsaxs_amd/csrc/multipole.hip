// Spherical-multipole kernels for gfx950 (spherical.rs:41-565 restated for
// the GPU).
//
//  multipole_geometry  once per (point set, l_max): FP64 real spherical
//                      harmonics by the associated-Legendre recurrence of
//                      spherical.rs:54-93, stored FP32 in MFMA B-operand tiles
//                      Y[tile][point][16]: each l owns ceil((2l+1)/16) tiles of
//                      16 columns (column c of l is m = c - l; padding is 0).
//  multipole_coeff     coefficients C_c[q][lm] = sum_i A_c,i(q) j_l(q r_i) Y_lm,i
//                      as an FP32 MFMA contraction (v_mfma_f32_16x16x4_f32):
//                      rows = 16 q values, columns = 16 m of one l, k = 4
//                      points.  A workgroup owns a chunk of points, a block of
//                      16 q and a group of column tiles; per 16-point slab its
//                      256 threads first generate j_l(q r_i) for (16 points x
//                      16 q) into LDS -- FP32 sin/cos, FP64 upward recurrence
//                      for l <= q r, FP64 downward ratio recurrence (Miller)
//                      above, valid for every x -- and the amplitudes; then
//                      each wave runs the MFMAs of its tiles.  Partial
//                      coefficients per chunk are written in FP32.
//  multipole_reduce    FP64 sum of the chunk partials per (q, column) and the
//                      quadratic forms sum_lm C_a C_b (1 or 6 per q).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace csx {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kSlab = 16;    // points per j_l generation step
constexpr int kFlush = 4;    // slabs per MFMA accumulation run
constexpr int kQB = 16;      // q values per MFMA row block
constexpr int kThreads = 256;

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
  auto store = [&](int l, int col, double v) {
    const int t = tile_first[l] + col / 16;
    Y[(size_t(t) * npad + i) * 16 + (col & 15)] = float(v);
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

// j_l(x) for l in [l_lo, l_hi] (zero above lmax) into w[l - l_lo].
__device__ inline void spherical_bessel_range(float xf, int lmax, int l_lo, int l_hi, double* w,
                                              int stride) {
  for (int l = l_lo; l <= l_hi; ++l) w[(l - l_lo) * stride] = 0.0;
  if (lmax < l_lo) return;
  if (xf < 1e-8f) {
    if (l_lo == 0) w[0] = 1.0;
    return;
  }
  const double x = double(xf);
  float sf, cf;
  sincosf(xf, &sf, &cf);
  const double s = double(sf), c = double(cf);
  const double ix = 1.0 / x;
  const double j0 = s * ix;
  const double j1 = s * ix * ix - c * ix;
  const int top = min(lmax, l_hi);
  const int lu = min(lmax, int(floorf(xf)));
  // Upward recurrence, stable for l <= x.
  double jm = j0, jc = j1, jlu = j0;
  if (l_lo == 0) w[0] = j0;
  if (lu >= 1) {
    if (l_lo <= 1 && 1 <= top) w[(1 - l_lo) * stride] = j1;
    jlu = j1;
    for (int l = 1; l < lu; ++l) {
      const double jn = double(2 * l + 1) * ix * jc - jm;
      jm = jc;
      jc = jn;
      if (l + 1 >= l_lo && l + 1 <= top) w[(l + 1 - l_lo) * stride] = jn;
    }
    jlu = jc;
  }
  if (top <= lu) return;
  // Downward ratio recurrence rho_l = j_l / j_{l-1} from lmax + 14, stable for
  // l > x; the ratios of the stored range are parked in w, those below it
  // are folded into one product.
  const int start = max(l_lo, lu + 1);
  double rho = x / double(2 * (lmax + 14) + 1);
  double below = 1.0;
  for (int l = lmax + 13; l > lu; --l) {
    rho = x / (double(2 * l + 1) - x * rho);
    if (l <= top && l >= start) w[(l - l_lo) * stride] = rho;
    else if (l < start) below *= rho;
  }
  double v = jlu * below;  // j_{start-1}
  for (int l = start; l <= top; ++l) {
    v *= w[(l - l_lo) * stride];
    w[(l - l_lo) * stride] = v;
  }
}

// partial[((chunk * NCH + c) * nqp + q) * ncol + col]
template <int NCH, int TPW>
__global__ __launch_bounds__(kThreads) void multipole_coeff(
    const float* __restrict__ radius, const int* __restrict__ type, const float* __restrict__ occ,
    int npts, int npad, const float* __restrict__ Y, int ntiles, const int* __restrict__ tile_l,
    const float* __restrict__ amp /* [ntypes][nqp][NCH] */, const float* __restrict__ qv,
    const int* __restrict__ qlmax, int nqp, int chunk, float* __restrict__ partial) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * chunk;
  const int c1 = min(c0 + chunk, npad);
  const int qb = blockIdx.y * kQB;
  const int tpg = 4 * TPW;
  const int g0 = blockIdx.z * tpg;
  const int g1 = min(g0 + tpg, ntiles);
  const int l_lo = tile_l[g0], l_hi = tile_l[g1 - 1];
  const int lr = l_hi - l_lo + 1;
  const int stride = (lr | 1);  // odd stride: conflict-free 8-byte LDS rows
  double* W = smem;                                           // [16 pt][16 q][stride]
  float* A = reinterpret_cast<float*>(smem + kSlab * kQB * stride);  // [16 pt][16 q][NCH]

  int my_tiles[TPW];
  int my_l[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = g0 + wave + 4 * t;
    my_tiles[t] = tile < g1 ? tile : -1;
    my_l[t] = tile < g1 ? tile_l[tile] : 0;
  }
  // Two-level accumulation: the MFMA accumulator covers kFlush slabs (64
  // points) and is then added into an FP32 running sum, so the rounding of a
  // long same-sign accumulation (forward scattering) does not build up.
  v4f acc[TPW][NCH], tot[TPW][NCH];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      acc[t][c] = v4f{0.f, 0.f, 0.f, 0.f};
      tot[t][c] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  int slabs = 0;

  // phase-1 roles: one (point, q) pair per thread
  const int gq = threadIdx.x % kQB, gp = threadIdx.x / kQB;
  const float q_gen = qv[qb + gq];
  const int lq_gen = qlmax[qb + gq];
  // phase-2 roles (MFMA lane layout): row = q, k = point
  const int mrow = lane % 16, mk = lane / 16;

  for (int s0 = c0; s0 < c1; s0 += kSlab) {
    {
      const int i = s0 + gp;
      double* w = W + (gp * kQB + gq) * stride;
      if (i < npts) {
        spherical_bessel_range(q_gen * radius[i], lq_gen, l_lo, l_hi, w, 1);
        const float o = occ[i];
        const int ty = type[i];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          A[(gp * kQB + gq) * NCH + c] = o * amp[(size_t(ty) * nqp + qb + gq) * NCH + c];
      } else {
        for (int l = 0; l < lr; ++l) w[l] = 0.0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) A[(gp * kQB + gq) * NCH + c] = 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kSlab / 4; ++ks) {
      const int p = 4 * ks + mk;
      float am[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) am[c] = A[(p * kQB + mrow) * NCH + c];
      const double* wrow = W + (p * kQB + mrow) * stride - l_lo;
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        if (my_tiles[t] < 0) continue;
        const float jv = float(wrow[my_l[t]]);
        const float b = Y[(size_t(my_tiles[t]) * npad + s0 + p) * 16 + mrow];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(jv * am[c], b, acc[t][c], 0, 0, 0);
      }
    }
    __syncthreads();
    if (++slabs == kFlush) {
      slabs = 0;
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          tot[t][c] += acc[t][c];
          acc[t][c] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int c = 0; c < NCH; ++c) tot[t][c] += acc[t][c];
  const int ncol = ntiles * 16;
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    if (my_tiles[t] < 0) continue;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qb + 4 * mk + r;
        partial[((size_t(blockIdx.x) * NCH + c) * nqp + q) * ncol + my_tiles[t] * 16 + mrow] =
            tot[t][c][r];
      }
  }
}

// out[q * NT + k]: NCH = 1 -> sum C^2; NCH = 3 -> vv, ee, hh, ve, vh, eh.
template <int NCH>
__global__ __launch_bounds__(256) void multipole_reduce(const float* __restrict__ partial,
                                                        int nchunk, int nqp, int ncol,
                                                        double* __restrict__ out) {
  constexpr int NT = NCH == 1 ? 1 : 6;
  __shared__ double scratch[4 * NT];
  const int q = blockIdx.x;
  double s[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) s[k] = 0.0;
  for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
    double cc[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      double v = 0.0;
      for (int ch = 0; ch < nchunk; ++ch)
        v += double(partial[((size_t(ch) * NCH + c) * nqp + q) * ncol + col]);
      cc[c] = v;
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

constexpr int kTPW = 8;

}  // namespace

int multipole_tiles_per_group() { return 4 * kTPW; }

size_t multipole_coeff_lds_bytes(int l_range, int nch) {
  const int stride = l_range | 1;
  return size_t(kSlab) * kQB * stride * sizeof(double) + size_t(kSlab) * kQB * nch * sizeof(float);
}

hipError_t launch_multipole_geometry(const double* xyz, int npts, const double* centre,
                                     int l_max, const double* norms, const int* tile_first,
                                     int npad, float* radius, float* Y, hipStream_t stream) {
  if (npts <= 0) return hipSuccess;
  multipole_geometry<<<(npts + 127) / 128, 128, 0, stream>>>(xyz, npts, centre, l_max, norms,
                                                             tile_first, npad, radius, Y);
  return hipGetLastError();
}

hipError_t launch_multipole_coeff(const float* radius, const int* type, const float* occ,
                                  int npts, int npad, const float* Y, int ntiles,
                                  const int* tile_l, int max_l_range, const float* amp,
                                  const float* qv, const int* qlmax, int nqp, int nch, int chunk,
                                  float* partial, double* out, hipStream_t stream,
                                  KernelTimer* timer) {
  if (nqp % kQB != 0 || chunk % kSlab != 0 || npad % kSlab != 0) return hipErrorInvalidValue;
  const int nchunk = (npad + chunk - 1) / chunk;
  const int ngroups = (ntiles + 4 * kTPW - 1) / (4 * kTPW);
  dim3 grid(nchunk, nqp / kQB, ngroups);
  const size_t lds = multipole_coeff_lds_bytes(max_l_range, nch);
  {
    TimerScope scope(timer, "multipole_coeff", stream);
    if (nch == 1)
      multipole_coeff<1, kTPW><<<grid, kThreads, lds, stream>>>(
          radius, type, occ, npts, npad, Y, ntiles, tile_l, amp, qv, qlmax, nqp, chunk, partial);
    else if (nch == 3)
      multipole_coeff<3, kTPW><<<grid, kThreads, lds, stream>>>(
          radius, type, occ, npts, npad, Y, ntiles, tile_l, amp, qv, qlmax, nqp, chunk, partial);
    else
      return hipErrorInvalidValue;
  }
  const int ncol = ntiles * 16;
  if (nch == 1)
    multipole_reduce<1><<<nqp, 256, 0, stream>>>(partial, nchunk, nqp, ncol, out);
  else
    multipole_reduce<3><<<nqp, 256, 0, stream>>>(partial, nchunk, nqp, ncol, out);
  return hipGetLastError();
}

}  // namespace csx

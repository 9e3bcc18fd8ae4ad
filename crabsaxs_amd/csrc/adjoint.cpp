// Host side of the Debye adjoint: W' tables, launch, chunk reduction.
#include <algorithm>
#include <cmath>

#include "capi_internal.hpp"

namespace csx {

hipError_t launch_debye_jacobian(const float4* pts, const int* orig, const int* type, int npts,
                                 const int* seg_start, int ntypes, const float* qrev,
                                 const float* Wp, int nq_pad, int sin_mode, double* out,
                                 hipStream_t stream, KernelTimer* timer);
hipError_t launch_debye_vjp(const float4* pts, const int4* blocks, int nb, const int2* items,
                            size_t nitems, const int* orig, int ntypes, const float* qrev,
                            float qstep, const float* AB, int nq_pad, int qt, int sin_mode,
                            int qsplit, int tsplit, size_t items_per_launch, double* part,
                            double* gacc, double* out, hipStream_t stream, KernelTimer* timer);
hipError_t launch_debye_vjp_radial(const float4* pts, const int4* blocks, int nb, const int2* items,
                                   size_t nitems, const int* orig, int ntypes, const double* qv,
                                   const double* B, int nq, double h, int nk, int tsplit,
                                   size_t items_per_launch, float* coef, double* part, double* gacc,
                                   double* out, hipStream_t stream, KernelTimer* timer);
int jacobian_q_tile();
int vjp_q_tile(int tuned, size_t nq, int sin_mode, size_t nitems, int cus, int max_tile);
int vjp_qsplit(size_t nqc, size_t nitems);

namespace {

constexpr int kBlock = 256;

// Per-type amplitudes A_t(q_k) of the intensity being differentiated.
std::vector<double> type_amplitudes_for(const csx_structure* s, const csx_solvent_params* p,
                                        bool implicit_h, const double* q, size_t nq, int T) {
  std::vector<double> A(static_cast<size_t>(T) * nq);
  for (size_t k = 0; k < nq; ++k) {
    const double g = p ? excluded_volume_scale(q[k], p->c1) : 0.0;
    for (int t = 0; t < T; ++t) {
      if (!p) {
        A[t * nq + k] = type_form_factor(s->types[t], implicit_h, q[k]);
      } else {
        const Amplitudes a = debye_amplitudes(s->types[t], implicit_h, q[k],
                                              p->shell_thickness_angstrom, p->bulk_density_e_per_a3);
        A[t * nq + k] = a.v - g * a.e + p->c2 * a.h;
      }
    }
  }
  return A;
}

}  // namespace

// Sine mode 4: F_ab(r) per radial bin from Chebyshev coefficients (the
// kernel file describes it); B_ab(q) = 2 w_q A_a A_b / q at the exact q.
static csx_status vjp_radial(Call& call, const csx_structure* s, PointSet* ps, const double* q,
                             size_t nq, const double* weights, const csx_solvent_params* p,
                             bool implicit_h, const std::vector<int>& blocks,
                             const std::vector<int>& items, double h, int nk, double* out) {
  const int T = ps->ntypes;
  const size_t n = s->n;
  const int nb = static_cast<int>(blocks.size() / 4);
  const size_t nitems = items.size() / 2;
  const std::vector<double> A = type_amplitudes_for(s, p, implicit_h, q, nq, T);
  std::vector<double> B(static_cast<size_t>(T) * T * nq, 0.0);
  for (size_t k = 0; k < nq; ++k) {
    if (q[k] == 0.0) continue;
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b)
        B[(static_cast<size_t>(a) * T + b) * nq + k] =
            2.0 * weights[k] * A[a * nq + k] * A[b * nq + k] / q[k];
  }
  const size_t fill = 4 * static_cast<size_t>(std::max(call.ctx()->cu_count, 1));
  int tsplit = 1;
  while (tsplit < 8 && 2 * nitems * static_cast<size_t>(tsplit) <= fill) tsplit *= 2;
  int split;
  size_t budget;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    split = call.ctx()->tuning.vjp_split;
    budget = call.ctx()->tuning.workspace_limit;
  }
  if (split > 0) tsplit = split;
  const size_t per_item = static_cast<size_t>(tsplit) * 2 * 3 * kBlock * sizeof(double);
  const size_t ipl = std::max<size_t>(1, std::min(nitems, budget / per_item));
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    int* last = call.ctx()->last_vjp;
    last[0] = 0;  // no q tile: the radial sweep
    last[1] = 1;
    last[2] = tsplit;
    last[3] = static_cast<int>((nitems + ipl - 1) / ipl);
  }
  const hipStream_t st = call.stream();
  Scratch d_q(call), d_b(call), d_coef(call), d_blocks(call), d_items(call), d_part(call),
      d_acc(call), d_out(call);
  CSX_HIP(d_q.alloc(nq * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_b.alloc(B.size() * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_coef.alloc(static_cast<size_t>(T) * T * nk * 8 * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_blocks.alloc(blocks.size() * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(d_items.alloc(items.size() * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(d_part.alloc(ipl * per_item, st), "hipMallocAsync");
  CSX_HIP(d_acc.alloc(static_cast<size_t>(nb) * 3 * kBlock * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(n * 3 * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), q, nq * sizeof(double), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_b.as<void>(), B.data(), B.size() * sizeof(double),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_blocks.as<void>(), blocks.data(), blocks.size() * sizeof(int),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_items.as<void>(), items.data(), items.size() * sizeof(int),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(launch_debye_vjp_radial(ps->pts32.as<float4>(), d_blocks.as<int4>(), nb,
                                  d_items.as<int2>(), nitems, ps->orig32.as<int>(), T,
                                  d_q.as<double>(), d_b.as<double>(), static_cast<int>(nq), h, nk,
                                  tsplit, ipl, d_coef.as<float>(), d_part.as<double>(),
                                  d_acc.as<double>(), d_out.as<double>(), st, call.timer()),
          "debye_vjp_radial");
  CSX_HIP(hipMemcpyAsync(out, d_out.as<void>(), n * 3 * sizeof(double), hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  d_q.reset();
  d_b.reset();
  d_coef.reset();
  d_blocks.reset();
  d_items.reset();
  d_part.reset();
  d_acc.reset();
  d_out.reset();
  return call.finish();
}

static csx_status vjp(Call& call, const csx_structure* s, PointSet* ps, const double* q, size_t nq,
                      const double* weights, const csx_solvent_params* p, bool implicit_h,
                      double* out) {
  const int T = ps->ntypes;
  const size_t n = s->n;
  // type-pure blocks of 256 sorted atoms and the block pairs I <= J
  std::vector<int> blocks;
  for (int t = 0; t < T; ++t)
    for (int i0 = ps->h_seg[t]; i0 < ps->h_seg[t + 1]; i0 += kBlock)
      blocks.insert(blocks.end(), {i0, std::min(kBlock, ps->h_seg[t + 1] - i0), t, 0});
  const int nb = static_cast<int>(blocks.size() / 4);
  std::vector<int> items;
  items.reserve(static_cast<size_t>(nb) * (nb + 1));
  for (int I = 0; I < nb; ++I)
    for (int J = I; J < nb; ++J) items.insert(items.end(), {I, J});
  const size_t nitems = items.size() / 2;
  int tune_sm;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    tune_sm = call.ctx()->tuning.sin_mode;
  }
  double rh = 0.0;
  int rnk = 0;
  if (tune_sm == 4 && radial_distance_bins(q, nq, ps->max_radius, kVjpRadialMaxBins, &rh, &rnk))
    return vjp_radial(call, s, ps, q, nq, weights, p, implicit_h, blocks, items, rh, rnk, out);
  // The unreduced hardware sine takes |q r / 2 pi| <= 256 revolutions; on
  // a uniform grid within that range the sines and cosines come from the
  // recurrence (sine mode 3).
  double umax = 0.0;
  for (size_t k = 0; k < nq; ++k)
    umax = std::max(umax, std::fabs(static_cast<double>(static_cast<float>(q[k] / (2.0 * kPi)))) *
                              2.0 * ps->max_radius);
  const float qstep = umax <= 250.0 ? uniform_q_step(q, nq, ps->max_radius) : 0.f;
  const int sin_mode = umax <= 250.0 ? (qstep != 0.f ? 3 : 2) : 1;
  int tuned, split;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    tuned = call.ctx()->tuning.vjp_q_tile;
    split = call.ctx()->tuning.vjp_split;
  }
  const int qt = vjp_q_tile(tuned, nq, sin_mode, nitems, call.ctx()->cu_count,
                            sin_mode == 3 ? rec_max_tile(qstep, ps->max_radius) : 1 << 30);
  const size_t nq_pad = (nq + qt - 1) / qt * qt;
  const std::vector<double> A = type_amplitudes_for(s, p, implicit_h, q, nq, T);
  // AB[a][b][0][k] = qe_k B_ab(k), AB[a][b][1][k] = B_ab(k) = w_k 2 A_a A_b / qe_k, where
  // qe_k is the q the kernel's phase uses: 2 pi (q_k / 2 pi)_f32, or under the
  // recurrence 2 pi times the chunk's first such value plus the steps since.
  std::vector<float> qrev(nq_pad, 0.0f), AB(static_cast<size_t>(T) * T * 2 * nq_pad, 0.0f);
  for (size_t k = 0; k < nq; ++k) qrev[k] = static_cast<float>(q[k] / (2.0 * kPi));
  for (size_t k = 0; k < nq; ++k) {
    const size_t kt = k % static_cast<size_t>(qt);
    const double qe = 2.0 * kPi *
                      (sin_mode == 3 ? static_cast<double>(qrev[k - kt]) +
                                           static_cast<double>(kt) * static_cast<double>(qstep)
                                     : static_cast<double>(qrev[k]));
    if (qe == 0.0) continue;
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b) {
        const double B = 2.0 * weights[k] * A[a * nq + k] * A[b * nq + k] / qe;
        float* row = &AB[((static_cast<size_t>(a) * T + b) * 2) * nq_pad];
        row[k] = static_cast<float>(qe * B);
        row[nq_pad + k] = static_cast<float>(B);
      }
  }
  const int nqc = static_cast<int>(nq_pad / qt);
  // enough workgroups to fill the device: split the q chunks of small problems
  const int qsplit = vjp_qsplit(static_cast<size_t>(nqc), nitems);
  // and, where the q chunks alone leave most CUs idle, the 256 steps of each
  // sweep over 2, 4 or 8 workgroups (a structure of a few thousand atoms has
  // tens of block pairs; its time is otherwise one whole sweep)
  const size_t wgs = nitems * static_cast<size_t>(qsplit);
  const size_t fill = 4 * static_cast<size_t>(std::max(call.ctx()->cu_count, 1));
  int tsplit = 1;
  while (tsplit < 8 && 2 * wgs * static_cast<size_t>(tsplit) <= fill) tsplit *= 2;
  if (split > 0) tsplit = split;
  const size_t per_item =
      static_cast<size_t>(qsplit) * static_cast<size_t>(tsplit) * 2 * 3 * kBlock * sizeof(double);
  // partial sums of one launch stay within the context's workspace limit
  size_t budget;
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    budget = call.ctx()->tuning.workspace_limit;
  }
  const size_t ipl = std::max<size_t>(1, std::min(nitems, budget / per_item));
  {
    std::lock_guard<std::mutex> lock(call.ctx()->mu);
    int* last = call.ctx()->last_vjp;
    last[0] = qt;
    last[1] = qsplit;
    last[2] = tsplit;
    last[3] = static_cast<int>((nitems + ipl - 1) / ipl);
  }
  const hipStream_t st = call.stream();
  Scratch d_q(call), d_ab(call), d_blocks(call), d_items(call), d_part(call), d_acc(call),
      d_out(call);
  CSX_HIP(d_q.alloc(nq_pad * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_ab.alloc(AB.size() * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_blocks.alloc(blocks.size() * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(d_items.alloc(items.size() * sizeof(int), st), "hipMallocAsync");
  CSX_HIP(d_part.alloc(ipl * per_item, st), "hipMallocAsync");
  CSX_HIP(d_acc.alloc(static_cast<size_t>(nb) * 3 * kBlock * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(n * 3 * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), qrev.data(), qrev.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_ab.as<void>(), AB.data(), AB.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_blocks.as<void>(), blocks.data(), blocks.size() * sizeof(int),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_items.as<void>(), items.data(), items.size() * sizeof(int),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(launch_debye_vjp(ps->pts32.as<float4>(), d_blocks.as<int4>(), nb, d_items.as<int2>(),
                           nitems, ps->orig32.as<int>(), T, d_q.as<float>(), qstep,
                           d_ab.as<float>(),
                           static_cast<int>(nq_pad), qt, sin_mode, qsplit, tsplit, ipl,
                           d_part.as<double>(),
                           d_acc.as<double>(), d_out.as<double>(), st, call.timer()),
          "debye_vjp_f32");
  CSX_HIP(hipMemcpyAsync(out, d_out.as<void>(), n * 3 * sizeof(double), hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
  d_q.reset();
  d_ab.reset();
  d_blocks.reset();
  d_items.reset();
  d_part.reset();
  d_acc.reset();
  d_out.reset();
  return call.finish();
}

static csx_status jacobian_run(Call& call, const csx_structure* s, PointSet* ps, const double* q,
                               size_t nq, const csx_solvent_params* p, bool implicit_h,
                               double* out) {
  const int T = ps->ntypes;
  const size_t n = s->n;
  const int qt = jacobian_q_tile();
  const size_t nq_pad = (nq + qt - 1) / qt * qt;
  const std::vector<double> A = type_amplitudes_for(s, p, implicit_h, q, nq, T);
  std::vector<float> qrev(nq_pad, 0.0f), Wp(static_cast<size_t>(T) * T * nq_pad, 0.0f);
  for (size_t k = 0; k < nq; ++k) {
    qrev[k] = static_cast<float>(q[k] / (2.0 * kPi));
    // The kernel's u = 2 pi (q/2pi)_f32 r: use the same effective q in W'.
    const double qe = 2.0 * kPi * static_cast<double>(qrev[k]);
    if (qe == 0.0) continue;
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b)
        Wp[(static_cast<size_t>(a) * T + b) * nq_pad + k] =
            static_cast<float>(2.0 * A[a * nq + k] * A[b * nq + k] / qe);
  }
  const hipStream_t st = call.stream();
  const size_t out_elems = nq_pad * 3 * n;
  Scratch d_q(call), d_w(call), d_out(call);
  CSX_HIP(d_q.alloc(nq_pad * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_w.alloc(Wp.size() * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(out_elems * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), qrev.data(), qrev.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_w.as<void>(), Wp.data(), Wp.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  // The unreduced hardware sine takes |q r / 2 pi| <= 256 revolutions.
  double umax = 0.0;
  for (size_t k = 0; k < nq; ++k)
    umax = std::max(umax, std::fabs(static_cast<double>(qrev[k])) * 2.0 * ps->max_radius);
  CSX_HIP(launch_debye_jacobian(ps->pts32.as<float4>(), ps->orig32.as<int>(), ps->type32.as<int>(),
                                static_cast<int>(n), ps->seg32.as<int>(), T, d_q.as<float>(),
                                d_w.as<float>(), static_cast<int>(nq_pad), umax <= 250.0 ? 2 : 1,
                                d_out.as<double>(), st, call.timer()),
          "debye_jacobian_f32");
  CSX_HIP(hipMemcpyAsync(out, d_out.as<void>(), nq * 3 * n * sizeof(double), hipMemcpyDeviceToHost,
                         st), "hipMemcpyAsync");
  d_q.reset();
  d_w.reset();
  d_out.reset();
  return call.finish();
}

// out: jacobian ? nq x n x 3 : n x 3 (weights applied).
static csx_status adjoint(csx_structure* s, const double* q, size_t nq, const double* weights,
                          const csx_solvent_params* p, bool implicit_h, bool jacobian,
                          double* out) {
  if (!s) return fail(CSX_INVALID_INPUT, "structure handle is null");
  if (nq > 0 && (!q || !out)) return fail(CSX_INVALID_INPUT, "q and output buffers are required");
  if (!jacobian && nq > 0 && !weights) return fail(CSX_INVALID_INPUT, "weights are required");
  if (p && p->hydration_model != CSX_HYDRATION_ANALYTIC)
    return fail(CSX_INVALID_INPUT, "the coordinate adjoint supports the Analytic hydration model");
  const size_t n = s->n;
  std::fill(out, out + (jacobian ? nq : 1) * 3 * n, 0.0);
  if (n == 0 || nq == 0) return CSX_OK;
  Call call(s->ctx);
  CSX_TRY(call.begin());
  std::shared_ptr<PointSet> ps;
  CSX_TRY(get_point_set(s, nullptr, &ps));
  return jacobian ? jacobian_run(call, s, ps.get(), q, nq, p, implicit_h, out)
                  : vjp(call, s, ps.get(), q, nq, weights, p, implicit_h, out);
}

}  // namespace csx

using namespace csx;

extern "C" {

csx_status csx_debye_grad_x(csx_structure* s, const double* q, size_t nq, const double* weights,
                            const csx_solvent_params* p, int implicit_h, double* grad) {
  CSX_API_RANGE();
  return adjoint(s, q, nq, weights, p, implicit_h != 0, false, grad);
}

csx_status csx_debye_jacobian(csx_structure* s, const double* q, size_t nq,
                              const csx_solvent_params* p, int implicit_h, double* jac) {
  CSX_API_RANGE();
  return adjoint(s, q, nq, nullptr, p, implicit_h != 0, true, jac);
}

}  // extern "C"

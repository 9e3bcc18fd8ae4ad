// Host side of the Debye adjoint: W' tables, launch, chunk reduction.
#include <algorithm>
#include <cmath>

#include "capi_internal.hpp"

namespace csx {

hipError_t launch_debye_adjoint(const float4* pts, const int* orig, const int* type, int npts,
                                const int* seg_start, int ntypes, const float* qrev,
                                const float* Wp, int nq_pad, bool jacobian, double* out,
                                hipStream_t stream, KernelTimer* timer);
int adjoint_q_tile(bool jacobian);

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
  PointSet* ps = nullptr;
  CSX_TRY(get_point_set(s, nullptr, &ps));
  const int T = ps->ntypes;
  const int qt = adjoint_q_tile(jacobian);
  const size_t nq_pad = (nq + qt - 1) / qt * qt;
  std::vector<float> qrev(nq_pad, 0.0f), Wp(static_cast<size_t>(T) * T * nq_pad, 0.0f);
  std::vector<double> A(static_cast<size_t>(T) * nq);
  for (size_t k = 0; k < nq; ++k) {
    qrev[k] = static_cast<float>(q[k] / (2.0 * kPi));
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
    // The kernel's u = 2 pi (q/2pi)_f32 r: use the same effective q in W'.
    const double qe = 2.0 * kPi * static_cast<double>(qrev[k]);
    if (qe == 0.0) continue;
    const double w = jacobian ? 1.0 : weights[k];
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b)
        Wp[(static_cast<size_t>(a) * T + b) * nq_pad + k] =
            static_cast<float>(2.0 * w * A[a * nq + k] * A[b * nq + k] / qe);
  }
  const hipStream_t st = call.stream();
  const size_t nqc = nq_pad / qt;
  const size_t out_elems = jacobian ? nq_pad * 3 * n : nqc * 3 * n;
  DevBuf d_q, d_w, d_out;
  CSX_HIP(d_q.alloc(nq_pad * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_w.alloc(Wp.size() * sizeof(float), st), "hipMallocAsync");
  CSX_HIP(d_out.alloc(out_elems * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_q.as<void>(), qrev.data(), qrev.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_w.as<void>(), Wp.data(), Wp.size() * sizeof(float),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(launch_debye_adjoint(ps->pts32.as<float4>(), ps->orig32.as<int>(), ps->type32.as<int>(),
                               static_cast<int>(n), ps->seg32.as<int>(), T, d_q.as<float>(),
                               d_w.as<float>(), static_cast<int>(nq_pad), jacobian,
                               d_out.as<double>(), st, call.timer()),
          "debye_adjoint_f32");
  std::vector<double> res(out_elems);
  CSX_HIP(hipMemcpyAsync(res.data(), d_out.as<void>(), out_elems * sizeof(double),
                         hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  d_q.reset();
  d_w.reset();
  d_out.reset();
  CSX_TRY(call.finish());
  if (jacobian) {
    std::copy(res.begin(), res.begin() + nq * 3 * n, out);
  } else {
    for (size_t c = 0; c < nqc; ++c)
      for (size_t k = 0; k < 3 * n; ++k) out[k] += res[c * 3 * n + k];
  }
  return CSX_OK;
}

}  // namespace csx

using namespace csx;

extern "C" {

csx_status csx_debye_grad_x(csx_structure* s, const double* q, size_t nq, const double* weights,
                            const csx_solvent_params* p, int implicit_h, double* grad) {
  return adjoint(s, q, nq, weights, p, implicit_h != 0, false, grad);
}

csx_status csx_debye_jacobian(csx_structure* s, const double* q, size_t nq,
                              const csx_solvent_params* p, int implicit_h, double* jac) {
  return adjoint(s, q, nq, nullptr, p, implicit_h != 0, true, jac);
}

}  // extern "C"

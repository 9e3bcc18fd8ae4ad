// Fitting layer (src/fit.rs) and ensemble weights (src/ensemble.rs) on the
// host, in FP64, over the GPU solvent bases.
//
// Nelder-Mead restates the algorithm of the argmin crate (0.10.0; not part of
// the reference sources) that fit.rs:401-418 drives: reflection 1, expansion
// 2, inside contraction 1/2, shrink 1/2, termination when the standard
// deviation of the simplex costs falls below sd_tolerance or at max_iters.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "capi_internal.hpp"
#include "../../include/crabsaxs_tables.h"

namespace csx {

// ----------------------------------------------------- linear solves --

struct Lin { double scale, bg, chi; };

// solve_linear / solve_scale_background, fit.rs:137-244.
static Lin solve_scale_background(const double* x, const double* y, const double* s, size_t n,
                                  int scale_mode, double scale_value, int bg_mode,
                                  double bg_value) {
  double scale = 0.0, bg = 0.0;
  if (scale_mode == CSX_SCALE_FIT) {
    double sw = 0, sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (size_t i = 0; i < n; ++i) {
      const double w = 1.0 / (s[i] * s[i]);
      sw += w;
      sx += w * x[i];
      sy += w * y[i];
      sxx += w * x[i] * x[i];
      sxy += w * x[i] * y[i];
    }
    if (bg_mode == CSX_BACKGROUND_FIT) {
      const double det = sxx * sw - sx * sx;
      scale = std::max(std::fabs(det) > 1e-20 ? (sxy * sw - sx * sy) / det : 0.0, 0.0);
      bg = sw > 0.0 ? (sy - scale * sx) / sw : 0.0;
    } else if (bg_mode == CSX_BACKGROUND_FIXED_ZERO) {
      scale = std::max(sxx > 0.0 ? sxy / sxx : 0.0, 0.0);
    } else {
      scale = std::max(sxx > 0.0 ? (sxy - bg_value * sx) / sxx : 0.0, 0.0);
      bg = bg_value;
    }
  } else {
    scale = std::max(scale_value, 0.0);
    if (bg_mode == CSX_BACKGROUND_FIT) {
      double wt = 0.0, w = 0.0;
      for (size_t i = 0; i < n; ++i) {
        const double ww = 1.0 / (s[i] * s[i]);
        wt += ww * (y[i] - scale * x[i]);
        w += ww;
      }
      bg = w > 0.0 ? wt / w : 0.0;
    } else {
      bg = bg_mode == CSX_BACKGROUND_FIXED_ZERO ? 0.0 : bg_value;
    }
  }
  const double dof = n > 1 ? static_cast<double>(n - 1) : 1.0;
  double chi = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double r = (scale * x[i] + bg - y[i]) / s[i];
    chi += r * r;
  }
  return {scale, bg, chi / dof};
}

// sampled_indices, fit.rs:84-91.
static std::vector<size_t> sampled_indices(size_t length, size_t stride) {
  if (length < 64) stride = 1;
  stride = std::max<size_t>(stride, 1);
  std::vector<size_t> idx;
  for (size_t i = 0; i < length; i += stride) idx.push_back(i);
  if (idx.empty() || idx.back() != length - 1) idx.push_back(length - 1);
  return idx;
}

// cubic_hermite_interpolate, fit.rs:93-123.
static std::vector<double> cubic_hermite(const std::vector<double>& x, const double* y,
                                         const double* t, size_t nt) {
  std::vector<double> out(nt);
  if (x.size() == 1) {
    std::fill(out.begin(), out.end(), y[0]);
    return out;
  }
  const size_t last = x.size() - 1;
  std::vector<double> sl(x.size());
  sl[0] = (y[1] - y[0]) / (x[1] - x[0]);
  sl[last] = (y[last] - y[last - 1]) / (x[last] - x[last - 1]);
  for (size_t i = 1; i < last; ++i) sl[i] = (y[i + 1] - y[i - 1]) / (x[i + 1] - x[i - 1]);
  size_t seg = 0;
  for (size_t k = 0; k < nt; ++k) {
    while (seg + 1 < last && t[k] > x[seg + 1]) ++seg;
    const double h = x[seg + 1] - x[seg];
    const double u = std::min(std::max((t[k] - x[seg]) / h, 0.0), 1.0);
    const double u2 = u * u, u3 = u2 * u;
    out[k] = (2.0 * u3 - 3.0 * u2 + 1.0) * y[seg] + (u3 - 2.0 * u2 + u) * h * sl[seg] +
             (-2.0 * u3 + 3.0 * u2) * y[seg + 1] + (u3 - u2) * h * sl[seg + 1];
  }
  return out;
}

// combine_solvent_basis and its derivatives in c1 and c2 (intensity clamped
// at zero has zero derivative).
static void combine_with_grad(const double* b, const double* q, size_t n, double c1, double c2,
                              double* I, double* dI1, double* dI2) {
  const double r2 = CSX_MEAN_ATOMIC_RADIUS * CSX_MEAN_ATOMIC_RADIUS;
  for (size_t i = 0; i < n; ++i) {
    const double g = excluded_volume_scale(q[i], c1);
    const double v = b[i] + g * g * b[n + i] + c2 * c2 * b[2 * n + i] - 2.0 * g * b[3 * n + i] +
                     2.0 * c2 * b[4 * n + i] - 2.0 * g * c2 * b[5 * n + i];
    I[i] = std::max(v, 0.0);
    if (dI1) {
      const double dg = g * (3.0 / c1 - c1 * r2 * q[i] * q[i] / 2.0);
      const double dIdg = 2.0 * g * b[n + i] - 2.0 * b[3 * n + i] - 2.0 * c2 * b[5 * n + i];
      dI1[i] = v > 0.0 ? dIdg * dg : 0.0;
      dI2[i] = v > 0.0 ? 2.0 * c2 * b[2 * n + i] + 2.0 * b[4 * n + i] - 2.0 * g * b[5 * n + i]
                       : 0.0;
    }
  }
}

// ----------------------------------------------------------- objective --

struct Objective {
  const std::vector<double>* basis;  // 6 * n on the experimental grid
  const double *q, *y, *s;
  size_t n;
  double b1[2], b2[2];
  int scale_mode, bg_mode;
  double scale_value, bg_value;
  mutable std::vector<double> I, d1, d2;

  // SolventObjective::cost, fit.rs:260-272.
  double cost(double c1, double c2) const {
    if (!(c1 >= b1[0] && c1 <= b1[1]) || !(c2 >= b2[0] && c2 <= b2[1])) {
      const double d1v = std::max(b1[0] - c1, 0.0) + std::max(c1 - b1[1], 0.0);
      const double d2v = std::max(b2[0] - c2, 0.0) + std::max(c2 - b2[1], 0.0);
      return 1e20 + 1e20 * (d1v * d1v + d2v * d2v);
    }
    I.resize(n);
    combine_with_grad(basis->data(), q, n, c1, c2, I.data(), nullptr, nullptr);
    return solve_scale_background(I.data(), y, s, n, scale_mode, scale_value, bg_mode, bg_value).chi;
  }

  // chi^2 at the optimal scale/background and its gradient in (c1, c2)
  // (envelope theorem: the inner least squares is stationary).
  double cost_grad(double c1, double c2, double* g) const {
    I.resize(n);
    d1.resize(n);
    d2.resize(n);
    combine_with_grad(basis->data(), q, n, c1, c2, I.data(), d1.data(), d2.data());
    const Lin lin = solve_scale_background(I.data(), y, s, n, scale_mode, scale_value, bg_mode,
                                           bg_value);
    const double dof = n > 1 ? static_cast<double>(n - 1) : 1.0;
    double g1 = 0.0, g2 = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double r = (lin.scale * I[i] + lin.bg - y[i]) / (s[i] * s[i]);
      g1 += r * d1[i];
      g2 += r * d2[i];
    }
    g[0] = 2.0 * lin.scale * g1 / dof;
    g[1] = 2.0 * lin.scale * g2 / dof;
    return lin.chi;
  }
};

struct OptResult {
  double c1, c2;
  int64_t iterations;
  bool converged;
  std::string termination;
};

// argmin NelderMead with sd_tolerance; the 2-D simplex of fit.rs:387-400.
static OptResult nelder_mead(const Objective& f, double x1, double x2, double step1, double step2,
                             int max_iters, double sd_tol) {
  auto offset = [](double v, double step, const double* b) { return v + step <= b[1] ? v + step : v - step; };
  struct V { double x[2]; double c; };
  V s[3] = {{{x1, x2}, 0}, {{offset(x1, step1, f.b1), x2}, 0}, {{x1, offset(x2, step2, f.b2)}, 0}};
  for (auto& v : s) v.c = f.cost(v.x[0], v.x[1]);
  auto sort = [&]() { std::stable_sort(s, s + 3, [](const V& a, const V& b) { return a.c < b.c; }); };
  sort();
  auto sd = [&]() {
    const double mean = (s[0].c + s[1].c + s[2].c) / 3.0;
    double var = 0.0;
    for (auto& v : s) var += (v.c - mean) * (v.c - mean);
    return std::sqrt(var / 3.0);
  };
  int64_t it = 0;
  OptResult r{x1, x2, 0, false, "Maximum number of iterations reached"};
  while (true) {
    if (!std::isfinite(s[0].c)) {
      r.termination = "objective was not finite";
      break;
    }
    if (sd() < sd_tol) {
      r.converged = true;
      r.termination = "Solver converged";
      break;
    }
    if (it >= max_iters) break;
    const double c0[2] = {(s[0].x[0] + s[1].x[0]) / 2.0, (s[0].x[1] + s[1].x[1]) / 2.0};
    V xr{{c0[0] + (c0[0] - s[2].x[0]), c0[1] + (c0[1] - s[2].x[1])}, 0};
    xr.c = f.cost(xr.x[0], xr.x[1]);
    if (xr.c < s[1].c && xr.c >= s[0].c) {
      s[2] = xr;
    } else if (xr.c < s[0].c) {
      V xe{{c0[0] + 2.0 * (xr.x[0] - c0[0]), c0[1] + 2.0 * (xr.x[1] - c0[1])}, 0};
      xe.c = f.cost(xe.x[0], xe.x[1]);
      s[2] = xe.c < xr.c ? xe : xr;
    } else {
      V xc{{c0[0] + 0.5 * (s[2].x[0] - c0[0]), c0[1] + 0.5 * (s[2].x[1] - c0[1])}, 0};
      xc.c = f.cost(xc.x[0], xc.x[1]);
      if (xc.c < s[2].c) {
        s[2] = xc;
      } else {
        for (int k = 1; k < 3; ++k) {
          for (int d = 0; d < 2; ++d) s[k].x[d] = s[0].x[d] + 0.5 * (s[k].x[d] - s[0].x[d]);
          s[k].c = f.cost(s[k].x[0], s[k].x[1]);
        }
      }
    }
    sort();
    ++it;
  }
  r.c1 = s[0].x[0];
  r.c2 = s[0].x[1];
  r.iterations = it;
  return r;
}

// Projected quasi-Newton (BFGS on the box, Armijo backtracking) with the
// analytic gradient of chi^2 in (c1, c2).
static OptResult gradient_fit(const Objective& f, double x1, double x2, int max_iters) {
  const double lo[2] = {f.b1[0], f.b2[0]}, hi[2] = {f.b1[1], f.b2[1]};
  const double span[2] = {hi[0] - lo[0], hi[1] - lo[1]};
  // Start from the best of the initial point and a 5 x 5 grid over the box:
  // where the fitted scale is clamped at zero chi^2 is flat in (c1, c2) and a
  // gradient method cannot leave that region by itself.
  double best = f.cost(x1, x2);
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      const double a = lo[0] + span[0] * i / 4.0, b = lo[1] + span[1] * j / 4.0;
      const double c = f.cost(a, b);
      if (c < best) {
        best = c;
        x1 = a;
        x2 = b;
      }
    }
  double x[2] = {x1, x2}, g[2];
  double fx = f.cost_grad(x[0], x[1], g);
  double H[2][2] = {{span[0] * span[0], 0}, {0, span[1] * span[1]}};  // inverse Hessian
  const double hn = std::max(std::fabs(g[0]) * span[0], std::fabs(g[1]) * span[1]);
  if (hn > 0) {
    H[0][0] *= 0.1 / hn;
    H[1][1] *= 0.1 / hn;
  }
  OptResult r{x1, x2, 0, false, "Maximum number of iterations reached"};
  int64_t it = 0;
  for (; it < max_iters; ++it) {
    // projected gradient norm (scaled to the box)
    double pg = 0.0;
    for (int d = 0; d < 2; ++d) {
      const double xp = std::min(std::max(x[d] - g[d], lo[d]), hi[d]);
      pg = std::max(pg, std::fabs(xp - x[d]) / span[d]);
    }
    if (pg < 1e-10 || !std::isfinite(fx)) {
      r.converged = std::isfinite(fx);
      r.termination = r.converged ? "Solver converged" : "objective was not finite";
      break;
    }
    double p[2] = {-(H[0][0] * g[0] + H[0][1] * g[1]), -(H[1][0] * g[0] + H[1][1] * g[1])};
    if (p[0] * g[0] + p[1] * g[1] >= 0) {  // not a descent direction: reset
      H[0][0] = span[0] * span[0] * 1e-3;
      H[1][1] = span[1] * span[1] * 1e-3;
      H[0][1] = H[1][0] = 0.0;
      p[0] = -H[0][0] * g[0];
      p[1] = -H[1][1] * g[1];
    }
    double t = 1.0, xn[2], gn[2], fn = fx;
    bool accepted = false;
    for (int ls = 0; ls < 40; ++ls, t *= 0.5) {
      for (int d = 0; d < 2; ++d) xn[d] = std::min(std::max(x[d] + t * p[d], lo[d]), hi[d]);
      fn = f.cost_grad(xn[0], xn[1], gn);
      const double dec = g[0] * (xn[0] - x[0]) + g[1] * (xn[1] - x[1]);
      if (fn <= fx + 1e-4 * dec) {
        accepted = true;
        break;
      }
    }
    const double sx[2] = {xn[0] - x[0], xn[1] - x[1]};
    if (!accepted || (std::fabs(sx[0]) < 1e-14 * span[0] && std::fabs(sx[1]) < 1e-14 * span[1])) {
      r.converged = true;
      r.termination = "Solver converged";
      break;
    }
    const double yv[2] = {gn[0] - g[0], gn[1] - g[1]};
    const double sy = sx[0] * yv[0] + sx[1] * yv[1];
    if (sy > 1e-300) {
      // BFGS update of the inverse Hessian
      const double Hy[2] = {H[0][0] * yv[0] + H[0][1] * yv[1], H[1][0] * yv[0] + H[1][1] * yv[1]};
      const double yHy = yv[0] * Hy[0] + yv[1] * Hy[1];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          H[a][b] += (sy + yHy) * sx[a] * sx[b] / (sy * sy) - (Hy[a] * sx[b] + sx[a] * Hy[b]) / sy;
    }
    x[0] = xn[0];
    x[1] = xn[1];
    g[0] = gn[0];
    g[1] = gn[1];
    fx = fn;
  }
  r.c1 = x[0];
  r.c2 = x[1];
  r.iterations = it;
  return r;
}

static bool fit_options_valid(const csx_fit_options* o) {
  const bool scale_ok = o->scale_mode == CSX_SCALE_FIT ||
                        (o->scale_mode == CSX_SCALE_FIXED && std::isfinite(o->scale_value) &&
                         o->scale_value >= 0.0);
  const bool bg_ok = o->background_mode == CSX_BACKGROUND_FIT ||
                     o->background_mode == CSX_BACKGROUND_FIXED_ZERO ||
                     (o->background_mode == CSX_BACKGROUND_FIXED && std::isfinite(o->background_value));
  return o->q_sampling_stride > 0 && o->max_iterations > 0 && std::isfinite(o->c1_bounds[0]) &&
         std::isfinite(o->c1_bounds[1]) && o->c1_bounds[0] > 0.0 &&
         o->c1_bounds[0] < o->c1_bounds[1] && std::isfinite(o->c2_bounds[0]) &&
         std::isfinite(o->c2_bounds[1]) && o->c2_bounds[0] < o->c2_bounds[1] && scale_ok &&
         bg_ok && (o->optimizer == CSX_OPTIMIZER_NELDER_MEAD || o->optimizer == CSX_OPTIMIZER_GRADIENT);
}

static bool curve_valid(const double* q, const double* y, const double* s, size_t n) {
  if (n == 0 || !q || !y || !s) return false;
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(q[i]) || q[i] < 0.0 || (i > 0 && q[i] < q[i - 1])) return false;
    if (!std::isfinite(y[i]) || !std::isfinite(s[i]) || s[i] <= 0.0) return false;
  }
  return true;
}

// The method choice of fit.rs:352-359 / scattering.rs:145-156.
static int choose_method(int method, const csx_solvent_params& p, size_t n) {
  if (method != CSX_METHOD_AUTO) return method;
  if (p.hydration_model != CSX_HYDRATION_ANALYTIC) return CSX_METHOD_MULTIPOLE;
  return n < 500 ? CSX_METHOD_DEBYE : CSX_METHOD_MULTIPOLE;
}

static int fit_l_max(const csx_structure* s, const double* q, size_t n) {
  double qmax = 0.0;
  for (size_t i = 0; i < n; ++i) qmax = std::max(qmax, q[i]);
  const double o = std::ceil(qmax * s->max_radius);
  return static_cast<int>(std::min(std::max(std::min(o, 1e6) + 8.0, 8.0), 64.0));
}

// Six-term basis of one structure on the experimental grid, evaluated at the
// stride-sampled q values and Hermite-interpolated (fit.rs:344-384).
static csx_status fit_basis(csx_structure* s, const double* q, size_t n, const csx_fit_options* o,
                            std::vector<double>& basis) {
  const std::vector<size_t> idx = sampled_indices(n, static_cast<size_t>(o->q_sampling_stride));
  std::vector<double> sq;
  for (size_t k : idx)
    if (sq.empty() || q[k] != sq.back()) sq.push_back(q[k]);
  const int method = choose_method(o->method, o->solvent, s->n);
  std::vector<double> sb(6 * sq.size());
  if (method == CSX_METHOD_DEBYE) {
    CSX_TRY(csx_debye_solvent_basis(s, sq.data(), sq.size(), &o->solvent, o->implicit_hydrogen,
                                    o->precision, sb.data()));
  } else {
    CSX_TRY(csx_multipole_solvent_basis(s, sq.data(), sq.size(), fit_l_max(s, q, n),
                                        o->implicit_hydrogen, &o->solvent, sb.data()));
  }
  basis.assign(6 * n, 0.0);
  if (sq.size() == n) {
    basis = sb;
    return CSX_OK;
  }
  for (int t = 0; t < 6; ++t) {
    const std::vector<double> v = cubic_hermite(sq, sb.data() + t * sq.size(), q, n);
    std::copy(v.begin(), v.end(), basis.begin() + t * n);
  }
  return CSX_OK;
}

static csx_status fit_from_basis(const std::vector<double>& basis, const double* q, const double* y,
                                 const double* sg, size_t n, const csx_fit_params* initial,
                                 const csx_fit_options* o, csx_fit_result* r, double* fitted) {
  Objective f;
  f.basis = &basis;
  f.q = q;
  f.y = y;
  f.s = sg;
  f.n = n;
  f.b1[0] = o->c1_bounds[0];
  f.b1[1] = o->c1_bounds[1];
  f.b2[0] = o->c2_bounds[0];
  f.b2[1] = o->c2_bounds[1];
  f.scale_mode = o->scale_mode;
  f.bg_mode = o->background_mode;
  f.scale_value = o->scale_value;
  f.bg_value = o->background_value;
  const double c1 = std::min(std::max(initial->c1, f.b1[0]), f.b1[1]);
  const double c2 = std::min(std::max(initial->c2, f.b2[0]), f.b2[1]);
  OptResult opt = o->optimizer == CSX_OPTIMIZER_GRADIENT
                      ? gradient_fit(f, c1, c2, o->max_iterations)
                      : nelder_mead(f, c1, c2, (f.b1[1] - f.b1[0]) * 0.1,
                                    (f.b2[1] - f.b2[0]) * 0.1, o->max_iterations, 1e-8);
  const double bc1 = std::min(std::max(opt.c1, f.b1[0]), f.b1[1]);
  const double bc2 = std::min(std::max(opt.c2, f.b2[0]), f.b2[1]);
  std::vector<double> raw(n);
  combine_with_grad(basis.data(), q, n, bc1, bc2, raw.data(), nullptr, nullptr);
  const Lin lin = solve_scale_background(raw.data(), y, sg, n, o->scale_mode, o->scale_value,
                                         o->background_mode, o->background_value);
  if (!std::isfinite(lin.chi)) return fail(CSX_NOT_CONVERGED, "objective was not finite");
  r->params = {lin.scale, lin.bg, bc1, bc2};
  r->chi2 = lin.chi;
  r->iterations = opt.iterations;
  r->converged = opt.converged ? 1 : 0;
  std::snprintf(r->termination, sizeof r->termination, "%s", opt.termination.c_str());
  if (fitted)
    for (size_t i = 0; i < n; ++i) fitted[i] = lin.scale * raw[i] + lin.bg;
  return CSX_OK;
}

// ------------------------------------------------------------ ensemble --

// project_simplex, ensemble.rs:449-465.
static std::vector<double> project_simplex(const std::vector<double>& v) {
  std::vector<double> sorted = v;
  std::sort(sorted.begin(), sorted.end(), [](double a, double b) { return a > b; });
  double cum = 0.0;
  size_t rho = 0;
  for (size_t i = 0; i < sorted.size(); ++i) {
    cum += sorted[i];
    if (sorted[i] + (1.0 - cum) / static_cast<double>(i + 1) > 0.0) rho = i + 1;
  }
  double top = 0.0;
  for (size_t i = 0; i < rho; ++i) top += sorted[i];
  const double theta = (top - 1.0) / static_cast<double>(std::max<size_t>(rho, 1));
  std::vector<double> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = std::max(v[i] - theta, 0.0);
  return out;
}

// solve_linear_fixed, ensemble.rs:383-434.
static Lin linear_fixed(const double* x, const double* y, const double* s, size_t n,
                        const csx_fit_options* o) {
  double scale;
  if (o->scale_mode == CSX_SCALE_FIT) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
      num += x[i] * y[i] / (s[i] * s[i]);
      den += x[i] * x[i] / (s[i] * s[i]);
    }
    scale = den > 0.0 ? num / den : 0.0;
  } else {
    scale = o->scale_value;
  }
  scale = std::max(scale, 0.0);
  double bg;
  if (o->background_mode == CSX_BACKGROUND_FIT) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
      num += (y[i] - scale * x[i]) / (s[i] * s[i]);
      den += 1.0 / (s[i] * s[i]);
    }
    bg = den > 0.0 ? num / den : 0.0;
  } else {
    bg = o->background_mode == CSX_BACKGROUND_FIXED_ZERO ? 0.0 : o->background_value;
  }
  double chi = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double r = (scale * x[i] + bg - y[i]) / s[i];
    chi += r * r;
  }
  return {scale, bg, chi / (n > 1 ? static_cast<double>(n - 1) : 1.0)};
}

struct WeightFit {
  std::vector<double> weights, fitted;
  Lin lin;
};

// optimize_weights, ensemble.rs:317-381.
static WeightFit optimize_weights(const std::vector<const double*>& curves, const double* y,
                                  const double* s, size_t n, const csx_fit_options* o, bool uniform,
                                  double reg) {
  const size_t m = curves.size();
  std::vector<double> w(m, 1.0 / static_cast<double>(m)), mixture(n), res(n), g(m), prop(m);
  auto mix = [&]() {
    for (size_t p = 0; p < n; ++p) {
      double v = 0.0;
      for (size_t c = 0; c < m; ++c) v += curves[c][p] * w[c];
      mixture[p] = v;
    }
  };
  if (!uniform) {
    for (int it = 0; it < 250; ++it) {
      mix();
      const Lin lin = linear_fixed(mixture.data(), y, s, n, o);
      for (size_t p = 0; p < n; ++p) res[p] = (lin.scale * mixture[p] + lin.bg - y[p]) / s[p];
      for (size_t c = 0; c < m; ++c) {
        double v = 0.0;
        for (size_t p = 0; p < n; ++p) v += res[p] * lin.scale * curves[c][p] / s[p];
        g[c] = 2.0 * v + 2.0 * reg * (w[c] - 1.0 / static_cast<double>(m));
      }
      double norm = 0.0;
      for (double x : g) norm += x * x;
      norm = std::max(std::sqrt(norm), 1.0);
      for (size_t c = 0; c < m; ++c) prop[c] = w[c] - 0.05 * g[c] / norm;
      prop = project_simplex(prop);
      double delta = 0.0;
      for (size_t c = 0; c < m; ++c) delta += std::fabs(prop[c] - w[c]);
      w = prop;
      if (delta < 1e-8) break;
    }
  }
  mix();
  WeightFit r;
  r.lin = linear_fixed(mixture.data(), y, s, n, o);
  r.weights = w;
  r.fitted.resize(n);
  for (size_t p = 0; p < n; ++p) r.fitted[p] = r.lin.scale * mixture[p] + r.lin.bg;
  return r;
}

// Least squares min |E[:, P] z - f| over the columns P of the n x m
// column-major E, by Householder QR of the n x |P| submatrix.  Returns false
// when a column is (numerically) dependent on the others.
static bool passive_lsq(const std::vector<double>& E, size_t n, const std::vector<size_t>& P,
                        const std::vector<double>& f, std::vector<double>& z) {
  const size_t k = P.size();
  if (k > n) return false;
  std::vector<double> A(n * k), b(f);
  for (size_t c = 0; c < k; ++c) std::copy(&E[P[c] * n], &E[P[c] * n] + n, &A[c * n]);
  std::vector<double> diag(k);
  for (size_t c = 0; c < k; ++c) {
    double norm = 0.0;
    for (size_t r = c; r < n; ++r) norm += A[c * n + r] * A[c * n + r];
    norm = std::sqrt(norm);
    if (!(norm > 1e-12)) return false;  // columns are unit-norm: rank loss
    const double alpha = A[c * n + c] > 0.0 ? -norm : norm;
    A[c * n + c] -= alpha;
    double vnorm = 0.0;
    for (size_t r = c; r < n; ++r) vnorm += A[c * n + r] * A[c * n + r];
    for (size_t cc = c + 1; cc < k; ++cc) {
      double dot = 0.0;
      for (size_t r = c; r < n; ++r) dot += A[c * n + r] * A[cc * n + r];
      const double t = 2.0 * dot / vnorm;
      for (size_t r = c; r < n; ++r) A[cc * n + r] -= t * A[c * n + r];
    }
    double dot = 0.0;
    for (size_t r = c; r < n; ++r) dot += A[c * n + r] * b[r];
    const double t = 2.0 * dot / vnorm;
    for (size_t r = c; r < n; ++r) b[r] -= t * A[c * n + r];
    diag[c] = alpha;
  }
  z.assign(k, 0.0);
  for (size_t c = k; c-- > 0;) {
    double v = b[c];
    for (size_t cc = c + 1; cc < k; ++cc) v -= A[cc * n + c] * z[cc];
    z[c] = v / diag[c];
  }
  return true;
}

// Lawson-Hanson active-set NNLS (Lawson & Hanson, "Solving Least Squares
// Problems", 1974, ch. 23): min |E a - f| subject to a >= 0.  E is n x m
// column-major with unit-norm (or zero) columns.
static std::vector<double> nnls(const std::vector<double>& E, size_t n, size_t m,
                                const std::vector<double>& f) {
  std::vector<double> a(m, 0.0), res(f), grad(m), z;
  std::vector<char> passive(m, 0), blocked(m, 0);
  double fnorm = 0.0;
  for (double v : f) fnorm += v * v;
  const double tol = 1e-12 * std::max(std::sqrt(fnorm), 1e-300);
  auto gradient = [&]() {
    for (size_t p = 0; p < n; ++p) {
      double v = f[p];
      for (size_t c = 0; c < m; ++c)
        if (a[c] != 0.0) v -= E[c * n + p] * a[c];
      res[p] = v;
    }
    for (size_t c = 0; c < m; ++c) {
      double v = 0.0;
      for (size_t p = 0; p < n; ++p) v += E[c * n + p] * res[p];
      grad[c] = v;
    }
  };
  gradient();
  for (size_t outer = 0; outer < 3 * m + 10; ++outer) {
    size_t t = m;
    double best = tol;
    for (size_t c = 0; c < m; ++c)
      if (!passive[c] && !blocked[c] && grad[c] > best) best = grad[c], t = c;
    if (t == m) break;
    passive[t] = 1;
    for (size_t inner = 0; inner < 3 * m + 10; ++inner) {
      std::vector<size_t> P;
      for (size_t c = 0; c < m; ++c)
        if (passive[c]) P.push_back(c);
      if (!passive_lsq(E, n, P, f, z)) {
        // the new column is dependent on the passive set: keep it out
        passive[t] = 0;
        blocked[t] = 1;
        break;
      }
      bool feasible = true;
      for (size_t k = 0; k < P.size(); ++k)
        if (z[k] <= 0.0) feasible = false;
      if (feasible) {
        for (size_t k = 0; k < P.size(); ++k) a[P[k]] = z[k];
        break;
      }
      if (inner == 0 && z[std::find(P.begin(), P.end(), t) - P.begin()] <= 0.0) {
        // rounding made the entering coefficient non-positive: drop it
        passive[t] = 0;
        blocked[t] = 1;
        break;
      }
      // step from a towards z until the first passive coefficient hits 0
      double alpha = 2.0;
      size_t kmin = 0;
      for (size_t k = 0; k < P.size(); ++k)
        if (z[k] <= 0.0) {
          const double r = a[P[k]] / (a[P[k]] - z[k]);
          if (r < alpha) alpha = r, kmin = k;
        }
      for (size_t k = 0; k < P.size(); ++k) {
        a[P[k]] += alpha * (z[k] - a[P[k]]);
        if (k == kmin || a[P[k]] <= 0.0) {
          a[P[k]] = 0.0;
          passive[P[k]] = 0;
        }
      }
    }
    gradient();
  }
  return a;
}

// Weights by NNLS: with background b eliminated (fitted: projected out of
// the sigma-weighted system; fixed: subtracted), a = scale * weights >= 0
// minimises sum ((X a + b - y) / sigma)^2; scale = sum a.
static csx_status nnls_weights(const std::vector<const double*>& curves, const double* y,
                               const double* s, size_t n, const csx_fit_options* o,
                               WeightFit& out) {
  if (o->scale_mode != CSX_SCALE_FIT)
    return fail(CSX_INVALID_INPUT, "NNLS weights need a fitted scale");
  const size_t m = curves.size();
  std::vector<double> E(n * m), f(n), u(n);
  const bool fit_bg = o->background_mode == CSX_BACKGROUND_FIT;
  const double bg0 = o->background_mode == CSX_BACKGROUND_FIXED ? o->background_value : 0.0;
  double uu = 0.0;
  for (size_t p = 0; p < n; ++p) {
    u[p] = 1.0 / s[p];
    uu += u[p] * u[p];
  }
  auto project = [&](double* v) {  // remove the component along u
    if (!fit_bg) return;
    double d = 0.0;
    for (size_t p = 0; p < n; ++p) d += v[p] * u[p];
    for (size_t p = 0; p < n; ++p) v[p] -= d / uu * u[p];
  };
  for (size_t p = 0; p < n; ++p) f[p] = (y[p] - bg0) / s[p];
  project(f.data());
  std::vector<double> norm(m, 0.0);
  for (size_t c = 0; c < m; ++c) {
    double* col = &E[c * n];
    for (size_t p = 0; p < n; ++p) col[p] = curves[c][p] / s[p];
    project(col);
    double v = 0.0;
    for (size_t p = 0; p < n; ++p) v += col[p] * col[p];
    norm[c] = std::sqrt(v);
    for (size_t p = 0; p < n; ++p) col[p] = norm[c] > 0.0 ? col[p] / norm[c] : 0.0;
  }
  std::vector<double> a = nnls(E, n, m, f);
  double scale = 0.0;
  for (size_t c = 0; c < m; ++c) {
    a[c] = norm[c] > 0.0 ? a[c] / norm[c] : 0.0;
    scale += a[c];
  }
  out.weights.assign(m, 1.0 / static_cast<double>(m));
  if (scale > 0.0)
    for (size_t c = 0; c < m; ++c) out.weights[c] = a[c] / scale;
  // Joint optimum: scale = sum a and the background that goes with it (not
  // the sequential scale-then-background solve of the projected-gradient
  // path, which is biased when the background is non-zero).
  std::vector<double> model(n, 0.0);
  double num = 0.0;
  for (size_t p = 0; p < n; ++p) {
    double v = 0.0;
    for (size_t c = 0; c < m; ++c) v += curves[c][p] * a[c];
    model[p] = v;
    num += (y[p] - v) * u[p] * u[p];
  }
  const double bg = fit_bg ? num / uu : bg0;
  double chi = 0.0;
  out.fitted.resize(n);
  for (size_t p = 0; p < n; ++p) {
    out.fitted[p] = model[p] + bg;
    const double r = (out.fitted[p] - y[p]) / s[p];
    chi += r * r;
  }
  out.lin = {scale, bg, chi / (n > 1 ? static_cast<double>(n - 1) : 1.0)};
  return CSX_OK;
}

}  // namespace csx

using namespace csx;

extern "C" {

void csx_fit_options_default(csx_fit_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->method = CSX_METHOD_AUTO;
  o->implicit_hydrogen = 1;
  o->c1_bounds[0] = 0.95;
  o->c1_bounds[1] = 1.05;
  o->c2_bounds[0] = -2.0;
  o->c2_bounds[1] = 4.0;
  o->max_iterations = 64;
  o->q_sampling_stride = 20;
  csx_solvent_default(&o->solvent);
  o->background_mode = CSX_BACKGROUND_FIT;
  o->scale_mode = CSX_SCALE_FIT;
  o->optimizer = CSX_OPTIMIZER_NELDER_MEAD;
  o->precision = CSX_FP32_COMPENSATED;
}

csx_status csx_solve_scale_background(const double* theory, const double* intensity,
                                      const double* sigma, size_t n, int scale_mode,
                                      double scale_value, int bg_mode, double bg_value,
                                      double* out3) {
  if (!out3 || (n > 0 && (!theory || !intensity || !sigma)))
    return fail(CSX_INVALID_INPUT, "buffers are required");
  const Lin l = solve_scale_background(theory, intensity, sigma, n, scale_mode, scale_value,
                                       bg_mode, bg_value);
  out3[0] = l.scale;
  out3[1] = l.bg;
  out3[2] = l.chi;
  return CSX_OK;
}

csx_status csx_cubic_hermite_interpolate(const double* x, const double* y, size_t n,
                                         const double* targets, size_t nt, double* out) {
  if (n == 0 || !x || !y || (nt > 0 && (!targets || !out)))
    return fail(CSX_INVALID_INPUT, "buffers are required");
  const std::vector<double> v = cubic_hermite(std::vector<double>(x, x + n), y, targets, nt);
  std::copy(v.begin(), v.end(), out);
  return CSX_OK;
}

csx_status csx_chi2_grad(const double* basis, const double* q, const double* intensity,
                         const double* sigma, size_t n, const csx_fit_params* p, double* chi2,
                         double* grad4) {
  CSX_API_RANGE();
  if (!basis || !q || !intensity || !sigma || !p || n == 0)
    return fail(CSX_INVALID_INPUT, "buffers are required");
  std::vector<double> I(n), d1(n), d2(n);
  combine_with_grad(basis, q, n, p->c1, p->c2, I.data(), d1.data(), d2.data());
  const double dof = n > 1 ? static_cast<double>(n - 1) : 1.0;
  double c = 0.0, gs = 0.0, gb = 0.0, g1 = 0.0, g2 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double r = (p->scale * I[i] + p->background - intensity[i]) / sigma[i];
    const double rw = r / sigma[i];
    c += r * r;
    gs += rw * I[i];
    gb += rw;
    g1 += rw * p->scale * d1[i];
    g2 += rw * p->scale * d2[i];
  }
  if (chi2) *chi2 = c / dof;
  if (grad4) {
    grad4[0] = 2.0 * gs / dof;
    grad4[1] = 2.0 * gb / dof;
    grad4[2] = 2.0 * g1 / dof;
    grad4[3] = 2.0 * g2 / dof;
  }
  return CSX_OK;
}

csx_status csx_fit_basis(csx_structure* s, const double* q, size_t n, const csx_fit_options* o,
                         double* basis_out) {
  CSX_API_RANGE();
  if (!s || !o || !basis_out || !q || n == 0) return fail(CSX_INVALID_INPUT, "arguments are required");
  if (!fit_options_valid(o))
    return fail(CSX_INVALID_INPUT, "fit bounds, iteration limit, and q sampling stride are invalid");
  std::vector<double> basis;
  CSX_TRY(fit_basis(s, q, n, o, basis));
  std::copy(basis.begin(), basis.end(), basis_out);
  return CSX_OK;
}

csx_status csx_fit_to_experimental(csx_structure* s, const double* q, const double* intensity,
                                   const double* sigma, size_t n, const csx_fit_params* initial,
                                   const csx_fit_options* o, csx_fit_result* result,
                                   double* fitted) {
  CSX_API_RANGE();
  if (!s || !o || !result || !initial) return fail(CSX_INVALID_INPUT, "arguments are required");
  if (!curve_valid(q, intensity, sigma, n) || s->n == 0)
    return fail(CSX_INVALID_INPUT,
                "fit requires a non-empty structure and equal-length finite q, intensity, and "
                "positive-sigma arrays on a non-decreasing q grid");
  if (!fit_options_valid(o))
    return fail(CSX_INVALID_INPUT, "fit bounds, iteration limit, and q sampling stride are invalid");
  std::vector<double> basis;
  CSX_TRY(fit_basis(s, q, n, o, basis));
  return fit_from_basis(basis, q, intensity, sigma, n, initial, o, result, fitted);
}

csx_status csx_fit_from_basis(const double* basis, const double* q, const double* intensity,
                              const double* sigma, size_t n, const csx_fit_params* initial,
                              const csx_fit_options* o, csx_fit_result* result, double* fitted) {
  CSX_API_RANGE();
  if (!basis || !o || !result || !initial) return fail(CSX_INVALID_INPUT, "arguments are required");
  if (!curve_valid(q, intensity, sigma, n))
    return fail(CSX_INVALID_INPUT, "fit requires finite q, intensity and positive sigma arrays");
  if (!fit_options_valid(o))
    return fail(CSX_INVALID_INPUT, "fit bounds, iteration limit, and q sampling stride are invalid");
  const std::vector<double> b(basis, basis + 6 * n);
  return fit_from_basis(b, q, intensity, sigma, n, initial, o, result, fitted);
}

csx_status csx_project_simplex(const double* values, size_t m, double* out) {
  if (m == 0 || !values || !out) return fail(CSX_INVALID_INPUT, "buffers are required");
  const std::vector<double> v = project_simplex(std::vector<double>(values, values + m));
  std::copy(v.begin(), v.end(), out);
  return CSX_OK;
}

csx_status csx_optimize_weights(const double* curves, size_t m, const double* intensity,
                                const double* sigma, size_t n, const csx_fit_options* o,
                                int uniform, double regularization, double* weights,
                                double* out3, double* fitted) {
  CSX_API_RANGE();
  if (!curves || !intensity || !sigma || !o || !weights || !out3 || m == 0 || n == 0)
    return fail(CSX_INVALID_INPUT, "buffers are required");
  std::vector<const double*> cv(m);
  for (size_t c = 0; c < m; ++c) cv[c] = curves + c * n;
  const WeightFit w = optimize_weights(cv, intensity, sigma, n, o, uniform != 0, regularization);
  std::copy(w.weights.begin(), w.weights.end(), weights);
  out3[0] = w.lin.scale;
  out3[1] = w.lin.bg;
  out3[2] = w.lin.chi;
  if (fitted) std::copy(w.fitted.begin(), w.fitted.end(), fitted);
  return CSX_OK;
}

csx_status csx_nnls_weights(const double* curves, size_t m, const double* intensity,
                            const double* sigma, size_t n, const csx_fit_options* o,
                            double* weights, double* out3, double* fitted) {
  CSX_API_RANGE();
  if (!curves || !intensity || !sigma || !o || !weights || !out3 || m == 0 || n == 0)
    return fail(CSX_INVALID_INPUT, "buffers are required");
  std::vector<const double*> cv(m);
  for (size_t c = 0; c < m; ++c) cv[c] = curves + c * n;
  WeightFit w;
  CSX_TRY(nnls_weights(cv, intensity, sigma, n, o, w));
  std::copy(w.weights.begin(), w.weights.end(), weights);
  out3[0] = w.lin.scale;
  out3[1] = w.lin.bg;
  out3[2] = w.lin.chi;
  if (fitted) std::copy(w.fitted.begin(), w.fitted.end(), fitted);
  return CSX_OK;
}

// EnsembleFitter::fit_core (ensemble.rs:234-299) from per-conformer bases:
// every (c1, c2) candidate of the 5 x 5 grid re-contracts the cached bases
// instead of recomputing profiles.  bases: m * 6 * n.
csx_status csx_ensemble_fit_core(const double* bases, size_t m, const double* q,
                                 const double* intensity, const double* sigma, size_t n,
                                 const csx_fit_options* o, int uniform, double regularization,
                                 int weight_solver, double* weights, csx_fit_params* params,
                                 double* chi2, double* fitted) {
  CSX_API_RANGE();
  if (!bases || !q || !intensity || !sigma || !o || !weights || !params || !chi2 || m == 0 || n == 0)
    return fail(CSX_INVALID_INPUT, "buffers are required");
  if (weight_solver != CSX_WEIGHTS_PROJECTED_GRADIENT && weight_solver != CSX_WEIGHTS_NNLS)
    return fail(CSX_INVALID_INPUT, "unknown weight solver");
  const bool use_nnls = weight_solver == CSX_WEIGHTS_NNLS && !uniform;
  if (use_nnls && regularization != 0.0)
    return fail(CSX_INVALID_INPUT, "NNLS weights take no regularization");
  std::vector<double> curves(m * n);
  std::vector<const double*> cv(m);
  double best = INFINITY;
  bool have = false;
  for (int i = 0; i < 5; ++i) {
    const double c1 = o->c1_bounds[0] + (o->c1_bounds[1] - o->c1_bounds[0]) * i / 4.0;
    for (int j = 0; j < 5; ++j) {
      const double c2 = o->c2_bounds[0] + (o->c2_bounds[1] - o->c2_bounds[0]) * j / 4.0;
      for (size_t c = 0; c < m; ++c) {
        combine_with_grad(bases + c * 6 * n, q, n, c1, c2, curves.data() + c * n, nullptr, nullptr);
        cv[c] = curves.data() + c * n;
      }
      WeightFit w;
      if (use_nnls)
        CSX_TRY(nnls_weights(cv, intensity, sigma, n, o, w));
      else
        w = optimize_weights(cv, intensity, sigma, n, o, uniform != 0, regularization);
      if (!have || w.lin.chi < best) {
        have = true;
        best = w.lin.chi;
        std::copy(w.weights.begin(), w.weights.end(), weights);
        *params = {w.lin.scale, w.lin.bg, c1, c2};
        if (fitted) std::copy(w.fitted.begin(), w.fitted.end(), fitted);
      }
    }
  }
  *chi2 = best;
  return CSX_OK;
}

csx_status csx_ensemble_fit_core_device(csx_ctx* ctx, const double* bases, size_t m,
                                        const double* q, const double* intensity,
                                        const double* sigma, size_t n, const csx_fit_options* o,
                                        int uniform, double regularization, int weight_solver,
                                        double* weights, csx_fit_params* params, double* chi2,
                                        double* fitted) {
  CSX_API_RANGE();
  if (!ctx) return fail(CSX_INVALID_INPUT, "context is null");
  if (weight_solver == CSX_WEIGHTS_NNLS || (weight_solver == CSX_WEIGHTS_PROJECTED_GRADIENT &&
                                            (!bases || !q || !intensity || !sigma || !o ||
                                             !weights || !params || !chi2 || m == 0 || n == 0)))
    return csx_ensemble_fit_core(bases, m, q, intensity, sigma, n, o, uniform, regularization,
                                 weight_solver, weights, params, chi2, fitted);
  if (weight_solver != CSX_WEIGHTS_PROJECTED_GRADIENT)
    return fail(CSX_INVALID_INPUT, "unknown weight solver");
  if (m > (1u << 20) || n > (1u << 20) || m * n > (size_t(1) << 28))
    return fail(CSX_INVALID_INPUT, "ensemble too large for the device fit");
  constexpr int kCand = 25;
  std::vector<double> g(kCand * n), c2v(kCand), c1v(kCand);
  for (int i = 0; i < 5; ++i) {
    const double c1 = o->c1_bounds[0] + (o->c1_bounds[1] - o->c1_bounds[0]) * i / 4.0;
    for (int j = 0; j < 5; ++j) {
      const double c2 = o->c2_bounds[0] + (o->c2_bounds[1] - o->c2_bounds[0]) * j / 4.0;
      const int k = 5 * i + j;
      c1v[k] = c1;
      c2v[k] = c2;
      for (size_t p = 0; p < n; ++p) g[k * n + p] = excluded_volume_scale(q[p], c1);
    }
  }
  int mpad = 2;
  while (static_cast<size_t>(mpad) < m) mpad <<= 1;
  Call call(ctx);
  CSX_TRY(call.begin());
  const hipStream_t st = call.stream();
  Scratch d_b(call), d_g(call), d_c2(call), d_y(call), d_s(call), d_curves(call), d_ws(call),
      d_sort(call), d_w(call), d_lin(call), d_fit(call);
  CSX_HIP(d_b.alloc(m * 6 * n * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_g.alloc(g.size() * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_c2.alloc(kCand * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_y.alloc(n * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_s.alloc(n * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_curves.alloc(kCand * m * n * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_ws.alloc(kCand * (3 * m + 2 * n) * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_sort.alloc(static_cast<size_t>(kCand) * mpad * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_w.alloc(kCand * m * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_lin.alloc(3 * kCand * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(d_fit.alloc(kCand * n * sizeof(double), st), "hipMallocAsync");
  CSX_HIP(hipMemcpyAsync(d_b.as<void>(), bases, m * 6 * n * sizeof(double), hipMemcpyHostToDevice,
                         st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_g.as<void>(), g.data(), g.size() * sizeof(double),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_c2.as<void>(), c2v.data(), kCand * sizeof(double),
                         hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_y.as<void>(), intensity, n * sizeof(double), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(d_s.as<void>(), sigma, n * sizeof(double), hipMemcpyHostToDevice, st),
          "hipMemcpyAsync");
  CSX_HIP(launch_ensemble_candidates(d_b.as<double>(), d_g.as<double>(), d_c2.as<double>(),
                                     static_cast<int>(m), static_cast<int>(n), kCand,
                                     d_y.as<double>(), d_s.as<double>(), uniform != 0,
                                     regularization, o->scale_mode, o->scale_value,
                                     o->background_mode, o->background_value,
                                     d_curves.as<double>(), d_ws.as<double>(), d_sort.as<double>(),
                                     mpad, d_w.as<double>(), d_lin.as<double>(),
                                     d_fit.as<double>(), st),
          "ensemble_fit_core");
  std::vector<double> w(kCand * m), lin(3 * kCand), fit(kCand * n);
  CSX_HIP(hipMemcpyAsync(w.data(), d_w.as<void>(), w.size() * sizeof(double),
                         hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(lin.data(), d_lin.as<void>(), lin.size() * sizeof(double),
                         hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  CSX_HIP(hipMemcpyAsync(fit.data(), d_fit.as<void>(), fit.size() * sizeof(double),
                         hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  for (Scratch* b : {&d_b, &d_g, &d_c2, &d_y, &d_s, &d_curves, &d_ws, &d_sort, &d_w, &d_lin, &d_fit})
    b->reset();
  CSX_TRY(call.finish());
  // the first candidate with the lowest chi^2, as the host loop picks it
  int best = 0;
  for (int k = 1; k < kCand; ++k)
    if (lin[3 * k + 2] < lin[3 * best + 2]) best = k;
  std::copy(w.begin() + best * m, w.begin() + (best + 1) * m, weights);
  *params = {lin[3 * best], lin[3 * best + 1], c1v[best], c2v[best]};
  *chi2 = lin[3 * best + 2];
  if (fitted) std::copy(fit.begin() + best * n, fit.begin() + (best + 1) * n, fitted);
  return CSX_OK;
}

}  // extern "C"

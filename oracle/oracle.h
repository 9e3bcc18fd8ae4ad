/*
 * oracle.h -- FP64 CPU restatement of crabSAXS's profile path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the correctness reference for
 * the HIP library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or, in bench.py,
 * as the timed CPU baseline).  The product (crabsaxs_amd/) never links,
 * imports or calls it.
 *
 * Every function follows the cited reference loop in its operation order
 * (compiled with -ffp-contract=off, no fast-math).  Parity status of each
 * function is recorded in the header comment of oracle.c.
 *
 * Atom arrays: xyz[3n] (A), element[n] (CSX_EL_*), group[n] (CSX_UA_* or
 * -1), occ[n].  Per-q loops run on `threads` POSIX threads (<= 0: all
 * online CPUs), split over q like the reference's rayon par_iter.
 */
#ifndef CSX_ORACLE_H
#define CSX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

double or_sinc(double x);
int or_united_atom_group(const char *residue, const char *atom_name);
double or_form_factor(int element, int group, double occ, int implicit_h, double q);
double or_excluded_volume(int element, int group, int implicit_h);
double or_water_form_factor(double q);
double or_excluded_volume_scale(double q, double c1);

/* debye_intensity_with_hydrogens, scattering.rs:86-110. out[nq]. */
void or_debye_intensity(const double *xyz, const int *element, const int *group,
                        const double *occ, size_t n, const double *q, size_t nq,
                        int implicit_h, int threads, double *out);

/* debye_solvent_basis, solvent.rs:504-584, with the hydration sites given
 * explicitly (nsites = 0 and model analytic -> analytic shell).  out[6*nq]. */
void or_debye_solvent_basis(const double *xyz, const int *element, const int *group,
                            const double *occ, size_t n, const double *q, size_t nq,
                            double c1, double c2, double thickness, double rho,
                            int grid_model, const double *sites, size_t nsites,
                            double voxel_volume, int implicit_h, int threads,
                            double *out);

/* Coordinate derivatives of the Debye intensity (vacuum, or Analytic solvent
 * with amplitude v - G(c1) e + c2 h): jac[k][i][c] = dI(q_k)/dx_{i,c} (may be
 * NULL) and grad[i][c] = sum_k weights[k] jac[k][i][c] (may be NULL). */
void or_debye_grad_x(const double *xyz, const int *element, const int *group,
                     const double *occ, size_t n, const double *q, size_t nq,
                     const double *weights, int solvent, double c1, double c2,
                     double thickness, double rho, int implicit_h, int threads,
                     double *grad, double *jac);

/* coordinate_features (metrics.rs:236-270): weighted *rg and *dmax, and the
 * unnormalised pair histogram of coordinate_pr (metrics.rs:475-548) in its
 * chunked order into hist[max(bins,16)] (untouched when dmax is 0). */
void or_coordinate_features(const double *xyz, const int *element, const int *group,
                            const double *occ, size_t n, size_t bins, double *rg,
                            double *dmax, double *hist);

/* combine_solvent_basis, spherical.rs:567-584. */
void or_combine_solvent_basis(const double *basis, const double *q, size_t nq,
                              double c1, double c2, double *out);

/* build_adaptive_hydration_grid (solvent.rs:121-223) / build_hydration_grid
 * (solvent.rs:259-401).  Return the site count (-1 on invalid input / too
 * large); write min(count, capacity) sites. */
long or_adaptive_grid(const double *xyz, size_t n, double spacing,
                      double clearance, double padding, double *sites,
                      size_t capacity, double *voxel_volume);
long or_exterior_grid(const double *xyz, const int *element, size_t n,
                      double spacing, double probe, double shell, double *sites,
                      size_t capacity, double *voxel_volume);

/* real_harmonics_all (spherical.rs:54-93) for one direction; out[(L+1)^2]. */
void or_real_harmonics(int l_max, double cos_theta, double phi, double *out);
/* spherical_bessel_all_into (spherical.rs:95-133): Miller from L+24. */
void or_bessel_reference(int l_max, double x, double *out);
/* j_0..j_L accurate for every x (Miller from max(L, x) + margin). */
void or_bessel_exact(int l_max, double x, double *out);

/* multipole_intensity (spherical.rs:266-312) when exact_bessel = 0;
 * with exact_bessel = 1 the same loop with or_bessel_exact. */
void or_multipole_intensity(const double *xyz, const int *element, const int *group,
                            const double *occ, size_t n, const double *q, size_t nq,
                            int l_max, int implicit_h, int exact_bessel, int threads,
                            double *out);
/* multipole_solvent_intensity (spherical.rs:408-470) and
 * multipole_solvent_basis (spherical.rs:474-565, out[6*nq]). */
void or_multipole_solvent(const double *xyz, const int *element, const int *group,
                          const double *occ, size_t n, const double *q, size_t nq,
                          int l_max, int implicit_h, double c1, double c2,
                          double thickness, double rho, int grid_model,
                          const double *sites, size_t nsites, double voxel_volume,
                          int basis, int exact_bessel, int threads, double *out);

/* solve_linear / solve_scale_background (fit.rs:137-244).
 * scale_mode: 0 Fit, 1 Fixed(scale_value); bg_mode: 0 Fit, 1 FixedZero,
 * 2 Fixed(bg_value).  out3 = (scale, background, chi2_reduced). */
void or_solve_scale_background(const double *theory, const double *intensity,
                               const double *sigma, size_t nq, int scale_mode,
                               double scale_value, int bg_mode, double bg_value,
                               double *out3);
/* cubic_hermite_interpolate (fit.rs:93-123). */
void or_cubic_hermite(const double *x, const double *y, size_t n,
                      const double *targets, size_t nt, double *out);
/* sampled_indices (fit.rs:84-91); returns count. */
size_t or_sampled_indices(size_t length, size_t stride, size_t *out);

/* optimize_weights (ensemble.rs:317-381) with solve_linear_fixed
 * (383-434) and project_simplex (449-465).  curves[m*nq]; out: weights[m],
 * out3 = (scale, background, chi2_reduced). */
void or_optimize_weights(const double *curves, size_t m, const double *intensity,
                         const double *sigma, size_t nq, int scale_mode,
                         double scale_value, int bg_mode, double bg_value,
                         int uniform, double regularization, double *weights,
                         double *out3);
void or_project_simplex(const double *values, size_t m, double *out);

#ifdef __cplusplus
}
#endif

#endif /* CSX_ORACLE_H */

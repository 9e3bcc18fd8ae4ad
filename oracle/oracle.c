/*
 * oracle.c -- FP64 CPU restatement of crabSAXS's profile path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, smoke() and
 * bench.py's cpu_baseline leg, never by the product.
 *
 * Parity pins (the tests that pin each function; tests/test_oracle.py):
 *   or_sinc                     closed form; mpmath                (pinned)
 *   or_form_factor / typing     f(0) electron counts of formfactor.rs:445-524;
 *                               table transcription vs the reference (pinned)
 *   or_debye_intensity          two-atom closed form (scattering.rs:250-257),
 *                               mpmath brute force golden vectors,
 *                               q = 0 -> (sum f)^2, invariance       (pinned)
 *   or_debye_solvent_basis      recombination identity (solvent.rs:651-665),
 *                               mpmath golden vectors of the six terms (pinned)
 *   or_adaptive_grid / exterior determinism, exterior bounds, shell rule,
 *                               fast-grid weight, translation invariance
 *                               (solvent.rs:689-771)                  (pinned)
 *   or_real_harmonics           scipy sph_harm in the real convention,
 *                               addition theorem                      (pinned)
 *   or_bessel_reference         scipy spherical_jn where the reference is
 *                               accurate (x < L+~15); bessel_zero_limit (pinned
 *                               there; its inaccurate regime is unpinned)
 *   or_bessel_exact             scipy spherical_jn and mpmath         (pinned)
 *   or_multipole_*              multipole vs Debye (spherical.rs:596-787,
 *                               scattering.rs:312-365)                (pinned)
 *   or_solve_scale_background   exact recovery (fit.rs:490-533)       (pinned)
 *   or_cubic_hermite            reproduces cubics exactly              (pinned)
 *   or_optimize_weights         simplex projection brute force, weight
 *                               recovery on noiseless mixtures        (pinned)
 *   or_debye_grad_x             central differences of or_debye_*     (pinned)
 *   or_coordinate_features      numpy brute force (exact bins, sums to 1e-12),
 *                               two-atom Dmax = 4 of metrics.rs:678-684,
 *                               sum of the histogram = sum of w_i w_j  (pinned)
 */
#include "oracle.h"

#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../include/crabsaxs_tables.h"

#define PI 3.14159265358979323846

/* ------------------------------------------------------------- helpers -- */

typedef void (*q_task)(void *ctx, size_t k);

struct pool_arg {
    q_task fn;
    void *ctx;
    size_t begin, end;
};

static void *pool_run(void *p) {
    struct pool_arg *a = (struct pool_arg *)p;
    for (size_t k = a->begin; k < a->end; ++k) a->fn(a->ctx, k);
    return NULL;
}

/* Run fn(ctx, k) for k in [0, nq) on `threads` threads (contiguous blocks). */
static void parallel_q(q_task fn, void *ctx, size_t nq, int threads) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if ((size_t)threads > nq) threads = (int)(nq ? nq : 1);
    if (threads == 1) {
        for (size_t k = 0; k < nq; ++k) fn(ctx, k);
        return;
    }
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    struct pool_arg *args = (struct pool_arg *)malloc(sizeof(struct pool_arg) * threads);
    for (int t = 0; t < threads; ++t) {
        args[t].fn = fn;
        args[t].ctx = ctx;
        args[t].begin = nq * t / threads;
        args[t].end = nq * (t + 1) / threads;
        pthread_create(&tid[t], NULL, pool_run, &args[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    free(tid);
    free(args);
}

static double dist3(const double *a, const double *b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* ------------------------------------------------------- form factors -- */

double or_sinc(double x) { return fabs(x) < 1e-8 ? 1.0 : sin(x) / x; }

static double gauss_sum(const double *a, const double *b, int terms, double c, double q) {
    const double s = q / (4.0 * PI);
    const double s2 = s * s;
    double sum = 0.0;
    for (int k = 0; k < terms; ++k) sum += a[k] * exp(-b[k] * s2);
    return sum + c;
}

static double cm(int el, double q) { return gauss_sum(CSX_CM_A[el], CSX_CM_B[el], 4, CSX_CM_C[el], q); }

static void upper_trim(const char *in, char *out, size_t cap) {
    while (*in && isspace((unsigned char)*in)) ++in;
    size_t len = strlen(in);
    while (len > 0 && isspace((unsigned char)in[len - 1])) --len;
    if (len >= cap) len = cap - 1;
    for (size_t i = 0; i < len; ++i) out[i] = (char)toupper((unsigned char)in[i]);
    out[len] = 0;
}

static int eq(const char *a, const char *b) { return strcmp(a, b) == 0; }
static int in2(const char *s, const char *a, const char *b) { return eq(s, a) || eq(s, b); }
static int in4(const char *s, const char *a, const char *b, const char *c, const char *d) {
    return eq(s, a) || eq(s, b) || eq(s, c) || eq(s, d);
}

/* formfactor.rs:314-323 */
static char nuc_base(const char *r) {
    if (in4(r, "A", "DA", "RA", "ADE")) return 'A';
    if (in4(r, "C", "DC", "RC", "CYT")) return 'C';
    if (in4(r, "G", "DG", "RG", "GUA")) return 'G';
    if (in4(r, "U", "DU", "RU", "URA")) return 'U';
    if (in4(r, "T", "DT", "RT", "THY")) return 'T';
    return 0;
}

/* formfactor.rs:325-355, same match order. */
static int nuc_group(const char *residue, char base, const char *name_in) {
    char name[64];
    strncpy(name, name_in, sizeof name - 1);
    name[sizeof name - 1] = 0;
    for (char *p = name; *p; ++p)
        if (*p == '*') *p = '\'';
    const int deoxy = residue[0] == 'D' || base == 'T';
    if (eq(name, "C1'") || eq(name, "C3'") || eq(name, "C4'")) return CSX_UA_CSP3H;
    if (eq(name, "C2'")) return deoxy ? CSX_UA_CSP3H2 : CSX_UA_CSP3H;
    if (eq(name, "C5'")) return CSX_UA_CSP3H2;
    if (eq(name, "O2'") && !deoxy) return CSX_UA_OALCOHOLH;
    if (in4(name, "OP1", "OP2", "O1P", "O2P")) return CSX_UA_ORESONANCE;
    if (eq(name, "C8") && (base == 'A' || base == 'G')) return CSX_UA_CAROMATICH;
    if (base == 'A' && eq(name, "C2")) return CSX_UA_CAROMATICH;
    if (base == 'A' && eq(name, "N6")) return CSX_UA_NH2;
    if (base == 'G' && eq(name, "N1")) return CSX_UA_NH;
    if (base == 'G' && eq(name, "N2")) return CSX_UA_NH2;
    if (base == 'C' && eq(name, "N4")) return CSX_UA_NH2;
    if ((base == 'C' && in2(name, "C5", "C6")) || (base == 'U' && in2(name, "C5", "C6")) ||
        (base == 'T' && eq(name, "C6")))
        return CSX_UA_CAROMATICH;
    if ((base == 'U' || base == 'T') && eq(name, "N3")) return CSX_UA_NH;
    if (base == 'T' && in2(name, "C7", "C5M")) return CSX_UA_CSP3H3;
    return -1;
}

/* formfactor.rs:357-417, same match order. */
int or_united_atom_group(const char *residue_in, const char *name_in) {
    char r[64], n[64];
    upper_trim(residue_in, r, sizeof r);
    upper_trim(name_in, n, sizeof n);
    const char base = nuc_base(r);
    if (base) return nuc_group(r, base, n);
    static const char *aa[] = {"ALA", "ARG", "ASN", "ASP", "ASH", "CYS", "CYX", "GLN", "GLU", "GLH",
                               "GLY", "HIS", "HID", "HIE", "HIP", "ILE", "LEU", "LYS", "MET", "MSE",
                               "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL", "PCA"};
    int is_aa = 0;
    for (size_t i = 0; i < sizeof aa / sizeof aa[0]; ++i) is_aa |= eq(r, aa[i]);
    if (!is_aa) return -1;
    if (eq(n, "N")) return in2(r, "PRO", "PCA") ? -1 : CSX_UA_NH;
    if (eq(n, "CA")) return eq(r, "GLY") ? CSX_UA_CSP3H2 : CSX_UA_CSP3H;
    if (eq(r, "ALA") && eq(n, "CB")) return CSX_UA_CSP3H3;
    if (eq(r, "ARG") && (eq(n, "CB") || eq(n, "CG") || eq(n, "CD"))) return CSX_UA_CSP3H2;
    if (eq(r, "ARG") && eq(n, "NE")) return CSX_UA_GUANIDINENH;
    if (eq(r, "ARG") && in2(n, "NH1", "NH2")) return CSX_UA_GUANIDINENH2;
    if ((eq(r, "ASN") && eq(n, "CB")) || (in2(r, "ASP", "ASH") && eq(n, "CB")) ||
        (in2(r, "CYS", "CYX") && eq(n, "CB")))
        return CSX_UA_CSP3H2;
    if (eq(r, "ASN") && eq(n, "ND2")) return CSX_UA_NH2;
    if (in2(r, "ASP", "GLU") && in4(n, "OD1", "OD2", "OE1", "OE2")) return CSX_UA_ORESONANCE;
    if ((eq(r, "ASH") && eq(n, "OD2")) || (eq(r, "GLH") && eq(n, "OE2"))) return CSX_UA_OACIDH;
    if (eq(r, "CYS") && eq(n, "SG")) return CSX_UA_SH;
    if ((eq(r, "GLN") && in2(n, "CB", "CG")) || (in2(r, "GLU", "GLH") && in2(n, "CB", "CG")))
        return CSX_UA_CSP3H2;
    if (eq(r, "GLN") && eq(n, "NE2")) return CSX_UA_NH2;
    if (in4(r, "HIS", "HID", "HIE", "HIP") && eq(n, "CB")) return CSX_UA_CSP3H2;
    if ((in2(r, "HIS", "HID") && eq(n, "ND1")) || (eq(r, "HIE") && eq(n, "NE2"))) return CSX_UA_NH;
    if (eq(r, "HIP") && in2(n, "ND1", "NE2")) return CSX_UA_NHPOSITIVE;
    if (in4(r, "HIS", "HID", "HIE", "HIP") && in2(n, "CD2", "CE1")) return CSX_UA_CAROMATICH;
    if ((eq(r, "ILE") && eq(n, "CB")) || (eq(r, "LEU") && eq(n, "CG")) ||
        (eq(r, "THR") && eq(n, "CB")) || (eq(r, "VAL") && eq(n, "CB")))
        return CSX_UA_CSP3H;
    if ((eq(r, "ILE") && eq(n, "CG1")) || (eq(r, "LEU") && eq(n, "CB"))) return CSX_UA_CSP3H2;
    if ((eq(r, "ILE") && in2(n, "CG2", "CD1")) || (eq(r, "LEU") && in2(n, "CD1", "CD2")))
        return CSX_UA_CSP3H3;
    if (eq(r, "LYS") && in4(n, "CB", "CG", "CD", "CE")) return CSX_UA_CSP3H2;
    if (eq(r, "LYS") && eq(n, "NZ")) return CSX_UA_NH3POSITIVE;
    if (in2(r, "MET", "MSE") && in2(n, "CB", "CG")) return CSX_UA_CSP3H2;
    if (in2(r, "MET", "MSE") && eq(n, "CE")) return CSX_UA_CSP3H3;
    if ((in2(r, "PHE", "TYR") && eq(n, "CB")) || (eq(r, "TRP") && eq(n, "CB"))) return CSX_UA_CSP3H2;
    if ((eq(r, "PHE") && (in4(n, "CD1", "CD2", "CE1", "CE2") || eq(n, "CZ"))) ||
        (eq(r, "TYR") && in4(n, "CD1", "CD2", "CE1", "CE2")) ||
        (eq(r, "TRP") && (in4(n, "CD1", "CE3", "CZ2", "CZ3") || eq(n, "CH2"))))
        return CSX_UA_CAROMATICH;
    if (eq(r, "PRO") && (eq(n, "CB") || eq(n, "CG") || eq(n, "CD"))) return CSX_UA_CSP3H2;
    if (eq(r, "SER") && eq(n, "CB")) return CSX_UA_CSP3H2;
    if ((eq(r, "SER") && eq(n, "OG")) || (eq(r, "THR") && eq(n, "OG1")) || (eq(r, "TYR") && eq(n, "OH")))
        return CSX_UA_OALCOHOLH;
    if ((eq(r, "THR") && eq(n, "CG2")) || (eq(r, "VAL") && in2(n, "CG1", "CG2"))) return CSX_UA_CSP3H3;
    if (eq(r, "TRP") && eq(n, "NE1")) return CSX_UA_NH;
    if (eq(r, "PCA") && in2(n, "CB", "CG")) return CSX_UA_CSP3H2;
    return -1;
}

/* formfactor.rs:425-439 */
double or_form_factor(int el, int group, double occ, int implicit_h, double q) {
    if (implicit_h && el == CSX_EL_H) return 0.0;
    double value;
    if (implicit_h && group >= 0)
        value = gauss_sum(CSX_UA_A[group], CSX_UA_B[group], 5, CSX_UA_C[group], q);
    else
        value = cm(el, q);
    return occ * value;
}

/* solvent.rs:428-438 */
double or_excluded_volume(int el, int group, int implicit_h) {
    if (implicit_h && el == CSX_EL_H) return 0.0;
    const double nh = group >= 0 ? CSX_UA_HYDROGENS[group] : 0.0;
    return CSX_EXCLUDED_VOLUME[el] + (implicit_h ? nh * CSX_EXCLUDED_VOLUME[CSX_EL_H] : 0.0);
}

/* solvent.rs:227-231 */
double or_water_form_factor(double q) {
    return cm(CSX_EL_O, q) + 2.0 * cm(CSX_EL_H, q) * or_sinc(q * 0.9572);
}

/* solvent.rs:441-443 */
static double fraser(double q, double v) { return v * exp(-pow(v, 2.0 / 3.0) * q * q / (4.0 * PI)); }

/* solvent.rs:448-452 */
double or_excluded_volume_scale(double q, double c1) {
    return c1 * c1 * c1 * exp(-(c1 * c1 - 1.0) * (1.62 * 1.62) * q * q / 4.0);
}

/* ------------------------------------------------------------- Debye -- */

struct debye_ctx {
    const double *xyz, *occ, *q;
    const int *el, *grp;
    size_t n;
    int implicit_h;
    double *out;
    double *f; /* per-thread scratch is allocated inside the task */
};

static void debye_task(void *p, size_t k) {
    struct debye_ctx *c = (struct debye_ctx *)p;
    const double q = c->q[k];
    double *f = (double *)malloc(sizeof(double) * (c->n ? c->n : 1));
    for (size_t i = 0; i < c->n; ++i) f[i] = or_form_factor(c->el[i], c->grp[i], c->occ[i], c->implicit_h, q);
    double intensity = 0.0;
    for (size_t i = 0; i < c->n; ++i) {
        intensity += f[i] * f[i];
        for (size_t j = i + 1; j < c->n; ++j) {
            const double r = dist3(c->xyz + 3 * i, c->xyz + 3 * j);
            intensity += 2.0 * f[i] * f[j] * or_sinc(q * r);
        }
    }
    free(f);
    c->out[k] = intensity > 0.0 ? intensity : 0.0;
}

void or_debye_intensity(const double *xyz, const int *element, const int *group, const double *occ,
                        size_t n, const double *q, size_t nq, int implicit_h, int threads,
                        double *out) {
    struct debye_ctx c = {xyz, occ, q, element, group, n, implicit_h, out, NULL};
    parallel_q(debye_task, &c, nq, threads);
}

struct basis_ctx {
    const double *xyz, *occ, *q, *sites;
    const int *el, *grp;
    size_t n, nsites, nq;
    double c1, c2, t, rho, vox;
    int grid, implicit_h;
    double *out;
};

/* component_amplitudes, solvent.rs:454-488 */
static void amplitudes(const struct basis_ctx *c, double q, double *v, double *e, double *h) {
    for (size_t i = 0; i < c->n; ++i) {
        v[i] = or_form_factor(c->el[i], c->grp[i], c->occ[i], c->implicit_h, q);
        const double volume = or_excluded_volume(c->el[i], c->grp[i], c->implicit_h);
        const double radius = cbrt(3.0 * volume / (4.0 * PI));
        e[i] = c->occ[i] * c->rho * fraser(q, volume);
        const double rs = radius + c->t / 2.0;
        const double shell_volume = 4.0 * PI * (rs * rs) * c->t;
        h[i] = c->occ[i] * c->rho * 0.1 * fraser(q, shell_volume);
    }
}

static void basis_task(void *p, size_t k) {
    const struct basis_ctx *c = (const struct basis_ctx *)p;
    const double q = c->q[k];
    const size_t n = c->n;
    double *v = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
    double *e = v + n, *h = v + 2 * n;
    amplitudes(c, q, v, e, h);
    double t[6] = {0, 0, 0, 0, 0, 0};
    const int nogrid = !c->grid;
    for (size_t i = 0; i < n; ++i) {
        t[0] += v[i] * v[i];
        t[1] += e[i] * e[i];
        t[3] += v[i] * e[i];
        if (nogrid) {
            t[2] += h[i] * h[i];
            t[4] += v[i] * h[i];
            t[5] += e[i] * h[i];
        }
        for (size_t j = i + 1; j < n; ++j) {
            const double s = or_sinc(q * dist3(c->xyz + 3 * i, c->xyz + 3 * j));
            t[0] += 2.0 * v[i] * v[j] * s;
            t[1] += 2.0 * e[i] * e[j] * s;
            t[3] += (v[i] * e[j] + v[j] * e[i]) * s;
            if (nogrid) {
                t[2] += 2.0 * h[i] * h[j] * s;
                t[4] += (v[i] * h[j] + v[j] * h[i]) * s;
                t[5] += (e[i] * h[j] + e[j] * h[i]) * s;
            }
        }
    }
    if (c->grid) {
        const double sa = c->rho / 10.0 * c->vox * or_water_form_factor(q);
        for (size_t i = 0; i < c->nsites; ++i) {
            const double *site = c->sites + 3 * i;
            t[2] += sa * sa;
            for (size_t o = i + 1; o < c->nsites; ++o)
                t[2] += 2.0 * sa * sa * or_sinc(q * dist3(site, c->sites + 3 * o));
            for (size_t a = 0; a < n; ++a) {
                const double s = or_sinc(q * dist3(c->xyz + 3 * a, site));
                t[4] += v[a] * sa * s;
                t[5] += e[a] * sa * s;
            }
        }
    }
    for (int m = 0; m < 6; ++m) c->out[m * c->nq + k] = t[m];
    free(v);
}

void or_debye_solvent_basis(const double *xyz, const int *element, const int *group, const double *occ,
                            size_t n, const double *q, size_t nq, double c1, double c2,
                            double thickness, double rho, int grid_model, const double *sites,
                            size_t nsites, double voxel_volume, int implicit_h, int threads,
                            double *out) {
    struct basis_ctx c = {xyz, occ, q, sites, element, group, n, nsites, nq, c1, c2,
                          thickness, rho, voxel_volume, grid_model != 0, implicit_h, out};
    parallel_q(basis_task, &c, nq, threads);
}

void or_combine_solvent_basis(const double *b, const double *q, size_t nq, double c1, double c2,
                              double *out) {
    for (size_t i = 0; i < nq; ++i) {
        const double g = or_excluded_volume_scale(q[i], c1);
        const double v = b[i] + g * g * b[nq + i] + c2 * c2 * b[2 * nq + i] - 2.0 * g * b[3 * nq + i] +
                         2.0 * c2 * b[4 * nq + i] - 2.0 * g * c2 * b[5 * nq + i];
        out[i] = v > 0.0 ? v : 0.0;
    }
}

/* ---------------------------------------------------- Debye gradient -- */

struct grad_ctx {
    const double *xyz, *occ, *q, *w;
    const int *el, *grp;
    size_t n;
    int implicit_h, solvent;
    double c1, c2, t, rho;
    double *per_q; /* nq * 3n */
};

/* d I(q)/dx of the Debye sum with amplitudes A_i = v - G e + c2 h (solvent)
 * or A_i = v: 2 sum_j A_i A_j q sinc'(q r_ij) (x_i - x_j) / r_ij, restated
 * from the derivative of scattering.rs:100-106 / solvent.rs:533-558. */
static void grad_task(void *p, size_t k) {
    const struct grad_ctx *c = (const struct grad_ctx *)p;
    const double q = c->q[k];
    const size_t n = c->n;
    double *g = c->per_q + k * 3 * n;
    for (size_t i = 0; i < 3 * n; ++i) g[i] = 0.0;
    if (q == 0.0) return;
    double *A = (double *)malloc(sizeof(double) * (n ? n : 1));
    struct basis_ctx b = {c->xyz, c->occ, c->q, NULL, c->el, c->grp, n, 0, 0, c->c1, c->c2,
                          c->t, c->rho, 0.0, 0, c->implicit_h, NULL};
    double *v = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
    amplitudes(&b, q, v, v + n, v + 2 * n);
    const double G = or_excluded_volume_scale(q, c->c1);
    for (size_t i = 0; i < n; ++i)
        A[i] = c->solvent ? v[i] - G * v[n + i] + c->c2 * v[2 * n + i] : v[i];
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            const double d[3] = {c->xyz[3 * i] - c->xyz[3 * j], c->xyz[3 * i + 1] - c->xyz[3 * j + 1],
                                 c->xyz[3 * i + 2] - c->xyz[3 * j + 2]};
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r < 1e-10) continue;
            const double u = q * r;
            const double dsinc = (u * cos(u) - sin(u)) / (u * u);
            const double f = 2.0 * A[i] * A[j] * q * dsinc / r;
            for (int a = 0; a < 3; ++a) {
                g[3 * i + a] += f * d[a];
                g[3 * j + a] -= f * d[a];
            }
        }
    free(A);
    free(v);
}

void or_debye_grad_x(const double *xyz, const int *element, const int *group, const double *occ,
                     size_t n, const double *q, size_t nq, const double *weights, int solvent,
                     double c1, double c2, double thickness, double rho, int implicit_h,
                     int threads, double *grad, double *jac) {
    double *per_q = jac ? jac : (double *)malloc(sizeof(double) * (nq ? nq : 1) * 3 * (n ? n : 1));
    struct grad_ctx c = {xyz, occ, q, weights, element, group, n, implicit_h, solvent,
                         c1, c2, thickness, rho, per_q};
    parallel_q(grad_task, &c, nq, threads);
    if (grad) {
        for (size_t i = 0; i < 3 * n; ++i) grad[i] = 0.0;
        for (size_t k = 0; k < nq; ++k)
            for (size_t i = 0; i < 3 * n; ++i) grad[i] += weights[k] * per_q[k * 3 * n + i];
    }
    if (!jac) free(per_q);
}

/* -------------------------------------------------------------- grids -- */

struct box {
    double lo[3];
    size_t nx, ny, nz;
};

static int make_box(const double *xyz, size_t n, double margin, double h, struct box *b) {
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (size_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            if (xyz[3 * i + a] < lo[a]) lo[a] = xyz[3 * i + a];
            if (xyz[3 * i + a] > hi[a]) hi[a] = xyz[3 * i + a];
        }
    size_t dim[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] -= margin;
        hi[a] += margin;
        const double cells = ceil((hi[a] - lo[a]) / h - 1e-10);
        if (!(cells < 1e9)) return -1;
        dim[a] = (size_t)cells + 1;
        b->lo[a] = lo[a];
    }
    b->nx = dim[0];
    b->ny = dim[1];
    b->nz = dim[2];
    return 0;
}

typedef void (*voxel_fn)(void *ctx, size_t cell);

/* solvent.rs:184-205 / 319-344 bounds and test. */
static void rasterize(const struct box *b, double h, const double *p, double radius, voxel_fn fn, void *ctx) {
    size_t lo[3], hi[3];
    const size_t dim[3] = {b->nx, b->ny, b->nz};
    for (int a = 0; a < 3; ++a) {
        const double lower = (p[a] - radius - b->lo[a]) / h;
        const double upper = (p[a] + radius - b->lo[a]) / h;
        double l = floor(lower - 1e-10);
        if (l < 0.0) l = 0.0;
        const double u = ceil(upper + 1e-10);
        lo[a] = (size_t)l;
        hi[a] = u <= 0.0 ? 0 : (size_t)u;
        if (hi[a] > dim[a] - 1) hi[a] = dim[a] - 1;
    }
    const double r2 = radius * radius;
    for (size_t z = lo[2]; z <= hi[2]; ++z)
        for (size_t y = lo[1]; y <= hi[1]; ++y)
            for (size_t x = lo[0]; x <= hi[0]; ++x) {
                const double dx = (b->lo[0] + (double)x * h) - p[0];
                const double dy = (b->lo[1] + (double)y * h) - p[1];
                const double dz = (b->lo[2] + (double)z * h) - p[2];
                if (dx * dx + dy * dy + dz * dz <= r2 + 1e-10) fn(ctx, (z * b->ny + y) * b->nx + x);
            }
}

static void set_one(void *ctx, size_t c) { ((unsigned char *)ctx)[c] = 1; }

static long emit(const struct box *b, double h, const unsigned char *keep, double *sites, size_t cap) {
    long count = 0;
    for (size_t z = 0; z < b->nz; ++z)
        for (size_t y = 0; y < b->ny; ++y)
            for (size_t x = 0; x < b->nx; ++x)
                if (keep[(z * b->ny + y) * b->nx + x]) {
                    if ((size_t)count < cap && sites) {
                        sites[3 * count] = b->lo[0] + (double)x * h;
                        sites[3 * count + 1] = b->lo[1] + (double)y * h;
                        sites[3 * count + 2] = b->lo[2] + (double)z * h;
                    }
                    ++count;
                }
    return count;
}

long or_adaptive_grid(const double *xyz, size_t n, double h, double clearance, double padding,
                      double *sites, size_t cap, double *volume) {
    *volume = h > 4.0 ? 2.0 * ((h / 2.0) * (h / 2.0) * (h / 2.0)) : h * h * h;
    if (n == 0) return 0;
    if (!(h > 0.0) || !isfinite(h) || !(clearance > 0.0) || !isfinite(clearance) || !(padding >= 0.0) ||
        !isfinite(padding))
        return -1;
    double c[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) c[a] += xyz[3 * i + a];
    for (int a = 0; a < 3; ++a) c[a] /= (double)n;
    double ss = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double dx = xyz[3 * i] - c[0], dy = xyz[3 * i + 1] - c[1], dz = xyz[3 * i + 2] - c[2];
        ss += dx * dx + dy * dy + dz * dz;
    }
    const double rg = sqrt(ss / (double)n);
    const double width = rg <= 15.0 ? 3.0 : rg >= 20.0 ? 5.0 : 3.0 + 0.4 * (rg - 15.0);
    struct box b;
    if (make_box(xyz, n, padding, h, &b)) return -1;
    const size_t count = b.nx * b.ny * b.nz;
    unsigned char *shell = (unsigned char *)calloc(count, 1);
    unsigned char *mol = (unsigned char *)calloc(count, 1);
    if (!shell || !mol) {
        free(shell);
        free(mol);
        return -1;
    }
    for (size_t i = 0; i < n; ++i) rasterize(&b, h, xyz + 3 * i, clearance + width, set_one, shell);
    for (size_t i = 0; i < n; ++i) rasterize(&b, h, xyz + 3 * i, clearance, set_one, mol);
    for (size_t k = 0; k < count; ++k) shell[k] = shell[k] && !mol[k];
    const long ns = emit(&b, h, shell, sites, cap);
    free(shell);
    free(mol);
    return ns;
}

struct ext_ctx {
    unsigned char *state;
    int value, only_exterior;
};

static void set_state(void *p, size_t c) {
    struct ext_ctx *e = (struct ext_ctx *)p;
    if (!e->only_exterior || e->state[c] == 2) e->state[c] = (unsigned char)e->value;
}

long or_exterior_grid(const double *xyz, const int *element, size_t n, double h, double probe, double shell,
                      double *sites, size_t cap, double *volume) {
    *volume = h * h * h;
    if (n == 0) return 0;
    if (!(h > 0.0) || !isfinite(h) || !(probe >= 0.0) || !isfinite(probe) || !(shell > 0.0) ||
        !isfinite(shell))
        return -1;
    double vmax = 0.0;
    for (size_t i = 0; i < n; ++i)
        if (CSX_VDW_RADIUS[element[i]] > vmax) vmax = CSX_VDW_RADIUS[element[i]];
    struct box b;
    if (make_box(xyz, n, vmax + probe + shell + h, h, &b)) return -1;
    const size_t count = b.nx * b.ny * b.nz;
    if (count > 250000000u) return -1;
    unsigned char *state = (unsigned char *)calloc(count, 1);
    size_t *queue = (size_t *)malloc(sizeof(size_t) * count);
    if (!state || !queue) {
        free(state);
        free(queue);
        return -1;
    }
    struct ext_ctx ctx = {state, 1, 0};
    for (size_t i = 0; i < n; ++i) rasterize(&b, h, xyz + 3 * i, CSX_VDW_RADIUS[element[i]] + probe, set_state, &ctx);
    size_t head = 0, tail = 0;
    const size_t nx = b.nx, ny = b.ny, nz = b.nz;
    for (size_t z = 0; z < nz; ++z)
        for (size_t y = 0; y < ny; ++y)
            for (size_t x = 0; x < nx; ++x) {
                if (x != 0 && x + 1 != nx && y != 0 && y + 1 != ny && z != 0 && z + 1 != nz) continue;
                const size_t c = (z * ny + y) * nx + x;
                if (state[c] == 0) {
                    state[c] = 2;
                    queue[tail++] = c;
                }
            }
    while (head < tail) {
        const size_t c = queue[head++];
        const size_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
        size_t nb[6];
        int k = 0;
        if (x > 0) nb[k++] = c - 1;
        if (x + 1 < nx) nb[k++] = c + 1;
        if (y > 0) nb[k++] = c - nx;
        if (y + 1 < ny) nb[k++] = c + nx;
        if (z > 0) nb[k++] = c - nx * ny;
        if (z + 1 < nz) nb[k++] = c + nx * ny;
        for (int m = 0; m < k; ++m)
            if (state[nb[m]] == 0) {
                state[nb[m]] = 2;
                queue[tail++] = nb[m];
            }
    }
    ctx.value = 3;
    ctx.only_exterior = 1;
    /* radius_extra = probe + shell is summed first (solvent.rs:386-391, 319) */
    const double extra = probe + shell;
    for (size_t i = 0; i < n; ++i)
        rasterize(&b, h, xyz + 3 * i, CSX_VDW_RADIUS[element[i]] + extra, set_state, &ctx);
    for (size_t c = 0; c < count; ++c) state[c] = state[c] == 3;
    const long ns = emit(&b, h, state, sites, cap);
    free(state);
    free(queue);
    return ns;
}

/* ---------------------------------------------------------- multipole -- */

/* harmonic_norms + factorial_ratio, spherical.rs:33-51 */
static void harmonic_norms(int L, double *norms) {
    for (int l = 0; l <= L; ++l)
        for (int m = 0; m <= l; ++m) {
            double ratio = 1.0;
            for (int k = l - m + 1; k <= l + m; ++k) ratio /= (double)k;
            norms[l * l + l + m] = sqrt(((double)(2 * l + 1) / (4.0 * PI)) * ratio);
        }
}

/* real_harmonics_all, spherical.rs:54-93 */
static void real_harmonics(int L, double x, double phi, const double *norms, double *out) {
    const int nc = (L + 1) * (L + 1);
    for (int k = 0; k < nc; ++k) out[k] = 0.0;
    double st = 1.0 - x * x;
    st = sqrt(st > 0.0 ? st : 0.0);
    const double sp = sin(phi), cp = cos(phi);
    double sm = 0.0, cmv = 1.0, pmm = 1.0;
    for (int m = 0; m <= L; ++m) {
        if (m > 0) {
            pmm *= -((double)(2 * m - 1)) * st;
            const double nc2 = cmv * cp - sm * sp;
            sm = sm * cp + cmv * sp;
            cmv = nc2;
        }
#define EMIT(l_, p_)                                                   \
    do {                                                               \
        const int pos_ = (l_) * (l_) + (l_) + m;                       \
        if (m == 0) {                                                  \
            out[pos_] = norms[pos_] * (p_);                            \
        } else {                                                       \
            const double val_ = sqrt(2.0) * norms[pos_] * (p_);        \
            out[pos_] = val_ * cmv;                                    \
            out[(l_) * (l_) + (l_) - m] = val_ * sm;                   \
        }                                                              \
    } while (0)
        EMIT(m, pmm);
        if (m == L) continue;
        const double pm1 = x * (double)(2 * m + 1) * pmm;
        EMIT(m + 1, pm1);
        double prev = pmm, cur = pm1;
        for (int l = m + 2; l <= L; ++l) {
            const double next = ((double)(2 * l - 1) * x * cur - (double)(l + m - 1) * prev) / (double)(l - m);
            EMIT(l, next);
            prev = cur;
            cur = next;
        }
#undef EMIT
    }
}

void or_real_harmonics(int L, double cos_theta, double phi, double *out) {
    double *norms = (double *)malloc(sizeof(double) * (L + 1) * (L + 1));
    harmonic_norms(L, norms);
    real_harmonics(L, cos_theta, phi, norms, out);
    free(norms);
}

/* Miller downward recurrence from `start`, normalised by the analytic j0 or
 * j1 -- spherical_bessel_all_into, spherical.rs:95-133, when start = L+24. */
static void bessel_miller(int L, int start, double x, double *j) {
    for (int l = 0; l <= L; ++l) j[l] = 0.0;
    if (fabs(x) < 1e-8) {
        j[0] = 1.0;
        return;
    }
    if (L == 0) {
        j[0] = sin(x) / x;
        return;
    }
    double *w = (double *)calloc((size_t)start + 2, sizeof(double));
    w[start] = 1.0;
    for (int l = start; l >= 1; --l) {
        w[l - 1] = (double)(2 * l + 1) * w[l] / x - w[l + 1];
        if (fabs(w[l - 1]) > 1e100)
            for (int k = l - 1; k <= start + 1; ++k) w[k] *= 1e-100;
    }
    const double j0 = sin(x) / x;
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    const double scale = fabs(j0) >= fabs(j1) ? j0 / w[0] : j1 / w[1];
    for (int l = 0; l <= L; ++l) j[l] = w[l] * scale;
    free(w);
}

void or_bessel_reference(int L, double x, double *out) { bessel_miller(L, L + 24, x, out); }

void or_bessel_exact(int L, double x, double *out) {
    const double top = (double)L > fabs(x) ? (double)L : fabs(x);
    const int start = (int)ceil(top + 30.0 + 6.0 * cbrt(top + 1.0) + 0.05 * top);
    bessel_miller(L, start, x, out);
}

struct geom {
    size_t n;
    int nc;
    double *radii, *harm, max_radius;
};

/* multipole_geometry_for_points, spherical.rs:165-199 */
static void geometry(const double *pts, size_t n, const double *centre, int L, struct geom *g) {
    g->n = n;
    g->nc = (L + 1) * (L + 1);
    g->radii = (double *)malloc(sizeof(double) * (n ? n : 1));
    g->harm = (double *)malloc(sizeof(double) * (n ? n : 1) * g->nc);
    double *norms = (double *)malloc(sizeof(double) * g->nc);
    harmonic_norms(L, norms);
    g->max_radius = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double px = pts[3 * i] - centre[0], py = pts[3 * i + 1] - centre[1], pz = pts[3 * i + 2] - centre[2];
        const double r = sqrt(px * px + py * py + pz * pz);
        g->radii[i] = r;
        double ct = 1.0, phi = 0.0;
        if (!(r < 1e-12)) {
            ct = pz / r;
            if (ct < -1.0) ct = -1.0;
            if (ct > 1.0) ct = 1.0;
            phi = atan2(py, px);
        }
        real_harmonics(L, ct, phi, norms, g->harm + i * g->nc);
        if (r > g->max_radius) g->max_radius = r;
    }
    free(norms);
}

static void free_geom(struct geom *g) {
    free(g->radii);
    free(g->harm);
}

/* adaptive_l_max, spherical.rs:201-203 */
static int adaptive_lmax(double q, double max_radius, int L) {
    const double o = ceil(q * max_radius);
    long v = (o > 1e6 ? 1000000 : (long)(o < 0 ? 0 : o)) + 8;
    const long lo = L < 8 ? L : 8;
    if (v < lo) v = lo;
    if (v > L) v = L;
    return (int)v;
}

struct mp_ctx {
    const double *xyz, *occ, *q, *sites;
    const int *el, *grp;
    size_t n, nsites, nq;
    int L, implicit_h, exact, mode; /* mode 0 vacuum, 1 solvent intensity, 2 basis */
    double c1, c2, t, rho, vox;
    int grid;
    struct geom atoms, sgeom;
    double *out;
};

/* solvent_components_for_atom, spherical.rs:205-231 */
static void mp_components(const struct mp_ctx *c, size_t i, double q, double *a) {
    const double volume = or_excluded_volume(c->el[i], c->grp[i], c->implicit_h);
    const double radius = cbrt(3.0 * volume / (4.0 * PI));
    const double occ = c->occ[i];
    const double excluded = occ * c->rho * volume * exp(-pow(volume, 2.0 / 3.0) * q * q / (4.0 * PI));
    const double rs = radius + c->t / 2.0;
    const double sv = 4.0 * PI * (rs * rs) * c->t;
    const double hydration = occ * c->rho * 0.1 * sv * exp(-q * q * rs * rs / 4.0);
    a[0] = or_form_factor(c->el[i], c->grp[i], occ, c->implicit_h, q);
    a[1] = excluded;
    a[2] = hydration;
}

static void mp_task(void *p, size_t k) {
    const struct mp_ctx *c = (const struct mp_ctx *)p;
    const double q = c->q[k];
    int ql = adaptive_lmax(q, c->atoms.max_radius, c->L);
    if (c->grid) {
        const int g2 = adaptive_lmax(q, c->sgeom.max_radius, c->L);
        if (g2 > ql) ql = g2;
    }
    const int nc = c->atoms.nc;
    const int nch = c->mode == 2 ? 3 : 1;
    double *coeff = (double *)calloc((size_t)nch * nc, sizeof(double));
    double *js = (double *)malloc(sizeof(double) * (ql + 1));
    for (size_t i = 0; i < c->n; ++i) {
        double a[3] = {0.0, 0.0, 0.0};
        double f = 0.0;
        if (c->mode == 0) {
            f = or_form_factor(c->el[i], c->grp[i], c->occ[i], c->implicit_h, q);
        } else {
            mp_components(c, i, q, a);
            if (c->mode == 1)
                f = a[0] - or_excluded_volume_scale(q, c->c1) * a[1] + (!c->grid ? c->c2 * a[2] : 0.0);
        }
        if (c->exact)
            or_bessel_exact(ql, q * c->atoms.radii[i], js);
        else
            bessel_miller(ql, ql + 24, q * c->atoms.radii[i], js);
        const double *hm = c->atoms.harm + i * nc;
        for (int l = 0; l <= ql; ++l)
            for (int m = l * l; m < (l + 1) * (l + 1); ++m) {
                if (c->mode == 2) {
                    const double cc = js[l] * hm[m];
                    coeff[m] += a[0] * cc;
                    coeff[nc + m] += a[1] * cc;
                    if (!c->grid) coeff[2 * nc + m] += a[2] * cc;
                } else {
                    coeff[m] += f * js[l] * hm[m];
                }
            }
    }
    if (c->grid && c->mode != 0) {
        const double sa = c->rho / 10.0 * c->vox * or_water_form_factor(q);
        for (size_t i = 0; i < c->nsites; ++i) {
            if (c->exact)
                or_bessel_exact(ql, q * c->sgeom.radii[i], js);
            else
                bessel_miller(ql, ql + 24, q * c->sgeom.radii[i], js);
            const double *hm = c->sgeom.harm + i * nc;
            for (int l = 0; l <= ql; ++l)
                for (int m = l * l; m < (l + 1) * (l + 1); ++m) {
                    if (c->mode == 2)
                        coeff[2 * nc + m] += sa * js[l] * hm[m];
                    else
                        coeff[m] += c->c2 * sa * js[l] * hm[m];
                }
        }
    }
    const double scale = 4.0 * PI;
    if (c->mode == 2) {
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int m = 0; m < nc; ++m) {
            s[0] += coeff[m] * coeff[m];
            s[1] += coeff[nc + m] * coeff[nc + m];
            s[2] += coeff[2 * nc + m] * coeff[2 * nc + m];
            s[3] += coeff[m] * coeff[nc + m];
            s[4] += coeff[m] * coeff[2 * nc + m];
            s[5] += coeff[nc + m] * coeff[2 * nc + m];
        }
        for (int t = 0; t < 6; ++t) c->out[t * c->nq + k] = scale * s[t];
    } else {
        double s = 0.0;
        for (int m = 0; m < nc; ++m) s += coeff[m] * coeff[m];
        const double v = scale * s;
        c->out[k] = v > 0.0 ? v : 0.0;
    }
    free(coeff);
    free(js);
}

static void centroid(const double *xyz, size_t n, double *c) {
    c[0] = c[1] = c[2] = 0.0;
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) c[a] += xyz[3 * i + a];
    for (int a = 0; a < 3; ++a) c[a] /= (double)n;
}

void or_multipole_intensity(const double *xyz, const int *element, const int *group, const double *occ,
                            size_t n, const double *q, size_t nq, int L, int implicit_h, int exact,
                            int threads, double *out) {
    if (n == 0) {
        for (size_t k = 0; k < nq; ++k) out[k] = 0.0;
        return;
    }
    struct mp_ctx c;
    memset(&c, 0, sizeof c);
    c.xyz = xyz;
    c.occ = occ;
    c.q = q;
    c.el = element;
    c.grp = group;
    c.n = n;
    c.nq = nq;
    c.L = L;
    c.implicit_h = implicit_h;
    c.exact = exact;
    c.mode = 0;
    c.out = out;
    double centre[3];
    centroid(xyz, n, centre);
    geometry(xyz, n, centre, L, &c.atoms);
    parallel_q(mp_task, &c, nq, threads);
    free_geom(&c.atoms);
}

void or_multipole_solvent(const double *xyz, const int *element, const int *group, const double *occ,
                          size_t n, const double *q, size_t nq, int L, int implicit_h, double c1,
                          double c2, double thickness, double rho, int grid_model, const double *sites,
                          size_t nsites, double voxel_volume, int basis, int exact, int threads,
                          double *out) {
    if (n == 0) {
        for (size_t k = 0; k < (basis ? 6 : 1) * nq; ++k) out[k] = 0.0;
        return;
    }
    struct mp_ctx c;
    memset(&c, 0, sizeof c);
    c.xyz = xyz;
    c.occ = occ;
    c.q = q;
    c.sites = sites;
    c.el = element;
    c.grp = group;
    c.n = n;
    c.nsites = nsites;
    c.nq = nq;
    c.L = L;
    c.implicit_h = implicit_h;
    c.exact = exact;
    c.mode = basis ? 2 : 1;
    c.c1 = c1;
    c.c2 = c2;
    c.t = thickness;
    c.rho = rho;
    c.vox = voxel_volume;
    c.grid = grid_model != 0;
    c.out = out;
    double centre[3];
    centroid(xyz, n, centre);
    geometry(xyz, n, centre, L, &c.atoms);
    if (c.grid) geometry(sites, nsites, centre, L, &c.sgeom);
    parallel_q(mp_task, &c, nq, threads);
    free_geom(&c.atoms);
    if (c.grid) free_geom(&c.sgeom);
}

/* ---------------------------------------------------------------- fit -- */

void or_solve_scale_background(const double *x, const double *y, const double *sg, size_t nq,
                               int scale_mode, double scale_value, int bg_mode, double bg_value,
                               double *out3) {
    double scale = 0.0, bg = 0.0;
    if (scale_mode == 0) {
        /* solve_linear, fit.rs:137-198 */
        double sw = 0, sx = 0, sy = 0, sxx = 0, sxy = 0;
        for (size_t i = 0; i < nq; ++i) {
            const double w = 1.0 / (sg[i] * sg[i]);
            sw += w;
            sx += w * x[i];
            sy += w * y[i];
            sxx += w * x[i] * x[i];
            sxy += w * x[i] * y[i];
        }
        if (bg_mode == 0) {
            const double det = sxx * sw - sx * sx;
            scale = fabs(det) > 1e-20 ? (sxy * sw - sx * sy) / det : 0.0;
            if (scale < 0.0) scale = 0.0;
            bg = sw > 0.0 ? (sy - scale * sx) / sw : 0.0;
        } else if (bg_mode == 1) {
            scale = sxx > 0.0 ? sxy / sxx : 0.0;
            if (scale < 0.0) scale = 0.0;
            bg = 0.0;
        } else {
            scale = sxx > 0.0 ? (sxy - bg_value * sx) / sxx : 0.0;
            if (scale < 0.0) scale = 0.0;
            bg = bg_value;
        }
    } else {
        /* solve_scale_background with a fixed scale, fit.rs:206-243 */
        scale = scale_value > 0.0 ? scale_value : 0.0;
        if (bg_mode == 0) {
            double wt = 0.0, w = 0.0;
            for (size_t i = 0; i < nq; ++i) {
                const double ww = 1.0 / (sg[i] * sg[i]);
                wt += ww * (y[i] - scale * x[i]);
                w += ww;
            }
            bg = w > 0.0 ? wt / w : 0.0;
        } else {
            bg = bg_mode == 1 ? 0.0 : bg_value;
        }
    }
    const double dof = nq > 1 ? (double)(nq - 1) : 1.0;
    double chi = 0.0;
    for (size_t i = 0; i < nq; ++i) {
        const double r = (scale * x[i] + bg - y[i]) / sg[i];
        chi += r * r;
    }
    out3[0] = scale;
    out3[1] = bg;
    out3[2] = chi / dof;
}

void or_cubic_hermite(const double *x, const double *y, size_t n, const double *t, size_t nt, double *out) {
    if (n == 1) {
        for (size_t k = 0; k < nt; ++k) out[k] = y[0];
        return;
    }
    double *sl = (double *)malloc(sizeof(double) * n);
    const size_t last = n - 1;
    sl[0] = (y[1] - y[0]) / (x[1] - x[0]);
    sl[last] = (y[last] - y[last - 1]) / (x[last] - x[last - 1]);
    for (size_t i = 1; i < last; ++i) sl[i] = (y[i + 1] - y[i - 1]) / (x[i + 1] - x[i - 1]);
    size_t seg = 0;
    for (size_t k = 0; k < nt; ++k) {
        while (seg + 1 < last && t[k] > x[seg + 1]) ++seg;
        const double h = x[seg + 1] - x[seg];
        double u = (t[k] - x[seg]) / h;
        if (u < 0.0) u = 0.0;
        if (u > 1.0) u = 1.0;
        const double u2 = u * u, u3 = u2 * u;
        out[k] = (2.0 * u3 - 3.0 * u2 + 1.0) * y[seg] + (u3 - 2.0 * u2 + u) * h * sl[seg] +
                 (-2.0 * u3 + 3.0 * u2) * y[seg + 1] + (u3 - u2) * h * sl[seg + 1];
    }
    free(sl);
}

size_t or_sampled_indices(size_t length, size_t stride, size_t *out) {
    if (length < 64) stride = 1;
    if (stride < 1) stride = 1;
    size_t c = 0;
    for (size_t i = 0; i < length; i += stride) out[c++] = i;
    if (c == 0 || out[c - 1] != length - 1) out[c++] = length - 1;
    return c;
}

/* --------------------------------------------------------- ensemble -- */

static int cmp_desc(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return (x < y) - (x > y);
}

void or_project_simplex(const double *v, size_t m, double *out) {
    double *s = (double *)malloc(sizeof(double) * (m ? m : 1));
    memcpy(s, v, sizeof(double) * m);
    qsort(s, m, sizeof(double), cmp_desc);
    double cum = 0.0;
    size_t rho = 0;
    for (size_t i = 0; i < m; ++i) {
        cum += s[i];
        if (s[i] + (1.0 - cum) / (double)(i + 1) > 0.0) rho = i + 1;
    }
    double top = 0.0;
    for (size_t i = 0; i < rho; ++i) top += s[i];
    const double theta = (top - 1.0) / (double)(rho > 0 ? rho : 1);
    for (size_t i = 0; i < m; ++i) {
        const double d = v[i] - theta;
        out[i] = d > 0.0 ? d : 0.0;
    }
    free(s);
}

/* solve_linear_fixed, ensemble.rs:383-434 */
static void linear_fixed(const double *x, const double *y, const double *sg, size_t nq, int scale_mode,
                         double scale_value, int bg_mode, double bg_value, double *out3) {
    double scale;
    if (scale_mode == 0) {
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < nq; ++i) {
            num += x[i] * y[i] / (sg[i] * sg[i]);
            den += x[i] * x[i] / (sg[i] * sg[i]);
        }
        scale = den > 0.0 ? num / den : 0.0;
    } else {
        scale = scale_value;
    }
    if (scale < 0.0) scale = 0.0;
    double bg;
    if (bg_mode == 0) {
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < nq; ++i) {
            num += (y[i] - scale * x[i]) / (sg[i] * sg[i]);
            den += 1.0 / (sg[i] * sg[i]);
        }
        bg = den > 0.0 ? num / den : 0.0;
    } else {
        bg = bg_mode == 1 ? 0.0 : bg_value;
    }
    double chi = 0.0;
    for (size_t i = 0; i < nq; ++i) {
        const double r = (scale * x[i] + bg - y[i]) / sg[i];
        chi += r * r;
    }
    out3[0] = scale;
    out3[1] = bg;
    out3[2] = chi / (nq > 1 ? (double)(nq - 1) : 1.0);
}

static void mix(const double *curves, const double *w, size_t m, size_t nq, double *out) {
    for (size_t p = 0; p < nq; ++p) {
        double s = 0.0;
        for (size_t c = 0; c < m; ++c) s += curves[c * nq + p] * w[c];
        out[p] = s;
    }
}

void or_optimize_weights(const double *curves, size_t m, const double *y, const double *sg, size_t nq,
                         int scale_mode, double scale_value, int bg_mode, double bg_value, int uniform,
                         double reg, double *w, double *out3) {
    for (size_t c = 0; c < m; ++c) w[c] = 1.0 / (double)m;
    double *mixture = (double *)malloc(sizeof(double) * nq);
    double *res = (double *)malloc(sizeof(double) * nq);
    double *g = (double *)malloc(sizeof(double) * (m ? m : 1));
    double *prop = (double *)malloc(sizeof(double) * (m ? m : 1));
    double lin[3];
    if (!uniform) {
        for (int it = 0; it < 250; ++it) {
            mix(curves, w, m, nq, mixture);
            linear_fixed(mixture, y, sg, nq, scale_mode, scale_value, bg_mode, bg_value, lin);
            for (size_t p = 0; p < nq; ++p) res[p] = (lin[0] * mixture[p] + lin[1] - y[p]) / sg[p];
            for (size_t c = 0; c < m; ++c) {
                double s = 0.0;
                for (size_t p = 0; p < nq; ++p) s += res[p] * lin[0] * curves[c * nq + p] / sg[p];
                g[c] = 2.0 * s + 2.0 * reg * (w[c] - 1.0 / (double)m);
            }
            double norm = 0.0;
            for (size_t c = 0; c < m; ++c) norm += g[c] * g[c];
            norm = sqrt(norm);
            if (norm < 1.0) norm = 1.0;
            for (size_t c = 0; c < m; ++c) g[c] = w[c] - 0.05 * g[c] / norm;
            or_project_simplex(g, m, prop);
            double delta = 0.0;
            for (size_t c = 0; c < m; ++c) delta += fabs(prop[c] - w[c]);
            memcpy(w, prop, sizeof(double) * m);
            if (delta < 1e-8) break;
        }
    }
    mix(curves, w, m, nq, mixture);
    linear_fixed(mixture, y, sg, nq, scale_mode, scale_value, bg_mode, bg_value, out3);
    free(mixture);
    free(res);
    free(g);
    free(prop);
}

/* ------------------------------------------------- coordinate features -- */

void or_coordinate_features(const double *xyz, const int *element, const int *group,
                            const double *occ, size_t n, size_t bins, double *rg,
                            double *dmax, double *hist) {
    /* metrics.rs:245-266: weights, centre, Rg */
    double *w = (double *)malloc(sizeof(double) * (n ? n : 1));
    double total = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double f = fabs(or_form_factor(element[i], group ? group[i] : -1, occ[i], 1, 0.0));
        w[i] = f > 1.0e-6 ? f : 1.0e-6;
        total += w[i];
    }
    double c[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) c[a] = c[a] + xyz[3 * i + a] * w[i];
    for (int a = 0; a < 3; ++a) c[a] = c[a] / total;
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double dx = xyz[3 * i] - c[0], dy = xyz[3 * i + 1] - c[1], dz = xyz[3 * i + 2] - c[2];
        acc += w[i] * (dx * dx + dy * dy + dz * dz);
    }
    *rg = sqrt(acc / total);
    /* metrics.rs:481-491: Dmax from squared distances */
    double dmax2 = 0.0;
    for (size_t l = 0; l < n; ++l)
        for (size_t r = l + 1; r < n; ++r) {
            double dx = xyz[3 * l] - xyz[3 * r], dy = xyz[3 * l + 1] - xyz[3 * r + 1],
                   dz = xyz[3 * l + 2] - xyz[3 * r + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > dmax2) dmax2 = d2;
        }
    if (!(dmax2 > 0.0)) {
        *dmax = 0.0;
        free(w);
        return;
    }
    *dmax = sqrt(dmax2);
    if (bins < 16) bins = 16;
    double step = *dmax / (double)bins;
    /* metrics.rs:508-531: 64-row chunks, each histogram in pair order, then
     * the chunks summed in order */
    double *part = (double *)calloc(bins, sizeof(double));
    for (size_t b = 0; b < bins; ++b) hist[b] = 0.0;
    for (size_t start = 0; start < n; start += 64) {
        size_t end = start + 64 < n ? start + 64 : n;
        for (size_t b = 0; b < bins; ++b) part[b] = 0.0;
        for (size_t l = start; l < end; ++l)
            for (size_t r = l + 1; r < n; ++r) {
                double dx = xyz[3 * l] - xyz[3 * r], dy = xyz[3 * l + 1] - xyz[3 * r + 1],
                       dz = xyz[3 * l + 2] - xyz[3 * r + 2];
                double d = sqrt(dx * dx + dy * dy + dz * dz);
                double f = floor(d / step);
                size_t b = f < (double)(bins - 1) ? (size_t)f : bins - 1;
                part[b] += w[l] * w[r];
            }
        for (size_t b = 0; b < bins; ++b) hist[b] += part[b];
    }
    free(part);
    free(w);
}

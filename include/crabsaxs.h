/*
 * crabsaxs.h -- C ABI of the MI355X-native crabSAXS profile engine.
 *
 * The reference (crabSAXS, Rust) has no FFI layer: its seam is the set of
 * `pub` functions of src/scattering.rs, src/solvent.rs and src/spherical.rs
 * that the fitting and analysis layers call.  Each compute entry point below
 * replaces one of them; the comment on each names the function it replaces
 * (file:line in the reference).  A Rust `-sys` crate binds these 1:1 with
 * `extern "C"` declarations that mirror this header.
 *
 * Conventions
 *   - Every function returns a csx_status; on failure a message is kept in a
 *     thread-local buffer, read with csx_last_error().
 *   - Inputs are borrowed (const pointers), outputs are caller-owned buffers
 *     of the documented size.  The library never hands out memory it
 *     allocated except through opaque handles released by *_destroy.
 *   - Status codes mirror `SaxsError` (src/error.rs:4-25).
 *   - Intensities are clamped at >= 0 exactly where the reference clamps;
 *     basis terms and pair tables are not clamped.
 *   - Functions are reentrant: handles may be shared between host threads
 *     (the reference calls its seam from rayon workers, ensemble.rs:244-256).
 *   - Coordinates are Angstrom, q is 1/Angstrom.
 *   - Every call that does work is wrapped in a roctx range named after the
 *     entry point (rocprofv3 --marker-trace shows them).
 *   - CSX_DEVICE_MEMORY_LIMIT=<bytes> in the environment (read once per
 *     process) caps the device memory the library holds: an allocation
 *     beyond it fails with CSX_OUT_OF_MEMORY, as a real one would.
 */
#ifndef CRABSAXS_H
#define CRABSAXS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CSX_OK = 0,
    CSX_INVALID_INPUT = 1,    /* SaxsError::InvalidInput */
    CSX_NOT_CONVERGED = 2,    /* SaxsError::FitDidNotConverge */
    CSX_NO_DEVICE = 3,        /* no usable HIP device */
    CSX_DEVICE_ERROR = 4,     /* a HIP runtime call failed */
    CSX_OUT_OF_MEMORY = 5,    /* device or host allocation failed */
    CSX_PARSE_ERROR = 6,      /* SaxsError::StructureParse / ExperimentalDataParse */
    CSX_UNSUPPORTED_ELEMENT = 7, /* SaxsError::UnsupportedElement */
    CSX_IO_ERROR = 8          /* SaxsError::Io (a file could not be read) */
} csx_status;

/* Pair-sum arithmetic. */
typedef enum {
    /* FP64 per-atom form factors, r_ij, sin and division per pair and q:
     * the reference loop of scattering.rs:100-106 run on the GPU. */
    CSX_FP64_PLAIN = 0,
    /* FP32 pair terms swept over all q for each r_ij, cascade-compensated
     * (FP32 inside a tile of <= 256 pairs, FP64 across tiles and blocks),
     * accumulated into per-type tables S_ab(q). */
    CSX_FP32_COMPENSATED = 1
} csx_precision;

/* HydrationModel, src/solvent.rs:17-25, plus one value with no reference
 * counterpart: the hydration sites attached to the structure with
 * csx_structure_set_hydration_sites, entering every path exactly like the
 * sites of a Grid model (uniform amplitude rho / 10 * V_voxel * f_w(q),
 * solvent.rs:560-573 and spherical.rs:449-451). */
typedef enum {
    CSX_HYDRATION_ANALYTIC = 0,
    CSX_HYDRATION_GRID = 1,
    CSX_HYDRATION_ADAPTIVE_GRID = 2,
    CSX_HYDRATION_EXPLICIT_SITES = 3
} csx_hydration_model;

/* ScatteringMethod, src/scattering.rs:16-21. */
typedef enum {
    CSX_METHOD_AUTO = 0,
    CSX_METHOD_DEBYE = 1,
    CSX_METHOD_MULTIPOLE = 2
} csx_method;

/* SolventParams, src/solvent.rs:27-50 (defaults: csx_solvent_default). */
typedef struct {
    double c1;
    double c2;
    double shell_thickness_angstrom;
    double bulk_density_e_per_a3;
    int32_t hydration_model;               /* csx_hydration_model */
    double hydration_grid_spacing_angstrom;
} csx_solvent_params;

/* ComputeOptions, src/scattering.rs:28-47.  l_max = 0 means "not given". */
typedef struct {
    int32_t method;           /* csx_method */
    int32_t implicit_hydrogen;/* HydrogenMode::Implicit = 1, Explicit = 0 */
    int32_t has_solvent;      /* Option<SolventParams>::is_some */
    csx_solvent_params solvent;
    int32_t l_max;            /* MultipoleOptions::l_max; 0 = None */
    int32_t precision;        /* csx_precision of the pair sums */
} csx_compute_options;

typedef struct csx_ctx csx_ctx;
typedef struct csx_structure csx_structure;

/* ---- library / context ------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char *csx_last_error(void);

/* SolventParams::default() (solvent.rs:39-50). */
void csx_solvent_default(csx_solvent_params *out);

/* ComputeOptions::default() (scattering.rs:37-47) with FP32_COMPENSATED. */
void csx_compute_options_default(csx_compute_options *out);

/* Number of visible HIP devices (0 when none; never fails). */
int csx_device_count(void);

/* Create a context bound to one HIP device.  CSX_NO_DEVICE when absent.
 * A context keeps, per stream, the device temporaries of its calls (grown on
 * demand, each at most 1 GiB; larger ones are freed after the call) until it
 * is destroyed.  Structures created on a context are not used after its
 * destroy; they may be destroyed before or after it. */
csx_status csx_ctx_create(int device, csx_ctx **out);
void csx_ctx_destroy(csx_ctx *ctx);

/* Kernel timing: when enabled, the library brackets every launch of the
 * named kernels with HIP events on the stream they run on and accumulates
 * their durations.  `name` is one of "debye_pairs_f32" (every FP32 pair
 * table launch, batches included), "debye_plain_f64", "multipole_coeff",
 * "debye_vjp_f32", "debye_jacobian_f32". */
csx_status csx_ctx_set_timing(csx_ctx *ctx, int enabled);
csx_status csx_ctx_kernel_time(csx_ctx *ctx, const char *name,
                               int64_t *launches, double *total_ms);
csx_status csx_ctx_reset_timing(csx_ctx *ctx);

/* Launch shape of the FP32 pair kernel (no reference counterpart; chosen by
 * the autotuner): q_tile q values per pair sweep, 0 (automatic: the tile
 * with the least modelled time for the call's q count and work items, the
 * work spread over the CUs plus twice the longest item, so small structures
 * take narrow tiles and more workgroups), a multiple of 8 up
 * to 64, 10, 20, 25, 50, 63, 100, 125 or 128 (sin mode 2 has 8, 10, 16, 20, 24, 25, 32,
 * 40, 48, 50, 56, 63 and 64;
 * sin modes 0 and 1 have 8, 16, 32 and 64; another tile falls back to the
 * automatic choice); sin_mode 0 (software
 * polynomial after exact reduction in revolutions), 1 (hardware V_SIN_F32
 * after the same reduction), 2 (V_SIN_F32 on the unreduced argument; used
 * only when |q| r_max / pi <= 250 and otherwise replaced by 1) or 3, the
 * default (on a uniform q grid whose step advances the phase of the largest
 * pair by at most 1/4 revolution, a sine recurrence along each tile's q
 * values, restarted from V_SIN_F32/V_COS_F32 per tile, for calls of at
 * least 2e7 pair.q or with a tuned recurrence tile; tiles 16, 25, 32, 50,
 * 64, 100, 125 and 128, held to those whose q span turns the largest pair's
 * phase by at most 4 revolutions; otherwise mode 2, with the same range
 * fallback) or 4 (radial interpolation of the pair distances: each pair
 * adds its 8 Chebyshev moments occ_i occ_j T_p(u) on its distance bin of
 * width h, the largest power of two with h max|q| <= 1.04, so that j_0(q r)
 * at the bin's 8 Chebyshev nodes interpolates it to within 1.0e-9, into
 * 64-bit fixed-point sums whose bits do not depend on the order; the
 * moments map to the nodes per bin, and the tables are the node sums of
 * sin(q rho) / rho in FP64; q_tile then sets the LDS accumulator copies
 * per workgroup: 0 (automatic), 8, 16 or 32, fewer where they do not fit;
 * where the distance bins do not fit a workgroup's LDS, mode 2.  About 5x the throughput of
 * mode 3 on 20,000 atoms x 500 q at the same accuracy; the no-allocation
 * entry point computes with mode 2 instead; the coordinate VJP follows it,
 * see csx_debye_grad_x);
 * j_chunk
 * atoms per work item: 0 (automatic: the largest of 512, 256, 128, 64 and
 * 32 giving at least one full 256 x j_chunk item per CU, so small structures
 * spread over the device), 32, 64, 128 or a multiple of 256.  The chunk orders
 * the FP64 sums: a structure's tables are bitwise reproducible for a given
 * chunk (and, with the automatic chunk, a given CU count), and under the
 * recurrence for a given q tile as well. */
csx_status csx_ctx_set_tuning(csx_ctx *ctx, int q_tile, int sin_mode,
                              int j_chunk);

/* j chunk and work-item count of the most recent FP32 pair-kernel launch on
 * the context (no reference counterpart; for reports and tests); 0s before
 * the first launch.  Any output may be NULL. */
csx_status csx_ctx_last_pair_items(csx_ctx *ctx, int *j_chunk, int *items);

/* Shape of the most recent FP32 pair-kernel launch on the context (after
 * the automatic tile choice and the sine-mode fallback); 0s before the
 * first launch.  Any output may be NULL. */
csx_status csx_ctx_last_launch(csx_ctx *ctx, int *q_tile, int *sin_mode,
                               int *persistent_workgroups);

/* Results cached per structure and solvent model (no reference
 * counterpart): the FP32 pair tables S_ab(q) of a q grid, and the per-type
 * multipole coefficients sum_{i of type t} occ_i j_l(q r_i) Y_lm,i of a q
 * grid and l_max (device memory).  Every Debye and multipole entry point
 * reuses them, so calls that change only (c1, c2), the shell thickness,
 * the bulk density or the hydrogen mode re-contract the cached tables
 * instead of re-running the pair or coefficient kernel, with bitwise the
 * results of a first call.  The key is the exact q grid (and l_max), the
 * sine mode and j chunk of the pair kernel; each structure keeps its most
 * recently used results up to `bytes` per solvent model (default 256 MiB;
 * 0 turns caching off and frees nothing already cached until it is
 * replaced).  Replaces the 25 x M compute_curve calls of
 * EnsembleFitter::fit_core (ensemble.rs:234-261) and the repeated calls of
 * SaxsScorer::score (analysis.rs:156-177) by re-contractions, with the
 * callers unchanged. */
csx_status csx_ctx_set_cache_limit(csx_ctx *ctx, size_t bytes);

/* Upper bound (bytes, >= 1 MiB; default 2 GiB) on the partial-sum
 * workspace of one coordinate-adjoint launch; larger problems run as
 * several launches.  No reference counterpart. */
csx_status csx_ctx_set_workspace_limit(csx_ctx *ctx, size_t bytes);

/* q values per pair sweep of the coordinate adjoint (csx_debye_grad_x and
 * the gradient fitter): 0 (automatic, the default: the tile with the least
 * modelled time for the call's q count and block pairs, narrow for small
 * structures), 16, 32, 48, 60 or 64 (48 and up need the unreduced sine,
 * |q| r_max / pi <= 250; otherwise the automatic choice among 16 and 32).
 * No reference counterpart. */
csx_status csx_ctx_set_vjp_tile(csx_ctx *ctx, int q_tile);

/* Column tiles (16 (l, m) columns each) per workgroup of the multipole
 * coefficient kernel: 32, 48, 64 or 128 (8, 12, 16 or 32 accumulator tiles
 * per wave), or 0 (the default: one group of 64 where 33 to 64 tiles cover
 * the expansion, groups of 32 otherwise).  The sums are the
 * same for every choice up to FP32 accumulation order.  No reference
 * counterpart. */
csx_status csx_ctx_set_multipole_group(csx_ctx *ctx, int tiles);

/* Radial interpolation of the multipole coefficients: 0 (the default:
 * automatic), 1 (every point enters the coefficient kernel), 2 (always
 * interpolate).  Interpolation groups each type's points into radial bins
 * of width h, the largest power of two with h * max|q| <= 1.04, and
 * replaces j_l(q r) inside a bin by its interpolant at the bin's 8
 * Chebyshev nodes (error at most 1.0e-9 of |j_l| <= 1 per point); the
 * kernel then contracts the nodes with the bins' harmonic moments.  The
 * automatic choice interpolates where the nodes number at most half the
 * points.  No reference counterpart (spherical.rs:266-565 sums every point). */
csx_status csx_ctx_set_multipole_radial(csx_ctx *ctx, int mode);

/* The radial bin width (0: per point) and the number of points (or nodes)
 * contracted by this context's last multipole coefficient computation.  No
 * reference counterpart. */
csx_status csx_ctx_last_multipole(csx_ctx *ctx, double *bin_width, int *points);

/* Workgroups sharing each 256-step sweep of the coordinate adjoint: 0 (the
 * default: automatic, 2, 4 or 8 where the block pairs and q chunks leave
 * most CUs idle, else 1), or 1, 2, 4 or 8.  Each part writes its own
 * partial sums, added in a fixed order.  No reference counterpart. */
csx_status csx_ctx_set_vjp_split(csx_ctx *ctx, int split);

/* Shape of the most recent coordinate-adjoint call on the context: q tile,
 * q chunks per workgroup split, step split, and the number of launches its
 * block pairs took under the workspace limit.  Any output may be NULL. */
csx_status csx_ctx_last_vjp_launch(csx_ctx *ctx, int *q_tile, int *q_split,
                                   int *step_split, int *launches);

/* Persistent launch of the FP32 pair kernel: a fixed grid of
 * workgroups_per_cu x CU workgroups drains a global task counter (0 = one
 * workgroup per task; -1, the default = automatic: 4 per CU when the call
 * has at least twice that many tasks, else one workgroup per task).  Every
 * launch form gives bitwise the same sums.  No reference counterpart. */
csx_status csx_ctx_set_persistence(csx_ctx *ctx, int workgroups_per_cu);

/* Compute-unit count and architecture name (e.g. "gfx950:...") of the
 * context's device; either output may be NULL. */
csx_status csx_ctx_device_info(csx_ctx *ctx, int *cu_count, char *arch,
                               size_t arch_len);

/* ---- host typing (src/formfactor.rs) ---------------------------------- */

/* Element::from_symbol (src/structure.rs:109-127): CSX_EL_* code. */
int csx_element_from_symbol(const char *symbol);

/* united_atom_group (src/formfactor.rs:357-417): CSX_UA_* or -1. */
int csx_united_atom_group(const char *residue_name, const char *atom_name);

/* form_factor_for_atom (src/formfactor.rs:425-439) for one atom. */
double csx_form_factor(int element, int group, double occupancy,
                       int implicit_hydrogen, double q);

/* excluded_volume_for_atom (src/solvent.rs:428-438). */
double csx_excluded_volume(int element, int group, int implicit_hydrogen);

/* water_form_factor (src/solvent.rs:227-231). */
double csx_water_form_factor(double q);

/* excluded_volume_scale (src/solvent.rs:448-452). */
double csx_excluded_volume_scale(double q, double radius_scale);

/* ---- structures --------------------------------------------------------- */

/* Upload one structure (SoA).  xyz: n*3 doubles; element: CSX_EL_*;
 * group: CSX_UA_* or -1 (from csx_united_atom_group); occupancy: n doubles.
 * The structure keeps an FP64 and a type-sorted FP32 copy in HBM.
 * Replaces the borrowed `&Structure` of every seam function. */
csx_status csx_structure_create(csx_ctx *ctx, size_t n, const double *xyz,
                                const uint8_t *element, const int8_t *group,
                                const double *occupancy,
                                csx_structure **out);
void csx_structure_destroy(csx_structure *s);
size_t csx_structure_atom_count(const csx_structure *s);

/* Attach hydration sites to a structure: `count` points (xyz: 3*count
 * doubles, A) of one voxel volume (A^3), used by every call whose solvent
 * model is CSX_HYDRATION_EXPLICIT_SITES.  They replace the sites a Grid or
 * AdaptiveGrid model builds from the structure (solvent.rs:121-401) with a
 * list the caller made (the proposed csx_sites_create of the seam: a site
 * list from the host grid builders or a generator).  A later call replaces
 * the list; calls running meanwhile finish with the previous one.  No
 * reference counterpart. */
csx_status csx_structure_set_hydration_sites(csx_structure *s,
                                             const double *xyz, size_t count,
                                             double voxel_volume);

/* ---- structure and curve ingest (src/structure.rs, src/io.rs) ---------- */

/* ResidueKind (structure.rs:24-33) and residue_kind (structure.rs:34-88). */
enum {
    CSX_RESIDUE_POLYMER = 0,
    CSX_RESIDUE_GLYCAN = 1,
    CSX_RESIDUE_WATER = 2,
    CSX_RESIDUE_OTHER = 3
};
int csx_residue_kind(const char *residue_name);

/* AltlocPolicy (structure.rs:13-21), OccupancyPolicy (structure.rs:166-175). */
enum { CSX_ALTLOC_HIGHEST = 0, CSX_ALTLOC_FIRST = 1, CSX_ALTLOC_ALL = 2 };
enum { CSX_OCCUPANCY_REPORTED = 0, CSX_OCCUPANCY_SELECTED_FULL = 1 };
/* Coordinate file formats of csx_parse_structure. */
enum { CSX_FORMAT_AUTO = 0, CSX_FORMAT_PDB = 1, CSX_FORMAT_MMCIF = 2 };

/* ParseOptions (structure.rs:177-202).  chains: NULL for every chain, else
 * a comma-separated list of chain ids (ParseOptions::chains). */
typedef struct {
    int include_hetatm;
    int include_waters;
    int include_glycans;
    size_t model_index;
    int allow_unknown_elements;
    int occupancy_policy;
    int altloc_policy;
    const char *chains;
} csx_parse_options;
void csx_parse_options_default(csx_parse_options *out);

/* One selected atom (Atom, structure.rs:150-160), with its united-atom group
 * (csx_united_atom_group of its residue and atom names). */
typedef struct {
    double x, y, z, occupancy;
    int element;
    int group;
    int is_hetatm;
    int64_t residue_id;
    char atom_name[16];
    char residue_name[16];
    char chain_id[16];
} csx_atom_record;

typedef struct csx_atoms csx_atoms;

/* Structure::from_pdb_file_with_options (structure.rs:209-326): PDB or
 * mmCIF (by extension .pdb/.ent or .cif/.mmcif, else by content), model
 * `model_index`, the reference's residue filters, element check, altloc and
 * occupancy policies and (chain, residue number, atom name) order.
 * Errors: CSX_IO_ERROR, CSX_PARSE_ERROR (malformed records, missing model),
 * CSX_UNSUPPORTED_ELEMENT.  Free with csx_atoms_destroy. */
csx_status csx_read_structure_file(const char *path,
                                   const csx_parse_options *options,
                                   csx_atoms **out);
/* The same from text in memory (format: CSX_FORMAT_*). */
csx_status csx_parse_structure(const char *text, size_t len, int format,
                               const csx_parse_options *options,
                               csx_atoms **out);
size_t csx_atoms_count(const csx_atoms *a);
/* Copy the csx_atoms_count(a) records. */
csx_status csx_atoms_get(const csx_atoms *a, csx_atom_record *out);
/* The SoA upload arrays of csx_structure_create (any pointer may be NULL). */
csx_status csx_atoms_arrays(const csx_atoms *a, double *xyz, uint8_t *element,
                            int8_t *group, double *occupancy);
void csx_atoms_destroy(csx_atoms *a);

/* ExperimentalCurve::from_dat_file (io.rs:19-112): comment and header
 * lines, ',' separators, the 1/nm header and q > 1 heuristics (q scaled by
 * 0.1), I = sigma = 0 sentinel rows, sigma 1 for two-column rows
 * (*has_errors 0), validation with the reference's messages.  Writes at most
 * `capacity` rows and the true count to *count (capacity 0 sizes). */
csx_status csx_read_dat_file(const char *path, double *q, double *intensity,
                             double *sigma, size_t capacity, size_t *count,
                             int *has_errors);
csx_status csx_parse_dat(const char *text, size_t len, double *q,
                         double *intensity, double *sigma, size_t capacity,
                         size_t *count, int *has_errors);

/* ---- hydration grids (src/solvent.rs:121-401) ---------------------------- */

/* build_adaptive_hydration_grid (solvent.rs:121-223) with
 * AdaptiveHydrationOptions{spacing, clearance, padding}, rasterised on the
 * structure's device.  Sites are bitwise the reference's, in its (z, y, x)
 * scan order.  Writes at most `capacity` sites (3 doubles each) and the true
 * count to *count; call with capacity 0 to size the buffer (the grid is kept
 * with the structure, so the second call does not rebuild it).  Grids of
 * more than 2^31 - 2^16 voxels are rejected. */
csx_status csx_build_adaptive_hydration_grid(
    csx_structure *s, double spacing, double clearance, double padding,
    double *sites, size_t capacity, size_t *count, double *voxel_volume);

/* build_hydration_grid (solvent.rs:259-401) with SolventGridOptions: the
 * probe-expanded rasterisation, the exterior flood fill and the shell, on
 * the structure's device; same output protocol. */
csx_status csx_build_hydration_grid(
    csx_structure *s, double spacing, double probe_radius,
    double shell_thickness, double *sites, size_t capacity, size_t *count,
    double *voxel_volume);

/* ---- Debye path ---------------------------------------------------------- */

/* debye_intensity_with_hydrogens (src/scattering.rs:86-110).  out: nq. */
csx_status csx_debye_intensity(csx_structure *s, const double *q, size_t nq,
                               int implicit_hydrogen, int precision,
                               double *out);

/* debye_solvent_basis (src/solvent.rs:504-584).  out: 6*nq, rows vv, ee,
 * hh, ve, vh, eh (term-major, like the reference's [Vec<f64>; 6]). */
csx_status csx_debye_solvent_basis(csx_structure *s, const double *q,
                                   size_t nq, const csx_solvent_params *p,
                                   int implicit_hydrogen, int precision,
                                   double *out);

/* solvent_corrected_intensity (src/solvent.rs:490-500).  out: nq. */
csx_status csx_solvent_corrected_intensity(csx_structure *s, const double *q,
                                           size_t nq,
                                           const csx_solvent_params *p,
                                           int implicit_hydrogen,
                                           int precision, double *out);

/* Per-type pair tables S_ab(q) (no reference counterpart; the reusable
 * half of debye_solvent_basis, solvent.rs:533-577).  Atoms are typed by
 * (element, united-atom group); with a Grid/AdaptiveGrid solvent model the
 * hydration sites of that model form one extra, last type (occupancy 1).
 * For the T types, out[b*nq + k] =
 *   sum_{i<j, type(i)=a, type(j)=c} occ_i occ_j sinc(q_k r_ij)
 * for bin b of (a <= c) in row-major upper-triangular order (pairs ordered
 * by type), computed with CSX_FP32_COMPENSATED arithmetic, and
 * self_terms[a] = sum_{type(i)=a} occ_i^2.  *n_types receives T.
 * Sizes: out >= nq*T*(T+1)/2, self_terms >= T.  p may be NULL (atoms only);
 * pass out = NULL to query T. */
csx_status csx_debye_pair_tables(csx_structure *s, const double *q,
                                 size_t nq, const csx_solvent_params *p,
                                 double *out, double *self_terms,
                                 size_t *n_types);

/* The pair tables of csx_debye_pair_tables (atoms only, p = NULL) written
 * to device memory on the caller's stream, with nothing allocated: no
 * hipMalloc, hipFree, pinned staging or stream creation.  No reference
 * counterpart; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  q: nq values in host memory (handed to the
 * device as kernel arguments); tables: device memory for T(T+1)/2 * nq
 * doubles, the bins of csx_debye_pair_tables (query T with it, out = NULL);
 * stream: a hipStream_t (NULL: the null stream).  The call is asynchronous:
 * the tables are complete once the stream has run its kernels.  Its
 * temporaries (task counter, q, per-work-item partials, raw tables) live in
 * one 64 MiB device array per GPU, held until the process ends, which limits
 * it to  256 + 4 nq' + 8 (items + bins) nq'  bytes (nq' the q count padded to
 * the q tile, items the pair kernel's work items: about n^2 / (2 * 256 *
 * j_chunk)); larger calls fail with CSX_INVALID_INPUT and a message.  Every
 * call on a GPU shares that array, so calls that overlap on one GPU (two
 * streams or two threads at once) race: serialise them.  The structure's
 * work items for the context's j chunk must exist: csx_structure_create
 * makes them for the j chunk of that moment.  Same sums, bit for bit, as
 * csx_debye_pair_tables with the same tuning. */
csx_status csx_debye_pair_tables_noalloc(csx_structure *s, const double *q,
                                         size_t nq, double *tables,
                                         void *stream);

/* Batched conformer profiles (no reference counterpart; replaces the per-
 * conformer compute_curve calls of EnsembleFitter::fit_core,
 * ensemble.rs:244-261): the work items of all m structures (any sizes, one
 * context) go through one launch of the FP32 pair kernel.
 * Vacuum intensities, out: m*nq (debye_intensity_with_hydrogens each). */
csx_status csx_batch_debye_intensity(csx_structure *const *structures,
                                     size_t m, const double *q, size_t nq,
                                     int implicit_hydrogen, double *out);
/* Six-term bases, out: m*6*nq (debye_solvent_basis each). */
csx_status csx_batch_debye_solvent_basis(csx_structure *const *structures,
                                         size_t m, const double *q, size_t nq,
                                         const csx_solvent_params *p,
                                         int implicit_hydrogen, double *out);

/* combine_solvent_basis (src/spherical.rs:567-584).  basis: 6*nq. */
csx_status csx_combine_solvent_basis(const double *basis, const double *q,
                                     size_t nq, double c1, double c2,
                                     double *out);

/* ---- multipole path ------------------------------------------------------ */
/* All three run the FP32 MFMA contraction with an exact j_l (FP32 upward
 * recurrence for l <= q r, FP64 Miller downward recurrence above) and the
 * per-q expansion order clamp(ceil(q R_max) + 8, min(8, l_max), l_max) of
 * spherical.rs:201-203.  0 <= l_max <= 64.  Empty structure -> zeros. */

/* multipole_intensity (src/spherical.rs:266-312).  out: nq. */
csx_status csx_multipole_intensity(csx_structure *s, const double *q,
                                   size_t nq, int l_max,
                                   int implicit_hydrogen, double *out);

/* multipole_solvent_intensity (src/spherical.rs:408-470).  out: nq. */
csx_status csx_multipole_solvent_intensity(csx_structure *s, const double *q,
                                           size_t nq, int l_max,
                                           int implicit_hydrogen,
                                           const csx_solvent_params *p,
                                           double *out);

/* multipole_solvent_basis (src/spherical.rs:474-565).  out: 6*nq rows vv,
 * ee, hh, ve, vh, eh. */
csx_status csx_multipole_solvent_basis(csx_structure *s, const double *q,
                                       size_t nq, int l_max,
                                       int implicit_hydrogen,
                                       const csx_solvent_params *p,
                                       double *out);

/* compute_curve (src/scattering.rs:112-212): validation, Auto method
 * choice, adaptive l_max and dispatch.  out: nq; *method_used receives
 * CSX_METHOD_DEBYE or CSX_METHOD_MULTIPOLE. */
csx_status csx_compute_curve(csx_structure *s, const double *q, size_t nq,
                             const csx_compute_options *options, double *out,
                             int *method_used);

/* ---- adjoint (no reference counterpart) -------------------------------- */
/* Derivatives of the Debye intensity of debye_intensity_with_hydrogens
 * (p == NULL) or of solvent_corrected_intensity with Analytic hydration
 * (solvent.rs:490-500, c1 and c2 fixed) with respect to the atom
 * coordinates, before the >= 0 clamp: FP32 pair terms, FP64 accumulation.
 * Grid hydration models are rejected (their sites are not differentiable).
 * With csx_ctx_set_tuning's sin_mode 4 the VJP takes each pair's q sum
 * sum_q B_ab(q) (qr cos qr - sin qr) from 8 Chebyshev coefficients of its
 * distance bin (width h, h max|q| <= 1.04; up to 2,048 bins, else the q
 * loop) instead of a q loop. */

/* grad[3*i + c] = sum_q weights[q] dI(q)/dx_{i,c}  (vector-Jacobian product;
 * with weights = d chi^2 / d I(q) this is d chi^2 / dx). */
csx_status csx_debye_grad_x(csx_structure *s, const double *q, size_t nq,
                            const double *weights,
                            const csx_solvent_params *p,
                            int implicit_hydrogen, double *grad);

/* jac[(k*n + i)*3 + c] = dI(q_k)/dx_{i,c}: nq * n * 3 doubles. */
csx_status csx_debye_jacobian(csx_structure *s, const double *q, size_t nq,
                              const csx_solvent_params *p,
                              int implicit_hydrogen, double *jac);

/* ---- coordinate features (src/metrics.rs) ------------------------------- */

/* coordinate_features (metrics.rs:236-270) with coordinate_pr
 * (metrics.rs:475-548) and normalize_distribution (metrics.rs:550-584):
 * weights |form_factor_for_atom(atom, implicit, 0)| floored at 1e-6;
 * *rg, *dmax the weighted radius of gyration and maximum pair distance;
 * the distribution on max(bins, 16) bins (at most 1536): *n_out points
 * (1 in the degenerate cases of the reference) written to r and p (each of
 * max(bins, 16) doubles), with its own rg / dmax in *pr_rg / *pr_dmax and
 * *pr_flags bit 0 / bit 1 set when they are present (Some).  Dmax and every
 * pair's bin are exactly the reference's (FP64 distances in its order); the
 * histogram's FP64 sums run in a different order. */
csx_status csx_coordinate_features(csx_structure *s, size_t bins, double *rg,
                                   double *dmax, double *r, double *p,
                                   size_t *n_out, double *pr_rg,
                                   double *pr_dmax, int *pr_flags);

/* ---- fitting (src/fit.rs) ---------------------------------------------- */

/* FitParams, fit.rs:11-17. */
typedef struct {
    double scale;
    double background;
    double c1;
    double c2;
} csx_fit_params;

/* BackgroundMode (fit.rs:41-47) and ScaleMode (fit.rs:29-34). */
typedef enum {
    CSX_BACKGROUND_FIT = 0,
    CSX_BACKGROUND_FIXED_ZERO = 1,
    CSX_BACKGROUND_FIXED = 2
} csx_background_mode;
typedef enum { CSX_SCALE_FIT = 0, CSX_SCALE_FIXED = 1 } csx_scale_mode;

/* Optimiser over (c1, c2): the reference's Nelder-Mead (fit.rs:385-418) or
 * a projected quasi-Newton on the analytic gradient d chi^2/d(c1, c2) (no
 * reference counterpart). */
typedef enum {
    CSX_OPTIMIZER_NELDER_MEAD = 0,
    CSX_OPTIMIZER_GRADIENT = 1
} csx_optimizer;

/* FitOptions, fit.rs:49-82 (defaults: csx_fit_options_default). */
typedef struct {
    int32_t method;              /* csx_method */
    int32_t implicit_hydrogen;
    double c1_bounds[2];
    double c2_bounds[2];
    int32_t max_iterations;
    int32_t q_sampling_stride;
    csx_solvent_params solvent;
    int32_t background_mode;     /* csx_background_mode */
    double background_value;     /* for CSX_BACKGROUND_FIXED */
    int32_t scale_mode;          /* csx_scale_mode */
    double scale_value;          /* for CSX_SCALE_FIXED */
    int32_t optimizer;           /* csx_optimizer */
    int32_t precision;           /* csx_precision of Debye pair sums */
} csx_fit_options;

/* FitResult, fit.rs:19-27 (fitted curve returned separately). */
typedef struct {
    csx_fit_params params;
    double chi2;
    int64_t iterations;
    int32_t converged;
    char termination[96];
} csx_fit_result;

void csx_fit_options_default(csx_fit_options *out);

/* solve_scale_background (fit.rs:200-244) incl. solve_linear (137-198):
 * out3 = (scale, background, reduced chi^2). */
csx_status csx_solve_scale_background(const double *theory,
                                      const double *intensity,
                                      const double *sigma, size_t n,
                                      int scale_mode, double scale_value,
                                      int background_mode,
                                      double background_value, double *out3);

/* cubic_hermite_interpolate (fit.rs:93-123). */
csx_status csx_cubic_hermite_interpolate(const double *x, const double *y,
                                         size_t n, const double *targets,
                                         size_t nt, double *out);

/* The stride-sampled, Hermite-interpolated six-term basis that
 * fit_to_experimental_with_options minimises over (fit.rs:329-384).
 * basis_out: 6*n on the experimental q grid. */
csx_status csx_fit_basis(csx_structure *s, const double *q, size_t n,
                         const csx_fit_options *options, double *basis_out);

/* fit_to_experimental_with_options (fit.rs:288-454).  fitted: n or NULL. */
csx_status csx_fit_to_experimental(csx_structure *s, const double *q,
                                   const double *intensity,
                                   const double *sigma, size_t n,
                                   const csx_fit_params *initial,
                                   const csx_fit_options *options,
                                   csx_fit_result *result, double *fitted);

/* The same fit from a precomputed basis (6*n); no reference counterpart. */
csx_status csx_fit_from_basis(const double *basis, const double *q,
                              const double *intensity, const double *sigma,
                              size_t n, const csx_fit_params *initial,
                              const csx_fit_options *options,
                              csx_fit_result *result, double *fitted);

/* chi^2 (reduced, N-1 dof as fit.rs:185-197) and its gradient
 * d chi^2 / d(scale, background, c1, c2) at the given parameters, from the
 * six-term basis (no reference counterpart).  grad4: 4 doubles. */
csx_status csx_chi2_grad(const double *basis, const double *q,
                         const double *intensity, const double *sigma,
                         size_t n, const csx_fit_params *params,
                         double *chi2, double *grad4);

/* ---- ensembles (src/ensemble.rs) ---------------------------------------- */

/* project_simplex (ensemble.rs:449-465). */
csx_status csx_project_simplex(const double *values, size_t m, double *out);

/* optimize_weights (ensemble.rs:317-381) on m curves (m*n, row-major):
 * weights[m]; out3 = (scale, background, reduced chi^2); fitted: n or NULL. */
csx_status csx_optimize_weights(const double *curves, size_t m,
                                const double *intensity, const double *sigma,
                                size_t n, const csx_fit_options *options,
                                int uniform, double regularization,
                                double *weights, double *out3,
                                double *fitted);

/* Weight solvers of an ensemble fit. */
typedef enum {
    CSX_WEIGHTS_PROJECTED_GRADIENT = 0, /* optimize_weights (ensemble.rs:317-381) */
    CSX_WEIGHTS_NNLS = 1 /* exact non-negative least squares (no reference
                          * counterpart): scale * weights = the non-negative
                          * coefficients a >= 0 minimising chi^2 with the
                          * background eliminated, by the Lawson-Hanson active
                          * set method; needs a fitted scale and zero
                          * regularization */
} csx_weight_solver;

/* Non-negative least-squares weights (CSX_WEIGHTS_NNLS) with the outputs of
 * csx_optimize_weights: weights[m] (sum 1, or uniform when every
 * coefficient is zero), out3 = (scale, background, reduced chi^2), fitted: n
 * or NULL.  No reference counterpart. */
csx_status csx_nnls_weights(const double *curves, size_t m,
                            const double *intensity, const double *sigma,
                            size_t n, const csx_fit_options *options,
                            double *weights, double *out3, double *fitted);

/* EnsembleFitter::fit_core (ensemble.rs:234-299): the 5 x 5 (c1, c2)
 * candidate grid re-contracts per-conformer six-term bases (m*6*n) instead
 * of recomputing profiles; best candidate's weights[m], params, reduced
 * chi^2 and fitted curve (n or NULL).  weight_solver: csx_weight_solver
 * (the reference uses CSX_WEIGHTS_PROJECTED_GRADIENT). */
csx_status csx_ensemble_fit_core(const double *bases, size_t m,
                                 const double *q, const double *intensity,
                                 const double *sigma, size_t n,
                                 const csx_fit_options *options, int uniform,
                                 double regularization, int weight_solver,
                                 double *weights, csx_fit_params *params,
                                 double *chi2, double *fitted);

/* The bootstrap resampling indices of EnsembleFitter::fit
 * (ensemble.rs:184-216): out[r*n + i] = the r-th replicate's draw for point
 * i from ChaCha8Rng::seed_from_u64(seed) and gen_range(0..n), in the
 * reference's order (replicate-major).  out: replicates * n. */
csx_status csx_bootstrap_indices(uint64_t seed, size_t n, size_t replicates,
                                 uint64_t *out);

/* The ChaCha block function under that stream (RFC 7539 section 2.3 with a
 * 64-bit block counter in state words 12-13 and a 64-bit stream in 14-15):
 * 16 output words for key[8], `rounds` (even) rounds. */
csx_status csx_chacha_block(const uint32_t *key, uint64_t counter,
                            uint64_t stream, int rounds, uint32_t *out);

/* The PCG32 generator under seed_from_u64 (O'Neill 2014): from `state`, n
 * steps state = state * 6364136223846793005 + increment, each giving the
 * XSH-RR output of the state before the step (output_new_state = 0, the
 * reference generator's convention) or after it (1, the seed expansion of
 * SeedableRng::seed_from_u64, which fills the ChaCha8 key with 8 outputs
 * for increment 11634580027462260723).  out: n words. */
csx_status csx_pcg32(uint64_t state, uint64_t increment, size_t n,
                     int output_new_state, uint32_t *out);

/* csx_ensemble_fit_core on the GPU of `ctx`: the 25 candidate profile sets
 * and their projected-gradient weight loops run on the device (one
 * workgroup per candidate), with every order-dependent sum in the host
 * order, so the outputs are bitwise those of csx_ensemble_fit_core.
 * weight_solver CSX_WEIGHTS_NNLS runs on the host (it is the same call).
 * No reference counterpart. */
csx_status csx_ensemble_fit_core_device(csx_ctx *ctx, const double *bases,
                                        size_t m, const double *q,
                                        const double *intensity,
                                        const double *sigma, size_t n,
                                        const csx_fit_options *options,
                                        int uniform, double regularization,
                                        int weight_solver, double *weights,
                                        csx_fit_params *params, double *chi2,
                                        double *fitted);

/* ---- synthetic workloads (no reference counterpart) -------------------- */

/* Seeded synthetic globular protein: self-avoiding C-alpha random walk with
 * 3.8 A steps (3.0 A exclusion) confined to an ellipsoid with semi-axes
 * (radius, radius, radius*elongation), 7 further atoms per residue at
 * 1.2-1.8 A from their C-alpha, elements drawn C 0.63 / N 0.17 / O 0.19 /
 * S 0.01, occupancy 1, centred on the centroid.  Deterministic for a seed
 * on every platform.  xyz: 3n, element: n. */
csx_status csx_generate_globule(size_t n, double radius, double elongation,
                                uint64_t seed, double *xyz,
                                uint8_t *element);

/* Multi-domain chain: `domains` globules (2-4) joined by flexible linkers of
 * C-alpha random walk; same atom placement and element rules. */
csx_status csx_generate_domain_chain(size_t n, int domains, uint64_t seed,
                                     double *xyz, uint8_t *element);

/* Hydration pseudo-atoms: grid points of spacing `spacing` whose nearest
 * atom lies in (inner, outer].  Same sizing protocol as the grid builders. */
csx_status csx_generate_shell_sites(const double *xyz, size_t n,
                                    double spacing, double inner,
                                    double outer, double *sites,
                                    size_t capacity, size_t *count);

#ifdef __cplusplus
}
#endif

#endif /* CRABSAXS_H */

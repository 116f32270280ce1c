/*
 * synthe_hip.h -- C ABI of libsynthe_hip.so, the MI355X (gfx950) hot path of
 * the synthe_py spectrum synthesiser: line opacity (TRANSP + ASYNTH), continuum
 * opacity, the JOSH formal solve with its scattering iteration, emergent flux.
 *
 * Conventions
 *  - Every array argument is a device pointer (hipMalloc'd or a torch CUDA
 *    tensor's data_ptr()) unless the comment says "host".  Shapes are row
 *    major; [D][W] means depth-major (depth d, wavelength w at d*W + w).
 *  - Outputs are written, never accumulated into; the caller allocates them.
 *  - Calls are asynchronous on `stream` (a hipStream_t; NULL = default
 *    stream) unless noted.  Calls on one context are serialised by its
 *    internal workspace; use one context per host thread for concurrency.
 *  - Return value: SYNTHE_OK (0) or a negative SYNTHE_E* code;
 *    synthe_last_error() gives the message of the calling thread's last
 *    failure.  No C++ exception crosses the ABI.
 *  - Pairs (line, depth) that fail a gate are skipped, never an error, as in
 *    the reference (physics/line_opacity.py:111-184).
 *
 * Reference interfaces each entry point replaces are cited as
 * path:line relative to the reference's synthe_py/ package.
 */
#ifndef SYNTHE_HIP_H
#define SYNTHE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SYNTHE_ABI_VERSION 2

#define SYNTHE_OK 0
#define SYNTHE_EINVAL (-1) /* bad size, pointer or argument */
#define SYNTHE_EHIP (-2)   /* a HIP runtime call failed */
#define SYNTHE_ENOMEM (-3) /* device allocation failed */
#define SYNTHE_ERANGE (-4) /* input exceeds a compiled limit (e.g. depth count) */

#define SYNTHE_MAX_DEPTH 128 /* largest depth count the RT kernels accept */
#define SYNTHE_NOALLOC_MAX_WL 65536 /* largest n_wl of synthe_lte_spectrum_noalloc */

typedef struct synthe_ctx synthe_ctx;
typedef void *synthe_stream; /* hipStream_t */

/* Per-line arrays, length n_lines, in catalog order (device pointers).
 * Mirrors the SoA built by physics/line_opacity.py:740-784. */
typedef struct {
  long n_lines;
  const double *wl;          /* index wavelength, nm */
  const double *gf;          /* 10**log gf */
  const double *cgf;         /* 0.026538/1.77245 * gf / nu */
  const double *gamma_rad;   /* normalised by 12.5664 nu */
  const double *gamma_stark;
  const double *gamma_vdw;
  const double *elo;         /* lower-level energy, cm^-1 */
  const int32_t *elem;       /* Z - 1, or -1 */
  const int32_t *ion;        /* ion stage, 1-based */
  const int32_t *ltype;      /* line type (0 = plain Voigt line) */
  const uint8_t *process;    /* 1 if the line takes the plain-line path */
  const int64_t *center_t;   /* centre index, engine/opacity.py:550-592 */
  const int64_t *center_w;   /* wing centre index, physics/line_opacity.py:1030-1059 */
} synthe_lines;

/* Per-depth arrays, length n_depths (device pointers). */
typedef struct {
  long n_depths;
  const double *temperature; /* K */
  const double *rhox;        /* column mass, g cm^-2 */
  const double *rho;         /* mass density, g cm^-3 */
  const double *xne;         /* electron density */
  const double *txnxn;       /* van der Waals perturber density */
  const double *hckt;        /* hc/kT, cm */
  const double *pop;         /* [D][n_ion][n_elem] N/U */
  const double *dop;         /* [D][n_ion][n_elem] Doppler width / c */
  long n_ion, n_elem;
} synthe_depths;

/* Message of the calling thread's last failure ("" if none). */
const char *synthe_last_error(void);
int synthe_abi_version(void);

/* Create a context on the current HIP device: builds the Voigt (H0/H1/H2,
 * physics/tables.py:214-246), FASTEX (physics/tables.py:13-59) and JOSH
 * (physics/josh_tables.py) tables and uploads them.  No reference
 * counterpart (the reference builds these lazily in Python). */
int synthe_create(synthe_ctx **out);
int synthe_destroy(synthe_ctx *ctx);

/* Host copies of the context's tables (host pointers; any may be NULL):
 * h0/h1/h2 [2001], extab/extabf [1001]. */
int synthe_copy_tables(const synthe_ctx *ctx, double *h0, double *h1, double *h2, double *extab,
                       double *extabf);

/* Profile precision of synthe_asynth's far-wing terms X / k^2 (no reference
 * counterpart; the reference is FP64 throughout).  SYNTHE_PREC_FP64 (the
 * default): each term within 32 ulp of the IEEE quotient.  SYNTHE_PREC_FP32:
 * each term to single precision (~2^-22 relative) through the hardware
 * reciprocal.  Both accumulate in FP64; populations, continuum, Voigt-zone
 * values and the RT stay FP64.  Applies to later calls on the context. */
#define SYNTHE_PREC_FP64 0
#define SYNTHE_PREC_FP32 1
int synthe_set_profile_precision(synthe_ctx *ctx, int precision);

/* Kernel timing (no reference counterpart; for benchmarks).  While enabled,
 * each synthe_asynth call records a pair of HIP events on its stream around
 * its wing-gather kernel (up to 256 launches; later ones are not recorded).
 * Enabling (or disabling) clears the record.  synthe_kernel_times waits for
 * the recorded events, writes min(count, cap) durations in milliseconds to
 * the host array ms (may be NULL) and the number recorded to *count. */
int synthe_set_kernel_timing(synthe_ctx *ctx, int enable);
int synthe_kernel_times(synthe_ctx *ctx, float *ms, int cap, int *count);

/* Geometric wavelength grid, computed on the HOST exactly as
 * engine/opacity.py:526-547 (_build_wavelength_grid).  Writes at most `cap`
 * values to the host array `out` (may be NULL) and returns the count
 * (negative on error). */
long synthe_wavelength_grid(double start, double end, double resolution, double *out, long cap);

/* physics/continuum.py:73-303 (build_depth_continuum, coefficient path).
 * wl [W]; wledge [E] (absolute value taken); half_edge, delta_edge [E-1] or
 * NULL to derive them; coeff_abs, coeff_scat [D][E-1][3] (log10 when
 * coeff_log10).  Outputs cont_abs, cont_scat [D][W]. */
int synthe_continuum(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl, long n_edges,
                     const double *wledge, const double *half_edge, const double *delta_edge,
                     const double *coeff_abs, const double *coeff_scat, int coeff_log10,
                     double *cont_abs, double *cont_scat, synthe_stream stream);

/* physics/line_opacity.py:620-860 (compute_transp) with the Boltzmann factor
 * FASTEX(E_low * HCKT) of physics/populations.py:172 fused in.
 * kapmin [D][W] = cont_abs + cont_scat.  Outputs transp [L][D] and
 * valid [L][D] (1/0). */
int synthe_transp(synthe_ctx *ctx, const synthe_lines *lines, const synthe_depths *depths,
                  long n_wl, const double *kapmin, double cutoff, double microturb_kms,
                  double *transp, uint8_t *valid, synthe_stream stream);

/* physics/line_opacity.py:946-1340 (compute_asynth_from_transp): centre adds,
 * wing accumulation for plain lines, times the stimulated-emission factor.
 * wl [W] is the grid; transp/valid [L][D] from synthe_transp;
 * center_order [L] = lines stably sorted by center_t and center_start [W+1]
 * its CSR offsets per pixel; wing_order [L] = lines stably sorted by
 * center_w (all host-computed index arrays, device copies).  Output
 * asynth [D][W].  stats (device, may be NULL) receives 5 doubles:
 * line x depth x lambda evaluations, wing pairs, sum of MAXSTEP, max MAXSTEP,
 * and the evaluations that took the per-pixel general (near-wing) path. */
int synthe_asynth(synthe_ctx *ctx, const synthe_lines *lines, const synthe_depths *depths,
                  long n_wl, const double *wl, const double *transp, const uint8_t *valid,
                  const double *kapmin, double cutoff, const int64_t *center_order,
                  const int64_t *center_start, const int64_t *wing_order, double *asynth,
                  double *stats, synthe_stream stream);

/* physics/line_opacity.py:620-860 and 946-1340 in one call, as
 * engine/opacity.py:2320-2363 runs compute_transp and then
 * compute_asynth_from_transp: TRANSP and its valid mask are computed in the
 * wing set-up's pass over the (line, depth) pairs and written to transp,
 * valid [L][D] (outputs here), then ASYNTH as synthe_asynth.  No
 * microturbulence (synthe_transp's microturb_kms = 0). */
int synthe_transp_asynth(synthe_ctx *ctx, const synthe_lines *lines, const synthe_depths *depths,
                         long n_wl, const double *wl, const double *kapmin, double cutoff,
                         const int64_t *center_order, const int64_t *center_start,
                         const int64_t *wing_order, double *transp, uint8_t *valid, double *asynth,
                         double *stats, synthe_stream stream);

/* Hydrogen lines (HPROF4 profiles), engine/opacity.py:1706-1884
 * (_compute_hydrogen_line_opacity) with :1547-1704 (_accumulate_hydrogen_profile)
 * and physics/profiles/hydrogen.py:1541-1709 (hydrogen_line_profile); the
 * result AHLINE joins ASYNTH as engine/opacity.py:2483-2487 adds it before the
 * FSCAT split.  All host arrays. */
typedef struct {
  int64_t n_lines;          /* the catalog's line type -1/-2, ion stage 1 lines, catalog order */
  const double *wavelength; /* [n] the record's wavelength, nm */
  const int64_t *center;    /* [n] nearest grid index of the index wavelength; -1 / n_wl off the grid */
  const double *gf;         /* [n] linear gf */
  const double *elo;        /* [n] lower level energy, cm^-1 */
  const int32_t *n_lower;   /* [n] NBLO as the catalog holds it */
  const int32_t *n_upper;   /* [n] NBUP */
} synthe_hlines;

typedef struct {
  int32_t n_depths;
  const double *temperature, *electron_density, *mass_density, *hckt; /* [D] */
  const double *pop_h1, *dop_h1;       /* [D] H I population and Doppler width (population_per_ion[:, 0, 0]) */
  const double *xnf_he1, *xnf_h2;      /* [D] or NULL; the hydrogen state exists when either is given */
  const double *xnfph;                 /* [D][2] or NULL (zeros) */
  const double *dopph;                 /* [D] or NULL: thermal H + turbulent widths */
  const double *turbulent_velocity;    /* [D] cm/s or NULL (for the dopph default) */
} synthe_hdepths;

/* Adds AHLINE (post-STIM, [D][W]) into asynth (device, [D][W]): per pixel
 * the lines' terms in catalog order, summed from zero, then added.  cont
 * (device [D][W]) is the total continuum (absorption + scattering) the
 * reference's cutoffs compare with; stim (device [D][W], or NULL for
 * 1 - exp(-h nu / k T) as engine/opacity.py:2215-2228 computes it) the
 * stimulated-emission factor; wl (device [W]) the grid.  Returns
 * SYNTHE_ERANGE (nothing added) if a line's n_upper + 2 exceeds the
 * radiative-damping table (96 levels), where the reference raises. */
int synthe_hydrogen_lines(synthe_ctx *ctx, const synthe_hlines *lines, const synthe_hdepths *depths,
                          double microturb_kms, double cutoff, long n_wl, const double *wl,
                          const double *cont, const double *stim, double *asynth, synthe_stream stream);

/* fort.19 lines, engine/opacity.py:1335-1546 (_add_fort19_asynth) with
 * :760-963 (_accumulate_metal_profile_kernel, _accumulate_merged_continuum,
 * _accumulate_autoionizing_profile) and :646-757 (WCON/WTAIL, merged
 * continuum limits): the records rgfall routes to fort.19
 * (io/lines/fort19.py:233-420) that are normal lines with NBLO/NBUP,
 * autoionising lines (Shore profiles) or merged-continuum ramps, in fort.19
 * order.  Kind 3 is the engine's helium pass instead (engine/opacity.py
 * :2534-2546, 2686-3066, 3551-3590, 3651-3659, with the TYPE < -2 branch of
 * _apply_fort19_profile, :998-1053): the catalog's He lines (TYPE -3/-4/-6),
 * each a Voigt profile walked to the cutoff, the sum per pixel times STIM;
 * a call takes either fort.19 records or helium lines, as the engine adds
 * them at different points.  All host arrays. */
typedef struct {
  int64_t n_lines;
  const int32_t *kind;      /* [n] 0 normal, 1 autoionising, 2 merged continuum, 3 helium line */
  const double *wavelength; /* [n] the record's vacuum wavelength, nm */
  const int64_t *center;    /* [n] nearest grid index of that wavelength, >= -1 and < n_wl */
  const double *gf;         /* [n] the record's float32 gf (CGF for a normal line), as double; CGF of a helium line */
  const double *gamma_rad, *gamma_stark, *gamma_vdw; /* [n] the record's float32 damping, as double */
  const int32_t *elem;      /* [n] Z - 1 of the matched catalog line */
  const int32_t *ion;       /* [n] its ion stage (1-based) */
  const int32_t *nelion;    /* [n] the record's NELION (a helium line's ion stage) */
  const int32_t *nlast;     /* [n] the record's TYPE (NLAST of a merged continuum; -4 is 3He) */
  const double *cont_val;   /* [n] CONTX(NCON, NELIONX), 0 where the record has no WCON */
  const double *cat_wavelength; /* [n] matched catalog line's wavelength (NULL if depths->doppler given) */
  const double *elo;        /* [n] its lower level energy, cm^-1 (NULL if depths->boltz given) */
} synthe_f19lines;

typedef struct {
  int32_t n_depths, n_ion, n_elem;
  const double *pop;        /* [D][n_ion][n_elem] population_per_ion */
  const double *mass_density, *electron_density, *txnxn, *temperature; /* [D] */
  const double *emerge, *emerge_h; /* [D] the engine's merge energies (engine/opacity.py:2369-2372) */
  const double *boltz;      /* [n][D] FASTEX(E_low HCKT) of each record's catalog line, or NULL: from elo, hckt */
  const double *hckt;       /* [D] (used when boltz is NULL) */
  const double *doppler;    /* [n][D] Doppler width, nm, of each record's catalog line, or NULL;
                               for a helium line DOPPLE (v/c) of its stage, required */
  const double *vscale;     /* [D] Doppler width / catalog wavelength (used when doppler is NULL) */
} synthe_f19depths;

/* Adds the fort.19 terms into asynth (device, [D][W]): per pixel the records'
 * terms in fort.19 order, summed from zero, times stim, added (the reference
 * sums a depth's records into a buffer and adds buffer * STIM).  cont
 * (device [D][W]) is the total continuum the cutoffs compare with, stim
 * (device [D][W], or NULL for 1 - exp(-h nu / k T)) and wl (device [W]) the
 * grid.  SYNTHE_ERANGE (nothing added) for a record whose element or stage
 * lies outside the population table, where the reference's lookup raises. */
int synthe_fort19_lines(synthe_ctx *ctx, const synthe_f19lines *lines, const synthe_f19depths *depths,
                        double cutoff, long n_wl, const double *wl, const double *cont, const double *stim,
                        double *asynth, synthe_stream stream);

/* engine/radiative.py:282-494 (solve_lte_spectrum) with
 * physics/josh_solver.py:367-767 (solve_josh_flux) per wavelength: the
 * continuum-only and the total JOSH solves.  All [D][W] except wl [W],
 * temperature and rhox [D]; line_source may be NULL (Planck).  Outputs
 * flux_total, flux_cont [W] per Hz; iters [W] (may be NULL) gets the summed
 * JOSH iteration count of the two solves, +1000 if the total solve took the
 * label-401 branch. */
int synthe_lte_spectrum(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl,
                        const double *temperature, const double *rhox, const double *cont_abs,
                        const double *cont_scat, const double *line_abs, const double *line_scat,
                        const double *line_source, double *flux_total, double *flux_cont,
                        int32_t *iters, synthe_stream stream);

/* No counterpart in the reference: synthe_lte_spectrum for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as synthe_lte_spectrum, but the RT workspace is a static device
 * array of the library (77 MiB per GPU, held until the process ends) instead
 * of the context's growable workspace: the call makes no hipMalloc, hipFree
 * or other allocation.  n_wl above SYNTHE_NOALLOC_MAX_WL returns
 * SYNTHE_ERANGE.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other contexts, other host threads) race. */
int synthe_lte_spectrum_noalloc(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl,
                                const double *temperature, const double *rhox,
                                const double *cont_abs, const double *cont_scat,
                                const double *line_abs, const double *line_scat,
                                const double *line_source, double *flux_total, double *flux_cont,
                                int32_t *iters, synthe_stream stream);

/* physics/josh_solver.py:367-767 (solve_josh_flux) over n_cols independent
 * columns: each input is [n][n_cols] (column c at j*n_cols + c); rhox [n]
 * shared.  Output flux [n_cols]; iters [n_cols] may be NULL. */
int synthe_josh_flux_batch(synthe_ctx *ctx, long n, long n_cols, const double *acont,
                           const double *scont, const double *aline, const double *sline,
                           const double *sigmac, const double *sigmal, const double *rhox,
                           double *flux, int32_t *iters, synthe_stream stream);

/* engine/opacity.py:2215-2228 (frequency quantities) and
 * physics/bfudge.py:72-168 in LTE (BFUDGE = 1): slinec [D][W] =
 * BNU * STIM / (1 - EHVKT).  Any output may be NULL. */
int synthe_source_function(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl,
                           const double *temperature, double *bnu, double *stim, double *ehvkt,
                           double *slinec, synthe_stream stream);

/* physics/bfudge.py:72-168 (compute_bfudge_and_slinec) with a departure
 * column: as synthe_source_function, with slinec = BNU * STIM / (BFUDGE(d) -
 * EHVKT) for bfudge [D] (device; NULL = 1, the LTE case).  BFUDGE itself is
 * BHYD^PH1 (BC1/BC2)^PC1 (BSI1/BSI2)^PSI1 per depth (host: physics.bfudge_column). */
int synthe_source_function_bfudge(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl,
                                  const double *temperature, const double *bfudge, double *bnu, double *stim,
                                  double *ehvkt, double *slinec, synthe_stream stream);

/* engine/opacity.py:3852-3861: flux[W] *= 2.99792458e17 / max(wl^2, 1e-40),
 * in place. */
int synthe_flux_to_lambda(synthe_ctx *ctx, long n_wl, const double *wl, double *flux,
                          synthe_stream stream);

/* Diagnostic (no reference counterpart): evaluates a far-wing quotient
 * routine q_fast = X / b and IEEE q_ieee = X / b on n device inputs, so tests
 * can compare them.  variant 0 = the correctly rounded reciprocal sequence;
 * 1, 2 = it with one or no Newton step before the remainder correction;
 * 3 = X * rcp(b); 4 = X * (rcp(b) after one Newton step), the wing gather's
 * edge tiles; 5, 6 = the shared-reciprocal form of its interior tiles with one
 * or no Newton steps, where b holds k (q = X / k^2) and each aligned group of
 * 8 inputs holds consecutive k of one sign, |k| >= 1. */
int synthe_diag_div_pos(long n, int variant, const double *x, const double *b, double *q_fast,
                        double *q_ieee, synthe_stream stream);

/* ------------------------------------------------------------------ batches
 * Many spectra (atmospheres of different depth counts, different windows)
 * in one replayable unit.  No reference counterpart: the reference runs one
 * spectrum per process (engine/opacity.py:2084 run_synthesis); a batch runs
 * that forward pass (continuum, TRANSP, ASYNTH, LTE RT, F_lambda) for every
 * instance, in LTE with no line scattering, with one persistent launch of the
 * wing gather over all instances' (depth, tile) items.
 *
 * Inputs are device arrays owned by the caller and must stay valid and
 * unchanged in shape for the batch's life (their contents may change between
 * runs).  Outputs flux and cont [n_wl] receive F_lambda of the total and the
 * continuum-only solve.  The library allocates one arena for everything
 * else at create time; synthe_batch_run allocates nothing. */
typedef struct {
  synthe_lines lines;      /* this instance's line SoA (its own centre indices) */
  synthe_depths depths;    /* n_depths <= SYNTHE_MAX_DEPTH */
  long n_wl;
  const double *wl;        /* [n_wl] grid */
  long n_edges;            /* continuum coefficients as synthe_continuum */
  const double *wledge, *half_edge, *delta_edge, *coeff_abs, *coeff_scat;
  int coeff_log10;
  const int64_t *center_order, *center_start, *wing_order; /* as synthe_asynth */
  double cutoff;
  double *flux, *cont;     /* outputs [n_wl], F_lambda */
  int32_t *iters;          /* output [n_wl] or NULL: as synthe_lte_spectrum's iters
                            * (+1000 where the total solve took the label-401 branch) */
} synthe_instance;

typedef struct synthe_batch synthe_batch;

#define SYNTHE_BATCH_AUTOTUNE 1 /* time the gather's tile / grid candidates on this batch, keep the fastest */
#define SYNTHE_BATCH_GRAPH 2    /* capture the whole batch into a hipGraph replayed by synthe_batch_run */

/* Plan a batch of n instances (host array).  Uses the context's profile
 * precision at create time.  Synchronous. */
int synthe_batch_create(synthe_ctx *ctx, int n, const synthe_instance *inst, int flags, synthe_batch **out);
/* Run the batch on stream (graph replay when created with SYNTHE_BATCH_GRAPH). */
int synthe_batch_run(synthe_batch *b, synthe_stream stream);
/* The gather's pixels per lane (tile = 64 x that, as synthe_asynth's) and
 * chosen grid (persistent, in workgroups of 64; 0 = one workgroup per tile and
 * depth), the autotuner's candidate times (ms: persistent grids of 2, 4, 8
 * waves per SIMD, then one workgroup per tile; up to tune_cap) and the arena
 * size.  Any pointer may be NULL. */
int synthe_batch_info(const synthe_batch *b, int *pixels_per_lane, int *gather_blocks, double *tune_ms,
                      int tune_cap, int *n_tuned, double *arena_bytes);
/* Device arrays of instance i of a batch after synthe_batch_run, for the
 * adjoint of a batched forward pass (synthe_lte_vjp and synthe_asynth_vjp take
 * them per instance): the continuum [D][W] (absorption, scattering, their sum
 * KAPMIN), the line source, ASYNTH (post-STIM, the line absorption of the
 * LTE solve), the all-zero line scattering, and TRANSP / valid [L][D].  Valid
 * until the next synthe_batch_run or synthe_batch_destroy.  No counterpart in
 * the reference (which has no batches). */
typedef struct {
  double *cont_abs, *cont_scat, *kapmin, *slinec, *asynth, *line_scat, *transp;
  uint8_t *valid;
} synthe_batch_view;

int synthe_batch_arrays(const synthe_batch *b, int i, synthe_batch_view *out);

int synthe_batch_destroy(synthe_batch *b);

/* ------------------------------------------------------------------ adjoint
 * Reverse-mode derivatives of the forward pass for the autograd wrapper (no
 * reference counterpart: the reference has no derivatives).  Branches -- the
 * KAPMIN gates, MAXSTEP, N10DOP, table indices, label 401 -- are held at their
 * forward values.
 *
 * synthe_asynth_vjp: g_asynth [D][W] = dL/d(asynth) of a synthe_asynth call
 * with the same inputs (ltype-0 lines).  Writes per (line, depth) pair [L][D]:
 * s_kappa0 = dL/d(kappa0), s_adamp = dL/d(a) and s_dopple = dL/d(Doppler
 * width / c) through the profile shape (kappa0 = cgf N/U FASTEX / (rho dop),
 * a = gamma / dop; the caller chains those) and, if kappa0 is not NULL,
 * kappa0 [L][D] itself (0 for pairs TRANSP rejected).  Recomputes the wing set-up and
 * every profile value; accumulates with floating-point atomics, so the last
 * bits can vary run to run.
 *
 * synthe_lte_vjp: v_flux [W] = dL/d(flux_total) (per Hz) of a
 * synthe_lte_spectrum call with the same inputs; writes dL/d(cont_abs,
 * cont_scat, line_abs) [D][W], dL/d(B) [D][W] for the Planck function at T(d)
 * (continuum source, and line source when line_source is NULL) and
 * dL/d(line_source) [D][W] (outputs may be NULL).  The scattering fixed point
 * and the label-401 iteration are differentiated implicitly at FP64
 * convergence, O(depths) per wavelength.  At most 96 depths; uses the
 * context's workspace, about 13 KB per wavelength (the per-wavelength state
 * between its kernel stages; 0.19 GB for 14,637 wavelengths). */
int synthe_asynth_vjp(synthe_ctx *ctx, const synthe_lines *lines, const synthe_depths *depths, long n_wl,
                      const double *wl, const double *transp, const uint8_t *valid, const double *kapmin,
                      double cutoff, const int64_t *center_order, const int64_t *center_start,
                      const int64_t *wing_order, const double *g_asynth, double *s_kappa0, double *s_adamp,
                      double *s_dopple, double *kappa0, synthe_stream stream);
int synthe_lte_vjp(synthe_ctx *ctx, long n_depths, long n_wl, const double *wl, const double *temperature,
                   const double *rhox, const double *cont_abs, const double *cont_scat, const double *line_abs,
                   const double *line_scat, const double *line_source, const double *v_flux, double *g_cont_abs,
                   double *g_cont_scat, double *g_line_abs, double *g_planck, double *g_line_source,
                   synthe_stream stream);

/* synthe_lte_vjp of n spectra in one launch per stage: each stage's lanes
 * run over every spectrum's wavelengths (a global index; each lane finds its
 * spectrum), so many small spectra fill the GPU together; per spectrum the
 * results are those of synthe_lte_vjp.  All arrays device pointers, as
 * synthe_lte_vjp takes them.  No counterpart in the reference (it has no
 * gradients or batches). */
typedef struct {
  long n_depths, n_wl;
  const double *wl, *temperature, *rhox, *cont_abs, *cont_scat, *line_abs, *line_scat, *line_source, *v_flux;
  double *g_cont_abs, *g_cont_scat, *g_line_abs, *g_planck, *g_line_source;
} synthe_vjp_instance;

int synthe_lte_vjp_batch(synthe_ctx *ctx, int n, const synthe_vjp_instance *inst, synthe_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* SYNTHE_HIP_H */

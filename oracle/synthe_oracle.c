/*
 * synthe_oracle.c -- CPU oracle for the synthe_py hot path (TEST INFRASTRUCTURE).
 *
 * A plain-C FP64 restatement of the reference algorithm, function by function,
 * each citing the reference file:line it follows (paths relative to the
 * reference's synthe_py/ package).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline; the product path never calls it.
 *
 * Numerics: compiled with -ffp-contract=off (no FMA contraction), SSE2 double,
 * so every + - * / and the operation order match the reference's Python/Numba
 * expressions.  libm exp/log/pow/sqrt stand in for numpy's (numpy's SIMD
 * exp/pow may differ from libm in the last ulp; noted where it matters).
 * The JOSH Gauss-Seidel sweep reproduces Numba's typing: float32 arrays,
 * float32 products accumulated in float64, float64 update rounded to float32
 * on store (physics/josh_solver.py:24-92, driven at 655-699).
 *
 * Parallelism: ASYNTH runs depth-parallel (each depth row is written only by
 * its own depth, so the result is bitwise equal to the reference's serial
 * line order); TRANSP is line-parallel; RT is wavelength-parallel.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NXTAU 51
#define MAX_PROFILE_STEPS 1000000

/* engine/opacity.py:39-45, engine/radiative.py:14-17 */
static const double H_PLANCK = 6.62607015e-27;
static const double K_BOLTZ = 1.380649e-16;
static const double C_LIGHT_CM = 2.99792458e10;
static const double C_LIGHT_NM = 2.99792458e17;
static const double NM_TO_CM = 1e-7;
static const double EPS38 = 1.0e-38; /* physics/josh_solver.py:16 */

/* ------------------------------------------------------------------ threads */
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------- MAP1 (table)
 * physics/tables.py:275-391 -- the MAP1 variant used to build H1TAB. */
void orc_map1_table(const double *xold, const double *fold, int nold,
                    const double *xnew, double *fnew, int nnew) {
#define XO(i) xold[(i) - 1]
#define FO(i) fold[(i) - 1]
    int l = 2, ll = 0;
    double a = 0, b = 0, c = 0, abac = 0, bbac = 0, cbac = 0, afor = 0, bfor = 0, cfor = 0;
    for (int k = 1; k <= nnew; ++k) {
        double x = xnew[k - 1];
        for (;;) { /* label 10 */
            if (x < XO(l)) break;
            l += 1;
            if (l > nold) break;
        }
        if (l > nold) { /* label 30 */
            if (l != ll) {
                l = nold < l ? nold : l;
                c = 0.0;
                b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
                a = FO(l) - XO(l) * b;
                ll = l;
            }
            fnew[k - 1] = a + (b + c * x) * x;
            continue;
        }
        if (l == ll) { fnew[k - 1] = a + (b + c * x) * x; continue; }
        if (l == 2) {
            c = 0.0;
            b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
            a = FO(l) - XO(l) * b;
            ll = l;
            fnew[k - 1] = a + (b + c * x) * x;
            continue;
        }
        int l1 = l - 1;
        if (l > ll + 1 || l == 3) { /* label 21 */
            int l2 = l - 2;
            double d = (FO(l1) - FO(l2)) / (XO(l1) - XO(l2));
            cbac = FO(l) / ((XO(l) - XO(l1)) * (XO(l) - XO(l2))) +
                   (FO(l2) / (XO(l) - XO(l2)) - FO(l1) / (XO(l) - XO(l1))) / (XO(l1) - XO(l2));
            bbac = d - (XO(l1) + XO(l2)) * cbac;
            abac = FO(l2) - XO(l2) * d + XO(l1) * XO(l2) * cbac;
            if (l == nold) { /* label 22 */
                c = cbac; b = bbac; a = abac; ll = l;
                fnew[k - 1] = a + (b + c * x) * x;
                continue;
            }
        } else {
            cbac = cfor; bbac = bfor; abac = afor;
            if (l == nold) {
                c = cbac; b = bbac; a = abac; ll = l;
                fnew[k - 1] = a + (b + c * x) * x;
                continue;
            }
        }
        /* label 25 */
        double d = (FO(l) - FO(l1)) / (XO(l) - XO(l1));
        cfor = FO(l + 1) / ((XO(l + 1) - XO(l)) * (XO(l + 1) - XO(l1))) +
               (FO(l1) / (XO(l + 1) - XO(l1)) - FO(l) / (XO(l + 1) - XO(l))) / (XO(l) - XO(l1));
        bfor = d - (XO(l) + XO(l1)) * cfor;
        afor = FO(l1) - XO(l1) * d + XO(l) * XO(l1) * cfor;
        double wt = 0.0;
        if (fabs(cfor) != 0.0) wt = fabs(cfor) / (fabs(cfor) + fabs(cbac));
        a = afor + wt * (abac - afor);
        b = bfor + wt * (bbac - bfor);
        c = cfor + wt * (cbac - cfor);
        ll = l;
        fnew[k - 1] = a + (b + c * x) * x;
    }
#undef XO
#undef FO
}

/* ------------------------------------------------------------ Voigt tables
 * physics/tables.py:224-241 (VoigtTables.build, vsteps=200, size=2001). */
void orc_build_voigt_tables(const double *tabvi, const double *tabh1, int nold,
                            double *h0, double *h1, double *h2, int n) {
    for (int i = 0; i < n; ++i) h0[i] = (double)i / 200.0;
    orc_map1_table(tabvi, tabh1, nold, h0, h1, n);
    for (int i = 0; i < n; ++i) {
        double q = (double)i / 200.0;
        double vv = q * q; /* Python ** 2 on a float: pow(q, 2) == q*q exactly */
        h0[i] = exp(-vv);
        h2[i] = h0[i] - (vv + vv) * h0[i];
    }
}

/* physics/tables.py:20-27 (ExpTables.build, size 1001). */
void orc_build_fastex_tables(double *extab, double *extabf, int n) {
    for (int i = 0; i < n; ++i) {
        extab[i] = exp(-(double)i);
        extabf[i] = exp(-(double)i * 0.001);
    }
}

/* physics/tables.py:29-42 (ExpTables.fast_ex). */
double orc_fastex(double x, const double *extab, const double *extabf, int n) {
    if (x <= 0.0) return x == 0.0 ? 1.0 : exp(-x);
    double fl = floor(x);
    if (fl >= (double)n) return exp(-x);
    int i = (int)fl;
    double frac = x - (double)i;
    int j = (int)(frac * 1000.0 + 0.5);
    if (j < 0) j = 0;
    if (j > n - 1) j = n - 1;
    return extab[i] * extabf[j];
}

/* physics/voigt_jit.py:16-69 (voigt_profile_jit); the pure-Python twin
 * physics/profiles/voigt.py:10-70 performs the same operations. */
double orc_voigt(double v, double a, const double *h0t, const double *h1t, const double *h2t, int n) {
    double av = fabs(v);
    double fiv = av * 200.0 + 0.5;
    int iv;
    if (fiv >= (double)(n - 1)) iv = n - 1;
    else iv = (int)fiv;
    if (iv < 0) iv = 0;
    if (a < 0.2) {
        if (av > 10.0) return 0.5642 * a / (v * v);
        return (h2t[iv] * a + h1t[iv]) * a + h0t[iv];
    } else if (a > 1.4 || (a + av) > 3.2) {
        double aa = a * a, vv = v * v;
        double u = (aa + vv) * 1.4142;
        double voigt = a * 0.79788 / u;
        if (a <= 100.0) {
            double aau = aa / u, vvu = vv / u, uu = u * u;
            voigt = ((((aau - 10.0 * vvu) * aau * 3.0 + 15.0 * vvu * vvu) + 3.0 * vv - aa) / uu + 1.0) * voigt;
        }
        return voigt;
    } else {
        double vv = v * v;
        double H0 = h0t[iv];
        double H1 = h1t[iv] + H0 * 1.12838;
        double H2 = h2t[iv] + H1 * 1.12838 - H0;
        double H3 = (1.0 - h2t[iv]) * 0.37613 - H1 * 0.66667 * vv + H2 * 1.12838;
        double H4 = (3.0 * H3 - H1) * 0.37613 + H0 * 0.66667 * vv * vv;
        double pa = (((H4 * a + H3) * a + H2) * a + H1) * a + H0;
        double pb = ((-0.122727278 * a + 0.532770573) * a - 0.96284325) * a + 0.979895032;
        return pa * pb;
    }
}

/* -------------------------------------------------------- wavelength grid
 * engine/opacity.py:526-547 (_build_wavelength_grid).  Returns the count;
 * writes at most cap values (call with out=NULL to count). */
long orc_wavelength_grid(double start, double end, double resolution, double *out, long cap) {
    double ratio = 1.0 + 1.0 / resolution;
    double rlog = log(ratio);
    double ix_start = log(start) / rlog;
    double ix_floor = floor(ix_start);
    if (exp(ix_floor * rlog) < start) ix_floor += 1.0;
    double wl = exp(ix_floor * rlog);
    long n = 0;
    double lim = end * (1.0 + 1e-9);
    while (wl <= lim) {
        if (out && n < cap) out[n] = wl;
        ++n;
        wl *= ratio;
    }
    return n;
}

/* engine/opacity.py:550-592 (_nearest_grid_indices). */
void orc_nearest_grid_indices(const double *grid, long W, const double *values, long n, int64_t *out) {
    if (W < 2) { for (long i = 0; i < n; ++i) out[i] = 0; return; }
    double ratio = grid[1] / grid[0];
    double ratiolg = log(ratio);
    int64_t ix_start = (int64_t)(log(grid[0]) / ratiolg + 0.5);
    for (long i = 0; i < n; ++i) {
        double v = values[i];
        int64_t ix = (int64_t)(log(v) / ratiolg + 0.5);
        int64_t idx = ix - ix_start;
        if (v < grid[0]) idx = -1;
        if (v > grid[W - 1]) idx = W;
        out[i] = idx;
    }
}

/* physics/line_opacity.py:1030-1059 (_nearest_grid_indices_raw with the
 * grid_origin re-basing).  origin <= 0 means "no origin" (grid[0]). */
void orc_wing_center_indices(const double *grid, long W, const double *values, long n,
                             double origin, int64_t *out) {
    if (W < 2) { for (long i = 0; i < n; ++i) out[i] = 0; return; }
    double ratio = grid[1] / grid[0];
    double ratiolg = log(ratio);
    double start_val = origin > 0.0 ? origin : grid[0];
    double ix_floor = floor(log(start_val) / ratiolg);
    double wbegin = exp(ix_floor * ratiolg);
    if (wbegin < start_val) { ix_floor += 1.0; wbegin = exp(ix_floor * ratiolg); }
    int64_t off = 0;
    if (origin > 0.0) {
        double fl = floor(log(origin) / ratiolg);
        double wb = exp(fl * ratiolg);
        if (wb < origin) { fl += 1.0; wb = exp(fl * ratiolg); }
        off = (int64_t)rint(log(grid[0] / wb) / ratiolg);
    }
    for (long i = 0; i < n; ++i)
        out[i] = (int64_t)rint(log(values[i] / wbegin) / ratiolg) - off;
}

/* ------------------------------------------------------------- continuum
 * physics/continuum.py:73-303 (build_depth_continuum, coefficient path).
 * coeff arrays are [D][E-1][3]; wledge [E] (abs taken); half/delta [E-1]
 * (pass NULL to derive them as the reference does). */
void orc_continuum(long D, long W, const double *wl, long E, const double *wledge_in,
                   const double *half_in, const double *delta_in, const double *cabs,
                   const double *cscat, int coeff_log10, double *out_abs, double *out_scat) {
    double *wledge = (double *)malloc(sizeof(double) * E);
    double *half = (double *)malloc(sizeof(double) * (E - 1));
    double *delta = (double *)malloc(sizeof(double) * (E - 1));
    for (long e = 0; e < E; ++e) wledge[e] = fabs(wledge_in[e]);
    for (long e = 0; e < E - 1; ++e) {
        half[e] = half_in ? half_in[e] : 0.5 * (wledge[e] + wledge[e + 1]);
        if (delta_in) delta[e] = delta_in[e];
        else { double dd = wledge[e + 1] - wledge[e]; delta[e] = 0.5 * dd * dd; }
        if (delta[e] == 0.0) delta[e] = 1e-20;
    }
#pragma omp parallel for schedule(static)
    for (long w = 0; w < W; ++w) {
        double x = wl[w];
        double ax = fabs(x);
        /* searchsorted(wledge, |wl|, side='right') - 1, clipped to [0, E-2] */
        long lo = 0, hi = E; /* first index with wledge > ax */
        while (lo < hi) { long mid = (lo + hi) >> 1; if (wledge[mid] <= ax) lo = mid + 1; else hi = mid; }
        long e = lo - 1;
        if (e < 0) e = 0;
        if (e > E - 2) e = E - 2;
        double left = wledge[e], right = wledge[e + 1], h = half[e], dl = delta[e];
        double c1 = (x - h) * (x - right) / dl;
        double c2 = (left - x) * (x - right) * 2.0 / dl;
        double c3 = (x - left) * (x - h) / dl;
        for (long d = 0; d < D; ++d) {
            const double *ca = cabs + (d * (E - 1) + e) * 3;
            const double *cs = cscat + (d * (E - 1) + e) * 3;
            double la = ca[0] * c1 + ca[1] * c2 + ca[2] * c3;
            double ls = cs[0] * c1 + cs[1] * c2 + cs[2] * c3;
            double va, vs;
            if (coeff_log10) {
                if (la > 38.0) la = 38.0;
                if (ls > 38.0) ls = 38.0;
                va = pow(10.0, la);
                vs = pow(10.0, ls);
                if (va < 0.0) va = 0.0;
                if (va > 3.4e38) va = 3.4e38;
                if (vs < 0.0) vs = 0.0;
                if (vs > 3.4e38) vs = 3.4e38;
                if (!isfinite(va)) va = 3.4e38;
                if (!isfinite(vs)) vs = 3.4e38;
            } else {
                va = la; vs = ls; /* the reference raises here; callers never pass 0 */
            }
            out_abs[d * W + w] = va;
            out_scat[d * W + w] = vs;
        }
    }
    free(wledge); free(half); free(delta);
}

/* --------------------------------------------------- frequency quantities
 * engine/opacity.py:2215-2228 (freq, ehvkt, stim, bnu) and
 * physics/bfudge.py:72-168 in LTE (BFUDGE = 1): slinec = bnu*stim/(1-ehvkt).
 * Writes [D][W] arrays; any output pointer may be NULL. */
void orc_freq_quantities(long D, long W, const double *wl, const double *temp,
                         double *ehvkt_o, double *stim_o, double *bnu_o, double *slinec_o) {
#pragma omp parallel for schedule(static)
    for (long d = 0; d < D; ++d) {
        double hkt = H_PLANCK / (K_BOLTZ * temp[d]);
        for (long w = 0; w < W; ++w) {
            double freq = C_LIGHT_CM / (wl[w] * NM_TO_CM);
            double ehvkt = exp(-freq * hkt);
            double stim = 1.0 - ehvkt;
            double freq15 = freq / 1.0e15;
            double bnu = 1.47439e-02 * pow(freq15, 3.0) * ehvkt / stim;
            long i = d * W + w;
            if (ehvkt_o) ehvkt_o[i] = ehvkt;
            if (stim_o) stim_o[i] = stim;
            if (bnu_o) bnu_o[i] = bnu;
            if (slinec_o) slinec_o[i] = (bnu * stim) / (1.0 - ehvkt);
        }
    }
}

/* ----------------------------------------------------------------- TRANSP
 * physics/line_opacity.py:56-184 (_compute_transp_numba_kernel) with the
 * Boltzmann factor of physics/populations.py:172 (FASTEX of elo*hckt) fused
 * in.  Per-line arrays [L]; pop/dop [D][nion][nelem]; kapmin [D][W];
 * outputs transp [L][D] (zeroed here) and valid [L][D]. */
typedef struct {
    const double *h0, *h1, *h2;
    int nvoigt;
    const double *extab, *extabf;
    int nfastex;
} orc_tables;

void orc_transp(long L, long D, long W, const uint8_t *process, const int32_t *elem,
                const int32_t *ion, const int32_t *ltype, const double *wl, const double *gf,
                const double *cgf, const double *grad, const double *gstark, const double *gvdw,
                const double *elo, const int64_t *center, const double *hckt,
                const double *pop, const double *dop, long nion, long nelem,
                const double *rho, const double *xne, const double *txnxn,
                const double *kapmin_grid, double cutoff, double microturb_kms,
                const orc_tables *tb, double *transp, uint8_t *valid) {
    double micro = microturb_kms > 0.0 ? microturb_kms / 299792.458 : 0.0;
#pragma omp parallel for schedule(dynamic, 256)
    for (long l = 0; l < L; ++l) {
        for (long d = 0; d < D; ++d) { transp[l * D + d] = 0.0; valid[l * D + d] = 0; }
        if (!process[l]) continue;
        int ei = elem[l];
        if (ei < 0 || ei >= nelem) continue;
        int nel = ion[l];
        if (nel <= 0 || nel > nion) continue;
        double lw = wl[l];
        long c = center[l];
        long cc = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
        for (long d = 0; d < D; ++d) {
            double pv = pop[(d * nion + (nel - 1)) * nelem + ei];
            double dv = dop[(d * nion + (nel - 1)) * nelem + ei];
            if (micro > 0.0) dv = sqrt(dv * dv + micro * micro);
            if (pv <= 0.0 || dv <= 0.0) continue;
            double r = rho[d];
            if (r <= 0.0) continue;
            double xnfdop = pv / (r * dv);
            double dw = dv * lw;
            double boltz = orc_fastex(elo[l] * hckt[d], tb->extab, tb->extabf, tb->nfastex);
            double k0pre = ltype[l] == 1 ? gvdw[l] * gf[l] * (pv / r) : cgf[l] * xnfdop;
            double kapmin = kapmin_grid[d * W + cc] * cutoff;
            if (k0pre < kapmin) continue;
            double post = k0pre * boltz;
            if (post < kapmin) continue;
            double kappa0 = post;
            if (kappa0 <= 0.0) continue;
            double dopple = lw > 0 ? dw / lw : 1e-6;
            double adamp;
            if (dw > 0 && dopple > 0) {
                double gt = grad[l] + gstark[l] * xne[d] + gvdw[l] * txnxn[d];
                adamp = gt / dopple;
            } else adamp = 0.0;
            if (adamp >= 0.0 && kappa0 > 0.0) {
                double kapcen;
                if (ltype[l] == 1) kapcen = kappa0;
                else if (adamp < 0.2) kapcen = kappa0 * (1.0 - 1.128 * adamp);
                else kapcen = kappa0 * orc_voigt(0.0, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
                transp[l * D + d] = kapcen;
                valid[l * D + d] = 1;
            }
        }
    }
}

/* Per-(line, depth) wing parameters: the ASYNTH prep (line_opacity.py:1160-1272)
 * and phases 1-2 of the wing kernel (line_opacity.py:410-537). */
typedef struct {
    double kappa0, adamp, dopple, dvoigt, x_far, kapmin_ref;
    long n10dop, maxstep;
    int use_far;
} orc_wing;

static int wing_setup(long l, long d, long W, double resolu, const int32_t *elem,
                      const int32_t *ion, const double *wl, const double *grad,
                      const double *gstark, const double *gvdw, const int64_t *center_w,
                      double tv, const double *pop, const double *dop, long nion, long nelem,
                      double xne_d, double txn_d, const double *kapmin_grid, double cutoff,
                      const orc_tables *tb, orc_wing *out, double *st) {
    int ei = elem[l], nel = ion[l];
    if (ei < 0 || ei >= nelem || nel <= 0 || nel > nion) return 0;
    double pv = pop[(d * nion + (nel - 1)) * nelem + ei];
    double dv = dop[(d * nion + (nel - 1)) * nelem + ei];
    if (pv <= 0.0 || dv <= 0.0) return 0;
    double lw = wl[l];
    double dw = dv * lw;
    double dopple_p = lw > 0 ? dw / lw : 1e-6;
    double adamp;
    if (dopple_p > 0 && lw > 0) {
        double gt = grad[l] + gstark[l] * xne_d + gvdw[l] * txn_d;
        adamp = gt / dopple_p;
    } else adamp = 0.0;
    if (adamp < 1e-12) adamp = 1e-12;
    long c = center_w[l];
    long cw = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
    double kapmin_ref = kapmin_grid[d * W + cw] * cutoff;
    double vc = adamp < 0.2 ? 1.0 - 1.128 * adamp
                            : orc_voigt(0.0, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
    double kappa0 = vc > 0 ? tv / vc : tv;
    if (dw <= 0.0) return 0;
    double dopple = lw > 0.0 ? dw / lw : 1e-10;
    long n10dop = (long)(10.0 * dopple * resolu);
    double dvoigt = dopple > 0 ? 1.0 / (dopple * resolu) : 1.0;
    double tabstep = 200.0 * dvoigt;
    int nv = tb->nvoigt;
    long nstep_cutoff = -1;
    double p_n10 = 0.0;
    double tabi = 0.5;
    for (long k = 1; k <= n10dop; ++k) {
        double p;
        if (adamp < 0.2) {
            tabi += tabstep;
            long idx = (long)tabi;
            if (idx < 0) idx = 0;
            double x = (double)k * dvoigt;
            if (x > 10.0) p = kappa0 * (0.5642 * adamp / (x * x));
            else {
                if (idx >= nv) idx = nv - 1;
                p = kappa0 * (tb->h0[idx] + adamp * tb->h1[idx]);
            }
        } else {
            double x = (double)k * dvoigt;
            p = kappa0 * orc_voigt(x, adamp, tb->h0, tb->h1, tb->h2, nv);
        }
        if (st) st[1] += 1;
        if (k == n10dop) p_n10 = p;
        if (p < kapmin_ref) { nstep_cutoff = k; break; }
    }
    long maxstep;
    int use_far;
    double x_far;
    if (nstep_cutoff != -1) { maxstep = nstep_cutoff; use_far = 0; x_far = 0.0; }
    else {
        use_far = 1;
        if (n10dop > 0 && p_n10 > 0.0 && kapmin_ref > 0.0) {
            x_far = p_n10 * ((double)n10dop * (double)n10dop);
            double ms = sqrt(x_far / kapmin_ref) + 1.0;
            maxstep = ms < 2.0e6 ? (long)ms : MAX_PROFILE_STEPS;
        } else { x_far = 0.0; maxstep = 0; }
        if (maxstep > MAX_PROFILE_STEPS) maxstep = MAX_PROFILE_STEPS;
    }
    if (st && maxstep > 0) {
        st[2] += 1;
        st[3] += (double)maxstep;
        if ((double)maxstep > st[4]) st[4] = (double)maxstep;
    }
    out->kappa0 = kappa0; out->adamp = adamp; out->dopple = dopple; out->dvoigt = dvoigt;
    out->x_far = x_far; out->kapmin_ref = kapmin_ref; out->n10dop = n10dop;
    out->maxstep = maxstep; out->use_far = use_far;
    return 1;
}

/* ----------------------------------------------------------------- ASYNTH
 * physics/line_opacity.py:946-1340 (compute_asynth_from_transp): centre adds
 * (1091-1102), per-(l,d) prep (1141-1294, WCON/WTAIL zero), the serial wing
 * kernel for line_type 0 (187-617, phases 1-3), then asynth *= stim_grid
 * (1338).  line_type 1 (autoionising) wings are not restated: the hot path's
 * line lists carry none (they are fort.19 lines, out of scope).
 *
 * stats (may be NULL) receives, summed over depths:
 *   [0] E_line  = centre adds + wing adds (line x depth x lambda evaluations)
 *   [1] P_line  = near-wing probe evaluations (phase 1)
 *   [2] pairs with wings (maxstep > 0)
 *   [3] sum of maxstep over those pairs
 *   [4] max maxstep */
void orc_asynth(long L, long D, long W, const double *grid, const int32_t *ltype,
                const int32_t *elem, const int32_t *ion, const double *wl, const double *grad,
                const double *gstark, const double *gvdw, const int64_t *center_t,
                const int64_t *center_w, const double *transp, const uint8_t *valid,
                const double *pop, const double *dop, long nion, long nelem,
                const double *xne, const double *txnxn, const double *temp,
                const double *kapmin_grid, double cutoff, const orc_tables *tb,
                double *asynth, double *stats) {
    double resolu = 300000.0;
    if (W > 1) {
        double ratio = grid[1] / grid[0];
        if (ratio > 1.0) resolu = 1.0 / (ratio - 1.0);
    }
    int nth = orc_get_threads();
    double *tstats = (double *)calloc((size_t)nth * 5, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (long d = 0; d < D; ++d) {
#ifdef _OPENMP
        double *st = tstats + 5 * omp_get_thread_num();
#else
        double *st = tstats;
#endif
        double *row = asynth + d * W;
        for (long w = 0; w < W; ++w) row[w] = 0.0;
        /* centre contributions, catalog order (line_opacity.py:1091-1102) */
        for (long l = 0; l < L; ++l) {
            int t = ltype[l];
            if (t == -2 || t == -1 || t == 1 || t == 2 || t == 3 || t == 4) continue;
            long c = center_t[l];
            if (c >= 0 && c < W && valid[l * D + d]) { row[c] += transp[l * D + d]; st[0] += 1; }
        }
        double xne_d = xne[d], txn_d = txnxn[d];
        for (long l = 0; l < L; ++l) {
            if (ltype[l] != 0) continue;
            long c = center_w[l];
            if (c < -MAX_PROFILE_STEPS || c > W - 1 + MAX_PROFILE_STEPS) continue;
            if (!valid[l * D + d]) continue;
            double tv = transp[l * D + d];
            if (tv <= 0.0) continue;
            orc_wing wg;
            if (!wing_setup(l, d, W, resolu, elem, ion, wl, grad, gstark, gvdw,
                            center_w, tv, pop, dop, nion, nelem, xne_d, txn_d, kapmin_grid,
                            cutoff, tb, &wg, st))
                continue;
            long n10dop = wg.n10dop, maxstep = wg.maxstep;
            int use_far = wg.use_far;
            double x_far = wg.x_far, kappa0 = wg.kappa0, adamp = wg.adamp, dvoigt = wg.dvoigt;
            double tabstep = 200.0 * dvoigt;
            int nv = tb->nvoigt;
            /* phase 3: red and blue wings, no per-step cutoff */
            int red = 1, blue = 1;
            double tabo = 0.5;
            for (long k = 1; k <= maxstep && (red || blue); ++k) {
                double p;
                if (use_far && k > n10dop) p = x_far / ((double)k * (double)k);
                else if (adamp < 0.2) {
                    tabo += tabstep;
                    long idx = (long)tabo;
                    if (idx < 0) idx = 0;
                    double x = (double)k * dvoigt;
                    if (x > 10.0) p = kappa0 * (0.5642 * adamp / (x * x));
                    else {
                        if (idx >= nv) idx = nv - 1;
                        p = kappa0 * (tb->h0[idx] + adamp * tb->h1[idx]);
                    }
                } else {
                    double x = (double)k * dvoigt;
                    p = kappa0 * orc_voigt(x, adamp, tb->h0, tb->h1, tb->h2, nv);
                }
                p = p * 1.0; /* stim_factors are ones (line_opacity.py:1070-1071) */
                if (red) {
                    long i = c + k;
                    if (i >= W) red = 0;
                    else if (i >= 0) { row[i] += p; st[0] += 1; }
                }
                if (blue) {
                    long i = c - k;
                    if (i < 0) blue = 0;
                    else if (i < W) { row[i] += p; st[0] += 1; }
                }
            }
        }
        /* asynth *= stim_grid (line_opacity.py:1005-1011, 1067, 1338) */
        double T = temp[d] > 1.0 ? temp[d] : 1.0;
        double hkt = H_PLANCK / (K_BOLTZ * T);
        for (long w = 0; w < W; ++w) {
            double fg = C_LIGHT_NM / grid[w];
            row[w] *= 1.0 - exp(-fg * hkt);
        }
    }
    if (stats) {
        for (int i = 0; i < 5; ++i) stats[i] = 0.0;
        for (int t = 0; t < nth; ++t) {
            for (int i = 0; i < 4; ++i) stats[i] += tstats[5 * t + i];
            if (tstats[5 * t + 4] > stats[4]) stats[4] = tstats[5 * t + 4];
        }
    }
    free(tstats);
}


/* Wing parameters of every (line, depth) pair, for checking the GPU probe:
 * out_d [L][D][6] = kappa0, adamp, dvoigt, x_far, kapmin_ref, dopple;
 * out_i [L][D][3] = n10dop, maxstep, use_far (all -1 where no wings). */
void orc_wing_params(long L, long D, long W, const double *grid, const int32_t *ltype,
                     const int32_t *elem, const int32_t *ion, const double *wl, const double *grad,
                     const double *gstark, const double *gvdw, const int64_t *center_w,
                     const double *transp, const uint8_t *valid, const double *pop,
                     const double *dop, long nion, long nelem, const double *xne,
                     const double *txnxn, const double *kapmin_grid, double cutoff,
                     const orc_tables *tb, double *out_d, int64_t *out_i) {
    double resolu = 300000.0;
    if (W > 1) {
        double ratio = grid[1] / grid[0];
        if (ratio > 1.0) resolu = 1.0 / (ratio - 1.0);
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (long l = 0; l < L; ++l) {
        for (long d = 0; d < D; ++d) {
            double *od = out_d + (l * D + d) * 6;
            int64_t *oi = out_i + (l * D + d) * 3;
            for (int q = 0; q < 6; ++q) od[q] = 0.0;
            oi[0] = oi[1] = oi[2] = -1;
            if (ltype[l] != 0) continue;
            long c = center_w[l];
            if (c < -MAX_PROFILE_STEPS || c > W - 1 + MAX_PROFILE_STEPS) continue;
            if (!valid[l * D + d]) continue;
            double tv = transp[l * D + d];
            if (tv <= 0.0) continue;
            orc_wing wg;
            if (!wing_setup(l, d, W, resolu, elem, ion, wl, grad, gstark, gvdw, center_w, tv, pop,
                            dop, nion, nelem, xne[d], txnxn[d], kapmin_grid, cutoff, tb, &wg, NULL))
                continue;
            od[0] = wg.kappa0; od[1] = wg.adamp; od[2] = wg.dvoigt; od[3] = wg.x_far;
            od[4] = wg.kapmin_ref; od[5] = wg.dopple;
            oi[0] = wg.n10dop; oi[1] = wg.maxstep; oi[2] = wg.use_far;
        }
    }
}

/* ------------------------------------------------------------ JOSH pieces */
static double g_xtau[NXTAU], g_ch[NXTAU], g_coefj[NXTAU * NXTAU];
static float g_coefj32[NXTAU * NXTAU], g_diag32[NXTAU];

/* physics/josh_tables.py (XTAU_GRID, CH_WEIGHTS, COEFJ_MATRIX) */
void orc_set_josh_tables(const double *xtau, const double *ch, const double *coefj) {
    memcpy(g_xtau, xtau, sizeof(g_xtau));
    memcpy(g_ch, ch, sizeof(g_ch));
    memcpy(g_coefj, coefj, sizeof(g_coefj));
    for (int i = 0; i < NXTAU * NXTAU; ++i) g_coefj32[i] = (float)coefj[i];
    for (int k = 0; k < NXTAU; ++k) g_diag32[k] = (float)coefj[k * NXTAU + k];
}

/* physics/josh_solver.py:95-167 (_parcoe) */
void orc_parcoe(const double *f, const double *x, long n, double *a, double *b, double *c) {
    for (long i = 0; i < n; ++i) a[i] = b[i] = c[i] = 0.0;
    if (n == 0) return;
    if (n == 1) { a[0] = f[0]; return; }
    c[0] = 0.0;
    double denom = x[1] - x[0];
    b[0] = denom != 0.0 ? (f[1] - f[0]) / denom : 0.0;
    a[0] = f[0] - x[0] * b[0];
    long n1 = n - 1;
    c[n - 1] = 0.0;
    denom = x[n - 1] - x[n1 - 1];
    b[n - 1] = denom != 0.0 ? (f[n - 1] - f[n1 - 1]) / denom : 0.0;
    a[n - 1] = f[n - 1] - x[n - 1] * b[n - 1];
    if (n == 2) return;
    for (long j = 1; j < n1; ++j) {
        long j1 = j - 1;
        denom = x[j] - x[j1];
        double d = denom != 0.0 ? (f[j] - f[j1]) / denom : 0.0;
        double denom1 = (x[j + 1] - x[j]) * (x[j + 1] - x[j1]);
        double term1 = denom1 != 0.0 ? f[j + 1] / denom1 : 0.0;
        double denom2 = x[j + 1] - x[j1];
        double denom3 = x[j + 1] - x[j];
        double denom4 = x[j] - x[j1];
        double part = 0.0;
        if (denom4 != 0.0) {
            double t1 = denom2 != 0.0 ? f[j1] / denom2 : 0.0;
            double t2 = denom3 != 0.0 ? f[j] / denom3 : 0.0;
            part = (t1 - t2) / denom4;
        }
        c[j] = term1 + part;
        b[j] = d - (x[j] + x[j1]) * c[j];
        a[j] = f[j1] - x[j1] * d + x[j] * x[j1] * c[j];
    }
    c[1] = 0.0;
    denom = x[2] - x[1];
    b[1] = denom != 0.0 ? (f[2] - f[1]) / denom : 0.0;
    a[1] = f[1] - x[1] * b[1];
    if (n > 3) {
        c[2] = 0.0;
        denom = x[3] - x[2];
        b[2] = denom != 0.0 ? (f[3] - f[2]) / denom : 0.0;
        a[2] = f[2] - x[2] * b[2];
    }
    for (long j = 1; j < n1; ++j) {
        if (c[j] == 0.0) continue;
        long j1 = j + 1 < n - 1 ? j + 1 : n - 1;
        double dn = fabs(c[j1]) + fabs(c[j]);
        double wt = dn > 0.0 ? fabs(c[j1]) / dn : 0.0;
        a[j] = a[j1] + wt * (a[j] - a[j1]);
        b[j] = b[j1] + wt * (b[j] - b[j1]);
        c[j] = c[j1] + wt * (c[j] - c[j1]);
    }
    a[n1 - 1] = a[n - 1];
    b[n1 - 1] = b[n - 1];
    c[n1 - 1] = c[n - 1];
}

/* physics/josh_solver.py:170-189 (_integ) */
void orc_integ(const double *x, const double *f, long n, double start, double *fint) {
    if (n == 0) return;
    double *a = (double *)malloc(sizeof(double) * 3 * n);
    double *b = a + n, *c = a + 2 * n;
    orc_parcoe(f, x, n, a, b, c);
    fint[0] = start;
    for (long i = 0; i + 1 < n; ++i) {
        double dx = x[i + 1] - x[i];
        double term = a[i] + 0.5 * b[i] * (x[i + 1] + x[i]);
        term += (c[i] / 3.0) * ((x[i + 1] + x[i]) * x[i + 1] + x[i] * x[i]);
        fint[i + 1] = fint[i] + term * dx;
    }
    free(a);
}

/* physics/josh_solver.py:192-217 (_deriv) */
void orc_deriv(const double *x, const double *f, long n, double *dfdx) {
    for (long i = 0; i < n; ++i) dfdx[i] = 0.0;
    if (n < 2) return;
    dfdx[0] = (f[1] - f[0]) / (x[1] - x[0]);
    dfdx[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
    if (n == 2) return;
    double s = x[1] != x[0] ? fabs(x[1] - x[0]) / (x[1] - x[0]) : 1.0;
    for (long j = 1; j < n - 1; ++j) {
        double scale = fabs(f[j - 1]);
        if (fabs(f[j]) > scale) scale = fabs(f[j]);
        if (fabs(f[j + 1]) > scale) scale = fabs(f[j + 1]);
        scale = x[j] != 0.0 ? scale / fabs(x[j]) : scale;
        if (scale == 0.0) scale = 1.0;
        double d1 = (f[j + 1] - f[j]) / (x[j + 1] - x[j]) / scale;
        double d0 = (f[j] - f[j - 1]) / (x[j] - x[j - 1]) / scale;
        double tan1 = d1 / (s * sqrt(1.0 + d1 * d1) + 1.0);
        double tan0 = d0 / (s * sqrt(1.0 + d0 * d0) + 1.0);
        dfdx[j] = (tan1 + tan0) / (1.0 - tan1 * tan0) * scale;
    }
}

/* physics/josh_solver.py:220-353 (_map1_kernel).  Returns maxj = ll - 1. */
int orc_map1_josh(const double *xold, const double *fold, int nold, const double *xnew,
                  double *fnew, int nnew) {
#define XO(i) xold[(i) - 1]
#define FO(i) fold[(i) - 1]
    for (int k = 0; k < nnew; ++k) fnew[k] = 0.0;
    if (nold == 0 || nnew == 0) return 0;
    int l = 2, ll = 0;
    double cfor = 0, bfor = 0, afor = 0, cbac = 0, bbac = 0, abac = 0, a = 0, b = 0, c = 0;
    for (int k = 1; k <= nnew; ++k) {
        double xk = xnew[k - 1];
        for (;;) {
            if (xk < XO(l)) { /* label 20 */
                if (l == ll) break;
                if (l == 2 || l == 3) { /* label 30 */
                    l = nold < l ? nold : l;
                    c = 0.0;
                    b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
                    a = FO(l) - XO(l) * b;
                    ll = l;
                    break;
                }
                int l1 = l - 1;
                if (l > ll + 1 || l == 4) { /* label 21 (the first test's l==3 arm is unreachable) */
                    int l2 = l - 2;
                    double d = (FO(l1) - FO(l2)) / (XO(l1) - XO(l2));
                    cbac = FO(l) / ((XO(l) - XO(l1)) * (XO(l) - XO(l2))) +
                           (FO(l2) / (XO(l) - XO(l2)) - FO(l1) / (XO(l) - XO(l1))) / (XO(l1) - XO(l2));
                    bbac = d - (XO(l1) + XO(l2)) * cbac;
                    abac = FO(l2) - XO(l2) * d + XO(l1) * XO(l2) * cbac;
                    if (!(l < nold)) { c = cbac; b = bbac; a = abac; ll = l; break; } /* label 22 */
                } else {
                    cbac = cfor; bbac = bfor; abac = afor;
                    if (l == nold) { c = cbac; b = bbac; a = abac; ll = l; break; }
                }
                /* label 25 */
                double d = (FO(l) - FO(l1)) / (XO(l) - XO(l1));
                cfor = FO(l + 1) / ((XO(l + 1) - XO(l)) * (XO(l + 1) - XO(l1))) +
                       (FO(l1) / (XO(l + 1) - XO(l1)) - FO(l) / (XO(l + 1) - XO(l))) / (XO(l) - XO(l1));
                bfor = d - (XO(l) + XO(l1)) * cfor;
                afor = FO(l1) - XO(l1) * d + XO(l) * XO(l1) * cfor;
                double wt = 0.0;
                if (fabs(cfor) != 0.0) wt = fabs(cfor) / (fabs(cfor) + fabs(cbac));
                a = afor + wt * (abac - afor);
                b = bfor + wt * (bbac - bfor);
                c = cfor + wt * (cbac - cfor);
                ll = l;
                break;
            }
            l += 1;
            if (l > nold) { /* label 30 */
                l = nold < l ? nold : l;
                c = 0.0;
                b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
                a = FO(l) - XO(l) * b;
                ll = l;
                break;
            }
        }
        fnew[k - 1] = a + (b + c * xk) * xk;
    }
    return ll - 1;
#undef XO
#undef FO
}

/* physics/josh_solver.py:24-92 (_josh_iteration_kernel) as Numba types it
 * when driven from 655-680: float32 arrays, float64 accumulator. */
/* Test-only smooth mode (orc_set_josh_fp64): the scattering sweep in FP64
 * on the same float32 COEFJ values and the label-401 iteration, both to a
 * relative change below tol.  The reference's float32 sweep stops at 1e-5,
 * which makes finite differences of the flux noisy; this mode is the FP64
 * reference the adjoint's finite-difference checks use.  Off by default. */
static int g_josh_fp64 = 0;
static double g_josh_tol = 1e-13;

void orc_set_josh_fp64(int on, double tol) {
    g_josh_fp64 = on;
    g_josh_tol = tol;
}

static int josh_iterate64(double *xs, const double *xa, const double *xb) {
    for (int k = 0; k < NXTAU; ++k) xs[k] = xb[k];
    for (int it = 0; it < 400; ++it) {
        double err = 0.0;
        for (int k = NXTAU - 1; k >= 0; --k) {
            double cx = 0.0;
            for (int m = 0; m < NXTAU; ++m) cx += (double)g_coefj32[k * NXTAU + m] * xs[m];
            double ckk = (double)g_coefj32[k * NXTAU + k];
            double nv = (xb[k] * (1.0 - xa[k]) + xa[k] * (cx - ckk * xs[k])) / (1.0 - xa[k] * ckk);
            double e = fabs(nv - xs[k]) / (fabs(nv) > EPS38 ? fabs(nv) : EPS38);
            if (e > err) err = e;
            xs[k] = nv;
        }
        if (err < g_josh_tol) return it + 1;
    }
    return 400;
}

int orc_josh_iterate(float *xs, const float *xalpha, const float *xsbar_mod, int max_iter) {
    float diag[NXTAU];
    for (int k = 0; k < NXTAU; ++k) diag[k] = 1.0f - xalpha[k] * g_diag32[k];
    const double tol = (double)(float)1.0e-5;
    const double eps = (double)(float)EPS38;
    for (int it = 0; it < max_iter; ++it) {
        int iferr = 0;
        for (int k = NXTAU - 1; k >= 0; --k) {
            double delxs = 0.0;
            const float *row = g_coefj32 + k * NXTAU;
            for (int m = 0; m < NXTAU; ++m) delxs += (double)(row[m] * xs[m]);
            delxs = (delxs * (double)xalpha[k] + (double)xsbar_mod[k] - (double)xs[k]) / (double)diag[k];
            double errorx = xs[k] != 0.0f ? fabs(delxs / (double)xs[k]) : INFINITY;
            if (errorx > tol) iferr = 1;
            double nv = (double)xs[k] + delxs;
            if (eps > nv) nv = eps; /* max(a, b) keeps a unless b > a */
            xs[k] = (float)nv;
        }
        if (iferr == 0) return it + 1;
    }
    return max_iter;
}

/* numpy's pairwise summation for n <= 128 (8 partial accumulators), used by
 * the label-401 convergence sum (physics/josh_solver.py:712). */
static double np_pairwise_sum(const double *a, long n) {
    if (n < 8) {
        double res = 0.0;
        for (long i = 0; i < n; ++i) res += a[i];
        return res;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        long i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    long n2 = n / 2;
    n2 -= n2 % 8;
    return np_pairwise_sum(a, n2) + np_pairwise_sum(a + n2, n - n2);
}

/* physics/josh_solver.py:367-767 (solve_josh_flux).  Arrays of length n in
 * the caller's depth order.  info (may be NULL): [0] iterations, [1] 1 if
 * the label-401 branch ran. */
double orc_solve_josh_flux(long n, const double *acont, const double *scont, const double *aline,
                           const double *sline, const double *sigmac, const double *sigmal,
                           const double *rhox, int *info) {
    if (info) { info[0] = 0; info[1] = 0; }
    if (n == 0) return 0.0;
    double *buf = (double *)malloc(sizeof(double) * 10 * n);
    double *abtot = buf, *alpha = buf + n, *snubar = buf + 2 * n, *rho = buf + 3 * n;
    double *ab_i = buf + 4 * n, *taunu = buf + 5 * n, *snu = buf + 6 * n, *hnu = buf + 7 * n;
    double *jnu = buf + 8 * n, *tmp = buf + 9 * n;
    for (long j = 0; j < n; ++j) {
        double ab = acont[j] + aline[j] + sigmac[j] + sigmal[j];
        abtot[j] = ab > EPS38 ? ab : (ab != ab ? ab : EPS38);
        double sc = sigmac[j] + sigmal[j];
        double al = abtot[j] > 0.0 ? sc / abtot[j] : 0.0;
        if (al < 0.0) al = 0.0;
        if (al > 1.0) al = 1.0;
        alpha[j] = al;
        double dn = acont[j] + aline[j];
        snubar[j] = dn > 0 ? (acont[j] * scont[j] + aline[j] * sline[j]) / dn : scont[j];
    }
    int rev = n > 1 && rhox[0] > rhox[n - 1];
    double start;
    if (rev) {
        for (long j = 0; j < n; ++j) { rho[j] = rhox[n - 1 - j]; ab_i[j] = abtot[n - 1 - j]; }
        start = ab_i[n - 1] * rho[n - 1];
    } else {
        for (long j = 0; j < n; ++j) { rho[j] = rhox[j]; ab_i[j] = abtot[j]; }
        start = abtot[0] * rho[0];
    }
    orc_integ(rho, ab_i, n, start, taunu);
    if (rev) {
        for (long j = 0; j < n / 2; ++j) {
            double t = snubar[j]; snubar[j] = snubar[n - 1 - j]; snubar[n - 1 - j] = t;
            t = alpha[j]; alpha[j] = alpha[n - 1 - j]; alpha[n - 1 - j] = t;
        }
    }
    double flux;
    if (taunu[0] > g_xtau[NXTAU - 1]) {
        /* label-401 branch (josh_solver.py:700-727) */
        for (long j = 0; j < n; ++j) snu[j] = snubar[j];
        int it = 0;
        const int max401 = g_josh_fp64 ? 300 : NXTAU;
        const double tol401 = g_josh_fp64 ? g_josh_tol * 10.0 : 1.0e-5;
        for (int l = 0; l < max401; ++l) {
            orc_deriv(taunu, snu, n, hnu);
            for (long j = 0; j < n; ++j) hnu[j] = hnu[j] / 3.0;
            orc_deriv(taunu, hnu, n, jnu);
            for (long j = 0; j < n; ++j) {
                double jn = jnu[j] + snu[j];
                double sn = (1.0 - alpha[j]) * snubar[j] + alpha[j] * jn;
                double den = fabs(sn) > EPS38 ? fabs(sn) : EPS38;
                tmp[j] = fabs(sn - snu[j]) / den;
                snu[j] = sn;
            }
            double err = np_pairwise_sum(tmp, n);
            it = l + 1;
            if (err < tol401) break;
        }
        flux = hnu[0];
        if (info) { info[0] = it; info[1] = 1; }
    } else {
        double xsbar[NXTAU], xalpha[NXTAU];
        orc_map1_josh(taunu, snubar, (int)n, g_xtau, xsbar, NXTAU);
        orc_map1_josh(taunu, alpha, (int)n, g_xtau, xalpha, NXTAU);
        float xs32[NXTAU], xal32[NXTAU], xsb32[NXTAU];
        double s0 = snubar[0] > EPS38 ? snubar[0] : EPS38;
        double a0 = alpha[0] < 0.0 ? 0.0 : (alpha[0] > 1.0 ? 1.0 : alpha[0]);
        for (int k = 0; k < NXTAU; ++k) {
            double xb = xsbar[k] > EPS38 ? xsbar[k] : EPS38;
            double xa = xalpha[k] < 0.0 ? 0.0 : (xalpha[k] > 1.0 ? 1.0 : xalpha[k]);
            if (g_xtau[k] < taunu[0]) { xb = s0; xa = a0; }
            xs32[k] = (float)xb;
            xal32[k] = (float)xa;
            xsb32[k] = (float)(xb * (1.0 - xa));
        }
        int it;
        flux = 0.0;
        if (g_josh_fp64) {
            double xa64[NXTAU], xb64[NXTAU], xs64[NXTAU];
            for (int k = 0; k < NXTAU; ++k) {
                double xb = xsbar[k] > EPS38 ? xsbar[k] : EPS38;
                double xa = xalpha[k] < 0.0 ? 0.0 : (xalpha[k] > 1.0 ? 1.0 : xalpha[k]);
                if (g_xtau[k] < taunu[0]) { xb = s0; xa = a0; }
                xa64[k] = xa;
                xb64[k] = xb;
            }
            it = josh_iterate64(xs64, xa64, xb64);
            for (int k = 0; k < NXTAU; ++k) flux += g_ch[k] * xs64[k];
        } else {
            it = orc_josh_iterate(xs32, xal32, xsb32, NXTAU);
            for (int k = 0; k < NXTAU; ++k) flux += g_ch[k] * (double)xs32[k];
        }
        if (info) { info[0] = it; info[1] = 0; }
    }
    free(buf);
    return flux;
}

/* engine/radiative.py:20-39 (_planck_nu) for one frequency over depths. */
static void planck_nu(double freq, const double *temp, long n, double *out) {
    double freq15 = freq / 1.0e15;
    double f3 = pow(freq15, 3.0);
    for (long j = 0; j < n; ++j) {
        double hkt = H_PLANCK / (K_BOLTZ * temp[j]);
        double ehvkt = exp(-freq * hkt);
        double stim = 1.0 - ehvkt;
        double b = 1.47439e-2 * f3 * ehvkt / stim;
        out[j] = b != b ? 0.0 : b;
    }
}

/* engine/radiative.py:42-218 (solve_lte_frequency) for wavelength index w of
 * [D][W] arrays.  line_source may be NULL.  Returns flux_total; *fc gets the
 * continuum flux; info2 (may be NULL) gets iterations of both solves. */
double orc_lte_frequency(long D, long W, long w, double wavelength, const double *temp,
                         const double *rhox, const double *cont_abs, const double *cont_scat,
                         const double *line_abs, const double *line_scat,
                         const double *line_source, double *fc, int *info2) {
    double *b = (double *)malloc(sizeof(double) * 10 * D);
    double *m = b, *t = b + D, *ca = b + 2 * D, *cs = b + 3 * D, *la = b + 4 * D, *ls = b + 5 * D;
    double *src = b + 6 * D, *pl = b + 7 * D, *zero = b + 8 * D, *plf = b + 9 * D;
    const double MAXR4 = 3.4e38;
    long n = 0;
    double freq = C_LIGHT_NM / (wavelength > 1e-12 ? wavelength : 1e-12);
    if (line_source) planck_nu(freq, temp, D, plf);
    for (long d = 0; d < D; ++d) {
        double r = rhox[d], a1 = cont_abs[d * W + w], a2 = cont_scat[d * W + w];
        double a3 = line_abs[d * W + w], a4 = line_scat[d * W + w];
        int ok = r >= 0.0 && isfinite(r) && isfinite(a1) && isfinite(a2) && isfinite(a3) &&
                 isfinite(a4) && a1 < MAXR4 && a2 < MAXR4 && a3 < MAXR4 && a4 < MAXR4;
        if (!ok || !(r > 0)) continue;
        m[n] = r; t[n] = temp[d]; ca[n] = a1; cs[n] = a2; la[n] = a3; ls[n] = a4;
        if (line_source) {
            double s = line_source[d * W + w];
            src[n] = isfinite(s) ? s : plf[d];
        }
        zero[n] = 0.0;
        ++n;
    }
    double ft = 0.0;
    *fc = 0.0;
    if (n > 0) {
        if (n > 1 && !(m[0] < m[n - 1])) {
            for (long j = 0; j < n / 2; ++j) {
                double *arrs[6] = {m, t, ca, cs, la, ls};
                for (int q = 0; q < 6; ++q) { double x = arrs[q][j]; arrs[q][j] = arrs[q][n - 1 - j]; arrs[q][n - 1 - j] = x; }
                if (line_source) { double x = src[j]; src[j] = src[n - 1 - j]; src[n - 1 - j] = x; }
            }
        }
        planck_nu(freq, t, n, pl);
        int i1[2] = {0, 0}, i2[2] = {0, 0};
        *fc = orc_solve_josh_flux(n, ca, pl, zero, pl, cs, zero, m, i1);
        ft = orc_solve_josh_flux(n, ca, pl, la, line_source ? src : pl, cs, ls, m, i2);
        if (info2) { info2[0] = i1[0]; info2[1] = i2[0]; info2[2] = i1[1] + 2 * i2[1]; }
    }
    free(b);
    return ft;
}

/* engine/radiative.py:282-494 (solve_lte_spectrum), wavelength-parallel.
 * iters (may be NULL) gets the summed JOSH iterations of both solves; iters_w
 * (may be NULL) per wavelength, +1000 when the total solve took label 401. */
void orc_lte_spectrum(long D, long W, const double *wl, const double *temp, const double *rhox,
                      const double *cont_abs, const double *cont_scat, const double *line_abs,
                      const double *line_scat, const double *line_source, double *flux_total,
                      double *flux_cont, double *iters, int32_t *iters_w) {
    double itsum = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : itsum)
    for (long w = 0; w < W; ++w) {
        int info2[3] = {0, 0, 0};
        double fc;
        flux_total[w] = orc_lte_frequency(D, W, w, wl[w], temp, rhox, cont_abs, cont_scat,
                                          line_abs, line_scat, line_source, &fc, info2);
        flux_cont[w] = fc;
        itsum += info2[0] + info2[1];
        if (iters_w) iters_w[w] = info2[0] + info2[1] + (info2[2] >= 2 ? 1000 : 0);
    }
    if (iters) *iters = itsum;
}

/* engine/opacity.py:3852-3861: F_lambda = F_nu * c / lambda^2. */
void orc_flux_to_lambda(long W, const double *wl, double *flux) {
    for (long w = 0; w < W; ++w) {
        double w2 = wl[w] * wl[w];
        flux[w] = flux[w] * (C_LIGHT_NM / (w2 > 1e-40 ? w2 : 1e-40));
    }
}

/* fort.19 lines in ASYNTH: C restatement for the CPU oracle (TEST
 * INFRASTRUCTURE ONLY -- the checker of synthe_fort19_lines, never part of
 * the product path).
 *
 * engine/opacity.py:1335-1546 (_add_fort19_asynth), one depth at a time: a
 * zeroed buffer takes every record's terms in fort.19 order, then
 * ASYNTH += buffer * STIM when any entry is positive.  The record kinds:
 *   normal (TYPE 0 with NBLO/NBUP)  opacity.py:1407-1478, with the metal wing
 *                                   loop :74-288 and WCON/WTAIL :646-686
 *   merged continuum (TYPE > 3)     opacity.py:1480-1513, :689-757, :850-892
 *   autoionising (TYPE 1)           opacity.py:1515-1545, :895-963
 * The loops are the reference's own, including the Python index -1 of a
 * buffer entry for a record whose centre lies below the grid.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const double *h0, *h1, *h2;
    int nvoigt;
    const double *extab, *extabf;
    int nfastex;
} f19_tables;

double orc_voigt(double v, double a, const double *h0t, const double *h1t, const double *h2t, int n);

#define F19_MAXPROF 1000000L
static const double C_LIGHT_CM = 2.99792458e10;
static const double NM_TO_CM = 1e-7;

/* engine/opacity.py:646-686 with ifvac = 1; -1 for None */
static void continuum_limits(int ncon, int nelion, int nelionx, double emerge, double emerge_h,
                             const double *contx, double *wcon_out, double *wtail_out) {
    *wcon_out = *wtail_out = -1.0;
    if (ncon <= 0 || nelionx <= 0) return;
    if (nelionx > 17 || ncon > 26) return;
    double cont_val = contx[(ncon - 1) * 17 + (nelionx - 1)];
    if (cont_val <= 0.0) return;
    double emerge_line = nelion == 1 ? emerge_h : emerge;
    double denom = cont_val - emerge_line;
    if (fabs(denom) <= 1e-8) return;
    double wcon = 1.0e7 / denom;
    double denom_tail = cont_val - emerge_line - 500.0;
    double wtail = -1.0;
    if (fabs(denom_tail) > 1e-8) {
        wtail = 1.0e7 / denom_tail;
        if (wtail < 0.0) wtail = 2.0 * wcon;
        if (2.0 * wcon < wtail) wtail = 2.0 * wcon;
    }
    if (wtail > 0.0 && wtail <= wcon) wtail = -1.0;
    *wcon_out = wcon > 0.0 ? wcon : -1.0;
    *wtail_out = wtail > 0.0 ? wtail : -1.0;
}

/* engine/opacity.py:74-288 (_accumulate_metal_profile_kernel) */
static void metal_profile(double *buf, const double *cont, const double *grid, long W, long c, double lw,
                          double kappa0, double damping, double dw, double cutoff, double wcon, double wtail,
                          const f19_tables *tb, double *profile) {
    if (dw <= 0.0 || kappa0 <= 0.0) return;
    double adamp = damping > 1e-12 ? damping : 1e-12;
    long cc = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
    double kapmin = cutoff * cont[cc];
    double resolu;
    if (cc < W - 1) {
        resolu = 1.0 / (grid[cc + 1] / grid[cc] - 1.0);
    } else if (cc > 0) {
        resolu = 1.0 / (grid[cc] / grid[cc - 1] - 1.0);
    } else {
        resolu = 300000.0;
    }
    double dopple = lw > 0 ? dw / lw : 1e-6;
    long n10 = (long)(10.0 * dopple * resolu);
    if (n10 > F19_MAXPROF) n10 = F19_MAXPROF;
    int early = 0;
    long nfinal = 0;
    if (adamp < 0.2) {
        double tabstep = dopple * resolu > 0 ? 200.0 / (dopple * resolu) : 200.0;
        double tabi = 0.5;
        for (long k = 1; k <= n10; ++k) {
            tabi = tabi + tabstep;
            long it = (long)tabi;
            if (it < 0) it = 0;
            if (it > tb->nvoigt - 1) it = tb->nvoigt - 1;
            profile[k] = kappa0 * (tb->h0[it] + adamp * tb->h1[it]);
            if (profile[k] < kapmin) {
                nfinal = k;
                early = 1;
                break;
            }
        }
    } else {
        double dvoigt = dopple * resolu > 0 ? 1.0 / dopple / resolu : 1e-6;
        for (long k = 1; k <= n10; ++k) {
            double x = (double)k * dvoigt;
            profile[k] = kappa0 * orc_voigt(x, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
            if (profile[k] < kapmin) {
                nfinal = k;
                early = 1;
                break;
            }
        }
    }
    if (!early) {
        double xf = (n10 > 0 && profile[n10] > 0) ? profile[n10] * ((double)n10 * (double)n10) : 0.0;
        long maxstep;
        if (xf > 0 && kapmin > 0) {
            double ms = sqrt(xf / kapmin) + 1.0;
            maxstep = ms < 2.0e6 ? (long)ms : F19_MAXPROF;
            if (maxstep > F19_MAXPROF) maxstep = F19_MAXPROF;
        } else {
            maxstep = n10;
        }
        for (long k = n10 + 1; k <= maxstep; ++k) profile[k] = xf / ((double)k * (double)k);
        nfinal = maxstep;
    }
    if (c + nfinal < 0 || c - nfinal >= W) return;
    int use_wcon = wcon > 0.0, use_wtail = wtail > 0.0;
    if (c < W - 1) {
        long maxred = c >= 0 ? (W - 1 - c < nfinal ? W - 1 - c : nfinal) : (W - 1 < nfinal ? W - 1 : nfinal);
        long minred = -c > 1 ? -c : 1;
        for (long k = minred; k <= maxred; ++k) {
            long idx = c + k;
            if (idx < 0 || idx >= W) continue;
            if (use_wcon && grid[idx] <= wcon) continue;
            double v = profile[k];
            if (use_wtail) {
                double wave = grid[idx];
                double base = use_wcon ? wcon : lw;
                if (wave < wtail) {
                    double span = wtail - base;
                    v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
                }
            }
            buf[idx] += v;
        }
    }
    if (c <= 0) return;
    long maxblue = c < nfinal ? c : nfinal;
    long minblue = c + 1 - W > 1 ? c + 1 - W : 1;
    for (long k = minblue; k <= maxblue; ++k) {
        long idx = c - k;
        if (idx < 0 || idx >= W) continue;
        if (use_wcon && grid[idx] <= wcon) break;
        double v = profile[k];
        if (use_wtail) {
            double wave = grid[idx];
            double base = use_wcon ? wcon : lw;
            if (wave < wtail) {
                double span = wtail - base;
                v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
            }
        }
        buf[idx] += v;
    }
}

/* numpy.searchsorted(grid, x, side) */
static long search(const double *grid, long W, double x, int right) {
    long lo = 0, hi = W;
    while (lo < hi) {
        long mid = (lo + hi) / 2;
        if (right ? grid[mid] <= x : grid[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* Python's buffer[i] for -W <= i < W */
#define PYIDX(i, W) ((i) < 0 ? (W) + (i) : (i))

void orc_fort19_asynth(long N, long D, long W, const double *grid, const int32_t *kind, const double *wl,
                       const int64_t *center, const double *gf, const double *gr, const double *gs,
                       const double *gw, const int32_t *elem, const int32_t *ion, const int32_t *nelion,
                       const int32_t *nlast, const int32_t *ncon, const int32_t *nelionx, const double *contx,
                       const double *boltz, const double *doppler, const double *pop, long nion, long nelem,
                       const double *rho_d, const double *xne, const double *txnxn, const double *emerge,
                       const double *emerge_h, const double *cont_all, const double *stim, double cutoff,
                       const f19_tables *tb, double *asynth) {
    double *buf = (double *)malloc(sizeof(double) * (size_t)W);
    double *profile = (double *)calloc((size_t)F19_MAXPROF + 1, sizeof(double));
    for (long d = 0; d < D; ++d) {
        const double *cont = cont_all + d * W;
        memset(buf, 0, sizeof(double) * (size_t)W);
        for (long f = 0; f < N; ++f) {
            double pop_val = pop[(d * nion + (ion[f] - 1)) * nelem + elem[f]];
            if (pop_val <= 0.0) continue;
            double bz = boltz[f * D + d];
            double lw = wl[f];
            long c = center[f];
            double rho = rho_d[d];
            if (kind[f] == 0) {
                if (rho <= 0.0) continue;
                double dw = doppler[f * D + d];
                if (lw <= 0.0 || dw <= 0.0) continue;
                double dopple = dw / lw;
                if (dopple <= 0.0) continue;
                double xnfdop = pop_val / (rho * dopple);
                double kappa0_pre = gf[f] * xnfdop;
                long cc = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
                double kapmin = cont[cc] * cutoff;
                if (kappa0_pre < kapmin) continue;
                double kappa0 = kappa0_pre * bz;
                if (kappa0 < kapmin) continue;
                double gamma_total = gr[f] + gs[f] * xne[d] + gw[f] * txnxn[d];
                double adamp = dopple > 0.0 ? gamma_total / dopple : 0.0;
                double kapcen;
                if (adamp < 0.2) kapcen = kappa0 * (1.0 - 1.128 * adamp);
                else kapcen = kappa0 * orc_voigt(0.0, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
                if (kapcen >= kapmin) buf[cc] += kapcen;
                double wcon = -1.0, wtail = -1.0;
                if (ncon[f] > 0 && nelionx[f] > 0)
                    continuum_limits(ncon[f], nelion[f], nelionx[f], emerge[d], emerge_h[d], contx, &wcon, &wtail);
                metal_profile(buf, cont, grid, W, c, lw, kappa0, adamp, dw, cutoff, wcon, wtail, tb, profile);
            } else if (kind[f] == 2) {
                if (rho <= 0.0) continue;
                double kappa = gf[f] * (pop_val / rho) * bz;
                /* opacity.py:689-727 */
                double wmerge = lw, wtail = lw;
                int nl = nlast[f];
                if (lw > 0.0 && nl > 0) {
                    double ryd = nelion[f] == 1 ? 109677.576 : 109737.312;
                    double ds = 1.0e7 / lw - ryd / (double)(nl * nl);
                    double eline = nelion[f] == 1 ? emerge_h[d] : emerge[d];
                    double dm = 1.0e7 / lw - eline;
                    if (fabs(ds) > 1e-12 && fabs(dm) > 1e-12) {
                        double wshift = 1.0e7 / ds;
                        double wm = 1.0e7 / dm;
                        if (wm < 0.0) wm = wshift + wshift;
                        if (wshift > wm) wm = wshift;
                        if (wshift + wshift < wm) wm = wshift + wshift;
                        wmerge = wm;
                        double dt = 1.0e7 / wm - 500.0;
                        if (fabs(dt) > 1e-12) {
                            double wt = 1.0e7 / dt;
                            if (wt < 0.0) wt = wm + wm;
                            wtail = wm + wm < wt ? wm + wm : wt;
                        }
                    }
                }
                /* opacity.py:850-892 */
                if (kappa <= 0.0) continue;
                long start = c > 0 ? c : 0;
                long imerge = search(grid, W, wmerge, 0);
                long itail = search(grid, W, wtail, 1);
                if (itail > W) itail = W;
                if (itail <= start) continue;
                long m0 = imerge > start ? imerge : start;
                long denom = itail - m0 > 1 ? itail - m0 : 1;
                for (long i = start; i < itail; ++i) {
                    double wave = grid[i];
                    if (wave < lw) continue;
                    double v = kappa;
                    if (i >= imerge) v *= (double)(itail - i > 0 ? itail - i : 0) / (double)denom;
                    if (v < cont[i] * cutoff) break;
                    buf[i] += v;
                }
            } else if (kind[f] == 1) {
                if (rho <= 0.0) continue;
                double kappa0 = gw[f] * gf[f] * (pop_val / rho) * bz;
                /* opacity.py:895-963 */
                double ccut = cont[PYIDX(c, W)] * cutoff;
                if (kappa0 < ccut || kappa0 <= 0.0) continue;
                double fc = C_LIGHT_CM / (lw * NM_TO_CM);
                double gamma = fabs(gr[f]) > 1e-30 ? fabs(gr[f]) : 1e-30;
                double ashore = gs[f], bshore = gw[f];
                if (fabs(bshore) < 1e-30) bshore = 1e-30;
                buf[PYIDX(c, W)] += kappa0;
                int red = 1, blue = 1;
                long max_steps = c > W - c - 1 ? c : W - c - 1;
                for (long o = 1; o <= max_steps && (red || blue); ++o) {
                    if (red) {
                        long idx = c + o;
                        if (idx >= W) {
                            red = 0;
                        } else {
                            double fr = C_LIGHT_CM / (grid[idx] * NM_TO_CM);
                            double eps = 2.0 * (fr - fc) / gamma;
                            double v = kappa0 * (ashore * eps + bshore) / (eps * eps + 1.0) / bshore;
                            if (v <= 0.0 || v < cont[idx] * cutoff) red = 0;
                            else buf[idx] += v;
                        }
                    }
                    if (blue) {
                        long idx = c - o;
                        if (idx < 0) {
                            blue = 0;
                        } else {
                            double fr = C_LIGHT_CM / (grid[idx] * NM_TO_CM);
                            double eps = 2.0 * (fr - fc) / gamma;
                            double v = kappa0 * (ashore * eps + bshore) / (eps * eps + 1.0) / bshore;
                            if (v <= 0.0 || v < cont[idx] * cutoff) blue = 0;
                            else buf[idx] += v;
                        }
                    }
                }
            }
        }
        int any = 0;
        for (long w = 0; w < W; ++w) any |= buf[w] > 0.0;
        if (any)
            for (long w = 0; w < W; ++w) asynth[d * W + w] += buf[w] * stim[d * W + w];
    }
    free(profile);
    free(buf);
}

/* The engine's helium pass (engine/opacity.py:2534-2546, 2686-3066 with the
 * TYPE < -2 branch of _apply_fort19_profile, :998-1053), one depth at a
 * time: every helium line in catalog order adds its Voigt wings to
 * HELIUM_WINGS; then line opacity += HELIUM_WINGS * STIM (:3651-3659).
 * Per line: cgf (CGF of the catalog gf), its TYPE (-4: 3He), the matched
 * fort.19 record's NCON / NELIONX (0 without one), the ion stage; per (line,
 * depth): the population of that stage, its DOPPLE and FASTEX(E HCKT). */
void orc_helium_wings(long N, long D, long W, const double *grid, const double *wl, const int64_t *center,
                      const double *cgf, const double *gr, const double *gs, const double *gw,
                      const int32_t *ltype, const int32_t *ion, const int32_t *ncon, const int32_t *nelionx,
                      const double *contx, const double *pop_nd, const double *dopple_nd,
                      const double *boltz_nd, const double *rho_d, const double *xne, const double *txnxn,
                      const double *emerge, const double *emerge_h, const double *cont_all, const double *stim,
                      double cutoff, const f19_tables *tb, double *out) {
    double *wings = (double *)malloc(sizeof(double) * (size_t)W);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)W);
    for (long d = 0; d < D; ++d) {
        const double *cont = cont_all + d * W;
        memset(wings, 0, sizeof(double) * (size_t)W);
        for (long l = 0; l < N; ++l) {
            long c = center[l];
            if (c < 0 || c >= W) continue;
            double pop_val = pop_nd[l * D + d], dop_val = dopple_nd[l * D + d];
            if (pop_val <= 0.0 || dop_val <= 0.0) continue;
            double rho = rho_d[d];
            if (!(rho > 0.0)) continue;
            double lw = wl[l];
            double xnfdop = pop_val / (rho * dop_val);
            double dw = dop_val * lw;
            double kmin = cont[c] * cutoff;
            double k0pre = cgf[l] * xnfdop;
            if (k0pre < kmin || dw <= 0.0) continue;
            double kappa0 = k0pre * boltz_nd[l * D + d];
            if (kappa0 < kmin) continue;
            double wcon = -1.0, wtail = -1.0;
            continuum_limits(ncon[l], ion[l], nelionx[l], emerge[d], emerge_h[d], contx, &wcon, &wtail);
            /* _apply_fort19_profile, TYPE < -2 */
            memset(tmp, 0, sizeof(double) * (size_t)W);
            double doppler = dw > 1e-12 ? dw : 1e-12;
            double keff = kappa0;
            if (ltype[l] == -4) {
                keff /= 1.155;
                doppler *= 1.155;
            }
            double dl = doppler / lw;
            double damping = (gr[l] + gs[l] * xne[d] + gw[l] * txnxn[d]) / (dl > 1e-40 ? dl : 1e-40);
            double adamp = damping > 1e-12 ? damping : 1e-12;
            long cc = c;
            double base = wcon > 0.0 ? wcon : lw;
            if (lw <= grid[W - 1]) {
                for (long i = cc; i < W; ++i) {
                    double wave = grid[i];
                    if (wcon > 0.0 && wave <= wcon) continue;
                    double x = fabs(wave - lw) / doppler;
                    double v = keff * orc_voigt(x, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
                    if (wtail > 0.0 && wave < wtail) {
                        double span = wtail - base;
                        v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
                    }
                    if (v < cont[i] * cutoff) break;
                    tmp[i] += v;
                }
            }
            if (cc > 0 && lw >= grid[0]) {
                for (long i = cc - 1; i >= 0; --i) {
                    double wave = grid[i];
                    if (wcon > 0.0 && wave <= wcon) break;
                    double x = fabs(wave - lw) / doppler;
                    double v = keff * orc_voigt(x, adamp, tb->h0, tb->h1, tb->h2, tb->nvoigt);
                    if (wtail > 0.0 && wave < wtail) {
                        double span = wtail - base;
                        v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
                    }
                    tmp[i] += v;
                    if (v < cont[i] * cutoff) break;
                }
            }
            for (long i = 0; i < W; ++i) wings[i] += tmp[i];
        }
        for (long i = 0; i < W; ++i) out[d * W + i] += wings[i] * stim[d * W + i];
    }
    free(tmp);
    free(wings);
}

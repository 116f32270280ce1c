/* CPU oracle for the hydrogen-line opacity (HPROF4 profiles), test
 * infrastructure only: a plain FP64 C restatement of
 *   physics/profiles/hydrogen.py:1408-1841 (VCSE1F, SOFBETA, HFNM, fine
 *       structure, hydrogen_line_profile, the Lyman-alpha cutoffs),
 *   physics/populations.py:96-128 (the per-depth hydrogen state),
 *   engine/opacity.py:1547-1704 (_accumulate_hydrogen_profile) and
 *   engine/opacity.py:1706-1884 (_compute_hydrogen_line_opacity),
 * one statement per reference statement and in the reference's order, with
 * C's pow/log/sqrt (the functions Python's float operators call).  Only the
 * tests, the smoke check and the benchmark's CPU baseline load it. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

double orc_fastex(double x, const double *extab, const double *extabf, int n);

static const double RYDH = 3.2880515e15;
static const double C_ANG = 2.99792458e18; /* hydrogen.py C_LIGHT (Angstrom/s) */
static const double C_CM = 2.99792458e10;
static const double C_NM = 2.99792458e17;
static const double LYA_WN = 82259.10;
static const double PI = 3.14159265358979323846;
static const double HYD_RYD_CM = 109677.576;
static const double HYD_EINF_CM = 109678.764;
static const double CGF_CONSTANT = 0.026538 / 1.77245;
static const double KBOLTZ = 1.380649e-16;
static const double AMU = 1.66054e-24;
static const double C_KMS = 299792.458;

/* tables (kurucz_tables.npz h_*, conth, ehyd_cm), row-major as numpy holds them */
static const double *t_propbm;  /* [7][5][15] */
static const double *t_c, *t_d; /* [5][7] */
static const double *t_pp;      /* [5] */
static const double *t_beta;    /* [15] */
static const double *t_stalph, *t_stwtal; /* [34] */
static const double *t_istal, *t_lnghal, *t_lncomp; /* [4] */
static const double *t_stcomp, *t_stcpwt; /* [5][4] */
static const double *t_cut_h2p, *t_cut_h2; /* [111], [91] */
static int n_cut_h2p, n_cut_h2;
static const double *t_asum_lyman, *t_asum; /* [100], [96] */
static int n_asum_lyman, n_asum;
static const double *t_y1wtm;  /* [2][2] */
static const double *t_xknmtb; /* [4][3] */
static const double *t_conth;  /* [15] */
static int n_conth;
static const double *t_ehyd;   /* [8] */
static const double *t_extab, *t_extabf;
static int n_fastex;

void orc_set_hydrogen_tables(const double *propbm, const double *c, const double *d, const double *pp,
                             const double *beta, const double *stalph, const double *stwtal,
                             const double *istal, const double *lnghal, const double *stcomp,
                             const double *stcpwt, const double *lncomp, const double *cut_h2p, int ncut_h2p,
                             const double *cut_h2, int ncut_h2, const double *asum_lyman, int nasum_lyman,
                             const double *asum, int nasum, const double *y1wtm, const double *xknmtb,
                             const double *conth, int nconth, const double *ehyd, const double *extab,
                             const double *extabf, int nfastex) {
    t_propbm = propbm; t_c = c; t_d = d; t_pp = pp; t_beta = beta; t_stalph = stalph; t_stwtal = stwtal;
    t_istal = istal; t_lnghal = lnghal; t_stcomp = stcomp; t_stcpwt = stcpwt; t_lncomp = lncomp;
    t_cut_h2p = cut_h2p; n_cut_h2p = ncut_h2p; t_cut_h2 = cut_h2; n_cut_h2 = ncut_h2;
    t_asum_lyman = asum_lyman; n_asum_lyman = nasum_lyman; t_asum = asum; n_asum = nasum;
    t_y1wtm = y1wtm; t_xknmtb = xknmtb; t_conth = conth; n_conth = nconth; t_ehyd = ehyd;
    t_extab = extab; t_extabf = extabf; n_fastex = nfastex;
}

static double fex(double x) { return orc_fastex(x, t_extab, t_extabf, n_fastex); }

/* hydrogen.py:1408-1429 */
double orc_vcse1f(double x) {
    if (x <= 0) return 0.0;
    if (x <= 0.01) return -log(x) - 0.577215 + x;
    if (x <= 1.0)
        return -log(x) - 0.57721566 +
               x * (0.99999193 + x * (-0.24991055 + x * (0.05519968 + x * (-0.00976004 + x * 0.00107857))));
    if (x > 30.0) return 0.0;
    double numerator = x * (x + 2.334733) + 0.25062;
    double denominator = (x * (x + 3.330657) + 1.681534) * x;
    return numerator / denominator * fex(x);
}

/* numpy searchsorted(a, v, side="right") */
static long search_right(const double *a, long n, double v) {
    long lo = 0, hi = n;
    while (lo < hi) {
        long mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* hydrogen.py:1451-1496 */
double orc_sofbeta(double beta, double p, int n, int m) {
    double corr = 1.0;
    double b2 = beta * beta;
    double sb = sqrt(beta);
    if (beta <= 500.0) {
        int indx = 7;
        int mmn = m - n;
        if (n <= 3 && mmn <= 2) indx = 2 * (n - 1) + mmn;
        indx = indx < 1 ? 1 : (indx > 7 ? 7 : indx);
        int im = (int)(5.0 * p) + 1; /* Python int(): truncation */
        im = im < 4 ? im : 4;
        im = im > 1 ? im : 1;
        int ip = im + 1;
        double wtp = 5.0 * (p - t_pp[im - 1]);
        wtp = wtp > 1.0 ? 1.0 : wtp;
        wtp = wtp < 0.0 ? 0.0 : wtp;
        double wtm = 1.0 - wtp;
        if (beta <= 25.12) {
            long j = search_right(t_beta, 15, beta);
            j = j < 1 ? 1 : (j > 14 ? 14 : j);
            long jm = j - 1, jp = j;
            double denom = t_beta[jp] - t_beta[jm];
            double wtb = denom <= 0 ? 0.0 : (beta - t_beta[jm]) / denom;
            double wtbm = 1.0 - wtb;
            const double *prop = t_propbm + (long)(indx - 1) * 75;
            double cbp = prop[(ip - 1) * 15 + jp] * wtp + prop[(im - 1) * 15 + jp] * wtm;
            double cbm = prop[(ip - 1) * 15 + jm] * wtp + prop[(im - 1) * 15 + jm] * wtm;
            corr = 1.0 + cbp * wtb + cbm * wtbm;
            double pr1 = 0.0, pr2 = 0.0;
            double wt = 0.5 * (10.0 - beta);
            wt = wt < 1.0 ? wt : 1.0;
            wt = wt > 0.0 ? wt : 0.0;
            if (beta <= 10.0) pr1 = 8.0 / (83.0 + (2.0 + 0.95 * b2) * beta);
            if (beta >= 8.0) pr2 = (1.5 / sb + 27.0 / b2) / b2;
            return (pr1 * wt + pr2 * (1.0 - wt)) * corr;
        }
        double cc = t_c[(im - 1) * 7 + (indx - 1)] * wtp + t_c[(ip - 1) * 7 + (indx - 1)] * wtm;
        double dd = t_d[(im - 1) * 7 + (indx - 1)] * wtp + t_d[(ip - 1) * 7 + (indx - 1)] * wtm;
        corr = 1.0 + dd / (cc + beta * sb);
    }
    return (1.5 / sb + 27.0 / b2) / b2 * corr;
}

/* hydrogen.py:1499-1513 */
double orc_hf_nm(int n, int m) {
    if (m <= n) return 0.0;
    double xn = (double)n;
    double ginf = 0.2027 / pow(xn, 0.71);
    double gca = 0.124 / xn;
    double fkn = xn * 1.9603;
    double wtc = 0.45 - 2.4 / pow(xn, 3.0) * (xn - 1.0);
    double xm = (double)m;
    double xmn = xm - xn;
    double fk = fkn * pow(xm / (xmn * (xm + xn)), 3.0);
    double xmn12 = pow(xmn, 1.2);
    double wt = (xmn12 - 1.0) / (xmn12 + wtc);
    return fk * (1.0 - wt * ginf - (0.222 + gca / xm) * (1.0 - wt));
}

/* hydrogen.py:1516-1534: fine-structure components (offsets in Hz, weights);
 * returns their number (Python slices clip at the table's end). */
int orc_fine_structure(int n, int m, double *off, double *wt) {
    int mmn = m - n;
    double xn = (double)n, xn2 = xn * xn;
    if (n > 4 || m > 10) {
        off[0] = 0.0;
        wt[0] = 1.0;
        return 1;
    }
    if (mmn != 1) {
        int ifins = (int)t_lncomp[n - 1];
        for (int i = 0; i < ifins; ++i) {
            off[i] = t_stcomp[i * 4 + (n - 1)] * 1.0e7;
            wt[i] = t_stcpwt[i * 4 + (n - 1)] / xn2;
        }
        return ifins;
    }
    int ifins = (int)t_lnghal[n - 1];
    int ipos = (int)t_istal[n - 1];
    int end = ipos + ifins < 34 ? ipos + ifins : 34;
    int k = 0;
    for (int i = ipos; i < end; ++i, ++k) {
        off[k] = t_stalph[i] * 1.0e7;
        wt[k] = t_stwtal[i] / xn2 / 3.0;
    }
    return k;
}

/* Hydrogen state of one depth (populations.py:18-29), flattened:
 * pp fo y1b y1s t3nhe t3nh2 c1d c2d gcon1 gcon2 xnfph0 xnfph1 dopph xne (the
 * depth's electron density as given) has (1 when the state exists). */
enum { HS_PP, HS_FO, HS_Y1B, HS_Y1S, HS_T3NHE, HS_T3NH2, HS_C1D, HS_C2D, HS_GCON1, HS_GCON2, HS_XNFPH0,
       HS_XNFPH1, HS_DOPPH, HS_XNE, HS_HAS, HS_N };

/* populations.py:96-128 (_hydrogen_state) and :210-231 (the dopph default and
 * the condition for a state).  has_he1_h2: the atmosphere carries xnf_he1 or
 * xnf_h2; xnfph [D][2] and dopph [D] may be NULL (absent). */
void orc_hydrogen_states(long D, const double *temperature, const double *electron_density,
                         const double *xnf_he1, const double *xnf_h2, const double *xnfph, const double *dopph,
                         const double *turbulent_velocity, double microturb_kms, int has_he1_h2, double *out) {
    for (long i = 0; i < D; ++i) {
        double *o = out + i * HS_N;
        memset(o, 0, sizeof(double) * HS_N);
        if (!has_he1_h2) continue;
        double temp = temperature[i] > 1.0 ? temperature[i] : 1.0;
        double he1 = xnf_he1 ? xnf_he1[i] : 0.0;
        double h2 = xnf_h2 ? xnf_h2[i] : 0.0;
        double ph0 = xnfph ? xnfph[2 * i] : 0.0, ph1 = xnfph ? xnfph[2 * i + 1] : 0.0;
        double dph;
        if (dopph) dph = dopph[i];
        else {
            double vt = turbulent_velocity ? turbulent_velocity[i] / 1e5 / C_KMS : 0.0;
            double micro = microturb_kms > 0 ? microturb_kms / C_KMS : 0.0;
            double total_turb = sqrt(vt * vt + micro * micro);
            double thermal = sqrt(2.0 * KBOLTZ * temp / (1.008 * AMU)) / C_CM;
            dph = sqrt(thermal * thermal + total_turb * total_turb);
        }
        double xne = electron_density[i] > 1e-40 ? electron_density[i] : 1e-40;
        double xne16 = pow(xne, 1.0 / 6.0);
        double pp = xne16 * 0.08989 / sqrt(temp > 1.0 ? temp : 1.0);
        double fo = pow(xne16, 4.0) * 1.25e-9;
        double tt = temp > 1.0 ? temp : 1.0;
        double y1b = 2.0 / (1.0 + 0.012 / tt * sqrt(xne / tt));
        double t4 = temp / 10000.0;
        double t43 = pow(t4, 0.3);
        double y1s = t43 / xne16;
        double c1d = fo * 78940.0 / tt;
        double c2d = pow(fo, 2.0) / 5.96e-23 / xne;
        double gcon1 = 0.2 + 0.09 * sqrt(t4 > 1e-12 ? t4 : 1e-12) / (1.0 + xne / 1.0e13);
        double gcon2 = 0.2 / (1.0 + xne / 1.0e15);
        o[HS_PP] = pp; o[HS_FO] = fo; o[HS_Y1B] = y1b; o[HS_Y1S] = y1s;
        o[HS_T3NHE] = t43 * he1; o[HS_T3NH2] = t43 * h2; o[HS_C1D] = c1d; o[HS_C2D] = c2d;
        o[HS_GCON1] = gcon1; o[HS_GCON2] = gcon2; o[HS_XNFPH0] = ph0; o[HS_XNFPH1] = ph1;
        o[HS_DOPPH] = dph; o[HS_XNE] = electron_density[i]; o[HS_HAS] = 1.0;
    }
}

/* hydrogen.py:1712-1728; returns 0 for None. */
static int interp_cutoff(double dwn, const double *table, int size, double start, double step, double *out) {
    double max_delta = start + step * (double)(size - 1);
    if (dwn > max_delta) return 0;
    if (dwn <= start) {
        if (size < 2) {
            if (size) { *out = table[0]; return 1; }
            return 0;
        }
        double frac = (dwn - start) / step;
        *out = table[0] + (table[1] - table[0]) * frac;
        return 1;
    }
    double position = (dwn - start) / step;
    long index = (long)floor(position);
    double frac = position - (double)index;
    if (index >= size - 1) { *out = table[size - 1]; return 1; }
    *out = table[index] + (table[index + 1] - table[index]) * frac;
    return 1;
}

/* hydrogen.py:1731-1797 */
static double lyman_alpha_lorentz(double freq, double freqnm, double del_freq, double dop, double hwres,
                                  double hwvdw, double hwrad, const double *h) {
    if (dop <= 0.0) return 0.0;
    double hwres_near = hwres * 4.0;
    double hwlor_near = hwres_near + hwvdw + hwrad;
    double hhw_near = freqnm * (hwlor_near > 0.0 ? hwlor_near : 0.0);
    double freq_threshold = (LYA_WN - 4000.0) * C_CM;
    double wavenumber = freq / C_CM;
    double delta_wn = wavenumber - LYA_WN;
    double hres_term, hhw_use;
    if (freq > freq_threshold && hhw_near > 0.0) {
        hres_term = hwres_near * freqnm / PI / (del_freq * del_freq + hhw_near * hhw_near) * 1.77245 * dop;
        hhw_use = hhw_near;
    } else {
        double cutoff_val = 0.0, cl;
        if (interp_cutoff(delta_wn, t_cut_h2, n_cut_h2, -22000.0, 200.0, &cl))
            cutoff_val = pow(10.0, cl - 14.0) * h[HS_XNFPH0] * 2.0 / C_CM;
        hres_term = cutoff_val * 1.77245 * dop;
        double hwlor = hwres + hwvdw + hwrad;
        hhw_use = freqnm * (hwlor > 0.0 ? hwlor : 0.0);
    }
    double hrad_term = 0.0;
    if (hwrad > 0.0 && hhw_use > 0.0) {
        double freq_low = 2.4190611e15, freq_high = 0.77 * RYDH;
        if (freq > freq_low && freq < freq_high)
            hrad_term = hwrad * freqnm / PI / (del_freq * del_freq + hhw_use * hhw_use) * 1.77245 * dop;
    }
    double hvdw_term = 0.0;
    if (hwvdw > 0.0 && hhw_use > 0.0 && freq >= 1.8e15)
        hvdw_term = hwvdw * freqnm / PI / (del_freq * del_freq + hhw_use * hhw_use) * 1.77245 * dop;
    return hres_term + hrad_term + hvdw_term;
}

/* hydrogen.py:1800-1841 */
static double lyman_quasistatic(double freq, double prqs, const double *h, double dbeta, double dop, int n, int m) {
    if (h[HS_FO] <= 0.0) return 0.0;
    double wavenumber = freq / C_CM;
    double delta_wn = wavenumber - LYA_WN;
    if (delta_wn < -20000.0) return 0.0;
    double extra = 0.0;
    double fo = h[HS_FO] > 1e-30 ? h[HS_FO] : 1e-30;
    if (delta_wn <= -4000.0) {
        double cl;
        if (interp_cutoff(delta_wn, t_cut_h2p, n_cut_h2p, -15000.0, 100.0, &cl)) {
            double cutoff_val = pow(10.0, cl - 14.0) * h[HS_XNFPH1] / C_CM;
            extra += cutoff_val * 1.77245 * dop;
        }
    } else {
        double beta4000 = 4000.0 * C_CM / fo * dbeta;
        double prqs4000 = orc_sofbeta(beta4000, h[HS_PP], n, m) * 0.5;
        double normalization = prqs4000 / fo * dbeta;
        double cutoff4000 = pow(10.0, -11.07 - 14.0) * h[HS_XNFPH1] / C_CM;
        if (normalization > 0.0) extra += cutoff4000 / normalization * (prqs / fo * dbeta) * 1.77245 * dop;
    }
    return extra;
}

/* hydrogen.py:1541-1709 (hydrogen_line_profile); h is one depth's state.
 * Returns -1 where the reference would raise (n or m beyond the tables). */
double orc_hydrogen_profile(int n, int m, const double *h, double delta_lambda_nm) {
    if (!h[HS_HAS]) return 0.0;
    int mmn = m - n;
    double xn = (double)n, xm = (double)m;
    double xn2 = xn * xn, xm2 = xm * xm;
    double xm2mn2 = xm2 - xn2;
    double xmn2 = xm2 * xn2;
    double gnm = xm2mn2 / xmn2;
    if (mmn <= 0) return 0.0;
    if (n < 1 || m > n_asum || (n == 1 && m > n_asum_lyman)) return -1.0;
    double xknm;
    if (mmn <= 3 && n <= 4) xknm = t_xknmtb[(n - 1) * 3 + (mmn - 1)];
    else xknm = 5.5e-5 / gnm * xmn2 / (1.0 + 0.13 / (double)mmn);
    double freqnm = RYDH * gnm;
    double wavenm = C_ANG / freqnm;
    double dbeta = C_ANG / (freqnm * freqnm * xknm);
    double c1con = xknm / wavenm * gnm * xm2mn2;
    double c2con = pow(xknm / wavenm, 2.0);
    double radamp = t_asum[n - 1] + t_asum[m - 1];
    if (n == 1) radamp = t_asum_lyman[m - 1];
    radamp /= 12.5664;
    radamp /= freqnm;
    double resont = orc_hf_nm(1, m) / xm / (1.0 - 1.0 / xm2);
    if (n != 1) resont += orc_hf_nm(1, n) / xn / (1.0 - 1.0 / xn2);
    resont *= 3.579e-24 / gnm;
    double vdw = 4.45e-26 / gnm * pow(xm2 * (7.0 * xm2 + 5.0), 0.4);
    double hwvdw = vdw * h[HS_T3NHE] + 2.0 * vdw * h[HS_T3NH2];
    double hwrad = radamp;
    double stark = 1.6678e-18 * freqnm * xknm;
    double hwres = resont * h[HS_XNFPH0] * 2.0;
    double hwstk = stark * h[HS_FO];
    double hwlor = hwres + hwvdw + hwrad;
    double off[40], wts[40];
    int nf = orc_fine_structure(n, m, off, wts);
    double wl0 = wavenm;
    double wl = wl0 + delta_lambda_nm * 10.0;
    double freq = C_ANG / wl;
    double del_freq = fabs(freq - freqnm);
    double dopph = h[HS_DOPPH] > 1e-40 ? h[HS_DOPPH] : 1e-40;
    double dop = freqnm * dopph;
    double wmax = dopph > hwlor ? dopph : hwlor;
    wmax = wmax > hwstk ? wmax : hwstk;
    double hfwidth = freqnm * wmax;
    int ifcore = del_freq <= hfwidth;
    int nwid = 1;
    if (!(dopph >= hwstk && dopph >= hwlor)) {
        nwid = 2;
        if (hwlor < hwstk) nwid = 3;
    }
    double core = 0.0;
    double dop30 = dop > 1e-30 ? dop : 1e-30;
    for (int k = 0; k < nf; ++k) {
        double component_freq = freqnm + off[k];
        double dd = fabs(freq - component_freq) / dop30;
        if (dd <= 7.0) core += fex(dd * dd) * wts[k];
    }
    double lorentz = 0.0;
    double hhw = freqnm * hwlor;
    if (n == 1 && m == 2) {
        lorentz = lyman_alpha_lorentz(freq, freqnm, del_freq, dop, hwres, hwvdw, hwrad, h);
    } else {
        double top = hhw;
        if (n == 1 && (m == 3 || m == 4 || m == 5)) {
            double freq_ratio = freq / RYDH;
            double t2 = hhw - freqnm * hwrad;
            t2 = t2 > 0.0 ? t2 : 0.0;
            if (m == 3 && 0.885 <= freq_ratio && freq_ratio <= 0.890) top = t2;
            else if (m == 4 && 0.936 <= freq_ratio && freq_ratio <= 0.938) top = t2;
            else if (m == 5 && 0.959 <= freq_ratio && freq_ratio <= 0.961) top = t2;
        }
        if (hhw > 0.0) lorentz = top / PI / (del_freq * del_freq + hhw * hhw) * 1.77245 * dop;
    }
    double y1num = 320.0;
    if (m == 2) y1num = 550.0;
    else if (m == 3) y1num = 380.0;
    double y1wht = 1.0e13;
    if (mmn <= 3) y1wht = 1.0e14;
    if (mmn <= 2 && n <= 2) y1wht = t_y1wtm[(n - 1) * 2 + (mmn - 1)];
    double xne0 = h[HS_XNE] > 0.0 ? h[HS_XNE] : 0.0;
    double wty1 = 1.0 / (1.0 + xne0 / (y1wht > 1e-30 ? y1wht : 1e-30));
    double y1_scal = y1num * h[HS_Y1S] * wty1 + h[HS_Y1B] * (1.0 - wty1);
    double c1 = h[HS_C1D] * c1con * y1_scal;
    double c2 = h[HS_C2D] * c2con;
    double fo30 = h[HS_FO] > 1e-30 ? h[HS_FO] : 1e-30;
    double beta = del_freq / fo30 * dbeta;
    double y1 = c1 * beta;
    double y2 = c2 * beta * beta;
    double g1 = 6.77 * sqrt(c1 > 1e-30 ? c1 : 1e-30);
    double ratio = 0.0;
    if (c1 > 0.0 && c2 > 0.0) ratio = sqrt(c2) / (c1 > 1e-30 ? c1 : 1e-30);
    double log_term = 0.0;
    if (ratio > 0.0) log_term = log(ratio > 1e-30 ? ratio : 1e-30);
    double lt = 0.2114 + log_term;
    double gnot = g1 * (lt > 0.0 ? lt : 0.0) * (1.0 - h[HS_GCON1] - h[HS_GCON2]);
    double gamma = gnot;
    if (y2 > 1e-4 && y1 > 1e-5)
        gamma = g1 * (0.5 * fex(y1 < 80.0 ? y1 : 80.0) + orc_vcse1f(y1) - 0.5 * orc_vcse1f(y2)) *
                (1.0 - h[HS_GCON1] / (1.0 + pow(90.0 * y1, 3.0)) - h[HS_GCON2] / (1.0 + 2000.0 * y1));
    double f = 0.0;
    if (gamma > 0) f = gamma / PI / (gamma * gamma + beta * beta);
    double prqs = orc_sofbeta(beta, h[HS_PP], n, m);
    double stark_extra = 0.0;
    if (m <= 2) {
        prqs *= 0.5;
        stark_extra = lyman_quasistatic(freq, prqs, h, dbeta, dop, n, m);
    }
    double p1 = pow(0.9 * y1, 2.0);
    double fns = (p1 + 0.03 * sqrt(y1 > 0.0 ? y1 : 0.0)) / (p1 + 1.0);
    double stark_core = (prqs * (1.0 + fns) + f) / fo30 * dbeta * 1.77245 * dop;
    if (ifcore) {
        if (nwid == 1) return core > 0.0 ? core : 0.0;
        if (nwid == 2) return lorentz > 0.0 ? lorentz : 0.0;
        double s = stark_core + stark_extra;
        return s > 0.0 ? s : 0.0;
    }
    double tot = core + lorentz + stark_core + stark_extra;
    return tot > 0.0 ? tot : 0.0;
}

/* engine/opacity.py:1758-1765 */
static double ehyd_cm(int n) {
    if (n <= 0) return 0.0;
    if (n - 1 < 8) return t_ehyd[n - 1];
    return HYD_EINF_CM - HYD_RYD_CM / (double)(n * n);
}

/* engine/opacity.py:1547-1704 (_accumulate_hydrogen_profile) into row[W]. */
static long accumulate(double *row, const double *cont, const double *stim, const double *grid, long W, long c,
                       double line_wl, double kappa0, const double *h, int n_lower, int n_upper, double wcon,
                       double wtail, double wlminus1, double wlminus2, double wlplus1, double wlplus2,
                       double redcut, double bluecut, double cutoff) {
    if (!h[HS_HAS]) return 0;
    long calls = 0;
    int simple = n_upper <= n_lower + 2;
    int use_taper = !simple && wtail > wcon;
    int upper_minus2 = n_upper - 2 > n_lower + 1 ? n_upper - 2 : n_lower + 1;
    int upper_plus2 = n_upper + 2;
    int red = 1, blue = 1;
    long offset = 1;
    long max_steps = c > W - c - 1 ? c : W - c - 1;
    if (0 <= c && c < W) {
        double wave_center = grid[c];
        if (!(!simple && wave_center < wcon)) {
            double delta_center = wave_center - line_wl;
            ++calls;
            double value_center = kappa0 * orc_hydrogen_profile(n_lower, n_upper, h, delta_center) * stim[c];
            if (use_taper && wave_center < wtail) value_center *= (wave_center - wcon) / (wtail - wcon);
            if (value_center >= cont[c] * cutoff) row[c] += value_center;
        }
    } else if (c >= W) {
        red = 0;
        offset = c - (W - 1) > 1 ? c - (W - 1) : 1;
    } else {
        blue = 0;
        offset = -c > 1 ? -c : 1;
    }
    while (offset <= max_steps && (red || blue)) {
        if (red) {
            long idx = c + offset;
            if (idx >= W) red = 0;
            else {
                double wave = grid[idx];
                if (!simple) {
                    if (wave > wlminus1) red = 0;
                    else if (wave < wcon) {
                        /* skipped, the walk continues */
                    } else {
                        double sv = stim[idx];
                        ++calls;
                        double value = kappa0 * orc_hydrogen_profile(n_lower, n_upper, h, wave - line_wl) * sv;
                        if (use_taper && wave < wtail) value *= (wave - wcon) / (wtail - wcon);
                        if (wave > redcut) {
                            ++calls;
                            double v2 = kappa0 * orc_hydrogen_profile(n_lower, upper_minus2, h, wave - wlminus2) * sv;
                            if (use_taper && wave < wtail) v2 *= (wave - wcon) / (wtail - wcon);
                            if (v2 >= value) {
                                red = 0;
                                value = 0.0;
                            }
                        }
                        if (value <= 0.0 || value < cont[idx] * cutoff) red = 0;
                        else row[idx] += value;
                    }
                } else {
                    double sv = stim[idx];
                    ++calls;
                    double value = kappa0 * orc_hydrogen_profile(n_lower, n_upper, h, wave - line_wl) * sv;
                    if (value <= 0.0 || value < cont[idx] * cutoff) red = 0;
                    else row[idx] += value;
                }
            }
        }
        if (blue) {
            long idx = c - offset;
            if (idx < 0) blue = 0;
            else {
                double wave = grid[idx];
                if (!simple && (wave < wcon || wave < wlplus1)) blue = 0;
                else {
                    double sv = stim[idx];
                    ++calls;
                    double value = kappa0 * orc_hydrogen_profile(n_lower, n_upper, h, wave - line_wl) * sv;
                    if (!simple) {
                        if (use_taper && wave < wtail) value *= (wave - wcon) / (wtail - wcon);
                        if (wave < bluecut) {
                            ++calls;
                            double v2 = kappa0 * orc_hydrogen_profile(n_lower, upper_plus2, h, wave - wlplus2) * sv;
                            if (use_taper && wave < wtail) v2 *= (wave - wcon) / (wtail - wcon);
                            if (v2 >= value) {
                                blue = 0;
                                value = 0.0;
                            }
                        }
                    }
                    if (value <= 0.0 || value < cont[idx] * cutoff) blue = 0;
                    else row[idx] += value;
                }
            }
        }
        ++offset;
    }
    return calls;
}

/* engine/opacity.py:1706-1884 (_compute_hydrogen_line_opacity) for the NH
 * hydrogen lines already selected (line type -1/-2, ion stage 1), in catalog
 * order: wl = the record's wavelength, center = nearest grid index of the
 * index wavelength (-1 / W off the grid), gf linear, elo in cm^-1, n_lower /
 * n_upper as the catalog holds them.  Per depth: pop_h1 / dop_h1 (H I
 * population and Doppler width), rho, hckt, electron density (for the
 * merge limit), states from orc_hydrogen_states.  cont and stim [D][W];
 * ahline [D][W] is added to.  stats[0] += profile evaluations.  Returns -1
 * (nothing added) if a line's n_upper lies beyond the damping tables. */
int orc_hydrogen_lines(long NH, long D, long W, const double *grid, const double *wl, const int64_t *center,
                       const double *gf, const double *elo, const int32_t *n_lower_in, const int32_t *n_upper_in,
                       const double *pop_h1, const double *dop_h1, const double *rho, const double *hckt,
                       const double *electron_density, const double *states, const double *cont,
                       const double *stim, double cutoff, double microturb_kms, double *ahline, double *stats) {
    for (long l = 0; l < NH; ++l) {
        int nl = n_lower_in[l], nu = n_upper_in[l];
        int nle = nl > 1 ? nl : 1;
        int nue = nu > nl + 1 ? nu : nl + 1;
        int nmax = nue + 2;
        if (nmax > n_asum || (nle == 1 && nmax > n_asum_lyman)) return -1;
    }
    double micro = microturb_kms > 0.0 ? microturb_kms / C_KMS : 0.0;
    double *emerge_h = (double *)malloc(sizeof(double) * (D > 0 ? D : 1));
    for (long d = 0; d < D; ++d) {
        double xne = electron_density[d] > 1e-40 ? electron_density[d] : 1e-40;
        double inglis = 1600.0 / pow(xne, 2.0 / 15.0);
        double nmerge = inglis - 1.5 > 1.0 ? inglis - 1.5 : 1.0;
        double nm2 = nmerge * nmerge;
        emerge_h[d] = HYD_RYD_CM / (nm2 > 1e-12 ? nm2 : 1e-12);
    }
    long calls = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : calls)
    for (long d = 0; d < D; ++d) {
        const double *h = states + d * HS_N;
        for (long l = 0; l < NH; ++l) {
            double line_wavelength = wl[l];
            long center_idx = center[l];
            double freq_hz = C_NM / line_wavelength;
            double cgf = CGF_CONSTANT * gf[l] / freq_hz;
            int n_lower = n_lower_in[l], n_upper = n_upper_in[l];
            int ncon = n_lower < n_conth ? n_lower : n_conth;
            ncon = (ncon > 1 ? ncon : 1) - 1;
            double conth_val = t_conth[ncon];
            int n_lower_eff = n_lower > 1 ? n_lower : 1;
            int n_upper_eff = n_upper > n_lower + 1 ? n_upper : n_lower + 1;
            double ehyd_lower = ehyd_cm(n_lower);
            double wlminus1 = n_upper - 1 > n_lower ? 1.0e7 / (ehyd_cm(n_upper - 1) - ehyd_lower) : line_wavelength;
            double wlminus2 = n_upper - 2 > n_lower ? 1.0e7 / (ehyd_cm(n_upper - 2) - ehyd_lower) : line_wavelength;
            double wlplus1 = 1.0e7 / (ehyd_cm(n_upper + 1) - ehyd_lower);
            double wlplus2 = 1.0e7 / (ehyd_cm(n_upper + 2) - ehyd_lower);
            double redcut = 1.0e7 / (t_conth[0] - HYD_RYD_CM / pow((double)n_upper - 0.8, 2.0) - ehyd_lower);
            double bluecut = 1.0e7 / (t_conth[0] - HYD_RYD_CM / pow((double)n_upper + 0.8, 2.0) - ehyd_lower);
            long cc = center_idx < 0 ? 0 : (center_idx > W - 1 ? W - 1 : center_idx);
            double wshift = 1.0e7 / (conth_val - HYD_RYD_CM / pow(81.0, 2.0));
            double pop_val = pop_h1[d], dop_val = dop_h1[d];
            if (micro > 0.0) dop_val = sqrt(dop_val * dop_val + micro * micro);
            double r = rho[d];
            if (pop_val <= 0.0 || dop_val <= 0.0 || r <= 0.0) continue;
            double xnfdop = pop_val / (r * dop_val);
            double boltz = fex(elo[l] * hckt[d]);
            double kappa0_pre = cgf * xnfdop;
            double kapmin = cont[d * W + cc] * cutoff;
            if (kappa0_pre < kapmin) continue;
            double kappa0 = kappa0_pre * boltz;
            if (kappa0 < kapmin) continue;
            double wmerge = 1.0e7 / (conth_val - emerge_h[d]);
            if (wmerge < 0.0) wmerge = wshift + wshift;
            double wcon = wshift > wmerge ? wshift : wmerge;
            double wtail = wcon > 0.0 ? 1.0e7 / (1.0e7 / wcon - 500.0) : wcon + wcon;
            wcon = wshift + wshift < wcon ? wshift + wshift : wcon;
            if (wtail < 0.0) wtail = wcon + wcon;
            wtail = wcon + wcon < wtail ? wcon + wcon : wtail;
            calls += accumulate(ahline + d * W, cont + d * W, stim + d * W, grid, W, center_idx, line_wavelength,
                                kappa0, h, n_lower_eff, n_upper_eff, wcon, wtail, wlminus1, wlminus2, wlplus1,
                                wlplus2, redcut, bluecut, cutoff);
        }
    }
    free(emerge_h);
    if (stats) stats[0] += (double)calls;
    return 0;
}

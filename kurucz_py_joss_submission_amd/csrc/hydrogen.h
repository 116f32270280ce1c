// Hydrogen-line opacity (HPROF4 profiles): the per-transition constants
// (host), the per-depth hydrogen state (host) and the profile (device), as
// physics/profiles/hydrogen.py:1408-1841, physics/populations.py:96-128 and
// engine/opacity.py:1547-1884 define them.  Included by synthe_hip.hip after
// _tables_gen.h (kH_* host tables, kD_* device tables) and synthe_device.h.
#pragma once

#include <math.h>
#include <stdint.h>

namespace synthe {
namespace hyd {

constexpr double kRydH = 3.2880515e15;
constexpr double kCAng = 2.99792458e18;  // hydrogen.py C_LIGHT (Angstrom/s)
constexpr double kCCm = 2.99792458e10;
constexpr double kCNm = 2.99792458e17;
constexpr double kLyaWn = 82259.10;
constexpr double kPi = 3.14159265358979323846;
constexpr double kRydCm = 109677.576;
constexpr double kEinfCm = 109678.764;
constexpr double kCgf = 0.026538 / 1.77245;
constexpr int kMaxFine = 16;  // fine-structure components (the tables hold at most 14)
constexpr int kMaxUpper = 94; // n_upper + 2 must index ASUM (96 entries)

// One transition n -> m: everything of hydrogen_line_profile that does not
// depend on the depth or the wavelength.
struct Trans {
  int n, m, mmn, nf;
  double xknm, freqnm, wavenm, dbeta, c1con, c2con, hwrad, resont, vdw, stark, y1num, y1wht;
  double off[kMaxFine], wt[kMaxFine];
};

// The per-depth hydrogen state (populations.py:18-29), in this order.
enum { kPP, kFO, kY1B, kY1S, kT3NHE, kT3NH2, kC1D, kC2D, kGCON1, kGCON2, kXNFPH0, kXNFPH1, kDOPPH, kXNE, kHAS,
       kStateN };

// hydrogen.py:1499-1513 (host)
inline double hf_nm(int n, int m) {
  if (m <= n) return 0.0;
  const double xn = (double)n;
  const double ginf = 0.2027 / pow(xn, 0.71);
  const double gca = 0.124 / xn;
  const double fkn = xn * 1.9603;
  const double wtc = 0.45 - 2.4 / pow(xn, 3.0) * (xn - 1.0);
  const double xm = (double)m;
  const double xmn = xm - xn;
  const double fk = fkn * pow(xm / (xmn * (xm + xn)), 3.0);
  const double xmn12 = pow(xmn, 1.2);
  const double wt = (xmn12 - 1.0) / (xmn12 + wtc);
  return fk * (1.0 - wt * ginf - (0.222 + gca / xm) * (1.0 - wt));
}

// hydrogen.py:1541-1575 and 1516-1534 (host).  Returns false where the
// reference would index past its damping tables.
inline bool make_trans(int n, int m, Trans* t) {
  t->n = n;
  t->m = m;
  t->mmn = m - n;
  const double xn = (double)n, xm = (double)m;
  const double xn2 = xn * xn, xm2 = xm * xm;
  const double xm2mn2 = xm2 - xn2;
  const double xmn2 = xm2 * xn2;
  const double gnm = xm2mn2 / xmn2;
  if (t->mmn <= 0 || n < 1) return t->mmn <= 0;
  if (m > 96 || (n == 1 && m > 100)) return false;
  const int mmn = t->mmn;
  double xknm;
  if (mmn <= 3 && n <= 4) xknm = kH_xknmtb[(n - 1) * 3 + (mmn - 1)];
  else xknm = 5.5e-5 / gnm * xmn2 / (1.0 + 0.13 / (double)mmn);
  t->xknm = xknm;
  t->freqnm = kRydH * gnm;
  t->wavenm = kCAng / t->freqnm;
  t->dbeta = kCAng / (t->freqnm * t->freqnm * xknm);
  t->c1con = xknm / t->wavenm * gnm * xm2mn2;
  t->c2con = pow(xknm / t->wavenm, 2.0);
  double radamp = kH_asum[n - 1] + kH_asum[m - 1];
  if (n == 1) radamp = kH_asum_lyman[m - 1];
  radamp /= 12.5664;
  radamp /= t->freqnm;
  t->hwrad = radamp;
  double resont = hf_nm(1, m) / xm / (1.0 - 1.0 / xm2);
  if (n != 1) resont += hf_nm(1, n) / xn / (1.0 - 1.0 / xn2);
  resont *= 3.579e-24 / gnm;
  t->resont = resont;
  t->vdw = 4.45e-26 / gnm * pow(xm2 * (7.0 * xm2 + 5.0), 0.4);
  t->stark = 1.6678e-18 * t->freqnm * xknm;
  // fine structure: Python slices clip at the table's end
  if (n > 4 || m > 10) {
    t->nf = 1;
    t->off[0] = 0.0;
    t->wt[0] = 1.0;
  } else if (mmn != 1) {
    const int ifins = (int)kH_lncomp[n - 1];
    t->nf = ifins;
    for (int i = 0; i < ifins; ++i) {
      t->off[i] = kH_stcomp[i * 4 + (n - 1)] * 1.0e7;
      t->wt[i] = kH_stcpwt[i * 4 + (n - 1)] / xn2;
    }
  } else {
    const int ifins = (int)kH_lnghal[n - 1], ipos = (int)kH_istal[n - 1];
    const int end = ipos + ifins < 34 ? ipos + ifins : 34;
    t->nf = end - ipos;
    for (int i = ipos; i < end; ++i) {
      t->off[i - ipos] = kH_stalph[i] * 1.0e7;
      t->wt[i - ipos] = kH_stwtal[i] / xn2 / 3.0;
    }
  }
  t->y1num = 320.0;
  if (m == 2) t->y1num = 550.0;
  else if (m == 3) t->y1num = 380.0;
  t->y1wht = 1.0e13;
  if (mmn <= 3) t->y1wht = 1.0e14;
  if (mmn <= 2 && n <= 2) t->y1wht = kH_y1wtm[(n - 1) * 2 + (mmn - 1)];
  return true;
}

// populations.py:96-128 and 210-231 (host): one depth's state into s[kStateN].
inline void make_state(double temperature, double electron_density, const double* xnf_he1, const double* xnf_h2,
                       const double* xnfph, const double* dopph, const double* turbulent_velocity,
                       double microturb_kms, bool has, int i, double* s) {
  for (int k = 0; k < kStateN; ++k) s[k] = 0.0;
  if (!has) return;
  const double temp = temperature > 1.0 ? temperature : 1.0;
  const double he1 = xnf_he1 ? xnf_he1[i] : 0.0, h2 = xnf_h2 ? xnf_h2[i] : 0.0;
  double dph;
  if (dopph) {
    dph = dopph[i];
  } else {
    const double ckms = 299792.458;
    const double vt = turbulent_velocity ? turbulent_velocity[i] / 1e5 / ckms : 0.0;
    const double micro = microturb_kms > 0 ? microturb_kms / ckms : 0.0;
    const double total_turb = sqrt(vt * vt + micro * micro);
    const double thermal = sqrt(2.0 * 1.380649e-16 * temp / (1.008 * 1.66054e-24)) / kCCm;
    dph = sqrt(thermal * thermal + total_turb * total_turb);
  }
  const double xne = electron_density > 1e-40 ? electron_density : 1e-40;
  const double xne16 = pow(xne, 1.0 / 6.0);
  const double tt = temp > 1.0 ? temp : 1.0;
  const double pp = xne16 * 0.08989 / sqrt(tt);
  const double fo = pow(xne16, 4.0) * 1.25e-9;
  const double y1b = 2.0 / (1.0 + 0.012 / tt * sqrt(xne / tt));
  const double t4 = temp / 10000.0;
  const double t43 = pow(t4, 0.3);
  s[kPP] = pp;
  s[kFO] = fo;
  s[kY1B] = y1b;
  s[kY1S] = t43 / xne16;
  s[kT3NHE] = t43 * he1;
  s[kT3NH2] = t43 * h2;
  s[kC1D] = fo * 78940.0 / tt;
  s[kC2D] = pow(fo, 2.0) / 5.96e-23 / xne;
  s[kGCON1] = 0.2 + 0.09 * sqrt(t4 > 1e-12 ? t4 : 1e-12) / (1.0 + xne / 1.0e13);
  s[kGCON2] = 0.2 / (1.0 + xne / 1.0e15);
  s[kXNFPH0] = xnfph ? xnfph[2 * i] : 0.0;
  s[kXNFPH1] = xnfph ? xnfph[2 * i + 1] : 0.0;
  s[kDOPPH] = dph;
  s[kXNE] = electron_density;
  s[kHAS] = 1.0;
}

// engine/opacity.py:1758-1765
inline double ehyd_cm(int n) {
  if (n <= 0) return 0.0;
  if (n - 1 < 8) return kH_ehyd_cm[n - 1];
  return kEinfCm - kRydCm / (double)(n * n);
}

#if defined(__HIPCC__)
// hydrogen.py:1408-1429
__device__ inline double vcse1f(double x, const double* extab, const double* extabf) {
  if (x <= 0) return 0.0;
  if (x <= 0.01) return -log(x) - 0.577215 + x;
  if (x <= 1.0)
    return -log(x) - 0.57721566 +
           x * (0.99999193 + x * (-0.24991055 + x * (0.05519968 + x * (-0.00976004 + x * 0.00107857))));
  if (x > 30.0) return 0.0;
  const double numerator = x * (x + 2.334733) + 0.25062;
  const double denominator = (x * (x + 3.330657) + 1.681534) * x;
  return numerator / denominator * fastex(x, extab, extabf);
}

// hydrogen.py:1451-1496
__device__ inline double sofbeta(double beta, double p, int n, int m) {
  double corr = 1.0;
  const double b2 = beta * beta;
  const double sb = sqrt(beta);
  if (beta <= 500.0) {
    int indx = 7;
    const int mmn = m - n;
    if (n <= 3 && mmn <= 2) indx = 2 * (n - 1) + mmn;
    indx = indx < 1 ? 1 : (indx > 7 ? 7 : indx);
    int im = (int)(5.0 * p) + 1;
    im = im < 4 ? im : 4;
    im = im > 1 ? im : 1;
    const int ip = im + 1;
    double wtp = 5.0 * (p - kD_pp[im - 1]);
    wtp = wtp > 1.0 ? 1.0 : wtp;
    wtp = wtp < 0.0 ? 0.0 : wtp;
    const double wtm = 1.0 - wtp;
    if (beta <= 25.12) {
      int j = 0;  // searchsorted(beta_grid, beta, side="right")
      while (j < 15 && kD_beta_grid[j] <= beta) ++j;
      j = j < 1 ? 1 : (j > 14 ? 14 : j);
      const int jm = j - 1, jp = j;
      const double denom = kD_beta_grid[jp] - kD_beta_grid[jm];
      const double wtb = denom <= 0 ? 0.0 : (beta - kD_beta_grid[jm]) / denom;
      const double wtbm = 1.0 - wtb;
      const double* prop = kD_propbm + (indx - 1) * 75;
      const double cbp = prop[(ip - 1) * 15 + jp] * wtp + prop[(im - 1) * 15 + jp] * wtm;
      const double cbm = prop[(ip - 1) * 15 + jm] * wtp + prop[(im - 1) * 15 + jm] * wtm;
      corr = 1.0 + cbp * wtb + cbm * wtbm;
      double pr1 = 0.0, pr2 = 0.0;
      double wt = 0.5 * (10.0 - beta);
      wt = wt < 1.0 ? wt : 1.0;
      wt = wt > 0.0 ? wt : 0.0;
      if (beta <= 10.0) pr1 = 8.0 / (83.0 + (2.0 + 0.95 * b2) * beta);
      if (beta >= 8.0) pr2 = (1.5 / sb + 27.0 / b2) / b2;
      return (pr1 * wt + pr2 * (1.0 - wt)) * corr;
    }
    const double cc = kD_c[(im - 1) * 7 + (indx - 1)] * wtp + kD_c[(ip - 1) * 7 + (indx - 1)] * wtm;
    const double dd = kD_d[(im - 1) * 7 + (indx - 1)] * wtp + kD_d[(ip - 1) * 7 + (indx - 1)] * wtm;
    corr = 1.0 + dd / (cc + beta * sb);
  }
  return (1.5 / sb + 27.0 / b2) / b2 * corr;
}

// hydrogen.py:1712-1728; false for None
__device__ inline bool interp_cutoff(double dwn, const double* table, int size, double start, double step,
                                     double* out) {
  const double max_delta = start + step * (double)(size - 1);
  if (dwn > max_delta) return false;
  if (dwn <= start) {
    const double frac = (dwn - start) / step;
    *out = table[0] + (table[1] - table[0]) * frac;
    return true;
  }
  const double position = (dwn - start) / step;
  const long index = (long)floor(position);
  const double frac = position - (double)index;
  if (index >= size - 1) {
    *out = table[size - 1];
    return true;
  }
  *out = table[index] + (table[index + 1] - table[index]) * frac;
  return true;
}

// hydrogen.py:1541-1709 for transition t at one depth state s.
__device__ inline double profile(const Trans& t, const double* s, double delta_nm, const double* extab,
                                 const double* extabf) {
  if (s[kHAS] == 0.0) return 0.0;
  if (t.mmn <= 0) return 0.0;
  const int n = t.n, m = t.m;
  const double freqnm = t.freqnm, hwrad = t.hwrad;
  const double hwvdw = t.vdw * s[kT3NHE] + 2.0 * t.vdw * s[kT3NH2];
  const double hwstk = t.stark * s[kFO];
  const double hwres = t.resont * s[kXNFPH0] * 2.0;
  const double hwlor = hwres + hwvdw + hwrad;
  const double wl = t.wavenm + delta_nm * 10.0;
  const double freq = kCAng / wl;
  const double del_freq = fabs(freq - freqnm);
  const double dopph = s[kDOPPH] > 1e-40 ? s[kDOPPH] : 1e-40;
  const double dop = freqnm * dopph;
  double wmax = dopph > hwlor ? dopph : hwlor;
  wmax = wmax > hwstk ? wmax : hwstk;
  const bool ifcore = del_freq <= freqnm * wmax;
  int nwid = 1;
  if (!(dopph >= hwstk && dopph >= hwlor)) nwid = hwlor < hwstk ? 3 : 2;
  double core = 0.0;
  const double dop30 = dop > 1e-30 ? dop : 1e-30;
  for (int k = 0; k < t.nf; ++k) {
    const double dd = fabs(freq - (freqnm + t.off[k])) / dop30;
    if (dd <= 7.0) core += fastex(dd * dd, extab, extabf) * t.wt[k];
  }
  double lorentz = 0.0;
  const double hhw = freqnm * hwlor;
  if (n == 1 && m == 2) {  // hydrogen.py:1731-1797
    if (dop > 0.0) {
      const double hwres_near = hwres * 4.0;
      const double hwlor_near = hwres_near + hwvdw + hwrad;
      const double hhw_near = freqnm * (hwlor_near > 0.0 ? hwlor_near : 0.0);
      const double delta_wn = freq / kCCm - kLyaWn;
      double hres_term, hhw_use;
      if (freq > (kLyaWn - 4000.0) * kCCm && hhw_near > 0.0) {
        hres_term = hwres_near * freqnm / kPi / (del_freq * del_freq + hhw_near * hhw_near) * 1.77245 * dop;
        hhw_use = hhw_near;
      } else {
        double cutoff_val = 0.0, cl;
        if (interp_cutoff(delta_wn, kD_cutoff_h2, 91, -22000.0, 200.0, &cl))
          cutoff_val = pow(10.0, cl - 14.0) * s[kXNFPH0] * 2.0 / kCCm;
        hres_term = cutoff_val * 1.77245 * dop;
        hhw_use = freqnm * (hwlor > 0.0 ? hwlor : 0.0);
      }
      double hrad_term = 0.0, hvdw_term = 0.0;
      if (hwrad > 0.0 && hhw_use > 0.0 && freq > 2.4190611e15 && freq < 0.77 * kRydH)
        hrad_term = hwrad * freqnm / kPi / (del_freq * del_freq + hhw_use * hhw_use) * 1.77245 * dop;
      if (hwvdw > 0.0 && hhw_use > 0.0 && freq >= 1.8e15)
        hvdw_term = hwvdw * freqnm / kPi / (del_freq * del_freq + hhw_use * hhw_use) * 1.77245 * dop;
      lorentz = hres_term + hrad_term + hvdw_term;
    }
  } else {
    double top = hhw;
    if (n == 1 && (m == 3 || m == 4 || m == 5)) {
      const double fr = freq / kRydH;
      double t2 = hhw - freqnm * hwrad;
      t2 = t2 > 0.0 ? t2 : 0.0;
      if ((m == 3 && 0.885 <= fr && fr <= 0.890) || (m == 4 && 0.936 <= fr && fr <= 0.938) ||
          (m == 5 && 0.959 <= fr && fr <= 0.961))
        top = t2;
    }
    if (hhw > 0.0) lorentz = top / kPi / (del_freq * del_freq + hhw * hhw) * 1.77245 * dop;
  }
  const double xne0 = s[kXNE] > 0.0 ? s[kXNE] : 0.0;
  const double wty1 = 1.0 / (1.0 + xne0 / (t.y1wht > 1e-30 ? t.y1wht : 1e-30));
  const double y1_scal = t.y1num * s[kY1S] * wty1 + s[kY1B] * (1.0 - wty1);
  const double c1 = s[kC1D] * t.c1con * y1_scal;
  const double c2 = s[kC2D] * t.c2con;
  const double fo30 = s[kFO] > 1e-30 ? s[kFO] : 1e-30;
  const double beta = del_freq / fo30 * t.dbeta;
  const double y1 = c1 * beta;
  const double y2 = c2 * beta * beta;
  const double g1 = 6.77 * sqrt(c1 > 1e-30 ? c1 : 1e-30);
  double ratio = 0.0;
  if (c1 > 0.0 && c2 > 0.0) ratio = sqrt(c2) / (c1 > 1e-30 ? c1 : 1e-30);
  double log_term = 0.0;
  if (ratio > 0.0) log_term = log(ratio > 1e-30 ? ratio : 1e-30);
  const double lt = 0.2114 + log_term;
  double gamma = g1 * (lt > 0.0 ? lt : 0.0) * (1.0 - s[kGCON1] - s[kGCON2]);
  if (y2 > 1e-4 && y1 > 1e-5)
    gamma = g1 *
            (0.5 * fastex(y1 < 80.0 ? y1 : 80.0, extab, extabf) + vcse1f(y1, extab, extabf) -
             0.5 * vcse1f(y2, extab, extabf)) *
            (1.0 - s[kGCON1] / (1.0 + pow(90.0 * y1, 3.0)) - s[kGCON2] / (1.0 + 2000.0 * y1));
  double f = 0.0;
  if (gamma > 0) f = gamma / kPi / (gamma * gamma + beta * beta);
  double prqs = sofbeta(beta, s[kPP], n, m);
  double stark_extra = 0.0;
  if (m <= 2) {  // hydrogen.py:1800-1841
    prqs *= 0.5;
    const double delta_wn = freq / kCCm - kLyaWn;
    if (s[kFO] > 0.0 && delta_wn >= -20000.0) {
      if (delta_wn <= -4000.0) {
        double cl;
        if (interp_cutoff(delta_wn, kD_cutoff_h2_plus, 111, -15000.0, 100.0, &cl))
          stark_extra += pow(10.0, cl - 14.0) * s[kXNFPH1] / kCCm * 1.77245 * dop;
      } else {
        const double beta4000 = 4000.0 * kCCm / fo30 * t.dbeta;
        const double prqs4000 = sofbeta(beta4000, s[kPP], n, m) * 0.5;
        const double normalization = prqs4000 / fo30 * t.dbeta;
        const double cutoff4000 = pow(10.0, -11.07 - 14.0) * s[kXNFPH1] / kCCm;
        if (normalization > 0.0)
          stark_extra += cutoff4000 / normalization * (prqs / fo30 * t.dbeta) * 1.77245 * dop;
      }
    }
  }
  const double p1 = pow(0.9 * y1, 2.0);
  const double fns = (p1 + 0.03 * sqrt(y1 > 0.0 ? y1 : 0.0)) / (p1 + 1.0);
  const double stark_core = (prqs * (1.0 + fns) + f) / fo30 * t.dbeta * 1.77245 * dop;
  double v;
  if (ifcore) v = nwid == 1 ? core : (nwid == 2 ? lorentz : stark_core + stark_extra);
  else v = core + lorentz + stark_core + stark_extra;
  return v > 0.0 ? v : 0.0;
}
#endif

}  // namespace hyd
}  // namespace synthe

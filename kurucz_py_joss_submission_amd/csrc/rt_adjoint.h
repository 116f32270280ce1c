// Vector-Jacobian product of the total-flux JOSH solve of one wavelength
// (physics/josh_solver.py:367-767) with respect to its per-depth inputs
// abtot, alpha, snubar, for the adjoint of the synthesis.
//
// The solve is: TAUNU = INTEG(rho, abtot) (PARCOE smoothing, _integ), then
// either the XTAU branch -- MAP1 of snubar and alpha onto the 51 XTAU points,
// surface clamping, the scattering fixed point XS = XSBAR (1 - XALPHA) +
// XALPHA (COEFJ . XS) and flux = CH . XS -- or, when TAUNU(0) > XTAU(51),
// the label-401 DERIV iteration on the depth grid with flux = HNU(0).
//
// Derivatives:
//  - the scattering fixed point is differentiated implicitly: with M = I -
//    diag(XA) C, the adjoint lam solves M^T lam = CH (a transposed sweep
//    iteration in FP64), giving dF/dXB_k = lam_k (1 - XA_k) and dF/dXA_k =
//    lam_k ((C XS)_k - XB_k) at the FP64 fixed point XS;
//  - INTEG, PARCOE and MAP1 (whose smoothing weights depend on the data) are
//    differentiated in forward mode with dual numbers, one input at a time,
//    and contracted with those sensitivities;
//  - the label-401 branch is differentiated in forward mode through its
//    iteration, run (values and derivatives) to a relative change of 1e-12
//    instead of the reference's 1e-5, so that both branches differentiate
//    the converged solve.
// Branch choices (label 401, the surface set XTAU < TAUNU(0), clamps) are
// held fixed, as their derivative is zero almost everywhere.
//
// The header compiles as HIP device code and as plain C++ (for CPU checks).
#pragma once

#include <math.h>

#include "rt_core.h"

namespace synthe {
namespace adj {

constexpr int kXt = 51;
constexpr double kEps = 1.0e-38;

struct Dual {
  double v, d;
  SYNTHE_HD Dual() : v(0.0), d(0.0) {}
  SYNTHE_HD Dual(double x) : v(x), d(0.0) {}
  SYNTHE_HD Dual(double x, double y) : v(x), d(y) {}
};

SYNTHE_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
SYNTHE_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
SYNTHE_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
SYNTHE_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
SYNTHE_HD Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
SYNTHE_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
SYNTHE_HD Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
SYNTHE_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
SYNTHE_HD Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
SYNTHE_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
SYNTHE_HD Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
SYNTHE_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
SYNTHE_HD Dual operator/(double a, Dual b) {
  double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
SYNTHE_HD Dual& operator+=(Dual& a, Dual b) { return a = a + b; }

SYNTHE_HD double val(double x) { return x; }
SYNTHE_HD double val(const Dual& x) { return x.v; }
SYNTHE_HD double tabs(double x) { return fabs(x); }
SYNTHE_HD Dual tabs(Dual x) { return x.v < 0.0 ? -x : x; }
SYNTHE_HD double tsqrt(double x) { return sqrt(x); }
SYNTHE_HD Dual tsqrt(Dual x) {
  double s = sqrt(x.v);
  return Dual(s, s > 0.0 ? 0.5 * x.d / s : 0.0);
}

// physics/josh_solver.py:95-167 (_parcoe) on f(x), x fixed.
template <class T>
SYNTHE_HD void parcoe_t(const T* f, const double* x, int n, T* a, T* b, T* c) {
  for (int i = 0; i < n; ++i) a[i] = b[i] = c[i] = T(0.0);
  if (n == 0) return;
  if (n == 1) {
    a[0] = f[0];
    return;
  }
  double denom = x[1] - x[0];
  b[0] = denom != 0.0 ? (f[1] - f[0]) / denom : T(0.0);
  a[0] = f[0] - x[0] * b[0];
  int n1 = n - 1;
  c[n - 1] = T(0.0);
  denom = x[n - 1] - x[n1 - 1];
  b[n - 1] = denom != 0.0 ? (f[n - 1] - f[n1 - 1]) / denom : T(0.0);
  a[n - 1] = f[n - 1] - x[n - 1] * b[n - 1];
  if (n == 2) return;
  for (int j = 1; j < n1; ++j) {
    int j1 = j - 1;
    denom = x[j] - x[j1];
    T d = denom != 0.0 ? (f[j] - f[j1]) / denom : T(0.0);
    double denom1 = (x[j + 1] - x[j]) * (x[j + 1] - x[j1]);
    T term1 = denom1 != 0.0 ? f[j + 1] / denom1 : T(0.0);
    double denom2 = x[j + 1] - x[j1];
    double denom3 = x[j + 1] - x[j];
    double denom4 = x[j] - x[j1];
    T part = T(0.0);
    if (denom4 != 0.0) {
      T t1 = denom2 != 0.0 ? f[j1] / denom2 : T(0.0);
      T t2 = denom3 != 0.0 ? f[j] / denom3 : T(0.0);
      part = (t1 - t2) / denom4;
    }
    c[j] = term1 + part;
    b[j] = d - (x[j] + x[j1]) * c[j];
    a[j] = f[j1] - x[j1] * d + x[j] * x[j1] * c[j];
  }
  c[1] = T(0.0);
  denom = x[2] - x[1];
  b[1] = denom != 0.0 ? (f[2] - f[1]) / denom : T(0.0);
  a[1] = f[1] - x[1] * b[1];
  if (n > 3) {
    c[2] = T(0.0);
    denom = x[3] - x[2];
    b[2] = denom != 0.0 ? (f[3] - f[2]) / denom : T(0.0);
    a[2] = f[2] - x[2] * b[2];
  }
  for (int j = 1; j < n1; ++j) {
    if (val(c[j]) == 0.0) continue;
    int j1 = j + 1 < n - 1 ? j + 1 : n - 1;
    T dn = tabs(c[j1]) + tabs(c[j]);
    T wt = val(dn) > 0.0 ? tabs(c[j1]) / dn : T(0.0);
    a[j] = a[j1] + wt * (a[j] - a[j1]);
    b[j] = b[j1] + wt * (b[j] - b[j1]);
    c[j] = c[j1] + wt * (c[j] - c[j1]);
  }
  a[n1 - 1] = a[n - 1];
  b[n1 - 1] = b[n - 1];
  c[n1 - 1] = c[n - 1];
}

// physics/josh_solver.py:170-189 (_integ): TAUNU from abtot over rho (increasing).
template <class T>
SYNTHE_HD void integ_t(const T* abtot, const double* rho, int n, T* taunu, T* pa, T* pb, T* pc) {
  parcoe_t(abtot, rho, n, pa, pb, pc);
  taunu[0] = abtot[0] * rho[0];
  for (int i = 0; i + 1 < n; ++i) {
    double dx = rho[i + 1] - rho[i];
    T term = pa[i] + 0.5 * pb[i] * (rho[i + 1] + rho[i]);
    term += (pc[i] / 3.0) * ((rho[i + 1] + rho[i]) * rho[i + 1] + rho[i] * rho[i]);
    taunu[i + 1] = taunu[i] + term * dx;
  }
}

// physics/josh_solver.py:192-217 (_deriv)
template <class T>
SYNTHE_HD void deriv_t(const T* x, const T* f, int n, T* dfdx) {
  for (int i = 0; i < n; ++i) dfdx[i] = T(0.0);
  if (n < 2) return;
  dfdx[0] = (f[1] - f[0]) / (x[1] - x[0]);
  dfdx[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
  if (n == 2) return;
  double s = val(x[1]) != val(x[0]) ? fabs(val(x[1]) - val(x[0])) / (val(x[1]) - val(x[0])) : 1.0;
  for (int j = 1; j < n - 1; ++j) {
    T scale = tabs(f[j - 1]);
    if (val(tabs(f[j])) > val(scale)) scale = tabs(f[j]);
    if (val(tabs(f[j + 1])) > val(scale)) scale = tabs(f[j + 1]);
    scale = val(x[j]) != 0.0 ? scale / tabs(x[j]) : scale;
    if (val(scale) == 0.0) scale = T(1.0);
    T d1 = (f[j + 1] - f[j]) / (x[j + 1] - x[j]) / scale;
    T d0 = (f[j] - f[j - 1]) / (x[j] - x[j - 1]) / scale;
    T tan1 = d1 / (s * tsqrt(1.0 + d1 * d1) + 1.0);
    T tan0 = d0 / (s * tsqrt(1.0 + d0 * d0) + 1.0);
    dfdx[j] = (tan1 + tan0) / (1.0 - tan1 * tan0) * scale;
  }
}

// physics/josh_solver.py:220-353 (_map1_kernel) onto the 51 XTAU points.
template <class T>
SYNTHE_HD void map1_t(const T* xold, const T* fold, int nold, const double* xnew, T* fnew) {
#define XO(i) xold[(i)-1]
#define FO(i) fold[(i)-1]
  for (int k = 0; k < kXt; ++k) fnew[k] = T(0.0);
  if (nold == 0) return;
  int l = 2, ll = 0;
  T cfor(0.0), bfor(0.0), afor(0.0), cbac(0.0), bbac(0.0), abac(0.0), a(0.0), b(0.0), c(0.0);
  for (int k = 1; k <= kXt; ++k) {
    double xk = xnew[k - 1];
    for (;;) {
      if (xk < val(XO(l))) {
        if (l == ll) break;
        if (l == 2 || l == 3) {
          l = nold < l ? nold : l;
          c = T(0.0);
          b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
          a = FO(l) - XO(l) * b;
          ll = l;
          break;
        }
        int l1 = l - 1;
        if (l > ll + 1 || l == 4) {
          int l2 = l - 2;
          T d = (FO(l1) - FO(l2)) / (XO(l1) - XO(l2));
          cbac = FO(l) / ((XO(l) - XO(l1)) * (XO(l) - XO(l2))) +
                 (FO(l2) / (XO(l) - XO(l2)) - FO(l1) / (XO(l) - XO(l1))) / (XO(l1) - XO(l2));
          bbac = d - (XO(l1) + XO(l2)) * cbac;
          abac = FO(l2) - XO(l2) * d + XO(l1) * XO(l2) * cbac;
          if (!(l < nold)) {
            c = cbac;
            b = bbac;
            a = abac;
            ll = l;
            break;
          }
        } else {
          cbac = cfor;
          bbac = bfor;
          abac = afor;
          if (l == nold) {
            c = cbac;
            b = bbac;
            a = abac;
            ll = l;
            break;
          }
        }
        T d = (FO(l) - FO(l1)) / (XO(l) - XO(l1));
        cfor = FO(l + 1) / ((XO(l + 1) - XO(l)) * (XO(l + 1) - XO(l1))) +
               (FO(l1) / (XO(l + 1) - XO(l1)) - FO(l) / (XO(l + 1) - XO(l))) / (XO(l) - XO(l1));
        bfor = d - (XO(l) + XO(l1)) * cfor;
        afor = FO(l1) - XO(l1) * d + XO(l) * XO(l1) * cfor;
        T wt(0.0);
        if (val(tabs(cfor)) != 0.0) wt = tabs(cfor) / (tabs(cfor) + tabs(cbac));
        a = afor + wt * (abac - afor);
        b = bfor + wt * (bbac - bfor);
        c = cfor + wt * (cbac - cfor);
        ll = l;
        break;
      }
      l += 1;
      if (l > nold) {
        l = nold < l ? nold : l;
        c = T(0.0);
        b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
        a = FO(l) - XO(l) * b;
        ll = l;
        break;
      }
    }
    fnew[k - 1] = a + (b + c * xk) * xk;
  }
#undef XO
#undef FO
}

// Label-401 branch (physics/josh_solver.py:700-727): returns HNU(0) after the
// iteration; snu, hnu, jnu, tmp are work arrays of n.
// The reference stops at a summed relative change below 1e-5 within 51
// iterations; the adjoint's smooth model iterates to tol (max_iter).
template <class T>
SYNTHE_HD T josh401_t(int n, const T* taunu, const T* snubar, const T* alpha, T* snu, T* hnu, T* jnu, T* tmp,
                      int* iters, double tol = 1.0e-5, int max_iter = kXt) {
  for (int j = 0; j < n; ++j) snu[j] = snubar[j];
  int it = 0;
  for (int l = 0; l < max_iter; ++l) {
    deriv_t(taunu, snu, n, hnu);
    for (int j = 0; j < n; ++j) hnu[j] = hnu[j] / 3.0;
    deriv_t(taunu, hnu, n, jnu);
    double err = 0.0;
    for (int j = 0; j < n; ++j) {
      T jn = jnu[j] + snu[j];
      T sn = (1.0 - alpha[j]) * snubar[j] + alpha[j] * jn;
      double den = fabs(val(sn)) > kEps ? fabs(val(sn)) : kEps;
      err += fabs(val(sn) - val(snu[j])) / den;
      snu[j] = sn;
    }
    it = l + 1;
    if (err < tol) break;
  }
  if (iters) *iters = it;
  return hnu[0];
}

constexpr double kTol401 = 1.0e-12;  // label-401 convergence of the smooth model
constexpr int kMax401 = 300;

struct JoshTab {
  const double* xtau;    // [51]
  const double* ch;      // [51]
  const float* coefj32;  // [51][51], the float32 COEFJ of the sweep
};

// FP64 scattering fixed point XS = XB (1 - XA) + XA (C XS) by Gauss-Seidel
// sweeps (descending k, as physics/josh_solver.py:24-92) to relative change
// tol; returns the sweep count.
SYNTHE_HD int fixed_point64(const JoshTab& tb, const double* xb, const double* xa, double* xs, double tol,
                            int max_sweeps) {
  for (int k = 0; k < kXt; ++k) xs[k] = xb[k];
  for (int it = 0; it < max_sweeps; ++it) {
    double err = 0.0;
    for (int k = kXt - 1; k >= 0; --k) {
      double cx = 0.0;
      for (int m = 0; m < kXt; ++m) cx += (double)tb.coefj32[k * kXt + m] * xs[m];
      double ckk = (double)tb.coefj32[k * kXt + k];
      double nv = (xb[k] * (1.0 - xa[k]) + xa[k] * (cx - ckk * xs[k])) / (1.0 - xa[k] * ckk);
      double e = fabs(nv - xs[k]) / (fabs(nv) > kEps ? fabs(nv) : kEps);
      err = e > err ? e : err;
      xs[k] = nv;
    }
    if (err < tol) return it + 1;
  }
  return max_sweeps;
}

// Adjoint of the fixed point: lam = CH + C^T (XA o lam), Gauss-Seidel.
SYNTHE_HD int adjoint64(const JoshTab& tb, const double* xa, double* lam, double tol, int max_sweeps) {
  for (int k = 0; k < kXt; ++k) lam[k] = tb.ch[k];
  for (int it = 0; it < max_sweeps; ++it) {
    double err = 0.0;
    for (int m = 0; m < kXt; ++m) {
      double s = 0.0;
      for (int k = 0; k < kXt; ++k)
        if (k != m) s += (double)tb.coefj32[k * kXt + m] * xa[k] * lam[k];
      double cmm = (double)tb.coefj32[m * kXt + m];
      double nv = (tb.ch[m] + s) / (1.0 - cmm * xa[m]);
      double e = fabs(nv - lam[m]) / (fabs(nv) > kEps ? fabs(nv) : kEps);
      err = e > err ? e : err;
      lam[m] = nv;
    }
    if (err < tol) return it + 1;
  }
  return max_sweeps;
}

// Per-depth inputs of one total-flux solve in increasing rho order (n >= 4).
template <int kMax>
struct Column {
  int n;
  double rho[kMax], abtot[kMax], alpha[kMax], snubar[kMax];
};

// Work space of the solve (caller-provided, e.g. private memory).
template <int kMax>
struct Work {
  double taunu[kMax], pa[kMax], pb[kMax], pc[kMax];
  double xsbar[kXt], xalpha[kXt], xb[kXt], xa[kXt], xs[kXt], lam[kXt], gxb[kXt], gxa[kXt];
  Dual d_in[kMax], d_tau[kMax], d_a[kMax], d_b[kMax], d_c[kMax], d_f[kMax], d_out[kXt], d_out2[kXt];
  Dual d_s[kMax], d_h[kMax], d_j[kMax], d_t[kMax], d_al[kMax], d_sb[kMax];
};

// FP64 total flux of the column with the fixed point solved to tol (the
// smooth model the vjp differentiates; the float32 sweep of the forward pass
// agrees with it to the forward's tolerance).  Returns the flux; *took401.
template <int kMax>
SYNTHE_HD double total_flux64(const Column<kMax>& col, const JoshTab& tb, Work<kMax>& w, double tol,
                              int* took401) {
  const int n = col.n;
  integ_t(col.abtot, col.rho, n, w.taunu, w.pa, w.pb, w.pc);
  if (w.taunu[0] > tb.xtau[kXt - 1]) {
    *took401 = 1;
    return josh401_t(n, w.taunu, col.snubar, col.alpha, w.pa, w.pb, w.pc, w.xs, nullptr, kTol401, kMax401);
  }
  *took401 = 0;
  map1_t(w.taunu, col.snubar, n, tb.xtau, w.xsbar);
  map1_t(w.taunu, col.alpha, n, tb.xtau, w.xalpha);
  const double s0 = col.snubar[0] > kEps ? col.snubar[0] : kEps;
  const double a0 = col.alpha[0] < 0.0 ? 0.0 : (col.alpha[0] > 1.0 ? 1.0 : col.alpha[0]);
  for (int k = 0; k < kXt; ++k) {
    double xb = w.xsbar[k] > kEps ? w.xsbar[k] : kEps;
    double xa = w.xalpha[k] < 0.0 ? 0.0 : (w.xalpha[k] > 1.0 ? 1.0 : w.xalpha[k]);
    if (tb.xtau[k] < w.taunu[0]) {
      xb = s0;
      xa = a0;
    }
    w.xb[k] = xb;
    w.xa[k] = xa;
  }
  fixed_point64(tb, w.xb, w.xa, w.xs, tol, 400);
  double f = 0.0;
  for (int k = 0; k < kXt; ++k) f += tb.ch[k] * w.xs[k];
  return f;
}

// g_abtot, g_alpha, g_snubar [n] = v * dF/d(input) for the inputs j = part,
// part + nparts, ... (all of them by default; others are left untouched);
// returns F.
template <int kMax>
SYNTHE_HD double total_flux_vjp(const Column<kMax>& col, const JoshTab& tb, Work<kMax>& w, double v,
                                double* g_abtot, double* g_alpha, double* g_snubar, int part = 0,
                                int nparts = 1) {
  const int n = col.n;
  int took401 = 0;
  const double F = total_flux64(col, tb, w, 1e-13, &took401);
  if (took401) {
    // forward mode through INTEG and the label-401 iteration, one input at a time
    for (int q = 0; q < 3; ++q) {
      double* g = q == 0 ? g_abtot : (q == 1 ? g_alpha : g_snubar);
      for (int j = part; j < n; j += nparts) {
        for (int i = 0; i < n; ++i) {
          w.d_in[i] = Dual(col.abtot[i], q == 0 && i == j ? 1.0 : 0.0);
          w.d_al[i] = Dual(col.alpha[i], q == 1 && i == j ? 1.0 : 0.0);
          w.d_sb[i] = Dual(col.snubar[i], q == 2 && i == j ? 1.0 : 0.0);
        }
        integ_t(w.d_in, col.rho, n, w.d_tau, w.d_a, w.d_b, w.d_c);
        Dual f = josh401_t(n, w.d_tau, w.d_sb, w.d_al, w.d_s, w.d_h, w.d_j, w.d_t, nullptr, kTol401, kMax401);
        g[j] = v * f.d;
      }
    }
    return F;
  }
  // sensitivities of F to the fixed point's inputs XB, XA
  adjoint64(tb, w.xa, w.lam, 1e-13, 400);
  double gs0 = 0.0, ga0 = 0.0;
  for (int k = 0; k < kXt; ++k) {
    double cx = 0.0;
    for (int m = 0; m < kXt; ++m) cx += (double)tb.coefj32[k * kXt + m] * w.xs[m];
    double gxb = w.lam[k] * (1.0 - w.xa[k]);
    double gxa = w.lam[k] * (cx - w.xb[k]);
    if (tb.xtau[k] < w.taunu[0]) {
      gs0 += gxb;
      ga0 += gxa;
      gxb = gxa = 0.0;
    } else {
      if (!(w.xsbar[k] > kEps)) gxb = 0.0;
      if (w.xalpha[k] < 0.0 || w.xalpha[k] > 1.0) gxa = 0.0;
    }
    w.gxb[k] = gxb;
    w.gxa[k] = gxa;
  }
  if (!(col.snubar[0] > kEps)) gs0 = 0.0;
  if (col.alpha[0] < 0.0 || col.alpha[0] > 1.0) ga0 = 0.0;
  // snubar and alpha enter MAP1 as values on fixed nodes
  for (int i = 0; i < n; ++i) w.d_tau[i] = Dual(w.taunu[i]);
  for (int j = part; j < n; j += nparts) {
    for (int i = 0; i < n; ++i) w.d_f[i] = Dual(col.snubar[i], i == j ? 1.0 : 0.0);
    map1_t(w.d_tau, w.d_f, n, tb.xtau, w.d_out);
    double s = j == 0 ? gs0 : 0.0;
    for (int k = 0; k < kXt; ++k) s += w.gxb[k] * w.d_out[k].d;
    g_snubar[j] = v * s;
    for (int i = 0; i < n; ++i) w.d_f[i] = Dual(col.alpha[i], i == j ? 1.0 : 0.0);
    map1_t(w.d_tau, w.d_f, n, tb.xtau, w.d_out);
    s = j == 0 ? ga0 : 0.0;
    for (int k = 0; k < kXt; ++k) s += w.gxa[k] * w.d_out[k].d;
    g_alpha[j] = v * s;
  }
  // abtot moves the nodes TAUNU
  for (int j = part; j < n; j += nparts) {
    for (int i = 0; i < n; ++i) w.d_in[i] = Dual(col.abtot[i], i == j ? 1.0 : 0.0);
    integ_t(w.d_in, col.rho, n, w.d_tau, w.d_a, w.d_b, w.d_c);
    for (int i = 0; i < n; ++i) w.d_f[i] = Dual(col.snubar[i]);
    map1_t(w.d_tau, w.d_f, n, tb.xtau, w.d_out);
    for (int i = 0; i < n; ++i) w.d_f[i] = Dual(col.alpha[i]);
    map1_t(w.d_tau, w.d_f, n, tb.xtau, w.d_out2);
    double s = 0.0;
    for (int k = 0; k < kXt; ++k) s += w.gxb[k] * w.d_out[k].d + w.gxa[k] * w.d_out2[k].d;
    g_abtot[j] = v * s;
  }
  return F;
}

}  // namespace adj
}  // namespace synthe

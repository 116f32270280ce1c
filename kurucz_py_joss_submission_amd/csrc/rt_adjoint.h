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
// Derivatives (reverse mode, O(n) per column):
//  - the scattering fixed point is differentiated implicitly: with M = I -
//    diag(XA) C, the adjoint lam solves M^T lam = CH (a transposed sweep
//    iteration in FP64), giving dF/dXB_k = lam_k (1 - XA_k) and dF/dXA_k =
//    lam_k ((C XS)_k - XB_k) at the FP64 fixed point XS;
//  - PARCOE and MAP1 (whose smoothing weights depend on the data) are local:
//    an output depends on at most 4 consecutive inputs.  Their Jacobians are
//    therefore exact from 4 dual-number passes each, seeding every input
//    j = r mod 4 in pass r (no output sees two seeded inputs), and are
//    contracted with the sensitivities; INTEG's running sum is reversed as a
//    suffix sum;
//  - the label-401 branch is differentiated implicitly at its converged
//    iterate S = PHI(S) = (1 - alpha) snubar + alpha (J(S) + S), where J is
//    two DERIVs: PHI's Jacobians in S and TAUNU are 5-banded (5 dual passes
//    each) and the adjoint solves (I - dPHI/dS)^T lam = dF/dS by the same
//    (transposed) iteration.  The value iteration runs to a relative change
//    of 1e-12 instead of the reference's 1e-5, so that both branches
//    differentiate the converged solve.
// Branch choices (label 401, the surface set XTAU < TAUNU(0), clamps, MAP1
// brackets) are held fixed, as their derivative is zero almost everywhere.
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
  Dual() = default;  // trivial, so work arrays of duals cost nothing to create
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
// base[k] (optional): the 1-based bracket index L at which fnew[k]'s parabola
// was formed; fnew[k] depends on xold and fold at 0-based L-3 .. L only.
template <class T>
SYNTHE_HD void map1_t(const T* xold, const T* fold, int nold, const double* xnew, T* fnew, int* base = nullptr) {
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
    if (base) base[k - 1] = ll;
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

// A vector of XTAU values with a stride: a private array (stride 1) or one
// lane's column of a workgroup array interleaved by lane (stride = lanes).
struct Vec {
  double* p;
  int s;
  SYNTHE_HD double& operator[](int k) const { return p[k * s]; }
};

constexpr int kGsSweeps = 60;      // Gauss-Seidel sweeps before the Krylov finish
constexpr int kKrylovIters = 200;  // BiCGSTAB iterations at most
constexpr double kKrylovTol = 1e-15;  // relative residual of the Krylov finish

// y = A v for the fixed point's matrix A = I - diag(XA) C, or for its
// transpose I - C^T diag(XA) (the adjoint's).
template <class VA, class VV>
SYNTHE_HD void fp_matvec(const JoshTab& tb, const VA& xa, const VV& v, double* y, bool transpose) {
  for (int k = 0; k < kXt; ++k) {
    double s = 0.0;
    if (transpose) {
      for (int m = 0; m < kXt; ++m) s += (double)tb.coefj32[m * kXt + k] * xa[m] * v[m];
      y[k] = v[k] - s;
    } else {
      for (int m = 0; m < kXt; ++m) s += (double)tb.coefj32[k * kXt + m] * v[m];
      y[k] = v[k] - xa[k] * s;
    }
  }
}

SYNTHE_HD double dot51(const double* a, const double* b) {
  double s = 0.0;
  for (int k = 0; k < kXt; ++k) s += a[k] * b[k];
  return s;
}

// BiCGSTAB (van der Vorst 1992) for A x = b, A as fp_matvec, from the
// iterate in x; b is given in kr[0].  kr: 6 vectors of kXt (b/r, r0, p, v,
// s, t).  Stops at |r| <= kKrylovTol |b|; returns the iteration count.
template <class VA, class VX>
SYNTHE_HD int bicgstab51(const JoshTab& tb, const VA& xa, const VX& x, bool transpose, double (*kr)[kXt]) {
  double *r = kr[0], *r0 = kr[1], *p = kr[2], *v = kr[3], *sv = kr[4], *t = kr[5];
  const double bn = sqrt(dot51(r, r));
  fp_matvec(tb, xa, x, v, transpose);
  for (int k = 0; k < kXt; ++k) {
    r[k] -= v[k];
    r0[k] = r[k];
    p[k] = v[k] = 0.0;
  }
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int it = 0;
  while (it < kKrylovIters && sqrt(dot51(r, r)) > kKrylovTol * bn) {
    ++it;
    const double rho1 = dot51(r0, r);
    if (rho1 == 0.0 || omega == 0.0) break;
    const double beta = (rho1 / rho) * (alpha / omega);
    for (int k = 0; k < kXt; ++k) p[k] = r[k] + beta * (p[k] - omega * v[k]);
    fp_matvec(tb, xa, p, v, transpose);
    const double r0v = dot51(r0, v);
    if (r0v == 0.0) break;
    alpha = rho1 / r0v;
    for (int k = 0; k < kXt; ++k) sv[k] = r[k] - alpha * v[k];
    fp_matvec(tb, xa, sv, t, transpose);
    const double tt = dot51(t, t);
    omega = tt > 0.0 ? dot51(t, sv) / tt : 0.0;
    for (int k = 0; k < kXt; ++k) {
      x[k] = x[k] + alpha * p[k] + omega * sv[k];
      r[k] = sv[k] - omega * t[k];
    }
    rho = rho1;
  }
  return it;
}

// FP64 scattering fixed point XS = XB (1 - XA) + XA (C XS) by Gauss-Seidel
// sweeps (descending k, as physics/josh_solver.py:24-92) to relative change
// tol.  Strong scattering (XA near 1) makes the sweeps contract slowly (by
// about max XA each): after kGsSweeps, if kr (6 x kXt scratch) is given, the
// solve is finished by BiCGSTAB.  Returns sweeps + Krylov iterations.
// V: double* or Vec.
template <class V>
SYNTHE_HD int fixed_point64(const JoshTab& tb, const double* xb, const V& xa, const V& xs, double tol,
                            int max_sweeps, double (*kr)[kXt] = nullptr) {
  for (int k = 0; k < kXt; ++k) xs[k] = xb[k];
  const int gs = kr && max_sweeps > kGsSweeps ? kGsSweeps : max_sweeps;
  for (int it = 0; it < gs; ++it) {
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
  if (gs == max_sweeps) return max_sweeps;
  for (int k = 0; k < kXt; ++k) kr[0][k] = xb[k] * (1.0 - xa[k]);
  return gs + bicgstab51(tb, xa, xs, false, kr);
}

// Adjoint of the fixed point: lam = CH + C^T (XA o lam), Gauss-Seidel, with
// the same Krylov finish.
template <class V>
SYNTHE_HD int adjoint64(const JoshTab& tb, const V& xa, const V& lam, double tol, int max_sweeps,
                        double (*kr)[kXt] = nullptr) {
  for (int k = 0; k < kXt; ++k) lam[k] = tb.ch[k];
  const int gs = kr && max_sweeps > kGsSweeps ? kGsSweeps : max_sweeps;
  for (int it = 0; it < gs; ++it) {
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
  if (gs == max_sweeps) return max_sweeps;
  for (int k = 0; k < kXt; ++k) kr[0][k] = tb.ch[k];
  return gs + bicgstab51(tb, xa, lam, true, kr);
}

// Per-depth inputs of one total-flux solve in increasing rho order (n >= 4).
template <int kMax>
struct Column {
  int n;
  double rho[kMax], abtot[kMax], alpha[kMax], snubar[kMax];
};

// Work space of the solve (caller-provided, e.g. private memory); the two
// branches' arrays share storage, and d0..d3 are the dual-number arrays of
// whichever pass runs.
template <int kMax>
struct Work {
  double taunu[kMax], pa[kMax], pb[kMax], pc[kMax], g_tau[kMax];
  struct Xtau {
    double xsbar[kXt], xalpha[kXt], xb[kXt], xa[kXt], xs[kXt], lam[kXt], gxb[kXt], gxa[kXt];
    int base[kXt];  // MAP1 bracket of each XTAU point
    Dual out1[kXt], out2[kXt];
    double kr[6][kXt];  // Krylov vectors of the fixed point's strong-scattering finish
  };
  struct Label401 {
    double lam[kMax], tmp[kMax], band[kMax][5];
  };
  union {
    Xtau x;
    Label401 s;
  };
  Dual d0[kMax], d1[kMax], d2[kMax], d3[kMax];
  // the fixed point's vectors XS, XA and its adjoint: x's own arrays after
  // bind_private(), or storage the caller binds (e.g. workgroup memory)
  Vec xs, xa, lam;
  SYNTHE_HD void bind_private() {
    xs = Vec{x.xs, 1};
    xa = Vec{x.xa, 1};
    lam = Vec{x.lam, 1};
  }
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
    return josh401_t(n, w.taunu, col.snubar, col.alpha, w.pa, w.pb, w.pc, w.g_tau, nullptr, kTol401, kMax401);
  }
  *took401 = 0;
  typename Work<kMax>::Xtau& x = w.x;
  map1_t(w.taunu, col.snubar, n, tb.xtau, x.xsbar, x.base);
  map1_t(w.taunu, col.alpha, n, tb.xtau, x.xalpha);
  const double s0 = col.snubar[0] > kEps ? col.snubar[0] : kEps;
  const double a0 = col.alpha[0] < 0.0 ? 0.0 : (col.alpha[0] > 1.0 ? 1.0 : col.alpha[0]);
  for (int k = 0; k < kXt; ++k) {
    double xb = x.xsbar[k] > kEps ? x.xsbar[k] : kEps;
    double xa = x.xalpha[k] < 0.0 ? 0.0 : (x.xalpha[k] > 1.0 ? 1.0 : x.xalpha[k]);
    if (tb.xtau[k] < w.taunu[0]) {
      xb = s0;
      xa = a0;
    }
    x.xb[k] = xb;
    w.xa[k] = xa;
  }
  fixed_point64(tb, x.xb, w.xa, w.xs, tol, 400, x.kr);
  double f = 0.0;
  for (int k = 0; k < kXt; ++k) f += tb.ch[k] * w.xs[k];
  return f;
}

// The index j = r (mod period) in [lo, lo + period).
SYNTHE_HD int seeded_in(int lo, int r, int period) { return lo + ((r - lo) % period + period) % period; }

// The vjp in phases.  Each colour pass r writes only the entries j = r
// (mod 4, or mod 5 for the label-401 Jacobians), so the passes of one column
// can run in any order or concurrently; total_flux_vjp below runs them in
// sequence.  Outputs g_* are accumulated unscaled (dF/d input); the caller
// zeroes them first.

// Suffix sums of g_tau: g_tau[i] <- sum_{m >= i} dF/dTAUNU[m].
template <int kMax>
SYNTHE_HD void tau_suffix(int n, Work<kMax>& w) {
  for (int i = n - 2; i >= 0; --i) w.g_tau[i] += w.g_tau[i + 1];
}

// INTEG, colour r: g_abtot[j] += dF/dabtot[j] for j = r (mod 4), given the
// suffix sums in w.g_tau.  TAUNU[0] = abtot[0] rho[0] and TAUNU[i+1] =
// TAUNU[i] + u_i, where u_i depends on abtot[i-1 .. i+2] through PARCOE.
template <int kMax>
SYNTHE_HD void integ_pass(const Column<kMax>& col, Work<kMax>& w, int r, double* g_abtot) {
  const int n = col.n;
  if (r == 0) g_abtot[0] += w.g_tau[0] * col.rho[0];
  for (int i = 0; i < n; ++i) w.d0[i] = Dual(col.abtot[i], (i & 3) == r ? 1.0 : 0.0);
  parcoe_t(w.d0, col.rho, n, w.d1, w.d2, w.d3);
  for (int i = 0; i + 1 < n; ++i) {
    const int j = seeded_in(i - 1, r, 4);
    if (j < 0 || j >= n) continue;
    const double r0 = col.rho[i], r1 = col.rho[i + 1];
    // the derivative part of INTEG's increment (pa + pb (r0 + r1) / 2 + pc (...) / 3) (r1 - r0)
    const double dterm = w.d1[i].d + 0.5 * w.d2[i].d * (r1 + r0) +
                         (w.d3[i].d / 3.0) * ((r1 + r0) * r1 + r0 * r0);
    g_abtot[j] += w.g_tau[i + 1] * dterm * (r1 - r0);
  }
}

// One label-401 update PHI(S) = (1 - alpha) snubar + alpha (J(S) + S) with
// J = DERIV(TAUNU, DERIV(TAUNU, S) / 3) in dual numbers, TAUNU in d0 and S
// in d1: d3 <- J(S).
template <int kMax>
SYNTHE_HD void josh401_j(int n, Work<kMax>& w) {
  deriv_t(w.d0, w.d1, n, w.d2);
  for (int j = 0; j < n; ++j) w.d2[j] = w.d2[j] / 3.0;
  deriv_t(w.d0, w.d2, n, w.d3);
}

// XTAU branch, after total_flux64: the fixed point's adjoint and the
// sensitivities gxb, gxa of F to MAP1's outputs (into w.x); returns the
// sensitivities to snubar[0] and alpha[0] through the surface clamp.
template <int kMax>
SYNTHE_HD void xtau_adjoint(const Column<kMax>& col, const JoshTab& tb, Work<kMax>& w, double* gs0,
                            double* ga0) {
  adjoint64(tb, w.xa, w.lam, 1e-13, 400, w.x.kr);
  double s0 = 0.0, a0 = 0.0;
  for (int k = 0; k < kXt; ++k) {
    double cx = 0.0;
    for (int m = 0; m < kXt; ++m) cx += (double)tb.coefj32[k * kXt + m] * w.xs[m];
    double gxb = w.lam[k] * (1.0 - w.xa[k]);
    double gxa = w.lam[k] * (cx - w.x.xb[k]);
    if (tb.xtau[k] < w.taunu[0]) {
      s0 += gxb;
      a0 += gxa;
      gxb = gxa = 0.0;
    } else {
      if (!(w.x.xsbar[k] > kEps)) gxb = 0.0;
      if (w.x.xalpha[k] < 0.0 || w.x.xalpha[k] > 1.0) gxa = 0.0;
    }
    w.x.gxb[k] = gxb;
    w.x.gxa[k] = gxa;
  }
  *gs0 = col.snubar[0] > kEps ? s0 : 0.0;
  *ga0 = (col.alpha[0] < 0.0 || col.alpha[0] > 1.0) ? 0.0 : a0;
}

// XTAU branch, colour r: MAP1 of snubar and alpha on the nodes TAUNU, values
// then nodes (needs w.taunu, w.x.base, gxb, gxa): g_snubar, g_alpha, g_tau
// entries j = r (mod 4).
template <int kMax>
SYNTHE_HD void xtau_pass(const Column<kMax>& col, const JoshTab& tb, Work<kMax>& w, int r, double* g_snubar,
                         double* g_alpha, double* g_tau) {
  const int n = col.n;
  for (int i = 0; i < n; ++i) {
    w.d0[i] = Dual(w.taunu[i]);
    w.d1[i] = Dual(col.snubar[i], (i & 3) == r ? 1.0 : 0.0);
  }
  map1_t(w.d0, w.d1, n, tb.xtau, w.x.out1);
  for (int i = 0; i < n; ++i) w.d1[i] = Dual(col.alpha[i], (i & 3) == r ? 1.0 : 0.0);
  map1_t(w.d0, w.d1, n, tb.xtau, w.x.out2);
  for (int k = 0; k < kXt; ++k) {
    const int j = seeded_in(w.x.base[k] - 3, r, 4);
    if (j < 0 || j >= n) continue;
    g_snubar[j] += w.x.gxb[k] * w.x.out1[k].d;
    g_alpha[j] += w.x.gxa[k] * w.x.out2[k].d;
  }
  for (int i = 0; i < n; ++i) {
    w.d0[i] = Dual(w.taunu[i], (i & 3) == r ? 1.0 : 0.0);
    w.d1[i] = Dual(col.snubar[i]);
  }
  map1_t(w.d0, w.d1, n, tb.xtau, w.x.out1);
  for (int i = 0; i < n; ++i) w.d1[i] = Dual(col.alpha[i]);
  map1_t(w.d0, w.d1, n, tb.xtau, w.x.out2);
  for (int k = 0; k < kXt; ++k) {
    const int j = seeded_in(w.x.base[k] - 3, r, 4);
    if (j < 0 || j >= n) continue;
    g_tau[j] += w.x.gxb[k] * w.x.out1[k].d + w.x.gxa[k] * w.x.out2[k].d;
  }
}

// Label-401 branch, colour r of 5: the band entries band[j][i - j + 2] =
// dPHI_j / dS[i] for i = r (mod 5), at the converged iterate S (w.pa).
template <int kMax>
SYNTHE_HD void l401_band(const Column<kMax>& col, Work<kMax>& w, int r, double (*band)[5]) {
  const int n = col.n;
  for (int i = 0; i < n; ++i) {
    w.d0[i] = Dual(w.taunu[i]);
    w.d1[i] = Dual(w.pa[i], i % 5 == r ? 1.0 : 0.0);
  }
  josh401_j(n, w);
  for (int j = 0; j < n; ++j) {
    const int i = seeded_in(j - 2, r, 5);
    if (i >= 0 && i < n) band[j][i - j + 2] = col.alpha[j] * (w.d3[j].d + w.d1[j].d);
  }
}

// Label-401 branch, given the whole band (zero outside the column): the
// adjoint lam = dF/dS + (dPHI/dS)^T lam (the forward's iteration transposed,
// into w.s.lam), g_snubar and g_alpha for every j, and the direct TAUNU terms
// of F = (S[1] - S[0]) / (TAUNU[1] - TAUNU[0]) / 3 into g_tau.
template <int kMax>
SYNTHE_HD void l401_adjoint(const Column<kMax>& col, Work<kMax>& w, const double (*band)[5], double* g_snubar,
                            double* g_alpha, double* g_tau) {
  const int n = col.n;
  const double* S = w.pa;
  const double dt = w.taunu[1] - w.taunu[0];
  const double dS = S[1] - S[0];
  g_tau[0] += dS / (3.0 * dt * dt);
  g_tau[1] += -dS / (3.0 * dt * dt);
  // J(S) in double (the value parts of the dual passes)
  for (int i = 0; i < n; ++i) w.s.tmp[i] = 0.0;
  deriv_t(w.taunu, S, n, w.pb);
  for (int j = 0; j < n; ++j) w.pb[j] = w.pb[j] / 3.0;
  deriv_t(w.taunu, w.pb, n, w.pc);
  for (int i = 0; i < n; ++i) w.s.lam[i] = 0.0;
  w.s.lam[0] = -1.0 / (3.0 * dt);
  w.s.lam[1] = 1.0 / (3.0 * dt);
  for (int it = 0; it < 4 * kMax401; ++it) {
    double err = 0.0;
    for (int i = 0; i < n; ++i) {
      double s = i == 0 ? -1.0 / (3.0 * dt) : (i == 1 ? 1.0 / (3.0 * dt) : 0.0);
      for (int j = i - 2; j <= i + 2; ++j)
        if (j >= 0 && j < n) s += band[j][i - j + 2] * w.s.lam[j];
      w.s.tmp[i] = s;
    }
    for (int i = 0; i < n; ++i) {
      const double den = fabs(w.s.tmp[i]) > kEps ? fabs(w.s.tmp[i]) : kEps;
      err += fabs(w.s.tmp[i] - w.s.lam[i]) / den;
      w.s.lam[i] = w.s.tmp[i];
    }
    if (err < 1e-13) break;
  }
  for (int j = 0; j < n; ++j) {
    g_snubar[j] += w.s.lam[j] * (1.0 - col.alpha[j]);
    g_alpha[j] += (w.pc[j] + S[j] - col.snubar[j]) * w.s.lam[j];  // dPHI_j/dalpha_j lam_j
  }
}

// Label-401 branch, colour r of 5: TAUNU enters PHI through both DERIVs
// (5-banded): g_tau entries i = r (mod 5), given lam (w.s.lam) and S.
template <int kMax>
SYNTHE_HD void l401_tau(const Column<kMax>& col, Work<kMax>& w, int r, double* g_tau) {
  const int n = col.n;
  for (int i = 0; i < n; ++i) {
    w.d0[i] = Dual(w.taunu[i], i % 5 == r ? 1.0 : 0.0);
    w.d1[i] = Dual(w.pa[i]);
  }
  josh401_j(n, w);
  for (int j = 0; j < n; ++j) {
    const int i = seeded_in(j - 2, r, 5);
    if (i >= 0 && i < n) g_tau[i] += w.s.lam[j] * col.alpha[j] * w.d3[j].d;
  }
}

// g_abtot, g_alpha, g_snubar [n] = v * dF/d(input) for every input; returns F.
template <int kMax>
SYNTHE_HD double total_flux_vjp(const Column<kMax>& col, const JoshTab& tb, Work<kMax>& w, double v,
                                double* g_abtot, double* g_alpha, double* g_snubar) {
  const int n = col.n;
  int took401 = 0;
  const double F = total_flux64(col, tb, w, 1e-13, &took401);
  for (int i = 0; i < n; ++i) g_abtot[i] = g_alpha[i] = g_snubar[i] = w.g_tau[i] = 0.0;
  if (took401) {
    for (int j = 0; j < n; ++j)
      for (int o = 0; o < 5; ++o) w.s.band[j][o] = 0.0;
    for (int r = 0; r < 5; ++r) l401_band(col, w, r, w.s.band);
    l401_adjoint(col, w, w.s.band, g_snubar, g_alpha, w.g_tau);
    for (int r = 0; r < 5; ++r) l401_tau(col, w, r, w.g_tau);
  } else {
    xtau_adjoint(col, tb, w, &g_snubar[0], &g_alpha[0]);
    for (int r = 0; r < 4; ++r) xtau_pass(col, tb, w, r, g_snubar, g_alpha, w.g_tau);
  }
  tau_suffix(n, w);
  for (int r = 0; r < 4; ++r) integ_pass(col, w, r, g_abtot);
  for (int i = 0; i < n; ++i) {
    g_abtot[i] *= v;
    g_alpha[i] *= v;
    g_snubar[i] *= v;
  }
  return F;
}

}  // namespace adj
}  // namespace synthe

// Streaming form of the JOSH set-up: INTEG (PARCOE + quadrature) and MAP1
// from TAUNU onto the 51 XTAU points, in one pass over depth with O(1) state.
//
// physics/josh_solver.py:95-189 (_parcoe, _integ) and 220-353 (_map1_kernel)
// restated so that no per-depth array is stored: PARCOE's smoothed
// coefficients at interval i depend only on f[i-1..i+2], and MAP1's result at
// a point depends only on its bracket (the first TAUNU index above it) --
// the Fortran code's reuse of the previous forward parabola as the next
// backward parabola evaluates the identical expression on the identical
// operands, so recomputing it is bit-for-bit the same.  Every expression
// keeps the reference's operation order (compile with -ffp-contract=off).
//
// The header compiles as HIP device code and as plain C++ (for CPU checks).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define SYNTHE_HD __host__ __device__ __forceinline__
#else
#define SYNTHE_HD inline
#endif

namespace synthe {
namespace rt {

constexpr int kXt = 51;

struct Coef {
  double a, b, c;
};

// Pass-1 PARCOE coefficients at m (0 <= m < n), n >= 4, from f/x at m-1..m+1.
// fm1/xm1 are unused for m == 0 and fp1/xp1 for m == n - 1.
SYNTHE_HD Coef parcoe_p1(int m, int n, double fm1, double f0, double fp1, double xm1, double x0,
                         double xp1) {
  Coef r;
  if (m == 0) {
    double denom = xp1 - x0;
    r.c = 0.0;
    r.b = denom != 0.0 ? (fp1 - f0) / denom : 0.0;
    r.a = f0 - x0 * r.b;
  } else if (m == n - 1) {
    double denom = x0 - xm1;
    r.c = 0.0;
    r.b = denom != 0.0 ? (f0 - fm1) / denom : 0.0;
    r.a = f0 - x0 * r.b;
  } else if (m == 1 || (m == 2 && n > 3)) {
    // boundary overrides applied after the general loop (josh_solver.py:147-158)
    double denom = xp1 - x0;
    r.c = 0.0;
    r.b = denom != 0.0 ? (fp1 - f0) / denom : 0.0;
    r.a = f0 - x0 * r.b;
  } else {
    double denom = x0 - xm1;
    double d = denom != 0.0 ? (f0 - fm1) / denom : 0.0;
    double denom1 = (xp1 - x0) * (xp1 - xm1);
    double term1 = denom1 != 0.0 ? fp1 / denom1 : 0.0;
    double denom2 = xp1 - xm1;
    double denom3 = xp1 - x0;
    double denom4 = x0 - xm1;
    double part = 0.0;
    if (denom4 != 0.0) {
      double t1 = denom2 != 0.0 ? fm1 / denom2 : 0.0;
      double t2 = denom3 != 0.0 ? f0 / denom3 : 0.0;
      part = (t1 - t2) / denom4;
    }
    r.c = term1 + part;
    r.b = d - (x0 + xm1) * r.c;
    r.a = fm1 - xm1 * d + x0 * xm1 * r.c;
  }
  return r;
}

// Final PARCOE coefficients at m (smoothing of josh_solver.py:160-167 and the
// a[n-2] = a[n-1] override), from the pass-1 values at m and m + 1.
SYNTHE_HD Coef parcoe_final(int m, int n, Coef p, Coef pn) {
  if (m == n - 2) return pn;
  if (m == 0 || m == n - 1) return p;
  if (p.c == 0.0) return p;
  double dn = fabs(pn.c) + fabs(p.c);
  double wt = dn > 0.0 ? fabs(pn.c) / dn : 0.0;
  Coef r;
  r.a = pn.a + wt * (p.a - pn.a);
  r.b = pn.b + wt * (p.b - pn.b);
  r.c = pn.c + wt * (p.c - pn.c);
  return r;
}

// One INTEG step (josh_solver.py:184-188).
SYNTHE_HD double integ_step(double tau, Coef s, double x0, double x1) {
  double dx = x1 - x0;
  double term = s.a + 0.5 * s.b * (x1 + x0);
  term += (s.c / 3.0) * ((x1 + x0) * x1 + x0 * x0);
  return tau + term * dx;
}

// MAP1 evaluation at x for bracket p (0-based first TAUNU index > x, p >= 1),
// given TAUNU and f at p-2..p+1 (t[0..3] = p-2..p+1; entries outside [0, n)
// are not read).  nold = n.
SYNTHE_HD double map1_eval(double x, int p, int nold, const double* t, const double* f) {
  // 1-based l = p + 1; t[1] = tau[p-1], t[2] = tau[p], t[3] = tau[p+1], t[0] = tau[p-2]
  double a, b, c;
  int l = p + 1;
  if (l == 2 || l == 3) {
    c = 0.0;
    b = (f[2] - f[1]) / (t[2] - t[1]);
    a = f[2] - t[2] * b;
  } else {
    // backward parabola through l-2, l-1, l (1-based) = t[0], t[1], t[2]
    double d = (f[1] - f[0]) / (t[1] - t[0]);
    double cbac = f[2] / ((t[2] - t[1]) * (t[2] - t[0])) +
                  (f[0] / (t[2] - t[0]) - f[1] / (t[2] - t[1])) / (t[1] - t[0]);
    double bbac = d - (t[1] + t[0]) * cbac;
    double abac = f[0] - t[0] * d + t[1] * t[0] * cbac;
    if (l == nold) {
      c = cbac;
      b = bbac;
      a = abac;
    } else {
      // forward parabola through l-1, l, l+1 = t[1], t[2], t[3]
      double d2 = (f[2] - f[1]) / (t[2] - t[1]);
      double cfor = f[3] / ((t[3] - t[2]) * (t[3] - t[1])) +
                    (f[1] / (t[3] - t[1]) - f[2] / (t[3] - t[2])) / (t[2] - t[1]);
      double bfor = d2 - (t[2] + t[1]) * cfor;
      double afor = f[1] - t[1] * d2 + t[2] * t[1] * cfor;
      double wt = 0.0;
      if (fabs(cfor) != 0.0) wt = fabs(cfor) / (fabs(cfor) + fabs(cbac));
      a = afor + wt * (abac - afor);
      b = bfor + wt * (bbac - bfor);
      c = cfor + wt * (cbac - cfor);
    }
  }
  return a + (b + c * x) * x;
}

// Linear extrapolation past the last TAUNU (MAP1 label 30 with l > nold).
SYNTHE_HD double map1_extrap(double x, double t0, double t1, double f0, double f1) {
  double b = (f1 - f0) / (t1 - t0);
  double a = f1 - t1 * b;
  return a + (b + 0.0 * x) * x;
}

// INTEG + MAP1 of one column as a stepper, fed one depth at a time in
// increasing rho order (abtot f, rho x, snubar, alpha): start() with depths 0
// and 1, step() for j = 1 .. n-1 with the values of depth j+1 (zeros past the
// last), then finish().  Out receives the MAP1 results
//   void put(int k, double xsbar, double xalpha)
// for k = 0..50 in increasing order.  Several columns can advance in
// lockstep over one read of the inputs.  Requires n >= 4.
struct IntegMap1 {
  int n, k;
  double fb, fc, xb, xc, sb_n, al_n, tau, x;
  // MAP1 window: tau/snubar/alpha at p-2..p+1 (slot 3 = newest)
  double T[4], S[4], A[4];
  Coef prev;
  const double* xtau;

  SYNTHE_HD double start(int n_, const double* xtau_, double f0, double x0, double sb0, double al0, double f1,
                         double x1, double sb1, double al1) {
    n = n_;
    xtau = xtau_;
    fb = f0;
    xb = x0;
    fc = f1;
    xc = x1;
    sb_n = sb1;
    al_n = al1;
    tau = fb * xb;
    T[3] = tau;
    S[3] = sb0;
    A[3] = al0;
    T[0] = T[1] = T[2] = 0.0;
    S[0] = S[1] = S[2] = 0.0;
    A[0] = A[1] = A[2] = 0.0;
    prev = parcoe_p1(0, n, 0.0, fb, fc, 0.0, xb, xc);
    k = 0;
    x = xtau[0];
    return tau;  // TAUNU(0), the label-401 gate
  }

  // window: fb = f[j-1], fc = f[j]; fd.. = depth j+1 (zeros past the last)
  template <class Out>
  SYNTHE_HD void step(int j, double fd, double xd, double sbd, double ald, Out& out) {
    Coef cur = parcoe_p1(j, n, fb, fc, fd, xb, xc, xd);
    Coef s = parcoe_final(j - 1, n, prev, cur);
    tau = integ_step(tau, s, xb, xc);
    prev = cur;
    fb = fc;
    xb = xc;
    fc = fd;
    xc = xd;
    T[0] = T[1]; T[1] = T[2]; T[2] = T[3]; T[3] = tau;
    S[0] = S[1]; S[1] = S[2]; S[2] = S[3]; S[3] = sb_n;
    A[0] = A[1]; A[1] = A[2]; A[2] = A[3]; A[3] = al_n;
    sb_n = sbd;
    al_n = ald;
    // bracket p = j - 1 (needs tau[p+1] = tau[j]) for points below tau[p]
    const int p = j - 1;
    if (p >= 1) {
      while (k < kXt && x < T[2]) {
        out.put(k, map1_eval(x, p, n, T, S), map1_eval(x, p, n, T, A));
        ++k;
        x = k < kXt ? xtau[k] : 0.0;
      }
    }
  }

  template <class Out>
  SYNTHE_HD void finish(Out& out) {
    // bracket p = n - 1 (l == nold): window slots 1..3 hold n-3..n-1
    double Tb[4] = {T[1], T[2], T[3], 0.0};
    double Sb[4] = {S[1], S[2], S[3], 0.0};
    double Ab[4] = {A[1], A[2], A[3], 0.0};
    while (k < kXt && x < T[3]) {
      out.put(k, map1_eval(x, n - 1, n, Tb, Sb), map1_eval(x, n - 1, n, Tb, Ab));
      ++k;
      x = k < kXt ? xtau[k] : 0.0;
    }
    while (k < kXt) {
      out.put(k, map1_extrap(x, T[2], T[3], S[2], S[3]), map1_extrap(x, T[2], T[3], A[2], A[3]));
      ++k;
      x = k < kXt ? xtau[k] : 0.0;
    }
  }
};

// One column through the stepper.  Src must provide
//   void get(int j, double* abtot, double* rho, double* snubar, double* alpha)
// for 0 <= j < n, in increasing rho order.  Returns TAUNU(0) through *tau0.
template <class Src, class Out>
SYNTHE_HD void integ_map1_stream(int n, const double* xtau, Src& src, Out& out, double* tau0) {
  double f0, x0, sb0, al0, f1, x1, sb1, al1;
  src.get(0, &f0, &x0, &sb0, &al0);
  src.get(1, &f1, &x1, &sb1, &al1);
  IntegMap1 st;
  *tau0 = st.start(n, xtau, f0, x0, sb0, al0, f1, x1, sb1, al1);
  for (int j = 1; j < n; ++j) {
    double fd = 0.0, xd = 0.0, sbd = 0.0, ald = 0.0;
    if (j + 1 < n) src.get(j + 1, &fd, &xd, &sbd, &ald);
    st.step(j, fd, xd, sbd, ald, out);
  }
  st.finish(out);
}

}  // namespace rt
}  // namespace synthe

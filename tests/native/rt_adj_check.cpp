// CPU build of the RT adjoint (kurucz_py_joss_submission_amd/csrc/rt_adjoint.h)
// for checking it against finite differences, the oracle, and forward-mode
// derivatives taken one input at a time.  Test infrastructure.
#include <string.h>

#include "rt_adjoint.h"

namespace {
constexpr int kMax = 128;
using Col = synthe::adj::Column<kMax>;
using Wk = synthe::adj::Work<kMax>;
using synthe::adj::Dual;

void fill(Col& c, int n, const double* rho, const double* abtot, const double* alpha, const double* snubar) {
  c.n = n;
  for (int i = 0; i < n; ++i) {
    c.rho[i] = rho[i];
    c.abtot[i] = abtot[i];
    c.alpha[i] = alpha[i];
    c.snubar[i] = snubar[i];
  }
}
}  // namespace

extern "C" {
double adj_flux64(int n, const double* rho, const double* abtot, const double* alpha, const double* snubar,
                  const double* xtau, const double* ch, const float* coefj32, double tol, int* took401) {
  static Col c;
  static Wk w;
  fill(c, n, rho, abtot, alpha, snubar);
  synthe::adj::JoshTab tb{xtau, ch, coefj32};
  w.bind_private();
  return synthe::adj::total_flux64(c, tb, w, tol, took401);
}

double adj_vjp(int n, const double* rho, const double* abtot, const double* alpha, const double* snubar,
               const double* xtau, const double* ch, const float* coefj32, double v, double* g_abtot,
               double* g_alpha, double* g_snubar) {
  static Col c;
  static Wk w;
  fill(c, n, rho, abtot, alpha, snubar);
  synthe::adj::JoshTab tb{xtau, ch, coefj32};
  w.bind_private();
  return synthe::adj::total_flux_vjp(c, tb, w, v, g_abtot, g_alpha, g_snubar);
}

// The same vjp by forward mode, one input at a time: dual numbers through
// INTEG and MAP1 (or through the whole label-401 iteration), contracted with
// the fixed point's adjoint in the XTAU branch.  O(n^2) per column.
double adj_vjp_by_inputs(int n, const double* rho, const double* abtot, const double* alpha,
                         const double* snubar, const double* xtau, const double* ch, const float* coefj32,
                         double v, double* g_abtot, double* g_alpha, double* g_snubar) {
  namespace A = synthe::adj;
  static Col c;
  static Wk w;
  static Dual in[kMax], al[kMax], sb[kMax], tau[kMax], pa[kMax], pb[kMax], pc[kMax], f[kMax];
  static Dual s[kMax], h[kMax], jn[kMax], t[kMax], o1[A::kXt], o2[A::kXt];
  fill(c, n, rho, abtot, alpha, snubar);
  A::JoshTab tb{xtau, ch, coefj32};
  w.bind_private();
  int took401 = 0;
  const double F = A::total_flux64(c, tb, w, 1e-13, &took401);
  if (took401) {
    for (int q = 0; q < 3; ++q) {
      double* g = q == 0 ? g_abtot : (q == 1 ? g_alpha : g_snubar);
      for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
          in[i] = Dual(c.abtot[i], q == 0 && i == j ? 1.0 : 0.0);
          al[i] = Dual(c.alpha[i], q == 1 && i == j ? 1.0 : 0.0);
          sb[i] = Dual(c.snubar[i], q == 2 && i == j ? 1.0 : 0.0);
        }
        A::integ_t(in, c.rho, n, tau, pa, pb, pc);
        Dual fl = A::josh401_t(n, tau, sb, al, s, h, jn, t, nullptr, A::kTol401, A::kMax401);
        g[j] = v * fl.d;
      }
    }
    return F;
  }
  A::adjoint64(tb, w.xa, w.lam, 1e-13, 400, w.x.kr);
  double gs0 = 0.0, ga0 = 0.0;
  for (int k = 0; k < A::kXt; ++k) {
    double cx = 0.0;
    for (int m = 0; m < A::kXt; ++m) cx += (double)tb.coefj32[k * A::kXt + m] * w.xs[m];
    double gxb = w.lam[k] * (1.0 - w.xa[k]);
    double gxa = w.lam[k] * (cx - w.x.xb[k]);
    if (tb.xtau[k] < w.taunu[0]) {
      gs0 += gxb;
      ga0 += gxa;
      gxb = gxa = 0.0;
    } else {
      if (!(w.x.xsbar[k] > A::kEps)) gxb = 0.0;
      if (w.x.xalpha[k] < 0.0 || w.x.xalpha[k] > 1.0) gxa = 0.0;
    }
    w.x.gxb[k] = gxb;
    w.x.gxa[k] = gxa;
  }
  if (!(c.snubar[0] > A::kEps)) gs0 = 0.0;
  if (c.alpha[0] < 0.0 || c.alpha[0] > 1.0) ga0 = 0.0;
  for (int i = 0; i < n; ++i) tau[i] = Dual(w.taunu[i]);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) f[i] = Dual(c.snubar[i], i == j ? 1.0 : 0.0);
    A::map1_t(tau, f, n, tb.xtau, o1);
    double acc = j == 0 ? gs0 : 0.0;
    for (int k = 0; k < A::kXt; ++k) acc += w.x.gxb[k] * o1[k].d;
    g_snubar[j] = v * acc;
    for (int i = 0; i < n; ++i) f[i] = Dual(c.alpha[i], i == j ? 1.0 : 0.0);
    A::map1_t(tau, f, n, tb.xtau, o1);
    acc = j == 0 ? ga0 : 0.0;
    for (int k = 0; k < A::kXt; ++k) acc += w.x.gxa[k] * o1[k].d;
    g_alpha[j] = v * acc;
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) in[i] = Dual(c.abtot[i], i == j ? 1.0 : 0.0);
    A::integ_t(in, c.rho, n, tau, pa, pb, pc);
    for (int i = 0; i < n; ++i) f[i] = Dual(c.snubar[i]);
    A::map1_t(tau, f, n, tb.xtau, o1);
    for (int i = 0; i < n; ++i) f[i] = Dual(c.alpha[i]);
    A::map1_t(tau, f, n, tb.xtau, o2);
    double acc = 0.0;
    for (int k = 0; k < A::kXt; ++k) acc += w.x.gxb[k] * o1[k].d + w.x.gxa[k] * o2[k].d;
    g_abtot[j] = v * acc;
  }
  return F;
}

// MAP1 of fold on xold onto xtau: values, and d(values)/d(fold[j]) by dual numbers.
void adj_map1(int n, const double* xold, const double* fold, const double* xtau, int j, double* val_out,
              double* der_out) {
  static Dual x[128], f[128], o[51];
  for (int i = 0; i < n; ++i) {
    x[i] = Dual(xold[i]);
    f[i] = Dual(fold[i], i == j ? 1.0 : 0.0);
  }
  synthe::adj::map1_t(x, f, n, xtau, o);
  for (int k = 0; k < 51; ++k) {
    val_out[k] = o[k].v;
    der_out[k] = o[k].d;
  }
}
}

// CPU build of the RT adjoint (kurucz_py_joss_submission_amd/csrc/rt_adjoint.h)
// for checking it against finite differences and the oracle.  Test infrastructure.
#include <string.h>

#include "rt_adjoint.h"

namespace {
constexpr int kMax = 128;
using Col = synthe::adj::Column<kMax>;
using Wk = synthe::adj::Work<kMax>;

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
  return synthe::adj::total_flux64(c, tb, w, tol, took401);
}

double adj_vjp(int n, const double* rho, const double* abtot, const double* alpha, const double* snubar,
               const double* xtau, const double* ch, const float* coefj32, double v, double* g_abtot,
               double* g_alpha, double* g_snubar) {
  static Col c;
  static Wk w;
  fill(c, n, rho, abtot, alpha, snubar);
  synthe::adj::JoshTab tb{xtau, ch, coefj32};
  return synthe::adj::total_flux_vjp(c, tb, w, v, g_abtot, g_alpha, g_snubar);
}
}

extern "C" {
// MAP1 of fold on xold onto xtau: values, and d(values)/d(fold[j]) by dual numbers.
void adj_map1(int n, const double* xold, const double* fold, const double* xtau, int j, double* val_out,
              double* der_out) {
  static synthe::adj::Dual x[128], f[128], o[51];
  for (int i = 0; i < n; ++i) {
    x[i] = synthe::adj::Dual(xold[i]);
    f[i] = synthe::adj::Dual(fold[i], i == j ? 1.0 : 0.0);
  }
  synthe::adj::map1_t(x, f, n, xtau, o);
  for (int k = 0; k < 51; ++k) {
    val_out[k] = o[k].v;
    der_out[k] = o[k].d;
  }
}
}

// CPU build of the device RT core (kurucz_py_joss_submission_amd/csrc/rt_core.h)
// for checking it against the oracle without a GPU.  Test infrastructure.
#include <math.h>
#include <string.h>
#include <vector>

#include "rt_core.h"

namespace {
struct ArrSrc {
  const double *abtot, *rho, *snubar, *alpha;
  void get(int j, double* f, double* x, double* s, double* a) {
    *f = abtot[j];
    *x = rho[j];
    *s = snubar[j];
    *a = alpha[j];
  }
};
struct ArrOut {
  double *xsbar, *xalpha;
  int last = -1;
  bool ordered = true;
  void put(int k, double sb, double al) {
    if (k != last + 1) ordered = false;
    last = k;
    xsbar[k] = sb;
    xalpha[k] = al;
  }
};
}  // namespace

extern "C" {
// Streaming INTEG + MAP1 for one column; returns TAUNU(0), -1 if puts were out of order.
double rtc_integ_map1(int n, const double* xtau, const double* abtot, const double* rho,
                      const double* snubar, const double* alpha, double* xsbar, double* xalpha) {
  ArrSrc s{abtot, rho, snubar, alpha};
  ArrOut o{xsbar, xalpha};
  double tau0 = 0.0;
  synthe::rt::integ_map1_stream(n, xtau, s, o, &tau0);
  if (!o.ordered || o.last != 50) return -1.0;
  return tau0;
}
}

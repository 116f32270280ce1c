// Device-side numerics of the synthe hot path (gfx950).
//
// Each function restates one routine of the reference (paths relative to its
// synthe_py/ package) with the same operation order.  Kernels that include this
// header are compiled with -ffp-contract=off so every multiply and add rounds
// separately, as the reference's Python/Numba arithmetic does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace synthe {

constexpr int kNxtau = 51;
constexpr int kVoigtN = 2001;
constexpr int kFastexN = 1001;
constexpr long kMaxProfileSteps = 1000000;
constexpr int kMaxDepth = 128;  // per-wavelength RT arrays are sized for this

constexpr double kHPlanck = 6.62607015e-27;
constexpr double kKBoltz = 1.380649e-16;
constexpr double kCLightCm = 2.99792458e10;
constexpr double kCLightNm = 2.99792458e17;
constexpr double kNmToCm = 1e-7;
constexpr double kEps38 = 1.0e-38;

// Tables resident in device memory, shared by all kernels of a context.
struct DeviceTables {
  const double* h0;      // [2001] exp(-v^2)
  const double* h1;      // [2001] MAP1 of TABH1
  const double* h2;      // [2001] (1-2v^2) exp(-v^2)
  const double* extab;   // [1001] exp(-i)
  const double* extabf;  // [1001] exp(-i/1000)
  const double* xtau;    // [51]
  const double* ch;      // [51]
  const float* coefj32;  // [51*51] row-major, float32
  const float* diag32;   // [51] float32 diagonal of COEFJ
};

// physics/tables.py:29-42 (ExpTables.fast_ex)
__device__ __forceinline__ double fastex(double x, const double* extab, const double* extabf) {
  if (x <= 0.0) return x == 0.0 ? 1.0 : exp(-x);
  double fl = floor(x);
  if (fl >= (double)kFastexN) return exp(-x);
  int i = (int)fl;
  double frac = x - (double)i;
  int j = (int)(frac * 1000.0 + 0.5);
  j = j < 0 ? 0 : (j > kFastexN - 1 ? kFastexN - 1 : j);
  return extab[i] * extabf[j];
}

// physics/voigt_jit.py:16-69 (voigt_profile_jit)
__device__ __forceinline__ double voigt(double v, double a, const double* h0t, const double* h1t,
                                        const double* h2t) {
  double av = fabs(v);
  double fiv = av * 200.0 + 0.5;
  int iv = fiv >= (double)(kVoigtN - 1) ? kVoigtN - 1 : (int)fiv;
  if (a < 0.2) {
    if (av > 10.0) return 0.5642 * a / (v * v);
    return (h2t[iv] * a + h1t[iv]) * a + h0t[iv];
  } else if (a > 1.4 || (a + av) > 3.2) {
    double aa = a * a, vv = v * v;
    double u = (aa + vv) * 1.4142;
    double vo = a * 0.79788 / u;
    if (a <= 100.0) {
      double aau = aa / u, vvu = vv / u, uu = u * u;
      vo = ((((aau - 10.0 * vvu) * aau * 3.0 + 15.0 * vvu * vvu) + 3.0 * vv - aa) / uu + 1.0) * vo;
    }
    return vo;
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

// Wing profile value at offset k >= 1 (line_opacity.py:470-487 near wing,
// 541-570 phase 3).  The a < 0.2 table index is the k-fold sequential sum
// 0.5 + tabstep + ... + tabstep, as the reference accumulates TABI.
struct WingParams {
  double kappa0, adamp, dvoigt, x_far;
  int32_t c, maxstep, n10dop, use_far;
};

// int(tabi) after the reference's k sequential adds tabi = 0.5; tabi += tabstep
// (physics/line_opacity.py:468-490).  The closed form 0.5 + k * tabstep is
// within (k/2 + 1) * 2^-52 * t of the sequential sum, so whenever t is farther
// than (k + 4) * 2^-52 * t from an integer both truncate to the same index;
// otherwise the sequential sum is formed.
__device__ __forceinline__ long tabi_index(double tabstep, long k) {
  double t = 0.5 + (double)k * tabstep;
  if (tabstep >= 0.0 && t < 0x1p40) {
    double fl = floor(t);
    double f = t - fl;
    double eps = (double)(k + 4) * t * 0x1p-52;
    if (f > eps && 1.0 - f > eps) return (long)fl;
  }
  double s = 0.5;
  for (long i = 0; i < k; ++i) s += tabstep;
  return (long)s;
}

__device__ __forceinline__ double near_wing(const WingParams& p, long k, const double* h0t,
                                            const double* h1t, const double* h2t) {
  double x = (double)k * p.dvoigt;
  if (p.adamp < 0.2) {
    if (x > 10.0) return p.kappa0 * (0.5642 * p.adamp / (x * x));
    long idx = tabi_index(200.0 * p.dvoigt, k);
    idx = idx < 0 ? 0 : (idx > kVoigtN - 1 ? kVoigtN - 1 : idx);
    return p.kappa0 * (h0t[idx] + p.adamp * h1t[idx]);
  }
  return p.kappa0 * voigt(x, p.adamp, h0t, h1t, h2t);
}

// x / b for x in [2^-900, 2^900] and b in [1, 2^60] (the far-wing X / k^2):
// reciprocal, two Newton steps and a Markstein remainder correction, which is
// the IEEE division sequence without its operand scaling and special-value
// fix-up -- both are identities in this range -- so the quotient is the
// correctly rounded one.  Outside the range it falls back to division.
__device__ __forceinline__ double div_pos(double x, double b) {
  if (!(x >= 0x1p-900 && x <= 0x1p900 && b >= 1.0 && b <= 0x1p60)) return x / b;
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  double q = x * r;
  double rem = __builtin_fma(-b, q, x);
  return __builtin_fma(rem, r, q);
}

// The same quotient with one Newton step, for callers that guarantee the
// operand ranges above; measured bit-identical to IEEE division on the GPU
// over X in [1e-250, 1e250] and b = k^2, k <= 10^6 (synthe_diag_div_pos).
__device__ __forceinline__ double div_pos_inrange(double x, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  double q = x * r;
  return __builtin_fma(__builtin_fma(-b, q, x), r, q);
}

// x / b within 32 ulp (measured <= 19 on gfx950 for b = k^2 <= 10^12):
// hardware reciprocal (a ~23-bit approximation), one Newton step, product.
__device__ __forceinline__ double div_fast(double x, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  return x * r;
}

// x / b to single-precision accuracy (~2^-22 relative): the hardware
// reciprocal alone.
__device__ __forceinline__ double div_single(double x, double b) {
  return x * __builtin_amdgcn_rcp(b);
}

// out[p] = x / (kd + p0 + p)^2 for p < C with one reciprocal for the C
// quotients (Montgomery's simultaneous inversion).  With q_p = (kd + p0 + p)^2
// and running products P_p = q_0 ... q_p, y = x / P_{C-1}; then for p = C-1
// down to 1, x / q_p = y * P_{p-1} and y <- y * q_p = x / P_{p-1}; finally
// x / q_0 = y.  The rounding of each P_p cancels in P_{p-1} / P_p, so a
// quotient carries the reciprocal's error plus two roundings per later pixel.
// kNewton = 2 gives the reciprocal to within an ulp (x / q within 16 ulp for
// C = 8), kNewton = 1 squares the hardware reciprocal's ~2^-23 error (x / q
// within 32 ulp, as div_fast), kNewton = 0 leaves it (~2^-22).  kMask:
// out[p] = +0 unless lo_k < |k| <= hi_k, and q_p = max(k^2, 1) so the centre
// pixel k = 0 is harmless.  Every |k| must be < 2^31, |k| >= 1 unless kMask (so
// P_{C-1} <= 2^496 for C <= 8), and x must be 0 or lie in [2^-500, 2^1023]
// (or be infinite or NaN, which propagate): then no product leaves the normal
// range.
template <int C, int kNewton, bool kMask>
__device__ __forceinline__ void far_chunk(double x, double kd, int p0, double lo_k, double hi_k,
                                          double* __restrict__ out) {
  static_assert(C >= 1 && C <= 8, "chunk of 1..8 pixels");
  double q[C], P[C];
  bool keep[C];
#pragma unroll
  for (int p = 0; p < C; ++p) {
    const double k = p0 + p == 0 ? kd : kd + (double)(p0 + p);
    const double kq = k * k;  // exact: integers below 2^62
    keep[p] = !kMask || (fabs(k) > lo_k && fabs(k) <= hi_k);
    q[p] = kMask ? fmax(kq, 1.0) : kq;
    P[p] = p == 0 ? q[0] : P[p - 1] * q[p];
  }
  double r = __builtin_amdgcn_rcp(P[C - 1]);
#pragma unroll
  for (int i = 0; i < kNewton; ++i) r = __builtin_fma(r, __builtin_fma(-P[C - 1], r, 1.0), r);
  double y = x * r;
#pragma unroll
  for (int p = C - 1; p > 0; --p) {
    const double v = y * P[p - 1];
    out[p] = keep[p] ? v : 0.0;
    y *= q[p];
  }
  out[0] = keep[0] ? y : 0.0;
}

// far_chunk without a mask, each quotient added to acc[p] with one rounding.
template <int C, int kNewton>
__device__ __forceinline__ void far_chunk_acc(double x, double kd, int p0, double* __restrict__ acc) {
  static_assert(C >= 1 && C <= 8, "chunk of 1..8 pixels");
  double q[C], P[C];
#pragma unroll
  for (int p = 0; p < C; ++p) {
    const double k = p0 + p == 0 ? kd : kd + (double)(p0 + p);
    q[p] = k * k;
    P[p] = p == 0 ? q[0] : P[p - 1] * q[p];
  }
  double r = __builtin_amdgcn_rcp(P[C - 1]);
#pragma unroll
  for (int i = 0; i < kNewton; ++i) r = __builtin_fma(r, __builtin_fma(-P[C - 1], r, 1.0), r);
  double y = x * r;
#pragma unroll
  for (int p = C - 1; p > 0; --p) {
    acc[p] = __builtin_fma(y, P[p - 1], acc[p]);
    y *= q[p];
  }
  acc[0] += y;
}

// Newton steps on the shared reciprocal of the FP64 far-wing form: one step
// squares the hardware reciprocal's ~2^-23 error, which leaves each quotient
// within 32 ulp of IEEE division (as div_fast).
constexpr int kFarNewton = 1;

// Smallest X for which far_chunk's products stay normal.
constexpr double kChunkMinX = 0x1p-500;

__device__ __forceinline__ double wing_value(const WingParams& p, long k, const double* h0t,
                                             const double* h1t, const double* h2t) {
  if (p.use_far && k > p.n10dop) return div_pos(p.x_far, (double)k * (double)k);
  return near_wing(p, k, h0t, h1t, h2t);
}

// ---------------------------------------------------------------- JOSH pieces
// physics/josh_solver.py:95-167 (_parcoe), coefficients written into a, b, c.
__device__ inline void parcoe(const double* f, const double* x, int n, double* a, double* b,
                              double* c) {
  for (int i = 0; i < n; ++i) a[i] = b[i] = c[i] = 0.0;
  if (n == 0) return;
  if (n == 1) { a[0] = f[0]; return; }
  double denom = x[1] - x[0];
  b[0] = denom != 0.0 ? (f[1] - f[0]) / denom : 0.0;
  a[0] = f[0] - x[0] * b[0];
  int n1 = n - 1;
  c[n - 1] = 0.0;
  denom = x[n - 1] - x[n1 - 1];
  b[n - 1] = denom != 0.0 ? (f[n - 1] - f[n1 - 1]) / denom : 0.0;
  a[n - 1] = f[n - 1] - x[n - 1] * b[n - 1];
  if (n == 2) return;
  for (int j = 1; j < n1; ++j) {
    int j1 = j - 1;
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
  for (int j = 1; j < n1; ++j) {
    if (c[j] == 0.0) continue;
    int j1 = j + 1 < n - 1 ? j + 1 : n - 1;
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

// physics/josh_solver.py:192-217 (_deriv)
__device__ inline void deriv(const double* x, const double* f, int n, double* dfdx) {
  for (int i = 0; i < n; ++i) dfdx[i] = 0.0;
  if (n < 2) return;
  dfdx[0] = (f[1] - f[0]) / (x[1] - x[0]);
  dfdx[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
  if (n == 2) return;
  double s = x[1] != x[0] ? fabs(x[1] - x[0]) / (x[1] - x[0]) : 1.0;
  for (int j = 1; j < n - 1; ++j) {
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

// physics/josh_solver.py:220-353 (_map1_kernel), fnew for the 51 XTAU points.
__device__ inline void map1_xtau(const double* xold, const double* fold, int nold,
                                 const double* xnew, double* fnew) {
#define XO(i) xold[(i)-1]
#define FO(i) fold[(i)-1]
  for (int k = 0; k < kNxtau; ++k) fnew[k] = 0.0;
  if (nold == 0) return;
  int l = 2, ll = 0;
  double cfor = 0, bfor = 0, afor = 0, cbac = 0, bbac = 0, abac = 0, a = 0, b = 0, c = 0;
  for (int k = 1; k <= kNxtau; ++k) {
    double xk = xnew[k - 1];
    for (;;) {
      if (xk < XO(l)) {
        if (l == ll) break;
        if (l == 2 || l == 3) {
          l = nold < l ? nold : l;
          c = 0.0;
          b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
          a = FO(l) - XO(l) * b;
          ll = l;
          break;
        }
        int l1 = l - 1;
        if (l > ll + 1 || l == 4) {
          int l2 = l - 2;
          double d = (FO(l1) - FO(l2)) / (XO(l1) - XO(l2));
          cbac = FO(l) / ((XO(l) - XO(l1)) * (XO(l) - XO(l2))) +
                 (FO(l2) / (XO(l) - XO(l2)) - FO(l1) / (XO(l) - XO(l1))) / (XO(l1) - XO(l2));
          bbac = d - (XO(l1) + XO(l2)) * cbac;
          abac = FO(l2) - XO(l2) * d + XO(l1) * XO(l2) * cbac;
          if (!(l < nold)) { c = cbac; b = bbac; a = abac; ll = l; break; }
        } else {
          cbac = cfor; bbac = bfor; abac = afor;
          if (l == nold) { c = cbac; b = bbac; a = abac; ll = l; break; }
        }
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
      if (l > nold) {
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
#undef XO
#undef FO
}

}  // namespace synthe

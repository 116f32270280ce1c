// libsynthe_hip: HIP kernels for gfx950 and the C ABI of include/synthe_hip.h.
//
// Pipeline for one spectrum (all device-resident):
//   k_continuum      build_depth_continuum            [D][W]
//   k_source         frequency quantities, SLINEC     [D][W]
//   k_transp_wing    TRANSP + per-(line,depth) wing setup (probe), one thread per pair
//   k_class_count / k_class_scan / k_regions / k_class_scatter
//                    per-depth lists of wing pairs, bucketed by reach class and
//                    ordered by centre: a deterministic, atomic-free transpose
//                    from line-major to wavelength-major
//   k_gather         centre adds + wing gather per (depth, pixel), times STIM
//   k_lte            JOSH formal solve, two per wavelength, one wavelength per lane
//
// Numerics follow the reference operation by operation (see synthe_device.h);
// the build uses -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "_tables_gen.h"
#include "rt_adjoint.h"
#include "gather_queue.h"
#include "rt_core.h"
#include "synthe_device.h"
#include "hydrogen.h"
#include "synthe_hip.h"

using namespace synthe;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(SYNTHE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Reach classes: class k holds pairs with 16 * 2^(k-1) < MAXSTEP <= 16 * 2^k
// (class 0: MAXSTEP <= 16); the last class reaches 16 * 2^16 > 10^6 steps.
constexpr int kClassShift = 1;
constexpr int kClasses = 17;

// Per list entry, what every candidate test and the far-wing path read (32
// bytes, one scalar load):
//   lo, hi    pixels c - MAXSTEP and c + MAXSTEP of a far wing; lo > hi (an
//             empty range) for a pair without one;
//   nlo, nhi  the Voigt zone c - r .. c + r, r = N10DOP with a far wing and
//             MAXSTEP without;
//   xs        the far wing's X, negated when it lies below far_chunk's range
//             (those wings take one quotient per pixel);
//   cd        the centre c as a double.
// Pixel indices fit in 32 bits: n_wl <= 2^30 and centres within 10^6 pixels
// of the grid.
struct __align__(32) WingHead {
  int32_t lo, hi, nlo, nhi;
  double xs, cd;
};
__device__ __forceinline__ int head_c(const WingHead& h) { return (int)h.cd; }
constexpr int kBlock = 256;

// ---------------------------------------------------------------- host tables
// physics/tables.py:275-391 (_map1_interpolate): the MAP1 variant that builds H1TAB.
void map1_table(const double* xold, const double* fold, int nold, const double* xnew, double* fnew,
                int nnew) {
#define XO(i) xold[(i)-1]
#define FO(i) fold[(i)-1]
  int l = 2, ll = 0;
  double a = 0, b = 0, c = 0, abac = 0, bbac = 0, cbac = 0, afor = 0, bfor = 0, cfor = 0;
  for (int k = 1; k <= nnew; ++k) {
    double x = xnew[k - 1];
    while (true) {
      if (x < XO(l)) break;
      ++l;
      if (l > nold) break;
    }
    if (l > nold) {
      if (l != ll) {
        l = nold;
        c = 0.0;
        b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
        a = FO(l) - XO(l) * b;
        ll = l;
      }
      fnew[k - 1] = a + (b + c * x) * x;
      continue;
    }
    if (l == ll) {
      fnew[k - 1] = a + (b + c * x) * x;
      continue;
    }
    if (l == 2) {
      c = 0.0;
      b = (FO(l) - FO(l - 1)) / (XO(l) - XO(l - 1));
      a = FO(l) - XO(l) * b;
      ll = l;
      fnew[k - 1] = a + (b + c * x) * x;
      continue;
    }
    int l1 = l - 1;
    bool done = false;
    if (l > ll + 1 || l == 3) {
      int l2 = l - 2;
      double d = (FO(l1) - FO(l2)) / (XO(l1) - XO(l2));
      cbac = FO(l) / ((XO(l) - XO(l1)) * (XO(l) - XO(l2))) +
             (FO(l2) / (XO(l) - XO(l2)) - FO(l1) / (XO(l) - XO(l1))) / (XO(l1) - XO(l2));
      bbac = d - (XO(l1) + XO(l2)) * cbac;
      abac = FO(l2) - XO(l2) * d + XO(l1) * XO(l2) * cbac;
      if (l == nold) done = true;
    } else {
      cbac = cfor;
      bbac = bfor;
      abac = afor;
      if (l == nold) done = true;
    }
    if (done) {
      c = cbac;
      b = bbac;
      a = abac;
      ll = l;
      fnew[k - 1] = a + (b + c * x) * x;
      continue;
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
    fnew[k - 1] = a + (b + c * x) * x;
  }
#undef XO
#undef FO
}

// ------------------------------------------------------------------- kernels
struct LinesDev {
  long n;
  const double *wl, *gf, *cgf, *grad, *gstark, *gvdw, *elo;
  const int32_t *elem, *ion, *ltype;
  const uint8_t* process;
  const int64_t *center_t, *center_w;
};

struct DepthsDev {
  int n;
  const double *temp, *rhox, *rho, *xne, *txnxn, *hckt, *pop, *dop;
  int nion, nelem;
};

__device__ __forceinline__ int reach_class(int m) {
  int k = 0;
  long r = 16;
  while (m > r && k < kClasses - 1) {
    r <<= kClassShift;
    ++k;
  }
  return k;
}

// physics/continuum.py:73-303, coefficient path, one thread per wavelength.
constexpr int kContinuumDepths = 8;  // depths per thread (grid rows cover the rest)

// The edge interval of wavelength x and its three quadratic weights.
struct ContWeights {
  int e;
  double c1, c2, c3;
};

__device__ __forceinline__ ContWeights continuum_weights(double x, int E, const double* __restrict__ wledge,
                                                         const double* __restrict__ half_in,
                                                         const double* __restrict__ delta_in) {
  double ax = fabs(x);
  int lo = 0, hi = E;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (fabs(wledge[mid]) <= ax) lo = mid + 1;
    else hi = mid;
  }
  int e = lo - 1;
  e = e < 0 ? 0 : (e > E - 2 ? E - 2 : e);
  double left = fabs(wledge[e]), right = fabs(wledge[e + 1]);
  double h = half_in ? half_in[e] : 0.5 * (left + right);
  double dl;
  if (delta_in) dl = delta_in[e];
  else {
    double dd = right - left;
    dl = 0.5 * dd * dd;
  }
  if (dl == 0.0) dl = 1e-20;
  ContWeights c;
  c.e = e;
  c.c1 = (x - h) * (x - right) / dl;
  c.c2 = (left - x) * (x - right) * 2.0 / dl;
  c.c3 = (x - left) * (x - h) / dl;
  return c;
}

// Continuum absorption and scattering at depth d.
__device__ __forceinline__ void continuum_value(const ContWeights& c, int d, int E, const double* __restrict__ cabs,
                                                const double* __restrict__ cscat, int log10flag, double* va_o,
                                                double* vs_o) {
  const double* ca = cabs + ((long)d * (E - 1) + c.e) * 3;
  const double* cs = cscat + ((long)d * (E - 1) + c.e) * 3;
  double la = ca[0] * c.c1 + ca[1] * c.c2 + ca[2] * c.c3;
  double ls = cs[0] * c.c1 + cs[1] * c.c2 + cs[2] * c.c3;
  double va = la, vs = ls;
  if (log10flag) {
    if (la > 38.0) la = 38.0;
    if (ls > 38.0) ls = 38.0;
    va = pow(10.0, la);
    vs = pow(10.0, ls);
    if (va < 0.0) va = 0.0;
    if (va > 3.4e38) va = 3.4e38;
    if (vs < 0.0) vs = 0.0;
    if (vs > 3.4e38) vs = 3.4e38;
    if (!isfinite(va)) va = 3.4e38;
    if (!isfinite(vs)) vs = 3.4e38;
  }
  *va_o = va;
  *vs_o = vs;
}

__global__ void k_continuum(int D, long W, const double* __restrict__ wl, int E,
                            const double* __restrict__ wledge, const double* __restrict__ half_in,
                            const double* __restrict__ delta_in, const double* __restrict__ cabs,
                            const double* __restrict__ cscat, int log10flag,
                            double* __restrict__ out_abs, double* __restrict__ out_scat) {
  long w = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (w >= W) return;
  const ContWeights c = continuum_weights(wl[w], E, wledge, half_in, delta_in);
  const int d1 = min(D, (int)(blockIdx.y + 1) * kContinuumDepths);
  for (int d = blockIdx.y * kContinuumDepths; d < d1; ++d) {
    double va, vs;
    continuum_value(c, d, E, cabs, cscat, log10flag, &va, &vs);
    out_abs[(long)d * W + w] = va;
    out_scat[(long)d * W + w] = vs;
  }
}

// engine/opacity.py:2215-2228: EHVKT, STIM and BNU at one (depth, wavelength).
struct Planck {
  double bnu, stim, ehvkt;
};

__device__ __forceinline__ Planck planck_terms(double wl_nm, double temp) {
  double hkt = kHPlanck / (kKBoltz * temp);
  double freq = kCLightCm / (wl_nm * kNmToCm);
  Planck p;
  p.ehvkt = exp(-freq * hkt);
  p.stim = 1.0 - p.ehvkt;
  double freq15 = freq / 1.0e15;
  p.bnu = 1.47439e-02 * pow(freq15, 3.0) * p.ehvkt / p.stim;
  return p;
}

// engine/opacity.py:2215-2228 and physics/bfudge.py:72-168 (LTE).
__global__ void k_source(int D, long W, const double* __restrict__ wl,
                         const double* __restrict__ temp, const double* __restrict__ bfudge, double* bnu_o,
                         double* stim_o, double* ehvkt_o, double* slinec_o) {
  long w = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int d = blockIdx.y;
  if (w >= W) return;
  const Planck p = planck_terms(wl[w], temp[d]);
  long i = (long)d * W + w;
  if (bnu_o) bnu_o[i] = p.bnu;
  if (stim_o) stim_o[i] = p.stim;
  if (ehvkt_o) ehvkt_o[i] = p.ehvkt;
  // physics/bfudge.py:72-168: SLINEC = BNU STIM / (BFUDGE - EHVKT), BFUDGE = 1 in LTE
  const double bf = bfudge ? bfudge[d] : 1.0;
  if (slinec_o) slinec_o[i] = (p.bnu * p.stim) / (bf - p.ehvkt);
}

// One spectrum of a batch for k_cont_src_batch.
struct ContSrcArgs {
  long W;
  long blk0;  // first x-block of this spectrum in the flat grid
  int D, E, log10flag;
  const double *wl, *wledge, *half, *delta, *cabs, *cscat, *temp;
  double *out_abs, *out_scat, *kapmin, *slinec;
};

// Continuum, KAPMIN = abs + scat and the LTE line source of every spectrum of
// a batch in one launch: x-blocks are laid end to end over the spectra
// (a[i].blk0 ascending, a[0].blk0 = 0), grid rows cover kContinuumDepths
// depths each.  Each value is the expression of k_continuum, the KAPMIN sum
// and k_source, so batched spectra equal single ones bit for bit.
__global__ void k_cont_src_batch(const ContSrcArgs* __restrict__ a, int n) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid].blk0 <= (long)blockIdx.x) lo = mid;
    else hi = mid;
  }
  const ContSrcArgs& A = a[lo];
  const long W = A.W;
  const long w = ((long)blockIdx.x - A.blk0) * blockDim.x + threadIdx.x;
  const int d0 = blockIdx.y * kContinuumDepths;
  if (w >= W || d0 >= A.D) return;
  const double x = A.wl[w];
  const ContWeights c = continuum_weights(x, A.E, A.wledge, A.half, A.delta);
  const int d1 = min(A.D, d0 + kContinuumDepths);
  for (int d = d0; d < d1; ++d) {
    double va, vs;
    continuum_value(c, d, A.E, A.cabs, A.cscat, A.log10flag, &va, &vs);
    const long i = (long)d * W + w;
    A.out_abs[i] = va;
    A.out_scat[i] = vs;
    A.kapmin[i] = va + vs;
    const Planck p = planck_terms(x, A.temp[d]);
    A.slinec[i] = (p.bnu * p.stim) / (1.0 - p.ehvkt);
  }
}

// TRANSP (physics/line_opacity.py:56-184) and the ASYNTH wing setup
// (line_opacity.py:1160-1272 prep, 410-537 phases 1-2) for one (line, depth)
// pair per thread.  mode bit 0: compute transp/valid (else read transp_in /
// valid_in); bit 1: compute wing parameters and the reach class.
__device__ __forceinline__ void k_transp_wing_at(long bx_, int by_, long gx_, LinesDev ln, DepthsDev dp, long W, const double* __restrict__ kapmin,
                              double cutoff, double micro, const double* __restrict__ wl, DeviceTables tb,
                              int mode,
                              double* transp, uint8_t* valid, const double* transp_in,
                              const uint8_t* valid_in, WingParams* wp, uint8_t* cls) {
  long i = bx_ * (long)blockDim.x + threadIdx.x;
  long L = ln.n;
  int D = dp.n;
  if (i >= L * D) return;
  long l = i / D;
  int d = (int)(i - l * D);
  double tv = 0.0;
  bool ok = false;
  if (mode & 1) {
    double kapcen = 0.0;
    int ei = ln.elem[l], nel = ln.ion[l];
    if (ln.process[l] && ei >= 0 && ei < dp.nelem && nel > 0 && nel <= dp.nion) {
      long c = ln.center_t[l];
      long cc = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
      long pi = ((long)d * dp.nion + (nel - 1)) * dp.nelem + ei;
      double pv = dp.pop[pi];
      double dv = dp.dop[pi];
      double lw = ln.wl[l];
      if (micro > 0.0) dv = sqrt(dv * dv + micro * micro);
      double r = dp.rho[d];
      if (pv > 0.0 && dv > 0.0 && r > 0.0) {
        double xnfdop = pv / (r * dv);
        double dw = dv * lw;
        double boltz = fastex(ln.elo[l] * dp.hckt[d], tb.extab, tb.extabf);
        int lt = ln.ltype[l];
        double k0pre = lt == 1 ? ln.gvdw[l] * ln.gf[l] * (pv / r) : ln.cgf[l] * xnfdop;
        double km = kapmin[(long)d * W + cc] * cutoff;
        double post = k0pre * boltz;
        if (!(k0pre < km) && !(post < km) && post > 0.0) {
          double dopple = lw > 0 ? dw / lw : 1e-6;
          double adamp = 0.0;
          if (dw > 0 && dopple > 0) {
            double gt = ln.grad[l] + ln.gstark[l] * dp.xne[d] + ln.gvdw[l] * dp.txnxn[d];
            adamp = gt / dopple;
          }
          if (adamp >= 0.0) {
            if (lt == 1) kapcen = post;
            else if (adamp < 0.2) kapcen = post * (1.0 - 1.128 * adamp);
            else kapcen = post * voigt(0.0, adamp, tb.h0, tb.h1, tb.h2);
            ok = true;
          }
        }
      }
    }
    tv = ok ? kapcen : 0.0;
    transp[i] = tv;
    valid[i] = ok ? 1 : 0;
  } else {
    tv = transp_in[i];
    ok = valid_in[i] != 0;
  }
  if (!(mode & 2)) return;
  // resolu from the grid's first ratio (line_opacity.py:233-238)
  double resolu = 300000.0;
  if (wl && W > 1) {
    double ratio = wl[1] / wl[0];
    if (ratio > 1.0) resolu = 1.0 / (ratio - 1.0);
  }
  uint8_t k_out = 0xFF;
  long c = ln.center_w[l];
  int ei = ln.elem[l], nel = ln.ion[l];
  if (ok && tv > 0.0 && ln.ltype[l] == 0 && c >= -kMaxProfileSteps && c <= W - 1 + kMaxProfileSteps &&
      ei >= 0 && ei < dp.nelem && nel > 0 && nel <= dp.nion) {
    long pi = ((long)d * dp.nion + (nel - 1)) * dp.nelem + ei;
    double pv = dp.pop[pi];
    double dv = dp.dop[pi];
    double lw = ln.wl[l];
    if (pv > 0.0 && dv > 0.0) {
      double dw = dv * lw;
      double dopple_p = lw > 0 ? dw / lw : 1e-6;
      double adamp = 0.0;
      if (dopple_p > 0 && lw > 0) {
        double gt = ln.grad[l] + ln.gstark[l] * dp.xne[d] + ln.gvdw[l] * dp.txnxn[d];
        adamp = gt / dopple_p;
      }
      if (adamp < 1e-12) adamp = 1e-12;
      long cw = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
      double kref = kapmin[(long)d * W + cw] * cutoff;
      double vc = adamp < 0.2 ? 1.0 - 1.128 * adamp : voigt(0.0, adamp, tb.h0, tb.h1, tb.h2);
      double kappa0 = vc > 0 ? tv / vc : tv;
      if (dw > 0.0) {
        double dopple = lw > 0.0 ? dw / lw : 1e-10;
        long n10dop = (long)(10.0 * dopple * resolu);
        // MAXPROF bounds every wing; the Voigt zone past it is never added
        // unless the reference's near-zone cut falls beyond it, which needs a
        // grid finer than R ~ 10^5 / (v/c) (the Python layer rejects such grids)
        if (n10dop > kMaxProfileSteps) n10dop = kMaxProfileSteps;
        double dvoigt = dopple > 0 ? 1.0 / (dopple * resolu) : 1.0;
        double tabstep = 200.0 * dvoigt;
        long cut = -1;
        double p_n10 = 0.0;
        double tabi = 0.5;
        for (long k = 1; k <= n10dop; ++k) {
          double p;
          double x = (double)k * dvoigt;
          if (adamp < 0.2) {
            tabi += tabstep;
            long idx = (long)tabi;
            if (x > 10.0) p = kappa0 * (0.5642 * adamp / (x * x));
            else {
              idx = idx < 0 ? 0 : (idx > kVoigtN - 1 ? kVoigtN - 1 : idx);
              p = kappa0 * (tb.h0[idx] + adamp * tb.h1[idx]);
            }
          } else {
            p = kappa0 * voigt(x, adamp, tb.h0, tb.h1, tb.h2);
          }
          if (k == n10dop) p_n10 = p;
          if (p < kref) {
            cut = k;
            break;
          }
        }
        long maxstep;
        int use_far;
        double x_far;
        if (cut != -1) {
          maxstep = cut;
          use_far = 0;
          x_far = 0.0;
        } else {
          use_far = 1;
          if (n10dop > 0 && p_n10 > 0.0 && kref > 0.0) {
            x_far = p_n10 * ((double)n10dop * (double)n10dop);
            double ms = sqrt(x_far / kref) + 1.0;
            maxstep = ms < 2.0e6 ? (long)ms : kMaxProfileSteps;
          } else {
            x_far = 0.0;
            maxstep = 0;
          }
          if (maxstep > kMaxProfileSteps) maxstep = kMaxProfileSteps;
        }
        if (maxstep > 0) {
          WingParams p;
          p.kappa0 = kappa0;
          p.adamp = adamp;
          p.dvoigt = dvoigt;
          p.x_far = x_far;
          p.c = (int32_t)c;
          p.maxstep = (int32_t)maxstep;
          p.n10dop = (int32_t)n10dop;
          p.use_far = use_far;
          wp[i] = p;
          k_out = (uint8_t)reach_class((int)maxstep);
        }
      }
    }
  }
  cls[i] = k_out;
}

__global__ void k_transp_wing(LinesDev ln, DepthsDev dp, long W, const double* __restrict__ kapmin,
                              double cutoff, double micro, const double* __restrict__ wl, DeviceTables tb,
                              int mode,
                              double* transp, uint8_t* valid, const double* transp_in,
                              const uint8_t* valid_in, WingParams* wp, uint8_t* cls) {
  k_transp_wing_at(blockIdx.x, blockIdx.y, gridDim.x, ln, dp, W, kapmin, cutoff, micro, wl, tb, mode, transp, valid, transp_in, valid_in, wp, cls);
}

constexpr int kListDepths = 8;  // depths per workgroup row of the class count and scatter

// Per (block of 256 centre-sorted lines, depth): pair counts per reach class.
// A lane's reach classes for the depths [d0, d1) of its line l (0xFF past the
// row, or for l < 0), read before the depth loop of the class count and
// scatter: read once per depth, a line's bytes were fetched again after L2
// had evicted them.
__device__ __forceinline__ void load_row_classes(long l, int D, int d0, int d1, const uint8_t* __restrict__ cls,
                                                 uint8_t (&kc)[kListDepths]) {
#pragma unroll
  for (int t = 0; t < kListDepths; ++t) kc[t] = (l >= 0 && d0 + t < d1) ? cls[l * D + d0 + t] : (uint8_t)0xFF;
}

__device__ __forceinline__ void k_class_count_at(long bx_, int by_, long gx_, long L, int D, const int64_t* __restrict__ order,
                                                        const uint8_t* __restrict__ cls,
                                                        int32_t* __restrict__ counts) {
  __shared__ int sh[kBlock / 64][kClasses];
  long s = bx_ * (long)kBlock + threadIdx.x;
  long l = s < L ? order[s] : -1;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // depths are independent: each workgroup row takes kListDepths of them
  const int d0 = by_ * kListDepths;
  const int d1 = min(D, d0 + kListDepths);
  uint8_t kc[kListDepths];
  load_row_classes(l, D, d0, d1, cls, kc);
#pragma unroll
  for (int t = 0; t < kListDepths; ++t) {
    const int d = d0 + t;
    if (d >= d1) break;  // uniform over the workgroup
    int k = kc[t];
    for (int kk = 0; kk < kClasses; ++kk) {
      unsigned long long m = __ballot(k == kk);
      if (lane == 0) sh[wave][kk] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) {
      int t = 0;
      for (int q = 0; q < kBlock / 64; ++q) t += sh[q][threadIdx.x];
      counts[((long)bx_ * D + d) * kClasses + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_class_count(long L, int D, const int64_t* __restrict__ order,
                                                        const uint8_t* __restrict__ cls,
                                                        int32_t* __restrict__ counts) {
  k_class_count_at(blockIdx.x, blockIdx.y, gridDim.x, L, D, order, cls, counts);
}

// Exclusive scan over blocks of counts[blk][d][k] for one (d, k) per workgroup;
// totals[d][k] gets the sum.
__device__ __forceinline__ void k_class_scan_at(long bx_, int by_, long gx_, long nblk, int D, int32_t* __restrict__ counts,
                                                       int64_t* __restrict__ totals) {
  __shared__ int64_t sh[kBlock];
  int dk = bx_;
  int d = dk / kClasses, kk = dk % kClasses;
  int64_t carry = 0;
  for (long base = 0; base < nblk; base += kBlock) {
    long b = base + threadIdx.x;
    int64_t v = b < nblk ? counts[(b * D + d) * kClasses + kk] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    int64_t incl = sh[threadIdx.x];
    if (b < nblk) counts[(b * D + d) * kClasses + kk] = (int32_t)(carry + incl - v);
    carry += sh[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[dk] = carry;
}

__global__ __launch_bounds__(kBlock) void k_class_scan(long nblk, int D, int32_t* __restrict__ counts,
                                                       int64_t* __restrict__ totals) {
  k_class_scan_at(blockIdx.x, blockIdx.y, gridDim.x, nblk, D, counts, totals);
}

// regions[d][k] = start of the (d, k) list; regions[d][kClasses] = end.
__device__ __forceinline__ void k_regions_at(long bx_, int by_, long gx_, int D, long L, const int64_t* __restrict__ totals, int64_t* __restrict__ regions) {
  int d = bx_ * blockDim.x + threadIdx.x;
  if (d >= D) return;
  int64_t pos = (int64_t)d * L;
  for (int k = 0; k < kClasses; ++k) {
    regions[d * (kClasses + 1) + k] = pos;
    pos += totals[d * kClasses + k];
  }
  regions[d * (kClasses + 1) + kClasses] = pos;
}

__global__ void k_regions(int D, long L, const int64_t* __restrict__ totals, int64_t* __restrict__ regions) {
  k_regions_at(blockIdx.x, blockIdx.y, gridDim.x, D, L, totals, regions);
}

// Stable scatter of wing pairs into their (depth, class) lists in centre order.
__device__ __forceinline__ void k_class_scatter_at(long bx_, int by_, long gx_, 
    long L, int D, const int64_t* __restrict__ order, const uint8_t* __restrict__ cls,
    const int32_t* __restrict__ offsets, const int64_t* __restrict__ regions,
    const WingParams* __restrict__ wp, WingHead* __restrict__ list_cm, WingParams* __restrict__ list_p,
    int32_t* __restrict__ list_l) {
  __shared__ int sh[kBlock / 64][kClasses];
  long s = bx_ * (long)kBlock + threadIdx.x;
  long l = s < L ? order[s] : -1;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // depths are independent: each workgroup row takes kListDepths of them
  const int d0 = by_ * kListDepths;
  const int d1 = min(D, d0 + kListDepths);
  uint8_t kc[kListDepths];
  load_row_classes(l, D, d0, d1, cls, kc);
#pragma unroll
  for (int t = 0; t < kListDepths; ++t) {
    const int d = d0 + t;
    if (d >= d1) break;  // uniform over the workgroup
    int k = kc[t];
    int rank = 0;
    for (int kk = 0; kk < kClasses; ++kk) {
      unsigned long long m = __ballot(k == kk);
      if (lane == 0) sh[wave][kk] = __popcll(m);
      if (k == kk) rank = __popcll(m & ltmask);
    }
    __syncthreads();
    if (k != 0xFF) {
      int before = 0;
      for (int q = 0; q < wave; ++q) before += sh[q][k];
      int64_t pos = regions[d * (kClasses + 1) + k] + offsets[((long)bx_ * D + d) * kClasses + k] +
                    before + rank;
      const WingParams p = wp[l * D + d];
      // the zone ends at MAXSTEP when MAXPROF caps a far wing below N10DOP
      const int r = p.use_far ? (p.n10dop < p.maxstep ? p.n10dop : p.maxstep) : p.maxstep;
      const bool chunk = !(p.x_far > 0.0 && p.x_far < kChunkMinX);  // X >= 0
      WingHead h;
      h.lo = p.use_far ? p.c - p.maxstep : INT32_MAX;
      h.hi = p.use_far ? p.c + p.maxstep : INT32_MIN;
      h.nlo = p.c - r;
      h.nhi = p.c + r;
      h.xs = p.use_far ? (chunk ? p.x_far : -p.x_far) : 0.0;
      h.cd = (double)p.c;
      list_cm[pos] = h;
      list_p[pos] = p;
      if (list_l) list_l[pos] = (int32_t)l;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_class_scatter(
    long L, int D, const int64_t* __restrict__ order, const uint8_t* __restrict__ cls,
    const int32_t* __restrict__ offsets, const int64_t* __restrict__ regions,
    const WingParams* __restrict__ wp, WingHead* __restrict__ list_cm, WingParams* __restrict__ list_p,
    int32_t* __restrict__ list_l) {
  k_class_scatter_at(blockIdx.x, blockIdx.y, gridDim.x, L, D, order, cls, offsets, regions, wp, list_cm, list_p, list_l);
}

// Per depth: the largest Voigt-zone reach over the list entries, i.e.
// N10DOP for pairs with a far wing and MAXSTEP (<= N10DOP) for the others.
// near_reach[] must be zero on entry; the result is a max, so it does not
// depend on the order of the atomics.
__device__ __forceinline__ void k_near_reach_at(long bx_, int by_, long gx_, int D, const int64_t* __restrict__ regions,
                                                       const WingHead* __restrict__ list_cm,
                                                       int32_t* __restrict__ near_reach) {
  const int d = by_;
  const long beg = regions[d * (kClasses + 1)], end = regions[d * (kClasses + 1) + kClasses];
  int m = 0;
  for (long j = beg + bx_ * (long)kBlock + threadIdx.x; j < end; j += (long)gx_ * kBlock) {
    const WingHead h = list_cm[j];
    const int r = h.nhi - head_c(h);
    m = r > m ? r : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_down(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(near_reach + d, m);
}

__global__ __launch_bounds__(kBlock) void k_near_reach(int D, const int64_t* __restrict__ regions,
                                                       const WingHead* __restrict__ list_cm,
                                                       int32_t* __restrict__ near_reach) {
  k_near_reach_at(blockIdx.x, blockIdx.y, gridDim.x, D, regions, list_cm, near_reach);
}

__global__ void k_fill_u64(long n, unsigned long long* p) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_zero_i32(int n, int32_t* p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__device__ __forceinline__ long uniform_long(long v) {
  int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (long)(((unsigned long)(unsigned)hi << 32) | (unsigned)lo);
}

// For each class kk < kClasses, lanes 2kk and 2kk+1 find the list window of
// entries whose centre lies in [lo_px - r_kk, hi_px + r_kk], r_kk = min(reach
// of class kk, cap); the result is read back with window_bounds().
__device__ __forceinline__ long class_window_search(int lane, int d, long lo_px, long hi_px, long cap,
                                                    const int64_t* __restrict__ regions,
                                                    const WingHead* __restrict__ list_cm) {
  long bound = 0;
  if (lane < 2 * kClasses) {
    int kk = lane >> 1;
    bool upper = lane & 1;
    long r = 16L << (kClassShift * kk);
    r = r < cap ? r : cap;
    long lo = regions[d * (kClasses + 1) + kk], hi = regions[d * (kClasses + 1) + kk + 1];
    long key = upper ? hi_px + r : lo_px - r;
    while (lo < hi) {  // first entry with c > key (upper) or c >= key (lower)
      long mid = (lo + hi) >> 1;
      long c = head_c(list_cm[mid]);
      if (upper ? c <= key : c < key) lo = mid + 1;
      else hi = mid;
    }
    bound = lo;
  }
  return bound;
}

// Voigt zone of every wing: pixels 1 <= |k| <= N10DOP of pairs with a far
// wing and 1 <= |k| <= MAXSTEP of the others, valued as
// physics/line_opacity.py:410-560 (near_wing).  One wave per 64 pixels of one
// depth, one pixel per lane.  The class windows are walked 64 entries per step
// (one coalesced load of their zone bounds); the wing parameters of the
// entries whose zone meets the window are staged in LDS, each lane builds the
// mask of those entries that reach its own pixel (two broadcasts and a range
// test per entry), then adds their values in list order.  Each pixel's sum
// runs in class-then-centre order, entry by entry.  Writes the pre-STIM sum
// to out [D][W] (0 where no wing reaches), which k_gather continues from.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool kStats>
__device__ __forceinline__ void k_near_at(long bx_, int by_, long gx_, long W, int D, const WingHead* __restrict__ list_cm,
                                              const WingParams* __restrict__ list_p,
                                              const int64_t* __restrict__ regions,
                                              const int32_t* __restrict__ near_reach, DeviceTables tb,
                                              double* __restrict__ out, unsigned long long* __restrict__ evals) {
  __shared__ WingParams s_p[4][64];  // one step's meeting entries, per wave
  const int d = by_;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long v0 = ((long)bx_ * 4 + wv) * 64;
  if (v0 >= W) return;  // whole wave; the waves of a block share no LDS
  const long w = v0 + lane;
  // pixel indices and centres fit 32 bits (n_wl <= 2^30): 32-bit entry tests
  const int v0i = (int)v0, wi = v0i + lane;
  const bool on_grid = w < W;
  const long bound = class_window_search(lane, d, v0, v0 + 63, (long)near_reach[d], regions, list_cm);
  WingParams* sp = s_p[wv];
  double acc = 0.0;
  unsigned long long nev = 0;
  for (int kk = 0; kk < kClasses; ++kk) {
    const long lo = uniform_long(__shfl(bound, 2 * kk));
    const long hi = uniform_long(__shfl(bound, 2 * kk + 1));
    for (long j0 = lo; j0 < hi; j0 += 64) {
      const long jl = j0 + lane;
      int2 z = make_int2(INT32_MAX, INT32_MIN);
      if (jl < hi) {
        const WingHead hv = list_cm[jl];
        z = make_int2(hv.nlo, hv.nhi);
      }
      const bool mine = !(v0i + 63 < z.x || v0i > z.y);  // my entry's zone meets the window
      unsigned long long meet = __ballot(mine);
      if (meet == 0ull) continue;
      if (mine) sp[lane] = list_p[jl];
      // the entries that reach this lane's pixel, 1 <= |k| <= the zone's half-width
      unsigned long long reach = 0ull;
      while (meet) {
        const int src = __builtin_ctzll(meet);
        meet &= meet - 1;
        const int nlo = __builtin_amdgcn_readlane(z.x, src);
        const int nhi = __builtin_amdgcn_readlane(z.y, src);
        const int lim = (nhi - nlo) >> 1;  // the zone's half-width
        const int ka = abs(wi - (nlo + lim));
        reach |= ((unsigned)(ka - 1) < (unsigned)lim && on_grid) ? (1ull << src) : 0ull;
      }
      wave_lds_sync();  // the staged parameters before their reads
      while (reach) {
        const int src = __builtin_ctzll(reach);
        reach &= reach - 1;
        const WingParams pr = sp[src];
        acc += near_wing(pr, abs(wi - pr.c), tb.h0, tb.h1, tb.h2);
        if (kStats) ++nev;
      }
      wave_lds_sync();  // the reads before the next step's writes
    }
  }
  if (w < W) out[(long)d * W + w] = acc;
  if (kStats) {
    for (int off = 32; off > 0; off >>= 1) nev += __shfl_down(nev, off);
    if (lane == 0) {
      atomicAdd(evals, nev);
      atomicAdd(evals + 1, nev);
    }
  }
}

template <bool kStats>
__global__ __launch_bounds__(256) void k_near(long W, int D, const WingHead* __restrict__ list_cm,
                                              const WingParams* __restrict__ list_p,
                                              const int64_t* __restrict__ regions,
                                              const int32_t* __restrict__ near_reach, DeviceTables tb,
                                              double* __restrict__ out, unsigned long long* __restrict__ evals) {
  k_near_at<kStats>(blockIdx.x, blockIdx.y, gridDim.x, W, D, list_cm, list_p, regions, near_reach, tb, out, evals);
}

#ifndef SYNTHE_GATHER_PIXELS
#define SYNTHE_GATHER_PIXELS 8
#endif
constexpr int kWP = SYNTHE_GATHER_PIXELS;  // consecutive pixels per lane
constexpr int kWT = 64 * kWP;       // pixels per wave (one wave per workgroup)

// Inputs and output of the wing gather for one spectrum (device pointers).
struct GatherArgs {
  long W;
  int D;
  const int64_t* corder;
  const int64_t* cstart;
  const double* transp;
  const uint8_t* valid;
  const WingHead* list_cm;
  const WingParams* list_p;
  const int64_t* regions;
  const double* grid;
  const double* temp;
  double* asynth;
};

// Profile precision of the far-wing term X / k^2 (accumulation is FP64 in
// both): kPrecFP64 within 32 ulp of the IEEE quotient (2^-1074 absolute where
// the quotient is subnormal), kPrecFP32 at single precision.
constexpr int kPrecFP64 = 0;
constexpr int kPrecFP32 = 1;

template <int kPrec>
__device__ __forceinline__ double far_quotient(double x, double b) {
  if constexpr (kPrec == kPrecFP32) return div_single(x, b);
  else return div_fast(x, b);
}

// waves per SIMD the gather's register budget is sized for: the tile and row
// expansions (24 moments, their wave sums and Horner evaluation) need about
// 130 VGPRs, so four waves (128 VGPRs each) spill little and five spill
// heavily
#ifndef SYNTHE_GATHER_WAVES
#define SYNTHE_GATHER_WAVES 4
#endif
#ifndef SYNTHE_GATHER_WAVES32
#define SYNTHE_GATHER_WAVES32 4
#endif
template <int kPrec>
constexpr int gather_waves() {
  return kPrec == 1 ? SYNTHE_GATHER_WAVES32 : SYNTHE_GATHER_WAVES;
}

// Local expansion of the far wings of entries far from a tile.  For an entry
// whose far zone covers the whole tile, with the tile centre at distance
// K = w_m - c from the line centre and pixel w = w_m + u, |u| <= h (half the
// tile),
//   X / (K + u)^2 = (X / K^2) sum_n (n + 1) (-u/K)^n.
// With rho = h / |K|, the terms from n = T on sum to at most
// (T + 1) rho^T (1 + rho)^2 / (1 - rho)^2 of the exact value.  Entries with
// |K| >= kFarExpandRatio h take the expansion: rho <= 1/4.5, so T = 24 terms
// leave 1.2e-14 relative per entry (FP64 form).  The FP32 form, whose bar is
// 1e-4 against the FP64 form, expands from |K| >= 2.5 h with T = 20 terms
// (rho <= 0.4: 1.3e-6 relative per entry), which leaves 18% fewer tiles to
// the exact pass.  Every entry's exact contribution X / k^2 is positive, so a pixel's
// sum of expansions keeps the same relative bound.
#ifndef SYNTHE_FAR_TERMS
#define SYNTHE_FAR_TERMS 24
#endif
constexpr int kFarTerms = SYNTHE_FAR_TERMS;
constexpr int kFarTerms32 = 20;
constexpr double kFarExpandRatio = 4.5;
constexpr double kFarExpandRatio32 = 2.5;

__device__ __forceinline__ double wave_sum_butterfly(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;  // every lane holds the same sum (fixed pairing order)
}

// The tile lies wholly in the entry's far zone N10DOP < |k| <= MAXSTEP.
__device__ __forceinline__ bool far_interior(const WingHead& h, int t0i, int t1i) {
  return (t0i > h.nhi && t1i <= h.hi) || (t1i < h.nlo && t0i >= h.lo);
}

// Continues each pixel's sum from the Voigt-zone sums of k_near (in asynth),
// adds the line centres (catalog order, line_opacity.py:1091-1102), then the
// far wings X / k^2 (N10DOP < |k| <= MAXSTEP, line_opacity.py:562-617) of
// every list entry that reaches the tile, and multiplies by STIM
// (line_opacity.py:1338).  One wave per (tile, depth); a tile is kWP rows of
// 64 pixels and lane l holds pixel t0 + 64 r + l of row r, so a row is one
// coalesced access and one pixel per lane.  The far wings are summed at three
// granularities:
//  1. tile: the class windows are walked 64 entries per step with one
//     coalesced load of their heads; entries whose far zone covers the tile
//     at |K| >= 4.5 h (K: tile centre - line centre, h: half the tile) add
//     their expansion moments (one entry per lane), reduced over the wave and
//     evaluated per pixel (Horner in u/h).  The other entries that reach the
//     tile are queued in LDS (their list positions, in the order met);
//  2. row: for each row, the queued entries whose far zone covers the row at
//     |K_r| >= 4.5 h_r (h_r = 32) add the same expansion about the row
//     centre;
//  3. exact: each queued entry, broadcast to the wave, adds X / k^2 on the
//     rows it reaches and could not expand, one quotient per pixel; on rows
//     holding a wing's N10DOP or MAXSTEP edge, pixels outside the far zone
//     add +0.
// Steps 2 and 3 run when the queue is full (after the tile moments met so far
// are added to the pixels) and at the end.  Each pixel's sum has a fixed
// order that depends only on the lists, no atomics, so results are
// run-to-run identical.
// The queue holds one 16-byte record per entry (gather_queue.h): X as the
// head stores it, and the centre relative to the tile with the two
// half-widths packed in 61 bits.  640 records are 10 KiB of LDS per wave, so
// the 16 waves a CU holds at four per SIMD fill its 160 KiB (a full queue is
// flushed: 384, 512, 576 and 640 records measured 21.9, 20.7, 20.5 and 20.0 ms
// on the full window).
#ifndef SYNTHE_GATHER_QUEUE
#define SYNTHE_GATHER_QUEUE 640
#endif
constexpr int kGatherQueue = SYNTHE_GATHER_QUEUE;
static_assert(kGatherQueue >= 64 && kGatherQueue % 64 == 0, "queue holds whole steps of 64 entries");
static_assert(kMaxProfileSteps == gq::kMaxZone, "queue record fields hold MAXPROF");
static_assert(64 * kWP <= gq::kMaxTile, "queue record centre field holds a tile");

struct QueueRec {
  double xs;
  uint64_t pk;
};

__device__ __forceinline__ QueueRec queue_pack(const WingHead& h, int t0i) {
  QueueRec q;
  q.xs = h.xs;
  q.pk = gq::pack(head_c(h) - t0i, (h.nhi - h.nlo) >> 1, (h.hi - h.lo) >> 1);
  return q;
}

// The head of a queued entry in the tile's own coordinates (pixel t0 is 0).
__device__ __forceinline__ WingHead queue_unpack(const QueueRec& q) {
  const gq::Fields f = gq::unpack(q.pk);
  WingHead h;
  h.lo = f.dc - f.m;
  h.hi = f.dc + f.m;
  h.nlo = f.dc - f.r;
  h.nhi = f.dc + f.r;
  h.xs = q.xs;
  h.cd = (double)f.dc;
  return h;
}

// Tile expansion: adds (n + 1) * sum_lanes(mom[n]) t^n to every pixel of the
// tile's row segment, with t = (pixel - tile centre) / h (exact: h is a power
// of two).
template <int kT, int kWP>
__device__ __forceinline__ void add_tile_expansion(const double* mom, double* __restrict__ row, long t0, long W,
                                                   int lane) {
  constexpr double kCentre = 0.5 * (64 * kWP - 1), kInvH = 1.0 / (32.0 * kWP);
  double coef[kT];
#pragma unroll
  for (int n = 0; n < kT; ++n) coef[n] = (double)(n + 1) * wave_sum_butterfly(mom[n]);
#pragma unroll
  for (int r = 0; r < kWP; ++r) {
    const double t = ((double)(64 * r + lane) - kCentre) * kInvH;
    double s = coef[kT - 1];
#pragma unroll
    for (int n = kT - 2; n >= 0; --n) s = __builtin_fma(s, t, coef[n]);
    const long w = t0 + 64 * r + lane;
    if (w < W) row[w] += s;
  }
}

// Row expansion: the same about the centre of a row (h = 32), the value at
// the lane's pixel of that row.
template <int kT>
__device__ __forceinline__ double row_expansion(const double* mom, int lane) {
  const double t = ((double)lane - 31.5) * (1.0 / 32.0);
  double s = (double)kT * wave_sum_butterfly(mom[kT - 1]);
#pragma unroll
  for (int n = kT - 2; n >= 0; --n) s = __builtin_fma(s, t, (double)(n + 1) * wave_sum_butterfly(mom[n]));
  return s;
}

// The value of v in lane src, for every lane.
__device__ __forceinline__ double bcast_double(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One expansion step: with y = 1/K, adds X/K^2 (-h/K)^n to mom[n] (x = 0
// adds nothing; K must be nonzero).
template <int kT>
__device__ __forceinline__ void expansion_moments(double x, double kq, double half, double* mom) {
  double y = __builtin_amdgcn_rcp(kq);  // 1/K to an ulp: two Newton steps
  y = __builtin_fma(y, __builtin_fma(-kq, y, 1.0), y);
  y = __builtin_fma(y, __builtin_fma(-kq, y, 1.0), y);
  double q = x * y * y;
  const double rr = -half * y;
  mom[0] += q;
#pragma unroll
  for (int n = 1; n < kT; ++n) {
    q *= rr;
    mom[n] += q;
  }
}

// The row r0..r1 lies in the entry's far zone and its centre is at least
// kmin_r from the line centre (the row-expansion test).
__device__ __forceinline__ bool row_expands(const WingHead& h, int r0, int r1, double wr, double kmin_r) {
  return far_interior(h, r0, r1) && fabs(wr - h.cd) >= kmin_r && h.xs >= 0.0;
}

template <bool kStats, int kPrec, int kWP>
__device__ __forceinline__ void gather_tile(const GatherArgs& g, int d, long t0, int lane,
                                            unsigned long long* __restrict__ evals) {
  constexpr int kWT = 64 * kWP;
  constexpr int kT = kPrec == kPrecFP32 ? kFarTerms32 : kFarTerms;
  constexpr double kRatio = kPrec == kPrecFP32 ? kFarExpandRatio32 : kFarExpandRatio;
  const long W = g.W;
  const int D = g.D;
  const int64_t* __restrict__ corder = g.corder;
  const int64_t* __restrict__ cstart = g.cstart;
  const double* __restrict__ transp = g.transp;
  const uint8_t* __restrict__ valid = g.valid;
  const WingHead* __restrict__ list_cm = g.list_cm;
  const int64_t* __restrict__ regions = g.regions;
  const double* __restrict__ grid = g.grid;
  const double* __restrict__ temp = g.temp;
  double* __restrict__ asynth = g.asynth;
  // Each pixel's running sum lives in its own element of the tile's asynth
  // row (L2-resident), read and written back only by its lane, so no
  // register holds the tile's sums across the steps below; pixels off the
  // grid have no element and their sums are not formed.
  double* row = asynth + (long)d * W;
  unsigned long long nev = 0;
#pragma unroll
  for (int r = 0; r < kWP; ++r) {
    const long w = t0 + 64 * r + lane;
    if (w < W) {
      double a = row[w];
      for (long j = cstart[w]; j < cstart[w + 1]; ++j) {
        long l = corder[j];
        if (valid[l * D + d]) {
          a += transp[l * D + d];
          if (kStats) ++nev;
        }
      }
      row[w] = a;
    }
  }
  // Pixel indices, centres and offsets fit in 32 bits (W <= 2^30, centres
  // within 10^6 of the grid), so the per-entry tests are 32-bit ones.
  const int t0i = (int)t0, t1i = (int)t0 + kWT - 1;
  const long bound = class_window_search(lane, d, t0, t0 + kWT - 1, 1L << 40, regions, list_cm);
  const long npix = W - t0 < kWT ? W - t0 : kWT;  // tile pixels on the grid
  constexpr double kHalf = 0.5 * kWT;             // powers of two: u / h is exact
  constexpr double kHalfRow = 32.0;
  const double wm = (double)t0 + 0.5 * (kWT - 1);
  const double kmin = kRatio * kHalf;
  const double kmin_r = kRatio * kHalfRow;
  WingHead none;
  none.lo = INT32_MAX;
  none.hi = INT32_MIN;
  none.nlo = INT32_MAX;
  none.nhi = INT32_MIN;
  none.xs = 0.0;
  none.cd = 0.0;

  __shared__ QueueRec s_q[kGatherQueue];
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int qn = 0;             // queued entries (wave-uniform)
  bool tile_any = false;  // a tile expansion since the moments were last added
  double mom[kT];
#pragma unroll
  for (int n = 0; n < kT; ++n) mom[n] = 0.0;

  // Steps 2 and 3 on the qn queued entries, then an empty queue.  The
  // queued heads are in the tile's own coordinates (queue_unpack), and so are
  // the rows here (row r: pixels 64 r .. 64 r + 63).  `ln` is the lane index
  // made opaque, so that the per-row pixel values below are formed at each
  // flush instead of being hoisted out of the scan loop and held in
  // registers throughout it.
  auto flush = [&]() {
    int ln = lane;
    __asm__ volatile("" : "+v"(ln));
    __syncthreads();  // the queue's writes before its reads (one wave per workgroup)
    // ---- step 2: row expansions
#pragma unroll 1
    for (int r = 0; r < kWP; ++r) {
      const int r0 = 64 * r, r1 = r0 + 63;
      const double wr = (double)r0 + 31.5;
      const long wa = t0 + r0;  // the row's first pixel on the grid
      double rm[kT];
#pragma unroll
      for (int n = 0; n < kT; ++n) rm[n] = 0.0;
      bool any = false;
#pragma unroll 1
      for (int i0 = 0; i0 < qn; i0 += 64) {
        const int i = i0 + ln;
        const WingHead h = i < qn ? queue_unpack(s_q[i]) : none;
        const bool ex = row_expands(h, r0, r1, wr, kmin_r);
        const unsigned long long em = __ballot(ex);
        if (em == 0ull) continue;
        any = true;
        expansion_moments<kT>(ex ? h.xs : 0.0, ex ? wr - h.cd : kmin_r, kHalfRow, rm);
        if (kStats && ex) {
          const long nr = W - wa < 64 ? W - wa : 64;
          nev += nr > 0 ? nr : 0;
        }
      }
      if (any) {
        const double s = row_expansion<kT>(rm, ln);
        if (wa + ln < W) row[wa + ln] += s;
      }
    }
    // ---- step 3: exact rows.  Each lane first classifies its entry's rows
    // (bit r of ex_int: the row lies in one far side and does not expand, so
    // every pixel adds X / k^2; of ex_edge: the row holds a zone edge, or X
    // lies below 2^-500, so pixels are masked); then, row by row, the
    // entries with that row's bit are broadcast in queue order (interior
    // rows first), two per iteration: their quotients are independent and
    // are added one after the other (a missing second entry adds +0, which
    // leaves the non-negative sum unchanged).  The row sums stay in
    // registers for the whole step.
    double acc[kWP];
#pragma unroll
    for (int r = 0; r < kWP; ++r) {
      const long w = t0 + 64 * r + ln;
      acc[r] = w < W ? row[w] : 0.0;
    }
#pragma unroll 1
    for (int i0 = 0; i0 < qn; i0 += 64) {
      const int i = i0 + ln;
      const WingHead hv = i < qn ? queue_unpack(s_q[i]) : none;
      const bool chunk_l = __double_as_longlong(hv.xs) >= 0;  // X >= 2^-500 (or 0)
      unsigned ex_int = 0u, ex_edge = 0u;
#pragma unroll
      for (int r = 0; r < kWP; ++r) {
        const int r0 = 64 * r, r1 = r0 + 63;
        // the row meets the far zone [lo, nlo - 1] u [nhi + 1, hi]
        const bool meets = (r0 < hv.nlo && r1 >= hv.lo) || (r1 > hv.nhi && r0 <= hv.hi);
        const bool inside = far_interior(hv, r0, r1);
        const bool expands = inside && fabs((double)r0 + 31.5 - hv.cd) >= kmin_r && hv.xs >= 0.0;
        ex_int |= (meets && inside && chunk_l && !expands) ? (1u << r) : 0u;
        ex_edge |= (meets && !(inside && chunk_l)) ? (1u << r) : 0u;
      }
      if (i >= qn) ex_int = ex_edge = 0u;
      // the values other lanes read with readlane: formed here in every lane
      // (the opaque asm keeps the compiler from sinking them into the
      // branches below, where the lane read may be inactive)
      double x_l = chunk_l ? hv.xs : -hv.xs;
      double cd_l = hv.cd;
      double n10_l = (double)((hv.nhi - hv.nlo) >> 1), m_l = (double)((hv.hi - hv.lo) >> 1);
      __asm__ volatile("" : "+v"(x_l), "+v"(cd_l), "+v"(n10_l), "+v"(m_l));
#pragma unroll
      for (int r = 0; r < kWP; ++r) {
        const double px = (double)(64 * r + ln);  // the lane's pixel in row r
        const bool on = t0 + 64 * r + ln < W;
        unsigned long long todo = __ballot((ex_int >> r) & 1u);
        while (todo) {
          const int s1 = __builtin_ctzll(todo);
          todo &= todo - 1;
          const bool two = todo != 0ull;
          const int s2 = two ? __builtin_ctzll(todo) : s1;
          if (two) todo &= todo - 1;
          const double k1 = px - bcast_double(cd_l, s1);  // |k| >= N10DOP + 1 >= 1
          const double k2 = px - bcast_double(cd_l, s2);
          const double x1 = bcast_double(x_l, s1), x2 = bcast_double(x_l, s2);
          const double v1 = far_quotient<kPrec>(x1, k1 * k1);
          const double v2 = far_quotient<kPrec>(two ? x2 : 0.0, k2 * k2);
          acc[r] += v1;
          acc[r] += v2;
          if (kStats) nev += on ? (two ? 2 : 1) : 0;
        }
        todo = __ballot((ex_edge >> r) & 1u);
        while (todo) {
          const int s1 = __builtin_ctzll(todo);
          todo &= todo - 1;
          const bool two = todo != 0ull;
          const int s2 = two ? __builtin_ctzll(todo) : s1;
          if (two) todo &= todo - 1;
          // every broadcast before any lane-dependent choice
          const double k1 = px - bcast_double(cd_l, s1);
          const double k2 = px - bcast_double(cd_l, s2);
          const double x1 = bcast_double(x_l, s1), x2 = bcast_double(x_l, s2);
          const double n1 = bcast_double(n10_l, s1), n2 = bcast_double(n10_l, s2);
          const double m1 = bcast_double(m_l, s1), m2 = bcast_double(m_l, s2);
          // bitwise, not short-circuit: no branch between the broadcasts
          const bool keep1 = ((int)(fabs(k1) > n1) & (int)(fabs(k1) <= m1)) != 0;
          const bool keep2 = ((int)two & (int)(fabs(k2) > n2) & (int)(fabs(k2) <= m2)) != 0;
          const double v1 = far_quotient<kPrec>(keep1 ? x1 : 0.0, fmax(k1 * k1, 1.0));
          const double v2 = far_quotient<kPrec>(keep2 ? x2 : 0.0, fmax(k2 * k2, 1.0));
          acc[r] += v1;
          acc[r] += v2;
          if (kStats) nev += on ? (keep1 ? 1 : 0) + (keep2 ? 1 : 0) : 0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kWP; ++r) {
      const long w = t0 + 64 * r + ln;
      if (w < W) row[w] = acc[r];
    }
    __syncthreads();  // the queue's reads before its next writes
    qn = 0;
  };

  // ---- step 1: tile expansions, queueing the other entries that reach the
  // tile; steps 2 and 3 whenever the next step could overflow the queue.
  // (kk, j0, hi): the class, first entry and window end of the current step.
  // The next step's heads are loaded while this one is processed, as long as
  // the queue has room for both steps; a step that may be followed by a
  // flush loads its own heads, so no prefetched head is live across one.
  auto settle = [&](int& c, long& s0, long& e) {  // the next non-empty window
    while (s0 >= e && c < kClasses) {
      if (++c < kClasses) {
        s0 = uniform_long(__shfl(bound, 2 * c));
        e = uniform_long(__shfl(bound, 2 * c + 1));
      }
    }
  };
  int kk = 0;
  long j0 = uniform_long(__shfl(bound, 0));
  long hi = uniform_long(__shfl(bound, 1));
  settle(kk, j0, hi);
  for (;;) {
    const bool done = kk == kClasses;
    if (!done && qn + 64 <= kGatherQueue) {
      WingHead h = j0 + lane < hi ? list_cm[j0 + lane] : none;
      for (;;) {
        int kk2 = kk;
        long j02 = j0 + 64, hi2 = hi;
        settle(kk2, j02, hi2);
        // the next step runs without a flush in between only if this step
        // leaves room for 64 more entries whatever it queues
        const bool ahead = kk2 < kClasses && qn + 128 <= kGatherQueue;
        const WingHead hn = ahead && j02 + lane < hi2 ? list_cm[j02 + lane] : none;
        const double K = wm - h.cd;
        const bool expand = far_interior(h, t0i, t1i) && fabs(K) >= kmin && h.xs >= 0.0;
        const bool exact = !(t1i < h.lo || t0i > h.hi) && !expand;
        const unsigned long long em = __ballot(exact);
        if (exact) s_q[qn + __popcll(em & below)] = queue_pack(h, t0i);
        qn += __popcll(em);
        const unsigned long long xm = __ballot(expand);
        if (xm != 0ull) {
          tile_any = true;
          expansion_moments<kT>(expand ? h.xs : 0.0, expand ? K : kmin, kHalf, mom);
          if (kStats && expand) nev += npix > 0 ? npix : 0;
        }
        kk = kk2;
        j0 = j02;
        hi = hi2;
        if (!ahead) break;
        h = hn;
      }
      if (kk < kClasses && qn + 64 <= kGatherQueue) continue;  // the next step loads its own heads
    }
    // the tile moments met so far go to the pixels first, so that they are
    // not live across steps 2 and 3
    if (tile_any) {
      int ln = lane;  // opaque, as in flush
      __asm__ volatile("" : "+v"(ln));
      add_tile_expansion<kT, kWP>(mom, row, t0, W, ln);
    }
    flush();
    tile_any = false;
#pragma unroll
    for (int n = 0; n < kT; ++n) mom[n] = 0.0;
    if (kk == kClasses) break;
  }
  double T = temp[d] > 1.0 ? temp[d] : 1.0;
  double hkt = kHPlanck / (kKBoltz * T);
#pragma unroll
  for (int r = 0; r < kWP; ++r) {
    const long w = t0 + 64 * r + lane;
    if (w < W) {
      double fg = kCLightNm / grid[w];
      row[w] = row[w] * (1.0 - exp(-fg * hkt));
    }
  }
  if (kStats) {
    for (int off = 32; off > 0; off >>= 1) nev += __shfl_down(nev, off);
    if (lane == 0) atomicAdd(evals, nev);
  }
}

// XCD-aware order of the gather's (depth, tile) items: the hardware hands
// workgroup L to XCD L mod 8, so item v(L) gives each XCD one contiguous run
// of items (adjacent tiles of a depth), whose common list heads then stay in
// that XCD's L2.  A bijection on [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ long xcd_item(long L, long n) {
  const long q = n / kXcds, r = n % kXcds;
  const long x = L % kXcds, i = L / kXcds;
  return x * q + (x < r ? x : r) + i;
}


template <bool kStats, int kPrec>
__global__ __launch_bounds__(64, gather_waves<kPrec>()) void k_gather(GatherArgs g,
                                                                    unsigned long long* __restrict__ evals) {
  const long ntiles = gridDim.x;
  const long v = xcd_item(blockIdx.y * ntiles + blockIdx.x, ntiles * gridDim.y);
  const int d = (int)(v / ntiles);
  gather_tile<kStats, kPrec, kWP>(g, d, (v - (long)d * ntiles) * kWT, threadIdx.x, evals);
}

// Persistent form over many spectra: each wave takes (instance, depth, tile)
// items from a queue until none is left.  items[i] .. items[i+1] are
// instance order[i]'s items, depth-major; every item writes only its own
// tile's pixels, so the result does not depend on which wave takes it.
template <int kPrec, int kWPt>
__global__ __launch_bounds__(64, gather_waves<kPrec>()) void k_gather_batch(const GatherArgs* __restrict__ inst,
                                                     const int32_t* __restrict__ order,
                                                     const int64_t* __restrict__ items, int n_inst,
                                                     unsigned long long* __restrict__ counter) {
  constexpr long kTile = 64L * kWPt;
  const int lane = threadIdx.x;
  const long n_items = items[n_inst];
  for (;;) {
    unsigned long long it = 0;
    if (lane == 0) it = atomicAdd(counter, 1ull);
    const long item = uniform_long((long)__shfl(it, 0));
    if (item >= n_items) break;
    int lo = 0, hi = n_inst - 1;  // last i with items[i] <= item
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (items[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const GatherArgs& g = inst[order[lo]];
    const long local = item - items[lo];
    const long ntiles = (g.W + kTile - 1) / kTile;
    const int d = (int)(local / ntiles);
    const long t = local - (long)d * ntiles;
    gather_tile<false, kPrec, kWPt>(g, d, t * kTile, lane, nullptr);
  }
}

// The same items with one workgroup each (item = blockIdx.x), in the queue's
// order: no loop around the tile, so nothing of it stays live between items.
template <int kPrec, int kWPt>
__global__ __launch_bounds__(64, gather_waves<kPrec>()) void k_gather_items(
    const GatherArgs* __restrict__ inst, const int32_t* __restrict__ order, const int64_t* __restrict__ items,
    int n_inst) {
  constexpr long kTile = 64L * kWPt;
  const long item = blockIdx.x;
  int lo = 0, hi = n_inst - 1;  // last i with items[i] <= item
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (items[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  const GatherArgs& g = inst[order[lo]];
  const long local = item - items[lo];
  const long ntiles = (g.W + kTile - 1) / kTile;
  const int d = (int)(local / ntiles);
  const long t = local - (long)d * ntiles;
  gather_tile<false, kPrec, kWPt>(g, d, t * kTile, threadIdx.x, nullptr);
}

// Wing statistics over the lists: pairs, sum and max of MAXSTEP.
__global__ void k_wing_stats(long n_entries_cap, int D, long L, const int64_t* __restrict__ regions,
                             const WingHead* __restrict__ list_cm, unsigned long long* out) {
  int d = blockIdx.y;
  long beg = regions[d * (kClasses + 1)], end = regions[d * (kClasses + 1) + kClasses];
  unsigned long long cnt = 0, sum = 0, mx = 0;
  for (long j = beg + blockIdx.x * (long)blockDim.x + threadIdx.x; j < end;
       j += (long)gridDim.x * blockDim.x) {
    const WingHead h = list_cm[j];
    unsigned long long m = (unsigned long long)(h.lo <= h.hi ? h.hi - h.lo : h.nhi - h.nlo) / 2;
    ++cnt;
    sum += m;
    mx = m > mx ? m : mx;
  }
  atomicAdd(out + 0, cnt);
  atomicAdd(out + 1, sum);
  atomicMax(out + 2, mx);
}

__global__ void k_stats_to_double(const unsigned long long* in, double* out) {
  if (threadIdx.x == 0) {
    out[0] = (double)in[3];
    out[1] = (double)in[0];
    out[2] = (double)in[1];
    out[3] = (double)in[2];
    out[4] = (double)in[4];
  }
}

// --------------------------------------------------------------- JOSH solve
// Inputs of one solve_josh_flux call, read through a depth index list so the
// masked/reversed arrays of solve_lte_frequency are never materialised.
struct SolveIn {
  const double *acont, *scont, *aline, *sline, *sigmac, *sigmal;  // strided by `stride`
  const double* rhox;                                             // per depth
  long stride;
  const double* scont_local;  // per-lane Planck values [n] (NULL: use scont)
  const double* sline_local;  // per-lane line source [n] (NULL: use sline)
  bool zero_line;             // aline = sigmal = 0 (continuum-only solve)
};

// One wavelength per lane.  xs lives in LDS ([k][lane]) so the Gauss-Seidel
// sweep can index it dynamically; COEFJ is read with wave-uniform addresses.
template <int kLanes>
__device__ double josh_solve(const SolveIn& in, const int* idx, int n, DeviceTables tb,
                             float (*s_xs)[kLanes], int lane, double* abtot, double* alpha,
                             double* snubar, double* rho, double* taunu, double* pa, double* pb,
                             double* pc, int* iters, int* took401) {
  for (int j = 0; j < n; ++j) {
    long di = (long)idx[j] * in.stride;
    double ac = in.acont[di];
    double al = in.zero_line ? 0.0 : in.aline[di];
    double sc = in.sigmac[di];
    double sl = in.zero_line ? 0.0 : in.sigmal[di];
    double scn = in.scont_local ? in.scont_local[j] : in.scont[di];
    double sln = in.sline_local ? in.sline_local[j] : in.sline[di];
    double ab = ac + al + sc + sl;
    ab = ab > kEps38 ? ab : (ab != ab ? ab : kEps38);
    abtot[j] = ab;
    double sct = sc + sl;
    double a = ab > 0.0 ? sct / ab : 0.0;
    a = a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
    alpha[j] = a;
    double dn = ac + al;
    snubar[j] = dn > 0 ? (ac * scn + al * sln) / dn : scn;
    rho[j] = in.rhox[idx[j]];
  }
  bool rev = n > 1 && rho[0] > rho[n - 1];
  double start;
  if (rev) {
    for (int j = 0; j < n / 2; ++j) {
      double t = rho[j]; rho[j] = rho[n - 1 - j]; rho[n - 1 - j] = t;
      t = abtot[j]; abtot[j] = abtot[n - 1 - j]; abtot[n - 1 - j] = t;
      t = snubar[j]; snubar[j] = snubar[n - 1 - j]; snubar[n - 1 - j] = t;
      t = alpha[j]; alpha[j] = alpha[n - 1 - j]; alpha[n - 1 - j] = t;
    }
    start = abtot[n - 1] * rho[n - 1];
  } else {
    start = abtot[0] * rho[0];
  }
  // physics/josh_solver.py:170-189 (_integ)
  parcoe(abtot, rho, n, pa, pb, pc);
  taunu[0] = start;
  for (int i = 0; i + 1 < n; ++i) {
    double dx = rho[i + 1] - rho[i];
    double term = pa[i] + 0.5 * pb[i] * (rho[i + 1] + rho[i]);
    term += (pc[i] / 3.0) * ((rho[i + 1] + rho[i]) * rho[i + 1] + rho[i] * rho[i]);
    taunu[i + 1] = taunu[i] + term * dx;
  }
  if (taunu[0] > tb.xtau[kNxtau - 1]) {
    // label-401 branch (josh_solver.py:700-727): DERIV iteration on TAUNU.
    double* snu = pa;
    double* hnu = pb;
    double* jnu = pc;
    double* tmp = abtot;
    for (int j = 0; j < n; ++j) snu[j] = snubar[j];
    int it = 0;
    for (int l = 0; l < kNxtau; ++l) {
      deriv(taunu, snu, n, hnu);
      for (int j = 0; j < n; ++j) hnu[j] = hnu[j] / 3.0;
      deriv(taunu, hnu, n, jnu);
      for (int j = 0; j < n; ++j) {
        double jn = jnu[j] + snu[j];
        double sn = (1.0 - alpha[j]) * snubar[j] + alpha[j] * jn;
        double den = fabs(sn) > kEps38 ? fabs(sn) : kEps38;
        tmp[j] = fabs(sn - snu[j]) / den;
        snu[j] = sn;
      }
      // numpy pairwise summation (n <= 128: eight partial sums)
      double err;
      if (n < 8) {
        err = 0.0;
        for (int j = 0; j < n; ++j) err += tmp[j];
      } else {
        double r0 = tmp[0], r1 = tmp[1], r2 = tmp[2], r3 = tmp[3], r4 = tmp[4], r5 = tmp[5],
               r6 = tmp[6], r7 = tmp[7];
        int j = 8;
        for (; j < n - (n % 8); j += 8) {
          r0 += tmp[j]; r1 += tmp[j + 1]; r2 += tmp[j + 2]; r3 += tmp[j + 3];
          r4 += tmp[j + 4]; r5 += tmp[j + 5]; r6 += tmp[j + 6]; r7 += tmp[j + 7];
        }
        err = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
        for (; j < n; ++j) err += tmp[j];
      }
      it = l + 1;
      if (err < 1.0e-5) break;
    }
    *iters = it;
    *took401 = 1;
    return hnu[0];
  }
  *took401 = 0;
  double xsbar[kNxtau], xalpha[kNxtau];
  map1_xtau(taunu, snubar, n, tb.xtau, xsbar);
  map1_xtau(taunu, alpha, n, tb.xtau, xalpha);
  double s0 = snubar[0] > kEps38 ? snubar[0] : kEps38;
  double a0 = alpha[0] < 0.0 ? 0.0 : (alpha[0] > 1.0 ? 1.0 : alpha[0]);
  float xal32[kNxtau], xsb32[kNxtau], dg32[kNxtau];
#pragma unroll
  for (int k = 0; k < kNxtau; ++k) {
    double xb = xsbar[k] > kEps38 ? xsbar[k] : kEps38;
    double xa = xalpha[k] < 0.0 ? 0.0 : (xalpha[k] > 1.0 ? 1.0 : xalpha[k]);
    if (tb.xtau[k] < taunu[0]) {
      xb = s0;
      xa = a0;
    }
    s_xs[k][lane] = (float)xb;
    xal32[k] = (float)xa;
    xsb32[k] = (float)(xb * (1.0 - xa));
    dg32[k] = 1.0f - xal32[k] * tb.diag32[k];
  }
  // physics/josh_solver.py:24-92: float32 arrays, float64 accumulation.
  const double tol = (double)(float)1.0e-5;
  const double eps = (double)(float)kEps38;
  int it = kNxtau;
  for (int sweep = 0; sweep < kNxtau; ++sweep) {
    bool err = false;
#pragma unroll
    for (int kr = 0; kr < kNxtau; ++kr) {
      const int k = kNxtau - 1 - kr;
      const float* row = tb.coefj32 + k * kNxtau;
      double delxs = 0.0;
      for (int m = 0; m < kNxtau; ++m) delxs += (double)(row[m] * s_xs[m][lane]);
      float xk = s_xs[k][lane];
      delxs = (delxs * (double)xal32[k] + (double)xsb32[k] - (double)xk) / (double)dg32[k];
      double ex = xk != 0.0f ? fabs(delxs / (double)xk) : INFINITY;
      if (ex > tol) err = true;
      double nv = (double)xk + delxs;
      if (eps > nv) nv = eps;
      s_xs[k][lane] = (float)nv;
    }
    if (!err) {
      it = sweep + 1;
      break;
    }
  }
  *iters = it;
  double flux = 0.0;
  for (int k = 0; k < kNxtau; ++k) flux += tb.ch[k] * (double)s_xs[k][lane];
  return flux;
}

__device__ __forceinline__ double planck_one(double freq, double f3, double T) {
  double hkt = kHPlanck / (kKBoltz * T);
  double ehvkt = exp(-freq * hkt);
  double stim = 1.0 - ehvkt;
  double b = 1.47439e-2 * f3 * ehvkt / stim;
  return b != b ? 0.0 : b;
}

constexpr int kRtLanes = 64;

// engine/radiative.py:42-218 (solve_lte_frequency) per wavelength, general
// form with per-lane depth arrays: runs only for the wavelengths the streaming
// kernels list (masked or non-increasing depths, fewer than 4 depths, or a
// label-401 solve), packed densely into waves.  list == NULL runs every one.
// RT inputs and outputs of one spectrum within a set of spectra that share
// a global wavelength index g = g0 + w (w < W).  to_lambda: store F_lambda
// (engine/opacity.py:3852-3861) instead of F_nu.
struct RtArgs {
  long g0, W;
  int D, to_lambda;
  const double *wl, *temp, *rhox, *ca, *cs, *la, *ls, *src;
  double *ft, *fc;
  int32_t* iters;
};

// Index of the spectrum holding global wavelength g (rt[] ascending in g0).
__device__ __forceinline__ int rt_find(const RtArgs* __restrict__ rt, int n, long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (rt[mid].g0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void rt_store(const RtArgs& a, long w, double ftot, double fcont, int it) {
  if (a.to_lambda) {
    double w2 = a.wl[w] * a.wl[w];
    double f = kCLightNm / (w2 > 1e-40 ? w2 : 1e-40);
    ftot = ftot * f;
    fcont = fcont * f;
  }
  a.ft[w] = ftot;
  a.fc[w] = fcont;
  if (a.iters) a.iters[w] = it;
}

__global__ void k_store_rtargs(RtArgs a, RtArgs* dst) {
  if (threadIdx.x == 0) *dst = a;
}

__global__ __launch_bounds__(kRtLanes) void k_lte(const RtArgs* __restrict__ rt, int n_rt,
                                                  const int64_t* __restrict__ list,
                                                  const int* __restrict__ list_n, DeviceTables tb) {
  __shared__ float s_xs[kNxtau][kRtLanes];
  const long e = blockIdx.x * (long)kRtLanes + threadIdx.x;
  if (e >= *list_n) return;
  const long g = list[e];
  const RtArgs& a = rt[rt_find(rt, n_rt, g)];
  const long w = g - a.g0, W = a.W;
  const int D = a.D;
  const double *wl = a.wl, *temp = a.temp, *rhox = a.rhox, *ca = a.ca, *cs = a.cs, *la = a.la, *ls = a.ls,
               *src = a.src;
  int lane = threadIdx.x;
  const double maxr4 = 3.4e38;
  int idx[kMaxDepth];
  int n = 0;
  for (int d = 0; d < D; ++d) {
    long i = (long)d * W + w;
    double r = rhox[d], a1 = ca[i], a2 = cs[i], a3 = la[i], a4 = ls[i];
    bool okd = r >= 0.0 && isfinite(r) && isfinite(a1) && isfinite(a2) && isfinite(a3) &&
               isfinite(a4) && a1 < maxr4 && a2 < maxr4 && a3 < maxr4 && a4 < maxr4;
    if (okd && r > 0) idx[n++] = d;
  }
  if (n == 0) {
    rt_store(a, w, 0.0, 0.0, 0);
    return;
  }
  if (n > 1 && !(rhox[idx[0]] < rhox[idx[n - 1]])) {
    for (int j = 0; j < n / 2; ++j) {
      int t = idx[j];
      idx[j] = idx[n - 1 - j];
      idx[n - 1 - j] = t;
    }
  }
  double wv = wl[w];
  double freq = kCLightNm / (wv > 1e-12 ? wv : 1e-12);
  double f3 = pow(freq / 1.0e15, 3.0);
  double pl[kMaxDepth], sl[kMaxDepth];
  for (int j = 0; j < n; ++j) {
    int d = idx[j];
    pl[j] = planck_one(freq, f3, temp[d]);
    if (src) {
      double s = src[(long)d * W + w];
      sl[j] = isfinite(s) ? s : pl[j];
    } else {
      sl[j] = pl[j];
    }
  }
  double abtot[kMaxDepth], alpha[kMaxDepth], snubar[kMaxDepth], rho[kMaxDepth], taunu[kMaxDepth];
  double pa[kMaxDepth], pb[kMaxDepth], pc[kMaxDepth];
  SolveIn cin{ca + w, nullptr, la + w, nullptr, cs + w, ls + w, rhox, W, pl, pl, true};
  int it_c = 0, it_t = 0, b401c = 0, b401t = 0;
  double fcont = josh_solve<kRtLanes>(cin, idx, n, tb, s_xs, lane, abtot, alpha, snubar, rho, taunu,
                                      pa, pb, pc, &it_c, &b401c);
  SolveIn tin{ca + w, nullptr, la + w, nullptr, cs + w, ls + w, rhox, W, pl, sl, false};
  double ftot = josh_solve<kRtLanes>(tin, idx, n, tb, s_xs, lane, abtot, alpha, snubar, rho, taunu,
                                     pa, pb, pc, &it_t, &b401t);
  rt_store(a, w, ftot, fcont, it_c + it_t + (b401t ? 1000 : 0));
}

// ----------------------------------------------------- streaming RT kernels
// Per-depth values of one solve_josh_flux call read straight from the [D][W]
// inputs (josh_solver.py:411-441): abtot, alpha, snubar; rho = rhox.
struct ColSrc {
  const double *ca, *cs, *la, *ls, *src, *temp, *rhox;
  long W;
  double freq, f3;
  bool total;  // false: continuum-only solve (aline = sigmal = 0, sline = B)
  bool check;  // test every depth read (the inputs finite and below 3.4e38, rho > 0) into ok
  bool ok;
  __device__ __forceinline__ void get(int j, double* f, double* x, double* sb, double* al) {
    long i = (long)j * W;
    double ac = ca[i], sc = cs[i];
    double aln = total ? la[i] : 0.0;
    double sln = total ? ls[i] : 0.0;
    if (check) {
      const double r = rhox[j], maxr4 = 3.4e38;
      ok = ok && r > 0.0 && isfinite(r) && isfinite(ac) && isfinite(sc) && isfinite(aln) && isfinite(sln) &&
           ac < maxr4 && sc < maxr4 && aln < maxr4 && sln < maxr4;
    }
    double B = planck_one(freq, f3, temp[j]);
    double sl = B;
    if (total && src) {
      double s = src[i];
      sl = isfinite(s) ? s : B;
    }
    double ab = ac + aln + sc + sln;
    ab = ab > kEps38 ? ab : (ab != ab ? ab : kEps38);
    double a = ab > 0.0 ? (sc + sln) / ab : 0.0;
    a = a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
    double dn = ac + aln;
    *f = ab;
    *x = rhox[j];
    *sb = dn > 0 ? (ac * B + aln * sl) / dn : B;
    *al = a;
  }
};

// MAP1 results -> the float32 iteration inputs (josh_solver.py:617-680):
// XSBAR = max(., EPS), XALPHA clipped, surface masking for XTAU < TAUNU(0).
struct ColOut {
  float *xs0, *xal, *xsbm;  // [51][W], column w already applied
  long W;
  const double* xtau;
  double tau0, s0, a0;
  __device__ __forceinline__ void put(int k, double xsbar, double xalpha) {
    double xb = xsbar > kEps38 ? xsbar : kEps38;
    double xa = xalpha < 0.0 ? 0.0 : (xalpha > 1.0 ? 1.0 : xalpha);
    if (xtau[k] < tau0) {
      xb = s0;
      xa = a0;
    }
    long i = (long)k * W;
    xs0[i] = (float)xb;
    xal[i] = (float)xa;
    xsbm[i] = (float)(xb * (1.0 - xa));
  }
};

struct TauOnlyOut {
  __device__ __forceinline__ void put(int, double, double) {}
};

constexpr int kRtBlock = 256;

// Fast-path set-up of both JOSH solves per wavelength: the streaming INTEG +
// MAP1 of the total solve first, whose reads of the five inputs also check
// every depth, then the continuum-only solve (two of the five again): 7
// reads per depth.  flags[w]: 1 = needs the general kernel (a depth fails
// the checks, D < 4, rho not increasing, or either solve takes label 401).
__global__ __launch_bounds__(kRtBlock, 3) void k_rt_prep(const RtArgs* __restrict__ rt, int n_rt, long G,
                                                      const double* __restrict__ xtau, float* ws,
                                                      uint8_t* __restrict__ flags,
                                                      int64_t* __restrict__ list, int* __restrict__ list_n) {
  const long g = blockIdx.x * (long)kRtBlock + threadIdx.x;
  if (g >= G) return;
  const RtArgs& a = rt[rt_find(rt, n_rt, g)];
  const long w = g - a.g0, W = a.W;
  const int D = a.D;
  const double *wl = a.wl, *temp = a.temp, *rhox = a.rhox, *ca = a.ca, *cs = a.cs, *la = a.la, *ls = a.ls,
               *src = a.src;
  bool fl = !(D >= 4 && rhox[0] < rhox[D - 1]);
  if (!fl) {
    double wv = wl[w];
    double freq = kCLightNm / (wv > 1e-12 ? wv : 1e-12);
    double f3 = pow(freq / 1.0e15, 3.0);
    for (int s = 1; s >= 0 && !fl; --s) {
      ColSrc in{ca + w, cs + w, la + w, ls + w, src ? src + w : nullptr, temp, rhox, W, freq, f3, s == 1, s == 1,
                true};
      double f0, x0, sb0, al0;
      in.get(0, &f0, &x0, &sb0, &al0);
      float* base = ws + (long)s * 3 * kNxtau * G + g;
      ColOut out{base, base + (long)kNxtau * G, base + 2L * kNxtau * G, G, xtau, 0.0,
                 sb0 > kEps38 ? sb0 : kEps38, al0 < 0.0 ? 0.0 : (al0 > 1.0 ? 1.0 : al0)};
      out.tau0 = f0 * x0;
      if (out.tau0 > xtau[kNxtau - 1] || !in.ok) {
        fl = true;  // label 401 or a failed check: general kernel
        break;
      }
      double tau0;
      rt::integ_map1_stream(D, xtau, in, out, &tau0);
      fl = !in.ok;
    }
  }
  flags[g] = fl ? 1 : 0;
  if (fl) list[atomicAdd(list_n, 1)] = g;  // order is irrelevant: one slot per wavelength
}

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0) *p = 0;
}

// Diagnostic: far-wing quotient variants against IEEE division.  variant 0 is
// the kernels' div_pos; 1 and 2 use one and zero Newton steps on the reciprocal
// before the remainder correction.
__global__ void k_check_div(long n, int variant, const double* x, const double* b, double* q_fast,
                            double* q_ieee) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xv = x[i], bv = b[i], q;
  if (variant == 5 || variant == 6) {
    // b holds k; each aligned group of 8 holds consecutive k (far_chunk)
    const int p = (int)(i & 7);
    double out[8];
    if (variant == 5) far_chunk<8, kFarNewton, false>(xv, bv - (double)p, 0, 0.0, 0.0, out);
    else far_chunk<8, 0, false>(xv, bv - (double)p, 0, 0.0, 0.0, out);
    q = out[0];
#pragma unroll
    for (int j = 1; j < 8; ++j) q = j == p ? out[j] : q;
    q_fast[i] = q;
    q_ieee[i] = xv / (bv * bv);
    return;
  }
  if (variant == 0) {
    q = div_pos(xv, bv);
  } else if (variant == 3) {
    q = xv * __builtin_amdgcn_rcp(bv);
  } else if (variant == 4) {
    double r = __builtin_amdgcn_rcp(bv);
    r = __builtin_fma(r, __builtin_fma(-bv, r, 1.0), r);
    q = xv * r;
  } else {
    double r = __builtin_amdgcn_rcp(bv);
    if (variant == 1) r = __builtin_fma(r, __builtin_fma(-bv, r, 1.0), r);
    double qq = xv * r;
    q = __builtin_fma(__builtin_fma(-bv, qq, xv), r, qq);
  }
  q_fast[i] = q;
  q_ieee[i] = xv / bv;
}

// Gauss-Seidel scattering iteration (josh_solver.py:24-92, Numba's float32
// arrays with float64 accumulation) and flux = CH . XS, both solves.
// Four waves per workgroup share one LDS copy of COEFJ (10.6 KB) beside their
// own XS columns (51 floats per lane): 62.8 KB per workgroup, two per CU, so
// the CU holds the 8 waves its registers allow (64-lane workgroups, each with
// its own COEFJ, fitted 6).
constexpr int kRtIterBlock = 256;

__global__ __launch_bounds__(kRtIterBlock) void k_rt_iter(const RtArgs* __restrict__ rt, int n_rt, long G,
                                                          const uint8_t* __restrict__ flags,
                                                          const float* __restrict__ ws, DeviceTables tb) {
  __shared__ float s_xs[kNxtau][kRtIterBlock];
  __shared__ float s_cj[kNxtau * kNxtau + kNxtau];  // COEFJ rows, then its diagonal
  for (int i = threadIdx.x; i < kNxtau * kNxtau; i += kRtIterBlock) s_cj[i] = tb.coefj32[i];
  if (threadIdx.x < kNxtau) s_cj[kNxtau * kNxtau + threadIdx.x] = tb.diag32[threadIdx.x];
  __syncthreads();
  const long g = blockIdx.x * (long)kRtIterBlock + threadIdx.x;
  if (g >= G) return;
  if (flags[g] != 0) return;
  const int lane = threadIdx.x;
  const double tol = (double)(float)1.0e-5;
  const double eps = (double)(float)kEps38;
  int it_sum = 0;
  double fsol[2];
  for (int s = 0; s < 2; ++s) {
    const float* base = ws + (long)s * 3 * kNxtau * G + g;
    float xal[kNxtau], xsb[kNxtau];
#pragma unroll
    for (int k = 0; k < kNxtau; ++k) {
      s_xs[k][lane] = base[(long)k * G];
      xal[k] = base[(long)(kNxtau + k) * G];
      xsb[k] = base[(long)(2 * kNxtau + k) * G];
    }
    int it = kNxtau;
    for (int sweep = 0; sweep < kNxtau; ++sweep) {
      bool err = false;
#pragma unroll
      for (int kr = 0; kr < kNxtau; ++kr) {
        const int k = kNxtau - 1 - kr;
        // compiler-only fence: keeps the COEFJ row reads inside the sweep
        // instead of hoisting all 2601 of them into registers
        __asm__ volatile("" ::: "memory");
        const float* row = s_cj + k * kNxtau;
        double delxs = 0.0;
#pragma unroll 3
        for (int m = 0; m < kNxtau; ++m) delxs += (double)(row[m] * s_xs[m][lane]);
        const float dg = 1.0f - xal[k] * s_cj[kNxtau * kNxtau + k];
        const float xk = s_xs[k][lane];
        delxs = (delxs * (double)xal[k] + (double)xsb[k] - (double)xk) / (double)dg;
        double ex = xk != 0.0f ? fabs(delxs / (double)xk) : INFINITY;
        if (ex > tol) err = true;
        double nv = (double)xk + delxs;
        if (eps > nv) nv = eps;
        s_xs[k][lane] = (float)nv;
      }
      if (!err) {
        it = sweep + 1;
        break;
      }
    }
    double flux = 0.0;
    for (int k = 0; k < kNxtau; ++k) flux += tb.ch[k] * (double)s_xs[k][lane];
    fsol[s] = flux;
    it_sum += it;
  }
  const RtArgs& a = rt[rt_find(rt, n_rt, g)];
  rt_store(a, g - a.g0, fsol[1], fsol[0], it_sum);
}

// solve_josh_flux over independent columns (physics/josh_solver.py:367-767).
__global__ __launch_bounds__(kRtLanes) void k_josh_batch(int n, long ncol, const double* acont,
                                                         const double* scont, const double* aline,
                                                         const double* sline, const double* sigmac,
                                                         const double* sigmal, const double* rhox,
                                                         double* flux, int32_t* iters,
                                                         DeviceTables tb) {
  __shared__ float s_xs[kNxtau][kRtLanes];
  long c = blockIdx.x * (long)kRtLanes + threadIdx.x;
  if (c >= ncol) return;
  int idx[kMaxDepth];
  for (int j = 0; j < n; ++j) idx[j] = j;
  double abtot[kMaxDepth], alpha[kMaxDepth], snubar[kMaxDepth], rho[kMaxDepth], taunu[kMaxDepth];
  double pa[kMaxDepth], pb[kMaxDepth], pc[kMaxDepth];
  SolveIn in{acont + c, scont + c, aline + c, sline + c, sigmac + c, sigmal + c, rhox, ncol,
             nullptr, nullptr, false};
  int it = 0, b401 = 0;
  double f = 0.0;
  if (n > 0)
    f = josh_solve<kRtLanes>(in, idx, n, tb, s_xs, threadIdx.x, abtot, alpha, snubar, rho, taunu,
                             pa, pb, pc, &it, &b401);
  flux[c] = f;
  if (iters) iters[c] = it + (b401 ? 1000 : 0);
}

__global__ void k_flux_to_lambda(long W, const double* __restrict__ wl, double* flux) {
  long w = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (w >= W) return;
  double w2 = wl[w] * wl[w];
  flux[w] = flux[w] * (kCLightNm / (w2 > 1e-40 ? w2 : 1e-40));
}

unsigned int blocks_for(long n, int b) { return (unsigned int)((n + b - 1) / b); }

}  // namespace

// ----------------------------------------------------------------- context
struct synthe_ctx {
  int device = 0;
  double* d_tab = nullptr;  // h0 h1 h2 extab extabf xtau ch
  float* d_tab32 = nullptr; // coefj32 diag32
  DeviceTables tb{};
  std::vector<double> h0, h1, h2, extab, extabf;
  void* ws = nullptr;
  size_t ws_size = 0;
  void* scratch = nullptr;  // device address of the static RT scratch on this device
  void* hstage = nullptr;   // hydrogen-line staging (synthe_hydrogen_lines), freed by hstage_free
  void (*hstage_free)(void*) = nullptr;
  void* f19stage = nullptr;  // fort.19 staging (synthe_fort19_lines), freed by f19stage_free
  void (*f19stage_free)(void*) = nullptr;
  // kernel timing: event pairs around each wing-gather launch while enabled
  static constexpr int kTimedMax = 256;
  int precision = SYNTHE_PREC_FP64;  // profile precision of synthe_asynth
  bool timing = false;
  int n_timed = 0;
  hipEvent_t ev_beg[kTimedMax] = {};
  hipEvent_t ev_end[kTimedMax] = {};
};

// Records ev_beg[i] (before = true) or ev_end[i] around the next timed launch.
static int timing_mark(synthe_ctx* ctx, hipStream_t st, bool before) {
  if (!ctx->timing || ctx->n_timed >= synthe_ctx::kTimedMax) return SYNTHE_OK;
  int i = ctx->n_timed;
  hipEvent_t* e = before ? &ctx->ev_beg[i] : &ctx->ev_end[i];
  if (!*e) HIP_TRY(hipEventCreate(e));
  HIP_TRY(hipEventRecord(*e, st));
  if (!before) ctx->n_timed = i + 1;
  return SYNTHE_OK;
}

static int ensure_ws(synthe_ctx* ctx, size_t bytes) {
  if (ctx->ws_size >= bytes) return SYNTHE_OK;
  if (ctx->ws) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_size = 0;
  }
  size_t sz = bytes + bytes / 8 + (1 << 20);
  if (hipMalloc(&ctx->ws, sz) != hipSuccess) {
    ctx->ws = nullptr;
    return fail(SYNTHE_ENOMEM, "workspace allocation of " + std::to_string(sz) + " bytes failed");
  }
  ctx->ws_size = sz;
  return SYNTHE_OK;
}

static constexpr size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }


static int check_lines_depths(const synthe_lines* ln, const synthe_depths* dp, const char* who) {
  if (!ln || !dp) return fail(SYNTHE_EINVAL, std::string(who) + ": lines/depths is NULL");
  if (ln->n_lines < 0 || dp->n_depths <= 0 || dp->n_depths > 65535 || dp->n_ion <= 0 || dp->n_elem <= 0)
    return fail(SYNTHE_EINVAL, std::string(who) + ": bad sizes");
  if (ln->n_lines > 0x7fffffffL)  // list entries of a depth are indexed in 32 bits
    return fail(SYNTHE_ERANGE, std::string(who) + ": more than 2^31 lines");
  if (ln->n_lines > 0 &&
      (!ln->wl || !ln->gf || !ln->cgf || !ln->gamma_rad || !ln->gamma_stark || !ln->gamma_vdw ||
       !ln->elo || !ln->elem || !ln->ion || !ln->ltype || !ln->process || !ln->center_t ||
       !ln->center_w))
    return fail(SYNTHE_EINVAL, std::string(who) + ": NULL line array");
  if (!dp->temperature || !dp->rhox || !dp->rho || !dp->xne || !dp->txnxn || !dp->hckt || !dp->pop ||
      !dp->dop)
    return fail(SYNTHE_EINVAL, std::string(who) + ": NULL depth array");
  return SYNTHE_OK;
}

static LinesDev to_dev(const synthe_lines* l) {
  return LinesDev{l->n_lines, l->wl,   l->gf,     l->cgf,   l->gamma_rad, l->gamma_stark, l->gamma_vdw,
                  l->elo,     l->elem, l->ion,    l->ltype, l->process,   l->center_t,    l->center_w};
}

static DepthsDev to_dev(const synthe_depths* d) {
  return DepthsDev{(int)d->n_depths, d->temperature, d->rhox, d->rho, d->xne, d->txnxn, d->hckt,
                   d->pop,           d->dop,         (int)d->n_ion, (int)d->n_elem};
}


// ------------------------------------------------------------------ adjoint
// Reverse-mode derivatives of the synthesis for the autograd wrapper.  The
// forward pass is not taped: the backward recomputes the wing set-up, the
// lists and every profile value it needs.

constexpr int kAdjMaxDepth = 96;  // depth limit of the RT vjp (private memory per lane)

// Voigt-zone profile H at offset k (near_wing's formula, times kappa0 there)
// and its partial derivatives in the damping a and in x = k * dvoigt: exact
// for the table and asymptotic forms (the table index held fixed; dH/dx uses
// d/dx exp(-x^2) and the centred difference of H1), central differences of
// voigt() for a >= 0.2.
__device__ void near_wing_partials(const WingParams& p, long k, const DeviceTables& tb, double* H, double* dHda,
                                   double* dHdx) {
  const double x = (double)k * p.dvoigt;
  const double a = p.adamp;
  if (a < 0.2) {
    if (x > 10.0) {
      *H = 0.5642 * a / (x * x);
      *dHda = 0.5642 / (x * x);
      *dHdx = -2.0 * 0.5642 * a / (x * x * x);
      return;
    }
    long idx = tabi_index(200.0 * p.dvoigt, k);
    idx = idx < 0 ? 0 : (idx > kVoigtN - 1 ? kVoigtN - 1 : idx);
    *H = tb.h0[idx] + a * tb.h1[idx];
    *dHda = tb.h1[idx];
    const long i0 = idx > 0 ? idx - 1 : 0, i1 = idx < kVoigtN - 1 ? idx + 1 : kVoigtN - 1;
    const double dh1 = (tb.h1[i1] - tb.h1[i0]) * 200.0 / (double)(i1 - i0);
    *dHdx = -2.0 * x * exp(-x * x) + a * dh1;
    return;
  }
  *H = voigt(x, a, tb.h0, tb.h1, tb.h2);
  const double ea = 1e-6 * (a > 1e-3 ? a : 1e-3), ex = 1e-6 * (x > 1e-3 ? x : 1e-3);
  *dHda = (voigt(x, a + ea, tb.h0, tb.h1, tb.h2) - voigt(x, a - ea, tb.h0, tb.h1, tb.h2)) / (2.0 * ea);
  *dHdx = (voigt(x + ex, a, tb.h0, tb.h1, tb.h2) - voigt(x > ex ? x - ex : 0.0, a, tb.h0, tb.h1, tb.h2)) /
          (x + ex - (x > ex ? x - ex : 0.0));
}

__device__ __forceinline__ double stim_at(const double* wl, const double* temp, int d, long w) {
  const double T = temp[d] > 1.0 ? temp[d] : 1.0;
  const double hkt = kHPlanck / (kKBoltz * T);
  return 1.0 - exp(-(kCLightNm / wl[w]) * hkt);
}

// u[d][w] = dL/d(asynth) * STIM: the weight of each pre-STIM contribution.
__global__ void k_adj_weights(int D, long W, const double* __restrict__ wl, const double* __restrict__ temp,
                              const double* __restrict__ g_asy, double* __restrict__ u) {
  const long w = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (w >= W) return;
  u[(long)d * W + w] = g_asy[(long)d * W + w] * stim_at(wl, temp, d, w);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  return v;
}

// Line centres (line_opacity.py:1091-1102): transp = kappa0 * vc(a) at the
// centre pixel, vc = 1 - 1.128 a or H(0, a) (line_opacity.py:123-184).
__global__ void k_adj_centre(LinesDev ln, DepthsDev dp, long W, const double* __restrict__ transp,
                             const uint8_t* __restrict__ valid, const double* __restrict__ u, DeviceTables tb,
                             double* __restrict__ s_k0, double* __restrict__ s_a, double* __restrict__ k0_out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long L = ln.n;
  const int D = dp.n;
  if (i >= L * D) return;
  if (!valid[i]) return;
  const long l = i / D;
  const int d = (int)(i - l * D);
  const long c = ln.center_t[l];
  if (ln.ltype[l] != 0) return;
  const int ei = ln.elem[l], nel = ln.ion[l];
  const long pi = ((long)d * dp.nion + (nel - 1)) * dp.nelem + ei;
  const double dv = dp.dop[pi];
  const double gt = ln.grad[l] + ln.gstark[l] * dp.xne[d] + ln.gvdw[l] * dp.txnxn[d];
  const double a = gt / dv;
  double vc, dvc;
  if (a < 0.2) {
    vc = 1.0 - 1.128 * a;
    dvc = -1.128;
  } else {
    vc = voigt(0.0, a, tb.h0, tb.h1, tb.h2);
    const double ea = 1e-6 * a;
    dvc = (voigt(0.0, a + ea, tb.h0, tb.h1, tb.h2) - voigt(0.0, a - ea, tb.h0, tb.h1, tb.h2)) / (2.0 * ea);
  }
  const double post = transp[i] / vc;
  if (k0_out) k0_out[i] = post;
  if (c < 0 || c >= W) return;
  const double uc = u[(long)d * W + c];
  s_k0[i] += uc * vc;
  s_a[i] += uc * post * dvc;
}

// Per entry (line, depth): dL/dkappa0, dL/da and dL/d(Doppler width) summed
// over its Voigt-zone pixels 1 <= |k| <= r (the k_near set, r = min(N10DOP,
// MAXSTEP) with a far wing, MAXSTEP without; c = kappa0 H(k dvoigt, a)).  One
// entry per lane, the depth's lists in order: the profile's partials depend
// on |k| only, so each is evaluated once and weighted by u at both c - k and
// c + k (pixels off the grid weigh 0).  Each pair has one list entry, and
// the kernels writing the pair's sums run in stream order, so the sums are
// plain adds: run-to-run identical.
__global__ __launch_bounds__(64) void k_near_adj(long W, int D, const WingHead* __restrict__ list_cm,
                                                 const WingParams* __restrict__ list_p,
                                                 const int32_t* __restrict__ list_l,
                                                 const int64_t* __restrict__ regions, DeviceTables tb,
                                                 const double* __restrict__ u, const double* __restrict__ wl,
                                                 double* __restrict__ s_k0, double* __restrict__ s_a,
                                                 double* __restrict__ s_dv) {
  const int d = blockIdx.y;
  const long beg = regions[(long)d * (kClasses + 1)], end = regions[(long)d * (kClasses + 1) + kClasses];
  const double resolu = W > 1 && wl[1] / wl[0] > 1.0 ? 1.0 / (wl[1] / wl[0] - 1.0) : 300000.0;
  const double* ud = u + (long)d * W;
  for (long j = beg + (long)blockIdx.x * 64 + threadIdx.x; j < end; j += (long)gridDim.x * 64) {
    const WingHead h = list_cm[j];
    const int lim = (h.nhi - h.nlo) >> 1;  // the zone's half-width
    const long c = head_c(h);
    if (lim < 1) continue;
    const WingParams pr = list_p[j];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int ka = 1; ka <= lim; ++ka) {
      const long wb = c - ka, wr = c + ka;
      const double uw = (wb >= 0 && wb < W ? ud[wb] : 0.0) + (wr >= 0 && wr < W ? ud[wr] : 0.0);
      if (uw == 0.0) continue;
      double H, dHa, dHx;
      near_wing_partials(pr, ka, tb, &H, &dHa, &dHx);
      const double x = (double)ka * pr.dvoigt;
      s0 += uw * H;
      s1 += uw * pr.kappa0 * dHa;
      s2 += uw * pr.kappa0 * dHx * (-x * pr.dvoigt * resolu);
    }
    const long pair = (long)list_l[j] * D + d;
    s_k0[pair] += s0;
    s_a[pair] += s1;
    s_dv[pair] += s2;
  }
}

// Tile moments of the adjoint weights for the far-wing expansion: for each
// (512-pixel tile, depth), m_n = sum over the tile's pixels of u(w) t^n with
// t = (w - wm) / h (wm the tile's centre, h half its width), n < kFarTerms.
// One wave per tile, 8 consecutive pixels per lane as k_gather.
__global__ __launch_bounds__(64) void k_adj_moments(long W, const double* __restrict__ u, double* __restrict__ mom) {
  const long t = blockIdx.x, ntiles = gridDim.x;
  const int d = blockIdx.y, lane = threadIdx.x;
  const long t0 = t * kWT;
  const double wm = (double)t0 + 0.5 * (kWT - 1);
  constexpr double kHalf = 0.5 * kWT;
  double m[kFarTerms];
#pragma unroll
  for (int n = 0; n < kFarTerms; ++n) m[n] = 0.0;
#pragma unroll
  for (int p = 0; p < kWP; ++p) {
    const long w = t0 + (long)kWP * lane + p;
    double pw = w < W ? u[(long)d * W + w] : 0.0;
    const double tt = ((double)w - wm) / kHalf;
#pragma unroll
    for (int n = 0; n < kFarTerms; ++n) {
      m[n] += pw;
      pw *= tt;
    }
  }
  double mine = 0.0;
#pragma unroll
  for (int n = 0; n < kFarTerms; ++n) {
    const double sn = wave_sum_butterfly(m[n]);
    mine = lane == n ? sn : mine;
  }
  if (lane < kFarTerms) mom[((long)d * ntiles + t) * kFarTerms + lane] = mine;
}

// Far wings of the adjoint: per list entry, Q = sum over its far pixels
// (N10DOP < |k| <= MAXSTEP) of u / k^2, the transpose of k_gather's far-wing
// sum, with the same split of the entry's tiles:
//  * tiles wholly in the far zone at |K| >= 4.5 h (K the tile centre's
//    offset from the line centre) take the expansion
//    u / (K + x)^2 summed = (1/K^2) sum_n (n+1) (-h/K)^n m_n over the tile's
//    moments (k_adj_moments), one entry per lane;
//  * the others (the line-centre tiles, the N10DOP and MAXSTEP edges) are
//    summed pixel by pixel, one entry at a time over the whole wave (8
//    consecutive pixels per lane, coalesced reads of u), then one wave sum.
// Each entry's Q is written once, in a fixed order: no atomics, run-to-run
// identical.  One thread per entry of a depth's lists, 64 entries per wave.
__global__ __launch_bounds__(64) void k_far_adj(long W, int D, const int64_t* __restrict__ regions,
                                                const WingHead* __restrict__ list_cm, const double* __restrict__ u,
                                                const double* __restrict__ mom, long ntiles,
                                                double* __restrict__ q_entry) {
  const int d = blockIdx.y, lane = threadIdx.x;
  const long beg = regions[(long)d * (kClasses + 1)], end = regions[(long)d * (kClasses + 1) + kClasses];
  constexpr double kHalf = 0.5 * kWT;
  const double kmin = kFarExpandRatio * kHalf;
  const long Wl = W;
  for (long j0 = beg + (long)blockIdx.x * 64; j0 < end; j0 += (long)gridDim.x * 64) {
    const long j = j0 + lane;
    WingHead h;
    h.lo = INT32_MAX;
    h.hi = INT32_MIN;
    h.nlo = 0;
    h.nhi = 0;
    h.xs = 0.0;
    h.cd = 0.0;
    if (j < end) h = list_cm[j];
    const bool far = h.lo <= h.hi;
    // expansion part, one entry per lane
    double qe = 0.0;
    long t_first = 0, t_last = -1;
    if (far) {
      const long lo_px = h.lo > 0 ? h.lo : 0, hi_px = h.hi < Wl - 1 ? h.hi : Wl - 1;
      if (lo_px <= hi_px) {
        t_first = lo_px / kWT;
        t_last = hi_px / kWT;
      }
    }
    const double c = h.cd;
    for (long t = t_first; t <= t_last; ++t) {
      const int t0i = (int)(t * kWT), t1i = t0i + kWT - 1;
      const double K = (double)t0i + 0.5 * (kWT - 1) - c;
      if (!(far_interior(h, t0i, t1i) && fabs(K) >= kmin)) continue;
      double y = __builtin_amdgcn_rcp(K);
      y = __builtin_fma(y, __builtin_fma(-K, y, 1.0), y);
      y = __builtin_fma(y, __builtin_fma(-K, y, 1.0), y);
      const double r = -kHalf * y;
      const double* m = mom + ((long)d * ntiles + t) * kFarTerms;
      double acc = (double)kFarTerms * m[kFarTerms - 1];
#pragma unroll
      for (int n = kFarTerms - 2; n >= 0; --n) acc = __builtin_fma(acc, r, (double)(n + 1) * m[n]);
      qe += acc * y * y;
    }
    // exact part, the wave on one entry at a time
    unsigned long long todo = __ballot(far && t_first <= t_last);
    double qx = 0.0;
    while (todo) {
      const int src = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int lo = __builtin_amdgcn_readlane(h.lo, src), hi = __builtin_amdgcn_readlane(h.hi, src);
      const int nlo = __builtin_amdgcn_readlane(h.nlo, src), nhi = __builtin_amdgcn_readlane(h.nhi, src);
      const int tf = __builtin_amdgcn_readlane((int)t_first, src), tl = __builtin_amdgcn_readlane((int)t_last, src);
      const int M = (hi - lo) >> 1, cc = lo + M, n10 = (nhi - nlo) >> 1;
      // the candidate tiles: the ends, the tiles within kmin of the centre,
      // and those holding the Voigt zone's edges; the others expand or hold
      // no far pixel
      const int tA0 = (int)floor(((double)cc - kmin) / kWT) - 1, tB0 = (int)floor(((double)cc + kmin) / kWT) + 1;
      const int tA = tA0 > tf ? tA0 : tf, tB = tB0 < tl ? tB0 : tl;
      const int zl = (int)floor((double)(cc - n10 - 1) / kWT), zr = (int)floor((double)(cc + n10 + 1) / kWT);
      // the candidate tiles in increasing order (tf <= zl' <= tA..tB <= zr' <=
      // tl, zl' = min(zl, tA), zr' = max(zr, tB)), each visited once
      const int ia[5] = {tf, zl < tA ? zl : tA, tA, zr > tB ? zr : tB, tl};
      const int ib[5] = {tf, zl < tA ? zl : tA, tB, zr > tB ? zr : tB, tl};
      double part = 0.0;
      int last = tf - 1;
      WingHead hh;
      hh.lo = lo;
      hh.hi = hi;
      hh.nlo = nlo;
      hh.nhi = nhi;
      const double n10d = (double)n10, md = (double)M;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int ta = ia[k] > last + 1 ? ia[k] : last + 1;
        const int tb = ib[k] < tl ? ib[k] : tl;
        for (int t = ta; t <= tb; ++t) {
          const int t0i = t * kWT, t1i = t0i + kWT - 1;
          const double K = (double)t0i + 0.5 * (kWT - 1) - (double)cc;
          if (far_interior(hh, t0i, t1i) && fabs(K) >= kmin) continue;  // expanded above
          const long w0 = (long)t0i + (long)kWP * lane;
          // u is padded by a tile past the last depth's row: the 8 loads are
          // in bounds and pixels past the grid carry weight 0 below
          const double* up = u + (long)d * W + w0;
          double uv[kWP];
#pragma unroll
          for (int p = 0; p < kWP; ++p) uv[p] = up[p];
          double inv[kWP];  // 1 / k^2 on the far pixels, 0 elsewhere (one shared reciprocal)
          far_chunk<kWP, kFarNewton, true>(1.0, (double)(w0 - cc), 0, n10d, md, inv);
#pragma unroll
          for (int p = 0; p < kWP; ++p) part = __builtin_fma(w0 + p < Wl ? uv[p] : 0.0, inv[p], part);
        }
        last = tb > last ? tb : last;
      }
      const double tot = wave_sum_butterfly(part);
      qx = lane == src ? tot : qx;
    }
    if (j < end) q_entry[j] = qe + qx;
  }
}

// Far-wing sums to the entry's pair: X = kappa0 H(x10, a) N10DOP^2, x10 =
// N10DOP dvoigt (line_opacity.py:562-617).
__global__ void k_far_finish(int D, const int64_t* __restrict__ regions, const WingParams* __restrict__ list_p,
                             const int32_t* __restrict__ list_l, const double* __restrict__ q_entry,
                             DeviceTables tb, const double* __restrict__ wl, long W, double* __restrict__ s_k0,
                             double* __restrict__ s_a, double* __restrict__ s_dv) {
  const int d = blockIdx.y;
  const long beg = regions[d * (kClasses + 1)], end = regions[d * (kClasses + 1) + kClasses];
  const double resolu = W > 1 && wl[1] / wl[0] > 1.0 ? 1.0 / (wl[1] / wl[0] - 1.0) : 300000.0;
  for (long j = beg + blockIdx.x * (long)blockDim.x + threadIdx.x; j < end; j += (long)gridDim.x * blockDim.x) {
    const double q = q_entry[j];
    if (q == 0.0) continue;
    const WingParams pr = list_p[j];
    if (!pr.use_far) continue;
    const long n10 = pr.n10dop;
    double H, dHa, dHx;
    near_wing_partials(pr, n10, tb, &H, &dHa, &dHx);
    const double n2 = (double)n10 * (double)n10;
    const double x10 = (double)n10 * pr.dvoigt;
    const long pair = (long)list_l[j] * D + d;
    atomicAdd(s_k0 + pair, q * H * n2);
    atomicAdd(s_a + pair, q * pr.kappa0 * n2 * dHa);
    atomicAdd(s_dv + pair, q * pr.kappa0 * n2 * dHx * (-x10 * pr.dvoigt * resolu));
  }
}

// RT vjp per wavelength (total-flux solve, physics/josh_solver.py:367-767
// through engine/radiative.py:42-218 masks): v[w] = dL/dF_nu(w); outputs
// dL/d(cont_abs, cont_scat, line_abs) [D][W], dL/dB (the Planck function at
// T(d) used as the continuum source) and dL/d(line_source), 0 on masked depths.
// O(depths) per column, in the stages below.  In k_vjp_solve the 51x51 loops
// of the scattering fixed point and its adjoint read COEFJ and the vectors
// XS, XA, lam for every product: COEFJ is staged in LDS once per workgroup and
// the three vectors live in LDS interleaved by lane (88 KB per 64-lane
// workgroup), the rest in private memory.
constexpr int kVjpBlock = 64;
// From this many wavelengths in a call, the colour stages run one lane per
// wavelength (1,024 waves and more fill the GPU without a lane per colour).
constexpr long kVjpSeqMin = 65536;

// The depths of wavelength w that enter the solve, in increasing rho (the
// masks of engine/radiative.py:42-218): writes idx[0..n), returns n.
__device__ int vjp_depths(int D, long W, long w, const double* __restrict__ rhox, const double* __restrict__ ca,
                          const double* __restrict__ cs, const double* __restrict__ la,
                          const double* __restrict__ ls, int* idx) {
  const double maxr4 = 3.4e38;
  int n = 0;
  for (int d = 0; d < D && n < kAdjMaxDepth; ++d) {
    const long i = (long)d * W + w;
    const double r = rhox[d], a1 = ca[i], a2 = cs[i], a3 = la[i], a4 = ls[i];
    const bool okd = r >= 0.0 && isfinite(r) && isfinite(a1) && isfinite(a2) && isfinite(a3) && isfinite(a4) &&
                     a1 < maxr4 && a2 < maxr4 && a3 < maxr4 && a4 < maxr4;
    if (okd && r > 0) idx[n++] = d;
  }
  if (n >= 4 && !(rhox[idx[0]] < rhox[idx[n - 1]])) {
    for (int j = 0; j < n / 2; ++j) {
      int t = idx[j];
      idx[j] = idx[n - 1 - j];
      idx[n - 1 - j] = t;
    }
  }
  return n;
}

// Processing order of the vjp: wavelengths whose solve takes the XTAU branch
// (or none) first, label-401 ones (TAUNU(0) = abtot rho at the top above
// XTAU(51)) last, so that a wave runs one branch.  Only the order depends on
// the atomics; each wavelength's result does not.
// One spectrum of an RT vjp (device pointers): D depths, W wavelengths, its
// first slot g0 in the batch's global wavelength index, inputs and outputs.
struct VjpArgs {
  int D;
  long W, g0;
  const double *wl, *temp, *rhox, *ca, *cs, *la, *ls, *src, *v;
  double *g_ca, *g_cs, *g_la, *g_b, *g_src;
};

// The spectrum holding global slot g: the last i with A[i].g0 <= g.
__device__ __forceinline__ int vjp_inst(const VjpArgs* __restrict__ A, int n, long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A[mid].g0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// XTAU-branch wavelengths first, label-401 ones last (counts[0], counts[1]).
__device__ __forceinline__ void vjp_order_at(const VjpArgs& a, long w, const double* xtau, int* __restrict__ counts,
                                             int32_t* __restrict__ order) {
  const int D = a.D;
  const long W = a.W;
  bool b401 = false;
  if (a.v[w] != 0.0) {
    int idx[kAdjMaxDepth];
    const int n = vjp_depths(D, W, w, a.rhox, a.ca, a.cs, a.la, a.ls, idx);
    if (n >= 4) {
      const long i = (long)idx[0] * W + w;
      double ab = a.ca[i] + a.la[i] + a.cs[i] + a.ls[i];
      ab = ab > kEps38 ? ab : kEps38;
      b401 = ab * a.rhox[idx[0]] > xtau[adj::kXt - 1];
    }
  }
  if (b401)
    order[W - 1 - atomicAdd(&counts[1], 1)] = (int32_t)w;
  else
    order[atomicAdd(&counts[0], 1)] = (int32_t)w;
}

__global__ void k_vjp_order(VjpArgs a, const double* xtau, int* __restrict__ counts, int32_t* __restrict__ order) {
  const long w = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (w < a.W) vjp_order_at(a, w, xtau, counts, order);
}

// Every spectrum of a batch (counts [2n], order [G] in each spectrum's slots).
__global__ void k_vjp_order_batch(const VjpArgs* __restrict__ A, int n, long G, const double* xtau,
                                  int* __restrict__ counts, int32_t* __restrict__ order) {
  const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int i = vjp_inst(A, n, g);
  const VjpArgs a = A[i];
  vjp_order_at(a, g - a.g0, xtau, counts + 2 * i, order + a.g0);
}

// The vjp of one wavelength in stages, each with the parallelism it has (the
// per-column solve is one long dependent chain; the colour passes of
// rt_adjoint.h are independent and write disjoint entries):
//   k_vjp_solve   one lane per wavelength: the column, the FP64 solve and, in
//                 the XTAU branch, the fixed point's adjoint;
//   k_vjp_pass    one lane per (wavelength, colour of 5): the XTAU MAP1
//                 passes (colours 0-3) or the label-401 band passes;
//   k_vjp_adj401  one lane per label-401 wavelength: its adjoint iteration;
//   k_vjp_tau401  one lane per (label-401 wavelength, colour of 5): TAUNU;
//   k_vjp_finish  one lane per (wavelength, colour of 4): INTEG/PARCOE and the
//                 chain to the outputs for depths j = colour (mod 4).
// A wavelength's state lives in its record between stages; each entry is
// owned by one lane of a stage, which reads, updates and writes it back, so
// every sum runs in the order adj::total_flux_vjp runs it.
// The per-wavelength records the vjp stages pass on (about 13 KB each), stored
// field-major: of field f, record t's element j is at f's region + j G + t
// (G records), so the lanes of a wave, which hold consecutive records (or
// share one), read and write each element with one coalesced access.  Held
// whole, one record per lane, every access was a transaction of its own and
// the stages waited on memory 56-87% of their cycles.
enum VjpField {
  kFVw, kFRho, kFAbtot, kFAlpha, kFSnubar, kFTaunu, kFS, kFLam, kFGxb, kFGxa, kFBand, kFGTau, kFGSnubar,
  kFGAlpha, kFN, kFFlag, kFIdx, kFBase, kFCount
};

// elements per record and element size of field f
__host__ __device__ constexpr size_t vjp_field_elems(int f) {
  return f == kFVw || f == kFN || f == kFFlag ? 1
         : f == kFGxb || f == kFGxa || f == kFBase ? (size_t)adj::kXt
         : f == kFBand ? (size_t)5 * kAdjMaxDepth
                       : (size_t)kAdjMaxDepth;
}
__host__ __device__ constexpr size_t vjp_field_size(int f) { return f >= kFN ? sizeof(int) : sizeof(double); }

// byte offset of field f's region for G records (regions 256-byte aligned)
__host__ __device__ inline size_t vjp_field_offset(int f, long G) {
  size_t off = 0;
  for (int i = 0; i < f; ++i) off += (vjp_field_elems(i) * vjp_field_size(i) * (size_t)G + 255) & ~(size_t)255;
  return off;
}

struct VjpStore {
  char* base;  // 256-byte aligned
  long G;      // records
};

static size_t vjp_store_bytes(long G) { return vjp_field_offset(kFCount, G); }

template <typename T>
struct FieldRef {
  T* r;
  long g, t;
  __device__ __forceinline__ T& operator[](int j) const { return r[(long)j * g + t]; }
};

// band[j][o], o < 5
struct BandRef {
  double* r;
  long g, t;
  struct Row {
    double* r;
    long g, t;
    __device__ __forceinline__ double& operator[](int o) const { return r[(long)o * g + t]; }
  };
  __device__ __forceinline__ Row operator[](int j) const { return Row{r + (long)j * 5 * g, g, t}; }
};

// Record t of a store, with the fields of the former per-wavelength struct.
struct VjpRecord {
  int& n;     // depths in the solve
  int& flag;  // 0 XTAU branch, 1 label 401, -1 no work
  double& vw;
  FieldRef<int> idx, base;
  FieldRef<double> rho, abtot, alpha, snubar, taunu, S, lam, gxb, gxa, g_tau, g_snubar, g_alpha;
  BandRef band;
  template <typename T>
  __device__ __forceinline__ static T* field(const VjpStore& s, int f) {
    return (T*)(s.base + vjp_field_offset(f, s.G));
  }
  __device__ __forceinline__ VjpRecord(const VjpStore& s, long t)
      : n(field<int>(s, kFN)[t]),
        flag(field<int>(s, kFFlag)[t]),
        vw(field<double>(s, kFVw)[t]),
        idx{field<int>(s, kFIdx), s.G, t},
        base{field<int>(s, kFBase), s.G, t},
        rho{field<double>(s, kFRho), s.G, t},
        abtot{field<double>(s, kFAbtot), s.G, t},
        alpha{field<double>(s, kFAlpha), s.G, t},
        snubar{field<double>(s, kFSnubar), s.G, t},
        taunu{field<double>(s, kFTaunu), s.G, t},
        S{field<double>(s, kFS), s.G, t},
        lam{field<double>(s, kFLam), s.G, t},
        gxb{field<double>(s, kFGxb), s.G, t},
        gxa{field<double>(s, kFGxa), s.G, t},
        g_tau{field<double>(s, kFGTau), s.G, t},
        g_snubar{field<double>(s, kFGSnubar), s.G, t},
        g_alpha{field<double>(s, kFGAlpha), s.G, t},
        band{field<double>(s, kFBand), s.G, t} {}
};

using VjpColumn = adj::Column<kAdjMaxDepth>;
using VjpWork = adj::Work<kAdjMaxDepth>;

__device__ __forceinline__ void vjp_load(const VjpRecord& R, VjpColumn& col, VjpWork& wk) {
  col.n = R.n;
  for (int j = 0; j < R.n; ++j) {
    col.rho[j] = R.rho[j];
    col.abtot[j] = R.abtot[j];
    col.alpha[j] = R.alpha[j];
    col.snubar[j] = R.snubar[j];
    wk.taunu[j] = R.taunu[j];
  }
}

// The column of one wavelength, its FP64 solve and (XTAU branch) the fixed
// point's adjoint, into record R; t indexes the spectrum's order.
__device__ __forceinline__ void vjp_solve_at(const VjpArgs& a, long t, const int32_t* __restrict__ order,
                                             const DeviceTables& tb, const float* s_cj, const VjpRecord& R) {
  const int D = a.D;
  const long W = a.W;
  const long w = order[t];
  const double vw = a.v[w];
  int idx[kAdjMaxDepth];
  const int n = vw != 0.0 ? vjp_depths(D, W, w, a.rhox, a.ca, a.cs, a.la, a.ls, idx) : 0;
  R.flag = -1;
  R.n = n;
  R.vw = vw;
  if (vw == 0.0 || n < 4) return;
  const double wv = a.wl[w];
  const double freq = kCLightNm / (wv > 1e-12 ? wv : 1e-12);
  const double f3 = pow(freq / 1.0e15, 3.0);
  VjpColumn col;
  VjpWork wk;
  col.n = n;
  for (int j = 0; j < n; ++j) {
    const int d = idx[j];
    const long i = (long)d * W + w;
    const double B = planck_one(freq, f3, a.temp[d]);
    double S = B;
    if (a.src) {
      const double s0 = a.src[i];
      S = isfinite(s0) ? s0 : B;
    }
    const double ac = a.ca[i], al = a.la[i], sc = a.cs[i], sgl = a.ls[i];
    double ab = ac + al + sc + sgl;
    ab = ab > kEps38 ? ab : kEps38;
    double alpha = ab > 0.0 ? (sc + sgl) / ab : 0.0;
    alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
    const double dn = ac + al;
    col.rho[j] = a.rhox[d];
    col.abtot[j] = ab;
    col.alpha[j] = alpha;
    col.snubar[j] = dn > 0 ? (ac * B + al * S) / dn : B;
    R.idx[j] = d;
    R.rho[j] = col.rho[j];
    R.abtot[j] = ab;
    R.alpha[j] = alpha;
    R.snubar[j] = col.snubar[j];
  }
  const adj::JoshTab jt{tb.xtau, tb.ch, s_cj};
  // the fixed point's vectors in the lane's own memory: holding them in LDS
  // (78 KiB per 64 lanes) left one wave per CU (176.7 against 197.7 ms for
  // the full-window backward)
  wk.bind_private();
  int took401 = 0;
  adj::total_flux64(col, jt, wk, 1e-13, &took401);
  for (int j = 0; j < n; ++j) {
    R.taunu[j] = wk.taunu[j];
    R.g_tau[j] = R.g_snubar[j] = R.g_alpha[j] = 0.0;
  }
  if (took401) {
    for (int j = 0; j < n; ++j) {
      R.S[j] = wk.pa[j];
      for (int o = 0; o < 5; ++o) R.band[j][o] = 0.0;
    }
  } else {
    double gs0, ga0;
    adj::xtau_adjoint(col, jt, wk, &gs0, &ga0);
    R.g_snubar[0] = gs0;
    R.g_alpha[0] = ga0;
    for (int k = 0; k < adj::kXt; ++k) {
      R.base[k] = wk.x.base[k];
      R.gxb[k] = wk.x.gxb[k];
      R.gxa[k] = wk.x.gxa[k];
    }
  }
  R.flag = took401;
}

__device__ __forceinline__ void vjp_stage_coefj(const DeviceTables& tb, float* s_cj) {
  for (int i = threadIdx.x; i < adj::kXt * adj::kXt; i += kVjpBlock) s_cj[i] = tb.coefj32[i];
  __syncthreads();  // the only barrier: lanes leave independently afterwards
}

__global__ __launch_bounds__(kVjpBlock) void k_vjp_solve(VjpArgs a, const int32_t* __restrict__ order,
                                                         DeviceTables tb, VjpStore recs) {
  __shared__ float s_cj[adj::kXt * adj::kXt];
  vjp_stage_coefj(tb, s_cj);
  const long t = blockIdx.x * (long)kVjpBlock + threadIdx.x;
  if (t < a.W) vjp_solve_at(a, t, order, tb, s_cj, VjpRecord(recs, t));
}

__global__ __launch_bounds__(kVjpBlock) void k_vjp_solve_batch(const VjpArgs* __restrict__ A, int n, long G,
                                                               const int32_t* __restrict__ order, DeviceTables tb,
                                                               VjpStore recs) {
  __shared__ float s_cj[adj::kXt * adj::kXt];
  vjp_stage_coefj(tb, s_cj);
  const long g = blockIdx.x * (long)kVjpBlock + threadIdx.x;
  if (g >= G) return;
  const VjpArgs a = A[vjp_inst(A, n, g)];
  vjp_solve_at(a, g - a.g0, order + a.g0, tb, s_cj, VjpRecord(recs, g));
}

// The colour passes: one lane per (wavelength, colour), or with seq one lane
// per wavelength running its colours in turn on the column it loaded once
// (the CPU order; each pass writes only its own entries, so the two forms
// give the same results).  seq suits calls whose wavelengths alone fill the
// GPU: the column is then read once instead of once per colour.
__global__ __launch_bounds__(64) void k_vjp_pass(long W, DeviceTables tb, VjpStore recs, int seq) {
  const long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= W * (seq ? 1 : 5)) return;
  const VjpRecord R(recs, seq ? u : u / 5);
  const int r = seq ? -1 : (int)(u % 5);  // -1: every colour
  const int flag = R.flag;
  if (flag < 0 || (flag == 0 && r == 4)) return;
  VjpColumn col;
  VjpWork wk;
  vjp_load(R, col, wk);
  const int n = col.n;
  if (flag == 0) {
    for (int k = 0; k < adj::kXt; ++k) {
      wk.x.base[k] = R.base[k];
      wk.x.gxb[k] = R.gxb[k];
      wk.x.gxa[k] = R.gxa[k];
    }
    const int j0 = r < 0 ? 0 : r, js = r < 0 ? 1 : 4;
    double gs[kAdjMaxDepth], ga[kAdjMaxDepth], gt[kAdjMaxDepth];
    for (int j = j0; j < n; j += js) {
      gs[j] = R.g_snubar[j];
      ga[j] = R.g_alpha[j];
      gt[j] = R.g_tau[j];
    }
    const adj::JoshTab jt{tb.xtau, tb.ch, tb.coefj32};
    if (r < 0) {
      for (int c = 0; c < 4; ++c) adj::xtau_pass(col, jt, wk, c, gs, ga, gt);
    } else {
      adj::xtau_pass(col, jt, wk, r, gs, ga, gt);
    }
    for (int j = j0; j < n; j += js) {
      R.g_snubar[j] = gs[j];
      R.g_alpha[j] = ga[j];
      R.g_tau[j] = gt[j];
    }
  } else {
    for (int j = 0; j < n; ++j) wk.pa[j] = R.S[j];
    if (r < 0) {
      for (int c = 0; c < 5; ++c) adj::l401_band(col, wk, c, wk.s.band);
    } else {
      adj::l401_band(col, wk, r, wk.s.band);
    }
    for (int j = 0; j < n; ++j) {
      for (int c = 0; c < 5; ++c) {
        if (r >= 0 && c != r) continue;
        const int i = adj::seeded_in(j - 2, c, 5);
        if (i >= 0 && i < n) R.band[j][i - j + 2] = wk.s.band[j][i - j + 2];
      }
    }
  }
}

__global__ __launch_bounds__(64) void k_vjp_adj401(long W, VjpStore recs) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= W) return;
  const VjpRecord R(recs, t);
  if (R.flag != 1) return;
  VjpColumn col;
  VjpWork wk;
  vjp_load(R, col, wk);
  const int n = col.n;
  double gs[kAdjMaxDepth], ga[kAdjMaxDepth], gt[kAdjMaxDepth];
  for (int j = 0; j < n; ++j) {
    wk.pa[j] = R.S[j];
    gs[j] = R.g_snubar[j];
    ga[j] = R.g_alpha[j];
    gt[j] = R.g_tau[j];
  }
  for (int j = 0; j < n; ++j)  // the band, into the lane's own memory
    for (int o = 0; o < 5; ++o) wk.s.band[j][o] = R.band[j][o];
  adj::l401_adjoint(col, wk, wk.s.band, gs, ga, gt);
  for (int j = 0; j < n; ++j) {
    R.g_snubar[j] = gs[j];
    R.g_alpha[j] = ga[j];
    R.g_tau[j] = gt[j];
    R.lam[j] = wk.s.lam[j];
  }
}

__global__ __launch_bounds__(64) void k_vjp_tau401(long W, VjpStore recs, int seq) {
  const long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= W * (seq ? 1 : 5)) return;
  const VjpRecord R(recs, seq ? u : u / 5);
  const int r = seq ? -1 : (int)(u % 5);  // -1: every colour in turn
  if (R.flag != 1) return;
  VjpColumn col;
  VjpWork wk;
  vjp_load(R, col, wk);
  const int n = col.n;
  double gt[kAdjMaxDepth];
  for (int j = 0; j < n; ++j) {
    wk.pa[j] = R.S[j];
    wk.s.lam[j] = R.lam[j];
  }
  const int i0 = r < 0 ? 0 : r, is = r < 0 ? 1 : 5;
  for (int i = i0; i < n; i += is) gt[i] = R.g_tau[i];
  if (r < 0) {
    for (int c = 0; c < 5; ++c) adj::l401_tau(col, wk, c, gt);
  } else {
    adj::l401_tau(col, wk, r, gt);
  }
  for (int i = i0; i < n; i += is) R.g_tau[i] = gt[i];
}

// INTEG/PARCOE and the chain to the outputs of one wavelength for the depths
// j = r (mod 4), or for every depth with the colours in turn when r < 0; t
// indexes the spectrum's order.
__device__ __forceinline__ void vjp_finish_at(const VjpArgs& a, long t, int r, const int32_t* __restrict__ order,
                                              const VjpRecord& R) {
  const long W = a.W;
  const double *wl = a.wl, *temp = a.temp, *ca = a.ca, *la = a.la, *src = a.src;
  double *g_ca = a.g_ca, *g_cs = a.g_cs, *g_la = a.g_la, *g_b = a.g_b, *g_src = a.g_src;
  if (R.flag < 0) return;
  VjpColumn col;
  VjpWork wk;
  vjp_load(R, col, wk);
  const int n = col.n;
  for (int j = 0; j < n; ++j) wk.g_tau[j] = R.g_tau[j];
  adj::tau_suffix(n, wk);
  double gab[kAdjMaxDepth];
  const int j0 = r < 0 ? 0 : r, js = r < 0 ? 1 : 4;
  for (int j = j0; j < n; j += js) gab[j] = 0.0;
  if (r < 0) {
    for (int c = 0; c < 4; ++c) adj::integ_pass(col, wk, c, gab);
  } else {
    adj::integ_pass(col, wk, r, gab);
  }
  const long w = order[t];
  const double vw = R.vw;
  const double wv = wl[w];
  const double freq = kCLightNm / (wv > 1e-12 ? wv : 1e-12);
  const double f3 = pow(freq / 1.0e15, 3.0);
  for (int j = j0; j < n; j += js) {
    const int d = R.idx[j];
    const long i = (long)d * W + w;
    const double g_ab = gab[j] * vw, g_al = R.g_alpha[j] * vw, g_sb = R.g_snubar[j] * vw;
    const double ac = ca[i], al = la[i];
    const double ab = col.abtot[j], alpha = col.alpha[j], sb = col.snubar[j];
    const double dn = ac + al;
    const double galpha = (alpha > 0.0 && alpha < 1.0) ? g_al : 0.0;
    const double da_dab = -alpha / ab;  // d alpha / d(ac or al)
    double gsb_ac = 0.0, gsb_al = 0.0, gB = g_sb, gS = 0.0;
    if (dn > 0) {
      // the continuum source B(T(d)) and the line source S at depth d
      const double B = planck_one(freq, f3, temp[d]);
      double S = B;
      if (src) {
        const double s0 = src[i];
        S = isfinite(s0) ? s0 : B;
      }
      gsb_ac = g_sb * (B - sb) / dn;
      gsb_al = g_sb * (S - sb) / dn;
      gB = g_sb * ac / dn;
      gS = g_sb * al / dn;
    }
    if (g_ca) g_ca[i] = g_ab + galpha * da_dab + gsb_ac;
    if (g_la) g_la[i] = g_ab + galpha * da_dab + gsb_al;
    if (g_cs) g_cs[i] = g_ab + galpha * (1.0 - alpha) / ab;
    if (src) {
      if (g_b) g_b[i] = gB;
      if (g_src) g_src[i] = gS;
    } else if (g_b) {
      g_b[i] = gB + gS;
    }
  }
}

__global__ __launch_bounds__(64) void k_vjp_finish(VjpArgs a, const int32_t* __restrict__ order, VjpStore recs,
                                                   int seq) {
  const long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int per = seq ? 1 : 4;
  if (u >= a.W * per) return;
  vjp_finish_at(a, u / per, seq ? -1 : (int)(u % 4), order, VjpRecord(recs, u / per));
}

__global__ __launch_bounds__(64) void k_vjp_finish_batch(const VjpArgs* __restrict__ A, int n, long G,
                                                         const int32_t* __restrict__ order, VjpStore recs,
                                                         int seq) {
  const long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int per = seq ? 1 : 4;
  if (u >= G * per) return;
  const long g = u / per;
  const VjpArgs a = A[vjp_inst(A, n, g)];
  vjp_finish_at(a, g - a.g0, seq ? -1 : (int)(u % 4), order + a.g0, VjpRecord(recs, g));
}

// ------------------------------------------------------------ ASYNTH pieces
// Workspace of one ASYNTH call (offsets from a 256-byte aligned base).
struct AsynthLayout {
  size_t o_wp, o_cls, o_cnt, o_tot, o_reg, o_lcm, o_lp, o_ev, o_nr, o_ll, total;
};

static AsynthLayout asynth_layout(long L, int D, bool line_index = false) {
  AsynthLayout y{};
  const long pairs = L * (long)D;
  const long nblk = (L + kBlock - 1) / kBlock;
  size_t off = 0;
  y.o_wp = off; off = align_up(off + (size_t)pairs * sizeof(WingParams));
  y.o_cls = off; off = align_up(off + (size_t)pairs);
  y.o_cnt = off; off = align_up(off + (size_t)(nblk > 0 ? nblk : 1) * D * kClasses * sizeof(int32_t));
  y.o_tot = off; off = align_up(off + (size_t)D * kClasses * sizeof(int64_t));
  y.o_reg = off; off = align_up(off + (size_t)D * (kClasses + 1) * sizeof(int64_t));
  y.o_lcm = off; off = align_up(off + (size_t)pairs * sizeof(WingHead));
  y.o_lp = off; off = align_up(off + (size_t)pairs * sizeof(WingParams));
  y.o_ev = off; off = align_up(off + 5 * sizeof(unsigned long long));
  y.o_nr = off; off = align_up(off + (size_t)D * sizeof(int32_t));
  y.o_ll = line_index ? off : 0; off = line_index ? align_up(off + (size_t)pairs * sizeof(int32_t)) : off;
  y.total = off;
  return y;
}

struct AsynthBufs {
  WingParams* wp;
  uint8_t* cls;
  int32_t* cnt;
  int64_t* tot;
  int64_t* reg;
  WingHead* lcm;
  WingParams* lp;
  unsigned long long* ev;
  int32_t* near_reach;
  int32_t* list_l;  // line of each list entry, or NULL
};

static AsynthBufs asynth_bufs(char* base, const AsynthLayout& y) {
  return AsynthBufs{(WingParams*)(base + y.o_wp), (uint8_t*)(base + y.o_cls), (int32_t*)(base + y.o_cnt),
                    (int64_t*)(base + y.o_tot),   (int64_t*)(base + y.o_reg), (WingHead*)(base + y.o_lcm),
                    (WingParams*)(base + y.o_lp), (unsigned long long*)(base + y.o_ev),
                    (int32_t*)(base + y.o_nr), y.o_ll ? (int32_t*)(base + y.o_ll) : nullptr};
}

static GatherArgs gather_args(const synthe_lines* lines, const synthe_depths* depths, long n_wl, const double* wl,
                              const double* transp, const uint8_t* valid, const int64_t* center_order,
                              const int64_t* center_start, double* asynth, const AsynthBufs& b) {
  (void)lines;
  return GatherArgs{n_wl,   (int)depths->n_depths, center_order, center_start, transp, valid, b.lcm, b.lp,
                    b.reg, wl,                     depths->temperature, asynth};
}

static int check_asynth_args(const synthe_lines* lines, const synthe_depths* depths, long n_wl, const double* wl,
                             const double* transp, const uint8_t* valid, const double* kapmin,
                             const int64_t* center_order, const int64_t* center_start,
                             const int64_t* wing_order, const double* asynth, const char* who) {
  int rc = check_lines_depths(lines, depths, who);
  if (rc) return rc;
  if (n_wl <= 0 || !wl || !asynth || !center_start)
    return fail(SYNTHE_EINVAL, std::string(who) + ": wl/asynth/center_start required and n_wl > 0");
  if (n_wl > (1L << 30))
    return fail(SYNTHE_EINVAL, std::string(who) + ": n_wl above 2^30 pixels");
  if (lines->n_lines > 0 && (!transp || !valid || !kapmin || !center_order || !wing_order))
    return fail(SYNTHE_EINVAL, std::string(who) + ": NULL transp/valid/kapmin/order array");
  return SYNTHE_OK;
}

// Everything of ASYNTH before the gather: wing set-up, class lists, Voigt
// zones (into asynth).  Kernel launches only (no allocation, copy, memset or
// synchronisation), so it can be captured into a graph.
static int asynth_prepare(const synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths,
                          long n_wl, const double* wl, const double* transp, const uint8_t* valid,
                          const double* kapmin, double cutoff, const int64_t* wing_order, double* asynth,
                          bool stats, const AsynthBufs& b, hipStream_t st, bool near_values = true,
                          double* transp_out = nullptr, uint8_t* valid_out = nullptr) {
  const long L = lines->n_lines;
  const int D = (int)depths->n_depths;
  const long pairs = L * (long)D;
  const long nblk = (L + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, 5L, b.ev);
  HIP_TRY(hipGetLastError());
  if (pairs > 0) {
    // transp_out: TRANSP and valid are computed here in the same pass (mode 3,
    // no microturbulence) instead of being read from an earlier mode-1 pass
    if (transp_out)
      hipLaunchKernelGGL(k_transp_wing, dim3(blocks_for(pairs, kBlock)), dim3(kBlock), 0, st, to_dev(lines),
                         to_dev(depths), n_wl, kapmin, cutoff, 0.0, wl, ctx->tb, 3, transp_out, valid_out, nullptr,
                         nullptr, b.wp, b.cls);
    else
      hipLaunchKernelGGL(k_transp_wing, dim3(blocks_for(pairs, kBlock)), dim3(kBlock), 0, st, to_dev(lines),
                         to_dev(depths), n_wl, kapmin, cutoff, 0.0, wl, ctx->tb, 2, nullptr, nullptr, transp, valid,
                         b.wp, b.cls);
    HIP_TRY(hipGetLastError());
    const unsigned rows = (unsigned)((D + kListDepths - 1) / kListDepths);
    hipLaunchKernelGGL(k_class_count, dim3((unsigned)nblk, rows), dim3(kBlock), 0, st, L, D, wing_order,
                       b.cls, b.cnt);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_class_scan, dim3((unsigned)(D * kClasses)), dim3(kBlock), 0, st, nblk, D, b.cnt, b.tot);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_regions, dim3(blocks_for(D, 64)), dim3(64), 0, st, D, L, b.tot, b.reg);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_class_scatter, dim3((unsigned)nblk, rows), dim3(kBlock), 0, st, L, D, wing_order,
                       b.cls, b.cnt,
                       b.reg, b.wp, b.lcm, b.lp, b.list_l);
    HIP_TRY(hipGetLastError());
  } else {
    const long n = (long)D * (kClasses + 1);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, (unsigned long long*)b.reg);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(D, 64)), dim3(64), 0, st, D, b.near_reach);
  HIP_TRY(hipGetLastError());
  if (pairs > 0) {
    hipLaunchKernelGGL(k_near_reach, dim3(64, (unsigned)D), dim3(kBlock), 0, st, D, b.reg, b.lcm, b.near_reach);
    HIP_TRY(hipGetLastError());
  }
  if (!near_values) return SYNTHE_OK;
  const dim3 g_near((unsigned)blocks_for(n_wl, 256), (unsigned)D);
  if (stats)
    hipLaunchKernelGGL(k_near<true>, g_near, dim3(256), 0, st, n_wl, D, b.lcm, b.lp, b.reg, b.near_reach, ctx->tb,
                       asynth, b.ev + 3);
  else
    hipLaunchKernelGGL(k_near<false>, g_near, dim3(256), 0, st, n_wl, D, b.lcm, b.lp, b.reg, b.near_reach,
                       ctx->tb, asynth, nullptr);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

// ------------------------------------------------- batched prepare stages
// One instance's arguments for the stages asynth_prepare launches per
// spectrum (TRANSP with the wing set-up, class lists, Voigt zones).  The
// batch launches each stage once over all instances: block B of stage k
// belongs to instance i with off[k][i] <= B < off[k][i+1] and runs that
// instance's block B - off[k][i] of the per-spectrum grid, through the same
// device body, so each instance's results are those of its own launches.
struct PrepInst {
  LinesDev ln;
  DepthsDev dp;
  long W, L, nblk;
  int D;
  const double *kapmin, *wl;
  double cutoff;
  double* transp;
  uint8_t* valid;
  const int64_t* wing_order;
  double* asynth;
  AsynthBufs b;
};

constexpr int kPrepStages = 8;  // zero, TRANSP + set-up, count, scan, regions, scatter, reach, Voigt zones

// Blocks of stage k for one instance (per-spectrum grid, flattened x-major).
static long prep_blocks(int k, const PrepInst& P) {
  const long pairs = P.L * (long)P.D;
  const long rows = (P.D + kListDepths - 1) / kListDepths;
  switch (k) {
    case 0: return 1;
    case 1: return pairs > 0 ? (pairs + kBlock - 1) / kBlock : 0;
    case 2: return pairs > 0 ? P.nblk * rows : 0;
    case 3: return pairs > 0 ? (long)P.D * kClasses : 0;
    case 4: return pairs > 0 ? (P.D + kBlock - 1) / kBlock : 0;
    case 5: return pairs > 0 ? P.nblk * rows : 0;
    case 6: return pairs > 0 ? 64L * P.D : 0;
    default: return ((P.W + 255) / 256) * P.D;
  }
}

template <int kStage>
__global__ __launch_bounds__(kBlock) void k_prep_batch(const PrepInst* __restrict__ inst,
                                                       const int64_t* __restrict__ off, int n, DeviceTables tb) {
  int lo = 0, hi = n - 1;  // last instance with off[i] <= block
  const long B = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= B) lo = mid;
    else hi = mid - 1;
  }
  const PrepInst& P = inst[lo];
  const long local = B - off[lo];
  const AsynthBufs& b = P.b;
  if constexpr (kStage == 0) {  // the counters, Voigt reaches and (without pairs) the list regions
    for (int t = threadIdx.x; t < 5; t += kBlock) b.ev[t] = 0ull;
    for (int t = threadIdx.x; t < P.D; t += kBlock) b.near_reach[t] = 0;
    if (P.L * (long)P.D == 0)
      for (long t = threadIdx.x; t < (long)P.D * (kClasses + 1); t += kBlock) b.reg[t] = 0;
  } else if constexpr (kStage == 1) {
    k_transp_wing_at(local, 0, local + 1, P.ln, P.dp, P.W, P.kapmin, P.cutoff, 0.0, P.wl, tb, 3, P.transp, P.valid,
                     nullptr, nullptr, b.wp, b.cls);
  } else if constexpr (kStage == 2) {
    k_class_count_at(local % P.nblk, (int)(local / P.nblk), P.nblk, P.L, P.D, P.wing_order, b.cls, b.cnt);
  } else if constexpr (kStage == 3) {
    k_class_scan_at(local, 0, (long)P.D * kClasses, P.nblk, P.D, b.cnt, b.tot);
  } else if constexpr (kStage == 4) {
    k_regions_at(local, 0, local + 1, P.D, P.L, b.tot, b.reg);
  } else if constexpr (kStage == 5) {
    k_class_scatter_at(local % P.nblk, (int)(local / P.nblk), P.nblk, P.L, P.D, P.wing_order, b.cls, b.cnt, b.reg,
                       b.wp, b.lcm, b.lp, b.list_l);
  } else if constexpr (kStage == 6) {
    k_near_reach_at(local % 64, (int)(local / 64), 64, P.D, b.reg, b.lcm, b.near_reach);
  } else {
    const long gx = (P.W + 255) / 256;
    k_near_at<false>(local % gx, (int)(local / gx), gx, P.W, P.D, b.lcm, b.lp, b.reg, b.near_reach, tb, P.asynth,
                     nullptr);
  }
}

// ---------------------------------------------------------------- RT pieces
struct RtLayout {
  size_t o_flags, o_list, o_cnt, o_args, total;
};

// Workspace of the RT over G wavelengths of n_rt spectra.
static constexpr RtLayout rt_layout(long G, int n_rt = 1) {
  RtLayout y{};
  y.o_flags = align_up((size_t)2 * 3 * kNxtau * G * sizeof(float));
  y.o_list = y.o_flags + align_up((size_t)G);
  y.o_cnt = y.o_list + align_up((size_t)G * sizeof(int64_t));
  y.o_args = y.o_cnt + 256;
  y.total = y.o_args + align_up((size_t)n_rt * sizeof(RtArgs));
  return y;
}

// Both JOSH solves for every wavelength of n_rt spectra (rt: device array in
// the workspace, filled by the caller); kernel launches only (capturable).
static int rt_launch_set(const synthe_ctx* ctx, const RtArgs* rt, int n_rt, long G, char* base, const RtLayout& y,
                         hipStream_t st) {
  float* ws = (float*)base;
  uint8_t* flags = (uint8_t*)base + y.o_flags;
  int64_t* list = (int64_t*)(base + y.o_list);
  int* list_n = (int*)(base + y.o_cnt);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, list_n);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_rt_prep, dim3(blocks_for(G, kRtBlock)), dim3(kRtBlock), 0, st, rt, n_rt, G, ctx->tb.xtau, ws,
                     flags, list, list_n);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_rt_iter, dim3(blocks_for(G, kRtIterBlock)), dim3(kRtIterBlock), 0, st, rt, n_rt, G, flags, ws,
                     ctx->tb);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_lte, dim3(blocks_for(G, kRtLanes)), dim3(kRtLanes), 0, st, rt, n_rt, list, list_n, ctx->tb);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

// One spectrum: its RtArgs are stored into the workspace by a kernel.
static int rt_launch(const synthe_ctx* ctx, long n_depths, long n_wl, const double* wl, const double* temperature,
                     const double* rhox, const double* cont_abs, const double* cont_scat, const double* line_abs,
                     const double* line_scat, const double* line_source, double* flux_total, double* flux_cont,
                     int32_t* iters, char* base, const RtLayout& y, hipStream_t st) {
  RtArgs a{0, n_wl, (int)n_depths, 0, wl, temperature, rhox, cont_abs, cont_scat, line_abs, line_scat,
           line_source, flux_total, flux_cont, iters};
  RtArgs* d_args = (RtArgs*)(base + y.o_args);
  hipLaunchKernelGGL(k_store_rtargs, dim3(1), dim3(64), 0, st, a, d_args);
  HIP_TRY(hipGetLastError());
  return rt_launch_set(ctx, d_args, 1, n_wl, base, y, st);
}

// ------------------------------------------------- allocation-free RT scratch
// The RT workspace of synthe_lte_spectrum_noalloc for up to
// SYNTHE_NOALLOC_MAX_WL wavelengths (77 MiB), in a file-scope device array:
// every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory:
// kernels get the device address from hipGetSymbolAddress, taken once per
// device in synthe_create and kept in g_scratch_dev.
constexpr size_t kNoAllocScratchBytes = rt_layout(SYNTHE_NOALLOC_MAX_WL).total;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device).
static int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(SYNTHE_ERANGE, "device ordinal above the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SYNTHE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return SYNTHE_OK;
}

extern "C" {

const char* synthe_last_error(void) { return g_err.c_str(); }

int synthe_abi_version(void) { return SYNTHE_ABI_VERSION; }

int synthe_create(synthe_ctx** out) {
  if (!out) return fail(SYNTHE_EINVAL, "synthe_create: out is NULL");
  *out = nullptr;
  synthe_ctx* ctx = new synthe_ctx();
  if (hipGetDevice(&ctx->device) != hipSuccess) {
    delete ctx;
    return fail(SYNTHE_EHIP, "synthe_create: no HIP device");
  }
  // physics/tables.py:224-241 (Voigt) and 20-27 (FASTEX)
  ctx->h0.assign(kVoigtN, 0.0);
  ctx->h1.assign(kVoigtN, 0.0);
  ctx->h2.assign(kVoigtN, 0.0);
  for (int i = 0; i < kVoigtN; ++i) ctx->h0[i] = (double)i / 200.0;
  map1_table(kTabvi, kTabh1, 81, ctx->h0.data(), ctx->h1.data(), kVoigtN);
  for (int i = 0; i < kVoigtN; ++i) {
    double q = (double)i / 200.0;
    double vv = q * q;
    ctx->h0[i] = exp(-vv);
    ctx->h2[i] = ctx->h0[i] - (vv + vv) * ctx->h0[i];
  }
  ctx->extab.assign(kFastexN, 0.0);
  ctx->extabf.assign(kFastexN, 0.0);
  for (int i = 0; i < kFastexN; ++i) {
    ctx->extab[i] = exp(-(double)i);
    ctx->extabf[i] = exp(-(double)i * 0.001);
  }
  std::vector<double> host;
  host.insert(host.end(), ctx->h0.begin(), ctx->h0.end());
  host.insert(host.end(), ctx->h1.begin(), ctx->h1.end());
  host.insert(host.end(), ctx->h2.begin(), ctx->h2.end());
  host.insert(host.end(), ctx->extab.begin(), ctx->extab.end());
  host.insert(host.end(), ctx->extabf.begin(), ctx->extabf.end());
  host.insert(host.end(), kXtau, kXtau + kNxtau);
  host.insert(host.end(), kCh, kCh + kNxtau);
  std::vector<float> h32(kNxtau * kNxtau + kNxtau);
  for (int i = 0; i < kNxtau * kNxtau; ++i) h32[i] = (float)kCoefj[i];
  for (int k = 0; k < kNxtau; ++k) h32[kNxtau * kNxtau + k] = (float)kCoefj[k * kNxtau + k];
  if (hipMalloc(&ctx->d_tab, host.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&ctx->d_tab32, h32.size() * sizeof(float)) != hipSuccess) {
    synthe_destroy(ctx);
    return fail(SYNTHE_ENOMEM, "synthe_create: table allocation failed");
  }
  if (hipMemcpy(ctx->d_tab, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(ctx->d_tab32, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice) !=
          hipSuccess) {
    synthe_destroy(ctx);
    return fail(SYNTHE_EHIP, "synthe_create: table upload failed");
  }
  double* p = ctx->d_tab;
  ctx->tb.h0 = p;
  ctx->tb.h1 = p + kVoigtN;
  ctx->tb.h2 = p + 2 * kVoigtN;
  ctx->tb.extab = p + 3 * kVoigtN;
  ctx->tb.extabf = p + 3 * kVoigtN + kFastexN;
  ctx->tb.xtau = p + 3 * kVoigtN + 2 * kFastexN;
  ctx->tb.ch = ctx->tb.xtau + kNxtau;
  ctx->tb.coefj32 = ctx->d_tab32;
  ctx->tb.diag32 = ctx->d_tab32 + kNxtau * kNxtau;
  if (int rc = scratch_address(ctx->device, &ctx->scratch)) {
    synthe_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return SYNTHE_OK;
}

int synthe_destroy(synthe_ctx* ctx) {
  if (!ctx) return SYNTHE_OK;
  if (ctx->d_tab) (void)hipFree(ctx->d_tab);
  if (ctx->d_tab32) (void)hipFree(ctx->d_tab32);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->hstage_free) ctx->hstage_free(ctx->hstage);
  if (ctx->f19stage_free) ctx->f19stage_free(ctx->f19stage);
  for (int i = 0; i < synthe_ctx::kTimedMax; ++i) {
    if (ctx->ev_beg[i]) (void)hipEventDestroy(ctx->ev_beg[i]);
    if (ctx->ev_end[i]) (void)hipEventDestroy(ctx->ev_end[i]);
  }
  delete ctx;
  return SYNTHE_OK;
}

int synthe_copy_tables(const synthe_ctx* ctx, double* h0, double* h1, double* h2, double* extab,
                       double* extabf) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_copy_tables: ctx is NULL");
  if (h0) memcpy(h0, ctx->h0.data(), kVoigtN * sizeof(double));
  if (h1) memcpy(h1, ctx->h1.data(), kVoigtN * sizeof(double));
  if (h2) memcpy(h2, ctx->h2.data(), kVoigtN * sizeof(double));
  if (extab) memcpy(extab, ctx->extab.data(), kFastexN * sizeof(double));
  if (extabf) memcpy(extabf, ctx->extabf.data(), kFastexN * sizeof(double));
  return SYNTHE_OK;
}

int synthe_set_profile_precision(synthe_ctx* ctx, int precision) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_set_profile_precision: ctx is NULL");
  if (precision != SYNTHE_PREC_FP64 && precision != SYNTHE_PREC_FP32)
    return fail(SYNTHE_EINVAL, "synthe_set_profile_precision: unknown precision " + std::to_string(precision));
  ctx->precision = precision;
  return SYNTHE_OK;
}

int synthe_set_kernel_timing(synthe_ctx* ctx, int enable) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_set_kernel_timing: ctx is NULL");
  ctx->timing = enable != 0;
  ctx->n_timed = 0;
  return SYNTHE_OK;
}

int synthe_kernel_times(synthe_ctx* ctx, float* ms, int cap, int* count) {
  if (!ctx || !count) return fail(SYNTHE_EINVAL, "synthe_kernel_times: ctx or count is NULL");
  int n = ctx->n_timed;
  for (int i = 0; i < n && i < cap && ms; ++i) {
    HIP_TRY(hipEventSynchronize(ctx->ev_end[i]));
    HIP_TRY(hipEventElapsedTime(&ms[i], ctx->ev_beg[i], ctx->ev_end[i]));
  }
  *count = n;
  return SYNTHE_OK;
}

long synthe_wavelength_grid(double start, double end, double resolution, double* out, long cap) {
  if (!(resolution > 0.0)) return fail(SYNTHE_EINVAL, "resolution must be positive");
  if (!(start > 0.0) || !(end >= start)) return fail(SYNTHE_EINVAL, "need 0 < start <= end");
  double ratio = 1.0 + 1.0 / resolution;
  double rlog = log(ratio);
  double ix_floor = floor(log(start) / rlog);
  if (exp(ix_floor * rlog) < start) ix_floor += 1.0;
  double wl = exp(ix_floor * rlog);
  long n = 0;
  double lim = end * (1.0 + 1e-9);
  while (wl <= lim) {
    if (out && n < cap) out[n] = wl;
    ++n;
    wl *= ratio;
  }
  return n;
}

int synthe_continuum(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl, long n_edges,
                     const double* wledge, const double* half_edge, const double* delta_edge,
                     const double* coeff_abs, const double* coeff_scat, int coeff_log10,
                     double* cont_abs, double* cont_scat, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_continuum: ctx is NULL");
  if (n_depths <= 0 || n_wl < 0 || n_edges < 2)
    return fail(SYNTHE_EINVAL, "synthe_continuum: need n_depths > 0 and at least two edges");
  if (n_wl == 0) return SYNTHE_OK;
  if (!wl || !wledge || !coeff_abs || !coeff_scat || !cont_abs || !cont_scat)
    return fail(SYNTHE_EINVAL, "synthe_continuum: NULL array");
  const unsigned rows = (unsigned)((n_depths + kContinuumDepths - 1) / kContinuumDepths);
  hipLaunchKernelGGL(k_continuum, dim3(blocks_for(n_wl, kBlock), rows), dim3(kBlock), 0, (hipStream_t)stream,
                     (int)n_depths, n_wl, wl, (int)n_edges, wledge, half_edge, delta_edge, coeff_abs,
                     coeff_scat, coeff_log10, cont_abs, cont_scat);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_source_function_bfudge(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl,
                                  const double* temperature, const double* bfudge, double* bnu, double* stim,
                                  double* ehvkt, double* slinec, synthe_stream stream) {
  if (!ctx || !wl || !temperature) return fail(SYNTHE_EINVAL, "synthe_source_function: NULL argument");
  if (n_depths <= 0 || n_wl < 0 || n_depths > 65535)
    return fail(SYNTHE_EINVAL, "synthe_source_function: bad sizes");
  if (n_wl == 0) return SYNTHE_OK;
  hipLaunchKernelGGL(k_source, dim3(blocks_for(n_wl, kBlock), (unsigned)n_depths), dim3(kBlock), 0,
                     (hipStream_t)stream, (int)n_depths, n_wl, wl, temperature, bfudge, bnu, stim, ehvkt,
                     slinec);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_source_function(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl,
                           const double* temperature, double* bnu, double* stim, double* ehvkt,
                           double* slinec, synthe_stream stream) {
  return synthe_source_function_bfudge(ctx, n_depths, n_wl, wl, temperature, nullptr, bnu, stim, ehvkt, slinec,
                                       stream);
}


int synthe_transp(synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths, long n_wl,
                  const double* kapmin, double cutoff, double microturb_kms, double* transp,
                  uint8_t* valid, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_transp: ctx is NULL");
  int rc = check_lines_depths(lines, depths, "synthe_transp");
  if (rc) return rc;
  if (n_wl <= 0 || !kapmin || !transp || !valid)
    return fail(SYNTHE_EINVAL, "synthe_transp: kapmin/transp/valid required and n_wl > 0");
  long pairs = lines->n_lines * depths->n_depths;
  if (pairs == 0) return SYNTHE_OK;
  double micro = microturb_kms > 0.0 ? microturb_kms / 299792.458 : 0.0;
  hipLaunchKernelGGL(k_transp_wing, dim3(blocks_for(pairs, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, to_dev(lines), to_dev(depths), n_wl, kapmin, cutoff, micro,
                     nullptr, ctx->tb, 1, transp, valid, nullptr, nullptr, nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

// synthe_asynth, or with transp_out the fused form: TRANSP and valid are
// computed in the wing set-up's pass (k_transp_wing mode 3) and written out.
static int asynth_call(synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths, long n_wl,
                       const double* wl, const double* transp, const uint8_t* valid, const double* kapmin,
                       double cutoff, const int64_t* center_order, const int64_t* center_start,
                       const int64_t* wing_order, double* asynth, double* stats, hipStream_t st,
                       double* transp_out, uint8_t* valid_out) {
  const long L = lines->n_lines;
  const int D = (int)depths->n_depths;
  const AsynthLayout lay = asynth_layout(L, D);
  int rc = ensure_ws(ctx, lay.total);
  if (rc) return rc;
  const AsynthBufs bufs = asynth_bufs((char*)ctx->ws, lay);
  rc = asynth_prepare(ctx, lines, depths, n_wl, wl, transp, valid, kapmin, cutoff, wing_order, asynth,
                      stats != nullptr, bufs, st, true, transp_out, valid_out);
  if (rc) return rc;
  rc = timing_mark(ctx, st, true);
  if (rc) return rc;
  const dim3 g_gather((unsigned)blocks_for(n_wl, kWT), (unsigned)D);
  auto gather = stats ? (ctx->precision == SYNTHE_PREC_FP32 ? k_gather<true, kPrecFP32> : k_gather<true, kPrecFP64>)
                      : (ctx->precision == SYNTHE_PREC_FP32 ? k_gather<false, kPrecFP32> : k_gather<false, kPrecFP64>);
  const GatherArgs ga = gather_args(lines, depths, n_wl, wl, transp_out ? transp_out : transp,
                                    valid_out ? valid_out : valid, center_order, center_start, asynth, bufs);
  hipLaunchKernelGGL(gather, g_gather, dim3(64), 0, st, ga, stats ? bufs.ev + 3 : nullptr);
  HIP_TRY(hipGetLastError());
  rc = timing_mark(ctx, st, false);
  if (rc) return rc;
  if (stats) {
    if (L * (long)D > 0) {
      hipLaunchKernelGGL(k_wing_stats, dim3(64, (unsigned)D), dim3(kBlock), 0, st, L * (long)D, D, L, bufs.reg,
                         bufs.lcm, bufs.ev);
      HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_stats_to_double, dim3(1), dim3(64), 0, st, bufs.ev, stats);
    HIP_TRY(hipGetLastError());
  }
  return SYNTHE_OK;
}

int synthe_asynth(synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths, long n_wl,
                  const double* wl, const double* transp, const uint8_t* valid, const double* kapmin,
                  double cutoff, const int64_t* center_order, const int64_t* center_start,
                  const int64_t* wing_order, double* asynth, double* stats, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_asynth: ctx is NULL");
  int rc = check_asynth_args(lines, depths, n_wl, wl, transp, valid, kapmin, center_order, center_start,
                             wing_order, asynth, "synthe_asynth");
  if (rc) return rc;
  return asynth_call(ctx, lines, depths, n_wl, wl, transp, valid, kapmin, cutoff, center_order, center_start,
                     wing_order, asynth, stats, (hipStream_t)stream, nullptr, nullptr);
}

int synthe_transp_asynth(synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths, long n_wl,
                         const double* wl, const double* kapmin, double cutoff, const int64_t* center_order,
                         const int64_t* center_start, const int64_t* wing_order, double* transp, uint8_t* valid,
                         double* asynth, double* stats, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_transp_asynth: ctx is NULL");
  if (lines && lines->n_lines > 0 && (!transp || !valid))
    return fail(SYNTHE_EINVAL, "synthe_transp_asynth: transp/valid outputs required");
  // transp/valid are outputs here; the argument check wants non-NULL inputs
  int rc = check_asynth_args(lines, depths, n_wl, wl, transp, valid, kapmin, center_order, center_start,
                             wing_order, asynth, "synthe_transp_asynth");
  if (rc) return rc;
  return asynth_call(ctx, lines, depths, n_wl, wl, nullptr, nullptr, kapmin, cutoff, center_order, center_start,
                     wing_order, asynth, stats, (hipStream_t)stream, transp, valid);
}

int synthe_lte_spectrum(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl,
                        const double* temperature, const double* rhox, const double* cont_abs,
                        const double* cont_scat, const double* line_abs, const double* line_scat,
                        const double* line_source, double* flux_total, double* flux_cont,
                        int32_t* iters, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_lte_spectrum: ctx is NULL");
  if (n_depths <= 0 || n_wl < 0) return fail(SYNTHE_EINVAL, "synthe_lte_spectrum: bad sizes");
  if (n_depths > kMaxDepth)
    return fail(SYNTHE_ERANGE, "synthe_lte_spectrum: more than SYNTHE_MAX_DEPTH depths");
  if (n_wl == 0) return SYNTHE_OK;
  if (!wl || !temperature || !rhox || !cont_abs || !cont_scat || !line_abs || !line_scat ||
      !flux_total || !flux_cont)
    return fail(SYNTHE_EINVAL, "synthe_lte_spectrum: NULL array");
  if (n_wl > 0x7fffffffL) return fail(SYNTHE_ERANGE, "synthe_lte_spectrum: more than 2^31 wavelengths");
  const RtLayout y = rt_layout(n_wl);
  int rc = ensure_ws(ctx, y.total);
  if (rc) return rc;
  return rt_launch(ctx, n_depths, n_wl, wl, temperature, rhox, cont_abs, cont_scat, line_abs, line_scat,
                   line_source, flux_total, flux_cont, iters, (char*)ctx->ws, y, (hipStream_t)stream);
}

int synthe_lte_spectrum_noalloc(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl,
                                const double* temperature, const double* rhox, const double* cont_abs,
                                const double* cont_scat, const double* line_abs, const double* line_scat,
                                const double* line_source, double* flux_total, double* flux_cont,
                                int32_t* iters, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_lte_spectrum_noalloc: ctx is NULL");
  if (n_depths <= 0 || n_wl < 0) return fail(SYNTHE_EINVAL, "synthe_lte_spectrum_noalloc: bad sizes");
  if (n_depths > kMaxDepth)
    return fail(SYNTHE_ERANGE, "synthe_lte_spectrum_noalloc: more than SYNTHE_MAX_DEPTH depths");
  if (n_wl > SYNTHE_NOALLOC_MAX_WL)
    return fail(SYNTHE_ERANGE, "synthe_lte_spectrum_noalloc: " + std::to_string(n_wl) +
                                   " wavelengths, above the static scratch's SYNTHE_NOALLOC_MAX_WL = " +
                                   std::to_string((long)SYNTHE_NOALLOC_MAX_WL));
  if (n_wl == 0) return SYNTHE_OK;
  if (!wl || !temperature || !rhox || !cont_abs || !cont_scat || !line_abs || !line_scat ||
      !flux_total || !flux_cont)
    return fail(SYNTHE_EINVAL, "synthe_lte_spectrum_noalloc: NULL array");
  if (!ctx->scratch) return fail(SYNTHE_EHIP, "synthe_lte_spectrum_noalloc: no scratch address");
  const RtLayout y = rt_layout(n_wl);
  if (y.total > kNoAllocScratchBytes)
    return fail(SYNTHE_ERANGE, "synthe_lte_spectrum_noalloc: workspace exceeds the static scratch");
  // every byte of the layout that a kernel reads is written earlier in the
  // same call (the list counter by k_zero_int), so the scratch needs no clearing
  return rt_launch(ctx, n_depths, n_wl, wl, temperature, rhox, cont_abs, cont_scat, line_abs, line_scat,
                   line_source, flux_total, flux_cont, iters, (char*)ctx->scratch, y, (hipStream_t)stream);
}

int synthe_josh_flux_batch(synthe_ctx* ctx, long n, long n_cols, const double* acont,
                           const double* scont, const double* aline, const double* sline,
                           const double* sigmac, const double* sigmal, const double* rhox,
                           double* flux, int32_t* iters, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_josh_flux_batch: ctx is NULL");
  if (n < 0 || n_cols < 0) return fail(SYNTHE_EINVAL, "synthe_josh_flux_batch: bad sizes");
  if (n > kMaxDepth) return fail(SYNTHE_ERANGE, "synthe_josh_flux_batch: more than SYNTHE_MAX_DEPTH depths");
  if (n_cols == 0) return SYNTHE_OK;
  if (!flux || (n > 0 && (!acont || !scont || !aline || !sline || !sigmac || !sigmal || !rhox)))
    return fail(SYNTHE_EINVAL, "synthe_josh_flux_batch: NULL array");
  hipLaunchKernelGGL(k_josh_batch, dim3(blocks_for(n_cols, kRtLanes)), dim3(kRtLanes), 0,
                     (hipStream_t)stream, (int)n, n_cols, acont, scont, aline, sline, sigmac, sigmal, rhox,
                     flux, iters, ctx->tb);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_diag_div_pos(long n, int variant, const double* x, const double* b, double* q_fast,
                        double* q_ieee, synthe_stream stream) {
  if (n < 0 || (n > 0 && (!x || !b || !q_fast || !q_ieee)))
    return fail(SYNTHE_EINVAL, "synthe_diag_div_pos: bad arguments");
  if (n == 0) return SYNTHE_OK;
  hipLaunchKernelGGL(k_check_div, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, n,
                     variant, x, b, q_fast, q_ieee);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_flux_to_lambda(synthe_ctx* ctx, long n_wl, const double* wl, double* flux,
                          synthe_stream stream) {
  if (!ctx || !wl || !flux) return fail(SYNTHE_EINVAL, "synthe_flux_to_lambda: NULL argument");
  if (n_wl <= 0) return n_wl == 0 ? SYNTHE_OK : fail(SYNTHE_EINVAL, "synthe_flux_to_lambda: bad size");
  hipLaunchKernelGGL(k_flux_to_lambda, dim3(blocks_for(n_wl, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                     n_wl, wl, flux);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

}  // extern "C"

// ------------------------------------------------------------ hydrogen lines
namespace {

// One hydrogen line's depth-independent quantities (engine/opacity.py:1768-1819)
// and its three transitions: n_lower -> n_upper and the neighbours the
// wing cut-offs compare with (n_upper - 2 on the red side, + 2 on the blue).
struct HLine {
  hyd::Trans main, minus2, plus2;
  double wl, cgf, elo, wlminus1, wlminus2, wlplus1, wlplus2, redcut, bluecut, wshift, conth_val;
  long c;
  int simple;  // n_upper <= n_lower + 2: the plain walk (labels 620/630)
};

// A (line, depth) pair after the gates: its KAPPA0, WCON/WTAIL and how far
// each wing's walk ran (the last offset it passed; 0 for none).
struct HPair {
  double kappa0, wcon, wtail;
  int red_ext, blue_ext, center_add, active;
};

struct HDepthDev {
  const double *pop, *dop, *rho, *hckt, *emerge_h, *state, *temp;
};

// The value a wing pixel adds, before the cut-off tests (opacity.py:1620-1700).
__device__ __forceinline__ double h_value(const HLine& L, const HPair& pr, const hyd::Trans& t, double shift,
                                          const double* st, double wave, double stim, bool taper,
                                          const DeviceTables& tb) {
  double v = pr.kappa0 * hyd::profile(t, st, wave - shift, tb.extab, tb.extabf) * stim;
  if (taper && wave < pr.wtail) v *= (wave - pr.wcon) / (pr.wtail - pr.wcon);
  return v;
}

// Gates, WCON/WTAIL and the extent of each wing's walk for one (line, depth)
// pair (opacity.py:1820-1881 and 1547-1704).  One wave per pair; each walk
// takes 64 offsets per step and stops at the first offset the reference
// stops at, so the pixels it passes are exactly the reference's.
__global__ __launch_bounds__(64) void k_h_extent(const HLine* __restrict__ lines, HDepthDev dp, int D, long W,
                                                 const double* __restrict__ grid, const double* __restrict__ cont,
                                                 const double* __restrict__ stim, double cutoff, DeviceTables tb,
                                                 HPair* __restrict__ pairs) {
  const int l = blockIdx.x, d = blockIdx.y, lane = threadIdx.x;
  const HLine& L = lines[l];
  HPair pr{0.0, 0.0, 0.0, 0, 0, 0, 0};
  const double* st = dp.state + (long)d * hyd::kStateN;
  const double pop_val = dp.pop[d], dop_val = dp.dop[d], r = dp.rho[d];
  bool ok = !(pop_val <= 0.0 || dop_val <= 0.0 || r <= 0.0) && st[hyd::kHAS] != 0.0;
  if (ok) {
    const double xnfdop = pop_val / (r * dop_val);
    const double boltz = fastex(L.elo * dp.hckt[d], tb.extab, tb.extabf);
    const double kappa0_pre = L.cgf * xnfdop;
    const long cc = L.c < 0 ? 0 : (L.c > W - 1 ? W - 1 : L.c);
    const double kapmin = cont[(long)d * W + cc] * cutoff;
    const double kappa0 = kappa0_pre * boltz;
    ok = !(kappa0_pre < kapmin) && !(kappa0 < kapmin);
    if (ok) {
      double wmerge = 1.0e7 / (L.conth_val - dp.emerge_h[d]);
      if (wmerge < 0.0) wmerge = L.wshift + L.wshift;
      double wcon = L.wshift > wmerge ? L.wshift : wmerge;
      double wtail = wcon > 0.0 ? 1.0e7 / (1.0e7 / wcon - 500.0) : wcon + wcon;
      wcon = L.wshift + L.wshift < wcon ? L.wshift + L.wshift : wcon;
      if (wtail < 0.0) wtail = wcon + wcon;
      wtail = wcon + wcon < wtail ? wcon + wcon : wtail;
      pr.kappa0 = kappa0;
      pr.wcon = wcon;
      pr.wtail = wtail;
    }
  }
  if (!ok) {
    if (lane == 0) pairs[(long)l * D + d] = pr;
    return;
  }
  pr.active = 1;
  const double temp = dp.temp[d];
  // STIM at a pixel: the caller's, else the engine's 1 - exp(-h nu / k T)
  auto stim_at = [&](long idx, double wave) {
    return stim ? stim[(long)d * W + idx] : planck_terms(wave, temp).stim;
  };
  const bool simple = L.simple != 0;
  const bool taper = !simple && pr.wtail > pr.wcon;
  const long c = L.c;
  const long max_steps = c > W - c - 1 ? c : W - c - 1;
  if (c >= 0 && c < W) {
    const double wave = grid[c];
    if (!(!simple && wave < pr.wcon)) {
      const double v = h_value(L, pr, L.main, L.wl, st, wave, stim_at(c, wave), taper, tb);
      pr.center_add = v >= cont[(long)d * W + c] * cutoff;
    }
  }
  // red wing
  long ext = 0;
  if (c < W) {
    const long start = c < 0 ? (-c > 1 ? -c : 1) : 1;
    for (long o0 = start;; o0 += 64) {
      const long o = o0 + lane;
      bool stop = o > max_steps;
      if (!stop) {
        const long idx = c + o;
        if (idx >= W) {
          stop = true;
        } else {
          const double wave = grid[idx];
          const double cut = cont[(long)d * W + idx] * cutoff;
          if (!simple) {
            if (wave > L.wlminus1) {
              stop = true;
            } else if (!(wave < pr.wcon)) {
              const double sv = stim_at(idx, wave);
              double v = h_value(L, pr, L.main, L.wl, st, wave, sv, taper, tb);
              if (wave > L.redcut && h_value(L, pr, L.minus2, L.wlminus2, st, wave, sv, taper, tb) >= v) v = 0.0;
              stop = v <= 0.0 || v < cut;
            }
          } else {
            const double v = h_value(L, pr, L.main, L.wl, st, wave, stim_at(idx, wave), false, tb);
            stop = v <= 0.0 || v < cut;
          }
        }
      }
      const unsigned long long m = __ballot(stop);
      if (m) {
        ext = o0 + __builtin_ctzll(m) - 1;
        break;
      }
    }
  }
  pr.red_ext = (int)ext;
  // blue wing
  ext = 0;
  if (c >= 0) {
    const long start = c >= W ? (c - (W - 1) > 1 ? c - (W - 1) : 1) : 1;
    for (long o0 = start;; o0 += 64) {
      const long o = o0 + lane;
      bool stop = o > max_steps;
      if (!stop) {
        const long idx = c - o;
        if (idx < 0) {
          stop = true;
        } else {
          const double wave = grid[idx];
          if (!simple && (wave < pr.wcon || wave < L.wlplus1)) {
            stop = true;
          } else {
            const double sv = stim_at(idx, wave);
            double v = h_value(L, pr, L.main, L.wl, st, wave, sv, taper, tb);
            if (!simple && wave < L.bluecut && h_value(L, pr, L.plus2, L.wlplus2, st, wave, sv, taper, tb) >= v)
              v = 0.0;
            stop = v <= 0.0 || v < cont[(long)d * W + idx] * cutoff;
          }
        }
      }
      const unsigned long long m = __ballot(stop);
      if (m) {
        ext = o0 + __builtin_ctzll(m) - 1;
        break;
      }
    }
  }
  pr.blue_ext = (int)ext;
  if (lane == 0) pairs[(long)l * D + d] = pr;
}

// AHLINE per pixel: every active pair whose walk passed the pixel adds its
// value, in catalog order, from zero; the sum joins ASYNTH.
__global__ __launch_bounds__(256) void k_h_gather(const HLine* __restrict__ lines, const HPair* __restrict__ pairs,
                                                  int NH, int D, long W, const double* __restrict__ grid,
                                                  const double* __restrict__ temp, const double* __restrict__ state,
                                                  const double* __restrict__ stim, DeviceTables tb,
                                                  double* __restrict__ asynth) {
  const long w = blockIdx.x * 256L + threadIdx.x;
  const int d = blockIdx.y;
  if (w >= W) return;
  const double wave = grid[w];
  const double sv = stim ? stim[(long)d * W + w] : planck_terms(wave, temp[d]).stim;
  const double* st = state + (long)d * hyd::kStateN;
  double sum = 0.0;
  for (int l = 0; l < NH; ++l) {
    const HPair& pr = pairs[(long)l * D + d];
    if (!pr.active) continue;
    const HLine& L = lines[l];
    const long o = w - L.c;
    const bool simple = L.simple != 0;
    const bool taper = !simple && pr.wtail > pr.wcon;
    bool add;
    if (o == 0) add = pr.center_add;
    else if (o > 0) add = o <= pr.red_ext && !(!simple && wave < pr.wcon);
    else add = -o <= pr.blue_ext;
    if (add) sum += h_value(L, pr, L.main, L.wl, st, wave, sv, taper, tb);
  }
  asynth[(long)d * W + w] += sum;
}

// Host-staged inputs of one call, kept until the next call on the context.
struct HydrogenStage {
  std::vector<HLine> lines;
  std::vector<double> depth;  // pop dop rho hckt emerge_h temp [D] each, then state [D][kStateN]
  void* dev = nullptr;
  size_t dev_size = 0;
};
void hstage_free(void* p) {
  HydrogenStage* S = (HydrogenStage*)p;
  if (S->dev) (void)hipFree(S->dev);
  delete S;
}

HydrogenStage* hstage(synthe_ctx* ctx) {
  if (!ctx->hstage) {
    ctx->hstage = new HydrogenStage();
    ctx->hstage_free = hstage_free;
  }
  return (HydrogenStage*)ctx->hstage;
}

}  // namespace

extern "C" {

int synthe_hydrogen_lines(synthe_ctx* ctx, const synthe_hlines* hl, const synthe_hdepths* hd, double microturb_kms,
                          double cutoff, long n_wl, const double* wl, const double* cont, const double* stim,
                          double* asynth, synthe_stream stream) {
  const char* who = "synthe_hydrogen_lines";
  if (!ctx || !hl || !hd || !wl || !cont || !asynth) return fail(SYNTHE_EINVAL, std::string(who) + ": NULL argument");
  const long NH = hl->n_lines, D = hd->n_depths;
  if (NH < 0 || D <= 0 || D > 65535 || n_wl <= 0 || n_wl > 0x7fffffffL)
    return fail(SYNTHE_EINVAL, std::string(who) + ": bad sizes");
  if (NH == 0) return SYNTHE_OK;
  if (NH > 65535) return fail(SYNTHE_ERANGE, std::string(who) + ": more than 65535 hydrogen lines");
  if (!hl->wavelength || !hl->center || !hl->gf || !hl->elo || !hl->n_lower || !hl->n_upper || !hd->temperature ||
      !hd->electron_density || !hd->mass_density || !hd->hckt || !hd->pop_h1 || !hd->dop_h1)
    return fail(SYNTHE_EINVAL, std::string(who) + ": NULL array");
  hipStream_t st = (hipStream_t)stream;
  HydrogenStage* S = hstage(ctx);
  HIP_TRY(hipStreamSynchronize(st));  // the previous call's copies have left the staging vectors
  S->lines.resize(NH);
  for (long l = 0; l < NH; ++l) {  // engine/opacity.py:1768-1819
    HLine& L = S->lines[l];
    const int n_lower = hl->n_lower[l], n_upper = hl->n_upper[l];
    const int nle = n_lower > 1 ? n_lower : 1;
    const int nue = n_upper > n_lower + 1 ? n_upper : n_lower + 1;
    const int um2 = nue - 2 > nle + 1 ? nue - 2 : nle + 1;
    if (nue + 2 > 96 || !hyd::make_trans(nle, nue, &L.main) || !hyd::make_trans(nle, um2, &L.minus2) ||
        !hyd::make_trans(nle, nue + 2, &L.plus2))
      return fail(SYNTHE_ERANGE, std::string(who) + ": n_upper + 2 beyond the 96-level damping table (line " +
                                     std::to_string(l) + ")");
    const double line_wavelength = hl->wavelength[l];
    L.wl = line_wavelength;
    L.c = hl->center[l];
    L.cgf = hyd::kCgf * hl->gf[l] / (hyd::kCNm / line_wavelength);
    L.elo = hl->elo[l];
    int ncon = n_lower < 15 ? n_lower : 15;
    ncon = (ncon > 1 ? ncon : 1) - 1;
    L.conth_val = kH_conth[ncon];
    const double ehyd_lower = hyd::ehyd_cm(n_lower);
    L.wlminus1 = n_upper - 1 > n_lower ? 1.0e7 / (hyd::ehyd_cm(n_upper - 1) - ehyd_lower) : line_wavelength;
    L.wlminus2 = n_upper - 2 > n_lower ? 1.0e7 / (hyd::ehyd_cm(n_upper - 2) - ehyd_lower) : line_wavelength;
    L.wlplus1 = 1.0e7 / (hyd::ehyd_cm(n_upper + 1) - ehyd_lower);
    L.wlplus2 = 1.0e7 / (hyd::ehyd_cm(n_upper + 2) - ehyd_lower);
    L.redcut = 1.0e7 / (kH_conth[0] - hyd::kRydCm / pow((double)n_upper - 0.8, 2.0) - ehyd_lower);
    L.bluecut = 1.0e7 / (kH_conth[0] - hyd::kRydCm / pow((double)n_upper + 0.8, 2.0) - ehyd_lower);
    L.wshift = 1.0e7 / (L.conth_val - hyd::kRydCm / pow(81.0, 2.0));
    L.simple = nue <= nle + 2;
  }
  const double micro = microturb_kms > 0.0 ? microturb_kms / 299792.458 : 0.0;
  S->depth.assign(D * (6 + hyd::kStateN), 0.0);
  double* pop = S->depth.data();
  double *dop = pop + D, *rho = dop + D, *hckt = rho + D, *emerge = hckt + D, *temp = emerge + D,
         *state = temp + D;
  const bool has = hd->xnf_he1 != nullptr || hd->xnf_h2 != nullptr;
  for (long d = 0; d < D; ++d) {
    pop[d] = hd->pop_h1[d];
    const double dv = hd->dop_h1[d];
    dop[d] = micro > 0.0 ? sqrt(dv * dv + micro * micro) : dv;
    rho[d] = hd->mass_density[d];
    hckt[d] = hd->hckt[d];
    const double xne = hd->electron_density[d] > 1e-40 ? hd->electron_density[d] : 1e-40;
    const double inglis = 1600.0 / pow(xne, 2.0 / 15.0);
    const double nmerge = inglis - 1.5 > 1.0 ? inglis - 1.5 : 1.0;
    const double nm2 = nmerge * nmerge;
    emerge[d] = hyd::kRydCm / (nm2 > 1e-12 ? nm2 : 1e-12);
    temp[d] = hd->temperature[d];
    hyd::make_state(hd->temperature[d], hd->electron_density[d], hd->xnf_he1, hd->xnf_h2, hd->xnfph, hd->dopph,
                    hd->turbulent_velocity, microturb_kms, has, (int)d, state + d * hyd::kStateN);
  }
  const size_t o_lines = 0, o_depth = align_up(NH * sizeof(HLine)),
               o_pairs = o_depth + align_up(S->depth.size() * sizeof(double)),
               total = o_pairs + align_up((size_t)NH * D * sizeof(HPair));
  if (S->dev_size < total) {
    if (S->dev) {
      HIP_TRY(hipFree(S->dev));
      S->dev = nullptr;
      S->dev_size = 0;
    }
    if (hipMalloc(&S->dev, total) != hipSuccess) {
      S->dev = nullptr;
      return fail(SYNTHE_ENOMEM, std::string(who) + ": staging allocation failed");
    }
    S->dev_size = total;
  }
  char* base = (char*)S->dev;
  HLine* d_lines = (HLine*)(base + o_lines);
  double* d_depth = (double*)(base + o_depth);
  HPair* d_pairs = (HPair*)(base + o_pairs);
  HIP_TRY(hipMemcpyAsync(d_lines, S->lines.data(), NH * sizeof(HLine), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_depth, S->depth.data(), S->depth.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HDepthDev dp{d_depth, d_depth + D, d_depth + 2 * D, d_depth + 3 * D, d_depth + 4 * D, d_depth + 6 * D,
               d_depth + 5 * D};
  hipLaunchKernelGGL(k_h_extent, dim3((unsigned)NH, (unsigned)D), dim3(64), 0, st, d_lines, dp, (int)D, n_wl, wl,
                     cont, stim, cutoff, ctx->tb, d_pairs);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_h_gather, dim3((unsigned)((n_wl + 255) / 256), (unsigned)D), dim3(256), 0, st, d_lines,
                     d_pairs, (int)NH, (int)D, n_wl, wl, dp.temp, dp.state, stim, ctx->tb, asynth);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

}  // extern "C"

// ------------------------------------------------------------ fort.19 lines
namespace {

// One fort.19 record the ASYNTH add takes (engine/opacity.py:1364-1546),
// packed on the host: the record's wavelength, float32 gf and damping (as
// doubles), centre index, the matched catalog line's Z - 1 and stage, the
// record's NELION and NLAST, CONTX(NCON, NELIONX) (0 where there is no WCON),
// and the matched line's wavelength and lower level energy for the Doppler
// width and Boltzmann factor when the caller does not give them.
struct F19Line {
  double wl, gf, gr, gs, gw, cont_val, lam_cat, elo;
  long c;
  int kind, elem, ion, nelion, nlast;
};

// A (record, depth) pair after its gates: what the per-pixel sum needs.
struct F19Pair {
  double kappa0;  // KAPPA0 (normal, autoionising) or KAPPA (merged continuum)
  double adamp, dopple, resolu, x_far, kapcen, wcon, wtail;
  long cpix;      // normal: the clamped centre; autoionising: the pixel its centre term lands on
  long lo, hi;    // normal: wing steps [1, hi] red (from lo) and [1, blue] blue; continuum: pixels [lo, hi)
  long blue;      // normal: last blue step; autoionising: blue walk extent
  long red;       // autoionising: red walk extent
  long merge, tail, denom;  // continuum: IDX_MERGE, IDX_TAIL, the ramp's denominator
  int active, center_add, n10dop, nstep, early, wing;
};

struct F19DepthDev {
  const double *pop, *rho, *xne, *txnxn, *emerge, *emerge_h, *temp, *hckt, *vscale, *boltz, *doppler;
  int n_ion, n_elem;
};

// profile(k) of the metal wing loop (engine/opacity.py:150-198): the near
// wing by the TABI walk (a < 0.2) or the Voigt function, the far wing X / k^2.
__device__ __forceinline__ double f19_profile(const F19Pair& p, long k, const DeviceTables& tb) {
  if (p.early || k <= p.n10dop) {
    if (p.adamp < 0.2) {
      const double dr = p.dopple * p.resolu;
      const double tabstep = dr > 0.0 ? 200.0 / dr : 200.0;
      long itab = tabi_index(tabstep, k);
      itab = itab < 0 ? 0 : (itab > kVoigtN - 1 ? kVoigtN - 1 : itab);
      return p.kappa0 * (tb.h0[itab] + p.adamp * tb.h1[itab]);
    }
    const double dvoigt = p.dopple * p.resolu > 0.0 ? 1.0 / p.dopple / p.resolu : 1e-6;
    return p.kappa0 * voigt((double)k * dvoigt, p.adamp, tb.h0, tb.h1, tb.h2);
  }
  return p.x_far / ((double)k * (double)k);
}

// The normal record's wing value at a pixel step k, with WTAIL's taper.
__device__ __forceinline__ double f19_normal_value(const F19Pair& p, long k, double wave, double line_wl,
                                                   const DeviceTables& tb) {
  double v = f19_profile(p, k, tb);
  if (p.wtail > 0.0 && wave < p.wtail) {
    const double base = p.wcon > 0.0 ? p.wcon : line_wl;
    const double span = p.wtail - base;
    v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
  }
  return v;
}

// A helium line's wing value (engine/opacity.py:998-1053, the TYPE < -2
// branch of _apply_fort19_profile): KAPPA0 H(a, |dλ| / Δλ_D), with WTAIL's
// taper (base WCON, else the line).
__device__ __forceinline__ double f19_he_value(const F19Pair& p, double wave, double line_wl,
                                               const DeviceTables& tb) {
  double v = p.kappa0 * voigt(fabs(wave - line_wl) / p.dopple, p.adamp, tb.h0, tb.h1, tb.h2);
  if (p.wtail > 0.0 && wave < p.wtail) {
    const double base = p.wcon > 0.0 ? p.wcon : line_wl;
    const double span = p.wtail - base;
    v = v * (wave - base) / (span > 1e-12 ? span : 1e-12);
  }
  return v;
}

// Shore profile of an autoionising line (engine/opacity.py:895-963).
__device__ __forceinline__ double f19_auto_value(const F19Line& L, double kappa0, double wave) {
  const double fc = kCLightCm / (L.wl * kNmToCm);
  const double gamma = fabs(L.gr) > 1e-30 ? fabs(L.gr) : 1e-30;
  const double bshore = fabs(L.gw) < 1e-30 ? 1e-30 : L.gw;
  const double eps = 2.0 * (kCLightCm / (wave * kNmToCm) - fc) / gamma;
  return kappa0 * (L.gs * eps + bshore) / (eps * eps + 1.0) / bshore;
}

// First index with grid[i] >= x (side "left") or > x (side "right").
__device__ __forceinline__ long f19_search(const double* grid, long W, double x, bool right) {
  long lo = 0, hi = W;
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (right ? grid[mid] <= x : grid[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// WCON / WTAIL of a record at depth d (opacity.py:646-686), vacuum
// wavelengths; -1 where the reference has None.
__device__ __forceinline__ void f19_limits(const F19Line& L, const F19DepthDev& dp, int d, F19Pair* p) {
  p->wcon = p->wtail = -1.0;
  if (!(L.cont_val > 0.0)) return;
  const double emerge_line = L.nelion == 1 ? dp.emerge_h[d] : dp.emerge[d];
  const double den = L.cont_val - emerge_line;
  if (!(fabs(den) > 1e-8)) return;
  const double wcon = 1.0e7 / den;
  const double dt = den - 500.0;
  double wtail = -1.0;
  if (fabs(dt) > 1e-8) {
    wtail = 1.0e7 / dt;
    if (wtail < 0.0) wtail = 2.0 * wcon;
    wtail = 2.0 * wcon < wtail ? 2.0 * wcon : wtail;
  }
  if (wtail > 0.0 && wtail <= wcon) wtail = -1.0;
  p->wcon = wcon > 0.0 ? wcon : -1.0;
  p->wtail = wtail > 0.0 ? wtail : -1.0;
}

// Gates and set-up of one (record, depth) pair (engine/opacity.py:1383-1546),
// one wave per pair: lane 0 runs the normal record's near-wing probe; the
// autoionising and merged-continuum walks take 64 offsets per step and stop
// where the reference's loops stop.
__global__ __launch_bounds__(64) void k_f19_extent(const F19Line* __restrict__ lines, F19DepthDev dp, int N, int D,
                                                   long W, const double* __restrict__ grid,
                                                   const double* __restrict__ cont, double cutoff, DeviceTables tb,
                                                   F19Pair* __restrict__ pairs) {
  const int r = blockIdx.x, d = blockIdx.y, lane = threadIdx.x;
  const F19Line& L = lines[r];
  F19Pair p{};
  const double* crow = cont + (long)d * W;
  const double pop_val = dp.pop[((long)d * dp.n_ion + (L.ion - 1)) * dp.n_elem + L.elem];
  const double boltz = dp.boltz ? dp.boltz[(long)r * D + d] : fastex(L.elo * dp.hckt[d], tb.extab, tb.extabf);
  const double rho = dp.rho[d];
  bool ok = pop_val > 0.0 && rho > 0.0;
  if (ok && L.kind == 0) {  // normal (opacity.py:1407-1478)
    const double dw = dp.doppler ? dp.doppler[(long)r * D + d] : L.lam_cat * dp.vscale[d];
    ok = L.wl > 0.0 && dw > 0.0;
    const double dopple = ok ? dw / L.wl : 0.0;
    ok = ok && dopple > 0.0;
    if (ok) {
      const double xnfdop = pop_val / (rho * dopple);
      const double kappa0_pre = L.gf * xnfdop;
      const long cc = L.c < 0 ? 0 : (L.c > W - 1 ? W - 1 : L.c);
      const double kapmin = crow[cc] * cutoff;
      const double kappa0 = kappa0_pre * boltz;
      ok = !(kappa0_pre < kapmin) && !(kappa0 < kapmin);
      if (ok) {
        const double gamma_total = L.gr + L.gs * dp.xne[d] + L.gw * dp.txnxn[d];
        const double adamp = gamma_total / dopple;
        const double kapcen = adamp < 0.2 ? kappa0 * (1.0 - 1.128 * adamp) : kappa0 * voigt(0.0, adamp, tb.h0, tb.h1, tb.h2);
        p.active = 1;
        p.cpix = cc;
        p.kapcen = kapcen;
        p.center_add = kapcen >= kapmin;
        p.kappa0 = kappa0;
        p.dopple = dopple;
        f19_limits(L, dp, d, &p);
        // the metal wing loop's set-up and near-wing probe (opacity.py:128-198)
        p.adamp = adamp > 1e-12 ? adamp : 1e-12;
        if (cc < W - 1) p.resolu = 1.0 / (grid[cc + 1] / grid[cc] - 1.0);
        else p.resolu = cc > 0 ? 1.0 / (grid[cc] / grid[cc - 1] - 1.0) : 300000.0;
        long n10 = (long)(10.0 * dopple * p.resolu);
        n10 = n10 < kMaxProfileSteps ? n10 : kMaxProfileSteps;
        p.n10dop = (int)n10;
        long nstep = 0;
        if (lane != 0) return;
        if (kappa0 > 0.0) {  // else no wing (opacity.py:128-129)
          const double kmin = cutoff * crow[cc];
          double last = 0.0;
          for (long k = 1; k <= n10; ++k) {
            p.early = 1;  // provisional: f19_profile takes the near wing for any k
            last = f19_profile(p, k, tb);
            p.early = 0;
            if (last < kmin) {
              nstep = k;
              p.early = 1;
              break;
            }
          }
          if (!p.early) {
            const double xf = n10 > 0 && last > 0.0 ? last * ((double)n10 * (double)n10) : 0.0;
            long maxstep = n10;
            if (xf > 0.0 && kmin > 0.0) {
              const double ms = sqrt(xf / kmin) + 1.0;  // int() of it, capped at MAX_PROFILE_STEPS
              maxstep = ms < 2.0e6 ? (long)ms : kMaxProfileSteps;
              maxstep = maxstep < kMaxProfileSteps ? maxstep : kMaxProfileSteps;
            }
            p.x_far = xf;
            nstep = maxstep;
          }
        }
        const long c = L.c;
        p.nstep = (int)nstep;
        p.wing = !(c + nstep < 0 || c - nstep >= W) && nstep > 0;
        p.lo = c >= 0 ? 1 : -c;  // MINRED
        p.hi = c < W - 1 ? (c >= 0 ? (W - 1 - c < nstep ? W - 1 - c : nstep) : (W - 1 < nstep ? W - 1 : nstep)) : 0;
        p.blue = c > 0 ? (c < nstep ? c : nstep) : 0;
        pairs[(long)r * D + d] = p;
        return;
      }
    }
  } else if (ok && L.kind == 2) {  // merged continuum (opacity.py:1480-1513, 850-892)
    const double kappa = L.gf * (pop_val / rho) * boltz;
    double wmerge = L.wl, wtail = L.wl;
    if (L.wl > 0.0 && L.nlast > 0) {
      const double ryd = L.nelion == 1 ? 109677.576 : 109737.312;
      const double ds = 1.0e7 / L.wl - ryd / (double)(L.nlast * L.nlast);
      const double emerge_line = L.nelion == 1 ? dp.emerge_h[d] : dp.emerge[d];
      const double dm = 1.0e7 / L.wl - emerge_line;
      if (fabs(ds) > 1e-12 && fabs(dm) > 1e-12) {
        const double wshift = 1.0e7 / ds;
        double wm = 1.0e7 / dm;
        if (wm < 0.0) wm = wshift + wshift;
        wm = wm > wshift ? wm : wshift;
        wm = wshift + wshift < wm ? wshift + wshift : wm;
        wmerge = wm;
        const double dt = 1.0e7 / wm - 500.0;
        if (fabs(dt) > 1e-12) {
          double wt = 1.0e7 / dt;
          if (wt < 0.0) wt = wm + wm;
          wtail = wm + wm < wt ? wm + wm : wt;
        }
      }
    }
    ok = kappa > 0.0;
    const long start = L.c > 0 ? L.c : 0;
    const long merge = f19_search(grid, W, wmerge, false);
    long tail = f19_search(grid, W, wtail, true);
    tail = tail < W ? tail : W;
    ok = ok && tail > start;
    if (ok) {
      const long m0 = merge > start ? merge : start;
      p.denom = tail - m0 > 1 ? tail - m0 : 1;
      p.kappa0 = kappa;
      p.merge = merge;
      p.tail = tail;
      p.lo = start;
      long stop = tail;
      for (long i0 = start; i0 < tail; i0 += 64) {
        const long i = i0 + lane;
        bool s = false;
        if (i < tail) {
          const double wave = grid[i];
          if (!(wave < L.wl)) {
            double v = kappa;
            if (i >= merge) v *= (double)(tail - i > 0 ? tail - i : 0) / (double)p.denom;
            s = v < crow[i] * cutoff;
          }
        }
        const unsigned long long m = __ballot(s);
        if (m) {
          stop = i0 + __builtin_ctzll(m);
          break;
        }
      }
      p.hi = stop;
      p.active = stop > start;
    }
  } else if (ok && L.kind == 1) {  // autoionising (opacity.py:1515-1545, 895-963)
    const double kappa0 = L.gw * L.gf * (pop_val / rho) * boltz;
    const long c = L.c;
    const long cpix = c >= 0 ? c : W + c;  // the reference indexes buffer[-1] for a centre below the grid
    ok = !(kappa0 < crow[cpix] * cutoff) && kappa0 > 0.0;
    if (ok) {
      p.active = 1;
      p.kappa0 = kappa0;
      p.cpix = cpix;
      p.center_add = 1;
      const long max_steps = c > W - c - 1 ? c : W - c - 1;
      for (int side = 0; side < 2; ++side) {
        long ext = 0;
        for (long o0 = 1;; o0 += 64) {
          const long o = o0 + lane;
          bool stop = o > max_steps;
          if (!stop) {
            const long idx = side == 0 ? c + o : c - o;
            if (idx >= W || idx < 0) {
              stop = true;
            } else {
              const double v = f19_auto_value(L, kappa0, grid[idx]);
              stop = v <= 0.0 || v < crow[idx] * cutoff;
            }
          }
          const unsigned long long m = __ballot(stop);
          if (m) {
            ext = o0 + __builtin_ctzll(m) - 1;
            break;
          }
        }
        if (side == 0) p.red = ext;
        else p.blue = ext;
      }
    }
  } else if (ok && L.kind == 3) {  // helium line (opacity.py:2716-3066, then :998-1053)
    const double dop_val = dp.doppler[(long)r * D + d];  // DOPPLE of the line's stage
    const double dw = dop_val * L.wl;
    const long c = L.c;
    const double* grow = grid;
    ok = dop_val > 0.0 && !(dw <= 0.0);
    const double xnfdop = pop_val / (rho * dop_val);
    const double kapmin = crow[c] * cutoff;
    const double kappa0_pre = L.gf * xnfdop;
    const double kappa0 = kappa0_pre * boltz;
    ok = ok && !(kappa0_pre < kapmin) && !(kappa0 < kapmin);
    if (ok) {
      f19_limits(L, dp, d, &p);
      double doppler = dw > 1e-12 ? dw : 1e-12;
      double keff = kappa0;
      if (L.nlast == -4) {  // 3He
        keff /= 1.155;
        doppler *= 1.155;
      }
      const double dl = doppler / L.wl;
      const double damping = (L.gr + L.gs * dp.xne[d] + L.gw * dp.txnxn[d]) / (dl > 1e-40 ? dl : 1e-40);
      p.active = 1;
      p.kappa0 = keff;
      p.dopple = doppler;
      p.adamp = damping > 1e-12 ? damping : 1e-12;
      p.cpix = c;
      // red wing: c .. W-1; a pixel at or below WCON is passed over, the
      // first value below the cutoff ends the wing (not added)
      long red_end = c;
      if (L.wl <= grow[W - 1]) {
        red_end = W;
        for (long i0 = c; i0 < W; i0 += 64) {
          const long i = i0 + lane;
          bool s = false;
          if (i < W) {
            const double wave = grow[i];
            if (!(p.wcon > 0.0 && wave <= p.wcon)) s = f19_he_value(p, wave, L.wl, tb) < crow[i] * cutoff;
          }
          const unsigned long long m = __ballot(s);
          if (m) {
            red_end = i0 + __builtin_ctzll(m);
            break;
          }
        }
      }
      p.hi = red_end;
      // blue wing: c - 1 .. 0; ends at a pixel at or below WCON (not added) or
      // after the first value below the cutoff (added)
      long blue_lo = c;
      if (c > 0 && L.wl >= grow[0]) {
        blue_lo = 0;
        for (long i0 = c - 1; i0 >= 0; i0 -= 64) {
          const long i = i0 - lane;
          int s = 0;  // 1: stop before i, 2: stop after i
          if (i >= 0) {
            const double wave = grow[i];
            if (p.wcon > 0.0 && wave <= p.wcon) s = 1;
            else if (f19_he_value(p, wave, L.wl, tb) < crow[i] * cutoff) s = 2;
          }
          const unsigned long long m = __ballot(s != 0);
          if (m) {
            const int first = __builtin_ctzll(m);
            const int how = __shfl(s, first);
            blue_lo = i0 - first + (how == 1 ? 1 : 0);
            break;
          }
        }
      }
      p.lo = blue_lo;
    }
  }
  if (lane == 0) pairs[(long)r * D + d] = p;
}

// The records' sum per pixel, in fort.19 order from zero (engine/opacity.py
// :1362-1546), then ASYNTH += sum * STIM (:1545-1546).
__global__ __launch_bounds__(256) void k_f19_gather(const F19Line* __restrict__ lines,
                                                    const F19Pair* __restrict__ pairs, int N, int D, long W,
                                                    const double* __restrict__ grid, const double* __restrict__ temp,
                                                    const double* __restrict__ stim, DeviceTables tb,
                                                    double* __restrict__ asynth) {
  const long w = blockIdx.x * 256L + threadIdx.x;
  const int d = blockIdx.y;
  if (w >= W) return;
  const double wave = grid[w];
  double sum = 0.0;
  for (int r = 0; r < N; ++r) {
    const F19Pair& p = pairs[(long)r * D + d];
    if (!p.active) continue;
    const F19Line& L = lines[r];
    if (L.kind == 0) {
      if (p.center_add && w == p.cpix) sum += p.kapcen;
      if (p.wing && !(p.wcon > 0.0 && wave <= p.wcon)) {
        const long k = w - L.c;
        if (k > 0 ? (k >= p.lo && k <= p.hi) : (-k >= 1 && -k <= p.blue)) sum += f19_normal_value(p, k > 0 ? k : -k, wave, L.wl, tb);
      }
    } else if (L.kind == 1) {
      if (w == p.cpix) sum += p.kappa0;
      const long o = w - L.c;
      if ((o > 0 && o <= p.red) || (o < 0 && -o <= p.blue)) sum += f19_auto_value(L, p.kappa0, wave);
    } else if (L.kind == 3) {
      if ((w >= p.cpix && w < p.hi && !(p.wcon > 0.0 && wave <= p.wcon)) || (w < p.cpix && w >= p.lo))
        sum += f19_he_value(p, wave, L.wl, tb);
    } else {
      if (w >= p.lo && w < p.hi && !(wave < L.wl)) {
        double v = p.kappa0;
        if (w >= p.merge) v *= (double)(p.tail - w > 0 ? p.tail - w : 0) / (double)p.denom;
        sum += v;
      }
    }
  }
  if (sum != 0.0) {
    const double sv = stim ? stim[(long)d * W + w] : planck_terms(wave, temp[d]).stim;
    asynth[(long)d * W + w] += sum * sv;
  }
}

struct Fort19Stage {
  std::vector<F19Line> lines;
  std::vector<double> depth;
  void* dev = nullptr;
  size_t dev_size = 0;
};
void f19stage_free(void* p) {
  Fort19Stage* S = (Fort19Stage*)p;
  if (S->dev) (void)hipFree(S->dev);
  delete S;
}

}  // namespace

extern "C" {

int synthe_fort19_lines(synthe_ctx* ctx, const synthe_f19lines* fl, const synthe_f19depths* fd, double cutoff,
                        long n_wl, const double* wl, const double* cont, const double* stim, double* asynth,
                        synthe_stream stream) {
  const char* who = "synthe_fort19_lines";
  if (!ctx || !fl || !fd || !wl || !cont || !asynth) return fail(SYNTHE_EINVAL, std::string(who) + ": NULL argument");
  const long N = fl->n_lines, D = fd->n_depths;
  if (N < 0 || D <= 0 || D > 65535 || n_wl <= 0 || n_wl > 0x7fffffffL || fd->n_ion <= 0 || fd->n_elem <= 0)
    return fail(SYNTHE_EINVAL, std::string(who) + ": bad sizes");
  if (N == 0) return SYNTHE_OK;
  if (N > 65535) return fail(SYNTHE_ERANGE, std::string(who) + ": more than 65535 fort.19 records");
  if (!fl->kind || !fl->wavelength || !fl->center || !fl->gf || !fl->gamma_rad || !fl->gamma_stark ||
      !fl->gamma_vdw || !fl->elem || !fl->ion || !fl->nelion || !fl->nlast || !fl->cont_val || !fd->pop ||
      !fd->mass_density || !fd->electron_density || !fd->txnxn || !fd->emerge || !fd->emerge_h ||
      !fd->temperature)
    return fail(SYNTHE_EINVAL, std::string(who) + ": NULL array");
  if (!fd->boltz && (!fd->hckt || !fl->elo)) return fail(SYNTHE_EINVAL, std::string(who) + ": no Boltzmann factors");
  if (!fd->doppler && (!fd->vscale || !fl->cat_wavelength))
    return fail(SYNTHE_EINVAL, std::string(who) + ": no Doppler widths");
  hipStream_t st = (hipStream_t)stream;
  if (!ctx->f19stage) {
    ctx->f19stage = new Fort19Stage();
    ctx->f19stage_free = f19stage_free;
  }
  Fort19Stage* S = (Fort19Stage*)ctx->f19stage;
  HIP_TRY(hipStreamSynchronize(st));  // the previous call's copies have left the staging vectors
  S->lines.resize(N);
  for (long r = 0; r < N; ++r) {
    F19Line& L = S->lines[r];
    L.kind = fl->kind[r];
    if (L.kind < 0 || L.kind > 3) return fail(SYNTHE_EINVAL, std::string(who) + ": bad record kind");
    if (L.kind == 3 && (!fd->doppler || fl->center[r] < 0))
      return fail(SYNTHE_EINVAL, std::string(who) + ": a helium line needs DOPPLE (depths->doppler) and a centre on "
                                                    "the grid");
    L.elem = fl->elem[r];
    L.ion = fl->ion[r];
    if (L.elem < 0 || L.elem >= fd->n_elem || L.ion < 1 || L.ion > fd->n_ion)
      return fail(SYNTHE_ERANGE, std::string(who) + ": element or stage outside the population table (record " +
                                     std::to_string(r) + ")");
    L.c = fl->center[r];
    if (L.c < -1 || L.c >= n_wl)
      return fail(SYNTHE_ERANGE, std::string(who) + ": centre index outside [-1, n_wl) (record " + std::to_string(r) + ")");
    L.wl = fl->wavelength[r];
    L.gf = fl->gf[r];
    L.gr = fl->gamma_rad[r];
    L.gs = fl->gamma_stark[r];
    L.gw = fl->gamma_vdw[r];
    L.cont_val = fl->cont_val[r];
    L.lam_cat = fl->cat_wavelength ? fl->cat_wavelength[r] : 0.0;
    L.elo = fl->elo ? fl->elo[r] : 0.0;
    L.nelion = fl->nelion[r];
    L.nlast = fl->nlast[r];
  }
  const long P = D * fd->n_ion * fd->n_elem;
  const long nd = 9 * D + P + (fd->boltz ? N * D : 0) + (fd->doppler ? N * D : 0);
  S->depth.assign(nd, 0.0);
  double* h = S->depth.data();
  const double* srcs[9] = {fd->mass_density, fd->electron_density, fd->txnxn, fd->emerge, fd->emerge_h,
                           fd->temperature, fd->hckt, fd->vscale, nullptr};
  for (int a = 0; a < 8; ++a)
    if (srcs[a]) std::copy(srcs[a], srcs[a] + D, h + a * D);
  std::copy(fd->pop, fd->pop + P, h + 9 * D);
  long off = 9 * D + P;
  if (fd->boltz) {
    std::copy(fd->boltz, fd->boltz + N * D, h + off);
    off += N * D;
  }
  if (fd->doppler) std::copy(fd->doppler, fd->doppler + N * D, h + off);
  const size_t o_depth = align_up(N * sizeof(F19Line)), o_pairs = o_depth + align_up(nd * sizeof(double)),
               total = o_pairs + align_up((size_t)N * D * sizeof(F19Pair));
  if (S->dev_size < total) {
    if (S->dev) {
      HIP_TRY(hipFree(S->dev));
      S->dev = nullptr;
      S->dev_size = 0;
    }
    if (hipMalloc(&S->dev, total) != hipSuccess) {
      S->dev = nullptr;
      return fail(SYNTHE_ENOMEM, std::string(who) + ": staging allocation failed");
    }
    S->dev_size = total;
  }
  char* base = (char*)S->dev;
  F19Line* d_lines = (F19Line*)base;
  double* dd = (double*)(base + o_depth);
  F19Pair* d_pairs = (F19Pair*)(base + o_pairs);
  HIP_TRY(hipMemcpyAsync(d_lines, S->lines.data(), N * sizeof(F19Line), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(dd, S->depth.data(), nd * sizeof(double), hipMemcpyHostToDevice, st));
  F19DepthDev dp;
  dp.rho = dd;
  dp.xne = dd + D;
  dp.txnxn = dd + 2 * D;
  dp.emerge = dd + 3 * D;
  dp.emerge_h = dd + 4 * D;
  dp.temp = dd + 5 * D;
  dp.hckt = dd + 6 * D;
  dp.vscale = dd + 7 * D;
  dp.pop = dd + 9 * D;
  off = 9 * D + P;
  dp.boltz = fd->boltz ? dd + off : nullptr;
  if (fd->boltz) off += N * D;
  dp.doppler = fd->doppler ? dd + off : nullptr;
  dp.n_ion = fd->n_ion;
  dp.n_elem = fd->n_elem;
  hipLaunchKernelGGL(k_f19_extent, dim3((unsigned)N, (unsigned)D), dim3(64), 0, st, d_lines, dp, (int)N, (int)D,
                     n_wl, wl, cont, cutoff, ctx->tb, d_pairs);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_f19_gather, dim3((unsigned)((n_wl + 255) / 256), (unsigned)D), dim3(256), 0, st, d_lines,
                     d_pairs, (int)N, (int)D, n_wl, wl, dp.temp, stim, ctx->tb, asynth);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
namespace {

struct BatchInst {
  synthe_instance in;  // the caller's pointers and sizes
  size_t o_cabs, o_cscat, o_kapmin, o_slinec, o_asynth, o_transp, o_valid, o_asy;
  AsynthLayout alay;
  double cost;  // ordering key for the persistent gather (larger first)
};

constexpr int kBatchStreams = 4;
// Pixels per lane of the batch gather: the single-spectrum tile.  A far-wing
// sum depends on the tile (which wings have an edge in it, and the groups of 8
// pixels sharing a reciprocal), so the batch keeps the single-spectrum tile
// and its spectra are bit-identical to single ones; the autotuner picks the
// grid.
constexpr int kTilePixels[1] = {kWP};
// persistent grids of the gather candidates; 0: one workgroup per item
constexpr int kWavesPerSimd[4] = {2, 4, 8, 0};

}  // namespace

struct synthe_batch {
  synthe_ctx* ctx = nullptr;
  std::vector<BatchInst> inst;
  char* arena = nullptr;
  size_t arena_size = 0;
  double* zeros = nullptr;  // [max D*W] line scattering (rhoxj = 0)
  char* rt_base = nullptr;  // RT workspace over all instances' wavelengths
  RtLayout rlay{};
  long rt_g = 0;            // total wavelengths
  RtArgs* d_rt = nullptr;   // [n], ascending g0 (in b->order)
  GatherArgs* d_args = nullptr;
  ContSrcArgs* d_cs = nullptr;  // [n] continuum/source stage, in instance order
  long cs_blocks = 0;           // x-blocks of that stage over all spectra
  unsigned cs_rows = 0;         // its depth rows
  int32_t* d_order = nullptr;
  int64_t* d_items = nullptr;
  unsigned long long* d_counter = nullptr;
  std::vector<int32_t> order;
  int kwp = kWP;
  int gather_blocks = 4096;
  int n_cu = 256;
  int precision = SYNTHE_PREC_FP64;
  long n_items = 0;              // gather items of the current tile
  std::vector<double> tune_ms;  // [tiles][4 grids], ms, or empty
  PrepInst* d_prep = nullptr;     // [n] prepare-stage arguments, instance order
  int64_t* d_poff = nullptr;       // [kPrepStages][n + 1] block offsets of each stage
  long prep_total[kPrepStages] = {};
  hipStream_t streams[kBatchStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_mid = nullptr;
  hipEvent_t ev_join[kBatchStreams] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Items of the persistent gather for tiles of 64 * kwp pixels, instances in
// b.order; uploads the prefix table.
int batch_items(synthe_batch* b, int kwp) {
  const long tile = 64L * kwp;
  std::vector<int64_t> items(b->inst.size() + 1, 0);
  for (size_t i = 0; i < b->order.size(); ++i) {
    const synthe_instance& in = b->inst[b->order[i]].in;
    long nt = (in.n_wl + tile - 1) / tile;
    items[i + 1] = items[i] + nt * in.depths.n_depths;
  }
  if (items.back() > INT32_MAX) return fail(SYNTHE_EINVAL, "synthe_batch_create: too many gather tiles");
  b->n_items = items.back();
  HIP_TRY(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  return SYNTHE_OK;
}

// blocks > 0: persistent grid of that many workgroups; 0: one per item.
int launch_gather_batch(synthe_batch* b, int kwp, int blocks, hipStream_t st) {
  const int n = (int)b->inst.size();
  const bool f32 = b->precision == SYNTHE_PREC_FP32;
  if (blocks > 0) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, 1L, b->d_counter);
    HIP_TRY(hipGetLastError());
  } else if (b->n_items == 0) {
    return SYNTHE_OK;
  }
#define SYNTHE_GB(P, K)                                                                                       \
  if (blocks > 0)                                                                                             \
    hipLaunchKernelGGL((k_gather_batch<P, K>), dim3(blocks), dim3(64), 0, st, b->d_args, b->d_order,         \
                       b->d_items, n, b->d_counter);                                                         \
  else                                                                                                        \
    hipLaunchKernelGGL((k_gather_items<P, K>), dim3((unsigned)b->n_items), dim3(64), 0, st, b->d_args,       \
                       b->d_order, b->d_items, n)
  if (kwp != kWP) return fail(SYNTHE_EINVAL, "synthe_batch: unsupported gather tile");
  if (f32) SYNTHE_GB(kPrecFP32, kWP);
  else SYNTHE_GB(kPrecFP64, kWP);
#undef SYNTHE_GB
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

// Continuum, KAPMIN and the line source of every instance: one launch.
int batch_cont_src(synthe_batch* b, hipStream_t st) {
  if (b->cs_blocks == 0) return SYNTHE_OK;
  hipLaunchKernelGGL(k_cont_src_batch, dim3((unsigned)b->cs_blocks, b->cs_rows), dim3(kBlock), 0, st, b->d_cs,
                     (int)b->inst.size());
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

// The prepare stages of every instance, one launch per stage (k_prep_batch).
int batch_prepare(synthe_batch* b, hipStream_t st) {
  const int n = (int)b->inst.size();
#define SYNTHE_PREP(K)                                                                                      \
  if (b->prep_total[K] > 0) {                                                                               \
    hipLaunchKernelGGL(k_prep_batch<K>, dim3((unsigned)b->prep_total[K]), dim3(kBlock), 0, st, b->d_prep,   \
                       b->d_poff + (long)(K) * (n + 1), n, b->ctx->tb);                                    \
    HIP_TRY(hipGetLastError());                                                                             \
  }
  SYNTHE_PREP(0)
  SYNTHE_PREP(1)
  SYNTHE_PREP(2)
  SYNTHE_PREP(3)
  SYNTHE_PREP(4)
  SYNTHE_PREP(5)
  SYNTHE_PREP(6)
  SYNTHE_PREP(7)
#undef SYNTHE_PREP
  return SYNTHE_OK;
}

// The whole batch: each prepare stage once over all instances, one persistent
// gather over all instances, one RT over all instances' wavelengths
// (F_lambda stored directly).  Launches only.
int batch_enqueue(synthe_batch* b, hipStream_t s0, bool prepare, bool gather, bool finish) {
  const int n = (int)b->inst.size();
  int rc;
  if (prepare) {
    if ((rc = batch_cont_src(b, s0))) return rc;
    if ((rc = batch_prepare(b, s0))) return rc;
  }
  if (gather && (rc = launch_gather_batch(b, b->kwp, b->gather_blocks, s0))) return rc;
  if (finish && (rc = rt_launch_set(b->ctx, b->d_rt, n, b->rt_g, b->rt_base, b->rlay, s0))) return rc;
  return SYNTHE_OK;
}

void batch_free(synthe_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int k = 0; k < kBatchStreams; ++k) {
    if (b->streams[k]) (void)hipStreamDestroy(b->streams[k]);
    if (b->ev_join[k]) (void)hipEventDestroy(b->ev_join[k]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_mid) (void)hipEventDestroy(b->ev_mid);
  if (b->arena) (void)hipFree(b->arena);
  delete b;
}

}  // namespace

extern "C" {

int synthe_batch_create(synthe_ctx* ctx, int n, const synthe_instance* inst, int flags, synthe_batch** out) {
  if (!ctx || !out || n <= 0 || !inst) return fail(SYNTHE_EINVAL, "synthe_batch_create: bad arguments");
  *out = nullptr;
  synthe_batch* b = new synthe_batch();
  b->ctx = ctx;
  b->precision = ctx->precision;
  size_t off = 0, zeros_n = 1;
  for (int i = 0; i < n; ++i) {
    const synthe_instance& in = inst[i];
    const char* who = "synthe_batch_create";
    int rc = check_asynth_args(&in.lines, &in.depths, in.n_wl, in.wl, (const double*)1, (const uint8_t*)1,
                               (const double*)1, in.center_order, in.center_start, in.wing_order, in.flux, who);
    if (!rc && (in.depths.n_depths > kMaxDepth || in.n_wl > 0x7fffffffL))
      rc = fail(SYNTHE_ERANGE, "synthe_batch_create: instance exceeds SYNTHE_MAX_DEPTH or 2^31 wavelengths");
    if (!rc && (!in.cont || !in.wledge || !in.coeff_abs || !in.coeff_scat || in.n_edges < 2 ||
                !in.depths.rhox || !in.depths.temperature))
      rc = fail(SYNTHE_EINVAL, "synthe_batch_create: instance " + std::to_string(i) + " lacks an input or output");
    if (rc) {
      batch_free(b);
      return rc;
    }
    BatchInst bi{};
    bi.in = in;
    const long D = in.depths.n_depths, W = in.n_wl, L = in.lines.n_lines;
    const size_t dw = (size_t)D * W * sizeof(double);
    bi.o_cabs = off; off = align_up(off + dw);
    bi.o_cscat = off; off = align_up(off + dw);
    bi.o_kapmin = off; off = align_up(off + dw);
    bi.o_slinec = off; off = align_up(off + dw);
    bi.o_asynth = off; off = align_up(off + dw);
    bi.o_transp = off; off = align_up(off + (size_t)L * D * sizeof(double));
    bi.o_valid = off; off = align_up(off + (size_t)L * D);
    bi.alay = asynth_layout(L, (int)D);
    bi.o_asy = off; off = align_up(off + bi.alay.total);
    b->rt_g += W;
    bi.cost = (double)L * (double)D + (double)W * (double)D;
    zeros_n = std::max(zeros_n, (size_t)D * W);
    b->inst.push_back(bi);
  }
  const size_t o_zero = off; off = align_up(off + zeros_n * sizeof(double));
  b->rlay = rt_layout(b->rt_g, n);
  const size_t o_rt = off; off = align_up(off + b->rlay.total);
  const size_t o_args = off; off = align_up(off + n * sizeof(GatherArgs));
  const size_t o_cs = off; off = align_up(off + n * sizeof(ContSrcArgs));
  const size_t o_order = off; off = align_up(off + n * sizeof(int32_t));
  const size_t o_items = off; off = align_up(off + (n + 1) * sizeof(int64_t));
  const size_t o_counter = off; off = align_up(off + sizeof(unsigned long long));
  const size_t o_prep = off; off = align_up(off + n * sizeof(PrepInst));
  const size_t o_poff = off; off = align_up(off + (size_t)kPrepStages * (n + 1) * sizeof(int64_t));
  if (hipMalloc(&b->arena, off) != hipSuccess) {
    b->arena = nullptr;
    batch_free(b);
    return fail(SYNTHE_ENOMEM, "synthe_batch_create: arena of " + std::to_string(off) + " bytes failed");
  }
  b->arena_size = off;
  b->zeros = (double*)(b->arena + o_zero);
  b->rt_base = b->arena + o_rt;
  b->d_rt = (RtArgs*)(b->rt_base + b->rlay.o_args);
  b->d_args = (GatherArgs*)(b->arena + o_args);
  b->d_cs = (ContSrcArgs*)(b->arena + o_cs);
  b->d_order = (int32_t*)(b->arena + o_order);
  b->d_items = (int64_t*)(b->arena + o_items);
  b->d_counter = (unsigned long long*)(b->arena + o_counter);
  b->d_prep = (PrepInst*)(b->arena + o_prep);
  b->d_poff = (int64_t*)(b->arena + o_poff);
  int rc = SYNTHE_OK;
  auto bail = [&](int code) {
    batch_free(b);
    return code;
  };
  // instance order: largest estimated work first
  b->order.resize(n);
  for (int i = 0; i < n; ++i) b->order[i] = i;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return b->inst[x].cost > b->inst[y].cost; });
  std::vector<GatherArgs> args(n);
  for (int i = 0; i < n; ++i) {
    BatchInst& bi = b->inst[i];
    const synthe_instance& in = bi.in;
    args[i] = gather_args(&in.lines, &in.depths, in.n_wl, in.wl, (double*)(b->arena + bi.o_transp),
                          (uint8_t*)(b->arena + bi.o_valid), in.center_order, in.center_start,
                          (double*)(b->arena + bi.o_asynth), asynth_bufs(b->arena + bi.o_asy, bi.alay));
  }
  std::vector<PrepInst> prep(n);
  std::vector<int64_t> poff((size_t)kPrepStages * (n + 1), 0);
  for (int i = 0; i < n; ++i) {
    BatchInst& bi = b->inst[i];
    const synthe_instance& in = bi.in;
    PrepInst& P = prep[i];
    P.ln = to_dev(&in.lines);
    P.dp = to_dev(&in.depths);
    P.W = in.n_wl;
    P.L = in.lines.n_lines;
    P.D = (int)in.depths.n_depths;
    P.nblk = (P.L + kBlock - 1) / kBlock;
    P.kapmin = (const double*)(b->arena + bi.o_kapmin);
    P.wl = in.wl;
    P.cutoff = in.cutoff;
    P.transp = (double*)(b->arena + bi.o_transp);
    P.valid = (uint8_t*)(b->arena + bi.o_valid);
    P.wing_order = in.wing_order;
    P.asynth = (double*)(b->arena + bi.o_asynth);
    P.b = asynth_bufs(b->arena + bi.o_asy, bi.alay);
    for (int k = 0; k < kPrepStages; ++k)
      poff[(size_t)k * (n + 1) + i + 1] = poff[(size_t)k * (n + 1) + i] + prep_blocks(k, P);
  }
  for (int k = 0; k < kPrepStages; ++k) {
    b->prep_total[k] = poff[(size_t)k * (n + 1) + n];
    if (b->prep_total[k] > 0x7fffffffL) return bail(fail(SYNTHE_ERANGE, "synthe_batch_create: prepare grid too large"));
  }
  std::vector<ContSrcArgs> cs(n);
  for (int i = 0; i < n; ++i) {
    const BatchInst& bi = b->inst[i];
    const synthe_instance& in = bi.in;
    const int D = (int)in.depths.n_depths;
    cs[i] = ContSrcArgs{in.n_wl, b->cs_blocks, D, (int)in.n_edges, in.coeff_log10, in.wl, in.wledge, in.half_edge,
                        in.delta_edge, in.coeff_abs, in.coeff_scat, in.depths.temperature,
                        (double*)(b->arena + bi.o_cabs), (double*)(b->arena + bi.o_cscat),
                        (double*)(b->arena + bi.o_kapmin), (double*)(b->arena + bi.o_slinec)};
    b->cs_blocks += blocks_for(in.n_wl, kBlock);
    b->cs_rows = std::max(b->cs_rows, (unsigned)((D + kContinuumDepths - 1) / kContinuumDepths));
  }
  if (b->cs_blocks > 0x7fffffffL) return bail(fail(SYNTHE_ERANGE, "synthe_batch_create: too many wavelengths"));
  std::vector<RtArgs> rts(n);
  long g0 = 0;
  for (int i = 0; i < n; ++i) {
    const BatchInst& bi = b->inst[b->order[i]];
    const synthe_instance& in = bi.in;
    rts[i] = RtArgs{g0, in.n_wl, (int)in.depths.n_depths, 1, in.wl, in.depths.temperature, in.depths.rhox,
                    (const double*)(b->arena + bi.o_cabs), (const double*)(b->arena + bi.o_cscat),
                    (const double*)(b->arena + bi.o_asynth), b->zeros, (const double*)(b->arena + bi.o_slinec),
                    in.flux, in.cont, in.iters};
    g0 += in.n_wl;
  }
  if (hipMemcpy(b->d_rt, rts.data(), n * sizeof(RtArgs), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_args, args.data(), n * sizeof(GatherArgs), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_cs, cs.data(), n * sizeof(ContSrcArgs), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_order, b->order.data(), n * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_prep, prep.data(), n * sizeof(PrepInst), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_poff, poff.data(), poff.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->zeros, 0, zeros_n * sizeof(double)) != hipSuccess)
    return bail(fail(SYNTHE_EHIP, "synthe_batch_create: upload failed"));
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return bail(fail(SYNTHE_EHIP, "synthe_batch_create: device query failed"));
  b->n_cu = prop.multiProcessorCount;
  b->gather_blocks = b->n_cu * 4 * 4;
  for (int k = 0; k < kBatchStreams; ++k) {
    if (hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming) != hipSuccess)
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: stream creation failed"));
  }
  if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_mid, hipEventDisableTiming) != hipSuccess)
    return bail(fail(SYNTHE_EHIP, "synthe_batch_create: event creation failed"));
  hipStream_t s0 = b->streams[0];
  if (flags & SYNTHE_BATCH_AUTOTUNE) {
    // Time every (tile, persistent grid) candidate of the gather on this
    // batch, after the stages it depends on, and keep the fastest.
    if ((rc = batch_enqueue(b, s0, true, false, false))) return bail(rc);
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: event creation failed"));
    double best = 1e300;
    for (int a = 0; a < (int)(sizeof(kTilePixels) / sizeof(int)) && !rc; ++a) {
      if ((rc = batch_items(b, kTilePixels[a]))) break;
      for (int c = 0; c < 4 && !rc; ++c) {
        const int blocks = b->n_cu * 4 * kWavesPerSimd[c];
        float ms = 0.f, best_c = 1e30f;
        for (int rep = 0; rep < 2 && !rc; ++rep) {
          if (hipEventRecord(t0, s0) != hipSuccess) rc = fail(SYNTHE_EHIP, "event record failed");
          if (!rc) rc = launch_gather_batch(b, kTilePixels[a], blocks, s0);
          if (!rc && (hipEventRecord(t1, s0) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                      hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
            rc = fail(SYNTHE_EHIP, "synthe_batch_create: timing failed");
          best_c = std::min(best_c, ms);
        }
        b->tune_ms.push_back(best_c);
        if (!rc && best_c < best) {
          best = best_c;
          b->kwp = kTilePixels[a];
          b->gather_blocks = blocks;
        }
      }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return bail(rc);
  }
  if ((rc = batch_items(b, b->kwp))) return bail(rc);
  if (flags & SYNTHE_BATCH_GRAPH) {
    if (hipStreamSynchronize(s0) != hipSuccess)
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: synchronisation failed"));
    hipStream_t cap;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: stream creation failed"));
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cap);
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: graph capture failed to start"));
    }
    rc = batch_enqueue(b, cap, true, true, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return bail(rc);
    }
    if (e != hipSuccess || !g) return bail(fail(SYNTHE_EHIP, "synthe_batch_create: graph capture failed"));
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
      b->exec = nullptr;
      return bail(fail(SYNTHE_EHIP, "synthe_batch_create: graph instantiation failed"));
    }
  }
  if (hipStreamSynchronize(s0) != hipSuccess)
    return bail(fail(SYNTHE_EHIP, "synthe_batch_create: synchronisation failed"));
  *out = b;
  return SYNTHE_OK;
}

int synthe_batch_run(synthe_batch* b, synthe_stream stream) {
  if (!b) return fail(SYNTHE_EINVAL, "synthe_batch_run: batch is NULL");
  hipStream_t st = (hipStream_t)stream;
  if (b->exec) {
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return SYNTHE_OK;
  }
  return batch_enqueue(b, st, true, true, true);
}

int synthe_batch_info(const synthe_batch* b, int* pixels_per_lane, int* gather_blocks, double* tune_ms,
                      int tune_cap, int* n_tuned, double* arena_bytes) {
  if (!b) return fail(SYNTHE_EINVAL, "synthe_batch_info: batch is NULL");
  if (pixels_per_lane) *pixels_per_lane = b->kwp;
  if (gather_blocks) *gather_blocks = b->gather_blocks;
  if (n_tuned) *n_tuned = (int)b->tune_ms.size();
  for (int i = 0; tune_ms && i < tune_cap && i < (int)b->tune_ms.size(); ++i) tune_ms[i] = b->tune_ms[i];
  if (arena_bytes) *arena_bytes = (double)b->arena_size;
  return SYNTHE_OK;
}

int synthe_batch_arrays(const synthe_batch* b, int i, synthe_batch_view* out) {
  if (!b || !out) return fail(SYNTHE_EINVAL, "synthe_batch_arrays: NULL argument");
  if (i < 0 || i >= (int)b->inst.size()) return fail(SYNTHE_EINVAL, "synthe_batch_arrays: no such instance");
  const BatchInst& bi = b->inst[i];
  out->cont_abs = (double*)(b->arena + bi.o_cabs);
  out->cont_scat = (double*)(b->arena + bi.o_cscat);
  out->kapmin = (double*)(b->arena + bi.o_kapmin);
  out->slinec = (double*)(b->arena + bi.o_slinec);
  out->asynth = (double*)(b->arena + bi.o_asynth);
  out->transp = (double*)(b->arena + bi.o_transp);
  out->valid = (uint8_t*)(b->arena + bi.o_valid);
  out->line_scat = b->zeros;
  return SYNTHE_OK;
}

int synthe_batch_destroy(synthe_batch* b) {
  if (b) {
    (void)hipDeviceSynchronize();
    batch_free(b);
  }
  return SYNTHE_OK;
}

}  // extern "C"

// ------------------------------------------------------------ adjoint ABI
extern "C" {

int synthe_asynth_vjp(synthe_ctx* ctx, const synthe_lines* lines, const synthe_depths* depths, long n_wl,
                      const double* wl, const double* transp, const uint8_t* valid, const double* kapmin,
                      double cutoff, const int64_t* center_order, const int64_t* center_start,
                      const int64_t* wing_order, const double* g_asynth, double* s_kappa0, double* s_adamp,
                      double* s_dopple, double* kappa0, synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_asynth_vjp: ctx is NULL");
  int rc = check_asynth_args(lines, depths, n_wl, wl, transp, valid, kapmin, center_order, center_start, wing_order,
                             g_asynth, "synthe_asynth_vjp");
  if (rc) return rc;
  if (!s_kappa0 || !s_adamp || !s_dopple) return fail(SYNTHE_EINVAL, "synthe_asynth_vjp: NULL output");
  hipStream_t st = (hipStream_t)stream;
  const long L = lines->n_lines;
  const int D = (int)depths->n_depths;
  const long pairs = L * (long)D;
  const AsynthLayout lay = asynth_layout(L, D, true);
  const size_t o_u = lay.total;
  const size_t o_q = align_up(o_u + ((size_t)D * n_wl + kWT) * sizeof(double));  // u, then one tile of padding
  const size_t o_mom = align_up(o_q + (size_t)(pairs > 0 ? pairs : 1) * sizeof(double));
  const size_t total = align_up(o_mom + (size_t)D * ((n_wl + kWT - 1) / kWT) * kFarTerms * sizeof(double));
  rc = ensure_ws(ctx, total);
  if (rc) return rc;
  char* base = (char*)ctx->ws;
  const AsynthBufs b = asynth_bufs(base, lay);
  double* u = (double*)(base + o_u);
  double* q = (double*)(base + o_q);
  for (double* z : {s_kappa0, s_adamp, s_dopple, q, kappa0}) {
    if (!z) continue;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(pairs > 0 ? pairs : 1, kBlock)), dim3(kBlock), 0, st,
                       pairs > 0 ? pairs : 1L, (unsigned long long*)z);
    HIP_TRY(hipGetLastError());
  }
  if (pairs == 0) return SYNTHE_OK;
  rc = asynth_prepare(ctx, lines, depths, n_wl, wl, transp, valid, kapmin, cutoff, wing_order, nullptr, false, b,
                      st, false);
  if (rc) return rc;
  hipLaunchKernelGGL(k_adj_weights, dim3(blocks_for(n_wl, kBlock), (unsigned)D), dim3(kBlock), 0, st, D, n_wl, wl,
                     depths->temperature, g_asynth, u);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_adj_centre, dim3(blocks_for(pairs, kBlock)), dim3(kBlock), 0, st, to_dev(lines),
                     to_dev(depths), n_wl, transp, valid, u, ctx->tb, s_kappa0, s_adamp, kappa0);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_near_adj, dim3(2048, (unsigned)D), dim3(64), 0, st, n_wl, D, b.lcm, b.lp, b.list_l, b.reg,
                     ctx->tb, u, wl, s_kappa0, s_adamp, s_dopple);
  HIP_TRY(hipGetLastError());
  const GatherArgs ga = gather_args(lines, depths, n_wl, wl, transp, valid, center_order, center_start, nullptr, b);
  (void)ga;
  const long ntiles = (n_wl + kWT - 1) / kWT;
  double* mom = (double*)(base + o_mom);
  hipLaunchKernelGGL(k_adj_moments, dim3((unsigned)ntiles, (unsigned)D), dim3(64), 0, st, n_wl, u, mom);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_far_adj, dim3(2048, (unsigned)D), dim3(64), 0, st, n_wl, D, b.reg, b.lcm, u, mom, ntiles, q);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_far_finish, dim3(64, (unsigned)D), dim3(kBlock), 0, st, D, b.reg, b.lp, b.list_l, q, ctx->tb,
                     wl, n_wl, s_kappa0, s_adamp, s_dopple);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_lte_vjp(synthe_ctx* ctx, long n_depths, long n_wl, const double* wl, const double* temperature,
                   const double* rhox, const double* cont_abs, const double* cont_scat, const double* line_abs,
                   const double* line_scat, const double* line_source, const double* v_flux, double* g_cont_abs,
                   double* g_cont_scat, double* g_line_abs, double* g_planck, double* g_line_source,
                   synthe_stream stream) {
  if (!ctx) return fail(SYNTHE_EINVAL, "synthe_lte_vjp: ctx is NULL");
  if (n_depths <= 0 || n_wl < 0) return fail(SYNTHE_EINVAL, "synthe_lte_vjp: bad sizes");
  if (n_depths > kAdjMaxDepth)
    return fail(SYNTHE_ERANGE, "synthe_lte_vjp: more than " + std::to_string(kAdjMaxDepth) + " depths");
  if (n_wl == 0) return SYNTHE_OK;
  if (n_wl > 0x7fffffffL) return fail(SYNTHE_ERANGE, "synthe_lte_vjp: more than 2^31 wavelengths");
  if (!wl || !temperature || !rhox || !cont_abs || !cont_scat || !line_abs || !line_scat || !v_flux)
    return fail(SYNTHE_EINVAL, "synthe_lte_vjp: NULL input");
  const long dw = n_depths * n_wl;
  for (double* z : {g_cont_abs, g_cont_scat, g_line_abs, g_planck, g_line_source}) {
    if (!z) continue;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(dw, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, dw,
                       (unsigned long long*)z);
    HIP_TRY(hipGetLastError());
  }
  const size_t o_cnt = align_up((size_t)n_wl * sizeof(int32_t));
  const size_t o_rec = o_cnt + 256;
  int rc = ensure_ws(ctx, o_rec + vjp_store_bytes(n_wl));
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  int32_t* order = (int32_t*)ctx->ws;
  int* counts = (int*)((char*)ctx->ws + o_cnt);
  const VjpStore recs{(char*)ctx->ws + o_rec, n_wl};
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, 2, counts);
  HIP_TRY(hipGetLastError());
  const VjpArgs a{(int)n_depths, n_wl, 0, wl, temperature, rhox, cont_abs, cont_scat, line_abs, line_scat,
                  line_source, v_flux, g_cont_abs, g_cont_scat, g_line_abs, g_planck, g_line_source};
  hipLaunchKernelGGL(k_vjp_order, dim3(blocks_for(n_wl, 256)), dim3(256), 0, st, a, ctx->tb.xtau, counts, order);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_solve, dim3(blocks_for(n_wl, kVjpBlock)), dim3(kVjpBlock), 0, st, a, order, ctx->tb,
                     recs);
  HIP_TRY(hipGetLastError());
  const int seq = n_wl >= kVjpSeqMin;  // one lane per wavelength, its colours in turn
  hipLaunchKernelGGL(k_vjp_pass, dim3(blocks_for(n_wl * (seq ? 1 : 5), 64)), dim3(64), 0, st, n_wl, ctx->tb, recs,
                     seq);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_adj401, dim3(blocks_for(n_wl, 64)), dim3(64), 0, st, n_wl, recs);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_tau401, dim3(blocks_for(n_wl * (seq ? 1 : 5), 64)), dim3(64), 0, st, n_wl, recs, seq);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_finish, dim3(blocks_for(n_wl * (seq ? 1 : 4), 64)), dim3(64), 0, st, a, order, recs,
                     seq);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

int synthe_lte_vjp_batch(synthe_ctx* ctx, int n, const synthe_vjp_instance* inst, synthe_stream stream) {
  const char* who = "synthe_lte_vjp_batch";
  if (!ctx || (n > 0 && !inst)) return fail(SYNTHE_EINVAL, std::string(who) + ": NULL argument");
  if (n < 0) return fail(SYNTHE_EINVAL, std::string(who) + ": bad count");
  if (n == 0) return SYNTHE_OK;
  hipStream_t st = (hipStream_t)stream;
  std::vector<VjpArgs> args(n);
  long G = 0;
  for (int i = 0; i < n; ++i) {
    const synthe_vjp_instance& q = inst[i];
    if (q.n_depths <= 0 || q.n_wl <= 0) return fail(SYNTHE_EINVAL, std::string(who) + ": bad sizes");
    if (q.n_depths > kAdjMaxDepth)
      return fail(SYNTHE_ERANGE, std::string(who) + ": more than " + std::to_string(kAdjMaxDepth) + " depths");
    if (!q.wl || !q.temperature || !q.rhox || !q.cont_abs || !q.cont_scat || !q.line_abs || !q.line_scat ||
        !q.v_flux)
      return fail(SYNTHE_EINVAL, std::string(who) + ": NULL input");
    args[i] = VjpArgs{(int)q.n_depths, q.n_wl, G, q.wl, q.temperature, q.rhox, q.cont_abs, q.cont_scat,
                      q.line_abs, q.line_scat, q.line_source, q.v_flux, q.g_cont_abs, q.g_cont_scat, q.g_line_abs,
                      q.g_planck, q.g_line_source};
    G += q.n_wl;
    if (G > 0x7fffffffL) return fail(SYNTHE_ERANGE, std::string(who) + ": more than 2^31 wavelengths");
    const long dw = q.n_depths * q.n_wl;
    for (double* z : {q.g_cont_abs, q.g_cont_scat, q.g_line_abs, q.g_planck, q.g_line_source}) {
      if (!z) continue;
      hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(dw, kBlock)), dim3(kBlock), 0, st, dw, (unsigned long long*)z);
      HIP_TRY(hipGetLastError());
    }
  }
  const size_t o_ord = align_up((size_t)n * sizeof(VjpArgs));
  const size_t o_cnt = o_ord + align_up((size_t)G * sizeof(int32_t));
  const size_t o_rec = o_cnt + align_up((size_t)2 * n * sizeof(int));
  int rc = ensure_ws(ctx, o_rec + vjp_store_bytes(G));
  if (rc) return rc;
  char* base = (char*)ctx->ws;
  VjpArgs* d_args = (VjpArgs*)base;
  int32_t* order = (int32_t*)(base + o_ord);
  int* counts = (int*)(base + o_cnt);
  const VjpStore recs{base + o_rec, G};
  HIP_TRY(hipMemcpyAsync(d_args, args.data(), (size_t)n * sizeof(VjpArgs), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // args (host) is freed when this call returns
  hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(2L * n, 64)), dim3(64), 0, st, 2 * n, counts);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_order_batch, dim3(blocks_for(G, 256)), dim3(256), 0, st, d_args, n, G, ctx->tb.xtau,
                     counts, order);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_solve_batch, dim3(blocks_for(G, kVjpBlock)), dim3(kVjpBlock), 0, st, d_args, n, G, order,
                     ctx->tb, recs);
  HIP_TRY(hipGetLastError());
  const int seq = G >= kVjpSeqMin;  // one lane per wavelength, its colours in turn
  hipLaunchKernelGGL(k_vjp_pass, dim3(blocks_for(G * (seq ? 1 : 5), 64)), dim3(64), 0, st, G, ctx->tb, recs, seq);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_adj401, dim3(blocks_for(G, 64)), dim3(64), 0, st, G, recs);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_tau401, dim3(blocks_for(G * (seq ? 1 : 5), 64)), dim3(64), 0, st, G, recs, seq);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_finish_batch, dim3(blocks_for(G * (seq ? 1 : 4), 64)), dim3(64), 0, st, d_args, n, G,
                     order, recs, seq);
  HIP_TRY(hipGetLastError());
  return SYNTHE_OK;
}

}  // extern "C"

// Memory and undefined-behaviour check of the CPU code that the tests trust:
// the oracle (oracle/synthe_oracle.c, oracle/hydrogen_oracle.c) and the CPU builds of the device headers
// (gather_queue_check.cpp, rt_core_check.cpp, rt_adj_check.cpp).  Built with
// -fsanitize=address,undefined into one executable by
// tests/test_sanitize_cpu.py, which then runs it; any report fails the run.
// Test infrastructure: every call runs on small seeded inputs whose array
// sizes follow the oracle's own calling conventions (oracle/oracle.py).
//
//   sanitize_main <tables.bin>
// tables.bin: int64 n_tabvi, tabvi[n], tabh1[n], xtau[51], ch[51], coefj[51*51],
// then the hydrogen tables, each as int64 size + doubles, in the order of
// orc_set_hydrogen_tables (tests/test_sanitize_cpu.py writes it).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

extern "C" {
typedef struct {
  const double *h0, *h1, *h2;
  int nvoigt;
  const double *extab, *extabf;
  int nfastex;
} orc_tables;
void orc_set_threads(int n);
void orc_map1_table(const double* xold, const double* fold, int nold, const double* xnew, double* fnew, int nnew);
void orc_build_voigt_tables(const double* tabvi, const double* tabh1, int nold, double* h0, double* h1,
                            double* h2, int n);
void orc_build_fastex_tables(double* extab, double* extabf, int n);
double orc_fastex(double x, const double* extab, const double* extabf, int n);
double orc_voigt(double v, double a, const double* h0t, const double* h1t, const double* h2t, int n);
long orc_wavelength_grid(double start, double end, double resolution, double* out, long cap);
void orc_nearest_grid_indices(const double* grid, long W, const double* values, long n, int64_t* out);
void orc_wing_center_indices(const double* grid, long W, const double* values, long n, double origin,
                             int64_t* out);
void orc_continuum(long D, long W, const double* wl, long E, const double* wledge_in, const double* half_in,
                   const double* delta_in, const double* cabs, const double* cscat, int coeff_log10,
                   double* out_abs, double* out_scat);
void orc_freq_quantities(long D, long W, const double* wl, const double* temp, double* ehvkt_o, double* stim_o,
                         double* bnu_o, double* slinec_o);
void orc_transp(long L, long D, long W, const uint8_t* process, const int32_t* elem, const int32_t* ion,
                const int32_t* ltype, const double* wl, const double* gf, const double* cgf, const double* grad,
                const double* gstark, const double* gvdw, const double* elo, const int64_t* center,
                const double* hckt, const double* pop, const double* dop, long nion, long nelem,
                const double* rho, const double* xne, const double* txnxn, const double* kapmin_grid,
                double cutoff, double microturb_kms, const orc_tables* tb, double* transp, uint8_t* valid);
void orc_asynth(long L, long D, long W, const double* grid, const int32_t* ltype, const int32_t* elem,
                const int32_t* ion, const double* wl, const double* grad, const double* gstark,
                const double* gvdw, const int64_t* center_t, const int64_t* center_w, const double* transp,
                const uint8_t* valid, const double* pop, const double* dop, long nion, long nelem,
                const double* xne, const double* txnxn, const double* temp, const double* kapmin_grid,
                double cutoff, const orc_tables* tb, double* asynth, double* stats);
void orc_wing_params(long L, long D, long W, const double* grid, const int32_t* ltype, const int32_t* elem,
                     const int32_t* ion, const double* wl, const double* grad, const double* gstark,
                     const double* gvdw, const int64_t* center_w, const double* transp, const uint8_t* valid,
                     const double* pop, const double* dop, long nion, long nelem, const double* xne,
                     const double* txnxn, const double* kapmin_grid, double cutoff, const orc_tables* tb,
                     double* out_d, int64_t* out_i);
void orc_set_josh_tables(const double* xtau, const double* ch, const double* coefj);
void orc_parcoe(const double* f, const double* x, long n, double* a, double* b, double* c);
void orc_integ(const double* x, const double* f, long n, double start, double* fint);
void orc_deriv(const double* x, const double* f, long n, double* dfdx);
int orc_map1_josh(const double* xold, const double* fold, int nold, const double* xnew, double* fnew, int nnew);
void orc_set_josh_fp64(int on, double tol);
double orc_solve_josh_flux(long n, const double* acont, const double* scont, const double* aline,
                           const double* sline, const double* sigmac, const double* sigmal, const double* rhox,
                           int* info);
void orc_lte_spectrum(long D, long W, const double* wl, const double* temp, const double* rhox,
                      const double* cont_abs, const double* cont_scat, const double* line_abs,
                      const double* line_scat, const double* line_source, double* flux_total, double* flux_cont,
                      double* iters, int32_t* iters_w);
void orc_flux_to_lambda(long W, const double* wl, double* flux);
void orc_set_hydrogen_tables(const double* propbm, const double* c, const double* d, const double* pp,
                             const double* beta, const double* stalph, const double* stwtal, const double* istal,
                             const double* lnghal, const double* stcomp, const double* stcpwt, const double* lncomp,
                             const double* cut_h2p, int ncut_h2p, const double* cut_h2, int ncut_h2,
                             const double* asum_lyman, int nasum_lyman, const double* asum, int nasum,
                             const double* y1wtm, const double* xknmtb, const double* conth, int nconth,
                             const double* ehyd, const double* extab, const double* extabf, int nfastex);
void orc_hydrogen_states(long D, const double* temperature, const double* electron_density, const double* xnf_he1,
                         const double* xnf_h2, const double* xnfph, const double* dopph,
                         const double* turbulent_velocity, double microturb_kms, int has_he1_h2, double* out);
double orc_hydrogen_profile(int n, int m, const double* h, double delta_lambda_nm);
int orc_hydrogen_lines(long NH, long D, long W, const double* grid, const double* wl, const int64_t* center,
                       const double* gf, const double* elo, const int32_t* n_lower_in, const int32_t* n_upper_in,
                       const double* pop_h1, const double* dop_h1, const double* rho, const double* hckt,
                       const double* electron_density, const double* states, const double* cont,
                       const double* stim, double cutoff, double microturb_kms, double* ahline, double* stats);

// CPU builds of the device headers
int gq_roundtrip(int dc, int r, int m, long* out);
long gq_max_zone(void);
double rtc_integ_map1(int n, const double* xtau, const double* abtot, const double* rho, const double* snubar,
                      const double* alpha, double* xsbar, double* xalpha);
double adj_flux64(int n, const double* rho, const double* abtot, const double* alpha, const double* snubar,
                  const double* xtau, const double* ch, const float* coefj32, double tol, int* took401);
double adj_vjp(int n, const double* rho, const double* abtot, const double* alpha, const double* snubar,
               const double* xtau, const double* ch, const float* coefj32, double v, double* g_abtot,
               double* g_alpha, double* g_snubar);
}

namespace {
std::mt19937_64 rng(7);
double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
long pick(long a, long b) { return std::uniform_int_distribution<long>(a, b)(rng); }

std::vector<double> read_doubles(FILE* f, size_t n) {
  std::vector<double> v(n);
  if (fread(v.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "short table file\n");
    exit(2);
  }
  return v;
}

// A small spectrum through every oracle stage, sized as oracle.synthesize sizes it.
void oracle_spectrum(const orc_tables& tb, double w0, double w1, long L, long D, int seed) {
  rng.seed(seed);
  const long W = orc_wavelength_grid(w0, w1, 300000.0, nullptr, 0);
  std::vector<double> grid(W);
  orc_wavelength_grid(w0, w1, 300000.0, grid.data(), W);
  const long nion = 6, nelem = 30, E = 5;
  std::vector<double> temp(D), rhox(D), rho(D), xne(D), txnxn(D), hckt(D);
  for (long d = 0; d < D; ++d) {
    temp[d] = 3500.0 + 5000.0 * d / (D - 1 > 0 ? D - 1 : 1);
    rhox[d] = 1e-4 * pow(10.0, 5.0 * d / (D - 1 > 0 ? D - 1 : 1));
    rho[d] = 1e-9 * pow(10.0, 2.0 * d / (D - 1 > 0 ? D - 1 : 1));
    xne[d] = 1e12 * pow(10.0, 3.0 * d / (D - 1 > 0 ? D - 1 : 1));
    txnxn[d] = 1e16 * uni(0.5, 2.0);
    hckt[d] = 1.4388 / temp[d];
  }
  std::vector<double> pop(D * nion * nelem), dop(D * nion * nelem);
  for (auto& v : pop) v = pow(10.0, uni(0.0, 12.0));
  for (auto& v : dop) v = uni(1e-6, 1e-5);
  // continuum edges bracketing the window, coefficients in log10
  std::vector<double> wledge(E), cabs(D * (E - 1) * 3), cscat(D * (E - 1) * 3);
  for (long e = 0; e < E; ++e) wledge[e] = w0 - 5.0 + (w1 - w0 + 10.0) * e / (E - 1);
  for (auto& v : cabs) v = uni(-3.0, -1.0);
  for (auto& v : cscat) v = uni(-5.0, -3.0);
  std::vector<double> ca(D * W), cs(D * W), kapmin(D * W);
  orc_continuum(D, W, grid.data(), E, wledge.data(), nullptr, nullptr, cabs.data(), cscat.data(), 1, ca.data(),
                cs.data());
  for (long i = 0; i < D * W; ++i) kapmin[i] = ca[i] + cs[i];
  std::vector<double> ehvkt(D * W), stim(D * W), bnu(D * W), slinec(D * W);
  orc_freq_quantities(D, W, grid.data(), temp.data(), ehvkt.data(), stim.data(), bnu.data(), slinec.data());

  // lines: inside, at and just beyond the window's ends
  std::vector<uint8_t> process(L);
  std::vector<int32_t> elem(L), ion(L), ltype(L);
  std::vector<double> wl(L), gf(L), cgf(L), grad(L), gstark(L), gvdw(L), elo(L);
  for (long l = 0; l < L; ++l) {
    process[l] = l % 11 != 0;
    elem[l] = (int32_t)pick(-1, nelem);  // -1 and nelem are out of range: skipped
    ion[l] = (int32_t)pick(0, nion + 1);
    ltype[l] = l % 13 == 0 ? 1 : 0;
    wl[l] = l % 17 == 0 ? (l % 2 ? w1 + 0.01 : w0 - 0.01) : uni(w0, w1);
    gf[l] = pow(10.0, uni(-3.0, 1.0));
    cgf[l] = 0.026538 / 1.77245 * gf[l] * wl[l] * 1e-7 / 2.99792458e17 * 1e9;
    grad[l] = pow(10.0, uni(-6.0, -3.0));
    gstark[l] = pow(10.0, uni(-14.0, -10.0));
    gvdw[l] = pow(10.0, uni(-20.0, -16.0));
    elo[l] = uni(0.0, 40000.0);
  }
  std::vector<int64_t> center_t(L), center_w(L);
  orc_nearest_grid_indices(grid.data(), W, wl.data(), L, center_t.data());
  orc_wing_center_indices(grid.data(), W, wl.data(), L, grid[0], center_w.data());
  std::vector<double> transp(L * D);
  std::vector<uint8_t> valid(L * D);
  orc_transp(L, D, W, process.data(), elem.data(), ion.data(), ltype.data(), wl.data(), gf.data(), cgf.data(),
             grad.data(), gstark.data(), gvdw.data(), elo.data(), center_t.data(), hckt.data(), pop.data(),
             dop.data(), nion, nelem, rho.data(), xne.data(), txnxn.data(), kapmin.data(), 1e-3, 2.0, &tb,
             transp.data(), valid.data());
  std::vector<double> od(L * D * 6);
  std::vector<int64_t> oi(L * D * 3);
  orc_wing_params(L, D, W, grid.data(), ltype.data(), elem.data(), ion.data(), wl.data(), grad.data(),
                  gstark.data(), gvdw.data(), center_w.data(), transp.data(), valid.data(), pop.data(), dop.data(),
                  nion, nelem, xne.data(), txnxn.data(), kapmin.data(), 1e-3, &tb, od.data(), oi.data());
  std::vector<double> asy(D * W), stats(5);
  orc_asynth(L, D, W, grid.data(), ltype.data(), elem.data(), ion.data(), wl.data(), grad.data(), gstark.data(),
             gvdw.data(), center_t.data(), center_w.data(), transp.data(), valid.data(), pop.data(), dop.data(),
             nion, nelem, xne.data(), txnxn.data(), temp.data(), kapmin.data(), 1e-3, &tb, asy.data(),
             stats.data());
  std::vector<double> aline(D * W), sigl(D * W), ft(W), fc(W), iters(1);
  std::vector<int32_t> iw(W);
  for (long d = 0; d < D; ++d)
    for (long w = 0; w < W; ++w) {
      const double f = seed % 2 ? exp(-rhox[d] / 1e-2) : 0.0;  // line scattering on odd seeds
      aline[d * W + w] = asy[d * W + w] * (1.0 - f);
      sigl[d * W + w] = asy[d * W + w] * f;
    }
  orc_lte_spectrum(D, W, grid.data(), temp.data(), rhox.data(), ca.data(), cs.data(), aline.data(), sigl.data(),
                   seed % 3 ? slinec.data() : nullptr, ft.data(), fc.data(), iters.data(), iw.data());
  orc_flux_to_lambda(W, grid.data(), ft.data());
  orc_flux_to_lambda(W, grid.data(), fc.data());
  double s = 0.0;
  for (long w = 0; w < W; ++w) s += ft[w] + fc[w];
  printf("oracle spectrum W=%ld L=%ld D=%ld e_line=%.0f checksum=%.6e\n", W, L, D, stats[0], s);
}

void oracle_kernels(const orc_tables& tb) {
  for (int t = 0; t < 200; ++t) {
    const double v = uni(0.0, 30.0), a = t % 2 ? uni(0.0, 0.2) : uni(0.2, 3.0);
    (void)orc_voigt(v, a, tb.h0, tb.h1, tb.h2, tb.nvoigt);
    (void)orc_fastex(uni(-1.0, 1200.0), tb.extab, tb.extabf, tb.nfastex);
  }
  for (int t = 0; t < 50; ++t) {
    const long n = pick(3, 90);
    std::vector<double> x(n), f(n), a(n), b(n), c(n), fi(n), df(n);
    double x0 = 1e-6;
    for (long i = 0; i < n; ++i) {
      x0 *= uni(1.05, 2.0);
      x[i] = x0;
      f[i] = pow(10.0, uni(-2.0, 4.0));
    }
    if (t % 5 == 0)
      for (long i = 0; i < n; ++i) f[i] = 3.0;  // constant: PARCOE's c = 0 branch
    orc_parcoe(f.data(), x.data(), n, a.data(), b.data(), c.data());
    orc_integ(x.data(), f.data(), n, f[0] * x[0], fi.data());
    orc_deriv(x.data(), f.data(), n, df.data());
    const int nn = (int)pick(1, 120);
    std::vector<double> xn(nn), fn(nn);
    for (int i = 0; i < nn; ++i) xn[i] = x[0] * 0.5 + (x[n - 1] * 1.5 - x[0] * 0.5) * i / (nn > 1 ? nn - 1 : 1);
    orc_map1_table(x.data(), f.data(), (int)n, xn.data(), fn.data(), nn);
    (void)orc_map1_josh(x.data(), f.data(), (int)n, xn.data(), fn.data(), nn);
    // one JOSH column
    std::vector<double> acont(n), scont(n), aline(n), sline(n), sigc(n), sigl(n);
    for (long i = 0; i < n; ++i) {
      acont[i] = pow(10.0, uni(-3.0, 1.0));
      scont[i] = uni(1e-6, 1e-4);
      aline[i] = t % 3 ? pow(10.0, uni(-2.0, 6.0)) : 0.0;
      sline[i] = scont[i] * uni(0.5, 1.5);
      sigc[i] = acont[i] * uni(0.0, 0.5);
      sigl[i] = t % 4 ? 0.0 : aline[i] * uni(0.0, 0.5);
    }
    int info[2] = {0, 0};
    orc_set_josh_fp64(t % 2, 1e-13);
    (void)orc_solve_josh_flux(n, acont.data(), scont.data(), aline.data(), sline.data(), sigc.data(), sigl.data(),
                              x.data(), info);
  }
  orc_set_josh_fp64(0, 0.0);
}

// Hydrogen lines: states with and without the optional arrays, profiles from
// Lyman alpha to high members, and the walk over a small grid with lines on,
// at and beyond its ends.
void hydrogen(const orc_tables& tb) {
  rng.seed(5);
  const long D = 6;
  std::vector<double> T(D), xne(D), he1(D), h2(D), xnfph(2 * D), dopph(D), vt(D), states(D * 15), st2(D * 15);
  for (long d = 0; d < D; ++d) {
    T[d] = uni(4000.0, 20000.0);
    xne[d] = pow(10.0, uni(10.0, 15.0));
    he1[d] = pow(10.0, uni(12.0, 15.0));
    h2[d] = pow(10.0, uni(0.0, 12.0));
    xnfph[2 * d] = pow(10.0, uni(12.0, 16.0));
    xnfph[2 * d + 1] = pow(10.0, uni(10.0, 14.0));
    dopph[d] = uni(1e-5, 4e-5);
    vt[d] = uni(0.0, 2e5);
  }
  orc_hydrogen_states(D, T.data(), xne.data(), he1.data(), h2.data(), xnfph.data(), dopph.data(), vt.data(), 1.0,
                      1, states.data());
  orc_hydrogen_states(D, T.data(), xne.data(), nullptr, h2.data(), nullptr, nullptr, nullptr, 0.0, 1, st2.data());
  double acc = 0.0;
  const int pairs[][2] = {{1, 2}, {1, 3}, {1, 5}, {2, 3}, {2, 4}, {2, 7}, {2, 30}, {3, 4}, {3, 9}, {4, 5}, {5, 11}};
  for (const auto& nm : pairs)
    for (long d = 0; d < D; ++d)
      for (double dl : {0.0, 0.003, -0.2, 5.0, -60.0}) acc += orc_hydrogen_profile(nm[0], nm[1], &states[d * 15], dl);
  const long W = orc_wavelength_grid(480.0, 481.0, 300000.0, nullptr, 0);
  std::vector<double> grid(W);
  orc_wavelength_grid(480.0, 481.0, 300000.0, grid.data(), W);
  const double wl[] = {479.0, 480.5, 486.1, 481.0, 434.0};
  const int64_t center[] = {-1, W / 2, W, W - 1, -1};
  const double gf[] = {0.1, 0.9, 0.5, 0.05, 0.3}, elo[] = {82259.0, 82259.0, 97492.0, 82259.0, 82259.0};
  const int32_t nlo[] = {2, 2, 3, 2, 2}, nup[] = {9, 4, 12, 6, 5};
  std::vector<double> pop(D), dop(D), rho(D), hckt(D), cont(D * W), stim(D * W), ah(D * W, 0.0);
  for (long d = 0; d < D; ++d) {
    pop[d] = pow(10.0, uni(12.0, 16.0));
    dop[d] = uni(1e-5, 4e-5);
    rho[d] = pow(10.0, uni(-9.0, -7.0));
    hckt[d] = 1.4388 / T[d];
    for (long w = 0; w < W; ++w) {
      cont[d * W + w] = pow(10.0, uni(-3.0, 0.0));
      stim[d * W + w] = uni(0.5, 1.0);
    }
  }
  double stats = 0.0;
  const int rc = orc_hydrogen_lines(5, D, W, grid.data(), wl, center, gf, elo, nlo, nup, pop.data(), dop.data(),
                                    rho.data(), hckt.data(), xne.data(), states.data(), cont.data(), stim.data(),
                                    1e-3, 1.5, ah.data(), &stats);
  for (double v : ah) acc += v;
  const int32_t bad[] = {95};  // n_upper + 2 beyond the damping tables: refused
  const int rc2 = orc_hydrogen_lines(1, D, W, grid.data(), wl, center, gf, elo, nlo, bad, pop.data(), dop.data(),
                                     rho.data(), hckt.data(), xne.data(), states.data(), cont.data(), stim.data(),
                                     1e-3, 0.0, ah.data(), &stats);
  (void)tb;
  printf("hydrogen rc=%d rc_bad=%d evals=%.0f checksum=%.6e\n", rc, rc2, stats, acc);
}

void header_builds(const std::vector<double>& xtau, const std::vector<double>& ch,
                   const std::vector<double>& coefj) {
  std::vector<float> coefj32(coefj.begin(), coefj.end());
  long out[3];
  const int zones[] = {0, 1, 999999, 1000000};
  for (int z : zones)
    for (int dc : {-1000000, -1, 0, 1004095})
      if (gq_roundtrip(dc, z, z, out) != 0 || out[0] != dc || out[1] != z || out[2] != z) {
        fprintf(stderr, "queue record did not round-trip\n");
        exit(3);
      }
  (void)gq_max_zone();
  for (int t = 0; t < 60; ++t) {
    const int n = (int)pick(4, 99);
    std::vector<double> rho(n), abtot(n), snubar(n), alpha(n), xs(51), xa(51), g1(n), g2(n), g3(n);
    double r = 1e-6;
    for (int i = 0; i < n; ++i) {
      r *= uni(1.05, 1.6);
      rho[i] = r;
      abtot[i] = pow(10.0, uni(-2.0, 5.0));
      snubar[i] = pow(10.0, uni(2.0, 6.0));
      alpha[i] = pow(uni(0.0, 1.0), 3.0);
    }
    (void)rtc_integ_map1(n, xtau.data(), abtot.data(), rho.data(), snubar.data(), alpha.data(), xs.data(),
                         xa.data());
    int took401 = 0;
    (void)adj_flux64(n, rho.data(), abtot.data(), alpha.data(), snubar.data(), xtau.data(), ch.data(),
                     coefj32.data(), 1e-12, &took401);
    (void)adj_vjp(n, rho.data(), abtot.data(), alpha.data(), snubar.data(), xtau.data(), ch.data(), coefj32.data(),
                  1.0, g1.data(), g2.data(), g3.data());
  }
  printf("device headers (CPU builds) ran\n");
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s tables.bin\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  int64_t nv = 0;
  if (fread(&nv, sizeof nv, 1, f) != 1 || nv < 2 || nv > 10000) {
    fprintf(stderr, "bad table header\n");
    return 2;
  }
  const std::vector<double> tabvi = read_doubles(f, nv), tabh1 = read_doubles(f, nv);
  const std::vector<double> xtau = read_doubles(f, 51), ch = read_doubles(f, 51), coefj = read_doubles(f, 51 * 51);
  std::vector<std::vector<double>> ht;
  for (int i = 0; i < 20; ++i) {  // the hydrogen tables, each size-prefixed
    int64_t sz = 0;
    if (fread(&sz, sizeof sz, 1, f) != 1 || sz < 1 || sz > 100000) {
      fprintf(stderr, "bad hydrogen table %d\n", i);
      return 2;
    }
    ht.push_back(read_doubles(f, (size_t)sz));
  }
  fclose(f);
  orc_set_threads(2);
  std::vector<double> h0(2001), h1(2001), h2(2001), ex(1001), exf(1001);
  orc_build_voigt_tables(tabvi.data(), tabh1.data(), (int)nv, h0.data(), h1.data(), h2.data(), 2001);
  orc_build_fastex_tables(ex.data(), exf.data(), 1001);
  orc_set_josh_tables(xtau.data(), ch.data(), coefj.data());
  const orc_tables tb = {h0.data(), h1.data(), h2.data(), 2001, ex.data(), exf.data(), 1001};
  oracle_kernels(tb);
  orc_set_hydrogen_tables(ht[0].data(), ht[1].data(), ht[2].data(), ht[3].data(), ht[4].data(), ht[5].data(),
                          ht[6].data(), ht[7].data(), ht[8].data(), ht[9].data(), ht[10].data(), ht[11].data(),
                          ht[12].data(), (int)ht[12].size(), ht[13].data(), (int)ht[13].size(), ht[14].data(),
                          (int)ht[14].size(), ht[15].data(), (int)ht[15].size(), ht[16].data(), ht[17].data(),
                          ht[18].data(), (int)ht[18].size(), ht[19].data(), ex.data(), exf.data(), 1001);
  hydrogen(tb);
  oracle_spectrum(tb, 500.0, 500.4, 60, 12, 1);
  oracle_spectrum(tb, 1200.0, 1200.2, 25, 5, 2);
  oracle_spectrum(tb, 300.0, 300.05, 8, 3, 3);
  header_builds(xtau, ch, coefj);
  printf("sanitize_main: done\n");
  return 0;
}

// gpumanypaths_dd of PHCpack's PHCtoC layer (Ada/PHCtoC/gpumanypaths_dd.cpp:18-57)
// over libphcmd.so: containers in, the batched double-double tracker, the
// endpoints back into the solutions container.  The container functions are
// PHCpack's C interface, resolved from the program that loads this library.
#include <cstdio>
#include <vector>

#include "../../include/phcmd.h"
#include "../../include/phcmd_phctoc.h"

extern "C" {
// PHCpack's C interface (C/phcpack.h, C/syscon.h, C/solcon.h)
int copy_dobldobl_target_system_to_container(void);
int copy_dobldobl_start_system_to_container(void);
int syscon_number_of_symbols(int *n);
int syscon_number_of_dobldobl_polynomials(int *length);
int syscon_number_of_dobldobl_terms(int i, int *nt);
int syscon_retrieve_dobldobl_term(int i, int j, int n, int *exp, double *c);
int solcon_number_of_dobldobl_solutions(int *length);
int solcon_retrieve_next_dobldobl_initialize(void);
int solcon_retrieve_next_dobldobl_solution(int n, int *k, int *m, double *sol);
int solcon_clear_dobldobl_solutions(void);
int solcon_append_dobldobl_solution(int n, int m, double *sol);
}

namespace {

// the terms of the system in the systems container: equation index,
// exponents, coefficient limbs (real high, real low, imaginary high,
// imaginary low)
struct Terms {
  std::vector<int> eq, exps;
  std::vector<double> c[4];
};

int read_system(int dim, Terms &T, int verbose, const char *what) {
  int neq = 0;
  if (syscon_number_of_dobldobl_polynomials(&neq) != 0 || neq != dim) {
    std::fprintf(stderr, "gpumanypaths_dd: the %s system has %d polynomials in %d unknowns\n", what, neq, dim);
    return 2;
  }
  std::vector<int> d(dim);
  for (int i = 1; i <= neq; ++i) {
    int nt = 0;
    if (syscon_number_of_dobldobl_terms(i, &nt) != 0) return 3;
    if (verbose > 0) std::printf(" #terms in polynomial %d of the %s system : %d\n", i, what, nt);
    for (int j = 1; j <= nt; ++j) {
      double c[4];
      if (syscon_retrieve_dobldobl_term(i, j, dim, d.data(), c) != 0) return 3;
      T.eq.push_back(i - 1);
      T.exps.insert(T.exps.end(), d.begin(), d.end());
      for (int q = 0; q < 4; ++q) T.c[q].push_back(c[q]);
    }
  }
  return 0;
}

bool ok(int rc, const char *fn) {
  if (rc == 0) return true;
  std::fprintf(stderr, "gpumanypaths_dd: %s: %s\n", fn, phc_last_error());
  return false;
}

// device buffer freed on scope exit
struct Dev {
  void *p = nullptr;
  explicit Dev(size_t bytes) {
    if (phc_malloc(&p, bytes > 0 ? bytes : 8) != 0) p = nullptr;
  }
  ~Dev() {
    if (p) phc_free(p);
  }
};

}  // namespace

extern "C" int gpumanypaths_dd(int mode, int verbose, double regamma, double imgamma) {
  if (mode == 1) {
    std::fprintf(stderr, "gpumanypaths_dd: mode 1 (CPU tracking) is not provided by this library\n");
    return 1;
  }
  int dim = 0, npaths = 0;
  if (copy_dobldobl_target_system_to_container() != 0) return 3;
  if (syscon_number_of_symbols(&dim) != 0 || dim < 1) {
    std::fprintf(stderr, "gpumanypaths_dd: no unknowns in the symbol table\n");
    return 2;
  }
  if (solcon_number_of_dobldobl_solutions(&npaths) != 0) return 3;
  if (verbose > 0) {
    std::printf("\nAcceleration of tracking many paths ...\n");
    std::printf("gamma = %.16e + i* %.16e\n", regamma, imgamma);
    std::printf("Mode of execution : %d\n", mode);
    std::printf("number of unknowns : %d\n", dim);
    std::printf("number of solutions : %d\n", npaths);
  }
  Terms target, start;
  int rc = read_system(dim, target, verbose, "target");
  if (rc != 0) return rc;
  if (copy_dobldobl_start_system_to_container() != 0) return 3;
  rc = read_system(dim, start, verbose, "start");
  if (rc != 0) return rc;
  // start solutions: t (4 doubles), dim complex double-doubles, err, rco, res
  const size_t pn = (size_t)npaths * dim;
  std::vector<double> planes(4 * pn);
  if (solcon_retrieve_next_dobldobl_initialize() != 0) return 3;
  std::vector<double> sol(4 * (size_t)dim + 10);
  for (int p = 0; p < npaths; ++p) {
    int k = 0, m = 0;
    if (solcon_retrieve_next_dobldobl_solution(dim, &k, &m, sol.data()) != 0) return 3;
    for (int j = 0; j < dim; ++j)
      for (int q = 0; q < 4; ++q) planes[q * pn + (size_t)p * dim + j] = sol[4 + 4 * j + q];
  }
  // the program of gamma (1 - t) q + t p: component 0 the target, 1 the start
  const double weights[12] = {0.0, 0.0, 1.0, 0.0, 0.0, 0.0, regamma, imgamma, -regamma, -imgamma, 0.0, 0.0};
  const int nterms[2] = {(int)target.eq.size(), (int)start.eq.size()};
  std::vector<int> eqs(target.eq), exps(target.exps);
  eqs.insert(eqs.end(), start.eq.begin(), start.eq.end());
  exps.insert(exps.end(), start.exps.begin(), start.exps.end());
  std::vector<double> coef;
  for (int q = 0; q < 4; ++q) {
    coef.insert(coef.end(), target.c[q].begin(), target.c[q].end());
    coef.insert(coef.end(), start.c[q].begin(), start.c[q].end());
  }
  const int deg = 8;
  phc_compiled *compiled = nullptr;
  if (!ok(phc_program_compile(&compiled, dim, deg, 64, 2, weights, nterms, eqs.data(), exps.data(), 2,
                              coef.data()),
          "phc_program_compile"))
    return 3;
  phc_program prog;
  phc_tracker *tracker = nullptr;
  rc = 3;
  if (ok(phc_compiled_program(compiled, &prog), "phc_compiled_program") &&
      ok(phc_tracker_create(&tracker, &prog), "phc_tracker_create")) {
    Dev d_start(sizeof(double) * 4 * pn), d_end(sizeof(double) * 4 * pn), d_tend(sizeof(double) * npaths),
        d_status(sizeof(int) * npaths), d_nsteps(sizeof(int) * npaths), d_nrej(sizeof(int) * npaths),
        d_dx(sizeof(double) * npaths);
    phc_track_params prm;
    prm.maxsteps = 1000;
    prm.corsteps = 4;
    prm.endsteps = 8;
    prm.fabry_offset = 2;
    prm.maxstep = 0.1;
    prm.minstep = 1e-6;
    prm.pbeta = 0.5;
    prm.alpha = 1e-3;
    prm.tolcor = 1e-8;
    prm.tolend = 1e-28;
    prm.divergence = 1e8;
    prm.pade_numdeg = 5;
    prm.pade_dendeg = 1;
    prm.cbeta = 0.005;
    std::vector<double> endp(4 * pn), tend(npaths), dx(npaths);
    std::vector<int> status(npaths);
    if (d_start.p && d_end.p && d_tend.p && d_status.p && d_nsteps.p && d_nrej.p && d_dx.p &&
        (npaths == 0 ||
         (ok(phc_memcpy_htod(d_start.p, planes.data(), sizeof(double) * 4 * pn), "phc_memcpy_htod") &&
          ok(phc_track(tracker, 2, npaths, (const double *)d_start.p, 2, (double *)d_end.p, (double *)d_tend.p,
                       (int *)d_status.p, (int *)d_nsteps.p, (int *)d_nrej.p, (double *)d_dx.p, nullptr, &prm,
                       nullptr, 8, nullptr),
             "phc_track") &&
          ok(phc_synchronize(), "phc_synchronize") &&
          ok(phc_memcpy_dtoh(endp.data(), d_end.p, sizeof(double) * 4 * pn), "phc_memcpy_dtoh") &&
          ok(phc_memcpy_dtoh(tend.data(), d_tend.p, sizeof(double) * npaths), "phc_memcpy_dtoh") &&
          ok(phc_memcpy_dtoh(dx.data(), d_dx.p, sizeof(double) * npaths), "phc_memcpy_dtoh") &&
          ok(phc_memcpy_dtoh(status.data(), d_status.p, sizeof(int) * npaths), "phc_memcpy_dtoh")))) {
      rc = 0;
      if (solcon_clear_dobldobl_solutions() != 0) rc = 3;
      for (int p = 0; p < npaths && rc == 0; ++p) {
        std::fill(sol.begin(), sol.end(), 0.0);
        sol[0] = tend[p];
        for (int j = 0; j < dim; ++j)
          for (int q = 0; q < 4; ++q) sol[4 + 4 * j + q] = endp[q * pn + (size_t)p * dim + j];
        sol[4 * (size_t)dim + 4] = dx[p];
        if (solcon_append_dobldobl_solution(dim, 1, sol.data()) != 0) rc = 3;
      }
      if (verbose > 0 && rc == 0) {
        int count[6] = {0, 0, 0, 0, 0, 0};
        for (int s : status) ++count[(s >= 0 && s < 6) ? s : 0];
        std::printf("paths: %d success, %d step below minstep, %d maxsteps, %d singular, %d diverged\n",
                    count[1], count[2], count[3], count[4], count[5]);
        std::printf("writing the solutions to the container ...\n");
      }
    } else if (!(d_start.p && d_end.p && d_tend.p && d_status.p && d_nsteps.p && d_nrej.p && d_dx.p)) {
      std::fprintf(stderr, "gpumanypaths_dd: device allocation failed: %s\n", phc_last_error());
    }
  }
  if (tracker) phc_tracker_destroy(tracker);
  phc_compiled_destroy(compiled);
  return rc;
}

// In-memory systems and solutions containers with the part of PHCpack's C
// interface that gpumanypaths_dd calls (C/phcpack.h, C/syscon.h,
// C/solcon.h: same names, arguments and layouts), for testing the PHCtoC
// entry without PHCpack's Ada library.  The ctr_* functions load and read
// the containers from a test.
//
// Layouts as in PHCpack: a term's coefficient is 4 doubles (real high, real
// low, imaginary high, imaginary low); a double-double solution of n
// unknowns is 4n + 10 doubles: t (4), the n coordinates (4 each), err, rco,
// res (2 each).
#include <cstring>
#include <vector>

namespace {

struct Sys {
  int n = 0;
  std::vector<std::vector<int>> exps;      // per equation: [terms][n]
  std::vector<std::vector<double>> coefs;  // per equation: [terms][4]
};

Sys g_target, g_start, *g_current = nullptr;
int g_nsym = 0;
std::vector<std::vector<double>> g_sols;
std::vector<int> g_mult;
size_t g_next = 0;

}  // namespace

extern "C" {

// ---- set-up and read-back for tests
void ctr_set_system(int which, int n, int neq, int nterms, const int *eq, const int *exps,
                    const double *c4) {
  Sys &s = which == 0 ? g_target : g_start;
  s.n = n;
  s.exps.assign(neq, {});
  s.coefs.assign(neq, {});
  for (int t = 0; t < nterms; ++t) {
    s.exps[eq[t]].insert(s.exps[eq[t]].end(), exps + (size_t)t * n, exps + (size_t)(t + 1) * n);
    s.coefs[eq[t]].insert(s.coefs[eq[t]].end(), c4 + 4 * (size_t)t, c4 + 4 * (size_t)(t + 1));
  }
  if (which == 0) g_nsym = n;
  g_current = nullptr;
}

void ctr_set_solutions(int n, int len, const double *sols) {
  g_sols.clear();
  g_mult.clear();
  for (int k = 0; k < len; ++k) {
    g_sols.emplace_back(sols + (size_t)k * (4 * n + 10), sols + (size_t)(k + 1) * (4 * n + 10));
    g_mult.push_back(1);
  }
  g_next = 0;
}

int ctr_number_of_solutions(void) { return (int)g_sols.size(); }

void ctr_get_solutions(int n, double *sols, int *mult) {
  for (size_t k = 0; k < g_sols.size(); ++k) {
    std::memcpy(sols + k * (4 * n + 10), g_sols[k].data(), sizeof(double) * (4 * n + 10));
    mult[k] = g_mult[k];
  }
}

// ---- PHCpack's C interface
int copy_dobldobl_target_system_to_container(void) {
  g_current = &g_target;
  return 0;
}

int copy_dobldobl_start_system_to_container(void) {
  g_current = &g_start;
  return 0;
}

int syscon_number_of_symbols(int *n) {
  *n = g_nsym;
  return 0;
}

int syscon_number_of_dobldobl_polynomials(int *length) {
  if (!g_current) return 1;
  *length = (int)g_current->exps.size();
  return 0;
}

int syscon_number_of_dobldobl_terms(int i, int *nt) {
  if (!g_current || i < 1 || i > (int)g_current->exps.size()) return 1;
  *nt = (int)(g_current->coefs[i - 1].size() / 4);
  return 0;
}

int syscon_retrieve_dobldobl_term(int i, int j, int n, int *exp, double *c) {
  if (!g_current || i < 1 || i > (int)g_current->exps.size() || n != g_current->n) return 1;
  const auto &cf = g_current->coefs[i - 1];
  if (j < 1 || (size_t)j > cf.size() / 4) return 1;
  std::memcpy(exp, g_current->exps[i - 1].data() + (size_t)(j - 1) * n, sizeof(int) * n);
  std::memcpy(c, cf.data() + 4 * (size_t)(j - 1), sizeof(double) * 4);
  return 0;
}

int solcon_number_of_dobldobl_solutions(int *length) {
  *length = (int)g_sols.size();
  return 0;
}

int solcon_retrieve_next_dobldobl_initialize(void) {
  g_next = 0;
  return 0;
}

int solcon_retrieve_next_dobldobl_solution(int n, int *k, int *m, double *sol) {
  if (g_next >= g_sols.size()) {
    *k = 0;
    return 1;
  }
  if (g_sols[g_next].size() != (size_t)(4 * n + 10)) return 1;
  std::memcpy(sol, g_sols[g_next].data(), sizeof(double) * (4 * n + 10));
  *m = g_mult[g_next];
  *k = (int)++g_next;
  return 0;
}

int solcon_clear_dobldobl_solutions(void) {
  g_sols.clear();
  g_mult.clear();
  g_next = 0;
  return 0;
}

int solcon_append_dobldobl_solution(int n, int m, double *sol) {
  g_sols.emplace_back(sol, sol + 4 * n + 10);
  g_mult.push_back(m);
  return 0;
}

}  // extern "C"

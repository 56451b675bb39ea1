// Prints the convolution and addition job schedules of seeded random
// supports through the public API of ConvolutionJobs and AdditionJobs, and
// of their complex (vectorized) versions ComplexConvolutionJobs,
// ComplexIncrementJobs and ComplexAdditionJobs.
// Built twice: against this library's classes (include/phcmd_compat.h,
// libphcmd.so) and, where the reference sources are present, against the
// reference's own classes (Polynomials/convolution_jobs.h, addition_jobs.h,
// compiled into oracle/_ref); the two outputs must be identical
// (tests/test_jobs.py, with the reference's output kept as the fixture
// tests/golden/jobs_dump.txt).
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <vector>

#ifdef REFERENCE_JOBS
#include "convolution_jobs.h"
#include "addition_jobs.h"
#include "complexconv_jobs.h"
#include "complexinc_jobs.h"
#include "complexadd_jobs.h"
#else
#include "phcmd_compat.h"
#endif

namespace {

unsigned long long state = 20261202ULL;
int draw(int n) {  // uniform in 0..n-1 from a 64-bit LCG
  state = state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (int)((state >> 33) % (unsigned long long)n);
}

void dump(int dim, int nbr, std::vector<int> &nvr, std::vector<std::vector<int>> &sup) {
  std::vector<int *> idx(nbr);
  for (int k = 0; k < nbr; ++k) idx[k] = sup[k].data();
  std::cout << "support dim " << dim << " nbr " << nbr << ":";
  for (int k = 0; k < nbr; ++k) {
    std::cout << " (";
    for (int i = 0; i < nvr[k]; ++i) std::cout << (i ? "," : "") << sup[k][i];
    std::cout << ")";
  }
  std::cout << "\n";
  ConvolutionJobs cnv(dim);
  cnv.make(nbr, nvr.data(), idx.data(), false);
  std::cout << "conv dim " << cnv.get_dimension() << " count " << cnv.get_count() << " depth "
            << cnv.get_depth() << "\n";
  for (int k = 0; k < cnv.get_depth(); ++k) {
    std::cout << " layer " << k << " count " << cnv.get_layer_count(k) << "\n";
    for (int i = 0; i < cnv.get_layer_count(k); ++i) {
      const ConvolutionJob j = cnv.get_job(k, i);
      std::cout << "  " << j.get_monomial_index() << " " << j.get_first_type() << " "
                << j.get_first_input() << " " << j.get_second_type() << " "
                << j.get_second_input() << " " << j.get_output_type() << " "
                << j.get_output_index() << " | " << j << "\n";
    }
  }
  AdditionJobs add(dim, nbr);
  add.make(nbr, nvr.data(), idx.data(), false);
  std::cout << "add vars " << add.get_number_of_variables() << " mons "
            << add.get_number_of_monomials() << " count " << add.get_count() << " depth "
            << add.get_depth() << "\n";
  for (int v = 0; v < dim; ++v) {
    std::cout << " var " << v << " difcount " << add.get_differential_count(v) << " :";
    for (int i = 0; i <= add.get_differential_count(v); ++i)
      std::cout << " " << add.get_differential_index(v, i);
    std::cout << "\n";
  }
  for (int k = 0; k < add.get_depth(); ++k) {
    std::cout << " layer " << k << " count " << add.get_layer_count(k) << "\n";
    for (int i = 0; i < add.get_layer_count(k); ++i) {
      const AdditionJob j = add.get_job(k, i);
      std::cout << "  " << j.get_addition_type() << " " << j.get_increment_type() << " "
                << j.get_update_monomial() << " " << j.get_update_index() << " "
                << j.get_increment_monomial() << " " << j.get_increment_index() << " | " << j
                << "\n";
    }
  }
  ComplexConvolutionJobs ccnv(dim);
  ccnv.make(nbr, nvr.data(), idx.data(), false);
  std::cout << "cplx conv dim " << ccnv.get_dimension() << " count " << ccnv.get_count()
            << " depth " << ccnv.get_depth() << "\n";
  for (int k = 0; k < ccnv.get_depth(); ++k) {
    std::cout << " layer " << k << " count " << ccnv.get_layer_count(k) << "\n";
    for (int i = 0; i < ccnv.get_layer_count(k); ++i) {
      const ComplexConvolutionJob j = ccnv.get_job(k, i);
      std::cout << "  " << j.get_monomial_index() << " " << j.get_first_type() << " "
                << j.get_first_input() << " " << j.get_second_type() << " "
                << j.get_second_input() << " " << j.get_output_type() << " "
                << j.get_output_index() << " | " << j << "\n";
    }
  }
  if (ccnv.get_depth() <= dim) {  // the reference's increment planner keeps dim layers
    ComplexIncrementJobs cinc(ccnv, false);
    std::cout << "cplx inc dim " << cinc.get_dimension() << " count " << cinc.get_count()
              << " depth " << cinc.get_depth() << "\n";
    for (int k = 0; k < cinc.get_depth(); ++k) {
      std::cout << " layer " << k << " count " << cinc.get_layer_count(k) << "\n";
      for (int i = 0; i < cinc.get_layer_count(k); ++i) {
        const ComplexIncrementJob j = cinc.get_job(k, i);
        std::cout << "  " << j.get_monomial_index() << " " << j.get_increment_kind() << " "
                  << j.get_increment_index() << " " << j.get_isreal() << " | " << j << "\n";
      }
    }
  }
  ComplexAdditionJobs cadd(dim, nbr);
  cadd.make(nbr, nvr.data(), idx.data(), false);
  std::cout << "cplx add vars " << cadd.get_number_of_variables() << " mons "
            << cadd.get_number_of_monomials() << " count " << cadd.get_count() << " depth "
            << cadd.get_depth() << "\n";
  for (int v = 0; v < dim; ++v) {
    std::cout << " var " << v << " difcount " << cadd.get_differential_count(v) << " :";
    for (int i = 0; i <= cadd.get_differential_count(v); ++i)
      std::cout << " " << cadd.get_differential_index(v, i);
    std::cout << "\n";
  }
  for (int k = 0; k < cadd.get_depth(); ++k) {
    std::cout << " layer " << k << " count " << cadd.get_layer_count(k) << "\n";
    for (int i = 0; i < cadd.get_layer_count(k); ++i) {
      const ComplexAdditionJob j = cadd.get_job(k, i);
      std::cout << "  " << j.get_addition_type() << " " << j.get_increment_type() << " "
                << j.get_update_monomial() << " " << j.get_update_index() << " "
                << j.get_increment_monomial() << " " << j.get_increment_index() << " | " << j
                << "\n";
    }
  }
}

// nbr monomials of sorted distinct variables of 0..dim-1, sizes in lo..hi
void random_support(int dim, int nbr, int lo, int hi) {
  std::vector<int> nvr(nbr);
  std::vector<std::vector<int>> sup(nbr);
  for (int k = 0; k < nbr; ++k) {
    nvr[k] = lo + draw(hi - lo + 1);
    std::vector<int> pick(dim);
    for (int i = 0; i < dim; ++i) pick[i] = i;
    for (int i = 0; i < nvr[k]; ++i) std::swap(pick[i], pick[i + draw(dim - i)]);
    std::vector<int> s(pick.begin(), pick.begin() + nvr[k]);
    std::sort(s.begin(), s.end());
    sup[k] = s;
  }
  dump(dim, nbr, nvr, sup);
}

}  // namespace

int main() {
  // the reference's test shapes: every monomial size from 1 up to dim
  for (int dim = 1; dim <= 8; ++dim)
    for (int nbr = 1; nbr <= 9; nbr += 2) random_support(dim, nbr, 1, dim);
  // three distinct variables per monomial (the row Newton test systems)
  random_support(8, 8, 3, 3);
  random_support(16, 24, 3, 3);
  // wide and long monomials, many terms
  random_support(12, 40, 1, 12);
  random_support(20, 17, 5, 20);
  random_support(6, 64, 1, 2);
  return 0;
}

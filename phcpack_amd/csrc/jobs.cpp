// The reference's job planners for polynomial evaluation and
// differentiation, with the reference's public API (declared in
// include/phcmd_compat.h), so that a caller that builds the schedules and
// passes them to GPU_dbl2_poly_evaldiff keeps its code:
//
//   ConvolutionJob(s)  Polynomials/convolution_job.h:9-77,
//                      convolution_jobs.h, convolution_jobs.cpp:10-218
//   AdditionJob(s)     Polynomials/addition_job.h, addition_jobs.h:12-218,
//                      addition_jobs.cpp:16-458
//
// The schedules list the same jobs in the same layers and the same order as
// the reference's (callers may walk them), but the evaluation entry points
// of this library do not read them: a polynomial is evaluated from a plan
// built once (phc_poly_create), not from per-call job layers.
//
// Convolution jobs of one monomial with nvr distinct variables x[i0..]
// (Speelpenning products: forward f, backward b, cross c, coefficient cff):
//   nvr = 1   layer 0: f0 = cff*x0
//   nvr = 2   layer 0: f0 = cff*x0, b0 = cff*x1; layer 1: f1 = f0*x1
//   nvr >= 3  forward chain f0 = cff*x0, f_i = f_{i-1}*x_i at layer i;
//             backward chain b0 = x_{n-1}*x_{n-2}, b_i = b_{i-1}*x_{n-2-i} at
//             layer i, b_{n-2} = b_{n-3}*cff at layer n-2; cross products
//             c_i = f_i*b_{n-4-i} at layer max(i, n-4-i)+1 and
//             c_{n-3} = f_{n-3}*x_{n-1} at layer n-2 (for n = 3: c0 = f0*x2
//             at layer 1).
// Addition jobs sum, in a binary tree over the monomials, the values (the
// last forward product of every monomial, plus the constant) and, for
// every variable, the derivative contributions of the monomials that
// contain it (a backward, forward or cross product, or the coefficient of
// a monomial of one variable).
#include <algorithm>
#include <iostream>
#include <ostream>
#include <sstream>
#include <vector>

#include "../../include/phcmd_compat.h"

// ------------------------------------------------------- ConvolutionJob --

ConvolutionJob::ConvolutionJob(int monomialix, int input1tp, int input1ix, int input2tp,
                               int input2ix, int outputtp, int outputix)
    : monidx(monomialix), inp1tp(input1tp), inp1ix(input1ix), inp2tp(input2tp),
      inp2ix(input2ix), outptp(outputtp), outidx(outputix) {}

int ConvolutionJob::get_monomial_index(void) const { return monidx; }
int ConvolutionJob::get_first_type(void) const { return inp1tp; }
int ConvolutionJob::get_first_input(void) const { return inp1ix; }
int ConvolutionJob::get_second_type(void) const { return inp2tp; }
int ConvolutionJob::get_second_input(void) const { return inp2ix; }
int ConvolutionJob::get_output_type(void) const { return outptp; }
int ConvolutionJob::get_output_index(void) const { return outidx; }

namespace {
// names of the operand types of a convolution job (-1 coefficient,
// 0 input, 1 forward, 2 backward, 3 cross), as the reference prints them
void put_operand(std::ostream &os, int tp, int ix, const char *after) {
  static const char *const name[] = {"input", "f", "b", "c"};
  if (tp == -1) os << "cff" << after;
  else if (tp >= 0 && tp <= 3) os << name[tp] << "[" << ix << "]" << after;
}
}  // namespace

std::ostream &operator<<(std::ostream &os, const ConvolutionJob &job) {
  os << "monomial " << job.monidx << " : ";
  put_operand(os, job.inp1tp, job.inp1ix, " * ");
  put_operand(os, job.inp2tp, job.inp2ix, " to ");
  if (job.outptp >= 1 && job.outptp <= 3) put_operand(os, job.outptp, job.outidx, "");
  return os;
}

// ------------------------------------------------------ ConvolutionJobs --

ConvolutionJobs::ConvolutionJobs(int dim) : dimension(dim), jobcount(0), laydepth(0) {}

void ConvolutionJobs::add_job(int layer, const ConvolutionJob &job, bool verbose) {
  ++jobcount;
  ++freqlaycnt[layer];
  jobs[layer].push_back(job);
  if (verbose) std::cout << jobcount << " : " << job << " : layer " << layer << std::endl;
}

void ConvolutionJobs::make_monomial(int nvr, int *idx, int monidx, bool verbose) {
  // forward products f_i at layer i
  add_job(0, ConvolutionJob(monidx, -1, -1, 0, idx[0], 1, 0), verbose);
  for (int i = 1; i < nvr; ++i)
    add_job(i, ConvolutionJob(monidx, 1, i - 1, 0, idx[i], 1, i), verbose);
  if (nvr > laydepth) laydepth = nvr;
  if (nvr <= 2) return;
  // backward products b_i at layer i, the last one times the coefficient
  add_job(0, ConvolutionJob(monidx, 0, idx[nvr - 1], 0, idx[nvr - 2], 2, 0), verbose);
  for (int i = 1; i < nvr - 2; ++i)
    add_job(i, ConvolutionJob(monidx, 2, i - 1, 0, idx[nvr - 2 - i], 2, i), verbose);
  add_job(nvr - 2, ConvolutionJob(monidx, 2, nvr - 3, -1, -1, 2, nvr - 2), verbose);
  // cross products
  if (nvr == 3) {
    add_job(1, ConvolutionJob(monidx, 1, 0, 0, idx[2], 3, 0), verbose);
    return;
  }
  for (int i = 0; i < nvr - 3; ++i) {
    const int j = nvr - 4 - i;  // c_i = f_i * b_j once both exist
    add_job((i > j ? i : j) + 1, ConvolutionJob(monidx, 1, i, 2, j, 3, i), verbose);
  }
  add_job(nvr - 2, ConvolutionJob(monidx, 1, nvr - 3, 0, idx[nvr - 1], 3, nvr - 3), verbose);
}

void ConvolutionJobs::make(int nbr, int *nvr, int **idx, bool verbose) {
  freqlaycnt.assign(dimension, 0);
  jobs.assign(dimension, std::vector<ConvolutionJob>());
  for (int i = 0; i < nbr; ++i) {
    const int n = nvr[i];
    if (n == 1) {
      add_job(0, ConvolutionJob(i, 0, idx[i][0], -1, -1, 1, 0), verbose);
      if (laydepth < 1) laydepth = 1;
    } else if (n == 2) {
      add_job(0, ConvolutionJob(i, -1, -1, 0, idx[i][0], 1, 0), verbose);
      add_job(0, ConvolutionJob(i, -1, -1, 0, idx[i][1], 2, 0), verbose);
      add_job(1, ConvolutionJob(i, 1, 0, 0, idx[i][1], 1, 1), verbose);
      if (laydepth < 2) laydepth = 2;
    } else if (n > 2) {
      make_monomial(n, idx[i], i, verbose);
    }
  }
}

int ConvolutionJobs::get_dimension(void) const { return dimension; }
int ConvolutionJobs::get_count(void) const { return jobcount; }
int ConvolutionJobs::get_layer_count(int k) const {
  return (k < 0 || k >= (int)freqlaycnt.size()) ? 0 : freqlaycnt[k];
}
int ConvolutionJobs::get_depth(void) const { return laydepth; }
ConvolutionJob ConvolutionJobs::get_job(int k, int i) const { return jobs[k][i]; }
ConvolutionJobs::~ConvolutionJobs(void) {}

// ---------------------------------------------------------- AdditionJob --

AdditionJob::AdditionJob(int atp, int itp, int monidx1, int monidx2, int ix1, int ix2)
    : adtype(atp), intype(itp), updmon(monidx1), incmon(monidx2), updidx(ix1), incidx(ix2) {}

int AdditionJob::get_addition_type(void) const { return adtype; }
int AdditionJob::get_increment_type(void) const { return intype; }
int AdditionJob::get_update_monomial(void) const { return updmon; }
int AdditionJob::get_update_index(void) const { return updidx; }
int AdditionJob::get_increment_monomial(void) const { return incmon; }
int AdditionJob::get_increment_index(void) const { return incidx; }

std::ostream &operator<<(std::ostream &os, const AdditionJob &job) {
  static const char *const name[] = {"", "f", "b", "c"};
  const int a = (job.adtype >= 1 && job.adtype <= 3) ? job.adtype : 3;
  os << name[a] << "[" << job.updmon << "," << job.updidx << "] += ";
  if (job.incmon < 0) {
    // a forward update without an increment index adds the constant
    if (a == 1 && job.incidx < 0) os << "cst";
    else os << "cff[" << job.incidx << "]";
  } else {
    const int t = (job.intype == 1 || job.intype == 2) ? job.intype : 3;
    os << name[t] << "[" << job.incmon << "," << job.incidx << "]";
  }
  return os;
}

// --------------------------------------------------------- AdditionJobs --

AdditionJobs::AdditionJobs(int dim, int nbr)
    : nbrvar(dim), nbrmon(nbr), jobcount(0), laydepth(0) {}

namespace {
// the top of the addition tree over nbr terms: level floor(log2 nbr) and
// stride 2^level (nbr >= 1)
void tree_start(int nbr, int *level, int *stride) {
  int lvl = 0, pwr = 1;
  while (pwr <= nbr) {
    ++lvl;
    pwr *= 2;
  }
  *level = lvl - 1;
  *stride = pwr / 2;
}
}  // namespace

void AdditionJobs::add_job(int level, const AdditionJob &job, bool verbose) {
  if (verbose) std::cout << "adding " << job << " to layer " << level << std::endl;
  jobs[level].push_back(job);
  ++jobcount;
  ++freqlaycnt[level];
}

// The addition tree: at (level, stride, nbr) the term at position nbr
// receives the term at position nbr - stride, then the subtree of the same
// nbr and, when nbr > stride, the subtree of nbr - stride take their turn
// one level down (depth first, so jobs land in their layers in the order of
// this walk).  kind 0 sums the values of the monomials 0..nbr-1 and the
// constant; kind 1 the derivative terms of variable 0; kind 2 those of
// variable var >= 1.
void AdditionJobs::tree(int kind, int level, int stride, int nbr, int *nvr, int **idx, int var,
                        bool verbose) {
  const int ix0 = nbr - stride;
  if (kind == 0) {
    // value: monomial nbr-1 receives monomial ix0-1, or the constant
    if (ix0 >= 1)
      add_job(level, AdditionJob(1, 1, nbr - 1, ix0 - 1, nvr[nbr - 1] - 1, nvr[ix0 - 1] - 1),
              verbose);
    else if (ix0 == 0)
      add_job(level, AdditionJob(1, 1, nbr - 1, -1, nvr[nbr - 1] - 1, -1), verbose);
  } else if (kind == 1) {
    // variable 0 comes first in its monomials: backward products
    const int m1 = difidx[0][nbr];
    if (ix0 > 0) {
      const int m2 = difidx[0][ix0];
      add_job(level, AdditionJob(2, 2, m1, m2, nvr[m1] - 2, nvr[m2] - 2), verbose);
    } else if (ix0 == 0 && difidx[0][0] != -1) {
      add_job(level, AdditionJob(2, 2, m1, -1, nvr[m1] - 2, difidx[0][0]), verbose);
    }
  } else {
    // the derivative to var of monomial m: its backward product when var
    // comes first, its forward product when var comes last, else its cross
    // product at var's position - 1
    auto term = [&](int m, int *tp, int *ix) {
      if (idx[m][0] == var) { *tp = 2; *ix = nvr[m] - 2; }
      else if (idx[m][nvr[m] - 1] == var) { *tp = 1; *ix = nvr[m] - 2; }
      else { *tp = 3; *ix = position(nvr[m], idx[m], var) - 1; }
    };
    const int m1 = difidx[var][nbr];
    int t1, i1;
    term(m1, &t1, &i1);
    if (verbose)
      std::cout << "nbr : " << nbr << ", stride : " << stride << ", ix0 : " << ix0
                << ", ix1 : " << m1;
    if (ix0 > 0) {
      const int m2 = difidx[var][ix0];
      int t2, i2;
      term(m2, &t2, &i2);
      if (verbose) std::cout << ", ix2 = " << m2 << ", ix4 = " << nvr[m2] - 1 << std::endl;
      add_job(level, AdditionJob(t1, t2, m1, m2, i1, i2), verbose);
    } else if (ix0 == 0) {
      // the coefficient of a monomial x_var alone, if there is one
      const int m2 = difidx[var][0];
      if (verbose) std::cout << ", ix2 = " << m2 << std::endl;
      if (m2 != -1) add_job(level, AdditionJob(t1, 0, m1, -1, i1, m2), verbose);
    }
  }
  if (level > 0) {
    tree(kind, level - 1, stride / 2, nbr, nvr, idx, var, verbose);
    if (nbr > stride) tree(kind, level - 1, stride / 2, nbr - stride, nvr, idx, var, verbose);
  }
}

int AdditionJobs::position(int n, int *idx, int k) {
  for (int i = 0; i < n; ++i)
    if (idx[i] == k) return i;
  return -1;
}

void AdditionJobs::make(int nbr, int *nvr, int **idx, bool verbose) {
  freqlaycnt.assign(nbrmon, 0);
  // per variable: the monomials that contain it, in order; position 0 holds
  // the monomial of that variable alone (-1: none), which adds its
  // coefficient, and is not counted
  difcnt.assign(nbrvar, 0);
  difidx.assign(nbrvar, std::vector<int>());
  for (int v = 0; v < nbrvar; ++v) {
    std::vector<int> &d = difidx[v];
    d.push_back(-1);
    int seen = 0;
    for (int m = 0; m < nbr; ++m) {
      if (position(nvr[m], idx[m], v) < 0) continue;
      ++seen;
      if (nvr[m] == 1) d[0] = m;
      else d.push_back(m);
    }
    difcnt[v] = (int)d.size() - 1;
    if (verbose) std::cout << "Variable " << v << " occurs in " << seen << " monomials" << std::endl;
  }
  jobs.assign(nbrmon, std::vector<AdditionJob>());
  int level, stride;
  tree_start(nbr, &level, &stride);
  tree(0, level, stride, nbr, nvr, idx, 0, verbose);
  laydepth = level + 1;  // the value involves every monomial: the deepest tree
  for (int v = 0; v < nbrvar; ++v) {
    if (difcnt[v] <= 0) continue;
    tree_start(difcnt[v], &level, &stride);
    tree(v == 0 ? 1 : 2, level, stride, difcnt[v], nvr, idx, v, verbose);
  }
}

int AdditionJobs::get_number_of_variables(void) const { return nbrvar; }
int AdditionJobs::get_number_of_monomials(void) const { return nbrmon; }
int AdditionJobs::get_count(void) const { return jobcount; }
int AdditionJobs::get_layer_count(int k) const {
  return (k < 0 || k >= (int)freqlaycnt.size()) ? 0 : freqlaycnt[k];
}
int AdditionJobs::get_depth(void) const { return laydepth; }
int AdditionJobs::get_differential_count(int k) const {
  return (k < 0 || k >= nbrvar || k >= (int)difcnt.size()) ? 0 : difcnt[k];
}
int AdditionJobs::get_differential_index(int k, int i) const {
  if (k < 0 || k >= nbrvar || k >= (int)difcnt.size() || i < 0 || i > difcnt[k]) return -1;
  return difidx[k][i];
}
AdditionJob AdditionJobs::get_job(int k, int i) const { return jobs[k][i]; }
AdditionJobs::~AdditionJobs(void) {
  nbrvar = 0;
  nbrmon = 0;
}

// ----------------------------------------------------- complex planners --
//
// The complex planners of the vectorized evaluation (Polynomials/
// complexconv_jobs.cpp, complexinc_jobs.cpp, complexadd_jobs.cpp) follow
// the real ones job for job: each real convolution job becomes, in place,
// the four real convolutions of its real and imaginary parts, and each real
// addition job one update of the real parts followed by one of the
// imaginary parts.  They are therefore built here from the real schedules.

ComplexConvolutionJob::ComplexConvolutionJob(int monomialix, int input1tp, int input1ix,
                                             int input2tp, int input2ix, int outputtp,
                                             int outputix)
    : monidx(monomialix), inp1tp(input1tp), inp1ix(input1ix), inp2tp(input2tp),
      inp2ix(input2ix), outptp(outputtp), outidx(outputix) {}

int ComplexConvolutionJob::get_monomial_index(void) const { return monidx; }
int ComplexConvolutionJob::get_first_type(void) const { return inp1tp; }
int ComplexConvolutionJob::get_first_input(void) const { return inp1ix; }
int ComplexConvolutionJob::get_second_type(void) const { return inp2tp; }
int ComplexConvolutionJob::get_second_input(void) const { return inp2ix; }
int ComplexConvolutionJob::get_output_type(void) const { return outptp; }
int ComplexConvolutionJob::get_output_index(void) const { return outidx; }

namespace {
// names of the kinds 1 forward, 2 backward, 3 cross
const char *kind_name(int k) { return k == 1 ? "f" : k == 2 ? "b" : k == 3 ? "c" : "?"; }

// an operand of a complex convolution job: -1/-2 coefficient re/im, 0/4
// input re/im, 1-3/5-7 forward, backward, cross re/im
void put_cplx_operand(std::ostream &os, int tp, int ix, const char *after) {
  if (tp == -2) os << "cffim" << after;
  else if (tp == -1) os << "cffre" << after;
  else if (tp == 0 || tp == 4) os << "input[" << ix << "]" << (tp == 0 ? "re" : "im") << after;
  else if (tp >= 1 && tp <= 7)
    os << kind_name((tp - 1) % 4 + 1) << "[" << ix << "]" << (tp < 4 ? "re" : "im") << after;
}
}  // namespace

std::ostream &operator<<(std::ostream &os, const ComplexConvolutionJob &job) {
  static const char *const part[] = {"re^a", "re^b", "im^a", "im^b"};
  os << "monomial " << job.monidx << " : ";
  put_cplx_operand(os, job.inp1tp, job.inp1ix, " * ");
  put_cplx_operand(os, job.inp2tp, job.inp2ix, " to ");
  if (job.outptp >= 1 && job.outptp <= 12)
    os << kind_name((job.outptp - 1) % 3 + 1) << "[" << job.outidx << "]"
       << part[(job.outptp - 1) / 3];
  return os;
}

ComplexConvolutionJobs::ComplexConvolutionJobs(int dim)
    : dimension(dim), jobcount(0), laydepth(0) {}

void ComplexConvolutionJobs::make(int nbr, int *nvr, int **idx, bool verbose) {
  // operand types of the real and imaginary parts: coefficient -1/-2,
  // input 0/4, forward 1/5, backward 2/6, cross 3/7; output types of
  // re*re, im*im, re*im, im*re: kind, kind + 3, kind + 6, kind + 9
  auto re = [](int tp) { return tp; };
  auto im = [](int tp) { return tp == -1 ? -2 : tp + 4; };
  ConvolutionJobs real(dimension);
  real.make(nbr, nvr, idx, false);
  freqlaycnt.assign(dimension, 0);
  jobs.assign(dimension, std::vector<ComplexConvolutionJob>());
  jobcount = 0;
  laydepth = real.get_depth();
  for (int k = 0; k < real.get_depth(); ++k) {
    for (int i = 0; i < real.get_layer_count(k); ++i) {
      const ConvolutionJob j = real.get_job(k, i);
      const int m = j.get_monomial_index();
      int t1 = j.get_first_type(), i1 = j.get_first_input();
      int t2 = j.get_second_type(), i2 = j.get_second_input();
      // a monomial of one variable: the complex jobs take the coefficient
      // first (complexconv_jobs.cpp:315-345), the real one the input
      if (t1 == 0 && t2 == -1) {
        std::swap(t1, t2);
        std::swap(i1, i2);
      }
      const int to = j.get_output_type(), io = j.get_output_index();
      const ComplexConvolutionJob four[4] = {
          ComplexConvolutionJob(m, re(t1), i1, re(t2), i2, to, io),
          ComplexConvolutionJob(m, im(t1), i1, im(t2), i2, to + 3, io),
          ComplexConvolutionJob(m, re(t1), i1, im(t2), i2, to + 6, io),
          ComplexConvolutionJob(m, im(t1), i1, re(t2), i2, to + 9, io)};
      for (const ComplexConvolutionJob &c : four) {
        jobs[k].push_back(c);
        ++jobcount;
        if (verbose) std::cout << jobcount << " : " << c << " at layer " << k << std::endl;
      }
      freqlaycnt[k] += 4;
    }
  }
}

int ComplexConvolutionJobs::get_dimension(void) const { return dimension; }
int ComplexConvolutionJobs::get_count(void) const { return jobcount; }
int ComplexConvolutionJobs::get_layer_count(int k) const {
  return (k < 0 || k >= (int)freqlaycnt.size()) ? 0 : freqlaycnt[k];
}
int ComplexConvolutionJobs::get_depth(void) const { return laydepth; }
ComplexConvolutionJob ComplexConvolutionJobs::get_job(int k, int i) const { return jobs[k][i]; }
ComplexConvolutionJobs::~ComplexConvolutionJobs(void) {}

ComplexIncrementJob::ComplexIncrementJob(int monomix, int kind, int ix, bool isreal)
    : monidx(monomix), inkind(kind), incidx(ix), rornot(isreal) {}
int ComplexIncrementJob::get_monomial_index(void) const { return monidx; }
int ComplexIncrementJob::get_increment_kind(void) const { return inkind; }
int ComplexIncrementJob::get_increment_index(void) const { return incidx; }
bool ComplexIncrementJob::get_isreal(void) const { return rornot; }

std::ostream &operator<<(std::ostream &os, const ComplexIncrementJob &job) {
  os << "monomial " << job.monidx << " : ";
  if (job.inkind >= 1 && job.inkind <= 3) os << kind_name(job.inkind) << "[" << job.incidx << "]";
  os << (job.rornot ? "re^a += " : "im^a += ");
  if (job.inkind >= 1 && job.inkind <= 3) os << kind_name(job.inkind) << "[" << job.incidx << "]";
  os << (job.rornot ? "re^b" : "im^b");
  return os;
}

// one increment per real part (output types 1-3, the re*re convolution
// takes the -im*im one) and per imaginary part (types 7-9, plus im*re),
// in the layer of the convolution (complexinc_jobs.cpp:10-60); the layers
// are as many as the dimension
ComplexIncrementJobs::ComplexIncrementJobs(ComplexConvolutionJobs cnvjobs, bool verbose)
    : dimension(cnvjobs.get_dimension()), jobcount(0), laydepth(cnvjobs.get_dimension()) {
  const int layers = std::max(dimension, cnvjobs.get_depth());
  freqlaycnt.assign(layers, 0);
  jobs.assign(layers, std::vector<ComplexIncrementJob>());
  for (int k = 0; k < cnvjobs.get_depth(); ++k)
    for (int i = 0; i < cnvjobs.get_layer_count(k); ++i) {
      const ComplexConvolutionJob c = cnvjobs.get_job(k, i);
      const int tp = c.get_output_type();
      if (tp < 4 || (tp > 6 && tp < 10)) {
        const ComplexIncrementJob job = make_job(c);
        jobs[k].push_back(job);
        ++jobcount;
        ++freqlaycnt[k];
        if (verbose) std::cout << jobcount << " : " << job << " at layer " << k << std::endl;
      }
    }
}

ComplexIncrementJob ComplexIncrementJobs::make_job(ComplexConvolutionJob cnvjob) const {
  const int tp = cnvjob.get_output_type();
  return tp < 4 ? ComplexIncrementJob(cnvjob.get_monomial_index(), tp, cnvjob.get_output_index(),
                                      true)
                : ComplexIncrementJob(cnvjob.get_monomial_index(), tp - 6,
                                      cnvjob.get_output_index(), false);
}

int ComplexIncrementJobs::get_dimension(void) const { return dimension; }
int ComplexIncrementJobs::get_count(void) const { return jobcount; }
int ComplexIncrementJobs::get_layer_count(int k) const {
  return (k < 0 || k >= dimension || k >= (int)freqlaycnt.size()) ? 0 : freqlaycnt[k];
}
int ComplexIncrementJobs::get_depth(void) const { return laydepth; }
ComplexIncrementJob ComplexIncrementJobs::get_job(int k, int i) const { return jobs[k][i]; }
ComplexIncrementJobs::~ComplexIncrementJobs(void) {}

ComplexAdditionJob::ComplexAdditionJob(int atp, int itp, int monidx1, int monidx2, int ix1,
                                       int ix2)
    : adtype(atp), intype(itp), updmon(monidx1), incmon(monidx2), updidx(ix1), incidx(ix2) {}
int ComplexAdditionJob::get_addition_type(void) const { return adtype; }
int ComplexAdditionJob::get_increment_type(void) const { return intype; }
int ComplexAdditionJob::get_update_monomial(void) const { return updmon; }
int ComplexAdditionJob::get_update_index(void) const { return updidx; }
int ComplexAdditionJob::get_increment_monomial(void) const { return incmon; }
int ComplexAdditionJob::get_increment_index(void) const { return incidx; }

std::ostream &operator<<(std::ostream &os, const ComplexAdditionJob &job) {
  if (job.adtype < 1 || job.adtype > 6) return os << "invalid operation";
  const bool real = job.adtype <= 3;
  const int kind = (job.adtype - 1) % 3 + 1;
  const char *part = real ? "re" : "im";
  // the real updates print their target with the increment, the imaginary
  // ones before it (an invalid increment type then follows the target)
  std::ostringstream target;
  target << kind_name(kind) << "[" << job.updmon << "," << job.updidx << "]" << part;
  if (!real) os << target.str() << " += ";
  if (job.incmon < 0) {
    if (real) os << target.str() << " += ";
    if (kind == 1 && job.incidx < 0) os << "cst" << part;
    else os << "cff[" << job.incidx << "]" << part;
    return os;
  }
  const int base = real ? 1 : 7;  // increment types base.. base+5
  const int t = job.intype - base;
  if (t < 0 || t > 5) return os << "invalid operation";
  if (real) os << target.str() << (t < 3 ? " += " : " -= ");
  return os << kind_name(t % 3 + 1) << "[" << job.incmon << "," << job.incidx << "]" << part
            << (t < 3 ? "^a" : "^b");
}

ComplexAdditionJobs::ComplexAdditionJobs(int dim, int nbr)
    : nbrvar(dim), nbrmon(nbr), jobcount(0), laydepth(0) {}

void ComplexAdditionJobs::make(int nbr, int *nvr, int **idx, bool verbose) {
  AdditionJobs real(nbrvar, nbrmon);
  real.make(nbr, nvr, idx, false);
  freqlaycnt.assign(nbrmon, 0);
  jobs.assign(nbrmon, std::vector<ComplexAdditionJob>());
  difcnt.assign(nbrvar, 0);
  difidx.assign(nbrvar, std::vector<int>());
  for (int v = 0; v < nbrvar; ++v) {
    difcnt[v] = real.get_differential_count(v);
    for (int i = 0; i <= difcnt[v]; ++i) difidx[v].push_back(real.get_differential_index(v, i));
  }
  jobcount = 0;
  laydepth = real.get_depth();
  for (int k = 0; k < real.get_depth(); ++k)
    for (int i = 0; i < real.get_layer_count(k); ++i) {
      const AdditionJob j = real.get_job(k, i);
      const int at = j.get_addition_type(), it = j.get_increment_type();
      const int um = j.get_update_monomial(), ui = j.get_update_index();
      int im = j.get_increment_monomial(), ii = j.get_increment_index();
      const ComplexAdditionJob first(at, it, um, im, ui, ii);
      // the imaginary update: the constant's imaginary part is index -2,
      // a coefficient's imaginary part monomial -2 (complexadd_jobs.cpp:56-63,
      // 108-115, 345-387)
      if (im == -1 && ii == -1) ii = -2;
      else if (im == -1) im = -2;
      const ComplexAdditionJob second(at + 3, it == 0 ? 0 : it + 6, um, im, ui, ii);
      for (const ComplexAdditionJob *c : {&first, &second}) {
        jobs[k].push_back(*c);
        ++jobcount;
        if (verbose) std::cout << "adding " << *c << " at layer " << k << std::endl;
      }
      freqlaycnt[k] += 2;
    }
}

int ComplexAdditionJobs::get_number_of_variables(void) const { return nbrvar; }
int ComplexAdditionJobs::get_number_of_monomials(void) const { return nbrmon; }
int ComplexAdditionJobs::get_count(void) const { return jobcount; }
int ComplexAdditionJobs::get_layer_count(int k) const {
  return (k < 0 || k >= (int)freqlaycnt.size()) ? 0 : freqlaycnt[k];
}
int ComplexAdditionJobs::get_depth(void) const { return laydepth; }
int ComplexAdditionJobs::get_differential_count(int k) const {
  return (k < 0 || k >= nbrvar || k >= (int)difcnt.size()) ? 0 : difcnt[k];
}
int ComplexAdditionJobs::get_differential_index(int k, int i) const {
  if (k < 0 || k >= nbrvar || k >= (int)difcnt.size() || i < 0 || i > difcnt[k]) return -1;
  return difidx[k][i];
}
int ComplexAdditionJobs::position(int n, int *idx, int k) {
  for (int i = 0; i < n; ++i)
    if (idx[i] == k) return i;
  return -1;
}
ComplexAdditionJob ComplexAdditionJobs::get_job(int k, int i) const { return jobs[k][i]; }
ComplexAdditionJobs::~ComplexAdditionJobs(void) {}

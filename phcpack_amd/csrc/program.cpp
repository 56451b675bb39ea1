// Host compiler of a homotopy H(x,t) = sum_r w_r(t) P_r(x), given term by
// term, into the flat tracker program (phc_program, include/phcmd.h).
//
// It builds the same arrays as compile_homotopy of phcpack_amd/homotopy.py
// (the Python host path), so that callers without Python - the
// gpumanypaths_dd entry of the PHCtoC layer, C and C++ programs - track the
// same program:
//   * the monomials of every (component, equation) are merged (coefficients
//     added in the order the terms come);
//   * product nodes x^e = x^(e - e_j) * x_j, j the last variable of e, for
//     the monomials, their first derivatives and (equations of degree > 2)
//     their second derivatives, ordered by DAG level;
//   * factor nodes keep the whole series, the others a ring of the last
//     orders the weights reach;
//   * terms segmented per equation (G segments, the terms of one monomial
//     adjacent), Jacobian and Hessian term lists, the largest singular
//     value of the constant Hessian of every quadratic equation.
// Coefficients may come in double-double (CL = 2): the low part of a merged
// coefficient becomes a second term of the same monomial right after the
// high part, so that the tracker, whose terms carry complex doubles, sums
// c_hi * u + c_lo * u in its own precision.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <string>
#include <vector>

#include "../../include/phcmd.h"
#include "common.hpp"

namespace {

typedef std::vector<int> Exp;
typedef std::complex<double> cd;

struct Coef {
  double re[2], im[2];  // high and low parts
};

// error-free sum of doubles and the double-double sum (QDlib's
// dd_add: two-sum of the high parts, low parts added, renormalized)
inline void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
inline void dd_add(const double *a, const double *b, double *c) {
  double s, e;
  two_sum(a[0], b[0], s, e);
  e += a[1] + b[1];
  const double hi = s + e;
  c[1] = e - (hi - s);
  c[0] = hi;
}

// largest singular value of the complex n x n matrix H (row major): one-sided
// Jacobi (Hestenes) on the real 2n x 2n matrix [Re -Im; Im Re], whose
// singular values are those of H, each twice; sweeps until no column pair
// has a cosine above 1e-15
double sigma_max(const std::vector<cd> &H, int n) {
  const int m = 2 * n;
  std::vector<double> M((size_t)m * m);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const cd h = H[(size_t)i * n + j];
      M[(size_t)i * m + j] = h.real();
      M[(size_t)i * m + n + j] = -h.imag();
      M[(size_t)(n + i) * m + j] = h.imag();
      M[(size_t)(n + i) * m + n + j] = h.real();
    }
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = 0; i < m; ++i) {
          const double x = M[(size_t)i * m + p], y = M[(size_t)i * m + q];
          a += x * x;
          b += y * y;
          g += x * y;
        }
        if (g == 0.0 || std::fabs(g) <= 1e-15 * std::sqrt(a * b)) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; ++i) {
          const double x = M[(size_t)i * m + p], y = M[(size_t)i * m + q];
          M[(size_t)i * m + p] = c * x - s * y;
          M[(size_t)i * m + q] = s * x + c * y;
        }
      }
    if (!rotated) break;
  }
  double best = 0.0;
  for (int j = 0; j < m; ++j) {
    double a = 0.0;
    for (int i = 0; i < m; ++i) a += M[(size_t)i * m + j] * M[(size_t)i * m + j];
    best = std::max(best, std::sqrt(a));
  }
  return best;
}

// (a, b, multiplicity) of the nonzero second derivatives of x^e, a <= b
struct HPair {
  int a, b, mult;
};
std::vector<HPair> hessian_pairs(const Exp &e) {
  std::vector<HPair> out;
  const int n = (int)e.size();
  for (int a = 0; a < n; ++a)
    for (int b = a; b < n; ++b) {
      if (a == b && e[a] >= 2) out.push_back({a, b, e[a] * (e[a] - 1)});
      else if (a != b && e[a] >= 1 && e[b] >= 1) out.push_back({a, b, e[a] * e[b]});
    }
  return out;
}

struct CTerm {
  int r, i;
  Exp e;
  Coef c;
};

}  // namespace

struct phc_compiled {
  int n = 0, deg = 0, ncomp = 0, wdeg = 0, G = 1, nprod = 0, nlev = 0, nslots = 0, nfactor = 0;
  std::vector<int> nodes, lev_ptr, pairs, seg_ptr, term_pair, jpairs, jent_ptr, jterms, hent_ptr,
      hterms, hpairs;
  std::vector<double> term_coef, weights, hsig;
};

namespace {

struct Compiler {
  int n;
  std::map<Exp, int> node_of;
  std::vector<Exp> node_exps;
  std::vector<int> left, right;
  std::vector<int> level;

  int ensure(const Exp &e) {
    auto it = node_of.find(e);
    if (it != node_of.end()) return it->second;
    int j = -1;
    for (int v = 0; v < n; ++v)
      if (e[v] > 0) j = v;
    Exp rest = e;
    rest[j] -= 1;
    const int a = ensure(rest);
    const int idx = (int)node_exps.size();
    node_of[e] = idx;
    node_exps.push_back(e);
    left.push_back(a);
    right.push_back(j);
    level.push_back(level[a] + 1);
    return idx;
  }
};

int compile(phc_compiled &P, int n, int deg, int lanes, int ncomp, const double *weights,
            const int *nterms, const int *eqs, const int *exps, int CL, const double *coef) {
  const int D = deg + 1;
  Compiler C;
  C.n = n;
  for (int i = 0; i < n; ++i) {
    Exp e(n, 0);
    e[i] = 1;
    C.node_of[e] = i;
    C.node_exps.push_back(e);
  }
  const Exp one(n, 0);
  C.node_of[one] = n;
  C.node_exps.push_back(one);
  C.level.assign(n + 1, 0);
  // merged terms per (component, equation, monomial), in the order of first
  // appearance; coefficients added in the coefficient's precision
  int total = 0;
  for (int r = 0; r < ncomp; ++r) total += nterms[r];
  std::vector<CTerm> cterms;
  {
    int t0 = 0;
    for (int r = 0; r < ncomp; ++r) {
      for (int i = 0; i < n; ++i) {
        std::map<Exp, int> pos;
        std::vector<CTerm> merged;
        for (int t = t0; t < t0 + nterms[r]; ++t) {
          if (eqs[t] != i) continue;
          Exp e(exps + (size_t)t * n, exps + (size_t)(t + 1) * n);
          Coef c{};
          c.re[0] = coef[(size_t)0 * total + t];
          c.im[0] = coef[(size_t)CL * total + t];
          if (CL == 2) {
            c.re[1] = coef[(size_t)1 * total + t];
            c.im[1] = coef[(size_t)3 * total + t];
          }
          auto it = pos.find(e);
          if (it == pos.end()) {
            // 0 + c: exact in either precision, and a zero of either sign
            // becomes +0
            for (int q = 0; q < 2; ++q) {
              c.re[q] = 0.0 + c.re[q];
              c.im[q] = 0.0 + c.im[q];
            }
            pos[e] = (int)merged.size();
            merged.push_back({r, i, e, c});
          } else {
            Coef &m = merged[it->second].c;
            if (CL == 2) {
              dd_add(m.re, c.re, m.re);
              dd_add(m.im, c.im, m.im);
            } else {
              m.re[0] += c.re[0];
              m.im[0] += c.im[0];
            }
          }
        }
        for (auto &m : merged) cterms.push_back(m);
      }
      t0 += nterms[r];
    }
  }
  auto nonzero = [](const Coef &c) { return c.re[0] != 0.0 || c.im[0] != 0.0 || c.re[1] != 0.0 || c.im[1] != 0.0; };
  auto has_lo = [](const Coef &c) { return c.re[1] != 0.0 || c.im[1] != 0.0; };
  auto tdeg = [](const Exp &e) {
    int s = 0;
    for (int v : e) s += v;
    return s;
  };
  std::vector<int> pdg(n, 0);
  for (auto &ct : cterms)
    if (nonzero(ct.c)) pdg[ct.i] = std::max(pdg[ct.i], tdeg(ct.e));
  {
    std::vector<Exp> monos;
    for (auto &ct : cterms) monos.push_back(ct.e);
    std::sort(monos.begin(), monos.end(), [&](const Exp &a, const Exp &b) {
      const int da = tdeg(a), db = tdeg(b);
      return da != db ? da < db : a < b;
    });
    monos.erase(std::unique(monos.begin(), monos.end()), monos.end());
    for (auto &e : monos) {
      C.ensure(e);
      for (int j = 0; j < n; ++j)
        if (e[j] > 0) {
          Exp d = e;
          d[j] -= 1;
          C.ensure(d);
        }
    }
  }
  for (auto &ct : cterms) {
    if (pdg[ct.i] <= 2) continue;
    for (auto &hp : hessian_pairs(ct.e)) {
      Exp d = ct.e;
      d[hp.a] -= 1;
      d[hp.b] -= 1;
      C.ensure(d);
    }
  }
  const int nprod = (int)C.left.size();
  std::vector<int> order(nprod);
  for (int i = 0; i < nprod; ++i) order[i] = n + 1 + i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return C.level[a] < C.level[b]; });
  std::vector<int> remap(n + 1 + nprod);
  for (int i = 0; i <= n; ++i) remap[i] = i;
  for (int k = 0; k < nprod; ++k) remap[order[k]] = n + 1 + k;
  std::vector<int> L(nprod), Rt(nprod), levs(nprod);
  std::vector<Exp> new_exps(n + 1 + nprod);
  for (int i = 0; i <= n; ++i) new_exps[i] = C.node_exps[i];
  for (int k = 0; k < nprod; ++k) {
    const int old = order[k];
    L[k] = remap[C.left[old - n - 1]];
    Rt[k] = remap[C.right[old - n - 1]];
    levs[k] = C.level[old];
    new_exps[remap[old]] = C.node_exps[old];
  }
  std::map<Exp, int> node_of;
  for (int i = 0; i < (int)new_exps.size(); ++i) node_of[new_exps[i]] = i;
  const int nlev = nprod ? *std::max_element(levs.begin(), levs.end()) : 0;
  P.lev_ptr.assign(nlev + 1, 0);
  for (int lv = 1; lv <= nlev; ++lv)
    P.lev_ptr[lv] = P.lev_ptr[lv - 1] + (int)std::count(levs.begin(), levs.end(), lv);
  // weight degree and the ring of the non-factor nodes
  int wdeg = 0;
  for (int r = 0; r < ncomp; ++r)
    for (int l = 0; l < 3; ++l)
      if (std::abs(cd(weights[(r * 3 + l) * 2], weights[(r * 3 + l) * 2 + 1])) > 0) wdeg = std::max(wdeg, l);
  int ring = 1;
  while (ring < std::min(wdeg + 1, D)) ring *= 2;
  const int nn = n + 1 + nprod;
  std::vector<char> factor(nn, 0);
  for (int i = 0; i <= n; ++i) factor[i] = 1;
  for (int a : L) factor[a] = 1;
  std::vector<long> off(nn);
  long acc = 0;
  for (int i = 0; i < nn; ++i) {
    off[i] = acc;
    acc += factor[i] ? D : ring;
  }
  PHC_REQUIRE(acc < (1L << 30), "phc_program_compile: too many series slots");
  std::vector<int> mask(nn), dn(nn, -1);
  for (int i = 0; i < nn; ++i) mask[i] = factor[i] ? -1 : ring - 1;
  int nfactor = 0;
  for (int i = n + 1; i < nn; ++i)
    if (factor[i]) dn[i] = nfactor++;
  P.nodes.resize((size_t)8 * nprod);
  for (int i = 0; i < nprod; ++i) {
    const int a = L[i], b = Rt[i], s = n + 1 + i;
    const int row[8] = {(int)off[a], mask[a], (int)off[b], (int)off[s], mask[s], a > n ? dn[a] : -1, dn[s], a};
    std::copy(row, row + 8, P.nodes.begin() + (size_t)8 * i);
  }
  // pairs (component, node) and the terms of every equation; a coefficient
  // with a low part gives two terms
  struct ETerm {
    int pair;
    double re, im;
    int ct;  // index into cterms
  };
  std::map<std::pair<int, int>, int> pair_of;
  std::vector<std::vector<ETerm>> eq_terms(n);
  for (int k = 0; k < (int)cterms.size(); ++k) {
    const CTerm &ct = cterms[k];
    const int m = node_of[ct.e];
    const auto key = std::make_pair(ct.r, m);
    if (!pair_of.count(key)) {
      pair_of[key] = (int)(P.pairs.size() / 4);
      P.pairs.insert(P.pairs.end(), {ct.r, (int)off[m], mask[m], m});
    }
    const int q = pair_of[key];
    eq_terms[ct.i].push_back({q, ct.c.re[0], ct.c.im[0], k});
    if (has_lo(ct.c)) eq_terms[ct.i].push_back({q, ct.c.re[1], ct.c.im[1], k});
  }
  int G = 1;
  while (n * G * 2 <= lanes) G *= 2;
  std::vector<std::vector<int>> term_of(cterms.size());  // term indices of a merged term
  std::vector<double> cre, cim;
  P.seg_ptr.assign(1, 0);
  for (int i = 0; i < n; ++i) {
    std::vector<ETerm> ts = eq_terms[i];
    std::stable_sort(ts.begin(), ts.end(), [&](const ETerm &a, const ETerm &b) {
      return node_of[cterms[a.ct].e] < node_of[cterms[b.ct].e];
    });
    std::vector<std::pair<int, int>> groups;  // [begin, end) in ts
    for (int k = 0; k < (int)ts.size(); ++k) {
      if (!groups.empty() && node_of[cterms[ts[groups.back().first].ct].e] == node_of[cterms[ts[k].ct].e])
        groups.back().second = k + 1;
      else
        groups.push_back({k, k + 1});
    }
    // G chunks of consecutive groups, the first (#groups mod G) one longer
    const int ng = (int)groups.size(), each = ng / G, extra = ng % G;
    int g0 = 0;
    for (int s = 0; s < G; ++s) {
      const int len = each + (s < extra ? 1 : 0);
      for (int g = g0; g < g0 + len; ++g)
        for (int k = groups[g].first; k < groups[g].second; ++k) {
          term_of[ts[k].ct].push_back((int)P.term_pair.size());
          P.term_pair.push_back(ts[k].pair);
          cre.push_back(ts[k].re);
          cim.push_back(ts[k].im);
        }
      g0 += len;
      P.seg_ptr.push_back((int)P.term_pair.size());
    }
  }
  const int nterm = (int)P.term_pair.size();
  P.term_coef.resize((size_t)2 * nterm);
  std::copy(cre.begin(), cre.end(), P.term_coef.begin());
  std::copy(cim.begin(), cim.end(), P.term_coef.begin() + nterm);
  // Jacobian entry (i, j): coef[term] * e_j * W_r0 * node(x^e / x_j)_0
  std::map<std::pair<int, int>, int> jpair_of;
  std::vector<std::vector<unsigned>> ent((size_t)n * n);
  for (int k = 0; k < (int)cterms.size(); ++k) {
    const CTerm &ct = cterms[k];
    for (int j = 0; j < n; ++j) {
      if (ct.e[j] == 0) continue;
      Exp d = ct.e;
      d[j] -= 1;
      const int m = node_of[d];
      const auto key = std::make_pair(ct.r, m);
      if (!jpair_of.count(key)) {
        jpair_of[key] = (int)(P.jpairs.size() / 2);
        P.jpairs.insert(P.jpairs.end(), {ct.r, (int)off[m]});
      }
      const unsigned q = (unsigned)jpair_of[key], ex = (unsigned)ct.e[j];
      for (int t : term_of[k]) {
        PHC_REQUIRE(t < 65536 && q < 4096 && ex < 16,
                    "phc_program_compile: more than 65536 terms, 4096 Jacobian pairs or an exponent above 15");
        ent[(size_t)ct.i * n + j].push_back((unsigned)t | (q << 16) | (ex << 28));
      }
    }
  }
  P.jent_ptr.assign(1, 0);
  for (auto &lst : ent) {
    for (unsigned v : lst) P.jterms.push_back((int)v);
    P.jent_ptr.push_back((int)P.jterms.size());
  }
  // Hessians: the largest singular value at t = 0 of the constant Hessian
  // of a quadratic equation, term lists for equations of degree > 2
  const int nh = n * (n + 1) / 2;
  P.hsig.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    if (pdg[i] <= 1) continue;
    if (pdg[i] > 2) {
      P.hsig[i] = -1.0;
      continue;
    }
    std::vector<cd> Hm((size_t)n * n, cd(0.0, 0.0));
    for (int r = 0; r < ncomp; ++r) {
      const cd w0(weights[r * 6], weights[r * 6 + 1]);
      for (auto &ct : cterms) {
        if (ct.r != r || ct.i != i) continue;
        const cd c(ct.c.re[0], ct.c.im[0]);
        for (auto &hp : hessian_pairs(ct.e)) {
          const cd v = w0 * c * cd((double)hp.mult, 0.0);
          Hm[(size_t)hp.a * n + hp.b] += v;
          if (hp.a != hp.b) Hm[(size_t)hp.b * n + hp.a] += v;
        }
      }
    }
    P.hsig[i] = sigma_max(Hm, n);
  }
  std::map<std::pair<int, int>, int> hpair_of;
  std::vector<std::vector<int>> hent((size_t)n * nh);
  for (int k = 0; k < (int)cterms.size(); ++k) {
    const CTerm &ct = cterms[k];
    if (pdg[ct.i] <= 2) continue;
    for (auto &hp : hessian_pairs(ct.e)) {
      Exp d = ct.e;
      d[hp.a] -= 1;
      d[hp.b] -= 1;
      const int m = node_of[d];
      const auto key = std::make_pair(ct.r, m);
      if (!hpair_of.count(key)) {
        hpair_of[key] = (int)(P.hpairs.size() / 2);
        P.hpairs.insert(P.hpairs.end(), {ct.r, (int)off[m]});
      }
      auto &lst = hent[(size_t)ct.i * nh + hp.a * n - hp.a * (hp.a - 1) / 2 + (hp.b - hp.a)];
      for (int t : term_of[k]) lst.insert(lst.end(), {t, hpair_of[key], hp.mult});
    }
  }
  P.hent_ptr.assign(1, 0);
  for (auto &lst : hent) {
    P.hterms.insert(P.hterms.end(), lst.begin(), lst.end());
    P.hent_ptr.push_back((int)(P.hterms.size() / 3));
  }
  // weights [2][8][3*ncomp]: the doubles in the leading limb
  const int nw = std::max(3 * ncomp, 1);
  P.weights.assign((size_t)2 * 8 * nw, 0.0);
  for (int k = 0; k < 3 * ncomp; ++k) {
    P.weights[k] = weights[2 * k];
    P.weights[(size_t)8 * nw + k] = weights[2 * k + 1];
  }
  P.n = n;
  P.deg = deg;
  P.ncomp = ncomp;
  P.wdeg = wdeg;
  P.G = G;
  P.nprod = nprod;
  P.nlev = nlev;
  P.nslots = (int)acc;
  P.nfactor = nfactor;
  return 0;
}

// pointer to the data of v, or to a zero when v is empty
template <class T> const T *data_or_zero(const std::vector<T> &v) {
  static const T zero[2] = {T(0), T(0)};
  return v.empty() ? zero : v.data();
}

}  // namespace

extern "C" {

int phc_program_compile(phc_compiled **out, int n, int deg, int lanes, int ncomp,
                        const double *weights, const int *nterms, const int *eqs, const int *exps,
                        int CL, const double *coef) {
  PHC_REQUIRE(out != nullptr, "phc_program_compile: null output");
  *out = nullptr;
  PHC_REQUIRE(n >= 1 && n <= 64 && deg >= 0 && deg <= 64 && lanes >= 1 && ncomp >= 1,
              "phc_program_compile: bad sizes");
  PHC_REQUIRE(CL == 1 || CL == 2, "phc_program_compile: coefficient limbs CL must be 1 or 2");
  PHC_REQUIRE(weights && nterms, "phc_program_compile: null weights or term counts");
  int total = 0;
  for (int r = 0; r < ncomp; ++r) {
    PHC_REQUIRE(nterms[r] >= 0, "phc_program_compile: negative term count");
    total += nterms[r];
  }
  PHC_REQUIRE(total == 0 || (eqs && exps && coef), "phc_program_compile: null term arrays");
  for (int t = 0; t < total; ++t) {
    PHC_REQUIRE(eqs[t] >= 0 && eqs[t] < n, "phc_program_compile: equation index out of range");
    for (int j = 0; j < n; ++j)
      PHC_REQUIRE(exps[(size_t)t * n + j] >= 0 && exps[(size_t)t * n + j] <= 15,
                  "phc_program_compile: exponents must be in 0..15");
  }
  bool anyw = false;
  for (int k = 0; k < 6 * ncomp; ++k) anyw = anyw || weights[k] != 0.0;
  PHC_REQUIRE(anyw, "phc_program_compile: every weight is zero");
  phc_compiled *P = new phc_compiled();
  const int rc = compile(*P, n, deg, lanes, ncomp, weights, nterms, eqs, exps, CL, coef);
  if (rc != 0) {
    delete P;
    return rc;
  }
  *out = P;
  return 0;
}

int phc_compiled_program(const phc_compiled *c, phc_program *p) {
  PHC_REQUIRE(c != nullptr && p != nullptr, "phc_compiled_program: null argument");
  p->n = c->n;
  p->deg = c->deg;
  p->ncomp = c->ncomp;
  p->wdeg = c->wdeg;
  p->G = c->G;
  p->nprod = c->nprod;
  p->nlev = c->nlev;
  p->nslots = c->nslots;
  p->nfactor = c->nfactor;
  p->nodes = data_or_zero(c->nodes);
  p->lev_ptr = data_or_zero(c->lev_ptr);
  p->npair = (int)(c->pairs.size() / 4);
  p->pairs = data_or_zero(c->pairs);
  p->nterm = (int)c->term_pair.size();
  p->seg_ptr = data_or_zero(c->seg_ptr);
  p->term_pair = data_or_zero(c->term_pair);
  p->term_coef = data_or_zero(c->term_coef);
  p->njpair = (int)(c->jpairs.size() / 2);
  p->jpairs = data_or_zero(c->jpairs);
  p->njterm = (int)c->jterms.size();
  p->jent_ptr = data_or_zero(c->jent_ptr);
  p->jterms = data_or_zero(c->jterms);
  p->weights = data_or_zero(c->weights);
  p->hsig = data_or_zero(c->hsig);
  p->nhterm = (int)(c->hterms.size() / 3);
  p->hent_ptr = data_or_zero(c->hent_ptr);
  p->hterms = data_or_zero(c->hterms);
  p->nhpair = (int)(c->hpairs.size() / 2);
  p->hpairs = data_or_zero(c->hpairs);
  return 0;
}

int phc_compiled_destroy(phc_compiled *c) {
  delete c;
  return 0;
}

}  // extern "C"

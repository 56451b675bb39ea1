"""Polynomial systems, total-degree start systems and homotopies, compiled to
the flat program the batched tracker evaluates.

Reference behaviour restated here:
  total-degree start system q_i = x_i^{d_i} - c_i and its roots
      Root_Counts/Product/total_degree_start_systems.adb:458-483
  homotopy H(x,t) = gamma*(1-t)^k*q(x) + t^k*p(x)
      Deformations/Homotopy/standard_homotopy.ads:21, .adb:331-360
  value and gradient of polynomials at power series
      Polynomials/dbl2_polynomials_host.cpp:259 (CPU_dbl2_poly_evaldiff),
      with exponents > 1 as in Newton/dbl2_systems_host.cpp:12-111.

Program layout (all index arrays int32).  A homotopy is
    H_i(x, t) = sum_r w_r(t) * sum_m c[r,i,m] * x^m,
a sum of components r with constant coefficients and a weight polynomial
w_r(t) of degree <= 2 (component p with w = t^k, component q with
w = gamma*(1-t)^k).  Its monomials are nodes of a product DAG: nodes
0..n-1 are the variables, node n is the constant 1, node n+1+i is product i
= left[i] * right[i] (both earlier nodes).  At series order k every node's
coefficient k is sum_a left_a * right_{k-a}; the weighted component values
u[q] = sum_l W[r,l] * node[m]_{k-l} (W the Taylor coefficients of w_r at the
current t) feed the residual H_i = sum over terms c * u[q], and the Jacobian
entries A_ij = sum over terms c * e * W[r,0] * node[m/x_j]_0.
"""
import itertools
import math
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------- systems --

@dataclass
class System:
    """n polynomials in n variables: eqs[i] = list of (complex coef, exps tuple)."""
    n: int
    eqs: list

    def degrees(self):
        return [max(sum(e) for _, e in eq) for eq in self.eqs]

    def eval(self, x):
        """complex128 value at a point (for checks)."""
        return np.array([sum(c * np.prod(np.power(x, e)) for c, e in eq) for eq in self.eqs])


def unit_circle(rng, size=None):
    return np.exp(2j * np.pi * rng.uniform(0.0, 1.0, size))


def dense_system(n, degree, rng, rank_deficient=0):
    """n dense polynomials of the given degree with complex coefficients on
    the unit circle: every monomial of degree <= degree appears.
    rank_deficient > 0 replaces that many equations by random combinations
    of two others (a singular target family)."""
    exps = [e for tot in range(degree + 1) for e in _exponents(n, tot)]
    eqs = []
    for _ in range(n):
        cf = unit_circle(rng, len(exps))
        eqs.append([(complex(c), e) for c, e in zip(cf, exps)])
    for k in range(rank_deficient):
        i = n - 1 - k
        a, b = rng.choice(n - rank_deficient, 2, replace=False)
        la, lb = unit_circle(rng), unit_circle(rng)
        terms = {}
        for c, e in eqs[a]:
            terms[e] = terms.get(e, 0) + la * c
        for c, e in eqs[b]:
            terms[e] = terms.get(e, 0) + lb * c
        eqs[i] = [(complex(c), e) for e, c in terms.items()]
    return System(n, eqs)


def _exponents(n, total):
    out = []
    for combo in itertools.combinations_with_replacement(range(n), total):
        e = [0] * n
        for j in combo:
            e[j] += 1
        out.append(tuple(e))
    return out


def total_degree_start(degrees, rng):
    """q_i = x_i^{d_i} - c_i with random c_i on the unit circle, and all
    prod(d_i) start roots (total_degree_start_systems.adb:458-483)."""
    n = len(degrees)
    c = unit_circle(rng, n)
    eqs = []
    for i, d in enumerate(degrees):
        e = [0] * n
        e[i] = d
        eqs.append([(1.0 + 0j, tuple(e)), (complex(-c[i]), (0,) * n)])
    roots_per = [c[i] ** (1.0 / d) * np.exp(2j * np.pi * np.arange(d) / d)
                 for i, d in enumerate(degrees)]
    sols = np.array(list(itertools.product(*roots_per)), dtype=np.complex128)
    return System(n, eqs), sols, c


@dataclass
class Homotopy:
    """H(x,t) = gamma*(1-t)^k*q + t^k*p as weighted components."""
    target: System
    start: System
    gamma: complex
    k: int = 1
    components: list = field(default_factory=list)

    def __post_init__(self):
        k = self.k
        # weight polynomials in t, coefficients of 1, t, t^2
        wp = [0j] * 3
        wq = [0j] * 3
        if k == 1:
            wp[1] = 1.0
            wq[0], wq[1] = self.gamma, -self.gamma
        elif k == 2:
            wp[2] = 1.0
            wq[0], wq[1], wq[2] = self.gamma, -2 * self.gamma, self.gamma
        else:
            raise ValueError("relaxation power k must be 1 or 2")
        self.components = [(np.array(wp), self.target), (np.array(wq), self.start)]

    @property
    def n(self):
        return self.target.n

    def eval(self, x, t):
        return sum(np.polyval(w[::-1], t) * s.eval(x) for w, s in self.components)


def total_degree_homotopy(target, rng, k=1, gamma=None):
    q, sols, _ = total_degree_start(target.degrees(), rng)
    g = complex(unit_circle(rng)) if gamma is None else gamma
    return Homotopy(target, q, g, k), sols


# ---------------------------------------------------------------- program --

@dataclass
class Program:
    n: int
    deg: int
    left: np.ndarray
    right: np.ndarray
    lev_ptr: np.ndarray
    off: np.ndarray        # storage offset (complex units) of every node
    slen: np.ndarray       # storage length: D for series-kept nodes, 2 for ring nodes
    nslots: int
    pair_comp: np.ndarray
    pair_node: np.ndarray
    seg_ptr: np.ndarray    # terms of segment s: term range seg_ptr[s]..seg_ptr[s+1]
    eq_seg: np.ndarray     # segments of equation i: eq_seg[i]..eq_seg[i+1]
    term_pair: np.ndarray
    term_coef: np.ndarray  # [2][8][nterm] complex coefficient, octo-double planes
    jpair_comp: np.ndarray
    jpair_node: np.ndarray
    jent_ptr: np.ndarray   # terms of Jacobian entry i*n+j
    jterm_pair: np.ndarray
    jterm_coef: np.ndarray  # [2][8][njterm]
    weights: np.ndarray    # [2][8][ncomp*3] weight polynomial coefficients
    ncomp: int
    wdeg: int = 1
    # for checks and for the oracle's independent evaluation
    monomials: list = None
    node_exps: list = None


def _cplx_planes(vals):
    """complex double values -> [2][8][m] octo-double planes (exact)."""
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1)
    out = np.zeros((2, 8, max(len(vals), 1)))
    out[0, 0, :len(vals)] = vals.real
    out[1, 0, :len(vals)] = vals.imag
    return out


def compile_homotopy(hom, deg, lanes=64):
    """Build the tracker program for series of degree deg (D = deg+1)."""
    n = hom.n
    D = deg + 1
    # distinct monomials over all components, plus the derivative monomials
    node_of = {}
    node_exps = [tuple(1 if j == i else 0 for j in range(n)) for i in range(n)]
    for i, e in enumerate(node_exps):
        node_of[e] = i
    one = (0,) * n
    node_of[one] = n
    node_exps.append(one)
    left, right, level = [], [], {i: 0 for i in range(n + 1)}

    def ensure(e):
        if e in node_of:
            return node_of[e]
        j = max(v for v in range(n) if e[v] > 0)
        rest = list(e)
        rest[j] -= 1
        rest = tuple(rest)
        a = ensure(rest)
        idx = len(node_exps)
        node_of[e] = idx
        node_exps.append(e)
        left.append(a)
        right.append(j)
        level[idx] = level[a] + 1
        return idx

    monos = set()
    for _, sysm in hom.components:
        for eq in sysm.eqs:
            for _, e in eq:
                monos.add(e)
    for e in sorted(monos, key=lambda e: (sum(e), e)):
        ensure(e)
        for j in range(n):
            if e[j] > 0:
                d = list(e)
                d[j] -= 1
                ensure(tuple(d))
    nprod = len(left)
    # order product nodes by level (stable); renumber
    prod_ids = list(range(n + 1, n + 1 + nprod))
    order = sorted(prod_ids, key=lambda i: (level[i], i))
    remap = {i: i for i in range(n + 1)}
    for new, old in enumerate(order):
        remap[old] = n + 1 + new
    L = np.zeros(nprod, np.int32)
    Rt = np.zeros(nprod, np.int32)
    levs = []
    new_exps = [None] * (n + 1 + nprod)
    for i in range(n + 1):
        new_exps[i] = node_exps[i]
    for old in order:
        new = remap[old] - n - 1
        L[new] = remap[left[old - n - 1]]
        Rt[new] = remap[right[old - n - 1]]
        levs.append(level[old])
        new_exps[remap[old]] = node_exps[old]
    node_exps = new_exps
    node_of = {e: i for i, e in enumerate(node_exps)}
    nlev = max(levs) if levs else 0
    lev_ptr = np.zeros(nlev + 1, np.int32)
    for lv in range(1, nlev + 1):
        lev_ptr[lv] = lev_ptr[lv - 1] + sum(1 for x in levs if x == lv)
    # storage: variables, ONE and every node used as a factor keep D
    # coefficients; other nodes keep a ring of 2 (orders k and k-1)
    nn = n + 1 + nprod
    factor = np.zeros(nn, bool)
    factor[: n + 1] = True
    for a in L:
        factor[a] = True
    wdeg = max(int(np.max(np.nonzero(np.abs(w) > 0)[0])) for w, _ in hom.components)
    ring = min(wdeg + 1, D)
    slen = np.where(factor, D, ring).astype(np.int32)
    off = np.zeros(nn, np.int32)
    acc = 0
    for i in range(nn):
        off[i] = acc
        acc += slen[i]
    nslots = int(acc)

    # weighted component values: pairs (component, node)
    pair_of = {}
    pair_comp, pair_node = [], []
    # residual terms per equation
    eq_terms = [[] for _ in range(n)]
    for r, (w, sysm) in enumerate(hom.components):
        for i, eq in enumerate(sysm.eqs):
            merged = {}
            for c, e in eq:
                merged[e] = merged.get(e, 0) + c
            for e, c in merged.items():
                key = (r, node_of[e])
                if key not in pair_of:
                    pair_of[key] = len(pair_comp)
                    pair_comp.append(r)
                    pair_node.append(node_of[e])
                eq_terms[i].append((pair_of[key], c))
    # segments: split every equation's terms into chunks for the lanes
    nterm_tot = sum(len(t) for t in eq_terms)
    segs_per_eq = max(1, lanes // n)
    seg_ptr, eq_seg, term_pair, term_cf = [0], [0], [], []
    for i in range(n):
        terms = eq_terms[i]
        ns = min(segs_per_eq, max(1, len(terms)))
        chunks = np.array_split(np.arange(len(terms)), ns)
        for ch in chunks:
            for t in ch:
                term_pair.append(terms[t][0])
                term_cf.append(terms[t][1])
            seg_ptr.append(len(term_pair))
        eq_seg.append(len(seg_ptr) - 1)
    assert len(term_pair) == nterm_tot
    # Jacobian: entry (i,j) terms c*e_j*W0*node(m/x_j)
    jpair_of = {}
    jpair_comp, jpair_node = [], []
    ent_terms = [[] for _ in range(n * n)]
    for r, (w, sysm) in enumerate(hom.components):
        for i, eq in enumerate(sysm.eqs):
            for c, e in eq:
                for j in range(n):
                    if e[j] == 0:
                        continue
                    d = list(e)
                    d[j] -= 1
                    key = (r, node_of[tuple(d)])
                    if key not in jpair_of:
                        jpair_of[key] = len(jpair_comp)
                        jpair_comp.append(r)
                        jpair_node.append(node_of[tuple(d)])
                    ent_terms[i * n + j].append((jpair_of[key], c * e[j]))
    jent_ptr, jterm_pair, jterm_cf = [0], [], []
    for ent in ent_terms:
        for q, c in ent:
            jterm_pair.append(q)
            jterm_cf.append(c)
        jent_ptr.append(len(jterm_pair))
    weights = _cplx_planes(np.concatenate([w for w, _ in hom.components]))
    return Program(
        n=n, deg=deg, left=L, right=Rt, lev_ptr=lev_ptr, off=off, slen=slen, nslots=nslots,
        pair_comp=np.array(pair_comp, np.int32), pair_node=np.array(pair_node, np.int32),
        seg_ptr=np.array(seg_ptr, np.int32), eq_seg=np.array(eq_seg, np.int32),
        term_pair=np.array(term_pair, np.int32), term_coef=_cplx_planes(term_cf),
        jpair_comp=np.array(jpair_comp, np.int32), jpair_node=np.array(jpair_node, np.int32),
        jent_ptr=np.array(jent_ptr, np.int32), jterm_pair=np.array(jterm_pair, np.int32),
        jterm_coef=_cplx_planes(jterm_cf), weights=weights, ncomp=len(hom.components), wdeg=wdeg,
        monomials=sorted(monos), node_exps=node_exps)


def flops_per_order(prog):
    """complex multiply-adds per series order (products at order k use k+1)."""
    return dict(products=len(prog.left), pairs=len(prog.pair_comp),
                terms=len(prog.term_pair), jterms=len(prog.jterm_pair))

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

The step rule needs two more things per equation.  The mixed residual
(standard_mixed_residuals.adb:472-484) evaluates every equation with the
moduli of its coefficients at |x|: the terms of an equation are grouped by
monomial (the components of one monomial are adjacent terms of one
segment), so a group's coefficient sum_r w_r(t) c[r,i,m] is formed before
its modulus is taken.  The curvature step (Hesse_Pade,
standard_predictor_convolutions.adb:648-900) needs the largest singular
value of the Hessian of every equation: for equations of degree <= 2 the
reference computes it once, at the first point it meets, and keeps it
(Cached_Singular_Values, :575-646), so here it is computed once on the host
at t = 0 (where a Hessian of degree <= 2 does not depend on x); equations
of higher degree get Hessian terms that the tracker evaluates at every step.
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
    rank_deficient > 0 makes the top-degree forms of that many equations
    random combinations of the top-degree forms of two others (their lower
    degree terms stay random): the leading forms of the system are then
    rank-deficient, the system has common zeros at infinity and part of the
    total-degree paths diverge or end at singular points."""
    exps = [e for tot in range(degree + 1) for e in _exponents(n, tot)]
    eqs = []
    for _ in range(n):
        cf = unit_circle(rng, len(exps))
        eqs.append([(complex(c), e) for c, e in zip(cf, exps)])
    for k in range(rank_deficient):
        i = n - 1 - k
        a, b = rng.choice(n - rank_deficient, 2, replace=False)
        la, lb = unit_circle(rng), unit_circle(rng)
        top = {}
        for c, e in eqs[a]:
            if sum(e) == degree:
                top[e] = top.get(e, 0) + la * c
        for c, e in eqs[b]:
            if sum(e) == degree:
                top[e] = top.get(e, 0) + lb * c
        eqs[i] = [(c, e) for c, e in eqs[i] if sum(e) < degree] + \
                 [(complex(c), e) for e, c in top.items()]
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

JT_TERM_BITS, JT_PAIR_BITS, JT_EXP_BITS = 16, 12, 4


@dataclass
class Program:
    """Flat tracker program (layout of phc_program in include/phcmd.h).

    nodes [8*nprod]: per product node (ordered by DAG level)
        aoff, amask, boff, soff, smask, adn, sdn, left
      aoff/amask: storage of the left factor (a variable or a product node);
      boff: series of the right factor (a variable); soff/smask: own storage
      (coefficient a at soff + (a & smask)); adn/sdn: index of the left
      factor's / own order-k correction in the dnode buffer (-1: none);
      left: node number of the left factor.
    pairs [4*npair]: component r, node storage off, mask, node number
    seg_ptr [n*G+1]: terms of segment s = i*G + g (G segments per equation);
      the terms of one monomial of an equation are adjacent in one segment
    term_pair [nterm], term_coef [2][nterm] (real, imaginary parts)
    jpairs [2*njpair]: component r, node storage off (order 0)
    jent_ptr [n*n+1]: Jacobian terms of entry i*n+j
    jterms [njterm]: term index | jpair << 16 | exponent << 28
    weights [2][8][3*ncomp]: coefficients of 1, t, t^2 of every w_r
    hsig [n]: largest singular value of the Hessian of equation i at t = 0
      when its degree is <= 2, -1 when the tracker evaluates it every step
    hent_ptr [n*n*(n+1)/2+1]: Hessian terms of entry (i; a, b), a <= b, at
      index i*n*(n+1)/2 + a*n - a*(a-1)/2 + (b-a)
    hterms [3*nhterm]: term index, hpair, multiplicity e_a*e_b (a != b) or
      e_a*(e_a-1) (a = b) of d^2 x^m / dx_a dx_b = mult * x^(m - e_a - e_b)
    hpairs [2*nhpair]: component r, storage offset of the node x^(m-e_a-e_b)
    """
    n: int
    deg: int
    ncomp: int
    wdeg: int
    G: int
    nodes: np.ndarray
    lev_ptr: np.ndarray
    nslots: int
    nfactor: int
    pairs: np.ndarray
    seg_ptr: np.ndarray
    term_pair: np.ndarray
    term_coef: np.ndarray
    jpairs: np.ndarray
    jent_ptr: np.ndarray
    jterms: np.ndarray
    weights: np.ndarray
    hsig: np.ndarray = None
    hent_ptr: np.ndarray = None
    hterms: np.ndarray = None
    hpairs: np.ndarray = None
    node_exps: list = None
    var_off: np.ndarray = None

    @property
    def nprod(self):
        return len(self.nodes) // 8

    @property
    def nlev(self):
        return len(self.lev_ptr) - 1


def _cplx_planes(vals):
    """complex double values -> [2][8][m] octo-double planes (exact)."""
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1)
    out = np.zeros((2, 8, max(len(vals), 1)))
    out[0, 0, :len(vals)] = vals.real
    out[1, 0, :len(vals)] = vals.imag
    return out


def segments_per_equation(n, lanes=64):
    """largest power of two G with n*G <= lanes (at least 1)"""
    G = 1
    while n * G * 2 <= lanes:
        G *= 2
    return G


def compile_homotopy(hom, deg, lanes=64):
    """Build the tracker program for series of degree deg (D = deg+1)."""
    n = hom.n
    D = deg + 1
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

    # merged terms per (component, equation, monomial)
    comp_terms = []
    for r, (w, sysm) in enumerate(hom.components):
        for i, eq in enumerate(sysm.eqs):
            merged = {}
            for c, e in eq:
                merged[e] = merged.get(e, 0) + c
            for e, c in merged.items():
                comp_terms.append((r, i, e, complex(c)))
    # polynomial degree of every equation of the homotopy
    pdg = [0] * n
    for _, i, e, c in comp_terms:
        if c != 0:
            pdg[i] = max(pdg[i], sum(e))
    monos = sorted({e for _, _, e, _ in comp_terms}, key=lambda e: (sum(e), e))
    for e in monos:
        ensure(e)
        for j in range(n):
            if e[j] > 0:
                d = list(e)
                d[j] -= 1
                ensure(tuple(d))
    for _, i, e, _ in comp_terms:
        if pdg[i] > 2:
            for _, div, _ in _hessian_divisors(e):
                ensure(div)
    nprod = len(left)
    order = sorted(range(n + 1, n + 1 + nprod), key=lambda i: (level[i], i))
    remap = {i: i for i in range(n + 1)}
    for new, old in enumerate(order):
        remap[old] = n + 1 + new
    L = np.zeros(nprod, np.int64)
    Rt = np.zeros(nprod, np.int64)
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
    wdeg = max(int(np.max(np.nonzero(np.abs(w) > 0)[0])) for w, _ in hom.components)
    ring = 1
    while ring < min(wdeg + 1, D):
        ring *= 2
    # storage: variables, ONE and factor nodes keep D coefficients, the
    # others a power-of-two ring of the last orders
    nn = n + 1 + nprod
    factor = np.zeros(nn, bool)
    factor[: n + 1] = True
    for a in L:
        factor[a] = True
    slen = np.where(factor, D, ring)
    off = np.zeros(nn, np.int64)
    acc = 0
    for i in range(nn):
        off[i] = acc
        acc += slen[i]
    mask = np.where(factor, -1, ring - 1)
    dn = -np.ones(nn, np.int64)
    nfactor = 0
    for i in range(n + 1, nn):
        if factor[i]:
            dn[i] = nfactor
            nfactor += 1
    nodes = np.zeros((nprod, 8), np.int32)
    for i in range(nprod):
        a, b, sidx = int(L[i]), int(Rt[i]), n + 1 + i
        nodes[i] = (off[a], mask[a], off[b], off[sidx], mask[sidx], dn[a] if a > n else -1,
                    dn[sidx], a)
    # pairs (component, node) and terms, segmented per equation; the terms of
    # one monomial (one per component) are adjacent and never split
    pair_of, pairs = {}, []
    eq_terms = [[] for _ in range(n)]
    term_key = {}
    for r, i, e, c in comp_terms:
        key = (r, node_of[e])
        if key not in pair_of:
            pair_of[key] = len(pairs)
            m = node_of[e]
            pairs.append((r, off[m], mask[m], m))
        eq_terms[i].append((pair_of[key], c, (r, i, e)))
    G = segments_per_equation(n, lanes)
    seg_ptr, term_pair, term_cf = [0], [], []
    for i in range(n):
        terms = sorted(eq_terms[i], key=lambda tm: node_of[tm[2][2]])  # stable
        groups = []
        for tm in terms:
            if groups and node_of[groups[-1][-1][2][2]] == node_of[tm[2][2]]:
                groups[-1].append(tm)
            else:
                groups.append([tm])
        for ch in np.array_split(np.arange(len(groups)), G):
            for gi in ch:
                for q, c, key in groups[gi]:
                    term_key[key] = len(term_pair)
                    term_pair.append(q)
                    term_cf.append(c)
            seg_ptr.append(len(term_pair))
    # Jacobian: entry (i,j) sums coef[term] * e_j * W_r0 * node(m/x_j)_0
    jpair_of, jpairs = {}, []
    ent = [[] for _ in range(n * n)]
    for r, i, e, c in comp_terms:
        for j in range(n):
            if e[j] == 0:
                continue
            d = list(e)
            d[j] -= 1
            m = node_of[tuple(d)]
            key = (r, m)
            if key not in jpair_of:
                jpair_of[key] = len(jpairs)
                jpairs.append((r, off[m]))
            ent[i * n + j].append((term_key[(r, i, e)], jpair_of[key], e[j]))
    jent_ptr, jterms = [0], []
    for lst in ent:
        for t, q, ex in lst:
            assert t < (1 << JT_TERM_BITS) and q < (1 << JT_PAIR_BITS) and ex < (1 << JT_EXP_BITS)
            jterms.append(t | (q << JT_TERM_BITS) | (ex << (JT_TERM_BITS + JT_PAIR_BITS)))
        jent_ptr.append(len(jterms))
    cf = np.array(term_cf, np.complex128)
    # Hessians: equations of degree <= 2 get the largest singular value of
    # their (constant) Hessian at t = 0, the others Hessian terms
    nn = n * (n + 1) // 2
    hsig = np.zeros(n)
    hent = [[] for _ in range(n * nn)]
    hpair_of, hpairs = {}, []
    for i in range(n):
        if pdg[i] <= 1:
            continue
        if pdg[i] == 2:
            Hm = np.zeros((n, n), np.complex128)
            for r, (w, _) in enumerate(hom.components):
                for rr, ii, e, c in comp_terms:
                    if rr != r or ii != i:
                        continue
                    for a, b, mult in _hessian_pairs(e):
                        Hm[a, b] += w[0] * c * mult
                        if a != b:
                            Hm[b, a] += w[0] * c * mult
            hsig[i] = float(np.linalg.svd(Hm, compute_uv=False)[0])
            continue
        hsig[i] = -1.0
    for r, i, e, c in comp_terms:
        if pdg[i] <= 2:
            continue
        for (a, b, mult), (_, div, _) in zip(_hessian_pairs(e), _hessian_divisors(e)):
            m = node_of[div]
            key = (r, m)
            if key not in hpair_of:
                hpair_of[key] = len(hpairs)
                hpairs.append((r, off[m]))
            hent[i * nn + a * n - a * (a - 1) // 2 + (b - a)].append(
                (term_key[(r, i, e)], hpair_of[key], mult))
    hent_ptr, hterms = [0], []
    for lst in hent:
        for trio in lst:
            hterms.extend(trio)
        hent_ptr.append(len(hterms) // 3)
    return Program(
        n=n, deg=deg, ncomp=len(hom.components), wdeg=wdeg, G=G,
        nodes=nodes.reshape(-1), lev_ptr=lev_ptr, nslots=int(acc), nfactor=nfactor,
        pairs=np.array(pairs, np.int32).reshape(-1), seg_ptr=np.array(seg_ptr, np.int32),
        term_pair=np.array(term_pair, np.int32), term_coef=np.stack([cf.real, cf.imag]),
        jpairs=np.array(jpairs, np.int32).reshape(-1), jent_ptr=np.array(jent_ptr, np.int32),
        jterms=np.array(jterms, np.uint32).view(np.int32),
        weights=_cplx_planes(np.concatenate([w for w, _ in hom.components])),
        hsig=hsig, hent_ptr=np.array(hent_ptr, np.int32), hterms=np.array(hterms, np.int32),
        hpairs=np.array(hpairs, np.int32).reshape(-1),
        node_exps=node_exps, var_off=off[:n + 1].astype(np.int32))


def _hessian_pairs(e):
    """(a, b, multiplicity) of the nonzero second derivatives
    d^2 x^e / dx_a dx_b, a <= b"""
    n = len(e)
    out = []
    for a in range(n):
        for b in range(a, n):
            if a == b and e[a] >= 2:
                out.append((a, b, e[a] * (e[a] - 1)))
            elif a != b and e[a] >= 1 and e[b] >= 1:
                out.append((a, b, e[a] * e[b]))
    return out


def _hessian_divisors(e):
    """(a, x^(e - e_a - e_b) exponents, b) for the pairs of _hessian_pairs"""
    out = []
    for a, b, _ in _hessian_pairs(e):
        d = list(e)
        d[a] -= 1
        d[b] -= 1
        out.append((a, tuple(d), b))
    return out


def compile_terms(n, deg, weights, comps, lanes=64):
    """The tracker program built by the C++ compiler of libphcmd.so
    (phc_program_compile), for callers that hold the terms as arrays.
    weights: complex [ncomp][3] (coefficients of 1, t, t^2 of every w_r);
    comps: per component (eqs [m] int, exps [m][n] int, coef) with coef
    complex [m] (double coefficients) or float [4][m] double-double limb
    planes (real high, real low, imaginary high, imaginary low)."""
    import ctypes
    from ._lib import P_double, P_int, PhcProgram, check, load
    lib = load()
    w = np.asarray(weights, np.complex128).reshape(-1, 3)
    ncomp = w.shape[0]
    wflat = np.ascontiguousarray(np.stack([w.real, w.imag], axis=-1).reshape(-1))
    cl = 2 if any(np.asarray(c).dtype != np.complex128 for _, _, c in comps) else 1
    eqs, exps, planes, nterms = [], [], [], []
    for e_i, x_i, c in comps:
        e_i = np.asarray(e_i, np.int32).reshape(-1)
        nterms.append(len(e_i))
        eqs.append(e_i)
        exps.append(np.asarray(x_i, np.int32).reshape(len(e_i), n))
        c = np.asarray(c)
        if c.dtype == np.complex128:
            p = np.zeros((2 * cl, len(e_i)))
            p[0], p[cl] = c.real, c.imag
        else:
            p = np.asarray(c, np.float64).reshape(4, len(e_i))
        planes.append(p)
    eqs = np.ascontiguousarray(np.concatenate(eqs) if eqs else np.zeros(0, np.int32), np.int32)
    exps = np.ascontiguousarray(np.concatenate(exps) if exps else np.zeros((0, n), np.int32),
                                np.int32)
    coef = np.ascontiguousarray(np.concatenate(planes, axis=1) if planes else np.zeros((2 * cl, 0)))
    nt = np.ascontiguousarray(nterms, np.int32)
    handle = ctypes.c_void_p()

    def ptr(a, typ):
        return a.ctypes.data_as(typ) if a.size else None

    check(lib.phc_program_compile(ctypes.byref(handle), n, deg, lanes, ncomp,
                                  wflat.ctypes.data_as(P_double), nt.ctypes.data_as(P_int),
                                  ptr(eqs, P_int), ptr(exps, P_int), cl, ptr(coef, P_double)))
    try:
        s = PhcProgram()
        check(lib.phc_compiled_program(handle, ctypes.byref(s)))

        def arr(p, size, dtype):
            if size == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(p, shape=(size,)).astype(dtype, copy=True)

        nw = max(3 * ncomp, 1)
        nh = n * (n + 1) // 2
        return Program(
            n=s.n, deg=s.deg, ncomp=s.ncomp, wdeg=s.wdeg, G=s.G,
            nodes=arr(s.nodes, 8 * s.nprod, np.int32), lev_ptr=arr(s.lev_ptr, s.nlev + 1, np.int32),
            nslots=s.nslots, nfactor=s.nfactor,
            pairs=arr(s.pairs, 4 * s.npair, np.int32),
            seg_ptr=arr(s.seg_ptr, n * s.G + 1, np.int32),
            term_pair=arr(s.term_pair, s.nterm, np.int32),
            term_coef=arr(s.term_coef, 2 * s.nterm, np.float64).reshape(2, s.nterm),
            jpairs=arr(s.jpairs, 2 * s.njpair, np.int32),
            jent_ptr=arr(s.jent_ptr, n * n + 1, np.int32),
            jterms=arr(s.jterms, s.njterm, np.int32),
            weights=arr(s.weights, 2 * 8 * nw, np.float64).reshape(2, 8, nw),
            hsig=arr(s.hsig, n, np.float64),
            hent_ptr=arr(s.hent_ptr, n * nh + 1, np.int32),
            hterms=arr(s.hterms, 3 * s.nhterm, np.int32),
            hpairs=arr(s.hpairs, 2 * s.nhpair, np.int32),
            var_off=np.arange(n + 1, dtype=np.int32) * (deg + 1))
    finally:
        lib.phc_compiled_destroy(handle)


def homotopy_terms(hom):
    """(weights [ncomp][3], comps) of a Homotopy in the arrays of compile_terms"""
    comps = []
    for _, sysm in hom.components:
        eqs, exps, cf = [], [], []
        for i, eq in enumerate(sysm.eqs):
            for c, e in eq:
                eqs.append(i)
                exps.append(e)
                cf.append(c)
        comps.append((np.array(eqs, np.int32), np.array(exps, np.int32).reshape(len(eqs), hom.n),
                      np.array(cf, np.complex128)))
    return np.array([w for w, _ in hom.components], np.complex128), comps


def compile_homotopy_native(hom, deg, lanes=64):
    """compile_homotopy through the C++ compiler (the same arrays; the
    Hessian singular values by a Jacobi SVD, to rounding)"""
    w, comps = homotopy_terms(hom)
    return compile_terms(hom.n, deg, w, comps, lanes)

"""PHCpack's text formats for polynomial systems and solution lists, and a
total-degree solve of a system read from them on the GPU.

Formats restated (the reference's readers and writers, read for behaviour):
  polynomial systems   Math_Lib/Polynomials/standard_complex_polynomials_io.ads
      (syntax: `x**2*y + 1/2*z*y**2 - 2*z + 3;`, `^` or `**` for powers,
      `i` or `I` the imaginary unit, `(1.01 + 2.8*i)*x1**2`, products of
      parenthesized polynomials, `;` ends a polynomial) and
      standard_complex_poly_systems_io.adb (the first line holds the number
      of equations, optionally followed by the number of unknowns); unknowns
      enter the symbol table in the order of their first appearance.
  solution lists       Deformations/Solutions/standard_complex_solutions_io.adb
      (`put` :591-596 writes "len n", a bar of '=' and per solution
      "solution k :", "t : re im", "m : m", "the solution for t :", one line
      " name : re im" per unknown and the diagnostics line
      "== err : e = rco : r = res : s =" (:332-338) closed by "=";
      `get` :207-231 reads the same, skipping text after "solution k :").
The example systems with stored solutions are the reference's Demo files
(src/Demo/katsura5, cyclic5, ...), kept as fixtures under tests/golden/demo.
"""
import re

import numpy as np

from .homotopy import System

BAR = "=" * 59


class ParseError(ValueError):
    pass


# ------------------------------------------------------------ polynomials --

_TOKEN = re.compile(r"""
    (?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?) |
    (?P<name>[A-Za-z][A-Za-z0-9_]*) |
    (?P<pow>\*\*|\^) |
    (?P<op>[-+*/();]) |
    (?P<ws>\s+)
""", re.VERBOSE)


def _tokens(text):
    pos, out = 0, []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ParseError(f"unexpected character {text[pos]!r} at {pos}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        out.append((kind, m.group(kind)))
    return out


class _Poly:
    """sparse polynomial: {exponent tuple over the growing symbol list: coef}"""

    def __init__(self, terms=None):
        self.t = dict(terms or {})

    @staticmethod
    def const(c):
        return _Poly({(): complex(c)})

    def _pad(self, n):
        return {e + (0,) * (n - len(e)): c for e, c in self.t.items()}

    def add(self, other, sign=1.0):
        n = max([len(e) for e in list(self.t) + list(other.t)] + [0])
        r = self._pad(n)
        for e, c in other._pad(n).items():
            r[e] = r.get(e, 0j) + sign * c
        return _Poly(r)

    def mul(self, other):
        n = max([len(e) for e in list(self.t) + list(other.t)] + [0])
        a, b = self._pad(n), other._pad(n)
        r = {}
        for ea, ca in a.items():
            for eb, cb in b.items():
                e = tuple(x + y for x, y in zip(ea, eb))
                r[e] = r.get(e, 0j) + ca * cb
        return _Poly(r)

    def is_const(self):
        return all(sum(e) == 0 for e in self.t)

    def const_value(self):
        return sum(self.t.values(), 0j)


class _Parser:
    """recursive descent over the tokens of one polynomial (up to ';')"""

    def __init__(self, toks, symbols):
        self.toks, self.i, self.sym = toks, 0, symbols

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expr(self):
        kind, val = self.peek()
        sign = 1.0
        if kind == "op" and val in "+-":
            self.take()
            sign = -1.0 if val == "-" else 1.0
        p = self.term()
        if sign < 0:
            p = _Poly.const(-1).mul(p)
        while True:
            kind, val = self.peek()
            if kind == "op" and val in "+-":
                self.take()
                p = p.add(self.term(), -1.0 if val == "-" else 1.0)
            else:
                return p

    def term(self):
        p = self.factor()
        while True:
            kind, val = self.peek()
            if kind == "op" and val == "*":
                self.take()
                p = p.mul(self.factor())
            elif kind == "op" and val == "/":
                self.take()
                d = self.factor()
                if not d.is_const() or d.const_value() == 0:
                    raise ParseError("division by a nonconstant or zero")
                p = _Poly.const(1.0 / d.const_value()).mul(p)
            elif kind in ("num", "name") or (kind == "op" and val == "("):
                p = p.mul(self.factor())  # implicit product
            else:
                return p

    def factor(self):
        kind, val = self.peek()
        if kind == "op" and val in "+-":
            self.take()
            f = self.factor()
            return _Poly.const(-1).mul(f) if val == "-" else f
        base = self.primary()
        kind, val = self.peek()
        if kind == "pow":
            self.take()
            k2, v2 = self.take()
            if k2 != "num" or not v2.isdigit():
                raise ParseError("exponent must be a natural number")
            r = _Poly.const(1)
            for _ in range(int(v2)):
                r = r.mul(base)
            return r
        return base

    def primary(self):
        kind, val = self.take()
        if kind == "num":
            return _Poly.const(float(val))
        if kind == "name":
            if val in ("i", "I"):
                return _Poly.const(1j)
            if val not in self.sym:
                self.sym.append(val)
            k = self.sym.index(val)
            return _Poly({(0,) * k + (1,): 1.0 + 0j})
        if kind == "op" and val == "(":
            p = self.expr()
            k2, v2 = self.take()
            if (k2, v2) != ("op", ")"):
                raise ParseError("missing ')'")
            return p
        raise ParseError(f"unexpected token {val!r}")


def parse_system(text, symbols=None):
    """(System, symbols) from PHCpack's input format: the number of
    equations (and optionally unknowns) on the first line, then the
    polynomials, each ended by ';'.  Text after the last polynomial (TITLE,
    ROOT COUNTS, THE SOLUTIONS, ...) is ignored.  symbols: an initial symbol
    table (the order of the unknowns), as the reference's reader takes an
    initialized table; otherwise unknowns are numbered as they appear."""
    lines = text.strip().splitlines()
    head = lines[0].split()
    neq = int(head[0])
    nvar = int(head[1]) if len(head) > 1 else neq
    toks = _tokens(_poly_part(lines[1:], neq))
    symbols, polys = list(symbols or []), []
    start = 0
    for _ in range(neq):
        end = start
        while end < len(toks) and toks[end] != ("op", ";"):
            end += 1
        if end == len(toks):
            raise ParseError("missing ';'")
        pr = _Parser(toks[start:end], symbols)
        p = pr.expr()
        if pr.i != end - start:
            raise ParseError(f"trailing tokens in polynomial {len(polys) + 1}")
        polys.append(p)
        start = end + 1
    if len(symbols) != nvar:
        raise ParseError(f"{len(symbols)} unknowns, expected {nvar}")
    if neq != nvar:
        raise ParseError("only square systems are tracked")
    eqs = []
    for p in polys:
        terms = [(c, tuple(list(e) + [0] * (nvar - len(e)))) for e, c in p._pad(nvar).items() if c != 0]
        eqs.append(sorted(terms, key=lambda ce: (sum(ce[1]), ce[1])))
    return System(nvar, eqs), symbols


def _poly_part(lines, neq):
    """the text of the first neq polynomials (up to the neq-th ';')"""
    out, count = [], 0
    for ln in lines:
        out.append(ln)
        count += ln.count(";")
        if count >= neq:
            break
    return "\n".join(out)


def write_system(system, symbols):
    """the system in PHCpack's input format (coefficients as (re + im*i))"""
    out = [f"{system.n}"]
    for eq in system.eqs:
        terms = []
        for c, e in eq:
            mono = "*".join(f"{symbols[j]}" + (f"^{k}" if k > 1 else "")
                            for j, k in enumerate(e) if k > 0)
            cf = f"({c.real:.17e} + {c.imag:.17e}*i)"
            terms.append(cf + ("*" + mono if mono else ""))
        out.append(" " + "\n + ".join(terms) + ";")
    return "\n".join(out) + "\n"


# -------------------------------------------------------------- solutions --

def _fmt(x):
    """Ada's default put of a double: sign or blank, d.dddddddddddddd, E+dd"""
    s = f"{x:.14E}"
    mant, ex = s.split("E")
    e = int(ex)
    return ("-" if mant.startswith("-") else " ") + mant.lstrip("-") + \
        "E" + ("-" if e < 0 else "+") + f"{abs(e):02d}"


def _fmt_diag(x):
    """put(x, 2, 3, 3): two positions before the point, three after, a
    three-character exponent"""
    s = f"{x:.3E}"
    mant, ex = s.split("E")
    e = int(ex)
    return f"{mant:>6}" + "E" + ("-" if e < 0 else "+") + f"{abs(e):02d}"


def write_solutions(sols, symbols, header=True):
    """PHCpack's solution list: sols is a list of dicts with keys x (complex
    vector) and optionally t, m, err, rco, res."""
    n = len(symbols)
    out = []
    if header:
        out.append("THE SOLUTIONS :")
    out.append(f"{len(sols)} {n}")
    out.append(BAR)
    for k, s in enumerate(sols, 1):
        t = complex(s.get("t", 1.0))
        out.append(f"solution {k} :")
        out.append(f"t : {_fmt(t.real)}  {_fmt(t.imag)}")
        out.append(f"m : {int(s.get('m', 1))}")
        out.append("the solution for t :")
        for name, v in zip(symbols, np.asarray(s["x"], dtype=np.complex128)):
            out.append(f" {name} : {_fmt(v.real)}  {_fmt(v.imag)}")
        out.append(f"== err : {_fmt_diag(s.get('err', 0.0))} = rco : {_fmt_diag(s.get('rco', 0.0))}"
                   f" = res : {_fmt_diag(s.get('res', 0.0))} ==")
    return "\n".join(out) + "\n"


_SOLNUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?"


def parse_solutions(text, symbols=None):
    """the solution list after "THE SOLUTIONS" (or the whole text when it
    starts with "len n"): a list of dicts t, m, x (complex128 in the order of
    the listed unknowns), names, err, rco, res."""
    pos = text.find("THE SOLUTIONS")
    body = text[pos:].split("\n", 1)[1] if pos >= 0 else text
    lines = [ln for ln in body.splitlines()]
    i = 0
    while i < len(lines) and not lines[i].strip():
        i += 1
    head = lines[i].split()
    nsol, n = int(head[0]), int(head[1])
    i += 1
    sols = []
    num = re.compile(_SOLNUM)
    while len(sols) < nsol:
        # a solution starts at its "t :" line (the banner before it is
        # "solution k :" or, in the path trackers' output, "== k = ..."),
        # then "m :"
        while i < len(lines) and not lines[i].lstrip().startswith(("t :", "m :")):
            i += 1
        if i + 1 >= len(lines):
            raise ParseError(f"{len(sols)} of {nsol} solutions found")
        tv = [1.0, 0.0]  # a block without its "t :" line (one Demo file has one)
        if lines[i].lstrip().startswith("t :"):
            tv = [float(v) for v in num.findall(lines[i].split(":", 1)[1])]
            i += 1
        m = int(lines[i].split(":", 1)[1].split()[0])
        i += 1
        while i < len(lines) and not lines[i].lstrip().startswith("the solution"):
            i += 1
        i += 1
        x, names = [], []
        for _ in range(n):
            name, vals = lines[i].split(":", 1)
            re_im = [float(v) for v in num.findall(vals)]
            x.append(complex(re_im[0], re_im[1]))
            names.append(name.strip())
            i += 1
        diag = {}
        if i < len(lines) and lines[i].lstrip().startswith("==") and "err" in lines[i]:
            for key in ("err", "rco", "res"):
                mm = re.search(key + r"\s*:\s*(" + _SOLNUM + ")", lines[i])
                diag[key] = float(mm.group(1)) if mm else 0.0
            i += 1
        x = np.array(x, dtype=np.complex128)
        if symbols is not None and names != list(symbols):
            if sorted(names) != sorted(symbols):
                raise ParseError(f"unknowns {names} do not match {list(symbols)}")
            order = [names.index(s) for s in symbols]
            x, names = x[order], list(symbols)
        sols.append(dict(t=complex(tv[0], tv[1]), m=m, x=x, names=names, **diag))
    return sols


# ------------------------------------------------------------------ solve --

def jacobian(system, x):
    n = system.n
    J = np.zeros((n, n), dtype=np.complex128)
    for i, eq in enumerate(system.eqs):
        for c, e in eq:
            for j in range(n):
                if e[j] == 0:
                    continue
                d = list(e)
                d[j] -= 1
                J[i, j] += c * e[j] * np.prod(np.power(x, d))
    return J


def solve_total_degree(system, seed=0, K=2, params=None, iters=4, tol=1e-28):
    """all solutions the total-degree homotopy reaches: FP64 tracking on the
    GPU, then Newton on the GPU at the endpoints in K limbs (K = 2: double-
    double, 4: quad-double); returns (solutions, tracker output).  Each
    solution has x (complex128, the leading limbs), err (last Newton
    correction), res (max |p_i(x)| in complex128 at the refined point), rco
    (1 / condition number of the Jacobian, from numpy: an estimate in the
    sense of the reference's rco, not its LU-based estimator), the tracker's
    status and the refinement status; failed paths are included with their
    statuses so the caller can count them."""
    from .homotopy import total_degree_homotopy
    from .tracker import TrackParams, Tracker
    rng = np.random.default_rng(seed)
    hom, starts = total_degree_homotopy(system, rng)
    tr = Tracker(hom, deg=8)
    prm = params or TrackParams()
    out = tr.track(starts, K=1, params=prm)
    ref = tr.refine(out["endp"], K, iters=iters, tol=tol)
    xs = ref["x"][0] + 1j * ref["x"][K]  # leading limbs [npaths][n]
    sols = []
    for p in range(xs.shape[0]):
        x = xs[p]
        res = float(np.max(np.abs(system.eval(x)))) if np.all(np.isfinite(x)) else np.inf
        try:
            sv = np.linalg.svd(jacobian(system, x), compute_uv=False)
            rco = float(sv[-1] / sv[0]) if sv[0] > 0 else 0.0
        except np.linalg.LinAlgError:
            rco = 0.0
        sols.append(dict(t=1.0, m=1, x=x, err=float(ref["dx"][p]), rco=rco, res=res,
                         track_status=int(out["status"][p]), refine_status=int(ref["status"][p])))
    return sols, out

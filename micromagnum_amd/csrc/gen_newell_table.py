#!/usr/bin/env python3
"""Generate newell_table.inc: derivatives of the generalized Newell potential
and the large-distance asymptotic sums used by demag_tensor.cpp.

Mathematics (Newell, Williams & Dunlop, J. Geophys. Res. 98 (1993) 9551; the
generalized form of Krueger used by the reference,
src/magneto/mmm/demag/tensor_integrals.h:113-585):

  F(x, y, z) is a 6-fold antiderivative of 1/R with d^2x d^2y d^2z F = 1/R.
  f(a,b,c) = d^a_x d^b_y d^c_z F.  The two base functions are
     f(0,0,2) = Newell's f (roles of x and z swapped)  -> Nzz-type entries
     f(0,1,1) = Newell's g (x -> y -> z -> x)           -> Nyz-type entries
  Every other f(a,b,c) with a <= b <= c is obtained here by symbolic
  differentiation of those two; d_y f(0,0,2) = d_z f(0,1,1) holds exactly, so
  both routes agree.

  Far field: the 27-point second difference of f(0,0,2) / f(0,1,1) around a
  point far out along x is expanded in inverse powers of x.  For x > 0
     D = sum_k  Dx[x^-k] * Dy Dz [c_k(y, z)]
  where c_k are the (polynomial) coefficients of x^-k in the large-x
  expansion, Dx, Dy, Dz are exact centred second differences with steps
  lx, ly, lz, and Dx[x^-k] is put over the common denominator
  x^k (x^2 - lx^2)^k so that it is evaluated without cancellation.  Terms
  that the differences annihilate (polynomials of degree < 2 in an axis,
  x log x times such polynomials, log(rho) and atan(y/z) times x^{0,1}) are
  dropped before differencing.  The first three non-vanishing orders are
  kept, the truncation the reference's asymptotic sums use.

Output: C++ source included by demag_tensor.cpp (long double evaluation).
Run:  python gen_newell_table.py  (about a minute; needs sympy).
"""
import os
import re
import sys
import time

import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "newell_table.inc")

x, y, z = sp.symbols("x y z", positive=True)
Rs = sp.Symbol("R", positive=True)
R = sp.sqrt(x ** 2 + y ** 2 + z ** 2)

# Newell f for the zz entry (Newell 1993 eq. 27 with x <-> z)
F002 = (-(R * (x ** 2 + y ** 2 - 2 * z ** 2)) / 6 - x * y * z * sp.atan(x * y / (R * z))
        + x * (y ** 2 - z ** 2) / 2 * sp.atanh(x / R) + y * (x ** 2 - z ** 2) / 2 * sp.atanh(y / R))
# Newell g for the yz entry (Newell 1993 eq. 32 with x -> y -> z -> x)
F011 = (-2 * R * y * z - 3 * x * z ** 2 * sp.atan(x * y / (R * z)) - 3 * x * y ** 2 * sp.atan(x * z / (R * y))
        - x ** 3 * sp.atan(y * z / (R * x)) + 6 * x * y * z * sp.atanh(x / R)
        + (3 * x ** 2 * z - z ** 3) * sp.atanh(y / R) - y * (y ** 2 - 3 * x ** 2) * sp.atanh(z / R)) / 6


# ---------------------------------------------------------------------------
# which f(a,b,c) the tensor algorithm evaluates (a <= b <= c after sorting)
# ---------------------------------------------------------------------------
def series_orders():
    """(O,P,Q) offsets used by the Taylor-expanded second differences."""
    out = []
    out += [(2, 2, 2), (4, 2, 2), (2, 4, 2), (2, 2, 4)]                      # 3 axes expanded
    out += [(o, 2, 0) for o in range(2, 7)] + [(2, 4, 0)]                      # 2 axes, border
    out += [(2, 2, 0), (4, 2, 0), (2, 4, 0)]                                   # 2 axes, centre
    out += [(o, 0, 0) for o in range(2, 8)]                                    # 1 axis, border
    out += [(2, 0, 0), (4, 0, 0), (6, 0, 0)]                                   # 1 axis, centre
    return out


def needed():
    comps = set()
    for base in ((2, 0, 0), (1, 1, 0)):
        for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
            comps.add(tuple(base[p] for p in perm))
    s = set()
    for (o, p, q) in comps:
        for (O, P, Q) in series_orders():
            a, b, c = sorted((o + O, p + P, q + Q))
            s.add((a, b, c))
    return sorted(s)


_memo = {}


def simp(e):
    if e.has(sp.atan) or e.has(sp.atanh):
        return sp.simplify(e)
    return sp.factor(sp.cancel(sp.together(e)))


def deriv(a, b, c):
    """f(0,0,c), c >= 2, by z-derivatives of f(0,0,2)."""
    key = (a, b, c)
    if key in _memo:
        return _memo[key]
    assert a == 0 and b == 0 and c >= 2
    e = F002 if c == 2 else simp(sp.diff(deriv(0, 0, c - 1), z))
    _memo[key] = e
    return e


def deriv_any(a, b, c):
    """f(a,b,c) for sorted indices, choosing a differentiation path."""
    key = (a, b, c)
    if key in _memo:
        return _memo[key]
    if a >= 2 and b >= 2 and c >= 2:
        e = sp.factor(sp.together(sp.diff(1 / R, x, a - 2, y, b - 2, z, c - 2)))
        _memo[key] = e
        return e
    if b == 0:
        return deriv(a, b, c)
    # reach (a,b,c) from (0,1,1) by z, then y, then x derivatives
    cur = (0, 1, 1)
    e = F011
    _memo[cur] = e
    path = [("z", c - 1), ("y", b - 1), ("x", a)]
    for var, cnt in path:
        for _ in range(cnt):
            if var == "z":
                cur = (cur[0], cur[1], cur[2] + 1)
                sym = z
            elif var == "y":
                cur = (cur[0], cur[1] + 1, cur[2])
                sym = y
            else:
                cur = (cur[0] + 1, cur[1], cur[2])
                sym = x
            if cur in _memo:
                e = _memo[cur]
            else:
                e = simp(sp.diff(e, sym))
                _memo[cur] = e
    return e


# ---------------------------------------------------------------------------
# large-x asymptotic sums
# ---------------------------------------------------------------------------
e_ = sp.Symbol("e", positive=True)          # e = 1/x
lx, ly, lz = sp.symbols("lx ly lz", positive=True)


def large_x_coeffs(base, kmax):
    """Coefficients c_k(y,z) of x^-k (k = 1..kmax) of the base function for
    x -> +inf, after dropping terms annihilated by the differences."""
    rho2 = y ** 2 + z ** 2
    X = 1 / e_
    Rx = X * sp.sqrt(1 + e_ ** 2 * rho2)
    # atanh(x/R) = log(R + x) - log(rho) = log(2x) + log((1+sqrt(1+e^2 rho^2))/2) - log(rho)
    at_x = sp.log((1 + sp.sqrt(1 + e_ ** 2 * rho2)) / 2)        # log(2x), log(rho) dropped
    sub = {sp.atanh(x / R): at_x}
    expr = base.subs(sub)
    expr = expr.subs({sp.sqrt(x ** 2 + y ** 2 + z ** 2): Rx}).subs(x, X)
    ser = sp.series(expr, e_, 0, kmax + 1).removeO()
    ser = sp.expand(ser)
    out = {}
    for k in range(1, kmax + 1):
        ck = sp.simplify(ser.coeff(e_, k))
        out[k] = sp.expand(ck)
    return out


def d2(expr, var, step):
    return sp.expand(expr.subs(var, var + step) - 2 * expr + expr.subs(var, var - step))


def large_x_terms(base, nterms=3):
    coeffs = large_x_coeffs(base, 2 * nterms + 5)
    terms = []
    for k in sorted(coeffs):
        c = coeffs[k]
        if c == 0:
            continue
        dyz = sp.factor(d2(d2(c, y, ly), z, lz))
        if dyz == 0:
            continue
        dxk = sp.factor(sp.together(1 / (x - lx) ** k - 2 / x ** k + 1 / (x + lx) ** k))
        terms.append((k, dxk, dyz))
        if len(terms) == nterms:
            break
    return terms


# ---------------------------------------------------------------------------
# code emission
# ---------------------------------------------------------------------------
class LPrinter(C99CodePrinter):
    def _print_Pow(self, expr):
        b, ex = expr.as_base_exp()
        if ex.is_Integer:
            n = int(ex)
            if n == 0:
                return "1.0L"
            inv = n < 0
            n = abs(n)
            bs = self._print(b)
            if n == 1:
                s = "(%s)" % bs
            else:
                s = "ipow(%s, %d)" % (bs, n)
            return "(1.0L/%s)" % s if inv else s
        if ex == sp.Rational(1, 2):
            return "sqrtl(%s)" % self._print(b)
        if ex == -sp.Rational(1, 2):
            return "(1.0L/sqrtl(%s))" % self._print(b)
        return "powl(%s, %s)" % (self._print(b), self._print(ex))

    def _print_Rational(self, expr):
        return "(%d.0L/%d.0L)" % (expr.p, expr.q)

    def _print_Integer(self, expr):
        return "%d.0L" % int(expr)

    def _print_Float(self, expr):
        return "%sL" % C99CodePrinter._print_Float(self, expr)

    def _print_Function(self, expr):
        name = expr.func.__name__
        m = {"atan": "atanl", "atanh": "atanhl", "log": "logl", "sqrt": "sqrtl"}
        if name in m:
            return "%s(%s)" % (m[name], ", ".join(self._print(a) for a in expr.args))
        return super()._print_Function(expr)


P = LPrinter()


def emit_function(name, expr, args="long double x, long double y, long double z"):
    expr = expr.subs(sp.sqrt(x ** 2 + y ** 2 + z ** 2), Rs)
    repl, red = sp.cse(expr, symbols=sp.numbered_symbols("t"), optimizations="basic")
    lines = ["static long double %s(%s)" % (name, args), "{"]
    if "R" in [str(s) for s in expr.free_symbols]:
        lines.append("    const long double R = sqrtl(x*x + y*y + z*z);")
    for s, v in repl:
        lines.append("    const long double %s = %s;" % (s, P.doprint(v)))
    lines.append("    return %s;" % P.doprint(red[0]))
    lines.append("}")
    return "\n".join(lines)


def main():
    t0 = time.time()
    need = needed()
    print("need %d functions: %s" % (len(need), need), flush=True)
    parts = ["/* Generated by gen_newell_table.py -- do not edit. */", ""]
    parts.append("static inline long double ipow(long double b, int n) "
                 "{ long double r = 1.0L; for (int i = 0; i < n; ++i) r *= b; return r; }")
    parts.append("")
    names = []
    for (a, b, c) in need:
        t = time.time()
        ex = deriv_any(a, b, c)
        nm = "nf_%d%d%d" % (a, b, c)
        parts.append(emit_function(nm, ex))
        parts.append("")
        names.append(((a, b, c), nm))
        print("  f(%d,%d,%d) %.1fs" % (a, b, c, time.time() - t), flush=True)
    # dispatcher
    parts.append("static bool newell_f_sorted(int a, int b, int c, long double x, long double y, "
                 "long double z, long double *out)")
    parts.append("{")
    for (a, b, c), nm in names:
        parts.append("    if (a == %d && b == %d && c == %d) { *out = %s(x, y, z); return true; }"
                     % (a, b, c, nm))
    parts.append("    return false;")
    parts.append("}")
    parts.append("")
    # large-x sums
    for label, base in (("002", F002), ("011", F011)):
        t = time.time()
        terms = large_x_terms(base, 3)
        body = ["static long double nlarge_%s(long double lx, long double ly, long double lz, "
                "long double x, long double y, long double z)" % label, "{", "    long double s = 0.0L;"]
        for (k, dxk, dyz) in terms:
            body.append("    /* order x^-%d */" % k)
            body.append("    s += (%s) * (%s);" % (P.doprint(dxk), P.doprint(dyz)))
        body.append("    return s;")
        body.append("}")
        parts.append("\n".join(body))
        parts.append("")
        print("  large-x %s: orders %s  %.1fs" % (label, [k for k, _, _ in terms], time.time() - t),
              flush=True)
    text = "\n".join(parts) + "\n"
    # long double versions of the libm calls the C99 printer emits
    for fn in ("atanh", "atan", "sqrt", "log"):
        text = re.sub(r"\b%s\(" % fn, "%sl(" % fn, text)
    with open(OUT, "w") as f:
        f.write(text)
    print("wrote %s in %.1fs" % (OUT, time.time() - t0))


if __name__ == "__main__":
    sys.setrecursionlimit(100000)
    main()

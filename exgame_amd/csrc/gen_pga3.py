"""Generate pga3_gen.h: closed-form PGA3 operations for the device kernels.

Run `python gen_pga3.py` to rewrite pga3_gen.h next to this script. Each formula
is the exact polynomial expansion of the reference's multivector expression
(no unit-norm simplification), derived symbolically from the sign law

  * e1, e2, e3 = bits 0..2, degenerate e0 = bit 3
  * blade product = permutation parity of merging into ascending order, e0² = 0
  * dual sign: b ∧ dual(b) = +I
  * screw markup: linear on e0∧eᵢ, angular on the Euclidean complement × σ = −1

and emitted as templates over the scalar type (double, float and the dual
numbers the Jacobian kernel uses).

Layouts: motor = 8 even blades ascending {1,e12,e13,e23,e01,e02,e03,e0123};
screw = (linear|force xyz, angular|torque xyz); a point is carried as its
weight w (e123) and the three sign-corrected trivector coordinates q (position
= q / w).
"""
from __future__ import annotations

import os

import sympy as sp

ZERO = 0b1000


def popc(x):
    return bin(x).count("1")


def blade_prod(a, b):
    if a & b & ZERO:
        return 0, 0
    sw, aa = 0, a >> 1
    while aa:
        sw += popc(aa & b)
        aa >>= 1
    return (1 if sw % 2 == 0 else -1), a ^ b


def csign(a, b):
    inv = 0
    for bi in range(4):
        if (a >> bi) & 1:
            inv += popc(b & ((1 << bi) - 1))
    return 1 if inv % 2 == 0 else -1


def euclid(i):
    v = 1 << i
    c = 0b0111 ^ v
    return c, csign(v, c)


def radical(i):
    s, m = blade_prod(ZERO, 1 << i)
    return m, s


def pointb(i):
    em, es = euclid(i)
    s, m = blade_prod(ZERO, em)
    return m, es * s


def sigma():
    s_w, _ = blade_prod(4, 1)
    _, s_e1 = euclid(1)
    mb, sb = euclid(2)
    mp, sp_ = pointb(0)
    _, sd = pointb(1)
    s1, _ = blade_prod(mb, mp)
    s2, _ = blade_prod(mp, mb)
    return s_e1 * s_w * sd * ((s1 - s2) // 2) * sb * sp_


SIG = sigma()
EVEN = [0, 3, 5, 6, 9, 10, 12, 15]


def gp(a, b):
    out = [sp.Integer(0)] * 16
    for i in range(16):
        if a[i] == 0:
            continue
        for j in range(16):
            if b[j] == 0:
                continue
            s, c = blade_prod(i, j)
            if s:
                out[c] += s * a[i] * b[j]
    return [sp.expand(x) for x in out]


def rev(a):
    return [(-x if ((popc(b) * (popc(b) - 1)) // 2) % 2 else x) for b, x in enumerate(a)]


def dual(a):
    out = [sp.Integer(0)] * 16
    for s in range(16):
        out[15 ^ s] += csign(s, 15 ^ s) * a[s]
    return out


def undual(a):
    out = [sp.Integer(0)] * 16
    for s in range(16):
        out[15 ^ s] += csign(15 ^ s, s) * a[s]
    return out


def wedge(a, b):
    out = [sp.Integer(0)] * 16
    for i in range(16):
        if a[i] == 0:
            continue
        for j in range(16):
            if b[j] == 0:
                continue
            s, c = blade_prod(i, j)
            if s and popc(c) == popc(i) + popc(j):
                out[c] += s * a[i] * b[j]
    return [sp.expand(x) for x in out]


def motor(sym):
    m = [sp.Integer(0)] * 16
    for k, b in enumerate(EVEN):
        m[b] = sym[k]
    return m


def screw(lin, ang):
    m = [sp.Integer(0)] * 16
    for i in range(3):
        mk, s = radical(i)
        m[mk] += s * lin[i]
        mk, s = euclid(i)
        m[mk] += s * SIG * ang[i]
    return m


def screw_read(m):
    out = []
    for i in range(3):
        mk, s = radical(i)
        out.append(s * m[mk])
    for i in range(3):
        mk, s = euclid(i)
        out.append(s * SIG * m[mk])
    return out


def point(w, q):
    m = [sp.Integer(0)] * 16
    m[7] = w
    for i in range(3):
        mk, s = pointb(i)
        m[mk] += s * q[i]
    return m


def point_read(m):
    return m[7], [pointb(i)[1] * m[pointb(i)[0]] for i in range(3)]


# ───────────────────────── C emission ─────────────────────────
def cexpr(e, names):
    e = sp.expand(e)
    if e == 0:
        return "S(0.0)"
    s = sp.ccode(e)
    for k, v in names.items():
        s = s.replace(k, v)
    return s


class Emitter:
    def __init__(self):
        self.lines = []

    def fn(self, sig, body_lines, tmpl="template <class S>", comment=None):
        if comment:
            self.lines.extend("// " + c for c in comment)
        self.lines.append(tmpl)
        self.lines.append(f"__host__ __device__ inline {sig} {{")
        for l in body_lines:
            self.lines.append("    " + l)
        self.lines.append("}")
        self.lines.append("")


def poly_c(e):
    """Polynomial to C with products written as a*b (sympy uses pow for squares)."""
    e = sp.expand(e)
    if e == 0:
        return "S(0.0)"
    terms = []
    for t in sp.Add.make_args(e):
        coeff, factors = t.as_coeff_mul()
        parts = []
        for f in factors:
            if f.is_Pow:
                base, ex = f.args
                parts += [str(base)] * int(ex)
            else:
                parts.append(str(f))
        c = int(coeff) if coeff == int(coeff) else None
        if c is None:
            raise ValueError(t)
        body = "*".join(parts)
        if abs(c) != 1:
            body = f"S({float(abs(c))})*" + body
        terms.append(("-" if c < 0 else "+", body))
    s = ""
    for i, (sg, body) in enumerate(terms):
        if i == 0:
            s += ("-" if sg == "-" else "") + body
        else:
            s += f" {sg} {body}"
    return s


def sub_index(s, var, n):
    for k in reversed(range(n)):
        s = s.replace(f"{var}{k}", f"{var}[{k}]")
    return s


def main():
    em = Emitter()
    a = sp.symbols("a0:8")
    b = sp.symbols("b0:8")
    m = sp.symbols("m0:8")

    # motor product c = a·b
    c = gp(motor(a), motor(b))
    body = [f"c[{k}] = {sub_index(sub_index(poly_c(c[blade]), 'a', 8), 'b', 8)};" for k, blade in enumerate(EVEN)]
    em.fn("void motor_mul(const S a[8], const S b[8], S c[8])", body)

    # reverse
    r = rev(motor(a))
    body = [f"r[{k}] = {sub_index(poly_c(r[blade]), 'a', 8)};" for k, blade in enumerate(EVEN)]
    em.fn("void motor_rev(const S a[8], S r[8])", body)

    # adjoint matrix: contra(t) = narrow(M t M~) as 6x6 acting on (lin, ang)
    M = motor(m)
    Mr = rev(M)
    body = []
    for col in range(6):
        lin = [1 if col == i else 0 for i in range(3)]
        ang = [1 if col == 3 + i else 0 for i in range(3)]
        t = screw(lin, ang)
        out = screw_read(gp(gp(M, t), Mr))
        for row in range(6):
            body.append(f"ad[{row}][{col}] = {sub_index(poly_c(out[row]), 'm', 8)};")
    em.fn("void ad_matrix(const S m[8], S ad[6][6])", body)

    # coadjoint matrix: co(w) = narrow(undual(M dual(w) M~))
    body = []
    for col in range(6):
        lin = [1 if col == i else 0 for i in range(3)]
        ang = [1 if col == 3 + i else 0 for i in range(3)]
        w = screw(lin, ang)
        out = screw_read(undual(gp(gp(M, dual(w)), Mr)))
        for row in range(6):
            body.append(f"cad[{row}][{col}] = {sub_index(poly_c(out[row]), 'm', 8)};")
    em.fn("void coad_matrix(const S m[8], S cad[6][6])", body)

    # point sandwich: M·P(1, p)·M~ for a weight-1 point; returns weight and q
    p = sp.symbols("p0:3")
    P = point(1, list(p))
    out = gp(gp(M, P), Mr)
    w, q = point_read(out)
    body = [f"w = {sub_index(poly_c(w), 'm', 8)};"]
    for i in range(3):
        s = poly_c(q[i])
        s = sub_index(sub_index(s, "m", 8), "p", 3)
        body.append(f"q[{i}] = {s};")
    body = ["using R = typename RealOf<S>::type;"] + [b.replace("S(2.0)", "R(2.0)") for b in body]
    em.fn("void point_sandwich(const S m[8], const P p[3], S& w, S q[3])", body, tmpl="template <class S, class P = S>",
          comment=["The point p may be of a plain type P while the motor carries tangents (S = Dual):",
                   "its terms then multiply by a constant."])

    # join of two points (weights wa, wb; coords qa, qb) → line bivector (6 grade-2 comps, ascending blade)
    wa, wb = sp.symbols("wa wb")
    qa = sp.symbols("qa0:3")
    qb = sp.symbols("qb0:3")
    L = undual(wedge(dual(point(wa, list(qa))), dual(point(wb, list(qb)))))
    G2 = [3, 5, 6, 9, 10, 12]
    body = []
    for k, blade in enumerate(G2):
        s = poly_c(L[blade])
        s = sub_index(sub_index(s, "qa", 3), "qb", 3)
        body.append(f"l[{k}] = {s};")
    em.fn("void join_points(const S& wa, const S qa[3], const S& wb, const S qb[3], S l[6])", body)

    # line norm² = <L L~>_0 ; from_line: narrow(dual(L)) read as a wrench; direction = euclid reading
    ls = sp.symbols("l0:6")
    Lm = [sp.Integer(0)] * 16
    for k, blade in enumerate(G2):
        Lm[blade] = ls[k]
    n2 = gp(Lm, rev(Lm))[0]
    em.fn("S line_norm2(const S l[6])", [f"return {sub_index(poly_c(n2), 'l', 6)};"])
    wl = screw_read(dual(Lm))
    em.fn("void wrench_from_line(const S l[6], S w[6])",
          [f"w[{i}] = {sub_index(poly_c(wl[i]), 'l', 6)};" for i in range(6)])
    dirs = [euclid(i)[1] * Lm[euclid(i)[0]] for i in range(3)]
    em.fn("void line_direction(const S l[6], S d[3])", [f"d[{i}] = {sub_index(poly_c(dirs[i]), 'l', 6)};" for i in range(3)])

    # pairing(w, t) = <w ∧ dual(t)>_I
    ws = sp.symbols("w0:6")
    ts = sp.symbols("t0:6")
    W = screw(ws[:3], ws[3:])
    T = screw(ts[:3], ts[3:])
    pr = wedge(W, dual(T))[15]
    em.fn("S pairing(const S w[6], const S t[6])", [f"return {sub_index(sub_index(poly_c(pr), 'w', 6), 't', 6)};"])

    # velocity_at(t, P): ½[T,P] read as a Direction (no weight division)
    pw = sp.symbols("pw")
    pq = sp.symbols("pq0:3")
    Pp = point(pw, list(pq))
    comm = [sp.Rational(1, 2) * (x - y) for x, y in zip(gp(T, Pp), gp(Pp, T))]
    _, v = point_read(comm)
    body = []
    for i in range(3):
        e = sp.expand(2 * v[i])  # emit 2·v, scale by ½ after
        s = poly_c(e)
        s = sub_index(sub_index(s, "t", 6), "pq", 3)
        body.append(f"v[{i}] = ({s}) * S(0.5);")
    em.fn("void velocity_at(const S t[6], const S& pw, const S pq[3], S v[3])", body)

    # coadjoint bracket ½·undual([t, dual(w)])
    dw = dual(W)
    raw = undual([x - y for x, y in zip(gp(T, dw), gp(dw, T))])
    out = screw_read(raw)
    body = [f"o[{i}] = ({sub_index(sub_index(poly_c(out[i]), 'w', 6), 't', 6)}) * S(0.5);" for i in range(6)]
    em.fn("void coad_bracket(const S w[6], const S t[6], S o[6])", body)

    # exp pieces for a bivector given as a screw (lin, ang): u = -(B²)_0, p = (B²)_I, I·B
    Bm = screw(ts[:3], ts[3:])
    b2 = gp(Bm, Bm)
    I = [sp.Integer(0)] * 16
    I[15] = sp.Integer(1)
    ib = gp(I, Bm)
    body = [f"u = {sub_index(poly_c(-b2[0]), 't', 6)};", f"p = {sub_index(poly_c(b2[15]), 't', 6)};"]
    for k, blade in enumerate(EVEN):
        body.append(f"bv[{k}] = {sub_index(poly_c(Bm[blade]), 't', 6)};")
        body.append(f"ibv[{k}] = {sub_index(poly_c(ib[blade]), 't', 6)};")
    em.fn("void exp_parts(const S t[6], S& u, S& p, S bv[8], S ibv[8])", body)

    # the reference's Pod layout of a screw: its grade-2 store, blades ascending
    # {e12, e13, e23, e01, e02, e03} with the markup signs (what
    # crates/newton/build.rs:106-152 probes at build time), to and from (lin, ang)
    cs = sp.symbols("c0:6")
    Sm = screw(cs[:3], cs[3:])
    em.fn("void screw_to_pod(const S c[6], S pod[6])",
          [f"pod[{k}] = {sub_index(poly_c(Sm[blade]), 'c', 6)};" for k, blade in enumerate(G2)])
    qs = sp.symbols("q0:6")
    Qm = [sp.Integer(0)] * 16
    for k, blade in enumerate(G2):
        Qm[blade] = qs[k]
    back = screw_read(Qm)
    em.fn("void screw_from_pod(const S q[6], S c[6])",
          [f"c[{i}] = {sub_index(poly_c(back[i]), 'q', 6)};" for i in range(6)])

    hdr = [
        "// GENERATED by gen_pga3.py — do not edit by hand.",
        "// Closed-form PGA3 operations (exact polynomial expansions of the reference's",
        "// multivector expressions) for the device kernels, templated on the scalar.",
        "#pragma once",
        '#include "dual.h"',
        "",
        "namespace exg {",
        "",
    ]
    text = "\n".join(hdr + em.lines + ["}  // namespace exg", ""])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pga3_gen.h")
    with open(path, "w") as f:
        f.write(text)
    print("wrote", path, len(text), "bytes")


if __name__ == "__main__":
    main()

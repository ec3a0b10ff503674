// ORACLE — test infrastructure only. A CPU restatement, in plain C++ f64, of the
// reference's projective geometric algebra PGA3 (crates/peano, crates/clifford).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// anything built from oracle/; the product never does.
//
// Every sign here is computed from the reference's one sign law, not tabulated
// by hand:
//   * generator chain Gen<0,Gen<1,Gen<1,Gen<1,Nil>>>>: e1 = bit 0, e2 = bit 1,
//     e3 = bit 2, e0 = bit 3 (degenerate)           crates/clifford/src/algebra/mod.rs:84
//   * blade_prod: permutation parity x metric        crates/clifford/src/algebra/mod.rs:99-120
//   * complement_sign with DUAL_ORIENTATION = +1     crates/clifford/src/algebra/mod.rs:124-147
//   * dual markup: axis_vec / euclid / radical / point / angular_embed_sign
//                                                    crates/clifford/src/algebra/markup.rs:13-100
// Products loop blade i ascending, then blade j ascending, exactly as gp_store
// (crates/clifford/src/algebra/store.rs:439-456). The full-Mv gp of the reference
// (Cayley–Dickson recursion) sums in another order; bit-exactness against Rust is
// therefore "parity unpinned" and every test uses a tolerance.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace oracle {

// ───────────────────────── sign law on bitmasks ─────────────────────────
constexpr int ZERO_MASK = 0b1000;  // e0 is the outermost generator: bit 3
constexpr int NEG_MASK = 0;
constexpr int DIM = 16;
constexpr int TOP = 15;

constexpr int popc(int x) { int n = 0; while (x) { n += x & 1; x >>= 1; } return n; }

struct BladeProd { int sign; int mask; };

// crates/clifford/src/algebra/mod.rs:99-120
constexpr BladeProd blade_prod(int a, int b) {
    int shared = a & b;
    if (shared & ZERO_MASK) return {0, 0};
    int neg = popc(shared & NEG_MASK) & 1;
    int swaps = 0;
    int aa = a >> 1;
    while (aa) { swaps += popc(aa & b); aa >>= 1; }
    return {((swaps + neg) & 1) == 0 ? 1 : -1, a ^ b};
}

// crates/clifford/src/algebra/mod.rs:128-147 (DUAL_ORIENTATION = +1)
constexpr int complement_sign(int a, int b) {
    int inv = 0;
    for (int bi = 0; bi < 32; ++bi)
        if ((a >> bi) & 1) inv += popc(b & ((1 << bi) - 1));
    return (inv % 2 == 0) ? 1 : -1;
}

// ───────────────────────── dual markup (markup.rs) ─────────────────────────
struct MS { int mask; int sign; };

constexpr int axis_vec(int i) {  // markup.rs:13-26
    int e = (DIM - 1) & ~ZERO_MASK;
    int k = i;
    while (true) {
        int low = e & -e;
        if (k == 0) return low;
        e ^= low;
        --k;
    }
}
constexpr MS euclid_blade(int i) {  // markup.rs:28-35
    int i3 = (DIM - 1) & ~ZERO_MASK;
    int v = axis_vec(i);
    int comp = i3 ^ v;
    return {comp, complement_sign(v, comp)};
}
constexpr MS radical_blade(int i) {  // markup.rs:37-43
    int v = axis_vec(i);
    auto p = blade_prod(ZERO_MASK, v);
    return {p.mask, p.sign};
}
constexpr MS point_blade(int i) {  // markup.rs:45-50
    auto ec = euclid_blade(i);
    auto p = blade_prod(ZERO_MASK, ec.mask);
    return {p.mask, ec.sign * p.sign};
}
constexpr int angular_embed_sign() {  // markup.rs:75-93
    int vz = axis_vec(2), vx = axis_vec(0);
    auto w = blade_prod(vz, vx);
    auto e1 = euclid_blade(1);
    auto b = euclid_blade(2);
    auto p = point_blade(0);
    auto d = point_blade(1);
    auto s1 = blade_prod(b.mask, p.mask);
    auto s2 = blade_prod(p.mask, b.mask);
    int hs = (s1.sign - s2.sign) / 2;
    (void)w.mask; (void)e1.mask; (void)d.mask;
    return e1.sign * w.sign * d.sign * hs * b.sign * p.sign;
}
constexpr MS angular_blade(int i) {  // markup.rs:96-99
    auto e = euclid_blade(i);
    return {e.mask, e.sign * angular_embed_sign()};
}
constexpr int I3_BLADE = (DIM - 1) & ~ZERO_MASK;  // e123: the point weight

static_assert(angular_embed_sign() == -1, "sigma derived from the +I anchor");

// blade-presence masks (bit b set = blade b may be nonzero)
constexpr unsigned EVEN = []() { unsigned m = 0; for (int b = 0; b < DIM; ++b) if (popc(b) % 2 == 0) m |= 1u << b; return m; }();
constexpr unsigned GRADE2 = []() { unsigned m = 0; for (int b = 0; b < DIM; ++b) if (popc(b) == 2) m |= 1u << b; return m; }();
constexpr unsigned GRADE3 = []() { unsigned m = 0; for (int b = 0; b < DIM; ++b) if (popc(b) == 3) m |= 1u << b; return m; }();
constexpr unsigned FULL = 0xFFFFu;

// ───────────────────────── multivector ─────────────────────────
template <class S>
struct Mv {
    S c[DIM];
    Mv() { for (auto& x : c) x = S(0.0); }
    static Mv basis(int blade, S v) { Mv m; m.c[blade] = v; return m; }
    Mv operator+(const Mv& o) const { Mv r; for (int i = 0; i < DIM; ++i) r.c[i] = c[i] + o.c[i]; return r; }
    Mv operator-(const Mv& o) const { Mv r; for (int i = 0; i < DIM; ++i) r.c[i] = c[i] - o.c[i]; return r; }
    Mv operator-() const { Mv r; for (int i = 0; i < DIM; ++i) r.c[i] = -c[i]; return r; }
    Mv scale(const S& s) const { Mv r; for (int i = 0; i < DIM; ++i) r.c[i] = c[i] * s; return r; }
};

// Geometric product restricted to the blades each operand can hold: a loop over
// i (blade of a) ascending, j (blade of b) ascending, out[c] += ±a_i·b_j — the
// order of gp_store.
template <unsigned AM = FULL, unsigned BM = FULL, class S>
Mv<S> gp(const Mv<S>& a, const Mv<S>& b) {
    Mv<S> out;
    for (int i = 0; i < DIM; ++i) {
        if (!((AM >> i) & 1)) continue;
        for (int j = 0; j < DIM; ++j) {
            if (!((BM >> j) & 1)) continue;
            auto p = blade_prod(i, j);
            if (p.sign == 0) continue;
            S term = a.c[i] * b.c[j];
            out.c[p.mask] = out.c[p.mask] + ((p.sign < 0) ? -term : term);
        }
    }
    return out;
}

template <class S>
Mv<S> reverse(const Mv<S>& a) {  // (-1)^{g(g-1)/2}, mv.rs:144-146
    Mv<S> r;
    for (int b = 0; b < DIM; ++b) {
        int g = popc(b);
        r.c[b] = (((g * (g - 1)) / 2) & 1) ? -a.c[b] : a.c[b];
    }
    return r;
}

template <class S>
Mv<S> dual(const Mv<S>& a) {  // mv.rs:221-233
    Mv<S> out;
    for (int s = 0; s < DIM; ++s) {
        int comp = TOP ^ s;
        out.c[comp] = complement_sign(s, comp) < 0 ? -a.c[s] : a.c[s];
    }
    return out;
}

template <class S>
Mv<S> undual(const Mv<S>& a) {  // mv.rs:238-249
    Mv<S> out;
    for (int s = 0; s < DIM; ++s) {
        int comp = TOP ^ s;
        out.c[comp] = complement_sign(comp, s) < 0 ? -a.c[s] : a.c[s];
    }
    return out;
}

// Outer product: the grade-(ga+gb) section of gp (mv.rs:186-214).
template <class S>
Mv<S> wedge(const Mv<S>& a, const Mv<S>& b) {
    Mv<S> out;
    for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) {
            auto p = blade_prod(i, j);
            if (p.sign != 0 && popc(p.mask) == popc(i) + popc(j)) {
                S term = a.c[i] * b.c[j];
                out.c[p.mask] = out.c[p.mask] + (p.sign < 0 ? -term : term);
            }
        }
    return out;
}

// Regressive product: undual(dual(a) ∧ dual(b)) (mv.rs:251-253).
template <class S>
Mv<S> meet(const Mv<S>& a, const Mv<S>& b) { return undual(wedge(dual(a), dual(b))); }

template <class S>
S norm_squared(const Mv<S>& a) { return gp(a, reverse(a)).c[0]; }  // mv.rs:447-449

// ───────────────────────── V3 helpers ─────────────────────────
template <class S>
struct V3 {
    S v[3];
    V3() { v[0] = v[1] = v[2] = S(0.0); }
    V3(S a, S b, S c) { v[0] = a; v[1] = b; v[2] = c; }
    S& operator[](int i) { return v[i]; }
    const S& operator[](int i) const { return v[i]; }
    V3 operator+(const V3& o) const { return V3(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
    V3 operator-(const V3& o) const { return V3(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
    V3 operator-() const { return V3(-v[0], -v[1], -v[2]); }
    V3 scale(const S& s) const { return V3(v[0] * s, v[1] * s, v[2] * s); }
    // peano's Dot recurses from the right: x0·y0 + (x1·y1 + (x2·y2 + 0)) (peano/src/vec.rs:129-138)
    S dot(const V3& o) const { return v[0] * o[0] + (v[1] * o[1] + (v[2] * o[2] + S(0.0))); }
};

template <class S, class F>
V3<S> read_dual(const Mv<S>& m, F blade) {  // mv.rs:402-412
    V3<S> out;
    for (int i = 0; i < 3; ++i) {
        MS ms = blade(i);
        out[i] = ms.sign < 0 ? -m.c[ms.mask] : m.c[ms.mask];
    }
    return out;
}
template <class S, class F>
void write_dual(Mv<S>& m, const V3<S>& v, F blade) {  // mv.rs:540-553
    for (int i = 0; i < 3; ++i) {
        MS ms = blade(i);
        m.c[ms.mask] = m.c[ms.mask] + (ms.sign < 0 ? -v[i] : v[i]);
    }
}

// ───────────────────────── Point / Line / Direction ─────────────────────────
template <class S>
Mv<S> make_point(const V3<S>& v) {  // mv.rs:432-442: origin e123 + point markup
    Mv<S> p = Mv<S>::basis(I3_BLADE, S(1.0));
    write_dual(p, v, point_blade);
    return p;
}
template <class S>
V3<S> point_coords(const Mv<S>& p) {  // mv.rs:417-430: projective normalisation
    S winv = S(1.0) / p.c[I3_BLADE];
    V3<S> out;
    for (int i = 0; i < 3; ++i) {
        MS ms = point_blade(i);
        S c = p.c[ms.mask] * winv;
        out[i] = ms.sign < 0 ? -c : c;
    }
    return out;
}
template <class S>
V3<S> direction_coords(const Mv<S>& m) { return read_dual(m, point_blade); }  // mv.rs:489-491
template <class S>
V3<S> line_direction(const Mv<S>& m) { return read_dual(m, euclid_blade); }  // mv.rs:493-495
template <class S>
Mv<S> join(const Mv<S>& p, const Mv<S>& q) { return meet(p, q); }  // objects.rs:46-48

// ───────────────────────── Motor (even carrier) ─────────────────────────
constexpr std::array<int, 8> EVEN_BLADES = {0, 3, 5, 6, 9, 10, 12, 15};  // motor.rs:119-152
constexpr std::array<int, 6> G2_BLADES = {3, 5, 6, 9, 10, 12};

template <class S>
Mv<S> motor_from8(const S* c8) { Mv<S> m; for (int k = 0; k < 8; ++k) m.c[EVEN_BLADES[k]] = c8[k]; return m; }
template <class S>
void motor_to8(const Mv<S>& m, S* c8) { for (int k = 0; k < 8; ++k) c8[k] = m.c[EVEN_BLADES[k]]; }
template <class S>
Mv<S> motor_identity() { return Mv<S>::basis(0, S(1.0)); }

// compose(self, rhs) = rhs·self: "first self, then rhs" (motor.rs:149-158)
template <class S>
Mv<S> compose(const Mv<S>& self, const Mv<S>& rhs) { return gp<EVEN, EVEN>(rhs, self); }
template <class S>
Mv<S> motor_inverse(const Mv<S>& m) { return reverse(m); }  // motor.rs:159-165
// sandwich M·X·M̃ (mv.rs:286-289) — M even, X arbitrary
template <class S>
Mv<S> conjugate_mv(const Mv<S>& m, const Mv<S>& x) { return gp<FULL, EVEN>(gp<EVEN, FULL>(m, x), reverse(m)); }

// ───────────────────────── Study-form scalars (peano/src/scalar.rs:60-135) ─────────────────────────
inline double standard_part(double x) { return x; }
inline bool is_effective_zero(double x) { return std::fabs(x) < 1e-9; }  // support.rs:38-42

template <class S> S sqrt_s(const S& x);
template <class S> S sin_s(const S& x);
template <class S> S cos_s(const S& x);
template <> inline double sqrt_s(const double& x) { return std::sqrt(x); }
template <> inline double sin_s(const double& x) { return std::sin(x); }
template <> inline double cos_s(const double& x) { return std::cos(x); }

template <class S>
S sinc_sq(const S& u) {  // scalar.rs:67-81
    if (is_effective_zero(standard_part(u))) {
        const S c1 = S(1.0 / 6.0), c2 = S(1.0 / 120.0), c3 = S(1.0 / 5040.0);
        return S(1.0) - u * (c1 - u * (c2 - u * c3));
    }
    S l = sqrt_s(u);
    return sin_s(l) / l;
}
template <class S>
S cos_sq(const S& u) {  // scalar.rs:84-97
    if (is_effective_zero(standard_part(u))) {
        const S c1 = S(1.0 / 2.0), c2 = S(1.0 / 24.0), c3 = S(1.0 / 720.0);
        return S(1.0) - u * (c1 - u * (c2 - u * c3));
    }
    return cos_s(sqrt_s(u));
}
template <class S>
S dsinc_sq(const S& u) {  // scalar.rs:100-118
    if (is_effective_zero(standard_part(u * S(1.0 / 100000.0)))) {
        const S c0 = S(1.0 / 6.0), c1 = S(1.0 / 60.0), c2 = S(1.0 / 1680.0), c3 = S(1.0 / 90720.0);
        return -c0 + u * (c1 - u * (c2 - u * c3));
    }
    S l = sqrt_s(u);
    return (l * cos_s(l) - sin_s(l)) / (S(2.0) * l * l * l);
}

// Motor::exp of a bivector via Study numbers (motor.rs:194-221).
template <class S>
Mv<S> motor_exp(const Mv<S>& bs) {
    const S half = S(1.0 / 2.0);
    Mv<S> b2 = gp<GRADE2, GRADE2>(bs, bs);
    S u = -b2.c[0];
    S p = b2.c[TOP];
    S cu = cos_sq(u), su = sinc_sq(u), du = dsinc_sq(u);
    Mv<S> i_mv = Mv<S>::basis(TOP, S(1.0));
    Mv<S> ib = gp<(1u << TOP), GRADE2>(i_mv, bs);
    Mv<S> scalar_part = Mv<S>::basis(0, cu) + i_mv.scale(p * half * su);
    Mv<S> bivector_part = bs.scale(su) - ib.scale(p * du);
    Mv<S> out = scalar_part + bivector_part;
    // the even carrier holds only even blades
    for (int b = 0; b < DIM; ++b) if (popc(b) % 2) out.c[b] = S(0.0);
    return out;
}

// ───────────────────────── Screw: Twist (Contra) / Wrench (Co) ─────────────────────────
// Stored as the grade-2 section of an Mv. Constructor coordinate order (linear|force,
// angular|torque) is the order of Twist::new / Wrench::new (screw.rs:235-249, 279-290).
template <class S>
Mv<S> screw_new(const V3<S>& lin, const V3<S>& ang) {
    Mv<S> m;
    write_dual(m, lin, radical_blade);   // translation_store (mv.rs:560-562)
    write_dual(m, ang, angular_blade);   // angular_store (mv.rs:565-567)
    return m;
}
template <class S>
V3<S> screw_linear(const Mv<S>& m) { return read_dual(m, radical_blade); }   // line_moment_store
template <class S>
V3<S> screw_angular(const Mv<S>& m) { return read_dual(m, angular_blade); }  // angular_from_store
template <class S>
Mv<S> narrow2(const Mv<S>& m) { Mv<S> r; for (int b : G2_BLADES) r.c[b] = m.c[b]; return r; }

template <class S>
void screw_to6(const Mv<S>& m, S* o6) {
    V3<S> l = screw_linear(m), a = screw_angular(m);
    for (int i = 0; i < 3; ++i) { o6[i] = l[i]; o6[3 + i] = a[i]; }
}
template <class S>
Mv<S> screw_from6(const S* i6) { return screw_new(V3<S>(i6[0], i6[1], i6[2]), V3<S>(i6[3], i6[4], i6[5])); }

// Contra: Adjoint M·t·M̃ (screw.rs:90-94). Co: undual(M·dual(w)·M̃) (screw.rs:96-101).
template <class S>
Mv<S> transport_contra(const Mv<S>& t, const Mv<S>& m) { return narrow2(conjugate_mv(m, narrow2(t))); }
template <class S>
Mv<S> transport_co(const Mv<S>& w, const Mv<S>& m) { return narrow2(undual(conjugate_mv(m, dual(narrow2(w))))); }

// Twist::exp(dt) = Motor::exp(self·(dt·½)) (screw.rs:254-260)
template <class S>
Mv<S> twist_exp(const Mv<S>& t, const S& dt) { return motor_exp(t.scale(dt * S(1.0 / 2.0))); }

// velocity_at: ½·[T, P] read as a Direction (screw.rs:276-281)
template <class S>
V3<S> velocity_at(const Mv<S>& t, const Mv<S>& p) {
    Mv<S> tw = narrow2(t);
    Mv<S> comm = gp(tw, p) - gp(p, tw);
    return direction_coords(comm.scale(S(1.0 / 2.0)));
}
// pairing ⟨w ∧ dual(t)⟩_I (screw.rs:350-352)
template <class S>
S pairing(const Mv<S>& w, const Mv<S>& t) { return wedge(narrow2(w), dual(narrow2(t))).c[TOP]; }
// from_line: narrow(dual(line)) (screw.rs:313-315)
template <class S>
Mv<S> wrench_from_line(const Mv<S>& line) { return narrow2(dual(line)); }
// coadjoint bracket ½·undual([t, dual(w)]) (screw.rs:317-328)
template <class S>
Mv<S> coad_bracket(const Mv<S>& w, const Mv<S>& t) {
    Mv<S> tw = narrow2(t), dw = dual(narrow2(w));
    Mv<S> raw = undual(gp(tw, dw) - gp(dw, tw));
    return narrow2(raw).scale(S(1.0 / 2.0));
}
// Lie bracket ½·(ab − ba) (screw.rs:262-274)
template <class S>
Mv<S> twist_bracket(const Mv<S>& a, const Mv<S>& b) {
    Mv<S> wa = narrow2(a), wb = narrow2(b);
    return narrow2(gp(wa, wb) - gp(wb, wa)).scale(S(1.0 / 2.0));
}
// Twist::basis(dof): one unit coordinate in the order Tx,Ty,Tz,Rx,Ry,Rz (screw.rs:186-202)
template <class S>
Mv<S> twist_basis(int dof) {
    V3<S> lin, ang;
    if (dof < 3) lin[dof] = S(1.0); else ang[dof - 3] = S(1.0);
    return screw_new(lin, ang);
}

}  // namespace oracle

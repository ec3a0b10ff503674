// A symbolic scalar for tracing the scalar-generic physics templates
// (exgame_amd/csrc/physics_dev.h, pga3_gen.h) into an expression DAG — the role
// `viete::Sym` plays for the reference's build-time codegen (crates/viete/src/api.rs:9-290,
// crates/viete/src/ir.rs:1-170).
//
// Every operation on a Sym appends a node to the current Tape, or returns an
// existing one: nodes are hash-consed (the same operation on the same operands is
// one node), constant operands are folded in IEEE double, and only identities that
// hold exactly in IEEE arithmetic are applied (x·1, 1·x, x/1, x − (+0), x + (−0),
// −(−x)); x·0 and x + 0 are kept, as the kernels compute them. Inputs are the
// kernel's value inputs, params its runtime parameter vector.
//
// Data-dependent branches (the EffectiveZero seams of the Study-form functions,
// crates/peano/src/scalar.rs:67-118) are forks: `effective_zero(value_of(x))` on a
// Sym records the condition |x| < 1e-9 and answers from the current decision plan
// (a condition already decided on the path keeps its answer).
// trace_paths() re-runs a trace once per reachable decision sequence (depth first,
// `true` before `false`), which enumerates every fork as viete's tracer does
// (crates/viete/src/engine.rs:917-1019); the emitter turns the paths into a
// decision trie.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace viete {

enum class Op : uint8_t { Const, Input, Param, Add, Sub, Mul, Div, Neg, Sqrt, Sin, Cos, Abs, Lt, Fma };

struct Node {
    Op op;
    int a = -1, b = -1, c = -1;
    double k = 0.0;  // Const: the value
    int idx = -1;    // Input / Param: the index
};

inline uint64_t bits_of(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
}

struct NodeKey {
    uint8_t op;
    int a, b, c, idx;
    uint64_t k;
    bool operator==(const NodeKey& o) const {
        return op == o.op && a == o.a && b == o.b && c == o.c && idx == o.idx && k == o.k;
    }
};
struct NodeKeyHash {
    size_t operator()(const NodeKey& x) const {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) {
            h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        };
        mix(x.op);
        mix((uint64_t)(uint32_t)x.a);
        mix((uint64_t)(uint32_t)x.b);
        mix((uint64_t)(uint32_t)x.c);
        mix((uint64_t)(uint32_t)x.idx);
        mix(x.k);
        return (size_t)h;
    }
};

// The host function a traced sin/cos stands for; the emitter calls the same-named
// device function, so both evaluate one definition.
using UnaryFn = double (*)(double);

struct Tape {
    std::vector<Node> nodes;
    std::unordered_map<NodeKey, int, NodeKeyHash> memo;
    UnaryFn sin_fn = nullptr, cos_fn = nullptr;

    int intern(const Node& n) {
        NodeKey key{(uint8_t)n.op, n.a, n.b, n.c, n.idx, n.op == Op::Const ? bits_of(n.k) : 0};
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        nodes.push_back(n);
        const int id = (int)nodes.size() - 1;
        memo.emplace(key, id);
        return id;
    }
    bool is_const(int id) const { return nodes[id].op == Op::Const; }
    bool is_const(int id, double v) const {
        return nodes[id].op == Op::Const && bits_of(nodes[id].k) == bits_of(v);
    }
    double value(int id) const { return nodes[id].k; }

    int constant(double v) {
        Node n{Op::Const};
        n.k = v;
        return intern(n);
    }
    int input(int i) {
        Node n{Op::Input};
        n.idx = i;
        return intern(n);
    }
    int param(int i) {
        Node n{Op::Param};
        n.idx = i;
        return intern(n);
    }
    int unary(Op op, int a) {
        if (is_const(a)) {
            const double x = value(a);
            switch (op) {
                case Op::Neg: return constant(-x);
                case Op::Sqrt: return constant(std::sqrt(x));
                case Op::Abs: return constant(std::fabs(x));
                case Op::Sin:
                    if (sin_fn) return constant(sin_fn(x));
                    break;
                case Op::Cos:
                    if (cos_fn) return constant(cos_fn(x));
                    break;
                default: break;
            }
        }
        if (op == Op::Neg && nodes[a].op == Op::Neg) return nodes[a].a;
        Node n{op};
        n.a = a;
        return intern(n);
    }
    int binary(Op op, int a, int b) {
        if (is_const(a) && is_const(b)) {
            const double x = value(a), y = value(b);
            switch (op) {
                case Op::Add: return constant(x + y);
                case Op::Sub: return constant(x - y);
                case Op::Mul: return constant(x * y);
                case Op::Div: return constant(x / y);
                case Op::Lt: return constant(x < y ? 1.0 : 0.0);
                default: break;
            }
        }
        switch (op) {
            case Op::Mul:
                if (is_const(b, 1.0)) return a;
                if (is_const(a, 1.0)) return b;
                break;
            case Op::Div:
                if (is_const(b, 1.0)) return a;
                break;
            case Op::Sub:
                if (is_const(b, 0.0)) return a;
                break;
            case Op::Add:
                if (is_const(b, -0.0)) return a;
                if (is_const(a, -0.0)) return b;
                break;
            default: break;
        }
        Node n{op};
        n.a = a;
        n.b = b;
        return intern(n);
    }
    int fma(int a, int b, int c) {
        Node n{Op::Fma};
        n.a = a;
        n.b = b;
        n.c = c;
        return intern(n);
    }
};

// ─────────── the current trace ───────────
struct TraceState {
    Tape* tape = nullptr;
    std::vector<char> plan;                  // decisions to take at the forks, in order
    size_t pos = 0;                          // forks met so far in this run
    std::vector<std::pair<int, char>> taken; // (condition node, decision) of this run
};
inline TraceState& state() {
    static TraceState s;
    return s;
}
inline Tape& tape() {
    if (!state().tape) throw std::runtime_error("viete: no active tape");
    return *state().tape;
}

struct FromId {};
struct Sym {
    int id;
    Sym() : id(tape().constant(0.0)) {}
    Sym(double v) : id(tape().constant(v)) {}
    Sym(int v) : id(tape().constant((double)v)) {}
    Sym(FromId, int i) : id(i) {}
};
inline Sym sym_of(int id) { return Sym(FromId{}, id); }
inline Sym input(int i) { return sym_of(tape().input(i)); }
inline Sym param(int i) { return sym_of(tape().param(i)); }

inline Sym operator+(const Sym& a, const Sym& b) { return sym_of(tape().binary(Op::Add, a.id, b.id)); }
inline Sym operator-(const Sym& a, const Sym& b) { return sym_of(tape().binary(Op::Sub, a.id, b.id)); }
inline Sym operator*(const Sym& a, const Sym& b) { return sym_of(tape().binary(Op::Mul, a.id, b.id)); }
inline Sym operator/(const Sym& a, const Sym& b) { return sym_of(tape().binary(Op::Div, a.id, b.id)); }
inline Sym operator-(const Sym& a) { return sym_of(tape().unary(Op::Neg, a.id)); }
inline Sym s_sqrt(const Sym& a) { return sym_of(tape().unary(Op::Sqrt, a.id)); }
inline Sym s_sin(const Sym& a) { return sym_of(tape().unary(Op::Sin, a.id)); }
inline Sym s_cos(const Sym& a) { return sym_of(tape().unary(Op::Cos, a.id)); }

// value_of(x) of a symbolic x: only its EffectiveZero test is defined
struct SymValue {
    int id;
};
inline SymValue value_of(const Sym& x) { return SymValue{x.id}; }
// EffectiveZero on f64: |x| < 1e-9 (crates/peano/src/support.rs:38-47), a fork
inline bool effective_zero(const SymValue& v) {
    Tape& t = tape();
    const int cond = t.binary(Op::Lt, t.unary(Op::Abs, v.id), t.constant(1e-9));
    if (t.is_const(cond)) return t.value(cond) != 0.0;
    TraceState& s = state();
    for (const auto& tk : s.taken)  // the same condition again on this path: same answer
        if (tk.first == cond) return tk.second != 0;
    char d;
    if (s.pos < s.plan.size()) {
        d = s.plan[s.pos];
    } else {
        d = 1;
        s.plan.push_back(d);
    }
    s.pos++;
    s.taken.emplace_back(cond, d);
    return d != 0;
}

// One decision sequence of a trace and the output nodes it reached.
struct Path {
    std::vector<std::pair<int, char>> decisions;
    std::vector<int> outputs;
};

// Runs `body` (which fills `outs` with output nodes) once per reachable decision
// sequence of its forks, depth first with `true` first.
inline std::vector<Path> trace_paths(Tape& t, const std::function<void(std::vector<int>&)>& body) {
    std::vector<Path> paths;
    TraceState& s = state();
    s.tape = &t;
    std::vector<char> plan;
    for (;;) {
        s.plan = plan;
        s.pos = 0;
        s.taken.clear();
        Path p;
        body(p.outputs);
        p.decisions = s.taken;
        paths.push_back(p);
        // next plan: flip the last `true` decision taken, drop what followed it
        std::vector<char> next(s.plan.begin(), s.plan.begin() + (long)s.pos);
        while (!next.empty() && next.back() == 0) next.pop_back();
        if (next.empty()) break;
        next.back() = 0;
        plan = next;
    }
    s.tape = nullptr;
    return paths;
}

}  // namespace viete

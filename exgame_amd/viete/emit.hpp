// Passes over a traced Tape and the HIP emitter (the role of the reference's GLSL
// emitter, crates/viete/src/glsl.rs:12-80, and its pass list, crates/viete/src/engine.rs):
//
//   dead-code elimination  only nodes an output or a fork condition reaches are emitted;
//   FMA contraction        a product used once, added to or subtracted from another
//                          value, becomes one fused multiply-add (fma(a, b, c),
//                          fma(a, b, −c), fma(−a, b, c));
//   decision trie          the traced paths, merged on their common decision prefixes,
//                          become nested if/else on the fork conditions;
//   hoisting               a node both sides of a fork need is emitted before the fork.
//
// The emitted body turns FP contraction off, so the fused operations are exactly the
// ones the FMA pass chose, whatever the compiler would fuse on its own.
#pragma once
#include <algorithm>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>
#include "sym.hpp"

namespace viete {

struct Program {
    const Tape* tape;
    std::vector<Path> paths;
    int n_out = 0;
    // filled by analyse()
    std::vector<int> uses;        // uses by live nodes, outputs and conditions
    std::vector<char> fused_mul;  // a Mul consumed by an FMA
    std::vector<int> fma_of;      // Add/Sub node → 0 (none), 1 (a is the product), 2 (b is the product)
};

inline void operands(const Node& n, int out[3]) {
    out[0] = n.a;
    out[1] = n.b;
    out[2] = n.c;
}

inline void analyse(Program& p, bool contract) {
    const auto& nodes = p.tape->nodes;
    const int N = (int)nodes.size();
    std::vector<char> live(N, 0);
    std::vector<int> stack;
    for (const auto& path : p.paths) {
        for (int o : path.outputs) stack.push_back(o);
        for (auto& d : path.decisions) stack.push_back(d.first);
    }
    while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        if (x < 0 || live[x]) continue;
        live[x] = 1;
        int ops[3];
        operands(nodes[x], ops);
        for (int o : ops)
            if (o >= 0) stack.push_back(o);
    }
    p.uses.assign(N, 0);
    for (int x = 0; x < N; ++x) {
        if (!live[x]) continue;
        int ops[3];
        operands(nodes[x], ops);
        for (int o : ops)
            if (o >= 0) p.uses[o]++;
    }
    std::set<int> roots;
    for (const auto& path : p.paths) {
        for (int o : path.outputs) roots.insert(o);
        for (auto& d : path.decisions) roots.insert(d.first);
    }
    for (int r : roots) p.uses[r]++;
    p.fused_mul.assign(N, 0);
    p.fma_of.assign(N, 0);
    if (!contract) return;
    for (int x = 0; x < N; ++x) {
        if (!live[x]) continue;
        const Node& n = nodes[x];
        if (n.op != Op::Add && n.op != Op::Sub) continue;
        auto single_mul = [&](int y) { return nodes[y].op == Op::Mul && p.uses[y] == 1 && !p.fused_mul[y]; };
        if (single_mul(n.a)) {
            p.fma_of[x] = 1;
            p.fused_mul[n.a] = 1;
        } else if (single_mul(n.b)) {
            p.fma_of[x] = 2;
            p.fused_mul[n.b] = 1;
        }
    }
}

inline std::string literal(double v) {
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    std::string s(buf);
    return v < 0 ? "(" + s + ")" : s;
}

struct Emitter {
    const Program& p;
    std::ostringstream os;
    explicit Emitter(const Program& prog) : p(prog) {}

    std::string ref(int x) const {
        const Node& n = p.tape->nodes[x];
        switch (n.op) {
            case Op::Const: return literal(n.k);
            case Op::Input: return "in[" + std::to_string(n.idx) + "]";
            case Op::Param: return "prm[" + std::to_string(n.idx) + "]";
            default: return "t" + std::to_string(x);
        }
    }
    std::string expr(int x) const {
        const Node& n = p.tape->nodes[x];
        switch (n.op) {
            case Op::Add:
            case Op::Sub: {
                const bool sub = n.op == Op::Sub;
                if (p.fma_of[x] == 1) {
                    const Node& m = p.tape->nodes[n.a];
                    return "fma(" + ref(m.a) + ", " + ref(m.b) + ", " + (sub ? "-" : "") + ref(n.b) + ")";
                }
                if (p.fma_of[x] == 2) {
                    const Node& m = p.tape->nodes[n.b];
                    return "fma(" + std::string(sub ? "-" : "") + ref(m.a) + ", " + ref(m.b) + ", " + ref(n.a) + ")";
                }
                return ref(n.a) + (sub ? " - " : " + ") + ref(n.b);
            }
            case Op::Mul: return ref(n.a) + " * " + ref(n.b);
            case Op::Div: return ref(n.a) + " / " + ref(n.b);
            case Op::Neg: return "-" + ref(n.a);
            case Op::Sqrt: return "sqrt(" + ref(n.a) + ")";
            case Op::Sin: return "exg::s_sin(" + ref(n.a) + ")";
            case Op::Cos: return "exg::s_cos(" + ref(n.a) + ")";
            case Op::Abs: return "fabs(" + ref(n.a) + ")";
            case Op::Lt: return ref(n.a) + " < " + ref(n.b);
            case Op::Fma: return "fma(" + ref(n.a) + ", " + ref(n.b) + ", " + ref(n.c) + ")";
            default: return ref(x);
        }
    }
    // the nodes `x` needs that are emitted as statements (not inputs, constants or
    // fused products)
    void deps(int x, std::set<int>& out) const {
        std::vector<int> st{x};
        while (!st.empty()) {
            const int y = st.back();
            st.pop_back();
            const Node& n = p.tape->nodes[y];
            if (n.op == Op::Const || n.op == Op::Input || n.op == Op::Param) continue;
            if (!p.fused_mul[y]) {
                if (out.count(y)) continue;
                out.insert(y);
            }
            int ops[3];
            operands(n, ops);
            for (int o : ops)
                if (o >= 0) st.push_back(o);
        }
    }

    struct Trie {
        int cond = -1;
        int child[2] = {-1, -1};  // [false, true]
        int leaf = -1;            // path index
    };
    std::vector<Trie> trie;

    int build(const std::vector<int>& members, size_t depth) {
        Trie t;
        const Path& first = p.paths[members[0]];
        if (members.size() == 1 && depth >= first.decisions.size()) {
            t.leaf = members[0];
            trie.push_back(t);
            return (int)trie.size() - 1;
        }
        t.cond = first.decisions[depth].first;
        std::vector<int> side[2];
        for (int m : members) side[(int)p.paths[m].decisions[depth].second].push_back(m);
        trie.push_back(t);
        const int me = (int)trie.size() - 1;
        for (int s = 0; s < 2; ++s) {
            const int c = side[s].empty() ? -1 : build(side[s], depth + 1);
            trie[me].child[s] = c;
        }
        return me;
    }
    void needed(int node, std::set<int>& out) const {
        const Trie& t = trie[node];
        if (t.leaf >= 0) {
            for (int o : p.paths[t.leaf].outputs) deps(o, out);
            return;
        }
        deps(t.cond, out);
        for (int s = 0; s < 2; ++s)
            if (t.child[s] >= 0) needed(t.child[s], out);
    }
    void statements(const std::set<int>& todo, std::set<int>& done, int indent) {
        for (int x : todo) {  // ascending ids: operands come first
            if (done.count(x)) continue;
            const Node& n = p.tape->nodes[x];
            os << std::string(indent, ' ') << (n.op == Op::Lt ? "const bool t" : "const double t") << x << " = "
               << expr(x) << ";\n";
            done.insert(x);
        }
    }
    void scope(int node, std::set<int> done, int indent) {
        const Trie& t = trie[node];
        if (t.leaf >= 0) {
            std::set<int> need;
            for (int o : p.paths[t.leaf].outputs) deps(o, need);
            statements(need, done, indent);
            const auto& outs = p.paths[t.leaf].outputs;
            for (size_t k = 0; k < outs.size(); ++k)
                os << std::string(indent, ' ') << "out(" << k << ", " << ref(outs[k]) << ");\n";
            return;
        }
        // hoist what both sides need, and the condition
        std::set<int> sides[2], hoist;
        for (int s = 0; s < 2; ++s)
            if (t.child[s] >= 0) needed(t.child[s], sides[s]);
        deps(t.cond, hoist);
        if (t.child[0] >= 0 && t.child[1] >= 0)
            std::set_intersection(sides[0].begin(), sides[0].end(), sides[1].begin(), sides[1].end(),
                                  std::inserter(hoist, hoist.begin()));
        statements(hoist, done, indent);
        const std::string pad(indent, ' ');
        if (t.child[1] >= 0 && t.child[0] >= 0) {
            os << pad << "if (" << ref(t.cond) << ") {\n";
            scope(t.child[1], done, indent + 4);
            os << pad << "} else {\n";
            scope(t.child[0], done, indent + 4);
            os << pad << "}\n";
        } else {
            scope(t.child[1] >= 0 ? t.child[1] : t.child[0], done, indent);
        }
    }

    std::string function(const std::string& name, const std::string& doc, bool has_params) {
        std::vector<int> all(p.paths.size());
        for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
        trie.clear();
        const int root = build(all, 0);
        os.str("");
        os << doc;
        os << "template <class Out>\n__device__ __forceinline__ void " << name << "(const double* __restrict__ in, "
           << (has_params ? "const double* __restrict__ prm, " : "") << "Out&& out) {\n";
        os << "#pragma clang fp contract(off)\n";
        scope(root, {}, 4);
        os << "}\n";
        return os.str();
    }
};

}  // namespace viete

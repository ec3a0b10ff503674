// Build-time code generator: traces the scalar-generic joint laws and the Pre
// stage of exgame_amd/csrc/physics_dev.h over a symbolic scalar and emits them as
// straight-line HIP __device__ functions in the reference's kernel I/O layout
// (crates/newton/build.rs:13-48, 89-104):
//
//   input  in[30]  0..12 velocity values (end a 0..6, end b 6..12) → gradient axes 12..24
//                  12..20 base pose of end a, 20..28 of end b (8 motor components)
//                  28 dt, 29 warp
//          prm[]   the joint's canonical parameter vector
//   output _full   out[300]: component c (a.f, a.τ, b.f, b.τ; 12) at out[c·25 .. c·25+25] =
//                  [value, ∂/∂axis 0 … ∂/∂axis 23]; pose axes 0..12 perturb base ∘ exp(δ)
//          _plain  out[12]: the twelve values
//   Pre    in[15]  twist 6, pose 8, retraction r → out[14]: pose ∘ exp(r·V) 8, V 6
//
// Usage: gen <header>         writes the generated header
//        gen --check <n>      evaluates every traced tape on n seeded inputs and
//                             compares it, bit for bit, with the same templates
//                             evaluated directly in double (exit status 1 on a mismatch)
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../csrc/physics_dev.h"
#include "emit.hpp"
#include "sym.hpp"

using viete::Sym;
using SD = exg::Dual<Sym>;

namespace {

const char* const LAW_NAME[4] = {"simple_spring_damper", "critically_damped_warped", "perpendicular_damper_warped",
                                 "torsional_damper_warped"};

// ───────────── the traced bodies (templates of the kernels' own code) ─────────────
template <class S, class P>
void joint_full_body(int kind, const S in[30], const P prm[10], S out[300]) {
    using D = exg::Dual<S>;
    for (int ax = 0; ax < 24; ++ax) {
        D pp[2][8], vv[2][6];
        for (int e = 0; e < 2; ++e) {
            S base[8];
            for (int k = 0; k < 8; ++k) base[k] = in[12 + 8 * e + k];
            if (ax < 12 && ax / 6 == e) exg::seed_pose(base, ax % 6, pp[e]);
            else
                for (int k = 0; k < 8; ++k) pp[e][k] = D(base[k]);
            for (int k = 0; k < 6; ++k) vv[e][k] = D(in[6 * e + k], (ax >= 12 && ax - 12 == 6 * e + k) ? S(1.0) : S(0.0));
        }
        D wa[6], wb[6];
        const P dt = P(in[28]), warp = P(in[29]);
        exg::joint_wrench(kind, prm, pp[0], pp[1], vv[0], vv[1], dt, warp, wa, wb);
        for (int c = 0; c < 6; ++c) {
            out[c * 25 + 1 + ax] = wa[c].d;
            out[(6 + c) * 25 + 1 + ax] = wb[c].d;
            if (ax == 0) {
                out[c * 25] = wa[c].v;
                out[(6 + c) * 25] = wb[c].v;
            }
        }
    }
}
template <class S, class P>
void joint_plain_body(int kind, const S in[30], const P prm[10], S out[12]) {
    S pa[8], pb[8], va[6], vb[6];
    for (int k = 0; k < 8; ++k) {
        pa[k] = in[12 + k];
        pb[k] = in[20 + k];
    }
    for (int k = 0; k < 6; ++k) {
        va[k] = in[k];
        vb[k] = in[6 + k];
    }
    const P dt = P(in[28]), warp = P(in[29]);
    exg::joint_wrench(kind, prm, pa, pb, va, vb, dt, warp, out, out + 6);
}
template <class S>
void pre_body(const S in[15], S out[14]) {
    S v[6], pose[8], e[8];
    for (int k = 0; k < 6; ++k) v[k] = in[k];
    for (int k = 0; k < 8; ++k) pose[k] = in[6 + k];
    exg::twist_exp(v, in[14], e);
    exg::compose(pose, e, out);
    for (int k = 0; k < 6; ++k) out[8 + k] = v[k];
}

struct Traced {
    std::string name, doc;
    bool has_params;
    int n_in, n_out;
    viete::Tape tape;
    std::vector<viete::Path> paths;
};

void trace_joint(Traced& t, int kind, bool full) {
    t.tape.sin_fn = &exg::s_sin;
    t.tape.cos_fn = &exg::s_cos;
    t.has_params = true;
    t.n_in = 30;
    t.n_out = full ? 300 : 12;
    t.paths = viete::trace_paths(t.tape, [&](std::vector<int>& outs) {
        Sym in[30], prm[10];
        for (int i = 0; i < 30; ++i) in[i] = viete::input(i);
        for (int i = 0; i < 10; ++i) prm[i] = viete::param(i);
        std::vector<Sym> o(t.n_out);
        if (full) joint_full_body<Sym, Sym>(kind, in, prm, o.data());
        else joint_plain_body<Sym, Sym>(kind, in, prm, o.data());
        outs.clear();
        for (auto& s : o) outs.push_back(s.id);
    });
}
void trace_pre(Traced& t) {
    t.tape.sin_fn = &exg::s_sin;
    t.tape.cos_fn = &exg::s_cos;
    t.has_params = false;
    t.n_in = 15;
    t.n_out = 14;
    t.paths = viete::trace_paths(t.tape, [&](std::vector<int>& outs) {
        Sym in[15];
        for (int i = 0; i < 15; ++i) in[i] = viete::input(i);
        Sym o[14];
        pre_body<Sym>(in, o);
        outs.clear();
        for (auto& s : o) outs.push_back(s.id);
    });
}

std::vector<Traced> trace_all() {
    std::vector<Traced> all(9);
    for (int kind = 0; kind < 4; ++kind) {
        for (int full = 0; full < 2; ++full) {
            Traced& t = all[kind * 2 + full];
            t.name = std::string("viete_") + LAW_NAME[kind] + (full ? "_full" : "_plain");
            trace_joint(t, kind, full != 0);
            char buf[256];
            std::snprintf(buf, sizeof buf, "// %s, %s: %zu nodes traced, %zu path(s)\n", LAW_NAME[kind],
                          full ? "value + 24 partials per component (out[300])" : "values (out[12])",
                          t.tape.nodes.size(), t.paths.size());
            t.doc = buf;
        }
    }
    Traced& p = all[8];
    p.name = "viete_pre";
    trace_pre(p);
    char buf[256];
    std::snprintf(buf, sizeof buf, "// Pre: midpoint pose ∘ exp(r·V) and V (out[14]): %zu nodes traced, %zu path(s)\n",
                  p.tape.nodes.size(), p.paths.size());
    p.doc = buf;
    return all;
}

// ───────────── the interpreter of a traced program ─────────────
bool eval(const Traced& t, const double* in, const double* prm, std::vector<double>& out) {
    const auto& nodes = t.tape.nodes;
    std::vector<double> v(nodes.size());
    for (size_t x = 0; x < nodes.size(); ++x) {
        const viete::Node& n = nodes[x];
        auto A = [&] { return v[n.a]; };
        auto B = [&] { return v[n.b]; };
        switch (n.op) {
            case viete::Op::Const: v[x] = n.k; break;
            case viete::Op::Input: v[x] = in[n.idx]; break;
            case viete::Op::Param: v[x] = prm[n.idx]; break;
            case viete::Op::Add: v[x] = A() + B(); break;
            case viete::Op::Sub: v[x] = A() - B(); break;
            case viete::Op::Mul: v[x] = A() * B(); break;
            case viete::Op::Div: v[x] = A() / B(); break;
            case viete::Op::Neg: v[x] = -A(); break;
            case viete::Op::Sqrt: v[x] = std::sqrt(A()); break;
            case viete::Op::Sin: v[x] = exg::s_sin(A()); break;
            case viete::Op::Cos: v[x] = exg::s_cos(A()); break;
            case viete::Op::Abs: v[x] = std::fabs(A()); break;
            case viete::Op::Lt: v[x] = A() < B() ? 1.0 : 0.0; break;
            case viete::Op::Fma: v[x] = std::fma(A(), B(), v[n.c]); break;
        }
    }
    for (const auto& path : t.paths) {
        bool match = true;
        for (auto& d : path.decisions) match = match && ((v[d.first] != 0.0) == (d.second != 0));
        if (!match) continue;
        out.resize(path.outputs.size());
        for (size_t k = 0; k < out.size(); ++k) out[k] = v[path.outputs[k]];
        return true;
    }
    return false;
}

void random_motor(std::mt19937_64& rng, double m[8], double rot_scale = 1.0) {
    std::normal_distribution<double> g(0.0, 1.0);
    double t6[6];
    for (int k = 0; k < 3; ++k) t6[k] = 3.0 * g(rng);
    for (int k = 3; k < 6; ++k) t6[k] = rot_scale * g(rng);
    exg::twist_exp(t6, 1.0, m);
}

int check(int n) {
    std::vector<Traced> all = trace_all();
    std::mt19937_64 rng(12345);
    std::normal_distribution<double> g(0.0, 1.0);
    std::uniform_real_distribution<double> u(0.5, 2.0);
    int bad = 0;
    long long compared = 0;
    for (int trial = 0; trial < n; ++trial) {
        double in[30], prm[10];
        for (int k = 0; k < 12; ++k) in[k] = g(rng);
        random_motor(rng, in + 12);
        random_motor(rng, in + 20);
        in[28] = 1e-2 * u(rng);
        in[29] = trial % 3 == 0 ? 1.0 : u(rng);
        for (int k = 0; k < 6; ++k) prm[k] = 0.3 * g(rng);
        prm[6] = u(rng);
        prm[7] = 1e3 * u(rng);
        prm[8] = 10.0 * u(rng);
        prm[9] = 1e-6;
        for (int kind = 0; kind < 4; ++kind) {
            double p[10];
            for (int k = 0; k < 10; ++k) p[k] = prm[k];
            if (kind == exg::PERPENDICULAR_DAMPER_WARPED) {
                p[0] = 5.0 * u(rng);
                p[1] = 1e-6;
            }
            if (kind == exg::TORSIONAL_DAMPER_WARPED) p[0] = 5.0 * u(rng);
            for (int full = 0; full < 2; ++full) {
                const Traced& t = all[kind * 2 + full];
                std::vector<double> got;
                if (!eval(t, in, p, got)) {
                    std::printf("%s: no path matched\n", t.name.c_str());
                    ++bad;
                    continue;
                }
                std::vector<double> want(t.n_out);
                if (full) joint_full_body<double, double>(kind, in, p, want.data());
                else joint_plain_body<double, double>(kind, in, p, want.data());
                for (int k = 0; k < t.n_out; ++k) {
                    ++compared;
                    if (viete::bits_of(got[k]) != viete::bits_of(want[k]) && !(got[k] == 0.0 && want[k] == 0.0)) {
                        if (bad < 10)
                            std::printf("%s out[%d]: traced %.17g direct %.17g\n", t.name.c_str(), k, got[k], want[k]);
                        ++bad;
                    }
                }
            }
        }
        // Pre at and around its seams: no rotation, tiny rotations, ordinary ones
        double pin[15];
        const double scales[4] = {0.0, 1e-6, 1e-3, 1.0};
        for (int k = 0; k < 3; ++k) pin[k] = g(rng);
        for (int k = 3; k < 6; ++k) pin[k] = scales[trial % 4] * g(rng);
        random_motor(rng, pin + 6);
        pin[14] = trial % 5 == 0 ? 0.5 : 1e-2 * u(rng);
        const Traced& t = all[8];
        std::vector<double> got;
        if (!eval(t, pin, nullptr, got)) {
            std::printf("viete_pre: no path matched\n");
            ++bad;
            continue;
        }
        double want[14];
        pre_body<double>(pin, want);
        for (int k = 0; k < 14; ++k) {
            ++compared;
            if (viete::bits_of(got[k]) != viete::bits_of(want[k]) && !(got[k] == 0.0 && want[k] == 0.0)) {
                if (bad < 10) std::printf("viete_pre out[%d]: traced %.17g direct %.17g\n", k, got[k], want[k]);
                ++bad;
            }
        }
    }
    std::printf("compared %lld outputs over %d inputs: %d mismatches\n", compared, n, bad);
    for (const auto& t : all) std::printf("%s: %zu nodes, %zu paths\n", t.name.c_str(), t.tape.nodes.size(), t.paths.size());
    return bad == 0 ? 0 : 1;
}

int write_header(const char* path) {
    std::vector<Traced> all = trace_all();
    std::ostringstream os;
    os << "// GENERATED by exgame_amd/viete/gen.cpp from the templates of physics_dev.h — do not edit.\n"
          "// The joint laws and the Pre stage traced over a symbolic scalar and emitted as\n"
          "// straight-line device code in the kernel I/O layout of crates/newton/build.rs:13-48\n"
          "// (in[30]: velocities of ends a, b; base poses of a, b; dt; warp. prm: the joint's\n"
          "// parameters. _full: output c·25 + 0 the value, c·25 + 1 + axis ∂/∂axis for the 12\n"
          "// wrench components c; _plain: outputs 0..12. Pre: in[15] twist, pose, r → outputs\n"
          "// 0..14). Output k with value v is delivered as out(k, v), k a constant.\n"
          "#pragma once\n#include <hip/hip_runtime.h>\n#include \"physics_dev.h\"\n\nnamespace exg {\n\n";
    for (auto& t : all) {
        viete::Program prog{&t.tape, t.paths, t.n_out};
        viete::analyse(prog, true);
        viete::Emitter em(prog);
        os << em.function(t.name, t.doc, t.has_params) << "\n";
    }
    os << "}  // namespace exg\n";
    std::ofstream f(path);
    f << os.str();
    return f.good() ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "--check") return check(argc >= 3 ? std::atoi(argv[2]) : 200);
    if (argc != 2) {
        std::fprintf(stderr, "usage: gen <header> | gen --check [n]\n");
        return 2;
    }
    return write_header(argv[1]);
}

// spec_emit.hpp — the text of a compiled-program specialisation: a struct of
// constexpr members (the lowered micro-op program, leaf types and access
// modes) that the kernel templates are instantiated on.  Written by specgen
// into the ahead-of-time table and by the run-time compiler for programs
// outside that table.  Host only.
#pragma once

#include <string>

#include "planner.hpp"
#include "reduce.hpp"

namespace xg {

inline std::string emit_program_text(const Program& prog, const LeafView* leaves) {
    std::string s;
    auto add = [&](const char* fmt, auto... a) {
        char buf[256];
        std::snprintf(buf, sizeof buf, fmt, a...);
        s += buf;
    };
    add("    static constexpr int nops = %d;\n", prog.nops);
    add("    static constexpr int nregs = %d;\n", prog.nregs);
    add("    static constexpr int nleaves = %d;\n", prog.nleaves);
    add("    static constexpr int result_reg = %d;\n", prog.result_reg);
    add("    static constexpr xg::MicroOp ops[%d] = {", prog.nops > 0 ? prog.nops : 1);
    for (int i = 0; i < prog.nops; ++i) {
        const MicroOp& m = prog.ops[i];
        add("%s{%d, %d, %d, %d, %d, %d, %d}", i ? ", " : "", m.op, m.t, m.dst, m.a, m.b, m.c, m.aux);
    }
    s += "};\n";
    const int nl = prog.nleaves > 0 ? prog.nleaves : 1;
    add("    static constexpr uint8_t leaf_dtype[%d] = {", nl);
    for (int l = 0; l < nl; ++l) add("%s%d", l ? ", " : "", l < prog.nleaves ? leaves[l].dtype : 0);
    add("};\n    static constexpr uint8_t leaf_mode[%d] = {", nl);
    for (int l = 0; l < nl; ++l) add("%s%d", l ? ", " : "", l < prog.nleaves ? lm_base(leaves[l].mode) : 0);
    s += "};\n";
    return s;
}

// specialisation of an elementwise plan at vector width V
inline std::string emit_ew_spec(const std::string& name, const EwPlan& p, int V) {
    return "struct " + name + " {\n    static constexpr bool is_static = true;\n    static constexpr int V = " +
           std::to_string(V) + ";\n    static constexpr int out_dtype = " + std::to_string(p.out_dtype) +
           ";\n    static constexpr int out_mode = " + std::to_string(lm_base(p.out_mode)) + ";\n" +
           emit_program_text(p.prog, p.leaves) + "};\n";
}

// specialisation of a reduction plan at vector width V
inline std::string emit_red_spec(const std::string& name, const RedPlan& p, int V) {
    return "struct " + name + " {\n    static constexpr bool is_static = true;\n    static constexpr int V = " +
           std::to_string(V) + ";\n    static constexpr int acc = " + std::to_string(p.acc) +
           ";\n    static constexpr int op = " + std::to_string(p.op) + ";\n" + emit_program_text(p.prog, p.leaves) +
           "};\n";
}

}  // namespace xg

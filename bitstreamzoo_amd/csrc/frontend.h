// frontend.h — the .bs DSL front-end: tokens, AST with interned names, parser
// and 3-address validator.
//
// Contract kept from the reference (grammar DSL.md:303-357):
//   tokenize          simulator/csim/src/lexer.cpp:62-142
//   parse_program     parser.cpp:36-79 (output int → IntLocalDecl 0,
//                     output stream x[N] → ArrayDecl, params / inputs lists)
//   expressions       parser.cpp:234-321 (precedence ~, << >>, &, +, ^, |;
//                     a shift amount is a full int_expr)
//   validate_3addr    parser.cpp:375-447
// Names are interned to dense ids so the control plane indexes vectors
// instead of hashing strings per operand (interpreter.h:85 hashes per access).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace bsz {

using Sym = int32_t;

struct SymTab {
    std::unordered_map<std::string, Sym> ids;
    std::vector<std::string> names;
    Sym intern(const std::string& s) {
        auto it = ids.find(s);
        if (it != ids.end()) return it->second;
        Sym id = (Sym)names.size();
        names.push_back(s);
        ids.emplace(s, id);
        return id;
    }
    Sym find(const std::string& s) const {
        auto it = ids.find(s);
        return it == ids.end() ? -1 : it->second;
    }
};

struct ParseError : std::runtime_error { using std::runtime_error::runtime_error; };
struct RunError : std::runtime_error { using std::runtime_error::runtime_error; };

// Expression nodes live in per-program arenas and refer to each other by index.
struct IExpr {
    enum Kind : uint8_t { Lit, Var, Add, Sub, Mul, Pop } k;
    int64_t lit = 0;
    Sym var = -1;
    int32_t l = -1, r = -1;   // Add/Sub/Mul operands (IExpr ids)
    int32_t arg = -1;         // Pop: SExpr id
};

struct SExpr {
    enum Kind : uint8_t { Var, Zero, Ones, Arr, Not, And, Or, Xor, Plus, Shl, Shr } k;
    Sym name = -1;            // Var, Arr
    int32_t a = -1, b = -1;   // operands (SExpr ids)
    int32_t idx = -1;         // Arr index / shift amount (IExpr id)
    bool primary() const { return k == Var || k == Zero || k == Ones || k == Arr; }
};

struct Stmt {
    enum Kind : uint8_t { Assign, ArrAssign, If, While, For, Local, ArrDecl, IntDecl, IntAssign } k;
    Sym name = -1;            // target / loop var / declared name / cond var
    int32_t e = -1;           // stream rhs (SExpr) or cond (SExpr)
    int32_t i1 = -1, i2 = -1; // index / size / int rhs / for bounds (IExpr)
    std::vector<Stmt> body;
};

struct Program {
    SymTab syms;
    std::vector<IExpr> iex;
    std::vector<SExpr> sex;
    std::vector<Stmt> stmts;
    std::vector<Sym> inputs, outputs, params, out_ints;
    std::vector<bool> is_array;       // by Sym: declared with [...]
    size_t source_bytes = 0;
};

// tokenize + parse + validate (parser.cpp:451-457).  Throws ParseError.
std::unique_ptr<Program> parse_program(const std::string& source);

// Static op count of the reference's OpCounter (ast_liveness.cpp:629-719):
// `for` bodies times trip count, `if`/`while` bodies once.
int64_t count_static_ops(const Program& p, const std::unordered_map<std::string, int64_t>& params);

}  // namespace bsz

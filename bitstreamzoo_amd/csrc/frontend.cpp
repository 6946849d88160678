// frontend.cpp — lexer, parser and 3-address validator for the .bs DSL.
// Reference behaviour: simulator/csim/src/lexer.cpp:62-142,
// parser.cpp:36-457; grammar DSL.md:303-357.
#include "frontend.h"

#include <cstring>

namespace bsz {
namespace {

enum class T : uint8_t {
    Input, Output, Param, Stream, Int, If, While, For, In, Zero, Ones, Popcount,
    Ident, Lit, LBrace, RBrace, LParen, RParen, LBracket, RBracket, Assign,
    Tilde, Amp, Pipe, Caret, LShift, RShift, Plus, Minus, Star, DotDot, Eof
};

const char* tname(T t) {
    static const char* n[] = {"INPUT", "OUTPUT", "PARAM", "STREAM", "INT", "IF", "WHILE", "FOR", "IN",
                              "ZERO", "ONES", "POPCOUNT", "IDENT", "INT_LIT", "LBRACE", "RBRACE",
                              "LPAREN", "RPAREN", "LBRACKET", "RBRACKET", "ASSIGN", "TILDE", "AMP",
                              "PIPE", "CARET", "LSHIFT", "RSHIFT", "PLUS", "MINUS", "STAR", "DOTDOT", "EOF"};
    return n[(int)t];
}

struct Tok {
    T t;
    uint32_t off, len;  // span in the source
    int line;
};

// Keyword lookup without a map: the 12 reserved words by length then bytes.
T keyword_or_ident(const char* s, uint32_t n) {
    struct KW { const char* w; T t; };
    static const KW kws[] = {{"input", T::Input}, {"output", T::Output}, {"param", T::Param},
                             {"stream", T::Stream}, {"int", T::Int}, {"if", T::If},
                             {"while", T::While}, {"for", T::For}, {"in", T::In},
                             {"ZERO", T::Zero}, {"ONES", T::Ones}, {"popcount", T::Popcount}};
    for (const KW& k : kws)
        if (std::strlen(k.w) == n && std::memcmp(k.w, s, n) == 0) return k.t;
    return T::Ident;
}

std::vector<Tok> lex(const std::string& src) {
    std::vector<Tok> out;
    out.reserve(src.size() / 3 + 8);
    const char* s = src.data();
    const uint32_t n = (uint32_t)src.size();
    uint32_t i = 0;
    int line = 1;
    auto is_alpha = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; };
    auto is_dig = [](char c) { return c >= '0' && c <= '9'; };
    while (i < n) {
        char c = s[i];
        if (c == '\n') { ++line; ++i; continue; }
        if (c == ' ' || c == '\t' || c == '\r') { ++i; continue; }
        if (c == '/' && i + 1 < n && s[i + 1] == '/') { while (i < n && s[i] != '\n') ++i; continue; }
        if (i + 1 < n) {
            char d = s[i + 1];
            T two = T::Eof;
            if (c == '<' && d == '<') two = T::LShift;
            else if (c == '>' && d == '>') two = T::RShift;
            else if (c == '.' && d == '.') two = T::DotDot;
            if (two != T::Eof) { out.push_back({two, i, 2, line}); i += 2; continue; }
        }
        T one = T::Eof;
        switch (c) {
            case '{': one = T::LBrace; break;   case '}': one = T::RBrace; break;
            case '(': one = T::LParen; break;   case ')': one = T::RParen; break;
            case '[': one = T::LBracket; break; case ']': one = T::RBracket; break;
            case '=': one = T::Assign; break;   case '~': one = T::Tilde; break;
            case '&': one = T::Amp; break;      case '|': one = T::Pipe; break;
            case '^': one = T::Caret; break;    case '+': one = T::Plus; break;
            case '-': one = T::Minus; break;    case '*': one = T::Star; break;
            default: break;
        }
        if (one != T::Eof) { out.push_back({one, i, 1, line}); ++i; continue; }
        if (is_dig(c)) {
            uint32_t b = i;
            while (i < n && is_dig(s[i])) ++i;
            out.push_back({T::Lit, b, i - b, line});
            continue;
        }
        if (is_alpha(c)) {
            uint32_t b = i;
            while (i < n && (is_alpha(s[i]) || is_dig(s[i]))) ++i;
            out.push_back({keyword_or_ident(s + b, i - b), b, i - b, line});
            continue;
        }
        throw ParseError("line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'");
    }
    out.push_back({T::Eof, n, 0, line});
    return out;
}

class Parser {
public:
    Parser(const std::string& src, Program& p) : src_(src), p_(p), toks_(lex(src)) {}

    void run() {
        while (cur().t != T::Eof) {
            switch (cur().t) {
                case T::Input: decl_input(); break;
                case T::Output: decl_output(); break;
                case T::Param: {
                    next(); want(T::Int);
                    Sym s = name(want(T::Ident));
                    int_names_.push_back(s);
                    mark_int(s);
                    p_.params.push_back(s);
                    break;
                }
                case T::Stream: case T::Int: case T::If: case T::While: case T::For: case T::Ident:
                    p_.stmts.push_back(stmt());
                    break;
                default:
                    throw ParseError("line " + std::to_string(cur().line) + ": unexpected token " +
                                     tname(cur().t) + " ('" + text(cur()) + "')");
            }
        }
    }

private:
    const std::string& src_;
    Program& p_;
    std::vector<Tok> toks_;
    size_t pos_ = 0;
    std::vector<bool> is_int_;       // by Sym: int-typed at this point of the parse
    std::vector<Sym> int_names_;

    const Tok& cur() const { return toks_[pos_]; }
    Tok next() { return toks_[pos_++]; }
    std::string text(const Tok& t) const { return src_.substr(t.off, t.len); }
    Tok want(T t) {
        Tok k = next();
        if (k.t != t)
            throw ParseError("line " + std::to_string(k.line) + ": expected " + tname(t) + ", got " +
                             tname(k.t) + " ('" + text(k) + "')");
        return k;
    }
    Sym name(const Tok& t) {
        Sym s = p_.syms.intern(text(t));
        if ((size_t)s >= p_.is_array.size()) p_.is_array.resize(s + 1, false);
        return s;
    }
    void mark_int(Sym s) {
        if ((size_t)s >= is_int_.size()) is_int_.resize(s + 1, false);
        is_int_[s] = true;
    }
    bool is_int(Sym s) const { return (size_t)s < is_int_.size() && is_int_[s]; }

    int32_t new_i(IExpr e) { p_.iex.push_back(e); return (int32_t)p_.iex.size() - 1; }
    int32_t new_s(SExpr e) { p_.sex.push_back(e); return (int32_t)p_.sex.size() - 1; }

    void decl_input() {
        next(); want(T::Stream);
        Sym s = name(want(T::Ident));
        if (cur().t == T::LBracket) { next(); iexpr(); want(T::RBracket); p_.is_array[s] = true; }
        p_.inputs.push_back(s);
    }

    void decl_output() {
        next();
        if (cur().t == T::Int) {
            next();
            Sym s = name(want(T::Ident));
            mark_int(s);
            p_.out_ints.push_back(s);
            Stmt st; st.k = Stmt::IntDecl; st.name = s;
            IExpr zero; zero.k = IExpr::Lit; zero.lit = 0;
            st.i1 = new_i(zero);
            p_.stmts.push_back(std::move(st));
            return;
        }
        want(T::Stream);
        Sym s = name(want(T::Ident));
        if (cur().t == T::LBracket) {
            next();
            Stmt st; st.k = Stmt::ArrDecl; st.name = s; st.i1 = iexpr();
            want(T::RBracket);
            p_.is_array[s] = true;
            p_.stmts.push_back(std::move(st));
        }
        p_.outputs.push_back(s);
    }

    Stmt stmt() {
        Stmt st;
        switch (cur().t) {
            case T::Stream: {
                next();
                st.name = name(want(T::Ident));
                if (cur().t == T::LBracket) {
                    next(); st.k = Stmt::ArrDecl; st.i1 = iexpr(); want(T::RBracket);
                    p_.is_array[st.name] = true;
                    return st;
                }
                st.k = Stmt::Local;
                if (cur().t == T::Assign) { next(); st.e = sexpr(0); }
                return st;
            }
            case T::Int: {
                next();
                st.k = Stmt::IntDecl;
                st.name = name(want(T::Ident));
                mark_int(st.name);
                want(T::Assign);
                st.i1 = iexpr();
                return st;
            }
            case T::If: case T::While: {
                st.k = cur().t == T::If ? Stmt::If : Stmt::While;
                next(); want(T::LParen); st.e = sexpr(0); want(T::RParen);
                st.body = block();
                return st;
            }
            case T::For: {
                next();
                st.k = Stmt::For;
                st.name = name(want(T::Ident));
                mark_int(st.name);
                want(T::In); st.i1 = iexpr(); want(T::DotDot); st.i2 = iexpr();
                st.body = block();
                return st;
            }
            case T::Ident: {
                st.name = name(next());
                if (cur().t == T::LBracket) {
                    next(); st.k = Stmt::ArrAssign; st.i1 = iexpr(); want(T::RBracket);
                    want(T::Assign); st.e = sexpr(0);
                    return st;
                }
                want(T::Assign);
                if (is_int(st.name)) { st.k = Stmt::IntAssign; st.i1 = iexpr(); }
                else { st.k = Stmt::Assign; st.e = sexpr(0); }
                return st;
            }
            default:
                throw ParseError("line " + std::to_string(cur().line) + ": unexpected token in block: " +
                                 tname(cur().t) + " ('" + text(cur()) + "')");
        }
    }

    std::vector<Stmt> block() {
        want(T::LBrace);
        std::vector<Stmt> body;
        while (cur().t != T::RBrace) {
            switch (cur().t) {
                case T::Stream: case T::Int: case T::If: case T::While: case T::For: case T::Ident:
                    body.push_back(stmt());
                    break;
                default:
                    throw ParseError("line " + std::to_string(cur().line) + ": unexpected token in block: " +
                                     tname(cur().t) + " ('" + text(cur()) + "')");
            }
        }
        want(T::RBrace);
        return body;
    }

    // Binary stream operators by binding strength: | 1, ^ 2, + 3, & 4.
    static int bin_level(T t) {
        switch (t) {
            case T::Pipe: return 1;
            case T::Caret: return 2;
            case T::Plus: return 3;
            case T::Amp: return 4;
            default: return 0;
        }
    }
    static SExpr::Kind bin_kind(T t) {
        return t == T::Pipe ? SExpr::Or : t == T::Caret ? SExpr::Xor : t == T::Plus ? SExpr::Plus : SExpr::And;
    }

    // Left-associative precedence climbing over | ^ + &, above the shift level.
    int32_t sexpr(int min_level) {
        int32_t lhs = shift_expr();
        for (;;) {
            int lv = bin_level(cur().t);
            if (lv == 0 || lv <= min_level) break;
            T op = next().t;
            int32_t rhs = sexpr(lv);
            SExpr e; e.k = bin_kind(op); e.a = lhs; e.b = rhs;
            lhs = new_s(e);
        }
        return lhs;
    }

    int32_t shift_expr() {
        int32_t lhs = unary();
        while (cur().t == T::LShift || cur().t == T::RShift) {
            SExpr e; e.k = next().t == T::LShift ? SExpr::Shl : SExpr::Shr;
            e.a = lhs;
            e.idx = iexpr();
            lhs = new_s(e);
        }
        return lhs;
    }

    int32_t unary() {
        if (cur().t == T::Tilde) {
            next();
            SExpr e; e.k = SExpr::Not; e.a = unary();
            return new_s(e);
        }
        return primary();
    }

    int32_t primary() {
        SExpr e;
        switch (cur().t) {
            case T::Zero: next(); e.k = SExpr::Zero; return new_s(e);
            case T::Ones: next(); e.k = SExpr::Ones; return new_s(e);
            case T::LParen: { next(); int32_t in = sexpr(0); want(T::RParen); return in; }
            case T::Ident: {
                e.name = name(next());
                if (cur().t == T::LBracket) { next(); e.k = SExpr::Arr; e.idx = iexpr(); want(T::RBracket); }
                else e.k = SExpr::Var;
                return new_s(e);
            }
            default:
                throw ParseError("line " + std::to_string(cur().line) + ": expected stream expression, got " +
                                 tname(cur().t) + " ('" + text(cur()) + "')");
        }
    }

    int32_t iexpr() {
        int32_t lhs = iterm();
        while (cur().t == T::Plus || cur().t == T::Minus) {
            IExpr e; e.k = next().t == T::Plus ? IExpr::Add : IExpr::Sub;
            e.l = lhs; e.r = iterm();
            lhs = new_i(e);
        }
        return lhs;
    }

    int32_t iterm() {
        int32_t lhs = iatom();
        while (cur().t == T::Star) {
            next();
            IExpr e; e.k = IExpr::Mul; e.l = lhs; e.r = iatom();
            lhs = new_i(e);
        }
        return lhs;
    }

    int32_t iatom() {
        IExpr e;
        switch (cur().t) {
            case T::Lit: {
                std::string v = text(next());
                // std::stoi (parser.cpp:323-326): an int32 literal or an error.
                size_t nz = v.find_first_not_of('0');
                std::string sig = nz == std::string::npos ? "0" : v.substr(nz);
                if (sig.size() > 10 || std::stoll(sig) > 2147483647LL) throw ParseError("stoi");
                e.k = IExpr::Lit; e.lit = std::stoll(sig);
                return new_i(e);
            }
            case T::LParen: { next(); int32_t in = iexpr(); want(T::RParen); return in; }
            case T::Popcount: {
                next(); want(T::LParen);
                e.k = IExpr::Pop; e.arg = sexpr(0);
                want(T::RParen);
                return new_i(e);
            }
            case T::Ident: e.k = IExpr::Var; e.var = name(next()); return new_i(e);
            default:
                throw ParseError("line " + std::to_string(cur().line) + ": expected integer expression, got " +
                                 tname(cur().t) + " ('" + text(cur()) + "')");
        }
    }
};

void check_rhs(const Program& p, int32_t id, const std::string& ctx) {
    const SExpr& e = p.sex[id];
    if (e.primary()) return;
    switch (e.k) {
        case SExpr::Not:
            if (!p.sex[e.a].primary()) throw ParseError(ctx + ": unary operand must be a primary");
            return;
        case SExpr::Shl: case SExpr::Shr:
            if (!p.sex[e.a].primary()) throw ParseError(ctx + ": shift operand must be a primary");
            return;
        default:
            if (!p.sex[e.a].primary()) throw ParseError(ctx + ": binary left operand must be a primary");
            if (!p.sex[e.b].primary()) throw ParseError(ctx + ": binary right operand must be a primary");
    }
}

void check_block(const Program& p, const std::vector<Stmt>& ss, const std::string& ctx) {
    for (const Stmt& s : ss) {
        const std::string& nm = p.syms.names[s.name];
        switch (s.k) {
            case Stmt::Local: if (s.e >= 0) check_rhs(p, s.e, ctx + ": stream " + nm); break;
            case Stmt::Assign: check_rhs(p, s.e, ctx + ": " + nm); break;
            case Stmt::ArrAssign: check_rhs(p, s.e, ctx + ": " + nm + "[...]"); break;
            case Stmt::IntDecl: case Stmt::IntAssign: {
                const IExpr& ie = p.iex[s.i1];
                if (ie.k == IExpr::Pop && p.sex[ie.arg].k != SExpr::Var)
                    throw ParseError(ctx + ": popcount arg must be a variable");
                break;
            }
            case Stmt::If:
                if (p.sex[s.e].k != SExpr::Var) throw ParseError(ctx + ": if condition must be a variable");
                check_block(p, s.body, ctx);
                break;
            case Stmt::While:
                if (p.sex[s.e].k != SExpr::Var) throw ParseError(ctx + ": while condition must be a variable");
                check_block(p, s.body, ctx);
                break;
            case Stmt::For: check_block(p, s.body, ctx); break;
            default: break;
        }
    }
}

struct StaticCounter {
    const Program& p;
    std::unordered_map<Sym, int64_t> env;
    const std::unordered_map<std::string, int64_t>& params;

    int64_t ev(int32_t id) {
        const IExpr& e = p.iex[id];
        switch (e.k) {
            case IExpr::Lit: return e.lit;
            case IExpr::Var: {
                auto it = env.find(e.var);
                if (it != env.end()) return it->second;
                auto pi = params.find(p.syms.names[e.var]);
                return pi == params.end() ? 0 : pi->second;
            }
            case IExpr::Add: return ev(e.l) + ev(e.r);
            case IExpr::Sub: return ev(e.l) - ev(e.r);
            case IExpr::Mul: return ev(e.l) * ev(e.r);
            default: return 0;
        }
    }
    bool compute(int32_t sid) const { return !p.sex[sid].primary(); }
    int64_t count(const std::vector<Stmt>& ss) {
        int64_t t = 0;
        for (const Stmt& s : ss) {
            switch (s.k) {
                case Stmt::Local: t += (s.e >= 0 && compute(s.e)) ? 1 : 0; break;
                case Stmt::Assign: case Stmt::ArrAssign: t += compute(s.e) ? 1 : 0; break;
                case Stmt::IntDecl: case Stmt::IntAssign: t += p.iex[s.i1].k == IExpr::Pop ? 1 : 0; break;
                case Stmt::If: case Stmt::While: t += count(s.body); break;
                case Stmt::For: {
                    int64_t lo = ev(s.i1), hi = ev(s.i2);
                    if (hi <= lo) break;
                    auto saved = env;
                    env[s.name] = lo;
                    int64_t b = count(s.body);
                    env = saved;
                    t += b * (hi - lo);
                    break;
                }
                default: break;
            }
        }
        return t;
    }
};

}  // namespace

std::unique_ptr<Program> parse_program(const std::string& source) {
    auto p = std::make_unique<Program>();
    p->source_bytes = source.size();
    Parser(source, *p).run();
    p->is_array.resize(p->syms.names.size(), false);
    check_block(*p, p->stmts, p->inputs.empty() ? "<program>" : p->syms.names[p->inputs[0]]);
    return p;
}

int64_t count_static_ops(const Program& p, const std::unordered_map<std::string, int64_t>& params) {
    StaticCounter c{p, {}, params};
    return c.count(p.stmts);
}

}  // namespace bsz

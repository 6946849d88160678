// bsz_oracle.cpp — CPU oracle and CPU baseline for the bitstream execution path.
//
// TEST INFRASTRUCTURE ONLY (see bsz_oracle.h).  Loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
// shipped library.
//
// Each part names the reference code it restates.  Parity pinning
// (tests/test_oracle_*.py):
//   tokenize/parse/validate  — every committed reference fixture program
//                              parses; malformed programs rejected with the
//                              same error class (parser.cpp:375-447).
//   run() control plane      — the reference's own committed fixtures
//                              (benchmark/apps/*/datasets/tests/*.bsdata,
//                              expected outputs inside) and oracle/_ref
//                              (the reference's Interpreter<CpuBackend>,
//                              compiled from its sources) on random programs.
//   word ops                 — closed forms: shl/shr against big-int shifts,
//                              add against Python big-int addition, popcount
//                              against bin().count; the test_ops.cpp shift
//                              and carry edge cases; cross-variant equality.
// No function is left unpinned.

#include "bsz_oracle.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ════════════════════════════════════════════════════════════════════════
// Word kernels
// ════════════════════════════════════════════════════════════════════════

bool cpu_has_avx512() {
    static int v = -1;
    if (v < 0) {
        __builtin_cpu_init();
        v = (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vpopcntdq")) ? 1 : 0;
    }
    return v == 1;
}

constexpr uint64_t kOmpMinWords = 1u << 14;  // below this OpenMP costs more than it saves

// ── pointwise: scalar ──
enum PwOp { PW_AND, PW_OR, PW_XOR, PW_NOT };

void pw_scalar(PwOp op, uint64_t* d, const uint64_t* a, const uint64_t* b, uint64_t n) {
    switch (op) {
        case PW_AND: for (uint64_t i = 0; i < n; ++i) d[i] = a[i] & b[i]; break;
        case PW_OR:  for (uint64_t i = 0; i < n; ++i) d[i] = a[i] | b[i]; break;
        case PW_XOR: for (uint64_t i = 0; i < n; ++i) d[i] = a[i] ^ b[i]; break;
        case PW_NOT: for (uint64_t i = 0; i < n; ++i) d[i] = ~a[i]; break;
    }
}

// ── pointwise: AVX-512, one 8-word vector per step (bitstream_ops.cpp:403-418) ──
__attribute__((target("avx512f")))
void pw_avx512_range(PwOp op, uint64_t* d, const uint64_t* a, const uint64_t* b,
                     uint64_t lo, uint64_t hi) {
    uint64_t i = lo;
    const __m512i ones = _mm512_set1_epi64(-1);
    for (; i + 8 <= hi; i += 8) {
        __m512i va = _mm512_loadu_si512(a + i), r;
        switch (op) {
            case PW_AND: r = _mm512_and_si512(va, _mm512_loadu_si512(b + i)); break;
            case PW_OR:  r = _mm512_or_si512(va, _mm512_loadu_si512(b + i)); break;
            case PW_XOR: r = _mm512_xor_si512(va, _mm512_loadu_si512(b + i)); break;
            default:     r = _mm512_xor_si512(va, ones); break;
        }
        _mm512_storeu_si512(d + i, r);
    }
    if (i < hi) pw_scalar(op, d + i, a + i, b ? b + i : nullptr, hi - i);
}

void pw_parallel(PwOp op, uint64_t* d, const uint64_t* a, const uint64_t* b, uint64_t n) {
    const bool avx = cpu_has_avx512();
    if (n < kOmpMinWords) {
        if (avx) pw_avx512_range(op, d, a, b, 0, n); else pw_scalar(op, d, a, b, n);
        return;
    }
    const int64_t nblk = (int64_t)((n + 4095) / 4096);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nblk; ++blk) {
        uint64_t lo = (uint64_t)blk * 4096, hi = std::min<uint64_t>(n, lo + 4096);
        if (avx) pw_avx512_range(op, d, a, b, lo, hi);
        else pw_scalar(op, d + lo, a + lo, b ? b + lo : nullptr, hi - lo);
    }
}

// ── shifts (bitstream_ops.cpp:120-159: backward shl so dst may alias src) ──
void shl_serial(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) {
    if (n == 0) return;
    if (k == 0) { if (dst != src) std::memmove(dst, src, n * 8); return; }
    const uint64_t woff = k / 64, boff = k % 64;
    if (woff >= n) { std::memset(dst, 0, n * 8); return; }
    if (boff == 0) {
        for (uint64_t i = n; i-- > woff;) dst[i] = src[i - woff];
    } else {
        for (uint64_t i = n; i-- > woff + 1;)
            dst[i] = (src[i - woff] << boff) | (src[i - woff - 1] >> (64 - boff));
        dst[woff] = src[0] << boff;
    }
    std::memset(dst, 0, woff * 8);
}

void shr_serial(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) {
    if (n == 0) return;
    if (k == 0) { if (dst != src) std::memmove(dst, src, n * 8); return; }
    const uint64_t woff = k / 64, boff = k % 64;
    if (woff >= n) { std::memset(dst, 0, n * 8); return; }
    const uint64_t limit = n - woff;
    if (boff == 0) {
        for (uint64_t i = 0; i < limit; ++i) dst[i] = src[i + woff];
    } else {
        for (uint64_t i = 0; i + 1 < limit; ++i)
            dst[i] = (src[i + woff] >> boff) | (src[i + woff + 1] << (64 - boff));
        dst[limit - 1] = src[n - 1] >> boff;
    }
    std::memset(dst + limit, 0, woff * 8);
}

// Out-of-place parallel shifts (dst != src), each output word from two inputs.
void shl_parallel(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) {
    const uint64_t woff = k / 64, boff = k % 64;
#pragma omp parallel for schedule(static) if (n >= kOmpMinWords)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        uint64_t hi = (i >= woff) ? src[i - woff] : 0;
        uint64_t lo = (i >= woff + 1) ? src[i - woff - 1] : 0;
        dst[i] = boff ? ((hi << boff) | (lo >> (64 - boff))) : hi;
    }
}

void shr_parallel(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) {
    const uint64_t woff = k / 64, boff = k % 64;
#pragma omp parallel for schedule(static) if (n >= kOmpMinWords)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        uint64_t lo = (i + woff < n) ? src[i + woff] : 0;
        uint64_t hi = (i + woff + 1 < n) ? src[i + woff + 1] : 0;
        dst[i] = boff ? ((lo >> boff) | (hi << (64 - boff))) : lo;
    }
}

// ── add: ripple __int128 carry chain (bitstream_ops.cpp:59-66) ──
void add_ripple(uint64_t* s, const uint64_t* a, const uint64_t* b, uint64_t n) {
    uint64_t carry = 0;
    for (uint64_t i = 0; i < n; ++i) {
        unsigned __int128 t = (unsigned __int128)a[i] + b[i] + carry;
        s[i] = (uint64_t)t;
        carry = (uint64_t)(t >> 64);
    }
}

// ── add: chunked parallel carry resolution (best-CPU mode) ──
// Each thread adds its chunk with carry-in 0 and records (carry-out, whether
// every sum word is all-ones).  A serial scan over the chunk summaries gives
// each chunk its carry-in; a chunk with carry-in 1 adds it from its first
// word until a word does not wrap.
void add_chunked(uint64_t* s, const uint64_t* a, const uint64_t* b, uint64_t n) {
    if (n < kOmpMinWords) { add_ripple(s, a, b, n); return; }
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    const uint64_t chunk = (n + nt - 1) / nt;
    std::vector<uint8_t> cout(nt, 0), allones(nt, 0), cin(nt, 0);
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        uint64_t lo = std::min<uint64_t>(n, (uint64_t)t * chunk), hi = std::min<uint64_t>(n, lo + chunk);
        uint64_t carry = 0, all = ~0ull;
        for (uint64_t i = lo; i < hi; ++i) {
            unsigned __int128 v = (unsigned __int128)a[i] + b[i] + carry;
            s[i] = (uint64_t)v;
            carry = (uint64_t)(v >> 64);
            all &= s[i];
        }
        cout[t] = (uint8_t)carry;
        allones[t] = (lo < hi && all == ~0ull) ? 1 : 0;
#pragma omp barrier
#pragma omp single
        {
            uint8_t c = 0;
            for (int j = 0; j < nt; ++j) {
                cin[j] = c;
                c = (uint8_t)(cout[j] | (allones[j] & c));
            }
        }
        if (cin[t]) {
            for (uint64_t i = lo; i < hi; ++i) {
                s[i] += 1;
                if (s[i] != 0) break;
            }
        }
    }
}

// ── popcount (bitstream_ops.cpp:474-487 VPOPCNTQ + OpenMP reduction) ──
__attribute__((target("avx512f,avx512vpopcntdq")))
int64_t popcnt_avx512_range(const uint64_t* a, uint64_t lo, uint64_t hi) {
    __m512i acc = _mm512_setzero_si512();
    uint64_t i = lo;
    for (; i + 8 <= hi; i += 8) acc = _mm512_add_epi64(acc, _mm512_popcnt_epi64(_mm512_loadu_si512(a + i)));
    int64_t c = _mm512_reduce_add_epi64(acc);
    for (; i < hi; ++i) c += __builtin_popcountll(a[i]);
    return c;
}

int64_t popcnt_words(const uint64_t* a, uint64_t n, bool parallel) {
    const bool avx = cpu_has_avx512();
    if (!parallel || n < kOmpMinWords) {
        if (avx) return popcnt_avx512_range(a, 0, n);
        int64_t c = 0;
        for (uint64_t i = 0; i < n; ++i) c += __builtin_popcountll(a[i]);
        return c;
    }
    int64_t total = 0;
    const int64_t nblk = (int64_t)((n + 4095) / 4096);
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int64_t blk = 0; blk < nblk; ++blk) {
        uint64_t lo = (uint64_t)blk * 4096, hi = std::min<uint64_t>(n, lo + 4096);
        if (avx) total += popcnt_avx512_range(a, lo, hi);
        else for (uint64_t i = lo; i < hi; ++i) total += __builtin_popcountll(a[i]);
    }
    return total;
}

bool nonzero_words(const uint64_t* a, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) if (a[i]) return true;
    return false;
}

// Reductions bounded to the logical bitlength (cpu_backend.h:111-129).
int64_t popcount_bounded(const uint64_t* a, int64_t bitlength, bool parallel) {
    if (bitlength <= 0) return 0;
    uint64_t lw = (uint64_t)(bitlength + 63) / 64, rem = (uint64_t)bitlength % 64;
    if (rem == 0) return popcnt_words(a, lw, parallel);
    int64_t c = lw > 1 ? popcnt_words(a, lw - 1, parallel) : 0;
    return c + __builtin_popcountll(a[lw - 1] & ((1ull << rem) - 1));
}

bool is_nonzero_bounded(const uint64_t* a, int64_t bitlength) {
    if (bitlength <= 0) return false;
    uint64_t lw = (uint64_t)(bitlength + 63) / 64, rem = (uint64_t)bitlength % 64;
    if (rem == 0) return nonzero_words(a, lw);
    if (lw > 1 && nonzero_words(a, lw - 1)) return true;
    return (a[lw - 1] & ((1ull << rem) - 1)) != 0;
}

// ════════════════════════════════════════════════════════════════════════
// Front-end: tokens, AST, parser, 3-address validator
// ════════════════════════════════════════════════════════════════════════

enum Tok {
    T_INPUT, T_OUTPUT, T_PARAM, T_STREAM, T_INT, T_IF, T_WHILE, T_FOR, T_IN, T_ZERO, T_ONES,
    T_POPCOUNT, T_IDENT, T_LIT, T_LBRACE, T_RBRACE, T_LPAREN, T_RPAREN, T_LBRACKET,
    T_RBRACKET, T_ASSIGN, T_TILDE, T_AMP, T_PIPE, T_CARET, T_LSHIFT, T_RSHIFT, T_PLUS,
    T_MINUS, T_STAR, T_DOTDOT, T_EOF
};

const char* tok_name(Tok t) {
    static const char* names[] = {
        "INPUT", "OUTPUT", "PARAM", "STREAM", "INT", "IF", "WHILE", "FOR", "IN", "ZERO", "ONES",
        "POPCOUNT", "IDENT", "INT_LIT", "LBRACE", "RBRACE", "LPAREN", "RPAREN", "LBRACKET",
        "RBRACKET", "ASSIGN", "TILDE", "AMP", "PIPE", "CARET", "LSHIFT", "RSHIFT", "PLUS",
        "MINUS", "STAR", "DOTDOT", "EOF"};
    return names[t];
}

struct Token { Tok t; std::string v; int line; };

struct ParseError : std::runtime_error { using std::runtime_error::runtime_error; };
struct RunError : std::runtime_error { using std::runtime_error::runtime_error; };

std::vector<Token> tokenize(const std::string& s) {
    static const std::map<std::string, Tok> kw = {
        {"input", T_INPUT}, {"output", T_OUTPUT}, {"param", T_PARAM}, {"stream", T_STREAM},
        {"int", T_INT}, {"if", T_IF}, {"while", T_WHILE}, {"for", T_FOR}, {"in", T_IN},
        {"ZERO", T_ZERO}, {"ONES", T_ONES}, {"popcount", T_POPCOUNT}};
    std::vector<Token> out;
    int line = 1;
    size_t i = 0, n = s.size();
    auto ident_start = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; };
    auto digit = [](char c) { return c >= '0' && c <= '9'; };
    while (i < n) {
        char c = s[i];
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') { if (c == '\n') ++line; ++i; continue; }
        if (c == '/' && i + 1 < n && s[i + 1] == '/') { while (i < n && s[i] != '\n') ++i; continue; }
        if (i + 1 < n) {
            std::string two = s.substr(i, 2);
            if (two == "<<") { out.push_back({T_LSHIFT, two, line}); i += 2; continue; }
            if (two == ">>") { out.push_back({T_RSHIFT, two, line}); i += 2; continue; }
            if (two == "..") { out.push_back({T_DOTDOT, two, line}); i += 2; continue; }
        }
        Tok single = T_EOF;
        switch (c) {
            case '{': single = T_LBRACE; break;   case '}': single = T_RBRACE; break;
            case '(': single = T_LPAREN; break;   case ')': single = T_RPAREN; break;
            case '[': single = T_LBRACKET; break; case ']': single = T_RBRACKET; break;
            case '=': single = T_ASSIGN; break;   case '~': single = T_TILDE; break;
            case '&': single = T_AMP; break;      case '|': single = T_PIPE; break;
            case '^': single = T_CARET; break;    case '+': single = T_PLUS; break;
            case '-': single = T_MINUS; break;    case '*': single = T_STAR; break;
            default: break;
        }
        if (single != T_EOF) { out.push_back({single, std::string(1, c), line}); ++i; continue; }
        if (digit(c)) {
            size_t b = i; while (i < n && digit(s[i])) ++i;
            out.push_back({T_LIT, s.substr(b, i - b), line}); continue;
        }
        if (ident_start(c)) {
            size_t b = i; while (i < n && (ident_start(s[i]) || digit(s[i]))) ++i;
            std::string w = s.substr(b, i - b);
            auto it = kw.find(w);
            out.push_back({it == kw.end() ? T_IDENT : it->second, w, line}); continue;
        }
        throw ParseError("line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'");
    }
    out.push_back({T_EOF, "", line});
    return out;
}

struct SExpr;
struct IExpr {
    enum K { LIT, VAR, BIN, POP } k;
    int64_t lit = 0;
    std::string name;
    char op = 0;
    std::shared_ptr<IExpr> l, r;
    std::shared_ptr<SExpr> arg;
};
struct SExpr {
    enum K { VAR, CONST, ARR, NOT, BIN, SHIFT } k;
    std::string name;          // VAR, ARR
    bool ones = false;         // CONST
    std::shared_ptr<IExpr> idx;  // ARR index / SHIFT amount
    char op = 0;               // BIN: & | ^ +   SHIFT: < (<<) or > (>>)
    std::shared_ptr<SExpr> a, b;
};
struct Stmt {
    enum K { ASSIGN, ARR_ASSIGN, IF, WHILE, FOR, LOCAL, ARR_DECL, INT_DECL, INT_ASSIGN } k;
    std::string name;
    std::shared_ptr<SExpr> e;        // stream rhs / cond
    std::shared_ptr<IExpr> i1, i2;   // index / size / int rhs / for bounds
    std::vector<Stmt> body;
};
struct Program {
    std::vector<Stmt> stmts;
    std::vector<std::string> inputs, outputs, params, out_ints;
};

class Parser {
public:
    explicit Parser(std::vector<Token> t) : t_(std::move(t)) {}

    Program program() {
        Program p;
        while (!at(T_EOF)) {
            if (at(T_INPUT)) {
                take(); expect(T_STREAM);
                std::string name = expect(T_IDENT).v;
                if (at(T_LBRACKET)) { take(); iexpr(); expect(T_RBRACKET); }
                p.inputs.push_back(name);
            } else if (at(T_OUTPUT)) {
                take();
                if (at(T_INT)) {
                    take();
                    std::string name = expect(T_IDENT).v;
                    ints_.insert(name);
                    p.out_ints.push_back(name);
                    Stmt s; s.k = Stmt::INT_DECL; s.name = name; s.i1 = lit(0);
                    p.stmts.push_back(s);
                } else {
                    expect(T_STREAM);
                    std::string name = expect(T_IDENT).v;
                    if (at(T_LBRACKET)) {
                        take();
                        Stmt s; s.k = Stmt::ARR_DECL; s.name = name; s.i1 = iexpr();
                        expect(T_RBRACKET);
                        p.stmts.push_back(s);
                    }
                    p.outputs.push_back(name);
                }
            } else if (at(T_PARAM)) {
                take(); expect(T_INT);
                std::string name = expect(T_IDENT).v;
                ints_.insert(name);
                p.params.push_back(name);
            } else if (at(T_STREAM) || at(T_INT) || at(T_IF) || at(T_WHILE) || at(T_FOR) || at(T_IDENT)) {
                p.stmts.push_back(statement(false));
            } else {
                throw ParseError("line " + std::to_string(peek().line) + ": unexpected token " +
                                 tok_name(peek().t) + " ('" + peek().v + "')");
            }
        }
        return p;
    }

private:
    std::vector<Token> t_;
    size_t pos_ = 0;
    std::set<std::string> ints_;

    const Token& peek() const { return t_[pos_]; }
    bool at(Tok t) const { return peek().t == t; }
    Token take() { return t_[pos_++]; }
    Token expect(Tok t) {
        Token k = take();
        if (k.t != t)
            throw ParseError("line " + std::to_string(k.line) + ": expected " + tok_name(t) + ", got " +
                             tok_name(k.t) + " ('" + k.v + "')");
        return k;
    }
    static std::shared_ptr<IExpr> lit(int64_t v) {
        auto e = std::make_shared<IExpr>(); e->k = IExpr::LIT; e->lit = v; return e;
    }

    Stmt statement(bool in_block) {
        Stmt s;
        if (at(T_STREAM)) {
            take();
            s.name = expect(T_IDENT).v;
            if (at(T_LBRACKET)) {
                take(); s.k = Stmt::ARR_DECL; s.i1 = iexpr(); expect(T_RBRACKET);
                return s;
            }
            s.k = Stmt::LOCAL;
            if (at(T_ASSIGN)) { take(); s.e = sexpr(); }
            return s;
        }
        if (at(T_INT)) {
            take();
            s.k = Stmt::INT_DECL; s.name = expect(T_IDENT).v;
            ints_.insert(s.name);
            expect(T_ASSIGN);
            s.i1 = iexpr();
            return s;
        }
        if (at(T_IF) || at(T_WHILE)) {
            s.k = at(T_IF) ? Stmt::IF : Stmt::WHILE;
            take(); expect(T_LPAREN); s.e = sexpr(); expect(T_RPAREN);
            s.body = block();
            return s;
        }
        if (at(T_FOR)) {
            take(); s.k = Stmt::FOR;
            s.name = expect(T_IDENT).v;
            ints_.insert(s.name);
            expect(T_IN); s.i1 = iexpr(); expect(T_DOTDOT); s.i2 = iexpr();
            s.body = block();
            return s;
        }
        if (at(T_IDENT)) {
            s.name = take().v;
            if (at(T_LBRACKET)) {
                take(); s.k = Stmt::ARR_ASSIGN; s.i1 = iexpr(); expect(T_RBRACKET);
                expect(T_ASSIGN); s.e = sexpr();
                return s;
            }
            expect(T_ASSIGN);
            if (ints_.count(s.name)) { s.k = Stmt::INT_ASSIGN; s.i1 = iexpr(); }
            else { s.k = Stmt::ASSIGN; s.e = sexpr(); }
            return s;
        }
        (void)in_block;
        throw ParseError("line " + std::to_string(peek().line) + ": unexpected token in block: " +
                         tok_name(peek().t) + " ('" + peek().v + "')");
    }

    std::vector<Stmt> block() {
        expect(T_LBRACE);
        std::vector<Stmt> out;
        while (!at(T_RBRACE)) {
            if (at(T_STREAM) || at(T_INT) || at(T_IF) || at(T_WHILE) || at(T_FOR) || at(T_IDENT))
                out.push_back(statement(true));
            else
                throw ParseError("line " + std::to_string(peek().line) + ": unexpected token in block: " +
                                 tok_name(peek().t) + " ('" + peek().v + "')");
        }
        expect(T_RBRACE);
        return out;
    }

    // Stream expressions, lowest to highest precedence: | ^ + & (<< >>) ~
    std::shared_ptr<SExpr> bin(char op, std::shared_ptr<SExpr> a, std::shared_ptr<SExpr> b) {
        auto e = std::make_shared<SExpr>(); e->k = SExpr::BIN; e->op = op; e->a = a; e->b = b; return e;
    }
    std::shared_ptr<SExpr> sexpr() { return s_or(); }
    std::shared_ptr<SExpr> s_or() { auto l = s_xor(); while (at(T_PIPE)) { take(); l = bin('|', l, s_xor()); } return l; }
    std::shared_ptr<SExpr> s_xor() { auto l = s_add(); while (at(T_CARET)) { take(); l = bin('^', l, s_add()); } return l; }
    std::shared_ptr<SExpr> s_add() { auto l = s_and(); while (at(T_PLUS)) { take(); l = bin('+', l, s_and()); } return l; }
    std::shared_ptr<SExpr> s_and() { auto l = s_shift(); while (at(T_AMP)) { take(); l = bin('&', l, s_shift()); } return l; }
    std::shared_ptr<SExpr> s_shift() {
        auto l = s_unary();
        while (at(T_LSHIFT) || at(T_RSHIFT)) {
            char op = take().t == T_LSHIFT ? '<' : '>';
            auto e = std::make_shared<SExpr>(); e->k = SExpr::SHIFT; e->op = op; e->a = l; e->idx = iexpr();
            l = e;
        }
        return l;
    }
    std::shared_ptr<SExpr> s_unary() {
        if (at(T_TILDE)) {
            take();
            auto e = std::make_shared<SExpr>(); e->k = SExpr::NOT; e->a = s_unary(); return e;
        }
        return s_primary();
    }
    std::shared_ptr<SExpr> s_primary() {
        auto e = std::make_shared<SExpr>();
        if (at(T_ZERO) || at(T_ONES)) { e->k = SExpr::CONST; e->ones = take().t == T_ONES; return e; }
        if (at(T_LPAREN)) { take(); auto in = sexpr(); expect(T_RPAREN); return in; }
        if (at(T_IDENT)) {
            e->name = take().v;
            if (at(T_LBRACKET)) { take(); e->k = SExpr::ARR; e->idx = iexpr(); expect(T_RBRACKET); }
            else e->k = SExpr::VAR;
            return e;
        }
        throw ParseError("line " + std::to_string(peek().line) + ": expected stream expression, got " +
                         tok_name(peek().t) + " ('" + peek().v + "')");
    }

    // Integer expressions: + - over *
    std::shared_ptr<IExpr> ibin(char op, std::shared_ptr<IExpr> a, std::shared_ptr<IExpr> b) {
        auto e = std::make_shared<IExpr>(); e->k = IExpr::BIN; e->op = op; e->l = a; e->r = b; return e;
    }
    std::shared_ptr<IExpr> iexpr() {
        auto l = i_mul();
        while (at(T_PLUS) || at(T_MINUS)) { char op = take().v[0]; l = ibin(op, l, i_mul()); }
        return l;
    }
    std::shared_ptr<IExpr> i_mul() {
        auto l = i_primary();
        while (at(T_STAR)) { take(); l = ibin('*', l, i_primary()); }
        return l;
    }
    std::shared_ptr<IExpr> i_primary() {
        if (at(T_LIT)) {
            std::string v = take().v;
            // std::stoi semantics (parser.cpp:323-326): values beyond int32 are rejected.
            size_t nz = v.find_first_not_of('0');
            std::string sig = nz == std::string::npos ? "0" : v.substr(nz);
            if (sig.size() > 10 || std::stoll(sig) > 2147483647LL) throw ParseError("stoi");
            return lit(std::stoll(sig));
        }
        if (at(T_LPAREN)) { take(); auto e = iexpr(); expect(T_RPAREN); return e; }
        if (at(T_POPCOUNT)) {
            take(); expect(T_LPAREN);
            auto e = std::make_shared<IExpr>(); e->k = IExpr::POP; e->arg = sexpr();
            expect(T_RPAREN);
            return e;
        }
        if (at(T_IDENT)) {
            auto e = std::make_shared<IExpr>(); e->k = IExpr::VAR; e->name = take().v; return e;
        }
        throw ParseError("line " + std::to_string(peek().line) + ": expected integer expression, got " +
                         tok_name(peek().t) + " ('" + peek().v + "')");
    }
};

// 3-address validation (parser.cpp:375-447).
bool is_primary(const SExpr& e) { return e.k == SExpr::VAR || e.k == SExpr::CONST || e.k == SExpr::ARR; }

void validate_rhs(const SExpr& e, const std::string& ctx) {
    if (is_primary(e)) return;
    if (e.k == SExpr::NOT) { if (!is_primary(*e.a)) throw ParseError(ctx + ": unary operand must be a primary"); return; }
    if (e.k == SExpr::BIN) {
        if (!is_primary(*e.a)) throw ParseError(ctx + ": binary left operand must be a primary");
        if (!is_primary(*e.b)) throw ParseError(ctx + ": binary right operand must be a primary");
        return;
    }
    if (e.k == SExpr::SHIFT) { if (!is_primary(*e.a)) throw ParseError(ctx + ": shift operand must be a primary"); return; }
}

void validate_stmts(const std::vector<Stmt>& ss, const std::string& ctx) {
    for (const Stmt& s : ss) {
        switch (s.k) {
            case Stmt::LOCAL: if (s.e) validate_rhs(*s.e, ctx + ": stream " + s.name); break;
            case Stmt::ASSIGN: validate_rhs(*s.e, ctx + ": " + s.name); break;
            case Stmt::ARR_ASSIGN: validate_rhs(*s.e, ctx + ": " + s.name + "[...]"); break;
            case Stmt::INT_DECL: case Stmt::INT_ASSIGN:
                if (s.i1->k == IExpr::POP && s.i1->arg->k != SExpr::VAR)
                    throw ParseError(ctx + ": popcount arg must be a variable");
                break;
            case Stmt::IF: case Stmt::WHILE:
                if (s.e->k != SExpr::VAR)
                    throw ParseError(ctx + (s.k == Stmt::IF ? ": if condition must be a variable" : ": while condition must be a variable"));
                validate_stmts(s.body, ctx);
                break;
            case Stmt::FOR: validate_stmts(s.body, ctx); break;
            default: break;
        }
    }
}

// ════════════════════════════════════════════════════════════════════════
// Interpreter (interpreter.inl:17-364, non-reuse mode)
// ════════════════════════════════════════════════════════════════════════

struct Buf {
    uint64_t* p = nullptr;
    uint64_t n = 0;
    Buf() = default;
    explicit Buf(uint64_t nw) : n(nw) {
        size_t bytes = std::max<uint64_t>(64, ((nw * 8 + 63) / 64) * 64);
        p = static_cast<uint64_t*>(std::aligned_alloc(64, bytes));
        std::memset(p, 0, bytes);
    }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    Buf& operator=(Buf&& o) noexcept { if (this != &o) { std::free(p); p = o.p; n = o.n; o.p = nullptr; o.n = 0; } return *this; }
    ~Buf() { std::free(p); }
};

constexpr int kMaxWhileIters = 10000;  // interpreter.h:109

struct Machine {
    int64_t bitlength = 0;
    uint64_t lw = 1;       // logical words
    uint64_t nw = 8;       // allocated words (lw rounded up to 8, cpu_backend.h:39)
    int sem = BSZO_SEM_CLEAN;
    int variant = BSZO_SCALAR;
    int64_t ops = 0;

    std::unordered_map<std::string, Buf> vars;
    std::map<std::string, std::map<int64_t, Buf>> arrays;
    std::unordered_map<std::string, int64_t> ints;
    std::map<std::string, std::vector<bool>> bcast;
    Buf zero, ones;

    void init_consts() {
        zero = Buf(nw);
        ones = Buf(nw);
        fill_ones(ones);
    }

    // Clean semantics: nothing at or above bitlength (DSL.md:273-275).
    // BSIM32: bsim_cuda with 32-bit words pads each stream to a multiple of 8
    // such words (cuda_backend.cuh:23-25), 256 bits, so the dirty region ends
    // there rather than at the 512-bit multiple of the 64-bit layout.
    void clean(Buf& b) const {
        if (sem == BSZO_SEM_BSIM32) {
            uint64_t w32 = bitlength > 0 ? ((uint64_t)bitlength + 31) / 32 : 2 * lw;
            uint64_t keep = ((w32 + 7) & ~uint64_t(7)) / 2;
            if (nw > keep) std::memset(b.p + keep, 0, (nw - keep) * 8);
            return;
        }
        if (sem != BSZO_SEM_CLEAN || bitlength <= 0) return;
        uint64_t rem = (uint64_t)bitlength % 64;
        if (rem) b.p[lw - 1] &= (1ull << rem) - 1;
        if (nw > lw) std::memset(b.p + lw, 0, (nw - lw) * 8);
    }
    void fill_ones(Buf& b) const { std::memset(b.p, 0xff, nw * 8); clean(b); }

    Buf& var(const std::string& name) {
        auto it = vars.find(name);
        if (it != vars.end()) return it->second;
        return vars.emplace(name, Buf(nw)).first->second;
    }

    int64_t eval_int(const IExpr& e) {
        switch (e.k) {
            case IExpr::LIT: return e.lit;
            case IExpr::VAR: {
                auto it = ints.find(e.name);
                if (it == ints.end()) throw RunError("undefined int variable: " + e.name);
                return it->second;
            }
            case IExpr::BIN: {
                int64_t l = eval_int(*e.l), r = eval_int(*e.r);
                if (e.op == '+') return (int64_t)((uint64_t)l + (uint64_t)r);
                if (e.op == '-') return (int64_t)((uint64_t)l - (uint64_t)r);
                return (int64_t)((uint64_t)l * (uint64_t)r);
            }
            case IExpr::POP: {
                if (e.arg->k != SExpr::VAR) throw RunError("bad variant access");
                Buf& v = var(e.arg->name);
                ++ops;
                return popcount_bounded(v.p, bitlength, variant != BSZO_SCALAR);
            }
        }
        return 0;
    }

    // get_operand (interpreter.inl:60-90): broadcast metadata first, then arrays.
    const Buf& operand(const SExpr& e) {
        if (e.k == SExpr::VAR) return var(e.name);
        if (e.k == SExpr::CONST) return e.ones ? ones : zero;
        if (e.k == SExpr::ARR) {
            int64_t idx = (int32_t)eval_int(*e.idx);
            auto b = bcast.find(e.name);
            if (b != bcast.end()) {
                if (idx >= 0 && idx < (int64_t)b->second.size()) return b->second[idx] ? ones : zero;
                return zero;
            }
            auto a = arrays.find(e.name);
            if (a == arrays.end()) throw RunError("undefined array: " + e.name);
            auto it = a->second.find(idx);
            return it != a->second.end() ? it->second : zero;
        }
        throw RunError("get_operand: not a primary expression");
    }

    void pointwise(PwOp op, Buf& d, const Buf& a, const Buf* b) {
        if (variant == BSZO_SCALAR) pw_scalar(op, d.p, a.p, b ? b->p : nullptr, nw);
        else pw_parallel(op, d.p, a.p, b ? b->p : nullptr, nw);
    }

    // exec_rhs (interpreter.inl:93-157).
    void exec_rhs(const SExpr& e, Buf& dst) {
        switch (e.k) {
            case SExpr::VAR: {
                Buf& src = var(e.name);
                if (&src != &dst) std::memcpy(dst.p, src.p, nw * 8);
                return;
            }
            case SExpr::CONST:
                if (e.ones) fill_ones(dst); else std::memset(dst.p, 0, nw * 8);
                return;
            case SExpr::ARR: {
                int64_t idx = (int32_t)eval_int(*e.idx);
                auto a = arrays.find(e.name);
                if (a != arrays.end()) {
                    auto it = a->second.find(idx);
                    if (it != a->second.end()) { if (&it->second != &dst) std::memcpy(dst.p, it->second.p, nw * 8); return; }
                }
                auto b = bcast.find(e.name);
                if (b != bcast.end() && idx >= 0 && idx < (int64_t)b->second.size()) {
                    if (b->second[idx]) fill_ones(dst); else std::memset(dst.p, 0, nw * 8);
                    return;
                }
                std::memset(dst.p, 0, nw * 8);
                return;
            }
            case SExpr::NOT: {
                const Buf& s = operand(*e.a);
                pointwise(PW_NOT, dst, s, nullptr);
                clean(dst);
                ++ops;
                return;
            }
            case SExpr::BIN: {
                const Buf& l = operand(*e.a);
                const Buf& r = operand(*e.b);
                if (e.op == '&') pointwise(PW_AND, dst, l, &r);
                else if (e.op == '|') pointwise(PW_OR, dst, l, &r);
                else if (e.op == '^') pointwise(PW_XOR, dst, l, &r);
                else {
                    if (variant == BSZO_BEST) {
                        if (&dst == &l || &dst == &r) { Buf t(nw); add_chunked(t.p, l.p, r.p, nw); std::memcpy(dst.p, t.p, nw * 8); }
                        else add_chunked(dst.p, l.p, r.p, nw);
                    } else {
                        add_ripple(dst.p, l.p, r.p, nw);
                    }
                    clean(dst);
                }
                ++ops;
                return;
            }
            case SExpr::SHIFT: {
                const Buf& s = operand(*e.a);
                uint64_t k = (uint64_t)eval_int(*e.idx);  // static_cast<size_t> (interpreter.inl:140)
                if (e.op == '<') {
                    if (variant == BSZO_BEST && &s != &dst) shl_parallel(dst.p, s.p, k, nw);
                    else shl_serial(dst.p, s.p, k, nw);
                    clean(dst);
                } else {
                    if (variant == BSZO_SCALAR || &s == &dst) shr_serial(dst.p, s.p, k, nw);
                    else shr_parallel(dst.p, s.p, k, nw);
                }
                ++ops;
                return;
            }
        }
    }

    void exec(const Stmt& s) {
        switch (s.k) {
            case Stmt::LOCAL: {
                Buf& b = var(s.name);
                if (s.e) exec_rhs(*s.e, b); else std::memset(b.p, 0, nw * 8);
                return;
            }
            case Stmt::ARR_DECL: {
                int64_t size = (int32_t)eval_int(*s.i1);
                auto& arr = arrays[s.name];
                for (int64_t i = 0; i < size; ++i) {
                    auto it = arr.find(i);
                    if (it != arr.end()) std::memset(it->second.p, 0, nw * 8);
                    else arr.emplace(i, Buf(nw));
                }
                return;
            }
            case Stmt::ASSIGN: exec_rhs(*s.e, var(s.name)); return;
            case Stmt::ARR_ASSIGN: {
                int64_t idx = (int32_t)eval_int(*s.i1);
                auto& arr = arrays[s.name];
                auto it = arr.find(idx);
                if (it == arr.end()) it = arr.emplace(idx, Buf(nw)).first;
                exec_rhs(*s.e, it->second);
                return;
            }
            case Stmt::IF: {
                Buf& c = var(s.e->name);
                if (is_nonzero_bounded(c.p, bitlength)) for (const Stmt& b : s.body) exec(b);
                return;
            }
            case Stmt::WHILE: {
                for (int it = 0; it < kMaxWhileIters; ++it) {
                    Buf& c = var(s.e->name);
                    if (!is_nonzero_bounded(c.p, bitlength)) return;
                    for (const Stmt& b : s.body) exec(b);
                }
                throw RunError("while loop exceeded max iterations");
            }
            case Stmt::FOR: {
                int32_t lo = (int32_t)eval_int(*s.i1), hi = (int32_t)eval_int(*s.i2);
                for (int32_t i = lo; i < hi; ++i) {
                    ints[s.name] = i;
                    for (const Stmt& b : s.body) exec(b);
                }
                return;
            }
            case Stmt::INT_DECL: case Stmt::INT_ASSIGN:
                ints[s.name] = eval_int(*s.i1);
                return;
        }
    }
};

// Static op count (ast_liveness.cpp:629-719).
struct OpCounter {
    std::unordered_map<std::string, int64_t> params, env;
    int64_t eval(const IExpr& e) {
        switch (e.k) {
            case IExpr::LIT: return e.lit;
            case IExpr::VAR: {
                auto it = env.find(e.name); if (it != env.end()) return it->second;
                auto p = params.find(e.name); if (p != params.end()) return p->second;
                return 0;
            }
            case IExpr::BIN: {
                int64_t l = eval(*e.l), r = eval(*e.r);
                return e.op == '+' ? l + r : e.op == '-' ? l - r : l * r;
            }
            default: return 0;
        }
    }
    static bool compute(const SExpr& e) { return !(e.k == SExpr::VAR || e.k == SExpr::CONST || e.k == SExpr::ARR); }
    int64_t count(const Stmt& s) {
        switch (s.k) {
            case Stmt::LOCAL: return (s.e && compute(*s.e)) ? 1 : 0;
            case Stmt::ASSIGN: case Stmt::ARR_ASSIGN: return compute(*s.e) ? 1 : 0;
            case Stmt::INT_DECL: case Stmt::INT_ASSIGN: return s.i1->k == IExpr::POP ? 1 : 0;
            case Stmt::FOR: {
                int64_t lo = eval(*s.i1), hi = eval(*s.i2);
                if (hi <= lo) return 0;
                auto saved = env;
                env[s.name] = lo;
                int64_t b = 0;
                for (const Stmt& x : s.body) b += count(x);
                env = saved;
                return b * (hi - lo);
            }
            case Stmt::IF: case Stmt::WHILE: {
                int64_t b = 0;
                for (const Stmt& x : s.body) b += count(x);
                return b;
            }
            default: return 0;
        }
    }
};

}  // namespace

// ════════════════════════════════════════════════════════════════════════
// C API
// ════════════════════════════════════════════════════════════════════════

struct bszo_case {
    Program prog;
    int64_t bitlength = 0;
    int sem = BSZO_SEM_CLEAN;
    std::unordered_map<std::string, int64_t> params;
    std::unordered_map<std::string, std::vector<uint64_t>> inputs;
    std::map<std::string, std::map<int64_t, std::vector<uint64_t>>> input_arrays;
    std::map<std::string, std::vector<bool>> bcast;
    std::unique_ptr<Machine> m;
    int64_t op_count = 0, static_ops = 0;
    double exec_ms = 0;
    std::string out_streams, out_ints;
};

static void set_err(char* err, size_t len, const std::string& msg) {
    if (err && len) { std::snprintf(err, len, "%s", msg.c_str()); }
}

extern "C" {

bszo_case* bszo_case_new(const char* source, char* err, size_t errlen) {
    try {
        auto c = std::make_unique<bszo_case>();
        Parser p(tokenize(source));
        c->prog = p.program();
        validate_stmts(c->prog.stmts, c->prog.inputs.empty() ? "<program>" : c->prog.inputs[0]);
        for (auto& s : c->prog.outputs) c->out_streams += s + "\n";
        for (auto& s : c->prog.out_ints) c->out_ints += s + "\n";
        return c.release();
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

void bszo_case_free(bszo_case* c) { delete c; }

int bszo_set_bitlength(bszo_case* c, int64_t bl) { c->bitlength = bl; return 0; }
int bszo_set_semantics(bszo_case* c, int sem) { c->sem = sem; return 0; }
int bszo_set_param(bszo_case* c, const char* name, int64_t v) { c->params[name] = v; return 0; }
int bszo_set_input(bszo_case* c, const char* name, const uint64_t* w, uint64_t n) {
    c->inputs[name].assign(w, w + n); return 0;
}
int bszo_set_array_elem(bszo_case* c, const char* name, int64_t idx, const uint64_t* w, uint64_t n) {
    c->input_arrays[name][idx].assign(w, w + n); return 0;
}
int bszo_set_broadcast(bszo_case* c, const char* name, const uint8_t* bits, int64_t count) {
    std::vector<bool> v(count);
    for (int64_t i = 0; i < count; ++i) v[i] = (bits[i / 8] >> (i % 8)) & 1;
    c->bcast[name] = std::move(v);
    return 0;
}

int bszo_run(bszo_case* c, int variant, char* err, size_t errlen) {
    try {
        auto m = std::make_unique<Machine>();
        m->bitlength = c->bitlength;
        m->sem = c->sem;
        m->variant = variant;
        // n_words from bitlength, else from the longest input (main.cu:40-49).
        uint64_t lw = 0;
        if (c->bitlength > 0) lw = (uint64_t)(c->bitlength + 63) / 64;
        else {
            lw = 1;
            for (auto& kv : c->inputs) lw = std::max<uint64_t>(lw, kv.second.size());
        }
        if (lw == 0) lw = 1;
        m->lw = lw;
        m->nw = (lw + 7) & ~uint64_t(7);
        m->bcast = c->bcast;
        m->init_consts();
        for (auto& kv : c->params) m->ints[kv.first] = kv.second;

        OpCounter oc;
        oc.params = c->params;
        int64_t so = 0;
        for (const Stmt& s : c->prog.stmts) so += oc.count(s);
        c->static_ops = so;

        auto t0 = std::chrono::steady_clock::now();
        // Input loading (interpreter.inl:288-312); inputs are truncated to the
        // stream length (bsdata_reader.h:54-58, bytes_to_words): to the
        // bitlength in clean semantics, to whole 64-bit words in bsim's, whose
        // reader keeps bits above the bitlength in the last word.
        const uint64_t in_words = c->sem == BSZO_SEM_CLEAN ? m->nw : m->lw;
        for (const std::string& name : c->prog.inputs) {
            auto it = c->inputs.find(name);
            if (it != c->inputs.end()) {
                Buf& b = m->var(name);
                uint64_t k = std::min<uint64_t>(it->second.size(), in_words);
                std::memcpy(b.p, it->second.data(), k * 8);
                if (k < m->nw) std::memset(b.p + k, 0, (m->nw - k) * 8);
                m->clean(b);
            } else if (c->bcast.count(name)) {
                // broadcast: served by operand()
            } else {
                auto ai = c->input_arrays.find(name);
                if (ai != c->input_arrays.end()) {
                    for (auto& e : ai->second) {
                        auto& arr = m->arrays[name];
                        auto slot = arr.find(e.first);
                        if (slot == arr.end()) slot = arr.emplace(e.first, Buf(m->nw)).first;
                        uint64_t k = std::min<uint64_t>(e.second.size(), in_words);
                        std::memcpy(slot->second.p, e.second.data(), k * 8);
                        if (k < m->nw) std::memset(slot->second.p + k, 0, (m->nw - k) * 8);
                        m->clean(slot->second);
                    }
                } else {
                    Buf& b = m->var(name);
                    std::memset(b.p, 0, m->nw * 8);
                }
            }
        }
        for (const Stmt& s : c->prog.stmts) m->exec(s);
        c->exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->op_count = m->ops;
        c->m = std::move(m);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        c->m.reset();
        return -1;
    }
}

uint64_t bszo_n_words(const bszo_case* c) { return c->m ? c->m->lw : 0; }
int64_t bszo_op_count(const bszo_case* c) { return c->op_count; }
int64_t bszo_static_ops(const bszo_case* c) { return c->static_ops; }
double bszo_exec_ms(const bszo_case* c) { return c->exec_ms; }
const char* bszo_output_streams(const bszo_case* c) { return c->out_streams.c_str(); }
const char* bszo_output_ints(const bszo_case* c) { return c->out_ints.c_str(); }

static void copy_masked(const Machine& m, const Buf& b, uint64_t* out, uint64_t n) {
    uint64_t k = std::min<uint64_t>(n, m.lw);
    std::memcpy(out, b.p, k * 8);
    if (k < n) std::memset(out + k, 0, (n - k) * 8);
    // mask to bitlength (word_utils.h:110-115), and nothing past the logical words
    if (m.bitlength > 0 && m.lw <= n) {
        uint64_t rem = (uint64_t)m.bitlength % 64;
        if (rem) out[m.lw - 1] &= (1ull << rem) - 1;
    }
}

int bszo_get_stream(const bszo_case* c, const char* name, uint64_t* out, uint64_t n) {
    if (!c->m) return 0;
    auto it = c->m->vars.find(name);
    if (it == c->m->vars.end()) return 0;
    copy_masked(*c->m, it->second, out, n);
    return 1;
}

int64_t bszo_array_count(const bszo_case* c, const char* name) {
    if (!c->m) return 0;
    auto it = c->m->arrays.find(name);
    return it == c->m->arrays.end() ? 0 : (int64_t)it->second.size();
}

int bszo_array_indices(const bszo_case* c, const char* name, int64_t* idx, int64_t cap) {
    if (!c->m) return 0;
    auto it = c->m->arrays.find(name);
    if (it == c->m->arrays.end()) return 0;
    int64_t k = 0;
    for (auto& e : it->second) { if (k < cap) idx[k] = e.first; ++k; }
    return (int)k;
}

int bszo_get_array_elem(const bszo_case* c, const char* name, int64_t idx, uint64_t* out, uint64_t n) {
    if (!c->m) return 0;
    auto it = c->m->arrays.find(name);
    if (it == c->m->arrays.end()) return 0;
    auto e = it->second.find(idx);
    if (e == it->second.end()) return 0;
    copy_masked(*c->m, e->second, out, n);
    return 1;
}

int bszo_get_int(const bszo_case* c, const char* name, int64_t* out) {
    if (!c->m) return 0;
    auto it = c->m->ints.find(name);
    if (it == c->m->ints.end()) return 0;
    *out = it->second;
    return 1;
}

void bszo_shl(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) { shl_serial(dst, src, k, n); }
void bszo_shr(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n) { shr_serial(dst, src, k, n); }
void bszo_add(uint64_t* dst, const uint64_t* a, const uint64_t* b, uint64_t n, int variant) {
    if (variant == BSZO_BEST) add_chunked(dst, a, b, n); else add_ripple(dst, a, b, n);
}
int64_t bszo_popcount_bounded(const uint64_t* a, uint64_t n_words, int64_t bitlength) {
    (void)n_words;
    return popcount_bounded(a, bitlength, false);
}
int bszo_is_nonzero_bounded(const uint64_t* a, uint64_t n_words, int64_t bitlength) {
    (void)n_words;
    return is_nonzero_bounded(a, bitlength) ? 1 : 0;
}

int bszo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}
int bszo_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int bszo_has_avx512(void) { return cpu_has_avx512() ? 1 : 0; }

}  // extern "C"

// ref_driver.cpp — a C entry layer over the reference's own CPU interpreter.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ and
// linked against the reference's sources where they lie under
// /root/reference/simulator/csim/src (lexer.cpp, parser.cpp, ast_liveness.cpp,
// bitstream_ops.cpp and the header-only Interpreter<CpuBackend>).  Nothing of
// the reference is copied into this repository.  The reference's bsim driver
// (main.cpp) is not built: it needs nlohmann/json, which is not available
// offline; this file stands in for its run_case (main.cpp:21-161) with the
// same n_words rule and backend selection, and takes its inputs as arrays.
//
// Used by tests to pin the oracle (oracle/bsz_oracle.cpp) against the real
// reference on programs and inputs the committed fixtures do not cover.

#include "ast_liveness.h"
#include "cpu_backend.h"
#include "interpreter.h"
#include "parser.h"
#include "word_utils.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

struct bsref_case {
    bs::Program prog;
    int64_t bitlength = 0;
    std::unordered_map<std::string, int64_t> params;
    std::unordered_map<std::string, std::vector<uint64_t>> inputs;
    std::map<std::string, std::map<int, std::vector<uint64_t>>> input_arrays;
    std::map<std::string, std::vector<bool>> bcast;
    bs::Interpreter<bs::CpuBackend>::Result res;
    bool have_res = false;
    size_t n_words = 1;
    double exec_ms = 0;
};

static void set_err(char* err, size_t len, const char* msg) {
    if (err && len) std::snprintf(err, len, "%s", msg);
}

extern "C" {

bsref_case* bsref_case_new(const char* src, char* err, size_t errlen) {
    try {
        auto* c = new bsref_case;
        c->prog = bs::parse(src);
        return c;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

void bsref_case_free(bsref_case* c) { delete c; }
int bsref_set_bitlength(bsref_case* c, int64_t bl) { c->bitlength = bl; return 0; }
int bsref_set_param(bsref_case* c, const char* n, int64_t v) { c->params[n] = v; return 0; }
int bsref_set_input(bsref_case* c, const char* n, const uint64_t* w, uint64_t k) {
    c->inputs[n].assign(w, w + k); return 0;
}
int bsref_set_array_elem(bsref_case* c, const char* n, int64_t idx, const uint64_t* w, uint64_t k) {
    c->input_arrays[n][(int)idx].assign(w, w + k); return 0;
}
int bsref_set_broadcast(bsref_case* c, const char* n, const uint8_t* bits, int64_t count) {
    std::vector<bool> v(count);
    for (int64_t i = 0; i < count; ++i) v[i] = (bits[i / 8] >> (i % 8)) & 1;
    c->bcast[n] = std::move(v);
    return 0;
}

// variant: 0 scalar, 1 simd, 2 simd_omp (main.cpp:50-57); add: 0 ripple, 1 kogge-stone.
int bsref_run(bsref_case* c, int variant, int add_mode, int reuse_mem, char* err, size_t errlen) {
    try {
        size_t n_words = 0;
        if (c->bitlength > 0) {
            n_words = (size_t)((c->bitlength + 63) / 64);
        } else if (!c->inputs.empty()) {
            size_t mx = 1;
            for (auto& kv : c->inputs) mx = std::max(mx, kv.second.size());
            n_words = mx;
        }
        if (n_words == 0) n_words = 1;
        // bsdata_reader.h sizes every decoded stream to n_words words.
        auto inputs = c->inputs;
        for (auto& kv : inputs) kv.second.resize(n_words, 0);
        auto arrays = c->input_arrays;
        for (auto& a : arrays) for (auto& e : a.second) e.second.resize(n_words, 0);
        auto v = variant == 2 ? bs::CpuBackend::Variant::SIMD_OMP
               : variant == 1 ? bs::CpuBackend::Variant::SIMD
                              : bs::CpuBackend::Variant::SCALAR;
        auto am = add_mode == 1 ? bs::CpuBackend::AddMode::KOGGE_STONE : bs::CpuBackend::AddMode::RIPPLE;
        bs::CpuBackend backend(n_words, (size_t)c->bitlength, v, am);
        bs::Interpreter<bs::CpuBackend> interp(std::move(backend), reuse_mem != 0);
        auto t0 = std::chrono::steady_clock::now();
        c->res = interp.run(c->prog, inputs, c->params, arrays, c->bcast);
        c->exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->have_res = true;
        c->n_words = n_words;
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        c->have_res = false;
        return -1;
    }
}

int64_t bsref_op_count(const bsref_case* c) { return c->res.op_count; }

// The reference's static op count (count_ops_ast, ast_liveness.cpp:723-727)
// of the parsed program under the case's params; needs no run.
int64_t bsref_static_ops(const bsref_case* c) { return count_ops_ast(c->prog, c->params); }
double bsref_exec_ms(const bsref_case* c) { return c->exec_ms; }

static void out_masked(const bsref_case* c, const std::vector<uint64_t>& v, uint64_t* out, uint64_t n) {
    std::vector<uint64_t> m(v.begin(), v.begin() + std::min(v.size(), c->n_words));
    bsdata::mask_words(m.data(), m.size(), c->bitlength);
    for (uint64_t i = 0; i < n; ++i) out[i] = i < m.size() ? m[i] : 0;
}

int bsref_get_stream(const bsref_case* c, const char* name, uint64_t* out, uint64_t n) {
    if (!c->have_res) return 0;
    auto it = c->res.streams.find(name);
    if (it == c->res.streams.end()) return 0;
    out_masked(c, it->second, out, n);
    return 1;
}

int64_t bsref_array_count(const bsref_case* c, const char* name) {
    if (!c->have_res) return 0;
    auto it = c->res.arrays.find(name);
    return it == c->res.arrays.end() ? 0 : (int64_t)it->second.size();
}

int bsref_array_indices(const bsref_case* c, const char* name, int64_t* idx, int64_t cap) {
    if (!c->have_res) return 0;
    auto it = c->res.arrays.find(name);
    if (it == c->res.arrays.end()) return 0;
    int64_t k = 0;
    for (auto& e : it->second) { if (k < cap) idx[k] = e.first; ++k; }
    return (int)k;
}

int bsref_get_array_elem(const bsref_case* c, const char* name, int64_t idx, uint64_t* out, uint64_t n) {
    if (!c->have_res) return 0;
    auto it = c->res.arrays.find(name);
    if (it == c->res.arrays.end()) return 0;
    auto e = it->second.find((int)idx);
    if (e == it->second.end()) return 0;
    out_masked(c, e->second, out, n);
    return 1;
}

int bsref_get_int(const bsref_case* c, const char* name, int64_t* out) {
    if (!c->have_res) return 0;
    auto it = c->res.ints.find(name);
    if (it == c->res.ints.end()) return 0;
    *out = it->second;
    return 1;
}

}  // extern "C"

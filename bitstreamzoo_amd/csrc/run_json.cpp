// run_json.cpp — see run_json.h.
#include "run_json.h"

#include <chrono>
#include <cstdio>
#include <cstring>

namespace bsz {

RunOutcome run_case_iterations(const Program& prog, const CaseData& c, Engine& eng, double parse_ms) {
    RunOutcome o;
    for (Sym s : prog.outputs) o.out_streams.push_back(prog.syms.names[s]);
    for (Sym s : prog.out_ints) o.out_ints.push_back(prog.syms.names[s]);
    o.bitlength = c.bitlength;
    uint64_t mw = 1;
    for (auto& kv : c.inputs) mw = std::max<uint64_t>(mw, kv.second.size());
    o.n_words = Geom::make(c.bitlength, mw).lw;
    o.static_ops = count_static_ops(prog, c.params);
    o.parse_ms = parse_ms;
    o.iterations = std::max(1, c.iterations);

    CaseData cur = c;
    hipEvent_t e0, e1;
    BSZ_HIP(hipEventCreate(&e0));
    BSZ_HIP(hipEventCreate(&e1));
    const int64_t launches0 = eng.n_kernel_launches();
    for (int it = 0; it < o.iterations; ++it) {
        auto t0 = std::chrono::steady_clock::now();
        BSZ_HIP(hipEventRecord(e0, eng.stream()));
        o.res = execute_program(prog, cur, eng);
        BSZ_HIP(hipEventRecord(e1, eng.stream()));
        BSZ_HIP(hipEventSynchronize(e1));
        auto t1 = std::chrono::steady_clock::now();
        float ms = 0;
        BSZ_HIP(hipEventElapsedTime(&ms, e0, e1));
        o.kernel_ms += ms;
        o.exec_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        for (auto& fb : c.feedback) {
            auto sit = o.res.streams.find(fb.first);
            if (sit != o.res.streams.end()) cur.inputs[fb.second] = sit->second;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    o.dynamic_ops = o.res.op_count * o.iterations;
    o.n_kernel_launches = eng.n_kernel_launches() - launches0;
    o.n_mallocs = eng.n_mallocs();
    o.peak_bytes = eng.peak_bytes();
    o.engine_stats = eng.stats_json();
    return o;
}

static std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.6f", v);
    return b;
}

namespace {

std::string bin_ref(std::vector<uint8_t>* blob, const std::vector<uint64_t>& w) {
    const size_t off = blob->size(), nb = w.size() * 8;
    blob->resize(off + nb);
    if (nb) std::memcpy(blob->data() + off, w.data(), nb);
    return "{\"_bin\":[" + std::to_string(off) + "," + std::to_string(nb) + "]}";
}

}  // namespace

std::string result_json(const RunOutcome& o, std::vector<uint8_t>* blob) {
    std::string s = "{\"dynamic_ops\":" + std::to_string(o.dynamic_ops);
    s += ",\"engine\":" + o.engine_stats;
    s += ",\"exec_ms\":" + num(o.exec_ms);
    double gops = o.exec_ms > 0 ? (double)o.dynamic_ops * (double)o.n_words * 2.0 / (o.exec_ms * 1e-3) / 1e9 : 0.0;
    s += ",\"int32_gops\":" + num(gops);
    const uint64_t stream_bytes = o.n_words * 8;
    const int64_t max_live = stream_bytes ? (int64_t)(o.peak_bytes / (((o.n_words + 63) / 64) * 512)) : 0;
    s += ",\"max_live\":" + std::to_string(max_live);
    s += ",\"n_kernel_launches\":" + std::to_string(o.n_kernel_launches);
    s += ",\"n_mallocs\":" + std::to_string(o.n_mallocs);
    s += ",\"n_variables\":" + std::to_string(o.res.n_variables);
    s += ",\"n_words\":" + std::to_string(o.n_words);
    s += ",\"op_count\":" + std::to_string(o.dynamic_ops);
    s += ",\"outputs\":{";
    bool first = true;
    auto key = [&](const std::string& k) { s += (first ? "" : ",") + json_escape(k) + ":"; first = false; };
    for (const std::string& n : o.out_streams) {
        auto it = o.res.streams.find(n);
        if (it != o.res.streams.end()) {
            key(n);
            if (blob) s += bin_ref(blob, it->second);
            else s += "\"" + words_to_hex(it->second.data(), it->second.size()) + "\"";
        }
        auto ai = o.res.arrays.find(n);
        if (ai != o.res.arrays.end()) {
            key(n);
            s += "{";
            bool f2 = true;
            for (auto& e : ai->second) {
                s += (f2 ? "\"" : ",\"") + std::to_string(e.first) + "\":";
                if (blob) s += bin_ref(blob, e.second);
                else s += "\"" + words_to_hex(e.second.data(), e.second.size()) + "\"";
                f2 = false;
            }
            s += "}";
        }
    }
    for (const std::string& n : o.out_ints) {
        auto it = o.res.ints.find(n);
        if (it != o.res.ints.end()) { key(n); s += std::to_string(it->second); }
    }
    s += "}";
    s += ",\"peak_bytes\":" + std::to_string((uint64_t)o.res.n_variables * stream_bytes);
    s += ",\"reuse_bytes\":" + std::to_string((uint64_t)o.peak_bytes);
    s += ",\"static_ops\":" + std::to_string(o.static_ops);
    s += ",\"timing\":{\"alloc_ms\":0.0,\"compile_ms\":" + num(o.compile_ms) + ",\"exec_ms\":" + num(o.exec_ms) +
         ",\"kernel_ms\":" + num(o.kernel_ms) + ",\"liveness_ms\":0.0,\"parse_ms\":" + num(o.parse_ms) +
         ",\"runs\":1,\"warmup_ms\":0.0}";
    s += "}";
    return s;
}

}  // namespace bsz

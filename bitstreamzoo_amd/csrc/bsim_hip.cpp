// bsim_hip.cpp — command-line driver with the bsim_cuda contract
// (simulator/csim/src/main.cu:209-331):
//
//   bsim_hip program.bs [--trace] [--word-bits=32|64] [--reuse-mem] [--warmup]
//                       [--input file.bsdata ...] [--dump-input]
//                       [--executor=fused|perop] [--device=N]
//                       [--output-bsdata out.bsdata] [--padding=clean|bsim]
//
// Without --input one JSON case is read from stdin (keys inputs, params,
// input_arrays, bitlength / n_vectors, iterations, feedback).  stdout carries
// one JSON object per case, or an array with a "name" per case when there are
// several.  Exit 1 with "Parse error: ...", "Error reading bsdata: ..." or
// "Error: ..." on stderr.  --reuse-mem and --trace are accepted for
// compatibility (memory is always planned; there is no trace output).
// --padding=bsim gives bits at or above the bitlength the values bsim_cuda
// gives them with the chosen --word-bits (its streams are padded to a
// multiple of 8 words; ~, ONES, + and << leave padding bits set and >> can
// bring them back); the default, clean, keeps them zero as DSL.md specifies.
// --output-bsdata writes the output streams in binary to a .bsdata-style
// file (the JSON printed on stdout, a NUL, then the words) and prints that
// JSON with {"_bin": [offset, bytes]} in place of each hex string: the hex
// of large outputs (2.2 GB for 33 streams of 2^28 bits) is never built.
#include "bsdata_io.h"
#include "executor.h"
#include "fused.h"
#include "frontend.h"
#include "run_json.h"

#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

using namespace bsz;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "Usage: bsim_hip program.bs [--trace] [--word-bits=32|64] [--reuse-mem] "
                     "[--warmup] [--input file.bsdata ...] [--dump-input] [--executor=fused|perop] "
                     "[--device=N]\n";
        return 1;
    }
    std::string bs_path = argv[1];
    bool dump_input = false, warmup = false;
    int word_bits = 64, device = 0;
    std::string executor = "fused";
    std::vector<std::string> inputs;
    std::string out_bsdata, padding = "clean";
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--trace" || a == "--reuse-mem") continue;
        if (a == "--output-bsdata" && i + 1 < argc) { out_bsdata = argv[++i]; continue; }
        if (a.rfind("--output-bsdata=", 0) == 0) { out_bsdata = a.substr(16); continue; }
        if (a.rfind("--padding=", 0) == 0) { padding = a.substr(10); continue; }
        if (a == "--warmup") warmup = true;
        else if (a.rfind("--word-bits=", 0) == 0) word_bits = std::atoi(a.c_str() + 12);
        else if (a == "--input" && i + 1 < argc) inputs.push_back(argv[++i]);
        else if (a == "--dump-input") dump_input = true;
        else if (a.rfind("--executor=", 0) == 0) executor = a.substr(11);
        else if (a.rfind("--device=", 0) == 0) device = std::atoi(a.c_str() + 9);
    }
    if (word_bits != 32 && word_bits != 64) {
        std::cerr << "Error: --word-bits must be 32 or 64\n";
        return 1;
    }
    if (padding != "clean" && padding != "bsim") {
        std::cerr << "Error: --padding must be clean or bsim\n";
        return 1;
    }
    if (executor != "fused" && executor != "perop") {
        std::cerr << "Error: --executor must be fused or perop\n";
        return 1;
    }
    std::ifstream fs(bs_path);
    if (!fs) {
        std::cerr << "Error: cannot open " << bs_path << "\n";
        return 1;
    }
    std::stringstream ss;
    ss << fs.rdbuf();
    const std::string source = ss.str();

    std::vector<CaseData> cases;
    if (!inputs.empty()) {
        for (const std::string& f : inputs) {
            try {
                auto v = read_bsdata_file(f);
                for (auto& c : v) cases.push_back(std::move(c));
            } catch (const std::exception& e) {
                std::cerr << "Error reading bsdata: " << e.what() << "\n";
                return 1;
            }
        }
    } else {
        try {
            std::stringstream in;
            in << std::cin.rdbuf();
            cases.push_back(case_from_json(json_parse(in.str())));
        } catch (const std::exception& e) {
            std::cerr << "JSON parse error: " << e.what() << "\n";
            return 1;
        }
    }

    if (dump_input) {
        if (cases.size() == 1) {
            std::cout << dump_case_json(cases[0], 2) << "\n";
        } else {
            std::cout << "[";
            for (size_t i = 0; i < cases.size(); ++i) std::cout << (i ? ",\n" : "\n") << dump_case_json(cases[i], 2);
            std::cout << "\n]\n";
        }
        return 0;
    }

    if (padding == "bsim")
        for (CaseData& c : cases) c.bsim_padding = word_bits;

    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<Program> prog;
    try {
        prog = parse_program(source);
    } catch (const std::exception& e) {
        std::cerr << "Parse error: " << e.what() << "\n";
        return 1;
    }
    double parse_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    try {
        BSZ_HIP(hipSetDevice(device));
        hipStream_t s;
        BSZ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::unique_ptr<Engine> eng =
            executor == "perop" ? make_perop_engine(s) : make_fused_engine(s, FusedOptions::from_env());
        const bool batch = cases.size() > 1;
        std::string out = batch ? "[" : "";
        std::vector<uint8_t> blob;
        for (size_t i = 0; i < cases.size(); ++i) {
            if (warmup) {
                CaseData w = cases[i];
                w.iterations = 1;
                (void)run_case_iterations(*prog, w, *eng, parse_ms);
            }
            RunOutcome o = run_case_iterations(*prog, cases[i], *eng, parse_ms);
            if (word_bits == 32) o.n_words = o.bitlength > 0 ? (uint64_t)(o.bitlength + 31) / 32 : o.n_words * 2;
            std::string j = result_json(o, out_bsdata.empty() ? nullptr : &blob);
            if (batch) {
                j.pop_back();
                j += ",\"name\":" + json_escape(cases[i].name) + "}";
                out += (i ? "," : "") + j;
            } else {
                out += j;
            }
        }
        if (batch) out += "]";
        if (!out_bsdata.empty()) {
            std::ofstream of(out_bsdata, std::ios::binary);
            if (!of) throw std::runtime_error("cannot write " + out_bsdata);
            of.write(out.data(), (std::streamsize)out.size());
            of.put('\0');
            of.write(reinterpret_cast<const char*>(blob.data()), (std::streamsize)blob.size());
            if (!of) throw std::runtime_error("write failed: " + out_bsdata);
        }
        std::cout << out << "\n";
        eng.reset();
        (void)hipStreamDestroy(s);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}

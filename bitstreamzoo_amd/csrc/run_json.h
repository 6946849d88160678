// run_json.h — one case end to end (iterations + feedback, main.cu:84-140)
// and the bsim_cuda per-case JSON object (main.cu:146-205).
#pragma once

#include "executor.h"

#include <string>
#include <vector>

namespace bsz {

struct RunOutcome {
    RunResult res;                 // outputs of the last iteration
    std::vector<std::string> out_streams, out_ints;  // declaration order
    int64_t bitlength = 0;
    uint64_t n_words = 1;          // ceil(bitlength/64) (main.cu:40-49)
    int64_t static_ops = 0, dynamic_ops = 0;
    int iterations = 1;
    double parse_ms = 0, exec_ms = 0, kernel_ms = 0, compile_ms = 0;
    int64_t n_kernel_launches = 0, n_mallocs = 0;
    size_t peak_bytes = 0;
    std::string engine_stats = "{}";
};

RunOutcome run_case_iterations(const Program& prog, const CaseData& c, Engine& eng, double parse_ms);
// With `blob`, every output stream's words (little-endian, masked to the
// bitlength) are appended to it and the JSON holds {"_bin": [offset, bytes]}
// in place of the hex string (the .bsdata encoding of a stream).
std::string result_json(const RunOutcome& o, std::vector<uint8_t>* blob = nullptr);

}  // namespace bsz

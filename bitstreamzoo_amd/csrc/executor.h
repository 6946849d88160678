// executor.h — the control plane: a walk over the program that hands every
// stream operation to an Engine.
//
// Replaces bs::Interpreter<Backend> (interpreter.h:27-123, interpreter.inl:17-364)
// with the same observable semantics: variables referenced before assignment
// read as ZERO and then exist; arrays are sparse (ArrayDecl zeroes/creates
// 0..N-1, unwritten elements read as ZERO); broadcast arrays serve ONES/ZERO
// by index; `for` bounds and array indices are int32; `while` stops after
// 10000 iterations with a RuntimeError; ops are counted as DSL.md:277-285.
//
// Stream values are immutable and reference counted, so `x = y` shares and no
// operation ever writes into one of its operands.  Two engines implement it:
//   PerOpEngine  — one HIP kernel per DSL op (the plain comparator);
//   FusedEngine  — records ops into a segment DAG and, at each point where
//                  the host must observe a value (if/while conditions, int
//                  expressions that read a popcount, the end of the program),
//                  partitions the segment into fused kernels (fused.h).
// popcount results stay on the device until an int expression or an output
// needs them.
#pragma once

#include "bsdata_io.h"
#include "device.h"
#include "frontend.h"

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace bsz {

// A stream value: either materialized in a device buffer, or (FusedEngine)
// a node of the segment being recorded.
struct StreamVal {
    DevBufP buf;
    int32_t node = -1;      // node id in the engine's open segment, -1 if none
    uint64_t segment = 0;   // id of that segment
    int32_t source = -1;    // recorded plan input (no data yet), -1 if none
    int8_t konst = -1;      // 0 ZERO, 1 ONES, -1 not a constant
    virtual ~StreamVal() = default;
};
using SV = std::shared_ptr<StreamVal>;

// A popcount result that may still be on the device.
struct PopCell {
    bool ready = false;
    int64_t value = 0;
    virtual ~PopCell() = default;
};
using PopRef = std::shared_ptr<PopCell>;

enum class BinOp : int { And = 0, Or = 1, Xor = 2, Add = 3 };

class Engine {
public:
    virtual ~Engine() = default;
    virtual void begin(const Geom& g) = 0;
    virtual SV constant(bool ones) = 0;
    virtual SV input(const uint64_t* host, uint64_t n) = 0;
    virtual SV op_not(const SV& a) = 0;
    virtual SV op_bin(BinOp op, const SV& a, const SV& b) = 0;
    virtual SV op_shift(bool left, const SV& a, uint64_t k) = 0;
    virtual PopRef popcount(const SV& a) = 0;
    virtual int64_t resolve(const PopRef& p) = 0;
    virtual bool is_nonzero(const SV& a) = 0;
    // Copy the lw logical words of `a` to host (masked to the bitlength).
    virtual void download(const SV& a, uint64_t* host) = 0;
    // Make everything recorded so far execute; called before outputs are read.
    virtual void flush() = 0;
    virtual void sync() = 0;
    virtual int64_t n_kernel_launches() const = 0;
    virtual int64_t n_mallocs() const = 0;
    virtual size_t peak_bytes() const = 0;
    virtual hipStream_t stream() const = 0;
    virtual std::string stats_json() const { return "{}"; }
    // Device-side `if` (if-conversion).  An engine that supports it keeps an
    // `if` body on the device: predicate(c) is "c has a bit set within the
    // bitlength", left on the device as the popcount of c (true when it is
    // not zero), and select(p, t, f) takes t where p holds, else f, word by
    // word.  The host reads predicates only when it needs them.
    virtual bool supports_predicates() const { return false; }
    virtual PopRef predicate(const SV& /*cond*/) { throw RunError("internal: this engine has no device predicates"); }
    virtual SV select(const PopRef& /*pred*/, const SV& /*t*/, const SV& /*f*/) {
        throw RunError("internal: this engine has no device predicates");
    }
    // Called before input() with the DSL name (and array index, or -1).
    virtual void note_input(const std::string& /*name*/, int64_t /*idx*/) {}
};

struct RunResult {
    std::map<std::string, std::vector<uint64_t>> streams;
    std::map<std::string, std::map<int64_t, std::vector<uint64_t>>> arrays;
    std::map<std::string, int64_t> ints;
    int64_t op_count = 0;
    std::map<std::string, int64_t> op_mix;
    int64_t n_variables = 0;
    double exec_ms = 0, kernel_ms = 0;
};

// Walks `prog` over `c` on `eng` (one iteration; feedback is the caller's).
RunResult execute_program(const Program& prog, const CaseData& c, Engine& eng);

// The values a walk leaves in the program's outputs, without reading them
// back (used to record a reusable plan).
struct RecordedOutputs {
    std::vector<std::pair<std::string, SV>> streams;
    std::vector<std::tuple<std::string, int64_t, SV>> array_elems;
    struct IntOut { std::string name; PopRef pop; int64_t value = 0; };
    std::vector<IntOut> ints;
    int64_t op_count = 0;
};
RecordedOutputs record_program(const Program& prog, const CaseData& c, Engine& eng);

std::unique_ptr<Engine> make_perop_engine(hipStream_t s);

}  // namespace bsz

// fused_graph.h — segment DAG, fusion partitioner, stream memory planner and
// kernel code generation for the fused executor.
//
// Execution model of a fused kernel (one per partition):
//   * a stream is cut into rows of 64 words; one wavefront owns a chunk of
//     `chunk_rows` consecutive rows and walks them in order, one 64-bit word
//     per lane per row, every live value in registers;
//   * an advance by k < 64 of a value computed in the same kernel takes the
//     lower neighbour word with a DPP wave_shr:1 lane move, lane 0 taking the
//     previous row's lane 63 word, kept in a scalar register;
//   * stream `+` forms the 64 lane sums, ballots their generate / propagate
//     predicates and resolves the row's carries with scalar arithmetic on the
//     two masks (C = (G + (G|P) + cin) ^ G ^ (G|P)); the carry out of the row
//     is the next row's carry in;
//   * those scalars (shift neighbours and carries) are the kernel's state.  A
//     chunk other than the first starts `warmup_rows` rows early with a zero
//     state (speculation).  During the warm-up the wave also tracks, per
//     value, the lowest lane from which it is certainly exact, and per state
//     word whether it is: a wrong start state damages only the low lanes of a
//     row, an advance's state is its operand's lane 63, and a carry is
//     settled by any lane that generates or kills it above the damage.  A
//     chunk whose warm-up proves every state exact (or starts at the
//     instance's first row) is final.  An unproven chunk runs its rows from
//     the speculative state, then looks back: it waits for its
//     predecessor's end state of this run and, when that differs from the
//     state it assumed, runs its rows again from it.  Every chunk publishes
//     its final end state as 64-bit words that each carry one 32-bit half
//     under the run's tag, so a reader checks the tags and needs no fence.
//     Wave w of the grid owns chunk w; workgroups are dispatched in order,
//     so the lowest unfinished chunk never waits for an unfinished one, and
//     each wait is bounded (a wait that reaches the bound sets the plan's
//     error word).  One launch per kernel; the result is exact for every
//     input, and speculation only decides how much work is redone;
//   * an advance of a value computed by an earlier kernel (any k) and `>>` are
//     direct loads at the shifted word offset, needing no state;
//   * popcounts accumulate per lane (v_bcnt), are reduced per wave and added
//     with one atomic per chunk into one of 2^spread_log slots (chunk mod
//     2^spread_log, up to kMaxPopSpread slots, about one per 4 chunks of the
//     longest instance), so no address serializes more than a few atomics;
//     the host sums the slots.  Slots live in two banks: a run adds into
//     bank (run & 1), where `run` counts the plan's completed runs over the
//     chunk, and clears the other for the next run; the plan's last kernel
//     advances each chunk's run number, so no separate clearing or epoch
//     launch is needed.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bsz {

// Sel: a where predicate `pred` holds, else b, word by word.  A predicate is
// one of the graph's popcounts read as "not zero"; it is produced by an
// earlier kernel (its slots are final once that kernel's launch ends) and
// read by every wave of the consuming kernel before its first row.
enum class NOp : uint8_t { Input, Zero, Ones, Not, And, Or, Xor, Add, Shl, Shr, Sel };

struct GNode {
    NOp op;
    int32_t a = -1, b = -1;
    uint64_t k = 0;      // shift amount
    int32_t src = -1;    // Input: index of the external source
    int32_t pred = -1;   // Sel: index into SegGraph::pops of its predicate
};

struct SegGraph {
    std::vector<GNode> nodes;        // topological (creation) order
    std::vector<int32_t> outputs;    // nodes to materialize
    std::vector<int32_t> pops;       // nodes to popcount
    int32_t n_sources = 0;
    std::string key() const;         // structural identity (no buffers, no sizes)
};

struct PlanOptions {
    int chunk_rows = 16;
    int warmup_rows = 1;
    int waves_per_block = 4;
    int rows_per_step = 8;           // upper bound on rows whose loads are issued together
    int words_per_lane = 1;          // 2: lanes own word pairs in kernels without cross-lane work
    bool store_order = true;         // scheduling barrier after each output's store
    int nontemporal = 0;             // 1: nontemporal stream stores, 2: loads too
    bool interleave = false;         // rows cc, cc+n, ... per chunk in kernels without state
    int max_ops = 4000;
    int max_live = 96;
    int occupancy = 0;               // waves per SIMD the register budget allows; 0 = from an estimate
    bool persistent = false;         // grid of resident waves, each striding over the chunks
    bool agprs = false;              // occupancy 1 only: values may also live in the 256 AGPRs
    bool lazy_loads = false;         // load each input word at its first use in a row, not at the row's start
    bool join_warmup = false;        // a stateful kernel whose step spans its chunk loads warm-up and chunk together
    bool bitop3 = true;              // boolean nodes fused into three-input functions (v_bitop3_b32)
    int order = 0;                   // 0: partition in recorded order; 1: pressure-aware list schedule
    std::string key() const;
};

// Where a kernel reads an input from.
struct KInput {
    enum Kind : uint8_t { Source, Slot, SegOut } kind = Source;
    int32_t id = -1;                 // source index, intermediate slot, or graph node of a segment output
};

struct KernelSpec {
    std::string name;
    std::vector<int32_t> nodes;      // computed nodes, in order
    std::vector<KInput> inputs;
    std::vector<int32_t> input_node; // graph node id of each input (Input node or producer)
    std::vector<int32_t> outputs;    // nodes this kernel stores
    std::vector<int32_t> out_slot;   // per output: -1 segment output, else intermediate slot
    std::vector<int32_t> pops;       // index into SegGraph::pops accumulated here
    std::vector<int32_t> preds;      // index into SegGraph::pops read as predicates (after the inputs in the table)
    int n_state = 0;                 // 64-bit state words
    int peak_live = 0;               // most in-kernel values live at once (partitioner's count)
    int rows_per_step = 1;           // as generated (capped by register budget)
    int words_per_lane = 1;          // as generated
    bool interleave = false;         // as generated
    int chunk_rows = 16;             // as generated
    int occupancy = 4;               // waves per SIMD the register budget was set for
    int launch_bounds = 256;         // largest workgroup the kernel accepts
    std::string source;              // generated HIP
    std::string entry;               // its entry point: bsz_fk_<hash of the unit>
    std::string unit;                // self-contained compilation unit: prelude + source
};

struct PlanSpec {
    SegGraph g;
    PlanOptions opt;
    std::vector<KernelSpec> kernels;
    int32_t n_slots = 0;             // intermediate buffers (after interval reuse)
    std::vector<int32_t> out_node_kernel;  // per g.outputs entry: kernel storing it, -1 if it is a source
    std::string module_source;       // all kernels, for inspection (each compiles as its own unit)
};

// Dead-code elimination, partitioning, slot assignment and code generation.
// emit=false stops after the kernel shapes (no HIP source): enough to
// describe a candidate schedule.
PlanSpec build_plan(const SegGraph& g, const PlanOptions& opt, bool emit = true);

// Regenerates kernel k of P with a smaller register footprint and the same
// nodes: half the rows per step, then 8-byte lanes, then (automatic occupancy
// only) one wave fewer per SIMD.  False when nothing is left to shrink.
bool shrink_kernel(PlanSpec& P, size_t k);

// Argument block every generated kernel takes by value (layout shared with the
// generated source, see kKernelPrelude in fused_codegen.cpp).
struct InstDesc {
    uint64_t rows;
    uint64_t lw;
    uint64_t lastmask;
    uint64_t opw;        // words an op result may occupy (Geom::opw)
    uint64_t opmask;     // mask of word opw-1
    uint32_t chunk_begin;
    uint32_t n_chunks;
};

// Most popcount accumulator slots per (instance, popcount).
constexpr int kMaxPopSpreadLog = 8;
// Polls of a look-back wait before it gives up and flags the plan (each poll
// is a memory round trip and a short sleep: seconds in all).
constexpr uint32_t kSpinLimit = 1u << 22;

// ctl block of a plan: [0] error word (bit 0: a look-back wait reached its
// bound); per kernel k, at 4 + 8k, two banks of four counters, bank (run & 1)
// counting the run: [0] chunks run again from their predecessor's end state,
// [1] chunks whose warm-up did not prove their start state, [2] look-back
// waits that reached the bound.
struct KArgs {
    const uint64_t* const* in;        // [inst][n_in]
    uint64_t* const* out;             // [inst][n_out]
    unsigned long long* pop_acc;      // bank 0: [inst][n_pop_k][2^spread_log]; bank 1 at + pop_stride
    const InstDesc* inst;
    const uint32_t* chunk_inst;       // [n_chunks] instance of each chunk
    unsigned long long* st_t;         // [n_chunks][2 * n_state] tagged end-state halves
    unsigned* runs;                   // [n_chunks] completed runs of the plan over each chunk
    unsigned* kstat;                  // this kernel's counters in the ctl block
    unsigned* err;                    // the plan's error word
    uint64_t pop_stride;              // words between the two banks
    uint32_t n_chunks;
    uint32_t n_inst;
    uint32_t last;                    // 1: the plan's last kernel (advances the run numbers)
    uint32_t spread_log;              // log2 of the popcount slots per (instance, popcount)
    uint32_t spin_limit;              // polls of a look-back wait before it gives up
    uint32_t grid_waves;              // waves in the grid (a persistent kernel strides by it)
};

}  // namespace bsz

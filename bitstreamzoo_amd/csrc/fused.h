// fused.h — the fused executor: segment DAG, partitioner, stream memory
// planner, HIP code generation, hipRTC compilation and the launch scheduler.
#pragma once

#include "executor.h"

#include <string>

namespace bsz {

struct Json;

struct FusedOptions {
    int chunk_rows = 16;        // 64-word rows per wave-chunk
    int warmup_rows = 1;        // speculative warm-up rows before a chunk
    int waves_per_block = 4;    // independent wavefronts per workgroup
    int rows_per_step = 8;      // upper bound on rows whose loads are issued together
    int words_per_lane = 1;     // 2: 16-byte loads/stores in kernels without cross-lane work
    int store_order = 1;        // 1: each output's store stays where its value is computed
    int nontemporal = 0;        // 1: nontemporal stores of stream words, 2: loads too
    int interleave = 0;         // 1: kernels without state take every n-th row instead of a contiguous chunk
    int max_ops = 4000;         // fusion depth: DAG nodes per kernel
    int max_live = 96;          // live 64-bit values per lane per kernel
    int occupancy = 0;          // waves per SIMD the register budget allows (512/occupancy VGPRs); 0 = auto
    int persistent = 0;         // 1: each kernel's grid holds only resident waves, each striding over the chunks
    int agprs = 0;              // 1 (with occupancy 1): kernels may keep values in AGPRs, 512 registers per wave
    int lazy_loads = 0;         // 1: each input word is loaded at its first use in a row (partitions count it
                                //   as a value from there to its last use) instead of at the row's start
    int join_warmup = 0;        // 1: a stateful kernel whose step spans its chunk loads the warm-up rows and the
                                //   chunk in one step; 0: the warm-up rows run in a loop of their own first
    int bitop3 = 1;             // 1: chains of NOT/AND/OR/XOR fused into three-input functions (v_bitop3_b32)
    int order = 0;              // partition order: 0 the recorded order, 1 a list schedule that keeps few
                                //   values live (programs without device-side `if` selections)
    std::string cache_dir;      // code-object cache ("" = $BSZ_CACHE_DIR or ~/.cache/bsz_hip)
    int compile = 1;            // compile-only: 0 partitions and generates without running hipRTC
    int compile_shard = 0;      // compile-only: compile the kernels k with k % compile_shards == compile_shard
    int compile_shards = 1;     //   (several processes fill the cache of one plan in parallel)
    int device_if = 0;          // compile-only: record `if` bodies as device selections (inspection of the plan
                                //   bsz_run would build; a recorded plan with selections is never run)
    int64_t spin_limit = -1;    // polls of a look-back wait before it gives up (-1: the library's bound);
                                //   a run whose wait gave up reports an error instead of outputs
    bool verbose = false;
    static FusedOptions from_env();
    // Env defaults overridden by the keys of a flat JSON object
    // ({"chunk_rows":16,"warmup_rows":1,...}); "" leaves the defaults.
    static FusedOptions from_json(const std::string& json);
    static FusedOptions from_json_obj(const Json& j);
};

std::unique_ptr<Engine> make_fused_engine(hipStream_t s, const FusedOptions& opt);

// A straight-line program recorded once into a fused plan and executed over
// any number of instances (each its own bitlength and buffers) in one launch
// per kernel: the tiles of every instance form one device-side work queue.
class FusedPlan {
public:
    // Walks `prog` with `shape`'s params (inputs need only be present, any
    // length).  Fails if the program needs host decisions (if/while, an int
    // expression reading a popcount).  compile=true runs hipRTC now (no GPU
    // needed); the module is loaded on the device by set_instances().
    static std::unique_ptr<FusedPlan> record(hipStream_t s, const Program& prog, const CaseData& shape,
                                             const FusedOptions& opt, bool compile);
    ~FusedPlan();
    void set_instances(const std::vector<int64_t>& bitlengths);
    int n_instances() const;
    // `dev` is a row-padded stream buffer of the instance's bitlength.
    void bind_input(int inst, const std::string& name, int64_t idx, const uint64_t* dev);
    uint64_t* output(int inst, const std::string& name, int64_t idx);
    // 1: fused kernels (default); 0: one kernel per DSL op over the same
    // recorded DAG, with the same resident inputs (the plain comparator).
    void set_executor(int executor);
    void prepare_now();  // upload binding tables (host sync) so run() only launches
    void run();        // asynchronous on the plan's stream
    void capture();    // record run() into a hipGraph
    void replay();     // launch the captured graph
    int64_t read_int(int inst, const std::string& name, bool refresh);
    uint32_t mismatches(int kernel);
    uint32_t unproven(int kernel);   // chunks whose warm-up left a state unproven in the last run
    std::string info_json() const;
    // Kernel shapes of the recorded program under each schedule of a JSON
    // array of option objects (partitioning only, no code generation).
    std::string describe_json(const std::string& options_array) const;
    const std::string& module_source() const;
    int64_t op_count() const;
    int n_kernels() const;

private:
    FusedPlan();
    void load();
    void prepare();
    struct Impl;
    std::unique_ptr<Impl> impl_;
    void* graph_ = nullptr;  // per-op executor's captured graph
};

}  // namespace bsz

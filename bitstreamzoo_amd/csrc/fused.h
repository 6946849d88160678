// fused.h — the fused executor: segment DAG, partitioner, stream memory
// planner, HIP code generation, hipRTC compilation and the launch scheduler.
#pragma once

#include "executor.h"

#include <string>

namespace bsz {

struct FusedOptions {
    int chunk_rows = 16;        // 64-word rows per wave-chunk
    int warmup_rows = 1;        // speculative warm-up rows before a chunk
    int waves_per_block = 4;    // independent wavefronts per workgroup
    int max_ops = 4000;         // fusion depth: DAG nodes per kernel
    int max_live = 96;          // live 64-bit values per lane per kernel
    std::string cache_dir;      // code-object cache ("" = $BSZ_CACHE_DIR or ~/.cache/bsz_hip)
    bool verbose = false;
    static FusedOptions from_env();
};

std::unique_ptr<Engine> make_fused_engine(hipStream_t s, const FusedOptions& opt);

}  // namespace bsz

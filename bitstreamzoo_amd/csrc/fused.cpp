// fused.cpp — placeholder until the fused engine lands.
#include "fused.h"

namespace bsz {

FusedOptions FusedOptions::from_env() { return FusedOptions{}; }

std::unique_ptr<Engine> make_fused_engine(hipStream_t, const FusedOptions&) {
    throw RunError("fused executor not available in this build");
}

}  // namespace bsz

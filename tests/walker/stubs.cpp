// stubs.cpp -- device entry points the walker's per-op engine links against;
// the CPU harness never calls them (test infrastructure).
#include "device.h"
namespace bsz {
int64_t kernel_launch_count() { return 0; }
uint64_t read_popcount_slots(hipStream_t, const unsigned long long*) { abort(); }
DevPool::~DevPool() {}
void* DevPool::get(size_t) { abort(); }
void DevPool::put(void*, size_t) { abort(); }
void DevPool::trim() {}
void launch_fill(hipStream_t, uint64_t*, bool, const Geom&) { abort(); }
void launch_shl(hipStream_t, uint64_t*, const uint64_t*, uint64_t, const Geom&) { abort(); }
void launch_shr(hipStream_t, uint64_t*, const uint64_t*, uint64_t, const Geom&) { abort(); }
void launch_pointwise(hipStream_t, PwOp, uint64_t*, const uint64_t*, const uint64_t*, const Geom&) { abort(); }
void launch_add(hipStream_t, uint64_t*, const uint64_t*, const uint64_t*, const Geom&, void*) { abort(); }
void launch_popcount(hipStream_t, unsigned long long*, const uint64_t*, const Geom&) { abort(); }
void launch_nonzero(hipStream_t, unsigned long long*, const uint64_t*, const Geom&) { abort(); }
void launch_zero_u64(hipStream_t, unsigned long long*, uint64_t) { abort(); }
void upload_stream(hipStream_t, uint64_t*, const uint64_t*, uint64_t, const Geom&) { abort(); }
}

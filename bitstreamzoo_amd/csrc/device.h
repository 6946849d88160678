// device.h — stream geometry, device buffers and the one-kernel-per-op HIP
// data plane (the counterpart of CudaBackend::op_*, cuda_backend.cuh:53-153,
// and of kernels.cuh / add_carry.cuh, written for gfx950).
//
// Layout in HBM: a stream of bitlength N occupies lw = ceil(N/64) 64-bit
// words, bit 0 = LSB of word 0 (bsdata.py:25-34).  Buffers are allocated in
// whole rows of 64 words (one wavefront-wide coalesced access of 512 B); every
// word at or above lw and every bit at or above N is kept zero, so any kernel
// may read a whole row without bounds checks and `>>` never sees stale bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bsz {

struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };

#define BSZ_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::bsz::HipError(std::string(#call) + ": " + hipGetErrorString(e_) +  \
                                  (e_ == hipErrorOutOfMemory ? " (out of memory)" : "")); \
    } while (0)

constexpr uint64_t kRowWords = 64;

struct Geom {
    int64_t bitlength = 0;   // logical length in bits (0: unbounded, as bsim)
    uint64_t lw = 1;         // logical words
    uint64_t rows = 1;       // 64-word rows
    uint64_t lastmask = ~0ull;  // valid bits of word lw-1

    // Words an op result may occupy, and the mask of the last of them.  By
    // default (clean values, DSL.md:273-275) they are lw and lastmask: no bit
    // at or above the bitlength is ever set.  bsim_padding() widens them to
    // the reference executor's padded stream, whose padding bits ~, ONES, +
    // and << leave set and >> can bring back.
    uint64_t opw = 1;
    uint64_t opmask = ~0ull;

    static Geom make(int64_t bitlength, uint64_t words_if_unbounded) {
        Geom g;
        g.bitlength = bitlength;
        g.lw = bitlength > 0 ? (uint64_t)(bitlength + 63) / 64 : std::max<uint64_t>(1, words_if_unbounded);
        g.rows = (g.lw + kRowWords - 1) / kRowWords;
        uint64_t rem = bitlength > 0 ? (uint64_t)bitlength % 64 : 0;
        g.lastmask = rem ? ((1ull << rem) - 1) : ~0ull;
        g.opw = g.lw;
        g.opmask = g.lastmask;
        return g;
    }
    // The padded stream of bsim_cuda with `word_bits`-bit words: its word
    // count rounded up to a multiple of 8 (cuda_backend.cuh:23-25), i.e.
    // dirty bits up to a multiple of 512 (64-bit) or 256 (32-bit) bits.
    // word_bits 0 keeps clean values.
    Geom bsim_padding(int word_bits) const {
        Geom g = *this;
        if (word_bits != 32 && word_bits != 64) return g;
        const uint64_t bits = bitlength > 0 ? (uint64_t)bitlength : lw * 64;
        const uint64_t words = word_bits == 32 && bitlength > 0 ? (bits + 31) / 32 : (word_bits == 32 ? lw * 2 : lw);
        const uint64_t padded = (words + 7) & ~uint64_t(7);
        g.opw = std::min<uint64_t>(alloc_words(), word_bits == 32 ? padded / 2 : padded);
        g.opmask = ~0ull;
        return g;
    }
    uint64_t alloc_words() const { return rows * kRowWords; }
    uint64_t bytes() const { return alloc_words() * 8; }
};

// Caching device allocator: freed blocks are kept per size and handed back.
class DevPool {
public:
    ~DevPool();
    void* get(size_t bytes);
    void put(void* p, size_t bytes);
    void trim();
    int64_t n_mallocs() const { return n_mallocs_; }
    size_t live_bytes() const { return live_; }
    size_t peak_bytes() const { return peak_; }

private:
    std::multimap<size_t, void*> free_;
    int64_t n_mallocs_ = 0;
    size_t live_ = 0, peak_ = 0;
};

// A device stream buffer (clean padding), returned to its pool when dropped.
struct DevBuf {
    DevPool* pool = nullptr;
    uint64_t* ptr = nullptr;
    size_t bytes = 0;
    DevBuf(DevPool* p, size_t b) : pool(p), ptr((uint64_t*)p->get(b)), bytes(b) {}
    ~DevBuf() { if (ptr) pool->put(ptr, bytes); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};
using DevBufP = std::shared_ptr<DevBuf>;

// ── per-op kernels (kernels.hip); all asynchronous on `s` ──
enum class PwOp : int { And = 0, Or = 1, Xor = 2, Not = 3, Copy = 4 };

void launch_pointwise(hipStream_t s, PwOp op, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g);
void launch_fill(hipStream_t s, uint64_t* dst, bool ones, const Geom& g);
void launch_shl(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g);
void launch_shr(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g);
// Stream addition: a main launch (one wave per segment of kRows rows, each
// settling its own carry-in from at most kWarm rows below it) and a fix-up
// launch that does nothing unless some segment's rows below all propagated.
// `scan_state` holds AddScanState::bytes(g) zero-initialised bytes (a count,
// a carry per segment, a bitmap of unsettled segments); the two launches
// leave it reusable for the next addition of the same geometry.
struct AddScanState {
    static constexpr int kRows = 4;
    static constexpr int kWarm = 4;
    static uint64_t segments(const Geom& g) { return (g.rows + kRows - 1) / kRows; }
    static size_t bytes(const Geom& g) { return 64 + segments(g) * 4 + (segments(g) + 31) / 32 * 4; }
};
void launch_add(hipStream_t s, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g, void* scan_state);
// The same addition with its scratch in a fixed device array of the library
// (no allocation): streams of at most kNoAllocMaxBits bits.  The array is
// shared by every call on a device, so such calls must not overlap there.
// setup_noalloc_scratch(device) must run once with `device` current.
constexpr uint64_t kNoAllocMaxBits = 1ull << 34;
void setup_noalloc_scratch(int device);
void launch_add_noalloc(hipStream_t s, int device, uint64_t* dst, const uint64_t* a, const uint64_t* b,
                        const Geom& g);
// A popcount's partial sums: one u64 per block of its fixed grid of
// kPopSlotWords blocks, each written (not added) by its block, so the buffer
// needs no clearing and takes no atomics.
constexpr size_t kPopSlotWords = 1024;
// popcount bounded to the bitlength, as partial sums into `slots`.
void launch_popcount(hipStream_t s, unsigned long long* slots, const uint64_t* src, const Geom& g);
// The sum of the partials (synchronous).
uint64_t read_popcount_slots(hipStream_t s, const unsigned long long* slots);
// is_nonzero bounded to the bitlength, or-ed into *slot (zeroed by the caller).
void launch_nonzero(hipStream_t s, unsigned long long* slot, const uint64_t* src, const Geom& g);
void launch_zero_u64(hipStream_t s, unsigned long long* p, uint64_t n);
// Load `n` host-side words into a row-padded buffer, truncated to the stream
// (bytes_to_words, word_utils.h:91-108): masked to the bitlength for clean
// values, kept to whole words under bsim padding, as bsim_cuda's reader does.
void upload_stream(hipStream_t s, uint64_t* dst, const uint64_t* host, uint64_t n, const Geom& g);
// Host-to-device copy, asynchronous on s; copies of 4 MB and more pass through
// two pinned staging buffers of 32 MB (BSZ_PINNED_STAGING=0: direct copy).
// The host buffer may be reused as soon as this returns.
void stage_h2d(hipStream_t s, void* dst, const void* src, size_t bytes);
// Device-to-host copy through the same buffers; synchronous.
void stage_d2h(hipStream_t s, void* dst, const void* src, size_t bytes);
// Copy-engine bandwidth probe kernel: dst[i] = src[i] over n words.
void launch_copy_words(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t n);

int64_t kernel_launch_count();

// Kernel launches one captured hipGraph may hold.  Capturing 131k per-op
// launches (2048 Shift-And instances) into one graph ended in a host segfault
// on MI355X / ROCm 7 whose cause was not found; graphs of up to 20k launches
// have worked, so larger captures are refused with an error.
constexpr int64_t kMaxGraphLaunches = 20000;

// Ends a capture on `s` that is still open and drops its partial graph, so a
// failed capture never leaves the stream in capture mode.
void abandon_capture(hipStream_t s);

}  // namespace bsz

// kernels.hip — the one-kernel-per-op data plane for gfx950.
//
// Counterparts in the reference: kern_not/and/or/xor (kernels.cuh:7-36),
// kern_shl/kern_shr (kernels.cuh:54-102), kern_add + gen/prop + CUB scan +
// carry-apply (kernels.cuh:106-112, add_carry.cuh:67-135), popcount per word +
// CUB sum (kernels.cuh:116-128, cuda_backend.cuh:87-120), CUB max for
// is_nonzero (cuda_backend.cuh:122-153).  Here:
//   * pointwise kernels move two words per lane (16-byte accesses);
//   * stream `+` is ONE kernel: a wave computes per-lane sums, turns the lane
//     generate/propagate predicates into 64-bit ballots and resolves carries
//     across its 64 lanes with scalar arithmetic on the ballot masks
//     (C = (G + (G|P) + cin) ^ G ^ (G|P)); tiles chain through a
//     decoupled-lookback scan of 2-bit carry transfer functions;
//   * popcount is v_bcnt per lane, a wavefront reduction and one atomic add
//     per block; no host round trip is needed unless the value is read.
#include "device.h"

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace bsz {

namespace {

std::atomic<int64_t> g_launches{0};

__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Masked value of word w of a stream with `lw` logical words.
__device__ __forceinline__ uint64_t clamp_word(uint64_t v, uint64_t w, uint64_t lw, uint64_t lastmask) {
    return w + 1 < lw ? v : (w + 1 == lw ? (v & lastmask) : 0ull);
}

// Carry out of the 64-bit sum G + GP + c (c in {0,1}).
__device__ __forceinline__ unsigned carry_out(uint64_t G, uint64_t GP, unsigned c) {
    uint64_t t = G + GP;
    unsigned c1 = t < G;
    uint64_t u = t + (uint64_t)c;
    return c1 | (unsigned)(u < t);
}

template <int OP>
__global__ __launch_bounds__(256) void k_pointwise(uint64_t* __restrict__ dst, const uint64_t* __restrict__ a,
                                                   const uint64_t* __restrict__ b, uint64_t n_pairs,
                                                   uint64_t lw, uint64_t lastmask) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pairs) return;
    const ulonglong2 x = reinterpret_cast<const ulonglong2*>(a)[i];
    ulonglong2 r;
    if constexpr (OP == (int)PwOp::Not) {
        r.x = clamp_word(~x.x, 2 * i, lw, lastmask);
        r.y = clamp_word(~x.y, 2 * i + 1, lw, lastmask);
    } else if constexpr (OP == (int)PwOp::Copy) {
        r = x;
    } else {
        const ulonglong2 y = reinterpret_cast<const ulonglong2*>(b)[i];
        if constexpr (OP == (int)PwOp::And) { r.x = x.x & y.x; r.y = x.y & y.y; }
        if constexpr (OP == (int)PwOp::Or)  { r.x = x.x | y.x; r.y = x.y | y.y; }
        if constexpr (OP == (int)PwOp::Xor) { r.x = x.x ^ y.x; r.y = x.y ^ y.y; }
    }
    reinterpret_cast<ulonglong2*>(dst)[i] = r;
}

__global__ __launch_bounds__(256) void k_fill(uint64_t* __restrict__ dst, uint64_t v, uint64_t n_pairs,
                                              uint64_t lw, uint64_t lastmask) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pairs) return;
    ulonglong2 r;
    r.x = clamp_word(v, 2 * i, lw, lastmask);
    r.y = clamp_word(v, 2 * i + 1, lw, lastmask);
    reinterpret_cast<ulonglong2*>(dst)[i] = r;
}

// dst[w] = src[w-q] << bs | src[w-q-1] >> (64-bs)   (advance by k = 64q + bs)
__global__ __launch_bounds__(256) void k_shl(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src,
                                             uint64_t q, unsigned bs, uint64_t n, uint64_t lw, uint64_t lastmask) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    uint64_t hi = w >= q ? src[w - q] : 0ull;
    uint64_t r = hi;
    if (bs) {
        uint64_t lo = w >= q + 1 ? src[w - q - 1] : 0ull;
        r = (hi << bs) | (lo >> (64 - bs));
    }
    dst[w] = clamp_word(r, w, lw, lastmask);
}

// dst[w] = src[w+q] >> bs | src[w+q+1] << (64-bs); words past the buffer are 0.
__global__ __launch_bounds__(256) void k_shr(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src,
                                             uint64_t q, unsigned bs, uint64_t n) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    uint64_t lo = (q < n && w < n - q) ? src[w + q] : 0ull;
    uint64_t r = lo;
    if (bs) {
        uint64_t hi = (q + 1 < n && w < n - q - 1) ? src[w + q + 1] : 0ull;
        r = (lo >> bs) | (hi << (64 - bs));
    }
    dst[w] = r;
}

// Stream addition, one segment of kAddRows rows (64 words each) per wave.
// Lane sums, a ballot of the lanes that generate a carry (G) and of those
// that generate or propagate one (GP), and the carries into all 64 lanes of
// a row from the scalar identity C = (G + GP + cin) ^ G ^ GP.  A wave fixes
// its carry-in from the rows just below its segment: the highest lane there
// that does not propagate (it generates or kills a carry) decides the row's
// carry-out, whatever came from further below.  Up to kAddWarm rows are
// examined, nearest first; the start of the stream has carry 0.  A segment
// whose examined rows all propagate speculates carry-in 0 and is marked;
// the fix-up launch walks the marked segments in order.  No atomics on a
// shared address, no waits, no cross-wave dependency.
__device__ __forceinline__ uint64_t sum_with_carries(uint64_t s, uint64_t G, uint64_t GP, unsigned cin, unsigned lane,
                                                     unsigned* cout) {
    const uint64_t t = G + GP;
    const unsigned c1 = t < G;
    const uint64_t u = t + (uint64_t)cin;
    *cout = c1 | (unsigned)(u < t);
    return s + (((u ^ G ^ GP) >> lane) & 1ull);
}

__global__ __launch_bounds__(256) void k_add(uint64_t* __restrict__ dst, const uint64_t* __restrict__ a,
                                             const uint64_t* __restrict__ b, uint64_t n_rows, uint64_t lw,
                                             uint64_t lastmask, unsigned* __restrict__ ctl,
                                             unsigned* __restrict__ seg_carry, unsigned* __restrict__ unproven,
                                             uint64_t n_segs) {
    const unsigned lane = lane_id();
    const uint64_t seg = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (seg >= n_segs) return;
    const uint64_t r0 = seg * AddScanState::kRows;
    const uint64_t nr = n_rows - r0 < (uint64_t)AddScanState::kRows ? n_rows - r0 : (uint64_t)AddScanState::kRows;
    uint64_t x[AddScanState::kRows], y[AddScanState::kRows];
#pragma unroll
    for (int r = 0; r < AddScanState::kRows; ++r) {
        const uint64_t w = (r0 + r) * 64 + lane;
        x[r] = (uint64_t)r < nr ? a[w] : 0ull;
        y[r] = (uint64_t)r < nr ? b[w] : 0ull;
    }
    unsigned cin = 0;
    bool proven = true;
    if (r0 > 0) {
        proven = false;
        for (int k = 1; k <= AddScanState::kWarm && !proven; ++k) {
            if ((uint64_t)k > r0) { proven = true; break; }  // the stream starts below: carry 0
            const uint64_t w = (r0 - k) * 64 + lane;
            const uint64_t xa = a[w], s = xa + b[w];
            const uint64_t G = ballot(s < xa);
            const uint64_t NP = ballot(s != ~0ull);  // lanes that generate or kill
            if (NP) {
                cin = (unsigned)((G >> (63 - __builtin_clzll(NP))) & 1ull);
                proven = true;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < AddScanState::kRows; ++r) {
        if ((uint64_t)r >= nr) break;
        const uint64_t w = (r0 + r) * 64 + lane;
        const uint64_t s = x[r] + y[r];
        const uint64_t G = ballot(s < x[r]);
        const uint64_t GP = G | ballot(s == ~0ull);
        unsigned cout;
        const uint64_t v = sum_with_carries(s, G, GP, cin, lane, &cout);
        cin = cout;
        dst[w] = clamp_word(v, w, lw, lastmask);
    }
    if (lane == 0) {
        seg_carry[seg] = cin;
        if (!proven) {
            __hip_atomic_fetch_or(&unproven[seg >> 5], 1u << (seg & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Fix-up of k_add (one wave): ends at once when no segment was marked;
// otherwise walks the marked segments in increasing order, gives each the
// carry-out of the segment below (exact: proven by its own warm-up, or
// corrected earlier in this walk), recomputes it when that carry is 1, and
// clears the marks for the next launch.
__global__ __launch_bounds__(64) void k_add_fix(uint64_t* __restrict__ dst, const uint64_t* __restrict__ a,
                                                const uint64_t* __restrict__ b, uint64_t n_rows, uint64_t lw,
                                                uint64_t lastmask, unsigned* __restrict__ ctl,
                                                unsigned* __restrict__ seg_carry, unsigned* __restrict__ unproven,
                                                uint64_t n_segs) {
    const unsigned lane = lane_id();
    if (__hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
    const uint64_t nbw = (n_segs + 31) / 32;
    uint64_t last_seg = ~0ull;
    unsigned last_carry = 0;
    for (uint64_t wb = 0; wb < nbw; wb += 64) {
        const uint64_t wi = wb + lane;
        unsigned bits = wi < nbw ? unproven[wi] : 0u;
        if (wi < nbw) unproven[wi] = 0u;
        uint64_t hit = ballot(bits != 0u);
        while (hit) {
            const unsigned l = (unsigned)__builtin_ctzll(hit);
            hit &= hit - 1;
            unsigned word = (unsigned)__builtin_amdgcn_readlane((int)bits, (int)l);
            while (word) {
                const uint64_t seg = (wb + l) * 32 + (unsigned)__builtin_ctz(word);
                word &= word - 1;
                unsigned cin = seg == last_seg + 1 ? last_carry : seg_carry[seg - 1];
                if (cin) {
                    const uint64_t r0 = seg * AddScanState::kRows;
                    for (uint64_t r = r0; r < r0 + AddScanState::kRows && r < n_rows; ++r) {
                        const uint64_t w = r * 64 + lane;
                        const uint64_t xa = a[w], s = xa + b[w];
                        const uint64_t G = ballot(s < xa);
                        const uint64_t GP = G | ballot(s == ~0ull);
                        unsigned cout;
                        const uint64_t v = sum_with_carries(s, G, GP, cin, lane, &cout);
                        cin = cout;
                        dst[w] = clamp_word(v, w, lw, lastmask);
                    }
                    if (lane == 0) seg_carry[seg] = cin;
                } else {
                    cin = seg_carry[seg];  // speculation was right: the stored carry-out stands
                }
                last_seg = seg;
                last_carry = cin;
            }
        }
    }
    if (lane == 0) __hip_atomic_store(&ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Bounded popcount over a fixed grid of kPopSlotWords blocks: 16 bytes per
// lane per load, two loads in flight per lane, v_bcnt, a wave and a block
// reduction, and one plain store of the block's sum into slots[block].
// Every block stores (0 when it had no words), so the partials need no
// clearing and no address takes an atomic.
__device__ __forceinline__ uint64_t pop_pair(const ulonglong2 v, uint64_t i, uint64_t lw, uint64_t lastmask) {
    const uint64_t w = 2 * i;
    uint64_t c = (uint64_t)__popcll(w + 1 == lw ? (v.x & lastmask) : v.x);
    if (w + 1 < lw) c += (uint64_t)__popcll(w + 2 == lw ? (v.y & lastmask) : v.y);
    return c;
}

__global__ __launch_bounds__(256) void k_popcount(unsigned long long* __restrict__ slots,
                                                  const uint64_t* __restrict__ src, uint64_t lw,
                                                  uint64_t lastmask) {
    __shared__ uint64_t part[4];
    uint64_t acc = 0;
    const uint64_t pairs = (lw + 1) / 2;   // the buffer is row-padded: word lw (if odd) exists and is 0
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const ulonglong2* p = reinterpret_cast<const ulonglong2*>(src);
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + stride < pairs; i += 2 * stride) {
        const ulonglong2 v0 = p[i], v1 = p[i + stride];
        acc += pop_pair(v0, i, lw, lastmask) + pop_pair(v1, i + stride, lw, lastmask);
    }
    if (i < pairs) acc += pop_pair(p[i], i, lw, lastmask);
    acc = wave_sum(acc);
    if (lane_id() == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) slots[blockIdx.x] = (unsigned long long)(part[0] + part[1] + part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_nonzero(unsigned long long* __restrict__ slot,
                                                 const uint64_t* __restrict__ src, uint64_t lw,
                                                 uint64_t lastmask) {
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < lw; w += stride)
        acc |= (w + 1 == lw ? (src[w] & lastmask) : src[w]);
    if (ballot(acc != 0) != 0 && lane_id() == 0) atomicOr(slot, 1ull);
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// After a host copy of `k` words: zero words [k, n) and mask word lw-1.
__global__ __launch_bounds__(256) void k_tail(uint64_t* dst, uint64_t k, uint64_t n, uint64_t lw, uint64_t lastmask) {
    uint64_t w = k + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n) dst[w] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0 && lw >= 1 && lw - 1 < k) dst[lw - 1] &= lastmask;
}

// Copies n_pairs 16-byte word pairs.
__global__ __launch_bounds__(256) void k_copy_words(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src,
                                                    uint64_t n_pairs) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pairs) reinterpret_cast<ulonglong2*>(dst)[i] = reinterpret_cast<const ulonglong2*>(src)[i];
}

inline unsigned blocks_for(uint64_t items, unsigned per_block) {
    return (unsigned)((items + per_block - 1) / per_block);
}

void check_launch() {
    g_launches.fetch_add(1, std::memory_order_relaxed);
    BSZ_HIP(hipGetLastError());
}

}  // namespace

int64_t kernel_launch_count() { return g_launches.load(); }

void launch_pointwise(hipStream_t s, PwOp op, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g) {
    const uint64_t pairs = g.alloc_words() / 2;
    const unsigned grid = blocks_for(pairs, 256);
    switch (op) {
        case PwOp::And: k_pointwise<(int)PwOp::And><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.opw, g.opmask); break;
        case PwOp::Or:  k_pointwise<(int)PwOp::Or><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.opw, g.opmask); break;
        case PwOp::Xor: k_pointwise<(int)PwOp::Xor><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.opw, g.opmask); break;
        case PwOp::Not: k_pointwise<(int)PwOp::Not><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.opw, g.opmask); break;
        case PwOp::Copy: k_pointwise<(int)PwOp::Copy><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.opw, g.opmask); break;
    }
    check_launch();
}

void launch_fill(hipStream_t s, uint64_t* dst, bool ones, const Geom& g) {
    const uint64_t pairs = g.alloc_words() / 2;
    k_fill<<<blocks_for(pairs, 256), 256, 0, s>>>(dst, ones ? ~0ull : 0ull, pairs, g.opw, g.opmask);
    check_launch();
}

void launch_shl(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g) {
    const uint64_t n = g.alloc_words();
    const uint64_t q = k / 64;
    if (q >= n) { launch_fill(s, dst, false, g); return; }
    k_shl<<<blocks_for(n, 256), 256, 0, s>>>(dst, src, q, (unsigned)(k % 64), n, g.opw, g.opmask);
    check_launch();
}

void launch_shr(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g) {
    const uint64_t n = g.alloc_words();
    const uint64_t q = k / 64;
    if (q >= n) { launch_fill(s, dst, false, g); return; }
    k_shr<<<blocks_for(n, 256), 256, 0, s>>>(dst, src, q, (unsigned)(k % 64), n);
    check_launch();
}

namespace {

void launch_add_parts(hipStream_t s, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g,
                      unsigned* ctl, unsigned* carry, unsigned* unproven) {
    const uint64_t segs = AddScanState::segments(g);
    const uint64_t blocks = (segs + 3) / 4;
    if (blocks > 0x7fffffffull) throw HipError("stream too long for one add launch");
    k_add<<<(unsigned)blocks, 256, 0, s>>>(dst, a, b, g.rows, g.opw, g.opmask, ctl, carry, unproven, segs);
    check_launch();
    k_add_fix<<<1, 64, 0, s>>>(dst, a, b, g.rows, g.opw, g.opmask, ctl, carry, unproven, segs);
    check_launch();
}

// Scratch of the allocation-free addition: 16 control words, the bitmap of
// unsettled segments and one carry per segment, for streams of up to
// kNoAllocMaxBits bits (4.3 MB).  Each device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for that device.
constexpr uint64_t kNoAllocSegs = kNoAllocMaxBits / 64 / kRowWords / AddScanState::kRows;
constexpr uint64_t kNoAllocBitmapWords = (kNoAllocSegs + 31) / 32;

std::mutex g_scratch_mu;
std::vector<unsigned*> g_scratch_ptr;  // per device; null until set up

}  // namespace

static __device__ __align__(256) unsigned scratch[16 + kNoAllocBitmapWords + kNoAllocSegs];

void launch_add(hipStream_t s, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g, void* scan_state) {
    unsigned* ctl = reinterpret_cast<unsigned*>(scan_state);
    unsigned* carry = ctl + 16;
    launch_add_parts(s, dst, a, b, g, ctl, carry, carry + AddScanState::segments(g));
}

void setup_noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) throw HipError("bad device index");
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[(size_t)device]) return;
    void* p = nullptr;
    BSZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the add scratch");
    g_scratch_ptr[(size_t)device] = static_cast<unsigned*>(p);
}

void launch_add_noalloc(hipStream_t s, int device, uint64_t* dst, const uint64_t* a, const uint64_t* b,
                        const Geom& g) {
    if (g.bitlength <= 0 || (uint64_t)g.bitlength > kNoAllocMaxBits)
        throw std::invalid_argument("bsz_op_add_noalloc: bitlength must be in [1, 2^34] (the size of its fixed "
                                    "scratch); use bsz_op_add for longer or unbounded streams");
    unsigned* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= 0 && (size_t)device < g_scratch_ptr.size()) base = g_scratch_ptr[(size_t)device];
    }
    if (!base) throw HipError("bsz_op_add_noalloc: the scratch of this device was not set up");
    const uint64_t segs = AddScanState::segments(g);
    // the control words and this geometry's bitmap start from zero (a fill
    // kernel, not a memset, so the call may be captured into a graph)
    const uint64_t clear_u32 = 16 + (segs + 31) / 32;
    launch_zero_u64(s, reinterpret_cast<unsigned long long*>(base), (clear_u32 + 1) / 2);
    launch_add_parts(s, dst, a, b, g, base, base + 16 + kNoAllocBitmapWords, base + 16);
}

void launch_popcount(hipStream_t s, unsigned long long* slots, const uint64_t* src, const Geom& g) {
    // bsim: popcount of an unbounded case is 0 (cuda_backend.cuh:88); the
    // grid still runs, over no words, so that every partial is written
    const uint64_t lw = g.bitlength > 0 ? g.lw : 0;
    k_popcount<<<(unsigned)kPopSlotWords, 256, 0, s>>>(slots, src, lw, g.lastmask);
    check_launch();
}

uint64_t read_popcount_slots(hipStream_t s, const unsigned long long* slots) {
    std::vector<unsigned long long> h(kPopSlotWords);
    BSZ_HIP(hipMemcpyAsync(h.data(), slots, h.size() * 8, hipMemcpyDeviceToHost, s));
    BSZ_HIP(hipStreamSynchronize(s));
    uint64_t t = 0;
    for (unsigned long long v : h) t += v;
    return t;
}

void launch_nonzero(hipStream_t s, unsigned long long* slot, const uint64_t* src, const Geom& g) {
    if (g.bitlength <= 0) return;
    const unsigned grid = std::min<unsigned>(2048u, blocks_for(g.lw, 256));
    k_nonzero<<<grid, 256, 0, s>>>(slot, src, g.lw, g.lastmask);
    check_launch();
}

void launch_zero_u64(hipStream_t s, unsigned long long* p, uint64_t n) {
    if (!n) return;
    k_zero_u64<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    check_launch();
}

namespace {

// Two pinned host buffers that large host-to-device copies pass through: the
// copy of the next piece into one buffer overlaps the DMA of the other, and
// the DMA engine reads page-locked memory at full PCIe rate.
struct PinnedStage {
    static constexpr size_t kBytes = size_t(32) << 20;
    std::mutex mu;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    int device = -1;
};
PinnedStage g_stage;

// Buffers and events for the current device (events belong to a device);
// called with the mutex held.
void stage_ready() {
    int dev = 0;
    BSZ_HIP(hipGetDevice(&dev));
    if (g_stage.device == dev) return;
    for (int b = 0; b < 2; ++b) {
        if (g_stage.done[b]) { (void)hipEventSynchronize(g_stage.done[b]); (void)hipEventDestroy(g_stage.done[b]); }
        g_stage.done[b] = nullptr;
        if (!g_stage.buf[b]) BSZ_HIP(hipHostMalloc(&g_stage.buf[b], PinnedStage::kBytes, hipHostMallocPortable));
        BSZ_HIP(hipEventCreateWithFlags(&g_stage.done[b], hipEventDisableTiming));
    }
    g_stage.device = dev;
}

bool staging_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("BSZ_PINNED_STAGING");
        return !(v && v[0] == '0');
    }();
    return on;
}

}  // namespace

void stage_h2d(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (bytes < (size_t(4) << 20) || !staging_enabled()) {
        BSZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return;
    }
    std::lock_guard<std::mutex> lk(g_stage.mu);
    stage_ready();
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    for (size_t off = 0, i = 0; off < bytes; off += PinnedStage::kBytes, ++i) {
        const int b = (int)(i & 1);
        const size_t n = std::min(PinnedStage::kBytes, bytes - off);
        BSZ_HIP(hipEventSynchronize(g_stage.done[b]));  // the buffer's previous transfer has finished
        std::memcpy(g_stage.buf[b], from + off, n);
        BSZ_HIP(hipMemcpyAsync(to + off, g_stage.buf[b], n, hipMemcpyHostToDevice, s));
        BSZ_HIP(hipEventRecord(g_stage.done[b], s));
    }
}

void stage_d2h(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (bytes < (size_t(4) << 20) || !staging_enabled()) {
        BSZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        BSZ_HIP(hipStreamSynchronize(s));
        return;
    }
    std::lock_guard<std::mutex> lk(g_stage.mu);
    stage_ready();
    for (int b = 0; b < 2; ++b) BSZ_HIP(hipEventSynchronize(g_stage.done[b]));  // earlier uploads have left
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    // piece i goes to buffer i & 1; piece i - 1 is copied out while piece i transfers
    const size_t pieces = (bytes + PinnedStage::kBytes - 1) / PinnedStage::kBytes;
    for (size_t i = 0; i <= pieces; ++i) {
        if (i < pieces) {
            const size_t off = i * PinnedStage::kBytes, n = std::min(PinnedStage::kBytes, bytes - off);
            BSZ_HIP(hipMemcpyAsync(g_stage.buf[i & 1], from + off, n, hipMemcpyDeviceToHost, s));
            BSZ_HIP(hipEventRecord(g_stage.done[i & 1], s));
        }
        if (i > 0) {
            const size_t j = i - 1, off = j * PinnedStage::kBytes, n = std::min(PinnedStage::kBytes, bytes - off);
            BSZ_HIP(hipEventSynchronize(g_stage.done[j & 1]));
            std::memcpy(to + off, g_stage.buf[j & 1], n);
        }
    }
}

void upload_stream(hipStream_t s, uint64_t* dst, const uint64_t* host, uint64_t n, const Geom& g) {
    const uint64_t k = std::min<uint64_t>(n, g.lw);
    if (k) stage_h2d(s, dst, host, k * 8);
    const uint64_t total = g.alloc_words();
    // opmask is lastmask for clean values and all ones under bsim padding
    k_tail<<<std::max(1u, blocks_for(total - k, 256)), 256, 0, s>>>(dst, k, total, g.lw, g.opmask);
    check_launch();
}

void launch_copy_words(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t n) {
    const uint64_t pairs = n / 2;
    k_copy_words<<<blocks_for(pairs, 256), 256, 0, s>>>(dst, src, pairs);
    check_launch();
}

void abandon_capture(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st == hipStreamCaptureStatusNone) return;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
}

// ── DevPool ──

DevPool::~DevPool() { trim(); }

void* DevPool::get(size_t bytes) {
    auto it = free_.find(bytes);
    void* p = nullptr;
    if (it != free_.end()) {
        p = it->second;
        free_.erase(it);
    } else {
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            trim();
            BSZ_HIP(hipMalloc(&p, bytes));
        }
        ++n_mallocs_;
    }
    live_ += bytes;
    peak_ = std::max(peak_, live_);
    return p;
}

void DevPool::put(void* p, size_t bytes) {
    live_ -= bytes;
    free_.emplace(bytes, p);
}

void DevPool::trim() {
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
}

}  // namespace bsz

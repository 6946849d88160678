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

// One-pass stream addition.  Block = 4 waves; wave v owns rows 4v..4v+3 of
// the tile (row = 64 consecutive words, one per lane).  Status words of the
// decoupled lookback are tagged with the launch epoch so they never need
// clearing:  bits 4.. tag, bits 2-3 kind (1 aggregate, 2 inclusive prefix),
// bits 0-1 the carry transfer function (f(0), f(1)) or the carry-out.
// ctl[0] tile ticket, ctl[1] finished tiles, ctl[2] epoch; the last tile to
// finish resets ctl[0..1] and advances the epoch.
__global__ __launch_bounds__(256) void k_add(uint64_t* __restrict__ dst, const uint64_t* __restrict__ a,
                                             const uint64_t* __restrict__ b, uint64_t n, uint64_t lw,
                                             uint64_t lastmask, unsigned* __restrict__ ctl,
                                             unsigned* __restrict__ status, unsigned n_tiles) {
    __shared__ unsigned s_tile, s_epoch, s_cin;
    __shared__ unsigned s_wf[4];
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        s_tile = atomicAdd(&ctl[0], 1u);
        s_epoch = __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const unsigned tile = s_tile;
    const unsigned tag = ((s_epoch & 0x0fffffffu) + 1u) << 4;
    const uint64_t base = (uint64_t)tile * 1024 + (uint64_t)wave * 256;

    uint64_t sum[4], G[4], GP[4];
    unsigned f0 = 0, f1 = 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint64_t w = base + (uint64_t)r * 64 + lane;
        const uint64_t x = w < n ? a[w] : 0ull;
        const uint64_t y = w < n ? b[w] : 0ull;
        sum[r] = x + y;
        G[r] = ballot(sum[r] < x);
        GP[r] = G[r] | ballot(sum[r] == ~0ull);
        f0 = carry_out(G[r], GP[r], f0);
        f1 = carry_out(G[r], GP[r], f1);
    }
    if (lane == 0) s_wf[wave] = f0 | (f1 << 1);
    __syncthreads();

    if (threadIdx.x == 0) {
        unsigned t0 = 0, t1 = 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const unsigned wf = s_wf[v];
            t0 = (wf >> t0) & 1u;
            t1 = (wf >> t1) & 1u;
        }
        unsigned cin = 0;
        if (tile > 0) {
            __hip_atomic_store(&status[tile], tag | (1u << 2) | t0 | (t1 << 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            unsigned h0 = 0, h1 = 1;  // carry into this tile as a function of carry into tile j+1
            for (int j = (int)tile - 1;; --j) {
                unsigned st;
                for (;;) {
                    st = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((st & ~15u) == tag && (st & 12u)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                if (((st >> 2) & 3u) == 2u) { cin = (st & 1u) ? h1 : h0; break; }
                const unsigned g0 = st & 1u, g1 = (st >> 1) & 1u;
                const unsigned n0 = g0 ? h1 : h0, n1 = g1 ? h1 : h0;
                h0 = n0; h1 = n1;
                if (h0 == h1) { cin = h0; break; }
            }
        }
        __hip_atomic_store(&status[tile], tag | (2u << 2) | (cin ? t1 : t0), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        s_cin = cin;
    }
    __syncthreads();

    unsigned cin = s_cin;
    for (unsigned v = 0; v < wave; ++v) cin = (s_wf[v] >> cin) & 1u;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint64_t w = base + (uint64_t)r * 64 + lane;
        const uint64_t t = G[r] + GP[r];
        const unsigned c1 = t < G[r];
        const uint64_t u = t + (uint64_t)cin;
        const unsigned c2 = u < t;
        const uint64_t C = u ^ G[r] ^ GP[r];       // carry into each lane
        const uint64_t v = sum[r] + ((C >> lane) & 1ull);
        cin = c1 | c2;
        if (w < n) dst[w] = clamp_word(v, w, lw, lastmask);
    }

    __syncthreads();
    if (threadIdx.x == 0) {
        // Every block took its ticket before it counts itself done, so the
        // last one can reset the counters; statuses are device-scope atomics
        // and need no cache writeback.
        if (atomicAdd(&ctl[1], 1u) == n_tiles - 1) {
            __hip_atomic_store(&ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_popcount(unsigned long long* __restrict__ slot,
                                                  const uint64_t* __restrict__ src, uint64_t lw,
                                                  uint64_t lastmask) {
    __shared__ uint64_t part[4];
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < lw; w += stride)
        acc += (uint64_t)__popcll(w + 1 == lw ? (src[w] & lastmask) : src[w]);
    acc = wave_sum(acc);
    if (lane_id() == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(slot, (unsigned long long)t);
    }
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

// Ends a fused plan run whose last kernel has no fix-up launch: the run
// counter selects the popcount bank the next run adds into.
__global__ __launch_bounds__(64) void k_plan_epoch(unsigned* ctl) {
    if (threadIdx.x == 0) ctl[0] += 1u;
}

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
        case PwOp::And: k_pointwise<(int)PwOp::And><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.lw, g.lastmask); break;
        case PwOp::Or:  k_pointwise<(int)PwOp::Or><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.lw, g.lastmask); break;
        case PwOp::Xor: k_pointwise<(int)PwOp::Xor><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.lw, g.lastmask); break;
        case PwOp::Not: k_pointwise<(int)PwOp::Not><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.lw, g.lastmask); break;
        case PwOp::Copy: k_pointwise<(int)PwOp::Copy><<<grid, 256, 0, s>>>(dst, a, b, pairs, g.lw, g.lastmask); break;
    }
    check_launch();
}

void launch_fill(hipStream_t s, uint64_t* dst, bool ones, const Geom& g) {
    const uint64_t pairs = g.alloc_words() / 2;
    k_fill<<<blocks_for(pairs, 256), 256, 0, s>>>(dst, ones ? ~0ull : 0ull, pairs, g.lw, g.lastmask);
    check_launch();
}

void launch_shl(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g) {
    const uint64_t n = g.alloc_words();
    const uint64_t q = k / 64;
    if (q >= n) { launch_fill(s, dst, false, g); return; }
    k_shl<<<blocks_for(n, 256), 256, 0, s>>>(dst, src, q, (unsigned)(k % 64), n, g.lw, g.lastmask);
    check_launch();
}

void launch_shr(hipStream_t s, uint64_t* dst, const uint64_t* src, uint64_t k, const Geom& g) {
    const uint64_t n = g.alloc_words();
    const uint64_t q = k / 64;
    if (q >= n) { launch_fill(s, dst, false, g); return; }
    k_shr<<<blocks_for(n, 256), 256, 0, s>>>(dst, src, q, (unsigned)(k % 64), n);
    check_launch();
}

void launch_add(hipStream_t s, uint64_t* dst, const uint64_t* a, const uint64_t* b, const Geom& g, void* scan_state) {
    const uint64_t tiles = AddScanState::tiles(g);
    if (tiles > 0xffffffffull) throw HipError("stream too long for one add launch");
    unsigned* ctl = reinterpret_cast<unsigned*>(scan_state);
    unsigned* status = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(scan_state) + 64);
    k_add<<<(unsigned)tiles, 256, 0, s>>>(dst, a, b, g.alloc_words(), g.lw, g.lastmask, ctl, status, (unsigned)tiles);
    check_launch();
}

void launch_popcount(hipStream_t s, unsigned long long* slot, const uint64_t* src, const Geom& g) {
    if (g.bitlength <= 0) return;  // bsim: popcount of an unbounded case is 0 (cuda_backend.cuh:88)
    const unsigned grid = std::min<unsigned>(2048u, blocks_for(g.lw, 256));
    k_popcount<<<grid, 256, 0, s>>>(slot, src, g.lw, g.lastmask);
    check_launch();
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

void upload_stream(hipStream_t s, uint64_t* dst, const uint64_t* host, uint64_t n, const Geom& g) {
    const uint64_t k = std::min<uint64_t>(n, g.lw);
    if (k) BSZ_HIP(hipMemcpyAsync(dst, host, k * 8, hipMemcpyHostToDevice, s));
    const uint64_t total = g.alloc_words();
    k_tail<<<std::max(1u, blocks_for(total - k, 256)), 256, 0, s>>>(dst, k, total, g.lw, g.lastmask);
    check_launch();
}

void launch_plan_epoch(hipStream_t s, unsigned* ctl) {
    k_plan_epoch<<<1, 64, 0, s>>>(ctl);
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

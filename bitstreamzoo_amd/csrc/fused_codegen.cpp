// fused_codegen.cpp — dead-code elimination, partitioning, intermediate
// buffer planning and HIP source generation for fused kernels (see
// fused_graph.h for the execution model).
#include "fused_graph.h"

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <queue>
#include <set>
#include <sstream>
#include <stdexcept>

namespace bsz {

namespace {

bool is_free(NOp op) { return op == NOp::Input || op == NOp::Zero || op == NOp::Ones; }
int arity(NOp op) {
    switch (op) {
        case NOp::Not: case NOp::Shl: case NOp::Shr: return 1;
        case NOp::And: case NOp::Or: case NOp::Xor: case NOp::Add: case NOp::Sel: return 2;
        default: return 0;
    }
}
// An advance that needs the previous row's word of an in-kernel value.
bool stateful_shift(const GNode& n) { return n.op == NOp::Shl && n.k >= 1 && n.k <= 63; }
// Ops that may not take an operand computed in the same kernel.
bool needs_external_operand(const GNode& n) {
    return n.op == NOp::Shr ? n.k != 0 : (n.op == NOp::Shl && n.k >= 64);
}

uint64_t fnv1a(const std::string& s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
    return h;
}

// The argument block and helpers every generated module starts with.
const char* kKernelPrelude = R"HIP(
typedef unsigned long long u64;
typedef unsigned int u32;
struct InstDesc { u64 rows; u64 lw; u64 lastmask; u64 opw; u64 opmask; u32 chunk_begin; u32 n_chunks; };
struct KArgs {
  const u64* const* in; u64* const* out; u64* pop_acc;
  const InstDesc* inst; const u32* chunk_inst;
  u64* st_t; u32* runs; u32* kstat; u32* err;
  u64 pop_stride; u32 n_chunks; u32 n_inst; u32 last; u32 spread_log; u32 spin_limit; u32 grid_waves;
};
#define AGENT __HIP_MEMORY_SCOPE_AGENT
// Binding tables are never written while a kernel runs: reading them through
// the constant address space lets the compiler use scalar loads, so stream
// pointers live in SGPRs (or VGPR lanes) instead of 2 VGPRs each.
typedef __attribute__((address_space(4))) const void* cptr_t;
__device__ __forceinline__ const u64* tab_in(const u64* const* t, u64 i) {
  return ((const u64* __attribute__((address_space(4))) const*)(cptr_t)t)[i];
}
__device__ __forceinline__ u64* tab_out(u64* const* t, u64 i) {
  return ((u64* __attribute__((address_space(4))) const*)(cptr_t)t)[i];
}
__device__ __forceinline__ InstDesc inst_at(const InstDesc* t, u32 i) {
  return ((const InstDesc __attribute__((address_space(4)))*)(cptr_t)t)[i];
}
__device__ __forceinline__ u32 u32_at(const u32* t, u32 i) {
  return ((const u32 __attribute__((address_space(4)))*)(cptr_t)t)[i];
}
__device__ __forceinline__ u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ u64 ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ u32 uni(u32 x) { return (u32)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ u64 lane63(u64 x) {
  u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)x, 63);
  u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(x >> 32), 63);
  return ((u64)hi << 32) | lo;
}
// x from lane-1; lane 0 receives `low` (the previous row's lane-63 word).
__device__ __forceinline__ u64 prev_word(u64 x, u64 low) {
  u32 lo = (u32)__builtin_amdgcn_update_dpp((int)(u32)low, (int)(u32)x, 0x138, 0xf, 0xf, false);
  u32 hi = (u32)__builtin_amdgcn_update_dpp((int)(u32)(low >> 32), (int)(u32)(x >> 32), 0x138, 0xf, 0xf, false);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 word_mask(long long w, u64 lw, u64 lastmask) {
  return (w < 0) ? 0ull : ((u64)w + 1 < lw ? ~0ull : ((u64)w + 1 == lw ? lastmask : 0ull));
}
// A wave-uniform row pointer made opaque to the optimizer, so that every row
// access is a scalar base plus the lane offset (global_load v, v_off, s[base])
// rather than a per-lane 64-bit pointer kept per stream.
#define GAS __attribute__((address_space(1)))
__device__ __forceinline__ const GAS u64* urow(const u64* p) {
  const u64 v = (u64)p;
  return (const GAS u64*)(((u64)uni((u32)(v >> 32)) << 32) | uni((u32)v));
}
__device__ __forceinline__ GAS u64* urow_w(u64* p) {
  const u64 v = (u64)p;
  return (GAS u64*)(((u64)uni((u32)(v >> 32)) << 32) | uni((u32)v));
}
// Streaming hints (set per unit): 1 = stores bypass the caches' normal
// retention (nontemporal), 2 = loads too.
#ifndef BSZ_NT
#define BSZ_NT 0
#endif
// rowbase[lane] as scalar base + 32-bit byte offset (the saddr form)
__device__ __forceinline__ u64 lane_ld(const GAS u64* rowbase, u32 off8) {
  const GAS u64* p = (const GAS u64*)((const GAS char*)rowbase + off8);
#if BSZ_NT >= 2
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void lane_st(GAS u64* rowbase, u32 off8, u64 v) {
  GAS u64* p = (GAS u64*)((GAS char*)rowbase + off8);
#if BSZ_NT >= 1
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
typedef u64 u64x2 __attribute__((ext_vector_type(2)));
// words 2l, 2l+1 of a row pair: 16-byte load / store at row base + lane * 16
__device__ __forceinline__ u64x2 lane_ld16(const GAS u64* rowbase, u32 off16) {
  const GAS u64x2* p = (const GAS u64x2*)((const GAS char*)rowbase + off16);
#if BSZ_NT >= 2
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void lane_st16(GAS u64* rowbase, u32 off16, u64x2 v) {
  GAS u64x2* p = (GAS u64x2*)((GAS char*)rowbase + off16);
#if BSZ_NT >= 1
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
// word `w` of a stream of n words (0 outside), read at row base + per-lane offset
__device__ __forceinline__ u64 ldr(const GAS u64* rowbase, long long off, long long w, u64 n) {
  return (w >= 0 && (u64)w < n) ? rowbase[off] : 0ull;
}
// An empty asm makes a value opaque to the optimizer.  The AMDGPU
// SelectionDAG backend of this compiler (ROCm 7.2 LLVM; gfx90a, gfx942 and
// gfx950, at -O1 and -O3) miscompiles a 64-bit OR whose operand has a 32-bit
// half known to be all ones and more than one use: the OR takes that half as
// zero.  `~(x >> 61)` used twice, `x | 0xffffffff00000000` used twice, both
// give a wrong half.  Such halves arise here only from values the compiler
// can partly evaluate: an advance or a sum of a value it has proved constant
// (say `b ^ b`).  Every advance and sum is therefore passed through opq(), so
// that no stream value carries a known half.
__device__ __forceinline__ u64 opq(u64 x) { __asm__("" : "+v"(x)); return x; }
// Any boolean function of three 64-bit words in two v_bitop3_b32 (one per
// 32-bit half): bit i of the result is bit (a_i << 2 | b_i << 1 | c_i) of T.
template <unsigned T>
__device__ __forceinline__ u64 bop3(u64 a, u64 b, u64 c) {
  const u32 lo = __builtin_amdgcn_bitop3_b32((u32)a, (u32)b, (u32)c, T);
  const u32 hi = __builtin_amdgcn_bitop3_b32((u32)(a >> 32), (u32)(b >> 32), (u32)(c >> 32), T);
  return opq(((u64)hi << 32) | lo);
}
__device__ __forceinline__ u64 fsl(u64 cur, u64 lower, u32 b) { return b ? ((cur << b) | (lower >> (64 - b))) : cur; }
__device__ __forceinline__ u64 fsr(u64 cur, u64 upper, u32 b) { return b ? ((cur >> b) | (upper << (64 - b))) : cur; }
__device__ __forceinline__ u32 wave_sum(u32 v) {
  const u32 l = lane_id();
  for (u32 m = 32; m > 0; m >>= 1) v += (u32)__builtin_amdgcn_ds_bpermute((int)(((l ^ m) & 63u) << 2), (int)v);
  return v;
}
// A predicate: the popcount slots of one (instance, popcount) in bank 0 (bank
// 1 lies `stride` words on); lane i reads slot i of this run's bank.
__device__ __forceinline__ bool pred_flag(const u64* slots, u64 stride, u32 ep, u32 spread_log) {
  const u32 l = lane_id();
  const u64 v = l < (1u << spread_log) ? ((const volatile u64*)slots)[(ep ? stride : 0ull) + l] : 0ull;
  return ballot(v != 0ull) != 0ull;
}
__device__ __forceinline__ u32 aload32(u32* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT); }
__device__ __forceinline__ u64 aload64(u64* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT); }
__device__ __forceinline__ void astore32(u32* p, u32 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, AGENT); }
__device__ __forceinline__ void astore64(u64* p, u64 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, AGENT); }
)HIP";

struct Gen {
    const SegGraph& g;
    const PlanOptions& o;
    std::ostringstream s;
    Gen(const SegGraph& gg, const PlanOptions& oo) : g(gg), o(oo) {}
};

std::string hex64(uint64_t v) {
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
    return b;
}

// State words (carries, advance neighbours) one kernel may carry from row to
// row; they live in scalar registers, spilled to VGPR lanes when many.  A
// kernel with the AGPRs has the registers for more of them.
int max_state_words(const PlanOptions& opt) { return opt.agprs ? 128 : 32; }
// Kernels of at most this many nodes prove their chunks exact in the warm-up.
constexpr int kTrackNodes = 128;

// Registers a generated kernel may use: the 256 architectural VGPRs at most
// (values kept in AGPRs or scratch cost extra moves on every row), fewer at
// a higher requested occupancy; at occupancy 1 with AGPRs allowed, all 512
// registers of the wave (the compiler moves values between the two files).
int kernel_registers(const PlanOptions& opt) {
    const int occ = opt.occupancy > 0 ? opt.occupancy : 1;
    if (opt.agprs && occ == 1) return 512;
    return std::min(256, 512 / occ);
}

// Every kernel reads the run number of its chunk (how many runs of the plan
// have completed over it): its parity selects this run's popcount bank, and
// a stateful kernel tags its published end states with run + 1.  The wave
// of the chunk clears its share of the other bank for the next run, and the
// wave of chunk 0 clears the next run's counters of a stateful kernel.
void emit_run_prologue(std::ostringstream& s, int NPOP, bool has_state) {
    s << "  const u32 run = u32_at(A.runs, c);\n";
    s << "  const u32 ep = run & 1u;\n";
    if (NPOP) s << "  u64* const bank = A.pop_acc + (ep ? A.pop_stride : 0ull);\n";
    if (has_state) {
        s << "  const u32 tag = run + 1u;\n";
        s << "  u32* const kst = A.kstat + (ep ? 4u : 0u);\n";
        s << "  if (c == 0u && lane < 4u) A.kstat[(ep ? 0u : 4u) + lane] = 0u;\n";
    }
}

// Which chunks a wave runs.  By default wave w of the grid owns chunk w.
// A persistent kernel's grid holds no more waves than the GPU keeps resident
// at the kernel's occupancy, and wave w runs chunks w, w + G, w + 2G, ...
// in increasing order: every wave is resident, so the lowest unfinished
// chunk, whose predecessor is finished, always makes progress, whatever the
// order in which the hardware dispatches workgroups.
void emit_chunk_loop(std::ostringstream& s, const PlanOptions& opt, int WPB) {
    if (opt.persistent) {
        s << "  for (u32 c = blockIdx.x * " << WPB << "u + uni(threadIdx.x >> 6); c < A.n_chunks; c += A.grid_waves) {\n";
    } else {
        s << "  {\n";
        s << "  const u32 c = blockIdx.x * " << WPB << "u + uni(threadIdx.x >> 6);\n";
        s << "  if (c >= A.n_chunks) return;\n";
    }
}

// The plan's last kernel ends the run over each chunk by advancing the
// chunk's run number, after everything else the wave wrote.
void emit_run_epilogue(std::ostringstream& s) {
    s << "  if (A.last && lane == 0u) A.runs[c] = run + 1u;\n";
}

// Generates one kernel; fills ks.source and ks.n_state.
void gen_kernel(const SegGraph& g, const PlanOptions& opt, KernelSpec& ks, bool emit) {
    std::set<int32_t> in_kernel(ks.nodes.begin(), ks.nodes.end());
    std::map<int32_t, int> input_index;     // graph node -> kernel input slot
    for (size_t j = 0; j < ks.input_node.size(); ++j) input_index[ks.input_node[j]] = (int)j;

    // state assignment
    std::map<int32_t, int> state_of;        // node -> state word
    int n_state = 0;
    for (int32_t n : ks.nodes) {
        const GNode& x = g.nodes[n];
        if (x.op == NOp::Add || (stateful_shift(x) && in_kernel.count(x.a))) state_of[n] = n_state++;
    }
    ks.n_state = n_state;
    const bool has_state = n_state > 0;
    // Lane-exactness tracking in warm-up rows costs a scalar per node; large
    // kernels skip it (their chunks after the first warm-up rows all look
    // back at their predecessor's end state), unless they have the AGPRs.
    const bool track = has_state && ((int)ks.nodes.size() <= kTrackNodes || opt.agprs);

    // which graph nodes need a plain value of an input (loaded per row)
    std::set<int32_t> plain_inputs;
    std::vector<int32_t> ext_shifts;        // shift nodes whose operand is not in the kernel
    auto note_operand = [&](int32_t a) {
        if (a >= 0 && !in_kernel.count(a) && g.nodes[a].op != NOp::Zero && g.nodes[a].op != NOp::Ones)
            plain_inputs.insert(a);
    };
    for (int32_t n : ks.nodes) {
        const GNode& x = g.nodes[n];
        if ((x.op == NOp::Shl || x.op == NOp::Shr) && !in_kernel.count(x.a)) {
            if (x.k == 0) note_operand(x.a);
            else ext_shifts.push_back(n);
            continue;
        }
        if (arity(x.op) >= 1) note_operand(x.a);
        if (arity(x.op) >= 2) note_operand(x.b);
    }
    for (int32_t p : ks.pops) note_operand(g.pops[p]);
    for (int32_t n : ks.outputs)
        if (!in_kernel.count(n)) note_operand(n);
    // predicates read by this kernel's selections, each once per chunk
    std::map<int32_t, int> pred_index;
    for (size_t j = 0; j < ks.preds.size(); ++j) pred_index[ks.preds[j]] = (int)j;

    // Rows whose loads are issued together: as many as the option allows while
    // the prefetched words stay within ~96 VGPRs (2 per 64-bit word), so a
    // kernel with many inputs does not spill.
    // Words at a signed offset from the row's own word that external shifts
    // read: one load per distinct (input, offset), shared by every shift.
    std::set<std::pair<int32_t, long long>> off_words;
    for (int32_t n : ext_shifts) {
        const GNode& x = g.nodes[n];
        const GNode& src = g.nodes[x.a];
        if (src.op == NOp::Zero || src.op == NOp::Ones) continue;
        const long long q = (long long)(x.k / 64);
        const bool part = (x.k % 64) != 0;
        if (x.op == NOp::Shl) {
            if (q == 0) plain_inputs.insert(x.a); else off_words.insert({x.a, -q});
            if (part) off_words.insert({x.a, -q - 1});
        } else {
            if (q == 0) plain_inputs.insert(x.a); else off_words.insert({x.a, q});
            if (part) off_words.insert({x.a, q + 1});
        }
    }
    // The chunk length is the plan's, shared by all its kernels.
    const int loads_per_row = (int)(plain_inputs.size() + off_words.size());
    const int CR = std::max(1, opt.chunk_rows);
    // Prefetched words (2 VGPRs each, per row in flight) share the kernel's
    // register file with its live values.
    const int NPOP_ = (int)ks.pops.size();
    const int live_ = ks.peak_live;
    // With an explicit occupancy the prefetch may use that occupancy's whole
    // register file (the compiled kernel's register use is checked by the
    // autotuner); by default it stays within 256 registers.
    const int regs_for_loads = opt.occupancy > 0 ? 512 / opt.occupancy : 256;
    const int load_budget = std::max(0, regs_for_loads - 48 - NPOP_ - 2 * live_);
    const int u_cap = std::max(1, std::min(opt.rows_per_step, load_budget / std::max(1, 2 * loads_per_row)));
    const int WR = has_state ? std::max(1, opt.warmup_rows) : 0;
    // Rows per step divide the chunk length, so a step never runs past its
    // chunk and needs no per-row bound inside the step.  (One kernel with 97
    // state words gave wrong results in the form that ran past the chunk,
    // with the same signature as the OR miscompile opq() guards against.)
    int U = 1;
    for (int d = 1; d <= std::min(u_cap, CR); ++d)
        if (CR % d == 0) U = d;
    // Two words per lane (16-byte loads and stores) where nothing crosses
    // lanes: no carries, no in-kernel advances, no loads at a shifted offset.
    // A lane then owns words 2l, 2l+1 of a pair of rows.
    // Only where a row pair's loads (4 VGPRs each) fit the register budget.
    const bool wide = opt.words_per_lane >= 2 && !has_state && ext_shifts.empty() && CR % 2 == 0 &&
                      4 * loads_per_row + 2 * live_ <= kernel_registers(opt) - 48 - NPOP_;
    if (wide) {
        const int pairs = CR / 2;
        const int cap2 = std::max(1, std::min(opt.rows_per_step, load_budget / std::max(1, 4 * loads_per_row)));
        U = 1;
        for (int d = 1; d <= std::min(cap2, pairs); ++d)
            if (pairs % d == 0) U = d;
    }
    ks.words_per_lane = wide ? 2 : 1;
    ks.rows_per_step = wide ? 2 * U : U;
    // On request, a stateful kernel whose step covers the whole chunk and
    // whose loads for the warm-up rows and the chunk fit together issues
    // them in one step, so the chunk's loads do not wait for the warm-up's.
    const bool joined = opt.join_warmup && has_state && !opt.interleave && U == CR && WR + CR <= u_cap &&
                        !opt.lazy_loads;
    ks.chunk_rows = CR;

    std::ostringstream s;
    const int NIN = (int)ks.inputs.size(), NOUT = (int)ks.outputs.size(), NPOP = (int)ks.pops.size();
    const int NPRED = (int)ks.preds.size(), NINT = NIN + NPRED;  // input table: streams, then predicates
    // Register budget: at least `occupancy` waves per SIMD (512/occupancy VGPRs),
    // which also bounds how many row loads the scheduler hoists.
    // estimated VGPRs: prefetched words, popcount accumulators, live values
    const int est_vgprs = (wide ? 4 : 2) * loads_per_row * U + NPOP + (wide ? 4 : 2) * live_ + 24;
    // Automatic occupancy never goes below 2 waves per SIMD (256 registers),
    // so an untuned kernel cannot spill into AGPRs; occupancy 1 is for the
    // autotuner, which checks the compiled register use.
    int occ = opt.occupancy;
    if (occ <= 0) {
        occ = 4;
        while (occ > 2 && est_vgprs > 512 / occ) --occ;
    }
    ks.occupancy = occ;
    const int WPB = opt.waves_per_block;
    const int LB = 64 * WPB;
    ks.launch_bounds = LB;
    if (!emit) return;  // shape only (register budget, rows per step), no source
    if (wide) {
        s << "extern \"C\" __global__ __launch_bounds__(" << LB << ") __attribute__((amdgpu_waves_per_eu(" << occ
          << "))) void bsz_kernel(KArgs A) {\n";
        s << "  const u32 lane = lane_id();\n";
        s << "  const u32 off16 = lane * 16u;\n";
        s << "  const u64x2 zero2 = {0ull, 0ull};\n";
        emit_chunk_loop(s, opt, WPB);
        emit_run_prologue(s, NPOP, false);
        if (NPOP) {
            s << "  {\n    u64* const other = A.pop_acc + (ep ? 0ull : A.pop_stride);\n";
            s << "    const u64 n_slots = ((u64)A.n_inst * " << NPOP << "ull) << A.spread_log;\n";
            s << "    for (u64 i = (u64)c * 64ull + lane; i < n_slots; i += (u64)A.n_chunks * 64ull) other[i] = 0ull;\n  }\n";
        }
        s << "  const u32 ii = A.chunk_inst ? u32_at(A.chunk_inst, c) : 0u;\n";
        s << "  const InstDesc D = inst_at(A.inst, ii);\n";
        s << "  const u32 cc = c - D.chunk_begin;\n";
        s << "  const u64 NW = D.rows * 64ull;\n";
        for (int j = 0; j < NIN; ++j) s << "  const u64* __restrict__ in" << j << " = tab_in(A.in, (u64)ii * " << NINT << "u + " << j << "u);\n";
        for (int j = 0; j < NOUT; ++j) s << "  u64* __restrict__ out" << j << " = tab_out(A.out, (u64)ii * " << NOUT << "u + " << j << "u);\n";
        for (int j = 0; j < NPRED; ++j)
            s << "  const bool pf" << j << " = pred_flag(tab_in(A.in, (u64)ii * " << NINT << "u + " << (NIN + j)
              << "u), A.pop_stride, ep, A.spread_log);\n";
        s << "  const u64 r_begin = (u64)cc * " << CR << "ull;\n";
        s << "  const u64 r_end = (r_begin + " << CR << "ull < D.rows) ? r_begin + " << CR << "ull : D.rows;\n";
        for (int p = 0; p < NPOP; ++p) s << "  u32 pc" << p << " = 0u;\n";
        s << "  for (u64 r = r_begin; r < r_end; r += " << 2 * U << "ull) {\n";
        for (int u = 0; u < U; ++u) {
            // rows ru, ru+1; the second may not exist at the end of a stream
            s << "    const u64 ru" << u << " = (u64)uni((u32)(r + " << 2 * u << "ull));\n";
            s << "    const long long w" << u << " = (long long)(ru" << u << " * 64ull + 2u * lane);\n";
            s << "    const bool ok" << u << " = r + " << 2 * u << "ull < r_end;\n";
            s << "    const bool lv" << u << " = ok" << u << " && (u64)w" << u << " < NW;\n";
            for (int32_t a : plain_inputs)
                s << "    const u64x2 x" << a << "_" << u << " = lv" << u << " ? lane_ld16(urow(in" << input_index.at(a)
                  << " + ru" << u << " * 64ull), off16) : zero2;\n";
        }
        std::map<int32_t, std::vector<int>> out_of, pop_of;
        for (int j = 0; j < NOUT; ++j) out_of[ks.outputs[j]].push_back(j);
        for (int p = 0; p < NPOP; ++p) pop_of[g.pops[ks.pops[p]]].push_back(p);
        for (int u = 0; u < U; ++u) {
            s << "    if (ok" << u << ") {\n";
            // m: the words an op result may occupy; mb: the bitlength (popcounts)
            s << "      const u64x2 m = {word_mask(w" << u << ", D.opw, D.opmask), word_mask(w" << u << " + 1, D.opw, D.opmask)};\n";
            if (NPOP)
                s << "      const u64x2 mb = {word_mask(w" << u << ", D.lw, D.lastmask), word_mask(w" << u
                  << " + 1, D.lw, D.lastmask)};\n";
            auto ref = [&](int32_t a) -> std::string {
                const GNode& x = g.nodes[a];
                if (in_kernel.count(a)) return "v" + std::to_string(a);
                if (x.op == NOp::Zero) return "zero2";
                if (x.op == NOp::Ones) return "m";
                return "x" + std::to_string(a) + "_" + std::to_string(u);
            };
            auto sink = [&](int32_t n) {
                auto oi = out_of.find(n);
                auto pi = pop_of.find(n);
                if (oi == out_of.end() && pi == pop_of.end()) return;
                s << "      {\n        const u64x2 t = " << ref(n) << " & m;\n";
                if (oi != out_of.end())
                    for (int j : oi->second)
                        s << "        if (lv" << u << ") lane_st16(urow_w(out" << j << " + ru" << u << " * 64ull), off16, t);\n";
                if (pi != pop_of.end())
                    for (int p : pi->second)
                        s << "        pc" << p << " += (u32)(__builtin_popcountll(t.x & mb.x) + __builtin_popcountll(t.y & mb.y));\n";
                s << "      }\n";
                if (opt.store_order) s << "      __builtin_amdgcn_sched_barrier(0);\n";
            };
            for (int32_t n : ks.nodes) {
                const GNode& x = g.nodes[n];
                const std::string v = "v" + std::to_string(n);
                switch (x.op) {
                    case NOp::Zero: s << "      const u64x2 " << v << " = zero2;\n"; break;
                    case NOp::Ones: s << "      const u64x2 " << v << " = m;\n"; break;
                    case NOp::Not: s << "      const u64x2 " << v << " = ~" << ref(x.a) << ";\n"; break;
                    case NOp::And: s << "      const u64x2 " << v << " = " << ref(x.a) << " & " << ref(x.b) << ";\n"; break;
                    case NOp::Or: s << "      const u64x2 " << v << " = " << ref(x.a) << " | " << ref(x.b) << ";\n"; break;
                    case NOp::Xor: s << "      const u64x2 " << v << " = " << ref(x.a) << " ^ " << ref(x.b) << ";\n"; break;
                    case NOp::Shl: case NOp::Shr:  // only shifts by 0 reach a wide kernel
                        s << "      const u64x2 " << v << " = " << ref(x.a) << ";\n";
                        break;
                    case NOp::Sel:
                        s << "      const u64x2 " << v << " = pf" << pred_index.at(x.pred) << " ? " << ref(x.a) << " : "
                          << ref(x.b) << ";\n";
                        break;
                    default: throw std::logic_error("wide kernel with a cross-lane op");
                }
                sink(n);
            }
            std::set<int32_t> done(ks.nodes.begin(), ks.nodes.end()), rest;
            for (auto& kv : out_of) if (!done.count(kv.first)) rest.insert(kv.first);
            for (auto& kv : pop_of) if (!done.count(kv.first)) rest.insert(kv.first);
            for (int32_t n : rest) sink(n);
            s << "    }\n";
        }
        s << "  }\n";
        for (int p = 0; p < NPOP; ++p) {
            s << "  {\n    const u32 tot = uni(wave_sum(pc" << p << "));\n";
            s << "    if (lane == 0) __hip_atomic_fetch_add(&bank[(((u64)ii * " << NPOP << "u + " << p
              << "u) << A.spread_log) + (c & ((1u << A.spread_log) - 1u))], (u64)tot, __ATOMIC_RELAXED, AGENT);\n  }\n";
        }
        emit_run_epilogue(s);
        s << "  }\n";  // chunk
        s << "}\n";
        ks.source = s.str();
        return;
    }
    const std::string NS = std::to_string(n_state) + "u";
    const std::string NS2 = std::to_string(2 * n_state) + "u";
    s << "extern \"C\" __global__ __launch_bounds__(" << LB << ") __attribute__((amdgpu_waves_per_eu(" << occ
      << "))) void bsz_kernel(KArgs A) {\n";
    s << "  const u32 lane = lane_id();\n";
    s << "  const u32 off8 = lane * 8u;\n";
    emit_chunk_loop(s, opt, WPB);
    emit_run_prologue(s, NPOP, has_state);
    if (NPOP) {
        s << "  {\n    u64* const other = A.pop_acc + (ep ? 0ull : A.pop_stride);\n";
        s << "    const u64 n_slots = ((u64)A.n_inst * " << NPOP << "ull) << A.spread_log;\n";
        s << "    for (u64 i = (u64)c * 64ull + lane; i < n_slots; i += (u64)A.n_chunks * 64ull) other[i] = 0ull;\n  }\n";
    }
    for (auto& kv : state_of) s << "  u64 st" << kv.second << " = 0ull;\n";
    s << "  const u32 ii = A.chunk_inst ? u32_at(A.chunk_inst, c) : 0u;\n";
    s << "  const InstDesc D = inst_at(A.inst, ii);\n";
    s << "  const u32 cc = c - D.chunk_begin;\n";
    s << "  const u64 NW = D.rows * 64ull;\n";
    for (int j = 0; j < NIN; ++j) s << "  const u64* __restrict__ in" << j << " = tab_in(A.in, (u64)ii * " << NINT << "u + " << j << "u);\n";
    for (int j = 0; j < NOUT; ++j) s << "  u64* __restrict__ out" << j << " = tab_out(A.out, (u64)ii * " << NOUT << "u + " << j << "u);\n";
    for (int j = 0; j < NPRED; ++j)
        s << "  const bool pf" << j << " = pred_flag(tab_in(A.in, (u64)ii * " << NINT << "u + " << (NIN + j)
          << "u), A.pop_stride, ep, A.spread_log);\n";
    s << "  const u64 r_begin = (u64)cc * " << CR << "ull;\n";
    s << "  const u64 r_end = (r_begin + " << CR << "ull < D.rows) ? r_begin + " << CR << "ull : D.rows;\n";
    // Interleaved rows (kernels without state only): chunk cc of an instance
    // with n chunks takes rows cc, cc + n, cc + 2n, ..., so the waves running
    // at one time read one contiguous window of every stream rather than one
    // distant segment each.
    const bool inter = opt.interleave && !has_state;
    ks.interleave = inter;
    std::map<int32_t, std::vector<int>> out_of, pop_of;
    for (int j = 0; j < NOUT; ++j) out_of[ks.outputs[j]].push_back(j);
    for (int p = 0; p < NPOP; ++p) pop_of[g.pops[ks.pops[p]]].push_back(p);

    // Three-input boolean functions.  A NOT/AND/OR/XOR node absorbs the
    // boolean nodes of this kernel that feed only it (no other consumer, not
    // stored, not counted) while its leaves stay within three values, and is
    // emitted as one bop3 (two v_bitop3_b32) over the leaves with the truth
    // table of the absorbed expression.  The absorbed nodes are still
    // emitted; their values are then unused and the compiler drops them.
    std::map<int32_t, std::pair<std::vector<int32_t>, unsigned>> fused3;   // node -> (leaves, table)
    if (opt.bitop3) {
        auto is_bool = [&](NOp op) { return op == NOp::Not || op == NOp::And || op == NOp::Or || op == NOp::Xor; };
        std::map<int32_t, int> kuse;
        for (int32_t n : ks.nodes) {
            const GNode& x = g.nodes[n];
            if (arity(x.op) >= 1 && in_kernel.count(x.a)) ++kuse[x.a];
            if (arity(x.op) >= 2 && in_kernel.count(x.b)) ++kuse[x.b];
        }
        auto absorbable = [&](int32_t m) {
            return in_kernel.count(m) && is_bool(g.nodes[m].op) && kuse[m] == 1 && !out_of.count(m) && !pop_of.count(m);
        };
        auto operands = [&](int32_t m) {
            std::vector<int32_t> v;
            const GNode& x = g.nodes[m];
            if (arity(x.op) >= 1 && g.nodes[x.a].op != NOp::Zero) v.push_back(x.a);
            if (arity(x.op) >= 2 && g.nodes[x.b].op != NOp::Zero) v.push_back(x.b);
            return v;
        };
        for (int32_t n : ks.nodes) {
            if (!is_bool(g.nodes[n].op)) continue;
            std::set<int32_t> inl{n};
            std::vector<int32_t> leaves = operands(n);
            std::sort(leaves.begin(), leaves.end());
            leaves.erase(std::unique(leaves.begin(), leaves.end()), leaves.end());
            for (bool grew = true; grew;) {
                grew = false;
                for (int32_t l : leaves) {
                    if (!absorbable(l)) continue;
                    std::vector<int32_t> nl;
                    for (int32_t y : leaves) if (y != l) nl.push_back(y);
                    for (int32_t y : operands(l)) nl.push_back(y);
                    std::sort(nl.begin(), nl.end());
                    nl.erase(std::unique(nl.begin(), nl.end()), nl.end());
                    if (nl.size() > 3) continue;
                    inl.insert(l);
                    leaves = nl;
                    grew = true;
                    break;
                }
            }
            if (inl.size() < 2) continue;
            // truth table: leaf j takes bit (2 - j) of the index
            unsigned table = 0;
            for (unsigned idx = 0; idx < 8; ++idx) {
                std::function<bool(int32_t)> ev = [&](int32_t m) -> bool {
                    for (size_t j = 0; j < leaves.size(); ++j)
                        if (leaves[j] == m) return (idx >> (2 - j)) & 1u;
                    const GNode& x = g.nodes[m];
                    if (x.op == NOp::Zero) return false;
                    switch (x.op) {
                        case NOp::Not: return !ev(x.a);
                        case NOp::And: return ev(x.a) && ev(x.b);
                        case NOp::Or: return ev(x.a) || ev(x.b);
                        case NOp::Xor: return ev(x.a) != ev(x.b);
                        default: throw std::logic_error("internal: bop3 over a non-boolean node");
                    }
                };
                if (ev(n)) table |= 1u << idx;
            }
            // a constant function would hand the compiler a known value
            if (table == 0u || table == 0xffu) continue;
            fused3[n] = {leaves, table};
        }
    }

    // Loads of U rows starting at row r (row u: `r + u`, or `r + u * istride`
    // when interleaved); row u exists while its index is below `limit`.
    // Row u's words of the kernel's inputs.  `guard`: the row may not exist
    // (the word is 0 then); inside a row's `if (ok)` block it does.
    auto emit_x = [&](int u, int32_t a, bool guard) {
        const std::string rb = "urow(in" + std::to_string(input_index.at(a)) + " + ru" + std::to_string(u) + " * 64ull)";
        s << "      const u64 x" << a << "_" << u << " = ";
        if (guard) s << "ok" << u << " ? lane_ld(" << rb << ", off8) : 0ull;\n";
        else s << "lane_ld(" << rb << ", off8);\n";
    };
    auto emit_z = [&](int u, const std::pair<int32_t, long long>& ow, bool guard) {
        const std::string d = ow.second < 0 ? " - " + std::to_string(-ow.second) + "ll" : " + " + std::to_string(ow.second) + "ll";
        s << "      const u64 z" << ow.first << "_" << (ow.second < 0 ? "m" : "p") << std::llabs(ow.second) << "_" << u << " = ";
        const std::string e = "ldr(urow(in" + std::to_string(input_index.at(ow.first)) + " + ru" + std::to_string(u) +
                              " * 64ull), (long long)lane" + d + ", w" + std::to_string(u) + d + ", NW)";
        if (guard) s << "ok" << u << " ? " << e << " : 0ull;\n";
        else s << e << ";\n";
    };
    // the words an external shift of node n reads, as (input, word offset)
    auto shift_words = [&](int32_t n) {
        std::vector<std::pair<int32_t, long long>> out;
        const GNode& x = g.nodes[n];
        const GNode& src = g.nodes[x.a];
        if (src.op == NOp::Zero || src.op == NOp::Ones) return out;
        const long long q = (long long)(x.k / 64);
        const bool part = (x.k % 64) != 0;
        if (x.op == NOp::Shl) {
            out.push_back({x.a, -q});
            if (part) out.push_back({x.a, -q - 1});
        } else {
            out.push_back({x.a, q});
            if (part) out.push_back({x.a, q + 1});
        }
        return out;
    };
    auto emit_y = [&](int u, int32_t n, bool guard) {
        const GNode& x = g.nodes[n];
        const long long q = (long long)(x.k / 64);
        const unsigned b = (unsigned)(x.k % 64);
        const GNode& src = g.nodes[x.a];
        auto word = [&](long long d) -> std::string {
            if (src.op == NOp::Zero) return "0ull";
            std::string wd = "w" + std::to_string(u) + (d < 0 ? " - " + std::to_string(-d) + "ll" : " + " + std::to_string(d) + "ll");
            if (src.op == NOp::Ones) return "word_mask(" + wd + ", D.opw, D.opmask)";
            if (d == 0) return "x" + std::to_string(x.a) + "_" + std::to_string(u);
            return "z" + std::to_string(x.a) + "_" + (d < 0 ? "m" : "p") + std::to_string(std::llabs(d)) + "_" + std::to_string(u);
        };
        std::string e = x.op == NOp::Shl
            ? "fsl(" + word(-q) + ", " + (b ? word(-q - 1) : std::string("0ull")) + ", " + std::to_string(b) + "u)"
            : "fsr(" + word(q) + ", " + (b ? word(q + 1) : std::string("0ull")) + ", " + std::to_string(b) + "u)";
        if (guard) s << "      const u64 y" << n << "_" << u << " = ok" << u << " ? opq(" << e << ") : 0ull;\n";
        else s << "      const u64 y" << n << "_" << u << " = opq(" << e << ");\n";
    };

    // Row indices of U rows starting at row r (row u: `r + u`, or
    // `r + u * istride` when interleaved; row u exists while its index is
    // below `limit`), and, unless loads are lazy, every input word the rows
    // need, issued together.
    auto emit_loads = [&](int UU, const std::string& limit) {
        for (int u = 0; u < UU; ++u) {
            // The row index goes through readfirstlane so that addresses are a
            // scalar row base plus the lane offset (no per-lane pointer per stream).
            const std::string rexpr = inter ? "r + " + std::to_string(u) + "ull * istride" : "r + " + std::to_string(u) + "ull";
            s << "      const u64 ru" << u << " = (u64)uni((u32)(" << rexpr << "));\n";
            s << "      const long long w" << u << " = (long long)(ru" << u << " * 64ull + lane);\n";
            if (UU > 1) s << "      const bool ok" << u << " = " << rexpr << " < " << limit << ";\n";
            else s << "      const bool ok" << u << " = true;\n";
            if (opt.lazy_loads) continue;
            for (int32_t a : plain_inputs) emit_x(u, a, UU > 1);
            for (auto& ow : off_words) emit_z(u, ow, true);
            for (int32_t n : ext_shifts) emit_y(u, n, true);
        }
    };

    // The nodes of row u.  warm = true: a warm-up row before the chunk; it
    // stores nothing and tracks, per value, the lowest lane from which the
    // value is certainly exact (L, 0..64) and, per state word, whether it is
    // exact (xs): an inexact start state damages at most the low lanes of a
    // row, and an advance's state (the operand's lane 63) or a carry out of
    // a lane that neither propagates nor lies in the damaged range is exact.
    auto emit_compute = [&](int u, bool warm) {
        s << "      if (ok" << u << ") {\n";
        // m: the words an op result may occupy; mb: the bitlength (popcounts)
        s << "        const u64 m = word_mask(w" << u << ", D.opw, D.opmask);\n";
        if (NPOP && !warm) s << "        const u64 mb = word_mask(w" << u << ", D.lw, D.lastmask);\n";
        auto ref = [&](int32_t a) -> std::string {
            const GNode& x = g.nodes[a];
            if (in_kernel.count(a)) return "v" + std::to_string(a);
            if (x.op == NOp::Zero) return "0ull";
            if (x.op == NOp::Ones) return "m";
            return "x" + std::to_string(a) + "_" + std::to_string(u);
        };
        auto lref = [&](int32_t a) -> std::string { return in_kernel.count(a) ? "L" + std::to_string(a) : "0u"; };
        // lazy loads: each input word right before its first use in the row
        std::set<int32_t> have_x, have_y;
        std::set<std::pair<int32_t, long long>> have_z;
        auto need_x = [&](int32_t a) {
            if (!opt.lazy_loads || !plain_inputs.count(a) || !have_x.insert(a).second) return;
            emit_x(u, a, false);
        };
        auto need = [&](int32_t n) {
            if (!opt.lazy_loads) return;
            const GNode& x = g.nodes[n];
            if ((x.op == NOp::Shl || x.op == NOp::Shr) && !in_kernel.count(x.a)) {
                if (x.k == 0) { need_x(x.a); return; }
                if (!have_y.insert(n).second) return;
                for (auto& wd : shift_words(n)) {
                    if (wd.second == 0) need_x(wd.first);
                    else if (have_z.insert(wd).second) emit_z(u, wd, false);
                }
                emit_y(u, n, false);
                return;
            }
            if (x.op == NOp::Input) { need_x(n); return; }
            if (arity(x.op) >= 1 && !in_kernel.count(x.a)) need_x(x.a);
            if (arity(x.op) >= 2 && !in_kernel.count(x.b)) need_x(x.b);
        };
        auto lmax = [&](const std::string& x, const std::string& y) -> std::string {
            if (x == "0u") return y;
            if (y == "0u" || x == y) return x;
            return "(" + x + " > " + y + " ? " + x + " : " + y + ")";
        };
        // outputs and popcounts of each node are emitted right after it, so
        // its register is free again at once
        auto sink = [&](int32_t n) {
            if (warm) return;
            auto oi = out_of.find(n);
            auto pi = pop_of.find(n);
            if (oi == out_of.end() && pi == pop_of.end()) return;
            if (oi != out_of.end())
                for (int j : oi->second)
                    s << "        lane_st(urow_w(out" << j << " + ru" << u << " * 64ull), off8, " << ref(n) << " & m);\n";
            if (pi != pop_of.end())
                for (int p : pi->second) s << "        pc" << p << " += (u32)__builtin_popcountll(" << ref(n) << " & mb);\n";
            // without it the scheduler may sink every store to the end of the
            // step and keep all the outputs of a row live at once
            if (opt.store_order) s << "        __builtin_amdgcn_sched_barrier(0);\n";
        };
        for (int32_t n : ks.nodes) {
            const GNode& x = g.nodes[n];
            need(n);
            const std::string v = "v" + std::to_string(n), L = "L" + std::to_string(n);
            switch (x.op) {
                case NOp::Zero: s << "        const u64 " << v << " = 0ull;\n"; break;
                case NOp::Ones: s << "        const u64 " << v << " = m;\n"; break;
                case NOp::Not: case NOp::And: case NOp::Or: case NOp::Xor: {
                    auto f3 = fused3.find(n);
                    if (f3 != fused3.end()) {
                        const std::vector<int32_t>& lv = f3->second.first;
                        s << "        const u64 " << v << " = bop3<" << f3->second.second << "u>(";
                        for (size_t j = 0; j < 3; ++j) {
                            if (j < lv.size()) need_x(lv[j]);
                            s << (j ? ", " : "") << (j < lv.size() ? ref(lv[j]) : std::string("0ull"));
                        }
                        s << ");\n";
                    } else if (x.op == NOp::Not) {
                        s << "        const u64 " << v << " = ~" << ref(x.a) << ";\n";
                    } else {
                        const char* o = x.op == NOp::And ? " & " : x.op == NOp::Or ? " | " : " ^ ";
                        s << "        const u64 " << v << " = " << ref(x.a) << o << ref(x.b) << ";\n";
                    }
                    break;
                }
                case NOp::Add: {
                    const std::string st = "st" + std::to_string(state_of.at(n)), t = "_" + std::to_string(n);
                    s << "        const u64 a" << t << " = " << ref(x.a) << ", s" << t << " = a" << t << " + " << ref(x.b) << ";\n";
                    s << "        const u64 G" << t << " = ballot(s" << t << " < a" << t << ");\n";
                    s << "        const u64 P" << t << " = G" << t << " | ballot(s" << t << " == ~0ull);\n";
                    s << "        const u64 t" << t << " = G" << t << " + P" << t << ";\n";
                    s << "        const u64 u" << t << " = t" << t << " + " << st << ";\n";
                    s << "        const u64 " << v << " = opq(s" << t << " + (((u" << t << " ^ G" << t << " ^ P" << t
                      << ") >> lane) & 1ull));\n";
                    s << "        " << st << " = (u64)((t" << t << " < G" << t << ") | (u" << t << " < t" << t << "));\n";
                    if (warm && track) {
                        // lanes that do not propagate a carry (they generate
                        // or kill it) at or above the damaged range end it
                        const std::string xs = "xs" + std::to_string(state_of.at(n));
                        const std::string li = lmax(lref(x.a), lref(x.b));
                        s << "        const u64 np" << t << " = (~P" << t << " | G" << t << ") & (" << li
                          << " >= 64u ? 0ull : (~0ull << " << li << "));\n";
                        s << "        const bool ex" << t << " = " << xs << " != 0u && " << li << " == 0u;\n";
                        s << "        const u32 " << L << " = ex" << t << " ? 0u : (np" << t << " ? (u32)__builtin_ctzll(np"
                          << t << ") + 1u : 64u);\n";
                        s << "        " << xs << " = (ex" << t << " || np" << t << " != 0ull) ? 1u : 0u;\n";
                    }
                    break;
                }
                case NOp::Shl: case NOp::Shr: {
                    if (!in_kernel.count(x.a)) {
                        if (x.k == 0) s << "        const u64 " << v << " = " << ref(x.a) << ";\n";
                        else s << "        const u64 " << v << " = y" << n << "_" << u << ";\n";
                    } else if (x.k == 0) {
                        s << "        const u64 " << v << " = " << ref(x.a) << ";\n";
                    } else {
                        const std::string st = "st" + std::to_string(state_of.at(n));
                        s << "        const u64 " << v << " = opq(fsl(" << ref(x.a) << ", prev_word(" << ref(x.a) << ", " << st
                          << "), " << x.k << "u));\n";
                        s << "        " << st << " = lane63(" << ref(x.a) << ");\n";
                        if (warm && track) {
                            // lane 0 takes the state; lane i > 0 takes operand lanes i and i-1
                            const std::string xs = "xs" + std::to_string(state_of.at(n)), la = lref(x.a);
                            s << "        const u32 " << L << " = " << la << " == 0u ? (" << xs << " ? 0u : 1u) : (" << la
                              << " < 64u ? " << la << " + 1u : 64u);\n";
                            s << "        " << xs << " = " << la << " <= 63u ? 1u : 0u;\n";
                        }
                    }
                    break;
                }
                case NOp::Sel:
                    s << "        const u64 " << v << " = pf" << pred_index.at(x.pred) << " ? " << ref(x.a) << " : "
                      << ref(x.b) << ";\n";
                    break;
                case NOp::Input: break;
            }
            if (warm && track) {
                switch (x.op) {
                    case NOp::Zero: case NOp::Ones: s << "        const u32 " << L << " = 0u;\n"; break;
                    case NOp::Not: s << "        const u32 " << L << " = " << lref(x.a) << ";\n"; break;
                    case NOp::And: case NOp::Or: case NOp::Xor: case NOp::Sel:
                        s << "        const u32 " << L << " = " << lmax(lref(x.a), lref(x.b)) << ";\n";
                        break;
                    case NOp::Shl: case NOp::Shr:
                        if (!in_kernel.count(x.a)) s << "        const u32 " << L << " = 0u;\n";
                        else if (x.k == 0) s << "        const u32 " << L << " = " << lref(x.a) << ";\n";
                        break;
                    default: break;
                }
            }
            sink(n);
        }
        // outputs / popcounts of values this kernel does not compute (inputs, constants)
        if (!warm) {
            std::set<int32_t> done(ks.nodes.begin(), ks.nodes.end());
            std::set<int32_t> rest;
            for (auto& kv : out_of) if (!done.count(kv.first)) rest.insert(kv.first);
            for (auto& kv : pop_of) if (!done.count(kv.first)) rest.insert(kv.first);
            for (int32_t n : rest) {
                need(n);
                sink(n);
            }
        }
        s << "      }\n";
    };

    // The state words as this chunk publishes them: in group g (state words
    // 32g .. 32g+31), lane 2j holds the low and lane 2j+1 the high half of
    // state word 32g+j, each under the run's tag in the upper 32 bits, so one
    // 64-bit store per lane publishes a word half with the run it belongs to
    // and a reader needs no fence.
    const int NG = (n_state + 31) / 32;
    auto group_lanes = [&](int g) { return std::to_string(std::min(64, 2 * n_state - 64 * g)) + "u"; };
    auto emit_tagged = [&](const std::string& var) {
        for (int g = 0; g < NG; ++g) s << "    u32 " << var << "_v" << g << " = 0u;\n";
        for (auto& kv : state_of) {
            const int i = kv.second, g = i / 32, j = i % 32;
            const std::string vv = var + "_v" + std::to_string(g);
            s << "    " << vv << " = lane == " << 2 * j << "u ? (u32)st" << i << " : " << vv << ";\n";
            s << "    " << vv << " = lane == " << 2 * j + 1 << "u ? (u32)(st" << i << " >> 32) : " << vv << ";\n";
        }
        for (int g = 0; g < NG; ++g)
            s << "    const u64 " << var << g << " = ((u64)tag << 32) | " << var << "_v" << g << ";\n";
    };
    // Look-back (an unproven chunk, after its rows ran from the speculative
    // state): wait for the predecessor's end state of this run; when it
    // differs from the assumed one, load it into the state words and leave
    // the emitted code inside `if (differs) {`, where the caller runs the
    // chunk's rows again.
    auto emit_lookback = [&]() {
        s << "      if (lane == 0u) __hip_atomic_fetch_add(&kst[1], 1u, __ATOMIC_RELAXED, AGENT);\n";
        s << "      const u64* const prev = A.st_t + (u64)(c - 1u) * " << NS2 << ";\n";
        for (int g = 0; g < NG; ++g) s << "      u64 w" << g << " = spec" << g << ";\n";
        s << "      for (u32 it = 0u;; ++it) {\n";
        s << "        bool stale = false;\n";
        for (int g = 0; g < NG; ++g) {
            s << "        const u64 x" << g << " = lane < " << group_lanes(g) << " ? aload64((u64*)prev + " << 64 * g
              << "u + lane) : ((u64)tag << 32);\n";
            s << "        stale |= ballot((u32)(x" << g << " >> 32) != tag) != 0ull;\n";
        }
        s << "        if (!stale) {\n";
        for (int g = 0; g < NG; ++g) s << "          w" << g << " = x" << g << ";\n";
        s << "          break;\n";
        s << "        }\n";
        s << "        if (it >= A.spin_limit) {\n";
        s << "          if (lane == 0u) {\n";
        s << "            __hip_atomic_fetch_or(A.err, 1u, __ATOMIC_RELAXED, AGENT);\n";
        s << "            __hip_atomic_fetch_add(&kst[2], 1u, __ATOMIC_RELAXED, AGENT);\n";
        s << "          }\n";
        s << "          break;\n";
        s << "        }\n";
        s << "        __builtin_amdgcn_s_sleep(2);\n";
        s << "      }\n";
        s << "      bool differs = false;\n";
        for (int g = 0; g < NG; ++g) s << "      differs |= ballot(w" << g << " != spec" << g << ") != 0ull;\n";
        s << "      if (differs) {\n";
        s << "        if (lane == 0u) __hip_atomic_fetch_add(&kst[0], 1u, __ATOMIC_RELAXED, AGENT);\n";
        for (auto& kv : state_of) {
            const int i = kv.second, g = i / 32, j = i % 32;
            s << "        st" << i << " = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)w" << g << ", " << 2 * j + 1
              << ") << 32) | (u64)(u32)__builtin_amdgcn_readlane((int)(u32)w" << g << ", " << 2 * j << ");\n";
        }
    };
    // the end state, final: this chunk's successor may be waiting for it
    auto emit_publish = [&]() {
        emit_tagged("fin");
        for (int g = 0; g < NG; ++g)
            s << "    if (lane < " << group_lanes(g) << ") astore64(&A.st_t[(u64)c * " << NS2 << " + " << 64 * g
              << "u + lane], fin" << g << ");\n";
    };
    if (joined) {
        // One step: rows r_begin - WR .. r_begin + CR - 1, loads issued
        // together.  Warm-up rows before the instance's first row (or any,
        // for its first chunk) do not exist; a warm-up that starts at the
        // instance's first row reaches the true state.
        s << "    u32 proven = 1u;\n";
        if (track)
            for (auto& kv : state_of) s << "    u32 xs" << kv.second << " = 0u;\n";
        for (int j = 0; j < WR + CR; ++j) {
            const std::string row = "(long long)r_begin - " + std::to_string(WR - j) + "ll";
            if (j < WR)
                s << "      const bool ok" << j << " = cc > 0u && " << row << " >= 0ll;\n";
            else
                s << "      const bool ok" << j << " = r_begin + " << (j - WR) << "ull < r_end;\n";
            s << "      const u64 ru" << j << " = ok" << j << " ? (u64)uni((u32)(" << row << ")) : 0ull;\n";
            s << "      const long long w" << j << " = (long long)(ru" << j << " * 64ull + lane);\n";
            for (int32_t a : plain_inputs) emit_x(j, a, true);
            for (auto& ow : off_words) emit_z(j, ow, true);
            for (int32_t n : ext_shifts) emit_y(j, n, true);
        }
        for (int j = 0; j < WR; ++j) emit_compute(j, true);
        s << "    proven = (cc == 0u || r_begin <= " << WR << "ull) ? 1u : " << (track ? "1u" : "0u") << ";\n";
        if (track) {
            s << "    if (cc > 0u && r_begin > " << WR << "ull) {\n";
            for (auto& kv : state_of) s << "      proven &= xs" << kv.second << ";\n";
            s << "    }\n";
        }
        emit_tagged("spec");
        for (int p = 0; p < NPOP; ++p) s << "    u32 pc" << p << " = 0u;\n";
        for (int j = WR; j < WR + CR; ++j) emit_compute(j, false);
        // an unproven chunk looks back; it runs its rows again from its
        // predecessor's end state when that differs from the one assumed
        s << "    if (!proven) {\n";
        emit_lookback();
        for (int p = 0; p < NPOP; ++p) s << "        pc" << p << " = 0u;\n";
        s << "    for (u64 r = r_begin; r < r_end; r += " << U << "ull) {\n";
        emit_loads(U, "r_end");
        for (int u = 0; u < U; ++u) emit_compute(u, false);
        s << "    }\n";  // row loop
        s << "      }\n";  // differs
        s << "    }\n";  // !proven
        emit_publish();
    } else if (has_state) {
        // Warm-up (chunks after an instance's first): the rows before the
        // chunk, from a zero state, storing nothing.  A warm-up that starts at
        // the instance's first row reaches the true state; otherwise the
        // tracking below may prove the state it reaches exact.
        s << "    const u64 warm = cc > 0u ? (r_begin < " << WR << "ull ? r_begin : " << WR << "ull) : 0ull;\n";
        s << "    u32 proven = 1u;\n";
        s << "    if (warm) {\n";
        if (track)
            for (auto& kv : state_of) s << "      u32 xs" << kv.second << " = 0u;\n";
        s << "      for (u64 r = r_begin - warm; r < r_begin; r += " << U << "ull) {\n";
        emit_loads(U, "r_begin");
        for (int u = 0; u < U; ++u) emit_compute(u, true);
        s << "      }\n";
        // a kernel too large to track its lanes proves only warm-ups that
        // start at the instance's first row
        s << "      proven = r_begin <= " << WR << "ull ? 1u : " << (track ? "1u" : "0u") << ";\n";
        if (track) {
            s << "      if (r_begin > " << WR << "ull) {\n";
            for (auto& kv : state_of) s << "        proven &= xs" << kv.second << ";\n";
            s << "      }\n";
        }
        s << "    }\n";
        emit_tagged("spec");
        for (int p = 0; p < NPOP; ++p) s << "    u32 pc" << p << " = 0u;\n";
        // Pass 0 runs the chunk from the speculative state.  An unproven
        // chunk then looks back: it waits for its predecessor's end state of
        // this run (final once published) and, if that differs from the
        // state it assumed, runs the chunk again from it (pass 1).  Chunks
        // are owned by waves in increasing order and the hardware dispatches
        // workgroups in order, so the lowest unfinished chunk never waits on
        // an unfinished one; the wait is bounded all the same.
        s << "    for (u32 pass = 0u;; ++pass) {\n";
        for (int p = 0; p < NPOP; ++p) s << "    pc" << p << " = 0u;\n";
    } else {
        for (int p = 0; p < NPOP; ++p) s << "    u32 pc" << p << " = 0u;\n";
    }
    if (!joined) {
    if (inter) {
        s << "    const u64 istride = (u64)D.n_chunks;\n";
        s << "    for (u64 r = cc; r < D.rows; r += istride * " << U << "ull) {\n";
        emit_loads(U, "D.rows");
    } else {
        s << "    for (u64 r = r_begin; r < r_end; r += " << U << "ull) {\n";
        emit_loads(U, "r_end");
    }
    for (int u = 0; u < U; ++u) emit_compute(u, false);
    s << "    }\n";  // row loop
    if (has_state) {
        s << "    if (pass == 0u && !proven) {\n";
        emit_lookback();
        s << "        continue;\n";
        s << "      }\n";
        s << "    }\n";
        s << "    break;\n";
        s << "    }\n";  // pass loop
        emit_publish();
    }
    }

    // popcounts: one atomic per chunk into slot c mod 2^spread_log
    for (int p = 0; p < NPOP; ++p) {
        s << "    {\n      const u32 tot = uni(wave_sum(pc" << p << "));\n";
        s << "      if (lane == 0) __hip_atomic_fetch_add(&bank[(((u64)ii * " << NPOP << "u + " << p
          << "u) << A.spread_log) + (c & ((1u << A.spread_log) - 1u))], (u64)tot, __ATOMIC_RELAXED, AGENT);\n";
        s << "    }\n";
    }
    emit_run_epilogue(s);
    s << "  }\n";  // chunk
    s << "}\n";
    ks.source = s.str();
}

// The compilation unit of a generated kernel, and its entry point named
// after the unit's hash (bsz_fk_<16 hex digits>), so profiles tell the
// kernels of different plans and schedules apart.
void finish_unit(const PlanOptions& opt, KernelSpec& ks) {
    const std::string prefix =
        opt.nontemporal > 0 ? "#define BSZ_NT " + std::to_string(std::min(2, opt.nontemporal)) + "\n" : std::string();
    const std::string generic = "void bsz_kernel(KArgs A)";
    const uint64_t h = fnv1a(prefix + kKernelPrelude + ks.source);
    char nm[40];
    std::snprintf(nm, sizeof nm, "bsz_fk_%016llx", (unsigned long long)h);
    ks.entry = nm;
    const size_t at = ks.source.find(generic);
    if (at != std::string::npos) ks.source.replace(at, generic.size(), "void " + ks.entry + "(KArgs A)");
    ks.unit = prefix + std::string(kKernelPrelude) + ks.source;
}

}  // namespace

std::string SegGraph::key() const {
    std::ostringstream s;
    s << "S" << n_sources << ";";
    for (const GNode& n : nodes) s << (int)n.op << "," << n.a << "," << n.b << "," << n.k << "," << n.src << "," << n.pred << ";";
    s << "|O";
    for (int32_t o : outputs) s << o << ",";
    s << "|P";
    for (int32_t p : pops) s << p << ",";
    return s.str();
}

std::string PlanOptions::key() const {
    std::ostringstream s;
    s << chunk_rows << "," << warmup_rows << "," << waves_per_block << "," << rows_per_step << "," << max_ops << ","
      << max_live << "," << occupancy << "," << words_per_lane << "," << store_order << "," << nontemporal << ","
      << interleave << "," << persistent << "," << agprs << "," << lazy_loads << "," << join_warmup << "," << bitop3;
    if (order) s << ",o" << order;  // keys of recorded-order plans stay as they were
    return s.str();
}

// Partition order that keeps few values live (option order 1).  A list
// schedule over the needed nodes: among the nodes whose operands are all
// placed, take the one that ends the most live ranges (operands it is the
// last unplaced consumer of) net of the value it starts, the most recently
// readied first on ties, so a value's consumers tend to follow it while it is
// still in the kernel being filled.  Inputs and constants are loaded, not
// held, and are placed first.  Returns a topological order of all N nodes;
// nodes that are not needed come last (the partitioner skips them).
static std::vector<int32_t> pressure_order(const SegGraph& g, const std::vector<char>& needed) {
    const int32_t N = (int32_t)g.nodes.size();
    auto held = [&](int32_t a) { return a >= 0 && needed[a] && !is_free(g.nodes[a].op); };
    auto operands = [&](int32_t n, int32_t (&o)[2]) {
        const GNode& x = g.nodes[n];
        int m = 0;
        if (arity(x.op) >= 1 && held(x.a)) o[m++] = x.a;
        if (arity(x.op) >= 2 && held(x.b) && x.b != x.a) o[m++] = x.b;
        return m;
    };
    // consumers of each node (CSR)
    std::vector<int32_t> cstart(N + 1, 0), cons;
    for (int32_t n = 0; n < N; ++n) {
        if (!held(n)) continue;
        int32_t o[2];
        const int m = operands(n, o);
        for (int i = 0; i < m; ++i) ++cstart[o[i] + 1];
    }
    for (int32_t n = 0; n < N; ++n) cstart[n + 1] += cstart[n];
    cons.resize(cstart[N]);
    {
        std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
        for (int32_t n = 0; n < N; ++n) {
            if (!held(n)) continue;
            int32_t o[2];
            const int m = operands(n, o);
            for (int i = 0; i < m; ++i) cons[fill[o[i]]++] = n;
        }
    }
    std::vector<int32_t> rem(N), waiting(N, 0), ready_at(N, 0);
    for (int32_t n = 0; n < N; ++n) {
        rem[n] = cstart[n + 1] - cstart[n];
        int32_t o[2];
        if (held(n)) waiting[n] = operands(n, o);
    }
    std::vector<char> placed(N, 0);
    std::vector<int32_t> order;
    order.reserve(N);
    for (int32_t n = 0; n < N; ++n)
        if (needed[n] && is_free(g.nodes[n].op)) { order.push_back(n); placed[n] = 1; }
    auto score = [&](int32_t n) {
        int32_t o[2];
        const int m = operands(n, o);
        int s = rem[n] > 0 ? -1 : 0;
        for (int i = 0; i < m; ++i) s += rem[o[i]] == 1;
        return s;
    };
    struct Entry { int score; int32_t ready; int32_t node; };
    auto worse = [](const Entry& a, const Entry& b) {
        if (a.score != b.score) return a.score < b.score;
        if (a.ready != b.ready) return a.ready < b.ready;
        return a.node > b.node;
    };
    std::priority_queue<Entry, std::vector<Entry>, decltype(worse)> pq(worse);
    for (int32_t n = 0; n < N; ++n)
        if (held(n) && waiting[n] == 0) pq.push({score(n), 0, n});
    int32_t t = 0;
    while (!pq.empty()) {
        const Entry e = pq.top();
        pq.pop();
        const int32_t n = e.node;
        if (placed[n] || e.score != score(n)) continue;  // placed, or a stale score (scores only rise)
        placed[n] = 1;
        order.push_back(n);
        ++t;
        int32_t o[2];
        const int m = operands(n, o);
        for (int i = 0; i < m; ++i) {
            if (--rem[o[i]] != 1) continue;
            // the operand's one consumer left now ends its live range
            for (int32_t j = cstart[o[i]]; j < cstart[o[i] + 1]; ++j) {
                const int32_t c = cons[j];
                if (!placed[c] && waiting[c] == 0) pq.push({score(c), ready_at[c], c});
            }
        }
        for (int32_t j = cstart[n]; j < cstart[n + 1]; ++j) {
            const int32_t c = cons[j];
            if (--waiting[c] == 0) {
                ready_at[c] = t;
                pq.push({score(c), t, c});
            }
        }
    }
    for (int32_t n = 0; n < N; ++n)
        if (!placed[n]) order.push_back(n);
    return order;
}

PlanSpec build_plan(const SegGraph& g, const PlanOptions& opt, bool emit) {
    PlanSpec P;
    P.g = g;
    P.opt = opt;
    const int32_t N = (int32_t)g.nodes.size();

    // ── liveness / dead-code elimination ──
    std::vector<char> needed(N, 0), is_out(N, 0);
    for (int32_t o : g.outputs) { needed[o] = 1; is_out[o] = 1; }
    for (int32_t p : g.pops) needed[p] = 1;
    for (int32_t n = N - 1; n >= 0; --n) {
        if (!needed[n]) continue;
        const GNode& x = g.nodes[n];
        if (arity(x.op) >= 1) needed[x.a] = 1;
        if (arity(x.op) >= 2) needed[x.b] = 1;
    }
    // Partition order.  Without selections it is the recorded order, or with
    // option order 1 the list schedule of pressure_order.  With selections,
    // nodes are stably sorted by predicate level (1 + the level of the
    // popcount a selection reads, the maximum over operands otherwise), so
    // the selections of independent `if`s share kernels instead of each
    // forcing a cut after its predicate's kernel.  Always a topological order.
    std::vector<int32_t> order(N), pos(N);
    for (int32_t n = 0; n < N; ++n) order[n] = n;
    {
        bool has_sel = false;
        for (int32_t n = 0; n < N; ++n) has_sel |= needed[n] && g.nodes[n].op == NOp::Sel;
        if (!has_sel && opt.order == 1) order = pressure_order(g, needed);
        if (has_sel) {
            std::vector<int32_t> level(N, 0);
            for (int32_t n = 0; n < N; ++n) {
                const GNode& x = g.nodes[n];
                int32_t l = 0;
                if (arity(x.op) >= 1) l = std::max(l, level[x.a]);
                if (arity(x.op) >= 2) l = std::max(l, level[x.b]);
                if (x.op == NOp::Sel) l = std::max(l, level[g.pops.at(x.pred)] + 1);
                level[n] = l;
            }
            std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return level[a] < level[b]; });
        }
    }
    for (int32_t i = 0; i < N; ++i) pos[order[i]] = i;
    // last_use in partition positions
    std::vector<int32_t> last_use(N, -1);
    for (int32_t n = 0; n < N; ++n) {
        if (!needed[n]) continue;
        const GNode& x = g.nodes[n];
        if (arity(x.op) >= 1) last_use[x.a] = std::max(last_use[x.a], pos[n]);
        if (arity(x.op) >= 2) last_use[x.b] = std::max(last_use[x.b], pos[n]);
    }
    // register liveness: outputs and popcounts are stored / counted right
    // after their producer, so only consumer nodes keep a value in registers
    const std::vector<int32_t> node_last_use = last_use;
    for (int32_t o : g.outputs) last_use[o] = N;
    for (int32_t p : g.pops) last_use[p] = N;

    // ── greedy linear partitioning ──
    // A kernel holds, per row, its live in-kernel values and every word it
    // loads (each distinct external operand, and each shift of one); all of
    // them count against max_live, which is also capped by the register file
    // at the requested occupancy, so a generated kernel does not spill.
    const int value_budget = std::max(8, std::min(opt.max_live, (kernel_registers(opt) - 64) / 2));
    std::vector<int32_t> kern_of(N, -1);
    std::vector<std::vector<int32_t>> kernels;
    std::vector<int> peak_live;                    // per kernel: most in-kernel values live at once
    std::set<std::pair<int32_t, uint64_t>> loaded;  // (external operand, shift or 0) words of this kernel
    // With lazy loads a loaded word takes a register only from its first use
    // in the kernel to its operand's last use: the positions at which the
    // words counted in `lload` are freed.
    std::priority_queue<int32_t, std::vector<int32_t>, std::greater<int32_t>> lload_free;
    int lload = 0;
    int live = 0, n_state = 0, k_pops = 0;
    std::vector<int> popped(N, 0);  // popcounts taken of each node
    for (int32_t p : g.pops) ++popped[p];
    auto is_const = [&](int32_t a) { return g.nodes[a].op == NOp::Zero || g.nodes[a].op == NOp::Ones; };
    for (int32_t i = 0; i < N; ++i) {
        const int32_t n = order[i];
        if (!needed[n]) continue;
        const GNode& x = g.nodes[n];
        if (is_free(x.op)) continue;
        while (!lload_free.empty() && lload_free.top() < i) {
            lload_free.pop();
            --lload;
        }
        const int K = (int)kernels.size() - 1;
        bool cut = kernels.empty();
        if (!cut) {
            auto inside = [&](int32_t a) { return a >= 0 && kern_of[a] == K; };
            if (needs_external_operand(x) && inside(x.a)) cut = true;
            // a selection reads its predicate from an earlier kernel's slots
            if (x.op == NOp::Sel && kern_of[g.pops.at(x.pred)] == K) cut = true;
            if ((int)kernels.back().size() >= opt.max_ops) cut = true;
            int delta = node_last_use[n] > i ? 1 : 0;
            if (arity(x.op) >= 1 && inside(x.a) && node_last_use[x.a] == i) --delta;
            if (arity(x.op) >= 2 && x.b != x.a && inside(x.b) && node_last_use[x.b] == i) --delta;
            int new_loads = 0;
            const uint64_t sk = (x.op == NOp::Shl || x.op == NOp::Shr) ? (x.op == NOp::Shr ? (1ull << 63) : 0) | x.k : 0;
            if (arity(x.op) >= 1 && !inside(x.a) && !is_const(x.a) && !loaded.count({x.a, sk})) ++new_loads;
            if (arity(x.op) >= 2 && !inside(x.b) && !is_const(x.b) && x.b != x.a && !loaded.count({x.b, 0})) ++new_loads;
            // a popcount keeps a 32-bit counter per lane: half a value
            const int held = opt.lazy_loads ? lload : (int)loaded.size();
            if (live + delta + held + new_loads + (k_pops + popped[n] + 1) / 2 > value_budget)
                cut = true;
            if (n_state >= max_state_words(opt) && (x.op == NOp::Add || (stateful_shift(x) && inside(x.a)))) cut = true;
        }
        if (cut) {
            kernels.emplace_back();
            peak_live.push_back(0);
            live = 0;
            n_state = 0;
            k_pops = 0;
            loaded.clear();
            lload = 0;
            lload_free = {};
        }
        k_pops += popped[n];
        const int KK = (int)kernels.size() - 1;
        auto inside = [&](int32_t a) { return a >= 0 && kern_of[a] == KK; };
        {
            const uint64_t sk = (x.op == NOp::Shl || x.op == NOp::Shr) ? (x.op == NOp::Shr ? (1ull << 63) : 0) | x.k : 0;
            auto load = [&](int32_t a, uint64_t shift) {
                if (!loaded.insert({a, shift}).second || node_last_use[a] <= i) return;
                ++lload;   // held until the operand's last use
                lload_free.push(node_last_use[a]);
            };
            if (arity(x.op) >= 1 && !inside(x.a) && !is_const(x.a)) load(x.a, sk);
            if (arity(x.op) >= 2 && !inside(x.b) && !is_const(x.b)) load(x.b, 0);
        }
        if (arity(x.op) >= 1 && inside(x.a) && node_last_use[x.a] == i) --live;
        if (arity(x.op) >= 2 && x.b != x.a && inside(x.b) && node_last_use[x.b] == i) --live;
        if (node_last_use[n] > i) ++live;
        peak_live.back() = std::max(peak_live.back(), live);
        if (x.op == NOp::Add || (stateful_shift(x) && inside(x.a))) ++n_state;
        kern_of[n] = KK;
        kernels.back().push_back(n);
    }
    // popcounts and outputs of free nodes (inputs / constants) ride on the last kernel
    bool free_work = false;
    for (int32_t p : g.pops) if (kern_of[p] < 0) free_work = true;
    for (int32_t o : g.outputs) if (kern_of[o] < 0 && g.nodes[o].op != NOp::Input) free_work = true;
    if (kernels.empty() && free_work) kernels.emplace_back();

    const int NK = (int)kernels.size();
    P.kernels.resize(NK);
    for (int k = 0; k < NK; ++k) {
        char nm[64];
        std::snprintf(nm, sizeof nm, "bsz_k%d", k);
        P.kernels[k].name = nm;
        P.kernels[k].nodes = kernels[k];
        P.kernels[k].peak_live = k < (int)peak_live.size() ? peak_live[k] : 0;
    }

    // consumers per kernel, for materialization and slot intervals
    std::vector<int> last_kernel_use(N, -1);
    auto use_in = [&](int32_t a, int k) {
        if (a >= 0 && kern_of[a] != k) last_kernel_use[a] = std::max(last_kernel_use[a], k);
    };
    for (int k = 0; k < NK; ++k) {
        for (int32_t n : kernels[k]) {
            const GNode& x = g.nodes[n];
            if (arity(x.op) >= 1) use_in(x.a, k);
            if (arity(x.op) >= 2) use_in(x.b, k);
        }
    }
    std::vector<int> pop_kernel(g.pops.size());
    for (size_t p = 0; p < g.pops.size(); ++p) {
        int32_t n = g.pops[p];
        int k = kern_of[n] >= 0 ? kern_of[n] : NK - 1;
        pop_kernel[p] = k;
        P.kernels[k].pops.push_back((int32_t)p);
        use_in(n, k);
    }
    // outputs of free nodes (constants) are stored by the last kernel
    std::vector<int> out_kernel(N, -1);
    for (int32_t o : g.outputs) {
        if (kern_of[o] >= 0) out_kernel[o] = kern_of[o];
        else if (g.nodes[o].op != NOp::Input) { out_kernel[o] = NK - 1; use_in(o, NK - 1); }
    }

    // ── materialized intermediates and slot reuse (interval colouring) ──
    std::vector<int32_t> slot_of(N, -1);
    std::vector<std::pair<int, int32_t>> free_slots;  // (free after kernel, slot)
    std::vector<int> slot_busy_until;
    int n_slots = 0;
    for (int k = 0; k < NK; ++k) {
        for (int32_t n : kernels[k]) {
            if (is_out[n] || last_kernel_use[n] <= k) continue;  // segment outputs own their buffers
            int32_t slot = -1;
            for (int s = 0; s < n_slots; ++s)
                if (slot_busy_until[s] < k) { slot = s; break; }
            if (slot < 0) { slot = n_slots++; slot_busy_until.push_back(0); }
            slot_busy_until[slot] = last_kernel_use[n];
            slot_of[n] = slot;
        }
    }
    P.n_slots = n_slots;

    // ── kernel interfaces ──
    for (int k = 0; k < NK; ++k) {
        KernelSpec& ks = P.kernels[k];
        std::set<int32_t> ins;
        auto ext = [&](int32_t a) {
            if (a < 0 || kern_of[a] == k) return;
            if (g.nodes[a].op == NOp::Zero || g.nodes[a].op == NOp::Ones) return;
            ins.insert(a);
        };
        for (int32_t n : ks.nodes) {
            const GNode& x = g.nodes[n];
            if (arity(x.op) >= 1) ext(x.a);
            if (arity(x.op) >= 2) ext(x.b);
        }
        for (int32_t p : ks.pops) ext(g.pops[p]);
        {
            std::set<int32_t> pr;
            for (int32_t n : ks.nodes)
                if (g.nodes[n].op == NOp::Sel) pr.insert(g.nodes[n].pred);
            ks.preds.assign(pr.begin(), pr.end());
        }
        for (int32_t a : ins) {
            KInput ki;
            const GNode& x = g.nodes[a];
            if (x.op == NOp::Input) { ki.kind = KInput::Source; ki.id = x.src; }
            else if (is_out[a]) { ki.kind = KInput::SegOut; ki.id = a; }
            else { ki.kind = KInput::Slot; ki.id = slot_of[a]; }
            ks.inputs.push_back(ki);
            ks.input_node.push_back(a);
        }
        for (int32_t n : ks.nodes) {
            if (is_out[n]) { ks.outputs.push_back(n); ks.out_slot.push_back(-1); }
            else if (slot_of[n] >= 0) { ks.outputs.push_back(n); ks.out_slot.push_back(slot_of[n]); }
        }
        for (int32_t o : g.outputs)
            if (out_kernel[o] == k && kern_of[o] < 0) { ks.outputs.push_back(o); ks.out_slot.push_back(-1); }
    }
    P.out_node_kernel.clear();
    for (int32_t o : g.outputs) P.out_node_kernel.push_back(out_kernel[o]);
    // every selection's predicate is complete before its kernel starts
    for (int k = 0; k < NK; ++k)
        for (int32_t p : P.kernels[k].preds)
            if (pop_kernel.at(p) >= k)
                throw std::logic_error("internal: a predicate is produced by the kernel that reads it, or later");

    // One compilation unit per kernel (prelude + kernel), so identical kernels
    // of different plans share a cached code object and units compile in
    // parallel; module_source lists them all for inspection.
    std::ostringstream mod;
    mod << kKernelPrelude;
    for (KernelSpec& ks : P.kernels) {
        gen_kernel(g, opt, ks, emit);
        if (!emit) continue;
        finish_unit(opt, ks);
        mod << "// ---- " << ks.name << " ----\n" << ks.source;
    }
    P.module_source = mod.str();
    return P;
}

bool shrink_kernel(PlanSpec& P, size_t k) {
    KernelSpec& ks = P.kernels.at(k);
    PlanOptions o = P.opt;
    const bool wide = ks.words_per_lane == 2;
    const int U = wide ? ks.rows_per_step / 2 : ks.rows_per_step;  // rows (or row pairs) per step
    // under automatic occupancy the kernel keeps (at most) the occupancy it
    // had, so fewer rows never buy a smaller register budget instead
    if (P.opt.occupancy <= 0) o.occupancy = ks.occupancy;
    if (U > 1) {
        o.rows_per_step = std::max(1, U / 2);
        o.words_per_lane = ks.words_per_lane;
    } else if (wide) {
        o.rows_per_step = 1;
        o.words_per_lane = 1;
    } else if (P.opt.occupancy <= 0 && ks.occupancy > 2) {
        o.rows_per_step = 1;
        o.occupancy = ks.occupancy - 1;
    } else {
        return false;
    }
    gen_kernel(P.g, o, ks, true);
    finish_unit(P.opt, ks);
    return true;
}

uint64_t source_hash(const std::string& s) { return fnv1a(s); }

}  // namespace bsz

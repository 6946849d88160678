// fused.cpp — the fused executor: segment recording, plan cache, hipRTC
// compilation with an on-disk code-object cache, plan launch (per partition
// one kernel and one launch per run, wave w owning chunk w of the batch; a
// chunk whose warm-up leaves its start state unproven looks back at its
// predecessor's end state inside the same launch; optionally captured into
// a hipGraph) and the batched plan interface.
#include "fused.h"

#include "bsdata_io.h"
#include "fused_graph.h"

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <set>
#include <thread>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>
#include <unordered_map>

namespace bsz {

uint64_t source_hash(const std::string& s);

namespace {

int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

}  // namespace

FusedOptions FusedOptions::from_env() {
    FusedOptions o;
    o.chunk_rows = env_int("BSZ_CHUNK_ROWS", o.chunk_rows);
    o.warmup_rows = env_int("BSZ_WARMUP_ROWS", o.warmup_rows);
    o.waves_per_block = env_int("BSZ_WAVES_PER_BLOCK", o.waves_per_block);
    o.rows_per_step = env_int("BSZ_ROWS_PER_STEP", o.rows_per_step);
    o.words_per_lane = env_int("BSZ_WORDS_PER_LANE", o.words_per_lane);
    o.store_order = env_int("BSZ_STORE_ORDER", o.store_order);
    o.nontemporal = env_int("BSZ_NONTEMPORAL", o.nontemporal);
    o.interleave = env_int("BSZ_INTERLEAVE", o.interleave);
    o.persistent = env_int("BSZ_PERSISTENT", o.persistent);
    o.max_ops = env_int("BSZ_MAX_OPS", o.max_ops);
    o.max_live = env_int("BSZ_MAX_LIVE", o.max_live);
    o.occupancy = env_int("BSZ_OCCUPANCY", o.occupancy);
    if (const char* d = std::getenv("BSZ_CACHE_DIR")) o.cache_dir = d;
    o.verbose = env_int("BSZ_VERBOSE", 0) != 0;
    return o;
}

FusedOptions FusedOptions::from_json(const std::string& text) {
    if (text.empty()) return from_env();
    return from_json_obj(json_parse(text));
}

FusedOptions FusedOptions::from_json_obj(const Json& j) {
    FusedOptions o = from_env();
    auto get = [&](const char* k, int& dst) {
        if (const Json* v = j.get(k)) dst = (int)v->as_int(dst);
    };
    get("chunk_rows", o.chunk_rows);
    get("warmup_rows", o.warmup_rows);
    get("waves_per_block", o.waves_per_block);
    get("rows_per_step", o.rows_per_step);
    get("words_per_lane", o.words_per_lane);
    get("store_order", o.store_order);
    get("nontemporal", o.nontemporal);
    get("interleave", o.interleave);
    get("max_ops", o.max_ops);
    get("max_live", o.max_live);
    get("occupancy", o.occupancy);
    get("persistent", o.persistent);
    get("agprs", o.agprs);
    get("lazy_loads", o.lazy_loads);
    get("join_warmup", o.join_warmup);
    get("bitop3", o.bitop3);
    get("order", o.order);
    get("compile", o.compile);
    get("compile_shard", o.compile_shard);
    get("compile_shards", o.compile_shards);
    get("device_if", o.device_if);
    if (const Json* v = j.get("spin_limit")) o.spin_limit = v->as_int(o.spin_limit);
    if (const Json* v = j.get("cache_dir")) o.cache_dir = v->s;
    return o;
}

PlanOptions to_plan_options(const FusedOptions& f) {
    PlanOptions p;
    p.chunk_rows = std::max(1, f.chunk_rows);
    p.warmup_rows = std::max(1, f.warmup_rows);
    p.waves_per_block = std::min(16, std::max(1, f.waves_per_block));
    p.rows_per_step = std::max(1, f.rows_per_step);
    p.words_per_lane = f.words_per_lane >= 2 ? 2 : 1;
    p.store_order = f.store_order != 0;
    p.nontemporal = std::max(0, std::min(2, f.nontemporal));
    p.interleave = f.interleave != 0;
    p.max_ops = std::max(1, f.max_ops);
    p.max_live = std::max(4, f.max_live);
    p.occupancy = f.occupancy <= 0 ? 0 : std::min(8, f.occupancy);
    p.persistent = f.persistent != 0;
    p.agprs = f.agprs != 0 && p.occupancy == 1;
    p.lazy_loads = f.lazy_loads != 0;
    p.join_warmup = f.join_warmup != 0;
    p.bitop3 = f.bitop3 != 0;
    p.order = f.order == 1 ? 1 : 0;
    return p;
}

// ════════════════════════════════════════════════════════════════════════
// hipRTC
// ════════════════════════════════════════════════════════════════════════

std::string default_arch() {
    if (const char* a = std::getenv("BSZ_OFFLOAD_ARCH")) return a;
    return "gfx950";
}

// hipRTC options of every generated unit (after --offload-arch).  Packed
// FP32 instructions stay disabled, as in the rest of the library.
// BSZ_RTC_OPTIONS appends space-separated options (for experiments with the
// compiler; they enter the cache key like the fixed ones).
const std::vector<std::string>& rtc_options() {
    static const std::vector<std::string> o = [] {
        std::vector<std::string> v = {"-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                                      "-packed-fp32-ops"};
        if (const char* e = std::getenv("BSZ_RTC_OPTIONS")) {
            std::istringstream in(e);
            for (std::string t; in >> t;) v.push_back(t);
        }
        return v;
    }();
    return o;
}

// Identifies the compiler and its options in the code-object cache key, so a
// ROCm upgrade or a changed flag never reuses an old code object.
const std::string& rtc_identity() {
    static const std::string id = [] {
        int major = 0, minor = 0;
        (void)hiprtcVersion(&major, &minor);
        std::string s = "hiprtc " + std::to_string(major) + "." + std::to_string(minor) + " |";
        for (const std::string& o : rtc_options()) s += " " + o;
        return s;
    }();
    return id;
}

std::vector<char> compile_module(const std::string& src, const std::string& arch, std::string* log) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "bsz_fused.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        throw RunError("hiprtcCreateProgram failed");
    std::string archopt = "--offload-arch=" + arch;
    std::vector<const char*> opts{archopt.c_str()};
    for (const std::string& o : rtc_options()) opts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t logsz = 0;
    hiprtcGetProgramLogSize(prog, &logsz);
    std::string lg(logsz, '\0');
    if (logsz) hiprtcGetProgramLog(prog, &lg[0]);
    if (log) *log = lg;
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        throw RunError("hipRTC compile failed: " + lg.substr(0, 4000));
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    std::vector<char> code(sz);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

namespace {

std::string cache_root(const std::string& configured) {
    if (!configured.empty()) return configured;
    if (const char* d = std::getenv("BSZ_CACHE_DIR")) return d;
    if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/bsz_hip";
    return "/tmp/bsz_hip_cache";
}

void mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(path.c_str(), 0755);
}

// Register use of a compiled kernel, read from its code object's metadata
// (the msgpack note of an AMDGPU code object: ".vgpr_count", ".agpr_count",
// ".private_segment_fixed_size").
struct KernelRegs {
    int vgprs = -1, agprs = -1, scratch = -1;
};

int64_t msgpack_uint_after(const std::vector<char>& code, const std::string& key) {
    // a msgpack fixstr key (length < 32) followed by a non-negative integer
    std::string pat;
    pat.push_back((char)(0xa0 | (key.size() & 31)));
    pat += key;
    auto it = std::search(code.begin(), code.end(), pat.begin(), pat.end());
    if (it == code.end()) return -1;
    size_t i = (size_t)(it - code.begin()) + pat.size();
    if (i >= code.size()) return -1;
    const uint8_t t = (uint8_t)code[i];
    auto be = [&](size_t n) -> int64_t {
        if (i + n >= code.size()) return -1;
        uint64_t v = 0;
        for (size_t k = 1; k <= n; ++k) v = (v << 8) | (uint8_t)code[i + k];
        return (int64_t)v;
    };
    if (t < 0x80) return t;
    if (t == 0xcc) return be(1);
    if (t == 0xcd) return be(2);
    if (t == 0xce) return be(4);
    return -1;
}

KernelRegs kernel_regs(const std::vector<char>& code) {
    KernelRegs r;
    r.vgprs = (int)msgpack_uint_after(code, ".vgpr_count");
    r.agprs = (int)msgpack_uint_after(code, ".agpr_count");
    r.scratch = (int)msgpack_uint_after(code, ".private_segment_fixed_size");
    return r;
}

// A generated kernel runs only if its code object shows that every value stays
// in the architectural VGPRs: no AGPRs, no scratch, and all three counts
// present in the metadata (a missing count is treated as a refusal).  A kernel
// that overflows moves values through AGPRs or scratch on every row and loses
// the streaming rate the partition was sized for, so it is regenerated or
// refused rather than run.  (Two wide random-DAG partitions whose kernels
// overflowed also gave 64-bit values with correct low and wrong high halves;
// that is the signature of the OR miscompile the generated code now guards
// against with opq(), see fused_codegen.cpp.)  A schedule that asks for the
// AGPRs ({"occupancy": 1, "agprs": 1}) lets a kernel keep values there, the
// second half of the 512 registers a wave has at occupancy 1; scratch is
// never allowed.
bool regs_clean(const KernelRegs& r, bool allow_agprs) {
    return r.vgprs > 0 && r.agprs >= 0 && (allow_agprs || r.agprs == 0) && r.scratch == 0;
}

void refuse_unclean(const PlanSpec& p, const std::vector<KernelRegs>& regs) {
    for (size_t k = 0; k < regs.size() && k < p.kernels.size(); ++k) {
        const KernelRegs& r = regs[k];
        if (regs_clean(r, p.opt.agprs)) continue;
        std::ostringstream m;
        m << "schedule refused: generated kernel " << p.kernels[k].name << " (" << p.kernels[k].nodes.size()
          << " nodes, " << p.kernels[k].inputs.size() << " inputs) ";
        if (r.vgprs < 0 || r.agprs < 0 || r.scratch < 0)
            m << "has a code object without register metadata";
        else
            m << "uses " << r.agprs << " AGPRs and " << r.scratch << " bytes of scratch per lane beside "
              << r.vgprs << " VGPRs";
        m << "; kernels that overflow the architectural VGPRs are not run (choose a higher occupancy, fewer "
             "rows_per_step or a smaller max_live)";
        throw RunError(m.str());
    }
}

// The entry points of a plan's kernels (code objects owned by the Jit).
struct Module {
    std::vector<hipFunction_t> fns;
    std::vector<size_t> scratch;   // private (spill) bytes per lane of each kernel
    std::vector<KernelRegs> regs;  // from the code-object metadata
};
using ModuleP = std::shared_ptr<Module>;

// Code objects cached in memory and on disk by the hash of their unit, so a
// kernel compiled once (by any plan, in this process or an earlier one) is
// loaded without hipRTC.  Missing units of a plan are compiled in parallel.
class Jit {
public:
    explicit Jit(std::string dir) : dir_(cache_root(dir)) {}
    ~Jit() {
        for (auto& kv : loaded_)
            if (kv.second.first) (void)hipModuleUnload(kv.second.first);
    }

    // Compiles (or finds cached) every kernel unit of `p`; no device needed.
    void compile_all(const PlanSpec& p, double* compile_ms, int shard = 0, int shards = 1) {
        const std::string arch = default_arch();
        std::vector<const KernelSpec*> todo;
        std::set<uint64_t> seen;
        for (size_t ki = 0; ki < p.kernels.size(); ++ki) {
            const KernelSpec& k = p.kernels[ki];
            if (shards > 1 && (int)(ki % (size_t)shards) != shard) continue;
            const uint64_t h = unit_hash(k.unit, arch);
            dump(k.unit, h);
            if (seen.count(h) || have_code(h)) continue;
            seen.insert(h);
            todo.push_back(&k);
        }
        if (todo.empty()) return;
        auto t0 = std::chrono::steady_clock::now();
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        // BSZ_COMPILE_THREADS caps the hipRTC threads of this process
        const unsigned cap = (unsigned)std::max(1, env_int("BSZ_COMPILE_THREADS", 16));
        const size_t nthr = std::min<size_t>(todo.size(), std::min(cap, hw));
        std::atomic<size_t> next{0};
        std::mutex err_mu;
        std::string err;
        auto work = [&] {
            for (;;) {
                const size_t i = next.fetch_add(1);
                if (i >= todo.size()) return;
                try {
                    std::vector<char> code = compile_module(todo[i]->unit, arch, nullptr);
                    store(unit_hash(todo[i]->unit, arch), code);
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> g(err_mu);
                    if (err.empty()) err = e.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nthr; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        if (compile_ms) *compile_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!err.empty()) throw RunError(err);
    }

    // Register use of each kernel of `p` from its compiled code object
    // (vgprs -1 where the unit has not been compiled).
    std::vector<KernelRegs> regs_of(const PlanSpec& p) {
        const std::string arch = default_arch();
        std::vector<KernelRegs> out;
        for (const KernelSpec& k : p.kernels) {
            const std::vector<char> code = read_code(unit_hash(k.unit, arch));
            out.push_back(code.empty() ? KernelRegs{} : kernel_regs(code));
        }
        return out;
    }

    // Compiles every kernel of `p`; a kernel whose code object overflows the
    // architectural VGPRs is regenerated smaller (shrink_kernel) and compiled
    // again, and the plan is refused if one still overflows.
    std::vector<KernelRegs> compile_clean(PlanSpec& p, double* compile_ms) {
        compile_all(p, compile_ms);
        std::vector<KernelRegs> regs = regs_of(p);
        for (int round = 0; round < 12; ++round) {
            bool changed = false;
            for (size_t k = 0; k < p.kernels.size(); ++k)
                if (!regs_clean(regs[k], p.opt.agprs) && shrink_kernel(p, k)) changed = true;
            if (!changed) break;
            compile_all(p, compile_ms);
            regs = regs_of(p);
        }
        refuse_unclean(p, regs);
        return regs;
    }

    ModuleP load(PlanSpec& p, double* compile_ms) {
        compile_clean(p, compile_ms);
        const std::string arch = default_arch();
        auto m = std::make_shared<Module>();
        for (const KernelSpec& k : p.kernels) {
            const uint64_t h = unit_hash(k.unit, arch);
            auto it = loaded_.find(h);
            if (it == loaded_.end()) {
                std::vector<char> code = read_code(h);
                if (code.empty()) throw RunError("internal: kernel code object missing after compilation");
                hipModule_t mod;
                BSZ_HIP(hipModuleLoadData(&mod, code.data()));
                hipFunction_t f;
                BSZ_HIP(hipModuleGetFunction(&f, mod, k.entry.c_str()));
                it = loaded_.emplace(h, std::make_pair(mod, f)).first;
                regs_[h] = kernel_regs(code);
            }
            m->fns.push_back(it->second.second);
            m->regs.push_back(regs_[h]);
            int local = 0;
            m->scratch.push_back(
                hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, it->second.second) == hipSuccess
                    ? (size_t)std::max(0, local)
                    : 0);
            // the loader's view of private memory must agree with the metadata
            if (m->scratch.back() > 0 && m->regs.back().scratch == 0) m->regs.back().scratch = (int)m->scratch.back();
        }
        refuse_unclean(p, m->regs);
        return m;
    }

private:
    std::string dir_;
    std::mutex mu_;
    std::unordered_map<uint64_t, std::vector<char>> codes_;                        // compiled this process
    std::unordered_map<uint64_t, std::pair<hipModule_t, hipFunction_t>> loaded_;    // on the device
    std::unordered_map<uint64_t, KernelRegs> regs_;

    static uint64_t unit_hash(const std::string& unit, const std::string& arch) {
        return source_hash(unit + "|" + arch + "|" + rtc_identity());
    }
    // A cache file is a 16-byte header (magic, the unit's key) and the code
    // object; a file whose header does not match its name is not used.
    static constexpr char kMagic[8] = {'B', 'S', 'Z', 'C', 'O', '0', '0', '2'};
    std::string path_of(uint64_t h) const {
        char name[64];
        std::snprintf(name, sizeof name, "/%016llx.co", (unsigned long long)h);
        return dir_ + name;
    }
    void dump(const std::string& unit, uint64_t h) const {
        if (const char* d = std::getenv("BSZ_DUMP_DIR")) {
            mkdirs(d);
            char name[64];
            std::snprintf(name, sizeof name, "/%016llx.hip", (unsigned long long)h);
            std::ofstream f(std::string(d) + name);
            f << unit;
        }
    }
    bool have_code(uint64_t h) {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (codes_.count(h)) return true;
        }
        return !read_code(h).empty();
    }
    std::vector<char> read_code(uint64_t h) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = codes_.find(h);
            if (it != codes_.end()) return it->second;
        }
        std::ifstream f(path_of(h), std::ios::binary);
        if (!f) return {};
        std::vector<char> all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        uint64_t key = 0;
        if (all.size() <= 16 || std::memcmp(all.data(), kMagic, 8) != 0) return {};
        std::memcpy(&key, all.data() + 8, 8);
        if (key != h) return {};
        return std::vector<char>(all.begin() + 16, all.end());
    }
    void store(uint64_t h, const std::vector<char>& code) {
        {
            std::lock_guard<std::mutex> g(mu_);
            codes_[h] = code;
        }
        mkdirs(dir_);
        const std::string path = path_of(h);
        std::ostringstream tmpn;
        tmpn << path << ".tmp" << (long)getpid() << "_" << std::this_thread::get_id();
        {
            std::ofstream f(tmpn.str(), std::ios::binary);
            f.write(kMagic, 8);
            f.write(reinterpret_cast<const char*>(&h), 8);
            f.write(code.data(), (std::streamsize)code.size());
        }
        std::rename(tmpn.str().c_str(), path.c_str());
    }
};

// A plan whose kernels all compile within the architectural VGPRs.  Under
// automatic occupancy a partition that is still refused after every kernel
// was shrunk is partitioned again with half the fusion budget (max_live,
// max_ops), up to three times; an explicit occupancy is the caller's choice
// and its refusal stands.  Any partition gives the same bits.
std::shared_ptr<PlanSpec> clean_plan(Jit& jit, const SegGraph& g, PlanOptions o, double* compile_ms,
                                     std::vector<KernelRegs>* regs_out) {
    for (int attempt = 0;; ++attempt) {
        auto spec = std::make_shared<PlanSpec>(build_plan(g, o));
        if (spec->kernels.empty()) return spec;
        try {
            std::vector<KernelRegs> r = jit.compile_clean(*spec, compile_ms);
            if (regs_out) *regs_out = std::move(r);
            return spec;
        } catch (const RunError& e) {
            if (o.occupancy > 0 || attempt >= 3 || std::string(e.what()).find("schedule refused") == std::string::npos)
                throw;
            o.max_live = std::max(16, o.max_live / 2);
            o.max_ops = std::max(16, o.max_ops / 2);
        }
    }
}

// ════════════════════════════════════════════════════════════════════════
// Plan runtime: device tables and state for one plan over a set of instances
// ════════════════════════════════════════════════════════════════════════

struct Arena {
    size_t bytes = 0;
    size_t take(size_t n) {
        size_t o = (bytes + 255) & ~size_t(255);
        bytes = o + n;
        return o;
    }
};

// Records the work `enqueue` puts on `s` into an executable graph.  Refuses a
// run of more than kMaxGraphLaunches launches before capturing anything, and
// ends the capture if `enqueue` throws, so the stream never stays in capture
// mode.
template <class F>
hipGraphExec_t capture_graph(hipStream_t s, int64_t launches, F&& enqueue) {
    if (launches > kMaxGraphLaunches)
        throw RunError("graph capture of " + std::to_string(launches) + " kernel launches refused (at most " +
                       std::to_string(kMaxGraphLaunches) +
                       " per graph: a 131k-launch capture crashed the host); run such a plan without capture");
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    BSZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        abandon_capture(s);
        throw;
    }
    BSZ_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ex;
}

class PlanRT {
public:
    std::shared_ptr<PlanSpec> spec;
    ModuleP mod;
    uint32_t n_inst = 0, n_chunks = 0;
    uint64_t chunk_rows = 16;
    uint32_t spin_limit = kSpinLimit;   // polls of a look-back wait before it gives up
    unsigned* err_ext = nullptr;        // error word kept outside the arena (an engine's, across flushes)

    ~PlanRT() { release(); }

    void release() {
        if (graph_) { (void)hipGraphExecDestroy(graph_); graph_ = nullptr; }
        if (dev_) { (void)hipFree(dev_); dev_ = nullptr; }
        dev_bytes_ = 0;
    }

    // Lays out the device arena for `geoms` (one per instance).
    void configure(const std::vector<Geom>& geoms, hipStream_t s) {
        const PlanSpec& P = *spec;
        chunk_rows = (uint64_t)std::max(1, P.opt.chunk_rows);
        n_inst = (uint32_t)geoms.size();
        inst_h_.assign(n_inst, InstDesc{});
        uint64_t nc = 0;
        for (uint32_t i = 0; i < n_inst; ++i) {
            InstDesc& d = inst_h_[i];
            d.rows = geoms[i].rows;
            d.lw = geoms[i].lw;
            d.lastmask = geoms[i].lastmask;
            d.opw = geoms[i].opw;
            d.opmask = geoms[i].opmask;
            d.chunk_begin = (uint32_t)nc;
            d.n_chunks = (uint32_t)((d.rows + chunk_rows - 1) / chunk_rows);
            nc += d.n_chunks;
        }
        if (nc >= 0xffffffffull) throw RunError("too many chunks in one launch");
        n_chunks = (uint32_t)nc;
        chunk_inst_h_.clear();
        if (n_inst > 1) {
            chunk_inst_h_.resize(n_chunks);
            for (uint32_t i = 0; i < n_inst; ++i)
                for (uint32_t c = 0; c < inst_h_[i].n_chunks; ++c) chunk_inst_h_[inst_h_[i].chunk_begin + c] = i;
        }
        Arena a;
        const size_t K = P.kernels.size();
        kres_.assign(K, KRes{});
        size_t total_pop = 0;
        off_inst_ = a.take(sizeof(InstDesc) * n_inst);
        off_chunk_inst_ = n_inst > 1 ? a.take(4ull * n_chunks) : 0;
        off_ctl_ = a.take(ctl_bytes());
        // run numbers of the chunks, cleared with the ctl block whenever the
        // layout is set
        off_runs_ = a.take(4ull * std::max<uint32_t>(1, n_chunks));
        for (size_t k = 0; k < K; ++k) total_pop += P.kernels[k].pops.size() * n_inst;
        uint32_t max_chunks = 1;
        for (const InstDesc& d : inst_h_) max_chunks = std::max(max_chunks, d.n_chunks);
        spread_log_ = 0;
        while (spread_log_ < (uint32_t)kMaxPopSpreadLog && (4u << spread_log_) < max_chunks) ++spread_log_;
        pop_stride_ = total_pop << spread_log_;
        off_pop_ = a.take(8ull * std::max<size_t>(1, 2 * pop_stride_));
        size_t pop_cursor = 0;
        for (size_t k = 0; k < K; ++k) {
            const KernelSpec& ks = P.kernels[k];
            KRes& r = kres_[k];
            r.nin = ks.inputs.size() + ks.preds.size();
            r.nout = ks.outputs.size();
            r.npop = ks.pops.size();
            r.nst = ks.n_state;
            r.in_tab = a.take(8ull * std::max<size_t>(1, r.nin * n_inst));
            r.out_tab = a.take(8ull * std::max<size_t>(1, r.nout * n_inst));
            r.pop = off_pop_ + (8ull << spread_log_) * pop_cursor;
            pop_cursor += (size_t)r.npop * n_inst;
            // tagged end-state halves: a stale tag never matches a run
            // (tags are run + 1, and the region starts at zero)
            if (r.nst) r.st_t = a.take(16ull * n_chunks * r.nst);
        }
        total_pop_ = total_pop;
        if (a.bytes > dev_bytes_) {
            if (graph_) { (void)hipGraphExecDestroy(graph_); graph_ = nullptr; }
            if (dev_) { BSZ_HIP(hipStreamSynchronize(s)); BSZ_HIP(hipFree(dev_)); }
            BSZ_HIP(hipMalloc(&dev_, a.bytes));
            BSZ_HIP(hipMemsetAsync(dev_, 0, a.bytes, s));
            dev_bytes_ = a.bytes;
        } else {
            // no run yet, no error, both popcount banks clear, and no stale
            // tag that a run could match, for the new layout
            BSZ_HIP(hipMemsetAsync(ptr(off_ctl_), 0, ctl_bytes(), s));
            BSZ_HIP(hipMemsetAsync(ptr(off_runs_), 0, 4ull * std::max<uint32_t>(1, n_chunks), s));
            BSZ_HIP(hipMemsetAsync(ptr(off_pop_), 0, 8ull * std::max<size_t>(1, 2 * pop_stride_), s));
            for (const KRes& r : kres_)
                if (r.nst) BSZ_HIP(hipMemsetAsync(ptr(r.st_t), 0, 16ull * n_chunks * r.nst, s));
        }
        in_tab_h_.assign(K, {});
        out_tab_h_.assign(K, {});
        for (size_t k = 0; k < K; ++k) {
            in_tab_h_[k].assign((size_t)kres_[k].nin * n_inst, nullptr);
            out_tab_h_[k].assign((size_t)kres_[k].nout * n_inst, nullptr);
        }
        BSZ_HIP(hipMemcpyAsync(ptr(off_inst_), inst_h_.data(), sizeof(InstDesc) * n_inst, hipMemcpyHostToDevice, s));
        if (n_inst > 1)
            BSZ_HIP(hipMemcpyAsync(ptr(off_chunk_inst_), chunk_inst_h_.data(), 4ull * n_chunks, hipMemcpyHostToDevice, s));
        tables_dirty_ = true;
    }

    void set_in(size_t k, uint32_t inst, size_t j, const uint64_t* p) {
        in_tab_h_[k][(size_t)inst * kres_[k].nin + j] = p;
        tables_dirty_ = true;
    }
    // entry j of kernel k's input table: bank 0 of the slots of popcount
    // `pop` of instance `inst`, read as a predicate
    void set_pred(size_t k, uint32_t inst, size_t j, size_t pop) {
        const size_t idx = pop_index(pop, inst);
        in_tab_h_[k][(size_t)inst * kres_[k].nin + j] = (const uint64_t*)ptr(off_pop_ + 8ull * (idx << spread_log_));
        tables_dirty_ = true;
    }
    void set_out(size_t k, uint32_t inst, size_t j, uint64_t* p) {
        out_tab_h_[k][(size_t)inst * kres_[k].nout + j] = p;
        tables_dirty_ = true;
    }

    void upload_tables(hipStream_t s) {
        if (!tables_dirty_) return;
        for (size_t k = 0; k < kres_.size(); ++k) {
            if (kres_[k].nin) BSZ_HIP(hipMemcpyAsync(ptr(kres_[k].in_tab), in_tab_h_[k].data(), 8ull * in_tab_h_[k].size(), hipMemcpyHostToDevice, s));
            if (kres_[k].nout) BSZ_HIP(hipMemcpyAsync(ptr(kres_[k].out_tab), out_tab_h_[k].data(), 8ull * out_tab_h_[k].size(), hipMemcpyHostToDevice, s));
        }
        BSZ_HIP(hipStreamSynchronize(s));  // host tables may change after this call
        tables_dirty_ = false;
    }

    // Every kernel, asynchronously on s, one launch each; the last advances
    // the run numbers of the chunks.
    void launch(hipStream_t s) {
        const PlanSpec& P = *spec;
        const int wpb = P.opt.waves_per_block;
        const uint32_t all_blocks = (n_chunks + wpb - 1) / wpb;
        const size_t K = P.kernels.size();
        const int64_t before = launches_;
        for (size_t k = 0; k < K; ++k) {
            const KRes& r = kres_[k];
            KArgs A{};
            A.in = (const uint64_t* const*)ptr(r.in_tab);
            A.out = (uint64_t* const*)ptr(r.out_tab);
            A.pop_acc = (unsigned long long*)ptr(r.pop);
            A.inst = (const InstDesc*)ptr(off_inst_);
            A.chunk_inst = n_inst > 1 ? (const uint32_t*)ptr(off_chunk_inst_) : nullptr;
            A.st_t = r.nst ? (unsigned long long*)ptr(r.st_t) : nullptr;
            A.runs = (unsigned*)ptr(off_runs_);
            A.kstat = (unsigned*)ptr(off_ctl_ + 4ull * (4 + 8 * k));
            A.err = err_word();
            A.pop_stride = pop_stride_;
            A.n_chunks = n_chunks;
            A.n_inst = n_inst;
            A.last = k + 1 == K ? 1u : 0u;
            A.spread_log = spread_log_;
            A.spin_limit = spin_limit;
            const uint32_t blocks = P.opt.persistent ? std::min(all_blocks, resident_blocks(k, wpb)) : all_blocks;
            A.grid_waves = blocks * (uint32_t)wpb;
            void* params[] = {&A};
            BSZ_HIP(hipModuleLaunchKernel(mod->fns[k], blocks, 1, 1, 64 * wpb, 1, 1, 0, s, params, nullptr));
            ++launches_;
        }
        per_run_launches_ = launches_ - before;
    }

    int64_t launches_per_run() const { return (int64_t)spec->kernels.size(); }

    // Workgroups of kernel k the GPU keeps resident at once: whole
    // workgroups of `wpb` waves per CU, at the waves per SIMD its compiled
    // VGPR count allows (8-register granules, at most 8 waves).
    uint32_t resident_blocks(size_t k, int wpb) const {
        const int vg = k < mod->regs.size() && mod->regs[k].vgprs > 0 ? mod->regs[k].vgprs : 256;
        const int per_simd = std::max(1, std::min(8, 512 / ((vg + 7) & ~7)));
        const uint32_t per_cu = (uint32_t)std::max(1, 4 * per_simd / wpb);
        return (uint32_t)device_cus() * per_cu;
    }
    static int device_cus() {
        static const int n = [] {
            int dev = 0, v = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
                return 256;
            return v;
        }();
        return n;
    }

    void capture(hipStream_t s) {
        if (graph_) { (void)hipGraphExecDestroy(graph_); graph_ = nullptr; }
        graph_ = capture_graph(s, launches_per_run(), [&] { launch(s); });
    }

    void replay(hipStream_t s) {
        if (!graph_) throw RunError("plan has no captured graph");
        BSZ_HIP(hipGraphLaunch(graph_, s));
        launches_ += per_run_launches_;
    }

    // popcount of SegGraph pop `p` for instance `inst`, after a run (sync)
    void read_pops(hipStream_t s, std::vector<unsigned long long>& out) {
        out.assign(total_pop_, 0);
        if (!total_pop_) return;
        const uint32_t runs = completed_runs(s);
        if (runs == 0) return;  // no run yet
        const size_t bank = (runs - 1) & 1u;
        std::vector<unsigned long long> slots(pop_stride_);
        BSZ_HIP(hipMemcpyAsync(slots.data(), ptr(off_pop_ + 8ull * bank * pop_stride_), 8ull * pop_stride_,
                               hipMemcpyDeviceToHost, s));
        BSZ_HIP(hipStreamSynchronize(s));
        const size_t S = size_t(1) << spread_log_;
        for (size_t q = 0; q < total_pop_; ++q)
            for (size_t j = 0; j < S; ++j) out[q] += slots[q * S + j];
    }
    // index into read_pops() output of graph pop p for instance i
    size_t pop_index(size_t p, uint32_t inst) const {
        size_t base = 0;
        for (size_t k = 0; k < spec->kernels.size(); ++k) {
            const KernelSpec& ks = spec->kernels[k];
            for (size_t j = 0; j < ks.pops.size(); ++j)
                if ((size_t)ks.pops[j] == p) return base + (size_t)inst * ks.pops.size() + j;
            base += ks.pops.size() * n_inst;
        }
        throw RunError("popcount not in plan");
    }
    // Runs completed over chunk 0 (every chunk's count is the same between
    // runs), after checking the plan's error word (sync).
    uint32_t completed_runs(hipStream_t s) {
        uint32_t v[2] = {0, 0};
        BSZ_HIP(hipMemcpyAsync(&v[0], err_word(), 4, hipMemcpyDeviceToHost, s));
        BSZ_HIP(hipMemcpyAsync(&v[1], ptr(off_runs_), 4, hipMemcpyDeviceToHost, s));
        BSZ_HIP(hipStreamSynchronize(s));
        if (v[0] & 1u)
            throw RunError("a fused kernel's look-back wait for its predecessor chunk reached its bound; "
                           "the outputs of that run are not exact");
        return v[1];
    }
    // counter `which` of kernel k for the last run (sync): 0 chunks run again
    // from their predecessor's end state, 1 chunks whose warm-up did not
    // prove their start state, 2 look-back waits that reached the bound
    uint32_t kernel_stat(hipStream_t s, size_t k, int which) {
        if (k >= kres_.size()) throw RunError("no kernel " + std::to_string(k) + " in this plan");
        const uint32_t runs = completed_runs(s);
        if (runs == 0) return 0;
        uint32_t v = 0;
        const size_t at = 4 + 8 * k + 4 * ((runs - 1) & 1u) + (size_t)which;
        BSZ_HIP(hipMemcpyAsync(&v, ptr(off_ctl_ + 4ull * at), 4, hipMemcpyDeviceToHost, s));
        BSZ_HIP(hipStreamSynchronize(s));
        return v;
    }
    uint32_t mismatches(hipStream_t s, size_t k) { return kernel_stat(s, k, 0); }
    int64_t launches() const { return launches_; }
    size_t device_bytes() const { return dev_bytes_; }

private:
    struct KRes {
        size_t nin = 0, nout = 0;
        int npop = 0, nst = 0;
        size_t in_tab = 0, out_tab = 0, pop = 0, st_t = 0;
    };
    void* dev_ = nullptr;
    size_t dev_bytes_ = 0;
    size_t off_inst_ = 0, off_chunk_inst_ = 0, off_ctl_ = 0, off_runs_ = 0, off_pop_ = 0, total_pop_ = 0,
           pop_stride_ = 0;
    uint32_t spread_log_ = 0;
    std::vector<KRes> kres_;
    std::vector<InstDesc> inst_h_;
    std::vector<uint32_t> chunk_inst_h_;
    std::vector<std::vector<const uint64_t*>> in_tab_h_;
    std::vector<std::vector<uint64_t*>> out_tab_h_;
    bool tables_dirty_ = true;
    hipGraphExec_t graph_ = nullptr;
    int64_t launches_ = 0, per_run_launches_ = 0;

    char* ptr(size_t off) const { return static_cast<char*>(dev_) + off; }
    // error word, 3 spare, then 8 counters per kernel
    size_t ctl_bytes() const { return 4ull * (4 + 8 * spec->kernels.size()); }
    unsigned* err_word() const { return err_ext ? err_ext : (unsigned*)ptr(off_ctl_); }
};

// Binds the buffers of one instance: sources, segment outputs and slots.
void bind_instance(PlanRT& rt, uint32_t inst, const std::vector<const uint64_t*>& sources,
                   const std::vector<uint64_t*>& seg_out_by_node, const std::vector<uint64_t*>& slots) {
    const PlanSpec& P = *rt.spec;
    for (size_t k = 0; k < P.kernels.size(); ++k) {
        const KernelSpec& ks = P.kernels[k];
        for (size_t j = 0; j < ks.inputs.size(); ++j) {
            const KInput& in = ks.inputs[j];
            const uint64_t* p = in.kind == KInput::Source ? sources.at(in.id)
                              : in.kind == KInput::Slot ? slots.at(in.id)
                                                        : seg_out_by_node.at(in.id);
            rt.set_in(k, inst, j, p);
        }
        for (size_t q = 0; q < ks.preds.size(); ++q) rt.set_pred(k, inst, ks.inputs.size() + q, (size_t)ks.preds[q]);
        for (size_t j = 0; j < ks.outputs.size(); ++j) {
            uint64_t* p = ks.out_slot[j] >= 0 ? slots.at(ks.out_slot[j]) : seg_out_by_node.at(ks.outputs[j]);
            rt.set_out(k, inst, j, p);
        }
    }
}

// ════════════════════════════════════════════════════════════════════════
// FusedEngine
// ════════════════════════════════════════════════════════════════════════

struct FusedPop : PopCell {
    int32_t node = -1;
    uint64_t segment = 0;
    int32_t index = -1;   // position among the open segment's popcounts
};

struct CachedPlan {
    std::shared_ptr<PlanSpec> spec;
    ModuleP mod;
    std::unique_ptr<PlanRT> rt;
};

class FusedEngine final : public Engine {
public:
    FusedEngine(hipStream_t s, const FusedOptions& opt, bool recording)
        : s_(s), fopt_(opt), popt_(to_plan_options(opt)), jit_(opt.cache_dir), recording_(recording) {}

    ~FusedEngine() override {
        if (!recording_) (void)hipStreamSynchronize(s_);
        if (slot_) (void)hipFree(slot_);
        if (err_) (void)hipFree(err_);
    }

    // ── Engine ──
    void begin(const Geom& g) override {
        if (!recording_) flush();
        g_ = g;
        zero_.reset();
        ones_.reset();
        launches0_ = launches_;
    }

    SV constant(bool ones) override {
        SV& c = ones ? ones_ : zero_;
        if (!c) {
            c = std::make_shared<StreamVal>();
            c->konst = ones ? 1 : 0;
        }
        return c;
    }

    SV input(const uint64_t* host, uint64_t n) override {
        auto v = std::make_shared<StreamVal>();
        if (recording_) {
            v->source = (int32_t)rec_sources_.size();
            rec_sources_.push_back(pending_input_);
            return v;
        }
        v->buf = std::make_shared<DevBuf>(&pool_, g_.bytes());
        upload_stream(s_, v->buf->ptr, host, n, g_);
        BSZ_HIP(hipStreamSynchronize(s_));
        return v;
    }

    void note_input(const std::string& name, int64_t idx) override { pending_input_ = {name, idx}; }

    SV op_not(const SV& a) override { return make(NOp::Not, node_of(a), -1, 0); }
    SV op_bin(BinOp op, const SV& a, const SV& b) override {
        NOp o = op == BinOp::And ? NOp::And : op == BinOp::Or ? NOp::Or : op == BinOp::Xor ? NOp::Xor : NOp::Add;
        int32_t x = node_of(a), y = node_of(b);
        return make(o, x, y, 0);
    }
    SV op_shift(bool left, const SV& a, uint64_t k) override {
        return make(left ? NOp::Shl : NOp::Shr, node_of(a), -1, k);
    }

    PopRef popcount(const SV& a) override {
        auto p = std::make_shared<FusedPop>();
        p->node = node_of(a);
        p->segment = seg_;
        p->index = (int32_t)pops_.size();
        pops_.push_back(p);
        return p;
    }

    bool supports_predicates() const override { return !recording_ || fopt_.device_if != 0; }

    PopRef predicate(const SV& c) override {
        int32_t n = node_of(c);
        // a predicate is accumulated by the kernel that computes its stream;
        // an input or constant gets a computed copy so that kernel exists
        if (nodes_[n].op == NOp::Input || nodes_[n].op == NOp::Zero || nodes_[n].op == NOp::Ones)
            n = make(NOp::Or, n, n, 0)->node;
        auto p = std::make_shared<FusedPop>();
        p->node = n;
        p->segment = seg_;
        p->index = (int32_t)pops_.size();
        pops_.push_back(p);
        return p;
    }

    SV select(const PopRef& pr, const SV& t, const SV& f) override {
        if (t == f) return t;
        auto* p = static_cast<FusedPop*>(pr.get());
        if (p->ready) return p->value != 0 ? t : f;  // resolved by an earlier flush
        if (p->segment != seg_) throw RunError("internal: predicate of a closed segment is not resolved");
        const int32_t a = node_of(t), b = node_of(f);
        SV v = make(NOp::Sel, a, b, 0);
        nodes_[v->node].pred = p->index;
        return v;
    }

    int64_t resolve(const PopRef& p) override {
        if (!p->ready) {
            if (recording_) throw RunError("program reads a popcount on the host; it cannot be recorded as one plan");
            flush();
            if (!p->ready) throw RunError("internal: popcount not produced by flush");
        }
        return p->value;
    }

    bool is_nonzero(const SV& a) override {
        if (recording_) throw RunError("program branches on stream contents; it cannot be recorded as one plan");
        flush();
        if (a->konst >= 0) return a->konst == 1 && g_.bitlength > 0;
        if (!slot_) BSZ_HIP(hipMalloc(&slot_, 8));
        launch_zero_u64(s_, slot_, 1);
        launch_nonzero(s_, slot_, a->buf->ptr, g_);
        unsigned long long v = 0;
        BSZ_HIP(hipMemcpyAsync(&v, slot_, 8, hipMemcpyDeviceToHost, s_));
        BSZ_HIP(hipStreamSynchronize(s_));
        launches_ += 2;
        return v != 0;
    }

    void download(const SV& a, uint64_t* host) override {
        flush();
        if (a->konst >= 0) {
            for (uint64_t i = 0; i < g_.lw; ++i) host[i] = a->konst ? ~0ull : 0ull;
            if (a->konst && g_.lw) host[g_.lw - 1] &= g_.lastmask;
            return;
        }
        stage_d2h(s_, host, a->buf->ptr, g_.lw * 8);
        check_waits();
    }

    // A look-back wait of any kernel since the engine began that gave up
    // leaves outputs that are not exact: report it (sync).
    void check_waits() {
        if (!err_) return;
        unsigned v = 0;
        BSZ_HIP(hipMemcpyAsync(&v, err_, 4, hipMemcpyDeviceToHost, s_));
        BSZ_HIP(hipStreamSynchronize(s_));
        if (v & 1u)
            throw RunError("a fused kernel's look-back wait for its predecessor chunk reached its bound; "
                           "the outputs of that run are not exact");
    }

    void flush() override;
    void sync() override { BSZ_HIP(hipStreamSynchronize(s_)); }
    int64_t n_kernel_launches() const override { return launches_; }
    int64_t n_mallocs() const override { return pool_.n_mallocs(); }
    size_t peak_bytes() const override { return pool_.peak_bytes(); }
    hipStream_t stream() const override { return s_; }
    std::string stats_json() const override {
        std::ostringstream o;
        o << "{\"executor\":\"fused\",\"flushes\":" << flushes_ << ",\"plans\":" << plans_.size()
          << ",\"kernels\":" << kernels_run_ << ",\"compile_ms\":" << compile_ms_
          << ",\"rerun_chunks\":" << reruns_ << "}";
        return o.str();
    }

    // ── recording (plan interface) ──
    SegGraph take_graph(std::vector<SV>& outputs_in_order, std::vector<int32_t>& out_nodes,
                        std::vector<std::shared_ptr<FusedPop>>& pops) {
        SegGraph g;
        g.nodes = nodes_;
        g.n_sources = (int32_t)rec_sources_.size();
        for (const SV& v : outputs_in_order) {
            int32_t n = node_of(v);
            out_nodes.push_back(n);
        }
        std::vector<char> seen(g.nodes.size(), 0);
        for (int32_t n : out_nodes)
            if (!seen[n]) { seen[n] = 1; g.outputs.push_back(n); }
        for (auto& p : pops_) { g.pops.push_back(p->node); pops.push_back(p); }
        g.nodes = nodes_;  // node_of() above may have added input/constant nodes
        return g;
    }
    const std::vector<std::pair<std::string, int64_t>>& recorded_sources() const { return rec_sources_; }
    Jit& jit() { return jit_; }
    const PlanOptions& plan_options() const { return popt_; }
    DevPool& pool() { return pool_; }
    double& compile_ms() { return compile_ms_; }

private:
    hipStream_t s_;
    FusedOptions fopt_;
    PlanOptions popt_;
    Jit jit_;
    bool recording_;
    Geom g_;
    DevPool pool_;
    SV zero_, ones_;
    unsigned long long* slot_ = nullptr;
    unsigned* err_ = nullptr;   // sticky error word of every plan this engine runs

    uint64_t seg_ = 1;
    std::vector<GNode> nodes_;
    std::vector<std::weak_ptr<StreamVal>> node_val_;
    std::vector<std::shared_ptr<FusedPop>> pops_;
    std::vector<DevBufP> sources_;
    std::unordered_map<const void*, int32_t> src_node_;
    int32_t zero_node_ = -1, ones_node_ = -1;
    std::vector<std::pair<std::string, int64_t>> rec_sources_;
    std::pair<std::string, int64_t> pending_input_{"", -1};
    std::unordered_map<std::string, CachedPlan> plans_;

    int64_t launches_ = 0, launches0_ = 0, flushes_ = 0, kernels_run_ = 0, reruns_ = 0;
    double compile_ms_ = 0;

    int32_t add_node(const GNode& n, const SV& v) {
        nodes_.push_back(n);
        node_val_.emplace_back(v);
        return (int32_t)nodes_.size() - 1;
    }

    int32_t node_of(const SV& v) {
        if (v->node >= 0 && v->segment == seg_) return v->node;
        if (v->konst >= 0) {
            int32_t& id = v->konst ? ones_node_ : zero_node_;
            if (id < 0) {
                GNode n;
                n.op = v->konst ? NOp::Ones : NOp::Zero;
                id = add_node(n, nullptr);
            }
            return id;
        }
        if (recording_ && v->source >= 0) {
            GNode n;
            n.op = NOp::Input;
            n.src = v->source;
            int32_t id = add_node(n, v);
            v->node = id;
            v->segment = seg_;
            return id;
        }
        if (!v->buf) throw RunError("internal: stream value without data");
        auto it = src_node_.find(v->buf->ptr);
        if (it != src_node_.end()) return it->second;
        GNode n;
        n.op = NOp::Input;
        n.src = (int32_t)sources_.size();
        sources_.push_back(v->buf);
        int32_t id = add_node(n, v);
        src_node_[v->buf->ptr] = id;
        return id;
    }

    SV make(NOp op, int32_t a, int32_t b, uint64_t k) {
        auto v = std::make_shared<StreamVal>();
        GNode n;
        n.op = op;
        n.a = a;
        n.b = b;
        n.k = k;
        v->node = add_node(n, v);
        v->segment = seg_;
        return v;
    }

    void reset_segment() {
        nodes_.clear();
        node_val_.clear();
        pops_.clear();
        sources_.clear();
        src_node_.clear();
        zero_node_ = ones_node_ = -1;
        ++seg_;
    }

public:
    CachedPlan& plan_for(const SegGraph& g) {
        std::string key = g.key() + "#" + popt_.key();
        auto it = plans_.find(key);
        if (it != plans_.end()) return it->second;
        CachedPlan cp;
        cp.spec = clean_plan(jit_, g, popt_, &compile_ms_, nullptr);
        if (!cp.spec->kernels.empty()) cp.mod = jit_.load(*cp.spec, &compile_ms_);
        cp.rt = std::make_unique<PlanRT>();
        cp.rt->spec = cp.spec;
        cp.rt->mod = cp.mod;
        return plans_.emplace(key, std::move(cp)).first->second;
    }
};

void FusedEngine::flush() {
    if (recording_) return;
    if (nodes_.empty() && pops_.empty()) return;
    ++flushes_;
    SegGraph g;
    g.nodes = nodes_;
    g.n_sources = (int32_t)sources_.size();
    std::vector<SV> out_vals;
    for (size_t i = 0; i < nodes_.size(); ++i) {
        if (nodes_[i].op == NOp::Input || nodes_[i].op == NOp::Zero || nodes_[i].op == NOp::Ones) continue;
        if (SV v = node_val_[i].lock()) {
            g.outputs.push_back((int32_t)i);
            out_vals.push_back(v);
        }
    }
    for (auto& p : pops_) g.pops.push_back(p->node);
    if (!g.outputs.empty() || !g.pops.empty()) {
        CachedPlan& cp = plan_for(g);
        PlanRT& rt = *cp.rt;
        if (!err_) {
            BSZ_HIP(hipMalloc(&err_, 4));
            BSZ_HIP(hipMemsetAsync(err_, 0, 4, s_));
        }
        rt.err_ext = err_;
        rt.configure({g_}, s_);
        std::vector<const uint64_t*> src(sources_.size());
        for (size_t i = 0; i < sources_.size(); ++i) src[i] = sources_[i]->ptr;
        std::vector<uint64_t*> seg_out(nodes_.size(), nullptr);
        std::vector<DevBufP> out_bufs;
        for (size_t i = 0; i < g.outputs.size(); ++i) {
            auto b = std::make_shared<DevBuf>(&pool_, g_.bytes());
            seg_out[g.outputs[i]] = b->ptr;
            out_bufs.push_back(b);
        }
        std::vector<DevBufP> slot_bufs;
        std::vector<uint64_t*> slots;
        for (int i = 0; i < cp.spec->n_slots; ++i) {
            slot_bufs.push_back(std::make_shared<DevBuf>(&pool_, g_.bytes()));
            slots.push_back(slot_bufs.back()->ptr);
        }
        if (!cp.spec->kernels.empty()) {
            bind_instance(rt, 0, src, seg_out, slots);
            rt.upload_tables(s_);
            rt.launch(s_);
            launches_ += (int64_t)cp.spec->kernels.size();
            kernels_run_ += (int64_t)cp.spec->kernels.size();
        }
        if (!pops_.empty()) {
            std::vector<unsigned long long> vals;
            rt.read_pops(s_, vals);
            for (size_t p = 0; p < pops_.size(); ++p) {
                pops_[p]->value = g_.bitlength > 0 ? (int64_t)vals[rt.pop_index(p, 0)] : 0;
                pops_[p]->ready = true;
            }
        }
        if (fopt_.verbose) {
            for (size_t k = 0; k < cp.spec->kernels.size(); ++k)
                if (cp.spec->kernels[k].n_state) reruns_ += rt.mismatches(s_, k);
        }
        for (size_t i = 0; i < out_vals.size(); ++i) {
            out_vals[i]->buf = out_bufs[i];
            out_vals[i]->node = -1;
        }
    }
    for (size_t i = 0; i < node_val_.size(); ++i)
        if (SV v = node_val_[i].lock()) v->node = -1;
    reset_segment();
}

}  // namespace

std::unique_ptr<Engine> make_fused_engine(hipStream_t s, const FusedOptions& opt) {
    return std::make_unique<FusedEngine>(s, opt, false);
}

// ════════════════════════════════════════════════════════════════════════
// Recorded plans (the batched interface behind bsz_plan_*)
// ════════════════════════════════════════════════════════════════════════

struct FusedPlan::Impl {
    hipStream_t s = nullptr;
    FusedOptions fopt;
    std::unique_ptr<FusedEngine> rec;       // the recording engine (owns the JIT and the pool)
    std::shared_ptr<PlanSpec> spec;
    ModuleP mod;
    std::vector<KernelRegs> compiled_regs;                  // from the code objects, before any load
    std::unique_ptr<PlanRT> rt;
    std::vector<std::pair<std::string, int64_t>> sources;   // per source: (name, idx)
    struct Out { std::string name; int64_t idx; int kind; int32_t ref; };  // kind 0 node, 1 source alias, 2 const
    std::vector<Out> outs;
    struct IntOut { std::string name; int32_t pop; int64_t value; };
    std::vector<IntOut> ints;
    int64_t op_count = 0;
    std::vector<Geom> geoms;
    std::vector<std::vector<const uint64_t*>> bound;       // [inst][source]
    std::vector<std::vector<DevBufP>> out_bufs;             // [inst][g.outputs index]
    std::vector<std::vector<DevBufP>> slot_bufs;            // [inst][slot]
    std::vector<std::vector<DevBufP>> const_bufs;           // [inst][outs index] for constant outputs
    std::vector<unsigned long long> pops_host;
    bool bound_dirty = true;

    // one kernel per DSL op over the same recorded DAG (the plain comparator)
    int executor = 1;                                       // 1 fused, 0 per-op
    std::vector<int32_t> perop_order;                       // needed computed nodes, in order
    std::vector<int32_t> perop_buf;                         // node -> scratch buffer index (-1: output/source)
    int perop_nbufs = 0;
    struct PerOpInst {
        std::vector<DevBufP> scratch;
        DevBufP zero, ones;
        void* add_state = nullptr;
        unsigned long long* pops = nullptr;
    };
    std::vector<PerOpInst> perop;
    std::vector<unsigned long long> perop_pops_host;

    ~Impl() {
        for (auto& p : perop) {
            if (p.add_state) (void)hipFree(p.add_state);
            if (p.pops) (void)hipFree(p.pops);
        }
    }

    void plan_perop() {
        const SegGraph& g = spec->g;
        const int32_t N = (int32_t)g.nodes.size();
        std::vector<char> need(N, 0), is_out(N, 0);
        for (int32_t o : g.outputs) { need[o] = 1; is_out[o] = 1; }
        for (int32_t p : g.pops) need[p] = 1;
        auto ar = [](NOp op) {
            return (op == NOp::Not || op == NOp::Shl || op == NOp::Shr) ? 1
                 : (op == NOp::And || op == NOp::Or || op == NOp::Xor || op == NOp::Add) ? 2 : 0;
        };
        for (int32_t n = N - 1; n >= 0; --n) {
            if (!need[n]) continue;
            if (ar(g.nodes[n].op) >= 1) need[g.nodes[n].a] = 1;
            if (ar(g.nodes[n].op) >= 2) need[g.nodes[n].b] = 1;
        }
        std::vector<int32_t> last(N, -1);
        for (int32_t n = 0; n < N; ++n) {
            if (!need[n]) continue;
            if (ar(g.nodes[n].op) >= 1) last[g.nodes[n].a] = n;
            if (ar(g.nodes[n].op) >= 2) last[g.nodes[n].b] = std::max(last[g.nodes[n].b], n);
        }
        for (int32_t p : g.pops) last[p] = N;
        perop_buf.assign(N, -1);
        perop_order.clear();
        std::vector<int> free_list;
        int nb = 0;
        for (int32_t n = 0; n < N; ++n) {
            if (!need[n] || ar(g.nodes[n].op) == 0) continue;
            perop_order.push_back(n);
            if (!is_out[n]) {
                int b;
                if (!free_list.empty()) { b = free_list.back(); free_list.pop_back(); }
                else b = nb++;
                perop_buf[n] = b;
            }
            // operands whose last use is this node hand their buffer back (after
            // this node's own buffer was taken, so dst never aliases an operand)
            const GNode& x = g.nodes[n];
            const int32_t ops2[2] = {x.a, (ar(x.op) >= 2 && x.b != x.a) ? x.b : -1};
            for (int32_t a : ops2)
                if (a >= 0 && last[a] == n && perop_buf[a] >= 0) free_list.push_back(perop_buf[a]);
        }
        perop_nbufs = nb;
    }
};

FusedPlan::FusedPlan() : impl_(new Impl) {}
FusedPlan::~FusedPlan() {
    if (impl_ && impl_->s) (void)hipStreamSynchronize(impl_->s);
    if (graph_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_);
}

std::unique_ptr<FusedPlan> FusedPlan::record(hipStream_t s, const Program& prog, const CaseData& shape,
                                             const FusedOptions& opt, bool compile) {
    std::unique_ptr<FusedPlan> fp(new FusedPlan);
    Impl& I = *fp->impl_;
    I.s = s;
    I.fopt = opt;
    I.rec = std::make_unique<FusedEngine>(s, opt, true);
    RecordedOutputs ro = record_program(prog, shape, *I.rec);
    I.op_count = ro.op_count;
    std::vector<SV> ovals;
    for (auto& kv : ro.streams) ovals.push_back(kv.second);
    for (auto& t : ro.array_elems) ovals.push_back(std::get<2>(t));
    std::vector<int32_t> out_nodes;
    std::vector<std::shared_ptr<FusedPop>> pops;
    SegGraph g = I.rec->take_graph(ovals, out_nodes, pops);
    I.sources = I.rec->recorded_sources();
    // outputs that are inputs or constants need no kernel
    std::vector<int32_t> computed;
    std::vector<char> seen(g.nodes.size(), 0);
    auto out_kind = [&](int32_t n) {
        if (g.nodes[n].op == NOp::Input) return 1;
        if (g.nodes[n].op == NOp::Zero || g.nodes[n].op == NOp::Ones) return 2;
        return 0;
    };
    size_t oi = 0;
    auto add_out = [&](const std::string& name, int64_t idx) {
        int32_t n = out_nodes[oi++];
        Impl::Out o{name, idx, out_kind(n), n};
        if (o.kind == 0 && !seen[n]) { seen[n] = 1; computed.push_back(n); }
        I.outs.push_back(o);
    };
    for (auto& kv : ro.streams) add_out(kv.first, -1);
    for (auto& t : ro.array_elems) add_out(std::get<0>(t), std::get<1>(t));
    g.outputs = computed;
    for (auto& io : ro.ints) {
        Impl::IntOut x{io.name, -1, io.value};
        if (io.pop) {
            for (size_t p = 0; p < pops.size(); ++p)
                if (pops[p].get() == io.pop.get()) x.pop = (int32_t)p;
            if (x.pop < 0) throw RunError("internal: int output popcount not recorded");
        }
        I.ints.push_back(x);
    }
    // kernel sources only when they will be compiled (partition and kernel
    // shapes are all a description needs)
    if (compile && std::max(1, opt.compile_shards) == 1) {
        I.spec = clean_plan(I.rec->jit(), g, to_plan_options(opt), &I.rec->compile_ms(), &I.compiled_regs);
    } else {
        I.spec = std::make_shared<PlanSpec>(build_plan(g, to_plan_options(opt), compile));
        // a shard job (one of several processes filling the cache) compiles
        // its share and reports no registers: other kernels may not exist yet
        if (compile && !I.spec->kernels.empty())
            I.rec->jit().compile_all(*I.spec, &I.rec->compile_ms(), std::max(0, opt.compile_shard),
                                     std::max(1, opt.compile_shards));
    }
    return fp;
}

void FusedPlan::load() {
    Impl& I = *impl_;
    if (!I.mod && !I.spec->kernels.empty()) I.mod = I.rec->jit().load(*I.spec, &I.rec->compile_ms());
    if (!I.rt) {
        I.rt = std::make_unique<PlanRT>();
        I.rt->spec = I.spec;
        I.rt->mod = I.mod;
        if (I.fopt.spin_limit >= 0) I.rt->spin_limit = (uint32_t)std::min<int64_t>(I.fopt.spin_limit, kSpinLimit);
    }
}

void FusedPlan::set_instances(const std::vector<int64_t>& bitlengths) {
    Impl& I = *impl_;
    load();
    I.geoms.clear();
    for (int64_t bl : bitlengths) I.geoms.push_back(Geom::make(bl, 1));
    I.rt->configure(I.geoms, I.s);
    const size_t n = I.geoms.size();
    I.bound.assign(n, std::vector<const uint64_t*>(I.sources.size(), nullptr));
    I.out_bufs.assign(n, {});
    I.slot_bufs.assign(n, {});
    I.const_bufs.assign(n, std::vector<DevBufP>(I.outs.size()));
    for (size_t i = 0; i < n; ++i) {
        for (size_t o = 0; o < I.outs.size(); ++o) {
            if (I.outs[o].kind != 2) continue;
            auto b = std::make_shared<DevBuf>(&I.rec->pool(), I.geoms[i].bytes());
            launch_fill(I.s, b->ptr, I.spec->g.nodes[I.outs[o].ref].op == NOp::Ones, I.geoms[i]);
            I.const_bufs[i][o] = b;
        }
        for (size_t o = 0; o < I.spec->g.outputs.size(); ++o)
            I.out_bufs[i].push_back(std::make_shared<DevBuf>(&I.rec->pool(), I.geoms[i].bytes()));
        for (int s = 0; s < I.spec->n_slots; ++s)
            I.slot_bufs[i].push_back(std::make_shared<DevBuf>(&I.rec->pool(), I.geoms[i].bytes()));
    }
    I.bound_dirty = true;
}

int FusedPlan::n_instances() const { return (int)impl_->geoms.size(); }

void FusedPlan::bind_input(int inst, const std::string& name, int64_t idx, const uint64_t* dev) {
    Impl& I = *impl_;
    bool hit = false;
    for (size_t j = 0; j < I.sources.size(); ++j)
        if (I.sources[j].first == name && I.sources[j].second == idx) { I.bound.at(inst)[j] = dev; hit = true; }
    if (!hit) throw RunError("plan has no input " + name + (idx >= 0 ? "[" + std::to_string(idx) + "]" : ""));
    I.bound_dirty = true;
}

uint64_t* FusedPlan::output(int inst, const std::string& name, int64_t idx) {
    Impl& I = *impl_;
    for (size_t oi = 0; oi < I.outs.size(); ++oi) {
        const Impl::Out& o = I.outs[oi];
        if (o.name != name || o.idx != idx) continue;
        if (o.kind == 1) {
            int32_t src = I.spec->g.nodes[o.ref].src;
            return const_cast<uint64_t*>(I.bound.at(inst).at(src));
        }
        if (o.kind == 2) return I.const_bufs.at(inst).at(oi)->ptr;
        for (size_t k = 0; k < I.spec->g.outputs.size(); ++k)
            if (I.spec->g.outputs[k] == o.ref) return I.out_bufs.at(inst)[k]->ptr;
    }
    throw RunError("plan has no output " + name);
}

void FusedPlan::prepare() {
    Impl& I = *impl_;
    if (!I.bound_dirty) return;
    for (size_t i = 0; i < I.geoms.size(); ++i) {
        for (size_t j = 0; j < I.sources.size(); ++j)
            if (!I.bound[i][j]) throw RunError("plan input " + I.sources[j].first + " not bound for instance " + std::to_string(i));
        std::vector<uint64_t*> seg_out(I.spec->g.nodes.size(), nullptr);
        for (size_t k = 0; k < I.spec->g.outputs.size(); ++k) seg_out[I.spec->g.outputs[k]] = I.out_bufs[i][k]->ptr;
        std::vector<uint64_t*> slots;
        for (auto& b : I.slot_bufs[i]) slots.push_back(b->ptr);
        if (!I.spec->kernels.empty()) bind_instance(*I.rt, (uint32_t)i, I.bound[i], seg_out, slots);
    }
    I.rt->upload_tables(I.s);
    I.bound_dirty = false;
}

void FusedPlan::set_executor(int executor) {
    Impl& I = *impl_;
    I.executor = executor;
    if (executor == 0 && I.perop.size() != I.geoms.size()) {
        I.plan_perop();
        I.perop.assign(I.geoms.size(), {});
        for (size_t i = 0; i < I.geoms.size(); ++i) {
            auto& P = I.perop[i];
            const Geom& g = I.geoms[i];
            for (int b = 0; b < I.perop_nbufs; ++b) P.scratch.push_back(std::make_shared<DevBuf>(&I.rec->pool(), g.bytes()));
            P.zero = std::make_shared<DevBuf>(&I.rec->pool(), g.bytes());
            P.ones = std::make_shared<DevBuf>(&I.rec->pool(), g.bytes());
            launch_fill(I.s, P.zero->ptr, false, g);
            launch_fill(I.s, P.ones->ptr, true, g);
            const size_t st = AddScanState::bytes(g);
            BSZ_HIP(hipMalloc(&P.add_state, st));
            BSZ_HIP(hipMemsetAsync(P.add_state, 0, st, I.s));
            BSZ_HIP(hipMalloc(&P.pops, 8 * kPopSlotWords * std::max<size_t>(1, I.spec->g.pops.size())));
        }
        BSZ_HIP(hipStreamSynchronize(I.s));
    }
}

void FusedPlan::run() {
    prepare();
    Impl& I = *impl_;
    if (I.executor == 1) {
        if (!I.spec->kernels.empty()) I.rt->launch(I.s);
        return;
    }
    const SegGraph& g = I.spec->g;
    for (size_t i = 0; i < I.geoms.size(); ++i) {
        auto& P = I.perop[i];
        const Geom& geo = I.geoms[i];
        std::vector<uint64_t*> seg_out(g.nodes.size(), nullptr);
        for (size_t k = 0; k < g.outputs.size(); ++k) seg_out[g.outputs[k]] = I.out_bufs[i][k]->ptr;
        auto ptr_of = [&](int32_t n) -> uint64_t* {
            const GNode& x = g.nodes[n];
            if (x.op == NOp::Input) return const_cast<uint64_t*>(I.bound[i][x.src]);
            if (x.op == NOp::Zero) return P.zero->ptr;
            if (x.op == NOp::Ones) return P.ones->ptr;
            if (seg_out[n]) return seg_out[n];
            return P.scratch.at(I.perop_buf[n])->ptr;
        };
        for (int32_t n : I.perop_order) {
            const GNode& x = g.nodes[n];
            uint64_t* d = ptr_of(n);
            switch (x.op) {
                case NOp::Not: launch_pointwise(I.s, PwOp::Not, d, ptr_of(x.a), nullptr, geo); break;
                case NOp::And: launch_pointwise(I.s, PwOp::And, d, ptr_of(x.a), ptr_of(x.b), geo); break;
                case NOp::Or: launch_pointwise(I.s, PwOp::Or, d, ptr_of(x.a), ptr_of(x.b), geo); break;
                case NOp::Xor: launch_pointwise(I.s, PwOp::Xor, d, ptr_of(x.a), ptr_of(x.b), geo); break;
                case NOp::Add: launch_add(I.s, d, ptr_of(x.a), ptr_of(x.b), geo, P.add_state); break;
                case NOp::Shl: launch_shl(I.s, d, ptr_of(x.a), x.k, geo); break;
                case NOp::Shr: launch_shr(I.s, d, ptr_of(x.a), x.k, geo); break;
                default: break;
            }
        }
        if (!g.pops.empty()) {
            for (size_t p = 0; p < g.pops.size(); ++p)
                launch_popcount(I.s, P.pops + p * kPopSlotWords, ptr_of(g.pops[p]), geo);
        }
    }
}

void FusedPlan::capture() {
    prepare();
    Impl& I = *impl_;
    if (I.executor == 0) {
        if (graph_) { (void)hipGraphExecDestroy((hipGraphExec_t)graph_); graph_ = nullptr; }
        const int64_t pops = I.spec->g.pops.empty() ? 0 : 1 + (int64_t)I.spec->g.pops.size();
        int64_t adds = 0;  // an addition is a main and a fix-up launch
        for (int32_t n : I.perop_order) adds += I.spec->g.nodes[n].op == NOp::Add ? 1 : 0;
        const int64_t per_inst = (int64_t)I.perop_order.size() + adds + pops;
        graph_ = capture_graph(I.s, per_inst * (int64_t)I.geoms.size(), [&] { run(); });
        return;
    }
    if (!I.spec->kernels.empty()) I.rt->capture(I.s);
}

void FusedPlan::replay() {
    Impl& I = *impl_;
    if (I.executor == 0) {
        if (!graph_) throw RunError("plan has no captured graph");
        BSZ_HIP(hipGraphLaunch((hipGraphExec_t)graph_, I.s));
        return;
    }
    if (!I.spec->kernels.empty()) I.rt->replay(I.s);
}

void FusedPlan::prepare_now() { prepare(); }

int64_t FusedPlan::read_int(int inst, const std::string& name, bool refresh) {
    Impl& I = *impl_;
    for (auto& io : I.ints) {
        if (io.name != name) continue;
        if (io.pop < 0) return io.value;
        if (I.executor == 0) {
            const uint64_t v = read_popcount_slots(I.s, I.perop.at(inst).pops + (size_t)io.pop * kPopSlotWords);
            return I.geoms.at(inst).bitlength <= 0 ? 0 : (int64_t)v;
        }
        if (refresh || I.pops_host.empty()) I.rt->read_pops(I.s, I.pops_host);
        if (I.geoms.at(inst).bitlength <= 0) return 0;
        return (int64_t)I.pops_host.at(I.rt->pop_index(io.pop, inst));
    }
    throw RunError("plan has no int output " + name);
}

uint32_t FusedPlan::mismatches(int kernel) { return impl_->rt->mismatches(impl_->s, kernel); }
uint32_t FusedPlan::unproven(int kernel) { return impl_->rt->kernel_stat(impl_->s, kernel, 1); }

std::string FusedPlan::describe_json(const std::string& options_array) const {
    const Json arr = json_parse(options_array);
    if (arr.k != Json::Arr) throw RunError("describe: expected a JSON array of option objects");
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < arr.a.size(); ++i) {
        const PlanSpec ps = build_plan(impl_->spec->g, to_plan_options(FusedOptions::from_json_obj(arr.a[i])), false);
        o << (i ? "," : "") << "{\"kernels\":[";
        for (size_t k = 0; k < ps.kernels.size(); ++k) {
            const KernelSpec& ks = ps.kernels[k];
            o << (k ? "," : "") << "{\"nodes\":" << ks.nodes.size() << ",\"inputs\":" << ks.inputs.size()
              << ",\"outputs\":" << ks.outputs.size() << ",\"pops\":" << ks.pops.size()
              << ",\"state_words\":" << ks.n_state << ",\"rows_per_step\":" << ks.rows_per_step
              << ",\"occupancy\":" << ks.occupancy << ",\"words_per_lane\":" << ks.words_per_lane << "}";
        }
        o << "],\"slots\":" << ps.n_slots << ",\"chunk_rows\":" << ps.opt.chunk_rows;
        if (const Json* d = arr.a[i].get("detail"); d && d->as_int(0)) {
            // the whole plan, for checking a partition on the host
            const SegGraph& g = impl_->spec->g;
            o << ",\"graph\":[";
            for (size_t n = 0; n < g.nodes.size(); ++n)
                o << (n ? "," : "") << "[" << (int)g.nodes[n].op << "," << g.nodes[n].a << "," << g.nodes[n].b << ","
                  << g.nodes[n].k << "," << g.nodes[n].src << "]";
            o << "],\"graph_outputs\":[";
            for (size_t n = 0; n < g.outputs.size(); ++n) o << (n ? "," : "") << g.outputs[n];
            o << "],\"named_outputs\":[";
            for (size_t n = 0; n < impl_->outs.size(); ++n)
                o << (n ? "," : "") << "[\"" << impl_->outs[n].name << "\"," << impl_->outs[n].ref << "]";
            o << "],\"sources\":[";
            for (size_t n = 0; n < impl_->sources.size(); ++n)
                o << (n ? "," : "") << "[\"" << impl_->sources[n].first << "\"," << impl_->sources[n].second << "]";
            o << "],\"kernel_detail\":[";
            for (size_t k = 0; k < ps.kernels.size(); ++k) {
                const KernelSpec& ks = ps.kernels[k];
                o << (k ? "," : "") << "{\"nodes\":[";
                for (size_t j = 0; j < ks.nodes.size(); ++j) o << (j ? "," : "") << ks.nodes[j];
                o << "],\"inputs\":[";
                for (size_t j = 0; j < ks.inputs.size(); ++j)
                    o << (j ? "," : "") << "[" << (int)ks.inputs[j].kind << "," << ks.inputs[j].id << "," << ks.input_node[j] << "]";
                o << "],\"outputs\":[";
                for (size_t j = 0; j < ks.outputs.size(); ++j)
                    o << (j ? "," : "") << "[" << ks.outputs[j] << "," << ks.out_slot[j] << "]";
                o << "]}";
            }
            o << "]";
        }
        o << "}";
    }
    o << "]";
    return o.str();
}

std::string FusedPlan::info_json() const {
    const Impl& I = *impl_;
    std::ostringstream o;
    o << "{\"kernels\":[";
    for (size_t k = 0; k < I.spec->kernels.size(); ++k) {
        const KernelSpec& ks = I.spec->kernels[k];
        o << (k ? "," : "") << "{\"name\":\"" << ks.name << "\",\"entry\":\"" << ks.entry
          << "\",\"nodes\":" << ks.nodes.size()
          << ",\"inputs\":" << ks.inputs.size() << ",\"outputs\":" << ks.outputs.size() << ",\"pops\":" << ks.pops.size()
          << ",\"state_words\":" << ks.n_state << ",\"rows_per_step\":" << ks.rows_per_step
          << ",\"occupancy\":" << ks.occupancy << ",\"words_per_lane\":" << ks.words_per_lane
          << ",\"interleave\":" << (ks.interleave ? 1 : 0);
        if (I.mod && k < I.mod->scratch.size()) o << ",\"scratch_bytes\":" << I.mod->scratch[k];
        if (I.mod && k < I.mod->regs.size()) {
            o << ",\"vgprs\":" << I.mod->regs[k].vgprs << ",\"agprs\":" << I.mod->regs[k].agprs;
        } else if (k < I.compiled_regs.size() && I.compiled_regs[k].vgprs >= 0) {
            // compiled here but not loaded: the metadata's private segment
            // size is the per-lane scratch the loader would report
            const KernelRegs& r = I.compiled_regs[k];
            o << ",\"scratch_bytes\":" << r.scratch << ",\"vgprs\":" << r.vgprs << ",\"agprs\":" << r.agprs;
        }
        o << "}";
    }
    o << "],\"sources\":[";
    for (size_t j = 0; j < I.sources.size(); ++j)
        o << (j ? "," : "") << "[\"" << I.sources[j].first << "\"," << I.sources[j].second << "]";
    o << "],\"outputs\":[";
    for (size_t j = 0; j < I.outs.size(); ++j)
        o << (j ? "," : "") << "[\"" << I.outs[j].name << "\"," << I.outs[j].idx << "," << I.outs[j].kind << "]";
    o << "],\"ints\":[";
    for (size_t j = 0; j < I.ints.size(); ++j) o << (j ? "," : "") << "\"" << I.ints[j].name << "\"";
    o << "],\"slots\":" << I.spec->n_slots << ",\"graph_nodes\":" << I.spec->g.nodes.size()
      << ",\"op_count\":" << I.op_count << ",\"compile_ms\":" << I.rec->compile_ms()
      << ",\"chunk_rows\":" << I.spec->opt.chunk_rows << ",\"warmup_rows\":" << I.spec->opt.warmup_rows
      << ",\"waves_per_block\":" << I.spec->opt.waves_per_block << ",\"rows_per_step\":" << I.spec->opt.rows_per_step
      << ",\"max_ops\":" << I.spec->opt.max_ops << ",\"max_live\":" << I.spec->opt.max_live
      << ",\"occupancy\":" << I.spec->opt.occupancy << ",\"persistent\":" << (I.spec->opt.persistent ? 1 : 0)
      << ",\"agprs\":" << (I.spec->opt.agprs ? 1 : 0)
      << ",\"n_chunks\":" << (I.rt ? I.rt->n_chunks : 0) << ",\"device_bytes\":" << (I.rt ? I.rt->device_bytes() : 0)
      << "}";
    return o.str();
}

const std::string& FusedPlan::module_source() const { return impl_->spec->module_source; }
int64_t FusedPlan::op_count() const { return impl_->op_count; }
int FusedPlan::n_kernels() const { return (int)impl_->spec->kernels.size(); }

}  // namespace bsz

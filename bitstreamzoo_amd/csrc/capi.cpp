// capi.cpp — the extern "C" boundary declared in include/bsz_hip.h.
#include "../../include/bsz_hip.h"

#include "bsdata_io.h"
#include "device.h"
#include "executor.h"
#include "fused.h"
#include "frontend.h"
#include "mlp.h"
#include "run_json.h"

#include <chrono>
#include <cstring>
#include <new>

using namespace bsz;

struct bsz_ctx {
    int device = 0;
    hipStream_t s = nullptr;
    std::string err;
    DevPool pool;
    std::unordered_map<uint64_t*, size_t> bufs;
    std::unique_ptr<Engine> perop, fused;
    void* add_state = nullptr;
    size_t add_state_bytes = 0;
    unsigned long long* slot = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipEvent_t order_ev = nullptr;   // recorded by bsz_stream_order on this context's stream
    bool capturing = false;          // between bsz_capture_begin and bsz_capture_end
    int64_t capture_launch0 = 0;     // kernel_launch_count() at bsz_capture_begin
};

struct bsz_program {
    std::unique_ptr<Program> prog;
    std::string source;
};

struct bsz_case {
    CaseData c;
};

struct bsz_result {
    RunOutcome out;
};

struct bsz_plan {
    bsz_ctx* ctx = nullptr;
    std::unique_ptr<FusedPlan> plan;
};

namespace {

thread_local std::string t_err;

int fail(bsz_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    t_err = msg;
    return code;
}

template <class F>
int guard(bsz_ctx* ctx, F&& f) {
    try {
        if (ctx) BSZ_HIP(hipSetDevice(ctx->device));
        f();
        return BSZ_OK;
    } catch (const ParseError& e) {
        return fail(ctx, BSZ_E_PARSE, e.what());
    } catch (const RunError& e) {
        return fail(ctx, BSZ_E_RUNTIME, e.what());
    } catch (const HipError& e) {
        std::string m = e.what();
        return fail(ctx, m.find("out of memory") != std::string::npos ? BSZ_E_OOM : BSZ_E_HIP, m);
    } catch (const std::bad_alloc&) {
        return fail(ctx, BSZ_E_OOM, "host out of memory");
    } catch (const std::invalid_argument& e) {
        return fail(ctx, BSZ_E_INVALID, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, BSZ_E_RUNTIME, e.what());
    }
}

Geom geom_of(int64_t bitlength) { return Geom::make(bitlength, 1); }

// An op enqueued while the context captures a graph: refused once the graph
// would exceed kMaxGraphLaunches launches (each op launches at most one kernel).
void capture_gate(bsz_ctx* ctx) {
    if (ctx->capturing && kernel_launch_count() - ctx->capture_launch0 >= kMaxGraphLaunches)
        throw RunError("graph capture refused: more than " + std::to_string(kMaxGraphLaunches) +
                       " kernel launches in one graph");
}

void ensure_scalar_state(bsz_ctx* ctx, const Geom& g) {
    size_t need = AddScanState::bytes(g);
    if (need > ctx->add_state_bytes) {
        if (ctx->add_state) { BSZ_HIP(hipStreamSynchronize(ctx->s)); BSZ_HIP(hipFree(ctx->add_state)); }
        BSZ_HIP(hipMalloc(&ctx->add_state, need));
        BSZ_HIP(hipMemsetAsync(ctx->add_state, 0, need, ctx->s));
        ctx->add_state_bytes = need;
    }
    if (!ctx->slot) BSZ_HIP(hipMalloc(&ctx->slot, 8 * kPopSlotWords));
}

}  // namespace

extern "C" {

int bsz_ctx_create(bsz_ctx** out, int device) {
    if (!out) return fail(nullptr, BSZ_E_INVALID, "null out");
    auto* c = new (std::nothrow) bsz_ctx;
    if (!c) return fail(nullptr, BSZ_E_OOM, "host out of memory");
    c->device = device;
    int rc = guard(c, [&] {
        BSZ_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        setup_noalloc_scratch(device);  // once per device, with the device current
    });
    if (rc != BSZ_OK) { t_err = c->err; delete c; return rc; }
    *out = c;
    return BSZ_OK;
}

int bsz_ctx_destroy(bsz_ctx* ctx) {
    if (!ctx) return BSZ_OK;
    int rc = guard(ctx, [&] {
        BSZ_HIP(hipStreamSynchronize(ctx->s));
        ctx->perop.reset();
        ctx->fused.reset();
        for (auto& kv : ctx->bufs) ctx->pool.put(kv.first, kv.second);
        ctx->bufs.clear();
        ctx->pool.trim();
        if (ctx->add_state) (void)hipFree(ctx->add_state);
        if (ctx->slot) (void)hipFree(ctx->slot);
        if (ctx->t0) { (void)hipEventDestroy(ctx->t0); (void)hipEventDestroy(ctx->t1); }
        if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
        (void)hipStreamDestroy(ctx->s);
    });
    delete ctx;
    return rc;
}

const char* bsz_last_error(const bsz_ctx* ctx) { return ctx ? ctx->err.c_str() : t_err.c_str(); }

int bsz_sync(bsz_ctx* ctx) { return guard(ctx, [&] { BSZ_HIP(hipStreamSynchronize(ctx->s)); }); }

void* bsz_ctx_stream(bsz_ctx* ctx) { return ctx ? (void*)ctx->s : nullptr; }

uint64_t bsz_stream_alloc_words(int64_t bitlength) { return geom_of(bitlength).alloc_words(); }

int bsz_buf_alloc(bsz_ctx* ctx, int64_t bitlength, uint64_t** dev_out) {
    return guard(ctx, [&] {
        Geom g = geom_of(bitlength);
        auto* p = (uint64_t*)ctx->pool.get(g.bytes());
        BSZ_HIP(hipMemsetAsync(p, 0, g.bytes(), ctx->s));
        ctx->bufs[p] = g.bytes();
        *dev_out = p;
    });
}

int bsz_buf_free(bsz_ctx* ctx, uint64_t* dev) {
    return guard(ctx, [&] {
        auto it = ctx->bufs.find(dev);
        if (it == ctx->bufs.end()) throw std::invalid_argument("bsz_buf_free: unknown buffer");
        BSZ_HIP(hipStreamSynchronize(ctx->s));
        ctx->pool.put(it->first, it->second);
        ctx->bufs.erase(it);
    });
}

int bsz_h2d(bsz_ctx* ctx, uint64_t* dev, const uint64_t* host, uint64_t n, int64_t bitlength) {
    return guard(ctx, [&] {
        upload_stream(ctx->s, dev, host, n, geom_of(bitlength));
        BSZ_HIP(hipStreamSynchronize(ctx->s));
    });
}

int bsz_d2h(bsz_ctx* ctx, uint64_t* host, const uint64_t* dev, uint64_t n) {
    return guard(ctx, [&] {
        stage_d2h(ctx->s, host, dev, n * 8);
    });
}

int bsz_op_fill(bsz_ctx* ctx, uint64_t* dst, int ones, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_fill(ctx->s, dst, ones != 0, geom_of(bl)); });
}
int bsz_op_copy(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_pointwise(ctx->s, PwOp::Copy, dst, src, nullptr, geom_of(bl)); });
}
int bsz_op_not(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_pointwise(ctx->s, PwOp::Not, dst, a, nullptr, geom_of(bl)); });
}
int bsz_op_and(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_pointwise(ctx->s, PwOp::And, dst, a, b, geom_of(bl)); });
}
int bsz_op_or(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_pointwise(ctx->s, PwOp::Or, dst, a, b, geom_of(bl)); });
}
int bsz_op_xor(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bl) {
    return guard(ctx, [&] { capture_gate(ctx); launch_pointwise(ctx->s, PwOp::Xor, dst, a, b, geom_of(bl)); });
}
int bsz_op_add(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bl) {
    return guard(ctx, [&] {
        if (dst == a || dst == b) throw std::invalid_argument("bsz_op_add: dst aliases an operand");
        capture_gate(ctx);
        Geom g = geom_of(bl);
        ensure_scalar_state(ctx, g);
        launch_add(ctx->s, dst, a, b, g, ctx->add_state);
    });
}
int bsz_op_add_noalloc(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bl) {
    return guard(ctx, [&] {
        if (dst == a || dst == b) throw std::invalid_argument("bsz_op_add_noalloc: dst aliases an operand");
        capture_gate(ctx);
        launch_add_noalloc(ctx->s, ctx->device, dst, a, b, geom_of(bl));
    });
}
int bsz_op_shl(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, uint64_t k, int64_t bl) {
    return guard(ctx, [&] {
        if (dst == src) throw std::invalid_argument("bsz_op_shl: dst aliases src");
        capture_gate(ctx);
        launch_shl(ctx->s, dst, src, k, geom_of(bl));
    });
}
int bsz_op_shr(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, uint64_t k, int64_t bl) {
    return guard(ctx, [&] {
        if (dst == src) throw std::invalid_argument("bsz_op_shr: dst aliases src");
        capture_gate(ctx);
        launch_shr(ctx->s, dst, src, k, geom_of(bl));
    });
}
int bsz_op_popcount(bsz_ctx* ctx, const uint64_t* src, int64_t bl, int64_t* out) {
    return guard(ctx, [&] {
        Geom g = geom_of(bl);
        ensure_scalar_state(ctx, g);
        launch_popcount(ctx->s, ctx->slot, src, g);
        *out = (int64_t)read_popcount_slots(ctx->s, ctx->slot);
    });
}
int bsz_op_is_nonzero(bsz_ctx* ctx, const uint64_t* src, int64_t bl, int* out) {
    return guard(ctx, [&] {
        Geom g = geom_of(bl);
        ensure_scalar_state(ctx, g);
        launch_zero_u64(ctx->s, ctx->slot, 1);
        launch_nonzero(ctx->s, ctx->slot, src, g);
        unsigned long long v = 0;
        BSZ_HIP(hipMemcpyAsync(&v, ctx->slot, 8, hipMemcpyDeviceToHost, ctx->s));
        BSZ_HIP(hipStreamSynchronize(ctx->s));
        *out = v ? 1 : 0;
    });
}

int bsz_program_parse(const char* source, bsz_program** out) {
    return guard(nullptr, [&] {
        auto p = std::make_unique<bsz_program>();
        p->source = source;
        p->prog = parse_program(p->source);
        *out = p.release();
    });
}

int bsz_program_free(bsz_program* p) { delete p; return BSZ_OK; }

int64_t bsz_program_static_ops(const bsz_program* p, const bsz_case* c) {
    return count_static_ops(*p->prog, c->c.params);
}

int bsz_case_create(bsz_case** out, int64_t bitlength) {
    auto* c = new (std::nothrow) bsz_case;
    if (!c) return fail(nullptr, BSZ_E_OOM, "host out of memory");
    c->c.bitlength = bitlength;
    *out = c;
    return BSZ_OK;
}
int bsz_case_free(bsz_case* c) { delete c; return BSZ_OK; }
int bsz_case_set_param(bsz_case* c, const char* n, int64_t v) { c->c.params[n] = v; return BSZ_OK; }
int bsz_case_set_input(bsz_case* c, const char* n, const uint64_t* w, uint64_t k) {
    c->c.inputs[n].assign(w, w + k);
    return BSZ_OK;
}
int bsz_case_set_array_elem(bsz_case* c, const char* n, int64_t idx, const uint64_t* w, uint64_t k) {
    c->c.input_arrays[n][idx].assign(w, w + k);
    return BSZ_OK;
}
int bsz_case_set_broadcast(bsz_case* c, const char* n, const uint8_t* bits, int64_t count) {
    std::vector<bool> v((size_t)std::max<int64_t>(0, count));
    for (int64_t i = 0; i < count; ++i) v[i] = (bits[i / 8] >> (i % 8)) & 1;
    c->c.broadcast[n] = std::move(v);
    return BSZ_OK;
}
int bsz_case_set_iterations(bsz_case* c, int it) { c->c.iterations = it; return BSZ_OK; }
int bsz_case_set_padding(bsz_case* c, int word_bits) {
    if (!c || (word_bits != 0 && word_bits != 32 && word_bits != 64))
        return fail(nullptr, BSZ_E_INVALID, "bsz_case_set_padding: word_bits must be 0, 32 or 64");
    c->c.bsim_padding = word_bits;
    return BSZ_OK;
}
int bsz_case_set_feedback(bsz_case* c, const char* o, const char* i) { c->c.feedback[o] = i; return BSZ_OK; }

int bsz_case_read_bsdata(const char* path, bsz_case*** cases_out, int* n_out) {
    try {
        auto v = read_bsdata_file(path);
        auto** arr = new bsz_case*[v.size() ? v.size() : 1];
        for (size_t i = 0; i < v.size(); ++i) {
            arr[i] = new bsz_case;
            arr[i]->c = std::move(v[i]);
        }
        *cases_out = arr;
        *n_out = (int)v.size();
        return BSZ_OK;
    } catch (const std::exception& e) {
        return fail(nullptr, BSZ_E_IO, e.what());
    }
}

int bsz_case_array_free(bsz_case** cases, int n) {
    for (int i = 0; i < n; ++i) delete cases[i];
    delete[] cases;
    return BSZ_OK;
}

int bsz_run(bsz_ctx* ctx, const bsz_program* p, const bsz_case* c, int executor, bsz_result** out) {
    return guard(ctx, [&] {
        Engine* eng = nullptr;
        if (executor == BSZ_EXEC_FUSED) {
            if (!ctx->fused) ctx->fused = make_fused_engine(ctx->s, FusedOptions::from_env());
            eng = ctx->fused.get();
        } else {
            if (!ctx->perop) ctx->perop = make_perop_engine(ctx->s);
            eng = ctx->perop.get();
        }
        auto r = std::make_unique<bsz_result>();
        r->out = run_case_iterations(*p->prog, c->c, *eng, 0.0);
        *out = r.release();
    });
}

int bsz_result_free(bsz_result* r) { delete r; return BSZ_OK; }

int bsz_result_stream(const bsz_result* r, const char* name, uint64_t* out, uint64_t n) {
    auto it = r->out.res.streams.find(name);
    if (it == r->out.res.streams.end()) return 0;
    for (uint64_t i = 0; i < n; ++i) out[i] = i < it->second.size() ? it->second[i] : 0;
    return 1;
}

int64_t bsz_result_array_count(const bsz_result* r, const char* name) {
    auto it = r->out.res.arrays.find(name);
    return it == r->out.res.arrays.end() ? 0 : (int64_t)it->second.size();
}

int bsz_result_array_indices(const bsz_result* r, const char* name, int64_t* idx, int64_t cap) {
    auto it = r->out.res.arrays.find(name);
    if (it == r->out.res.arrays.end()) return 0;
    int64_t k = 0;
    for (auto& e : it->second) { if (k < cap) idx[k] = e.first; ++k; }
    return (int)k;
}

int bsz_result_array_elem(const bsz_result* r, const char* name, int64_t idx, uint64_t* out, uint64_t n) {
    auto it = r->out.res.arrays.find(name);
    if (it == r->out.res.arrays.end()) return 0;
    auto e = it->second.find(idx);
    if (e == it->second.end()) return 0;
    for (uint64_t i = 0; i < n; ++i) out[i] = i < e->second.size() ? e->second[i] : 0;
    return 1;
}

int bsz_result_int(const bsz_result* r, const char* name, int64_t* out) {
    auto it = r->out.res.ints.find(name);
    if (it == r->out.res.ints.end()) return 0;
    *out = it->second;
    return 1;
}

int64_t bsz_result_op_count(const bsz_result* r) { return r->out.dynamic_ops; }
uint64_t bsz_result_n_words(const bsz_result* r) { return r->out.n_words; }

int64_t bsz_result_json(const bsz_result* r, char* buf, uint64_t cap) {
    std::string j = result_json(r->out);
    if (buf && cap) {
        size_t k = std::min<size_t>(j.size(), cap - 1);
        std::memcpy(buf, j.data(), k);
        buf[k] = 0;
    }
    return (int64_t)j.size();
}


static int64_t put_str(const std::string& j, char* buf, uint64_t cap) {
    if (buf && cap) {
        size_t k = std::min<size_t>(j.size(), cap - 1);
        std::memcpy(buf, j.data(), k);
        buf[k] = 0;
    }
    return (int64_t)j.size();
}

int bsz_plan_create(bsz_ctx* ctx, const bsz_program* p, const bsz_case* shape, const char* options_json,
                    bsz_plan** out) {
    return guard(ctx, [&] {
        if (!ctx || !p || !shape || !out) throw std::invalid_argument("bsz_plan_create: null argument");
        auto h = std::make_unique<bsz_plan>();
        h->ctx = ctx;
        const FusedOptions o = FusedOptions::from_json(options_json ? options_json : "");
        if (o.device_if) throw std::invalid_argument("device_if is an inspection option of bsz_plan_compile_only");
        h->plan = FusedPlan::record(ctx->s, *p->prog, shape->c, o, true);
        *out = h.release();
    });
}

int64_t bsz_plan_compile_only(const bsz_program* p, const bsz_case* shape, const char* options_json, char* info,
                              uint64_t cap) {
    std::string j;
    int rc = guard(nullptr, [&] {
        const FusedOptions o = FusedOptions::from_json(options_json ? options_json : "");
        auto fp = FusedPlan::record(nullptr, *p->prog, shape->c, o, o.compile != 0);
        j = fp->info_json();
    });
    if (rc != BSZ_OK) return rc;
    return put_str(j, info, cap);
}

int64_t bsz_plan_describe(const bsz_program* p, const bsz_case* shape, const char* options_array_json, char* out,
                          uint64_t cap) {
    std::string j;
    int rc = guard(nullptr, [&] {
        FusedOptions o = FusedOptions::from_env();
        auto fp = FusedPlan::record(nullptr, *p->prog, shape->c, o, false);
        j = fp->describe_json(options_array_json ? options_array_json : "[]");
    });
    if (rc != BSZ_OK) return rc;
    return put_str(j, out, cap);
}

int bsz_plan_free(bsz_plan* plan) {
    if (!plan) return BSZ_OK;
    int rc = guard(plan->ctx, [&] { plan->plan.reset(); });
    delete plan;
    return rc;
}

int bsz_plan_set_instances(bsz_plan* plan, int n, const int64_t* bitlengths) {
    return guard(plan->ctx, [&] { plan->plan->set_instances(std::vector<int64_t>(bitlengths, bitlengths + n)); });
}

int bsz_plan_bind_input(bsz_plan* plan, int inst, const char* name, int64_t idx, const uint64_t* dev) {
    return guard(plan->ctx, [&] { plan->plan->bind_input(inst, name, idx, dev); });
}

int bsz_plan_output(bsz_plan* plan, int inst, const char* name, int64_t idx, uint64_t** dev_out) {
    return guard(plan->ctx, [&] { *dev_out = plan->plan->output(inst, name, idx); });
}

int bsz_plan_run(bsz_plan* plan) { return guard(plan->ctx, [&] { plan->plan->run(); }); }
int bsz_plan_capture(bsz_plan* plan) { return guard(plan->ctx, [&] { plan->plan->capture(); }); }
int bsz_plan_replay(bsz_plan* plan, int times) {
    return guard(plan->ctx, [&] { for (int i = 0; i < times; ++i) plan->plan->replay(); });
}

int bsz_plan_read_int(bsz_plan* plan, int inst, const char* name, int64_t* out) {
    return guard(plan->ctx, [&] { *out = plan->plan->read_int(inst, name, true); });
}

int bsz_plan_mismatches(bsz_plan* plan, int kernel, int64_t* out) {
    return guard(plan->ctx, [&] { *out = plan->plan->mismatches(kernel); });
}

int bsz_plan_unproven(bsz_plan* plan, int kernel, int64_t* out) {
    return guard(plan->ctx, [&] { *out = plan->plan->unproven(kernel); });
}

int bsz_plan_set_executor(bsz_plan* plan, int executor) {
    return guard(plan->ctx, [&] { plan->plan->set_executor(executor == BSZ_EXEC_PEROP ? 0 : 1); });
}

int bsz_plan_prepare(bsz_plan* plan) { return guard(plan->ctx, [&] { plan->plan->prepare_now(); }); }

int bsz_timer_start(bsz_ctx* ctx) {
    return guard(ctx, [&] {
        if (!ctx->t0) { BSZ_HIP(hipEventCreate(&ctx->t0)); BSZ_HIP(hipEventCreate(&ctx->t1)); }
        BSZ_HIP(hipEventRecord(ctx->t0, ctx->s));
    });
}

int bsz_timer_stop(bsz_ctx* ctx, double* ms) {
    return guard(ctx, [&] {
        if (!ctx->t0) throw std::invalid_argument("bsz_timer_stop without bsz_timer_start");
        BSZ_HIP(hipEventRecord(ctx->t1, ctx->s));
        BSZ_HIP(hipEventSynchronize(ctx->t1));
        float f = 0;
        BSZ_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
        *ms = f;
    });
}

int bsz_capture_begin(bsz_ctx* ctx) {
    return guard(ctx, [&] {
        BSZ_HIP(hipStreamBeginCapture(ctx->s, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        ctx->capture_launch0 = kernel_launch_count();
    });
}

int bsz_capture_end(bsz_ctx* ctx, void** graph_exec) {
    return guard(ctx, [&] {
        ctx->capturing = false;
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        BSZ_HIP(hipStreamEndCapture(ctx->s, &g));
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        *graph_exec = ex;
    });
}

int bsz_stream_order(bsz_ctx* from, bsz_ctx* to) {
    return guard(to, [&] {
        if (!from || from == to) throw std::invalid_argument("bsz_stream_order: needs two different contexts");
        if (from->device != to->device) throw std::invalid_argument("bsz_stream_order: contexts on different devices");
        if (!from->order_ev) BSZ_HIP(hipEventCreateWithFlags(&from->order_ev, hipEventDisableTiming));
        // the wait takes the event as recorded at this point; a later record
        // on `from` does not change what `to` waits for
        BSZ_HIP(hipEventRecord(from->order_ev, from->s));
        BSZ_HIP(hipStreamWaitEvent(to->s, from->order_ev, 0));
    });
}

int bsz_graph_launch(bsz_ctx* ctx, void* graph_exec, int times) {
    return guard(ctx, [&] {
        for (int i = 0; i < times; ++i) BSZ_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, ctx->s));
    });
}

int bsz_graph_destroy(void* graph_exec) {
    if (graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
    return BSZ_OK;
}

int64_t bsz_plan_info(const bsz_plan* plan, char* buf, uint64_t cap) { return put_str(plan->plan->info_json(), buf, cap); }
int64_t bsz_plan_source(const bsz_plan* plan, char* buf, uint64_t cap) {
    return put_str(plan->plan->module_source(), buf, cap);
}

}  // extern "C"

// ── schedule cost model ──────────────────────────────────────────────────

struct bsz_mlp {
    bsz_ctx* ctx = nullptr;
    std::unique_ptr<Mlp> m;
};

extern "C" {

int bsz_mlp_create(bsz_ctx* ctx, int n_in, int h1, int h2, uint64_t seed, bsz_mlp** out) {
    if (!ctx || !out) return fail(ctx, BSZ_E_INVALID, "null argument");
    return guard(ctx, [&] {
        auto* m = new bsz_mlp;
        m->ctx = ctx;
        try {
            m->m = std::make_unique<Mlp>(ctx->s, n_in, h1, h2, seed);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int bsz_mlp_free(bsz_mlp* m) {
    if (!m) return BSZ_OK;
    int rc = guard(m->ctx, [&] { m->m.reset(); });
    delete m;
    return rc;
}

int64_t bsz_mlp_n_params(const bsz_mlp* m) { return m ? (int64_t)m->m->n_params() : 0; }

int bsz_mlp_get_params(bsz_mlp* m, float* out) {
    return guard(m->ctx, [&] { m->m->get_params(out); });
}

int bsz_mlp_set_params(bsz_mlp* m, const float* in) {
    return guard(m->ctx, [&] { m->m->set_params(in); });
}

int bsz_mlp_predict(bsz_mlp* m, const float* x, int n, float* out) {
    return guard(m->ctx, [&] { m->m->predict(x, n, out); });
}

int bsz_mlp_gradients(bsz_mlp* m, const float* x, const float* y, int n, float* grad, float* loss) {
    return guard(m->ctx, [&] {
        const float l = m->m->gradients(x, y, n, grad);
        if (loss) *loss = l;
    });
}

int bsz_mlp_train_step(bsz_mlp* m, const float* x, const float* y, int n, float lr, float* loss) {
    return guard(m->ctx, [&] {
        const float l = m->m->train_step(x, y, n, lr);
        if (loss) *loss = l;
    });
}

int bsz_mlp_train(bsz_mlp* m, const float* x, const float* y, int n, int epochs, int batch, float lr, uint64_t seed,
                  float* loss) {
    return guard(m->ctx, [&] {
        const float l = m->m->train(x, y, n, epochs, batch, lr, seed);
        if (loss) *loss = l;
    });
}

}  // extern "C"

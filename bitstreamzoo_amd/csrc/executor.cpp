// executor.cpp — program walk (control plane) and the one-kernel-per-op engine.
#include "executor.h"

#include <algorithm>
#include <chrono>
#include <set>

namespace bsz {

namespace {

constexpr int kMaxWhileIters = 10000;  // interpreter.h:109

const char* kOpNames[] = {"~", "&", "|", "^", "+", "<<", ">>", "popcount"};
enum OpIdx { OP_NOT, OP_AND, OP_OR, OP_XOR, OP_ADD, OP_SHL, OP_SHR, OP_POP, OP_N };

class Walker {
public:
    Walker(const Program& p, const CaseData& c, Engine& e) : P(p), C(c), E(e) {
        const size_t ns = P.syms.names.size();
        vars.resize(ns);
        ints.resize(ns);
        for (auto& kv : C.broadcast) {
            Sym s = P.syms.find(kv.first);
            if (s >= 0) bcast[s] = &kv.second;
        }
    }

    RunResult run() {
        walk();
        // The variable count is taken now; every stream that is not a program
        // output is then released, so the last flush materialises the outputs
        // alone (the same graph a recorded plan of the program builds) rather
        // than every variable still in scope.
        const int64_t nv = count_variables();
        release_non_outputs();
        E.flush();
        settle_predicated();
        return collect(nv);
    }

    RecordedOutputs record() {
        walk();
        RecordedOutputs r;
        for (Sym o : P.outputs) {
            const std::string& nm = P.syms.names[o];
            if (vars[o]) r.streams.emplace_back(nm, vars[o]);
            auto ai = arrays.find(o);
            if (ai != arrays.end())
                for (auto& e : ai->second) r.array_elems.emplace_back(nm, e.first, e.second);
        }
        for (Sym o : P.out_ints) {
            IntSlot& is = ints[o];
            if (!is.set) continue;
            RecordedOutputs::IntOut io;
            io.name = P.syms.names[o];
            io.pop = is.pop;
            io.value = is.v;
            r.ints.push_back(std::move(io));
        }
        r.op_count = ops;
        return r;
    }

private:
    void walk() {
        uint64_t max_words = 1;
        for (auto& kv : C.inputs) max_words = std::max<uint64_t>(max_words, kv.second.size());
        g = Geom::make(C.bitlength, max_words).bsim_padding(C.bsim_padding);
        E.begin(g);
        zero_c = E.constant(false);
        ones_c = E.constant(true);
        for (auto& kv : C.params) {
            Sym s = P.syms.find(kv.first);
            if (s >= 0) { ints[s].set = true; ints[s].v = kv.second; }
            else extra_params[kv.first] = kv.second;
        }
        load_inputs();
        for (const Stmt& s : P.stmts) exec(s);
    }

    int64_t count_variables() const {
        int64_t nv = 0;
        for (auto& v : vars) nv += v ? 1 : 0;
        for (auto& a : arrays) nv += (int64_t)a.second.size();
        return nv;
    }

    void release_non_outputs() {
        std::vector<char> keep(vars.size(), 0);
        for (Sym o : P.outputs) keep[o] = 1;
        for (size_t s = 0; s < vars.size(); ++s)
            if (!keep[s]) vars[s].reset();
        for (auto& a : arrays)
            if (!keep[a.first])
                for (auto& e : a.second) e.second.reset();
    }

    RunResult collect(int64_t nv) {
        RunResult r;
        for (Sym o : P.outputs) {
            const std::string& nm = P.syms.names[o];
            if (vars[o] && !absent_vars.count(o)) {
                std::vector<uint64_t> w(g.lw);
                E.download(vars[o], w.data());
                if (g.bitlength > 0) w[g.lw - 1] &= g.lastmask;  // padding bits are never output
                r.streams[nm] = std::move(w);
            }
            auto ai = arrays.find(o);
            if (ai != arrays.end()) {
                auto& dst = r.arrays[nm];
                for (auto& e : ai->second) {
                    if (absent_elems.count({o, e.first})) continue;
                    std::vector<uint64_t> w(g.lw);
                    E.download(e.second, w.data());
                    if (g.bitlength > 0) w[g.lw - 1] &= g.lastmask;
                    dst[e.first] = std::move(w);
                }
            }
        }
        for (Sym o : P.out_ints) {
            IntSlot& is = ints[o];
            if (!is.set) continue;
            r.ints[P.syms.names[o]] = value_of(is);
        }
        r.op_count = ops;
        for (int i = 0; i < OP_N; ++i)
            if (mix[i]) r.op_mix[kOpNames[i]] = mix[i];
        r.n_variables = nv - (int64_t)absent_vars.size() - (int64_t)absent_elems.size();
        return r;
    }

    struct IntSlot {
        bool set = false;
        int64_t v = 0;
        PopRef pop;  // non-null while the value is still a device popcount
    };

    const Program& P;
    const CaseData& C;
    Engine& E;
    Geom g;
    std::vector<SV> vars;
    std::unordered_map<Sym, std::map<int64_t, SV>> arrays;
    std::unordered_map<Sym, const std::vector<bool>*> bcast;
    std::vector<IntSlot> ints;
    std::unordered_map<std::string, int64_t> extra_params;
    SV zero_c, ones_c;
    int64_t ops = 0;
    int64_t mix[OP_N] = {};

    void count(OpIdx i) { ++ops; ++mix[i]; }

    // ── device-side `if` (if-conversion) ──
    // An `if` whose body only assigns streams (no int assignments, loops or
    // declarations of arrays) runs on the device when the engine supports
    // it: the body is walked unconditionally, every stream it assigns
    // becomes select(p, new, old), and the ops it counts, the variables and
    // elements it creates, count only where every enclosing predicate holds,
    // which the host learns once, after the final flush.  A body that raises
    // a run-time error is undone and the `if` is decided on the host instead
    // (the reference raises only where the body runs).
    struct Frame {
        PopRef pred;
        std::vector<std::pair<Sym, SV>> old_vars;                      // first old value (null: did not exist)
        std::vector<std::tuple<Sym, int64_t, SV, bool>> old_elems;     // (array, idx, old, existed)
        std::unordered_map<Sym, char> seen_vars;
        std::set<std::pair<Sym, int64_t>> seen_elems;
        size_t cond_created0 = 0, cond_ops0 = 0;
        int64_t ops0 = 0;
        int64_t mix0[OP_N] = {};
    };
    struct CondOps {
        std::vector<PopRef> preds;
        int64_t ops = 0;
        int64_t mix[OP_N] = {};
    };
    struct Created {   // a variable or element first created inside predicated frames
        Sym sym;
        int64_t idx;   // -1: a scalar variable
        std::vector<PopRef> preds;
    };
    std::vector<Frame> frames;
    std::vector<CondOps> cond_ops;
    std::vector<Created> cond_created;
    std::set<Sym> absent_vars;
    std::set<std::pair<Sym, int64_t>> absent_elems;

    std::vector<PopRef> pred_stack() const {
        std::vector<PopRef> v;
        for (const Frame& f : frames) v.push_back(f.pred);
        return v;
    }

    // Before vars[s] changes (or is created) inside predicated frames.
    void note_var(Sym s) {
        for (Frame& f : frames) {
            if (f.seen_vars.emplace(s, 1).second) f.old_vars.emplace_back(s, vars[s]);
        }
        if (!frames.empty() && !vars[s]) {
            cond_created.push_back({s, -1, pred_stack()});
            if ((size_t)s >= cond_only.size()) cond_only.resize(vars.size(), 0);
            cond_only[s] = 1;
        }
    }
    void note_elem(Sym a, int64_t idx, bool exists, const SV& cur) {
        for (Frame& f : frames) {
            if (f.seen_elems.insert({a, idx}).second) f.old_elems.emplace_back(a, idx, cur, exists);
        }
        if (!frames.empty() && !exists) cond_created.push_back({a, idx, pred_stack()});
    }

    static bool convertible(const std::vector<Stmt>& body) {
        for (const Stmt& b : body) {
            switch (b.k) {
                case Stmt::Assign: case Stmt::ArrAssign: case Stmt::Local: break;
                case Stmt::If: if (!convertible(b.body)) return false; break;
                default: return false;
            }
        }
        return true;
    }

    void exec_if(const Stmt& s) {
        SV c = var(P.sex[s.e].name);
        if (c == zero_c || g.bitlength <= 0) {  // never true on the device-side path's terms
            if (E.is_nonzero(c)) for (const Stmt& b : s.body) exec(b);
            return;
        }
        if (!E.supports_predicates() || !convertible(s.body) || c == ones_c) {
            if (E.is_nonzero(c)) for (const Stmt& b : s.body) exec(b);
            return;
        }
        Frame f;
        f.pred = E.predicate(c);
        f.ops0 = ops;
        std::copy(mix, mix + OP_N, f.mix0);
        f.cond_created0 = cond_created.size();
        f.cond_ops0 = cond_ops.size();
        frames.push_back(std::move(f));
        try {
            for (const Stmt& b : s.body) exec(b);
        } catch (const RunError&) {
            undo_frame();
            if (E.is_nonzero(c)) for (const Stmt& b : s.body) exec(b);
            return;
        }
        // ops counted directly in this body (nested bodies already moved
        // theirs out), conditional on every predicate up to this one
        CondOps co;
        co.preds = pred_stack();
        Frame fr = std::move(frames.back());
        co.ops = ops - fr.ops0;
        for (int i = 0; i < OP_N; ++i) co.mix[i] = mix[i] - fr.mix0[i];
        cond_ops.push_back(std::move(co));
        ops = fr.ops0;
        std::copy(fr.mix0, fr.mix0 + OP_N, mix);
        frames.pop_back();
        for (auto& ov : fr.old_vars) {
            Sym sv = ov.first;
            SV now = vars[sv];
            SV merged = E.select(fr.pred, now ? now : zero_c, ov.second ? ov.second : zero_c);
            for (Frame& outer : frames)  // the merge is an assignment of the enclosing bodies
                if (outer.seen_vars.emplace(sv, 1).second) outer.old_vars.emplace_back(sv, ov.second);
            vars[sv] = merged;
        }
        for (auto& oe : fr.old_elems) {
            Sym a = std::get<0>(oe);
            int64_t idx = std::get<1>(oe);
            const SV& old = std::get<2>(oe);
            SV& slot = arrays[a][idx];
            SV merged = E.select(fr.pred, slot ? slot : zero_c, old ? old : zero_c);
            for (Frame& outer : frames)
                if (outer.seen_elems.insert({a, idx}).second) outer.old_elems.emplace_back(a, idx, old, std::get<3>(oe));
            slot = merged;
        }
    }

    // Restores everything the innermost frame changed and drops it.
    void undo_frame() {
        Frame fr = std::move(frames.back());
        frames.pop_back();
        for (auto it = fr.old_vars.rbegin(); it != fr.old_vars.rend(); ++it) vars[it->first] = it->second;
        for (auto it = fr.old_elems.rbegin(); it != fr.old_elems.rend(); ++it) {
            Sym a = std::get<0>(*it);
            int64_t idx = std::get<1>(*it);
            if (std::get<3>(*it)) arrays[a][idx] = std::get<2>(*it);
            else {
                auto ai = arrays.find(a);
                if (ai != arrays.end()) ai->second.erase(idx);
            }
        }
        cond_created.resize(fr.cond_created0);
        cond_ops.resize(fr.cond_ops0);  // records of nested bodies that completed inside this one
        ops = fr.ops0;
        std::copy(fr.mix0, fr.mix0 + OP_N, mix);
    }

    bool all_true(const std::vector<PopRef>& ps) {
        for (const PopRef& p : ps)
            if (E.resolve(p) == 0) return false;
        return true;
    }

    // After the final flush: conditional op counts, and which conditionally
    // created variables and elements exist.
    void settle_predicated() {
        for (const CondOps& co : cond_ops) {
            if (!all_true(co.preds)) continue;
            ops += co.ops;
            for (int i = 0; i < OP_N; ++i) mix[i] += co.mix[i];
        }
        std::map<std::pair<Sym, int64_t>, bool> exists;
        for (const Created& c : cond_created) {
            bool& e = exists[{c.sym, c.idx}];
            e = e || all_true(c.preds);
        }
        for (auto& kv : exists) {
            if (kv.second || unconditional.count(kv.first)) continue;
            if (kv.first.second < 0) absent_vars.insert(kv.first.first);
            else absent_elems.insert(kv.first);
        }
    }
    // variables / elements written or created outside any predicated body
    // (they exist for certain), and the variables created only inside one so
    // far (a later unconditional read makes them exist too)
    std::set<std::pair<Sym, int64_t>> unconditional;
    std::vector<char> cond_only;

    int64_t value_of(IntSlot& is) {
        if (is.pop) {
            is.v = E.resolve(is.pop);
            is.pop.reset();
        }
        return is.v;
    }

    const std::string& nm(Sym s) const { return P.syms.names[s]; }

    SV& var(Sym s) {
        if (!vars[s]) {
            if (!frames.empty()) note_var(s);
            else unconditional.insert({s, -1});
            vars[s] = zero_c;
        } else if (frames.empty() && (size_t)s < cond_only.size() && cond_only[s]) {
            unconditional.insert({s, -1});
            cond_only[s] = 0;
        }
        return vars[s];
    }

    void set_var(Sym s, SV v) {
        if (!frames.empty()) note_var(s);
        else unconditional.insert({s, -1});
        vars[s] = std::move(v);
    }

    void load_inputs() {
        for (Sym s : P.inputs) {
            auto it = C.inputs.find(nm(s));
            if (it != C.inputs.end()) {
                E.note_input(nm(s), -1);
                vars[s] = E.input(it->second.data(), it->second.size());
                continue;
            }
            if (bcast.count(s)) continue;
            auto ai = C.input_arrays.find(nm(s));
            if (ai != C.input_arrays.end()) {
                auto& arr = arrays[s];
                for (auto& e : ai->second) {
                    E.note_input(nm(s), (int64_t)(int32_t)e.first);
                    arr[(int64_t)(int32_t)e.first] = E.input(e.second.data(), e.second.size());
                }
                continue;
            }
            vars[s] = zero_c;
        }
    }

    int64_t eval_int(int32_t id) {
        const IExpr& e = P.iex[id];
        switch (e.k) {
            case IExpr::Lit: return e.lit;
            case IExpr::Var: {
                IntSlot& is = ints[e.var];
                if (!is.set) throw RunError("undefined int variable: " + nm(e.var));
                return value_of(is);
            }
            case IExpr::Add: return (int64_t)((uint64_t)eval_int(e.l) + (uint64_t)eval_int(e.r));
            case IExpr::Sub: return (int64_t)((uint64_t)eval_int(e.l) - (uint64_t)eval_int(e.r));
            case IExpr::Mul: return (int64_t)((uint64_t)eval_int(e.l) * (uint64_t)eval_int(e.r));
            case IExpr::Pop: return E.resolve(popcount_of(e));
        }
        return 0;
    }

    PopRef popcount_of(const IExpr& e) {
        const SExpr& a = P.sex[e.arg];
        if (a.k != SExpr::Var) throw RunError("bad variant access");
        SV v = var(a.name);
        count(OP_POP);
        return E.popcount(v);
    }

    // Assignment of an int expression: a bare popcount stays on the device.
    void assign_int(Sym target, int32_t id) {
        const IExpr& e = P.iex[id];
        IntSlot is;
        is.set = true;
        if (e.k == IExpr::Pop) is.pop = popcount_of(e);
        else is.v = eval_int(id);
        ints[target] = std::move(is);
    }

    // get_operand (interpreter.inl:60-90): broadcast metadata first.
    SV operand(int32_t id) {
        const SExpr& e = P.sex[id];
        switch (e.k) {
            case SExpr::Var: return var(e.name);
            case SExpr::Zero: return zero_c;
            case SExpr::Ones: return ones_c;
            case SExpr::Arr: {
                int64_t idx = (int32_t)eval_int(e.idx);
                auto b = bcast.find(e.name);
                if (b != bcast.end()) {
                    if (idx >= 0 && idx < (int64_t)b->second->size()) return (*b->second)[idx] ? ones_c : zero_c;
                    return zero_c;
                }
                auto a = arrays.find(e.name);
                if (a == arrays.end()) throw RunError("undefined array: " + nm(e.name));
                auto it = a->second.find(idx);
                return it != a->second.end() ? it->second : zero_c;
            }
            default: throw RunError("get_operand: not a primary expression");
        }
    }

    // exec_rhs (interpreter.inl:93-157); a plain copy shares the value.
    SV rhs(int32_t id) {
        const SExpr& e = P.sex[id];
        switch (e.k) {
            case SExpr::Var: return var(e.name);
            case SExpr::Zero: return zero_c;
            case SExpr::Ones: return ones_c;
            case SExpr::Arr: {
                int64_t idx = (int32_t)eval_int(e.idx);
                auto a = arrays.find(e.name);
                if (a != arrays.end()) {
                    auto it = a->second.find(idx);
                    if (it != a->second.end()) return it->second;
                }
                auto b = bcast.find(e.name);
                if (b != bcast.end() && idx >= 0 && idx < (int64_t)b->second->size())
                    return (*b->second)[idx] ? ones_c : zero_c;
                return zero_c;
            }
            case SExpr::Not: {
                SV a = operand(e.a);
                count(OP_NOT);
                return E.op_not(a);
            }
            case SExpr::And: case SExpr::Or: case SExpr::Xor: case SExpr::Plus: {
                SV a = operand(e.a);
                SV b = operand(e.b);
                BinOp op = e.k == SExpr::And ? BinOp::And : e.k == SExpr::Or ? BinOp::Or
                         : e.k == SExpr::Xor ? BinOp::Xor : BinOp::Add;
                count(e.k == SExpr::And ? OP_AND : e.k == SExpr::Or ? OP_OR : e.k == SExpr::Xor ? OP_XOR : OP_ADD);
                return E.op_bin(op, a, b);
            }
            case SExpr::Shl: case SExpr::Shr: {
                SV a = operand(e.a);
                uint64_t k = (uint64_t)eval_int(e.idx);  // static_cast<size_t> (interpreter.inl:140)
                count(e.k == SExpr::Shl ? OP_SHL : OP_SHR);
                return E.op_shift(e.k == SExpr::Shl, a, k);
            }
        }
        return zero_c;
    }

    void exec(const Stmt& s) {
        switch (s.k) {
            case Stmt::Local:
                set_var(s.name, s.e >= 0 ? rhs(s.e) : zero_c);
                return;
            case Stmt::ArrDecl: {
                int64_t n = (int32_t)eval_int(s.i1);
                auto& arr = arrays[s.name];
                for (int64_t i = 0; i < n; ++i) arr[i] = zero_c;
                return;
            }
            case Stmt::Assign: {
                var(s.name);
                SV v = rhs(s.e);
                set_var(s.name, std::move(v));
                return;
            }
            case Stmt::ArrAssign: {
                int64_t idx = (int32_t)eval_int(s.i1);
                auto& arr = arrays[s.name];
                auto it = arr.find(idx);
                const bool existed = it != arr.end() && it->second;
                if (!frames.empty()) note_elem(s.name, idx, existed, existed ? it->second : SV());
                else unconditional.insert({s.name, idx});
                SV& slot = arr[idx];
                if (!slot) slot = zero_c;
                SV v = rhs(s.e);
                arrays[s.name][idx] = v;
                return;
            }
            case Stmt::If:
                exec_if(s);
                return;
            case Stmt::While: {
                Sym cs = P.sex[s.e].name;
                for (int it = 0; it < kMaxWhileIters; ++it) {
                    SV c = var(cs);
                    if (!E.is_nonzero(c)) return;
                    for (const Stmt& b : s.body) exec(b);
                }
                throw RunError("while loop exceeded max iterations");
            }
            case Stmt::For: {
                int32_t lo = (int32_t)eval_int(s.i1), hi = (int32_t)eval_int(s.i2);
                for (int32_t i = lo; i < hi; ++i) {
                    IntSlot& is = ints[s.name];
                    is.set = true; is.v = i; is.pop.reset();
                    for (const Stmt& b : s.body) exec(b);
                }
                return;
            }
            case Stmt::IntDecl: case Stmt::IntAssign:
                assign_int(s.name, s.i1);
                return;
        }
    }
};

// ── one kernel per op ──

struct PerOpPop : PopCell {
    unsigned long long* dev = nullptr;
    hipStream_t s = nullptr;
};

class PerOpEngine final : public Engine {
public:
    explicit PerOpEngine(hipStream_t s) : s_(s) {}
    ~PerOpEngine() override {
        (void)hipStreamSynchronize(s_);
        for (void* p : slot_blocks_) (void)hipFree(p);
        if (add_state_) (void)hipFree(add_state_);
    }

    void begin(const Geom& g) override {
        g_ = g;
        size_t need = AddScanState::bytes(g);
        if (need > add_state_bytes_) {
            if (add_state_) { BSZ_HIP(hipStreamSynchronize(s_)); BSZ_HIP(hipFree(add_state_)); }
            BSZ_HIP(hipMalloc(&add_state_, need));
            BSZ_HIP(hipMemsetAsync(add_state_, 0, need, s_));
            add_state_bytes_ = need;
        }
        zero_.reset();
        ones_.reset();
        launches0_ = kernel_launch_count();
    }

    SV constant(bool ones) override {
        SV& c = ones ? ones_ : zero_;
        if (!c) {
            c = fresh();
            launch_fill(s_, c->buf->ptr, ones, g_);
        }
        return c;
    }

    SV input(const uint64_t* host, uint64_t n) override {
        SV v = fresh();
        upload_stream(s_, v->buf->ptr, host, n, g_);
        BSZ_HIP(hipStreamSynchronize(s_));  // host words may be freed by the caller
        return v;
    }

    SV op_not(const SV& a) override {
        SV v = fresh();
        launch_pointwise(s_, PwOp::Not, v->buf->ptr, a->buf->ptr, nullptr, g_);
        return v;
    }

    SV op_bin(BinOp op, const SV& a, const SV& b) override {
        SV v = fresh();
        switch (op) {
            case BinOp::And: launch_pointwise(s_, PwOp::And, v->buf->ptr, a->buf->ptr, b->buf->ptr, g_); break;
            case BinOp::Or: launch_pointwise(s_, PwOp::Or, v->buf->ptr, a->buf->ptr, b->buf->ptr, g_); break;
            case BinOp::Xor: launch_pointwise(s_, PwOp::Xor, v->buf->ptr, a->buf->ptr, b->buf->ptr, g_); break;
            case BinOp::Add: launch_add(s_, v->buf->ptr, a->buf->ptr, b->buf->ptr, g_, add_state_); break;
        }
        return v;
    }

    SV op_shift(bool left, const SV& a, uint64_t k) override {
        SV v = fresh();
        if (left) launch_shl(s_, v->buf->ptr, a->buf->ptr, k, g_);
        else launch_shr(s_, v->buf->ptr, a->buf->ptr, k, g_);
        return v;
    }

    PopRef popcount(const SV& a) override {
        auto p = std::make_shared<PerOpPop>();
        p->dev = next_slots(kPopSlotWords);
        p->s = s_;
        launch_popcount(s_, p->dev, a->buf->ptr, g_);
        return p;
    }

    int64_t resolve(const PopRef& pr) override {
        if (!pr->ready) {
            auto* p = static_cast<PerOpPop*>(pr.get());
            p->value = (int64_t)read_popcount_slots(s_, p->dev);
            p->ready = true;
        }
        return pr->value;
    }

    bool is_nonzero(const SV& a) override {
        unsigned long long* slot = next_slots(1);
        launch_zero_u64(s_, slot, 1);
        launch_nonzero(s_, slot, a->buf->ptr, g_);
        unsigned long long v = 0;
        BSZ_HIP(hipMemcpyAsync(&v, slot, 8, hipMemcpyDeviceToHost, s_));
        BSZ_HIP(hipStreamSynchronize(s_));
        return v != 0;
    }

    void download(const SV& a, uint64_t* host) override {
        stage_d2h(s_, host, a->buf->ptr, g_.lw * 8);
    }

    void flush() override {}
    void sync() override { BSZ_HIP(hipStreamSynchronize(s_)); }
    int64_t n_kernel_launches() const override { return kernel_launch_count() - launches0_; }
    int64_t n_mallocs() const override { return pool_.n_mallocs(); }
    size_t peak_bytes() const override { return pool_.peak_bytes(); }
    hipStream_t stream() const override { return s_; }

private:
    hipStream_t s_;
    Geom g_;
    DevPool pool_;
    void* add_state_ = nullptr;
    size_t add_state_bytes_ = 0;
    std::vector<void*> slot_blocks_;
    size_t slot_used_ = 0;
    SV zero_, ones_;
    int64_t launches0_ = 0;
    static constexpr size_t kSlotsPerBlock = 1 << 16;

    SV fresh() {
        auto v = std::make_shared<StreamVal>();
        v->buf = std::make_shared<DevBuf>(&pool_, g_.bytes());
        return v;
    }

    // n consecutive device words for a popcount accumulator or a flag
    unsigned long long* next_slots(size_t n) {
        if (slot_blocks_.empty() || slot_used_ + n > kSlotsPerBlock) {
            void* p = nullptr;
            BSZ_HIP(hipMalloc(&p, kSlotsPerBlock * 8));
            slot_blocks_.push_back(p);
            slot_used_ = 0;
        }
        unsigned long long* r = reinterpret_cast<unsigned long long*>(slot_blocks_.back()) + slot_used_;
        slot_used_ += n;
        return r;
    }
};

}  // namespace

RunResult execute_program(const Program& prog, const CaseData& c, Engine& eng) {
    return Walker(prog, c, eng).run();
}

RecordedOutputs record_program(const Program& prog, const CaseData& c, Engine& eng) {
    return Walker(prog, c, eng).record();
}

std::unique_ptr<Engine> make_perop_engine(hipStream_t s) { return std::make_unique<PerOpEngine>(s); }

}  // namespace bsz

/*
 * bsz_hip.h — C ABI of the MI355X bitstream executor (libbsz_hip.so).
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * BSZ_OK (0) or a negative bsz_status; bsz_last_error() gives the message.
 * A context owns one HIP stream on one device; calls on a context are
 * asynchronous on that stream unless they return host data.  One context per
 * host thread.  Host pointers are borrowed for the duration of the call.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference tree).
 */
#ifndef BSZ_HIP_H
#define BSZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BSZ_OK = 0,
    BSZ_E_INVALID = -1,   /* bad argument / unknown name */
    BSZ_E_PARSE = -2,     /* .bs lexer, parser or 3-address error (main.cu:301-306) */
    BSZ_E_RUNTIME = -3,   /* bs::RuntimeError (interpreter.h:22-25), e.g. while limit */
    BSZ_E_HIP = -4,       /* HIP runtime error */
    BSZ_E_OOM = -5,       /* device allocation failed ("out of memory", base.py:219-223) */
    BSZ_E_IO = -6         /* .bsdata / file error (main.cu:267-270) */
} bsz_status;

typedef struct bsz_ctx bsz_ctx;
typedef struct bsz_program bsz_program;
typedef struct bsz_case bsz_case;
typedef struct bsz_result bsz_result;
typedef struct bsz_plan bsz_plan;

/* ── context ─────────────────────────────────────────────────────────── */

/* Replaces the per-case CudaBackend construction (main.cu:84-86). */
int bsz_ctx_create(bsz_ctx** out, int device);
int bsz_ctx_destroy(bsz_ctx* ctx);
/* Last error message of ctx (or of the calling thread when ctx is NULL). */
const char* bsz_last_error(const bsz_ctx* ctx);
int bsz_sync(bsz_ctx* ctx);
/* The context's hipStream_t, as an opaque pointer. */
void* bsz_ctx_stream(bsz_ctx* ctx);

/* ── device streams and one-kernel-per-op (the Backend concept,
 *    cuda_backend.cuh:16-229 / cpu_backend.h:13-151) ─────────────────── */

/* Words a stream of `bitlength` bits occupies on the device: ceil(bitlength/64)
 * rounded up to a whole 64-word row.  Replaces GpuStream::alloc sizing
 * (common.cuh:75-80, cuda_backend.cuh:23-25). */
uint64_t bsz_stream_alloc_words(int64_t bitlength);
/* alloc_stream (cuda_backend.cuh:39-44): zero-filled. */
int bsz_buf_alloc(bsz_ctx* ctx, int64_t bitlength, uint64_t** dev_out);
int bsz_buf_free(bsz_ctx* ctx, uint64_t* dev);
/* load_from_words (cuda_backend.cuh:156-172): n host words, truncated or
 * zero-padded to the stream and masked to the bitlength. */
int bsz_h2d(bsz_ctx* ctx, uint64_t* dev, const uint64_t* host, uint64_t n, int64_t bitlength);
/* store_to_words (cuda_backend.cuh:174-189): first n words to host (sync). */
int bsz_d2h(bsz_ctx* ctx, uint64_t* host, const uint64_t* dev, uint64_t n);
/* zero_stream / ones_stream (cuda_backend.cuh:46-51). */
int bsz_op_fill(bsz_ctx* ctx, uint64_t* dst, int ones, int64_t bitlength);
/* copy_stream (cuda_backend.cuh:45). */
int bsz_op_copy(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, int64_t bitlength);
/* op_not / op_and / op_or / op_xor (cuda_backend.cuh:53-68). */
int bsz_op_not(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, int64_t bitlength);
int bsz_op_and(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bitlength);
int bsz_op_or(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bitlength);
int bsz_op_xor(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bitlength);
/* op_add (cuda_backend.cuh:69-74 + add_carry.cuh:110-135): one kernel,
 * wave ballot carries + decoupled lookback.  dst must not alias a or b. */
int bsz_op_add(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bitlength);
/* op_add with no allocation: the same two kernels as bsz_op_add, with their
 * scratch (a count, a carry per 256-word segment and a bitmap) in a fixed
 * 4.3 MB device array of the library instead of memory bsz_op_add allocates
 * on first use.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under graph capture, serving loops).  bitlength must
 * be in [1, 2^34]; BSZ_E_INVALID otherwise.  Every call on one device shares
 * the array, so calls that overlap on a device race. */
int bsz_op_add_noalloc(bsz_ctx* ctx, uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t bitlength);
/* op_shl / op_shr (cuda_backend.cuh:75-84).  dst must not alias src. */
int bsz_op_shl(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, uint64_t k, int64_t bitlength);
int bsz_op_shr(bsz_ctx* ctx, uint64_t* dst, const uint64_t* src, uint64_t k, int64_t bitlength);
/* op_popcount / op_is_nonzero (cuda_backend.cuh:87-153), bounded to the
 * bitlength; synchronous, results to host. */
int bsz_op_popcount(bsz_ctx* ctx, const uint64_t* src, int64_t bitlength, int64_t* out);
int bsz_op_is_nonzero(bsz_ctx* ctx, const uint64_t* src, int64_t bitlength, int* out);

/* ── programs and cases (bs::parse, parser.cpp:451-457; BsdataInput,
 *    bsdata_reader.h:31-40) ─────────────────────────────────────────── */

int bsz_program_parse(const char* source, bsz_program** out);
int bsz_program_free(bsz_program* p);
/* count_ops_ast (ast_liveness.cpp:723-727) for the given case's params. */
int64_t bsz_program_static_ops(const bsz_program* p, const bsz_case* c);

int bsz_case_create(bsz_case** out, int64_t bitlength);
int bsz_case_free(bsz_case* c);
int bsz_case_set_param(bsz_case* c, const char* name, int64_t value);
int bsz_case_set_input(bsz_case* c, const char* name, const uint64_t* words, uint64_t n);
int bsz_case_set_array_elem(bsz_case* c, const char* name, int64_t idx, const uint64_t* words, uint64_t n);
/* count elements, one bit each (1 = ONES), LSB first (bsdata.py:96-102). */
int bsz_case_set_broadcast(bsz_case* c, const char* name, const uint8_t* bits, int64_t count);
int bsz_case_set_iterations(bsz_case* c, int iterations);
/* Value semantics of bits at or above the bitlength.  0 (default): clean,
 * as DSL.md:273-275 and the reference's Python simulator.  64 or 32: those of
 * bsim_cuda with --word-bits=64|32, whose streams are padded to a multiple of
 * 8 words (cuda_backend.cuh:23-25, common.cuh:35-45): ~, ONES, + and << leave
 * the padding bits set and >> brings them back into range.  Outputs,
 * popcounts and if/while conditions are bounded by the bitlength either way. */
int bsz_case_set_padding(bsz_case* c, int word_bits);
int bsz_case_set_feedback(bsz_case* c, const char* out_name, const char* in_name);
/* All cases of a .bsdata file (read_bsdata_multi, bsdata_reader.h:262-299). */
int bsz_case_read_bsdata(const char* path, bsz_case*** cases_out, int* n_out);
int bsz_case_array_free(bsz_case** cases, int n);

/* ── execution (Interpreter<Backend>::run, interpreter.inl:234-364, with
 *    the iterations/feedback loop of main.cu:84-140) ─────────────────── */

typedef enum { BSZ_EXEC_PEROP = 0, BSZ_EXEC_FUSED = 1 } bsz_executor;

int bsz_run(bsz_ctx* ctx, const bsz_program* p, const bsz_case* c, int executor, bsz_result** out);
int bsz_result_free(bsz_result* r);
/* Stream output: returns 1 and writes min(n, lw) words masked to the
 * bitlength, 0 when the program never produced it. */
int bsz_result_stream(const bsz_result* r, const char* name, uint64_t* out, uint64_t n);
int64_t bsz_result_array_count(const bsz_result* r, const char* name);
int bsz_result_array_indices(const bsz_result* r, const char* name, int64_t* idx, int64_t cap);
int bsz_result_array_elem(const bsz_result* r, const char* name, int64_t idx, uint64_t* out, uint64_t n);
int bsz_result_int(const bsz_result* r, const char* name, int64_t* out);
int64_t bsz_result_op_count(const bsz_result* r);
uint64_t bsz_result_n_words(const bsz_result* r);
/* The per-case JSON object bsim_cuda prints (main.cu:146-205); writes at most
 * cap bytes (NUL-terminated) and returns the full length. */
int64_t bsz_result_json(const bsz_result* r, char* buf, uint64_t cap);

/* ── fused plans: a straight-line program recorded once, then run over many
 *    instances (configuration batches of uneven lengths) in one launch per
 *    kernel; the reference runs one process per .bsdata file and one kernel
 *    per op (base.py:228-254, interpreter.inl:314-316), so these have no
 *    direct counterpart there ──────────────────────────────────────────── */

/* Records the program with `shape`'s params (every input it reads must be
 * present in `shape`, any length) and compiles it.  `options_json` is "" or a
 * flat object with chunk_rows, warmup_rows, waves_per_block, rows_per_step,
 * max_ops, max_live, occupancy (0 = automatic), words_per_lane (1 or 2),
 * store_order (1 = default), nontemporal (0, 1 stores, 2 loads and
 * stores), interleave, persistent, lazy_loads, agprs, join_warmup, bitop3
 * (1 = default: three-input boolean folding), order (0 = default: partition
 * in recorded order; 1: a list schedule that keeps few values live, so fewer
 * streams cross kernels) and spin_limit.  BSZ_E_RUNTIME
 * if the program needs host decisions (if/while, int expressions that read a
 * popcount). */
int bsz_plan_create(bsz_ctx* ctx, const bsz_program* p, const bsz_case* shape, const char* options_json,
                    bsz_plan** out);
/* Record + generate + hipRTC-compile only (no device needed); writes the plan
 * description JSON, with each compiled kernel's VGPR, AGPR and scratch use
 * read from its code object.  Option "compile":0 stops before hipRTC
 * (partition and kernel shapes only); "compile_shard"/"compile_shards" compile
 * a subset of the kernels, so several processes can fill the code-object cache
 * in parallel.  Returns the full length of the JSON; it may differ between
 * calls while other processes fill the cache, so read it from a call whose
 * `cap` exceeded the length. */
int64_t bsz_plan_compile_only(const bsz_program* p, const bsz_case* shape, const char* options_json,
                              char* info, uint64_t cap);
/* Kernel shapes (per kernel: nodes, inputs, outputs, popcounts, state words,
 * rows per step, occupancy) of the program under each schedule of a JSON
 * array of option objects; partitioning only (no code, no device).  Writes a
 * JSON array, one element per schedule.  Features for the cost model. */
int64_t bsz_plan_describe(const bsz_program* p, const bsz_case* shape, const char* options_array_json, char* out,
                          uint64_t cap);
int bsz_plan_free(bsz_plan* plan);
/* One instance per bitlength; allocates the plan's output and scratch buffers. */
int bsz_plan_set_instances(bsz_plan* plan, int n, const int64_t* bitlengths);
/* Binds a device stream (from bsz_buf_alloc, instance's bitlength) as input
 * `name` (array element `idx`, or -1) of instance `inst`. */
int bsz_plan_bind_input(bsz_plan* plan, int inst, const char* name, int64_t idx, const uint64_t* dev);
/* Device buffer of output `name` (array element idx or -1) of instance inst. */
int bsz_plan_output(bsz_plan* plan, int inst, const char* name, int64_t idx, uint64_t** dev_out);
/* Launch asynchronously on the context stream. */
int bsz_plan_run(bsz_plan* plan);
/* Capture one run into a hipGraph; bsz_plan_replay launches it `times` times. */
int bsz_plan_capture(bsz_plan* plan);
int bsz_plan_replay(bsz_plan* plan, int times);
/* Int output (a popcount) of instance inst after the last run (synchronous).
 * BSZ_E_RUNTIME when a chunk's look-back wait of that run reached its bound
 * (option "spin_limit"): the run's outputs are then not exact. */
int bsz_plan_read_int(bsz_plan* plan, int inst, const char* name, int64_t* out);
/* Chunks of kernel k whose speculative start state was wrong in the last run
 * (each ran again from its predecessor's end state, in the same launch). */
int bsz_plan_mismatches(bsz_plan* plan, int kernel, int64_t* out);
/* Chunks of kernel k whose warm-up rows did not prove their start state
 * exact in the last run (only these wait for their predecessor's end state). */
int bsz_plan_unproven(bsz_plan* plan, int kernel, int64_t* out);
/* BSZ_EXEC_FUSED (default) or BSZ_EXEC_PEROP: one kernel per DSL op over the
 * same recorded DAG and the same bound buffers (the plain comparator). */
int bsz_plan_set_executor(bsz_plan* plan, int executor);
/* Upload binding tables now (host sync), so that run/capture only launch. */
int bsz_plan_prepare(bsz_plan* plan);
int64_t bsz_plan_info(const bsz_plan* plan, char* buf, uint64_t cap);
int64_t bsz_plan_source(const bsz_plan* plan, char* buf, uint64_t cap);

/* ── device timing and graph capture on the context stream (hipEvent /
 *    hipGraph; replaces GpuTimer, gpu_timer.cuh:6-29) ───────────────────── */

int bsz_timer_start(bsz_ctx* ctx);
/* Elapsed milliseconds between the start event and now (synchronous). */
int bsz_timer_stop(bsz_ctx* ctx, double* ms);
int bsz_capture_begin(bsz_ctx* ctx);
int bsz_capture_end(bsz_ctx* ctx, void** graph_exec);
/* Orders the work enqueued after this call on `to`'s stream after the work
 * enqueued so far on `from`'s stream (an event recorded on `from`, waited on
 * by `to`; the host does not wait).  While `from` captures, this forks the
 * capture into `to`'s stream, and a later bsz_stream_order(to, from) joins
 * it back before bsz_capture_end(from): independent plans bound to
 * contexts of one device then run concurrently inside one graph.  No
 * counterpart in the reference, which runs one process per input file on
 * one stream (base.py:228-254). */
int bsz_stream_order(bsz_ctx* from, bsz_ctx* to);
int bsz_graph_launch(bsz_ctx* ctx, void* graph_exec, int times);
int bsz_graph_destroy(void* graph_exec);

/* ── schedule cost model: a small MLP (n_in -> h1 -> h2 -> 1, ReLU) trained
 *    and evaluated on the GPU with BF16 MFMA kernels (FP32 accumulation,
 *    FP32 master weights, Adam).  It predicts the log run time of a
 *    (program, schedule) pair from partition features, to rank schedules of
 *    programs too large to autotune; it never touches program outputs.  The
 *    reference has no schedule search (one kernel per op, kernels.cuh:7-36),
 *    so these have no counterpart there. ─────────────────────────────────── */

typedef struct bsz_mlp bsz_mlp;
int bsz_mlp_create(bsz_ctx* ctx, int n_in, int h1, int h2, uint64_t seed, bsz_mlp** out);
int bsz_mlp_free(bsz_mlp* m);
/* Parameters as one float vector: W1[n_in][h1], b1[h1], W2[h1][h2], b2[h2], w3[h2], b3. */
int64_t bsz_mlp_n_params(const bsz_mlp* m);
int bsz_mlp_get_params(bsz_mlp* m, float* out);
int bsz_mlp_set_params(bsz_mlp* m, const float* in);
/* x: n rows of n_in floats (row-major); synchronous. */
int bsz_mlp_predict(bsz_mlp* m, const float* x, int n, float* out);
/* Mean squared error on a batch and its gradient (parameter layout), no update. */
int bsz_mlp_gradients(bsz_mlp* m, const float* x, const float* y, int n, float* grad, float* loss);
int bsz_mlp_train_step(bsz_mlp* m, const float* x, const float* y, int n, float lr, float* loss);
/* `epochs` shuffled passes in minibatches of `batch` rows; *loss = last epoch's mean batch loss. */
int bsz_mlp_train(bsz_mlp* m, const float* x, const float* y, int n, int epochs, int batch, float lr, uint64_t seed,
                  float* loss);

#ifdef __cplusplus
}
#endif
#endif

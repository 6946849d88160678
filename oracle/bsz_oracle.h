/*
 * bsz_oracle.h — CPU oracle for the bitstream execution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links this; it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and only as the checker / the CPU baseline, never as the measured path.
 *
 * It is a from-scratch restatement of the reference's CPU path:
 *   - front-end: simulator/csim/src/lexer.cpp:62-142, parser.cpp:36-457
 *     (grammar of DSL.md:303-357, precedence ~ << >> & + ^ |);
 *   - control plane: simulator/csim/src/interpreter.inl:17-364
 *     (get_var / get_operand / exec_rhs / exec_stmt / run, MAX_WHILE_ITERS);
 *   - data plane: simulator/csim/src/cpu_backend.h:36-151 and
 *     bitstream_ops.cpp:29-66 (ripple add), 120-159 (shl/shr edge cases),
 *     403-487 (AVX-512 + OpenMP pointwise and popcount);
 *   - op counting: DSL.md:277-285, ast_liveness.cpp:629-727 (static ops).
 *
 * Two value semantics are offered:
 *   BSZO_SEM_CLEAN (default): DSL.md:273-275 bounded execution — inputs are
 *     truncated to bitlength, and ~, ONES, + and << never leave a bit at or
 *     above bitlength (pythonsim interpreter.py:176-210 masks the same ops).
 *   BSZO_SEM_BSIM: cpu_backend.h:90-96 — ONES/NOT/+/<< leave padding bits
 *     dirty and only the bounded reductions and the final output mask hide
 *     them.  The two differ only when `>>` pulls padding into [0, bitlength).
 *   BSZO_SEM_BSIM32: the same with bsim_cuda's --word-bits=32 layout, whose
 *     padding ends at a multiple of 256 bits instead of 512.
 *
 * Parity status of each function is listed in the header comment of
 * bsz_oracle.cpp.
 */
#ifndef BSZ_ORACLE_H
#define BSZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bszo_case bszo_case;

enum { BSZO_SEM_CLEAN = 0, BSZO_SEM_BSIM = 1, BSZO_SEM_BSIM32 = 2 };

/* Execution variants of the word kernels (all give identical bits).
 *   SCALAR      plain loops, ripple add — the reference model
 *   FAITHFUL    AVX-512 + OpenMP pointwise/popcount, serial backward shl and
 *               serial __int128 ripple add, as bsim --backend simd_omp --add ripple
 *   BEST        AVX-512 + OpenMP everywhere, chunked parallel carry resolution
 */
enum { BSZO_SCALAR = 0, BSZO_FAITHFUL = 1, BSZO_BEST = 2 };

/* Parse `source`; returns NULL and fills err on a parse / 3-address error. */
bszo_case* bszo_case_new(const char* source, char* err, size_t errlen);
void bszo_case_free(bszo_case* c);

/* Per-case inputs (mirrors the .bsdata case fields, bsdata_reader.h:31-40). */
int bszo_set_bitlength(bszo_case* c, int64_t bitlength);
int bszo_set_semantics(bszo_case* c, int sem);
int bszo_set_param(bszo_case* c, const char* name, int64_t value);
int bszo_set_input(bszo_case* c, const char* name, const uint64_t* words, uint64_t n);
int bszo_set_array_elem(bszo_case* c, const char* name, int64_t idx,
                        const uint64_t* words, uint64_t n);
int bszo_set_broadcast(bszo_case* c, const char* name, const uint8_t* bits, int64_t count);

/* Execute.  Returns 0, or -1 with a message (RuntimeError text) in err. */
int bszo_run(bszo_case* c, int variant, char* err, size_t errlen);

/* Results. */
uint64_t bszo_n_words(const bszo_case* c);          /* ceil(bitlength/64), >= 1 */
int64_t  bszo_op_count(const bszo_case* c);         /* dynamic ops */
int64_t  bszo_static_ops(const bszo_case* c);       /* count_ops_ast equivalent */
double   bszo_exec_ms(const bszo_case* c);
/* Newline-separated lists of declared names. */
const char* bszo_output_streams(const bszo_case* c);
const char* bszo_output_ints(const bszo_case* c);
/* 1 if the output exists after the run; words are masked to bitlength. */
int bszo_get_stream(const bszo_case* c, const char* name, uint64_t* out, uint64_t n);
/* Array outputs: number of elements and their indices. */
int64_t bszo_array_count(const bszo_case* c, const char* name);
int bszo_array_indices(const bszo_case* c, const char* name, int64_t* idx, int64_t cap);
int bszo_get_array_elem(const bszo_case* c, const char* name, int64_t idx,
                        uint64_t* out, uint64_t n);
int bszo_get_int(const bszo_case* c, const char* name, int64_t* out);

/* Word-level kernels, exported for the unit tests that mirror test_ops.cpp.
 * n is the word count; k the shift in bits.  dst may alias src. */
void bszo_shl(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n);
void bszo_shr(uint64_t* dst, const uint64_t* src, uint64_t k, uint64_t n);
void bszo_add(uint64_t* dst, const uint64_t* a, const uint64_t* b, uint64_t n, int variant);
int64_t bszo_popcount_bounded(const uint64_t* a, uint64_t n_words, int64_t bitlength);
int bszo_is_nonzero_bounded(const uint64_t* a, uint64_t n_words, int64_t bitlength);

/* Threads OpenMP will use for FAITHFUL / BEST (0 = runtime default). */
int bszo_set_threads(int n);
int bszo_get_threads(void);
int bszo_has_avx512(void);

#ifdef __cplusplus
}
#endif
#endif

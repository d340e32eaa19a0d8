/*
 * q1tsim MI355X statevector backend -- state-level C ABI.
 *
 * One opaque q1t_gvs is one GpuVectorState: the replacement for q1tsim's
 * VectorState behind the QuState trait (src/qustate.rs:5-89).  A Rust
 * `GpuVectorState` implements QuState by calling these functions (the
 * binding is shown in INTEGRATION.md); gates arrive as their FP64
 * Gate::matrix() plus qubit indices, exactly what the trait hands a backend.
 *
 * Conventions (identical to the reference):
 *  - qubit q is index bit n-1-q (qubit 0 is the most significant bit);
 *  - a k-qubit gate matrix is 2^k x 2^k, row-major, complex interleaved
 *    (re, im) doubles, the first listed qubit being the most significant bit
 *    of the local index (gates.rs:176-183);
 *  - runs of column j are the contiguous range sum(counts[:j]) .. of the
 *    result array (vectorstate.rs:283-297).
 *
 * All functions return 0 on success or a Q1T_ERR_* code; q1t_last_error()
 * then returns the message (the Display text of the matching src/error.rs
 * variant where one exists).  Gates are queued and applied lazily: any
 * measurement, read-out or q1t_gvs_flush applies the queue as planned sweeps.
 */
#ifndef Q1TSIM_GVS_H
#define Q1TSIM_GVS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    Q1T_OK = 0,
    Q1T_ERR_INVALID_NR_BITS = 1,              /* Error::InvalidNrBits */
    Q1T_ERR_INVALID_QBIT = 2,                 /* Error::InvalidQBit */
    Q1T_ERR_INVALID_CBIT = 3,                 /* Error::InvalidCBit */
    Q1T_ERR_INVALID_NR_CONTROL_BITS = 4,      /* Error::InvalidNrControlBits */
    Q1T_ERR_INVALID_NR_MEASUREMENT_BITS = 5,  /* Error::InvalidNrMeasurementBits */
    Q1T_ERR_NOT_ENOUGH_SPACE = 6,             /* Error::NotEnoughSpace */
    Q1T_ERR_NOT_EXECUTED = 7,                 /* Error::NotExecuted */
    Q1T_ERR_INVALID_WEIGHTS = 8,              /* WeightedIndex::new failure (vectorstate.rs:123-125) */
    Q1T_ERR_UNKNOWN_GATE = 9,                 /* ParseError::UnknownGate */
    Q1T_ERR_INVALID_NR_ARGUMENTS = 10,        /* ParseError::InvalidNrArguments */
    Q1T_ERR_HIP = 11,
    Q1T_ERR_OUT_OF_MEMORY = 12,
    Q1T_ERR_INVALID_ARGUMENT = 13,
    Q1T_ERR_NOT_SUPPORTED = 14
};

enum { Q1T_PREC_C128 = 0, Q1T_PREC_C64 = 1 };

typedef struct q1t_gvs q1t_gvs;

/* Random source: raw 64-bit draws (a Rust `R: Rng` behind a trampoline, or
 * q1t_xoshiro_next).  Uniforms are (x >> 12) * 2^-52, as rand 0.7 builds a
 * Uniform<f64>; a Binomial(count, p) draw is the number of runs with u < p. */
typedef struct {
    void* ctx;
    uint64_t (*next_u64)(void* ctx);
} q1t_rng;

typedef struct {
    double gate_seconds;   /* wall time of the last gate flush */
    double bytes;          /* algorithmic HBM bytes moved by it */
    uint64_t launches;     /* kernels launched */
    uint64_t sweeps;       /* HBM sweeps over the state */
    uint64_t gates;        /* gates applied */
    uint64_t blocks;       /* fused dense blocks applied (per column) */
    uint64_t phases;       /* in-tile phases (LDS exchanges) over all sweeps */
    double mfma_flops;     /* FP32 MFMA flops of the fused blocks: 2^(n+k+3) per k-qubit block and column */
} q1t_stats;

const char* q1t_last_error(void);

/* VectorState::new (vectorstate.rs:41-53): |0...0>, one column, nr_shots runs. */
int q1t_gvs_new(size_t nr_bits, size_t nr_shots, int precision, q1t_gvs** out);
void q1t_gvs_free(q1t_gvs* s);
/* VectorState::from_qubit_coefs (vectorstate.rs:62-83): 2*nr_bits complex coefficients. */
int q1t_gvs_from_qubit_coefs(q1t_gvs* s, const double* coefs, size_t ncoefs);
/* Build-added: upload a 2^n amplitude vector (no reference counterpart). */
int q1t_gvs_set_amplitudes(q1t_gvs* s, const double* amps, size_t len);
/* Build-added: re, im i.i.d. N(0,1) (SplitMix64 counter hash + Box-Muller), normalised. */
int q1t_gvs_set_random_normal(q1t_gvs* s, uint64_t seed);

size_t q1t_gvs_nr_bits(const q1t_gvs* s);
size_t q1t_gvs_nr_cols(const q1t_gvs* s);
int q1t_gvs_counts(const q1t_gvs* s, size_t* out, size_t len);

/* QuState::apply_gate (qustate.rs:8, vectorstate.rs:166-178).  Gates on up to
 * 12 qubits (wider ones return ERR_NOT_SUPPORTED). */
int q1t_gvs_apply_gate(q1t_gvs* s, const double* mat, size_t nr_affected, const size_t* bits, size_t nbits,
                       const char* desc);
/* QuState::apply_unary_gate_all (qustate.rs:13, vectorstate.rs:180-189). */
int q1t_gvs_apply_unary_gate_all(q1t_gvs* s, const double* mat2, const char* desc);
/* QuState::apply_conditional_gate (qustate.rs:17, vectorstate.rs:193-227). */
int q1t_gvs_apply_conditional_gate(q1t_gvs* s, const uint8_t* control, size_t ncontrol, const double* mat,
                                   size_t nr_affected, const size_t* bits, size_t nbits, const char* desc);

/* Two-step measure for hosts that draw their own binomials
 * (vectorstate.rs:249-275 then 277-327): w0 per column, then the collapse. */
int q1t_gvs_marginal_w0(q1t_gvs* s, size_t qbit, double* w0s, size_t len);
int q1t_gvs_measure_with_n0(q1t_gvs* s, size_t qbit, size_t cbit, const size_t* n0s, const double* w0s,
                            size_t ncols, uint64_t* res, size_t res_len, int collapse);
/* No counterpart in the reference.  q1t_gvs_marginal_w0 for callers that
 * cannot allocate device memory (code under graph capture, tight serving
 * loops): the call makes no hipMalloc/hipFree.  Its partial sums live in a
 * fixed 516 KiB device array per GPU, held until the process ends; at most
 * 256 branch columns (a larger state is rejected with
 * Q1T_ERR_INVALID_ARGUMENT), no pending gates (flush first), and the device
 * current at the call must be the state's.  Results are bit-identical to
 * q1t_gvs_marginal_w0.  Calls that overlap on one GPU share the array and
 * race: serialise them. */
int q1t_gvs_marginal_w0_noalloc(q1t_gvs* s, size_t qbit, double* w0s, size_t len);
/* QuState::measure_into (qustate.rs:35, vectorstate.rs:237-329). */
int q1t_gvs_measure_into(q1t_gvs* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len, const q1t_rng* rng);
/* QuState::peek_into (qustate.rs:62, vectorstate.rs:346-393). */
int q1t_gvs_peek_into(q1t_gvs* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len, const q1t_rng* rng);
/* Build-added: measure (collapse = 1) or peek one qubit in the 'Z', 'X' or
 * 'Y' basis, the op loop's basis change (circuit.rs:669-718: H / Sdg.H
 * before, H / H.S after) folded into the marginal and collapse kernels: one
 * read pass for the marginal, one pass for the collapse, no gate passes. */
int q1t_gvs_measure_basis_into(q1t_gvs* s, size_t qbit, size_t cbit, char basis, uint64_t* res, size_t res_len,
                               const q1t_rng* rng, int collapse);
/* QuState::measure_all_into (qustate.rs:50, vectorstate.rs:340-344). */
int q1t_gvs_measure_all_into(q1t_gvs* s, const size_t* cbits, size_t ncbits, uint64_t* res, size_t res_len,
                             const q1t_rng* rng);
/* QuState::peek_all_into (qustate.rs:74, vectorstate.rs:395-400). */
int q1t_gvs_peek_all_into(q1t_gvs* s, const size_t* cbits, size_t ncbits, uint64_t* res, size_t res_len,
                          const q1t_rng* rng);
/* measure_all_into_helper (vectorstate.rs:106-161) with one injected uniform
 * in [0,1) per run: the chosen weight is u * total, mapped to an index by
 * the canonical F=16 tree walk (FP64).  collapse = 0 is peek_all. */
int q1t_gvs_measure_all_uniforms(q1t_gvs* s, const size_t* cbits, size_t ncbits, const double* u01,
                                 uint64_t* res, size_t res_len, int collapse);
/* QuState::reset (qustate.rs:82, vectorstate.rs:402-408). */
int q1t_gvs_reset(q1t_gvs* s, size_t bit, const q1t_rng* rng);
/* QuState::reset_all (qustate.rs:88, vectorstate.rs:410-415). */
int q1t_gvs_reset_all(q1t_gvs* s);

/* Build-added read-out: column `col` as 2^n interleaved complex doubles. */
int q1t_gvs_read_amplitudes(q1t_gvs* s, size_t col, double* out, size_t len);
/* Build-added: apply queued gates now. */
int q1t_gvs_flush(q1t_gvs* s);
/* Build-added: fusion_k = 0 keeps per-gate FP64/FP32 arithmetic (plain);
 * 2..6 composes gate runs into dense k-qubit blocks; tiling = 1 applies
 * runs that fit an LDS tile in one HBM sweep. */
int q1t_gvs_set_plan(q1t_gvs* s, int fusion_k, int tiling);
/* Build-added: local search of the sweep tiles and register phases;
 * mode -1 (default) for states of 512 MiB and more, 0 never, 1 always.
 * Changes only the schedule, not the gates applied. */
int q1t_gvs_set_plan_search(q1t_gvs* s, int mode);
/* Build-added, no counterpart in the reference: flush graphs (on = 1).  The
 * kernel launches of a planned flush are captured into one hipGraph; a later
 * flush whose launch signature (tiles, program sizes, single-gate kernels
 * with their matrices, columns, buffer addresses) is the same replays it
 * after uploading its own sweep programs, so new gate parameters reach the
 * replay.  Off while profiling or recording; a flush with a 3+-target
 * single-gate kernel runs eagerly.  Results equal the eager path's. */
int q1t_gvs_set_graph(q1t_gvs* s, int on);
/* Build-added, no counterpart in the reference: sweep kernels specialised at
 * run time (on = 1).  Each sweep program's structure (op types, register
 * slots, controls, phases) becomes straight-line code compiled with hipRTC
 * on the first flush that needs it and cached per process; the gate
 * matrices are still read from the uploaded program, so one kernel serves
 * every program of the same structure.  Same results as the interpreter
 * kernels (the same op bodies).  Diagonal ops on bits outside a sweep's tile
 * are folded into one factor per tile, applied as the tile is stored.
 * on = 2: the same kernels with each workgroup's next tile loaded into
 * registers while the current one runs.  on = 3: specialised kernels
 * without tile factors -- every op where the interpreter applies it, so the
 * amplitudes equal the interpreter kernels' bit for bit. */
int q1t_gvs_set_jit(q1t_gvs* s, int on);
/* Build-added: graphs captured and replays launched over the state's life. */
int q1t_gvs_graph_stats(q1t_gvs* s, uint64_t* captures, uint64_t* replays);
/* Build-added: tile geometry of the sweeps (M tile bits, R register bits,
 * L low bits per tile); only the compiled shapes are accepted
 * (complex128: M=12 with R=3 or 4; complex64: M=12 with R=3 or 4, M=13 with R=4). */
int q1t_gvs_set_geometry(q1t_gvs* s, int M, int R, int L);
int q1t_gvs_last_stats(const q1t_gvs* s, q1t_stats* out);

/* Build-added measurement: with profiling on, every kernel launch of a flush
 * is bracketed by HIP events on the state's stream; per-kernel totals. */
typedef struct {
    char name[48];
    uint64_t launches;
    double total_ms;
    double bytes;      /* algorithmic HBM bytes of those launches */
} q1t_kernel_profile;
int q1t_gvs_set_profiling(q1t_gvs* s, int on);
int q1t_gvs_kernel_profile(const q1t_gvs* s, q1t_kernel_profile* out, size_t cap, size_t* count);
/* Build-added measurement: best-of-`reps` time of a device-to-device copy of
 * `bytes` (HIP events), i.e. the read+write streaming rate of the part. */
int q1t_copy_bandwidth(size_t bytes, int reps, double* best_ms);
/* Build-added, no counterpart in the reference: the process-wide cache of
 * released device blocks (state buffers are reused by size class).  Limit
 * in bytes (0: 7/8 of device memory, the default); lowering it frees the
 * oldest cached blocks at once.  A host that shares the GPU with other
 * allocators lowers it. */
int q1t_device_cache_limit(size_t bytes);
int q1t_device_cache_bytes(size_t* out);

/* Build-added accuracy report: <a|b>, max_i |a_i - b_i| and both squared
 * norms over one column of each state, FP64 on the device (either state may
 * be complex64).  Infidelity = 1 - |<a|b>|^2 / (|a|^2 |b|^2). */
int q1t_gvs_compare(q1t_gvs* a, size_t col_a, q1t_gvs* b, size_t col_b, double* out5);

/* Gate::matrix() of a named gate (names of ffi.rs:335-366 and
 * composite.rs:289-445, case-insensitive): writes 2^k x 2^k interleaved
 * complex doubles (out must hold 2 * 4^k doubles, cap counts doubles) and k. */
int q1t_gate_matrix(const char* name, const double* params, size_t nparams, double* out, size_t cap,
                    size_t* nbits);

/* Built-in xoshiro256** stream, seeded through SplitMix64. */
typedef struct q1t_xoshiro q1t_xoshiro;
q1t_xoshiro* q1t_xoshiro_new(uint64_t seed);
void q1t_xoshiro_free(q1t_xoshiro* x);
uint64_t q1t_xoshiro_next(void* ctx);

/* ---- learned sweep cost model (no reference counterpart: q1tsim applies
 * gates one by one, src/circuit.rs:643-762, and has no planner) ----
 * A 16 -> 64 -> 64 -> 1 tanh MLP in FP32 on standardised features that
 * predicts log(microseconds per 1024 tiles) of a sweep program.  Training
 * and inference run in HIP kernels; the model only ranks equivalent
 * programs, it never changes amplitudes. */
#define Q1T_CM_FEATURES 20
#define Q1T_CM_PARAMS 5569
typedef struct q1t_cost_model q1t_cost_model;
int q1t_cm_new(uint64_t seed, q1t_cost_model** out);
void q1t_cm_free(q1t_cost_model* cm);
int q1t_cm_get_params(const q1t_cost_model* cm, float* out, size_t len);
int q1t_cm_set_params(q1t_cost_model* cm, const float* p, size_t len);
/* x' = (x - mu) / sigma per feature, sigma > 0 */
int q1t_cm_set_norm(q1t_cost_model* cm, const float* mu, const float* sigma, size_t nfeat);
int q1t_cm_get_norm(const q1t_cost_model* cm, float* mu, float* sigma, size_t nfeat);
int q1t_cm_predict(q1t_cost_model* cm, const float* x, size_t nrows, float* y);
/* mean squared error and its gradient (layout of the parameters) */
int q1t_cm_loss_grad(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, float* grad, size_t len,
                     double* loss);
int q1t_cm_adam_step(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, float lr, double* loss);
/* minibatch Adam over the set; *loss = final full-set loss; marks the model trained */
int q1t_cm_train(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, int epochs, size_t batch,
                 float lr, uint64_t seed, double* loss);
int q1t_cm_set_trained(q1t_cost_model* cm, int trained);

/* planner mode of a state's fused complex64 flushes: 0 greedy fixed-width
 * fusion, 1 learned (cm ranks the candidate programs; greedy while cm is
 * untrained or NULL), 2 seeded random choice among the candidates */
int q1t_gvs_set_cost_model(q1t_gvs* s, q1t_cost_model* cm, int mode, uint64_t seed);
/* record every sweep's features and measured time (turns profiling on) */
int q1t_gvs_set_recording(q1t_gvs* s, int on);
/* move up to cap recorded samples out, oldest first (the rest stay):
 * feat[cap][16], tiles[cap] (tiles over all columns), ms[cap] */
int q1t_gvs_take_samples(q1t_gvs* s, float* feat, double* tiles, double* ms, size_t cap, size_t* count);

#ifdef __cplusplus
}
#endif
#endif

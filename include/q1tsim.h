/*
 * q1tsim MI355X statevector backend -- circuit-level C ABI.
 *
 * The symbols, struct layouts and result conventions of q1tsim's src/ffi.rs,
 * so python/q1tsimffi.py's cdef (q1tsimffi.py:35-82) binds this library
 * (libq1tsim.so) unchanged.  Every circuit executes on the GPU statevector
 * (q1tsim_gvs.h); the q1t_circuit_* functions are build-added controls.
 */
#ifndef Q1TSIM_H
#define Q1TSIM_H

#include <stddef.h>
#include <stdint.h>

#include "q1tsim_gvs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* result types (ffi.rs:5-9) */
#define RESULT_ERROR 0
#define RESULT_EMPTY 1
#define RESULT_STRING 2
#define RESULT_HISTOGRAM 3
#define RESULT_CSTATE 5

typedef struct circuit circuit_t;

/* CHistElem (ffi.rs:25-30) */
typedef struct {
    const char* key;
    size_t count;
} histelem_t;

/* CParameter (ffi.rs:40-46): a non-null value_ptr is read at execute time */
typedef struct {
    double value;
    const double* value_ptr;
} parameter_t;

/* CResult (ffi.rs:63-70) */
typedef struct {
    void* data;
    size_t length;
    size_t size;
    uint32_t restype;
} result_t;

void result_free(result_t res);                                       /* ffi.rs:165-169 */
circuit_t* circuit_new(size_t nr_qbits, size_t nr_cbits);             /* ffi.rs:171-176 */
void circuit_free(circuit_t* ptr);                                    /* ffi.rs:178-186 */
size_t circuit_nr_qbits(const circuit_t* ptr);                        /* ffi.rs:188-194 */
size_t circuit_nr_cbits(const circuit_t* ptr);                        /* ffi.rs:196-202 */
result_t circuit_cstate(const circuit_t* ptr);                        /* ffi.rs:204-214 */
result_t circuit_add_gate(circuit_t* ptr, const char* gate, const size_t* qbits, size_t nr_qbits,
                          const parameter_t* params, size_t nr_params);  /* ffi.rs:307-380 */
result_t circuit_add_conditional_gate(circuit_t* ptr, const size_t* control, size_t nr_control, uint64_t target,
                                      const char* gate, const size_t* qbits, size_t nr_qbits,
                                      const parameter_t* params, size_t nr_params);  /* ffi.rs:382-459 */
result_t circuit_reset(circuit_t* ptr, size_t qbit);                  /* ffi.rs:461-477 */
result_t circuit_reset_all(circuit_t* ptr);                           /* ffi.rs:479-492 */
result_t circuit_measure(circuit_t* ptr, size_t qbit, size_t cbit, char dir, uint8_t collapse);  /* ffi.rs:494-522 */
result_t circuit_measure_all(circuit_t* ptr, const size_t* cbits, size_t nr_cbits, char dir,
                             uint8_t collapse);                       /* ffi.rs:524-557 */
result_t circuit_execute(circuit_t* ptr, size_t nr_shots);            /* ffi.rs:559-575 */
result_t circuit_reexecute(circuit_t* ptr);                           /* ffi.rs:577-593 */
result_t circuit_histogram(const circuit_t* ptr);                     /* ffi.rs:595-611 */
result_t circuit_latex(const circuit_t* ptr);                         /* ffi.rs:613-629 (export: not provided) */
result_t circuit_open_qasm(const circuit_t* ptr);                     /* ffi.rs:632-648 (export: not provided) */
result_t circuit_c_qasm(const circuit_t* ptr);                        /* ffi.rs:650-666 (export: not provided) */

/* ---- build-added controls (no ffi.rs counterpart) ---- */
/* precision Q1T_PREC_C128 (plain, FP64) or Q1T_PREC_C64; see q1t_gvs_set_plan */
int q1t_circuit_set_backend(circuit_t* ptr, int precision, int fusion_k, int tiling);
/* Build-added, no counterpart in the reference: flush graphs of the
 * circuit's state (q1t_gvs_set_graph).  `reexecute` keeps the state, so its
 * gate flushes replay the previous execute's graphs, with reference
 * parameters re-read; graph_stats counts over the current state. */
int q1t_circuit_set_graph(circuit_t* ptr, int on);
int q1t_circuit_graph_stats(circuit_t* ptr, uint64_t* captures, uint64_t* replays);
/* Build-added, no counterpart in the reference: sweep kernels specialised
 * to each sweep program of the circuit's state (q1t_gvs_set_jit). */
int q1t_circuit_set_jit(circuit_t* ptr, int on);
/* Build-added: plan a gate-only circuit on the host and compile the kernels
 * specialised to its sweeps with hipRTC (no GPU needed); precision as
 * q1t_circuit_set_backend, fusion_k > 0 greedy fusion, mode as
 * q1t_gvs_set_jit (1, 2 or 3), geometry that of the specialised kernels.  With out_dir, each
 * code object and source is written there as sweep_<i>.co / .hip. */
int q1t_circuit_jit_compile(const circuit_t* ptr, int precision, int fusion_k, int mode, const char* out_dir,
                            size_t* nkernels, double* seconds);
/* Build-added: specialised kernels compiled in this process, launches of
 * already loaded ones, and host seconds spent compiling. */
int q1t_jit_stats(uint64_t* compiled, uint64_t* hits, double* compile_s);
/* seed of the built-in xoshiro256** stream used by execute/reexecute */
int q1t_circuit_seed(circuit_t* ptr, uint64_t seed);
/* initial state for execute: 2^n interleaved amplitudes, or a seeded N(0,1) state */
int q1t_circuit_set_initial_amplitudes(circuit_t* ptr, const double* amps, size_t len);
int q1t_circuit_set_initial_random(circuit_t* ptr, uint64_t seed);
int q1t_circuit_barrier(circuit_t* ptr, const size_t* qbits, size_t n);     /* Circuit::barrier (circuit.rs:543-554) */
int q1t_circuit_nr_columns(circuit_t* ptr, size_t* out);
int q1t_circuit_counts(circuit_t* ptr, size_t* out, size_t len);
int q1t_circuit_read_amplitudes(circuit_t* ptr, size_t col, double* out, size_t len);
int q1t_circuit_last_stats(const circuit_t* ptr, q1t_stats* out);
/* histogram_vec (circuit.rs:800-815): counts of length 2^nr_cbits */
int q1t_circuit_histogram_vec(const circuit_t* ptr, size_t* out, size_t len);
/* The order in which the plain (per-gate) GPU plan of this gate-only circuit
 * applies its gates (indices in add order; identity gates never run and are
 * left out), for the default tile shape of `precision`; search != 0 refines
 * tiles and register phases by local search.  A planning report, computed on
 * the host: it needs no GPU. */
int q1t_circuit_plan_order(const circuit_t* ptr, int precision, int search, size_t* out, size_t cap,
                           size_t* count);

/* ---- batches of circuits (no ffi.rs counterpart: the reference runs one
 * circuit per Circuit::execute, circuit.rs:562-600) ----
 * Every added circuit is executed as circuit_execute(ptr, nr_shots) would
 * execute it (its own seed, the same classical state), in lockstep with the
 * others: the gate runs of all circuits between their measurements are
 * flushed together, one launch per sweep round.  Circuits must outlive the
 * batch's execute call; their backend is set to the batch's. */
typedef struct q1t_batch q1t_batch;
typedef struct q1t_batch_stats {
    q1t_stats gates;        /* summed over all gate flushes */
    uint64_t waves;         /* groups of circuits resident at once */
    uint64_t rounds;        /* lockstep rounds */
    uint64_t sync_ops;      /* measurements, conditional gates, resets */
    double seconds;         /* wall time of the execute call */
    double start_s, queue_s, flush_s, measure_s, other_s;   /* its parts: state set-up, gate queueing,
                                                               batched flushes, batched measurements,
                                                               other ops (conditional gates, resets) */
    uint64_t column_launches;   /* kernels and copies of the batched column work (measurements,
                                   collapses, conditional gates, basis columns) */
    uint64_t solo_ops;          /* ops run circuit by circuit (outside the batched kinds) */
    double plan_s;              /* host planning inside the batched flushes */
    double pack_s;              /* host table building and upload inside the batched flushes */
} q1t_batch_stats;
/* mem_budget: device bytes one wave of states may take (0: 3/4 of free memory) */
int q1t_batch_new(int precision, int fusion_k, size_t mem_budget, q1t_batch** out);
void q1t_batch_free(q1t_batch* b);
int q1t_batch_add(q1t_batch* b, circuit_t* c);
/* keep_states != 0: leave each end state on the device (circuit reads work) */
int q1t_batch_execute(q1t_batch* b, size_t nr_shots, int keep_states);
int q1t_batch_stats_get(const q1t_batch* b, q1t_batch_stats* out);
/* Autotuner: executes the first `sample` circuits under every compiled tile
 * shape (and, for complex64, fusion widths 0 and 4), keeps the fastest, and
 * restores each sampled circuit's random stream.  Writes up to cap timed
 * configurations; *count = configurations timed. */
typedef struct q1t_batch_config {
    int M, R, L;            /* tile bits, register bits, low bits per tile */
    int fusion_k;           /* 0: per-gate arithmetic; 3..6: dense blocks */
    double seconds;         /* wall time of the sample (0 for get_config) */
} q1t_batch_config;
int q1t_batch_autotune(q1t_batch* b, size_t nr_shots, size_t sample, q1t_batch_config* out, size_t cap,
                       size_t* count);
int q1t_batch_set_config(q1t_batch* b, const q1t_batch_config* cfg);
int q1t_batch_get_config(const q1t_batch* b, q1t_batch_config* out);
/* Build-added, no counterpart in the reference: consecutive tiles (1..1024)
 * a workgroup of a batched sweep takes at a time.  Schedule only; results
 * are unchanged. */
int q1t_batch_set_tile_chunk(q1t_batch* b, unsigned chunk);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU FP64 oracle for the q1tsim statevector path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of q1tsim 0.5.0's complex128 VectorState
 * (src/vectorstate.rs), its gate application (src/gates.rs, the per-gate files under src/gates/,
 * src/permutation.rs), the conditional run ranges (src/qustate.rs:100-127)
 * and the measurement bit mapping (src/support.rs:50-75).
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load this library, and only as the checker (or as the timed
 * CPU baseline); the product (q1tsim_amd/libq1tsim.so) never links it.
 *
 * Layout follows the reference exactly: the state is a row-major
 * (2^n x ncols) complex matrix, so amplitude i of every branch column is
 * contiguous (cmatrix.rs:20-24, vectorstate.rs:43).
 */
#ifndef Q1T_ORACLE_VECTORSTATE_REF_H
#define Q1T_ORACLE_VECTORSTATE_REF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_cplx;

/* Gate kinds: one per reference gate body (src/gates/<gate>.rs). */
enum orc_kind {
    ORC_I = 0, ORC_H, ORC_X, ORC_Y, ORC_Z, ORC_S, ORC_SDG, ORC_T, ORC_TDG,
    ORC_RX, ORC_RY, ORC_RZ, ORC_U1, ORC_U2, ORC_U3, ORC_V, ORC_VDG,
    ORC_SWAP,
    ORC_MATRIX   /* a gate known only through matrix(): default bodies */
};

/* C<...<G>> with nctrl control layers (controlled.rs:23-82). */
typedef struct {
    int kind;
    int nctrl;
    double p[3];            /* parameters in the reference's constructor order */
    size_t mat_bits;        /* ORC_MATRIX only: nr of affected bits */
    const orc_cplx* mat;    /* ORC_MATRIX only: 2^k x 2^k row-major */
} orc_gate;

typedef struct orc_state orc_state;

/* xoshiro256** (Blackman & Vigna) seeded through SplitMix64. */
typedef struct { uint64_t s[4]; } orc_rng;

void     orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);
double   orc_rng_uniform(orc_rng* r);   /* (x>>12) * 2^-52, in [0,1) */

orc_state* orc_new(size_t nr_bits, size_t nr_shots);
orc_state* orc_from_qubit_coefs(const orc_cplx* coefs, size_t nr_bits, size_t nr_shots);
orc_state* orc_from_amplitudes(const orc_cplx* amps, size_t nr_bits, size_t nr_shots);
void       orc_free(orc_state* s);

size_t orc_nr_bits(const orc_state* s);
size_t orc_nr_cols(const orc_state* s);
void   orc_counts(const orc_state* s, size_t* out);
void   orc_column(const orc_state* s, size_t col, orc_cplx* out);

/* Full 2^k x 2^k matrix of a gate, as Gate::matrix() builds it. */
size_t orc_gate_nbits(const orc_gate* g);
void   orc_gate_matrix(const orc_gate* g, orc_cplx* out);

/* Error codes mirror src/error.rs variants. 0 = Ok. */
enum { ORC_OK = 0, ORC_ERR_NR_BITS = 1, ORC_ERR_QBIT = 2, ORC_ERR_CBIT = 3,
       ORC_ERR_NR_CONTROL = 4, ORC_ERR_NR_MEAS = 5, ORC_ERR_SPACE = 6,
       ORC_ERR_WEIGHTS = 7 };

int orc_apply_gate(orc_state* s, const orc_gate* g, const size_t* bits, size_t nbits);
int orc_apply_unary_gate_all(orc_state* s, const orc_gate* g);
int orc_apply_conditional_gate(orc_state* s, const uint8_t* control, size_t ncontrol,
                               const orc_gate* g, const size_t* bits, size_t nbits);

/* Single-qubit measurement. n0 for each column is drawn as the number of
 * successes among `count` Bernoulli(min(w0,1)) trials, u < p, with u from
 * the rng (one draw per run, columns in order). */
int orc_marginal_w0(const orc_state* s, size_t qbit, double* w0s);
int orc_measure_into(orc_state* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len,
                     orc_rng* rng, int collapse);

/* measure_all / peek_all with one uniform u in [0,1) per run (runs in
 * column order); chosen weight t = u * T.  mode 0 = canonical F-ary tree
 * walk (the order the GPU uses), mode 1 = rand 0.7 WeightedIndex
 * (sequential cumulative sums, #cw <= t).  Results are grouped per column
 * by ascending outcome index. */
int orc_measure_all_uniforms(orc_state* s, const size_t* cbits, size_t ncbits,
                             const double* u01, uint64_t* res, size_t res_len,
                             int collapse, int mode);
int orc_measure_all_into(orc_state* s, const size_t* cbits, size_t ncbits,
                         uint64_t* res, size_t res_len, orc_rng* rng, int collapse);

int  orc_reset(orc_state* s, size_t bit, orc_rng* rng);
void orc_reset_all(orc_state* s);

/* Sampling primitives, exported for unit tests. */
#define ORC_TREE_FANOUT 16
size_t orc_tree_sample(const double* w, size_t len, double u, double* total_out);
void   orc_tree_sample_many(const double* w, size_t len, const double* us, size_t nu, size_t* out,
                            double* total_out);
size_t orc_weighted_index_seq(const double* w, size_t len, double u);
double orc_norm_sqr(orc_cplx a);
void   orc_canonical_weights(const orc_cplx* a, size_t n, double* out);

/* Support bit helpers (support.rs:50-75). */
uint64_t orc_reverse_bits(uint64_t idx, size_t nr_bits);
uint64_t orc_shuffle_bits(uint64_t idx, const size_t* bits, size_t nbits);

/* Conditional ranges (qustate.rs:100-127); returns the number of triplets. */
size_t orc_collect_conditional_ranges(const size_t* counts, size_t ncols, const uint8_t* control,
                                      size_t* out_col, size_t* out_len, uint8_t* out_apply);

/* Seeded normal random state: re, im i.i.d. N(0,1) from a SplitMix64
 * counter hash and Box-Muller, then FP64 normalised. */
void orc_random_normal_state(orc_cplx* out, size_t nr_bits, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU FP64 oracle for the q1tsim statevector path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates q1tsim 0.5.0's VectorState and gate application in plain C so that
 * the HIP backend can be checked against it.  Every function cites the
 * reference file:line it follows.  Compiled with -ffp-contract=off: Rust does
 * not contract a*b+c into fma, so neither does this restatement, except where
 * the canonical sampling order explicitly calls fma() (orc_norm_sqr).
 *
 * Parity notes (what pins each part is listed in the test files):
 *  - gate bodies, permutation path, conditional ranges, collapse: restated
 *    one-to-one; pinned by the reference's known-answer tests and by a dense
 *    numpy operator built from Kronecker products.
 *  - Binomial draws (rand_distr 0.2) and the RNG stream (rand 0.7
 *    thread_rng): parity unpinned -- replaced by a Bernoulli count over the
 *    xoshiro256** stream, identical in oracle and product.
 *  - WeightedIndex (rand 0.7): restated in orc_weighted_index_seq; the GPU
 *    uses the canonical F-ary tree walk (orc_tree_sample), which is also
 *    restated here so the two can be compared on the same uniforms.
 */
#include "vectorstate_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_state {
    size_t nr_bits;
    size_t nr_shots;
    size_t ncols;
    size_t* counts;
    orc_cplx* states;   /* row-major: states[i * ncols + c] */
};

/* ------------------------------------------------------------------ RNG -- */

static uint64_t splitmix64_next(uint64_t* x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void orc_rng_seed(orc_rng* r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64_next(&x);
}

uint64_t orc_rng_next(orc_rng* r)
{
    uint64_t* s = r->s;
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* rand 0.7 Uniform<f64>: 52 random mantissa bits -> [1,2) -> minus 1. */
double orc_rng_uniform(orc_rng* r)
{
    return (double)(orc_rng_next(r) >> 12) * (1.0 / 4503599627370496.0);
}

/* -------------------------------------------------------------- complex -- */

static inline orc_cplx cmk(double re, double im) { orc_cplx c = {re, im}; return c; }
/* num_complex 0.2 Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline orc_cplx cmul(orc_cplx a, orc_cplx b)
{
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline orc_cplx cadd(orc_cplx a, orc_cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline orc_cplx csub(orc_cplx a, orc_cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline orc_cplx cneg(orc_cplx a) { return cmk(-a.re, -a.im); }
/* num_complex 0.2 from_polar: (r*cos(theta), r*sin(theta)) */
static inline orc_cplx cpolar(double r, double th) { return cmk(r * cos(th), r * sin(th)); }

#define HSQRT2 0.70710678118654752440

/* Canonical weight |a|^2 used by the sampling order shared with the GPU. */
double orc_norm_sqr(orc_cplx a) { return fma(a.re, a.re, a.im * a.im); }

/* ---------------------------------------------------------- gate matrix -- */

size_t orc_gate_nbits(const orc_gate* g)
{
    size_t base;
    if (g->kind == ORC_MATRIX) base = g->mat_bits;
    else if (g->kind == ORC_SWAP) base = 2;
    else base = 1;
    return base + (size_t)g->nctrl;
}

/* Gate::matrix() of the uncontrolled gate (src/gates/<gate>.rs matrix()). */
static void base_matrix(const orc_gate* g, orc_cplx* m)
{
    const orc_cplx z = cmk(0, 0), o = cmk(1, 0), i = cmk(0, 1);
    const orc_cplx x = cmk(HSQRT2, 0);
    switch (g->kind) {
    case ORC_I: m[0] = o; m[1] = z; m[2] = z; m[3] = o; break;             /* identity.rs */
    case ORC_H: m[0] = x; m[1] = x; m[2] = x; m[3] = cneg(x); break;       /* hadamard.rs:89-93 */
    case ORC_X: m[0] = z; m[1] = o; m[2] = o; m[3] = z; break;             /* x.rs:79-84 */
    case ORC_Y: m[0] = z; m[1] = cneg(i); m[2] = i; m[3] = z; break;       /* y.rs:53-58 */
    case ORC_Z: m[0] = o; m[1] = z; m[2] = z; m[3] = cneg(o); break;       /* z.rs:53-58 */
    case ORC_S: m[0] = o; m[1] = z; m[2] = z; m[3] = i; break;             /* s.rs:60-66 */
    case ORC_SDG: m[0] = o; m[1] = z; m[2] = z; m[3] = cneg(i); break;     /* s.rs:186-192 */
    case ORC_T: m[0] = o; m[1] = z; m[2] = z; m[3] = cadd(x, cmul(x, i)); break; /* t.rs:52-59 */
    case ORC_TDG: m[0] = o; m[1] = z; m[2] = z; m[3] = csub(x, cmul(x, i)); break;
    case ORC_RX: {                                                        /* rx.rs:64-70 */
        double h = 0.5 * g->p[0];
        orc_cplx c = cmk(cos(h), 0), si = cmk(0, sin(h));
        m[0] = c; m[1] = cneg(si); m[2] = cneg(si); m[3] = c; break;
    }
    case ORC_RY: {                                                        /* ry.rs:64-70 */
        double h = 0.5 * g->p[0];
        orc_cplx c = cmk(cos(h), 0), s = cmk(sin(h), 0);
        m[0] = c; m[1] = cneg(s); m[2] = s; m[3] = c; break;
    }
    case ORC_RZ: {                                                        /* rz.rs:65-70 */
        orc_cplx p = cpolar(1.0, 0.5 * g->p[0]);
        m[0] = cmk(p.re, -p.im); m[1] = z; m[2] = z; m[3] = p; break;
    }
    case ORC_U1: m[0] = o; m[1] = z; m[2] = z; m[3] = cpolar(1.0, g->p[0]); break; /* u1.rs:69-75 */
    case ORC_U2: {                                                        /* u2.rs:70-79 */
        double phi = g->p[0], lambda = g->p[1];
        m[0] = cmk(HSQRT2, 0); m[1] = cneg(cpolar(HSQRT2, lambda));
        m[2] = cpolar(HSQRT2, phi); m[3] = cpolar(HSQRT2, phi + lambda); break;
    }
    case ORC_U3: {                                                        /* u3.rs:74-84 */
        double h = 0.5 * g->p[0], phi = g->p[1], lambda = g->p[2];
        double c = cos(h), s = sin(h);
        m[0] = cmk(c, 0); m[1] = cneg(cpolar(s, lambda));
        m[2] = cpolar(s, phi); m[3] = cpolar(c, phi + lambda); break;
    }
    case ORC_V: {                                                         /* v.rs:58-63 */
        orc_cplx a = cmk(0.5, 0.5), b = cmk(0.5, -0.5);
        m[0] = a; m[1] = b; m[2] = b; m[3] = a; break;
    }
    case ORC_VDG: {                                                       /* v.rs:157-162 */
        orc_cplx a = cmk(0.5, 0.5), b = cmk(0.5, -0.5);
        m[0] = b; m[1] = a; m[2] = a; m[3] = b; break;
    }
    case ORC_SWAP:                                                        /* swap.rs:78-88 */
        for (int k = 0; k < 16; ++k) m[k] = z;
        m[0] = o; m[1 * 4 + 2] = o; m[2 * 4 + 1] = o; m[15] = o; break;
    case ORC_MATRIX: {
        size_t d = (size_t)1 << g->mat_bits;
        memcpy(m, g->mat, d * d * sizeof(orc_cplx));
        break;
    }
    default: break;
    }
}

/* C<G>::matrix(): eye(2g) with G in the lower-right block (controlled.rs:60-69). */
void orc_gate_matrix(const orc_gate* g, orc_cplx* out)
{
    orc_gate inner = *g;
    inner.nctrl = 0;
    size_t d = (size_t)1 << orc_gate_nbits(&inner);
    orc_cplx* cur = (orc_cplx*)malloc(d * d * sizeof(orc_cplx));
    base_matrix(&inner, cur);
    for (int c = 0; c < g->nctrl; ++c) {
        size_t nd = 2 * d;
        orc_cplx* res = (orc_cplx*)calloc(nd * nd, sizeof(orc_cplx));
        for (size_t k = 0; k < d; ++k) res[k * nd + k] = cmk(1, 0);
        for (size_t r = 0; r < d; ++r)
            for (size_t q = 0; q < d; ++q) res[(d + r) * nd + d + q] = cur[r * d + q];
        free(cur);
        cur = res;
        d = nd;
    }
    memcpy(out, cur, d * d * sizeof(orc_cplx));
    free(cur);
}

/* ----------------------------------------------------------- gate bodies -- */
/*
 * Every body works on a flat, contiguous segment of `len` complex values:
 * a row range of the row-major state is contiguous over all columns, so the
 * reference's "rows [a, b), all columns" slices are flat sub-ranges.
 */

static void scale_range(orc_cplx* s, size_t len, orc_cplx f)
{
    for (size_t k = 0; k < len; ++k) s[k] = cmul(s[k], f);
}

/* Default dense apply (gates.rs:210-263 apply_slice / 273-326 apply_mat_slice). */
static void body_default(const orc_cplx* mat, size_t nbits, orc_cplx* s, size_t len)
{
    size_t d = (size_t)1 << nbits;
    size_t n = len >> nbits;
    orc_cplx* copy = (orc_cplx*)malloc(len * sizeof(orc_cplx));
    memcpy(copy, s, len * sizeof(orc_cplx));
    if (nbits == 1 || nbits == 2) {
        /* explicit unrolls: left-to-right sum of s_j * m[i][j] */
        for (size_t i = 0; i < d; ++i)
            for (size_t k = 0; k < n; ++k) {
                orc_cplx acc = cmul(copy[k], mat[i * d + 0]);
                for (size_t j = 1; j < d; ++j) acc = cadd(acc, cmul(copy[j * n + k], mat[i * d + j]));
                s[i * n + k] = acc;
            }
    } else {
        /* res = 0; res_i += s_j * m[i][j] for j in order */
        for (size_t i = 0; i < d; ++i)
            for (size_t k = 0; k < n; ++k) {
                orc_cplx acc = cmk(0, 0);
                for (size_t j = 0; j < d; ++j) acc = cadd(acc, cmul(copy[j * n + k], mat[i * d + j]));
                s[i * n + k] = acc;
            }
    }
    free(copy);
}

/* One (possibly controlled) gate body on a flat slice; vec selects the
 * apply_slice variant, otherwise apply_mat_slice. */
static void body(const orc_gate* g, int ctrl_left, int vec, orc_cplx* s, size_t len)
{
    if (ctrl_left > 0) {
        /* C<G>: apply G to the lower half (controlled.rs:71-81) */
        body(g, ctrl_left - 1, vec, s + len / 2, len / 2);
        return;
    }
    size_t n = len / 2;
    switch (g->kind) {
    case ORC_H: {   /* hadamard.rs:36-64 */
        orc_cplx* s1c = (orc_cplx*)malloc(n * sizeof(orc_cplx));
        memcpy(s1c, s + n, n * sizeof(orc_cplx));
        orc_cplx mx = cneg(cmk(HSQRT2, 0)), x = cmk(HSQRT2, 0);
        for (size_t k = 0; k < n; ++k) s[n + k] = cmul(csub(s[n + k], s[k]), mx);
        for (size_t k = 0; k < n; ++k) s[k] = cmul(cadd(s[k], s1c[k]), x);
        free(s1c);
        return;
    }
    case ORC_X:     /* x.rs:35-59 */
        for (size_t k = 0; k < n; ++k) { orc_cplx t = s[k]; s[k] = s[n + k]; s[n + k] = t; }
        return;
    case ORC_Y:     /* y.rs:60-86 */
        for (size_t k = 0; k < n; ++k) { orc_cplx t = s[k]; s[k] = s[n + k]; s[n + k] = t; }
        scale_range(s, n, cneg(cmk(0, 1)));
        scale_range(s + n, n, cmk(0, 1));
        return;
    case ORC_Z:     /* z.rs:60-75 */
        for (size_t k = 0; k < n; ++k) s[n + k] = cneg(s[n + k]);
        return;
    case ORC_S: scale_range(s + n, n, cmk(0, 1)); return;            /* s.rs:68-84 */
    case ORC_SDG: scale_range(s + n, n, cneg(cmk(0, 1))); return;    /* s.rs:194-210 */
    case ORC_T: scale_range(s + n, n, cpolar(1.0, M_PI / 4)); return;   /* t.rs:61-77 */
    case ORC_TDG: scale_range(s + n, n, cpolar(1.0, -M_PI / 4)); return;
    case ORC_RX:    /* rx.rs:72-112 */
    case ORC_RY: {  /* ry.rs:72-112 */
        double h = 0.5 * g->p[0];
        orc_cplx c = cmk(cos(h), 0);
        orc_cplx f = g->kind == ORC_RX ? cmk(0, sin(h)) : cmk(sin(h), 0);
        orc_cplx* t = (orc_cplx*)malloc(len * sizeof(orc_cplx));
        for (size_t k = 0; k < len; ++k) t[k] = cmul(s[k], f);
        scale_range(s, len, c);
        for (size_t k = 0; k < n; ++k) s[k] = csub(s[k], t[n + k]);
        if (g->kind == ORC_RX)
            for (size_t k = 0; k < n; ++k) s[n + k] = csub(s[n + k], t[k]);
        else
            for (size_t k = 0; k < n; ++k) s[n + k] = cadd(s[n + k], t[k]);
        free(t);
        return;
    }
    case ORC_RZ: {  /* rz.rs:72-102 */
        double h = 0.5 * g->p[0];
        scale_range(s, n, cpolar(1.0, -h));
        scale_range(s + n, n, cpolar(1.0, h));
        return;
    }
    case ORC_SWAP: {    /* swap.rs:34-58 */
        size_t q = len / 4;
        for (size_t k = q; k < 2 * q; ++k) { orc_cplx t = s[k]; s[k] = s[k + q]; s[k + q] = t; }
        return;
    }
    case ORC_U1:
        if (vec) { scale_range(s + n, n, cpolar(1.0, g->p[0])); return; }   /* u1.rs:77-84 */
        break;  /* mat path: default dense */
    case ORC_I:
        if (vec) return;   /* identity.rs:57-60 */
        break;
    default:
        break;
    }
    /* U1(mat), U2, U3, V, Vdg, I(mat), MATRIX: default dense body */
    orc_gate base = *g;
    base.nctrl = 0;
    size_t nb = orc_gate_nbits(&base);
    size_t d = (size_t)1 << nb;
    orc_cplx* m = (orc_cplx*)malloc(d * d * sizeof(orc_cplx));
    base_matrix(&base, m);
    body_default(m, nb, s, len);
    free(m);
}

/* ---------------------------------------------------------- permutation -- */

/* bit_permutation (gates.rs:53-80): the returned table is the inverse of
 * perm1, so that work[new] = v[perm[new]] puts the affected bits on top. */
static size_t* bit_permutation(size_t nr_bits, const size_t* bits, size_t nbits)
{
    size_t N = (size_t)1 << nr_bits;
    size_t* perm1 = (size_t*)malloc(N * sizeof(size_t));
    for (size_t i = 0; i < N; ++i) perm1[i] = i;
    size_t* ab = (size_t*)malloc(nbits * sizeof(size_t));
    memcpy(ab, bits, nbits * sizeof(size_t));
    size_t nab = nbits;
    while (nab > 0) {
        size_t s = ab[--nab];
        size_t idx = nr_bits - s - 1;
        size_t bit = (size_t)1 << idx;
        size_t lmask = bit - 1;
        size_t umask = ~(bit | lmask);
        for (size_t i = 0; i < N; ++i) {
            size_t v = perm1[i];
            perm1[i] = ((v & umask) >> 1) | (v & lmask) | ((v & bit) << s);
        }
        for (size_t k = 0; k < nab; ++k)
            if (ab[k] < s) ab[k] += 1;
    }
    free(ab);
    /* Permutation::inverse (permutation.rs:81-89) */
    size_t* inv = (size_t*)malloc(N * sizeof(size_t));
    for (size_t i = 0; i < N; ++i) inv[perm1[i]] = i;
    free(perm1);
    return inv;
}

/* ------------------------------------------------------------ the state -- */

static orc_state* alloc_state(size_t nr_bits, size_t nr_shots, size_t ncols)
{
    orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
    s->nr_bits = nr_bits;
    s->nr_shots = nr_shots;
    s->ncols = ncols;
    s->counts = (size_t*)calloc(ncols ? ncols : 1, sizeof(size_t));
    s->states = (orc_cplx*)calloc(((size_t)1 << nr_bits) * (ncols ? ncols : 1), sizeof(orc_cplx));
    return s;
}

/* VectorState::new (vectorstate.rs:41-53) */
orc_state* orc_new(size_t nr_bits, size_t nr_shots)
{
    orc_state* s = alloc_state(nr_bits, nr_shots, 1);
    s->states[0] = cmk(1, 0);
    s->counts[0] = nr_shots;
    return s;
}

/* VectorState::from_qubit_coefs (vectorstate.rs:62-83), kron_mat (cmatrix.rs:39-51) */
orc_state* orc_from_qubit_coefs(const orc_cplx* coefs, size_t nr_bits, size_t nr_shots)
{
    orc_state* s = alloc_state(nr_bits, nr_shots, 1);
    size_t len = 1;
    orc_cplx* cur = (orc_cplx*)malloc(sizeof(orc_cplx) << nr_bits);
    cur[0] = cmk(1, 0);
    for (size_t q = 0; q < nr_bits; ++q) {
        orc_cplx c0 = coefs[2 * q], c1 = coefs[2 * q + 1];
        double norm = sqrt((c0.re * c0.re + c0.im * c0.im) + (c1.re * c1.re + c1.im * c1.im));
        /* Complex / f64 divides both parts */
        orc_cplx b0 = cmk(c0.re / norm, c0.im / norm), b1 = cmk(c1.re / norm, c1.im / norm);
        /* kron(states, bit): row i*2 + j = states[i] * bit[j] (a1 * a0[[i,j]]) */
        for (size_t i = len; i-- > 0;) {
            orc_cplx a = cur[i];
            cur[2 * i] = cmul(b0, a);
            cur[2 * i + 1] = cmul(b1, a);
        }
        len *= 2;
    }
    memcpy(s->states, cur, len * sizeof(orc_cplx));
    free(cur);
    s->counts[0] = nr_shots;
    return s;
}

orc_state* orc_from_amplitudes(const orc_cplx* amps, size_t nr_bits, size_t nr_shots)
{
    orc_state* s = alloc_state(nr_bits, nr_shots, 1);
    memcpy(s->states, amps, sizeof(orc_cplx) << nr_bits);
    s->counts[0] = nr_shots;
    return s;
}

void orc_free(orc_state* s)
{
    if (!s) return;
    free(s->counts);
    free(s->states);
    free(s);
}

size_t orc_nr_bits(const orc_state* s) { return s->nr_bits; }
size_t orc_nr_cols(const orc_state* s) { return s->ncols; }
void orc_counts(const orc_state* s, size_t* out) { memcpy(out, s->counts, s->ncols * sizeof(size_t)); }

void orc_column(const orc_state* s, size_t col, orc_cplx* out)
{
    size_t N = (size_t)1 << s->nr_bits;
    for (size_t i = 0; i < N; ++i) out[i] = s->states[i * s->ncols + col];
}

/* apply_gate_mat_slice (gates.rs:121-152) on the whole row-major matrix. */
static void apply_gate_mat(orc_state* st, const orc_gate* g, const size_t* bits, size_t nbits)
{
    size_t n = st->nr_bits, nc = st->ncols, N = (size_t)1 << n;
    if (nbits == 1) {
        size_t bit = bits[0];
        size_t block = (size_t)1 << (n - bit);
        size_t nblocks = (size_t)1 << bit;
        for (size_t i = 0; i < nblocks; ++i)
            body(g, g->nctrl, 0, st->states + i * block * nc, block * nc);
        return;
    }
    size_t* perm = bit_permutation(n, bits, nbits);
    orc_cplx* work = (orc_cplx*)calloc(N * nc, sizeof(orc_cplx));
    /* per column: apply_vec_into (permutation.rs:141-149) */
    for (size_t c = 0; c < nc; ++c)
        for (size_t k = 0; k < N; ++k) work[k * nc + c] = st->states[perm[k] * nc + c];
    body(g, g->nctrl, 0, work, N * nc);
    /* per column: apply_inverse_vec_into (permutation.rs:155-163) */
    for (size_t c = 0; c < nc; ++c)
        for (size_t k = 0; k < N; ++k) st->states[perm[k] * nc + c] = work[k * nc + c];
    free(work);
    free(perm);
}

/* apply_gate_slice (gates.rs:86-115) on one contiguous column vector. */
static void apply_gate_vec(orc_cplx* v, size_t n, const orc_gate* g, const size_t* bits, size_t nbits)
{
    size_t N = (size_t)1 << n;
    if (nbits == 1) {
        size_t bit = bits[0];
        size_t block = (size_t)1 << (n - bit);
        size_t nblocks = (size_t)1 << bit;
        for (size_t i = 0; i < nblocks; ++i) body(g, g->nctrl, 1, v + i * block, block);
        return;
    }
    size_t* perm = bit_permutation(n, bits, nbits);
    orc_cplx* work = (orc_cplx*)calloc(N, sizeof(orc_cplx));
    for (size_t k = 0; k < N; ++k) work[k] = v[perm[k]];
    body(g, g->nctrl, 1, work, N);
    for (size_t k = 0; k < N; ++k) v[perm[k]] = work[k];
    free(work);
    free(perm);
}

/* QuState::apply_gate (vectorstate.rs:166-178) */
int orc_apply_gate(orc_state* s, const orc_gate* g, const size_t* bits, size_t nbits)
{
    if (orc_gate_nbits(g) != nbits) return ORC_ERR_NR_BITS;
    apply_gate_mat(s, g, bits, nbits);
    return ORC_OK;
}

/* apply_unary_gate_all (vectorstate.rs:180-189) */
int orc_apply_unary_gate_all(orc_state* s, const orc_gate* g)
{
    for (size_t b = 0; b < s->nr_bits; ++b) {
        int rc = orc_apply_gate(s, g, &b, 1);
        if (rc) return rc;
    }
    return ORC_OK;
}

/* collect_conditional_ranges (qustate.rs:100-127) */
size_t orc_collect_conditional_ranges(const size_t* counts, size_t ncols, const uint8_t* control,
                                      size_t* out_col, size_t* out_len, uint8_t* out_apply)
{
    size_t nr = 0, off = 0;
    for (size_t icol = 0; icol < ncols; ++icol) {
        size_t count = counts[icol];
        size_t begin = off;
        uint8_t prev = control[off] ? 1 : 0;
        for (size_t ibit = off + 1; ibit < off + count; ++ibit) {
            uint8_t cb = control[ibit] ? 1 : 0;
            if (cb != prev) {
                out_col[nr] = icol; out_len[nr] = ibit - begin; out_apply[nr] = prev; ++nr;
                begin = ibit;
                prev = !prev;
            }
        }
        if (begin < off + count) {
            out_col[nr] = icol; out_len[nr] = off + count - begin; out_apply[nr] = prev; ++nr;
        }
        off += count;
    }
    return nr;
}

/* apply_conditional_gate (vectorstate.rs:193-227) */
int orc_apply_conditional_gate(orc_state* s, const uint8_t* control, size_t ncontrol,
                               const orc_gate* g, const size_t* bits, size_t nbits)
{
    if (ncontrol != s->nr_shots) return ORC_ERR_NR_CONTROL;
    if (orc_gate_nbits(g) != nbits) return ORC_ERR_NR_BITS;
    size_t N = (size_t)1 << s->nr_bits;
    size_t cap = s->nr_shots + s->ncols;
    size_t* rcol = (size_t*)malloc(cap * sizeof(size_t));
    size_t* rlen = (size_t*)malloc(cap * sizeof(size_t));
    uint8_t* rapp = (uint8_t*)malloc(cap);
    size_t nr = orc_collect_conditional_ranges(s->counts, s->ncols, control, rcol, rlen, rapp);
    orc_cplx* ns = (orc_cplx*)calloc(N * (nr ? nr : 1), sizeof(orc_cplx));
    orc_cplx* colv = (orc_cplx*)malloc(N * sizeof(orc_cplx));
    for (size_t r = 0; r < nr; ++r) {
        for (size_t i = 0; i < N; ++i) colv[i] = s->states[i * s->ncols + rcol[r]];
        if (rapp[r]) apply_gate_vec(colv, s->nr_bits, g, bits, nbits);
        for (size_t i = 0; i < N; ++i) ns[i * nr + r] = colv[i];
    }
    free(colv);
    free(s->states);
    s->states = ns;
    free(s->counts);
    s->counts = (size_t*)malloc((nr ? nr : 1) * sizeof(size_t));
    for (size_t r = 0; r < nr; ++r) s->counts[r] = rlen[r];
    s->ncols = nr;
    free(rcol); free(rlen); free(rapp);
    return ORC_OK;
}

/* --------------------------------------------------------- measurement -- */

/* w0 per column (vectorstate.rs:249-261): sum of |a|^2 over qubit = 0 rows. */
int orc_marginal_w0(const orc_state* s, size_t qbit, double* w0s)
{
    if (qbit >= s->nr_bits) return ORC_ERR_QBIT;
    size_t nc = s->ncols;
    size_t block = (size_t)1 << (s->nr_bits - qbit - 1);
    size_t nblocks = (size_t)1 << qbit;
    double* part = (double*)malloc(nc * sizeof(double));
    for (size_t c = 0; c < nc; ++c) w0s[c] = 0.0;
    size_t off = 0;
    for (size_t b = 0; b < nblocks; ++b) {
        for (size_t c = 0; c < nc; ++c) part[c] = 0.0;
        for (size_t r = off; r < off + block; ++r)
            for (size_t c = 0; c < nc; ++c) {
                orc_cplx a = s->states[r * nc + c];
                part[c] += a.re * a.re + a.im * a.im;
            }
        for (size_t c = 0; c < nc; ++c) w0s[c] += part[c];
        off += 2 * block;
    }
    free(part);
    return ORC_OK;
}

/* collapse (vectorstate.rs:91-104) on one column vector. */
static void collapse_col(orc_cplx* v, size_t block, size_t nblocks, size_t offset, double norm_sq)
{
    size_t off = offset;
    for (size_t b = 0; b < nblocks; ++b) {
        for (size_t k = off; k < off + block; ++k) v[k] = cmk(0, 0);
        off += 2 * block;
    }
    orc_cplx f = cmk(1.0 / sqrt(norm_sq), 0.0);
    size_t N = 2 * block * nblocks;
    for (size_t k = 0; k < N; ++k) v[k] = cmul(v[k], f);
}

static size_t bernoulli_count(orc_rng* rng, size_t count, double p)
{
    size_t n0 = 0;
    for (size_t r = 0; r < count; ++r)
        if (orc_rng_uniform(rng) < p) ++n0;
    return n0;
}

/* measure_into / peek_into (vectorstate.rs:237-329, 346-393) */
int orc_measure_into(orc_state* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len,
                     orc_rng* rng, int collapse)
{
    if (qbit >= s->nr_bits) return ORC_ERR_QBIT;
    if (res_len < s->nr_shots) return ORC_ERR_SPACE;
    size_t nc = s->ncols, N = (size_t)1 << s->nr_bits;
    size_t block = (size_t)1 << (s->nr_bits - qbit - 1);
    size_t nblocks = (size_t)1 << qbit;
    double* w0s = (double*)malloc(nc * sizeof(double));
    orc_marginal_w0(s, qbit, w0s);
    size_t* n0s = (size_t*)malloc(nc * sizeof(size_t));
    size_t new_nr = 0;
    for (size_t c = 0; c < nc; ++c) {
        double p = w0s[c] < 1.0 ? w0s[c] : 1.0;
        n0s[c] = bernoulli_count(rng, s->counts[c], p);
        new_nr += (n0s[c] == 0 || n0s[c] == s->counts[c]) ? 1 : 2;
    }
    uint64_t one_mask = (uint64_t)1 << cbit, zero_mask = ~one_mask;
    size_t rs = 0;
    for (size_t c = 0; c < nc; ++c) {
        for (size_t r = rs; r < rs + n0s[c]; ++r) res[r] &= zero_mask;
        for (size_t r = rs + n0s[c]; r < rs + s->counts[c]; ++r) res[r] |= one_mask;
        rs += s->counts[c];
    }
    if (collapse) {
        orc_cplx* ns = (orc_cplx*)calloc(N * new_nr, sizeof(orc_cplx));
        size_t* ncounts = (size_t*)malloc(new_nr * sizeof(size_t));
        orc_cplx* v = (orc_cplx*)malloc(N * sizeof(orc_cplx));
        size_t ni = 0;
        for (size_t c = 0; c < nc; ++c) {
            double w0 = w0s[c];
            size_t n0 = n0s[c], count = s->counts[c];
            for (size_t i = 0; i < N; ++i) v[i] = s->states[i * nc + c];
            if (n0 == count) {
                collapse_col(v, block, nblocks, block, w0);
                ncounts[ni] = count;
            } else if (n0 == 0) {
                collapse_col(v, block, nblocks, 0, 1.0 - w0);
                ncounts[ni] = count;
            } else {
                collapse_col(v, block, nblocks, block, w0);
                ncounts[ni] = n0;
                for (size_t i = 0; i < N; ++i) ns[i * new_nr + ni] = v[i];
                ++ni;
                for (size_t i = 0; i < N; ++i) v[i] = s->states[i * nc + c];
                collapse_col(v, block, nblocks, 0, 1.0 - w0);
                ncounts[ni] = count - n0;
            }
            for (size_t i = 0; i < N; ++i) ns[i * new_nr + ni] = v[i];
            ++ni;
        }
        free(v);
        free(s->states);
        free(s->counts);
        s->states = ns;
        s->counts = ncounts;
        s->ncols = new_nr;
    }
    free(w0s);
    free(n0s);
    return ORC_OK;
}

/* rand 0.7 WeightedIndex::new + sample (third-party; used at
 * vectorstate.rs:123-126): cumulative_weights = [w0, w0+w1, ...] (N-1
 * entries), total = sum of all, chosen = u*total, index = #{cw <= chosen}. */
size_t orc_weighted_index_seq(const double* w, size_t len, double u)
{
    double total = w[0];
    double* cw = (double*)malloc((len > 1 ? len - 1 : 1) * sizeof(double));
    for (size_t i = 1; i < len; ++i) {
        cw[i - 1] = total;
        total += w[i];
    }
    double t = u * total;
    size_t lo = 0, hi = len - 1;   /* partition point of cw <= t */
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (cw[mid] <= t) lo = mid + 1; else hi = mid;
    }
    free(cw);
    return lo;
}

/*
 * Canonical F-ary tree walk (the order the GPU sampler uses).
 * Level 0 = weights; level l+1 entry g = left fold of level-l entries
 * [gF, min(gF+F, size_l)).  Levels are built until size <= F; the total T is
 * the left fold of the top level.  The walk starts at the top with offset 0;
 * within a group it keeps offset += S_j while offset + S_j <= t and descends
 * into the first child where that fails (or into the last positive child).
 */
size_t orc_tree_sample(const double* w, size_t len, double u, double* total_out)
{
    size_t idx = 0;
    orc_tree_sample_many(w, len, &u, 1, &idx, total_out);
    return idx;
}

void orc_tree_sample_many(const double* w, size_t len, const double* us, size_t nu, size_t* out,
                          double* total_out)
{
    const size_t F = ORC_TREE_FANOUT;
    size_t nlev = 1;
    size_t sizes[64];
    const double* lev[64];
    double* owned[64];
    sizes[0] = len;
    lev[0] = w;
    owned[0] = NULL;
    while (sizes[nlev - 1] > F) {
        size_t ps = sizes[nlev - 1];
        size_t ns = (ps + F - 1) / F;
        double* a = (double*)malloc(ns * sizeof(double));
        for (size_t g = 0; g < ns; ++g) {
            size_t e = (g + 1) * F < ps ? (g + 1) * F : ps;
            double acc = lev[nlev - 1][g * F];
            for (size_t k = g * F + 1; k < e; ++k) acc += lev[nlev - 1][k];
            a[g] = acc;
        }
        sizes[nlev] = ns;
        lev[nlev] = a;
        owned[nlev] = a;
        ++nlev;
    }
    size_t top = nlev - 1;
    double total = lev[top][0];
    for (size_t k = 1; k < sizes[top]; ++k) total += lev[top][k];
    if (total_out) *total_out = total;
    for (size_t r = 0; r < nu; ++r) {
        double t = us[r] * total;
        double offset = 0.0;
        size_t gstart = 0, glen = sizes[top], idx = 0;
        for (size_t l = top + 1; l-- > 0;) {
            const double* v = lev[l] + gstart;
            size_t last = 0;
            for (size_t j = 0; j < glen; ++j)
                if (v[j] > 0.0) last = j;
            size_t chosen = last;
            for (size_t j = 0; j < last; ++j) {
                double nx = offset + v[j];
                if (nx <= t) offset = nx;
                else { chosen = j; break; }
            }
            idx = gstart + chosen;
            if (l == 0) break;
            gstart = idx * F;
            size_t rem = sizes[l - 1] - gstart;
            glen = rem < F ? rem : F;
        }
        out[r] = idx;
    }
    for (size_t l = 1; l < nlev; ++l) free(owned[l]);
}

/* rand 0.7 WeightedIndex on many uniforms with one cumulative table. */
static void weighted_index_seq_many(const double* w, size_t len, const double* us, size_t nu, size_t* out)
{
    double total = w[0];
    double* cw = (double*)malloc((len > 1 ? len - 1 : 1) * sizeof(double));
    for (size_t i = 1; i < len; ++i) {
        cw[i - 1] = total;
        total += w[i];
    }
    for (size_t r = 0; r < nu; ++r) {
        double t = us[r] * total;
        size_t lo = 0, hi = len - 1;
        while (lo < hi) {
            size_t mid = lo + (hi - lo) / 2;
            if (cw[mid] <= t) lo = mid + 1; else hi = mid;
        }
        out[r] = lo;
    }
    free(cw);
}

/* reverse_bits / shuffle_bits (support.rs:50-75) */
uint64_t orc_reverse_bits(uint64_t idx, size_t nr_bits)
{
    uint64_t res = 0, sidx = idx;
    for (size_t i = 0; i < nr_bits; ++i) {
        res |= (sidx & 1) << (nr_bits - 1 - i);
        sidx >>= 1;
    }
    return res;
}

uint64_t orc_shuffle_bits(uint64_t idx, const size_t* bits, size_t nbits)
{
    uint64_t res = 0, sidx = idx;
    for (size_t k = 0; k < nbits; ++k) {
        res |= (sidx & 1) << bits[k];
        sidx >>= 1;
    }
    return res;
}

static int cmp_size(const void* a, const void* b)
{
    size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* measure_all_into_helper (vectorstate.rs:106-161) with injected uniforms. */
int orc_measure_all_uniforms(orc_state* s, const size_t* cbits, size_t ncbits,
                             const double* u01, uint64_t* res, size_t res_len,
                             int collapse, int mode)
{
    if (res_len < s->nr_shots) return ORC_ERR_SPACE;
    if (ncbits != s->nr_bits) return ORC_ERR_NR_MEAS;
    size_t nc = s->ncols, N = (size_t)1 << s->nr_bits;
    double* w = (double*)malloc(N * sizeof(double));
    size_t* samples = (size_t*)malloc((s->nr_shots ? s->nr_shots : 1) * sizeof(size_t));
    size_t* u_idx = (size_t*)malloc((s->nr_shots + 1) * sizeof(size_t));
    size_t* u_cnt = (size_t*)malloc((s->nr_shots + 1) * sizeof(size_t));
    size_t* u_col = (size_t*)malloc((s->nr_shots + 1) * sizeof(size_t));
    size_t nu = 0, run = 0;
    for (size_t c = 0; c < nc; ++c) {
        double total = 0.0;
        for (size_t i = 0; i < N; ++i) {
            orc_cplx a = s->states[i * nc + c];
            w[i] = mode == 0 ? orc_norm_sqr(a) : a.re * a.re + a.im * a.im;
            total += w[i];
        }
        if (!(total > 0.0)) {  /* WeightedIndex::new(...).unwrap() panics */
            free(w); free(samples); free(u_idx); free(u_cnt); free(u_col);
            return ORC_ERR_WEIGHTS;
        }
        size_t count = s->counts[c];
        if (mode == 0)
            orc_tree_sample_many(w, N, u01 + run, count, samples + run, NULL);
        else
            weighted_index_seq_many(w, N, u01 + run, count, samples + run);
        qsort(samples + run, count, sizeof(size_t), cmp_size);
        for (size_t r = 0; r < count;) {
            size_t e = r;
            while (e < count && samples[run + e] == samples[run + r]) ++e;
            u_idx[nu] = samples[run + r];
            u_cnt[nu] = e - r;
            u_col[nu] = c;
            ++nu;
            r = e;
        }
        run += count;
    }
    uint64_t mask = 0;
    for (size_t k = 0; k < ncbits; ++k) mask |= (uint64_t)1 << cbits[k];
    mask = ~mask;
    size_t off = 0;
    for (size_t k = 0; k < nu; ++k) {
        uint64_t rev = orc_reverse_bits(u_idx[k], s->nr_bits);
        uint64_t perm = orc_shuffle_bits(rev, cbits, ncbits);
        for (size_t r = off; r < off + u_cnt[k]; ++r) res[r] = (res[r] & mask) | perm;
        off += u_cnt[k];
    }
    if (collapse) {
        orc_cplx* ns = (orc_cplx*)calloc(N * (nu ? nu : 1), sizeof(orc_cplx));
        size_t* ncounts = (size_t*)malloc((nu ? nu : 1) * sizeof(size_t));
        for (size_t k = 0; k < nu; ++k) {
            ns[u_idx[k] * nu + k] = cmk(1, 0);
            ncounts[k] = u_cnt[k];
        }
        free(s->states);
        free(s->counts);
        s->states = ns;
        s->counts = ncounts;
        s->ncols = nu;
    }
    free(w); free(samples); free(u_idx); free(u_cnt); free(u_col);
    return ORC_OK;
}

int orc_measure_all_into(orc_state* s, const size_t* cbits, size_t ncbits,
                         uint64_t* res, size_t res_len, orc_rng* rng, int collapse)
{
    if (res_len < s->nr_shots) return ORC_ERR_SPACE;
    if (ncbits != s->nr_bits) return ORC_ERR_NR_MEAS;
    double* u = (double*)malloc((s->nr_shots ? s->nr_shots : 1) * sizeof(double));
    for (size_t r = 0; r < s->nr_shots; ++r) u[r] = orc_rng_uniform(rng);
    int rc = orc_measure_all_uniforms(s, cbits, ncbits, u, res, res_len, collapse, 0);
    free(u);
    return rc;
}

/* reset (vectorstate.rs:402-408): measure, then X where the result is 1. */
int orc_reset(orc_state* s, size_t bit, orc_rng* rng)
{
    uint64_t* m = (uint64_t*)calloc(s->nr_shots ? s->nr_shots : 1, sizeof(uint64_t));
    int rc = orc_measure_into(s, bit, 0, m, s->nr_shots, rng, 1);
    if (rc) { free(m); return rc; }
    uint8_t* control = (uint8_t*)malloc(s->nr_shots ? s->nr_shots : 1);
    for (size_t r = 0; r < s->nr_shots; ++r) control[r] = m[r] != 0;
    orc_gate x = {ORC_X, 0, {0, 0, 0}, 0, NULL};
    rc = orc_apply_conditional_gate(s, control, s->nr_shots, &x, &bit, 1);
    free(control);
    free(m);
    return rc;
}

/* reset_all (vectorstate.rs:410-415) */
void orc_reset_all(orc_state* s)
{
    free(s->states);
    free(s->counts);
    s->ncols = 1;
    s->states = (orc_cplx*)calloc((size_t)1 << s->nr_bits, sizeof(orc_cplx));
    s->states[0] = cmk(1, 0);
    s->counts = (size_t*)malloc(sizeof(size_t));
    s->counts[0] = s->nr_shots;
}

/* ---------------------------------------------------- random input state -- */

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void orc_random_normal_state(orc_cplx* out, size_t nr_bits, uint64_t seed)
{
    size_t N = (size_t)1 << nr_bits;
    double ss = 0.0;
    for (size_t i = 0; i < N; ++i) {
        uint64_t a = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * (uint64_t)i + 1));
        uint64_t b = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * (uint64_t)i + 2));
        double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0);  /* (0,1] */
        double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);        /* [0,1) */
        double r = sqrt(-2.0 * log(u1));
        double th = 2.0 * M_PI * u2;
        out[i] = cmk(r * cos(th), r * sin(th));
        ss += out[i].re * out[i].re + out[i].im * out[i].im;
    }
    double f = 1.0 / sqrt(ss);
    for (size_t i = 0; i < N; ++i) { out[i].re *= f; out[i].im *= f; }
}

/* Canonical sampling weights of an amplitude vector (orc_norm_sqr each). */
void orc_canonical_weights(const orc_cplx* a, size_t n, double* out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_norm_sqr(a[i]);
}

/*
 * oracle.h — CPU restatement of apogee-vm's per-shard proving hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * entry and bench.py's cpu_baseline leg load it; the product library never
 * links, loads or calls it.
 *
 * Representation: a field element is the reference's in-memory form, four
 * little-endian u64 Montgomery limbs with R = 2^256 (32 bytes). A G1 affine
 * point is 64 bytes, Montgomery x || y, all-zero for infinity. A Jacobian
 * point is 96 bytes X || Y || Z with Z = 0 for the identity.
 */
#ifndef APOGEE_ORACLE_H
#define APOGEE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fields: which = 0 for Fr, 1 for Fq ---- */
void orc_mul(int which, const uint64_t *a, const uint64_t *b, uint64_t *out);
void orc_add(int which, const uint64_t *a, const uint64_t *b, uint64_t *out);
void orc_sub(int which, const uint64_t *a, const uint64_t *b, uint64_t *out);
void orc_neg(int which, const uint64_t *a, uint64_t *out);
void orc_square(int which, const uint64_t *a, uint64_t *out);
int orc_inverse(int which, const uint64_t *a, uint64_t *out); /* 0 for zero */
void orc_pow(int which, const uint64_t *a, const uint64_t *exp, uint64_t *out);
void orc_from_u64(int which, uint64_t x, uint64_t *out);
/* canonical little-endian 32 bytes <-> Montgomery limbs; from_bytes returns 0 if >= modulus */
void orc_to_bytes(int which, const uint64_t *a, uint8_t *out32);
int orc_from_bytes(int which, const uint8_t *in32, uint64_t *out);

/* ---- G1 ---- */
void orc_g1_add(const uint64_t *p, const uint64_t *q, uint64_t *out);        /* jac + jac */
void orc_g1_add_affine(const uint64_t *p, const uint64_t *q_aff, uint64_t *out);
void orc_g1_double(const uint64_t *p, uint64_t *out);
void orc_g1_mul(const uint64_t *p, const uint64_t *scalar_fr, uint64_t *out); /* jac * Fr */
void orc_g1_mul_le_bytes(const uint64_t *p, const uint8_t *k32, uint64_t *out);
void orc_g1_to_affine(const uint64_t *p, uint64_t *out_aff);
void orc_g1_from_affine(const uint64_t *aff, uint64_t *out);
void orc_g1_generator(uint64_t *out_aff);
/* wire form: canonical x || y, all-zero infinity */
void orc_g1_affine_to_wire(const uint64_t *aff, uint8_t *out64);
int orc_g1_affine_from_wire(const uint8_t *in64, uint64_t *aff); /* validates */

/* ---- MSM (crates/curve/src/msm.rs) ---- */
size_t orc_msm_window_bits(size_t n);
/* threads = the rayon pool size the task split is computed for (>= 1).
 * workers = OS threads actually used to run the tasks (>= 1). */
int orc_msm(const uint64_t *bases_aff, const uint64_t *scalars, size_t n, size_t threads,
            size_t workers, uint64_t *out_jac);
int orc_msm_small_u32(const uint64_t *bases_aff, const uint32_t *scalars, size_t n,
                      size_t threads, size_t workers, uint64_t *out_jac);
/* [tau^i]G for i in [0, n), affine (the test toy SRS construction) */
void orc_g1_gen_mul_batch(const uint64_t *scalars_fr, size_t n, size_t workers, uint64_t *out_aff);
void orc_srs_powers(const uint64_t *tau, size_t n, size_t workers, uint64_t *out_aff);

/* ---- poly (crates/poly/src/lib.rs) ---- */
void orc_bind(uint64_t *table, size_t len, const uint64_t *r); /* in place, len -> len/2 */
void orc_eq_table(const uint64_t *r, size_t n, uint64_t *out); /* out: 2^n */
void orc_eq_eval(const uint64_t *r, const uint64_t *y, size_t n, uint64_t *out);
void orc_evaluate(const uint64_t *table, size_t num_vars, const uint64_t *point, uint64_t *out);

/* ---- transcript (crates/transcript/src/lib.rs) ---- */
typedef struct {
    uint64_t state[3][4];
    uint64_t input[2][4];
    uint64_t input_len;
    uint64_t output[2][4];
    uint64_t output_len;
} orc_transcript;
void orc_poseidon2_permute(uint64_t *state /* 3 x 4 limbs */);
void orc_tr_init(orc_transcript *t);
void orc_tr_observe(orc_transcript *t, const uint64_t *x);
void orc_tr_sample(orc_transcript *t, uint64_t *out);
void orc_tr_append_scalars(orc_transcript *t, uint64_t tag, const uint64_t *xs, size_t n);
void orc_tr_append_bytes(orc_transcript *t, uint64_t tag, const uint8_t *bytes, size_t len);
void orc_tr_challenge(orc_transcript *t, uint64_t tag, uint64_t *out);
void orc_g1_limbs(const uint8_t *wire64, uint64_t *out /* 4 x 4 limbs */);
void orc_tr_append_g1(orc_transcript *t, uint64_t tag, const uint8_t *wire64s, size_t k);

/* ---- sumcheck ----
 * A gate is a list of terms coef * x[f0] * x[f1] * x[f2]; a factor index of -1
 * is absent. The reference gate (crates/sumcheck/src/lib.rs:88-112) has at
 * most two factors; a third factor is the degree-4 extension. */
typedef struct {
    uint64_t coef[4];
    int32_t f[3];
} orc_term;
/* round polynomial of degree d from its values at 0..d, ascending coefficients */
void orc_interpolate(const uint64_t *vals, size_t d, uint64_t *coeffs);
/* sumcheck::prove_zerocheck: tables are consumed (left bound); r drawn from t. */
void orc_prove_zerocheck(const orc_term *terms, size_t n_terms, uint64_t **tables, size_t k,
                         size_t num_vars, orc_transcript *t, uint64_t *rounds_out,
                         uint64_t *final_evals_out);
/* gkr::prove_sumcheck with a gate summand, from a given eq point. degree = max
 * factor count + 1. rounds_out: num_vars x (degree+1) elements; point_out:
 * num_vars elements. workers: threads for the pair loop. */
void orc_prove_sumcheck(const uint64_t *eq_point, size_t num_vars, const orc_term *terms,
                        size_t n_terms, uint64_t **tables, size_t k, orc_transcript *t,
                        size_t workers, uint64_t *rounds_out, uint64_t *point_out);
void orc_witness_digest(uint64_t **tables, size_t k, size_t num_vars, uint64_t *out);
/* out_o[i] = sum_c w[o k + c] * cols[c n + i] for o < outs */
void orc_axpy(size_t n, const uint64_t *a, const uint64_t *x, uint64_t *y);
void orc_product_claim(const uint64_t *r, size_t n, const uint64_t *const *tables, size_t f, uint64_t *out);
size_t orc_msm_recode(const uint64_t *scalar_mont, size_t w, int32_t *digits);
uint64_t orc_msm_nonzero_digits(const uint64_t *scalars, size_t n, size_t w);
void orc_combine(const uint64_t *cols, size_t n, size_t k, const uint64_t *w, size_t outs,
                 uint64_t *out);

/* ---- seeded inputs ---- */
uint64_t orc_splitmix_next(uint64_t *state);
void orc_splitmix_next_fr(uint64_t *state, uint64_t *out); /* draw-mask-reject */
typedef struct {
    uint32_t key[8];
    uint32_t counter;
    uint8_t buf[64];
    uint32_t pos;
} orc_chacha;
void orc_chacha_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3],
                      uint8_t out[64]);
void orc_chacha_init(orc_chacha *c, uint64_t seed);
uint32_t orc_chacha_u32(orc_chacha *c);
void orc_chacha_fr(orc_chacha *c, uint64_t *out);
/* witness-like mix: 20% zero, 10% one, 50% u32, 20% uniform Fr */
void orc_chacha_witness(orc_chacha *c, uint64_t *out);
void orc_fill_uniform(uint64_t seed, uint64_t *out, size_t n);
void orc_fill_witness(uint64_t seed, uint64_t *out, size_t n);
/* counter-mode draws (element i independent of the others); kind 0 uniform
 * Fr, 1 witness mix */
void orc_fill_ctr(uint64_t seed, int kind, uint64_t *out, size_t n);

/* ---- Mercury opening (crates/pcs/src/lib.rs:313-492), oracle/mercury.c ----
 * srs: srs_len affine powers; table: 2^num_vars Fr (the evaluation table read
 * as coefficients); cm_wire: the commitment as passed (64 B wire form); u:
 * num_vars Fr; tr: the transcript, advanced as the reference advances it.
 * Out: v (Fr) and the 704-byte proof (8 wire points, then 6 canonical Fr).
 * Returns 0, or -2 unsupported num_vars, -3 SRS too small, -4 degenerate
 * challenge, -5..-8 a broken internal identity. */
int orc_mercury_open(const uint64_t *srs_aff, size_t srs_len, const uint64_t *table, size_t num_vars,
                     const uint8_t *cm_wire, const uint64_t *u, orc_transcript *tr, size_t workers,
                     uint64_t *v_out, uint8_t *proof_out);

/* batch_open_stacked (crates/pcs/src/lib.rs:701-789): cols holds the stacks'
 * columns back to back in slot order, 2^width Fr each; stack_sizes[s] columns
 * in stack s (at most 2^nr); cms_wire: one 64-byte commitment per stack; the
 * opening point is u (width) || r (nr). Out: vs (one Fr per stack) and the
 * 704-byte proof. Returns 0, -9 EmptyBatch, -10 BatchLengthMismatch, -11 an
 * undecodable commitment, or open's codes. */
int orc_mercury_batch_open_stacked(const uint64_t *srs_aff, size_t srs_len, const uint64_t *cols,
                                   const uint32_t *stack_sizes, size_t nstacks, size_t width,
                                   const uint8_t *cms_wire, const uint64_t *u, const uint64_t *r, size_t nr,
                                   orc_transcript *tr, size_t workers, uint64_t *vs_out, uint8_t *proof_out);

/* ---- the general GKR layer, oracle/gkr.c ----
 * Gate shapes carry the wire tags of constraints::GateDef; virtual kinds the
 * wire order of constraints::VirtualKind (crates/constraints/src/lib.rs:
 * 215-250,310-360). A source is (kind << 24) | index: lower[index] (a column
 * of the layer, or child 0 of one), upper[index] (child 1), virtuals[index],
 * or cached entry index. Coefficients are Montgomery limbs, per gate in
 * GateDef::coefficients order with challenges substituted; split is an
 * AffineProduct's left length or a Quadratic's linear length. */
enum {
    ORC_GATE_LINEAR = 0,
    ORC_GATE_PRODUCT = 1,
    ORC_GATE_MASK = 2,
    ORC_GATE_AFFINE_PRODUCT = 3,
    ORC_GATE_TREE_PRODUCT = 4,
    ORC_GATE_QUADRATIC = 5,
    ORC_GATE_TREE_CROSS = 6,
};
enum {
    ORC_VIRT_ROW_INDEX = 0,
    ORC_VIRT_RAM_LIVE = 1,
    ORC_VIRT_RANGE19 = 2,
    ORC_VIRT_RANGE16 = 3,
    ORC_VIRT_XOR8_A = 4,
    ORC_VIRT_XOR8_B = 5,
    ORC_VIRT_XOR8_OUT = 6,
};
enum { ORC_SRC_LOWER = 0, ORC_SRC_UPPER = 1, ORC_SRC_VIRTUAL = 2, ORC_SRC_CACHED = 3 };
#define ORC_SRC_KIND_SHIFT 24
#define ORC_SRC_INDEX_MASK 0xffffffu
#define ORC_TAG_SUMCHECK_ROUND 4
#define ORC_TAG_SUMCHECK_CHALLENGE 5
typedef struct {
    uint32_t shape, src_begin, src_end, coef_begin, split;
} orc_gkr_gate;
typedef struct {
    const orc_gkr_gate *gates; /* cached entries, producing, enforcing */
    uint32_t ncached, nproducing, nenforcing;
    const uint32_t *sources;
    uint32_t nsources;
    const uint64_t *coeffs;
    uint32_t ncoeffs;
    const uint32_t *virtual_kinds;
    uint32_t nvirtual;
    uint32_t halving;
} orc_gkr_list;
void orc_gkr_eval_gate(uint32_t shape, const uint64_t *coeffs, uint32_t split, const uint64_t *values,
                       uint32_t nvalues, uint64_t *out);
void orc_gkr_virtual_at_row(uint32_t kind, uint64_t row, uint64_t *out);
void orc_gkr_virtual_at_point(uint32_t kind, const uint64_t *point, size_t n, uint64_t *out);
void orc_gkr_gate_values(const orc_gkr_list *l, const uint64_t *lower, size_t nlower, const uint64_t *upper,
                         size_t nupper, const uint64_t *virtuals, uint64_t *out);
void orc_gkr_summand(const orc_gkr_list *l, const uint64_t *weights, const uint64_t *lower, size_t nlower,
                     const uint64_t *upper, size_t nupper, const uint64_t *virtuals, uint64_t *out);
void orc_gkr_forward_layer(const orc_gkr_list *l, uint64_t *const *cols, size_t ncols, size_t num_vars,
                           uint64_t **out);
void orc_gkr_prove_sumcheck(const orc_gkr_list *l, const uint64_t *weights, const uint64_t *eq_point, size_t n,
                            uint64_t *const *lower, uint64_t *const *upper, size_t ncols, orc_transcript *tr,
                            uint64_t *rounds_out, uint64_t *point_out);

#ifdef __cplusplus
}
#endif
#endif

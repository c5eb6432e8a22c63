/*
 * apogee_gpu.h — C ABI of the MI355X (gfx950) per-shard proving path:
 * the commitment MSM over BN254 G1 and the sumcheck round evaluation + fold.
 *
 * Everything crosses as plain bytes:
 *   - Fr / Fq: 32 bytes, the reference's in-memory Montgomery image
 *     (Fr::to_memory_bytes, crates/field/src/lib.rs:321): eight little-endian
 *     u32 words of x * 2^256 mod m, fully reduced.
 *   - G1 affine: 64 bytes, Montgomery x || y; all-zero is the point at
 *     infinity ((0,0) is off the curve, crates/curve/src/g1.rs:106-111).
 *   - "wire" G1: canonical x || y (G1Affine::to_bytes, g1.rs:92-100).
 * Inputs are borrowed for the call, outputs go to caller-owned buffers.
 * Entries return APG_OK (0) or a negative code and never abort or throw;
 * apg_last_error() describes the last failure on the calling thread.
 * Calls on one apg_ctx are serialised; use one context per worker thread
 * for concurrency. Results never depend on concurrency.
 */
#ifndef APOGEE_GPU_H
#define APOGEE_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APG_OK 0
#define APG_ERR_ARG (-2)
#define APG_ERR_HIP (-3)
#define APG_ERR_OOM (-4)
#define APG_ERR_INTERNAL (-5)
#define APG_ERR_CALLBACK (-6)
#define APG_ERR_DEGENERATE (-7) /* pcs::open: PcsError::DegenerateChallenge */

typedef struct apg_ctx apg_ctx;
typedef struct apg_srs apg_srs;
typedef struct apg_transcript apg_transcript;

const char *apg_last_error(void);
int apg_version(void);

/* ---- context: one HIP stream + scratch arena on one device ---- */
int apg_ctx_create(int device, apg_ctx **out);
void apg_ctx_destroy(apg_ctx *ctx);
int apg_ctx_synchronize(apg_ctx *ctx);

/* ---- kernel timing, recorded with HIP events on the context's stream.
 * No counterpart in the reference (its timing is tools/bench's wall clock,
 * tools/bench/src/timing.rs:9-47); for benchmarks and roofline reports. */
typedef struct {
    double accumulate_ms;          /* MSM bucket-accumulation kernel */
    uint64_t accumulate_launches;
    uint64_t accumulate_entries;   /* nonzero signed digits it processed */
    double msm_ms;                 /* whole commits, host wall clock */
    uint64_t msm_calls;
    uint64_t msm_points;           /* scalars committed (points x columns) */
    double sumcheck_kernel_ms;     /* fused fold + round-evaluation kernels */
    uint64_t sumcheck_launches;
    uint64_t sumcheck_bytes;       /* algorithmic HBM bytes of those launches */
    double sumcheck_ms;            /* whole sumchecks, host wall clock */
} apg_stats;
int apg_ctx_set_profiling(apg_ctx *ctx, int on);
int apg_ctx_stats(apg_ctx *ctx, apg_stats *out);
int apg_ctx_reset_stats(apg_ctx *ctx);

/* ---- seeded synthetic inputs: counter-mode ChaCha20 keyed by seed, element
 * i from block counter i (nonce = attempt, domain); kind 0 = uniform Fr
 * (32-byte candidates, top two bits cleared, reject >= p: the draw rule of
 * crates/pcs/tests/common/mod.rs:114-123), kind 1 = witness-like mix (20%
 * zero, 10% one, 50% u32, 20% uniform Fr). n Montgomery elements, on the host
 * or into device memory (same values). No counterpart in the reference (its
 * benches draw from splitmix64). */
int apg_synthetic_fill(uint64_t seed, int kind, size_t n, uint8_t *out_mont);
int apg_synthetic_fill_dev(apg_ctx *ctx, uint64_t seed, int kind, size_t n, void *out_dev);

/* ---- the device's Montgomery multiplication (the 32-bit-word CIOS the
 * kernels run), compiled for the host so a CPU-only test can check it on
 * many inputs: out[i] = a[i] * b[i] * 2^-256 mod m, which = 0 for Fr, 1 for
 * Fq. No counterpart in the reference. */
int apg_debug_mul32_batch(int which, const uint8_t *a_mont, const uint8_t *b_mont, size_t n,
                          uint8_t *out_mont);
/* The transcript's Poseidon2 permutation (crates/transcript/src/lib.rs:
 * 273-292) on n states of three Montgomery words (96 B each): form 0 the one
 * the transcript uses, 1 the portable 64-bit multiply, 2 the x86-64
 * mulx/adcx/adox multiply (APG_ERR_ARG where the CPU lacks BMI2 or ADX). For
 * CPU-only checks that every form agrees with the reference's vectors; no
 * counterpart in the reference. */
int apg_debug_poseidon2_batch(int form, const uint8_t *states_mont, size_t n, uint8_t *out_mont);
/* pcs::open (crates/pcs/src/lib.rs:313-492): the Mercury opening of the
 * commitment cm_wire (64-byte wire form, absorbed as passed) at u (num_vars
 * Montgomery Fr) for the 2^num_vars evaluation table, read as coefficients.
 * The transcript advances exactly as the reference's, including the final
 * PAIRING_MERGE squeeze. Out: v = fhat(u) (Montgomery Fr) and the 704-byte
 * MercuryProof::to_bytes (8 wire points h, q, g, s, d, pi_z, w, w', then the
 * canonical g(z), g(1/z), h(z), h(1/z), S(z), S(1/z)). num_vars must be even
 * and in [2, 30] (check_num_vars, crates/pcs-verify/src/lib.rs:97) and the SRS
 * hold 2^num_vars powers (PcsError::SrsTooSmall): APG_ERR_ARG otherwise;
 * APG_ERR_DEGENERATE for the reference's DegenerateChallenge.
 * _dev: the table is device memory (2^num_vars x 32 B). */
int apg_pcs_open(apg_ctx *ctx, apg_srs *srs, const uint8_t *table_mont, size_t num_vars,
                 const uint8_t cm_wire[64], const uint8_t *u_mont, apg_transcript *tr, uint8_t v_mont[32],
                 uint8_t proof[704]);
int apg_pcs_open_dev(apg_ctx *ctx, apg_srs *srs, const void *table_dev, size_t num_vars,
                     const uint8_t cm_wire[64], const uint8_t *u_mont, apg_transcript *tr, uint8_t v_mont[32],
                     uint8_t proof[704]);
/* batch_open_stacked (crates/pcs/src/lib.rs:701-789; batch_open, :689, is
 * nr = 0 with one column per stack): cols holds the stacks' columns back to
 * back in slot order, 2^width Montgomery Fr each; stack s has stack_sizes[s]
 * columns (1 .. 2^nr); cms_wire one 64-byte commitment per stack, absorbed as
 * passed. Opens cm* = sum rho^i cm_i at u (width) || r (nr). Out: vs_mont, one
 * value per stack (sum_j eq(r, j) col_j(u)), and the 704-byte proof.
 * APG_ERR_ARG for the reference's EmptyBatch, BatchLengthMismatch,
 * UnsupportedNumVars (width + nr) and SrsTooSmall; APG_ERR_DEGENERATE as open.
 * _dev: cols is device memory. */
int apg_pcs_batch_open_stacked(apg_ctx *ctx, apg_srs *srs, const uint8_t *cols_mont, const uint32_t *stack_sizes,
                               size_t nstacks, size_t width, const uint8_t *cms_wire, const uint8_t *u_mont,
                               const uint8_t *r_mont, size_t nr, apg_transcript *tr, uint8_t *vs_mont,
                               uint8_t proof[704]);
int apg_pcs_batch_open_stacked_dev(apg_ctx *ctx, apg_srs *srs, const void *cols_dev, const uint32_t *stack_sizes,
                                   size_t nstacks, size_t width, const uint8_t *cms_wire, const uint8_t *u_mont,
                                   const uint8_t *r_mont, size_t nr, apg_transcript *tr, uint8_t *vs_mont,
                                   uint8_t proof[704]);

/* The MSM accumulation's Fq arithmetic in nine 29-bit words with lazy
 * reduction (field29.cuh), on the host (ctx NULL) or in a kernel on ctx's
 * device (host buffers in and out): op 0 = a b, 1 = a + b, 2 = a - b,
 * 3 = ((a + b)(a - b) + a^2) - b, 4 = a^2 (the squaring), 5 = a b + a a (two
 * products under one reduction); outputs fully reduced Montgomery Fq. */
int apg_debug_fq29_batch(apg_ctx *ctx, int op, const uint8_t *a_mont, const uint8_t *b_mont, size_t n,
                         uint8_t *out_mont);
/* The MSM bucket loop's chains: for each case c, the sum over points
 * i in [offsets[c], offsets[c+1]) of (neg[i] ? -P_i : P_i), P_i affine (64 B
 * Montgomery x || y, all-zero = infinity), as an affine point (all-zero for
 * the identity). form 0 = lazy 29-bit arithmetic with mixed adds, 1 = 32-bit
 * with mixed adds, 2 = lazy 29-bit with general XYZZ adds of the points
 * scaled to Z = i + 2; ctx NULL runs on the host. */
int apg_debug_accumulate29(apg_ctx *ctx, int form, const uint8_t *points, const uint8_t *neg,
                           const uint64_t *offsets, size_t ncases, uint8_t *out_affine);
/* The device's issue rate of the 32 x 32 -> 64 multiply-add (v_mad_u64_u32)
 * that the field multiplies are built from, the denominator of the MSM and
 * sumcheck rooflines: 8 blocks of 256 threads per CU, each thread running 32
 * independent chains of `iters` multiply-adds; the best of `reps` timed
 * launches (HIP events) after one untimed launch. No reference counterpart
 * (a measurement aid). */
int apg_debug_imul_peak(apg_ctx *ctx, int iters, int reps, double *best_ms, double *imul_per_s);

/* ---- SRS ingestion (crates/srs/src/lib.rs:175-240, docs/spec/srs.md §4).
 * The archive: magic "APOGESRS", version u32 = 1, power u32, count u64 =
 * 2^power, the two G2 points (128 B each, canonical x.c0 x.c1 y.c0 y.c1),
 * then count canonical G1 points (64 B each, x || y, all-zero = infinity).
 * srs_error takes SrsError's variant: */
#define APG_SRS_IO 1
#define APG_SRS_TRUNCATED 2
#define APG_SRS_BAD_MAGIC 3
#define APG_SRS_BAD_VERSION 4       /* detail = the version read */
#define APG_SRS_BAD_SECTION 5       /* count is not 2^power, or power > 30 */
#define APG_SRS_INVALID_POINT 6     /* detail = index; count, count + 1 = the G2 points */
/* Srs::load: every G1 power decoded and validated on the device
 * (G1Affine::from_bytes: canonical coordinates, on the curve), the header
 * rules and both G2 points (on the twist, r P = O) on the host, in the
 * reference's order. On an SrsError, APG_ERR_ARG with *srs_error and
 * *detail set; with several invalid G1 points, detail is the lowest index.
 * g2_out receives the two G2 points as stored. */
int apg_srs_load(apg_ctx *ctx, const char *path, apg_srs **out, uint8_t g2_out[256], int *srs_error,
                 uint64_t *detail);
/* The header rules and the two G2 points alone (host only, no device). */
int apg_srs_archive_check(const char *path, uint32_t *power, uint8_t g2_out[256], int *srs_error,
                          uint64_t *detail);
/* Srs::save of a resident SRS whose length is a power of two, with the
 * caller's G2 points (stored as given). */
int apg_srs_save(apg_ctx *ctx, const apg_srs *srs, const uint8_t g2[256], const char *path);
/* n canonical 64-byte G1 points decoded and validated on the device into a
 * resident SRS (ptau::read_canonical_g1, crates/srs/src/ptau.rs:307-328).
 * APG_ERR_ARG and *bad_index (the lowest) for an invalid point. */
int apg_srs_upload_wire(apg_ctx *ctx, const uint8_t *wire, size_t n, apg_srs **out, uint64_t *bad_index);
/* G2Affine::from_bytes validity of one 128-byte point (host): 1 or 0. */
int apg_g2_wire_valid(const uint8_t wire[128]);
/* G1Projective::batch_to_affine (crates/curve/src/g1.rs:328-347) on the
 * device: n Jacobian points (96 B: x, y, z Montgomery) to affine (64 B,
 * all-zero for z = 0). */
int apg_g1_batch_to_affine(apg_ctx *ctx, const uint8_t *jac_mont, size_t n, uint8_t *out_affine);

/* ---- device memory for callers without their own allocator ---- */
int apg_dev_alloc(apg_ctx *ctx, size_t bytes, void **out);
int apg_dev_free(apg_ctx *ctx, void *p);
int apg_dev_upload(apg_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int apg_dev_download(apg_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
/* device-to-device copy by the library's own streaming kernel (k_copy16:
 * 16-byte loads and stores, bytes a multiple of 16); it reads and writes
 * exactly `bytes` each, the known-size reference for HBM traffic counters */
int apg_dev_copy(apg_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* ---- SRS residency.
 * Replaces the &Srs every pcs::commit borrows (crates/srs/src/lib.rs:96-276,
 * Srs::g1 :244): the powers [tau^i]_1 stay in HBM for the life of the handle.
 * apg_srs_generate builds [tau^i]G on the device (the toy SRS construction of
 * crates/pcs/tests/common/mod.rs:43-60, for any tau).
 * apg_srs_prepare builds the window-shifted base tables for commits of up to
 * n points ahead of time (otherwise the first commit of a size class does). */
int apg_srs_upload(apg_ctx *ctx, const uint8_t *bases_mont, size_t n, apg_srs **out);
int apg_srs_generate(apg_ctx *ctx, const uint8_t tau_mont[32], size_t n, apg_srs **out);
int apg_srs_download(apg_ctx *ctx, const apg_srs *srs, size_t offset, size_t n, uint8_t *out);
size_t apg_srs_len(const apg_srs *srs);
int apg_srs_prepare(apg_ctx *ctx, apg_srs *srs, size_t n);
size_t apg_srs_device_bytes(const apg_srs *srs);
/* Pippenger window width per size class (floor(log2 n)); results never
 * depend on it. apg_msm_set_window_bits overrides one class (c = 0 restores
 * the built-in table; 3 <= c <= 17), for the autotuner
 * (apogee-vm_amd/autotune.py). Replaces the fixed rule of
 * crates/curve/src/msm.rs:112-118. A new width builds its own SRS table on
 * the next commit of that class. */
int apg_msm_set_window_bits(int log2n, int c);
int apg_msm_window_bits(size_t n);
/* Accumulation slices per resident thread per size class (the bucket tile:
 * more rounds, shorter slices); results never depend on it.
 * apg_msm_set_slice_rounds overrides one class (0 restores the built-in
 * table; 1 <= rounds <= 16), for the autotuner. No counterpart in the
 * reference, whose tasks are windows x chunks (crates/curve/src/msm.rs:242-245). */
int apg_msm_set_slice_rounds(int log2n, int rounds);
int apg_msm_slice_rounds(size_t n);
void apg_srs_release(apg_srs *srs);

/* ---- MSM over explicit bases.
 * apg_msm_g1_fr replaces curve::msm::msm (crates/curve/src/msm.rs:48);
 * apg_msm_g1_u32 replaces curve::msm::msm_small_u32 (msm.rs:79). n == 0
 * gives the identity (msm.rs:55-57); the reference's only error,
 * LengthMismatch (msm.rs:29-33), cannot arise with one length. */
int apg_msm_g1_fr(apg_ctx *ctx, const uint8_t *bases_mont, const uint8_t *scalars_mont, size_t n,
                  uint8_t out_mont[64]);
int apg_msm_g1_u32(apg_ctx *ctx, const uint8_t *bases_mont, const uint32_t *scalars, size_t n,
                   uint8_t out_mont[64]);

/* ---- commitments over the resident SRS.
 * k columns of n scalars each, column-major, against srs powers [0, n):
 * pcs::commit / column_msm (crates/pcs/src/lib.rs:237-247, :280-298) for each
 * column; one call covers a whole commit_all (crates/prover/src/lib.rs:444).
 * No size rule is applied here (that is kzg_commit, crates/srs/src/kzg.rs:20);
 * pcs::commit's check_num_vars stays with the caller.
 * The _dev forms take device pointers (scalars resident in HBM). */
int apg_commit_fr(apg_ctx *ctx, apg_srs *srs, const uint8_t *scalars_mont, size_t n, size_t k,
                  uint8_t *out_mont);
int apg_commit_u32(apg_ctx *ctx, apg_srs *srs, const uint32_t *scalars, size_t n, size_t k,
                   uint8_t *out_mont);
int apg_commit_fr_dev(apg_ctx *ctx, apg_srs *srs, const void *scalars_dev, size_t n, size_t k,
                      uint8_t *out_mont);

/* ---- sumcheck.
 * A gate term is coef * x[factor[0]] * x[factor[1]] * x[factor[2]], factor
 * -1 = absent (crates/sumcheck/src/lib.rs:88-112 allows two factors; a third
 * is the degree-4 extension). Round polynomials have d + 1 coefficients,
 * d = max(3, 1 + max factors). */
typedef struct {
    uint8_t coef_mont[32];
    int32_t factor[3];
} apg_term;

/* Per-round hook: receives the round's d + 1 ascending coefficients and
 * writes the challenge. Return 0, or nonzero to abort the proof. */
typedef int (*apg_round_fn)(void *user, const uint8_t *coeffs_mont, size_t n_coeffs,
                            uint8_t challenge_mont[32]);

/* gkr::prove_sumcheck (crates/gkr/src/lib.rs:792) with a gate summand: the
 * rounds of sum_x eq(eq_point, x) * G(tables(x)). tables: k host pointers
 * to 2^num_vars elements each (not modified). Outputs: rounds_out
 * num_vars * (d + 1) elements, point_out num_vars, final_evals_out k (the
 * tables bound to the point). */
int apg_sumcheck_prove(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                       const uint8_t *const *tables_mont, size_t k, size_t num_vars,
                       const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                       uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out);
/* the same with device-resident tables (consumed: contents are left undefined) */
int apg_sumcheck_prove_dev(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                           void *const *tables_dev, size_t k, size_t num_vars,
                           const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                           uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out);

/* ---- the general GKR layer.
 * A gate list of a circuit artifact, flattened as gkr_verify::ResolvedList::new
 * resolves it (crates/gkr-verify/src/lib.rs:330-395): gates[] holds the
 * cached entries, then the producing gates, then the enforcing gates. Gate i
 * has shape = the GateDef wire tag (0 Linear, 1 Product, 2 MaskIntoIdentity,
 * 3 AffineProduct, 4 TreeProduct, 5 Quadratic, 6 TreeCross;
 * crates/constraints/src/lib.rs:310-360), reads sources[src_begin, src_end)
 * in eval_gate's value order, and takes its coefficients from
 * coeffs_mont[coef_begin ..] in GateDef::coefficients order
 * (crates/constraints/src/lib.rs:393-430) with every challenge slot
 * substituted; split is an AffineProduct's left length or a Quadratic's
 * linear length. A source is (kind << 24) | index with kind 0 = a column of
 * the layer (child 0 of it for a halving list), 1 = child 1, 2 = virtual
 * table index (virtual_kinds[], VirtualKind wire order 0 RowIndex ..
 * 6 Xor8Out), 3 = cached entry index. */
typedef struct {
    uint32_t shape, src_begin, src_end, coef_begin, split;
} apg_gkr_gate;
typedef struct {
    const apg_gkr_gate *gates;
    uint32_t ncached, nproducing, nenforcing;
    const uint32_t *sources;
    uint32_t nsources;
    const uint8_t *coeffs_mont; /* 32 B each */
    uint32_t ncoeffs;
    const uint32_t *virtual_kinds;
    uint32_t nvirtual;
    uint32_t halving;
} apg_gkr_list;

/* gkr::produce (crates/gkr/src/lib.rs:300-336): the producing gates of the
 * list at every row. cols: ncols columns of layer k, 2^num_vars elements
 * each; out: one column per producing gate, 2^num_vars rows for a row-wise
 * list, 2^(num_vars - 1) for a halving one (row y reads rows y and
 * y + 2^(num_vars - 1)). APG_ERR_ARG for a malformed list. */
int apg_gkr_forward_layer(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *const *cols_mont, size_t ncols,
                          size_t num_vars, uint8_t *const *out_mont);
int apg_gkr_forward_layer_dev(apg_ctx *ctx, const apg_gkr_list *list, const void *const *cols_dev, size_t ncols,
                              size_t num_vars, void *const *out_dev);

/* gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899) for LayerSummand
 * { weights } (crates/gkr/src/lib.rs:687-701): the rounds of
 * sum_y eq(eq_point, y) * sum_j weights[j] gate_j, weights over the producing
 * then the enforcing gates. lower: ncols tables of 2^num_vars elements (child
 * 0 of each column for a halving list); upper: child 1 of each (halving
 * lists only, else NULL). Outputs: rounds_out num_vars x 4 coefficients,
 * point_out num_vars challenges, finals_out the tables bound to the point
 * (lower, then upper). round_fn absorbs each round and returns the challenge
 * (apg_transcript_round does SUMCHECK_ROUND / SUMCHECK_CHALLENGE).
 * APG_ERR_ARG for a malformed list or a weighted gate of degree above 2 (the
 * artifact's degree law). The _dev form takes device tables and leaves them
 * unmodified (the folds go to the context's scratch). */
int apg_gkr_prove_sumcheck(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *weights_mont,
                           const uint8_t *eq_point_mont, size_t num_vars, const uint8_t *const *lower_mont,
                           const uint8_t *const *upper_mont, size_t ncols, apg_round_fn round_fn, void *user,
                           uint8_t *rounds_out, uint8_t *point_out, uint8_t *finals_out);
int apg_gkr_prove_sumcheck_dev(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *weights_mont,
                               const uint8_t *eq_point_mont, size_t num_vars, void *const *lower_dev,
                               void *const *upper_dev, size_t ncols, apg_round_fn round_fn, void *user,
                               uint8_t *rounds_out, uint8_t *point_out, uint8_t *finals_out);

/* Table backings (crates/poly/src/lib.rs:46-58, PolyBacking): Fr entries
 * (32 B Montgomery), or narrow entries kept at their width on the device and
 * lifted to Fr as they are read: U1 (u64 limbs, entry i = bit i % 64 of limb
 * i / 64), U8, U16, U32 (an entry v is v in Fr). */
enum { APG_BACKING_FR = 0, APG_BACKING_U1 = 1, APG_BACKING_U8 = 2, APG_BACKING_U16 = 3, APG_BACKING_U32 = 4 };
/* apg_sumcheck_prove_dev over tables of the given backings (one per table):
 * the first round reads narrow tables through the lift and the second folds
 * them into Fr scratch, as MultilinearPoly::bind lifts on its first call
 * (crates/poly/src/lib.rs:156-185), so a narrow table crosses PCIe and sits
 * in HBM at its own width and no lifted copy is made. APG_ERR_ARG for an
 * unknown backing. */
int apg_sumcheck_prove_dev_lifted(apg_ctx *ctx, const apg_term *terms, size_t n_terms, void *const *tables_dev,
                                  const int *backings, size_t k, size_t num_vars, const uint8_t *eq_point_mont,
                                  apg_round_fn round_fn, void *user, uint8_t *rounds_out, uint8_t *point_out,
                                  uint8_t *final_evals_out);

/* The product summand eq(r, x) t_0(x) ... t_(k-1)(x), 1 <= k <= 3, over device
 * tables of 2^num_vars entries (read, not modified), with every round and its
 * transcript step (SUMCHECK_ROUND, SUMCHECK_CHALLENGE) on the device: the
 * rounds are captured as one HIP graph and launched once, with no host step
 * between rounds. The transcript is the library's own (its state moves to the
 * device and back); rounds (num_vars x max(4, k + 2) coefficients), point and
 * final evaluations equal apg_sumcheck_prove_dev's for the gate with one
 * term of factors 0..k-1 driven by apg_transcript_round on the same
 * transcript. Replaces the round loop of gkr::prove_sumcheck
 * (crates/gkr/src/lib.rs:792-899) for callers that cannot run a host step per
 * round; a device permutation runs ~70x slower than a host one, so the host-
 * driven entry is the faster one for a single proof. */
int apg_sumcheck_prove_graph_dev(apg_ctx *ctx, void *const *tables_dev, size_t k, size_t num_vars,
                                 const uint8_t *eq_point_mont, apg_transcript *t, uint8_t *rounds_out,
                                 uint8_t *point_out, uint8_t *final_evals_out);

/* apg_sumcheck_prove_dev with no device allocation in the call. No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Its scratch is one 64 MiB array per device, held from
 * the first context on that device until the process ends; it needs
 * 24 B x k x 2^num_vars (the two fold buffers) plus about 0.5 MiB and the
 * suffix-eq tables, so k x 2^num_vars up to 2^21 entries fit, and a
 * larger call is APG_ERR_ARG with the sizes in apg_last_error(). Calls on one
 * device share that array and are serialised by a lock; each starts from
 * what the last call left, which needs no clearing: every round writes the
 * partial sums and node sums it reads before reading them. Results equal
 * apg_sumcheck_prove_dev's. */
int apg_sumcheck_prove_dev_noalloc(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                                   void *const *tables_dev, size_t k, size_t num_vars,
                                   const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                                   uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out);

/* sumcheck::prove_zerocheck (crates/sumcheck/src/lib.rs:359): n
 * SUMCHECK_CHALLENGE draws for r, the rounds, then final_evals absorbed under
 * SUMCHECK_FINAL_EVALS, all on the given transcript. */
int apg_zerocheck_prove(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                        const uint8_t *const *tables_mont, size_t k, size_t num_vars,
                        apg_transcript *t, uint8_t *rounds_out, uint8_t *final_evals_out);

/* sumcheck::witness_digest (crates/sumcheck/src/lib.rs:228-256): a Poseidon2
 * sponge: append_scalars(WITNESS_DIGEST, [k, num_vars]), then every column's
 * cells as one append_scalars(WITNESS_DIGEST, ...) each, then sample. A
 * strictly sequential chain of permutations, so it runs on the host CPU
 * (the _dev form copies device columns down first; k * 2^num_vars elements,
 * column-major). absorb_witness_digest (:262) is
 * apg_transcript_append_scalars(t, WITNESS_DIGEST, digest, 1). k == 0 is
 * APG_ERR_ARG (the reference panics). */
int apg_witness_digest(const uint8_t *const *columns_mont, size_t k, size_t num_vars, uint8_t out_mont[32]);
int apg_witness_digest_dev(apg_ctx *ctx, const void *columns_dev, size_t k, size_t num_vars,
                           uint8_t out_mont[32]);

/* ---- one shard: commit k columns, then the batched sumcheck.
 * The build-defined composition of reference primitives used by the
 * benchmarks: commitments absorbed as one COMMITMENT message (pcs
 * append_g1_list, crates/pcs/src/lib.rs:220); rho = MERCURY_BATCH challenge;
 * f = degree - 1 combined tables t_o = sum_c rho^(o k + c) col_c; eq point =
 * num_vars SUMCHECK_CHALLENGE draws; rounds of eq * t_0 * ... * t_(f-1)
 * (SUMCHECK_ROUND / SUMCHECK_CHALLENGE per round); final evals absorbed under
 * SUMCHECK_FINAL_EVALS. columns_dev: k * 2^num_vars elements, device.
 * degree is 3 or 4 (one gate term of at most three factors); anything else,
 * num_vars outside 1..30 or k outside 1..24 is APG_ERR_ARG. */
int apg_prove_shard_dev(apg_ctx *ctx, apg_srs *srs, const void *columns_dev, size_t num_vars,
                        size_t k, int degree, apg_transcript *t, uint8_t *commitments_wire,
                        uint8_t *rounds_out, uint8_t *final_evals_out);

/* ---- the shard engine: continuous batching of shard jobs on one device.
 * Mirrors the streaming prover's worker model (max_in_flight scoped workers
 * claiming shards, crates/prover/src/streaming.rs:601-680): `workers` HIP
 * streams, each with its own scratch; jobs wait in one queue (arrival order
 * unless apg_engine_set_policy says otherwise) and start when a worker is
 * free and their HBM charge fits hbm_budget_bytes: each worker is charged
 * the device memory its scratch arena holds (live and outgrown buffers), or
 * while it runs a job the larger of that and the job's estimate; with nothing
 * running the next job is admitted whatever the arenas hold.
 * Each job runs apg_prove_shard_dev on its own copy of the
 * seed transcript; its bytes do not depend on what else runs. Columns stay
 * caller-owned device memory until the job has been waited for. */
typedef struct apg_engine apg_engine;
typedef struct {
    double queue_ms; /* submit -> start */
    double run_ms;   /* start -> end */
    double total_ms; /* submit -> end */
    int worker;
} apg_job_stats;
typedef struct {
    size_t queued, running, completed;
    size_t hbm_in_use, hbm_peak, hbm_budget; /* the workers' charges (see above) */
    size_t arena_bytes;                      /* device scratch the workers hold */
    size_t workers;
} apg_engine_info;
int apg_engine_create(int device, int workers, size_t hbm_budget_bytes, apg_engine **out);
void apg_engine_destroy(apg_engine *e);
int apg_engine_submit_shard(apg_engine *e, apg_srs *srs, const void *columns_dev, size_t num_vars,
                            size_t k, int degree, const apg_transcript *seed, uint64_t *job_id);
/* 1 done, 0 pending, APG_ERR_ARG for an id never issued or already waited for */
int apg_engine_poll(apg_engine *e, uint64_t job_id);
/* blocks until the job is done; outputs as apg_prove_shard_dev, plus the final
 * transcript (nullable) and the job's timing (nullable); APG_ERR_ARG at once
 * for an id never issued or already waited for */
int apg_engine_wait(apg_engine *e, uint64_t job_id, uint8_t *commitments_wire, uint8_t *rounds_out,
                    uint8_t *final_evals_out, apg_transcript *transcript_out, apg_job_stats *stats);
int apg_engine_stats(apg_engine *e, apg_engine_info *out);
/* Queue policy: APG_POLICY_FIFO (arrival order, the default),
 * APG_POLICY_SHORTEST (smallest rows x columns first among jobs whose HBM
 * estimate fits) or APG_POLICY_LONGEST (largest first among those: a batch
 * submitted at once starts its long shards early). With reserve > 0 the first `reserve` workers are a fast
 * lane that takes only jobs of at most small_scalars rows x columns; the
 * other workers leave such jobs to an idle fast-lane worker. With
 * small_scalars > 0, jobs above it run their MSM bucket accumulation at a
 * lower wave priority than every other kernel of the shard path, so a short
 * job's kernels issue first beside a long one. No counterpart
 * in the reference, whose workers claim shards in emission order
 * (crates/prover/src/streaming.rs:601-680). */
enum { APG_POLICY_FIFO = 0, APG_POLICY_SHORTEST = 1, APG_POLICY_LONGEST = 2 };
int apg_engine_set_policy(apg_engine *e, int policy, int reserve, double small_scalars);
/* An engine whose first lane_workers workers are a fast lane for jobs of at
 * most small_scalars rows x columns (shortest-first policy), with their
 * streams on CUs [0, lane_cus) and every other worker's streams on the
 * remaining CUs (lane_cus = 0: no CU split), so a short shard's kernels never
 * wait for execution slots behind a long shard's. */
int apg_engine_create_lanes(int device, int workers, size_t hbm_budget_bytes, int lane_workers, int lane_cus,
                            double small_scalars, apg_engine **out);
/* Preemption at phase boundaries. With ratio >= 1, a job running on a worker
 * outside the fast lane pauses at its next phase boundary (between the MSM's
 * sort passes, accumulation chunks and bucket phases, after the column
 * combination, between sumcheck rounds) while another worker (fast-lane
 * workers count when the lane has no CUs of its own) runs a job it gives way
 * to, and resumes when none does:
 *   APG_PREEMPT_SIZE:    a job whose rows x columns times `ratio` is below its own;
 *   APG_PREEMPT_URGENCY: a job whose urgency exceeds its own times `ratio`,
 *     urgency = (time since submission + estimated remaining time) /
 *     estimated time alone, so a job falling behind its isolated time gains
 *     priority (the estimates: a linear fit of shard times on MI355X, and the
 *     fraction of its phases a job has done).
 * The job's host thread drains its stream at each boundary, so the job given
 * way to has the CUs to itself once the phase in flight ends. ratio 0 (the
 * default) turns it off. Results do not change: a job's bytes depend only on
 * its inputs. *pauses (nullable) receives the pauses so far. No counterpart
 * in the reference, whose shard workers run to completion
 * (crates/prover/src/streaming.rs:601-680). */
enum { APG_PREEMPT_SIZE = 0, APG_PREEMPT_URGENCY = 1 };
int apg_engine_set_preemption(apg_engine *e, double ratio, int mode, uint64_t *pauses);
/* kernel timing (apg_stats) on every worker's stream, summed as jobs finish */
int apg_engine_set_profiling(apg_engine *e, int on);
int apg_engine_kernel_stats(apg_engine *e, apg_stats *out);
int apg_engine_reset_kernel_stats(apg_engine *e);

/* ---- the Poseidon2 transcript (crates/transcript/src/lib.rs) ---- */
int apg_transcript_new(apg_transcript **out);
void apg_transcript_free(apg_transcript *t);
int apg_transcript_clone(const apg_transcript *t, apg_transcript **out);
void apg_transcript_observe(apg_transcript *t, const uint8_t x_mont[32]);
void apg_transcript_sample(apg_transcript *t, uint8_t out_mont[32]);
void apg_transcript_append_scalars(apg_transcript *t, uint64_t tag, const uint8_t *xs_mont,
                                   size_t n);
void apg_transcript_append_bytes(apg_transcript *t, uint64_t tag, const uint8_t *bytes, size_t len);
void apg_transcript_challenge(apg_transcript *t, uint64_t tag, uint8_t out_mont[32]);
void apg_transcript_append_g1(apg_transcript *t, uint64_t tag, const uint8_t *wire64s, size_t k);
/* an apg_round_fn: append_scalars(SUMCHECK_ROUND, coeffs), then
 * challenge_scalar(SUMCHECK_CHALLENGE) (crates/sumcheck/src/lib.rs:392-395) */
int apg_transcript_round(void *transcript, const uint8_t *coeffs_mont, size_t n_coeffs,
                         uint8_t challenge_mont[32]);
/* sponge state: state[3], input[2], output[2] (Montgomery, 32 B each), then
 * input_len and output_len as u32: 232 bytes */
int apg_transcript_state(const apg_transcript *t, uint8_t out[232]);
/* ---- the transcript on the device (one transcript per thread) ----
 * apg_poseidon2_permute_dev permutes n states of three 32-byte Montgomery Fr
 * elements in place (host buffer, canonical elements): the reference's
 * poseidon2_permute (crates/transcript/src/lib.rs:121-146) for many states at
 * once. apg_transcript_run_dev runs, for each i < count, the program
 * ops[offsets[i] .. offsets[i + 1]) on transcript ts[i] on the device and
 * leaves each transcript in its new state. An op is four uint32 words:
 *   {0, a, 0, 0}   observe operands[a]               (Transcript::observe, lib.rs:241-250)
 *   {1, lo, hi, 0} observe the element lo + 2^32 hi  (tags and lengths of the typed layer)
 *   {2, o, 0, 0}   outputs[o] = sample()              (Transcript::sample, lib.rs:253-262)
 * so append_scalars(tag, xs) is {1, tag}, {1, n}, then {0, .} per scalar, and
 * challenge(tag) is {1, tag}, {2, .} (lib.rs:306-346). Operand or output
 * indices out of range, unknown ops, decreasing offsets or non-canonical
 * operands are APG_ERR_ARG. Results equal the host transcript's. */
int apg_poseidon2_permute_dev(apg_ctx *ctx, uint8_t *states_mont, size_t n);
int apg_transcript_run_dev(apg_ctx *ctx, apg_transcript *const *ts, size_t count, const uint32_t *ops, size_t n_ops,
                           const uint32_t *offsets, const uint8_t *operands_mont, size_t n_operands,
                           uint8_t *outputs_mont, size_t n_outputs);
/* The inverse of apg_transcript_state: sets the sponge state, both buffers
 * and their fill counts from the same 232-byte layout, so a caller can hand
 * its own transcript's state across the boundary and take it back
 * (crates/transcript/src/lib.rs:197-346 keeps these fields). APG_ERR_ARG for
 * a non-canonical element or a fill count above 2. */
int apg_transcript_set_state(apg_transcript *t, const uint8_t in[232]);

/* ---- encodings ---- */
int apg_fr_to_bytes(const uint8_t mont[32], uint8_t canonical[32]);
int apg_fr_from_bytes(const uint8_t canonical[32], uint8_t mont[32]); /* APG_ERR_ARG if >= p */
int apg_fq_to_bytes(const uint8_t mont[32], uint8_t canonical[32]);
int apg_fq_from_bytes(const uint8_t canonical[32], uint8_t mont[32]);
void apg_fr_from_u64(uint64_t x, uint8_t mont[32]);
int apg_g1_to_wire(const uint8_t mont[64], uint8_t wire[64]);
int apg_g1_from_wire(const uint8_t wire[64], uint8_t mont[64]); /* checks canonicity + curve */

#ifdef __cplusplus
}
#endif
#endif

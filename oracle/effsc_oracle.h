/*
 * effsc_oracle.h — CPU oracle for the Goldilocks MSB sumcheck path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * crate `effsc` (compsec-epfl/efficient-sumcheck, Rust) used as the checker
 * for the HIP library and as the timed CPU baseline.  Only tests/, the smoke
 * entry and the benchmark's cpu_baseline leg may load it; the product library
 * never links or calls it.
 *
 * Data at this interface: every field element is `ext` canonical u64 words
 * (AoS, value in [0, p)), p = 2^64 - 2^32 + 1.  Internally the oracle keeps
 * arkworks-style Montgomery words (R = 2^64), as the reference's F64 does
 * (src/tests/fields.rs:26-30, src/field.rs:99-116).
 */
#ifndef EFFSC_ORACLE_H
#define EFFSC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field (canonical in/out) ---- */
uint64_t or_gl_add(uint64_t a, uint64_t b);
uint64_t or_gl_sub(uint64_t a, uint64_t b);
uint64_t or_gl_mul(uint64_t a, uint64_t b);       /* via Montgomery CIOS */
uint64_t or_gl_mul_i128(uint64_t a, uint64_t b);  /* via 128-bit remainder */
uint64_t or_gl_inv(uint64_t a);
uint64_t or_mont_mul(uint64_t a, uint64_t b);     /* raw Montgomery words */
uint64_t or_to_mont(uint64_t a);
uint64_t or_from_mont(uint64_t a);
/* out = a*b in F_p[X]/(X^d - w), all canonical, d in {1,2,3} */
void or_ext_mul(int d, uint64_t w, const uint64_t *a, const uint64_t *b, uint64_t *out);
void or_ext_inv(int d, uint64_t w, const uint64_t *a, uint64_t *out);

/* ---- synthetic inputs: counter-based uniform values in [0, p) ---- */
uint64_t or_uniform(uint64_t seed, uint64_t index);
void or_fill_uniform(uint64_t *out, size_t n, uint64_t seed);

/* ---- Keccak / SHAKE256 ---- */
void or_keccak_f1600(uint64_t st[25]);
void or_shake256(const uint8_t *msg, size_t len, uint8_t *out, size_t outlen);

/* ---- transcripts ---- */
typedef struct or_transcript or_transcript;
/* Fiat-Shamir transcript: SHAKE256 over a stream of little-endian u64 words
 *   "effsc-hip/fs/v3\0" (2 words) || len || domsep zero-padded to 8 bytes
 *   send(x):      word 1, then the canonical components
 *   challenge(d): word 2 is appended; output = d x (u128le mod p) from the
 *                 first 16 d bytes of SHAKE256(stream); then the stream is
 *                 padded to the next 136-byte boundary with the SHAKE
 *                 padding (0x1F, zeros, 0x80), so a challenge costs one
 *                 permutation. */
or_transcript *or_tr_new_shake(const uint8_t *domsep, size_t len);
/* Test transcript: challenge = from_u64(splitmix64 next) lifted into ext. */
or_transcript *or_tr_new_rng(uint64_t seed);
void or_tr_send(or_transcript *t, int d, const uint64_t *value);
void or_tr_challenge(or_transcript *t, int d, uint64_t *out);
void or_tr_free(or_transcript *t);

/* ---- provers ---- */
enum { OR_ML = 1, OR_IP = 2, OR_PRODUCT = 3, OR_EQ = 4, OR_GKR = 5 };
typedef struct or_prover or_prover;
/* ext: extension degree 1..3, w: canonical non-residue (ignored for ext=1) */
or_prover *or_prover_new_ml(int ext, uint64_t w, const uint64_t *evals, size_t n);
or_prover *or_prover_new_ip(int ext, uint64_t w, const uint64_t *a, const uint64_t *b, size_t n);
/* product of k pairwise tables, each given as base-field words and lifted
 * into the extension (as the reference's CoefficientProver<F64Ext3> runs on
 * lifted inputs); tables: k pointers to n words each */
or_prover *or_prover_new_product(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n);
/* Same, but each table already holds n extension elements (ext words each). */
or_prover *or_prover_new_product_ext(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n);
or_prover *or_prover_new_eq(int ext, uint64_t w, const uint64_t *point, size_t v, const uint64_t *p, size_t n);
/* LSB (pair-split) provers: pairs (2k, 2k+1) */
or_prover *or_prover_new_ml_lsb(int ext, uint64_t w, const uint64_t *evals, size_t n);
or_prover *or_prover_new_ip_lsb(int ext, uint64_t w, const uint64_t *a, const uint64_t *b, size_t n);
or_prover *or_prover_new_product_lsb(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n);
/* Same, each table already holding n extension elements (ext words each). */
or_prover *or_prover_new_product_lsb_ext(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n);
/* GkrProver: add and mult hold n*n elements over (b, c), wit holds n
 * elements W(x); tables 0..3 are add, mult, w_b, w_c (ext words each). */
or_prover *or_prover_new_gkr(int ext, uint64_t w, const uint64_t *add, const uint64_t *mult, const uint64_t *wit,
                             size_t n);
int or_prover_degree(const or_prover *p);
void or_prover_round(or_prover *p, const uint64_t *chal_or_null, uint64_t *out);
void or_prover_finalize(or_prover *p, const uint64_t *chal);
void or_prover_final_value(const or_prover *p, uint64_t *out);
size_t or_prover_table_len(const or_prover *p, int table);
void or_prover_read_table(const or_prover *p, int table, uint64_t *out);
void or_prover_free(or_prover *p);

/* ---- runner / verifier ---- */
/* round_polys: num_rounds*degree*ext words; challenges: num_rounds*ext; fin: ext */
void or_sumcheck(or_prover *p, size_t num_rounds, or_transcript *t,
                 uint64_t *round_polys, uint64_t *challenges, uint64_t *fin);
/* Returns 0 on success, or 1 + round for a transcript error (proof too short).
 * challenges: num_rounds*ext; final_claim: ext */
int or_sumcheck_verify(int ext, uint64_t w, const uint64_t *claimed_sum, size_t degree,
                       size_t num_rounds, or_transcript *t, const uint64_t *proof_words,
                       size_t proof_len_words, uint64_t *challenges, uint64_t *final_claim);

/* ---- CPU baseline (reference InnerProductProver::round structure) ---- */
/* Runs a full inner-product proof on base-field tables with the SHAKE
 * transcript; threads > 1 parallelises the two folds as the reference's rayon
 * `recurse_both` does and keeps the evaluate loop serial, as
 * src/provers/inner_product.rs:94-103 does. Returns wall seconds. */
double or_bench_ip(uint64_t *a, uint64_t *b, size_t n, size_t num_rounds, int threads,
                   uint64_t *round_polys, uint64_t *challenges, uint64_t *fin);
/* The same proof with the transcript's domain separator given, and with
 * fused = 1 the reference's faster legacy path instead
 * (inner_product_sumcheck_partial + fused_fold_and_compute_polynomial,
 * src/inner_product_sumcheck.rs:199-360: fold and evaluate in one parallel
 * sweep per round).  Both produce the InnerProductProver wire. */
double or_bench_ip_ex(uint64_t *a, uint64_t *b, size_t n, size_t num_rounds, int threads, int fused,
                      const uint8_t *domsep, size_t domsep_len, uint64_t *round_polys, uint64_t *challenges,
                      uint64_t *fin);
/* MultilinearProver with the reference's parallel fused kernels (base field,
 * [s0] per round); v is consumed.  Returns the seconds of the proof. */
double or_bench_ml_ex(uint64_t *v, size_t n, size_t num_rounds, int threads, const uint8_t *domsep,
                      size_t domsep_len, uint64_t *round_polys, uint64_t *challenges, uint64_t *fin);

#ifdef __cplusplus
}
#endif
#endif

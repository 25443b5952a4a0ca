/*
 * effsc_hip.h — C ABI of the MI355X (gfx950) Goldilocks sumcheck prover.
 *
 * Drop-in boundary for the reference crate `effsc` 0.0.2
 * (compsec-epfl/efficient-sumcheck).  The reference's boundary is the Rust
 * trait `SumcheckProver<F>` (src/sumcheck_prover.rs:36-71) consumed by
 * `runner::sumcheck` (src/runner.rs:29-69); a foreign binding implements that
 * trait by forwarding each method to the entry point named beside it below.
 *
 * Conventions
 *  - Field elements are u64 words: `ext` words per element (ext = 1 for
 *    F_p, 2 for F_p[X]/(X^2 - w), 3 for F_p[X]/(X^3 - w): the reference's
 *    F64, F64Ext2 and F64Ext3, src/tests/fields.rs:26-100), components in
 *    order c0, c1, c2 (the arkworks layout the reference reinterprets
 *    through _as_u64_slice, src/field.rs:118-145).
 *  - `repr` selects the word encoding at this boundary: canonical values in
 *    [0, p), or arkworks Montgomery words (R = 2^64), which is what the
 *    reference's F64 stores (src/tests/fields.rs:26-30).  Device tables are
 *    always canonical.
 *  - Every function returns an effsc_hip_status; no C++ exception crosses the
 *    boundary.  On failure effsc_hip_last_error() describes it (thread-local).
 *    EINVAL is where the reference panics (e.g. src/provers/inner_product.rs:34,
 *    src/provers/eq_factored.rs:93-99).
 *  - A prover handle is not thread-safe; distinct handles may be driven from
 *    distinct threads.  Each handle owns one HIP stream on its device.
 *  - Input tables are copied to the device at construction (host or device
 *    source pointers are both accepted) unless EFFSC_HIP_BORROW_INPUT is set.
 *    Folding never writes the input tables: the first fold writes a separate
 *    half-size work table, so effsc_hip_reset() restarts a proof for free.
 */
#ifndef EFFSC_HIP_H
#define EFFSC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EFFSC_HIP_OK = 0,
  EFFSC_HIP_EINVAL = 1,   /* bad argument (the reference panics here) */
  EFFSC_HIP_ENOMEM = 2,   /* device or pinned host allocation failed */
  EFFSC_HIP_EDEVICE = 3,  /* HIP runtime error */
  EFFSC_HIP_ESTATE = 4,   /* call not valid in the prover's current state */
  EFFSC_HIP_EVERIFY = 5   /* verifier: transcript error (proof too short) or hook error */
} effsc_hip_status;

typedef enum { EFFSC_HIP_CANONICAL = 0, EFFSC_HIP_MONT64 = 1 } effsc_hip_repr;

/* Input flag: `tables` are device pointers to canonical words in the device
 * layout (base: one plane; ext2 / ext3: two / three planes of `n` words each), owned by the
 * caller and kept alive and unmodified until effsc_hip_free. */
#define EFFSC_HIP_BORROW_INPUT 0x1u

/* The reference's extension non-residue F64::from_raw(7), canonical value
 * 7 * 2^-64 mod p, for both F64Ext2 and F64Ext3 (src/tests/fields.rs:49,80). */
#define EFFSC_HIP_W_FROM_RAW7 0xFFFFFFF800000001ull

typedef struct {
  int ext;          /* field of tables, challenges and round polynomials: 1, 2 or 3 */
  int input_ext;    /* field of the input words: 1 (lifted when ext > 1) or ext */
  int repr;         /* effsc_hip_repr of every word crossing this ABI */
  uint64_t ext_w;   /* canonical non-residue w of X^ext - w (ignored when ext = 1) */
  int device;       /* HIP device ordinal */
  unsigned flags;   /* EFFSC_HIP_BORROW_INPUT */
} effsc_hip_config;

typedef struct effsc_hip_prover effsc_hip_prover;
typedef struct effsc_hip_transcript effsc_hip_transcript;

/* Fills `cfg` with ext = input_ext = 1, canonical words, w = from_raw(7),
 * device 0, no flags. */
void effsc_hip_config_default(effsc_hip_config* cfg);

/* ---------------------------------------------------------------- provers */

/* MultilinearProver::new(evals) — src/provers/multilinear.rs:31-35.
 * Degree 1; any n >= 0 (non-powers of two are implicitly zero-padded). */
int effsc_hip_multilinear_new(const effsc_hip_config* cfg, const uint64_t* evals, size_t n,
                              effsc_hip_prover** out);

/* InnerProductProver::new(a, b) — src/provers/inner_product.rs:33-36.
 * Degree 2; a and b have equal length n. */
int effsc_hip_inner_product_new(const effsc_hip_config* cfg, const uint64_t* a, const uint64_t* b, size_t n,
                                effsc_hip_prover** out);

/* CoefficientProver::new(&evaluator, vec![], pairwise) with the built-in
 * product-of-k evaluator, sum_x prod_i f_i(x) (degree k, 1 <= k <= 4) —
 * src/provers/coefficient.rs:28-40; the k = 3 evaluator is the reference's
 * TripleProductEval (tests/adversarial_verifier.rs:444-465).  All k tables
 * have length n, any n: each round pairs k with k + n/2 and the fold keeps
 * n/2 elements, dropping an odd last one (msb_fold_vec,
 * src/provers/coefficient.rs:131-140).  Tablewise tables belong to
 * user-defined evaluators, which do not cross this ABI. */
int effsc_hip_product_new(const effsc_hip_config* cfg, const uint64_t* const* tables, int k, size_t n,
                          effsc_hip_prover** out);

/* LSB (pair-split) provers: pairs (2k, 2k+1), the least significant variable
 * bound first.  MultilinearProverLSB::new (src/provers/multilinear_lsb.rs:33-36),
 * InnerProductProverLSB::new (src/provers/inner_product_lsb.rs),
 * CoefficientProverLSB::new with the product-of-k evaluator
 * (src/provers/coefficient_lsb.rs).  For power-of-two tables the prover
 * bit-reverses them once into its own device planes (so
 * EFFSC_HIP_BORROW_INPUT sources are read, never kept) and runs the MSB path;
 * post-state reads return the reference's pair-split order.  The multilinear
 * prover also takes any other length, as the reference does
 * (src/provers/multilinear_lsb.rs:59-126): the round value sums the
 * even-indexed elements, the fold drops an unpaired last element; the
 * inner-product and product provers reject such lengths (EINVAL), where the
 * reference panics. */
int effsc_hip_multilinear_lsb_new(const effsc_hip_config* cfg, const uint64_t* evals, size_t n,
                                  effsc_hip_prover** out);
int effsc_hip_inner_product_lsb_new(const effsc_hip_config* cfg, const uint64_t* a, const uint64_t* b, size_t n,
                                    effsc_hip_prover** out);
int effsc_hip_product_lsb_new(const effsc_hip_config* cfg, const uint64_t* const* tables, int k, size_t n,
                              effsc_hip_prover** out);
/* reorder_vec_msb (src/streams/memory/core.rs:20-24) on the device:
 * dst[i] = src[bit_reverse(i)] for `planes` consecutive planes of n = 2^v
 * words each (out of place; device pointers). */
int effsc_hip_bit_reverse(int device, uint64_t* dst, const uint64_t* src, size_t n, int planes);

/* GkrProver::new(add_evals, mult_evals, w_evals) — src/provers/gkr.rs:58-87.
 * Degree 2 over 2k variables: add and mult hold n * n elements over (b, c)
 * (index b n + c), wit holds the n witness values W(x).  Tables 0..3 of the
 * post-state are add, mult, w_b and w_c (the witness broadcast over c and
 * over b, as in the reference). */
int effsc_hip_gkr_new(const effsc_hip_config* cfg, const uint64_t* add, const uint64_t* mult, const uint64_t* wit,
                      size_t n, effsc_hip_prover** out);
/* claimed_w_values() — src/provers/gkr.rs:89-97: (W(b*), W(c*)), 2 * ext
 * words, once every variable is bound (zeros before). */
int effsc_hip_gkr_claimed_w(effsc_hip_prover* h, uint64_t* out);

/* EqFactoredProver::new(w, p) — src/provers/eq_factored.rs:90-114.
 * Degree 2; `point` holds v elements (host memory), p holds n <= 2^v elements
 * (zero-padded).  With EFFSC_HIP_BORROW_INPUT, p is a canonical device table
 * of exactly 2^v elements that the prover never writes. */
int effsc_hip_eq_factored_new(const effsc_hip_config* cfg, const uint64_t* point, size_t v, const uint64_t* p,
                              size_t n, effsc_hip_prover** out);

/* SumcheckProver::degree — src/sumcheck_prover.rs:37-38. */
int effsc_hip_degree(const effsc_hip_prover* h, size_t* out);

/* Extension degree of the prover's field (words per element). */
int effsc_hip_ext(const effsc_hip_prover* h, size_t* out);

/* SumcheckProver::round — src/sumcheck_prover.rs:40-59.  `challenge` is NULL
 * for Option::None, else `ext` words; `out` receives max(degree, 1) * ext
 * words in the EvalsInfty wire format. */
int effsc_hip_round(effsc_hip_prover* h, const uint64_t* challenge, uint64_t* out);

/* SumcheckProver::finalize — src/sumcheck_prover.rs:61-66. */
int effsc_hip_finalize(effsc_hip_prover* h, const uint64_t* challenge);

/* SumcheckProver::final_value — src/sumcheck_prover.rs:68-70 (`ext` words). */
int effsc_hip_final_value(effsc_hip_prover* h, uint64_t* out);

/* Post-state accessors: evals() (src/provers/multilinear.rs:45-47),
 * evaluations() (src/provers/inner_product.rs:39-41), pairwise()
 * (src/provers/coefficient.rs:46-48).  EqFactored exposes p, eq_l, eq_r as
 * tables 0, 1, 2 (final_factors, src/provers/eq_factored.rs:117-123). */
int effsc_hip_table_count(const effsc_hip_prover* h, size_t* out);
int effsc_hip_table_len(const effsc_hip_prover* h, size_t table, size_t* out);
/* Copies table `table` (len * ext words, AoS, in cfg.repr) into dst; fails
 * with EINVAL if cap_words is too small. */
int effsc_hip_read_table(effsc_hip_prover* h, size_t table, uint64_t* dst, size_t cap_words, size_t* len);

/* Restarts the prover from its (unmodified) input tables; an EqFactored
 * prover also rebuilds its eq tables from the point.  No reference
 * counterpart: the reference consumes its tables. */
int effsc_hip_reset(effsc_hip_prover* h);

/* Drop. */
void effsc_hip_free(effsc_hip_prover* h);

/* Kernel timing, no reference counterpart.  effsc_hip_profile(h, 1) clears
 * the counters and brackets every later multi-workgroup round launch with HIP
 * events on the prover's stream; effsc_hip_profile_read returns, per kernel
 * class (0 fused fold+evaluate, 1 evaluate, 2 fold, 3 two-round pass; arrays
 * of 4), the launch count, the summed event time in ms and the summed
 * algorithmic bytes (each input element read once, each output element
 * written once); effsc_hip_profile_read_muls the summed Goldilocks products
 * (an F_p^3 product counts 8: Karatsuba's 6 plus 2 by w). */
int effsc_hip_profile(effsc_hip_prover* h, int enable);
int effsc_hip_profile_read(const effsc_hip_prover* h, uint64_t* launches, double* ms, double* bytes);
int effsc_hip_profile_read_muls(const effsc_hip_prover* h, double* muls);

/* hipDeviceSynchronize on `device`. */
int effsc_hip_synchronize(int device);

/* Thread-local description of the last failure ("" if none). */
const char* effsc_hip_last_error(void);

/* ------------------------------------------------------------ free fold */

/* fold(&mut Vec<F>, weight) — src/fold.rs:14, src/multilinear_sumcheck.rs:132-178:
 * in-place half-split fold of n elements (device pointer, canonical, base
 * field) with implicit zero padding; *new_len = next_pow2(n) / 2 (or n if
 * n <= 1). */
int effsc_hip_fold(int device, uint64_t* values, size_t n, uint64_t weight, size_t* new_len);
/* The same fold over an extension field (fold<F64Ext2> / fold<F64Ext3>,
 * src/fold.rs:14; the reference benchmarks fold/F64Ext3,
 * benches/sumcheck.rs:156-180): n elements of `ext` canonical words each, AoS
 * (the layout of the reference's Vec<F> reinterpreted as u64 words), in
 * place on the device; `weight` is ext host words, w the non-residue. */
int effsc_hip_fold_ext(int device, int ext, uint64_t w, uint64_t* values, size_t n, const uint64_t* weight,
                       size_t* new_len);

/* --------------------------------------------------------- transcripts */

/* Fiat-Shamir transcript (SHAKE256 duplex over a fixed byte encoding; the
 * reference binds spongefish here, src/transcript/spongefish.rs:17-44). */
effsc_hip_transcript* effsc_hip_transcript_new(const uint8_t* domain_separator, size_t len);
/* TestTranscript (src/transcript/sanity.rs:30-61): sends are no-ops and each
 * challenge is from_u64(splitmix64 next) lifted into the extension. */
effsc_hip_transcript* effsc_hip_transcript_new_test(uint64_t seed);
/* ProverTranscript::send / VerifierTranscript::receive absorption
 * (src/transcript/transcript.rs:11-43); `value` is ext canonical words. */
int effsc_hip_transcript_send(effsc_hip_transcript* t, int ext, const uint64_t* value);
int effsc_hip_transcript_challenge(effsc_hip_transcript* t, int ext, uint64_t* out);
void effsc_hip_transcript_free(effsc_hip_transcript* t);

/* ------------------------------------------------- runner and verifier */

/* Per-round hook: called after the round's values are sent and before the
 * challenge is read; a non-zero return from a verifier hook is
 * SumcheckError::HookError (src/verifier.rs:86-87).  The prover runner
 * ignores the return value, as the reference's FnMut hook returns nothing. */
typedef int (*effsc_hip_hook)(void* ctx, size_t round, effsc_hip_transcript* t);

/* runner::sumcheck — src/runner.rs:29-69.  round_polys: num_rounds *
 * max(degree,1) * ext words; challenges: num_rounds * ext; final_value: ext.
 * Words are canonical.  hook may be NULL (noop_hook, src/lib.rs:28-30). */
int effsc_hip_sumcheck(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, effsc_hip_hook hook,
                       void* hook_ctx, uint64_t* round_polys, uint64_t* challenges, uint64_t* final_value);

/* A whole MultilinearProver (kind 1) or InnerProductProver (kind 2) proof that
 * allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It is effsc_hip_sumcheck over all
 * v = log2(n) rounds with no hook, on borrowed canonical device tables a and b
 * (b ignored for kind 1; 16-byte aligned, n a power of two, 2 <= n <= 2^20),
 * run on the device's null stream.  Every buffer beyond the inputs and outputs
 * lives in one static device array of the library (about 9 MiB, held on each
 * device it runs on until the process ends); larger n is EINVAL.  Outputs as
 * effsc_hip_sumcheck (host words, canonical); the transcript state is read
 * and updated.  Calls that overlap on one device share that array and race:
 * serialise them. */
int effsc_hip_sumcheck_noalloc(int device, int kind, const uint64_t* a, const uint64_t* b, size_t n,
                               effsc_hip_transcript* t, uint64_t* round_polys, uint64_t* challenges,
                               uint64_t* final_value);

/* verifier::sumcheck_verify — src/verifier.rs:61-139.  The prover messages
 * are read from `proof` (proof_words canonical words).  On success returns
 * OK with challenges (num_rounds * ext) and final_claim (ext); the oracle
 * check final_claim == g(r) stays with the caller, as in the reference.
 * Returns EVERIFY with *err_round set for SumcheckError::TranscriptError or
 * ::HookError (*err_kind = 1 transcript, 2 hook). */
int effsc_hip_sumcheck_verify(int ext, uint64_t w, const uint64_t* claimed_sum, size_t degree, size_t num_rounds,
                              effsc_hip_transcript* t, const uint64_t* proof, size_t proof_words,
                              effsc_hip_hook hook, void* hook_ctx, uint64_t* challenges, uint64_t* final_claim,
                              size_t* err_round, int* err_kind);

/* ------------------------------------------------------- batched proofs */

/* Many independent multilinear (kind 1) or inner-product (kind 2) instances
 * of uneven power-of-two size proved together: every round of every
 * instance runs in one launch, each instance's reduction and Fiat-Shamir step
 * in one workgroup of a second launch, and the round sequence is captured
 * once into a hipGraph (flags & 0x2 disables the graph).  Tables are
 * borrowed device pointers (canonical base-field words; b ignored for kind
 * 1), unmodified, alive until effsc_hip_batch_free.  No reference
 * counterpart: the reference proves one instance per runner::sumcheck call
 * (src/runner.rs:29-69); each instance's proof here equals that call's. */
typedef struct effsc_hip_batch effsc_hip_batch;
int effsc_hip_batch_new(int device, int kind, size_t count, const size_t* lens, const uint64_t* const* a,
                        const uint64_t* const* b, unsigned flags, effsc_hip_batch** out);
/* Rounds v_i = log2(len_i) of instance i and its offset in the round arrays. */
int effsc_hip_batch_rounds(const effsc_hip_batch* b, size_t instance, size_t* v, size_t* first_round);
/* Proves every instance with its own transcript (state read and updated).
 * round_polys: sum_i v_i * kind words (instance i at first_round * kind);
 * challenges: sum_i v_i words; finals: count words.  Canonical words. */
int effsc_hip_batch_prove(effsc_hip_batch* b, effsc_hip_transcript* const* transcripts, uint64_t* round_polys,
                          uint64_t* challenges, uint64_t* finals);
/* Tile autotuner: times `reps` whole batched proofs for each candidate
 * number of units (pairs in round 0, quadruples later) per workgroup, keeps
 * the fastest (its chunk tables and graph are rebuilt) and reports it in
 * `chosen`; ms_out (ncand doubles, may be NULL) receives each candidate's time
 * per proof.  Proofs are the same for every tile size. */
int effsc_hip_batch_autotune(effsc_hip_batch* b, const size_t* candidates, size_t ncand, int reps, size_t* chosen,
                             double* ms_out);
void effsc_hip_batch_free(effsc_hip_batch* b);

/* ------------------------------------------------------------ utilities */

/* ---------------------------------------------------------------------------
 * Streamed input.  FileStream (src/streams/file/core.rs:13-76): a file of
 * 2^v serialized elements, each component one little-endian canonical u64
 * (ark-serialize uncompressed), mapped read-only.
 * ------------------------------------------------------------------------- */
typedef struct effsc_hip_stream effsc_hip_stream;
/* FileStream::new (file/core.rs:29-43): EINVAL unless the file holds a
 * power-of-two number of ext-word elements. */
int effsc_hip_stream_open(const char* path, int ext, effsc_hip_stream** out);
int effsc_hip_stream_num_variables(const effsc_hip_stream* s, size_t* v);
/* FileStream::evaluation for `count` consecutive points (canonical words). */
int effsc_hip_stream_read(const effsc_hip_stream* s, size_t first, size_t count, uint64_t* dst);
void effsc_hip_stream_close(effsc_hip_stream* s);
/* multivariate_claim (k = 1) / multivariate_product_claim
 * (src/streams/stream.rs:3-41): sum_i prod_j s_j[i] on the device, the
 * streams moved in chunks through pinned buffers (ext words out; w is the
 * extension non-residue when ext > 1). */
int effsc_hip_stream_claim(int device, effsc_hip_stream* const* streams, int k, uint64_t w, uint64_t* out);
/* MultilinearProver (kind 1) / InnerProductProver (kind 2) over base-field
 * tables read from streams.  The upload runs chunk by chunk on a copy stream
 * while the first two-round pass over the landed chunks runs on the prover's
 * stream; effsc_hip_sumcheck then starts from that pass.  cfg.input_ext = 1. */
int effsc_hip_stream_prover_new(const effsc_hip_config* cfg, int kind, effsc_hip_stream* const* streams,
                                effsc_hip_prover** out);
/* The same provers for inputs larger than device memory (FileStream's
 * purpose, src/streams/file/core.rs:13-76; the reference reads such a stream
 * on the CPU).  mode EFFSC_HIP_STREAM_RESIDENT is effsc_hip_stream_prover_new;
 * EFFSC_HIP_STREAM_OUT_OF_CORE keeps the input in the streams: rounds 0-1 are
 * accumulated here while it passes through pinned and device staging chunk by
 * chunk, and the proof streams it once more for rounds 2-3, folding it on the
 * fly into level-n/4 tables, the only resident ones (K n/4 elements of device
 * memory instead of K (n + n/2)).  The streams must stay open for the
 * prover's lifetime.  Such a prover runs whole proofs only (effsc_hip_sumcheck
 * over all rounds, no hook; ESTATE otherwise until that proof has run) and
 * needs n/4 above the one-workgroup tail (n >= 2^17 by default).
 * EFFSC_HIP_STREAM_AUTO picks resident when the input fits with a margin. */
#define EFFSC_HIP_STREAM_RESIDENT 0
#define EFFSC_HIP_STREAM_OUT_OF_CORE 1
#define EFFSC_HIP_STREAM_AUTO 2
int effsc_hip_stream_prover_new_ex(const effsc_hip_config* cfg, int kind, effsc_hip_stream* const* streams, int mode,
                                   effsc_hip_prover** out);

/* ---------------------------------------------------------------------------
 * Served proofs: a request loop with round-level continuous batching.  No
 * reference counterpart: each request is one runner::sumcheck
 * (src/runner.rs:29-69) call with the no-op hook, and its proof is exactly
 * that call's for the request alone.
 *
 * A scheduler thread admits requests in priority order (lower first, then
 * submission order) while a slot and room in the server's HBM arena are free;
 * every admitted request advances one round per tick alongside all others,
 * with tables, transcript state and challenges resident on the device.  A
 * finished request frees its slot and arena bytes, so requests that arrive
 * while others run join the batch at the next tick.
 * ------------------------------------------------------------------------- */
typedef struct effsc_hip_server effsc_hip_server;

#define EFFSC_HIP_SERVE_MULTILINEAR 1   /* MultilinearProver, base field, degree 1 */
#define EFFSC_HIP_SERVE_INNER_PRODUCT 2 /* InnerProductProver, base field, degree 2 */
#define EFFSC_HIP_SERVE_PRODUCT3_EXT3 3 /* CoefficientProver product of 3, base inputs, F_p^3 challenges */

typedef struct {
  int kind;                         /* EFFSC_HIP_SERVE_* */
  unsigned v;                       /* variables: tables of 2^v elements */
  uint64_t seeds[3];                /* table i = uniform(seeds[i]) made on the device when tables[i] is NULL */
  const uint64_t* tables[3];        /* or caller-owned canonical device tables of 2^v words, 16-byte
                                       aligned (not counted in the footprint, kept alive until the
                                       request completes) */
  const effsc_hip_transcript* transcript; /* its state is copied at submit */
  double priority;                  /* admission order: lower first; ties in submission order */
} effsc_hip_request;

/* hbm_budget: bytes of the arena holding every admitted request's tables and
 * proof words; max_active: requests on the device at once (<= 4096). */
int effsc_hip_server_new(int device, size_t hbm_budget, size_t max_active, effsc_hip_server** out);
/* Arena bytes a request occupies while admitted. */
int effsc_hip_server_footprint(const effsc_hip_request* req, size_t* bytes);
/* Queues a request; EINVAL if its footprint exceeds the budget. */
int effsc_hip_server_submit(effsc_hip_server* s, const effsc_hip_request* req, uint64_t* id);
/* Blocks until request `id` completes and copies its proof: round_polys
 * (v * K * ext words, K = 1, 2, 3 values per round), challenges (v * ext),
 * final_value (ext); times (3 doubles, server clock in seconds): submitted,
 * admitted, completed.  Each id is waited for once. */
int effsc_hip_server_wait(effsc_hip_server* s, uint64_t id, uint64_t* round_polys, uint64_t* challenges,
                          uint64_t* final_value, double* times);
int effsc_hip_server_stats(effsc_hip_server* s, uint64_t* ticks, uint64_t* peak_active, uint64_t* peak_bytes,
                           uint64_t* completed);
/* The server clock (seconds since effsc_hip_server_new). */
int effsc_hip_server_now(effsc_hip_server* s, double* t);
/* Completes every submitted request, then stops the scheduler and frees. */
void effsc_hip_server_free(effsc_hip_server* s);

/* ---------------------------------------------------------------------------
 * Learned service-time model for serving admission.  No reference
 * counterpart; proofs never depend on it (it only orders admissions).
 * A fp32 network x (inputs) -> tanh (hidden) -> tanh (hidden) -> y, trained
 * on the device with its own forward / backward / Adam kernels on the mean
 * squared error.  Parameters are one flat array: W1 [hidden][inputs], b1,
 * W2 [hidden][hidden], b2, w3 [hidden], b3.  x is row-major [n][inputs].
 * ------------------------------------------------------------------------- */
typedef struct effsc_hip_mlp effsc_hip_mlp;
int effsc_hip_mlp_new(int device, int inputs, int hidden, uint64_t seed, effsc_hip_mlp** out);
int effsc_hip_mlp_param_count(const effsc_hip_mlp* m, size_t* out);
int effsc_hip_mlp_get_params(effsc_hip_mlp* m, float* out);
/* Also resets the optimiser state. */
int effsc_hip_mlp_set_params(effsc_hip_mlp* m, const float* in);
int effsc_hip_mlp_forward(effsc_hip_mlp* m, const float* x, size_t n, float* y);
/* Mean squared error over the n samples and its gradient (param_count floats). */
int effsc_hip_mlp_loss_grad(effsc_hip_mlp* m, const float* x, const float* t, size_t n, float* loss, float* grad);
/* `steps` full-batch Adam steps at learning rate lr; the loss afterwards. */
int effsc_hip_mlp_train(effsc_hip_mlp* m, const float* x, const float* t, size_t n, int steps, float lr,
                        float* final_loss);
void effsc_hip_mlp_free(effsc_hip_mlp* m);

/* Device buffers for callers that build tables on the device. */
int effsc_hip_device_alloc(int device, size_t words, uint64_t** out);
int effsc_hip_device_free(int device, uint64_t* ptr);
int effsc_hip_copy(int device, uint64_t* dst, const uint64_t* src, size_t words);
/* Counter-based uniform values in [0, p): dst[i] = U(seed, first + i). */
int effsc_hip_fill_uniform(int device, uint64_t* dst, size_t n, uint64_t seed, uint64_t first);
/* Measured Goldilocks multiply throughput of the device (products per
 * second over independent chains filling every CU): the compute roofline of
 * the degree >= 3 rounds.  No reference counterpart. */
int effsc_hip_mul_rate(int device, double* muls_per_s);
/* Measured attainable HBM rate (bytes per second, best of five launches) of
 * a kernel with no field arithmetic that moves the bytes of one access
 * pattern over two tables of n words (n a power of two, >= 1024; the call
 * allocates 2.5 n words): a stream copy of one table (16 n bytes), the read
 * pass of the two-round passes (both tables read in four quarter regions,
 * 16 n bytes), or the fold pass with two pending folds (both tables read in
 * sixteen regions, level-n/4 tables written, 20 n bytes).  The HBM side of
 * the roofline.  No reference counterpart. */
enum { EFFSC_HIP_HBM_COPY = 0, EFFSC_HIP_HBM_READ_PASS = 1, EFFSC_HIP_HBM_FOLD_PASS = 2 };
int effsc_hip_hbm_rate(int device, int pattern, size_t n, double* bytes_per_s);

#ifdef __cplusplus
}
#endif
#endif /* EFFSC_HIP_H */

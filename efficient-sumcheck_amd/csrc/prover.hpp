// Internal interface of the prover library, shared by its translation units
// (capi.hip, devloop.hip, eq.hip, gkr.hip, stream.hip, ops_f*.hip): the
// prover handle, the helpers every section calls and the round launchers'
// declarations (definitions in ops.hpp, instantiated once per output field
// in ops_f1.hip, ops_f2.hip and ops_f3.hip so that the heavy kernel
// instantiations compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "effsc_hip.h"
#include "kernels.hpp"
#include "pass2.hpp"
#include "prove.hpp"
#include "shake.hpp"

namespace effsc_hip {
namespace {
// Import / export / synthetic-input kernels (used by this file only).
// AoS caller words -> canonical SoA planes (optionally from Montgomery form).
[[maybe_unused]] __global__ __launch_bounds__(kBlock) void k_import(uint64_t* dst, size_t stride, const uint64_t* src, size_t n,
                                                   int d, int mont) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < n; i += step) {
    for (int c = 0; c < d; c++) {
      uint64_t x = src[i * d + c];
      x = mont ? gl::mul(x, gl::MONT_TO_CANON) : gl::canon(x);
      dst[c * stride + i] = x;
    }
  }
}

// Canonical SoA planes -> AoS caller words (optionally to Montgomery form).
[[maybe_unused]] __global__ __launch_bounds__(kBlock) void k_export(uint64_t* dst, const uint64_t* src, size_t stride, size_t n,
                                                   int d, int mont) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < n; i += step) {
    for (int c = 0; c < d; c++) {
      uint64_t x = src[c * stride + i];
      dst[i * d + c] = mont ? gl::mul(x, gl::CANON_TO_MONT) : x;
    }
  }
}

}  // namespace
}  // namespace effsc_hip

using effsc_hip::DevFs;
using effsc_hip::kMaxK;
using effsc_hip::Reduce;

// ---------------------------------------------------------------------------
// Prover handle
// ---------------------------------------------------------------------------
struct effsc_hip_prover {
  int kind = 0;
  int ext = 1, in_ext = 1, K = 1, deg = 1, repr = 0, dev = 0;
  uint64_t w = EFFSC_HIP_W_FROM_RAW7;
  size_t n = 0, len = 0;
  bool in_work = false, own_in = true;
  uint64_t* in[kMaxK] = {};
  size_t in_stride = 0;
  uint64_t* work[kMaxK] = {};
  size_t work_stride = 0;
  hipStream_t stream = nullptr;
  Reduce red = {};
  uint64_t* h_res = nullptr;  // pinned, kMaxK * 3 words
  int grid_max = 2048;
  // EqFactored state (src/provers/eq_factored.rs:67-83)
  size_t v = 0, v_l = 0, rounds_elapsed = 0;
  uint64_t* eq_l = nullptr;  // ext planes, stride eq_l_stride
  uint64_t* eq_r = nullptr;
  size_t eq_l_stride = 0, eq_r_stride = 0, eq_l_len = 0, eq_r_len = 0;
  uint64_t* eq_point = nullptr;  // canonical point, v * ext words (rebuilds the eq tables on reset)
  uint64_t* eq_full = nullptr;   // EQ = eq_l x eq_r at the register-tail level (at most 2^14 words)
  // fresh eq_l / eq_r built behind a device-loop proof (after its proof
  // words are copied out): reset swaps them in instead of rebuilding
  uint64_t* eq_l_spare = nullptr;
  uint64_t* eq_r_spare = nullptr;
  bool eq_spare_ready = false;
  // the other eq_l buffer of the fused base-field passes (k_eq_pass_sv_fs
  // writes the folded table there; the two swap after the launch)
  uint64_t* eq_l_alt = nullptr;
  // eq_r with its coefficient sums for Karatsuba products (extension-field
  // passes): 3 planes over F_p^2, 6 over F_p^3, stride eq_r_stride
  uint64_t* eq_rp = nullptr;
  size_t eq_rp_words = 0;
  // LSB (pair-split) provers hold their tables bit-reversed and run the MSB
  // path; post-state reads reverse back (src/streams/memory/core.rs:20-47)
  bool lsb = false;
  // MultilinearProverLSB over a table whose length is not a power of two
  // (src/provers/multilinear_lsb.rs:59-126): no bit reversal; native
  // pair-split rounds ping-pong between work and work_alt
  bool lsb_native = false;
  uint64_t* work_alt[effsc_hip::kMaxK] = {};
  // Provers built from streams: the first two-round pass (round 0 values and
  // round 1 coefficients over the input) was accumulated during the upload
  // and its reduced values live in d_pass0 (the input is never written, so
  // they stay valid across resets)
  uint64_t* d_pass0 = nullptr;
  bool pass0_ready = false;
  // Out-of-core stream provers: the input stays in the host streams; the two
  // passes that read it (rounds 0-1 at construction, rounds 2-3 in the proof)
  // move it through pinned and device staging chunk by chunk, and only the
  // level-2 tables (n/4 elements each) are ever resident
  bool ooc = false;
  const effsc_hip_stream* ooc_src[kMaxK] = {};
  uint64_t* ooc_parts = nullptr;  // partials of the streamed fold pass
  uint64_t* gkr_b = nullptr;  // GKR factored witness: two ping-pong planesets of n/2 elements (3 planes)
  size_t ooc_nparts = 0;
  // Device-resident proof loop (effsc_hip_sumcheck without a hook)
  bool dev_mode = false;       // Ops skip the plain reduction; grid in last_grid
  unsigned last_grid = 0;
  // one device block, polys | chals | fin | fs, the layout of h_proof
  uint64_t* d_polys = nullptr;  // kMaxRounds * kMaxK * 3 words (block start)
  uint64_t* d_chals = nullptr;  // kMaxRounds * 3 words
  uint64_t* d_fin = nullptr;    // 4 words
  DevFs* d_fs = nullptr;
  uint64_t* d_cur = nullptr;    // 3 words: current challenge (own allocation)
  uint64_t* h_proof = nullptr;  // staging: polys | chals | fin | fs
  bool h_proof_pinned = false;  // h_proof from hipHostMalloc (device-readable)
  // set when the last kernel of a device-loop proof wrote the proof block into
  // h_proof itself (k_tail_pipe), so no copy is needed
  bool proof_by_tail = false;
  // set when the proof block's copy to h_proof is already enqueued (work
  // for the post-state only may follow it on the stream); ev_proof marks it
  bool proof_copied = false;
  hipEvent_t ev_proof = nullptr;
  // where the current transcript state is when it is not in d_fs: the pinned
  // staging of a starting device-loop proof, or d_fs_alt after a fused pass;
  // the next transcript kernel loads it from there (fs_in), any other next
  // user gets it copied to d_fs (fs_flush)
  const DevFs* fs_pending = nullptr;
  // fused passes (k_pass_fs2): the second sponge slot and partials buffer
  // (kFuseMaxGrid blocks of up to 8 words), allocated with the proof block;
  // null (no fused passes) for the allocation-free entry point
  DevFs* d_fs_alt = nullptr;
  uint64_t* fuse_parts = nullptr;
  // Kernel timing (effsc_hip_profile): HIP events around each launch on the
  // prover's stream; slot 0 fused fold+evaluate, 1 evaluate, 2 fold,
  // 3 two-round pass.
  bool prof_on = false;
  std::vector<hipEvent_t> ev;             // pairs (2i, 2i+1)
  struct Pend {
    int slot;
    double bytes, muls;
  };
  std::vector<Pend> pend;  // recorded event pairs
  uint64_t prof_launches[4] = {0, 0, 0, 0};
  double prof_ms[4] = {0, 0, 0, 0};
  double prof_bytes[4] = {0, 0, 0, 0};
  double prof_muls[4] = {0, 0, 0, 0};  // Goldilocks products (ext3 product = 8)
  // Two-round passes (base-field ML / IP): the logical table is the
  // materialised one (len) folded by the npend challenges in fpend; when
  // have_next, the next round's values are the polynomials in next_coef
  // (pass layout of pass2.hpp) evaluated at the next challenge.
  bool twostep = false;
  int npend = 0;
  uint64_t fpend[2] = {0, 0};
  bool have_next = false;
  uint64_t next_coef[effsc_hip::kMaxK + effsc_hip::kMaxK * (effsc_hip::kMaxK + 1)] = {};
};

namespace effsc_hip {
namespace pv {

using gl::Fp;
using gl::Fp2;
using gl::Fp3;

int fail(int code, const std::string& msg);

enum Kind { K_ML = 1, K_IP = 2, K_PRODUCT = 3, K_EQ = 4, K_GKR = 5 };

inline size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}
inline bool is_pow2(size_t n) { return n && !(n & (n - 1)); }
constexpr size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }
int grid_max_for(int dev);

constexpr size_t kMaxRounds = 64;
constexpr int kPassGridScale = 4;
// planes of at least this many bytes are streamed past the caches by the
// passes that read or write them once (nontemporal loads and stores)
constexpr double kNtMinBytes = 134217728.0;
// Words of the proof block: polys | chals | fin (4) | transcript state.  The
// device block and the pinned host staging share this layout, so one copy
// brings a whole proof back (each small copy is a blit of about 5 us on the
// stream).
constexpr size_t kProofWords = kMaxRounds * kMaxK * 3 + kMaxRounds * 3 + 4 + (sizeof(DevFs) + 7) / 8;

int alloc_dev(uint64_t** p, size_t words);
int prof_begin(effsc_hip_prover* h);
int prof_end(effsc_hip_prover* h, int slot, double bytes, double muls);
int prof_collect(effsc_hip_prover* h);
unsigned blocks_for(const effsc_hip_prover* h, size_t units);

// Algorithmic Goldilocks products per field product and per fold product
// c * (hi - lo) with c in FO, the difference in FI.  An F_p^3 product counts
// as Karatsuba's 8 full products (6 + 2 by w) and an F_p^2 product as 4
// (3 + 1 by w) whatever the kernels execute: they run unreduced 64x64
// products with one reduction per coefficient (gl::mul(FpN, FpNPre)), so the
// count is a fixed unit of work.
template <class F>
constexpr double mul_cost() {
  return F::D == 1 ? 1.0 : F::D == 2 ? 4.0 : 8.0;
}
template <class FI, class FO>
constexpr double fold_cost() {
  return FI::D == FO::D ? mul_cost<FO>() : (double)FO::D;
}
// products of one EvalsInfty pair evaluation: K values of K - 1 products
constexpr double eval_products(int K) { return K == 1 ? 0.0 : (double)K * (K - 1); }

template <class F, int K>
Views<F, K> views(uint64_t* const* planes, size_t stride) {
  Views<F, K> v;
  for (int i = 0; i < K; i++) v.t[i] = View<F>{planes[i], stride};
  return v;
}

template <class F>
F from_words(const uint64_t* w) {
  return gl::load_words<F>(w);
}

template <class F>
F fmul(const effsc_hip_prover* h, F a, F b) {
  return gl::mul(a, b, h->w);
}

// Calls fn(F{}) for the field of `ext` words per element (1, 2 or 3).
template <class Fn>
int with_field(int ext, Fn&& fn) {
  switch (ext) {
    case 1: return fn(Fp{});
    case 2: return fn(Fp2{});
    case 3: return fn(Fp3{});
  }
  return fail(EFFSC_HIP_EINVAL, "ext must be 1, 2 or 3");
}

// Host products of field elements given as `ext` canonical words:
// out = a b, and out = a b c.
void ext_mul(int ext, uint64_t w, const uint64_t* a, const uint64_t* b, uint64_t* out);
void ext_mul3(int ext, uint64_t w, const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* out);

// Sums the last launch's partials into red.result (one workgroup), unless the
// device loop runs its own reduction-and-transcript kernel next.
template <class F, int K>
int reduce(effsc_hip_prover* h, unsigned nblocks) {
  h->last_grid = nblocks;
  if (h->dev_mode) return EFFSC_HIP_OK;  // the caller launches k_reduce_fs
  hipLaunchKernelGGL((k_reduce<F, K>), dim3(1), dim3(1024), 0, h->stream, h->red, nblocks);
  HIPCHK(hipGetLastError());
  return EFFSC_HIP_OK;
}

// Round launchers of the product family over K tables, input field FI and
// output field FO (ops.hpp).  eval: round polynomial without a fold; fused:
// fold with the challenge (host words, or device words at cptr) and evaluate;
// fold: fold only.
template <class FI, class FO, int K>
struct Ops {
  using In = FI;
  using Out = FO;
  static int eval(effsc_hip_prover* h, uint64_t* const* planes, size_t stride, size_t half, size_t paired);
  static int fused(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal, size_t q,
                   const uint64_t* cptr = nullptr);
  static int fold(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal, size_t half,
                  size_t paired, const uint64_t* cptr = nullptr);
};

// Device-loop launchers (ops.hpp): the reduction-and-transcript step and the
// one-workgroup tail.
template <class FI, class F, int K>
struct DevOps {
  static int reduce_fs(effsc_hip_prover* h, unsigned g, size_t round);
  static int tail(effsc_hip_prover* h, uint64_t* const* src, size_t sst, bool from_input, size_t L, int rounds,
                  bool have, size_t first_round, const uint64_t* cur);
};

template <template <class, class, int> class O, class FI, class FO, class Fn>
int with_k(int K, Fn&& fn) {
  switch (K) {
    case 1: return fn(O<FI, FO, 1>{});
    case 2: return fn(O<FI, FO, 2>{});
    case 3: return fn(O<FI, FO, 3>{});
    case 4: return fn(O<FI, FO, 4>{});
  }
  return fail(EFFSC_HIP_EINVAL, "unsupported table count");
}

// Calls fn(O<FI, FO, K>{}) for the runtime field degrees and table count.
template <template <class, class, int> class O = Ops, class Fn>
int dispatch(int fi, int fo, int K, Fn&& fn) {
  if (fi == 1 && fo == 1) return with_k<O, Fp, Fp>(K, fn);
  if (fi == 1 && fo == 2) return with_k<O, Fp, Fp2>(K, fn);
  if (fi == 2 && fo == 2) return with_k<O, Fp2, Fp2>(K, fn);
  if (fi == 1 && fo == 3) return with_k<O, Fp, Fp3>(K, fn);
  if (fi == 3 && fo == 3) return with_k<O, Fp3, Fp3>(K, fn);
  return fail(EFFSC_HIP_EINVAL, "unsupported field combination");
}

size_t tail_max(const effsc_hip_prover* h);
// The pipelined tail (k_tail_pipe): base-field tables, K <= 2, L0 (the length
// after the pending folds) a power of two in [4, pipe_max<K>()], 1 <= rounds
// <= log2(L0); EFFSC_HIP_TAIL_REG or EFFSC_HIP_TAIL_PLAIN keep the earlier
// tails.
bool pipe_tail_ok(const effsc_hip_prover* h, size_t L0, size_t rounds);
// The sponge a transcript kernel starts from: the pending state (taken; the
// kernel writes its result to d_fs) or d_fs.
const DevFs* fs_in(effsc_hip_prover* h);
// Copies a pending transcript state to d_fs (before a kernel that only reads
// d_fs, and before the proof block is copied back).
int fs_flush(effsc_hip_prover* h);
int launch_tail_pipe(effsc_hip_prover* h, uint64_t* const* src, size_t sst, size_t L0, int pend, int rounds,
                     size_t first_round, const uint64_t* cur);
int src_field(const effsc_hip_prover* h);
uint64_t* const* src_planes(const effsc_hip_prover* h);
size_t src_stride(const effsc_hip_prover* h);
void geometry(const effsc_hip_prover* h, size_t L, size_t* half, size_t* paired);
int read_elements(effsc_hip_prover* h, size_t idx, uint64_t* out);
int fetch_result(effsc_hip_prover* h, int fi, uint64_t* out_vals, int count);
int pass_value_count(const effsc_hip_prover* h);
int fetch_pass(effsc_hip_prover* h, uint64_t* vals);
int fold_step(effsc_hip_prover* h, const uint64_t* c, bool fuse, bool* evaluated);
int evaluate(effsc_hip_prover* h, bool already, uint64_t* vals);
void to_wire(const effsc_hip_prover* h, const uint64_t* vals, uint64_t* out);
int launch_pass(effsc_hip_prover* h, int p, size_t q, const uint64_t* cptr);
int materialize(effsc_hip_prover* h);
size_t logical_len(const effsc_hip_prover* h);
int round_impl(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out);
int finalize_impl(effsc_hip_prover* h, const uint64_t* chal);
int final_value_impl(effsc_hip_prover* h, uint64_t* out);
void words_in(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n);
void words_out(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n);
int check_cfg(const effsc_hip_config* cfg);
void destroy(effsc_hip_prover* h);
int setup_common(effsc_hip_prover* h, size_t cap);
int import_tables(effsc_hip_prover* h, const effsc_hip_config* cfg, const uint64_t* const* tables, size_t n);
int new_product_like(const effsc_hip_config* cfg, int kind, int K, int deg, const uint64_t* const* tables, size_t n,
                     effsc_hip_prover** out);
int ooc_guard(const effsc_hip_prover* h);

inline uint64_t gl_brev64(uint64_t x) {
  uint64_t r = 0;
  for (int b = 0; b < 64; b++) r |= ((x >> b) & 1) << (63 - b);
  return r;
}

// device-resident proof loop (devloop.hip)
bool device_loop_applies(const effsc_hip_prover* h, size_t num_rounds);
int ensure_dev_bufs(effsc_hip_prover* h);
int copy_proof_out(effsc_hip_prover* h);
int prove_device(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                 uint64_t* challenges, uint64_t* final_value);
struct DevModeGuard {
  effsc_hip_prover* h;
  explicit DevModeGuard(effsc_hip_prover* p) : h(p) { h->dev_mode = true; }
  ~DevModeGuard() { h->dev_mode = false; }
};

// two-round passes over extension fields in the device loop (pass_ext.hip)
bool ext_twostep(const effsc_hip_prover* h);
int launch_pass_ext(effsc_hip_prover* h, int p, size_t q, const uint64_t* cptr);
int launch_reduce_fs2_ext(effsc_hip_prover* h, const Reduce& red, unsigned g, size_t round);
// Two-round EqFactored passes over F_p^2 and F_p^3 while both variables of a
// pass are left variables (split-value form): the prepared eq_r table, one
// pass over p (p pending folds, q units, chunk units per workgroup), and the
// two transcript rounds with the folds of eq_l (halves h0 then h1; 0 skips).
int eq_prep_r_ext(effsc_hip_prover* h);
int eq_pass_ext(effsc_hip_prover* h, int p, size_t q, size_t chunk, unsigned m_log, unsigned g);
int eq_reduce_fs2_ext(effsc_hip_prover* h, unsigned g, size_t round, size_t h0, size_t h1);

// EqFactoredProver (eq.hip)
int eq_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out);
int eq_finalize(effsc_hip_prover* h, const uint64_t* chal);
int eq_final_value(effsc_hip_prover* h, uint64_t* out);
int build_eq(effsc_hip_prover* h, const uint64_t* wpt_canon_dev, size_t off, size_t vv, uint64_t** planes,
             size_t* stride, size_t* len);
bool eq_device_applies(const effsc_hip_prover* h, size_t num_rounds);
int prove_device_eq(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                    uint64_t* challenges, uint64_t* final_value);

// GkrProver (gkr.hip)
int gkr_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out);
int gkr_final_value(effsc_hip_prover* h, uint64_t* out);
bool gkr_device_applies(const effsc_hip_prover* h, size_t num_rounds);
int prove_device_gkr(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                     uint64_t* challenges, uint64_t* final_value);

// streamed input (stream.hip)
int ooc_fold_pass(effsc_hip_prover* h, size_t q, unsigned* nblocks);

}  // namespace pv
}  // namespace effsc_hip

// effsc.hpp — C++ host API over the gfx950 library (include/effsc_hip.h).
//
// Mirrors the reference crate's protocol layer: the SumcheckProver trait
// (src/sumcheck_prover.rs:36-71), runner::sumcheck (src/runner.rs:29-69),
// verifier::sumcheck_verify (src/verifier.rs:61-139), SumcheckProof /
// SumcheckError (src/proof.rs:17-53), the transcript traits
// (src/transcript/transcript.rs:11-43) and noop hooks (src/lib.rs:28-40).
// The GPU provers forward every trait method to the C ABI.  Header-only.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../effsc_hip.h"

namespace effsc {

// ------------------------------------------------------------------ fields
// Elements are canonical words: F64 one word, F64Ext2 two (c0, c1) of
// F_p[X]/(X^2 - w), F64Ext3 three (c0, c1, c2) of F_p[X]/(X^3 - w).  Arithmetic lives on the device; the host types are
// plain carriers (the verifier's arithmetic runs inside the library).
struct F64 {
  static constexpr int EXT = 1;
  uint64_t v = 0;
  F64() = default;
  explicit F64(uint64_t x) : v(x % 0xFFFFFFFF00000001ull) {}
  const uint64_t* words() const { return &v; }
  static F64 from_words(const uint64_t* w) { return F64(w[0]); }
  bool operator==(const F64& o) const { return v == o.v; }
  bool operator!=(const F64& o) const { return v != o.v; }
};

struct F64Ext2 {
  static constexpr int EXT = 2;
  std::array<uint64_t, 2> c{0, 0};
  F64Ext2() = default;
  F64Ext2(uint64_t c0, uint64_t c1) : c{c0, c1} {}
  const uint64_t* words() const { return c.data(); }
  static F64Ext2 from_words(const uint64_t* w) { return F64Ext2(w[0], w[1]); }
  bool operator==(const F64Ext2& o) const { return c == o.c; }
  bool operator!=(const F64Ext2& o) const { return c != o.c; }
};

struct F64Ext3 {
  static constexpr int EXT = 3;
  std::array<uint64_t, 3> c{0, 0, 0};
  F64Ext3() = default;
  F64Ext3(uint64_t c0, uint64_t c1, uint64_t c2) : c{c0, c1, c2} {}
  const uint64_t* words() const { return c.data(); }
  static F64Ext3 from_words(const uint64_t* w) { return F64Ext3(w[0], w[1], w[2]); }
  bool operator==(const F64Ext3& o) const { return c == o.c; }
  bool operator!=(const F64Ext3& o) const { return c != o.c; }
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != EFFSC_HIP_OK) throw Error(rc, effsc_hip_last_error());
}

// ------------------------------------------------------------- proof types
template <class F>
struct SumcheckProof {
  std::vector<std::vector<F>> round_polys;  // EvalsInfty values per round
  std::vector<F> challenges;
  F final_value{};
};

template <class F>
struct SumcheckResult {
  std::vector<F> challenges;
  F final_claim{};
};

struct SumcheckError : std::runtime_error {
  enum Kind { ConsistencyCheck, DegreeMismatch, FinalEvaluation, TranscriptError, HookError } kind;
  size_t round;
  SumcheckError(Kind k, size_t r)
      : std::runtime_error(k == TranscriptError ? "round " + std::to_string(r) + ": transcript error"
                           : k == HookError     ? "round " + std::to_string(r) + ": hook error"
                                                : "sumcheck error"),
        kind(k),
        round(r) {}
};

// ------------------------------------------------------------ transcripts
// Prover/verifier transcript over the library's native transcripts
// (SHAKE256 Fiat-Shamir, or the test RNG transcript).  A verifier-side
// transcript additionally reads prover messages from a proof.
template <class F>
class Transcript {
 public:
  static Transcript fiat_shamir(const std::string& domain_separator) {
    return Transcript(effsc_hip_transcript_new(reinterpret_cast<const uint8_t*>(domain_separator.data()),
                                               domain_separator.size()));
  }
  static Transcript test(uint64_t seed) { return Transcript(effsc_hip_transcript_new_test(seed)); }

  Transcript(Transcript&& o) noexcept : t_(std::exchange(o.t_, nullptr)), tape_(std::move(o.tape_)), pos_(o.pos_) {}
  Transcript& operator=(Transcript&& o) noexcept {
    std::swap(t_, o.t_);
    tape_ = std::move(o.tape_);
    pos_ = o.pos_;
    return *this;
  }
  Transcript(const Transcript&) = delete;
  ~Transcript() {
    if (t_) effsc_hip_transcript_free(t_);
  }

  // ProverTranscript::send / ::challenge
  void send(const F& v) { check(effsc_hip_transcript_send(t_, F::EXT, v.words())); }
  F challenge() {
    uint64_t w[3];
    check(effsc_hip_transcript_challenge(t_, F::EXT, w));
    return F::from_words(w);
  }
  // VerifierTranscript::receive: next prover message of the attached proof
  void attach(const SumcheckProof<F>& proof) {
    tape_.clear();
    for (auto& rp : proof.round_polys)
      for (auto& x : rp) tape_.push_back(x);
    pos_ = 0;
  }
  std::optional<F> receive() {
    if (pos_ >= tape_.size()) return std::nullopt;
    F v = tape_[pos_++];
    send(v);
    return v;
  }
  effsc_hip_transcript* raw() { return t_; }
  const std::vector<F>& tape() const { return tape_; }

 private:
  explicit Transcript(effsc_hip_transcript* t) : t_(t) {
    if (!t_) throw Error(EFFSC_HIP_ENOMEM, "transcript allocation failed");
  }
  effsc_hip_transcript* t_ = nullptr;
  std::vector<F> tape_;
  size_t pos_ = 0;
};

// ------------------------------------------------------------ prover trait
template <class F>
class SumcheckProver {
 public:
  virtual ~SumcheckProver() = default;
  virtual size_t degree() const = 0;
  virtual std::vector<F> round(std::optional<F> challenge) = 0;
  virtual void finalize(F last_challenge) = 0;
  virtual F final_value() const = 0;
};

// ------------------------------------------------------------- GPU provers
template <class F>
class GpuProver : public SumcheckProver<F> {
 public:
  GpuProver(GpuProver&& o) noexcept : h_(std::exchange(o.h_, nullptr)), deg_(o.deg_) {}
  GpuProver(const GpuProver&) = delete;
  ~GpuProver() override {
    if (h_) effsc_hip_free(h_);
  }
  size_t degree() const override { return deg_; }
  std::vector<F> round(std::optional<F> challenge) override {
    std::vector<uint64_t> out(std::max<size_t>(deg_, 1) * F::EXT);
    check(effsc_hip_round(h_, challenge ? challenge->words() : nullptr, out.data()));
    std::vector<F> r;
    for (size_t i = 0; i < out.size(); i += F::EXT) r.push_back(F::from_words(out.data() + i));
    return r;
  }
  void finalize(F c) override { check(effsc_hip_finalize(h_, c.words())); }
  F final_value() const override {
    uint64_t w[3];
    check(effsc_hip_final_value(h_, w));
    return F::from_words(w);
  }
  // post-state: table i (evals(), evaluations(), pairwise(), final_factors())
  std::vector<F> table(size_t i) const {
    size_t n = 0;
    check(effsc_hip_table_len(h_, i, &n));
    std::vector<uint64_t> w(std::max<size_t>(n, 1) * F::EXT);
    check(effsc_hip_read_table(h_, i, w.data(), n * F::EXT, &n));
    std::vector<F> r;
    for (size_t k = 0; k < n; k++) r.push_back(F::from_words(w.data() + k * F::EXT));
    return r;
  }
  void reset() { check(effsc_hip_reset(h_)); }
  effsc_hip_prover* raw() const { return h_; }

 protected:
  GpuProver() = default;
  void adopt(effsc_hip_prover* h) {
    h_ = h;
    check(effsc_hip_degree(h_, &deg_));
  }
  static effsc_hip_config config(int input_ext, int device) {
    effsc_hip_config c;
    effsc_hip_config_default(&c);
    c.ext = F::EXT;
    c.input_ext = input_ext;
    c.device = device;
    return c;
  }
  effsc_hip_prover* h_ = nullptr;
  size_t deg_ = 0;
};

// Words of a table of elements (AoS).
template <class F>
std::vector<uint64_t> flatten(const std::vector<F>& t) {
  std::vector<uint64_t> w;
  w.reserve(t.size() * F::EXT);
  for (auto& x : t) w.insert(w.end(), x.words(), x.words() + F::EXT);
  return w;
}

// MultilinearProver (src/provers/multilinear.rs:25-80)
template <class F>
class MultilinearProver : public GpuProver<F> {
 public:
  explicit MultilinearProver(const std::vector<F>& evals, int device = 0) {
    auto cfg = this->config(F::EXT, device);
    auto w = flatten(evals);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_multilinear_new(&cfg, w.data(), evals.size(), &h));
    this->adopt(h);
  }
  std::vector<F> evals() const { return this->table(0); }
};

// InnerProductProver (src/provers/inner_product.rs:26-127)
template <class F>
class InnerProductProver : public GpuProver<F> {
 public:
  InnerProductProver(const std::vector<F>& a, const std::vector<F>& b, int device = 0) {
    if (a.size() != b.size()) throw Error(EFFSC_HIP_EINVAL, "a and b must have equal length");
    auto cfg = this->config(F::EXT, device);
    auto wa = flatten(a), wb = flatten(b);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_inner_product_new(&cfg, wa.data(), wb.data(), a.size(), &h));
    this->adopt(h);
  }
  std::pair<F, F> final_evaluations() const {
    auto a = this->table(0), b = this->table(1);
    if (a.size() == 1) return {a[0], b[0]};
    return {F{}, F{}};
  }
};

// CoefficientProver with the built-in product-of-k evaluator
// (src/provers/coefficient.rs:18-330); base-field input tables are lifted.
template <class F>
class CoefficientProver : public GpuProver<F> {
 public:
  explicit CoefficientProver(const std::vector<std::vector<F64>>& pairwise, int device = 0) {
    auto cfg = this->config(1, device);
    std::vector<std::vector<uint64_t>> w;
    std::vector<const uint64_t*> ptrs;
    for (auto& t : pairwise) w.push_back(flatten(t));
    for (auto& t : w) ptrs.push_back(t.data());
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_product_new(&cfg, ptrs.data(), (int)ptrs.size(), pairwise.empty() ? 0 : pairwise[0].size(), &h));
    this->adopt(h);
  }
};

// EqFactoredProver (src/provers/eq_factored.rs:67-289)
template <class F>
class EqFactoredProver : public GpuProver<F> {
 public:
  EqFactoredProver(const std::vector<F>& w, const std::vector<F>& p, int device = 0) {
    auto cfg = this->config(F::EXT, device);
    auto ww = flatten(w), pw = flatten(p);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_eq_factored_new(&cfg, ww.data(), w.size(), pw.data(), p.size(), &h));
    this->adopt(h);
  }
};

// GkrProver (src/provers/gkr.rs:38-202): add and mult over (b, c) in
// {0,1}^{2k} (n * n elements, index b n + c), w the n witness values.
template <class F>
class GkrProver : public GpuProver<F> {
 public:
  GkrProver(const std::vector<F>& add, const std::vector<F>& mult, const std::vector<F>& w, int device = 0) {
    if (add.size() != w.size() * w.size()) throw Error(EFFSC_HIP_EINVAL, "add_evals must have len w^2");
    if (mult.size() != w.size() * w.size()) throw Error(EFFSC_HIP_EINVAL, "mult_evals must have len w^2");
    auto cfg = this->config(F::EXT, device);
    auto wa = flatten(add), wm = flatten(mult), ww = flatten(w);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_gkr_new(&cfg, wa.data(), wm.data(), ww.data(), w.size(), &h));
    this->adopt(h);
  }
  // claimed_w_values (src/provers/gkr.rs:89-97): (W(b*), W(c*))
  std::pair<F, F> claimed_w_values() const {
    uint64_t out[6];
    check(effsc_hip_gkr_claimed_w(this->h_, out));
    return {F::from_words(out), F::from_words(out + F::EXT)};
  }
};

// LSB (pair-split) provers (src/provers/{multilinear,inner_product}_lsb.rs):
// power-of-two tables, pairs (2k, 2k+1).
template <class F>
class MultilinearProverLSB : public GpuProver<F> {
 public:
  explicit MultilinearProverLSB(const std::vector<F>& evals, int device = 0) {
    auto cfg = this->config(F::EXT, device);
    auto w = flatten(evals);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_multilinear_lsb_new(&cfg, w.data(), evals.size(), &h));
    this->adopt(h);
  }
  std::vector<F> evals() const { return this->table(0); }
};

template <class F>
class InnerProductProverLSB : public GpuProver<F> {
 public:
  InnerProductProverLSB(const std::vector<F>& a, const std::vector<F>& b, int device = 0) {
    if (a.size() != b.size()) throw Error(EFFSC_HIP_EINVAL, "a and b must have equal length");
    auto cfg = this->config(F::EXT, device);
    auto wa = flatten(a), wb = flatten(b);
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_inner_product_lsb_new(&cfg, wa.data(), wb.data(), a.size(), &h));
    this->adopt(h);
  }
};

// FileStream (src/streams/file/core.rs:13-76) of base-field elements.
class FileStream {
 public:
  explicit FileStream(const std::string& path, int ext = 1) { check(effsc_hip_stream_open(path.c_str(), ext, &s_)); }
  FileStream(FileStream&& o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
  FileStream(const FileStream&) = delete;
  ~FileStream() {
    if (s_) effsc_hip_stream_close(s_);
  }
  size_t num_variables() const {
    size_t v = 0;
    check(effsc_hip_stream_num_variables(s_, &v));
    return v;
  }
  F64 evaluation(size_t point) const {
    uint64_t w = 0;
    check(effsc_hip_stream_read(s_, point, 1, &w));
    return F64{w};
  }
  effsc_hip_stream* raw() const { return s_; }

 private:
  effsc_hip_stream* s_ = nullptr;
};

// multivariate_claim (src/streams/stream.rs:3-14), on the device.
inline F64 multivariate_claim(const FileStream& s, int device = 0) {
  effsc_hip_stream* h[1] = {s.raw()};
  uint64_t out[3] = {0, 0, 0};
  check(effsc_hip_stream_claim(device, h, 1, EFFSC_HIP_W_FROM_RAW7, out));
  return F64{out[0]};
}

// An InnerProductProver over two streams' tables: the upload overlaps the
// first pass over the data.
template <class F>
class StreamInnerProductProver : public GpuProver<F> {
 public:
  StreamInnerProductProver(const FileStream& a, const FileStream& b, int device = 0) {
    auto cfg = this->config(1, device);
    effsc_hip_stream* hs[2] = {a.raw(), b.raw()};
    effsc_hip_prover* h = nullptr;
    check(effsc_hip_stream_prover_new(&cfg, 2, hs, &h));
    this->adopt(h);
  }
};

// ------------------------------------------------------------------- hooks
struct NoopHook {
  template <class T>
  void operator()(size_t, T&) const {}
};
inline constexpr NoopHook noop_hook{};

// ------------------------------------------------------------------ runner
// runner::sumcheck (src/runner.rs:29-69) for any prover and transcript.
template <class F, class T, class Hook = NoopHook>
SumcheckProof<F> sumcheck(SumcheckProver<F>& prover, size_t num_rounds, T& transcript, Hook hook = Hook{}) {
  SumcheckProof<F> proof;
  std::optional<F> prev;
  for (size_t round = 0; round < num_rounds; round++) {
    auto evals = prover.round(prev);
    for (auto& v : evals) transcript.send(v);
    proof.round_polys.push_back(evals);
    hook(round, transcript);
    F r = transcript.challenge();
    proof.challenges.push_back(r);
    prev = r;
  }
  if (prev) prover.finalize(*prev);
  proof.final_value = prover.final_value();
  return proof;
}

// GPU prover + native transcript with no hook: the library runs the whole
// loop, on the device when the prover allows it.
template <class F>
SumcheckProof<F> sumcheck(GpuProver<F>& prover, size_t num_rounds, Transcript<F>& transcript, NoopHook = {}) {
  size_t nw = std::max<size_t>(prover.degree(), 1);
  std::vector<uint64_t> polys(std::max<size_t>(num_rounds, 1) * nw * F::EXT), chals(std::max<size_t>(num_rounds, 1) * F::EXT);
  uint64_t fin[3];
  check(effsc_hip_sumcheck(prover.raw(), num_rounds, transcript.raw(), nullptr, nullptr, polys.data(), chals.data(),
                           fin));
  SumcheckProof<F> proof;
  for (size_t r = 0; r < num_rounds; r++) {
    std::vector<F> rp;
    for (size_t i = 0; i < nw; i++) rp.push_back(F::from_words(polys.data() + (r * nw + i) * F::EXT));
    proof.round_polys.push_back(rp);
    proof.challenges.push_back(F::from_words(chals.data() + r * F::EXT));
  }
  proof.final_value = F::from_words(fin);
  return proof;
}

// ---------------------------------------------------------------- verifier
// verifier::sumcheck_verify (src/verifier.rs:61-139) over a verifier-side
// native transcript (Transcript::attach).  Throws SumcheckError.
template <class F>
SumcheckResult<F> sumcheck_verify(F claimed_sum, size_t expected_degree, size_t num_rounds, Transcript<F>& transcript,
                                  uint64_t ext_w = EFFSC_HIP_W_FROM_RAW7) {
  auto words = flatten(transcript.tape());
  std::vector<uint64_t> chals(std::max<size_t>(num_rounds, 1) * F::EXT);
  uint64_t fin[3];
  size_t err_round = 0;
  int err_kind = 0;
  int rc = effsc_hip_sumcheck_verify(F::EXT, ext_w, claimed_sum.words(), expected_degree, num_rounds,
                                     transcript.raw(), words.data(), words.size(), nullptr, nullptr, chals.data(), fin,
                                     &err_round, &err_kind);
  if (rc == EFFSC_HIP_EVERIFY)
    throw SumcheckError(err_kind == 1 ? SumcheckError::TranscriptError : SumcheckError::HookError, err_round);
  check(rc);
  SumcheckResult<F> res;
  for (size_t r = 0; r < num_rounds; r++) res.challenges.push_back(F::from_words(chals.data() + r * F::EXT));
  res.final_claim = F::from_words(fin);
  return res;
}

}  // namespace effsc

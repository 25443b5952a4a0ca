// Exercises the C++ host API (include/effsc/effsc.hpp) the way the
// reference's own tests exercise its Rust API (tests/canonical_api.rs,
// tests/adversarial_verifier.rs), and prints the proofs as JSON lines for the
// pytest driver to compare against the oracle.
//   host_api_test cpu   transcripts and verifier only (no GPU)
//   host_api_test gpu   GPU provers through the trait runner and the library runner
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "effsc/effsc.hpp"

using namespace effsc;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// counter-based uniform element of [0, p) (same rule as effsc_hip_fill_uniform)
static uint64_t uniform(uint64_t seed, uint64_t i) {
  uint64_t base = mix64(seed) + i * 0x9E3779B97F4A7C15ull;
  for (uint64_t a = 0;; a++) {
    uint64_t c = mix64(base + a * 0xD1B54A32D192ED03ull);
    if (c < 0xFFFFFFFF00000001ull) return c;
  }
}

static int failures = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                  \
    }                                                              \
  } while (0)

template <class F>
static void print_proof(const char* name, const SumcheckProof<F>& p) {
  std::printf("{\"name\": \"%s\", \"round_polys\": [", name);
  for (size_t r = 0; r < p.round_polys.size(); r++) {
    std::printf("%s[", r ? ", " : "");
    for (size_t i = 0; i < p.round_polys[r].size(); i++)
      std::printf("%s%llu", i ? ", " : "", (unsigned long long)p.round_polys[r][i].words()[0]);
    std::printf("]");
  }
  std::printf("], \"challenges\": [");
  for (size_t r = 0; r < p.challenges.size(); r++)
    std::printf("%s%llu", r ? ", " : "", (unsigned long long)p.challenges[r].words()[0]);
  std::printf("], \"final\": %llu}\n", (unsigned long long)p.final_value.words()[0]);
}

static int cpu_tests() {
  // transcript challenges, printed for comparison with the oracle
  auto t = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  t.send(F64(1));
  t.send(F64(2));
  F64 c1 = t.challenge();
  F64 c2 = t.challenge();
  std::printf("{\"name\": \"transcript\", \"challenges\": [%llu, %llu]}\n", (unsigned long long)c1.v,
              (unsigned long long)c2.v);
  // verifier: a proof shorter than the rounds is a TranscriptError at the
  // first round that runs out (src/verifier.rs:76-80)
  SumcheckProof<F64> short_proof;
  short_proof.round_polys = {{F64(1), F64(2)}, {F64(3)}};
  auto tv = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  tv.attach(short_proof);
  try {
    sumcheck_verify(F64(5), 2, 2, tv);
    EXPECT(false);
  } catch (const SumcheckError& e) {
    EXPECT(e.kind == SumcheckError::TranscriptError && e.round == 1);
  }
  return failures;
}

static int gpu_tests() {
  const size_t v = 10, n = size_t(1) << v;
  std::vector<F64> a(n), b(n);
  for (size_t i = 0; i < n; i++) {
    a[i] = F64(uniform(0xC0FFEE01, i));
    b[i] = F64(uniform(0xC0FFEE02, i));
  }
  uint64_t claim = 0;
  {
    // claimed sum on the host: sum a_i b_i mod p
    unsigned __int128 acc = 0;
    const unsigned __int128 p = 0xFFFFFFFF00000001ull;
    for (size_t i = 0; i < n; i++) acc = (acc + (unsigned __int128)a[i].v * b[i].v % p) % p;
    claim = (uint64_t)acc;
  }
  // (1) trait runner: per-round calls, hook order (tests/canonical_api.rs:180-197)
  InnerProductProver<F64> p1(a, b);
  auto t1 = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  std::vector<size_t> hook_rounds;
  auto proof1 = sumcheck<F64>(p1, v, t1, [&](size_t r, Transcript<F64>&) { hook_rounds.push_back(r); });
  EXPECT(hook_rounds.size() == v);
  for (size_t r = 0; r < hook_rounds.size(); r++) EXPECT(hook_rounds[r] == r);
  // (2) library runner (device-resident loop)
  InnerProductProver<F64> p2(a, b);
  auto t2 = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  auto proof2 = sumcheck(p2, v, t2);
  EXPECT(proof1.round_polys == proof2.round_polys);
  EXPECT(proof1.challenges == proof2.challenges);
  EXPECT(proof1.final_value == proof2.final_value);
  // final value = f(r) g(r) and post-state (src/provers/inner_product.rs:44-50)
  auto fe = p2.final_evaluations();
  {
    const unsigned __int128 p = 0xFFFFFFFF00000001ull;
    EXPECT((uint64_t)((unsigned __int128)fe.first.v * fe.second.v % p) == proof2.final_value.v);
  }
  // verifier accepts (tests/adversarial_verifier.rs:193-202)
  auto tv = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  tv.attach(proof2);
  auto res = sumcheck_verify(F64(claim), 2, v, tv);
  EXPECT(res.final_claim == proof2.final_value);
  EXPECT(res.challenges == proof2.challenges);
  // partial then continue (tests/canonical_api.rs:258-290)
  InnerProductProver<F64> p3(a, b);
  auto t3 = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  auto first = sumcheck(p3, 4, t3);
  auto second = sumcheck(p3, v - 4, t3);
  EXPECT(first.round_polys.size() == 4 && second.round_polys.size() == v - 4);
  for (size_t r = 0; r < 4; r++) EXPECT(first.round_polys[r] == proof2.round_polys[r]);
  for (size_t r = 4; r < v; r++) EXPECT(second.round_polys[r - 4] == proof2.round_polys[r]);
  EXPECT(second.final_value == proof2.final_value);
  // degree-3 product over F_p^3 with base inputs, accepted by the verifier
  std::vector<std::vector<F64>> tabs(3, std::vector<F64>(n));
  for (int k = 0; k < 3; k++)
    for (size_t i = 0; i < n; i++) tabs[k][i] = F64(uniform(0xC0FFEE10 + k, i));
  CoefficientProver<F64Ext3> p4(tabs);
  auto t4 = Transcript<F64Ext3>::fiat_shamir("effsc-hip/cpp3");
  auto proof4 = sumcheck(p4, v, t4);
  EXPECT(proof4.round_polys.size() == v && proof4.round_polys[0].size() == 3);
  print_proof("ip", proof2);

  // GKR layer with k = 4 (src/provers/gkr.rs:208-266): claim chain through the
  // verifier, claimed W values = the witness MLE at the two challenge halves
  {
    const size_t k = 4, m = size_t(1) << k;
    const unsigned __int128 P = 0xFFFFFFFF00000001ull;
    std::vector<F64> add(m * m), mult(m * m), w(m);
    for (size_t i = 0; i < m * m; i++) {
      add[i] = F64(uniform(0xC0FFEE20, i));
      mult[i] = F64(uniform(0xC0FFEE21, i));
    }
    for (size_t i = 0; i < m; i++) w[i] = F64(uniform(0xC0FFEE22, i));
    unsigned __int128 gc = 0;
    for (size_t bb = 0; bb < m; bb++)
      for (size_t cc = 0; cc < m; cc++) {
        const unsigned __int128 wb = w[bb].v, wc = w[cc].v;
        gc = (gc + (unsigned __int128)add[bb * m + cc].v * ((wb + wc) % P) % P +
              (unsigned __int128)mult[bb * m + cc].v * (wb * wc % P) % P) % P;
      }
    GkrProver<F64> g(add, mult, w);
    auto tg = Transcript<F64>::fiat_shamir("effsc-hip/cpp-gkr");
    auto pg = sumcheck(g, 2 * k, tg);
    auto tgv = Transcript<F64>::fiat_shamir("effsc-hip/cpp-gkr");
    tgv.attach(pg);
    auto rg = sumcheck_verify(F64((uint64_t)gc), 2, 2 * k, tgv);
    EXPECT(rg.final_claim == pg.final_value);
    auto mle = [&](size_t lo) {
      std::vector<unsigned __int128> t(m);
      for (size_t i = 0; i < m; i++) t[i] = w[i].v;
      for (size_t r = 0; r < k; r++) {
        const unsigned __int128 c = pg.challenges[lo + r].v;
        const size_t h = t.size() / 2;
        for (size_t j = 0; j < h; j++) t[j] = (t[j] + c * ((t[j + h] + P - t[j]) % P)) % P;
        t.resize(h);
      }
      return (uint64_t)t[0];
    };
    auto cw = g.claimed_w_values();
    EXPECT(cw.first.v == mle(0));
    EXPECT(cw.second.v == mle(k));
  }
  // LSB inner product over the same tables, printed for the oracle
  InnerProductProverLSB<F64> pl(a, b);
  auto tl = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
  auto proofl = sumcheck(pl, v, tl);
  print_proof("ip_lsb", proofl);
  // the same inner product from FileStreams equals the in-memory proof
  {
    const std::string fa = "/tmp/effsc_cpp_stream_a.bin", fb = "/tmp/effsc_cpp_stream_b.bin";
    for (auto pr : {std::make_pair(fa, &a), std::make_pair(fb, &b)}) {
      FILE* f = std::fopen(pr.first.c_str(), "wb");
      EXPECT(f != nullptr);
      if (!f) break;
      for (auto& x : *pr.second) std::fwrite(&x.v, 8, 1, f);
      std::fclose(f);
    }
    {
      FileStream sa(fa), sb(fb);
      EXPECT(sa.num_variables() == v);
      EXPECT(sa.evaluation(5) == a[5]);
      unsigned __int128 sum = 0;
      for (auto& x : a) sum = (sum + x.v) % 0xFFFFFFFF00000001ull;
      EXPECT(multivariate_claim(sa).v == (uint64_t)sum);
      StreamInnerProductProver<F64> ps(sa, sb);
      auto ts = Transcript<F64>::fiat_shamir("effsc-hip/cpp");
      auto proofs = sumcheck(ps, v, ts);
      EXPECT(proofs.round_polys == proof2.round_polys);
      EXPECT(proofs.final_value == proof2.final_value);
    }
    std::remove(fa.c_str());
    std::remove(fb.c_str());
  }
  return failures;
}

int main(int argc, char** argv) {
  std::string mode = argc > 1 ? argv[1] : "cpu";
  try {
    int f = mode == "gpu" ? gpu_tests() : cpu_tests();
    std::fprintf(stderr, "%s: %d failure(s)\n", mode.c_str(), f);
    return f ? 1 : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
}

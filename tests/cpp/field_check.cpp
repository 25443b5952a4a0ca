// CPU check of the device field arithmetic (efficient-sumcheck_amd/csrc/field.hpp,
// compiled here for the host): the Goldilocks product and reduction, the F_p^2
// and F_p^3 products in their prepared-multiplier schoolbook form, and the
// F_p^3 fold form, against independent big-integer definitions (unsigned
// __int128 remainders, the schoolbook product mod X^d - w reduced
// coefficient by coefficient, and Karatsuba's form with every operation
// reduced), on edge operands crossed with each other and on random ones.
// Prints one line per check and exits non-zero on the first mismatch.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "field.hpp"

typedef unsigned __int128 u128;
static const uint64_t P = gl::P;

static uint64_t rmul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
static uint64_t radd(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + b) % P); }
static uint64_t rsub(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + P - b) % P); }

// schoolbook product mod X^d - w, each coefficient summed exactly then reduced
static void school(int d, uint64_t w, const uint64_t* a, const uint64_t* b, uint64_t* out) {
  uint64_t prod[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) prod[i + j] = radd(prod[i + j], rmul(a[i], b[j]));
  for (int k = 2 * d - 2; k >= d; k--) prod[k - d] = radd(prod[k - d], rmul(w, prod[k]));
  for (int i = 0; i < d; i++) out[i] = prod[i];
}

// Karatsuba's F_p^3 product (6 products + 2 by w), every operation reduced
static void kara3(uint64_t w, const uint64_t* a, const uint64_t* b, uint64_t* out) {
  const uint64_t v0 = rmul(a[0], b[0]), v1 = rmul(a[1], b[1]), v2 = rmul(a[2], b[2]);
  const uint64_t t12 = rsub(rsub(rmul(radd(a[1], a[2]), radd(b[1], b[2])), v1), v2);
  const uint64_t t01 = rsub(rsub(rmul(radd(a[0], a[1]), radd(b[0], b[1])), v0), v1);
  const uint64_t t02 = rsub(rsub(rmul(radd(a[0], a[2]), radd(b[0], b[2])), v0), v2);
  out[0] = radd(v0, rmul(w, t12));
  out[1] = radd(t01, rmul(w, v2));
  out[2] = radd(t02, v1);
}

static uint64_t state = 0x243F6A8885A308D3ull;
static uint64_t rnd() {
  state += 0x9E3779B97F4A7C15ull;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z % P;
}

int main(int argc, char** argv) {
  const long randoms = argc > 1 ? atol(argv[1]) : 2000000;
  const uint64_t specials[] = {0, 1, 2, P - 1, P - 2, 0xFFFFFFFFull, 1ull << 32, (1ull << 32) + 1, 1ull << 63,
                               P >> 1, EFFSC_W, 0xFFFFFFFE00000001ull, 0x8000000080000000ull, 3};
  const int ns = sizeof(specials) / sizeof(specials[0]);
  const uint64_t ws[2] = {7, EFFSC_W};
  long checked = 0;
  auto fail = [](const char* what, long i) {
    printf("{\"name\": \"%s\", \"ok\": false, \"case\": %ld}\n", what, i);
    exit(1);
  };
  // base field: product, reduction of arbitrary 128-bit values, add, sub
  for (long i = 0; i < ns * ns + randoms; i++) {
    uint64_t a, b;
    if (i < ns * ns) { a = specials[i / ns]; b = specials[i % ns]; } else { a = rnd(); b = rnd(); }
    if (gl::mul(a, b) != rmul(a, b)) fail("mul", i);
    if (gl::add(a, b) != radd(a, b) || gl::sub(a, b) != rsub(a, b)) fail("add_sub", i);
    const uint64_t lo = i < ns * ns ? ~a : rnd() * 0x9E3779B97F4A7C15ull, hi = i < ns * ns ? ~b : rnd() << 1;
    if (gl::reduce128(lo, hi) != (uint64_t)((((u128)hi << 64) | lo) % P)) fail("reduce128", i);
    checked++;
  }
  printf("{\"name\": \"base\", \"ok\": true, \"cases\": %ld}\n", checked);
  // extensions: every (w, a, b) with components from the edge values, then random
  for (int d = 2; d <= 3; d++) {
    long n = 0;
    for (int wi = 0; wi < 2; wi++) {
      const uint64_t w = ws[wi];
      for (long i = 0; i < (long)ns * ns * ns + randoms / 4; i++) {
        uint64_t a[3], b[3];
        if (i < (long)ns * ns * ns) {
          a[0] = specials[i % ns]; a[1] = specials[(i / ns) % ns]; a[2] = specials[(i / ns / ns) % ns];
          b[0] = specials[(i / 7) % ns]; b[1] = specials[(i * 5) % ns]; b[2] = specials[(i / 3) % ns];
        } else {
          for (int c = 0; c < 3; c++) { a[c] = rnd(); b[c] = rnd(); }
        }
        uint64_t ref[3], got[3] = {0, 0, 0};
        school(d, w, a, b, ref);
        if (d == 2) {
          gl::store_words(gl::mul(gl::Fp2{a[0], a[1]}, gl::Fp2{b[0], b[1]}, w), got);
          if (got[0] != ref[0] || got[1] != ref[1]) fail("fp2_mul", i);
          // inverse: a a^-1 = 1 for nonzero a
          if (a[0] | a[1]) {
            uint64_t one[3] = {0, 0, 0};
            gl::store_words(gl::mul(gl::Fp2{a[0], a[1]}, gl::inv(gl::Fp2{a[0], a[1]}, w), w), one);
            if (one[0] != 1 || one[1] != 0) fail("fp2_inv", i);
          }
        } else {
          gl::store_words(gl::mul(gl::Fp3{a[0], a[1], a[2]}, gl::Fp3{b[0], b[1], b[2]}, w), got);
          uint64_t k[3];
          kara3(w, a, b, k);
          if (got[0] != ref[0] || got[1] != ref[1] || got[2] != ref[2]) fail("fp3_mul", i);
          if (k[0] != ref[0] || k[1] != ref[1] || k[2] != ref[2]) fail("fp3_karatsuba", i);
          // fold form: c (hi - lo) + lo with c prepared once
          uint64_t f[3], diff[3] = {rsub(b[0], a[0]), rsub(b[1], a[1]), rsub(b[2], a[2])}, cd[3];
          const gl::Fp3 c{b[2], a[1], b[0]};
          const uint64_t cw[3] = {b[2], a[1], b[0]};
          gl::store_words(gl::add(gl::mul(gl::Fp3{diff[0], diff[1], diff[2]}, gl::prepare(c, w)),
                                  gl::Fp3{a[0], a[1], a[2]}), f);
          school(3, w, cw, diff, cd);
          if (f[0] != radd(cd[0], a[0]) || f[1] != radd(cd[1], a[1]) || f[2] != radd(cd[2], a[2])) fail("fp3_fold", i);
        }
        n++;
      }
    }
    printf("{\"name\": \"fp%d\", \"ok\": true, \"cases\": %ld}\n", d, n);
  }
  // lazy Karatsuba dot products (gl::KDot): sums of 1..4096 terms a_i r_i
  // against the sum of schoolbook products, edge-valued and random terms
  for (int d = 2; d <= 3; d++) {
    long n = 0;
    for (int wi = 0; wi < 2; wi++) {
      const uint64_t w = ws[wi];
      for (long t = 0; t < 200 + randoms / 20000; t++) {
        const long terms = t < 64 ? 1 + t : 1 + (long)(rnd() % 4096);
        uint64_t ref[3] = {0, 0, 0}, got[3] = {0, 0, 0};
        gl::KDot<gl::Fp2> k2;
        gl::KDot<gl::Fp3> k3;
        k2.zero();
        k3.zero();
        for (long i = 0; i < terms; i++) {
          uint64_t a[3], r[3], pr[3];
          for (int c = 0; c < 3; c++) {
            // every third term takes edge components (p - 1 maximises the carries)
            a[c] = (t + i) % 3 == 0 ? specials[(t * 7 + i * 3 + c) % ns] : rnd();
            r[c] = (t + i) % 3 == 1 ? specials[(t * 5 + i + c * 11) % ns] : rnd();
          }
          school(d, w, a, r, pr);
          for (int c = 0; c < d; c++) ref[c] = radd(ref[c], pr[c]);
          uint64_t rp[6];
          if (d == 2) {
            gl::kdot_prep(gl::Fp2{r[0], r[1]}, rp);
            k2.add_term(gl::Fp2{a[0], a[1]}, rp);
          } else {
            gl::kdot_prep(gl::Fp3{r[0], r[1], r[2]}, rp);
            k3.add_term(gl::Fp3{a[0], a[1], a[2]}, rp);
          }
        }
        if (d == 2)
          gl::store_words(k2.value(w), got);
        else
          gl::store_words(k3.value(w), got);
        for (int c = 0; c < d; c++)
          if (got[c] != ref[c]) fail(d == 2 ? "kdot2" : "kdot3", t);
        n++;
      }
    }
    printf("{\"name\": \"kdot%d\", \"ok\": true, \"sums\": %ld}\n", d, n);
  }
  // two pending folds in one lazy sum (gl::fold2) against the nested
  // schoolbook folds lo + c (hi - lo)
  for (int d = 2; d <= 3; d++) {
    long n = 0;
    for (int wi = 0; wi < 2; wi++) {
      const uint64_t w = ws[wi];
      for (long i = 0; i < (long)ns * ns + randoms / 8; i++) {
        uint64_t x[4][3], c1[3], c2[3];
        for (int c = 0; c < 3; c++) {
          for (int k = 0; k < 4; k++) x[k][c] = i < ns * ns && (k + c) % 2 ? specials[(i + 3 * k + c) % ns] : rnd();
          c1[c] = i < ns * ns ? specials[(i / ns + c) % ns] : rnd();
          c2[c] = i < ns * ns ? specials[(i % ns + 2 * c) % ns] : rnd();
        }
        auto fold = [&](const uint64_t* lo, const uint64_t* hi, const uint64_t* c, uint64_t* out) {
          uint64_t df[3], pr[3];
          for (int k = 0; k < 3; k++) df[k] = rsub(hi[k], lo[k]);
          school(d, w, c, df, pr);
          for (int k = 0; k < d; k++) out[k] = radd(pr[k], lo[k]);
        };
        uint64_t a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, ref[3] = {0, 0, 0}, got[3] = {0, 0, 0};
        fold(x[0], x[2], c1, a);
        fold(x[1], x[3], c1, b);
        fold(a, b, c2, ref);
        if (d == 2) {
          auto e = [&](int k) { return gl::Fp2{x[k][0], x[k][1]}; };
          const auto p = gl::prepare_fold2(gl::Fp2{c1[0], c1[1]}, gl::Fp2{c2[0], c2[1]}, w);
          gl::store_words(gl::fold2(e(0), e(1), e(2), e(3), p, w), got);
        } else {
          auto e = [&](int k) { return gl::Fp3{x[k][0], x[k][1], x[k][2]}; };
          const auto p = gl::prepare_fold2(gl::Fp3{c1[0], c1[1], c1[2]}, gl::Fp3{c2[0], c2[1], c2[2]}, w);
          gl::store_words(gl::fold2(e(0), e(1), e(2), e(3), p, w), got);
        }
        for (int c = 0; c < d; c++)
          if (got[c] != ref[c]) fail(d == 2 ? "fold2_fp2" : "fold2_fp3", i);
        // base-field x0..x3 (first fold of base inputs): the same folds with
        // x lifted to (x, 0, 0)
        uint64_t xb[4][3] = {{x[0][0], 0, 0}, {x[1][0], 0, 0}, {x[2][0], 0, 0}, {x[3][0], 0, 0}};
        fold(xb[0], xb[2], c1, a);
        fold(xb[1], xb[3], c1, b);
        fold(a, b, c2, ref);
        auto eb = [&](int k) { return gl::Fp{x[k][0]}; };
        if (d == 2) {
          const auto p = gl::prepare_fold2(gl::Fp2{c1[0], c1[1]}, gl::Fp2{c2[0], c2[1]}, w);
          gl::store_words(gl::fold2(eb(0), eb(1), eb(2), eb(3), p, w), got);
        } else {
          const auto p = gl::prepare_fold2(gl::Fp3{c1[0], c1[1], c1[2]}, gl::Fp3{c2[0], c2[1], c2[2]}, w);
          gl::store_words(gl::fold2(eb(0), eb(1), eb(2), eb(3), p, w), got);
        }
        for (int c = 0; c < d; c++)
          if (got[c] != ref[c]) fail(d == 2 ? "fold2_base_fp2" : "fold2_base_fp3", i);
        n++;
      }
    }
    printf("{\"name\": \"fold2_fp%d\", \"ok\": true, \"cases\": %ld}\n", d, n);
  }
  // base field: two pending folds in one 160-bit sum (gl::fold2 with
  // c12 = c1 c2) against the nested folds; the 160-bit sum of a word and up
  // to three products (gl::Acc3) against the exact sum; the column sums
  // (gl::AccW) of up to 4096 products and words against the exact sum
  {
    long n = 0;
    for (long i = 0; i < (long)ns * ns * ns + randoms / 4; i++) {
      uint64_t x[4], c1, c2;
      if (i < (long)ns * ns * ns) {
        for (int k = 0; k < 4; k++) x[k] = specials[(i * (k + 1) + k) % ns];
        c1 = specials[i % ns];
        c2 = specials[(i / ns) % ns];
        if ((i / ns / ns) % 2) x[3] = specials[(i / ns / ns) % ns];
      } else {
        for (int k = 0; k < 4; k++) x[k] = rnd();
        c1 = rnd();
        c2 = rnd();
      }
      const uint64_t a = radd(x[0], rmul(c1, rsub(x[2], x[0]))), b = radd(x[1], rmul(c1, rsub(x[3], x[1])));
      const uint64_t ref = radd(a, rmul(c2, rsub(b, a)));
      if (gl::fold2(x[0], x[1], x[2], x[3], c1, c2, gl::mul(c1, c2)) != ref) fail("fold2_base", i);
      // x0 + x1 x2 + x3 c1 + c2 c2 as one Acc3
      gl::Acc3 s = gl::wide(x[0]);
      gl::mac(s, x[1], x[2]);
      gl::mac(s, x[3], c1);
      gl::mac(s, c2, c2);
      const uint64_t sref = radd(radd(x[0], rmul(x[1], x[2])), radd(rmul(x[3], c1), rmul(c2, c2)));
      if (gl::reduce(s) != sref) fail("acc3_sum", i);
      n++;
    }
    for (long t = 0; t < 200 + randoms / 20000; t++) {
      const long terms = t < 64 ? 1 + t : 1 + (long)(rnd() % 4096);
      gl::AccW acc;
      gl::zero(acc);
      uint64_t ref = 0;
      for (long i = 0; i < terms; i++) {
        // edge operands on most terms: p - 1 maximises every column's carries
        const uint64_t x = (t + i) % 4 ? specials[(t * 3 + i) % ns] : rnd();
        const uint64_t y = (t + i) % 4 != 1 ? specials[(t + i * 5) % ns] : rnd();
        gl::macw(acc, x, y);
        ref = radd(ref, rmul(x, y));
        if (i % 3 == 0) {
          gl::addw(acc, x);
          ref = radd(ref, x);
        }
      }
      if (gl::reduce(acc) != ref) fail("accw_sum", t);
      n++;
    }
    printf("{\"name\": \"fold2_base\", \"ok\": true, \"cases\": %ld}\n", n);
  }
  return 0;
}

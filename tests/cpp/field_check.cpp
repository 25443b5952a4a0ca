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
  return 0;
}

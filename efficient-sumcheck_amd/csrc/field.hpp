// Goldilocks field p = 2^64 - 2^32 + 1 and its quadratic and cubic extensions
// (F_p[X]/(X^2 - w), F_p[X]/(X^3 - w), src/tests/fields.rs:41-100), canonical form,
// shared by the gfx950 kernels and the host protocol code.
//
// Every stored value is canonical (in [0, p)).  The reference keeps arkworks
// Montgomery words (src/simd_fields/goldilocks/avx512.rs:1-35); words are
// converted at the C-ABI boundary when the caller asks for that
// representation, so the arithmetic here never sees Montgomery form.
//
// Multiplication: a 64x64 -> 128 product from four 32x32 -> 64 partial
// products (v_mad_u64_u32 on gfx950), then the Goldilocks reduction
// 2^64 = 2^32 - 1, 2^96 = -1 (mod p): x = lo + 2^64 (hl + 2^32 hh) is
// congruent to lo - hh + hl (2^32 - 1).  No Montgomery step, no division.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GLHD __host__ __device__ __forceinline__

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p
// 2^-64 mod p = -2^32 mod p: multiplying an arkworks Montgomery word by this
// constant yields the canonical value.
constexpr uint64_t MONT_TO_CANON = 0xFFFFFFFE00000001ull;
// 2^64 mod p: canonical -> Montgomery word.
constexpr uint64_t CANON_TO_MONT = EPS;

GLHD uint64_t add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  uint64_t r = (s < a) ? s + EPS : s;  // wrapped: a + b - 2^64 + eps = a + b - p
  return (r >= P) ? r - P : r;
}

GLHD uint64_t sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  return (a < b) ? d + P : d;  // wrapping add of p
}

GLHD uint64_t neg(uint64_t a) { return a ? P - a : 0; }

GLHD void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  uint64_t p00 = (uint64_t)a0 * b0;
  uint64_t p01 = (uint64_t)a0 * b1;
  uint64_t p10 = (uint64_t)a1 * b0;
  uint64_t p11 = (uint64_t)a1 * b1;
  // (2^32-1)^2 + 2 (2^32-1) = 2^64 - 1: no overflow
  uint64_t mid = p01 + (p00 >> 32) + (uint32_t)p10;
  lo = (mid << 32) | (uint32_t)p00;
  hi = p11 + (mid >> 32) + (p10 >> 32);
}

// Reduce lo + 2^64 hi (any 128-bit value) to [0, p).
GLHD uint64_t reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hh = hi >> 32;
  uint64_t hl = hi & EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= EPS;           // borrow: lo - hh + p
  uint64_t t1 = (hl << 32) - hl;    // hl * (2^32 - 1) < 2^64
  uint64_t r = t0 + t1;
  if (r < t1) r += EPS;             // carry: cannot overflow again
  return (r >= P) ? r - P : r;
}

GLHD uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  return reduce128(lo, hi);
}

// Any u64 word to its canonical residue.
GLHD uint64_t canon(uint64_t a) { return a >= P ? a - P : a; }

GLHD uint64_t pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return r;
}

GLHD uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// ---------------------------------------------------------------------------
// Field element types.  Fp: base field.  Fp2: F_p[X]/(X^2 - w).  Fp3:
// F_p[X]/(X^3 - w).  w is a runtime value (the reference's default is
// F64::from_raw(7) for both extensions, src/tests/fields.rs:49,80).
// ---------------------------------------------------------------------------
struct Fp {
  uint64_t v;
  static constexpr int D = 1;
};

struct Fp2 {
  uint64_t c0, c1;
  static constexpr int D = 2;
};

struct Fp3 {
  uint64_t c0, c1, c2;
  static constexpr int D = 3;
};

GLHD Fp add(Fp a, Fp b) { return {add(a.v, b.v)}; }
GLHD Fp sub(Fp a, Fp b) { return {sub(a.v, b.v)}; }
GLHD Fp mul(Fp a, Fp b, uint64_t) { return {mul(a.v, b.v)}; }
GLHD Fp2 add(Fp2 a, Fp2 b) { return {add(a.c0, b.c0), add(a.c1, b.c1)}; }
GLHD Fp2 sub(Fp2 a, Fp2 b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1)}; }
GLHD Fp3 add(Fp3 a, Fp3 b) { return {add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
GLHD Fp3 sub(Fp3 a, Fp3 b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }

// Sum of up to three 64x64 products without intermediate reduction:
// value = lo + 2^64 hi + 2^128 top, top <= 2.
struct Acc3 {
  uint64_t lo, hi;
  uint32_t top;
};
GLHD Acc3 prod(uint64_t x, uint64_t y) {
  Acc3 a;
  mul128(x, y, a.lo, a.hi);
  a.top = 0;
  return a;
}
GLHD void mac(Acc3& a, uint64_t x, uint64_t y) {
  uint64_t plo, phi;
  mul128(x, y, plo, phi);
  uint64_t s = a.lo + plo;
  uint64_t c = s < plo;
  a.lo = s;
  uint64_t h = a.hi + phi;
  uint32_t t = h < phi;
  uint64_t h2 = h + c;
  t += h2 < h;
  a.hi = h2;
  a.top += t;
}
// 2^128 = -2^32 (mod p)
GLHD uint64_t reduce(const Acc3& a) { return sub(reduce128(a.lo, a.hi), (uint64_t)a.top << 32); }

// A multiplier prepared for products mod X^2 - w: b and w b1.
struct Fp2Pre {
  uint64_t b0, b1, wb1;
};
GLHD Fp2Pre prepare(Fp2 b, uint64_t w) { return {b.c0, b.c1, mul(w, b.c1)}; }

// (a0 + a1 X)(b0 + b1 X) mod X^2 - w = (a0 b0 + a1 w b1) + (a0 b1 + a1 b0) X,
// schoolbook with one reduction per coefficient: 4 products and 2
// reductions (each coefficient a sum of two unreduced products < 2^129).
GLHD Fp2 mul(Fp2 a, const Fp2Pre& b) {
  Acc3 r0 = prod(a.c0, b.b0), r1 = prod(a.c0, b.b1);
  mac(r0, a.c1, b.wb1);
  mac(r1, a.c1, b.b0);
  return {reduce(r0), reduce(r1)};
}

// General product: the first operand is prepared (1 product by w), then the
// prepared form; a loop-invariant first operand (a fold's challenge) is
// prepared once, outside the loop.
GLHD Fp2 mul(Fp2 a, Fp2 b, uint64_t w) { return mul(b, prepare(a, w)); }

// A multiplier prepared for products mod X^3 - w: b and w b1, w b2.
struct Fp3Pre {
  uint64_t b0, b1, b2, wb1, wb2;
};
GLHD Fp3Pre prepare(Fp3 b, uint64_t w) { return {b.c0, b.c1, b.c2, mul(w, b.c1), mul(w, b.c2)}; }

// (a0 + a1 X + a2 X^2)(b0 + b1 X + b2 X^2) mod X^3 - w, schoolbook with one
// reduction per coefficient: 9 products and 3 reductions.
GLHD Fp3 mul(Fp3 a, const Fp3Pre& b) {
  Acc3 r0 = prod(a.c0, b.b0), r1 = prod(a.c0, b.b1), r2 = prod(a.c0, b.b2);
  mac(r0, a.c1, b.wb2);
  mac(r0, a.c2, b.wb1);
  mac(r1, a.c1, b.b0);
  mac(r1, a.c2, b.wb2);
  mac(r2, a.c1, b.b1);
  mac(r2, a.c2, b.b0);
  return {reduce(r0), reduce(r1), reduce(r2)};
}

// General product: the first operand is prepared (2 products by w), then the
// prepared form; 11 products and 5 reductions against Karatsuba's 8 full
// products and 15 additions.  A loop-invariant first operand (a fold's
// challenge) is prepared once, outside the loop.
GLHD Fp3 mul(Fp3 a, Fp3 b, uint64_t w) { return mul(b, prepare(a, w)); }

// Mixed products: extension element times base element.
GLHD Fp2 mul(Fp2 a, Fp b, uint64_t) { return {mul(a.c0, b.v), mul(a.c1, b.v)}; }
GLHD Fp2 mul(Fp a, Fp2 b, uint64_t) { return {mul(a.v, b.c0), mul(a.v, b.c1)}; }
GLHD Fp3 mul(Fp3 a, Fp b, uint64_t) { return {mul(a.c0, b.v), mul(a.c1, b.v), mul(a.c2, b.v)}; }
GLHD Fp3 mul(Fp a, Fp3 b, uint64_t) { return {mul(a.v, b.c0), mul(a.v, b.c1), mul(a.v, b.c2)}; }

// Lift base -> extension.
template <class F>
GLHD F lift(Fp x);
template <>
GLHD Fp lift<Fp>(Fp x) { return x; }
template <>
GLHD Fp2 lift<Fp2>(Fp x) { return {x.v, 0}; }
template <>
GLHD Fp3 lift<Fp3>(Fp x) { return {x.v, 0, 0}; }

// Lift into a (possibly) larger field.
GLHD void lift_into(Fp x, Fp& o) { o = x; }
GLHD void lift_into(Fp x, Fp2& o) { o = {x.v, 0}; }
GLHD void lift_into(Fp2 x, Fp2& o) { o = x; }
GLHD void lift_into(Fp x, Fp3& o) { o = {x.v, 0, 0}; }
GLHD void lift_into(Fp3 x, Fp3& o) { o = x; }
template <class FO, class FI>
GLHD FO lift_to(FI x) {
  FO o;
  lift_into(x, o);
  return o;
}

template <class F>
GLHD F zero();
template <>
GLHD Fp zero<Fp>() { return {0}; }
template <>
GLHD Fp2 zero<Fp2>() { return {0, 0}; }
template <>
GLHD Fp3 zero<Fp3>() { return {0, 0, 0}; }

template <class F>
GLHD F one();
template <>
GLHD Fp one<Fp>() { return {1}; }
template <>
GLHD Fp2 one<Fp2>() { return {1, 0}; }
template <>
GLHD Fp3 one<Fp3>() { return {1, 0, 0}; }

GLHD bool is_zero(Fp a) { return a.v == 0; }
GLHD bool is_zero(Fp2 a) { return (a.c0 | a.c1) == 0; }
GLHD bool is_zero(Fp3 a) { return (a.c0 | a.c1 | a.c2) == 0; }

GLHD Fp inv(Fp a, uint64_t) { return {a.v ? inv(a.v) : 0}; }
// (a0 + a1 X)^-1 = (a0 - a1 X) / (a0^2 - w a1^2), the norm being nonzero for
// a nonzero element when w is a non-residue.
GLHD Fp2 inv(Fp2 a, uint64_t w) {
  const uint64_t n = sub(mul(a.c0, a.c0), mul(w, mul(a.c1, a.c1)));
  const uint64_t ni = n ? inv(n) : 0;
  return {mul(a.c0, ni), mul(neg(a.c1), ni)};
}
// Inverse through the adjugate of the multiplication matrix.
GLHD Fp3 inv(Fp3 a, uint64_t w) {
  uint64_t t0 = sub(mul(a.c0, a.c0), mul(w, mul(a.c1, a.c2)));
  uint64_t t1 = sub(mul(w, mul(a.c2, a.c2)), mul(a.c0, a.c1));
  uint64_t t2 = sub(mul(a.c1, a.c1), mul(a.c0, a.c2));
  uint64_t n = add(mul(a.c0, t0), mul(w, add(mul(a.c2, t1), mul(a.c1, t2))));
  uint64_t ni = n ? inv(n) : 0;
  return {mul(t0, ni), mul(t1, ni), mul(t2, ni)};
}

// Component access (element -> words).
GLHD void store_words(Fp a, uint64_t* w) { w[0] = a.v; }
GLHD void store_words(Fp2 a, uint64_t* w) {
  w[0] = a.c0;
  w[1] = a.c1;
}
GLHD void store_words(Fp3 a, uint64_t* w) {
  w[0] = a.c0;
  w[1] = a.c1;
  w[2] = a.c2;
}
template <class F>
GLHD F load_words(const uint64_t* w);
template <>
GLHD Fp load_words<Fp>(const uint64_t* w) { return {canon(w[0])}; }
template <>
GLHD Fp2 load_words<Fp2>(const uint64_t* w) { return {canon(w[0]), canon(w[1])}; }
template <>
GLHD Fp3 load_words<Fp3>(const uint64_t* w) { return {canon(w[0]), canon(w[1]), canon(w[2])}; }

}  // namespace gl

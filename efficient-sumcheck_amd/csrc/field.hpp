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

// A sum of a few 64x64 products and words, reduced once:
//   value = w0 + 2^32 w1 + 2^64 w2 + 2^96 w3 + 2^128 top.
// Adding a 128-bit product is one carry chain of five 32-bit adds on gfx950
// (the compiler's 64-bit form compares for every carry: 24 against 15
// instructions per product).  The reduction treats H = top:w3 as one word:
// with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p), 2^96 H + 2^64 w2 is congruent
// to w2 (2^32 - 1) - H, which is reduce128's step with hh widened to 64 bits;
// it stays exact while H <= p, i.e. top <= 2^32 - 2.
struct Acc3 {
  uint32_t w0, w1, w2, w3, top;
};
GLHD Acc3 wide(uint64_t x) { return {(uint32_t)x, (uint32_t)(x >> 32), 0, 0, 0}; }
GLHD void add_wide(Acc3& a, uint64_t lo, uint64_t hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  // a carry written to VCC is read two wait states later (s_nop 1), as the
  // compiler schedules its own add-with-carry pairs on gfx950
  asm("v_add_co_u32 %0, vcc, %0, %5\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %1, vcc, %1, %6, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %2, vcc, %2, %7, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %3, vcc, %3, %8, vcc\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %4, vcc, %4, 0, vcc"
      : "+v"(a.w0), "+v"(a.w1), "+v"(a.w2), "+v"(a.w3), "+v"(a.top)
      : "v"((uint32_t)lo), "v"((uint32_t)(lo >> 32)), "v"((uint32_t)hi), "v"((uint32_t)(hi >> 32))
      : "vcc");
#else
  const uint64_t alo = ((uint64_t)a.w1 << 32) | a.w0, ahi = ((uint64_t)a.w3 << 32) | a.w2;
  const uint64_t s = alo + lo;
  const uint64_t c = s < lo;
  const uint64_t h = ahi + hi;
  uint32_t t = h < hi;
  const uint64_t h2 = h + c;
  t += h2 < h;
  a = {(uint32_t)s, (uint32_t)(s >> 32), (uint32_t)h2, (uint32_t)(h2 >> 32), a.top + t};
#endif
}
GLHD Acc3 prod(uint64_t x, uint64_t y) {
  uint64_t lo, hi;
  mul128(x, y, lo, hi);
  return {(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32), 0};
}
GLHD void mac(Acc3& a, uint64_t x, uint64_t y) {
  uint64_t lo, hi;
  mul128(x, y, lo, hi);
  add_wide(a, lo, hi);
}
GLHD uint64_t reduce(const Acc3& a) {
  const uint64_t lo = ((uint64_t)a.w1 << 32) | a.w0;
  const uint64_t H = ((uint64_t)a.top << 32) | a.w3;
  const uint64_t hl = a.w2;
  uint64_t t0 = lo - H;
  if (lo < H) t0 -= EPS;             // borrow: lo - H + p (>= 0 since H <= p)
  const uint64_t t1 = (hl << 32) - hl;
  uint64_t r = t0 + t1;
  if (r < t1) r += EPS;              // carry: cannot overflow again
  return (r >= P) ? r - P : r;
}

// Two pending folds of the half-split table in one reduction: the element
// of the twice-folded table from x0 = T[e], x1 = T[e + L], x2 = T[e + 2L],
// x3 = T[e + 3L] (first fold c1 over the top variable, then c2) is
//   a + c2 (b - a), a = x0 + c1 (x2 - x0), b = x1 + c1 (x3 - x1)
//   = x0 + c1 (x2 - x0) + c2 (x1 - x0) + c1 c2 ((x3 - x1) - (x2 - x0)),
// three products summed unreduced (c12 = c1 c2 is prepared once).
GLHD uint64_t fold2(uint64_t x0, uint64_t x1, uint64_t x2, uint64_t x3, uint64_t c1, uint64_t c2, uint64_t c12) {
  const uint64_t d20 = sub(x2, x0), d10 = sub(x1, x0), dd = sub(sub(x3, x1), d20);
  Acc3 s = wide(x0);
  mac(s, c1, d20);
  mac(s, c2, d10);
  mac(s, c12, dd);
  return reduce(s);
}

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

// a + b as a word congruent to a + b (mod p), not necessarily below p (a, b
// in [0, p)): a product operand only.  A wrapped sum a + b - 2^64 + eps stays
// below 2^64 - eps.
GLHD uint64_t add_lazy(uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return s < a ? s + EPS : s;
}

// A sum of 64 x 64-bit products kept in 32-bit columns with carry counts:
//   value = s0 + 2^32 s1 + 2^64 (s2 + k0) + 2^96 k1 + 2^128 k2,
// s0 += x0 y0, s1 += x0 y1 + x1 y0, s2 += x1 y1 (xi, yi the 32-bit halves)
// and each k counts the carries out of its column.  On gfx950 one product is
// four v_mad_u64_u32 with the column as addend and carry out, and four
// v_addc_co_u32 into the counts; at most 2^31 products per sum.
struct AccW {
  uint64_t s0, s1, s2;
  uint32_t k0, k1, k2;
};
GLHD void zero(AccW& a) { a = {0, 0, 0, 0, 0, 0}; }
// a word x into the low column
GLHD void addw(AccW& a, uint64_t x) {
  a.s0 += x;
  a.k0 += a.s0 < x;
}
GLHD void macw(AccW& a, uint64_t x, uint64_t y) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32), y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
  // every carry is read three instructions after the mad that writes it
  uint64_t c0, c1, c2, c3;
  asm("v_mad_u64_u32 %0, %6, %10, %12, %0\n\t"
      "v_mad_u64_u32 %1, %7, %10, %13, %1\n\t"
      "v_mad_u64_u32 %1, %8, %11, %12, %1\n\t"
      "v_mad_u64_u32 %2, %9, %11, %13, %2\n\t"
      "v_addc_co_u32_e64 %3, %6, %3, 0, %6\n\t"
      "v_addc_co_u32_e64 %4, %7, %4, 0, %7\n\t"
      "v_addc_co_u32_e64 %4, %8, %4, 0, %8\n\t"
      "v_addc_co_u32_e64 %5, %9, %5, 0, %9"
      : "+v"(a.s0), "+v"(a.s1), "+v"(a.s2), "+v"(a.k0), "+v"(a.k1), "+v"(a.k2), "=&s"(c0), "=&s"(c1), "=&s"(c2),
        "=&s"(c3)
      : "v"(x0), "v"(x1), "v"(y0), "v"(y1));
#else
  auto col = [](uint64_t& s, uint32_t& k, uint64_t p) {
    s += p;
    k += s < p;
  };
  col(a.s0, a.k0, (uint64_t)x0 * y0);
  col(a.s1, a.k1, (uint64_t)x0 * y1);
  col(a.s1, a.k1, (uint64_t)x1 * y0);
  col(a.s2, a.k2, (uint64_t)x1 * y1);
#endif
}
// 2^128 = -2^32 (mod p)
GLHD uint64_t reduce(const AccW& a) {
  typedef unsigned __int128 u128;
  const uint64_t sl = a.s1 << 32, lo = a.s0 + sl;
  const u128 mid = (u128)a.s2 + a.k0 + (a.s1 >> 32) + (uint64_t)(lo < sl) + ((u128)a.k1 << 32);  // < 2^66
  return sub(reduce128(lo, (uint64_t)mid), ((uint64_t)a.k2 + (uint64_t)(mid >> 64)) << 32);
}

// a b accumulated coefficient-wise into r (b prepared: mod X^d - w).
GLHD void mac_pre(AccW* r, Fp2 a, const Fp2Pre& b) {
  macw(r[0], a.c0, b.b0);
  macw(r[0], a.c1, b.wb1);
  macw(r[1], a.c0, b.b1);
  macw(r[1], a.c1, b.b0);
}
GLHD void mac_pre(AccW* r, Fp3 a, const Fp3Pre& b) {
  macw(r[0], a.c0, b.b0);
  macw(r[0], a.c1, b.wb2);
  macw(r[0], a.c2, b.wb1);
  macw(r[1], a.c0, b.b1);
  macw(r[1], a.c1, b.b0);
  macw(r[1], a.c2, b.wb2);
  macw(r[2], a.c0, b.b2);
  macw(r[2], a.c1, b.b1);
  macw(r[2], a.c2, b.b0);
}

// Sums of products a r (a, r in F_p^2 or F_p^3) by Karatsuba without any
// reduction in the sum: each partial-product class accumulates in its own
// column sum (AccW) and the classes are combined once, in value().  r comes
// prepared (kdot_prep): its coefficients and their pairwise sums.
//   F_p^2: m0 = a0 r0, m1 = a1 r1, m2 = (a0 + a1)(r0 + r1);
//          a r = (m0 + w m1) + (m2 - m0 - m1) X.
//   F_p^3: m0..m2 = ai ri, m3 = (a0 + a1)(r0 + r1), m4 = (a0 + a2)(r0 + r2),
//          m5 = (a1 + a2)(r1 + r2); with s1 = m3 - m0 - m1,
//          s2 = m4 - m0 - m2 + m1, s3 = m5 - m1 - m2:
//          a r = (m0 + w s3) + (s1 + w m2) X + s2 X^2.
// 3 (F_p^2) or 6 (F_p^3) products per term against 4 or 9 in the schoolbook
// form.
template <class F>
struct KDot;
template <>
struct KDot<Fp2> {
  static constexpr int R = 3;  // prepared words per multiplier
  AccW m[3];
  GLHD void zero() {
#pragma unroll
    for (int i = 0; i < 3; i++) gl::zero(m[i]);
  }
  GLHD void add_term(Fp2 a, const uint64_t* r) {
    macw(m[0], a.c0, r[0]);
    macw(m[1], a.c1, r[1]);
    macw(m[2], add_lazy(a.c0, a.c1), r[2]);
  }
  GLHD Fp2 value(uint64_t w) const {
    const uint64_t m0 = gl::reduce(m[0]), m1 = gl::reduce(m[1]), m2 = gl::reduce(m[2]);
    return {add(m0, mul(w, m1)), sub(sub(m2, m0), m1)};
  }
};
template <>
struct KDot<Fp3> {
  static constexpr int R = 6;
  AccW m[6];
  GLHD void zero() {
#pragma unroll
    for (int i = 0; i < 6; i++) gl::zero(m[i]);
  }
  GLHD void add_term(Fp3 a, const uint64_t* r) {
    macw(m[0], a.c0, r[0]);
    macw(m[1], a.c1, r[1]);
    macw(m[2], a.c2, r[2]);
    macw(m[3], add_lazy(a.c0, a.c1), r[3]);
    macw(m[4], add_lazy(a.c0, a.c2), r[4]);
    macw(m[5], add_lazy(a.c1, a.c2), r[5]);
  }
  GLHD Fp3 value(uint64_t w) const {
    uint64_t v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) v[i] = gl::reduce(m[i]);
    const uint64_t s1 = sub(sub(v[3], v[0]), v[1]);
    const uint64_t s2 = add(sub(sub(v[4], v[0]), v[2]), v[1]);
    const uint64_t s3 = sub(sub(v[5], v[1]), v[2]);
    return {add(v[0], mul(w, s3)), add(s1, mul(w, v[2])), s2};
  }
};
GLHD void kdot_prep(Fp2 r, uint64_t* o) {
  o[0] = r.c0;
  o[1] = r.c1;
  o[2] = add(r.c0, r.c1);
}
GLHD void kdot_prep(Fp3 r, uint64_t* o) {
  o[0] = r.c0;
  o[1] = r.c1;
  o[2] = r.c2;
  o[3] = add(r.c0, r.c1);
  o[4] = add(r.c0, r.c2);
  o[5] = add(r.c1, r.c2);
}

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

// Two pending folds of four elements x0..x3 (x_k at offset k L of the level
// 4L table) with challenges c1 (the older) and c2:
//   fold(fold(x0, x2, c1), fold(x1, x3, c1), c2)
//     = x0 + c1 (x2 - x0) + c2 (x1 - x0) + c1 c2 (x3 - x1 - x2 + x0),
// three products by prepared multipliers (c1, c2, c1 c2) summed in one
// column sum per coefficient and reduced once.  (The form with four
// constants, x0 (1 - c1 - c2 + c1 c2) + ..., saves the three subtractions but
// holds a fourth prepared constant: the F_p^2 inner-product pass then fell
// from three to two waves per SIMD and ran 5% slower.)
template <class F>
struct Fold2Pre {
  decltype(prepare(F{}, uint64_t{})) e[3];
};
template <class F>
GLHD Fold2Pre<F> prepare_fold2(F c1, F c2, uint64_t w) {
  Fold2Pre<F> p;
  p.e[0] = prepare(c1, w);
  p.e[1] = prepare(c2, w);
  p.e[2] = prepare(mul(c1, c2, w), w);
  return p;
}
template <class F>
GLHD F fold2(F x0, F x1, F x2, F x3, const Fold2Pre<F>& p, uint64_t) {
  const F d20 = sub(x2, x0), d10 = sub(x1, x0), dd = sub(sub(x3, x1), d20);
  AccW r[F::D];
  uint64_t x0w[3];
  store_words(x0, x0w);
#pragma unroll
  for (int i = 0; i < F::D; i++) {
    zero(r[i]);
    addw(r[i], x0w[i]);
  }
  mac_pre(r, d20, p.e[0]);
  mac_pre(r, d10, p.e[1]);
  mac_pre(r, dd, p.e[2]);
  uint64_t wd[3];
#pragma unroll
  for (int i = 0; i < F::D; i++) wd[i] = reduce(r[i]);
  return load_words<F>(wd);
}
// The same with base-field x0..x3 (a first fold of base inputs under
// extension challenges): the differences are base elements, so coefficient i
// is x0 (i = 0) plus the column sum of the differences times coefficient i of
// c1, c2, c1 c2.
GLHD uint64_t pre_coeff(const Fp2Pre& b, int i) { return i == 0 ? b.b0 : b.b1; }
GLHD uint64_t pre_coeff(const Fp3Pre& b, int i) { return i == 0 ? b.b0 : i == 1 ? b.b1 : b.b2; }
template <class F>
GLHD F fold2(Fp x0, Fp x1, Fp x2, Fp x3, const Fold2Pre<F>& p, uint64_t) {
  const uint64_t d20 = sub(x2.v, x0.v), d10 = sub(x1.v, x0.v), dd = sub(sub(x3.v, x1.v), d20);
  uint64_t wd[3];
#pragma unroll
  for (int i = 0; i < F::D; i++) {
    AccW r;
    zero(r);
    if (i == 0) addw(r, x0.v);
    macw(r, d20, pre_coeff(p.e[0], i));
    macw(r, d10, pre_coeff(p.e[1], i));
    macw(r, dd, pre_coeff(p.e[2], i));
    wd[i] = reduce(r);
  }
  return load_words<F>(wd);
}

}  // namespace gl

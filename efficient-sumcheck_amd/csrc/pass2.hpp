// Two-round passes for the base-field multilinear (K = 1) and inner-product
// (K = 2) provers.
//
// The round-(s+1) polynomial is, per pair, a product of K functions that are
// linear in the round-s challenge x, so its EvalsInfty values are polynomials
// of degree K in x.  One pass over the level-s table therefore yields round s's
// values and the K+1 coefficients of each round-(s+1) value; once the
// transcript produces x the host (or k_reduce_fs2) evaluates them.  The next
// pass then folds twice on the fly (with r_s and r_{s+1}) and materialises
// only every other level.  For the v = 28 inner product that is 11.4 GB of
// HBM traffic instead of the 17.2 GB of one fold per round
// (src/inner_product_sumcheck.rs:199-293 reads and writes every level), and
// round 1 becomes a pure read pass.  All outputs are the same field elements.
//
// Geometry of one pass: the materialised table has length M; p in {0,1,2}
// pending challenges fold it to the level-L table, L = M / 2^p, without
// storing intermediate levels.  Unit u in [0, L/4) owns the level-L elements
// E_j = T_L[u + j L/4], j = 0..3; each needs level-M elements at
// u + (j + 4 t) L/4, t < 2^p.  Round s pairs (E_0, E_2) and (E_1, E_3);
// round s+1 pairs (E_0 + x (E_2 - E_0), E_1 + x (E_3 - E_1)).
// Position u + m L/4 is read and written by unit u only, so the pass may run
// in place.
#pragma once
#include "kernels.hpp"

namespace effsc_hip {

// Unreduced sum of products and words kept in 32-bit columns (gl::AccW):
// one product is four v_mad_u64_u32 with the column as addend and carry out
// plus four carry counts (8 instructions against 24 for a 160-bit sum with
// compares), one word an add and a carry count.
struct Acc128 {
  gl::AccW w;
  __device__ __forceinline__ void zero() { gl::zero(w); }
  __device__ __forceinline__ void mac(uint64_t x, uint64_t y) { gl::macw(w, x, y); }
  __device__ __forceinline__ void add(uint64_t x) { gl::addw(w, x); }
  __device__ __forceinline__ Fp reduce() const { return {gl::reduce(w)}; }
};

// Number of values a pass produces: K for round s, K (K+1) coefficients for
// round s+1 (value i, coefficient j at K + i (K+1) + j).
template <int K>
constexpr int pass_values() {
  return K + K * (K + 1);
}

// Raw accumulators per unit.  K = 1: S0, S1, S2 = sums of E0, E1, E2.
// K = 2 (products of table a with table b, A = E[0], B = E[1]):
//   P1 = A0 B0, P2 = A1 B1, P3 = dA0 dB0, P4 = dA1 dB1, P5 = A2 B2,
//   P6 = eA eB, P7 = (A3 - A2)(B3 - B2), P8 = fA fB
// with dA0 = A2 - A0, dA1 = A3 - A1, eA = A1 - A0, fA = dA1 - dA0
// (so eA + fA = A3 - A2).  Then
//   round s:      q(0) = P1 + P2, q(inf) = P3 + P4
//   round s+1:    q(0)(x)   = (A0 + x dA0)(B0 + x dB0) = P1 + x (P5 - P1 - P3) + x^2 P3
//                 q(inf)(x) = (eA + x fA)(eB + x fB)   = P6 + x (P7 - P6 - P8) + x^2 P8
// (A0 dB0 + dA0 B0 = A2 B2 - A0 B0 - dA0 dB0 since A2 = A0 + dA0; likewise
// for e, f): 8 products per unit instead of 12.
template <int K>
constexpr int pass_raw() {
  return K == 1 ? 3 : 8;
}

template <int K>
struct PassAcc {
  Acc128 a[pass_raw<K>()];
};

// Accumulates one unit (level-L elements E[i][j], table i, j = 0..3).
template <int K>
__device__ __forceinline__ void pass_accumulate(const uint64_t (*E)[4], PassAcc<K>& acc) {
  if constexpr (K == 1) {
    acc.a[0].add(E[0][0]);
    acc.a[1].add(E[0][1]);
    acc.a[2].add(E[0][2]);
  } else {
    const uint64_t *A = E[0], *B = E[1];
    uint64_t dA0 = gl::sub(A[2], A[0]), dA1 = gl::sub(A[3], A[1]);
    uint64_t dB0 = gl::sub(B[2], B[0]), dB1 = gl::sub(B[3], B[1]);
    uint64_t eA = gl::sub(A[1], A[0]), eB = gl::sub(B[1], B[0]);
    uint64_t fA = gl::sub(dA1, dA0), fB = gl::sub(dB1, dB0);
    acc.a[0].mac(A[0], B[0]);
    acc.a[1].mac(A[1], B[1]);
    acc.a[2].mac(dA0, dB0);
    acc.a[3].mac(dA1, dB1);
    acc.a[4].mac(A[2], B[2]);
    acc.a[5].mac(eA, eB);
    acc.a[6].mac(gl::sub(A[3], A[2]), gl::sub(B[3], B[2]));
    acc.a[7].mac(fA, fB);
  }
}

// Raw sums -> pass values (layout of pass_values).
template <int K>
__device__ __forceinline__ void pass_values_from_raw(const Fp* s, Fp* v) {
  if constexpr (K == 1) {
    v[0] = gl::add(s[0], s[1]);              // round s: sum of low halves
    v[1] = s[0];                             // round s+1: E0 + x (E2 - E0)
    v[2] = gl::sub(s[2], s[0]);
  } else {
    v[0] = gl::add(s[0], s[1]);
    v[1] = gl::add(s[2], s[3]);
    v[2] = s[0];
    v[3] = gl::sub(gl::sub(s[4], s[0]), s[2]);
    v[4] = s[2];
    v[5] = s[5];
    v[6] = gl::sub(gl::sub(s[6], s[5]), s[7]);
    v[7] = s[7];
  }
}

// Level-L element at index e from the level-M table with P pending folds
// (c12 = c1 c2, used when P = 2).
template <int P>
__device__ __forceinline__ uint64_t fold_chain(const uint64_t* src, size_t e, size_t L, uint64_t c1, uint64_t c2,
                                               uint64_t c12) {
  if constexpr (P == 0) {
    return src[e];
  } else if constexpr (P == 1) {
    uint64_t x0 = src[e], x1 = src[e + L];
    return gl::add(x0, gl::mul(c1, gl::sub(x1, x0)));
  } else {
    return gl::fold2(src[e], src[e + L], src[e + 2 * L], src[e + 3 * L], c1, c2, c12);
  }
}

// Vector form: elements e and e+1 (e even, 16-byte aligned planes).
template <int P, bool NT>
__device__ __forceinline__ void fold_chain2(const uint64_t* src, size_t e, size_t L, uint64_t c1, uint64_t c2,
                                            uint64_t c12, uint64_t& o0, uint64_t& o1) {
  auto ld = [&](size_t i) -> v2u64 {
    if constexpr (NT)
      return __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(src + i));
    else
      return *reinterpret_cast<const v2u64*>(src + i);
  };
  if constexpr (P == 0) {
    v2u64 x = ld(e);
    o0 = x.x;
    o1 = x.y;
  } else if constexpr (P == 1) {
    v2u64 x0 = ld(e), x1 = ld(e + L);
    o0 = gl::add(x0.x, gl::mul(c1, gl::sub(x1.x, x0.x)));
    o1 = gl::add(x0.y, gl::mul(c1, gl::sub(x1.y, x0.y)));
  } else {
    v2u64 x0 = ld(e), x1 = ld(e + L), x2 = ld(e + 2 * L), x3 = ld(e + 3 * L);
    o0 = gl::fold2(x0.x, x1.x, x2.x, x3.x, c1, c2, c12);
    o1 = gl::fold2(x0.y, x1.y, x2.y, x3.y, c1, c2, c12);
  }
}

// One pass: P pending folds from the level-M tables `src` to level L = 4 q,
// written to `dst` when P > 0; accumulates the pass values.  Challenges come
// from `cptr` (device words c1, c2) when non-null, else from c1, c2.
//
// Unit order: grid-stride (span == 0), or wave chunks: each wave owns
// `span` consecutive units (a multiple of 64 VEC) at a time, so a wave reads
// every region of the pass in runs of span/64 consecutive 1 KB lines.  For the
// v = 28 inner product, runs of four lines took the read pass from 688 to
// 655 us and the first fold pass from 1030 to 965 us (runs of one line are the
// grid-stride order; longer runs measured no faster).  The chunk loop's unit
// index goes through an empty asm statement: otherwise loop strength
// reduction keeps one 64-bit pointer per region live across the loop (269
// VGPRs, one wave per SIMD, for the fold pass against 122).
template <int K, int P, int VEC, int NTM = 0>
__global__ __launch_bounds__(kBlock) void k_pass(Views<Fp, K> src, Views<Fp, K> dst, uint64_t c1, uint64_t c2,
                                                 const uint64_t* cptr, size_t q, Reduce r, size_t span = 0) {
  if (cptr) {
    c1 = cptr[0];
    c2 = cptr[1];
  }
  const uint64_t c12 = P == 2 ? gl::mul(c1, c2) : 0;
  const size_t L = 4 * q;
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) acc.a[i].zero();
  auto unit = [&](size_t u) {
    uint64_t E0[K][4], E1[K][4];
#pragma unroll
    for (int i = 0; i < K; i++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if constexpr (VEC == 2) {
          // a pass without folds reads each element exactly once: stream it
          fold_chain2<P, P == 0 || (NTM & 1)>(src.t[i].p, u + j * q, L, c1, c2, c12, E0[i][j], E1[i][j]);
        } else {
          E0[i][j] = fold_chain<P>(src.t[i].p, u + j * q, L, c1, c2, c12);
        }
        // one region's loads and folds at a time: bounds the live registers
        if constexpr (P > 0) __builtin_amdgcn_sched_barrier(0);
      }
    }
    if constexpr (P > 0) {
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if constexpr (VEC == 2 && (NTM & 2))
            __builtin_nontemporal_store(v2u64{E0[i][j], E1[i][j]}, reinterpret_cast<v2u64*>(dst.t[i].p + u + j * q));
          else if constexpr (VEC == 2)
            *reinterpret_cast<ulonglong2*>(dst.t[i].p + u + j * q) = make_ulonglong2(E0[i][j], E1[i][j]);
          else
            dst.t[i].p[u + j * q] = E0[i][j];
        }
    }
    pass_accumulate<K>(E0, acc);
    if constexpr (VEC == 2) pass_accumulate<K>(E1, acc);
  };
  if (span) {
    const size_t lane = threadIdx.x & 63;
    const size_t wv = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const size_t nwv = (size_t)gridDim.x * (blockDim.x >> 6);
    for (size_t b0 = wv * span; b0 < q; b0 += nwv * span) {
      const size_t bend = min(q, b0 + span);
#pragma unroll 1
      for (size_t u = b0 + VEC * lane; u < bend; u += 64 * VEC) {
        size_t uo = u;
        asm volatile("" : "+v"(uo));
        unit(uo);
      }
    }
  } else {
    for (size_t u = gtid() * VEC; u < q; u += (size_t)gridDim.x * blockDim.x * VEC) unit(u);
  }
  Fp raw[pass_raw<K>()], vals[pass_values<K>()];
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  finish<Fp, pass_values<K>()>(vals, r);
}

// Round-(s+1) value i at x from its K+1 coefficients.
template <int K>
__host__ __device__ __forceinline__ uint64_t pass_eval_next(const uint64_t* coeffs, int i, uint64_t x) {
  const uint64_t* c = coeffs + K + i * (K + 1);
  uint64_t v = c[K];
  for (int j = K - 1; j >= 0; j--) v = gl::add(gl::mul(v, x), c[j]);
  return v;
}

}  // namespace effsc_hip

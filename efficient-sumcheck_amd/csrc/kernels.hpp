// gfx950 kernels for the half-split (MSB) sumcheck round.
//
// One family covers the multilinear (K = 1), inner-product (K = 2) and
// product-of-K coefficient provers: the round polynomial of
// prod_i f_i(x) over a pair (lo_i, hi_i) of every table is evaluated at the
// EvalsInfty points {0, inf, 2, ..., K-1} (src/sumcheck_prover.rs:40-59):
//   h(0)   = prod lo_i
//   h(inf) = prod (hi_i - lo_i)          (leading coefficient)
//   h(t)   = prod (lo_i + t (hi_i - lo_i)), t = 2..K-1
// For K = 1 that is [sum lo] (MultilinearProver, src/provers/multilinear.rs:58-67)
// and for K = 2 it is [q(0), q(inf)] (InnerProductProver,
// src/provers/inner_product.rs:67-112).
//
// Kernels:
//   k_eval   - round polynomial of the current table, no fold (round 0, or
//              after a separate fold); pairs (k, k+half), implicit-zero high
//              half for k >= paired (src/provers/inner_product.rs:87-110).
//   k_fused  - fold with the previous challenge and evaluate in one pass over
//              the quadruple (k, k+L/4, k+L/2, k+3L/4)
//              (src/inner_product_sumcheck.rs:199-293,
//               src/multilinear_sumcheck.rs:189-254).
//   k_fold   - fold only, with the (1 - w) rule for the unpaired tail
//              (src/inner_product_sumcheck.rs:143-176).
// Field addition is exact, so the order of the cross-block reduction does not
// change any output bit.
#pragma once
#include <hip/hip_runtime.h>

#include "field.hpp"

namespace effsc_hip {

using gl::Fp;
using gl::Fp2;
using gl::Fp3;

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kMaxK = 4;             // tables per product
constexpr int kMaxWaves = 16;              // block_reduce supports up to 1024 threads

// ---------------------------------------------------------------------------
// Table views.  Base tables are one u64 plane; Fp2 and Fp3 tables are two and
// three planes (structure of arrays) `stride` words apart, so every load is
// coalesced.
// ---------------------------------------------------------------------------
typedef uint64_t v2u64 __attribute__((ext_vector_type(2)));

template <class F>
struct View;

template <>
struct View<Fp> {
  uint64_t* p;
  size_t stride;  // unused
  __device__ __forceinline__ Fp ld(size_t i) const { return {p[i]}; }
  __device__ __forceinline__ void st(size_t i, Fp x) const { p[i] = x.v; }
  __device__ __forceinline__ void ld2(size_t i, Fp& a, Fp& b) const {
    ulonglong2 q = *reinterpret_cast<const ulonglong2*>(p + i);
    a.v = q.x;
    b.v = q.y;
  }
  __device__ __forceinline__ void st2(size_t i, Fp a, Fp b) const {
    *reinterpret_cast<ulonglong2*>(p + i) = make_ulonglong2(a.v, b.v);
  }
  // streaming read: no cache allocation (data read once)
  __device__ __forceinline__ void ld2_nt(size_t i, Fp& a, Fp& b) const {
    v2u64 q = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + i));
    a.v = q.x;
    b.v = q.y;
  }
};

template <>
struct View<Fp2> {
  uint64_t* p;
  size_t stride;
  __device__ __forceinline__ Fp2 ld(size_t i) const { return {p[i], p[stride + i]}; }
  __device__ __forceinline__ void st(size_t i, Fp2 x) const {
    p[i] = x.c0;
    p[stride + i] = x.c1;
  }
  __device__ __forceinline__ void ld2(size_t i, Fp2& a, Fp2& b) const {
    ulonglong2 q0 = *reinterpret_cast<const ulonglong2*>(p + i);
    ulonglong2 q1 = *reinterpret_cast<const ulonglong2*>(p + stride + i);
    a = {q0.x, q1.x};
    b = {q0.y, q1.y};
  }
  __device__ __forceinline__ void st2(size_t i, Fp2 a, Fp2 b) const {
    *reinterpret_cast<ulonglong2*>(p + i) = make_ulonglong2(a.c0, b.c0);
    *reinterpret_cast<ulonglong2*>(p + stride + i) = make_ulonglong2(a.c1, b.c1);
  }
  __device__ __forceinline__ void st2_nt(size_t i, Fp2 a, Fp2 b) const {
    __builtin_nontemporal_store(v2u64{a.c0, b.c0}, reinterpret_cast<v2u64*>(p + i));
    __builtin_nontemporal_store(v2u64{a.c1, b.c1}, reinterpret_cast<v2u64*>(p + stride + i));
  }
  __device__ __forceinline__ void ld2_nt(size_t i, Fp2& a, Fp2& b) const {
    v2u64 q0 = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + i));
    v2u64 q1 = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + stride + i));
    a = {q0.x, q1.x};
    b = {q0.y, q1.y};
  }
};

template <>
struct View<Fp3> {
  uint64_t* p;
  size_t stride;
  __device__ __forceinline__ Fp3 ld(size_t i) const { return {p[i], p[stride + i], p[2 * stride + i]}; }
  __device__ __forceinline__ void st(size_t i, Fp3 x) const {
    p[i] = x.c0;
    p[stride + i] = x.c1;
    p[2 * stride + i] = x.c2;
  }
  __device__ __forceinline__ void ld2(size_t i, Fp3& a, Fp3& b) const {
    ulonglong2 q0 = *reinterpret_cast<const ulonglong2*>(p + i);
    ulonglong2 q1 = *reinterpret_cast<const ulonglong2*>(p + stride + i);
    ulonglong2 q2 = *reinterpret_cast<const ulonglong2*>(p + 2 * stride + i);
    a = {q0.x, q1.x, q2.x};
    b = {q0.y, q1.y, q2.y};
  }
  __device__ __forceinline__ void st2(size_t i, Fp3 a, Fp3 b) const {
    *reinterpret_cast<ulonglong2*>(p + i) = make_ulonglong2(a.c0, b.c0);
    *reinterpret_cast<ulonglong2*>(p + stride + i) = make_ulonglong2(a.c1, b.c1);
    *reinterpret_cast<ulonglong2*>(p + 2 * stride + i) = make_ulonglong2(a.c2, b.c2);
  }
  __device__ __forceinline__ void st2_nt(size_t i, Fp3 a, Fp3 b) const {
    __builtin_nontemporal_store(v2u64{a.c0, b.c0}, reinterpret_cast<v2u64*>(p + i));
    __builtin_nontemporal_store(v2u64{a.c1, b.c1}, reinterpret_cast<v2u64*>(p + stride + i));
    __builtin_nontemporal_store(v2u64{a.c2, b.c2}, reinterpret_cast<v2u64*>(p + 2 * stride + i));
  }
  __device__ __forceinline__ void ld2_nt(size_t i, Fp3& a, Fp3& b) const {
    v2u64 q0 = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + i));
    v2u64 q1 = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + stride + i));
    v2u64 q2 = __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(p + 2 * stride + i));
    a = {q0.x, q1.x, q2.x};
    b = {q0.y, q1.y, q2.y};
  }
};

template <class F, int K>
struct Views {
  View<F> t[K];
};

// Reduction scratch owned by one prover: per-block partials and the final K
// values.
struct Reduce {
  uint64_t* partials;  // gridDim * K * D words
  uint64_t* result;    // K * D words
};

// ---------------------------------------------------------------------------
// Product-of-K pair evaluation (accumulates into acc[0..K)).
// ---------------------------------------------------------------------------
template <class F, int K>
__device__ __forceinline__ void eval_pair(const F* lo, const F* hi, F* acc, uint64_t w) {
  if constexpr (K == 1) {
    acc[0] = gl::add(acc[0], lo[0]);
  } else {
    F d[K];
#pragma unroll
    for (int i = 0; i < K; i++) d[i] = gl::sub(hi[i], lo[i]);
    F p0 = lo[0], pinf = d[0];
#pragma unroll
    for (int i = 1; i < K; i++) {
      p0 = gl::mul(p0, lo[i], w);
      pinf = gl::mul(pinf, d[i], w);
    }
    acc[0] = gl::add(acc[0], p0);
    acc[1] = gl::add(acc[1], pinf);
    if constexpr (K >= 3) {
      // x_t = lo + t d, stepped from x_1 = hi
      F x[K];
#pragma unroll
      for (int i = 0; i < K; i++) x[i] = hi[i];
#pragma unroll
      for (int t = 2; t < K; t++) {
#pragma unroll
        for (int i = 0; i < K; i++) x[i] = gl::add(x[i], d[i]);
        F pt = x[0];
#pragma unroll
        for (int i = 1; i < K; i++) pt = gl::mul(pt, x[i], w);
        acc[t] = gl::add(acc[t], pt);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Wave / block reduction of K field elements.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  int lo = __shfl_xor((int)(uint32_t)v, m, 64);
  int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ Fp shfl_xor(Fp a, int m) { return {shfl_xor64(a.v, m)}; }
__device__ __forceinline__ Fp2 shfl_xor(Fp2 a, int m) { return {shfl_xor64(a.c0, m), shfl_xor64(a.c1, m)}; }
__device__ __forceinline__ Fp3 shfl_xor(Fp3 a, int m) {
  return {shfl_xor64(a.c0, m), shfl_xor64(a.c1, m), shfl_xor64(a.c2, m)};
}

constexpr int ilog2c(int p) { return p <= 1 ? 0 : 1 + ilog2c(p / 2); }
constexpr int pow2_ceil(int k) { return k <= 1 ? 1 : 2 * pow2_ceil((k + 1) / 2); }

__device__ __forceinline__ Fp select(bool c, Fp a, Fp b) { return {c ? a.v : b.v}; }
__device__ __forceinline__ Fp2 select(bool c, Fp2 a, Fp2 b) { return {c ? a.c0 : b.c0, c ? a.c1 : b.c1}; }
__device__ __forceinline__ Fp3 select(bool c, Fp3 a, Fp3 b) {
  return {c ? a.c0 : b.c0, c ? a.c1 : b.c1, c ? a.c2 : b.c2};
}
__device__ __forceinline__ uint64_t readlane64(uint64_t x, int l) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l), hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ Fp readlane(Fp a, int l) { return {readlane64(a.v, l)}; }
__device__ __forceinline__ Fp2 readlane(Fp2 a, int l) { return {readlane64(a.c0, l), readlane64(a.c1, l)}; }
__device__ __forceinline__ Fp3 readlane(Fp3 a, int l) {
  return {readlane64(a.c0, l), readlane64(a.c1, l), readlane64(a.c2, l)};
}

// The wave's sums of K values, reduced transposed: at the step with lane mask
// m, lanes with that bit set keep the upper half of their values and lanes
// without it the lower half, and each receives its partner's copy of the half
// it keeps.  After log2(P) steps (P = K rounded up to a power of two) every
// lane holds one value, index lane >> (6 - log2 P), and the remaining masks
// finish a plain butterfly: lane l returns the sum of value
// l >> (6 - log2 P) (zero for indices >= K).  P - 1 + 6 - log2 P exchanges
// instead of 6 K.
template <class F, int K>
__device__ __forceinline__ F wave_reduce_t(const F* acc) {
  constexpr int P = pow2_ceil(K), LP = ilog2c(P);
  const int lane = threadIdx.x & 63;
  F v[P];
#pragma unroll
  for (int i = 0; i < P; i++) v[i] = i < K ? acc[i] : gl::zero<F>();
#pragma unroll
  for (int l = 0; l < LP; l++) {
    const int m = 32 >> l, half = P >> (l + 1);
    const bool hi = lane & m;
#pragma unroll
    for (int j = 0; j < half; j++) {
      F keep = select(hi, v[j + half], v[j]);
      F send = select(hi, v[j], v[j + half]);
      v[j] = gl::add(keep, shfl_xor(send, m));
    }
  }
#pragma unroll
  for (int m = 32 >> LP; m >= 1; m >>= 1) v[0] = gl::add(v[0], shfl_xor(v[0], m));
  return v[0];
}

// Reduces acc over the block; the result is valid in every lane of wave 0.
//
// Within a wave the K values are reduced transposed (wave_reduce_t; measured
// for K = 8 over 1024 threads: 2.0 against 8.8 us for a butterfly per value).
// One lane per (wave, value) writes LDS; wave 0 sums the waves, lane i value
// i, and broadcasts with readlane.
template <class F, int K>
__device__ __forceinline__ void block_reduce(F* acc) {
  // up to 24 words per wave: the pass values of the two-round passes over
  // F_p^3 (8 values of 3 words)
  __shared__ uint64_t lds[kMaxWaves][24];
  static_assert(K * F::D <= 24, "block_reduce: too many words");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (K == 1) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc[0] = gl::add(acc[0], shfl_xor(acc[0], m));
    if (lane == 0) gl::store_words(acc[0], &lds[wave][0]);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = blockDim.x >> 6;
      for (int wv = 1; wv < nw; wv++) acc[0] = gl::add(acc[0], gl::load_words<F>(&lds[wv][0]));
    }
  } else {
    constexpr int LP = ilog2c(pow2_ceil(K));
    const F v0 = wave_reduce_t<F, K>(acc);
    const int idx = lane >> (6 - LP);
    if ((lane & ((64 >> LP) - 1)) == 0 && idx < K) gl::store_words(v0, &lds[wave][idx * F::D]);
    __syncthreads();
    if (wave == 0) {
      const int nw = blockDim.x >> 6;
      F s = gl::zero<F>();
      if (lane < K)
        for (int wv = 0; wv < nw; wv++) s = gl::add(s, gl::load_words<F>(&lds[wv][lane * F::D]));
#pragma unroll
      for (int i = 0; i < K; i++) acc[i] = readlane(s, i);
    }
  }
  __syncthreads();
}

// Writes this block's partial (K values) to r.partials[blockIdx.x].  The
// cross-block sum is a separate one-workgroup kernel (k_reduce) launched next
// on the same stream: a last-block-arrives scheme would need a device-scope
// release fence per workgroup, which on gfx950 writes back the XCD's L2 and
// costs far more than one extra dispatch.
template <class F, int K>
__device__ __forceinline__ void finish(F* acc, Reduce r) {
  block_reduce<F, K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < K; i++) gl::store_words(acc[i], r.partials + ((size_t)blockIdx.x * K + i) * F::D);
  }
}

// Adds this thread's share of `nblocks` partials of K values into s.  Four
// blocks per step with every load issued before the adds: the partials come
// from HBM, and one load-add chain per block left the single reducing
// workgroup latency-bound (8192 blocks of 8 values took about 8 us).
template <class F, int K>
__device__ __forceinline__ void sum_partials(const uint64_t* partials, unsigned nblocks, F* s) {
  // two blocks per step for F_p^3 sums of three or more values: four spilled
  // registers in the 1024-thread reduction kernels
  // (one block per step when the values span more than 12 words)
  constexpr int U = (F::D * K > 12) ? 1 : (F::D == 3 && K >= 3) ? 2 : 4;
  const unsigned st = blockDim.x;
  unsigned b = threadIdx.x;
  for (; b + (U - 1) * st < nblocks; b += U * st) {
    F v[U][K];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int i = 0; i < K; i++) v[u][i] = gl::load_words<F>(partials + ((size_t)(b + u * st) * K + i) * F::D);
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int i = 0; i < K; i++) s[i] = gl::add(s[i], v[u][i]);
  }
  for (; b < nblocks; b += st) {
#pragma unroll
    for (int i = 0; i < K; i++) s[i] = gl::add(s[i], gl::load_words<F>(partials + ((size_t)b * K + i) * F::D));
  }
}

// Sums `nblocks` partials of K values each into r.result (one workgroup).
template <class F, int K>
__global__ __launch_bounds__(1024) void k_reduce(Reduce r, unsigned nblocks) {
  F s[K];
#pragma unroll
  for (int i = 0; i < K; i++) s[i] = gl::zero<F>();
  sum_partials<F, K>(r.partials, nblocks, s);
  block_reduce<F, K>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < K; i++) gl::store_words(s[i], r.result + i * F::D);
  }
}

// ---------------------------------------------------------------------------
// Fold helper: lo + c (hi - lo), mixed input/output fields.
// ---------------------------------------------------------------------------
template <class FO, class FI>
__device__ __forceinline__ FO fold1(FI lo, FI hi, FO c, uint64_t w) {
  return gl::add(gl::mul(c, gl::sub(hi, lo), w), gl::lift_to<FO>(lo));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t gtid() { return (size_t)blockIdx.x * blockDim.x + threadIdx.x; }

// Round polynomial of K tables of length L without folding.  Pairs
// (k, k + half) for k < paired; (k, 0) for paired <= k < half.
// VEC = 2 requires paired == half and even half.
template <class F, int K, int VEC>
__global__ __launch_bounds__(kBlock) void k_eval(Views<F, K> t, size_t half, size_t paired, uint64_t w,
                                                 Reduce r) {
  F acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = gl::zero<F>();
  const size_t step = (size_t)gridDim.x * blockDim.x * VEC;
  if constexpr (VEC == 2) {
    for (size_t k = gtid() * 2; k < half; k += step) {
      F lo0[K], lo1[K], hi0[K], hi1[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        t.t[i].ld2_nt(k, lo0[i], lo1[i]);
        t.t[i].ld2_nt(k + half, hi0[i], hi1[i]);
      }
      eval_pair<F, K>(lo0, hi0, acc, w);
      eval_pair<F, K>(lo1, hi1, acc, w);
    }
  } else {
    for (size_t k = gtid(); k < half; k += step) {
      F lo[K], hi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        lo[i] = t.t[i].ld(k);
        hi[i] = k < paired ? t.t[i].ld(k + half) : gl::zero<F>();
      }
      eval_pair<F, K>(lo, hi, acc, w);
    }
  }
  finish<F, K>(acc, r);
}

// F_p^3 products of K >= 3 tables: the same fold and evaluation as k_fused
// (fold the length-L tables `in` with c into `out`, length L/2, and evaluate
// the pairs (k, k + L/4)), one table at a time, keeping running products
// h(0) = prod lo, h(inf) = prod (hi - lo), h(t) = prod (lo + t (hi - lo))
// instead of every folded pair.  At most 256 VGPRs (two waves per SIMD): the
// all-tables-at-once form needs one wave.  VEC = 2: two quadruples per
// thread step with 16-byte loads and stores per plane (even q and strides);
// measured 131 VGPRs without spills for F_p^3 x 3 tables, where VEC = 1 took
// 256 with spills, and 4% faster rounds.  For the larger instantiations the
// compiler keeps the table loop rolled (its unrolled size exceeds the pragma
// threshold); the table pointers then come from the kernel arguments by index.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <class FI, class FO, int K, int VEC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_fused_seq(
    Views<FI, K> in, Views<FO, K> out, FO c, const uint64_t* cptr, size_t q, uint64_t w, Reduce r) {
  if (cptr) c = gl::load_words<FO>(cptr);
  FO acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = gl::zero<FO>();
  const size_t step = (size_t)gridDim.x * blockDim.x * VEC;
  // the table loop stays rolled for extension outputs: unrolled, the compiler
  // hoisted every table's loads, and the one-quadruple F_p^3 form spilled at
  // the two-wave cap of 256 VGPRs
  constexpr int UT = FO::D >= 2 ? 1 : K;
  for (size_t k = gtid() * VEC; k < q; k += step) {
    FO p[VEC][K];
#pragma unroll UT
    for (int i = 0; i < K; i++) {
      FI x0[VEC], x1[VEC], x2[VEC], x3[VEC];
      if constexpr (VEC == 2) {
        in.t[i].ld2(k, x0[0], x0[1]);
        in.t[i].ld2(k + q, x1[0], x1[1]);
        in.t[i].ld2(k + 2 * q, x2[0], x2[1]);
        in.t[i].ld2(k + 3 * q, x3[0], x3[1]);
      } else {
        x0[0] = in.t[i].ld(k);
        x1[0] = in.t[i].ld(k + q);
        x2[0] = in.t[i].ld(k + 2 * q);
        x3[0] = in.t[i].ld(k + 3 * q);
      }
      FO lo[VEC], hi[VEC];
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        lo[e] = fold1<FO>(x0[e], x2[e], c, w);
        hi[e] = fold1<FO>(x1[e], x3[e], c, w);
      }
      if constexpr (VEC == 2) {
        out.t[i].st2(k, lo[0], lo[1]);
        out.t[i].st2(k + q, hi[0], hi[1]);
      } else {
        out.t[i].st(k, lo[0]);
        out.t[i].st(k + q, hi[0]);
      }
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const FO d = gl::sub(hi[e], lo[e]);
        FO x = hi[e];
        if (i == 0) {
          p[e][0] = lo[e];
          p[e][1] = d;
#pragma unroll
          for (int t = 2; t < K; t++) {
            x = gl::add(x, d);
            p[e][t] = x;
          }
        } else {
          p[e][0] = gl::mul(p[e][0], lo[e], w);
          p[e][1] = gl::mul(p[e][1], d, w);
#pragma unroll
          for (int t = 2; t < K; t++) {
            x = gl::add(x, d);
            p[e][t] = gl::mul(p[e][t], x, w);
          }
        }
      }
    }
#pragma unroll
    for (int e = 0; e < VEC; e++)
#pragma unroll
      for (int t = 0; t < K; t++) acc[t] = gl::add(acc[t], p[e][t]);
  }
  finish<FO, K>(acc, r);
}

// Base-field terms of a product of three values a_i + b_i c (i = 0..2):
// adds e0 = a0 a1 a2, e1 = b0 a1 a2 + a0 (a1 b2 + b1 a2),
// e2 = b0 (a1 b2 + b1 a2) + a0 b1 b2 and e3 = b0 b1 b2 into e[0..3].
__device__ __forceinline__ void base3_terms(const Fp* a, const Fp* b, Fp* e) {
  const Fp a12 = gl::mul(a[1], a[2], 0), b12 = gl::mul(b[1], b[2], 0);
  const Fp m12 = gl::add(gl::mul(a[1], b[2], 0), gl::mul(b[1], a[2], 0));
  e[0] = gl::add(e[0], gl::mul(a[0], a12, 0));
  e[1] = gl::add(e[1], gl::add(gl::mul(b[0], a12, 0), gl::mul(a[0], m12, 0)));
  e[2] = gl::add(e[2], gl::add(gl::mul(b[0], m12, 0), gl::mul(a[0], b12, 0)));
  e[3] = gl::add(e[3], gl::mul(b[0], b12, 0));
}

// The evaluation-point coefficients of one base quadruple (x0, x1, x2, x3)
// of a table folded with c: h(0) = lo', h(inf) = hi' - lo', h(2) = 2 hi' -
// lo', each a[t] + b[t] c.
__device__ __forceinline__ void base3_points(Fp x0, Fp x1, Fp x2, Fp x3, Fp* a, Fp* b) {
  const Fp dl = gl::sub(x2, x0), dh = gl::sub(x3, x1);
  const Fp ai = gl::sub(x1, x0), bi = gl::sub(dh, dl);
  a[0] = x0;
  b[0] = dl;
  a[1] = ai;
  b[1] = bi;
  a[2] = gl::add(x1, ai);  // 2 x1 - x0
  b[2] = gl::add(dh, bi);  // 2 dh - dl
}

// sum_k e[t][k] c^k for the three points, added into acc.
__device__ __forceinline__ void base3_finish(const Fp (*e)[4], Fp3 c, uint64_t w, Fp3* acc) {
  const Fp3 c2 = gl::mul(c, c, w), c3 = gl::mul(c2, c, w);
#pragma unroll
  for (int t = 0; t < 3; t++)
    acc[t] = gl::add(acc[t], gl::add(gl::add(gl::lift_to<Fp3>(e[t][0]), gl::mul(c, e[t][1], w)),
                                     gl::add(gl::mul(c2, e[t][2], w), gl::mul(c3, e[t][3], w))));
}

// The first fused round of a product of three base-field tables under an
// F_p^3 challenge c.  Every folded value is then affine in c with base-field
// coefficients: lo' = x0 + c (x2 - x0), and the evaluation points of a pair
// are h(0) = lo', h(inf) = hi' - lo', h(2) = 2 hi' - lo', each a_i + b_i c.
// Their product over the three tables is e0 + e1 c + e2 c^2 + e3 c^3 with
//   e0 = a0 a1 a2,  e1 = b0 a1 a2 + a0 (a1 b2 + b1 a2),
//   e2 = b0 (a1 b2 + b1 a2) + a0 b1 b2,  e3 = b0 b1 b2,
// so the kernel sums e_k over its pairs in the base field (10 products per
// point instead of two F_p^3 products) and multiplies by c^k once.  The sums
// are the same field elements as k_fused_seq's; the folded tables are
// written as by k_fused_seq.  NT: nontemporal loads and stores (VEC = 2),
// for planes too large to be reused from the caches.
template <int VEC, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_fused_base3(Views<Fp, 3> in, Views<Fp3, 3> out, Fp3 c, const uint64_t* cptr,
                                                        size_t q, uint64_t w, Reduce r) {
  if (cptr) c = gl::load_words<Fp3>(cptr);
  Fp e[3][4];
#pragma unroll
  for (int t = 0; t < 3; t++)
#pragma unroll
    for (int k = 0; k < 4; k++) e[t][k] = gl::zero<Fp>();
  const size_t step = (size_t)gridDim.x * blockDim.x * VEC;
  for (size_t k = gtid() * VEC; k < q; k += step) {
    Fp a[VEC][3][3], b[VEC][3][3];  // [element][point][table]
#pragma unroll
    for (int i = 0; i < 3; i++) {
      Fp x0[VEC], x1[VEC], x2[VEC], x3[VEC];
      if constexpr (VEC == 2 && NT) {
        in.t[i].ld2_nt(k, x0[0], x0[1]);
        in.t[i].ld2_nt(k + q, x1[0], x1[1]);
        in.t[i].ld2_nt(k + 2 * q, x2[0], x2[1]);
        in.t[i].ld2_nt(k + 3 * q, x3[0], x3[1]);
      } else if constexpr (VEC == 2) {
        in.t[i].ld2(k, x0[0], x0[1]);
        in.t[i].ld2(k + q, x1[0], x1[1]);
        in.t[i].ld2(k + 2 * q, x2[0], x2[1]);
        in.t[i].ld2(k + 3 * q, x3[0], x3[1]);
      } else {
        x0[0] = in.t[i].ld(k);
        x1[0] = in.t[i].ld(k + q);
        x2[0] = in.t[i].ld(k + 2 * q);
        x3[0] = in.t[i].ld(k + 3 * q);
      }
      Fp3 lo[VEC], hi[VEC];
#pragma unroll
      for (int u = 0; u < VEC; u++) {
        lo[u] = fold1<Fp3>(x0[u], x2[u], c, w);
        hi[u] = fold1<Fp3>(x1[u], x3[u], c, w);
        Fp pa[3], pb[3];
        base3_points(x0[u], x1[u], x2[u], x3[u], pa, pb);
#pragma unroll
        for (int t = 0; t < 3; t++) {
          a[u][t][i] = pa[t];
          b[u][t][i] = pb[t];
        }
      }
      if constexpr (VEC == 2 && NT) {
        out.t[i].st2_nt(k, lo[0], lo[1]);
        out.t[i].st2_nt(k + q, hi[0], hi[1]);
      } else if constexpr (VEC == 2) {
        out.t[i].st2(k, lo[0], lo[1]);
        out.t[i].st2(k + q, hi[0], hi[1]);
      } else {
        out.t[i].st(k, lo[0]);
        out.t[i].st(k + q, hi[0]);
      }
    }
#pragma unroll
    for (int u = 0; u < VEC; u++)
#pragma unroll
      for (int t = 0; t < 3; t++) base3_terms(a[u][t], b[u][t], e[t]);
  }
  Fp3 acc[3] = {gl::zero<Fp3>(), gl::zero<Fp3>(), gl::zero<Fp3>()};
  base3_finish(e, c, w, acc);
  finish<Fp3, 3>(acc, r);
}
#pragma clang diagnostic pop

// alias: positions k and k+q are read and written by the same thread only,
// positions k+2q and k+3q are only read.  q = L/4.  VEC = 2 requires even q.
template <class FI, class FO, int K, int VEC>
__global__ __launch_bounds__(kBlock) void k_fused(Views<FI, K> in, Views<FO, K> out, FO c, const uint64_t* cptr,
                                                  size_t q, uint64_t w, Reduce r) {
  if (cptr) c = gl::load_words<FO>(cptr);
  FO acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = gl::zero<FO>();
  const size_t step = (size_t)gridDim.x * blockDim.x * VEC;
  if constexpr (VEC == 2) {
    for (size_t k = gtid() * 2; k < q; k += step) {
      FI x0a[K], x0b[K], x1a[K], x1b[K], x2a[K], x2b[K], x3a[K], x3b[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        in.t[i].ld2(k, x0a[i], x0b[i]);
        in.t[i].ld2(k + q, x1a[i], x1b[i]);
        in.t[i].ld2(k + 2 * q, x2a[i], x2b[i]);
        in.t[i].ld2(k + 3 * q, x3a[i], x3b[i]);
      }
      FO loa[K], lob[K], hia[K], hib[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        loa[i] = fold1<FO>(x0a[i], x2a[i], c, w);
        lob[i] = fold1<FO>(x0b[i], x2b[i], c, w);
        hia[i] = fold1<FO>(x1a[i], x3a[i], c, w);
        hib[i] = fold1<FO>(x1b[i], x3b[i], c, w);
        out.t[i].st2(k, loa[i], lob[i]);
        out.t[i].st2(k + q, hia[i], hib[i]);
      }
      eval_pair<FO, K>(loa, hia, acc, w);
      eval_pair<FO, K>(lob, hib, acc, w);
    }
  } else {
    for (size_t k = gtid(); k < q; k += step) {
      FO lo[K], hi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        FI x0 = in.t[i].ld(k), x1 = in.t[i].ld(k + q), x2 = in.t[i].ld(k + 2 * q), x3 = in.t[i].ld(k + 3 * q);
        lo[i] = fold1<FO>(x0, x2, c, w);
        hi[i] = fold1<FO>(x1, x3, c, w);
        out.t[i].st(k, lo[i]);
        out.t[i].st(k + q, hi[i]);
      }
      eval_pair<FO, K>(lo, hi, acc, w);
    }
  }
  finish<FO, K>(acc, r);
}

// Fold only: out[k] = in[k] + c (in[k+half] - in[k]) for k < paired and
// in[k] (1 - c) for paired <= k < half.  May run in place (thread k reads k
// and k+half and writes k).
template <class FI, class FO, int K>
__global__ __launch_bounds__(kBlock) void k_fold(Views<FI, K> in, Views<FO, K> out, FO c, const uint64_t* cptr,
                                                 size_t half, size_t paired, uint64_t w) {
  if (cptr) c = gl::load_words<FO>(cptr);
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < half; k += step) {
#pragma unroll
    for (int i = 0; i < K; i++) {
      FI lo = in.t[i].ld(k);
      FI hi = k < paired ? in.t[i].ld(k + half) : gl::zero<FI>();
      out.t[i].st(k, fold1<FO>(lo, hi, c, w));
    }
  }
}

// Counter-based uniform values in [0, p): element i of seed s is the first
// candidate below p of mix(mix(s) + i G + a K), a = 0, 1, ... (splitmix64
// finaliser).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t uniform_key(uint64_t seed) { return mix64(seed); }
__device__ __forceinline__ uint64_t uniform_word(uint64_t key, uint64_t i) {
  uint64_t base = key + i * 0x9E3779B97F4A7C15ull;
  uint64_t c = mix64(base);
  for (uint64_t a = 1; c >= gl::P; a++) c = mix64(base + a * 0xD1B54A32D192ED03ull);
  return c;
}

}  // namespace effsc_hip

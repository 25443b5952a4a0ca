// Two-round passes over the extension fields F_p^2 and F_p^3 for the
// multilinear (K = 1) and inner-product (K = 2) provers in the device loop.
//
// The same algebra as the base-field passes (pass2.hpp): the round-(s+1)
// values are polynomials of degree K in the round-s challenge x, so one pass
// over the level-s table gives round s's values and the K + 1 coefficients of
// each round-(s+1) value, and the next pass folds the stored level with both
// challenges on the fly, materialising every other level only.  Here the
// table elements, challenges and sums are extension elements (inputs may be
// base-field words, lifted).  The two pending folds of a level-L element are
// one double fold by prepared constants (gl::fold2: one reduction per
// coefficient); the raw sums are kept reduced (one F product per raw sum and
// unit); k_reduce_fs2_ext sums the workgroups' values and runs the two
// transcript steps.  All outputs are the same field elements as one round per
// launch.
#include <utility>

#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

namespace effsc_hip {
namespace pv {
namespace {

// Compile-time loop: fn(std::integral_constant<int, i>) for i < N, expanded in
// the source.  With #pragma unroll the compiler declined the pass's unit loops
// and kept the per-unit arrays, indexed by the loop counter, in scratch.
template <class Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn&& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

// The two pending folds' constants (gl::prepare_fold2) from the challenges
// at cptr (c1 then c2), kept in scalar registers.
template <class F>
__device__ __forceinline__ gl::Fold2Pre<F> uniform_fold2(const uint64_t* cptr, uint64_t w) {
  gl::Fold2Pre<F> p = gl::prepare_fold2(gl::load_words<F>(cptr), gl::load_words<F>(cptr + F::D), w);
  constexpr int NW = sizeof(p.e[0]) / 8, NE = sizeof(p.e) / sizeof(p.e[0]);
#pragma unroll
  for (int k = 0; k < NE; k++) {
    uint64_t* wd = reinterpret_cast<uint64_t*>(&p.e[k]);
#pragma unroll
    for (int i = 0; i < NW; i++) wd[i] = uniform64(wd[i]);
  }
  return p;
}

template <bool NT, class FI>
__device__ __forceinline__ void ld2x(const View<FI>& t, size_t i, FI& a, FI& b) {
  if constexpr (NT)
    t.ld2_nt(i, a, b);
  else
    t.ld2(i, a, b);
}

// VEC consecutive level-L elements from e (e even and 16-byte aligned planes
// when VEC = 2), P = 0 (the stored level itself, lifted) or P = 2 pending folds.
template <int P, int VEC, bool NT, class FI, class F>
__device__ __forceinline__ void load_level(const View<FI>& t, size_t e, size_t L, const gl::Fold2Pre<F>& c,
                                           uint64_t w, F (&o)[VEC]) {
  if constexpr (VEC == 2) {
    if constexpr (P == 0) {
      FI a, b;
      ld2x<NT>(t, e, a, b);
      o[0] = gl::lift_to<F>(a);
      o[1] = gl::lift_to<F>(b);
    } else {
      FI x0[2], x1[2], x2[2], x3[2];
      ld2x<NT>(t, e, x0[0], x0[1]);
      ld2x<NT>(t, e + L, x1[0], x1[1]);
      ld2x<NT>(t, e + 2 * L, x2[0], x2[1]);
      ld2x<NT>(t, e + 3 * L, x3[0], x3[1]);
      o[0] = gl::fold2(x0[0], x1[0], x2[0], x3[0], c, w);
      o[1] = gl::fold2(x0[1], x1[1], x2[1], x3[1], c, w);
    }
  } else {
    if constexpr (P == 0)
      o[0] = gl::lift_to<F>(t.ld(e));
    else
      o[0] = gl::fold2(t.ld(e), t.ld(e + L), t.ld(e + 2 * L), t.ld(e + 3 * L), c, w);
  }
}

template <int VEC, bool NT, class F>
__device__ __forceinline__ void store_level(const View<F>& t, size_t e, const F (&x)[VEC]) {
  if constexpr (VEC == 2 && NT)
    t.st2_nt(e, x[0], x[1]);
  else if constexpr (VEC == 2)
    t.st2(e, x[0], x[1]);
  else
    t.st(e, x[0]);
}

// Raw sums -> pass values (pass2.hpp layout: K values of round s, then
// value i's coefficient j at K + i (K + 1) + j).
template <class F, int K>
__device__ __forceinline__ void values_from_raw(const F* s, F* v) {
  if constexpr (K == 1) {
    v[0] = gl::add(s[0], s[1]);
    v[1] = s[0];
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

// Raw sums of one unit (pass2.hpp: pass_accumulate); A from table 0, B from
// table 1, element v of each VEC group.
template <class F, int K, int VEC, int V>
__device__ __forceinline__ void accumulate_unit(const F (&E)[K][4][VEC], F* s, uint64_t w) {
  if constexpr (K == 1) {
    s[0] = gl::add(s[0], E[0][0][V]);
    s[1] = gl::add(s[1], E[0][1][V]);
    s[2] = gl::add(s[2], E[0][2][V]);
  } else {
    const F A0 = E[0][0][V], A1 = E[0][1][V], A2 = E[0][2][V], A3 = E[0][3][V];
    const F B0 = E[1][0][V], B1 = E[1][1][V], B2 = E[1][2][V], B3 = E[1][3][V];
    const F dA0 = gl::sub(A2, A0), dA1 = gl::sub(A3, A1), dB0 = gl::sub(B2, B0), dB1 = gl::sub(B3, B1);
    s[0] = gl::add(s[0], gl::mul(A0, B0, w));
    s[1] = gl::add(s[1], gl::mul(A1, B1, w));
    s[2] = gl::add(s[2], gl::mul(dA0, dB0, w));
    s[3] = gl::add(s[3], gl::mul(dA1, dB1, w));
    s[4] = gl::add(s[4], gl::mul(A2, B2, w));
    s[5] = gl::add(s[5], gl::mul(gl::sub(A1, A0), gl::sub(B1, B0), w));
    s[6] = gl::add(s[6], gl::mul(gl::sub(A3, A2), gl::sub(B3, B2), w));
    s[7] = gl::add(s[7], gl::mul(gl::sub(dA1, dA0), gl::sub(dB1, dB0), w));
  }
}

// One pass over K tables: P pending folds from the level-M tables `src`
// (field FI) to level L = 4 q, written to `dst` (field F) when P > 0.  Unit u
// owns E_j = T_L[u + j q], j < 4 (VEC consecutive units per thread step);
// round s pairs (E0, E2) and (E1, E3), round s + 1 pairs
// (E0 + x (E2 - E0), E1 + x (E3 - E1)).  Raw sums per unit:
// K = 1: E0, E1, E2; K = 2 (A from table 0, B from table 1):
//   A0 B0, A1 B1, dA0 dB0, dA1 dB1, A2 B2, eA eB, (A3 - A2)(B3 - B2), fA fB
// with dA0 = A2 - A0, dA1 = A3 - A1, eA = A1 - A0, fA = dA1 - dA0.
// Challenges: 2 D words at cptr (c1 then c2).  One region's loads and folds
// at a time (sched_barrier), at most 256 VGPRs (two waves per SIMD); NT
// streams the source and the written level past the caches.
template <class FI, class F, int K, int P, int VEC, bool NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_pass_ext(
    Views<FI, K> src, Views<F, K> dst, const uint64_t* cptr, size_t q, uint64_t w, Reduce r) {
  constexpr int NR = pass_raw<K>(), NV = pass_values<K>();
  gl::Fold2Pre<F> cc{};
  if constexpr (P > 0) cc = uniform_fold2<F>(cptr, w);
  const size_t L = 4 * q;
  F s[NR];
  static_for<NR>([&](auto i) { s[i] = gl::zero<F>(); });
  for (size_t u0 = gtid() * VEC; u0 < q; u0 += (size_t)gridDim.x * blockDim.x * VEC) {
    // the unit index through an empty asm: otherwise loop strength reduction
    // keeps a 64-bit pointer per region and plane live across the loop
    size_t u = u0;
    asm volatile("" : "+v"(u));
    F E[K][4][VEC];
    static_for<K>([&](auto i) {
      static_for<4>([&](auto j) {
        load_level<P, VEC, NT>(src.t[i], u + j * q, L, cc, w, E[i][j]);
        if constexpr (P > 0) __builtin_amdgcn_sched_barrier(0);
      });
    });
    if constexpr (P > 0)
      static_for<K>([&](auto i) { static_for<4>([&](auto j) { store_level<VEC, NT>(dst.t[i], u + j * q, E[i][j]); }); });
    static_for<VEC>([&](auto v) { accumulate_unit<F, K, VEC, decltype(v)::value>(E, s, w); });
  }
  F v[NV];
  values_from_raw<F, K>(s, v);
  finish<F, NV>(v, r);
}

// Round-(s+1) value i at x from its K + 1 coefficients (Horner).
template <class F, int K>
__device__ __forceinline__ F eval_next(const F* vals, int i, F x, uint64_t w) {
  const F* c = vals + K + i * (K + 1);
  F v = c[K];
#pragma unroll
  for (int j = K - 1; j >= 0; j--) v = gl::add(gl::mul(v, x, w), c[j]);
  return v;
}

template <class F>
__device__ __forceinline__ F uniform_elem(const F& a) {
  uint64_t wd[3] = {0, 0, 0};
  gl::store_words(a, wd);
#pragma unroll
  for (int c = 0; c < 3; c++) wd[c] = uniform64(wd[c]);
  return gl::load_words<F>(wd);
}

// The two transcript steps of a pass from its summed values s (wave 0,
// uniform values): round s values, challenge x; then round s + 1 values (the
// pass polynomials at x), challenge y.  poly receives 2 K elements, chals
// (x, y), cur (x, y): the pending folds of the next pass; xy (shared memory)
// x after the first step and y after the second.
template <class F, int K>
__device__ __forceinline__ F emit_first_ext(const F* s, LaneFs& fs, uint64_t* poly, uint64_t* xy) {
  constexpr int D = F::D;
  const bool lane0 = (threadIdx.x & 63) == 0;
  uint64_t wd[3];
#pragma unroll
  for (int i = 0; i < K; i++) {
    gl::store_words(s[i], wd);
    if (lane0)
#pragma unroll
      for (int c = 0; c < D; c++) poly[i * D + c] = wd[c];
    fs.send(D, wd);
  }
  uint64_t x[3];
  fs.challenge(D, x);
  if (lane0)
#pragma unroll
    for (int c = 0; c < D; c++) xy[c] = x[c];
  return gl::load_words<F>(x);
}
template <class F, int K>
__device__ __forceinline__ void emit_second_ext(const F* s, const F& X, LaneFs& fs, DevFs* t, uint64_t* poly,
                                                uint64_t* chals, uint64_t* cur, uint64_t w, uint64_t* xy) {
  constexpr int D = F::D;
  const int lane = threadIdx.x & 63;
  const bool lane0 = lane == 0;
  // lane i < K evaluates value i; the others repeat value 0
  F cf[K + 1];
#pragma unroll
  for (int j = 0; j <= K; j++) {
    cf[j] = s[K + j];
#pragma unroll
    for (int i = 1; i < K; i++) cf[j] = select(lane == i, s[K + i * (K + 1) + j], cf[j]);
  }
  F mine = cf[K];
#pragma unroll
  for (int j = K - 1; j >= 0; j--) mine = gl::add(gl::mul(mine, X, w), cf[j]);
  uint64_t wd[3];
#pragma unroll
  for (int i = 0; i < K; i++) {
    gl::store_words(readlane(mine, i), wd);
    if (lane0)
#pragma unroll
      for (int c = 0; c < D; c++) poly[(K + i) * D + c] = wd[c];
    fs.send(D, wd);
  }
  uint64_t y[3], x[3];
  gl::store_words(X, x);
  fs.challenge(D, y);
  if (lane0) {
#pragma unroll
    for (int c = 0; c < D; c++) {
      chals[c] = x[c];
      chals[D + c] = y[c];
      cur[c] = x[c];
      cur[D + c] = y[c];
      if (xy) xy[D + c] = y[c];
    }
  }
  fs.store(t);
}

// The summed pass values (sum_partials_words) in wave 0.
template <class F, int NV>
__device__ __forceinline__ void load_values(const uint64_t* sh, F* s) {
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = gl::load_words<F>(sh + i * F::D);
}

// Sums the pass values of `nblocks` workgroups and runs the two transcript
// steps.
template <class F, int K>
__global__ __launch_bounds__(kReduceBlock) void k_reduce_fs2_ext(Reduce r, unsigned nblocks, DevFs* t, uint64_t* poly,
                                                                 uint64_t* chals, uint64_t* cur, uint64_t w) {
  constexpr int NV = pass_values<K>();
  __shared__ uint64_t sh[kReduceBlock];
  __shared__ uint64_t xy[6];
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  sum_partials_words<NV * F::D>(r.partials, nblocks, sh);
  if (threadIdx.x >= 64) return;
  F s[NV];
  load_values<F, NV>(sh, s);
  const F X = emit_first_ext<F, K>(s, fs, poly, xy);
  emit_second_ext<F, K>(s, X, fs, t, poly, chals, cur, w, nullptr);
}

// ---------------------------------------------------------------------------
// EqFactored over F_p^2 / F_p^3 in two-round passes (src/provers/eq_factored.rs:
// 139-172, the left phase).  The proof is the inner product of p with
// EQ(x) = eq_l[x >> m] eq_r[x & (2^m - 1)] (m = log2 |eq_r|); while both
// variables of a pass are left variables, q = |level| / 4 is a multiple of
// 2^m and unit u + j q (j < 4) has EQ = L_j R(u) with L_j = eq_l[(u >> m) +
// j (q >> m)] and R(u) = eq_r[u & (2^m - 1)].  A workgroup owns a chunk of
// units inside one block of 2^m, so its L_j are fixed, and the eight raw sums
// of the inner-product pass (pass values of K = 2) are L-combinations times
// T-combinations, T_j = sum over the chunk of A_j R:
//   L0 T0, L1 T1, (L2 - L0)(T2 - T0), (L3 - L1)(T3 - T1), L2 T2,
//   (L1 - L0)(T1 - T0), (L3 - L2)(T3 - T2), (L3 - L2 - L1 + L0)(T3 - T2 - T1 + T0).
// Each T_j is a lazy Karatsuba sum (gl::KDot) against eq_r in its prepared
// form (gl::kdot_prep), so an element costs 6 F_p^3 (3 F_p^2) word products
// and no reduction; eq_r does not change while the left variables are bound.
// p enters with P pending folds (the previous pass's challenges), applied on
// the fly, and the folded level is written (in place is safe: element e is
// written only by the thread that reads e, e + L, e + 2 L, e + 3 L).
// ---------------------------------------------------------------------------

// lo + c (hi - lo) with a prepared challenge
template <class F, class Pre>
__device__ __forceinline__ F fold_pre(const F& lo, const F& hi, const Pre& c) {
  return gl::add(gl::mul(gl::sub(hi, lo), c), lo);
}


// The level-4L elements behind level-L elements e, e + 1 (e even): x[k] at
// e + k L (P = 2), or the level-L pair itself (P = 0).
template <int P, class F>
struct Quad2 {
  F x[P > 0 ? 4 : 1][2];
};
template <int P, class F>
__device__ __forceinline__ void quad_load(const View<F>& t, size_t e, size_t L, Quad2<P, F>& d) {
#pragma unroll
  for (int k = 0; k < (P > 0 ? 4 : 1); k++) t.ld2(e + k * L, d.x[k][0], d.x[k][1]);
}
// The two pending folds (gl::fold2) of a loaded quad.
template <int P, class F>
__device__ __forceinline__ void quad_fold(const Quad2<P, F>& d, const gl::Fold2Pre<F>& p, uint64_t w, F& o0, F& o1) {
  if constexpr (P == 0) {
    o0 = d.x[0][0];
    o1 = d.x[0][1];
  } else {
    o0 = gl::fold2(d.x[0][0], d.x[1][0], d.x[2][0], d.x[3][0], p, w);
    o1 = gl::fold2(d.x[0][1], d.x[1][1], d.x[2][1], d.x[3][1], p, w);
  }
}

// eq_r -> its prepared form, KDot<F>::R planes of stride st.
template <class F>
__global__ __launch_bounds__(kBlock) void k_eq_prep_r(View<F> eqr, size_t n, uint64_t* rp, size_t st) {
  constexpr int R = gl::KDot<F>::R;
  for (size_t b = gtid(); b < n; b += (size_t)gridDim.x * blockDim.x) {
    uint64_t o[R];
    gl::kdot_prep(eqr.ld(b), o);
#pragma unroll
    for (int i = 0; i < R; i++) rp[i * st + b] = o[i];
  }
}

// One pass over units [blockIdx.x chunk, + chunk) of level L = 4 q: wave j
// (kBlock = 4 waves) owns quarter j, two units per lane step (16-byte loads
// and stores; q, chunk and the strides even, chunk a multiple of 128), so
// each lane keeps one lazy sum.  The eight raw products are formed by lanes
// 0..7 of wave 0, one each.
template <class F, int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(P == 0 ? 4 : 2, 8))) void k_eq_pass_ext(
    View<F> src, View<F> dst, const uint64_t* cptr, size_t q, size_t chunk, View<F> eql, const uint64_t* rp,
    size_t rst, unsigned m_log, uint64_t w, Reduce r) {
  constexpr int R = gl::KDot<F>::R, D = F::D, NV = pass_values<2>();
  static_assert(kBlock == 256, "one wave per quarter");
  __shared__ uint64_t tsh[4][3], rsh[8][3];
  gl::Fold2Pre<F> pf{};
  if constexpr (P > 0) pf = uniform_fold2<F>(cptr, w);
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  const size_t u0 = (size_t)blockIdx.x * chunk;
  const unsigned j = threadIdx.x >> 6, lane = threadIdx.x & 63;
  gl::KDot<F> acc;
  acc.zero();
  // the next step's table loads are issued before this step's arithmetic
  size_t u = u0 + 2 * lane;
  const size_t uend = u0 + chunk;
  Quad2<P, F> cur;
  quad_load<P>(src, u + j * q, L, cur);
  for (;;) {
    const size_t un = u + 128;
    const bool more = un < uend;  // the same for every lane: chunk is a multiple of 128
    Quad2<P, F> nxt;
    if (more) quad_load<P>(src, un + j * q, L, nxt);
    F e0, e1;
    quad_fold<P>(cur, pf, w, e0, e1);
    if constexpr (P > 0) dst.st2(u + j * q, e0, e1);
    const size_t b = u & mask;
    uint64_t ra[R], rb[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(rp + i * rst + b);
      ra[i] = v.x;
      rb[i] = v.y;
    }
    acc.add_term(e0, ra);
    acc.add_term(e1, rb);
    if (!more) break;
    cur = nxt;
    u = un;
  }
  F t = acc.value(w);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) t = gl::add(t, shfl_xor(t, m));
  if (lane == 0) gl::store_words(t, tsh[j]);
  __syncthreads();
  if (threadIdx.x < 8) {
    const int i = threadIdx.x;
    const size_t b = u0 >> m_log, qs = q >> m_log;
    F l[4], T[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      l[k] = eql.ld(b + k * qs);
      T[k] = gl::load_words<F>(tsh[k]);
    }
    // raw product i: the same combination of the L_k and of the T_k
    auto d = [](const F* v, int a, int b) { return gl::sub(v[a], v[b]); };
    F x, y;
    switch (i) {
      case 0: x = l[0]; y = T[0]; break;
      case 1: x = l[1]; y = T[1]; break;
      case 2: x = d(l, 2, 0); y = d(T, 2, 0); break;
      case 3: x = d(l, 3, 1); y = d(T, 3, 1); break;
      case 4: x = l[2]; y = T[2]; break;
      case 5: x = d(l, 1, 0); y = d(T, 1, 0); break;
      case 6: x = d(l, 3, 2); y = d(T, 3, 2); break;
      default:
        x = gl::sub(d(l, 3, 1), d(l, 2, 0));
        y = gl::sub(d(T, 3, 1), d(T, 2, 0));
    }
    gl::store_words(gl::mul(x, y, w), rsh[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    F raw[8], vals[NV];
#pragma unroll
    for (int i = 0; i < 8; i++) raw[i] = gl::load_words<F>(rsh[i]);
    values_from_raw<F, 2>(raw, vals);
#pragma unroll
    for (int i = 0; i < NV; i++) gl::store_words(vals[i], r.partials + ((size_t)blockIdx.x * NV + i) * D);
  }
}

// Sums the EQ pass values and runs the two transcript steps, then folds eq_l
// (both variables of the pass are left ones): half h0 with x, by waves
// 1..15 while wave 0 runs the second transcript step, then half h1 with y
// (0 skips a fold).
template <class F>
__global__ __launch_bounds__(kReduceBlock) void k_eq_reduce_fs2_ext(Reduce r, unsigned nblocks, DevFs* t,
                                                                    uint64_t* poly, uint64_t* chals, uint64_t* cur,
                                                                    View<F> tab, size_t h0, size_t h1, uint64_t w) {
  constexpr int NV = pass_values<2>(), D = F::D;
  __shared__ uint64_t sh[kReduceBlock];
  __shared__ uint64_t xy[6];
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  sum_partials_words<NV * D>(r.partials, nblocks, sh);
  F s[NV], X;
  if (threadIdx.x < 64) {
    load_values<F, NV>(sh, s);
    X = emit_first_ext<F, 2>(s, fs, poly, xy);
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    emit_second_ext<F, 2>(s, X, fs, t, poly, chals, cur, w, xy);
  } else {
    const auto x = gl::prepare(gl::load_words<F>(xy), w);
    for (size_t k = threadIdx.x - 64; k < h0; k += blockDim.x - 64) tab.st(k, fold_pre(tab.ld(k), tab.ld(k + h0), x));
  }
  __syncthreads();
  const auto y = gl::prepare(gl::load_words<F>(xy + D), w);
  for (size_t k = threadIdx.x; k < h1; k += blockDim.x) tab.st(k, fold_pre(tab.ld(k), tab.ld(k + h1), y));
}

template <class FI, class F, int K, int P>
int launch_pass_ext_t(effsc_hip_prover* h, size_t q, const uint64_t* cptr) {
  const Views<FI, K> in = views<FI, K>(src_planes(h), src_stride(h));
  const Views<F, K> out = views<F, K>(h->work, h->work_stride);
  const bool vec = q % 2 == 0 && src_stride(h) % 2 == 0 && h->work_stride % 2 == 0;
  const size_t L = 4 * q;
  // fold passes over planes of 128 MiB and more: nontemporal loads and stores
  const bool nt = vec && P > 0 && 8.0 * (double)(L << P) >= kNtMinBytes;
  const unsigned g = blocks_for(h, vec ? q / 2 : q);
  TRY(prof_begin(h));
  if (nt)
    hipLaunchKernelGGL((k_pass_ext<FI, F, K, P, 2, true>), dim3(g), dim3(kBlock), 0, h->stream, in, out, cptr, q, h->w,
                       h->red);
  else if (vec)
    hipLaunchKernelGGL((k_pass_ext<FI, F, K, P, 2, false>), dim3(g), dim3(kBlock), 0, h->stream, in, out, cptr, q,
                       h->w, h->red);
  else
    hipLaunchKernelGGL((k_pass_ext<FI, F, K, P, 1, false>), dim3(g), dim3(kBlock), 0, h->stream, in, out, cptr, q,
                       h->w, h->red);
  HIPCHK(hipGetLastError());
  // read the stored level (2^P L elements), write level L when P > 0
  const double bytes = 8.0 * K * ((double)(1 << P) * (double)L * FI::D + (P > 0 ? (double)L * F::D : 0.0));
  const double muls = (double)q * (4.0 * K * (P == 0 ? 0.0 : 3.0 * mul_cost<F>()) + (K == 2 ? 8.0 * mul_cost<F>() : 0.0));
  TRY(prof_end(h, 3, bytes, muls));
  h->last_grid = g;
  return EFFSC_HIP_OK;
}

}  // namespace

// Used over F_p^2, and for the multilinear prover over F_p^3 (measured with
// the SHAKE transcript, v = 26 over F_p^2: inner product 2.41 -> 2.12 ms with
// three separate folds, 1.66 ms with gl::fold2, multilinear 1.06 -> 0.88 -> 0.73
// ms; v = 24 over F_p^3: multilinear 0.669 -> 0.644 -> 0.471 ms).  The F_p^3
// inner-product pass (8 F_p^3 products per unit at two waves per SIMD)
// measured 1.071 against 1.078 ms for one round per launch: no gain, so it
// runs only with EFFSC_HIP_EXT3_PASS=1.
bool ext_twostep(const effsc_hip_prover* h) {
  const bool field_ok = h->ext == 2 || (h->ext == 3 && (h->K == 1 || std::getenv("EFFSC_HIP_EXT3_PASS")));
  return (h->kind == K_ML || h->kind == K_IP) && field_ok && h->K <= 2 && !h->lsb_native && !h->ooc &&
         !std::getenv("EFFSC_HIP_ONE_ROUND");
}

int launch_pass_ext(effsc_hip_prover* h, int p, size_t q, const uint64_t* cptr) {
  if (p != 0 && p != 2) return fail(EFFSC_HIP_EINVAL, "extension passes take 0 or 2 pending folds");
  const int fi = src_field(h);
  auto go = [&](auto fo, auto kc) -> int {
    using F = decltype(fo);
    constexpr int K = decltype(kc)::value;
    if (fi == 1)
      return p == 0 ? launch_pass_ext_t<Fp, F, K, 0>(h, q, cptr) : launch_pass_ext_t<Fp, F, K, 2>(h, q, cptr);
    return p == 0 ? launch_pass_ext_t<F, F, K, 0>(h, q, cptr) : launch_pass_ext_t<F, F, K, 2>(h, q, cptr);
  };
  return with_field(h->ext, [&](auto fo) -> int {
    if constexpr (decltype(fo)::D == 1) {
      return fail(EFFSC_HIP_EINVAL, "extension passes need an extension field");
    } else {
      if (h->K == 1) return go(fo, std::integral_constant<int, 1>{});
      return go(fo, std::integral_constant<int, 2>{});
    }
  });
}

int launch_reduce_fs2_ext(effsc_hip_prover* h, const Reduce& red, unsigned g, size_t round) {
  const int E = h->ext, K = h->K;
  uint64_t* poly = h->d_polys + round * K * E;
  uint64_t* chals = h->d_chals + round * E;
  return with_field(E, [&](auto fo) -> int {
    using F = decltype(fo);
    if constexpr (F::D == 1) {
      return fail(EFFSC_HIP_EINVAL, "extension passes need an extension field");
    } else {
      if (K == 1)
        hipLaunchKernelGGL((k_reduce_fs2_ext<F, 1>), dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, h->d_fs, poly,
                           chals, h->d_cur, h->w);
      else
        hipLaunchKernelGGL((k_reduce_fs2_ext<F, 2>), dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, h->d_fs, poly,
                           chals, h->d_cur, h->w);
      HIPCHK(hipGetLastError());
      return EFFSC_HIP_OK;
    }
  });
}

int eq_prep_r_ext(effsc_hip_prover* h) {
  const int R = h->ext == 2 ? gl::KDot<Fp2>::R : gl::KDot<Fp3>::R;
  const size_t words = (size_t)R * h->eq_r_stride;
  if (h->eq_rp_words < words) {
    if (h->eq_rp) (void)hipFree(h->eq_rp);
    h->eq_rp = nullptr;
    h->eq_rp_words = 0;
    TRY(alloc_dev(&h->eq_rp, words));
    h->eq_rp_words = words;
  }
  return with_field(h->ext, [&](auto fo) -> int {
    using F = decltype(fo);
    if constexpr (F::D == 1) {
      return fail(EFFSC_HIP_EINVAL, "extension passes need an extension field");
    } else {
      hipLaunchKernelGGL(k_eq_prep_r<F>, dim3(blocks_for(h, h->eq_r_len)), dim3(kBlock), 0, h->stream,
                         View<F>{h->eq_r, h->eq_r_stride}, h->eq_r_len, h->eq_rp, h->eq_r_stride);
      HIPCHK(hipGetLastError());
      return EFFSC_HIP_OK;
    }
  });
}

int eq_pass_ext(effsc_hip_prover* h, int p, size_t q, size_t chunk, unsigned m_log, unsigned g) {
  if (p != 0 && p != 2) return fail(EFFSC_HIP_EINVAL, "extension passes take 0 or 2 pending folds");
  if (q % 2 || chunk % 128 || src_stride(h) % 2 || h->work_stride % 2 || h->eq_r_stride % 2)
    return fail(EFFSC_HIP_EINVAL, "eq pass: odd sizes");
  return with_field(h->ext, [&](auto fo) -> int {
    using F = decltype(fo);
    if constexpr (F::D == 1) {
      return fail(EFFSC_HIP_EINVAL, "extension passes need an extension field");
    } else {
      const View<F> src{src_planes(h)[0], src_stride(h)}, dst{h->work[0], h->work_stride};
      const View<F> eql{h->eq_l, h->eq_l_stride};
      TRY(prof_begin(h));
      if (p == 0)
        hipLaunchKernelGGL((k_eq_pass_ext<F, 0>), dim3(g), dim3(kBlock), 0, h->stream, src, dst, h->d_cur, q, chunk,
                           eql, h->eq_rp, h->eq_r_stride, m_log, h->w, h->red);
      else
        hipLaunchKernelGGL((k_eq_pass_ext<F, 2>), dim3(g), dim3(kBlock), 0, h->stream, src, dst, h->d_cur, q, chunk,
                           eql, h->eq_rp, h->eq_r_stride, m_log, h->w, h->red);
      HIPCHK(hipGetLastError());
      // p read at level 2^P L and level L written when P > 0; word products:
      // R per level-L element (Karatsuba) and 3 prepared folds of D^2 each
      const double L = 4.0 * (double)q;
      const double bytes = 8.0 * F::D * (p ? 5.0 * L : L);
      const double muls = L * ((F::D == 2 ? 3.0 : 6.0) + (p ? 3.0 * F::D * F::D : 0.0));
      TRY(prof_end(h, 3, bytes, muls));
      return EFFSC_HIP_OK;
    }
  });
}

int eq_reduce_fs2_ext(effsc_hip_prover* h, unsigned g, size_t round, size_t h0, size_t h1) {
  const int E = h->ext;
  uint64_t* poly = h->d_polys + round * 2 * E;
  uint64_t* chals = h->d_chals + round * E;
  return with_field(E, [&](auto fo) -> int {
    using F = decltype(fo);
    if constexpr (F::D == 1) {
      return fail(EFFSC_HIP_EINVAL, "extension passes need an extension field");
    } else {
      hipLaunchKernelGGL(k_eq_reduce_fs2_ext<F>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, poly,
                         chals, h->d_cur, View<F>{h->eq_l, h->eq_l_stride}, h0, h1, h->w);
      HIPCHK(hipGetLastError());
      return EFFSC_HIP_OK;
    }
  });
}

}  // namespace pv
}  // namespace effsc_hip

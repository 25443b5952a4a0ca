// EqFactoredProver (src/provers/eq_factored.rs:67-289): per-round path,
// device-resident proofs and the constructor.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

// ===========================================================================
// EqFactoredProver: sum_x eq(w, x) p(x) with eq split into eq(w_L, .) and
// eq(w_R, .) (src/provers/eq_factored.rs:67-289).  Tables in the handle:
// p -> in[0]/work[0] (K = 1), eq_l, eq_r.
// ===========================================================================
namespace effsc_hip {
namespace pv {

template <class F>
__global__ __launch_bounds__(kBlock) void k_eq_table(uint64_t* planes, size_t stride, const uint64_t* wpt, size_t v,
                                                     uint64_t w) {
  // table[x] = prod_j (x_j ? w_j : 1 - w_j), x_j = bit (v-1-j) of x
  // (build_eq_table, src/provers/eq_factored.rs:217-239, MSB layout)
  View<F> t{planes, stride};
  const size_t n = (size_t)1 << v;
  for (size_t x = gtid(); x < n; x += (size_t)gridDim.x * blockDim.x) {
    F acc = gl::one<F>();
    for (size_t j = 0; j < v; j++) {
      F wj = gl::load_words<F>(wpt + j * F::D);
      F f = ((x >> (v - 1 - j)) & 1) ? wj : gl::sub(gl::one<F>(), wj);
      acc = gl::mul(acc, f, w);
    }
    t.st(x, acc);
  }
}

// Left phase (rounds_elapsed < v_L): q0 = sum_a eqL_lo[a] sum_b eqR[b] p_lo[a m + b]
// and q_inf = sum_a (eqL_hi[a] - eqL_lo[a]) sum_b eqR[b] (p_hi - p_lo)[a m + b]
// (round_poly_left, src/provers/eq_factored.rs:139-172).  Each thread handles
// index i = a m + b of the low half and contributes eqL-weighted terms
// directly: eqL_lo[a] eqR[b] p_lo and (eqL_hi - eqL_lo)[a] eqR[b] (p_hi - p_lo).
template <class F>
__global__ __launch_bounds__(kBlock) void k_eq_left(View<F> p, size_t p_half, View<F> eql, size_t eql_half,
                                                    View<F> eqr, size_t m_log, uint64_t w, Reduce r) {
  F acc[2] = {gl::zero<F>(), gl::zero<F>()};
  const size_t mask = ((size_t)1 << m_log) - 1;
  for (size_t i = gtid(); i < p_half; i += (size_t)gridDim.x * blockDim.x) {
    size_t a = i >> m_log, b = i & mask;
    F er = eqr.ld(b);
    F pl = p.ld(i), ph = p.ld(p_half + i);
    F el = eql.ld(a), eh = eql.ld(eql_half + a);
    F t0 = gl::mul(er, pl, w);
    F td = gl::mul(er, gl::sub(ph, pl), w);
    acc[0] = gl::add(acc[0], gl::mul(el, t0, w));
    acc[1] = gl::add(acc[1], gl::mul(gl::sub(eh, el), td, w));
  }
  finish<F, 2>(acc, r);
}

int eq_dispatch_left(effsc_hip_prover* h) {
  uint64_t* const* src = src_planes(h);
  size_t sst = src_stride(h);
  size_t p_half = h->len >> 1;
  size_t eql_half = h->eq_l_len >> 1;
  size_t m_log = 0;
  while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
  unsigned g = blocks_for(h, p_half);
  return with_field(h->ext, [&](auto f) -> int {
    using F = decltype(f);
    hipLaunchKernelGGL(k_eq_left<F>, dim3(g), dim3(kBlock), 0, h->stream, View<F>{src[0], sst}, p_half,
                       View<F>{h->eq_l, h->eq_l_stride}, eql_half, View<F>{h->eq_r, h->eq_r_stride}, m_log, h->w,
                       h->red);
    HIPCHK(hipGetLastError());
    return reduce<F, 2>(h, g);
  });
}

// Folds one eq half-table in place (ip::fold semantics on a pow2 table).
int eq_fold_table(effsc_hip_prover* h, uint64_t* planes, size_t stride, size_t* len, const uint64_t* c) {
  if (*len <= 1) return EFFSC_HIP_OK;
  size_t half = *len >> 1;
  unsigned g = blocks_for(h, half);
  uint64_t* pl[kMaxK] = {planes};
  TRY(with_field(h->ext, [&](auto f) -> int {
    using F = decltype(f);
    Views<F, 1> v = views<F, 1>(pl, stride);
    hipLaunchKernelGGL((k_fold<F, F, 1>), dim3(g), dim3(kBlock), 0, h->stream, v, v, from_words<F>(c), nullptr, half,
                       half, h->w);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }));
  *len = half;
  return EFFSC_HIP_OK;
}

// Reads element 0 of an ext-plane table.
// ---------------------------------------------------------------------------

// Device-resident EqFactored proof.  A round is the inner-product round of
// (p, E) with E = eq_l (x) eq_r while the left variables are bound and
// E = eq_l[0] eq_r afterwards (round_poly_left / round_poly_right,
// src/provers/eq_factored.rs:139-209): q0 = sum E_lo p_lo and
// q_inf = sum (E_hi - E_lo)(p_hi - p_lo).  The eq tables are tiny next to p
// and stay in L2; each is folded in place right after the challenge of its
// variable is squeezed.
// ---------------------------------------------------------------------------

// Weights of pair k (half H2 of the evaluated table): left phase, index
// k = a m + b with m = 2^m_log = |eq_r|; right phase, eq_r alone.
template <class F>
__device__ __forceinline__ void eq_weights(View<F> eql, size_t eql_len, View<F> eqr, size_t m_log, int left, size_t k,
                                           size_t H2, uint64_t w, F& el, F& eh) {
  if (left) {
    const size_t a = k >> m_log, b = k & (((size_t)1 << m_log) - 1);
    const F er = eqr.ld(b);
    el = gl::mul(eql.ld(a), er, w);
    eh = gl::mul(eql.ld(a + (eql_len >> 1)), er, w);
  } else {
    el = eqr.ld(k);
    eh = eqr.ld(k + H2);
  }
}

// One round over p of length L before the round's fold.  With `fold`, p is
// folded with the device challenge on the fly (quadruples, in place when pin
// and pout alias: thread k alone reads k, k+q, k+2q, k+3q and writes k, k+q).
template <class FI, class F>
__global__ __launch_bounds__(kBlock) void k_eq_round(View<FI> pin, View<F> pout, size_t L, int fold,
                                                     const uint64_t* cptr, View<F> eql, size_t eql_len, View<F> eqr,
                                                     size_t m_log, int left, uint64_t w, Reduce r) {
  const F c = fold ? gl::load_words<F>(cptr) : gl::zero<F>();
  const size_t H2 = fold ? L / 4 : L / 2;
  F acc[2] = {gl::zero<F>(), gl::zero<F>()};
  for (size_t k = gtid(); k < H2; k += (size_t)gridDim.x * blockDim.x) {
    F lo, hi;
    if (fold) {
      FI x0 = pin.ld(k), x1 = pin.ld(k + H2), x2 = pin.ld(k + 2 * H2), x3 = pin.ld(k + 3 * H2);
      lo = fold1<F>(x0, x2, c, w);
      hi = fold1<F>(x1, x3, c, w);
      pout.st(k, lo);
      pout.st(k + H2, hi);
    } else {
      lo = gl::lift_to<F>(pin.ld(k));
      hi = gl::lift_to<F>(pin.ld(k + H2));
    }
    F el, eh;
    eq_weights<F>(eql, eql_len, eqr, m_log, left, k, H2, w, el, eh);
    acc[0] = gl::add(acc[0], gl::mul(el, lo, w));
    acc[1] = gl::add(acc[1], gl::mul(gl::sub(eh, el), gl::sub(hi, lo), w));
  }
  finish<F, 2>(acc, r);
}

// Split-value form of k_eq_round for the left phase: pair k = a 2^m + b has
// weights L_lo(a) R(b) and L_hi(a) R(b), and a workgroup owns one chunk of
// pairs inside one block of 2^m, so its a is fixed:
//   q(0)   = L_lo(a) sum R lo,   q(inf) = (L_hi(a) - L_lo(a)) sum R (hi - lo),
// two products per pair instead of four, and two per workgroup.  The same
// partials as k_eq_round.
template <class FI, class F>
__global__ __launch_bounds__(kBlock) void k_eq_round_sv(View<FI> pin, View<F> pout, size_t L, int fold,
                                                        const uint64_t* cptr, View<F> eql, size_t eql_len,
                                                        View<F> eqr, size_t m_log, size_t chunk, uint64_t w,
                                                        Reduce r) {
  const F c = fold ? gl::load_words<F>(cptr) : gl::zero<F>();
  const size_t H2 = fold ? L / 4 : L / 2;
  const size_t mask = ((size_t)1 << m_log) - 1;
  F T[2] = {gl::zero<F>(), gl::zero<F>()};
  const size_t k0 = (size_t)blockIdx.x * chunk;
  for (size_t k = k0 + threadIdx.x; k < k0 + chunk; k += blockDim.x) {
    F lo, hi;
    if (fold) {
      FI x0 = pin.ld(k), x1 = pin.ld(k + H2), x2 = pin.ld(k + 2 * H2), x3 = pin.ld(k + 3 * H2);
      lo = fold1<F>(x0, x2, c, w);
      hi = fold1<F>(x1, x3, c, w);
      pout.st(k, lo);
      pout.st(k + H2, hi);
    } else {
      lo = gl::lift_to<F>(pin.ld(k));
      hi = gl::lift_to<F>(pin.ld(k + H2));
    }
    const F R = eqr.ld(k & mask);
    T[0] = gl::add(T[0], gl::mul(R, lo, w));
    T[1] = gl::add(T[1], gl::mul(R, gl::sub(hi, lo), w));
  }
  block_reduce<F, 2>(T);
  if (threadIdx.x == 0) {
    const size_t a = k0 >> m_log;
    const F llo = eql.ld(a), lhi = eql.ld(a + (eql_len >> 1));
    const F v[2] = {gl::mul(llo, T[0], w), gl::mul(gl::sub(lhi, llo), T[1], w)};
#pragma unroll
    for (int i = 0; i < 2; i++) gl::store_words(v[i], r.partials + ((size_t)blockIdx.x * 2 + i) * F::D);
  }
}

// Sums the round's partials, scales by eq_l[0] in the right phase, runs the
// transcript step, then folds the eq table of this round's variable in
// place (fold_half > 0) with the new challenge.
template <class F>
__global__ __launch_bounds__(kReduceBlock) void k_eq_reduce_fs(Reduce r, unsigned nblocks, View<F> eql, int right,
                                                               DevFs* t, int ext, uint64_t* poly, uint64_t* chal_out,
                                                               uint64_t* chal_cur, View<F> ft, size_t fold_half,
                                                               uint64_t w) {
  __shared__ uint64_t chal_sh[3];
  __shared__ uint64_t sh[kReduceBlock];
  sum_partials_words<2 * F::D>(r.partials, nblocks, sh);
  F s[2] = {gl::load_words<F>(sh), gl::load_words<F>(sh + F::D)};
  if (right) {
    const F sc = eql.ld(0);
    s[0] = gl::mul(s[0], sc, w);
    s[1] = gl::mul(s[1], sc, w);
  }
  if (threadIdx.x < 64) emit_round<F, 2>(s, ext, t, poly, chal_out, chal_sh);
  __syncthreads();
  if (threadIdx.x < (unsigned)ext) chal_cur[threadIdx.x] = chal_sh[threadIdx.x];
  const F c = gl::load_words<F>(chal_sh);
  for (size_t k = threadIdx.x; k < fold_half; k += blockDim.x) ft.st(k, fold1<F>(ft.ld(k), ft.ld(k + fold_half), c, w));
}

// EqFactored over the base field in two-round passes: the proof is the inner
// product of p with the eq table EQ(x) = eq_l[x >> m] eq_r[x & (2^m - 1)]
// (m = log2 of the current eq_r length; eq_l has length 1 once its
// variables are bound), so the two-round pass of pass2.hpp applies with the
// second table computed from the two small eq tables instead of read.
// k_eq_pass<P> folds p with P pending challenges on the fly (writing the
// level when P > 0) and accumulates the pass values against EQ;
// k_eq_reduce_fs2 runs the two transcript rounds and folds the eq tables of
// both variables (the current eq tables are EQ at the next level).
template <int P, bool NT>
__global__ __launch_bounds__(kBlock) void k_eq_pass(const uint64_t* src, uint64_t* dst, const uint64_t* cptr, size_t q,
                                                    const uint64_t* eql, const uint64_t* eqr, unsigned m_log,
                                                    Reduce r) {
  const uint64_t c1 = P ? cptr[0] : 0, c2 = P ? cptr[1] : 0, c12 = P == 2 ? gl::mul(c1, c2) : 0;
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  PassAcc<2> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<2>(); i++) acc.a[i].zero();
  const size_t step = (size_t)gridDim.x * blockDim.x * 2;
  for (size_t u = gtid() * 2; u < q; u += step) {
    uint64_t E0[2][4], E1[2][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      fold_chain2<P, NT || P == 0>(src, u + j * q, L, c1, c2, c12, E0[0][j], E1[0][j]);
      const size_t x = u + j * q;
      E0[1][j] = gl::mul(eql[x >> m_log], eqr[x & mask]);
      E1[1][j] = gl::mul(eql[(x + 1) >> m_log], eqr[(x + 1) & mask]);
    }
    if constexpr (P > 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if constexpr (NT)
          __builtin_nontemporal_store(v2u64{E0[0][j], E1[0][j]}, reinterpret_cast<v2u64*>(dst + u + j * q));
        else
          *reinterpret_cast<ulonglong2*>(dst + u + j * q) = make_ulonglong2(E0[0][j], E1[0][j]);
      }
    }
    pass_accumulate<2>(E0, acc);
    pass_accumulate<2>(E1, acc);
  }
  Fp raw[pass_raw<2>()], vals[pass_values<2>()];
#pragma unroll
  for (int i = 0; i < pass_raw<2>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<2>(raw, vals);
  finish<Fp, pass_values<2>()>(vals, r);
}

// Split-value form of k_eq_pass, for q a multiple of 2^m (a left variable
// still free): EQ(u + j q) = L_j R(u) with L_j = eql[(u >> m) + j (q >> m)]
// and R(u) = eqr[u & (2^m - 1)], and the units of one workgroup (one chunk
// inside a block of 2^m units) share the L_j.  With A = p, every raw sum of
// pass_accumulate<2> is then an L-combination times a T-combination,
// T_j = sum over the workgroup's units of A_j R:
//   P1 = L0 T0, P2 = L1 T1, P3 = (L2 - L0)(T2 - T0), P4 = (L3 - L1)(T3 - T1),
//   P5 = L2 T2, P6 = (L1 - L0)(T1 - T0), P7 = (L3 - L2)(T3 - T2),
//   P8 = (L3 - L2 - L1 + L0)(T3 - T2 - T1 + T0),
// so a unit costs 4 products (A_j R) instead of 4 EQ products and 8 pass
// products; the 8 raw sums are formed once per workgroup.  The same field
// elements as k_eq_pass.
template <int P, bool NT>
__global__ __launch_bounds__(kBlock) void k_eq_pass_sv(const uint64_t* src, uint64_t* dst, const uint64_t* cptr,
                                                       size_t q, size_t chunk, const uint64_t* eql,
                                                       const uint64_t* eqr, unsigned m_log, Reduce r) {
  const uint64_t c1 = P ? cptr[0] : 0, c2 = P ? cptr[1] : 0, c12 = P == 2 ? gl::mul(c1, c2) : 0;
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  Acc128 T[4];
#pragma unroll
  for (int j = 0; j < 4; j++) T[j].zero();
  const size_t u0 = (size_t)blockIdx.x * chunk;
  for (size_t u = u0 + threadIdx.x * 2; u < u0 + chunk; u += (size_t)blockDim.x * 2) {
    uint64_t E0[4], E1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) fold_chain2<P, NT || P == 0>(src, u + j * q, L, c1, c2, c12, E0[j], E1[j]);
    if constexpr (P > 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if constexpr (NT)
          __builtin_nontemporal_store(v2u64{E0[j], E1[j]}, reinterpret_cast<v2u64*>(dst + u + j * q));
        else
          *reinterpret_cast<ulonglong2*>(dst + u + j * q) = make_ulonglong2(E0[j], E1[j]);
      }
    }
    // u even: R(u) and R(u + 1) are one aligned pair of eqr
    const v2u64 R = *reinterpret_cast<const v2u64*>(eqr + (u & mask));
#pragma unroll
    for (int j = 0; j < 4; j++) {
      T[j].mac(E0[j], R.x);
      T[j].mac(E1[j], R.y);
    }
  }
  Fp t[4];
#pragma unroll
  for (int j = 0; j < 4; j++) t[j] = T[j].reduce();
  block_reduce<Fp, 4>(t);
  if (threadIdx.x == 0) {
    const size_t b = u0 >> m_log, qs = q >> m_log;
    uint64_t l[4];
#pragma unroll
    for (int j = 0; j < 4; j++) l[j] = eql[b + j * qs];
    auto m = [](uint64_t x, uint64_t y) { return Fp{gl::mul(x, y)}; };
    const uint64_t l20 = gl::sub(l[2], l[0]), l31 = gl::sub(l[3], l[1]), l10 = gl::sub(l[1], l[0]),
                   l32 = gl::sub(l[3], l[2]);
    const uint64_t t20 = gl::sub(t[2].v, t[0].v), t31 = gl::sub(t[3].v, t[1].v), t10 = gl::sub(t[1].v, t[0].v),
                   t32 = gl::sub(t[3].v, t[2].v);
    const Fp raw[8] = {m(l[0], t[0].v), m(l[1], t[1].v), m(l20, t20),         m(l31, t31),
                       m(l[2], t[2].v), m(l10, t10),     m(l32, t32),         m(gl::sub(l31, l20), gl::sub(t31, t20))};
    Fp vals[pass_values<2>()];
    pass_values_from_raw<2>(raw, vals);
#pragma unroll
    for (int i = 0; i < pass_values<2>(); i++) r.partials[(size_t)blockIdx.x * pass_values<2>() + i] = vals[i].v;
  }
}

// k_eq_pass_sv<2, false> with the previous pass's transcript step in the same
// launch (as k_pass_fs2, csrc/prove.hpp): every workgroup issues its loads,
// sums the previous pass's partials (rin, nin blocks) and runs the two
// transcript steps on wave 0, deriving the same (x, y) as k_eq_reduce_fs2.
// Both variables of that step are left ones, so this pass's eq_l is the
// double fold of the current table eql (4 h1 entries) with x then y, entry i
// from eql[i + k h1] (k < 4), and eq_r is unchanged; each workgroup folds the
// four entries it uses, and workgroup 0 writes the whole folded table to
// eql_out (the next table; eql itself is read by every workgroup).  Workgroup
// 0 alone writes the proof words, the challenges and the sponge (t_out).
// Chunks of 512 units: one unit pair per thread.
__global__ __launch_bounds__(kBlock) void k_eq_pass_sv_fs(const uint64_t* src, uint64_t* dst, size_t q,
                                                          const uint64_t* eql, size_t h1, uint64_t* eql_out,
                                                          const uint64_t* eqr, unsigned m_log, Reduce rin,
                                                          unsigned nin, Reduce rout, const DevFs* t_in, DevFs* t_out,
                                                          uint64_t* poly, uint64_t* chals, uint64_t* cur) {
  constexpr int NV = pass_values<2>();
  __shared__ uint64_t sh[kBlock];
  __shared__ uint64_t xy[2];
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  const size_t u0 = (size_t)blockIdx.x * (2 * kBlock), u = u0 + threadIdx.x * 2;
  const size_t b = u0 >> m_log, qs = q >> m_log;
  // p at level 4L, both units of the pair; the eq_r pair; thread 0: the eq_l
  // entries under L_j
  v2u64 X[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int k = 0; k < 4; k++) X[j][k] = *reinterpret_cast<const v2u64*>(src + u + j * q + k * L);
  const v2u64 R = *reinterpret_cast<const v2u64*>(eqr + (u & mask));
  uint64_t El[4][4];
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 4; k++) El[j][k] = eql[b + j * qs + k * h1];
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t_in);
  sum_partials_words<NV>(rin.partials, nin, sh);
  if (threadIdx.x < 64) {
    Fp sv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) sv[i] = Fp{sh[i]};
    emit_two_rounds<2>(sv, fs, t_out, poly, chals, cur, blockIdx.x == 0, xy);
  }
  __syncthreads();
  const uint64_t c1 = xy[0], c2 = xy[1], c12 = gl::mul(c1, c2);
  if (blockIdx.x == 0)
    for (size_t i = threadIdx.x; i < h1; i += blockDim.x)
      eql_out[i] = gl::fold2(eql[i], eql[i + h1], eql[i + 2 * h1], eql[i + 3 * h1], c1, c2, c12);
  Acc128 T[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint64_t e0 = gl::fold2(X[j][0].x, X[j][1].x, X[j][2].x, X[j][3].x, c1, c2, c12);
    const uint64_t e1 = gl::fold2(X[j][0].y, X[j][1].y, X[j][2].y, X[j][3].y, c1, c2, c12);
    *reinterpret_cast<ulonglong2*>(dst + u + j * q) = make_ulonglong2(e0, e1);
    T[j].zero();
    T[j].mac(e0, R.x);
    T[j].mac(e1, R.y);
  }
  Fp t[4];
#pragma unroll
  for (int j = 0; j < 4; j++) t[j] = T[j].reduce();
  block_reduce<Fp, 4>(t);
  if (threadIdx.x == 0) {
    uint64_t l[4];
#pragma unroll
    for (int j = 0; j < 4; j++) l[j] = gl::fold2(El[j][0], El[j][1], El[j][2], El[j][3], c1, c2, c12);
    auto m = [](uint64_t x, uint64_t y) { return Fp{gl::mul(x, y)}; };
    const uint64_t l20 = gl::sub(l[2], l[0]), l31 = gl::sub(l[3], l[1]), l10 = gl::sub(l[1], l[0]),
                   l32 = gl::sub(l[3], l[2]);
    const uint64_t t20 = gl::sub(t[2].v, t[0].v), t31 = gl::sub(t[3].v, t[1].v), t10 = gl::sub(t[1].v, t[0].v),
                   t32 = gl::sub(t[3].v, t[2].v);
    const Fp raw[8] = {m(l[0], t[0].v), m(l[1], t[1].v), m(l20, t20),         m(l31, t31),
                       m(l[2], t[2].v), m(l10, t10),     m(l32, t32),         m(gl::sub(l31, l20), gl::sub(t31, t20))};
    Fp vals[NV];
    pass_values_from_raw<2>(raw, vals);
#pragma unroll
    for (int i = 0; i < NV; i++) rout.partials[(size_t)blockIdx.x * NV + i] = vals[i].v;
  }
}

// Two transcript rounds from the pass values, then the MSB folds of the eq
// table of round s's variable (t0, half h0) with x and of round s+1's (t1,
// half h1) with y; t0 and t1 may be the same table (h1 = h0 / 2 then).
// The sponge is loaded from t_in (t, or the pinned host staging of a starting
// proof) and stored to t.
__global__ __launch_bounds__(kReduceBlock) void k_eq_reduce_fs2(Reduce r, unsigned nblocks, const DevFs* t_in,
                                                                DevFs* t, uint64_t* poly, uint64_t* chals,
                                                                uint64_t* cur, uint64_t* t0, size_t h0, uint64_t* t1,
                                                                size_t h1) {
  constexpr int NV = pass_values<2>();
  __shared__ uint64_t sh[kReduceBlock];
  __shared__ uint64_t xy[3];
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t_in);
  sum_partials_words<NV>(r.partials, nblocks, sh);
  if (threadIdx.x < 64) {
    Fp sv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) sv[i] = Fp{sh[i]};
    emit_two_rounds<2>(sv, fs, t, poly, chals, xy);
  }
  __syncthreads();
  const uint64_t x = xy[0], y = xy[1];
  if (threadIdx.x < 2) cur[threadIdx.x] = xy[threadIdx.x];
  for (size_t k = threadIdx.x; k < h0; k += blockDim.x) t0[k] = gl::add(t0[k], gl::mul(x, gl::sub(t0[k + h0], t0[k])));
  __syncthreads();
  for (size_t k = threadIdx.x; k < h1; k += blockDim.x) t1[k] = gl::add(t1[k], gl::mul(y, gl::sub(t1[k + h1], t1[k])));
}

template <class F>
constexpr int eq_tail_block() {
  return F::D == 3 ? 256 : 1024;
}

// Rounds [j0, v) in one workgroup, then the final fold and the final value
// p(r) eq_l(r_L) eq_r(r_R) (src/provers/eq_factored.rs:272-288).
template <class FI, class F>
__global__ __launch_bounds__((eq_tail_block<F>())) void k_eq_tail(View<FI> pin, View<F> work, int from_input,
                                                                   size_t L, int j0, int v, int v_l, View<F> eql,
                                                                   size_t eql_len, View<F> eqr, size_t eqr_len,
                                                                   int have_chal, const uint64_t* cur_in, DevFs* t,
                                                                   int ext, uint64_t* polys, uint64_t* chals,
                                                                   uint64_t* cur_out, uint64_t* fin, uint64_t w) {
  __shared__ uint64_t chal_sh[3];
  if (threadIdx.x < 3) chal_sh[threadIdx.x] = have_chal && (int)threadIdx.x < ext ? cur_in[threadIdx.x] : 0;
  // the sponge stays in wave 0's registers for the whole tail
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  __syncthreads();
  const unsigned tid = threadIdx.x, nt = blockDim.x;
  bool src_in = from_input != 0;
  for (int j = j0; j < v; j++) {
    const int left = j < v_l;
    size_t m_log = 0;
    while (((size_t)1 << m_log) < eqr_len) m_log++;
    F acc[2] = {gl::zero<F>(), gl::zero<F>()};
    if (have_chal) {
      const F c = gl::load_words<F>(chal_sh);
      const size_t q = L / 4;
      for (size_t k = tid; k < q; k += nt) {
        F lo, hi;
        if (src_in) {
          FI x0 = pin.ld(k), x1 = pin.ld(k + q), x2 = pin.ld(k + 2 * q), x3 = pin.ld(k + 3 * q);
          lo = fold1<F>(x0, x2, c, w);
          hi = fold1<F>(x1, x3, c, w);
        } else {
          F x0 = work.ld(k), x1 = work.ld(k + q), x2 = work.ld(k + 2 * q), x3 = work.ld(k + 3 * q);
          lo = fold1<F>(x0, x2, c, w);
          hi = fold1<F>(x1, x3, c, w);
        }
        work.st(k, lo);
        work.st(k + q, hi);
        F el, eh;
        eq_weights<F>(eql, eql_len, eqr, m_log, left, k, q, w, el, eh);
        acc[0] = gl::add(acc[0], gl::mul(el, lo, w));
        acc[1] = gl::add(acc[1], gl::mul(gl::sub(eh, el), gl::sub(hi, lo), w));
      }
      L /= 2;
      src_in = false;
    } else {
      const size_t half = L / 2;
      for (size_t k = tid; k < half; k += nt) {
        F lo, hi;
        if (src_in) {
          lo = gl::lift_to<F>(pin.ld(k));
          hi = gl::lift_to<F>(pin.ld(k + half));
        } else {
          lo = work.ld(k);
          hi = work.ld(k + half);
        }
        F el, eh;
        eq_weights<F>(eql, eql_len, eqr, m_log, left, k, half, w, el, eh);
        acc[0] = gl::add(acc[0], gl::mul(el, lo, w));
        acc[1] = gl::add(acc[1], gl::mul(gl::sub(eh, el), gl::sub(hi, lo), w));
      }
    }
    block_reduce<F, 2>(acc);
    if (!left) {
      const F sc = eql.ld(0);
      acc[0] = gl::mul(acc[0], sc, w);
      acc[1] = gl::mul(acc[1], sc, w);
    }
    if (tid < 64)
      emit_round_fs<F, 2>(acc, ext, fs, polys + (size_t)(j - j0) * 2 * ext, chals + (size_t)(j - j0) * ext, chal_sh);
    __syncthreads();
    // fold the eq table of variable j with its challenge
    const F c = gl::load_words<F>(chal_sh);
    if (left) {
      const size_t hh = eql_len >> 1;
      for (size_t k = tid; k < hh; k += nt) eql.st(k, fold1<F>(eql.ld(k), eql.ld(k + hh), c, w));
      eql_len = hh;
    } else {
      const size_t hh = eqr_len >> 1;
      for (size_t k = tid; k < hh; k += nt) eqr.st(k, fold1<F>(eqr.ld(k), eqr.ld(k + hh), c, w));
      eqr_len = hh;
    }
    __syncthreads();
    have_chal = 1;
  }
  if (tid == 0) {
    // finalize: p folds 2 -> 1 with the last challenge (the eq table of the
    // last variable was folded above)
    F p0;
    if (have_chal && L >= 2) {
      const F c = gl::load_words<F>(chal_sh);
      F lo = src_in ? gl::lift_to<F>(pin.ld(0)) : work.ld(0);
      F hi = src_in ? gl::lift_to<F>(pin.ld(1)) : work.ld(1);
      p0 = fold1<F>(lo, hi, c, w);
      work.st(0, p0);
    } else {
      p0 = src_in ? gl::lift_to<F>(pin.ld(0)) : work.ld(0);
    }
    gl::store_words(gl::mul(gl::mul(p0, eql.ld(0), w), eqr.ld(0), w), fin);
    for (int i = 0; i < ext; i++) cur_out[i] = chal_sh[i];
  }
  if (tid < 64) fs.store(t);
}

int read0(effsc_hip_prover* h, const uint64_t* planes, size_t stride, int d, uint64_t* out) {
  for (int c = 0; c < h->ext; c++) {
    out[c] = 0;
    if (c >= d) continue;
    HIPCHK(hipMemcpyAsync(h->h_res, planes + c * stride, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    out[c] = h->h_res[0];
  }
  return EFFSC_HIP_OK;
}


int eq_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  const int E = h->ext;
  if (chal) {
    if (h->rounds_elapsed == 0) return fail(EFFSC_HIP_ESTATE, "challenge given before the first round");
    size_t prev_var = h->rounds_elapsed - 1;
    bool ev = false;
    TRY(fold_step(h, chal, false, &ev));
    if (prev_var < h->v_l)
      TRY(eq_fold_table(h, h->eq_l, h->eq_l_stride, &h->eq_l_len, chal));
    else
      TRY(eq_fold_table(h, h->eq_r, h->eq_r_stride, &h->eq_r_len, chal));
  }
  std::fill(out, out + 2 * E, 0);
  if (h->rounds_elapsed < h->v_l) {
    TRY(eq_dispatch_left(h));
    uint64_t vals[kMaxK * 3];
    TRY(fetch_result(h, E, vals, 2));
    std::copy(vals, vals + 2 * E, out);
  } else {
    // round_poly_right (src/provers/eq_factored.rs:178-209): the inner-product
    // round of (eq_r, p) scaled by the folded scalar eq_l[0].
    uint64_t scalar[3];
    TRY(read0(h, h->eq_l, h->eq_l_stride, E, scalar));
    size_t n = h->len;
    if (n <= 1) {
      if (n == 1) {
        uint64_t er[3], p0[3];
        TRY(read0(h, h->eq_r, h->eq_r_stride, E, er));
        TRY(read0(h, src_planes(h)[0], src_stride(h), src_field(h), p0));
        ext_mul3(E, h->w, scalar, er, p0, out);
      }
    } else {
      size_t half = n >> 1;
      uint64_t* planes[kMaxK] = {h->eq_r, src_planes(h)[0]};
      // both tables must share a stride for the K = 2 view; eq_r and p use
      // their own strides, so evaluate through per-table views
      unsigned g = blocks_for(h, half);
      TRY(with_field(E, [&](auto f) -> int {
        using F = decltype(f);
        Views<F, 2> t;
        t.t[0] = View<F>{planes[0], h->eq_r_stride};
        t.t[1] = View<F>{planes[1], src_stride(h)};
        hipLaunchKernelGGL((k_eval<F, 2, 1>), dim3(g), dim3(kBlock), 0, h->stream, t, half, half, h->w, h->red);
        HIPCHK(hipGetLastError());
        return reduce<F, 2>(h, g);
      }));
      uint64_t vals[kMaxK * 3];
      TRY(fetch_result(h, E, vals, 2));
      ext_mul(E, h->w, vals, scalar, out);
      ext_mul(E, h->w, vals + E, scalar, out + E);
    }
  }
  h->rounds_elapsed++;
  return EFFSC_HIP_OK;
}

int eq_finalize(effsc_hip_prover* h, const uint64_t* chal) {
  // src/provers/eq_factored.rs:272-280
  bool ev = false;
  TRY(fold_step(h, chal, false, &ev));
  size_t last_var = h->v == 0 ? (size_t)-1 : h->v - 1;
  if (h->v > 0 && last_var < h->v_l)
    TRY(eq_fold_table(h, h->eq_l, h->eq_l_stride, &h->eq_l_len, chal));
  else
    TRY(eq_fold_table(h, h->eq_r, h->eq_r_stride, &h->eq_r_len, chal));
  return EFFSC_HIP_OK;
}

int eq_final_value(effsc_hip_prover* h, uint64_t* out) {
  // src/provers/eq_factored.rs:282-288
  const int E = h->ext;
  std::fill(out, out + E, 0);
  if (h->len != 1) return EFFSC_HIP_OK;
  uint64_t p0[3], l0[3], r0[3];
  TRY(read0(h, src_planes(h)[0], src_stride(h), src_field(h), p0));
  TRY(read0(h, h->eq_l, h->eq_l_stride, E, l0));
  TRY(read0(h, h->eq_r, h->eq_r_stride, E, r0));
  ext_mul3(E, h->w, p0, l0, r0, out);
  return EFFSC_HIP_OK;
}

int build_eq(effsc_hip_prover* h, const uint64_t* wpt_canon_dev, size_t off, size_t vv, uint64_t** planes,
             size_t* stride, size_t* len) {
  size_t n = (size_t)1 << vv;
  *stride = round_up(n, 64);
  *len = n;
  if (!*planes) TRY(alloc_dev(planes, *stride * h->ext));
  unsigned g = blocks_for(h, n);
  return with_field(h->ext, [&](auto f) -> int {
    using F = decltype(f);
    hipLaunchKernelGGL(k_eq_table<F>, dim3(g), dim3(kBlock), 0, h->stream, *planes, *stride,
                       wpt_canon_dev + off * h->ext, vv, h->w);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  });
}

}  // namespace pv
}  // namespace effsc_hip

namespace effsc_hip {
namespace pv {
// EqFactored: a full proof from the fresh state on the device.  Rounds on
// large p: k_eq_round (fold p with the device challenge, evaluate with eq
// weights) and k_eq_reduce_fs (sum, transcript, fold the eq table of the
// round's variable in place); once p is small, k_eq_tail runs the remaining
// rounds, the final fold and the final value in one workgroup.
bool eq_device_applies(const effsc_hip_prover* h, size_t num_rounds) {
  return h->kind == K_EQ && h->rounds_elapsed == 0 && num_rounds == h->v && h->v >= 1 && h->v <= kMaxRounds &&
         h->len == ((size_t)1 << h->v) && !h->in_work && h->eq_l_len == ((size_t)1 << h->v_l) &&
         h->eq_r_len == ((size_t)1 << (h->v - h->v_l));
}

// Base-field EqFactored rounds in the register tail: the proof is the inner
// product of p with EQ = eq_l x eq_r (src/provers/eq_factored.rs:411-446),
// so once p is down to Lt = 2^11..2^14 elements the remaining rounds are
// k_tail_reg<2> over p and EQ written out at that level.  The final value
// p(r) EQ(r) equals p(r) eq_l(r_L) eq_r(r_R) (the MLE of a tensor product is
// the product of the MLEs); k_eq_post_fold then folds eq_l and eq_r with the
// tail's challenges for the post-state.
// k_eq_tail_prep: p folded with the two pending challenges (level 4 Lt ->
// Lt, in place when src == dst: element e reads e + k Lt, k < 4, and only
// its own thread writes e) and EQ[e] = eq_l[e >> m] eq_r[e & (2^m - 1)].
__global__ __launch_bounds__(kBlock) void k_eq_tail_prep(const uint64_t* src, uint64_t* dst, const uint64_t* cptr,
                                                         size_t Lt, const uint64_t* eql, const uint64_t* eqr,
                                                         unsigned m_log, uint64_t* eqf) {
  const uint64_t c1 = cptr[0], c2 = cptr[1], c12 = gl::mul(c1, c2);
  const size_t mask = ((size_t)1 << m_log) - 1;
  for (size_t e = gtid(); e < Lt; e += (size_t)gridDim.x * blockDim.x) {
    dst[e] = fold_chain<2>(src, e, Lt, c1, c2, c12);
    eqf[e] = gl::mul(eql[e >> m_log], eqr[e & mask]);
  }
}

// One workgroup: folds eq_l with the challenges of the first `left` rounds
// and eq_r with the rest (MSB folds in place, one barrier per round), and
// leaves the last challenge in cur.
__global__ __launch_bounds__(kReduceBlock) void k_eq_post_fold(uint64_t* eql, size_t eql_len, uint64_t* eqr,
                                                               size_t eqr_len, const uint64_t* chals, int rounds,
                                                               int left, uint64_t* cur) {
  for (int r = 0; r < rounds; r++) {
    const uint64_t c = chals[r];
    uint64_t* t = r < left ? eql : eqr;
    size_t& len = r < left ? eql_len : eqr_len;
    const size_t h = len / 2;
    for (size_t k = threadIdx.x; k < h; k += blockDim.x) t[k] = gl::add(t[k], gl::mul(c, gl::sub(t[k + h], t[k])));
    len = h;
    __syncthreads();
  }
  if (threadIdx.x == 0 && rounds > 0) cur[0] = chals[rounds - 1];
}

// F_p^3 EqFactored rounds in the register tail of the F_p^3 product
// (k_tail_reg3<2>) over p and EQ, as the base-field form above.  p enters
// with its pending fold; k_eq_expand_dup writes EQ at p's level 2 L1 as two
// copies of eq_l x eq_r at level L1, so the tail's pending fold
// (lo + c (hi - lo) with hi = lo) leaves EQ as it is.
__global__ __launch_bounds__(kBlock) void k_eq_expand_dup(View<Fp3> eql, View<Fp3> eqr, unsigned m_log, size_t L1,
                                                          View<Fp3> eqf, uint64_t w) {
  const size_t mask = ((size_t)1 << m_log) - 1;
  for (size_t e = gtid(); e < L1; e += (size_t)gridDim.x * blockDim.x) {
    const Fp3 x = gl::mul(eql.ld(e >> m_log), eqr.ld(e & mask), w);
    eqf.st(e, x);
    eqf.st(e + L1, x);
  }
}

// k_eq_post_fold over F_p^3 tables; leaves the last challenge (3 words) in
// cur.
__global__ __launch_bounds__(kReduceBlock) void k_eq_post_fold3(View<Fp3> eql, size_t eql_len, View<Fp3> eqr,
                                                                size_t eqr_len, const uint64_t* chals, int rounds,
                                                                int left, uint64_t* cur, uint64_t w) {
  for (int r = 0; r < rounds; r++) {
    const Fp3 c = gl::load_words<Fp3>(chals + 3 * r);
    const View<Fp3>& t = r < left ? eql : eqr;
    size_t& len = r < left ? eql_len : eqr_len;
    const size_t h = len / 2;
    for (size_t k = threadIdx.x; k < h; k += blockDim.x) t.st(k, fold1<Fp3>(t.ld(k), t.ld(k + h), c, w));
    len = h;
    __syncthreads();
  }
  if (threadIdx.x < 3 && rounds > 0) cur[threadIdx.x] = chals[3 * (rounds - 1) + threadIdx.x];
}

template <class FI, class F>
int eq_device_run(effsc_hip_prover* h, size_t v) {
  const int E = h->ext;
  const size_t v_l = h->v_l;
  // rounds on the device until p is small enough for one workgroup; eq
  // tables folded inside the reduction when small, else by k_fold
  const size_t tail = h->ext == 3 ? ((size_t)1 << 10) : h->ext == 2 ? ((size_t)1 << 11) : ((size_t)1 << 12);
  size_t fold_in_reduce = (size_t)1 << 14;
  if (const char* e = std::getenv("EFFSC_HIP_EQ_FOLD_LIMIT")) fold_in_reduce = std::strtoull(e, nullptr, 10);
  size_t L = h->len, j = 0;
  bool have = false;
  View<F> eql{h->eq_l, h->eq_l_stride}, eqr{h->eq_r, h->eq_r_stride};
  const uint64_t* cur_in = h->d_cur;  // the pending challenge when entering the tail
  if constexpr (std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value) {
    if (!std::getenv("EFFSC_HIP_ONE_ROUND")) {
      // two rounds per pass while the level evaluated is above the tail
      int pend = 0;
      // a pass's transcript step and eq folds wait for the next pass, which
      // takes them into its launch when it is small (k_eq_pass_sv_fs); the
      // eq lengths below already count the deferred folds
      struct Deferred {
        Reduce red;
        unsigned g;
        uint64_t *t0, *t1;
        size_t h0, h1;
      } dfr{};
      bool deferred = false;
      auto reduce_now = [&]() -> int {
        hipLaunchKernelGGL(k_eq_reduce_fs2, dim3(1), dim3(kReduceBlock), 0, h->stream, dfr.red, dfr.g, fs_in(h),
                           h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, dfr.t0, dfr.h0, dfr.t1, dfr.h1);
        HIPCHK(hipGetLastError());
        j += 2;
        deferred = false;
        return EFFSC_HIP_OK;
      };
      while (j + (deferred ? 2 : 0) + 2 <= v) {
        const size_t Lv = L >> pend;
        if (Lv <= tail || Lv < 8) break;
        size_t m_log = 0;
        while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
        const size_t q = Lv / 4;
        // split-value form while q is a multiple of 2^m: chunks of at least
        // 512 units (one step of the workgroup), inside one block of 2^m, and
        // at most grid_max workgroups (measured v = 24, first pass: 8192
        // one-step workgroups 38.9 us, 2048 four-step workgroups 29.1 us)
        const size_t gmax = (size_t)h->grid_max;
        size_t chunk = 2 * kBlock;
        while (q / chunk > gmax) chunk *= 2;
        const bool sv = q % ((size_t)1 << m_log) == 0 && chunk <= ((size_t)1 << m_log) && q % chunk == 0 &&
                        !std::getenv("EFFSC_HIP_EQ_DIRECT");
        const unsigned g = sv ? (unsigned)(q / chunk)
                              : (unsigned)std::max<size_t>(
                                    1, std::min<size_t>((q / 2 + kBlock - 1) / kBlock, gmax * kPassGridScale));
        const uint64_t* src = h->in_work ? h->work[0] : h->in[0];
        const bool nt = 8.0 * (double)L >= kNtMinBytes;
        // fused with the deferred step: both of its variables are left ones
        // (one eq_l buffer folded into the other), a split-value pass of
        // one-step workgroups, few enough to be resident together
        const bool fuse = deferred && h->fuse_parts && sv && pend == 2 && !nt && chunk == 2 * (size_t)kBlock && g <= kFuseMaxGrid &&
                          dfr.g <= kFuseMaxIn && dfr.t0 == h->eq_l && dfr.t1 == h->eq_l &&
                          !std::getenv("EFFSC_HIP_NO_FUSE");
        if (deferred && !fuse) TRY(reduce_now());
        if (fuse && !h->eq_l_alt) TRY(alloc_dev(&h->eq_l_alt, h->eq_l_stride * (size_t)h->ext));
        Reduce pred = h->red;
        if (fuse) {
          const Reduce ro = dfr.red.partials == h->fuse_parts ? h->red : Reduce{h->fuse_parts, h->red.result};
          const DevFs* tin = fs_in(h);
          DevFs* tout = tin == h->d_fs ? h->d_fs_alt : h->d_fs;
          hipLaunchKernelGGL(k_eq_pass_sv_fs, dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], q, h->eq_l, dfr.h1,
                             h->eq_l_alt, h->eq_r, (unsigned)m_log, dfr.red, dfr.g, ro, tin, tout, h->d_polys + j * 2,
                             h->d_chals + j, h->d_cur);
          HIPCHK(hipGetLastError());
          if (tout == h->d_fs_alt) h->fs_pending = h->d_fs_alt;
          std::swap(h->eq_l, h->eq_l_alt);
          j += 2;
          deferred = false;
          pred = ro;
        } else {
        TRY(prof_begin(h));
        if (sv && pend == 0)
          hipLaunchKernelGGL((k_eq_pass_sv<0, true>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], h->d_cur,
                             q, chunk, h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        else if (sv && nt)
          hipLaunchKernelGGL((k_eq_pass_sv<2, true>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], h->d_cur,
                             q, chunk, h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        else if (sv)
          hipLaunchKernelGGL((k_eq_pass_sv<2, false>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0],
                             h->d_cur, q, chunk, h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        else if (pend == 0)
          hipLaunchKernelGGL((k_eq_pass<0, true>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], h->d_cur, q,
                             h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        else if (nt)
          hipLaunchKernelGGL((k_eq_pass<2, true>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], h->d_cur, q,
                             h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        else
          hipLaunchKernelGGL((k_eq_pass<2, false>), dim3(g), dim3(kBlock), 0, h->stream, src, h->work[0], h->d_cur,
                             q, h->eq_l, h->eq_r, (unsigned)m_log, h->red);
        HIPCHK(hipGetLastError());
        TRY(prof_end(h, 3, 8.0 * ((double)L + (pend ? (double)Lv : 0.0)),
                     (double)q * ((sv ? 4.0 : 8.0 + 8.0) + (pend ? 12.0 : 0.0))));
        }
        if (pend) {
          L = Lv;
          h->len = Lv;
          h->in_work = true;
        }
        // the eq tables of variables j and j + 1, folded after the two rounds
        size_t* l0 = j < v_l ? &h->eq_l_len : &h->eq_r_len;
        uint64_t* t0 = j < v_l ? h->eq_l : h->eq_r;
        const size_t h0 = *l0 >> 1;
        *l0 = h0;
        size_t* l1 = j + 1 < v_l ? &h->eq_l_len : &h->eq_r_len;
        uint64_t* t1 = j + 1 < v_l ? h->eq_l : h->eq_r;
        const size_t h1 = *l1 >> 1;
        *l1 = h1;
        dfr = Deferred{pred, g, t0, t1, h0, h1};
        deferred = true;
        pend = 2;
      }
      if (deferred) TRY(reduce_now());
      // a fused pass leaves eq_l in the other buffer
      eql = View<F>{h->eq_l, h->eq_l_stride};
      const size_t Lt = L >> pend;
      if (pend == 2 && Lt >= 2048 && Lt <= 16384 && !std::getenv("EFFSC_HIP_EQ_TAIL")) {
        // the remaining v - j rounds in the register tail over p and EQ
        if (!h->eq_full) TRY(alloc_dev(&h->eq_full, (size_t)1 << 14));
        size_t m_log = 0;
        while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
        const uint64_t* src = h->in_work ? h->work[0] : h->in[0];
        Views<Fp, 2> tv{};
        tv.t[0] = View<Fp>{h->work[0], 0};
        tv.t[1] = View<Fp>{h->eq_full, 0};
        const int rounds = (int)(v - j);
        size_t lg = 0;
        while (((size_t)1 << lg) < Lt) lg++;
        if (Lt <= (size_t)pipe_max<2>() && rounds >= 1 && (size_t)rounds <= lg && !std::getenv("EFFSC_HIP_TAIL_REG")) {
          // the pipelined tail over p and EQ: p's two pending folds and EQ
          // at level Lt as it loads them
          Views<Fp, 2> sv{};
          sv.t[0] = View<Fp>{const_cast<uint64_t*>(src), 0};
          sv.t[1] = View<Fp>{h->eq_full, 0};
          // the tail also copies the proof block into a pinned staging
          // (EFFSC_HIP_PROOF_COPY=1: a separate copy after the tail)
          uint64_t* ph = h->h_proof_pinned && !std::getenv("EFFSC_HIP_PROOF_COPY") ? h->h_proof : nullptr;
          hipLaunchKernelGGL((k_tail_pipe<2, true>), dim3(1), dim3(kPipeThreads), 0, h->stream, sv, tv, (int)Lt, 2,
                             rounds, fs_in(h), h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, h->d_fin, h->eq_l,
                             h->eq_r, (unsigned)m_log, ph, h->d_polys, (int)kProofWords);
          h->proof_by_tail = ph != nullptr;
        } else {
        TRY(fs_flush(h));
        hipLaunchKernelGGL(k_eq_tail_prep, dim3(blocks_for(h, Lt)), dim3(kBlock), 0, h->stream, src, h->work[0],
                           h->d_cur, Lt, h->eq_l, h->eq_r, (unsigned)m_log, h->eq_full);
        HIPCHK(hipGetLastError());
        switch (Lt) {
          case 2048:
            hipLaunchKernelGGL((k_tail_reg<2, 2>), dim3(1), dim3(kTailRegThreads), 0, h->stream, tv, tv, 0, rounds, 0,
                               h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, h->d_fin);
            break;
          case 4096:
            hipLaunchKernelGGL((k_tail_reg<2, 4>), dim3(1), dim3(kTailRegThreads), 0, h->stream, tv, tv, 0, rounds, 0,
                               h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, h->d_fin);
            break;
          case 8192:
            hipLaunchKernelGGL((k_tail_reg<2, 8>), dim3(1), dim3(kTailRegThreads), 0, h->stream, tv, tv, 0, rounds, 0,
                               h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, h->d_fin);
            break;
          default:
            hipLaunchKernelGGL((k_tail_reg<2, 16>), dim3(1), dim3(kTailRegThreads), 0, h->stream, tv, tv, 0, rounds,
                               0, h->d_fs, h->d_polys + j * 2, h->d_chals + j, h->d_cur, h->d_fin);
        }
        }
        HIPCHK(hipGetLastError());
        // the proof block is complete: copy it out first (unless the tail
        // wrote it to the staging), then fold eq_l and eq_r for the
        // post-state (no proof word depends on them), so the caller waits for
        // the copy only
        if (h->proof_by_tail)
          h->proof_by_tail = false;
        else
          TRY(copy_proof_out(h));
        if (!h->ev_proof) HIPCHK(hipEventCreateWithFlags(&h->ev_proof, hipEventDisableTiming));
        HIPCHK(hipEventRecord(h->ev_proof, h->stream));
        h->proof_copied = true;
        const int left = j < v_l ? (int)(v_l - j) : 0;
        hipLaunchKernelGGL(k_eq_post_fold, dim3(1), dim3(kReduceBlock), 0, h->stream, h->eq_l, h->eq_l_len, h->eq_r,
                           h->eq_r_len, h->d_chals + j, rounds, left, h->d_cur);
        HIPCHK(hipGetLastError());
        // fresh tables for the next proof, behind the copy as well
        size_t st = 0, ln = 0;
        TRY(build_eq(h, h->eq_point, 0, v_l, &h->eq_l_spare, &st, &ln));
        TRY(build_eq(h, h->eq_point, v_l, v - v_l, &h->eq_r_spare, &st, &ln));
        h->eq_spare_ready = true;
        h->len = 1;
        h->in_work = true;
        h->eq_l_len = 1;
        h->eq_r_len = 1;
        h->rounds_elapsed = v;
        return EFFSC_HIP_OK;
      }
      if (pend == 2) {
        // the older pending challenge folds p with a multi-block fold; the
        // newer one stays pending for the rounds below
        Views<Fp, 1> a{}, b{};
        a.t[0] = View<Fp>{h->in_work ? h->work[0] : h->in[0], 0};
        b.t[0] = View<Fp>{h->work[0], 0};
        hipLaunchKernelGGL((k_fold<Fp, Fp, 1>), dim3(blocks_for(h, L / 2)), dim3(kBlock), 0, h->stream, a, b,
                           gl::zero<Fp>(), h->d_cur, L / 2, L / 2, h->w);
        HIPCHK(hipGetLastError());
        L /= 2;
        h->len = L;
        h->in_work = true;
        have = true;
        cur_in = h->d_cur + 1;
      }
    }
  }
  // every kernel below reads the transcript from d_fs
  TRY(fs_flush(h));
  if constexpr (F::D >= 2 && std::is_same<FI, F>::value) {
    if (!std::getenv("EFFSC_HIP_ONE_ROUND") && !std::getenv("EFFSC_HIP_EQ_DIRECT")) {
      // two rounds per pass while both variables of a pass are left ones
      // (split-value form, q a multiple of 2^m) and the level evaluated is
      // above the tail
      int pend = 0;
      bool prepped = false;
      while (j + 2 <= v) {
        const size_t Lv = L >> pend;
        if (Lv <= tail || Lv < 8) break;
        size_t m_log = 0;
        while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
        const size_t q = Lv / 4;
        // a wave per quarter, 128 units per wave step; at most grid_max
        // workgroups (8 extension values each in the partials buffer)
        const size_t gmax = (size_t)h->grid_max;
        size_t chunk = 128;
        while (q / chunk > gmax) chunk *= 2;
        if (q % ((size_t)1 << m_log) != 0 || chunk > ((size_t)1 << m_log) || q % chunk != 0) break;
        if (!prepped) {
          TRY(eq_prep_r_ext(h));  // eq_r is fixed while left variables are bound
          prepped = true;
        }
        const unsigned g = (unsigned)(q / chunk);
        TRY(eq_pass_ext(h, pend, q, chunk, (unsigned)m_log, g));
        if (pend) {
          L = Lv;
          h->len = Lv;
          h->in_work = true;
        }
        const size_t h0 = h->eq_l_len >> 1, h1 = h0 >> 1;
        const bool in_reduce = h0 <= fold_in_reduce;
        TRY(eq_reduce_fs2_ext(h, g, j, in_reduce ? h0 : 0, in_reduce ? h1 : 0));
        if (!in_reduce) {
          Views<F, 1> t1{};
          t1.t[0] = eql;
          hipLaunchKernelGGL((k_fold<F, F, 1>), dim3(blocks_for(h, h0)), dim3(kBlock), 0, h->stream, t1, t1,
                             gl::zero<F>(), h->d_cur, h0, h0, h->w);
          HIPCHK(hipGetLastError());
          hipLaunchKernelGGL((k_fold<F, F, 1>), dim3(blocks_for(h, h1)), dim3(kBlock), 0, h->stream, t1, t1,
                             gl::zero<F>(), h->d_cur + E, h1, h1, h->w);
          HIPCHK(hipGetLastError());
        }
        h->eq_l_len = h1;
        j += 2;
        pend = 2;
      }
      if (pend == 2) {
        // the older pending challenge folds p with a multi-block fold; the
        // newer one stays pending for the rounds below
        Views<F, 1> a{}, b{};
        a.t[0] = View<F>{src_planes(h)[0], src_stride(h)};
        b.t[0] = View<F>{h->work[0], h->work_stride};
        hipLaunchKernelGGL((k_fold<F, F, 1>), dim3(blocks_for(h, L / 2)), dim3(kBlock), 0, h->stream, a, b,
                           gl::zero<F>(), h->d_cur, L / 2, L / 2, h->w);
        HIPCHK(hipGetLastError());
        L /= 2;
        h->len = L;
        h->in_work = true;
        have = true;
        cur_in = h->d_cur + E;
      }
    }
  }
  while (j < v && L > tail && L >= 4) {
    const int left = j < v_l;
    size_t m_log = 0;
    while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
    const size_t units = have ? L / 4 : L / 2;
    // split-value form in the left phase: chunks of at least 256 pairs inside
    // one block of 2^m
    size_t chunk = kBlock;
    while (units / chunk > (size_t)h->grid_max * kPassGridScale) chunk *= 2;
    const bool sv = left && chunk <= ((size_t)1 << m_log) && units % chunk == 0 &&
                    !std::getenv("EFFSC_HIP_EQ_DIRECT");
    const unsigned g = sv ? (unsigned)(units / chunk) : blocks_for(h, units);
    const View<FI> pin{h->in_work ? nullptr : h->in[0], h->in_stride};
    const View<F> pw{h->work[0], h->work_stride};
    TRY(prof_begin(h));
    if (sv && have && h->in_work) {
      hipLaunchKernelGGL((k_eq_round_sv<F, F>), dim3(g), dim3(kBlock), 0, h->stream, pw, pw, L, 1, cur_in, eql,
                         h->eq_l_len, eqr, m_log, chunk, h->w, h->red);
    } else if (sv) {
      hipLaunchKernelGGL((k_eq_round_sv<FI, F>), dim3(g), dim3(kBlock), 0, h->stream, pin, pw, L, have ? 1 : 0, cur_in,
                         eql, h->eq_l_len, eqr, m_log, chunk, h->w, h->red);
    } else if (have && h->in_work) {
      hipLaunchKernelGGL((k_eq_round<F, F>), dim3(g), dim3(kBlock), 0, h->stream, pw, pw, L, 1, cur_in, eql,
                         h->eq_l_len, eqr, m_log, left, h->w, h->red);
    } else {
      hipLaunchKernelGGL((k_eq_round<FI, F>), dim3(g), dim3(kBlock), 0, h->stream, pin, pw, L, have ? 1 : 0,
                         cur_in, eql, h->eq_l_len, eqr, m_log, left, h->w, h->red);
    }
    HIPCHK(hipGetLastError());
    {
      // algorithmic bytes: p read once (and its folded half written); the
      // eq tables are L2-resident.  Products per evaluated pair: 2, plus 2
      // for the left-phase weights; 2 folds per pair when folding.
      const int din = (have && h->in_work) ? F::D : FI::D;
      const double pairs = (double)units;
      const double bytes = have ? 8.0 * (4.0 * pairs * din + 2.0 * pairs * F::D) : 8.0 * 2.0 * pairs * din;
      const double muls = pairs * ((left && !sv ? 4.0 : 2.0) + (have ? 2.0 : 0.0)) * mul_cost<F>();
      TRY(prof_end(h, have ? 0 : 1, bytes, muls));
    }
    if (have) {
      L /= 2;
      h->in_work = true;
    }
    size_t* tl = left ? &h->eq_l_len : &h->eq_r_len;
    View<F> tv = left ? eql : eqr;
    const size_t half = *tl >> 1;
    hipLaunchKernelGGL(k_eq_reduce_fs<F>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, eql, left ? 0 : 1,
                       h->d_fs, E, h->d_polys + j * 2 * E, h->d_chals + j * E, h->d_cur, tv,
                       half <= fold_in_reduce ? half : 0, h->w);
    HIPCHK(hipGetLastError());
    if (half > fold_in_reduce) {
      Views<F, 1> t1{};
      t1.t[0] = tv;
      hipLaunchKernelGGL((k_fold<F, F, 1>), dim3(blocks_for(h, half)), dim3(kBlock), 0, h->stream, t1, t1,
                         gl::zero<F>(), h->d_cur, half, half, h->w);
      HIPCHK(hipGetLastError());
    }
    *tl = half;
    have = true;
    cur_in = h->d_cur;  // k_eq_reduce_fs wrote this round's challenge there
    j++;
  }
  if constexpr (std::is_same<F, Fp3>::value) {
    // the loop above leaves L = 1024 with a pending fold whenever it ran a
    // fold (v >= 11)
    if (have && h->in_work && j < v && L == 1024 && !std::getenv("EFFSC_HIP_EQ_TAIL")) {
      // the remaining v - j = log2(L / 2) rounds in the F_p^3 register tail
      if (!h->eq_full) TRY(alloc_dev(&h->eq_full, (size_t)1 << 14));
      const size_t L1 = L / 2;
      size_t m_log = 0;
      while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
      const View<Fp3> eqf{h->eq_full, 1024};  // three planes of 2 L1 = 1024 elements
      hipLaunchKernelGGL(k_eq_expand_dup, dim3(blocks_for(h, L1)), dim3(kBlock), 0, h->stream, eql, eqr,
                         (unsigned)m_log, L1, eqf, h->w);
      HIPCHK(hipGetLastError());
      Views<Fp3, 2> tv{};
      tv.t[0] = View<Fp3>{h->work[0], h->work_stride};
      tv.t[1] = eqf;
      const int rounds = (int)(v - j);
      hipLaunchKernelGGL((k_tail_reg3<2, 2>), dim3(1), dim3(kTailReg3Threads), 0, h->stream, tv, rounds, h->d_fs,
                         h->d_polys + j * 2 * 3, h->d_chals + j * 3, cur_in, h->d_fin, h->w);
      HIPCHK(hipGetLastError());
      const int left = j < v_l ? (int)(v_l - j) : 0;
      hipLaunchKernelGGL(k_eq_post_fold3, dim3(1), dim3(kReduceBlock), 0, h->stream, eql, h->eq_l_len, eqr,
                         h->eq_r_len, h->d_chals + j * 3, rounds, left, h->d_cur, h->w);
      HIPCHK(hipGetLastError());
      h->len = 1;
      h->eq_l_len = 1;
      h->eq_r_len = 1;
      h->rounds_elapsed = v;
      return EFFSC_HIP_OK;
    }
  }
  // remaining rounds, finalize and final value
  const View<FI> pin{h->in[0], h->in_stride};
  const View<F> pw{h->work[0], h->work_stride};
  hipLaunchKernelGGL((k_eq_tail<FI, F>), dim3(1), dim3(eq_tail_block<F>()), 0, h->stream, pin, pw,
                     h->in_work ? 0 : 1, L, (int)j, (int)v, (int)v_l, eql, h->eq_l_len, eqr, h->eq_r_len,
                     have ? 1 : 0, cur_in, h->d_fs, E, h->d_polys + j * 2 * E, h->d_chals + j * E, h->d_cur,
                     h->d_fin, h->w);
  HIPCHK(hipGetLastError());
  // post-state: every variable bound
  h->len = 1;
  h->in_work = true;
  h->eq_l_len = 1;
  h->eq_r_len = 1;
  h->rounds_elapsed = v;
  return EFFSC_HIP_OK;
}

int prove_device_eq(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                    uint64_t* challenges, uint64_t* final_value) {
  TRY(ensure_dev_bufs(h));
  const int E = h->ext;
  DevFs* hs = reinterpret_cast<DevFs*>(h->h_proof + kMaxRounds * kMaxK * 3 + kMaxRounds * 3 + 4);
  std::memset(hs, 0, sizeof(DevFs));
  hs->kind = (uint32_t)t->kind;
  if (t->kind == 0) {
    for (int i = 0; i < 25; i++) hs->st[i] = t->fs.sp.st[i];
    hs->pos = t->fs.sp.pos;
  }
  hs->rng = t->rng;
  // a pinned staging is read by the first transcript kernel itself (fs_in);
  // otherwise the state is uploaded now
  h->fs_pending = hs;
  h->proof_copied = false;
  h->proof_by_tail = false;
  if (!h->h_proof_pinned) TRY(fs_flush(h));
  if (E == 1)
    TRY((eq_device_run<Fp, Fp>(h, num_rounds)));
  else if (E == 2)
    TRY((eq_device_run<Fp2, Fp2>(h, num_rounds)));
  else
    TRY((eq_device_run<Fp3, Fp3>(h, num_rounds)));
  uint64_t* hp = h->h_proof;
  uint64_t* hc = hp + kMaxRounds * kMaxK * 3;
  uint64_t* hf = hc + kMaxRounds * 3;
  if (h->proof_copied) {
    // post-state work follows the copy on the stream; the proof is ready
    // once the copy is
    h->proof_copied = false;
    HIPCHK(hipEventSynchronize(h->ev_proof));
  } else {
    TRY(copy_proof_out(h));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  TRY(prof_collect(h));
  std::copy(hp, hp + num_rounds * 2 * E, round_polys);
  std::copy(hc, hc + num_rounds * E, challenges);
  std::copy(hf, hf + E, final_value);
  if (t->kind == 0) {
    for (int i = 0; i < 25; i++) t->fs.sp.st[i] = hs->st[i];
    t->fs.sp.pos = hs->pos;
  }
  t->rng = hs->rng;
  return EFFSC_HIP_OK;
}

}  // namespace pv
}  // namespace effsc_hip

extern "C" {

int effsc_hip_eq_factored_new(const effsc_hip_config* cfg, const uint64_t* point, size_t v, const uint64_t* p,
                              size_t n, effsc_hip_prover** out) {
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (v >= 48) return fail(EFFSC_HIP_EINVAL, "too many variables");
  if (v > 0 && !point) return fail(EFFSC_HIP_EINVAL, "null point");
  if (n > 0 && !p) return fail(EFFSC_HIP_EINVAL, "null p");
  if (n > ((size_t)1 << v))
    return fail(EFFSC_HIP_EINVAL, "p_evals length exceeds 2^v");  // src/provers/eq_factored.rs:93-99
  if (cfg->input_ext != cfg->ext) return fail(EFFSC_HIP_EINVAL, "eq-factored input must be in the prover's field");
  const bool borrow = (cfg->flags & EFFSC_HIP_BORROW_INPUT) != 0;
  if (borrow && n != ((size_t)1 << v))
    return fail(EFFSC_HIP_EINVAL, "a borrowed p must hold exactly 2^v elements (no padding)");
  if (borrow && (!p || reinterpret_cast<uintptr_t>(p) % 16))
    return fail(EFFSC_HIP_EINVAL, "borrowed table null or not 16-byte aligned");
  auto* h = new effsc_hip_prover();
  h->kind = K_EQ;
  h->ext = h->in_ext = cfg->ext;
  h->K = 1;
  h->deg = 2;
  h->repr = cfg->repr;
  h->dev = cfg->device;
  h->w = cfg->ext > 1 ? gl::canon(cfg->ext_w) : 0;
  size_t N = (size_t)1 << v;
  h->n = N;
  h->len = N;
  h->v = v;
  h->v_l = v / 2;
  int rc = setup_common(h, N >> 1);
  if (rc == EFFSC_HIP_OK && borrow) {
    h->own_in = false;
    h->in_stride = N;
    h->in[0] = const_cast<uint64_t*>(p);
  }
  // p is zero-padded to 2^v: import the given prefix, clear the rest
  if (rc == EFFSC_HIP_OK && !borrow) {
    h->own_in = true;
    h->in_stride = round_up(N, 64);
    rc = alloc_dev(&h->in[0], h->in_stride * h->ext);
  }
  if (rc == EFFSC_HIP_OK && !borrow) {
    hipError_t e = hipMemsetAsync(h->in[0], 0, h->in_stride * h->ext * 8, h->stream);
    if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
  }
  uint64_t* staging = nullptr;
  if (rc == EFFSC_HIP_OK && !borrow) rc = alloc_dev(&staging, std::max<size_t>(n, 1) * h->ext + 1);
  if (rc == EFFSC_HIP_OK) rc = alloc_dev(&h->eq_point, std::max<size_t>(v, 1) * h->ext);
  if (rc == EFFSC_HIP_OK && n && !borrow) {
    hipError_t e = hipMemcpyAsync(staging, p, n * h->ext * 8, hipMemcpyDefault, h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_import, dim3(blocks_for(h, n)), dim3(kBlock), 0, h->stream, h->in[0], h->in_stride,
                         staging, n, h->ext, h->repr == EFFSC_HIP_MONT64 ? 1 : 0);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, std::string("import p: ") + hipGetErrorString(e));
  }
  if (rc == EFFSC_HIP_OK) {
    // the point, canonical AoS on the device
    std::vector<uint64_t> wc(std::max<size_t>(v, 1) * h->ext, 0);
    for (size_t i = 0; i < v * h->ext; i++) {
      uint64_t x = point[i];
      wc[i] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(x, gl::MONT_TO_CANON) : gl::canon(x);
    }
    hipError_t e = hipMemcpyAsync(h->eq_point, wc.data(), wc.size() * 8, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
    if (rc == EFFSC_HIP_OK) rc = build_eq(h, h->eq_point, 0, h->v_l, &h->eq_l, &h->eq_l_stride, &h->eq_l_len);
    if (rc == EFFSC_HIP_OK)
      rc = build_eq(h, h->eq_point, h->v_l, v - h->v_l, &h->eq_r, &h->eq_r_stride, &h->eq_r_len);
    if (rc == EFFSC_HIP_OK) {
      e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
    }
  }
  if (staging) (void)hipFree(staging);
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  *out = h;
  return EFFSC_HIP_OK;
}

}  // extern "C"

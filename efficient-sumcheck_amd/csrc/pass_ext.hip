// Two-round passes over the extension fields F_p^2 and F_p^3 for the
// multilinear (K = 1) and inner-product (K = 2) provers in the device loop.
//
// The same algebra as the base-field passes (pass2.hpp): the round-(s+1)
// values are polynomials of degree K in the round-s challenge x, so one pass
// over the level-s table gives round s's values and the K + 1 coefficients of
// each round-(s+1) value, and the next pass folds the stored level with both
// challenges on the fly, materialising every other level only.  Here the
// table elements, challenges and sums are extension elements (inputs may be
// base-field words, lifted).  The raw sums are kept reduced (one F product
// per raw sum and unit); k_reduce_fs2_ext sums the workgroups' values and runs
// the two transcript steps.  All outputs are the same field elements as one
// round per launch.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

namespace effsc_hip {
namespace pv {
namespace {

// Level-L element e from the level-M table (M = 2^P L) with P pending folds
// c1 (the older) and c2.
template <int P, class FI, class F>
__device__ __forceinline__ F chain(const View<FI>& t, size_t e, size_t L, const F& c1, const F& c2, uint64_t w) {
  if constexpr (P == 0) {
    return gl::lift_to<F>(t.ld(e));
  } else {
    const FI x0 = t.ld(e), x1 = t.ld(e + L), x2 = t.ld(e + 2 * L), x3 = t.ld(e + 3 * L);
    const F a = fold1<F>(x0, x2, c1, w), b = fold1<F>(x1, x3, c1, w);
    return fold1<F>(a, b, c2, w);
  }
}

// Two consecutive level-L elements e, e + 1 (e even, 16-byte aligned planes).
template <int P, class FI, class F>
__device__ __forceinline__ void chain2(const View<FI>& t, size_t e, size_t L, const F& c1, const F& c2, uint64_t w,
                                       F& o0, F& o1) {
  if constexpr (P == 0) {
    FI x0, x1;
    t.ld2(e, x0, x1);
    o0 = gl::lift_to<F>(x0);
    o1 = gl::lift_to<F>(x1);
  } else {
    FI x0[2], x1[2], x2[2], x3[2];
    t.ld2(e, x0[0], x0[1]);
    t.ld2(e + L, x1[0], x1[1]);
    t.ld2(e + 2 * L, x2[0], x2[1]);
    t.ld2(e + 3 * L, x3[0], x3[1]);
    o0 = fold1<F>(fold1<F>(x0[0], x2[0], c1, w), fold1<F>(x1[0], x3[0], c1, w), c2, w);
    o1 = fold1<F>(fold1<F>(x0[1], x2[1], c1, w), fold1<F>(x1[1], x3[1], c1, w), c2, w);
  }
}

// The multilinear pass with two units per thread step and 16-byte loads and
// stores (q even, strides even).
template <class FI, class F, int P>
__global__ __launch_bounds__(kBlock) void k_pass_ext_ml2(View<FI> src, View<F> dst, const uint64_t* cptr, size_t q,
                                                         uint64_t w, Reduce r) {
  F c1 = gl::zero<F>(), c2 = gl::zero<F>();
  if constexpr (P > 0) {
    c1 = gl::load_words<F>(cptr);
    c2 = gl::load_words<F>(cptr + F::D);
  }
  const size_t L = 4 * q;
  F s[3] = {gl::zero<F>(), gl::zero<F>(), gl::zero<F>()};
  for (size_t u = gtid() * 2; u < q; u += (size_t)gridDim.x * blockDim.x * 2) {
    F E0[4], E1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) chain2<P>(src, u + j * q, L, c1, c2, w, E0[j], E1[j]);
    if constexpr (P > 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) dst.st2(u + j * q, E0[j], E1[j]);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) s[j] = gl::add(s[j], gl::add(E0[j], E1[j]));
  }
  F v[3] = {gl::add(s[0], s[1]), s[0], gl::sub(s[2], s[0])};
  finish<F, 3>(v, r);
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

// One pass over K tables: P pending folds from the level-M tables `src`
// (field FI) to level L = 4 q, written to `dst` (field F) when P > 0.  Unit u
// owns E_j = T_L[u + j q], j < 4; round s pairs (E0, E2) and (E1, E3), round
// s + 1 pairs (E0 + x (E2 - E0), E1 + x (E3 - E1)).  Raw sums per unit:
// K = 1: E0, E1, E2; K = 2 (A from table 0, B from table 1):
//   A0 B0, A1 B1, dA0 dB0, dA1 dB1, A2 B2, eA eB, (A3 - A2)(B3 - B2), fA fB
// with dA0 = A2 - A0, dA1 = A3 - A1, eA = A1 - A0, fA = dA1 - dA0.
// Challenges: 2 D words at cptr (c1 then c2).
// At most 256 VGPRs (two waves per SIMD): unconstrained, the compiler issued
// every level-M load of a unit at once and ran at one wave per SIMD.
template <class FI, class F, int K, int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_pass_ext(
    Views<FI, K> src, Views<F, K> dst, const uint64_t* cptr, size_t q, uint64_t w, Reduce r) {
  constexpr int NR = pass_raw<K>(), NV = pass_values<K>();
  F c1 = gl::zero<F>(), c2 = gl::zero<F>();
  if constexpr (P > 0) {
    c1 = gl::load_words<F>(cptr);
    c2 = gl::load_words<F>(cptr + F::D);
  }
  const size_t L = 4 * q;
  F s[NR];
#pragma unroll
  for (int i = 0; i < NR; i++) s[i] = gl::zero<F>();
  for (size_t u = gtid(); u < q; u += (size_t)gridDim.x * blockDim.x) {
    F E[K][4];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) E[i][j] = chain<P>(src.t[i], u + j * q, L, c1, c2, w);
    if constexpr (P > 0) {
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) dst.t[i].st(u + j * q, E[i][j]);
    }
    if constexpr (K == 1) {
      s[0] = gl::add(s[0], E[0][0]);
      s[1] = gl::add(s[1], E[0][1]);
      s[2] = gl::add(s[2], E[0][2]);
    } else {
      const F* A = E[0];
      const F* B = E[1];
      const F dA0 = gl::sub(A[2], A[0]), dA1 = gl::sub(A[3], A[1]), dB0 = gl::sub(B[2], B[0]),
              dB1 = gl::sub(B[3], B[1]);
      const F eA = gl::sub(A[1], A[0]), eB = gl::sub(B[1], B[0]);
      s[0] = gl::add(s[0], gl::mul(A[0], B[0], w));
      s[1] = gl::add(s[1], gl::mul(A[1], B[1], w));
      s[2] = gl::add(s[2], gl::mul(dA0, dB0, w));
      s[3] = gl::add(s[3], gl::mul(dA1, dB1, w));
      s[4] = gl::add(s[4], gl::mul(A[2], B[2], w));
      s[5] = gl::add(s[5], gl::mul(eA, eB, w));
      s[6] = gl::add(s[6], gl::mul(gl::sub(A[3], A[2]), gl::sub(B[3], B[2]), w));
      s[7] = gl::add(s[7], gl::mul(gl::sub(dA1, dA0), gl::sub(dB1, dB0), w));
    }
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

// Sums the pass values of `nblocks` workgroups and runs the two transcript
// steps by wave 0: round s values, challenge x, round s + 1 values (the pass
// polynomials at x), challenge y.  poly receives 2 K elements, chals (x, y),
// cur (x, y): the pending folds of the next pass.
template <class F, int K>
__global__ __launch_bounds__(kReduceBlock) void k_reduce_fs2_ext(Reduce r, unsigned nblocks, DevFs* t, uint64_t* poly,
                                                                 uint64_t* chals, uint64_t* cur, uint64_t w) {
  constexpr int NV = pass_values<K>(), D = F::D;
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  F s[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = gl::zero<F>();
  sum_partials<F, NV>(r.partials, nblocks, s);
  block_reduce<F, NV>(s);
  if (threadIdx.x >= 64) return;
  const bool lane0 = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = uniform_elem(s[i]);
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
  const F X = gl::load_words<F>(x);
#pragma unroll
  for (int i = 0; i < K; i++) {
    gl::store_words(eval_next<F, K>(s, i, X, w), wd);
    if (lane0)
#pragma unroll
      for (int c = 0; c < D; c++) poly[(K + i) * D + c] = wd[c];
    fs.send(D, wd);
  }
  uint64_t y[3];
  fs.challenge(D, y);
  if (lane0) {
#pragma unroll
    for (int c = 0; c < D; c++) {
      chals[c] = x[c];
      chals[D + c] = y[c];
      cur[c] = x[c];
      cur[D + c] = y[c];
    }
  }
  fs.store(t);
}

template <class FI, class F, int K, int P>
int launch_pass_ext_t(effsc_hip_prover* h, size_t q, const uint64_t* cptr) {
  const Views<FI, K> in = views<FI, K>(src_planes(h), src_stride(h));
  const Views<F, K> out = views<F, K>(h->work, h->work_stride);
  const bool vec = K == 1 && q % 2 == 0 && src_stride(h) % 2 == 0 && h->work_stride % 2 == 0;
  const unsigned g = blocks_for(h, vec ? q / 2 : q);
  TRY(prof_begin(h));
  if (vec)
    hipLaunchKernelGGL((k_pass_ext_ml2<FI, F, P>), dim3(g), dim3(kBlock), 0, h->stream, in.t[0], out.t[0], cptr, q,
                       h->w, h->red);
  else
    hipLaunchKernelGGL((k_pass_ext<FI, F, K, P>), dim3(g), dim3(kBlock), 0, h->stream, in, out, cptr, q, h->w, h->red);
  HIPCHK(hipGetLastError());
  const double L = 4.0 * (double)q;
  // read the stored level (2^P L elements), write level L when P > 0
  const double bytes = 8.0 * K * ((double)(1 << P) * L * FI::D + (P > 0 ? L * F::D : 0.0));
  const double muls = (double)q * (4.0 * K * (P == 0 ? 0.0 : 3.0 * mul_cost<F>()) + (K == 2 ? 8.0 * mul_cost<F>() : 0.0));
  TRY(prof_end(h, 3, bytes, muls));
  h->last_grid = g;
  return EFFSC_HIP_OK;
}

}  // namespace

// Used over F_p^2, and for the multilinear prover over F_p^3 (measured with
// the SHAKE transcript, v = 26 over F_p^2: inner product 2.41 -> 2.12 ms,
// multilinear 1.06 -> 0.88 ms; v = 24 over F_p^3: multilinear 0.669 -> 0.644
// ms).  The F_p^3 inner-product pass (8 F_p^3 products per unit at two waves
// per SIMD) measured 1.349 against 1.307 ms for one round per launch, so it
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

}  // namespace pv
}  // namespace effsc_hip

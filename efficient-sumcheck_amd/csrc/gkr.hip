// GkrProver (src/provers/gkr.rs:38-202): per-round path, device-resident
// proofs and the constructor.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

// ===========================================================================
// GkrProver: f(b, c) = add(b, c) (W(b) + W(c)) + mult(b, c) W(b) W(c) over
// (b, c) in {0,1}^{2k} (src/provers/gkr.rs:38-202).  Four tables of n^2
// elements as in the reference: add, mult and the witness broadcast over the
// hypercube, w_b[b n + c] = W(b) and w_c[b n + c] = W(c) (built on the device
// from W).  Every round folds the four tables with ip::fold semantics
// (k_fold, K = 4) and evaluates q(0), q(1), q(2) over the half-split pairs
// (k_gkr_eval); the wire's second value is q(inf) = (q(0) + q(2) - 2 q(1)) / 2
// (src/provers/gkr.rs:176-181).
// ===========================================================================
namespace effsc_hip {
namespace pv {

template <class F>
__global__ __launch_bounds__(kBlock) void k_gkr_broadcast(View<F> wit, size_t n, View<F> wb, View<F> wc) {
  const size_t nbc = n * n;
  for (size_t i = gtid(); i < nbc; i += (size_t)gridDim.x * blockDim.x) {
    wb.st(i, wit.ld(i / n));
    wc.st(i, wit.ld(i % n));
  }
}

template <class F>
__host__ __device__ __forceinline__ F gkr_f(F a, F m, F b, F c, uint64_t w) {
  return gl::add(gl::mul(a, gl::add(b, c), w), gl::mul(m, gl::mul(b, c, w), w));
}

// q(0), q(1), q(2) of the current tables; pairs k >= paired have hi = 0
// (src/provers/gkr.rs:121-171).
template <class F>
__global__ __launch_bounds__(kBlock) void k_gkr_eval(Views<F, 4> t, size_t half, size_t paired, uint64_t w,
                                                     Reduce r) {
  F acc[3] = {gl::zero<F>(), gl::zero<F>(), gl::zero<F>()};
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < half; k += step) {
    F lo[4], hi[4], x2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      lo[j] = t.t[j].ld(k);
      hi[j] = k < paired ? t.t[j].ld(k + half) : gl::zero<F>();
      x2[j] = gl::sub(gl::add(hi[j], hi[j]), lo[j]);
    }
    acc[0] = gl::add(acc[0], gkr_f(lo[0], lo[1], lo[2], lo[3], w));
    acc[1] = gl::add(acc[1], gkr_f(hi[0], hi[1], hi[2], hi[3], w));
    acc[2] = gl::add(acc[2], gkr_f(x2[0], x2[1], x2[2], x2[3], w));
  }
  finish<F, 3>(acc, r);
}

// f at element 0 of the four tables (lifted words), as ext words.
void gkr_f_host(const effsc_hip_prover* h, const uint64_t* el, uint64_t* out) {
  const int E = h->ext;
  (void)with_field(E, [&](auto f) -> int {
    using F = decltype(f);
    const F v = gkr_f<F>(from_words<F>(el), from_words<F>(el + E), from_words<F>(el + 2 * E),
                         from_words<F>(el + 3 * E), h->w);
    gl::store_words(v, out);
    return EFFSC_HIP_OK;
  });
}

int gkr_eval(effsc_hip_prover* h, uint64_t* out) {
  const int E = h->ext;
  std::fill(out, out + 2 * E, 0);
  const size_t L = h->len;
  if (L <= 1) {
    if (L == 1) {
      uint64_t el[kMaxK * 3];
      TRY(read_elements(h, 0, el));
      gkr_f_host(h, el, out);
    }
    return EFFSC_HIP_OK;
  }
  size_t half, paired;
  geometry(h, L, &half, &paired);
  const int fi = src_field(h);
  uint64_t* const* src = src_planes(h);
  const size_t sst = src_stride(h);
  const unsigned g = blocks_for(h, half);
  TRY(with_field(fi, [&](auto f) -> int {
    using F = decltype(f);
    hipLaunchKernelGGL(k_gkr_eval<F>, dim3(g), dim3(kBlock), 0, h->stream, views<F, 4>(src, sst), half, paired, h->w,
                       h->red);
    HIPCHK(hipGetLastError());
    return reduce<F, 3>(h, g);
  }));
  uint64_t q[3 * 3];
  TRY(fetch_result(h, fi, q, 3));  // lifted to E words each
  const uint64_t inv2 = (gl::P + 1) / 2;
  for (int c = 0; c < E; c++) {
    uint64_t q0 = q[c], q1 = q[E + c], q2 = q[2 * E + c];
    out[c] = q0;
    out[E + c] = gl::mul(gl::sub(gl::add(q0, q2), gl::add(q1, q1)), inv2);
  }
  return EFFSC_HIP_OK;
}

// Device-resident GKR rounds.  k_gkr_fused folds the four tables with the
// device challenge (quadruples k, k + q, k + 2q, k + 3q, q = L/4, as
// k_fused) and evaluates q(0), q(1), q(2) of the folded pairs;
// k_gkr_reduce_fs sums the partials, writes the wire [q(0), q(inf)] and runs
// the transcript step; k_gkr_tail runs the remaining rounds, the final fold
// and f at the last element in one workgroup.
template <class FI, class F>
__global__ __launch_bounds__(kBlock) void k_gkr_fused(Views<FI, 4> in, Views<F, 4> out, const uint64_t* cptr, size_t q,
                                                      uint64_t w, Reduce r) {
  const F c = gl::load_words<F>(cptr);
  F acc[3] = {gl::zero<F>(), gl::zero<F>(), gl::zero<F>()};
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < q; k += step) {
    F lo[4], hi[4], x2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const FI x0 = in.t[j].ld(k), x1 = in.t[j].ld(k + q), y0 = in.t[j].ld(k + 2 * q), y1 = in.t[j].ld(k + 3 * q);
      lo[j] = fold1<F>(x0, y0, c, w);
      hi[j] = fold1<F>(x1, y1, c, w);
      out.t[j].st(k, lo[j]);
      out.t[j].st(k + q, hi[j]);
      x2[j] = gl::sub(gl::add(hi[j], hi[j]), lo[j]);
    }
    acc[0] = gl::add(acc[0], gkr_f(lo[0], lo[1], lo[2], lo[3], w));
    acc[1] = gl::add(acc[1], gkr_f(hi[0], hi[1], hi[2], hi[3], w));
    acc[2] = gl::add(acc[2], gkr_f(x2[0], x2[1], x2[2], x2[3], w));
  }
  finish<F, 3>(acc, r);
}

// q(0), q(1), q(2) -> the wire [q(0), q(inf)], q(inf) = (q(0) + q(2) - 2 q(1)) / 2
template <class F>
__device__ __forceinline__ void gkr_wire(const F* q, F* wire) {
  wire[0] = q[0];
  wire[1] = gl::mul(gl::sub(gl::add(q[0], q[2]), gl::add(q[1], q[1])), Fp{(gl::P + 1) / 2}, 0);
}

template <class F>
__global__ __launch_bounds__(kReduceBlock) void k_gkr_reduce_fs(Reduce r, unsigned nblocks, DevFs* t, int ext,
                                                                uint64_t* poly, uint64_t* chal_out, uint64_t* chal_cur) {
  F s[3] = {gl::zero<F>(), gl::zero<F>(), gl::zero<F>()};
  sum_partials<F, 3>(r.partials, nblocks, s);
  block_reduce<F, 3>(s);
  if (threadIdx.x >= 64) return;
  F wire[2];
  gkr_wire<F>(s, wire);
  emit_round<F, 2>(wire, ext, t, poly, chal_out, chal_cur);
}

template <class FI, class F>
__global__ __launch_bounds__((tail_block<F, 4>())) void k_gkr_tail(Views<FI, 4> in, Views<F, 4> work, int from_input,
                                                                  size_t L, int rounds, int have_chal, DevFs* t,
                                                                  int ext, uint64_t* polys, uint64_t* chals,
                                                                  const uint64_t* chal_cur, uint64_t* fin, uint64_t w) {
  __shared__ uint64_t chal_sh[3];
  if (threadIdx.x < 3) chal_sh[threadIdx.x] = have_chal ? chal_cur[threadIdx.x] : 0;
  // the sponge stays in wave 0's registers for the whole tail
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  __syncthreads();
  const unsigned tid = threadIdx.x, nt = blockDim.x;
  bool src_in = from_input != 0;
  for (int r = 0; r < rounds; r++) {
    F acc[3] = {gl::zero<F>(), gl::zero<F>(), gl::zero<F>()};
    F lo[4], hi[4], x2[4];
    if (have_chal) {
      const F c = gl::load_words<F>(chal_sh);
      const size_t q = L / 4;
      for (size_t k = tid; k < q; k += nt) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (src_in) {
            lo[j] = fold1<F>(in.t[j].ld(k), in.t[j].ld(k + 2 * q), c, w);
            hi[j] = fold1<F>(in.t[j].ld(k + q), in.t[j].ld(k + 3 * q), c, w);
          } else {
            lo[j] = fold1<F>(work.t[j].ld(k), work.t[j].ld(k + 2 * q), c, w);
            hi[j] = fold1<F>(work.t[j].ld(k + q), work.t[j].ld(k + 3 * q), c, w);
          }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
          work.t[j].st(k, lo[j]);
          work.t[j].st(k + q, hi[j]);
          x2[j] = gl::sub(gl::add(hi[j], hi[j]), lo[j]);
        }
        acc[0] = gl::add(acc[0], gkr_f(lo[0], lo[1], lo[2], lo[3], w));
        acc[1] = gl::add(acc[1], gkr_f(hi[0], hi[1], hi[2], hi[3], w));
        acc[2] = gl::add(acc[2], gkr_f(x2[0], x2[1], x2[2], x2[3], w));
      }
      L /= 2;
      src_in = false;
    } else {
      const size_t half = L / 2;
      for (size_t k = tid; k < half; k += nt) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          lo[j] = src_in ? gl::lift_to<F>(in.t[j].ld(k)) : work.t[j].ld(k);
          hi[j] = src_in ? gl::lift_to<F>(in.t[j].ld(k + half)) : work.t[j].ld(k + half);
          x2[j] = gl::sub(gl::add(hi[j], hi[j]), lo[j]);
        }
        acc[0] = gl::add(acc[0], gkr_f(lo[0], lo[1], lo[2], lo[3], w));
        acc[1] = gl::add(acc[1], gkr_f(hi[0], hi[1], hi[2], hi[3], w));
        acc[2] = gl::add(acc[2], gkr_f(x2[0], x2[1], x2[2], x2[3], w));
      }
    }
    block_reduce<F, 3>(acc);
    if (tid < 64) {
      F wire[2];
      gkr_wire<F>(acc, wire);
      emit_round_fs<F, 2>(wire, ext, fs, polys + (size_t)r * 2 * ext, chals + (size_t)r * ext, chal_sh);
    }
    __syncthreads();
    have_chal = 1;
  }
  // finalize: fold with the last challenge (src/runner.rs:57-60)
  if (have_chal && L >= 2) {
    const F c = gl::load_words<F>(chal_sh);
    const size_t half = L / 2;
    for (size_t k = tid; k < half; k += nt) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const F v = src_in ? fold1<F>(in.t[j].ld(k), in.t[j].ld(k + half), c, w)
                           : fold1<F>(work.t[j].ld(k), work.t[j].ld(k + half), c, w);
        work.t[j].st(k, v);
      }
    }
    L = half;
    __syncthreads();
  }
  if (tid == 0) {
    // src/provers/gkr.rs:192-201: f at the single remaining point, else 0
    F v = gl::zero<F>();
    if (L == 1 && src_in)
      v = gkr_f(gl::lift_to<F>(in.t[0].ld(0)), gl::lift_to<F>(in.t[1].ld(0)), gl::lift_to<F>(in.t[2].ld(0)),
                gl::lift_to<F>(in.t[3].ld(0)), w);
    else if (L == 1)
      v = gkr_f(work.t[0].ld(0), work.t[1].ld(0), work.t[2].ld(0), work.t[3].ld(0), w);
    uint64_t wd[3] = {0, 0, 0};
    gl::store_words(v, wd);
    for (int c = 0; c < ext; c++) fin[c] = wd[c];
  }
  if (tid < 64) fs.store(t);
}

// Factored witness (n = 2^k).  While the b variables are bound (j <= k folds
// applied), the level-2^{2k-j} broadcast tables are w_b[x] = B_j[x >> k]
// and w_c[x] = W[x & (n - 1)], with B_0 = W and B_{j+1} the MSB fold of B_j:
// the pair (x, x + L/2) of w_b is (B_j[x >> k], B_j[(x >> k) + 2^{k-j-1}]),
// and the pair of w_c is two equal elements, which fold to the same element
// (lo + r (lo - lo) = lo).  So those rounds fold only add and mult through
// HBM, half the traffic of the four tables, and read B_j (2^{k-j}
// elements) and W (n elements) from cache.  The tables left behind are the
// same field elements as the four-table rounds.
//
// In every round of a power-of-two layer one witness table is constant over
// each pair (w_c while b is bound, w_b after), so with v that constant and u
// the other, f(t) = a(t) (u(t) + v) + m(t) u(t) v has degree 2 and
//   q(0)   = u0 (a0 + m0 v) + a0 v,
//   q(inf) = (u1 - u0) ((a1 - a0) + v (m1 - m0)),
// the leading coefficient, which equals the reference's
// (q(0) + q(2) - 2 q(1)) / 2 (src/provers/gkr.rs:176-181) as a field element:
// 5 products per pair instead of 9 for q(0), q(1), q(2).
template <class F>
__device__ __forceinline__ void gkr_pair(F a0, F a1, F m0, F m1, F u0, F u1, F v, uint64_t w, F& q0, F& qi) {
  q0 = gl::add(gl::mul(u0, gl::add(a0, gl::mul(m0, v, w)), w), gl::mul(a0, v, w));
  qi = gl::mul(gl::sub(u1, u0), gl::add(gl::sub(a1, a0), gl::mul(v, gl::sub(m1, m0), w)), w);
}

// k_gkr_eval_fw is round 0 over the input (w_c constant over the pairs);
// k_gkr_fused_fw folds add and mult with the pending challenge, writes
// B_{j+1} into the other of two planesets and evaluates the folded pairs,
// with w_b folded on the fly.  Both leave the wire [q(0), q(inf)] per
// workgroup.
// Two elements per thread: the GKR prover owns its planes, whose strides are
// multiples of 64 words, so (x, x + 1) with x even is one 16-byte load per
// plane, and with k >= 1 both share x >> k and (x + L/2) >> k.
template <class FI>
__global__ __launch_bounds__(kBlock) void k_gkr_eval_fw(Views<FI, 2> t, View<FI> wit, unsigned k, size_t half,
                                                        uint64_t w, Reduce r) {
  FI acc[2] = {gl::zero<FI>(), gl::zero<FI>()};
  const size_t mask = ((size_t)1 << k) - 1;
  const size_t step = (size_t)gridDim.x * blockDim.x * 2;
  for (size_t x = gtid() * 2; x < half; x += step) {
    FI a0[2], a1[2], m0[2], m1[2], v[2];
    t.t[0].ld2(x, a0[0], a0[1]);
    t.t[0].ld2(x + half, a1[0], a1[1]);
    t.t[1].ld2(x, m0[0], m0[1]);
    t.t[1].ld2(x + half, m1[0], m1[1]);
    wit.ld2(x & mask, v[0], v[1]);
    const FI u0 = wit.ld(x >> k), u1 = wit.ld((x + half) >> k);
#pragma unroll
    for (int e = 0; e < 2; e++) {
      FI q0, qi;
      gkr_pair(a0[e], a1[e], m0[e], m1[e], u0, u1, v[e], w, q0, qi);
      acc[0] = gl::add(acc[0], q0);
      acc[1] = gl::add(acc[1], qi);
    }
  }
  finish<FI, 2>(acc, r);
}

// The previous round's transcript step taken into a k_gkr_fused_fw launch
// (FS = true): the partial wires of the previous launch (rin, nin blocks),
// the sponge (from t_in, to t_out) and the proof words, written by
// workgroup 0 alone; the partials and the sponge alternate between two
// buffers, so no launch reads what it writes.
struct GkrStep {
  Reduce rin;
  unsigned nin;
  const DevFs* t_in;
  DevFs* t_out;
  uint64_t *poly, *chals, *cur;
};

// From level 4q with j folds applied to level 2q with j + 1 (j + 1 <= k):
// bsrc holds B_j (2 hb elements), bdst receives B_{j+1} (hb elements).
// cconst: the level-2q pairs bind a b variable (j + 1 < k), so w_c is the
// constant of each pair; else w_b is (B_k has one element).  With FS the
// launch first runs the previous round's transcript step (every workgroup
// derives the same challenge, as k_gkr_wire_fs), base field only.
template <class FI, class FW, class F, bool FS = false>
__global__ __launch_bounds__(kBlock) void k_gkr_fused_fw(Views<FI, 2> in, Views<F, 2> out, View<FI> bsrc,
                                                         View<F> bdst, size_t hb, View<FW> wit, unsigned k, int cconst,
                                                         const uint64_t* cptr, size_t q, uint64_t w, Reduce r,
                                                         GkrStep stp = {}) {
  F c;
  if constexpr (FS) {
    static_assert(F::D == 1, "fused transcript step: base field");
    __shared__ uint64_t sh[kBlock];
    __shared__ uint64_t csh[3];
    LaneFs fs;
    if (threadIdx.x < 64) fs.load(stp.t_in);
    sum_partials_words<2>(stp.rin.partials, stp.nin, sh);
    if (threadIdx.x < 64) {
      const F s2[2] = {F{sh[0]}, F{sh[1]}};
      const bool w0 = blockIdx.x == 0;
      emit_round_fs<F, 2>(s2, 1, fs, stp.poly, stp.chals, csh, w0);
      if (w0) {
        if ((threadIdx.x & 63) == 0) stp.cur[0] = csh[0];
        fs.store(stp.t_out);
      }
    }
    __syncthreads();
    c = gl::load_words<F>(csh);
  } else {
    c = gl::load_words<F>(cptr);
  }
  const size_t mask = ((size_t)1 << k) - 1;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < hb; i += step) bdst.st(i, fold1<F>(bsrc.ld(i), bsrc.ld(i + hb), c, w));
  F acc[2] = {gl::zero<F>(), gl::zero<F>()};
  for (size_t x = gtid() * 2; x < q; x += step * 2) {
    F lo[2][2], hi[2][2], cl[2], ch[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      FI x0[2], x1[2], y0[2], y1[2];
      in.t[j].ld2(x, x0[0], x0[1]);
      in.t[j].ld2(x + q, x1[0], x1[1]);
      in.t[j].ld2(x + 2 * q, y0[0], y0[1]);
      in.t[j].ld2(x + 3 * q, y1[0], y1[1]);
#pragma unroll
      for (int e = 0; e < 2; e++) {
        lo[j][e] = fold1<F>(x0[e], y0[e], c, w);
        hi[j][e] = fold1<F>(x1[e], y1[e], c, w);
      }
      out.t[j].st2(x, lo[j][0], lo[j][1]);
      out.t[j].st2(x + q, hi[j][0], hi[j][1]);
    }
    // positions x, x + 1 and x + q, x + q + 1 of the level-2q tables
    const size_t bl = x >> k;
    const F blo = fold1<F>(bsrc.ld(bl), bsrc.ld(bl + hb), c, w);
    FW w0, w1;
    wit.ld2(x & mask, w0, w1);
    cl[0] = gl::lift_to<F>(w0);
    cl[1] = gl::lift_to<F>(w1);
    if (cconst) {
      const size_t bh = (x + q) >> k;
      const F bhi = fold1<F>(bsrc.ld(bh), bsrc.ld(bh + hb), c, w);
#pragma unroll
      for (int e = 0; e < 2; e++) {
        F q0, qi;
        gkr_pair(lo[0][e], hi[0][e], lo[1][e], hi[1][e], blo, bhi, cl[e], w, q0, qi);
        acc[0] = gl::add(acc[0], q0);
        acc[1] = gl::add(acc[1], qi);
      }
    } else {
      wit.ld2((x + q) & mask, w0, w1);
      ch[0] = gl::lift_to<F>(w0);
      ch[1] = gl::lift_to<F>(w1);
#pragma unroll
      for (int e = 0; e < 2; e++) {
        F q0, qi;
        gkr_pair(lo[0][e], hi[0][e], lo[1][e], hi[1][e], cl[e], ch[e], blo, w, q0, qi);
        acc[0] = gl::add(acc[0], q0);
        acc[1] = gl::add(acc[1], qi);
      }
    }
  }
  finish<F, 2>(acc, r);
}

// Sums the per-workgroup wires and runs the transcript step.
template <class F>
// The sponge is loaded from t_in (t, or the other buffer after a fused
// launch) and stored to t.
__global__ __launch_bounds__(kReduceBlock) void k_gkr_wire_fs(Reduce r, unsigned nblocks, const DevFs* t_in, DevFs* t,
                                                              int ext, uint64_t* poly, uint64_t* chal_out,
                                                              uint64_t* chal_cur) {
  F s[2] = {gl::zero<F>(), gl::zero<F>()};
  sum_partials<F, 2>(r.partials, nblocks, s);
  block_reduce<F, 2>(s);
  if (threadIdx.x >= 64) return;
  LaneFs fs;
  fs.load(t_in);
  emit_round_fs<F, 2>(s, ext, fs, poly, chal_out, chal_cur);
  fs.store(t);
}

// The level-M broadcast tables from the factored witness (j >= 1 folds).
template <class FW, class F>
__global__ __launch_bounds__(kBlock) void k_gkr_expand(View<F> b, View<FW> wit, unsigned k, size_t M, View<F> wb,
                                                       View<F> wc) {
  const size_t mask = ((size_t)1 << k) - 1;
  for (size_t x = gtid(); x < M; x += (size_t)gridDim.x * blockDim.x) {
    wb.st(x, b.ld(x >> k));
    wc.st(x, gl::lift_to<F>(wit.ld(x & mask)));
  }
}

int gkr_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  // src/provers/gkr.rs:93-101: fold all four tables with the challenge
  if (chal) {
    bool ev = false;
    TRY(fold_step(h, chal, false, &ev));
  }
  return gkr_eval(h, out);
}

int gkr_final_value(effsc_hip_prover* h, uint64_t* out) {
  // src/provers/gkr.rs:192-201
  std::fill(out, out + h->ext, 0);
  if (h->len != 1) return EFFSC_HIP_OK;
  uint64_t el[kMaxK * 3];
  TRY(read_elements(h, 0, el));
  gkr_f_host(h, el, out);
  return EFFSC_HIP_OK;
}

}  // namespace pv
}  // namespace effsc_hip

namespace effsc_hip {
namespace pv {
// GkrProver proofs on the device (no hook, power-of-two tables): round
// launches while the tables are large, then one-workgroup rounds; the same
// wire, transcript and post-state as the per-round path.
bool gkr_device_applies(const effsc_hip_prover* h, size_t num_rounds) {
  if (h->kind != K_GKR || num_rounds == 0 || num_rounds > kMaxRounds || !is_pow2(h->len)) return false;
  size_t logl = 0;
  while (((size_t)1 << logl) < h->len) logl++;
  return num_rounds <= logl;
}

template <class FI, class F>
int gkr_device_run(effsc_hip_prover* h, size_t num_rounds, bool* tail_used) {
  const int E = h->ext;
  const size_t tail = F::D == 3 ? ((size_t)1 << 10) : F::D == 2 ? ((size_t)1 << 11) : ((size_t)1 << 12);
  size_t r = 0;
  bool have = false;
  // factored-witness rounds from the input of a power-of-two layer, then the
  // broadcast tables are written out at the level reached
  unsigned kb = 0;
  while (((size_t)1 << (2 * kb)) < h->n) kb++;
  if (!h->in_work && h->len == h->n && ((size_t)1 << (2 * kb)) == h->n && kb >= 1 && h->len > tail &&
      h->in_stride % 2 == 0 && h->work_stride % 2 == 0 && !std::getenv("EFFSC_HIP_GKR_BROADCAST")) {
    const size_t hcap = ((size_t)1 << kb) / 2;
    if (!h->gkr_b) TRY(alloc_dev(&h->gkr_b, 2 * hcap * 3));
    auto bplanes = [&](size_t j) { return View<F>{h->gkr_b + (j & 1) * hcap * 3, hcap}; };
    const View<FI> wit{h->in[3], h->in_stride};
    const Views<FI, 2> in2 = views<FI, 2>(h->in, h->in_stride);
    const Views<F, 2> w2 = views<F, 2>(h->work, h->work_stride);
    size_t M = h->len;
    unsigned g = blocks_for(h, M / 4);
    TRY(prof_begin(h));
    hipLaunchKernelGGL(k_gkr_eval_fw<FI>, dim3(g), dim3(kBlock), 0, h->stream, in2, wit, kb, M / 2, h->w, h->red);
    HIPCHK(hipGetLastError());
    TRY(prof_end(h, 1, 8.0 * 2 * (double)M * FI::D, 0.0));
    // a round's transcript step waits for the next launch, which takes it
    // into its own when that launch is small (base field, k_gkr_fused_fw<FS>)
    Reduce dred = h->red;
    unsigned dg = g;
    bool deferred = true;
    auto wire_now = [&]() -> int {
      if (r == 0)
        hipLaunchKernelGGL(k_gkr_wire_fs<FI>, dim3(1), dim3(kReduceBlock), 0, h->stream, dred, dg, fs_in(h), h->d_fs,
                           E, h->d_polys, h->d_chals, h->d_cur);
      else
        hipLaunchKernelGGL(k_gkr_wire_fs<F>, dim3(1), dim3(kReduceBlock), 0, h->stream, dred, dg, fs_in(h), h->d_fs,
                           E, h->d_polys + r * 2 * E, h->d_chals + r * E, h->d_cur);
      HIPCHK(hipGetLastError());
      r++;
      deferred = false;
      return EFFSC_HIP_OK;
    };
    have = true;
    size_t j = 0;  // folds applied to the tables
    while (r + (deferred ? 1 : 0) < num_rounds && h->len > tail && h->len >= 4 && j + 1 <= kb) {
      M = h->len;
      g = blocks_for(h, M / 8);
      const size_t hb = ((size_t)1 << kb) >> (j + 1);
      const int cconst = j + 1 < kb ? 1 : 0;
      bool fused = false;
      Reduce red = h->red;
      if constexpr (std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value) {
        if (deferred && j >= 1 && h->fuse_parts && g <= kFuseMaxGrid && dg <= kFuseMaxIn &&
            !std::getenv("EFFSC_HIP_NO_FUSE")) {
          red = dred.partials == h->fuse_parts ? h->red : Reduce{h->fuse_parts, h->red.result};
          GkrStep stp{dred, dg, fs_in(h), nullptr, h->d_polys + r * 2, h->d_chals + r, h->d_cur};
          stp.t_out = stp.t_in == h->d_fs ? h->d_fs_alt : h->d_fs;
          hipLaunchKernelGGL((k_gkr_fused_fw<F, FI, F, true>), dim3(g), dim3(kBlock), 0, h->stream, w2, w2,
                             bplanes(j), bplanes(j + 1), hb, wit, kb, cconst, h->d_cur, M / 4, h->w, red, stp);
          HIPCHK(hipGetLastError());
          if (stp.t_out == h->d_fs_alt) h->fs_pending = h->d_fs_alt;
          r++;
          deferred = false;
          fused = true;
        }
      }
      if (!fused) {
        if (deferred) TRY(wire_now());
        TRY(prof_begin(h));
        if (j == 0)
          hipLaunchKernelGGL((k_gkr_fused_fw<FI, FI, F>), dim3(g), dim3(kBlock), 0, h->stream, in2, w2, wit,
                             bplanes(1), hb, wit, kb, cconst, h->d_cur, M / 4, h->w, h->red);
        else
          hipLaunchKernelGGL((k_gkr_fused_fw<F, FI, F>), dim3(g), dim3(kBlock), 0, h->stream, w2, w2, bplanes(j),
                             bplanes(j + 1), hb, wit, kb, cconst, h->d_cur, M / 4, h->w, h->red);
        HIPCHK(hipGetLastError());
        const int din = j == 0 ? FI::D : F::D;
        TRY(prof_end(h, 0, 8.0 * 2 * ((double)M * din + (double)(M / 2) * F::D), 0.0));
      }
      h->len = M / 2;
      j++;
      dred = red;
      dg = g;
      deferred = true;
    }
    if (deferred) TRY(wire_now());
    if (j >= 1) {
      hipLaunchKernelGGL((k_gkr_expand<FI, F>), dim3(blocks_for(h, h->len)), dim3(kBlock), 0, h->stream, bplanes(j),
                         wit, kb, h->len, View<F>{h->work[2], h->work_stride}, View<F>{h->work[3], h->work_stride});
      HIPCHK(hipGetLastError());
      h->in_work = true;
    }
  }
  while (r < num_rounds && h->len > tail && h->len >= 4) {
    const size_t M = h->len;
    const Views<FI, 4> in = views<FI, 4>(h->in, h->in_stride);
    const Views<F, 4> wv = views<F, 4>(h->work, h->work_stride);
    unsigned g;
    if (have) {
      g = blocks_for(h, M / 4);
      TRY(prof_begin(h));
      if (h->in_work)
        hipLaunchKernelGGL((k_gkr_fused<F, F>), dim3(g), dim3(kBlock), 0, h->stream, wv, wv, h->d_cur, M / 4, h->w,
                           h->red);
      else
        hipLaunchKernelGGL((k_gkr_fused<FI, F>), dim3(g), dim3(kBlock), 0, h->stream, in, wv, h->d_cur, M / 4, h->w,
                           h->red);
      HIPCHK(hipGetLastError());
      const int din = h->in_work ? F::D : FI::D;
      TRY(prof_end(h, 0, 8.0 * 4 * ((double)M * din + (double)(M / 2) * F::D), 0.0));
      h->len = M / 2;
      h->in_work = true;
    } else {
      g = blocks_for(h, M / 2);
      TRY(prof_begin(h));
      if (h->in_work)
        hipLaunchKernelGGL(k_gkr_eval<F>, dim3(g), dim3(kBlock), 0, h->stream, wv, M / 2, M / 2, h->w, h->red);
      else
        hipLaunchKernelGGL(k_gkr_eval<FI>, dim3(g), dim3(kBlock), 0, h->stream, in, M / 2, M / 2, h->w, h->red);
      HIPCHK(hipGetLastError());
      TRY(prof_end(h, 1, 8.0 * 4 * (double)M * (h->in_work ? F::D : FI::D), 0.0));
    }
    if (have || h->in_work)
      hipLaunchKernelGGL(k_gkr_reduce_fs<F>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, E,
                         h->d_polys + r * 2 * E, h->d_chals + r * E, h->d_cur);
    else
      hipLaunchKernelGGL(k_gkr_reduce_fs<FI>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, E,
                         h->d_polys + r * 2 * E, h->d_chals + r * E, h->d_cur);
    HIPCHK(hipGetLastError());
    r++;
    have = true;
  }
  const size_t rounds_left = num_rounds - r;
  const size_t M = h->len;
  const Views<FI, 4> in = views<FI, 4>(h->in, h->in_stride);
  const Views<F, 4> wv = views<F, 4>(h->work, h->work_stride);
  if (rounds_left > 0 || M <= tail) {
    hipLaunchKernelGGL((k_gkr_tail<FI, F>), dim3(1), dim3(tail_block<F, 4>()), 0, h->stream, in, wv,
                       h->in_work ? 0 : 1, M, (int)rounds_left, have ? 1 : 0, h->d_fs, E, h->d_polys + r * 2 * E,
                       h->d_chals + r * E, h->d_cur, h->d_fin, h->w);
    HIPCHK(hipGetLastError());
    const size_t folds = rounds_left == 0 ? 0 : (have ? rounds_left : rounds_left - 1);
    size_t Lt = M >> folds;
    if ((have || rounds_left > 0) && Lt >= 2) Lt /= 2;
    h->len = Lt;
    h->in_work = true;
    *tail_used = true;
  } else if (have) {
    // partial run ending on a large table: finalize with the multi-block fold
    // (of the input when no round has folded yet)
    if (h->in_work)
      hipLaunchKernelGGL((k_fold<F, F, 4>), dim3(blocks_for(h, M / 2)), dim3(kBlock), 0, h->stream, wv, wv,
                         gl::zero<F>(), h->d_cur, M / 2, M / 2, h->w);
    else
      hipLaunchKernelGGL((k_fold<FI, F, 4>), dim3(blocks_for(h, M / 2)), dim3(kBlock), 0, h->stream, in, wv,
                         gl::zero<F>(), h->d_cur, M / 2, M / 2, h->w);
    HIPCHK(hipGetLastError());
    h->len = M / 2;
    h->in_work = true;
  }
  return EFFSC_HIP_OK;
}

int prove_device_gkr(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                     uint64_t* challenges, uint64_t* final_value) {
  TRY(ensure_dev_bufs(h));
  DevModeGuard guard(h);
  const int E = h->ext;
  DevFs* hs = reinterpret_cast<DevFs*>(h->h_proof + kMaxRounds * kMaxK * 3 + kMaxRounds * 3 + 4);
  std::memset(hs, 0, sizeof(DevFs));
  hs->kind = (uint32_t)t->kind;
  if (t->kind == 0) {
    for (int i = 0; i < 25; i++) hs->st[i] = t->fs.sp.st[i];
    hs->pos = t->fs.sp.pos;
  }
  hs->rng = t->rng;
  h->fs_pending = nullptr;
  HIPCHK(hipMemcpyAsync(h->d_fs, hs, sizeof(DevFs), hipMemcpyHostToDevice, h->stream));
  bool tail_used = false;
  const int fi = src_field(h);
  if (fi == 1 && E == 1)
    TRY((gkr_device_run<Fp, Fp>(h, num_rounds, &tail_used)));
  else if (fi == 1 && E == 2)
    TRY((gkr_device_run<Fp, Fp2>(h, num_rounds, &tail_used)));
  else if (E == 2)
    TRY((gkr_device_run<Fp2, Fp2>(h, num_rounds, &tail_used)));
  else if (fi == 1)
    TRY((gkr_device_run<Fp, Fp3>(h, num_rounds, &tail_used)));
  else
    TRY((gkr_device_run<Fp3, Fp3>(h, num_rounds, &tail_used)));
  uint64_t* hp = h->h_proof;
  uint64_t* hc = hp + kMaxRounds * kMaxK * 3;
  uint64_t* hf = hc + kMaxRounds * 3;
  TRY(copy_proof_out(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  TRY(prof_collect(h));
  std::copy(hp, hp + num_rounds * 2 * E, round_polys);
  std::copy(hc, hc + num_rounds * E, challenges);
  if (t->kind == 0) {
    for (int i = 0; i < 25; i++) t->fs.sp.st[i] = hs->st[i];
    t->fs.sp.pos = hs->pos;
  }
  t->rng = hs->rng;
  if (tail_used) {
    std::copy(hf, hf + E, final_value);
  } else {
    TRY(final_value_impl(h, final_value));
  }
  return EFFSC_HIP_OK;
}
}  // namespace pv
}  // namespace effsc_hip

extern "C" {

int effsc_hip_gkr_new(const effsc_hip_config* cfg, const uint64_t* add, const uint64_t* mult, const uint64_t* wit,
                      size_t n, effsc_hip_prover** out) {
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (cfg->flags & EFFSC_HIP_BORROW_INPUT) return fail(EFFSC_HIP_EINVAL, "the GKR prover copies its inputs");
  if (n > ((size_t)1 << 20)) return fail(EFFSC_HIP_EINVAL, "witness too long");
  if (n && (!add || !mult || !wit)) return fail(EFFSC_HIP_EINVAL, "null table");
  const size_t nbc = n * n;
  auto* h = new effsc_hip_prover();
  h->kind = K_GKR;
  h->ext = cfg->ext;
  h->in_ext = cfg->input_ext;
  h->K = 4;
  h->deg = 2;
  h->repr = cfg->repr;
  h->dev = cfg->device;
  h->w = cfg->ext > 1 ? gl::canon(cfg->ext_w) : 0;
  h->n = nbc;
  h->len = nbc;
  int rc = setup_common(h, next_pow2(std::max<size_t>(nbc, 1)) >> 1);
  // add and mult as tables 0, 1 (src/provers/gkr.rs:58-63)
  const uint64_t* gates[2] = {add, mult};
  if (rc == EFFSC_HIP_OK) {
    h->K = 2;
    rc = import_tables(h, cfg, gates, nbc);
    h->K = 4;
  }
  // the witness, then its broadcast tables 2, 3 (src/provers/gkr.rs:65-76)
  const int d = h->in_ext;
  uint64_t *staging = nullptr, *wplanes = nullptr;
  for (int i = 2; i < 4 && rc == EFFSC_HIP_OK; i++) rc = alloc_dev(&h->in[i], h->in_stride * d);
  if (rc == EFFSC_HIP_OK && n) rc = alloc_dev(&staging, n * d);
  if (rc == EFFSC_HIP_OK && n) rc = alloc_dev(&wplanes, n * d);
  if (rc == EFFSC_HIP_OK && n) {
    hipError_t e = hipMemcpyAsync(staging, wit, n * d * 8, hipMemcpyDefault, h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_import, dim3(blocks_for(h, n)), dim3(kBlock), 0, h->stream, wplanes, n, staging, n, d,
                         h->repr == EFFSC_HIP_MONT64 ? 1 : 0);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      const unsigned g = blocks_for(h, nbc);
      (void)with_field(d, [&](auto f) -> int {
        using F = decltype(f);
        hipLaunchKernelGGL(k_gkr_broadcast<F>, dim3(g), dim3(kBlock), 0, h->stream, View<F>{wplanes, n}, n,
                           View<F>{h->in[2], h->in_stride}, View<F>{h->in[3], h->in_stride});
        return EFFSC_HIP_OK;
      });
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, std::string("gkr witness: ") + hipGetErrorString(e));
  }
  if (staging) (void)hipFree(staging);
  if (wplanes) (void)hipFree(wplanes);
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  *out = h;
  return EFFSC_HIP_OK;
}

int effsc_hip_gkr_claimed_w(effsc_hip_prover* h, uint64_t* out) {
  // claimed_w_values, src/provers/gkr.rs:89-97: (W(b*), W(c*)) once the
  // tables have length 1, else zeros
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  if (h->kind != K_GKR) return fail(EFFSC_HIP_EINVAL, "not a GKR prover");
  HIPCHK(hipSetDevice(h->dev));
  const int E = h->ext;
  std::fill(out, out + 2 * E, 0);
  if (h->len != 1) return EFFSC_HIP_OK;
  uint64_t el[kMaxK * 3];
  TRY(read_elements(h, 0, el));
  words_out(h, el + 2 * E, out, 2 * E);
  return EFFSC_HIP_OK;
}

}  // extern "C"

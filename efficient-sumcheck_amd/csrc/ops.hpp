// Round launchers of the product-family provers: Ops (evaluate, fused fold +
// evaluate, fold) and DevOps (device-loop reduction-and-transcript step and
// one-workgroup tails).  Included by ops_f1.hip, ops_f2.hip and ops_f3.hip,
// which instantiate them for one output field each.
#pragma once
#include "prover.hpp"

namespace effsc_hip {
namespace pv {

// Lane I's value in every quad of lanes (DPP quad_perm [I, I, I, I]).
template <int I>
__device__ __forceinline__ gl::Fp2 quad_bcast(const gl::Fp2& x) {
  constexpr int ctrl = I | (I << 2) | (I << 4) | (I << 6);
  return {dpp64<ctrl>(x.c0), dpp64<ctrl>(x.c1)};
}
template <int I>
__device__ __forceinline__ gl::Fp3 quad_bcast(const gl::Fp3& x) {
  constexpr int ctrl = I | (I << 2) | (I << 4) | (I << 6);
  return {dpp64<ctrl>(x.c0), dpp64<ctrl>(x.c1), dpp64<ctrl>(x.c2)};
}

// Fold + evaluate of an extension-field product of K = 3 or 4 tables with
// the work of a quadruple split over a quad of lanes: lane j folds table j
// (both elements of the folded pair, written as by k_fused_seq), the quad
// exchanges the folded pairs by DPP, and lane j evaluates point j (0,
// infinity, 2, 3) of the pair's product.  A lane's dependent chain is two
// folds and K - 1 products instead of 2 K folds and K (K - 1) products, for
// the latency-bound small levels.  Partials: K values per workgroup, as
// finish<F, K>.
template <class F, int K>
__global__ __launch_bounds__(kBlock) void k_fused_split(Views<F, K> in, Views<F, K> out, F c, const uint64_t* cptr,
                                                        size_t q, uint64_t w, Reduce r) {
  static_assert(K >= 3 && K <= 4, "a lane per table and per point within a quad");
  if (cptr) c = gl::load_words<F>(cptr);
  const auto cp = gl::prepare(c, w);
  const unsigned j = threadIdx.x & 3;
  View<F> tin = in.t[0], tout = out.t[0];
#pragma unroll
  for (int i = 1; i < K; i++)
    if (j == (unsigned)i) {
      tin = in.t[i];
      tout = out.t[i];
    }
  F acc = gl::zero<F>();
  const size_t step = ((size_t)gridDim.x * blockDim.x) >> 2;
  for (size_t g = gtid() >> 2; g < q; g += step) {
    F lo = gl::zero<F>(), hi = gl::zero<F>();
    if (j < (unsigned)K) {
      const F x0 = tin.ld(g), x1 = tin.ld(g + q), x2 = tin.ld(g + 2 * q), x3 = tin.ld(g + 3 * q);
      lo = gl::add(gl::mul(gl::sub(x2, x0), cp), x0);
      hi = gl::add(gl::mul(gl::sub(x3, x1), cp), x1);
      tout.st(g, lo);
      tout.st(g + q, hi);
    }
    F prod = gl::zero<F>();
    auto factor = [&](const F& l, const F& h) {
      const F d = gl::sub(h, l);
      F x = j == 0 ? l : j == 1 ? d : gl::add(h, d);
      if (j == 3) x = gl::add(x, d);
      return x;
    };
    prod = factor(quad_bcast<0>(lo), quad_bcast<0>(hi));
    prod = gl::mul(prod, factor(quad_bcast<1>(lo), quad_bcast<1>(hi)), w);
    prod = gl::mul(prod, factor(quad_bcast<2>(lo), quad_bcast<2>(hi)), w);
    if constexpr (K == 4) prod = gl::mul(prod, factor(quad_bcast<3>(lo), quad_bcast<3>(hi)), w);
    if (j < (unsigned)K) acc = gl::add(acc, prod);
  }
  // point j: lanes j, j + 4, ... of each wave, then the waves
#pragma unroll
  for (int m = 32; m >= 4; m >>= 1) acc = gl::add(acc, shfl_xor(acc, m));
  __shared__ uint64_t ps[kBlock / 64][4][3];
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < 4) {
    uint64_t wd[3] = {0, 0, 0};
    gl::store_words(acc, wd);
#pragma unroll
    for (int k = 0; k < 3; k++) ps[wave][lane][k] = wd[k];
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)K) {
    F s = gl::zero<F>();
#pragma unroll
    for (int v = 0; v < kBlock / 64; v++) s = gl::add(s, gl::load_words<F>(ps[v][threadIdx.x]));
    gl::store_words(s, r.partials + ((size_t)blockIdx.x * K + threadIdx.x) * F::D);
  }
}

template <class FI, class FO, int K>
struct OpsImpl {
  static int eval(effsc_hip_prover* h, uint64_t* const* planes, size_t stride, size_t half, size_t paired) {
    Views<FI, K> t = views<FI, K>(planes, stride);
    bool vec = (paired == half) && (half % 2 == 0) && (stride % 2 == 0);
    TRY(prof_begin(h));
    if (vec) {
      unsigned g = blocks_for(h, half / 2);
      hipLaunchKernelGGL((k_eval<FI, K, 2>), dim3(g), dim3(kBlock), 0, h->stream, t, half, paired, h->w, h->red);
    } else {
      unsigned g = blocks_for(h, half);
      hipLaunchKernelGGL((k_eval<FI, K, 1>), dim3(g), dim3(kBlock), 0, h->stream, t, half, paired, h->w, h->red);
    }
    HIPCHK(hipGetLastError());
    // algorithmic bytes: every element of the K tables read once
    TRY(prof_end(h, 1, 8.0 * K * FI::D * (double)(half + paired), eval_products(K) * mul_cost<FI>() * half));
    return reduce<FI, K>(h, vec ? blocks_for(h, half / 2) : blocks_for(h, half));
  }
  // chal: host words of the challenge, or nullptr with cptr the device words
  static int fused(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal, size_t q,
                   const uint64_t* cptr = nullptr) {
    Views<FI, K> in = views<FI, K>(src, src_stride);
    Views<FO, K> out = views<FO, K>(h->work, h->work_stride);
    FO c = chal ? from_words<FO>(chal) : gl::zero<FO>();
    bool vec = (q % 2 == 0) && (src_stride % 2 == 0) && (h->work_stride % 2 == 0);
    // F_p^3 products: one quadruple per thread below 2^18 quadruples.  There
    // the launch is latency-bound (few waves, each a chain of about a dozen
    // F_p^3 products per quadruple), and twice the threads halve the chain:
    // 12.8-16.3 us against 22.5-25.1 us for 2^9..2^16 quadruples, 25.0
    // against 28.7 us at 2^17; from 2^18 on two per thread are faster.
    if constexpr (FO::D >= 2 && K >= 3)
      vec = vec && q >= ((size_t)1 << 18) && !std::getenv("EFFSC_HIP_FSEQ_SCALAR");
    // below 2^17 quadruples the quadruple's work split over a quad of lanes
    // (k_fused_split; EFFSC_HIP_FSEQ_SCALAR keeps one quadruple per thread)
    const bool split = FO::D >= 2 && K >= 3 && K <= 4 && std::is_same<FI, FO>::value && !vec &&
                       q < ((size_t)1 << 17) && !std::getenv("EFFSC_HIP_FSEQ_SCALAR");
    TRY(prof_begin(h));
    if constexpr (std::is_same<FI, Fp>::value && std::is_same<FO, Fp3>::value && K == 3) {
      // base tables folded into F_p^3: sums of base-field coefficients of
      // c^k (EFFSC_HIP_FSEQ_GENERIC keeps k_fused_seq)
      const bool generic = std::getenv("EFFSC_HIP_FSEQ_GENERIC") != nullptr;
      // planes of 128 MiB and more: nontemporal (v = 24: 1.424-1.432 against
      // 1.446-1.452 ms per proof)
      const bool nt = 8.0 * 4.0 * (double)q >= kNtMinBytes;
      if (vec && nt && !generic)
        hipLaunchKernelGGL((k_fused_base3<2, true>), dim3(blocks_for(h, q / 2)), dim3(kBlock), 0, h->stream, in, out, c,
                           cptr, q, h->w, h->red);
      else if (vec && !generic)
        hipLaunchKernelGGL((k_fused_base3<2>), dim3(blocks_for(h, q / 2)), dim3(kBlock), 0, h->stream, in, out, c, cptr,
                           q, h->w, h->red);
      else if (!generic)
        hipLaunchKernelGGL((k_fused_base3<1>), dim3(blocks_for(h, q)), dim3(kBlock), 0, h->stream, in, out, c, cptr, q,
                           h->w, h->red);
      else if (vec)
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 2>), dim3(blocks_for(h, q / 2)), dim3(kBlock), 0, h->stream, in,
                           out, c, cptr, q, h->w, h->red);
      else
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 1>), dim3(blocks_for(h, q)), dim3(kBlock), 0, h->stream, in, out,
                           c, cptr, q, h->w, h->red);
    } else if constexpr (FO::D >= 2 && K >= 3) {
      if (vec) {
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 2>), dim3(blocks_for(h, q / 2)), dim3(kBlock), 0, h->stream, in,
                           out, c, cptr, q, h->w, h->red);
      } else if constexpr (std::is_same<FI, FO>::value && K <= 4) {
        if (split) {
          hipLaunchKernelGGL((k_fused_split<FO, K>), dim3(blocks_for(h, 4 * q)), dim3(kBlock), 0, h->stream, in, out,
                             c, cptr, q, h->w, h->red);
        } else {
          hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 1>), dim3(blocks_for(h, q)), dim3(kBlock), 0, h->stream, in,
                             out, c, cptr, q, h->w, h->red);
        }
      } else {
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 1>), dim3(blocks_for(h, q)), dim3(kBlock), 0, h->stream, in, out,
                           c, cptr, q, h->w, h->red);
      }
    } else if (vec) {
      unsigned g = blocks_for(h, q / 2);
      hipLaunchKernelGGL((k_fused<FI, FO, K, 2>), dim3(g), dim3(kBlock), 0, h->stream, in, out, c, cptr, q, h->w, h->red);
    } else {
      unsigned g = blocks_for(h, q);
      hipLaunchKernelGGL((k_fused<FI, FO, K, 1>), dim3(g), dim3(kBlock), 0, h->stream, in, out, c, cptr, q, h->w, h->red);
    }
    HIPCHK(hipGetLastError());
    // algorithmic bytes: read 4q elements, write 2q elements, per table
    // per quadruple: 2 folded elements per table, one evaluated pair
    TRY(prof_end(h, 0, 8.0 * K * (4.0 * q * FI::D + 2.0 * q * FO::D),
                 (double)q * (2.0 * K * fold_cost<FI, FO>() + eval_products(K) * mul_cost<FO>())));
    return reduce<FO, K>(h, vec ? blocks_for(h, q / 2) : split ? blocks_for(h, 4 * q) : blocks_for(h, q));
  }
  static int fold(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal, size_t half,
                  size_t paired, const uint64_t* cptr = nullptr) {
    Views<FI, K> in = views<FI, K>(src, src_stride);
    Views<FO, K> out = views<FO, K>(h->work, h->work_stride);
    FO c = chal ? from_words<FO>(chal) : gl::zero<FO>();
    unsigned g = blocks_for(h, half);
    TRY(prof_begin(h));
    hipLaunchKernelGGL((k_fold<FI, FO, K>), dim3(g), dim3(kBlock), 0, h->stream, in, out, c, cptr, half, paired, h->w);
    HIPCHK(hipGetLastError());
    TRY(prof_end(h, 2, 8.0 * K * ((double)(half + paired) * FI::D + (double)half * FO::D),
                 (double)half * K * fold_cost<FI, FO>()));
    return EFFSC_HIP_OK;
  }
};


template <class FI, class F, int K>
struct DevOpsImpl {
  static int reduce_fs(effsc_hip_prover* h, unsigned g, size_t round) {
    TRY(fs_flush(h));
    hipLaunchKernelGGL((k_reduce_fs<FI, K>), dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, h->ext,
                       h->d_polys + round * K * h->ext, h->d_chals + round * h->ext, h->d_cur);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }
  static int tail(effsc_hip_prover* h, uint64_t* const* src, size_t sst, bool from_input, size_t L, int rounds,
                  bool have, size_t first_round, const uint64_t* cur) {
    if constexpr (!(std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value && K <= 2)) TRY(fs_flush(h));
    Views<FI, K> in = views<FI, K>(src, sst);
    Views<F, K> work = views<F, K>(h->work, h->work_stride);
    if constexpr (std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value && K <= 2) {
      // base-field tables of 4 .. 4096 (K = 2) or 8192 (K = 1) elements after
      // the pending fold: the pipelined tail; 2^11 .. 2^14: the
      // register-resident tail
      const size_t L1 = have ? L / 2 : L;
      if (pipe_tail_ok(h, L1, (size_t)std::max(rounds, 0)))
        return launch_tail_pipe(h, src, sst, L1, have ? 1 : 0, rounds, first_round, cur);
      TRY(fs_flush(h));
      if (rounds > 0 && L1 >= 2048 && L1 <= 16384 && !std::getenv("EFFSC_HIP_TAIL_PLAIN")) {
        auto go = [&](auto m0) {
          constexpr int M0 = decltype(m0)::value;
          hipLaunchKernelGGL((k_tail_reg<K, M0>), dim3(1), dim3(kTailRegThreads), 0, h->stream, in, work,
                             from_input ? 1 : 0, rounds, have ? 1 : 0, h->d_fs, h->d_polys + first_round * K,
                             h->d_chals + first_round, cur, h->d_fin);
        };
        switch (L1 / 1024) {
          case 2: go(std::integral_constant<int, 2>{}); break;
          case 4: go(std::integral_constant<int, 4>{}); break;
          case 8: go(std::integral_constant<int, 8>{}); break;
          default: go(std::integral_constant<int, 16>{}); break;
        }
        HIPCHK(hipGetLastError());
        return EFFSC_HIP_OK;
      }
    }
    if constexpr (std::is_same<FI, Fp3>::value && std::is_same<F, Fp3>::value && K >= 3 && K <= 4) {
      // F_p^3 products of 2^8 or 2^9 work elements (after the pending
      // fold): the register-resident tail (at 2^10 it spills)
      const size_t L1 = L / 2;
      if (have && !from_input && rounds > 0 && L1 >= 256 && L1 <= 512 && !std::getenv("EFFSC_HIP_TAIL_PLAIN")) {
        auto go = [&](auto m0) {
          constexpr int M0 = decltype(m0)::value;
          hipLaunchKernelGGL((k_tail_reg3<K, M0>), dim3(1), dim3(kTailReg3Threads), 0, h->stream, work, rounds,
                             h->d_fs, h->d_polys + first_round * K * 3, h->d_chals + first_round * 3, cur, h->d_fin,
                             h->w);
        };
        if (L1 == 256)
          go(std::integral_constant<int, 1>{});
        else
          go(std::integral_constant<int, 2>{});
        HIPCHK(hipGetLastError());
        return EFFSC_HIP_OK;
      }
    }
    hipLaunchKernelGGL((k_tail<FI, F, K>), dim3(1), dim3(tail_block<F, K>()), 0, h->stream, in, work, from_input ? 1 : 0, L,
                       rounds, have ? 1 : 0, h->d_fs, h->ext, h->d_polys + first_round * K * h->ext,
                       h->d_chals + first_round * h->ext, const_cast<uint64_t*>(cur), h->d_fin, h->w);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }
};

template <class FI, class FO, int K>
int Ops<FI, FO, K>::eval(effsc_hip_prover* h, uint64_t* const* planes, size_t stride, size_t half, size_t paired) {
  return OpsImpl<FI, FO, K>::eval(h, planes, stride, half, paired);
}
template <class FI, class FO, int K>
int Ops<FI, FO, K>::fused(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal, size_t q,
                          const uint64_t* cptr) {
  return OpsImpl<FI, FO, K>::fused(h, src, src_stride, chal, q, cptr);
}
template <class FI, class FO, int K>
int Ops<FI, FO, K>::fold(effsc_hip_prover* h, uint64_t* const* src, size_t src_stride, const uint64_t* chal,
                         size_t half, size_t paired, const uint64_t* cptr) {
  return OpsImpl<FI, FO, K>::fold(h, src, src_stride, chal, half, paired, cptr);
}
template <class FI, class F, int K>
int DevOps<FI, F, K>::reduce_fs(effsc_hip_prover* h, unsigned g, size_t round) {
  return DevOpsImpl<FI, F, K>::reduce_fs(h, g, round);
}
template <class FI, class F, int K>
int DevOps<FI, F, K>::tail(effsc_hip_prover* h, uint64_t* const* src, size_t sst, bool from_input, size_t L,
                           int rounds, bool have, size_t first_round, const uint64_t* cur) {
  return DevOpsImpl<FI, F, K>::tail(h, src, sst, from_input, L, rounds, have, first_round, cur);
}

}  // namespace pv
}  // namespace effsc_hip

#define EFFSC_INSTANTIATE_OPS(FI, FO)                \
  template struct effsc_hip::pv::Ops<FI, FO, 1>;     \
  template struct effsc_hip::pv::Ops<FI, FO, 2>;     \
  template struct effsc_hip::pv::Ops<FI, FO, 3>;     \
  template struct effsc_hip::pv::Ops<FI, FO, 4>;     \
  template struct effsc_hip::pv::DevOps<FI, FO, 1>;  \
  template struct effsc_hip::pv::DevOps<FI, FO, 2>;  \
  template struct effsc_hip::pv::DevOps<FI, FO, 3>;  \
  template struct effsc_hip::pv::DevOps<FI, FO, 4>;

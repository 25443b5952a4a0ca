// Served proofs: a request loop with round-level continuous batching.
//
// Requests (multilinear, inner product, or product of three multilinears over
// F_p^3, each with its own transcript) arrive at any time.  A scheduler thread
// admits them, in priority order, while a slot and room in a fixed HBM arena
// are free: a request's footprint is its input tables (unless borrowed), its
// work tables and its proof words.  Every admitted request then advances at
// every tick, whatever its size and whenever it arrived: ML and IP requests
// two rounds (the two-round passes of pass2.hpp), F_p^3 requests one:
//
//   k_serve_pass        one workgroup per chunk of a base request's two-round
//                       pass (step 0 reads the input, later steps fold twice
//                       with the request's two challenges and evaluate)
//   k_serve_fs2         one workgroup per base request: sums its partials and
//                       runs two transcript rounds, or at the final step folds
//                       what is left, runs the odd last round and writes the
//                       final value
//   k_serve_round_ext   one workgroup per chunk of a product-of-3 request's
//                       round over F_p^3 (round 0 evaluates the base inputs,
//                       round j >= 1 folds with the challenge and evaluates)
//   k_serve_fs          one workgroup per F_p^3 request: sums its partials,
//                       writes the round polynomial, runs the Fiat-Shamir step
//                       on the request's device transcript, and after the last
//                       round folds the tables once more into the final value
//
// Tables, transcripts and challenges stay in device memory from admission to
// completion; only the proof words travel back.  A finished request releases
// its slot and arena bytes at once, so later arrivals join the running batch
// at the next tick.  Each request's proof is exactly runner::sumcheck's for
// that request alone (src/runner.rs:29-69).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "effsc_hip.h"
#include "kernels.hpp"
#include "prove.hpp"

using namespace effsc_hip;
using gl::Fp;
using gl::Fp3;

namespace {

constexpr uint32_t kServeMaxV = 34;
constexpr int kPartWords = 9;              // K * ext <= 9 words per partial
constexpr uint64_t kChunkMin = 2048;       // units per workgroup, at least
constexpr uint64_t kBlocksTarget = 32768;  // workgroups per tick, about
constexpr uint32_t kFinalOnly = 0xFFFFFFFFu;
constexpr uint32_t kFinalStep = 0x80000000u;  // base requests: the final step flag on FItem::j

// A request's device record.
struct SInst {
  const uint64_t* in[3];  // base input tables, 2^v canonical words each
  uint64_t* work[3];      // work tables: ext planes of ws words each
  uint64_t ws;
  uint64_t* polys;  // v * K * ext words, then chals (v * ext), then fin (ext)
  uint64_t* chals;
  uint64_t* fin;
  uint64_t cur[3];  // current challenge
  uint64_t w;       // extension non-residue
  uint32_t v, K, ext, pad;
  DevFs fs;
};

// One request's share of a round launch: blocks [chunk0, chunk0 + ceil(units /
// per)) of the tick's global block numbering.
struct RItem {
  uint32_t slot, j;
  uint64_t units;  // pairs (j = 0) or quadruples (j >= 1)
  uint64_t chunk0;
  uint64_t per;
};

// One request's transcript workgroup.
struct FItem {
  uint32_t slot, j;  // j = the round just evaluated, or kFinalOnly (v = 0)
  uint64_t part0, nparts;
};

template <int K, class FW>
__device__ __forceinline__ void serve_chunk(const SInst* ip, uint32_t j, uint64_t begin, uint64_t end, FW* acc) {
  const uint64_t w = ip->w;
  const uint32_t v = ip->v;
  if (j == 0) {
    // round 0 evaluates the base inputs; products of base elements are the
    // same field elements as products of their lifts
    const uint64_t half = (uint64_t)1 << (v - 1);
    Fp a0[K];
#pragma unroll
    for (int i = 0; i < K; i++) a0[i] = gl::zero<Fp>();
    for (uint64_t k = begin + threadIdx.x; k < end; k += blockDim.x) {
      Fp lo[K], hi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        lo[i] = {ip->in[i][k]};
        hi[i] = {ip->in[i][k + half]};
      }
      eval_pair<Fp, K>(lo, hi, a0, 0);
    }
#pragma unroll
    for (int i = 0; i < K; i++) acc[i] = gl::lift_to<FW>(a0[i]);
    return;
  }
  // round j >= 1: fold quadruples of the level-(j-1) table (the input when
  // j = 1, else the work table, in place) and evaluate the folded pairs
  const uint64_t q = (uint64_t)1 << (v - j - 1);
  const FW c = gl::load_words<FW>(ip->cur);
  if constexpr (FW::D == 3 && K == 3) {
    if (j == 1) {
      // base inputs folded into F_p^3: the evaluation as base-field
      // coefficients of c^k, converted once (as k_fused_base3)
      Fp e[3][4];
#pragma unroll
      for (int t = 0; t < 3; t++)
#pragma unroll
        for (int m = 0; m < 4; m++) e[t][m] = gl::zero<Fp>();
      for (uint64_t k = begin + threadIdx.x; k < end; k += blockDim.x) {
        Fp a[3][3], b[3][3];  // [point][table]
#pragma unroll
        for (int i = 0; i < K; i++) {
          View<FW> d{ip->work[i], ip->ws};
          const uint64_t* s = ip->in[i];
          const Fp x0 = {s[k]}, x1 = {s[k + q]}, x2 = {s[k + 2 * q]}, x3 = {s[k + 3 * q]};
          d.st(k, fold1<FW>(x0, x2, c, w));
          d.st(k + q, fold1<FW>(x1, x3, c, w));
          Fp pa[3], pb[3];
          base3_points(x0, x1, x2, x3, pa, pb);
#pragma unroll
          for (int t = 0; t < 3; t++) {
            a[t][i] = pa[t];
            b[t][i] = pb[t];
          }
        }
#pragma unroll
        for (int t = 0; t < 3; t++) base3_terms(a[t], b[t], e[t]);
      }
      base3_finish(e, c, w, acc);
      return;
    }
  }
  if constexpr (FW::D == 3 && K >= 3) {
    // one table at a time with running products (as k_fused_seq): the
    // all-tables-at-once form needs more registers than two waves allow
    for (uint64_t k = begin + threadIdx.x; k < end; k += blockDim.x) {
      FW p[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        View<FW> d{ip->work[i], ip->ws};
        FW lo, hi;
        if (j == 1) {
          const uint64_t* s = ip->in[i];
          Fp x0 = {s[k]}, x1 = {s[k + q]}, x2 = {s[k + 2 * q]}, x3 = {s[k + 3 * q]};
          lo = fold1<FW>(x0, x2, c, w);
          hi = fold1<FW>(x1, x3, c, w);
        } else {
          FW x0 = d.ld(k), x1 = d.ld(k + q), x2 = d.ld(k + 2 * q), x3 = d.ld(k + 3 * q);
          lo = fold1<FW>(x0, x2, c, w);
          hi = fold1<FW>(x1, x3, c, w);
        }
        d.st(k, lo);
        d.st(k + q, hi);
        const FW dd = gl::sub(hi, lo);
        FW x = hi;
        if (i == 0) {
          p[0] = lo;
          p[1] = dd;
#pragma unroll
          for (int t = 2; t < K; t++) {
            x = gl::add(x, dd);
            p[t] = x;
          }
        } else {
          p[0] = gl::mul(p[0], lo, w);
          p[1] = gl::mul(p[1], dd, w);
#pragma unroll
          for (int t = 2; t < K; t++) {
            x = gl::add(x, dd);
            p[t] = gl::mul(p[t], x, w);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < K; t++) acc[t] = gl::add(acc[t], p[t]);
    }
    return;
  }
  for (uint64_t k = begin + threadIdx.x; k < end; k += blockDim.x) {
    FW lo[K], hi[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
      View<FW> d{ip->work[i], ip->ws};
      if (j == 1) {
        const uint64_t* s = ip->in[i];
        Fp x0 = {s[k]}, x1 = {s[k + q]}, x2 = {s[k + 2 * q]}, x3 = {s[k + 3 * q]};
        lo[i] = fold1<FW>(x0, x2, c, w);
        hi[i] = fold1<FW>(x1, x3, c, w);
      } else {
        FW x0 = d.ld(k), x1 = d.ld(k + q), x2 = d.ld(k + 2 * q), x3 = d.ld(k + 3 * q);
        lo[i] = fold1<FW>(x0, x2, c, w);
        hi[i] = fold1<FW>(x1, x3, c, w);
      }
      d.st(k, lo[i]);
      d.st(k + q, hi[i]);
    }
    eval_pair<FW, K>(lo, hi, acc, w);
  }
}

template <int K, class FW>
__device__ __forceinline__ void serve_block(const SInst* ip, const RItem& it, uint64_t b, uint64_t* partials) {
  FW acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = gl::zero<FW>();
  const uint64_t begin = b * it.per, end = std::min(begin + it.per, it.units);
  serve_chunk<K, FW>(ip, it.j, begin, end, acc);
  block_reduce<FW, K>(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < K; i++) gl::store_words(acc[i], partials + (it.chunk0 + b) * kPartWords + i * FW::D);
}

// The item of global block gb: the last one with chunk0 <= gb (items are
// sorted by chunk0).
__device__ __forceinline__ uint32_t find_item(const RItem* items, uint32_t nitems, uint64_t gb) {
  uint32_t lo = 0, hi = nitems;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (items[mid].chunk0 <= gb)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Base-field requests (ML, IP): two rounds per tick with the two-round
// passes of pass2.hpp.  Step s < v/2 is a pass over the level-2^(v-2s) table
// (units q = 2^(v-2s-2)): step 0 reads the input, step 1 folds the input
// twice with the step-0 challenges in cur and stores level 2^(v-2), later
// steps fold the stored level in place; k_serve_fs2 sums the partials and
// runs two transcript rounds.  The final step (s = v/2) folds what is left,
// runs the last round when v is odd, and writes the final value.
template <int K>
__device__ __forceinline__ void serve_pass_block(const SInst* ip, const RItem& it, uint64_t b, uint64_t* partials) {
  const uint32_t s = it.j, v = ip->v;
  const size_t L = (size_t)1 << (v - 2 * s), q = L / 4;
  const uint64_t c1 = s ? ip->cur[0] : 0, c2 = s ? ip->cur[1] : 0, c12 = gl::mul(c1, c2);
  const uint64_t* src[2] = {s <= 1 ? ip->in[0] : ip->work[0], s <= 1 ? ip->in[1] : ip->work[1]};
  const uint64_t begin = b * it.per, end = std::min<uint64_t>(begin + it.per, it.units);
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) acc.a[i].zero();
  if (q >= 2) {
    // pairs of units with 16-byte loads (q, per and the planes are even)
    for (uint64_t u = begin + 2 * threadIdx.x; u < end; u += 2 * blockDim.x) {
      uint64_t E0[K][4], E1[K][4];
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (s == 0)
            fold_chain2<0, true>(src[i], u + j * q, L, c1, c2, c12, E0[i][j], E1[i][j]);
          else
            fold_chain2<2, true>(src[i], u + j * q, L, c1, c2, c12, E0[i][j], E1[i][j]);
        }
      if (s > 0) {
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
          for (int j = 0; j < 4; j++)
            __builtin_nontemporal_store(v2u64{E0[i][j], E1[i][j]}, reinterpret_cast<v2u64*>(ip->work[i] + u + j * q));
      }
      pass_accumulate<K>(E0, acc);
      pass_accumulate<K>(E1, acc);
    }
  } else {
    for (uint64_t u = begin + threadIdx.x; u < end; u += blockDim.x) {
      uint64_t E[K][4];
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          E[i][j] = s == 0 ? fold_chain<0>(src[i], u + j * q, L, c1, c2, c12)
                          : fold_chain<2>(src[i], u + j * q, L, c1, c2, c12);
      if (s > 0) {
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) ip->work[i][u + j * q] = E[i][j];
      }
      pass_accumulate<K>(E, acc);
    }
  }
  constexpr int NV = pass_values<K>();
  Fp raw[pass_raw<K>()], vals[NV];
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  block_reduce<Fp, NV>(vals);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; i++) partials[(it.chunk0 + b) * kPartWords + i] = vals[i].v;
}

__global__ __launch_bounds__(kBlock) void k_serve_pass(const SInst* insts, const RItem* items, uint32_t nitems,
                                                       uint64_t block0, uint64_t* partials) {
  const uint64_t gb = block0 + blockIdx.x;
  const RItem it = items[find_item(items, nitems, gb)];
  const SInst* ip = insts + it.slot;
  if (ip->K == 1)
    serve_pass_block<1>(ip, it, gb - it.chunk0, partials);
  else
    serve_pass_block<2>(ip, it, gb - it.chunk0, partials);
}

// The final step s of a base request: the stored level (the input when
// s <= 1, else the work table) with the step's two pending folds, at most 8
// elements per table, is folded to 1 or 2; the odd last round, the last fold
// and the product follow.
template <int K>
__device__ __forceinline__ void serve_final2(SInst* ip, uint32_t s) {
  __shared__ uint64_t chal_sh[3];
  const uint32_t v = ip->v;
  if (v == 0) {
    if (threadIdx.x == 0) ip->fin[0] = K == 2 ? gl::mul(ip->in[0][0], ip->in[1][0]) : ip->in[0][0];
    return;
  }
  const size_t M = s == 0 ? ((size_t)1 << v) : ((size_t)1 << (v - 2 * s + 2));
  const uint64_t* src[2] = {s <= 1 ? ip->in[0] : ip->work[0], s <= 1 ? ip->in[1] : ip->work[1]};
  const uint64_t x = s ? ip->cur[0] : 0, y = s ? ip->cur[1] : 0;
  const size_t n2 = s ? M / 4 : M;  // 1 or 2 elements per table after the pending folds
  uint64_t t[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int i = 0; i < K; i++)
    for (size_t e = 0; e < n2; e++) t[i][e] = s == 0 ? src[i][e] : fold_chain<2>(src[i], e, M / 4, x, y, gl::mul(x, y));
  uint64_t f[2] = {t[0][0], t[1][0]};
  if (n2 == 2) {
    // round v - 1 over the pair (t[.][0], t[.][1]): ML [h(0)], IP [h(0), h(inf)]
    Fp vals[K];
    if constexpr (K == 1) {
      vals[0] = {t[0][0]};
    } else {
      vals[0] = {gl::mul(t[0][0], t[1][0])};
      vals[1] = {gl::mul(gl::sub(t[0][1], t[0][0]), gl::sub(t[1][1], t[1][0]))};
    }
    if (threadIdx.x < 64) emit_round<Fp, K>(vals, 1, &ip->fs, ip->polys + (size_t)(v - 1) * K, ip->chals + (v - 1), chal_sh);
    __syncthreads();
    const uint64_t z = chal_sh[0];
#pragma unroll
    for (int i = 0; i < K; i++) f[i] = gl::add(t[i][0], gl::mul(z, gl::sub(t[i][1], t[i][0])));
  }
  if (threadIdx.x == 0) ip->fin[0] = K == 2 ? gl::mul(f[0], f[1]) : f[0];
}

template <int K>
__device__ __forceinline__ void serve_fs2(SInst* ip, const FItem& it, const uint64_t* partials) {
  const uint32_t s = it.j & ~kFinalStep;
  if (it.j & kFinalStep) {
    serve_final2<K>(ip, s);
    return;
  }
  constexpr int NV = pass_values<K>();
  // the sponge load overlaps the partial sums
  LaneFs lf;
  if (threadIdx.x < 64) lf.load(&ip->fs);
  Fp sum[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) sum[i] = gl::zero<Fp>();
  for (uint64_t b = threadIdx.x; b < it.nparts; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; i++) sum[i] = gl::add(sum[i], Fp{partials[(it.part0 + b) * kPartWords + i]});
  block_reduce<Fp, NV>(sum);
  if (threadIdx.x < 64)
    emit_two_rounds<K>(sum, lf, &ip->fs, ip->polys + (size_t)2 * s * K, ip->chals + 2 * s, ip->cur);
}

__global__ __launch_bounds__(kBlock) void k_serve_fs2(SInst* insts, const FItem* items, const uint64_t* partials) {
  const FItem it = items[blockIdx.x];
  SInst* ip = insts + it.slot;
  if (ip->K == 1)
    serve_fs2<1>(ip, it, partials);
  else
    serve_fs2<2>(ip, it, partials);
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_serve_round_ext(
    const SInst* insts, const RItem* items, uint32_t nitems,
                                                            uint64_t block0, uint64_t* partials) {
  const uint64_t gb = block0 + blockIdx.x;
  const RItem it = items[find_item(items, nitems, gb)];
  serve_block<3, Fp3>(insts + it.slot, it, gb - it.chunk0, partials);
}

template <int K, class FW>
__device__ __forceinline__ void serve_fs(SInst* ip, const FItem& it, const uint64_t* partials) {
  constexpr int D = FW::D;
  if (it.j != kFinalOnly) {
    FW s[K];
#pragma unroll
    for (int i = 0; i < K; i++) s[i] = gl::zero<FW>();
    for (uint64_t b = threadIdx.x; b < it.nparts; b += blockDim.x)
#pragma unroll
      for (int i = 0; i < K; i++)
        s[i] = gl::add(s[i], gl::load_words<FW>(partials + (it.part0 + b) * kPartWords + i * D));
    block_reduce<FW, K>(s);
    if (threadIdx.x < 64)
      emit_round<FW, K>(s, D, &ip->fs, ip->polys + (size_t)it.j * K * D, ip->chals + (size_t)it.j * D, ip->cur);
    if (it.j + 1 < ip->v) return;
  }
  // final fold with the last challenge (src/runner.rs:57-60) and the final
  // value, the product of the folded tables; thread 0 wrote cur above
  if (threadIdx.x != 0) return;
  const uint64_t w = ip->w;
  FW val = gl::one<FW>();
  if (ip->v == 0) {
#pragma unroll
    for (int i = 0; i < K; i++) val = gl::mul(val, gl::lift_to<FW>(Fp{ip->in[i][0]}), w);
  } else {
    const FW c = gl::load_words<FW>(ip->cur);
#pragma unroll
    for (int i = 0; i < K; i++) {
      FW f;
      if (ip->v == 1) {
        f = fold1<FW>(Fp{ip->in[i][0]}, Fp{ip->in[i][1]}, c, w);
      } else {
        View<FW> d{ip->work[i], ip->ws};
        f = fold1<FW>(d.ld(0), d.ld(1), c, w);
      }
      val = gl::mul(val, f, w);
    }
  }
  gl::store_words(val, ip->fin);
}

__global__ __launch_bounds__(kBlock) void k_serve_fs(SInst* insts, const FItem* items, const uint64_t* partials) {
  const FItem it = items[blockIdx.x];
  SInst* ip = insts + it.slot;
  serve_fs<3, Fp3>(ip, it, partials);  // F_p^3 requests; base requests use k_serve_fs2
}

__global__ __launch_bounds__(kBlock) void k_serve_fill(uint64_t* dst, uint64_t n, uint64_t key) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gtid(); i < n; i += step) dst[i] = uniform_word(key, i);
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit allocator over the server's HBM arena (offsets in bytes).
class Arena {
 public:
  size_t cap = 0, used = 0, peak = 0;
  void init(size_t c) {
    cap = c;
    free_.clear();
    if (c) free_[0] = c;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      size_t o = it->first, rem = it->second - bytes;
      free_.erase(it);
      if (rem) free_[o + bytes] = rem;
      *off = o;
      used += bytes;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align256(bytes);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;
};

struct Layout {
  size_t in_off[3], work_off, proof_off, bytes;
  size_t ws;  // work plane stride in words
};

// Arena layout of one request: [inputs][work planes][polys | chals | fin].
Layout layout(int K, int ext, uint32_t v, const uint64_t* const* tables) {
  Layout L{};
  size_t n = (size_t)1 << v, o = 0;
  for (int i = 0; i < K; i++) {
    L.in_off[i] = o;
    if (!tables[i]) o += align256(n * 8);
  }
  // base-field requests (ML, IP) run two rounds per tick and store every
  // other level, from n/4 down; F_p^3 requests store each level from n/2
  const bool two_round = K <= 2 && ext == 1;
  L.ws = (std::max<size_t>(two_round ? n / 4 : n / 2, 2) + 31) & ~(size_t)31;
  L.work_off = o;
  o += align256((size_t)K * ext * L.ws * 8);
  L.proof_off = o;
  o += align256(((size_t)v * K * ext + (size_t)v * ext + ext) * 8);
  L.bytes = o;
  return L;
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Req {
  uint64_t id = 0;
  int kind = 0, K = 0, ext = 1;
  uint32_t v = 0;
  uint64_t seeds[3] = {};
  const uint64_t* tables[3] = {};
  double priority = 0;
  DevFs fs0{};
  Layout lay{};
  size_t arena_off = 0;
  int slot = -1;
  uint32_t j = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool done = false;
  int status = EFFSC_HIP_OK;
  std::vector<uint64_t> proof;  // polys | chals | fin
  hipEvent_t ev = nullptr;      // retire event
};

}  // namespace

struct effsc_hip_server {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::chrono::steady_clock::time_point t0;
  uint8_t* arena_base = nullptr;
  Arena arena;
  size_t max_active = 0;
  SInst* d_insts = nullptr;
  uint8_t* d_items = nullptr;
  size_t items_bytes = 0;
  uint64_t* d_partials = nullptr;
  size_t partial_blocks = 0;
  static constexpr int kRing = 3;
  uint8_t* h_stage[kRing] = {};
  hipEvent_t stage_ev[kRing] = {};
  bool stage_armed[kRing] = {};
  size_t stage_bytes = 0;
  uint64_t tick_no = 0;
  uint64_t* h_res = nullptr;  // pinned proof words per slot
  size_t res_words = 0;
  std::vector<hipEvent_t> ev_pool;

  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::vector<std::shared_ptr<Req>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> reqs;
  std::vector<int> free_slots;
  uint64_t next_id = 1;
  bool stop = false;
  std::string fatal;
  std::atomic<uint64_t> stat_ticks{0}, stat_peak_active{0}, stat_completed{0};

  std::vector<std::shared_ptr<Req>> active, retiring;  // scheduler thread only
  std::thread th;
};

namespace {

int record_fatal(effsc_hip_server* s, const std::string& m) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fatal.empty()) s->fatal = m;
  for (auto& kv : s->reqs) {
    if (!kv.second->done) {
      kv.second->done = true;
      kv.second->status = EFFSC_HIP_EDEVICE;
    }
  }
  s->pending.clear();
  s->cv_done.notify_all();
  return EFFSC_HIP_EDEVICE;
}

#define SCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return record_fatal(s, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Completes retiring requests whose proof words have arrived; with `block`,
// waits for the oldest one first.
int poll_retiring(effsc_hip_server* s, bool block) {
  if (s->retiring.empty()) return EFFSC_HIP_OK;
  if (block) SCHK(hipEventSynchronize(s->retiring.front()->ev));
  std::vector<std::shared_ptr<Req>> keep;
  for (auto& r : s->retiring) {
    hipError_t q = hipEventQuery(r->ev);
    if (q == hipErrorNotReady) {
      keep.push_back(r);
      continue;
    }
    if (q != hipSuccess) return record_fatal(s, std::string("retire: ") + hipGetErrorString(q));
    const size_t words = (size_t)r->v * r->K * r->ext + (size_t)r->v * r->ext + r->ext;
    r->proof.assign(s->h_res + (size_t)r->slot * s->res_words, s->h_res + (size_t)r->slot * s->res_words + words);
    s->ev_pool.push_back(r->ev);
    r->ev = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_done = now_s(s->t0);
    s->arena.release(r->arena_off, r->lay.bytes);
    s->free_slots.push_back(r->slot);
    r->done = true;
    s->stat_completed++;
    s->cv_done.notify_all();
    s->cv_submit.notify_all();  // memory and a slot were freed
  }
  s->retiring.swap(keep);
  return EFFSC_HIP_OK;
}

// One tick: device set-up of newly admitted requests, then one round of every
// active request.
int tick(effsc_hip_server* s, const std::vector<std::shared_ptr<Req>>& admitted) {
  const int R = effsc_hip_server::kRing;
  const int slot_r = (int)(s->tick_no % R);
  if (s->stage_armed[slot_r]) SCHK(hipEventSynchronize(s->stage_ev[slot_r]));
  uint8_t* stage = s->h_stage[slot_r];

  // new requests: records (staged), table generation
  std::vector<SInst> recs;
  for (auto& r : admitted) {
    SInst in{};
    uint8_t* base = s->arena_base + r->arena_off;
    for (int i = 0; i < 3; i++) {
      in.in[i] = nullptr;
      in.work[i] = nullptr;
    }
    for (int i = 0; i < r->K; i++) {
      in.in[i] = r->tables[i] ? r->tables[i] : reinterpret_cast<const uint64_t*>(base + r->lay.in_off[i]);
      in.work[i] = reinterpret_cast<uint64_t*>(base + r->lay.work_off) + (size_t)i * r->ext * r->lay.ws;
    }
    in.ws = r->lay.ws;
    in.polys = reinterpret_cast<uint64_t*>(base + r->lay.proof_off);
    in.chals = in.polys + (size_t)r->v * r->K * r->ext;
    in.fin = in.chals + (size_t)r->v * r->ext;
    in.w = EFFSC_HIP_W_FROM_RAW7;
    in.v = r->v;
    in.K = (uint32_t)r->K;
    in.ext = (uint32_t)r->ext;
    in.fs = r->fs0;
    recs.push_back(in);
  }
  // round and transcript items
  std::vector<RItem> base_items, ext_items;
  std::vector<FItem> fs_items, fs2_items;
  // base requests: step s = j / 2 is a two-round pass while s < v / 2, then
  // the final step; F_p^3 requests: one round per tick
  auto base_units = [](const Req& r) -> uint64_t {
    const uint32_t st = r.j / 2;
    return st < r.v / 2 ? (uint64_t)1 << (r.v - 2 * st - 2) : 0;
  };
  uint64_t total_units = 0;
  for (auto& r : s->active) {
    if (r->K != 3)
      total_units += base_units(*r);
    else if (r->v > 0)
      total_units += r->j == 0 ? (uint64_t)1 << (r->v - 1) : (uint64_t)1 << (r->v - r->j - 1);
  }
  // even: a pass chunk takes pairs of units
  uint64_t per = std::max<uint64_t>(kChunkMin, (total_units + kBlocksTarget - 1) / kBlocksTarget);
  per += per & 1;
  uint64_t gb = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (auto& r : s->active) {
      const bool ext_class = r->K == 3;
      if (ext_class != (pass == 1)) continue;
      if (!ext_class) {
        const uint32_t st = r->j / 2;
        const uint64_t units = base_units(*r);
        if (units == 0) {
          fs2_items.push_back({(uint32_t)r->slot, st | kFinalStep, 0, 0});
          continue;
        }
        const uint64_t nb = (units + per - 1) / per;
        base_items.push_back({(uint32_t)r->slot, st, units, gb, per});
        fs2_items.push_back({(uint32_t)r->slot, st, gb, nb});
        gb += nb;
        continue;
      }
      if (r->v == 0) {
        fs_items.push_back({(uint32_t)r->slot, kFinalOnly, 0, 0});
        continue;
      }
      uint64_t units = r->j == 0 ? (uint64_t)1 << (r->v - 1) : (uint64_t)1 << (r->v - r->j - 1);
      uint64_t nb = (units + per - 1) / per;
      RItem it{(uint32_t)r->slot, r->j, units, gb, per};
      (ext_class ? ext_items : base_items).push_back(it);
      fs_items.push_back({(uint32_t)r->slot, r->j, gb, nb});
      gb += nb;
    }
  }
  if (gb > s->partial_blocks) return record_fatal(s, "serve: partials overflow");
  const uint64_t nbase = base_items.empty() ? 0 : ext_items.empty() ? gb : ext_items.front().chunk0;
  const uint64_t next = gb - nbase;
  // stage: records | base items | ext items | fs items
  size_t off = 0;
  auto put = [&](const void* p, size_t bytes) {
    std::memcpy(stage + off, p, bytes);
    size_t o = off;
    off += (bytes + 15) & ~(size_t)15;
    return o;
  };
  size_t rec_off = put(recs.data(), recs.size() * sizeof(SInst));
  size_t items_start = off;
  size_t base_off = put(base_items.data(), base_items.size() * sizeof(RItem)) - items_start;
  size_t ext_off = put(ext_items.data(), ext_items.size() * sizeof(RItem)) - items_start;
  size_t fs_off = put(fs_items.data(), fs_items.size() * sizeof(FItem)) - items_start;
  size_t fs2_off = put(fs2_items.data(), fs2_items.size() * sizeof(FItem)) - items_start;
  if (off > s->stage_bytes || off - items_start > s->items_bytes) return record_fatal(s, "serve: staging overflow");
  for (size_t i = 0; i < admitted.size(); i++)
    SCHK(hipMemcpyAsync(s->d_insts + admitted[i]->slot, stage + rec_off + i * sizeof(SInst), sizeof(SInst),
                        hipMemcpyHostToDevice, s->stream));
  if (off > items_start)
    SCHK(hipMemcpyAsync(s->d_items, stage + items_start, off - items_start, hipMemcpyHostToDevice, s->stream));
  SCHK(hipEventRecord(s->stage_ev[slot_r], s->stream));
  s->stage_armed[slot_r] = true;
  s->tick_no++;
  for (auto& r : admitted) {
    const size_t n = (size_t)1 << r->v;
    for (int i = 0; i < r->K; i++) {
      if (r->tables[i]) continue;
      unsigned g = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
      uint64_t* dst = reinterpret_cast<uint64_t*>(s->arena_base + r->arena_off + r->lay.in_off[i]);
      hipLaunchKernelGGL(k_serve_fill, dim3(g), dim3(kBlock), 0, s->stream, dst, (uint64_t)n,
                         uniform_key(r->seeds[i]));
      SCHK(hipGetLastError());
    }
  }
  if (!base_items.empty()) {
    hipLaunchKernelGGL(k_serve_pass, dim3((unsigned)nbase), dim3(kBlock), 0, s->stream, s->d_insts,
                       reinterpret_cast<const RItem*>(s->d_items + base_off), (uint32_t)base_items.size(),
                       (uint64_t)0, s->d_partials);
    SCHK(hipGetLastError());
  }
  if (!ext_items.empty()) {
    hipLaunchKernelGGL(k_serve_round_ext, dim3((unsigned)next), dim3(kBlock), 0, s->stream, s->d_insts,
                       reinterpret_cast<const RItem*>(s->d_items + ext_off), (uint32_t)ext_items.size(), nbase,
                       s->d_partials);
    SCHK(hipGetLastError());
  }
  if (!fs_items.empty()) {
    hipLaunchKernelGGL(k_serve_fs, dim3((unsigned)fs_items.size()), dim3(kBlock), 0, s->stream, s->d_insts,
                       reinterpret_cast<const FItem*>(s->d_items + fs_off), s->d_partials);
    SCHK(hipGetLastError());
  }
  if (!fs2_items.empty()) {
    hipLaunchKernelGGL(k_serve_fs2, dim3((unsigned)fs2_items.size()), dim3(kBlock), 0, s->stream, s->d_insts,
                       reinterpret_cast<const FItem*>(s->d_items + fs2_off), s->d_partials);
    SCHK(hipGetLastError());
  }
  s->stat_ticks++;
  // advance; finished requests copy their proof words back
  std::vector<std::shared_ptr<Req>> still;
  for (auto& r : s->active) {
    if (r->K != 3) {
      // a base request retires after its final step
      if (r->j / 2 < r->v / 2) {
        r->j += 2;
        still.push_back(r);
        continue;
      }
    } else {
      r->j++;
      if (r->v > 0 && r->j < r->v) {
        still.push_back(r);
        continue;
      }
    }
    const size_t words = (size_t)r->v * r->K * r->ext + (size_t)r->v * r->ext + r->ext;
    SCHK(hipMemcpyAsync(s->h_res + (size_t)r->slot * s->res_words, s->arena_base + r->arena_off + r->lay.proof_off,
                        words * 8, hipMemcpyDeviceToHost, s->stream));
    if (s->ev_pool.empty()) {
      hipEvent_t e;
      SCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      s->ev_pool.push_back(e);
    }
    r->ev = s->ev_pool.back();
    s->ev_pool.pop_back();
    SCHK(hipEventRecord(r->ev, s->stream));
    s->retiring.push_back(r);
  }
  s->active.swap(still);
  return EFFSC_HIP_OK;
}

void scheduler(effsc_hip_server* s) {
  if (hipSetDevice(s->dev) != hipSuccess) {
    record_fatal(s, "scheduler: hipSetDevice failed");
    return;
  }
  for (;;) {
    if (poll_retiring(s, false) != EFFSC_HIP_OK) return;
    std::vector<std::shared_ptr<Req>> admitted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (!s->fatal.empty()) return;
      while (!s->pending.empty() && !s->free_slots.empty()) {
        auto best = std::min_element(s->pending.begin(), s->pending.end(), [](const auto& a, const auto& b) {
          return a->priority != b->priority ? a->priority < b->priority : a->id < b->id;
        });
        size_t off = 0;
        if (!s->arena.alloc((*best)->lay.bytes, &off)) break;  // wait for memory, keeping the order
        auto r = *best;
        s->pending.erase(best);
        r->arena_off = off;
        r->slot = s->free_slots.back();
        s->free_slots.pop_back();
        r->t_admit = now_s(s->t0);
        admitted.push_back(r);
      }
      if (s->active.empty() && admitted.empty()) {
        if (!s->retiring.empty()) {
          lk.unlock();
          if (poll_retiring(s, true) != EFFSC_HIP_OK) return;
          continue;
        }
        if (s->stop && s->pending.empty()) return;
        s->cv_submit.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->fatal.empty(); });
        continue;
      }
    }
    for (auto& r : admitted) s->active.push_back(r);
    if (s->active.size() > s->stat_peak_active) s->stat_peak_active = s->active.size();
    if (tick(s, admitted) != EFFSC_HIP_OK) return;
  }
}

void destroy_server(effsc_hip_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->dev);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& kv : s->reqs)
    if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  for (int i = 0; i < effsc_hip_server::kRing; i++) {
    if (s->stage_ev[i]) (void)hipEventDestroy(s->stage_ev[i]);
    if (s->h_stage[i]) (void)hipHostFree(s->h_stage[i]);
  }
  if (s->h_res) (void)hipHostFree(s->h_res);
  for (void* p : {(void*)s->arena_base, (void*)s->d_insts, (void*)s->d_items, (void*)s->d_partials})
    if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int request_shape(const effsc_hip_request* q, int* K, int* ext) {
  if (!q) return set_error(EFFSC_HIP_EINVAL, "null request");
  if (q->kind == EFFSC_HIP_SERVE_MULTILINEAR) {
    *K = 1;
    *ext = 1;
  } else if (q->kind == EFFSC_HIP_SERVE_INNER_PRODUCT) {
    *K = 2;
    *ext = 1;
  } else if (q->kind == EFFSC_HIP_SERVE_PRODUCT3_EXT3) {
    *K = 3;
    *ext = 3;
  } else {
    return set_error(EFFSC_HIP_EINVAL, "request kind must be multilinear, inner product or product-of-3 over F_p^3");
  }
  if (q->v > kServeMaxV) return set_error(EFFSC_HIP_EINVAL, "request has too many variables");
  return EFFSC_HIP_OK;
}

}  // namespace

extern "C" {

int effsc_hip_server_footprint(const effsc_hip_request* req, size_t* bytes) {
  int K = 0, ext = 1;
  TRY(request_shape(req, &K, &ext));
  if (!bytes) return set_error(EFFSC_HIP_EINVAL, "null out");
  *bytes = layout(K, ext, req->v, req->tables).bytes;
  return EFFSC_HIP_OK;
}

int effsc_hip_server_new(int device, size_t hbm_budget, size_t max_active, effsc_hip_server** out) {
  if (!out) return set_error(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  if (max_active == 0 || max_active > 4096) return set_error(EFFSC_HIP_EINVAL, "max_active must be in [1, 4096]");
  if (hbm_budget < 4096) return set_error(EFFSC_HIP_EINVAL, "HBM budget too small");
  HIPCHK(hipSetDevice(device));
  auto* s = new effsc_hip_server();
  s->dev = device;
  s->t0 = std::chrono::steady_clock::now();
  s->max_active = max_active;
  s->arena.init(align256(hbm_budget));
  s->partial_blocks = kBlocksTarget + 2 * max_active;
  s->items_bytes = 2 * max_active * sizeof(RItem) + max_active * sizeof(FItem) + 64;
  s->stage_bytes = max_active * sizeof(SInst) + s->items_bytes + 64;
  s->res_words = (size_t)kServeMaxV * 9 + kServeMaxV * 3 + 3;
  auto fail_new = [&](hipError_t e, const char* what) {
    int rc = set_error(EFFSC_HIP_ENOMEM, std::string(what) + ": " + hipGetErrorString(e));
    destroy_server(s);
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail_new(e, "stream");
  if ((e = hipMalloc((void**)&s->arena_base, s->arena.cap)) != hipSuccess) return fail_new(e, "arena");
  if ((e = hipMalloc((void**)&s->d_insts, max_active * sizeof(SInst))) != hipSuccess) return fail_new(e, "records");
  if ((e = hipMalloc((void**)&s->d_items, s->items_bytes)) != hipSuccess) return fail_new(e, "items");
  if ((e = hipMalloc((void**)&s->d_partials, s->partial_blocks * kPartWords * 8)) != hipSuccess)
    return fail_new(e, "partials");
  for (int i = 0; i < effsc_hip_server::kRing; i++) {
    if ((e = hipHostMalloc((void**)&s->h_stage[i], s->stage_bytes, hipHostMallocDefault)) != hipSuccess)
      return fail_new(e, "staging");
    if ((e = hipEventCreateWithFlags(&s->stage_ev[i], hipEventDisableTiming)) != hipSuccess)
      return fail_new(e, "event");
  }
  if ((e = hipHostMalloc((void**)&s->h_res, max_active * s->res_words * 8, hipHostMallocDefault)) != hipSuccess)
    return fail_new(e, "results");
  for (size_t i = max_active; i-- > 0;) s->free_slots.push_back((int)i);
  s->th = std::thread(scheduler, s);
  *out = s;
  return EFFSC_HIP_OK;
}

int effsc_hip_server_submit(effsc_hip_server* s, const effsc_hip_request* req, uint64_t* id) {
  if (!s || !id) return set_error(EFFSC_HIP_EINVAL, "null argument");
  int K = 0, ext = 1;
  TRY(request_shape(req, &K, &ext));
  if (!req->transcript) return set_error(EFFSC_HIP_EINVAL, "null transcript");
  auto r = std::make_shared<Req>();
  r->kind = req->kind;
  r->K = K;
  r->ext = ext;
  r->v = req->v;
  for (int i = 0; i < 3; i++) {
    r->seeds[i] = req->seeds[i];
    r->tables[i] = i < K ? req->tables[i] : nullptr;
    // the base-field passes load pairs of words
    if (r->tables[i] && reinterpret_cast<uintptr_t>(r->tables[i]) % 16)
      return set_error(EFFSC_HIP_EINVAL, "borrowed request table not 16-byte aligned");
  }
  r->lay = layout(K, ext, r->v, r->tables);
  if (align256(r->lay.bytes) > s->arena.cap)
    return set_error(EFFSC_HIP_EINVAL, "request footprint exceeds the server's HBM budget");
  const effsc_hip_transcript* t = req->transcript;
  r->fs0.kind = (uint32_t)t->kind;
  if (t->kind == 0) {
    for (int k = 0; k < 25; k++) r->fs0.st[k] = t->fs.sp.st[k];
    r->fs0.pos = t->fs.sp.pos;
  }
  r->fs0.rng = t->rng;
  r->priority = req->priority;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->fatal.empty()) return set_error(EFFSC_HIP_EDEVICE, s->fatal);
  if (s->stop) return set_error(EFFSC_HIP_ESTATE, "server is stopping");
  r->id = s->next_id++;
  r->t_submit = now_s(s->t0);
  s->reqs[r->id] = r;
  s->pending.push_back(r);
  *id = r->id;
  s->cv_submit.notify_all();
  return EFFSC_HIP_OK;
}

int effsc_hip_server_wait(effsc_hip_server* s, uint64_t id, uint64_t* round_polys, uint64_t* challenges,
                          uint64_t* final_value, double* times) {
  if (!s) return set_error(EFFSC_HIP_EINVAL, "null server");
  std::shared_ptr<Req> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(EFFSC_HIP_EINVAL, "unknown request id");
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->reqs.erase(it);
    if (r->status != EFFSC_HIP_OK) return set_error(r->status, s->fatal.empty() ? "request failed" : s->fatal);
  }
  const size_t np = (size_t)r->v * r->K * r->ext, nc = (size_t)r->v * r->ext;
  if (np && round_polys) std::copy(r->proof.begin(), r->proof.begin() + np, round_polys);
  if (nc && challenges) std::copy(r->proof.begin() + np, r->proof.begin() + np + nc, challenges);
  if (final_value) std::copy(r->proof.begin() + np + nc, r->proof.begin() + np + nc + r->ext, final_value);
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_admit;
    times[2] = r->t_done;
  }
  return EFFSC_HIP_OK;
}

int effsc_hip_server_stats(effsc_hip_server* s, uint64_t* ticks, uint64_t* peak_active, uint64_t* peak_bytes,
                           uint64_t* completed) {
  if (!s) return set_error(EFFSC_HIP_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (ticks) *ticks = s->stat_ticks;
  if (peak_active) *peak_active = s->stat_peak_active;
  if (peak_bytes) *peak_bytes = s->arena.peak;
  if (completed) *completed = s->stat_completed;
  return EFFSC_HIP_OK;
}

int effsc_hip_server_now(effsc_hip_server* s, double* t) {
  if (!s || !t) return set_error(EFFSC_HIP_EINVAL, "null argument");
  *t = now_s(s->t0);
  return EFFSC_HIP_OK;
}

void effsc_hip_server_free(effsc_hip_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_submit.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  destroy_server(s);
}

}  // extern "C"

// Device-resident proof loop: the Fiat-Shamir transcript lives in device
// memory, so a whole sumcheck (src/runner.rs:29-69 with the no-op hook) is
// enqueued without a host round trip per round.
//
// Per round while the tables are large:
//   k_fused / k_eval          fold with the challenge read from device memory
//                             and evaluate; one partial per workgroup
//   k_reduce_fs               one workgroup: sum the partials, write the round
//                             polynomial, absorb it into the transcript and
//                             squeeze the next challenge into device memory
// Once the tables are small, k_tail runs every remaining round, the final fold
// and the final value in a single workgroup (tables stay in L2).
#pragma once
#include "kernels.hpp"
#include "pass2.hpp"
#include "shake.hpp"

namespace effsc_hip {

// Transcript state in device memory (host FsTranscript / test RNG mirror).
struct DevFs {
  uint64_t st[25];
  uint32_t pos;
  uint32_t kind;  // 0: SHAKE256 Fiat-Shamir, 1: test RNG (splitmix64)
  uint64_t rng;
};

// Wave-uniform copy of a 64-bit value held by lane 0.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  int lo = __shfl((int)(uint32_t)v, src, 64);
  int hi = __shfl((int)(uint32_t)(v >> 32), src, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Cross-lane moves on the VALU.  DPP row shifts (row_shl:n: lane l reads
// l + n, row_shr:n: l - n, row_ror:n: rotate, all within 16-lane rows) and
// the gfx950 permlane swaps, used for sums over lanes l, l ^ 32 and l, l ^ 16.
// A DPP read from a lane that is off in EXEC returns zero, so these are only
// used where the whole wave is active.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)x, CTRL, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, true);
  return ((uint64_t)hi << 32) | lo;
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}

// Keccak-f[1600] spread over a wave.  Interface: lane i < 25 holds state word
// (x, y) = (i % 5, i / 5) on entry and on return (lanes >= 25: ignored on
// entry, word 0 on return).  All 64 lanes must call.
//
// Inside, row y of the state occupies the 8-lane group g = y, each word as two
// 32-bit halves (lo, hi).  Position p of a group holds x = (p + 4) mod 5:
// p = 1..5 are x = 0..4, p = 0 a copy of x = 4 and p = 6, 7 copies of x = 0, 1,
// so that for p = 1..5 the neighbours x - 1, x + 1, x + 2 are the positions
// p - 1, p + 1, p + 2 (plain DPP row shifts, no selects).  Groups 5..7 are zero.
//   theta: the column sums are one packed reduction: row_ror:8 xors the two
//          groups of each 16-lane row; permlane32_swap of (lo, hi) and a xor
//          leave the lo sums of rows r, r ^ 2 in rows 0, 1 and the hi sums in
//          rows 2, 3; permlane16_swap and a xor finish both (lo parity in rows
//          0, 1, hi parity in rows 2, 3); a last permlane32_swap spreads each
//          half to all rows.  D = C[x - 1] ^ rotl(C[x + 1], 1) is row_shr:1 and
//          row_shl:1 and a three-input xor (v_bitop3).
//   rho+pi: one ds_bpermute per half from the source word's lane (every
//          position, copies included, reads a main lane, so the copies are
//          exact afterwards), rotated by the source's offset r as two
//          v_alignbit_b32 on the halves, taken in swapped order when r = 0 or
//          r > 32
//   chi:   row_shl:1 and row_shl:2 and one v_bitop3 per half (exact for
//          p = 0..5); iota one v_bitop3 on lane (0, 0); row_shr:5 refreshes the
//          copy at p = 6, the only copy theta reads that chi leaves stale
//          (p = 7 is read by chi only, and pi rewrites it)
// About 35 VALU operations and two ds_bpermute per round, against about 60
// with 64-bit lanes (one select per theta, two-instruction chi and iota).
__device__ __forceinline__ uint64_t keccak_lanes(uint64_t a_in) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 3, p = lane & 7;
  const bool active = g < 5;
  const int x = (p + 4) % 5, y = active ? g : 0;
  const int rot[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
  // pi: destination (x, y) reads source (3 (y - 3x) mod 5, x) at its main
  // position sx + 1
  const int sx = (3 * (y - 3 * x) % 5 + 25) % 5, sy = x;
  const int src4 = 4 * (8 * sy + sx + 1);
  const uint32_t r = rot[sx + 5 * sy];
  // rotl by r in (0, 32]: hi' = alignbit(hi, lo, t), lo' = alignbit(lo, hi, t)
  // with t = (32 - r) mod 32; for r in (32, 64) and r = 0 the operands swap
  const uint32_t tsh = (32 - (r & 31)) & 31;
  const bool swp = r == 0 || r > 32;
  const uint32_t m_iota = lane == 1 ? ~0u : 0u;
  const bool copy = p == 6;
  const uint64_t a0 = shfl64(a_in, x + 5 * y);
  uint32_t lo = active ? (uint32_t)a0 : 0u, hi = active ? (uint32_t)(a0 >> 32) : 0u;
  for (int rnd = 0; rnd < 24; rnd++) {
    // theta
    const uint32_t tl = lo ^ dpp32<0x128>(lo), th = hi ^ dpp32<0x128>(hi);  // row_ror:8
    const auto s1 = __builtin_amdgcn_permlane32_swap(tl, th, false, false);
    const uint32_t u = s1[0] ^ s1[1];
    const auto s2 = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    const uint32_t w = s2[0] ^ s2[1];
    const auto s3 = __builtin_amdgcn_permlane32_swap(w, w, false, false);
    const uint32_t cl = s3[0], ch = s3[1];
    const uint32_t pl = dpp32<0x101>(cl), ph = dpp32<0x101>(ch);  // row_shl:1: C[x + 1]
    lo = __builtin_amdgcn_bitop3_b32(lo, dpp32<0x111>(cl), __builtin_amdgcn_alignbit(pl, ph, 31), 0x96);
    hi = __builtin_amdgcn_bitop3_b32(hi, dpp32<0x111>(ch), __builtin_amdgcn_alignbit(ph, pl, 31), 0x96);
    // rho + pi
    const uint32_t sl = (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)lo);
    const uint32_t sh = (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)hi);
    const uint32_t P = swp ? sl : sh, Q = swp ? sh : sl;
    const uint32_t bh = __builtin_amdgcn_alignbit(P, Q, tsh), bl = __builtin_amdgcn_alignbit(Q, P, tsh);
    // chi: b ^ (~b[x + 1] & b[x + 2]) (truth table 0xD2), then iota
    // a ^ (rc & m) (0x78)
    const uint64_t rc = keccak_rc(rnd);
    lo = __builtin_amdgcn_bitop3_b32(bl, dpp32<0x101>(bl), dpp32<0x102>(bl), 0xD2);
    hi = __builtin_amdgcn_bitop3_b32(bh, dpp32<0x101>(bh), dpp32<0x102>(bh), 0xD2);
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)rc, m_iota, 0x78);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(rc >> 32), m_iota, 0x78);
    const uint32_t cpl = dpp32<0x115>(lo), cph = dpp32<0x115>(hi);  // row_shr:5
    lo = active ? (copy ? cpl : lo) : 0u;
    hi = active ? (copy ? cph : hi) : 0u;
  }
  const int i = lane < 25 ? lane : 0;
  return shfl64(((uint64_t)hi << 32) | lo, 8 * (i / 5) + i % 5 + 1);
}

// Sponge with the state spread over a wave (one word per lane); items are
// whole words, so absorbing xors the word into the lane at the position.
struct LaneFs {
  uint64_t a;
  uint32_t pos, kind;  // pos in bytes, a multiple of 8
  uint64_t rng;

  __device__ __forceinline__ void load(const DevFs* t) {
    const int lane = threadIdx.x & 63;
    a = t->st[lane < 25 ? lane : 0];
    pos = __builtin_amdgcn_readfirstlane(t->pos);
    kind = __builtin_amdgcn_readfirstlane(t->kind);
    rng = uniform64(t->rng);
  }
  __device__ __forceinline__ void store(DevFs* t) const {
    const int lane = threadIdx.x & 63;
    if (lane < 25) t->st[lane] = a;
    if (lane == 0) {
      t->pos = pos;
      t->kind = kind;
      t->rng = rng;
    }
  }
  __device__ __forceinline__ void absorb_word(uint64_t x) {
    if ((uint32_t)(threadIdx.x & 63) == (pos >> 3)) a ^= x;
    pos += 8;
    if (pos == kShakeRate) {
      a = keccak_lanes(a);
      pos = 0;
    }
  }
  __device__ __forceinline__ void send(int ext, const uint64_t* w) {
    if (kind != 0) return;
    absorb_word(1);
#pragma unroll
    for (int i = 0; i < 3; i++)
      if (i < ext) absorb_word(w[i]);
  }
  // Writes out[0 .. ext): callers read no further.
  __device__ __forceinline__ void challenge(int ext, uint64_t* out) {
    if (kind == 1) {
      rng += 0x9E3779B97F4A7C15ull;
      out[0] = gl::canon(mix64(rng));
      if (ext > 1) out[1] = 0;
      if (ext > 2) out[2] = 0;
      return;
    }
    // pad the sponge itself to the block boundary and permute: the padding
    // stays in the stream, the next word starts a fresh block
    absorb_word(2);
    const uint32_t lane = threadIdx.x & 63;
    if (lane == (pos >> 3)) a ^= 0x1F;  // SHAKE padding
    if (lane == (kShakeRate - 1) / 8) a ^= (uint64_t)0x80 << (8 * ((kShakeRate - 1) % 8));
    a = keccak_lanes(a);
    pos = 0;
    const uint64_t c = a;
    uint64_t w0 = shfl64(c, 0), w1 = shfl64(c, 1), w2 = shfl64(c, 2), w3 = shfl64(c, 3), w4 = shfl64(c, 4),
             w5 = shfl64(c, 5);
    out[0] = uniform64(gl::reduce128(w0, w1));
    if (ext > 1) out[1] = uniform64(gl::reduce128(w2, w3));
    if (ext > 2) out[2] = uniform64(gl::reduce128(w4, w5));
  }
};

// Called by all 64 lanes of one wave; lane 0 holds the K round values.
// Writes them (lifted to ext words) to `poly`, absorbs them, squeezes the
// challenge into chal_out and chal_cur; lane 0 does the stores.
template <class F, int K>
__device__ __forceinline__ void emit_round_fs(const F* s, int ext, LaneFs& fs, uint64_t* poly, uint64_t* chal_out,
                                              uint64_t* chal_cur, bool write = true) {
  // write false: only chal_cur receives the challenge (poly and chal_out
  // are left alone)
  const bool lane0 = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    uint64_t wd[3] = {0, 0, 0};
    gl::store_words(s[i], wd);
#pragma unroll
    for (int c = 0; c < 3; c++) wd[c] = uniform64(wd[c]);
    if (lane0 && write)
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (c < ext) poly[i * ext + c] = wd[c];
    fs.send(ext, wd);
  }
  uint64_t ch[3];
  fs.challenge(ext, ch);
  if (lane0) {
    // exactly ext words each: chal_cur may be one slot of a per-instance array
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (c < ext) {
        if (write) chal_out[c] = ch[c];
        chal_cur[c] = ch[c];
      }
    }
  }
}

// emit_round_fs with the sponge loaded from and stored back to t.
template <class F, int K>
__device__ __forceinline__ void emit_round(const F* s, int ext, DevFs* t, uint64_t* poly, uint64_t* chal_out,
                                           uint64_t* chal_cur) {
  LaneFs fs;
  fs.load(t);
  emit_round_fs<F, K>(s, ext, fs, poly, chal_out, chal_cur);
  fs.store(t);
}

// Cross-block sum of one round plus the transcript step.
constexpr int kReduceBlock = 1024;

// Sums `nblocks` partials of W canonical words each, word by word, into
// sh[0..W) (sh: blockDim.x words of shared memory; every thread calls, the
// result is valid after the return).  Thread t adds word t % W of blocks
// t / W, t / W + R, ... (R = blockDim.x / W rows; coalesced: consecutive
// threads read consecutive words) in a 128-bit sum, four loads in flight,
// then the R rows are added pairwise in shared memory.  Against a sum per
// block and a block-wide reduction of W words: one word per thread and no
// cross-lane exchanges (F_p^3 pass values, 24 words of 2048 workgroups:
// 3.6 + 7.8 us for the partial and block sums before).
template <int W>
__device__ __forceinline__ void sum_partials_words(const uint64_t* partials, unsigned nblocks, uint64_t* sh) {
  const unsigned R = blockDim.x / W, t = threadIdx.x, c = t % W, row = t / W;
  uint64_t lo = 0, hi = 0;
  auto acc = [&](uint64_t x) {
    lo += x;
    hi += lo < x;
  };
  if (row < R) {
    unsigned b = row;
    for (; b + 3 * R < nblocks; b += 4 * R) {
      const uint64_t x0 = partials[(size_t)b * W + c], x1 = partials[(size_t)(b + R) * W + c];
      const uint64_t x2 = partials[(size_t)(b + 2 * R) * W + c], x3 = partials[(size_t)(b + 3 * R) * W + c];
      acc(x0);
      acc(x1);
      acc(x2);
      acc(x3);
    }
    for (; b < nblocks; b += R) acc(partials[(size_t)b * W + c]);
  }
  sh[t] = row < R ? gl::reduce128(lo, hi) : 0;
  __syncthreads();
  // rows [0, n - h) take rows [h, n): disjoint readers and read rows
  for (unsigned n = R; n > 1;) {
    const unsigned h = (n + 1) >> 1;
    if (row < n - h) sh[t] = gl::add(sh[t], sh[t + h * W]);
    __syncthreads();
    n = h;
  }
}

template <class F, int K>
__global__ __launch_bounds__(kReduceBlock) void k_reduce_fs(Reduce r, unsigned nblocks, DevFs* t, int ext, uint64_t* poly,
                                                      uint64_t* chal_out, uint64_t* chal_cur) {
  __shared__ uint64_t sh[kReduceBlock];
  sum_partials_words<K * F::D>(r.partials, nblocks, sh);
  if (threadIdx.x >= 64) return;
  F s[K];
#pragma unroll
  for (int i = 0; i < K; i++) s[i] = gl::load_words<F>(sh + i * F::D);
  emit_round<F, K>(s, ext, t, poly, chal_out, chal_cur);
}

// Two transcript steps of a two-round pass, by the 64 lanes of one wave
// whose lane 0 holds the reduced pass values: round s values, challenge x,
// round s+1 values (the pass polynomials at x), challenge y.  poly receives
// 2K words, chals (x, y), cur (x, y): the pending folds of the next pass.
// `fs` holds the sponge loaded from t (callers issue that load early).  With
// write false nothing goes to memory but (x, y) to xy_sh (when given), which
// lane 0 writes in either case.
template <int K>
__device__ __forceinline__ void emit_two_rounds(const Fp* s, LaneFs& fs, DevFs* t, uint64_t* poly, uint64_t* chals,
                                                uint64_t* cur, bool write = true, uint64_t* xy_sh = nullptr) {
  constexpr int NV = pass_values<K>();
  const bool lane0 = (threadIdx.x & 63) == 0 && write;
  uint64_t v[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = uniform64(s[i].v);
#pragma unroll
  for (int i = 0; i < K; i++) {
    uint64_t w[3] = {v[i], 0, 0};
    if (lane0) poly[i] = v[i];
    fs.send(1, w);
  }
  uint64_t x[3];
  fs.challenge(1, x);
#pragma unroll
  for (int i = 0; i < K; i++) {
    uint64_t w[3] = {pass_eval_next<K>(v, i, x[0]), 0, 0};
    if (lane0) poly[K + i] = w[0];
    fs.send(1, w);
  }
  uint64_t y[3];
  fs.challenge(1, y);
  if (lane0) {
    chals[0] = x[0];
    chals[1] = y[0];
    cur[0] = x[0];
    cur[1] = y[0];
  }
  if (xy_sh && (threadIdx.x & 63) == 0) {
    xy_sh[0] = x[0];
    xy_sh[1] = y[0];
  }
  if (write) fs.store(t);
}

// Two-round pass (pass2.hpp) over base-field tables: sums the pass values,
// then runs the two transcript steps (emit_two_rounds).
// The sponge is loaded from t_in (t, or the pinned host staging of a starting
// proof) and stored to t.
template <int K>
__global__ __launch_bounds__(kReduceBlock) void k_reduce_fs2(Reduce r, unsigned nblocks, const DevFs* t_in, DevFs* t,
                                                       uint64_t* poly, uint64_t* chals, uint64_t* cur) {
  constexpr int NV = pass_values<K>();
  __shared__ uint64_t sh[kReduceBlock];
  // the sponge load overlaps the partial sums
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t_in);
  sum_partials_words<NV>(r.partials, nblocks, sh);
  if (threadIdx.x >= 64) return;
  Fp s[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = Fp{sh[i]};
  emit_two_rounds<K>(s, fs, t, poly, chals, cur);
}

// A small two-round fold pass (pass2.hpp) with the previous pass's transcript
// step in the same launch, for grids of at most kFuseMaxGrid workgroups (all
// resident at once, so the step below costs the launch its latency once).
// Every workgroup issues its units' source loads first, sums the previous
// pass's partials (rin, nin blocks) and runs the two transcript steps on wave
// 0 (the steps of k_reduce_fs2, so every workgroup derives the same two
// challenges), then applies both folds, stores the folded level and
// accumulates this pass's values into rout.  Workgroup 0 alone writes the
// proof words, the challenges, the pending folds and the sponge; the sponge
// goes to t_out and the partials to rout, neither of which this launch reads.
// One unit pair per thread: the grid covers q / 2 threads.  Every workgroup
// reads all nin partials, so the previous pass's grid is bounded too
// (kFuseMaxIn): the inner product's 2^20-element level (512 workgroups after
// a 2048-block pass) took 38.8 us fused against 36.5 us as a pass plus a
// transcript launch, the 2^18 level 21.3 against 28.8 us.
constexpr unsigned kFuseMaxGrid = 128, kFuseMaxIn = 1024;
template <int K>
__global__ __launch_bounds__(kBlock) void k_pass_fs2(Views<Fp, K> src, Views<Fp, K> dst, size_t q, Reduce rin,
                                                     unsigned nin, Reduce rout, const DevFs* t_in, DevFs* t_out,
                                                     uint64_t* poly, uint64_t* chals, uint64_t* cur) {
  constexpr int NV = pass_values<K>(), NR = pass_raw<K>();
  __shared__ uint64_t sh[kBlock];
  __shared__ uint64_t xy[2];
  const size_t L = 4 * q, u = gtid() * 2;
  const bool act = u < q;
  // level-4L source: element e + kL (k < 4) of region j for units u, u + 1
  v2u64 X[K][4][4];
  if (act) {
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) X[i][j][k] = *reinterpret_cast<const v2u64*>(src.t[i].p + u + j * q + k * L);
  }
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t_in);
  sum_partials_words<NV>(rin.partials, nin, sh);
  if (threadIdx.x < 64) {
    Fp s[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) s[i] = Fp{sh[i]};
    emit_two_rounds<K>(s, fs, t_out, poly, chals, cur, blockIdx.x == 0, xy);
  }
  __syncthreads();
  const uint64_t c1 = xy[0], c2 = xy[1], c12 = gl::mul(c1, c2);
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < NR; i++) acc.a[i].zero();
  if (act) {
    uint64_t E0[K][4], E1[K][4];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        E0[i][j] = gl::fold2(X[i][j][0].x, X[i][j][1].x, X[i][j][2].x, X[i][j][3].x, c1, c2, c12);
        E1[i][j] = gl::fold2(X[i][j][0].y, X[i][j][1].y, X[i][j][2].y, X[i][j][3].y, c1, c2, c12);
        *reinterpret_cast<ulonglong2*>(dst.t[i].p + u + j * q) = make_ulonglong2(E0[i][j], E1[i][j]);
      }
    pass_accumulate<K>(E0, acc);
    pass_accumulate<K>(E1, acc);
  }
  Fp raw[NR], vals[NV];
#pragma unroll
  for (int i = 0; i < NR; i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  finish<Fp, NV>(vals, rout);
}

constexpr int kTailBlock = 1024;
// Extension-field products of two or more tables need more than the 128
// VGPRs a 1024-thread workgroup allows (they spilled there): 256 threads; a
// single F_p^3 table 512 (about 140 VGPRs).
template <class F, int K>
constexpr int tail_block() {
  return (F::D >= 2 && K >= 2) ? 256 : (F::D == 3 ? 512 : kTailBlock);
}

// All remaining rounds in one workgroup.  On entry the K tables have pow2
// length L (>= 2); they are in `in` (field FI) when from_input, else in
// `work` (field F).  Round r (0 <= r < rounds) folds with the current
// challenge when have_chal (fused with the evaluation, L >= 4), evaluates
// pairs (k, k + L/2) and runs the transcript step; afterwards the last
// challenge folds the tables once more and the final value
// (prod_i table_i[0] when the length reached 1, else 0) goes to fin.
template <class FI, class F, int K>
__global__ __launch_bounds__((tail_block<F, K>())) void k_tail(Views<FI, K> in, Views<F, K> work, int from_input, size_t L,
                                                     int rounds, int have_chal, DevFs* t, int ext, uint64_t* polys,
                                                     uint64_t* chals, uint64_t* chal_cur, uint64_t* fin, uint64_t w) {
  __shared__ uint64_t chal_sh[3];
  if (threadIdx.x < 3) chal_sh[threadIdx.x] = have_chal ? chal_cur[threadIdx.x] : 0;
  // the sponge stays in wave 0's registers for the whole tail
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  __syncthreads();
  const unsigned tid = threadIdx.x, nt = blockDim.x;
  bool src_in = from_input != 0;
  for (int r = 0; r < rounds; r++) {
    F acc[K];
#pragma unroll
    for (int i = 0; i < K; i++) acc[i] = gl::zero<F>();
    if (have_chal) {
      F c = gl::load_words<F>(chal_sh);
      size_t q = L / 4;
      // thread k reads positions k, k+q, k+2q, k+3q and alone writes k and
      // k+q; positions >= 2q are only read, so the fold runs in place
      for (size_t k = tid; k < q; k += nt) {
        F lo[K], hi[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          if (src_in) {
            FI x0 = in.t[i].ld(k), x1 = in.t[i].ld(k + q), x2 = in.t[i].ld(k + 2 * q), x3 = in.t[i].ld(k + 3 * q);
            lo[i] = fold1<F>(x0, x2, c, w);
            hi[i] = fold1<F>(x1, x3, c, w);
          } else {
            F x0 = work.t[i].ld(k), x1 = work.t[i].ld(k + q), x2 = work.t[i].ld(k + 2 * q), x3 = work.t[i].ld(k + 3 * q);
            lo[i] = fold1<F>(x0, x2, c, w);
            hi[i] = fold1<F>(x1, x3, c, w);
          }
        }
#pragma unroll
        for (int i = 0; i < K; i++) {
          work.t[i].st(k, lo[i]);
          work.t[i].st(k + q, hi[i]);
        }
        eval_pair<F, K>(lo, hi, acc, w);
      }
      L /= 2;
      src_in = false;
    } else {
      size_t half = L / 2;
      for (size_t k = tid; k < half; k += nt) {
        F lo[K], hi[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          if (src_in) {
            lo[i] = gl::lift_to<F>(in.t[i].ld(k));
            hi[i] = gl::lift_to<F>(in.t[i].ld(k + half));
          } else {
            lo[i] = work.t[i].ld(k);
            hi[i] = work.t[i].ld(k + half);
          }
        }
        eval_pair<F, K>(lo, hi, acc, w);
      }
    }
    block_reduce<F, K>(acc);
    if (tid < 64) emit_round_fs<F, K>(acc, ext, fs, polys + (size_t)r * K * ext, chals + (size_t)r * ext, chal_sh);
    __syncthreads();
    have_chal = 1;
  }
  // finalize: fold with the last challenge (src/runner.rs:57-60)
  if (have_chal && L >= 2) {
    F c = gl::load_words<F>(chal_sh);
    size_t half = L / 2;
    for (size_t k = tid; k < half; k += nt) {
#pragma unroll
      for (int i = 0; i < K; i++) {
        F v;
        if (src_in)
          v = fold1<F>(in.t[i].ld(k), in.t[i].ld(k + half), c, w);
        else
          v = fold1<F>(work.t[i].ld(k), work.t[i].ld(k + half), c, w);
        work.t[i].st(k, v);
      }
    }
    L = half;
    __syncthreads();
  }
  if (tid == 0) {
    F v = gl::zero<F>();
    if (L == 1) {
      v = work.t[0].ld(0);
#pragma unroll
      for (int i = 1; i < K; i++) v = gl::mul(v, work.t[i].ld(0), w);
    }
    uint64_t wd[3] = {0, 0, 0};
    gl::store_words(v, wd);
    for (int c = 0; c < ext; c++) fin[c] = wd[c];
  }
  if (tid < 64) fs.store(t);
}

// Transcript step of one round with the sponge held in registers by the 64
// lanes of one wave (lane 0 holds the K base-field values): writes poly
// (K words) and chal_out, returns the challenge (wave-uniform).
template <int K>
__device__ __forceinline__ uint64_t emit_round_regs(LaneFs& fs, const Fp* s, uint64_t* poly, uint64_t* chal_out) {
  const bool lane0 = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    uint64_t w[3] = {uniform64(s[i].v), 0, 0};
    if (lane0) poly[i] = w[0];
    fs.send(1, w);
  }
  uint64_t ch[3];
  fs.challenge(1, ch);
  if (lane0) chal_out[0] = ch[0];
  return ch[0];
}

// Sum of K values over one wave, valid in every lane.
template <int K>
__device__ __forceinline__ void wave_sum(Fp* acc) {
#pragma unroll
  for (int i = 0; i < K; i++)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc[i] = gl::add(acc[i], shfl_xor(acc[i], m));
}

constexpr int kTailRegThreads = 1024;

// k_tail for base-field tables with K <= 2 and 2^11 <= L1 <= 2^14 elements
// after the pending fold (L1 = M0 * 1024), with the tables in registers.
//
// Thread t holds elements t + 1024 m (m < Mc) of every table: while Mc >= 2
// the pairs (k, k + L/2) of a round are (m, m + Mc/2) of one thread, so a
// round is an in-register evaluation, a block reduction, the transcript step
// by wave 0 (sponge kept in its registers for the whole tail) and an
// in-register fold.  At 1024 elements the table moves through LDS to wave 0,
// lane l holding l + 64 m, and the remaining rounds run in that wave alone:
// in-lane pairs down to 64 elements, then pairs (l, l + h) by shuffles; no
// workgroup barrier after the hand-over.  The results are those of k_tail:
// same round polynomials, challenges, final tables (written back to `work`
// when the rounds stop early) and final value.
template <int K, int M0>
__global__ __launch_bounds__(kTailRegThreads) void k_tail_reg(Views<Fp, K> in, Views<Fp, K> work, int from_input,
                                                             int rounds, int have_chal, DevFs* t, uint64_t* polys,
                                                             uint64_t* chals, const uint64_t* chal_cur, uint64_t* fin) {
  static_assert(M0 >= 2 && M0 <= 16 && (M0 & (M0 - 1)) == 0, "M0: 2, 4, 8 or 16");
  constexpr int LG0 = ilog2c(M0);
  // loops below count log2 sizes so that they unroll and index registers
  __shared__ uint64_t tab[K][kTailRegThreads];
  __shared__ uint64_t chal_sh;
  const unsigned tid = threadIdx.x;
  const int lane = tid & 63;
  const uint64_t w = 0;
  const Views<Fp, K>& src = from_input ? in : work;
  Fp x[K][M0];
  if (have_chal) {
    // the pending fold: length 2 L1 -> L1, pairs (m, m + M0) of this thread
    const Fp c{chal_cur[0]};
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int m = 0; m < M0; m++)
        x[i][m] = fold1<Fp>(src.t[i].ld(tid + 1024 * m), src.t[i].ld(tid + 1024 * (m + M0)), c, w);
  } else {
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int m = 0; m < M0; m++) x[i][m] = src.t[i].ld(tid + 1024 * m);
  }
  LaneFs fs;
  if (tid < 64) fs.load(t);
  int r = 0;
  int Lc = 1024 * M0;  // current table length
  // ---- all waves: rounds with in-thread pairs
#pragma unroll
  for (int lg = LG0; lg >= 1; lg--) {
    const int Mc = 1 << lg;
    if (r < rounds) {
      Fp acc[K];
#pragma unroll
      for (int i = 0; i < K; i++) acc[i] = gl::zero<Fp>();
#pragma unroll
      for (int m = 0; m < Mc / 2; m++) {
        Fp lo[K], hi[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          lo[i] = x[i][m];
          hi[i] = x[i][m + Mc / 2];
        }
        eval_pair<Fp, K>(lo, hi, acc, w);
      }
      block_reduce<Fp, K>(acc);
      if (tid < 64) {
        const uint64_t c = emit_round_regs<K>(fs, acc, polys + (size_t)r * K, chals + r);
        if (tid == 0) chal_sh = c;
      }
      __syncthreads();
      const Fp c{chal_sh};
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int m = 0; m < Mc / 2; m++) x[i][m] = fold1<Fp>(x[i][m], x[i][m + Mc / 2], c, w);
      Lc /= 2;
      r++;
    }
  }
  if (r < rounds) {
    // ---- hand-over: 1024 elements to wave 0 through LDS; lane l owns
    // positions l + 64 m of `tab` (no other lane touches them, so the wave
    // needs no barrier)
#pragma unroll
    for (int i = 0; i < K; i++) tab[i][tid] = x[i][0].v;
    __syncthreads();
    if (tid >= 64) return;
#pragma unroll
    for (int lg = 4; lg >= 1; lg--) {
      const int Mw = 1 << lg;
      if (r < rounds) {
        Fp acc[K];
#pragma unroll
        for (int i = 0; i < K; i++) acc[i] = gl::zero<Fp>();
#pragma unroll
        for (int m = 0; m < Mw / 2; m++) {
          Fp lo[K], hi[K];
#pragma unroll
          for (int i = 0; i < K; i++) {
            lo[i] = Fp{tab[i][lane + 64 * m]};
            hi[i] = Fp{tab[i][lane + 64 * (m + Mw / 2)]};
          }
          eval_pair<Fp, K>(lo, hi, acc, w);
        }
        wave_sum<K>(acc);
        const Fp c{emit_round_regs<K>(fs, acc, polys + (size_t)r * K, chals + r)};
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
          for (int m = 0; m < Mw / 2; m++)
            tab[i][lane + 64 * m] = fold1<Fp>(Fp{tab[i][lane + 64 * m]}, Fp{tab[i][lane + 64 * (m + Mw / 2)]}, c, w).v;
        Lc /= 2;
        r++;
      }
    }
    // 64 elements, one per lane: pairs (l, l + h)
    Fp z[K];
#pragma unroll
    for (int i = 0; i < K; i++) z[i] = Fp{tab[i][lane]};
#pragma unroll
    for (int lg = 5; lg >= 0; lg--) {
      const int h = 1 << lg;
      if (r < rounds) {
        Fp lo[K], hi[K], acc[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          lo[i] = z[i];
          hi[i] = shfl_xor(z[i], h);  // lanes l < h read l + h
          acc[i] = gl::zero<Fp>();
        }
        if (lane < h) eval_pair<Fp, K>(lo, hi, acc, w);
        wave_sum<K>(acc);
        const Fp c{emit_round_regs<K>(fs, acc, polys + (size_t)r * K, chals + r)};
#pragma unroll
        for (int i = 0; i < K; i++) z[i] = fold1<Fp>(lo[i], hi[i], c, w);
        Lc /= 2;
        r++;
      }
    }
    // write back the tables (length Lc; lanes >= Lc hold no element)
    if (Lc >= 64) {
      const int Mw = Lc / 64;
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int m = 0; m < 16; m++)
          if (m < Mw) work.t[i].st(lane + 64 * m, Fp{tab[i][lane + 64 * m]});
    } else {
#pragma unroll
      for (int i = 0; i < K; i++)
        if (lane < Lc) work.t[i].st(lane, z[i]);
    }
    if (lane == 0) {
      Fp v = gl::zero<Fp>();
      if (Lc == 1) {
        v = z[0];
#pragma unroll
        for (int i = 1; i < K; i++) v = gl::mul(v, z[i], w);
      }
      fin[0] = v.v;
    }
    fs.store(t);
    return;
  }
  // rounds ended with the tables still spread over the workgroup
  const int Mc = Lc / 1024;
#pragma unroll
  for (int i = 0; i < K; i++)
#pragma unroll
    for (int m = 0; m < M0; m++)
      if (m < Mc) work.t[i].st(tid + 1024 * m, x[i][m]);
  if (tid == 0) fin[0] = 0;  // more than one element left
  if (tid < 64) fs.store(t);
}

// ---------------------------------------------------------------------------
// Pipelined tail for base-field tables (K <= 2) of L0 <= pipe_max<K>()
// elements after the pending folds.  Round s+1's values are polynomials of
// degree K in round s's challenge (the two-round pass algebra, pass2.hpp), so
// while wave 0 runs round s's transcript step the worker waves (1..8) fold the
// tables with c_{s-1} and accumulate the raw sums R_s of the folded level,
// whose coefficient part gives round s+1's values at c_s.  A round then costs
// one transcript step and K polynomial evaluations on wave 0; the fold and
// the sums run beside the Keccak instead of before it.
//   init     workers: T_0 (the pending folds applied) into LDS, R_0
//   step 0   wave 0: round 0's values from R_0, absorb, c_0
//   step s   wave 0: round s's values = R_{s-1}'s coefficients at c_{s-1},
//            absorb, c_s;  workers: T_s = fold(T_{s-1}, c_{s-1}), R_s (only
//            while a round s + 1 follows)
//   end      workers: the final fold with the last challenge, the tables back
//            to `work`, the final value
// One barrier per round.  The tables alternate between two LDS buffers
// (T_s in A for even s, in B for odd s), the raw sums and challenges between
// two slots, so a step's reads and writes never meet.  Same round
// polynomials, challenges, final tables and final value as k_tail_reg.
// Needs 1 <= rounds <= log2(L0).
constexpr int kPipeWorkers = 512;
constexpr int kPipeThreads = 64 + kPipeWorkers;
template <int K>
constexpr int pipe_max() {
  return K == 1 ? 8192 : 4096;
}

// The workers' raw sums, reduced over each wave into part[wave - 1].
template <int K>
__device__ __forceinline__ void pipe_part(const PassAcc<K>& acc, uint64_t (*part)[pass_raw<K>()]) {
  constexpr int NR = pass_raw<K>(), LP = ilog2c(pow2_ceil(NR));
  Fp raw[NR];
#pragma unroll
  for (int i = 0; i < NR; i++) raw[i] = acc.a[i].reduce();
  const Fp v = wave_reduce_t<Fp, NR>(raw);
  const int lane = threadIdx.x & 63, idx = lane >> (6 - LP);
  if ((lane & ((64 >> LP) - 1)) == 0 && idx < NR) part[(threadIdx.x >> 6) - 1][idx] = v.v;
}

// EQ (EqFactoredProver, K = 2): table 1 of T_0 is EQ[e] = eql[e >> m_log]
// eqr[e & (2^m_log - 1)], computed as it is loaded (table 0 is p with its
// pending folds), so no separate preparation launch writes it first.
template <int K, bool EQ = false>
__global__ __launch_bounds__(kPipeThreads) void k_tail_pipe(Views<Fp, K> src, Views<Fp, K> work, int L0, int pend,
                                                            int rounds, const DevFs* t_in, DevFs* t, uint64_t* polys,
                                                            uint64_t* chals,
                                                            const uint64_t* chal_cur, uint64_t* fin,
                                                            const uint64_t* eql = nullptr,
                                                            const uint64_t* eqr = nullptr, unsigned m_log = 0,
                                                            uint64_t* proof_host = nullptr,
                                                            const uint64_t* proof_dev = nullptr,
                                                            int proof_words = 0) {
  static_assert(!EQ || K == 2, "EQ: p and EQ");
  constexpr int NR = pass_raw<K>(), NV = pass_values<K>(), NW = kPipeWorkers / 64, LA = pipe_max<K>();
  __shared__ uint64_t tabA[K][LA], tabB[K][LA / 2];
  __shared__ uint64_t part[2][NW][NR];
  __shared__ uint64_t cs[2];
  const int tid = threadIdx.x, lane = tid & 63, wk = tid - 64;
  const bool w0 = tid < 64;
  LaneFs fs;
  if (w0) fs.load(t_in);
  if (!w0) {
    // T_0: the level-L0 tables from the stored level with pend pending folds
    auto init = [&](auto pc) {
      constexpr int P = decltype(pc)::value;
      const uint64_t c1 = P >= 1 ? chal_cur[0] : 0, c2 = P == 2 ? chal_cur[1] : 0;
      const uint64_t c12 = P == 2 ? gl::mul(c1, c2) : 0;
      const int q = L0 / 4;
      PassAcc<K> acc;
#pragma unroll
      for (int i = 0; i < NR; i++) acc.a[i].zero();
      for (int u0 = wk; u0 < q; u0 += kPipeWorkers) {
        // the unit index through an empty asm: otherwise loop strength
        // reduction keeps a 64-bit pointer per region and table live
        int u = u0;
        asm volatile("" : "+v"(u));
        uint64_t E[K][4];
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int e = u + j * q;
            if (EQ && i == 1)
              E[i][j] = gl::mul(eql[e >> m_log], eqr[e & ((1 << m_log) - 1)]);
            else
              E[i][j] = fold_chain<P>(src.t[i].p, e, L0, c1, c2, c12);
            tabA[i][e] = E[i][j];
            if constexpr (P > 0) __builtin_amdgcn_sched_barrier(0);
          }
        pass_accumulate<K>(E, acc);
      }
      pipe_part<K>(acc, part[0]);
    };
    if (pend == 0)
      init(std::integral_constant<int, 0>{});
    else if (pend == 1)
      init(std::integral_constant<int, 1>{});
    else
      init(std::integral_constant<int, 2>{});
  }
  __syncthreads();
  uint64_t cprev = 0;
  for (int s = 0; s < rounds; s++) {
    if (w0) {
      // raw sums R_{s-1} (R_0 at s = 0): lane i < NR sums value i over the
      // worker waves, then every lane takes all of them
      const int slot = s == 0 ? 0 : (s - 1) & 1;
      Fp sv = gl::zero<Fp>();
      if (lane < NR)
#pragma unroll
        for (int w = 0; w < NW; w++) sv = gl::add(sv, Fp{part[slot][w][lane]});
      Fp raw[NR], vals[NV], V[K];
#pragma unroll
      for (int i = 0; i < NR; i++) raw[i] = readlane(sv, i);
      pass_values_from_raw<K>(raw, vals);
      if (s == 0) {
#pragma unroll
        for (int i = 0; i < K; i++) V[i] = vals[i];
      } else {
        uint64_t cw[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) cw[i] = vals[i].v;
#pragma unroll
        for (int i = 0; i < K; i++) V[i] = Fp{pass_eval_next<K>(cw, i, cprev)};
      }
      cprev = emit_round_regs<K>(fs, V, polys + (size_t)s * K, chals + s);
      if (lane == 0) cs[s & 1] = cprev;
    } else if (s >= 1) {
      const uint64_t c = cs[(s - 1) & 1];
      const int Ls = L0 >> s;  // length of T_s; T_{s-1} has 2 Ls
      // T_{s-1} in A for even s - 1, in B for odd; T_s into the other
      const bool src_b = (s - 1) & 1;
      auto ld = [&](int i, int e) -> uint64_t { return src_b ? tabB[i][e] : tabA[i][e]; };
      auto st = [&](int i, int e, uint64_t x) {
        if (src_b)
          tabA[i][e] = x;
        else
          tabB[i][e] = x;
      };
      if (s + 1 < rounds) {
        const int q = Ls / 4;
        PassAcc<K> acc;
#pragma unroll
        for (int i = 0; i < NR; i++) acc.a[i].zero();
        for (int u = wk; u < q; u += kPipeWorkers) {
          uint64_t E[K][4];
#pragma unroll
          for (int i = 0; i < K; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const int e = u + j * q;
              const uint64_t lo = ld(i, e), hi = ld(i, e + Ls);
              const uint64_t x = gl::add(lo, gl::mul(c, gl::sub(hi, lo)));
              E[i][j] = x;
              st(i, e, x);
            }
          pass_accumulate<K>(E, acc);
        }
        pipe_part<K>(acc, part[s & 1]);
      } else {
        for (int e = wk; e < Ls; e += kPipeWorkers)
#pragma unroll
          for (int i = 0; i < K; i++) {
            const uint64_t lo = ld(i, e), hi = ld(i, e + Ls);
            st(i, e, gl::add(lo, gl::mul(c, gl::sub(hi, lo))));
          }
      }
    }
    __syncthreads();
  }
  if (w0) {
    fs.store(t);
  } else {
    // the final fold with c_{rounds-1}: T_{rounds-1} (length 2 Lf) -> work
    const uint64_t c = cs[(rounds - 1) & 1];
    const int Lf = L0 >> rounds;
    const bool inB = (rounds - 1) & 1;
    for (int e = wk; e < Lf; e += kPipeWorkers) {
      uint64_t prod = 1;
#pragma unroll
      for (int i = 0; i < K; i++) {
        const uint64_t lo = inB ? tabB[i][e] : tabA[i][e], hi = inB ? tabB[i][e + Lf] : tabA[i][e + Lf];
        const uint64_t x = gl::add(lo, gl::mul(c, gl::sub(hi, lo)));
        work.t[i].p[e] = x;
        prod = i == 0 ? x : gl::mul(prod, x);
      }
      if (e == 0) fin[0] = Lf == 1 ? prod : 0;
    }
  }
  if (proof_host) {
    // the proof block (this launch's words and the earlier launches') straight
    // into the pinned host staging, so the host waits for this kernel only
    __syncthreads();
    for (int i = tid; i < proof_words; i += kPipeThreads) proof_host[i] = proof_dev[i];
  }
}

// ---------------------------------------------------------------------------
// k_tail for products of K >= 3 F_p^3 tables of 2^8 or 2^9 elements after the
// pending fold, with the tables in registers (as k_tail_reg for the base
// field): 256 threads, thread t holding elements t + 256 m (m < M0) of every
// table.  While a thread holds two or more elements a round's pairs are in
// the thread; at 256 elements the tables move through LDS to wave 0 (lane l
// holding l + 64 m) and the remaining rounds run in that wave alone, in-lane
// down to 64 elements, then by shuffles.  The sponge stays in wave 0's
// registers.  Same round polynomials, challenges, final tables (written back
// when the rounds stop early) and final value as k_tail.
constexpr int kTailReg3Threads = 256;

// Transcript step of one F_p^3 round with the sponge in registers: writes
// poly (K x 3 words) and chal_out (3 words), returns the challenge
// (wave-uniform).
template <int K>
__device__ __forceinline__ gl::Fp3 emit_round_regs3(LaneFs& fs, const gl::Fp3* s, uint64_t* poly, uint64_t* chal_out) {
  const bool lane0 = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    uint64_t wd[3] = {uniform64(s[i].c0), uniform64(s[i].c1), uniform64(s[i].c2)};
    if (lane0)
#pragma unroll
      for (int c = 0; c < 3; c++) poly[3 * i + c] = wd[c];
    fs.send(3, wd);
  }
  uint64_t ch[3];
  fs.challenge(3, ch);
  if (lane0)
#pragma unroll
    for (int c = 0; c < 3; c++) chal_out[c] = ch[c];
  return gl::Fp3{ch[0], ch[1], ch[2]};
}

template <int K, int M0>
__global__ __launch_bounds__(kTailReg3Threads) void k_tail_reg3(Views<gl::Fp3, K> work, int rounds, DevFs* t,
                                                               uint64_t* polys, uint64_t* chals,
                                                               const uint64_t* chal_cur, uint64_t* fin, uint64_t w) {
  using gl::Fp3;
  static_assert(M0 == 1 || M0 == 2, "M0: 1 or 2");
  constexpr int LG0 = ilog2c(M0);
  constexpr int T = kTailReg3Threads;
  __shared__ uint64_t tab[K][3][T];
  __shared__ uint64_t chal_sh[3];
  const unsigned tid = threadIdx.x;
  const int lane = tid & 63;
  Fp3 x[K][M0];
  {
    // the pending fold: length 2 L1 -> L1, pairs (m, m + M0) of this thread
    const Fp3 c = gl::load_words<Fp3>(chal_cur);
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int m = 0; m < M0; m++)
        x[i][m] = fold1<Fp3>(work.t[i].ld(tid + T * m), work.t[i].ld(tid + T * (m + M0)), c, w);
  }
  LaneFs fs;
  if (tid < 64) fs.load(t);
  int r = 0;
  int Lc = T * M0;
  // ---- all waves: rounds with in-thread pairs
#pragma unroll
  for (int lg = LG0; lg >= 1; lg--) {
    const int Mc = 1 << lg;
    if (r < rounds) {
      Fp3 acc[K];
#pragma unroll
      for (int i = 0; i < K; i++) acc[i] = gl::zero<Fp3>();
#pragma unroll
      for (int m = 0; m < Mc / 2; m++) {
        Fp3 lo[K], hi[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          lo[i] = x[i][m];
          hi[i] = x[i][m + Mc / 2];
        }
        eval_pair<Fp3, K>(lo, hi, acc, w);
      }
      block_reduce<Fp3, K>(acc);
      if (tid < 64) {
        const Fp3 c = emit_round_regs3<K>(fs, acc, polys + (size_t)r * K * 3, chals + (size_t)r * 3);
        if (tid == 0) {
          chal_sh[0] = c.c0;
          chal_sh[1] = c.c1;
          chal_sh[2] = c.c2;
        }
      }
      __syncthreads();
      const Fp3 c = gl::load_words<Fp3>(chal_sh);
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int m = 0; m < Mc / 2; m++) x[i][m] = fold1<Fp3>(x[i][m], x[i][m + Mc / 2], c, w);
      Lc /= 2;
      r++;
    }
  }
  if (r < rounds) {
    // ---- hand-over: 256 elements per table to shared memory, then the
    // lane-split rounds: slot s = p K + k of the H = Lc / 2 pairs evaluates
    // point k of pair p (one product of K factors) and, after the challenge,
    // folds table k of pair p, so a round's dependent chain is one slot's
    // work instead of one pair's K-point evaluation and K folds (measured
    // product of 3 over F_p^3, nine rounds from 512 elements: 183 us with a
    // pair per lane of wave 0)
#pragma unroll
    for (int i = 0; i < K; i++) {
      tab[i][0][tid] = x[i][0].c0;
      tab[i][1][tid] = x[i][0].c1;
      tab[i][2][tid] = x[i][0].c2;
    }
    __shared__ uint64_t val[3][2 * T];  // slot values, H K <= 128 K <= 2 T
    __shared__ uint64_t psum[K][3];
    static_assert(K <= 4, "one wave per evaluation point");
    auto ld = [&](int i, int pos) { return Fp3{tab[i][0][pos], tab[i][1][pos], tab[i][2][pos]}; };
    auto st = [&](int i, int pos, Fp3 v) {
      tab[i][0][pos] = v.c0;
      tab[i][1][pos] = v.c1;
      tab[i][2][pos] = v.c2;
    };
    __syncthreads();
    for (; r < rounds; r++) {
      const int H = Lc / 2, S = H * K;
      for (int s = tid; s < S; s += T) {
        const int p = s / K, k = s % K;
        // point k: 0, infinity (the leading coefficient), then 2, 3, ...
        Fp3 prod = gl::zero<Fp3>();
#pragma unroll
        for (int i = 0; i < K; i++) {
          const Fp3 lo = ld(i, p), hi = ld(i, p + H), d = gl::sub(hi, lo);
          Fp3 xi = k == 0 ? lo : k == 1 ? d : gl::add(hi, d);
          for (int t = 3; t <= k; t++) xi = gl::add(xi, d);
          prod = i == 0 ? xi : gl::mul(prod, xi, w);
        }
        val[0][s] = prod.c0;
        val[1][s] = prod.c1;
        val[2][s] = prod.c2;
      }
      __syncthreads();
      if (tid < 64 * K) {
        // wave k sums point k over the pairs
        const int k = tid >> 6;
        Fp3 a = gl::zero<Fp3>();
        for (int pp = lane; pp < H; pp += 64) a = gl::add(a, Fp3{val[0][pp * K + k], val[1][pp * K + k], val[2][pp * K + k]});
#pragma unroll
        for (int mm = 32; mm >= 1; mm >>= 1) a = gl::add(a, shfl_xor(a, mm));
        if (lane == 0) {
          psum[k][0] = a.c0;
          psum[k][1] = a.c1;
          psum[k][2] = a.c2;
        }
      }
      __syncthreads();
      if (tid < 64) {
        Fp3 acc[K];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = Fp3{psum[k][0], psum[k][1], psum[k][2]};
        const Fp3 c = emit_round_regs3<K>(fs, acc, polys + (size_t)r * K * 3, chals + (size_t)r * 3);
        if (tid == 0) {
          chal_sh[0] = c.c0;
          chal_sh[1] = c.c1;
          chal_sh[2] = c.c2;
        }
      }
      __syncthreads();
      const gl::Fp3Pre cp = gl::prepare(gl::load_words<Fp3>(chal_sh), w);
      for (int s = tid; s < S; s += T) {
        const int p = s / K, i = s % K;
        const Fp3 lo = ld(i, p);
        st(i, p, gl::add(gl::mul(gl::sub(ld(i, p + H), lo), cp), lo));
      }
      __syncthreads();
      Lc = H;
    }
    // write back the tables (length Lc) and the final value
    for (int s = tid; s < Lc * K; s += T) work.t[s % K].st(s / K, ld(s % K, s / K));
    if (tid == 0) {
      Fp3 v = gl::zero<Fp3>();
      if (Lc == 1) {
        v = ld(0, 0);
#pragma unroll
        for (int i = 1; i < K; i++) v = gl::mul(v, ld(i, 0), w);
      }
      gl::store_words(v, fin);
    }
    if (tid < 64) fs.store(t);
    return;
  }
  // rounds ended with the tables still spread over the workgroup
  const int Mc = Lc / T;
#pragma unroll
  for (int i = 0; i < K; i++)
#pragma unroll
    for (int m = 0; m < M0; m++)
      if (m < Mc) work.t[i].st(tid + T * m, x[i][m]);
  if (tid == 0) gl::store_words(gl::zero<Fp3>(), fin);  // more than one element left
  if (tid < 64) fs.store(t);
}

}  // namespace effsc_hip

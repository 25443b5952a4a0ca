// Device-resident proof loop of the product-family provers (multilinear,
// inner product, product of k) and the two-round pass launchers.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp3;

namespace effsc_hip {
namespace pv {
// ---------------------------------------------------------- two-round passes

template <int K, int P>
int launch_pass_kp(effsc_hip_prover* h, size_t q, const uint64_t* cptr) {
  uint64_t* const* src = src_planes(h);
  size_t sst = src_stride(h);
  Views<Fp, K> in = views<Fp, K>(src, sst);
  Views<Fp, K> out = views<Fp, K>(h->work, h->work_stride);
  const size_t L = 4 * q;
  bool vec = (q % 2 == 0) && (sst % 2 == 0) && (h->work_stride % 2 == 0);
  const unsigned gmax = (unsigned)h->grid_max * kPassGridScale;
  const size_t units = vec ? q / 2 : q;
  unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((units + kBlock - 1) / kBlock, gmax));
  TRY(prof_begin(h));
  // fold passes over planes of 128 MiB and more stream their source and the
  // written level past the caches (nontemporal loads and stores): neither is
  // reused before it would be evicted; measured 2.75 against 2.85 ms for the
  // v = 28 inner product, no gain below that size
  const bool nt = vec && P > 0 && 8.0 * (double)(L << P) >= kNtMinBytes;
  // wave chunks (k_pass): C consecutive 128-unit steps per wave, C enough to
  // cover the table with gmax workgroups and at most 8 (read pass) or 4 (fold
  // passes); the workgroups are as many as the chunks fill
  size_t span = 0;
  if (vec) {
    const size_t steps = (q + 127) / 128, waves = (size_t)gmax * (kBlock / 64);
    const size_t C = std::min<size_t>(P == 0 ? 8 : 4, std::max<size_t>(1, (steps + waves - 1) / waves));
    span = C * 128;
    const size_t chunks = (q + span - 1) / span;
    g = (unsigned)std::max<size_t>(1, std::min<size_t>((chunks + kBlock / 64 - 1) / (kBlock / 64), gmax));
  }
  if (nt)
    hipLaunchKernelGGL((k_pass<K, P, 2, 3>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1],
                       cptr, q, h->red, span);
  else if (vec)
    hipLaunchKernelGGL((k_pass<K, P, 2>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1], cptr,
                       q, h->red, span);
  else
    hipLaunchKernelGGL((k_pass<K, P, 1>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1], cptr,
                       q, h->red, (size_t)0);
  HIPCHK(hipGetLastError());
  double bytes = 8.0 * K * ((double)(L << P) + (P > 0 ? (double)L : 0.0));
  // per unit: 4 level-L elements per table, each 0, 1 or 3 fold products
  // away from the stored level; IP accumulates 8 products
  double muls = (double)q * (4.0 * K * (P == 0 ? 0 : P == 1 ? 1 : 3) + (K == 2 ? 8 : 0));
  TRY(prof_end(h, 3, bytes, muls));
  h->last_grid = g;
  return EFFSC_HIP_OK;
}

// One pass over the current tables with p pending folds (from fpend, or from
// device memory at cptr); leaves the reduced pass values in red.result.
int launch_pass(effsc_hip_prover* h, int p, size_t q, const uint64_t* cptr) {
  auto go = [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    switch (p) {
      case 0: return launch_pass_kp<K, 0>(h, q, cptr);
      case 1: return launch_pass_kp<K, 1>(h, q, cptr);
      case 2: return launch_pass_kp<K, 2>(h, q, cptr);
    }
    return fail(EFFSC_HIP_EINVAL, "bad pending fold count");
  };
  if (h->K == 1) return go(std::integral_constant<int, 1>{});
  if (h->K == 2) return go(std::integral_constant<int, 2>{});
  return fail(EFFSC_HIP_EINVAL, "two-round pass needs K <= 2");
}

}  // namespace pv
}  // namespace effsc_hip

// ===========================================================================
// Device-resident proof loop (effsc_hip_sumcheck with no hook)
// ===========================================================================
namespace effsc_hip {
namespace pv {

bool device_loop_applies(const effsc_hip_prover* h, size_t num_rounds) {
  if (h->kind == K_EQ || h->kind == K_GKR || h->lsb_native || num_rounds == 0 || num_rounds > kMaxRounds) return false;
  if (!is_pow2(h->len)) return false;
  size_t logl = 0;
  while (((size_t)1 << logl) < h->len) logl++;
  return num_rounds <= logl;  // every round evaluates a table of length >= 2
}

int ensure_dev_bufs(effsc_hip_prover* h) {
  if (h->d_fs) return EFFSC_HIP_OK;
  uint64_t* blk = nullptr;
  TRY(alloc_dev(&blk, kProofWords));
  h->d_polys = blk;
  h->d_chals = blk + kMaxRounds * kMaxK * 3;
  h->d_fin = h->d_chals + kMaxRounds * 3;
  h->d_fs = reinterpret_cast<DevFs*>(h->d_fin + 4);
  TRY(alloc_dev(&h->d_cur, 8));  // two pending challenges of up to 3 words
  TRY(alloc_dev(&h->fuse_parts, (size_t)kFuseMaxGrid * 8 + (sizeof(DevFs) + 7) / 8));
  h->d_fs_alt = reinterpret_cast<DevFs*>(h->fuse_parts + (size_t)kFuseMaxGrid * 8);
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h->h_proof), kProofWords * 8, hipHostMallocDefault));
  h->h_proof_pinned = true;
  return EFFSC_HIP_OK;
}

// The proof block back to the host staging (polys, challenges, final value
// and transcript state in one copy).
int copy_proof_out(effsc_hip_prover* h) {
  TRY(fs_flush(h));
  HIPCHK(hipMemcpyAsync(h->h_proof, h->d_polys, kProofWords * 8, hipMemcpyDeviceToHost, h->stream));
  return EFFSC_HIP_OK;
}


// A base-field fold pass that can take the previous pass's transcript step
// into its own launch (k_pass_fs2): two pending folds, vector-aligned
// planes, a grid small enough to be resident at once and a previous pass of
// at most kFuseMaxIn workgroups (nin).
bool fuse_ok(const effsc_hip_prover* h, int p, size_t q, unsigned nin) {
  if (!h->fuse_parts || h->ext != 1 || h->K > 2 || p != 2 || nin > kFuseMaxIn || std::getenv("EFFSC_HIP_NO_FUSE"))
    return false;
  if (q < 2 || q % 2 || src_stride(h) % 2 || h->work_stride % 2) return false;
  return (q / 2 + kBlock - 1) / kBlock <= kFuseMaxGrid;
}

// k_pass_fs2 over level 4q -> q units: runs the transcript step of the pass
// whose values are in (rin, nin) as rounds r, r + 1, then this pass; its
// partials go to the other buffer, returned in rout / g.
int launch_pass_fs(effsc_hip_prover* h, size_t q, Reduce rin, unsigned nin, size_t r, Reduce* rout, unsigned* g) {
  uint64_t* const* src = src_planes(h);
  const size_t sst = src_stride(h);
  const unsigned grid = (unsigned)((q / 2 + kBlock - 1) / kBlock);
  const Reduce ro = rin.partials == h->fuse_parts ? h->red : Reduce{h->fuse_parts, h->red.result};
  const DevFs* tin = fs_in(h);
  DevFs* tout = tin == h->d_fs ? h->d_fs_alt : h->d_fs;
  auto go = [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    hipLaunchKernelGGL((k_pass_fs2<K>), dim3(grid), dim3(kBlock), 0, h->stream, views<Fp, K>(src, sst),
                       views<Fp, K>(h->work, h->work_stride), q, rin, nin, ro, tin, tout, h->d_polys + K * r,
                       h->d_chals + r, h->d_cur);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  };
  TRY(h->K == 1 ? go(std::integral_constant<int, 1>{}) : go(std::integral_constant<int, 2>{}));
  // the next transcript kernel starts from d_fs_alt (or it is copied to d_fs)
  if (tout == h->d_fs_alt) h->fs_pending = h->d_fs_alt;
  *rout = ro;
  *g = grid;
  h->last_grid = grid;
  return EFFSC_HIP_OK;
}

// Runs rounds [0, num_rounds), finalize and final_value on the device with
// the transcript state copied in and out of `t`.  Words out are canonical.
int prove_device(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, uint64_t* round_polys,
                 uint64_t* challenges, uint64_t* final_value) {
  TRY(ensure_dev_bufs(h));
  DevModeGuard guard(h);
  const int E = h->ext, K = h->K;
  // transcript state -> device
  DevFs* hs = reinterpret_cast<DevFs*>(h->h_proof + kMaxRounds * kMaxK * 3 + kMaxRounds * 3 + 4);
  std::memset(hs, 0, sizeof(DevFs));
  hs->kind = (uint32_t)t->kind;
  if (t->kind == 0) {
    for (int i = 0; i < 25; i++) hs->st[i] = t->fs.sp.st[i];
    hs->pos = t->fs.sp.pos;
  }
  hs->rng = t->rng;
  // a pinned staging is read by the first transcript kernel itself (fs_in);
  // otherwise (a caller's host buffer) the state is uploaded now
  h->fs_pending = hs;
  h->proof_by_tail = false;
  if (!h->h_proof_pinned) TRY(fs_flush(h));

  // host-side pending folds are applied first; from here on the pending
  // challenges live on the device (d_cur[0..p))
  TRY(materialize(h));
  h->have_next = false;
  size_t r = 0;
  int p = 0;
  bool tail_used = false;
  const bool ext2 = ext_twostep(h);
  if (h->twostep || ext2) {
    // two rounds per pass while the tables are large (base field: pass2.hpp;
    // extension fields: pass_ext.hip)
    // a base-field pass's transcript step waits for the next pass, which
    // takes it into its launch when it is small enough (k_pass_fs2)
    bool deferred = false;
    Reduce dred{};
    unsigned dg = 0;
    auto reduce_now = [&](Reduce red, unsigned g) -> int {
      if (K == 1)
        hipLaunchKernelGGL(k_reduce_fs2<1>, dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, fs_in(h), h->d_fs,
                           h->d_polys + r, h->d_chals + r, h->d_cur);
      else
        hipLaunchKernelGGL(k_reduce_fs2<2>, dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, fs_in(h), h->d_fs,
                           h->d_polys + 2 * r, h->d_chals + r, h->d_cur);
      HIPCHK(hipGetLastError());
      r += 2;
      return EFFSC_HIP_OK;
    };
    while (num_rounds >= r + (deferred ? 2 : 0) + 2) {
      size_t L = h->len >> p;
      if (L <= tail_max(h) || L < 4) break;
      Reduce red = h->red;
      unsigned g = 0;
      const bool from_pass0 = p == 0 && h->pass0_ready && !h->in_work && h->len == h->n;
      const bool plain = !from_pass0 && !(h->ooc && !h->in_work) && !ext2;
      if (deferred && !(plain && fuse_ok(h, p, L / 4, dg))) {
        TRY(reduce_now(dred, dg));
        deferred = false;
      }
      if (deferred) {
        TRY(launch_pass_fs(h, L / 4, dred, dg, r, &red, &g));
        r += 2;
        deferred = false;
      } else if (from_pass0) {
        // the pass over the input was accumulated while it streamed in
        red = Reduce{h->d_pass0, h->red.result};
        g = 1;
      } else if (h->ooc && !h->in_work) {
        // out-of-core input: the fold pass streams it from the host again
        if (p != 2) return fail(EFFSC_HIP_ESTATE, "out-of-core pass order");
        TRY(ooc_fold_pass(h, L / 4, &g));
        red = Reduce{h->ooc_parts, h->red.result};
      } else if (ext2) {
        TRY(launch_pass_ext(h, p, L / 4, p ? h->d_cur : nullptr));
        g = h->last_grid;
      } else {
        TRY(launch_pass(h, p, L / 4, p ? h->d_cur : nullptr));
        g = h->last_grid;
      }
      if (p > 0) {
        h->len = L;
        h->in_work = true;
      }
      if (ext2) {
        TRY(fs_flush(h));
        TRY(launch_reduce_fs2_ext(h, red, g, r));
        r += 2;
      } else if (plain) {
        deferred = true;
        dred = red;
        dg = g;
      } else {
        TRY(reduce_now(red, g));
      }
      p = 2;
    }
    if (deferred) TRY(reduce_now(dred, dg));
  } else {
    while (r < num_rounds) {
      size_t M = h->len;
      if (M <= tail_max(h)) break;
      int fi = src_field(h);
      uint64_t* const* src = src_planes(h);
      size_t sst = src_stride(h);
      int fr;  // field of the reduced values
      if (p == 1) {
        TRY(dispatch(fi, E, K, [&](auto op) { return decltype(op)::fused(h, src, sst, nullptr, M / 4, h->d_cur); }));
        h->len = M / 2;
        h->in_work = true;
        fr = E;
      } else {
        TRY(dispatch(fi, fi, K, [&](auto op) { return decltype(op)::eval(h, src, sst, M / 2, M / 2); }));
        fr = fi;
      }
      unsigned g = h->last_grid;
      TRY(fs_flush(h));
      TRY(dispatch<DevOps>(fr, fr, K, [&](auto op) { return decltype(op)::reduce_fs(h, g, r); }));
      r++;
      p = 1;
    }
  }
  const uint64_t* cur = h->d_cur;
  if (p == 2 && pipe_tail_ok(h, h->len / 4, num_rounds - r)) {
    // the pipelined tail applies both pending folds as it loads the tables
    const size_t L0 = h->len / 4, rounds_left = num_rounds - r;
    TRY(launch_tail_pipe(h, src_planes(h), src_stride(h), L0, 2, (int)rounds_left, r, h->d_cur));
    h->len = L0 >> rounds_left;
    h->in_work = true;
    tail_used = true;
    p = 0;
    r = num_rounds;
  } else if (p == 2) {
    // apply the older of two pending challenges with a multi-block fold
    size_t M = h->len;
    uint64_t* const* src = src_planes(h);
    size_t sst = src_stride(h);
    TRY(dispatch(src_field(h), E, K,
                 [&](auto op) { return decltype(op)::fold(h, src, sst, nullptr, M / 2, M / 2, h->d_cur); }));
    h->len = M / 2;
    h->in_work = true;
    p = 1;
    cur = h->d_cur + E;
  }
  size_t rounds_left = num_rounds - r;
  size_t M = h->len;
  const bool have = p == 1;
  if (tail_used) {
    // done by the pipelined tail
  } else if (rounds_left > 0 || M <= tail_max(h)) {
    // remaining rounds, finalize and final value in one workgroup
    int fi = src_field(h);
    bool from_input = !h->in_work;
    TRY(dispatch<DevOps>(fi, E, K, [&](auto op) {
      return decltype(op)::tail(h, src_planes(h), src_stride(h), from_input, M, (int)rounds_left, have, r, cur);
    }));
    // folds inside the tail: one per round that has a challenge, then the
    // final fold with the last challenge
    size_t folds = rounds_left == 0 ? 0 : (have ? rounds_left : rounds_left - 1);
    size_t Lt = M >> folds;
    if ((have || rounds_left > 0) && Lt >= 2) Lt /= 2;
    h->len = Lt;
    h->in_work = true;
    tail_used = true;
  } else if (have) {
    // partial run ending on a large table: finalize with the multi-block fold
    int fi = src_field(h);
    uint64_t* const* src = src_planes(h);
    size_t sst = src_stride(h);
    TRY(dispatch(fi, E, K, [&](auto op) { return decltype(op)::fold(h, src, sst, nullptr, M / 2, M / 2, cur); }));
    h->len = M / 2;
    h->in_work = true;
  }
  // proof and transcript state -> host
  uint64_t* hp = h->h_proof;
  uint64_t* hc = hp + kMaxRounds * kMaxK * 3;
  uint64_t* hf = hc + kMaxRounds * 3;
  if (h->proof_by_tail)
    h->proof_by_tail = false;  // the pipelined tail wrote the block into h_proof
  else
    TRY(copy_proof_out(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  TRY(prof_collect(h));
  std::copy(hp, hp + num_rounds * K * E, round_polys);
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


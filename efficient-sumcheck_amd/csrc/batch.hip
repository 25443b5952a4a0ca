// Batched proofs: many independent multilinear / inner-product sumcheck
// instances of uneven size advance together, two rounds per step.
//
// Instance i with v_i variables runs floor(v_i / 2) pass steps and one final
// step.  Step s (two-round passes, pass2.hpp):
//   k_bpass   one workgroup per chunk of one instance's pass units: step 0
//             reads the input once (round 0 values and round 1
//             coefficients); step s >= 1 folds the stored level twice on the
//             fly with the instance's two pending challenges, stores the
//             level-2s table and accumulates rounds 2s and 2s+1
//   k_bfs     one workgroup per instance: after a pass, sums its chunks'
//             partials and runs two transcript steps on the instance's device
//             sponge (emit_two_rounds); at an instance's final step it folds
//             the last (at most 8-element) table, runs the odd last round if
//             there is one, and writes the final value
// Levels after the input are materialised every other round, so the batch
// moves about two thirds of the bytes of one fold per round.  The static
// chunk tables are built once; the whole step sequence is captured into a
// hipGraph and replayed for every prove call.  Each instance's proof is the
// one runner::sumcheck (src/runner.rs:29-69) produces for that instance alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "effsc_hip.h"
#include "kernels.hpp"
#include "pass2.hpp"
#include "prove.hpp"

using namespace effsc_hip;
using gl::Fp;

namespace {

struct BInst {
  const uint64_t* in[2];  // input tables (borrowed, canonical), n = 2^v words
  uint64_t* work[2];      // level-2 (then deeper levels, in place) tables, n/4 words
  uint32_t v;             // rounds
  uint32_t first_round;   // offset of this instance in the round/challenge arrays
};

struct BChunk {
  uint32_t inst;
  uint32_t begin;  // first pass unit
  uint32_t count;
  uint32_t pad;
};

// Per-instance item of a step: kind 0 = after a pass (range of this step's
// partials), 1 = the instance's final step.
struct BItem {
  uint32_t inst, kind, part0, nparts;
};

inline uint32_t pass_steps(uint32_t v) { return v / 2; }

// Pass units of an instance with v variables at step s < pass_steps(v): the
// evaluated level has L = 2^(v - 2s) elements, L / 4 units.
inline size_t pass_units(uint32_t v, uint32_t s) { return (size_t)1 << (v - 2 * s - 2); }

template <int K>
__global__ __launch_bounds__(kBlock) void k_bpass(const BInst* insts, const BChunk* chunks, uint32_t s,
                                                  const uint64_t* cur, uint64_t* partials) {
  const BChunk ch = chunks[blockIdx.x];
  const BInst& in = insts[ch.inst];
  const size_t L = (size_t)1 << (in.v - 2 * s), q = L / 4;
  const uint64_t c1 = s ? cur[2 * ch.inst] : 0, c2 = s ? cur[2 * ch.inst + 1] : 0;
  const uint64_t c12 = gl::mul(c1, c2);
  // the stored level: the input for steps 0 and 1, else level 2s - 2
  const uint64_t* src[2] = {s <= 1 ? in.in[0] : in.work[0], s <= 1 ? in.in[1] : in.work[1]};
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) acc.a[i].zero();
  const size_t end = (size_t)ch.begin + ch.count;
  if (q >= 2) {
    // pairs of units with 16-byte loads (q and the chunk bounds are even)
    for (size_t u = ch.begin + 2 * threadIdx.x; u < end; u += 2 * blockDim.x) {
      uint64_t E0[K][4], E1[K][4];
#pragma unroll
      for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          // every stored level is read once per step and the written level
          // only in the next step, after the other instances' traffic: both
          // bypass the caches (measured 12.0 against 12.2 ms for 512 instances)
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
            __builtin_nontemporal_store(v2u64{E0[i][j], E1[i][j]}, reinterpret_cast<v2u64*>(in.work[i] + u + j * q));
      }
      pass_accumulate<K>(E0, acc);
      pass_accumulate<K>(E1, acc);
    }
  } else {
    for (size_t u = ch.begin + threadIdx.x; u < end; u += blockDim.x) {
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
          for (int j = 0; j < 4; j++) in.work[i][u + j * q] = E[i][j];
      }
      pass_accumulate<K>(E, acc);
    }
  }
  Fp raw[pass_raw<K>()], vals[pass_values<K>()];
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  finish<Fp, pass_values<K>()>(vals, Reduce{partials, nullptr});
}

// Final step of an instance (the whole workgroup calls it).  The stored level
// (the input at step 0 or 1, else the work table) has M = 2^(v - 2s + 2)
// elements (M = 2^v at s = 0) and, for s >= 1, two pending challenges.
// Folding them leaves one element per table (v even) or two (v odd: that
// pair is round v - 1, then the final fold with its challenge).
template <int K>
__device__ __forceinline__ void batch_final(const BInst& in, uint32_t inst, uint32_t s, const uint64_t* cur, DevFs* fs,
                                            uint64_t* polys, uint64_t* chals, uint64_t* finals) {
  __shared__ uint64_t chal_sh[3];
  const uint32_t v = in.v;
  if (v == 0) {
    if (threadIdx.x == 0) finals[inst] = K == 2 ? gl::mul(in.in[0][0], in.in[1][0]) : in.in[0][0];
    return;
  }
  const size_t M = s == 0 ? ((size_t)1 << v) : ((size_t)1 << (v - 2 * s + 2));
  const uint64_t* src[2] = {s <= 1 ? in.in[0] : in.work[0], s <= 1 ? in.in[1] : in.work[1]};
  const uint64_t x = s ? cur[2 * inst] : 0, y = s ? cur[2 * inst + 1] : 0;
  const size_t n2 = s ? M / 4 : M;  // 1 or 2 elements per table after the pending folds
  uint64_t t[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int i = 0; i < K; i++)
    for (size_t e = 0; e < n2; e++) {
      if (s == 0) {
        t[i][e] = src[i][e];
      } else {
        // level M -> M/2 with x (pairs (k, k + M/2)), then M/2 -> M/4 with y
        const size_t Q = M / 4;
        const uint64_t a = gl::add(src[i][e], gl::mul(x, gl::sub(src[i][e + 2 * Q], src[i][e])));
        const uint64_t b = gl::add(src[i][e + Q], gl::mul(x, gl::sub(src[i][e + 3 * Q], src[i][e + Q])));
        t[i][e] = gl::add(a, gl::mul(y, gl::sub(b, a)));
      }
    }
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
    const size_t r = in.first_round + v - 1;
    if (threadIdx.x < 64) emit_round<Fp, K>(vals, 1, fs + inst, polys + r * K, chals + r, chal_sh);
    __syncthreads();
    const uint64_t z = chal_sh[0];
#pragma unroll
    for (int i = 0; i < K; i++) f[i] = gl::add(t[i][0], gl::mul(z, gl::sub(t[i][1], t[i][0])));
  }
  if (threadIdx.x == 0) finals[inst] = K == 2 ? gl::mul(f[0], f[1]) : f[0];
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_bfs(const BInst* insts, const BItem* items, uint32_t s,
                                                const uint64_t* partials, DevFs* fs, uint64_t* polys, uint64_t* chals,
                                                uint64_t* cur, uint64_t* finals) {
  const BItem it = items[blockIdx.x];
  const BInst& in = insts[it.inst];
  if (it.kind == 1) {
    batch_final<K>(in, it.inst, s, cur, fs, polys, chals, finals);
    return;
  }
  constexpr int NV = pass_values<K>();
  // the sponge load overlaps the partial sums
  LaneFs lf;
  if (threadIdx.x < 64) lf.load(fs + it.inst);
  Fp sum[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) sum[i] = gl::zero<Fp>();
  sum_partials<Fp, NV>(partials + (size_t)it.part0 * NV, it.nparts, sum);
  block_reduce<Fp, NV>(sum);
  if (threadIdx.x < 64) {
    const size_t r = in.first_round + 2 * s;
    emit_two_rounds<K>(sum, lf, fs + it.inst, polys + r * K, chals + r, cur + 2 * it.inst);
  }
}

}  // namespace

struct effsc_hip_batch {
  int dev = 0, K = 2;
  size_t count = 0, total_rounds = 0;
  uint32_t steps = 0;
  std::vector<uint32_t> v, first_round;
  hipStream_t stream = nullptr;
  BInst* d_insts = nullptr;
  uint64_t* d_work = nullptr;
  BChunk* d_chunks = nullptr;
  BItem* d_items = nullptr;
  std::vector<size_t> chunk_off, nchunks, item_off, nitems;  // per step
  uint64_t* d_partials = nullptr;
  DevFs* d_fs = nullptr;
  uint64_t* d_polys = nullptr;
  uint64_t* d_chals = nullptr;
  uint64_t* d_cur = nullptr;  // two pending challenges per instance
  uint64_t* d_finals = nullptr;
  DevFs* h_fs = nullptr;  // pinned staging
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  size_t chunk_units = 0;
};

namespace {

int balloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) return set_error(EFFSC_HIP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return EFFSC_HIP_OK;
}

void bdestroy(effsc_hip_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : {(void*)b->d_insts, (void*)b->d_work, (void*)b->d_chunks, (void*)b->d_items, (void*)b->d_partials,
                  (void*)b->d_fs, (void*)b->d_polys, (void*)b->d_chals, (void*)b->d_cur, (void*)b->d_finals})
    if (p) (void)hipFree(p);
  if (b->h_fs) (void)hipHostFree(b->h_fs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Enqueues every step of the batch on the stream.
int enqueue_steps(effsc_hip_batch* b) {
  for (uint32_t s = 0; s < b->steps; s++) {
    if (b->nchunks[s]) {
      const BChunk* ch = b->d_chunks + b->chunk_off[s];
      if (b->K == 1)
        hipLaunchKernelGGL(k_bpass<1>, dim3((unsigned)b->nchunks[s]), dim3(kBlock), 0, b->stream, b->d_insts, ch, s,
                           b->d_cur, b->d_partials);
      else
        hipLaunchKernelGGL(k_bpass<2>, dim3((unsigned)b->nchunks[s]), dim3(kBlock), 0, b->stream, b->d_insts, ch, s,
                           b->d_cur, b->d_partials);
      HIPCHK(hipGetLastError());
    }
    if (b->nitems[s]) {
      const BItem* it = b->d_items + b->item_off[s];
      if (b->K == 1)
        hipLaunchKernelGGL(k_bfs<1>, dim3((unsigned)b->nitems[s]), dim3(kBlock), 0, b->stream, b->d_insts, it, s,
                           b->d_partials, b->d_fs, b->d_polys, b->d_chals, b->d_cur, b->d_finals);
      else
        hipLaunchKernelGGL(k_bfs<2>, dim3((unsigned)b->nitems[s]), dim3(kBlock), 0, b->stream, b->d_insts, it, s,
                           b->d_partials, b->d_fs, b->d_polys, b->d_chals, b->d_cur, b->d_finals);
      HIPCHK(hipGetLastError());
    }
  }
  return EFFSC_HIP_OK;
}

// (Re)builds the per-step chunk and item tables for b->chunk_units units per
// workgroup and, when graphs are on, captures the step sequence.
int build_plan(effsc_hip_batch* b) {
  const size_t count = b->count;
  if (b->stream) HIPCHK(hipStreamSynchronize(b->stream));  // no launch may still use the old tables
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (void** p : {(void**)&b->d_chunks, (void**)&b->d_items, (void**)&b->d_partials}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  b->chunk_off.clear();
  b->nchunks.clear();
  b->item_off.clear();
  b->nitems.clear();
  // even chunk sizes: a pair of units never straddles two chunks
  const size_t cu = std::max<size_t>(2, b->chunk_units & ~(size_t)1);
  std::vector<BChunk> chunks;
  std::vector<BItem> items;
  for (uint32_t s = 0; s < b->steps; s++) {
    b->chunk_off.push_back(chunks.size());
    b->item_off.push_back(items.size());
    for (size_t i = 0; i < count; i++) {
      const uint32_t np = pass_steps(b->v[i]);
      if (s < np) {
        const size_t u = pass_units(b->v[i], s);
        // partials are indexed by the chunk's position in this step's launch
        const uint32_t first = (uint32_t)(chunks.size() - b->chunk_off[s]);
        for (size_t x = 0; x < u; x += cu)
          chunks.push_back({(uint32_t)i, (uint32_t)x, (uint32_t)std::min(cu, u - x), 0});
        items.push_back({(uint32_t)i, 0u, first, (uint32_t)(chunks.size() - b->chunk_off[s] - first)});
      } else if (s == np) {
        items.push_back({(uint32_t)i, 1u, 0u, 0u});
      }
    }
    b->nchunks.push_back(chunks.size() - b->chunk_off.back());
    b->nitems.push_back(items.size() - b->item_off.back());
  }
  size_t max_chunks = 0;
  for (size_t x : b->nchunks) max_chunks = std::max(max_chunks, x);
  TRY(balloc((void**)&b->d_chunks, chunks.size() * sizeof(BChunk)));
  TRY(balloc((void**)&b->d_items, items.size() * sizeof(BItem)));
  TRY(balloc((void**)&b->d_partials, max_chunks * pass_values<2>() * 8));
  hipError_t e = hipSuccess;
  if (!chunks.empty()) e = hipMemcpy(b->d_chunks, chunks.data(), chunks.size() * sizeof(BChunk), hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
  if (!b->use_graph) return EFFSC_HIP_OK;
  // capture the step sequence once
  e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, std::string("begin capture: ") + hipGetErrorString(e));
  int rr = enqueue_steps(b);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  b->graph = g;
  if (rr != EFFSC_HIP_OK) return rr;
  if (e2 != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e2));
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
  return EFFSC_HIP_OK;
}

// Enqueues one whole batched proof from the staged transcripts.
int run_once(effsc_hip_batch* b) {
  HIPCHK(hipMemcpyAsync(b->d_fs, b->h_fs, b->count * sizeof(DevFs), hipMemcpyHostToDevice, b->stream));
  if (b->use_graph)
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
  else
    TRY(enqueue_steps(b));
  return EFFSC_HIP_OK;
}

}  // namespace

extern "C" {

int effsc_hip_batch_autotune(effsc_hip_batch* b, const size_t* candidates, size_t ncand, int reps, size_t* chosen,
                             double* ms_out) {
  if (!b || !candidates || ncand == 0 || reps < 1) return set_error(EFFSC_HIP_EINVAL, "bad argument");
  HIPCHK(hipSetDevice(b->dev));
  // any transcript state gives the same work: time proofs with a zero sponge
  std::memset(b->h_fs, 0, b->count * sizeof(DevFs));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  size_t best = b->chunk_units;
  double best_ms = 1e300;
  int rc = EFFSC_HIP_OK;
  for (size_t c = 0; c < ncand && rc == EFFSC_HIP_OK; c++) {
    if (candidates[c] < 1) {
      rc = set_error(EFFSC_HIP_EINVAL, "chunk size must be >= 1");
      break;
    }
    b->chunk_units = candidates[c];
    rc = build_plan(b);
    if (rc == EFFSC_HIP_OK) rc = run_once(b);  // warm-up
    float ms = 0;
    if (rc == EFFSC_HIP_OK && hipEventRecord(e0, b->stream) != hipSuccess) rc = set_error(EFFSC_HIP_EDEVICE, "event");
    for (int r = 0; r < reps && rc == EFFSC_HIP_OK; r++) rc = run_once(b);
    if (rc == EFFSC_HIP_OK && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                               hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = set_error(EFFSC_HIP_EDEVICE, "autotune timing");
    if (rc != EFFSC_HIP_OK) break;
    double per = ms / reps;
    if (ms_out) ms_out[c] = per;
    if (per < best_ms) {
      best_ms = per;
      best = candidates[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  TRY(rc);
  b->chunk_units = best;
  TRY(build_plan(b));
  if (chosen) *chosen = best;
  return EFFSC_HIP_OK;
}

int effsc_hip_batch_new(int device, int kind, size_t count, const size_t* lens, const uint64_t* const* a,
                        const uint64_t* const* b_tabs, unsigned flags, effsc_hip_batch** out) {
  if (!out) return set_error(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  if (kind != 1 && kind != 2) return set_error(EFFSC_HIP_EINVAL, "batch kind must be 1 (multilinear) or 2 (inner product)");
  if (count == 0 || !lens || !a || (kind == 2 && !b_tabs)) return set_error(EFFSC_HIP_EINVAL, "null argument");
  HIPCHK(hipSetDevice(device));
  auto* b = new effsc_hip_batch();
  b->dev = device;
  b->K = kind;
  b->count = count;
  b->use_graph = !(flags & 0x2u);
  b->chunk_units = (size_t)kBlock * 8;
  if (const char* e = std::getenv("EFFSC_HIP_BATCH_CHUNK")) {
    long long x = std::atoll(e);
    if (x >= 1) b->chunk_units = (size_t)x;
  }
  int rc = EFFSC_HIP_OK;
  std::vector<BInst> hi(count);
  // work tables hold level 2 (n/4 elements) and deeper levels in place; an
  // even number of words each keeps every table 16-byte aligned
  auto work_len = [](size_t n) { return (std::max<size_t>(n / 4, 1) + 1) & ~(size_t)1; };
  size_t work_words = 0;
  for (size_t i = 0; i < count && rc == EFFSC_HIP_OK; i++) {
    size_t n = lens[i];
    if (n == 0 || (n & (n - 1)) || n > ((size_t)1 << 31)) rc = set_error(EFFSC_HIP_EINVAL, "batch tables must have power-of-two length <= 2^31");
    uint32_t v = 0;
    while (((size_t)1 << v) < n) v++;
    b->v.push_back(v);
    b->first_round.push_back((uint32_t)b->total_rounds);
    b->total_rounds += v;
    b->steps = std::max(b->steps, pass_steps(v) + 1);
    work_words += kind * work_len(n);
  }
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_work, work_words * 8);
  size_t off = 0;
  for (size_t i = 0; i < count && rc == EFFSC_HIP_OK; i++) {
    size_t n = lens[i];
    for (int t = 0; t < 2; t++) {
      hi[i].in[t] = t < kind ? (t == 0 ? a[i] : b_tabs[i]) : nullptr;
      if (t < kind) {
        hi[i].work[t] = b->d_work + off;
        off += work_len(n);
      } else {
        hi[i].work[t] = nullptr;
      }
      if (t < kind && !hi[i].in[t]) rc = set_error(EFFSC_HIP_EINVAL, "null table");
      if (t < kind && hi[i].in[t] && reinterpret_cast<uintptr_t>(hi[i].in[t]) % 16)
        rc = set_error(EFFSC_HIP_EINVAL, "batch tables must be 16-byte aligned");
    }
    hi[i].v = b->v[i];
    hi[i].first_round = b->first_round[i];
  }
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_insts, count * sizeof(BInst));
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_fs, count * sizeof(DevFs));
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_polys, std::max<size_t>(b->total_rounds, 1) * kind * 8);
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_chals, std::max<size_t>(b->total_rounds, 1) * 8);
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_cur, count * 2 * 8);
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_finals, count * 8);
  if (rc == EFFSC_HIP_OK) {
    hipError_t e = hipHostMalloc((void**)&b->h_fs, count * sizeof(DevFs), hipHostMallocDefault);
    if (e != hipSuccess) rc = set_error(EFFSC_HIP_ENOMEM, hipGetErrorString(e));
  }
  if (rc == EFFSC_HIP_OK) {
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpy(b->d_insts, hi.data(), count * sizeof(BInst), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = set_error(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
  }
  if (rc == EFFSC_HIP_OK) rc = build_plan(b);
  if (rc != EFFSC_HIP_OK) {
    bdestroy(b);
    return rc;
  }
  *out = b;
  return EFFSC_HIP_OK;
}

int effsc_hip_batch_rounds(const effsc_hip_batch* b, size_t instance, size_t* v, size_t* first_round) {
  if (!b || instance >= b->count) return set_error(EFFSC_HIP_EINVAL, "bad instance");
  if (v) *v = b->v[instance];
  if (first_round) *first_round = b->first_round[instance];
  return EFFSC_HIP_OK;
}

int effsc_hip_batch_prove(effsc_hip_batch* b, effsc_hip_transcript* const* transcripts, uint64_t* round_polys,
                          uint64_t* challenges, uint64_t* finals) {
  if (!b || !transcripts || !finals || (b->total_rounds && (!round_polys || !challenges)))
    return set_error(EFFSC_HIP_EINVAL, "null argument");
  HIPCHK(hipSetDevice(b->dev));
  for (size_t i = 0; i < b->count; i++) {
    const effsc_hip_transcript* t = transcripts[i];
    if (!t) return set_error(EFFSC_HIP_EINVAL, "null transcript");
    DevFs& d = b->h_fs[i];
    std::memset(&d, 0, sizeof(DevFs));
    d.kind = (uint32_t)t->kind;
    if (t->kind == 0) {
      for (int k = 0; k < 25; k++) d.st[k] = t->fs.sp.st[k];
      d.pos = t->fs.sp.pos;
    }
    d.rng = t->rng;
  }
  TRY(run_once(b));
  if (b->total_rounds) {
    HIPCHK(hipMemcpyAsync(round_polys, b->d_polys, b->total_rounds * b->K * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(challenges, b->d_chals, b->total_rounds * 8, hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipMemcpyAsync(finals, b->d_finals, b->count * 8, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(b->h_fs, b->d_fs, b->count * sizeof(DevFs), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->count; i++) {
    effsc_hip_transcript* t = transcripts[i];
    if (t->kind == 0) {
      for (int k = 0; k < 25; k++) t->fs.sp.st[k] = b->h_fs[i].st[k];
      t->fs.sp.pos = b->h_fs[i].pos;
    }
    t->rng = b->h_fs[i].rng;
  }
  return EFFSC_HIP_OK;
}

void effsc_hip_batch_free(effsc_hip_batch* b) { bdestroy(b); }

}  // extern "C"

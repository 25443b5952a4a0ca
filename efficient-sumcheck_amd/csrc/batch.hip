// Batched proofs: many independent multilinear / inner-product sumcheck
// instances of uneven size advance round by round together.
//
// Per round two launches serve every active instance:
//   k_batch_round  one workgroup per chunk of one instance's work (static
//                  chunk tables built at setup): round 0 evaluates, round
//                  j >= 1 folds with the instance's challenge and evaluates
//                  (the fused quadruple pass of kernels.hpp), round j = v_i
//                  applies the last fold and writes the final value.
//   k_batch_fs     one workgroup per active instance: sums its chunks'
//                  partials, writes the round polynomial and runs the
//                  Fiat-Shamir step on its own device transcript.
// So the latency of the transcript step (one Keccak permutation per round) is
// paid once per round for the whole batch instead of once per instance.  The
// whole round sequence is captured once into a hipGraph and replayed for
// every prove call; every kernel reads its challenges from device memory.
// Each instance's proof is the one runner::sumcheck (src/runner.rs:29-69)
// produces for that instance alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "effsc_hip.h"
#include "kernels.hpp"
#include "prove.hpp"

using namespace effsc_hip;
using gl::Fp;

namespace {

struct BInst {
  const uint64_t* in[2];  // input tables (borrowed, canonical)
  uint64_t* work[2];      // work tables, n/2 elements
  uint32_t v;             // rounds
  uint32_t first_round;   // offset of this instance in the round/challenge arrays
};

struct BChunk {
  uint32_t inst;
  uint32_t begin;  // first unit (pair in round 0, quadruple otherwise)
  uint32_t count;
  uint32_t pad;
};

// Units an instance processes in round j (j <= v): round 0 evaluates n/2
// pairs, round j in [1, v) folds and evaluates L/4 quadruples (L = 2^(v-j+1)),
// round v applies the last fold (one unit).
inline size_t units(uint32_t v, uint32_t j) {
  if (j == v) return 1;
  if (j == 0) return (size_t)1 << (v - 1);
  return (size_t)1 << (v - j - 1);
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_batch_round(const BInst* insts, const BChunk* chunks, uint32_t j,
                                                        const uint64_t* cur, uint64_t* finals, uint64_t* partials) {
  const BChunk ch = chunks[blockIdx.x];
  const BInst in = insts[ch.inst];
  Fp acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = gl::zero<Fp>();
  const size_t n = (size_t)1 << in.v;
  if (j == in.v) {
    // finalize (src/runner.rs:57-60) and final value
    if (threadIdx.x == 0) {
      uint64_t val;
      if (in.v == 0) {
        val = in.in[0][0];
        if (K == 2) val = gl::mul(val, in.in[1][0]);
      } else {
        const uint64_t* src[2] = {in.v == 1 ? in.in[0] : in.work[0], in.v == 1 ? in.in[1] : in.work[1]};
        uint64_t c = cur[ch.inst];
        uint64_t f[2];
#pragma unroll
        for (int i = 0; i < K; i++) f[i] = gl::add(src[i][0], gl::mul(c, gl::sub(src[i][1], src[i][0])));
        val = K == 2 ? gl::mul(f[0], f[1]) : f[0];
#pragma unroll
        for (int i = 0; i < K; i++) in.work[i][0] = f[i];
      }
      finals[ch.inst] = val;
    }
  } else if (j == 0) {
    const size_t half = n / 2;
    for (size_t k = ch.begin + threadIdx.x; k < (size_t)ch.begin + ch.count; k += blockDim.x) {
      Fp lo[K], hi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        lo[i] = {in.in[i][k]};
        hi[i] = {in.in[i][k + half]};
      }
      eval_pair<Fp, K>(lo, hi, acc, 0);
    }
  } else {
    const size_t q = n >> (j + 1);  // L / 4 with L = 2^(v - j + 1)
    const Fp c = {cur[ch.inst]};
    for (size_t k = ch.begin + threadIdx.x; k < (size_t)ch.begin + ch.count; k += blockDim.x) {
      Fp lo[K], hi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        const uint64_t* s = j == 1 ? in.in[i] : in.work[i];
        Fp x0 = {s[k]}, x1 = {s[k + q]}, x2 = {s[k + 2 * q]}, x3 = {s[k + 3 * q]};
        lo[i] = fold1<Fp>(x0, x2, c, 0);
        hi[i] = fold1<Fp>(x1, x3, c, 0);
        in.work[i][k] = lo[i].v;
        in.work[i][k + q] = hi[i].v;
      }
      eval_pair<Fp, K>(lo, hi, acc, 0);
    }
  }
  block_reduce<Fp, K>(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < K; i++) partials[(size_t)blockIdx.x * K + i] = acc[i].v;
}

// One workgroup per active instance of round j: its chunks are
// [chunk_begin[w], chunk_begin[w] + chunk_count[w]) of this round's table.
template <int K>
__global__ __launch_bounds__(kBlock) void k_batch_fs(const BInst* insts, const uint32_t* act_inst,
                                                     const uint32_t* chunk_begin, const uint32_t* chunk_count,
                                                     uint32_t j, const uint64_t* partials, DevFs* fs,
                                                     uint64_t* polys, uint64_t* chals, uint64_t* cur) {
  const uint32_t inst = act_inst[blockIdx.x];
  const uint32_t b0 = chunk_begin[blockIdx.x], nb = chunk_count[blockIdx.x];
  Fp s[K];
#pragma unroll
  for (int i = 0; i < K; i++) s[i] = gl::zero<Fp>();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < K; i++) s[i] = gl::add(s[i], Fp{partials[(size_t)(b0 + b) * K + i]});
  block_reduce<Fp, K>(s);
  if (threadIdx.x < 64) {
    const uint32_t r = insts[inst].first_round + j;
    emit_round<Fp, K>(s, 1, fs + inst, polys + (size_t)r * K, chals + r, cur + inst);
  }
}

}  // namespace

struct effsc_hip_batch {
  int dev = 0, K = 2;
  size_t count = 0, total_rounds = 0;
  uint32_t max_v = 0;
  std::vector<uint32_t> v, first_round;
  hipStream_t stream = nullptr;
  BInst* d_insts = nullptr;
  uint64_t* d_work = nullptr;
  BChunk* d_chunks = nullptr;
  uint32_t* d_act = nullptr;      // per round: active instances
  uint32_t* d_cbegin = nullptr;   // per round: chunk range per active instance
  uint32_t* d_ccount = nullptr;
  std::vector<size_t> chunk_off, nchunks, act_off, nact;  // per round offsets
  uint64_t* d_partials = nullptr;
  DevFs* d_fs = nullptr;
  uint64_t* d_polys = nullptr;
  uint64_t* d_chals = nullptr;
  uint64_t* d_cur = nullptr;
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
  for (void* p : {(void*)b->d_insts, (void*)b->d_work, (void*)b->d_chunks, (void*)b->d_act, (void*)b->d_cbegin,
                  (void*)b->d_ccount, (void*)b->d_partials, (void*)b->d_fs, (void*)b->d_polys, (void*)b->d_chals,
                  (void*)b->d_cur, (void*)b->d_finals})
    if (p) (void)hipFree(p);
  if (b->h_fs) (void)hipHostFree(b->h_fs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Enqueues every round of the batch on the stream.
int enqueue_rounds(effsc_hip_batch* b) {
  for (uint32_t j = 0; j <= b->max_v; j++) {
    if (b->nchunks[j] == 0) continue;
    const BChunk* ch = b->d_chunks + b->chunk_off[j];
    uint64_t* part = b->d_partials;
    if (b->K == 1)
      hipLaunchKernelGGL(k_batch_round<1>, dim3((unsigned)b->nchunks[j]), dim3(kBlock), 0, b->stream, b->d_insts, ch,
                         j, b->d_cur, b->d_finals, part);
    else
      hipLaunchKernelGGL(k_batch_round<2>, dim3((unsigned)b->nchunks[j]), dim3(kBlock), 0, b->stream, b->d_insts, ch,
                         j, b->d_cur, b->d_finals, part);
    HIPCHK(hipGetLastError());
    if (b->nact[j] == 0) continue;
    const uint32_t* act = b->d_act + b->act_off[j];
    const uint32_t* cb = b->d_cbegin + b->act_off[j];
    const uint32_t* cc = b->d_ccount + b->act_off[j];
    if (b->K == 1)
      hipLaunchKernelGGL(k_batch_fs<1>, dim3((unsigned)b->nact[j]), dim3(kBlock), 0, b->stream, b->d_insts, act, cb,
                         cc, j, part, b->d_fs, b->d_polys, b->d_chals, b->d_cur);
    else
      hipLaunchKernelGGL(k_batch_fs<2>, dim3((unsigned)b->nact[j]), dim3(kBlock), 0, b->stream, b->d_insts, act, cb,
                         cc, j, part, b->d_fs, b->d_polys, b->d_chals, b->d_cur);
    HIPCHK(hipGetLastError());
  }
  return EFFSC_HIP_OK;
}

// (Re)builds the per-round chunk tables for b->chunk_units units per
// workgroup and, when graphs are on, captures the round sequence.
int build_plan(effsc_hip_batch* b) {
  const size_t count = b->count;
  if (b->stream) HIPCHK(hipStreamSynchronize(b->stream));  // no launch may still use the old tables
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (void** p : {(void**)&b->d_chunks, (void**)&b->d_act, (void**)&b->d_cbegin, (void**)&b->d_ccount,
                   (void**)&b->d_partials}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  b->chunk_off.clear();
  b->act_off.clear();
  b->nchunks.clear();
  b->nact.clear();
  std::vector<BChunk> chunks;
  std::vector<uint32_t> act, cb, cc;
  for (uint32_t j = 0; j <= b->max_v; j++) {
    b->chunk_off.push_back(chunks.size());
    b->act_off.push_back(act.size());
    for (size_t i = 0; i < count; i++) {
      if (j > b->v[i]) continue;
      size_t u = units(b->v[i], j);
      // partials of round j are indexed by the chunk's position in that
      // round's launch (blockIdx.x), i.e. relative to chunk_off[j]
      uint32_t first = (uint32_t)(chunks.size() - b->chunk_off[j]);
      for (size_t s = 0; s < u; s += b->chunk_units)
        chunks.push_back({(uint32_t)i, (uint32_t)s, (uint32_t)std::min(b->chunk_units, u - s), 0});
      if (j < b->v[i]) {
        act.push_back((uint32_t)i);
        cb.push_back(first);
        cc.push_back((uint32_t)(chunks.size() - b->chunk_off[j] - first));
      }
    }
    b->nchunks.push_back(chunks.size() - b->chunk_off.back());
    b->nact.push_back(act.size() - b->act_off.back());
  }
  size_t max_chunks = 0;
  for (size_t x : b->nchunks) max_chunks = std::max(max_chunks, x);
  TRY(balloc((void**)&b->d_chunks, chunks.size() * sizeof(BChunk)));
  TRY(balloc((void**)&b->d_act, act.size() * 4));
  TRY(balloc((void**)&b->d_cbegin, act.size() * 4));
  TRY(balloc((void**)&b->d_ccount, act.size() * 4));
  TRY(balloc((void**)&b->d_partials, max_chunks * b->K * 8));
  hipError_t e = hipSuccess;
  if (!chunks.empty()) e = hipMemcpy(b->d_chunks, chunks.data(), chunks.size() * sizeof(BChunk), hipMemcpyHostToDevice);
  if (e == hipSuccess && !act.empty()) e = hipMemcpy(b->d_act, act.data(), act.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && !act.empty()) e = hipMemcpy(b->d_cbegin, cb.data(), cb.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && !act.empty()) e = hipMemcpy(b->d_ccount, cc.data(), cc.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, hipGetErrorString(e));
  if (!b->use_graph) return EFFSC_HIP_OK;
  // capture the round sequence once
  e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return set_error(EFFSC_HIP_EDEVICE, std::string("begin capture: ") + hipGetErrorString(e));
  int rr = enqueue_rounds(b);
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
    TRY(enqueue_rounds(b));
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
  size_t work_words = 0;
  for (size_t i = 0; i < count && rc == EFFSC_HIP_OK; i++) {
    size_t n = lens[i];
    if (n == 0 || (n & (n - 1)) || n > ((size_t)1 << 31)) rc = set_error(EFFSC_HIP_EINVAL, "batch tables must have power-of-two length <= 2^31");
    uint32_t v = 0;
    while (((size_t)1 << v) < n) v++;
    b->v.push_back(v);
    b->first_round.push_back((uint32_t)b->total_rounds);
    b->total_rounds += v;
    b->max_v = std::max(b->max_v, v);
    work_words += kind * std::max<size_t>(n / 2, 1);
  }
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_work, work_words * 8);
  size_t off = 0;
  for (size_t i = 0; i < count && rc == EFFSC_HIP_OK; i++) {
    size_t n = lens[i];
    for (int t = 0; t < 2; t++) {
      hi[i].in[t] = t < kind ? (t == 0 ? a[i] : b_tabs[i]) : nullptr;
      if (t < kind) {
        hi[i].work[t] = b->d_work + off;
        off += std::max<size_t>(n / 2, 1);
      } else {
        hi[i].work[t] = nullptr;
      }
      if (t < kind && !hi[i].in[t]) rc = set_error(EFFSC_HIP_EINVAL, "null table");
    }
    hi[i].v = b->v[i];
    hi[i].first_round = b->first_round[i];
  }
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_insts, count * sizeof(BInst));
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_fs, count * sizeof(DevFs));
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_polys, std::max<size_t>(b->total_rounds, 1) * kind * 8);
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_chals, std::max<size_t>(b->total_rounds, 1) * 8);
  if (rc == EFFSC_HIP_OK) rc = balloc((void**)&b->d_cur, count * 8);
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

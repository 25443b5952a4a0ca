// C ABI of the gfx950 sumcheck prover: prover handles, round dispatch, the
// host protocol loop (runner, verifier, transcripts) and utilities.
// Declarations and reference citations: include/effsc_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "effsc_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "pass2.hpp"
#include "prove.hpp"
#include "shake.hpp"

using namespace effsc_hip;
using gl::Fp;
using gl::Fp3;

namespace effsc_hip {
thread_local std::string g_err;
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* last_error() { return g_err.c_str(); }

// Import / export / synthetic-input kernels (used by this file only).
// AoS caller words -> canonical SoA planes (optionally from Montgomery form).
__global__ __launch_bounds__(kBlock) void k_import(uint64_t* dst, size_t stride, const uint64_t* src, size_t n,
                                                   int d, int mont) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < n; i += step) {
    for (int c = 0; c < d; c++) {
      uint64_t x = src[i * d + c];
      x = mont ? gl::mul(x, gl::MONT_TO_CANON) : gl::canon(x);
      dst[c * stride + i] = x;
    }
  }
}

// Canonical SoA planes -> AoS caller words (optionally to Montgomery form).
__global__ __launch_bounds__(kBlock) void k_export(uint64_t* dst, const uint64_t* src, size_t stride, size_t n,
                                                   int d, int mont) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < n; i += step) {
    for (int c = 0; c < d; c++) {
      uint64_t x = src[c * stride + i];
      dst[i * d + c] = mont ? gl::mul(x, gl::CANON_TO_MONT) : x;
    }
  }
}

// Bit-reversal permutation of planes of n = 2^v words: dst[i] = src[rev_v(i)]
// (reorder_vec_msb, src/streams/memory/core.rs:20-47).  The index splits
// into (a, m, b) with t = 5 top and bottom bits; a workgroup owns one middle
// value m and moves the 32 x 32 tile of (a, b) through LDS, so both the reads
// (b contiguous) and the writes (rev(a) contiguous) are coalesced:
// dst[(rho, rev(m), kappa)] = src[(rev(kappa), m, rev(rho))].
constexpr int kRevT = 5;
__device__ __forceinline__ uint32_t rev5(uint32_t x) { return __brev(x) >> 27; }

__global__ __launch_bounds__(kBlock) void k_bitrev_tiled(uint64_t* dst, size_t dstr, const uint64_t* src, size_t sstr,
                                                          int v) {
  __shared__ uint64_t tile[32][33];
  src += (size_t)blockIdx.y * sstr;
  dst += (size_t)blockIdx.y * dstr;
  const int mbits = v - 2 * kRevT;
  const uint64_t m = blockIdx.x;
  const uint64_t mrev = mbits ? (__brevll(m) >> (64 - mbits)) : 0;
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (uint32_t a = ty; a < 32; a += kBlock / 32)
    tile[a][tx] = src[((uint64_t)a << (v - kRevT)) | (m << kRevT) | tx];
  __syncthreads();
  for (uint32_t rho = ty; rho < 32; rho += kBlock / 32)
    dst[((uint64_t)rho << (v - kRevT)) | (mrev << kRevT) | tx] = tile[rev5(tx)][rev5(rho)];
}

__global__ __launch_bounds__(kBlock) void k_bitrev_small(uint64_t* dst, size_t dstr, const uint64_t* src, size_t sstr,
                                                          int v) {
  src += (size_t)blockIdx.y * sstr;
  dst += (size_t)blockIdx.y * dstr;
  const uint64_t n = (uint64_t)1 << v;
  for (uint64_t i = gtid(); i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[v ? (__brevll(i) >> (64 - v)) : 0];
}

__global__ __launch_bounds__(kBlock) void k_fill_uniform(uint64_t* dst, size_t n, uint64_t key, uint64_t first) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = gtid(); i < n; i += step) dst[i] = uniform_word(key, first + i);
}

}  // namespace effsc_hip

namespace {

int fail(int code, const std::string& msg) { return effsc_hip::set_error(code, msg); }

enum Kind { K_ML = 1, K_IP = 2, K_PRODUCT = 3, K_EQ = 4, K_GKR = 5 };

size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}
bool is_pow2(size_t n) { return n && !(n & (n - 1)); }
size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

int g_grid_max = 0;  // 0: derived from the device's CU count

int grid_max_for(int dev) {
  if (g_grid_max > 0) return g_grid_max;
  if (const char* e = std::getenv("EFFSC_HIP_GRID_MAX")) {
    int v = std::atoi(e);
    if (v > 0) return v;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  return cus * 8;
}

}  // namespace

// ---------------------------------------------------------------------------
// Prover handle
// ---------------------------------------------------------------------------
struct effsc_hip_prover {
  int kind = 0;
  int ext = 1, in_ext = 1, K = 1, deg = 1, repr = 0, dev = 0;
  uint64_t w = EFFSC_HIP_W_FROM_RAW7;
  size_t n = 0, len = 0;
  bool in_work = false, own_in = true;
  uint64_t* in[kMaxK] = {};
  size_t in_stride = 0;
  uint64_t* work[kMaxK] = {};
  size_t work_stride = 0;
  hipStream_t stream = nullptr;
  Reduce red = {};
  uint64_t* h_res = nullptr;  // pinned, kMaxK * 3 words
  int grid_max = 2048;
  // EqFactored state (src/provers/eq_factored.rs:67-83)
  size_t v = 0, v_l = 0, rounds_elapsed = 0;
  uint64_t* eq_l = nullptr;  // ext planes, stride eq_l_stride
  uint64_t* eq_r = nullptr;
  size_t eq_l_stride = 0, eq_r_stride = 0, eq_l_len = 0, eq_r_len = 0;
  uint64_t* eq_point = nullptr;  // canonical point, v * ext words (rebuilds the eq tables on reset)
  uint64_t* eq_full = nullptr;   // EQ = eq_l x eq_r at the register-tail level (at most 2^14 words)
  // LSB (pair-split) provers hold their tables bit-reversed and run the MSB
  // path; post-state reads reverse back (src/streams/memory/core.rs:20-47)
  bool lsb = false;
  // Provers built from streams: the first two-round pass (round 0 values and
  // round 1 coefficients over the input) was accumulated during the upload
  // and its reduced values live in d_pass0 (the input is never written, so
  // they stay valid across resets)
  uint64_t* d_pass0 = nullptr;
  bool pass0_ready = false;
  // Out-of-core stream provers: the input stays in the host streams; the two
  // passes that read it (rounds 0-1 at construction, rounds 2-3 in the proof)
  // move it through pinned and device staging chunk by chunk, and only the
  // level-2 tables (n/4 elements each) are ever resident
  bool ooc = false;
  const effsc_hip_stream* ooc_src[kMaxK] = {};
  uint64_t* ooc_parts = nullptr;  // partials of the streamed fold pass
  uint64_t* gkr_b = nullptr;  // GKR factored witness: two ping-pong planesets of n/2 elements (3 planes)
  size_t ooc_nparts = 0;
  // Device-resident proof loop (effsc_hip_sumcheck without a hook)
  bool dev_mode = false;       // Ops skip the plain reduction; grid in last_grid
  unsigned last_grid = 0;
  // one device block, polys | chals | fin | fs, the layout of h_proof
  uint64_t* d_polys = nullptr;  // kMaxRounds * kMaxK * 3 words (block start)
  uint64_t* d_chals = nullptr;  // kMaxRounds * 3 words
  uint64_t* d_fin = nullptr;    // 4 words
  DevFs* d_fs = nullptr;
  uint64_t* d_cur = nullptr;    // 3 words: current challenge (own allocation)
  uint64_t* h_proof = nullptr;  // pinned staging: polys | chals | fin | fs
  // Kernel timing (effsc_hip_profile): HIP events around each launch on the
  // prover's stream; slot 0 fused fold+evaluate, 1 evaluate, 2 fold,
  // 3 two-round pass.
  bool prof_on = false;
  std::vector<hipEvent_t> ev;             // pairs (2i, 2i+1)
  struct Pend {
    int slot;
    double bytes, muls;
  };
  std::vector<Pend> pend;  // recorded event pairs
  uint64_t prof_launches[4] = {0, 0, 0, 0};
  double prof_ms[4] = {0, 0, 0, 0};
  double prof_bytes[4] = {0, 0, 0, 0};
  double prof_muls[4] = {0, 0, 0, 0};  // Goldilocks products (ext3 product = 8)
  // Two-round passes (base-field ML / IP): the logical table is the
  // materialised one (len) folded by the npend challenges in fpend; when
  // have_next, the next round's values are the polynomials in next_coef
  // (pass layout of pass2.hpp) evaluated at the next challenge.
  bool twostep = false;
  int npend = 0;
  uint64_t fpend[2] = {0, 0};
  bool have_next = false;
  uint64_t next_coef[kMaxK + kMaxK * (kMaxK + 1)] = {};
};

namespace {

constexpr size_t kMaxRounds = 64;

constexpr int kPassGridScale = 4;

int alloc_dev(uint64_t** p, size_t words) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(words, 1) * 8);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(EFFSC_HIP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return EFFSC_HIP_OK;
}

// Brackets a kernel launch with events when profiling; the elapsed time is
// collected at the next stream synchronisation (prof_collect).
int prof_begin(effsc_hip_prover* h) {
  if (!h->prof_on) return EFFSC_HIP_OK;
  size_t i = h->pend.size();
  while (h->ev.size() < 2 * (i + 1)) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    h->ev.push_back(e);
  }
  HIPCHK(hipEventRecord(h->ev[2 * i], h->stream));
  return EFFSC_HIP_OK;
}
int prof_end(effsc_hip_prover* h, int slot, double bytes, double muls) {
  if (!h->prof_on) return EFFSC_HIP_OK;
  size_t i = h->pend.size();
  HIPCHK(hipEventRecord(h->ev[2 * i + 1], h->stream));
  h->pend.push_back({slot, bytes, muls});
  return EFFSC_HIP_OK;
}

// Algorithmic Goldilocks products per field product and per fold product
// c * (hi - lo) with c in FO, the difference in FI.  An F_p^3 product counts
// as Karatsuba's 8 full products (6 + 2 by w) whatever the kernels execute:
// they run 9 or 11 unreduced 64x64 products with one reduction per
// coefficient (gl::mul(Fp3, Fp3Pre)), so the count is a fixed unit of work.
template <class F>
constexpr double mul_cost() {
  return F::D == 1 ? 1.0 : 8.0;
}
template <class FI, class FO>
constexpr double fold_cost() {
  return FI::D == FO::D ? mul_cost<FO>() : 3.0;
}
// products of one EvalsInfty pair evaluation: K values of K - 1 products
constexpr double eval_products(int K) { return K == 1 ? 0.0 : (double)K * (K - 1); }
int prof_collect(effsc_hip_prover* h) {
  if (!h->prof_on) return EFFSC_HIP_OK;
  for (size_t i = 0; i < h->pend.size(); i++) {
    HIPCHK(hipEventSynchronize(h->ev[2 * i + 1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->ev[2 * i], h->ev[2 * i + 1]));
    int s = h->pend[i].slot;
    h->prof_launches[s]++;
    h->prof_ms[s] += ms;
    h->prof_bytes[s] += h->pend[i].bytes;
    h->prof_muls[s] += h->pend[i].muls;
  }
  h->pend.clear();
  return EFFSC_HIP_OK;
}

unsigned blocks_for(const effsc_hip_prover* h, size_t units) {
  size_t b = (units + kBlock - 1) / kBlock;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(b, (size_t)h->grid_max));
}

template <class F, int K>
Views<F, K> views(uint64_t* const* planes, size_t stride) {
  Views<F, K> v;
  for (int i = 0; i < K; i++) v.t[i] = View<F>{planes[i], stride};
  return v;
}

template <class F>
F from_words(const uint64_t* w) {
  return gl::load_words<F>(w);
}

// -------------------------------------------------------------- launchers
template <class F, int K>
int reduce(effsc_hip_prover* h, unsigned nblocks) {
  h->last_grid = nblocks;
  if (h->dev_mode) return EFFSC_HIP_OK;  // the caller launches k_reduce_fs
  hipLaunchKernelGGL((k_reduce<F, K>), dim3(1), dim3(1024), 0, h->stream, h->red, nblocks);
  HIPCHK(hipGetLastError());
  return EFFSC_HIP_OK;
}

// planes of at least this many bytes are streamed past the caches by the
// passes that read or write them once (nontemporal loads and stores)
constexpr double kNtMinBytes = 134217728.0;

template <class FI, class FO, int K>
struct Ops {
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
    if constexpr (FO::D == 3 && K >= 3)
      vec = vec && q >= ((size_t)1 << 18) && !std::getenv("EFFSC_HIP_FSEQ_SCALAR");
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
    } else if constexpr (FO::D == 3 && K >= 3) {
      if (vec)
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 2>), dim3(blocks_for(h, q / 2)), dim3(kBlock), 0, h->stream, in,
                           out, c, cptr, q, h->w, h->red);
      else
        hipLaunchKernelGGL((k_fused_seq<FI, FO, K, 1>), dim3(blocks_for(h, q)), dim3(kBlock), 0, h->stream, in, out,
                           c, cptr, q, h->w, h->red);
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
    return reduce<FO, K>(h, vec ? blocks_for(h, q / 2) : blocks_for(h, q));
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

// Tail threshold: once the tables have at most this many elements, one
// workgroup runs every remaining round (k_tail).
size_t tail_max(const effsc_hip_prover* h) {
  if (const char* e = std::getenv("EFFSC_HIP_TAIL")) {
    long long x = std::atoll(e);
    if (x >= 2) return (size_t)x;
  }
  // one workgroup is one CU: hand it only what a CU finishes faster than a
  // multi-workgroup round plus its reduction (about 20-30 us).  A base-field
  // pair costs a few products, an F_p^3 product-of-K pair about 12 K.
  if (h->ext == 3) return h->K >= 3 ? (size_t)1 << 10 : (size_t)1 << 12;
  return (size_t)1 << 14;
}

template <class FI, class F, int K>
struct DevOps {
  static int reduce_fs(effsc_hip_prover* h, unsigned g, size_t round) {
    hipLaunchKernelGGL((k_reduce_fs<FI, K>), dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, h->ext,
                       h->d_polys + round * K * h->ext, h->d_chals + round * h->ext, h->d_cur);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }
  static int tail(effsc_hip_prover* h, uint64_t* const* src, size_t sst, bool from_input, size_t L, int rounds,
                  bool have, size_t first_round, const uint64_t* cur) {
    Views<FI, K> in = views<FI, K>(src, sst);
    Views<F, K> work = views<F, K>(h->work, h->work_stride);
    if constexpr (std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value && K <= 2) {
      // base-field tables of 2^11 .. 2^14 elements (after the pending fold):
      // the register-resident tail
      const size_t L1 = have ? L / 2 : L;
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

template <template <class, class, int> class O, class FI, class FO, class Fn>
int with_k(int K, Fn&& fn) {
  switch (K) {
    case 1: return fn(O<FI, FO, 1>{});
    case 2: return fn(O<FI, FO, 2>{});
    case 3: return fn(O<FI, FO, 3>{});
    case 4: return fn(O<FI, FO, 4>{});
  }
  return fail(EFFSC_HIP_EINVAL, "unsupported table count");
}

// Calls fn(O<FI, FO, K>{}) for the runtime field degrees and table count.
template <template <class, class, int> class O = Ops, class Fn>
int dispatch(int fi, int fo, int K, Fn&& fn) {
  if (fi == 1 && fo == 1) return with_k<O, Fp, Fp>(K, fn);
  if (fi == 1 && fo == 3) return with_k<O, Fp, Fp3>(K, fn);
  if (fi == 3 && fo == 3) return with_k<O, Fp3, Fp3>(K, fn);
  return fail(EFFSC_HIP_EINVAL, "unsupported field combination");
}

// ------------------------------------------------------------ round logic
int src_field(const effsc_hip_prover* h) { return h->in_work ? h->ext : h->in_ext; }
uint64_t* const* src_planes(const effsc_hip_prover* h) { return h->in_work ? h->work : h->in; }
size_t src_stride(const effsc_hip_prover* h) { return h->in_work ? h->work_stride : h->in_stride; }

// Half-split geometry of the current length: ML/IP pad implicitly to the next
// power of two (src/inner_product_sumcheck.rs:165-172); the coefficient
// prover pairs (k, k + n/2) (src/provers/coefficient.rs:131-140).
void geometry(const effsc_hip_prover* h, size_t L, size_t* half, size_t* paired) {
  if (h->kind == K_PRODUCT) {
    *half = L / 2;
    *paired = *half;
  } else {
    *half = next_pow2(L) >> 1;
    *paired = L - *half;
  }
}

// Reads element `idx` of every table (current state), lifted to the prover's
// field, as canonical words: out[i * ext + c].
int read_elements(effsc_hip_prover* h, size_t idx, uint64_t* out) {
  int fs = src_field(h);
  uint64_t* const* planes = src_planes(h);
  size_t stride = src_stride(h);
  for (int i = 0; i < h->K; i++) {
    for (int c = 0; c < h->ext; c++) {
      uint64_t x = 0;
      if (c < fs) HIPCHK(hipMemcpyAsync(&h->h_res[0], planes[i] + c * stride + idx, 8, hipMemcpyDeviceToHost, h->stream));
      if (c < fs) {
        HIPCHK(hipStreamSynchronize(h->stream));
        x = h->h_res[0];
      }
      out[i * h->ext + c] = x;
    }
  }
  return EFFSC_HIP_OK;
}

// Copies `count` reduced values (fi words each on the device) into out_vals,
// lifted to the prover's field (ext words each).
int fetch_result(effsc_hip_prover* h, int fi, uint64_t* out_vals, int count) {
  HIPCHK(hipMemcpyAsync(h->h_res, h->red.result, (size_t)count * fi * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  TRY(prof_collect(h));
  for (int i = 0; i < count; i++)
    for (int c = 0; c < h->ext; c++) out_vals[i * h->ext + c] = c < fi ? h->h_res[i * fi + c] : 0;
  return EFFSC_HIP_OK;
}

int pass_value_count(const effsc_hip_prover* h) { return h->K + h->K * (h->K + 1); }

// Reduced pass values (pass2.hpp layout) -> host.
int fetch_pass(effsc_hip_prover* h, uint64_t* vals) {
  int n = pass_value_count(h);
  HIPCHK(hipMemcpyAsync(h->h_res, h->red.result, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  TRY(prof_collect(h));
  std::copy(h->h_res, h->h_res + n, vals);
  return EFFSC_HIP_OK;
}

// Fold the current tables with challenge c (ext words).  If `fuse` and the
// length allows, also evaluates the next round polynomial into red.result
// and sets *evaluated.
int fold_step(effsc_hip_prover* h, const uint64_t* c, bool fuse, bool* evaluated) {
  *evaluated = false;
  size_t L = h->len;
  if (L <= 1) return EFFSC_HIP_OK;
  size_t half, paired;
  geometry(h, L, &half, &paired);
  int fi = src_field(h);
  uint64_t* const* src = src_planes(h);
  size_t sst = src_stride(h);
  if (fuse && is_pow2(L) && L >= 4) {
    TRY(dispatch(fi, h->ext, h->K, [&](auto op) { return decltype(op)::fused(h, src, sst, c, L / 4); }));
    *evaluated = true;
  } else {
    TRY(dispatch(fi, h->ext, h->K, [&](auto op) { return decltype(op)::fold(h, src, sst, c, half, paired); }));
  }
  h->len = half;
  h->in_work = true;
  return EFFSC_HIP_OK;
}

// Round polynomial values (K values, ext words each) of the current state.
int evaluate(effsc_hip_prover* h, bool already, uint64_t* vals) {
  size_t L = h->len;
  const int E = h->ext;
  std::fill(vals, vals + (size_t)h->K * E, 0);
  if (already) return fetch_result(h, h->ext, vals, h->K);
  if (h->kind != K_PRODUCT && L <= 1) {
    if (L == 0) return EFFSC_HIP_OK;
    // n == 1: ML -> [v0]; IP -> [a0 b0, 0] (src/provers/inner_product.rs:74-81)
    uint64_t el[kMaxK * 3];
    TRY(read_elements(h, 0, el));
    if (h->kind == K_ML) {
      std::copy(el, el + E, vals);
    } else if (E == 1) {
      vals[0] = gl::mul(el[0], el[1]);
    } else {
      Fp3 p = gl::mul(from_words<Fp3>(el), from_words<Fp3>(el + 3), h->w);
      gl::store_words(p, vals);
    }
    return EFFSC_HIP_OK;
  }
  size_t half, paired;
  geometry(h, L, &half, &paired);
  if (half == 0) return EFFSC_HIP_OK;
  int fi = src_field(h);
  uint64_t* const* src = src_planes(h);
  size_t sst = src_stride(h);
  TRY(dispatch(fi, fi, h->K, [&](auto op) { return decltype(op)::eval(h, src, sst, half, paired); }));
  return fetch_result(h, fi, vals, h->K);
}

// EvalsInfty wire from the K product values (h(0), h(inf), h(2), ...): the
// kernels already produce exactly that order; ML keeps one value.
void to_wire(const effsc_hip_prover* h, const uint64_t* vals, uint64_t* out) {
  size_t nw = (size_t)std::max(h->deg, 1);
  std::copy(vals, vals + nw * h->ext, out);
}

int eq_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out);
int ooc_fold_pass(effsc_hip_prover* h, size_t q, unsigned* nblocks);
int eq_finalize(effsc_hip_prover* h, const uint64_t* chal);
int eq_final_value(effsc_hip_prover* h, uint64_t* out);
int gkr_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out);
int gkr_final_value(effsc_hip_prover* h, uint64_t* out);

// ---------------------------------------------------------- two-round passes

template <int K, int P>
int launch_pass_kp(effsc_hip_prover* h, size_t q, const uint64_t* cptr) {
  uint64_t* const* src = src_planes(h);
  size_t sst = src_stride(h);
  Views<Fp, K> in = views<Fp, K>(src, sst);
  Views<Fp, K> out = views<Fp, K>(h->work, h->work_stride);
  const size_t L = 4 * q;
  bool vec = (q % 2 == 0) && (sst % 2 == 0) && (h->work_stride % 2 == 0);
  // a pass streams the whole table once: four workgroups per CU-slot of
  // grid_max keep more loads in flight (measured 5.3 against 5.05 TB/s for
  // the v = 28 inner-product read pass)
  const size_t units = vec ? q / 2 : q;
  unsigned g = (unsigned)std::max<size_t>(
      1, std::min<size_t>((units + kBlock - 1) / kBlock, (size_t)h->grid_max * kPassGridScale));
  TRY(prof_begin(h));
  // fold passes over planes of 128 MiB and more stream their source and the
  // written level past the caches (nontemporal loads and stores): neither is
  // reused before it would be evicted; measured 2.75 against 2.85 ms for the
  // v = 28 inner product, no gain below that size
  const bool nt = vec && P > 0 && 8.0 * (double)(L << P) >= kNtMinBytes;
  // the read pass (P = 0) gives each workgroup one contiguous chunk of units;
  // the fold passes keep the grid-stride order.  Measured for v = 28: the
  // chunked read pass took the proof from 2.52-2.62 to 2.38 ms, chunked fold
  // passes were slower (2.66 ms)
  size_t chunk = 0;
  unsigned gpb = 1;
  // groups of 8..2048 workgroups per chunk for the read pass, or 2..128 for
  // the fold passes, measured within noise of these two choices
  const unsigned G = P == 0 ? g : 1;
  if (vec && G > 1) {
    // G groups of gpb workgroups; each group strides over its own chunk
    const unsigned groups = std::min<unsigned>((unsigned)G, g);
    gpb = std::max<unsigned>(1, g / groups);
    const size_t per_step = (size_t)kBlock * 2 * gpb;
    chunk = std::max<size_t>(1, (q + per_step * groups - 1) / (per_step * groups)) * per_step;
    g = (unsigned)((q + chunk - 1) / chunk) * gpb;
  }
  if (nt)
    hipLaunchKernelGGL((k_pass<K, P, 2, 3>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1],
                       cptr, q, h->red, chunk, gpb);
  else if (vec)
    hipLaunchKernelGGL((k_pass<K, P, 2>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1], cptr,
                       q, h->red, chunk, gpb);
  else
    hipLaunchKernelGGL((k_pass<K, P, 1>), dim3(g), dim3(kBlock), 0, h->stream, in, out, h->fpend[0], h->fpend[1], cptr,
                       q, h->red);
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

// Applies the pending folds to the materialised tables.
int materialize(effsc_hip_prover* h) {
  for (int i = 0; i < h->npend; i++) {
    bool ev = false;
    uint64_t c = h->fpend[i];
    TRY(fold_step(h, &c, false, &ev));
  }
  h->npend = 0;
  return EFFSC_HIP_OK;
}

size_t logical_len(const effsc_hip_prover* h) { return h->len >> h->npend; }

int round_twostep(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  const int K = h->K;
  if (chal && h->have_next) {
    // round s+1 from the coefficients of the previous pass
    for (int i = 0; i < K; i++)
      out[i] = K == 1 ? pass_eval_next<1>(h->next_coef, i, chal[0]) : pass_eval_next<2>(h->next_coef, i, chal[0]);
    h->fpend[h->npend++] = chal[0];
    h->have_next = false;
    return EFFSC_HIP_OK;
  }
  h->have_next = false;
  if (chal) {
    if (h->npend == 2) TRY(materialize(h));
    h->fpend[h->npend++] = chal[0];
  }
  size_t M = h->len;
  int p = h->npend;
  if (is_pow2(M) && (M >> p) >= 4) {
    size_t L = M >> p;
    TRY(launch_pass(h, p, L / 4, nullptr));
    TRY((K == 1 ? reduce<Fp, pass_values<1>()>(h, h->last_grid) : reduce<Fp, pass_values<2>()>(h, h->last_grid)));
    uint64_t vals[kMaxK + kMaxK * (kMaxK + 1)];
    TRY(fetch_pass(h, vals));
    if (p > 0) {
      h->len = L;
      h->in_work = true;
      h->npend = 0;
    }
    std::copy(vals, vals + pass_value_count(h), h->next_coef);
    h->have_next = true;
    std::copy(vals, vals + K, out);
    return EFFSC_HIP_OK;
  }
  // small or non-power-of-two tables: single-round path
  TRY(materialize(h));
  uint64_t vals[kMaxK * 3];
  TRY(evaluate(h, false, vals));
  to_wire(h, vals, out);
  return EFFSC_HIP_OK;
}

int round_impl(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  if (h->kind == K_EQ) return eq_round(h, chal, out);
  if (h->kind == K_GKR) return gkr_round(h, chal, out);
  if (h->twostep) return round_twostep(h, chal, out);
  bool evaluated = false;
  if (chal) TRY(fold_step(h, chal, true, &evaluated));
  uint64_t vals[kMaxK * 3];
  TRY(evaluate(h, evaluated, vals));
  to_wire(h, vals, out);
  return EFFSC_HIP_OK;
}

int finalize_impl(effsc_hip_prover* h, const uint64_t* chal) {
  if (h->kind == K_EQ) return eq_finalize(h, chal);
  h->have_next = false;
  TRY(materialize(h));
  bool evaluated = false;
  return fold_step(h, chal, false, &evaluated);
}

int final_value_impl(effsc_hip_prover* h, uint64_t* out) {
  if (h->kind == K_EQ) return eq_final_value(h, out);
  if (h->kind == K_GKR) return gkr_final_value(h, out);
  TRY(materialize(h));
  const int E = h->ext;
  std::fill(out, out + E, 0);
  if (h->len != 1) return EFFSC_HIP_OK;
  uint64_t el[kMaxK * 3];
  TRY(read_elements(h, 0, el));
  // ML: evals[0]; IP: a0 b0; product: prod of singletons
  // (src/provers/multilinear.rs:73-79, inner_product.rs:120-126,
  //  coefficient.rs:300-329 with pairs (x_i, 0): h(0) + h(1) = prod x_i)
  if (E == 1) {
    uint64_t acc = el[0];
    for (int i = 1; i < h->K; i++) acc = gl::mul(acc, el[i]);
    out[0] = acc;
  } else {
    Fp3 acc = from_words<Fp3>(el);
    for (int i = 1; i < h->K; i++) acc = gl::mul(acc, from_words<Fp3>(el + 3 * i), h->w);
    gl::store_words(acc, out);
  }
  return EFFSC_HIP_OK;
}

// ----------------------------------------------------- word conversions
inline uint64_t gl_brev64(uint64_t x) {
  uint64_t r = 0;
  for (int b = 0; b < 64; b++) r |= ((x >> b) & 1) << (63 - b);
  return r;
}

void words_in(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(src[i], gl::MONT_TO_CANON) : gl::canon(src[i]);
}
void words_out(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(src[i], gl::CANON_TO_MONT) : src[i];
}

// ---------------------------------------------------------- construction
int check_cfg(const effsc_hip_config* cfg) {
  if (!cfg) return fail(EFFSC_HIP_EINVAL, "null config");
  if (cfg->ext != 1 && cfg->ext != 3) return fail(EFFSC_HIP_EINVAL, "ext must be 1 or 3");
  if (cfg->input_ext != 1 && cfg->input_ext != cfg->ext)
    return fail(EFFSC_HIP_EINVAL, "input_ext must be 1 or equal to ext");
  if (cfg->repr != EFFSC_HIP_CANONICAL && cfg->repr != EFFSC_HIP_MONT64) return fail(EFFSC_HIP_EINVAL, "bad repr");
  if ((cfg->flags & EFFSC_HIP_BORROW_INPUT) && cfg->repr != EFFSC_HIP_CANONICAL)
    return fail(EFFSC_HIP_EINVAL, "borrowed input must be canonical");
  return EFFSC_HIP_OK;
}

void destroy(effsc_hip_prover* h) {
  if (!h) return;
  (void)hipSetDevice(h->dev);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int i = 0; i < kMaxK; i++) {
    if (h->own_in && h->in[i]) (void)hipFree(h->in[i]);
    if (h->work[i]) (void)hipFree(h->work[i]);
  }
  if (h->red.partials) (void)hipFree(h->red.partials);
  if (h->red.result) (void)hipFree(h->red.result);
  if (h->eq_l) (void)hipFree(h->eq_l);
  if (h->eq_r) (void)hipFree(h->eq_r);
  if (h->eq_point) (void)hipFree(h->eq_point);
  if (h->eq_full) (void)hipFree(h->eq_full);
  if (h->d_pass0) (void)hipFree(h->d_pass0);
  if (h->ooc_parts) (void)hipFree(h->ooc_parts);
  if (h->gkr_b) (void)hipFree(h->gkr_b);
  if (h->h_res) (void)hipHostFree(h->h_res);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  for (uint64_t* p : {h->d_polys, h->d_cur}) if (p) (void)hipFree(p);  // d_chals, d_fin, d_fs: inside d_polys
  if (h->h_proof) (void)hipHostFree(h->h_proof);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

// Allocates stream, reduction scratch and work tables (K tables, ext planes of
// `cap` elements).
int setup_common(effsc_hip_prover* h, size_t cap) {
  HIPCHK(hipSetDevice(h->dev));
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->grid_max = grid_max_for(h->dev);
  // the two-round passes launch up to kPassGridScale * grid_max workgroups
  TRY(alloc_dev(&h->red.partials, (size_t)h->grid_max * kPassGridScale * kMaxK * 3));
  TRY(alloc_dev(&h->red.result, kMaxK * 3));
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h->h_res), kMaxK * 3 * 8, hipHostMallocDefault));
  h->work_stride = round_up(std::max<size_t>(cap, 1), 64);
  for (int i = 0; i < h->K; i++) TRY(alloc_dev(&h->work[i], h->work_stride * h->ext));
  return EFFSC_HIP_OK;
}

// Imports K input tables of n elements (in_ext words each, AoS, cfg repr)
// into canonical device planes.
int import_tables(effsc_hip_prover* h, const effsc_hip_config* cfg, const uint64_t* const* tables, size_t n) {
  if (cfg->flags & EFFSC_HIP_BORROW_INPUT) {
    h->own_in = false;
    h->in_stride = n;
    for (int i = 0; i < h->K; i++) {
      if (!tables[i] && n) return fail(EFFSC_HIP_EINVAL, "null table");
      if (reinterpret_cast<uintptr_t>(tables[i]) % 16) return fail(EFFSC_HIP_EINVAL, "borrowed table not 16-byte aligned");
      h->in[i] = const_cast<uint64_t*>(tables[i]);
    }
    return EFFSC_HIP_OK;
  }
  h->own_in = true;
  h->in_stride = round_up(std::max<size_t>(n, 1), 64);
  const int d = h->in_ext;
  uint64_t* staging = nullptr;
  if (d > 1 && n) TRY(alloc_dev(&staging, n * d));
  int rc = EFFSC_HIP_OK;
  for (int i = 0; i < h->K && rc == EFFSC_HIP_OK; i++) {
    rc = alloc_dev(&h->in[i], h->in_stride * d);
    if (rc != EFFSC_HIP_OK || n == 0) continue;
    if (!tables[i]) {
      rc = fail(EFFSC_HIP_EINVAL, "null table");
      break;
    }
    uint64_t* dst = d > 1 ? staging : h->in[i];
    hipError_t e = hipMemcpyAsync(dst, tables[i], n * d * 8, hipMemcpyDefault, h->stream);
    if (e != hipSuccess) {
      rc = fail(EFFSC_HIP_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
      break;
    }
    unsigned g = blocks_for(h, n);
    hipLaunchKernelGGL(k_import, dim3(g), dim3(kBlock), 0, h->stream, h->in[i], h->in_stride, dst, n, d,
                       h->repr == EFFSC_HIP_MONT64 ? 1 : 0);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, std::string("import: ") + hipGetErrorString(e));
  }
  if (staging) (void)hipFree(staging);
  return rc;
}

int new_product_like(const effsc_hip_config* cfg, int kind, int K, int deg, const uint64_t* const* tables, size_t n,
                     effsc_hip_prover** out) {
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (K < 1 || K > kMaxK) return fail(EFFSC_HIP_EINVAL, "table count must be in 1..4");
  auto* h = new effsc_hip_prover();
  h->kind = kind;
  h->ext = cfg->ext;
  h->in_ext = cfg->input_ext;
  h->K = K;
  h->deg = deg;
  h->repr = cfg->repr;
  h->dev = cfg->device;
  h->w = cfg->ext == 3 ? gl::canon(cfg->ext_w) : 0;
  h->n = n;
  h->len = n;
  h->twostep = (kind == K_ML || kind == K_IP) && h->ext == 1 && h->in_ext == 1 && !std::getenv("EFFSC_HIP_ONE_ROUND");
  int rc = setup_common(h, next_pow2(std::max<size_t>(n, 1)) >> 1);
  if (rc == EFFSC_HIP_OK) rc = import_tables(h, cfg, tables, n);
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  *out = h;
  return EFFSC_HIP_OK;
}

}  // namespace

// ===========================================================================
// EqFactoredProver: sum_x eq(w, x) p(x) with eq split into eq(w_L, .) and
// eq(w_R, .) (src/provers/eq_factored.rs:67-289).  Tables in the handle:
// p -> in[0]/work[0] (K = 1), eq_l, eq_r.
// ===========================================================================
namespace {

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
  if (h->ext == 1) {
    hipLaunchKernelGGL(k_eq_left<Fp>, dim3(g), dim3(kBlock), 0, h->stream, View<Fp>{src[0], sst}, p_half,
                       View<Fp>{h->eq_l, h->eq_l_stride}, eql_half, View<Fp>{h->eq_r, h->eq_r_stride}, m_log, h->w,
                       h->red);
  } else {
    hipLaunchKernelGGL(k_eq_left<Fp3>, dim3(g), dim3(kBlock), 0, h->stream, View<Fp3>{src[0], sst}, p_half,
                       View<Fp3>{h->eq_l, h->eq_l_stride}, eql_half, View<Fp3>{h->eq_r, h->eq_r_stride}, m_log,
                       h->w, h->red);
  }
  HIPCHK(hipGetLastError());
  return h->ext == 1 ? reduce<Fp, 2>(h, g) : reduce<Fp3, 2>(h, g);
}

// Folds one eq half-table in place (ip::fold semantics on a pow2 table).
int eq_fold_table(effsc_hip_prover* h, uint64_t* planes, size_t stride, size_t* len, const uint64_t* c) {
  if (*len <= 1) return EFFSC_HIP_OK;
  size_t half = *len >> 1;
  unsigned g = blocks_for(h, half);
  uint64_t* pl[kMaxK] = {planes};
  if (h->ext == 1) {
    Views<Fp, 1> v = views<Fp, 1>(pl, stride);
    hipLaunchKernelGGL((k_fold<Fp, Fp, 1>), dim3(g), dim3(kBlock), 0, h->stream, v, v, from_words<Fp>(c), nullptr,
                       half, half, h->w);
  } else {
    Views<Fp3, 1> v = views<Fp3, 1>(pl, stride);
    hipLaunchKernelGGL((k_fold<Fp3, Fp3, 1>), dim3(g), dim3(kBlock), 0, h->stream, v, v, from_words<Fp3>(c),
                       nullptr, half, half, h->w);
  }
  HIPCHK(hipGetLastError());
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
  F s[2] = {gl::zero<F>(), gl::zero<F>()};
  for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < 2; i++) s[i] = gl::add(s[i], gl::load_words<F>(r.partials + ((size_t)b * 2 + i) * F::D));
  block_reduce<F, 2>(s);
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
  const uint64_t c1 = P ? cptr[0] : 0, c2 = P ? cptr[1] : 0;
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  PassAcc<2> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<2>(); i++) acc.a[i].zero();
  const size_t step = (size_t)gridDim.x * blockDim.x * 2;
  for (size_t u = gtid() * 2; u < q; u += step) {
    uint64_t E0[2][4], E1[2][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      fold_chain2<P, NT || P == 0>(src, u + j * q, L, c1, c2, E0[0][j], E1[0][j]);
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
  const uint64_t c1 = P ? cptr[0] : 0, c2 = P ? cptr[1] : 0;
  const size_t L = 4 * q, mask = ((size_t)1 << m_log) - 1;
  Acc128 T[4];
#pragma unroll
  for (int j = 0; j < 4; j++) T[j].zero();
  const size_t u0 = (size_t)blockIdx.x * chunk;
  for (size_t u = u0 + threadIdx.x * 2; u < u0 + chunk; u += (size_t)blockDim.x * 2) {
    uint64_t E0[4], E1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) fold_chain2<P, NT || P == 0>(src, u + j * q, L, c1, c2, E0[j], E1[j]);
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

// Two transcript rounds from the pass values, then the MSB folds of the eq
// table of round s's variable (t0, half h0) with x and of round s+1's (t1,
// half h1) with y; t0 and t1 may be the same table (h1 = h0 / 2 then).
__global__ __launch_bounds__(kReduceBlock) void k_eq_reduce_fs2(Reduce r, unsigned nblocks, DevFs* t, uint64_t* poly,
                                                                uint64_t* chals, uint64_t* cur, uint64_t* t0,
                                                                size_t h0, uint64_t* t1, size_t h1) {
  constexpr int NV = pass_values<2>();
  __shared__ uint64_t xy[3];
  LaneFs fs;
  if (threadIdx.x < 64) fs.load(t);
  Fp sv[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) sv[i] = gl::zero<Fp>();
  sum_partials<Fp, NV>(r.partials, nblocks, sv);
  block_reduce<Fp, NV>(sv);
  if (threadIdx.x < 64) emit_two_rounds<2>(sv, fs, t, poly, chals, xy);
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

template <class F>
F fmul(const effsc_hip_prover* h, F a, F b) {
  return gl::mul(a, b, h->w);
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
        if (E == 1) {
          out[0] = gl::mul(gl::mul(scalar[0], er[0]), p0[0]);
        } else {
          Fp3 v = fmul(h, fmul(h, from_words<Fp3>(scalar), from_words<Fp3>(er)), from_words<Fp3>(p0));
          gl::store_words(v, out);
        }
      }
    } else {
      size_t half = n >> 1;
      uint64_t* planes[kMaxK] = {h->eq_r, src_planes(h)[0]};
      // both tables must share a stride for the K = 2 view; eq_r and p use
      // their own strides, so evaluate through per-table views
      unsigned g = blocks_for(h, half);
      if (E == 1) {
        Views<Fp, 2> t;
        t.t[0] = View<Fp>{planes[0], h->eq_r_stride};
        t.t[1] = View<Fp>{planes[1], src_stride(h)};
        hipLaunchKernelGGL((k_eval<Fp, 2, 1>), dim3(g), dim3(kBlock), 0, h->stream, t, half, half, h->w, h->red);
      } else {
        Views<Fp3, 2> t;
        t.t[0] = View<Fp3>{planes[0], h->eq_r_stride};
        t.t[1] = View<Fp3>{planes[1], src_stride(h)};
        hipLaunchKernelGGL((k_eval<Fp3, 2, 1>), dim3(g), dim3(kBlock), 0, h->stream, t, half, half, h->w, h->red);
      }
      HIPCHK(hipGetLastError());
      TRY((E == 1 ? reduce<Fp, 2>(h, g) : reduce<Fp3, 2>(h, g)));
      uint64_t vals[kMaxK * 3];
      TRY(fetch_result(h, E, vals, 2));
      if (E == 1) {
        out[0] = gl::mul(vals[0], scalar[0]);
        out[1] = gl::mul(vals[1], scalar[0]);
      } else {
        Fp3 s = from_words<Fp3>(scalar);
        gl::store_words(fmul(h, from_words<Fp3>(vals), s), out);
        gl::store_words(fmul(h, from_words<Fp3>(vals + 3), s), out + 3);
      }
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
  if (E == 1) {
    out[0] = gl::mul(gl::mul(p0[0], l0[0]), r0[0]);
  } else {
    Fp3 v = fmul(h, fmul(h, from_words<Fp3>(p0), from_words<Fp3>(l0)), from_words<Fp3>(r0));
    gl::store_words(v, out);
  }
  return EFFSC_HIP_OK;
}

int build_eq(effsc_hip_prover* h, const uint64_t* wpt_canon_dev, size_t off, size_t vv, uint64_t** planes,
             size_t* stride, size_t* len) {
  size_t n = (size_t)1 << vv;
  *stride = round_up(n, 64);
  *len = n;
  if (!*planes) TRY(alloc_dev(planes, *stride * h->ext));
  unsigned g = blocks_for(h, n);
  if (h->ext == 1)
    hipLaunchKernelGGL(k_eq_table<Fp>, dim3(g), dim3(kBlock), 0, h->stream, *planes, *stride,
                       wpt_canon_dev + off * h->ext, vv, h->w);
  else
    hipLaunchKernelGGL(k_eq_table<Fp3>, dim3(g), dim3(kBlock), 0, h->stream, *planes, *stride,
                       wpt_canon_dev + off * h->ext, vv, h->w);
  HIPCHK(hipGetLastError());
  return EFFSC_HIP_OK;
}

}  // namespace

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
namespace {

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
  if (h->ext == 1) {
    out[0] = gkr_f<Fp>(Fp{el[0]}, Fp{el[1]}, Fp{el[2]}, Fp{el[3]}, 0).v;
  } else {
    Fp3 v = gkr_f<Fp3>(from_words<Fp3>(el), from_words<Fp3>(el + 3), from_words<Fp3>(el + 6),
                       from_words<Fp3>(el + 9), h->w);
    gl::store_words(v, out);
  }
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
  if (fi == 1) {
    hipLaunchKernelGGL(k_gkr_eval<Fp>, dim3(g), dim3(kBlock), 0, h->stream, views<Fp, 4>(src, sst), half, paired,
                       h->w, h->red);
    HIPCHK(hipGetLastError());
    TRY((reduce<Fp, 3>(h, g)));
  } else {
    hipLaunchKernelGGL(k_gkr_eval<Fp3>, dim3(g), dim3(kBlock), 0, h->stream, views<Fp3, 4>(src, sst), half, paired,
                       h->w, h->red);
    HIPCHK(hipGetLastError());
    TRY((reduce<Fp3, 3>(h, g)));
  }
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

// From level 4q with j folds applied to level 2q with j + 1 (j + 1 <= k):
// bsrc holds B_j (2 hb elements), bdst receives B_{j+1} (hb elements).
// cconst: the level-2q pairs bind a b variable (j + 1 < k), so w_c is the
// constant of each pair; else w_b is (B_k has one element).
template <class FI, class FW, class F>
__global__ __launch_bounds__(kBlock) void k_gkr_fused_fw(Views<FI, 2> in, Views<F, 2> out, View<FI> bsrc,
                                                         View<F> bdst, size_t hb, View<FW> wit, unsigned k, int cconst,
                                                         const uint64_t* cptr, size_t q, uint64_t w, Reduce r) {
  const F c = gl::load_words<F>(cptr);
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
__global__ __launch_bounds__(kReduceBlock) void k_gkr_wire_fs(Reduce r, unsigned nblocks, DevFs* t, int ext,
                                                              uint64_t* poly, uint64_t* chal_out, uint64_t* chal_cur) {
  F s[2] = {gl::zero<F>(), gl::zero<F>()};
  sum_partials<F, 2>(r.partials, nblocks, s);
  block_reduce<F, 2>(s);
  if (threadIdx.x >= 64) return;
  emit_round<F, 2>(s, ext, t, poly, chal_out, chal_cur);
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

}  // namespace

// ===========================================================================
// Device-resident proof loop (effsc_hip_sumcheck with no hook)
// ===========================================================================
namespace {

bool device_loop_applies(const effsc_hip_prover* h, size_t num_rounds) {
  if (h->kind == K_EQ || h->kind == K_GKR || num_rounds == 0 || num_rounds > kMaxRounds) return false;
  if (!is_pow2(h->len)) return false;
  size_t logl = 0;
  while (((size_t)1 << logl) < h->len) logl++;
  return num_rounds <= logl;  // every round evaluates a table of length >= 2
}

// Words of the proof block: polys | chals | fin (4) | transcript state.  The
// device block and the pinned host staging share this layout, so one copy
// brings a whole proof back (each small copy is a blit of about 5 us on the
// stream).
constexpr size_t kProofWords = kMaxRounds * kMaxK * 3 + kMaxRounds * 3 + 4 + (sizeof(DevFs) + 7) / 8;

int ensure_dev_bufs(effsc_hip_prover* h) {
  if (h->d_fs) return EFFSC_HIP_OK;
  uint64_t* blk = nullptr;
  TRY(alloc_dev(&blk, kProofWords));
  h->d_polys = blk;
  h->d_chals = blk + kMaxRounds * kMaxK * 3;
  h->d_fin = h->d_chals + kMaxRounds * 3;
  h->d_fs = reinterpret_cast<DevFs*>(h->d_fin + 4);
  TRY(alloc_dev(&h->d_cur, 4));
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h->h_proof), kProofWords * 8, hipHostMallocDefault));
  return EFFSC_HIP_OK;
}

// The proof block back to the host staging (polys, challenges, final value
// and transcript state in one copy).
int copy_proof_out(effsc_hip_prover* h) {
  HIPCHK(hipMemcpyAsync(h->h_proof, h->d_polys, kProofWords * 8, hipMemcpyDeviceToHost, h->stream));
  return EFFSC_HIP_OK;
}

struct DevModeGuard {
  effsc_hip_prover* h;
  explicit DevModeGuard(effsc_hip_prover* p) : h(p) { h->dev_mode = true; }
  ~DevModeGuard() { h->dev_mode = false; }
};

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
  HIPCHK(hipMemcpyAsync(h->d_fs, hs, sizeof(DevFs), hipMemcpyHostToDevice, h->stream));

  // host-side pending folds are applied first; from here on the pending
  // challenges live on the device (d_cur[0..p))
  TRY(materialize(h));
  h->have_next = false;
  size_t r = 0;
  int p = 0;
  bool tail_used = false;
  if (h->twostep) {
    // two rounds per pass while the tables are large
    while (num_rounds - r >= 2) {
      size_t L = h->len >> p;
      if (L <= tail_max(h) || L < 4) break;
      Reduce red = h->red;
      unsigned g = 0;
      if (p == 0 && h->pass0_ready && !h->in_work && h->len == h->n) {
        // the pass over the input was accumulated while it streamed in
        red = Reduce{h->d_pass0, h->red.result};
        g = 1;
      } else if (h->ooc && !h->in_work) {
        // out-of-core input: the fold pass streams it from the host again
        if (p != 2) return fail(EFFSC_HIP_ESTATE, "out-of-core pass order");
        TRY(ooc_fold_pass(h, L / 4, &g));
        red = Reduce{h->ooc_parts, h->red.result};
      } else {
        TRY(launch_pass(h, p, L / 4, p ? h->d_cur : nullptr));
        g = h->last_grid;
      }
      if (p > 0) {
        h->len = L;
        h->in_work = true;
      }
      if (K == 1)
        hipLaunchKernelGGL(k_reduce_fs2<1>, dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, h->d_fs,
                           h->d_polys + r, h->d_chals + r, h->d_cur);
      else
        hipLaunchKernelGGL(k_reduce_fs2<2>, dim3(1), dim3(kReduceBlock), 0, h->stream, red, g, h->d_fs,
                           h->d_polys + 2 * r, h->d_chals + r, h->d_cur);
      HIPCHK(hipGetLastError());
      r += 2;
      p = 2;
    }
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
      TRY(dispatch<DevOps>(fr, fr, K, [&](auto op) { return decltype(op)::reduce_fs(h, g, r); }));
      r++;
      p = 1;
    }
  }
  const uint64_t* cur = h->d_cur;
  if (p == 2) {
    // apply the older of two pending challenges with a multi-block fold
    size_t M = h->len;
    uint64_t* const* src = src_planes(h);
    size_t sst = src_stride(h);
    TRY(dispatch(1, 1, K, [&](auto op) { return decltype(op)::fold(h, src, sst, nullptr, M / 2, M / 2, h->d_cur); }));
    h->len = M / 2;
    h->in_work = true;
    p = 1;
    cur = h->d_cur + 1;
  }
  size_t rounds_left = num_rounds - r;
  size_t M = h->len;
  const bool have = p == 1;
  if (rounds_left > 0 || M <= tail_max(h)) {
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
  const uint64_t c1 = cptr[0], c2 = cptr[1];
  const size_t mask = ((size_t)1 << m_log) - 1;
  for (size_t e = gtid(); e < Lt; e += (size_t)gridDim.x * blockDim.x) {
    dst[e] = fold_chain<2>(src, e, Lt, c1, c2);
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
  const size_t tail = h->ext == 3 ? ((size_t)1 << 10) : ((size_t)1 << 12);
  const size_t fold_in_reduce = (size_t)1 << 14;
  size_t L = h->len, j = 0;
  bool have = false;
  View<F> eql{h->eq_l, h->eq_l_stride}, eqr{h->eq_r, h->eq_r_stride};
  const uint64_t* cur_in = h->d_cur;  // the pending challenge when entering the tail
  if constexpr (std::is_same<FI, Fp>::value && std::is_same<F, Fp>::value) {
    if (!std::getenv("EFFSC_HIP_ONE_ROUND")) {
      // two rounds per pass while the level evaluated is above the tail
      int pend = 0;
      while (j + 2 <= v) {
        const size_t Lv = L >> pend;
        if (Lv <= tail || Lv < 8) break;
        size_t m_log = 0;
        while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
        const size_t q = Lv / 4;
        const size_t gmax = (size_t)h->grid_max * kPassGridScale;
        // split-value form while q is a multiple of 2^m: chunks of at least
        // 512 units (one step of the workgroup), inside one block of 2^m
        size_t chunk = 2 * kBlock;
        while (q / chunk > gmax) chunk *= 2;
        const bool sv = q % ((size_t)1 << m_log) == 0 && chunk <= ((size_t)1 << m_log) && q % chunk == 0 &&
                        !std::getenv("EFFSC_HIP_EQ_DIRECT");
        const unsigned g = sv ? (unsigned)(q / chunk)
                              : (unsigned)std::max<size_t>(1, std::min<size_t>((q / 2 + kBlock - 1) / kBlock, gmax));
        const uint64_t* src = h->in_work ? h->work[0] : h->in[0];
        const bool nt = 8.0 * (double)L >= kNtMinBytes;
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
        hipLaunchKernelGGL(k_eq_reduce_fs2, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs,
                           h->d_polys + j * 2, h->d_chals + j, h->d_cur, t0, h0, t1, h1);
        HIPCHK(hipGetLastError());
        j += 2;
        pend = 2;
      }
      const size_t Lt = L >> pend;
      if (pend == 2 && Lt >= 2048 && Lt <= 16384 && !std::getenv("EFFSC_HIP_EQ_TAIL")) {
        // the remaining v - j rounds in the register tail over p and EQ
        if (!h->eq_full) TRY(alloc_dev(&h->eq_full, (size_t)1 << 14));
        size_t m_log = 0;
        while (((size_t)1 << m_log) < h->eq_r_len) m_log++;
        const uint64_t* src = h->in_work ? h->work[0] : h->in[0];
        hipLaunchKernelGGL(k_eq_tail_prep, dim3(blocks_for(h, Lt)), dim3(kBlock), 0, h->stream, src, h->work[0],
                           h->d_cur, Lt, h->eq_l, h->eq_r, (unsigned)m_log, h->eq_full);
        HIPCHK(hipGetLastError());
        Views<Fp, 2> tv{};
        tv.t[0] = View<Fp>{h->work[0], 0};
        tv.t[1] = View<Fp>{h->eq_full, 0};
        const int rounds = (int)(v - j);
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
        HIPCHK(hipGetLastError());
        const int left = j < v_l ? (int)(v_l - j) : 0;
        hipLaunchKernelGGL(k_eq_post_fold, dim3(1), dim3(kReduceBlock), 0, h->stream, h->eq_l, h->eq_l_len, h->eq_r,
                           h->eq_r_len, h->d_chals + j, rounds, left, h->d_cur);
        HIPCHK(hipGetLastError());
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
  const size_t tail = F::D == 3 ? ((size_t)1 << 10) : ((size_t)1 << 12);
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
    hipLaunchKernelGGL(k_gkr_wire_fs<FI>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, E,
                       h->d_polys, h->d_chals, h->d_cur);
    HIPCHK(hipGetLastError());
    r = 1;
    have = true;
    size_t j = 0;  // folds applied to the tables
    while (r < num_rounds && h->len > tail && h->len >= 4 && j + 1 <= kb) {
      M = h->len;
      g = blocks_for(h, M / 8);
      const size_t hb = ((size_t)1 << kb) >> (j + 1);
      TRY(prof_begin(h));
      const int cconst = j + 1 < kb ? 1 : 0;
      if (j == 0)
        hipLaunchKernelGGL((k_gkr_fused_fw<FI, FI, F>), dim3(g), dim3(kBlock), 0, h->stream, in2, w2, wit,
                           bplanes(1), hb, wit, kb, cconst, h->d_cur, M / 4, h->w, h->red);
      else
        hipLaunchKernelGGL((k_gkr_fused_fw<F, FI, F>), dim3(g), dim3(kBlock), 0, h->stream, w2, w2, bplanes(j),
                           bplanes(j + 1), hb, wit, kb, cconst, h->d_cur, M / 4, h->w, h->red);
      HIPCHK(hipGetLastError());
      const int din = j == 0 ? FI::D : F::D;
      TRY(prof_end(h, 0, 8.0 * 2 * ((double)M * din + (double)(M / 2) * F::D), 0.0));
      h->len = M / 2;
      j++;
      hipLaunchKernelGGL(k_gkr_wire_fs<F>, dim3(1), dim3(kReduceBlock), 0, h->stream, h->red, g, h->d_fs, E,
                         h->d_polys + r * 2 * E, h->d_chals + r * E, h->d_cur);
      HIPCHK(hipGetLastError());
      r++;
    }
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
  HIPCHK(hipMemcpyAsync(h->d_fs, hs, sizeof(DevFs), hipMemcpyHostToDevice, h->stream));
  bool tail_used = false;
  const int fi = src_field(h);
  if (fi == 1 && E == 1)
    TRY((gkr_device_run<Fp, Fp>(h, num_rounds, &tail_used)));
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
  HIPCHK(hipMemcpyAsync(h->d_fs, hs, sizeof(DevFs), hipMemcpyHostToDevice, h->stream));
  if (E == 1)
    TRY((eq_device_run<Fp, Fp>(h, num_rounds)));
  else
    TRY((eq_device_run<Fp3, Fp3>(h, num_rounds)));
  uint64_t* hp = h->h_proof;
  uint64_t* hc = hp + kMaxRounds * kMaxK * 3;
  uint64_t* hf = hc + kMaxRounds * 3;
  TRY(copy_proof_out(h));
  HIPCHK(hipStreamSynchronize(h->stream));
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

}  // namespace

// ===========================================================================
// Exported entry points
// ===========================================================================
extern "C" {

void effsc_hip_config_default(effsc_hip_config* cfg) {
  if (!cfg) return;
  cfg->ext = 1;
  cfg->input_ext = 1;
  cfg->repr = EFFSC_HIP_CANONICAL;
  cfg->ext_w = EFFSC_HIP_W_FROM_RAW7;
  cfg->device = 0;
  cfg->flags = 0;
}

const char* effsc_hip_last_error(void) { return effsc_hip::last_error(); }

namespace {

// Launches the bit reversal of `planes` planes of n = 2^v words.
int launch_bitrev(hipStream_t st, uint64_t* dst, size_t dstr, const uint64_t* src, size_t sstr, size_t n, int planes) {
  int v = 0;
  while (((size_t)1 << v) < n) v++;
  if (v >= 2 * kRevT) {
    hipLaunchKernelGGL(k_bitrev_tiled, dim3((unsigned)((size_t)1 << (v - 2 * kRevT)), (unsigned)planes), dim3(kBlock),
                       0, st, dst, dstr, src, sstr, v);
  } else {
    hipLaunchKernelGGL(k_bitrev_small, dim3(1, (unsigned)planes), dim3(kBlock), 0, st, dst, dstr, src, sstr, v);
  }
  HIPCHK(hipGetLastError());
  return EFFSC_HIP_OK;
}

// LSB provers (src/provers/{multilinear,inner_product,coefficient}_lsb.rs):
// the MSB prover of the same kind over the bit-reversed tables.  Pairing
// (2k, 2k+1) of the table is pairing (rev k, rev k + n/2) of its reversal, and
// folding keeps the two tables bit-reversals of each other, so every round
// polynomial, challenge and final value is the same field element.
int new_lsb(const effsc_hip_config* cfg, int kind, int K, int deg, const uint64_t* const* tables, size_t n,
            effsc_hip_prover** out) {
  if (out) *out = nullptr;
  if (!cfg) return fail(EFFSC_HIP_EINVAL, "null config");
  if (!is_pow2(n)) return fail(EFFSC_HIP_EINVAL, "LSB provers take power-of-two tables");
  effsc_hip_config c2 = *cfg;
  c2.flags &= ~EFFSC_HIP_BORROW_INPUT;  // the reordered tables are the prover's own
  effsc_hip_prover* h = nullptr;
  TRY(new_product_like(&c2, kind, K, deg, tables, n, &h));
  int rc = EFFSC_HIP_OK;
  for (int i = 0; i < K && rc == EFFSC_HIP_OK; i++) {
    uint64_t* tmp = nullptr;
    rc = alloc_dev(&tmp, h->in_stride * h->in_ext);
    if (rc == EFFSC_HIP_OK) rc = launch_bitrev(h->stream, tmp, h->in_stride, h->in[i], h->in_stride, n, h->in_ext);
    if (rc == EFFSC_HIP_OK) {
      hipError_t e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, std::string("bit reverse: ") + hipGetErrorString(e));
    }
    if (rc == EFFSC_HIP_OK) {
      (void)hipFree(h->in[i]);
      h->in[i] = tmp;
    } else if (tmp) {
      (void)hipFree(tmp);
    }
  }
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  h->lsb = true;
  *out = h;
  return EFFSC_HIP_OK;
}

}  // namespace

int effsc_hip_multilinear_lsb_new(const effsc_hip_config* cfg, const uint64_t* evals, size_t n, effsc_hip_prover** out) {
  const uint64_t* t[1] = {evals};
  return new_lsb(cfg, K_ML, 1, 1, t, n, out);
}

int effsc_hip_inner_product_lsb_new(const effsc_hip_config* cfg, const uint64_t* a, const uint64_t* b, size_t n,
                                    effsc_hip_prover** out) {
  const uint64_t* t[2] = {a, b};
  return new_lsb(cfg, K_IP, 2, 2, t, n, out);
}

int effsc_hip_product_lsb_new(const effsc_hip_config* cfg, const uint64_t* const* tables, int k, size_t n,
                              effsc_hip_prover** out) {
  if (out) *out = nullptr;
  if (!tables) return fail(EFFSC_HIP_EINVAL, "null tables");
  return new_lsb(cfg, K_PRODUCT, k, k, tables, n, out);
}

int effsc_hip_bit_reverse(int device, uint64_t* dst, const uint64_t* src, size_t n, int planes) {
  if (!dst || !src || planes < 1) return fail(EFFSC_HIP_EINVAL, "bad argument");
  if (!is_pow2(n)) return fail(EFFSC_HIP_EINVAL, "bit reversal needs a power-of-two length");
  if (dst == src) return fail(EFFSC_HIP_EINVAL, "bit reversal is out of place");
  HIPCHK(hipSetDevice(device));
  TRY(launch_bitrev(nullptr, dst, n, src, n, n, planes));
  HIPCHK(hipDeviceSynchronize());
  return EFFSC_HIP_OK;
}

int effsc_hip_multilinear_new(const effsc_hip_config* cfg, const uint64_t* evals, size_t n, effsc_hip_prover** out) {
  const uint64_t* t[1] = {evals};
  return new_product_like(cfg, K_ML, 1, 1, t, n, out);
}

int effsc_hip_inner_product_new(const effsc_hip_config* cfg, const uint64_t* a, const uint64_t* b, size_t n,
                                effsc_hip_prover** out) {
  const uint64_t* t[2] = {a, b};
  return new_product_like(cfg, K_IP, 2, 2, t, n, out);
}

int effsc_hip_product_new(const effsc_hip_config* cfg, const uint64_t* const* tables, int k, size_t n,
                          effsc_hip_prover** out) {
  if (out) *out = nullptr;
  if (!tables) return fail(EFFSC_HIP_EINVAL, "null tables");
  if (n && !is_pow2(n)) return fail(EFFSC_HIP_EINVAL, "product prover tables must have power-of-two length");
  return new_product_like(cfg, K_PRODUCT, k, k, tables, n, out);
}

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
  h->w = cfg->ext == 3 ? gl::canon(cfg->ext_w) : 0;
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
      if (d == 1)
        hipLaunchKernelGGL(k_gkr_broadcast<Fp>, dim3(g), dim3(kBlock), 0, h->stream, View<Fp>{wplanes, n}, n,
                           View<Fp>{h->in[2], h->in_stride}, View<Fp>{h->in[3], h->in_stride});
      else
        hipLaunchKernelGGL(k_gkr_broadcast<Fp3>, dim3(g), dim3(kBlock), 0, h->stream, View<Fp3>{wplanes, n}, n,
                           View<Fp3>{h->in[2], h->in_stride}, View<Fp3>{h->in[3], h->in_stride});
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

int effsc_hip_eq_factored_new(const effsc_hip_config* cfg, const uint64_t* point, size_t v, const uint64_t* p,
                              size_t n, effsc_hip_prover** out) {
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (v >= 48) return fail(EFFSC_HIP_EINVAL, "too many variables");
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
  h->w = cfg->ext == 3 ? gl::canon(cfg->ext_w) : 0;
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

int effsc_hip_degree(const effsc_hip_prover* h, size_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  *out = (size_t)h->deg;
  return EFFSC_HIP_OK;
}

int effsc_hip_ext(const effsc_hip_prover* h, size_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  *out = (size_t)h->ext;
  return EFFSC_HIP_OK;
}

// An out-of-core prover's input is only read by whole device-resident proofs.
static int ooc_guard(const effsc_hip_prover* h) {
  if (h->ooc && !h->in_work)
    return fail(EFFSC_HIP_ESTATE,
                "out-of-core stream prover: the input stays in the streams; run effsc_hip_sumcheck over all rounds "
                "without a hook first");
  return EFFSC_HIP_OK;
}

int effsc_hip_round(effsc_hip_prover* h, const uint64_t* challenge, uint64_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  TRY(ooc_guard(h));
  HIPCHK(hipSetDevice(h->dev));
  uint64_t c[3];
  if (challenge) words_in(h, challenge, c, h->ext);
  size_t nw = (size_t)std::max(h->deg, 1) * h->ext;
  uint64_t res[kMaxK * 3];
  TRY(round_impl(h, challenge ? c : nullptr, res));
  words_out(h, res, out, nw);
  return EFFSC_HIP_OK;
}

int effsc_hip_finalize(effsc_hip_prover* h, const uint64_t* challenge) {
  if (!h || !challenge) return fail(EFFSC_HIP_EINVAL, "null argument");
  TRY(ooc_guard(h));
  HIPCHK(hipSetDevice(h->dev));
  uint64_t c[3];
  words_in(h, challenge, c, h->ext);
  TRY(finalize_impl(h, c));
  HIPCHK(hipStreamSynchronize(h->stream));
  TRY(prof_collect(h));
  return EFFSC_HIP_OK;
}

int effsc_hip_final_value(effsc_hip_prover* h, uint64_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  TRY(ooc_guard(h));
  HIPCHK(hipSetDevice(h->dev));
  uint64_t v[3];
  TRY(final_value_impl(h, v));
  words_out(h, v, out, h->ext);
  return EFFSC_HIP_OK;
}

int effsc_hip_table_count(const effsc_hip_prover* h, size_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  *out = h->kind == K_EQ ? 3 : (size_t)h->K;
  return EFFSC_HIP_OK;
}

int effsc_hip_table_len(const effsc_hip_prover* h, size_t table, size_t* out) {
  if (!h || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  if (h->kind == K_EQ) {
    if (table > 2) return fail(EFFSC_HIP_EINVAL, "table index out of range");
    *out = table == 0 ? h->len : table == 1 ? h->eq_l_len : h->eq_r_len;
    return EFFSC_HIP_OK;
  }
  if (table >= (size_t)h->K) return fail(EFFSC_HIP_EINVAL, "table index out of range");
  *out = logical_len(h);
  return EFFSC_HIP_OK;
}

int effsc_hip_read_table(effsc_hip_prover* h, size_t table, uint64_t* dst, size_t cap_words, size_t* len) {
  if (!h) return fail(EFFSC_HIP_EINVAL, "null prover");
  size_t L = 0;
  TRY(effsc_hip_table_len(h, table, &L));
  if (len) *len = L;
  const int E = h->ext;
  if (!dst) return cap_words == 0 ? EFFSC_HIP_OK : fail(EFFSC_HIP_EINVAL, "null destination");
  if (cap_words < L * E) return fail(EFFSC_HIP_EINVAL, "destination too small");
  TRY(ooc_guard(h));
  HIPCHK(hipSetDevice(h->dev));
  TRY(materialize(h));
  const uint64_t* planes;
  size_t stride;
  int d;
  if (h->kind == K_EQ && table > 0) {
    planes = table == 1 ? h->eq_l : h->eq_r;
    stride = table == 1 ? h->eq_l_stride : h->eq_r_stride;
    d = E;
  } else {
    planes = src_planes(h)[table];
    stride = src_stride(h);
    d = src_field(h);
  }
  std::vector<uint64_t> buf(std::max<size_t>(L, 1) * d);
  for (int c = 0; c < d && L; c++)
    HIPCHK(hipMemcpyAsync(buf.data() + c * L, planes + c * stride, L * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  // LSB provers: element i of the reference's table is element rev(i) here
  int lv = 0;
  while (h->lsb && ((size_t)1 << lv) < L) lv++;
  for (size_t i = 0; i < L; i++) {
    const size_t j = (h->lsb && lv) ? (size_t)(gl_brev64(i) >> (64 - lv)) : i;
    for (int c = 0; c < E; c++) {
      uint64_t x = c < d ? buf[c * L + j] : 0;
      dst[i * E + c] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(x, gl::CANON_TO_MONT) : x;
    }
  }
  return EFFSC_HIP_OK;
}

int effsc_hip_reset(effsc_hip_prover* h) {
  if (!h) return fail(EFFSC_HIP_EINVAL, "null prover");
  if (h->kind == K_EQ) {
    // the eq tables are folded in place: rebuild them from the stored point
    HIPCHK(hipSetDevice(h->dev));
    TRY(build_eq(h, h->eq_point, 0, h->v_l, &h->eq_l, &h->eq_l_stride, &h->eq_l_len));
    TRY(build_eq(h, h->eq_point, h->v_l, h->v - h->v_l, &h->eq_r, &h->eq_r_stride, &h->eq_r_len));
    h->rounds_elapsed = 0;
  }
  h->len = h->n;
  h->in_work = false;
  h->npend = 0;
  h->have_next = false;
  return EFFSC_HIP_OK;
}

void effsc_hip_free(effsc_hip_prover* h) { destroy(h); }

int effsc_hip_profile(effsc_hip_prover* h, int enable) {
  if (!h) return fail(EFFSC_HIP_EINVAL, "null prover");
  HIPCHK(hipSetDevice(h->dev));
  h->prof_on = enable != 0;
  h->pend.clear();
  for (int i = 0; i < 4; i++) {
    h->prof_launches[i] = 0;
    h->prof_ms[i] = 0;
    h->prof_bytes[i] = 0;
    h->prof_muls[i] = 0;
  }
  return EFFSC_HIP_OK;
}

int effsc_hip_profile_read(const effsc_hip_prover* h, uint64_t* launches, double* ms, double* bytes) {
  if (!h || !launches || !ms || !bytes) return fail(EFFSC_HIP_EINVAL, "null argument");
  for (int i = 0; i < 4; i++) {
    launches[i] = h->prof_launches[i];
    ms[i] = h->prof_ms[i];
    bytes[i] = h->prof_bytes[i];
  }
  return EFFSC_HIP_OK;
}

int effsc_hip_profile_read_muls(const effsc_hip_prover* h, double* muls) {
  if (!h || !muls) return fail(EFFSC_HIP_EINVAL, "null argument");
  for (int i = 0; i < 4; i++) muls[i] = h->prof_muls[i];
  return EFFSC_HIP_OK;
}

int effsc_hip_synchronize(int device) {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipDeviceSynchronize());
  return EFFSC_HIP_OK;
}

int effsc_hip_fold(int device, uint64_t* values, size_t n, uint64_t weight, size_t* new_len) {
  if (!values && n) return fail(EFFSC_HIP_EINVAL, "null values");
  HIPCHK(hipSetDevice(device));
  if (n <= 1) {
    if (new_len) *new_len = n;
    return EFFSC_HIP_OK;
  }
  size_t half = next_pow2(n) >> 1, paired = n - half;
  uint64_t* pl[kMaxK] = {values};
  Views<Fp, 1> v = views<Fp, 1>(pl, n);
  size_t b = std::min<size_t>((half + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL((k_fold<Fp, Fp, 1>), dim3((unsigned)std::max<size_t>(b, 1)), dim3(kBlock), 0, nullptr, v, v,
                     Fp{gl::canon(weight)}, nullptr, half, paired, (uint64_t)0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  if (new_len) *new_len = half;
  return EFFSC_HIP_OK;
}

// ----------------------------------------------------------------- transcripts
effsc_hip_transcript* effsc_hip_transcript_new(const uint8_t* domsep, size_t len) {
  auto* t = new effsc_hip_transcript();
  t->kind = 0;
  t->fs.init(domsep, domsep ? len : 0);
  return t;
}

effsc_hip_transcript* effsc_hip_transcript_new_test(uint64_t seed) {
  auto* t = new effsc_hip_transcript();
  t->kind = 1;
  t->rng = seed;
  return t;
}

int effsc_hip_transcript_send(effsc_hip_transcript* t, int ext, const uint64_t* value) {
  if (!t || !value || ext < 1 || ext > 3) return fail(EFFSC_HIP_EINVAL, "bad transcript argument");
  if (t->kind == 0) t->fs.send(ext, value);
  return EFFSC_HIP_OK;
}

int effsc_hip_transcript_challenge(effsc_hip_transcript* t, int ext, uint64_t* out) {
  if (!t || !out || ext < 1 || ext > 3) return fail(EFFSC_HIP_EINVAL, "bad transcript argument");
  if (t->kind == 0) {
    t->fs.challenge(ext, out);
  } else {
    // splitmix64; F::from_u64 reduces mod p (src/transcript/sanity.rs:42-44)
    t->rng += 0x9E3779B97F4A7C15ull;
    uint64_t z = t->rng;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[0] = gl::canon(z);
    for (int i = 1; i < ext; i++) out[i] = 0;
  }
  return EFFSC_HIP_OK;
}

void effsc_hip_transcript_free(effsc_hip_transcript* t) { delete t; }

// --------------------------------------------------------------------- runner
int effsc_hip_sumcheck(effsc_hip_prover* h, size_t num_rounds, effsc_hip_transcript* t, effsc_hip_hook hook,
                       void* hook_ctx, uint64_t* round_polys, uint64_t* challenges, uint64_t* final_value) {
  // src/runner.rs:29-69
  if (!h || !t || !final_value || (num_rounds && (!round_polys || !challenges)))
    return fail(EFFSC_HIP_EINVAL, "null argument");
  HIPCHK(hipSetDevice(h->dev));
  if (h->ooc && !h->in_work) {
    // the input is read by the two streamed passes of a whole device proof
    const bool whole = num_rounds == h->v && !hook && device_loop_applies(h, num_rounds) &&
                       !std::getenv("EFFSC_HIP_HOST_LOOP") && h->len == h->n && h->n / 4 > tail_max(h);
    if (!whole) TRY(ooc_guard(h));
  }
  if (!hook && device_loop_applies(h, num_rounds) && !std::getenv("EFFSC_HIP_HOST_LOOP"))
    return prove_device(h, num_rounds, t, round_polys, challenges, final_value);
  if (!hook && eq_device_applies(h, num_rounds) && !std::getenv("EFFSC_HIP_HOST_LOOP"))
    return prove_device_eq(h, num_rounds, t, round_polys, challenges, final_value);
  if (!hook && gkr_device_applies(h, num_rounds) && !std::getenv("EFFSC_HIP_HOST_LOOP"))
    return prove_device_gkr(h, num_rounds, t, round_polys, challenges, final_value);
  const int E = h->ext;
  const size_t nw = (size_t)std::max(h->deg, 1);
  uint64_t prev[3];
  bool have = false;
  for (size_t r = 0; r < num_rounds; r++) {
    uint64_t* evals = round_polys + r * nw * E;
    TRY(round_impl(h, have ? prev : nullptr, evals));
    for (size_t i = 0; i < nw; i++) TRY(effsc_hip_transcript_send(t, E, evals + i * E));
    if (hook) hook(hook_ctx, r, t);
    TRY(effsc_hip_transcript_challenge(t, E, prev));
    std::copy(prev, prev + E, challenges + r * E);
    have = true;
  }
  if (have) {
    TRY(finalize_impl(h, prev));
    HIPCHK(hipStreamSynchronize(h->stream));
    TRY(prof_collect(h));
  }
  TRY(final_value_impl(h, final_value));
  return EFFSC_HIP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------- verifier
namespace {

template <class F>
F fconst(uint64_t x) {
  return gl::lift_to<F>(Fp{gl::canon(x)});
}

// evaluate_from_evals + barycentric_weight: src/verifier.rs:146-200
template <class F>
F evaluate_from_evals(const std::vector<F>& ev, F r, uint64_t w) {
  size_t d = ev.size();
  if (d == 0) return gl::zero<F>();
  if (d == 1) return ev[0];
  if (d == 2) return gl::add(ev[0], gl::mul(r, gl::sub(ev[1], ev[0]), w));
  std::vector<F> rm(d), prefix(d + 1, gl::one<F>()), suffix(d + 1, gl::one<F>());
  for (size_t j = 0; j < d; j++) rm[j] = gl::sub(r, fconst<F>(j));
  for (size_t i = 0; i < d; i++) prefix[i + 1] = gl::mul(prefix[i], rm[i], w);
  for (size_t i = d; i-- > 0;) suffix[i] = gl::mul(suffix[i + 1], rm[i], w);
  F result = gl::zero<F>();
  for (size_t i = 0; i < d; i++) {
    F num = gl::mul(prefix[i], suffix[i + 1], w);
    F den = gl::one<F>();
    for (size_t j = 0; j < d; j++) {
      if (j == i) continue;
      long long diff = (long long)i - (long long)j;
      F df = diff > 0 ? fconst<F>((uint64_t)diff) : gl::sub(gl::zero<F>(), fconst<F>((uint64_t)(-diff)));
      den = gl::mul(den, df, w);
    }
    if (!gl::is_zero(den)) result = gl::add(result, gl::mul(gl::mul(ev[i], num, w), gl::inv(den, w), w));
  }
  return result;
}

template <class F>
int verify_t(uint64_t w, const uint64_t* claimed, size_t d, size_t num_rounds, effsc_hip_transcript* t,
             const uint64_t* proof, size_t proof_words, effsc_hip_hook hook, void* ctx, uint64_t* challenges,
             uint64_t* final_claim, size_t* err_round, int* err_kind) {
  // src/verifier.rs:61-139
  const int E = F::D;
  F claim = gl::load_words<F>(claimed);
  size_t cursor = 0;
  for (size_t round = 0; round < num_rounds; round++) {
    size_t n_wire = std::max<size_t>(d, 1);
    std::vector<F> recv;
    for (size_t i = 0; i < n_wire; i++) {
      if (cursor + E > proof_words) {
        if (err_round) *err_round = round;
        if (err_kind) *err_kind = 1;
        return fail(EFFSC_HIP_EVERIFY, "transcript error: proof exhausted");
      }
      recv.push_back(gl::load_words<F>(proof + cursor));
      TRY(effsc_hip_transcript_send(t, E, proof + cursor));
      cursor += E;
    }
    F h0 = recv[0];
    F h1 = gl::sub(claim, h0);
    if (hook && hook(ctx, round, t) != 0) {
      if (err_round) *err_round = round;
      if (err_kind) *err_kind = 2;
      return fail(EFFSC_HIP_EVERIFY, "hook error");
    }
    uint64_t rw[3];
    TRY(effsc_hip_transcript_challenge(t, E, rw));
    std::copy(rw, rw + E, challenges + round * E);
    F r = gl::load_words<F>(rw);
    if (d == 0) {
      claim = h0;
      continue;
    }
    F h_inf = d >= 2 ? recv[1] : gl::sub(h1, h0);
    std::vector<F> q;
    q.push_back(h0);
    q.push_back(gl::sub(h1, h_inf));
    for (size_t i = 2; i < d; i++) {
      F i_f = fconst<F>(i), i_d = gl::one<F>();
      for (size_t j = 0; j < d; j++) i_d = gl::mul(i_d, i_f, w);
      q.push_back(gl::sub(recv[i], gl::mul(h_inf, i_d, w)));
    }
    F q_r = evaluate_from_evals<F>(q, r, w);
    F r_d = gl::one<F>();
    for (size_t j = 0; j < d; j++) r_d = gl::mul(r_d, r, w);
    claim = gl::add(q_r, gl::mul(h_inf, r_d, w));
  }
  gl::store_words(claim, final_claim);
  return EFFSC_HIP_OK;
}

}  // namespace

extern "C" int effsc_hip_sumcheck_verify(int ext, uint64_t w, const uint64_t* claimed_sum, size_t degree,
                                         size_t num_rounds, effsc_hip_transcript* t, const uint64_t* proof,
                                         size_t proof_words, effsc_hip_hook hook, void* hook_ctx,
                                         uint64_t* challenges, uint64_t* final_claim, size_t* err_round,
                                         int* err_kind) {
  if (!claimed_sum || !t || !final_claim || (num_rounds && !challenges) || (proof_words && !proof))
    return fail(EFFSC_HIP_EINVAL, "null argument");
  if (err_kind) *err_kind = 0;
  if (ext == 1)
    return verify_t<Fp>(0, claimed_sum, degree, num_rounds, t, proof, proof_words, hook, hook_ctx, challenges,
                        final_claim, err_round, err_kind);
  if (ext == 3)
    return verify_t<Fp3>(gl::canon(w), claimed_sum, degree, num_rounds, t, proof, proof_words, hook, hook_ctx,
                         challenges, final_claim, err_round, err_kind);
  return fail(EFFSC_HIP_EINVAL, "ext must be 1 or 3");
}

// ------------------------------------------------------------------ utilities
// Throughput probe: each thread runs 8 independent chains x <- x * y of
// Goldilocks products (the VALU roofline of the degree >= 3 rounds).
__global__ __launch_bounds__(kBlock) void k_mul_rate(uint64_t* out, uint64_t y, int iters) {
  uint64_t x[8];
#pragma unroll
  for (int i = 0; i < 8; i++) x[i] = gtid() * 8 + i + 1;
  for (int it = 0; it < iters; it++)
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = gl::mul(x[i], y);
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) s ^= x[i];
  out[gtid()] = s;
}

extern "C" int effsc_hip_mul_rate(int device, double* muls_per_s) {
  if (!muls_per_s) return fail(EFFSC_HIP_EINVAL, "null out");
  HIPCHK(hipSetDevice(device));
  const unsigned grid = 8192;
  const int iters = 2048;
  uint64_t* out = nullptr;
  TRY(alloc_dev(&out, (size_t)grid * kBlock));
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float ms = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mul_rate, dim3(grid), dim3(kBlock), 0, st, out, 0x123456789ABCDEFull, iters);
    e = hipEventRecord(e0, st);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mul_rate, dim3(grid), dim3(kBlock), 0, st, out, 0xFEDCBA987654321ull, iters);
    e = hipEventRecord(e1, st);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(EFFSC_HIP_EDEVICE, std::string("mul rate: ") + hipGetErrorString(e));
  *muls_per_s = (double)grid * kBlock * 8 * iters / (ms * 1e-3);
  return EFFSC_HIP_OK;
}

// Attainable HBM rates: kernels with no field arithmetic that move the bytes
// of a stream copy, of the read pass (P = 0: two tables, four quarter
// regions each) and of the fold pass (P = 2: sixteen sixteenth regions read
// and four quarter regions of the level-n/4 table written per table), in the
// order and with the load/store kinds of k_pass.  The read kernels keep
// their xor live through a store that no input makes.
__global__ __launch_bounds__(kBlock) void k_bw_copy(const v2u64* s, v2u64* d, size_t n2) {
  for (size_t i = gtid(); i < n2; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}

__global__ __launch_bounds__(kBlock) void k_bw_read(const uint64_t* a, const uint64_t* b, size_t q, size_t chunk,
                                                    uint64_t* sink) {
  v2u64 acc = {0, 0};
  const size_t uend = min(q, (blockIdx.x + 1) * chunk);
  for (size_t u = blockIdx.x * chunk + threadIdx.x * 2; u < uend; u += (size_t)blockDim.x * 2) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      acc ^= __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(a + u + j * q));
      acc ^= __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(b + u + j * q));
    }
  }
  if (acc.x == 0x9E3779B97F4A7C15ull && acc.y == 0x243F6A8885A308D3ull) sink[0] = acc.x;
}

__global__ __launch_bounds__(kBlock) void k_bw_fold(const uint64_t* a, const uint64_t* b, uint64_t* oa, uint64_t* ob,
                                                    size_t q) {
  for (size_t u = gtid() * 2; u < q; u += (size_t)gridDim.x * blockDim.x * 2) {
    v2u64 x[2][4] = {};
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int m = 0; m < 4; m++) {
        x[0][j] ^= __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(a + u + (j + 4 * m) * q));
        x[1][j] ^= __builtin_nontemporal_load(reinterpret_cast<const v2u64*>(b + u + (j + 4 * m) * q));
      }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      __builtin_nontemporal_store(x[0][j], reinterpret_cast<v2u64*>(oa + u + j * q));
      __builtin_nontemporal_store(x[1][j], reinterpret_cast<v2u64*>(ob + u + j * q));
    }
  }
}

extern "C" int effsc_hip_hbm_rate(int device, int pattern, size_t n, double* bytes_per_s) {
  if (!bytes_per_s) return fail(EFFSC_HIP_EINVAL, "null out");
  if (pattern < EFFSC_HIP_HBM_COPY || pattern > EFFSC_HIP_HBM_FOLD_PASS)
    return fail(EFFSC_HIP_EINVAL, "unknown access pattern");
  if (!is_pow2(n) || n < 1024) return fail(EFFSC_HIP_EINVAL, "n must be a power of two >= 1024");
  HIPCHK(hipSetDevice(device));
  const size_t gmax = (size_t)grid_max_for(device) * kPassGridScale;
  // two tables of n words, their level-n/4 outputs, one sink word
  uint64_t* buf = nullptr;
  TRY(alloc_dev(&buf, 2 * n + n / 2 + 32));
  uint64_t *a = buf, *b = buf + n, *oa = buf + 2 * n, *ob = oa + n / 4, *sink = ob + n / 4;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(buf, 0, (2 * n + n / 2 + 32) * 8, st);
  double bytes = 0;
  auto launch = [&] {
    if (pattern == EFFSC_HIP_HBM_COPY) {
      bytes = 16.0 * n;
      hipLaunchKernelGGL(k_bw_copy, dim3((unsigned)std::min<size_t>(n / 2 / kBlock, gmax / 8)), dim3(kBlock), 0, st,
                         reinterpret_cast<const v2u64*>(a), reinterpret_cast<v2u64*>(b), n / 2);
    } else if (pattern == EFFSC_HIP_HBM_READ_PASS) {
      bytes = 16.0 * n;
      const size_t q = n / 4, step = (size_t)kBlock * 2;
      const size_t g = std::min<size_t>((q / 2 + kBlock - 1) / kBlock, gmax);
      const size_t chunk = (q + step * g - 1) / (step * g) * step;
      hipLaunchKernelGGL(k_bw_read, dim3((unsigned)((q + chunk - 1) / chunk)), dim3(kBlock), 0, st, a, b, q, chunk,
                         sink);
    } else {
      bytes = 20.0 * n;
      const size_t q = n / 16;
      hipLaunchKernelGGL(k_bw_fold, dim3((unsigned)std::min<size_t>((q / 2 + kBlock - 1) / kBlock, gmax)), dim3(kBlock),
                         0, st, a, b, oa, ob, q);
    }
  };
  float best = 0;
  for (int rep = 0; rep < 6 && e == hipSuccess; rep++) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) launch();
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && rep > 0 && (best == 0 || ms < best)) best = ms;  // the first launch warms up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(EFFSC_HIP_EDEVICE, std::string("hbm rate: ") + hipGetErrorString(e));
  *bytes_per_s = bytes / (best * 1e-3);
  return EFFSC_HIP_OK;
}

extern "C" int effsc_hip_device_alloc(int device, size_t words, uint64_t** out) {
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  HIPCHK(hipSetDevice(device));
  return alloc_dev(out, words);
}

extern "C" int effsc_hip_device_free(int device, uint64_t* ptr) {
  HIPCHK(hipSetDevice(device));
  if (ptr) HIPCHK(hipFree(ptr));
  return EFFSC_HIP_OK;
}

extern "C" int effsc_hip_copy(int device, uint64_t* dst, const uint64_t* src, size_t words) {
  HIPCHK(hipSetDevice(device));
  if (words) HIPCHK(hipMemcpy(dst, src, words * 8, hipMemcpyDefault));
  return EFFSC_HIP_OK;
}

extern "C" int effsc_hip_fill_uniform(int device, uint64_t* dst, size_t n, uint64_t seed, uint64_t first) {
  if (!dst && n) return fail(EFFSC_HIP_EINVAL, "null destination");
  HIPCHK(hipSetDevice(device));
  if (!n) return EFFSC_HIP_OK;
  uint64_t z = uniform_key(seed);
  size_t b = std::min<size_t>((n + kBlock - 1) / kBlock, 16384);
  hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)b), dim3(kBlock), 0, nullptr, dst, n, z, first);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return EFFSC_HIP_OK;
}

// ===========================================================================
// Streamed input.  FileStream (src/streams/file/core.rs:13-76): a file of
// 2^v serialized field elements (ark-serialize uncompressed: each component
// one little-endian canonical u64), mapped read-only.  Tables move to the
// device in chunks through pinned double buffers on a copy stream while the
// compute stream works on the previous chunk, so the transfer hides the
// first pass over the data.
// ===========================================================================
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <thread>

struct effsc_hip_stream {
  int fd = -1;
  const uint8_t* map = nullptr;
  size_t bytes = 0;
  int ext = 1;
  size_t n = 0;
  unsigned v = 0;
};

namespace {

// Host copy of several segments (the mapped files are pageable, so they are
// staged through pinned memory by the CPU) with up to 8 threads sharing the
// total bytes: a chunk is 4 or 16 regions per table of a few MiB each, too
// small to split one at a time.
struct CopySeg {
  void* dst;
  const void* src;
  size_t bytes;
};
void par_copy_segs(const std::vector<CopySeg>& segs) {
  size_t total = 0;
  for (const CopySeg& sg : segs) total += sg.bytes;
  const size_t per = (size_t)4 << 20;
  static const int max_threads = [] {
    const char* e = std::getenv("EFFSC_HIP_COPY_THREADS");
    const int x = e ? std::atoi(e) : 0;
    return x >= 1 && x <= 64 ? x : 8;
  }();
  const int nt = (int)std::min<size_t>((size_t)max_threads, (total + per - 1) / per);
  if (nt <= 1) {
    for (const CopySeg& sg : segs) std::memcpy(sg.dst, sg.src, sg.bytes);
    return;
  }
  const size_t share = (total + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    const size_t lo = (size_t)t * share, hi = std::min(total, lo + share);
    if (lo >= hi) break;
    th.emplace_back([&segs, lo, hi] {
      size_t base = 0;
      for (const CopySeg& sg : segs) {
        const size_t a = std::max(lo, base), b = std::min(hi, base + sg.bytes);
        if (a < b) std::memcpy((uint8_t*)sg.dst + (a - base), (const uint8_t*)sg.src + (a - base), b - a);
        base += sg.bytes;
        if (base >= hi) break;
      }
    });
  }
  for (auto& x : th) x.join();
}

size_t stream_chunk_units() {
  if (const char* e = std::getenv("EFFSC_HIP_STREAM_CHUNK")) {
    long long x = std::atoll(e);
    if (x >= 1 && !(x & (x - 1))) return (size_t)x;
  }
  return (size_t)1 << 20;
}

// Round 0 values and round 1 coefficients (the P = 0 two-round pass) of the
// units [u0, u0 + count) of resident base tables, one partial per block;
// flags any word >= p (a serialized element must be canonical).
// Unit u's element j of table i is t.t[i].p[(u - base) + j * stride]: the
// resident table (base 0, stride q) or a staged chunk (base u0, stride B).
template <int K>
__global__ __launch_bounds__(kBlock) void k_stream_pass(Views<Fp, K> t, size_t stride, size_t base, size_t u0,
                                                        size_t count, uint64_t* partials, unsigned* bad) {
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) acc.a[i].zero();
  unsigned b = 0;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < count; k += step) {
    const size_t u = u0 + k - base;
    uint64_t E[K][4];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        E[i][j] = t.t[i].p[u + j * stride];
        b |= E[i][j] >= gl::P;
      }
    pass_accumulate<K>(E, acc);
  }
  if (b) atomicOr(bad, 1u);
  Fp raw[pass_raw<K>()], vals[pass_values<K>()];
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  finish<Fp, pass_values<K>()>(vals, Reduce{partials, nullptr});
}

// The streamed fold pass of an out-of-core prover (the P = 2 two-round pass
// from the input, pass2.hpp): the staged chunk holds, per table, the 16
// regions m q + [u0, u0 + count) of the input (M = 16 q elements) at m B;
// unit u folds input elements u + (j + 4t) q (t < 4) with c1 = cptr[0],
// c2 = cptr[1] into the level-4q element u + j q of `work` and accumulates
// the pass values, one partial per block.
template <int K>
__global__ __launch_bounds__(kBlock) void k_stream_fold(Views<Fp, K> stage, size_t B, size_t u0, size_t count,
                                                        Views<Fp, K> work, size_t q, const uint64_t* cptr,
                                                        uint64_t* partials) {
  const uint64_t c1 = cptr[0], c2 = cptr[1];
  PassAcc<K> acc;
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) acc.a[i].zero();
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < count; k += step) {
    const size_t u = u0 + k;
    uint64_t E[K][4];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t* x = stage.t[i].p + k;
        const uint64_t x0 = x[j * B], x1 = x[(j + 4) * B], x2 = x[(j + 8) * B], x3 = x[(j + 12) * B];
        const uint64_t a = gl::add(x0, gl::mul(c1, gl::sub(x2, x0)));
        const uint64_t b = gl::add(x1, gl::mul(c1, gl::sub(x3, x1)));
        E[i][j] = gl::add(a, gl::mul(c2, gl::sub(b, a)));
        work.t[i].p[u + j * q] = E[i][j];
      }
    pass_accumulate<K>(E, acc);
  }
  Fp raw[pass_raw<K>()], vals[pass_values<K>()];
#pragma unroll
  for (int i = 0; i < pass_raw<K>(); i++) raw[i] = acc.a[i].reduce();
  pass_values_from_raw<K>(raw, vals);
  finish<Fp, pass_values<K>()>(vals, Reduce{partials, nullptr});
}

// Chunked transfer of an out-of-core prover's input: the n elements of each
// of the K streams are cut into nreg regions of R = n / nreg; chunk c moves
// elements [c B, c B + cnt) of every region (host pages -> pinned buffer ->
// device staging, double-buffered, on a copy stream) and then `consume(c, u0,
// cnt, staged)` enqueues the chunk's kernel on the prover's stream, which
// sees table i region m at staged + (i nreg + m) B.  Returns after the last
// kernel has finished reading the staging.
struct OocChunks {
  size_t B = 0, nch = 0;
  unsigned G = 0;  // workgroups per chunk kernel
};
OocChunks ooc_chunks(size_t R) {
  OocChunks c;
  c.B = std::max<size_t>(1, std::min(R, stream_chunk_units() >> 2));
  c.nch = (R + c.B - 1) / c.B;
  c.G = (unsigned)std::min<size_t>((c.B + kBlock - 1) / kBlock, 1024);
  return c;
}

template <class Consume>
int ooc_stream(effsc_hip_prover* h, size_t nreg, Consume&& consume) {
  const int K = h->K;
  const size_t R = h->n / nreg;
  const OocChunks ch = ooc_chunks(R);
  const size_t words = (size_t)K * nreg * ch.B;
  int rc = EFFSC_HIP_OK;
  hipStream_t cs = nullptr;
  uint8_t* pin[2] = {};
  uint64_t* stage[2] = {};
  hipEvent_t copied[2] = {}, used[2] = {};
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == EFFSC_HIP_OK) rc = fail(EFFSC_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return rc == EFFSC_HIP_OK;
  };
  chk(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
  for (int b = 0; b < 2 && rc == EFFSC_HIP_OK; b++) {
    chk(hipHostMalloc((void**)&pin[b], words * 8, hipHostMallocDefault), "pinned staging");
    if (rc == EFFSC_HIP_OK) rc = alloc_dev(&stage[b], words);
    chk(hipEventCreateWithFlags(&copied[b], hipEventDisableTiming), "event");
    chk(hipEventCreateWithFlags(&used[b], hipEventDisableTiming), "event");
  }
  for (size_t c = 0; c < ch.nch && rc == EFFSC_HIP_OK; c++) {
    const int b = (int)(c & 1);
    const size_t u0 = c * ch.B, cnt = std::min(ch.B, R - u0);
    // pinned buffer b was last read by the copy of chunk c - 2
    if (c >= 2 && !chk(hipEventSynchronize(copied[b]), "wait")) break;
    std::vector<CopySeg> segs;
    for (int i = 0; i < K; i++)
      for (size_t m = 0; m < nreg; m++)
        segs.push_back({pin[b] + ((size_t)i * nreg + m) * ch.B * 8, h->ooc_src[i]->map + (m * R + u0) * 8, cnt * 8});
    par_copy_segs(segs);
    // device staging b was last read by the kernel of chunk c - 2
    if (c >= 2) chk(hipStreamWaitEvent(cs, used[b], 0), "wait");
    chk(hipMemcpyAsync(stage[b], pin[b], words * 8, hipMemcpyHostToDevice, cs), "copy");
    chk(hipEventRecord(copied[b], cs), "event");
    chk(hipStreamWaitEvent(h->stream, copied[b], 0), "wait");
    if (rc != EFFSC_HIP_OK) break;
    TRY(consume(c, u0, cnt, stage[b], ch));
    chk(hipEventRecord(used[b], h->stream), "event");
  }
  if (cs) chk(hipStreamSynchronize(cs), "sync");
  chk(hipStreamSynchronize(h->stream), "sync");
  for (int b = 0; b < 2; b++) {
    if (pin[b]) (void)hipHostFree(pin[b]);
    if (stage[b]) (void)hipFree(stage[b]);
    if (copied[b]) (void)hipEventDestroy(copied[b]);
    if (used[b]) (void)hipEventDestroy(used[b]);
  }
  if (cs) (void)hipStreamDestroy(cs);
  return rc;
}

template <int K>
Views<Fp, K> staged_views(uint64_t* staged, size_t nreg, size_t B) {
  Views<Fp, K> v{};
  for (int i = 0; i < K; i++) v.t[i] = View<Fp>{staged + (size_t)i * nreg * B, B};
  return v;
}

// Rounds 0-1 of an out-of-core prover: round 0 values and round 1
// coefficients of the input into d_pass0, with the canonical-word check.
int ooc_pass0(effsc_hip_prover* h) {
  const size_t q = h->n / 4;
  const OocChunks ch = ooc_chunks(q);
  constexpr int NVmax = kMaxK + kMaxK * (kMaxK + 1);
  uint64_t* parts = nullptr;
  unsigned* bad = nullptr;
  TRY(alloc_dev(&parts, ch.nch * ch.G * NVmax));
  int rc = alloc_dev(reinterpret_cast<uint64_t**>(&bad), 1);
  if (rc == EFFSC_HIP_OK && hipMemsetAsync(bad, 0, 4, h->stream) != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, "flag");
  if (rc == EFFSC_HIP_OK)
    rc = ooc_stream(h, 4, [&](size_t c, size_t u0, size_t cnt, uint64_t* st, const OocChunks& cc) -> int {
      if (h->K == 1)
        hipLaunchKernelGGL(k_stream_pass<1>, dim3(cc.G), dim3(kBlock), 0, h->stream, staged_views<1>(st, 4, cc.B), cc.B,
                           u0, u0, cnt, parts + c * cc.G * pass_values<1>(), bad);
      else
        hipLaunchKernelGGL(k_stream_pass<2>, dim3(cc.G), dim3(kBlock), 0, h->stream, staged_views<2>(st, 4, cc.B), cc.B,
                           u0, u0, cnt, parts + c * cc.G * pass_values<2>(), bad);
      HIPCHK(hipGetLastError());
      return EFFSC_HIP_OK;
    });
  if (rc == EFFSC_HIP_OK) {
    if (h->K == 1)
      hipLaunchKernelGGL((k_reduce<Fp, pass_values<1>()>), dim3(1), dim3(1024), 0, h->stream,
                         Reduce{parts, h->d_pass0}, (unsigned)(ch.nch * ch.G));
    else
      hipLaunchKernelGGL((k_reduce<Fp, pass_values<2>()>), dim3(1), dim3(1024), 0, h->stream,
                         Reduce{parts, h->d_pass0}, (unsigned)(ch.nch * ch.G));
    if (hipGetLastError() != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, "reduce");
  }
  unsigned hbad = 0;
  if (rc == EFFSC_HIP_OK && hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    rc = fail(EFFSC_HIP_EDEVICE, "copy back");
  if (rc == EFFSC_HIP_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(EFFSC_HIP_EDEVICE, "sync");
  if (rc == EFFSC_HIP_OK && hbad) rc = fail(EFFSC_HIP_EINVAL, "non-canonical serialized element");
  (void)hipStreamSynchronize(h->stream);
  if (parts) (void)hipFree(parts);
  if (bad) (void)hipFree(bad);
  return rc;
}

// Rounds 2-3 of an out-of-core prover: the P = 2 pass over the streamed
// input with the two challenges in d_cur; writes the level-n/4 tables into
// `work` and leaves nblocks partials in ooc_parts.
int ooc_fold_pass(effsc_hip_prover* h, size_t q, unsigned* nblocks) {
  const OocChunks ch = ooc_chunks(q);
  if (ch.nch * ch.G > h->ooc_nparts) return fail(EFFSC_HIP_ESTATE, "out-of-core partials too small");
  Views<Fp, 2> wv{};
  for (int i = 0; i < h->K; i++) wv.t[i] = View<Fp>{h->work[i], h->work_stride};
  TRY(ooc_stream(h, 16, [&](size_t c, size_t u0, size_t cnt, uint64_t* st, const OocChunks& cc) -> int {
    if (h->K == 1) {
      Views<Fp, 1> w1{};
      w1.t[0] = wv.t[0];
      hipLaunchKernelGGL(k_stream_fold<1>, dim3(cc.G), dim3(kBlock), 0, h->stream, staged_views<1>(st, 16, cc.B), cc.B,
                         u0, cnt, w1, q, h->d_cur, h->ooc_parts + c * cc.G * pass_values<1>());
    } else {
      hipLaunchKernelGGL(k_stream_fold<2>, dim3(cc.G), dim3(kBlock), 0, h->stream, staged_views<2>(st, 16, cc.B), cc.B,
                         u0, cnt, wv, q, h->d_cur, h->ooc_parts + c * cc.G * pass_values<2>());
    }
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }));
  *nblocks = (unsigned)(ch.nch * ch.G);
  return EFFSC_HIP_OK;
}

// Flags non-canonical words of a resident base table.
__global__ __launch_bounds__(kBlock) void k_check_canonical(const uint64_t* t, size_t n, unsigned* bad) {
  unsigned b = 0;
  for (size_t i = gtid(); i < n; i += (size_t)gridDim.x * blockDim.x) b |= t[i] >= gl::P;
  if (b) atomicOr(bad, 1u);
}

// sum_i prod_j s_j[i] over a chunk of `count` elements held AoS (D words) in
// the device staging of each stream; one partial per block.
template <class F>
__global__ __launch_bounds__(kBlock) void k_stream_claim(const uint64_t* const* st, int k, size_t count,
                                                         uint64_t w, uint64_t* partials, unsigned* bad) {
  F acc[1] = {gl::zero<F>()};
  unsigned b = 0;
  for (size_t i = gtid(); i < count; i += (size_t)gridDim.x * blockDim.x) {
    F p = gl::one<F>();
    for (int j = 0; j < k; j++) {
      uint64_t wd[3] = {0, 0, 0};
#pragma unroll
      for (int c = 0; c < F::D; c++) {
        wd[c] = st[j][i * F::D + c];
        b |= wd[c] >= gl::P;
      }
      p = gl::mul(p, gl::load_words<F>(wd), w);
    }
    acc[0] = gl::add(acc[0], p);
  }
  if (b) atomicOr(bad, 1u);
  finish<F, 1>(acc, Reduce{partials, nullptr});
}

}  // namespace

extern "C" {

int effsc_hip_stream_open(const char* path, int ext, effsc_hip_stream** out) {
  if (!path || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  *out = nullptr;
  if (ext != 1 && ext != 3) return fail(EFFSC_HIP_EINVAL, "ext must be 1 or 3");
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(EFFSC_HIP_EINVAL, std::string("cannot open ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return fail(EFFSC_HIP_EINVAL, "fstat failed");
  }
  const size_t bytes = (size_t)st.st_size, el = 8 * (size_t)ext;
  const size_t n = bytes / el;
  // FileStream::new asserts a power-of-two element count (file/core.rs:34-36)
  if (bytes % el || !is_pow2(n)) {
    close(fd);
    return fail(EFFSC_HIP_EINVAL, "a stream holds a power-of-two number of serialized elements");
  }
  void* m = mmap(nullptr, bytes, PROT_READ, MAP_SHARED, fd, 0);
  if (m == MAP_FAILED) {
    close(fd);
    return fail(EFFSC_HIP_EINVAL, "mmap failed");
  }
  auto* s = new effsc_hip_stream();
  s->fd = fd;
  s->map = static_cast<const uint8_t*>(m);
  s->bytes = bytes;
  s->ext = ext;
  s->n = n;
  while (((size_t)1 << s->v) < n) s->v++;
  *out = s;
  return EFFSC_HIP_OK;
}

int effsc_hip_stream_num_variables(const effsc_hip_stream* s, size_t* v) {
  if (!s || !v) return fail(EFFSC_HIP_EINVAL, "null argument");
  *v = s->v;
  return EFFSC_HIP_OK;
}

int effsc_hip_stream_read(const effsc_hip_stream* s, size_t first, size_t count, uint64_t* dst) {
  // FileStream::evaluation (file/core.rs:45-49, 66-68), `count` elements
  if (!s || (count && !dst)) return fail(EFFSC_HIP_EINVAL, "null argument");
  if (first > s->n || count > s->n - first) return fail(EFFSC_HIP_EINVAL, "read past the end of the stream");
  std::memcpy(dst, s->map + first * 8 * s->ext, count * 8 * s->ext);
  for (size_t i = 0; i < count * s->ext; i++)
    if (dst[i] >= gl::P) return fail(EFFSC_HIP_EINVAL, "non-canonical serialized element");
  return EFFSC_HIP_OK;
}

void effsc_hip_stream_close(effsc_hip_stream* s) {
  if (!s) return;
  if (s->map) munmap(const_cast<uint8_t*>(s->map), s->bytes);
  if (s->fd >= 0) close(s->fd);
  delete s;
}

int effsc_hip_stream_claim(int device, effsc_hip_stream* const* streams, int k, uint64_t w, uint64_t* out) {
  // multivariate_claim / multivariate_product_claim (src/streams/stream.rs:3-41)
  if (!streams || k < 1 || k > 8 || !out) return fail(EFFSC_HIP_EINVAL, "bad argument");
  const int ext = streams[0]->ext;
  const size_t n = streams[0]->n;
  for (int j = 0; j < k; j++) {
    if (!streams[j]) return fail(EFFSC_HIP_EINVAL, "null stream");
    if (streams[j]->n != n || streams[j]->ext != ext)
      return fail(EFFSC_HIP_EINVAL, "streams differ in length or field");  // stream.rs:26-30
  }
  HIPCHK(hipSetDevice(device));
  const size_t B = std::min(n, stream_chunk_units()), nch = (n + B - 1) / B;
  const unsigned G = (unsigned)std::min<size_t>((B + kBlock - 1) / kBlock, 1024);
  const size_t el = 8 * (size_t)ext;
  hipStream_t cs = nullptr, ks = nullptr;
  hipEvent_t copied[2] = {}, used[2] = {};
  uint8_t* pin[2] = {};
  uint64_t *dst[2] = {}, *parts = nullptr, *res = nullptr;
  const uint64_t** dptrs = nullptr;
  unsigned* bad = nullptr;
  int rc = EFFSC_HIP_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == EFFSC_HIP_OK) rc = fail(EFFSC_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return rc == EFFSC_HIP_OK;
  };
  chk(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
  chk(hipStreamCreateWithFlags(&ks, hipStreamNonBlocking), "stream");
  for (int b = 0; b < 2 && rc == EFFSC_HIP_OK; b++) {
    chk(hipHostMalloc((void**)&pin[b], k * B * el, hipHostMallocDefault), "pinned staging");
    chk(hipMalloc((void**)&dst[b], k * B * el), "device staging");
    chk(hipEventCreateWithFlags(&copied[b], hipEventDisableTiming), "event");
    chk(hipEventCreateWithFlags(&used[b], hipEventDisableTiming), "event");
  }
  if (rc == EFFSC_HIP_OK) chk(hipMalloc((void**)&parts, nch * G * 3 * 8), "partials");
  if (rc == EFFSC_HIP_OK) chk(hipMalloc((void**)&res, 3 * 8), "result");
  if (rc == EFFSC_HIP_OK) chk(hipMalloc((void**)&dptrs, 2 * 8 * sizeof(uint64_t*)), "pointers");
  if (rc == EFFSC_HIP_OK) chk(hipMalloc((void**)&bad, 4), "flag");
  if (rc == EFFSC_HIP_OK) chk(hipMemsetAsync(bad, 0, 4, ks), "flag");
  if (rc == EFFSC_HIP_OK) {
    std::vector<const uint64_t*> hp(16);
    for (int b = 0; b < 2; b++)
      for (int j = 0; j < k; j++) hp[b * 8 + j] = dst[b] + (size_t)j * B * ext;
    chk(hipMemcpy(dptrs, hp.data(), 16 * sizeof(uint64_t*), hipMemcpyHostToDevice), "pointers");
  }
  for (size_t c = 0; c < nch && rc == EFFSC_HIP_OK; c++) {
    const int b = (int)(c & 1);
    const size_t first = c * B, cnt = std::min(B, n - first);
    if (c >= 2 && !chk(hipEventSynchronize(used[b]), "wait")) break;  // staging b is free again
    std::vector<CopySeg> segs;
    for (int j = 0; j < k; j++) segs.push_back({pin[b] + (size_t)j * B * el, streams[j]->map + first * el, cnt * el});
    par_copy_segs(segs);
    for (int j = 0; j < k && rc == EFFSC_HIP_OK; j++)
      chk(hipMemcpyAsync(dst[b] + (size_t)j * B * ext, pin[b] + (size_t)j * B * el, cnt * el, hipMemcpyHostToDevice,
                         cs), "copy");
    chk(hipEventRecord(copied[b], cs), "event");
    chk(hipStreamWaitEvent(ks, copied[b], 0), "wait");
    if (ext == 1)
      hipLaunchKernelGGL(k_stream_claim<Fp>, dim3(G), dim3(kBlock), 0, ks, dptrs + b * 8, k, cnt, w,
                         parts + c * G, bad);
    else
      hipLaunchKernelGGL(k_stream_claim<Fp3>, dim3(G), dim3(kBlock), 0, ks, dptrs + b * 8, k, cnt, w,
                         parts + c * G * 3, bad);
    chk(hipGetLastError(), "claim kernel");
    chk(hipEventRecord(used[b], ks), "event");
  }
  if (rc == EFFSC_HIP_OK) {
    if (ext == 1)
      hipLaunchKernelGGL((k_reduce<Fp, 1>), dim3(1), dim3(1024), 0, ks, Reduce{parts, res}, (unsigned)(nch * G));
    else
      hipLaunchKernelGGL((k_reduce<Fp3, 1>), dim3(1), dim3(1024), 0, ks, Reduce{parts, res}, (unsigned)(nch * G));
    chk(hipGetLastError(), "reduce");
    uint64_t hres[4] = {0, 0, 0, 0};
    unsigned hbad = 0;
    chk(hipMemcpyAsync(hres, res, 8 * ext, hipMemcpyDeviceToHost, ks), "copy back");
    chk(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, ks), "copy back");
    chk(hipStreamSynchronize(ks), "sync");
    if (rc == EFFSC_HIP_OK && hbad) rc = fail(EFFSC_HIP_EINVAL, "non-canonical serialized element");
    if (rc == EFFSC_HIP_OK) std::copy(hres, hres + ext, out);
  }
  if (ks) (void)hipStreamSynchronize(ks);
  if (cs) (void)hipStreamSynchronize(cs);
  for (int b = 0; b < 2; b++) {
    if (pin[b]) (void)hipHostFree(pin[b]);
    if (dst[b]) (void)hipFree(dst[b]);
    if (copied[b]) (void)hipEventDestroy(copied[b]);
    if (used[b]) (void)hipEventDestroy(used[b]);
  }
  for (void* p : {(void*)parts, (void*)res, (void*)dptrs, (void*)bad})
    if (p) (void)hipFree(p);
  if (cs) (void)hipStreamDestroy(cs);
  if (ks) (void)hipStreamDestroy(ks);
  return rc;
}

int effsc_hip_stream_prover_new_ex(const effsc_hip_config* cfg, int kind, effsc_hip_stream* const* streams, int mode,
                                   effsc_hip_prover** out) {
  if (mode != EFFSC_HIP_STREAM_RESIDENT && mode != EFFSC_HIP_STREAM_OUT_OF_CORE && mode != EFFSC_HIP_STREAM_AUTO)
    return fail(EFFSC_HIP_EINVAL, "stream prover mode: resident (0), out of core (1) or auto (2)");
  if (mode == EFFSC_HIP_STREAM_RESIDENT) return effsc_hip_stream_prover_new(cfg, kind, streams, out);
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (kind != K_ML && kind != K_IP) return fail(EFFSC_HIP_EINVAL, "stream provers: multilinear (1) or inner product (2)");
  if (cfg->input_ext != 1 || cfg->ext != 1)
    return fail(EFFSC_HIP_EINVAL, "out-of-core stream provers: base-field streams and challenges");
  const int K = kind;
  if (!streams) return fail(EFFSC_HIP_EINVAL, "null streams");
  for (int i = 0; i < K; i++)
    if (!streams[i] || streams[i]->ext != 1) return fail(EFFSC_HIP_EINVAL, "stream provers read base-field streams");
  const size_t n = streams[0]->n;
  if (K == 2 && streams[1]->n != n) return fail(EFFSC_HIP_EINVAL, "a and b must have equal length");
  HIPCHK(hipSetDevice(cfg->device));
  if (mode == EFFSC_HIP_STREAM_AUTO) {
    // resident when the input, its level-1 work tables and a margin fit
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)K * (double)n * 8.0 * 1.5 + (double)(1 << 30);
    if (need < (double)free_b) return effsc_hip_stream_prover_new(cfg, kind, streams, out);
  }
  auto* h = new effsc_hip_prover();
  h->kind = kind;
  h->ext = 1;
  h->in_ext = 1;
  h->K = K;
  h->deg = K;
  h->repr = cfg->repr;
  h->dev = cfg->device;
  h->n = n;
  h->len = n;
  h->twostep = !std::getenv("EFFSC_HIP_ONE_ROUND");
  while (((size_t)1 << h->v) < n) h->v++;
  h->ooc = true;
  h->own_in = false;
  for (int i = 0; i < K; i++) h->ooc_src[i] = streams[i];
  int rc = EFFSC_HIP_OK;
  // the two streamed passes end at level n/4, which must still be above the
  // one-workgroup tail
  if (!h->twostep || n / 4 <= tail_max(h) || n < 64)
    rc = fail(EFFSC_HIP_EINVAL, "out-of-core stream provers need the two-round passes and n/4 above the tail size");
  if (rc == EFFSC_HIP_OK) rc = setup_common(h, n / 4);
  if (rc == EFFSC_HIP_OK) rc = alloc_dev(&h->d_pass0, kMaxK + kMaxK * (kMaxK + 1));
  if (rc == EFFSC_HIP_OK) {
    const OocChunks ch = ooc_chunks(n / 16);
    h->ooc_nparts = ch.nch * ch.G;
    rc = alloc_dev(&h->ooc_parts, h->ooc_nparts * (kMaxK + kMaxK * (kMaxK + 1)));
  }
  if (rc == EFFSC_HIP_OK) rc = ooc_pass0(h);
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  h->pass0_ready = true;
  *out = h;
  return EFFSC_HIP_OK;
}

int effsc_hip_stream_prover_new(const effsc_hip_config* cfg, int kind, effsc_hip_stream* const* streams,
                                effsc_hip_prover** out) {
  // MultilinearProver / InnerProductProver over tables read from streams
  if (!out) return fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  TRY(check_cfg(cfg));
  if (kind != K_ML && kind != K_IP) return fail(EFFSC_HIP_EINVAL, "stream provers: multilinear (1) or inner product (2)");
  if (cfg->input_ext != 1) return fail(EFFSC_HIP_EINVAL, "stream provers read base-field streams");
  const int K = kind;
  if (!streams) return fail(EFFSC_HIP_EINVAL, "null streams");
  for (int i = 0; i < K; i++)
    if (!streams[i] || streams[i]->ext != 1) return fail(EFFSC_HIP_EINVAL, "stream provers read base-field streams");
  const size_t n = streams[0]->n;
  if (K == 2 && streams[1]->n != n) return fail(EFFSC_HIP_EINVAL, "a and b must have equal length");
  auto* h = new effsc_hip_prover();
  h->kind = kind;
  h->ext = cfg->ext;
  h->in_ext = 1;
  h->K = K;
  h->deg = K;
  h->repr = cfg->repr;
  h->dev = cfg->device;
  h->w = cfg->ext == 3 ? gl::canon(cfg->ext_w) : 0;
  h->n = n;
  h->len = n;
  h->twostep = h->ext == 1 && !std::getenv("EFFSC_HIP_ONE_ROUND");
  int rc = setup_common(h, std::max<size_t>(n, 1) >> 1);
  h->own_in = true;
  h->in_stride = round_up(std::max<size_t>(n, 1), 64);
  for (int i = 0; i < K && rc == EFFSC_HIP_OK; i++) rc = alloc_dev(&h->in[i], h->in_stride);
  unsigned* bad = nullptr;
  if (rc == EFFSC_HIP_OK) rc = alloc_dev(reinterpret_cast<uint64_t**>(&bad), 1);
  if (rc == EFFSC_HIP_OK) rc = alloc_dev(&h->d_pass0, kMaxK + kMaxK * (kMaxK + 1));
  hipStream_t cs = nullptr;
  hipEvent_t copied[2] = {};
  uint8_t* pin[2] = {};
  uint64_t* parts = nullptr;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == EFFSC_HIP_OK) rc = fail(EFFSC_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return rc == EFFSC_HIP_OK;
  };
  if (rc == EFFSC_HIP_OK) chk(hipMemsetAsync(bad, 0, 4, h->stream), "flag");
  const size_t q = n / 4;
  const bool pass = h->twostep && n >= 16;
  if (rc == EFFSC_HIP_OK && !pass) {
    // small tables: one copy, no overlap worth having
    for (int i = 0; i < K && rc == EFFSC_HIP_OK; i++) {
      chk(hipMemcpy(h->in[i], streams[i]->map, n * 8, hipMemcpyHostToDevice), "copy");
      hipLaunchKernelGGL(k_check_canonical, dim3(blocks_for(h, n)), dim3(kBlock), 0, h->stream, h->in[i], n, bad);
      chk(hipGetLastError(), "check");
    }
  } else if (rc == EFFSC_HIP_OK) {
    // chunk c moves the units [c B, (c + 1) B): four quarter slices of each
    // table, so the pass over those units can run as soon as they land
    const size_t B = std::min(q, stream_chunk_units()), nch = (q + B - 1) / B;
    const unsigned G = (unsigned)std::min<size_t>((B + kBlock - 1) / kBlock, 1024);
    constexpr int NVmax = kMaxK + kMaxK * (kMaxK + 1);
    chk(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
    for (int b = 0; b < 2 && rc == EFFSC_HIP_OK; b++) {
      chk(hipHostMalloc((void**)&pin[b], (size_t)K * 4 * B * 8, hipHostMallocDefault), "pinned staging");
      chk(hipEventCreateWithFlags(&copied[b], hipEventDisableTiming), "event");
    }
    if (rc == EFFSC_HIP_OK) chk(hipMalloc((void**)&parts, nch * G * NVmax * 8), "partials");
    Views<Fp, 2> tv{};
    for (int i = 0; i < K; i++) tv.t[i] = View<Fp>{h->in[i], h->in_stride};
    for (size_t c = 0; c < nch && rc == EFFSC_HIP_OK; c++) {
      const int b = (int)(c & 1);
      const size_t u0 = c * B, cnt = std::min(B, q - u0);
      // pinned buffer b was last read by the copy of chunk c - 2
      if (c >= 2 && !chk(hipEventSynchronize(copied[b]), "wait")) break;
      std::vector<CopySeg> segs;
      for (int i = 0; i < K; i++)
        for (int j = 0; j < 4; j++)
          segs.push_back({pin[b] + ((size_t)i * 4 + j) * B * 8, streams[i]->map + (j * q + u0) * 8, cnt * 8});
      par_copy_segs(segs);
      for (int i = 0; i < K && rc == EFFSC_HIP_OK; i++)
        for (int j = 0; j < 4 && rc == EFFSC_HIP_OK; j++)
          chk(hipMemcpyAsync(h->in[i] + j * q + u0, pin[b] + ((size_t)i * 4 + j) * B * 8, cnt * 8,
                             hipMemcpyHostToDevice, cs), "copy");
      chk(hipEventRecord(copied[b], cs), "event");
      chk(hipStreamWaitEvent(h->stream, copied[b], 0), "wait");
      if (K == 1) {
        Views<Fp, 1> t1{};
        t1.t[0] = tv.t[0];
        hipLaunchKernelGGL(k_stream_pass<1>, dim3(G), dim3(kBlock), 0, h->stream, t1, q, (size_t)0, u0, cnt,
                           parts + c * G * pass_values<1>(), bad);
      } else {
        hipLaunchKernelGGL(k_stream_pass<2>, dim3(G), dim3(kBlock), 0, h->stream, tv, q, (size_t)0, u0, cnt,
                           parts + c * G * pass_values<2>(), bad);
      }
      chk(hipGetLastError(), "stream pass");
    }
    if (rc == EFFSC_HIP_OK) {
      if (K == 1)
        hipLaunchKernelGGL((k_reduce<Fp, pass_values<1>()>), dim3(1), dim3(1024), 0, h->stream,
                           Reduce{parts, h->d_pass0}, (unsigned)(nch * G));
      else
        hipLaunchKernelGGL((k_reduce<Fp, pass_values<2>()>), dim3(1), dim3(1024), 0, h->stream,
                           Reduce{parts, h->d_pass0}, (unsigned)(nch * G));
      chk(hipGetLastError(), "reduce");
    }
  }
  unsigned hbad = 0;
  if (rc == EFFSC_HIP_OK) chk(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, h->stream), "copy back");
  if (rc == EFFSC_HIP_OK) chk(hipStreamSynchronize(h->stream), "sync");
  if (cs) (void)hipStreamSynchronize(cs);
  if (rc == EFFSC_HIP_OK && hbad) rc = fail(EFFSC_HIP_EINVAL, "non-canonical serialized element");
  for (int b = 0; b < 2; b++) {
    if (pin[b]) (void)hipHostFree(pin[b]);
    if (copied[b]) (void)hipEventDestroy(copied[b]);
  }
  if (parts) (void)hipFree(parts);
  if (bad) (void)hipFree(bad);
  if (cs) (void)hipStreamDestroy(cs);
  if (rc != EFFSC_HIP_OK) {
    destroy(h);
    return rc;
  }
  h->pass0_ready = pass;
  *out = h;
  return EFFSC_HIP_OK;
}

}  // extern "C"

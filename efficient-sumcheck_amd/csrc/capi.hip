// C ABI of the gfx950 sumcheck prover: prover handles, round dispatch, the
// host protocol loop (runner, verifier, transcripts) and utilities.
// Declarations and reference citations: include/effsc_hip.h.  The
// EqFactored, GKR, device-loop and stream sections live in eq.hip, gkr.hip,
// devloop.hip and stream.hip; the round launchers (Ops, DevOps) in ops.hpp.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

namespace effsc_hip {
thread_local std::string g_err;
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* last_error() { return g_err.c_str(); }

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

namespace effsc_hip {
namespace pv {

int fail(int code, const std::string& msg) { return effsc_hip::set_error(code, msg); }

void ext_mul(int ext, uint64_t w, const uint64_t* a, const uint64_t* b, uint64_t* out) {
  uint64_t r[3] = {0, 0, 0};
  if (ext == 1) {
    r[0] = gl::mul(a[0], b[0]);
  } else if (ext == 2) {
    gl::store_words(gl::mul(gl::load_words<Fp2>(a), gl::load_words<Fp2>(b), w), r);
  } else {
    gl::store_words(gl::mul(gl::load_words<Fp3>(a), gl::load_words<Fp3>(b), w), r);
  }
  std::copy(r, r + ext, out);
}

void ext_mul3(int ext, uint64_t w, const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* out) {
  uint64_t ab[3];
  ext_mul(ext, w, a, b, ab);
  ext_mul(ext, w, ab, c, out);
}

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
  if (h->ext == 2) return h->K >= 3 ? (size_t)1 << 11 : (size_t)1 << 13;
  // base field, K <= 2: the pipelined tail takes up to 4096 (K = 2) or 8192
  // (K = 1) elements, the register tail 2^14
  if (h->K <= 2 && !std::getenv("EFFSC_HIP_TAIL_REG") && !std::getenv("EFFSC_HIP_TAIL_PLAIN"))
    return h->K == 1 ? (size_t)8192 : (size_t)4096;
  return (size_t)1 << 14;
}

const DevFs* fs_in(effsc_hip_prover* h) {
  const DevFs* t = h->fs_pending ? h->fs_pending : h->d_fs;
  h->fs_pending = nullptr;
  return t;
}

int fs_flush(effsc_hip_prover* h) {
  if (h->fs_pending) {
    // from the pinned staging or from d_fs_alt
    HIPCHK(hipMemcpyAsync(h->d_fs, h->fs_pending, sizeof(DevFs), hipMemcpyDefault, h->stream));
    h->fs_pending = nullptr;
  }
  return EFFSC_HIP_OK;
}

bool pipe_tail_ok(const effsc_hip_prover* h, size_t L0, size_t rounds) {
  if (h->ext != 1 || h->K > 2 || std::getenv("EFFSC_HIP_TAIL_REG") || std::getenv("EFFSC_HIP_TAIL_PLAIN")) return false;
  const size_t lmax = h->K == 1 ? (size_t)pipe_max<1>() : (size_t)pipe_max<2>();
  if (L0 < 4 || L0 > lmax || (L0 & (L0 - 1)) || rounds < 1) return false;
  size_t lg = 0;
  while (((size_t)1 << lg) < L0) lg++;
  return rounds <= lg;
}

int launch_tail_pipe(effsc_hip_prover* h, uint64_t* const* src, size_t sst, size_t L0, int pend, int rounds,
                     size_t first_round, const uint64_t* cur) {
  auto go = [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    const Views<Fp, K> in = views<Fp, K>(src, sst), work = views<Fp, K>(h->work, h->work_stride);
    // the tail also copies the proof block into a pinned staging
    // (EFFSC_HIP_PROOF_COPY=1: a separate copy after the tail)
    uint64_t* ph = h->h_proof_pinned && !std::getenv("EFFSC_HIP_PROOF_COPY") ? h->h_proof : nullptr;
    hipLaunchKernelGGL((k_tail_pipe<K>), dim3(1), dim3(kPipeThreads), 0, h->stream, in, work, (int)L0, pend, rounds,
                       fs_in(h), h->d_fs, h->d_polys + first_round * K, h->d_chals + first_round, cur, h->d_fin,
                       nullptr, nullptr, 0u, ph, h->d_polys, (int)kProofWords);
    HIPCHK(hipGetLastError());
    h->proof_by_tail = ph != nullptr;
    return EFFSC_HIP_OK;
  };
  if (h->K == 1) return go(std::integral_constant<int, 1>{});
  return go(std::integral_constant<int, 2>{});
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
    if (h->kind == K_ML)
      std::copy(el, el + E, vals);
    else
      ext_mul(E, h->w, el, el + E, vals);
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

// ------------------------------------------------ native pair-split rounds
// MultilinearProverLSB over a table of any length (src/provers/
// multilinear_lsb.rs:59-126): the round value is s0 = sum of the
// even-indexed elements (an unpaired last element included), the fold is
// new[k] = f[2k] + w (f[2k+1] - f[2k]) for k < n/2 (an odd trailing element
// is dropped).  Out of place: thread k reads 2k, 2k+1 and writes k.
template <class FI, class FO>
__global__ __launch_bounds__(kBlock) void k_lsb_fold(View<FI> in, View<FO> out, FO c, size_t half, uint64_t w) {
  for (size_t k = gtid(); k < half; k += (size_t)gridDim.x * blockDim.x)
    out.st(k, fold1<FO>(in.ld(2 * k), in.ld(2 * k + 1), c, w));
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_lsb_even_sum(View<F> t, size_t n, Reduce r) {
  F acc[1] = {gl::zero<F>()};
  for (size_t k = gtid(); 2 * k < n; k += (size_t)gridDim.x * blockDim.x) acc[0] = gl::add(acc[0], t.ld(2 * k));
  finish<F, 1>(acc, r);
}

int lsb_native_fold(effsc_hip_prover* h, const uint64_t* chal) {
  const size_t L = h->len;
  if (L <= 1) return EFFSC_HIP_OK;
  const size_t half = L / 2;
  const int fi = src_field(h);
  uint64_t* src = src_planes(h)[0];
  const size_t sst = src_stride(h);
  uint64_t* dst = h->in_work ? h->work_alt[0] : h->work[0];
  const unsigned g = blocks_for(h, half);
  TRY(dispatch(fi, h->ext, 1, [&](auto op) -> int {
    using FI = typename decltype(op)::In;
    using FO = typename decltype(op)::Out;
    hipLaunchKernelGGL((k_lsb_fold<FI, FO>), dim3(g), dim3(kBlock), 0, h->stream, View<FI>{src, sst},
                       View<FO>{dst, h->work_stride}, from_words<FO>(chal), half, h->w);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }));
  if (h->in_work) std::swap(h->work[0], h->work_alt[0]);
  h->len = half;
  h->in_work = true;
  return EFFSC_HIP_OK;
}

int lsb_native_round(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  if (chal) TRY(lsb_native_fold(h, chal));
  const int E = h->ext;
  std::fill(out, out + E, 0);
  if (h->len == 0) return EFFSC_HIP_OK;
  const int fi = src_field(h);
  const unsigned g = blocks_for(h, (h->len + 1) / 2);
  TRY(with_field(fi, [&](auto f) -> int {
    using F = decltype(f);
    hipLaunchKernelGGL(k_lsb_even_sum<F>, dim3(g), dim3(kBlock), 0, h->stream, View<F>{src_planes(h)[0], src_stride(h)},
                       h->len, h->red);
    HIPCHK(hipGetLastError());
    return reduce<F, 1>(h, g);
  }));
  return fetch_result(h, fi, out, 1);
}

int round_impl(effsc_hip_prover* h, const uint64_t* chal, uint64_t* out) {
  if (h->kind == K_EQ) return eq_round(h, chal, out);
  if (h->lsb_native) return lsb_native_round(h, chal, out);
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
  if (h->lsb_native) return lsb_native_fold(h, chal);
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
  std::copy(el, el + E, out);
  for (int i = 1; i < h->K; i++) ext_mul(E, h->w, out, el + (size_t)i * E, out);
  return EFFSC_HIP_OK;
}

// ----------------------------------------------------- word conversions

void words_in(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(src[i], gl::MONT_TO_CANON) : gl::canon(src[i]);
}
void words_out(const effsc_hip_prover* h, const uint64_t* src, uint64_t* dst, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = h->repr == EFFSC_HIP_MONT64 ? gl::mul(src[i], gl::CANON_TO_MONT) : src[i];
}

// ---------------------------------------------------------- construction
int check_cfg(const effsc_hip_config* cfg) {
  if (!cfg) return fail(EFFSC_HIP_EINVAL, "null config");
  if (cfg->ext < 1 || cfg->ext > 3) return fail(EFFSC_HIP_EINVAL, "ext must be 1, 2 or 3");
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
    if (h->work_alt[i]) (void)hipFree(h->work_alt[i]);
  }
  if (h->red.partials) (void)hipFree(h->red.partials);
  if (h->red.result) (void)hipFree(h->red.result);
  if (h->eq_l) (void)hipFree(h->eq_l);
  if (h->eq_r) (void)hipFree(h->eq_r);
  if (h->eq_point) (void)hipFree(h->eq_point);
  if (h->eq_full) (void)hipFree(h->eq_full);
  if (h->fuse_parts) (void)hipFree(h->fuse_parts);  // d_fs_alt: inside fuse_parts
  if (h->eq_l_spare) (void)hipFree(h->eq_l_spare);
  if (h->eq_l_alt) (void)hipFree(h->eq_l_alt);
  if (h->eq_r_spare) (void)hipFree(h->eq_r_spare);
  if (h->eq_rp) (void)hipFree(h->eq_rp);
  if (h->d_pass0) (void)hipFree(h->d_pass0);
  if (h->ooc_parts) (void)hipFree(h->ooc_parts);
  if (h->gkr_b) (void)hipFree(h->gkr_b);
  if (h->h_res) (void)hipHostFree(h->h_res);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  if (h->ev_proof) (void)hipEventDestroy(h->ev_proof);
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
  h->w = cfg->ext > 1 ? gl::canon(cfg->ext_w) : 0;
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

}  // namespace pv
}  // namespace effsc_hip

// An out-of-core prover's input is only read by whole device-resident proofs.
namespace effsc_hip {
namespace pv {
int ooc_guard(const effsc_hip_prover* h) {
  if (h->ooc && !h->in_work)
    return fail(EFFSC_HIP_ESTATE,
                "out-of-core stream prover: the input stays in the streams; run effsc_hip_sumcheck over all rounds "
                "without a hook first");
  return EFFSC_HIP_OK;
}
}  // namespace pv
}  // namespace effsc_hip

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

namespace effsc_hip {
namespace pv {

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
  effsc_hip_config c2 = *cfg;
  c2.flags &= ~EFFSC_HIP_BORROW_INPUT;  // the reordered tables are the prover's own
  effsc_hip_prover* h = nullptr;
  if (!is_pow2(n)) {
    // the reference's IP and coefficient LSB provers panic on such lengths;
    // MultilinearProverLSB drops the unpaired element at each fold
    if (kind != K_ML) return fail(EFFSC_HIP_EINVAL, "LSB inner-product and product provers take power-of-two tables");
    TRY(new_product_like(&c2, kind, K, deg, tables, n, &h));
    h->twostep = false;
    h->lsb_native = true;
    int rc = alloc_dev(&h->work_alt[0], h->work_stride * h->ext);
    if (rc != EFFSC_HIP_OK) {
      destroy(h);
      return rc;
    }
    *out = h;
    return EFFSC_HIP_OK;
  }
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

}  // namespace pv
}  // namespace effsc_hip

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
  return new_product_like(cfg, K_PRODUCT, k, k, tables, n, out);
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
  while (h->lsb && !h->lsb_native && ((size_t)1 << lv) < L) lv++;
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
    // the eq tables are folded in place: take the fresh ones a device-loop
    // proof built behind its proof words, or rebuild them from the point
    HIPCHK(hipSetDevice(h->dev));
    if (h->eq_spare_ready) {
      std::swap(h->eq_l, h->eq_l_spare);
      std::swap(h->eq_r, h->eq_r_spare);
      h->eq_l_len = (size_t)1 << h->v_l;
      h->eq_r_len = (size_t)1 << (h->v - h->v_l);
      h->eq_spare_ready = false;
    } else {
      TRY(build_eq(h, h->eq_point, 0, h->v_l, &h->eq_l, &h->eq_l_stride, &h->eq_l_len));
      TRY(build_eq(h, h->eq_point, h->v_l, h->v - h->v_l, &h->eq_r, &h->eq_r_stride, &h->eq_r_len));
    }
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

}  // extern "C"

namespace effsc_hip {
namespace pv {
// fold over AoS extension words: new[k] = v[k] + c (v[k + half] - v[k]) for
// k < paired, v[k] (1 - c) for paired <= k < half; in place (thread k reads k
// and k + half, writes k only).
template <class F>
__global__ __launch_bounds__(kBlock) void k_fold_aos(uint64_t* v, F c, size_t half, size_t paired, uint64_t w) {
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t k = gtid(); k < half; k += step) {
    const F lo = gl::load_words<F>(v + k * F::D);
    const F hi = k < paired ? gl::load_words<F>(v + (k + half) * F::D) : gl::zero<F>();
    gl::store_words(fold1<F>(lo, hi, c, w), v + k * F::D);
  }
}
}  // namespace pv
}  // namespace effsc_hip

extern "C" {

int effsc_hip_fold_ext(int device, int ext, uint64_t w, uint64_t* values, size_t n, const uint64_t* weight,
                       size_t* new_len) {
  if ((!values && n) || !weight) return fail(EFFSC_HIP_EINVAL, "null argument");
  if (ext < 1 || ext > 3) return fail(EFFSC_HIP_EINVAL, "ext must be 1, 2 or 3");
  HIPCHK(hipSetDevice(device));
  if (n <= 1) {
    if (new_len) *new_len = n;
    return EFFSC_HIP_OK;
  }
  const size_t half = next_pow2(n) >> 1, paired = n - half;
  const unsigned b = (unsigned)std::max<size_t>(1, std::min<size_t>((half + kBlock - 1) / kBlock, 8192));
  TRY(with_field(ext, [&](auto f) -> int {
    using F = decltype(f);
    hipLaunchKernelGGL(k_fold_aos<F>, dim3(b), dim3(kBlock), 0, nullptr, values, gl::load_words<F>(weight), half,
                       paired, ext > 1 ? gl::canon(w) : 0);
    HIPCHK(hipGetLastError());
    return EFFSC_HIP_OK;
  }));
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
namespace effsc_hip {
namespace pv {

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

}  // namespace pv
}  // namespace effsc_hip

extern "C" int effsc_hip_sumcheck_verify(int ext, uint64_t w, const uint64_t* claimed_sum, size_t degree,
                                         size_t num_rounds, effsc_hip_transcript* t, const uint64_t* proof,
                                         size_t proof_words, effsc_hip_hook hook, void* hook_ctx,
                                         uint64_t* challenges, uint64_t* final_claim, size_t* err_round,
                                         int* err_kind) {
  if (!claimed_sum || !t || !final_claim || (num_rounds && !challenges) || (proof_words && !proof))
    return fail(EFFSC_HIP_EINVAL, "null argument");
  if (err_kind) *err_kind = 0;
  return with_field(ext, [&](auto f) -> int {
    using F = decltype(f);
    return verify_t<F>(F::D == 1 ? 0 : gl::canon(w), claimed_sum, degree, num_rounds, t, proof, proof_words, hook,
                       hook_ctx, challenges, final_claim, err_round, err_kind);
  });
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


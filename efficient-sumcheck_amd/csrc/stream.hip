// Streamed input (src/streams/): FileStream mapping, on-device claims and
// provers fed from streams, resident and out of core.
#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;
using gl::Fp;
using gl::Fp2;
using gl::Fp3;

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

namespace effsc_hip {
namespace pv {

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
    // on failure fall through to the common tail: it waits for both streams
    // (copies may still be in flight) before freeing the buffers
    rc = consume(c, u0, cnt, stage[b], ch);
    if (rc != EFFSC_HIP_OK) break;
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

}  // namespace pv
}  // namespace effsc_hip

extern "C" {

int effsc_hip_stream_open(const char* path, int ext, effsc_hip_stream** out) {
  if (!path || !out) return fail(EFFSC_HIP_EINVAL, "null argument");
  *out = nullptr;
  if (ext < 1 || ext > 3) return fail(EFFSC_HIP_EINVAL, "ext must be 1, 2 or 3");
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
  if (!streams[0]) return fail(EFFSC_HIP_EINVAL, "null stream");
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
    (void)with_field(ext, [&](auto f) -> int {
      using F = decltype(f);
      hipLaunchKernelGGL(k_stream_claim<F>, dim3(G), dim3(kBlock), 0, ks, dptrs + b * 8, k, cnt, w,
                         parts + c * G * F::D, bad);
      return EFFSC_HIP_OK;
    });
    chk(hipGetLastError(), "claim kernel");
    chk(hipEventRecord(used[b], ks), "event");
  }
  if (rc == EFFSC_HIP_OK) {
    (void)with_field(ext, [&](auto f) -> int {
      using F = decltype(f);
      hipLaunchKernelGGL((k_reduce<F, 1>), dim3(1), dim3(1024), 0, ks, Reduce{parts, res}, (unsigned)(nch * G));
      return EFFSC_HIP_OK;
    });
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
  h->w = cfg->ext > 1 ? gl::canon(cfg->ext_w) : 0;
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

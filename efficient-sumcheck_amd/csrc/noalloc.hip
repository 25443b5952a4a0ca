// A whole multilinear or inner-product proof that allocates nothing: the
// device loop of effsc_hip_sumcheck (devloop.hip) run over caller-owned
// device tables, with every buffer it needs beyond the inputs and outputs
// (per-workgroup partials, reduced values, work tables, the proof block, the
// sponge state and the pending challenges) carved out of one file-scope
// device array.  For callers that cannot call hipMalloc (graph capture,
// serving loops, harnesses that forbid allocation).
#include <mutex>

#include "prover.hpp"

using namespace effsc_hip;
using namespace effsc_hip::pv;

namespace {

// Largest table this entry point accepts (elements per table).
constexpr size_t kNoAllocMaxN = (size_t)1 << 20;
// The two-round passes launch at most this many workgroups here.
constexpr int kNoAllocGridMax = 2048;
constexpr size_t kPartialWords = (size_t)kNoAllocGridMax * kPassGridScale * kMaxK * 3;
constexpr size_t kResultWords = 64;
constexpr size_t kBlockWords = round_up(kProofWords, 64);
constexpr size_t kCurWords = 64;
constexpr size_t kWorkStride = kNoAllocMaxN / 2;  // a multiple of 64
constexpr size_t kScratchWords = kPartialWords + kResultWords + kBlockWords + kCurWords + 2 * kWorkStride;

// The device scratch: kScratchWords words (about 9 MiB) on every device the
// library runs this entry point on, held until the process ends.  Host code
// uses only the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on `dev` (made current by the caller),
// looked up once per device.
int scratch_ptr(int dev, uint64_t** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(EFFSC_HIP_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(EFFSC_HIP_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_ptr[dev];
  return EFFSC_HIP_OK;
}

}  // namespace

extern "C" int effsc_hip_sumcheck_noalloc(int device, int kind, const uint64_t* a, const uint64_t* b, size_t n,
                                          effsc_hip_transcript* t, uint64_t* round_polys, uint64_t* challenges,
                                          uint64_t* final_value) {
  if (kind != 1 && kind != 2) return fail(EFFSC_HIP_EINVAL, "kind must be 1 (multilinear) or 2 (inner product)");
  if (!a || (kind == 2 && !b) || !t || !round_polys || !challenges || !final_value)
    return fail(EFFSC_HIP_EINVAL, "null argument");
  if (n < 2 || !is_pow2(n)) return fail(EFFSC_HIP_EINVAL, "tables must have a power-of-two length >= 2");
  if (n > kNoAllocMaxN)
    return fail(EFFSC_HIP_EINVAL, "tables longer than 2^20 elements exceed the static scratch of this entry point");
  if (reinterpret_cast<uintptr_t>(a) % 16 || (kind == 2 && reinterpret_cast<uintptr_t>(b) % 16))
    return fail(EFFSC_HIP_EINVAL, "tables must be 16-byte aligned");
  HIPCHK(hipSetDevice(device));
  uint64_t* s = nullptr;
  TRY(scratch_ptr(device, &s));

  // A prover over borrowed input tables whose buffers all lie in `scratch`;
  // host staging (the proof block copy, reduced values) lives on this stack.
  // Nothing needs clearing: within one call every scratch word is written
  // before it is read.
  uint64_t h_res[kMaxK * 3] = {};
  uint64_t h_proof[kBlockWords];
  effsc_hip_prover h;
  h.kind = kind == 1 ? K_ML : K_IP;
  h.K = kind;
  h.deg = kind;
  h.ext = h.in_ext = 1;
  h.repr = EFFSC_HIP_CANONICAL;
  h.dev = device;
  h.w = 0;
  h.n = h.len = n;
  h.own_in = false;
  h.in[0] = const_cast<uint64_t*>(a);
  if (kind == 2) h.in[1] = const_cast<uint64_t*>(b);
  h.in_stride = n;
  h.stream = nullptr;  // the device's null stream
  h.grid_max = std::min(grid_max_for(device), kNoAllocGridMax);
  h.red.partials = s;
  h.red.result = s + kPartialWords;
  h.d_polys = s + kPartialWords + kResultWords;
  h.d_chals = h.d_polys + kMaxRounds * kMaxK * 3;
  h.d_fin = h.d_chals + kMaxRounds * 3;
  h.d_fs = reinterpret_cast<DevFs*>(h.d_fin + 4);
  h.d_cur = h.d_polys + kBlockWords;
  h.work_stride = kWorkStride;
  h.work[0] = h.d_cur + kCurWords;
  h.work[1] = h.work[0] + kWorkStride;
  h.h_res = h_res;
  h.h_proof = h_proof;
  h.twostep = !std::getenv("EFFSC_HIP_ONE_ROUND");
  size_t v = 0;
  while (((size_t)1 << v) < n) v++;
  return prove_device(&h, v, t, round_polys, challenges, final_value);
}

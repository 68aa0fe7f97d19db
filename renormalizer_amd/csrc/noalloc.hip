// Allocation-free effective-Hamiltonian matvec: reno_hop_apply with its
// scratch (permuted MPO, the two stage intermediates, split-K partials) in a
// file-scope device array instead of a caller workspace, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops).  Every device the library runs on holds its own copy of the array
// until the process ends; its address on each device comes from
// hipGetSymbolAddress on the first call there and is used for every kernel
// argument afterwards.  All calls on one device share the array: calls that
// overlap on a device (two streams, two host threads) race on it and must be
// serialised by the caller.
#include <mutex>
#include <vector>

#include "contract.h"

namespace reno {

// 96 MiB: the FP64 one-site matvec at M_l = M_r = 256, d = 10, D = 3 needs
// 90 MiB (permuted MPO, two 31.5 MB stage intermediates, split-K partials)
constexpr size_t kScratchBytes = (size_t)96 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device, nullptr until resolved

// Device address of this device's copy of `scratch` (resolved once per
// device, with that device current).
static int scratch_address(void** out) {
  int dev = 0;
  RENO_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) {
    set_error("hop_apply_noalloc: invalid current device %d", dev);
    return RENO_EINVAL;
  }
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    RENO_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hop_apply_noalloc: no device address for the scratch array on device %d", dev);
      return RENO_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RENO_OK;
}

}  // namespace reno

using namespace reno;

extern "C" int reno_hop_apply_noalloc(const reno_hop_desc* op, const void* c_in, void* c_out,
                                      void* stream) {
  RENO_TRY(check_hop(op));
  if (!c_in || !c_out || c_in == c_out) {
    set_error("hop_apply_noalloc: null or aliased vectors");
    return RENO_EINVAL;
  }
  const HopLayout h = hop_layout(op);
  if (h.total > kScratchBytes) {
    set_error("hop_apply_noalloc: this operator needs %zu bytes of scratch, the static array "
              "holds %zu (use reno_hop_apply with a workspace)", h.total, kScratchBytes);
    return RENO_EINVAL;
  }
  void* ws = nullptr;
  RENO_TRY(scratch_address(&ws));
  hipStream_t s = (hipStream_t)stream;
  char* p = (char*)ws;
  void* wl = p;
  void* t1 = p + h.wl_bytes;
  void* t2 = p + h.wl_bytes + h.t1_bytes;
  void* part = p + h.wl_bytes + h.t1_bytes + h.t2_bytes;
  // the timing hook may allocate its marker array: keep it out of this call
  TimingHookSuspend no_hook;
  if (op->nsite == 1) RENO_TRY(permute_w(op->prec, op->W, wl, op->Dl, op->d, op->Dr, s));
  return hop_apply_core(op, wl, c_in, c_out, t1, t2, s, nullptr, part, h.part_bytes);
}

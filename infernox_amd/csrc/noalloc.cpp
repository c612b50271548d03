// F3 Forward window scores without any device allocation, for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// The sequence, the window descriptors and the scores live in one file-scope
// device array, `scratch`, of IXG_NOALLOC_BYTES; every GPU the library runs
// on holds its own copy of it until the process ends. Host code never uses
// the array's name as an address: at context set-up (ixg_init, with the
// device current) hipGetSymbolAddress gives this device's address of the
// array, which is kept per device and passed to the kernels.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; calls that overlap on one GPU (two host threads, or two contexts of
// the same device) race. Nothing in the array has to start at zero: the
// Forward kernels write the score of every window of the call.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>
#include "ctx.h"

static __device__ __align__(256) uint8_t scratch[IXG_NOALLOC_BYTES];

namespace ixg {

FwdDev p7_fwd_dev(const ixg_p7 *p);

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_ptr[kMaxDevices] = {};
}  // namespace

// Looks up this device's address of the array (device current); idempotent.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    g_err = "no-allocation scratch: device index out of range";
    return IXG_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return IXG_OK;
  void *ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));  // the array (ixg::scratch is the slot allocator)
  if (!ptr) {
    g_err = "no-allocation scratch: hipGetSymbolAddress returned a null address";
    return IXG_EDEVICE;
  }
  g_ptr[device] = (uint8_t *)ptr;
  return IXG_OK;
}

static uint8_t *noalloc_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace ixg

using namespace ixg;

extern "C" int ixg_fwd_scores_noalloc(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len,
                                      const int64_t *off, const int64_t *len, int64_t nwin, float *sc) {
  if (!c || !p || !seq || !off || !len || !sc || nwin < 0 || seq_len < 0) {
    g_err = "ixg_fwd_scores_noalloc: bad arguments";
    return IXG_EINVAL;
  }
  if (seq_len > IXG_NOALLOC_MAX_SEQ || nwin > IXG_NOALLOC_MAX_WINDOWS) {
    g_err = "ixg_fwd_scores_noalloc: sequence longer than IXG_NOALLOC_MAX_SEQ bytes or more than "
            "IXG_NOALLOC_MAX_WINDOWS windows";
    return IXG_ERANGE;
  }
  const FwdDev fd = p7_fwd_dev(p);
  if (!fd.tmain) {
    g_err = "ixg_fwd_scores_noalloc: host-only profile";
    return IXG_EINVAL;
  }
  for (int64_t i = 0; i < nwin; i++)
    if (off[i] < 0 || len[i] < 1 || off[i] + len[i] >= seq_len) {
      g_err = "window outside the sequence buffer";
      return IXG_EINVAL;
    }
  // residue codes index the 18-symbol tables (easel/alphabet.rs:46-157)
  for (int64_t i = 0; i < nwin; i++) {
    uint8_t acc = 0;
    for (int64_t k = 1; k <= len[i]; k++) acc |= (uint8_t)(seq[off[i] + k] >= KP);
    if (acc) {
      g_err = "digitized residue code >= 18 in a window";
      return IXG_EINVAL;
    }
  }
  if (nwin == 0) return IXG_OK;
  HIPCHK(hipSetDevice(c->device));
  uint8_t *base = noalloc_ptr(c->device);
  if (!base) {
    RCHK(noalloc_setup(c->device));
    base = noalloc_ptr(c->device);
  }
  // layout: [sequence | window descriptors | scores], each 256-byte aligned
  const size_t seq_bytes = ((size_t)seq_len + 255) & ~(size_t)255;
  const size_t win_bytes = (((size_t)nwin * sizeof(WinDev)) + 255) & ~(size_t)255;
  uint8_t *d_seq = base;
  WinDev *d_w = (WinDev *)(base + seq_bytes);
  float *d_sc = (float *)(base + seq_bytes + win_bytes);
  static_assert((size_t)IXG_NOALLOC_MAX_SEQ + 256 + (size_t)IXG_NOALLOC_MAX_WINDOWS * (sizeof(WinDev) + 4) + 512 <=
                    (size_t)IXG_NOALLOC_BYTES,
                "scratch array too small for its stated limits");
  // windows longest first (the launch order ixg_fwd_scores uses); idx keeps
  // each score at its window's position
  std::vector<WinDev> wd((size_t)nwin);
  for (int64_t i = 0; i < nwin; i++) {
    wd[i].off = off[i];
    wd[i].len = (int32_t)len[i];
    wd[i].idx = (int32_t)i;
    wd[i].a = wd[i].b = 0.f;
  }
  std::stable_sort(wd.begin(), wd.end(), [](const WinDev &a, const WinDev &b) { return a.len > b.len; });
  HIPCHK(hipMemcpyAsync(d_seq, seq, (size_t)seq_len, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_w, wd.data(), wd.size() * sizeof(WinDev), hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_fwd(fd, d_seq, d_w, nwin, d_sc, c->stream));
  HIPCHK(hipMemcpyAsync(sc, d_sc, (size_t)nwin * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return IXG_OK;
}

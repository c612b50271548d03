// Device context and scratch helpers shared by the C ABI translation units.
#pragma once
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>
#include "ixg_internal.h"

namespace ixg {

extern thread_local std::string g_err;

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::ixg::g_err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
      return e_ == hipErrorOutOfMemory ? IXG_ENOMEM : IXG_EDEVICE;                    \
    }                                                                                 \
  } while (0)
#define RCHK(expr)               \
  do {                           \
    int r_ = (expr);             \
    if (r_ != IXG_OK) return r_; \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
};

enum Slot { S_TASKS, S_ITEMS, S_MASK, S_RAW, S_COUNTER, S_WINS, S_SC0, S_SC1, S_SC2, S_OVF0, S_OVF1, S_SEQ,
            S_CM_ENV, S_CM_BANDS, S_CM_ALPHA, S_CM_SEQ, S_CINFO, S_CM_ROOT, S_CM_SCHED, S_SERVE_SEQ, S_NSLOTS };

int dmalloc(ixg_ctx *c, void **p, size_t n);
void dfree(ixg_ctx *c, void *p, size_t n);
int scratch(ixg_ctx *c, int slot, size_t n, void **out);
// Pinned host staging buffer `slot` of the context grown to at least n bytes
// (contents not preserved; a slot is reused only after the stream has
// drained past its previous copy).
enum PinSlot { P_PACK, P_UP, P_DOWN, P_SERVE, P_NSLOTS };
int64_t arena_elems(const ixg_ctx *c, int slot);
int pinned(ixg_ctx *c, int slot, size_t n, void **out);

// Runs f(k) for k in [0, n) on up to 16 host threads, in chunks of 4 items
// (serially when n is small). f must not touch g_err; return codes are the
// caller's business.
template <class F>
void par_for(int64_t n, F &&f) {
  const unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(16u, hw ? hw : 1u);
  if (n < 16) nt = 1;
  if (nt <= 1) {
    for (int64_t k = 0; k < n; k++) f(k);
    return;
  }
  std::atomic<int64_t> next(0);
  auto work = [&]() {
    for (;;) {
      const int64_t k0 = next.fetch_add(4);
      if (k0 >= n) break;
      for (int64_t k = k0; k < std::min<int64_t>(n, k0 + 4); k++) f(k);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
}

}  // namespace ixg

struct ixg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::vector<ixg::DevBuf> bufs;  // grow-only scratch slots
  void *hpin[ixg::P_NSLOTS] = {};   // pinned host staging (hipHostMalloc), grow-only
  size_t hpin_cap[ixg::P_NSLOTS] = {};
  hipEvent_t ev[8];
  double ms[5] = {0, 0, 0, 0, 0};
  int64_t cells[5] = {0, 0, 0, 0, 0};
  // per-kernel device time (HIP events around each launch), cumulative:
  // IXG_K_* index; kev is the event pair used for it
  hipEvent_t kev[2];
  hipEvent_t sev[2];  // scanner batch kernel: start/end
  double kms[IXG_NKERNELS] = {};
  int64_t kn[IXG_NKERNELS] = {};
  // candidate root cells per residue seen by the last QDB scans (sizes the
  // next call's candidate list so that it rarely has to be rerun)
  double qdb_cand_rate = 0.125, trq_cand_rate = 0.125;
};

namespace ixg {
// Records the start event of a timed launch on the context's stream.
inline hipError_t ktime_begin(ixg_ctx *c) { return hipEventRecord(c->kev[0], c->stream); }
// Records the end event, waits for it and adds the launch to kernel k.
inline hipError_t ktime_end(ixg_ctx *c, int k) {
  hipError_t e = hipEventRecord(c->kev[1], c->stream);
  if (e != hipSuccess) return e;
  if ((e = hipEventSynchronize(c->kev[1])) != hipSuccess) return e;
  float ms = 0;
  if ((e = hipEventElapsedTime(&ms, c->kev[0], c->kev[1])) != hipSuccess) return e;
  c->kms[k] += ms;
  c->kn[k] += 1;
  return hipSuccess;
}
}  // namespace ixg

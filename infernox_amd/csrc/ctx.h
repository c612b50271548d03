// Device context and scratch helpers shared by the C ABI translation units.
#pragma once
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <map>
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
            S_CM_ENV, S_CM_BANDS, S_CM_ALPHA, S_CM_SEQ, S_CINFO, S_CM_ROOT, S_CM_SCHED, S_SERVE_SEQ,
            S_GFB_ITEMS, S_GFB_X, S_GFB_M, S_GFB_SC, S_NSLOTS };

int dmalloc(ixg_ctx *c, void **p, size_t n);
void dfree(ixg_ctx *c, void *p, size_t n);
int scratch(ixg_ctx *c, int slot, size_t n, void **out);
// Pinned host staging buffer `slot` of the context grown to at least n bytes
// (contents not preserved). Growing waits for the context's stream first, so
// no copy still queued on it reads or writes the buffer being freed.
enum PinSlot { P_PACK, P_UP, P_DOWN, P_SERVE, P_WORD, P_NSLOTS };
int64_t arena_elems(const ixg_ctx *c, int slot);
// Device-to-host copy ordered after the work queued on the context's stream
// (a plain hipMemcpy runs on the null stream, which does not wait for a
// non-blocking stream); returns when the bytes are on the host.
int d2h(ixg_ctx *c, void *dst, const void *src, size_t n);
struct HostTask {
  int64_t base;  // offset of the task's dsq view (1-indexed) in the device sequence buffer
  int64_t L;
};
struct WinRef {
  int32_t task;
  int64_t start, end;  // chunk-local window
};
// The truncated-band fields of CP9Bands (cp9.rs:1471-1491): predicted
// start / end nodes, the R / L marginal i / j bands, per state the J, L, R, T
// candidacy (cp9_seq2bands_trunc's steps 3 and 4).
struct TrBand {
  float thresh1 = 0.01f, thresh2 = 0.98f;  // DEFAULT_CP9BANDS_THRESH1/2 (cp9.rs:1495)
  int32_t sp1 = -1, sp2 = -1, ep1 = -1, ep2 = -1;
  int32_t rmarg_imin = -1, rmarg_imax = -2, lmarg_jmin = -1, lmarg_jmax = -2;
  std::vector<uint8_t> jv, lv, rv, tv;
};
// A window of the envelope definition: its task (a sequence view), its
// coordinates in that view and its pass (IXG_PASS_*).
struct EnvWin {
  int32_t task;
  int64_t start, end;
  int32_t pass;
};
// Per-window F1b/F2/F3/F3b scores (kind 1 bias, 0 Forward, 2 Viterbi, 3
// MSV) of windows wl over tasks whose dsq views sit in d_seq.
int run_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
               const std::vector<WinRef> &wl, int kind, std::vector<float> &sc, std::vector<uint8_t> &ovf);
GmDev p7_gm_dev(const ixg_p7 *p);
const GmHost &p7_gm_host(const ixg_p7 *p);
// the truncated passes' profiles: 0 = 5', 1 = 3', 2 = 5'&3'
const GmHost &p7_gm_trunc(const ixg_p7 *p, int k);
bool p7_has_device(const ixg_p7 *p);
// Looks up the device address of the no-allocation scratch array for
// `device` (current) and keeps it; called by ixg_init.
int noalloc_setup(int device);
int pinned(ixg_ctx *c, int slot, size_t n, void **out);

}  // namespace ixg
bool ixg_target_view(const ixg_target *t, int64_t seq, int strand, int64_t *off, int64_t *L, const uint8_t **host);
int64_t ixg_target_nseq(const ixg_target *t);
namespace ixg {
// Runs f(k) for k in [0, n) on the calling thread and up to 15 pooled host
// workers, in chunks of 4 items (serially when n is small). f must not touch
// g_err; return codes are the caller's business. Calls from several threads
// at once share the pool.
void pool_run(int64_t n, void (*body)(void *, int64_t), void *arg);
template <class F>
void par_for(int64_t n, F &&f) {
  if (n < 16 || std::thread::hardware_concurrency() <= 1) {
    for (int64_t k = 0; k < n; k++) f(k);
    return;
  }
  pool_run(n, [](void *a, int64_t k) { (*static_cast<F *>(a))(k); }, &f);
}

// f(lo, hi) over [0, n) in blocks of `grain` indices on the host workers
// (for loops whose bodies are too small to hand out one index at a time)
template <class F>
void par_range(int64_t n, int64_t grain, F &&f) {
  const int64_t nb = (n + grain - 1) / grain;
  par_for(nb, [&](int64_t b) { f(b * grain, std::min(n, (b + 1) * grain)); });
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
  // per-kernel device time, cumulative (IXG_K_* index): each timed launch is
  // bracketed by a pair of events from kpool, recorded without waiting; the
  // pairs are read (after one wait on the last end event) when the pool is
  // full or the timings are queried, so timing adds no host round trip
  static constexpr int KT_POOL = 256;
  hipEvent_t kpool[2 * KT_POOL];
  int kpend_k[KT_POOL];
  int knext = 0;
  hipEvent_t sev[2];  // scanner batch kernel: start/end
  double kms[IXG_NKERNELS] = {};
  int64_t kn[IXG_NKERNELS] = {};
  // candidate root cells per residue seen by the last QDB scans (sizes the
  // next call's candidate list so that it rarely has to be rerun)
  double qdb_cand_rate = 0.125, trq_cand_rate = 0.125;
  // the last ixg_search_std call: host wall time per stage (IXG_SS_*) and
  // the HB cells its CYK and Inside launches computed
  double search_ms[9] = {};
  int64_t search_cells[2] = {};
  // launch-shape autotuning of the F3 Forward (ixg_set_autotune): 1 times
  // every applicable shape once per (stripe count, window-count class) and
  // keeps the fastest; 0 takes the fixed order; 2 + v forces shape v
  int autotune = 1;
  std::map<std::pair<int, int>, int> tuned;  // (stripe count, log2 windows) -> shape
  // auxiliary contexts (own stream and scratch) on which ixg_search runs the
  // truncated passes' CM stages concurrently with the STD pass's; created on
  // first use, freed with this context
  ixg_ctx *aux[3] = {};
};

namespace ixg {
// Reads every recorded event pair of the pool into kms/kn (one wait on the
// last end event) and empties the pool.
inline hipError_t ktime_flush(ixg_ctx *c) {
  if (c->knext == 0) return hipSuccess;
  hipError_t e = hipEventSynchronize(c->kpool[2 * c->knext - 1]);
  if (e != hipSuccess) return e;
  for (int t = 0; t < c->knext; t++) {
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, c->kpool[2 * t], c->kpool[2 * t + 1])) != hipSuccess) return e;
    c->kms[c->kpend_k[t]] += ms;
    c->kn[c->kpend_k[t]] += 1;
  }
  c->knext = 0;
  return hipSuccess;
}
// Records the start event of a timed launch on the context's stream.
inline hipError_t ktime_begin(ixg_ctx *c) {
  if (c->knext == ixg_ctx::KT_POOL) {
    hipError_t e = ktime_flush(c);
    if (e != hipSuccess) return e;
  }
  return hipEventRecord(c->kpool[2 * c->knext], c->stream);
}
// Records the end event of the launch begun last and files it under kernel k.
inline hipError_t ktime_end(ixg_ctx *c, int k) {
  hipError_t e = hipEventRecord(c->kpool[2 * c->knext + 1], c->stream);
  if (e != hipSuccess) return e;
  c->kpend_k[c->knext++] = k;
  return hipSuccess;
}
}  // namespace ixg

// C ABI implementation: device context, resident profiles and targets, and
// the host side of the filter cascade (window extension/merge/split, P-value
// gates, survivor merge, accounting), which mirrors cm_pipeline.rs line by
// line while every DP runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#include <chrono>
#include "ctx.h"

using namespace ixg;

namespace ixg {

thread_local std::string g_err;

int dmalloc(ixg_ctx *c, void **p, size_t n) {
  if (n == 0) n = 1;
  if (c->budget && c->in_use + n > c->budget) {
    g_err = "device memory budget exceeded";
    return IXG_ENOMEM;
  }
  HIPCHK(hipMalloc(p, n));
  c->in_use += n;
  return IXG_OK;
}

void dfree(ixg_ctx *c, void *p, size_t n) {
  if (!p) return;
  (void)hipFree(p);
  c->in_use -= std::min(c->in_use, n ? n : (size_t)1);
}

// Scratch slot `slot` grown to at least n bytes (contents not preserved).
int scratch(ixg_ctx *c, int slot, size_t n, void **out) {
  if ((int)c->bufs.size() <= slot) c->bufs.resize(slot + 1);
  DevBuf &b = c->bufs[slot];
  if (b.cap < n) {
    dfree(c, b.p, b.cap);
    b.p = nullptr;
    b.cap = 0;
    size_t want = std::max(n, (size_t)4096);
    want = std::max(want, n + n / 4);
    RCHK(dmalloc(c, &b.p, want));
    b.cap = want;
  }
  *out = b.p;
  return IXG_OK;
}

// 4-byte elements one launch group may put in scratch slot `slot`: 64 GiB
// (HBM is 288 GB, and a launch of fewer windows than the chip has CUs leaves
// most of it idle), or under a memory budget half of what the budget leaves
// (counting the slot's current allocation, which a regrow frees), at least
// 64 MiB.
int64_t arena_elems(const ixg_ctx *c, int slot) {
  const int64_t cap = (int64_t)1 << 34;
  if (!c->budget) return cap;
  const size_t held = (int)c->bufs.size() > slot ? c->bufs[slot].cap : 0;
  const int64_t left = (int64_t)c->budget - (int64_t)c->in_use + (int64_t)held;
  return std::max<int64_t>((int64_t)1 << 24, std::min<int64_t>(cap, left / 8));
}

int pinned(ixg_ctx *c, int slot, size_t n, void **out) {
  if (c->hpin_cap[slot] < n) {
    if (c->hpin[slot]) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipHostFree(c->hpin[slot]));
    }
    c->hpin[slot] = nullptr;
    c->hpin_cap[slot] = 0;
    const size_t want = std::max(n + n / 4, (size_t)1 << 20);
    HIPCHK(hipHostMalloc(&c->hpin[slot], want, hipHostMallocDefault));
    c->hpin_cap[slot] = want;
  }
  *out = c->hpin[slot];
  return IXG_OK;
}

int d2h(ixg_ctx *c, void *dst, const void *src, size_t n) {
  if (n == 0) return IXG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return IXG_OK;
}

template <typename T>
static int upload(ixg_ctx *c, int slot, const std::vector<T> &v, T **out) {
  void *p;
  RCHK(scratch(c, slot, v.size() * sizeof(T), &p));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  *out = (T *)p;
  return IXG_OK;
}

}  // namespace ixg

struct ixg_p7 {
  MsvHost msv;
  VitHost vit;
  FwdHost fwd;
  BiasHost bias;
  GmHost gm;                // glocal profile of the F4/F5 envelope definition
  GmHost gmt[3];            // the truncated passes' profiles (5', 3', 5'&3')
  float *d_gm = nullptr;    // tsc x 4 (gm, gmt[0..2]) | msc | isc | nats logsum table
  uint32_t *d_cost4 = nullptr;
  uint8_t *d_rbv = nullptr;
  float *d_fwd = nullptr;   // tmain | tddv | rfvv | tz
  int16_t *d_vit = nullptr; // rwv | tr(8*(m+1))
  size_t bytes = 0;
  double vit_lvmu = 0, vit_lvlambda = 0;
  SsvDev ssv_dev() const {
    SsvDev d;
    d.cost4 = d_cost4;
    d.rbv = d_rbv;
    d.m = msv.m;
    d.bias = msv.bias_b;
    d.base_b = msv.base_b;
    d.tjb_b = msv.tjb_b;
    d.tbm_b = msv.tbm_b;
    int tj = msv.tjb_b + msv.tbm_b;
    tj = tj > 255 ? 255 : tj;
    d.xb = (uint8_t)(msv.base_b > tj ? msv.base_b - tj : 0);
    d.scale_b = msv.scale_b;
    return d;
  }
  FwdDev fwd_dev() const {
    FwdDev d;
    d.tmain = d_fwd;
    d.tddv = d_fwd + fwd.tmain.size();
    d.rfvv = d_fwd + fwd.tmain.size() + fwd.tddv.size();
    d.tz = d_fwd + fwd.tmain.size() + fwd.tddv.size() + fwd.rfvv.size();
    d.m = fwd.m;
    d.nq = fwd.nq;
    return d;
  }
  GmDev gm_dev() const {
    GmDev d;
    const size_t M1 = (size_t)gm.M + 1;
    d.tsc = d_gm;
    d.msc = d_gm + 4 * 8 * M1;
    d.isc = d.msc + KP * M1;
    d.tbl = d.isc + KP * M1;
    d.m = gm.M;
    return d;
  }
  VitDev vit_dev() const {
    VitDev d;
    d.rwv = d_vit;
    d.tr = d_vit + vit.rwv.size();
    d.m = vit.m;
    d.base_w = vit.base_w;
    d.xw_e_move = vit.xw_e_move;
    d.xw_e_loop = vit.xw_e_loop;
    d.scale_w = vit.scale_w;
    return d;
  }
  BiasDev bias_dev() const {
    BiasDev d;
    memcpy(d.eo, bias.eo, sizeof(d.eo));
    memcpy(d.t1, bias.t1, sizeof(d.t1));
    memcpy(d.pi, bias.pi, sizeof(d.pi));
    return d;
  }
};
namespace ixg {
GmDev p7_gm_dev(const ixg_p7 *p) { return p->gm_dev(); }
const GmHost &p7_gm_host(const ixg_p7 *p) { return p->gm; }
const GmHost &p7_gm_trunc(const ixg_p7 *p, int k) { return p->gmt[k]; }
bool p7_has_device(const ixg_p7 *p) { return p->d_gm != nullptr; }
FwdDev p7_fwd_dev(const ixg_p7 *p) {
  if (!p->d_fwd) return FwdDev{};
  return p->fwd_dev();
}
}  // namespace ixg

// One sequence, or several back to back: per sequence s at soff[s] its
// [fwd (L_s+2) | revcomp (L_s+2)], sentinel-padded; L is sequence 0's length.
struct ixg_target {
  uint8_t *d_both = nullptr;
  int64_t L = 0;
  size_t bytes = 0;
  std::vector<uint8_t> h_both;  // host copy of the same buffer
  std::vector<int64_t> soff, sL;
};

const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L) {
  *L = t->L;
  return t->d_both;
}
const uint8_t *ixg_target_host_dsq(const ixg_target *t, int strand) {
  return t->h_both.data() + (strand ? t->L + 2 : 0);
}
// Strand `strand` of sequence `seq`: its offset in the device / host buffer
// and its length; false when seq is not one of the target's sequences.
bool ixg_target_view(const ixg_target *t, int64_t seq, int strand, int64_t *off, int64_t *L, const uint8_t **host) {
  if (seq < 0 || seq >= (int64_t)t->soff.size()) return false;
  *L = t->sL[seq];
  *off = t->soff[seq] + (strand ? t->sL[seq] + 2 : 0);
  if (host) *host = t->h_both.data() + *off;
  return true;
}
int64_t ixg_target_nseq(const ixg_target *t) { return (int64_t)t->soff.size(); }


extern "C" {

const char *ixg_last_error(void) { return g_err.c_str(); }

int ixg_init(int device, size_t mem_budget, ixg_ctx **out) {
  if (!out) return IXG_EINVAL;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    g_err = "no such device";
    return IXG_EINVAL;
  }
  HIPCHK(hipSetDevice(device));
  auto *c = new ixg_ctx();
  c->device = device;
  c->budget = mem_budget;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_err = hipGetErrorString(e);
    delete c;
    return IXG_EDEVICE;
  }
  for (auto &ev : c->ev) HIPCHK(hipEventCreate(&ev));
  for (auto &ev : c->kpool) HIPCHK(hipEventCreate(&ev));
  for (auto &ev : c->sev) HIPCHK(hipEventCreate(&ev));
  RCHK(noalloc_setup(device));
  *out = c;
  return IXG_OK;
}

int ixg_shutdown(ixg_ctx *c) {
  if (!c) return IXG_EINVAL;
  for (auto *a : c->aux)
    if (a) ixg_shutdown(a);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &b : c->bufs) dfree(c, b.p, b.cap);
  for (auto *h : c->hpin)
    if (h) (void)hipHostFree(h);
  for (auto &ev : c->ev) (void)hipEventDestroy(ev);
  for (auto &ev : c->kpool) (void)hipEventDestroy(ev);
  for (auto &ev : c->sev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return IXG_OK;
}

size_t ixg_mem_in_use(const ixg_ctx *c) { return c ? c->in_use : 0; }

int ixg_p7_create(ixg_ctx *c, const ixg_p7_hmm *h, int32_t max_length, ixg_p7 **out) {
  if (!h || !out || h->M < 1 || h->M > 1024 || !h->mat || !h->trans || max_length < 1) {
    g_err = "ixg_p7_create: bad arguments (1 <= M <= 1024 required)";
    return IXG_EINVAL;
  }
  auto p = std::make_unique<ixg_p7>();
  build_msv(*h, max_length, p->msv);
  build_vit(*h, p->vit);
  build_fwd(*h, p->fwd);
  build_bias(h->compo, h->M, p->bias);
  build_glocal(*h, p->gm);
  build_trunc_profiles(*h, p->gm, p->gmt);
  p->vit_lvmu = h->lvmu;
  p->vit_lvlambda = h->lvlambda;
  if (!c) {  // host tables only (inspection); no device copy
    *out = p.release();
    return IXG_OK;
  }
  HIPCHK(hipSetDevice(c->device));
  const int m = h->M;
  // cost words for k = 0..m, padded with cost 255 to a multiple of 4 (+4) so
  // the sweep reads whole groups of four positions
  std::vector<uint32_t> cost4((size_t)((m + 1 + 3) & ~3) + 4, 0xFFFFFFFFu);
  for (int k = 0; k <= m; k++) {
    uint32_t v = 0;
    for (int x = 0; x < 4; x++) v |= (uint32_t)p->msv.rbv[k * KP + x] << (8 * x);
    cost4[k] = v;
  }
  std::vector<float> fwd;
  fwd.insert(fwd.end(), p->fwd.tmain.begin(), p->fwd.tmain.end());
  fwd.insert(fwd.end(), p->fwd.tddv.begin(), p->fwd.tddv.end());
  fwd.insert(fwd.end(), p->fwd.rfvv.begin(), p->fwd.rfvv.end());
  // The striped transitions regrouped per stripe lane z ([4][nq][8]), then
  // the DD transitions alone per lane, zero-padded to a multiple of 8.
  const int nq = p->fwd.nq, nq8 = (nq + 7) & ~7;
  for (int z = 0; z < 4; z++)
    for (int q = 0; q < nq; q++) {
      for (int t = 0; t < 7; t++) fwd.push_back(p->fwd.tmain[q * 28 + t * 4 + z]);
      fwd.push_back(p->fwd.tddv[q * 4 + z]);
    }
  for (int z = 0; z < 4; z++)
    for (int q = 0; q < nq8; q++) fwd.push_back(q < nq ? p->fwd.tddv[q * 4 + z] : 0.f);
  std::vector<int16_t> vit(p->vit.rwv);
  for (auto *a : {&p->vit.tmm, &p->vit.tim, &p->vit.tdm, &p->vit.vbm, &p->vit.tmi, &p->vit.tii, &p->vit.tmd,
                  &p->vit.tdd})
    vit.insert(vit.end(), a->begin(), a->end());
  std::vector<float> gmv(p->gm.tsc);
  for (const GmHost &t : p->gmt) gmv.insert(gmv.end(), t.tsc.begin(), t.tsc.end());
  gmv.insert(gmv.end(), p->gm.msc.begin(), p->gm.msc.end());
  gmv.insert(gmv.end(), p->gm.isc.begin(), p->gm.isc.end());
  gmv.insert(gmv.end(), p7_logsum_table().begin(), p7_logsum_table().end());
  const size_t b4 = gmv.size() * 4;
  RCHK(dmalloc(c, (void **)&p->d_gm, b4));
  HIPCHK(hipMemcpy(p->d_gm, gmv.data(), b4, hipMemcpyHostToDevice));
  size_t b0 = cost4.size() * 4, b1 = p->msv.rbv.size(), b2 = fwd.size() * 4, b3 = vit.size() * 2;
  RCHK(dmalloc(c, (void **)&p->d_cost4, b0));
  RCHK(dmalloc(c, (void **)&p->d_rbv, b1));
  RCHK(dmalloc(c, (void **)&p->d_fwd, b2));
  RCHK(dmalloc(c, (void **)&p->d_vit, b3));
  p->bytes = b0 + b1 + b2 + b3 + b4;
  HIPCHK(hipMemcpy(p->d_cost4, cost4.data(), b0, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->d_rbv, p->msv.rbv.data(), b1, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->d_fwd, fwd.data(), b2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->d_vit, vit.data(), b3, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  *out = p.release();
  return IXG_OK;
}

int ixg_p7_free(ixg_ctx *c, ixg_p7 *p) {
  if (!p) return IXG_EINVAL;
  if (!c) {
    if (p->d_cost4) return IXG_EINVAL;  // device copy needs its context
    delete p;
    return IXG_OK;
  }
  dfree(c, p->d_cost4, (p->msv.m + 1) * 4);
  dfree(c, p->d_rbv, p->msv.rbv.size());
  dfree(c, p->d_fwd, (p->fwd.tmain.size() + p->fwd.tddv.size() + p->fwd.rfvv.size()) * 4);
  dfree(c, p->d_vit, (p->vit.rwv.size() + 8 * (p->vit.m + 1)) * 2);
  dfree(c, p->d_gm, (4 * p->gm.tsc.size() + p->gm.msc.size() + p->gm.isc.size() + GFB_TBL) * 4);
  delete p;
  return IXG_OK;
}

int ixg_p7_msv_tables(const ixg_p7 *p, uint8_t *rbv, uint8_t *bytes, float *prefix, float *suffix) {
  if (!p) return IXG_EINVAL;
  if (rbv) memcpy(rbv, p->msv.rbv.data(), p->msv.rbv.size());
  if (bytes) {
    memcpy(bytes, &p->msv.scale_b, 4);
    bytes[4] = p->msv.base_b;
    bytes[5] = p->msv.bias_b;
    bytes[6] = p->msv.tbm_b;
    bytes[7] = p->msv.tec_b;
    bytes[8] = p->msv.tjb_b;
  }
  if (prefix) memcpy(prefix, p->msv.prefix.data(), p->msv.prefix.size() * 4);
  if (suffix) memcpy(suffix, p->msv.suffix.data(), p->msv.suffix.size() * 4);
  return IXG_OK;
}

int ixg_p7_vit_tables(const ixg_p7 *p, int16_t *rwv, int16_t *tr8, int16_t *ddb) {
  if (!p) return IXG_EINVAL;
  if (rwv) memcpy(rwv, p->vit.rwv.data(), p->vit.rwv.size() * 2);
  if (tr8) {
    int16_t *o = tr8;
    for (auto *a : {&p->vit.tmm, &p->vit.tim, &p->vit.tdm, &p->vit.vbm, &p->vit.tmi, &p->vit.tii, &p->vit.tmd,
                    &p->vit.tdd}) {
      memcpy(o, a->data(), a->size() * 2);
      o += a->size();
    }
  }
  if (ddb) *ddb = p->vit.ddbound_w;
  return IXG_OK;
}

int ixg_p7_fwd_tables(const ixg_p7 *p, int32_t *nq, float *tmain, float *tddv, float *rfvv) {
  if (!p) return IXG_EINVAL;
  if (nq) *nq = p->fwd.nq;
  if (tmain) memcpy(tmain, p->fwd.tmain.data(), p->fwd.tmain.size() * 4);
  if (tddv) memcpy(tddv, p->fwd.tddv.data(), p->fwd.tddv.size() * 4);
  if (rfvv) memcpy(rfvv, p->fwd.rfvv.data(), p->fwd.rfvv.size() * 4);
  return IXG_OK;
}

static bool codes_ok_pub(const uint8_t *p, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (p[i] >= KP) {
      g_err = "digitized residue code >= 18 in the sequence";
      return false;
    }
  return true;
}

int ixg_target_create(ixg_ctx *c, const uint8_t *dsq, int64_t L, ixg_target **out) {
  if (!c || !dsq || L < 0 || !out) return IXG_EINVAL;
  if (!codes_ok_pub(dsq + 1, L)) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  auto t = std::make_unique<ixg_target>();
  t->L = L;
  t->soff = {0};
  t->sL = {L};
  t->bytes = (size_t)2 * (L + 2);
  RCHK(dmalloc(c, (void **)&t->d_both, t->bytes));
  t->h_both.assign(2 * (L + 2), 255);
  uint8_t *h = t->h_both.data();
  if (L) memcpy(h + 1, dsq + 1, L);
  for (int64_t j = 1; j <= L; j++) {  // reverse complement (cm_search.rs:1822-1827)
    const uint8_t d = dsq[L + 1 - j];
    h[L + 2 + j] = d < 4 ? (uint8_t)(3 - d) : d;
  }
  HIPCHK(hipMemcpyAsync(t->d_both, h, L + 2, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_revcomp(t->d_both, t->d_both + L + 2, L, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = t.release();
  return IXG_OK;
}

int ixg_target_create_multi(ixg_ctx *c, const uint8_t *const *dsq, const int64_t *L, int64_t nseq,
                            ixg_target **out) {
  if (!c || !dsq || !L || nseq < 1 || !out) {
    g_err = "ixg_target_create_multi: bad arguments";
    return IXG_EINVAL;
  }
  auto t = std::make_unique<ixg_target>();
  size_t tot = 0;
  for (int64_t s = 0; s < nseq; s++) {
    if (!dsq[s] || L[s] < 0 || !codes_ok_pub(dsq[s] + 1, L[s])) {
      g_err = "ixg_target_create_multi: bad sequence";
      return IXG_EINVAL;
    }
    t->soff.push_back((int64_t)tot);
    t->sL.push_back(L[s]);
    tot += (size_t)2 * (L[s] + 2);
  }
  t->L = L[0];
  t->bytes = tot;
  t->h_both.assign(tot, 255);
  par_for(nseq, [&](int64_t s) {
    uint8_t *h = t->h_both.data() + t->soff[s];
    const int64_t n = L[s];
    if (n) memcpy(h + 1, dsq[s] + 1, n);
    for (int64_t j = 1; j <= n; j++) {  // reverse complement (cm_search.rs:1822-1827)
      const uint8_t d = dsq[s][n + 1 - j];
      h[n + 2 + j] = d < 4 ? (uint8_t)(3 - d) : d;
    }
  });
  HIPCHK(hipSetDevice(c->device));
  RCHK(dmalloc(c, (void **)&t->d_both, t->bytes));
  HIPCHK(hipMemcpyAsync(t->d_both, t->h_both.data(), t->bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = t.release();
  return IXG_OK;
}

int ixg_target_free(ixg_ctx *c, ixg_target *t) {
  if (!c || !t) return IXG_EINVAL;
  dfree(c, t->d_both, t->bytes);
  delete t;
  return IXG_OK;
}

int64_t ixg_window_tasks(int64_t L, int64_t maxw, int do_top, int do_bot, ixg_task *out, int64_t cap) {
  // cm_search.rs:1774-1808 (CM_MAX_RESIDUE_COUNT = 100000, cm_pipeline.rs:858)
  int64_t n = 0, new_start = 1;
  bool first = true;
  while (new_start <= L) {
    int64_t ctx = first ? 0 : std::min(maxw, new_start - 1);
    int64_t wgs = new_start - ctx;
    int64_t new_count = std::min<int64_t>(100000, L - new_start + 1);
    int64_t wge = wgs + ctx + new_count - 1;
    for (int s = 0; s < 2; s++) {
      if ((s == 0 && !do_top) || (s == 1 && !do_bot)) continue;
      if (n < cap && out) {
        out[n].start = wgs;
        out[n].end = wge;
        out[n].strand = s;
        out[n].ctx = (int32_t)ctx;
      }
      n++;
    }
    new_start = wge + 1;
    first = false;
  }
  return n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cascade internals
// ---------------------------------------------------------------------------
namespace {


struct RawW {
  int64_t n, length, k;
  float score;
};
struct Win {
  int64_t start, end;
  float score;
};

// cm_pipeline.rs:722-793
std::vector<Win> extend_and_merge(const MsvHost &f, std::vector<RawW> raws, int64_t target_len) {
  std::vector<Win> out;
  if (raws.empty()) return out;
  const double maxl = (double)f.max_length;
  for (auto &w : raws) {
    int64_t ks = std::clamp<int64_t>(w.k - w.length + 1, 0, f.m);
    int64_t ke = std::clamp<int64_t>(w.k, 0, f.m);
    double pref = (double)f.prefix[ks], suff = (double)f.suffix[ke];
    double ws = std::max((double)w.n - maxl * (0.1 + pref), 1.0);
    double we = std::min((double)w.n + (double)w.length + maxl * (0.1 + suff), (double)target_len);
    int64_t s = (int64_t)ws, e = (int64_t)we;
    w.n = s;
    w.length = e - s + 1;
  }
  std::vector<RawW> kept;
  kept.reserve(raws.size());
  kept.push_back(raws[0]);
  for (size_t i = 1; i < raws.size(); i++) {
    const RawW cur = raws[i];
    RawW &prev = kept.back();
    int64_t ws = std::max(prev.n, cur.n);
    int64_t pe = prev.n + prev.length - 1, ce = cur.n + cur.length - 1;
    int64_t wl = std::min(pe, ce) - ws + 1;
    int64_t minlen = std::min(prev.length, cur.length);
    if ((float)wl / (float)minlen > 0.0f) {
      int64_t ms = std::min(prev.n, cur.n), me = std::max(pe, ce);
      prev.n = ms;
      prev.length = me - ms + 1;
    } else {
      kept.push_back(cur);
    }
  }
  for (auto &k : kept) out.push_back({k.n, k.n + k.length - 1, k.score});
  return out;
}

// cm_pipeline.rs:817-837
std::vector<Win> split_windows(const std::vector<Win> &in, int64_t cmw) {
  std::vector<Win> out;
  for (const auto &w : in) {
    int64_t wlen = w.end - w.start + 1;
    if (wlen > 2 * cmw) {
      int64_t nws = w.start, nwe = std::min(w.start + 2 * cmw - 1, w.end);
      out.push_back({nws, nwe, w.score});
      while (nwe < w.end) {
        nws = std::min(nws + cmw, w.end);
        nwe = std::min(nwe + cmw, w.end);
        out.push_back({nws, nwe, w.score});
      }
    } else {
      out.push_back(w);
    }
  }
  return out;
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Runs F1 for all tasks on the device; raws[t] gets task t's diagonals.
int run_f1(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks, double f1,
           std::vector<std::vector<RawW>> &raws) {
  const SsvDev sd = p->ssv_dev();
  const int thresh = ssv_thresh(p->msv, f1);
  std::vector<TaskDev> td(tasks.size());
  std::vector<SweepItem> items;
  int64_t mask_words = 0;
  for (size_t t = 0; t < tasks.size(); t++) {
    td[t].dsq = d_seq + tasks[t].base;
    td[t].L = tasks[t].L;
    td[t].mask_off = mask_words;
    mask_words += (tasks[t].L + 63) / 64;
    const int64_t nseg = (tasks[t].L + SWEEP_SEG - 1) / SWEEP_SEG;
    for (int64_t s = 0; s < nseg; s++) items.push_back({(int32_t)t, (int32_t)s});
  }
  TaskDev *d_tasks;
  SweepItem *d_items;
  RCHK(upload(c, S_TASKS, td, &d_tasks));
  RCHK(upload(c, S_ITEMS, items, &d_items));
  uint64_t *d_mask;
  RCHK(scratch(c, S_MASK, (size_t)std::max<int64_t>(mask_words, 1) * 8, (void **)&d_mask));
  uint32_t *d_cinfo;  // per flagged row: reset-free maximum << 16 | leftmost argmax
  RCHK(scratch(c, S_CINFO, (size_t)std::max<int64_t>(mask_words, 1) * 64 * 4, (void **)&d_cinfo));
  int64_t total_len = 0;
  for (auto &t : tasks) total_len += t.L;
  int64_t cap = total_len / 32 + 256 * (int64_t)tasks.size() + 1024;
  unsigned long long *d_counter;
  RCHK(scratch(c, S_COUNTER, 8, (void **)&d_counter));
  HIPCHK(hipEventRecord(c->ev[0], c->stream));
  HIPCHK(launch_ssv_sweep(sd, d_tasks, d_items, (int64_t)items.size(), d_mask, d_cinfo, thresh, c->stream));
  HIPCHK(hipEventRecord(c->ev[1], c->stream));
  RawWinDev *d_raw = nullptr;
  unsigned long long count = 0;
  unsigned long long *h_count;
  RCHK(pinned(c, P_WORD, 64, (void **)&h_count));
  for (int attempt = 0; attempt < 2; attempt++) {
    RCHK(scratch(c, S_RAW, (size_t)cap * sizeof(RawWinDev), (void **)&d_raw));
    HIPCHK(hipMemsetAsync(d_counter, 0, 8, c->stream));
    if (attempt) HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(hipEventRecord(c->ev[7], c->stream));
    HIPCHK(launch_ssv_resolve(sd, d_tasks, d_items, (int64_t)items.size(), d_mask, d_cinfo, thresh, d_raw, cap,
                              d_counter, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    HIPCHK(hipMemcpyAsync(h_count, d_counter, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    count = *h_count;
    if ((int64_t)count <= cap) break;
    cap = (int64_t)count + 1024;
  }
  RawWinDev *h = nullptr;
  RCHK(pinned(c, P_DOWN, std::max<size_t>(count, 1) * sizeof(RawWinDev), (void **)&h));
  if (count) {
    HIPCHK(hipMemcpyAsync(h, d_raw, count * sizeof(RawWinDev), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // records by task, then by trigger row within a task (blocks append in
  // trigger order, but blocks of one task interleave)
  // (a counting sort by task, then each task's records by (row, index) on
  // the host workers)
  const size_t ntk = tasks.size();
  raws.assign(ntk, {});
  std::vector<int64_t> tbeg(ntk + 1, 0);
  for (int64_t i = 0; i < (int64_t)count; i++) tbeg[h[i].task + 1]++;
  for (size_t t = 0; t < ntk; t++) tbeg[t + 1] += tbeg[t];
  std::vector<int32_t> ord(count);
  {
    std::vector<int64_t> at(tbeg.begin(), tbeg.end() - 1);
    for (int32_t i = 0; i < (int32_t)count; i++) ord[at[h[i].task]++] = i;
  }
  par_for((int64_t)ntk, [&](int64_t t) {
    int32_t *b = ord.data() + tbeg[t], *e = ord.data() + tbeg[t + 1];
    std::sort(b, e, [&](int32_t x, int32_t y) { return h[x].row != h[y].row ? h[x].row < h[y].row : x < y; });
    raws[t].reserve(e - b);
    for (const int32_t *q = b; q < e; q++) {
      const RawWinDev &r = h[*q];
      raws[t].push_back({r.n, r.length, r.k, r.score});
    }
  });
  float a = 0, b = 0, kr = 0;
  HIPCHK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
  HIPCHK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
  HIPCHK(hipEventElapsedTime(&kr, c->ev[7], c->ev[2]));
  c->ms[0] += a;
  c->ms[1] += b;
  c->kms[IXG_K_SSV_SWEEP] += a;
  c->kn[IXG_K_SSV_SWEEP] += 1;
  c->kms[IXG_K_SSV_RESOLVE] += kr;
  c->kn[IXG_K_SSV_RESOLVE] += 1;
  c->cells[0] += total_len * p->msv.m;
  return IXG_OK;
}


// The F3 Forward launch with the context's shape choice: forced, fixed, or
// autotuned: the first batch of a (stripe count, log2 window count) class of
// at least 256 windows runs every applicable shape twice, the second run
// timed with events on the stream (the first absorbs one-time launch costs),
// and the fastest is kept for the class; every shape writes the same scores.
static hipError_t fwd_launch_tuned(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const WinDev *d_w, int64_t n,
                                   float *d_sc) {
  const FwdDev fd = p->fwd_dev();
  if (c->autotune >= 2) return launch_fwd_variant(fd, d_seq, d_w, n, d_sc, c->autotune - 2, c->stream);
  if (c->autotune == 0 || n < 256) return launch_fwd(fd, d_seq, d_w, n, d_sc, c->stream);
  int cls = 0;
  while (((int64_t)2 << cls) <= n) cls++;
  const auto key = std::make_pair(fd.nq, cls);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) return launch_fwd_variant(fd, d_seq, d_w, n, d_sc, it->second, c->stream);
  int best = -1;
  float best_ms = 0;
  for (int v = 0; v < FWD_V_N; v++) {
    if (!fwd_variant_ok(fd, v)) continue;
    hipError_t e = launch_fwd_variant(fd, d_seq, d_w, n, d_sc, v, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[5], c->stream);
    if (e == hipSuccess) e = launch_fwd_variant(fd, d_seq, d_w, n, d_sc, v, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[6], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[6]);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[5], c->ev[6]);
    if (e != hipSuccess) return e;
    if (best < 0 || ms < best_ms) best = v, best_ms = ms;
  }
  if (best < 0) return hipErrorInvalidValue;
  c->tuned[key] = best;
  return hipSuccess;  // the last shape's scores are in d_sc, the same as the best's
}

// Scores for a list of windows: kind 0 fwd, 1 bias, 2 vit, 3 msv (F1b).
int scores_impl(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                const std::vector<WinRef> &wl, int kind, std::vector<float> &sc, std::vector<uint8_t> &ovf) {
  const int64_t n = (int64_t)wl.size();
  sc.assign(n, 0.f);
  ovf.assign(n, 0);
  if (n == 0) return IXG_OK;
  std::vector<WinDev> wd(n);
  par_range(n, 2048, [&](int64_t lo, int64_t hi) {
   for (int64_t i = lo; i < hi; i++) {
    const WinRef &w = wl[i];
    const int64_t len = w.end - w.start + 1;
    wd[i].off = tasks[w.task].base + w.start - 1;
    wd[i].len = (int32_t)len;
    wd[i].idx = (int32_t)i;
    wd[i].a = wd[i].b = 0.f;
    if (kind == 1) {  // bias tail: L*ln(p1) and ln(1-p1) in f32 (cm_pipeline.rs:1773)
      const float p1 = (float)len / ((float)len + 1.0f);
      wd[i].a = (float)len * std::log(p1);
      wd[i].b = std::log(1.0f - p1);
    } else if (kind == 2) {  // xw_move (p7_vitfilter.rs:311-312)
      const float pmove = 3.0f / ((float)len + 3.0f);
      wd[i].a = (float)wordify(p->vit.scale_w, std::log(pmove));
    } else if (kind == 3) {  // reconfigured tjb_b (cm_pipeline.rs:427) and tec_b
      wd[i].a = (float)unbiased_byteify(p->msv.scale_b, (float)std::log(3.0 / ((double)len + 3.0)));
      wd[i].b = (float)p->msv.tec_b;
    }
   }
  });
  // longest first (a stable counting sort by length), so windows that share
  // a wave or workgroup have similar lengths
  {
    int32_t lmax = 0;
    for (const auto &w : wd) lmax = std::max(lmax, w.len);
    std::vector<int64_t> cnt((size_t)lmax + 2, 0);
    for (const auto &w : wd) cnt[(size_t)(lmax - w.len) + 1]++;
    for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
    std::vector<WinDev> sorted(wd.size());
    for (const auto &w : wd) sorted[cnt[(size_t)(lmax - w.len)]++] = w;
    wd.swap(sorted);
  }
  WinDev *d_w, *h_w;
  RCHK(scratch(c, S_WINS, wd.size() * sizeof(WinDev), (void **)&d_w));
  RCHK(pinned(c, P_UP, wd.size() * sizeof(WinDev), (void **)&h_w));
  std::copy(wd.begin(), wd.end(), h_w);
  HIPCHK(hipMemcpyAsync(d_w, h_w, wd.size() * sizeof(WinDev), hipMemcpyHostToDevice, c->stream));
  float *d_sc;
  uint8_t *d_ovf;
  RCHK(scratch(c, S_SC0, n * 4, (void **)&d_sc));
  RCHK(scratch(c, S_OVF0, n, (void **)&d_ovf));
  hipError_t e = hipSuccess;
  int64_t cells = 0;
  for (auto &w : wd) cells += w.len;
  HIPCHK(ktime_begin(c));
  if (kind == 0) e = fwd_launch_tuned(c, p, d_seq, d_w, n, d_sc);
  else if (kind == 1) e = launch_bias(p->bias_dev(), d_seq, d_w, n, d_sc, c->stream);
  else if (kind == 2) e = launch_vit(p->vit_dev(), d_seq, d_w, n, d_sc, d_ovf, c->stream);
  else e = launch_msv(p->ssv_dev(), d_seq, d_w, n, d_sc, d_ovf, c->stream);
  HIPCHK(e);
  HIPCHK(ktime_end(c, kind == 0 ? IXG_K_FWD : kind == 1 ? IXG_K_BIAS : kind == 2 ? IXG_K_VIT : IXG_K_MSV));
  uint8_t *h_out;
  RCHK(pinned(c, P_DOWN, (size_t)n * 5, (void **)&h_out));
  HIPCHK(hipMemcpyAsync(h_out, d_sc, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (kind >= 2) HIPCHK(hipMemcpyAsync(h_out + n * 4, d_ovf, n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::copy((const float *)h_out, (const float *)h_out + n, sc.begin());
  if (kind >= 2) std::copy(h_out + n * 4, h_out + n * 5, ovf.begin());
  c->cells[1 + (kind == 0 ? 0 : (kind == 1 ? 1 : 2))] += cells * (kind == 1 ? 1 : p->msv.m);
  return IXG_OK;
}

static const double LN2 = 0.693147180559945309417232121458176568;

// Host side of f3_filter_sequence (cm_pipeline.rs:1863-2164) for many tasks.
int cascade(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
            const ixg_f3opts &o, std::vector<std::vector<Win>> &res, std::vector<ixg_acct> &accts) {
  const bool do_msv = o.opts & IXG_DO_MSV, do_fwd = o.opts & IXG_DO_FWD, do_fwdbias = o.opts & IXG_DO_FWDBIAS;
  const bool do_vit = o.opts & IXG_DO_VIT, do_vitbias = o.opts & IXG_DO_VITBIAS,
             do_msvbias = o.opts & IXG_DO_MSVBIAS;
  const size_t nt = tasks.size();
  std::vector<std::vector<Win>> wins(nt);
  if (do_msv) {
    std::vector<std::vector<RawW>> raws;
    RCHK(run_f1(c, p, d_seq, tasks, o.f1, raws));
    Timer tm;
    par_for((int64_t)nt,
            [&](int64_t t) { wins[t] = split_windows(extend_and_merge(p->msv, raws[t], tasks[t].L), o.cmw); });
    c->ms[3] += tm.ms();
  } else {  // deterministic 2*maxW tiling (cm_pipeline.rs:1893-1912)
    for (size_t t = 0; t < nt; t++) {
      const int64_t n = tasks[t].L, mw = o.maxw;
      int64_t nwin = 1;
      if (n > 2 * mw) {
        nwin += (n - 2 * mw) / (mw + 1);
        if ((n - 2 * mw) % (mw + 1) > 0) nwin++;
      }
      for (int64_t i = 0; i < nwin; i++) {
        int64_t ws = 1 + i * (mw + 1);
        wins[t].push_back({ws, std::min(ws + 2 * mw - 1, n), 0.f});
      }
    }
  }
  // flat window list + per-window survival flags
  std::vector<size_t> first(nt + 1, 0);
  for (size_t t = 0; t < nt; t++) first[t + 1] = first[t] + wins[t].size();
  std::vector<WinRef> all(first[nt]);
  par_range((int64_t)nt, 256, [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; t++)
      for (size_t q = 0; q < wins[t].size(); q++) all[first[t] + q] = {(int32_t)t, wins[t][q].start, wins[t][q].end};
  });
  const size_t N = all.size();
  std::vector<uint8_t> s_f1b(N, 1), s_f2(N, 1), s_f2b(N, 1), s_f3(N, 0), s_f3b(N, 0), alive(N, 1);
  std::vector<float> wb(N, -999.0f), biassc(N, 0.f), fwdsc(N, 0.f);
  std::vector<uint8_t> have_bias(N, 0);
  auto alive_list = [&](std::vector<size_t> &idx, std::vector<WinRef> &sub) {
    idx.clear();
    sub.clear();
    for (size_t i = 0; i < N; i++)
      if (alive[i]) {
        idx.push_back(i);
        sub.push_back(all[i]);
      }
  };
  auto ensure_bias = [&](const std::vector<size_t> &idx) -> int {
    std::vector<size_t> need;
    std::vector<WinRef> sub;
    for (size_t i : idx)
      if (!have_bias[i]) {
        need.push_back(i);
        sub.push_back(all[i]);
      }
    std::vector<float> sc;
    std::vector<uint8_t> ov;
    RCHK(scores_impl(c, p, d_seq, tasks, sub, 1, sc, ov));
    for (size_t j = 0; j < need.size(); j++) {
      biassc[need[j]] = sc[j];
      have_bias[need[j]] = 1;
    }
    return IXG_OK;
  };
  std::vector<size_t> idx;
  std::vector<WinRef> sub;
  std::vector<float> sc;
  std::vector<uint8_t> ov;
  hipEvent_t e0 = c->ev[3], e1 = c->ev[4];
  HIPCHK(hipEventRecord(e0, c->stream));
  if (do_msvbias) {  // F1b (cm_pipeline.rs:1947-1958)
    alive_list(idx, sub);
    RCHK(scores_impl(c, p, d_seq, tasks, sub, 3, sc, ov));
    RCHK(ensure_bias(idx));
    for (size_t j = 0; j < idx.size(); j++) {
      if (ov[j]) continue;
      const size_t i = idx[j];
      double w = ((double)sc[j] - (double)biassc[i]) / LN2;
      // the MSV-bias Gumbel uses the p7's local MSV parameters
      if (gumbel_surv(w, p->msv.mmu, p->msv.mlambda) > o.f1b) {
        s_f1b[i] = 0;
        alive[i] = 0;
      }
    }
  }
  if (do_vit) {  // F2/F2b (cm_pipeline.rs:1966-1997)
    alive_list(idx, sub);
    RCHK(scores_impl(c, p, d_seq, tasks, sub, 2, sc, ov));
    std::vector<uint8_t> fail(idx.size(), 0);  // the P-values on the host workers
    par_range((int64_t)idx.size(), 2048, [&](int64_t lo, int64_t hi) {
      for (int64_t j = lo; j < hi; j++) {
        if (ov[j]) continue;
        const int64_t wlen = all[idx[j]].end - all[idx[j]].start + 1;
        const double w = ((double)sc[j] - bg_null_one(wlen)) / LN2;
        fail[j] = gumbel_surv(w, p->vit_lvmu, p->vit_lvlambda) > o.f2;
      }
    });
    std::vector<size_t> pass;
    for (size_t j = 0; j < idx.size(); j++) {
      const size_t i = idx[j];
      if (ov[j]) continue;
      if (fail[j]) {
        s_f2[i] = s_f2b[i] = 0;
        alive[i] = 0;
        continue;
      }
      if (do_vitbias) pass.push_back(j);
    }
    if (do_vitbias) {
      std::vector<size_t> bidx;
      for (size_t j : pass) bidx.push_back(idx[j]);
      RCHK(ensure_bias(bidx));
      par_range((int64_t)pass.size(), 2048, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
          const size_t j = pass[q], i = idx[j];
          const double w = ((double)sc[j] - (double)biassc[i]) / LN2;
          if (gumbel_surv(w, p->vit_lvmu, p->vit_lvlambda) > o.f2b) {
            s_f2b[i] = 0;
            alive[i] = 0;
          }
        }
      });
    }
  }
  if (do_fwd) {  // F3/F3b (cm_pipeline.rs:2013-2052)
    alive_list(idx, sub);
    RCHK(scores_impl(c, p, d_seq, tasks, sub, 0, sc, ov));
    std::vector<size_t> pass;
    for (size_t j = 0; j < idx.size(); j++) {
      const size_t i = idx[j];
      const int64_t wlen = all[i].end - all[i].start + 1;
      fwdsc[i] = sc[j];
      double w = ((double)sc[j] - bg_null_one(wlen)) / LN2;
      double pv = exp_surv(w, p->fwd.ftau, p->fwd.flambda);
      wb[i] = (float)w;
      if (pv > o.f3) {
        alive[i] = 0;
        continue;
      }
      s_f3[i] = 1;
      pass.push_back(i);
    }
    if (do_fwdbias) {
      RCHK(ensure_bias(pass));
      for (size_t i : pass) {
        double w = ((double)fwdsc[i] - (double)biassc[i]) / LN2;
        double pv = exp_surv(w, p->fwd.ftau, p->fwd.flambda);
        wb[i] = (float)w;
        if (pv > o.f3b) {
          alive[i] = 0;
          continue;
        }
        s_f3b[i] = 1;
      }
    } else {
      for (size_t i : pass) s_f3b[i] = 1;
    }
  } else {
    for (size_t i = 0; i < N; i++)
      if (alive[i]) s_f3[i] = s_f3b[i] = 1;
  }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float wms = 0;
  HIPCHK(hipEventElapsedTime(&wms, e0, e1));
  c->ms[2] += wms;
  Timer tm;
  // survivors, merge of overlapping/abutting windows (cm_pipeline.rs:2056-2114)
  res.assign(nt, {});
  accts.assign(nt, ixg_acct{});
  par_for((int64_t)nt, [&](int64_t t) {
    std::vector<Win> sv;
    for (size_t i = first[t]; i < first[t + 1]; i++)
      if (s_f3b[i]) sv.push_back({all[i].start, all[i].end, wb[i]});
    size_t i = 0;
    while (i < sv.size()) {
      size_t i2 = i + 1;
      while (i2 < sv.size() && sv[i].end + 1 >= sv[i2].start) {
        sv[i].end = sv[i2].end;
        sv[i].score = std::max(sv[i].score, sv[i2].score);
        i2++;
      }
      res[t].push_back(sv[i]);
      i = i2;
    }
    ixg_acct &a = accts[t];  // cm_pipeline.rs:2116-2164
    a.n_past_msv = first[t + 1] - first[t];
    for (size_t k = first[t]; k < first[t + 1]; k++) {
      const uint64_t wl = (uint64_t)(all[k].end - all[k].start + 1);
      a.pos_past_msv += wl;
      if (s_f1b[k]) { a.n_past_msvbias++; a.pos_past_msvbias += wl; }
      if (s_f2[k]) { a.n_past_vit++; a.pos_past_vit += wl; }
      if (s_f2b[k]) { a.n_past_vitbias++; a.pos_past_vitbias += wl; }
      if (s_f3[k]) { a.n_past_fwd++; a.pos_past_fwd += wl; }
      if (s_f3b[k]) { a.n_past_fwdbias++; a.pos_past_fwdbias += wl; }
      if (k > first[t]) {
        const int64_t ovl = all[k - 1].end - all[k].start + 1;
        if (ovl > 0) {
          const uint64_t ovv = (uint64_t)ovl;
          a.pos_past_msv -= ovv;
          if (s_f1b[k] && s_f1b[k - 1]) a.pos_past_msvbias -= ovv;
          if (s_f2[k] && s_f2[k - 1]) a.pos_past_vit -= ovv;
          if (s_f2b[k] && s_f2b[k - 1]) a.pos_past_vitbias -= ovv;
          if (s_f3[k] && s_f3[k - 1]) a.pos_past_fwd -= ovv;
          if (s_f3b[k] && s_f3b[k - 1]) a.pos_past_fwdbias -= ovv;
        }
      }
    }
  });
  c->ms[3] += tm.ms();
  return IXG_OK;
}

// Residue codes must index the 18-symbol tables (easel/alphabet.rs:46-157).
bool codes_ok(const uint8_t *p, int64_t n) {
  uint8_t acc = 0;
  for (int64_t i = 0; i < n; i++) acc |= (uint8_t)(p[i] >= KP);
  if (acc) g_err = "digitized residue code >= 18 in the sequence";
  return !acc;
}

// Uploads a host 1-indexed sequence into a scratch slot (sentinel-padded).
int upload_seq(ixg_ctx *c, const uint8_t *dsq, int64_t L, uint8_t **out) {
  if (!codes_ok(dsq + 1, L)) return IXG_EINVAL;
  uint8_t *d;
  RCHK(scratch(c, S_SEQ, (size_t)L + 2, (void **)&d));
  std::vector<uint8_t> h(L + 2, 255);
  if (L) memcpy(h.data() + 1, dsq + 1, L);
  HIPCHK(hipMemcpyAsync(d, h.data(), L + 2, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = d;
  return IXG_OK;
}

}  // namespace

namespace ixg {
int run_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
               const std::vector<WinRef> &wl, int kind, std::vector<float> &sc, std::vector<uint8_t> &ovf) {
  return scores_impl(c, p, d_seq, tasks, wl, kind, sc, ovf);
}

// The cascade over tasks given as (base, L) views of a device sequence
// buffer; windows come back grouped by task with their task index, first[t]
// is where task t's windows begin (first[nt] = total).
int cascade_views(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const int64_t *base, const int64_t *len,
                  int64_t nt, const ixg_f3opts &o, std::vector<ixg_window> &out, std::vector<int64_t> &first,
                  std::vector<ixg_acct> &acct) {
  std::vector<HostTask> ht(nt);
  for (int64_t t = 0; t < nt; t++) ht[t] = {base[t], len[t]};
  std::vector<std::vector<Win>> res;
  RCHK(cascade(c, p, d_seq, ht, o, res, acct));
  out.clear();
  first.assign(nt + 1, 0);
  for (int64_t t = 0; t < nt; t++) {
    first[t] = (int64_t)out.size();
    for (auto &w : res[t]) out.push_back({w.start, w.end, w.score, (int32_t)t});
  }
  first[nt] = (int64_t)out.size();
  return IXG_OK;
}

}  // namespace ixg

extern "C" {

int ixg_ssv_diagonals(ixg_ctx *c, const ixg_p7 *p, const uint8_t *dsq, int64_t L, double f1, ixg_rawwin *out,
                      int64_t cap, int64_t *n) {
  if (!c || !p || !dsq || L < 0 || !n || (cap > 0 && !out)) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(upload_seq(c, dsq, L, &d));
  std::vector<std::vector<RawW>> raws;
  RCHK(run_f1(c, p, d, {{0, L}}, f1, raws));
  const auto &r = raws[0];
  *n = (int64_t)r.size();
  for (int64_t i = 0; i < std::min<int64_t>(cap, r.size()); i++)
    out[i] = {r[i].n, r[i].length, r[i].k, r[i].score, 0};
  return (int64_t)r.size() > cap ? IXG_ERANGE : IXG_OK;
}

static int scores_entry(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                        const int64_t *len, int64_t nwin, float *sc, uint8_t *ovf, int kind) {
  if (!c || !p || !seq || !off || !len || !sc || nwin < 0) return IXG_EINVAL;
  for (int64_t i = 0; i < nwin; i++)
    if (off[i] < 0 || len[i] < 1 || off[i] + len[i] >= seq_len || len[i] > INT32_MAX) {
      g_err = "window outside the sequence buffer";
      return IXG_EINVAL;
    }
  if (!codes_ok(seq, seq_len)) {
    // only window residues are read; check exactly those
    for (int64_t i = 0; i < nwin; i++)
      if (!codes_ok(seq + off[i] + 1, len[i])) return IXG_EINVAL;
    g_err.clear();
  }
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_SEQ, (size_t)seq_len + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, seq, seq_len, hipMemcpyHostToDevice, c->stream));
  std::vector<HostTask> tasks{{0, seq_len}};
  std::vector<WinRef> wl(nwin);
  for (int64_t i = 0; i < nwin; i++) wl[i] = {0, off[i] + 1, off[i] + len[i]};
  std::vector<float> s;
  std::vector<uint8_t> o;
  RCHK(scores_impl(c, p, d, tasks, wl, kind, s, o));
  memcpy(sc, s.data(), nwin * 4);
  if (ovf) memcpy(ovf, o.data(), nwin);
  return IXG_OK;
}

int ixg_vit_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                   const int64_t *len, int64_t nwin, float *sc, uint8_t *ovf) {
  return scores_entry(c, p, seq, seq_len, off, len, nwin, sc, ovf, 2);
}
int ixg_fwd_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                   const int64_t *len, int64_t nwin, float *sc) {
  return scores_entry(c, p, seq, seq_len, off, len, nwin, sc, nullptr, 0);
}
int ixg_bias_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                    const int64_t *len, int64_t nwin, float *sc) {
  return scores_entry(c, p, seq, seq_len, off, len, nwin, sc, nullptr, 1);
}
int ixg_msv_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                   const int64_t *len, int64_t nwin, float *sc, uint8_t *ovf) {
  return scores_entry(c, p, seq, seq_len, off, len, nwin, sc, ovf, 3);
}

int ixg_p7_forward_batch(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                         const int64_t *len, int64_t n, int unihit, float *mat, float *xmx, float *sc,
                         float *totscale) {
  if (!c || !p || !seq || (n && (!off || !len)) || n < 0 || !mat || !xmx || !sc || !totscale) return IXG_EINVAL;
  if (!p->d_fwd) {
    g_err = "ixg_p7_forward: host-only profile";
    return IXG_EINVAL;
  }
  const int M = p->fwd.m;
  std::vector<int64_t> moff(n), xoff(n);
  int64_t mtot = 0, xtot = 0;
  for (int64_t i = 0; i < n; i++) {
    if (off[i] < 0 || len[i] < 1 || off[i] + len[i] >= seq_len || len[i] > INT32_MAX) {
      g_err = "region outside the sequence buffer";
      return IXG_EINVAL;
    }
    if (!codes_ok(seq + off[i] + 1, len[i])) return IXG_EINVAL;
    moff[i] = mtot;
    xoff[i] = xtot;
    mtot += 3 * (len[i] + 1) * (int64_t)(M + 1);
    xtot += 6 * (len[i] + 1);
  }
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_SEQ, (size_t)seq_len + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, seq, seq_len, hipMemcpyHostToDevice, c->stream));
  std::vector<WinDev> wd(n);
  for (int64_t i = 0; i < n; i++) {
    wd[i].off = off[i];
    wd[i].len = (int32_t)len[i];
    wd[i].idx = (int32_t)i;
    wd[i].a = wd[i].b = 0.f;
  }
  std::vector<int64_t> offs(moff);
  offs.insert(offs.end(), xoff.begin(), xoff.end());
  WinDev *d_w;
  int64_t *d_off;
  float *d_mat, *d_x, *d_out;
  RCHK(upload(c, S_WINS, wd, &d_w));
  RCHK(upload(c, S_TASKS, offs, &d_off));
  RCHK(scratch(c, S_CM_ALPHA, (size_t)std::max<int64_t>(mtot, 1) * 4, (void **)&d_mat));
  RCHK(scratch(c, S_CM_ROOT, (size_t)std::max<int64_t>(xtot, 1) * 4, (void **)&d_x));
  RCHK(scratch(c, S_SC0, (size_t)std::max<int64_t>(2 * n, 1) * 4, (void **)&d_out));
  HIPCHK(ktime_begin(c));
  HIPCHK(launch_fwd_full(p->fwd_dev(), d, d_w, n, d_off, d_off + n, d_mat, d_x, d_out, unihit ? 1 : 0, c->stream));
  HIPCHK(ktime_end(c, IXG_K_FWD_FULL));
  std::vector<float> o(2 * n);
  if (n) HIPCHK(hipMemcpyAsync(o.data(), d_out, 2 * n * 4, hipMemcpyDeviceToHost, c->stream));
  if (mtot) HIPCHK(hipMemcpyAsync(mat, d_mat, mtot * 4, hipMemcpyDeviceToHost, c->stream));
  if (xtot) HIPCHK(hipMemcpyAsync(xmx, d_x, xtot * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; i++) {
    sc[i] = o[2 * i];
    totscale[i] = o[2 * i + 1];
  }
  return IXG_OK;
}

int ixg_p7_forward(ixg_ctx *c, const ixg_p7 *p, const uint8_t *dsq, int64_t L, int unihit, float *mmx, float *imx,
                   float *dmx, float *xmx, float *totscale, float *sc) {
  if (!c || !p || !dsq || L < 1 || !mmx || !imx || !dmx || !xmx || !totscale || !sc) return IXG_EINVAL;
  const int64_t off = 0, len = L;
  const size_t plane = (size_t)(L + 1) * (p->fwd.m + 1);
  std::vector<float> mat(3 * plane);
  RCHK(ixg_p7_forward_batch(c, p, dsq, L + 1, &off, &len, 1, unihit, mat.data(), xmx, sc, totscale));
  memcpy(mmx, mat.data(), plane * 4);
  memcpy(imx, mat.data() + plane, plane * 4);
  memcpy(dmx, mat.data() + 2 * plane, plane * 4);
  return IXG_OK;
}

static int emit(const std::vector<std::vector<Win>> &res, ixg_window *out, int64_t cap, int64_t *n) {
  int64_t k = 0;
  for (size_t t = 0; t < res.size(); t++)
    for (auto &w : res[t]) {
      if (k < cap && out) out[k] = {w.start, w.end, w.score, (int32_t)t};
      k++;
    }
  *n = k;
  return k > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_f3_filter_sequence(ixg_ctx *c, const ixg_p7 *p, const uint8_t *dsq, int64_t L, const ixg_f3opts *o,
                           ixg_window *out, int64_t cap, int64_t *n, ixg_acct *acct) {
  if (!c || !p || !dsq || !o || !n || L < 0) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(upload_seq(c, dsq, L, &d));
  std::vector<std::vector<Win>> res;
  std::vector<ixg_acct> accts;
  RCHK(cascade(c, p, d, {{0, L}}, *o, res, accts));
  if (acct) *acct = accts[0];
  return emit(res, out, cap, n);
}

int ixg_f3_filter_tasks(ixg_ctx *c, const ixg_p7 *p, const ixg_target *tgt, const ixg_task *tasks, int64_t nt,
                        const ixg_f3opts *o, ixg_window *out, int64_t cap, int64_t *n, ixg_acct *acct) {
  if (!c || !p || !tgt || (!tasks && nt) || !o || !n) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  Timer total;
  for (int i = 0; i < 5; i++) c->ms[i] = 0, c->cells[i] = 0;
  std::vector<HostTask> ht(nt);
  for (int64_t t = 0; t < nt; t++) {
    const ixg_task &k = tasks[t];
    if (k.start < 1 || k.end > tgt->L || k.end < k.start || (k.strand != 0 && k.strand != 1)) {
      g_err = "task outside the target";
      return IXG_EINVAL;
    }
    // top strand: dsq[1] = seq[start]; bottom: dsq[k] = comp(seq[end-k+1]) = rc[L-end+k]
    ht[t].base = k.strand == 0 ? (k.start - 1) : (tgt->L + 2) + (tgt->L - k.end);
    ht[t].L = k.end - k.start + 1;
  }
  std::vector<std::vector<Win>> res;
  std::vector<ixg_acct> accts;
  RCHK(cascade(c, p, tgt->d_both, ht, *o, res, accts));
  if (acct)
    for (int64_t t = 0; t < nt; t++) acct[t] = accts[t];
  c->ms[4] = total.ms();
  return emit(res, out, cap, n);
}

}  // extern "C"

namespace ixg {
// F1-F3b of many sequence views of a target (HostTask bases in its device
// buffer, any of its sequences and strands); windows carry their view's
// index. The stage timings (ms, cells) add up over calls: the caller resets.
int f3_filter_views(ixg_ctx *c, const ixg_p7 *p, const ixg_target *tgt, const std::vector<HostTask> &tasks,
                    const ixg_f3opts &o, std::vector<ixg_window> &wins, std::vector<ixg_acct> &accts) {
  HIPCHK(hipSetDevice(c->device));
  for (const HostTask &t : tasks)
    if (t.base < 0 || t.L < 0 || (size_t)(t.base + t.L + 2) > tgt->bytes) {
      g_err = "f3_filter_views: a view outside the target";
      return IXG_EINVAL;
    }
  Timer total;
  std::vector<std::vector<Win>> res;
  RCHK(cascade(c, p, tgt->d_both, tasks, o, res, accts));
  c->ms[4] += total.ms();
  int64_t n = 0;
  for (auto &r : res) n += (int64_t)r.size();
  wins.resize(n);
  return emit(res, wins.data(), n, &n);
}
}  // namespace ixg

extern "C" {

int ixg_set_autotune(ixg_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 1 + FWD_V_N) {
    g_err = "ixg_set_autotune: bad arguments";
    return IXG_EINVAL;
  }
  c->autotune = mode;
  c->tuned.clear();
  return IXG_OK;
}

int ixg_autotune_choices(ixg_ctx *c, int32_t *nq, int32_t *cls, int32_t *shape, int64_t cap, int64_t *n) {
  if (!c || !n || (cap > 0 && (!nq || !cls || !shape))) {
    g_err = "ixg_autotune_choices: bad arguments";
    return IXG_EINVAL;
  }
  int64_t k = 0;
  for (const auto &kv : c->tuned) {
    if (k < cap) nq[k] = kv.first.first, cls[k] = kv.first.second, shape[k] = kv.second;
    k++;
  }
  *n = k;
  return k > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_copy_probe(ixg_ctx *c, size_t bytes, int reps, double *ms_per_copy) {
  if (!c || bytes < 16 || reps < 1 || !ms_per_copy) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  RCHK(dmalloc(c, &a, n * 16));
  int rc = dmalloc(c, &b, n * 16);
  if (rc != IXG_OK) {
    dfree(c, a, n * 16);
    return rc;
  }
  double tot = 0;
  hipError_t e = hipMemsetAsync(a, 1, n * 16, c->stream);
  for (int r = 0; r <= reps && e == hipSuccess; r++) {  // the first copy is a warm-up
    if ((e = ktime_begin(c)) != hipSuccess) break;
    if ((e = launch_copy_probe(a, b, (int64_t)n, c->stream)) != hipSuccess) break;
    const double before = c->kms[IXG_K_COPY];
    if ((e = ktime_end(c, IXG_K_COPY)) != hipSuccess) break;
    if (r) tot += c->kms[IXG_K_COPY] - before;
  }
  dfree(c, a, n * 16);
  dfree(c, b, n * 16);
  if (e != hipSuccess) {
    g_err = std::string("ixg_copy_probe: ") + hipGetErrorString(e);
    return IXG_EDEVICE;
  }
  *ms_per_copy = tot / reps;
  return IXG_OK;
}

int ixg_kernel_timings(ixg_ctx *c, double *ms, int64_t *launches, int reset) {
  if (!c) return IXG_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(ktime_flush(c));
  if (ms) memcpy(ms, c->kms, sizeof(c->kms));
  if (launches) memcpy(launches, c->kn, sizeof(c->kn));
  if (reset)
    for (int k = 0; k < IXG_NKERNELS; k++) c->kms[k] = 0, c->kn[k] = 0;
  return IXG_OK;
}

int ixg_last_timings(const ixg_ctx *c, double *ms5, int64_t *cells5) {
  if (!c) return IXG_EINVAL;
  if (ms5) memcpy(ms5, c->ms, sizeof(c->ms));
  if (cells5) memcpy(cells5, c->cells, sizeof(c->cells));
  return IXG_OK;
}

}  // extern "C"

// ---- host worker pool of par_for ----
#include <condition_variable>
#include <deque>
#include <mutex>
namespace ixg {
namespace {
struct PoolJob {
  int64_t n;
  void (*body)(void *, int64_t);
  void *arg;
  std::atomic<int64_t> next{0}, done{0};
  int users = 0;  // pool workers inside work() on this job (guarded by the pool mutex)
};
struct Pool {
  std::mutex mu;
  std::condition_variable cv, fin;
  std::deque<PoolJob *> jobs;  // jobs that may have chunks left to claim
  Pool() {
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = (int)std::min<unsigned>(16u, hw ? hw : 1u) - 1;
    for (int t = 0; t < nt; t++) std::thread([this] { loop(); }).detach();  // live as long as the process
  }
  static void work(PoolJob *j) {
    for (;;) {
      const int64_t k0 = j->next.fetch_add(4);
      if (k0 >= j->n) break;
      const int64_t k1 = std::min<int64_t>(j->n, k0 + 4);
      for (int64_t k = k0; k < k1; k++) j->body(j->arg, k);
      j->done.fetch_add(k1 - k0);
    }
  }
  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [&] { return !jobs.empty(); });
      PoolJob *j = jobs.front();
      if (j->next.load() >= j->n) {  // fully claimed: retire it from the queue
        jobs.pop_front();
        continue;
      }
      j->users++;  // the owner does not return while a worker holds the job
      lk.unlock();
      work(j);
      lk.lock();
      j->users--;
      fin.notify_all();
    }
  }
};
Pool &pool() {
  static Pool *p = new Pool();  // never destroyed: workers may outlive static destruction
  return *p;
}
}  // namespace

void pool_run(int64_t n, void (*body)(void *, int64_t), void *arg) {
  Pool &p = pool();
  PoolJob j;
  j.n = n, j.body = body, j.arg = arg;
  {
    std::lock_guard<std::mutex> lk(p.mu);
    p.jobs.push_back(&j);
  }
  p.cv.notify_all();
  Pool::work(&j);
  std::unique_lock<std::mutex> lk(p.mu);
  p.fin.wait(lk, [&] { return j.done.load() == j.n && j.users == 0; });
  for (auto it = p.jobs.begin(); it != p.jobs.end(); ++it)  // still queued when no worker retired it
    if (*it == &j) {
      p.jobs.erase(it);
      break;
    }
}
}  // namespace ixg

// C ABI of the CP9 HMM banding DPs (cp9.rs:1275 cp9_forward, :1995
// cp9_backward, :1517 cp9_fb2hmmbands): the resident HMM with its local-end
// lists in both directions, one wave per window, launch groups of windows
// whose six matrices fit the matrix budget.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include "cm_handle.h"

using namespace ixg;

struct ixg_cp9 {
  int M = 0, el = 0, el_selfsc = 0;
  int32_t *d = nullptr;  // otsc | msc | isc | el_ct | el_off | el_idx | bel_* | ilog | has_el (bytes)
  size_t bytes = 0;
  Cp9Dev dev;
};

const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L);

namespace {

// cp9.rs:977-998 Prob2Score, the band threshold of cp9_fb2hmmbands (:1531)
int32_t band_thresh(double tau) {
  const float p = (float)((1.0 - (1.0 - tau)) / 2.0);
  if (p == 0.0f) return -INFTY_I;
  const double lg = (double)p > 0.0 ? std::log((double)p) * 1.44269504 : -1e36;
  const double v = std::floor(0.5 + 1000.0 * lg);
  if (!std::isfinite(v)) return -INFTY_I;
  return std::max((int32_t)v, -INFTY_I);
}

struct CWin {
  const uint8_t *dsq;  // device
  int32_t i0, j0;
};

// Runs Forward, Backward and the posterior bands of every window; fsc, bsc,
// pn are host arrays in the windows' order. mats (optional, n == 1): the six
// matrices of the single window in [(L+1)][(M+1)] order.
int cp9_run(ixg_ctx *c, const ixg_cp9 *h, const std::vector<CWin> &wins, double tau, int mode, float *fsc,
            float *bsc, int32_t *pn, int32_t **mats, int32_t *occ = nullptr) {
  const int M = h->M, M1 = M + 1;
  const int64_t n = (int64_t)wins.size();
  if (!(tau > 0.0 && tau < 1.0)) {
    g_err = "CP9 bands: tau must be in (0, 1)";
    return IXG_EINVAL;
  }
  if (mode & ~(IXG_CP9_FWD_SCAN | IXG_CP9_BCK_SCAN | IXG_CP9_ALIGN)) {
    g_err = "CP9 bands: unknown mode bits";
    return IXG_EINVAL;
  }
  const int32_t thresh = band_thresh(tau);
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return wins[a].j0 - wins[a].i0 > wins[b].j0 - wins[b].i0;
  });
  const int64_t budget = arena_elems(c, S_CM_ALPHA);
  std::vector<float> hf(std::max<int64_t>(n, 1)), hb(std::max<int64_t>(n, 1));
  std::vector<int32_t> hp((size_t)std::max<int64_t>(n, 1) * 6 * M1);
  int64_t g0 = 0;  // first window (in sorted order) of the launch group
  while (g0 < n) {
    // windows, longest first, while their matrices fit
    int64_t g1 = g0, elems = 0;
    std::vector<int64_t> woff;
    while (g1 < n) {
      const int L = wins[order[g1]].j0 - wins[order[g1]].i0 + 1;
      const int64_t need = 6 * (int64_t)(L + 1) * M1 + 64;
      if (g1 > g0 && elems + need > budget) break;
      woff.push_back(elems);
      elems += need;
      g1++;
    }
    const int64_t nw = g1 - g0;
    int32_t *d_mx;
    RCHK(scratch(c, S_CM_ALPHA, (size_t)elems * 4, (void **)&d_mx));
    std::vector<Cp9WinDev> wd(nw);
    for (int64_t k = 0; k < nw; k++) {
      const CWin &w = wins[order[g0 + k]];
      wd[k] = Cp9WinDev{w.dsq, w.i0, w.j0, k, d_mx + woff[k]};
    }
    const size_t wbytes = nw * sizeof(Cp9WinDev);
    const size_t obytes = (size_t)nw * (2 * 4 + 6 * M1 * 4) + (occ ? (size_t)nw * M1 * 4 : 0);
    char *d_meta;
    RCHK(scratch(c, S_CM_ENV, wbytes + obytes + 64, (void **)&d_meta));
    Cp9WinDev *d_w = (Cp9WinDev *)d_meta;
    float *d_f = (float *)(d_meta + ((wbytes + 15) & ~(size_t)15));
    float *d_b = d_f + nw;
    int32_t *d_pn = (int32_t *)(d_b + nw);
    int32_t *d_occ = occ ? d_pn + (size_t)nw * 6 * M1 : nullptr;
    HIPCHK(hipMemcpyAsync(d_w, wd.data(), wbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(ktime_begin(c));
    HIPCHK(launch_cp9_bands(h->dev, d_w, nw, mode, thresh, d_f, d_b, d_pn, d_occ, c->stream));
    HIPCHK(ktime_end(c, IXG_K_CP9));
    std::vector<float> tf(nw), tb(nw);
    std::vector<int32_t> tp((size_t)nw * 6 * M1);
    RCHK(d2h(c, tf.data(), d_f, nw * 4));
    RCHK(d2h(c, tb.data(), d_b, nw * 4));
    RCHK(d2h(c, tp.data(), d_pn, tp.size() * 4));
    std::vector<int32_t> to(occ ? (size_t)nw * M1 : 0);
    if (occ) RCHK(d2h(c, to.data(), d_occ, to.size() * 4));
    for (int64_t k = 0; k < nw; k++) {
      const int64_t o = order[g0 + k];
      hf[o] = tf[k];
      hb[o] = tb[k];
      std::copy(tp.begin() + k * 6 * M1, tp.begin() + (k + 1) * 6 * M1, hp.begin() + o * 6 * M1);
      if (occ) std::copy(to.begin() + k * M1, to.begin() + (k + 1) * M1, occ + o * M1);
    }
    if (mats && n == 1) {  // the single window's matrices, [(L+1)][(M+1)] each
      const int L = wins[0].j0 - wins[0].i0 + 1;
      const size_t cells = (size_t)(L + 1) * M1;
      for (int m = 0; m < 6; m++)
        if (mats[m]) RCHK(d2h(c, mats[m], d_mx + m * cells, cells * 4));
    }
    g0 = g1;
  }
  if (fsc) std::copy(hf.begin(), hf.begin() + n, fsc);
  if (bsc) std::copy(hb.begin(), hb.begin() + n, bsc);
  if (pn) std::copy(hp.begin(), hp.begin() + n * 6 * M1, pn);
  return IXG_OK;
}

int check_cp9_window(const ixg_cp9 *h, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0) {
  if (!h || !dsq || i0 < 1 || j0 > L || j0 < i0) {
    g_err = "CP9 bands: bad arguments (1 <= i0 <= j0 <= L)";
    return IXG_EINVAL;
  }
  for (int64_t i = i0; i <= j0; i++)
    if (dsq[i] >= KP) {
      g_err = "digitized residue code >= 18 in the window";
      return IXG_EINVAL;
    }
  return IXG_OK;
}

}  // namespace

extern "C" {

int ixg_cp9_create(ixg_ctx *c, const ixg_cp9_params *p, ixg_cp9 **out) {
  if (!c || !p || !out || p->M < 1 || !p->otsc || !p->msc || !p->isc || !p->has_el || !p->el_from_ct ||
      !p->el_from_off || (p->el && !p->el_from_idx)) {
    g_err = "ixg_cp9_create: bad arguments";
    return IXG_EINVAL;
  }
  const int M = p->M, M1 = M + 1;
  const int nidx = p->el_from_off[M + 2];
  for (int k = 0; k <= M + 1; k++)
    if (p->el_from_ct[k] < 0 || p->el_from_off[k] + p->el_from_ct[k] > nidx) {
      g_err = "ixg_cp9_create: local-end lists out of range";
      return IXG_EINVAL;
    }
  for (int x = 0; x < nidx; x++)
    if (p->el_from_idx[x] < 0 || p->el_from_idx[x] > M) {
      g_err = "ixg_cp9_create: local-end source outside 0..M";
      return IXG_EINVAL;
    }
  // A local end's source node precedes the node it ends at (cp9.rs:1042-1076
  // takes them from a CM node's left and right consensus positions); the
  // Backward completes each local-end cell before its node's match cell.
  for (int k = 1; k <= M + 1; k++)
    for (int c = 0; c < p->el_from_ct[k]; c++)
      if (p->el_from_idx[p->el_from_off[k] + c] >= k) {
        g_err = "ixg_cp9_create: a local-end source must precede the node it ends at";
        return IXG_EINVAL;
      }
  std::vector<int32_t> ilog(FLOGSUM_TBL);
  int nz = 1;
  for (int i = 0; i < FLOGSUM_TBL; i++) {  // cp9.rs:28-43 ILogsum table
    ilog[i] = (int32_t)std::nearbyint(1000.0 * std::log2(1.0 + std::pow(2.0, -(double)i / 1000.0)));
    if (ilog[i] != 0) nz = i + 1;
  }
  if (cp9_lds_bytes(M, nz) + 64 > 65536) {
    g_err = "ixg_cp9_create: the HMM has more nodes than the kernels' LDS holds (M <= 2100)";
    return IXG_EINVAL;
  }
  // Backward local-end lists per source node y (used only with local ends)
  std::vector<int32_t> bel_nM(M1, 0), bel_reset(M1, 0), bel_off(M1 + 1, 0), bel_k;
  if (p->el) {
    for (int c = 0; c < p->el_from_ct[M]; c++) bel_nM[p->el_from_idx[p->el_from_off[M] + c]]++;
    for (int c = 0; c < p->el_from_ct[M + 1]; c++) bel_reset[p->el_from_idx[p->el_from_off[M + 1] + c]] = 1;
    std::vector<std::vector<int32_t>> per(M1);
    for (int k = M - 1; k >= 1; k--)
      for (int c = 0; c < p->el_from_ct[k]; c++) per[p->el_from_idx[p->el_from_off[k] + c]].push_back(k);
    for (int y = 0; y <= M; y++) {
      bel_off[y] = (int32_t)bel_k.size();
      bel_k.insert(bel_k.end(), per[y].begin(), per[y].end());
    }
    bel_off[M1] = (int32_t)bel_k.size();
  }
  // LDS slots of the wavefront kernels, coloured in node order. The reader of
  // a slot runs one lane (one step) behind its writer, so a slot is free for a
  // new occupant two nodes after its occupant's last reader.
  std::vector<int32_t> fslot(M1, -1), bslot(M1, -1);
  int nsf = 0, nsb = 0;
  {  // Forward: a source node's local-end cell, read up to the last node it ends at (M+1: the row end)
    std::vector<int> last(M1, -1), ends;
    for (int k = 1; k <= M + 1; k++)
      for (int c = 0; c < p->el_from_ct[k]; c++) {
        const int y = p->el_from_idx[p->el_from_off[k] + c];
        last[y] = std::max(last[y], k);
      }
    for (int y = 0; y <= M; y++) {
      if (last[y] < 0) continue;
      int q = 0;
      while (q < (int)ends.size() && ends[q] + 2 > y) q++;
      if (q == (int)ends.size()) ends.push_back(0);
      ends[q] = last[y];
      fslot[y] = q;
    }
    nsf = (int)ends.size();
  }
  if (p->el) {  // Backward: an end node's match cell, read down to the lowest source ending there
    std::vector<int> low(M1, M1), ends;
    for (int k = 1; k <= M - 1; k++)
      for (int c = 0; c < p->el_from_ct[k]; c++) low[k] = std::min(low[k], p->el_from_idx[p->el_from_off[k] + c]);
    for (int e = M - 1; e >= 1; e--) {
      if (low[e] == M1) continue;
      int q = 0;
      while (q < (int)ends.size() && ends[q] - 2 < e) q++;
      if (q == (int)ends.size()) ends.push_back(0);
      ends[q] = low[e];
      bslot[e] = q;
    }
    nsb = (int)ends.size();
  }
  std::vector<int32_t> blob;
  auto put = [&](const int32_t *x, size_t k) {
    const size_t o = blob.size();
    blob.insert(blob.end(), x, x + k);
    while (blob.size() % 4) blob.push_back(0);
    return o;
  };
  const size_t o_t = put(p->otsc, (size_t)M1 * 12), o_m = put(p->msc, (size_t)M1 * KP),
               o_i = put(p->isc, (size_t)M1 * KP), o_c = put(p->el_from_ct, M + 2), o_o = put(p->el_from_off, M + 3);
  const int32_t zero = 0;
  const size_t o_x = nidx ? put(p->el_from_idx, nidx) : put(&zero, 1);
  const size_t o_bn = put(bel_nM.data(), M1), o_br = put(bel_reset.data(), M1), o_bo = put(bel_off.data(), M1 + 1);
  const size_t o_bk = bel_k.empty() ? put(&zero, 1) : put(bel_k.data(), bel_k.size());
  std::vector<int32_t> el_slot(std::max(nidx, 1), -1), bel_slot(std::max<size_t>(bel_k.size(), 1), -1);
  for (int x = 0; x < nidx; x++) el_slot[x] = fslot[p->el_from_idx[x]];
  for (size_t x = 0; x < bel_k.size(); x++) bel_slot[x] = bslot[bel_k[x]];
  if (M > 32767) {  // list entries are staged in LDS as 16-bit node numbers
    g_err = "ixg_cp9_create: more than 32767 nodes";
    return IXG_EINVAL;
  }
  // packed per-node words of the wavefront kernels (16-bit fields)
  std::vector<int32_t> finfo(2 * (M + 2), 0), binfo(2 * (M + 2), 0);
  int wave = (nidx < 65536 && nsf < 32768 ? 1 : 0) | (bel_k.size() < 65536 && nsb < 32768 ? 2 : 0);
  for (int k = 0; k <= M + 1; k++) {
    finfo[k] = (int32_t)((uint32_t)p->el_from_off[k] << 16 | (uint32_t)(p->el_from_ct[k] & 0xFFFF));
    const int sl = k <= M ? fslot[k] : -1, he = k <= M ? (p->has_el[k] != 0) : 0;
    finfo[M + 2 + k] = (int32_t)(((uint32_t)sl & 0xFFFF) | (uint32_t)he << 16);
    if (p->el_from_ct[k] > 0xFFFF) wave &= ~1;
  }
  for (int k = 0; k <= M; k++) {
    const int n = bel_k.empty() ? 0 : bel_off[k + 1] - bel_off[k];
    binfo[k] = (int32_t)((uint32_t)(bel_k.empty() ? 0 : bel_off[k]) << 16 | (uint32_t)n);
    binfo[M + 2 + k] = (int32_t)(((uint32_t)bslot[k] & 0xFFFF) | (uint32_t)(p->has_el[k] != 0) << 16 |
                                 (uint32_t)(bel_reset[k] != 0) << 17 | (uint32_t)bel_nM[k] << 18);
    if (bel_nM[k] >= (1 << 14)) wave &= ~2;
  }
  const size_t o_fs = put(fslot.data(), M1), o_bs = put(bslot.data(), M1);
  const size_t o_fi = put(finfo.data(), finfo.size()), o_bi = put(binfo.data(), binfo.size());
  const size_t o_es = put(el_slot.data(), el_slot.size()), o_bes = put(bel_slot.data(), bel_slot.size());
  const size_t o_l = put(ilog.data(), ilog.size());
  const size_t o_h = blob.size();
  blob.resize(o_h + (M1 + 3) / 4 + 1, 0);
  std::memcpy(blob.data() + o_h, p->has_el, M1);
  ixg_cp9 *h = new ixg_cp9;
  h->M = M;
  h->el = p->el ? 1 : 0;
  h->el_selfsc = p->el_selfsc;
  h->bytes = blob.size() * 4;
  if (int r = dmalloc(c, (void **)&h->d, h->bytes)) {
    delete h;
    return r;
  }
  if (hipMemcpy(h->d, blob.data(), h->bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamSynchronize(nullptr) != hipSuccess) {
    dfree(c, h->d, h->bytes);
    delete h;
    g_err = "ixg_cp9_create: upload failed";
    return IXG_EDEVICE;
  }
  Cp9Dev &d = h->dev;
  d.M = M;
  d.el = h->el;
  d.el_selfsc = h->el_selfsc;
  d.otsc = h->d + o_t;
  d.msc = h->d + o_m;
  d.isc = h->d + o_i;
  d.el_ct = h->d + o_c;
  d.el_off = h->d + o_o;
  d.el_idx = h->d + o_x;
  d.bel_nM = h->d + o_bn;
  d.bel_reset = h->d + o_br;
  d.bel_off = h->d + o_bo;
  d.bel_k = h->d + o_bk;
  d.nz = nz;
  d.fslot = h->d + o_fs;
  d.bslot = h->d + o_bs;
  d.nsf = nsf;
  d.nsb = nsb;
  d.el_slot = h->d + o_es;
  d.bel_slot = h->d + o_bes;
  d.nel = nidx;
  d.nbel = (int)bel_k.size();
  d.finfo = (const uint32_t *)(h->d + o_fi);
  d.binfo = (const uint32_t *)(h->d + o_bi);
  d.wave = wave;
  d.ilog = h->d + o_l;
  d.has_el = (const uint8_t *)(h->d + o_h);
  *out = h;
  return IXG_OK;
}

int ixg_cp9_free(ixg_ctx *c, ixg_cp9 *h) {
  if (!h) return IXG_OK;
  if (c) dfree(c, h->d, h->bytes);
  delete h;
  return IXG_OK;
}

int ixg_cp9_bands(ixg_ctx *c, const ixg_cp9 *h, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0, double tau,
                  int mode, int32_t *fmx, int32_t *fix, int32_t *fdx, int32_t *bmx, int32_t *bix, int32_t *bdx,
                  float *fwd_sc, float *bwd_sc, int32_t *pn) {
  if (!c) return IXG_EINVAL;
  RCHK(check_cp9_window(h, dsq, L, i0, j0));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)(L + 2), (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)(L + 2), hipMemcpyHostToDevice, c->stream));
  int32_t *mats[6] = {fmx, fix, fdx, bmx, bix, bdx};
  return cp9_run(c, h, {CWin{d, i0, j0}}, tau, mode, fwd_sc, bwd_sc, pn, mats);
}

int ixg_cp9_bands_batch(ixg_ctx *c, const ixg_cp9 *h, const ixg_target *tgt, const ixg_cp9win *win, int64_t n,
                        double tau, int mode, float *fwd_sc, float *bwd_sc, int32_t *pn) {
  if (!c || !h || !tgt || (n > 0 && (!win || !pn)) || n < 0) {
    g_err = "ixg_cp9_bands_batch: bad arguments";
    return IXG_EINVAL;
  }
  int64_t L0;
  const uint8_t *base = ixg_target_device_base(tgt, &L0);
  std::vector<CWin> w(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t off = 0, L = 0;
    if ((win[k].strand != 0 && win[k].strand != 1) || !ixg_target_view(tgt, win[k].seq, win[k].strand, &off, &L, nullptr) ||
        win[k].i0 < 1 || win[k].j0 > L || win[k].j0 < win[k].i0) {
      g_err = "ixg_cp9_bands_batch: window outside the target";
      return IXG_EINVAL;
    }
    w[k] = CWin{base + off, win[k].i0, win[k].j0};
  }
  return cp9_run(c, h, w, tau, mode, fwd_sc, bwd_sc, pn, nullptr);
}

}  // extern "C"

namespace ixg {
// Posterior HMM bands pn [n][6][M+1] of windows dsq[k][i0[k]..j0[k]] (device
// sequences), the batched DP of cp9_seq2bands (cp9.rs:4053).
// occ (optional): the per-node posterior occupancy [n][M+1] as well.
int cp9_posterior_bands(ixg_ctx *c, const ixg_cp9 *h, const std::vector<const uint8_t *> &dsq,
                        const std::vector<int32_t> &i0, const std::vector<int32_t> &j0, double tau, int mode,
                        std::vector<int32_t> &pn, std::vector<int32_t> *occ) {
  const size_t n = dsq.size();
  std::vector<CWin> w(n);
  for (size_t k = 0; k < n; k++) w[k] = CWin{dsq[k], i0[k], j0[k]};
  pn.assign(n * 6 * (size_t)(h->M + 1), 0);
  if (occ) occ->assign(n * (size_t)(h->M + 1), 0);
  if (n == 0) return IXG_OK;
  return cp9_run(c, h, w, tau, mode, nullptr, nullptr, pn.data(), nullptr, occ ? occ->data() : nullptr);
}
int cp9_nodes(const ixg_cp9 *h) { return h->M; }
}  // namespace ixg

// C ABI of the HMM-banded CYK / Inside scanners: resident CM score tables,
// band validation and deck layout (cp9.rs:4741-4763), batched launches, and
// the host side of the reference's reporting (vsc_root and the envelope,
// cp9.rs:5305-5325; ReportHitsGreedily with null3, cp9.rs:5950-5997).
#include <cmath>
#include <cstring>
#include <memory>
#include "cm_internal.h"
#include "ctx.h"
#include "cm_handle.h"

using namespace ixg;

namespace {

int sdelta_host(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }

// Deck layout of one envelope: roff per (v, jp) and the cell count, after
// checking that every cell stays inside [i0, j0] and that the d bands respect
// the states' emission counts (as CP9-derived bands do, cp9.rs:3815-3850).
int sldelta_host(int st) { return (st == MP_ST || st == ML_ST || st == IL_ST) ? 1 : 0; }
int srdelta_host(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

// Smallest d a band of state type st may start at: the emission count, or for
// a truncated scan the larger of the left and right counts.
int dmin_host(int st, bool trunc) {
  if (st == E_ST) return 0;
  return trunc ? std::max(sldelta_host(st), srdelta_host(st)) : sdelta_host(st);
}

// Checks one envelope's bands and lays out its deck: roff[r] = first cell
// of band row r (rows in state order), ncells, nrows. On error returns
// IXG_EINVAL with *err set (g_err is left alone: this runs on host workers).
int layout(const CmHost &h, const ixg_bands &b, int32_t i0, int32_t j0, bool trunc, std::vector<int32_t> &roff,
           int64_t &ncells, int32_t &nrows, const char **err) {
  const int M = h.M;
  if (!b.jmin || !b.jmax || !b.hd_off || !b.hdmin || !b.hdmax || i0 < 1 || j0 < i0 - 1) {
    *err = "bands: null array or bad envelope";
    return IXG_EINVAL;
  }
  nrows = b.hd_off[M];
  roff.assign(nrows > 0 ? nrows : 1, 0);
  int64_t nc = 0;
  for (int v = 0; v < M; v++) {
    const int nj = b.jmax[v] >= b.jmin[v] ? b.jmax[v] - b.jmin[v] + 1 : 0;
    if (b.hd_off[v] < 0 || b.hd_off[v] + nj > nrows || (v + 1 <= M && b.hd_off[v + 1] < b.hd_off[v] + nj)) {
      *err = "bands: hd_off inconsistent with the j bands";
      return IXG_EINVAL;
    }
    if (nj && (b.jmin[v] < i0 - 1 || b.jmax[v] > j0)) {
      *err = "bands: j band outside the envelope";
      return IXG_EINVAL;
    }
    const int st = h.sttype[v];
    for (int jp = 0; jp < nj; jp++) {
      const int r = b.hd_off[v] + jp, j = b.jmin[v] + jp;
      roff[r] = (int32_t)nc;
      const int lo = b.hdmin[r], hi = b.hdmax[r];
      if (hi >= lo) {
        if (lo < dmin_host(st, trunc) || j - hi + 1 < i0 || (st == E_ST && (lo != 0 || hi != 0))) {
          *err = "bands: d band outside the envelope or below the state's emission count";
          return IXG_EINVAL;
        }
        nc += hi - lo + 1;
      } else if (st == E_ST) {
        *err = "bands: empty row of an end state";
        return IXG_EINVAL;
      }
    }
  }
  if (nc >= (trunc ? INT32_MAX / 4 : INT32_MAX)) {
    *err = "bands: deck larger than 2^31 cells";
    return IXG_EINVAL;
  }
  ncells = nc;
  return IXG_OK;
}

// vsc_root and the envelope bounds from the root deck (cp9.rs:5305-5325).
void root_summary(const ixg_bands &b, const float *root, int32_t i0, int32_t j0, float env_cutoff, float &vsc,
                  int64_t &envi, int64_t &envj) {
  float vr = IMPOSSIBLE;
  int64_t ei = (int64_t)j0 + 1, ej = (int64_t)i0 - 1;
  int64_t off = 0;
  for (int j = b.jmin[0]; j <= b.jmax[0]; j++) {
    const int r = b.hd_off[0] + (j - b.jmin[0]);
    for (int d = b.hdmin[r]; d <= b.hdmax[r]; d++) {
      const float s = root[off + (d - b.hdmin[r])];
      if (s > vr) vr = s;
      if (s >= env_cutoff) {
        ei = std::min<int64_t>(ei, j - d + 1);
        ej = std::max<int64_t>(ej, j);
      }
    }
    if (b.hdmax[r] >= b.hdmin[r]) off += b.hdmax[r] - b.hdmin[r] + 1;
  }
  vsc = vr;
  envi = ei == (int64_t)j0 + 1 ? -1 : ei;
  envj = ej == (int64_t)i0 - 1 ? -1 : ej;
}

float kahan(const float *v, int n) {
  float sum = 0.f, c = 0.f;
  for (int i = 0; i < n; i++) {
    float y = v[i] - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

// cp9.rs:6516-6529 score_correction_null3
float null3(const ixg_cm &cm, const float *comp, int len) { return null3_score(cm.h, cm.ftbl_host, comp, len); }

// ReportHitsGreedily over the root deck (cp9.rs:5887-5900, 5950-5997).
void report_hits(const ixg_cm &cm, const ixg_bands &b, const uint8_t *dsq, int32_t i0, int32_t j0,
                 const float *root, float cutoff, int do_null3, int32_t env, std::vector<ixg_hit> &out) {
  const int W = j0 - i0 + 1;
  std::vector<double> act;
  if (do_null3) {
    act.assign((size_t)(W + 1) * 4, 0.0);
    for (int jp = 1; jp <= W; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      const int r = dsq[i0 + jp - 1];
      if (r < 4) act[jp * 4 + r] += 1.0;
    }
  }
  int64_t off = 0;
  for (int j = b.jmin[0]; j <= b.jmax[0]; j++) {
    const int r = b.hd_off[0] + (j - b.jmin[0]);
    const int dmin = b.hdmin[r], dmax = b.hdmax[r];
    if (dmin <= dmax) {
      float maxrep = IMPOSSIBLE;
      for (int d = std::max(dmin, 1); d <= dmax; d++) {
        const int i = j - d + 1;
        float hs = root[off + (d - dmin)];
        if (hs > maxrep && hs >= cutoff && hs > -9.999e35f) {
          bool rep = true;
          float bias = 0.f;
          if (do_null3) {
            float comp[4];
            const int ip = i - i0 + 1, jj = j - i0 + 1;
            for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
            const float s = kahan(comp, 4);
            if (s != 0.f)
              for (int a = 0; a < 4; a++) comp[a] /= s;
            else
              for (int a = 0; a < 4; a++) comp[a] = 1.0f / 4.0f;
            const float corr = null3(cm, comp, d);
            hs -= corr;
            bias = corr;
            rep = hs > maxrep && hs >= cutoff;
          }
          if (rep) {
            out.push_back({i, j, hs, bias, env});
            maxrep = hs;
          }
        }
      }
      off += dmax - dmin + 1;
    }
  }
}

struct EnvPlan {
  int64_t ncells;  // band cells (what a scan computes and reports)
  int32_t nrows;
  int64_t nroot;
  std::vector<int32_t> roff;
  // standard scans: the deck's state rectangles, root row offsets, deck
  // floats and wavefront steps
  std::vector<int32_t> rect, rootoff;
  int64_t deck = 0;
  int32_t nsteps = 0;
};

// The standard scan's deck and schedule for one envelope (see cm_kernels.hip):
// per state its band's rectangle, diagonal-major (d, x) with x counting j or
// i, whichever band width is smaller, and the number of wavefront steps (the
// largest d + 1 plus the root's level).
int rect_plan(const CmHost &h, const ixg_bands &b, EnvPlan &p, const char **err) {
  const int M = h.M;
  p.rect.assign((size_t)HB_RECT * M, 0);
  int64_t base = 0;
  int dmax = -1, maxlvl = 0;
  for (int v = 0; v < M; v++) {
    const int nj = b.jmax[v] >= b.jmin[v] ? b.jmax[v] - b.jmin[v] + 1 : 0;
    int ilo = INT32_MAX, ihi = INT32_MIN, dlo = INT32_MAX, dhi = INT32_MIN;
    for (int jp = 0; jp < nj; jp++) {
      const int r = b.hd_off[v] + jp, j = b.jmin[v] + jp, lo = b.hdmin[r], hi = b.hdmax[r];
      if (hi < lo) continue;
      ilo = std::min(ilo, j - hi + 1);
      ihi = std::max(ihi, j - lo + 1);
      dlo = std::min(dlo, lo);
      dhi = std::max(dhi, hi);
    }
    int32_t *R = p.rect.data() + (size_t)HB_RECT * v;
    R[0] = b.jmin[v];
    R[8] = b.hd_off[v];
    // type | child count << 4 | first child << 8 (a bifurcation's fields are
    // not used: the kernel reads its two children from the model)
    R[9] = (int32_t)((uint32_t)h.sttype[v] | (h.sttype[v] == B_ST ? 0u : (uint32_t)h.cnum[v] << 4 |
                                                                        (uint32_t)h.cfirst[v] << 8));
    if (dhi < dlo) {  // no band cells: no deck
      R[5] = h.lvl[v] * 4;
      continue;
    }
    const int ni = ihi - ilo + 1, mode = nj <= ni ? 1 : 0, nx = mode ? nj : ni, nd = dhi - dlo + 1;
    // a box band: every row's d band is its (i, j) box row clipped to
    // dlo..dhi (what ij2d_bands makes of i and j bands); then a rectangle
    // cell is a band cell exactly when its i is in ilo..ihi, with no row
    // lookup
    bool box = true;
    for (int jp = 0; jp < nj && box; jp++) {
      const int r = b.hd_off[v] + jp, j = b.jmin[v] + jp, lo = b.hdmin[r], hi = b.hdmax[r];
      const int elo = std::max(j - ihi + 1, dlo), ehi = std::min(j - ilo + 1, dhi);
      const bool emp = hi < lo, eemp = ehi < elo;
      box = emp == eemp && (emp || (lo == elo && hi == ehi));
    }
    R[1] = nj;
    R[2] = ilo;
    R[3] = nx;
    R[4] = (int32_t)base;
    R[5] = h.lvl[v] * 4 + (box ? 2 : 0) + mode;
    R[6] = dlo;
    R[7] = nd;
    R[10] = ni;
    base += (int64_t)nd * nx;
    if (base >= INT32_MAX) {
      *err = "bands: deck larger than 2^31 cells";
      return IXG_EINVAL;
    }
    dmax = std::max(dmax, dhi);
    maxlvl = std::max(maxlvl, h.lvl[v]);
  }
  p.deck = base;
  p.nsteps = dmax >= 0 ? dmax + 1 + maxlvl : 0;
  const int nj0 = p.rect[1];
  p.rootoff.assign(std::max(nj0, 1), 0);
  for (int jp = 0; jp < nj0; jp++) p.rootoff[jp] = p.roff[b.hd_off[0] + jp] - p.roff[b.hd_off[0]];
  if (M > 65535) {
    *err = "standard scan: more than 65535 states";
    return IXG_EINVAL;
  }
  return IXG_OK;
}

// Every band row's d band (hdmin, hdmax), 2 ints per row at out.
void rect_rows(const ixg_bands &b, int nrows, int32_t *out) {
  for (int r = 0; r < nrows; r++) {
    out[2 * (size_t)r] = b.hdmin[r];
    out[2 * (size_t)r + 1] = b.hdmax[r];
  }
}


// A truncated scan's pass: the mode bits it fills and its device penalties
// and entry-state list.
struct TrSel {
  int fill;  // BL | BR | BT bits enabled by the pass
  const float *d_pty;
  const int32_t *d_yl;
  int ny;
};

int mode_mask(const ixg_trvalid &val, int v, int fill) {
  return (val.jvalid[v] ? 1 : 0) | ((val.lvalid[v] && (fill & 2)) ? 2 : 0) | ((val.rvalid[v] && (fill & 4)) ? 4 : 0) |
         ((val.tvalid[v] && (fill & 8)) ? 8 : 0);
}

// Per state, the work units of its level step (cells; rows of an IL state;
// i diagonals of an IR state; 0 for a state the scan skips), then the first
// i of each IR state: 2*M values written at units.
void append_units(const CmHost &h, const ixg_bands &b, const ixg_trvalid *val, int fill, int32_t *units) {
  const int M = h.M;
  int32_t *ib = units + M;
  std::fill(ib, ib + M, 0);
  for (int v = 0; v < M; v++) {
    const int nj = b.jmax[v] >= b.jmin[v] ? b.jmax[v] - b.jmin[v] + 1 : 0;
    int64_t n = 0;
    const bool skip = val && (v == 0 || mode_mask(*val, v, fill) == 0);
    const int st = h.sttype[v];
    if (nj > 0 && !skip) {
      if (st == IL_ST) {
        n = nj;
      } else if (st == IR_ST) {
        int imin = INT32_MAX, imax = INT32_MIN;
        for (int jp = 0; jp < nj; jp++) {
          const int r = b.hd_off[v] + jp;
          if (b.hdmax[r] < b.hdmin[r]) continue;
          imin = std::min(imin, b.jmin[v] + jp - b.hdmax[r] + 1);
          imax = std::max(imax, b.jmin[v] + jp - b.hdmin[r] + 1);
        }
        if (imax >= imin) {
          n = imax - imin + 1;
          ib[v] = imin;
        }
      } else {
        for (int jp = 0; jp < nj; jp++) {
          const int r = b.hd_off[v] + jp;
          if (b.hdmax[r] >= b.hdmin[r]) n += b.hdmax[r] - b.hdmin[r] + 1;
        }
      }
    }
    units[v] = (int32_t)n;
  }
}

// One or more envelopes over a device sequence buffer: uploads bands and
// layouts, launches the scan in groups that fit the deck budget, and returns
// every envelope's root rows on the host (standard: nroot floats; truncated:
// the J, L, R, T root planes then the best entry per root cell, nroot each).
// One persistent launch over envelopes k0..k1 of several models: the
// distinct models' descriptors, each envelope's model index, and the claim
// order (largest deck first) go to the device; the grid is two workgroups
// per CU (fewer when there are fewer envelopes).
int launch_multi(ixg_ctx *c, const std::vector<const ixg_cm *> &percm, std::vector<EnvDev> &ed,
                 const std::vector<EnvPlan> &plan, int64_t k0, int64_t k1, EnvDev *d_env, int inside,
                 size_t lds_ints) {
  std::vector<const ixg_cm *> uniq;
  std::vector<CmDev> devs;
  for (int64_t k = k0; k < k1; k++) {
    const ixg_cm *m = percm[k];
    auto it = std::find(uniq.begin(), uniq.end(), m);
    if (it == uniq.end()) {
      uniq.push_back(m);
      devs.push_back(m->dev());
      it = uniq.end() - 1;
    }
    ed[k - k0].cmi = (int32_t)(it - uniq.begin());
  }
  std::vector<int32_t> order(k1 - k0);
  for (int64_t k = 0; k < k1 - k0; k++) order[k] = (int32_t)k;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return plan[k0 + a].ncells > plan[k0 + b].ncells; });
  const size_t dev_bytes = devs.size() * sizeof(CmDev), ord_bytes = order.size() * 4;
  char *d_meta;
  RCHK(scratch(c, S_CM_SCHED, dev_bytes + ord_bytes + 256, (void **)&d_meta));
  CmDev *d_cms = (CmDev *)d_meta;
  int32_t *d_order = (int32_t *)(d_meta + dev_bytes);
  uint32_t *d_counter = (uint32_t *)(d_meta + ((dev_bytes + ord_bytes + 127) / 128) * 128);
  HIPCHK(hipMemcpyAsync(d_env, ed.data(), ed.size() * sizeof(EnvDev), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_cms, devs.data(), dev_bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_order, order.data(), ord_bytes, hipMemcpyHostToDevice, c->stream));
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  HIPCHK(launch_hb_scan_queue(d_cms, d_env, d_order, k1 - k0, d_counter, std::max(cus, 1), uniq[0]->d_ftbl(),
                              inside, lds_ints, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // the copies above read host vectors local to this call
  return IXG_OK;
}

// percm (standard scans only): the model of each envelope, run as one
// persistent launch over a work queue; null runs every envelope against cm.
int run_envelopes(ixg_ctx *c, const ixg_cm *cm, const uint8_t *d_seq_base, const std::vector<int64_t> &seq_off,
                  const ixg_envelope *env, int64_t n, int inside, const TrSel *tr, const ixg_trvalid *valid,
                  std::vector<std::vector<float>> &roots, int64_t *cells_out,
                  const std::vector<const ixg_cm *> *percm = nullptr) {
  auto cmk = [&](int64_t k) { return percm ? (*percm)[k] : cm; };
  const bool trunc = tr != nullptr;
  const int planes = trunc ? 4 : 1, rper = trunc ? 5 : 1;
  std::vector<EnvPlan> plan(n);
  std::vector<const char *> err(n, nullptr);
  // per envelope, on host workers: band checks, deck layout, root row count
  par_for(n, [&](int64_t k) {
    if (layout(cmk(k)->h, env[k].bands, env[k].i0, env[k].j0, trunc, plan[k].roff, plan[k].ncells, plan[k].nrows,
               &err[k]) != IXG_OK)
      return;
    if (!trunc && rect_plan(cmk(k)->h, env[k].bands, plan[k], &err[k]) != IXG_OK) return;
    if (trunc && (!valid[k].jvalid || !valid[k].lvalid || !valid[k].rvalid || !valid[k].tvalid)) {
      err[k] = "truncated scan: null validity array";
      return;
    }
    const ixg_bands &b = env[k].bands;
    int64_t nr = 0;
    for (int j = b.jmin[0]; j <= b.jmax[0]; j++) {
      const int r = b.hd_off[0] + (j - b.jmin[0]);
      if (b.hdmax[r] >= b.hdmin[r]) nr += b.hdmax[r] - b.hdmin[r] + 1;
    }
    plan[k].nroot = nr;
  });
  int64_t total_cells = 0;
  for (int64_t k = 0; k < n; k++) {
    if (err[k]) {
      g_err = err[k];
      return IXG_EINVAL;
    }
    total_cells += plan[k].ncells;
  }
  if (cells_out) *cells_out = total_cells;
  roots.assign(n, {});
  const int64_t budget = arena_elems(c, S_CM_ALPHA);  // deck floats per launch group
  int64_t k0 = 0;
  while (k0 < n) {
    auto deck = [&](int64_t k) { return trunc ? planes * plan[k].ncells : plan[k].deck; };
    int64_t k1 = k0, floats = 0;
    while (k1 < n && (k1 == k0 || floats + deck(k1) <= budget)) floats += deck(k1++);
    // pack per envelope, staged in pinned memory and filled on host workers.
    // Truncated: jmin|jmax (M each), hd_off (M+1), hdmin|hdmax|roff (nrows
    // each), mode mask (M), units (M), first IR i (M). Standard: rectangles
    // (HB_RECT ints per state), begin states and scores (2 nbegin), root row
    // offsets (root rows), row d bands (2 nrows).
    std::vector<int64_t> at(k1 - k0 + 1), rat(k1 - k0);
    int64_t rtotal = 0;
    size_t lds_ints = 0, scan_ints = 0;
    at[0] = 0;
    for (int64_t k = k0; k < k1; k++) {
      const int M = cmk(k)->h.M;
      const int64_t nr = trunc ? plan[k].nrows : std::max<int64_t>(plan[k].nrows, 1);
      const int64_t nb = cmk(k)->h.local_begin ? (int64_t)cmk(k)->h.begin_list.size() : 0;
      at[k - k0 + 1] = at[k - k0] + (trunc ? 3 * (int64_t)M + 1 + 3 * nr + M + 2 * M
                                           : HB_RECT * (int64_t)M + 2 * nb + (int64_t)plan[k].rootoff.size() + 2 * nr);
      rat[k - k0] = rtotal;
      rtotal += rper * plan[k].nroot;
      if (!trunc) {
        const size_t sc = (size_t)hb_scan_ints(M) + hb_seq_ints(env[k].j0 - env[k].i0 + 1);
        scan_ints = std::max(scan_ints, sc);
        lds_ints = std::max(lds_ints, sc + HB_RECT * M + 2 * nb + (inside ? FLOGSUM_TBL : 0));
      }
    }
    // LDS always holds the per-step scan and the residues; the Inside FLogsum
    // table, rectangle table and begin list too when they fit (else the
    // kernel reads them from memory)
    constexpr size_t kMaxLdsInts = 156 * 1024 / 4;
    if (!trunc && scan_ints > kMaxLdsInts) {
      g_err = "standard scan: too many states for the LDS step scan";
      return IXG_EINVAL;
    }
    lds_ints = std::min(lds_ints, kMaxLdsInts);
    const int64_t pack_n = at[k1 - k0];
    int32_t *pack;
    RCHK(pinned(c, P_PACK, (size_t)pack_n * 4, (void **)&pack));
    int32_t *d_pack;
    float *d_alpha, *d_rout;
    EnvDev *d_env;
    RCHK(scratch(c, S_CM_BANDS, (size_t)pack_n * 4, (void **)&d_pack));
    RCHK(scratch(c, S_CM_ALPHA, (size_t)std::max<int64_t>(floats, 1) * 4, (void **)&d_alpha));
    RCHK(scratch(c, S_CM_ENV, (size_t)(k1 - k0) * sizeof(EnvDev), (void **)&d_env));
    RCHK(scratch(c, S_CM_ROOT, (size_t)std::max<int64_t>(rtotal, 1) * 4, (void **)&d_rout));
    std::vector<EnvDev> ed(k1 - k0);
    int64_t aoff = 0;
    for (int64_t k = k0; k < k1; k++) {
      const int M = cmk(k)->h.M;
      const int nr = trunc ? plan[k].nrows : std::max(plan[k].nrows, 1);
      const int32_t *base = d_pack + at[k - k0];
      EnvDev &e = ed[k - k0];
      e = EnvDev{};
      e.dsq = d_seq_base + seq_off[k];
      if (trunc) {
        e.jmin = base;
        e.jmax = base + M;
        e.hd_off = base + 2 * M;
        e.hdmin = base + 3 * M + 1;
        e.hdmax = e.hdmin + nr;
        e.roff = e.hdmax + nr;
        e.mask = e.roff + nr;
        e.units = e.roff + nr + M;
        e.ibase = e.units + M;
      } else {
        const int nb = cmk(k)->h.local_begin ? (int)cmk(k)->h.begin_list.size() : 0;
        e.rect = base;
        e.nbegin = nb;
        e.rootoff = base + HB_RECT * M + 2 * nb;
        e.rowx = e.rootoff + plan[k].rootoff.size();
        e.nsteps = plan[k].nsteps;
      }
      e.alpha = d_alpha + aoff;
      e.nc = (int32_t)plan[k].ncells;
      e.rout = d_rout + rat[k - k0];
      e.i0 = env[k].i0;
      e.j0 = env[k].j0;
      aoff += deck(k);
    }
    HIPCHK(hipMemcpyAsync(d_env, ed.data(), ed.size() * sizeof(EnvDev), hipMemcpyHostToDevice, c->stream));
    // the band upload in parts of about equal size: the host packs part s+1
    // (into its own part of the pinned buffer) while part s is copied
    constexpr int NSUB = 4;
    std::vector<int64_t> sub{0};
    for (int64_t kk = 0; kk < k1 - k0; kk++)
      if ((int)sub.size() < NSUB && at[kk + 1] * NSUB >= pack_n * (int64_t)sub.size() && kk + 1 < k1 - k0)
        sub.push_back(kk + 1);
    sub.push_back(k1 - k0);
    for (size_t g = 0; g + 1 < sub.size(); g++) {
      const int64_t a0 = sub[g], a1 = sub[g + 1];
      par_for(a1 - a0, [&](int64_t t) {
        const int64_t kk = a0 + t, k = k0 + kk;
        const ixg_bands &b = env[k].bands;
        const int M = cmk(k)->h.M;
        const int nr = plan[k].nrows;
        int32_t *o = pack + at[kk];
        if (!trunc) {
          const EnvPlan &pk = plan[k];
          const CmHost &h = cmk(k)->h;
          o = std::copy(pk.rect.begin(), pk.rect.end(), o);
          if (h.local_begin) {
            o = std::copy(h.begin_list.begin(), h.begin_list.end(), o);
            for (int32_t y : h.begin_list) {
              const float sc = h.beginsc[y];
              std::memcpy(o++, &sc, 4);
            }
          }
          o = std::copy(pk.rootoff.begin(), pk.rootoff.end(), o);
          rect_rows(b, nr, o);
          return;
        }
        o = std::copy(b.jmin, b.jmin + M, o);
        o = std::copy(b.jmax, b.jmax + M, o);
        o = std::copy(b.hd_off, b.hd_off + M + 1, o);
        o = std::copy(b.hdmin, b.hdmin + nr, o);
        o = std::copy(b.hdmax, b.hdmax + nr, o);
        o = std::copy(plan[k].roff.begin(), plan[k].roff.begin() + nr, o);
        for (int v = 0; v < M; v++) *o++ = mode_mask(valid[k], v, tr->fill);
        append_units(cmk(k)->h, b, &valid[k], tr->fill, o);
      });
      HIPCHK(hipMemcpyAsync(d_pack + at[a0], pack + at[a0], (size_t)(at[a1] - at[a0]) * 4, hipMemcpyHostToDevice,
                            c->stream));
    }
    HIPCHK(hipEventRecord(c->sev[0], c->stream));
    if (trunc)
      HIPCHK(launch_trhb_scan(cm->dev(), d_env, k1 - k0, cm->d_ftbl(), tr->d_pty, tr->d_yl, tr->ny, inside,
                              c->stream));
    else if (percm)
      RCHK(launch_multi(c, *percm, ed, plan, k0, k1, d_env, inside, lds_ints));
    else
      HIPCHK(launch_hb_scan(cm->dev(), d_env, k1 - k0, cm->d_ftbl(), inside, lds_ints, c->stream));
    HIPCHK(hipEventRecord(c->sev[1], c->stream));
    float *rhost;
    RCHK(pinned(c, P_DOWN, (size_t)std::max<int64_t>(rtotal, 1) * 4, (void **)&rhost));
    if (rtotal) HIPCHK(hipMemcpyAsync(rhost, d_rout, rtotal * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t k = k0; k < k1; k++)
      roots[k].assign(rhost + rat[k - k0], rhost + rat[k - k0] + rper * plan[k].nroot);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->sev[0], c->sev[1]));
    c->ms[2] += ms;
    c->kms[trunc ? IXG_K_TRHB_SCAN : IXG_K_HB_SCAN] += ms;
    c->kn[trunc ? IXG_K_TRHB_SCAN : IXG_K_HB_SCAN] += 1;
    k0 = k1;
  }
  return IXG_OK;
}

// ---- truncated scans: pass selection, root summary, reporting ----

int pty_index(int pass_idx) {  // cm_trunc.rs:99-107
  switch (pass_idx) {
    case IXG_PASS_5P_ONLY_FORCE: return 1;
    case IXG_PASS_3P_ONLY_FORCE: return 2;
    case IXG_PASS_5P_AND_3P_FORCE:
    case IXG_PASS_5P_AND_3P_ANY: return 0;
    default: return -1;
  }
}
int fill_bits(int pass_idx) {  // cm_dpsearch_trunc.rs:184-199
  if (pass_idx == IXG_PASS_5P_AND_3P_FORCE || pass_idx == IXG_PASS_5P_AND_3P_ANY) return 2 | 4 | 8;
  if (pass_idx == IXG_PASS_5P_ONLY_FORCE) return 4;
  if (pass_idx == IXG_PASS_3P_ONLY_FORCE) return 2;
  return 0;
}

int tr_select(const ixg_cm *cm, int pass_idx, int local, TrSel &sel) {
  const int idx = pty_index(pass_idx);
  if (idx < 0) {
    g_err = "truncated scan: pass_idx must be a truncated pass (2..5)";
    return IXG_EINVAL;
  }
  if (!cm->d_tr) {
    g_err = "truncated scan: ixg_cm_set_trpenalties has not been called";
    return IXG_EINVAL;
  }
  const int q = (local ? 3 : 0) + idx;
  sel.fill = fill_bits(pass_idx);
  sel.d_pty = cm->d_tr + (size_t)q * cm->h.M;
  sel.d_yl = cm->d_yl + cm->yoff[q];
  sel.ny = cm->yoff[q + 1] - cm->yoff[q];
  return IXG_OK;
}

// vsc_root / vmode over the four root planes and the envelope bounds
// (cm_dpsearch_trunc.rs:989-1022).
void tr_root_summary(const ixg_bands &b, const float *rr, int64_t nroot, int m0, int32_t i0, int32_t j0,
                     float env_cutoff, float &vsc, int32_t &vmode, int64_t &envi, int64_t &envj) {
  static const int modes[4] = {IXG_TRMODE_J, IXG_TRMODE_L, IXG_TRMODE_R, IXG_TRMODE_T};
  float vr = IMPOSSIBLE;
  int32_t vm = IXG_TRMODE_UNKNOWN;
  int64_t ei = (int64_t)j0 + 1, ej = (int64_t)i0 - 1;
  int64_t off = 0;
  for (int j = b.jmin[0]; j <= b.jmax[0]; j++) {
    const int r = b.hd_off[0] + (j - b.jmin[0]);
    for (int d = b.hdmin[r]; d <= b.hdmax[r]; d++) {
      const int64_t c = off + (d - b.hdmin[r]);
      bool in_env = false;
      for (int p = 0; p < 4; p++) {
        if (!(m0 & (1 << p))) continue;
        const float s = rr[p * nroot + c];
        if (s > vr) vr = s, vm = modes[p];
        if (s >= env_cutoff) in_env = true;
      }
      if (in_env) {
        ei = std::min<int64_t>(ei, j - d + 1);
        ej = std::max<int64_t>(ej, j);
      }
    }
    if (b.hdmax[r] >= b.hdmin[r]) off += b.hdmax[r] - b.hdmin[r] + 1;
  }
  vsc = vr;
  vmode = vm;
  envi = ei == (int64_t)j0 + 1 ? -1 : ei;
  envj = ej == (int64_t)i0 - 1 ? -1 : ej;
}

// cm_dpsearch_trunc.rs:2376-2422 allow_truncation
bool allow_truncation(const CmHost &h, int pass_idx, int local, int start, int stop, int i0, int j0, int mode, int b) {
  if (pass_idx == IXG_PASS_STD_ANY || pass_idx == IXG_PASS_5P_AND_3P_ANY) return true;
  if (start == i0 && stop == j0) return true;
  const int lp = h.lpos[b], rp = h.rpos[b];
  if (local) {
    if (mode == IXG_TRMODE_J) return true;
    if (mode == IXG_TRMODE_L) return stop == j0;
    if (mode == IXG_TRMODE_R) return start == i0;
    return false;
  }
  if (mode == IXG_TRMODE_J) return lp == 1 && rp == h.clen;
  if (mode == IXG_TRMODE_L) return stop == j0 && lp == 1;
  if (mode == IXG_TRMODE_R) return start == i0 && rp == h.clen;
  return false;
}

// Greedy truncated reporting per j (cm_dpsearch_trunc.rs:2322-2372) over the
// root planes, then greedy overlap removal (:2470-2495); appends to out.
void tr_report(const ixg_cm &cm, const ixg_bands &b, const uint8_t *dsq, int32_t i0, int32_t j0, const float *rr,
               int64_t nroot, int pass_idx, int local, float cutoff, int do_null3, int32_t env,
               std::vector<ixg_trhit> &out) {
  const int W = j0 - i0 + 1;
  std::vector<double> act;
  if (do_null3) {
    act.assign((size_t)(W + 1) * 4, 0.0);
    for (int jp = 1; jp <= W; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      const int r = dsq[i0 + jp - 1];
      if (r < 4) act[jp * 4 + r] += 1.0;
    }
  }
  std::vector<ixg_trhit> tmp;
  int64_t off = 0;
  for (int j = b.jmin[0]; j <= b.jmax[0]; j++) {
    const int r = b.hd_off[0] + (j - b.jmin[0]);
    const int dmin = b.hdmin[r], dmax = b.hdmax[r];
    if (dmin > dmax) continue;
    float maxrep = IMPOSSIBLE;
    for (int d = std::max(dmin, 1); d <= dmax; d++) {
      const int64_t c = off + (d - dmin);
      float hs = IMPOSSIBLE;
      for (int p = 0; p < 4; p++) hs = std::max(hs, rr[p * nroot + c]);
      int32_t code;
      memcpy(&code, &rr[4 * nroot + c], 4);
      const int mode = code & 7, root = code >> 3;
      const int i = j - d + 1;
      if (!(hs > maxrep && hs >= cutoff && hs > IMPOSSIBLE + 1.0f)) continue;
      bool rep = allow_truncation(cm.h, pass_idx, local, i, j, i0, j0, mode, root);
      float bias = 0.f;
      if (rep && do_null3) {
        float comp[4];
        const int ip = i - i0 + 1, jj = j - i0 + 1;
        for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
        float sum = 0.f;
        for (int a = 0; a < 4; a++) sum += comp[a];
        if (sum > 0.f)
          for (int a = 0; a < 4; a++) comp[a] /= sum;
        else
          for (int a = 0; a < 4; a++) comp[a] = 1.0f / 4.0f;
        const float corr = null3(cm, comp, d);
        hs -= corr;
        bias = corr;
        rep = hs > maxrep && hs >= cutoff;
      }
      if (rep) {
        tmp.push_back({i, j, hs, bias, mode, root, env});
        maxrep = hs;
      }
    }
    off += dmax - dmin + 1;
  }
  std::stable_sort(tmp.begin(), tmp.end(), [](const ixg_trhit &a, const ixg_trhit &b) {
    if (a.score != b.score) return a.score > b.score;
    return a.i < b.i;
  });
  std::vector<char> removed(tmp.size(), 0);
  for (size_t a = 0; a < tmp.size(); a++) {
    if (removed[a]) continue;
    for (size_t k = a + 1; k < tmp.size(); k++)
      if (!removed[k] && !(tmp[k].j < tmp[a].i) && !(tmp[a].j < tmp[k].i)) removed[k] = 1;
  }
  for (size_t a = 0; a < tmp.size(); a++)
    if (!removed[a]) out.push_back(tmp[a]);
}

}  // namespace

extern "C" {

int ixg_cm_create(ixg_ctx *c, const ixg_cm_params *p, ixg_cm **out) { return ixg_cm_create_cfg(c, p, 1, out); }

int ixg_cm_create_cfg(ixg_ctx *c, const ixg_cm_params *p, int local, ixg_cm **out) {
  if (!p || !out || p->M < 1 || p->nodes < 2 || !p->sttype || !p->ndtype || !p->nodemap || !p->cfirst || !p->cnum ||
      !p->t || !p->e) {
    g_err = "ixg_cm_create: bad arguments";
    return IXG_EINVAL;
  }
  for (int nd = 0; nd < p->nodes; nd++) {  // nodes in order, each owning a run of states
    if (p->ndtype[nd] < BIF_ND || p->ndtype[nd] > END_ND || p->nodemap[nd] < 0 || p->nodemap[nd] >= p->M ||
        (nd == 0 && p->nodemap[nd] != 0) || (nd > 0 && p->nodemap[nd] <= p->nodemap[nd - 1]) ||
        ((p->ndtype[nd] == BIF_ND) != (p->sttype[p->nodemap[nd]] == B_ST))) {
      g_err = "ixg_cm_create: bad node types or node map";
      return IXG_EINVAL;
    }
  }
  for (int v = 0; v < p->M; v++) {  // children must follow their parent
    const int st = p->sttype[v];
    if (st < 0 || st > B_ST) {
      g_err = "ixg_cm_create: bad state type";
      return IXG_EINVAL;
    }
    if (st == B_ST) {
      if (p->cfirst[v] <= v || p->cnum[v] <= v || p->cfirst[v] >= p->M || p->cnum[v] >= p->M) {
        g_err = "ixg_cm_create: bad bifurcation children";
        return IXG_EINVAL;
      }
    } else if (st != E_ST) {
      const bool self = st == IL_ST || st == IR_ST;
      if (p->cnum[v] < 1 || p->cnum[v] > MAXC || p->cfirst[v] < (self ? v : v + 1) ||
          p->cfirst[v] + p->cnum[v] > p->M) {
        g_err = "ixg_cm_create: bad children";
        return IXG_EINVAL;
      }
    }
  }
  auto cm = std::make_unique<ixg_cm>();
  RCHK(configure_cm(*p, local, cm->h));
  marginal_emissions(*p, cm->h);
  RCHK(state_spans(*p, cm->h));
  dependency_levels(*p, cm->h);
  RCHK(qdb_tables(*p, cm->h));
  cm->h.clen = p->clen;
  flogsum_table(cm->ftbl_host);
  if (!c) {
    *out = cm.release();
    return IXG_OK;
  }
  HIPCHK(hipSetDevice(c->device));
  const int M = p->M;
  std::vector<int32_t> i32(cm->h.cfirst);
  i32.insert(i32.end(), cm->h.cnum.begin(), cm->h.cnum.end());
  i32.insert(i32.end(), cm->h.begin_list.begin(), cm->h.begin_list.end());
  i32.insert(i32.end(), cm->h.lvl_states.begin(), cm->h.lvl_states.end());
  i32.insert(i32.end(), cm->h.lvl_off.begin(), cm->h.lvl_off.end());
  std::vector<float> f(cm->h.tsc);
  f.insert(f.end(), cm->h.oesc.begin(), cm->h.oesc.end());
  f.insert(f.end(), cm->h.beginsc.begin(), cm->h.beginsc.end());
  f.insert(f.end(), cm->h.endsc.begin(), cm->h.endsc.end());
  f.insert(f.end(), cm->ftbl_host.begin(), cm->ftbl_host.end());
  f.insert(f.end(), cm->h.lmesc.begin(), cm->h.lmesc.end());
  f.insert(f.end(), cm->h.rmesc.begin(), cm->h.rmesc.end());
  RCHK(dmalloc(c, (void **)&cm->d_st, M));
  RCHK(dmalloc(c, (void **)&cm->d_i32, i32.size() * 4));
  RCHK(dmalloc(c, (void **)&cm->d_f, f.size() * 4));
  cm->b_i32 = i32.size() * 4;
  cm->b_f = f.size() * 4;
  HIPCHK(hipMemcpy(cm->d_st, cm->h.sttype.data(), M, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(cm->d_i32, i32.data(), i32.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(cm->d_f, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  RCHK(qdb_upload(c, cm.get()));
  *out = cm.release();
  return IXG_OK;
}

int ixg_cm_free(ixg_ctx *c, ixg_cm *cm) {
  if (!cm) return IXG_EINVAL;
  if (cm->d_st) {
    if (!c) return IXG_EINVAL;
    dfree(c, cm->d_st, cm->h.M);
    dfree(c, cm->d_i32, cm->b_i32);
    dfree(c, cm->d_f, cm->b_f);
    dfree(c, cm->d_tr, cm->b_tr);
    dfree(c, cm->d_yl, cm->b_yl);
    dfree(c, cm->d_q, cm->b_q);
    dfree(c, cm->d_tq, cm->b_tq);
  }
  delete cm;
  return IXG_OK;
}

int ixg_cm_tables(const ixg_cm *cm, float *tsc, float *oesc, float *beginsc, float *endsc, float *el) {
  if (!cm) return IXG_EINVAL;
  const CmHost &h = cm->h;
  if (tsc) memcpy(tsc, h.tsc.data(), h.tsc.size() * 4);
  if (oesc) memcpy(oesc, h.oesc.data(), h.oesc.size() * 4);
  if (beginsc) memcpy(beginsc, h.beginsc.data(), h.beginsc.size() * 4);
  if (endsc) memcpy(endsc, h.endsc.data(), h.endsc.size() * 4);
  if (el) *el = h.el_selfsc;
  return IXG_OK;
}

static int single_env(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const uint8_t *dsq, int64_t L, int32_t i0,
                      int32_t j0, int inside, std::vector<float> &root) {
  if (!c || !cm || !b || !dsq || !cm->d_st || i0 < 1 || j0 > L || j0 < i0) {
    g_err = "scan: bad arguments (1 <= i0 <= j0 <= L, device CM)";
    return IXG_EINVAL;
  }
  for (int64_t i = i0; i <= j0; i++)
    if (dsq[i] >= KP) {
      g_err = "digitized residue code >= 18 in the envelope";
      return IXG_EINVAL;
    }
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)L + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)L + 1, hipMemcpyHostToDevice, c->stream));
  ixg_envelope e;
  e.i0 = i0;
  e.j0 = j0;
  e.strand = 0;
  e.seq = 0;
  e.bands = *b;
  std::vector<std::vector<float>> roots;
  RCHK(run_envelopes(c, cm, d, {0}, &e, 1, inside, nullptr, nullptr, roots, nullptr));
  root.swap(roots[0]);
  return IXG_OK;
}

int ixg_cyk_scan_hb(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const uint8_t *dsq, int64_t L, int32_t i0,
                    int32_t j0, float env_cutoff, float *vsc, int64_t *envi, int64_t *envj) {
  if (!vsc || !envi || !envj) return IXG_EINVAL;
  std::vector<float> root;
  RCHK(single_env(c, cm, b, dsq, L, i0, j0, 0, root));
  root_summary(*b, root.data(), i0, j0, env_cutoff, *vsc, *envi, *envj);
  return IXG_OK;
}

int ixg_inside_scan_hb(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const uint8_t *dsq, int64_t L, int32_t i0,
                       int32_t j0, int inside, float env_cutoff, float cutoff, int do_null3, float *vsc,
                       int64_t *envi, int64_t *envj, ixg_hit *hits, int64_t cap, int64_t *nhits) {
  if (!vsc || !envi || !envj || !nhits || (cap > 0 && !hits)) return IXG_EINVAL;
  std::vector<float> root;
  RCHK(single_env(c, cm, b, dsq, L, i0, j0, inside ? 1 : 0, root));
  root_summary(*b, root.data(), i0, j0, env_cutoff, *vsc, *envi, *envj);
  std::vector<ixg_hit> out;
  report_hits(*cm, *b, dsq, i0, j0, root.data(), cutoff, do_null3, 0, out);
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

}  // extern "C"

// Host view of a device target (defined with the target in api.cpp).
const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L);
const uint8_t *ixg_target_host_dsq(const ixg_target *t, int strand);

static int hb_batch(ixg_ctx *c, const ixg_cm *cm, const std::vector<const ixg_cm *> *percm, const ixg_target *tgt,
                    const ixg_envelope *env, int64_t n, int inside, float env_cutoff, int report, float cutoff,
                    int do_null3, float *vsc, int64_t *envi, int64_t *envj, ixg_hit *hits, int64_t cap,
                    int64_t *nhits, int64_t *cells) {
  HIPCHK(hipSetDevice(c->device));
  int64_t L0 = 0;
  const uint8_t *base = ixg_target_device_base(tgt, &L0);
  std::vector<int64_t> off(n);
  std::vector<const uint8_t *> host(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t L = 0;
    if ((env[k].strand != 0 && env[k].strand != 1) ||
        !ixg_target_view(tgt, env[k].seq, env[k].strand, &off[k], &L, &host[k]) || env[k].i0 < 1 ||
        env[k].j0 > L || env[k].j0 < env[k].i0) {
      g_err = "ixg_hb_scan_batch: envelope outside the target";
      return IXG_EINVAL;
    }
  }
  c->ms[2] = 0;
  std::vector<std::vector<float>> roots;
  RCHK(run_envelopes(c, cm, base, off, env, n, inside ? 1 : 0, nullptr, nullptr, roots, cells, percm));
  std::vector<std::vector<ixg_hit>> per(n);
  par_for(n, [&](int64_t k) {
    root_summary(env[k].bands, roots[k].data(), env[k].i0, env[k].j0, env_cutoff, vsc[k], envi[k], envj[k]);
    if (report)
      report_hits(percm ? *(*percm)[k] : *cm, env[k].bands, host[k], env[k].i0, env[k].j0, roots[k].data(), cutoff,
                  do_null3, (int32_t)k, per[k]);
  });
  std::vector<ixg_hit> out;
  for (auto &h : per) out.insert(out.end(), h.begin(), h.end());
  if (report) {
    *nhits = (int64_t)out.size();
    for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
    if ((int64_t)out.size() > cap) return IXG_ERANGE;
  }
  return IXG_OK;
}

extern "C" int ixg_hb_scan_batch(ixg_ctx *c, const ixg_cm *cm, const ixg_target *tgt, const ixg_envelope *env,
                                 int64_t n, int inside, float env_cutoff, int report, float cutoff, int do_null3,
                                 float *vsc, int64_t *envi, int64_t *envj, ixg_hit *hits, int64_t cap,
                                 int64_t *nhits, int64_t *cells) {
  if (!c || !cm || !cm->d_st || !tgt || (n && !env) || !vsc || !envi || !envj || (report && !nhits)) {
    g_err = "ixg_hb_scan_batch: bad arguments";
    return IXG_EINVAL;
  }
  return hb_batch(c, cm, nullptr, tgt, env, n, inside, env_cutoff, report, cutoff, do_null3, vsc, envi, envj, hits,
                  cap, nhits, cells);
}

extern "C" int ixg_hb_scan_multi(ixg_ctx *c, const ixg_cm *const *cms, int32_t ncm, const int32_t *env_cm,
                                 const ixg_target *tgt, const ixg_envelope *env, int64_t n, int inside,
                                 float env_cutoff, int report, float cutoff, int do_null3, float *vsc, int64_t *envi,
                                 int64_t *envj, ixg_hit *hits, int64_t cap, int64_t *nhits, int64_t *cells) {
  if (!c || !cms || ncm < 1 || (n && (!env || !env_cm)) || !tgt || !vsc || !envi || !envj || (report && !nhits) ||
      n > INT32_MAX) {
    g_err = "ixg_hb_scan_multi: bad arguments";
    return IXG_EINVAL;
  }
  for (int32_t m = 0; m < ncm; m++)
    if (!cms[m] || !cms[m]->d_st) {
      g_err = "ixg_hb_scan_multi: null model";
      return IXG_EINVAL;
    }
  std::vector<const ixg_cm *> percm(n);
  for (int64_t k = 0; k < n; k++) {
    if (env_cm[k] < 0 || env_cm[k] >= ncm) {
      g_err = "ixg_hb_scan_multi: model index out of range";
      return IXG_EINVAL;
    }
    percm[k] = cms[env_cm[k]];
  }
  return hb_batch(c, cms[0], &percm, tgt, env, n, inside, env_cutoff, report, cutoff, do_null3, vsc, envi, envj,
                  hits, cap, nhits, cells);
}

extern "C" {

int ixg_cm_set_trpenalties(ixg_ctx *c, ixg_cm *cm, const float *g, const float *l) {
  if (!cm || !g || !l || (cm->d_st && !c)) {
    g_err = "ixg_cm_set_trpenalties: bad arguments";
    return IXG_EINVAL;
  }
  const int M = cm->h.M;
  cm->h.trp_g.assign(g, g + 3 * (size_t)M);
  cm->h.trp_l.assign(l, l + 3 * (size_t)M);
  std::vector<float> all(cm->h.trp_g);
  all.insert(all.end(), cm->h.trp_l.begin(), cm->h.trp_l.end());
  std::vector<int32_t> yl;
  for (int q = 0; q < 6; q++) {
    cm->yoff[q] = (int32_t)yl.size();
    for (int y = 1; y < M; y++)
      if (all[(size_t)q * M + y] > IMPOSSIBLE + 1.0f) yl.push_back(y);
  }
  cm->yoff[6] = (int32_t)yl.size();
  if (!cm->d_st) return IXG_OK;
  HIPCHK(hipSetDevice(c->device));
  dfree(c, cm->d_tr, cm->b_tr);
  dfree(c, cm->d_yl, cm->b_yl);
  cm->d_tr = nullptr;
  cm->d_yl = nullptr;
  cm->b_tr = all.size() * 4;
  cm->b_yl = std::max<size_t>(yl.size(), 1) * 4;
  RCHK(dmalloc(c, (void **)&cm->d_tr, cm->b_tr));
  RCHK(dmalloc(c, (void **)&cm->d_yl, cm->b_yl));
  HIPCHK(hipMemcpy(cm->d_tr, all.data(), cm->b_tr, hipMemcpyHostToDevice));
  if (!yl.empty()) HIPCHK(hipMemcpy(cm->d_yl, yl.data(), yl.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  return IXG_OK;
}

int ixg_cm_trtables(const ixg_cm *cm, float *lmesc, float *rmesc, int32_t *lpos, int32_t *rpos) {
  if (!cm) return IXG_EINVAL;
  const CmHost &h = cm->h;
  if (lmesc) memcpy(lmesc, h.lmesc.data(), h.lmesc.size() * 4);
  if (rmesc) memcpy(rmesc, h.rmesc.data(), h.rmesc.size() * 4);
  if (lpos) memcpy(lpos, h.lpos.data(), h.lpos.size() * 4);
  if (rpos) memcpy(rpos, h.rpos.data(), h.rpos.size() * 4);
  return IXG_OK;
}

}  // extern "C"

static int tr_single(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const ixg_trvalid *valid, const uint8_t *dsq,
                     int64_t L, int32_t i0, int32_t j0, float cutoff, int do_null3, int pass_idx, float env_cutoff,
                     int local, int inside, float *vsc, int32_t *vmode, int64_t *envi, int64_t *envj,
                     ixg_trhit *hits, int64_t cap, int64_t *nhits) {
  if (!c || !cm || !b || !valid || !dsq || !cm->d_st || i0 < 1 || j0 > L || j0 < i0 || !vsc || !vmode || !envi ||
      !envj || !nhits || (cap > 0 && !hits)) {
    g_err = "truncated scan: bad arguments (1 <= i0 <= j0 <= L, device CM)";
    return IXG_EINVAL;
  }
  for (int64_t i = i0; i <= j0; i++)
    if (dsq[i] >= KP) {
      g_err = "digitized residue code >= 18 in the envelope";
      return IXG_EINVAL;
    }
  TrSel sel;
  RCHK(tr_select(cm, pass_idx, local, sel));
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)L + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)L + 1, hipMemcpyHostToDevice, c->stream));
  ixg_envelope e;
  e.i0 = i0;
  e.j0 = j0;
  e.strand = 0;
  e.seq = 0;
  e.bands = *b;
  std::vector<std::vector<float>> roots;
  RCHK(run_envelopes(c, cm, d, {0}, &e, 1, inside, &sel, valid, roots, nullptr));
  const int64_t nroot = (int64_t)roots[0].size() / 5;
  tr_root_summary(*b, roots[0].data(), nroot, mode_mask(*valid, 0, sel.fill), i0, j0, env_cutoff, *vsc, *vmode,
                  *envi, *envj);
  std::vector<ixg_trhit> out;
  tr_report(*cm, *b, dsq, i0, j0, roots[0].data(), nroot, pass_idx, local, cutoff, do_null3, 0, out);
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

extern "C" {

int ixg_trcyk_scan_hb(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const ixg_trvalid *valid,
                      const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0, float cutoff, int do_null3,
                      int pass_idx, float env_cutoff, int local, float *vsc, int32_t *vmode, int64_t *envi,
                      int64_t *envj, ixg_trhit *hits, int64_t cap, int64_t *nhits) {
  return tr_single(c, cm, b, valid, dsq, L, i0, j0, cutoff, do_null3, pass_idx, env_cutoff, local, 0, vsc, vmode,
                   envi, envj, hits, cap, nhits);
}

int ixg_trinside_scan_hb(ixg_ctx *c, const ixg_cm *cm, const ixg_bands *b, const ixg_trvalid *valid,
                         const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0, float cutoff, int do_null3,
                         int pass_idx, float env_cutoff, int local, float *vsc, int32_t *vmode, int64_t *envi,
                         int64_t *envj, ixg_trhit *hits, int64_t cap, int64_t *nhits) {
  return tr_single(c, cm, b, valid, dsq, L, i0, j0, cutoff, do_null3, pass_idx, env_cutoff, local, 1, vsc, vmode,
                   envi, envj, hits, cap, nhits);
}

int ixg_trhb_scan_batch(ixg_ctx *c, const ixg_cm *cm, const ixg_target *tgt, const ixg_trenvelope *env, int64_t n,
                        int inside, int pass_idx, int local, float env_cutoff, float cutoff, int do_null3,
                        float *vsc, int32_t *vmode, int64_t *envi, int64_t *envj, ixg_trhit *hits, int64_t cap,
                        int64_t *nhits, int64_t *cells) {
  if (!c || !cm || !cm->d_st || !tgt || (n && !env) || !vsc || !vmode || !envi || !envj || !nhits ||
      (cap > 0 && !hits)) {
    g_err = "ixg_trhb_scan_batch: bad arguments";
    return IXG_EINVAL;
  }
  TrSel sel;
  RCHK(tr_select(cm, pass_idx, local, sel));
  HIPCHK(hipSetDevice(c->device));
  int64_t L0 = 0;
  const uint8_t *base = ixg_target_device_base(tgt, &L0);
  std::vector<int64_t> off(n);
  std::vector<const uint8_t *> host(n);
  std::vector<ixg_envelope> ev(n);
  std::vector<ixg_trvalid> val(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t L = 0;
    if ((env[k].strand != 0 && env[k].strand != 1) ||
        !ixg_target_view(tgt, env[k].seq, env[k].strand, &off[k], &L, &host[k]) || env[k].i0 < 1 ||
        env[k].j0 > L || env[k].j0 < env[k].i0) {
      g_err = "ixg_trhb_scan_batch: envelope outside the target";
      return IXG_EINVAL;
    }
    ev[k].i0 = env[k].i0;
    ev[k].j0 = env[k].j0;
    ev[k].strand = env[k].strand;
    ev[k].seq = env[k].seq;
    ev[k].bands = env[k].bands;
    val[k] = env[k].valid;
  }
  c->ms[2] = 0;
  std::vector<std::vector<float>> roots;
  RCHK(run_envelopes(c, cm, base, off, ev.data(), n, inside ? 1 : 0, &sel, val.data(), roots, cells));
  std::vector<std::vector<ixg_trhit>> per(n);
  par_for(n, [&](int64_t k) {
    const int64_t nroot = (int64_t)roots[k].size() / 5;
    tr_root_summary(ev[k].bands, roots[k].data(), nroot, mode_mask(val[k], 0, sel.fill), ev[k].i0, ev[k].j0,
                    env_cutoff, vsc[k], vmode[k], envi[k], envj[k]);
    tr_report(*cm, ev[k].bands, host[k], ev[k].i0, ev[k].j0, roots[k].data(), nroot,
              pass_idx, local, cutoff, do_null3, (int32_t)k, per[k]);
  });
  std::vector<ixg_trhit> out;
  for (auto &h : per) out.insert(out.end(), h.begin(), h.end());
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

}  // extern "C"

namespace ixg {
bool tr_allow(const CmHost &h, int pass_idx, int local, int start, int stop, int i0, int j0, int mode, int b) {
  return allow_truncation(h, pass_idx, local, start, stop, i0, j0, mode, b);
}
}  // namespace ixg

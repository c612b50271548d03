// The device-resident configured CM handle shared by the scanner APIs.
#pragma once
#include "cm_internal.h"
#include "ctx.h"

struct ixg_cm {
  using CmHost = ixg::CmHost;
  using CmDev = ixg::CmDev;
  CmHost h;
  std::vector<float> ftbl_host;
  int8_t *d_st = nullptr;
  int32_t *d_i32 = nullptr;  // cfirst | cnum | begin_list | lvl_states | lvl_off
  float *d_f = nullptr;      // tsc | oesc | beginsc | endsc | flogsum table | lmesc | rmesc
  float *d_tr = nullptr;     // truncated-begin penalties: g[3][M] | l[3][M]
  int32_t *d_q = nullptr;    // QDB tables: itsc | ioesc | ibeginsc | iendsc | ilogsum table |
                             // begl index | base0 (int) | el_sca (int) | base0 (f32 bits) | el_sca (f32 bits) | begin list
  size_t b_q = 0;
  int32_t *d_tq = nullptr;   // truncated QDB tables: ilmesc | irmesc (M*18 each) | ig[3][M] | il[3][M]
  size_t b_tq = 0;
  int32_t *d_yl = nullptr;   // per (local, pass index): states with an allowed truncated begin
  int32_t yoff[7] = {0, 0, 0, 0, 0, 0, 0};
  size_t b_i32 = 0, b_f = 0, b_tr = 0, b_yl = 0;
  CmDev dev() const {
    CmDev c;
    const int M = h.M;
    c.sttype = d_st;
    c.cfirst = d_i32;
    c.cnum = d_i32 + M;
    c.begin_list = d_i32 + 2 * M;
    c.nbegin = (int)h.begin_list.size();
    c.lvl_states = c.begin_list + h.begin_list.size();
    c.lvl_off = c.lvl_states + M;
    c.nlev = (int)h.lvl_off.size() - 1;
    c.tsc = d_f;
    c.oesc = d_f + (size_t)M * ixg::MAXC;
    c.beginsc = c.oesc + (size_t)M * ixg::KP * ixg::KP;
    c.endsc = c.beginsc + M;
    c.lmesc = d_ftbl() + ixg::FLOGSUM_TBL;
    c.rmesc = c.lmesc + (size_t)M * ixg::KP;
    c.M = M;
    c.local_begin = h.local_begin;
    c.el_selfsc = h.el_selfsc;
    return c;
  }
  const float *d_ftbl() const { return d_f + (size_t)h.M * ixg::MAXC + (size_t)h.M * ixg::KP * ixg::KP + 2 * (size_t)h.M; }
};

// Uploads the QDB scanner's integer scores and tables (cm->d_q) and the
// truncated QDB scanner's integer marginal emissions (cm->d_tq).
int qdb_upload(ixg_ctx *c, ixg_cm *cm);

namespace ixg {
// cm_dpsearch_trunc.rs:2376-2422 cm_hit_AllowTruncation for a hit start..stop
// of marginal mode `mode` entered at state b.
bool tr_allow(const CmHost &h, int pass_idx, int local, int start, int stop, int i0, int j0, int mode, int b);
}  // namespace ixg


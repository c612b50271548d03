// Host-side covariance-model score configuration (once per model): local
// begins/ends, log-odds transitions and emissions, EL self score. Mirrors
// cp9.rs:4194-4545 (cm_config_local, cm_logoddsify, cm_configure_scores).
#include <algorithm>
#include <cmath>
#include <cstring>
#include "cm_internal.h"
#include "ctx.h"

namespace ixg {

static double sre_log2(double x) { return x > 0.0 ? std::log(x) * 1.44269504 : -1e36; }  // cp9.rs:961

static const unsigned char DEGEN[KP][4] = {  // cm.rs:24-43
    {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {0, 0, 0, 0}, {1, 0, 1, 0}, {0, 1, 0, 1},
    {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 1, 1, 0}, {1, 0, 0, 1}, {1, 1, 0, 1}, {0, 1, 1, 1}, {1, 1, 1, 0},
    {1, 0, 1, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}};
static int ndeg(int x) { return DEGEN[x][0] + DEGEN[x][1] + DEGEN[x][2] + DEGEN[x][3]; }
static float frac(int x, int l) { return DEGEN[x][l] ? 1.0f / (float)ndeg(x) : 0.0f; }
static float favg(int x, const float *sc) {
  if (!(x < 4 || (x > 4 && x < KP - 2))) return 0.0f;
  float r = 0.0f;
  for (int i = 0; i < 4; i++)
    if (DEGEN[x][i]) r += sc[i];
  return r / (float)ndeg(x);
}
static float kahan_sum(const float *v, int n) {  // cp9.rs:451 esl_vec_fsum
  float sum = 0.0f, c = 0.0f;
  for (int i = 0; i < n; i++) {
    float y = v[i] - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

// cm_nohmm.rs:27-33 prob2score: floor(0.5 + 1000 log2(p / null)), -INFTY for p == 0
static int32_t prob2score(float p, float null) {
  if (p == 0.0f) return -INFTY_I;
  return (int32_t)std::floor(0.5 + 1000.0 * std::log2((double)p / (double)null));
}
// cm.rs:104-120 abc_iavg_score: mean over a degenerate code, rounded half away from zero
static int32_t iavg(int x, const int32_t *sc) {
  if (!(x < 4 || (x > 4 && x < KP - 2))) return 0;
  float r = 0.0f;
  for (int i = 0; i < 4; i++)
    if (DEGEN[x][i]) r += (float)sc[i];
  r /= (float)ndeg(x);
  return r < 0.0f ? (int32_t)(r - 0.5f) : (int32_t)(r + 0.5f);
}

// Integer scores of the QDB Inside scanner (cm_nohmm.rs:279-368
// build_integer_scores) from the configured probabilities.
static void integer_scores(const ixg_cm_params &p, const float *t, const float *begin, const float *end, CmHost &h) {
  const int M = p.M;
  h.itsc.assign((size_t)M * MAXC, 0);
  h.ioesc.assign((size_t)M * KP * KP, 0);
  h.ibeginsc.assign(M, 0);
  h.iendsc.assign(M, 0);
  for (int v = 0; v < M; v++) {
    const int st = p.sttype[v];
    if (st != B_ST && st != E_ST)
      for (int x = 0; x < p.cnum[v]; x++) h.itsc[(size_t)v * MAXC + x] = prob2score(t[(size_t)v * MAXC + x], 1.0f);
    int32_t *oe = &h.ioesc[(size_t)v * KP * KP];
    const float *e = p.e + (size_t)v * 16;
    if (st == MP_ST) {
      int32_t ie[16];
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) ie[a * 4 + b] = prob2score(e[a * 4 + b], p.null[a] * p.null[b]);
      for (int i = 0; i < KP * KP; i++) oe[i] = -INFTY_I;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) oe[a * KP + b] = ie[a * 4 + b];
      for (int a = 5; a < KP - 1; a++)
        for (int b = 0; b < 4; b++) {
          float sc = 0.f;
          for (int l = 0; l < 4; l++) sc += (float)ie[l * 4 + b] * frac(a, l);
          oe[a * KP + b] = (int32_t)sc;
        }
      for (int a = 0; a < 4; a++)
        for (int b = 5; b < KP - 1; b++) {
          float sc = 0.f;
          for (int r = 0; r < 4; r++) sc += (float)ie[a * 4 + r] * frac(b, r);
          oe[a * KP + b] = (int32_t)sc;
        }
      for (int a = 5; a < KP - 1; a++)
        for (int b = 5; b < KP - 1; b++) {
          float sc = 0.f;
          for (int l = 0; l < 4; l++)
            for (int r = 0; r < 4; r++) sc += (float)ie[l * 4 + r] * frac(a, l) * frac(b, r);
          oe[a * KP + b] = (int32_t)sc;
        }
    } else if (st == ML_ST || st == MR_ST || st == IL_ST || st == IR_ST) {
      int32_t ie[4];
      for (int a = 0; a < 4; a++) ie[a] = prob2score(e[a], p.null[a]);
      for (int a = 0; a < KP; a++) oe[a] = -INFTY_I;
      for (int a = 0; a < 4; a++) oe[a] = ie[a];
      for (int a = 5; a < KP - 1; a++) oe[a] = iavg(a, ie);
    }
    h.ibeginsc[v] = prob2score(begin[v], 1.0f);
    h.iendsc[v] = prob2score(end[v], 1.0f);
  }
  h.iel_selfsc = prob2score(std::pow(2.0f, h.el_selfsc), 1.0f);
}

int configure_cm(const ixg_cm_params &p, int local, CmHost &h) {
  const int M = p.M;
  h.M = M;
  h.sttype.assign(p.sttype, p.sttype + M);
  h.cfirst.assign(p.cfirst, p.cfirst + M);
  h.cnum.assign(p.cnum, p.cnum + M);
  memcpy(h.null, p.null, sizeof(h.null));
  h.n3_omega = p.n3_omega;
  std::vector<float> t(p.t, p.t + (size_t)M * MAXC), begin(M, 0.f), end(M, 0.f);
  auto is_start = [&](int nd) {
    int ty = p.ndtype[nd];
    return ty == MATP_ND || ty == MATL_ND || ty == MATR_ND || ty == BIF_ND;
  };
  if (local) {  // cm_config_local: local begins and ends
    int nstarts = 0;
    for (int nd = 2; nd < p.nodes; nd++) nstarts += is_start(nd);
    begin[p.nodemap[1]] = 1.0f - p.pbegin;
    const float pb = p.pbegin / (float)nstarts;
    for (int nd = 2; nd < p.nodes; nd++)
      if (is_start(nd)) begin[p.nodemap[nd]] = pb;
    for (int y = 0; y < p.cnum[0]; y++) t[y] = 0.f;
    auto is_exit = [&](int nd) {
      int ty = p.ndtype[nd];
      return (ty == MATP_ND || ty == MATL_ND || ty == MATR_ND || ty == BEGL_ND || ty == BEGR_ND) &&
             p.ndtype[nd + 1] != END_ND;
    };
    int nexits = 0;
    for (int nd = 1; nd < p.nodes; nd++) nexits += is_exit(nd);
    for (int nd = 1; nd < p.nodes; nd++)
      if (is_exit(nd)) {
        const int v = p.nodemap[nd];
        end[v] = p.pend / (float)nexits;
        float denom = kahan_sum(&t[(size_t)v * MAXC], p.cnum[v]);
        denom += end[v];
        const float inv = 1.0f / denom;
        for (int x = 0; x < p.cnum[v]; x++) t[(size_t)v * MAXC + x] *= inv;
      }
  }
  h.local_begin = local ? 1 : 0;
  h.el_selfsc = p.el_selfsc;
  if (h.el_selfsc * (float)p.W < IMPOSSIBLE) h.el_selfsc = IMPOSSIBLE / ((float)p.W + 1.0f);
  h.tsc.assign((size_t)M * MAXC, 0.f);
  h.oesc.assign((size_t)M * KP * KP, 0.f);
  h.beginsc.assign(M, 0.f);
  h.endsc.assign(M, 0.f);
  for (int v = 0; v < M; v++) {
    const int st = p.sttype[v];
    if (st != B_ST && st != E_ST)
      for (int x = 0; x < p.cnum[v]; x++) h.tsc[(size_t)v * MAXC + x] = (float)sre_log2((double)t[(size_t)v * MAXC + x]);
    float *oe = &h.oesc[(size_t)v * KP * KP];
    const float *e = p.e + (size_t)v * 16;
    if (st == MP_ST) {
      float esc[16];
      for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) esc[x * 4 + y] = (float)sre_log2((double)(e[x * 4 + y] / (p.null[x] * p.null[y])));
      for (int i = 0; i < KP * KP; i++) oe[i] = IMPOSSIBLE;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) oe[a * KP + b] = esc[a * 4 + b];
      for (int a = 5; a < KP - 1; a++)
        for (int b = 0; b < 4; b++) {
          float sc = 0.f;
          for (int l = 0; l < 4; l++) sc += esc[l * 4 + b] * frac(a, l);
          oe[a * KP + b] = sc;
        }
      for (int a = 0; a < 4; a++)
        for (int b = 5; b < KP - 1; b++) {
          float sc = 0.f;
          for (int r = 0; r < 4; r++) sc += esc[a * 4 + r] * frac(b, r);
          oe[a * KP + b] = sc;
        }
      for (int a = 5; a < KP - 1; a++)
        for (int b = 5; b < KP - 1; b++) {
          float sc = 0.f;
          for (int l = 0; l < 4; l++)
            for (int r = 0; r < 4; r++) sc += esc[l * 4 + r] * frac(a, l) * frac(b, r);
          oe[a * KP + b] = sc;
        }
    } else if (st == ML_ST || st == MR_ST || st == IL_ST || st == IR_ST) {
      float esc[4];
      for (int x = 0; x < 4; x++) esc[x] = (float)sre_log2((double)(e[x] / p.null[x]));
      for (int a = 0; a < KP; a++) oe[a] = IMPOSSIBLE;
      for (int a = 0; a < 4; a++) oe[a] = esc[a];
      for (int a = 5; a < KP - 1; a++) oe[a] = favg(a, esc);
    }
    h.beginsc[v] = (float)sre_log2((double)begin[v]);
    h.endsc[v] = (float)sre_log2((double)end[v]);
  }
  integer_scores(p, t.data(), begin.data(), end.data(), h);
  h.begin_list.clear();
  for (int y = 1; y < M; y++)
    if (h.beginsc[y] > -9.999e35f) h.begin_list.push_back(y);
  return IXG_OK;
}

void flogsum_table(std::vector<float> &t) {  // cp9.rs:5391-5401
  t.resize(FLOGSUM_TBL);
  for (int i = 0; i < FLOGSUM_TBL; i++) t[i] = (float)sre_log2(1.0 + std::exp((-(double)i / 1000.0) * 0.69314718));
}

float host_flogsum(const std::vector<float> &t, float a, float b) {  // cp9.rs:5405
  const float mx = a > b ? a : b, mn = a < b ? a : b;
  if (mn == -INFINITY || (mx - mn) >= 23.0f) return mx;
  return mx + t[(size_t)((mx - mn) * 1000.0f)];
}

}  // namespace ixg

namespace ixg {

// Null-weighted mean score of a degenerate code (cp9.rs:4375-4388).
static float expect_score(int x, const float *sc, const float *p) {
  if (!(x < 4 || (x > 4 && x < KP - 2))) return 0.0f;
  float r = 0.0f, den = 0.0f;
  for (int i = 0; i < 4; i++)
    if (DEGEN[x][i]) {
      r += sc[i] * p[i];
      den += p[i];
    }
  return r / den;
}

// cp9.rs:4391-4408 i_expect_score (esl_abc_IExpectScore): the null-weighted
// mean of the canonical integer scores over a degenerate code, rounded half
// away from zero; 0 for non-residue codes.
static int32_t i_expect_score(int x, const int32_t *sc, const float *null) {
  if (!(x < 4 || (x > 4 && x < KP - 2))) return 0;
  float r = 0.0f, den = 0.0f;
  for (int i = 0; i < 4; i++)
    if (DEGEN[x][i]) {
      r += (float)sc[i] * null[i];
      den += null[i];
    }
  r /= den;
  return r < 0.0f ? (int32_t)(r - 0.5f) : (int32_t)(r + 0.5f);
}

// cp9.rs:4441-4527 fill_marginal_emissions, integer part (ilmesc/irmesc).
static void imarginal_emissions(const ixg_cm_params &p, CmHost &h) {
  const int M = p.M;
  h.ilmesc.assign((size_t)M * KP, 0);
  h.irmesc.assign((size_t)M * KP, 0);
  for (int v = 0; v < M; v++) {
    int32_t *lm = &h.ilmesc[(size_t)v * KP], *rm = &h.irmesc[(size_t)v * KP];
    lm[4] = rm[4] = lm[KP - 2] = rm[KP - 2] = lm[KP - 1] = rm[KP - 1] = -INFTY_I;
    const int st = p.sttype[v];
    const float *e = p.e + (size_t)v * 16;
    int32_t *sides[2] = {nullptr, nullptr};
    if (st == MP_ST) {
      float l4[4] = {0, 0, 0, 0}, r4[4] = {0, 0, 0, 0};
      for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) {
          l4[x] += e[x * 4 + y];
          r4[y] += e[x * 4 + y];
        }
      for (int x = 0; x < 4; x++) {
        lm[x] = prob2score(l4[x], p.null[x]);
        rm[x] = prob2score(r4[x], p.null[x]);
      }
      sides[0] = lm;
      sides[1] = rm;
    } else if (st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST) {
      int32_t *own = (st == ML_ST || st == IL_ST) ? lm : rm;
      for (int x = 0; x < 4; x++) own[x] = prob2score(e[x], p.null[x]);
      sides[0] = own;
    }
    for (int32_t *s : sides)
      if (s)
        for (int x = 5; x <= KP - 3; x++) s[x] = i_expect_score(x, s, p.null);
  }
}

// Marginal emission scores of the truncated scanners (cp9.rs:4441-4527): a
// pair state's left (right) score is the log-odds of its row (column) sum; a
// singlet state's own side is its emission score, the other side 0.
void marginal_emissions(const ixg_cm_params &p, CmHost &h) {
  imarginal_emissions(p, h);
  const int M = p.M;
  h.lmesc.assign((size_t)M * KP, 0.f);
  h.rmesc.assign((size_t)M * KP, 0.f);
  for (int v = 0; v < M; v++) {
    float *lm = &h.lmesc[(size_t)v * KP], *rm = &h.rmesc[(size_t)v * KP];
    lm[4] = rm[4] = lm[KP - 2] = rm[KP - 2] = lm[KP - 1] = rm[KP - 1] = IMPOSSIBLE;
    const int st = p.sttype[v];
    const float *e = p.e + (size_t)v * 16;
    float *sides[2] = {nullptr, nullptr};
    if (st == MP_ST) {
      float l4[4] = {0, 0, 0, 0}, r4[4] = {0, 0, 0, 0};
      for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) {
          l4[x] += e[x * 4 + y];
          r4[y] += e[x * 4 + y];
        }
      for (int x = 0; x < 4; x++) {
        lm[x] = (float)sre_log2((double)(l4[x] / p.null[x]));
        rm[x] = (float)sre_log2((double)(r4[x] / p.null[x]));
      }
      sides[0] = lm;
      sides[1] = rm;
    } else if (st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST) {
      float *own = (st == ML_ST || st == IL_ST) ? lm : rm;
      for (int x = 0; x < 4; x++) own[x] = (float)sre_log2((double)(e[x] / p.null[x]));
      sides[0] = own;
    }
    for (float *s : sides)
      if (s)
        for (int x = 5; x <= KP - 3; x++) s[x] = expect_score(x, s, p.null);
  }
}

// Consensus span [lpos, rpos] of every state's subtree: a left-to-right walk
// of the guide tree counting consensus columns (the emit map, cp9.rs:144-204),
// with the per-node-type offsets of cm_dpsearch_trunc.rs:2394-2403.
int state_spans(const ixg_cm_params &p, CmHost &h) {
  const int n = p.nodes, M = p.M;
  std::vector<int32_t> ndidx(M), lpos(n, -1), rpos(n, -1);
  for (int nd = 0, v = 0; v < M; v++) {
    while (nd + 1 < n && p.nodemap[nd + 1] <= v) nd++;
    ndidx[v] = nd;
  }
  std::vector<int32_t> stack;  // (node << 1) | right side
  stack.push_back(0);
  int cpos = 0;
  int64_t steps = 0;
  while (!stack.empty()) {
    const int it = stack.back(), nd = it >> 1, nt = p.ndtype[nd];
    stack.pop_back();
    if (++steps > 2 * (int64_t)n + 2 || (!(it & 1) && lpos[nd] != -1)) {
      g_err = "ixg_cm_create: node types and bifurcation children do not form a guide tree";
      return IXG_EINVAL;
    }
    if (it & 1) {
      rpos[nd] = cpos + 1;
      if (nt == MATP_ND || nt == MATR_ND) cpos++;
      continue;
    }
    if (nt == MATP_ND || nt == MATL_ND) cpos++;
    lpos[nd] = cpos;
    stack.push_back((nd << 1) | 1);
    if (nt == BIF_ND) {
      const int v = p.nodemap[nd];
      stack.push_back(ndidx[p.cnum[v]] << 1);
      stack.push_back(ndidx[p.cfirst[v]] << 1);
    } else if (nt != END_ND && nd + 1 < n) {
      stack.push_back((nd + 1) << 1);
    }
  }
  h.lpos.resize(M);
  h.rpos.resize(M);
  for (int v = 0; v < M; v++) {
    const int nd = ndidx[v], nt = p.ndtype[nd];
    h.lpos[v] = (nt == MATP_ND || nt == MATL_ND) ? lpos[nd] : lpos[nd] + 1;
    h.rpos[v] = (nt == MATP_ND || nt == MATR_ND) ? rpos[nd] : rpos[nd] - 1;
  }
  return IXG_OK;
}

// cp9.rs:752-837 cm_expected_state_occupancy of the un-localized CM: each
// state's expected visits, summed over its parents in increasing order (an
// insert's own self transition taken out and applied as a geometric factor
// afterwards); a parent x reaches child v with t[x][v - cfirst[x]].
void expected_occupancy(const ixg_cm_params &p, std::vector<double> &psi) {
  const int M = p.M;
  psi.assign(M, 0.0);
  // parents: every non-B, non-E state lists its children as a run
  std::vector<std::vector<int32_t>> par(M);
  for (int x = 0; x < M; x++) {
    const int st = p.sttype[x];
    if (st == B_ST || st == E_ST) continue;
    for (int y = p.cfirst[x]; y < p.cfirst[x] + p.cnum[x]; y++)
      if (y != x) par[y].push_back(x);
  }
  for (int v = 0; v < M; v++) {
    const int st = p.sttype[v];
    if (st == S_ST) {
      psi[v] = 1.0;
      continue;
    }
    double a = 0.0;
    for (int32_t x : par[v]) a += psi[x] * (double)p.t[(size_t)x * 6 + (v - p.cfirst[x])];
    psi[v] = a;
    if (st == IL_ST || st == IR_ST) {
      const double ts = (double)p.t[(size_t)v * 6];
      psi[v] += psi[v] * (ts / (1.0 - ts));
    }
  }
}

// cm_trunc.rs:164-182: log2 as ln x 1.44269504 and Prob2Score over it (a
// float-to-int conversion saturates there)
static float trunc_log2(float x) { return x > 0.0f ? (float)(std::log((double)x) * 1.44269504) : -1e36f; }
static int32_t trunc_prob2score(float x) {
  if (x == 0.0f) return -INFTY_I;
  const double r = std::floor(0.5 + 1000.0 * (x > 0.0f ? std::log((double)x) * 1.44269504 : -1e36));
  return r < (double)INT32_MIN ? INT32_MIN : (r > (double)INT32_MAX ? INT32_MAX : (int32_t)r);
}

// cm_trunc.rs:192-384 TrPenalties::new on the global CM: per node, the
// truncated-begin probability of a MATP/MATL/MATR/BIF node (global: uniform
// over the 5'&3' fragments, or over the 5'/3' fragments when the node's span
// reaches the 3'/5' end; local: the node's local begin probability over its
// subtree's fragments plus the running value carried down MATL / MATR
// chains) is split between the node's first state and the inserts before it
// by their occupancies, then taken to log2 and to integer scores.
int tr_penalties(const ixg_cm_params &p, float *g, float *l, int32_t *ig, int32_t *il) {
  const int M = p.M, n = p.nodes, clen = p.clen;
  CmHost h;
  RCHK(state_spans(p, h));
  std::vector<double> psi;
  expected_occupancy(p, psi);
  std::vector<int32_t> ndidx(M);
  for (int nd = 0, v = 0; v < M; v++) {
    while (nd + 1 < n && p.nodemap[nd + 1] <= v) nd++;
    ndidx[v] = nd;
  }
  for (int q = 0; q < 3 * M; q++) g[q] = l[q] = IMPOSSIBLE, ig[q] = il[q] = -INFTY_I;
  const float g53 = (float)(2.0 / (double)(clen * (clen + 1))), g5or3 = (float)(1.0 / (double)clen);
  auto starts = [&](int nd) {
    const int t = p.ndtype[nd];
    return t == MATP_ND || t == MATL_ND || t == MATR_ND || t == BIF_ND;
  };
  std::vector<float> begin(M, 0.0f);  // local begin probabilities (cm_modelconfig.c:397)
  int nstarts = 0;
  for (int nd = 2; nd < n; nd++) nstarts += starts(nd);
  if (n > 1) begin[p.nodemap[1]] = (float)(1.0 - (double)p.pbegin);
  const float pb = p.pbegin / (float)nstarts;
  for (int nd = 2; nd < n; nd++)
    if (starts(nd)) begin[p.nodemap[nd]] = pb;
  float prv5 = 0.f, prv3 = 0.f, prv53 = 0.f;
  for (int nd = 0; nd < n; nd++) {
    const int t = p.ndtype[nd], m = p.nodemap[nd];
    if (t == END_ND) {
      prv5 = prv3 = prv53 = 0.f;
    } else if (t == BEGL_ND || t == BEGR_ND) {
      int b = -1;  // the bifurcation whose child this S state is
      for (int x = 0; x < m; x++)
        if (p.sttype[x] == B_ST && (p.cfirst[x] == m || p.cnum[x] == m)) b = x;
      if (b < 0) {
        g_err = "ixg_cm_trpenalties: a BEGL/BEGR node without its bifurcation";
        return IXG_EINVAL;
      }
      prv5 = t == BEGL_ND ? 0.f : l[1 * M + b];
      prv3 = t == BEGR_ND ? 0.f : l[2 * M + b];
      prv53 = l[0 * M + b];
    } else if (starts(nd)) {
      int i1 = -1, i2 = -1;  // inserts_given_node(nd - 1), cm_trunc.rs:111
      if (nd >= 1) {
        const int u = p.nodemap[nd - 1], tu = p.ndtype[nd - 1];
        if (tu == MATP_ND) i1 = u + 4, i2 = u + 5;
        else if (tu == MATL_ND || tu == MATR_ND) i1 = u + 2;
        else if (tu == BEGR_ND) i1 = u + 1;
        else if (tu == ROOT_ND) i1 = u + 1, i2 = u + 2;
      }
      const float mp = (float)psi[m], p1 = i1 >= 0 ? (float)psi[i1] : 0.f, p2 = i2 >= 0 ? (float)psi[i2] : 0.f;
      const float sum = mp + p1 + p2;
      auto put = [&](float *a, int q, float prob) {
        a[q * M + m] = (mp / sum) * prob;
        if (i1 >= 0) a[q * M + i1] = (p1 / sum) * prob;
        if (i2 >= 0) a[q * M + i2] = (p2 / sum) * prob;
      };
      const int lp = h.lpos[m], rp = h.rpos[m];
      put(g, 0, g53);
      if (rp == clen) put(g, 1, g5or3);
      if (lp == 1) put(g, 2, g5or3);
      const int sub = rp - lp + 1, nf53 = (sub * (sub + 1)) / 2;
      const float c5 = begin[m] / (float)sub + prv5, c3 = begin[m] / (float)sub + prv3;
      const float c53 = begin[m] / (float)nf53 + prv53;
      put(l, 0, c53);
      put(l, 1, c5);
      put(l, 2, c3);
      prv5 = t == MATL_ND ? c5 : 0.f;
      prv3 = t == MATR_ND ? c3 : 0.f;
      prv53 = c53;
    }
  }
  for (int v = 0; v < M; v++) {
    const int st = p.sttype[v], nt = p.ndtype[ndidx[v]];
    const bool detached = v + 1 < M && p.sttype[v + 1] == E_ST;
    const bool q = st == MP_ST || (st == ML_ST && nt == MATL_ND) || (st == MR_ST && nt == MATR_ND) || st == B_ST ||
                   ((st == IL_ST || st == IR_ST) && !detached);
    if (!q) continue;
    if (st == IL_ST && nt == MATP_ND && ndidx[v] + 1 < n && p.ndtype[ndidx[v] + 1] == END_ND) continue;
    for (int k = 0; k < 3; k++) {
      float &gv = g[k * M + v], &lv = l[k * M + v];
      if (gv > -0.5e36f) ig[k * M + v] = trunc_prob2score(gv), gv = trunc_log2(gv);
      il[k * M + v] = trunc_prob2score(lv);
      lv = trunc_log2(lv);
    }
  }
  return IXG_OK;
}

}  // namespace ixg

extern "C" int ixg_cm_trpenalties(const ixg_cm_params *p, float *g, float *l, int32_t *ig, int32_t *il) {
  if (!p || !g || !l || !ig || !il || p->M < 1 || p->nodes < 2 || p->clen < 1 || !p->sttype || !p->ndtype ||
      !p->nodemap || !p->cfirst || !p->cnum || !p->t) {
    ixg::g_err = "ixg_cm_trpenalties: bad arguments";
    return IXG_EINVAL;
  }
  return ixg::tr_penalties(*p, g, l, ig, il);
}

namespace ixg {

void dependency_levels(const ixg_cm_params &p, CmHost &h) {
  const int M = p.M;
  std::vector<int32_t> lvl(M, 0);
  int nlev = 0;
  for (int v = M - 1; v >= 0; v--) {
    const int st = p.sttype[v];
    int l = 0;
    if (st == B_ST) {
      l = 1 + std::max(lvl[p.cfirst[v]], lvl[p.cnum[v]]);
    } else if (st != E_ST) {
      for (int k = 0; k < p.cnum[v]; k++) {
        const int y = p.cfirst[v] + k;
        if (y != v) l = std::max(l, lvl[y] + 1);
      }
    }
    lvl[v] = l;
    nlev = std::max(nlev, l + 1);
  }
  h.lvl_off.assign(nlev + 1, 0);
  for (int v = 0; v < M; v++) h.lvl_off[lvl[v] + 1]++;
  for (int l = 0; l < nlev; l++) h.lvl_off[l + 1] += h.lvl_off[l];
  h.lvl_states.assign(M, 0);
  std::vector<int32_t> at(h.lvl_off.begin(), h.lvl_off.end() - 1);
  for (int v = M - 1; v >= 0; v--) h.lvl_states[at[lvl[v]]++] = v;
  h.lvl = lvl;
}

}  // namespace ixg

namespace ixg {

static int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

int qdb_tables(const ixg_cm_params &p, CmHost &h) {
  const int M = p.M;
  h.W = p.W;
  if (h.W < 1) {
    g_err = "ixg_cm_create: W must be at least 1";
    return IXG_EINVAL;
  }
  std::vector<int32_t> ndidx(M);
  for (int nd = 0, v = 0; v < M; v++) {
    while (nd + 1 < p.nodes && p.nodemap[nd + 1] <= v) nd++;
    ndidx[v] = nd;
  }
  h.qbegl.assign(M, -1);
  int nbegl = 0;
  for (int v = 0; v < M; v++)
    if (p.sttype[v] == S_ST && p.ndtype[ndidx[v]] == BEGL_ND) h.qbegl[v] = nbegl++;
  for (int v = 0; v < M; v++)
    if (p.sttype[v] == B_ST && h.qbegl[p.cfirst[v]] < 0) {
      g_err = "ixg_cm_create: a bifurcation's left child is not the start state of a BEGL node";
      return IXG_EINVAL;
    }
  // an insert state's first child is itself (its self transition), as the
  // scanners' serial IL and IR walks assume
  for (int v = 0; v < M; v++)
    if ((p.sttype[v] == IL_ST || p.sttype[v] == IR_ST) && (p.cnum[v] < 1 || p.cfirst[v] != v)) {
      g_err = "ixg_cm_create: an insert state's first child is not itself";
      return IXG_EINVAL;
    }
  // el_sca[d] = el_self added d times (cm_nohmm.rs:466-470)
  h.qfel.assign(h.W + 1, 0.f);
  h.qiel.assign(h.W + 1, 0);
  for (int d = 1; d <= h.W; d++) {
    h.qfel[d] = h.qfel[d - 1] + h.el_selfsc;
    h.qiel[d] = wadd(h.qiel[d - 1], h.iel_selfsc);
  }
  // d = 0 base cases, states in descending order (cm_nohmm.rs:489-532)
  h.qfbase0.assign(M, IMPOSSIBLE);
  h.qibase0.assign(M, -INFTY_I);
  h.qfbbase0.assign(std::max(nbegl, 1), IMPOSSIBLE);
  h.qibbase0.assign(std::max(nbegl, 1), -INFTY_I);
  for (int v = M - 1; v >= 0; v--) {
    const int st = p.sttype[v];
    if (st == E_ST) {
      h.qfbase0[v] = 0.f;
      h.qibase0[v] = 0;
    } else if (st == S_ST || st == D_ST) {
      float fa = h.endsc[v];
      int32_t ia = h.iendsc[v];
      for (int k = 0; k < p.cnum[v]; k++) {
        const int y = p.cfirst[v] + k;
        const float fc = h.qfbase0[y] + h.tsc[(size_t)v * MAXC + k];
        const int32_t ic = wadd(h.qibase0[y], h.itsc[(size_t)v * MAXC + k]);
        fa = fa >= fc ? fa : fc;
        ia = ia >= ic ? ia : ic;
      }
      fa = fa >= IMPOSSIBLE ? fa : IMPOSSIBLE;
      ia = ia >= -INFTY_I ? ia : -INFTY_I;
      if (h.qbegl[v] >= 0) {  // a BEGL_S: the value lives in its deck
        h.qfbbase0[h.qbegl[v]] = fa;
        h.qibbase0[h.qbegl[v]] = ia;
      } else {
        h.qfbase0[v] = fa;
        h.qibase0[v] = ia;
      }
    } else if (st == B_ST) {
      const int bl = h.qbegl[p.cfirst[v]], y = p.cnum[v];
      h.qfbase0[v] = h.qfbbase0[bl] + h.qfbase0[y];
      h.qibase0[v] = wadd(h.qibbase0[bl], h.qibase0[y]);
    }
  }
  h.qblist.clear();
  for (int y = 1; y < M; y++)
    if (h.beginsc[y] > IMPOSSIBLE + 1.0f) h.qblist.push_back(y);
  h.ilog.resize(FLOGSUM_TBL);  // cp9.rs:28-43 ILogsum table
  for (int i = 0; i < FLOGSUM_TBL; i++)
    h.ilog[i] = (int32_t)std::nearbyint(1000.0 * std::log2(1.0 + std::pow(2.0, -(double)i / 1000.0)));
  return IXG_OK;
}

// cp9.rs:6516-6529 score_correction_null3
float null3_score(const CmHost &h, const std::vector<float> &ftbl, const float *comp, int len) {
  float score = 0.f;
  for (int a = 0; a < 4; a++) {
    const double term = sre_log2((double)(comp[a] / h.null[a])) * (double)comp[a] * (double)len;
    score = (float)((double)score + term);
  }
  score = (float)((double)score + sre_log2((double)(float)h.n3_omega));
  return host_flogsum(ftbl, 0.0f, score);
}

}  // namespace ixg

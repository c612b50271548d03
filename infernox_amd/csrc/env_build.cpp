// Host builders for the F4/F5 envelope definition: the glocal p7 profile and
// its length / hit-mode reconfigurations (p7_generic.rs:156-392, :1075-1094),
// the nats FLogsum table (p7_generic.rs:27-69), and posterior domain decoding
// from the Forward/Backward special states (p7_generic.rs:1096-1121).
#include <cmath>
#include "ctx.h"

namespace ixg {

const std::vector<float> &p7_logsum_table() {
  static const std::vector<float> t = [] {
    std::vector<float> v(GFB_TBL);
    for (int i = 0; i < GFB_TBL; i++) v[i] = (float)std::log(1.0 + std::exp(-(double)i / 1000.0));
    return v;
  }();
  return t;
}

float p7_logsum(float a, float b) {
  const std::vector<float> &t = p7_logsum_table();
  const float mx = a > b ? a : b, mn = a > b ? b : a;
  return (mn == -INFINITY || (mx - mn) >= 15.7f) ? mx : mx + t[(int)((mx - mn) * 1000.0f)];
}

// p7_generic.rs:156-257 (glocal entry by left-wing retraction, bg 0.25, multihit)
void build_glocal(const ixg_p7_hmm &h, GmHost &g) {
  enum { H_MM = 0, H_MI = 1, H_MD = 2, H_IM = 3, H_II = 4, H_DM = 5, H_DD = 6 };
  const int M = h.M;
  g.M = M;
  g.tsc.assign((size_t)(M + 1) * 8, 0.f);
  g.msc.assign((size_t)KP * (M + 1), 0.f);
  g.isc.assign((size_t)KP * (M + 1), 0.f);
  auto tr = [&](int k, int s) { return (double)h.trans[k * 7 + s]; };
  float z = (float)std::log(tr(0, H_MD));
  g.tsc[GmHost::BM] = (float)std::log(1.0 - tr(0, H_MD));
  for (int k = 1; k < M; k++) {
    g.tsc[(size_t)k * 8 + GmHost::BM] = (float)((double)z + std::log(tr(k, H_DM)));
    z = (float)((double)z + std::log(tr(k, H_DD)));
  }
  const int hmap[8] = {H_MM, H_IM, H_DM, -1, H_MD, H_DD, H_MI, H_II};  // by GmHost index
  for (int k = 1; k < M; k++)
    for (int s = 0; s < 8; s++)
      if (hmap[s] >= 0) g.tsc[(size_t)k * 8 + s] = (float)std::log(tr(k, hmap[s]));
  // match emissions with degenerate codes as the background-weighted mean
  // (esl_abc_FExpectScVec over a uniform background)
  static const unsigned char deg[KP] = {1, 2, 4, 8, 0, 5, 10, 3, 12, 6, 9, 11, 14, 7, 13, 15, 0, 0};
  for (int k = 1; k <= M; k++) {
    float sc[KP];
    for (int x = 0; x < 4; x++) sc[x] = (float)std::log((double)h.mat[k * 4 + x] / 0.25);
    sc[4] = sc[KP - 2] = sc[KP - 1] = -INFINITY;
    for (int x = 5; x <= KP - 3; x++) {
      float num = 0.f, den = 0.f;
      for (int r = 0; r < 4; r++)
        if (deg[x] >> r & 1) {
          num += sc[r] * 0.25f;
          den += 0.25f;
        }
      sc[x] = num / den;
    }
    for (int x = 0; x < KP; x++) g.msc[(size_t)x * (M + 1) + k] = sc[x];
  }
  for (int x = 0; x < KP; x++) {
    for (int k = 1; k < M; k++) g.isc[(size_t)x * (M + 1) + k] = 0.f;
    g.isc[(size_t)x * (M + 1) + M] = -INFINITY;
  }
  for (int k = 1; k <= M; k++)
    for (int x : {4, KP - 2, KP - 1}) g.isc[(size_t)x * (M + 1) + k] = -INFINITY;
  g.set_multihit();
}

// p7_generic.rs:379-392 ReconfigLength: the N, C, J loop/move bear L
void GmHost::set_length(int L) {
  const float pmove = (2.0f + nj) / ((float)L + 2.0f + nj);
  const float ploop = 1.0f - pmove;
  const float lo = (float)std::log((double)ploop), mv = (float)std::log((double)pmove);
  for (int s : {XN, XC, XJ}) {
    xsc[s][0] = lo;
    xsc[s][1] = mv;
  }
  len = L;
}
// p7_generic.rs:350-376: the truncated length models, in which only the C
// (5') or only the N (3') state, and J, bear the length
void GmHost::set_length_5p(int L) {
  const float pmove = (1.0f + nj) / ((float)L + 1.0f + nj);
  const float ploop = 1.0f - pmove;
  const float lo = (float)std::log((double)ploop), mv = (float)std::log((double)pmove);
  xsc[XC][0] = xsc[XJ][0] = lo;
  xsc[XC][1] = xsc[XJ][1] = mv;
  len = L;
}
void GmHost::set_length_3p(int L) {
  const float pmove = (1.0f + nj) / ((float)L + 1.0f + nj);
  const float ploop = 1.0f - pmove;
  const float lo = (float)std::log((double)ploop), mv = (float)std::log((double)pmove);
  xsc[XN][0] = xsc[XJ][0] = lo;
  xsc[XN][1] = xsc[XJ][1] = mv;
  len = L;
}

// p7_generic.rs:259-348 and the L = 100 configuration of cm_search.rs:1011-1016
void build_trunc_profiles(const ixg_p7_hmm &h, const GmHost &gm, GmHost out[3]) {
  enum { H_MM = 0, H_MI = 1, H_DM = 5 };
  const int M = gm.M;
  GmHost base = gm;  // build_glocal_profile(p7, 100): multihit, length 100
  base.set_multihit();
  base.set_length(100);
  // 5' (Rgm, p7_UNIGLOCAL): uniform begins log(1/M); N and E loops off
  GmHost &r = out[0];
  r = base;
  const float inv = (float)std::log(1.0 / (double)M);
  for (int k = 1; k <= M; k++) r.tsc[(size_t)(k - 1) * 8 + GmHost::BM] = inv;
  r.local_end = false;
  r.xsc[GmHost::XN][1] = 0.0f, r.xsc[GmHost::XE][1] = 0.0f;
  r.xsc[GmHost::XN][0] = -INFINITY, r.xsc[GmHost::XE][0] = -INFINITY;
  r.set_length_5p(100);
  // 3' (Lgm, p7_UNILOCAL): glocal begins, local ends; C and E loops off
  GmHost &l = out[1];
  l = base;
  l.local_end = true;
  l.xsc[GmHost::XC][1] = 0.0f, l.xsc[GmHost::XE][1] = 0.0f;
  l.xsc[GmHost::XC][0] = -INFINITY, l.xsc[GmHost::XE][0] = -INFINITY;
  l.set_length_3p(100);
  // 5'&3' (Tgm): the local profile's occupancy-weighted begins, local ends,
  // every loop off, nj = 0, no length model
  GmHost &t = out[2];
  t = base;
  auto tr = [&](int k, int s) { return h.trans[k * 7 + s]; };
  std::vector<float> mocc(M + 1, 0.0f);
  mocc[1] = tr(0, H_MI) + tr(0, H_MM);
  for (int k = 2; k <= M; k++)
    mocc[k] = mocc[k - 1] * (tr(k - 1, H_MM) + tr(k - 1, H_MI)) + (1.0f - mocc[k - 1]) * tr(k - 1, H_DM);
  float z = 0.0f;
  for (int k = 1; k <= M; k++) z += mocc[k] * (float)(M - k + 1);
  for (int k = 1; k <= M; k++) t.tsc[(size_t)(k - 1) * 8 + GmHost::BM] = (float)std::log((double)(mocc[k] / z));
  t.local_end = true;
  t.xsc[GmHost::XN][1] = t.xsc[GmHost::XC][1] = t.xsc[GmHost::XE][1] = 0.0f;
  t.xsc[GmHost::XN][0] = t.xsc[GmHost::XC][0] = t.xsc[GmHost::XE][0] = -INFINITY;
  t.nj = 0.0f;
  t.len = 100;
}

// p7_generic.rs:1075-1094
void GmHost::set_unihit() {
  xsc[XE][1] = 0.0f;
  xsc[XE][0] = -INFINITY;
  nj = 0.0f;
}
void GmHost::set_multihit() {
  const float l2 = (float)0.693147180559945309417232121458176568;
  xsc[XE][1] = -l2;
  xsc[XE][0] = -l2;
  nj = 1.0f;
}
// special-state scores in the kernels' order [E N J C][LOOP MOVE]
void GmHost::xsc8(float out[8]) const {
  for (int s = 0; s < 4; s++) {
    out[2 * s] = xsc[s][0];
    out[2 * s + 1] = xsc[s][1];
  }
}

// p7_generic.rs:1096-1121 p7_gdomain_decoding from the specials only:
// fx / bx [(L+1)][E N J B C]. btot/etot accumulate double exp, the N/J/C
// occupancy is in float.
void domain_decoding(const GmHost &g, int L, const float *fx, const float *bx, float *btot, float *etot,
                     float *mocc) {
  auto X = [](const float *m, int i, int s) { return m[(size_t)i * 5 + s]; };
  enum { E = 0, N = 1, J = 2, B = 3, C = 4 };
  const float overall = X(fx, L, C) + g.xsc[GmHost::XC][1];
  btot[0] = etot[0] = mocc[0] = 0.0f;
  for (int i = 1; i <= L; i++) {
    btot[i] = btot[i - 1] + (float)std::exp((double)(X(fx, i - 1, B) + X(bx, i - 1, B) - overall));
    etot[i] = etot[i - 1] + (float)std::exp((double)(X(fx, i, E) + X(bx, i, E) - overall));
    float njc = expf(X(fx, i - 1, N) + X(bx, i, N) + g.xsc[GmHost::XN][0] - overall);
    njc += expf(X(fx, i - 1, J) + X(bx, i, J) + g.xsc[GmHost::XJ][0] - overall);
    njc += expf(X(fx, i - 1, C) + X(bx, i, C) + g.xsc[GmHost::XC][0] - overall);
    mocc[i] = 1.0f - njc;
  }
}

}  // namespace ixg

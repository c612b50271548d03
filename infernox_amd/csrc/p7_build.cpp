// Host-side p7 filter profile builders. These run once per model, on the
// host, exactly as in the reference; their outputs are uploaded to HBM and
// stay resident. Float evaluation order follows the reference line by line
// (compiled with -ffp-contract=off).
#include <cmath>
#include <cstring>
#include "ixg_internal.h"

namespace ixg {

static const double LN2 = 0.693147180559945309417232121458176568;

// RNA IUPAC code -> canonical residues (cm_pipeline.rs:124-141).
static int degen(int code, int out[4]) {
  switch (code) {
    case 0: case 1: case 2: case 3: out[0] = code; return 1;
    case 5: out[0] = 0; out[1] = 2; return 2;
    case 6: out[0] = 1; out[1] = 3; return 2;
    case 7: out[0] = 0; out[1] = 1; return 2;
    case 8: out[0] = 2; out[1] = 3; return 2;
    case 9: out[0] = 1; out[1] = 2; return 2;
    case 10: out[0] = 0; out[1] = 3; return 2;
    case 11: out[0] = 0; out[1] = 1; out[2] = 3; return 3;
    case 12: out[0] = 1; out[1] = 2; out[2] = 3; return 3;
    case 13: out[0] = 0; out[1] = 1; out[2] = 2; return 3;
    case 14: out[0] = 0; out[1] = 2; out[2] = 3; return 3;
    case 15: out[0] = 0; out[1] = 1; out[2] = 2; out[3] = 3; return 4;
    default: return 0;
  }
}

// MSC[k][x] (modelconfig.c:141-151 as restated at cm_pipeline.rs:216-239).
static void msc_row(const ixg_p7_hmm &h, int k, float sc[KP]) {
  for (int x = 0; x < KP; x++) sc[x] = -INFINITY;
  for (int x = 0; x < 4; x++) sc[x] = (float)std::log((double)h.mat[k * 4 + x] / 0.25);
  for (int c = 4; c < KP; c++) {
    int set[4];
    int n = degen(c, set);
    if (!n) continue;
    float s = 0.f;
    for (int i = 0; i < n; i++) s += sc[set[i]];
    sc[c] = s / (float)n;
  }
}

uint8_t unbiased_byteify(float scale_b, float sc) {  // cm_pipeline.rs:157
  float v = -std::roundf(scale_b * sc);
  if (v >= 255.f) return 255;
  if (v <= 0.f) return 0;
  return (uint8_t)v;
}
static uint8_t biased_byteify(float scale_b, uint8_t bias_b, float sc) {  // cm_pipeline.rs:172
  float v = -std::roundf(scale_b * sc);
  if (v > (float)(255 - (int)bias_b)) return 255;
  int r = (int)v + (int)bias_b;
  return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
}

void build_msv(const ixg_p7_hmm &h, int max_length, MsvHost &f) {  // cm_pipeline.rs:204
  const int m = h.M;
  f.m = m;
  f.scale_b = (float)(3.0 / LN2);
  f.base_b = 190;
  f.max_length = max_length;
  f.mmu = h.lmmu;
  f.mlambda = h.lmlambda;
  std::vector<float> msc((size_t)(m + 1) * KP, -INFINITY);
  float maxsc = 0.f;
  for (int k = 1; k <= m; k++) {
    msc_row(h, k, &msc[(size_t)k * KP]);
    for (int x = 0; x < 4; x++) maxsc = msc[k * KP + x] > maxsc ? msc[k * KP + x] : maxsc;
  }
  f.bias_b = unbiased_byteify(f.scale_b, -maxsc);
  f.rbv.assign((size_t)(m + 1) * KP, 255);
  for (int k = 1; k <= m; k++)
    for (int x = 0; x < KP; x++) f.rbv[k * KP + x] = biased_byteify(f.scale_b, f.bias_b, msc[k * KP + x]);
  f.tbm_b = unbiased_byteify(f.scale_b, (float)std::log(2.0 / ((double)m * ((double)m + 1.0))));
  f.tec_b = unbiased_byteify(f.scale_b, std::log(0.5f));
  f.tjb_b = unbiased_byteify(f.scale_b, (float)std::log(3.0 / ((double)max_length + 3.0)));
  // prefix/suffix window-extension fractions (cm_pipeline.rs:281-324)
  f.prefix.assign(m + 1, 0.f);
  f.suffix.assign(m + 1, 0.f);
  float sum = 0.f;
  for (int k = 1; k < m; k++) {
    double t_mi = (double)h.trans[k * 7 + 1], t_ii = (double)h.trans[k * 7 + 4];
    float pk = 1.f;
    if (t_mi != 0.0) {
      double lii = std::log(t_ii);
      if (lii != 0.0 && std::isfinite(lii)) pk = (float)(1.0 + std::trunc(std::log(1e-7 / t_mi) / lii));
    }
    f.prefix[k] = pk;
    sum += pk;
  }
  f.prefix[0] = 0.f;
  f.prefix[m] = 0.f;
  if (sum > 0.f)
    for (int k = 1; k < m; k++) f.prefix[k] /= sum;
  if (m >= 1) {
    f.suffix[m] = f.prefix[m - 1];
    for (int k = m - 1; k >= 1; k--) f.suffix[k] = f.suffix[k + 1] + f.prefix[k - 1];
    for (int k = 2; k < m; k++) f.prefix[k] += f.prefix[k - 1];
  }
}

int16_t wordify(float scale_w, float sc) {  // p7_vitfilter.rs:57
  if (sc == -INFINITY) return NEGINF_W;
  float v = std::roundf(scale_w * sc);
  if (v >= 32767.f) return 32767;
  if (v <= -32768.f) return NEGINF_W;
  return (int16_t)v;
}

static void occupancy(const ixg_p7_hmm &h, std::vector<float> &occ, float &z) {
  const int m = h.M;
  occ.assign(m + 1, 0.f);
  if (m >= 1) occ[1] = h.trans[1] + h.trans[0];
  for (int k = 2; k <= m; k++) {
    const float *t = h.trans + (size_t)(k - 1) * 7;
    occ[k] = occ[k - 1] * (t[0] + t[1]) + (1.0f - occ[k - 1]) * t[5];
  }
  z = 0.f;
  for (int k = 1; k <= m; k++) z += occ[k] * (float)(m - k + 1);
}

void build_vit(const ixg_p7_hmm &h, VitHost &f) {  // p7_vitfilter.rs:121
  const int m = h.M;
  f.m = m;
  f.scale_w = (float)(500.0 / LN2);
  f.base_w = 12000;
  f.rwv.assign((size_t)(m + 1) * KP, NEGINF_W);
  for (int k = 1; k <= m; k++) {
    float sc[KP];
    msc_row(h, k, sc);
    for (int x = 0; x < KP; x++) f.rwv[k * KP + x] = wordify(f.scale_w, sc[x]);
  }
  for (auto *a : {&f.tmm, &f.tim, &f.tdm, &f.vbm, &f.tmi, &f.tii, &f.tmd, &f.tdd}) a->assign(m + 1, NEGINF_W);
  auto w0 = [&](float p, int16_t cap) {
    int16_t v = wordify(f.scale_w, (float)std::log((double)p));
    return v < cap ? v : cap;
  };
  for (int j = 1; j < m; j++) {
    const float *t = h.trans + (size_t)j * 7;
    f.tmm[j] = w0(t[0], 0);
    f.tim[j] = w0(t[3], 0);
    f.tdm[j] = w0(t[5], 0);
    f.tmi[j] = w0(t[1], 0);
    f.tii[j] = w0(t[4], -1);
    f.tmd[j] = w0(t[2], 0);
    f.tdd[j] = w0(t[6], 0);
  }
  std::vector<float> occ;
  float z;
  occupancy(h, occ, z);
  for (int k = 1; k <= m; k++) {
    int16_t v = wordify(f.scale_w, (float)std::log((double)(occ[k] / z)));
    f.vbm[k] = v < 0 ? v : 0;
  }
  f.xw_e_move = wordify(f.scale_w, -(float)LN2);
  f.xw_e_loop = f.xw_e_move;
  int16_t ddb = NEGINF_W;
  for (int k = 2; k < m - 1; k++) {
    int dd = (int)f.tdd[k] + (int)f.tdm[k + 1] - (int)f.vbm[k + 2];
    ddb = (int16_t)((int)ddb > dd ? (int)ddb : dd);
  }
  f.ddbound_w = ddb;
}

void build_fwd(const ixg_p7_hmm &h, FwdHost &f) {  // cm_pipeline.rs:969
  const int m = h.M;
  f.m = m;
  f.ftau = h.lftau;
  f.flambda = h.lflambda;
  f.rfv.assign((size_t)(m + 1) * KP, 0.f);
  for (int k = 1; k <= m; k++) {
    float sc[KP];
    msc_row(h, k, sc);
    for (int x = 0; x < KP; x++) f.rfv[k * KP + x] = std::exp(sc[x]);
  }
  for (auto *a : {&f.tbm, &f.amm, &f.aim, &f.adm, &f.tmi, &f.tii, &f.tmd, &f.tdd}) a->assign(m + 1, 0.f);
  std::vector<float> occ;
  float z;
  occupancy(h, occ, z);
  for (int k = 1; k <= m; k++) f.tbm[k] = (float)std::exp(std::log((double)(occ[k] / z)));
  for (int k = 1; k <= m; k++) {
    const float *t = h.trans + (size_t)(k - 1) * 7;
    f.amm[k] = t[0];
    f.aim[k] = t[3];
    f.adm[k] = t[5];
  }
  for (int k = 1; k < m; k++) {
    const float *t = h.trans + (size_t)k * 7;
    f.tmi[k] = t[1];
    f.tii[k] = t[4];
    f.tmd[k] = t[2];
    f.tdd[k] = t[6];
  }
  int nq = m >= 1 ? (m - 1) / 4 + 1 : 2;
  if (nq < 2) nq = 2;
  f.nq = nq;
  f.tmain.assign((size_t)nq * 28, 0.f);
  f.tddv.assign((size_t)nq * 4, 0.f);
  f.rfvv.assign((size_t)KP * nq * 4, 0.f);
  for (int q = 0; q < nq; q++)
    for (int zz = 0; zz < 4; zz++) {
      int p = q + 1 + zz * nq;
      bool ok = p <= m;
      const float *src[7] = {f.tbm.data(), f.amm.data(), f.aim.data(), f.adm.data(),
                             f.tmd.data(), f.tmi.data(), f.tii.data()};
      for (int t = 0; t < 7; t++) f.tmain[q * 28 + t * 4 + zz] = ok ? src[t][p] : 0.f;
      f.tddv[q * 4 + zz] = ok ? f.tdd[p] : 0.f;
      for (int x = 0; x < KP; x++) f.rfvv[x * (nq * 4) + q * 4 + zz] = ok ? f.rfv[p * KP + x] : 0.f;
    }
}

void build_bias(const float compo[4], int m, BiasHost &b) {  // cm_pipeline.rs:1683
  const float bg = 0.25f;
  const float l1 = (float)m / 8.0f;
  for (int c = 0; c < KP; c++) {
    b.eo[c][0] = 1.f;
    int set[4];
    int n = degen(c, set);
    if (!n) {
      b.eo[c][1] = 1.f;
    } else {
      float num = 0.f, den = 0.f;
      for (int i = 0; i < n; i++) {
        num += compo[set[i]];
        den += bg;
      }
      b.eo[c][1] = den > 0.f ? num / den : 0.f;
    }
  }
  b.t1[0] = 1.0f / (l1 + 1.0f);
  b.t1[1] = l1 / (l1 + 1.0f);
  b.t1[2] = 1.0f;
  b.pi[0] = 0.999f;
  b.pi[1] = 0.001f;
}

static double gumbel_invsurv(double p, double mu, double lambda) {  // easel/gumbel.rs:168
  double lp = p < 5e-9 ? (std::pow(p, p) - 1.0) / p : std::log(-std::log(1.0 - p));
  return mu - lp / lambda;
}
double gumbel_surv(double x, double mu, double lambda) {  // cm_pipeline.rs:1852
  double y = lambda * (x - mu);
  double ey = -std::exp(-y);
  if (std::fabs(ey) < 5e-9) return -ey;
  return 1.0 - std::exp(ey);
}
double exp_surv(double x, double mu, double lambda) {  // cp9.rs:6531
  return x < mu ? 1.0 : std::exp(-lambda * (x - mu));
}
double bg_null_one(int64_t L) {  // cm_pipeline.rs:1122
  double p1 = (double)L / ((double)L + 1.0);
  return (double)L * std::log(p1) + std::log(1.0 - p1);
}
int ssv_thresh(const MsvHost &f, double f1) {  // cm_pipeline.rs:522-533
  double p1 = (double)f.max_length / ((double)f.max_length + 1.0);
  double nullsc = (double)f.max_length * std::log(p1) + std::log(1.0 - p1);
  double invp = gumbel_invsurv(f1, f.mmu, f.mlambda);
  return (int)std::ceil((nullsc + invp * LN2 + 3.0) * (double)f.scale_b + (double)f.base_b +
                        (double)f.tec_b + (double)f.tjb_b);
}

}  // namespace ixg

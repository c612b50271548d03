// gfx950 kernels for the F4/F5 envelope definition: the generic glocal
// Forward and Backward of a p7 profile in nats with HMMER's FLogsum table
// (p7_generic.rs:559-676 p7_gforward, :939-1073 p7_gbackward), bit-exact
// (built with -ffp-contract=off, same operations in the same order).
//
// Mapping: a workgroup of four waves, each wave one item (a window, a region
// or an envelope with its own length configuration). The 16,000-entry
// FLogsum table and the profile's MD / DD / BM transitions sit in LDS, shared
// by the four waves. Lane l holds model positions k = 1 + l*KPL .. l*KPL+KPL.
// Within a row the M and I cells depend only on the row before and are
// computed in parallel over k; the D chain (and in the Backward the B sum,
// with local ends also the E terms) is a serial recurrence over k in a fixed
// order, which lane 0 runs out of LDS with one table lookup per step while
// the other lanes wait; what depends on the chain's values but not on the
// next step (the E sum in glocal mode, the Backward's M cells) is computed
// around it in parallel. The specials (E N J B C) are carried by lane 0.
#include "ixg_internal.h"

namespace ixg {

namespace {

constexpr int GFB_WAVES = 4;
enum { T_MM = 0, T_IM = 1, T_DM = 2, T_BM = 3, T_MD = 4, T_DD = 5, T_MI = 6, T_II = 7 };
enum { XS_EL = 0, XS_EM = 1, XS_NL = 2, XS_NM = 3, XS_JL = 4, XS_JM = 5, XS_CL = 6, XS_CM = 7 };

// FLogsum in nats (p7_generic.rs FLogsum): max + T[(int)(diff * 1000)] when
// the smaller term is finite and diff < 15.7, else the max. Branch-free with
// a short dependency chain (it is the serial step of the D recurrences): the
// LDS copy of the table holds 0 from index LSUM_CUT on and the index is
// clamped to LSUM_CUT, so every case that returns the max reads a 0:
//  * diff >= 15.7f gives diff * 1000 >= 15700 after float rounding (and
//    diff < 15.7f gives at most 15699.999, index 15699, as the reference);
//  * a -inf smaller term gives diff = +inf, converted to INT_MAX, clamped;
//  * both -inf gives NaN, converted to 0: -inf + T[0] = -inf = the max.
// |a - b| is exactly mx - mn (IEEE subtraction is sign-symmetric), and the
// sign of a zero max never matters (equal terms always take the table).
constexpr int LSUM_CUT = 15700;
__device__ __forceinline__ float lsum(const float *tbl, float a, float b) {
  const float mx = fmaxf(a, b);
  const int idx = min((int)(fabsf(a - b) * 1000.0f), LSUM_CUT);
  return mx + tbl[idx];
}

// Orders this wave's LDS writes before its later LDS reads (cross-lane).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The transition table of an item's profile: the profiles of one model
// (glocal, and the truncated passes' 5', 3' and 5'&3' profiles) differ only
// in their begin transitions and sit one after another; flags bits 2-3 pick
// one. MD and DD, staged in LDS, are the same in all of them.
__device__ __forceinline__ const float *item_tsc(const GmDev &g, const GfbItem &item) {
  return g.tsc + (size_t)((item.flags >> 2) & 3) * (size_t)(g.m + 1) * 8;
}

struct GfbShared {
  float *tbl, *md, *dd, *bm;  // table; per k = 0..M: MD, DD, BM (tsc row k)
  float *wave;                // per wave: 3 rows of M+2 floats
};

__device__ GfbShared gfb_stage(const GmDev &g, float *smem) {
  GfbShared s;
  const int M = g.m;
  s.tbl = smem;
  s.md = smem + GFB_TBL;
  s.dd = s.md + (M + 1);
  s.bm = s.dd + (M + 1);
  s.wave = s.bm + (M + 1);
  for (int t = threadIdx.x; t < GFB_TBL; t += blockDim.x) s.tbl[t] = t < LSUM_CUT ? g.tbl[t] : 0.0f;
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    s.md[k] = g.tsc[k * 8 + T_MD];
    s.dd[k] = g.tsc[k * 8 + T_DD];
    s.bm[k] = g.tsc[k * 8 + T_BM];
  }
  __syncthreads();
  return s;
}

// Per-lane copies of the transitions a lane's positions use (k clamped to M
// on the positions past M, whose results are discarded).
template <int KPL>
struct LaneTsc {
  float mm[KPL], im[KPL], dm[KPL], bm[KPL], mi[KPL], ii[KPL], md[KPL];
};

// Forward, one wave per item. Per row: (A) all lanes compute their M and I
// cells from the row before and leave M(i,k) + MD(k) for the D chain; (B)
// lane 0 runs the D chain, the only serial recurrence of the row (one FLogsum
// per position, its operands loaded a step ahead), then E and the specials;
// (C) all lanes pick up their D cells and the new B. In glocal mode (LE
// false) the reference's E sum over k < M adds only -inf terms, so E is
// FLogsum(M(i,M), D(i,M)) exactly; with local ends (LE) the E sum runs along
// the chain in the reference's order. The next row's emission scores are
// fetched while the chain runs.
template <int KPL, bool LE>
__device__ __forceinline__ void gfwd_item(const GmDev &g, const GfbShared &S, const uint8_t *seq,
                                          const GfbItem &item, int64_t it, float *xout, float *mout, float *sc) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int M = g.m, W = M + 2;
  const float *tsc = item_tsc(g, item);
  // sA[k] = M(i,k-1) + MD(k-1) (k = 1..M), sD[k] = D(i,k), sM[k] = M(i,k),
  // sM[M+1] = the row's B for the next row
  float *sA = S.wave + (size_t)wv * 3 * W, *sD = sA + W, *sM = sD + W;
  const int L = item.len;
  const bool keep = item.flags & 2;
  float xs[8];
#pragma unroll
  for (int t = 0; t < 8; t++) xs[t] = item.xsc[t];
  const uint8_t *dsq = seq + item.off;
  IXG_GLOBAL float *xm = gmem(xout) + item.xoff;
  IXG_GLOBAL float *dp = keep ? gmem(mout) + item.moff : nullptr;
  const int kb = 1 + lane * KPL;
  LaneTsc<KPL> T;
  int kk[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    kk[j] = kb + j <= M ? kb + j : M;
    const float *t = tsc + (size_t)(kk[j] - 1) * 8, *tk = tsc + (size_t)kk[j] * 8;
    T.mm[j] = t[T_MM];
    T.im[j] = t[T_IM];
    T.dm[j] = t[T_DM];
    T.bm[j] = t[T_BM];
    T.mi[j] = tk[T_MI];
    T.ii[j] = tk[T_II];
    T.md[j] = tk[T_MD];
  }
  float mp[KPL], ip[KPL], dpv[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) mp[j] = ip[j] = dpv[j] = -INFINITY;
  // row 0
  float xN = 0.0f, xB = xs[XS_NM], xE = -INFINITY, xJ = -INFINITY, xC = -INFINITY;
  if (lane == 0) {
    xm[0] = xE;
    xm[1] = xN;
    xm[2] = xJ;
    xm[3] = xB;
    xm[4] = xC;
    sA[1] = -INFINITY;  // M(i,0) = -inf
  }
  if (keep)
    for (int c = lane; c < 3 * (M + 1); c += 64) dp[c] = -INFINITY;
  float em[KPL], ei[KPL];
  {
    const int x = L >= 1 ? dsq[1] : 0;
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      em[j] = g.msc[(size_t)x * (M + 1) + kk[j]];
      ei[j] = g.isc[(size_t)x * (M + 1) + kk[j]];
    }
  }
  for (int i = 1; i <= L; i++) {
    const int xn = i < L ? dsq[i + 1] : 0;
    float mpl = __shfl_up(mp[KPL - 1], 1), ipl = __shfl_up(ip[KPL - 1], 1), dpl = __shfl_up(dpv[KPL - 1], 1);
    if (lane == 0) mpl = ipl = dpl = -INFINITY;
    float mc[KPL], ic[KPL];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      const float pm = j ? mp[j - 1] : mpl, pi = j ? ip[j - 1] : ipl, pd = j ? dpv[j - 1] : dpl;
      const float s = lsum(S.tbl, lsum(S.tbl, pm + T.mm[j], pi + T.im[j]), lsum(S.tbl, xB + T.bm[j], pd + T.dm[j]));
      mc[j] = k <= M ? s + em[j] : -INFINITY;
      ic[j] = k < M ? lsum(S.tbl, mp[j] + T.mi[j], ip[j] + T.ii[j]) + ei[j] : -INFINITY;
      if (k < M) sA[k + 1] = mc[j] + T.md[j];
      if (LE ? k <= M : k == M) sM[k] = mc[j];
    }
    // the next row's emission scores, in flight while the chain runs
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      em[j] = g.msc[(size_t)xn * (M + 1) + kk[j]];
      ei[j] = g.isc[(size_t)xn * (M + 1) + kk[j]];
    }
    wave_lds_sync();
    if (lane == 0) {
      // D chain (p7_generic.rs:93-117)
      float d = -INFINITY, e = -INFINITY;
      float an = sA[1], ddn = S.dd[0];
      for (int k = 1; k <= M; k++) {
        const float ak = an, ddk = ddn;
        an = sA[k + 1];  // k + 1 <= M + 1 < W; unused past M
        ddn = S.dd[k];
        d = lsum(S.tbl, ak, d + ddk);
        sD[k] = d;
        if (LE) {
          const float mk = sM[k];
          e = k < M ? lsum(S.tbl, lsum(S.tbl, mk + 0.0f, d + 0.0f), e) : lsum(S.tbl, lsum(S.tbl, mk, d), e);
        }
      }
      xE = LE ? e : lsum(S.tbl, sM[M], d);
      xJ = lsum(S.tbl, xJ + xs[XS_JL], xE + xs[XS_EL]);
      xC = lsum(S.tbl, xC + xs[XS_CL], xE + xs[XS_EM]);
      xN = xN + xs[XS_NL];
      xB = lsum(S.tbl, xN + xs[XS_NM], xJ + xs[XS_JM]);
      IXG_GLOBAL float *r = xm + (size_t)i * 5;
      r[0] = xE;
      r[1] = xN;
      r[2] = xJ;
      r[3] = xB;
      r[4] = xC;
      sM[M + 1] = xB;
    }
    wave_lds_sync();
    xB = sM[M + 1];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      dpv[j] = k <= M ? sD[k] : -INFINITY;
      mp[j] = mc[j];
      ip[j] = ic[j];
    }
    if (keep) {
      IXG_GLOBAL float *row = dp + (size_t)i * 3 * (M + 1);
#pragma unroll
      for (int j = 0; j < KPL; j++) {
        const int k = kb + j;
        if (k <= M) {
          row[k] = mp[j];
          row[(M + 1) + k] = ip[j];
          row[2 * (M + 1) + k] = dpv[j];
        }
      }
      if (lane == 0) row[0] = row[M + 1] = row[2 * (M + 1)] = -INFINITY;
    }
    wave_lds_sync();  // this row's LDS values are read before the next row rewrites them
  }
  if (lane == 0) sc[it] = xC + xs[XS_CM];
}

// Backward, one wave per item. Per row: (A) all lanes compute the B terms and
// the parts of M, I and D that depend only on the row after; (B) lane 0 folds
// B in increasing k, updates the specials and runs the D chain down from
// k = M-1 (one FLogsum per position in glocal mode: the E term of the
// reference's inner FLogsum is -inf there), operands loaded a step ahead;
// (C) all lanes finish their M cells from the chain's D values. With local
// ends (LE) the inner FLogsums with E stay, in the reference's order. The
// next row's emission scores are fetched while lane 0 works.
template <int KPL, bool LE>
__device__ __forceinline__ void gbwd_item(const GmDev &g, const GfbShared &S, const uint8_t *seq,
                                          const GfbItem &item, int64_t it, float *xout, float *sc) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int M = g.m, W = M + 2;
  const float *tsc = item_tsc(g, item);
  // sT: the B terms, then D(i,k) (sT[M] = E); sP1 / sP2: the M / D parts
  // from row i+1
  float *sT = S.wave + (size_t)wv * 3 * W, *sP1 = sT + W, *sP2 = sP1 + W;
  const int L = item.len;
  const float esc = LE ? 0.0f : -INFINITY;
  float xs[8];
#pragma unroll
  for (int t = 0; t < 8; t++) xs[t] = item.xsc[t];
  const uint8_t *dsq = seq + item.off;
  IXG_GLOBAL float *xm = gmem(xout) + item.xoff;
  const int kb = 1 + lane * KPL;
  // transitions at k (row k of tsc) and BM at k-1
  float tmm[KPL], tmi[KPL], tim[KPL], tii[KPL], tdm[KPL], tmd[KPL], tbm[KPL];
  int kk[KPL], kn[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    kk[j] = kb + j <= M ? kb + j : M;
    kn[j] = kb + j < M ? kb + j + 1 : M;  // k + 1, clamped
    const float *t = tsc + (size_t)kk[j] * 8;
    tmm[j] = t[T_MM];
    tmi[j] = t[T_MI];
    tim[j] = t[T_IM];
    tii[j] = t[T_II];
    tdm[j] = t[T_DM];
    tmd[j] = t[T_MD];
    tbm[j] = tsc[(size_t)(kk[j] - 1) * 8 + T_BM];
  }
  // row L (p7_generic.rs:946-962)
  float xJ = -INFINITY, xB = -INFINITY, xN = -INFINITY, xC = xs[XS_CM];
  float xE = xC + xs[XS_EM];
  if (lane == 0) {
    IXG_GLOBAL float *r = xm + (size_t)L * 5;
    r[0] = xE;
    r[1] = xN;
    r[2] = xJ;
    r[3] = xB;
    r[4] = xC;
    float dn = xE;
    sP1[M] = xE;
    sP2[M] = xE;
    for (int k = M - 1; k >= 1; k--) {
      sP1[k] = lsum(S.tbl, xE + esc, dn + S.md[k]);
      const float dk = lsum(S.tbl, xE + esc, dn + S.dd[k]);
      sP2[k] = dk;
      dn = dk;
    }
  }
  wave_lds_sync();
  float mn[KPL], in_[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    const int k = kb + j;
    mn[j] = k <= M ? sP1[k] : -INFINITY;
    in_[j] = -INFINITY;
  }
  wave_lds_sync();
  // emission scores of row i+1's residue: em at k, em1 at k+1, ei at k
  float em[KPL], em1[KPL], ei[KPL];
  {
    const int x = L >= 1 ? dsq[L] : 0;
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      em[j] = g.msc[(size_t)x * (M + 1) + kk[j]];
      em1[j] = g.msc[(size_t)x * (M + 1) + kn[j]];
      ei[j] = g.isc[(size_t)x * (M + 1) + kk[j]];
    }
  }
  for (int i = L - 1; i >= 1; i--) {
    const int xn = dsq[i];  // the residue of row i, for row i-1
    float mn1 = __shfl_down(mn[0], 1);  // M(i+1, k+1) for the last position of this lane
    if (lane == 63) mn1 = -INFINITY;
    float ic[KPL];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      ic[j] = -INFINITY;
      if (k <= M) {
        sT[k] = mn[j] + tbm[j] + em[j];
        if (k < M) {
          const float mk1 = (j + 1 < KPL) ? mn[j + 1] : mn1;
          sP1[k] = lsum(S.tbl, mk1 + tmm[j] + em1[j], in_[j] + tmi[j] + ei[j]);
          ic[j] = lsum(S.tbl, mk1 + tim[j] + em1[j], in_[j] + tii[j] + ei[j]);
          sP2[k] = mk1 + tdm[j] + em1[j];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      em[j] = g.msc[(size_t)xn * (M + 1) + kk[j]];
      em1[j] = g.msc[(size_t)xn * (M + 1) + kn[j]];
      ei[j] = g.isc[(size_t)xn * (M + 1) + kk[j]];
    }
    wave_lds_sync();
    if (lane == 0) {
      float b = sT[1], tn = sT[2];
      for (int k = 2; k <= M; k++) {
        const float tk = tn;
        tn = sT[k + 1];  // k + 1 <= M + 1 < W; unused past M
        b = lsum(S.tbl, b, tk);
      }
      xB = b;
      xJ = lsum(S.tbl, xJ + xs[XS_JL], xB + xs[XS_JM]);
      xC = xC + xs[XS_CL];
      xE = lsum(S.tbl, xJ + xs[XS_EL], xC + xs[XS_EM]);
      xN = lsum(S.tbl, xN + xs[XS_NL], xB + xs[XS_NM]);
      IXG_GLOBAL float *r = xm + (size_t)i * 5;
      r[0] = xE;
      r[1] = xN;
      r[2] = xJ;
      r[3] = xB;
      r[4] = xC;
      // D chain from k = M-1 down (p7_generic.rs:1012-1042)
      float dn = xE;
      sT[M] = xE;
      float pn = sP2[M - 1], ddn = S.dd[M - 1];
      for (int k = M - 1; k >= 1; k--) {
        const float p2 = pn, ddk = ddn;
        pn = sP2[k - 1];  // k - 1 >= 0; unused at k = 1
        ddn = S.dd[k - 1];
        const float dk = LE ? lsum(S.tbl, p2, lsum(S.tbl, dn + ddk, xE + esc)) : lsum(S.tbl, p2, dn + ddk);
        sT[k] = dk;
        dn = dk;
      }
    }
    wave_lds_sync();
    const float xe = sT[M];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      if (k < M) {
        const float dn = sT[k + 1];  // D(i, k+1)
        mn[j] = LE ? lsum(S.tbl, sP1[k], lsum(S.tbl, xe + esc, dn + tmd[j])) : lsum(S.tbl, sP1[k], dn + tmd[j]);
      } else {
        mn[j] = k == M ? xe : -INFINITY;
      }
      in_[j] = ic[j];
    }
    wave_lds_sync();
  }
  // row 0 (p7_generic.rs:1044-1066); em holds row 1's residue
  {
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      if (k <= M) sT[k] = mn[j] + tbm[j] + em[j];
    }
    wave_lds_sync();
    if (lane == 0) {
      float b = sT[1];
      for (int k = 2; k <= M; k++) b = lsum(S.tbl, b, sT[k]);
      xB = b;
      const float n0 = lsum(S.tbl, xN + xs[XS_NL], xB + xs[XS_NM]);
      xm[0] = -INFINITY;
      xm[1] = n0;
      xm[2] = -INFINITY;
      xm[3] = xB;
      xm[4] = -INFINITY;
      sc[it] = n0;
    }
  }
}

// One item per wave; an item's direction (flags bit 4: Backward) and
// local-end mode (flags bit 0) select the body, uniformly over the wave, so
// Forward and Backward items share a launch.
template <int KPL>
__global__ __launch_bounds__(64 * GFB_WAVES) void gfb_kernel(GmDev g, const uint8_t *seq, const GfbItem *items,
                                                             int32_t n, float *xout, float *mout, float *sc) {
  extern __shared__ float smem[];
  const GfbShared S = gfb_stage(g, smem);
  const int64_t it = (int64_t)blockIdx.x * GFB_WAVES + (threadIdx.x >> 6);
  if (it >= n) return;
  const GfbItem item = items[it];
  const bool le = item.flags & 1;
  if (item.flags & 16) {
    if (le) gbwd_item<KPL, true>(g, S, seq, item, it, xout, sc);
    else gbwd_item<KPL, false>(g, S, seq, item, it, xout, sc);
  } else {
    if (le) gfwd_item<KPL, true>(g, S, seq, item, it, xout, mout, sc);
    else gfwd_item<KPL, false>(g, S, seq, item, it, xout, mout, sc);
  }
}

template <int KPL>
hipError_t gfb_launch(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout,
                      float *mout, float *sc, hipStream_t s) {
  const size_t lds = ((size_t)GFB_TBL + 3 * (size_t)(g.m + 1) + (size_t)GFB_WAVES * 3 * (g.m + 2)) * sizeof(float);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((n + GFB_WAVES - 1) / GFB_WAVES);
  hipLaunchKernelGGL((gfb_kernel<KPL>), dim3(grid), dim3(64 * GFB_WAVES), lds, s, g, seq, items, (int32_t)n, xout,
                     mout, sc);
  return hipGetLastError();
}

hipError_t launch_gfb_kpl(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout,
                          float *mout, float *sc, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (g.m < 2 || n > INT32_MAX) return hipErrorInvalidValue;
  const int kpl = (g.m + 63) / 64;
  if (kpl <= 1) return gfb_launch<1>(g, seq, items, n, xout, mout, sc, s);
  if (kpl <= 2) return gfb_launch<2>(g, seq, items, n, xout, mout, sc, s);
  if (kpl <= 4) return gfb_launch<4>(g, seq, items, n, xout, mout, sc, s);
  if (kpl <= 8) return gfb_launch<8>(g, seq, items, n, xout, mout, sc, s);
  if (kpl <= 16) return gfb_launch<16>(g, seq, items, n, xout, mout, sc, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_gfb(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout, float *mout,
                      float *sc, hipStream_t s) {
  return launch_gfb_kpl(g, seq, items, n, xout, mout, sc, s);
}

}  // namespace ixg

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
// computed in parallel over k; the D cells (and, with local ends, the E sum;
// in the Backward the M cells through their D term, and the B sum) are
// serial chains over k in a fixed order, which lane 0 runs out of LDS while
// the other lanes wait. The specials (E N J B C) are wave-uniform scalars.
#include "ixg_internal.h"

namespace ixg {

namespace {

constexpr int GFB_WAVES = 4;
enum { T_MM = 0, T_IM = 1, T_DM = 2, T_BM = 3, T_MD = 4, T_DD = 5, T_MI = 6, T_II = 7 };
enum { XS_EL = 0, XS_EM = 1, XS_NL = 2, XS_NM = 3, XS_JL = 4, XS_JM = 5, XS_CL = 6, XS_CM = 7 };

__device__ __forceinline__ float lsum(const float *tbl, float a, float b) {
  const float mx = a > b ? a : b, mn = a > b ? b : a;
  return (mn == -INFINITY || (mx - mn) >= 15.7f) ? mx : mx + tbl[(int)((mx - mn) * 1000.0f)];
}

// Orders this wave's LDS writes before its later LDS reads (cross-lane).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
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
  for (int t = threadIdx.x; t < GFB_TBL; t += blockDim.x) s.tbl[t] = g.tbl[t];
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    s.md[k] = g.tsc[k * 8 + T_MD];
    s.dd[k] = g.tsc[k * 8 + T_DD];
    s.bm[k] = g.tsc[k * 8 + T_BM];
  }
  __syncthreads();
  return s;
}

template <int KPL>
__global__ __launch_bounds__(64 * GFB_WAVES) void gfwd_kernel(GmDev g, const uint8_t *seq, const GfbItem *items,
                                                              int32_t n, float *xout, float *mout, float *sc) {
  extern __shared__ float smem[];
  const GfbShared S = gfb_stage(g, smem);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t it = (int64_t)blockIdx.x * GFB_WAVES + wv;
  if (it >= n) return;
  const int M = g.m, W = M + 2;
  float *sM = S.wave + (size_t)wv * 3 * W, *sD = sM + W;
  const GfbItem item = items[it];
  const int L = item.len;
  const bool local_end = item.flags & 1, keep = item.flags & 2;
  const float esc = local_end ? 0.0f : -INFINITY;
  float xs[8];
#pragma unroll
  for (int t = 0; t < 8; t++) xs[t] = item.xsc[t];
  const uint8_t *dsq = seq + item.off;
  IXG_GLOBAL float *xm = gmem(xout) + item.xoff;
  IXG_GLOBAL float *dp = keep ? gmem(mout) + item.moff : nullptr;
  const int kb = 1 + lane * KPL;
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
  }
  if (keep)
    for (int c = lane; c < 3 * (M + 1); c += 64) dp[c] = -INFINITY;
  for (int i = 1; i <= L; i++) {
    const int x = dsq[i];
    const float *msx = g.msc + (size_t)x * (M + 1), *isx = g.isc + (size_t)x * (M + 1);
    float mpl = __shfl_up(mp[KPL - 1], 1), ipl = __shfl_up(ip[KPL - 1], 1), dpl = __shfl_up(dpv[KPL - 1], 1);
    if (lane == 0) mpl = ipl = dpl = -INFINITY;
    float mc[KPL], ic[KPL];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      mc[j] = ic[j] = -INFINITY;
      if (k <= M) {
        const float pm = j ? mp[j - 1] : mpl, pi = j ? ip[j - 1] : ipl, pd = j ? dpv[j - 1] : dpl;
        const float *t = g.tsc + (size_t)(k - 1) * 8;
        const float s = lsum(S.tbl, lsum(S.tbl, pm + t[T_MM], pi + t[T_IM]), lsum(S.tbl, xB + t[T_BM], pd + t[T_DM]));
        mc[j] = s + msx[k];
        if (k < M) {
          const float *tk = g.tsc + (size_t)k * 8;
          ic[j] = lsum(S.tbl, mp[j] + tk[T_MI], ip[j] + tk[T_II]) + isx[k];
        }
        sM[k] = mc[j];
      }
    }
    wave_lds_sync();
    // D chain and the E sum, serial over k (p7_generic.rs:93-117)
    float e = -INFINITY;
    if (lane == 0) {
      float d = -INFINITY, mprev = -INFINITY;
      for (int k = 1; k <= M; k++) {
        const float dk = lsum(S.tbl, mprev + S.md[k - 1], d + S.dd[k - 1]);
        sD[k] = dk;
        const float mk = sM[k];
        e = k < M ? lsum(S.tbl, lsum(S.tbl, mk + esc, dk + esc), e) : lsum(S.tbl, lsum(S.tbl, mk, dk), e);
        mprev = mk;
        d = dk;
      }
    }
    wave_lds_sync();
    xE = __shfl(e, 0);
    xJ = lsum(S.tbl, xJ + xs[XS_JL], xE + xs[XS_EL]);
    xC = lsum(S.tbl, xC + xs[XS_CL], xE + xs[XS_EM]);
    xN = xN + xs[XS_NL];
    xB = lsum(S.tbl, xN + xs[XS_NM], xJ + xs[XS_JM]);
    if (lane == 0) {
      IXG_GLOBAL float *r = xm + (size_t)i * 5;
      r[0] = xE;
      r[1] = xN;
      r[2] = xJ;
      r[3] = xB;
      r[4] = xC;
    }
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
    wave_lds_sync();  // sM/sD of this row are read before the next row rewrites them
  }
  if (lane == 0) sc[it] = xC + xs[XS_CM];
}

template <int KPL>
__global__ __launch_bounds__(64 * GFB_WAVES) void gbwd_kernel(GmDev g, const uint8_t *seq, const GfbItem *items,
                                                              int32_t n, float *xout, float *sc) {
  extern __shared__ float smem[];
  const GfbShared S = gfb_stage(g, smem);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t it = (int64_t)blockIdx.x * GFB_WAVES + wv;
  if (it >= n) return;
  const int M = g.m, W = M + 2;
  float *sT = S.wave + (size_t)wv * 3 * W, *sP1 = sT + W, *sP2 = sP1 + W;
  const GfbItem item = items[it];
  const int L = item.len;
  const bool local_end = item.flags & 1;
  const float esc = local_end ? 0.0f : -INFINITY;
  float xs[8];
#pragma unroll
  for (int t = 0; t < 8; t++) xs[t] = item.xsc[t];
  const uint8_t *dsq = seq + item.off;
  IXG_GLOBAL float *xm = gmem(xout) + item.xoff;
  const int kb = 1 + lane * KPL;
  // row L (p7_generic.rs:946-962); sP1 carries M and sP2 D across the chain
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
  for (int i = L - 1; i >= 1; i--) {
    const int x = dsq[i + 1];
    const float *msx = g.msc + (size_t)x * (M + 1), *isx = g.isc + (size_t)x * (M + 1);
    // B terms (k = 1..M) and the M/I/D pieces that depend only on row i+1
    float mn1 = __shfl_down(mn[0], 1);  // M(i+1, k+1) for the last position of this lane
    if (lane == 63) mn1 = -INFINITY;
    float ic[KPL];
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      ic[j] = -INFINITY;
      if (k <= M) {
        sT[k] = mn[j] + S.bm[k - 1] + msx[k];
        if (k < M) {
          const float mk1 = (j + 1 < KPL) ? mn[j + 1] : mn1;
          const float *t = g.tsc + (size_t)k * 8;
          const float ms1 = mk1;
          sP1[k] = lsum(S.tbl, ms1 + t[T_MM] + msx[k + 1], in_[j] + t[T_MI] + isx[k]);
          ic[j] = lsum(S.tbl, ms1 + t[T_IM] + msx[k + 1], in_[j] + t[T_II] + isx[k]);
          sP2[k] = ms1 + t[T_DM] + msx[k + 1];
        }
      }
    }
    wave_lds_sync();
    float b = 0.f;
    if (lane == 0) {
      b = sT[1];
      for (int k = 2; k <= M; k++) b = lsum(S.tbl, b, sT[k]);
    }
    xB = __shfl(b, 0);
    xJ = lsum(S.tbl, xJ + xs[XS_JL], xB + xs[XS_JM]);
    xC = xC + xs[XS_CL];
    xE = lsum(S.tbl, xJ + xs[XS_EL], xC + xs[XS_EM]);
    xN = lsum(S.tbl, xN + xs[XS_NL], xB + xs[XS_NM]);
    if (lane == 0) {
      IXG_GLOBAL float *r = xm + (size_t)i * 5;
      r[0] = xE;
      r[1] = xN;
      r[2] = xJ;
      r[3] = xB;
      r[4] = xC;
      // M and D chain from k = M-1 down (p7_generic.rs:1012-1042); results
      // overwrite sP1 (M) and sP2 (D) at the positions already consumed
      float dn = xE;
      sP1[M] = xE;
      sP2[M] = xE;
      for (int k = M - 1; k >= 1; k--) {
        const float mk = lsum(S.tbl, sP1[k], lsum(S.tbl, xE + esc, dn + S.md[k]));
        const float dk = lsum(S.tbl, sP2[k], lsum(S.tbl, dn + S.dd[k], xE + esc));
        sP1[k] = mk;
        sP2[k] = dk;
        dn = dk;
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      mn[j] = k <= M ? sP1[k] : -INFINITY;
      in_[j] = ic[j];
    }
    wave_lds_sync();
  }
  // row 0 (p7_generic.rs:1044-1066)
  {
    const int x = dsq[1];
    const float *msx = g.msc + (size_t)x * (M + 1);
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int k = kb + j;
      if (k <= M) sT[k] = mn[j] + S.bm[k - 1] + msx[k];
    }
    wave_lds_sync();
    float b = 0.f;
    if (lane == 0) {
      b = sT[1];
      for (int k = 2; k <= M; k++) b = lsum(S.tbl, b, sT[k]);
    }
    xB = __shfl(b, 0);
    const float n0 = lsum(S.tbl, xN + xs[XS_NL], xB + xs[XS_NM]);
    if (lane == 0) {
      xm[0] = -INFINITY;
      xm[1] = n0;
      xm[2] = -INFINITY;
      xm[3] = xB;
      xm[4] = -INFINITY;
      sc[it] = n0;
    }
  }
}

template <int KPL>
hipError_t gfb_launch(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout,
                      float *mout, float *sc, bool backward, hipStream_t s) {
  const size_t lds = ((size_t)GFB_TBL + 3 * (size_t)(g.m + 1) + (size_t)GFB_WAVES * 3 * (g.m + 2)) * sizeof(float);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((n + GFB_WAVES - 1) / GFB_WAVES);
  if (backward)
    hipLaunchKernelGGL(gbwd_kernel<KPL>, dim3(grid), dim3(64 * GFB_WAVES), lds, s, g, seq, items, (int32_t)n, xout,
                       sc);
  else
    hipLaunchKernelGGL(gfwd_kernel<KPL>, dim3(grid), dim3(64 * GFB_WAVES), lds, s, g, seq, items, (int32_t)n, xout,
                       mout, sc);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gfb(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout, float *mout,
                      float *sc, bool backward, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (g.m < 2 || n > INT32_MAX) return hipErrorInvalidValue;
  const int kpl = (g.m + 63) / 64;
  if (kpl <= 1) return gfb_launch<1>(g, seq, items, n, xout, mout, sc, backward, s);
  if (kpl <= 2) return gfb_launch<2>(g, seq, items, n, xout, mout, sc, backward, s);
  if (kpl <= 4) return gfb_launch<4>(g, seq, items, n, xout, mout, sc, backward, s);
  if (kpl <= 8) return gfb_launch<8>(g, seq, items, n, xout, mout, sc, backward, s);
  if (kpl <= 16) return gfb_launch<16>(g, seq, items, n, xout, mout, sc, backward, s);
  return hipErrorInvalidValue;
}

}  // namespace ixg

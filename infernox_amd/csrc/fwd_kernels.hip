// F3 Forward, striped order (cm_pipeline.rs:1268-1418). Four lanes emulate
// the reference's four SSE lanes for one window: lane z owns model positions
// q+1+z*nq and runs the q loop in the same order, so every add and multiply
// happens in the same sequence as on the CPU. A wave serves 16 windows.
//
// fwd_exact<NQ>: the stripe count is a compile-time constant, the M and D
// columns live in registers and the I column in LDS; the profile is staged
// in LDS with the transitions re-laid out per (q, z) so a lane reads its
// seven values with two 16-byte loads. Instantiated for every stripe count
// 2..FWD_NQ_EXACT_MAX. fwd_generic: the same recurrence for any stripe count
// with all three columns in LDS.
#include "ixg_internal.h"

namespace ixg {

#define DPP_QUAD_RSZ 0x90   // quad_perm [0,0,1,2]
#define DPP_QUAD_X1 0xB1    // quad_perm [1,0,3,2]
#define DPP_QUAD_X2 0x4E    // quad_perm [2,3,0,1]
constexpr int FWD_NQ_EXACT_MAX = 100;

template <int CTRL>
__device__ __forceinline__ float quad_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ int quad_any(int v) {
  int a = v | __builtin_amdgcn_update_dpp(0, v, DPP_QUAD_X1, 0xF, 0xF, false);
  return a | __builtin_amdgcn_update_dpp(0, a, DPP_QUAD_X2, 0xF, 0xF, false);
}
__device__ __forceinline__ float quad_rsz(float v, int z) {
  const float r = quad_dpp<DPP_QUAD_RSZ>(v);
  return z ? r : 0.f;
}

template <int NQ>
__global__ __launch_bounds__(64) void fwd_exact(FwdDev p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                                                float *sc) {
  extern __shared__ float4 lds4[];
  float4 *tr4 = lds4;                    // [NQ][4 z][2] : BM MM IM DM | MD MI II DD
  float *rf = (float *)(lds4 + NQ * 8);  // [18][NQ][4]
  float *imc = rf + KP * NQ * 4;         // [NQ][64] insert column
  for (int k = threadIdx.x; k < NQ * 4; k += 64) {
    const int q = k >> 2, z = k & 3;
    const float *t = p.tmain + q * 28 + z;
    tr4[k * 2] = make_float4(t[0], t[4], t[8], t[12]);
    tr4[k * 2 + 1] = make_float4(t[16], t[20], t[24], p.tddv[q * 4 + z]);
  }
  for (int k = threadIdx.x; k < KP * NQ * 4; k += 64) rf[k] = p.rfvv[k];
  const int lane = threadIdx.x, z = lane & 3;
  for (int q = 0; q < NQ; q++) imc[q * 64 + lane] = 0.f;
  __syncthreads();
  const int64_t widx = (int64_t)blockIdx.x * 16 + (lane >> 2);
  WinDev w;
  w.off = 0;
  w.len = 0;
  w.idx = 0;
  w.a = 0.f;
  w.b = 0.f;
  if (widx < nwin) w = wins[widx];
  float mm[NQ], dm[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) mm[q] = dm[q] = 0.f;
  int maxL = w.len;
  for (int off = 32; off >= 1; off >>= 1) maxL = max(maxL, __shfl_xor(maxL, off));
  const float L = (float)w.len;
  const float pmove = (2.0f + 1.0f) / (L + 2.0f + 1.0f);
  const float ploop = 1.0f - pmove;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  const bool lazy = p.m >= 100;
  int xnext = (w.len >= 1) ? (int)seq[w.off + 1] : 0;
  for (int i = 1; i <= maxL; i++) {
    const bool act = i <= w.len;
    const int x = act ? xnext : 0;
    if (i + 1 <= w.len) xnext = seq[w.off + i + 1];
    const float *rp = rf + x * (NQ * 4) + z;
    float mpv = quad_rsz(mm[NQ - 1], z);
    float dpv = quad_rsz(dm[NQ - 1], z);
    float ipv = quad_rsz(imc[(NQ - 1) * 64 + lane], z);
    float dcv = 0.f, xev = 0.f;
    const float xbv = xb;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const float4 ta = tr4[(q * 4 + z) * 2], tb = tr4[(q * 4 + z) * 2 + 1];
      float sv = xbv * ta.x;
      sv = sv + mpv * ta.y;
      sv = sv + ipv * ta.z;
      sv = sv + dpv * ta.w;
      sv = sv * rp[q * 4];
      xev = xev + sv;
      mpv = mm[q];
      dpv = dm[q];
      ipv = imc[q * 64 + lane];
      mm[q] = sv;
      dm[q] = dcv;
      dcv = sv * tb.x;
      const float svi = mpv * tb.y;
      imc[q * 64 + lane] = svi + ipv * tb.z;
    }
    // DD paths: first obligatory pass
    dcv = quad_rsz(dcv, z);
    dm[0] = 0.f;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const float d = dcv + dm[q];
      dm[q] = d;
      dcv = d * tr4[(q * 4 + z) * 2 + 1].w;
    }
    bool done = false;
    for (int j = 1; j < 4; j++) {
      // `done` is uniform within a quad, so the shuffles only mix one window
      dcv = quad_rsz(dcv, z);
      int cv = 0;
      if (!done) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const float old = dm[q];
          const float d = dcv + old;
          cv |= d > old;
          dm[q] = d;
          dcv = dcv * tr4[(q * 4 + z) * 2 + 1].w;
        }
      }
      if (lazy && !quad_any(cv)) done = true;
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) xev = xev + dm[q];
    const float p01 = xev + quad_dpp<DPP_QUAD_X1>(xev);
    const float xe = p01 + quad_dpp<DPP_QUAD_X2>(p01);
    if (act) {
      xn = xn * ploop;
      xc = (xc * ploop) + (xe * 0.5f);
      xj = (xj * ploop) + (xe * 0.5f);
      xb = (xj * pmove) + (xn * pmove);
      if (xe > 1.0e4f) {
        const float inv = 1.0f / xe;
        xn *= inv;
        xc *= inv;
        xj *= inv;
        xb *= inv;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          mm[q] *= inv;
          dm[q] *= inv;
          imc[q * 64 + lane] *= inv;
        }
        totscale += (float)log((double)xe);
      }
    }
  }
  if (widx < nwin && z == 0) sc[w.idx] = (float)((double)totscale + log((double)xc * (double)pmove));
}

__global__ __launch_bounds__(64) void fwd_generic(FwdDev p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                                                  float *sc) {
  extern __shared__ float col[];  // [3][nq][64]
  const int lane = threadIdx.x, z = lane & 3;
  const int64_t widx = (int64_t)blockIdx.x * 16 + (lane >> 2);
  WinDev w;
  w.off = 0;
  w.len = 0;
  w.idx = 0;
  w.a = 0.f;
  w.b = 0.f;
  if (widx < nwin) w = wins[widx];
  const int nq = p.nq;
  float *mm = col, *dm = col + nq * 64, *im = col + 2 * nq * 64;
  for (int q = 0; q < nq; q++) mm[q * 64 + lane] = dm[q * 64 + lane] = im[q * 64 + lane] = 0.f;
  int maxL = w.len;
  for (int off = 32; off >= 1; off >>= 1) maxL = max(maxL, __shfl_xor(maxL, off));
  const float L = (float)w.len;
  const float pmove = (2.0f + 1.0f) / (L + 2.0f + 1.0f);
  const float ploop = 1.0f - pmove;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  const float *tm = p.tmain;
  for (int i = 1; i <= maxL; i++) {
    const bool act = i <= w.len;
    const int x = act ? (int)seq[w.off + i] : 0;
    const float *rp = p.rfvv + (size_t)x * (nq * 4);
    float dcv = 0.f, xev = 0.f;
    const float xbv = xb;
    float mpv = quad_rsz(mm[(nq - 1) * 64 + lane], z);
    float dpv = quad_rsz(dm[(nq - 1) * 64 + lane], z);
    float ipv = quad_rsz(im[(nq - 1) * 64 + lane], z);
    for (int q = 0; q < nq; q++) {
      const float *tb = tm + q * 28 + z;
      float sv = xbv * tb[0];
      sv = sv + mpv * tb[4];
      sv = sv + ipv * tb[8];
      sv = sv + dpv * tb[12];
      sv = sv * rp[q * 4 + z];
      xev = xev + sv;
      mpv = mm[q * 64 + lane];
      dpv = dm[q * 64 + lane];
      ipv = im[q * 64 + lane];
      mm[q * 64 + lane] = sv;
      dm[q * 64 + lane] = dcv;
      dcv = sv * tb[16];
      const float svi = mpv * tb[20];
      im[q * 64 + lane] = svi + ipv * tb[24];
    }
    dcv = quad_rsz(dcv, z);
    dm[lane] = 0.f;
    for (int q = 0; q < nq; q++) {
      const float d = dcv + dm[q * 64 + lane];
      dm[q * 64 + lane] = d;
      dcv = d * p.tddv[q * 4 + z];
    }
    bool done = false;
    for (int j = 1; j < 4; j++) {
      dcv = quad_rsz(dcv, z);
      int cv = 0;
      if (!done) {
        for (int q = 0; q < nq; q++) {
          const float old = dm[q * 64 + lane];
          const float d = dcv + old;
          cv |= d > old;
          dm[q * 64 + lane] = d;
          dcv = dcv * p.tddv[q * 4 + z];
        }
      }
      if (p.m >= 100 && !quad_any(cv)) done = true;
    }
    for (int q = 0; q < nq; q++) xev = xev + dm[q * 64 + lane];
    const float p01 = xev + quad_dpp<DPP_QUAD_X1>(xev);
    const float xe = p01 + quad_dpp<DPP_QUAD_X2>(p01);
    if (act) {
      xn = xn * ploop;
      xc = (xc * ploop) + (xe * 0.5f);
      xj = (xj * ploop) + (xe * 0.5f);
      xb = (xj * pmove) + (xn * pmove);
      if (xe > 1.0e4f) {
        const float inv = 1.0f / xe;
        xn *= inv;
        xc *= inv;
        xj *= inv;
        xb *= inv;
        for (int q = 0; q < nq; q++) {
          mm[q * 64 + lane] *= inv;
          dm[q * 64 + lane] *= inv;
          im[q * 64 + lane] *= inv;
        }
        totscale += (float)log((double)xe);
      }
    }
  }
  if (widx < nwin && z == 0) sc[w.idx] = (float)((double)totscale + log((double)xc * (double)pmove));
}

// Full-matrix Forward (p7_fwdback.rs:301-532 p7_forward + forward_score):
// the fwd_generic recurrence with the length model of either the multihit or
// the unihit configuration, every row's M, I and D (after rescaling) stored
// in natural order k = 0..M at mat + moff[w] (M | I | D, (len+1)*(M+1)
// each), the specials E N J B C SCALE per row at xmx + xoff[w], and score
// and totscale at out[2*idx], out[2*idx+1]. A workgroup of 4 * wpw lanes
// holds wpw windows (4 lanes each): a batch of few windows is spread over
// more SIMDs instead of packed 16 to a wave.
__global__ __launch_bounds__(64) void fwd_full(FwdDev p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                                               const int64_t *moff, const int64_t *xoff, float *mat, float *xmx,
                                               float *out, int unihit) {
  // LDS: the transitions [nq][28], DD [nq][4] and emissions [KP][nq][4]
  // staged once, then the M, D, I columns [3][nq][S]
  extern __shared__ float lds[];
  const int lane = threadIdx.x, z = lane & 3, S = blockDim.x;
  const int nq = p.nq, M = p.m;
  float *tm = lds, *tdd = tm + nq * 28, *rf = tdd + nq * 4, *col = rf + KP * nq * 4;
  for (int k = lane; k < nq * 28; k += S) tm[k] = p.tmain[k];
  for (int k = lane; k < nq * 4; k += S) tdd[k] = p.tddv[k];
  for (int k = lane; k < KP * nq * 4; k += S) rf[k] = p.rfvv[k];
  const int64_t widx = (int64_t)blockIdx.x * (S >> 2) + (lane >> 2);
  WinDev w;
  w.off = 0;
  w.len = 0;
  w.idx = 0;
  w.a = 0.f;
  w.b = 0.f;
  if (widx < nwin) w = wins[widx];
  float *mm = col, *dm = col + nq * S, *im = col + 2 * nq * S;
  for (int q = 0; q < nq; q++) mm[q * S + lane] = dm[q * S + lane] = im[q * S + lane] = 0.f;
  int maxL = w.len;
  for (int off = S >> 1; off >= 1; off >>= 1) maxL = max(maxL, __shfl_xor(maxL, off));
  const float L = (float)w.len;
  const float nj = unihit ? 0.0f : 1.0f;
  const float pmove = (2.0f + nj) / (L + 2.0f + nj);  // p7_fwdback.rs:212-252
  const float ploop = 1.0f - pmove;
  const float emove = unihit ? 1.0f : 0.5f, eloop = unihit ? 0.0f : 0.5f;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  const size_t rowlen = (size_t)M + 1, plane = ((size_t)w.len + 1) * rowlen;
  float *om = mat + (widx < nwin ? moff[w.idx] : 0), *oi = om + plane, *od = om + 2 * plane;
  float *ox = xmx + (widx < nwin ? xoff[w.idx] : 0);
  if (widx < nwin) {  // row 0 (p7_fwdback.rs:314-322)
    for (int k = z; k <= M; k += 4) om[k] = oi[k] = od[k] = 0.f;
    if (z == 0) {
      ox[0] = 0.f;
      ox[1] = 1.f;
      ox[2] = 0.f;
      ox[3] = pmove;
      ox[4] = 0.f;
      ox[5] = 1.f;
    }
  }
  __syncthreads();
  for (int i = 1; i <= maxL; i++) {
    const bool act = i <= w.len;
    const int x = act ? (int)seq[w.off + i] : 0;
    const float *rp = rf + (size_t)x * (nq * 4);
    float dcv = 0.f, xev = 0.f;
    const float xbv = xb;
    float mpv = quad_rsz(mm[(nq - 1) * S + lane], z);
    float dpv = quad_rsz(dm[(nq - 1) * S + lane], z);
    float ipv = quad_rsz(im[(nq - 1) * S + lane], z);
    for (int q = 0; q < nq; q++) {
      const float *tb = tm + q * 28 + z;
      float sv = xbv * tb[0];
      sv = sv + mpv * tb[4];
      sv = sv + ipv * tb[8];
      sv = sv + dpv * tb[12];
      sv = sv * rp[q * 4 + z];
      xev = xev + sv;
      mpv = mm[q * S + lane];
      dpv = dm[q * S + lane];
      ipv = im[q * S + lane];
      mm[q * S + lane] = sv;
      dm[q * S + lane] = dcv;
      dcv = sv * tb[16];
      const float svi = mpv * tb[20];
      im[q * S + lane] = svi + ipv * tb[24];
    }
    dcv = quad_rsz(dcv, z);
    dm[lane] = 0.f;
    for (int q = 0; q < nq; q++) {
      const float d = dcv + dm[q * S + lane];
      dm[q * S + lane] = d;
      dcv = d * tdd[q * 4 + z];
    }
    bool done = false;
    for (int j = 1; j < 4; j++) {
      dcv = quad_rsz(dcv, z);
      int cv = 0;
      if (!done) {
        for (int q = 0; q < nq; q++) {
          const float old = dm[q * S + lane];
          const float d = dcv + old;
          cv |= d > old;
          dm[q * S + lane] = d;
          dcv = dcv * tdd[q * 4 + z];
        }
      }
      if (p.m >= 100 && !quad_any(cv)) done = true;
    }
    for (int q = 0; q < nq; q++) xev = xev + dm[q * S + lane];
    const float p01 = xev + quad_dpp<DPP_QUAD_X1>(xev);
    float xe = p01 + quad_dpp<DPP_QUAD_X2>(p01);
    if (act) {
      xn = xn * ploop;
      xc = (xc * ploop) + (xe * emove);
      xj = (xj * ploop) + (xe * eloop);
      xb = (xj * pmove) + (xn * pmove);
      float scale = 1.0f;
      if (xe > 1.0e4f) {
        const float inv = 1.0f / xe;
        xn *= inv;
        xc *= inv;
        xj *= inv;
        xb *= inv;
        for (int q = 0; q < nq; q++) {
          mm[q * S + lane] *= inv;
          dm[q * S + lane] *= inv;
          im[q * S + lane] *= inv;
        }
        totscale += (float)log((double)xe);
        scale = xe;
        xe = 1.0f;
      }
      const size_t r = (size_t)i * rowlen;
      if (z == 0) {
        om[r] = oi[r] = od[r] = 0.f;
        float *xr = ox + (size_t)i * 6;
        xr[0] = xe;
        xr[1] = xn;
        xr[2] = xj;
        xr[3] = xb;
        xr[4] = xc;
        xr[5] = scale;
      }
      for (int q = 0; q < nq; q++) {
        const int k = z * nq + q + 1;
        if (k <= M) {
          om[r + k] = mm[q * S + lane];
          oi[r + k] = im[q * S + lane];
          od[r + k] = dm[q * S + lane];
        }
      }
    }
  }
  if (widx < nwin && z == 0) {
    out[2 * w.idx] = (float)((double)totscale + log((double)xc * (double)pmove));
    out[2 * w.idx + 1] = totscale;
  }
}

hipError_t launch_fwd_full(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                           const int64_t *moff, const int64_t *xoff, float *mat, float *xmx, float *out, int unihit,
                           hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  // windows per workgroup: 16 once the batch gives every SIMD (1024) two
  // waves of 16, fewer (a power of two) below that
  int wpw = 16;
  while (wpw > 1 && nwin < (int64_t)wpw * 2048) wpw >>= 1;
  const size_t lds = ((size_t)p.nq * (28 + 4 + KP * 4) + (size_t)3 * p.nq * 4 * wpw) * sizeof(float);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fwd_full, dim3((unsigned)((nwin + wpw - 1) / wpw)), dim3(4 * wpw), lds, s, p, seq, wins, nwin,
                     moff, xoff, mat, xmx, out, unihit);
  return hipGetLastError();
}

template <int NQ>
static hipError_t fwd_exact_launch(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                                   hipStream_t s) {
  const size_t lds = (size_t)NQ * 8 * sizeof(float4) + (size_t)KP * NQ * 4 * sizeof(float) +
                     (size_t)NQ * 64 * sizeof(float);
  hipLaunchKernelGGL(fwd_exact<NQ>, dim3((unsigned)((nwin + 15) / 16)), dim3(64), lds, s, p, seq, wins, nwin, sc);
  return hipGetLastError();
}

typedef hipError_t (*fwd_fn)(const FwdDev &, const uint8_t *, const WinDev *, int64_t, float *, hipStream_t);
template <int... Q>
struct FwdTable {
  static constexpr fwd_fn fns[] = {fwd_exact_launch<Q>...};
};
template <int N, int... Q>
struct MakeFwdTable : MakeFwdTable<N - 1, N, Q...> {};
template <int... Q>
struct MakeFwdTable<1, Q...> {
  using type = FwdTable<Q...>;
};
using FwdAll = MakeFwdTable<FWD_NQ_EXACT_MAX>::type;  // stripe counts 2..FWD_NQ_EXACT_MAX

bool fwd_variant_ok(const FwdDev &p, int v) {
  if (v == FWD_V_ZW) return p.nq >= 2 && p.nq <= FWD_ZW_NQ_MAX && p.tz;
  if (v == FWD_V_EXACT) return p.nq >= 2 && p.nq <= FWD_NQ_EXACT_MAX;
  return v == FWD_V_GENERIC && (size_t)3 * p.nq * 64 * sizeof(float) <= 160 * 1024;
}

hipError_t launch_fwd_variant(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                              int v, hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  if (!fwd_variant_ok(p, v)) return hipErrorInvalidValue;
  if (v == FWD_V_ZW) return launch_fwd_zw(p, seq, wins, nwin, sc, s);
  if (v == FWD_V_EXACT) return FwdAll::fns[p.nq - 2](p, seq, wins, nwin, sc, s);
  const size_t lds = (size_t)3 * p.nq * 64 * sizeof(float);
  hipLaunchKernelGGL(fwd_generic, dim3((unsigned)((nwin + 15) / 16)), dim3(64), lds, s, p, seq, wins, nwin, sc);
  return hipGetLastError();
}

hipError_t launch_fwd(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                      hipStream_t s) {
  for (int v : {FWD_V_ZW, FWD_V_EXACT, FWD_V_GENERIC})
    if (fwd_variant_ok(p, v)) return launch_fwd_variant(p, seq, wins, nwin, sc, v, s);
  return hipErrorInvalidValue;
}

}  // namespace ixg

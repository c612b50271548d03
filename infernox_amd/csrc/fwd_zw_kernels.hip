// F3 Forward, striped order (cm_pipeline.rs:1268-1418), stripe-per-wave
// form. The reference's four SSE lanes z = 0..3 (lane z owns model positions
// q+1+z*nq) become the four waves of a 256-thread workgroup and the 64 lanes
// of a wave are 64 windows, so
//   * a wave's transitions are the same for all its lanes and are read with
//     scalar loads (no per-lane profile traffic); only the emission odds,
//     which depend on each window's residue, are read from LDS;
//   * everything the reference does inside one SSE lane (the q loop of M/I,
//     the DD passes, the xE partial sums) stays in one lane, in the same
//     order, on values held in registers;
//   * the steps that cross SSE lanes go through LDS between the four waves:
//     the end-of-row shift (lane z-1's last stripe feeds lane z's first),
//     the carried D value of every DD pass, the lazy-pass "changed" flags,
//     and the horizontal xE sum (x0 + x1) + (x2 + x3).
// Every add and multiply of the reference happens in the same sequence, so
// the score is bit-identical to the four-lanes-per-window kernels.
#include "ixg_internal.h"

namespace ixg {

// Constant address space: wave-uniform reads through such a pointer are
// scalar loads.
typedef __attribute__((address_space(4))) const float CFloat;

// The DD passes read their transitions through a pointer the compiler
// cannot relate to the M/I loop's, so it does not keep the M/I loop's copies
// of them alive (in spilled scalar registers) until the DD passes.
__device__ __forceinline__ const CFloat *dd_table(const float *p) {
  uint64_t a = (uint64_t)p;
  asm volatile("" : "+s"(a));
  return (const CFloat *)a;
}

template <int NQ>
__global__ __launch_bounds__(256) void fwd_zw(FwdDev p, const float *__restrict__ tz,
                                              const uint8_t *__restrict__ seq, const WinDev *__restrict__ wins,
                                              int64_t nwin, float *sc) {
  constexpr int NQP = (NQ + 3) & ~3;  // emission row stride (16-byte rows)
  extern __shared__ float4 sh4[];
  float *rf = (float *)sh4;            // [4 z][KP][NQP] emission odds
  float *x1 = rf + 4 * KP * NQP;       // [4][64] D carry after the M/I loop
  float *xl = x1 + 256;                // [3][4][64] D carry before lazy pass j
  int *cvx = (int *)(xl + 768);        // [2][4][64] "changed" flag of lazy pass j
  float4 *xx = (float4 *)(cvx + 512);  // [4][64] xE partial and the row's last stripe (M, D, I)
  const int tid = threadIdx.x, lane = tid & 63;
  const int z = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int k = tid; k < 4 * KP * NQP; k += 256) {
    const int zz = k / (KP * NQP), r = k - zz * KP * NQP, x = r / NQP, q = r - x * NQP;
    rf[k] = q < NQ ? p.rfvv[(size_t)x * (NQ * 4) + q * 4 + zz] : 0.f;
  }
  const int64_t widx = (int64_t)blockIdx.x * 64 + lane;
  WinDev w;
  w.off = 0;
  w.len = 0;
  w.idx = 0;
  w.a = 0.f;
  w.b = 0.f;
  if (widx < nwin) w = wins[widx];
  int maxL = w.len;
  for (int off = 32; off >= 1; off >>= 1) maxL = max(maxL, __shfl_xor(maxL, off));
  __syncthreads();
  constexpr int NQ8 = (NQ + 7) & ~7;
  const float *tz_z = tz + (size_t)z * NQ * 8;               // BM MM IM DM MD MI II DD per stripe position
  const float *tdd_z = tz + (size_t)4 * NQ * 8 + z * NQ8;    // DD alone, zero-padded to NQ8
  const float *rz = rf + z * KP * NQP;
  float mm[NQ], dm[NQ], im[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) mm[q] = dm[q] = im[q] = 0.f;
  const float L = (float)w.len;
  const float pmove = (2.0f + 1.0f) / (L + 2.0f + 1.0f);
  const float ploop = 1.0f - pmove;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  float mprev = 0.f, dprev = 0.f, iprev = 0.f;  // stripe z-1's last position, previous row
  const bool lazy = p.m >= 100;
  int xnext = (w.len >= 1) ? (int)seq[w.off + 1] : 0;
  for (int i = 1; i <= maxL; i++) {
    // The transition loads are loop-invariant; left to itself the compiler
    // hoists all 8*NQ of them out of the row loop and spills them. An empty
    // asm that "modifies" the table address keeps them inside the loop.
    uint64_t ta = (uint64_t)tz_z;
    asm volatile("" : "+s"(ta));
    const CFloat *tq = (const CFloat *)ta;
    const bool act = i <= w.len;
    const int x = act ? xnext : 0;
    if (i + 1 <= w.len) xnext = seq[w.off + i + 1];
    const float4 *rp4 = (const float4 *)(rz + x * NQP);
    const float xbv = xb;
    // M and I in descending q, in place: new M(q) reads the old M, I, D of
    // q-1 (not yet overwritten), new I(q) the old M(q) and I(q), and the new
    // D(q+1) = M(q)*MD(q) replaces a D(q+1) already consumed by M(q+2). Each
    // value is the reference's expression evaluated in its order; only the
    // order in which independent positions are visited differs. Transitions
    // are fetched two positions ahead and emissions one group of four ahead;
    // scheduling barriers between pairs of steps keep the steps in order so
    // that only the three columns and the prefetched values are live.
    float dcv = 0.f;  // D carry out of this stripe: M(NQ-1)*MD(NQ-1)
    float t0[8], t1[8], t2[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      t0[k] = tq[(NQ - 1) * 8 + k];
      t1[k] = NQ > 1 ? tq[(NQ - 2) * 8 + k] : 0.f;
    }
    float4 e4 = rp4[(NQ - 1) / 4], e4n = e4;
#pragma unroll
    for (int q = NQ - 1; q >= 0; q--) {
      if (q - 2 >= 0) {
#pragma unroll
        for (int k = 0; k < 8; k++) t2[k] = tq[(q - 2) * 8 + k];
      }
      if ((q & 3) == 3 || q == NQ - 1) {
        if (q / 4 > 0) e4n = rp4[q / 4 - 1];
      }
      const float ev = (q & 3) == 0 ? e4.x : (q & 3) == 1 ? e4.y : (q & 3) == 2 ? e4.z : e4.w;
      const float mo = q ? mm[q - 1] : mprev, io = q ? im[q - 1] : iprev, dold = q ? dm[q - 1] : dprev;
      const float svi = mm[q] * t0[5];
      im[q] = svi + im[q] * t0[6];
      float sv = xbv * t0[0];
      sv = sv + mo * t0[1];
      sv = sv + io * t0[2];
      sv = sv + dold * t0[3];
      sv = sv * ev;
      mm[q] = sv;
      if (q + 1 < NQ) {
        dm[q + 1] = sv * t0[4];
        asm volatile("" : "+v"(mm[q]), "+v"(im[q]), "+v"(dm[q + 1]));  // computed here, not sunk
      } else {
        dcv = sv * t0[4];
        asm volatile("" : "+v"(mm[q]), "+v"(im[q]), "+v"(dcv));
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        t0[k] = t1[k];
        t1[k] = t2[k];
      }
      if ((q & 3) == 0) e4 = e4n;
      if ((q & 1) == 0) __builtin_amdgcn_sched_barrier(0);  // pairs of steps may interleave
    }
    dm[0] = 0.f;
    // Lane z's xE partial is M over q ascending, then D over q ascending with
    // the final D values. The M sum runs inside DD pass 1 and the D sum
    // inside every lazy pass (the last pass that runs leaves the final
    // values), each as a second dependency chain beside the D carry.
    float xev_m = 0.f, xev = 0.f;
    // DD paths: first obligatory pass on the carry shifted in from stripe z-1
    x1[tid] = dcv;
    __syncthreads();
    dcv = z ? x1[tid - 64] : 0.f;
    {
      const CFloat *td = dd_table(tdd_z);
      float dd0[8], dd1[8];
#pragma unroll
      for (int k = 0; k < 8; k++) dd0[k] = td[k];
#pragma unroll
      for (int c = 0; c < NQ8 / 8; c++) {
        if (c + 1 < NQ8 / 8) {
#pragma unroll
          for (int k = 0; k < 8; k++) dd1[k] = td[(c + 1) * 8 + k];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int q = c * 8 + u;
          if (q < NQ) {
            xev_m = xev_m + mm[q];
            const float d = dcv + dm[q];
            dm[q] = d;
            dcv = d * dd0[u];
          }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) dd0[k] = dd1[k];
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    bool done = false;
    int cv = 0;
#pragma unroll 1
    for (int j = 1; j < 4; j++) {
      xl[(j - 1) * 256 + tid] = dcv;
      if (j > 1) cvx[(j - 2) * 256 + tid] = cv;
      __syncthreads();
      dcv = z ? xl[(j - 1) * 256 + tid - 64] : 0.f;
      if (j > 1 && lazy) {
        const int *c = cvx + (j - 2) * 256 + lane;
        if (!(c[0] | c[64] | c[128] | c[192])) done = true;
      }
      // all four waves hold the same flags for the same 64 windows, so this
      // exit is uniform across the workgroup
      if (__all(done)) break;
      // A finished window runs the pass with a zero carry, which leaves its
      // D column unchanged (0 + d == d for the non-negative D values) and
      // keeps the pass free of divergent branches.
      if (done) dcv = 0.f;
      cv = 0;
      xev = xev_m;
      {
        const CFloat *td = dd_table(tdd_z);
        float dd0[8], dd1[8];
#pragma unroll
        for (int k = 0; k < 8; k++) dd0[k] = td[k];
#pragma unroll
        for (int c = 0; c < NQ8 / 8; c++) {
          if (c + 1 < NQ8 / 8) {
#pragma unroll
            for (int k = 0; k < 8; k++) dd1[k] = td[(c + 1) * 8 + k];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int q = c * 8 + u;
            if (q < NQ) {
              const float old = dm[q];
              const float d = dcv + old;
              cv |= d > old;
              dm[q] = d;
              xev = xev + d;
              dcv = dcv * dd0[u];
            }
          }
#pragma unroll
          for (int k = 0; k < 8; k++) dd0[k] = dd1[k];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    xx[tid] = make_float4(xev, mm[NQ - 1], dm[NQ - 1], im[NQ - 1]);
    __syncthreads();
    const float xe = (xx[lane].x + xx[64 + lane].x) + (xx[128 + lane].x + xx[192 + lane].x);
    if (z) {
      const float4 pz = xx[tid - 64];
      mprev = pz.y;
      dprev = pz.z;
      iprev = pz.w;
    }
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
          im[q] *= inv;
        }
        mprev *= inv;
        dprev *= inv;
        iprev *= inv;
        totscale += (float)log((double)xe);
      }
    }
  }
  if (widx < nwin && z == 0) sc[w.idx] = (float)((double)totscale + log((double)xc * (double)pmove));
}

template <int NQ>
static hipError_t fwd_zw_launch(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                                hipStream_t s) {
  constexpr int NQP = (NQ + 3) & ~3;
  const size_t lds = (size_t)(4 * KP * NQP + 256 + 768 + 512) * 4 + 256 * sizeof(float4);
  hipLaunchKernelGGL(fwd_zw<NQ>, dim3((unsigned)((nwin + 63) / 64)), dim3(256), lds, s, p, p.tz, seq, wins, nwin,
                     sc);
  return hipGetLastError();
}

typedef hipError_t (*fwd_zw_fn)(const FwdDev &, const uint8_t *, const WinDev *, int64_t, float *, hipStream_t);
template <int... Q>
struct FwdZwTable {
  static constexpr fwd_zw_fn fns[] = {fwd_zw_launch<Q>...};
};
template <int N, int... Q>
struct MakeFwdZwTable : MakeFwdZwTable<N - 1, N, Q...> {};
template <int... Q>
struct MakeFwdZwTable<1, Q...> {
  using type = FwdZwTable<Q...>;
};
using FwdZwAll = MakeFwdZwTable<FWD_ZW_NQ_MAX>::type;  // stripe counts 2..FWD_ZW_NQ_MAX

hipError_t launch_fwd_zw(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                         hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  if (p.nq < 2 || p.nq > FWD_ZW_NQ_MAX || !p.tz) return hipErrorInvalidValue;
  return FwdZwAll::fns[p.nq - 2](p, seq, wins, nwin, sc, s);
}

}  // namespace ixg

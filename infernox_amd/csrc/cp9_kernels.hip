// gfx950 CP9 HMM banding: the integer Forward and Backward of the band
// producing HMM over a window and its posterior HMM bands
// (cp9.rs:1275-1453 cp9_forward, :1995-2208 cp9_backward, :1517-1737
// cp9_fb2hmmbands). One lane runs one window's whole recurrence in the
// reference's order, so every ILogsum is the reference's; 64 windows of
// similar length share a wave. Matrices are interleaved over the wave's
// lanes ([row][k][lane]), so each load and store of a wave is one coalesced
// 256-byte access; transitions and the local-end lists are the same for all
// lanes (scalar loads); the ILogsum table sits in LDS as 16-bit entries. The
// local-end (EL) rows are only read one row back and roll over two rows; the
// posterior pass recomputes each posterior cell in both of its scans
// instead of keeping the posterior matrices.
#include "cm_internal.h"

namespace ixg {

namespace {

constexpr int32_t NINF = -INFTY_I;
enum { O_MM = 0, O_IM, O_DM, O_BM, O_MI, O_II, O_DI, O_MD, O_ID, O_DD, O_ME, O_MEL, O_NT };

__device__ __forceinline__ int32_t wa(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t ws(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

// ILogsum (cp9.rs:62-74)
__device__ __forceinline__ int32_t lsum(int32_t a, int32_t b, const uint16_t *tbl) {
  const int32_t mx = max(NINF, max(a, b)), mn = min(a, b);
  const int32_t diff = mx - mn;
  const bool use = mn > NINF && diff < FLOGSUM_TBL;
  const int32_t t = tbl[use ? diff : 0];
  return use ? mx + t : mx;
}

__device__ __forceinline__ float scorify(int32_t sc) { return (float)sc / 1000.0f; }  // cp9.rs:1257

struct Lane {  // one lane's view of its group's interleaved matrices
  int M1, l;
  __device__ __forceinline__ size_t at(int ip, int k) const { return ((size_t)ip * M1 + k) * 64 + l; }
  __device__ __forceinline__ size_t at1(int k) const { return (size_t)k * 64 + l; }
};

__device__ const uint16_t *load_table(const int32_t *ilog, uint16_t *s) {
  for (int x = threadIdx.x; x < FLOGSUM_TBL; x += blockDim.x) s[x] = (uint16_t)ilog[x];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(64) void cp9_fwd_kernel(Cp9Dev h, const Cp9WinDev *wins, int64_t n,
                                                    const Cp9GrpDev *grps, int do_scan, float *best_out) {
  __shared__ uint16_t s_tbl[FLOGSUM_TBL];
  const uint16_t *tbl = load_table(h.ilog, s_tbl);
  const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (e >= n) return;
  const Cp9WinDev w = wins[e];
  const Cp9GrpDev g = grps[blockIdx.x];
  const int M = h.M, i0 = w.i0, j0 = w.j0;
  const Lane a{M + 1, (int)threadIdx.x};
  int32_t *mmx = g.fm, *imx = g.fi, *dmx = g.fd, *el = g.fel;
  const int32_t *T = h.otsc;
  const uint8_t *dsq = w.dsq;
#define TS(k, x) T[(size_t)(k) * O_NT + (x)]
  float best = -1e36f;
  mmx[a.at(0, 0)] = 0;
  imx[a.at(0, 0)] = NINF;
  dmx[a.at(0, 0)] = NINF;
  el[a.at1(0)] = NINF;
  for (int k = 1; k <= M; k++) {
    mmx[a.at(0, k)] = NINF;
    imx[a.at(0, k)] = NINF;
    el[a.at1(k)] = NINF;
    dmx[a.at(0, k)] = lsum(lsum(wa(mmx[a.at(0, k - 1)], TS(k - 1, O_MD)), wa(imx[a.at(0, k - 1)], TS(k - 1, O_ID)), tbl),
                           wa(dmx[a.at(0, k - 1)], TS(k - 1, O_DD)), tbl);
  }
  int32_t sca = wa(dmx[a.at(0, M)], TS(M, O_DM));
  if (scorify(sca) > best) best = scorify(sca);
  for (int j = i0; j <= j0; j++) {
    const int dj = dsq[j], cur = j - i0 + 1, prv = cur - 1;
    int32_t *elc = el + (size_t)(cur & 1) * (M + 1) * 64, *elp = el + (size_t)(prv & 1) * (M + 1) * 64;
    const int32_t *ms = h.msc, *is = h.isc;
    mmx[a.at(cur, 0)] = do_scan ? 0 : NINF;
    dmx[a.at(cur, 0)] = NINF;
    elc[a.at1(0)] = NINF;
    const int32_t m0p = mmx[a.at(prv, 0)];
    int32_t sc = lsum(lsum(wa(m0p, TS(0, O_MI)), wa(imx[a.at(prv, 0)], TS(0, O_II)), tbl),
                      wa(dmx[a.at(prv, 0)], TS(0, O_DI)), tbl);
    imx[a.at(cur, 0)] = max(wa(sc, is[dj]), NINF);
    int32_t endsc = NINF;
    int32_t mprv_k1 = m0p, iprv_k1 = imx[a.at(prv, 0)], dprv_k1 = dmx[a.at(prv, 0)];  // row prv at k-1
    int32_t mcur_k1 = mmx[a.at(cur, 0)], icur_k1 = imx[a.at(cur, 0)], dcur_k1 = NINF;  // row cur at k-1
    for (int k = 1; k <= M; k++) {
      const int32_t mp = mmx[a.at(prv, k)], ip = imx[a.at(prv, k)], dp = dmx[a.at(prv, k)];
      sc = lsum(lsum(wa(mprv_k1, TS(k - 1, O_MM)), wa(iprv_k1, TS(k - 1, O_IM)), tbl),
                lsum(wa(dprv_k1, TS(k - 1, O_DM)), wa(m0p, TS(k, O_BM)), tbl), tbl);
      const int nsrc = h.el_ct[k], o = h.el_off[k];
      for (int c = 0; c < nsrc; c++) sc = lsum(sc, elp[a.at1(h.el_idx[o + c])], tbl);
      const int32_t mc = max(wa(sc, ms[(size_t)k * KP + dj]), NINF);
      mmx[a.at(cur, k)] = mc;
      endsc = lsum(endsc, wa(mc, TS(k, O_ME)), tbl);
      sc = lsum(lsum(wa(mp, TS(k, O_MI)), wa(ip, TS(k, O_II)), tbl), wa(dp, TS(k, O_DI)), tbl);
      const int32_t ic = max(wa(sc, is[(size_t)k * KP + dj]), NINF);
      imx[a.at(cur, k)] = ic;
      const int32_t dc = lsum(lsum(wa(mcur_k1, TS(k - 1, O_MD)), wa(icur_k1, TS(k - 1, O_ID)), tbl),
                              wa(dcur_k1, TS(k - 1, O_DD)), tbl);
      dmx[a.at(cur, k)] = dc;
      int32_t ec = NINF;
      if (h.el && h.has_el[k]) ec = lsum(wa(mc, TS(k, O_MEL)), wa(elp[a.at1(k)], h.el_selfsc), tbl);
      elc[a.at1(k)] = ec;
      mprv_k1 = mp;
      iprv_k1 = ip;
      dprv_k1 = dp;
      mcur_k1 = mc;
      icur_k1 = ic;
      dcur_k1 = dc;
    }
    endsc = lsum(lsum(endsc, wa(dcur_k1, TS(M, O_DM)), tbl), wa(icur_k1, TS(M, O_IM)), tbl);
    const int nsrc = h.el_ct[M + 1], o = h.el_off[M + 1];
    for (int c = 0; c < nsrc; c++) endsc = lsum(endsc, elc[a.at1(h.el_idx[o + c])], tbl);
    sca = endsc;
    if (scorify(endsc) > best) best = scorify(endsc);
  }
  if (!do_scan) best = scorify(sca);  // doing_align: the score of the whole window
  best_out[w.out] = best;
#undef TS
}

__global__ __launch_bounds__(64) void cp9_bwd_kernel(Cp9Dev h, const Cp9WinDev *wins, int64_t n,
                                                    const Cp9GrpDev *grps, int do_scan, float *best_out) {
  __shared__ uint16_t s_tbl[FLOGSUM_TBL];
  const uint16_t *tbl = load_table(h.ilog, s_tbl);
  const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (e >= n) return;
  const Cp9WinDev w = wins[e];
  const Cp9GrpDev g = grps[blockIdx.x];
  const int M = h.M, i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const Lane a{M + 1, (int)threadIdx.x};
  int32_t *mmx = g.bm, *imx = g.bi, *dmx = g.bd, *el = g.bel;
  const int32_t *T = h.otsc, *ms = h.msc, *is = h.isc;
  const uint8_t *dsq = w.dsq;
  const bool elon = h.el != 0;
#define TS(k, x) T[(size_t)(k) * O_NT + (x)]
#define MSC(k, x) ms[(size_t)(k) * KP + (x)]
#define ISC(k, x) is[(size_t)(k) * KP + (x)]
  float best = -1e36f;
  int32_t sca0 = NINF;
  {
    const int cur = L, dj0 = dsq[j0];
    int32_t *elc = el + (size_t)(cur & 1) * (M + 1) * 64;
    for (int k = 1; k <= M; k++) elc[a.at1(k)] = NINF;
    if (elon)
      for (int c = 0; c < h.el_ct[M + 1]; c++) elc[a.at1(h.el_idx[h.el_off[M + 1] + c])] = 0;
    mmx[a.at(cur, M)] = wa(lsum(wa(elc[a.at1(M)], TS(M, O_MEL)), TS(M, O_ME), tbl), MSC(M, dj0));
    imx[a.at(cur, M)] = wa(TS(M, O_IM), ISC(M, dj0));
    int32_t dnext = TS(M, O_DM);
    dmx[a.at(cur, M)] = dnext;
    for (int k = M - 1; k >= 1; k--) {
      int32_t v = TS(k, O_ME);
      v = lsum(v, wa(dnext, TS(k, O_MD)), tbl);
      if (elon) v = lsum(v, wa(elc[a.at1(k)], TS(k, O_MEL)), tbl);
      mmx[a.at(cur, k)] = wa(v, MSC(k, dj0));
      imx[a.at(cur, k)] = wa(wa(dnext, TS(k, O_ID)), ISC(k, dj0));
      dnext = wa(dnext, TS(k, O_DD));
      dmx[a.at(cur, k)] = dnext;
    }
    mmx[a.at(cur, 0)] = wa(dnext, TS(0, O_MD));
    imx[a.at(cur, 0)] = wa(wa(dnext, TS(0, O_ID)), ISC(0, dj0));
    dmx[a.at(cur, 0)] = NINF;
    elc[a.at1(0)] = NINF;
  }
  for (int i = j0 - 1; i >= i0; i--) {
    const int cur = i - i0 + 1, prv = cur + 1, di = dsq[i];
    int32_t *elc = el + (size_t)(cur & 1) * (M + 1) * 64, *elp = el + (size_t)(prv & 1) * (M + 1) * 64;
    for (int k = 0; k <= M; k++) elc[a.at1(k)] = NINF;
    if (elon && h.has_el[M]) elc[a.at1(M)] = wa(elc[a.at1(M)], h.el_selfsc);
    const int32_t iprvM = imx[a.at(prv, M)];
    int32_t mcM = wa(wa(iprvM, TS(M, O_MI)), MSC(M, di));
    if (elon && h.has_el[M]) mcM = lsum(mcM, wa(elc[a.at1(M)], TS(M, O_MEL)), tbl);
    int32_t icM = wa(wa(iprvM, TS(M, O_II)), ISC(M, di));
    int32_t dcM = wa(iprvM, TS(M, O_DI));
    if (elon) {
      const int32_t mpM = mmx[a.at(prv, M)];
      for (int c = 0; c < h.el_ct[M]; c++) {
        const int src = h.el_idx[h.el_off[M] + c];
        elc[a.at1(src)] = lsum(elc[a.at1(src)], mpM, tbl);
      }
    }
    if (do_scan) {
      if (elon)
        for (int c = 0; c < h.el_ct[M + 1]; c++) elc[a.at1(h.el_idx[h.el_off[M + 1] + c])] = 0;
      mcM = lsum(mcM, lsum(wa(elc[a.at1(M)], TS(M, O_MEL)), TS(M, O_ME), tbl), tbl);
      icM = lsum(icM, TS(M, O_IM), tbl);
      dcM = lsum(dcM, TS(M, O_DM), tbl);
    }
    mmx[a.at(cur, M)] = mcM;
    imx[a.at(cur, M)] = icM;
    dmx[a.at(cur, M)] = dcM;
    int32_t dnext = dcM, mpn = mmx[a.at(prv, M)];  // row cur at k+1; row prv M at k+1
    for (int k = M - 1; k >= 1; k--) {
      const int32_t mpk = mmx[a.at(prv, k)], ipk = imx[a.at(prv, k)];
      if (elon)
        for (int c = 0; c < h.el_ct[k]; c++) {
          const int src = h.el_idx[h.el_off[k] + c];
          elc[a.at1(src)] = lsum(elc[a.at1(src)], mpk, tbl);
        }
      if (elon && h.has_el[k]) elc[a.at1(k)] = lsum(elc[a.at1(k)], wa(elp[a.at1(k)], h.el_selfsc), tbl);
      int32_t mv = lsum(lsum(wa(mpn, TS(k, O_MM)), wa(ipk, TS(k, O_MI)), tbl), wa(dnext, TS(k, O_MD)), tbl);
      if (elon && h.has_el[k]) mv = lsum(mv, wa(elc[a.at1(k)], TS(k, O_MEL)), tbl);
      int32_t mc = wa(mv, MSC(k, di));
      const int32_t ic = wa(lsum(lsum(wa(mpn, TS(k, O_IM)), wa(ipk, TS(k, O_II)), tbl), wa(dnext, TS(k, O_ID)), tbl),
                            ISC(k, di));
      if (do_scan) mc = lsum(mc, TS(k, O_ME), tbl);
      const int32_t dc = lsum(lsum(wa(mpn, TS(k, O_DM)), wa(ipk, TS(k, O_DI)), tbl), wa(dnext, TS(k, O_DD)), tbl);
      mmx[a.at(cur, k)] = mc;
      imx[a.at(cur, k)] = ic;
      dmx[a.at(cur, k)] = dc;
      dnext = dc;
      mpn = mpk;
    }
    const int32_t ip0 = imx[a.at(prv, 0)];
    imx[a.at(cur, 0)] = wa(lsum(lsum(wa(mpn, TS(0, O_IM)), wa(ip0, TS(0, O_II)), tbl), wa(dnext, TS(0, O_ID)), tbl),
                           ISC(0, di));
    dmx[a.at(cur, 0)] = NINF;
    elc[a.at1(0)] = NINF;
    int32_t b = NINF;
    for (int k = M; k >= 1; k--) b = lsum(b, wa(mmx[a.at(prv, k)], TS(k, O_BM)), tbl);
    b = lsum(b, wa(ip0, TS(0, O_MI)), tbl);
    b = lsum(b, wa(dnext, TS(0, O_MD)), tbl);
    mmx[a.at(cur, 0)] = b;
    if (scorify(b) > best) best = scorify(b);
  }
  {
    const int cur = 0, prv = 1;
    mmx[a.at(cur, M)] = NINF;
    imx[a.at(cur, M)] = NINF;
    int32_t dnext = wa(imx[a.at(prv, M)], TS(M, O_DI));
    if (do_scan) dnext = lsum(dnext, TS(M, O_DM), tbl);
    dmx[a.at(cur, M)] = dnext;
    for (int k = M - 1; k >= 1; k--) {
      mmx[a.at(cur, k)] = NINF;
      imx[a.at(cur, k)] = NINF;
      dnext = lsum(lsum(wa(mmx[a.at(prv, k + 1)], TS(k, O_DM)), wa(imx[a.at(prv, k)], TS(k, O_DI)), tbl),
                   wa(dnext, TS(k, O_DD)), tbl);
      dmx[a.at(cur, k)] = dnext;
    }
    imx[a.at(cur, 0)] = NINF;
    dmx[a.at(cur, 0)] = NINF;
    int32_t b = NINF;
    for (int k = M; k >= 1; k--) b = lsum(b, wa(mmx[a.at(prv, k)], TS(k, O_BM)), tbl);
    b = lsum(b, wa(imx[a.at(prv, 0)], TS(0, O_MI)), tbl);
    b = lsum(b, wa(dnext, TS(0, O_MD)), tbl);
    mmx[a.at(cur, 0)] = b;
    sca0 = b;
    if (scorify(b) > best) best = scorify(b);
  }
  if (!do_scan) best = scorify(sca0);
  best_out[w.out] = best;
#undef TS
#undef MSC
#undef ISC
}

// cp9.rs:1517-1737: posterior mass crept in from both ends per state; the
// posterior cells are recomputed in the maximum scan.
__global__ __launch_bounds__(64) void cp9_post_kernel(Cp9Dev h, const Cp9WinDev *wins, int64_t n,
                                                     const Cp9GrpDev *grps, int did_scan, int32_t thresh,
                                                     int32_t *pn_out) {
  __shared__ uint16_t s_tbl[FLOGSUM_TBL];
  const uint16_t *tbl = load_table(h.ilog, s_tbl);
  const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (e >= n) return;
  const Cp9WinDev w = wins[e];
  const Cp9GrpDev g = grps[blockIdx.x];
  const int M = h.M, i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const Lane a{M + 1, (int)threadIdx.x};
  const int32_t *fm = g.fm, *fi = g.fi, *fd = g.fd, *bm = g.bm, *bi = g.bi, *bd = g.bd;
  const int32_t *ms = h.msc, *is = h.isc;
  const uint8_t *dsq = w.dsq;
  // per state: mass (M, I, D) and band ends; a band end is set once (>= 0)
  int32_t *mass = g.post, *pn = pn_out + (size_t)w.out * 6 * (M + 1);
  int32_t *pmin_m = pn, *pmax_m = pn + (M + 1), *pmin_i = pn + 2 * (M + 1), *pmax_i = pn + 3 * (M + 1);
  int32_t *pmin_d = pn + 4 * (M + 1), *pmax_d = pn + 5 * (M + 1);
  auto MS = [&](int s, int k) -> int32_t & { return mass[((size_t)s * (M + 1) + k) * 64 + a.l]; };
  for (int k = 0; k <= M; k++) {
    MS(0, k) = MS(1, k) = MS(2, k) = NINF;
    pmin_m[k] = pmin_i[k] = pmin_d[k] = pmax_m[k] = pmax_i[k] = pmax_d[k] = -1;
  }
  int32_t sc;
  if (did_scan) {
    sc = NINF;
    for (int ip = 0; ip <= L; ip++) sc = lsum(sc, bm[a.at(ip, 0)], tbl);
  } else {
    sc = bm[a.at(0, 0)];
  }
  const int32_t first = i0 - 1 < 0 ? 0 : i0 - 1;
  {
    const int32_t p = ws(wa(fm[a.at(0, 0)], bm[a.at(0, 0)]), sc);
    MS(0, 0) = p;
    if (p > thresh) pmin_m[0] = first;
    for (int k = 1; k <= M; k++) {
      const int32_t q = ws(wa(fd[a.at(0, k)], bd[a.at(0, k)]), sc);
      MS(2, k) = q;
      if (q > thresh) pmin_d[k] = first;
    }
  }
  for (int ip = 1; ip <= L; ip++) {
    const int i = i0 + ip - 1, x = dsq[i];
    if (pmin_m[0] < 0) {
      const int32_t p = max(ws(wa(fm[a.at(ip, 0)], bm[a.at(ip, 0)]), sc), NINF);
      const int32_t s = lsum(MS(0, 0), p, tbl);
      MS(0, 0) = s;
      if (s > thresh) pmin_m[0] = i;
    }
    if (pmin_i[0] < 0) {
      const int32_t p = max(ws(ws(wa(fi[a.at(ip, 0)], bi[a.at(ip, 0)]), is[x]), sc), NINF);
      const int32_t s = lsum(MS(1, 0), p, tbl);
      MS(1, 0) = s;
      if (s > thresh) pmin_i[0] = i;
    }
    for (int k = 1; k <= M; k++) {
      if (pmin_m[k] < 0) {
        const int32_t p = max(ws(ws(wa(fm[a.at(ip, k)], bm[a.at(ip, k)]), ms[(size_t)k * KP + x]), sc), NINF);
        const int32_t s = lsum(MS(0, k), p, tbl);
        MS(0, k) = s;
        if (s > thresh) pmin_m[k] = i;
      }
      if (pmin_i[k] < 0) {
        const int32_t p = max(ws(ws(wa(fi[a.at(ip, k)], bi[a.at(ip, k)]), is[(size_t)k * KP + x]), sc), NINF);
        const int32_t s = lsum(MS(1, k), p, tbl);
        MS(1, k) = s;
        if (s > thresh) pmin_i[k] = i;
      }
      if (pmin_d[k] < 0) {
        const int32_t p = max(ws(wa(fd[a.at(ip, k)], bd[a.at(ip, k)]), sc), NINF);
        const int32_t s = lsum(MS(2, k), p, tbl);
        MS(2, k) = s;
        if (s > thresh) pmin_d[k] = i;
      }
    }
  }
  for (int k = 0; k <= M; k++) MS(0, k) = MS(1, k) = MS(2, k) = NINF;
  for (int ip = L; ip >= 1; ip--) {
    const int i = i0 + ip - 1, x = dsq[i];
    for (int k = 0; k <= M; k++) {
      if (pmax_m[k] < 0) {
        const int32_t em = k ? ms[(size_t)k * KP + x] : 0;
        const int32_t p = max(ws(ws(wa(fm[a.at(ip, k)], bm[a.at(ip, k)]), em), sc), NINF);
        const int32_t s = lsum(MS(0, k), p, tbl);
        MS(0, k) = s;
        if (s > thresh) pmax_m[k] = i;
      }
      if (pmax_i[k] < 0) {
        const int32_t p = max(ws(ws(wa(fi[a.at(ip, k)], bi[a.at(ip, k)]), is[(size_t)k * KP + x]), sc), NINF);
        const int32_t s = lsum(MS(1, k), p, tbl);
        MS(1, k) = s;
        if (s > thresh) pmax_i[k] = i;
      }
      if (pmax_d[k] < 0) {
        const int32_t p = k ? max(ws(wa(fd[a.at(ip, k)], bd[a.at(ip, k)]), sc), NINF) : NINF;
        const int32_t s = lsum(MS(2, k), p, tbl);
        MS(2, k) = s;
        if (s > thresh) pmax_d[k] = i;
      }
    }
  }
  if (pmax_m[0] < 0) {
    const int32_t s = lsum(MS(0, 0), ws(wa(fm[a.at(0, 0)], bm[a.at(0, 0)]), sc), tbl);
    if (s > thresh) pmax_m[0] = first;
  }
  for (int k = 1; k <= M; k++)
    if (pmax_d[k] < 0) {
      const int32_t s = lsum(MS(2, k), ws(wa(fd[a.at(0, k)], bd[a.at(0, k)]), sc), tbl);
      if (s > thresh) pmax_d[k] = first;
    }
  for (int k = 0; k <= M; k++) {
    if (pmin_m[k] < 0 || pmax_m[k] < 0 || pmax_m[k] < pmin_m[k]) pmin_m[k] = pmax_m[k] = -1;
    if (pmin_i[k] < 0 || pmax_i[k] < 0 || pmax_i[k] < pmin_i[k]) pmin_i[k] = pmax_i[k] = -1;
    if (pmin_d[k] < 0 || pmax_d[k] < 0 || pmax_d[k] < pmin_d[k]) pmin_d[k] = pmax_d[k] = -1;
  }
  pmin_d[0] = pmax_d[0] = -1;
}

}  // namespace

hipError_t launch_cp9_bands(const Cp9Dev &h, const Cp9WinDev *wins, int64_t n, const Cp9GrpDev *grps, int do_scan,
                            int32_t thresh, float *fsc, float *bsc, int32_t *pn, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 63) / 64)), block(64);
  hipLaunchKernelGGL(cp9_fwd_kernel, grid, block, 0, s, h, wins, n, grps, do_scan, fsc);
  hipLaunchKernelGGL(cp9_bwd_kernel, grid, block, 0, s, h, wins, n, grps, do_scan, bsc);
  hipLaunchKernelGGL(cp9_post_kernel, grid, block, 0, s, h, wins, n, grps, do_scan, thresh, pn);
  return hipGetLastError();
}

}  // namespace ixg

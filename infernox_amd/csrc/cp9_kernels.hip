// gfx950 CP9 HMM banding: the integer Forward and Backward of the band
// producing HMM over a window and its posterior HMM bands
// (cp9.rs:1275-1453 cp9_forward, :1995-2208 cp9_backward, :1517-1737
// cp9_fb2hmmbands). One wave per window, its lanes over the HMM nodes k:
//   * a row's match, insert and local-end (EL) cells read only the previous
//     row (Forward) or the previous row and this row's delete cells
//     (Backward): the lanes compute them for all k at once;
//   * the delete chain D(k) <- D(k -/+ 1) and the row's end (Forward) or
//     begin (Backward) sum are serial in k. Their terms that do not depend on
//     the chain are computed lane-parallel into LDS first; then lane 0 walks
//     the delete chain and lane 1 the sum in the same instruction stream, both
//     as acc = ILogsum(u[k], acc + z[k]);
//   * a Backward EL cell accumulates the next row's match cells of the nodes
//     it ends at, in the reference's decreasing-k order, then its own self
//     loop; a local end's source node precedes its end node, so the cell is
//     complete before the match cell at its node reads it. One lane per
//     source node walks a host-built list of the nodes it ends at;
//   * the posterior bands are per-node scans over the rows: one lane per node.
// Every ILogsum takes the reference's operands in its order, so matrices,
// scores and bands are bit-identical. The M/I/D matrices live in global
// memory (the posterior pass reads them all); the rolling EL rows, the chain
// terms and the ILogsum table (16-bit entries, up to its last nonzero one)
// live in LDS.
#include "cm_internal.h"

namespace ixg {

namespace {

constexpr int32_t NINF = -INFTY_I;
enum { O_MM = 0, O_IM, O_DM, O_BM, O_MI, O_II, O_DI, O_MD, O_ID, O_DD, O_ME, O_MEL, O_NT };

__device__ __forceinline__ int32_t wa(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t ws(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

// ILogsum (cp9.rs:65-74): max + T[max - min] when the smaller term is above
// -INFTY, else the max. The reference's table is zero from nz on; the LDS
// copy holds entries 0..nz (T[nz] = 0) and the index is clamped to nz, and
// to nz as well when the smaller term is at or below -INFTY, so every case
// that returns the max reads a 0 (a short dependency chain: this is the
// serial step of the delete chains).
__device__ __forceinline__ int32_t lsum(int32_t a, int32_t b, const uint16_t *tbl, int nz) {
  const int32_t mx = max(NINF, max(a, b)), mn = min(a, b);
  const uint32_t diff = (uint32_t)mx - (uint32_t)mn;
  const uint32_t idx = mn > NINF ? min(diff, (uint32_t)nz) : (uint32_t)nz;
  return wa(mx, (int32_t)tbl[idx]);
}

__device__ __forceinline__ float scorify(int32_t sc) { return (float)sc / 1000.0f; }  // cp9.rs:1257

struct Lds {
  int32_t *el, *ua, *ub, *z;  // EL rows [2][M+1]; chain terms; a transition column
  uint16_t *tbl;
};

__device__ __forceinline__ Lds lds_setup(const Cp9Dev &h, int zcol, int zshift) {
  extern __shared__ int32_t lds[];
  const int M1 = h.M + 1;
  Lds s{lds, lds + 2 * M1, lds + 3 * M1, lds + 4 * M1, (uint16_t *)(lds + 5 * M1)};
  const IXG_GLOBAL int32_t *ilog = gmem(h.ilog), *otsc = gmem(h.otsc);
  for (int x = threadIdx.x; x <= h.nz; x += 64) s.tbl[x] = x < h.nz ? (uint16_t)ilog[x] : (uint16_t)0;
  // z[k] = the transition zcol of node k - zshift
  for (int k = threadIdx.x; k <= h.M; k += 64) s.z[k] = k >= zshift ? otsc[(size_t)(k - zshift) * O_NT + zcol] : 0;
  return s;
}

#define TS(k, x) g.otsc[(size_t)(k) * O_NT + (x)]
#define MSC(k, x) g.msc[(size_t)(k) * KP + (x)]
#define ISC(k, x) g.isc[(size_t)(k) * KP + (x)]

// The HMM's tables as global pointers.
struct G {
  const IXG_GLOBAL int32_t *otsc, *msc, *isc, *el_ct, *el_off, *el_idx, *bel_nM, *bel_reset, *bel_off, *bel_k;
  const IXG_GLOBAL uint8_t *has_el, *dsq;
};
__device__ __forceinline__ G gtabs(const Cp9Dev &h, const Cp9WinDev &w) {
  return G{gmem(h.otsc),   gmem(h.msc),    gmem(h.isc),       gmem(h.el_ct),   gmem(h.el_off), gmem(h.el_idx),
           gmem(h.bel_nM), gmem(h.bel_reset), gmem(h.bel_off), gmem(h.bel_k), gmem(h.has_el), gmem(w.dsq)};
}
#define LS(a, b) lsum((a), (b), s.tbl, nz)

// acc = ILogsum(u[k], acc + z[k]) over k = k0, k0 + dk, ... (n steps); lane 0
// stores each acc to out[k] when out is set. z is used by lane 0 only.
__device__ __forceinline__ int32_t chain(const int32_t *u, const int32_t *z, int32_t acc, int k0, int dk, int n,
                                        IXG_GLOBAL int32_t *out, const uint16_t *tbl, int nz) {
  int k = k0;
  int c = 0;
  for (; c + 4 <= n; c += 4) {
    int32_t uu[4], zz[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uu[q] = u[k + q * dk];
      zz[q] = z ? z[k + q * dk] : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      acc = lsum(uu[q], wa(acc, zz[q]), tbl, nz);
      if (out) out[k + q * dk] = acc;
    }
    k += 4 * dk;
  }
  for (; c < n; c++, k += dk) {
    acc = lsum(u[k], wa(acc, z ? z[k] : 0), tbl, nz);
    if (out) out[k] = acc;
  }
  return acc;
}

__global__ __launch_bounds__(64) void cp9_fwd_kernel(Cp9Dev h, const Cp9WinDev *wins, int do_scan, int align,
                                                    float *best_out) {
  // z[k] = DD(k - 1): the delete chain's self term into node k
  const Lds s = lds_setup(h, O_DD, 1);
  const int t = threadIdx.x, M = h.M, M1 = M + 1, nz = h.nz;
  const Cp9WinDev w = wins[blockIdx.x];
  const int i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * M1;
  const G g = gtabs(h, w);
  IXG_GLOBAL int32_t *gm = gmem(w.mx), *gi = gm + cells, *gd = gi + cells;
  const bool el = h.el != 0;
  // row 0, position i0 - 1 (cp9.rs:1309-1333)
  for (int k = t; k <= M; k += 64) {
    gm[k] = k ? NINF : 0;
    gi[k] = NINF;
    s.el[k] = NINF;
  }
  __syncthreads();
  if (t == 0) {
    int32_t d = NINF;
    gd[0] = NINF;
    for (int k = 1; k <= M; k++) {
      d = LS(LS(wa(k == 1 ? 0 : NINF, TS(k - 1, O_MD)), wa(NINF, TS(k - 1, O_ID))), wa(d, s.z[k]));
      gd[k] = d;
    }
  }
  __syncthreads();
  float best = -1e36f;
  int32_t sca = wa(gd[M], TS(M, O_DM));
  if (t == 1) best = scorify(sca);
  for (int j = i0; j <= j0; j++) {
    const int dj = g.dsq[j], cur = j - i0 + 1;
    const IXG_GLOBAL int32_t *mp = gm + (size_t)(cur - 1) * M1, *ip = gi + (size_t)(cur - 1) * M1,
                             *dp = gd + (size_t)(cur - 1) * M1;
    IXG_GLOBAL int32_t *mc = gm + (size_t)cur * M1, *ic = gi + (size_t)cur * M1, *dc = gd + (size_t)cur * M1;
    const int32_t *ep = s.el + ((cur - 1) & 1) * M1;
    int32_t *ec = s.el + (cur & 1) * M1;
    const int32_t mp0 = mp[0];
    // match, insert and EL cells from the previous row (cp9.rs:1377-1418),
    // and the terms of this row's delete chain and end sum
    for (int k = t; k <= M; k += 64) {
      int32_t m, in;
      if (k == 0) {
        m = do_scan ? 0 : NINF;
        ec[0] = NINF;
        const int32_t sc = LS(LS(wa(mp0, TS(0, O_MI)), wa(ip[0], TS(0, O_II))), wa(dp[0], TS(0, O_DI)));
        in = max(wa(sc, ISC(0, dj)), NINF);
      } else {
        int32_t sc = LS(LS(wa(mp[k - 1], TS(k - 1, O_MM)), wa(ip[k - 1], TS(k - 1, O_IM))),
                        LS(wa(dp[k - 1], TS(k - 1, O_DM)), wa(mp0, TS(k, O_BM))));
        const int n = g.el_ct[k], o = g.el_off[k];
        for (int c = 0; c < n; c++) sc = LS(sc, ep[g.el_idx[o + c]]);
        m = max(wa(sc, MSC(k, dj)), NINF);
        sc = LS(LS(wa(mp[k], TS(k, O_MI)), wa(ip[k], TS(k, O_II))), wa(dp[k], TS(k, O_DI)));
        in = max(wa(sc, ISC(k, dj)), NINF);
        ec[k] = (el && g.has_el[k]) ? LS(wa(m, TS(k, O_MEL)), wa(ep[k], h.el_selfsc)) : NINF;
        s.ub[k] = max(wa(m, TS(k, O_ME)), NINF);  // ILogsum(x, -INFTY)
      }
      mc[k] = m;
      ic[k] = in;
      if (k < M) s.ua[k + 1] = LS(wa(m, TS(k, O_MD)), wa(in, TS(k, O_ID)));
    }
    __syncthreads();
    // lane 0: D(k) = ILogsum(ua[k], D(k-1) + DD(k-1)); lane 1: the end sum
    int32_t acc = NINF;
    if (t < 2) {
      if (t == 0) dc[0] = NINF;
      acc = chain(t ? s.ub : s.ua, t ? nullptr : s.z, NINF, 1, 1, M, t ? nullptr : dc, s.tbl, nz);
    }
    const int32_t dM = __shfl(acc, 0, 64);
    if (t == 1) {  // cp9.rs:1420-1430
      int32_t endsc = LS(LS(acc, wa(dM, TS(M, O_DM))), wa(ic[M], TS(M, O_IM)));
      const int n = g.el_ct[M + 1], o = g.el_off[M + 1];
      for (int c = 0; c < n; c++) endsc = LS(endsc, ec[g.el_idx[o + c]]);
      sca = endsc;
      if (scorify(endsc) > best) best = scorify(endsc);
    }
    __syncthreads();
  }
  if (t == 1) best_out[w.out] = align ? scorify(sca) : best;  // doing_align: the whole window's score
}

__global__ __launch_bounds__(64) void cp9_bwd_kernel(Cp9Dev h, const Cp9WinDev *wins, int do_scan, int align,
                                                    float *best_out) {
  // z[k] = DD(k): the delete chain's self term out of node k
  const Lds s = lds_setup(h, O_DD, 0);
  const int t = threadIdx.x, M = h.M, M1 = M + 1, nz = h.nz;
  const Cp9WinDev w = wins[blockIdx.x];
  const int i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * M1;
  const G g = gtabs(h, w);
  IXG_GLOBAL int32_t *gm = gmem(w.mx) + 3 * cells, *gi = gm + cells, *gd = gi + cells;
  const bool el = h.el != 0;
  __syncthreads();
  if (t == 0) {  // row L, position j0 (cp9.rs:2029-2055)
    const int dj0 = g.dsq[j0];
    int32_t *e = s.el + (L & 1) * M1;
    IXG_GLOBAL int32_t *m = gm + (size_t)L * M1, *i = gi + (size_t)L * M1, *d = gd + (size_t)L * M1;
    for (int k = 1; k <= M; k++) e[k] = NINF;
    if (el)
      for (int c = 0; c < g.el_ct[M + 1]; c++) e[g.el_idx[g.el_off[M + 1] + c]] = 0;
    e[0] = NINF;
    m[M] = wa(LS(wa(e[M], TS(M, O_MEL)), TS(M, O_ME)), MSC(M, dj0));
    i[M] = wa(TS(M, O_IM), ISC(M, dj0));
    int32_t dn = TS(M, O_DM);
    d[M] = dn;
    for (int k = M - 1; k >= 1; k--) {
      int32_t v = TS(k, O_ME);
      v = LS(v, wa(dn, TS(k, O_MD)));
      if (el) v = LS(v, wa(e[k], TS(k, O_MEL)));
      m[k] = wa(v, MSC(k, dj0));
      i[k] = wa(wa(dn, TS(k, O_ID)), ISC(k, dj0));
      dn = wa(dn, TS(k, O_DD));
      d[k] = dn;
    }
    m[0] = wa(dn, TS(0, O_MD));
    i[0] = wa(wa(dn, TS(0, O_ID)), ISC(0, dj0));
    d[0] = NINF;
  }
  __syncthreads();
  float best = -1e36f;
  int32_t b = NINF;
  // rows L-1 .. 1 (positions j0-1 .. i0, cp9.rs:2057-2169), then row 0
  // (position i0 - 1, cp9.rs:2171-2200)
  for (int cur = L - 1; cur >= 0; cur--) {
    const int di = cur ? g.dsq[i0 + cur - 1] : 0;
    const IXG_GLOBAL int32_t *mp = gm + (size_t)(cur + 1) * M1, *ip = gi + (size_t)(cur + 1) * M1;
    IXG_GLOBAL int32_t *mc = gm + (size_t)cur * M1, *ic = gi + (size_t)cur * M1, *dc = gd + (size_t)cur * M1;
    const int32_t *ep = s.el + ((cur + 1) & 1) * M1;
    int32_t *ec = s.el + (cur & 1) * M1;
    // EL cells (cp9.rs:2068-2119), and the terms of the delete chain and the
    // begin sum, all from the previous row
    for (int y = t; y <= M; y += 64) {
      if (cur > 0) {
        int32_t v = NINF;
        if (y == M && el && g.has_el[M]) v = wa(v, h.el_selfsc);
        const int nM = g.bel_nM[y];
        for (int c = 0; c < nM; c++) v = LS(v, mp[M]);
        if (do_scan && g.bel_reset[y]) v = 0;
        const int o = g.bel_off[y], n = g.bel_off[y + 1] - o;
        for (int c = 0; c < n; c++) v = LS(v, mp[g.bel_k[o + c]]);
        if (el && y >= 1 && y < M && g.has_el[y]) v = LS(v, wa(ep[y], h.el_selfsc));
        ec[y] = y ? v : NINF;
      }
      if (y >= 1 && y < M) s.ua[y] = LS(wa(mp[y + 1], TS(y, O_DM)), wa(ip[y], TS(y, O_DI)));
      if (y >= 2) s.ub[y] = max(wa(mp[y], TS(y, O_BM)), NINF);  // ILogsum(x, -INFTY)
    }
    __syncthreads();
    // lane 0: D(k) = ILogsum(ua[k], D(k+1) + DD(k)) for k = M-1 .. 1; lane 1:
    // the begin sum over M(k) + BM(k) for k = M .. 2
    int32_t acc = NINF;
    if (t < 2) {
      if (t == 0) {
        acc = wa(ip[M], TS(M, O_DI));
        if (do_scan) acc = LS(acc, TS(M, O_DM));
        dc[M] = acc;
        dc[0] = NINF;
        acc = chain(s.ua, s.z, acc, M - 1, -1, M - 1, dc, s.tbl, nz);
      } else {
        acc = chain(s.ub + 1, nullptr, NINF, M - 1, -1, M - 1, nullptr, s.tbl, nz);
      }
    }
    __syncthreads();
    // match and insert cells (cp9.rs:2078-2146); the begin cell
    for (int k = t; k <= M; k += 64) {
      int32_t m, in;
      if (cur == 0) {
        m = NINF;
        in = NINF;
      } else if (k == 0) {
        m = NINF;  // the begin cell, written by lane 1 below
        in = wa(LS(LS(wa(mp[1], TS(0, O_IM)), wa(ip[0], TS(0, O_II))), wa(dc[1], TS(0, O_ID))), ISC(0, di));
      } else if (k == M) {
        const bool sel = el && g.has_el[M];
        m = wa(wa(ip[M], TS(M, O_MI)), MSC(M, di));
        if (sel) m = LS(m, wa(wa(NINF, h.el_selfsc), TS(M, O_MEL)));
        in = wa(wa(ip[M], TS(M, O_II)), ISC(M, di));
        if (do_scan) {
          m = LS(m, LS(wa(ec[M], TS(M, O_MEL)), TS(M, O_ME)));
          in = LS(in, TS(M, O_IM));
        }
      } else {
        int32_t mv = LS(LS(wa(mp[k + 1], TS(k, O_MM)), wa(ip[k], TS(k, O_MI))), wa(dc[k + 1], TS(k, O_MD)));
        if (el && g.has_el[k]) mv = LS(mv, wa(ec[k], TS(k, O_MEL)));
        m = wa(mv, MSC(k, di));
        in = wa(LS(LS(wa(mp[k + 1], TS(k, O_IM)), wa(ip[k], TS(k, O_II))), wa(dc[k + 1], TS(k, O_ID))), ISC(k, di));
        if (do_scan) m = LS(m, TS(k, O_ME));
      }
      if (k) mc[k] = m;
      ic[k] = in;
    }
    if (t == 1) {  // cp9.rs:2147-2167: k = 1, then I(0), then D(1)
      b = LS(acc, wa(mp[1], TS(1, O_BM)));
      b = LS(b, wa(ip[0], TS(0, O_MI)));
      b = LS(b, wa(dc[1], TS(0, O_MD)));
      mc[0] = b;
      if (scorify(b) > best) best = scorify(b);
    }
    __syncthreads();
  }
  if (t == 1) best_out[w.out] = align ? scorify(b) : best;  // doing_align: the whole window's score
}

// cp9.rs:1517-1737: posterior mass crept in from both ends, one lane per
// node; the posterior cells are recomputed in the maximum scan.
// occ_out (may be null): per node the posterior occupancy ILogsum over rows
// 0..L of ILogsum(M, D) posterior (cp9.rs:1765-1776, the quantity the
// truncated bands' start / end prediction thresholds), [n][M+1], node 0
// -INFTY.
__global__ __launch_bounds__(64) void cp9_post_kernel(Cp9Dev h, const Cp9WinDev *wins, int did_scan, int32_t thresh,
                                                     int32_t *pn_out, int32_t *occ_out) {
  const Lds s = lds_setup(h, O_DD, 0);
  __shared__ int32_t s_sc;
  const int t = threadIdx.x, M = h.M, M1 = M + 1, nz = h.nz;
  const Cp9WinDev w = wins[blockIdx.x];
  const int i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * M1;
  const G g = gtabs(h, w);
  const IXG_GLOBAL int32_t *fm = gmem(w.mx), *fi = fm + cells, *fd = fi + cells, *bm = fd + cells, *bi = bm + cells,
                           *bd = bi + cells;
  __syncthreads();
  if (t == 0) {
    int32_t sc;
    if (did_scan) {
      sc = NINF;
      for (int r = 0; r <= L; r++) sc = LS(sc, bm[(size_t)r * M1]);
    } else {
      sc = bm[0];
    }
    s_sc = sc;
  }
  __syncthreads();
  const int32_t sc = s_sc;
  IXG_GLOBAL int32_t *pn = gmem(pn_out) + (size_t)w.out * 6 * M1;
  const int32_t first = i0 - 1 < 0 ? 0 : i0 - 1;
  for (int k = t; k <= M; k += 64) {
    int32_t mm = NINF, mi = NINF, md = NINF;
    int32_t nm = -1, ni = -1, nd = -1, xm = -1, xi = -1, xd = -1;
    // minimum scan: row 0, then rows 1 .. L
    if (k == 0) {
      mm = ws(wa(fm[0], bm[0]), sc);
      if (mm > thresh) nm = first;
    } else {
      md = ws(wa(fd[k], bd[k]), sc);
      if (md > thresh) nd = first;
    }
    for (int r = 1; r <= L; r++) {
      const int i = i0 + r - 1, x = g.dsq[i];
      const size_t o = (size_t)r * M1 + k;
      if (nm < 0) {
        mm = LS(mm, max(ws(ws(wa(fm[o], bm[o]), k ? MSC(k, x) : 0), sc), NINF));
        if (mm > thresh) nm = i;
      }
      if (ni < 0) {
        mi = LS(mi, max(ws(ws(wa(fi[o], bi[o]), ISC(k, x)), sc), NINF));
        if (mi > thresh) ni = i;
      }
      if (k > 0 && nd < 0) {
        md = LS(md, max(ws(wa(fd[o], bd[o]), sc), NINF));
        if (md > thresh) nd = i;
      }
    }
    // maximum scan: rows L .. 1, then row 0
    mm = mi = md = NINF;
    for (int r = L; r >= 1; r--) {
      const int i = i0 + r - 1, x = g.dsq[i];
      const size_t o = (size_t)r * M1 + k;
      if (xm < 0) {
        mm = LS(mm, max(ws(ws(wa(fm[o], bm[o]), k ? MSC(k, x) : 0), sc), NINF));
        if (mm > thresh) xm = i;
      }
      if (xi < 0) {
        mi = LS(mi, max(ws(ws(wa(fi[o], bi[o]), ISC(k, x)), sc), NINF));
        if (mi > thresh) xi = i;
      }
      if (k > 0 && xd < 0) {
        md = LS(md, max(ws(wa(fd[o], bd[o]), sc), NINF));
        if (md > thresh) xd = i;
      }
    }
    if (k == 0 && xm < 0) {
      mm = LS(mm, ws(wa(fm[0], bm[0]), sc));
      if (mm > thresh) xm = first;
    }
    if (k > 0 && xd < 0) {
      md = LS(md, ws(wa(fd[k], bd[k]), sc));
      if (md > thresh) xd = first;
    }
    if (nm < 0 || xm < 0 || xm < nm) nm = xm = -1;
    if (ni < 0 || xi < 0 || xi < ni) ni = xi = -1;
    if (k == 0 || nd < 0 || xd < 0 || xd < nd) nd = xd = -1;
    pn[k] = nm;
    pn[M1 + k] = xm;
    pn[2 * M1 + k] = ni;
    pn[3 * M1 + k] = xi;
    pn[4 * M1 + k] = nd;
    pn[5 * M1 + k] = xd;
    if (occ_out) {
      int32_t o = NINF;
      if (k > 0) {
        o = LS(NINF, LS(NINF, ws(wa(fd[k], bd[k]), sc)));
        for (int r = 1; r <= L; r++) {
          const int x = g.dsq[i0 + r - 1];
          const size_t c = (size_t)r * M1 + k;
          const int32_t m = max(ws(ws(wa(fm[c], bm[c]), MSC(k, x)), sc), NINF);
          const int32_t d = max(ws(wa(fd[c], bd[c]), sc), NINF);
          o = LS(o, LS(m, d));
        }
      }
      gmem(occ_out)[(size_t)w.out * M1 + k] = o;
    }
  }
}

// ---- anti-diagonal wavefront forms -----------------------------------------
// One wave per window, lane p holding one row of a 64-row strip: the Forward
// visits row s0 + 1 + p at node k = step - p, the Backward row L - 1 - s0 - p
// at node k = M - step + p. A cell's neighbours in the row before (the row
// after, for the Backward) were computed by lane p - 1 one step earlier and
// arrive by a one-lane DPP shift; lane 0 reads them from the matrices in
// global memory (the previous strip's last row) and from a local-end row that
// lane 63 left in LDS. The delete chain and the end/begin sum run along each
// lane's own row, one node per step, so a strip takes M + 64 steps instead of
// the row-at-a-time form's M dependent steps per row. Values a lane needs
// from the row before at an earlier node (the local-end cells a Forward match
// cell sums; the match cells a Backward local-end cell sums) go through LDS
// slots, one column per lane, that the host assigns so that no slot is
// rewritten before its last reader has read it.

// the row before's cell from lane p - 1; lane 0 gets `own`
__device__ __forceinline__ int32_t from_prev_lane(int32_t own, int32_t v) {
  return __builtin_amdgcn_update_dpp(own, v, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}

struct WaveLds {
  int32_t *slot;    // [nslots][65]: column c holds lane c - 1's values
  int32_t *elb;     // [2][M+1]: the local-end row of a strip's last row, for the next strip's lane 0
  int32_t *mb;      // [2][M+1]: the same row's match cells (Backward)
  uint32_t *info;   // [2][M+2]: the packed per-node words
  int16_t *lnode;   // the local-end list's entries: node ...
  int16_t *lslot;   // ... and its slot
  uint16_t *tbl;
};
__device__ __forceinline__ WaveLds wave_lds(const Cp9Dev &h, int nslots, const int32_t *node, const int32_t *slot,
                                            int nlist, const uint32_t *info) {
  extern __shared__ int32_t lds[];
  uint32_t *in = (uint32_t *)(lds + nslots * 65 + 4 * (h.M + 1));
  int16_t *ln = (int16_t *)(in + 2 * (h.M + 2));
  WaveLds s{lds, lds + nslots * 65, lds + nslots * 65 + 2 * (h.M + 1), in, ln, ln + nlist, (uint16_t *)(ln + 2 * nlist)};
  const IXG_GLOBAL int32_t *ilog = gmem(h.ilog), *gn = gmem(node), *gs = gmem(slot);
  const IXG_GLOBAL uint32_t *gi = gmem(info);
  for (int x = threadIdx.x; x < 2 * (h.M + 2); x += 64) s.info[x] = gi[x];
  for (int x = threadIdx.x; x <= h.nz; x += 64) s.tbl[x] = x < h.nz ? (uint16_t)ilog[x] : (uint16_t)0;
  for (int x = threadIdx.x; x < nlist; x += 64) {
    s.lnode[x] = (int16_t)gn[x];
    s.lslot[x] = (int16_t)gs[x];
  }
  return s;
}

// One node's transitions (cp9O order: MM IM DM BM | MI II DI MD | ID DD ME
// MEL) and emission scores, loaded a step ahead of their use.
struct NodeW {
  int4 a, b, c;
  int32_t ms, is;
};
__device__ __forceinline__ NodeW node_w(const G &g, int k, int x) {
  const IXG_GLOBAL int32_t *t = g.otsc + (size_t)k * O_NT;
  return NodeW{make_int4(t[0], t[1], t[2], t[3]), make_int4(t[4], t[5], t[6], t[7]),
               make_int4(t[8], t[9], t[10], t[11]), g.msc[(size_t)k * KP + x], g.isc[(size_t)k * KP + x]};
}
#define N_MM(n) (n).a.x
#define N_IM(n) (n).a.y
#define N_DM(n) (n).a.z
#define N_BM(n) (n).a.w
#define N_MI(n) (n).b.x
#define N_II(n) (n).b.y
#define N_DI(n) (n).b.z
#define N_MD(n) (n).b.w
#define N_ID(n) (n).c.x
#define N_DD(n) (n).c.y
#define N_ME(n) (n).c.z
#define N_MEL(n) (n).c.w

#define LW(a, b) lsum((a), (b), s.tbl, nz)

__device__ __forceinline__ void cp9_fwd_wave_body(const Cp9Dev &h, const Cp9WinDev *wins, int wi, int do_scan,
                                          int align, float *best_out) {
  const WaveLds s = wave_lds(h, h.nsf, h.el_idx, h.el_slot, h.nel, h.finfo);
  const int t = threadIdx.x, M = h.M, M1 = M + 1, nz = h.nz;
  const Cp9WinDev w = wins[wi];
  const G g = gtabs(h, w);
  const int i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * M1;
  IXG_GLOBAL int32_t *gm = gmem(w.mx), *gi = gm + cells, *gd = gi + cells;
  const bool el = h.el != 0;
  // row 0, position i0 - 1 (cp9.rs:1309-1333); its local-end row is -INFTY
  for (int k = t; k <= M; k += 64) {
    gm[k] = k ? NINF : 0;
    gi[k] = NINF;
    s.elb[k] = NINF;
  }
  __syncthreads();
  if (t == 0) {
    int32_t d = NINF;
    gd[0] = NINF;
    for (int k = 1; k <= M; k++) {
      d = LW(LW(wa(k == 1 ? 0 : NINF, TS(k - 1, O_MD)), wa(NINF, TS(k - 1, O_ID))), wa(d, TS(k - 1, O_DD)));
      gd[k] = d;
    }
  }
  __syncthreads();
  float best = t == 0 ? scorify(wa(gd[M], TS(M, O_DM))) : -1e36f;
  for (int s0 = 0; s0 < L; s0 += 64) {
    const int r = s0 + 1 + t, par = (s0 >> 6) & 1;
    const bool live = r <= L;
    const int x = live ? g.dsq[i0 + r - 1] : 0;
    const IXG_GLOBAL int32_t *um = gm + (size_t)s0 * M1, *ui = gi + (size_t)s0 * M1, *ud = gd + (size_t)s0 * M1;
    IXG_GLOBAL int32_t *om = gm + (size_t)(live ? r : 0) * M1, *oi = gi + (size_t)(live ? r : 0) * M1,
                       *od = gd + (size_t)(live ? r : 0) * M1;
    const int32_t *uel = s.elb + par * M1;
    int32_t *nel = s.elb + (par ^ 1) * M1;
    int32_t oM = NINF, oI = NINF, oD = NINF, oE = NINF;  // this lane's cells at its previous node
    int32_t lM = NINF, lI = NINF, lD = NINF;             // the row before at the previous node
    int32_t mp0 = NINF, endsc = NINF;
    NodeW pn = node_w(g, 0, x), nx = node_w(g, min(max(-t, 0), M), x);  // the previous and next nodes
    // lane 0: the row before's cells at its next node, read a step ahead
    // (every lane loads, at a clamped node, so that the step's memory
    // operations are the same on every path and a wait for them never has
    // to wait for the step's stores)
    int32_t nM = um[0], nI = ui[0], nD = ud[0];
    IXG_GLOBAL int32_t *sink = gm + 6 * cells + t;
    for (int step = 0; step <= M + 63; step++) {
      const int k = step - t;
      const bool act = live && k >= 0 && k <= M;
      const NodeW cn = nx;
      nx = node_w(g, min(max(k + 1, 0), M), x);
      const int32_t gM = nM, gI = nI, gD = nD, gE = (t == 0 && act) ? uel[k] : NINF;
      {
        const int kn = min(max(k + 1, 0), M);
        nM = um[kn];
        nI = ui[kn];
        nD = ud[kn];
      }
      int32_t cM = NINF, cI = NINF, cD = NINF;
      const int32_t uM = from_prev_lane(gM, oM), uI = from_prev_lane(gI, oI), uD = from_prev_lane(gD, oD),
                    uE = from_prev_lane(gE, oE);
      if (act) {
        int32_t cE;
        if (k == 0) {  // cp9.rs:1362-1371
          cM = do_scan ? 0 : NINF;
          cI = max(wa(LW(LW(wa(uM, N_MI(cn)), wa(uI, N_II(cn))), wa(uD, N_DI(cn))), cn.is), NINF);
          cD = NINF;
          cE = NINF;
          mp0 = uM;
        } else {  // cp9.rs:1377-1418
          int32_t sc = LW(LW(wa(lM, N_MM(pn)), wa(lI, N_IM(pn))), LW(wa(lD, N_DM(pn)), wa(mp0, N_BM(cn))));
          const uint32_t i0w = s.info[k];
          const int n = i0w & 0xFFFF, o = i0w >> 16;
          for (int c = 0; c < n; c++)
            sc = LW(sc, t == 0 ? uel[s.lnode[o + c]] : s.slot[s.lslot[o + c] * 65 + t]);
          cM = max(wa(sc, cn.ms), NINF);
          cI = max(wa(LW(LW(wa(uM, N_MI(cn)), wa(uI, N_II(cn))), wa(uD, N_DI(cn))), cn.is), NINF);
          cD = LW(LW(wa(oM, N_MD(pn)), wa(oI, N_ID(pn))), wa(oD, N_DD(pn)));
          cE = (el && (s.info[M + 2 + k] >> 16 & 1)) ? LW(wa(cM, N_MEL(cn)), wa(uE, h.el_selfsc)) : NINF;
          endsc = LW(endsc, wa(cM, N_ME(cn)));
        }
        const int sl = (int16_t)(s.info[M + 2 + k] & 0xFFFF);
        if (sl >= 0) s.slot[sl * 65 + t + 1] = cE;
        if (t == 63) nel[k] = cE;
        if (k == M) {  // cp9.rs:1420-1430
          endsc = LW(LW(endsc, wa(cD, N_DM(cn))), wa(cI, N_IM(cn)));
          const uint32_t iw = s.info[M + 1];
          const int n = iw & 0xFFFF, o = iw >> 16;
          for (int c = 0; c < n; c++) endsc = LW(endsc, s.slot[s.lslot[o + c] * 65 + t + 1]);
          if (scorify(endsc) > best) best = scorify(endsc);
          if (align && r == L) best_out[w.out] = scorify(endsc);  // doing_align: the whole window's score
        }
        oM = cM;
        oI = cI;
        oD = cD;
        oE = cE;
        lM = uM;
        lI = uI;
        lD = uD;
        pn = cn;
      }
      *(act ? om + k : sink) = cM;
      *(act ? oi + k : sink) = cI;
      *(act ? od + k : sink) = cD;
    }
    __syncthreads();
  }
  for (int off = 32; off >= 1; off >>= 1) best = fmaxf(best, __shfl_xor(best, off));
  if (!align && t == 0) best_out[w.out] = best;
}

__global__ __launch_bounds__(64) void cp9_fwd_wave(Cp9Dev h, const Cp9WinDev *wins, int do_scan, int align,
                                              float *best_out) {
  cp9_fwd_wave_body(h, wins, blockIdx.x, do_scan, align, best_out);
}

__device__ __forceinline__ void cp9_bwd_wave_body(const Cp9Dev &h, const Cp9WinDev *wins, int wi, int do_scan,
                                          int align, float *best_out) {
  const WaveLds s = wave_lds(h, h.nsb, h.bel_k, h.bel_slot, h.nbel, h.binfo);
  const int t = threadIdx.x, M = h.M, M1 = M + 1, nz = h.nz;
  const Cp9WinDev w = wins[wi];
  const G g = gtabs(h, w);
  const int i0 = w.i0, j0 = w.j0, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * M1;
  IXG_GLOBAL int32_t *gm = gmem(w.mx) + 3 * cells, *gi = gm + cells, *gd = gi + cells;
  const bool el = h.el != 0;
  __syncthreads();
  if (t == 0) {  // row L, position j0 (cp9.rs:2029-2055); its local-end row for strip 0's lane 0
    const int dj0 = g.dsq[j0];
    int32_t *e = s.elb;
    IXG_GLOBAL int32_t *m = gm + (size_t)L * M1, *i = gi + (size_t)L * M1, *d = gd + (size_t)L * M1;
    for (int k = 1; k <= M; k++) e[k] = NINF;
    if (el)
      for (int c = 0; c < g.el_ct[M + 1]; c++) e[g.el_idx[g.el_off[M + 1] + c]] = 0;
    e[0] = NINF;
    int32_t *mb = s.mb;  // row L's match cells for strip 0's lane 0
    m[M] = mb[M] = wa(LW(wa(e[M], TS(M, O_MEL)), TS(M, O_ME)), MSC(M, dj0));
    i[M] = wa(TS(M, O_IM), ISC(M, dj0));
    int32_t dn = TS(M, O_DM);
    d[M] = dn;
    for (int k = M - 1; k >= 1; k--) {
      int32_t v = TS(k, O_ME);
      v = LW(v, wa(dn, TS(k, O_MD)));
      if (el) v = LW(v, wa(e[k], TS(k, O_MEL)));
      m[k] = mb[k] = wa(v, MSC(k, dj0));
      i[k] = wa(wa(dn, TS(k, O_ID)), ISC(k, dj0));
      dn = wa(dn, TS(k, O_DD));
      d[k] = dn;
    }
    m[0] = mb[0] = wa(dn, TS(0, O_MD));
    i[0] = wa(wa(dn, TS(0, O_ID)), ISC(0, dj0));
    d[0] = NINF;
  }
  __syncthreads();
  float best = -1e36f;
  // rows L-1 .. 1 (positions j0-1 .. i0, cp9.rs:2057-2169), then row 0
  // (position i0 - 1, cp9.rs:2171-2200), 64 to a strip
  for (int s0 = 0; s0 < L; s0 += 64) {
    const int cur = L - 1 - s0 - t, par = (s0 >> 6) & 1;
    const bool live = cur >= 0;
    const int x = cur > 0 ? g.dsq[i0 + cur - 1] : 0;
    const int up = L - s0;  // lane 0's row after
    const IXG_GLOBAL int32_t *um = gm + (size_t)up * M1, *ui = gi + (size_t)up * M1;
    IXG_GLOBAL int32_t *om = gm + (size_t)(live ? cur : 0) * M1, *oi = gi + (size_t)(live ? cur : 0) * M1,
                       *od = gd + (size_t)(live ? cur : 0) * M1;
    const int32_t *uel = s.elb + par * M1, *umb = s.mb + par * M1;
    int32_t *nel = s.elb + (par ^ 1) * M1, *nmb = s.mb + (par ^ 1) * M1;
    int32_t oM = NINF, oI = NINF, oD = NINF, oE = NINF;  // this lane's cells at its previous node (k + 1)
    int32_t lM = NINF;                                   // the row after's match cell at k + 1
    int32_t mM = NINF, b = NINF;                         // the row after's M(M); the begin sum
    NodeW nx = node_w(g, min(M + t, M), x);              // the next node
    // lane 0: the row after's cells at its next node, read a step ahead
    // (every lane loads, at a clamped node, as in the Forward)
    int32_t nM = um[M], nI = ui[M];
    IXG_GLOBAL int32_t *sink = gmem(w.mx) + 6 * cells + t;
    for (int step = 0; step <= M + 63; step++) {
      const int k = M - step + t;
      const bool act = live && k >= 0 && k <= M;
      const NodeW cn = nx;
      nx = node_w(g, min(max(k - 1, 0), M), x);
      const int32_t gM = nM, gI = nI, gE = (t == 0 && act) ? uel[k] : NINF;
      {
        const int kn = min(max(k - 1, 0), M);
        nM = um[kn];
        nI = ui[kn];
      }
      int32_t cM = NINF, cI = NINF, cD = NINF;
      const int32_t uM = from_prev_lane(gM, oM), uI = from_prev_lane(gI, oI), uE = from_prev_lane(gE, oE);
      if (act) {
        int32_t cE = NINF;
        if (k == M) mM = uM;
        // local-end cell (cp9.rs:2068-2119): its end nodes' next-row match
        // cells in decreasing order, the scan reset, its self loop
        const uint32_t i1w = s.info[M + 2 + k];
        const bool hel = i1w >> 16 & 1;
        if (cur > 0 && k >= 1) {
          int32_t v = (k == M && el && hel) ? wa(NINF, h.el_selfsc) : NINF;
          const int nM = i1w >> 18;
          for (int c = 0; c < nM; c++) v = LW(v, mM);
          if (do_scan && (i1w >> 17 & 1)) v = 0;
          const uint32_t i0w = s.info[k];
          const int o = i0w >> 16, n = i0w & 0xFFFF;
          for (int c = 0; c < n; c++) v = LW(v, t == 0 ? umb[s.lnode[o + c]] : s.slot[s.lslot[o + c] * 65 + t]);
          if (el && k < M && hel) v = LW(v, wa(uE, h.el_selfsc));
          cE = v;
        }
        if (k == M) {
          cD = wa(uI, N_DI(cn));
          if (do_scan) cD = LW(cD, N_DM(cn));
          if (cur > 0) {
            cM = wa(wa(uI, N_MI(cn)), cn.ms);
            if (el && hel) cM = LW(cM, wa(wa(NINF, h.el_selfsc), N_MEL(cn)));
            cI = wa(wa(uI, N_II(cn)), cn.is);
            if (do_scan) {
              cM = LW(cM, LW(wa(cE, N_MEL(cn)), N_ME(cn)));
              cI = LW(cI, N_IM(cn));
            }
          }
        } else if (k >= 1) {
          cD = LW(LW(wa(lM, N_DM(cn)), wa(uI, N_DI(cn))), wa(oD, N_DD(cn)));
          if (cur > 0) {
            int32_t mv = LW(LW(wa(lM, N_MM(cn)), wa(uI, N_MI(cn))), wa(oD, N_MD(cn)));
            if (el && hel) mv = LW(mv, wa(cE, N_MEL(cn)));
            cM = wa(mv, cn.ms);
            cI = wa(LW(LW(wa(lM, N_IM(cn)), wa(uI, N_II(cn))), wa(oD, N_ID(cn))), cn.is);
            if (do_scan) cM = LW(cM, N_ME(cn));
          }
        } else {  // k = 0 (cp9.rs:2147-2167): I(0), D(0), the begin cell
          cD = NINF;
          if (cur > 0)
            cI = wa(LW(LW(wa(lM, N_IM(cn)), wa(uI, N_II(cn))), wa(oD, N_ID(cn))), cn.is);
          b = LW(b, wa(uI, N_MI(cn)));
          b = LW(b, wa(oD, N_MD(cn)));
          cM = b;
          if (scorify(b) > best) best = scorify(b);
          if (align && cur == 0) best_out[w.out] = scorify(b);  // doing_align: the whole window's score
        }
        if (k >= 1) b = LW(b, wa(uM, N_BM(cn)));
        const int sl = (int16_t)(i1w & 0xFFFF);
        if (sl >= 0) s.slot[sl * 65 + t + 1] = cM;
        if (t == 63) {
          nel[k] = cE;
          nmb[k] = cM;
        }
        oM = cM;
        oI = cI;
        oD = cD;
        oE = cE;
        lM = uM;
      }
      *(act ? om + k : sink) = cM;
      *(act ? oi + k : sink) = cI;
      *(act ? od + k : sink) = cD;
    }
    __syncthreads();
  }
  for (int off = 32; off >= 1; off >>= 1) best = fmaxf(best, __shfl_xor(best, off));
  if (!align && t == 0) best_out[w.out] = best;
}

__global__ __launch_bounds__(64) void cp9_bwd_wave(Cp9Dev h, const Cp9WinDev *wins, int do_scan, int align,
                                              float *best_out) {
  cp9_bwd_wave_body(h, wins, blockIdx.x, do_scan, align, best_out);
}

#undef LW
#undef N_MM
#undef N_IM
#undef N_DM
#undef N_BM
#undef N_MI
#undef N_II
#undef N_DI
#undef N_MD
#undef N_ID
#undef N_DD
#undef N_ME
#undef N_MEL

#undef TS
#undef MSC
#undef ISC
#undef LS

}  // namespace

// The Forward and Backward wavefront kernels in one launch: blocks [0, n)
// run the Forward of window b, blocks [n, 2n) the Backward of window b - n.
// The two write disjoint matrices (the Forward's three planes, the
// Backward's three) and each its own score array, so they need no ordering.
__global__ __launch_bounds__(64) void cp9_fb_wave(Cp9Dev h, const Cp9WinDev *wins, int n, int fscan, int bscan,
                                                 int align, float *fsc, float *bsc) {
  if ((int)blockIdx.x < n) cp9_fwd_wave_body(h, wins, (int)blockIdx.x, fscan, align, fsc);
  else cp9_bwd_wave_body(h, wins, (int)blockIdx.x - n, bscan, align, bsc);
}

hipError_t launch_cp9_bands(const Cp9Dev &h, const Cp9WinDev *wins, int64_t n, int mode, int32_t thresh,
                            float *fsc, float *bsc, int32_t *pn, int32_t *occ, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)n), block(64);
  const size_t lds = cp9_lds_bytes(h.M, h.nz);
  const int fscan = mode & IXG_CP9_FWD_SCAN ? 1 : 0, bscan = mode & IXG_CP9_BCK_SCAN ? 1 : 0;
  const int align = mode & IXG_CP9_ALIGN ? 1 : 0;
  // the wavefront forms when their local-end slots fit in LDS
  const size_t lf = cp9_wave_lds_bytes(h.M, h.nz, h.nsf, h.nel), lb = cp9_wave_lds_bytes(h.M, h.nz, h.nsb, h.nbel);
  const bool fw = (h.wave & 1) && lf <= CP9_LDS_MAX, bw = (h.wave & 2) && lb <= CP9_LDS_MAX;
  if (fw && bw && n <= (int64_t)1 << 30) {  // both directions of every window in one launch
    hipLaunchKernelGGL(cp9_fb_wave, dim3((unsigned)(2 * n)), block, std::max(lf, lb), s, h, wins, (int)n, fscan,
                       bscan, align, fsc, bsc);
  } else {
    if (fw) hipLaunchKernelGGL(cp9_fwd_wave, grid, block, lf, s, h, wins, fscan, align, fsc);
    else hipLaunchKernelGGL(cp9_fwd_kernel, grid, block, lds, s, h, wins, fscan, align, fsc);
    if (bw) hipLaunchKernelGGL(cp9_bwd_wave, grid, block, lb, s, h, wins, bscan, align, bsc);
    else hipLaunchKernelGGL(cp9_bwd_kernel, grid, block, lds, s, h, wins, bscan, align, bsc);
  }
  hipLaunchKernelGGL(cp9_post_kernel, grid, block, lds, s, h, wins, fscan, thresh, pn, occ);
  return hipGetLastError();
}

}  // namespace ixg

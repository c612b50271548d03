// Work schedule shared by the HMM-banded scanners: per-state band scalars
// staged in LDS, and the level-by-level distribution of a level's work units
// (cells, IL rows, IR diagonals of all its states, concatenated) round-robin
// over the workgroup's lanes.
#pragma once
#include "cm_internal.h"

namespace ixg {

// Copies jmin, jmax, hd_off, units and ibase of the M states into LDS (when
// M <= HB_LDS_STATES; the launch then provides 20 * M bytes) and points the
// envelope at the copies.
__device__ __forceinline__ void hb_stage_scalars(EnvDev &e, int M, int32_t *smem) {
  if (M > HB_LDS_STATES) return;
  int32_t *a = smem, *b = smem + M, *c = smem + 2 * M, *d = smem + 3 * M, *f = smem + 4 * M;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    a[i] = e.jmin[i];
    b[i] = e.jmax[i];
    c[i] = e.hd_off[i];
    d[i] = e.units[i];
    f[i] = e.ibase[i];
  }
  __syncthreads();
  e.jmin = a;
  e.jmax = b;
  e.hd_off = c;
  e.units = d;
  e.ibase = f;
}

inline size_t hb_lds_bytes(int M) { return M <= HB_LDS_STATES ? (size_t)20 * M : 0; }

// Calls f(v, u) for this lane's share of level l's work units: the units of
// the level's states in list order, dealt round-robin over the lanes.
template <class F>
__device__ __forceinline__ void hb_level_units(const CmDev &cm, const EnvDev &e, int l, F &&f) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int off = 0;  // units dealt so far, modulo nt
  for (int q = cm.lvl_off[l]; q < cm.lvl_off[l + 1]; q++) {
    const int v = cm.lvl_states[q];
    const int n = e.units[v];
    if (n <= 0) continue;
    int r = (tid - off) % nt;
    if (r < 0) r += nt;
    for (int u = r; u < n; u += nt) f(v, u);
    off = (int)(((int64_t)off + n) % nt);
  }
}

// Row jy of state y: cell d at base + d for lo <= d <= hi (lo > hi when jy
// is outside y's j band). The loads are unconditional (index 0 stands in for
// an absent row) so that the rows of several children are fetched together.
struct RowRef {
  int base, lo, hi;
};
__device__ __forceinline__ RowRef row_ref(const EnvDev &e, int y, int jy) {
  const int jmn = e.jmin[y];
  const bool in = jy >= jmn && jy <= e.jmax[y];
  const int h = in ? e.hd_off[y] + (jy - jmn) : 0;
  const int lo = e.hdmin[h], hi = e.hdmax[h], ro = e.roff[h];
  RowRef r;
  r.lo = in ? lo : 1;
  r.hi = in ? hi : 0;
  r.base = ro - lo;
  return r;
}

// Row of the flat cell index tgt among rows [0, last] of one state.
__device__ __forceinline__ int hb_row_of(const int32_t *ro, int last, int tgt) {
  int lo = 0, hi = last;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ro[mid] <= tgt) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace ixg

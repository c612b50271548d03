// Internal declarations shared by the host glue and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/ixg.h"

namespace ixg {

constexpr int KP = IXG_KP;
constexpr int16_t NEGINF_W = -32768;

// A device pointer the compiler cannot trace to a kernel argument (one read
// from a descriptor in memory) is generic: its loads and stores become flat
// instructions, which count against the LDS counter too, so every LDS wait
// after one also waits for the global access. gmem() states that the pointer
// addresses global memory.
#define IXG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ IXG_GLOBAL T *gmem(T *p) {
  return (IXG_GLOBAL T *)p;
}
// The same, handed back as a plain pointer: the compiler still sees where it
// came from and emits global instructions through it, so it can flow into
// code written for plain pointers.
template <class T>
__device__ __forceinline__ T *gplain(T *p) {
  return (T *)(IXG_GLOBAL T *)p;
}

// Host-built p7 filter profiles (natural order + striped packs).
struct MsvHost {
  int m = 0;
  float scale_b = 0.f;
  uint8_t base_b = 0, bias_b = 0, tbm_b = 0, tec_b = 0, tjb_b = 0;
  std::vector<uint8_t> rbv;          // (m+1)*KP
  std::vector<float> prefix, suffix; // m+1
  double mmu = 0, mlambda = 0;
  int max_length = 0;
};
struct VitHost {
  int m = 0;
  float scale_w = 0.f;
  int16_t base_w = 0;
  std::vector<int16_t> rwv;          // (m+1)*KP
  std::vector<int16_t> tmm, tim, tdm, vbm, tmi, tii, tmd, tdd;
  int16_t xw_e_move = 0, xw_e_loop = 0;
  int16_t ddbound_w = 0;
};
struct FwdHost {
  int m = 0, nq = 0;
  std::vector<float> rfv;            // (m+1)*KP
  std::vector<float> tbm, amm, aim, adm, tmi, tii, tmd, tdd;
  std::vector<float> tmain, tddv, rfvv;
  double ftau = 0, flambda = 0;
};
struct BiasHost {
  float eo[KP][2];
  float t1[3];
  float pi[2];
};

// Glocal p7 profile (p7_generic.rs:134-147): tsc [(M+1)][8] in the order
// below, msc / isc [KP][M+1], special states [E N J C][LOOP MOVE], nats.
struct GmHost {
  enum { MM = 0, IM = 1, DM = 2, BM = 3, MD = 4, DD = 5, MI = 6, II = 7 };
  enum { XE = 0, XN = 1, XJ = 2, XC = 3 };
  int M = 0;
  std::vector<float> tsc, msc, isc;
  float xsc[4][2] = {};
  float nj = 0.f;
  int len = 0;
  bool local_end = false;  // exits from every position (esc = 0)
  void set_length(int L);
  void set_length_5p(int L);
  void set_length_3p(int L);
  void set_unihit();
  void set_multihit();
  void xsc8(float out[8]) const;
};
void build_glocal(const ixg_p7_hmm &h, GmHost &g);
// The truncated passes' profiles from the glocal one (p7_generic.rs:259-348,
// as cm_search.rs:1011-1016 configures them at L = 100): [0] 5' (Rgm), [1]
// 3' (Lgm), [2] 5'&3' (Tgm).
void build_trunc_profiles(const ixg_p7_hmm &h, const GmHost &gm, GmHost out[3]);
const std::vector<float> &p7_logsum_table();
float p7_logsum(float a, float b);
void domain_decoding(const GmHost &g, int L, const float *fx, const float *bx, float *btot, float *etot,
                     float *mocc);

void build_msv(const ixg_p7_hmm &h, int max_length, MsvHost &f);
void build_vit(const ixg_p7_hmm &h, VitHost &f);
void build_fwd(const ixg_p7_hmm &h, FwdHost &f);
void build_bias(const float compo[4], int m, BiasHost &b);
int ssv_thresh(const MsvHost &f, double f1);
uint8_t unbiased_byteify(float scale_b, float sc);
int16_t wordify(float scale_w, float sc);
double gumbel_surv(double x, double mu, double lambda);
double exp_surv(double x, double mu, double lambda);
double bg_null_one(int64_t L);

// Device-side views (all pointers are device pointers).
struct SsvDev {
  const uint32_t *cost4;   // (m+1) rounded up to 4, + 4: bytes = rbv[k][A,C,G,U]; 0xFFFFFFFF past m
  const uint8_t *rbv;      // (m+1)*KP
  int m;
  uint8_t xb, bias, base_b, tjb_b, tbm_b;
  float scale_b;
};
struct FwdDev {
  const float *tmain, *tddv, *rfvv;
  const float *tz;  // [4 z][nq][8]: BM MM IM DM MD MI II DD of stripe position q+1+z*nq,
                    // then [4 z][nq rounded up to 8]: DD alone, zero-padded
  int m, nq;
};
struct VitDev {
  const int16_t *rwv;      // (m+1)*KP
  const int16_t *tr;       // 8*(m+1): tmm,tim,tdm,vbm,tmi,tii,tmd,tdd
  int m;
  int16_t base_w, xw_e_move, xw_e_loop;
  float scale_w;
};
struct BiasDev {
  float eo[KP][2];
  float t1[3];
  float pi[2];
};

// One F1 sweep work item: rows [s0, s0+seg) of task t, plus warm-up.
struct SweepItem {
  int32_t task;
  int32_t seg;      // segment index within the task
};
struct TaskDev {
  const uint8_t *dsq;  // 1-indexed view: dsq[1..L]
  int64_t L;
  int64_t mask_off;    // offset (in u64 words) of the task's candidate bitmask
};
struct RawWinDev {
  int64_t n, length, k;
  int64_t row;  // the row that triggered the diagonal
  float score;
  int32_t task;
};
struct WinDev {
  int64_t off;   // seq + off is the window's 1-indexed dsq
  int32_t len;
  int32_t idx;   // index into the output score arrays
  float a, b;    // per-window host-computed terms (length model pieces)
};

// Glocal p7 profile for the F4/F5 envelope definition (p7_generic.rs:134-147),
// device form: tsc [(M+1)][8] (MM IM DM BM MD DD MI II, nats), msc / isc
// [KP][M+1], and the 16,000-entry nats FLogsum table.
constexpr int GFB_TBL = 16000;
struct GmDev {
  const float *tsc, *msc, *isc, *tbl;
  int m;
};
// One glocal Forward / Backward item: the dsq at seq + off (1-indexed, len
// residues), its special-state scores [E N J C][LOOP MOVE], flags bit 0 =
// local ends (esc = 0), bit 1 = keep the Forward matrix, bits 2-3 = the
// profile (0 glocal, 1 5', 2 3', 3 5'&3'; GmDev.tsc holds the four transition
// tables one after another), bit 4 = a Backward item; the specials go to
// xout + xoff [(len+1)][E N J B C], the matrix to mout + moff
// [(len+1)][M | I | D][M+1].
struct GfbItem {
  int64_t off;
  int32_t len, flags;
  float xsc[8];
  int64_t xoff, moff;
};
hipError_t launch_gfb(const GmDev &g, const uint8_t *seq, const GfbItem *items, int64_t n, float *xout, float *mout,
                      float *sc, hipStream_t s);

// Kernel launchers (p7_kernels.hip).
constexpr int SWEEP_SEG = 4096;  // rows per F1 sweep work item (multiple of 64)
constexpr int SWEEP_TILES = 8;   // 64-row tiles a sweep wave carries in registers
hipError_t launch_ssv_sweep(const SsvDev &p, const TaskDev *tasks, const SweepItem *items,
                            int64_t nitems, uint64_t *mask, uint32_t *cinfo, int thresh, hipStream_t s);
hipError_t launch_ssv_resolve(const SsvDev &p, const TaskDev *tasks, const SweepItem *items, int64_t nitems,
                              const uint64_t *mask, const uint32_t *cinfo, int thresh, RawWinDev *out,
                              int64_t cap, unsigned long long *counter, hipStream_t s);
hipError_t launch_fwd(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                      float *sc, hipStream_t s);
constexpr int FWD_ZW_NQ_MAX = 80;  // largest stripe count of the stripe-per-wave Forward kernel
// The F3 Forward's launch shapes, all computing the same scores bit for bit:
// stripe per wave with 64 windows per wave (nq <= 80), 4 lanes per window with
// 16 windows per wave (nq <= 100), and the generic LDS form (any nq whose
// three stripe rows fit in LDS). launch_fwd takes the first that applies.
enum { FWD_V_ZW = 0, FWD_V_EXACT = 1, FWD_V_GENERIC = 2, FWD_V_N = 3 };
bool fwd_variant_ok(const FwdDev &p, int v);
hipError_t launch_fwd_variant(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                              int v, hipStream_t s);
hipError_t launch_fwd_zw(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                         hipStream_t s);
hipError_t launch_fwd_full(const FwdDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                           const int64_t *moff, const int64_t *xoff, float *mat, float *xmx, float *out, int unihit,
                           hipStream_t s);
hipError_t launch_bias(const BiasDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                       float *sc, hipStream_t s);
hipError_t launch_vit(const VitDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                      float *sc, uint8_t *ovf, hipStream_t s);
hipError_t launch_msv(const SsvDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin,
                      float *sc, uint8_t *ovf, hipStream_t s);
hipError_t launch_revcomp(const uint8_t *fwd, uint8_t *rc, int64_t L, hipStream_t s);
hipError_t launch_copy_probe(const void *a, void *b, int64_t n16, hipStream_t s);

}  // namespace ixg

// Device-side closed-shell CCSD context: integral blocks, work buffers, residual, update, DIIS.
#pragma once
#include "common.h"
#include "contract.h"
#include "gemm.h"
#include "kernels.h"
#include "split16.h"
#include <vector>

namespace ccx {

enum Mode : int { MODE_DP = 0, MODE_MIXED = 1, MODE_MIXED_BF16X3 = 2, MODE_MIXED_F16X3 = 3 };

struct Ctx {
  int64_t o = 0, v = 0, nmo = 0;
  int mode = MODE_DP;
  bool have_ints = false;
  Workspace ws;
  std::vector<void*> owned;          // every hipMalloc of the context, freed by destroy()
  int64_t bytes = 0;                 // device bytes held in `owned`

  // ---- integrals (FP64, chemist blocks (pq|rs) stored [p][q][r][s]) ----
  double* eps = nullptr;             // nmo orbital energies (diagonal of the setup Fock)
  double* Coooo = nullptr;           // (mi|nj)
  double* Cooov = nullptr;           // (mi|ne)
  double* Coovv = nullptr;           // (mn|ef)
  double* Covov = nullptr;           // (me|nf)  == K[(me),(nf)]
  double* Ioovv = nullptr;           // <mn|ef> = (me|nf) in Dirac order [m][n][e][f]
  double* Covvv = nullptr;           // (me|af)
  double* Loovv = nullptr;           // L_ijab = 2<ij|ab> - <ij|ba>, [i][j][a][b]
  double* Looov = nullptr;           // Lp[i][m][n][e] = 2(mi|ne) - (ni|me) = L_mnie = L_nmei
  double* J = nullptr;               // J[(me),(nf)] = (mf|ne)
  float* K32 = nullptr;              // FP32 copies for the mixed ring GEMMs
  float* J32 = nullptr;
  float* Io32 = nullptr;             // FP32 <mn|ef> for the mixed hole-hole ladder terms
  // packed ladder matrices: rows (e<=f) / (e<f), columns (a<=b)/(a<b) then (m, b)
  void* LP = nullptr;                // double or float
  void* LM = nullptr;
  // split ladder matrices [N][Kp]: bf16 (MODE_MIXED_BF16X3) or f16 with per-row scales lsP / lsM
  // (MODE_MIXED_F16X3)
  bf16 *LPh = nullptr, *LPl = nullptr, *LMh = nullptr, *LMl = nullptr;
  float *lsP = nullptr, *lsM = nullptr;
  // FP16x3 ring operands (MODE_MIXED_F16X3): split K and J, [ov][ovp] interleaved (lo = hi + 32)
  f16 *Kh = nullptr, *Kl = nullptr, *Jh = nullptr, *Jl = nullptr;
  f16 *Ioh = nullptr, *Iol = nullptr; // split <mn|ef> [(mn)][(ef) padded] (hole-hole term :603)
  // split Covvv: CvA[(m,e,a)][f] = (me|af) and CvT[(m,b,e)][f] = (mf|be), f padded to v32
  f16 *CvAh = nullptr, *CvAl = nullptr, *CvTh = nullptr, *CvTl = nullptr;
  float* xs = nullptr;               // operand scales: [0] ring amplitudes, [1] J, [2] K, [3] W,
                                     // [4] packed tau, [5] tau, [6] <mn|ef>, [7] Wmnij,
                                     // [8] Covvv, [9] t1, [10] Ud, [11] tau~, [12] L,
                                     // [13] t2, [14] off-diagonal G (:925)
  float* xpart = nullptr;            // absmax partials (2 x x3_partials())

  // ---- per-residual work buffers (allocated on first use) ----
  bool have_work = false;
  double *Fae = nullptr, *Fmi = nullptr, *Fme = nullptr, *G = nullptr, *Hm = nullptr;
  double *tau = nullptr, *Wmnij = nullptr, *Hh = nullptr, *Zt = nullptr, *Wd = nullptr, *Ud = nullptr;
  void *S1 = nullptr, *A1 = nullptr, *S2 = nullptr, *Tx = nullptr, *W = nullptr;
  void *Tp = nullptr, *Tm = nullptr, *CP = nullptr, *CM = nullptr;
  bf16 *Tph = nullptr, *Tpl = nullptr, *Tmh = nullptr, *Tml = nullptr;   // f16 in MODE_MIXED_F16X3
  float *tau32 = nullptr, *W32 = nullptr;   // mixed modes: FP32 tau and Wmnij (hole-hole terms)
  // FP16x3 ring operands (see ring_prep_x3): BJ [2ov][ovp], BK [ov][ovp], AP [ov][2ovp],
  // AQ [ov][ovp], BP = split W^T [ov][2ovp], all interleaved (lo = hi + 32, rows twice as long)
  f16 *BJh = nullptr, *BJl = nullptr, *BKh = nullptr, *BKl = nullptr, *APh = nullptr, *APl = nullptr;
  f16 *AQh = nullptr, *AQl = nullptr, *BPh = nullptr, *BPl = nullptr;
  // FP16x3 hole-hole operands: tau [(ij)][(ef)], tau^T [(ab)][(mn)], Wmnij^T [(ij)][(mn)]
  f16 *TAh = nullptr, *TAl = nullptr, *TTh = nullptr, *TTl = nullptr, *WTh = nullptr, *WTl = nullptr;
  // FP16x3 <ov|vv> terms: split t1 [o][v32], Ud as [i][(m,e)][v32]; PT[j][(m,x,y)] =
  // sum_z t_jz Covvv[m][x][y][z], P2T[j][(m,b,e)] = sum_f t_jf (mf|be)
  f16 *T1h = nullptr, *T1l = nullptr, *UTh = nullptr, *UTl = nullptr;
  // FP16x3 :925 operands (interleaved, lo = hi + 32): t2 as [(i,j,a)][e], off-diagonal G as [b][e];
  // G's off-diagonal part and diagonal in FP64
  f16 *T2h = nullptr, *T2l = nullptr, *GPh = nullptr, *GPl = nullptr;
  double *Gp = nullptr, *Gd = nullptr;
  float *PT = nullptr, *P2T = nullptr;
  // W0 = the integral part of [Wmbej | Wmbje] ((me|jb), -(mj|be)) in the W layout; split tau~ and L
  // as [(i,j,a)][f padded] for :497 / :533
  float* W0 = nullptr;
  f16 *TFh = nullptr, *TFl = nullptr, *LFh = nullptr, *LFl = nullptr;
  // side stream for the packed ladder in the split-GEMM modes: it depends only on t1, t2, so it runs
  // beside the rest of the residual (fork after the amplitude scales, join before zt_add)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double* red = nullptr;             // reduction partials
  double* scal = nullptr;            // small device scalars

  // ---- optional HIP-event timing of the GEMM phases (ccx_timing_start / ccx_timing_read) ----
  enum Phase : int { PH_LADDER = 0, PH_RING = 1, PH_RESIDUAL = 2, PH_COUNT = 3 };
  struct Span {
    hipEvent_t a, b;
    int phase;
    double flops;
  };
  bool timing = false;
  std::vector<Span> spans;
  std::vector<hipEvent_t> pool;
  hipEvent_t take_event() {
    if (pool.empty()) {
      hipEvent_t e;
      CCX_HIP(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // record a start event; returns its span index (or -1 when timing is off)
  int tbegin(hipStream_t s, int phase, double flops) {
    if (!timing) return -1;
    Span sp{take_event(), take_event(), phase, flops};
    CCX_HIP(hipEventRecord(sp.a, s));
    spans.push_back(sp);
    return (int)spans.size() - 1;
  }
  void tend(hipStream_t s, int k) {
    if (k >= 0) CCX_HIP(hipEventRecord(spans[k].b, s));
  }

  // ---- HBAR (cchbar) and Lambda work buffers (allocated on first use, FP64) ----
  bool have_hbar_bufs = false, have_hbar = false, hbar_stored = false;
  double *Ht1 = nullptr, *Ht2 = nullptr;               // amplitudes the HBAR was built from
  double *Hov = nullptr, *Hvv = nullptr, *Hoo = nullptr, *Hoooo = nullptr, *Hooov = nullptr;
  double *Hovvo = nullptr, *Hovov = nullptr, *Hovoo = nullptr;
  double *Hvvvv = nullptr, *Hvovv = nullptr, *Hvvvo = nullptr;   // stored HBAR only
  double *Ltmp = nullptr, *Goo = nullptr, *Gvv = nullptr, *Lh = nullptr, *Lsym = nullptr;
  double *LY = nullptr, *LVf = nullptr, *LVg = nullptr, *LU = nullptr, *LQ = nullptr, *Lsm = nullptr;
  double* Jsm = nullptr;
  double* Jv3 = nullptr;   // density energy: two v^3 slices and the term values

  // ---- DIIS (pycc utils.helper_diis) ----
  int max_diis = 0, cap = 0;
  std::vector<int> order;            // active slots, oldest first
  std::vector<std::vector<double>> bdots;   // cached <e_i, e_j> by slot
  double *dv = nullptr, *de = nullptr, *doldt = nullptr, *dcoef = nullptr;

  int64_t ov() const { return o * v; }
  int64_t n1() const { return o * v; }
  int64_t n2() const { return o * o * v * v; }
  int64_t npP() const { return v * (v + 1) / 2; }
  int64_t npM() const { return v * (v - 1) / 2; }
  int64_t noP() const { return o * (o + 1) / 2; }
  int64_t noM() const { return o * (o - 1) / 2; }
  int64_t ovp() const { return (o * v + 31) / 32 * 32; }   // K padding of the x3 ring operands
  int64_t o2p() const { return (o * o + 31) / 32 * 32; }
  int64_t v2p() const { return (v * v + 31) / 32 * 32; }
  int64_t v32() const { return (v + 31) / 32 * 32; }
  bool split_ladder() const { return mode == MODE_MIXED_BF16X3 || mode == MODE_MIXED_F16X3; }

  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    CCX_HIP(hipMalloc(&p, n * sizeof(T)));
    owned.push_back(p);
    bytes += (int64_t)(n * sizeof(T));
    return (T*)p;
  }
  void destroy();
};

// The packed ladder of the MIXED_F16X3 forward applied to l2 (tau = l2, (ab) columns only, FP32 into
// CP / CM), for the Lambda residual (split16.hip).
void ladder_x3_l2(Ctx& c, const double* l2, hipStream_t s);

void set_integrals_dense(Ctx& c, const double* fock, const double* eri, hipStream_t s);
void set_integrals_df(Ctx& c, const double* fock, const double* B, int64_t naux, hipStream_t s);
void residual(Ctx& c, const double* fock, const double* t1, const double* t2, double* r1, double* r2,
              hipStream_t s);
void alloc_work(Ctx& c);
// HBAR at (F, t1, t2): stored = all 11 blocks, else the o^2 v^2-sized ones (hbar.hip).
void hbar_build(Ctx& c, const double* F, const double* t1, const double* t2, bool stored, hipStream_t s);
void hbar_block(Ctx& c, int block, double* out, hipStream_t s);
// r_L1 and r2 = alpha * r_L2 at (l1, l2) on the current HBAR; without the source terms it is the
// vector-Jacobian product of the CCSD residual.
void lambda_residual(Ctx& c, const double* l1, const double* l2, double* r1, double* r2, bool with_source,
                     double alpha, hipStream_t s);
// out[i][j][a][b] = alpha * (h[i][j][a][b] + h[j][i][b][a])
void pair_symmetrize(int64_t o, int64_t v, double alpha, const double* h, double* out, hipStream_t s);
void pseudoenergy(Ctx& c, const double* l2, double* e_dev, hipStream_t s);
// dL/dF (full [nmo][nmo] D) and the two-particle densities of the Lagrangian at (t, l)
void densities(Ctx& c, const double* t1, const double* t2, const double* l1, const double* l2, double* D,
               double* Doooo, double* Dooov, double* Dovov, double* Doovv, double* Dvvvo, double* Dvvvv,
               hipStream_t s);
// ccdensity.compute_energy from the densities (Dvvvv term through the packed ladder of l2)
void density_energy(Ctx& c, const double* t1, const double* t2, const double* l2, const double* F,
                    const double* D, const double* Doooo, const double* Dooov, const double* Dovov,
                    const double* Doovv, const double* Dvvvo, double* e_dev, hipStream_t s);
// j1, j2 = the CCSD Jacobian applied to (x1, x2) on the current HBAR (x2 pair-symmetric)
void jacobian(Ctx& c, const double* x1, const double* x2, double* j1, double* j2, hipStream_t s);
// the closed-shell (T) correction at (F, t1, t2) over the triples i >= j >= k with i0 <= i < i1, added to
// *e_dev (accumulate) or written there (triples.hip); synchronises the stream
void triples_energy(Ctx& c, const double* F, const double* t1, const double* t2, int i0, int i1, bool accumulate,
                    double* e_dev, hipStream_t s);
// t3[x][y][z] (n^3, contiguous) divided by the orbital-energy denominator of the fixed triple (p, q, r) and the
// running orbitals e0 + x, e0 + y, e0 + z (occ_fixed: D = f_pp + f_qq + f_rr - f_xx - f_yy - f_zz, else -D)
void t3_denominator(const double* F, int64_t nmo, int p, int q, int r, int e0, int n, bool occ_fixed, double* t3,
                    hipStream_t s);
void timing_start(Ctx& c);
void timing_read(Ctx& c, double* out, int n);
void diis_reset(Ctx& c, const double* t1, const double* t2, int max_diis, int cap, hipStream_t s);
void diis_add(Ctx& c, const double* t1, const double* t2, hipStream_t s);
void diis_extrapolate(Ctx& c, double* t1, double* t2, double* coef_out, hipStream_t s);

}  // namespace ccx

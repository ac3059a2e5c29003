// Closed-shell CCSD residual on gfx950: integral setup, residual orchestration, DIIS.
//
// Reference algorithm: pycc/ccwfn.py (residuals :321-372, builders :432-761, r_T2 :764-791 and
// :886-944, energy :1122-1162) and pycc/utils.py helper_diis :257-361. Each step below cites the
// reference line it reproduces. Differences in form (never in value):
//   * integrals are kept as chemist blocks (pq|rs); the Dirac <pq|rs> = (pr|qs) of each reference
//     term is written next to it;
//   * the ladder  sum_ef tau_ijef <ab|ef>  (:931) and Z_mbij (:715) are one GEMM over the
//     (anti)symmetrised pairs e<=f, i<=j, a<=b (tau+/- packing), a quarter of the flops;
//   * the ring terms (:644-645, :683, :933-935) are ov x ov GEMMs on [i][a][j][b] operands;
//   * terms whose (ij)(ab) pair-swap gives the same symmetrised r2 (:790) are taken in the
//     form whose operands need no transposition (noted where used).
#include "ccsd.h"
#include <cmath>
#include <cstring>
#include <type_traits>

namespace ccx {

void Ctx::destroy() {
  for (auto& sp : spans) {
    (void)hipEventDestroy(sp.a);
    (void)hipEventDestroy(sp.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
  spans.clear();
  pool.clear();
  for (void* p : owned) (void)hipFree(p);
  owned.clear();
  ws.release();
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  ev_fork = ev_join = nullptr;
  side = nullptr;
}

// ---------------------------------------------------------------------------------------------
// Chemist block (pq|rs) = sum_Q B_Qpq B_Qrs as one GEMM over the auxiliary index.
static void df_block(Ctx& c, const double* B, int64_t naux, int64_t p0, int64_t np, int64_t q0,
                     int64_t nq, int64_t r0, int64_t nr, int64_t s0, int64_t ns, double* out,
                     hipStream_t s) {
  const int64_t nmo = c.nmo;
  GemmParams<double> g;
  g.M = np * nq;
  g.N = nr * ns;
  g.K = naux;
  g.A = B + p0 * nmo + q0;
  g.am = Idx2{nq, nmo, 1};
  g.ak = Idx2{naux, 0, nmo * nmo};
  g.B = B + r0 * nmo + s0;
  g.bk = Idx2{naux, 0, nmo * nmo};
  g.bn = Idx2{ns, nmo, 1};
  g.C = out;
  g.cm = Idx2{nq, nq * nr * ns, nr * ns};
  g.cn = Idx2{ns, ns, 1};
  g.alpha = 1.0;
  g.beta = 0.0;
  tc_gemm<double>(g, c.ws, s);
}

static void alloc_integrals(Ctx& c) {
  const int64_t o = c.o, v = c.v, ov = c.ov();
  c.eps = c.alloc<double>(c.nmo);
  c.Coooo = c.alloc<double>(o * o * o * o);
  c.Cooov = c.alloc<double>(o * o * o * v);
  c.Coovv = c.alloc<double>(o * o * v * v);
  c.Covov = c.alloc<double>(ov * ov);
  c.Ioovv = c.alloc<double>(ov * ov);
  c.Covvv = c.alloc<double>(ov * v * v);
  c.Loovv = c.alloc<double>(ov * ov);
  c.Looov = c.alloc<double>(o * o * o * v);
  c.J = c.alloc<double>(ov * ov);
  const int64_t szP = c.npP() * (c.npP() + ov), szM = c.npM() * (c.npM() + ov);
  if (c.mode == MODE_DP) {
    c.LP = c.alloc<double>(szP);
    c.LM = c.alloc<double>(szM);
  } else {
    if (c.mode == MODE_MIXED_F16X3) {
      const int64_t ovp = c.ovp();
      for (f16** q : {&c.Kh, &c.Jh}) {   // interleaved split copies: lo = hi + 32
        *q = c.alloc<f16>(2 * ov * ovp);
        CCX_HIP(hipMemset(*q, 0, 2 * ov * ovp * sizeof(f16)));   // K padding
      }
      c.Kl = c.Kh + 32;
      c.Jl = c.Jh + 32;
      for (f16** q : {&c.Ioh, &c.Iol}) {
        *q = c.alloc<f16>(o * o * c.v2p());
        CCX_HIP(hipMemset(*q, 0, o * o * c.v2p() * sizeof(f16)));   // K padding
      }
      c.xs = c.alloc<float>(64);
      c.xpart = c.alloc<float>(2 * x3_partials());
    } else {
      c.K32 = c.alloc<float>(ov * ov);
      c.J32 = c.alloc<float>(ov * ov);
      c.Io32 = c.alloc<float>(ov * ov);
    }
    if (c.mode == MODE_MIXED) {
      c.LP = c.alloc<float>(szP);
      c.LM = c.alloc<float>(szM);
    } else {
      const int64_t bP = (c.npP() + ov) * bf16_kpad(c.npP()), bM = (c.npM() + ov) * bf16_kpad(c.npM());
      if (c.mode == MODE_MIXED_F16X3) {
        // interleaved split rows (hi and lo of each 32-wide k-block side by side, lo = hi + 32)
        // (at least 64 elements, so that lo = hi + 32 stays inside an empty one)
        c.LPh = c.alloc<bf16>(std::max<int64_t>(2 * bP, 64));
        c.LMh = c.alloc<bf16>(std::max<int64_t>(2 * bM, 64));
        CCX_HIP(hipMemset(c.LPh, 0, 2 * bP * sizeof(bf16)));   // K padding
        if (bM) CCX_HIP(hipMemset(c.LMh, 0, 2 * bM * sizeof(bf16)));
        c.LPl = c.LPh + 32;
        c.LMl = c.LMh + 32;
      } else {
        c.LPh = c.alloc<bf16>(bP);
        c.LPl = c.alloc<bf16>(bP);
        c.LMh = c.alloc<bf16>(bM);
        c.LMl = c.alloc<bf16>(bM);
        for (bf16* q : {c.LPh, c.LPl}) CCX_HIP(hipMemset(q, 0, bP * sizeof(bf16)));   // K padding
        for (bf16* q : {c.LMh, c.LMl}) CCX_HIP(hipMemset(q, 0, bM * sizeof(bf16)));
      }
      if (c.mode == MODE_MIXED_F16X3) {
        c.lsP = c.alloc<float>(c.npP() + ov);
        c.lsM = c.alloc<float>(c.npM() + ov);
      }
    }
  }
  c.red = c.alloc<double>(1024 * 64);
  c.scal = c.alloc<double>(128);   // [0, 64): per-slot overlaps, [64, 72): diis_push overlaps
}

static void pack_ladder(Ctx& c, int64_t e0, int64_t e1, VVVVSrc src, hipStream_t s) {
  if (c.mode == MODE_DP)
    ladder_pack<double>(c.o, c.v, e0, e1, src, c.Covvv, (double*)c.LP, (double*)c.LM, s);
  else if (c.mode == MODE_MIXED)
    ladder_pack<float>(c.o, c.v, e0, e1, src, c.Covvv, (float*)c.LP, (float*)c.LM, s);
  else if (c.mode == MODE_MIXED_F16X3)
    ladder_pack_x3(c.v, e0, e1, src, (f16*)c.LPh, (f16*)c.LPl, (f16*)c.LMh, (f16*)c.LMl, c.lsP, c.lsM, s);
  else
    ladder_pack_bf16x3(c.o, c.v, e0, e1, src, c.LPh, c.LPl, c.LMh, c.LMl, s);
}

// Derived blocks shared by both integral sources.
static void finish_integrals(Ctx& c, const double* fock, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov(), nmo = c.nmo;
  // eps = diag(F) (hamiltonian.py:63), copied with a strided view
  axpby<double>(1.0, tvs(fock, "p", {nmo}, {nmo + 1}), 0.0, tv(c.eps, "p", {nmo}), s);
  // <ij|ab> = (ia|jb) in Dirac order
  axpby<double>(1.0, tv(c.Covov, "iajb", {o, v, o, v}), 0.0, tv(c.Ioovv, "ijab", {o, o, v, v}), s);
  // L_ijab = 2<ij|ab> - <ij|ba> = 2(ia|jb) - (ib|ja)   (hamiltonian.py:70)
  axpby<double>(2.0, tv(c.Covov, "iajb", {o, v, o, v}), 0.0, tv(c.Loovv, "ijab", {o, o, v, v}), s);
  axpby<double>(-1.0, tv(c.Covov, "ibja", {o, v, o, v}), 1.0, tv(c.Loovv, "ijab", {o, o, v, v}), s);
  // Lp[i][m][n][e] = 2(mi|ne) - (ni|me): L_mnie (Fmi, ccwfn.py:532) and L_nmei (r_T1, :760)
  axpby<double>(2.0, tv(c.Cooov, "mine", {o, o, o, v}), 0.0, tv(c.Looov, "imne", {o, o, o, v}), s);
  axpby<double>(-1.0, tv(c.Cooov, "nime", {o, o, o, v}), 1.0, tv(c.Looov, "imne", {o, o, o, v}), s);
  // J[(me),(nf)] = (mf|ne)
  axpby<double>(1.0, tv(c.Covov, "mfne", {o, v, o, v}), 0.0, tv(c.J, "menf", {o, v, o, v}), s);
  if (c.mode == MODE_MIXED_F16X3) {
    // split K and J once, each with its own power-of-two scale (xs[2], xs[1])
    const int64_t ovp = c.ovp();
    x3_absmax<double>(c.Covov, ov * ov, c.xpart, s);
    x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 2, s);
    x3_split_rows<double>(c.Covov, ov, ov, ov, c.xs + 2, c.Kh, c.Kl, ovp, s, true);
    x3_absmax<double>(c.J, ov * ov, c.xpart, s);
    x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 1, s);
    x3_split_rows<double>(c.J, ov, ov, ov, c.xs + 1, c.Jh, c.Jl, ovp, s, true);
    // <mn|ef> as [(mn)][(ef)] for the hole-hole term :603
    x3_absmax<double>(c.Ioovv, ov * ov, c.xpart, s);
    x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 6, s);
    x3_split_rows<double>(c.Ioovv, o * o, v * v, v * v, c.xs + 6, c.Ioh, c.Iol, c.v2p(), s);
    covvv_x3_setup(c, s);
  } else if (c.mode != MODE_DP) {
    cast<double, float>(ov * ov, c.Covov, c.K32, s);
    cast<double, float>(ov * ov, c.J, c.J32, s);
    cast<double, float>(ov * ov, c.Ioovv, c.Io32, s);
  }
  if (c.mode == MODE_MIXED_BF16X3) ladder_pack_o_bf16x3(o, v, c.Covvv, c.LPh, c.LPl, c.LMh, c.LMl, s);
  if (c.mode == MODE_MIXED_F16X3)
    ladder_pack_o_x3(o, v, c.Covvv, (f16*)c.LPh, (f16*)c.LPl, (f16*)c.LMh, (f16*)c.LMl, c.lsP, c.lsM, s);
  c.have_ints = true;
}

void set_integrals_dense(Ctx& c, const double* fock, const double* eri, hipStream_t s) {
  CCX_CHECK(fock && eri, ERR_ARG, "null integral pointer");
  CCX_CHECK(!c.have_ints, ERR_STATE, "integrals already set");
  const int64_t o = c.o, v = c.v, nmo = c.nmo;
  alloc_integrals(c);
  gather_chem(eri, nmo, 0, o, 0, o, 0, o, 0, o, c.Coooo, s);
  gather_chem(eri, nmo, 0, o, 0, o, 0, o, o, v, c.Cooov, s);
  gather_chem(eri, nmo, 0, o, 0, o, o, v, o, v, c.Coovv, s);
  gather_chem(eri, nmo, 0, o, o, v, 0, o, o, v, c.Covov, s);
  gather_chem(eri, nmo, 0, o, o, v, o, v, o, v, c.Covvv, s);
  // (ea|fb) = <ef|ab> = ERI[o+e][o+f][o+a][o+b]
  const int64_t n3 = nmo * nmo * nmo, n2 = nmo * nmo;
  VVVVSrc src{eri + o * (n3 + n2 + nmo + 1), n3, nmo, n2, 1};
  pack_ladder(c, 0, v, src, s);
  finish_integrals(c, fock, s);
}

void set_integrals_df(Ctx& c, const double* fock, const double* B, int64_t naux, hipStream_t s) {
  CCX_CHECK(fock && B && naux > 0, ERR_ARG, "bad DF integral arguments");
  CCX_CHECK(!c.have_ints, ERR_STATE, "integrals already set");
  const int64_t o = c.o, v = c.v;
  alloc_integrals(c);
  df_block(c, B, naux, 0, o, 0, o, 0, o, 0, o, c.Coooo, s);
  df_block(c, B, naux, 0, o, 0, o, 0, o, o, v, c.Cooov, s);
  df_block(c, B, naux, 0, o, 0, o, o, v, o, v, c.Coovv, s);
  df_block(c, B, naux, 0, o, o, v, 0, o, o, v, c.Covov, s);
  df_block(c, B, naux, 0, o, o, v, o, v, o, v, c.Covvv, s);
  // (vv|vv) in chunks of e: X[e - e0][a][f][b] = (ea|fb), then packed
  const int64_t v3 = v * v * v;
  int64_t ne = std::max<int64_t>(1, std::min<int64_t>(v, (int64_t)(4.0e9 / (8.0 * (double)v3))));
  double* X = nullptr;
  CCX_HIP(hipMalloc(&X, (size_t)ne * v3 * sizeof(double)));
  try {
    for (int64_t e0 = 0; e0 < v; e0 += ne) {
      int64_t e1 = std::min(v, e0 + ne);
      df_block(c, B, naux, o + e0, e1 - e0, o, v, o, v, o, v, X, s);
      VVVVSrc src{X - e0 * v3, v3, v * v, v, 1};
      pack_ladder(c, e0, e1, src, s);
    }
    CCX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(X);
    throw;
  }
  CCX_HIP(hipFree(X));
  finish_integrals(c, fock, s);
}

// ---------------------------------------------------------------------------------------------
void alloc_work(Ctx& c) {
  if (c.have_work) return;
  const int64_t o = c.o, v = c.v, ov = c.ov();
  c.Fae = c.alloc<double>(v * v);
  c.Fmi = c.alloc<double>(o * o);
  c.Fme = c.alloc<double>(ov);
  c.G = c.alloc<double>(v * v);
  c.Hm = c.alloc<double>(o * o);
  c.tau = c.alloc<double>(ov * ov);
  c.Wmnij = c.alloc<double>(o * o * o * o);
  c.Hh = c.alloc<double>(ov * ov);
  c.Zt = c.alloc<double>(ov * o * o);
  if (c.mode != MODE_MIXED_F16X3) c.Wd = c.alloc<double>(2 * ov * ov);   // FP16x3 builds W directly
  c.Ud = c.alloc<double>(ov * ov);
  const size_t es = c.mode == MODE_DP ? sizeof(double) : sizeof(float);
  auto talloc = [&](size_t n) -> void* { return es == 8 ? (void*)c.alloc<double>(n) : (void*)c.alloc<float>(n); };
  if (c.mode == MODE_MIXED_F16X3) {
    // FP32 ring outputs P (A1) and Q (S1); the operands are split FP16
    c.S1 = talloc(ov * ov);
    c.A1 = talloc(ov * ov);
    const int64_t ovp = c.ovp();
    auto halloc = [&](f16** q, int64_t n) {
      *q = c.alloc<f16>(n);
      CCX_HIP(hipMemset(*q, 0, n * sizeof(f16)));   // K padding stays zero
    };
    // the ring operands are interleaved (hi and lo of each 32-wide k-block side by side, lo = hi + 32)
    auto ialloc = [&](f16** h, f16** l, int64_t n) {
      halloc(h, 2 * n);
      *l = *h + 32;
    };
    ialloc(&c.BJh, &c.BJl, 2 * ov * ovp);
    ialloc(&c.BKh, &c.BKl, ov * ovp);
    ialloc(&c.APh, &c.APl, 2 * ov * ovp);
    ialloc(&c.AQh, &c.AQl, ov * ovp);
    ialloc(&c.BPh, &c.BPl, 2 * ov * ovp);
    if (hh_g_x3_fits(o, v)) {   // :925 as an FP16x3 GEMM
      ialloc(&c.T2h, &c.T2l, o * o * v * c.v32());
      ialloc(&c.GPh, &c.GPl, v * c.v32());
      c.Gp = c.alloc<double>(v * v);
      c.Gd = c.alloc<double>(v);
    }
    halloc(&c.TAh, o * o * c.v2p());
    halloc(&c.TAl, o * o * c.v2p());
    halloc(&c.TTh, v * v * c.o2p());
    halloc(&c.TTl, v * v * c.o2p());
    halloc(&c.WTh, o * o * c.o2p());
    halloc(&c.WTl, o * o * c.o2p());
  } else {
    c.S1 = talloc(2 * ov * ov);
    c.A1 = talloc(ov * ov);
    c.S2 = talloc(2 * ov * ov);
    c.Tx = talloc(ov * ov);
  }
  c.W = c.mode == MODE_DP ? (void*)c.Wd : talloc(2 * ov * ov);
  const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
  if (c.split_ladder()) {
    // default priority; ordered against the caller's stream by events only
    CCX_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    CCX_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    CCX_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    const int64_t aP = noP * bf16_kpad(npP), aM = noM * bf16_kpad(npM);
    if (c.mode == MODE_MIXED_F16X3) {   // interleaved, lo = hi + 32
      c.Tph = c.alloc<bf16>(std::max<int64_t>(2 * aP, 64));
      c.Tmh = c.alloc<bf16>(std::max<int64_t>(2 * aM, 64));
      CCX_HIP(hipMemset(c.Tph, 0, 2 * aP * sizeof(bf16)));   // K padding
      if (aM) CCX_HIP(hipMemset(c.Tmh, 0, 2 * aM * sizeof(bf16)));
      c.Tpl = c.Tph + 32;
      c.Tml = c.Tmh + 32;
    } else {
      c.Tph = c.alloc<bf16>(aP);
      c.Tpl = c.alloc<bf16>(aP);
      c.Tmh = c.alloc<bf16>(aM);
      c.Tml = c.alloc<bf16>(aM);
      for (bf16* q : {c.Tph, c.Tpl}) CCX_HIP(hipMemset(q, 0, aP * sizeof(bf16)));   // K padding
      for (bf16* q : {c.Tmh, c.Tml}) CCX_HIP(hipMemset(q, 0, aM * sizeof(bf16)));
    }
  } else {
    c.Tp = talloc(noP * npP);
    c.Tm = talloc(noM * npM);
  }
  if (c.mode != MODE_DP) {
    if (c.mode != MODE_MIXED_F16X3) c.tau32 = c.alloc<float>(ov * ov);
    c.W32 = c.alloc<float>(o * o * o * o);
  }
  c.CP = talloc(noP * (npP + ov));
  c.CM = talloc(noM * (npM + ov));
  c.have_work = true;
}

void ladder_bf16x3(Ctx& c, const double* t1, const double* t2, hipStream_t s);   // ladder_bf16x3.hip
void ladder_x3(Ctx& c, const double* t1, const double* t2, hipStream_t s);       // split16.hip

// The ring and ladder GEMMs run in T (double for DP, float for the mixed modes).
template <typename T>
static void residual_t(Ctx& c, const double* F, const double* t1, const double* t2, double* r1,
                       double* r2, hipStream_t s) {
  const int64_t o = c.o, v = c.v, nmo = c.nmo, ov = c.ov();
  Workspace& ws = c.ws;
  const std::vector<int64_t> OV{o, v}, OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OVVV{o, v, v, v}, OOOO{o, o, o, o};
  // Fock blocks, row-major (nmo x nmo)
  const TV Fov = tvs(F + o, "me", OV, {nmo, 1});
  const TV Fvv = tvs(F + o * nmo + o, "ae", VV, {nmo, 1});
  const TV Foo = tvs(F, "mi", OO, {nmo, 1});
  const TV Fvo = tvs(F + o * nmo, "ai", {v, o}, {nmo, 1});

  auto T1 = [&](const char* l) { return tv(t1, l, OV); };
  auto T2 = [&](const char* l) { return tv(t2, l, OOVV); };
  auto cd = [&](double a, const TV& A, const TV& B, double b, const TV& C) {
    contract<double>(a, A, B, b, C, ws, s);
  };

  const bool x3 = c.mode == MODE_MIXED_F16X3;
  if (x3) x3_amp_scales(c, t1, t2, s);   // every amplitude-derived FP16x3 operand scale
  // ---- ladder + Z: ccwfn.py:931 (1/2 tau_ijef <ab|ef>) and :715 (Z_mbij = <mb|ef> tau_ijef) ----
  // In the split-GEMM modes it needs only t1, t2 (and the scales above): it runs on the side stream
  // beside everything up to zt_add. Its kernels use no shared workspace.
  if (c.split_ladder()) {
    CCX_HIP(hipEventRecord(c.ev_fork, s));
    CCX_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
    if (x3) ladder_x3(c, t1, t2, c.side);
    else ladder_bf16x3(c, t1, t2, c.side);
    CCX_HIP(hipEventRecord(c.ev_join, c.side));
  }
  if (x3) covvv_x3_products(c, t1, s);   // the <ov|vv> products of t1 (Fae, W, Hh terms below)

  // ---- Fme: ccwfn.py:563-564 -------------------------------------------------------------
  const TV Fme = tv(c.Fme, "me", OV);
  axpby<double>(1.0, Fov, 0.0, Fme, s);
  // + t_nf L_mnef: Fme[m][e] += sum_(n,f) t[n][f] Loovv[m][n][e][f]
  wsum(o, v, o, v, 1.0, t1, v, c.Loovv, o * v * v, v, v * v, 1.0, c.Fme, v, s);

  // ---- Fae: ccwfn.py:494-497 (tau~ = t2 + 1/2 t1 t1) ----------------------------------------
  build_tau(o, v, 1.0, 0.5, t1, t2, c.tau, s);
  const TV Fae = tv(c.Fae, "ae", VV);
  axpby<double>(1.0, Fvv, 0.0, Fae, s);
  cd(-0.5, Fov, T1("ma"), 1.0, Fae);
  if (x3) {
    covvv_x3_fae(c, c.Fae, s);
  } else {
    cd(2.0, T1("mf"), tv(c.Covvv, "mfae", OVVV), 1.0, Fae);   // L_mafe = 2(mf|ae) - (me|af)
    cd(-1.0, T1("mf"), tv(c.Covvv, "meaf", OVVV), 1.0, Fae);
  }
  if (!x3) cd(-1.0, tv(c.tau, "mnaf", OOVV), tv(c.Loovv, "mnef", OOVV), 1.0, Fae);   // FP16x3: below

  // ---- Fmi: ccwfn.py:530-533 --------------------------------------------------------------
  const TV Fmi = tv(c.Fmi, "mi", OO);
  axpby<double>(1.0, Foo, 0.0, Fmi, s);
  cd(0.5, T1("ie"), Fov, 1.0, Fmi);
  cd(1.0, T1("ne"), tv(c.Looov, "imne", OOOV), 1.0, Fmi);     // L_mnie = 2(mi|ne) - (ni|me)
  cd(1.0, tv(c.tau, "inef", OOVV), tv(c.Loovv, "mnef", OOVV), 1.0, Fmi);
  if (x3) fae_tauL_x3(c, c.Fae, s);   // :497 as an FP16x3 GEMM (tau~ still in c.tau)

  // ---- Wmnij: ccwfn.py:596-603 (tau = t2 + t1 t1) -------------------------------------------
  build_tau(o, v, 1.0, 1.0, t1, t2, c.tau, s);
  const TV Wmnij = tv(c.Wmnij, "mnij", OOOO);
  axpby<double>(1.0, tv(c.Coooo, "minj", OOOO), 0.0, Wmnij, s);   // <mn|ij> = (mi|nj)
  cd(1.0, T1("je"), tv(c.Cooov, "mine", OOOV), 1.0, Wmnij);        // <mn|ie> = (mi|ne)
  cd(1.0, T1("ie"), tv(c.Cooov, "njme", OOOV), 1.0, Wmnij);        // <mn|ej> = (nj|me)
  if (std::is_same<T, double>::value) {
    cd(1.0, tv(c.tau, "ijef", OOVV), tv(c.Ioovv, "mnef", OOVV), 1.0, Wmnij);  // <mn|ef>
  } else if (c.mode == MODE_MIXED_F16X3) {
    // the two o^4 v^2 hole-hole contractions (:603 and :930) as FP16x3 GEMMs; tau is split once in
    // both orientations ([(ij)][(ef)] here, [(ab)][(mn)] for :930) with one scale
    x3_split_rows<double>(c.tau, o * o, v * v, v * v, c.xs + 5, c.TAh, c.TAl, c.v2p(), s);
    x3_split_t<double>(c.tau, o * o, v * v, 1, v * v, c.xs + 5, c.TTh, c.TTl, 0, c.o2p(), s);
    X3Gemm g;   // Wmnij[(mn)][(ij)] += sum_(ef) <mn|ef> tau_ijef
    g.M = o * o; g.N = o * o; g.Kp = c.v2p(); g.lda = c.v2p(); g.ldb = c.v2p(); g.ldc = o * o;
    g.Ah = c.Ioh; g.Al = c.Iol; g.Bh = c.TAh; g.Bl = c.TAl; g.C64 = c.Wmnij; g.beta = 1.f;
    g.sa = c.xs + 6; g.sb = c.xs + 5;
    x3_gemm<_Float16>(g, s, &ws);   // into Wmnij (FP64), through split K
  } else {
    // mixed modes: the two o^4 v^2 hole-hole ladder contractions (:603 and :930) run in FP32
    cast<double, float>(ov * ov, c.tau, c.tau32, s);
    contract<float>(1.0f, tv(c.tau32, "ijef", OOVV), tv(c.Io32, "mnef", OOVV), 0.0f, tv(c.W32, "mnij", OOOO),
                    ws, s);
    accumulate(o * o * o * o, 1.0, c.W32, c.Wmnij, s);
  }

  // ---- ring operands and the W intermediates: ccwfn.py:641-645 (Wmbej), :680-683 (Wmbje) ----
  // W[(m,e)][x][(j,b)]: x = 0 Wmbej, x = 1 Wmbje (built in FP64 Wd and cast, or in MIXED_F16X3
  // written in FP32 by one kernel).
  T* S1 = (T*)c.S1;
  T* A1 = (T*)c.A1;
  T* S2 = (T*)c.S2;
  T* Tx = (T*)c.Tx;
  T* W = (T*)c.W;
  if (x3) {
    // every ring operand is bounded by 3 max|t2| + max|t1|^2 (scale xs[0])
    ring_prep_x3(o, v, c.ovp(), t1, t2, c.xs, c.BJh, c.BJl, c.BKh, c.BKl, c.APh, c.APl, c.AQh, c.AQl, c.Ud, s);
  } else {
    ring_prep<T>(o, v, t1, t2, S1, A1, S2, Tx, c.Ud, s);
  }
  // W[(me)][x][(jb)]: x = 0 holds Wmbej, x = 1 holds Wmbje
  const std::vector<int64_t> WSTR{v * 2 * ov, 2 * ov, v, 1};
  if (x3) {
    w_build_x3(c, t1, (float*)W, s);   // :641-643, :680-682 in one pass, straight to FP32
  } else {
    const TV Wej = tvs(c.Wd, "mejb", OVOV, WSTR), Wje = tvs(c.Wd + ov, "mejb", OVOV, WSTR);
    axpby<double>(1.0, tv(c.Covov, "mejb", OVOV), 0.0, Wej, s);    // <mb|ej> = (me|jb)        :641
    cd(1.0, T1("jf"), tv(c.Covvv, "mebf", OVVV), 1.0, Wej);        // t_jf <mb|ef>, (me|bf)     :642
    axpby<double>(-1.0, tv(c.Coovv, "mjbe", OOVV), 0.0, Wje, s);   // -<mb|je> = -(mj|be)       :680
    cd(-1.0, T1("jf"), tv(c.Covvv, "mfbe", OVVV), 1.0, Wje);       // t_jf <mb|fe>, (mf|be)     :681
    cd(-1.0, T1("nb"), tv(c.Cooov, "njme", OOOV), 1.0, Wej);       // t_nb <mn|ej>, (nj|me)     :643
    cd(1.0, T1("nb"), tv(c.Cooov, "mjne", OOOV), 1.0, Wje);        // t_nb <mn|je>, (mj|ne)     :682
    if (!std::is_same<T, double>::value) cast<double, T>(2 * ov * ov, c.Wd, W, s);
  }
  const T* Kt = std::is_same<T, double>::value ? (const T*)c.Covov : (const T*)c.K32;
  const T* Jt = std::is_same<T, double>::value ? (const T*)c.J : (const T*)c.J32;
  const std::vector<int64_t> OVXOV{o, v, 2, o, v};
  // :644 -tau'_jnfb <mn|ef> (tau' = t2/2 + t1 t1), :645 +1/2 t_njfb L_mnef, :683 +tau'_jnfb <mn|fe>:
  //   Wej += K.(T - Tx/2 - P1) - 1/2 J.T ;  Wje += J.tau'
  const double unit = 2.0 * (double)ov * ov * ov;   // one ov x ov x ov GEMM
  const int64_t ovp = c.ovp();
  int kr = c.tbegin(s, Ctx::PH_RING, 2 * unit);
  if (x3) {
    X3Gemm g;   // interleaved operands: rows of 2 ovp, kb = 32, ska = skb = 64
    g.M = ov; g.N = 2 * ov; g.Kp = ovp; g.lda = 2 * ovp; g.ldb = 2 * ovp; g.ldc = 2 * ov;
    g.kb = 32; g.ska = 64; g.skb = 64;
    g.Ah = c.Jh; g.Al = c.Jl; g.Bh = c.BJh; g.Bl = c.BJl; g.C = (float*)W;
    g.sa = c.xs + 1; g.sb = c.xs; g.beta = 1.f;
    x3_gemm<_Float16>(g, s);
  } else {
    contract<T>(T(1), tv(Jt, "menf", OVOV), tv(S1, "nfxjb", OVXOV), T(1), tv(W, "mexjb", OVXOV), ws, s);
  }
  c.tend(s, kr);
  kr = c.tbegin(s, Ctx::PH_RING, unit);
  if (x3) {
    X3Gemm g;
    g.M = ov; g.N = ov; g.Kp = ovp; g.lda = 2 * ovp; g.ldb = 2 * ovp; g.ldc = 2 * ov;
    g.kb = 32; g.ska = 64; g.skb = 64;
    g.Ah = c.Kh; g.Al = c.Kl; g.Bh = c.BKh; g.Bl = c.BKl; g.C = (float*)W;
    g.sa = c.xs + 2; g.sb = c.xs; g.beta = 1.f;
    x3_gemm<_Float16>(g, s);
    // W (FP32 [me][x][jb]) -> split BP[jb][x][me] with its own scale xs[3]
    x3_absmax<float>((const float*)W, 2 * ov * ov, c.xpart, s);
    x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 3, s);
    x3_split_t((const float*)W, ov, ov, 2, 2 * ov, c.xs + 3, c.BPh, c.BPl, ovp, 2 * ovp, s, true);
  } else {
    contract<T>(T(1), tv(Kt, "menf", OVOV), tv(A1, "nfjb", OVOV), T(1), tvs(W, "mejb", OVOV, WSTR), ws, s);
  }
  c.tend(s, kr);

  // ---- ladder + Z in the FP64 / FP32 modes (through the contraction planner and its workspace) ----
  const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
  const double f_lad_P = 2.0 * noP * (npP + ov) * npP, f_lad_M = 2.0 * noM * (npM + ov) * npM;
  if (!c.split_ladder()) {
    tau_pack<T>(o, v, t1, t2, (T*)c.Tp, (T*)c.Tm, s);
    int k = c.tbegin(s, Ctx::PH_LADDER, f_lad_P);
    contract<T>(T(1), tv(c.Tp, "xk", {noP, npP}), tv(c.LP, "ky", {npP, npP + ov}), T(0),
                tv(c.CP, "xy", {noP, npP + ov}), ws, s);
    c.tend(s, k);
    if (noM > 0) {
      k = c.tbegin(s, Ctx::PH_LADDER, f_lad_M);
      contract<T>(T(1), tv(c.Tm, "xk", {noM, npM}), tv(c.LM, "ky", {npM, npM + ov}), T(0),
                  tv(c.CM, "xy", {noM, npM + ov}), ws, s);
      c.tend(s, k);
    }
  }

  // ---- ring GEMMs: ccwfn.py:933-935 ------------------------------------------------------------
  // P[(ia),(jb)] = (2t2 - t2^T)_imae Wmbej + t2_imae Wmbje      (:933 + :934)
  // Q[(ja),(ib)] = t2_mjae Wmbie                                (:935)
  T* P = A1;   // A1 is dead after the K GEMM
  T* Q = S1;   // S1 is dead after the J GEMM
  kr = c.tbegin(s, Ctx::PH_RING, 2 * unit);
  if (x3) {
    X3Gemm g;
    g.M = ov; g.N = ov; g.Kp = 2 * ovp; g.lda = 4 * ovp; g.ldb = 4 * ovp; g.ldc = ov;
    g.kb = 32; g.ska = 64; g.skb = 64;
    g.Ah = c.APh; g.Al = c.APl; g.Bh = c.BPh; g.Bl = c.BPl; g.C = (float*)P;
    g.sa = c.xs; g.sb = c.xs + 3;
    x3_gemm<_Float16>(g, s);
  } else {
    contract<T>(T(1), tv(S2, "iamex", {o, v, o, v, 2}), tv(W, "mexjb", OVXOV), T(0), tv(P, "iajb", OVOV), ws, s);
  }
  c.tend(s, kr);
  kr = c.tbegin(s, Ctx::PH_RING, unit);
  if (x3) {
    X3Gemm g;   // B = the x = 1 half of BP rows: k offset ovp, at 2 ovp in the interleaved row
    g.M = ov; g.N = ov; g.Kp = ovp; g.lda = 2 * ovp; g.ldb = 4 * ovp; g.ldc = ov;
    g.kb = 32; g.ska = 64; g.skb = 64;
    g.Ah = c.AQh; g.Al = c.AQl; g.Bh = c.BPh + 2 * ovp; g.Bl = c.BPl + 2 * ovp; g.C = (float*)Q;
    g.sa = c.xs; g.sb = c.xs + 3;
    x3_gemm<_Float16>(g, s);
  } else {
    contract<T>(T(1), tv(Tx, "jame", OVOV), tvs(W + ov, "meib", OVOV, WSTR), T(0), tv(Q, "jaib", OVOV), ws, s);
  }
  c.tend(s, kr);

  // ---- Zt_mbij = Z_mbij + <mb|ij> + t_ie <mb|ej> + t_je <mb|ie> -----------------------------------
  // :932 -t_ma Z_mbij, :940 -t_ma <mb|ij>, :937 -t_ie t_ma <mb|ej>, and :938 -t_ie t_mb <ma|je>
  // taken in its (ij)(ab)-swapped form -t_ma (t_je <mb|ie>) (same symmetrised r2, :790).
  const TV Zt = tv(c.Zt, "mijb", {o, o, o, v});                     // Z-like terms, [m][i][j][b]
  axpby<double>(1.0, tv(c.Cooov, "mijb", OOOV), 0.0, Zt, s);        // <mb|ij> = (mi|jb)
  cd(1.0, T1("ie"), tv(c.Covov, "mejb", OVOV), 1.0, Zt);            // <mb|ej> = (me|jb)
  cd(1.0, T1("je"), tv(c.Coovv, "mibe", OOVV), 1.0, Zt);            // <mb|ie> = (mi|be)
  if (c.split_ladder()) CCX_HIP(hipStreamWaitEvent(s, c.ev_join, 0));   // CP, CM from the side stream
  if (c.mode == MODE_DP) zt_add<double>(o, v, (const double*)c.CP, (const double*)c.CM, c.Zt, s);
  else zt_add<float>(o, v, (const float*)c.CP, (const float*)c.CM, c.Zt, s);

  // ---- half residual Hh (before P(ij,ab)): ccwfn.py:922-940 ------------------------------------
  const TV Hh = tv(c.Hh, "ijab", OOVV);
  // :923-925  t2_ijae (Fae_be - 1/2 t_mb Fme_me)
  const TV G = tv(c.G, "be", VV);
  axpby<double>(1.0, tv(c.Fae, "be", VV), 0.0, G, s);
  cd(-0.5, T1("mb"), Fme, 1.0, G);
  // In MIXED_F16X3 the FP32 terms :925 (off-diagonal G) and :930 go to Xr (the ring W buffer, dead
  // after the ring GEMMs), and <ij|ab> (:922) and the diagonal t2_ijab G_bb are added in FP64 by
  // assemble_r2; Hh then holds the -t2 Hm term only.
  float* Xr = x3 ? (float*)c.W : nullptr;
  const bool g_split = x3 && c.T2h;
  if (g_split) {
    hh_g_x3(c, t2, Xr, s);   // Xr = off-diagonal t2 G (:925) in FP32; Gd = G's diagonal
  } else {
    axpby<double>(0.5, tv(c.Ioovv, "ijab", OOVV), 0.0, Hh, s);      // 1/2 <ab|ij> = 1/2 <ij|ab> :922
    cd(1.0, T2("ijae"), G, 1.0, Hh);
  }
  // :926-928  -t2_imab (Fmi_mj + 1/2 t_je Fme_me)
  const TV Hm = tv(c.Hm, "mj", OO);
  axpby<double>(1.0, tv(c.Fmi, "mj", OO), 0.0, Hm, s);
  cd(0.5, T1("je"), Fme, 1.0, Hm);
  cd(-1.0, T2("imab"), Hm, g_split ? 0.0 : 1.0, Hh);
  // :930  1/2 tau_mnab Wmnij
  const float* X930 = nullptr;
  if (std::is_same<T, double>::value) {
    cd(0.5, tv(c.tau, "mnab", OOVV), Wmnij, 1.0, Hh);
  } else if (c.mode == MODE_MIXED_F16X3) {
    cast<double, float>(o * o * o * o, c.Wmnij, c.W32, s);
    x3_absmax<float>(c.W32, o * o * o * o, c.xpart, s);
    x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 7, s);
    x3_split_t<float>(c.W32, o * o, o * o, 1, o * o, c.xs + 7, c.WTh, c.WTl, 0, c.o2p(), s);   // [(ij)][(mn)]
    X3Gemm g;   // Xr[(ij)][(ab)] (+)= 1/2 sum_(mn) Wmnij tau_mnab in FP32
    g.M = o * o; g.N = v * v; g.Kp = c.o2p(); g.lda = c.o2p(); g.ldb = c.o2p(); g.ldc = v * v;
    g.Ah = c.WTh; g.Al = c.WTl; g.Bh = c.TTh; g.Bl = c.TTl; g.C = Xr; g.beta = g_split ? 1.f : 0.f;
    g.sa = c.xs + 7; g.sb = c.xs + 5; g.alpha = 0.5f;
    x3_gemm<_Float16>(g, s, &ws);
    X930 = Xr;
  } else {
    float* X32 = (float*)c.S2;   // S2 is dead after the P ring GEMM
    cast<double, float>(o * o * o * o, c.Wmnij, c.W32, s);
    contract<float>(0.5f, tv(c.tau32, "mnab", OOVV), tv(c.W32, "mnij", OOOO), 0.0f, tv(X32, "ijab", OOVV), ws, s);
    X930 = X32;
  }
  // :932, :937, :938, :940 through Zt: -t_ma Zt_mijb, summed inside assemble_r2
  // :939  t_ie <ab|ej> = t_ie (jb|ae): in MIXED_F16X3 the FP32 products PT[i][(j,b,a)], added inside
  // assemble_r2 (with X930, the FP32 :930 term of the mixed modes)
  if (!x3) cd(1.0, T1("ie"), tv(c.Covvv, "jbae", OVVV), 1.0, Hh);

  // ---- r2 = Hh + Hh^T + ladder + rings (ccwfn.py:790) --------------------------------------------
  if (c.mode == MODE_DP)
    assemble_r2<double>(o, v, c.Hh, (const double*)c.CP, (const double*)c.CM, (const double*)P,
                        (const double*)Q, t1, c.Zt, nullptr, nullptr, nullptr, nullptr, nullptr, r2, s);
  else
    assemble_r2<float>(o, v, c.Hh, (const float*)c.CP, (const float*)c.CM, (const float*)P,
                       (const float*)Q, t1, c.Zt, x3 ? c.PT : nullptr, X930, g_split ? c.Ioovv : nullptr,
                       g_split ? t2 : nullptr, g_split ? c.Gd : nullptr, r2, s);

  // ---- r1: ccwfn.py:754-760 ------------------------------------------------------------------
  const TV R1 = tv(r1, "ia", OV);
  axpby<double>(1.0, Fvo, 0.0, R1, s);                               // f_ai                 :754
  cd(1.0, T1("ie"), Fae, 1.0, R1);                                   //                      :755
  cd(-1.0, Fmi, T1("ma"), 1.0, R1);                                  //                      :756
  // U = 2 t2 - t2^T in ring layout: Ud[i][a][m][e] = 2 t2_imae - t2_imea
  // (2t2 - t2^T) Fme :757: r1[i][a] += sum_(m,e) Ud[i][a][m][e] Fme[m][e]
  wsum(o, v, o, v, 1.0, c.Fme, v, c.Ud, v * ov, ov, v, 1.0, r1, v, s);
  // L_nafi = 2(nf|ia) - (ni|af) :758: 2 sum_(nf) t_nf Covov[(nf)][(ia)] - sum_(n,f) t_nf Coovv[n][i][a][f]
  colsum(ov, ov, 2.0, t1, c.Covov, ov, 1.0, r1, (double*)ws.get((size_t)colsum_chunks(ov) * ov * sizeof(double)), s);
  wsum(o, v, o, v, -1.0, t1, v, c.Coovv, v * v, v, o * v * v, 1.0, r1, v, s);
  if (x3) covvv_x3_r1(c, c.Ud, r1, s);                              // <ma|ef> = (me|af)    :759
  else cd(1.0, tv(c.Ud, "meif", OVOV), tv(c.Covvv, "meaf", OVVV), 1.0, R1);
  cd(-1.0, T2("mnae"), tv(c.Looov, "imne", OOOV), 1.0, R1);         // L_nmei = 2(mi|ne) - (ni|me) :760
}

void residual(Ctx& c, const double* fock, const double* t1, const double* t2, double* r1, double* r2,
              hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "ccx_residual called before the integrals were set");
  CCX_CHECK(fock && t1 && t2 && r1 && r2, ERR_ARG, "null amplitude/residual pointer");
  alloc_work(c);
  int k = c.tbegin(s, Ctx::PH_RESIDUAL, 0.0);
  if (c.mode == MODE_DP) residual_t<double>(c, fock, t1, t2, r1, r2, s);
  else residual_t<float>(c, fock, t1, t2, r1, r2, s);
  c.tend(s, k);
}

void timing_start(Ctx& c) {
  for (auto& sp : c.spans) {
    c.pool.push_back(sp.a);
    c.pool.push_back(sp.b);
  }
  c.spans.clear();
  c.timing = true;
}

// out[3*p + 0..2] = launches, milliseconds, algorithmic flops of phase p; stops recording.
void timing_read(Ctx& c, double* out, int n) {
  std::vector<double> acc(3 * Ctx::PH_COUNT, 0.0);
  for (auto& sp : c.spans) {
    CCX_HIP(hipEventSynchronize(sp.b));
    float ms = 0.f;
    CCX_HIP(hipEventElapsedTime(&ms, sp.a, sp.b));
    acc[3 * sp.phase] += 1;
    acc[3 * sp.phase + 1] += ms;
    acc[3 * sp.phase + 2] += sp.flops;
  }
  for (int i = 0; i < n && i < (int)acc.size(); ++i) out[i] = acc[i];
  timing_start(c);
  c.timing = false;
}

// ---------------------------------------------------------------------------------------------
// DIIS, pycc/utils.py:257-361. Vectors are [t1 | t2] concatenated (utils.py:289-291).
void diis_reset(Ctx& c, const double* t1, const double* t2, int max_diis, int cap, hipStream_t s) {
  CCX_CHECK(max_diis >= 0 && cap >= max_diis + 1 && cap <= 64, ERR_ARG, "bad DIIS sizes (capacity 1..64)");
  const int64_t n = c.n1() + c.n2();
  if (c.cap < cap) {
    c.dv = c.alloc<double>((size_t)cap * n);
    c.de = c.alloc<double>((size_t)cap * n);
    if (!c.doldt) c.doldt = c.alloc<double>(n);
    if (!c.dcoef) c.dcoef = c.alloc<double>(64);
    c.cap = cap;
  }
  if (!c.red) c.red = c.alloc<double>(1024 * 64);
  c.max_diis = max_diis;
  c.order.clear();
  c.bdots.assign(c.cap, std::vector<double>(c.cap, 0.0));
  CCX_HIP(hipMemcpyAsync(c.doldt, t1, c.n1() * sizeof(double), hipMemcpyDeviceToDevice, s));
  CCX_HIP(hipMemcpyAsync(c.doldt + c.n1(), t2, c.n2() * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void diis_add(Ctx& c, const double* t1, const double* t2, hipStream_t s) {
  if (c.max_diis == 0) return;     // extrapolate() is a no-op then (utils.py:313-314)
  const int64_t n1 = c.n1(), n = n1 + c.n2();
  CCX_CHECK((int)c.order.size() < c.cap, ERR_STATE, "DIIS history full (raise the capacity)");
  int slot = 0;
  std::vector<bool> used(c.cap, false);
  for (int k : c.order) used[k] = true;
  while (used[slot]) ++slot;
  double* vv = c.dv + (int64_t)slot * n;
  double* ee = c.de + (int64_t)slot * n;
  c.order.push_back(slot);
  // one pass: vv = t, ee = t - t_old (utils.py:287-291), t_old = t, and the first (up to 8) entries
  // of the new overlap row <e_k, e_new> (the rest, for histories past 8, in further passes)
  DiisSlots d{};
  d.n = std::min<int>((int)c.order.size(), 8);
  for (int k = 0; k < d.n; ++k) d.slot[k] = c.order[k];
  diis_push(n1, c.n2(), t1, t2, c.doldt, vv, ee, c.de, d, c.red, c.scal + 64, s);
  std::vector<double> first(d.n);
  for (int k = (int)d.n; k < (int)c.order.size(); ++k)
    dots(n, 1, c.de + (int64_t)c.order[k] * n, n, ee, c.red, c.scal + c.order[k], s);
  CCX_HIP(hipMemcpyAsync(first.data(), c.scal + 64, d.n * sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<double> hrow(c.cap);
  CCX_HIP(hipMemcpyAsync(hrow.data(), c.scal, c.cap * sizeof(double), hipMemcpyDeviceToHost, s));
  CCX_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < d.n; ++k) hrow[c.order[k]] = first[k];
  for (int k : c.order) {
    c.bdots[k][slot] = hrow[k];
    c.bdots[slot][k] = hrow[k];
  }
}

void diis_extrapolate(Ctx& c, double* t1, double* t2, double* coef_out, hipStream_t s) {
  if (c.max_diis == 0) return;
  const int64_t n1 = c.n1(), n = n1 + c.n2();
  if ((int)c.order.size() > c.max_diis) c.order.erase(c.order.begin());   // utils.py:317-320
  const int m = (int)c.order.size();
  CCX_CHECK(m > 0, ERR_STATE, "DIIS extrapolation without stored vectors");
  // bordered Pulay system (utils.py:327-348), overlap block normalised by its largest magnitude
  std::vector<double> B((m + 1) * (m + 1), -1.0), rhs(m + 1, 0.0);
  B[m * (m + 1) + m] = 0.0;
  double mx = 0.0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) mx = std::max(mx, std::fabs(c.bdots[c.order[i]][c.order[j]]));
  CCX_CHECK(mx > 0.0 && std::isfinite(mx), ERR_NUMERIC, "DIIS overlap matrix is zero or not finite");
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) B[i * (m + 1) + j] = c.bdots[c.order[i]][c.order[j]] / mx;
  rhs[m] = -1.0;
  // LU with partial pivoting (as LAPACK gesv)
  const int N = m + 1;
  std::vector<int> piv(N);
  for (int k = 0; k < N; ++k) {
    int p = k;
    for (int i = k + 1; i < N; ++i)
      if (std::fabs(B[i * N + k]) > std::fabs(B[p * N + k])) p = i;
    CCX_CHECK(B[p * N + k] != 0.0, ERR_NUMERIC, "singular DIIS system");
    if (p != k) {
      for (int j = 0; j < N; ++j) std::swap(B[k * N + j], B[p * N + j]);
      std::swap(rhs[k], rhs[p]);
    }
    for (int i = k + 1; i < N; ++i) {
      double f = B[i * N + k] / B[k * N + k];
      for (int j = k; j < N; ++j) B[i * N + j] -= f * B[k * N + j];
      rhs[i] -= f * rhs[k];
    }
  }
  std::vector<double> x(N);
  for (int i = N - 1; i >= 0; --i) {
    double acc = rhs[i];
    for (int j = i + 1; j < N; ++j) acc -= B[i * N + j] * x[j];
    x[i] = acc / B[i * N + i];
  }
  // t = sum_k c_k t_k over the stored (non-contiguous) slots: gather coefficients per slot
  std::vector<double> coef(c.cap, 0.0);
  for (int i = 0; i < m; ++i) coef[c.order[i]] = x[i];
  if (coef_out)
    for (int i = 0; i < m; ++i) coef_out[i] = x[i];
  CCX_HIP(hipMemcpyAsync(c.dcoef, coef.data(), c.cap * sizeof(double), hipMemcpyHostToDevice, s));
  // t = sum_k c_k t_k (utils.py:351-355), written to t_old as well (utils.py:358-359)
  lincomb2(n1, c.cap, c.dcoef, c.dv, n, t1, c.doldt, s);
  lincomb2(c.n2(), c.cap, c.dcoef, c.dv + n1, n, t2, c.doldt + n1, s);
  CCX_HIP(hipStreamSynchronize(s));   // `coef` (host) must outlive the copy
}

}  // namespace ccx

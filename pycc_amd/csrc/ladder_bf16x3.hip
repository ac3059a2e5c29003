// BF16x3 split-product ladder for gfx950: the packed ladder of ccwfn.py:931 (and the Z columns,
// :715) on the split 16-bit GEMM of split16.hip with E = bf16.
//
// x = x_hi + x_lo with x_hi = bf16(x), x_lo = bf16(x - x_hi); the product a*b is taken as
// a_hi*b_hi + a_hi*b_lo + a_lo*b_hi (the lo*lo term, ~2^-16 relative, is dropped), accumulated in
// FP32. Both operands are stored K-contiguous with K padded to a multiple of 32 (zeros):
//   A  = packed tau+- rows  [M][Kp]
//   Bt = ladder matrix      [N][Kp]   (its (ab) block is symmetric, so it is packed untransposed)
//   C  = CP / CM            [M][N]    FP32, consumed by zt_add / assemble_r2
#include "ccsd.h"

namespace ccx {

namespace {

void bf16x3_gemm(int64_t M, int64_t N, int64_t Kp, const bf16* Ah, const bf16* Al, const bf16* Bh,
                 const bf16* Bl, float* C, int64_t ldc, hipStream_t s) {
  X3Gemm g;
  g.M = M; g.N = N; g.Kp = Kp; g.lda = Kp; g.ldb = Kp; g.ldc = ldc;
  g.Ah = Ah; g.Al = Al; g.Bh = Bh; g.Bl = Bl; g.C = C;
  x3_gemm<__bf16>(g, s);
}

}  // namespace

void ladder_bf16x3(Ctx& c, const double* t1, const double* t2, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov();
  (void)o;
  const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
  const int64_t KpP = bf16_kpad(npP), KpM = bf16_kpad(npM);
  tau_pack_bf16x3(o, v, KpP, KpM, t1, t2, c.Tph, c.Tpl, c.Tmh, c.Tml, s);
  // algorithmic flops counted as one GEMM (the three split products are the method's cost)
  int k = c.tbegin(s, Ctx::PH_LADDER, 2.0 * noP * (npP + ov) * npP);
  bf16x3_gemm(noP, npP + ov, KpP, c.Tph, c.Tpl, c.LPh, c.LPl, (float*)c.CP, npP + ov, s);
  c.tend(s, k);
  if (noM > 0) {
    k = c.tbegin(s, Ctx::PH_LADDER, 2.0 * noM * (npM + ov) * npM);
    bf16x3_gemm(noM, npM + ov, KpM, c.Tmh, c.Tml, c.LMh, c.LMl, (float*)c.CM, npM + ov, s);
    c.tend(s, k);
  }
}

}  // namespace ccx

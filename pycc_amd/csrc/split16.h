// Split 16-bit ("x3") GEMMs and their operand conversions (split16.hip).
#pragma once
#include "common.h"
#include "kernels.h"

namespace ccx {

typedef _Float16 f16;

// C[m][n] = alpha / (sa sb) * sum_k (Ah Bh + Ah Bl + Al Bh)[m][n][k] + beta C[m][n], with A stored
// [M][lda] and B stored [N][ldb], both K-contiguous, K padded to Kp (a multiple of 32, zeros in
// the padding of both operands). sa / sb are device scalars (operand scales) or both null.
struct X3Gemm {
  int64_t M = 0, N = 0, Kp = 0, lda = 0, ldb = 0, ldc = 0;
  const void *Ah = nullptr, *Al = nullptr, *Bh = nullptr, *Bl = nullptr;
  float* C = nullptr;
  const float *sa = nullptr, *sb = nullptr;
  const float* sbn = nullptr;        // per-row scales of B (one per n), in place of sb
  float alpha = 1.f, beta = 0.f;
  // blocked K (kb > 0): k = q kb + r lives at offset q ska + r in a row of A (q skb + r in B); kb
  // is a multiple of 32, so a 32-wide k-tile never straddles two blocks. Interleaved operands
  // (hi and lo of each 32-wide k-block side by side, lo = hi + 32) are kb = 32, ska / skb = 64.
  int64_t kb = 0, ska = 0, skb = 0;
  // FP64 destination (optional): C64[m ldc + n] = alpha / (sa sb) sum + beta C64. On grids under 512
  // tiles (needs ws) the split-K partials (FP32, each over a short K range) are summed in FP64; on
  // larger grids the kernel adds its FP32 sum into C64 directly.
  double* C64 = nullptr;
  // with C64 written straight from the kernel (full grids only): beta scales C64in instead of C64
  // when given, and D64[m ldc + n] dvec[n] is added in FP64 (a diagonal term kept out of the split)
  const double* C64in = nullptr;
  const double* D64 = nullptr;
  const double* dvec = nullptr;
  // split-K (set by x3_gemm): grid.y = splits K ranges of kchunk, raw partials in work[s][M][N]
  int64_t kchunk = 0;
  float* work = nullptr;
  // x3_gemm_kcat only (interleaved operands, FP32 C): k in [K1, Kp) reads A from Ah2 (lo = Ah2 +
  // 32) at k - K1 and row (m mod sw) sw + m / sw, the (x, y) -> (y, x) swap of a row m = x sw + y
  // (M = sw^2); batch b of nbatch adds b sA2 to Ah2, b sB to Bh and Bl, and b sC to C.
  const void* Ah2 = nullptr;
  int64_t K1 = 0, sw = 0, nbatch = 1, sA2 = 0, sB = 0, sC = 0;
  // ytile (x3_gemm_kcat only, = sw = N, a multiple of 8): C element (m = x sw + y, n) goes to the
  // tiled layout [x][y / 8][n / 8][n % 8][y % 8] (ldc unused), so that an 8 x 8 (y, n) block is one
  // run of 256 bytes, stored column by column: the offset is (m - m % 8) sw + 8 n + m % 8
  int64_t ytile = 0;
};
struct Workspace;
// ws (optional) enables split-K when the tile grid cannot fill the chip.
template <typename E>   // __bf16 or _Float16
void x3_gemm(const X3Gemm& p, hipStream_t s, Workspace* ws = nullptr);

// An FP64 GEMM of the contraction planner (batch 1) as an FP16x3 split product: A and B^T are
// split (one power-of-two scale per operand, from its largest magnitude) into scratch from ws,
// multiplied with FP32 accumulation and added into C (alpha, beta) in FP64. Returns false, doing
// nothing, when the tile grid is under one full round (the caller then runs the FP64 GEMM).
bool x3_contract(const GemmParams<double>& p, Workspace& ws, hipStream_t s);

// nbatch FP16x3 GEMMs C_b = alpha / (sa sb) [A | A2_b swapped] [B_b] with K = Kp concatenated
// from two operands (fields Ah2 .. sC above): the connected (T) numerators, two slot products per
// GEMM. Output tiles of 384 x 192, 256 x 256 or 256 x 192, whichever pads M x N least.
void x3_gemm_kcat(const X3Gemm& p, hipStream_t s);
// max |x| over n elements as x3_partials() per-block partial maxima in `part`.
template <typename T>
void x3_absmax(const T* x, int64_t n, float* part, hipStream_t s);
int x3_partials();
// *s_out = 2^k such that (ca max(pa) + cb max(pb)^2) * 2^k < 2^15 (pb may be null).
void x3_scale(const float* pa, float ca, const float* pb, float cb, float* s_out, hipStream_t s);
// *s_out = 2^k such that max(pa[0 .. n)) * 2^k < 2^15: the scale of the union of the operands whose
// partial maxima fill pa (n a multiple of x3_partials()).
void x3_scale_union(const float* pa, int n, float* s_out, hipStream_t s);

// hi/lo[r * ldd + c] = split(src[r * lds + c] * sc), r < rows, c < cols. il: interleaved layout
// (lo = hi + 32, ldd a multiple of 32; k-block q of a row at 64 q, the row at 2 ldd r).
template <typename Tin>
void x3_split_rows(const Tin* src, int64_t rows, int64_t cols, int64_t lds, const float* sc, f16* hi, f16* lo,
                   int64_t ldd, hipStream_t s, bool il = false);
// hi/lo[cc * ldd + x * Kb + r] = split(src[r * lds + x * Cb + cc] * sc), r < R, x < nx, cc < Cb.
template <typename Tin>
void x3_split_t(const Tin* src, int64_t R, int64_t Cb, int nx, int64_t lds, const float* sc, f16* hi, f16* lo,
                int64_t Kb, int64_t ldd, hipStream_t s, bool il = false);
// FP16x3 packed ladder matrices (the BF16x3 layouts of kernels.h, [N][Kp], interleaved: lo = hi + 32)
// with one power-of-two scale per row in sP / sM: the (ab) rows for e in [e0, e1) from src, and the
// (m, b) rows from Covvv.
void ladder_pack_x3(int64_t v, int64_t e0, int64_t e1, VVVVSrc src, f16* LPh, f16* LPl, f16* LMh, f16* LMl,
                    float* sP, float* sM, hipStream_t s);
void ladder_pack_o_x3(int64_t o, int64_t v, const double* Covvv, f16* LPh, f16* LPl, f16* LMh, f16* LMl,
                      float* sP, float* sM, hipStream_t s);
// FP16x3 packed tau rows (tau_pack_bf16x3 layout, interleaved: lo = hi + 32) scaled by *sc.
void tau_pack_x3(int64_t o, int64_t v, int64_t KpP, int64_t KpM, const double* t1, const double* t2, const float* sc,
                 f16* Tph, f16* Tpl, f16* Tmh, f16* Tml, hipStream_t s);
// C = alpha op(A) op(B) + beta C (FP32 row-major, op = transpose when ta / tb) through FP16x3 with
// copies of the split operands in ws and split-K partials in kws.
void gemm_f16x3(bool ta, bool tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                const float* B, int64_t ldb, float beta, float* C, int64_t ldc, Workspace& ws, Workspace& kws,
                hipStream_t s);
// The <ov|vv> terms of MIXED_F16X3 (csrc/split16.hip): split copies of Covvv, the three split
// GEMMs and their scatter into Fae, W, Hh and r1.
struct Ctx;
// max|t2| and max|t1| once per residual, and from them every amplitude-derived operand scale
void x3_amp_scales(Ctx& c, const double* t1, const double* t2, hipStream_t s);
// Fae -= tau~_mnaf L_mnef (:497) as an FP16x3 GEMM (tau~ in c.tau). Its twin in Fmi (:533) stays
// FP64: its ~1 Eh diagonal sums over o v^2 terms made the converged energy drift by ~1 uEh in FP16x3.
void fae_tauL_x3(Ctx& c, double* Fae, hipStream_t s);
void covvv_x3_setup(Ctx& c, hipStream_t s);
void covvv_x3_products(Ctx& c, const double* t1, hipStream_t s);   // PT, P2T from t1
void covvv_x3_fae(Ctx& c, double* Fae, hipStream_t s);              // + 2 t_mf (mf|ae) - t_mf (me|af)
// the FP32 ring W[(m,e)][x][(j,b)] (:641-643, :680-682) from W0, PT, P2T and t1 (Cooov terms)
void w_build_x3(Ctx& c, const double* t1, float* W, hipStream_t s);
void covvv_x3_r1(Ctx& c, const double* Ud, double* r1, hipStream_t s);   // :759
// X = t2_ijae G_be over e != b (the off-diagonal part of :925) as one FP16x3 GEMM with FP32 output
// [i][j][a][b]; the diagonal t2_ijab G_bb (orbital-energy sized, kept in c.Gd) and <ij|ab> (:922)
// are added in FP64 by the r2 assembly. Sized for grids of at least 512 tiles (hh_g_x3_fits).
bool hh_g_x3_fits(int64_t o, int64_t v);
void hh_g_x3(Ctx& c, const double* t2, float* X, hipStream_t s);
// The split ring operands BJ, BK, AP, AQ (K = ovp, or 2 ovp for AP; interleaved, lo = hi + 32) and Ud
// from t1, t2.
void ring_prep_x3(int64_t o, int64_t v, int64_t ovp, const double* t1, const double* t2, const float* sc, f16* BJh,
                  f16* BJl, f16* BKh, f16* BKl, f16* APh, f16* APl, f16* AQh, f16* AQl, double* Ud,
                  hipStream_t s);

}  // namespace ccx

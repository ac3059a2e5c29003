// Fused elementwise / packing / reduction kernels of the CCSD residual (see kernels.hip).
#pragma once
#include "common.h"
#include <hip/hip_bf16.h>

namespace ccx {

typedef __hip_bfloat16 bf16;

// Pair tables of the packed ladder: index p -> (lo, hi) with lo <= hi (P) or lo < hi (M), ordered
// by hi then lo: pP(a, b) = b(b+1)/2 + a, pM(a, b) = b(b-1)/2 + a.
__host__ __device__ inline int64_t pairP(int64_t a, int64_t b) { return b * (b + 1) / 2 + a; }
__host__ __device__ inline int64_t pairM(int64_t a, int64_t b) { return b * (b - 1) / 2 + a; }

// Chemist block (pq|rs) gathered from a dense Dirac ERI <pq|rs> = (pr|qs):
// out[p][q][r][s] = ERI[p0+p][r0+r][q0+q][s0+s].
void gather_chem(const double* eri, int64_t nmo, int64_t p0, int64_t np, int64_t q0, int64_t nq,
                 int64_t r0, int64_t nr, int64_t s0, int64_t ns, double* out, hipStream_t s);

// tau = f1 * t2 + f2 * t1 (x) t1, layout [i][j][a][b]
void build_tau(int64_t o, int64_t v, double f1, double f2, const double* t1, const double* t2,
               double* tau, hipStream_t s);

// Ring-layout operands ([i][a][j][b], i.e. ov x ov matrices), see ring_prep_kernel.
template <typename T>
void ring_prep(int64_t o, int64_t v, const double* t1, const double* t2, T* S1, T* A1, T* S2, T* Tx,
               double* Ud, hipStream_t s);

// Packed ladder amplitudes: Tp[(i<=j)][(e<=f)], Tm[(i<j)][(e<f)] from tau = t2 + t1 t1 (t1 may be
// null: tau = t2).
template <typename T>
void tau_pack(int64_t o, int64_t v, const double* t1, const double* t2, T* Tp, T* Tm, hipStream_t s);
// BF16x3 / FP16x3 ladder operands are K-contiguous with K padded to an odd number of 32-wide blocks
// (padding stays zero): interleaved rows are then an odd number of 128-byte lines long, which spreads
// the rows of one k-step over the L2 channels (measured at o=48, v=384: the tau- ladder GEMM fetched
// 60 GB instead of 78 GB from beyond L2, and the iteration took 1% less).
__host__ __device__ inline int64_t bf16_kpad(int64_t k) {
  const int64_t b = (k + 31) / 32;
  return (b + (b > 0 && b % 2 == 0 ? 1 : 0)) * 32;
}
void tau_pack_bf16x3(int64_t o, int64_t v, int64_t KpP, int64_t KpM, const double* t1, const double* t2,
                     bf16* Tp_hi, bf16* Tp_lo, bf16* Tm_hi, bf16* Tm_lo, hipStream_t s);

// Source of (ea|fb) for e in [e0, e1): address base + e*se + a*sa + f*sf + b*sb (absolute e).
struct VVVVSrc {
  const double* base;
  int64_t se, sa, sf, sb;
};
// Rows (e<=f / e<f) with e0 <= e < e1 of the packed ladder matrices; columns [0, npairs) hold
// 1/2(<ab|ef> +- <ab|fe>) and columns npairs + m*v + b hold 1/2(<mb|ef> +- <mb|fe>) from Covvv.
template <typename T>
void ladder_pack(int64_t o, int64_t v, int64_t e0, int64_t e1, VVVVSrc src, const double* Covvv,
                 T* LP, T* LM, hipStream_t s);
// BF16x3 ladder matrices stored [N][Kp]: the (ab) block is symmetric under (ab) <-> (ef), so its
// rows are written untransposed; the (m, b) rows are written by ladder_pack_o_bf16x3.
void ladder_pack_bf16x3(int64_t o, int64_t v, int64_t e0, int64_t e1, VVVVSrc src, bf16* LP_hi,
                        bf16* LP_lo, bf16* LM_hi, bf16* LM_lo, hipStream_t s);
void ladder_pack_o_bf16x3(int64_t o, int64_t v, const double* Covvv, bf16* LP_hi, bf16* LP_lo,
                          bf16* LM_hi, bf16* LM_lo, hipStream_t s);

// Zt[m][i][j][b] += Z_mbij from the packed ladder outputs.
template <typename T>
void zt_add(int64_t o, int64_t v, const T* CP, const T* CM, double* Zt, hipStream_t s);

// r2[i][j][a][b] = H[ijab] + H[jiba] + ladder[ijab] + P[ia,jb] + P[jb,ia] + Q[ja,ib] + Q[ib,ja]
//                  - sum_m (t1[m][a] Zt[m][i][j][b] + t1[m][b] Zt[m][j][i][a])
//                  [+ PTt[i][j][b][a] + PTt[j][i][a][b]] [+ X[ijab] + X[jiba]]   (FP32, each optional)
//                  [+ I[ijab]] [+ T2d[ijab] Gd[b] + T2d[jiba] Gd[a]]               (FP64, each optional)
// computed once per pair-symmetric couple and written to both (ij)(ab) partners
template <typename T>
void assemble_r2(int64_t o, int64_t v, const double* H, const T* CP, const T* CM, const T* P,
                 const T* Q, const double* t1, const double* Zt, const float* PTt, const float* X,
                 const double* I, const double* T2d, const double* Gd, double* r2, hipStream_t s);

// E = 2 f_ia t_ia + (t2_ijab + t_ia t_jb) L_ijab, written to e_out (device scalar).
void energy(int64_t o, int64_t v, const double* fock, int64_t nmo, const double* t1, const double* t2,
            const double* Loovv, double* partials, double* e_out, hipStream_t s);

// t1 += r1/Dia, t2 += r2/Dijab; rms2_out = sum (r/D)^2 (device scalar).
void jacobi_update(int64_t o, int64_t v, const double* eps, double* t1, double* t2, const double* r1,
                   const double* r2, double* partials, double* rms2_out, hipStream_t s);

// out[k] = <x_k, y> for k < nvec, vectors of length n spaced `stride` apart (device array out).
void dots(int64_t n, int nvec, const double* x0, int64_t stride, const double* y, double* partials,
          double* out, hipStream_t s);

// out = sum_k c[k] * x_k (c on the device), vectors spaced `stride` apart.
void lincomb(int64_t n, int nvec, const double* c, const double* x0, int64_t stride, double* out,
             hipStream_t s);

// out = x - y
void sub(int64_t n, const double* x, const double* y, double* out, hipStream_t s);

// DIIS push (one pass): v = [t1; t2], e = v - oldt, oldt = v, and out[k] = <e, E_k> for the nd stored
// error vectors E_k = E + slots[k] * n (a slot equal to the new one uses e itself); nd <= 8.
struct DiisSlots {
  int n;
  int slot[8];
};
void diis_push(int64_t n1, int64_t n2, const double* t1, const double* t2, double* oldt, double* v, double* e,
               const double* E, DiisSlots d, double* partials, double* out, hipStream_t s);
// lincomb into two outputs: out = out2 = sum_k c_k x_k (the extrapolated amplitudes and the new oldt)
void lincomb2(int64_t n, int nvec, const double* c, const double* x0, int64_t stride, double* out, double* out2,
              hipStream_t s);

template <typename Tin, typename Tout>
void cast(int64_t n, const Tin* in, Tout* out, hipStream_t s);

// out += alpha * in (FP32 partial results folded into FP64 sums)
void accumulate(int64_t n, double alpha, const float* in, double* out, hipStream_t s);

// out[p ldo + q] = beta out + alpha sum_{n < N, f < F} w[n ldw + f] X[p sp + q sq + n sn + f] (FP64,
// one wave per output; for contractions with a small output and a contiguous innermost index).
void wsum(int64_t P, int64_t Q, int64_t N, int64_t F, double alpha, const double* w, int64_t ldw, const double* X,
          int64_t sp, int64_t sq, int64_t sn, double beta, double* out, int64_t ldo, hipStream_t s);
// out[c] = beta out[c] + alpha sum_r w[r] X[r ldx + c] (FP64; partial needs colsum_chunks(R) * C doubles).
int colsum_chunks(int64_t R);
void colsum(int64_t R, int64_t C, double alpha, const double* w, const double* X, int64_t ldx, double beta,
            double* out, double* partial, hipStream_t s);

}  // namespace ccx

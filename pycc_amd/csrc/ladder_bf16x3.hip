// BF16x3 split-product ladder GEMM for gfx950 (v_mfma_f32_16x16x32_bf16).
//
// x = x_hi + x_lo with x_hi = bf16(x), x_lo = bf16(x - x_hi); the product a*b is taken as
// a_hi*b_hi + a_hi*b_lo + a_lo*b_hi (the lo*lo term, ~2^-16 relative, is dropped), accumulated in
// FP32. Both operands are stored K-contiguous with K padded to a multiple of 32 (zeros), so every
// lane loads 16-byte vectors of 8 consecutive k values:
//   A  = packed tau+- rows  [M][Kp]
//   Bt = ladder matrix      [N][Kp]   (its (ab) block is symmetric, so it is packed untransposed)
//   C  = CP / CM            [M][N]    FP32, consumed by zt_add / assemble_r2
// MFMA operand layout used: lane l holds A[l % 16][8*(l/16) .. +7] and Bt[l % 16][8*(l/16) .. +7];
// C register r of lane l is C[4*(l/16) + r][l % 16].
#include "ccsd.h"

namespace ccx {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int BM = 128, BN = 128, BK = 32, KPAD = 8;   // LDS row = 40 bf16 = 80 bytes

__device__ __forceinline__ int64_t xcd_remap_l(int64_t w, int64_t ntiles) {
  const int64_t nx = 8;
  int64_t q = ntiles / nx, r = ntiles % nx;
  int64_t x = w % nx, j = w / nx;
  return x * q + (x < r ? x : r) + j;
}

__global__ __launch_bounds__(256) void bf16x3_gemm_kernel(int64_t M, int64_t N, int64_t Kp,
                                                         const bf16* __restrict__ Ah, const bf16* __restrict__ Al,
                                                         const bf16* __restrict__ Bh, const bf16* __restrict__ Bl,
                                                         float* __restrict__ C, int64_t ldc) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][2][BM][BK + KPAD];   // [buf][hi/lo]
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][2][BN][BK + KPAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int64_t t = xcd_remap_l(blockIdx.x, tiles_m * tiles_n);
  // M is small (<= ~1200): walk all M tiles of one N panel consecutively
  const int64_t tm = t % tiles_m, tn = t / tiles_m;

  // each operand tile (128 rows x 32 k) = 512 vectors of 16 B; thread loads rows r and r + 64
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  const int64_t mA0 = tm * BM + lr, mA1 = mA0 + 64, nB0 = tn * BN + lr, nB1 = nB0 + 64;
  const bool okA0 = mA0 < M, okA1 = mA1 < M, okB0 = nB0 < N, okB1 = nB1 < N;
  // rows beyond the matrix read row 0 (valid memory) and are zeroed before the LDS store
  const bf16* pAh0 = Ah + (okA0 ? mA0 : 0) * Kp + lk;
  const bf16* pAh1 = Ah + (okA1 ? mA1 : 0) * Kp + lk;
  const bf16* pAl0 = Al + (okA0 ? mA0 : 0) * Kp + lk;
  const bf16* pAl1 = Al + (okA1 ? mA1 : 0) * Kp + lk;
  const bf16* pBh0 = Bh + (okB0 ? nB0 : 0) * Kp + lk;
  const bf16* pBh1 = Bh + (okB1 ? nB1 : 0) * Kp + lk;
  const bf16* pBl0 = Bl + (okB0 ? nB0 : 0) * Kp + lk;
  const bf16* pBl1 = Bl + (okB1 ? nB1 : 0) * Kp + lk;
  const bf16x8 zero8 = bf16x8{};
  bf16x8 rAh0, rAh1, rAl0, rAl1, rBh0, rBh1, rBl0, rBl1;

#define CCX_GLOAD(k0)                                        \
  do {                                                       \
    rAh0 = *(const bf16x8*)(pAh0 + (k0));                    \
    rAh1 = *(const bf16x8*)(pAh1 + (k0));                    \
    rAl0 = *(const bf16x8*)(pAl0 + (k0));                    \
    rAl1 = *(const bf16x8*)(pAl1 + (k0));                    \
    rBh0 = *(const bf16x8*)(pBh0 + (k0));                    \
    rBh1 = *(const bf16x8*)(pBh1 + (k0));                    \
    rBl0 = *(const bf16x8*)(pBl0 + (k0));                    \
    rBl1 = *(const bf16x8*)(pBl1 + (k0));                    \
  } while (0)
#define CCX_SSTORE(b)                                                    \
  do {                                                                   \
    *(bf16x8*)&sA[b][0][lr][lk] = okA0 ? rAh0 : zero8;                   \
    *(bf16x8*)&sA[b][0][lr + 64][lk] = okA1 ? rAh1 : zero8;              \
    *(bf16x8*)&sA[b][1][lr][lk] = okA0 ? rAl0 : zero8;                   \
    *(bf16x8*)&sA[b][1][lr + 64][lk] = okA1 ? rAl1 : zero8;              \
    *(bf16x8*)&sB[b][0][lr][lk] = okB0 ? rBh0 : zero8;                   \
    *(bf16x8*)&sB[b][0][lr + 64][lk] = okB1 ? rBh1 : zero8;              \
    *(bf16x8*)&sB[b][1][lr][lk] = okB0 ? rBl0 : zero8;                   \
    *(bf16x8*)&sB[b][1][lr + 64][lk] = okB1 ? rBl1 : zero8;              \
  } while (0)

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int fr = lane >> 4, fc = lane & 15;
  CCX_GLOAD(0);
  CCX_SSTORE(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
    const bool has_next = k0 + BK < Kp;
    if (has_next) CCX_GLOAD(k0 + BK);
    bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ah[i] = *(const bf16x8*)&sA[buf][0][wm * 64 + i * 16 + fc][fr * 8];
      al[i] = *(const bf16x8*)&sA[buf][1][wm * 64 + i * 16 + fc][fr * 8];
      bh[i] = *(const bf16x8*)&sB[buf][0][wn * 64 + i * 16 + fc][fr * 8];
      bl[i] = *(const bf16x8*)&sB[buf][1][wn * 64 + i * 16 + fc][fr * 8];
    }
    // the small products first; 16 independent accumulators between dependent MFMAs
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    if (has_next) CCX_SSTORE(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#undef CCX_GLOAD
#undef CCX_SSTORE

#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int64_t n = tn * BN + wn * 64 + j * 16 + fc;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t m = tm * BM + wm * 64 + i * 16 + fr * 4 + r;
        if (m < M) C[m * ldc + n] = acc[i][j][r];
      }
  }
}

// 256 x 256 tiles, 4 waves (one per SIMD), each computing 128 x 128 (8 x 8 MFMA blocks of 16 x 16).
// Per launch the operand panels are read (M/256)(N/256)(256 + 256) K times instead of
// (M/128)(N/128)(128 + 128) K times: half the L2 -> HBM traffic of the 128 x 128 kernel, which is
// what bounds this GEMM (K ~ 7e4, so each panel is tens of MB). LDS: 2 buffers x (A, B) x (hi, lo)
// x 256 rows x 40 bf16 = 160 KiB, one workgroup per CU.
constexpr int BT = 256;

__global__ __launch_bounds__(256, 1) void bf16x3_gemm256_kernel(int64_t M, int64_t N, int64_t Kp,
                                                               const bf16* __restrict__ Ah,
                                                               const bf16* __restrict__ Al,
                                                               const bf16* __restrict__ Bh,
                                                               const bf16* __restrict__ Bl,
                                                               float* __restrict__ C, int64_t ldc) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][2][BT][BK + KPAD];   // [buf][hi/lo]
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][2][BT][BK + KPAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t tiles_m = (M + BT - 1) / BT, tiles_n = (N + BT - 1) / BT;
  const int64_t t = xcd_remap_l(blockIdx.x, tiles_m * tiles_n);
  const int64_t tm = t % tiles_m, tn = t / tiles_m;   // the M tiles of one N panel run together

  // each operand tile (256 rows x 32 k) = 1024 vectors of 16 B; thread loads rows lr + 64 q, q < 4
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  int64_t offA[4], offB[4];   // element offsets of this thread's rows at k = 0
  bool okA[4], okB[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t m = tm * BT + lr + 64 * q, n = tn * BT + lr + 64 * q;
    okA[q] = m < M;
    okB[q] = n < N;
    // rows beyond the matrix read row 0 (valid memory) and are zeroed before the LDS store
    offA[q] = (okA[q] ? m : 0) * Kp + lk;
    offB[q] = (okB[q] ? n : 0) * Kp + lk;
  }
  const bf16x8 zero8 = bf16x8{};
  bf16x8 rA[2][4], rB[2][4];

  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rA[0][q] = *(const bf16x8*)(Ah + offA[q] + k0);
      rA[1][q] = *(const bf16x8*)(Al + offA[q] + k0);
      rB[0][q] = *(const bf16x8*)(Bh + offB[q] + k0);
      rB[1][q] = *(const bf16x8*)(Bl + offB[q] + k0);
    }
  };
  auto sstore = [&](int b) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        *(bf16x8*)&sA[b][h][lr + 64 * q][lk] = okA[q] ? rA[h][q] : zero8;
        *(bf16x8*)&sB[b][h][lr + 64 * q][lk] = okB[q] ? rB[h][q] : zero8;
      }
  };

  f4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int fr = lane >> 4, fc = lane & 15;
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
    const bool has_next = k0 + BK < Kp;
    if (has_next) gload(k0 + BK);
    bf16x8 a[8], b[8];
    // lo(A) x hi(B)
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = *(const bf16x8*)&sA[buf][1][wm * 128 + i * 16 + fc][fr * 8];
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = *(const bf16x8*)&sB[buf][0][wn * 128 + j * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    // hi(A) x hi(B)
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = *(const bf16x8*)&sA[buf][0][wm * 128 + i * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    // hi(A) x lo(B)
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = *(const bf16x8*)&sB[buf][1][wn * 128 + j * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (has_next) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t n = tn * BT + wn * 128 + j * 16 + fc;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = tm * BT + wm * 128 + i * 16 + fr * 4 + r;
        if (m < M) C[m * ldc + n] = acc[i][j][r];
      }
  }
}

void bf16x3_gemm(int64_t M, int64_t N, int64_t Kp, const bf16* Ah, const bf16* Al, const bf16* Bh,
                 const bf16* Bl, float* C, int64_t ldc, hipStream_t s) {
  if (M == 0 || N == 0) return;
  CCX_CHECK(Kp % BK == 0, ERR_INTERNAL, "bf16x3 K padding");
  int64_t tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  CCX_CHECK(tiles < (1ll << 31), ERR_ARG, "bf16x3 grid too large");
  if (Kp == 0) {   // empty contraction (v = 1): C = 0
    scale<float>(0.f, tvs(C, "mn", {M, N}, {ldc, 1}), s);
    return;
  }
  if (M >= BT) {
    const int64_t tiles256 = ((M + BT - 1) / BT) * ((N + BT - 1) / BT);
    bf16x3_gemm256_kernel<<<dim3((unsigned)tiles256), 256, 0, s>>>(M, N, Kp, Ah, Al, Bh, Bl, C, ldc);
  } else {
    bf16x3_gemm_kernel<<<dim3((unsigned)tiles), 256, 0, s>>>(M, N, Kp, Ah, Al, Bh, Bl, C, ldc);
  }
  CCX_HIP(hipGetLastError());
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

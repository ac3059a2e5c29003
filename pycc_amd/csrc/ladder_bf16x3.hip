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

union V16 {
  uint4 u;
  bf16x8 v;
};

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
  __shared__ __attribute__((aligned(16))) bf16 sA[2][2][BM][BK + KPAD];   // [buf][hi/lo]
  __shared__ __attribute__((aligned(16))) bf16 sB[2][2][BN][BK + KPAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int64_t t = xcd_remap_l(blockIdx.x, tiles_m * tiles_n);
  // M is small (<= ~1200): walk all M tiles of one N panel consecutively
  const int64_t tm = t % tiles_m, tn = t / tiles_m;

  // each operand tile (128 rows x 32 k) = 512 vectors of 16 B; thread loads rows r and r + 64
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  uint4 ra[2][2], rb[2][2];   // [row half][hi/lo]

  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int64_t m = tm * BM + lr + 64 * h;
      int64_t n = tn * BN + lr + 64 * h;
      uint4 z = make_uint4(0, 0, 0, 0);
      ra[h][0] = m < M ? *(const uint4*)(Ah + m * Kp + k0 + lk) : z;
      ra[h][1] = m < M ? *(const uint4*)(Al + m * Kp + k0 + lk) : z;
      rb[h][0] = n < N ? *(const uint4*)(Bh + n * Kp + k0 + lk) : z;
      rb[h][1] = n < N ? *(const uint4*)(Bl + n * Kp + k0 + lk) : z;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        *(uint4*)&sA[buf][p][lr + 64 * h][lk] = ra[h][p];
        *(uint4*)&sB[buf][p][lr + 64 * h][lk] = rb[h][p];
      }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int fr = lane >> 4, fc = lane & 15;
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
    const bool has_next = k0 + BK < Kp;
    if (has_next) gload(k0 + BK);
    V16 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ah[i].u = *(const uint4*)&sA[buf][0][wm * 64 + i * 16 + fc][fr * 8];
      al[i].u = *(const uint4*)&sA[buf][1][wm * 64 + i * 16 + fc][fr * 8];
      bh[i].u = *(const uint4*)&sB[buf][0][wn * 64 + i * 16 + fc][fr * 8];
      bl[i].u = *(const uint4*)&sB[buf][1][wn * 64 + i * 16 + fc][fr * 8];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i].v, bh[j].v, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i].v, bl[j].v, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i].v, bh[j].v, acc[i][j], 0, 0, 0);
      }
    if (has_next) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

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
  bf16x3_gemm_kernel<<<dim3((unsigned)tiles), 256, 0, s>>>(M, N, Kp, Ah, Al, Bh, Bl, C, ldc);
  CCX_HIP(hipGetLastError());
}

}  // namespace

void ladder_bf16x3(Ctx& c, const double* t1, const double* t2, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov();
  (void)o;
  const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
  const int64_t KpP = bf16_kpad(npP), KpM = bf16_kpad(npM);
  tau_pack_bf16x3(o, v, KpP, KpM, t1, t2, c.Tph, c.Tpl, c.Tmh, c.Tml, s);
  bf16x3_gemm(noP, npP + ov, KpP, c.Tph, c.Tpl, c.LPh, c.LPl, (float*)c.CP, npP + ov, s);
  if (noM > 0) bf16x3_gemm(noM, npM + ov, KpM, c.Tmh, c.Tml, c.LMh, c.LMl, (float*)c.CM, npM + ov, s);
}

}  // namespace ccx

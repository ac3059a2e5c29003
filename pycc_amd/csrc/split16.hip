// Split 16-bit ("x3") GEMMs for gfx950 and the conversions that feed them.
//
// An operand value x is carried as two 16-bit numbers, x*s = hi + lo + err with hi = E(x*s) and
// lo = E(x*s - hi), s a power-of-two scale of the whole operand. A product is taken as
// (hi_a hi_b + hi_a lo_b + lo_a hi_b) / (s_a s_b): three MFMAs accumulated in FP32, lo*lo dropped.
//  * E = _Float16 (11-bit significands): hi + lo hold 22 bits, so the dropped and rounded parts are
//    ~2^-22 |x|, four times FP32 rounding. s puts the operand's bound below 2^15 (FP16 max 65504);
//    values under 2^-14 (subnormal) keep an absolute error <= 2^-25, ~2^-40 of the bound.
//  * E = bf16 (8-bit significands, FP32 exponent range): 16 bits, ~2^-16 |x|; no scale needed.
// Both operands are K-contiguous ("NT": C[m][n] = sum_k A[m][k] B[n][k]) with row strides lda and
// ldb and K padded to a multiple of 32 with zeros.
#include "ccsd.h"
#include "split16.h"
#include <algorithm>
#include <cmath>

namespace ccx {

typedef float f4 __attribute__((ext_vector_type(4)));

template <typename E> struct X3;
template <> struct X3<__bf16> {
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct X3<_Float16> {
  typedef _Float16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

namespace {

constexpr int BK = 32, KPAD = 8;   // LDS row = 40 elements = 80 bytes
constexpr int GM = 8;              // M tiles per tile group (operand reuse in one XCD's L2)

// Hardware workgroup id (round-robin over the 8 XCDs) -> logical tile id, so that each XCD runs a
// contiguous range of logical tiles.
__device__ __forceinline__ int64_t xcd_remap_x3(int64_t w, int64_t ntiles) {
  const int64_t nx = 8;
  int64_t q = ntiles / nx, r = ntiles % nx;
  int64_t x = w % nx, j = w / nx;
  return x * q + (x < r ? x : r) + j;
}

// BT x BT output tiles, 4 waves in 2 x 2, each computing (BT/2) x (BT/2) as 16 x 16 MFMA blocks.
// Operand tiles (BT rows x 32 k, hi and lo) are double-buffered in LDS with one register stage:
// the next k-tile's global loads are issued before the current tile's MFMAs.
// MFMA operand layout: lane l holds A[l % 16][8 (l / 16) .. +7] and B[l % 16][8 (l / 16) .. +7];
// C register r of lane l is C[4 (l / 16) + r][l % 16].
template <typename E, int BT>
__global__ __launch_bounds__(256) void x3_gemm_kernel(X3Gemm p) {
  typedef typename X3<E>::v8 v8;
  constexpr int Q = BT / 64;     // rows per thread in one operand tile
  constexpr int WT = BT / 2;
  constexpr int NB = WT / 16;
  __shared__ __attribute__((aligned(16))) E sA[2][2][BT][BK + KPAD];   // [buf][hi/lo]
  __shared__ __attribute__((aligned(16))) E sB[2][2][BT][BK + KPAD];
  const E* __restrict__ Ah = (const E*)p.Ah;
  const E* __restrict__ Al = (const E*)p.Al;
  const E* __restrict__ Bh = (const E*)p.Bh;
  const E* __restrict__ Bl = (const E*)p.Bl;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t tiles_m = (p.M + BT - 1) / BT, tiles_n = (p.N + BT - 1) / BT;
  const int64_t t = xcd_remap_x3(blockIdx.x, tiles_m * tiles_n);
  const int64_t split = blockIdx.y;
  // groups of GM M tiles sweep all N tiles: consecutive tiles share B rows and GM A panels
  const int64_t group = t / (GM * tiles_n), first_m = group * GM;
  const int64_t gsize = tiles_m - first_m < GM ? tiles_m - first_m : GM;
  const int64_t within = t - group * GM * tiles_n;
  const int64_t tm = first_m + within % gsize, tn = within / gsize;

  const int lr = tid >> 2, lk = (tid & 3) * 8;
  int64_t offA[Q], offB[Q];
  bool okA[Q], okB[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int64_t m = tm * BT + lr + 64 * q, n = tn * BT + lr + 64 * q;
    okA[q] = m < p.M;
    okB[q] = n < p.N;
    // rows beyond the matrix read row 0 (valid memory) and are zeroed before the LDS store
    offA[q] = (okA[q] ? m : 0) * p.lda + lk;
    offB[q] = (okB[q] ? n : 0) * p.ldb + lk;
  }
  const v8 zero8 = v8{};
  v8 rA[2][Q], rB[2][Q];
  auto gload = [&](int64_t k0) {
    int64_t ka = k0, kb = k0;
    if (p.kb) {
      const int64_t q = k0 / p.kb, r = k0 - q * p.kb;
      ka = q * p.ska + r;
      kb = q * p.skb + r;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      rA[0][q] = *(const v8*)(Ah + offA[q] + ka);
      rA[1][q] = *(const v8*)(Al + offA[q] + ka);
      rB[0][q] = *(const v8*)(Bh + offB[q] + kb);
      rB[1][q] = *(const v8*)(Bl + offB[q] + kb);
    }
  };
  auto sstore = [&](int b) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        *(v8*)&sA[b][h][lr + 64 * q][lk] = okA[q] ? rA[h][q] : zero8;
        *(v8*)&sB[b][h][lr + 64 * q][lk] = okB[q] ? rB[h][q] : zero8;
      }
  };

  f4 acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int fr = lane >> 4, fc = lane & 15;
  const int64_t kbeg = p.work ? split * p.kchunk : 0;
  const int64_t kend = p.work ? (kbeg + p.kchunk < p.Kp ? kbeg + p.kchunk : p.Kp) : p.Kp;
  gload(kbeg);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    const bool has_next = k0 + BK < kend;
    if (has_next) gload(k0 + BK);
    v8 a[NB], b[NB];
    // lo(A) x hi(B)
#pragma unroll
    for (int i = 0; i < NB; ++i) a[i] = *(const v8*)&sA[buf][1][wm * WT + i * 16 + fc][fr * 8];
#pragma unroll
    for (int j = 0; j < NB; ++j) b[j] = *(const v8*)&sB[buf][0][wn * WT + j * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
    // hi(A) x hi(B)
#pragma unroll
    for (int i = 0; i < NB; ++i) a[i] = *(const v8*)&sA[buf][0][wm * WT + i * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
    // hi(A) x lo(B)
#pragma unroll
    for (int j = 0; j < NB; ++j) b[j] = *(const v8*)&sB[buf][1][wn * WT + j * 16 + fc][fr * 8];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
    if (has_next) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  if (p.work) {   // split-K: raw partial sums
    float* Wb = p.work + split * p.M * p.N;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t n = tn * BT + wn * WT + j * 16 + fc;
      if (n >= p.N) continue;
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t m = tm * BT + wm * WT + i * 16 + fr * 4 + r;
          if (m < p.M) Wb[m * p.N + n] = acc[i][j][r];
        }
    }
    return;
  }
  float f0 = p.alpha;
  if (p.sa) f0 = f0 / p.sa[0];
  if (p.sb) f0 = f0 / p.sb[0];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int64_t n = tn * BT + wn * WT + j * 16 + fc;
    if (n >= p.N) continue;
    const float f = p.sbn ? f0 / p.sbn[n] : f0;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = tm * BT + wm * WT + i * 16 + fr * 4 + r;
        if (m < p.M) {
          if (p.C64) {   // FP64 destination without split K
            const int64_t at = m * p.ldc + n;
            double val = (double)(f * acc[i][j][r]);
            if (p.beta != 0.f) val += (double)p.beta * (p.C64in ? p.C64in[at] : p.C64[at]);
            if (p.D64) val += p.D64[at] * p.dvec[n];
            p.C64[at] = val;
          } else {
            float* d = p.C + m * p.ldc + n;
            float val = f * acc[i][j][r];
            if (p.beta != 0.f) val += p.beta * *d;
            *d = val;
          }
        }
      }
  }
}

// Skinny split GEMM for M <= 64 (the t1 products of <ov|vv> and the :759 term): HBM-bound on B.
// One workgroup of 4 waves covers 64 rows (all of A) x 64 NBN columns; wave w owns the 16 NBN
// columns [16 NBN w, 16 NBN (w + 1)) of the tile. K goes in steps of 64: A's 64 x 64 k-tile (hi and
// lo) is staged through LDS (144-byte rows: conflict-free 16-byte reads), and B goes from global
// memory straight into registers, the next step's loads in flight during the current step's MFMAs.
// Lane l (row l % 16 of a 16-row block, group g = l / 16) holds the 16 k values 16 g .. 16 g + 15
// of the step as two 8-value halves u = 0, 1: one 32-byte run, so each row's 128-byte line of a
// 64-k step is read whole. MFMA u of a step sums k = 16 g + 8 u + [0, 8) over the four groups;
// A's fragments use the same k assignment, so the two MFMAs together cover the step's 64 k.
constexpr int SK_STEP = 64, SK_PITCH = SK_STEP + 8;
template <typename E, int NBN>
__global__ __launch_bounds__(256) void x3_skinny_kernel(X3Gemm p) {
  typedef typename X3<E>::v8 v8;
  __shared__ __attribute__((aligned(16))) E sA[2][2][64 * SK_PITCH];   // [stage][hi/lo]
  const E* __restrict__ Ah = (const E*)p.Ah;
  const E* __restrict__ Al = (const E*)p.Al;
  const E* __restrict__ Bh = (const E*)p.Bh;
  const E* __restrict__ Bl = (const E*)p.Bl;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int fr = lane >> 4, fc = lane & 15;
  const int64_t n0 = (int64_t)blockIdx.x * (64 * NBN) + wave * (16 * NBN);
  const int64_t kbeg = p.work ? (int64_t)blockIdx.y * p.kchunk : 0;
  const int64_t kend = p.work ? (kbeg + p.kchunk < p.Kp ? kbeg + p.kchunk : p.Kp) : p.Kp;
  const int nbm = (int)((p.M + 15) / 16);   // valid 16-row blocks of A
  // offset of k in a row (blocked K: k = q kb + r at q sk + r; kb is a multiple of 16, so a run of
  // 16 k starting at a multiple of 16 stays in one block)
  auto koff = [&](int64_t k, int64_t sk) -> int64_t {
    if (!p.kb) return k;
    const int64_t q = k / p.kb;
    return q * sk + (k - q * p.kb);
  };
  // B rows of this lane (columns past N read row N - 1 and are not stored)
  int64_t offB[NBN];
#pragma unroll
  for (int j = 0; j < NBN; ++j) {
    const int64_t n = n0 + 16 * j + fc;
    offB[j] = (n < p.N ? n : p.N - 1) * p.ldb;
  }
  // A: thread t stages 16 bytes (k chunk t % 8 of the step) of rows t / 8 and t / 8 + 32
  const int ar = tid >> 3, ac = tid & 7;
  int64_t offA[2];
  bool okA[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t m = ar + 32 * h;
    okA[h] = m < p.M;
    offA[h] = (okA[h] ? m : 0) * p.lda;
  }
  const v8 zero8 = v8{};
  v8 ra[2][2];   // [row half][hi/lo]
  auto aload = [&](int64_t k0) {
    const int64_t k = k0 + 8 * ac;
    const bool kok = k < kend;
    const int64_t o = kok ? koff(k, p.ska) : 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool ok = okA[h] && kok;
      ra[h][0] = ok ? *(const v8*)(Ah + offA[h] + o) : zero8;
      ra[h][1] = ok ? *(const v8*)(Al + offA[h] + o) : zero8;
    }
  };
  auto astore = [&](int st) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      *(v8*)&sA[st][0][(ar + 32 * h) * SK_PITCH + 8 * ac] = ra[h][0];
      *(v8*)&sA[st][1][(ar + 32 * h) * SK_PITCH + 8 * ac] = ra[h][1];
    }
  };
  v8 rb0[NBN][2][2], rb1[NBN][2][2];   // two register stages: [block][hi/lo][half]
  auto bload = [&](v8 (&rb)[NBN][2][2], int64_t k0) {
    const int64_t k = k0 + 16 * fr;
    const bool kok = k < kend;
    const int64_t o = kok ? koff(k, p.skb) : 0;
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        rb[j][0][u] = kok ? *(const v8*)(Bh + offB[j] + o + 8 * u) : zero8;
        rb[j][1][u] = kok ? *(const v8*)(Bl + offB[j] + o + 8 * u) : zero8;
      }
  };
  f4 acc[4][NBN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  auto compute = [&](int st, const v8 (&rb)[NBN][2][2]) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= nbm) break;   // wave-uniform
        const int at = (16 * i + fc) * SK_PITCH + 16 * fr + 8 * u;
        const v8 ah = *(const v8*)&sA[st][0][at];
        const v8 al = *(const v8*)&sA[st][1][at];
#pragma unroll
        for (int j = 0; j < NBN; ++j) {
          acc[i][j] = X3<E>::mfma(al, rb[j][0][u], acc[i][j]);
          acc[i][j] = X3<E>::mfma(ah, rb[j][0][u], acc[i][j]);
          acc[i][j] = X3<E>::mfma(ah, rb[j][1][u], acc[i][j]);
        }
      }
  };
  aload(kbeg);
  bload(rb0, kbeg);
  astore(0);
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += 2 * SK_STEP) {
    {
      const bool nxt = k0 + SK_STEP < kend;
      if (nxt) {
        aload(k0 + SK_STEP);
        bload(rb1, k0 + SK_STEP);
      }
      compute(0, rb0);
      if (nxt) astore(1);
      __syncthreads();
      if (!nxt) break;
    }
    {
      const bool nxt = k0 + 2 * SK_STEP < kend;
      if (nxt) {
        aload(k0 + 2 * SK_STEP);
        bload(rb0, k0 + 2 * SK_STEP);
      }
      compute(1, rb1);
      if (nxt) astore(0);
      __syncthreads();
    }
  }

  // C register r of lane l in block (i, j) is C[16 i + 4 (l / 16) + r][n0 + 16 j + l % 16]
  if (p.work) {   // split-K: raw partial sums
    float* Wb = p.work + (int64_t)blockIdx.y * p.M * p.N;
#pragma unroll
    for (int j = 0; j < NBN; ++j) {
      const int64_t n = n0 + 16 * j + fc;
      if (n >= p.N) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t m = 16 * i + 4 * fr + r;
          if (m < p.M) Wb[m * p.N + n] = acc[i][j][r];
        }
    }
    return;
  }
  float f0 = p.alpha;
  if (p.sa) f0 = f0 / p.sa[0];
  if (p.sb) f0 = f0 / p.sb[0];
#pragma unroll
  for (int j = 0; j < NBN; ++j) {
    const int64_t n = n0 + 16 * j + fc;
    if (n >= p.N) continue;
    const float f = p.sbn ? f0 / p.sbn[n] : f0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = 16 * i + 4 * fr + r;
        if (m < p.M) {
          if (p.C64) {
            const int64_t at = m * p.ldc + n;
            double val = (double)(f * acc[i][j][r]);
            if (p.beta != 0.f) val += (double)p.beta * (p.C64in ? p.C64in[at] : p.C64[at]);
            if (p.D64) val += p.D64[at] * p.dvec[n];
            p.C64[at] = val;
          } else {
            float* d = p.C + m * p.ldc + n;
            float val = f * acc[i][j][r];
            if (p.beta != 0.f) val += p.beta * *d;
            *d = val;
          }
        }
      }
  }
}

// 256 x 256 output tiles for large GEMMs, 8 waves (2 per SIMD) in 2 x 4, each computing 128 x 64.
// Operand k-tiles (256 rows x 32 k, hi and lo) go from global memory straight into LDS
// (global_load_lds_dwordx4, issued from inline assembly so that the only waits are the explicit
// ones below) and are double-buffered: the next tile's loads are in flight while the current one
// feeds the MFMAs. LDS rows are 32 elements (64 bytes) with the 16-byte chunk c of row r stored at
// chunk c ^ ((r >> 1) & 3), which keeps both the DMA writes and the fragment reads free of bank
// conflicts. An LDS-DMA instruction writes 16 bytes per lane at M0 + 16 lane, so lane l of an
// instruction covering rows R0 .. R0 + 15 loads row R0 + l / 4, chunk (l % 4) ^ ((row >> 1) & 3).
// Rows past M / N load a valid row (their results are not stored).
__device__ __forceinline__ int x3_swz(int r, int c) { return r * 32 + ((c ^ ((r >> 1) & 3)) * 8); }

template <typename E>
__device__ __forceinline__ void x3_dma16(const E* g, const E* lds) {
  const uint32_t base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const E*)lds;
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"
               :: "s"(__builtin_amdgcn_readfirstlane(base)), "v"(g) : "memory");
}

// BN = 192, 256 or 320 output columns per tile (320: 128 x 80 per wave, 144 KiB of LDS; B's 20 row
// blocks of 16 go to wave q % 8, so waves 0-3 load three and 4-7 two); BT = 256 or 384 rows (384
// with BN = 192: 192 x 48 per wave, 144 KiB). KCAT: the K-concatenated, batched form of
// x3_gemm_kcat (interleaved operands, FP32 C, no split K).
// WGM: waves along M (2: 2 x 4 waves; 4: 4 x 2, e.g. 96 x 96 per wave for 384 x 192 tiles, which
// holds fewer fragment registers than 192 x 48).
template <typename E, int BN, bool KCAT = false, int BT = 256, bool YT = false, int WGM = 2>
__global__ __launch_bounds__(512) void x3_gemm_dma_kernel(X3Gemm p) {
  typedef typename X3<E>::v8 v8;
  constexpr int WGN = 8 / WGM;
  constexpr int WTM = BT / WGM, WTN = BN / WGN, NBM = WTM / 16, NBN = WTN / 16, IPW = BT / 128;
  constexpr int NIB = BN / 16, IPB = (NIB + 7) / 8;   // B row blocks, per wave (rounded up)
  __shared__ __attribute__((aligned(1024))) E sA[2][2][BT * 32];   // [stage][hi/lo]
  __shared__ __attribute__((aligned(1024))) E sB[2][2][BN * 32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int64_t tiles_m = (p.M + BT - 1) / BT, tiles_n = (p.N + BN - 1) / BN;
  int64_t t = xcd_remap_x3(blockIdx.x, tiles_m * tiles_n * (KCAT ? p.nbatch : 1));
  int64_t batch = 0;
  if constexpr (KCAT) {
    batch = t / (tiles_m * tiles_n);
    t -= batch * tiles_m * tiles_n;
  }
  const E* __restrict__ Ah = (const E*)p.Ah;
  const E* __restrict__ Al = (const E*)p.Al;
  const E* __restrict__ Bh = (const E*)p.Bh + batch * p.sB;
  const E* __restrict__ Bl = (const E*)p.Bl + batch * p.sB;
  const E* __restrict__ Ah2 = KCAT ? (const E*)p.Ah2 + batch * p.sA2 : nullptr;
  const int64_t split = blockIdx.y;
  const int64_t group = t / (GM * tiles_n), first_m = group * GM;
  const int64_t gsize = tiles_m - first_m < GM ? tiles_m - first_m : GM;
  const int64_t within = t - group * GM * tiles_n;
  const int64_t tm = first_m + within % gsize, tn = within / gsize;

  int64_t gA[IPW], gB[IPB];
  int gA2[KCAT ? IPW : 1];   // M lda < 2^31 (checked by x3_gemm_kcat)
  int ldso[IPW], ldsb[IPB];
#pragma unroll
  for (int i = 0; i < IPW; ++i) {
    const int R0 = (wave * IPW + i) * 16;
    const int row = R0 + (lane >> 2), c = (lane & 3) ^ ((row >> 1) & 3);
    int64_t m = tm * BT + row;
    m = m < p.M ? m : p.M - 1;
    gA[i] = m * p.lda + c * 8;
    if constexpr (KCAT) gA2[i] = (int)(((m % p.sw) * p.sw + m / p.sw) * p.lda + c * 8);
    ldso[i] = R0 * 32;
  }
#pragma unroll
  for (int i = 0; i < IPB; ++i) {
    const int q = wave + 8 * i;
    const int R0 = (q < NIB ? q : 0) * 16;   // blocks past NIB are not loaded
    const int row = R0 + (lane >> 2), c = (lane & 3) ^ ((row >> 1) & 3);
    int64_t n = tn * BN + row;
    n = n < p.N ? n : p.N - 1;
    gB[i] = n * p.ldb + c * 8;
    ldsb[i] = R0 * 32;
  }
  auto issue = [&](int st, int64_t k0) {
    int64_t ka = k0, kb = k0;
    if (p.kb == BK) {   // interleaved operands and other 32-blocked layouts
      ka = (k0 >> 5) * p.ska;
      kb = (k0 >> 5) * p.skb;
    } else if (p.kb) {
      const int64_t q = k0 / p.kb, r = k0 - q * p.kb;
      ka = q * p.ska + r;
      kb = q * p.skb + r;
    }
    if (KCAT && k0 >= p.K1) {   // second operand of the concatenation (wave-uniform)
      const int64_t ka2 = ((k0 - p.K1) >> 5) * p.ska;
#pragma unroll
      for (int i = 0; i < IPW; ++i) {
        x3_dma16<E>(Ah2 + gA2[i] + ka2, &sA[st][0][ldso[i]]);
        x3_dma16<E>(Ah2 + 32 + gA2[i] + ka2, &sA[st][1][ldso[i]]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < IPW; ++i) {
        x3_dma16<E>(Ah + gA[i] + ka, &sA[st][0][ldso[i]]);
        x3_dma16<E>(Al + gA[i] + ka, &sA[st][1][ldso[i]]);
      }
    }
#pragma unroll
    for (int i = 0; i < IPB; ++i) {
      if (wave + 8 * i < NIB) {   // wave-uniform
        x3_dma16<E>(Bh + gB[i] + kb, &sB[st][0][ldsb[i]]);
        x3_dma16<E>(Bl + gB[i] + kb, &sB[st][1][ldsb[i]]);
      }
    }
  };
  const int fr = lane >> 4, fc = lane & 15;
  int ra[NBM], rb[NBN];
#pragma unroll
  for (int i = 0; i < NBM; ++i) ra[i] = x3_swz(wm * WTM + i * 16 + fc, fr);
#pragma unroll
  for (int j = 0; j < NBN; ++j) rb[j] = x3_swz(wn * WTN + j * 16 + fc, fr);
  f4 acc[NBM][NBN];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int64_t kbeg = p.work ? split * p.kchunk : 0;
  const int64_t kend = p.work ? (kbeg + p.kchunk < p.Kp ? kbeg + p.kchunk : p.Kp) : p.Kp;
  issue(0, kbeg);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    if (k0 + BK < kend) issue(buf ^ 1, k0 + BK);
    v8 a[NBM], b[NBN];
    // lo(A) x hi(B), hi(A) x hi(B), hi(A) x lo(B)
#pragma unroll
    for (int i = 0; i < NBM; ++i) a[i] = *(const v8*)&sA[buf][1][ra[i]];
#pragma unroll
    for (int j = 0; j < NBN; ++j) b[j] = *(const v8*)&sB[buf][0][rb[j]];
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < NBM; ++i) a[i] = *(const v8*)&sA[buf][0][ra[i]];
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
#pragma unroll
    for (int j = 0; j < NBN; ++j) b[j] = *(const v8*)&sB[buf][1][rb[j]];
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X3<E>::mfma(a[i], b[j], acc[i][j]);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wave's loads of the next k-tile
    __syncthreads();                                    // everyone's loads landed, reads of buf done
    buf ^= 1;
  }

  if (p.work) {   // split-K: raw partial sums
    float* Wb = p.work + split * p.M * p.N;
#pragma unroll
    for (int j = 0; j < NBN; ++j) {
      const int64_t n = tn * BN + wn * WTN + j * 16 + fc;
      if (n >= p.N) continue;
#pragma unroll
      for (int i = 0; i < NBM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t m = tm * BT + wm * WTM + i * 16 + fr * 4 + r;
          if (m < p.M) Wb[m * p.N + n] = acc[i][j][r];
        }
    }
    return;
  }
  float f0 = p.alpha;
  if (p.sa) f0 = f0 / p.sa[0];
  if (p.sb) f0 = f0 / p.sb[0];
  if constexpr (KCAT) {   // FP32 C, beta = 0
    float* Cb = p.C + batch * p.sC;
#pragma unroll
    for (int j = 0; j < NBN; ++j) {
      const int64_t nb = tn * BN + wn * WTN + j * 16;   // 16-aligned block of n
      if (nb + fc >= p.N) continue;
#pragma unroll
      for (int i = 0; i < NBM; ++i) {
        const int64_t mb = tm * BT + wm * WTM + i * 16;   // 16-aligned block of m
        if constexpr (YT) {   // x3_gemm_kcat's ytile layout (M, N multiples of 8): the lane's four
          // rows m = mb + 4 fr .. + 3 of column n are 16 contiguous bytes
          if (mb >= p.M) continue;
          typedef float f4v __attribute__((ext_vector_type(4)));
          *(f4v*)(Cb + (mb + 8 * (fr >> 1)) * p.ytile + 8 * (nb + fc) + 4 * (fr & 1)) =
              f4v{f0 * acc[i][j][0], f0 * acc[i][j][1], f0 * acc[i][j][2], f0 * acc[i][j][3]};
        } else {
          float* d = Cb + mb * p.ldc + nb + fc;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (mb + fr * 4 + r < p.M) d[(fr * 4 + r) * p.ldc] = f0 * acc[i][j][r];
        }
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < NBN; ++j) {
    const int64_t n = tn * BN + wn * WTN + j * 16 + fc;
    if (n >= p.N) continue;
    const float f = p.sbn ? f0 / p.sbn[n] : f0;
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = tm * BT + wm * WTM + i * 16 + fr * 4 + r;
        if (m < p.M) {
          if (p.C64) {   // FP64 destination without split K
            const int64_t at = m * p.ldc + n;
            double val = (double)(f * acc[i][j][r]);
            if (p.beta != 0.f) val += (double)p.beta * (p.C64in ? p.C64in[at] : p.C64[at]);
            if (p.D64) val += p.D64[at] * p.dvec[n];
            p.C64[at] = val;
          } else {
            float* d = p.C + m * p.ldc + n;
            float val = f * acc[i][j][r];
            if (p.beta != 0.f) val += p.beta * *d;
            *d = val;
          }
        }
      }
  }
}

// C = alpha / (sa sb sbn[n]) sum_s work[s] + beta C, summed in FP64 in a fixed order
// (deterministic); written to C64 (FP64) when given, else to C.
__global__ __launch_bounds__(256) void x3_reduce_kernel(X3Gemm p, int splits) {
  double f0 = p.alpha;
  if (p.sa) f0 = f0 / p.sa[0];
  if (p.sb) f0 = f0 / p.sb[0];
  const int64_t total = p.M * p.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < splits; ++k) acc += (double)p.work[k * total + e];
    const int64_t m = e / p.N, n = e - m * p.N;
    const double f = p.sbn ? f0 / p.sbn[n] : f0;
    double val = f * acc;
    if (p.C64) {
      double* d = p.C64 + m * p.ldc + n;
      if (p.beta != 0.f) val += (double)p.beta * *d;
      *d = val;
    } else {
      float* d = p.C + m * p.ldc + n;
      if (p.beta != 0.f) val += (double)p.beta * (double)*d;
      *d = (float)val;
    }
  }
}

// ---- operand conversions ------------------------------------------------------------------
constexpr int X3_RED_BLOCKS = 1024;

// 2^k with bound * 2^k < 2^15 (k clamped to [-40, 60]).
__device__ __forceinline__ float pow2_scale(float bound) {
  int e = 0;
  frexpf(bound, &e);
  int k = 15 - e;
  k = k < -40 ? -40 : (k > 60 ? 60 : k);
  return ldexpf(1.0f, k);
}

__device__ __forceinline__ float block_max(float x) {
  __shared__ float sh[256];
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  float r = sh[0];
  __syncthreads();
  return r;
}

// max |x[r * ld + c]| over r < rows, c < cols, one partial per workgroup
template <typename T>
__global__ __launch_bounds__(256) void absmax_kernel(const T* __restrict__ x, int64_t rows, int64_t cols, int64_t ld,
                                                     float* __restrict__ part) {
  double m = 0.0;
  const int64_t n = rows * cols;
  if (ld == cols) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
      m = fmax(m, fabs((double)x[e]));
  } else {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
      const int64_t r = e / cols;
      m = fmax(m, fabs((double)x[r * ld + (e - r * cols)]));
    }
  }
  float r = block_max((float)m);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// s = 2^(15 - e) with bound = ca max|a| + cb max|b|^2 < 2^e: bound * s < 2^15.
__global__ __launch_bounds__(256) void x3_scale_kernel(const float* __restrict__ pa, int na, float ca,
                                                       const float* __restrict__ pb, int nb, float cb,
                                                       float* __restrict__ s_out) {
  float ma = 0.f, mb = 0.f;
  for (int k = threadIdx.x; k < na; k += blockDim.x) ma = fmaxf(ma, pa[k]);
  for (int k = threadIdx.x; k < nb; k += blockDim.x) mb = fmaxf(mb, pb[k]);
  ma = block_max(ma);
  mb = block_max(mb);
  const float sc = pow2_scale(ca * ma + cb * mb * mb);
  if (threadIdx.x == 0) s_out[0] = sc;
}

template <typename E>
__device__ __forceinline__ void split_store(float x, E* hi, E* lo, int64_t at) {
  const E h = (E)x;
  hi[at] = h;
  lo[at] = (E)(x - (float)h);
}

// Interleaved ("IL") split operands: a row of K (padded to a multiple of 32) is stored as 2 K
// elements, k-block q of 32 as hi at 64 q and lo at 64 q + 32, so that one 128-byte line carries
// both halves of a row's k-tile. With hi = base and lo = base + 32, element `at` = row * ld + k
// (ld a multiple of 32) of the plain layout lives at il_at(at) in both.
__device__ __forceinline__ int64_t il_at(int64_t at) { return ((at >> 5) << 6) | (at & 31); }

template <typename Tin, typename E>
__global__ __launch_bounds__(256) void split_rows_kernel(const Tin* __restrict__ src, int64_t rows, int64_t cols,
                                                         int64_t lds, const float* __restrict__ sc, E* hi, E* lo,
                                                         int64_t ldd, bool il) {
  const double s = sc[0];
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / cols, c = e - r * cols;
    const int64_t at = r * ldd + c;
    split_store<E>((float)((double)src[r * lds + c] * s), hi, lo, il ? il_at(at) : at);
  }
}

// dst[cc * ldd + x * Kb + r] = split(src[r * lds + x * Cb + cc]) for r < R, x < nx, cc < Cb,
// through 32 x 33 LDS tiles (coalesced reads along cc, writes along r).
template <typename Tin, typename E>
__global__ __launch_bounds__(256) void split_t_kernel(const Tin* __restrict__ src, int64_t R, int64_t Cb, int nx,
                                                      int64_t lds, const float* __restrict__ sc, E* hi, E* lo,
                                                      int64_t Kb, int64_t ldd, bool il) {
  __shared__ float tile[32][33];
  const double s = sc[0];
  const int64_t tr = (R + 31) / 32, tc = (Cb + 31) / 32;
  const int64_t per_x = tr * tc;
  for (int64_t w = blockIdx.x; w < per_x * nx; w += gridDim.x) {
    const int64_t x = w / per_x, rem = w - x * per_x;
    const int64_t r0 = (rem / tc) * 32, c0 = (rem % tc) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int y = ty; y < 32; y += 8) {
      const int64_t r = r0 + y, cc = c0 + tx;
      tile[y][tx] = (r < R && cc < Cb) ? (float)((double)src[r * lds + x * Cb + cc] * s) : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      const int64_t cc = c0 + y, r = r0 + tx;
      const int64_t at = cc * ldd + x * Kb + r;
      if (cc < Cb && r < R) split_store<E>(tile[tx][y], hi, lo, il ? il_at(at) : at);
    }
    __syncthreads();
  }
}

// Split ring operands straight from t1, t2 (layouts of the FP32 ring_prep, regrouped so that
// every ring GEMM is K-contiguous on both sides; T = t2_ijab, X = t2_ijba, P = t_ja t_ib at
// (ia, jb), ov = o v, Kp = ovp):
//   BJ[x][(jb)][(ia)]  x = 0: -T/2, x = 1: X/2 + P     (B of the J GEMM, N = (x, jb), K = ia)
//   BK[(jb)][(ia)]     T - X/2 - P                      (B of the K GEMM)
//   AP[(ia)][x][(jb)]  x = 0: 2T - X, x = 1: T          (A of the ring GEMM, K = (x, me))
//   AQ[(ia)][(jb)]     X                                (A of the crossed ring GEMM)
//   Ud[(ia)][(jb)]     2T - X in FP64
// The split operands are written interleaved (il_at; the lo pointers are hi + 32).
template <typename E>
__global__ __launch_bounds__(256) void ring_prep_x3_kernel(int64_t o, int64_t v, int64_t ovp,
                                                           const double* __restrict__ t1,
                                                           const double* __restrict__ t2,
                                                           const float* __restrict__ sc, E* BJh, E* BJl,
                                                           E* BKh, E* BKl, E* APh, E* APl, E* AQh, E* AQl,
                                                           double* __restrict__ Ud) {
  __shared__ double tn[32][33];
  __shared__ double tx[32][33];
  const double s = sc[0];
  const int64_t ij = blockIdx.y;
  const int64_t i = ij / o, j = ij % o;
  const int64_t tiles_b = (v + 31) / 32;
  const int64_t a0 = (blockIdx.x / tiles_b) * 32, b0 = (blockIdx.x % tiles_b) * 32;
  const double* blk = t2 + ij * v * v;
  const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
  for (int y = y0; y < 32; y += 8) {
    int64_t a = a0 + y, b = b0 + x;
    tn[y][x] = (a < v && b < v) ? blk[a * v + b] : 0.0;   // T(a0 + y, b0 + x)
    int64_t bb = b0 + y, aa = a0 + x;
    tx[y][x] = (bb < v && aa < v) ? blk[bb * v + aa] : 0.0;   // X(a0 + x, b0 + y)
  }
  __syncthreads();
  const int64_t ov = o * v;
  // b fastest: AP, AQ, Ud
  for (int y = y0; y < 32; y += 8) {
    const int64_t a = a0 + y, b = b0 + x;
    if (a >= v || b >= v) continue;
    const double T = tn[y][x], X = tx[x][y];
    const int64_t ia = i * v + a, jb = j * v + b;
    split_store<E>((float)((2.0 * T - X) * s), APh, APl, il_at(ia * 2 * ovp + jb));
    split_store<E>((float)(T * s), APh, APl, il_at(ia * 2 * ovp + ovp + jb));
    split_store<E>((float)(X * s), AQh, AQl, il_at(ia * ovp + jb));
    Ud[ia * ov + jb] = 2.0 * T - X;
  }
  // a fastest: BJ, BK
  for (int y = y0; y < 32; y += 8) {
    const int64_t a = a0 + x, b = b0 + y;
    if (a >= v || b >= v) continue;
    const double T = tn[x][y], X = tx[y][x];
    const double P = t1[j * v + a] * t1[i * v + b];
    const int64_t ia = i * v + a, jb = j * v + b;
    split_store<E>((float)(-0.5 * T * s), BJh, BJl, il_at(jb * ovp + ia));
    split_store<E>((float)((0.5 * X + P) * s), BJh, BJl, il_at((ov + jb) * ovp + ia));
    split_store<E>((float)((T - 0.5 * X - P) * s), BKh, BKl, il_at(jb * ovp + ia));
  }
}

// ---- packed ladder operands (layouts of the BF16x3 ones in kernels.hip) ----------------------
__device__ __forceinline__ void decode_pair(int64_t c, int64_t& a, int64_t& b) {
  int64_t f = (int64_t)((sqrt(8.0 * (double)c + 1.0) - 1.0) * 0.5);
  while (f * (f + 1) / 2 > c) --f;
  while ((f + 1) * (f + 2) / 2 <= c) ++f;
  b = f;
  a = c - f * (f + 1) / 2;
}

// One workgroup per row (e <= f) of the (ab) block: row maxima first, then the scaled split values
// 1/2(<ab|ef> + <ab|fe>) (row pairP(e, f) of LP) and, for e < f, 1/2(<ab|ef> - <ab|fe>) (row
// pairM(e, f) of LM, columns a < b).
template <typename E>
__global__ __launch_bounds__(256) void ladder_pack_x3_kernel(int64_t v, int64_t e0, VVVVSrc src, E* LPh, E* LPl,
                                                             E* LMh, E* LMl, float* sP, float* sM, int64_t KpP,
                                                             int64_t KpM) {
  const int64_t npP = v * (v + 1) / 2;
  const int64_t e = e0 + blockIdx.y / v, f = blockIdx.y % v;
  if (f < e) return;
  const int64_t rowP = pairP(e, f);
  const bool hasM = e < f;
  const int64_t rowM = hasM ? pairM(e, f) : 0;
  auto vals = [&](int64_t c, double& x, double& y, bool& mcol, int64_t& cm) {
    int64_t a, b;
    decode_pair(c, a, b);
    x = src.base[e * src.se + a * src.sa + f * src.sf + b * src.sb];   // <ab|ef> = (ea|fb)
    y = src.base[e * src.se + b * src.sa + f * src.sf + a * src.sb];   // <ab|fe> = (eb|fa)
    mcol = a < b;
    cm = mcol ? pairM(a, b) : 0;
  };
  double mp = 0.0, mm = 0.0;
  for (int64_t c = threadIdx.x; c < npP; c += blockDim.x) {
    double x, y;
    bool mcol;
    int64_t cm;
    vals(c, x, y, mcol, cm);
    mp = fmax(mp, fabs(0.5 * (x + y)));
    if (mcol) mm = fmax(mm, fabs(0.5 * (x - y)));
  }
  const float scP = pow2_scale(block_max((float)mp));
  const float scM = pow2_scale(block_max((float)mm));
  if (threadIdx.x == 0) {
    sP[rowP] = scP;
    if (hasM) sM[rowM] = scM;
  }
  for (int64_t c = threadIdx.x; c < npP; c += blockDim.x) {
    double x, y;
    bool mcol;
    int64_t cm;
    vals(c, x, y, mcol, cm);
    split_store<E>((float)(0.5 * (x + y) * scP), LPh, LPl, il_at(rowP * KpP + c));
    if (hasM && mcol) split_store<E>((float)(0.5 * (x - y) * scM), LMh, LMl, il_at(rowM * KpM + cm));
  }
}

// One workgroup per (m, b): row npP + mb of LP holds 1/2(<mb|ef> + <mb|fe>) over (e <= f), row
// npM + mb of LM holds 1/2(<mb|ef> - <mb|fe>) over (e < f).
template <typename E>
__global__ __launch_bounds__(256) void ladder_pack_o_x3_kernel(int64_t o, int64_t v, const double* __restrict__ Covvv,
                                                               E* LPh, E* LPl, E* LMh, E* LMl, float* sP, float* sM) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  const int64_t KpP = bf16_kpad(npP), KpM = bf16_kpad(npM);
  const int64_t mb = blockIdx.y, m = mb / v, b = mb % v;
  auto vals = [&](int64_t k, double& x, double& y, int64_t& ef_m, bool& lt) {
    int64_t e, f;
    decode_pair(k, e, f);
    x = Covvv[((m * v + e) * v + b) * v + f];   // <mb|ef> = (me|bf)
    y = Covvv[((m * v + f) * v + b) * v + e];   // <mb|fe> = (mf|be)
    lt = e < f;
    ef_m = lt ? pairM(e, f) : 0;
  };
  double mp = 0.0, mm = 0.0;
  for (int64_t k = threadIdx.x; k < npP; k += blockDim.x) {
    double x, y;
    int64_t q;
    bool lt;
    vals(k, x, y, q, lt);
    mp = fmax(mp, fabs(0.5 * (x + y)));
    if (lt) mm = fmax(mm, fabs(0.5 * (x - y)));
  }
  const float scP = pow2_scale(block_max((float)mp));
  const float scM = pow2_scale(block_max((float)mm));
  if (threadIdx.x == 0) {
    sP[npP + mb] = scP;
    sM[npM + mb] = scM;
  }
  for (int64_t k = threadIdx.x; k < npP; k += blockDim.x) {
    double x, y;
    int64_t q;
    bool lt;
    vals(k, x, y, q, lt);
    split_store<E>((float)(0.5 * (x + y) * scP), LPh, LPl, il_at((npP + mb) * KpP + k));
    if (lt) split_store<E>((float)(0.5 * (x - y) * scM), LMh, LMl, il_at((npM + mb) * KpM + q));
  }
}

// Row (i <= j) of the packed tau+ (and, for i < j, tau-) amplitudes, scaled by *sc.
template <typename E>
__global__ __launch_bounds__(256) void tau_pack_x3_kernel(int64_t o, int64_t v, int64_t KpP, int64_t KpM,
                                                          const double* __restrict__ t1,
                                                          const double* __restrict__ t2,
                                                          const float* __restrict__ sc, E* Tph, E* Tpl, E* Tmh,
                                                          E* Tml) {
  const int64_t npP = v * (v + 1) / 2;
  const double s = sc[0];
  int64_t i, j;
  decode_pair(blockIdx.y, i, j);
  const int64_t rowP = blockIdx.y;
  const bool hasM = i < j;
  const int64_t rowM = hasM ? pairM(i, j) : 0;
  const double* blk = t2 + (i * o + j) * v * v;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npP; c += (int64_t)gridDim.x * blockDim.x) {
    int64_t e, f;
    decode_pair(c, e, f);
    const double tef = blk[e * v + f] + (t1 ? t1[i * v + e] * t1[j * v + f] : 0.0);   // t1 null: tau = t2
    const double tfe = blk[f * v + e] + (t1 ? t1[i * v + f] * t1[j * v + e] : 0.0);
    split_store<E>((float)((e < f ? tef + tfe : tef) * s), Tph, Tpl, il_at(rowP * KpP + c));
    if (hasM && e < f) split_store<E>((float)((tef - tfe) * s), Tmh, Tml, il_at(rowM * KpM + pairM(e, f)));
  }
}

inline int grid_x3(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

}  // namespace

template <typename E>
void x3_gemm(const X3Gemm& p0, hipStream_t s, Workspace* ws) {
  X3Gemm p = p0;
  if (p.M == 0 || p.N == 0) return;
  CCX_CHECK(p.Kp % BK == 0 && p.lda % 8 == 0 && p.ldb % 8 == 0, ERR_INTERNAL, "x3 GEMM operand padding");
  CCX_CHECK(p.kb ? (p.kb % BK == 0 && p.ska % 8 == 0 && p.skb % 8 == 0) : (p.lda >= p.Kp && p.ldb >= p.Kp),
            ERR_INTERNAL, "x3 GEMM K layout");
  if (p.Kp == 0) {   // empty contraction: C = beta C
    scale<float>(p.beta, tvs(p.C, "mn", {p.M, p.N}, {p.ldc, 1}), s);
    return;
  }
  // two 16-column blocks per wave (128 columns per workgroup): 158 registers, three waves per SIMD
  constexpr int nbn = 2;
  const int64_t sk_tiles = ceil_div(p.N, 64 * nbn);
  if (p.M <= 64 && (p.kb == 0 || p.kb % 16 == 0) && !p.C64in && !p.D64 && (!p.C64 || ws || sk_tiles >= 512)) {
    // skinny: all of A in one 64-row tile
    const int64_t tiles = sk_tiles;
    CCX_CHECK(tiles < (1ll << 31), ERR_ARG, "x3 GEMM grid too large");
    int64_t splits = 1;
    p.work = nullptr;
    if (ws && tiles < 512) {   // split K (at least 8 steps of 64 each) to about 1024 workgroups
      splits = std::max<int64_t>(1, std::min<int64_t>(1024 / tiles, p.Kp / (8 * SK_STEP)));
      if (splits > 1 || p.C64) {
        p.kchunk = ceil_div(ceil_div(p.Kp, splits), SK_STEP) * SK_STEP;
        splits = ceil_div(p.Kp, p.kchunk);
        p.work = (float*)ws->get((size_t)splits * p.M * p.N * sizeof(float));
      }
    }
    dim3 grid((unsigned)tiles, (unsigned)splits);
    x3_skinny_kernel<E, nbn><<<grid, 256, 0, s>>>(p);
    CCX_HIP(hipGetLastError());
    if (p.work) {
      x3_reduce_kernel<<<(unsigned)std::min<int64_t>(ceil_div(p.M * p.N, 256), 4096), 256, 0, s>>>(p, (int)splits);
      CCX_HIP(hipGetLastError());
    }
    return;
  }
  const bool big = p.M >= 256;
  const bool wide = big && p.N >= 8192;   // 320-column tiles: fewer operand panel fetches per output
  const int64_t bt = big ? 256 : 128, bn = wide ? 320 : bt;
  const int64_t tiles = ceil_div(p.M, bt) * ceil_div(p.N, bn);
  CCX_CHECK(tiles < (1ll << 31), ERR_ARG, "x3 GEMM grid too large");
  // split K into workspace partials when the tile grid cannot fill 256 CUs (keep >= 32 k-tiles each)
  int64_t splits = 1;
  p.work = nullptr;
  // An FP64 destination is written by the kernel directly when the tile grid fills the chip, else
  // through split-K partials summed in FP64 (which need a workspace).
  CCX_CHECK(!p.C64 || ws || tiles >= 512, ERR_INTERNAL, "an FP64 destination on a small grid needs a workspace");
  CCX_CHECK(!(p.C64in || p.D64) || (p.C64 && tiles >= 512 && (!p.D64 || p.dvec)), ERR_INTERNAL,
            "FP64 epilogue inputs need the direct FP64 destination");
  if (ws && tiles < 512) {
    splits = std::min<int64_t>({1024, ceil_div(1024, tiles), p.Kp / (32 * BK)});
    splits = std::max<int64_t>(splits, 1);
    if (big) {
      // one 512-thread workgroup per CU: among up to twice as many splits (shorter FP32 partial
      // sums), take the count whose last round of 256 workgroups is fullest
      const int64_t smax = std::min<int64_t>(2 * splits, std::max<int64_t>(splits, p.Kp / (32 * BK)));
      int64_t best = splits;
      double best_eff = 0.0;
      for (int64_t sc = splits; sc <= smax; ++sc) {
        const int64_t wg = tiles * sc;
        const double eff = (double)wg / (256.0 * (double)ceil_div(wg, 256));
        if (eff > best_eff + 1e-9) {
          best_eff = eff;
          best = sc;
        }
      }
      splits = best;
    }
    if (splits > 1 || p.C64) {
      p.kchunk = ceil_div(ceil_div(p.Kp, splits), BK) * BK;
      splits = ceil_div(p.Kp, p.kchunk);
      p.work = (float*)ws->get((size_t)splits * p.M * p.N * sizeof(float));
    }
  }
  dim3 grid((unsigned)tiles, (unsigned)splits);
  if (wide) x3_gemm_dma_kernel<E, 320><<<grid, 512, 0, s>>>(p);
  else if (big) x3_gemm_dma_kernel<E, 256><<<grid, 512, 0, s>>>(p);
  else x3_gemm_kernel<E, 128><<<grid, 256, 0, s>>>(p);
  CCX_HIP(hipGetLastError());
  if (p.work) {
    x3_reduce_kernel<<<(unsigned)std::min<int64_t>(ceil_div(p.M * p.N, 256), 4096), 256, 0, s>>>(p, (int)splits);
    CCX_HIP(hipGetLastError());
  }
}
template void x3_gemm<__bf16>(const X3Gemm&, hipStream_t, Workspace*);
template void x3_gemm<_Float16>(const X3Gemm&, hipStream_t, Workspace*);

// ---- FP64 contractions as FP16x3 split products (contract(..., split16 = true)) -------------------
namespace {

__device__ __forceinline__ int64_t idx2_at(const Idx2& d, int64_t i) {
  if (i < d.n2) return i * d.s2;
  if ((i | d.n2) < (int64_t(1) << 31)) {   // 32-bit division when it fits
    const int q = (int)i / (int)d.n2;
    return q * d.s1 + ((int)i - q * (int)d.n2) * d.s2;
  }
  return (i / d.n2) * d.s1 + (i % d.n2) * d.s2;
}

// The R x K view X(r, k) = src[idx2_at(dr, r) + idx2_at(dk, k)] in 64 x 64 tiles; a tile's loads run
// along r when r is the contiguous index of src (rfast), else along k.
// Pass 1 (part != null): per-work-group max |X| into part[blockIdx.x + gridDim.x blockIdx.y].
// Pass 2: hl[r][2 Kp] = interleaved FP16 hi/lo of X(r, k) sc (zero for K <= k < Kp), through LDS.
__global__ __launch_bounds__(256) void view_split_kernel(const double* __restrict__ src, int64_t R, int64_t K,
                                                         int64_t Kp, Idx2 dr, Idx2 dk, bool rfast,
                                                         const float* __restrict__ sc, float* __restrict__ part,
                                                         f16* __restrict__ hl) {
  __shared__ float tile[64][65];
  __shared__ int64_t roff[64], koff[64];   // the tile's row and column offsets in src (-1: outside)
  const double scale = part ? 1.0 : (double)sc[0];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
  const int64_t ntr = (R + 63) / 64, ntk = (Kp + 63) / 64;
  double mx = 0.0;
  for (int64_t tr = blockIdx.y; tr < ntr; tr += gridDim.y)
    for (int64_t tk = blockIdx.x; tk < ntk; tk += gridDim.x) {
      const int64_t r0 = tr * 64, k0 = tk * 64;
      __syncthreads();   // the previous tile's offsets and LDS tile are no longer read
      if (ty == 0) roff[tx] = r0 + tx < R ? idx2_at(dr, r0 + tx) : -1;
      else if (ty == 1) koff[tx] = k0 + tx < K ? idx2_at(dk, k0 + tx) : -1;
      __syncthreads();
      for (int j = ty; j < 64; j += 4) {
        const int rl = rfast ? tx : j, kl = rfast ? j : tx;
        const int64_t ro = roff[rl], ko = koff[kl];
        const double x = (ro >= 0 && ko >= 0) ? src[ro + ko] : 0.0;
        if (part) mx = fmax(mx, fabs(x));
        else tile[rl][kl] = (float)(x * scale);
      }
      if (part) continue;
      __syncthreads();
      for (int j = ty; j < 64; j += 4) {
        const int64_t r = r0 + j, k = k0 + tx;
        if (r < R && k < Kp) {
          const float x = tile[j][tx];
          const f16 h = (f16)x;
          const int64_t at = r * 2 * Kp + ((k >> 5) << 6) + (k & 31);
          hl[at] = h;
          hl[at + 32] = (f16)(x - (float)h);
        }
      }
    }
  if (part) {
    const float m = block_max((float)mx);
    if (threadIdx.x == 0) part[blockIdx.x + (int64_t)gridDim.x * blockIdx.y] = m;
  }
}

// C[idx2_at(cm, m) + idx2_at(cn, n)] = beta C + S[m][n] (S from the split GEMM, alpha applied)
__global__ __launch_bounds__(256) void scatter_add_kernel(const float* __restrict__ S, int64_t M, int64_t N, Idx2 cm,
                                                          Idx2 cn, double beta, double* __restrict__ C) {
  for (int64_t m = blockIdx.y; m < M; m += gridDim.y) {
    const int64_t cb = idx2_at(cm, m);
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x) {
      double* c = C + cb + idx2_at(cn, n);
      const double val = (double)S[m * N + n];
      *c = beta == 0.0 ? val : beta * *c + val;
    }
  }
}

}  // namespace

bool x3_contract(const GemmParams<double>& p, Workspace& ws, hipStream_t s) {
  const int64_t M = p.M, N = p.N, K = p.K, Kp = (K + 31) / 32 * 32;
  // A full round of 256 x 256 tiles at least (smaller grids would need split-K partials), and
  // M N / (M + N) >= 128: splitting costs about 16 bytes of traffic per element of A and B, which the
  // faster product repays only when each element takes part in enough products (a thin GEMM, e.g.
  // o x v^3 times v, is faster left in FP64)
  if (p.batch != 1 || K < 32 || ceil_div(M, 256) * ceil_div(N, 256) < 256 || (double)M * N < 128.0 * (M + N))
    return false;
  // A(m, k) = A[am(m) + ak(k)], B(k, n) = B[bk(k) + bn(n)]: B^T rows are (n, k). C is written by
  // the GEMM in FP64 when it is a matrix C[m ldc + n] and x3_gemm's grid has 512 tiles (its direct
  // FP64 epilogue), else through an FP32 copy scattered into C.
  const bool big = M >= 256, wide = big && N >= 8192;
  const int64_t tiles = M <= 64 ? ceil_div(N, 128) : ceil_div(M, big ? 256 : 128) * ceil_div(N, wide ? 320 : (big ? 256 : 128));
  const bool direct = p.cn.s2 == 1 && p.cn.n2 >= N && p.cm.n2 >= M && p.cm.s2 >= N && tiles >= 512 &&
                      (double)(float)p.beta == p.beta;   // the epilogue applies beta in FP32
  auto r256 = [](int64_t b) { return (b + 255) / 256 * 256; };
  constexpr int PB = 256;   // partial maxima per operand (a 16 x 16 grid of the absmax pass)
  const int64_t bA = r256(M * 2 * Kp * 2), bB = r256(N * 2 * Kp * 2), bC = direct ? 0 : r256(M * N * 4);
  char* base = (char*)ws.get((size_t)(bA + bB + bC + 2 * PB * 4 + 256));
  f16* Ahl = (f16*)base;
  f16* Bhl = (f16*)(base + bA);
  float* S = (float*)(base + bA + bB);
  float* part = (float*)(base + bA + bB + bC);
  float* sc = part + 2 * PB;   // sc[0] scale of A, sc[1] scale of B
  const dim3 gmax(16, 16);
  auto split = [&](const double* src, int64_t R, Idx2 dr, Idx2 dk, float* pmax, float* scl, f16* hl) {
    const bool rfast = dr.s2 == 1 && dk.s2 != 1;
    view_split_kernel<<<gmax, 256, 0, s>>>(src, R, K, Kp, dr, dk, rfast, nullptr, pmax, nullptr);
    CCX_HIP(hipGetLastError());
    x3_scale_kernel<<<1, 256, 0, s>>>(pmax, PB, 1.f, pmax, 0, 0.f, scl);
    CCX_HIP(hipGetLastError());
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(Kp, 64), 1024), (unsigned)std::min<int64_t>(ceil_div(R, 64), 65535));
    view_split_kernel<<<g, 256, 0, s>>>(src, R, K, Kp, dr, dk, rfast, scl, nullptr, hl);
    CCX_HIP(hipGetLastError());
  };
  split(p.A, M, p.am, p.ak, part, sc, Ahl);
  split(p.B, N, p.bn, p.bk, part + PB, sc + 1, Bhl);
  X3Gemm g;
  g.M = M; g.N = N; g.Kp = Kp; g.lda = 2 * Kp; g.ldb = 2 * Kp;
  g.kb = 32; g.ska = 64; g.skb = 64;
  g.Ah = Ahl; g.Al = Ahl + 32; g.Bh = Bhl; g.Bl = Bhl + 32;
  g.sa = sc; g.sb = sc + 1; g.alpha = (float)p.alpha;
  if (direct) {
    g.C64 = p.C; g.ldc = p.cm.s2; g.beta = (float)p.beta;
    x3_gemm<_Float16>(g, s);
  } else {
    g.C = S; g.ldc = N; g.beta = 0.f;
    x3_gemm<_Float16>(g, s);
    const dim3 gs((unsigned)std::min<int64_t>(ceil_div(N, 256), 64), (unsigned)std::min<int64_t>(M, 65535));
    scatter_add_kernel<<<gs, 256, 0, s>>>(S, M, N, p.cm, p.cn, p.beta, p.C);
    CCX_HIP(hipGetLastError());
  }
  return true;
}

void x3_gemm_kcat(const X3Gemm& p, hipStream_t s) {
  if (p.M == 0 || p.N == 0 || p.nbatch == 0) return;
  CCX_CHECK(p.kb == BK && p.ska == 2 * BK && p.skb == 2 * BK && p.Kp % BK == 0 && p.K1 % BK == 0 && p.K1 > 0 &&
                p.K1 < p.Kp && p.lda % 8 == 0 && p.ldb % 8 == 0 && p.Ah2 && p.C && !p.C64 && !p.work &&
                p.sw > 0 && p.sw * p.sw == p.M && p.M * p.lda < (int64_t(1) << 31) && p.beta == 0.f && !p.sbn &&
                (p.ytile == 0 || (p.ytile % 8 == 0 && p.ytile == p.sw && p.ytile == p.N &&
                                   (uintptr_t)p.C % 16 == 0 && p.sC % 4 == 0)),
            ERR_INTERNAL, "x3 K-concatenated GEMM layout");
  // the output tile (384 x 192, 256 x 256 or 256 x 192) that pads M x N least, in that order on a
  // tie (the 320-column tile spills registers in this form)
  const int shapes[3][2] = {{384, 192}, {256, 256}, {256, 192}};
  int pick = 0;
  double best = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double area = (double)(ceil_div(p.M, shapes[c][0]) * shapes[c][0]) *
                        (double)(ceil_div(p.N, shapes[c][1]) * shapes[c][1]);
    if (c == 0 || area < best) {
      best = area;
      pick = c;
    }
  }
  const int64_t tiles = ceil_div(p.M, shapes[pick][0]) * ceil_div(p.N, shapes[pick][1]) * p.nbatch;
  CCX_CHECK(tiles < (1ll << 31), ERR_ARG, "x3 GEMM grid too large");
  const bool yt = p.ytile > 0;
  if (pick == 0 && yt) x3_gemm_dma_kernel<_Float16, 192, true, 384, true, 4><<<(unsigned)tiles, 512, 0, s>>>(p);
  else if (pick == 0) x3_gemm_dma_kernel<_Float16, 192, true, 384, false, 4><<<(unsigned)tiles, 512, 0, s>>>(p);
  else if (pick == 1 && yt) x3_gemm_dma_kernel<_Float16, 256, true, 256, true><<<(unsigned)tiles, 512, 0, s>>>(p);
  else if (pick == 1) x3_gemm_dma_kernel<_Float16, 256, true><<<(unsigned)tiles, 512, 0, s>>>(p);
  else if (yt) x3_gemm_dma_kernel<_Float16, 192, true, 256, true><<<(unsigned)tiles, 512, 0, s>>>(p);
  else x3_gemm_dma_kernel<_Float16, 192, true><<<(unsigned)tiles, 512, 0, s>>>(p);
  CCX_HIP(hipGetLastError());
}

template <typename T>
void x3_absmax(const T* x, int64_t n, float* part, hipStream_t s) {
  absmax_kernel<T><<<X3_RED_BLOCKS, 256, 0, s>>>(x, 1, n, n, part);
  CCX_HIP(hipGetLastError());
}
template void x3_absmax<double>(const double*, int64_t, float*, hipStream_t);
template void x3_absmax<float>(const float*, int64_t, float*, hipStream_t);

void x3_scale(const float* pa, float ca, const float* pb, float cb, float* s_out, hipStream_t s) {
  x3_scale_kernel<<<1, 256, 0, s>>>(pa, X3_RED_BLOCKS, ca, pb ? pb : pa, pb ? X3_RED_BLOCKS : 0, cb, s_out);
  CCX_HIP(hipGetLastError());
}
void x3_scale_union(const float* pa, int n, float* s_out, hipStream_t s) {
  x3_scale_kernel<<<1, 256, 0, s>>>(pa, n, 1.f, pa, 0, 0.f, s_out);
  CCX_HIP(hipGetLastError());
}
int x3_partials() { return X3_RED_BLOCKS; }

template <typename Tin>
void x3_split_rows(const Tin* src, int64_t rows, int64_t cols, int64_t lds, const float* sc, f16* hi, f16* lo,
                   int64_t ldd, hipStream_t s, bool il) {
  CCX_CHECK(!il || (ldd % 32 == 0 && lo == hi + 32), ERR_INTERNAL, "interleaved split layout");
  split_rows_kernel<Tin, _Float16><<<grid_x3(rows * cols), 256, 0, s>>>(src, rows, cols, lds, sc, (_Float16*)hi,
                                                                        (_Float16*)lo, ldd, il);
  CCX_HIP(hipGetLastError());
}
template void x3_split_rows<double>(const double*, int64_t, int64_t, int64_t, const float*, f16*, f16*, int64_t,
                                    hipStream_t, bool);
template void x3_split_rows<float>(const float*, int64_t, int64_t, int64_t, const float*, f16*, f16*, int64_t,
                                   hipStream_t, bool);

template <typename Tin>
void x3_split_t(const Tin* src, int64_t R, int64_t Cb, int nx, int64_t lds, const float* sc, f16* hi, f16* lo,
                int64_t Kb, int64_t ldd, hipStream_t s, bool il) {
  CCX_CHECK(!il || (ldd % 32 == 0 && Kb % 32 == 0 && lo == hi + 32), ERR_INTERNAL, "interleaved split layout");
  const int64_t work = ceil_div(R, 32) * ceil_div(Cb, 32) * nx;
  split_t_kernel<Tin, _Float16><<<(unsigned)std::min<int64_t>(work, 1 << 20), 256, 0, s>>>(
      src, R, Cb, nx, lds, sc, (_Float16*)hi, (_Float16*)lo, Kb, ldd, il);
  CCX_HIP(hipGetLastError());
}
template void x3_split_t<float>(const float*, int64_t, int64_t, int, int64_t, const float*, f16*, f16*, int64_t,
                                int64_t, hipStream_t, bool);
template void x3_split_t<double>(const double*, int64_t, int64_t, int, int64_t, const float*, f16*, f16*, int64_t,
                                 int64_t, hipStream_t, bool);

void ladder_pack_x3(int64_t v, int64_t e0, int64_t e1, VVVVSrc src, f16* LPh, f16* LPl, f16* LMh, f16* LMl,
                    float* sP, float* sM, hipStream_t s) {
  if (e1 <= e0) return;
  CCX_CHECK(LPl == LPh + 32 && LMl == LMh + 32, ERR_INTERNAL, "ladder operands must be interleaved");
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  dim3 grid(1, (unsigned)((e1 - e0) * v));
  ladder_pack_x3_kernel<_Float16><<<grid, 256, 0, s>>>(v, e0, src, (_Float16*)LPh, (_Float16*)LPl, (_Float16*)LMh,
                                                       (_Float16*)LMl, sP, sM, bf16_kpad(npP), bf16_kpad(npM));
  CCX_HIP(hipGetLastError());
}

void ladder_pack_o_x3(int64_t o, int64_t v, const double* Covvv, f16* LPh, f16* LPl, f16* LMh, f16* LMl,
                      float* sP, float* sM, hipStream_t s) {
  CCX_CHECK(LPl == LPh + 32 && LMl == LMh + 32, ERR_INTERNAL, "ladder operands must be interleaved");
  dim3 grid(1, (unsigned)(o * v));
  ladder_pack_o_x3_kernel<_Float16><<<grid, 256, 0, s>>>(o, v, Covvv, (_Float16*)LPh, (_Float16*)LPl,
                                                         (_Float16*)LMh, (_Float16*)LMl, sP, sM);
  CCX_HIP(hipGetLastError());
}

void tau_pack_x3(int64_t o, int64_t v, int64_t KpP, int64_t KpM, const double* t1, const double* t2, const float* sc,
                 f16* Tph, f16* Tpl, f16* Tmh, f16* Tml, hipStream_t s) {
  CCX_CHECK(KpP % 32 == 0 && KpM % 32 == 0 && Tpl == Tph + 32 && Tml == Tmh + 32, ERR_INTERNAL,
            "packed tau must be interleaved");
  const int64_t npP = v * (v + 1) / 2;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(npP, 256), 1024), (unsigned)(o * (o + 1) / 2));
  tau_pack_x3_kernel<_Float16><<<grid, 256, 0, s>>>(o, v, KpP, KpM, t1, t2, sc, (_Float16*)Tph, (_Float16*)Tpl,
                                                    (_Float16*)Tmh, (_Float16*)Tml);
  CCX_HIP(hipGetLastError());
}

void ring_prep_x3(int64_t o, int64_t v, int64_t ovp, const double* t1, const double* t2, const float* sc, f16* BJh,
                  f16* BJl, f16* BKh, f16* BKl, f16* APh, f16* APl, f16* AQh, f16* AQl, double* Ud,
                  hipStream_t s) {
  CCX_CHECK(ovp % 32 == 0 && BJl == BJh + 32 && BKl == BKh + 32 && APl == APh + 32 && AQl == AQh + 32,
            ERR_INTERNAL, "ring operands must be interleaved");
  const int64_t tiles = ceil_div(v, 32);
  dim3 grid((unsigned)(tiles * tiles), (unsigned)(o * o));
  ring_prep_x3_kernel<_Float16><<<grid, 256, 0, s>>>(o, v, ovp, t1, t2, sc, (_Float16*)BJh, (_Float16*)BJl,
                                                     (_Float16*)BKh, (_Float16*)BKl, (_Float16*)APh,
                                                     (_Float16*)APl, (_Float16*)AQh, (_Float16*)AQl, Ud);
  CCX_HIP(hipGetLastError());
}

// ---- the <ov|vv> terms of MIXED_F16X3 ----------------------------------------------------------
// Six contractions stream Covvv = (me|af) [m][e][a][f] (ccwfn.py:496 twice, :642, :681, :939, :759).
// With t1 as the small operand they are three split GEMMs over two split copies of Covvv:
//   PT[j][(m,x,y)]  = sum_z t_jz Covvv[m][x][y][z]        (copy CvA = Covvv rows, K = z)
//   P2T[j][(m,b,e)] = sum_f t_jf (mf|be)                   (copy CvT, rows (m,b,e), K = f)
//   r1[i][a]       += sum_(m,e,f) Ud[m][e][i][f] (me|af)   (CvA with blocked K = (m,e), f)
// and their uses are permuted adds:
//   Fae[a][e] += 2 sum_m P2T[m][(m,a,e)] - sum_m PT[m][(m,e,a)]               (:496)
//   Wmbej[m][e][j][b] += PT[j][(m,e,b)],  Wmbje[m][e][j][b] -= P2T[j][(m,b,e)]  (:642, :681)
//   Hh[i][j][a][b] += PT[i][(j,b,a)]                      (:939, added inside the r2 assembly)
namespace {

// dst[(i, q)][f] (row stride ldd, q < nq) = split(src[q][i][f] * sc): Ud [(m,e)][(i,f)] -> [i][(m,e)][f]
__global__ __launch_bounds__(256) void split_udt_kernel(const double* __restrict__ Ud, int64_t o, int64_t v, int64_t nq,
                                                        const float* __restrict__ sc, _Float16* hi, _Float16* lo,
                                                        int64_t v32) {
  const double s = sc[0];
  const int64_t total = o * nq * v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = e % v, t = e / v, q = t % nq, i = t / nq;   // destination order: f fastest
    split_store<_Float16>((float)(Ud[(q * o + i) * v + f] * s), hi, lo, (i * nq + q) * v32 + f);
  }
}

// The FP32 ring W in one pass (ccwfn.py:641-643 Wmbej, :680-682 Wmbje), W[(m,e)][x][(j,b)]:
//   x = 0: W0 + PT[j][(m,e,b)]  - sum_n t_nb (nj|me)
//   x = 1: W0 - P2T[j][(m,b,e)] + sum_n t_nb (mj|ne)
// summed in FP64 and rounded once. For fixed (m, j) the two n sums are (v x o)(o x v) products,
// A0[e][n] = (nj|me), A1[e][n] = (mj|ne), T[n][b] = t_nb: one workgroup per (m, j, 32 e's, 128 b's),
// its 4 waves in 2 (e) x 2 (b), each wave 16 e x 64 b as four 16 x 16 blocks per sum on FP64 MFMA
// (v_mfma_f64_16x16x4f64, K = o), operands read straight from global memory in the lane layout
// (A: lane l holds A[l % 16][l / 16], B: lane l holds B[l / 16][l % 16], C: lane l register r holds
// C[4 r + l / 16][l % 16]). The epilogue adds W0, PT and P2T and writes W.
typedef double wb_d4 __attribute__((ext_vector_type(4)));
constexpr int WBE = 32, WBB = 128;
__global__ __launch_bounds__(256) void w_build_x3_kernel(int64_t o, int64_t v, const double* __restrict__ t1,
                                                         const double* __restrict__ Cooov,
                                                         const float* __restrict__ W0, const float* __restrict__ PT,
                                                         const float* __restrict__ P2T, float* __restrict__ W) {
  const int64_t ov = o * v, neb = (v + WBE - 1) / WBE, nbb = (v + WBB - 1) / WBB;
  int64_t q = blockIdx.x;
  const int64_t bb = q % nbb;
  q /= nbb;
  const int64_t eb = q % neb, mj = q / neb;
  const int64_t m = mj / o, j = mj - m * o;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t e0 = eb * WBE + 16 * (w >> 1), b0 = bb * WBB + 64 * (w & 1);
  const int64_t ea = e0 + (lane & 15);                       // A row of this lane
  const bool eok = ea < v;
  wb_d4 acc0[4], acc1[4];
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) acc0[jb] = acc1[jb] = wb_d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = 0; k0 < o; k0 += 4) {
    const int64_t n = k0 + (lane >> 4);
    const bool nok = n < o;
    const double a0 = (nok && eok) ? Cooov[((n * o + j) * o + m) * v + ea] : 0.0;   // (nj|me)
    const double a1 = (nok && eok) ? Cooov[((m * o + j) * o + n) * v + ea] : 0.0;   // (mj|ne)
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
      const int64_t bn = b0 + 16 * jb + (lane & 15);
      const double tb = (nok && bn < v) ? t1[n * v + bn] : 0.0;
      acc0[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, tb, acc0[jb], 0, 0, 0);
      acc1[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, tb, acc1[jb], 0, 0, 0);
    }
  }
  const float* __restrict__ p2src = P2T + (j * o + m) * v * v;
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) {
    const int64_t b = b0 + 16 * jb + (lane & 15);
    if (b >= v) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = e0 + 4 * r + (lane >> 4);
      if (e >= v) continue;
      const int64_t me = m * v + e;
      const int64_t at0 = me * 2 * ov + j * v + b, at1 = at0 + ov;
      W[at0] = (float)((double)W0[at0] + (double)PT[(j * ov + me) * v + b] - acc0[jb][r]);
      W[at1] = (float)((double)W0[at1] - (double)p2src[b * v + e] + acc1[jb][r]);
    }
  }
}

// Fae[a][e] += 2 sum_m P2T[m][(m,a,e)] - sum_m PT[m][(m,e,a)]
__global__ __launch_bounds__(256) void fae_x3_kernel(const float* __restrict__ PT, const float* __restrict__ P2T,
                                                     int64_t o, int64_t v, double* __restrict__ Fae) {
  const int64_t vv = v * v, ov2 = o * vv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < vv; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = e / v, ee = e - a * v;
    double acc = 0.0;
    for (int64_t m = 0; m < o; ++m) {
      const int64_t base = m * ov2 + m * vv;
      acc += 2.0 * (double)P2T[base + a * v + ee] - (double)PT[base + ee * v + a];
    }
    Fae[e] += acc;
  }
}

inline unsigned grid_e(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

}  // namespace

void covvv_x3_setup(Ctx& c, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov(), v32 = c.v32();
  auto halloc = [&](f16** q, int64_t n) {
    *q = c.alloc<f16>(n);
    CCX_HIP(hipMemsetAsync(*q, 0, n * sizeof(f16), s));   // K padding stays zero
  };
  halloc(&c.CvAh, ov * v * v32);
  halloc(&c.CvAl, ov * v * v32);
  halloc(&c.CvTh, ov * v * v32);
  halloc(&c.CvTl, ov * v * v32);
  halloc(&c.T1h, o * v32);
  halloc(&c.T1l, o * v32);
  halloc(&c.UTh, o * ov * v32);
  halloc(&c.UTl, o * ov * v32);
  c.PT = c.alloc<float>(o * ov * v);
  c.P2T = c.alloc<float>(o * ov * v);
  halloc(&c.TFh, o * ov * v32);
  halloc(&c.TFl, o * ov * v32);
  halloc(&c.LFh, o * ov * v32);
  halloc(&c.LFl, o * ov * v32);
  // W0[(m,e)][x][(j,b)] (FP32): x = 0 (me|jb) (:641), x = 1 -(mj|be) (:680), through an FP64 temporary
  c.W0 = c.alloc<float>(2 * ov * ov);
  {
    double* W0d = nullptr;
    CCX_HIP(hipMalloc(&W0d, 2 * ov * ov * sizeof(double)));
    const std::vector<int64_t> OVOV{o, v, o, v}, WSTR{v * 2 * ov, 2 * ov, v, 1};
    axpby<double>(1.0, tv(c.Covov, "mejb", OVOV), 0.0, tvs(W0d, "mejb", OVOV, WSTR), s);
    axpby<double>(-1.0, tv(c.Coovv, "mjbe", {o, o, v, v}), 0.0, tvs(W0d + ov, "mejb", OVOV, WSTR), s);
    cast<double, float>(2 * ov * ov, W0d, c.W0, s);
    CCX_HIP(hipStreamSynchronize(s));
    CCX_HIP(hipFree(W0d));
  }
  x3_absmax<double>(c.Loovv, ov * ov, c.xpart, s);
  x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 12, s);
  x3_split_rows<double>(c.Loovv, o * ov, v, v, c.xs + 12, c.LFh, c.LFl, v32, s);
  x3_absmax<double>(c.Covvv, ov * v * v, c.xpart, s);
  x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 8, s);
  x3_split_rows<double>(c.Covvv, ov * v, v, v, c.xs + 8, c.CvAh, c.CvAl, v32, s);
  for (int64_t m = 0; m < o; ++m)   // CvT[m][(b,e)][f] = Covvv[m][f][b][e]
    x3_split_t<double>(c.Covvv + m * v * v * v, v, v * v, 1, v * v, c.xs + 8, c.CvTh + m * v * v * v32,
                       c.CvTl + m * v * v * v32, 0, v32, s);
}

void covvv_x3_products(Ctx& c, const double* t1, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov(), v32 = c.v32();
  x3_split_rows<double>(t1, o, v, v, c.xs + 9, c.T1h, c.T1l, v32, s);   // scale from x3_amp_scales
  X3Gemm g;
  g.M = o; g.N = ov * v; g.Kp = v32; g.lda = v32; g.ldb = v32; g.ldc = ov * v;
  g.Ah = c.T1h; g.Al = c.T1l; g.Bh = c.CvAh; g.Bl = c.CvAl; g.C = c.PT;
  g.sa = c.xs + 9; g.sb = c.xs + 8;
  x3_gemm<_Float16>(g, s);
  g.Bh = c.CvTh; g.Bl = c.CvTl; g.C = c.P2T;
  x3_gemm<_Float16>(g, s);
}

void covvv_x3_fae(Ctx& c, double* Fae, hipStream_t s) {
  fae_x3_kernel<<<grid_e(c.v * c.v), 256, 0, s>>>(c.PT, c.P2T, c.o, c.v, Fae);
  CCX_HIP(hipGetLastError());
}

// W = the FP32 ring W (every element written)
void w_build_x3(Ctx& c, const double* t1, float* W, hipStream_t s) {
  const int64_t o = c.o, v = c.v;
  const int64_t blocks = o * o * ceil_div(v, WBE) * ceil_div(v, WBB);
  CCX_CHECK(blocks < (1ll << 31), ERR_ARG, "W build grid too large");
  w_build_x3_kernel<<<(unsigned)blocks, 256, 0, s>>>(o, v, t1, c.Cooov, c.W0, c.PT, c.P2T, W);
  CCX_HIP(hipGetLastError());
}

void covvv_x3_r1(Ctx& c, const double* Ud, double* r1, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov(), v32 = c.v32();
  split_udt_kernel<<<grid_e(o * ov * v), 256, 0, s>>>(Ud, o, v, ov, c.xs + 10, (_Float16*)c.UTh, (_Float16*)c.UTl,
                                                      v32);
  CCX_HIP(hipGetLastError());
  X3Gemm g;   // r1[i][a] += sum_(me, f) UT[i][(me)][f] CvA[(me)][a][f], partials summed in FP64
  g.M = o; g.N = v; g.Kp = ov * v32; g.lda = ov * v32; g.ldb = v32; g.ldc = v;
  g.kb = v32; g.ska = v32; g.skb = v * v32;
  g.Ah = c.UTh; g.Al = c.UTl; g.Bh = c.CvAh; g.Bl = c.CvAl; g.C64 = r1;
  g.sa = c.xs + 10; g.sb = c.xs + 8; g.beta = 1.f;
  x3_gemm<_Float16>(g, s, &c.ws);
}

void x3_amp_scales(Ctx& c, const double* t1, const double* t2, hipStream_t s) {
  const int np = x3_partials();
  const int64_t ov = c.ov();
  x3_absmax<double>(t2, ov * ov, c.xpart, s);
  x3_absmax<double>(t1, ov, c.xpart + np, s);
  const float* p2 = c.xpart;
  const float* p1 = c.xpart + np;
  x3_scale(p2, 3.f, p1, 1.f, c.xs + 0, s);     // ring operands: |2T - X|, |T - X/2 - P|, ...
  x3_scale(p2, 2.f, p1, 2.f, c.xs + 4, s);     // packed tau+-
  x3_scale(p2, 1.f, p1, 1.f, c.xs + 5, s);     // tau = t2 + t1 t1
  x3_scale(p1, 1.f, nullptr, 0.f, c.xs + 9, s);   // t1
  x3_scale(p2, 3.f, nullptr, 0.f, c.xs + 10, s);  // Ud = 2 t2 - t2^T
  x3_scale(p2, 1.f, p1, 0.5f, c.xs + 11, s);   // tau~ = t2 + 1/2 t1 t1
  x3_scale(p2, 1.f, nullptr, 0.f, c.xs + 13, s);  // t2
}

namespace {
// Gp = G with a zero diagonal, Gd = diag(G) (G is v x v)
__global__ __launch_bounds__(256) void g_offdiag_kernel(const double* __restrict__ G, int64_t v,
                                                        double* __restrict__ Gp, double* __restrict__ Gd) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < v * v; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / v, f = e - b * v;
    Gp[e] = b == f ? 0.0 : G[e];
    if (b == f) Gd[b] = G[e];
  }
}
}  // namespace

bool hh_g_x3_fits(int64_t o, int64_t v) {
  return o * o * v >= 256 && ceil_div(o * o * v, 256) * ceil_div(v, 256) >= 512;
}

void hh_g_x3(Ctx& c, const double* t2, float* X, hipStream_t s) {
  const int64_t o = c.o, v = c.v, v32 = c.v32();
  CCX_CHECK(hh_g_x3_fits(o, v) && c.T2h && c.GPh, ERR_INTERNAL, "hh_g_x3 on a grid it does not fit");
  g_offdiag_kernel<<<grid_e(v * v), 256, 0, s>>>(c.G, v, c.Gp, c.Gd);
  CCX_HIP(hipGetLastError());
  x3_absmax<double>(c.Gp, v * v, c.xpart, s);
  x3_scale(c.xpart, 1.f, nullptr, 0.f, c.xs + 14, s);
  x3_split_rows<double>(c.Gp, v, v, v, c.xs + 14, c.GPh, c.GPl, v32, s, true);
  x3_split_rows<double>(t2, o * o * v, v, v, c.xs + 13, c.T2h, c.T2l, v32, s, true);
  X3Gemm g;   // X[(ija)][b] = sum_(e != b) t2[(ija)][e] G[b][e] in FP32
  g.M = o * o * v; g.N = v; g.Kp = v32; g.lda = 2 * v32; g.ldb = 2 * v32; g.ldc = v;
  g.kb = 32; g.ska = 64; g.skb = 64;
  g.Ah = c.T2h; g.Al = c.T2l; g.Bh = c.GPh; g.Bl = c.GPl; g.sa = c.xs + 13; g.sb = c.xs + 14;
  g.C = X; g.beta = 0.f;
  x3_gemm<_Float16>(g, s);
}

void fae_tauL_x3(Ctx& c, double* Fae, hipStream_t s) {
  const int64_t o = c.o, v = c.v, ov = c.ov(), v32 = c.v32();
  x3_split_rows<double>(c.tau, o * ov, v, v, c.xs + 11, c.TFh, c.TFl, v32, s);   // tau~ [(m,n,a)][f]
  // the diagonal sums are of one sign (pair-energy-like): split K finely, sum the FP32 partials in
  // FP64 straight into Fae
  X3Gemm g;   // Fae[a][e] -= sum_(mn, f) tau~[(mn)][a][f] L[(mn)][e][f]
  g.M = v; g.N = v; g.Kp = o * o * v32; g.lda = v32; g.ldb = v32; g.ldc = v;
  g.kb = v32; g.ska = v * v32; g.skb = v * v32;
  g.Ah = c.TFh; g.Al = c.TFl; g.Bh = c.LFh; g.Bl = c.LFl; g.C64 = Fae;
  g.sa = c.xs + 11; g.sb = c.xs + 12; g.alpha = -1.f; g.beta = 1.f;
  x3_gemm<_Float16>(g, s, &c.ws);
}

// C = alpha op(A) op(B) + beta C for FP32 row-major matrices through FP16x3: op(A) (M x K) and
// op(B)^T (N x K) are split into K-contiguous workspace copies with their own scales.
void gemm_f16x3(bool ta, bool tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                const float* B, int64_t ldb, float beta, float* C, int64_t ldc, Workspace& ws, Workspace& kws,
                hipStream_t s) {
  CCX_CHECK(M >= 0 && N >= 0 && K >= 0, ERR_ARG, "negative GEMM extent");
  if (M == 0 || N == 0) return;
  const int64_t Kp = bf16_kpad(K);
  const int np = x3_partials();
  const size_t half = (size_t)(M + N) * Kp * 2 * sizeof(f16);
  char* w = (char*)ws.get(half + (2 * (size_t)np + 64) * sizeof(float));
  f16* Ah = (f16*)w;
  f16* Al = Ah + M * Kp;
  f16* Bh = Al + M * Kp;
  f16* Bl = Bh + N * Kp;
  float* part = (float*)(w + half);
  float* sc = part + 2 * np;
  CCX_HIP(hipMemsetAsync(w, 0, half, s));   // K padding
  // op(A)[m][k] = ta ? A[k][m] : A[m][k]; op(B)^T[n][k] = tb ? B[n][k] : B[k][n]
  absmax_kernel<float><<<np, 256, 0, s>>>(A, ta ? K : M, ta ? M : K, lda, part);
  absmax_kernel<float><<<np, 256, 0, s>>>(B, tb ? N : K, tb ? K : N, ldb, part + np);
  CCX_HIP(hipGetLastError());
  x3_scale(part, 1.f, nullptr, 0.f, sc, s);
  x3_scale(part + np, 1.f, nullptr, 0.f, sc + 1, s);
  if (K > 0) {
    if (ta) x3_split_t(A, K, M, 1, lda, sc, Ah, Al, 0, Kp, s);
    else x3_split_rows<float>(A, M, K, lda, sc, Ah, Al, Kp, s);
    if (tb) x3_split_rows<float>(B, N, K, ldb, sc + 1, Bh, Bl, Kp, s);
    else x3_split_t(B, K, N, 1, ldb, sc + 1, Bh, Bl, 0, Kp, s);
  }
  X3Gemm g;
  g.M = M; g.N = N; g.Kp = Kp; g.lda = Kp; g.ldb = Kp; g.ldc = ldc;
  g.Ah = Ah; g.Al = Al; g.Bh = Bh; g.Bl = Bl; g.C = C;
  g.sa = sc; g.sb = sc + 1; g.alpha = alpha; g.beta = beta;
  x3_gemm<_Float16>(g, s, &kws);
}

// The packed ladder (ccwfn.py:931 and Z, :715) as FP16x3 GEMMs: scaled tau rows times the
// per-row-scaled ladder matrices, FP32 outputs CP / CM.
// The packed ladder GEMMs (tau+- against the split <ab|ef> rows): with Z the (ab) and the ov Z
// columns, else the (ab) columns only; the outputs keep their rows of npP + ov (npM + ov) floats.
static void ladder_x3_gemms(Ctx& c, const float* sa, bool with_z, hipStream_t s) {
  const int64_t ov = c.ov();
  const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
  const int64_t KpP = bf16_kpad(npP), KpM = bf16_kpad(npM);
  X3Gemm g;
  // all operands interleaved: rows of 2 Kp, kb = 32, ska = skb = 64
  g.M = noP; g.N = npP + (with_z ? ov : 0); g.Kp = KpP; g.lda = 2 * KpP; g.ldb = 2 * KpP; g.ldc = npP + ov;
  g.kb = 32; g.ska = 64; g.skb = 64;
  g.Ah = c.Tph; g.Al = c.Tpl; g.Bh = c.LPh; g.Bl = c.LPl; g.C = (float*)c.CP;
  g.sa = sa; g.sbn = c.lsP;
  int k = c.tbegin(s, Ctx::PH_LADDER, 2.0 * noP * g.N * npP);
  x3_gemm<_Float16>(g, s);
  c.tend(s, k);
  if (noM > 0) {
    X3Gemm h = g;
    h.M = noM; h.N = npM + (with_z ? ov : 0); h.Kp = KpM; h.lda = 2 * KpM; h.ldb = 2 * KpM; h.ldc = npM + ov;
    h.Ah = c.Tmh; h.Al = c.Tml; h.Bh = c.LMh; h.Bl = c.LMl; h.C = (float*)c.CM; h.sbn = c.lsM;
    k = c.tbegin(s, Ctx::PH_LADDER, 2.0 * noM * h.N * npM);
    x3_gemm<_Float16>(h, s);
    c.tend(s, k);
  }
}

void ladder_x3(Ctx& c, const double* t1, const double* t2, hipStream_t s) {
  const int64_t o = c.o, v = c.v;
  const int64_t KpP = bf16_kpad(c.npP()), KpM = bf16_kpad(c.npM());
  // packed tau+- are bounded by 2 max|t2| + 2 max|t1|^2 (scale xs[4] from x3_amp_scales)
  tau_pack_x3(o, v, KpP, KpM, t1, t2, c.xs + 4, (f16*)c.Tph, (f16*)c.Tpl, (f16*)c.Tmh, (f16*)c.Tml, s);
  ladder_x3_gemms(c, c.xs + 4, true, s);
}

void ladder_x3_l2(Ctx& c, const double* l2, hipStream_t s) {
  const int64_t o = c.o, v = c.v;
  const int64_t KpP = bf16_kpad(c.npP()), KpM = bf16_kpad(c.npM());
  // l2's packed +- combinations are bounded by 2 max|l2|: scale xs[16]
  x3_absmax<double>(l2, o * o * v * v, c.xpart, s);
  x3_scale(c.xpart, 2.f, nullptr, 0.f, c.xs + 16, s);
  tau_pack_x3(o, v, KpP, KpM, nullptr, l2, c.xs + 16, (f16*)c.Tph, (f16*)c.Tpl, (f16*)c.Tmh, (f16*)c.Tml, s);
  ladder_x3_gemms(c, c.xs + 16, false, s);
}

}  // namespace ccx

// Plain-matrix MFMA GEMM for gfx950: C[M][N] = alpha * op(A) op(B) + beta * C, C row-major.
//
// Used for every contraction whose operands are ordinary strided matrices -- the packed ladder and
// the six ring GEMM units of the CCSD residual, which carry almost all of its flops. Compared with
// the 2-level-index kernel (gemm.hip) it loads 16-byte vectors, computes addresses once per tile
// and keeps the k loop free of index arithmetic.
//
//   A: A_KC = true  -> A[m * lda + k]  (k contiguous);  false -> A[k * lda + m]  (m contiguous)
//   B: B_NC = true  -> B[k * ldb + n]  (n contiguous);  false -> B[n * ldb + k]  (k contiguous)
//
// Tiles: 128 x 128 per workgroup (4 waves, 2 x 2, each 64 x 64 = 4 x 4 MFMA 16x16 blocks), BK = 16
// (f64) or 32 (f32) so that one tile row is 128 bytes. Operands are double-buffered: the next tile
// is loaded into registers while the current one is consumed from LDS.
#include "common.h"
#include "gemm.h"
#include <cstdlib>
#include <type_traits>

namespace ccx {

typedef double d4p __attribute__((ext_vector_type(4)));
typedef float f4p __attribute__((ext_vector_type(4)));

namespace {

// CCX_PLAIN_TILE=128 or 256 forces the FP32 tile size (measurement of the two variants).
int plain_tile_override() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("CCX_PLAIN_TILE");
    v = e ? std::atoi(e) : 0;
  }
  return v;
}

template <typename T> struct PAcc;
template <> struct PAcc<double> { typedef d4p type; };
template <> struct PAcc<float> { typedef f4p type; };

__device__ __forceinline__ d4p pmfma(double a, double b, d4p c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4p pmfma(float a, float b, f4p c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int64_t pxcd_remap(int64_t w, int64_t ntiles) {
  const int64_t nx = 8;
  int64_t q = ntiles / nx, r = ntiles % nx;
  int64_t x = w % nx, j = w / nx;
  return x * q + (x < r ? x : r) + j;
}

// One operand tile: R rows (m or n) x BK k's. KC: stored [R][BK + PAD] (k contiguous in LDS) and
// loaded along k; otherwise stored [BK][R + PAD] and loaded along the row index.
template <typename T, int R, int BK, bool KC>
struct alignas(16) PTile {
  static constexpr int VEC = 16 / sizeof(T);
  // Row padding (keeps every LDS row 16-byte aligned). A k-contiguous tile is read along its rows
  // (16 lanes on 16 rows) and VEC suffices; a row-index-contiguous tile is read by 4 lane groups on
  // 4 consecutive k rows, which must then start 16 banks apart: R + 16 elements per row.
  static constexpr int PAD = KC ? VEC : 16;
  static constexpr int NV = R * BK / VEC / 256;         // 16-byte vectors per thread
  T d[KC ? R : BK][(KC ? BK : R) + PAD];
};

template <typename T, int R, int BK, bool KC>
struct PLoader {
  static constexpr int VEC = 16 / sizeof(T);
  static constexpr int NV = R * BK / VEC / 256;
  static constexpr int SEGS = KC ? BK / VEC : R / VEC;  // vectors per stored row
  uint4 v[NV];
  int row[NV], seg[NV];                                 // stored-row index and vector index

  __device__ __forceinline__ void init(int tid) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      int q = tid + 256 * i;
      row[i] = q / SEGS;
      seg[i] = q % SEGS;
    }
  }
  // base points at the tile origin (row 0, k = k0); ld = leading dimension in elements
  __device__ __forceinline__ void load_fast(const T* __restrict__ base, int64_t ld) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
      v[i] = *(const uint4*)(base + (int64_t)row[i] * ld + seg[i] * VEC);
  }
  // guarded: rows beyond nrows / k beyond nk are zero (nrows, nk relative to the tile origin)
  __device__ __forceinline__ void load_safe(const T* __restrict__ base, int64_t ld, int64_t nrows, int64_t nk) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      union { uint4 u; T e[VEC]; } x;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        int64_t r = KC ? row[i] : (int64_t)seg[i] * VEC + e;    // tile row (m / n)
        int64_t k = KC ? (int64_t)seg[i] * VEC + e : row[i];    // tile k
        const T* p = KC ? base + (int64_t)row[i] * ld + seg[i] * VEC + e
                        : base + (int64_t)row[i] * ld + seg[i] * VEC + e;
        x.e[e] = (r < nrows && k < nk) ? *p : T(0);
      }
      v[i] = x.u;
    }
  }
  // row r of the tile at element offset roff[r] (an LDS table, relative to base) instead of r * ld;
  // guarded like load_safe unless FASTL (rows beyond nrows / k beyond nk are zero)
  template <bool FASTL>
  __device__ __forceinline__ void load_off(const T* __restrict__ base, const int* roff, int64_t nrows, int64_t nk) {
    if constexpr (!KC) return;   // row offsets are for k-contiguous tiles only
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int64_t o = roff[row[i]];
      if constexpr (FASTL) {
        v[i] = *(const uint4*)(base + o + seg[i] * VEC);
      } else {
        union { uint4 u; T e[VEC]; } x;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const int64_t k = (int64_t)seg[i] * VEC + e;
          x.e[e] = (row[i] < nrows && k < nk) ? base[o + k] : T(0);
        }
        v[i] = x.u;
      }
    }
  }
  __device__ __forceinline__ void store(PTile<T, R, BK, KC>& t) {
#pragma unroll
    for (int i = 0; i < NV; ++i) *(uint4*)&t.d[row[i]][seg[i] * VEC] = v[i];
  }
};

// BT = 128: 4 waves of 64 x 64, two workgroups per CU. BT = 256 (FP32 only): 4 waves of 128 x 128
// (256 accumulator registers each), one workgroup per CU, 4x the MFMAs per LDS read and per barrier.
// KCAT: the K-concatenated form of plain_gemm_kcat (A k-contiguous).
template <typename T, int BK, bool A_KC, bool B_NC, int BT, bool KCAT = false>
__global__ __launch_bounds__(256, BT == 128 ? 2 : 1) void plain_gemm_kernel(PlainGemm<T> p) {
  constexpr int BM = BT, BN = BT, WM = BT / 2, WN = BT / 2, NBI = WM / 16, NBJ = WN / 16;
  typedef typename PAcc<T>::type acc_t;
  __shared__ PTile<T, BM, BK, A_KC> As[2];
  __shared__ PTile<T, BN, BK, !B_NC> Bs[2];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int fr = lane >> 4, fc = lane & 15;

  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  int64_t t = pxcd_remap(blockIdx.x, tiles_m * tiles_n);
  const int64_t GM = 8;
  int64_t group = t / (GM * tiles_n);
  int64_t first_m = group * GM;
  int64_t gsize = tiles_m - first_m < GM ? tiles_m - first_m : GM;
  int64_t within = t - group * GM * tiles_n;
  const int64_t tm = first_m + within % gsize, tn = within / gsize;
  const int64_t m0 = tm * BM, n0 = tn * BN;

  // K of this grid row: a split-K chunk may be shorter than p.K
  const int64_t Kb = p.ktot > 0 ? (p.ktot - (int64_t)blockIdx.y * p.K < p.K ? p.ktot - (int64_t)blockIdx.y * p.K : p.K)
                                : p.K;
  const T* A = p.A + blockIdx.y * p.a_b;
  const T* B = p.B + blockIdx.y * p.b_b;
  T* C = p.C + blockIdx.y * p.c_b;

  // tile origins: A at (m0, k), B at (k, n0)
  auto a_at = [&](int64_t k) { return A_KC ? A + m0 * p.lda + k : A + k * p.lda + m0; };
  auto b_at = [&](int64_t k) { return B_NC ? B + k * p.ldb + n0 : B + n0 * p.ldb + k; };
  const int64_t a_rows = p.M - m0, b_rows = p.N - n0;   // valid tile rows (m / n)
  const bool interior = a_rows >= BM && b_rows >= BN;

  PLoader<T, BM, BK, A_KC> la;
  PLoader<T, BN, BK, !B_NC> lb;
  la.init(tid);
  lb.init(tid);
  // KCAT: the second operand's row offsets, one per tile row, in LDS (sw^2 lda < 2^31, checked on
  // the host; rows past M clamp to row M - 1 and are not stored). The barrier before the k loop
  // publishes them.
  // The first operand's rows get the same treatment (offsets r lda from the tile origin), so that
  // one load path serves both.
  const T* A2 = KCAT ? p.A2 + blockIdx.y * p.a2_b : nullptr;
  __shared__ int roff[2][KCAT ? BM : 1];
  if constexpr (KCAT) {
    if (tid < BM) {
      const int64_t m = m0 + tid < p.M ? m0 + tid : p.M - 1;
      roff[0][tid] = (int)(tid * p.lda);
      roff[1][tid] = (int)(((m % p.sw) * p.sw + m / p.sw) * p.lda);
    }
    __syncthreads();
  }

  acc_t acc[NBI][NBJ];
#pragma unroll
  for (int i = 0; i < NBI; ++i)
#pragma unroll
    for (int j = 0; j < NBJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // The k loop, compiled twice: FAST (an interior tile and K a multiple of BK, so every load is a
  // full in-bounds vector) and checked. Keeping the checked loads out of the FAST loop lets the
  // prefetch of tile k + 1 stay in flight while tile k is multiplied (a shared loop forces a full
  // wait on the loads at the join of the two load paths).
  auto kloop = [&](auto fast_tag) {
    constexpr bool FAST = decltype(fast_tag)::value;
    auto load = [&](int64_t k0) {
      const int64_t nk = Kb - k0;
      if constexpr (KCAT) {   // the k-tiles past K1 read the second operand (block-uniform)
        const bool second = k0 >= p.K1;
        la.template load_off<FAST>(second ? A2 + (k0 - p.K1) : a_at(k0), roff[second ? 1 : 0], a_rows, nk);
      } else if constexpr (FAST) {
        la.load_fast(a_at(k0), p.lda);
      } else {
        la.load_safe(a_at(k0), p.lda, a_rows, nk);
      }
      if constexpr (FAST) lb.load_fast(b_at(k0), p.ldb);
      else lb.load_safe(b_at(k0), p.ldb, b_rows, nk);
    };
    load(0);
    la.store(As[0]);
    lb.store(Bs[0]);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = 0; k0 < Kb; k0 += BK) {
      const bool has_next = k0 + BK < Kb;
      if (has_next) load(k0 + BK);
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T a[NBI], b[NBJ];
#pragma unroll
        for (int i = 0; i < NBI; ++i)
          a[i] = A_KC ? As[buf].d[wm * WM + i * 16 + fc][ks + fr] : As[buf].d[ks + fr][wm * WM + i * 16 + fc];
#pragma unroll
        for (int j = 0; j < NBJ; ++j)
          b[j] = B_NC ? Bs[buf].d[ks + fr][wn * WN + j * 16 + fc] : Bs[buf].d[wn * WN + j * 16 + fc][ks + fr];
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
          for (int j = 0; j < NBJ; ++j) acc[i][j] = pmfma(a[i], b[j], acc[i][j]);
      }
      if (has_next) {
        la.store(As[buf ^ 1]);
        lb.store(Bs[buf ^ 1]);
      }
      __syncthreads();
      buf ^= 1;
    }
  };
  if (Kb > 0) {
    if (interior && Kb % BK == 0) kloop(std::true_type{});
    else kloop(std::false_type{});
  }

#pragma unroll
  for (int j = 0; j < NBJ; ++j) {
    const int64_t n = n0 + wn * WN + j * 16 + fc;
    if (n >= p.N) continue;
#pragma unroll
    for (int i = 0; i < NBI; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // output row of register r (f64: 4r + fr, f32: 4 fr + r; see gemm.hip)
        const int64_t m = m0 + wm * WM + i * 16 + (sizeof(T) == 8 ? 4 * r + fr : 4 * fr + r);
        if (m >= p.M) continue;
        T* dst = C + m * p.ldc + n;
        T val = p.alpha * acc[i][j][r];
        if (p.beta != T(0)) val += p.beta * *dst;
        *dst = val;
      }
  }
}

// 256 x 256 tiles for large FP32 GEMMs (enough tiles to fill the 256 CUs several times over).
template <typename T>
bool plain_big_tiles(const PlainGemm<T>& p) {
  if (sizeof(T) != 4 || plain_tile_override() == 128) return false;
  return plain_tile_override() == 256 || ceil_div(p.M, 256) * ceil_div(p.N, 256) * p.batch >= 1024;
}

template <typename T, bool A_KC, bool B_NC>
void launch_plain(const PlainGemm<T>& p, hipStream_t s) {
  constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  if constexpr (sizeof(T) == 4) {
    if (plain_big_tiles(p)) {
      int64_t tiles = ceil_div(p.M, 256) * ceil_div(p.N, 256);
      plain_gemm_kernel<T, BK, A_KC, B_NC, 256><<<dim3((unsigned)tiles, (unsigned)p.batch), 256, 0, s>>>(p);
      CCX_HIP(hipGetLastError());
      return;
    }
  }
  {
    int64_t tiles = ceil_div(p.M, 128) * ceil_div(p.N, 128);
    plain_gemm_kernel<T, BK, A_KC, B_NC, 128><<<dim3((unsigned)tiles, (unsigned)p.batch), 256, 0, s>>>(p);
  }
  CCX_HIP(hipGetLastError());
}

}  // namespace

template <typename T>
bool plain_gemm_ok(const PlainGemm<T>& p) {
  const int64_t VEC = 16 / sizeof(T);
  if (p.M <= 0 || p.N <= 0 || p.batch < 1 || p.batch > 65535) return false;
  if (ceil_div(p.M, 128) * ceil_div(p.N, 128) >= (1ll << 31)) return false;
  auto aligned = [&](const void* q) { return ((uintptr_t)q % 16) == 0; };
  if (!aligned(p.A) || !aligned(p.B)) return false;
  if (p.lda % VEC || p.ldb % VEC) return false;
  if (p.batch > 1 && (p.a_b % VEC || p.b_b % VEC)) return false;
  return true;
}

template <typename T>
void plain_gemm(const PlainGemm<T>& p, hipStream_t s) {
  CCX_CHECK(plain_gemm_ok(p), ERR_INTERNAL, "plain_gemm called on an unsupported layout");
  if (p.a_kc && p.b_nc) launch_plain<T, true, true>(p, s);
  else if (p.a_kc) launch_plain<T, true, false>(p, s);
  else if (p.b_nc) launch_plain<T, false, true>(p, s);
  else launch_plain<T, false, false>(p, s);
}

void plain_gemm_kcat(const PlainGemm<double>& p, hipStream_t s) {
  if (p.M == 0 || p.N == 0 || p.batch == 0) return;
  CCX_CHECK(plain_gemm_ok(p) && p.a_kc && !p.b_nc && p.ktot == 0 && p.A2 && p.K1 > 0 && p.K1 % 16 == 0 &&
                p.K1 < p.K && p.sw > 0 && p.sw * p.sw == p.M && p.M * p.lda < (int64_t(1) << 31) &&
                ((uintptr_t)p.A2 % 16) == 0 && p.a2_b % 2 == 0,
            ERR_INTERNAL, "plain_gemm_kcat layout");
  const int64_t tiles = ceil_div(p.M, 128) * ceil_div(p.N, 128);
  plain_gemm_kernel<double, 16, true, false, 128, true><<<dim3((unsigned)tiles, (unsigned)p.batch), 256, 0, s>>>(p);
  CCX_HIP(hipGetLastError());
}

template bool plain_gemm_ok<double>(const PlainGemm<double>&);
template bool plain_gemm_ok<float>(const PlainGemm<float>&);
template void plain_gemm<double>(const PlainGemm<double>&, hipStream_t);
template void plain_gemm<float>(const PlainGemm<float>&, hipStream_t);

}  // namespace ccx

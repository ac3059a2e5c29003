// MFMA tensor-contraction GEMM for gfx950.
//
// One kernel serves every contraction of the CCSD residual: each GEMM dimension (M, N, K) is a
// 2-level index (two tensor indices with independent strides) and there is one batch index, so
// e.g. 'jnfb,mnef->mbej' runs without materialising permuted copies. Operands are staged through
// LDS in k-major tiles; each of the 4 waves owns a (BM/2) x (BN/2) block of 16x16 MFMA tiles.
//
// Lane layout of the 16x16x4 MFMAs used here (measured on gfx950):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   C (16x16): lane l, register r holds C[row][l % 16] with
//              row = 4 * (l / 16) + r   for v_mfma_f32_16x16x4f32
//              row = 4 * r + (l / 16)   for v_mfma_f64_16x16x4f64
#include "common.h"
#include "gemm.h"

namespace ccx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct AccT;
template <> struct AccT<double> { typedef d4 type; };
template <> struct AccT<float> { typedef f4 type; };

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int64_t off2(const Idx2& d, int64_t i) {
  // all extents are < 2^31, so 32-bit division is exact here
  uint32_t q = (uint32_t)i / (uint32_t)d.n2;
  uint32_t r = (uint32_t)i - q * (uint32_t)d.n2;
  return (int64_t)q * d.s1 + (int64_t)r * d.s2;
}

// Map the hardware workgroup id (dispatched round-robin over the 8 XCDs) to a logical tile id so
// that each XCD receives one contiguous run of logical tiles (and so shares operands in its L2).
__device__ __forceinline__ int64_t xcd_remap(int64_t w, int64_t ntiles) {
  const int64_t nx = 8;
  int64_t q = ntiles / nx, r = ntiles % nx;
  int64_t x = w % nx, j = w / nx;
  return x * q + (x < r ? x : r) + j;
}

template <typename T, int BM, int BN, int BK, bool A_KFAST, bool B_KFAST>
__global__ __launch_bounds__(256) void tc_gemm_kernel(GemmParams<T> p, T* __restrict__ work,
                                                      int splits, int64_t kchunk) {
  constexpr int NT = 256;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int NBI = WM / 16, NBJ = WN / 16;
  constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
  constexpr int NA = BM * BK / NT, NB = BN * BK / NT;
  static_assert(NA >= 1 && NB >= 1, "tile too small for 256 threads");
  static_assert(NT % BK == 0 && NT % BM == 0 && NT % BN == 0, "tile shape");
  typedef typename AccT<T>::type acc_t;

  __shared__ T As[2][BK][BM + PAD];
  __shared__ T Bs[2][BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  const int64_t ntiles = tiles_m * tiles_n;
  int64_t t = xcd_remap(blockIdx.x, ntiles);
  // group 8 tile-rows together so that concurrently running tiles share A and B panels
  const int64_t GM = 8;
  int64_t group = t / (GM * tiles_n);
  int64_t first_m = group * GM;
  int64_t gsize = tiles_m - first_m < GM ? tiles_m - first_m : GM;
  int64_t within = t - group * GM * tiles_n;
  const int64_t tm = first_m + within % gsize;
  const int64_t tn = within / gsize;

  const int64_t bz = blockIdx.y;
  const int64_t bidx = bz / splits, sidx = bz - bidx * splits;
  const int64_t k_begin = sidx * kchunk;
  const int64_t k_end = (k_begin + kchunk < p.K) ? k_begin + kchunk : p.K;

  const T* __restrict__ Ab = p.A + bidx * p.a_b;
  const T* __restrict__ Bb = p.B + bidx * p.b_b;

  // ---- per-thread load geometry ----
  // A tile: BM x BK. K-fast: thread owns one k column and NA rows; M-fast: one row, NA k's.
  int a_kl, a_ml;            // local coordinates of the first element
  int64_t a_moff[A_KFAST ? NA : 1];
  bool a_mok[A_KFAST ? NA : 1];
  if (A_KFAST) {
    a_kl = tid % BK; a_ml = tid / BK;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int64_t m = tm * BM + a_ml + i * (NT / BK);
      a_mok[i] = m < p.M;
      a_moff[i] = a_mok[i] ? off2(p.am, m) : 0;
    }
  } else {
    a_ml = tid % BM; a_kl = tid / BM;
    int64_t m = tm * BM + a_ml;
    a_mok[0] = m < p.M;
    a_moff[0] = a_mok[0] ? off2(p.am, m) : 0;
  }
  int b_kl, b_nl;
  int64_t b_noff[B_KFAST ? NB : 1];
  bool b_nok[B_KFAST ? NB : 1];
  if (B_KFAST) {
    b_kl = tid % BK; b_nl = tid / BK;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int64_t n = tn * BN + b_nl + i * (NT / BK);
      b_nok[i] = n < p.N;
      b_noff[i] = b_nok[i] ? off2(p.bn, n) : 0;
    }
  } else {
    b_nl = tid % BN; b_kl = tid / BN;
    int64_t n = tn * BN + b_nl;
    b_nok[0] = n < p.N;
    b_noff[0] = b_nok[0] ? off2(p.bn, n) : 0;
  }

  T ra[NA], rb[NB];

  auto load_a = [&](int64_t k0) {
    if (A_KFAST) {
      int64_t k = k0 + a_kl;
      bool kok = k < k_end;
      int64_t ko = kok ? off2(p.ak, k) : 0;
#pragma unroll
      for (int i = 0; i < NA; ++i) ra[i] = (kok && a_mok[i]) ? Ab[a_moff[i] + ko] : T(0);
    } else {
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        int64_t k = k0 + a_kl + i * (NT / BM);
        bool ok = a_mok[0] && k < k_end;
        ra[i] = ok ? Ab[a_moff[0] + off2(p.ak, k)] : T(0);
      }
    }
  };
  auto load_b = [&](int64_t k0) {
    if (B_KFAST) {
      int64_t k = k0 + b_kl;
      bool kok = k < k_end;
      int64_t ko = kok ? off2(p.bk, k) : 0;
#pragma unroll
      for (int i = 0; i < NB; ++i) rb[i] = (kok && b_nok[i]) ? Bb[b_noff[i] + ko] : T(0);
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int64_t k = k0 + b_kl + i * (NT / BN);
        bool ok = b_nok[0] && k < k_end;
        rb[i] = ok ? Bb[b_noff[0] + off2(p.bk, k)] : T(0);
      }
    }
  };
  auto store = [&](int buf) {
    if (A_KFAST) {
#pragma unroll
      for (int i = 0; i < NA; ++i) As[buf][a_kl][a_ml + i * (NT / BK)] = ra[i];
    } else {
#pragma unroll
      for (int i = 0; i < NA; ++i) As[buf][a_kl + i * (NT / BM)][a_ml] = ra[i];
    }
    if (B_KFAST) {
#pragma unroll
      for (int i = 0; i < NB; ++i) Bs[buf][b_kl][b_nl + i * (NT / BK)] = rb[i];
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) Bs[buf][b_kl + i * (NT / BN)][b_nl] = rb[i];
    }
  };

  acc_t acc[NBI][NBJ];
#pragma unroll
  for (int i = 0; i < NBI; ++i)
#pragma unroll
    for (int j = 0; j < NBJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int fr = lane >> 4;   // k row within an MFMA step
  const int fc = lane & 15;   // m / n column within a 16-wide block
  // output row of register r within a 16x16 block (see the layout note at the top)
  auto orow = [&](int r) { return sizeof(T) == 8 ? 4 * r + fr : 4 * fr + r; };

  if (k_begin < k_end) {
    load_a(k_begin);
    load_b(k_begin);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = k_begin; k0 < k_end; k0 += BK) {
      const bool has_next = k0 + BK < k_end;
      if (has_next) { load_a(k0 + BK); load_b(k0 + BK); }
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T a[NBI], b[NBJ];
#pragma unroll
        for (int i = 0; i < NBI; ++i) a[i] = As[buf][ks + fr][wm * WM + i * 16 + fc];
#pragma unroll
        for (int j = 0; j < NBJ; ++j) b[j] = Bs[buf][ks + fr][wn * WN + j * 16 + fc];
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
          for (int j = 0; j < NBJ; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
      }
      if (has_next) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }

  // ---- epilogue ----
  if (splits == 1) {
    T* Cb = p.C + bidx * p.c_b;
#pragma unroll
    for (int j = 0; j < NBJ; ++j) {
      int64_t n = tn * BN + wn * WN + j * 16 + fc;
      if (n >= p.N) continue;
      int64_t no = off2(p.cn, n);
#pragma unroll
      for (int i = 0; i < NBI; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int64_t m = tm * BM + wm * WM + i * 16 + orow(r);
          if (m >= p.M) continue;
          T* dst = Cb + off2(p.cm, m) + no;
          T v = p.alpha * acc[i][j][r];
          if (p.beta != T(0)) v += p.beta * *dst;
          *dst = v;
        }
      }
    }
  } else {
    // dense partial tile: work[s][b][m][n]
    T* Wb = work + (sidx * p.batch + bidx) * p.M * p.N;
#pragma unroll
    for (int j = 0; j < NBJ; ++j) {
      int64_t n = tn * BN + wn * WN + j * 16 + fc;
      if (n >= p.N) continue;
#pragma unroll
      for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int64_t m = tm * BM + wm * WM + i * 16 + orow(r);
          if (m < p.M) Wb[m * p.N + n] = acc[i][j][r];
        }
    }
  }
}

// C = alpha * sum_s work[s] + beta * C, in a fixed summation order (deterministic).
template <typename T>
__global__ void splitk_reduce_kernel(GemmParams<T> p, const T* __restrict__ work, int splits) {
  const int64_t total = p.batch * p.M * p.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = e / (p.M * p.N);
    int64_t r = e - b * p.M * p.N;
    int64_t m = r / p.N, n = r - m * p.N;
    T s = 0;
    for (int k = 0; k < splits; ++k) s += work[(k * p.batch) * p.M * p.N + e];
    T* dst = p.C + b * p.c_b + off2(p.cm, m) + off2(p.cn, n);
    T v = p.alpha * s;
    if (p.beta != T(0)) v += p.beta * *dst;
    *dst = v;
  }
}

template <typename T, int BM, int BN, int BK>
static void launch_tiles(const GemmParams<T>& p, bool akf, bool bkf, T* work, int splits,
                         int64_t kchunk, hipStream_t s) {
  int64_t ntiles = ceil_div(p.M, BM) * ceil_div(p.N, BN);
  dim3 grid((unsigned)ntiles, (unsigned)(p.batch * splits));
  if (akf && bkf) tc_gemm_kernel<T, BM, BN, BK, true, true><<<grid, 256, 0, s>>>(p, work, splits, kchunk);
  else if (akf) tc_gemm_kernel<T, BM, BN, BK, true, false><<<grid, 256, 0, s>>>(p, work, splits, kchunk);
  else if (bkf) tc_gemm_kernel<T, BM, BN, BK, false, true><<<grid, 256, 0, s>>>(p, work, splits, kchunk);
  else tc_gemm_kernel<T, BM, BN, BK, false, false><<<grid, 256, 0, s>>>(p, work, splits, kchunk);
  CCX_HIP(hipGetLastError());
}

template <typename T>
void tc_gemm(const GemmParams<T>& p, Workspace& ws, hipStream_t s) {
  if (p.M == 0 || p.N == 0 || p.batch == 0) return;
  CCX_CHECK(p.M < (1ll << 31) && p.N < (1ll << 31) && p.K < (1ll << 31), ERR_ARG, "gemm extent too large");
  CCX_CHECK(p.batch < 65536, ERR_ARG, "gemm batch too large");
  // A is "k-fast" when its inner k index is contiguous (coalesced loads along k); otherwise the
  // loader walks along m. Same for B with n.
  const bool akf = !(p.am.s2 == 1 && p.ak.s2 != 1);
  const bool bkf = !(p.bn.s2 == 1 && p.bk.s2 != 1);
  const bool small = p.M <= 64 || p.N <= 64;
  const int BM = small ? 64 : 128, BN = small ? 64 : 128, BKt = 16;
  int64_t tiles = ceil_div(p.M, BM) * ceil_div(p.N, BN) * p.batch;
  CCX_CHECK(tiles < (1ll << 31), ERR_ARG, "gemm grid too large");
  // split K when the tile grid cannot fill the 256 CUs and K is long
  int splits = 1;
  if (p.K > 0 && tiles < 512) {
    int64_t want = ceil_div(512, tiles);
    int64_t maxs = p.K / (BKt * 8);
    splits = (int)std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(maxs, 64)));
    splits = (int)std::max<int64_t>(1, std::min<int64_t>(splits, 65535 / p.batch));   // grid.y limit
  }
  int64_t kchunk = p.K;
  T* work = nullptr;
  if (splits > 1) {
    kchunk = ceil_div(ceil_div(p.K, splits), BKt) * BKt;
    splits = (int)ceil_div(p.K, kchunk);
    size_t bytes = (size_t)splits * p.batch * p.M * p.N * sizeof(T);
    if (bytes > ws.limit) {       // keep the workspace bounded; fall back to fewer splits
      splits = (int)std::max<int64_t>(1, ws.limit / ((size_t)p.batch * p.M * p.N * sizeof(T)));
      kchunk = ceil_div(ceil_div(p.K, splits), BKt) * BKt;
      splits = (int)ceil_div(p.K, kchunk);
    }
    if (splits > 1) work = (T*)ws.get((size_t)splits * p.batch * p.M * p.N * sizeof(T));
    else kchunk = p.K;
  }
  if (p.K == 0) kchunk = 0;
  if (small) launch_tiles<T, 64, 64, 16>(p, akf, bkf, work, splits, kchunk, s);
  else launch_tiles<T, 128, 128, 16>(p, akf, bkf, work, splits, kchunk, s);
  if (splits > 1) {
    int64_t total = p.batch * p.M * p.N;
    int blocks = (int)std::min<int64_t>(ceil_div(total, 256), 65536);
    splitk_reduce_kernel<T><<<blocks, 256, 0, s>>>(p, work, splits);
    CCX_HIP(hipGetLastError());
  }
}

template void tc_gemm<double>(const GemmParams<double>&, Workspace&, hipStream_t);
template void tc_gemm<float>(const GemmParams<float>&, Workspace&, hipStream_t);

}  // namespace ccx

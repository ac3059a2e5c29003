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
#include <cstdio>
#include <cstdlib>

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

// LDS tile of one operand: stored [rows][BK + 1] when the operand is loaded along k (k-fast) and
// [BK][rows + 1] when it is loaded along m / n. The +1 shifts consecutive rows by two 4-byte banks,
// so neither the stores nor the MFMA fragment reads conflict.
template <typename T, int R, int BK, bool KFAST>
struct LdsTile {
  T d[KFAST ? R : BK][(KFAST ? BK : R) + 1];
  __device__ __forceinline__ T& at(int r, int k) { return KFAST ? d[r][k] : d[k][r]; }
};

// Loader of one operand tile (R rows x BK k's) from a 2-level-indexed global array. Each thread
// owns NE = R * BK / 256 elements: k-fast -> one k column and NE rows; otherwise one row, NE k's.
template <typename T, int R, int BK, bool KFAST>
struct TileLoader {
  static constexpr int NT = 256;
  static constexpr int NE = R * BK / NT;
  int kl, rl;                      // local coordinates of the thread's first element
  int64_t roff[KFAST ? NE : 1];    // global offsets of the thread's rows
  bool rok[KFAST ? NE : 1];
  T v[NE];

  __device__ __forceinline__ void init(const Idx2& rix, int64_t row0, int64_t nrows, int tid) {
    if (KFAST) {
      kl = tid % BK; rl = tid / BK;
#pragma unroll
      for (int i = 0; i < NE; ++i) {
        int64_t r = row0 + rl + i * (NT / BK);
        rok[i] = r < nrows;
        roff[i] = rok[i] ? off2(rix, r) : 0;
      }
    } else {
      rl = tid % R; kl = tid / R;
      int64_t r = row0 + rl;
      rok[0] = r < nrows;
      roff[0] = rok[0] ? off2(rix, r) : 0;
    }
  }
  // Unguarded load of a k-tile that lies inside one k1 block (kix.n2 % BK == 0) and whose rows are
  // all valid: one offset computation per tile, then constant strides.
  __device__ __forceinline__ void load_fast(const T* __restrict__ base, const Idx2& kix, int64_t k0) {
    const int64_t kb = off2(kix, k0);
    if (KFAST) {
      const int64_t ko = kb + (int64_t)kl * kix.s2;
#pragma unroll
      for (int i = 0; i < NE; ++i) v[i] = base[roff[i] + ko];
    } else {
#pragma unroll
      for (int i = 0; i < NE; ++i) v[i] = base[roff[0] + kb + (int64_t)(kl + i * (NT / R)) * kix.s2];
    }
  }
  // Guarded load for edge tiles / k-tiles that straddle a k1 block.
  __device__ __forceinline__ void load_safe(const T* __restrict__ base, const Idx2& kix, int64_t k0, int64_t kend) {
    if (KFAST) {
      int64_t k = k0 + kl;
      bool kok = k < kend;
      int64_t ko = kok ? off2(kix, k) : 0;
#pragma unroll
      for (int i = 0; i < NE; ++i) v[i] = (kok && rok[i]) ? base[roff[i] + ko] : T(0);
    } else {
#pragma unroll
      for (int i = 0; i < NE; ++i) {
        int64_t k = k0 + kl + i * (NT / R);
        v[i] = (rok[0] && k < kend) ? base[roff[0] + off2(kix, k)] : T(0);
      }
    }
  }
  __device__ __forceinline__ void store(LdsTile<T, R, BK, KFAST>& t) {
    if (KFAST) {
#pragma unroll
      for (int i = 0; i < NE; ++i) t.at(rl + i * (NT / BK), kl) = v[i];
    } else {
#pragma unroll
      for (int i = 0; i < NE; ++i) t.at(rl, kl + i * (NT / R)) = v[i];
    }
  }
};

template <typename T, int BM, int BN, int BK, bool A_KFAST, bool B_KFAST>
__global__ __launch_bounds__(256, 2) void tc_gemm_kernel(GemmParams<T> p, T* __restrict__ work,
                                                         int splits, int64_t kchunk) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int NBI = WM / 16, NBJ = WN / 16;
  typedef typename AccT<T>::type acc_t;

  __shared__ LdsTile<T, BM, BK, A_KFAST> As[2];
  __shared__ LdsTile<T, BN, BK, B_KFAST> Bs[2];

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

  TileLoader<T, BM, BK, A_KFAST> la;
  TileLoader<T, BN, BK, B_KFAST> lb;
  la.init(p.am, tm * BM, p.M, tid);
  lb.init(p.bn, tn * BN, p.N, tid);
  // interior tile whose k-tiles never straddle a k1 block: unguarded loads
  const bool fast = (tm + 1) * BM <= p.M && (tn + 1) * BN <= p.N && (p.ak.n2 % BK == 0) &&
                    (k_begin % BK == 0);

  acc_t acc[NBI][NBJ];
#pragma unroll
  for (int i = 0; i < NBI; ++i)
#pragma unroll
    for (int j = 0; j < NBJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int fr = lane >> 4;   // k row within an MFMA step
  const int fc = lane & 15;   // m / n column within a 16-wide block
  // output row of register r within a 16x16 block (see the layout note at the top)
  auto orow = [&](int r) { return sizeof(T) == 8 ? 4 * r + fr : 4 * fr + r; };

  auto load = [&](int64_t k0) {
    if (fast && k0 + BK <= k_end) {
      la.load_fast(Ab, p.ak, k0);
      lb.load_fast(Bb, p.bk, k0);
    } else {
      la.load_safe(Ab, p.ak, k0, k_end);
      lb.load_safe(Bb, p.bk, k0, k_end);
    }
  };

  if (k_begin < k_end) {
    load(k_begin);
    la.store(As[0]);
    lb.store(Bs[0]);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = k_begin; k0 < k_end; k0 += BK) {
      const bool has_next = k0 + BK < k_end;
      if (has_next) load(k0 + BK);
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T a[NBI], b[NBJ];
#pragma unroll
        for (int i = 0; i < NBI; ++i) a[i] = As[buf].at(wm * WM + i * 16 + fc, ks + fr);
#pragma unroll
        for (int j = 0; j < NBJ; ++j) b[j] = Bs[buf].at(wn * WN + j * 16 + fc, ks + fr);
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
          for (int j = 0; j < NBJ; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
      }
      if (has_next) {
        la.store(As[buf ^ 1]);
        lb.store(Bs[buf ^ 1]);
      }
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

// C[m * ldc + n] = alpha * sum_s work[s][m][n] + beta * C, fixed summation order (deterministic).
template <typename T>
__global__ void plain_reduce_kernel(int64_t M, int64_t N, int splits, const T* __restrict__ work,
                                    T* __restrict__ C, int64_t ldc, T alpha, T beta) {
  const int64_t total = M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    T s = 0;
    for (int k = 0; k < splits; ++k) s += work[k * total + e];
    const int64_t m = e / N, n = e - m * N;
    T* dst = C + m * ldc + n;
    T v = alpha * s;
    if (beta != T(0)) v += beta * *dst;
    *dst = v;
  }
}

// Plain GEMM, split along K into workspace partials when its tile grid cannot fill the chip.
template <typename T>
static void plain_dispatch(const PlainGemm<T>& q, Workspace& ws, hipStream_t s) {
  const int64_t BKp = sizeof(T) == 8 ? 16 : 32;
  const int64_t tiles = ceil_div(q.M, 128) * ceil_div(q.N, 128);
  int64_t splits = 1;
  if (q.batch == 1 && tiles < 768) {
    splits = std::min<int64_t>(ceil_div(1024, tiles), std::min<int64_t>(q.K / (BKp * 16), 64));
    const size_t one = (size_t)q.M * q.N * sizeof(T);
    splits = std::min<int64_t>(splits, (int64_t)(ws.limit / one));
  }
  if (splits > 1) {
    const int64_t kchunk = ceil_div(ceil_div(q.K, splits), BKp) * BKp;
    splits = ceil_div(q.K, kchunk);
  }
  if (std::getenv("CCX_TRACE"))
    std::fprintf(stderr, "[ccx]   -> plain %s%s batch %lld splits %lld\n", q.a_kc ? "N" : "T", q.b_nc ? "N" : "T",
                 (long long)q.batch, (long long)splits);
  if (splits <= 1) {
    plain_gemm<T>(q, s);
    return;
  }
  const int64_t kchunk = ceil_div(ceil_div(q.K, splits), BKp) * BKp;
  T* work = (T*)ws.get((size_t)splits * q.M * q.N * sizeof(T));
  PlainGemm<T> w = q;
  w.batch = splits;
  w.ktot = q.K;
  w.K = kchunk;
  w.a_b = q.a_kc ? kchunk : kchunk * q.lda;
  w.b_b = q.b_nc ? kchunk * q.ldb : kchunk;
  w.C = work;
  w.ldc = q.N;
  w.c_b = q.M * q.N;
  w.alpha = T(1);
  w.beta = T(0);
  plain_gemm<T>(w, s);
  const int blocks = (int)std::min<int64_t>(ceil_div(q.M * q.N, 256), 65536);
  plain_reduce_kernel<T><<<blocks, 256, 0, s>>>(q.M, q.N, (int)splits, work, q.C, q.ldc, q.alpha, q.beta);
  CCX_HIP(hipGetLastError());
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

// Single stride of a 2-level index over `extent` values, or -1 when it is not a plain stride.
static int64_t plain_stride(const Idx2& d, int64_t extent) {
  if (extent <= 1) return 0;
  if (extent <= d.n2) return d.s2;
  if (d.s1 == d.n2 * d.s2) return d.s2;
  return -1;
}

// Express the contraction as a plain strided GEMM when every operand is an ordinary matrix.
template <typename T>
static bool as_plain(const GemmParams<T>& p, PlainGemm<T>& q) {
  const int64_t am = plain_stride(p.am, p.M), ak = plain_stride(p.ak, p.K);
  const int64_t bk = plain_stride(p.bk, p.K), bn = plain_stride(p.bn, p.N);
  const int64_t cm = plain_stride(p.cm, p.M), cn = plain_stride(p.cn, p.N);
  if (am < 0 || ak < 0 || bk < 0 || bn < 0 || cm < 0 || cn < 0 || p.K < 1) return false;
  q = PlainGemm<T>{};
  q.batch = p.batch;
  q.alpha = p.alpha;
  q.beta = p.beta;
  if (cn == 1 || p.N == 1) {          // C row-major
    q.M = p.M; q.N = p.N; q.K = p.K;
    q.A = p.A; q.B = p.B; q.C = p.C;
    q.a_b = p.a_b; q.b_b = p.b_b; q.c_b = p.c_b;
    q.ldc = cm;
    if (ak == 1) { q.a_kc = true; q.lda = am; } else if (am == 1) { q.a_kc = false; q.lda = ak; } else return false;
    if (bn == 1) { q.b_nc = true; q.ldb = bk; } else if (bk == 1) { q.b_nc = false; q.ldb = bn; } else return false;
  } else if (cm == 1) {               // C column-major: C^T = B^T A^T
    q.M = p.N; q.N = p.M; q.K = p.K;
    q.A = p.B; q.B = p.A; q.C = p.C;
    q.a_b = p.b_b; q.b_b = p.a_b; q.c_b = p.c_b;
    q.ldc = cn;
    if (bk == 1) { q.a_kc = true; q.lda = bn; } else if (bn == 1) { q.a_kc = false; q.lda = bk; } else return false;
    if (am == 1) { q.b_nc = true; q.ldb = ak; } else if (ak == 1) { q.b_nc = false; q.ldb = am; } else return false;
  } else {
    return false;
  }
  if (q.lda < 1 || q.ldb < 1 || q.ldc < 1) return false;
  return plain_gemm_ok(q);
}

template <typename T>
void tc_gemm(const GemmParams<T>& p, Workspace& ws, hipStream_t s) {
  if (p.M == 0 || p.N == 0 || p.batch == 0) return;
  {
    // enough 128 x 128 tiles to fill the 256 CUs: the plain-matrix kernel
    PlainGemm<T> q;
    if (ceil_div(p.M, 128) * ceil_div(p.N, 128) * p.batch >= 240 && as_plain(p, q)) {
      plain_dispatch<T>(q, ws, s);
      return;
    }
  }
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
  if (p.K > 0 && tiles < 1024) {
    // fill ~4 workgroups per CU; every split keeps at least 8 k-tiles
    int64_t want = ceil_div(1024, tiles);
    int64_t maxs = p.K / (BKt * 8);
    splits = (int)std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(maxs, 256)));
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
  if (std::getenv("CCX_TRACE"))
    std::fprintf(stderr, "[ccx]   -> generic %s tile%s%s batch %lld splits %d\n", small ? "64" : "128", akf ? " Ak" : " Am",
                 bkf ? " Bk" : " Bn", (long long)p.batch, splits);
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

// Fused elementwise, packing and reduction kernels of the CCSD residual (gfx950).
//
// These are HBM-bound: each reads its inputs once and writes its outputs once. Transposed reads
// of t2 blocks go through 32x33 LDS tiles so that global loads and stores stay coalesced.
#include "kernels.h"
#include <algorithm>
#include <cmath>

namespace ccx {

namespace {

constexpr int RED_BLOCKS = 1024;   // fixed grid for deterministic two-stage reductions

__device__ __forceinline__ void decodeP(int64_t c, int64_t& a, int64_t& b) {
  int64_t f = (int64_t)((sqrt(8.0 * (double)c + 1.0) - 1.0) * 0.5);
  while (f * (f + 1) / 2 > c) --f;
  while ((f + 1) * (f + 2) / 2 <= c) ++f;
  b = f;
  a = c - f * (f + 1) / 2;
}
__device__ __forceinline__ void decodeM(int64_t c, int64_t& a, int64_t& b) {
  int64_t f = (int64_t)((sqrt(8.0 * (double)c + 1.0) + 1.0) * 0.5);
  while (f * (f - 1) / 2 > c) --f;
  while ((f + 1) * f / 2 <= c) ++f;
  b = f;
  a = c - f * (f - 1) / 2;
}

__device__ __forceinline__ double block_sum(double x) {
  __shared__ double sh[256];
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void finish_sum_kernel(const double* __restrict__ partials, int n, double* out) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x == 0) *out = s;
}

inline int grid_for(int64_t n, int64_t cap = 65536) {
  int64_t b = ceil_div(n, 256);
  return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
__global__ void gather_chem_kernel(const double* __restrict__ eri, int64_t nmo, int64_t p0, int64_t np,
                                   int64_t q0, int64_t nq, int64_t r0, int64_t nr, int64_t s0, int64_t ns,
                                   double* __restrict__ out) {
  int64_t total = np * nq * nr * ns;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = e % ns, t = e / ns;
    int64_t r = t % nr; t /= nr;
    int64_t q = t % nq, p = t / nq;
    out[e] = eri[(((p0 + p) * nmo + (r0 + r)) * nmo + (q0 + q)) * nmo + (s0 + s)];
  }
}

void gather_chem(const double* eri, int64_t nmo, int64_t p0, int64_t np, int64_t q0, int64_t nq,
                 int64_t r0, int64_t nr, int64_t s0, int64_t ns, double* out, hipStream_t s) {
  int64_t total = np * nq * nr * ns;
  if (!total) return;
  gather_chem_kernel<<<grid_for(total), 256, 0, s>>>(eri, nmo, p0, np, q0, nq, r0, nr, s0, ns, out);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// tau = f1 t2 + f2 t1 (x) t1. One block per (pair ij, chunk of 256 VEC TAU_STEPS plane elements):
// the pair and the plane position come from 32-bit arithmetic on the block index, and each thread
// moves VEC consecutive doubles per step (16-byte loads and stores when the v x v plane has an even
// size, so that every plane starts on a 16-byte boundary); all loads of a thread are issued before
// its first store.
constexpr int TAU_STEPS = 4;
template <int VEC>
__global__ __launch_bounds__(256) void tau_kernel(int o, int v, int nchunk, double f1, double f2,
                                                  const double* __restrict__ t1, const double* __restrict__ t2,
                                                  double* __restrict__ tau) {
  const int ij = (int)(blockIdx.x / (unsigned)nchunk), chunk = (int)blockIdx.x - ij * nchunk;
  const int i = ij / o, j = ij - i * o;
  const uint32_t plane = (uint32_t)v * (uint32_t)v;
  const double* __restrict__ ti = t1 + (int64_t)i * v;
  const double* __restrict__ tj = t1 + (int64_t)j * v;
  const int64_t base = (int64_t)ij * plane;
  const uint32_t e0 = (uint32_t)chunk * (256u * VEC * TAU_STEPS) + threadIdx.x * VEC;
  if (VEC == 2) {
    double2 x[TAU_STEPS];
#pragma unroll
    for (int k = 0; k < TAU_STEPS; ++k) {
      const uint32_t e = e0 + k * 256u * VEC;
      if (e < plane) x[k] = *(const double2*)(t2 + base + e);
    }
#pragma unroll
    for (int k = 0; k < TAU_STEPS; ++k) {
      const uint32_t e = e0 + k * 256u * VEC;
      if (e < plane) {
        const uint32_t a0 = e / (uint32_t)v, b0 = e - a0 * v;
        const uint32_t a1 = (e + 1) / (uint32_t)v, b1 = e + 1 - a1 * v;
        double2 y;
        y.x = f1 * x[k].x + f2 * (ti[a0] * tj[b0]);
        y.y = f1 * x[k].y + f2 * (ti[a1] * tj[b1]);
        *(double2*)(tau + base + e) = y;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < TAU_STEPS; ++k) {
      const uint32_t e = e0 + k * 256u;
      if (e < plane) {
        const uint32_t a = e / (uint32_t)v, b = e - a * v;
        tau[base + e] = f1 * t2[base + e] + f2 * (ti[a] * tj[b]);
      }
    }
  }
}

void build_tau(int64_t o, int64_t v, double f1, double f2, const double* t1, const double* t2,
               double* tau, hipStream_t s) {
  if (o == 0 || v == 0) return;
  CCX_CHECK(v * v < (1ll << 31), ERR_ARG, "build_tau: v x v plane too large");
  const int vec = (v % 2 == 0) ? 2 : 1;
  const int64_t nchunk = ceil_div(v * v, 256 * vec * TAU_STEPS);
  CCX_CHECK(o * o * nchunk < (1ll << 31), ERR_ARG, "build_tau: grid too large");
  const unsigned grid = (unsigned)(o * o * nchunk);
  if (vec == 2) tau_kernel<2><<<grid, 256, 0, s>>>((int)o, (int)v, (int)nchunk, f1, f2, t1, t2, tau);
  else tau_kernel<1><<<grid, 256, 0, s>>>((int)o, (int)v, (int)nchunk, f1, f2, t1, t2, tau);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// For every (i, a, j, b), with T = t2_ijab, X = t2_ijba, P = t_ja t_ib and the compound indices
// (ia), (jb) of ov x ov matrices, writes
//   S1[(ia)][0][(jb)] = -T/2      S1[(ia)][1][(jb)] = X/2 + P   (B of the J GEMM -> Wej | Wje)
//   A1[(ia)][(jb)]    = T - X/2 - P                            (B of the K GEMM -> Wej)
//   S2[(ia)][(jb)][0] = 2T - X    S2[(ia)][(jb)][1] = T         (A of the ring GEMM, K = (me, x))
//   Tx[(ia)][(jb)]    = X                                      (A of the crossed ring GEMM)
//   Ud[(ia)][(jb)]    = 2T - X  in FP64                        (r_T1 terms with 2 t2 - t2^T)
// The interleavings make every ring GEMM a plain row-major matrix product.
template <typename T>
__global__ __launch_bounds__(256) void ring_prep_kernel(int64_t o, int64_t v, const double* __restrict__ t1,
                                                        const double* __restrict__ t2, T* __restrict__ S1,
                                                        T* __restrict__ A1, T* __restrict__ S2,
                                                        T* __restrict__ Tx, double* __restrict__ Ud) {
  __shared__ double tn[32][33];
  __shared__ double tx[32][33];
  const int64_t ij = blockIdx.y;
  const int64_t i = ij / o, j = ij % o;
  const int64_t tiles_b = (v + 31) / 32;
  const int64_t a0 = (blockIdx.x / tiles_b) * 32, b0 = (blockIdx.x % tiles_b) * 32;
  const double* blk = t2 + ij * v * v;
  const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
  for (int y = y0; y < 32; y += 8) {
    int64_t a = a0 + y, b = b0 + x;
    tn[y][x] = (a < v && b < v) ? blk[a * v + b] : 0.0;
    int64_t bb = b0 + y, aa = a0 + x;
    tx[y][x] = (bb < v && aa < v) ? blk[bb * v + aa] : 0.0;
  }
  __syncthreads();
  const int64_t ov = o * v;
  for (int y = y0; y < 32; y += 8) {
    int64_t a = a0 + y, b = b0 + x;
    if (a >= v || b >= v) continue;
    double tv = tn[y][x];
    double xv = tx[x][y];
    double pv = t1[j * v + a] * t1[i * v + b];
    const int64_t ia = i * v + a, jb = j * v + b;
    const int64_t out = ia * ov + jb;
    S1[2 * ia * ov + jb] = (T)(-0.5 * tv);
    S1[2 * ia * ov + ov + jb] = (T)(0.5 * xv + pv);
    A1[out] = (T)(tv - 0.5 * xv - pv);
    S2[2 * out] = (T)(2.0 * tv - xv);
    S2[2 * out + 1] = (T)tv;
    Tx[out] = (T)xv;
    Ud[out] = 2.0 * tv - xv;
  }
}

template <typename T>
void ring_prep(int64_t o, int64_t v, const double* t1, const double* t2, T* S1, T* A1, T* S2, T* Tx,
               double* Ud, hipStream_t s) {
  int64_t tiles = ceil_div(v, 32);
  dim3 grid((unsigned)(tiles * tiles), (unsigned)(o * o));
  ring_prep_kernel<T><<<grid, 256, 0, s>>>(o, v, t1, t2, S1, A1, S2, Tx, Ud);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Row (i <= j) of Tp and, when i < j, row pM(i, j) of Tm.
template <typename T>
__global__ void tau_pack_kernel(int64_t o, int64_t v, const double* __restrict__ t1,
                                const double* __restrict__ t2, T* __restrict__ Tp, T* __restrict__ Tm) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  int64_t i, j;
  decodeP(blockIdx.y, i, j);
  const int64_t rowP = blockIdx.y;
  const bool hasM = i < j;
  const int64_t rowM = hasM ? pairM(i, j) : 0;
  const double* blk = t2 + (i * o + j) * v * v;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npP;
       c += (int64_t)gridDim.x * blockDim.x) {
    int64_t e, f;
    decodeP(c, e, f);
    double tef = blk[e * v + f], tfe = blk[f * v + e];
    if (t1) {
      tef += t1[i * v + e] * t1[j * v + f];
      tfe += t1[i * v + f] * t1[j * v + e];
    }
    Tp[rowP * npP + c] = (T)(e < f ? tef + tfe : tef);
    if (hasM && e < f) Tm[rowM * npM + pairM(e, f)] = (T)(tef - tfe);
  }
}

template <typename T>
void tau_pack(int64_t o, int64_t v, const double* t1, const double* t2, T* Tp, T* Tm, hipStream_t s) {
  int64_t npP = v * (v + 1) / 2;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(npP, 256), 1024), (unsigned)(o * (o + 1) / 2));
  tau_pack_kernel<T><<<grid, 256, 0, s>>>(o, v, t1, t2, Tp, Tm);
  CCX_HIP(hipGetLastError());
}

__device__ __forceinline__ void split_bf16(double x, bf16& hi, bf16& lo) {
  hi = __float2bfloat16((float)x);
  lo = __float2bfloat16((float)(x - (double)__bfloat162float(hi)));
}

__global__ void tau_pack_bf16x3_kernel(int64_t o, int64_t v, int64_t KpP, int64_t KpM,
                                       const double* __restrict__ t1, const double* __restrict__ t2,
                                       bf16* Tp_hi, bf16* Tp_lo, bf16* Tm_hi, bf16* Tm_lo) {
  const int64_t npP = v * (v + 1) / 2;
  int64_t i, j;
  decodeP(blockIdx.y, i, j);
  const int64_t rowP = blockIdx.y;
  const bool hasM = i < j;
  const int64_t rowM = hasM ? pairM(i, j) : 0;
  const double* blk = t2 + (i * o + j) * v * v;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npP;
       c += (int64_t)gridDim.x * blockDim.x) {
    int64_t e, f;
    decodeP(c, e, f);
    double tef = blk[e * v + f] + t1[i * v + e] * t1[j * v + f];
    double tfe = blk[f * v + e] + t1[i * v + f] * t1[j * v + e];
    split_bf16(e < f ? tef + tfe : tef, Tp_hi[rowP * KpP + c], Tp_lo[rowP * KpP + c]);
    if (hasM && e < f) {
      int64_t k = rowM * KpM + pairM(e, f);
      split_bf16(tef - tfe, Tm_hi[k], Tm_lo[k]);
    }
  }
}

void tau_pack_bf16x3(int64_t o, int64_t v, int64_t KpP, int64_t KpM, const double* t1, const double* t2,
                     bf16* Tp_hi, bf16* Tp_lo, bf16* Tm_hi, bf16* Tm_lo, hipStream_t s) {
  int64_t npP = v * (v + 1) / 2;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(npP, 256), 1024), (unsigned)(o * (o + 1) / 2));
  tau_pack_bf16x3_kernel<<<grid, 256, 0, s>>>(o, v, KpP, KpM, t1, t2, Tp_hi, Tp_lo, Tm_hi, Tm_lo);
  CCX_HIP(hipGetLastError());
}

// (ab) block of the BF16x3 ladder matrices: row (e<=f) = column (e<=f) by symmetry.
__global__ void ladder_pack_bf16x3_kernel(int64_t v, int64_t e0, VVVVSrc src, bf16* LPh, bf16* LPl,
                                          bf16* LMh, bf16* LMl, int64_t KpP, int64_t KpM) {
  const int64_t npP = v * (v + 1) / 2;
  const int64_t e = e0 + blockIdx.y / v, f = blockIdx.y % v;
  if (f < e) return;
  const int64_t rowP = pairP(e, f);
  const bool hasM = e < f;
  const int64_t rowM = hasM ? pairM(e, f) : 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npP;
       c += (int64_t)gridDim.x * blockDim.x) {
    int64_t a, b;
    decodeP(c, a, b);
    double x = src.base[e * src.se + a * src.sa + f * src.sf + b * src.sb];   // <ab|ef> = (ea|fb)
    double y = src.base[e * src.se + b * src.sa + f * src.sf + a * src.sb];   // <ab|fe> = (eb|fa)
    split_bf16(0.5 * (x + y), LPh[rowP * KpP + c], LPl[rowP * KpP + c]);
    if (hasM && a < b) {
      int64_t k = rowM * KpM + pairM(a, b);
      split_bf16(0.5 * (x - y), LMh[k], LMl[k]);
    }
  }
}

void ladder_pack_bf16x3(int64_t o, int64_t v, int64_t e0, int64_t e1, VVVVSrc src, bf16* LP_hi,
                        bf16* LP_lo, bf16* LM_hi, bf16* LM_lo, hipStream_t s) {
  if (e1 <= e0) return;
  int64_t npP = v * (v + 1) / 2;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(npP, 256), 512), (unsigned)((e1 - e0) * v));
  ladder_pack_bf16x3_kernel<<<grid, 256, 0, s>>>(v, e0, src, LP_hi, LP_lo, LM_hi, LM_lo,
                                                  bf16_kpad(npP), bf16_kpad(v * (v - 1) / 2));
  CCX_HIP(hipGetLastError());
}

// (m, b) rows of the BF16x3 ladder matrices: row npP + mb holds 1/2(<mb|ef> + <mb|fe>) over (e<=f).
__global__ void ladder_pack_o_bf16x3_kernel(int64_t o, int64_t v, const double* __restrict__ Covvv,
                                            bf16* LPh, bf16* LPl, bf16* LMh, bf16* LMl) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  const int64_t KpP = bf16_kpad(npP), KpM = bf16_kpad(npM);
  const int64_t mb = blockIdx.y, m = mb / v, b = mb % v;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < npP;
       k += (int64_t)gridDim.x * blockDim.x) {
    int64_t e, f;
    decodeP(k, e, f);
    double x = Covvv[((m * v + e) * v + b) * v + f];                            // <mb|ef> = (me|bf)
    double y = Covvv[((m * v + f) * v + b) * v + e];                            // <mb|fe> = (mf|be)
    split_bf16(0.5 * (x + y), LPh[(npP + mb) * KpP + k], LPl[(npP + mb) * KpP + k]);
    if (e < f) {
      int64_t q = (npM + mb) * KpM + pairM(e, f);
      split_bf16(0.5 * (x - y), LMh[q], LMl[q]);
    }
  }
}

void ladder_pack_o_bf16x3(int64_t o, int64_t v, const double* Covvv, bf16* LP_hi, bf16* LP_lo,
                          bf16* LM_hi, bf16* LM_lo, hipStream_t s) {
  int64_t npP = v * (v + 1) / 2;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(npP, 256), 256), (unsigned)(o * v));
  ladder_pack_o_bf16x3_kernel<<<grid, 256, 0, s>>>(o, v, Covvv, LP_hi, LP_lo, LM_hi, LM_lo);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
template <typename Store>
__global__ void ladder_pack_kernel(int64_t o, int64_t v, int64_t e0, VVVVSrc src,
                                   const double* __restrict__ Covvv, Store st) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  const int64_t ov = o * v;
  const int64_t e = e0 + blockIdx.y / v, f = blockIdx.y % v;
  if (f < e) return;
  const int64_t rowP = pairP(e, f);
  const bool hasM = e < f;
  const int64_t rowM = hasM ? pairM(e, f) : 0;
  const int64_t ldP = npP + ov, ldM = npM + ov;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npP + ov;
       c += (int64_t)gridDim.x * blockDim.x) {
    double x, y;
    bool mcol;
    int64_t cm;
    if (c < npP) {
      int64_t a, b;
      decodeP(c, a, b);
      x = src.base[e * src.se + a * src.sa + f * src.sf + b * src.sb];   // <ab|ef> = (ea|fb)
      y = src.base[e * src.se + b * src.sa + f * src.sf + a * src.sb];   // <ab|fe> = (eb|fa)
      mcol = a < b;
      cm = mcol ? pairM(a, b) : 0;
    } else {
      int64_t mb = c - npP, m = mb / v, b = mb - (mb / v) * v;
      x = Covvv[((m * v + e) * v + b) * v + f];                          // <mb|ef> = (me|bf)
      y = Covvv[((m * v + f) * v + b) * v + e];                          // <mb|fe> = (mf|be)
      mcol = true;
      cm = npM + mb;
    }
    st.put(rowP * ldP + c, 0.5 * (x + y), true);
    if (hasM && mcol) st.put(rowM * ldM + cm, 0.5 * (x - y), false);
  }
}

template <typename T>
struct StorePM {
  T* P;
  T* M;
  __device__ void put(int64_t k, double val, bool plus) const { (plus ? P : M)[k] = (T)val; }
};
template <typename T>
void ladder_pack(int64_t o, int64_t v, int64_t e0, int64_t e1, VVVVSrc src, const double* Covvv, T* LP,
                 T* LM, hipStream_t s) {
  if (e1 <= e0) return;
  int64_t ncol = v * (v + 1) / 2 + o * v;
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(ncol, 256), 512), (unsigned)((e1 - e0) * v));
  ladder_pack_kernel<StorePM<T>><<<grid, 256, 0, s>>>(o, v, e0, src, Covvv, StorePM<T>{LP, LM});
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void zt_add_kernel(int64_t o, int64_t v, const T* __restrict__ CP, const T* __restrict__ CM,
                              double* __restrict__ Zt) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  const int64_t ldP = npP + o * v, ldM = npM + o * v;
  const int64_t total = o * o * o * v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    // Zt layout [m][i][j][b]
    int64_t b = e % v, t = e / v;
    int64_t j = t % o; t /= o;
    int64_t i = t % o, m = t / o;
    int64_t mb = m * v + b;
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    double z = (double)CP[pairP(lo, hi) * ldP + npP + mb];
    if (i != j) {
      double zm = (double)CM[pairM(lo, hi) * ldM + npM + mb];
      z += (i < j) ? zm : -zm;
    }
    Zt[e] += z;
  }
}

template <typename T>
void zt_add(int64_t o, int64_t v, const T* CP, const T* CM, double* Zt, hipStream_t s) {
  zt_add_kernel<T><<<grid_for(o * v * o * o), 256, 0, s>>>(o, v, CP, CM, Zt);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// One block per (i, j) pair and 32 x 32 tile of (a, b). The three transposed operands
// (H[j][i][b][a], P[(jb),(ia)], Q[(ib),(ja)]) are read along their contiguous index and transposed
// through LDS, so every global access is coalesced.
// r2 is symmetric under (ij)(ab): one workgroup per pair i <= j and 32 x 32 tile (a, b) computes
//   r2[i][j][a][b] = H[ijab] + H[jiba] + ladder[ijab] + P[(ia),(jb)] + P[(jb),(ia)] + Q[(ja),(ib)]
//                    + Q[(ib),(ja)] - sum_m (t_ma Zt[m][i][j][b] + t_mb Zt[m][j][i][a])
//                    [+ PTt[i][j][b][a] + PTt[j][i][a][b]] [+ X[ijab] + X[jiba]]
//                    [+ I[ijab] + t2[ijab] Gd[b] + t2[jiba] Gd[a]]
// (the optional FP32 half-residual terms: PTt the t_ie <ab|ej> products of :939 stored with (b, a)
// planes, X the 1/2 tau Wmnij term of :930 plus the off-diagonal t2 G of :925; and in FP64 the
// <ij|ab> of :922 (both halves: <ij|ab> = <ji|ba>) and the diagonal of :925) once and writes it to
// both r2[i][j][a][b] and r2[j][i][b][a], so every input element is read once. For i = j only the
// tiles with a-tile <= b-tile run (the diagonal tile writes itself only).
// The terms indexed (j, i, b, a) are read along a and summed into one LDS tile; the Zt term
// (ccwfn.py:932, :937, :938, :940: the -t_ma Zt contraction of both halves) is one FP64 MFMA product
// per wave, K = the two m sums side by side (2o), operands straight from global memory in the lane
// layout (A: lane l holds A[l % 16][l / 16], B: lane l holds B[l / 16][l % 16], C: lane l register r
// holds C[4 r + l / 16][l % 16]); each lane then finishes the four elements of its C registers.
typedef double asm_d4 __attribute__((ext_vector_type(4)));
template <typename T>
__global__ __launch_bounds__(256) void assemble_r2_kernel(int64_t o, int64_t v, const double* __restrict__ H,
                                                          const T* __restrict__ CP, const T* __restrict__ CM,
                                                          const T* __restrict__ P, const T* __restrict__ Q,
                                                          const double* __restrict__ t1,
                                                          const double* __restrict__ Zt,
                                                          const float* __restrict__ PTt,
                                                          const float* __restrict__ X,
                                                          const double* __restrict__ I,
                                                          const double* __restrict__ T2d,
                                                          const double* __restrict__ Gd, double* __restrict__ r2) {
  __shared__ double part[32][33];   // [b - b0][a - a0]: the (j, i, b, a) terms; later the output tile
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2;
  const int64_t ldP = npP + o * v, ldM = npM + o * v, ov = o * v;
  int64_t i, j;
  decodeP(blockIdx.y, i, j);   // i <= j
  const int64_t tiles_b = (v + 31) / 32;
  const int64_t at = blockIdx.x / tiles_b, bt = blockIdx.x % tiles_b;
  if (i == j && at > bt) return;   // covered by the transposed write of tile (bt, at)
  const int64_t a0 = at * 32, b0 = bt * 32;
  {
    const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
    for (int y = y0; y < 32; y += 8) {
      const int64_t a = a0 + x, b = b0 + y;
      double p = 0.0;
      if (a < v && b < v) {
        p = H[((j * o + i) * v + b) * v + a] + (double)P[(j * v + b) * ov + i * v + a] +
            (double)Q[(i * v + b) * ov + j * v + a];
        if (PTt) p += (double)PTt[((i * o + j) * v + b) * v + a];
        if (X) p += (double)X[((j * o + i) * v + b) * v + a];
        if (T2d) p += T2d[((j * o + i) * v + b) * v + a] * Gd[a];
      }
      part[y][x] = p;
    }
  }
  // Zt term of this wave's 16 x 16 block: sum over k < 2o of A[a][k] B[k][b] with
  // A[a][m] = t1[m][a], B[m][b] = Zt[m][i][j][b] for k = m < o, and
  // A[a][o + m] = Zt[m][j][i][a], B[o + m][b] = t1[m][b]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ab0 = 16 * (w >> 1), bb0 = 16 * (w & 1);   // the wave's block inside the tile
  asm_d4 acc = {0.0, 0.0, 0.0, 0.0};
  {
    const int64_t am = a0 + ab0 + (lane & 15), bn = b0 + bb0 + (lane & 15);
    const bool aok = am < v, bok = bn < v;
    for (int64_t k0 = 0; k0 < 2 * o; k0 += 4) {
      const int64_t k = k0 + (lane >> 4);
      double av = 0.0, bv = 0.0;
      if (k < o) {
        av = aok ? t1[k * v + am] : 0.0;
        bv = bok ? Zt[((k * o + i) * o + j) * v + bn] : 0.0;
      } else if (k < 2 * o) {
        const int64_t m = k - o;
        av = aok ? Zt[((m * o + j) * o + i) * v + am] : 0.0;
        bv = bok ? t1[m * v + bn] : 0.0;
      }
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
  }
  __syncthreads();
  const T* cpr = CP + pairP(i, j) * ldP;
  const T* cmr = (i != j) ? CM + pairM(i, j) * ldM : nullptr;
  const bool sij = i < j;
  const int bl = bb0 + (lane & 15);   // this lane's b - b0
  const int64_t b = b0 + bl;
  const bool bok = b < v;
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int al = ab0 + 4 * k + (lane >> 4);   // a - a0 of register k
    const int64_t a = a0 + al;
    double e = 0.0;
    if (a < v && bok) {
      const int64_t ijab = ((i * o + j) * v + a) * v + b;
      e = H[ijab] + part[bl][al];
      const int64_t alo = a < b ? a : b, ahi = a < b ? b : a;
      e += (double)cpr[pairP(alo, ahi)];
      if (cmr && a != b) {
        const double m = (double)cmr[pairM(alo, ahi)];
        e += (sij == (a < b)) ? m : -m;
      }
      e += (double)P[(i * v + a) * ov + j * v + b] + (double)Q[(j * v + a) * ov + i * v + b];
      if (PTt) e += (double)PTt[((j * o + i) * v + a) * v + b];
      if (X) e += (double)X[ijab];
      if (I) e += I[ijab];
      if (T2d) e += T2d[ijab] * Gd[b];
      e -= acc[k];
    }
    r[k] = e;
  }
  // r2[i][j][a][b] directly (b fastest)
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t a = a0 + ab0 + 4 * k + (lane >> 4);
    if (a < v && bok) r2[((i * o + j) * v + a) * v + b] = r[k];
  }
  if (i == j && at == bt) return;   // the diagonal tile of a diagonal pair is its own transpose
  // r2[j][i][b][a] = the same values, through an LDS transpose (a fastest)
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) part[ab0 + 4 * k + (lane >> 4)][bl] = r[k];   // part[a - a0][b - b0]
  __syncthreads();
  {
    const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
    for (int y = y0; y < 32; y += 8) {
      const int64_t bb = b0 + y, a = a0 + x;
      if (bb < v && a < v) r2[((j * o + i) * v + bb) * v + a] = part[x][y];
    }
  }
}

template <typename T>
void assemble_r2(int64_t o, int64_t v, const double* H, const T* CP, const T* CM, const T* P,
                 const T* Q, const double* t1, const double* Zt, const float* PTt, const float* X,
                 const double* I, const double* T2d, const double* Gd, double* r2, hipStream_t s) {
  CCX_CHECK((T2d == nullptr) == (Gd == nullptr), ERR_INTERNAL, "assemble_r2: t2 and G's diagonal go together");
  const int64_t tiles = ceil_div(v, 32);
  CCX_CHECK(o * (o + 1) / 2 < 65536 && tiles * tiles < (1ll << 31), ERR_ARG, "assemble_r2: grid too large");
  dim3 grid((unsigned)(tiles * tiles), (unsigned)(o * (o + 1) / 2));
  assemble_r2_kernel<T><<<grid, 256, 0, s>>>(o, v, H, CP, CM, P, Q, t1, Zt, PTt, X, I, T2d, Gd, r2);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
__global__ void energy_kernel(int64_t o, int64_t v, const double* __restrict__ fock, int64_t nmo,
                              const double* __restrict__ t1, const double* __restrict__ t2,
                              const double* __restrict__ L, double* __restrict__ partials) {
  const int64_t total = o * o * v * v;
  double s = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = e % v, t = e / v;
    int64_t a = t % v; t /= v;
    int64_t j = t % o, i = t / o;
    s += (t2[e] + t1[i * v + a] * t1[j * v + b]) * L[e];
  }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < o * v;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / v, a = e % v;
    s += 2.0 * fock[i * nmo + o + a] * t1[e];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

void energy(int64_t o, int64_t v, const double* fock, int64_t nmo, const double* t1, const double* t2,
            const double* Loovv, double* partials, double* e_out, hipStream_t s) {
  energy_kernel<<<RED_BLOCKS, 256, 0, s>>>(o, v, fock, nmo, t1, t2, Loovv, partials);
  CCX_HIP(hipGetLastError());
  finish_sum_kernel<<<1, 256, 0, s>>>(partials, RED_BLOCKS, e_out);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
__global__ void jacobi_kernel(int64_t o, int64_t v, const double* __restrict__ eps, double* __restrict__ t1,
                              double* __restrict__ t2, const double* __restrict__ r1,
                              const double* __restrict__ r2, double* __restrict__ partials) {
  const int64_t total = o * o * v * v;
  double s = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = e % v, t = e / v;
    int64_t a = t % v; t /= v;
    int64_t j = t % o, i = t / o;
    double d = eps[i] + eps[j] - eps[o + a] - eps[o + b];
    double x = r2[e] / d;
    t2[e] += x;
    s += x * x;
  }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < o * v;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / v, a = e % v;
    double x = r1[e] / (eps[i] - eps[o + a]);
    t1[e] += x;
    s += x * x;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

void jacobi_update(int64_t o, int64_t v, const double* eps, double* t1, double* t2, const double* r1,
                   const double* r2, double* partials, double* rms2_out, hipStream_t s) {
  jacobi_kernel<<<RED_BLOCKS, 256, 0, s>>>(o, v, eps, t1, t2, r1, r2, partials);
  CCX_HIP(hipGetLastError());
  finish_sum_kernel<<<1, 256, 0, s>>>(partials, RED_BLOCKS, rms2_out);
  CCX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
__global__ void dots_kernel(int64_t n, const double* __restrict__ x0, int64_t stride,
                            const double* __restrict__ y, double* __restrict__ partials) {
  const double* x = x0 + blockIdx.y * stride;
  double s = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    s += x[e] * y[e];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void dots_finish_kernel(const double* __restrict__ partials, int nb, double* out) {
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partials[blockIdx.x * nb + i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

void dots(int64_t n, int nvec, const double* x0, int64_t stride, const double* y, double* partials,
          double* out, hipStream_t s) {
  if (nvec <= 0) return;
  const int nb = 256;
  dots_kernel<<<dim3(nb, nvec), 256, 0, s>>>(n, x0, stride, y, partials);
  CCX_HIP(hipGetLastError());
  dots_finish_kernel<<<nvec, 256, 0, s>>>(partials, nb, out);
  CCX_HIP(hipGetLastError());
}

__global__ void lincomb_kernel(int64_t n, int nvec, const double* __restrict__ c,
                               const double* __restrict__ x0, int64_t stride, double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int k = 0; k < nvec; ++k)
      if (c[k] != 0.0) s += c[k] * x0[k * stride + e];   // unused slots carry c = 0
    out[e] = s;
  }
}

void lincomb(int64_t n, int nvec, const double* c, const double* x0, int64_t stride, double* out,
             hipStream_t s) {
  lincomb_kernel<<<grid_for(n), 256, 0, s>>>(n, nvec, c, x0, stride, out);
  CCX_HIP(hipGetLastError());
}

__global__ void sub_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = x[e] - y[e];
}

constexpr int DIIS_BLOCKS = 1024;

__global__ __launch_bounds__(256) void diis_push_kernel(int64_t n1, int64_t n2, const double* __restrict__ t1,
                                                        const double* __restrict__ t2, double* oldt,
                                                        double* __restrict__ v, double* __restrict__ e,
                                                        const double* __restrict__ E, int64_t stride, DiisSlots d,
                                                        int self_slot, double* __restrict__ partials) {
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  const int64_t n = n1 + n2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double t = i < n1 ? t1[i] : t2[i - n1];
    const double err = t - oldt[i];
    v[i] = t;
    e[i] = err;
    oldt[i] = t;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < d.n) acc[k] += err * (d.slot[k] == self_slot ? err : E[d.slot[k] * stride + i]);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k >= d.n) break;
    const double r = block_sum(acc[k]);
    if (threadIdx.x == 0) partials[k * gridDim.x + blockIdx.x] = r;
  }
}

void diis_push(int64_t n1, int64_t n2, const double* t1, const double* t2, double* oldt, double* v, double* e,
               const double* E, DiisSlots d, double* partials, double* out, hipStream_t s) {
  CCX_CHECK(d.n >= 1 && d.n <= 8, ERR_INTERNAL, "diis_push takes 1..8 overlaps");
  const int64_t stride = n1 + n2;
  const int self_slot = (int)((e - E) / stride);
  diis_push_kernel<<<DIIS_BLOCKS, 256, 0, s>>>(n1, n2, t1, t2, oldt, v, e, E, stride, d, self_slot, partials);
  CCX_HIP(hipGetLastError());
  dots_finish_kernel<<<d.n, 256, 0, s>>>(partials, DIIS_BLOCKS, out);
  CCX_HIP(hipGetLastError());
}

__global__ void lincomb2_kernel(int64_t n, int nvec, const double* __restrict__ c, const double* __restrict__ x0,
                                int64_t stride, double* __restrict__ out, double* __restrict__ out2) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int k = 0; k < nvec; ++k)
      if (c[k] != 0.0) s += c[k] * x0[k * stride + e];   // unused slots carry c = 0
    out[e] = s;
    out2[e] = s;
  }
}

void lincomb2(int64_t n, int nvec, const double* c, const double* x0, int64_t stride, double* out, double* out2,
              hipStream_t s) {
  lincomb2_kernel<<<grid_for(n), 256, 0, s>>>(n, nvec, c, x0, stride, out, out2);
  CCX_HIP(hipGetLastError());
}

void sub(int64_t n, const double* x, const double* y, double* out, hipStream_t s) {
  sub_kernel<<<grid_for(n), 256, 0, s>>>(n, x, y, out);
  CCX_HIP(hipGetLastError());
}

template <typename Tin, typename Tout>
__global__ void cast_kernel(int64_t n, const Tin* __restrict__ in, Tout* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = (Tout)in[e];
}

template <typename Tin, typename Tout>
void cast(int64_t n, const Tin* in, Tout* out, hipStream_t s) {
  if (!n) return;
  cast_kernel<Tin, Tout><<<grid_for(n), 256, 0, s>>>(n, in, out);
  CCX_HIP(hipGetLastError());
}

__global__ void accumulate_kernel(int64_t n, double alpha, const float* __restrict__ in, double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] += alpha * (double)in[e];
}

void accumulate(int64_t n, double alpha, const float* in, double* out, hipStream_t s) {
  if (!n) return;
  accumulate_kernel<<<grid_for(n), 256, 0, s>>>(n, alpha, in, out);
  CCX_HIP(hipGetLastError());
}

// ---- bandwidth-bound weighted sums (the GEMV-shaped FP64 contractions of the residual) ----
// out[p ldo + q] = beta out + alpha sum_{n < N, f < F} w[n ldw + f] X[p sp + q sq + n sn + f]:
// one wave per output. The (n, f) range is walked as segments of 64 VEC contiguous elements (one
// load per lane, VEC = 2 for 16-byte loads); WSUM_U segments are loaded before any is summed, so
// each wave keeps WSUM_U independent loads of X in flight. Lanes sum in a fixed order, then a
// fixed butterfly.
constexpr int WSUM_U = 8;
template <int VEC>
__global__ __launch_bounds__(256) void wsum_kernel(int64_t P, int64_t Q, int N, int F, double alpha,
                                                   const double* __restrict__ w, int64_t ldw,
                                                   const double* __restrict__ X, int64_t sp, int64_t sq, int64_t sn,
                                                   double beta, double* __restrict__ out, int64_t ldo) {
  const int lane = threadIdx.x & 63;
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= P * Q) return;
  const int64_t p = t / Q, q = t - p * Q;
  const double* __restrict__ x = X + p * sp + q * sq;
  constexpr int SEG = 64 * VEC;
  const int fc = (F + SEG - 1) / SEG;   // segments per n
  const int nseg = N * fc;
  double acc[WSUM_U];
#pragma unroll
  for (int u = 0; u < WSUM_U; ++u) acc[u] = 0.0;
  for (int c0 = 0; c0 < nseg; c0 += WSUM_U) {
    double xv[WSUM_U][VEC], wv[WSUM_U][VEC];
#pragma unroll
    for (int u = 0; u < WSUM_U; ++u) {
      const int c = c0 + u;
      const int n = c / fc;
      const int f = (c - n * fc) * SEG + lane * VEC;
      if (c < nseg && f < F) {
        if (VEC == 2) {
          const double2 a = *(const double2*)(x + n * sn + f);
          const double2 b = *(const double2*)(w + n * ldw + f);
          xv[u][0] = a.x; xv[u][VEC - 1] = a.y;
          wv[u][0] = b.x; wv[u][VEC - 1] = b.y;
        } else {
          xv[u][0] = x[n * sn + f];
          wv[u][0] = w[n * ldw + f];
        }
      } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) xv[u][k] = wv[u][k] = 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < WSUM_U; ++u)
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc[u] = fma(wv[u][k], xv[u][k], acc[u]);
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < WSUM_U; ++u) s += acc[u];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    double* d = out + p * ldo + q;
    *d = (beta == 0.0 ? 0.0 : beta * *d) + alpha * s;
  }
}

// partial[rc][c] = sum_{r in row chunk rc} w[r] X[r ldx + c]
__global__ __launch_bounds__(256) void colsum_partial_kernel(int64_t R, int64_t C, int64_t rchunk,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ X, int64_t ldx,
                                                             double* __restrict__ partial) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, rc = blockIdx.y;
  if (c >= C) return;
  const int64_t r0 = rc * rchunk, r1 = r0 + rchunk < R ? r0 + rchunk : R;
  double a0 = 0.0, a1 = 0.0;
  int64_t r = r0;
  for (; r + 1 < r1; r += 2) {
    a0 = fma(w[r], X[r * ldx + c], a0);
    a1 = fma(w[r + 1], X[(r + 1) * ldx + c], a1);
  }
  if (r < r1) a0 = fma(w[r], X[r * ldx + c], a0);
  partial[rc * C + c] = a0 + a1;
}

// out[c] = beta out[c] + alpha sum_rc partial[rc][c], rc in order
__global__ __launch_bounds__(256) void colsum_finish_kernel(int64_t C, int nrc, const double* __restrict__ partial,
                                                            double alpha, double beta, double* __restrict__ out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= C) return;
  double acc = 0.0;
  for (int rc = 0; rc < nrc; ++rc) acc += partial[rc * C + c];
  out[c] = (beta == 0.0 ? 0.0 : beta * out[c]) + alpha * acc;
}

void wsum(int64_t P, int64_t Q, int64_t N, int64_t F, double alpha, const double* w, int64_t ldw, const double* X,
          int64_t sp, int64_t sq, int64_t sn, double beta, double* out, int64_t ldo, hipStream_t s) {
  if (P * Q == 0) return;
  CCX_CHECK(N < (1ll << 31) / std::max<int64_t>(1, F) && ceil_div(P * Q, 4) < (1ll << 31), ERR_ARG,
            "wsum: extents too large");
  const int64_t blocks = ceil_div(P * Q, 4);   // 4 waves per block, one output per wave
  // 16-byte loads need even strides, an even F and 16-byte aligned bases
  const bool vec2 = F % 2 == 0 && ldw % 2 == 0 && sp % 2 == 0 && sq % 2 == 0 && sn % 2 == 0 &&
                    ((uintptr_t)w & 15) == 0 && ((uintptr_t)X & 15) == 0;
  if (vec2)
    wsum_kernel<2><<<(unsigned)blocks, 256, 0, s>>>(P, Q, (int)N, (int)F, alpha, w, ldw, X, sp, sq, sn, beta, out, ldo);
  else
    wsum_kernel<1><<<(unsigned)blocks, 256, 0, s>>>(P, Q, (int)N, (int)F, alpha, w, ldw, X, sp, sq, sn, beta, out, ldo);
  CCX_HIP(hipGetLastError());
}

int colsum_chunks(int64_t R) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(R, 128))); }

void colsum(int64_t R, int64_t C, double alpha, const double* w, const double* X, int64_t ldx, double beta,
            double* out, double* partial, hipStream_t s) {
  if (C == 0) return;
  const int nrc = colsum_chunks(R);
  const int64_t rchunk = ceil_div(std::max<int64_t>(R, 1), nrc);
  CCX_CHECK(ceil_div(C, 256) < (1ll << 31), ERR_ARG, "column sum too wide");
  colsum_partial_kernel<<<dim3((unsigned)ceil_div(C, 256), (unsigned)nrc), 256, 0, s>>>(R, C, rchunk, w, X, ldx,
                                                                                       partial);
  CCX_HIP(hipGetLastError());
  colsum_finish_kernel<<<(unsigned)ceil_div(C, 256), 256, 0, s>>>(C, nrc, partial, alpha, beta, out);
  CCX_HIP(hipGetLastError());
}

// explicit instantiations
template void ring_prep<double>(int64_t, int64_t, const double*, const double*, double*, double*, double*,
                                double*, double*, hipStream_t);
template void ring_prep<float>(int64_t, int64_t, const double*, const double*, float*, float*, float*,
                               float*, double*, hipStream_t);
template void tau_pack<double>(int64_t, int64_t, const double*, const double*, double*, double*, hipStream_t);
template void tau_pack<float>(int64_t, int64_t, const double*, const double*, float*, float*, hipStream_t);
template void ladder_pack<double>(int64_t, int64_t, int64_t, int64_t, VVVVSrc, const double*, double*,
                                  double*, hipStream_t);
template void ladder_pack<float>(int64_t, int64_t, int64_t, int64_t, VVVVSrc, const double*, float*,
                                 float*, hipStream_t);
template void zt_add<double>(int64_t, int64_t, const double*, const double*, double*, hipStream_t);
template void zt_add<float>(int64_t, int64_t, const float*, const float*, double*, hipStream_t);
template void assemble_r2<double>(int64_t, int64_t, const double*, const double*, const double*, const double*,
                                  const double*, const double*, const double*, const float*, const float*,
                                  const double*, const double*, const double*, double*, hipStream_t);
template void assemble_r2<float>(int64_t, int64_t, const double*, const float*, const float*, const float*,
                                 const float*, const double*, const double*, const float*, const float*,
                                 const double*, const double*, const double*, double*, hipStream_t);
template void cast<double, float>(int64_t, const double*, float*, hipStream_t);
template void cast<float, double>(int64_t, const float*, double*, hipStream_t);

}  // namespace ccx

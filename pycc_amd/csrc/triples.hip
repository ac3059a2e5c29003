// (T) correction of closed-shell CCSD(T) (pycc/cctriples.py:177-240, t_tjl): FP64, except in the
// MIXED_F16X3 mode, whose slot GEMMs are FP16x3 with FP32 output (triples_energy_x3 below).
//
// For every occupied triple i >= j >= k the connected numerator W[a][b][c] (cctriples.py:27-72,
// t3c_ijk without the denominator) comes from GEMMs against A_p = [Covvv[p] | t2[p]^T] (built once
// per call): the six slot products, each holding two of the twelve t3c_ijk terms, pair up along K
// into three GEMMs per triple (Y1, Y2, Y3 below), batched over up to T3_KC values of k for fixed
// (i, j). One kernel per chunk then forms W from the three blocks, orbit by orbit of 8^3 tiles, in
// LDS (W is never stored), adds the disconnected part (t3d_ijk, cctriples.py:108-146:
// <ij|ab> t1_kc + t2_ijab f_kc and their permutations) to form V and reduces
//
//   E_ijk = sum_abc V_abc (4 W_abc + W_bca + W_cab - 2 W_acb - 2 W_bac - 2 W_cba) / D_ijk^abc,
//   E(T)  = sum_{i>=j>=k} (2 - d_ij - d_ik - d_jk) E_ijk,
//
// which is t_tjl's sum over a >= b >= c of (Y - 2Z)(W_abc + W_bca + W_cab) + (Z - 2Y)(W_acb +
// W_bac + W_cba) + 3X with its 1/(1 + d_ab + d_ac + d_bc) weights, rewritten as a full sum over
// a, b, c (the bracket is symmetric under permutations of abc, so each unordered abc appears
// 6 / (1 + d_ab + d_ac + d_bc) times). The per-triple partial sums are reduced in a fixed order, so
// the result is deterministic.
#include "ccsd.h"
#include "contract.h"
#include "split16.h"

#include <algorithm>

namespace ccx {

namespace {

__device__ __forceinline__ double t3_block_sum(double x) {   // blockDim.x <= 1024, a power of two
  __shared__ double sh[1024];
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

__global__ void t3_zero_kernel(double* e) {
  if (threadIdx.x == 0) e[0] = 0.0;
}

// A[p][(x, y)][kk] = Covvv[p][x][y][kk] for kk < v, t2[p][kk - v][x][y] for v <= kk < v + o, 0 above:
// the two t3c_ijk products that share p as one GEMM operand with K = v + o (padded to Kp).
__global__ void t3_operand_a_kernel(int64_t o, int64_t v, int64_t Kp, int64_t nrow_p, const double* __restrict__ Covvv,
                                    const double* __restrict__ t2, double* __restrict__ A) {
  const int64_t total = nrow_p * v * v * Kp;   // rows p < nrow_p
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = e % Kp, row = e / Kp;   // row = (p, x, y)
    const int64_t xy = row % (v * v), p = row / (v * v);
    double val = 0.0;
    if (kk < v) val = Covvv[row * v + kk];
    else if (kk < v + o) val = t2[(p * o + (kk - v)) * v * v + xy];
    A[e] = val;
  }
}

// Device buffer owned by one call: freed on every exit, including an exception thrown after it was
// allocated (hipFree waits for the work queued on the device before it releases the memory).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(int64_t n) { alloc(n); }
  void alloc(int64_t n) {
    CCX_HIP(hipMalloc((void**)&p, (size_t)std::max<int64_t>(n, 1) * sizeof(T)));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// k values per chunk of a fixed (i, j)
constexpr int T3_KC = 8;

// ---- slot products paired along K ----------------------------------------------------------------
// The slot products X_pqr[x][y][z] = sum_kk A_p[(x, y)][kk] B_(q,r)[z][kk] hold the twelve t3c_ijk
// terms (cctriples.py:50-62), with B_(q,r)[z][kk] = t2[r][q][z][kk] for kk < v and -(m q|r z) =
// -<m z|q r> (m = kk - v) below v + o, and W[a][b][c] = X_ijk[a,b,c] + X_ikj[a,c,b] + X_jik[b,a,c] +
// X_jki[b,c,a] + X_kij[c,a,b] + X_kji[c,b,a]. With A_p^s the rows of A_p with (x, y) swapped, they
// pair up along K into three GEMMs (K = 2 Kp):
//   Y1[a][b][c] = X_ijk[a][b][c] + X_jik[b][a][c] = [A_i | A_j^s](a,b) . [B_(j,k) | B_(i,k)](c)
//   Y2[a][c][b] = X_ikj[a][c][b] + X_kij[c][a][b] = [A_i | A_k^s](a,c) . [B_(k,j) | B_(i,j)](b)
//   Y3[b][c][a] = X_jki[b][c][a] + X_kji[c][b][a] = [A_j | A_k^s](b,c) . [B_(k,i) | B_(j,i)](a)
// so that W[a][b][c] = Y1[a][b][c] + Y2[a][c][b] + Y3[b][c][a]: FP64 GEMMs (plain_gemm_kcat) in
// every mode but MIXED_F16X3, whose GEMMs are FP16x3 with FP32 output, in 8 x 8 (y, z) blocks
// when v is a multiple of 8 (x3_gemm_kcat with ytile).
//
// B rows of the chunk (i, j, k0 .. k0 + nk): block 3 e + s (s = 0, 1, 2 for Y1, Y2, Y3 of k = k0 + e)
// holds, for z < v, [B_(q1,r1)[z] | B_(q2,r2)[z]] (2 Kp wide) with the pairs named above.
__global__ void t3_pack_kcat_kernel(int64_t o, int64_t v, int64_t Kp, int i, int j, int k0, int nk,
                                    const double* __restrict__ t2, const double* __restrict__ Cooov,
                                    double* __restrict__ B) {
  const int64_t L = 2 * Kp, total = 3 * (int64_t)nk * v * L;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk2 = e % L, row = e / L, blk = row / v, z = row - blk * v;
    const int e3 = (int)(blk / 3), sl = (int)(blk - 3 * e3), k = k0 + e3;
    const bool h = kk2 >= Kp;
    const int64_t kk = h ? kk2 - Kp : kk2;
    int q, r;
    if (sl == 0) { q = h ? i : j; r = k; }        // (j, k) | (i, k)
    else if (sl == 1) { q = h ? i : k; r = j; }   // (k, j) | (i, j)
    else { q = h ? j : k; r = i; }                // (k, i) | (j, i)
    double val = 0.0;
    if (kk < v) val = t2[(((int64_t)r * o + q) * v + z) * v + kk];
    else if (kk < v + o) val = -Cooov[(((kk - v) * o + q) * o + r) * v + z];
    B[e] = val;
  }
}

// MIXED_F16X3: the same B rows split straight into interleaved FP16 hi/lo (hi of k-block q of a row
// at 64 q, lo at 64 q + 32; rows of 2 L halves) with the scale sc[0]; work-group (x, row) takes
// 256 consecutive kk2 of the row.
__global__ __launch_bounds__(256) void t3_pack_split_kernel(int64_t o, int64_t v, int64_t Kp, int i, int j, int k0,
                                                           const double* __restrict__ t2,
                                                           const double* __restrict__ Cooov,
                                                           const float* __restrict__ sc, _Float16* __restrict__ Bhl) {
  const int64_t L = 2 * Kp, row = blockIdx.y, blk = row / v, z = row - blk * v;
  const int e3 = (int)(blk / 3), sl = (int)(blk - 3 * e3), k = k0 + e3;
  const int64_t kk2 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (kk2 >= L) return;
  const bool h = kk2 >= Kp;
  const int64_t kk = h ? kk2 - Kp : kk2;
  int q, r;
  if (sl == 0) { q = h ? i : j; r = k; }        // (j, k) | (i, k)
  else if (sl == 1) { q = h ? i : k; r = j; }   // (k, j) | (i, j)
  else { q = h ? j : k; r = i; }                // (k, i) | (j, i)
  double val = 0.0;
  if (kk < v) val = t2[(((int64_t)r * o + q) * v + z) * v + kk];
  else if (kk < v + o) val = -Cooov[(((kk - v) * o + q) * o + r) * v + z];
  const float x = (float)(val * (double)sc[0]);
  const _Float16 hv = (_Float16)x;
  const int64_t at = row * 2 * L + ((kk2 >> 5) << 6) + (kk2 & 31);
  Bhl[at] = hv;
  Bhl[at + 32] = (_Float16)(x - (float)hv);
}

// One work-group per (orbit {A >= B >= C} of 8^3 tiles, triple k0 + blockIdx.y): W of the orbit's
// distinct tiles (one, three or six permutations of (A, B, C)) from Y1, Y2, Y3 into LDS (FP32),
// then the energy sum over those tiles, every permuted W read from LDS. HBM
// traffic is Y read once; W is never written. A work-group needs 36 KiB of LDS, so that four are
// resident on a CU and the loads of one overlap the sums of the others.
constexpr int OE = 8;                // tile edge
constexpr int OTW = OE * OE * OE;    // words per LDS tile
constexpr int OT_THREADS = 256;
// LDS holds a W tile (x, y, z) at a word address that is linear over GF(2) in the nine index bits:
// the XOR, over the set bits, of x bit k: 8, 6, 89; y bit k: 11, 34, 155; z bit k: 1, 52, 260
// (k = 0..2), a bijection onto 0..511. Every LDS access of W here is made by waves whose 64 lanes
// run over two whole coordinates, or over x bits 0-1 (lane bits 4-5), one whole coordinate and
// bit 2 of the third; for each such pattern the 64 lanes hit 64 distinct banks (address mod 64)
// and each half wave 32 distinct banks mod 32, so that no access has a bank conflict.
__device__ __forceinline__ constexpr int t3_ax(int x) {   // words
  return ((x & 1) * 8) ^ (((x >> 1) & 1) * 6) ^ (((x >> 2) & 1) * 89);
}
__device__ __forceinline__ constexpr int t3_ay(int y) {
  return ((y & 1) * 11) ^ (((y >> 1) & 1) * 34) ^ (((y >> 2) & 1) * 155);
}
__device__ __forceinline__ constexpr int t3_az(int z) {
  return ((z & 1) * 1) ^ (((z >> 1) & 1) * 52) ^ (((z >> 2) & 1) * 260);
}
// slot x holds the tile at the permutation (A, B, C), (A, C, B), (B, A, C), (B, C, A), (C, A, B),
// (C, B, A) of the orbit.

// which of (A, B, C) sits at position n of slot x's permutation: 2 bits at 6 x + 2 n of 0x192261624
// (the slots' permutations 012, 021, 102, 120, 201, 210), so that a run-time slot needs no array
__device__ __forceinline__ constexpr int t3_slot_src(int x, int n) {
  return (int)((0x192261624ull >> (6 * x + 2 * n)) & 3);
}
__device__ __forceinline__ int t3_slot_tile(int x, int n, int A, int B, int C) {
  const int src = t3_slot_src(x, n);
  return src == 0 ? A : (src == 1 ? B : C);
}

// 1/x: the hardware estimate (relative error below 2^-23) and NS Newton steps (one: below 2^-45)
template <int NS>
__device__ __forceinline__ double t3_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int n = 0; n < NS; ++n) r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// TY: the type of Y and of W in LDS (float in MIXED_F16X3, double in the FP64 path). FULL (float
// only): Y in the tiled layout of x3_gemm_kcat's ytile (v a multiple of 8, every tile whole); else
// Y is [x][y][z] with rows of ldy elements and the tiles at the edge are partial.
// The disconnected blocks come from PB (t3_pair_blocks_kernel) and t1, F. One work-group per (orbit,
// triple) pair t = orbit + norb e of the 1-D grid, taken in an order that gives each XCD a
// contiguous range of t (consecutive orbits share tiles, so their blocks meet in that XCD's L2).
template <typename TY, bool FULL>
__global__ __launch_bounds__(OT_THREADS) void t3_orbit_energy_kernel(
    int64_t o, int64_t v, int64_t ldy, int i, int j, int k0, int nk, const TY* __restrict__ Y1,
    const TY* __restrict__ Y2, const TY* __restrict__ Y3, const int* __restrict__ orbits, int64_t norb,
    const double* __restrict__ F, int64_t nmo, const double* __restrict__ t1, const double* __restrict__ PB,
    double* __restrict__ partials) {
  static_assert(!FULL || sizeof(TY) == 4, "the tiled Y layout is FP32");
  // for the tile (P0, P1, P2) of slot x, the slots of (P1, P2, P0), (P2, P0, P1), (P0, P2, P1),
  // (P1, P0, P2) and (P2, P1, P0)
  constexpr int t3_cyc[6] = {3, 5, 1, 4, 0, 2};
  constexpr int t3_cyc2[6] = {4, 2, 5, 0, 3, 1};
  constexpr int t3_sw23[6] = {1, 0, 3, 2, 5, 4};
  constexpr int t3_sw12[6] = {2, 4, 0, 5, 1, 3};
  constexpr int t3_rev[6] = {5, 3, 4, 1, 2, 0};
  typedef TY f4v __attribute__((ext_vector_type(4)));
  typedef double d2v __attribute__((ext_vector_type(2)));
  __shared__ TY sW[6 * OTW];
  // the disconnected factors of the tiles, in pairs: sP[x][0][a][b] = (<ij|ab>, t2_ijab),
  // sP[x][1][a][c] = (<ik|ac>, t2_ikac), sP[x][2][b][c] = (<jk|bc>, t2_jkbc); sTF[n][s][.] = (t1,
  // f row) of the occupied orbital (i, j, k)[n] over tile (A, B, C)[s] and sE[s][.] = eps over it
  __shared__ d2v sP[6][3][OE][OE];
  __shared__ d2v sTF[3][3][OE];
  __shared__ double sE[3][OE];
  int64_t tw;
  {   // hardware work-group id (round-robin over the 8 XCDs) -> t
    const int64_t n = norb * nk, q = n / 8, r = n % 8, xcd = blockIdx.x % 8, w = blockIdx.x / 8;
    tw = xcd * q + (xcd < r ? xcd : r) + w;
  }
  const int e = (int)(tw / norb), k = k0 + e;
  const int64_t ob = tw - (int64_t)e * norb;
  const int64_t nt = (v + OE - 1) / OE, vv = v * v, ys = vv * ldy;
  const int code = orbits[ob];
  const int A = (int)(code / (nt * nt)), B = (int)(code / nt % nt), C = (int)(code % nt);
  const int T[6][3] = {{A, B, C}, {A, C, B}, {B, A, C}, {B, C, A}, {C, A, B}, {C, B, A}};
  // rep[x]: the first slot with slot x's tile (A >= B >= C: only A == B or B == C repeat tiles);
  // bit x of kept: rep[x] == x
  int rep[6];
  {
    const bool ab = A == B, bc = B == C;
    rep[0] = 0;
    rep[1] = bc ? 0 : 1;
    rep[2] = ab ? 0 : 2;
    rep[3] = ab ? 1 : 3;
    rep[4] = ab ? 4 : (bc ? 2 : 4);
    rep[5] = ab ? 4 : (bc ? 3 : 5);
    if (ab && bc) rep[1] = rep[2] = rep[3] = rep[4] = rep[5] = 0;
  }
  int kept = 0;
#pragma unroll
  for (int x = 0; x < 6; ++x) kept |= (rep[x] == x) << x;
  const TY* y1 = Y1 + e * ys;
  const TY* y2 = Y2 + e * ys;
  const TY* y3 = Y3 + e * ys;
  // Box loads: the threads of half hf = t / 128 take the boxes of the slots 2 p + hf; thread (x, z,
  // yg) = (t / 16 % 8, t / 2 % 8, t % 2) the four elements y = 4 yg + u, u = 0..3, of the row x and
  // column z (in the tiled layout the 16 bytes at 32 z + 16 yg of the box's 256-byte run for x).
  const int hf = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);   // wave-uniform
  const int bx = (threadIdx.x >> 4) & 7, bz = (threadIdx.x >> 1) & 7, yg = threadIdx.x & 1;
  const int lbox = bx * (int)vv + 8 * bz + 4 * yg;   // the lane's offset in a whole box (v^2 < 2^28)
  auto box = [&](const TY* Y, int gx, int gy, int gz) -> f4v {   // tile origins in tiles
    if constexpr (FULL) {
      return *(const f4v*)(Y + ((OE * (int64_t)gx * v + OE * gy) * v + 64 * gz) + lbox);
    } else {
      const int64_t x = OE * gx + bx, z = OE * gz + bz;
      f4v val = f4v{0, 0, 0, 0};
      if (x < v && z < v) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t y = OE * gy + 4 * yg + u;
          if (y < v) val[u] = Y[(x * v + y) * ldy + z];
        }
      }
      return val;
    }
  };
  f4v ra[3], rb[3], rc[3];   // slot 2 p + hf
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int x = 2 * p + hf;
    if ((kept >> x) & 1) {   // wave-uniform
      const int P0 = t3_slot_tile(x, 0, A, B, C), P1 = t3_slot_tile(x, 1, A, B, C), P2 = t3_slot_tile(x, 2, A, B, C);
      ra[p] = box(y1, P0, P1, P2);   // Y1[a][b][c]
      rb[p] = box(y2, P0, P2, P1);   // Y2[a][c][b]
      rc[p] = box(y3, P1, P2, P0);   // Y3[b][c][a]
    }
  }
  // the disconnected factors: element t + OT_THREADS r of the 6 x 6 x 64 block elements (slot x =
  // el / 384, block d = el / 64 % 6: <ij|ab>, t2_ijab, <ik|ac>, t2_ikac, <jk|bc>, t2_jkbc) and of the
  // 6 x 9 x 8 vector elements (slot x, q = 3 n + u: u = 0 t1, 1 f row, 2 eps; n = a, b, c)
  // wave w of round rr takes block bi = 4 rr + w < 18: pair d = bi % 3 ((i, j), (i, k), (j, k)) of
  // slot x = bi / 3, whose 1 KiB in PB lane l reads as the element (l / 8, l % 8). Every wave loads
  // (bi clamped to 17; slots that repeat a tile load blocks that are never read), so that no load
  // sits in a branch.
  constexpr int NB = (18 + 3) / 4;
  const int lane = threadIdx.x & 63, w4 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  d2v pd[NB];
  double pv = 0.0;
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) {
    const int bi = min(4 * rr + w4, 17), x = bi / 3, d = bi - 3 * x;   // wave-uniform
    const int pp = d == 0 ? 0 : (d == 1 ? 1 + e : 1 + nk + e);
    const int rt = t3_slot_tile(x, d < 2 ? 0 : 1, A, B, C), ct = t3_slot_tile(x, d == 0 ? 1 : 2, A, B, C);
    pd[rr] = *(const d2v*)(PB + 2 * ((((int64_t)pp * nt + rt) * nt + ct) * 64 + lane));
  }
  // the vectors: row R = t / 8 < 21 (t1 of the occupied (i, j, k)[R / 3] over tile (A, B, C)[R % 3]
  // for R < 9, then the f rows for R < 18, then eps over tile R - 18), element t % 8
  {
    const int R = threadIdx.x >> 3;
    if (R < 21) {
      const int n = (R < 18 ? R % 9 : 0) / 3, sidx = R < 18 ? R % 3 : R - 18;
      const int64_t g = OE * (int64_t)(sidx == 0 ? A : (sidx == 1 ? B : C)) + (threadIdx.x & 7);
      if (FULL || g < v) {
        const int occ = n == 0 ? i : (n == 1 ? j : k);
        pv = R < 9 ? t1[occ * v + g] : (R < 18 ? F[occ * nmo + o + g] : F[(o + g) * nmo + o + g]);
      }
    }
  }
  // W = Y1[a][b][c] + Y2[a][c][b] + Y3[b][c][a] in three passes; thread (bx, bz, yg) holds Y1 (a, b,
  // c) = (bx, 4 yg + u, bz), Y2 (a, c, b) = (bx, 4 yg + u, bz), Y3 (b, c, a) = (bx, 4 yg + u, bz)
  const int a1 = t3_ax(bx) ^ t3_ay(4 * yg) ^ t3_az(bz);   // W (bx, 4 yg + u, bz)
  const int a2 = t3_ax(bx) ^ t3_ay(bz) ^ t3_az(4 * yg);   // W (bx, bz, 4 yg + u)
  const int a3 = t3_ax(bz) ^ t3_ay(bx) ^ t3_az(4 * yg);   // W (bz, bx, 4 yg + u)
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int x = 2 * p + hf;
    if ((kept >> x) & 1) {
      TY* w = sW + x * OTW;
#pragma unroll
      for (int u = 0; u < 4; ++u) w[a1 ^ t3_ay(u)] = ra[p][u];
    }
  }
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) {
    const int bi = 4 * rr + w4, x = bi / 3, d = bi - 3 * x;
    if (bi < 18) sP[x][d][lane >> 3][lane & 7] = pd[rr];
  }
  {
    const int R = threadIdx.x >> 3;
    if (R < 18) ((double*)&sTF[(R % 9) / 3][R % 3][threadIdx.x & 7])[R < 9 ? 0 : 1] = pv;
    else if (R < 21) sE[R - 18][threadIdx.x & 7] = pv;
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int x = 2 * p + hf;
    if ((kept >> x) & 1) {
      TY* w = sW + x * OTW;
#pragma unroll
      for (int u = 0; u < 4; ++u) w[a2 ^ t3_az(u)] += rb[p][u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int x = 2 * p + hf;
    if ((kept >> x) & 1) {
      TY* w = sW + x * OTW;
#pragma unroll
      for (int u = 0; u < 4; ++u) w[a3 ^ t3_az(u)] += rc[p][u];
    }
  }
  __syncthreads();

  // Energy: thread t takes the column (b, c) = (t / 8 % 8, t % 8) of each kept tile and a = t / 64 +
  // 4 ii, ii = 0, 1.
  const int bl = (threadIdx.x >> 3) & 7, cl = threadIdx.x & 7, al0 = threadIdx.x >> 6;
  // word addresses of W at (a, b, c) permuted, for ii = 0 (the XOR of t3_a*(4 ii) adds 4 ii)
  const int e0 = t3_ax(al0) ^ t3_ay(bl) ^ t3_az(cl);   // (a, b, c)
  const int e1 = t3_ax(bl) ^ t3_ay(cl) ^ t3_az(al0);   // (b, c, a)
  const int e2 = t3_ax(cl) ^ t3_ay(al0) ^ t3_az(bl);   // (c, a, b)
  const int e3 = t3_ax(al0) ^ t3_ay(cl) ^ t3_az(bl);   // (a, c, b)
  const int e4 = t3_ax(bl) ^ t3_ay(al0) ^ t3_az(cl);   // (b, a, c)
  const int e5 = t3_ax(cl) ^ t3_ay(bl) ^ t3_az(al0);   // (c, b, a)
  auto rd = [&](int off) { return (double)*(const TY*)((const char*)sW + off); };
  const double dijk = F[i * nmo + i] + F[j * nmo + j] + F[k * nmo + k];
  double s = 0.0;
#pragma unroll
  for (int x = 0; x < 6; ++x) {
    if (rep[x] != x) continue;   // block-uniform
    const int64_t a0 = OE * (int64_t)T[x][0], b = OE * (int64_t)T[x][1] + bl, c = OE * (int64_t)T[x][2] + cl;
    if (FULL || (b < v && c < v)) {
      // byte offsets in sW of the six permuted reads: slot (the bits above the tile) | element
      // address * sizeof(TY)
      constexpr int SZ = (int)sizeof(TY);
      const int q0 = (x * OTW + e0) * SZ, q1 = (rep[t3_cyc[x]] * OTW + e1) * SZ;
      const int q2 = (rep[t3_cyc2[x]] * OTW + e2) * SZ, q3 = (rep[t3_sw23[x]] * OTW + e3) * SZ;
      const int q4 = (rep[t3_sw12[x]] * OTW + e4) * SZ, q5 = (rep[t3_rev[x]] * OTW + e5) * SZ;
      const int s0 = t3_slot_src(x, 0), s1 = t3_slot_src(x, 1), s2 = t3_slot_src(x, 2);
      const double dbc = dijk - sE[s1][bl] - sE[s2][cl];
      const d2v pbc = sP[x][2][bl][cl], tfb = sTF[1][s1][bl], tfc = sTF[2][s2][cl];
#pragma unroll
      for (int ii = 0; ii < 2; ++ii) {
        const int al = al0 + 4 * ii;
        if (!FULL && a0 + al >= v) break;   // wave-uniform
        // 4 W_abc + W_bca + W_cab - 2 (W_acb + W_bac + W_cba)
        const double w = rd(q0 ^ SZ * t3_ax(4 * ii));
        const double z = fma(-2.0, rd(q3 ^ SZ * t3_ax(4 * ii)) + rd(q4 ^ SZ * t3_ay(4 * ii)) + rd(q5 ^ SZ * t3_az(4 * ii)),
                             fma(4.0, w, rd(q1 ^ SZ * t3_az(4 * ii)) + rd(q2 ^ SZ * t3_ay(4 * ii))));
        const d2v pab = sP[x][0][al][bl], pac = sP[x][1][al][cl], tfa = sTF[0][s0][al];
        const double vabc = w + pab.x * tfc.x + pac.x * tfb.x + pbc.x * tfa.x + pab.y * tfc.y + pac.y * tfb.y +
                            pbc.y * tfa.y;
        // FP32 W: one Newton step (relative error below 2^-45); FP64: two (to the last bits)
        s += vabc * z * t3_rcp<sizeof(TY) == 8 ? 2 : 1>(dbc - sE[s0][al]);
      }
    }
  }
  // work-group sum: within each wave, then over the four waves
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ double sw4[OT_THREADS / 64];
  if ((threadIdx.x & 63) == 0) sw4[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[tw] = (sw4[0] + sw4[1]) + (sw4[2] + sw4[3]);
}

// PB[pp][rt][ct][r][c] = (<p q|a b>, t2_pqab) at a = 8 rt + r, b = 8 ct + c (zero past v), for the pairs
// pp = 0: (i, j), 1 + e: (i, k0 + e), 1 + nk + e: (j, k0 + e), e < nk: the disconnected factors of a
// chunk as contiguous 8 x 8 blocks of (integral, amplitude) pairs.
__global__ void t3_pair_blocks_kernel(int64_t o, int64_t v, int64_t nt, int i, int j, int k0, int nk,
                                      const double* __restrict__ Ioovv, const double* __restrict__ t2,
                                      double* __restrict__ PB) {
  const int64_t per = nt * nt * 64, total = (1 + 2 * (int64_t)nk) * per;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
    const int pp = (int)(x / per);
    const int64_t rem = x - pp * per, blk = rem >> 6, rt = blk / nt, ct = blk - rt * nt;
    const int64_t a = 8 * rt + ((rem >> 3) & 7), b = 8 * ct + (rem & 7);
    const int p = pp <= nk ? i : j, q = pp == 0 ? j : k0 + (pp <= nk ? pp - 1 : pp - 1 - nk);
    double iv = 0.0, tv = 0.0;
    if (a < v && b < v) {
      const int64_t off = ((int64_t)p * o + q) * v * v + a * v + b;
      iv = Ioovv[off];
      tv = t2[off];
    }
    PB[2 * x] = iv;
    PB[2 * x + 1] = tv;
  }
}

// e[0] += sum over the chunk's triples (i, j, k0 + x) of weight * sum(partials[x][0 .. norb)), in a
// fixed order: work-group g of T3_NSUM sums the weighted partials g, g + T3_NSUM, ... into sums[g],
// then one work-group adds sums[0 .. T3_NSUM) to e[0].
constexpr int T3_NSUM = 256;
__global__ __launch_bounds__(256) void t3_chunk_sums_kernel(const double* __restrict__ partials, int64_t norb, int nk,
                                                            int i, int j, int k0, double* __restrict__ sums) {
  const int64_t n = (int64_t)nk * norb;
  double s = 0.0;
  for (int64_t t = blockIdx.x + (int64_t)T3_NSUM * threadIdx.x; t < n; t += (int64_t)T3_NSUM * blockDim.x) {
    const int k = k0 + (int)(t / norb);
    s += (2.0 - (double)((i == j) + (i == k) + (j == k))) * partials[t];   // cctriples.py:237
  }
  s = t3_block_sum(s);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}
__global__ __launch_bounds__(T3_NSUM) void t3_accumulate_chunk_kernel(const double* __restrict__ sums, double* e) {
  const double s = t3_block_sum(sums[threadIdx.x]);
  if (threadIdx.x == 0) e[0] += s;
}

// orbits {A >= B >= C} of the nt^3 tiles as (A nt + B) nt + C
std::vector<int> t3_orbits(int64_t nt) {
  std::vector<int> out;
  for (int64_t A = 0; A < nt; ++A)
    for (int64_t B = 0; B <= A; ++B)
      for (int64_t C = 0; C <= B; ++C) out.push_back((int)((A * nt + B) * nt + C));
  return out;
}

}  // namespace

// MIXED_F16X3: for fixed (i, j) the k values run in chunks of T3_KC; per chunk the K-concatenated
// B rows (FP64, then split with one scale), the batched Y1, Y2, Y3 GEMMs and one orbit-energy launch.
static void triples_energy_x3(Ctx& c, const double* F, const double* t1, const double* t2, int i0, int i1, bool accumulate,
                              double* e_dev, hipStream_t s) {
  const int64_t o = c.o, v = c.v, nmo = c.nmo, vv = v * v;
  const int64_t nt = ceil_div(v, OE), KC = T3_KC;
  const int64_t Kp = (v + o + 31) / 32 * 32, L = 2 * Kp;
  // Y in 8 x 8 (y, z) blocks when v is a multiple of 8, else rows padded to 4 floats
  const bool tiled = v % 8 == 0;
  const int64_t ldy = tiled ? v : (v + 3) / 4 * 4;
  const int64_t nA = (int64_t)i1 * vv * Kp, nBr = 3 * KC * v, np = x3_partials();
  CCX_CHECK(vv * ldy * KC < (int64_t(1) << 40) && ceil_div(vv, 256) < (int64_t(1) << 24) && nBr < 65536, ERR_ARG,
            "too many virtual orbitals for the (T) scratch");
  DevBuf<f16> Ahl(nA * 2);   // A split, [i1 v^2][2 Kp]
  DevBuf<float> sc(16 + 3 * np);   // sc[0] = scale of A, sc[1] = scale of B, then partials
  // The entries of A over every p are exactly those of Covvv and of t2 (plus zero padding), so its
  // scale is that of their union: the same for every range of i, whatever rows are built. Those
  // of every chunk's B rows are entries of t2 and of -Cooov: B takes the scale of that union.
  x3_absmax<double>(c.Covvv, o * vv * v, sc.p + 16, s);
  x3_absmax<double>(t2, o * o * vv, sc.p + 16 + np, s);
  x3_absmax<double>(c.Cooov, o * o * o * v, sc.p + 16 + 2 * np, s);
  x3_scale_union(sc.p + 16, (int)(2 * np), sc.p, s);
  x3_scale_union(sc.p + 16 + np, (int)(2 * np), sc.p + 1, s);
  if (nA > 0) {   // A split; its FP64 form lives only until then
    DevBuf<double> A64(nA);
    t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, i1, c.Covvv, t2,
                                                                                         A64.p);
    CCX_HIP(hipGetLastError());
    x3_split_rows<double>(A64.p, (int64_t)i1 * vv, Kp, Kp, sc.p, Ahl.p, Ahl.p + 32, Kp, s, true);
    CCX_HIP(hipStreamSynchronize(s));
  }
  const std::vector<int> orb = t3_orbits(nt);
  const int64_t norb = (int64_t)orb.size();
  DevBuf<int> dorb(norb);
  CCX_HIP(hipMemcpy(dorb.p, orb.data(), (size_t)norb * sizeof(int), hipMemcpyHostToDevice));
  DevBuf<f16> Bhl(nBr * 2 * L);      // the K-concatenated B rows of a chunk, split: rows of 2 L halves
  const int64_t ys = vv * ldy;          // one triple's Y block (see t3_orbit_energy_kernel)
  DevBuf<float> Y(3 * KC * ys);        // Y1, Y2, Y3 of the chunk's triples
  DevBuf<double> part(KC * norb + T3_NSUM);   // per (orbit, triple), then the chunk's T3_NSUM sums
  DevBuf<double> PB((1 + 2 * KC) * nt * nt * 64 * 2);   // the chunk's disconnected blocks
  float *Y1 = Y.p, *Y2 = Y1 + KC * ys, *Y3 = Y2 + KC * ys;
  if (!accumulate) {
    t3_zero_kernel<<<1, 64, 0, s>>>(e_dev);
    CCX_HIP(hipGetLastError());
  }
  auto Ap = [&](int p) { return Ahl.p + (int64_t)p * vv * 2 * Kp; };
  for (int i = i0; i < i1; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k0 = 0; k0 <= j; k0 += (int)KC) {
        const int nk = (int)std::min<int64_t>(KC, j + 1 - k0);
        const int64_t nrows = 3 * (int64_t)nk * v;
        t3_pack_split_kernel<<<dim3((unsigned)ceil_div(L, 256), (unsigned)nrows), 256, 0, s>>>(
            o, v, Kp, i, j, k0, t2, c.Cooov, sc.p + 1, Bhl.p);
        CCX_HIP(hipGetLastError());
        X3Gemm g;   // C_e[m][n] = sum_kk [A | A2_e swapped][m][kk] B_e[n][kk]
        g.M = vv; g.N = v; g.Kp = 2 * Kp; g.K1 = Kp; g.lda = 2 * Kp; g.ldb = 2 * L; g.ldc = ldy;
        g.kb = 32; g.ska = 64; g.skb = 64; g.sw = v; g.nbatch = nk; g.ytile = tiled ? v : 0;
        g.sa = sc.p; g.sb = sc.p + 1; g.sB = 3 * v * 2 * L; g.sC = ys;
        // Y1 = [A_i | A_j^s] [B_(j,k) | B_(i,k)]
        g.Ah = Ap(i); g.Al = Ap(i) + 32; g.Ah2 = Ap(j); g.sA2 = 0;
        g.Bh = Bhl.p; g.Bl = Bhl.p + 32; g.C = Y1;
        x3_gemm_kcat(g, s);
        // Y2 = [A_i | A_k^s] [B_(k,j) | B_(i,j)]
        g.Ah2 = Ap(k0); g.sA2 = vv * 2 * Kp;
        g.Bh = Bhl.p + v * 2 * L; g.Bl = (const f16*)g.Bh + 32; g.C = Y2;
        x3_gemm_kcat(g, s);
        // Y3 = [A_j | A_k^s] [B_(k,i) | B_(j,i)]
        g.Ah = Ap(j); g.Al = Ap(j) + 32;
        g.Bh = Bhl.p + 2 * v * 2 * L; g.Bl = (const f16*)g.Bh + 32; g.C = Y3;
        x3_gemm_kcat(g, s);
        {
          const int64_t nb = (1 + 2 * (int64_t)nk) * nt * nt * 64;
          t3_pair_blocks_kernel<<<(int)std::min<int64_t>(8192, ceil_div(nb, 256)), 256, 0, s>>>(
              o, v, nt, i, j, k0, nk, c.Ioovv, t2, PB.p);
          CCX_HIP(hipGetLastError());
        }
        if (tiled)
          t3_orbit_energy_kernel<float, true><<<(unsigned)(norb * nk), OT_THREADS, 0, s>>>(
              o, v, ldy, i, j, k0, nk, Y1, Y2, Y3, dorb.p, norb, F, nmo, t1, PB.p, part.p);
        else
          t3_orbit_energy_kernel<float, false><<<(unsigned)(norb * nk), OT_THREADS, 0, s>>>(
              o, v, ldy, i, j, k0, nk, Y1, Y2, Y3, dorb.p, norb, F, nmo, t1, PB.p, part.p);
        CCX_HIP(hipGetLastError());
        t3_chunk_sums_kernel<<<T3_NSUM, 256, 0, s>>>(part.p, norb, nk, i, j, k0, part.p + KC * norb);
        CCX_HIP(hipGetLastError());
        t3_accumulate_chunk_kernel<<<1, T3_NSUM, 0, s>>>(part.p + KC * norb, e_dev);
        CCX_HIP(hipGetLastError());
      }
  CCX_HIP(hipStreamSynchronize(s));
}

void triples_energy(Ctx& c, const double* F, const double* t1, const double* t2, int i0, int i1, bool accumulate,
                    double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "triples called before the integrals were set");
  CCX_CHECK(F && t1 && t2 && e_dev, ERR_ARG, "null argument");
  CCX_CHECK(0 <= i0 && i0 <= i1 && i1 <= c.o, ERR_ARG, "triples: need 0 <= i0 <= i1 <= no");
  const int64_t o = c.o, v = c.v, nmo = c.nmo, vv = v * v, vvv = vv * v;
  CCX_CHECK(c.Covvv && c.Cooov && c.Ioovv, ERR_STATE, "the context holds no FP64 <ov|vv>/<oo|ov> blocks");
  const int64_t nt8 = ceil_div(v, OE);
  CCX_CHECK(nt8 * nt8 * nt8 < (int64_t(1) << 31) && (nt8 + 2) * (nt8 + 1) * nt8 / 6 * T3_KC < (int64_t(1) << 31),
            ERR_ARG, "too many virtual orbitals for the (T) tiles");
  // In MIXED_F16X3 the slot GEMMs are FP16x3 split products (interleaved hi/lo operands, K padded
  // to 32, FP32 accumulation and output, two slot products per GEMM); every other mode runs them
  // in FP64.
  const bool x3 = c.mode == MODE_MIXED_F16X3;
  const int64_t Kp = x3 ? (v + o + 31) / 32 * 32 : (v + o + 15) / 16 * 16;
  const int64_t KC = T3_KC;
  // A: only the rows p < i1 are read by this range (p is one of i >= j >= k)
  const int64_t nA = (int64_t)i1 * vv * Kp;
  if (x3) {
    triples_energy_x3(c, F, t1, t2, i0, i1, accumulate, e_dev, s);
    return;
  }
  // For fixed (i, j) the k values run in chunks of KC. The six slot products pair up along K into
  // the three products Y1, Y2, Y3 of triples_energy_x3 (FP64 here); per chunk they are three
  // batched GEMMs (plain_gemm_kcat, a batch entry per k), then the chunk's disconnected blocks and
  // one orbit-energy launch (W formed in LDS from Y, never stored). Scratch: Y1, Y2, Y3 (KC v^3
  // each), the chunk's B rows (3 KC v x 2 Kp), its disconnected blocks, the orbit partials and A
  // (i1 v^2 x Kp). (every region starts on a 256-byte boundary: the GEMM loads its operands 16
  // bytes at a time)
  auto r32 = [](int64_t n) { return (n + 31) / 32 * 32; };
  const int64_t ot = ceil_div(v, OE);
  const std::vector<int> orb = t3_orbits(ot);
  const int64_t norb = (int64_t)orb.size();
  DevBuf<int> dorb(norb);
  CCX_HIP(hipMemcpy(dorb.p, orb.data(), (size_t)norb * sizeof(int), hipMemcpyHostToDevice));
  const int64_t L = 2 * Kp, nY = r32(KC * vvv), nB = r32(3 * KC * v * L), nP = r32(KC * norb + T3_NSUM);
  const int64_t nPB = r32((1 + 2 * KC) * ot * ot * 64 * 2);
  DevBuf<double> buf(3 * nY + nB + nP + nPB + nA);
  double *Y1 = buf.p, *Y2 = Y1 + nY, *Y3 = Y2 + nY, *B = Y3 + nY, *part = B + nB, *PB = part + nP;
  double* A = PB + nPB;
  if (nA > 0) {
    t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, i1, c.Covvv, t2, A);
    CCX_HIP(hipGetLastError());
  }
  if (!accumulate) {
    t3_zero_kernel<<<1, 64, 0, s>>>(e_dev);
    CCX_HIP(hipGetLastError());
  }
  auto Ap = [&](int p) { return A + (int64_t)p * vv * Kp; };
  for (int i = i0; i < i1; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k0 = 0; k0 <= j; k0 += (int)KC) {
        const int nk = (int)std::min<int64_t>(KC, j + 1 - k0);
        const int64_t nrows = 3 * (int64_t)nk * v;
        t3_pack_kcat_kernel<<<(int)std::min<int64_t>(4096, ceil_div(nrows * L, 256)), 256, 0, s>>>(
            o, v, Kp, i, j, k0, nk, t2, c.Cooov, B);
        CCX_HIP(hipGetLastError());
        PlainGemm<double> g;   // C_e[m][n] = sum_kk [A | A2_e swapped][m][kk] B_e[n][kk]
        g.M = vv; g.N = v; g.K = L; g.K1 = Kp; g.sw = v; g.batch = nk;
        g.lda = Kp; g.ldb = L; g.ldc = v; g.a_kc = true; g.b_nc = false;
        g.b_b = 3 * v * L; g.c_b = vvv;
        g.A = Ap(i); g.A2 = Ap(j); g.a2_b = 0; g.B = B; g.C = Y1;                   // [A_i | A_j^s]
        plain_gemm_kcat(g, s);
        g.A2 = Ap(k0); g.a2_b = vv * Kp; g.B = B + v * L; g.C = Y2;                // [A_i | A_k^s]
        plain_gemm_kcat(g, s);
        g.A = Ap(j); g.B = B + 2 * v * L; g.C = Y3;                                // [A_j | A_k^s]
        plain_gemm_kcat(g, s);
        const int64_t npb = (1 + 2 * (int64_t)nk) * ot * ot * 64;
        t3_pair_blocks_kernel<<<(int)std::min<int64_t>(8192, ceil_div(npb, 256)), 256, 0, s>>>(
            o, v, ot, i, j, k0, nk, c.Ioovv, t2, PB);
        CCX_HIP(hipGetLastError());
        t3_orbit_energy_kernel<double, false><<<(unsigned)(norb * nk), OT_THREADS, 0, s>>>(
            o, v, v, i, j, k0, nk, Y1, Y2, Y3, dorb.p, norb, F, nmo, t1, PB, part);
        CCX_HIP(hipGetLastError());
        t3_chunk_sums_kernel<<<T3_NSUM, 256, 0, s>>>(part, norb, nk, i, j, k0, part + KC * norb);
        CCX_HIP(hipGetLastError());
        t3_accumulate_chunk_kernel<<<1, T3_NSUM, 0, s>>>(part + KC * norb, e_dev);
        CCX_HIP(hipGetLastError());
      }
  CCX_HIP(hipStreamSynchronize(s));
}

namespace {

// t3[x][y][z] /= D for one fixed index triple (p, q, r) and the running (x, y, z) of one n^3 batch:
// D = (f_pp + f_qq + f_rr) - (f_xx + f_yy + f_zz) with running orbitals e0 + x, ... when the fixed
// triple is occupied (the i, j, k batches of cctriples.py:64-70, :139-145), its negative when it is
// virtual (the a, b, c batches, :97-103, :165-171); both sums are added left to right, as there.
// Grid (ceil(n / 256), n, n): z = blockIdx.x 256 + threadIdx.x, y = blockIdx.y, x = blockIdx.z.
__global__ void __launch_bounds__(256) t3_denominator_kernel(const double* __restrict__ F, int64_t nmo, int p,
                                                             int q, int r, int e0, int n, int occ_fixed,
                                                             double* __restrict__ t3) {
  const int z = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, x = blockIdx.z;
  if (z >= n) return;
  const int64_t d1 = nmo + 1;
  const double sf = (F[p * d1] + F[q * d1]) + F[r * d1];
  const double sr = (F[(e0 + x) * d1] + F[(e0 + y) * d1]) + F[(e0 + z) * d1];
  double* e = t3 + ((int64_t)x * n + y) * n + z;
  *e = *e / (occ_fixed ? sf - sr : sr - sf);
}

}  // namespace

void t3_denominator(const double* F, int64_t nmo, int p, int q, int r, int e0, int n, bool occ_fixed, double* t3,
                    hipStream_t s) {
  CCX_CHECK(F != nullptr && nmo >= 1 && n >= 0, ERR_ARG, "t3 denominator: null fock or negative extent");
  CCX_CHECK(p >= 0 && q >= 0 && r >= 0 && p < nmo && q < nmo && r < nmo, ERR_ARG,
            "t3 denominator: fixed orbital outside the Fock matrix");
  CCX_CHECK(e0 >= 0 && (int64_t)e0 + n <= nmo, ERR_ARG, "t3 denominator: running orbitals outside the Fock matrix");
  CCX_CHECK(n <= 65535, ERR_ARG, "t3 denominator: more than 65535 running orbitals");
  if (n == 0) return;
  CCX_CHECK(t3 != nullptr, ERR_ARG, "t3 denominator: null t3");
  const dim3 g((unsigned)ceil_div(n, 256), (unsigned)n, (unsigned)n);
  t3_denominator_kernel<<<g, 256, 0, s>>>(F, nmo, p, q, r, e0, n, occ_fixed ? 1 : 0, t3);
  CCX_HIP(hipGetLastError());
}

}  // namespace ccx

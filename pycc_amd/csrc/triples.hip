// (T) correction of closed-shell CCSD(T) (pycc/cctriples.py:177-240, t_tjl): FP64, except that the
// MIXED_F16X3 mode runs its three GEMMs per triple as FP16x3 split products with FP64 output.
//
// For every occupied triple i >= j >= k the connected numerator W[a][b][c] (cctriples.py:27-72,
// t3c_ijk without the denominator) comes from three GEMMs and one tiled permute-and-add: for each p
// in (i, j, k) the four terms that share p (two <ab|ep> t2 and two <mc|qr> t2 products) form one
// GEMM of v^2 x 2v x (v + o), whose A operand [Covvv[p] | t2[p]^T] is built once per call and whose
// B operand (t2 and -(mq|rc) columns of the two (q, r) orders side by side) is packed per triple;
// the three v^2 x 2v results are then added into W's [a][b][c] order through 16^3 LDS tiles.
// A tiled kernel then adds the disconnected part (t3d_ijk, cctriples.py:108-146: <ij|ab> t1_kc + t2_ijab f_kc and their
// permutations) to form V and reduces
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

namespace ccx {

namespace {

__device__ __forceinline__ double t3_block_sum(double x) {
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

__global__ void t3_zero_kernel(double* e) {
  if (threadIdx.x == 0) e[0] = 0.0;
}

// ---- 16^3 tiles of v^3 arrays: output (a, b, c) in [org_a, org_a + 16) x ... ----------------------
// A thread owns (b_l, c_l) = (tid / 16, tid % 16) and the 16 elements a_l = 0..15 of its column.
constexpr int TT = 16, TP = TT + 1;   // tile edge; LDS row pitch (padded against bank conflicts)

// Staging of one 16^3 box of a source array S[g0][g1][off + g2] whose dimension k is indexed by the
// output dimension Dk (0 = a, 1 = b, 2 = c). A thread loads the 16 elements (l0 = 0..15, l1 = tid / 16,
// l2 = tid % 16) -- 128-byte rows along the source's contiguous dimension -- into registers; they go
// to LDS as sh[l0][l1][l2] and are read back in the output's order. Loads of the next box are issued
// before the current box is read back, so they overlap the LDS work.
template <int D0, int D1, int D2>
__device__ __forceinline__ void t3_load(const double* __restrict__ S, int64_t s0, int64_t s1, int64_t off,
                                        int64_t v, const int64_t* org, double* r) {
  const int l1 = threadIdx.x / TT, l2 = threadIdx.x % TT;
  const int64_t g1 = org[D1] + l1, g2 = org[D2] + l2;
  const bool in = g1 < v && g2 < v;
  const double* p = S + org[D0] * s0 + g1 * s1 + off + g2;
  const int64_t n0 = v - org[D0];
#pragma unroll
  for (int l0 = 0; l0 < TT; ++l0) r[l0] = (in && l0 < n0) ? p[l0 * s0] : 0.0;
}

__device__ __forceinline__ void t3_store(double (*sh)[TT][TP], const double* r) {
  const int l1 = threadIdx.x / TT, l2 = threadIdx.x % TT;
#pragma unroll
  for (int l0 = 0; l0 < TT; ++l0) sh[l0][l1][l2] = r[l0];
}

// acc[a_l] += coef * (staged box at the output position (a_l, b_l, c_l))
template <int D0, int D1, int D2>
__device__ __forceinline__ void t3_accumulate(double (*sh)[TT][TP], double coef, double* acc) {
  const int bl = threadIdx.x / TT, cl = threadIdx.x % TT;
#pragma unroll
  for (int al = 0; al < TT; ++al) {
    const int L[3] = {al, bl, cl};
    acc[al] += coef * sh[L[D0]][L[D1]][L[D2]];
  }
}

__device__ __forceinline__ void t3_tile_origin(int64_t v, int64_t* org) {
  const int64_t nt = (v + TT - 1) / TT;
  const int64_t t = blockIdx.x;
  org[0] = t / (nt * nt) * TT;
  org[1] = t / nt % nt * TT;
  org[2] = t % nt * TT;
}

// A[p][(x, y)][kk] = Covvv[p][x][y][kk] for kk < v, t2[p][kk - v][x][y] for v <= kk < v + o, 0 above:
// the two t3c_ijk products that share p as one GEMM operand with K = v + o (padded to Kp).
__global__ void t3_operand_a_kernel(int64_t o, int64_t v, int64_t Kp, const double* __restrict__ Covvv,
                                    const double* __restrict__ t2, double* __restrict__ A) {
  const int64_t total = o * v * v * Kp;
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

// B[s][z'][kk] = t2[r][q][z][kk] for kk < v and -(m q|r z) = -<m z|q r> with m = kk - v below v + o,
// for the two (q, r) orders of slot s (z' < v: first order, z' >= v: second): slot 0 is p = i with
// (j, k), (k, j); slot 1 is p = j with (i, k), (k, i); slot 2 is p = k with (i, j), (j, i).
__global__ void t3_pack_kernel(int64_t o, int64_t v, int64_t Kp, int i, int j, int k,
                               const double* __restrict__ t2, const double* __restrict__ Cooov,
                               double* __restrict__ B) {
  const int64_t total = 3 * 2 * v * Kp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = e % Kp, zz = e / Kp % (2 * v), sl = e / (2 * v * Kp);
    const int h = zz >= v, z = (int)(zz - h * v);
    const int q0 = sl == 0 ? j : i, r0 = sl == 2 ? j : k;   // the slot's first (q, r)
    const int q = h ? r0 : q0, r = h ? q0 : r0;
    double val = 0.0;
    if (kk < v) val = t2[(((int64_t)r * o + q) * v + z) * v + kk];
    else if (kk < v + o) val = -Cooov[(((kk - v) * o + q) * o + r) * v + z];
    B[e] = val;
  }
}

// W[a][b][c] = X_ijk[a,b,c] + X_ikj[a,c,b] + X_jik[b,a,c] + X_jki[b,c,a] + X_kij[c,a,b] + X_kji[c,b,a]
// from the slot products X_s[x][y][z'] (v x v x 2v; z' >= v holds the slot's second order).
__global__ __launch_bounds__(256) void t3_assemble_kernel(int64_t v, const double* __restrict__ X,
                                                          double* __restrict__ W) {
  __shared__ double sh[TT][TT][TP];
  int64_t org[3];
  t3_tile_origin(v, org);
  const int bl = threadIdx.x / TT, cl = threadIdx.x % TT;
  const int64_t s0 = 2 * v * v, s1 = 2 * v, xs = v * v * 2 * v;
  const double *X0 = X, *X1 = X + xs, *X2 = X + 2 * xs;
  double acc[TT], r[TT];
#pragma unroll
  for (int al = 0; al < TT; ++al) acc[al] = 0.0;
  t3_load<0, 1, 2>(X0, s0, s1, 0, v, org, r);                      // X_ijk[a][b][c]
  t3_store(sh, r);
  __syncthreads();
  t3_load<0, 2, 1>(X0, s0, s1, v, v, org, r);                      // X_ikj[a][c][b]
  t3_accumulate<0, 1, 2>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 0, 2>(X1, s0, s1, 0, v, org, r);                      // X_jik[b][a][c]
  t3_accumulate<0, 2, 1>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 2, 0>(X1, s0, s1, v, v, org, r);                      // X_jki[b][c][a]
  t3_accumulate<1, 0, 2>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 0, 1>(X2, s0, s1, 0, v, org, r);                      // X_kij[c][a][b]
  t3_accumulate<1, 2, 0>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 1, 0>(X2, s0, s1, v, v, org, r);                      // X_kji[c][b][a]
  t3_accumulate<2, 0, 1>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_accumulate<2, 1, 0>(sh, 1.0, acc);
  const int64_t b = org[1] + bl, c = org[2] + cl;
  if (b < v && c < v)
    for (int al = 0; al < TT; ++al) {
      const int64_t a = org[0] + al;
      if (a < v) W[(a * v + b) * v + c] = acc[al];
    }
}

// partials[block] = sum over the block's tile of V_abc (4 W_abc + W_bca + W_cab - 2 (W_acb + W_bac +
// W_cba)) / D_ijk^abc with V = W + the disconnected numerator of t3d_ijk
__global__ __launch_bounds__(256) void t3_energy_kernel(int64_t o, int64_t v, int i, int j, int k,
                                                        const double* __restrict__ W,
                                                        const double* __restrict__ F, int64_t nmo,
                                                        const double* __restrict__ t1,
                                                        const double* __restrict__ t2,
                                                        const double* __restrict__ Ioovv,
                                                        double* __restrict__ partials) {
  __shared__ double sh[TT][TT][TP];
  int64_t org[3];
  t3_tile_origin(v, org);
  const int bl = threadIdx.x / TT, cl = threadIdx.x % TT;
  double q[TT], w[TT], r[TT];
#pragma unroll
  for (int al = 0; al < TT; ++al) q[al] = w[al] = 0.0;
  const int64_t vv = v * v;
  t3_load<0, 1, 2>(W, vv, v, 0, v, org, r);     // W_abc
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 2, 0>(W, vv, v, 0, v, org, r);     // W_bca = W[b][c][a]
  t3_accumulate<0, 1, 2>(sh, 1.0, w);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 0, 1>(W, vv, v, 0, v, org, r);     // W_cab
  t3_accumulate<1, 2, 0>(sh, 1.0, q);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<0, 2, 1>(W, vv, v, 0, v, org, r);     // W_acb
  t3_accumulate<2, 0, 1>(sh, 1.0, q);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 0, 2>(W, vv, v, 0, v, org, r);     // W_bac
  t3_accumulate<0, 2, 1>(sh, -2.0, q);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 1, 0>(W, vv, v, 0, v, org, r);     // W_cba
  t3_accumulate<1, 0, 2>(sh, -2.0, q);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_accumulate<2, 1, 0>(sh, -2.0, q);
  const int64_t b = org[1] + bl, c = org[2] + cl;
  double s = 0.0;
  if (b < v && c < v) {
    const double* Iij = Ioovv + ((int64_t)i * o + j) * vv;
    const double* Iik = Ioovv + ((int64_t)i * o + k) * vv;
    const double* Ijk = Ioovv + ((int64_t)j * o + k) * vv;
    const double* Tij = t2 + ((int64_t)i * o + j) * vv;
    const double* Tik = t2 + ((int64_t)i * o + k) * vv;
    const double* Tjk = t2 + ((int64_t)j * o + k) * vv;
    const double* Fi = F + (int64_t)i * nmo + o;   // f_ia row
    const double* Fj = F + (int64_t)j * nmo + o;
    const double* Fk = F + (int64_t)k * nmo + o;
    const int64_t ob = o + b, oc = o + c;
    const double dbc = F[i * nmo + i] + F[j * nmo + j] + F[k * nmo + k] - F[ob * nmo + ob] - F[oc * nmo + oc];
    const double ijbc = Ijk[b * v + c], tjbc = Tjk[b * v + c];
    const double t1kc = t1[k * v + c], t1jb = t1[j * v + b], fkc = Fk[c], fjb = Fj[b];
    for (int al = 0; al < TT; ++al) {
      const int64_t a = org[0] + al;
      if (a >= v) break;
      const double vabc = w[al] + Iij[a * v + b] * t1kc + Iik[a * v + c] * t1jb + ijbc * t1[i * v + a] +
                          Tij[a * v + b] * fkc + Tik[a * v + c] * fjb + tjbc * Fi[a];
      const int64_t oa = o + a;
      s += vabc * (4.0 * w[al] + q[al]) / (dbc - F[oa * nmo + oa]);
    }
  }
  s = t3_block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// e[0] += weight * sum(partials[0 .. n)), in a fixed order
__global__ void t3_accumulate_kernel(const double* __restrict__ partials, int64_t n, double weight, double* e) {
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < n; b += blockDim.x) s += partials[b];
  s = t3_block_sum(s);
  if (threadIdx.x == 0) e[0] += weight * s;
}

}  // namespace

void triples_energy(Ctx& c, const double* F, const double* t1, const double* t2, double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "triples called before the integrals were set");
  CCX_CHECK(F && t1 && t2 && e_dev, ERR_ARG, "null argument");
  const int64_t o = c.o, v = c.v, nmo = c.nmo, vv = v * v, vvv = vv * v;
  CCX_CHECK(c.Covvv && c.Cooov && c.Ioovv, ERR_STATE, "the context holds no FP64 <ov|vv>/<oo|ov> blocks");
  const int64_t nt = ceil_div(v, TT), tiles = nt * nt * nt;
  CCX_CHECK(tiles < (int64_t(1) << 31), ERR_ARG, "too many virtual orbitals for the (T) tiles");
  // In MIXED_F16X3 the slot GEMMs are FP16x3 split products (interleaved hi/lo operands, K padded
  // to 32, FP32 accumulation, FP64 output); every other mode runs them in FP64.
  const bool x3 = c.mode == MODE_MIXED_F16X3;
  const int64_t Kp = x3 ? (v + o + 31) / 32 * 32 : (v + o + 7) / 8 * 8;
  // scratch: X (3 slots of v x v x 2v), W (v^3), B (3 x 2v x Kp), tile partials, and A (o x v^2 x
  // Kp) in FP64 -- or, in MIXED_F16X3, split A and B (hi/lo interleaved, rows of 2 Kp halves) plus
  // absmax partials and the two operand scales
  const int64_t nA = o * vv * Kp, nX = 3 * 2 * vvv, nB = 3 * 2 * v * Kp;
  const int64_t np = x3_partials();
  const int64_t n64 = nX + vvv + nB + tiles + (x3 ? 0 : nA);
  const int64_t nx3 = x3 ? (nA + nB) * 2 * (int64_t)sizeof(f16) / 8 + (np + 8) / 2 + 1 : 0;   // in doubles
  double* base = nullptr;
  CCX_HIP(hipMalloc(&base, (n64 + nx3) * sizeof(double)));
  double *X = base, *W = X + nX, *B = W + vvv, *partials = B + nB, *A = partials + tiles;
  f16* Ahl = x3 ? (f16*)(base + n64) : nullptr;        // A split, [o v^2][2 Kp]
  f16* Bhl = x3 ? Ahl + nA * 2 : nullptr;             // B split, [3 x 2v][2 Kp]
  float* sc = x3 ? (float*)(Bhl + nB * 2) : nullptr;  // sc[0] = scale of A, sc[1] = scale of B
  float* part = x3 ? sc + 8 : nullptr;
  Workspace& ws = c.ws;
  const std::vector<int64_t> VVK{v, v, Kp}, VV2{v, v, 2 * v}, ZK{2 * v, Kp};

  if (x3) {   // A in FP64 only long enough to split it
    double* A64 = nullptr;
    CCX_HIP(hipMalloc(&A64, nA * sizeof(double)));
    t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, c.Covvv, t2, A64);
    CCX_HIP(hipGetLastError());
    x3_absmax<double>(A64, nA, part, s);
    x3_scale(part, 1.f, nullptr, 0.f, sc, s);
    x3_split_rows<double>(A64, o * vv, Kp, Kp, sc, Ahl, Ahl + 32, Kp, s, true);
    CCX_HIP(hipStreamSynchronize(s));
    CCX_HIP(hipFree(A64));
  } else {
    t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, c.Covvv, t2, A);
    CCX_HIP(hipGetLastError());
  }
  t3_zero_kernel<<<1, 64, 0, s>>>(e_dev);
  CCX_HIP(hipGetLastError());
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k) {
        t3_pack_kernel<<<(int)std::min<int64_t>(1024, ceil_div(nB, 256)), 256, 0, s>>>(o, v, Kp, i, j, k, t2,
                                                                                     c.Cooov, B);
        CCX_HIP(hipGetLastError());
        // X_s[x][y][z'] = sum_e <yx|ep> t_rq^{ze} - sum_m <mz|qr> t_pm^{xy}, with <yx|ep> = (px|ye) =
        // Covvv[p][x][y][e]: the twelve t3c_ijk terms of cctriples.py:50-62, four per GEMM
        const int ps[3] = {i, j, k};
        if (x3) {
          x3_absmax<double>(B, nB, part, s);
          x3_scale(part, 1.f, nullptr, 0.f, sc + 1, s);
          x3_split_rows<double>(B, 3 * 2 * v, Kp, Kp, sc + 1, Bhl, Bhl + 32, Kp, s, true);
          for (int sl = 0; sl < 3; ++sl) {
            X3Gemm g;   // X_s[(x,y)][z'] = sum_kk A_p[(x,y)][kk] B_s[z'][kk]
            g.M = vv; g.N = 2 * v; g.Kp = Kp; g.lda = 2 * Kp; g.ldb = 2 * Kp; g.ldc = 2 * v;
            g.kb = 32; g.ska = 64; g.skb = 64;
            g.Ah = Ahl + (int64_t)ps[sl] * vv * 2 * Kp; g.Al = (const f16*)g.Ah + 32;
            g.Bh = Bhl + (int64_t)sl * 2 * v * 2 * Kp; g.Bl = (const f16*)g.Bh + 32;
            g.sa = sc; g.sb = sc + 1; g.C64 = X + sl * 2 * vvv;
            x3_gemm<_Float16>(g, s, &ws);
          }
        } else {
          for (int sl = 0; sl < 3; ++sl)
            contract<double>(1.0, tv(A + (int64_t)ps[sl] * vv * Kp, "xyk", VVK), tv(B + sl * 2 * v * Kp, "zk", ZK),
                             0.0, tv(X + sl * 2 * vvv, "xyz", VV2), ws, s);
        }
        t3_assemble_kernel<<<(int)tiles, 256, 0, s>>>(v, X, W);
        CCX_HIP(hipGetLastError());
        t3_energy_kernel<<<(int)tiles, 256, 0, s>>>(o, v, i, j, k, W, F, nmo, t1, t2, c.Ioovv, partials);
        CCX_HIP(hipGetLastError());
        const double weight = 2.0 - (double)((i == j) + (i == k) + (j == k));   // cctriples.py:237
        t3_accumulate_kernel<<<1, 256, 0, s>>>(partials, tiles, weight, e_dev);
        CCX_HIP(hipGetLastError());
      }
  CCX_HIP(hipStreamSynchronize(s));
  CCX_HIP(hipFree(base));
}

}  // namespace ccx

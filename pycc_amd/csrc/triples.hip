// (T) correction of closed-shell CCSD(T) (pycc/cctriples.py:177-240, t_tjl): FP64, except that the
// MIXED_F16X3 mode runs its three GEMMs per triple as FP16x3 split products with FP64 output.
//
// For every occupied triple i >= j >= k the connected numerator W[a][b][c] (cctriples.py:27-72,
// t3c_ijk without the denominator) comes from GEMMs and one tiled permute-and-add: for each p in
// (i, j, k) the four terms that share p (two <ab|ep> t2 and two <mc|qr> t2 products) are one GEMM
// row block of v^2 x 2v x (v + o), whose A operand [Covvv[p] | t2[p]^T] is built once per call and
// whose B operand (t2 and -(mq|rc) columns of the two (q, r) orders side by side) is packed per
// triple. For fixed (i, j), up to T3_KC values of k share one GEMM per slot (wider N for p = i and
// p = j, taller M for p = k); the three v^2 x 2v blocks of a triple are then added into W's [a][b][c]
// order through 16^3 LDS tiles. A tiled kernel then adds the disconnected part (t3d_ijk,
// cctriples.py:108-146: <ij|ab> t1_kc + t2_ijab f_kc and their permutations) to form V and reduces
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

// Up to T3_KC (q, r) entries; entry e packs the two orders (q_e, r_e) and (r_e, q_e).
constexpr int T3_KC = 8;
struct T3Pairs {
  int n;
  int q[T3_KC], r[T3_KC];
};

// B[e][z'][kk] = t2[r][q][z][kk] for kk < v and -(m q|r z) = -<m z|q r> with m = kk - v below v + o,
// where (q, r) is entry e's first order for z' < v and its second order for z' >= v, z = z' mod v.
// For the triple (i, j, k) the slot of p = i takes the entry (j, k), that of p = j the entry (i, k)
// and that of p = k the entry (i, j).
__global__ void t3_pack_kernel(int64_t o, int64_t v, int64_t Kp, T3Pairs pr, const double* __restrict__ t2,
                               const double* __restrict__ Cooov, double* __restrict__ B) {
  const int64_t total = (int64_t)pr.n * 2 * v * Kp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = e % Kp, zz = e / Kp % (2 * v), en = e / (2 * v * Kp);
    const int h = zz >= v, z = (int)(zz - h * v);
    int q0 = pr.q[0], r0 = pr.r[0];
#pragma unroll
    for (int x = 1; x < T3_KC; ++x)
      if (en == x) { q0 = pr.q[x]; r0 = pr.r[x]; }
    const int q = h ? r0 : q0, r = h ? q0 : r0;
    double val = 0.0;
    if (kk < v) val = t2[(((int64_t)r * o + q) * v + z) * v + kk];
    else if (kk < v + o) val = -Cooov[(((kk - v) * o + q) * o + r) * v + z];
    B[e] = val;
  }
}

// W[a][b][c] = X_ijk[a,b,c] + X_ikj[a,c,b] + X_jik[b,a,c] + X_jki[b,c,a] + X_kij[c,a,b] + X_kji[c,b,a]
// from the slot products X_s[x][y][z'] (X_s[x * s0 + y * s1 + z'], z' in [0, 2v); z' >= v holds the
// slot's second order).
struct T3Slots {
  const double* X[3];
  int64_t s0[3], s1[3];
};
__global__ __launch_bounds__(256) void t3_assemble_kernel(int64_t v, T3Slots xs, double* __restrict__ W) {
  __shared__ double sh[TT][TT][TP];
  int64_t org[3];
  t3_tile_origin(v, org);
  const int bl = threadIdx.x / TT, cl = threadIdx.x % TT;
  const double *X0 = xs.X[0], *X1 = xs.X[1], *X2 = xs.X[2];
  double acc[TT], r[TT];
#pragma unroll
  for (int al = 0; al < TT; ++al) acc[al] = 0.0;
  t3_load<0, 1, 2>(X0, xs.s0[0], xs.s1[0], 0, v, org, r);          // X_ijk[a][b][c]
  t3_store(sh, r);
  __syncthreads();
  t3_load<0, 2, 1>(X0, xs.s0[0], xs.s1[0], v, v, org, r);          // X_ikj[a][c][b]
  t3_accumulate<0, 1, 2>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 0, 2>(X1, xs.s0[1], xs.s1[1], 0, v, org, r);          // X_jik[b][a][c]
  t3_accumulate<0, 2, 1>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<1, 2, 0>(X1, xs.s0[1], xs.s1[1], v, v, org, r);          // X_jki[b][c][a]
  t3_accumulate<1, 0, 2>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 0, 1>(X2, xs.s0[2], xs.s1[2], 0, v, org, r);          // X_kij[c][a][b]
  t3_accumulate<1, 2, 0>(sh, 1.0, acc);
  __syncthreads();
  t3_store(sh, r);
  __syncthreads();
  t3_load<2, 1, 0>(X2, xs.s0[2], xs.s1[2], v, v, org, r);          // X_kji[c][b][a]
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

void triples_energy(Ctx& c, const double* F, const double* t1, const double* t2, int i0, int i1, bool accumulate,
                    double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "triples called before the integrals were set");
  CCX_CHECK(F && t1 && t2 && e_dev, ERR_ARG, "null argument");
  CCX_CHECK(0 <= i0 && i0 <= i1 && i1 <= c.o, ERR_ARG, "triples: need 0 <= i0 <= i1 <= no");
  const int64_t o = c.o, v = c.v, nmo = c.nmo, vv = v * v, vvv = vv * v;
  CCX_CHECK(c.Covvv && c.Cooov && c.Ioovv, ERR_STATE, "the context holds no FP64 <ov|vv>/<oo|ov> blocks");
  const int64_t nt = ceil_div(v, TT), tiles = nt * nt * nt;
  CCX_CHECK(tiles < (int64_t(1) << 31), ERR_ARG, "too many virtual orbitals for the (T) tiles");
  // In MIXED_F16X3 the slot GEMMs are FP16x3 split products (interleaved hi/lo operands, K padded
  // to 32, FP32 accumulation, FP64 output); every other mode runs them in FP64.
  const bool x3 = c.mode == MODE_MIXED_F16X3;
  const int64_t Kp = x3 ? (v + o + 31) / 32 * 32 : (v + o + 7) / 8 * 8;
  const int64_t KC = T3_KC;
  // For fixed (i, j) the k values run in chunks of KC: the p = i and p = j GEMMs of a chunk share
  // their A operand and become one GEMM each with N = 2v nk, and the p = k GEMMs (one B operand,
  // consecutive A_k rows) one GEMM with M = nk v^2.
  // scratch: XI, XJ (v^2 x 2v KC each), XK (KC v^2 x 2v), W (v^3), B (rows of the i, j and k entries,
  // (2 KC + 1) 2v x Kp), tile partials, and A (o v^2 x Kp) in FP64 -- or, in MIXED_F16X3, A and B
  // split (hi/lo interleaved, rows of 2 Kp halves) plus absmax partials and the operand scales
  // A: only the rows p < i1 are read by this range (p is one of i >= j >= k)
  const int64_t nA = (int64_t)i1 * vv * Kp, nXc = vv * 2 * v * KC, nB = (2 * KC + 1) * 2 * v * Kp;
  const int64_t np = x3_partials();
  DevBuf<f16> Ahl;     // A split, [i1 v^2][2 Kp]
  DevBuf<float> scA;   // scA[0] = scale of A, then absmax partials
  if (x3) {   // A split first; its FP64 form lives only until then
    Ahl.alloc(nA * 2);
    scA.alloc(8 + 2 * np);
    // The entries of A over every p are exactly those of Covvv and of t2 (plus zero padding), so its
    // scale is that of their union: the same for every range of i, whatever rows are built.
    x3_absmax<double>(c.Covvv, o * vv * v, scA.p + 8, s);
    x3_absmax<double>(t2, o * o * vv, scA.p + 8 + np, s);
    x3_scale_union(scA.p + 8, (int)(2 * np), scA.p, s);
    if (nA > 0) {
      DevBuf<double> A64(nA);
      t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, i1, c.Covvv,
                                                                                           t2, A64.p);
      CCX_HIP(hipGetLastError());
      x3_split_rows<double>(A64.p, (int64_t)i1 * vv, Kp, Kp, scA.p, Ahl.p, Ahl.p + 32, Kp, s, true);
      CCX_HIP(hipStreamSynchronize(s));
    }
  }
  const int64_t n64 = 3 * nXc + vvv + nB + tiles + (x3 ? 0 : nA);
  const int64_t nx3 = x3 ? nB * 2 * (int64_t)sizeof(f16) / 8 + (np + 8) / 2 + 1 : 0;   // in doubles
  DevBuf<double> buf(n64 + nx3);
  double* base = buf.p;
  double *XI = base, *XJ = XI + nXc, *XK = XJ + nXc, *W = XK + nXc, *B = W + vvv, *partials = B + nB;
  double* A = partials + tiles;
  f16* Bhl = x3 ? (f16*)(base + n64) : nullptr;       // B split, [(2 KC + 1) 2v][2 Kp]
  float* sc = x3 ? (float*)(Bhl + nB * 2) : nullptr;  // sc[1] = scale of B
  float* part = x3 ? sc + 8 : nullptr;
  Workspace& ws = c.ws;

  if (!x3 && nA > 0) {
    t3_operand_a_kernel<<<(int)std::min<int64_t>(65536, ceil_div(nA, 256)), 256, 0, s>>>(o, v, Kp, i1, c.Covvv, t2, A);
    CCX_HIP(hipGetLastError());
  }
  if (!accumulate) {
    t3_zero_kernel<<<1, 64, 0, s>>>(e_dev);
    CCX_HIP(hipGetLastError());
  }
  const int64_t ldx = 2 * v * KC;   // row length of XI, XJ
  for (int i = i0; i < i1; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k0 = 0; k0 <= j; k0 += (int)KC) {
        const int nk = (int)std::min<int64_t>(KC, j + 1 - k0);
        // B rows: [0, nk 2v) the p = i entries (j, k), [nk 2v, 2 nk 2v) the p = j entries (i, k), then
        // the p = k entry (i, j); X_s[x][y][z'] = sum_e <yx|ep> t_rq^{ze} - sum_m <mz|qr> t_pm^{xy} with
        // <yx|ep> = (px|ye) = Covvv[p][x][y][e]: the twelve t3c_ijk terms of cctriples.py:50-62
        T3Pairs pi{nk, {}, {}}, pj{nk, {}, {}}, pk{1, {}, {}};
        for (int e = 0; e < nk; ++e) {
          pi.q[e] = j; pi.r[e] = k0 + e;
          pj.q[e] = i; pj.r[e] = k0 + e;
        }
        pk.q[0] = i; pk.r[0] = j;
        double* BI = B;
        double* BJ = BI + (int64_t)nk * 2 * v * Kp;
        double* BK = BJ + (int64_t)nk * 2 * v * Kp;
        const int64_t nrows = (2 * (int64_t)nk + 1) * 2 * v;
        const int pg = (int)std::min<int64_t>(1024, ceil_div((int64_t)nk * 2 * v * Kp, 256));
        t3_pack_kernel<<<pg, 256, 0, s>>>(o, v, Kp, pi, t2, c.Cooov, BI);
        CCX_HIP(hipGetLastError());
        t3_pack_kernel<<<pg, 256, 0, s>>>(o, v, Kp, pj, t2, c.Cooov, BJ);
        CCX_HIP(hipGetLastError());
        t3_pack_kernel<<<(int)std::min<int64_t>(1024, ceil_div(2 * v * Kp, 256)), 256, 0, s>>>(o, v, Kp, pk, t2,
                                                                                           c.Cooov, BK);
        CCX_HIP(hipGetLastError());
        if (x3) {
          x3_absmax<double>(B, nrows * Kp, part, s);
          x3_scale(part, 1.f, nullptr, 0.f, sc + 1, s);
          x3_split_rows<double>(B, nrows, Kp, Kp, sc + 1, Bhl, Bhl + 32, Kp, s, true);
          const f16* BhI = Bhl;
          const f16* BhJ = Bhl + (int64_t)nk * 2 * v * 2 * Kp;
          const f16* BhK = BhJ + (int64_t)nk * 2 * v * 2 * Kp;
          auto run = [&](const f16* Ah, int64_t M, const f16* Bh, int64_t N, double* C, int64_t ldc) {
            X3Gemm g;   // C[m][n] = sum_kk A[m][kk] B[n][kk]
            g.M = M; g.N = N; g.Kp = Kp; g.lda = 2 * Kp; g.ldb = 2 * Kp; g.ldc = ldc;
            g.kb = 32; g.ska = 64; g.skb = 64;
            g.Ah = Ah; g.Al = Ah + 32; g.Bh = Bh; g.Bl = Bh + 32;
            g.sa = scA.p; g.sb = sc + 1; g.C64 = C;
            x3_gemm<_Float16>(g, s, &ws);
          };
          run(Ahl.p + (int64_t)i * vv * 2 * Kp, vv, BhI, 2 * v * nk, XI, ldx);
          run(Ahl.p + (int64_t)j * vv * 2 * Kp, vv, BhJ, 2 * v * nk, XJ, ldx);
          run(Ahl.p + (int64_t)k0 * vv * 2 * Kp, nk * vv, BhK, 2 * v, XK, 2 * v);
        } else {
          const std::vector<int64_t> VVK{v, v, Kp};
          const TV XIv = tvs(XI, "xyz", {v, v, 2 * v * nk}, {v * ldx, ldx, 1});
          const TV XJv = tvs(XJ, "xyz", {v, v, 2 * v * nk}, {v * ldx, ldx, 1});
          contract<double>(1.0, tv(A + (int64_t)i * vv * Kp, "xyk", VVK), tv(BI, "zk", {2 * v * nk, Kp}), 0.0, XIv,
                           ws, s);
          contract<double>(1.0, tv(A + (int64_t)j * vv * Kp, "xyk", VVK), tv(BJ, "zk", {2 * v * nk, Kp}), 0.0, XJv,
                           ws, s);
          contract<double>(1.0, tv(A + (int64_t)k0 * vv * Kp, "mk", {nk * vv, Kp}), tv(BK, "zk", {2 * v, Kp}), 0.0,
                           tv(XK, "mz", {nk * vv, 2 * v}), ws, s);
        }
        for (int e = 0; e < nk; ++e) {
          const int k = k0 + e;
          T3Slots xs;
          xs.X[0] = XI + (int64_t)e * 2 * v; xs.s0[0] = v * ldx; xs.s1[0] = ldx;
          xs.X[1] = XJ + (int64_t)e * 2 * v; xs.s0[1] = v * ldx; xs.s1[1] = ldx;
          xs.X[2] = XK + (int64_t)e * vv * 2 * v; xs.s0[2] = 2 * vv; xs.s1[2] = 2 * v;
          t3_assemble_kernel<<<(int)tiles, 256, 0, s>>>(v, xs, W);
          CCX_HIP(hipGetLastError());
          t3_energy_kernel<<<(int)tiles, 256, 0, s>>>(o, v, i, j, k, W, F, nmo, t1, t2, c.Ioovv, partials);
          CCX_HIP(hipGetLastError());
          const double weight = 2.0 - (double)((i == j) + (i == k) + (j == k));   // cctriples.py:237
          t3_accumulate_kernel<<<1, 256, 0, s>>>(partials, tiles, weight, e_dev);
          CCX_HIP(hipGetLastError());
        }
      }
  CCX_HIP(hipStreamSynchronize(s));
}

}  // namespace ccx

// (T) correction of closed-shell CCSD(T) in FP64 (pycc/cctriples.py:177-240, t_tjl).
//
// For every occupied triple i >= j >= k the connected numerator W[a][b][c] (cctriples.py:27-72,
// t3c_ijk without the denominator) is accumulated into one v^3 buffer by twelve contractions: six
// <ab|ei> t2 terms (GEMMs of v^2 x v x v) and six <mc|jk> t2 terms (v^2 x v x o), written straight
// into W's [a][b][c] order by the strided contraction engine. A fused kernel then adds the
// disconnected part (t3d_ijk, cctriples.py:108-146: <ij|ab> t1_kc + t2_ijab f_kc and their
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

namespace ccx {

namespace {

constexpr int T3_BLOCKS = 512;

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

// partials[block] = sum over this block's abc of V (4W + W_bca + W_cab - 2(W_acb + W_bac + W_cba)) / D
__global__ __launch_bounds__(256) void t3_energy_kernel(int64_t o, int64_t v, int i, int j, int k,
                                                        const double* __restrict__ W,
                                                        const double* __restrict__ F, int64_t nmo,
                                                        const double* __restrict__ t1,
                                                        const double* __restrict__ t2,
                                                        const double* __restrict__ Ioovv,
                                                        double* __restrict__ partials) {
  const int64_t vv = v * v, total = vv * v;
  const double dijk = F[i * nmo + i] + F[j * nmo + j] + F[k * nmo + k];
  const double* Iij = Ioovv + ((int64_t)i * o + j) * vv;
  const double* Iik = Ioovv + ((int64_t)i * o + k) * vv;
  const double* Ijk = Ioovv + ((int64_t)j * o + k) * vv;
  const double* Tij = t2 + ((int64_t)i * o + j) * vv;
  const double* Tik = t2 + ((int64_t)i * o + k) * vv;
  const double* Tjk = t2 + ((int64_t)j * o + k) * vv;
  const double* Fi = F + (int64_t)i * nmo + o;   // f_ia row
  const double* Fj = F + (int64_t)j * nmo + o;
  const double* Fk = F + (int64_t)k * nmo + o;
  double s = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % v, ab = e / v;
    const int64_t b = ab % v, a = ab / v;
    const double w_abc = W[e];
    const double w_bca = W[(b * v + c) * v + a];
    const double w_cab = W[(c * v + a) * v + b];
    const double w_acb = W[(a * v + c) * v + b];
    const double w_bac = W[(b * v + a) * v + c];
    const double w_cba = W[(c * v + b) * v + a];
    const double vabc = w_abc + Iij[a * v + b] * t1[k * v + c] + Iik[a * v + c] * t1[j * v + b] +
                        Ijk[b * v + c] * t1[i * v + a] + Tij[a * v + b] * Fk[c] + Tik[a * v + c] * Fj[b] +
                        Tjk[b * v + c] * Fi[a];
    const int64_t oa = o + a, ob = o + b, oc = o + c;
    const double d = dijk - F[oa * nmo + oa] - F[ob * nmo + ob] - F[oc * nmo + oc];
    s += vabc * (4.0 * w_abc + w_bca + w_cab - 2.0 * (w_acb + w_bac + w_cba)) / d;
  }
  s = t3_block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// e[0] += weight * sum(partials[0 .. n)), in a fixed order
__global__ void t3_accumulate_kernel(const double* __restrict__ partials, int n, double weight, double* e) {
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) s += partials[b];
  s = t3_block_sum(s);
  if (threadIdx.x == 0) e[0] += weight * s;
}

}  // namespace

void triples_energy(Ctx& c, const double* F, const double* t1, const double* t2, double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "triples called before the integrals were set");
  CCX_CHECK(F && t1 && t2 && e_dev, ERR_ARG, "null argument");
  const int64_t o = c.o, v = c.v, nmo = c.nmo, ov = o * v, vvv = v * v * v;
  CCX_CHECK(c.Covvv && c.Cooov && c.Ioovv, ERR_STATE, "the context holds no FP64 <ov|vv>/<oo|ov> blocks");
  if (!c.red) c.red = c.alloc<double>(1024 * 64);
  double* partials = c.red + 1024 * 48;
  const int blocks = (int)std::min<int64_t>(T3_BLOCKS, std::max<int64_t>(1, ceil_div(vvv, 256)));
  double* W = nullptr;
  CCX_HIP(hipMalloc(&W, vvv * sizeof(double)));
  Workspace& ws = c.ws;
  const std::vector<int64_t> VV{v, v}, VVV{v, v, v}, OVV{o, v, v}, OV{o, v};
  const TV Wabc = tv(W, "abc", VVV);
  // Wvvvo[x][y][e][p] = <xy|ep> = (xe|yp) = Covvv[p][y][x][e]: `lab` names Covvv[p]'s dimensions
  auto Wv = [&](int p, const char* lab) { return tv(c.Covvv + (int64_t)p * vvv, lab, VVV); };
  // Wovoo[m][c][q][r] = <mc|qr> = (mq|cr) = Cooov[m][q][r][c]
  auto Wo = [&](int q, int r, const char* lab) {
    return tvs(c.Cooov + (int64_t)q * ov + (int64_t)r * v, lab, OV, {o * o * v, 1});
  };
  auto T2p = [&](int p, int q, const char* lab) { return tv(t2 + ((int64_t)p * o + q) * v * v, lab, VV); };
  auto T2r = [&](int p, const char* lab) { return tv(t2 + (int64_t)p * o * v * v, lab, OVV); };
  auto cd = [&](double a, const TV& A, const TV& B, double b) { contract<double>(a, A, B, b, Wabc, ws, s); };

  t3_zero_kernel<<<1, 64, 0, s>>>(e_dev);
  CCX_HIP(hipGetLastError());
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k) {
        // t3c_ijk (cctriples.py:50-62), each term in the reference's order
        cd(1.0, Wv(i, "abe"), T2p(k, j, "ce"), 0.0);     // bae,ce->abc  W[:,:,:,i] t2[k,j]
        cd(1.0, Wv(i, "ace"), T2p(j, k, "be"), 1.0);     // cae,be->abc  W[:,:,:,i] t2[j,k]
        cd(1.0, Wv(k, "cae"), T2p(j, i, "be"), 1.0);     // ace,be->abc  W[:,:,:,k] t2[j,i]
        cd(1.0, Wv(k, "cbe"), T2p(i, j, "ae"), 1.0);     // bce,ae->abc  W[:,:,:,k] t2[i,j]
        cd(1.0, Wv(j, "bce"), T2p(i, k, "ae"), 1.0);     // cbe,ae->abc  W[:,:,:,j] t2[i,k]
        cd(1.0, Wv(j, "bae"), T2p(k, i, "ce"), 1.0);     // abe,ce->abc  W[:,:,:,j] t2[k,i]
        cd(-1.0, Wo(j, k, "mc"), T2r(i, "mab"), 1.0);    // mc,mab->abc  Wovoo[:,:,j,k] t2[i]
        cd(-1.0, Wo(k, j, "mb"), T2r(i, "mac"), 1.0);    // mb,mac->abc  Wovoo[:,:,k,j] t2[i]
        cd(-1.0, Wo(i, j, "mb"), T2r(k, "mca"), 1.0);    // mb,mca->abc  Wovoo[:,:,i,j] t2[k]
        cd(-1.0, Wo(j, i, "ma"), T2r(k, "mcb"), 1.0);    // ma,mcb->abc  Wovoo[:,:,j,i] t2[k]
        cd(-1.0, Wo(k, i, "ma"), T2r(j, "mbc"), 1.0);    // ma,mbc->abc  Wovoo[:,:,k,i] t2[j]
        cd(-1.0, Wo(i, k, "mc"), T2r(j, "mba"), 1.0);    // mc,mba->abc  Wovoo[:,:,i,k] t2[j]
        t3_energy_kernel<<<blocks, 256, 0, s>>>(o, v, i, j, k, W, F, nmo, t1, t2, c.Ioovv, partials);
        CCX_HIP(hipGetLastError());
        const double weight = 2.0 - (double)((i == j) + (i == k) + (j == k));   // cctriples.py:237
        t3_accumulate_kernel<<<1, 256, 0, s>>>(partials, blocks, weight, e_dev);
        CCX_HIP(hipGetLastError());
      }
  CCX_HIP(hipStreamSynchronize(s));
  CCX_HIP(hipFree(W));
}

}  // namespace ccx

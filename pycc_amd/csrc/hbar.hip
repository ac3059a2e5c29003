// CCSD similarity-transformed Hamiltonian (HBAR) and the Lambda residual on gfx950, FP64.
//
// Reference: pycc/cchbar.py (spatial CCSD branches of the block builders, :128-823) and
// pycc/cclambda.py (build_Goo :281, build_Gvv :306, r_L1 :345-369, r_L2 :468-496,
// pseudoenergy :570). Every contraction cites its reference line; the Dirac integral of each term
// is read from a chemist block, <pq|rs> = (pr|qs).
//
// Two ways to hold HBAR:
//   stored    all 11 blocks, as the reference builds them (Hvvvv is v^4, Hvvvo and Hvovv ov^3);
//   factored  the blocks of size o^2 v^2 and below only. Every Lambda term that reads Hvvvv, Hvvvo
//             or Hvovv is expanded into the packed ladder GEMM of the forward residual (<ab|ef> in
//             its symmetric/antisymmetric pair form), o^3 v^3 contractions and smaller ones, so no
//             v^4 or ov^3 HBAR block is formed and the ladder costs o^2 v^4 / 2 flops per Lambda
//             residual instead of 2 o^2 v^4. The factored form needs l2 with pair symmetry
//             (l2_ijab = l2_jiba), which every Lambda amplitude has.
//
// The Lambda residual with the source terms (2 Hov in r_L1, L in r_L2) left out is the
// vector-Jacobian product of the CCSD residual: with the Lagrangian E + l1.r1 + 1/2 l2.r2,
// r_L1 = dL/dt1 and r_L2 = dL/dt2 + its (ij)(ab) transpose at any t and l, and the source terms
// are the energy gradient.
#include "ccsd.h"

namespace ccx {

namespace {

// <ab|ef> for every a, b, e, f from the packed FP64 ladder matrices (rows (e<=f), columns (a<=b)):
// <ab|ef> = Vp[(ef)][(ab)] + s(a,b) s(e,f) Vm[(ef)][(ab)], s(p,q) = +1 (p < q), -1 (p > q).
__global__ void unpack_vvvv_kernel(int64_t o, int64_t v, const double* __restrict__ LP,
                                   const double* __restrict__ LM, double* __restrict__ out) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2, ov = o * v;
  const int64_t ldP = npP + ov, ldM = npM + ov;
  const int64_t total = v * v * v * v;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
       x += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = x % v, t = x / v;
    int64_t e = t % v;
    t /= v;
    int64_t b = t % v, a = t / v;
    int64_t el = e < f ? e : f, eh = e < f ? f : e, al = a < b ? a : b, ah = a < b ? b : a;
    double r = LP[pairP(el, eh) * ldP + pairP(al, ah)];
    if (e != f && a != b) {
      double m = LM[pairM(el, eh) * ldM + pairM(al, ah)];
      r += ((e < f) == (a < b)) ? m : -m;
    }
    out[x] = r;
  }
}

// out[i][j][a][b] = alpha * (h[i][j][a][b] + h[j][i][b][a]), through 32 x 33 LDS tiles.
__global__ __launch_bounds__(256) void pair_sym_kernel(int64_t o, int64_t v, double alpha,
                                                       const double* __restrict__ h, double* __restrict__ out) {
  __shared__ double sh[32][33];
  const int64_t ij = blockIdx.y, i = ij / o, j = ij % o;
  const int64_t tiles_b = (v + 31) / 32;
  const int64_t a0 = (blockIdx.x / tiles_b) * 32, b0 = (blockIdx.x % tiles_b) * 32;
  const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
  for (int y = y0; y < 32; y += 8) {
    const int64_t a = a0 + x, b = b0 + y;
    sh[y][x] = (a < v && b < v) ? h[((j * o + i) * v + b) * v + a] : 0.0;
  }
  __syncthreads();
  for (int y = y0; y < 32; y += 8) {
    const int64_t a = a0 + y, b = b0 + x;
    if (a >= v || b >= v) continue;
    const int64_t e = ((i * o + j) * v + a) * v + b;
    out[e] = alpha * (h[e] + sh[x][y]);
  }
}

// Y[i][j][a][b] = CP[(ij)][(ab)] + s(i,j) s(a,b) CM[(ij)][(ab)]: the full ladder product
// sum_ef X_ijef <ab|ef> of a pair-symmetric X from its packed GEMM outputs.
template <typename TC>
__global__ void ladder_unpack_kernel(int64_t o, int64_t v, const TC* __restrict__ CP,
                                     const TC* __restrict__ CM, double* __restrict__ Y) {
  const int64_t npP = v * (v + 1) / 2, npM = v * (v - 1) / 2, ov = o * v;
  const int64_t ldP = npP + ov, ldM = npM + ov;
  const int64_t ij = blockIdx.y, i = ij / o, j = ij % o;
  const int64_t il = i < j ? i : j, ih = i < j ? j : i;
  const TC* cpr = CP + pairP(il, ih) * ldP;
  const TC* cmr = i != j ? CM + pairM(il, ih) * ldM : nullptr;
  const bool sij = i < j;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < v * v;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = x / v, b = x % v;
    const int64_t al = a < b ? a : b, ah = a < b ? b : a;
    double r = (double)cpr[pairP(al, ah)];
    if (cmr && a != b) {
      double m = (double)cmr[pairM(al, ah)];
      r += (sij == (a < b)) ? m : -m;
    }
    Y[ij * v * v + x] = r;
  }
}

// Z[i][j][e][f] += x_ie t_jf + t_ie x_jf (the t1 part of a Jacobian direction, pair-symmetric)
__global__ void pair_outer_kernel(int64_t o, int64_t v, const double* __restrict__ x, const double* __restrict__ t,
                                  double* __restrict__ Z) {
  const int64_t total = o * o * v * v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = e % v, r = e / v;
    int64_t a = r % v;
    r /= v;
    int64_t j = r % o, i = r / o;
    Z[e] += x[i * v + a] * t[j * v + f] + t[i * v + a] * x[j * v + f];
  }
}

// out[0] = sum of x[0..n) in a fixed order (one block).
__global__ void sum_terms_kernel(int n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double sh[256];
  double a = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) a += x[k];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace

void pair_symmetrize(int64_t o, int64_t v, double alpha, const double* h, double* out, hipStream_t s) {
  CCX_CHECK(o * o < 65536, ERR_ARG, "too many occupied orbitals for the pair grid");
  int64_t tiles = ceil_div(v, 32);
  pair_sym_kernel<<<dim3((unsigned)(tiles * tiles), (unsigned)(o * o)), 256, 0, s>>>(o, v, alpha, h, out);
  CCX_HIP(hipGetLastError());
}

static void alloc_hbar(Ctx& c, bool stored) {
  const int64_t o = c.o, v = c.v, ov = o * v;
  if (!c.have_hbar_bufs) {
    c.Ht1 = c.alloc<double>(ov);
    c.Ht2 = c.alloc<double>(ov * ov);
    c.Hov = c.alloc<double>(ov);
    c.Hvv = c.alloc<double>(v * v);
    c.Hoo = c.alloc<double>(o * o);
    c.Hoooo = c.alloc<double>(o * o * o * o);
    c.Hooov = c.alloc<double>(o * o * ov);
    c.Hovvo = c.alloc<double>(ov * ov);
    c.Hovov = c.alloc<double>(ov * ov);
    c.Hovoo = c.alloc<double>(o * o * ov);
    c.Ltmp = c.alloc<double>(ov * ov);
    c.Goo = c.alloc<double>(o * o);
    c.Gvv = c.alloc<double>(v * v);
    c.Lh = c.alloc<double>(ov * ov);
    c.Lsym = c.alloc<double>(ov * ov);
    c.LY = c.alloc<double>(ov * ov);
    c.LVf = c.alloc<double>(o * o * ov);
    c.LVg = c.alloc<double>(o * o * ov);
    c.LU = c.alloc<double>(o * o * o * o);
    c.LQ = c.alloc<double>(o * o * ov);
    c.Lsm = c.alloc<double>(2 * o * o + 2 * ov);
    c.Jsm = c.alloc<double>(v * v + 4 * o * o + 2 * ov);
    c.have_hbar_bufs = true;
  }
  if (stored && !c.Hvvvv) {
    c.Hvvvv = c.alloc<double>(v * v * v * v);
    c.Hvovv = c.alloc<double>(ov * v * v);
    c.Hvvvo = c.alloc<double>(ov * v * v);
  }
}

void hbar_build(Ctx& c, const double* F, const double* t1, const double* t2, bool stored, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "HBAR requested before the integrals were set");
  // DP: every contraction in FP64. MIXED_F16X3: the large ones as FP16x3 split products
  // (x3_contract), HBAR still held in FP64; the v^4 block is never stored there (the packed ladder
  // is split FP16).
  CCX_CHECK(c.mode == MODE_DP || (c.mode == MODE_MIXED_F16X3 && !stored), ERR_STATE,
            "HBAR and the Lambda equations run in a DP context, or in a MIXED_F16X3 one without the "
            "stored v^4 block");
  CCX_CHECK(F && t1 && t2, ERR_ARG, "null argument");
  CCX_CHECK(c.o * c.o < 65536, ERR_ARG, "too many occupied orbitals");
  alloc_work(c);
  alloc_hbar(c, stored);
  c.have_hbar = false;
  const int64_t o = c.o, v = c.v, nmo = c.nmo, ov = o * v;
  Workspace& ws = c.ws;
  const std::vector<int64_t> OV{o, v}, OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OVVV{o, v, v, v}, OOOO{o, o, o, o}, VVVV{v, v, v, v};
  const std::vector<int64_t> VOVV{v, o, v, v}, OVVO{o, v, v, o}, VVVO{v, v, v, o}, OVOO{o, v, o, o};
  const bool x3 = c.mode == MODE_MIXED_F16X3;
  auto cd = [&](double a, const TV& A, const TV& B, double b, const TV& C) {
    contract<double>(a, A, B, b, C, ws, s, x3);
  };
  CCX_HIP(hipMemcpyAsync(c.Ht1, t1, ov * sizeof(double), hipMemcpyDeviceToDevice, s));
  CCX_HIP(hipMemcpyAsync(c.Ht2, t2, ov * ov * sizeof(double), hipMemcpyDeviceToDevice, s));
  auto T1 = [&](const char* l) { return tv(c.Ht1, l, OV); };
  auto T2 = [&](const char* l) { return tv(c.Ht2, l, OOVV); };
  const TV Fov = tvs(F + o, "me", OV, {nmo, 1});
  const TV Fvv = tvs(F + o * nmo + o, "ae", VV, {nmo, 1});
  const TV Foo = tvs(F, "mi", OO, {nmo, 1});
  build_tau(o, v, 1.0, 1.0, c.Ht1, c.Ht2, c.tau, s);      // tau = t2 + t1 t1 (ccwfn.py:455)
  auto TAU = [&](const char* l) { return tv(c.tau, l, OOVV); };
  auto CVVV = [&](const char* l) { return tv(c.Covvv, l, OVVV); };   // (me|af) stored [m][e][a][f]
  auto COOV = [&](const char* l) { return tv(c.Cooov, l, OOOV); };   // (mi|ne) stored [m][i][n][e]
  auto COVV = [&](const char* l) { return tv(c.Coovv, l, OOVV); };   // (mn|ef) stored [m][n][e][f]
  auto COVO = [&](const char* l) { return tv(c.Covov, l, OVOV); };   // (me|nf) stored [m][e][n][f]
  auto IOVV = [&](const char* l) { return tv(c.Ioovv, l, OOVV); };   // <mn|ef> stored [m][n][e][f]
  auto LOVV = [&](const char* l) { return tv(c.Loovv, l, OOVV); };   // L_mnef
  auto LOOV = [&](const char* l) { return tv(c.Looov, l, OOOV); };   // L_mnie stored [i][m][n][e]

  // ---- Hov: cchbar.py:152-153 ----
  const TV Hov = tv(c.Hov, "me", OV);
  axpby<double>(1.0, Fov, 0.0, Hov, s);
  cd(1.0, T1("nf"), LOVV("mnef"), 1.0, Hov);

  // ---- Hvv: :206-209 (L_amef = 2(mf|ae) - (me|af)) ----
  const TV Hvv = tv(c.Hvv, "ae", VV);
  axpby<double>(1.0, Fvv, 0.0, Hvv, s);
  cd(-1.0, Fov, T1("ma"), 1.0, Hvv);
  cd(2.0, T1("mf"), CVVV("mfae"), 1.0, Hvv);
  cd(-1.0, T1("mf"), CVVV("meaf"), 1.0, Hvv);
  cd(-1.0, TAU("mnfa"), LOVV("mnfe"), 1.0, Hvv);

  // ---- Hoo: :269-272 ----
  const TV Hoo = tv(c.Hoo, "mi", OO);
  axpby<double>(1.0, Foo, 0.0, Hoo, s);
  cd(1.0, T1("ie"), Fov, 1.0, Hoo);
  cd(1.0, T1("ne"), LOOV("imne"), 1.0, Hoo);
  cd(1.0, TAU("inef"), LOVV("mnef"), 1.0, Hoo);

  // ---- Hoooo: :331-337 -- <mn|ij> + t_je <mn|ie> + t_ie <nm|je> + tau_ijef <mn|ef> ----
  const TV Hoooo = tv(c.Hoooo, "mnij", OOOO);
  axpby<double>(1.0, tv(c.Coooo, "minj", OOOO), 0.0, Hoooo, s);     // (mi|nj)
  cd(1.0, T1("je"), COOV("mine"), 1.0, Hoooo);                       // (mi|ne)
  cd(1.0, T1("ie"), COOV("njme"), 1.0, Hoooo);                       // (nj|me)
  cd(1.0, TAU("ijef"), IOVV("mnef"), 1.0, Hoooo);

  // ---- Hooov: :505-506 -- <mn|ie> + t_if <nm|ef> ----
  const TV Hooov = tv(c.Hooov, "mnie", OOOV);
  axpby<double>(1.0, COOV("mine"), 0.0, Hooov, s);
  cd(1.0, T1("if"), IOVV("nmef"), 1.0, Hooov);

  // ---- Hovvo: :563-568 -- <mb|ej> + t_jf <mb|ef> - t_nb <mn|ej> - tau_jnfb <mn|ef> + t_njfb L_mnef ----
  const TV Hovvo = tv(c.Hovvo, "mbej", OVVO);
  axpby<double>(1.0, COVO("mejb"), 0.0, Hovvo, s);                  // (me|jb)
  cd(1.0, T1("jf"), CVVV("mebf"), 1.0, Hovvo);                       // (me|bf)
  cd(-1.0, T1("nb"), COOV("njme"), 1.0, Hovvo);                      // (nj|me)
  cd(-1.0, TAU("jnfb"), IOVV("mnef"), 1.0, Hovvo);
  cd(1.0, T2("njfb"), LOVV("mnef"), 1.0, Hovvo);

  // ---- Hovov: :625-629 -- <mb|je> + t_jf <bm|ef> - t_nb <mn|je> - tau_jnfb <nm|ef> ----
  const TV Hovov = tv(c.Hovov, "mbje", OVOV);
  axpby<double>(1.0, COVV("mjbe"), 0.0, Hovov, s);                  // (mj|be)
  cd(1.0, T1("jf"), CVVV("mfbe"), 1.0, Hovov);                       // <bm|ef> = (mf|be)
  cd(-1.0, T1("nb"), COOV("mjne"), 1.0, Hovov);                      // <mn|je> = (mj|ne)
  cd(-1.0, TAU("jnfb"), IOVV("nmef"), 1.0, Hovov);

  // ---- Hovoo: :806-819 ----
  const TV Hovoo = tv(c.Hovoo, "mbij", OVOO);
  axpby<double>(1.0, COOV("mijb"), 0.0, Hovoo, s);                  // <mb|ij> = (mi|jb)    :806
  cd(1.0, Hov, T2("ijeb"), 1.0, Hovoo);                              //                      :807
  cd(-1.0, T1("nb"), Hoooo, 1.0, Hovoo);                             //                      :808
  cd(1.0, TAU("ijef"), CVVV("mebf"), 1.0, Hovoo);                    // <mb|ef> = (me|bf)    :809
  cd(-1.0, T2("ineb"), COOV("njme"), 1.0, Hovoo);                    // <nm|je> = (nj|me)    :810
  cd(-1.0, T2("jneb"), COOV("mine"), 1.0, Hovoo);                    // <mn|ie> = (mi|ne)    :811
  cd(1.0, T2("njeb"), LOOV("imne"), 1.0, Hovoo);                     // L_mnie               :812
  {
    const TV tmp = tv(c.Ltmp, "mbie", OVOV);                        //                      :813-815
    axpby<double>(1.0, COVV("mibe"), 0.0, tmp, s);                   // <mb|ie> = (mi|be)
    cd(-1.0, T2("infb"), IOVV("mnfe"), 1.0, tmp);
    cd(1.0, T1("je"), tmp, 1.0, Hovoo);
  }
  {
    const TV tmp = tv(c.Ltmp, "bmje", {v, o, o, v});                //                      :816-819
    axpby<double>(1.0, COVO("mejb"), 0.0, tmp, s);                   // <bm|je> = (me|jb)
    cd(-1.0, T2("jnfb"), IOVV("mnef"), 1.0, tmp);
    cd(1.0, T2("njfb"), LOVV("mnef"), 1.0, tmp);
    cd(1.0, T1("ie"), tmp, 1.0, Hovoo);
  }

  c.hbar_stored = stored;
  if (stored) {
    // ---- Hvvvv: :395-401 -- <ab|ef> - t_mb <am|ef> - t_ma <bm|fe> + tau_mnab <mn|ef> ----
    const TV Hvvvv = tv(c.Hvvvv, "abef", VVVV);
    const int64_t total = v * v * v * v;
    unpack_vvvv_kernel<<<(unsigned)std::min<int64_t>(ceil_div(total, 256), 65536), 256, 0, s>>>(
        o, v, (const double*)c.LP, (const double*)c.LM, c.Hvvvv);
    CCX_HIP(hipGetLastError());
    cd(-1.0, T1("mb"), CVVV("mfae"), 1.0, Hvvvv);                    // <am|ef> = (mf|ae)
    cd(-1.0, T1("ma"), CVVV("mebf"), 1.0, Hvvvv);                    // <bm|fe> = (me|bf)
    cd(1.0, TAU("mnab"), IOVV("mnef"), 1.0, Hvvvv);

    // ---- Hvovv: :455-456 -- <am|ef> - t_na <nm|ef> ----
    const TV Hvovv = tv(c.Hvovv, "amef", VOVV);
    axpby<double>(1.0, CVVV("mfae"), 0.0, Hvovv, s);
    cd(-1.0, T1("na"), IOVV("nmef"), 1.0, Hvovv);

    // ---- Hvvvo: :682-695 ----
    const TV Hvvvo = tv(c.Hvvvo, "abei", VVVO);
    axpby<double>(1.0, CVVV("ibae"), 0.0, Hvvvo, s);                // <ab|ei> = (ib|ae)    :682
    cd(-1.0, Hov, T2("miab"), 1.0, Hvvvo);                           //                      :683
    cd(1.0, T1("if"), Hvvvv, 1.0, Hvvvo);                            //                      :684
    cd(1.0, TAU("mnab"), COOV("nime"), 1.0, Hvvvo);                  // <mn|ei> = (ni|me)    :685
    cd(-1.0, T2("imfa"), CVVV("mebf"), 1.0, Hvvvo);                  // <bm|fe> = (me|bf)    :686
    cd(-1.0, T2("imfb"), CVVV("mfae"), 1.0, Hvvvo);                  // <am|ef> = (mf|ae)    :687
    cd(2.0, T2("mifb"), CVVV("mfae"), 1.0, Hvvvo);                   // L_amef = 2(mf|ae)    :688
    cd(-1.0, T2("mifb"), CVVV("meaf"), 1.0, Hvvvo);                  //        - (me|af)
    {
      const TV tmp = tv(c.Ltmp, "amei", {v, o, v, o});              //                      :689-691
      axpby<double>(1.0, COVV("miae"), 0.0, tmp, s);                 // <am|ei> = (mi|ae)
      cd(-1.0, T2("infa"), IOVV("mnfe"), 1.0, tmp);
      cd(-1.0, T1("mb"), tmp, 1.0, Hvvvo);
    }
    {
      const TV tmp = tv(c.Ltmp, "bmie", {v, o, o, v});              //                      :692-695
      axpby<double>(1.0, COVO("meib"), 0.0, tmp, s);                 // <bm|ie> = (me|ib)
      cd(-1.0, T2("infb"), IOVV("mnef"), 1.0, tmp);
      cd(1.0, T2("nifb"), LOVV("mnef"), 1.0, tmp);
      cd(-1.0, T1("ma"), tmp, 1.0, Hvvvo);
    }
  }
  c.have_hbar = true;
}

void hbar_block(Ctx& c, int block, double* out, hipStream_t s) {
  CCX_CHECK(c.have_hbar, ERR_STATE, "no HBAR has been built");
  CCX_CHECK(out, ERR_ARG, "null output");
  const int64_t o = c.o, v = c.v;
  const double* src = nullptr;
  int64_t n = 0;
  bool big = false;
  switch (block) {
    case 0: src = c.Hov; n = o * v; break;
    case 1: src = c.Hvv; n = v * v; break;
    case 2: src = c.Hoo; n = o * o; break;
    case 3: src = c.Hoooo; n = o * o * o * o; break;
    case 4: src = c.Hvvvv; n = v * v * v * v; big = true; break;
    case 5: src = c.Hvovv; n = v * o * v * v; big = true; break;
    case 6: src = c.Hooov; n = o * o * o * v; break;
    case 7: src = c.Hovvo; n = o * v * v * o; break;
    case 8: src = c.Hovov; n = o * v * o * v; break;
    case 9: src = c.Hvvvo; n = v * v * v * o; big = true; break;
    case 10: src = c.Hovoo; n = o * v * o * o; break;
    default: CCX_CHECK(false, ERR_ARG, "HBAR block index out of range");
  }
  CCX_CHECK(!big || c.hbar_stored, ERR_STATE, "Hvvvv, Hvovv and Hvvvo are held by a stored HBAR only");
  CCX_HIP(hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}

// r_L1 / r_L2 (cclambda.py:345-369, :468-496), r2 written as alpha * (H + H^T) (:496). With
// with_source = false the terms that do not depend on l (2 Hov and L) are left out.
void lambda_residual(Ctx& c, const double* l1, const double* l2, double* r1, double* r2, bool with_source,
                     double alpha, hipStream_t s) {
  CCX_CHECK(c.have_hbar, ERR_STATE, "the Lambda residual needs an HBAR (ccx_hbar_build)");
  CCX_CHECK(l1 && l2 && r1 && r2, ERR_ARG, "null argument");
  const int64_t o = c.o, v = c.v, ov = o * v;
  Workspace& ws = c.ws;
  const std::vector<int64_t> OV{o, v}, OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OOOO{o, o, o, o}, OVVV{o, v, v, v};
  const std::vector<int64_t> VOVV{v, o, v, v}, OVVO{o, v, v, o}, VVVO{v, v, v, o}, OVOO{o, v, o, o};
  const std::vector<int64_t> OOVO{o, o, v, o};
  const bool x3 = c.mode == MODE_MIXED_F16X3;
  auto cd = [&](double a, const TV& A, const TV& B, double b, const TV& C) {
    contract<double>(a, A, B, b, C, ws, s, x3);
  };
  auto L1 = [&](const char* l) { return tv(l1, l, OV); };
  auto L2 = [&](const char* l) { return tv(l2, l, OOVV); };
  auto T1 = [&](const char* l) { return tv(c.Ht1, l, OV); };
  auto T2 = [&](const char* l) { return tv(c.Ht2, l, OOVV); };
  auto CVVV = [&](const char* l) { return tv(c.Covvv, l, OVVV); };
  auto COOV = [&](const char* l) { return tv(c.Cooov, l, OOOV); };
  auto COVV = [&](const char* l) { return tv(c.Coovv, l, OOVV); };
  auto COVO = [&](const char* l) { return tv(c.Covov, l, OVOV); };
  auto IOVV = [&](const char* l) { return tv(c.Ioovv, l, OOVV); };
  auto LOVV = [&](const char* l) { return tv(c.Loovv, l, OOVV); };

  // ---- Goo, Gvv: cclambda.py:281, :306 ----
  const TV Goo = tv(c.Goo, "mi", OO), Gvv = tv(c.Gvv, "ae", VV);
  cd(1.0, T2("mjab"), L2("ijab"), 0.0, Goo);
  cd(-1.0, T2("ijeb"), L2("ijab"), 0.0, Gvv);

  // ---- r_L1: :348-369 ----
  const TV R1 = tv(r1, "ia", OV);
  if (with_source) axpby<double>(2.0, tv(c.Hov, "ia", OV), 0.0, R1, s);          // :348
  else scale<double>(0.0, R1, s);
  cd(1.0, L1("ie"), tv(c.Hvv, "ea", VV), 1.0, R1);                               // :354
  cd(-1.0, L1("ma"), tv(c.Hoo, "im", OO), 1.0, R1);                              // :355
  cd(-1.0, L2("mnae"), tv(c.Hovoo, "iemn", OVOO), 1.0, R1);                      // :357
  cd(2.0, L1("me"), tv(c.Hovvo, "ieam", OVVO), 1.0, R1);                         // :358
  cd(-1.0, L1("me"), tv(c.Hovov, "iema", OVOV), 1.0, R1);
  cd(-2.0, relabel(Goo, "mn"), tv(c.Hooov, "mina", OOOV), 1.0, R1);              // :368
  cd(1.0, relabel(Goo, "mn"), tv(c.Hooov, "imna", OOOV), 1.0, R1);               // :369

  // ---- r_L2 before P(ij,ab): :468-494 ----
  const TV H = tv(c.Lh, "ijab", OOVV);
  if (with_source) axpby<double>(1.0, LOVV("ijab"), 0.0, H, s);                  // :468
  else scale<double>(0.0, H, s);
  cd(2.0, L1("ia"), tv(c.Hov, "jb", OV), 1.0, H);                                // :476
  cd(-1.0, L1("ja"), tv(c.Hov, "ib", OV), 1.0, H);                               // :477
  cd(-2.0, L1("mb"), tv(c.Hooov, "jima", OOOV), 1.0, H);                         // :480
  cd(1.0, L1("mb"), tv(c.Hooov, "ijma", OOOV), 1.0, H);                          // :481
  cd(1.0, L2("ijeb"), tv(c.Hvv, "ea", VV), 1.0, H);                              // :486
  cd(-1.0, L2("mjab"), tv(c.Hoo, "im", OO), 1.0, H);                             // :487
  cd(0.5, L2("mnab"), tv(c.Hoooo, "ijmn", OOOO), 1.0, H);                        // :488
  cd(2.0, L2("mjeb"), tv(c.Hovvo, "ieam", OVVO), 1.0, H);                        // :490
  cd(-1.0, L2("mjeb"), tv(c.Hovov, "iema", OVOV), 1.0, H);
  cd(-1.0, L2("mibe"), tv(c.Hovov, "jema", OVOV), 1.0, H);                       // :491
  cd(-1.0, L2("mieb"), tv(c.Hovvo, "jeam", OVVO), 1.0, H);                       // :492
  cd(1.0, Gvv, LOVV("ijeb"), 1.0, H);                                            // :493
  cd(-1.0, Goo, LOVV("mjab"), 1.0, H);                                           // :494

  if (c.hbar_stored) {
    cd(1.0, L2("imef"), tv(c.Hvvvo, "efam", VVVO), 1.0, R1);                     // :356
    cd(-2.0, relabel(Gvv, "ef"), tv(c.Hvovv, "eifa", VOVV), 1.0, R1);            // :366
    cd(1.0, relabel(Gvv, "ef"), tv(c.Hvovv, "eiaf", VOVV), 1.0, R1);             // :367
    cd(2.0, L1("ie"), tv(c.Hvovv, "ejab", VOVV), 1.0, H);                        // :478
    cd(-1.0, L1("ie"), tv(c.Hvovv, "ejba", VOVV), 1.0, H);                       // :479
    cd(0.5, L2("ijef"), tv(c.Hvvvv, "efab", {v, v, v, v}), 1.0, H);              // :489
  } else {
    // Factored form of the terms that read Hvvvv, Hvovv and Hvvvo, with the intermediates
    //   Vf_ijem = l2_ijef t_mf,  Vg_ijfm = l2_ijef t_me,  U_ijmn = l2_ijef tau_mnef,
    //   Y_ijab  = l2_ijef Hvvvv_efab
    //           = ladder(l2)_ijab - Vf_ijem <em|ab> - Vg_ijfm <fm|ba> + U_ijmn <mn|ab>   (:395-401)
    build_tau(o, v, 1.0, 1.0, c.Ht1, c.Ht2, c.tau, s);
    auto TAU = [&](const char* l) { return tv(c.tau, l, OOVV); };
    const TV Vf = tv(c.LVf, "ijem", OOVO), Vg = tv(c.LVg, "ijfm", OOVO), U = tv(c.LU, "ijmn", OOOO);
    const TV Y = tv(c.LY, "ijab", OOVV);
    cd(1.0, L2("ijef"), T1("mf"), 0.0, Vf);
    cd(1.0, L2("ijef"), T1("me"), 0.0, Vg);
    cd(1.0, L2("ijef"), TAU("mnef"), 0.0, U);
    {
      // the packed ladder GEMM of the forward residual applied to l2, (ab) columns only
      const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
      dim3 grid((unsigned)std::min<int64_t>(ceil_div(v * v, 256), 64), (unsigned)(o * o));
      if (x3) {   // FP16x3 against the context's split ladder rows, FP32 output
        ladder_x3_l2(c, l2, s);
        ladder_unpack_kernel<float><<<grid, 256, 0, s>>>(o, v, (const float*)c.CP, (const float*)c.CM, c.LY);
      } else {
        tau_pack<double>(o, v, nullptr, l2, (double*)c.Tp, (double*)c.Tm, s);
        cd(1.0, tv(c.Tp, "xk", {noP, npP}), tvs(c.LP, "ky", {npP, npP}, {npP + ov, 1}), 0.0,
           tvs(c.CP, "xy", {noP, npP}, {npP + ov, 1}));
        if (noM > 0)
          cd(1.0, tv(c.Tm, "xk", {noM, npM}), tvs(c.LM, "ky", {npM, npM}, {npM + ov, 1}), 0.0,
             tvs(c.CM, "xy", {noM, npM}, {npM + ov, 1}));
        ladder_unpack_kernel<double><<<grid, 256, 0, s>>>(o, v, (const double*)c.CP, (const double*)c.CM, c.LY);
      }
      CCX_HIP(hipGetLastError());
    }
    cd(-1.0, Vf, CVVV("mbea"), 1.0, Y);                        // <em|ab> = (mb|ea)
    cd(-1.0, Vg, CVVV("mafb"), 1.0, Y);                        // <fm|ba> = (ma|fb)
    cd(1.0, U, IOVV("mnab"), 1.0, Y);                          // <mn|ab>
    axpby<double>(0.5, Y, 1.0, H, s);                                            // :489

    // :356  l2_imef Hvvvo_efam, term by term over :682-695 (p, q occupied and g virtual dummies)
    cd(1.0, L2("imef"), CVVV("mfea"), 1.0, R1);                // :682 <ef|am> = (mf|ea)
    const TV Xoo = tv(c.Lsm, "ip", OO);                        // :683 -Hov_pa l2_imef t_pmef
    cd(1.0, L2("imef"), T2("pmef"), 0.0, Xoo);
    cd(-1.0, Xoo, tv(c.Hov, "pa", OV), 1.0, R1);
    cd(1.0, T1("mg"), relabel(Y, "imag"), 1.0, R1);            // :684 t_mg Hvvvv_efag -> t_mg Y_imag
    cd(1.0, relabel(U, "impq"), COOV("qmpa"), 1.0, R1);        // :685 <pq|am> = (qm|pa)
    const TV W = tv(c.Ltmp, "ifpg", OVOV);                     // :686 -t_mpge <fp|ga>, (pa|fg)
    cd(1.0, L2("imef"), T2("mpge"), 0.0, W);
    cd(-1.0, W, CVVV("pafg"), 1.0, R1);
    const TV W6 = tv(c.Ltmp, "iepg", OVOV);                    // :687 -t_mpgf <ep|ag>, (pg|ea)
    cd(1.0, L2("imef"), T2("mpgf"), 0.0, W6);
    cd(-1.0, W6, CVVV("pgea"), 1.0, R1);
    cd(1.0, L2("imef"), T2("pmgf"), 0.0, W6);                  // :688 t_pmgf L_epag
    cd(2.0, W6, CVVV("pgea"), 1.0, R1);                        //   L_epag = 2(pg|ea) - (pa|eg)
    cd(-1.0, W6, CVVV("paeg"), 1.0, R1);
    // :689-691  -t_pf A_epam,  A_epam = <ep|am> - t_mqge <pq|ga>;  Vf_imep = l2_imef t_pf
    cd(-1.0, relabel(Vf, "imep"), COVV("pmea"), 1.0, R1);      //   <ep|am> = (pm|ea)
    const TV Q = tv(c.LQ, "ipqg", OOOV);
    cd(1.0, relabel(Vf, "imep"), T2("mqge"), 0.0, Q);
    cd(1.0, Q, IOVV("pqga"), 1.0, R1);
    // :692-695  -t_pe B_fpma,  B_fpma = <fp|ma> - t_mqgf <pq|ag> + t_qmgf L_pqag;  Vg_imfp = l2_imef t_pe
    cd(-1.0, relabel(Vg, "imfp"), COVO("mfpa"), 1.0, R1);      //   <fp|ma> = (mf|pa)
    cd(1.0, relabel(Vg, "imfp"), T2("mqgf"), 0.0, Q);
    cd(1.0, Q, IOVV("pqag"), 1.0, R1);
    cd(1.0, relabel(Vg, "imfp"), T2("qmgf"), 0.0, Q);
    cd(-1.0, Q, LOVV("pqag"), 1.0, R1);

    // :366-367  Gvv_ef Hvovv, Hvovv_eifa = <ei|fa> - t_ne <ni|fa>
    const TV Gef = relabel(Gvv, "ef");
    cd(-2.0, Gef, CVVV("iaef"), 1.0, R1);                      // <ei|fa> = (ia|ef)
    cd(1.0, Gef, CVVV("ifea"), 1.0, R1);                       // <ei|af> = (if|ea)
    const TV gt = tv(c.Lsm + o * o, "nf", OV);                 // gt_nf = Gvv_ef t_ne
    cd(1.0, Gef, T1("ne"), 0.0, gt);
    cd(2.0, gt, IOVV("nifa"), 1.0, R1);
    cd(-1.0, gt, IOVV("niaf"), 1.0, R1);

    // :478-479  l1_ie Hvovv_ejab, Hvovv_ejab = <ej|ab> - t_ne <nj|ab>
    cd(2.0, L1("ie"), CVVV("jbea"), 1.0, H);                   // <ej|ab> = (jb|ea)
    cd(-1.0, L1("ie"), CVVV("jaeb"), 1.0, H);                  // <ej|ba> = (ja|eb)
    const TV lt = tv(c.Lsm + o * o + ov, "in", OO);            // lt_in = l1_ie t_ne
    cd(1.0, L1("ie"), T1("ne"), 0.0, lt);
    cd(-2.0, lt, IOVV("njab"), 1.0, H);
    cd(1.0, lt, IOVV("njba"), 1.0, H);
  }

  pair_symmetrize(o, v, alpha, c.Lh, r2, s);                                     // :496
}

// The CCSD Jacobian applied to (x1, x2) on the current HBAR (CCderiv._ccsd_jacobian,
// ccderiv.py:149-180): j1 = sigma1, j2 = sigma2 + sigma2^T, the derivative of the residual along x.
// x2 must be pair-symmetric.
void jacobian(Ctx& c, const double* x1, const double* x2, double* j1, double* j2, hipStream_t s) {
  CCX_CHECK(c.have_hbar, ERR_STATE, "the Jacobian needs an HBAR (ccx_hbar_build)");
  CCX_CHECK(c.mode == MODE_DP, ERR_STATE, "the Jacobian runs in a DP context");
  CCX_CHECK(x1 && x2 && j1 && j2, ERR_ARG, "null argument");
  const int64_t o = c.o, v = c.v, ov = o * v;
  Workspace& ws = c.ws;
  const std::vector<int64_t> OV{o, v}, OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OOOO{o, o, o, o}, OVVV{o, v, v, v};
  const std::vector<int64_t> VOVV{v, o, v, v}, OVVO{o, v, v, o}, VVVO{v, v, v, o}, OVOO{o, v, o, o};
  auto cd = [&](double a, const TV& A, const TV& B, double b, const TV& C) {
    contract<double>(a, A, B, b, C, ws, s);
  };
  auto X1 = [&](const char* l) { return tv(x1, l, OV); };
  auto X2 = [&](const char* l) { return tv(x2, l, OOVV); };
  auto T1 = [&](const char* l) { return tv(c.Ht1, l, OV); };
  auto T2 = [&](const char* l) { return tv(c.Ht2, l, OOVV); };
  auto CVVV = [&](const char* l) { return tv(c.Covvv, l, OVVV); };
  auto COOV = [&](const char* l) { return tv(c.Cooov, l, OOOV); };
  auto COVV = [&](const char* l) { return tv(c.Coovv, l, OOVV); };
  auto COVO = [&](const char* l) { return tv(c.Covov, l, OVOV); };
  auto IOVV = [&](const char* l) { return tv(c.Ioovv, l, OOVV); };
  auto LOVV = [&](const char* l) { return tv(c.Loovv, l, OOVV); };
  auto HOOV = [&](const char* l) { return tv(c.Hooov, l, OOOV); };
  const TV Hov = tv(c.Hov, "me", OV);
  double* sm = c.Jsm;
  const TV Zvv = tv(sm, "ae", VV), Zoo = tv(sm + v * v, "mi", OO);

  // ---- sigma1: ccderiv.py:165-169 ----
  const TV R1 = tv(j1, "ia", OV);
  cd(1.0, X1("ie"), tv(c.Hvv, "ae", VV), 0.0, R1);                                    // :165
  cd(-1.0, X1("ma"), tv(c.Hoo, "mi", OO), 1.0, R1);
  cd(2.0, X1("me"), tv(c.Hovvo, "maei", OVVO), 1.0, R1);                              // :166
  cd(-1.0, X1("me"), tv(c.Hovov, "maie", OVOV), 1.0, R1);
  cd(2.0, Hov, X2("miea"), 1.0, R1);                                                  // :167
  cd(-1.0, Hov, X2("imea"), 1.0, R1);
  cd(-2.0, X2("mnae"), HOOV("mnie"), 1.0, R1);                                        // :169
  cd(1.0, X2("mnae"), HOOV("nmie"), 1.0, R1);
  // Zvv, Zoo: :170-171 (the Hvvmvo parts below, by HBAR form)
  cd(-1.0, LOVV("mnef"), X2("mnaf"), 0.0, Zvv);
  cd(-2.0, HOOV("mnie"), X1("ne"), 0.0, Zoo);
  cd(1.0, HOOV("nmie"), X1("ne"), 1.0, Zoo);
  cd(-1.0, LOVV("mnef"), X2("inef"), 1.0, Zoo);

  // ---- sigma2 before the final P(ij,ab): ccderiv.py:172-177 ----
  const TV H = tv(c.Lh, "ijab", OOVV);
  cd(-1.0, X1("ma"), tv(c.Hovoo, "mbij", OVOO), 0.0, H);                              // :172
  cd(1.0, X2("ijeb"), tv(c.Hvv, "ae", VV), 1.0, H);                                   // :174
  cd(-1.0, X2("mjab"), tv(c.Hoo, "mi", OO), 1.0, H);
  cd(0.5, X2("mnab"), tv(c.Hoooo, "mnij", OOOO), 1.0, H);                             // :175
  cd(-1.0, X2("imeb"), tv(c.Hovov, "maje", OVOV), 1.0, H);                            // :176
  cd(-1.0, X2("imea"), tv(c.Hovvo, "mbej", OVVO), 1.0, H);
  cd(2.0, X2("miea"), tv(c.Hovvo, "mbej", OVVO), 1.0, H);                             // :177
  cd(-1.0, X2("miea"), tv(c.Hovov, "mbje", OVOV), 1.0, H);

  if (c.hbar_stored) {
    // Hvvmvo = 2 Hvovv - Hvovv^(ef) (:163)
    cd(2.0, X2("imef"), tv(c.Hvovv, "amef", VOVV), 1.0, R1);                          // :168
    cd(-1.0, X2("imef"), tv(c.Hvovv, "amfe", VOVV), 1.0, R1);
    cd(2.0, tv(c.Hvovv, "amef", VOVV), X1("mf"), 1.0, Zvv);                           // :170
    cd(-1.0, tv(c.Hvovv, "amfe", VOVV), X1("mf"), 1.0, Zvv);
    cd(1.0, X1("ie"), tv(c.Hvvvo, "abej", VVVO), 1.0, H);                             // :172
    cd(0.5, X2("ijef"), tv(c.Hvvvv, "abef", {v, v, v, v}), 1.0, H);                   // :175
  } else {
    // Hvvmvo_amef = L_amef - t_na L_nmef with L_amef = 2 (mf|ae) - (me|af)
    const TV Win = tv(sm + v * v + o * o, "in", OO), q = tv(sm + v * v + 2 * o * o, "ne", OV);
    cd(2.0, X2("imef"), CVVV("mfae"), 1.0, R1);                                       // :168
    cd(-1.0, X2("imef"), CVVV("meaf"), 1.0, R1);
    cd(1.0, X2("imef"), LOVV("nmef"), 0.0, Win);
    cd(-1.0, T1("na"), Win, 1.0, R1);
    cd(2.0, CVVV("mfae"), X1("mf"), 1.0, Zvv);                                        // :170
    cd(-1.0, CVVV("meaf"), X1("mf"), 1.0, Zvv);
    cd(1.0, X1("mf"), LOVV("nmef"), 0.0, q);
    cd(-1.0, T1("na"), q, 1.0, Zvv);
    // x_ie Hvvvo_abej over cchbar.py:682-695 (its Hvvvv part joins the x2 ladder below)
    build_tau(o, v, 1.0, 1.0, c.Ht1, c.Ht2, c.tau, s);
    auto TAU = [&](const char* l) { return tv(c.tau, l, OOVV); };
    cd(1.0, X1("ie"), CVVV("jbae"), 1.0, H);                                          // <ab|ej> = (jb|ae)
    const TV u = tv(sm + v * v + 3 * o * o, "im", OO);                                // -Hov_me t_mjab
    cd(1.0, X1("ie"), Hov, 0.0, u);
    cd(-1.0, u, T2("mjab"), 1.0, H);
    const TV V = tv(c.LU, "mnij", OOOO);                                              // tau_mnab <mn|ej>
    cd(1.0, X1("ie"), COOV("njme"), 0.0, V);
    cd(1.0, TAU("mnab"), V, 1.0, H);
    const TV G = tv(c.Ltmp, "imbf", OOVV);                                            // -t_jmfa <bm|fe>
    cd(1.0, X1("ie"), CVVV("mebf"), 0.0, G);
    cd(-1.0, T2("jmfa"), G, 1.0, H);
    const TV G2 = tv(c.Ltmp, "imfa", OOVV), G4 = tv(c.LY, "imfa", OOVV);
    cd(1.0, X1("ie"), CVVV("mfae"), 0.0, G2);                                         // x_ie <am|ef>
    cd(-1.0, T2("jmfb"), G2, 1.0, H);                                                 // -t_jmfb <am|ef>
    cd(2.0, T2("mjfb"), G2, 1.0, H);                                                  // +t_mjfb L_amef
    cd(1.0, X1("ie"), CVVV("meaf"), 0.0, G4);
    cd(-1.0, T2("mjfb"), G4, 1.0, H);
    const TV K1 = tv(c.LVf, "imja", OOOV), K2 = tv(c.LVg, "imnf", OOOV), K3 = tv(c.LQ, "ijma", OOOV);
    cd(1.0, X1("ie"), COVV("mjae"), 0.0, K1);                                         // -t_mb A_amej
    cd(-1.0, T1("mb"), K1, 1.0, H);
    cd(1.0, X1("ie"), IOVV("mnfe"), 0.0, K2);
    cd(1.0, T2("jnfa"), K2, 0.0, K3);
    cd(1.0, T1("mb"), K3, 1.0, H);
    const TV K4 = tv(c.LVf, "imjb", OOOV), K5 = tv(c.LVg, "imnf", OOOV), K6 = tv(c.LQ, "ijmb", OOOV);
    cd(1.0, X1("ie"), COVO("mejb"), 0.0, K4);                                         // -t_ma B_bmje
    cd(-1.0, T1("ma"), K4, 1.0, H);
    cd(1.0, X1("ie"), IOVV("mnef"), 0.0, K5);
    cd(1.0, T2("jnfb"), K5, 0.0, K6);
    cd(1.0, T1("ma"), K6, 1.0, H);
    const TV K7 = tv(c.LVg, "imnf", OOOV), K8 = tv(c.LQ, "ijmb", OOOV);
    cd(1.0, X1("ie"), LOVV("mnef"), 0.0, K7);
    cd(1.0, T2("njfb"), K7, 0.0, K8);
    cd(-1.0, T1("ma"), K8, 1.0, H);
    // 1/2 x2 Hvvvv (:175) and x_ie t_jf Hvvvv_abef (:684 in :172) are, after P(ij,ab), the one
    // pair-symmetric contraction 1/2 Y(Z) with Z = x2 + x1 t1 + t1 x1 and Y(Z)_ijab = Z_ijef Hvvvv_abef
    //   = ladder(Z) - t_mb Z_ijef <am|ef> - t_ma Z_ijef <bm|fe> + tau_mnab Z_ijef <mn|ef>
    CCX_HIP(hipMemcpyAsync(c.LY, x2, ov * ov * sizeof(double), hipMemcpyDeviceToDevice, s));
    pair_outer_kernel<<<(unsigned)std::min<int64_t>(ceil_div(ov * ov, 256), 65536), 256, 0, s>>>(o, v, x1, c.Ht1,
                                                                                                 c.LY);
    CCX_HIP(hipGetLastError());
    const TV Ze = tv(c.LU, "ijmn", OOOO);
    cd(1.0, tv(c.LY, "ijef", OOVV), IOVV("mnef"), 0.0, Ze);
    const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
    tau_pack<double>(o, v, nullptr, c.LY, (double*)c.Tp, (double*)c.Tm, s);
    cd(1.0, tv(c.Tp, "xk", {noP, npP}), tv(c.LP, "ky", {npP, npP + ov}), 0.0, tv(c.CP, "xy", {noP, npP + ov}));
    if (noM > 0)
      cd(1.0, tv(c.Tm, "xk", {noM, npM}), tv(c.LM, "ky", {npM, npM + ov}), 0.0, tv(c.CM, "xy", {noM, npM + ov}));
    dim3 grid((unsigned)std::min<int64_t>(ceil_div(v * v, 256), 64), (unsigned)(o * o));
    ladder_unpack_kernel<double><<<grid, 256, 0, s>>>(o, v, (const double*)c.CP, (const double*)c.CM, c.LY);
    CCX_HIP(hipGetLastError());
    // Zf[m][i][j][b] = Z_ijef <mb|ef> from the (m, b) columns of the same GEMM
    const TV Zf = tv(c.LQ, "mijb", OOOV);
    scale<double>(0.0, Zf, s);
    zt_add<double>(o, v, (const double*)c.CP, (const double*)c.CM, c.LQ, s);
    const TV Y = tv(c.LY, "ijab", OOVV);
    cd(-1.0, T1("mb"), tv(c.LQ, "mjia", OOOV), 1.0, Y);                               // Z_ijef <am|ef> = Zf[m][j][i][a]
    cd(-1.0, T1("ma"), Zf, 1.0, Y);                                                   // Z_ijef <bm|fe> = Zf[m][i][j][b]
    cd(1.0, TAU("mnab"), Ze, 1.0, Y);
    axpby<double>(0.5, Y, 1.0, H, s);
  }
  cd(1.0, Zoo, T2("mjab"), 1.0, H);                                                   // :173
  cd(1.0, Zvv, T2("ijeb"), 1.0, H);
  pair_symmetrize(o, v, 1.0, c.Lh, j2, s);
}

// One- and two-particle densities of the CC Lagrangian E + l1.r1 + 1/2 l2.r2 at (t, l)
// (ccdensity.py, spatial CCSD builders :335-931): dL/dF and dL/d<pq|rs> at fixed t and l.
// D is written as the full [nmo][nmo] matrix (Doo, Dov, Dvo, Dvv blocks); Dvvvo and Dvvvv may be
// null (they are ov^3 and v^4).
void densities(Ctx& c, const double* t1, const double* t2, const double* l1, const double* l2, double* D,
               double* Doooo, double* Dooov, double* Dovov, double* Doovv, double* Dvvvo, double* Dvvvv,
               hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "densities requested before the integrals were set");
  CCX_CHECK(c.mode == MODE_DP, ERR_STATE, "the densities run in a DP context");
  CCX_CHECK(t1 && t2 && l1 && l2 && D && Doooo && Dooov && Dovov && Doovv, ERR_ARG, "null argument");
  alloc_work(c);
  alloc_hbar(c, false);
  const int64_t o = c.o, v = c.v, nmo = c.nmo;
  Workspace& ws = c.ws;
  const std::vector<int64_t> OV{o, v}, OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OOOO{o, o, o, o}, VVVO{v, v, v, o}, VOVO{v, o, v, o};
  auto cd = [&](double a, const TV& A, const TV& B, double b, const TV& C) {
    contract<double>(a, A, B, b, C, ws, s);
  };
  auto T1 = [&](const char* l) { return tv(t1, l, OV); };
  auto T2 = [&](const char* l) { return tv(t2, l, OOVV); };
  auto L1 = [&](const char* l) { return tv(l1, l, OV); };
  auto L2 = [&](const char* l) { return tv(l2, l, OOVV); };
  build_tau(o, v, 1.0, 1.0, t1, t2, c.tau, s);
  auto TAU = [&](const char* l) { return tv(c.tau, l, OOVV); };
  const TV Goo = tv(c.Goo, "mi", OO), Gvv = tv(c.Gvv, "ae", VV);
  cd(1.0, T2("mjab"), L2("ijab"), 0.0, Goo);                                        // cclambda.py:281
  cd(-1.0, T2("ijeb"), L2("ijab"), 0.0, Gvv);                                       // cclambda.py:306

  // ---- one-particle: :364-365, :401-402, :413, :454-461 ----
  const TV Doo = tvs(D, "ij", OO, {nmo, 1}), Dvv = tvs(D + o * nmo + o, "ab", VV, {nmo, 1});
  const TV Dov = tvs(D + o, "ia", OV, {nmo, 1}), Dvo = tvs(D + o * nmo, "ai", {v, o}, {nmo, 1});
  cd(-1.0, T1("ie"), L1("je"), 0.0, Doo);
  cd(-1.0, T2("imef"), L2("jmef"), 1.0, Doo);
  cd(1.0, T1("mb"), L1("ma"), 0.0, Dvv);
  cd(1.0, T2("mnbe"), L2("mnae"), 1.0, Dvv);
  axpby<double>(1.0, L1("ia"), 0.0, Dvo, s);
  axpby<double>(2.0, T1("ia"), 0.0, Dov, s);
  cd(2.0, L1("me"), T2("imae"), 1.0, Dov);
  cd(-1.0, L1("me"), TAU("miae"), 1.0, Dov);
  const TV soo = tv(c.Lsm, "mi", OO), svv = tv(c.LY, "ea", VV);
  cd(1.0, L2("mnef"), T2("inef"), 0.0, soo);
  cd(-1.0, soo, T1("ma"), 1.0, Dov);
  cd(1.0, L2("mnef"), T2("mnaf"), 0.0, svv);
  cd(-1.0, svv, T1("ie"), 1.0, Dov);

  // ---- Doooo :585, Dvvvv :609 ----
  cd(1.0, TAU("ijef"), L2("klef"), 0.0, tv(Doooo, "ijkl", OOOO));
  if (Dvvvv) cd(1.0, TAU("mnab"), L2("mncd"), 0.0, tv(Dvvvv, "abcd", {v, v, v, v}));

  // ---- Dooov :645-670 ----
  {
    const TV Dq = tv(Dooov, "ijka", OOOV);
    const TV tv4 = tv(c.Ltmp, "jake", OVOV), to4 = tv(c.LU, "ijkm", OOOO);
    cd(-2.0, L1("ke"), TAU("ijea"), 0.0, Dq);                                        // :646 (2 tau - tau^T)
    cd(1.0, L1("ke"), TAU("ijae"), 1.0, Dq);
    cd(-1.0, T1("ie"), L2("jkae"), 1.0, Dq);                                         // :647
    cd(-2.0, relabel(Goo, "ik"), T1("ja"), 1.0, Dq);                                 // :652
    cd(1.0, relabel(Goo, "jk"), T1("ia"), 1.0, Dq);                                  // :653
    cd(1.0, T2("jmaf"), L2("kmef"), 0.0, tv4);                                       // :654
    cd(-2.0, tv4, T1("ie"), 1.0, Dq);                                                // :655
    cd(1.0, relabel(tv4, "iake"), T1("je"), 1.0, Dq);                                // :656
    cd(1.0, T2("ijef"), L2("kmef"), 0.0, to4);                                       // :658
    cd(1.0, to4, T1("ma"), 1.0, Dq);                                                 // :659
    cd(1.0, T2("mjaf"), L2("kmef"), 0.0, tv4);                                       // :660
    cd(1.0, tv4, T1("ie"), 1.0, Dq);                                                 // :661
    const TV tk = tv(c.Ltmp, "iakf", OVOV);
    cd(1.0, T2("imea"), L2("kmef"), 0.0, tk);                                        // :662
    cd(1.0, tk, T1("jf"), 1.0, Dq);                                                  // :663
    const TV t3 = tv(c.LVf, "kmej", {o, o, v, o}), t4 = tv(c.LU, "kmij", OOOO);
    cd(1.0, L2("kmef"), T1("jf"), 0.0, t3);                                          // :668
    cd(1.0, t3, T1("ie"), 0.0, t4);                                                  // :669
    cd(1.0, t4, T1("ma"), 1.0, Dq);                                                  // :670
  }

  // ---- Dvvvo :716-741 (the v^4 intermediate of :729-730 taken as t_nmab (l_nmce t_ie)) ----
  if (Dvvvo) {
    const TV Dq = tv(Dvvvo, "abci", VVVO);
    cd(2.0, L1("mc"), TAU("miab"), 0.0, Dq);                                         // :717 (2 tau - tau^T)
    cd(-1.0, L1("mc"), TAU("miba"), 1.0, Dq);
    cd(1.0, T1("ma"), L2("imbc"), 1.0, Dq);                                          // :718
    cd(-2.0, relabel(Gvv, "ca"), T1("ib"), 1.0, Dq);                                 // :723
    cd(1.0, relabel(Gvv, "cb"), T1("ia"), 1.0, Dq);                                  // :724
    const TV tb = tv(c.Ltmp, "ibnc", OVOV);
    cd(1.0, T2("imbe"), L2("nmce"), 0.0, tb);                                        // :725
    cd(2.0, tb, T1("na"), 1.0, Dq);                                                  // :726
    cd(-1.0, relabel(tb, "ianc"), T1("nb"), 1.0, Dq);                                // :727
    const TV q = tv(c.LVf, "nmci", {o, o, v, o});
    cd(1.0, L2("nmce"), T1("ie"), 0.0, q);                                           // :729-730 as
    cd(-1.0, T2("nmab"), q, 1.0, Dq);                                                //   t_nmab (l_nmce t_ie)
    const TV ta = tv(c.Ltmp, "iamc", OVOV);
    cd(1.0, T2("niae"), L2("nmce"), 0.0, ta);                                        // :731
    cd(-1.0, ta, T1("mb"), 1.0, Dq);                                                 // :732
    cd(1.0, T2("mibe"), L2("nmce"), 0.0, tb);                                        // :733
    cd(-1.0, tb, T1("na"), 1.0, Dq);                                                 // :734
    const TV am = tv(c.Ltmp, "amci", VOVO);
    cd(1.0, q, T1("na"), 0.0, am);                                                   // :739-740
    cd(-1.0, am, T1("mb"), 1.0, Dq);                                                 // :741
  }

  // ---- Dovov :775-780 ----
  {
    const TV Dq = tv(Dovov, "iajb", OVOV);
    cd(-1.0, T1("ia"), L1("jb"), 0.0, Dq);
    cd(-1.0, TAU("mibe"), L2("jmea"), 1.0, Dq);
    cd(-1.0, T2("imbe"), L2("mjea"), 1.0, Dq);
  }

  // ---- Doovv :830-922 ----
  {
    const TV Dq = tv(Doovv, "ijab", OOVV);
    cd(4.0, T1("ia"), L1("jb"), 0.0, Dq);                                            // :860
    axpby<double>(4.0, TAU("ijab"), 1.0, Dq, s);                                     // :861  2 (2 tau - tau^T)
    axpby<double>(-2.0, TAU("ijba"), 1.0, Dq, s);
    axpby<double>(1.0, L2("ijab"), 1.0, Dq, s);                                      // :862
    const TV x2 = tv(c.Lsm, "jb", OV), x3 = tv(c.LVf, "ijmb", OOOV);
    cd(4.0, L1("me"), T2("jmbe"), 0.0, x2);                                          // :864-865  2 l_me (2 t - t^T)
    cd(-2.0, L1("me"), T2("jmeb"), 1.0, x2);
    cd(2.0, x2, T1("ia"), 1.0, Dq);                                                  // :866
    cd(-1.0, relabel(x2, "ja"), T1("ib"), 1.0, Dq);                                  // :867
    cd(4.0, T2("ijeb"), L1("me"), 0.0, x3);                                          // :868
    cd(-2.0, T2("ijbe"), L1("me"), 1.0, x3);
    cd(-1.0, x3, T1("ma"), 1.0, Dq);                                                 // :869
    // :870-871  -(2 l_me (2 tau_jmba - tau_jmab)) t_ie = -2 l_me t_ie (2 tau_jmba - tau_jmab)
    const TV li = tv(c.Lsm + o * v, "mi", OO);
    cd(1.0, L1("me"), T1("ie"), 0.0, li);
    cd(-4.0, li, TAU("jmba"), 1.0, Dq);
    cd(2.0, li, TAU("jmab"), 1.0, Dq);
    cd(4.0, T2("imae"), L2("mjeb"), 1.0, Dq);                                        // :876
    cd(-2.0, TAU("mjbe"), L2("imae"), 1.0, Dq);                                      // :877
    const TV too = tv(c.LU, "ijmn", OOOO);
    cd(1.0, T2("ijef"), L2("mnef"), 0.0, too);                                       // :879
    cd(1.0, too, T2("mnab"), 1.0, Dq);                                               // :880
    const TV ta = tv(c.Ltmp, "jbme", OVOV);
    cd(1.0, T2("njbf"), L2("mnef"), 0.0, ta);                                        // :881
    cd(1.0, ta, T2("miae"), 1.0, Dq);                                                // :882
    const TV tb = tv(c.Ltmp, "ibne", OVOV);
    cd(1.0, T2("imfb"), L2("mnef"), 0.0, tb);                                        // :883
    cd(1.0, tb, T2("njae"), 1.0, Dq);                                                // :884
    cd(4.0, relabel(Gvv, "eb"), TAU("ijae"), 1.0, Dq);                               // :886
    cd(-2.0, relabel(Gvv, "ea"), TAU("ijbe"), 1.0, Dq);                              // :887
    cd(-4.0, relabel(Goo, "jm"), TAU("imab"), 1.0, Dq);                              // :889
    cd(2.0, relabel(Goo, "jm"), TAU("imba"), 1.0, Dq);                               // :890
    const TV tc = tv(c.Ltmp, "iame", OVOV);
    cd(1.0, T2("inaf"), L2("mnef"), 0.0, tc);                                        // :891
    cd(-4.0, tc, TAU("mjbe"), 1.0, Dq);                                              // :892
    cd(2.0, relabel(tc, "ibme"), TAU("mjae"), 1.0, Dq);                              // :893
    cd(4.0, relabel(tc, "jbme"), T2("imae"), 1.0, Dq);                               // :894
    cd(-2.0, relabel(tc, "jame"), T2("imbe"), 1.0, Dq);                              // :895
    const TV tm = tv(c.LVf, "ijmb", OOOV);
    cd(1.0, T1("nb"), too, 0.0, tm);                                                 // :898
    cd(1.0, T1("ma"), tm, 1.0, Dq);                                                  // :899
    const TV y1 = tv(c.LVg, "mnif", OOOV), y2 = tv(c.LU, "mnij", OOOO);
    cd(1.0, T1("ie"), L2("mnef"), 0.0, y1);                                          // :900
    cd(1.0, T1("jf"), y1, 0.0, y2);                                                  // :901
    cd(1.0, y2, T2("mnab"), 1.0, Dq);                                                // :902
    const TV y3 = tv(c.LQ, "mijb", OOOV);
    cd(1.0, y1, T2("njbf"), 0.0, y3);                                                // :903-904
    cd(1.0, T1("ma"), y3, 1.0, Dq);                                                  // :905
    const TV y4 = tv(c.LVg, "mnej", {o, o, v, o}), y5 = tv(c.LQ, "njia", OOOV);
    cd(1.0, T1("jf"), L2("mnef"), 0.0, y4);                                          // :906
    cd(1.0, y4, T2("miae"), 0.0, y5);                                                // :907
    cd(1.0, T1("nb"), y5, 1.0, Dq);                                                  // :908
    const TV y6 = tv(c.LVg, "mnjf", OOOV), y7 = tv(c.LQ, "njib", OOOV);
    cd(1.0, T1("je"), L2("mnef"), 0.0, y6);                                          // :909
    cd(1.0, y6, T2("imfb"), 0.0, y7);                                                // :910
    cd(1.0, T1("na"), y7, 1.0, Dq);                                                  // :911
    const TV y8 = tv(c.LVg, "mnei", {o, o, v, o}), y9 = tv(c.LQ, "mija", OOOV);
    cd(1.0, T1("if"), L2("mnef"), 0.0, y8);                                          // :912
    cd(1.0, y8, T2("njae"), 0.0, y9);                                                // :913
    cd(1.0, T1("mb"), y9, 1.0, Dq);                                                  // :914
    const TV z1 = tv(c.LVg, "mnej", {o, o, v, o}), z2 = tv(c.LU, "mnij", OOOO);
    const TV z3 = tv(c.LQ, "mbij", {o, v, o, o});
    cd(1.0, T1("jf"), L2("mnef"), 0.0, z1);                                          // :919
    cd(1.0, T1("ie"), z1, 0.0, z2);                                                  // :920
    cd(1.0, T1("nb"), z2, 0.0, z3);                                                  // :921
    cd(1.0, T1("ma"), z3, 1.0, Dq);                                                  // :922
  }
}

// compute_energy (ccdensity.py:113-211): F_ij D_ij + F_ab D_ab + 1/2 <ij|kl> Doooo + <ij|ka> Dooov
// + <ab|ci> Dvvvo + <ia|jb> Dovov + 1/2 <ij|ab> Doovv + 1/2 <ab|cd> Dvvvv, the last one taken as
// 1/2 tau_mnab (l2_mncd <ab|cd>) through the packed ladder GEMM (no v^4 array). e_dev: device scalar.
void density_energy(Ctx& c, const double* t1, const double* t2, const double* l2, const double* F,
                    const double* D, const double* Doooo, const double* Dooov, const double* Dovov,
                    const double* Doovv, const double* Dvvvo, double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints && c.mode == MODE_DP, ERR_STATE, "density energy needs a DP context with integrals");
  CCX_CHECK(t1 && t2 && l2 && F && D && Doooo && Dooov && Dovov && Doovv && Dvvvo && e_dev, ERR_ARG,
            "null argument");
  alloc_work(c);
  alloc_hbar(c, false);
  const int64_t o = c.o, v = c.v, nmo = c.nmo, ov = o * v;
  const int64_t slab = std::max<int64_t>(v * v * v, o * o + v * v);   // one v^3 slice, or the Fock blocks
  if (!c.Jv3) c.Jv3 = c.alloc<double>(2 * slab + o + 16);
  Workspace& ws = c.ws;
  const std::vector<int64_t> OO{o, o}, VV{v, v}, OOVV{o, o, v, v}, OVOV{o, v, o, v};
  const std::vector<int64_t> OOOV{o, o, o, v}, OOOO{o, o, o, o}, VVV{v, v, v};
  double* part = c.Jv3 + 2 * slab;        // one value per term (o + 7 of them)
  double* buf = c.Ltmp;                   // the o^2 v^2 integral block permuted to the density layout
  int nt = 0;
  auto dot = [&](int64_t n, const double* x, const double* y, double w) {
    dots(n, 1, x, 0, y, c.red, part + nt, s);
    if (w != 1.0) scale<double>(w, tv(part + nt, "x", {1}), s);
    ++nt;
  };
  // Fock terms (blocks copied contiguous)
  axpby<double>(1.0, tvs(F, "ij", OO, {nmo, 1}), 0.0, tv(c.Jv3, "ij", OO), s);
  axpby<double>(1.0, tvs(D, "ij", OO, {nmo, 1}), 0.0, tv(c.Jv3 + slab, "ij", OO), s);
  dot(o * o, c.Jv3, c.Jv3 + slab, 1.0);
  axpby<double>(1.0, tvs(F + o * nmo + o, "ab", VV, {nmo, 1}), 0.0, tv(c.Jv3, "ab", VV), s);
  axpby<double>(1.0, tvs(D + o * nmo + o, "ab", VV, {nmo, 1}), 0.0, tv(c.Jv3 + slab, "ab", VV), s);
  dot(v * v, c.Jv3, c.Jv3 + slab, 1.0);
  axpby<double>(1.0, tv(c.Coooo, "ikjl", OOOO), 0.0, tv(c.LU, "ijkl", OOOO), s);  // <ij|kl> = (ik|jl)
  dot(o * o * o * o, c.LU, Doooo, 0.5);
  axpby<double>(1.0, tv(c.Cooov, "ikja", OOOV), 0.0, tv(c.LVf, "ijka", OOOV), s); // <ij|ka> = (ik|ja)
  dot(o * o * ov, c.LVf, Dooov, 1.0);
  axpby<double>(1.0, tv(c.Coovv, "ijab", OOVV), 0.0, tv(buf, "iajb", OVOV), s);   // <ia|jb> = (ij|ab)
  dot(ov * ov, buf, Dovov, 1.0);
  dot(ov * ov, c.Ioovv, Doovv, 0.5);                                              // <ij|ab>
  // <ab|ci> = (ac|bi) = (ib|ac) = Covvv[i][b][a][c], one i at a time against Dvvvo[a][b][c][i]
  for (int64_t i = 0; i < o; ++i) {
    axpby<double>(1.0, tv(c.Covvv + i * v * v * v, "bac", VVV), 0.0, tv(c.Jv3, "abc", VVV), s);
    axpby<double>(1.0, tvs(Dvvvo + i, "abc", VVV, {v * v * o, v * o, o}), 0.0, tv(c.Jv3 + slab, "abc", VVV), s);
    dot(v * v * v, c.Jv3, c.Jv3 + slab, 1.0);
  }
  {   // 1/2 <ab|cd> Dvvvv_abcd = 1/2 tau_mnab ladder(l2)_mnab
    const int64_t npP = c.npP(), npM = c.npM(), noP = c.noP(), noM = c.noM();
    tau_pack<double>(o, v, nullptr, l2, (double*)c.Tp, (double*)c.Tm, s);
    contract<double>(1.0, tv(c.Tp, "xk", {noP, npP}), tvs(c.LP, "ky", {npP, npP}, {npP + ov, 1}), 0.0,
                     tvs(c.CP, "xy", {noP, npP}, {npP + ov, 1}), ws, s);
    if (noM > 0)
      contract<double>(1.0, tv(c.Tm, "xk", {noM, npM}), tvs(c.LM, "ky", {npM, npM}, {npM + ov, 1}), 0.0,
                       tvs(c.CM, "xy", {noM, npM}, {npM + ov, 1}), ws, s);
    dim3 grid((unsigned)std::min<int64_t>(ceil_div(v * v, 256), 64), (unsigned)(o * o));
    ladder_unpack_kernel<double><<<grid, 256, 0, s>>>(o, v, (const double*)c.CP, (const double*)c.CM, c.LY);
    CCX_HIP(hipGetLastError());
    build_tau(o, v, 1.0, 1.0, t1, t2, c.tau, s);
    dot(ov * ov, c.tau, c.LY, 0.5);
  }
  sum_terms_kernel<<<1, 256, 0, s>>>(nt, part, e_dev);
  CCX_HIP(hipGetLastError());
}

// 1/2 <ij|ab> l2_ijab (cclambda.py:570) into a device scalar.
void pseudoenergy(Ctx& c, const double* l2, double* e_dev, hipStream_t s) {
  CCX_CHECK(c.have_ints, ERR_STATE, "pseudoenergy requested before the integrals were set");
  CCX_CHECK(l2 && e_dev, ERR_ARG, "null argument");
  dots(c.n2(), 1, c.Ioovv, 0, l2, c.red, e_dev, s);
  scale<double>(0.5, tv(e_dev, "x", {1}), s);
}

}  // namespace ccx

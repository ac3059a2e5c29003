// C ABI of the CCSD device library (declared in include/ccx.h).
#include "../../include/ccx.h"
#include "ccsd.h"
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

struct ccx_ctx {
  ccx::Ctx c;
};

// Scratch of ccx_contract_noalloc (split-K partial sums), CCX_NOALLOC_SCRATCH_BYTES per GPU.
static __device__ __align__(256) double scratch[CCX_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

thread_local std::string g_err;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return CCX_OK;
  } catch (const ccx::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CCX_ERR_INTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return CCX_ERR_INTERNAL;
  }
}

inline hipStream_t st(void* s) { return (hipStream_t)s; }

// Parse "ab,bc->ac" into three label strings.
void parse_subscripts(const char* sub, std::string& a, std::string& b, std::string& c) {
  CCX_CHECK(sub != nullptr, ccx::ERR_ARG, "null subscripts");
  std::string s(sub);
  s.erase(std::remove(s.begin(), s.end(), ' '), s.end());
  size_t comma = s.find(','), arrow = s.find("->");
  CCX_CHECK(comma != std::string::npos && arrow != std::string::npos && comma < arrow, ccx::ERR_ARG,
            "subscripts must look like 'ab,bc->ac'");
  a = s.substr(0, comma);
  b = s.substr(comma + 1, arrow - comma - 1);
  c = s.substr(arrow + 2);
  CCX_CHECK(!a.empty() && !b.empty() && a.size() <= 8 && b.size() <= 8 && c.size() <= 8, ccx::ERR_ARG,
            "operand rank must be 1..8");
}

ccx::TV make_view(const void* p, const std::string& lab, const int64_t* shape, const int64_t* strides) {
  CCX_CHECK(p && (lab.empty() || (shape && strides)), ccx::ERR_ARG, "null tensor pointer/shape");
  std::vector<int64_t> e(lab.size()), s(lab.size());
  for (size_t i = 0; i < lab.size(); ++i) {
    CCX_CHECK(shape[i] >= 0, ccx::ERR_ARG, "negative extent");
    e[i] = shape[i];
    s[i] = strides[i];
  }
  return ccx::tvs(p, lab.c_str(), e, s);
}

template <typename T>
int contract_any(const char* sub, T alpha, const T* A, const int64_t* as, const int64_t* ast, const T* B,
                 const int64_t* bs, const int64_t* bst, T beta, T* C, const int64_t* cs, const int64_t* cst,
                 void* stream, ccx::Workspace* fixed_ws = nullptr) {
  static thread_local ccx::Workspace ws;
  return guarded([&] {
    std::string la, lb, lc;
    parse_subscripts(sub, la, lb, lc);
    ccx::contract<T>(alpha, make_view(A, la, as, ast), make_view(B, lb, bs, bst), beta,
                     make_view(C, lc, cs, cst), fixed_ws ? *fixed_ws : ws, st(stream));
  });
}

// ---- the allocation-free contraction's scratch ----------------------------------------------------
// Split-K partial sums of ccx_contract_noalloc live in the file-scope device array `scratch` (one copy
// per GPU, held until the process ends). Host code never passes `scratch` itself to a kernel: on the
// host the name is only a stand-in symbol. Its device address is looked up on the first call on each
// device, with that device current, and kept in scratch_dev[]; kernels get that address.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[NOALLOC_MAX_DEVICES] = {};

// Device address of the current device's scratch (looked up on the first call on that device).
void* noalloc_scratch() {
  int dev = -1;
  CCX_HIP(hipGetDevice(&dev));
  CCX_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, ccx::ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    CCX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CCX_CHECK(p != nullptr, ccx::ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_dev[dev] = p;
  }
  return scratch_dev[dev];
}

}  // namespace

extern "C" {

int ccx_create(int no, int nv, int mode, ccx_ctx** out) {
  return guarded([&] {
    CCX_CHECK(out != nullptr, ccx::ERR_ARG, "null output handle");
    CCX_CHECK(no >= 1 && nv >= 1, ccx::ERR_ARG, "need no >= 1 and nv >= 1");
    CCX_CHECK(mode == CCX_MODE_DP || mode == CCX_MODE_MIXED || mode == CCX_MODE_MIXED_BF16X3 ||
                  mode == CCX_MODE_MIXED_F16X3,
              ccx::ERR_ARG, "mode must be CCX_MODE_DP, CCX_MODE_MIXED, CCX_MODE_MIXED_BF16X3 or CCX_MODE_MIXED_F16X3");
    int ndev = 0;
    CCX_HIP(hipGetDeviceCount(&ndev));
    CCX_CHECK(ndev > 0, ccx::ERR_HIP, "no HIP device");
    auto* h = new ccx_ctx;
    h->c.o = no;
    h->c.v = nv;
    h->c.nmo = no + nv;
    h->c.mode = mode;
    *out = h;
  });
}

int ccx_destroy(ccx_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipDeviceSynchronize();
    ctx->c.destroy();
    delete ctx;
  });
}

int ccx_set_integrals_dense(ccx_ctx* ctx, const double* fock, const double* eri, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::set_integrals_dense(ctx->c, fock, eri, st(stream));
    CCX_HIP(hipStreamSynchronize(st(stream)));
  });
}

int ccx_set_integrals_df(ccx_ctx* ctx, const double* fock, const double* B, int naux, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::set_integrals_df(ctx->c, fock, B, naux, st(stream));
    CCX_HIP(hipStreamSynchronize(st(stream)));
  });
}

int ccx_residual(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* r1,
                 double* r2, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::residual(ctx->c, fock, t1, t2, r1, r2, st(stream));
  });
}

int ccx_energy(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* e_dev,
               void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && fock && t1 && t2 && e_dev, ccx::ERR_ARG, "null argument");
    auto& c = ctx->c;
    CCX_CHECK(c.have_ints, ccx::ERR_STATE, "ccx_energy called before the integrals were set");
    ccx::energy(c.o, c.v, fock, c.nmo, t1, t2, c.Loovv, c.red, e_dev, st(stream));
  });
}

int ccx_update(ccx_ctx* ctx, double* t1, double* t2, const double* r1, const double* r2, double* rms2_dev,
               void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && t1 && t2 && r1 && r2 && rms2_dev, ccx::ERR_ARG, "null argument");
    auto& c = ctx->c;
    CCX_CHECK(c.have_ints, ccx::ERR_STATE, "ccx_update called before the integrals were set");
    ccx::jacobi_update(c.o, c.v, c.eps, t1, t2, r1, r2, c.red + 1024 * 32, rms2_dev, st(stream));
  });
}

int ccx_diis_reset(ccx_ctx* ctx, const double* t1, const double* t2, int max_diis, int capacity,
                   void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && t1 && t2, ccx::ERR_ARG, "null argument");
    ccx::diis_reset(ctx->c, t1, t2, max_diis, capacity, st(stream));
  });
}

int ccx_diis_add(ccx_ctx* ctx, const double* t1, const double* t2, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && t1 && t2, ccx::ERR_ARG, "null argument");
    ccx::diis_add(ctx->c, t1, t2, st(stream));
  });
}

int ccx_diis_extrapolate(ccx_ctx* ctx, double* t1, double* t2, double* coef_out, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && t1 && t2, ccx::ERR_ARG, "null argument");
    ccx::diis_extrapolate(ctx->c, t1, t2, coef_out, st(stream));
  });
}

int ccx_hbar_build(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, int flags,
                   void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    CCX_CHECK((flags & ~CCX_HBAR_STORED) == 0, ccx::ERR_ARG, "unknown ccx_hbar_build flag");
    ccx::hbar_build(ctx->c, fock, t1, t2, (flags & CCX_HBAR_STORED) != 0, st(stream));
  });
}

int ccx_hbar_block(ccx_ctx* ctx, int block, double* out, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::hbar_block(ctx->c, block, out, st(stream));
  });
}

int ccx_lambda_residual(ccx_ctx* ctx, const double* l1, const double* l2, double* r1, double* r2,
                        int with_source, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::lambda_residual(ctx->c, l1, l2, r1, r2, with_source != 0, 1.0, st(stream));
  });
}

int ccx_pseudoenergy(ccx_ctx* ctx, const double* l2, double* e_dev, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::pseudoenergy(ctx->c, l2, e_dev, st(stream));
  });
}

int ccx_vjp(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, const double* c1,
            const double* c2, double* g1, double* g2, int flags, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && c1 && c2 && g1 && g2, ccx::ERR_ARG, "null argument");
    CCX_CHECK((flags & ~(CCX_HBAR_STORED | CCX_VJP_REUSE_HBAR)) == 0, ccx::ERR_ARG, "unknown ccx_vjp flag");
    auto& c = ctx->c;
    hipStream_t s = st(stream);
    if (flags & CCX_VJP_REUSE_HBAR)
      CCX_CHECK(c.have_hbar, ccx::ERR_STATE, "CCX_VJP_REUSE_HBAR without a built HBAR");
    else
      ccx::hbar_build(c, fock, t1, t2, (flags & CCX_HBAR_STORED) != 0, s);
    // l1 = c1, l2 = c2 + c2^T (r2 is pair-symmetric, so only the symmetric part of c2 acts)
    ccx::pair_symmetrize(c.o, c.v, 1.0, c2, c.Lsym, s);
    ccx::lambda_residual(c, c1, c.Lsym, g1, g2, false, 0.5, s);
  });
}

int ccx_jvp(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, const double* x1,
            const double* x2, double* j1, double* j2, int flags, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx && x1 && x2 && j1 && j2, ccx::ERR_ARG, "null argument");
    CCX_CHECK((flags & ~(CCX_HBAR_STORED | CCX_VJP_REUSE_HBAR)) == 0, ccx::ERR_ARG, "unknown ccx_jvp flag");
    auto& c = ctx->c;
    hipStream_t s = st(stream);
    if (flags & CCX_VJP_REUSE_HBAR)
      CCX_CHECK(c.have_hbar, ccx::ERR_STATE, "CCX_VJP_REUSE_HBAR without a built HBAR");
    else
      ccx::hbar_build(c, fock, t1, t2, (flags & CCX_HBAR_STORED) != 0, s);
    // the pair-symmetric part of x2 (t2_ijab and t2_jiba are one parameter)
    ccx::pair_symmetrize(c.o, c.v, 0.5, x2, c.Lsym, s);
    ccx::jacobian(c, x1, c.Lsym, j1, j2, s);
  });
}

int ccx_densities(ccx_ctx* ctx, const double* t1, const double* t2, const double* l1, const double* l2,
                  double* D, double* Doooo, double* Dooov, double* Dovov, double* Doovv, double* Dvvvo,
                  double* Dvvvv, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::densities(ctx->c, t1, t2, l1, l2, D, Doooo, Dooov, Dovov, Doovv, Dvvvo, Dvvvv, st(stream));
  });
}

int ccx_density_energy(ccx_ctx* ctx, const double* t1, const double* t2, const double* l2, const double* fock,
                       const double* D, const double* Doooo, const double* Dooov, const double* Dovov,
                       const double* Doovv, const double* Dvvvo, double* e_dev, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::density_energy(ctx->c, t1, t2, l2, fock, D, Doooo, Dooov, Dovov, Doovv, Dvvvo, e_dev, st(stream));
  });
}

int ccx_triples(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* e_dev,
                void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::triples_energy(ctx->c, fock, t1, t2, 0, ctx->c.o, false, e_dev, st(stream));
  });
}

int ccx_triples_range(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, int i0, int i1,
                      double* e_dev, void* stream) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::triples_energy(ctx->c, fock, t1, t2, i0, i1, true, e_dev, st(stream));
  });
}

int ccx_t3_denominator(const double* fock, int64_t nmo, int p, int q, int r, int e0, int n, int occ_fixed,
                       double* t3, void* stream) {
  return guarded([&] { ccx::t3_denominator(fock, nmo, p, q, r, e0, n, occ_fixed != 0, t3, st(stream)); });
}

int ccx_timing_start(ccx_ctx* ctx) {
  return guarded([&] {
    CCX_CHECK(ctx, ccx::ERR_ARG, "null context");
    ccx::timing_start(ctx->c);
  });
}

int ccx_timing_read(ccx_ctx* ctx, double* out, int n) {
  return guarded([&] {
    CCX_CHECK(ctx && out && n >= 0, ccx::ERR_ARG, "bad timing arguments");
    ccx::timing_read(ctx->c, out, n);
  });
}

int64_t ccx_device_bytes(const ccx_ctx* ctx) {
  return ctx ? ctx->c.bytes + (int64_t)ctx->c.ws.size : 0;
}

int ccx_contract(const char* subscripts, double alpha, const double* A, const int64_t* a_shape,
                 const int64_t* a_strides, const double* B, const int64_t* b_shape, const int64_t* b_strides,
                 double beta, double* C, const int64_t* c_shape, const int64_t* c_strides, void* stream) {
  return contract_any<double>(subscripts, alpha, A, a_shape, a_strides, B, b_shape, b_strides, beta, C,
                              c_shape, c_strides, stream);
}

int ccx_contract_noalloc(const char* subscripts, double alpha, const double* A, const int64_t* a_shape,
                         const int64_t* a_strides, const double* B, const int64_t* b_shape,
                         const int64_t* b_strides, double beta, double* C, const int64_t* c_shape,
                         const int64_t* c_strides, void* stream) {
  void* p = nullptr;
  const int rc = guarded([&] { p = noalloc_scratch(); });
  if (rc != CCX_OK) return rc;
  ccx::Workspace ws = ccx::Workspace::wrap(p, CCX_NOALLOC_SCRATCH_BYTES);
  return contract_any<double>(subscripts, alpha, A, a_shape, a_strides, B, b_shape, b_strides, beta, C,
                              c_shape, c_strides, stream, &ws);
}

int ccx_contract_f32(const char* subscripts, float alpha, const float* A, const int64_t* a_shape,
                     const int64_t* a_strides, const float* B, const int64_t* b_shape,
                     const int64_t* b_strides, float beta, float* C, const int64_t* c_shape,
                     const int64_t* c_strides, void* stream) {
  return contract_any<float>(subscripts, alpha, A, a_shape, a_strides, B, b_shape, b_strides, beta, C,
                             c_shape, c_strides, stream);
}

int ccx_gemm_f16x3(int transa, int transb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                   int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, void* stream) {
  static thread_local ccx::Workspace ws, kws;
  return guarded([&] {
    CCX_CHECK(((A && B) || K == 0) && (C || M == 0 || N == 0), ccx::ERR_ARG, "null matrix pointer");
    CCX_CHECK(lda >= (transa ? M : K) && ldb >= (transb ? K : N) && ldc >= N, ccx::ERR_ARG, "leading dimension too small");
    ccx::gemm_f16x3(transa != 0, transb != 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ws, kws, st(stream));
  });
}

const char* ccx_last_error(void) { return g_err.c_str(); }

}  // extern "C"

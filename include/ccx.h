/*
 * ccx.h -- C ABI of the MI355X closed-shell CCSD library (libccx.so, gfx950).
 *
 * The drop-in boundary for pycc's spatial RHF-CCSD hot path. Each entry point replaces one piece of
 * the reference (CrawfordGroup/pycc) that a host binding calls:
 *
 *   pycc/device.py:38-86     ContractionBackend.__call__(subscripts, *operands)  -> ccx_contract
 *   pycc/device.py:89-173    DeviceManager(device, precision)                    -> ccx_create
 *   pycc/wavefunction.py:165-169, hamiltonian.py:67-70  (ERI/L seeding)          -> ccx_set_integrals_*
 *   pycc/ccwfn.py:321-372    CCwfn.residuals(F, t1, t2)                          -> ccx_residual
 *   pycc/ccwfn.py:1122-1162  CCwfn.cc_energy(o, v, F, L, t1, t2)                 -> ccx_energy
 *   pycc/ccwfn.py:281-284    Jacobi update t += r/D and the 2-norm "rms"         -> ccx_update
 *   pycc/utils.py:257-361    helper_diis (__init__, add_error_vector, extrapolate) -> ccx_diis_*
 *   pycc/cctriples.py:177-240  t_tjl(ccwfn), the (T) correction                   -> ccx_triples
 *
 * Conventions
 *   - All amplitude/residual/integral pointers are DEVICE pointers (FP64, C-contiguous):
 *       t1, r1 [no][nv];  t2, r2 [no][no][nv][nv];  fock [nmo][nmo] with nmo = no + nv (active
 *       orbitals only, occupied first -- a frozen core is removed by the caller).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream). All work of a call is
 *     enqueued on it; calls return without synchronising except where noted (scalar readbacks,
 *     DIIS solves, integral setup).
 *   - Return value: 0 on success, a negative CCX_ERR_* code on failure; ccx_last_error() returns the
 *     thread-local message of the last failure.
 *   - The caller owns amplitude and residual buffers; the context owns its integral copies and
 *     work space. One context per device and thread.
 */
#ifndef CCX_H
#define CCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccx_ctx ccx_ctx;

/* precision modes (pycc/device.py:113-125 validates 'SP'/'DP'; these are the MI355X modes) */
#define CCX_MODE_DP 0           /* every contraction in FP64 (v_mfma_f64_16x16x4f64)                */
#define CCX_MODE_MIXED 1        /* ladder + ring GEMMs in FP32 (v_mfma_f32_16x16x4f32, FP32 accum.); */
                                /* every other term, the update, DIIS, norms and energy in FP64      */
#define CCX_MODE_MIXED_BF16X3 2 /* as MIXED, with the ladder as BF16x3 split products                */
#define CCX_MODE_MIXED_F16X3 3  /* ladder, rings and every other large contraction as FP16x3 split  */
                                /* products (hi + lo, 22 bits, power-of-two operand scales, FP32     */
                                /* accumulation); update, DIIS, norms, energy and Fmi's tau~.L FP64  */

#define CCX_OK 0
#define CCX_ERR_ARG (-1)
#define CCX_ERR_HIP (-2)
#define CCX_ERR_STATE (-3)
#define CCX_ERR_NUMERIC (-4)
#define CCX_ERR_INTERNAL (-5)

/* Create a context for `no` occupied and `nv` virtual orbitals in precision `mode`.
 * Replaces DeviceManager(device='GPU', precision=...) (pycc/device.py:113-151). */
int ccx_create(int no, int nv, int mode, ccx_ctx** out);
int ccx_destroy(ccx_ctx* ctx);

/* Integrals from a dense Dirac ERI <pq|rs> [nmo]^4 (pycc Hamiltonian layout, hamiltonian.py:67-70)
 * and the Fock matrix. The context builds its own blocks; the inputs may be freed afterwards.
 * Orbital energies (denominators, ccwfn.py:195-198) are the diagonal of this Fock matrix. */
int ccx_set_integrals_dense(ccx_ctx* ctx, const double* fock, const double* eri, void* stream);

/* Integrals from density-fitting factors B[naux][nmo][nmo] with (pq|rs) = sum_Q B_Qpq B_Qrs; the
 * integral blocks are generated on the device (no nmo^4 array is formed). */
int ccx_set_integrals_df(ccx_ctx* ctx, const double* fock, const double* B, int naux, void* stream);

/* r1, r2 = CCwfn.residuals(F, t1, t2) (pycc/ccwfn.py:321-372): r2 is the full symmetrised
 * residual (ccwfn.py:790). `fock` may differ from the setup Fock (e.g. a perturbed one). */
int ccx_residual(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* r1,
                 double* r2, void* stream);

/* *e_dev = 2 f_ia t_ia + (t2_ijab + t_ia t_jb) L_ijab (pycc/ccwfn.py:1160-1161); e_dev is a device
 * scalar. */
int ccx_energy(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* e_dev,
               void* stream);

/* t1 += r1/Dia, t2 += r2/Dijab; *rms2_dev = sum (r1/Dia)^2 + sum (r2/Dijab)^2 (device scalar;
 * pycc/ccwfn.py:281-284 takes its square root). */
int ccx_update(ccx_ctx* ctx, double* t1, double* t2, const double* r1, const double* r2,
               double* rms2_dev, void* stream);

/* DIIS (pycc/utils.py:257-361).
 * ccx_diis_reset    = helper_diis(t1, t2, max_diis): seeds the previous amplitudes; `capacity`
 *                     bounds the stored vectors (>= max_diis + 1, <= 64).
 * ccx_diis_add      = add_error_vector(t1, t2)   (error e = t - t_previous); synchronises.
 * ccx_diis_extrapolate = extrapolate(t1, t2): overwrites t1, t2 with sum_k c_k t_k; if coef_out is
 *                     non-NULL it receives the coefficients (host array, >= capacity entries).
 *                     Synchronises. A singular Pulay system returns CCX_ERR_NUMERIC. */
int ccx_diis_reset(ccx_ctx* ctx, const double* t1, const double* t2, int max_diis, int capacity,
                   void* stream);
int ccx_diis_add(ccx_ctx* ctx, const double* t1, const double* t2, void* stream);
int ccx_diis_extrapolate(ccx_ctx* ctx, double* t1, double* t2, double* coef_out, void* stream);

/* ---- backward: HBAR, the Lambda residual and the vector-Jacobian product ----------------------
 *
 * DP contexts run them in FP64. CCX_MODE_MIXED_F16X3 contexts run them too, without
 * CCX_HBAR_STORED: HBAR is held in FP64, and its contractions of at least 1e11 flops and the Lambda
 * ladder run as FP16x3 split products (the Lambda residual within 5e-7 relative of DP at o=16,
 * v=256). ccx_jvp and
 * ccx_densities need a DP context.
 *
 * ccx_hbar_build = cchbar(ccwfn) (pycc/cchbar.py:54-98, spatial CCSD builders :128-823) at the given
 *   fock, t1, t2 (copied; the inputs may change afterwards). flags = 0 holds the blocks of size
 *   o^2 v^2 and below and evaluates every Lambda term that needs Hvvvv, Hvovv or Hvvvo in factored
 *   form from the packed <ab|ef> of the context (no v^4 block is formed); flags = CCX_HBAR_STORED
 *   also stores Hvvvv (v^4), Hvovv and Hvvvo as the reference does. */
#define CCX_HBAR_STORED 1
#define CCX_VJP_REUSE_HBAR 2
int ccx_hbar_build(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, int flags,
                   void* stream);

/* Copy HBAR block `block` into `out` (device, reference layout and order of cchbar.py:86-96):
 *   0 Hov[o][v]  1 Hvv[v][v]  2 Hoo[o][o]  3 Hoooo[o][o][o][o]  4 Hvvvv[v][v][v][v]  5 Hvovv[v][o][v][v]
 *   6 Hooov[o][o][o][v]  7 Hovvo[o][v][v][o]  8 Hovov[o][v][o][v]  9 Hvvvo[v][v][v][o]  10 Hovoo[o][v][o][o]
 * Blocks 4, 5 and 9 exist only after a CCX_HBAR_STORED build. */
int ccx_hbar_block(ccx_ctx* ctx, int block, double* out, void* stream);

/* r1, r2 = the Lambda residuals r_L1, r_L2 at (l1, l2) on the current HBAR
 * (cclambda.residuals, pycc/cclambda.py:202-256: build_Goo :281, build_Gvv :306, r_L1 :345-369,
 * r_L2 :468-496). with_source = 0 leaves out the l-independent terms 2 Hov and L. l2 must be
 * pair-symmetric (l2_ijab = l2_jiba) unless the HBAR is stored. Outputs must not alias inputs. */
int ccx_lambda_residual(ccx_ctx* ctx, const double* l1, const double* l2, double* r1, double* r2,
                        int with_source, void* stream);

/* *e_dev = 1/2 <ij|ab> l2_ijab (cclambda.pseudoenergy, pycc/cclambda.py:547-570). */
int ccx_pseudoenergy(ccx_ctx* ctx, const double* l2, double* e_dev, void* stream);

/* Vector-Jacobian product of ccx_residual at (fock, t1, t2) with cotangent (c1, c2):
 *   g1 = sum c1 . dr1/dt1 + sum c2 . dr2/dt1,   g2 = the same for t2, pair-symmetrised,
 * computed as the Lambda residual without source terms at l1 = c1, l2 = c2 + c2^T (the backward of
 * the residual; cclambda.py:345-369, :468-496 are its contractions). The fock matrix is held
 * constant. flags: CCX_HBAR_STORED as for ccx_hbar_build; CCX_VJP_REUSE_HBAR skips the HBAR build
 * and uses the last one (built from the same fock, t1, t2). Outputs must not alias inputs. */
int ccx_vjp(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, const double* c1,
            const double* c2, double* g1, double* g2, int flags, void* stream);

/* Jacobian-vector product of ccx_residual at (fock, t1, t2) along (x1, x2): j1, j2 = the derivative of
 * (r1, r2) in that direction, the CCSD Jacobian applied on HBAR (CCderiv._ccsd_jacobian,
 * pycc/ccderiv.py:149-180, with its doubles symmetrised as ccderiv.py:257-258 does). Only the
 * pair-symmetric part of x2 acts. flags as for ccx_vjp. Outputs must not alias inputs. */
int ccx_jvp(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, const double* x1,
            const double* x2, double* j1, double* j2, int flags, void* stream);

/* One- and two-particle densities of the CC Lagrangian E + l1.r1 + 1/2 l2.r2 at (t1, t2, l1, l2):
 * the spatial CCSD builders of ccdensity.py (build_Doo :335, Dvv :373, Dvo :410, Dov :415,
 * Doooo :563, Dvvvv :587, Dooov :611, Dvvvo :682, Dovov :753, Doovv :784). They are dL/dF and
 * dL/d<pq|rs> at fixed amplitudes, so with l = (c1, c2 + c2^T) and the energy part (l = 0)
 * subtracted they are the integral half of the vector-Jacobian product. Outputs (device, FP64):
 *   D [nmo][nmo] (Doo, Dov, Dvo, Dvv blocks), Doooo [o][o][o][o], Dooov [o][o][o][v],
 *   Dovov [o][v][o][v], Doovv [o][o][v][v], Dvvvo [v][v][v][o] (may be NULL), Dvvvv [v]^4 (may be NULL).
 * DP contexts. */
int ccx_densities(ccx_ctx* ctx, const double* t1, const double* t2, const double* l1, const double* l2,
                  double* D, double* Doooo, double* Dooov, double* Dovov, double* Doovv, double* Dvvvo,
                  double* Dvvvv, void* stream);

/* *e_dev = ccdensity.compute_energy (pycc/ccdensity.py:113-211) from the densities above:
 * F_ij D_ij + F_ab D_ab + 1/2 <ij|kl> Doooo + <ij|ka> Dooov + <ab|ci> Dvvvo + <ia|jb> Dovov
 * + 1/2 <ij|ab> Doovv + 1/2 <ab|cd> Dvvvv, with the Dvvvv term evaluated from (t1, t2, l2) through
 * the packed ladder (no v^4 array). Like the reference it omits the f_ov terms (canonical orbitals). */
int ccx_density_energy(ccx_ctx* ctx, const double* t1, const double* t2, const double* l2, const double* fock,
                       const double* D, const double* Doooo, const double* Dooov, const double* Dovov,
                       const double* Doovv, const double* Dvvvo, double* e_dev, void* stream);

/* *e_dev = the (T) correction of closed-shell CCSD(T) at (fock, t1, t2): pycc/cctriples.py:177-240 (t_tjl; the same energy as t_vikings :243 and t_vikings_inverted
 * :311), with the connected and disconnected T3 numerators of cctriples.py:27-72 and :108-146 and
 * the orbital-energy denominators from the diagonal of `fock`. FP64, except in CCX_MODE_MIXED_F16X3,
 * whose slot products are FP16x3 split products with FP32 accumulation and output (E(T) within
 * ~1e-9 relative of FP64 for tens of virtuals and more, ~5e-8 for two or three). Synchronises
 * `stream`. Device scratch for the duration of the call:
 * o v^2 (v + o) + 24 v^3 doubles in FP64; in CCX_MODE_MIXED_F16X3 o v^2 (v + o) doubles while the
 * operand is split (half that afterwards, as FP16 hi/lo) + 12 v^3 + 34 v^2 doubles. */
int ccx_triples(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, double* e_dev,
                void* stream);

/* e_dev[0] += the part of ccx_triples' E(T) from the occupied triples i >= j >= k with
 * i0 <= i < i1 (0 <= i0 <= i1 <= no). Zeroing e_dev[0] and calling it over consecutive ranges that
 * cover [0, no) in ascending order gives ccx_triples' value bit for bit. A long (T) can then report
 * progress, or stop and resume, between ranges. Same modes, scratch and synchronisation as
 * ccx_triples; no counterpart in the reference, whose t_tjl is one call (pycc/cctriples.py:177). */
int ccx_triples_range(ccx_ctx* ctx, const double* fock, const double* t1, const double* t2, int i0, int i1,
                      double* e_dev, void* stream);

/* t3[x][y][z] /= D for x, y, z < n (t3 a contiguous n^3 device array), the orbital-energy
 * denominator of the WithDenom=True branches of the batched T3 builders: pycc/cctriples.py:64-70
 * (t3c_ijk) and :139-145 (t3d_ijk) with occ_fixed = 1, fixed occupied orbitals (p, q, r) = (i, j, k)
 * and running virtuals e0 + a, e0 + b, e0 + c:  D = f_pp + f_qq + f_rr - (f_aa + f_bb + f_cc);
 * :97-103 (t3c_abc) and :165-171 (t3d_abc) with occ_fixed = 0, fixed virtuals (a, b, c) and
 * running occupied orbitals: D = (f_ii + f_jj + f_kk) - (f_pp + f_qq + f_rr). Orbital indices are
 * rows of the nmo x nmo device matrix `fock`, whose diagonal is read; each sum of three is added
 * left to right, as the reference's broadcasts add them. Asynchronous on `stream`. */
int ccx_t3_denominator(const double* fock, int64_t nmo, int p, int q, int r, int e0, int n, int occ_fixed,
                       double* t3, void* stream);

/* Measurement hooks (no counterpart in the reference, which has only wall-clock prints,
 * pycc/utils.py:229-238): HIP events recorded on the residual's stream around the ladder GEMMs,
 * the ring GEMMs and each whole residual.
 * ccx_timing_start clears and starts recording; ccx_timing_read synchronises, writes
 * out[3*p + {0,1,2}] = {launches, milliseconds, algorithmic flops} for p = 0 (ladder GEMMs),
 * 1 (ring GEMMs), 2 (whole residuals), n <= 9 entries, and stops recording. */
int ccx_timing_start(ccx_ctx* ctx);
int ccx_timing_read(ccx_ctx* ctx, double* out, int n);

/* Device bytes held by the context (integrals, work space, DIIS history). */
int64_t ccx_device_bytes(const ccx_ctx* ctx);

/* C = alpha * sum A*B + beta * C for an einsum string "ab,bc->ac" over strided device tensors
 * (ContractionBackend.__call__, pycc/device.py:64-86; two operands). Shapes and strides are in
 * elements, one entry per index letter of the corresponding operand. */
int ccx_contract(const char* subscripts, double alpha, const double* A, const int64_t* a_shape,
                 const int64_t* a_strides, const double* B, const int64_t* b_shape,
                 const int64_t* b_strides, double beta, double* C, const int64_t* c_shape,
                 const int64_t* c_strides, void* stream);
/* FP32 variant (FP32 MFMA, FP32 accumulation). */
int ccx_contract_f32(const char* subscripts, float alpha, const float* A, const int64_t* a_shape,
                     const int64_t* a_strides, const float* B, const int64_t* b_shape,
                     const int64_t* b_strides, float beta, float* C, const int64_t* c_shape,
                     const int64_t* c_strides, void* stream);

/* C = alpha op(A) op(B) + beta C for FP32 row-major device matrices (op = transpose when transa /
 * transb is nonzero; C is M x N, op(A) M x K, op(B) K x N) through FP16x3 split products: each
 * operand is held as hi + lo FP16 with a power-of-two scale (22 significant bits), three FP16 MFMA
 * products per product, FP32 accumulation. No counterpart in the reference: it is the split GEMM
 * of CCX_MODE_MIXED_F16X3 exposed on its own. Uses a per-thread device workspace. */
int ccx_gemm_f16x3(int transa, int transb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                   int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, void* stream);

/* ccx_contract without any device allocation: the same contraction, with its split-K partial sums in
 * a fixed device array of CCX_NOALLOC_SCRATCH_BYTES per GPU inside the library (no hipMalloc, no
 * hipFree, no allocating library call; the split count is chosen to fit, and a request beyond the
 * array is refused with CCX_ERR_ARG). No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it). Every call on a GPU shares that GPU's array: calls that overlap on one GPU (other threads or
 * other streams) race. The array's device address is looked up on the first call on each device. */
#define CCX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ccx_contract_noalloc(const char* subscripts, double alpha, const double* A, const int64_t* a_shape,
                         const int64_t* a_strides, const double* B, const int64_t* b_shape,
                         const int64_t* b_strides, double beta, double* C, const int64_t* c_shape,
                         const int64_t* c_strides, void* stream);

/* Message of the last failed call on this thread ("" if none). */
const char* ccx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CCX_H */

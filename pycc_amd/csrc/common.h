// Shared definitions for the CCSD device library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <stdexcept>

namespace ccx {

// Error carried from any internal call up to the C-ABI boundary, where it becomes a status code
// plus the thread-local message returned by ccx_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  OK = 0,
  ERR_ARG = -1,       // invalid argument (shape, null pointer, unsupported mode)
  ERR_HIP = -2,       // HIP runtime error (launch, allocation, copy)
  ERR_STATE = -3,     // call out of order (e.g. residual before integrals)
  ERR_NUMERIC = -4,   // singular DIIS system, non-finite values
  ERR_INTERNAL = -5,
};

#define CCX_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::ccx::Error(::ccx::ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define CCX_CHECK(cond, code, msg)                                   \
  do {                                                               \
    if (!(cond)) throw ::ccx::Error((code), std::string(msg));       \
  } while (0)

// A 2-level index: i -> (i / n2) * s1 + (i % n2) * s2. Lets one GEMM dimension span two tensor
// indices with unrelated strides (e.g. the compound (m, e) of a ring contraction).
struct Idx2 {
  int64_t n2 = 1;
  int64_t s1 = 0, s2 = 0;
};

// C[b][m][n] = alpha * sum_k A[b][m][k] * B[b][k][n] + beta * C[b][m][n], every operand addressed
// through 2-level indices. Split-K partials go to `work` and are reduced by a second kernel.
template <typename T>
struct GemmParams {
  int64_t M = 0, N = 0, K = 0;
  int64_t batch = 1;
  Idx2 am, ak, bk, bn, cm, cn;
  int64_t a_b = 0, b_b = 0, c_b = 0;   // batch strides
  T alpha = 1, beta = 0;
  const T* A = nullptr;
  const T* B = nullptr;
  T* C = nullptr;
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace ccx

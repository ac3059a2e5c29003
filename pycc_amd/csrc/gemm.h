// Tensor-contraction GEMM launcher (see gemm.hip).
#pragma once
#include "common.h"
#include <algorithm>
#include <vector>

namespace ccx {

// Grow-only device scratch owned by a context; `limit` bounds what split-K may request. A fixed
// workspace wraps memory it does not own (size = limit): it never allocates or frees, and a request
// beyond its size is an error.
struct Workspace {
  void* ptr = nullptr;
  size_t size = 0;
  size_t limit = size_t(1) << 30;
  bool fixed = false;
  static Workspace wrap(void* p, size_t bytes) {
    Workspace w;
    w.ptr = p;
    w.size = w.limit = bytes;
    w.fixed = true;
    return w;
  }
  void* get(size_t bytes) {
    if (bytes > size) {
      CCX_CHECK(!fixed, ERR_ARG,
                "workspace request of " + std::to_string(bytes) + " bytes exceeds the fixed scratch of " +
                    std::to_string(size) + " bytes");
      if (ptr) CCX_HIP(hipFree(ptr));
      ptr = nullptr;
      CCX_HIP(hipMalloc(&ptr, bytes));
      size = bytes;
    }
    return ptr;
  }
  void release() {
    if (ptr && !fixed) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
  }
};

template <typename T>
void tc_gemm(const GemmParams<T>& p, Workspace& ws, hipStream_t s);

// Plain strided matrices (gemm_plain.hip): C[b][m * ldc + n] = alpha op(A) op(B) + beta C.
template <typename T>
struct PlainGemm {
  int64_t M = 0, N = 0, K = 0, batch = 1;
  const T* A = nullptr;
  const T* B = nullptr;
  T* C = nullptr;
  int64_t lda = 0, ldb = 0, ldc = 0;
  int64_t a_b = 0, b_b = 0, c_b = 0;
  bool a_kc = true;   // A[m * lda + k] (else A[k * lda + m])
  bool b_nc = true;   // B[k * ldb + n] (else B[n * ldb + k])
  T alpha = 1, beta = 0;
  // split-K: when ktot > 0, grid.y runs over K chunks of length K (the last one shorter), each
  // chunk is a "batch" entry (a_b / b_b advance by one chunk) and ktot is the total K.
  int64_t ktot = 0;
  // K concatenation (plain_gemm_kcat; A k-contiguous, FP64): k in [K1, K) reads A2 at k - K1 and
  // row (m mod sw) sw + m / sw, the (x, y) -> (y, x) swap of a row m = x sw + y (M = sw^2); batch b
  // adds b a2_b to A2.
  const T* A2 = nullptr;
  int64_t K1 = 0, sw = 0, a2_b = 0;
};
template <typename T>
bool plain_gemm_ok(const PlainGemm<T>& p);
template <typename T>
void plain_gemm(const PlainGemm<T>& p, hipStream_t s);
// C_b = alpha [A | A2_b swapped] B_b^T + beta C_b over K = K concatenated at K1 (FP64, A and B
// k-contiguous, fields A2 .. a2_b above; K1 a multiple of 16).
void plain_gemm_kcat(const PlainGemm<double>& p, hipStream_t s);

}  // namespace ccx

// Tensor-contraction GEMM launcher (see gemm.hip).
#pragma once
#include "common.h"
#include <algorithm>
#include <vector>

namespace ccx {

// Grow-only device scratch owned by a context; `limit` bounds what split-K may request.
struct Workspace {
  void* ptr = nullptr;
  size_t size = 0;
  size_t limit = size_t(1) << 30;
  void* get(size_t bytes) {
    if (bytes > size) {
      if (ptr) CCX_HIP(hipFree(ptr));
      ptr = nullptr;
      CCX_HIP(hipMalloc(&ptr, bytes));
      size = bytes;
    }
    return ptr;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
  }
};

template <typename T>
void tc_gemm(const GemmParams<T>& p, Workspace& ws, hipStream_t s);

}  // namespace ccx

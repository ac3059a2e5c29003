// Einsum-style contraction planner over the tensor-contraction GEMM.
//
// contract(alpha, A, B, beta, C): C = alpha * sum A * B + beta * C for labelled strided views, the
// device counterpart of pycc's ContractionBackend.__call__(subscripts, *operands)
// (pycc/device.py:64-86). Indices are classified as batch (in A, B and C), M (A and C), N (B and C)
// or K (A and B). Indices of one class whose strides nest in every operand are fused; a class left
// with more than two indices (more than one for batch) is looped on the host.
#pragma once
#include "common.h"
#include "gemm.h"
#include <string>
#include <vector>

namespace ccx {

struct TV {
  const void* p = nullptr;
  std::string lab;               // one character per dimension
  std::vector<int64_t> ext, str; // extents and element strides
  int rank() const { return (int)lab.size(); }
  int find(char c) const {
    for (int i = 0; i < rank(); ++i)
      if (lab[i] == c) return i;
    return -1;
  }
};

// Row-major view.
TV tv(const void* p, const char* lab, std::vector<int64_t> ext);
// Explicit strides.
TV tvs(const void* p, const char* lab, std::vector<int64_t> ext, std::vector<int64_t> str);
// Same memory, dimensions relabelled/reordered: `lab` lists the view's labels for the storage
// dimensions in storage order (e.g. a [m][e][n][f] buffer seen as "menf").
inline TV relabel(TV v, const char* lab) { v.lab = lab; return v; }

// split16: an FP64 contraction whose GEMM has at least X3_ROUTE_FLOPS flops (and no batch index,
// and a full tile grid) runs as an FP16x3 split product (FP32 accumulation, about 2^-22 relative
// error per product) instead of in FP64; smaller ones, and every FP32 one, stay as they are.
constexpr double X3_ROUTE_FLOPS = 1e11;
template <typename T>
void contract(T alpha, const TV& A, const TV& B, T beta, const TV& C, Workspace& ws, hipStream_t s,
              bool split16 = false);

// C = alpha * A + beta * C elementwise, with A's labels a permutation of C's (up to rank 6).
template <typename T>
void axpby(T alpha, const TV& A, T beta, const TV& C, hipStream_t s);

// C = beta * C (beta = 0 writes zeros).
template <typename T>
void scale(T beta, const TV& C, hipStream_t s);

}  // namespace ccx

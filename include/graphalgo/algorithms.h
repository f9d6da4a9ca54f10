// Drop-in of the graph algorithms that live in the reference's CLI
// (src/main.cpp:11-146): the operator plug point mxmOp<T>, the three triangle
// counts and k_truss.  Each takes the multiplication operator as the
// reference does.  When the operator is one of this library's structural-mask
// products (mxmm_msa / mxmm_mca / mxmm_heap for int), the whole algorithm runs
// device-resident (ga_triangle_count, ga_ktruss, ...: no per-pass host
// round trip); any other operator is driven exactly as the reference drives
// it, one product per pass.
#pragma once

#include <chrono>
#include <cstdint>
#include <iostream>

#include "betweenness_centrality.h"

template <typename T>
using mxmOp = void (*)(bool, const spMtx<T>&, const spMtx<T>&, const spMtx<T>&, spMtx<T>&);

namespace graphalgo_detail {
inline bool is_gpu_op(mxmOp<int> op) {
  return op == static_cast<mxmOp<int>>(mxmm_msa<int, int>) || op == static_cast<mxmOp<int>>(mxmm_mca<int, int>) ||
         op == static_cast<mxmOp<int>>(mxmm_heap<int>);
}
}  // namespace graphalgo_detail

// src/main.cpp:65-86: C<L> = L*L, T = sum(C).
inline int64_t triangle_counting_masked_sandia(const spMtx<int>& A, mxmOp<int> matrixMult, bool isParallel) {
  int64_t T = 0;
  if (graphalgo_detail::is_gpu_op(matrixMult)) {
    graphalgo_detail::ensure_init();
    const ga_csr a = A.view(false);
    graphalgo_detail::check(ga_triangle_count(&a, &T));
    return T;
  }
  spMtx<int> L = extract_lower_triangle(A), C;
  matrixMult(isParallel, L, L, L, C);
  for (size_t j = 0; j < C.nz; ++j) T += C.Val[j];
  return T;
}

// src/main.cpp:40-62: C<A> = L*U, T = sum(C)/2.
inline int64_t triangle_counting_masked_lu(const spMtx<int>& A, mxmOp<int> matrixMult, bool isParallel) {
  int64_t T = 0;
  if (graphalgo_detail::is_gpu_op(matrixMult)) {
    graphalgo_detail::ensure_init();
    const ga_csr a = A.view(false);
    graphalgo_detail::check(ga_triangle_count_lu(&a, &T));
    return T;
  }
  spMtx<int> L = extract_lower_triangle(A), U = transpose(L), C;
  matrixMult(isParallel, L, U, A, C);
  for (size_t j = 0; j < C.nz; ++j) T += C.Val[j];
  return T >> 1;
}

// src/main.cpp:14-37: per-vertex triangles; caller owns the new[] array.
inline int* triangle_counting_vertex(const spMtx<int>& A, mxmOp<int> matrixMult, bool isParallel) {
  int* out = new int[A.m];
  if (graphalgo_detail::is_gpu_op(matrixMult)) {
    graphalgo_detail::ensure_init();
    std::vector<int64_t> pv(A.m);
    const ga_csr a = A.view(false);
    ga_status st = ga_triangle_count_vertex(&a, pv.data());
    if (st != GA_OK) {
      delete[] out;
      graphalgo_detail::check(st);
    }
    for (size_t i = 0; i < A.m; ++i) out[i] = (int)pv[i];
    return out;
  }
  spMtx<int> SQ;
  matrixMult(isParallel, A, A, A, SQ);
  for (size_t i = 0; i < A.m; ++i) {
    int s = 0;
    for (int j = SQ.Rst[i]; j < SQ.Rst[i + 1]; ++j) s += SQ.Val[j];
    out[i] = s >> 1;
  }
  return out;
}

// src/main.cpp:89-146: the k-truss of A (values 1); *iterations receives
// the number of support passes (the reference's totalIterationNum, which it
// prints).  The four-argument form below has the reference's exact type.
inline spMtx<int> k_truss(const spMtx<int>& A, int k, mxmOp<int> matrixMult, bool isParallel, int* iterations) {
  if (graphalgo_detail::is_gpu_op(matrixMult)) {
    graphalgo_detail::ensure_init();
    spMtx<int> C(A.m, A.n, A.nz);
    int64_t nnz = 0;
    int32_t it = 0;
    const ga_csr a = A.view(false);
    graphalgo_detail::check(ga_ktruss(&a, k, C.Rst, C.Col, &nnz, &it));
    C.nz = (size_t)nnz;
    for (size_t j = 0; j < C.nz; ++j) C.Val[j] = 1;
    if (iterations) *iterations = it;
    return C;
  }
  spMtx<int> C = A, Tmp;
  const int n = (int)A.m;
  std::vector<int> rst(n + 1, 0);
  int it = 0;
  for (int t = 0; t < n; ++t) {
    matrixMult(isParallel, C, C, C, Tmp);
    int pos = 0;
    for (int i = 0; i < n; ++i) {
      for (int j = Tmp.Rst[i]; j < Tmp.Rst[i + 1]; ++j)
        if (Tmp.Val[j] >= k - 2) {
          Tmp.Col[pos] = Tmp.Col[j];
          Tmp.Val[pos++] = 1;
        }
      rst[i + 1] = pos;
    }
    std::memcpy(Tmp.Rst, rst.data(), (n + 1) * sizeof(int));
    Tmp.nz = (size_t)Tmp.Rst[n];
    ++it;
    if (Tmp.nz == C.nz) break;
    std::swap(C, Tmp);
  }
  if (iterations) *iterations = it;
  return C;
}

inline spMtx<int> k_truss(const spMtx<int>& A, int k, mxmOp<int> matrixMult, bool isParallel) {
  return k_truss(A, k, matrixMult, isParallel, nullptr);
}

// Truss number of every stored entry of A (the loop over k around k_truss).
inline std::vector<int32_t> truss_decomposition(const spMtx<int>& A, int32_t* kmax = nullptr) {
  graphalgo_detail::ensure_init();
  std::vector<int32_t> truss(A.nz);
  const ga_csr a = A.view(false);
  graphalgo_detail::check(ga_truss_decomposition(&a, truss.data(), kmax, nullptr));
  return truss;
}

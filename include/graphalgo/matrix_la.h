// Drop-in operator surface of the reference's include/matrix_la.h.  Same
// names and signatures; every product runs on the GPU through the C ABI.
//
//   mxmm_msa / mxmm_mca / mxmm_heap   include/matrix_la.h:571-599, :700-713,
//                                     :998-1028 -> ga_mxm_masked
//   mxmm_msa_cmask                    include/matrix_la.h:788-798 -> ga_mxm_cmask
//   mxmm_spd                          include/matrix_la.h:239-270 -> ga_mxm_spd
//   eWiseAdd, eWiseMult, add_nointersect, multScalar
//                                     include/matrix_la.h:272-541 -> ga_ewise, ga_scale
//   fuseEWiseMultAdd, eWiseMult (dense, masked)
//                                     include/matrix_la.h:192-219
//                                     -> ga_dense_fuse_mult_add, ga_dense_mult_masked
//
// The three structural-mask variants are one GPU operator (the accumulator
// choice of the reference -- dense MSA, mask-indexed MCA, heap -- becomes a
// per-row decision inside the kernel), so binding any of them to
// mxmOp<T> (src/main.cpp:11-12) selects the same path.  `isParallel` is
// accepted and ignored.  Value types: int, int64_t, float, double.  On error the
// functions throw the C ABI's message (const char*), the reference's error
// style.
#pragma once

#include <type_traits>

#include "matrix.h"

namespace graphalgo_detail {

template <typename T>
constexpr ga_dtype dtype_of() {
  static_assert(std::is_same<T, int>::value || std::is_same<T, int64_t>::value || std::is_same<T, float>::value ||
                    std::is_same<T, double>::value,
                "the GPU operators support int, int64_t, float and double values");
  return std::is_same<T, int>::value       ? GA_INT32
         : std::is_same<T, int64_t>::value ? GA_INT64
         : std::is_same<T, float>::value   ? GA_FLOAT32
                                           : GA_FLOAT64;
}

inline void check(ga_status s) {
  if (s != GA_OK) throw ga_last_error();
}

inline void ensure_init() {
  static bool done = false;
  if (!done) {
    check(ga_init(0));
    done = true;
  }
}

}  // namespace graphalgo_detail

// C<M> = A*B, structural mask: C takes M's pattern, explicit zeros kept.
template <typename T, typename U>
void mxmm_msa(bool /*isParallel*/, const spMtx<T>& A, const spMtx<T>& B, const spMtx<U>& M, spMtx<T>& C) {
  graphalgo_detail::ensure_init();
  C.resizeRows(M.m);
  C.resizeVals(M.nz);
  C.n = M.n;
  if (M.nz) std::memcpy(C.Col, M.Col, M.nz * sizeof(int));
  std::memcpy(C.Rst, M.Rst, (M.m + 1) * sizeof(int));
  const ga_csr a = A.view(true), b = B.view(true), mm = M.view(false);
  graphalgo_detail::check(ga_mxm_masked(&a, &b, &mm, graphalgo_detail::dtype_of<T>(), C.Val));
}

template <typename T, typename U>
void mxmm_mca(bool isParallel, const spMtx<T>& A, const spMtx<T>& B, const spMtx<U>& M, spMtx<T>& C) {
  mxmm_msa(isParallel, A, B, M, C);
}

template <typename T, typename U>
spMtx<T> mxmm_mca(bool isParallel, const spMtx<T>& A, const spMtx<U>& B, const spMtx<T>& M) {
  spMtx<T> C;
  mxmm_msa(isParallel, A, B, M, C);
  return C;
}

template <typename T>
void mxmm_heap(bool isParallel, const spMtx<T>& A, const spMtx<T>& B, const spMtx<T>& M, spMtx<T>& C) {
  mxmm_msa(isParallel, A, B, M, C);
}

template <typename T>
spMtx<T> mxmm_heap(bool isParallel, const spMtx<T>& A, const spMtx<T>& B, const spMtx<T>& M) {
  spMtx<T> C;
  mxmm_msa(isParallel, A, B, M, C);
  return C;
}

// C<!M> = A*B with sorted output columns.
template <typename T, typename U>
void mxmm_msa_cmask(bool /*isParallel*/, const spMtx<T>& A, const spMtx<T>& B, const spMtx<U>& M, spMtx<T>& C) {
  graphalgo_detail::ensure_init();
  const ga_csr a = A.view(true), b = B.view(true), mm = M.view(false);
  int32_t* rp = nullptr;
  int32_t* col = nullptr;
  void* val = nullptr;
  int64_t nnz = 0;
  graphalgo_detail::check(ga_mxm_cmask(&a, &b, &mm, graphalgo_detail::dtype_of<T>(), &rp, &col, &val, &nnz));
  C.resizeRows(M.m);
  C.n = M.n;
  std::memcpy(C.Rst, rp, (M.m + 1) * sizeof(int));
  C.resizeVals((size_t)nnz);
  if (nnz) {
    std::memcpy(C.Col, col, nnz * sizeof(int));
    std::memcpy(C.Val, val, nnz * sizeof(T));
  }
  ga_free_host(rp);
  ga_free_host(col);
  ga_free_host(val);
}

// C<M> = A*B with B, C dense (the reference's 5-argument form swaps the
// result buffer into C, include/matrix_la.h:242-270; Cbuf is left unspecified).
template <typename T, typename U>
void mxmm_spd(const spMtx<T>& A, const denseMtx<T>& B, const spMtx<U>& M, denseMtx<T>& C, denseMtx<T>& Cbuf) {
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                "mxmm_spd runs in FP64 or FP32 on the GPU");
  graphalgo_detail::ensure_init();
  Cbuf.resize(M.m, B.n);
  const ga_csr a = A.view(true), mm = M.view(false);
  if constexpr (std::is_same<T, double>::value)
    graphalgo_detail::check(ga_mxm_spd(&a, B.Val, (int64_t)B.n, &mm, Cbuf.Val));
  else
    graphalgo_detail::check(ga_mxm_spd_f32(&a, B.Val, (int64_t)B.n, &mm, Cbuf.Val));
  std::swap(C.Val, Cbuf.Val);
  std::swap(C.capacity, Cbuf.capacity);
  C.m = M.m;
  C.n = B.n;
}

// ---- dense element-wise operators (include/matrix_la.h:192-219) ----
// C += A .* B over all m * n values.
template <typename T>
void fuseEWiseMultAdd(const denseMtx<T>& A, const denseMtx<T>& B, denseMtx<T>& C) {
  graphalgo_detail::ensure_init();
  if (A.m != B.m || A.n != B.n || A.m != C.m || A.n != C.n) throw -1;
  graphalgo_detail::check(
      ga_dense_fuse_mult_add((int64_t)(A.m * A.n), graphalgo_detail::dtype_of<T>(), A.Val, B.Val, C.Val));
}

// C<M> = A .* B: C is zero outside M's pattern.  C is taken by const
// reference and written through its value array, as in the reference.
template <typename T, typename U>
void eWiseMult(const denseMtx<T>& A, const denseMtx<T>& B, const spMtx<U>& M, const denseMtx<T>& C) {
  graphalgo_detail::ensure_init();
  if (A.m != B.m || A.n != B.n || A.m != C.m || A.n != C.n || M.m != C.m || M.n != C.n) throw -1;
  const ga_csr mm = M.view(false);
  graphalgo_detail::check(ga_dense_mult_masked(&mm, graphalgo_detail::dtype_of<T>(), A.Val, B.Val, C.Val));
}

// ---- element-wise operators (include/matrix_la.h:272-541, defects fixed) ----
namespace graphalgo_detail {
template <typename T>
spMtx<T> ewise(int op, const spMtx<T>& A, const spMtx<T>& B, const ga_csr* M) {
  ensure_init();
  const ga_csr a = A.view(true), b = B.view(true);
  int32_t* rp = nullptr;
  int32_t* col = nullptr;
  void* val = nullptr;
  int64_t nnz = 0;
  check(ga_ewise(op, &a, &b, M, dtype_of<T>(), &rp, &col, &val, &nnz));
  spMtx<T> C;
  C.resizeRows(A.m);
  C.n = A.n;
  std::memcpy(C.Rst, rp, (A.m + 1) * sizeof(int));
  C.resizeVals((size_t)nnz);
  if (nnz) {
    std::memcpy(C.Col, col, nnz * sizeof(int));
    std::memcpy(C.Val, val, nnz * sizeof(T));
  }
  ga_free_host(rp);
  ga_free_host(col);
  ga_free_host(val);
  return C;
}
}  // namespace graphalgo_detail

// C<M> = A + B on M's pattern; positions whose sum is zero are left out
template <typename T, typename U>
spMtx<T> eWiseAdd(const spMtx<T>& A, const spMtx<T>& B, const spMtx<U>& M) {
  if (A.m != B.m || A.n != B.n) throw -1;
  const ga_csr mm = M.view(false);
  return graphalgo_detail::ewise(GA_EW_ADD, A, B, &mm);
}

// the union of the patterns, values added where both hold an entry
template <typename T>
spMtx<T> eWiseAdd(const spMtx<T>& A, const spMtx<T>& B) {
  if (A.m != B.m || A.n != B.n) throw -1;
  return graphalgo_detail::ewise(GA_EW_ADD, A, B, nullptr);
}

// the union of two patterns that share no position
template <typename T>
spMtx<T> add_nointersect(const spMtx<T>& A, const spMtx<T>& B) {
  if (A.m != B.m || A.n != B.n) throw -1;
  return graphalgo_detail::ewise(GA_EW_ADD_DISJOINT, A, B, nullptr);
}

// the intersection of the patterns, values multiplied
template <typename T>
spMtx<T> eWiseMult(const spMtx<T>& A, const spMtx<T>& B) {
  if (A.m != B.m || A.n != B.n) throw -1;
  return graphalgo_detail::ewise(GA_EW_MULT, A, B, nullptr);
}

// the intersection restricted to M's pattern
template <typename T, typename U>
spMtx<T> eWiseMult(const spMtx<T>& A, const spMtx<T>& B, const spMtx<U>& M) {
  if (A.m != B.m || A.n != B.n) throw -1;
  const ga_csr mm = M.view(false);
  return graphalgo_detail::ewise(GA_EW_MULT, A, B, &mm);
}

// A's values times alpha, returned as a new matrix (the operand is not
// modified, unlike the reference's write through a const reference)
template <typename MatrixValT, typename ScalarT>
spMtx<MatrixValT> multScalar(const spMtx<MatrixValT>& A, const ScalarT& alpha) {
  graphalgo_detail::ensure_init();
  spMtx<MatrixValT> C = A;
  const ga_csr a = A.view(true);
  if (A.nz) graphalgo_detail::check(ga_scale(&a, graphalgo_detail::dtype_of<MatrixValT>(), (double)alpha, C.Val));
  return C;
}

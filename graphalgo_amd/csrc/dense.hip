// Dense element-wise operators of the reference's BC backward sweep
// (include/matrix_la.h:192-219): C += A .* B and C<M> = A .* B on row-major
// m x n arrays.  Both are streaming kernels bound by HBM bandwidth: the
// fused multiply-add reads A, B, C and writes C (32 B per FP64 value); the
// masked product reads M's row, A and B at M's entries and writes every C
// value once (zeros outside M), so C is never read.
//
// Products and sums are rounded separately (no contraction into fused
// multiply-adds), which is how the reference's scalar loops round them.
#include "common.hpp"

namespace ga {

template <typename T>
__global__ __launch_bounds__(256) void k_fuse_mult_add(const T* __restrict__ A, const T* __restrict__ B,
                                                       T* __restrict__ C, int64_t count) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    C[i] = C[i] + A[i] * B[i];
}

// One wavefront per row: every column of the row is written once, with
// A * B where M's row holds the column (binary search in the sorted row)
// and zero elsewhere.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_mult_masked(const int32_t* __restrict__ Mrp,
                                                           const int32_t* __restrict__ Mcol,
                                                           const T* __restrict__ A, const T* __restrict__ B,
                                                           T* __restrict__ C, int64_t m, int64_t n) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * wpb) {
    const int mb = Mrp[i], me = Mrp[i + 1];
    const int64_t row = i * n;
    for (int64_t c = lane; c < n; c += 64) {
      int lo = mb, hi = me;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Mcol[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      T v = T(0);
      if (lo < me && Mcol[lo] == c) v = A[row + c] * B[row + c];
      C[row + c] = v;
    }
  }
}

template <typename T>
void fuse_mult_add_run(int64_t count, const void* A, const void* B, void* C) {
  hipStream_t s = ctx().stream;
  DevBuf dA, dB, dC;
  const size_t bytes = sizeof(T) * count;
  dA.ensure(bytes);
  dB.ensure(bytes);
  dC.ensure(bytes);
  GA_HIP(hipMemcpyAsync(dA.p, A, bytes, hipMemcpyHostToDevice, s));
  GA_HIP(hipMemcpyAsync(dB.p, B, bytes, hipMemcpyHostToDevice, s));
  GA_HIP(hipMemcpyAsync(dC.p, C, bytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fuse_mult_add<T>, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, dA.as<T>(),
                     dB.as<T>(), dC.as<T>(), count);
  GA_HIP(hipGetLastError());
  GA_HIP(hipMemcpyAsync(C, dC.p, bytes, hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
}

template <typename T>
void dense_mult_masked_run(const ga_csr* M, const void* A, const void* B, void* C) {
  hipStream_t s = ctx().stream;
  const int64_t m = M->nrows, n = M->ncols, count = m * n;
  OwnedCsr dM;
  ga_csr pat = *M;
  pat.val = nullptr;
  upload_csr(&pat, GA_INT32, dM, s);
  DevBuf dA, dB, dC;
  const size_t bytes = sizeof(T) * count;
  dA.ensure(bytes);
  dB.ensure(bytes);
  dC.ensure(bytes);
  GA_HIP(hipMemcpyAsync(dA.p, A, bytes, hipMemcpyHostToDevice, s));
  GA_HIP(hipMemcpyAsync(dB.p, B, bytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_dense_mult_masked<T>, dim3(grid_for(m, 4, 16384)), dim3(256), 0, s, dM.rp.as<int32_t>(),
                     dM.col.as<int32_t>(), dA.as<T>(), dB.as<T>(), dC.as<T>(), m, n);
  GA_HIP(hipGetLastError());
  GA_HIP(hipMemcpyAsync(C, dC.p, bytes, hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
}

}  // namespace ga

using namespace ga;

extern "C" {

ga_status ga_dense_fuse_mult_add(int64_t count, ga_dtype dtype, const void* A, const void* B, void* C) {
  return guarded([&] {
    dtype_size(dtype);
    GA_REQUIRE(count >= 0, GA_ERR_ARG, "negative count");
    if (count == 0) return;
    GA_REQUIRE(A && B && C, GA_ERR_ARG, "null buffer");
    switch (dtype) {
      case GA_INT32: fuse_mult_add_run<int32_t>(count, A, B, C); break;
      case GA_INT64: fuse_mult_add_run<int64_t>(count, A, B, C); break;
      case GA_FLOAT32: fuse_mult_add_run<float>(count, A, B, C); break;
      default: fuse_mult_add_run<double>(count, A, B, C); break;
    }
  });
}

ga_status ga_dense_mult_masked(const ga_csr* M, ga_dtype dtype, const void* A, const void* B, void* C) {
  return guarded([&] {
    dtype_size(dtype);
    validate_host_csr(M, "M");
    if (M->nrows == 0 || M->ncols == 0) return;
    GA_REQUIRE(A && B && C, GA_ERR_ARG, "null buffer");
    switch (dtype) {
      case GA_INT32: dense_mult_masked_run<int32_t>(M, A, B, C); break;
      case GA_INT64: dense_mult_masked_run<int64_t>(M, A, B, C); break;
      case GA_FLOAT32: dense_mult_masked_run<float>(M, A, B, C); break;
      default: dense_mult_masked_run<double>(M, A, B, C); break;
    }
  });
}

}  // extern "C"

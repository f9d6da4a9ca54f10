// The reference's two other masked products, as general GPU operators.
//
// mxmm_spd (include/matrix_la.h:239-270): C<M> = A*B, B and C dense
//   row-major.  One wavefront per row; each lane owns output columns and
//   writes every element of the row exactly once (value on M's pattern,
//   zero elsewhere).  A mask column's value is summed over A's row in CSR
//   order, the reference's accumulation order, so FP64 results are
//   bit-identical to it.
//
// mxmm_msa_cmask (include/matrix_la.h:788-891): C<!M> = A*B with sorted
//   output columns, in the reference's two phases (symbolic count, prefix
//   sum, numeric fill).  One workgroup per row.  Rows whose product count
//   fits CM_HCAP use an LDS hash table (column -> accumulator) whose
//   surviving keys are bitonic-sorted in LDS before the write; longer rows
//   use a dense per-workgroup accumulator in global memory (the reference's
//   MSA layout) scanned in column order.  Accumulation order inside a row
//   follows the arrival of products, so FP64 values can differ from the
//   reference's left-to-right sums in the last bits; integer values are
//   exact.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>

namespace ga {

// ---- mxmm_spd ---------------------------------------------------------------
// Products and sums rounded separately (no contraction into fused
// multiply-adds), as the reference's scalar loop rounds them.
template <typename T>
__global__ __launch_bounds__(256) void k_spd(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol,
                                             const T* __restrict__ Aval, const T* __restrict__ B, int64_t ncols,
                                             const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol,
                                             T* __restrict__ C, int64_t m) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * wpb) {
    const int mb = Mrp[i], me = Mrp[i + 1];
    const int ab = Arp[i], ae = Arp[i + 1];
    T* crow = C + i * ncols;
    for (int64_t c = lane; c < ncols; c += 64) {
      int lo = mb, hi = me;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Mcol[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      T acc = T(0);
      if (lo < me && Mcol[lo] == c) {
        for (int t = ab; t < ae; ++t) acc += (Aval ? Aval[t] : T(1)) * B[(int64_t)Acol[t] * ncols + c];
      }
      crow[c] = acc;
    }
  }
}

// ---- mxmm_msa_cmask ---------------------------------------------------------
constexpr int CM_THREADS = 256;
constexpr int CM_HCAP = 2048;          // products per row handled in LDS
constexpr int CM_HSIZE = 2 * CM_HCAP;  // hash slots (power of two)
constexpr int32_t CM_EMPTY = -1;

__device__ inline bool member(const int32_t* __restrict__ Mcol, int mb, int me, int c) {
  int lo = mb, hi = me;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (Mcol[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return lo < me && Mcol[lo] == c;
}

// phase 0: per-row product counts (flops) to route rows
__global__ void k_cm_flops(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol,
                           const int32_t* __restrict__ Brp, int64_t m, int64_t* __restrict__ flops) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * wpb) {
    long long f = 0;
    for (int t = Arp[i] + lane; t < Arp[i + 1]; t += 64) f += Brp[Acol[t] + 1] - Brp[Acol[t]];
    f = wave_reduce_sum(f);
    if (lane == 0) flops[i] = f;
  }
}

// phase 1 (count) / phase 2 (fill) for rows with flops <= CM_HCAP
template <typename T, bool FILL>
__global__ __launch_bounds__(CM_THREADS) void k_cm_small(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
    const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol, const T* __restrict__ Bval,
    const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol, const int64_t* __restrict__ flops, int64_t m,
    int32_t* __restrict__ cnt, const int32_t* __restrict__ Crp, int32_t* __restrict__ Ccol, T* __restrict__ Cval) {
  __shared__ int32_t s_key[CM_HSIZE];
  __shared__ T s_val[FILL ? CM_HSIZE : 1];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  for (int64_t i = blockIdx.x; i < m; i += gridDim.x) {
    if (flops[i] > CM_HCAP) continue;  // uniform per workgroup
    for (int q = tid; q < CM_HSIZE; q += CM_THREADS) {
      s_key[q] = CM_EMPTY;
      if (FILL) s_val[q] = T(0);
    }
    if (tid == 0) s_n = 0;
    __syncthreads();
    const int ab = Arp[i], ae = Arp[i + 1], mb = Mrp[i], me = Mrp[i + 1];
    // one thread per A entry; the row has at most CM_HCAP products
    for (int t = ab + tid; t < ae; t += CM_THREADS) {
      const int k = Acol[t];
      const T a = Aval ? Aval[t] : T(1);
      for (int j = Brp[k]; j < Brp[k + 1]; ++j) {
        const int c = Bcol[j];
        if (member(Mcol, mb, me, c)) continue;
        unsigned h = ((unsigned)c * 2654435761u) & (CM_HSIZE - 1);
        while (true) {
          const int32_t prev = atomicCAS(&s_key[h], CM_EMPTY, c);
          if (prev == CM_EMPTY || prev == c) {
            if (FILL) atomic_add_val(&s_val[h], a * (Bval ? Bval[j] : T(1)));
            if (prev == CM_EMPTY) atomicAdd(&s_n, 1);
            break;
          }
          h = (h + 1) & (CM_HSIZE - 1);
        }
      }
    }
    __syncthreads();
    const int nk = s_n;
    if (!FILL) {
      if (tid == 0) cnt[i] = nk;
    } else {
      // bitonic sort of the slots by key (empty = INT_MAX sinks to the end)
      for (int q = tid; q < CM_HSIZE; q += CM_THREADS)
        if (s_key[q] == CM_EMPTY) s_key[q] = 0x7FFFFFFF;
      __syncthreads();
      for (int size = 2; size <= CM_HSIZE; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
          for (int q = tid; q < CM_HSIZE / 2; q += CM_THREADS) {
            const int lo = 2 * q - (q & (stride - 1));
            const int hi = lo + stride;
            const bool up = ((lo & size) == 0);
            const int32_t kl = s_key[lo], kh = s_key[hi];
            if ((kl > kh) == up) {
              s_key[lo] = kh;
              s_key[hi] = kl;
              const T tv = s_val[lo];
              s_val[lo] = s_val[hi];
              s_val[hi] = tv;
            }
          }
          __syncthreads();
        }
      }
      const int base = Crp[i];
      for (int q = tid; q < nk; q += CM_THREADS) {
        Ccol[base + q] = s_key[q];
        Cval[base + q] = s_val[q];
      }
    }
    __syncthreads();
  }
}

// rows with flops > CM_HCAP: dense accumulator per workgroup in global memory
template <typename T, bool FILL>
__global__ __launch_bounds__(CM_THREADS) void k_cm_big(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
    const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol, const T* __restrict__ Bval,
    const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol, const int64_t* __restrict__ flops, int64_t m,
    int64_t ncols, uint32_t* __restrict__ flag_ws, T* __restrict__ val_ws, int32_t* __restrict__ cnt,
    const int32_t* __restrict__ Crp, int32_t* __restrict__ Ccol, T* __restrict__ Cval) {
  __shared__ int s_run[CM_THREADS / 64];
  __shared__ int s_base;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t* flag = flag_ws + (int64_t)blockIdx.x * ncols;
  T* val = val_ws + (int64_t)blockIdx.x * ncols;
  for (int64_t c = tid; c < ncols; c += CM_THREADS) {
    flag[c] = 0u;
    if (FILL) val[c] = T(0);
  }
  __syncthreads();
  for (int64_t i = blockIdx.x; i < m; i += gridDim.x) {
    if (flops[i] <= CM_HCAP) continue;
    const int ab = Arp[i], ae = Arp[i + 1], mb = Mrp[i], me = Mrp[i + 1];
    for (int t = ab; t < ae; ++t) {
      const int k = Acol[t];
      const T a = Aval ? Aval[t] : T(1);
      for (int j = Brp[k] + tid; j < Brp[k + 1]; j += CM_THREADS) {
        const int c = Bcol[j];
        if (member(Mcol, mb, me, c)) continue;
        flag[c] = 1u;
        if (FILL) atomic_add_val(&val[c], a * (Bval ? Bval[j] : T(1)));
      }
    }
    __threadfence_block();
    __syncthreads();
    // ordered scan over the columns: emit (and clear) set columns
    if (tid == 0) s_base = FILL ? Crp[i] : 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < ncols; c0 += CM_THREADS) {
      const int64_t c = c0 + tid;
      const bool set = c < ncols && flag[c] != 0u;
      const unsigned long long bal = __ballot(set);
      const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
      if (lane == 0) s_run[wv] = __popcll(bal);
      __syncthreads();
      int off = s_base;
      for (int k2 = 0; k2 < wv; ++k2) off += s_run[k2];
      if (set) {
        if (FILL) {
          const int p = off + __popcll(bal & lt);
          Ccol[p] = (int32_t)c;
          Cval[p] = val[c];
          val[c] = T(0);
        }
        flag[c] = 0u;
      }
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int k2 = 0; k2 < CM_THREADS / 64; ++k2) tot += s_run[k2];
        s_base += tot;
      }
      __syncthreads();
    }
    if (!FILL && tid == 0) cnt[i] = s_base;
    __syncthreads();
  }
}

template <typename T>
static void cmask_run(const DCsr& A, const DCsr& B, const DCsr& M, int32_t* Crp_h, int32_t** Ccol_h, void** Cval_h,
                      int64_t* Cnnz, hipStream_t s) {
  const int64_t m = A.nrows, ncols = B.ncols;
  DevBuf flops, cnt, rp, tmp, fws, vws, ccol, cval;
  flops.ensure(sizeof(int64_t) * std::max<int64_t>(m, 1));
  cnt.ensure(sizeof(int32_t) * std::max<int64_t>(m, 1));
  rp.ensure(sizeof(int32_t) * (m + 1));
  fill_u32(cnt.p, 0u, m, s);
  hipLaunchKernelGGL(k_cm_flops, dim3(grid_for(m, 4, 16384)), dim3(256), 0, s, A.rp, A.col, B.rp, m,
                     flops.as<int64_t>());
  const unsigned gsmall = grid_for(m, 1, 8192);
  const unsigned gbig = 64;
  fws.ensure(sizeof(uint32_t) * gbig * std::max<int64_t>(ncols, 1));
  vws.ensure(sizeof(T) * gbig * std::max<int64_t>(ncols, 1));
  const T* Av = static_cast<const T*>(A.val);
  const T* Bv = static_cast<const T*>(B.val);
  hipLaunchKernelGGL((k_cm_small<T, false>), dim3(gsmall), dim3(CM_THREADS), 0, s, A.rp, A.col, Av, B.rp, B.col, Bv,
                     M.rp, M.col, (const int64_t*)flops.as<int64_t>(), m, cnt.as<int32_t>(), nullptr, nullptr,
                     nullptr);
  hipLaunchKernelGGL((k_cm_big<T, false>), dim3(gbig), dim3(CM_THREADS), 0, s, A.rp, A.col, Av, B.rp, B.col, Bv,
                     M.rp, M.col, (const int64_t*)flops.as<int64_t>(), m, ncols, fws.as<uint32_t>(), vws.as<T>(),
                     cnt.as<int32_t>(), nullptr, nullptr, nullptr);
  GA_HIP(hipGetLastError());
  scan_exclusive_i32(cnt.as<int32_t>(), rp.as<int32_t>(), m, tmp, s);
  GA_HIP(hipMemcpyAsync(Crp_h, rp.p, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  const int64_t nz = Crp_h[m];
  *Cnnz = nz;
  ccol.ensure(sizeof(int32_t) * std::max<int64_t>(nz, 1));
  cval.ensure(sizeof(T) * std::max<int64_t>(nz, 1));
  hipLaunchKernelGGL((k_cm_small<T, true>), dim3(gsmall), dim3(CM_THREADS), 0, s, A.rp, A.col, Av, B.rp, B.col, Bv,
                     M.rp, M.col, (const int64_t*)flops.as<int64_t>(), m, nullptr,
                     (const int32_t*)rp.as<int32_t>(), ccol.as<int32_t>(), cval.as<T>());
  hipLaunchKernelGGL((k_cm_big<T, true>), dim3(gbig), dim3(CM_THREADS), 0, s, A.rp, A.col, Av, B.rp, B.col, Bv,
                     M.rp, M.col, (const int64_t*)flops.as<int64_t>(), m, ncols, fws.as<uint32_t>(), vws.as<T>(),
                     nullptr, (const int32_t*)rp.as<int32_t>(), ccol.as<int32_t>(), cval.as<T>());
  GA_HIP(hipGetLastError());
  *Ccol_h = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<int64_t>(nz, 1)));
  *Cval_h = std::malloc(sizeof(T) * std::max<int64_t>(nz, 1));
  GA_REQUIRE(*Ccol_h && *Cval_h, GA_ERR_OOM, "host allocation failed");
  if (nz) {
    GA_HIP(hipMemcpyAsync(*Ccol_h, ccol.p, sizeof(int32_t) * nz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipMemcpyAsync(*Cval_h, cval.p, sizeof(T) * nz, hipMemcpyDeviceToHost, s));
  }
  GA_HIP(hipStreamSynchronize(s));
}

template <typename T>
void spd_run(const ga_csr* A, const T* B, int64_t ncols, const ga_csr* M, T* C, int dtype) {
  validate_host_csr(A, "A");
  validate_host_csr(M, "M");
  GA_REQUIRE(ncols >= 0 && M->ncols == ncols && A->nrows == M->nrows, GA_ERR_DIM,
             "shapes: A m x k, B k x ncols, M m x ncols");
  GA_REQUIRE((A->ncols == 0 || ncols == 0 || B) && (A->nrows == 0 || ncols == 0 || C), GA_ERR_ARG, "null buffer");
  hipStream_t s = ctx().stream;
  OwnedCsr dA, dM;
  upload_csr(A, dtype, dA, s);
  upload_csr(M, dtype, dM, s);
  const int64_t m = A->nrows, k = A->ncols;
  DevBuf dB, dC;
  dB.ensure(sizeof(T) * std::max<int64_t>(k * ncols, 1));
  dC.ensure(sizeof(T) * std::max<int64_t>(m * ncols, 1));
  if (k * ncols) GA_HIP(hipMemcpyAsync(dB.p, B, sizeof(T) * k * ncols, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_spd<T>, dim3(grid_for(m, 4, 16384)), dim3(256), 0, s, dA.rp.as<int32_t>(),
                     dA.col.as<int32_t>(), (const T*)dA.val.p, (const T*)dB.as<T>(), ncols, dM.rp.as<int32_t>(),
                     dM.col.as<int32_t>(), dC.as<T>(), m);
  GA_HIP(hipGetLastError());
  if (m * ncols) GA_HIP(hipMemcpyAsync(C, dC.p, sizeof(T) * m * ncols, hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
}

}  // namespace ga

using namespace ga;

extern "C" {

ga_status ga_mxm_spd(const ga_csr* A, const double* B, int64_t ncols, const ga_csr* M, double* C) {
  return guarded([&] { spd_run<double>(A, B, ncols, M, C, GA_FLOAT64); });
}

ga_status ga_mxm_spd_f32(const ga_csr* A, const float* B, int64_t ncols, const ga_csr* M, float* C) {
  return guarded([&] { spd_run<float>(A, B, ncols, M, C, GA_FLOAT32); });
}

ga_status ga_mxm_cmask(const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, int32_t** Crowptr,
                       int32_t** Ccol, void** Cval, int64_t* Cnnz) {
  return guarded([&] {
    dtype_size(dtype);
    validate_host_csr(A, "A");
    validate_host_csr(B, "B");
    validate_host_csr(M, "M");
    GA_REQUIRE(A->nrows == M->nrows && A->ncols == B->nrows && B->ncols == M->ncols, GA_ERR_DIM,
               "operand shapes do not match (A: m x k, B: k x n, M: m x n)");
    GA_REQUIRE(Crowptr && Ccol && Cval && Cnnz, GA_ERR_ARG, "null output");
    hipStream_t s = ctx().stream;
    OwnedCsr dA, dB, dM;
    upload_csr(A, dtype, dA, s);
    upload_csr(B, dtype, dB, s);
    upload_csr(M, dtype, dM, s);
    DCsr vM = dM.view();
    vM.val = nullptr;
    int32_t* rp = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (A->nrows + 1)));
    GA_REQUIRE(rp != nullptr, GA_ERR_OOM, "host allocation failed");
    *Ccol = nullptr;
    *Cval = nullptr;
    try {
      switch (dtype) {
        case GA_INT32: cmask_run<int32_t>(dA.view(), dB.view(), vM, rp, Ccol, Cval, Cnnz, s); break;
        case GA_INT64: cmask_run<int64_t>(dA.view(), dB.view(), vM, rp, Ccol, Cval, Cnnz, s); break;
        case GA_FLOAT32: cmask_run<float>(dA.view(), dB.view(), vM, rp, Ccol, Cval, Cnnz, s); break;
        default: cmask_run<double>(dA.view(), dB.view(), vM, rp, Ccol, Cval, Cnnz, s); break;
      }
    } catch (...) {
      std::free(rp);
      std::free(*Ccol);
      std::free(*Cval);
      *Ccol = nullptr;
      *Cval = nullptr;
      throw;
    }
    *Crowptr = rp;
  });
}

}  // extern "C"

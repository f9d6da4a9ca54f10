// Masked SpGEMM C<M> = A*B with a structural mask (C has M's pattern).
//
// Row-wise Gustavson product as in the reference's MSA path
// (_mxmm_msa_parallel, include/matrix_la.h:715-749), re-laid out for CDNA4:
//
//  * Work items.  Row i is split into items of at most MXM_TCAP entries of
//    A's row (hub rows become many items, short rows one).  A persistent grid
//    sized to the chip's occupancy pulls items from an atomic counter, so the
//    skewed R-MAT rows balance dynamically instead of through static chunks.
//  * Flattened wedge space.  Inside an item the B rows selected by A's
//    entries are concatenated (LDS prefix sum of their lengths) and the 256
//    threads stride over that wedge space, so consecutive lanes read
//    consecutive B columns (coalesced) whether B rows are long or short.
//  * Mask membership.  M's row is staged in LDS as a sorted column list plus
//    an n-bit bitmap (for n <= 2^18 columns); a wedge whose column misses the
//    bitmap costs one LDS read.  Hits locate their mask position by binary
//    search in the LDS list and accumulate into a mask-indexed LDS
//    accumulator (the reference's MCA layout, include/matrix_la.h:543-569),
//    flushed once per item.  Mask rows longer than MXM_MCAP entries search
//    M in global memory and accumulate with global atomics.
//  * Integer accumulation is exact in any order, so results are bit-identical
//    to the CPU path; FP64 sums of rows split across items are summed in
//    arrival order.
#include "common.hpp"

#include <algorithm>

namespace ga {

constexpr int64_t MXM_BITMAP_MAX_COLS = int64_t(1) << 18;  // 32 KB bitmap

// THREADS = workgroup size = A entries per item (one entry per thread in the
// entry scan); MCAP = mask entries staged in LDS.  BATCHED: the operands are
// many graphs stored back to back with graph-local column indices; row i
// belongs to graph row_group[i] whose rows start at gbase[row_group[i]].
template <typename T, bool PAT_A, bool PAT_B, int MODE, int THREADS, bool BATCHED>
__global__ __launch_bounds__(THREADS) void k_mxm_items(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
    const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol, const T* __restrict__ Bval,
    const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol, T* __restrict__ Cval,
    unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ item_row, const int32_t* __restrict__ item_t0,
    const int32_t* __restrict__ n_items_ptr, unsigned int* __restrict__ work_counter,
    unsigned long long* __restrict__ wedge_counter, int bitmap_words) {
  constexpr int MXM_THREADS = THREADS;
  constexpr int MXM_TCAP = THREADS;
  constexpr int MXM_MCAP = 8 * THREADS;
  extern __shared__ uint32_t s_bitmap[];
  __shared__ int32_t s_mcol[MXM_MCAP];
  __shared__ T s_acc[MODE == 0 ? MXM_MCAP : 1];
  __shared__ int32_t s_ebeg[MXM_TCAP];
  __shared__ long long s_eoff[MXM_TCAP + 1];
  __shared__ T s_aval[PAT_A ? 1 : MXM_TCAP];
  __shared__ long long s_wsum[MXM_THREADS / 64];
  __shared__ int s_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const bool use_bm = bitmap_words > 0;
  for (int q = tid; q < bitmap_words; q += MXM_THREADS) s_bitmap[q] = 0u;
  const int n_items = *n_items_ptr;

  while (true) {
    __syncthreads();  // previous item fully retired (LDS reuse)
    if (tid == 0) s_item = (int)atomicAdd(work_counter, 1u);
    __syncthreads();
    const int q = s_item;
    if (q >= n_items) break;

    const int i = item_row[q];
    const int t0 = item_t0[q];
    const int a_beg = Arp[i];
    const int a_end = Arp[i + 1];
    const int t1 = min(t0 + MXM_TCAP, a_end);
    const bool multi = (a_end - a_beg) > MXM_TCAP;
    const int ms = Mrp[i];
    const int mnz = Mrp[i + 1] - ms;
    const bool local = mnz <= MXM_MCAP;
    const int bbase = BATCHED ? gbase[row_group[i]] : 0;

    // stage the mask row
    for (int p = tid; p < mnz; p += MXM_THREADS) {
      const int c = Mcol[ms + p];
      if (local) {
        s_mcol[p] = c;
        if (MODE == 0) s_acc[p] = T(0);
      }
      if (use_bm) atomicOr(&s_bitmap[c >> 5], 1u << (c & 31));
    }
    // stage the item's entries: B row start and length, A value
    const int ne = t1 - t0;
    int len = 0;
    if (tid < ne) {
      const int k = bbase + Acol[t0 + tid];
      const int b0 = Brp[k];
      len = Brp[k + 1] - b0;
      s_ebeg[tid] = b0;
      if (!PAT_A) s_aval[tid] = Aval[t0 + tid];
    }
    const long long incl = wave_incl_scan((long long)len);
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    long long base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < MXM_THREADS / 64; ++k) {
      if (k < wv) base += s_wsum[k];
      total += s_wsum[k];
    }
    if (tid < ne) s_eoff[tid] = base + incl - len;
    if (tid == 0) s_eoff[ne] = total;
    __syncthreads();

    // stride over the flattened wedge space of the item
    long long lsum = 0;
    int e = 0;
    for (long long w = tid; w < total; w += MXM_THREADS) {
      if (s_eoff[e + 1] <= w) {  // advance to the entry holding wedge w
        int lo = e + 1, hi = ne - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_eoff[mid] <= w) lo = mid;
          else hi = mid - 1;
        }
        e = lo;
      }
      const int j = s_ebeg[e] + (int)(w - s_eoff[e]);
      const int c = Bcol[j];
      if (use_bm && !((s_bitmap[c >> 5] >> (c & 31)) & 1u)) continue;
      int lo = 0, hi = mnz;
      bool found;
      if (local) {
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_mcol[mid] < c) lo = mid + 1;
          else hi = mid;
        }
        found = lo < mnz && s_mcol[lo] == c;
      } else {
        const int32_t* mrow = Mcol + ms;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (mrow[mid] < c) lo = mid + 1;
          else hi = mid;
        }
        found = lo < mnz && mrow[lo] == c;
      }
      if (found) {
        const T v = (PAT_A ? T(1) : s_aval[e]) * (PAT_B ? T(1) : Bval[j]);
        if (MODE == 1) lsum += (long long)v;
        else if (local) atomic_add_val(&s_acc[lo], v);
        else atomic_add_val(&Cval[ms + lo], v);
      }
    }
    __syncthreads();

    if (MODE == 0) {
      if (local) {
        for (int p = tid; p < mnz; p += MXM_THREADS) {
          const T v = s_acc[p];
          if (multi) {
            if (v != T(0)) atomic_add_val(&Cval[ms + p], v);
          } else {
            Cval[ms + p] = v;
          }
        }
      }
    } else {
      const long long r = wave_reduce_sum(lsum);
      __syncthreads();
      if (lane == 0) s_wsum[wv] = r;
      __syncthreads();
      if (tid == 0) {
        long long t = 0;
#pragma unroll
        for (int k = 0; k < MXM_THREADS / 64; ++k) t += s_wsum[k];
        if (t) atomicAdd(&sums[row_group ? row_group[i] : 0], (unsigned long long)t);
      }
    }
    if (use_bm) {
      for (int p = tid; p < mnz; p += MXM_THREADS) {
        const int c = local ? s_mcol[p] : Mcol[ms + p];
        s_bitmap[c >> 5] = 0u;
      }
    }
    if (tid == 0 && wedge_counter) atomicAdd(wedge_counter, (unsigned long long)total);
  }
}

__global__ void k_item_count(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Mrp, int64_t m,
                             int tcap, int32_t* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = Arp[i + 1] - Arp[i];
    const int mz = Mrp[i + 1] - Mrp[i];
    cnt[i] = (a > 0 && mz > 0) ? (a + tcap - 1) / tcap : 0;
  }
}

__global__ void k_item_fill(const int32_t* __restrict__ Arp, const int32_t* __restrict__ cnt,
                            const int32_t* __restrict__ start, int64_t m, int tcap, int32_t* __restrict__ item_row,
                            int32_t* __restrict__ item_t0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = cnt[i], s0 = start[i], a0 = Arp[i];
    for (int q = 0; q < c; ++q) {
      item_row[s0 + q] = (int32_t)i;
      item_t0[s0 + q] = a0 + q * tcap;
    }
  }
}

template <typename T, bool PA, bool PB, int MODE, int THREADS, bool BATCHED>
static void launch_items(const DCsr& A, const DCsr& B, const DCsr& M, void* Cval, unsigned long long* sums,
                         const int32_t* row_group, const int32_t* gbase, MxmWork& w, int bitmap_words,
                         hipStream_t s) {
  constexpr int MXM_THREADS = THREADS;
  constexpr int MXM_TCAP = THREADS;
  auto kern = k_mxm_items<T, PA, PB, MODE, THREADS, BATCHED>;
  const size_t dyn = sizeof(uint32_t) * (size_t)bitmap_words;
  int occ = 0;
  GA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, MXM_THREADS, dyn));
  if (occ < 1) occ = 1;
  Context& c = ctx();
  const int64_t max_items = A.nrows + ceil_div(A.nnz, MXM_TCAP) + 1;
  int64_t grid = (int64_t)c.num_cu * occ;
  if (grid > max_items) grid = max_items;
  if (grid < 1) grid = 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c.prof) {
    auto take = [&]() {
      hipEvent_t ev;
      if (!c.event_pool.empty()) {
        ev = c.event_pool.back();
        c.event_pool.pop_back();
      } else {
        GA_HIP(hipEventCreate(&ev));
      }
      return ev;
    };
    e0 = take();
    e1 = take();
    GA_HIP(hipEventRecord(e0, s));
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(MXM_THREADS), dyn, s, A.rp, A.col,
                     static_cast<const T*>(A.val), B.rp, B.col, static_cast<const T*>(B.val), M.rp, M.col,
                     static_cast<T*>(Cval), sums, row_group, gbase, w.item_row.as<int32_t>(), w.item_t0.as<int32_t>(),
                     (const int32_t*)(w.start.as<int32_t>() + A.nrows), w.counter.as<unsigned int>(),
                     c.prof ? c.prof_dev.as<unsigned long long>() : nullptr, bitmap_words);
  GA_HIP(hipGetLastError());
  if (c.prof) {
    GA_HIP(hipEventRecord(e1, s));
    c.prof_events.emplace_back(e0, e1);
    c.prof_launches += 1;
    // host-known algorithmic bytes: A and M streamed, B row bounds, output
    const int64_t out = MODE == 0 ? (int64_t)sizeof(T) * M.nnz : 8;
    c.prof_bytes_host += 4 * (A.nrows + 1) + 4 * A.nnz + 8 * A.nnz + 4 * (M.nrows + 1) + 4 * M.nnz + out;
  }
}

template <typename T, int MODE>
static void dispatch_pat(const DCsr& A, const DCsr& B, const DCsr& M, void* Cval, unsigned long long* sums,
                         const int32_t* row_group, MxmWork& w, int bm, hipStream_t s) {
  const bool pa = A.val == nullptr, pb = B.val == nullptr;
  if (pa && pb) launch_items<T, true, true, MODE, 256, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s);
  else if (pa) launch_items<T, true, false, MODE, 256, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s);
  else if (pb) launch_items<T, false, true, MODE, 256, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s);
  else launch_items<T, false, false, MODE, 256, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s);
}

static void build_items(const DCsr& A, const DCsr& M, int tcap, MxmWork& w, hipStream_t s) {
  const int64_t m = A.nrows;
  const int64_t max_items = m + ceil_div(A.nnz, tcap) + 1;
  GA_REQUIRE(max_items < INT32_MAX, GA_ERR_LIMIT, "too many work items");
  w.cnt.ensure(sizeof(int32_t) * m);
  w.start.ensure(sizeof(int32_t) * (m + 1));
  w.item_row.ensure(sizeof(int32_t) * max_items);
  w.item_t0.ensure(sizeof(int32_t) * max_items);
  w.counter.ensure(256);
  fill_u32(w.counter.p, 0u, 1, s);
  hipLaunchKernelGGL(k_item_count, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, A.rp, M.rp, m, tcap,
                     w.cnt.as<int32_t>());
  GA_HIP(hipGetLastError());
  scan_exclusive_i32(w.cnt.as<int32_t>(), w.start.as<int32_t>(), m, w.scan_tmp, s);
  hipLaunchKernelGGL(k_item_fill, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, A.rp,
                     (const int32_t*)w.cnt.as<int32_t>(), (const int32_t*)w.start.as<int32_t>(), m, tcap,
                     w.item_row.as<int32_t>(), w.item_t0.as<int32_t>());
  GA_HIP(hipGetLastError());
}

void mxm_batched_sum_dev(const DCsr& L, const int32_t* row_group, const int32_t* gbase, int64_t max_cols,
                         int threads, unsigned long long* sums, MxmWork& w, hipStream_t s) {
  if (L.nrows == 0 || L.nnz == 0) return;
  build_items(L, L, threads, w, s);
  const int bm = (max_cols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(max_cols, 32) : 0;
  switch (threads) {
    case 64:
      launch_items<int64_t, true, true, 1, 64, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s);
      break;
    case 128:
      launch_items<int64_t, true, true, 1, 128, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s);
      break;
    case 256:
      launch_items<int64_t, true, true, 1, 256, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s);
      break;
    default: throw Error{GA_ERR_ARG, "workgroup size must be 64, 128 or 256"};
  }
}

void mxm_masked_dev(const DCsr& A, const DCsr& B, const DCsr& M, int dtype, int mode, void* Cval,
                    unsigned long long* sums, const int32_t* row_group, MxmWork& w, hipStream_t s) {
  GA_REQUIRE(A.nrows == M.nrows, GA_ERR_DIM, "A.nrows != M.nrows");
  GA_REQUIRE(A.ncols == B.nrows, GA_ERR_DIM, "A.ncols != B.nrows");
  GA_REQUIRE(B.ncols == M.ncols, GA_ERR_DIM, "B.ncols != M.ncols");
  const int64_t m = A.nrows;
  if (mode == 0 && M.nnz > 0) {
    const size_t es = dtype_size(dtype);
    if (es == 4) fill_u32(Cval, 0u, M.nnz, s);
    else fill_u64(Cval, 0ull, M.nnz, s);
  }
  if (m == 0 || A.nnz == 0 || M.nnz == 0) return;
  build_items(A, M, 256, w, s);
  const int bm = (M.ncols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(M.ncols, 32) : 0;
  if (mode == 1) {
    dispatch_pat<int64_t, 1>(A, B, M, nullptr, sums, row_group, w, bm, s);
    return;
  }
  switch (dtype) {
    case GA_INT32: dispatch_pat<int32_t, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    case GA_INT64: dispatch_pat<int64_t, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    case GA_FLOAT64: dispatch_pat<double, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    default: throw Error{GA_ERR_ARG, "unknown dtype"};
  }
}

}  // namespace ga

using namespace ga;

extern "C" ga_status ga_mxm_masked(const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, void* Cval) {
  return guarded([&] {
    dtype_size(dtype);
    validate_host_csr(A, "A");
    validate_host_csr(B, "B");
    validate_host_csr(M, "M");
    GA_REQUIRE(A->nrows == M->nrows && A->ncols == B->nrows && B->ncols == M->ncols, GA_ERR_DIM,
               "operand shapes do not match (A: m x k, B: k x n, M: m x n)");
    GA_REQUIRE(M->nnz == 0 || Cval != nullptr, GA_ERR_ARG, "null Cval");
    Context& c = ctx();
    hipStream_t s = c.stream;
    OwnedCsr dA, dB, dM;
    upload_csr(A, dtype, dA, s);
    upload_csr(B, dtype, dB, s);
    upload_csr(M, dtype, dM, s);
    DevBuf dC;
    const size_t es = dtype_size(dtype);
    dC.ensure(es * std::max<int64_t>(M->nnz, 1));
    MxmWork w;
    DCsr vM = dM.view();
    vM.val = nullptr;
    mxm_masked_dev(dA.view(), dB.view(), vM, dtype, 0, dC.p, nullptr, nullptr, w, s);
    if (M->nnz) GA_HIP(hipMemcpyAsync(Cval, dC.p, es * M->nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
  });
}

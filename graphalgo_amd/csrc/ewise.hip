// Element-wise operators of the reference's GraphBLAS layer, with their
// defects fixed (include/matrix_la.h:272-541):
//   eWiseAdd(A, B)        C = A (+) B over the union of the patterns, values
//                         added where both hold an entry (sums of zero kept);
//   eWiseAdd(A, B, M)     C<M> = A + B on M's pattern; entries whose sum is
//                         zero are dropped (the reference keeps row[c] != 0;
//                         its second loop reads A's columns for B, fixed);
//   eWiseMult(A, B)       C = A (x) B over the intersection, values multiplied
//                         (the reference's fill loop advances the row index);
//   eWiseMult(A, B, M)    the intersection restricted to M's pattern;
//   add_nointersect(A, B) the union of two disjoint patterns (checked here;
//                         the reference asserts sorted output);
//   multScalar(A, alpha)  values times alpha in the value type (integer
//                         values: converted back toward zero, like `v *= alpha`).
//
// Entry-parallel: the driving operand's entries (A's without a mask, M's
// with one) are cut into aligned groups of 64; a wavefront takes EW_Q groups
// (EW_Q entries per lane, their binary searches in lockstep), so a hub row
// is spread over as many wavefronts as it has tiles of 64 EW_Q entries.  Each
// lane finds its entry's row in the short range of rows its group spans, and
// the entry's column in the other operand(s) by binary search.  The mark
// pass stores one bit per driving entry (A: has a match in B; M: kept) and
// the bit count of each group; an exclusive scan of the group counts gives
// P(p), the number of set bits before entry p.  Everything else follows
// from global positions, without per-row counts:
//   union:        C's row pointer i = Arp[i] + Brp[i] - P(Arp[i]);
//                 A entry t goes to t + pb - P(t) (pb = its lower bound in B);
//                 an unmatched B entry u goes to u + pa - P(pa);
//   intersection: C's row pointer i = P(Arp[i]); a matched A entry t goes to P(t);
//   masked:       C's row pointer i = P(Mrp[i]); a kept M entry t goes to P(t).
// Without a mask the mark pass also stores each A entry's lower bound in B,
// so the fill reads A's side back instead of searching again; B's entries
// (and M's, with a mask) are searched again in the fill.  C is written in
// CSR order.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace ga {

constexpr int EW_ADD = GA_EW_ADD, EW_MULT = GA_EW_MULT, EW_DISJOINT = GA_EW_ADD_DISJOINT;
constexpr int EW_Q = 8;             // entries per lane
constexpr int EW_TILE = 64 * EW_Q;  // entries per wavefront

template <typename T>
__device__ inline T ew_val(const T* __restrict__ v, int64_t j) {
  return v ? v[j] : T(1);  // pattern operands hold ones
}

// tile_row[g] = the row that holds entry 64 g: the last row r with
// rp[r] <= 64 g (one thread per group)
__global__ void k_ew_group_rows(const int32_t* __restrict__ rp, int64_t m, int64_t ngroups,
                                int32_t* __restrict__ tile_row) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ngroups; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = g * 64;
    int lo = 0, hi = (int)(m - 1);
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    tile_row[g] = lo;
  }
}

// the row of entry t: the last row r in [tile_row[g], tile_row[g + 1]] with
// rp[r] <= t (entry 64 (g + 1) lies in or after t's row)
__device__ inline int ew_row_of(const int32_t* __restrict__ rp, const int32_t* __restrict__ tile_row, int64_t ngroups,
                                int64_t m, int64_t t) {
  const int64_t g = t >> 6;
  int lo = tile_row[g];
  int hi = (g + 1 < ngroups) ? tile_row[g + 1] : (int)(m - 1);
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rp[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// lockstep lower bounds: lo[q] = first position in col[lo[q], hi[q]) with a
// column >= c[q].  One loop for all Q searches, no branch inside: a finished
// search probes col[0] (in bounds: column buffers hold at least one entry)
// and keeps its state, so the Q loads of a step are independent and issue
// back to back.  (A `while (any)` loop with a conditional step per search
// compiled here into nested loops whose exit mask dropped lanes: searches
// ended early on the GPU while the same code on the host was exact.)
template <int Q>
__device__ inline void ew_lbound_q(const int32_t* __restrict__ col, int (&lo)[Q], const int (&hi)[Q],
                                   const int (&c)[Q]) {
  int len[Q];
  int live = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    len[q] = hi[q] - lo[q];
    live |= len[q];
  }
  while (live > 0) {
    int v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = col[len[q] > 0 ? lo[q] + (len[q] >> 1) : 0];
    live = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int half = len[q] >> 1;
      const bool right = len[q] > 0 && v[q] < c[q];
      lo[q] = right ? lo[q] + half + 1 : lo[q];
      len[q] = len[q] > 0 ? (right ? len[q] - half - 1 : half) : 0;
      live |= len[q];
    }
  }
}

// number of set bits before entry p (P has ngroups + 1 entries)
__device__ inline int ew_prefix(const uint64_t* __restrict__ bits, const int32_t* __restrict__ P, int64_t p) {
  const int64_t g = p >> 6;
  const int r = (int)(p & 63);
  return P[g] + (r ? __popcll(bits[g] & ((1ull << r) - 1ull)) : 0);
}

// The driving entries of one wavefront: rows, columns and the searches.
// Unmasked: X = A, searched in B (fa = the match, pb = the lower bound).
// Masked: X = M, searched in A and B.
template <typename T, int OP, bool MASKED>
struct EwTile {
  bool valid[EW_Q];
  int col[EW_Q];
  int pa[EW_Q], pb[EW_Q];
  bool fa[EW_Q], fb[EW_Q];
  bool flag[EW_Q];  // the stored bit
  T v[EW_Q];

  __device__ inline void run(int64_t base, int lane, const int32_t* __restrict__ Xrp, const int32_t* __restrict__ Xcol,
                             const int32_t* __restrict__ Xtr, int64_t nX, int64_t ngroups, int64_t m,
                             const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol,
                             const T* __restrict__ Aval, const int32_t* __restrict__ Brp,
                             const int32_t* __restrict__ Bcol, const T* __restrict__ Bval) {
    int lo_a[EW_Q], lo_b[EW_Q], end_a[EW_Q], end_b[EW_Q];
#pragma unroll
    for (int q = 0; q < EW_Q; ++q) {
      const int64_t t = base + q * 64 + lane;
      valid[q] = t < nX;
      lo_a[q] = lo_b[q] = end_a[q] = end_b[q] = 0;
      col[q] = 0;
      if (valid[q]) {
        const int i = ew_row_of(Xrp, Xtr, ngroups, m, t);
        col[q] = Xcol[t];
        lo_b[q] = Brp[i];
        end_b[q] = Brp[i + 1];
        if (MASKED) {
          lo_a[q] = Arp[i];
          end_a[q] = Arp[i + 1];
        }
      }
    }
    ew_lbound_q<EW_Q>(Bcol, lo_b, end_b, col);
    if (MASKED) ew_lbound_q<EW_Q>(Acol, lo_a, end_a, col);
#pragma unroll
    for (int q = 0; q < EW_Q; ++q) {
      const int64_t t = base + q * 64 + lane;
      pb[q] = lo_b[q];
      fb[q] = valid[q] && lo_b[q] < end_b[q] && Bcol[lo_b[q]] == col[q];
      if (MASKED) {
        pa[q] = lo_a[q];
        fa[q] = valid[q] && lo_a[q] < end_a[q] && Acol[lo_a[q]] == col[q];
        if (OP == EW_MULT) {
          flag[q] = fa[q] && fb[q];
          v[q] = flag[q] ? ew_val(Aval, pa[q]) * ew_val(Bval, pb[q]) : T(0);
        } else {
          v[q] = (fa[q] ? ew_val(Aval, pa[q]) : T(0)) + (fb[q] ? ew_val(Bval, pb[q]) : T(0));
          flag[q] = valid[q] && v[q] != T(0);
        }
      } else {
        pa[q] = (int)t;
        fa[q] = valid[q];
        flag[q] = fb[q];
        v[q] = T(0);
      }
    }
  }
};

template <typename T, int OP, bool MASKED>
__global__ __launch_bounds__(256) void k_ew_mark(const int32_t* __restrict__ Xrp, const int32_t* __restrict__ Xcol,
                                                 const int32_t* __restrict__ Xtr, int64_t nX, int64_t ngroups,
                                                 int64_t m, const int32_t* __restrict__ Arp,
                                                 const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
                                                 const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol,
                                                 const T* __restrict__ Bval, uint64_t* __restrict__ bits,
                                                 int32_t* __restrict__ gcnt, int32_t* __restrict__ bound) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t w = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); w * EW_TILE < nX; w += nwaves) {
    const int64_t base = w * EW_TILE;
    EwTile<T, OP, MASKED> e;
    e.run(base, lane, Xrp, Xcol, Xtr, nX, ngroups, m, Arp, Acol, Aval, Brp, Bcol, Bval);
#pragma unroll
    for (int q = 0; q < EW_Q; ++q) {
      const unsigned long long b = __ballot(e.flag[q]);
      const int64_t g = (base >> 6) + q;
      if (lane == 0 && g < ngroups) {
        bits[g] = b;
        gcnt[g] = __popcll(b);
      }
      if (!MASKED && e.valid[q]) bound[base + q * 64 + lane] = e.pb[q];  // the fill reads it back
    }
  }
}

// C's row pointers (i = 0..m) from the prefix of the marks
template <int OP, bool MASKED>
__global__ void k_ew_rowptr(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Brp,
                            const int32_t* __restrict__ Mrp, int64_t m, const uint64_t* __restrict__ bits,
                            const int32_t* __restrict__ P, int32_t* __restrict__ Crp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x) {
    if (MASKED) Crp[i] = ew_prefix(bits, P, Mrp[i]);
    else if (OP == EW_MULT) Crp[i] = ew_prefix(bits, P, Arp[i]);
    else Crp[i] = Arp[i] + Brp[i] - ew_prefix(bits, P, Arp[i]);
  }
}

// Fill: wavefronts take tiles of the driving entries; for a union, tiles of
// A and of B alternate (tile j of A, then tile j of B, ...: both write the
// same region of C at about the same time), and B's entries without a match
// in A are written.
template <typename T, int OP, bool MASKED>
__global__ __launch_bounds__(256) void k_ew_fill(const int32_t* __restrict__ Xrp, const int32_t* __restrict__ Xcol,
                                                 const int32_t* __restrict__ Xtr, int64_t nX, int64_t ngroups,
                                                 int64_t m, const int32_t* __restrict__ Arp,
                                                 const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
                                                 const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol,
                                                 const T* __restrict__ Bval, const int32_t* __restrict__ Btr,
                                                 int64_t ngroupsB, const uint64_t* __restrict__ bits,
                                                 const int32_t* __restrict__ P, const int32_t* __restrict__ bound,
                                                 int32_t* __restrict__ Ccol, T* __restrict__ Cval) {
  constexpr bool UNION = !MASKED && OP != EW_MULT;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t nB = UNION ? (int64_t)Brp[m] : 0;
  const int64_t wX = (nX + EW_TILE - 1) / EW_TILE, wB = UNION ? (nB + EW_TILE - 1) / EW_TILE : 0;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t both = wX < wB ? wX : wB;
  for (int64_t w = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); w < wX + wB; w += nwaves) {
    const bool on_x = w < 2 * both ? (w & 1) == 0 : wX > wB;
    const int64_t tile = w < 2 * both ? (w >> 1) : w - both;
    if (on_x && !MASKED) {
      // A's entries: lower bounds in B and match bits from the mark pass
      const int64_t base = tile * EW_TILE;
#pragma unroll
      for (int q = 0; q < EW_Q; ++q) {
        const int64_t t = base + q * 64 + lane;
        if (t >= nX) break;
        const int64_t g = t >> 6;
        const unsigned long long b = bits[g];
        const bool fb = (b >> lane) & 1ull;
        const int before = P[g] + __popcll(b & lt);  // matches before t
        const int pb = bound[t];
        if (UNION) {
          const int64_t r = t + pb - before;
          Ccol[r] = Xcol[t];
          Cval[r] = fb ? ew_val(Aval, t) + ew_val(Bval, pb) : ew_val(Aval, t);
        } else if (fb) {
          Ccol[before] = Xcol[t];
          Cval[before] = ew_val(Aval, t) * ew_val(Bval, pb);
        }
      }
    } else if (on_x) {
      const int64_t base = tile * EW_TILE;
      EwTile<T, OP, MASKED> e;
      e.run(base, lane, Xrp, Xcol, Xtr, nX, ngroups, m, Arp, Acol, Aval, Brp, Bcol, Bval);
#pragma unroll
      for (int q = 0; q < EW_Q; ++q) {
        const int64_t t = base + q * 64 + lane;
        const unsigned long long b = __ballot(e.flag[q]);
        const int64_t g = (base >> 6) + q;
        if (!e.valid[q]) continue;
        const int before = (g < ngroups ? P[g] : 0) + __popcll(b & lt);  // set bits before t
        if (UNION) {
          const int64_t r = t + e.pb[q] - before;
          Ccol[r] = e.col[q];
          Cval[r] = e.fb[q] ? ew_val(Aval, t) + ew_val(Bval, e.pb[q]) : ew_val(Aval, t);
        } else if (e.flag[q]) {
          Ccol[before] = e.col[q];
          Cval[before] = MASKED ? e.v[q] : ew_val(Aval, t) * ew_val(Bval, e.pb[q]);
        }
      }
    } else if (UNION) {
      // B's entries: their lower bounds in A; the unmatched ones are written
      const int64_t base = tile * EW_TILE;
      bool valid[EW_Q];
      int c[EW_Q], lo[EW_Q], end[EW_Q];
#pragma unroll
      for (int q = 0; q < EW_Q; ++q) {
        const int64_t u = base + q * 64 + lane;
        valid[q] = u < nB;
        c[q] = lo[q] = end[q] = 0;
        if (valid[q]) {
          const int i = ew_row_of(Brp, Btr, ngroupsB, m, u);
          c[q] = Bcol[u];
          lo[q] = Arp[i];
          end[q] = Arp[i + 1];
        }
      }
      ew_lbound_q<EW_Q>(Acol, lo, end, c);
#pragma unroll
      for (int q = 0; q < EW_Q; ++q) {
        const int64_t u = base + q * 64 + lane;
        if (!valid[q]) continue;
        const int pa = lo[q];
        if (pa < end[q] && Acol[pa] == c[q]) continue;  // written with its A entry
        const int64_t r = u + pa - ew_prefix(bits, P, pa);
        Ccol[r] = c[q];
        Cval[r] = ew_val(Bval, u);
      }
    }
  }
}

template <typename T>
__global__ void k_scale(const T* __restrict__ v, int64_t n, double alpha, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (T)((double)v[i] * alpha);
}

static hipEvent_t ew_event(Context& c) {
  hipEvent_t ev;
  if (!c.event_pool.empty()) {
    ev = c.event_pool.back();
    c.event_pool.pop_back();
  } else {
    GA_HIP(hipEventCreate(&ev));
  }
  return ev;
}

template <typename T, int OP, bool MASKED>
static void ew_launch_mark(const DCsr& X, const int32_t* Xtr, int64_t ngX, const DCsr& A, const DCsr& B,
                           uint64_t* bits, int32_t* gcnt, int32_t* bound, hipStream_t s) {
  const unsigned grid = grid_for(ceil_div(X.nnz, EW_TILE), 4, 65536);
  hipLaunchKernelGGL((k_ew_mark<T, OP, MASKED>), dim3(grid), dim3(256), 0, s, X.rp, X.col, Xtr, X.nnz, ngX,
                     A.nrows, A.rp, A.col, static_cast<const T*>(A.val), B.rp, B.col, static_cast<const T*>(B.val),
                     bits, gcnt, bound);
  GA_HIP(hipGetLastError());
}

template <typename T, int OP, bool MASKED>
static void ew_launch_fill(const DCsr& X, const int32_t* Xtr, int64_t ngX, const DCsr& A, const DCsr& B,
                           const int32_t* Btr, int64_t ngB, const uint64_t* bits, const int32_t* P,
                           const int32_t* bound, int32_t* Ccol, T* Cval, hipStream_t s) {
  constexpr bool UNION = !MASKED && OP != EW_MULT;
  const int64_t waves = ceil_div(X.nnz, EW_TILE) + (UNION ? ceil_div(B.nnz, EW_TILE) : 0);
  const unsigned grid = grid_for(waves, 4, 65536);
  hipLaunchKernelGGL((k_ew_fill<T, OP, MASKED>), dim3(grid), dim3(256), 0, s, X.rp, X.col, Xtr, X.nnz, ngX,
                     A.nrows, A.rp, A.col, static_cast<const T*>(A.val), B.rp, B.col, static_cast<const T*>(B.val),
                     Btr, ngB, bits, P, bound, Ccol, Cval);
  GA_HIP(hipGetLastError());
}

// Device-resident operands in, host arrays out (malloc).  With profiling on,
// the mark pass (group rows, marks, the scan of the group counts and C's row
// pointers) and the fill pass are two timed launches (the host reads C's
// size between them).  Algorithmic bytes: the first pass reads the row
// pointers and columns of A, B (and M) and writes C's row pointers; the fill
// pass reads the structure again, the values that reach C (all of them for
// the unmasked sum; for products and masked sums, the two values of each
// entry of C -- the values of dropped zero sums are not counted), and writes C.
template <typename T, int OP, bool MASKED>
static void ewise_run(const DCsr& A, const DCsr& B, const DCsr* M, int32_t* Crp_h, int32_t** Ccol_h, void** Cval_h,
                      int64_t* Cnnz, hipStream_t s) {
  constexpr bool UNION = !MASKED && OP != EW_MULT;
  const int64_t m = A.nrows;
  const DCsr& X = MASKED ? *M : A;
  const int64_t ngX = ceil_div(X.nnz, 64), ngB = UNION ? ceil_div(B.nnz, 64) : 0;
  Context& c = ctx();
  DevBuf trX, trB, bits, gcnt, P, rp, tmp, ccol, cval, bound;
  trX.ensure(sizeof(int32_t) * std::max<int64_t>(ngX, 1));
  trB.ensure(sizeof(int32_t) * std::max<int64_t>(ngB, 1));
  bits.ensure(sizeof(uint64_t) * std::max<int64_t>(ngX, 1));
  gcnt.ensure(sizeof(int32_t) * std::max<int64_t>(ngX, 1));
  P.ensure(sizeof(int32_t) * (ngX + 1));
  rp.ensure(sizeof(int32_t) * (m + 1));
  bound.ensure(sizeof(int32_t) * (MASKED ? 1 : std::max<int64_t>(A.nnz, 1)));  // A's lower bounds in B
  hipEvent_t e_open = nullptr;
  auto mark = [&](bool open) {
    if (!c.prof_events_on) return;
    hipEvent_t e = ew_event(c);
    GA_HIP(hipEventRecord(e, s));
    if (open) {
      e_open = e;
    } else {
      c.prof_events.emplace_back(e_open, e);
      c.prof_launches += 1;
    }
  };
  mark(true);
  hipLaunchKernelGGL(k_ew_group_rows, dim3(grid_for(ngX, 256, 16384)), dim3(256), 0, s, X.rp, m, ngX,
                     trX.as<int32_t>());
  if (UNION)
    hipLaunchKernelGGL(k_ew_group_rows, dim3(grid_for(ngB, 256, 16384)), dim3(256), 0, s, B.rp, m, ngB,
                       trB.as<int32_t>());
  GA_HIP(hipGetLastError());
  ew_launch_mark<T, OP, MASKED>(X, trX.as<int32_t>(), ngX, A, B, bits.as<uint64_t>(), gcnt.as<int32_t>(),
                                bound.as<int32_t>(), s);
  scan_exclusive_i32(gcnt.as<int32_t>(), P.as<int32_t>(), ngX, tmp, s);
  hipLaunchKernelGGL((k_ew_rowptr<OP, MASKED>), dim3(grid_for(m + 1, 256, 16384)), dim3(256), 0, s, A.rp, B.rp,
                     MASKED ? M->rp : nullptr, m, (const uint64_t*)bits.as<uint64_t>(), (const int32_t*)P.as<int32_t>(),
                     rp.as<int32_t>());
  GA_HIP(hipGetLastError());
  mark(false);
  GA_HIP(hipMemcpyAsync(Crp_h, rp.p, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  const int64_t nz = Crp_h[m];
  // a union holds nnz(A) + nnz(B) - (shared positions) entries
  GA_REQUIRE(OP != EW_DISJOINT || nz == A.nnz + B.nnz, GA_ERR_ARG, "add_nointersect: the operands share a position");
  *Cnnz = nz;
  ccol.ensure(sizeof(int32_t) * std::max<int64_t>(nz, 1));
  cval.ensure(sizeof(T) * std::max<int64_t>(nz, 1));
  mark(true);
  ew_launch_fill<T, OP, MASKED>(X, trX.as<int32_t>(), ngX, A, B, trB.as<int32_t>(), ngB, bits.as<uint64_t>(),
                                P.as<int32_t>(), bound.as<int32_t>(), ccol.as<int32_t>(), cval.as<T>(), s);
  mark(false);
  if (c.prof_bytes) {
    const int64_t vs = (int64_t)sizeof(T), va = A.val ? vs : 0, vb = B.val ? vs : 0;
    const int64_t st = 4 * (m + 1) * 2 + 4 * A.nnz + 4 * B.nnz + (MASKED ? 4 * (m + 1) + 4 * M->nnz : 0);
    const int64_t vals = UNION ? va * A.nnz + vb * B.nnz : (va + vb) * nz;
    c.prof_bytes_host += (st + 4 * (m + 1)) + (st + vals + (4 + vs) * nz);
    c.prof_bytes_launches += 2;
  }
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
static void ewise_dispatch(int op, const DCsr& A, const DCsr& B, const DCsr* M, int32_t* Crp_h, int32_t** Ccol_h,
                           void** Cval_h, int64_t* Cnnz, hipStream_t s) {
  if (M) {
    if (op == EW_MULT) ewise_run<T, EW_MULT, true>(A, B, M, Crp_h, Ccol_h, Cval_h, Cnnz, s);
    else ewise_run<T, EW_ADD, true>(A, B, M, Crp_h, Ccol_h, Cval_h, Cnnz, s);
  } else if (op == EW_MULT) {
    ewise_run<T, EW_MULT, false>(A, B, M, Crp_h, Ccol_h, Cval_h, Cnnz, s);
  } else if (op == EW_DISJOINT) {
    ewise_run<T, EW_DISJOINT, false>(A, B, M, Crp_h, Ccol_h, Cval_h, Cnnz, s);
  } else {
    ewise_run<T, EW_ADD, false>(A, B, M, Crp_h, Ccol_h, Cval_h, Cnnz, s);
  }
}

}  // namespace ga

using namespace ga;

extern "C" {

ga_status ga_ewise(int op, const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, int32_t** Crowptr,
                   int32_t** Ccol, void** Cval, int64_t* Cnnz) {
  return guarded([&] {
    dtype_size(dtype);
    GA_REQUIRE(op == GA_EW_ADD || op == GA_EW_MULT || op == GA_EW_ADD_DISJOINT, GA_ERR_ARG, "unknown operator");
    validate_host_csr(A, "A");
    validate_host_csr(B, "B");
    GA_REQUIRE(A->nrows == B->nrows && A->ncols == B->ncols, GA_ERR_DIM, "A and B must have the same shape");
    if (M) {
      validate_host_csr(M, "M");
      GA_REQUIRE(M->nrows == A->nrows && M->ncols == A->ncols, GA_ERR_DIM, "M must have the shape of A");
      GA_REQUIRE(op != GA_EW_ADD_DISJOINT, GA_ERR_ARG, "add_nointersect takes no mask");
    }
    // C's row pointers are int32: the union of A and B must fit them
    GA_REQUIRE(M || op == GA_EW_MULT || A->nnz + B->nnz < INT32_MAX, GA_ERR_LIMIT,
               "the union of A and B exceeds int32 CSR indexing");
    GA_REQUIRE(Crowptr && Ccol && Cval && Cnnz, GA_ERR_ARG, "null output");
    hipStream_t s = ctx().stream;
    OwnedCsr dA, dB, dM;
    upload_csr(A, dtype, dA, s);
    upload_csr(B, dtype, dB, s);
    DCsr vM{};
    if (M) {
      ga_csr pattern = *M;  // the mask's values are never read
      pattern.val = nullptr;
      upload_csr(&pattern, dtype, dM, s);
      vM = dM.view();
    }
    int32_t* rp = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (A->nrows + 1)));
    GA_REQUIRE(rp != nullptr, GA_ERR_OOM, "host allocation failed");
    *Ccol = nullptr;
    *Cval = nullptr;
    try {
      const DCsr* pm = M ? &vM : nullptr;
      switch (dtype) {
        case GA_INT32: ewise_dispatch<int32_t>(op, dA.view(), dB.view(), pm, rp, Ccol, Cval, Cnnz, s); break;
        case GA_INT64: ewise_dispatch<int64_t>(op, dA.view(), dB.view(), pm, rp, Ccol, Cval, Cnnz, s); break;
        case GA_FLOAT32: ewise_dispatch<float>(op, dA.view(), dB.view(), pm, rp, Ccol, Cval, Cnnz, s); break;
        default: ewise_dispatch<double>(op, dA.view(), dB.view(), pm, rp, Ccol, Cval, Cnnz, s); break;
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

ga_status ga_scale(const ga_csr* A, ga_dtype dtype, double alpha, void* Cval) {
  return guarded([&] {
    const size_t es = dtype_size(dtype);
    validate_host_csr(A, "A");
    GA_REQUIRE(A->val != nullptr && (A->nnz == 0 || Cval), GA_ERR_ARG, "null values");
    if (A->nnz == 0) return;
    hipStream_t s = ctx().stream;
    DevBuf v, out;
    v.ensure(es * A->nnz);
    out.ensure(es * A->nnz);
    GA_HIP(hipMemcpyAsync(v.p, A->val, es * A->nnz, hipMemcpyHostToDevice, s));
    const unsigned grid = grid_for(A->nnz, 256, 8192);
    switch (dtype) {
      case GA_INT32:
        hipLaunchKernelGGL(k_scale<int32_t>, dim3(grid), dim3(256), 0, s, v.as<int32_t>(), A->nnz, alpha,
                           out.as<int32_t>());
        break;
      case GA_INT64:
        hipLaunchKernelGGL(k_scale<int64_t>, dim3(grid), dim3(256), 0, s, v.as<int64_t>(), A->nnz, alpha,
                           out.as<int64_t>());
        break;
      case GA_FLOAT32:
        hipLaunchKernelGGL(k_scale<float>, dim3(grid), dim3(256), 0, s, v.as<float>(), A->nnz, alpha,
                           out.as<float>());
        break;
      default:
        hipLaunchKernelGGL(k_scale<double>, dim3(grid), dim3(256), 0, s, v.as<double>(), A->nnz, alpha,
                           out.as<double>());
        break;
    }
    GA_HIP(hipGetLastError());
    GA_HIP(hipMemcpyAsync(Cval, out.p, es * A->nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

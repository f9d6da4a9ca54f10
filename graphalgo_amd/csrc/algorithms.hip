// Triangle counting and k-truss on the masked SpGEMM.
//
// Triangle count (src/main.cpp:65-86): L = tril(A,-1) extracted on device,
// C<L> = L*L run in the kernel's sum mode (per-item int64 partial sums, one
// atomic per item), so C is never materialised.
//
// k-truss (src/main.cpp:89-146) and the truss decomposition (the caller-side
// loop over k): the graph lives on the device as a ping-pong pair of CSR
// buffers plus an edge-id array mapping each surviving entry to its slot in
// the input.  One step = support pass S<C> = C*C (skipped when the graph is
// unchanged since the last pass), then a wavefront-per-row filter that counts
// survivors with ballot/popcount, a device scan of the row counts, and a
// ballot-ranked compaction into the other buffer.  Only the survivor total
// returns to the host per step (4 bytes) to steer the loop.
#include "common.hpp"

#include <algorithm>
#include <chrono>

namespace ga {

__global__ void k_upper_count(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                              int32_t* __restrict__ cnt, int32_t* __restrict__ first) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = rp[i], hi = rp[i + 1];
    const int e = hi;
    while (lo < hi) {  // first column > i
      int mid = (lo + hi) >> 1;
      if (col[mid] <= i) lo = mid + 1;
      else hi = mid;
    }
    cnt[i] = e - lo;
    first[i] = lo;
  }
}

__global__ void k_upper_copy(const int32_t* __restrict__ col, int64_t n, const int32_t* __restrict__ first,
                             const int32_t* __restrict__ urp, int32_t* __restrict__ ucol) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * wpb) {
    const int src = first[i], dst = urp[i], len = urp[i + 1] - urp[i];
    for (int q = lane; q < len; q += 64) ucol[dst + q] = col[src + q];
  }
}

static void upper_triangle_dev(const DCsr& A, OwnedCsr& U, DevBuf& tmp, hipStream_t s) {
  const int64_t n = A.nrows;
  U.nrows = n;
  U.ncols = A.ncols;
  U.rp.ensure(sizeof(int32_t) * (n + 1));
  DevBuf cnt, first;
  cnt.ensure(sizeof(int32_t) * std::max<int64_t>(n, 1));
  first.ensure(sizeof(int32_t) * std::max<int64_t>(n, 1));
  hipLaunchKernelGGL(k_upper_count, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, A.rp, A.col, n,
                     cnt.as<int32_t>(), first.as<int32_t>());
  scan_exclusive_i32(cnt.as<int32_t>(), U.rp.as<int32_t>(), n, tmp, s);
  int32_t unnz = 0;
  GA_HIP(hipMemcpyAsync(&unnz, U.rp.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  U.nnz = unnz;
  U.col.ensure(sizeof(int32_t) * std::max<int64_t>(unnz, 1));
  hipLaunchKernelGGL(k_upper_copy, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.col, n,
                     (const int32_t*)first.as<int32_t>(), (const int32_t*)U.rp.as<int32_t>(), U.col.as<int32_t>());
  GA_HIP(hipGetLastError());
}

// ---- filter / compaction (entry-parallel) ---------------------------------
// The entries with support >= thr survive.  The CSR keeps its order, so the
// new position of a survivor is the number of survivors before it: one
// exclusive scan over the entries (tiles of KEEP_TILE) places every entry,
// and the new row pointer of row i is that count at rp[i].  Every thread
// owns KEEP_PER consecutive entries; all loads and stores are coalesced and
// the work does not depend on the row lengths.
// Peeling state of a device-driven run (one block of device memory).  The
// step kernels read it when they start; k_truss_state, the last kernel of a
// step, updates it.  A run needs no host decision between steps: steps
// enqueued after the run has finished see done = 1 and do nothing.
constexpr int TRUSS_MAX_K = 64;  // k values of one shared k-truss run
struct TrussDevState {
  long long nnz;     // entries of the current graph
  long long passes;  // support passes run
  long long steps;   // filter steps run
  long long ran;     // step kernels run, the no-ops after the end included
  int k;             // the current k (threshold k - 2)
  int valid;         // the supports of the current graph are computed
  int done;
  int kidx;          // k list: index of the current k
  int kmax;          // truss decomposition: largest k whose truss is nonempty
  int skip_support;  // the next step runs no support pass
  int skip_compact;  // the next step runs no compaction
  int assign_k;      // truss decomposition: k for the current entries (0 = none)
  int all_k;         // truss decomposition (else a k list)
  int nk;
  int error;         // a survivor count outside [0, nnz] ended the run
  int ks[TRUSS_MAX_K];
  int passes_at[TRUSS_MAX_K];
  long long nnz_at[TRUSS_MAX_K];
};

constexpr int KEEP_T = 256;
constexpr int KEEP_PER = 8;
constexpr int KEEP_TILE = KEEP_T * KEEP_PER;
constexpr int KEEP_FUSED_TILES = 1024;  // up to here a tile sums the tiles before it itself

// ds (device-driven steps): nnz and thr come from the state, the nnz
// argument is the first graph's (its tiles are walked with a grid stride,
// the ones past the current graph count zero), and the entries of a graph
// that was stable at the previous step get their truss number assign_k
// (eid, truss; truss decomposition).
__global__ __launch_bounds__(KEEP_T) void k_keep_tile_sums(const int32_t* __restrict__ sup, int64_t nnz, int thr,
                                                           long long* __restrict__ bsum,
                                                           const TrussDevState* __restrict__ ds,
                                                           const int32_t* __restrict__ eid,
                                                           int32_t* __restrict__ truss) {
  __shared__ long long red[KEEP_T / 64];
  int64_t ntiles = gridDim.x;
  int assign = 0;
  if (ds) {
    ntiles = (nnz + KEEP_TILE - 1) / KEEP_TILE;
    nnz = ds->skip_compact ? 0 : ds->nnz;
    thr = ds->k - 2;
    assign = truss ? ds->assign_k : 0;
  }
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t base = tile * KEEP_TILE;
    if (base >= nnz) {  // device-driven: past the current graph
      if (threadIdx.x == 0) bsum[tile] = 0;
      continue;
    }
    long long c = 0;
    for (int q = threadIdx.x; q < KEEP_TILE; q += KEEP_T) {
      const int64_t j = base + q;
      if (j < nnz) {
        c += sup[j] >= thr ? 1 : 0;
        if (assign) truss[eid[j]] = assign;
      }
    }
    c = wave_reduce_sum(c);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long t = 0;
      for (int w = 0; w < KEEP_T / 64; ++w) t += red[w];
      bsum[tile] = t;
    }
    __syncthreads();  // red is reused by the next tile
  }
}

// Survivors to their new positions (col, entry id, old position, the next
// pass's support: cleared, or carried with `carry`); newpos[j] = new position
// of a survivor, -1 - (survivors before j) for a removed entry.  bsum: tile
// sums, or exclusive tile offsets when offsets_ready.  Wavefront w of a tile
// owns its entries [w * 512, (w + 1) * 512) and reads them 64 at a time
// (coalesced); a ballot per 64 entries ranks the survivors, so the stores of
// consecutive survivors are consecutive too.  The last tile writes the
// survivor count to *kept.  ds: as in k_keep_tile_sums (tiles walked with a
// grid stride up to the current graph's last).
__global__ __launch_bounds__(KEEP_T) void k_keep_scatter(
    const int32_t* __restrict__ col, const int32_t* __restrict__ eid, const int32_t* __restrict__ sup, int64_t nnz,
    int thr, const long long* __restrict__ bsum, int offsets_ready, int32_t* __restrict__ ncol,
    int32_t* __restrict__ neid, int32_t* __restrict__ nold, int32_t* __restrict__ newpos, int32_t* __restrict__ nsup,
    int carry, int32_t* __restrict__ kept, const TrussDevState* __restrict__ ds, long long* __restrict__ host_kept,
    long long seq) {
  constexpr int NW = KEEP_T / 64;
  __shared__ long long prior[NW];
  __shared__ int wcount[NW];
  int64_t ntiles = gridDim.x;
  if (ds) {
    nnz = ds->skip_compact ? 0 : ds->nnz;
    thr = ds->k - 2;
    ntiles = (nnz + KEEP_TILE - 1) / KEEP_TILE;
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    long long before = 0;
    if (offsets_ready) {
      before = tid == 0 ? bsum[tile] : 0;
    } else {
      for (int64_t b = tid; b < tile; b += KEEP_T) before += bsum[b];
    }
    before = wave_reduce_sum(before);
    if (lane == 0) prior[wv] = before;
    const int64_t wbase = tile * KEEP_TILE + (int64_t)wv * 64 * KEEP_PER;
    unsigned long long mk[KEEP_PER];
    int32_t sv[KEEP_PER];
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < KEEP_PER; ++q) {
      const int64_t j = wbase + q * 64 + lane;
      sv[q] = j < nnz ? sup[j] : -1;
      mk[q] = __ballot(j < nnz && sv[q] >= thr);
      cnt += __popcll(mk[q]);
    }
    if (lane == 0) wcount[wv] = cnt;
    __syncthreads();
    long long off = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      off += prior[k];
      if (k < wv) off += wcount[k];
    }
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int q = 0; q < KEEP_PER; ++q) {
      const int64_t j = wbase + q * 64 + lane;
      if (j < nnz) {
        const bool keep = (mk[q] >> lane) & 1ull;
        const int r = (int)(off + __popcll(mk[q] & lt));
        if (keep) {
          ncol[r] = col[j];
          neid[r] = eid[j];
          nold[r] = (int32_t)j;
          nsup[r] = carry ? sv[q] : 0;
        }
        newpos[j] = keep ? r : -1 - r;
      }
      off += __popcll(mk[q]);
    }
    if ((tile + 1) * KEEP_TILE >= nnz && tid == 0) {  // the last tile
      long long t = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) t += prior[k] + wcount[k];
      *kept = (int32_t)t;
      if (host_kept) {  // the host's copy: the count, then the sequence number
        __hip_atomic_store(&host_kept[0], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&host_kept[1], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    __syncthreads();  // prior and wcount are reused by the next tile
  }
}

// The rest of a compaction, one launch: thread t is row t of the new row
// pointers (nrp[i] = survivors before row i, the count at rp[i]; dirty[i] =
// row i loses entries, only when dirty is given) and survivor t of the mirror
// index update (the transposed entry of a survivor survives too -- equal
// supports -- so the mirror index carries over through the old -> new map).
__global__ void k_keep_finish(const int32_t* __restrict__ rp, int64_t n, const int32_t* __restrict__ newpos,
                              int64_t nnz, const int32_t* __restrict__ kept, int32_t* __restrict__ nrp,
                              uint8_t* __restrict__ dirty, const int32_t* __restrict__ mirror_old,
                              const int32_t* __restrict__ nold, int32_t* __restrict__ mirror_new,
                              const TrussDevState* __restrict__ ds) {
  if (ds) {
    if (ds->skip_compact) return;
    nnz = ds->nnz;
  }
  const int32_t total = *kept;
  auto before = [&](int64_t j) -> int32_t {
    if (j >= nnz) return total;
    const int32_t p = newpos[j];
    return p >= 0 ? p : -1 - p;
  };
  const int64_t span = max<int64_t>(n + 1, total);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < span; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < n) {
      const int32_t a = before(rp[t]);
      nrp[t] = a;
      if (dirty) dirty[t] = (before(rp[t + 1]) - a) < (rp[t + 1] - rp[t]) ? 1 : 0;
    } else if (t == n) {
      nrp[n] = total;
    }
    if (t < total) {
      const int mo = mirror_old[nold[t]];
      mirror_new[t] = mo >= 0 ? newpos[mo] : -1;
    }
  }
}

__global__ void k_truss_assign(const int32_t* __restrict__ eid, int64_t nnz, int k, int32_t* __restrict__ truss) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    truss[eid[e]] = k;
}

__global__ void k_iota(int32_t* __restrict__ p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = (int32_t)e;
}

// ---- device-driven peeling steps -------------------------------------------
// A step is the body of the host loop in truss_loop: a support pass when the
// current graph has none, a compaction at threshold k - 2, and the state
// update.  Its kernels take the same arguments every time (the graph always
// lives in buffer set 0: a compaction that removes entries copies the
// survivors back from set 1), so one captured step replays for the whole run.

struct TrussInit {
  long long nnz;
  int all_k;
  int nk;
  int ks[TRUSS_MAX_K];
};

__global__ void k_truss_state_init(TrussDevState* __restrict__ ds, TrussInit in) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  TrussDevState& d = *ds;
  d.nnz = in.nnz;
  d.passes = 0;
  d.steps = 0;
  d.ran = 0;
  d.all_k = in.all_k;
  d.nk = in.nk;
  for (int i = 0; i < TRUSS_MAX_K; ++i) {
    d.ks[i] = i < in.nk ? in.ks[i] : 0;
    d.passes_at[i] = 0;
    d.nnz_at[i] = 0;
  }
  d.k = in.all_k ? 3 : in.ks[0];
  d.valid = 0;
  d.kidx = 0;
  d.kmax = 2;
  d.assign_k = 0;
  d.error = 0;
  d.done = (in.all_k && in.nnz == 0) ? 1 : 0;
  d.skip_support = (d.done || in.nnz == 0) ? 1 : 0;
  d.skip_compact = d.skip_support;
}

// Survivors back into buffer set 0 (row pointers, columns, edge ids, mirror
// positions); their supports cleared for the next pass.  Only when the
// step's compaction removed entries.
__global__ void k_keep_copyback(const TrussDevState* __restrict__ ds, const int32_t* __restrict__ kept, int64_t n,
                                const int32_t* __restrict__ rp1, const int32_t* __restrict__ col1,
                                const int32_t* __restrict__ eid1, const int32_t* __restrict__ mir1,
                                int32_t* __restrict__ rp0, int32_t* __restrict__ col0, int32_t* __restrict__ eid0,
                                int32_t* __restrict__ mir0, int32_t* __restrict__ sup0) {
  if (ds->skip_compact) return;
  const int64_t total = *kept;
  if (total == ds->nnz) return;
  const int64_t span = max<int64_t>(n + 1, total);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < span; t += (int64_t)gridDim.x * blockDim.x) {
    if (t <= n) rp0[t] = rp1[t];
    if (t < total) {
      col0[t] = col1[t];
      eid0[t] = eid1[t];
      mir0[t] = mir1[t];
      sup0[t] = 0;
    }
  }
}

static __device__ void truss_state_update(TrussDevState& d, const int32_t* __restrict__ kept) {
  auto next_k = [&]() {
    d.nnz_at[d.kidx] = d.nnz;
    d.passes_at[d.kidx] = (int)d.passes;
    if (++d.kidx == d.nk) d.done = 1;
    else d.k = d.ks[d.kidx];
  };
  if (!d.valid) {  // this step ran the support pass (none for an empty graph)
    d.passes += 1;
    d.valid = 1;
  }
  d.steps += 1;
  if (d.nnz != 0 && (*kept < 0 || (long long)*kept > d.nnz)) {  // cannot happen; never index with it
    d.error = 1;
    d.done = 1;
  } else if (d.nnz == 0) {  // nothing to filter: a k-truss that became empty
    if (d.all_k) d.done = 1;
    else next_k();
  } else if ((long long)*kept == d.nnz) {  // stable at k
    if (d.all_k) {
      d.assign_k = d.k;
      d.kmax = d.k;
      d.k += 1;
    } else {
      next_k();
    }
  } else {
    d.nnz = *kept;
    d.valid = 0;
  }
  if (d.all_k && d.nnz == 0) d.done = 1;
  d.skip_support = (d.done || d.valid || d.nnz == 0) ? 1 : 0;
  d.skip_compact = (d.done || d.nnz == 0) ? 1 : 0;
}

// k_truss_state runs the rest of one iteration of truss_loop's host loop
// (truss_state_update, in the same order: passes, steps and the per-k
// records count exactly as there), then the test at the top of the next
// iteration.  progress (host memory): the step
// kernels run so far and the done flag, for the host's flow control.
__global__ void k_truss_state(TrussDevState* __restrict__ ds, const int32_t* __restrict__ kept,
                              long long* __restrict__ progress) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  TrussDevState& d = *ds;
  d.assign_k = 0;
  d.ran += 1;
  if (!d.done) truss_state_update(d, kept);
  if (progress) {
    __hip_atomic_store(&progress[1], (long long)d.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&progress[0], d.ran, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

static bool selective_on(const Context& c, int64_t nnz0) {
  return c.selective_support > 0 || (c.selective_support < 0 && c.support_tri == 0 && nnz0 >= c.selective_min_nnz);
}

// The peeling loop of truss_loop as device-driven steps (buffers set up by
// truss_loop; selective support passes and byte accounting stay on the host
// loop).  graph: the step is captured once into a HIP graph and replayed;
// otherwise its kernels are launched step by step.  The host keeps a few
// steps queued ahead of the device (flow control below); the steps queued
// after the last one are no-ops.
static void truss_steps(const DCsr& A, bool all_k, const int32_t* ks, int nk, TrussWork& tw, hipStream_t s,
                        bool graph, int64_t* nnz_out, int32_t* kmax_out, ga_truss_stats* st, int64_t* nnz_at,
                        int32_t* passes_at) {
  const int64_t n = A.nrows, nnz0 = A.nnz;
  GA_REQUIRE(all_k || nk <= TRUSS_MAX_K, GA_ERR_LIMIT, "at most 64 k values in one k-truss run");
  Context& c = ctx();
  tw.dstate.ensure(sizeof(TrussDevState));
  auto* ds = tw.dstate.as<TrussDevState>();
  if (nnz0) fill_u32(tw.sup[0].p, 0u, nnz0, s);
  tw.mw.rwork[0].ensure(sizeof(unsigned long long) * std::max<int64_t>(n, 1));
  if (n) fill_u64(tw.mw.rwork[0].p, 0ull, n, s);
  TrussInit in{};
  in.nnz = nnz0;
  in.all_k = all_k ? 1 : 0;
  in.nk = all_k ? 0 : nk;
  for (int i = 0; i < in.nk; ++i) in.ks[i] = ks[i];
  hipLaunchKernelGGL(k_truss_state_init, dim3(1), dim3(64), 0, s, ds, in);
  GA_HIP(hipGetLastError());
  struct FixedArgs {  // the support passes take fixed arguments for the run only
    MxmWork& w;
    ~FixedArgs() {
      w.fixed_args = false;
      w.skip = nullptr;
    }
  } fixed{tw.mw};
  tw.mw.fixed_args = true;
  tw.mw.skip = &ds->skip_support;
  int32_t* kept_dev = tw.kept.as<int32_t>();
  long long* bsum = tw.scan_tmp.as<long long>();
  long long* prog_dev = nullptr;
  long long* prog = progress_words(&prog_dev);
  __atomic_store_n(&prog[0], 0ll, __ATOMIC_RELAXED);
  __atomic_store_n(&prog[1], 0ll, __ATOMIC_RELEASE);
  const int64_t nbk = ceil_div(nnz0, KEEP_TILE);
  const bool offsets = nbk > KEEP_FUSED_TILES;
  // grids of at most two workgroups per CU: the graph shrinks during the
  // run, and a step that has little left to do should not dispatch
  // thousands of workgroups that find no work
  const int64_t gcap = 2 * (int64_t)c.num_cu;
  const unsigned gtiles = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nbk, gcap));
  const unsigned gspan = (unsigned)std::min<int64_t>(grid_for(std::max<int64_t>(n + 1, nnz0), 256, 8192), gcap);
  int64_t support_launches = 0;  // enqueued item launches (profiling: the skipped ones are not passes)
  auto step = [&]() {
    if (nnz0 > 0) {
      const DCsr C{n, n, nnz0, tw.cur[0].rp.as<int32_t>(), tw.cur[0].col.as<int32_t>(), nullptr};
      support_symmetric_dev(C, tw.mirror[0].as<int32_t>(), tw.sup[0].as<int32_t>(), tw.mw, s, true);
      ++support_launches;
      hipLaunchKernelGGL(k_keep_tile_sums, dim3(gtiles), dim3(KEEP_T), 0, s,
                         (const int32_t*)tw.sup[0].as<int32_t>(), nnz0, 0, bsum, (const TrussDevState*)ds,
                         (const int32_t*)tw.eid[0].as<int32_t>(), all_k ? tw.truss.as<int32_t>() : nullptr);
      if (offsets) scan_tile_offsets(bsum, nbk, kept_dev, s);
      hipLaunchKernelGGL(k_keep_scatter, dim3(gtiles), dim3(KEEP_T), 0, s,
                         (const int32_t*)tw.cur[0].col.as<int32_t>(), (const int32_t*)tw.eid[0].as<int32_t>(),
                         (const int32_t*)tw.sup[0].as<int32_t>(), nnz0, 0, (const long long*)bsum, offsets ? 1 : 0,
                         tw.cur[1].col.as<int32_t>(), tw.eid[1].as<int32_t>(), tw.nold.as<int32_t>(),
                         tw.newpos.as<int32_t>(), tw.sup[1].as<int32_t>(), 0, kept_dev, (const TrussDevState*)ds,
                         (long long*)nullptr, 0ll);
      hipLaunchKernelGGL(k_keep_finish, dim3(gspan), dim3(256), 0, s, (const int32_t*)tw.cur[0].rp.as<int32_t>(), n,
                         (const int32_t*)tw.newpos.as<int32_t>(), nnz0, (const int32_t*)kept_dev,
                         tw.cur[1].rp.as<int32_t>(), (uint8_t*)nullptr, (const int32_t*)tw.mirror[0].as<int32_t>(),
                         (const int32_t*)tw.nold.as<int32_t>(), tw.mirror[1].as<int32_t>(), (const TrussDevState*)ds);
      hipLaunchKernelGGL(k_keep_copyback, dim3(gspan), dim3(256), 0, s, (const TrussDevState*)ds,
                         (const int32_t*)kept_dev, n, (const int32_t*)tw.cur[1].rp.as<int32_t>(),
                         (const int32_t*)tw.cur[1].col.as<int32_t>(), (const int32_t*)tw.eid[1].as<int32_t>(),
                         (const int32_t*)tw.mirror[1].as<int32_t>(), tw.cur[0].rp.as<int32_t>(),
                         tw.cur[0].col.as<int32_t>(), tw.eid[0].as<int32_t>(), tw.mirror[0].as<int32_t>(),
                         tw.sup[0].as<int32_t>());
    }
    // (a truss decomposition's entries get the k they were stable at from
    // the next step's tile sums: that step always runs, on the same graph)
    hipLaunchKernelGGL(k_truss_state, dim3(1), dim3(64), 0, s, ds, (const int32_t*)kept_dev, prog_dev);
    GA_HIP(hipGetLastError());
  };
  // the first step runs eagerly: it sizes every workspace buffer, so the
  // capture allocates nothing
  step();
  graph = graph && !c.prof_events_on;  // kernel timing needs events around every launch
  if (graph) {
    const MxmWork& w = tw.mw;
    std::vector<const void*> key = {
        (const void*)(intptr_t)n, (const void*)(intptr_t)nnz0, (const void*)(intptr_t)all_k,
        (const void*)(intptr_t)c.item_wedges, (const void*)(intptr_t)w.max_mask_row,
        (const void*)(intptr_t)w.max_b_row, tw.cur[0].rp.p, tw.cur[0].col.p, tw.cur[1].rp.p, tw.cur[1].col.p,
        tw.eid[0].p, tw.eid[1].p, tw.mirror[0].p, tw.mirror[1].p, tw.sup[0].p, tw.sup[1].p, tw.nold.p,
        tw.newpos.p, tw.kept.p, tw.scan_tmp.p, tw.dstate.p, tw.truss.p, prog_dev, w.cnt.p, w.start.p, w.item_row.p,
        w.item_t0.p, w.scan_tmp.p, w.counter.p, w.mend.p, w.rwork[0].p};
    if (tw.step_graph.exec == nullptr || tw.step_graph.key != key) {
      tw.step_graph.reset();
      hipGraph_t g = nullptr;
      GA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try {
        step();
      } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(s, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
      }
      GA_HIP(hipStreamEndCapture(s, &g));
      const hipError_t e = hipGraphInstantiate(&tw.step_graph.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      GA_HIP(e);
      tw.step_graph.key = std::move(key);
    }
  }
  auto launch = [&]() {
    if (graph) GA_HIP(hipGraphLaunch(tw.step_graph.exec, s));
    else step();
  };
  // flow control: at most kInFlight steps queued beyond the last one the
  // device has finished (it reports through host memory, read here without
  // a HIP call), so the device never waits for the host and at most
  // kInFlight no-op steps run after the end
  constexpr int64_t kInFlight = 3;
  int64_t queued = 1;
  const int64_t max_steps = nnz0 / 2 + n + nk + 64;  // a step removes an edge or moves k on
  auto t_seen = std::chrono::steady_clock::now();
  long long seen = -1;
  struct DrainOnThrow {  // an error leaves no step running on buffers the caller may free
    hipStream_t s;
    bool armed = true;
    ~DrainOnThrow() {
      if (armed) (void)hipStreamSynchronize(s);
    }
  } drain{s};
  while (true) {
    const long long ran = __atomic_load_n(&prog[0], __ATOMIC_ACQUIRE);
    if (__atomic_load_n(&prog[1], __ATOMIC_RELAXED)) break;
    if (queued - ran < kInFlight) {
      GA_REQUIRE(queued < max_steps, GA_ERR_STATE, "k-truss steps did not finish");
      launch();
      ++queued;
      continue;
    }
    const auto now = std::chrono::steady_clock::now();
    if (ran != seen) {
      seen = ran;
      t_seen = now;
    } else if (now - t_seen > std::chrono::seconds(60)) {
      GA_HIP(hipStreamSynchronize(s));  // a step that ran for a minute: report, do not spin on
      GA_REQUIRE(__atomic_load_n(&prog[0], __ATOMIC_ACQUIRE) != seen, GA_ERR_STATE, "k-truss step made no progress");
    }
  }
  TrussDevState h{};
  GA_HIP(hipMemcpyAsync(&h, ds, sizeof(h), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  drain.armed = false;
  GA_REQUIRE(h.done && !h.error, GA_ERR_STATE, "k-truss steps did not finish");
  if (c.prof_events_on) {
    // every enqueued step timed its item launch; the ones a skip flag emptied
    // are not support passes (their few microseconds stay in the total)
    // launches with work: every pass except the one a k list counts on a
    // k-truss that became empty (no launch there, as in the host loop)
    const int64_t real = (!all_k && h.nnz == 0 && h.passes > 0) ? h.passes - 1 : h.passes;
    c.prof_launches -= std::max<int64_t>(0, support_launches - real);
  }
  if (nnz_out) *nnz_out = h.nnz;
  if (kmax_out) *kmax_out = h.kmax;
  if (!all_k)
    for (int i = 0; i < nk; ++i) {
      if (nnz_at) nnz_at[i] = h.nnz_at[i];
      if (passes_at) passes_at[i] = h.passes_at[i];
    }
  if (st) {
    st->passes = h.passes;
    st->steps = h.steps;
  }
}

// ---- peeling by support updates -------------------------------------------
// The reference recomputes every support, S<C> = C*C, after each filter
// (src/main.cpp:98-129).  With C = C' + D (D the entries removed by a filter,
// C' the survivors) the product splits by distributivity:
//   (C*C)<C'> = (C'*C')<C'> + (D*C' + C'*D + D*D)<C'>,
// so the supports of C' are the old supports minus, for each surviving entry,
// the triangles of C that contain it and at least one entry of D.  A round
// therefore lists the triangles through the entries it removes -- one
// intersection of the two endpoint rows per removed edge -- and decrements the
// supports of their surviving edges; every support stays exactly the one the
// reference computes for the same graph, so the filters remove the same
// entries in the same rounds and the pass counts are the reference's
// (1 + the rounds that removed something).
//
// Per entry (both directions of an edge hold the same stamp):
//   sup    the support, kept on the lower entry (row > col), the edge's
//          canonical position; the upper entry's value is not maintained
//   stamp  the round that removes the edge (PEEL_ALIVE while none does): a
//          round r removes the edges stamped r; a stamp below r is an edge
//          removed earlier, above r an edge still alive in round r.
// A triangle of C is destroyed once, by its edge of least canonical position
// among those stamped r; triangles with an edge removed earlier were
// destroyed then.  The decrement that takes a support from thr to thr - 1
// stamps the edge r + 1 and appends it to the next round's list, so rounds
// never scan the graph: only a new k does (k_peel_scan, entry-parallel: an
// entry is canonical when its transposed entry lies before it).  Dead
// entries stay in the rows until a compaction at a k boundary drops them
// (when they are at least a third of the stored entries).
constexpr int32_t PEEL_ALIVE = 0x7fffffff;
// A tick reports its progress to pinned host memory every PEEL_REPORT ticks
// and when the run ends (every 8: 10.92 ms per R-MAT 16 decomposition, every
// 4: 11.01, every 2: 11.04).  The host keeps more ticks than that queued, or
// it would wait for a report that no queued tick makes.
constexpr int64_t PEEL_REPORT = 8;

// Counters of a peeling run that change from round to round.
struct PeelDyn {
  long long alive;   // alive entries (both directions)
  long long passes;  // support products of the reference's schedule (1 + rounds that removed)
  long long steps;   // filter steps (removal rounds + the step that finds a k-truss stable)
  int r;             // rounds run (a round r removes the edges stamped r)
  int k;             // current k (threshold k - 2)
  int kidx, fin, kmax, err;  // k list position; run ended; largest nonempty k; a list outgrew its capacity
};

// Removal lists.  Every wavefront that pushes edges and every scan tile
// claims list slots with a device-scope atomic, and claims on one address
// serialise (about 11 ns each on MI355X: 4096 of them take 48 us, spread over
// 32 counters 4.5 us).  A list is therefore PEEL_SEGS segments of segcap
// items, each with its own counter on its own 128-byte line (a workgroup
// appends to segment blockIdx % PEEL_SEGS), followed by an overflow segment
// of ovcap items for the claims past a full segment.  The items of a list are
// numbered segment by segment, the overflow's last.
constexpr int PEEL_SEGS = 32;
constexpr int PEEL_CNT_STRIDE = 32;  // ints from one counter to the next (128 bytes)

// A list item is an edge (i, j), i > j: e = (its canonical entry (row i,
// column j), the transposed entry, i, j).  Carrying the endpoints saves a
// round the dependent load of both columns before the row bounds.
struct alignas(16) PeelItem {
  int4 e;
};

struct PeelLists {
  PeelItem* items; // list l starts at items + l * stride
  int64_t stride;  // PEEL_SEGS * segcap + ovcap
  int segcap, ovcap;
};

// The peeling state in device memory.  Host-steered runs keep their counters
// in `h` and use lists 0 and 1 (`ic` the next round's); device-driven runs
// alternate `dyn[2]` (tick t starts from dyn[t & 1] and stores the next
// tick's start in dyn[(t + 1) & 1]) and rotate four lists: tick t consumes
// list t & 3, fills (t + 1) & 3 and clears (t + 2) & 3, which no kernel in
// flight reads.
struct PeelCtl {
  long long bytes;  // algorithmic bytes of the round launches (accounting on)
  PeelDyn h;
  int ic;           // host-steered: the list of the next round
  PeelDyn dyn[2];
  int cr, clist, ccnt;  // compaction view: rounds run, the pending list and its size
  int nk, all_k;
  int ks[TRUSS_MAX_K];
  int passes_at[TRUSS_MAX_K];
  long long nnz_at[TRUSS_MAX_K];
  // list counters (last: the host reads only the fields above)
  alignas(128) int lcnt[4][PEEL_SEGS + 1][PEEL_CNT_STRIDE];
};
constexpr size_t PEEL_CTL_HOST_BYTES = offsetof(PeelCtl, lcnt);

// Items per list for a graph of nnz stored entries: a round or a scan lists
// each edge at most once, so nnz / 2 + 64 items always fit the overflow
// segment alone.
inline PeelLists peel_lists_geometry(int64_t nnz) {
  PeelLists L{};
  const int64_t lcap = nnz / 2 + 64;
  L.segcap = (int)(lcap / (2 * PEEL_SEGS) + 64);
  L.ovcap = (int)lcap;
  L.stride = (int64_t)PEEL_SEGS * L.segcap + L.ovcap;
  return L;
}

__device__ inline int* peel_lc(PeelCtl* __restrict__ ctl, int l) { return &ctl->lcnt[l][0][0]; }

// A wavefront's view of list l: lane s <= PEEL_SEGS holds the number of items
// before segment s (segment PEEL_SEGS: the overflow); total and over (the
// overflow outgrew its capacity) are the same in every lane.
struct PeelView {
  int pre, total;
  bool over;
};

__device__ inline PeelView peel_view(const PeelCtl* __restrict__ ctl, int l, const PeelLists& L, int lane) {
  const int raw = lane <= PEEL_SEGS ? ctl->lcnt[l][lane][0] : 0;
  const int sz = lane < PEEL_SEGS ? min(raw, L.segcap) : (lane == PEEL_SEGS ? min(raw, L.ovcap) : 0);
  const int inc = wave_incl_scan(sz);
  PeelView v;
  v.pre = inc - sz;
  v.total = __shfl(inc, 63);
  v.over = __ballot(lane == PEEL_SEGS && raw > L.ovcap) != 0;
  return v;
}

// Item t (0 <= t < total, the same in every lane) of a list.
__device__ inline PeelItem peel_item(const PeelItem* __restrict__ list, const PeelView& v, const PeelLists& L,
                                     int64_t t, int lane) {
  const unsigned long long b = __ballot(lane <= PEEL_SEGS && v.pre <= t);  // lane 0 always (pre = 0)
  const int s = 63 - __clzll(b);
  return list[(int64_t)s * L.segcap + (t - __shfl(v.pre, s))];
}

// The slot of item t of a list, t different per lane (0 <= t < total): the
// segment by a binary search over the wavefront's prefix (every lane runs the
// same steps).
__device__ inline int64_t peel_slot_lane(const PeelView& v, const PeelLists& L, int64_t t) {
  int lo = 0;  // the last segment s with pre[s] <= t (segment PEEL_SEGS: the overflow)
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const int cand = min(lo + step, PEEL_SEGS);
    if (__shfl(v.pre, cand) <= t) lo = cand;
  }
  return (int64_t)lo * L.segcap + (t - __shfl(v.pre, lo));
}

__device__ inline PeelItem peel_item_lane(const PeelItem* __restrict__ list, const PeelView& v, const PeelLists& L,
                                          int64_t t) {
  return list[peel_slot_lane(v, L, t)];
}

// Sums of a list's segment sizes from one thread (one-thread kernels).
__device__ inline int peel_total1(const PeelCtl* __restrict__ ctl, int l, const PeelLists& L, bool* over) {
  int tot = 0;
  for (int s = 0; s < PEEL_SEGS; ++s) tot += min(ctl->lcnt[l][s][0], L.segcap);
  const int ov = ctl->lcnt[l][PEEL_SEGS][0];
  if (over) *over = ov > L.ovcap;
  return tot + min(ov, L.ovcap);
}

// The slot of the r-th item (0-based) of a claim of n items that got `base`
// in segment seg and `obase` in the overflow (used only when the segment had
// fewer than n slots left); -1 past the overflow's capacity.
__device__ inline int64_t peel_slot(int seg, int base, int obase, int r, const PeelLists& L) {
  const int room = max(L.segcap - base, 0);
  if (r < room) return (int64_t)seg * L.segcap + base + r;
  const int o = obase + r - room;
  return o < L.ovcap ? (int64_t)PEEL_SEGS * L.segcap + o : -1;
}

// Claims n slots of list counters lc for the calling workgroup's segment:
// base in the segment, obase in the overflow (claimed only for the part that
// does not fit).  One thread calls it.
__device__ inline void peel_claim(int* __restrict__ lc, int n, const PeelLists& L, int& seg, int& base, int& obase) {
  seg = blockIdx.x % PEEL_SEGS;
  base = atomicAdd(&lc[seg * PEEL_CNT_STRIDE], n);
  const int room = max(L.segcap - base, 0);
  obase = room < n ? atomicAdd(&lc[PEEL_SEGS * PEEL_CNT_STRIDE], n - room) : 0;
}

// Wavefront-aggregated append of one item per selected lane.
__device__ inline void peel_push(bool sel, const PeelItem& item, PeelItem* __restrict__ list, int* __restrict__ lc,
                                 const PeelLists& L, int lane) {
  const unsigned long long m = __ballot(sel);
  if (!m) return;
  int seg = 0, base = 0, obase = 0;
  if (lane == 0) peel_claim(lc, __popcll(m), L, seg, base, obase);
  seg = __shfl(seg, 0);
  base = __shfl(base, 0);
  obase = __shfl(obase, 0);
  if (sel) {
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t at = peel_slot(seg, base, obase, __popcll(m & lt), L);
    if (at >= 0) list[at] = item;
  }
}

// host words: [0] list size, [1] alive, [2] run ended, [3] ticks, [4] sequence (release)
__device__ inline void peel_publish(long long* __restrict__ host, long long size, long long alive, int fin,
                                    long long ticks, long long seq) {
  if (!host) return;
  __hip_atomic_store(&host[0], size, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&host[1], alive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&host[2], (long long)fin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&host[3], ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&host[4], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// alive canonical entries with support below thr: stamped r, listed.  A
// workgroup takes 1024 entries at a time (4 per thread, coalesced) and
// claims their list slots with one claim in its segment.  The body holds
// __syncthreads: every thread of the workgroup calls it.
constexpr int PEEL_SCAN_PER = 4;
__device__ inline void peel_scan_body(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                      const int32_t* __restrict__ mirror, const int32_t* __restrict__ sup,
                                      int32_t* __restrict__ stamp, int64_t nnz, int thr, int r,
                                      PeelItem* __restrict__ list, int* __restrict__ lc, const PeelLists& L) {
  __shared__ int s_wcnt[4];
  __shared__ int s_claim[3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tile = (int64_t)blockDim.x * PEEL_SCAN_PER;
  for (int64_t base = blockIdx.x * tile; base < nnz; base += (int64_t)gridDim.x * tile) {
    unsigned long long mk[PEEL_SCAN_PER];
    int mm[PEEL_SCAN_PER];
    int wc = 0;
#pragma unroll
    for (int u = 0; u < PEEL_SCAN_PER; ++u) {
      const int64_t p = base + u * blockDim.x + threadIdx.x;
      bool sel = false;
      int m = 0;
      if (p < nnz) {
        m = mirror[p];
        const int sv = sup[p], sp = stamp[p];  // independent loads
        sel = m < p && sv < thr && sp == PEEL_ALIVE;
      }
      mm[u] = m;
      mk[u] = __ballot(sel);
      wc += __popcll(mk[u]);
    }
    if (lane == 0) s_wcnt[wv] = wc;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += s_wcnt[w];
      int seg = 0, cb = 0, ob = 0;
      if (tot) peel_claim(lc, tot, L, seg, cb, ob);
      s_claim[0] = seg;
      s_claim[1] = cb;
      s_claim[2] = ob;
    }
    __syncthreads();
    const int seg = s_claim[0], cb = s_claim[1], ob = s_claim[2];
    int at = 0;
    for (int w = 0; w < wv; ++w) at += s_wcnt[w];
#pragma unroll
    for (int u = 0; u < PEEL_SCAN_PER; ++u) {
      const int64_t p = base + u * blockDim.x + threadIdx.x;
      if ((mk[u] >> lane) & 1ull) {
        const int64_t slot = peel_slot(seg, cb, ob, at + __popcll(mk[u] & lt), L);
        const int j = col[p], i = col[mm[u]];  // p: row i, column j
        stamp[p] = r;
        stamp[mm[u]] = r;
        if (slot >= 0)
          list[slot] = PeelItem{make_int4((int32_t)p, mm[u], i, j)};
      }
      at += __popcll(mk[u]);
    }
    __syncthreads();  // s_wcnt and s_claim are reused by the next tile
  }
}

// The k after a stable k-truss (decomposition: k + 1 until the graph is
// empty; k list: the next k until the list ends).
__device__ inline void peel_next_k(const PeelCtl* __restrict__ ctl, long long alive, int& k, int& kidx, int& fin) {
  if (ctl->all_k) {
    if (alive == 0) fin = 1;
    else ++k;
  } else {
    if (++kidx == ctl->nk) fin = 1;
    else k = ctl->ks[kidx];
  }
}

// The counters after a round that removed nc edges.
__device__ inline void peel_count_round(PeelDyn& d, int all_k, int nc) {
  d.alive -= 2ll * nc;
  d.steps += 1;
  if (!all_k || d.alive > 0) d.passes += 1;  // a decomposition stops without a product at an empty graph
  d.r += 1;
}

__global__ void k_peel_ctl_init(PeelCtl* __restrict__ ctl, long long alive, int all_k, int nk, int k0,
                                long long passes) {
  if (blockIdx.x != 0) return;
  for (int i = threadIdx.x; i < 4 * (PEEL_SEGS + 1); i += blockDim.x) ctl->lcnt[i / (PEEL_SEGS + 1)][i % (PEEL_SEGS + 1)][0] = 0;
  if (threadIdx.x == 0) {
    PeelDyn d{};
    d.alive = alive;
    d.passes = passes;
    d.steps = 0;
    d.r = 0;
    d.k = k0;
    d.kidx = 0;
    d.fin = alive == 0 && all_k;
    d.kmax = 2;
    d.err = 0;
    ctl->bytes = 0;
    ctl->h = d;
    ctl->dyn[0] = ctl->dyn[1] = d;
    ctl->ic = 0;
    ctl->cr = ctl->clist = ctl->ccnt = 0;
    ctl->nk = nk;
    ctl->all_k = all_k;
  }
}

// A round removes the edges of a list: the triangles through each edge
// (i, j) (i > j: row i is the column of the transposed entry).  The shorter
// row is probed against the longer one: the longer row, or every stride-th
// entry of it when it is longer than the LDS sample, is staged in LDS; a
// probed column is found by a binary search of the sample and, between two
// sampled entries, of at most stride - 1 entries in global memory.  Each
// triangle the edge owns decrements the supports of its surviving edges, and
// an edge pushed below thr is stamped r + 1 and joins the next list.  The
// edge's own truss number is tval (decomposition).
struct PeelEdge {
  int c, mc, ab, ae, arow, bb, be, brow, stride, m;
};

__device__ inline PeelEdge peel_edge(const PeelItem& it, const int32_t* __restrict__ rp, int cap) {
  PeelEdge e;
  e.c = it.e.x;
  e.mc = it.e.y;
  const int i = it.e.z, j = it.e.w;
  const int ib = rp[i], ie = rp[i + 1], jb = rp[j], je = rp[j + 1];
  const bool pi = (ie - ib) <= (je - jb);  // probe row i
  e.ab = pi ? ib : jb;
  e.ae = pi ? ie : je;
  e.arow = pi ? i : j;
  e.bb = pi ? jb : ib;
  e.be = pi ? je : ie;
  e.brow = pi ? j : i;
  const int tl = e.be - e.bb;
  e.stride = tl > cap ? (tl + cap - 1) / cap : 1;
  e.m = (tl + e.stride - 1) / e.stride;
  return e;
}

// The same edge searched directly: the sample is the whole target row.
__device__ inline PeelEdge peel_direct(PeelEdge e) {
  e.stride = 1;
  e.m = e.be - e.bb;
  return e;
}

// A probed entry's column, stamp and transposed position, loaded before the
// target row's sample is staged so that both sets of loads are in flight
// together.  A lane past the probe row gets stamp -1, which every round
// (r >= 1) skips.
struct PeelQ {
  int w, sa, mq;
};

__device__ inline PeelQ peel_qload(int q, int qe, const int32_t* __restrict__ col, const int32_t* __restrict__ stamp,
                                   const int32_t* __restrict__ mirror) {
  PeelQ v{-1, -1, -1};
  if (q < qe) {
    v.w = col[q];
    v.sa = stamp[q];
    v.mq = mirror[q];
  }
  return v;
}

// One probed entry q of the edge's probe row: the decrements it owns and the
// edges to push (x = -1: none).  st: the target row's sample (LDS) or the row
// itself (global memory).
__device__ inline void peel_probe(const PeelEdge& e, int q, const PeelQ& v, const int32_t* __restrict__ st,
                                  const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                  const int32_t* __restrict__ mirror, int32_t* __restrict__ sup,
                                  int32_t* __restrict__ stamp, int thr, int r, PeelItem& push_a, PeelItem& push_b,
                                  long long& found) {
  push_a.e.x = push_b.e.x = -1;
  const int w = v.w, sa = v.sa, mq = v.mq;
  if (sa < r || w == e.brow) return;
  int lo = 0, hi = e.m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (st[mid] < w) lo = mid + 1;
    else hi = mid;
  }
  int pos = -1;
  if (lo < e.m && st[lo] == w) {
    pos = e.bb + lo * e.stride;
  } else if (e.stride > 1 && lo > 0) {
    int glo = e.bb + (lo - 1) * e.stride + 1;
    int ghi = lo < e.m ? e.bb + lo * e.stride : e.be;
    while (glo < ghi) {
      const int mid = (glo + ghi) >> 1;
      if (col[mid] < w) glo = mid + 1;
      else ghi = mid;
    }
    if (glo < e.be && col[glo] == w) pos = glo;
  }
  if (pos < 0) return;
  const int sb = stamp[pos];
  if (sb < r) return;
  const int mp = mirror[pos];
  const int cid_a = w < e.arow ? q : mq;
  const int cid_b = w < e.brow ? pos : mp;
  if ((sa == r && cid_a < e.c) || (sb == r && cid_b < e.c)) return;  // another removed edge owns it
  ++found;
  const bool xa = sa > r && atomicSub(&sup[cid_a], 1) == thr;  // edge (arow, w): entry q lies in row arow
  const bool xb = sb > r && atomicSub(&sup[cid_b], 1) == thr;  // edge (brow, w): entry pos lies in row brow
  if (!xa && !xb) return;
  if (xa) {
    push_a.e = make_int4(cid_a, cid_a == q ? mq : q, max(e.arow, w), min(e.arow, w));
    stamp[q] = r + 1;
    stamp[mq] = r + 1;
  }
  if (xb) {
    push_b.e = make_int4(cid_b, cid_b == pos ? mp : pos, max(e.brow, w), min(e.brow, w));
    stamp[pos] = r + 1;
    stamp[mp] = r + 1;
  }
}

constexpr int PEEL_WSTAGE = 1024;         // sample entries per wavefront (large rounds)
constexpr int PEEL_BSTAGE = 4096;         // sample entries per workgroup (small rounds)
constexpr int PEEL_SPLIT = 8;             // least workgroups per edge in small rounds
constexpr int PEEL_SPLIT_MAX = 32;        // most workgroups per edge in small rounds
constexpr int PEEL_SPLIT_MAX_EDGES = 64;  // rounds of at most this many edges split them by workgroups
constexpr int PEEL_STAGE_ROUND = 4096;    // rounds up to this size stage every long target row

// The round itself.  Rounds of more than PEEL_SPLIT_MAX_EDGES edges: a
// wavefront per (edge, part); when the round has fewer edges than the grid has
// wavefronts, each edge's probe row is split into 64-entry chunks dealt
// round-robin over nw / nc wavefronts (a hub edge's chunks no longer run one
// after another on one wavefront).  A wavefront probing a chunk searches the
// target row in global memory, or stages its sample first when the row is
// long.  Small rounds (few edges, usually long rows): PEEL_SPLIT to
// PEEL_SPLIT_MAX workgroups per edge, each staging the sample once and probing
// every split-th run of 256 probe entries.
struct PeelGraph {
  const int32_t* rp;
  const int32_t* col;
  const int32_t* mirror;
  const int32_t* eid;
  int32_t* sup;
  int32_t* stamp;
  int32_t* truss;
};

__device__ inline void peel_round_body(const PeelItem* __restrict__ cur, const PeelView& cv,
                                       PeelItem* __restrict__ next,
                                       int* __restrict__ nlc, const PeelLists& L, int r, int thr, int tval,
                                       const PeelGraph& g, long long* __restrict__ bytes_ctr,
                                       int32_t* __restrict__ s_stage) {
  const int lane = threadIdx.x & 63;
  const int nc = cv.total;
  long long bytes = 0, found = 0;
  if (thr <= 1) {
    // supports below 1: no triangle holds an edge of the round, so there is
    // nothing to intersect or decrement (the first round of k = 3); only the
    // truss numbers are written, an edge per thread
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x; b < nc; b += stride) {  // uniform in the wavefront
      const int64_t t = b + threadIdx.x;
      const PeelItem it = peel_item_lane(cur, cv, L, min<int64_t>(t, nc - 1));
      if (t < nc) {
        if (g.truss) {
          g.truss[g.eid[it.e.x]] = tval;
          g.truss[g.eid[it.e.y]] = tval;
        }
        if (bytes_ctr) bytes += 28;
      }
    }
  } else if (nc > PEEL_SPLIT_MAX_EDGES) {
    int32_t* st = s_stage + (threadIdx.x >> 6) * PEEL_WSTAGE;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    const int wpe = nc < nw ? (int)(nw / nc) : 1;  // wavefronts per edge
    const int64_t nv = (int64_t)nc * wpe;
    for (int64_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v < nv; v += nw) {
      const int64_t t = v / wpe;
      const int part = (int)(v - t * wpe);
      const PeelEdge e = peel_edge(peel_item(cur, cv, L, t, lane), g.rp, PEEL_WSTAGE);
      if (part == 0 && lane == 0) {
        if (g.truss) {
          g.truss[g.eid[e.c]] = tval;
          g.truss[g.eid[e.mc]] = tval;
        }
        if (bytes_ctr) bytes += 28 + 4ll * ((e.ae - e.ab) + (e.be - e.bb));
      }
      const int q0 = e.ab + part * 64;
      if (q0 >= e.ae) continue;  // no chunk of this edge for this part (uniform in the wavefront)
      // a long target row is staged when several chunks probe it or the
      // round is small enough that the staging loads are cheaper than the
      // chains of a binary search in global memory
      const bool staged = e.be - e.bb > 64 && (e.ae - e.ab > 64 || nc <= PEEL_STAGE_ROUND);
      const PeelEdge pe = staged ? e : peel_direct(e);
      __threadfence_block();  // the previous edge's sample reads are done (waits for every load in flight)
      // the first chunk's probe loads, in flight with the staging loads
      PeelQ pq = peel_qload(q0 + lane, e.ae, g.col, g.stamp, g.mirror);
      if (staged)
        for (int u = lane; u < e.m; u += 64) st[u] = g.col[e.bb + u * e.stride];
      __threadfence_block();
      for (int qb = q0; qb < e.ae; qb += 64 * wpe) {
        const int q = qb + lane;
        if (qb != q0) pq = peel_qload(q, e.ae, g.col, g.stamp, g.mirror);
        PeelItem pa, pb;
        peel_probe(pe, q, pq, staged ? st : g.col + e.bb, g.rp, g.col, g.mirror, g.sup, g.stamp, thr, r, pa, pb,
                   found);
        peel_push(pa.e.x >= 0, pa, next, nlc, L, lane);
        peel_push(pb.e.x >= 0, pb, next, nlc, L, lane);
      }
    }
  } else {
    const int split = max(PEEL_SPLIT, min(PEEL_SPLIT_MAX, (int)(gridDim.x / max(nc, 1))));
    for (int64_t b = blockIdx.x; b < (int64_t)nc * split; b += gridDim.x) {
      const int t = (int)(b / split), part = (int)(b % split);
      const PeelEdge e = peel_edge(peel_item(cur, cv, L, t, lane), g.rp, PEEL_BSTAGE);
      if (part == 0 && threadIdx.x == 0) {
        if (g.truss) {
          g.truss[g.eid[e.c]] = tval;
          g.truss[g.eid[e.mc]] = tval;
        }
        if (bytes_ctr) bytes += 28 + 4ll * ((e.ae - e.ab) + (e.be - e.bb));
      }
      const int q0 = e.ab + part * 256;
      if (q0 >= e.ae) continue;  // uniform in the workgroup
      __syncthreads();  // the previous edge's sample reads are done (waits for every load in flight)
      PeelQ pq = peel_qload(q0 + (int)threadIdx.x, e.ae, g.col, g.stamp, g.mirror);  // in flight with the staging
      for (int u = threadIdx.x; u < e.m; u += blockDim.x) s_stage[u] = g.col[e.bb + u * e.stride];
      __syncthreads();
      for (int qb = q0; qb < e.ae; qb += split * 256) {
        const int q = qb + threadIdx.x;
        if (qb != q0) pq = peel_qload(q, e.ae, g.col, g.stamp, g.mirror);
        PeelItem pa, pb;
        peel_probe(e, q, pq, s_stage, g.rp, g.col, g.mirror, g.sup, g.stamp, thr, r, pa, pb, found);
        peel_push(pa.e.x >= 0, pa, next, nlc, L, lane);
        peel_push(pb.e.x >= 0, pb, next, nlc, L, lane);
      }
    }
  }
  if (bytes_ctr) {
    bytes = wave_reduce_sum(bytes + 16 * found);
    if (lane == 0 && bytes) atomicAdd((unsigned long long*)bytes_ctr, (unsigned long long)bytes);
  }
}

// ---- host-steered runs: a kernel per round or scan, then k_peel_post ------
__global__ __launch_bounds__(256) void k_peel_scan(PeelLists L, PeelCtl* __restrict__ ctl,
                                                   const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ mirror,
                                                   const int32_t* __restrict__ sup, int32_t* __restrict__ stamp,
                                                   int64_t nnz) {
  const int ic = ctl->ic;
  peel_scan_body(rp, col, mirror, sup, stamp, nnz, ctl->h.k - 2, ctl->h.r + 1, L.items + ic * L.stride,
                 peel_lc(ctl, ic), L);
}

__global__ __launch_bounds__(256) void k_peel_destroy(PeelLists L, PeelCtl* __restrict__ ctl, PeelGraph g,
                                                      int account) {
  __shared__ int32_t s_stage[PEEL_BSTAGE];
  const int ic = ctl->ic;
  const PeelView cv = peel_view(ctl, ic, L, threadIdx.x & 63);
  if (ctl->h.fin || cv.total == 0) return;
  peel_round_body(L.items + ic * L.stride, cv, L.items + (ic ^ 1) * L.stride, peel_lc(ctl, ic ^ 1), L,
                  ctl->h.r + 1, ctl->h.k - 2, ctl->h.k - 1, g, account ? &ctl->bytes : nullptr, s_stage);
}

// mode 0: a round ran -- counted, its list emptied, the lists swap; mode 1:
// a scan ran.  Then the list size of the next round to the host.
__global__ void k_peel_post(PeelCtl* __restrict__ ctl, int mode, PeelLists L, long long* __restrict__ host,
                            long long seq) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  bool over0 = false, over1 = false;
  if (mode == 0 && !ctl->h.fin) {
    const int nc = peel_total1(ctl, ctl->ic, L, &over0);
    if (nc > 0) {
      peel_count_round(ctl->h, ctl->all_k, nc);
      for (int s = 0; s <= PEEL_SEGS; ++s) ctl->lcnt[ctl->ic][s][0] = 0;  // consumed; the round after next fills it
      ctl->ic ^= 1;
    }
  }
  const int size = peel_total1(ctl, ctl->ic, L, &over1);
  if (over0 || over1) ctl->h.err = ctl->h.fin = 1;
  peel_publish(host, size, ctl->h.alive, ctl->h.fin, 0, seq);
}

// Host-steered runs: the k to scan next, and the filter step that found the
// previous k-truss stable.
__global__ void k_peel_set_k(PeelCtl* __restrict__ ctl, int k, int stable_step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->h.k = k;
    ctl->h.steps += stable_step;
  }
}

// ---- device-driven runs: one kernel per tick --------------------------------
// dyn[t & 1] is the state tick t starts from.  Tick t's workgroup 0 stores
// the state tick t + 1 starts from -- tick t's start state moved on by what
// tick t does with its list of nc items: a round (nc > 0: counted) or the
// stable step (nc == 0: recorded, k moved on; tick t scans the new
// threshold) -- so a tick reads one state and one list's counters.  Tick 0
// is the first threshold's scan; ticks 0 and 1 start from the initial state.
__device__ inline PeelDyn peel_after_tick(const PeelCtl* __restrict__ ctl, PeelDyn d, int nc, int& rec_kidx,
                                          long long& rec_alive, int& rec_passes) {
  rec_kidx = -1;
  if (d.fin) return d;
  if (nc > 0) {
    peel_count_round(d, ctl->all_k, nc);
  } else {
    if (ctl->all_k) {
      if (d.alive > 0) {
        d.kmax = d.k;
        d.steps += 1;
      }
    } else {
      rec_kidx = d.kidx;
      rec_alive = d.alive;
      rec_passes = (int)d.passes;
      d.steps += 1;
    }
    int k = d.k, kidx = d.kidx, fin = 0;
    peel_next_k(ctl, d.alive, k, kidx, fin);
    d.k = k;
    d.kidx = kidx;
    d.fin = fin;
  }
  return d;
}

__global__ __launch_bounds__(256) void k_peel_tick(int64_t t, PeelLists L, PeelCtl* __restrict__ ctl, PeelGraph g,
                                                   int64_t nnz, int account, long long* __restrict__ host) {
  __shared__ int32_t s_stage[PEEL_BSTAGE];
  const int lane = threadIdx.x & 63;
  const PeelDyn d = ctl->dyn[t & 1];
  const PeelView cv = peel_view(ctl, (int)(t & 3), L, lane);
  if (blockIdx.x == 0 && threadIdx.x <= PEEL_SEGS) ctl->lcnt[(t + 2) & 3][threadIdx.x][0] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int rec_kidx;
    long long rec_alive = 0;
    int rec_passes = 0;
    PeelDyn dn = peel_after_tick(ctl, d, cv.total, rec_kidx, rec_alive, rec_passes);
    if (cv.over) dn.err = dn.fin = 1;
    ctl->dyn[(t + 1) & 1] = dn;
    if (rec_kidx >= 0) {
      ctl->nnz_at[rec_kidx] = rec_alive;
      ctl->passes_at[rec_kidx] = rec_passes;
    }
    if ((t & (PEEL_REPORT - 1)) == 0 || d.fin) peel_publish(host, cv.total, d.alive, d.fin, t, 0);
  }
  if (d.fin || cv.over) return;
  const int nl = (int)((t + 1) & 3);
  PeelItem* next = L.items + nl * L.stride;
  if (cv.total > 0) {
    peel_round_body(L.items + (t & 3) * L.stride, cv, next, peel_lc(ctl, nl), L, d.r + 1, d.k - 2, d.k - 1, g,
                    account ? &ctl->bytes : nullptr, s_stage);
  } else {
    int k = d.k, kidx = d.kidx, fin = 0;  // the k-truss is stable: scan the next threshold
    peel_next_k(ctl, d.alive, k, kidx, fin);
    if (!fin && d.alive > 0) {
      peel_scan_body(g.rp, g.col, g.mirror, g.sup, g.stamp, nnz, k - 2, d.r + 1, next, peel_lc(ctl, nl), L);
      // accounting: the scan reads mirror, support and stamp of every stored entry
      if (account && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&ctl->bytes, 12ull * nnz);
    }
  }
}

// The device-driven queue drained after tick t: the state tick t + 1 will
// start from, to the host (sequence seq) and as the compaction view.  One
// wavefront.
__global__ void k_peel_drain(int64_t t, PeelCtl* __restrict__ ctl, PeelLists L, long long* __restrict__ host,
                             long long seq) {
  if (blockIdx.x != 0 || threadIdx.x >= 64) return;
  const PeelView cv = peel_view(ctl, (int)((t + 1) & 3), L, threadIdx.x);
  if (threadIdx.x != 0) return;
  const PeelDyn d = ctl->dyn[(t + 1) & 1];
  ctl->cr = d.r;
  ctl->clist = (int)((t + 1) & 3);
  ctl->ccnt = d.fin ? 0 : cv.total;
  peel_publish(host, ctl->ccnt, d.alive, d.fin || cv.over, t, seq);
}

// Host-steered runs before a compaction (a k boundary: no list pending).
__global__ void k_peel_cview_host(PeelCtl* __restrict__ ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->cr = ctl->h.r;
    ctl->clist = ctl->ic;
    ctl->ccnt = 0;
  }
}

// Entries of no round still pending or to come (stamp <= cr) lose their
// support value (-1), so a compaction at threshold 0 keeps exactly the alive
// ones, the ones of the pending list included.
__global__ void k_peel_drop_dead(const int32_t* __restrict__ stamp, int32_t* __restrict__ sup, int64_t nnz,
                                 const PeelCtl* __restrict__ ctl) {
  const int r = ctl->cr;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x)
    if (stamp[p] <= r) sup[p] = -1;
}

// After a compaction: the pending list's entries move to their new positions
// (rows of the compacted graph, row pointers nrp) and keep their stamp cr + 1
// (both directions); every other stamp was reset.  An item per thread.
__global__ void k_peel_remap(PeelLists L, const PeelCtl* __restrict__ ctl, const int32_t* __restrict__ newpos,
                             const int32_t* __restrict__ nrp, int32_t* __restrict__ stamp) {
  const int n = ctl->ccnt;
  if (n == 0) return;
  const int l = ctl->clist, r1 = ctl->cr + 1;
  PeelItem* list = L.items + l * L.stride;
  const PeelView v = peel_view(ctl, l, L, threadIdx.x & 63);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x; b < n; b += stride) {  // uniform in the wavefront
    const int64_t t = b + threadIdx.x;
    const int64_t slot = peel_slot_lane(v, L, min<int64_t>(t, n - 1));
    if (t < n) {
      const PeelItem it = list[slot];
      const int c = newpos[it.e.x], m = newpos[it.e.y], i = it.e.z, j = it.e.w;  // the endpoints do not move
      list[slot] = PeelItem{make_int4(c, m, i, j)};
      stamp[c] = r1;
      stamp[m] = r1;
    }
  }
}

// Tick 0 of a device-driven run: the first threshold's scan into list 1.
__global__ __launch_bounds__(256) void k_peel_tick0(PeelLists L, PeelCtl* __restrict__ ctl,
                                                    const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ mirror,
                                                    const int32_t* __restrict__ sup, int32_t* __restrict__ stamp,
                                                    int64_t nnz) {
  peel_scan_body(rp, col, mirror, sup, stamp, nnz, ctl->dyn[0].k - 2, ctl->dyn[0].r + 1, L.items + L.stride,
                 peel_lc(ctl, 1), L);
}

// The final compaction of a finished run (no list pending).
__global__ void k_peel_cview_end(PeelCtl* __restrict__ ctl, int r) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->cr = r;
    ctl->clist = 0;
    ctl->ccnt = 0;
  }
}

static hipEvent_t peel_event(Context& c) {
  hipEvent_t ev;
  if (!c.event_pool.empty()) {
    ev = c.event_pool.back();
    c.event_pool.pop_back();
  } else {
    GA_HIP(hipEventCreate(&ev));
  }
  return ev;
}

// Peeling with support updates (GA_OPT_TRUSS_LOOP 3 and 4): the buffers and
// the first graph are set up by truss_loop; arguments and results as there.
// host_steer (always with a k callback): the host reads each round's list
// size, answers each k as it is reached and takes on the callback's extra k
// values.  Otherwise the rounds run as device ticks queued a few ahead.
static void truss_peel(const DCsr& A, bool all_k, const int32_t* ks, int nk, TrussWork& tw, hipStream_t s,
                       bool host_steer, int* cur_out, int64_t* nnz_out, int32_t* kmax_out, ga_truss_stats* st,
                       int64_t* nnz_at, int32_t* passes_at, const KDone* on_k) {
  const int64_t n = A.nrows, nnz0 = A.nnz;
  Context& c = ctx();
  host_steer = host_steer || on_k != nullptr;
  GA_REQUIRE(all_k || host_steer || nk <= TRUSS_MAX_K, GA_ERR_LIMIT, "at most 64 k values in one k-truss run");
  PeelLists L = peel_lists_geometry(nnz0);
  const int64_t lcap = nnz0 / 2 + 64;  // canonical entries: every list holds at most this many
  tw.plist.ensure(sizeof(PeelItem) * 4 * L.stride);
  tw.pcnt.ensure(sizeof(PeelCtl));
  tw.stamp.ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
  L.items = tw.plist.as<PeelItem>();
  auto* ctl = tw.pcnt.as<PeelCtl>();
  int32_t* stamp = tw.stamp.as<int32_t>();
  const int64_t passes0 = (all_k && nnz0 == 0) ? 0 : 1;  // a decomposition of no entries runs no product
  hipLaunchKernelGGL(k_peel_ctl_init, dim3(1), dim3(64), 0, s, ctl, (long long)nnz0, all_k ? 1 : 0, all_k ? 0 : nk,
                     all_k ? 3 : ks[0], (long long)passes0);
  if (!all_k && !host_steer)
    GA_HIP(hipMemcpyAsync(reinterpret_cast<char*>(ctl) + offsetof(PeelCtl, ks), ks, sizeof(int32_t) * nk,
                          hipMemcpyHostToDevice, s));
  fill_u32(stamp, (uint32_t)PEEL_ALIVE, nnz0, s);
  const bool account = c.prof_bytes;
  // the first supports: one full product of the input graph (outside the
  // kernel timing and byte accounting, which cover the peeling launches)
  int cur = 0;
  {
    const DCsr C{n, n, nnz0, tw.cur[0].rp.as<int32_t>(), tw.cur[0].col.as<int32_t>(), nullptr};
    tw.mw.balance = false;
    struct Hooks {
      Context& c;
      bool ev, by;
      ~Hooks() {
        c.prof_events_on = ev;
        c.prof_bytes = by;
      }
    } hooks{c, c.prof_events_on, c.prof_bytes};
    c.prof_events_on = false;
    c.prof_bytes = false;
    support_symmetric_dev(C, tw.mirror[0].as<int32_t>(), tw.sup[0].as<int32_t>(), tw.mw, s, false);
  }
  int64_t stored = nnz0;  // entries in the current buffers (alive or not)
  auto graph = [&]() {
    return PeelGraph{tw.cur[cur].rp.as<int32_t>(), tw.cur[cur].col.as<int32_t>(), tw.mirror[cur].as<int32_t>(),
                     tw.eid[cur].as<int32_t>(), tw.sup[cur].as<int32_t>(), stamp,
                     all_k ? tw.truss.as<int32_t>() : nullptr};
  };
  long long* hw_dev = nullptr;
  long long* hw = progress_words(&hw_dev) + 16;  // words 16..20 (see peel_publish)
  hw_dev += 16;
  const unsigned cap = (unsigned)(8 * (int64_t)c.num_cu);
  const unsigned gscan = std::min<unsigned>(grid_for(std::max<int64_t>(stored, 1), 256, 1 << 20), cap);
  // round launches: every workgroup of a tick is queued before its list size
  // is known; rounds smaller than the grid spread each edge over several
  // wavefronts (peel_round_body)
  // (4 workgroups per CU measured best on R-MAT 16 against 2, 6 and 8)
  // six workgroups of 256 threads per CU (R-MAT 16 decomposition 10.81 ms;
  // three 11.25, four 10.91, five 10.84, eight 11.70)
  const unsigned gtick = std::min<unsigned>(gscan, 6 * (unsigned)c.num_cu);
  auto timed = [&](auto&& launch) {  // a round launch, bracketed by events when profiling
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c.prof_events_on) {
      e0 = peel_event(c);
      e1 = peel_event(c);
      GA_HIP(hipEventRecord(e0, s));
    }
    if (account) c.prof_bytes_launches += 1;
    launch();
    GA_HIP(hipGetLastError());
    if (c.prof_events_on) {
      GA_HIP(hipEventRecord(e1, s));
      c.prof_events.emplace_back(e0, e1);
      c.prof_launches += 1;
    }
  };
  // compaction (every kernel before it in stream order is done): the dead
  // entries go, supports carried; the pending list moves with its entries
  auto compact = [&](int64_t alive) {
    hipLaunchKernelGGL(k_peel_drop_dead, dim3(grid_for(stored, 256, 8192)), dim3(256), 0, s, (const int32_t*)stamp,
                       tw.sup[cur].as<int32_t>(), stored, (const PeelCtl*)ctl);
    int32_t* kept_dev = tw.kept.as<int32_t>();
    long long* bsum = tw.scan_tmp.as<long long>();
    const int64_t nbk = ceil_div(stored, KEEP_TILE);
    hipLaunchKernelGGL(k_keep_tile_sums, dim3((unsigned)nbk), dim3(KEEP_T), 0, s,
                       (const int32_t*)tw.sup[cur].as<int32_t>(), stored, 0, bsum, (const TrussDevState*)nullptr,
                       (const int32_t*)nullptr, (int32_t*)nullptr);
    const bool offsets = nbk > KEEP_FUSED_TILES;
    if (offsets) scan_tile_offsets(bsum, nbk, kept_dev, s);
    hipLaunchKernelGGL(k_keep_scatter, dim3((unsigned)nbk), dim3(KEEP_T), 0, s,
                       (const int32_t*)tw.cur[cur].col.as<int32_t>(), (const int32_t*)tw.eid[cur].as<int32_t>(),
                       (const int32_t*)tw.sup[cur].as<int32_t>(), stored, 0, (const long long*)bsum, offsets ? 1 : 0,
                       tw.cur[cur ^ 1].col.as<int32_t>(), tw.eid[cur ^ 1].as<int32_t>(), tw.nold.as<int32_t>(),
                       tw.newpos.as<int32_t>(), tw.sup[cur ^ 1].as<int32_t>(), 1, kept_dev,
                       (const TrussDevState*)nullptr, (long long*)nullptr, 0ll);
    hipLaunchKernelGGL(k_keep_finish, dim3(grid_for(std::max<int64_t>(n + 1, stored), 256, 8192)), dim3(256), 0, s,
                       (const int32_t*)tw.cur[cur].rp.as<int32_t>(), n, (const int32_t*)tw.newpos.as<int32_t>(),
                       stored, (const int32_t*)kept_dev, tw.cur[cur ^ 1].rp.as<int32_t>(), (uint8_t*)nullptr,
                       (const int32_t*)tw.mirror[cur].as<int32_t>(), (const int32_t*)tw.nold.as<int32_t>(),
                       tw.mirror[cur ^ 1].as<int32_t>(), (const TrussDevState*)nullptr);
    fill_u32(stamp, (uint32_t)PEEL_ALIVE, alive, s);
    hipLaunchKernelGGL(k_peel_remap, dim3(grid_for(lcap, 256, 4096)), dim3(256), 0, s, L,
                       (const PeelCtl*)ctl, (const int32_t*)tw.newpos.as<int32_t>(),
                       (const int32_t*)tw.cur[cur ^ 1].rp.as<int32_t>(), stamp);
    GA_HIP(hipGetLastError());
    cur ^= 1;
    stored = alive;
  };
  // compact once a third of the stored entries are dead (R-MAT 16
  // decomposition 11.2 ms when compacting at half, 11.0 at a third, 11.1 at
  // a quarter; 11.5 and 11.7 at two thirds and three quarters): rounds probe
  // rows with fewer dead entries in them, for a few more compactions
  auto worth_compacting = [&](int64_t alive) { return alive > 0 && 3 * alive <= 2 * stored; };
  struct DrainOnThrow {  // an error leaves no kernel running on buffers the caller may free
    hipStream_t s;
    bool armed = true;
    ~DrainOnThrow() {
      if (armed) (void)hipStreamSynchronize(s);
    }
  } drain{s};
  auto list_size = [&]() -> int64_t {
    const int64_t v = __atomic_load_n(&hw[0], __ATOMIC_RELAXED);
    GA_REQUIRE(v >= 0 && v <= lcap && !__atomic_load_n(&hw[2], __ATOMIC_RELAXED), GA_ERR_STATE,
               "peeling list overflow");
    return v;
  };
  PeelDyn out{};
  if (host_steer) {
    std::vector<int32_t> kv(ks, ks + (all_k ? 0 : nk));
    std::vector<int64_t> nz_v(kv.size(), 0);
    std::vector<int32_t> ps_v(kv.size(), 0);
    int kidx = 0, k = all_k ? 3 : kv[0];
    int32_t kmax = 2;
    int64_t alive = nnz0, passes = passes0, ncur = 0;
    auto scan = [&]() {
      const long long seq = next_host_seq();
      hipLaunchKernelGGL(k_peel_scan, dim3(gscan), dim3(256), 0, s, L, ctl,
                         (const int32_t*)tw.cur[cur].rp.as<int32_t>(), (const int32_t*)tw.cur[cur].col.as<int32_t>(),
                         (const int32_t*)tw.mirror[cur].as<int32_t>(),
                         (const int32_t*)tw.sup[cur].as<int32_t>(),
                         stamp, stored);
      hipLaunchKernelGGL(k_peel_post, dim3(1), dim3(64), 0, s, ctl, 1, L, hw_dev, seq);
      GA_HIP(hipGetLastError());
      wait_host_word(&hw[4], seq, s);
      ncur = list_size();
    };
    if (alive > 0) scan();
    const int64_t max_rounds = nnz0 / 2 + n + 4 * (int64_t)kv.size() + 64;  // a round removes an edge or moves k on
    int64_t rounds = 0;
    while (true) {
      GA_REQUIRE(++rounds <= max_rounds && k <= n + 3, GA_ERR_STATE, "peeling did not converge");
      if (ncur == 0) {  // the k-truss is stable: one filter step that removed nothing
        if (all_k && alive == 0) break;
        bool last = false;
        if (all_k) {
          kmax = k++;
        } else {
          nz_v[kidx] = alive;
          ps_v[kidx] = (int32_t)passes;
          if (on_k) {
            std::vector<int32_t> more;
            (*on_k)(kv[kidx], alive, (int32_t)passes, &more);
            for (int32_t x : more) {
              if (x <= kv[kidx]) continue;  // only larger k can continue from this truss
              auto itx = std::lower_bound(kv.begin() + kidx + 1, kv.end(), x);
              if (itx != kv.end() && *itx == x) continue;
              const auto at = itx - kv.begin();
              kv.insert(itx, x);
              nz_v.insert(nz_v.begin() + at, 0);
              ps_v.insert(ps_v.begin() + at, 0);
            }
          }
          last = ++kidx == (int)kv.size();
          if (!last) k = kv[kidx];
        }
        hipLaunchKernelGGL(k_peel_set_k, dim3(1), dim3(64), 0, s, ctl, k, 1);
        if (last) break;
        if (alive == 0) continue;  // (a k list past an empty truss: every later k is empty too)
        if (worth_compacting(alive)) {
          hipLaunchKernelGGL(k_peel_cview_host, dim3(1), dim3(64), 0, s, ctl);
          compact(alive);
        }
        scan();
        continue;
      }
      const long long seq = next_host_seq();
      const PeelGraph g = graph();
      const unsigned grid = ncur <= PEEL_SPLIT_MAX_EDGES ? (unsigned)(ncur * PEEL_SPLIT) : gtick;
      timed([&] {
        hipLaunchKernelGGL(k_peel_destroy, dim3(grid), dim3(256), 0, s, L, ctl, g,
                           account ? 1 : 0);
      });
      hipLaunchKernelGGL(k_peel_post, dim3(1), dim3(64), 0, s, ctl, 0, L, hw_dev, seq);
      GA_HIP(hipGetLastError());
      wait_host_word(&hw[4], seq, s);
      alive -= 2 * ncur;
      if (!all_k || alive > 0) ++passes;
      ncur = list_size();
    }
    PeelCtl h{};
    GA_HIP(hipMemcpyAsync(&h, ctl, PEEL_CTL_HOST_BYTES, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    GA_REQUIRE(!h.h.err && h.h.alive == alive && h.h.passes == passes, GA_ERR_STATE,
               "peeling state disagrees with the host");
    for (int i = 0; i < (all_k ? 0 : nk); ++i) {  // the caller's k values (joined ones went to on_k)
      const auto at = std::lower_bound(kv.begin(), kv.end(), ks[i]) - kv.begin();
      if (nnz_at) nnz_at[i] = nz_v[at];
      if (passes_at) passes_at[i] = ps_v[at];
    }
    out = h.h;
    out.kmax = kmax;
  } else {
    // device-driven: ticks queued a few ahead of the device, which reports
    // every PEEL_REPORT ticks and when the run ends; to compact, the host drains
    // the queue through k_peel_drain and compacts when the dead entries are
    // at least a third of the stored ones
    constexpr int64_t kInFlight = 12;
    static_assert(kInFlight > PEEL_REPORT && (PEEL_REPORT & (PEEL_REPORT - 1)) == 0, "a queued tick must report");
    const int64_t max_ticks = nnz0 / 2 + n + nk + 64;  // a tick removes an edge or moves k on
    __atomic_store_n(&hw[1], (long long)nnz0, __ATOMIC_RELAXED);
    __atomic_store_n(&hw[2], 0ll, __ATOMIC_RELAXED);
    __atomic_store_n(&hw[3], 0ll, __ATOMIC_RELEASE);
    if (nnz0 > 0)  // tick 0: the first threshold, into list 1
      hipLaunchKernelGGL(k_peel_tick0, dim3(gscan), dim3(256), 0, s, L, ctl,
                         (const int32_t*)tw.cur[0].rp.as<int32_t>(), (const int32_t*)tw.cur[0].col.as<int32_t>(),
                         (const int32_t*)tw.mirror[0].as<int32_t>(),
                         (const int32_t*)tw.sup[0].as<int32_t>(), stamp, stored);
    int64_t t = 0;  // the last tick queued
    auto t_seen = std::chrono::steady_clock::now();
    int64_t seen = -1;
    while (true) {
      if (__atomic_load_n(&hw[2], __ATOMIC_ACQUIRE)) break;
      const int64_t done_t = __atomic_load_n(&hw[3], __ATOMIC_ACQUIRE);
      const int64_t alive_seen = __atomic_load_n(&hw[1], __ATOMIC_RELAXED);
      if (worth_compacting(alive_seen)) {
        const long long seq = next_host_seq();
        hipLaunchKernelGGL(k_peel_drain, dim3(1), dim3(64), 0, s, t, ctl, L, hw_dev, seq);
        GA_HIP(hipGetLastError());
        wait_host_word(&hw[4], seq, s);
        if (__atomic_load_n(&hw[2], __ATOMIC_RELAXED)) break;
        const int64_t alive = __atomic_load_n(&hw[1], __ATOMIC_RELAXED);
        if (worth_compacting(alive)) compact(alive);
        continue;
      }
      if (t - done_t < kInFlight) {
        GA_REQUIRE(t < max_ticks, GA_ERR_STATE, "peeling did not converge");
        ++t;
        const PeelGraph g = graph();
        timed([&] {
          hipLaunchKernelGGL(k_peel_tick, dim3(gtick), dim3(256), 0, s, t, L, ctl, g, stored,
                             account ? 1 : 0, hw_dev);
        });
        continue;
      }
      const auto now = std::chrono::steady_clock::now();
      if (done_t != seen) {
        seen = done_t;
        t_seen = now;
      } else if (now - t_seen > std::chrono::seconds(60)) {
        GA_HIP(hipStreamSynchronize(s));  // a tick that ran for a minute: report, do not spin on
        GA_REQUIRE(__atomic_load_n(&hw[3], __ATOMIC_ACQUIRE) != seen || __atomic_load_n(&hw[2], __ATOMIC_ACQUIRE),
                   GA_ERR_STATE, "peeling made no progress");
      }
    }
    // one more tick stores the final state (and the last k's record) when the
    // end was seen by a drain; after a tick that ended the run it changes nothing
    ++t;
    {
      const PeelGraph g = graph();
      hipLaunchKernelGGL(k_peel_tick, dim3(gtick), dim3(256), 0, s, t, L, ctl, g, stored, 0,
                         (long long*)nullptr);
      GA_HIP(hipGetLastError());
    }
    PeelCtl h{};
    GA_HIP(hipMemcpyAsync(&h, ctl, PEEL_CTL_HOST_BYTES, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    out = h.dyn[(t + 1) & 1];  // the state after the last tick
    GA_REQUIRE(out.fin && !out.err, GA_ERR_STATE, "peeling did not finish (a removal list outgrew its capacity)");
    for (int i = 0; i < (all_k ? 0 : nk); ++i) {
      if (nnz_at) nnz_at[i] = h.nnz_at[i];
      if (passes_at) passes_at[i] = h.passes_at[i];
    }
  }
  if (cur_out && stored != out.alive) {
    hipLaunchKernelGGL(k_peel_cview_end, dim3(1), dim3(64), 0, s, ctl, out.r);
    compact(out.alive);
  }
  GA_HIP(hipStreamSynchronize(s));
  drain.armed = false;
  if (account) {
    long long bytes = 0;
    GA_HIP(hipMemcpy(&bytes, &ctl->bytes, sizeof(bytes), hipMemcpyDeviceToHost));
    c.prof_bytes_host += bytes;
  }
  if (cur_out) *cur_out = cur;
  if (nnz_out) *nnz_out = out.alive;
  if (kmax_out) *kmax_out = out.kmax;
  if (st) {
    st->passes = out.passes;
    st->steps = out.steps;
  }
}

// Runs the peeling loop on the device-resident graph A.  all_k: truss
// decomposition (truss[] per entry of A); otherwise k-truss for one k, and
// the surviving pattern is left in tw.cur[*cur_out].
// k-truss mode takes an ascending list of k: the loop peels at ks[0] until
// stable, records the surviving entries and passes so far, and continues
// from that truss to ks[1] (the k2-truss lies inside the k1-truss), ...
static void truss_loop(const DCsr& A, bool all_k, const int32_t* ks, int nk, TrussWork& tw, hipStream_t s,
                       int* cur_out, int64_t* nnz_out, int32_t* kmax_out, ga_truss_stats* st,
                       int64_t* nnz_at = nullptr, int32_t* passes_at = nullptr, const KDone* on_k = nullptr) {
  auto t_start = std::chrono::steady_clock::now();
  const int64_t n = A.nrows, nnz0 = A.nnz;
  GA_REQUIRE(A.nrows == A.ncols, GA_ERR_DIM, "k-truss needs a square matrix");
  for (int b = 0; b < 2; ++b) {
    tw.cur[b].nrows = n;
    tw.cur[b].ncols = n;
    tw.cur[b].rp.ensure(sizeof(int32_t) * (n + 1));
    tw.cur[b].col.ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
    tw.eid[b].ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
    tw.mirror[b].ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
    tw.sup[b].ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
  }
  tw.nold.ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
  tw.newpos.ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
  tw.dirty.ensure(std::max<int64_t>(n, 1));
  tw.kept.ensure(64);
  tw.scan_tmp.ensure(sizeof(long long) * std::max<int64_t>(ceil_div(nnz0, KEEP_TILE), 1));
  tw.mw.balance = true;  // the support passes split rows by the previous pass's work
  tw.mw.rwork_have = false;
  GA_HIP(hipMemcpyAsync(tw.cur[0].rp.p, A.rp, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, s));
  if (nnz0) GA_HIP(hipMemcpyAsync(tw.cur[0].col.p, A.col, sizeof(int32_t) * nnz0, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz0, 256, 8192)), dim3(256), 0, s, tw.eid[0].as<int32_t>(), nnz0);
  mirror_index_dev(DCsr{n, n, nnz0, tw.cur[0].rp.as<int32_t>(), tw.cur[0].col.as<int32_t>(), nullptr},
                   tw.mirror[0].as<int32_t>(), s);
  if (all_k) {
    tw.truss.ensure(sizeof(int32_t) * std::max<int64_t>(nnz0, 1));
    fill_u32(tw.truss.p, 2u, nnz0, s);
  }
  GA_REQUIRE(all_k || (ks && nk >= 1), GA_ERR_ARG, "no k given");
  {
    Context& c = ctx();
    const int mode = c.truss_loop >= 0 ? c.truss_loop : 3;
    if (mode >= 3) {
      truss_peel(A, all_k, ks, nk, tw, s, mode == 4, cur_out, nnz_out, kmax_out, st, nnz_at, passes_at, on_k);
      if (st)
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
      return;
    }
    if (mode > 0 && !selective_on(c, nnz0) && !c.prof_bytes) {
      std::vector<int64_t> nz_at(all_k ? 0 : nk);
      std::vector<int32_t> ps_at(all_k ? 0 : nk);
      truss_steps(A, all_k, ks, nk, tw, s, mode == 2 && !tw.no_graph, nnz_out, kmax_out, st, nz_at.data(),
                  ps_at.data());
      for (int i = 0; i < (int)nz_at.size(); ++i) {
        if (nnz_at) nnz_at[i] = nz_at[i];
        if (passes_at) passes_at[i] = ps_at[i];
        if (on_k) (*on_k)(ks[i], nz_at[i], ps_at[i], nullptr);
      }
      if (cur_out) *cur_out = 0;
      if (st)
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
      return;
    }
  }
  int cur = 0;
  int64_t nnz = nnz0;
  int kidx = 0;
  int k = all_k ? 3 : ks[0];
  // k list mode: the list (ascending) and its results; on_k may extend it
  std::vector<int32_t> kv(ks, ks + (all_k ? 0 : nk));
  std::vector<int64_t> nz_v(kv.size(), 0);
  std::vector<int32_t> ps_v(kv.size(), 0);
  bool valid = false;
  int64_t passes = 0, steps = 0, wedges_host = 0;
  auto record = [&](int64_t value) {
    nz_v[kidx] = value;
    ps_v[kidx] = (int32_t)passes;
    if (!on_k) return;
    std::vector<int32_t> more;
    (*on_k)(kv[kidx], value, (int32_t)passes, &more);
    for (int32_t x : more) {
      if (x <= kv[kidx]) continue;  // only larger k can continue from this truss
      auto it = std::lower_bound(kv.begin() + kidx + 1, kv.end(), x);
      if (it != kv.end() && *it == x) continue;
      const auto at = it - kv.begin();
      kv.insert(it, x);
      nz_v.insert(nz_v.begin() + at, 0);
      ps_v.insert(ps_v.begin() + at, 0);
    }
  };
  int32_t kmax = 2;
  // the host waits for the survivor count only: the scatter's last tile
  // stores it with a sequence number to pinned host memory (words 8 and 9),
  // read without a copy or an event; the compaction's last kernel is queued
  // before the host reads it, so the device never idles while the host
  // decides (when nothing is removed the compaction writes the spare
  // buffers, which the next real one rewrites)
  long long* hk_dev = nullptr;
  long long* hk = progress_words(&hk_dev) + 8;
  hk_dev += 8;
  long long kept_seq = 0;
  const bool selective = selective_on(ctx(), nnz0);
  bool sup_zeroed = false;  // tw.sup[cur] was cleared by the compaction that made cur
  {
    while (true) {
      if (all_k && nnz == 0) break;
      const DCsr C{n, n, nnz, tw.cur[cur].rp.as<int32_t>(), tw.cur[cur].col.as<int32_t>(), nullptr};
      int32_t* sup = tw.sup[cur].as<int32_t>();
      if (!valid) {
        // first pass: every support; later passes: only the entries with an
        // endpoint that lost entries in the last compaction (the rest carry over)
        if (passes == 0 || !selective)
          support_symmetric_dev(C, tw.mirror[cur].as<int32_t>(), sup, tw.mw, s, sup_zeroed);
        else
          support_selected_dev(C, tw.mirror[cur].as<int32_t>(), tw.dirty.as<uint8_t>(), sup, tw.sw, tw.mw, s);
        ++passes;
        valid = true;
      }
      const int thr = k - 2;
      // compaction at threshold thr, queued before the survivor count is known
      int32_t* nrp = tw.cur[cur ^ 1].rp.as<int32_t>();
      int32_t* kept_dev = tw.kept.as<int32_t>();
      long long* bsum = tw.scan_tmp.as<long long>();
      const int64_t nbk = ceil_div(nnz, KEEP_TILE);
      if (nnz == 0) {  // nothing to filter (a k-truss that became empty)
        ++steps;
        if (!all_k) {
          record(0);
          if (++kidx == (int)kv.size()) break;
          k = kv[kidx];
          continue;
        }
        break;
      }
      hipLaunchKernelGGL(k_keep_tile_sums, dim3((unsigned)nbk), dim3(KEEP_T), 0, s, (const int32_t*)sup, nnz, thr,
                         bsum, (const TrussDevState*)nullptr, (const int32_t*)nullptr, (int32_t*)nullptr);
      const bool offsets = nbk > KEEP_FUSED_TILES;
      if (offsets) scan_tile_offsets(bsum, nbk, kept_dev, s);
      kept_seq = next_host_seq();
      hipLaunchKernelGGL(k_keep_scatter, dim3((unsigned)nbk), dim3(KEEP_T), 0, s, (const int32_t*)C.col,
                         (const int32_t*)tw.eid[cur].as<int32_t>(), (const int32_t*)sup, nnz, thr,
                         (const long long*)bsum, offsets ? 1 : 0, tw.cur[cur ^ 1].col.as<int32_t>(),
                         tw.eid[cur ^ 1].as<int32_t>(), tw.nold.as<int32_t>(), tw.newpos.as<int32_t>(),
                         tw.sup[cur ^ 1].as<int32_t>(), selective ? 1 : 0, kept_dev, (const TrussDevState*)nullptr,
                         hk_dev, kept_seq);
      hipLaunchKernelGGL(k_keep_finish, dim3(grid_for(std::max<int64_t>(n + 1, nnz), 256, 8192)), dim3(256), 0, s,
                         C.rp, n, (const int32_t*)tw.newpos.as<int32_t>(), nnz, (const int32_t*)kept_dev, nrp,
                         selective ? tw.dirty.as<uint8_t>() : nullptr, (const int32_t*)tw.mirror[cur].as<int32_t>(),
                         (const int32_t*)tw.nold.as<int32_t>(), tw.mirror[cur ^ 1].as<int32_t>(),
                         (const TrussDevState*)nullptr);
      GA_HIP(hipGetLastError());
      wait_host_word(&hk[1], kept_seq, s);
      const int64_t kept = __atomic_load_n(&hk[0], __ATOMIC_RELAXED);
      ++steps;
      if (kept == nnz) {
        if (!all_k) {
          record(nnz);
          if (++kidx == (int)kv.size()) break;
          k = kv[kidx];
          continue;
        }
        hipLaunchKernelGGL(k_truss_assign, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s,
                           (const int32_t*)tw.eid[cur].as<int32_t>(), nnz, k, tw.truss.as<int32_t>());
        GA_HIP(hipGetLastError());
        kmax = k;
        ++k;
        continue;
      }
      cur ^= 1;
      nnz = kept;
      valid = false;
      sup_zeroed = !selective;
    }
  }
  GA_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < (all_k ? 0 : nk); ++i) {  // the caller's k values (joined ones went to on_k)
    const auto at = std::lower_bound(kv.begin(), kv.end(), ks[i]) - kv.begin();
    if (nnz_at) nnz_at[i] = nz_v[at];
    if (passes_at) passes_at[i] = ps_v[at];
  }
  if (cur_out) *cur_out = cur;
  if (nnz_out) *nnz_out = nnz;
  if (kmax_out) *kmax_out = kmax;
  if (st) {
    st->passes = passes;
    st->steps = steps;
    st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    (void)wedges_host;
  }
}

void ktruss_dev(const DCsr& A, int k, TrussWork& tw, hipStream_t s, int64_t* nnz_out, int32_t* passes_out) {
  GA_REQUIRE(k >= 3, GA_ERR_ARG, "k must be >= 3");
  int cur = 0;
  int64_t nnz = 0;
  ga_truss_stats st{};
  const int32_t kk = k;
  truss_loop(A, false, &kk, 1, tw, s, &cur, &nnz, nullptr, &st);
  if (nnz_out) *nnz_out = nnz;
  if (passes_out) *passes_out = (int32_t)st.passes;
}

void ktruss_multi_dev(const DCsr& A, const int32_t* ks, int nk, TrussWork& tw, hipStream_t s, int64_t* nnz_at,
                      int32_t* passes_at, const KDone* on_k) {
  GA_REQUIRE(nk >= 1, GA_ERR_ARG, "no k given");
  for (int i = 0; i < nk; ++i) {
    GA_REQUIRE(ks[i] >= 3, GA_ERR_ARG, "k must be >= 3");
    GA_REQUIRE(i == 0 || ks[i] > ks[i - 1], GA_ERR_ARG, "k list must be ascending and distinct");
  }
  ga_truss_stats st{};
  truss_loop(A, false, ks, nk, tw, s, nullptr, nullptr, nullptr, &st, nnz_at, passes_at, on_k);
}

// Every buffer truss_loop sizes for one k (no per-entry truss array), at the
// sizes of the first pass, which are the largest.
size_t ktruss_workspace_bytes(int64_t n, int64_t nnz) {
  const size_t e = lease_bytes(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
  size_t b = 2 * (lease_bytes(sizeof(int32_t) * (n + 1)) + 4 * e);  // cur rp/col, eid, mirror, sup
  b += 2 * e;                                                       // nold, newpos
  b += lease_bytes(std::max<int64_t>(n, 1));                                                  // dirty
  b += lease_bytes(64);                                                                         // kept
  b += lease_bytes(sizeof(TrussDevState));                                                      // dstate
  b += lease_bytes(sizeof(long long) * std::max<int64_t>(ceil_div(std::max<int64_t>(nnz, 1), KEEP_TILE), 1));
  b += mxm_workspace_bytes(n, nnz, true);
  b += sel_workspace_bytes(n, nnz);
  b += lease_bytes(sizeof(PeelItem) * 4 * peel_lists_geometry(nnz).stride) + lease_bytes(sizeof(PeelCtl));  // peeling lists
  b += e;                                                                                   // stamps
  return b;
}

}  // namespace ga

using namespace ga;

struct ga_graph {
  int device = -1;
  bool symmetric = false;
  OwnedCsr A;
  OwnedCsr L;
  bool have_L = false;
  bool sums_clean = false;  // the sum and the arrival count are zero (the last count left them so)
  TrussWork tw;
  MxmWork mw;
  BcWork bw;
  DevBuf sums, scan_tmp;
};

extern "C" {

ga_status ga_triangle_count(const ga_csr* A, int64_t* triangles) {
  return guarded([&] {
    GA_REQUIRE(triangles != nullptr, GA_ERR_ARG, "null output");
    validate_host_csr(A, "A");
    GA_REQUIRE(A->nrows == A->ncols, GA_ERR_DIM, "triangle counting needs a square matrix");
    hipStream_t s = ctx().stream;
    OwnedCsr dA, L;
    upload_csr(A, GA_INT32, dA, s);
    DCsr vA = dA.view();
    vA.val = nullptr;
    DevBuf tmp, sums;
    lower_triangle_dev(vA, L, tmp, s);
    sums.ensure(64);
    fill_u64(sums.p, 0ull, 1, s);
    MxmWork w;
    w.max_mask_row = w.max_b_row = host_max_row(A);
    const DCsr vL = L.view();
    mxm_masked_dev(vL, vL, vL, GA_INT64, 1, nullptr, sums.as<unsigned long long>(), nullptr, w, s);
    auto* t = static_cast<unsigned long long*>(pinned_scratch());
    GA_HIP(hipMemcpyAsync(t, sums.p, sizeof(*t), hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    *triangles = (int64_t)*t;
  });
}

ga_status ga_triangle_count_lu(const ga_csr* A, int64_t* triangles) {
  return guarded([&] {
    GA_REQUIRE(triangles != nullptr, GA_ERR_ARG, "null output");
    validate_host_csr(A, "A");
    GA_REQUIRE(A->nrows == A->ncols, GA_ERR_DIM, "triangle counting needs a square matrix");
    hipStream_t s = ctx().stream;
    OwnedCsr dA, L, U;
    upload_csr(A, GA_INT32, dA, s);
    DCsr vA = dA.view();
    vA.val = nullptr;
    DevBuf tmp, sums;
    lower_triangle_dev(vA, L, tmp, s);
    // U = L^T; for the symmetric input of the reference's TC this is the
    // strict upper triangle of A
    upper_triangle_dev(vA, U, tmp, s);
    sums.ensure(64);
    fill_u64(sums.p, 0ull, 1, s);
    MxmWork w;
    w.max_mask_row = w.max_b_row = host_max_row(A);
    mxm_masked_dev(L.view(), U.view(), vA, GA_INT64, 1, nullptr, sums.as<unsigned long long>(), nullptr, w, s);
    auto* t = static_cast<unsigned long long*>(pinned_scratch());
    GA_HIP(hipMemcpyAsync(t, sums.p, sizeof(*t), hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    *triangles = (int64_t)(*t >> 1);
  });
}

ga_status ga_triangle_count_vertex(const ga_csr* A, int64_t* per_vertex) {
  return guarded([&] {
    GA_REQUIRE(per_vertex != nullptr, GA_ERR_ARG, "null output");
    validate_host_csr(A, "A");
    GA_REQUIRE(A->nrows == A->ncols, GA_ERR_DIM, "triangle counting needs a square matrix");
    hipStream_t s = ctx().stream;
    const int64_t n = A->nrows;
    OwnedCsr dA;
    upload_csr(A, GA_INT32, dA, s);
    DCsr vA = dA.view();
    vA.val = nullptr;
    DevBuf sums, groups;
    sums.ensure(8 * std::max<int64_t>(n, 1));
    groups.ensure(4 * std::max<int64_t>(n, 1));
    fill_u64(sums.p, 0ull, n, s);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, groups.as<int32_t>(), n);
    MxmWork w;
    w.max_mask_row = w.max_b_row = host_max_row(A);
    mxm_masked_dev(vA, vA, vA, GA_INT64, 1, nullptr, sums.as<unsigned long long>(),
                   (const int32_t*)groups.as<int32_t>(), w, s);
    if (n) GA_HIP(hipMemcpyAsync(per_vertex, sums.p, 8 * n, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) per_vertex[i] >>= 1;
  });
}

ga_status ga_ktruss(const ga_csr* A, int k, int32_t* out_rowptr, int32_t* out_col, int64_t* out_nnz,
                    int32_t* iterations) {
  return guarded([&] {
    validate_host_csr(A, "A");
    GA_REQUIRE(host_csr_symmetric(A), GA_ERR_ARG, "k-truss needs a symmetric pattern");
    GA_REQUIRE(k >= 3, GA_ERR_ARG, "k must be >= 3");
    GA_REQUIRE(out_rowptr && out_nnz && (A->nnz == 0 || out_col), GA_ERR_ARG, "null output");
    hipStream_t s = ctx().stream;
    OwnedCsr dA;
    upload_csr(A, GA_INT32, dA, s);
    DCsr vA = dA.view();
    vA.val = nullptr;
    TrussWork tw;
    tw.mw.max_mask_row = tw.mw.max_b_row = host_max_row(A);
    tw.no_graph = true;  // one run per workspace: no capture to amortise
    int cur = 0;
    int64_t nnz = 0;
    ga_truss_stats st{};
    const int32_t kk = k;
    truss_loop(vA, false, &kk, 1, tw, s, &cur, &nnz, nullptr, &st);
    GA_HIP(hipMemcpyAsync(out_rowptr, tw.cur[cur].rp.p, sizeof(int32_t) * (A->nrows + 1), hipMemcpyDeviceToHost, s));
    if (nnz) GA_HIP(hipMemcpyAsync(out_col, tw.cur[cur].col.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    *out_nnz = nnz;
    if (iterations) *iterations = (int32_t)st.passes;
  });
}

ga_status ga_truss_decomposition(const ga_csr* A, int32_t* truss, int32_t* kmax, ga_truss_stats* stats) {
  return guarded([&] {
    validate_host_csr(A, "A");
    GA_REQUIRE(host_csr_symmetric(A), GA_ERR_ARG, "truss decomposition needs a symmetric pattern");
    GA_REQUIRE(A->nnz == 0 || truss != nullptr, GA_ERR_ARG, "null output");
    hipStream_t s = ctx().stream;
    OwnedCsr dA;
    upload_csr(A, GA_INT32, dA, s);
    DCsr vA = dA.view();
    vA.val = nullptr;
    TrussWork tw;
    tw.mw.max_mask_row = tw.mw.max_b_row = host_max_row(A);
    int32_t km = 2;
    ga_truss_stats st{};
    truss_loop(vA, true, nullptr, 0, tw, s, nullptr, nullptr, &km, &st);
    if (A->nnz) GA_HIP(hipMemcpyAsync(truss, tw.truss.p, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    if (kmax) *kmax = km;
    if (stats) *stats = st;
  });
}

ga_status ga_graph_create(const ga_csr* A, ga_graph** out) {
  return guarded([&] {
    GA_REQUIRE(out != nullptr, GA_ERR_ARG, "null output");
    validate_host_csr(A, "A");
    GA_REQUIRE(A->nrows == A->ncols, GA_ERR_DIM, "graphs are square");
    Context& c = ctx();
    auto* g = new ga_graph();
    try {
      g->device = c.device;
      g->symmetric = host_csr_symmetric(A);
      g->tw.mw.max_mask_row = g->tw.mw.max_b_row = host_max_row(A);
      g->mw.max_mask_row = g->mw.max_b_row = g->tw.mw.max_mask_row;
      g->mw.reuse_plan = true;  // the triangle-count product of a resident graph never changes
      upload_csr(A, GA_INT32, g->A, c.stream);
      g->A.val.release();
      GA_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

ga_status ga_graph_destroy(ga_graph* g) {
  return guarded([&] {
    if (!g) return;
    GA_HIP(hipStreamSynchronize(ctx().stream));
    delete g;
  });
}

ga_status ga_graph_triangle_count(ga_graph* g, int64_t* triangles) {
  return guarded([&] {
    GA_REQUIRE(g && triangles, GA_ERR_ARG, "null argument");
    hipStream_t s = ctx().stream;
    if (!g->have_L) {
      lower_triangle_dev(g->A.view(), g->L, g->scan_tmp, s);
      g->have_L = true;
    }
    g->sums.ensure(sizeof(unsigned long long) * (SUM_PUB_TICKET + 1));
    if (!g->sums_clean) fill_u64(g->sums.p, 0ull, SUM_PUB_TICKET + 1, s);
    g->sums_clean = false;
    const DCsr vL = g->L.view();
    // the count reaches the host through pinned memory: written by the count
    // launch itself when its kept plan is one statically split kernel, else
    // by a kernel that also clears the sum and the claim counters
    long long* hdev = nullptr;
    long long* h = progress_words(&hdev) + 4;  // words 4 (value) and 5 (sequence)
    const long long seq = next_host_seq();
    g->mw.pub_host = hdev + 4;
    g->mw.pub_seq = seq;
    g->mw.pub_done = false;
    struct PubOff {
      MxmWork& w;
      ~PubOff() { w.pub_host = nullptr; }
    } pub_off{g->mw};
    mxm_masked_dev(vL, vL, vL, GA_INT64, 1, nullptr, g->sums.as<unsigned long long>(), nullptr, g->mw, s);
    if (g->mw.pub_done) {
      wait_host_word(&h[1], seq, s);
      *triangles = (int64_t)__atomic_load_n(&h[0], __ATOMIC_RELAXED);
    } else {
      *triangles = (int64_t)publish_sum_and_wait(g->mw, g->sums.as<unsigned long long>(), s);
    }
    g->sums_clean = true;
  });
}

ga_status ga_graph_truss(ga_graph* g, int32_t* truss, int32_t* kmax, ga_truss_stats* stats) {
  return guarded([&] {
    GA_REQUIRE(g != nullptr, GA_ERR_ARG, "null graph");
    GA_REQUIRE(g->symmetric, GA_ERR_ARG, "truss decomposition needs a symmetric pattern");
    hipStream_t s = ctx().stream;
    int32_t km = 2;
    ga_truss_stats st{};
    truss_loop(g->A.view(), true, nullptr, 0, g->tw, s, nullptr, nullptr, &km, &st);
    if (truss && g->A.nnz)
      GA_HIP(hipMemcpyAsync(truss, g->tw.truss.p, sizeof(int32_t) * g->A.nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    if (kmax) *kmax = km;
    if (stats) *stats = st;
  });
}

ga_status ga_graph_ktruss(ga_graph* g, int k, int64_t* nnz, int32_t* passes) {
  return guarded([&] {
    GA_REQUIRE(g && nnz, GA_ERR_ARG, "null argument");
    GA_REQUIRE(g->symmetric, GA_ERR_ARG, "k-truss needs a symmetric pattern");
    ktruss_dev(g->A.view(), k, g->tw, ctx().stream, nnz, passes);
  });
}

ga_status ga_graph_bc(ga_graph* g, const int32_t* sources, int64_t nsources, double* bc) {
  return guarded([&] {
    GA_REQUIRE(g && sources && bc, GA_ERR_ARG, "null argument");
    bc_batch_dev(g->A.view(), sources, nsources, bc, g->bw, ctx().stream);
  });
}

}  // extern "C"

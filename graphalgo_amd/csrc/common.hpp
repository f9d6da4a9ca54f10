// Internal runtime of the gfx950 library: errors, device context, device
// buffers and the device-side primitives (wave scans, block scans) shared by
// the kernels.  Wavefronts are 64 lanes wide on CDNA4; every primitive here
// is written for that width.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "graphalgo.h"

namespace ga {

struct Error {
  int code;
  std::string msg;
};

void set_error(int code, const std::string& msg);

#define GA_HIP(x)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) {                                                                     \
      (void)hipGetLastError();                                                                  \
      throw ::ga::Error{GA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)};            \
    }                                                                                           \
  } while (0)

#define GA_REQUIRE(cond, code, msg)                \
  do {                                             \
    if (!(cond)) throw ::ga::Error{(code), (msg)}; \
  } while (0)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return GA_OK;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(GA_ERR_OOM, "host allocation failed");
    return GA_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(GA_ERR_HIP, e.what());
    return GA_ERR_HIP;
  }
}

// Grow-only device allocation owned by a workspace or handle.
// A query's device workspace inside a server arena: one contiguous block,
// handed out front to back.  While a worker thread runs a query it points
// t_lease at the block, and every DevBuf sized on that thread takes its
// memory from it instead of hipMalloc (no allocation, and no implicit device
// synchronisation of hipFree, while other queries run).  The block size is
// the query's footprint; running past it is an error, never a silent
// allocation.
struct Lease {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t b);
};
extern thread_local Lease* t_lease;
// bytes a buffer of b bytes occupies in a lease (256-byte aligned blocks)
inline size_t lease_bytes(size_t b) { return (b + 255) & ~size_t(255); }

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool own = false;  // allocated with hipMalloc (else a lease block or empty)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), own(o.own) {
    o.p = nullptr; o.bytes = 0; o.own = false;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; own = o.own;
      o.p = nullptr; o.bytes = 0; o.own = false;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    own = false;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (b == 0) return;
    if (t_lease) {
      p = t_lease->take(b);
      bytes = b;
      return;
    }
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error{GA_ERR_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed"};
    }
    bytes = b;
    own = true;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// 256 bytes of pinned host memory per thread, for the small device-to-host
// reads that steer the host loops (allocated on first use, kept for the
// thread's lifetime: no per-call hipHostMalloc / hipHostFree).
void* pinned_scratch();
// Per host thread: 32 words of coherent pinned host memory that kernels write
// with system-scope stores and the host reads without a HIP call (progress
// of device-driven loops); *dev = the device's address of the same words.
long long* progress_words(long long** dev);

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  int64_t hbm_bytes = 0;
  int64_t lds_per_block = 65536;  // the device's LDS limit per workgroup
  std::string name;
  // row binning default: rows with at most this many products (and a mask
  // of at most 128 entries) go to the wavefront-per-row kernel; GA_WAVE_FLOPS
  // overrides it at ga_init
  long long wave_flops = 65536;
  // item list construction when the wavefront bin is off: false = scanned
  // list of the non-empty rows' items; true = rows are their own first item
  // and only extra chunks are listed (GA_DIRECT_ITEMS=1), which pays a claim
  // for every empty row and loses on peeled graphs
  bool direct_items = false;
  // Selective support passes: k-truss passes after the first recompute only
  // the supports of entries with an endpoint that lost entries (1 = on,
  // 0 = off, -1 = on for graphs of at least selective_min_nnz stored entries
  // when triangle-listing passes are switched off, since those beat it;
  // GA_SELECTIVE_SUPPORT overrides it at ga_init).  On R-MAT 16 it
  // forms 19% fewer products but the pass time is set by per-item costs,
  // not products, and the selection launches cost more than the products
  // save (510 vs 493 ms per decomposition); on the R-MAT 18-20 graphs of
  // the served pool one k-truss with k = 8 or 16 runs 6-27% faster (R-MAT 18,
  // k = 8: 337 -> 247 ms), with k = 3 (two passes) 2-4% slower.
  int selective_support = -1;
  // support passes as triangle listing: C<L> = L*L over the strict lower
  // half, each hit adding 1 to its three edges (1 = on, 0 = off, -1 = on for
  // every pass whose graph holds at least tri_list_min_nnz stored entries;
  // GA_SUPPORT_TRI).  R-MAT 16 decomposition with the passes above 2^20
  // entries as triangle listing: 168.4 vs 170.8 ms (above 2^19: 200.5 ms).
  // Measured on MI355X: one k-truss on the served pool's R-MAT 18 / 20 runs
  // 2.1-2.2x faster than with the symmetric half pass plus selective passes
  // (k = 8: 224 -> 108 ms, 318 -> 147 ms; one full pass 31 -> 10 ms on
  // R-MAT 18), but the R-MAT 16 decomposition runs 40% slower (227 vs
  // 164 ms): its passes at large k are on dense cores where most wedges close
  // a triangle, and every closed one costs a global atomic for the B entry.
  int support_tri = -1;
  int64_t tri_list_min_nnz = int64_t(1) << 20;
  // served k-truss queries joining a peeling run on their graph that is
  // already going (GA_SERVE_JOIN): 0 = never, 1 = when the query's k lies
  // within the run's k range (it adds no work), 2 = also when it is above it
  // (the run goes on to it)
  int serve_join = 2;
  int64_t selective_min_nnz = int64_t(1) << 22;
  // k-truss peeling loop: 0 = host-driven (a host wait for the survivor count
  // after every filter step), 1 = device-driven steps launched one by one,
  // 2 = device-driven steps replayed from a captured HIP graph, -1 = 1 for
  // graphs of at most truss_device_max_nnz stored entries, else 0
  // (GA_TRUSS_LOOP overrides it at ga_init).  Measured on MI355X: the device
  // steps halve a small graph's k-truss time (R-MAT 12, k = 8: 1.49 -> 0.60
  // ms) and lose 2% on R-MAT 16 (the survivors are copied back into the
  // buffers the fixed launches name); graph replay is slower than launching
  // the same kernels (R-MAT 16 decomposition 170.8 vs 167.6 ms).
  int truss_loop = -1;
  int64_t truss_device_max_nnz = int64_t(1) << 18;
  // BC forward: switch to bottom-up steps once the level list has more than
  // 1/bc_alpha of the edges of the vertices with unvisited lanes
  // (GA_BC_BOTTOM_UP=0: top-down only; GA_BC_ALPHA sets the ratio).  R-MAT 18
  // x 256 sources: 14.9 ms per pass at alpha 1, 12.7 at 2, 12.3 at 4 and 14.
  bool bc_bottom_up = true;
  double bc_alpha = 14.0;
  // BC level loop (GA_BC_LOOP): 0 = host-steered (the host reads every
  // level list's size), 1 = device-driven (level lists sized on the device,
  // one host wait at the end of the forward sweep), 2 = the device-driven
  // steps replayed from captured HIP graphs
  int bc_loop = 1;
  // target wedges per work item when rows are split by their measured work
  // (GA_ITEM_WEDGES overrides it at ga_init)
  long long item_wedges = 32768;
  // rows per claim of the short-row kernels (0 = automatic; GA_ROW_CHUNK
  // overrides it at ga_init, MxmWork::row_chunk per workspace)
  int row_chunk = 0;
  // short rows of pattern triangle counts: 0 / 2 = rows packed per
  // wavefront with rank-located wedges, 512 / 256 wedges per trip
  // (k_sum_wave), 1 = 16-lane wedge runs (k_sum_subrows), 8/16/32/64 =
  // lanes per row walking its entries (k_sum_entries); GA_SUM_SUB overrides
  // it at ga_init
  int sum_sub = 0;
  // kernel timing (events around the masked-product launches) and
  // algorithmic byte accounting (extra counting kernels; kept out of timed runs)
  bool prof_events_on = false;
  bool prof_bytes = false;
  int64_t prof_bytes_launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
  std::vector<hipEvent_t> event_pool;
  int64_t prof_launches = 0;
  int64_t prof_bytes_host = 0;  // host-known part of the algorithmic bytes
  DevBuf prof_dev;  // u64: [0] wedges processed, [1] algorithmic wedges (4-byte gathers),
                    // [2] algorithmic bytes reported by BC, [3] BC wedges
};

Context& ctx();  // throws GA_ERR_STATE when ga_init has not run

// Device view of a CSR matrix.
struct DCsr {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  const int32_t* rp = nullptr;
  const int32_t* col = nullptr;
  const void* val = nullptr;
};

// An owned device CSR (pattern or valued).
struct OwnedCsr {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  DevBuf rp, col, val;
  DCsr view() const { return DCsr{nrows, ncols, nnz, rp.as<int32_t>(), col.as<int32_t>(), val.p}; }
};

size_t dtype_size(int dtype);
// resolves the device's copy of the allocation-free entry point's scratch
// array (noalloc.hip); called by ga_init with the device current
void noalloc_setup(int device);
void validate_host_csr(const ga_csr* A, const char* name);
// Longest row of a host CSR.
inline int64_t host_max_row(const ga_csr* A) {
  int64_t mx = 0;
  for (int64_t i = 0; i < A->nrows; ++i) mx = std::max<int64_t>(mx, A->rowptr[i + 1] - A->rowptr[i]);
  return mx;
}
// True when the (validated) pattern equals its transpose.
bool host_csr_symmetric(const ga_csr* A);
void upload_csr(const ga_csr* A, int dtype, OwnedCsr& out, hipStream_t s);

// ---- device primitives --------------------------------------------------
template <typename T>
__device__ inline T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

template <typename T>
__device__ inline T wave_reduce_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ inline void atomic_add_val(int32_t* p, int32_t v) { atomicAdd(p, v); }
__device__ inline void atomic_add_val(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}
__device__ inline void atomic_add_val(double* p, double v) { atomicAdd(p, v); }
__device__ inline void atomic_add_val(float* p, float v) { atomicAdd(p, v); }

// ---- host launchers shared between translation units -------------------
// Exclusive scan of n int32 counts into out[0..n]; out[n] = total.  Totals
// must fit in int32 (they index CSR arrays).  tmp is grown as needed.
void scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, DevBuf& tmp, hipStream_t s);
// In place: bsum[0..nb) (per-tile sums) becomes exclusive tile offsets; the
// total goes to *total (one workgroup).
void scan_tile_offsets(long long* bsum, int64_t nb, int32_t* total, hipStream_t s);
// Fill n 32-bit words with a value (graph-capture safe: a kernel, not a memset).
void fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s);
void fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s);

// Workspace of the masked SpGEMM (row items and counters).
struct MxmWork {
  DevBuf cnt, start, item_row, item_t0, scan_tmp, counter, mend;
  DevBuf flag, rstart, rows;  // wavefront-per-row bin
  long long wave_flops = -1;     // binning threshold (products per row) for the wavefront kernel;
                                 // -1 = the context default
  int threads = 256;             // workgroup width of the item kernel (64, 128 or 256)
  int row_chunk = 0;             // rows per claim of the short-row kernels (0 = automatic)
  int sum_sub = -1;              // short pattern sum rows: as Context::sum_sub
                                 // walking entries (-1 = the context default)
  int64_t max_mask_row = INT32_MAX;  // longest mask row (host-known bound); < 65536 enables the rank lookup
  int64_t max_b_row = INT32_MAX;     // longest B row (host-known bound; B.nnz is used when tighter)
  // Work balance over repeated products of one matrix (the k-truss support
  // passes): each launch records the wedges of every row, and the next
  // launch splits rows into items of about item_wedges wedges (at most
  // MXM_SPLIT_MAX times the entry-based item count).
  bool balance = false;
  DevBuf rwork[2];
  int rwork_cur = 0;
  bool rwork_have = false;
  // Device-driven launches (the k-truss steps replayed from a HIP graph):
  // the kernel arguments may not change between launches, so the per-row
  // work lives in one buffer (each row's previous count is read before it is
  // cleared; the caller zeroes rwork[0] first), every row goes to the listed
  // work items (no wavefront bin), and a nonzero *skip makes the launch a no-op.
  const int32_t* skip = nullptr;
  bool fixed_args = false;
  // the claim counters are known to be zero (a publish kernel cleared them
  // after the previous launch on this workspace), so the next launch skips
  // its clearing fill
  bool counters_clean = false;
  // support passes: the smallest graph (stored entries) that the default
  // triangle-listing choice takes (-1: the context's tri_list_min_nnz)
  int64_t tri_min_nnz = -1;
  // reuse_plan: the operands do not change between launches (a resident
  // batch), so the row bins and item lists of the first launch are kept
  bool reuse_plan = false;
  bool plan_valid = false;
  const int32_t* plan_rp = nullptr;
  int64_t plan_m = 0, plan_nnz = 0;
  int plan_tcap = 0;
  long long plan_wf = 0;
  // short rows and work items of the kept plan, read back once when it is
  // built (-1 = unknown): an empty kernel of a kept plan is not launched
  int64_t plan_rows = -1, plan_items = -1;
  // A one-target sum may publish its total itself: with pub_host set, a
  // launch whose kept plan is one statically split short-row kernel has its
  // last workgroup write the total to pub_host[0] and pub_seq to pub_host[1]
  // (the sum and an arrival count in the sums buffer: words 0 and
  // SUM_PUB_TICKET, zero before and after), and sets pub_done; otherwise the
  // caller publishes with a kernel of its own.
  long long* pub_host = nullptr;
  long long pub_seq = 0;
  bool pub_done = false;
};
constexpr int SUM_PUB_TICKET = 16;  // the arrival count's word: a 128-byte line after the sum

// C<M> = A*B (structural mask).  mode 0: per-entry values into Cval (dtype);
// mode 1: pattern inputs, int64 sum of all C entries accumulated into
// sums[row_group ? row_group[i] : 0] (caller zeroes sums).
void mxm_masked_dev(const DCsr& A, const DCsr& B, const DCsr& M, int dtype, int mode, void* Cval,
                    unsigned long long* sums, const int32_t* row_group, MxmWork& w, hipStream_t s);

// After a pattern sum launch (mode 1) on workspace w: one kernel hands
// sums[0] to the host through pinned host memory and clears sums[0] and the
// claim counters for the next launch; the calling thread waits for the value
// without a stream synchronisation.  Returns sums[0].
unsigned long long publish_sum_and_wait(MxmWork& w, unsigned long long* sums, hipStream_t s);
// Host-memory handoff: a sequence number unique on the calling thread, and a
// wait until a kernel has stored it to *w (system scope) -- a stream error or
// a drained stream without the store ends the wait with an error.
long long next_host_seq();
void wait_host_word(const long long* w, long long seq, hipStream_t s);

// S<C> = C*C (int32 per entry) for a symmetric loop-free pattern C, computed
// on the strict lower half and mirrored (the k-truss support pass).
// mirror[t] = position of entry t's transposed entry; kept valid through
// compaction by the k-truss loop.
void support_symmetric_dev(const DCsr& C, const int32_t* mirror, int32_t* S, MxmWork& w, hipStream_t s,
                           bool s_zeroed = false);
void mirror_index_dev(const DCsr& C, int32_t* mirror, hipStream_t s);
// Support pass over the affected entries only (entries with an endpoint in a
// row flagged dirty); the other entries of S keep their values.
struct SelWork {
  DevBuf cnt, mrp, mcol, opos, scan_tmp;
};
void support_selected_dev(const DCsr& C, const int32_t* mirror, const uint8_t* dirty, int32_t* S, SelWork& sw,
                          MxmWork& w, hipStream_t s);
size_t sel_workspace_bytes(int64_t n, int64_t nnz);

// Batched sum-mode product over many graphs stored back to back (graph-local
// column indices): sums[g] += sum of C<L_g> = L_g * L_g.  threads = 64, 128
// or 256 (workgroup size = A entries per work item).
void mxm_batched_sum_dev(const DCsr& L, const int32_t* row_group, const int32_t* gbase, int64_t max_cols,
                         unsigned long long* sums, MxmWork& w, hipStream_t s);

// Device bytes of the masked-product workspace for an m-row A with nnz_a
// entries (every buffer launch_binned sizes, plus the symmetric mask prefix
// when with_mend), as taken from a lease.
size_t mxm_workspace_bytes(int64_t m, int64_t nnz_a, bool with_mend);
size_t scan_tmp_bytes(int64_t n);

// k-truss workspace (algorithms.hip): ping-pong CSR, edge ids, mirror index.
// An instantiated HIP graph of one device-driven k-truss step; it stays
// valid while the buffers and sizes it was captured with (key) are unchanged.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
  StepGraph() = default;
  StepGraph(const StepGraph&) = delete;
  StepGraph& operator=(const StepGraph&) = delete;
  ~StepGraph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

struct TrussWork {
  OwnedCsr cur[2];
  DevBuf eid[2], mirror[2], sup[2];
  DevBuf cnt, scan_tmp, truss, kept, nold, newpos, dirty;
  DevBuf plist, pcnt, stamp;  // peeling by support updates: removal lists, their counters, round stamps
  DevBuf dstate;  // device-driven loop: the peeling state (k, entries, passes, ...)
  MxmWork mw;
  SelWork sw;
  StepGraph step_graph;
  // launch device-driven steps one by one even when the context replays a
  // captured graph (a workspace used for one run only: the capture would
  // cost more than it saves)
  bool no_graph = false;
};
// k-truss for one k of a symmetric loop-free pattern on the device; returns
// the surviving entry count (both triangles) and the support passes run.
void ktruss_dev(const DCsr& A, int k, TrussWork& tw, hipStream_t s, int64_t* nnz_out, int32_t* passes_out);
// Several k at once (ascending, distinct): one peeling run that stops at each
// k in turn; nnz_at[i] / passes_at[i] = entries of the ks[i]-truss and the
// support passes run up to it.
// on_k (optional): called with (k, entries, passes, more) as soon as the
// k-truss is known.  The host-driven loop calls it while the run can still go
// on: `more` is then non-null, and the callee may append larger k values,
// which the run takes on in ascending order (reported through on_k only).
// The device-driven steps call it at the end of the run with more = null.
using KDone = std::function<void(int32_t, int64_t, int32_t, std::vector<int32_t>*)>;
void ktruss_multi_dev(const DCsr& A, const int32_t* ks, int nk, TrussWork& tw, hipStream_t s, int64_t* nnz_at,
                      int32_t* passes_at, const KDone* on_k = nullptr);
size_t ktruss_workspace_bytes(int64_t n, int64_t nnz);

// Batched betweenness centrality workspace and driver (bc.hip).
// Instantiated HIP graphs of a BC pass's steps (captured once, replayed per
// level), with the launch arguments they were captured with.
struct BcGraphs {
  hipGraphExec_t fwd = nullptr, bwd = nullptr;
  std::vector<uintptr_t> key;
  BcGraphs() = default;
  BcGraphs(const BcGraphs&) = delete;
  BcGraphs& operator=(const BcGraphs&) = delete;
  BcGraphs(BcGraphs&& o) noexcept : fwd(o.fwd), bwd(o.bwd), key(std::move(o.key)) { o.fwd = o.bwd = nullptr; }
  BcGraphs& operator=(BcGraphs&& o) noexcept {
    if (this != &o) {
      reset();
      fwd = o.fwd;
      bwd = o.bwd;
      key = std::move(o.key);
      o.fwd = o.bwd = nullptr;
    }
    return *this;
  }
  ~BcGraphs() { reset(); }
  void reset() {
    if (fwd) (void)hipGraphExecDestroy(fwd);
    if (bwd) (void)hipGraphExecDestroy(bwd);
    fwd = bwd = nullptr;
    key.clear();
  }
};

struct BcWork {
  DevBuf depth, sigma, q, bc, flag, levels, cnt, off, scan_tmp, src;
  DevBuf icnt, hcnt, istart, hstart, it_v, it_t0, it_slot, partial;  // neighbour-chunked items
  DevBuf dsum;                                                        // edge counts of two lists
  // device-driven passes: candidate flags, list and items, list builder
  // counts, the control block
  DevBuf cflag, cit_v, cit_t0, cit_slot, chstart, gcnt, gsum, ctl, hpos, chpos;
  BcGraphs graphs;
};
void bc_batch_dev(const DCsr& A, const int32_t* src_h, int64_t ns, double* bc_h, BcWork& w, hipStream_t s);
size_t bc_workspace_bytes(int64_t n, int64_t nnz, int64_t ns);
// Independent source sets (each at most 64 sources) in one pass, one score
// vector per set, bit-identical to passes of the sets alone.
void bc_groups_dev(const DCsr& A, const std::vector<std::vector<int32_t>>& sets, double* const* bc_out, BcWork& w,
                   hipStream_t s);
size_t bc_groups_workspace_bytes(int64_t n, int64_t nnz, const std::vector<int64_t>& group_sizes);

// Strict lower triangle of a square pattern with sorted rows, on device.
void lower_triangle_dev(const DCsr& A, OwnedCsr& L, DevBuf& tmp, hipStream_t s);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = ceil_div(n, block);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace ga

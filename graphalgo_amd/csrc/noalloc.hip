// Triangle count that allocates nothing: for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).
//
// The path is ga_triangle_count's (C<L> = L*L summed, the masked-product
// kernels in sum mode), on operands the caller already holds in device
// memory.  Every work buffer of the path -- row bins, item lists, claim
// counters, scan tiles -- is taken from `scratch`, a file-scope device array
// of NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the process:
// the call points the thread's workspace lease at that array, so the
// buffers the path sizes come out of it by bump allocation, never from
// hipMalloc.  The total goes straight to the caller's device word (cleared
// first by a fill kernel).
//
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (two host threads, or two streams in flight at once) race; a call that
// follows another in stream order is safe.
#include "common.hpp"

#include <array>
#include <mutex>

namespace ga {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static_assert(NOALLOC_SCRATCH_BYTES == (size_t)GA_NOALLOC_SCRATCH_BYTES, "header and library disagree");

// The array itself.  In host code its name only stands for the symbol: the
// device address comes from hipGetSymbolAddress (one per GPU).
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::array<char*, kMaxDevices> g_scratch_ptr{};
}  // namespace

// Called by ga_init with `device` current: resolves that GPU's copy of the
// array once and keeps its address.
void noalloc_setup(int device) {
  GA_REQUIRE(device >= 0 && device < kMaxDevices, GA_ERR_LIMIT, "device id beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  GA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  GA_REQUIRE(p != nullptr, GA_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[device] = static_cast<char*>(p);
}

static char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  char* p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
  GA_REQUIRE(p != nullptr, GA_ERR_STATE, "the scratch array of this device was not set up (ga_init)");
  return p;
}

}  // namespace ga

using namespace ga;

extern "C" ga_status ga_triangle_count_noalloc(const ga_csr* L, int64_t max_row, int64_t* triangles,
                                               void* stream) {
  return guarded([&] {
    GA_REQUIRE(L && triangles, GA_ERR_ARG, "null argument");
    GA_REQUIRE(L->nrows >= 0 && L->nrows == L->ncols && L->nnz >= 0, GA_ERR_DIM, "L must be square");
    GA_REQUIRE(L->nrows < INT32_MAX && L->nnz < INT32_MAX, GA_ERR_LIMIT, "sizes must fit int32 CSR indexing");
    GA_REQUIRE(L->rowptr && (L->nnz == 0 || L->col), GA_ERR_ARG, "null rowptr or col");
    GA_REQUIRE(max_row >= 0 && max_row <= L->nrows, GA_ERR_ARG, "max_row must be in [0, nrows]");
    GA_REQUIRE((reinterpret_cast<uintptr_t>(triangles) & 7) == 0, GA_ERR_ARG, "triangles must be 8-byte aligned");
    Context& c = ctx();
    GA_REQUIRE(!c.prof_events_on && !c.prof_bytes, GA_ERR_STATE,
               "kernel timing hooks create events: disable them for the allocation-free entry point");
    const size_t need = mxm_workspace_bytes(L->nrows, L->nnz, false);
    GA_REQUIRE(need <= NOALLOC_SCRATCH_BYTES, GA_ERR_LIMIT,
               "the graph's work buffers (" + std::to_string(need) + " bytes) exceed the fixed scratch array of " +
                   std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
    Lease lease{scratch_of(c.device), NOALLOC_SCRATCH_BYTES, 0};
    t_lease = &lease;
    struct Unset {
      ~Unset() { t_lease = nullptr; }
    } unset;
    fill_u64(triangles, 0ull, 1, s);
    const DCsr vL{L->nrows, L->ncols, L->nnz, L->rowptr, L->col, nullptr};
    MxmWork w;
    w.max_mask_row = w.max_b_row = max_row;
    mxm_masked_dev(vL, vL, vL, GA_INT64, 1, nullptr, reinterpret_cast<unsigned long long*>(triangles), nullptr, w,
                   s);
  });
}

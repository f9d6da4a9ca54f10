// Library runtime: context, error reporting, uploads, scans, fills and the
// kernel-timing hooks.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace ga {

namespace {
thread_local std::string t_err;
std::mutex g_ctx_mu;
Context* g_ctx = nullptr;
}  // namespace

void set_error(int code, const std::string& msg) { t_err = "[" + std::to_string(code) + "] " + msg; }

thread_local Lease* t_lease = nullptr;

void* Lease::take(size_t b) {
  const size_t a = lease_bytes(b);
  if (used + a > cap)
    throw Error{GA_ERR_LIMIT, "workspace needs more than the query's footprint of " + std::to_string(cap) +
                                  " bytes (" + std::to_string(used + a) + " requested)"};
  void* p = base + used;
  used += a;
  return p;
}

namespace {
struct PinnedScratch {
  void* p = nullptr;
  void* dev = nullptr;  // the device's address of p (mapped allocations)
  ~PinnedScratch() {
    if (p) (void)hipHostFree(p);
  }
};
thread_local PinnedScratch t_pinned;
thread_local PinnedScratch t_progress;
}  // namespace

void* pinned_scratch() {
  if (!t_pinned.p) GA_HIP(hipHostMalloc(&t_pinned.p, 256, hipHostMallocDefault));
  return t_pinned.p;
}

long long* progress_words(long long** dev) {
  if (!t_progress.p) {
    GA_HIP(hipHostMalloc(&t_progress.p, 256, hipHostMallocMapped | hipHostMallocCoherent));
    GA_HIP(hipHostGetDevicePointer(&t_progress.dev, t_progress.p, 0));
  }
  *dev = static_cast<long long*>(t_progress.dev);
  return static_cast<long long*>(t_progress.p);
}

Context& ctx() {
  GA_REQUIRE(g_ctx != nullptr, GA_ERR_STATE, "ga_init has not been called");
  GA_HIP(hipSetDevice(g_ctx->device));
  return *g_ctx;
}

size_t dtype_size(int dtype) {
  switch (dtype) {
    case GA_INT32: return 4;
    case GA_INT64: return 8;
    case GA_FLOAT64: return 8;
    case GA_FLOAT32: return 4;
  }
  throw Error{GA_ERR_ARG, "unknown dtype " + std::to_string(dtype)};
}

void validate_host_csr(const ga_csr* A, const char* name) {
  std::string nm(name);
  GA_REQUIRE(A != nullptr, GA_ERR_ARG, nm + ": null matrix");
  GA_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, GA_ERR_ARG, nm + ": negative size");
  GA_REQUIRE(A->nrows < INT32_MAX && A->ncols < INT32_MAX && A->nnz < INT32_MAX, GA_ERR_LIMIT,
             nm + ": sizes must fit int32 CSR indexing");
  GA_REQUIRE(A->rowptr != nullptr, GA_ERR_ARG, nm + ": null rowptr");
  GA_REQUIRE(A->nnz == 0 || A->col != nullptr, GA_ERR_ARG, nm + ": null col");
  GA_REQUIRE(A->rowptr[0] == 0 && A->rowptr[A->nrows] == A->nnz, GA_ERR_ARG, nm + ": rowptr does not span nnz");
  for (int64_t i = 0; i < A->nrows; ++i) {
    int32_t b = A->rowptr[i], e = A->rowptr[i + 1];
    GA_REQUIRE(b <= e, GA_ERR_ARG, nm + ": rowptr not monotone at row " + std::to_string(i));
    for (int32_t t = b; t < e; ++t) {
      int32_t c = A->col[t];
      GA_REQUIRE(c >= 0 && c < A->ncols, GA_ERR_UNSORTED, nm + ": column out of range in row " + std::to_string(i));
      GA_REQUIRE(t == b || A->col[t - 1] < c, GA_ERR_UNSORTED,
                 nm + ": row " + std::to_string(i) + " not sorted / has duplicates");
    }
  }
}

bool host_csr_symmetric(const ga_csr* A) {
  if (A->nrows != A->ncols) return false;
  for (int64_t i = 0; i < A->nrows; ++i)
    for (int32_t t = A->rowptr[i]; t < A->rowptr[i + 1]; ++t) {
      const int32_t j = A->col[t];
      const int32_t* b = A->col + A->rowptr[j];
      const int32_t* e = A->col + A->rowptr[j + 1];
      const int32_t* p = std::lower_bound(b, e, (int32_t)i);
      if (p == e || *p != i) return false;
    }
  return true;
}

void upload_csr(const ga_csr* A, int dtype, OwnedCsr& out, hipStream_t s) {
  out.nrows = A->nrows;
  out.ncols = A->ncols;
  out.nnz = A->nnz;
  out.rp.ensure(sizeof(int32_t) * (A->nrows + 1));
  out.col.ensure(sizeof(int32_t) * std::max<int64_t>(A->nnz, 1));
  GA_HIP(hipMemcpyAsync(out.rp.p, A->rowptr, sizeof(int32_t) * (A->nrows + 1), hipMemcpyHostToDevice, s));
  if (A->nnz) GA_HIP(hipMemcpyAsync(out.col.p, A->col, sizeof(int32_t) * A->nnz, hipMemcpyHostToDevice, s));
  if (A->val) {
    size_t es = dtype_size(dtype);
    out.val.ensure(es * std::max<int64_t>(A->nnz, 1));
    if (A->nnz) GA_HIP(hipMemcpyAsync(out.val.p, A->val, es * A->nnz, hipMemcpyHostToDevice, s));
  } else {
    out.val.release();
  }
}

// ---- fills ---------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (uint32_t*)p, v, n);
  GA_HIP(hipGetLastError());
}
void fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (uint64_t*)p, v, n);
  GA_HIP(hipGetLastError());
}

// ---- exclusive scan (reduce, scan tile sums, apply) ----------------------
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_PER;

__global__ __launch_bounds__(SCAN_T) void k_scan_tile_sums(const int32_t* __restrict__ in, int64_t n,
                                                           long long* __restrict__ bsum) {
  __shared__ long long red[SCAN_T / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  long long s = 0;
  for (int q = threadIdx.x; q < SCAN_TILE; q += SCAN_T) {
    int64_t idx = base + q;
    if (idx < n) s += in[idx];
  }
  s = wave_reduce_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < SCAN_T / 64; ++w) t += red[w];
    bsum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_scan_tile_offsets(long long* __restrict__ bsum, int64_t nb,
                                                            int32_t* __restrict__ out_total) {
  __shared__ long long wtot[16];
  __shared__ long long carry;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
    const int64_t idx = c0 + tid;
    const long long v = idx < nb ? bsum[idx] : 0;
    const long long incl = wave_incl_scan(v);
    if (lane == 63) wtot[wv] = incl;
    __syncthreads();
    long long wbase = 0, ctot = 0;
    for (int k = 0; k < 16; ++k) {
      if (k < wv) wbase += wtot[k];
      ctot += wtot[k];
    }
    const long long c = carry;
    if (idx < nb) bsum[idx] = c + wbase + incl - v;
    __syncthreads();
    if (tid == 0) carry = c + ctot;
    __syncthreads();
  }
  if (tid == 0) *out_total = (int32_t)carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_tile_apply(const int32_t* __restrict__ in,
                                                            int32_t* __restrict__ out, int64_t n,
                                                            const long long* __restrict__ bsum) {
  __shared__ long long wtot[SCAN_T / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)tid * SCAN_PER;
  int32_t v[SCAN_PER];
  long long tsum = 0;
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    tsum += v[q];
  }
  const long long incl = wave_incl_scan(tsum);
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  long long off = bsum[blockIdx.x] + incl - tsum;
  for (int k = 0; k < wv; ++k) off += wtot[k];
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    if (base + q < n) out[base + q] = (int32_t)off;
    off += v[q];
  }
}

// Tile scan with the tile offsets summed in place: each workgroup adds the
// sums of the tiles before it (at most SCAN_FUSED_TILES of them), so short
// scans take two launches instead of three.  The last tile writes the total.
constexpr int SCAN_FUSED_TILES = 1024;
__global__ __launch_bounds__(SCAN_T) void k_scan_tile_apply_fused(const int32_t* __restrict__ in,
                                                                  int32_t* __restrict__ out, int64_t n,
                                                                  const long long* __restrict__ bsum) {
  __shared__ long long wtot[SCAN_T / 64];
  __shared__ long long prior[SCAN_T / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  long long before = 0;
  for (int b = tid; b < (int)blockIdx.x; b += SCAN_T) before += bsum[b];
  before = wave_reduce_sum(before);
  if (lane == 0) prior[wv] = before;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)tid * SCAN_PER;
  int32_t v[SCAN_PER];
  long long tsum = 0;
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    tsum += v[q];
  }
  const long long incl = wave_incl_scan(tsum);
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  long long off = incl - tsum;
  for (int k = 0; k < SCAN_T / 64; ++k) {
    off += prior[k];
    if (k < wv) off += wtot[k];
  }
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    if (base + q < n) out[base + q] = (int32_t)off;
    off += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && tid == SCAN_T - 1) out[n] = (int32_t)off;
}

void scan_tile_offsets(long long* bsum, int64_t nb, int32_t* total, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, s, bsum, nb, total);
  GA_HIP(hipGetLastError());
}

size_t scan_tmp_bytes(int64_t n) { return n <= 0 ? 0 : lease_bytes(sizeof(long long) * ceil_div(n, SCAN_TILE)); }

void scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, DevBuf& tmp, hipStream_t s) {
  if (n <= 0) {
    fill_u32(out, 0u, 1, s);
    return;
  }
  const int64_t nb = ceil_div(n, SCAN_TILE);
  tmp.ensure(sizeof(long long) * nb);
  long long* bsum = tmp.as<long long>();
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, n, bsum);
  if (nb <= SCAN_FUSED_TILES) {
    hipLaunchKernelGGL(k_scan_tile_apply_fused, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, out, n,
                       (const long long*)bsum);
  } else {
    hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, s, bsum, nb, out + n);
    hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, out, n,
                       (const long long*)bsum);
  }
  GA_HIP(hipGetLastError());
}

// ---- lower triangle --------------------------------------------------------
__global__ void k_lower_count(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                              int32_t* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = rp[i], hi = rp[i + 1];
    const int b = lo;
    while (lo < hi) {  // first column >= i
      int mid = (lo + hi) >> 1;
      if (col[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    cnt[i] = lo - b;
  }
}

__global__ void k_lower_copy(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                             const int32_t* __restrict__ lrp, int32_t* __restrict__ lcol) {
  // one wavefront per row
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * wpb) {
    const int src = rp[i], dst = lrp[i], len = lrp[i + 1] - lrp[i];
    for (int q = lane; q < len; q += 64) lcol[dst + q] = col[src + q];
  }
}

void lower_triangle_dev(const DCsr& A, OwnedCsr& L, DevBuf& tmp, hipStream_t s) {
  const int64_t n = A.nrows;
  L.nrows = n;
  L.ncols = A.ncols;
  L.rp.ensure(sizeof(int32_t) * (n + 1));
  DevBuf cnt;
  cnt.ensure(sizeof(int32_t) * std::max<int64_t>(n, 1));
  hipLaunchKernelGGL(k_lower_count, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, A.rp, A.col, n,
                     cnt.as<int32_t>());
  scan_exclusive_i32(cnt.as<int32_t>(), L.rp.as<int32_t>(), n, tmp, s);
  int32_t lnnz = 0;
  GA_HIP(hipMemcpyAsync(&lnnz, L.rp.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  L.nnz = lnnz;
  L.col.ensure(sizeof(int32_t) * std::max<int64_t>(lnnz, 1));
  hipLaunchKernelGGL(k_lower_copy, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.rp, A.col, n,
                     (const int32_t*)L.rp.as<int32_t>(), L.col.as<int32_t>());
  GA_HIP(hipGetLastError());
}

// Plain streaming copy (16 B per lane per iteration): the calibration
// reference for HBM byte counters and the achievable-bandwidth probe.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace ga

using namespace ga;

extern "C" {

const char* ga_last_error(void) { return t_err.c_str(); }

ga_status ga_init(int device) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_ctx && g_ctx->device == device) return;
    int count = 0;
    GA_HIP(hipGetDeviceCount(&count));
    GA_REQUIRE(device >= 0 && device < count, GA_ERR_ARG, "no HIP device " + std::to_string(device));
    GA_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    GA_HIP(hipGetDeviceProperties(&prop, device));
    auto* c = new Context();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->hbm_bytes = (int64_t)prop.totalGlobalMem;
    c->lds_per_block = (int64_t)prop.sharedMemPerBlock;
    c->name = prop.gcnArchName;
    GA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if (const char* wf = std::getenv("GA_WAVE_FLOPS")) c->wave_flops = std::atoll(wf);
    if (const char* di = std::getenv("GA_DIRECT_ITEMS")) c->direct_items = std::atoi(di) != 0;
    if (const char* ss = std::getenv("GA_SELECTIVE_SUPPORT"))
      c->selective_support = std::min(1, std::max(-1, std::atoi(ss)));
    if (const char* sj = std::getenv("GA_SERVE_JOIN")) c->serve_join = std::min(2, std::max(0, std::atoi(sj)));
    if (const char* bl = std::getenv("GA_BC_LOOP")) c->bc_loop = std::min(2, std::max(0, std::atoi(bl)));
    if (const char* st = std::getenv("GA_SUPPORT_TRI")) c->support_tri = std::min(1, std::max(-1, std::atoi(st)));
    if (const char* tl = std::getenv("GA_TRUSS_LOOP")) c->truss_loop = std::min(4, std::max(-1, std::atoi(tl)));
    if (const char* iw = std::getenv("GA_ITEM_WEDGES")) c->item_wedges = std::max(1LL, std::atoll(iw));
    if (const char* rc = std::getenv("GA_ROW_CHUNK")) c->row_chunk = std::max(0, std::atoi(rc));
    if (const char* bu = std::getenv("GA_BC_BOTTOM_UP")) c->bc_bottom_up = std::atoi(bu) != 0;
    if (const char* ba = std::getenv("GA_BC_ALPHA")) c->bc_alpha = std::max(0.01, std::atof(ba));
    if (const char* ss = std::getenv("GA_SUM_SUB")) {
      const int v = std::atoi(ss);
      c->sum_sub = (v == 1 || v == 2 || v == 8 || v == 16 || v == 32 || v == 64) ? v : 0;
    }
    c->prof_dev.ensure(64);
    GA_HIP(hipMemsetAsync(c->prof_dev.p, 0, 64, c->stream));
    GA_HIP(hipStreamSynchronize(c->stream));
    noalloc_setup(device);
    g_ctx = c;
  });
}

ga_status ga_set_option(int option, double value) {
  return guarded([&] {
    Context& c = ctx();
    const bool integral = value == (double)(long long)value;
    switch (option) {
      case GA_OPT_BC_BOTTOM_UP:
        GA_REQUIRE(value == 0.0 || value == 1.0, GA_ERR_ARG, "bc_bottom_up is 0 or 1");
        c.bc_bottom_up = value != 0.0;
        break;
      case GA_OPT_BC_ALPHA:
        GA_REQUIRE(value > 0.0, GA_ERR_ARG, "bc_alpha must be positive");
        c.bc_alpha = value;
        break;
      case GA_OPT_ROW_CLAIM:
        GA_REQUIRE(integral && value >= 0.0 && value <= 1 << 20, GA_ERR_ARG, "row claim is 0..2^20");
        c.row_chunk = (int)value;
        break;
      case GA_OPT_SUM_SUB:
        GA_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0 || value == 8.0 || value == 16.0 || value == 32.0 ||
                       value == 64.0,
                   GA_ERR_ARG, "sum_sub is 0, 1, 2, 8, 16, 32 or 64");
        c.sum_sub = (int)value;
        break;
      case GA_OPT_ITEM_WEDGES:
        GA_REQUIRE(integral && value >= 1.0, GA_ERR_ARG, "item_wedges must be a positive integer");
        c.item_wedges = (long long)value;
        break;
      case GA_OPT_SELECTIVE_SUPPORT:
        GA_REQUIRE(value == -1.0 || value == 0.0 || value == 1.0, GA_ERR_ARG, "selective_support is -1, 0 or 1");
        c.selective_support = (int)value;
        break;
      case GA_OPT_SERVE_JOIN:
        GA_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0, GA_ERR_ARG, "serve_join is 0, 1 or 2");
        c.serve_join = (int)value;
        break;
      case GA_OPT_BC_LOOP:
        GA_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0, GA_ERR_ARG, "bc_loop is 0, 1 or 2");
        c.bc_loop = (int)value;
        break;
      case GA_OPT_WAVE_FLOPS:
        GA_REQUIRE(integral && value >= -1.0 && value <= 1e12, GA_ERR_ARG, "wave_flops is -1 .. 1e12");
        c.wave_flops = (long long)value;
        break;
      case GA_OPT_SUPPORT_TRI:
        GA_REQUIRE(value == -1.0 || value == 0.0 || value == 1.0, GA_ERR_ARG, "support_tri is -1, 0 or 1");
        c.support_tri = (int)value;
        break;
      case GA_OPT_TRUSS_LOOP:
        GA_REQUIRE(value == -1.0 || value == 0.0 || value == 1.0 || value == 2.0 || value == 3.0 || value == 4.0,
                   GA_ERR_ARG, "truss_loop is -1, 0, 1, 2, 3 or 4");
        c.truss_loop = (int)value;
        break;
      default:
        GA_REQUIRE(false, GA_ERR_ARG, "unknown option " + std::to_string(option));
    }
  });
}

ga_status ga_get_option(int option, double* value) {
  return guarded([&] {
    GA_REQUIRE(value != nullptr, GA_ERR_ARG, "null argument");
    Context& c = ctx();
    switch (option) {
      case GA_OPT_BC_BOTTOM_UP: *value = c.bc_bottom_up ? 1.0 : 0.0; break;
      case GA_OPT_BC_ALPHA: *value = c.bc_alpha; break;
      case GA_OPT_ROW_CLAIM: *value = c.row_chunk; break;
      case GA_OPT_SUM_SUB: *value = c.sum_sub; break;
      case GA_OPT_ITEM_WEDGES: *value = (double)c.item_wedges; break;
      case GA_OPT_SELECTIVE_SUPPORT: *value = c.selective_support; break;
      case GA_OPT_TRUSS_LOOP: *value = c.truss_loop; break;
      case GA_OPT_SUPPORT_TRI: *value = c.support_tri; break;
      case GA_OPT_SERVE_JOIN: *value = c.serve_join; break;
      case GA_OPT_BC_LOOP: *value = c.bc_loop; break;
      case GA_OPT_WAVE_FLOPS: *value = (double)c.wave_flops; break;
      default: GA_REQUIRE(false, GA_ERR_ARG, "unknown option " + std::to_string(option));
    }
  });
}

ga_status ga_device_info(char* name, int name_len, int* compute_units, int64_t* hbm_bytes) {
  return guarded([&] {
    Context& c = ctx();
    if (name && name_len > 0) {
      std::strncpy(name, c.name.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (compute_units) *compute_units = c.num_cu;
    if (hbm_bytes) *hbm_bytes = c.hbm_bytes;
  });
}

ga_status ga_synchronize(void) {
  return guarded([&] { GA_HIP(hipStreamSynchronize(ctx().stream)); });
}

ga_status ga_check_csr(const ga_csr* A) {
  return guarded([&] { validate_host_csr(A, "A"); });
}

void ga_free_host(void* p) { std::free(p); }

ga_status ga_host_alloc(int64_t bytes, void** ptr) {
  return guarded([&] {
    GA_REQUIRE(ptr && bytes > 0, GA_ERR_ARG, "ga_host_alloc needs a positive size and an output pointer");
    *ptr = nullptr;
    ctx();  // the device is initialised
    GA_HIP(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
  });
}

ga_status ga_host_free(void* ptr) {
  return guarded([&] {
    if (ptr) GA_HIP(hipHostFree(ptr));
  });
}

ga_status ga_stream_copy(int64_t bytes, int reps, double* ms_per_copy) {
  return guarded([&] {
    GA_REQUIRE(bytes > 0 && bytes % 16 == 0 && reps > 0, GA_ERR_ARG, "bytes must be a positive multiple of 16");
    Context& c = ctx();
    DevBuf a, b;
    a.ensure(bytes);
    b.ensure(bytes);
    fill_u32(a.p, 0x01020304u, bytes / 4, c.stream);
    const int64_t n16 = bytes / 16;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n16, 256), (int64_t)c.num_cu * 32);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c.stream, (const uint4*)a.p, (uint4*)b.p, n16);
    hipEvent_t e0, e1;
    GA_HIP(hipEventCreate(&e0));
    GA_HIP(hipEventCreate(&e1));
    GA_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c.stream, (const uint4*)a.p, (uint4*)b.p, n16);
    GA_HIP(hipEventRecord(e1, c.stream));
    GA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GA_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
  });
}

ga_status ga_prof_enable(int flags) {
  return guarded([&] {
    ctx().prof_events_on = (flags & 1) != 0;
    ctx().prof_bytes = (flags & 2) != 0;
  });
}

ga_status ga_prof_read(ga_kernel_prof* out, int reset) {
  return guarded([&] {
    GA_REQUIRE(out != nullptr, GA_ERR_ARG, "null output");
    Context& c = ctx();
    GA_HIP(hipStreamSynchronize(c.stream));
    double ms = 0.0;
    for (auto& pr : c.prof_events) {
      float t = 0.f;
      GA_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
      ms += t;
    }
    unsigned long long w[4] = {0, 0, 0, 0};
    GA_HIP(hipMemcpy(w, c.prof_dev.p, sizeof(w), hipMemcpyDeviceToHost));
    out->launches = c.prof_launches;
    out->accounted_launches = c.prof_bytes_launches;
    out->kernel_ms = ms;
    out->wedges = (int64_t)(w[1] + w[3]);
    out->wedges_processed = (int64_t)w[0];
    // [1]: wedges of products with 4-byte gathers; [2]: algorithmic bytes
    // reported directly by kernels with another per-wedge cost (BC), whose
    // wedges are counted in [3]
    out->bytes_alg = c.prof_bytes_host + 4 * (int64_t)w[1] + (int64_t)w[2];
    if (reset) {
      for (auto& pr : c.prof_events) {
        c.event_pool.push_back(pr.first);
        c.event_pool.push_back(pr.second);
      }
      c.prof_events.clear();
      c.prof_launches = 0;
      c.prof_bytes_launches = 0;
      c.prof_bytes_host = 0;
      GA_HIP(hipMemsetAsync(c.prof_dev.p, 0, 64, c.stream));
      GA_HIP(hipStreamSynchronize(c.stream));
    }
  });
}

}  // extern "C"

// Batched betweenness centrality, FP64 path counts and dependencies.
//
// Restates the linear-algebraic Brandes of betweenness_centrality_batch
// (include/betweenness_centrality.h:125-237) as one tall-skinny multi-source
// pass.  The reference keeps a sparse frontier n x ns (CSR), grows the
// visited set with add_nointersect and advances with the complement-masked
// SpGEMM (AT * Front)<!Numsp>; backward it runs the masked sparse x dense
// product A * W on the previous level's pattern.  Here the per-(vertex,
// source) state is dense and vertex-major -- depth (uint16), sigma (FP64),
// q = (1 + delta) / sigma (FP64), row v holding all sources contiguously
// (padded to a multiple of 64) -- so a wavefront owning vertex v reads the
// state of 64 sources with one coalesced load, and the sparse structure is
// carried by per-level vertex lists:
//
//  forward, level d -> d+1 (push along out-edges, the frontier product):
//    for u in F_d, for v in N(u), for sources s with depth(u,s) = d and
//    depth(v,s) in {unvisited, d+1}: depth(v,s) = d+1, sigma(v,s) += sigma(u,s)
//    The complement mask is the depth test; sigma values are integers, so the
//    FP64 atomic additions are exact in any order (below 2^53).
//  backward, level k = dmax-1 .. 1 (pull along out-edges, the A * W product):
//    for v in F_k, for sources s with depth(v,s) = k:
//      delta(v,s) = sigma(v,s) * sum_{w in N(v), depth(w,s) = k+1} q(w,s)
//    summed in CSR order (deterministic), then q(v,s) = (1 + delta) / sigma
//    and bc[v] += sum_s delta(v,s) (a fixed-order wavefront reduction).
// The result is bc[v] = sum_{s in S} delta_s(v): unnormalised, not halved.
//
// Both sweeps run over work items of at most BC_CH neighbours of one level
// vertex, one wavefront each, so the adjacency list of a hub is spread over
// many wavefronts instead of serialising its level; the backward partial sums
// of a hub are added in chunk order by a second pass (fixed order: the FP64
// result is the same on every run).
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace ga {

constexpr uint16_t BC_UNVISITED = 0xFFFF;
constexpr int BC_MAX_LEVELS = 0xFFFE;

__global__ void k_bc_init(uint16_t* __restrict__ depth, double* __restrict__ sigma, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    depth[i] = BC_UNVISITED;
    sigma[i] = 0.0;
  }
}

__global__ void k_bc_sources(const int32_t* __restrict__ src, int ns, int nsp, uint16_t* __restrict__ depth,
                             double* __restrict__ sigma, int32_t* __restrict__ level0) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ns; j += gridDim.x * blockDim.x) {
    const int64_t idx = (int64_t)src[j] * nsp + j;
    depth[idx] = 0;
    sigma[idx] = 1.0;
    level0[j] = src[j];
  }
}

// frontier list of the next level from the flags (ascending vertex order)
__global__ void k_flag_count(const uint8_t* __restrict__ flag, int64_t n, int32_t* __restrict__ cnt, int chunk) {
  const int64_t i0 = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) * chunk;
  if (i0 >= n) return;
  int c = 0;
  for (int64_t i = i0; i < i0 + chunk && i < n; ++i) c += flag[i] != 0;
  cnt[blockIdx.x * blockDim.x + threadIdx.x] = c;
}

__global__ void k_flag_emit(uint8_t* __restrict__ flag, int64_t n, const int32_t* __restrict__ off, int chunk,
                            int32_t* __restrict__ out) {
  const int64_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i0 = g * chunk;
  if (i0 >= n) return;
  int p = off[g];
  for (int64_t i = i0; i < i0 + chunk && i < n; ++i) {
    if (flag[i]) {
      out[p++] = (int32_t)i;
      flag[i] = 0;
    }
  }
}

__global__ __launch_bounds__(256) void k_bc_last(const int32_t* __restrict__ front, int nfront, int nsp, int k,
                                                 const uint16_t* __restrict__ depth,
                                                 const double* __restrict__ sigma, double* __restrict__ q) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  for (int f = blockIdx.x * wpb + (threadIdx.x >> 6); f < nfront; f += gridDim.x * wpb) {
    const int v = front[f];
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] == (uint16_t)k) q[vi] = 1.0 / sigma[vi];
    }
  }
}

// ---- neighbour-chunked items ----------------------------------------------
// A frontier vertex with deg > BC_CH becomes ceil(deg / BC_CH) items so that
// hub adjacency lists are spread over many wavefronts; every vertex has at
// least one item (a backward vertex without successors still needs q).
constexpr int BC_CH = 256;
constexpr int BC_UNROLL = 4;

__global__ void k_bc_item_count(const int32_t* __restrict__ front, int nf, const int32_t* __restrict__ rp,
                                int32_t* __restrict__ cnt, int32_t* __restrict__ hcnt) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const int v = front[f];
    const int deg = rp[v + 1] - rp[v];
    const int nc = deg > BC_CH ? (deg + BC_CH - 1) / BC_CH : 1;
    cnt[f] = nc;
    hcnt[f] = nc > 1 ? nc : 0;
  }
}

__global__ void k_bc_item_fill(const int32_t* __restrict__ front, int nf, const int32_t* __restrict__ rp,
                               const int32_t* __restrict__ start, const int32_t* __restrict__ hstart,
                               int32_t* __restrict__ it_v, int32_t* __restrict__ it_t0,
                               int32_t* __restrict__ it_slot) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const int v = front[f];
    const int deg = rp[v + 1] - rp[v];
    const int nc = deg > BC_CH ? (deg + BC_CH - 1) / BC_CH : 1;
    for (int c = 0; c < nc; ++c) {
      const int q = start[f] + c;
      it_v[q] = v;
      it_t0[q] = rp[v] + c * BC_CH;
      it_slot[q] = nc > 1 ? hstart[f] + c : -1;
    }
  }
}

// forward, level d -> d+1: one wavefront per item (vertex u, <= BC_CH
// neighbours), lanes = 64 sources; BC_UNROLL neighbour states in flight
__global__ __launch_bounds__(256) void k_bc_forward_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ n_items_ptr, int nsp, int d,
    uint16_t* __restrict__ depth, double* __restrict__ sigma, uint8_t* __restrict__ next_flag) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int n_items = *n_items_ptr;
  const uint16_t dd = (uint16_t)d, dn = (uint16_t)(d + 1);
  for (int q = blockIdx.x * wpb + (threadIdx.x >> 6); q < n_items; q += gridDim.x * wpb) {
    const int u = it_v[q];
    const int t0 = it_t0[q];
    const int t1 = min(t0 + BC_CH, rp[u + 1]);
    for (int c = 0; c < nsp; c += 64) {
      const int64_t ui = (int64_t)u * nsp + c + lane;
      const bool act = depth[ui] == dd;
      if (__ballot(act) == 0ull) continue;
      const double su = act ? sigma[ui] : 0.0;
      for (int t = t0; t < t1; t += BC_UNROLL) {
        int vv[BC_UNROLL];
        uint16_t dv[BC_UNROLL];
#pragma unroll
        for (int k = 0; k < BC_UNROLL; ++k) vv[k] = (t + k < t1) ? col[t + k] : -1;
#pragma unroll
        for (int k = 0; k < BC_UNROLL; ++k)
          dv[k] = (act && vv[k] >= 0) ? depth[(int64_t)vv[k] * nsp + c + lane] : (uint16_t)0;
#pragma unroll
        for (int k = 0; k < BC_UNROLL; ++k) {
          if (act && vv[k] >= 0 && (dv[k] == BC_UNVISITED || dv[k] == dn)) {
            const int64_t vi = (int64_t)vv[k] * nsp + c + lane;
            if (dv[k] == BC_UNVISITED) depth[vi] = dn;
            atomicAdd(&sigma[vi], su);
            next_flag[vv[k]] = 1;
          }
        }
      }
    }
  }
}

// backward, level k: one wavefront per item; successors summed in CSR order
// within the item.  Single-item vertices finish delta, q and bc here; the
// items of a hub vertex write per-chunk partial sums that k_bc_combine adds
// in chunk order (a fixed order: results are reproducible run to run).
__global__ __launch_bounds__(256) void k_bc_backward_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ it_slot, const int32_t* __restrict__ n_items_ptr,
    int nsp, int k, const uint16_t* __restrict__ depth, const double* __restrict__ sigma, double* __restrict__ q,
    double* __restrict__ bc, double* __restrict__ partial) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int n_items = *n_items_ptr;
  const uint16_t dk = (uint16_t)k, dn = (uint16_t)(k + 1);
  for (int it = blockIdx.x * wpb + (threadIdx.x >> 6); it < n_items; it += gridDim.x * wpb) {
    const int v = it_v[it];
    const int t0 = it_t0[it];
    const int t1 = min(t0 + BC_CH, rp[v + 1]);
    const int slot = it_slot[it];
    double part = 0.0;
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      const bool act = depth[vi] == dk;
      if (__ballot(act) == 0ull) continue;
      double acc = 0.0;
      for (int t = t0; t < t1; t += BC_UNROLL) {
        int ww[BC_UNROLL];
        uint16_t dw[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) ww[j] = (t + j < t1) ? col[t + j] : -1;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j)
          dw[j] = (act && ww[j] >= 0) ? depth[(int64_t)ww[j] * nsp + c + lane] : (uint16_t)0;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j)
          if (act && ww[j] >= 0 && dw[j] == dn) acc += q[(int64_t)ww[j] * nsp + c + lane];
      }
      if (slot >= 0) {
        if (act) partial[(int64_t)slot * nsp + c + lane] = acc;
      } else if (act) {
        const double sg = sigma[vi];
        const double delta = sg * acc;
        q[vi] = (1.0 + delta) / sg;
        part += delta;
      }
    }
    if (slot < 0) {
      part = wave_reduce_sum(part);
      if (lane == 0) bc[v] += part;
    }
  }
}

__global__ __launch_bounds__(256) void k_bc_combine(const int32_t* __restrict__ front, int nf,
                                                    const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ hstart, int nsp, int k,
                                                    const uint16_t* __restrict__ depth,
                                                    const double* __restrict__ sigma, double* __restrict__ q,
                                                    double* __restrict__ bc, const double* __restrict__ partial) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const uint16_t dk = (uint16_t)k;
  for (int f = blockIdx.x * wpb + (threadIdx.x >> 6); f < nf; f += gridDim.x * wpb) {
    const int v = front[f];
    const int deg = rp[v + 1] - rp[v];
    if (deg <= BC_CH) continue;  // single-item vertex: finished in the item pass
    const int nc = (deg + BC_CH - 1) / BC_CH;
    const int slot0 = hstart[f];
    double part = 0.0;
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] != dk) continue;
      double acc = 0.0;
      for (int j = 0; j < nc; ++j) acc += partial[(int64_t)(slot0 + j) * nsp + c + lane];
      const double sg = sigma[vi];
      const double delta = sg * acc;
      q[vi] = (1.0 + delta) / sg;
      part += delta;
    }
    part = wave_reduce_sum(part);
    if (lane == 0) bc[v] += part;
  }
}

// item list of a frontier (device scans; the item count stays on the device)
static void bc_items(const DCsr& A, const int32_t* front, int nf, BcWork& w, hipStream_t s) {
  w.icnt.ensure(sizeof(int32_t) * std::max(nf, 1));
  w.hcnt.ensure(sizeof(int32_t) * std::max(nf, 1));
  w.istart.ensure(sizeof(int32_t) * (nf + 1));
  w.hstart.ensure(sizeof(int32_t) * (nf + 1));
  hipLaunchKernelGGL(k_bc_item_count, dim3(grid_for(nf, 256, 8192)), dim3(256), 0, s, front, nf, A.rp,
                     w.icnt.as<int32_t>(), w.hcnt.as<int32_t>());
  scan_exclusive_i32(w.icnt.as<int32_t>(), w.istart.as<int32_t>(), nf, w.scan_tmp, s);
  scan_exclusive_i32(w.hcnt.as<int32_t>(), w.hstart.as<int32_t>(), nf, w.scan_tmp, s);
  hipLaunchKernelGGL(k_bc_item_fill, dim3(grid_for(nf, 256, 8192)), dim3(256), 0, s, front, nf, A.rp,
                     (const int32_t*)w.istart.as<int32_t>(), (const int32_t*)w.hstart.as<int32_t>(),
                     w.it_v.as<int32_t>(), w.it_t0.as<int32_t>(), w.it_slot.as<int32_t>());
  GA_HIP(hipGetLastError());
}

// Algorithmic accounting of one batch: reached (vertex, source) pairs P and
// wedges W = sum over reached pairs of deg(v); bytes = 24 W + 12 P (forward
// and backward: source index + FP64 sigma / W per wedge; sigma / delta per pair).
__global__ void k_bc_account(const int32_t* __restrict__ rp, const uint16_t* __restrict__ depth, int64_t n, int nsp,
                             unsigned long long* __restrict__ prof) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  unsigned long long wedges = 0, pairs = 0;
  for (int64_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += (int64_t)gridDim.x * wpb) {
    int cnt = 0;
    for (int c = lane; c < nsp; c += 64) cnt += depth[v * nsp + c] != BC_UNVISITED;
    cnt = wave_reduce_sum(cnt);
    if (lane == 0) {
      wedges += (unsigned long long)cnt * (unsigned long long)(rp[v + 1] - rp[v]);
      pairs += (unsigned long long)cnt;
    }
  }
  if (lane == 0 && (wedges || pairs)) {
    atomicAdd(&prof[3], wedges);
    atomicAdd(&prof[2], 24ull * wedges + 12ull * pairs);
  }
}

static hipEvent_t bc_event(Context& c) {
  hipEvent_t ev;
  if (!c.event_pool.empty()) {
    ev = c.event_pool.back();
    c.event_pool.pop_back();
  } else {
    GA_HIP(hipEventCreate(&ev));
  }
  return ev;
}

static void frontier_from_flags(uint8_t* flag, int64_t n, BcWork& w, int32_t* out, int32_t* count_h,
                                hipStream_t s) {
  const int chunk = 64;
  const int64_t ng = ceil_div(n, chunk);
  w.cnt.ensure(sizeof(int32_t) * std::max<int64_t>(ng, 1));
  w.off.ensure(sizeof(int32_t) * (ng + 1));
  hipLaunchKernelGGL(k_flag_count, dim3(grid_for(ng, 256)), dim3(256), 0, s, flag, n, w.cnt.as<int32_t>(), chunk);
  scan_exclusive_i32(w.cnt.as<int32_t>(), w.off.as<int32_t>(), ng, w.scan_tmp, s);
  hipLaunchKernelGGL(k_flag_emit, dim3(grid_for(ng, 256)), dim3(256), 0, s, flag, n,
                     (const int32_t*)w.off.as<int32_t>(), chunk, out);
  GA_HIP(hipMemcpyAsync(count_h, w.off.as<int32_t>() + ng, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
}

// A: device pattern (out-edges); src: host array of distinct sources.
void bc_batch_dev(const DCsr& A, const int32_t* src_h, int64_t ns, double* bc_h, BcWork& w, hipStream_t s) {
  const int64_t n = A.nrows;
  GA_REQUIRE(A.nrows == A.ncols, GA_ERR_DIM, "betweenness centrality needs a square matrix");
  GA_REQUIRE(ns >= 1 && ns <= 4096, GA_ERR_LIMIT, "between 1 and 4096 sources per batch");
  {
    std::vector<char> seen(n, 0);
    for (int64_t j = 0; j < ns; ++j) {
      GA_REQUIRE(src_h[j] >= 0 && src_h[j] < n, GA_ERR_ARG, "source out of range");
      GA_REQUIRE(!seen[src_h[j]], GA_ERR_ARG, "duplicate source");
      seen[src_h[j]] = 1;
    }
  }
  const int nsp = (int)ceil_div(ns, 64) * 64;
  const int64_t total = n * nsp;
  w.depth.ensure(sizeof(uint16_t) * total);
  w.sigma.ensure(sizeof(double) * total);
  w.q.ensure(sizeof(double) * total);
  w.bc.ensure(sizeof(double) * std::max<int64_t>(n, 1));
  w.flag.ensure(4 * ceil_div(std::max<int64_t>(n, 1), 4));
  w.src.ensure(sizeof(int32_t) * ns);
  const int64_t item_cap = n + ceil_div(A.nnz, BC_CH) + 1;
  w.it_v.ensure(sizeof(int32_t) * item_cap);
  w.it_t0.ensure(sizeof(int32_t) * item_cap);
  w.it_slot.ensure(sizeof(int32_t) * item_cap);
  w.partial.ensure(sizeof(double) * (size_t)(2 * ceil_div(A.nnz, BC_CH) + 1) * nsp);
  GA_HIP(hipMemcpyAsync(w.src.p, src_h, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_bc_init, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, w.depth.as<uint16_t>(),
                     w.sigma.as<double>(), total);
  fill_u64(w.bc.p, 0ull, n, s);
  fill_u32(w.flag.p, 0u, ceil_div(std::max<int64_t>(n, 1), 4), s);
  // level lists, grown as the BFS proceeds
  std::vector<int64_t> loff{0, ns};
  std::vector<int32_t> lcnt{(int32_t)ns};
  int64_t cap = std::max<int64_t>(4 * n, ns + n);
  w.levels.ensure(sizeof(int32_t) * cap);
  hipLaunchKernelGGL(k_bc_sources, dim3(grid_for(ns, 256)), dim3(256), 0, s, (const int32_t*)w.src.as<int32_t>(),
                     (int)ns, nsp, w.depth.as<uint16_t>(), w.sigma.as<double>(), w.levels.as<int32_t>());
  GA_HIP(hipGetLastError());
  int32_t* count_h = nullptr;
  GA_HIP(hipHostMalloc((void**)&count_h, 64, hipHostMallocDefault));
  try {
    int d = 0;
    while (true) {
      GA_REQUIRE(d < BC_MAX_LEVELS - 1, GA_ERR_LIMIT, "BFS deeper than 65533 levels");
      const int32_t nf = lcnt[d];
      const int32_t* front = w.levels.as<int32_t>() + loff[d];
      Context& cx = ctx();
      hipEvent_t fe0 = nullptr, fe1 = nullptr;
      if (cx.prof_events_on) {
        fe0 = bc_event(cx);
        fe1 = bc_event(cx);
        GA_HIP(hipEventRecord(fe0, s));
      }
      bc_items(A, front, nf, w, s);
      hipLaunchKernelGGL(k_bc_forward_items, dim3(grid_for(nf + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0, s, A.rp,
                         A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                         (const int32_t*)(w.istart.as<int32_t>() + nf), nsp, d, w.depth.as<uint16_t>(),
                         w.sigma.as<double>(), w.flag.as<uint8_t>());
      GA_HIP(hipGetLastError());
      if (cx.prof_events_on) {
        GA_HIP(hipEventRecord(fe1, s));
        cx.prof_events.emplace_back(fe0, fe1);
        cx.prof_launches += 1;
      }
      if (loff[d + 1] + n > cap) {  // make room for the next list (at most n vertices)
        int64_t ncap = std::max<int64_t>(2 * cap, loff[d + 1] + n);
        DevBuf nb;
        nb.ensure(sizeof(int32_t) * ncap);
        GA_HIP(hipMemcpyAsync(nb.p, w.levels.p, sizeof(int32_t) * loff[d + 1], hipMemcpyDeviceToDevice, s));
        GA_HIP(hipStreamSynchronize(s));
        w.levels = std::move(nb);
        cap = ncap;
      }
      frontier_from_flags(w.flag.as<uint8_t>(), n, w, w.levels.as<int32_t>() + loff[d + 1], count_h, s);
      const int32_t nn = *count_h;
      if (nn == 0) break;
      lcnt.push_back(nn);
      loff.push_back(loff[d + 1] + nn);
      ++d;
    }
    // d = deepest level that holds vertices
    const int dmax = d;
    Context& cx = ctx();
    if (cx.prof_bytes) {
      hipLaunchKernelGGL(k_bc_account, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.rp,
                         (const uint16_t*)w.depth.as<uint16_t>(), n, nsp, cx.prof_dev.as<unsigned long long>());
      cx.prof_bytes_host += (int64_t)(dmax + 1) * (4 * (n + 1) + 4 * A.nnz);
      cx.prof_bytes_launches += 2 * dmax;
    }
    hipEvent_t be0 = nullptr, be1 = nullptr;
    if (cx.prof_events_on) {
      be0 = bc_event(cx);
      be1 = bc_event(cx);
      GA_HIP(hipEventRecord(be0, s));
    }
    if (dmax >= 1) {
      hipLaunchKernelGGL(k_bc_last, dim3(grid_for(lcnt[dmax], 4, 65536)), dim3(256), 0, s,
                         (const int32_t*)(w.levels.as<int32_t>() + loff[dmax]), lcnt[dmax], nsp, dmax,
                         (const uint16_t*)w.depth.as<uint16_t>(), (const double*)w.sigma.as<double>(),
                         w.q.as<double>());
      for (int k = dmax - 1; k >= 1; --k) {
        const int32_t* front = w.levels.as<int32_t>() + loff[k];
        const int nf = lcnt[k];
        bc_items(A, front, nf, w, s);
        hipLaunchKernelGGL(k_bc_backward_items, dim3(grid_for(nf + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0, s,
                           A.rp, A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                           (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(w.istart.as<int32_t>() + nf),
                           nsp, k, (const uint16_t*)w.depth.as<uint16_t>(), (const double*)w.sigma.as<double>(),
                           w.q.as<double>(), w.bc.as<double>(), w.partial.as<double>());
        hipLaunchKernelGGL(k_bc_combine, dim3(grid_for(nf, 4, 65536)), dim3(256), 0, s, front, nf, A.rp,
                           (const int32_t*)w.hstart.as<int32_t>(), nsp, k, (const uint16_t*)w.depth.as<uint16_t>(),
                           (const double*)w.sigma.as<double>(), w.q.as<double>(), w.bc.as<double>(),
                           (const double*)w.partial.as<double>());
        GA_HIP(hipGetLastError());
      }
    }
    if (cx.prof_events_on) {
      GA_HIP(hipEventRecord(be1, s));
      cx.prof_events.emplace_back(be0, be1);
      cx.prof_launches += 1;
    }
    if (n) GA_HIP(hipMemcpyAsync(bc_h, w.bc.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipHostFree(count_h);
    throw;
  }
  GA_HIP(hipHostFree(count_h));
}

}  // namespace ga

using namespace ga;

extern "C" ga_status ga_bc_batch(const ga_csr* A, const int32_t* sources, int64_t nsources, double* bc) {
  return guarded([&] {
    validate_host_csr(A, "A");
    GA_REQUIRE(sources && bc, GA_ERR_ARG, "null argument");
    hipStream_t s = ctx().stream;
    OwnedCsr dA;
    upload_csr(A, GA_INT32, dA, s);
    DCsr v = dA.view();
    v.val = nullptr;
    BcWork w;
    bc_batch_dev(v, sources, nsources, bc, w, s);
  });
}

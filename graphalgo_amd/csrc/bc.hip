// Batched betweenness centrality, FP64 path counts and dependencies.
//
// Restates the linear-algebraic Brandes of betweenness_centrality_batch
// (include/betweenness_centrality.h:125-237) as one tall-skinny multi-source
// pass.  The reference keeps a sparse frontier n x ns (CSR), grows the
// visited set with add_nointersect and advances with the complement-masked
// SpGEMM (AT * Front)<!Numsp>; backward it runs the masked sparse x dense
// product A * W on the previous level's pattern.  Here the per-(vertex,
// source) state is dense and vertex-major -- depth (8 bits, 16 past 253 levels), sigma (FP64),
// q = (1 + delta) / sigma (FP64), row v holding all sources contiguously
// (padded to a multiple of 64) -- so a wavefront owning vertex v reads the
// state of 64 sources with one coalesced load, and the sparse structure is
// carried by per-level vertex lists:
//
//  forward, level d -> d+1 (the frontier product), in two sweeps:
//    discovery, pushed along out-edges: for u in F_d, v in N(u), sources s with
//      depth(u,s) = d and depth(v,s) unvisited: depth(v,s) = d+1 (the
//      complement mask is the depth test);
//    path counts, pulled: for v in F_{d+1}, sources s with depth(v,s) = d+1:
//      sigma(v,s) = sum_{u in N(v), depth(u,s) = d} sigma(u,s), in CSR order.
//    Neither sweep needs an atomic operation (the push version's FP64 atomics
//    on sigma serialise on the rows of hub vertices).
//  backward, level k = dmax-1 .. 1 (pull along out-edges, the A * W product):
//    for v in F_k, for sources s with depth(v,s) = k:
//      delta(v,s) = sigma(v,s) * sum_{w in N(v), depth(w,s) = k+1} q(w,s)
//    summed in CSR order (deterministic), then q(v,s) = (1 + delta) / sigma
//    and bc[v] += sum_s delta(v,s) (a fixed-order wavefront reduction).
// The result is bc[v] = sum_{s in S} delta_s(v): unnormalised, not halved.
//
// Every sweep runs over work items of at most BC_CH neighbours of one level
// vertex, one wavefront each, so the adjacency list of a hub is spread over
// many wavefronts instead of serialising its level; the partial sums of a hub
// (sigma forward, delta backward) are added in chunk order by a second pass
// (fixed order: the FP64 result is the same on every run).
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace ga {

// Depths are stored in DT = uint8_t (one byte per (vertex, lane): half the
// bytes of the dominant gathers) and the pass is redone with uint16_t in the
// rare case of a BFS deeper than 253 levels.  The all-ones value marks an
// unvisited pair.
template <typename DT>
__host__ __device__ constexpr DT bc_unvisited() { return (DT)~DT(0); }

// State of a pass whose level lists are sized on the device (device-driven
// loop): the host queues level steps without reading any size, and each
// kernel of a step takes its level, its lists and their sizes from here; a
// step queued after the last level finds `done` and does nothing.  The
// kernels below take it as `ctl` (null on the host-steered path, which
// passes levels and sizes as arguments).
struct BcCtl {
  int32_t* cur;   // level-d list (nf vertices, ascending)
  int32_t* next;  // level-(d+1) list being built
  int d, nf, done, dmax, bu, over, allow_bu, bk;  // bk: level of the backward step
  double alpha;
  int next_tot[4];  // list builder totals: vertices, items, hub slots, hub vertices
  int cand_tot[4];  // the same for the bottom-up candidates
  int back_tot[4];  // the same for a backward level
  unsigned long long edges[2];
  long long seq;
};
// which steps of the forward sweep run (every one only while !done)
enum { BG_ANY = 0, BG_CAND = 1, BG_DECIDE = 2, BG_BU = 3, BG_TD = 4, BG_POST = 5 };
__device__ inline bool bc_live(const BcCtl* c, int g) {
  if (g == BG_POST) return true;  // after the sweep
  if (c->done) return false;
  switch (g) {
    case BG_CAND: return c->d >= 1 && (c->bu || c->allow_bu);
    case BG_DECIDE: return c->d >= 1 && !c->bu && c->allow_bu;
    case BG_BU: return c->bu != 0;
    case BG_TD: return c->bu == 0;
    default: return true;
  }
}

template <typename DT>
__global__ void k_bc_init(DT* __restrict__ depth, double* __restrict__ sigma, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    depth[i] = bc_unvisited<DT>();
    sigma[i] = 0.0;
  }
}

// source j sits on lane lane[j]; its vertex is flagged for the level-0 list
// (a vertex that is a source on several lanes is listed once)
template <typename DT>
__global__ void k_bc_sources(const int32_t* __restrict__ src, const int32_t* __restrict__ lane, int ns, int nsp,
                             DT* __restrict__ depth, double* __restrict__ sigma, uint8_t* __restrict__ flag) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ns; j += gridDim.x * blockDim.x) {
    const int64_t idx = (int64_t)src[j] * nsp + lane[j];
    depth[idx] = 0;
    sigma[idx] = 1.0;
    flag[src[j]] = 1;
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

// flag[v] = 1 when some lane of vertex v is at depth k
template <typename DT>
__global__ __launch_bounds__(256) void k_level_flags(const DT* __restrict__ depth, int64_t n, int nsp, int k,
                                                     uint8_t* __restrict__ flag, const BcCtl* __restrict__ ctl) {
  if (ctl) k = ctl->bk;
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += (int64_t)gridDim.x * wpb) {
    bool hit = false;
    for (int c = lane; c < nsp; c += 64) hit |= depth[v * nsp + c] == (DT)k;
    if (__ballot(hit) != 0ull && lane == 0) flag[v] = 1;
  }
}

template <typename DT>
__global__ __launch_bounds__(256) void k_bc_last(const int32_t* __restrict__ front, int nfront, int nsp, int k,
                                                 const DT* __restrict__ depth,
                                                 const double* __restrict__ sigma, double* __restrict__ q,
                                                 const BcCtl* __restrict__ ctl) {
  if (ctl) {  // the deepest level's list
    front = ctl->cur;
    nfront = ctl->nf;
    k = ctl->dmax;
  }
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  for (int f = blockIdx.x * wpb + (threadIdx.x >> 6); f < nfront; f += gridDim.x * wpb) {
    const int v = front[f];
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] == (DT)k) q[vi] = 1.0 / sigma[vi];
    }
  }
}

// ---- neighbour-chunked items ----------------------------------------------
// A frontier vertex with deg > BC_CH becomes ceil(deg / BC_CH) items so that
// hub adjacency lists are spread over many wavefronts; every vertex has at
// least one item (a backward vertex without successors still needs q).
constexpr int BC_CH = 256;
constexpr int BC_UNROLL = 8;
// pull sweeps (bottom-up, backward): BC_CG 64-lane chunks of sources are
// walked together, BC_NU neighbours per step, so the 64-byte depth pieces
// and 512-byte value pieces of one neighbour's row are read together (one
// row per neighbour, not one piece per pass over the neighbour list)
constexpr int BC_CG = 4;
constexpr int BC_NU = 4;
constexpr int BC_FLAG_CHUNK = 64;  // vertices per thread when flags become a list

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

// forward, level d -> d+1, in two steps that need no atomics.
//
// Discovery: one wavefront per item of a level-d vertex u (<= BC_CH
// neighbours), lanes = 64 sources.  The item's neighbour ids are staged in
// LDS once (one coalesced load) and reused for every 64-lane chunk; BC_UNROLL
// neighbour depths are in flight per lane.  A neighbour pair that is still
// unvisited gets depth d+1 (every writer stores the same byte) and its vertex
// is flagged for the level-(d+1) list.
// (nb: the wavefront's BC_CH-entry LDS slice)
template <typename DT>
__device__ __forceinline__ void bc_discover_body(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                 const int32_t* __restrict__ it_v, const int32_t* __restrict__ it_t0,
                                                 int n_items, int nsp, int d, DT* __restrict__ depth,
                                                 uint8_t* __restrict__ next_flag, int32_t* nb) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const DT dd = (DT)d, dn = (DT)(d + 1);
  for (int q = blockIdx.x * wpb + wv; q < n_items; q += gridDim.x * wpb) {
    const int u = it_v[q];
    const int t0 = it_t0[q];
    const int cnt = min(BC_CH, rp[u + 1] - t0);
    for (int t = lane; t < cnt; t += 64) nb[t] = col[t0 + t];
    __threadfence_block();
    for (int c = 0; c < nsp; c += 64) {
      const bool act = depth[(int64_t)u * nsp + c + lane] == dd;
      if (__ballot(act) == 0ull) continue;
      for (int t = 0; t < cnt; t += BC_UNROLL) {
        int vv[BC_UNROLL];
        DT dv[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) vv[j] = (t + j < cnt) ? nb[t + j] : -1;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j)
          dv[j] = (act && vv[j] >= 0) ? depth[(int64_t)vv[j] * nsp + c + lane] : (DT)0;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) {
          if (act && vv[j] >= 0 && dv[j] == bc_unvisited<DT>()) {
            depth[(int64_t)vv[j] * nsp + c + lane] = dn;
            next_flag[vv[j]] = 1;
          }
        }
      }
    }
    __threadfence_block();  // the neighbour slots are rewritten by the next item
  }
}

template <typename DT>
__global__ __launch_bounds__(256) void k_bc_discover_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ n_items_ptr, int nsp, int d,
    DT* __restrict__ depth, uint8_t* __restrict__ next_flag, const BcCtl* __restrict__ ctl, int guard) {
  __shared__ int32_t s_nb[4][BC_CH];
  if (ctl) {
    if (!bc_live(ctl, guard)) return;
    d = ctl->d;
  }
  bc_discover_body<DT>(rp, col, it_v, it_t0, *n_items_ptr, nsp, d, depth, next_flag, s_nb[threadIdx.x >> 6]);
}

// Path counts of level d (pull along the edges, the frontier product):
// one wavefront per item of a level-d vertex v; for the lanes with
// depth(v, s) = d, sigma(v, s) = sum over neighbours u with depth(u, s) = d-1
// of sigma(u, s), in CSR order.  The items of a hub vertex write per-chunk
// partial sums that k_bc_sigma_combine adds in chunk order.  Level-d pairs are
// written while level-(d-1) pairs are read: the two sets are disjoint.
template <typename DT>
__global__ __launch_bounds__(256) void k_bc_sigma_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ it_slot, const int32_t* __restrict__ n_items_ptr,
    int nsp, int d, const DT* __restrict__ depth, double* __restrict__ sigma, double* __restrict__ partial,
    const BcCtl* __restrict__ ctl, int guard) {
  __shared__ int32_t s_nb[4][BC_CH];
  if (ctl) {  // the level being added: d + 1
    if (!bc_live(ctl, guard)) return;
    d = ctl->d + 1;
  }
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int n_items = *n_items_ptr;
  const DT dd = (DT)d, dp = (DT)(d - 1);
  int32_t* nb = s_nb[wv];
  for (int it = blockIdx.x * wpb + wv; it < n_items; it += gridDim.x * wpb) {
    const int v = it_v[it];
    const int t0 = it_t0[it];
    const int cnt = min(BC_CH, rp[v + 1] - t0);
    const int slot = it_slot[it];
    for (int t = lane; t < cnt; t += 64) nb[t] = col[t0 + t];
    __threadfence_block();
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      const bool act = depth[vi] == dd;
      if (__ballot(act) == 0ull) continue;
      double acc = 0.0;
      for (int t = 0; t < cnt; t += BC_UNROLL) {
        int uu[BC_UNROLL];
        DT du[BC_UNROLL];
        double su[BC_UNROLL];
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) uu[j] = (t + j < cnt) ? nb[t + j] : -1;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j)
          du[j] = (act && uu[j] >= 0) ? depth[(int64_t)uu[j] * nsp + c + lane] : (DT)0;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j)
          su[j] = (act && uu[j] >= 0 && du[j] == dp) ? sigma[(int64_t)uu[j] * nsp + c + lane] : 0.0;
#pragma unroll
        for (int j = 0; j < BC_UNROLL; ++j) acc += su[j];
      }
      if (act) {
        if (slot >= 0)
          partial[(int64_t)slot * nsp + c + lane] = acc;
        else
          sigma[vi] = acc;
      }
    }
    __threadfence_block();  // the neighbour slots are rewritten by the next item
  }
}

// Bottom-up step of level d -> d+1 (direction-optimising BFS): one wavefront
// per item of a candidate vertex v (a vertex with an edge and an unvisited
// lane); for its unvisited lanes, sigma = the sum of sigma(u, s) over
// neighbours u with depth(u, s) = d, in CSR order -- the same terms in the
// same order as the pull above -- and a lane with a nonzero sum (path counts
// of reached pairs are >= 1) is at depth d+1.  It replaces the discovery push
// from the level-d list and the pull over the level-(d+1) list when the
// candidates have fewer edges than the frontier.  Hub items write per-chunk
// partial sums; k_bc_bottom_up_combine decides their lanes.  Reads of
// depth = d never see a lane written here (those become d+1).
template <typename DT>
__device__ __forceinline__ void bc_bottom_up_body(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ it_v, const int32_t* __restrict__ it_t0,
                                                  const int32_t* __restrict__ it_slot, int n_items, int nsp, int d,
                                                  DT* __restrict__ depth, double* __restrict__ sigma,
                                                  double* __restrict__ partial, uint8_t* __restrict__ next_flag,
                                                  int32_t* nb) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const DT dd = (DT)d, dn = (DT)(d + 1);
  for (int it = blockIdx.x * wpb + wv; it < n_items; it += gridDim.x * wpb) {
    const int v = it_v[it];
    const int t0 = it_t0[it];
    const int cnt = min(BC_CH, rp[v + 1] - t0);
    const int slot = it_slot[it];
    for (int t = lane; t < cnt; t += 64) nb[t] = col[t0 + t];
    __threadfence_block();
    bool reached = false;
    for (int c0 = 0; c0 < nsp; c0 += 64 * BC_CG) {
      bool act[BC_CG];
      bool any = false;
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) {
        const int c = c0 + 64 * g;
        act[g] = c < nsp && depth[(int64_t)v * nsp + c + lane] == bc_unvisited<DT>();
        any |= __ballot(act[g]) != 0ull;
      }
      if (!any) continue;
      double acc[BC_CG];
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) acc[g] = 0.0;
      for (int t = 0; t < cnt; t += BC_NU) {
        int uu[BC_NU];
        DT du[BC_NU][BC_CG];
        double su[BC_NU][BC_CG];
#pragma unroll
        for (int j = 0; j < BC_NU; ++j) uu[j] = (t + j < cnt) ? nb[t + j] : -1;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g)
            du[j][g] = (act[g] && uu[j] >= 0) ? depth[(int64_t)uu[j] * nsp + c0 + 64 * g + lane] : (DT)0;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g)
            su[j][g] = (act[g] && uu[j] >= 0 && du[j][g] == dd)
                           ? sigma[(int64_t)uu[j] * nsp + c0 + 64 * g + lane]
                           : 0.0;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g) acc[g] += su[j][g];
      }
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) {
        const int64_t vi = (int64_t)v * nsp + c0 + 64 * g + lane;
        if (slot >= 0) {
          if (act[g]) partial[(int64_t)slot * nsp + c0 + 64 * g + lane] = acc[g];
        } else if (act[g] && acc[g] > 0.0) {
          depth[vi] = dn;
          sigma[vi] = acc[g];
          reached = true;
        }
      }
    }
    if (__ballot(reached) != 0ull && lane == 0) next_flag[v] = 1;
    __threadfence_block();  // the neighbour slots are rewritten by the next item
  }
}

template <typename DT>
__global__ __launch_bounds__(256) void k_bc_bottom_up_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ it_slot, const int32_t* __restrict__ n_items_ptr,
    int nsp, int d, DT* __restrict__ depth, double* __restrict__ sigma, double* __restrict__ partial,
    uint8_t* __restrict__ next_flag, const BcCtl* __restrict__ ctl, int guard) {
  __shared__ int32_t s_nb[4][BC_CH];
  if (ctl) {
    if (!bc_live(ctl, guard)) return;
    d = ctl->d;
  }
  bc_bottom_up_body<DT>(rp, col, it_v, it_t0, it_slot, *n_items_ptr, nsp, d, depth, sigma, partial, next_flag,
                        s_nb[threadIdx.x >> 6]);
}

// Device-driven forward step, first sweep: the bottom-up sweep of the
// candidates or the top-down discovery from the level-d list, whichever
// direction the step has taken (one launch instead of one per direction).
template <typename DT>
__global__ __launch_bounds__(256) void k_bc_forward_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ cit_v, const int32_t* __restrict__ cit_t0,
    const int32_t* __restrict__ cit_slot, int nsp, DT* __restrict__ depth, double* __restrict__ sigma,
    double* __restrict__ partial, uint8_t* __restrict__ next_flag, const BcCtl* __restrict__ ctl) {
  __shared__ int32_t s_nb[4][BC_CH];
  if (ctl->done) return;
  if (ctl->bu)
    bc_bottom_up_body<DT>(rp, col, cit_v, cit_t0, cit_slot, ctl->cand_tot[1], nsp, ctl->d, depth, sigma, partial,
                          next_flag, s_nb[threadIdx.x >> 6]);
  else
    bc_discover_body<DT>(rp, col, it_v, it_t0, ctl->next_tot[1], nsp, ctl->d, depth, next_flag,
                         s_nb[threadIdx.x >> 6]);
}

// the lanes of the candidate hubs: partial sums in chunk order, then as above
template <typename DT>
__global__ __launch_bounds__(256) void k_bc_bottom_up_combine(const int32_t* __restrict__ cand, int nc,
                                                              const int32_t* __restrict__ rp,
                                                              const int32_t* __restrict__ hstart, int nsp, int d,
                                                              DT* __restrict__ depth, double* __restrict__ sigma,
                                                              const double* __restrict__ partial,
                                                              uint8_t* __restrict__ next_flag,
                                                              const int32_t* __restrict__ hpos,
                                                              const BcCtl* __restrict__ ctl, int guard) {
  if (ctl) {  // the candidate hubs only (hpos: their list positions)
    if (!bc_live(ctl, guard)) return;
    nc = ctl->cand_tot[3];
    d = ctl->d;
  }
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const DT dn = (DT)(d + 1);
  for (int j = blockIdx.x * wpb + (threadIdx.x >> 6); j < nc; j += gridDim.x * wpb) {
    const int f = ctl ? hpos[j] : j;
    const int v = cand[f];
    const int deg = rp[v + 1] - rp[v];
    if (deg <= BC_CH) continue;  // single-item vertex: finished in the item pass
    const int nchunk = (deg + BC_CH - 1) / BC_CH;
    const int slot0 = hstart[f];
    bool reached = false;
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] == bc_unvisited<DT>()) {
        double acc = 0.0;
        for (int j = 0; j < nchunk; ++j) acc += partial[(int64_t)(slot0 + j) * nsp + c + lane];
        if (acc > 0.0) {
          depth[vi] = dn;
          sigma[vi] = acc;
          reached = true;
        }
      }
    }
    if (__ballot(reached) != 0ull && lane == 0) next_flag[v] = 1;
  }
}

// flag[v] = 1 when vertex v has an edge and an unvisited lane (the
// candidates of a bottom-up step)
// (with ctl every flag is written, 0 or 1, so the array needs no clearing)
template <typename DT>
__global__ __launch_bounds__(256) void k_unvisited_flags(const int32_t* __restrict__ rp, const DT* __restrict__ depth,
                                                         int64_t n, int nsp, uint8_t* __restrict__ flag,
                                                         const BcCtl* __restrict__ ctl, int guard) {
  if (ctl && !bc_live(ctl, guard)) return;
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += (int64_t)gridDim.x * wpb) {
    bool hit = false;
    if (rp[v + 1] != rp[v])
      for (int c = lane; c < nsp; c += 64) hit |= depth[v * nsp + c] == bc_unvisited<DT>();
    const bool any = __ballot(hit) != 0ull;
    if (lane == 0 && (any || ctl)) flag[v] = any ? 1 : 0;
  }
}

// out[0] += the edges of list a (na entries), out[1] += those of list b
// (*nb_ptr entries, a device count)
__global__ void k_list_edges(const int32_t* __restrict__ a, int na, const int32_t* __restrict__ b,
                             const int32_t* __restrict__ nb_ptr, const int32_t* __restrict__ rp,
                             unsigned long long* __restrict__ out, const BcCtl* __restrict__ ctl, int guard) {
  if (ctl) {  // list a: the level-d list
    if (!bc_live(ctl, guard)) return;
    a = ctl->cur;
    na = ctl->nf;
  }
  const int nb = *nb_ptr;
  unsigned long long ea = 0, eb = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < max(na, nb); i += gridDim.x * blockDim.x) {
    if (i < na) ea += (unsigned long long)(rp[a[i] + 1] - rp[a[i]]);
    if (i < nb) eb += (unsigned long long)(rp[b[i] + 1] - rp[b[i]]);
  }
  ea = wave_reduce_sum(ea);
  eb = wave_reduce_sum(eb);
  if ((threadIdx.x & 63) == 0) {
    if (ea) atomicAdd(&out[0], ea);
    if (eb) atomicAdd(&out[1], eb);
  }
}

// sigma of the hub vertices of level d: their per-chunk partial sums in chunk order
template <typename DT>
__global__ __launch_bounds__(256) void k_bc_sigma_combine(const int32_t* __restrict__ front, int nf,
                                                          const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ hstart, int nsp, int d,
                                                          const DT* __restrict__ depth, double* __restrict__ sigma,
                                                          const double* __restrict__ partial,
                                                          const int32_t* __restrict__ hpos,
                                                          const BcCtl* __restrict__ ctl, int guard) {
  if (ctl) {  // the hubs of the list being added (level d + 1)
    if (!bc_live(ctl, guard)) return;
    front = ctl->next;
    nf = ctl->next_tot[3];
    d = ctl->d + 1;
  }
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const DT dd = (DT)d;
  for (int j = blockIdx.x * wpb + (threadIdx.x >> 6); j < nf; j += gridDim.x * wpb) {
    const int f = ctl ? hpos[j] : j;
    const int v = front[f];
    const int deg = rp[v + 1] - rp[v];
    if (deg <= BC_CH) continue;  // single-item vertex: finished in the item pass
    const int nc = (deg + BC_CH - 1) / BC_CH;
    const int slot0 = hstart[f];
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] == dd) {
        double acc = 0.0;
        for (int j = 0; j < nc; ++j) acc += partial[(int64_t)(slot0 + j) * nsp + c + lane];
        sigma[vi] = acc;
      }
    }
  }
}

// backward, level k: one wavefront per item; successors summed in CSR order
// within the item (a non-successor adds +0.0, which leaves the sum unchanged).
// Single-item vertices finish delta, q and bc here; the items of a hub vertex
// write per-chunk partial sums that k_bc_combine adds in chunk order (a fixed
// order: results are reproducible run to run).
template <typename DT>
__global__ __launch_bounds__(256) void k_bc_backward_items(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const int32_t* __restrict__ it_v,
    const int32_t* __restrict__ it_t0, const int32_t* __restrict__ it_slot, const int32_t* __restrict__ n_items_ptr,
    int nsp, int k, const DT* __restrict__ depth, const double* __restrict__ sigma, double* __restrict__ q,
    double* __restrict__ bc, double* __restrict__ partial, int64_t bc_stride, const BcCtl* __restrict__ ctl) {
  __shared__ int32_t s_nb[4][BC_CH];
  if (ctl) k = ctl->bk;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int n_items = *n_items_ptr;
  const DT dk = (DT)k, dn = (DT)(k + 1);
  int32_t* nb = s_nb[wv];
  for (int it = blockIdx.x * wpb + wv; it < n_items; it += gridDim.x * wpb) {
    const int v = it_v[it];
    const int t0 = it_t0[it];
    const int cnt = min(BC_CH, rp[v + 1] - t0);
    const int slot = it_slot[it];
    for (int t = lane; t < cnt; t += 64) nb[t] = col[t0 + t];
    __threadfence_block();
    double part = 0.0;
    for (int c0 = 0; c0 < nsp; c0 += 64 * BC_CG) {
      bool act[BC_CG];
      unsigned live = 0u;  // chunks with an active lane (wavefront-uniform)
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) {
        const int c = c0 + 64 * g;
        act[g] = c < nsp && depth[(int64_t)v * nsp + c + lane] == dk;
        if (__ballot(act[g]) != 0ull) live |= 1u << g;
      }
      if (live == 0u) continue;
      double acc[BC_CG];
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) acc[g] = 0.0;
      for (int t = 0; t < cnt; t += BC_NU) {
        int ww[BC_NU];
        DT dw[BC_NU][BC_CG];
        double qw[BC_NU][BC_CG];
#pragma unroll
        for (int j = 0; j < BC_NU; ++j) ww[j] = (t + j < cnt) ? nb[t + j] : -1;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g)
            dw[j][g] = (act[g] && ww[j] >= 0) ? depth[(int64_t)ww[j] * nsp + c0 + 64 * g + lane] : (DT)0;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g)
            qw[j][g] = (act[g] && ww[j] >= 0 && dw[j][g] == dn) ? q[(int64_t)ww[j] * nsp + c0 + 64 * g + lane] : 0.0;
#pragma unroll
        for (int j = 0; j < BC_NU; ++j)
#pragma unroll
          for (int g = 0; g < BC_CG; ++g) acc[g] += qw[j][g];
      }
      // per chunk, in chunk order, as a chunk-at-a-time walk would finish them
#pragma unroll
      for (int g = 0; g < BC_CG; ++g) {
        if (!((live >> g) & 1u)) continue;
        const int c = c0 + 64 * g;
        const int64_t vi = (int64_t)v * nsp + c + lane;
        if (slot >= 0) {
          if (act[g]) partial[(int64_t)slot * nsp + c + lane] = acc[g];
        } else {
          if (act[g]) {
            const double sg = sigma[vi];
            const double delta = sg * acc[g];
            q[vi] = (1.0 + delta) / sg;
            part += delta;
          }
          if (bc_stride) {  // one score vector per 64-lane group
            part = wave_reduce_sum(part);
            if (lane == 0) bc[(c >> 6) * bc_stride + v] += part;
            part = 0.0;
          }
        }
      }
    }
    if (slot < 0 && !bc_stride) {
      part = wave_reduce_sum(part);
      if (lane == 0) bc[v] += part;
    }
    __threadfence_block();  // the neighbour slots are rewritten by the next item
  }
}

template <typename DT>
__global__ __launch_bounds__(256) void k_bc_combine(const int32_t* __restrict__ front, int nf,
                                                    const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ hstart, int nsp, int k,
                                                    const DT* __restrict__ depth,
                                                    const double* __restrict__ sigma, double* __restrict__ q,
                                                    double* __restrict__ bc, const double* __restrict__ partial,
                                                    int64_t bc_stride, const int32_t* __restrict__ hpos,
                                                    const BcCtl* __restrict__ ctl) {
  if (ctl) {  // the hubs of the backward level's list (built into `front`)
    nf = ctl->back_tot[3];
    k = ctl->bk;
  }
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const DT dk = (DT)k;
  for (int j = blockIdx.x * wpb + (threadIdx.x >> 6); j < nf; j += gridDim.x * wpb) {
    const int f = ctl ? hpos[j] : j;
    const int v = front[f];
    const int deg = rp[v + 1] - rp[v];
    if (deg <= BC_CH) continue;  // single-item vertex: finished in the item pass
    const int nc = (deg + BC_CH - 1) / BC_CH;
    const int slot0 = hstart[f];
    double part = 0.0;
    for (int c = 0; c < nsp; c += 64) {
      const int64_t vi = (int64_t)v * nsp + c + lane;
      if (depth[vi] == dk) {
        double acc = 0.0;
        for (int j = 0; j < nc; ++j) acc += partial[(int64_t)(slot0 + j) * nsp + c + lane];
        const double sg = sigma[vi];
        const double delta = sg * acc;
        q[vi] = (1.0 + delta) / sg;
        part += delta;
      }
      if (bc_stride) {  // one score vector per 64-lane group
        part = wave_reduce_sum(part);
        if (lane == 0) bc[(c >> 6) * bc_stride + v] += part;
        part = 0.0;
      }
    }
    if (!bc_stride) {
      part = wave_reduce_sum(part);
      if (lane == 0) bc[v] += part;
    }
  }
}

// item list of a frontier (device scans; the item count stays on the device);
// the buffers are sized for nf = n by bc_batch_dev
static void bc_items(const DCsr& A, const int32_t* front, int nf, BcWork& w, hipStream_t s) {
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
template <typename DT>
__global__ void k_bc_account(const int32_t* __restrict__ rp, const DT* __restrict__ depth, int64_t n, int nsp,
                             unsigned long long* __restrict__ prof) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  unsigned long long wedges = 0, pairs = 0;
  for (int64_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += (int64_t)gridDim.x * wpb) {
    int cnt = 0;
    for (int c = lane; c < nsp; c += 64) cnt += depth[v * nsp + c] != bc_unvisited<DT>();
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

// ---- level lists sized on the device ------------------------------------
// Flags -> ascending vertex list + its neighbour-chunk items + hub slots, in
// two launches whose grids depend only on n.  A group is 64 consecutive
// vertices, one per lane of a wavefront; each workgroup owns a contiguous
// range of groups (its four wavefronts a quarter each, walked in order).
// The count pass writes every group's vertices, items and hub slots and its
// workgroup's sums; the emit pass adds the sums of the workgroups before its
// own and of its earlier wavefronts, then walks its groups placing each
// lane's vertex by a ballot count and its items and hub slots by wave scans:
// the list (hstart[f] = first hub slot of list vertex f), the items (same
// layout as bc_items), the list positions of its hub vertices (more than
// BC_CH neighbours, the only ones the combine passes touch) and, from the
// last wavefront, the four totals.  Flags are cleared.
constexpr int BCB_T = 256;
constexpr int BCB_MAX_BLOCKS = 1024;
__device__ inline int64_t ceil_div_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct BcGroupRange {
  int64_t g0, g1;  // the wavefront's groups
};
__device__ inline BcGroupRange bc_wave_groups(int64_t ng, int64_t per_block) {
  const int64_t b0 = (int64_t)blockIdx.x * per_block, b1 = min(ng, b0 + per_block);
  const int64_t per_wave = ceil_div_dev(per_block, BCB_T / 64);
  const int64_t w0 = b0 + (int64_t)(threadIdx.x >> 6) * per_wave;
  return {min(w0, b1), min(w0 + per_wave, b1)};
}
// lane's vertex of group g: flagged?, its items, its hub slots
__device__ inline void bc_group_lane(const uint8_t* flag, const int32_t* rp, int64_t n, int64_t g, bool& f, int& nc,
                                     int& hub) {  // hub: hub slots of the lane's vertex
  const int64_t v = g * BC_FLAG_CHUNK + (threadIdx.x & 63);
  f = v < n && flag[v] != 0;
  nc = 0;
  if (f) {
    const int deg = rp[v + 1] - rp[v];
    nc = deg > BC_CH ? (deg + BC_CH - 1) / BC_CH : 1;
  }
  hub = nc > 1 ? nc : 0;
}

__global__ __launch_bounds__(BCB_T) void k_bc_build_count(const uint8_t* __restrict__ flag, int64_t n,
                                                          const int32_t* __restrict__ rp, int32_t* __restrict__ gcnt,
                                                          long long* __restrict__ bsum, int64_t per_block,
                                                          const BcCtl* __restrict__ ctl, int guard) {
  if (ctl && !bc_live(ctl, guard)) return;
  __shared__ long long red[4][BCB_T / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t ng = ceil_div_dev(n, BC_FLAG_CHUNK);
  const BcGroupRange r = bc_wave_groups(ng, per_block);
  long long t[4] = {0, 0, 0, 0};
  for (int64_t g = r.g0; g < r.g1; ++g) {
    bool f;
    int nc, hub;
    bc_group_lane(flag, rp, n, g, f, nc, hub);
    const int cv = __popcll(__ballot(f)), cn = __popcll(__ballot(hub > 0));
    const int ci = wave_reduce_sum(nc), ch = wave_reduce_sum(hub);
    if (lane == 0) {
      gcnt[g] = cv;
      gcnt[ng + g] = ci;
      gcnt[2 * ng + g] = ch;
      gcnt[3 * ng + g] = cn;
    }
    t[0] += cv;
    t[1] += ci;
    t[2] += ch;
    t[3] += cn;
  }
  if (lane == 0)
    for (int k = 0; k < 4; ++k) red[k][wv] = t[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    long long u = 0;
    for (int w = 0; w < BCB_T / 64; ++w) u += red[threadIdx.x][w];
    bsum[4 * (int64_t)blockIdx.x + threadIdx.x] = u;
  }
}

// to_next: the list goes to ctl->next (the level being built); otherwise to `list`
__global__ __launch_bounds__(BCB_T) void k_bc_build_emit(uint8_t* __restrict__ flag, int64_t n,
                                                         const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ gcnt,
                                                         const long long* __restrict__ bsum, int64_t per_block,
                                                         int32_t* list, int32_t* __restrict__ hstart,
                                                         int32_t* __restrict__ it_v, int32_t* __restrict__ it_t0,
                                                         int32_t* __restrict__ it_slot, int32_t* __restrict__ hpos,
                                                         int32_t* __restrict__ tot, const BcCtl* __restrict__ ctl,
                                                         int guard, int to_next) {
  if (ctl && !bc_live(ctl, guard)) return;
  if (ctl && to_next) list = ctl->next;
  __shared__ long long prior[4][BCB_T / 64];
  __shared__ long long wtot[4][BCB_T / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t ng = ceil_div_dev(n, BC_FLAG_CHUNK);
  const BcGroupRange r = bc_wave_groups(ng, per_block);
  long long p[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += BCB_T)
    for (int k = 0; k < 4; ++k) p[k] += bsum[4 * (int64_t)b + k];
  for (int64_t g = r.g0 + lane; g < r.g1; g += 64)
    for (int k = 0; k < 4; ++k) q[k] += gcnt[k * ng + g];
  for (int k = 0; k < 4; ++k) {
    p[k] = wave_reduce_sum(p[k]);
    q[k] = wave_reduce_sum(q[k]);
    if (lane == 0) {
      prior[k][wv] = p[k];
      wtot[k][wv] = q[k];
    }
  }
  __syncthreads();
  long long run[4];
  for (int k = 0; k < 4; ++k) {
    run[k] = 0;
    for (int w = 0; w < BCB_T / 64; ++w) {
      run[k] += prior[k][w];
      if (w < wv) run[k] += wtot[k][w];
    }
  }
  for (int64_t g = r.g0; g < r.g1; ++g) {
    bool f;
    int nc, hub;
    bc_group_lane(flag, rp, n, g, f, nc, hub);
    const unsigned long long m = __ballot(f), mh = __ballot(hub > 0);
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const int hbelow =
        (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mh >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mh, 0u));
    const int ci = wave_incl_scan(nc), ch = wave_incl_scan(hub);
    if (f) {
      const int64_t v = g * BC_FLAG_CHUNK + lane;
      const int ov = (int)run[0] + below, oi = (int)run[1] + ci - nc, oh = (int)run[2] + ch - hub;
      const int r0 = rp[v];
      flag[v] = 0;
      list[ov] = (int32_t)v;
      hstart[ov] = oh;
      if (hub > 0) hpos[(int)run[3] + hbelow] = ov;
      for (int j = 0; j < nc; ++j) {
        it_v[oi + j] = (int32_t)v;
        it_t0[oi + j] = r0 + j * BC_CH;
        it_slot[oi + j] = nc > 1 ? oh + j : -1;
      }
    }
    run[0] += __popcll(m);
    run[1] += __shfl(ci, 63, 64);
    run[2] += __shfl(ch, 63, 64);
    run[3] += __popcll(mh);
  }
  if (blockIdx.x == gridDim.x - 1 && wv == BCB_T / 64 - 1 && lane == 0)
    for (int k = 0; k < 4; ++k) tot[k] = (int32_t)run[k];
}

// One thread steers the sweep between the launches of a step.
enum { BC_OP_INIT = 0, BC_OP_START = 1, BC_OP_DECIDE = 2, BC_OP_NEXT = 3, BC_OP_END = 4, BC_OP_BACK = 5 };
__global__ void k_bc_ctl(BcCtl* __restrict__ c, int op, int32_t* l0, int32_t* l1, int allow_bu, double alpha,
                         int limit, long long* __restrict__ host) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  switch (op) {
    case BC_OP_INIT:  // before the source list is built into `next`
      c->cur = l1;
      c->next = l0;
      c->d = c->nf = c->done = c->dmax = c->bu = c->over = 0;
      c->allow_bu = allow_bu;
      c->alpha = alpha;
      c->bk = 0;
      c->edges[0] = c->edges[1] = 0ull;
      c->seq = 0;
      break;
    case BC_OP_START: {  // level 0 = the source list
      int32_t* t = c->cur;
      c->cur = c->next;
      c->next = t;
      c->nf = c->next_tot[0];
      if (c->nf == 0) c->done = 1;
      break;
    }
    case BC_OP_DECIDE:  // after the candidates and the edges of both lists
      if (!c->done && c->d >= 1 && !c->bu && c->allow_bu)
        c->bu = (double)c->edges[0] * c->alpha > (double)c->edges[1] ? 1 : 0;
      c->edges[0] = c->edges[1] = 0ull;
      if (!c->done && c->bu && c->cand_tot[0] == 0) {  // no vertex can be reached any more
        c->done = 1;
        c->dmax = c->d;
      }
      break;
    case BC_OP_NEXT:  // after the level-(d+1) list
      if (!c->done && c->next_tot[0] == 0) {
        c->done = 1;
        c->dmax = c->d;
      }
      break;
    case BC_OP_END:  // level d + 1 complete
      if (!c->done) {
        int32_t* t = c->cur;
        c->cur = c->next;
        c->next = t;
        c->nf = c->next_tot[0];
        c->d += 1;
        if (c->d + 1 >= limit) {  // the next level would reach the unvisited marker
          c->over = 1;
          c->done = 1;
          c->dmax = c->d;
        }
      }
      c->seq += 1;
      if (host) {  // the last word (released after the others): 2 x steps + done
        __hip_atomic_store(&host[0], (long long)c->d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&host[1], (long long)c->dmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&host[2], (long long)c->over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&host[3], 2 * c->seq + (c->done ? 1 : 0), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      break;
    case BC_OP_BACK:  // a backward level done: the next one up (the first is set by the host)
      c->bk = limit > 0 ? limit : c->bk - 1;
      break;
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

// the flagged vertices as an ascending list (flags cleared); its size goes to
// count_h (pinned) and, on the device, to w.off[ceil(n / BC_FLAG_CHUNK)]
static void frontier_from_flags(uint8_t* flag, int64_t n, BcWork& w, int32_t* out, int32_t* count_h,
                                hipStream_t s, bool sync = true) {
  const int chunk = BC_FLAG_CHUNK;
  const int64_t ng = ceil_div(n, chunk);
  hipLaunchKernelGGL(k_flag_count, dim3(grid_for(ng, 256)), dim3(256), 0, s, flag, n, w.cnt.as<int32_t>(), chunk);
  scan_exclusive_i32(w.cnt.as<int32_t>(), w.off.as<int32_t>(), ng, w.scan_tmp, s);
  hipLaunchKernelGGL(k_flag_emit, dim3(grid_for(ng, 256)), dim3(256), 0, s, flag, n,
                     (const int32_t*)w.off.as<int32_t>(), chunk, out);
  GA_HIP(hipMemcpyAsync(count_h, w.off.as<int32_t>() + ng, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (sync) GA_HIP(hipStreamSynchronize(s));
}

// Sizes of every BcWork buffer for an n-vertex graph with nnz entries, ns
// sources on nsp lanes and `outs` score vectors.  Two frontier lists of at
// most n vertices each (the backward sweep rebuilds its lists from depths).
struct BcSizes {
  int64_t nsp, total, item_cap, partial, levels, ng, outs;
};
static BcSizes bc_sizes(int64_t n, int64_t nnz, int64_t ns, int64_t nsp, int64_t outs) {
  BcSizes z;
  z.nsp = nsp;
  z.total = n * z.nsp;
  z.item_cap = n + ceil_div(nnz, BC_CH) + 1;
  z.partial = (2 * ceil_div(nnz, BC_CH) + 1) * z.nsp;
  z.levels = 3 * std::max<int64_t>(n, 1);  // two frontier lists + the bottom-up candidates
  z.ng = ceil_div(std::max<int64_t>(n, 1), BC_FLAG_CHUNK);
  z.outs = outs;
  return z;
}

static void bc_reserve(BcWork& w, int64_t n, const BcSizes& z, int64_t ns) {
  const int64_t n1 = std::max<int64_t>(n, 1);
  w.depth.ensure(sizeof(uint16_t) * z.total);  // room for either depth width
  w.sigma.ensure(sizeof(double) * z.total);
  w.q.ensure(sizeof(double) * z.total);
  w.bc.ensure(sizeof(double) * n1 * z.outs);
  w.flag.ensure(4 * ceil_div(n1, 4));
  w.src.ensure(sizeof(int32_t) * 2 * ns);  // vertices, then lanes
  w.it_v.ensure(sizeof(int32_t) * z.item_cap);
  w.it_t0.ensure(sizeof(int32_t) * z.item_cap);
  w.it_slot.ensure(sizeof(int32_t) * z.item_cap);
  w.partial.ensure(sizeof(double) * z.partial);
  w.levels.ensure(sizeof(int32_t) * z.levels);
  w.dsum.ensure(sizeof(unsigned long long) * 2);
  w.cnt.ensure(sizeof(int32_t) * z.ng);
  w.off.ensure(sizeof(int32_t) * (z.ng + 1));
  w.icnt.ensure(sizeof(int32_t) * n1);
  w.hcnt.ensure(sizeof(int32_t) * n1);
  w.istart.ensure(sizeof(int32_t) * (n1 + 1));
  w.hstart.ensure(sizeof(int32_t) * (n1 + 1));
  // largest scan: a level list (at most n entries) or the flag groups
  w.scan_tmp.ensure(scan_tmp_bytes(std::max(n1, z.ng)));
  // device-driven passes
  w.cflag.ensure(4 * ceil_div(n1, 4));
  w.cit_v.ensure(sizeof(int32_t) * z.item_cap);
  w.cit_t0.ensure(sizeof(int32_t) * z.item_cap);
  w.cit_slot.ensure(sizeof(int32_t) * z.item_cap);
  w.chstart.ensure(sizeof(int32_t) * (n1 + 1));
  w.gcnt.ensure(sizeof(int32_t) * 4 * z.ng);
  w.gsum.ensure(sizeof(long long) * 4 * BCB_MAX_BLOCKS);
  w.hpos.ensure(sizeof(int32_t) * n1);
  w.chpos.ensure(sizeof(int32_t) * n1);
  w.ctl.ensure(sizeof(BcCtl));
}

static size_t bc_bytes(int64_t n, const BcSizes& z, int64_t ns) {
  const int64_t n1 = std::max<int64_t>(n, 1);
  size_t b = lease_bytes(sizeof(uint16_t) * z.total) + 2 * lease_bytes(sizeof(double) * z.total);
  b += lease_bytes(sizeof(double) * n1 * z.outs) + lease_bytes(4 * ceil_div(n1, 4)) +
       lease_bytes(sizeof(int32_t) * 2 * ns);
  b += 3 * lease_bytes(sizeof(int32_t) * z.item_cap) + lease_bytes(sizeof(double) * z.partial);
  b += lease_bytes(sizeof(int32_t) * z.levels) + lease_bytes(sizeof(unsigned long long) * 2);
  b += lease_bytes(sizeof(int32_t) * z.ng) + lease_bytes(sizeof(int32_t) * (z.ng + 1));
  b += 2 * lease_bytes(sizeof(int32_t) * n1) + 2 * lease_bytes(sizeof(int32_t) * (n1 + 1));
  b += lease_bytes(scan_tmp_bytes(std::max(n1, z.ng)));
  b += lease_bytes(4 * ceil_div(n1, 4)) + 3 * lease_bytes(sizeof(int32_t) * z.item_cap) +
       lease_bytes(sizeof(int32_t) * (n1 + 1)) + lease_bytes(sizeof(int32_t) * 4 * z.ng) +
       lease_bytes(sizeof(long long) * 4 * BCB_MAX_BLOCKS) + lease_bytes(sizeof(BcCtl)) +
       2 * lease_bytes(sizeof(int32_t) * n1);
  return b;
}

size_t bc_workspace_bytes(int64_t n, int64_t nnz, int64_t ns) {
  return bc_bytes(n, bc_sizes(n, nnz, ns, ceil_div(ns, 64) * 64, 1), ns);
}

size_t bc_groups_workspace_bytes(int64_t n, int64_t nnz, const std::vector<int64_t>& group_sizes) {
  int64_t ns = 0;
  for (int64_t g : group_sizes) ns += g;
  const int64_t G = (int64_t)group_sizes.size();
  return bc_bytes(n, bc_sizes(n, nnz, ns, 64 * G, G), ns);
}

// Runs the multi-source pass.  Source j is vertex src_h[j] on lane lane_h[j]
// (distinct lanes; nsp lanes in all).  outs == 1: bc_h[v] sums every lane;
// outs == nsp / 64: lane group g (lanes 64g .. 64g+63) sums into bc_h[g*n + v].
template <typename DT>
static bool bc_run(const DCsr& A, const int32_t* src_h, const int32_t* lane_h, int64_t ns, int64_t nsp, int64_t outs,
                   double* bc_h, BcWork& w, hipStream_t s);
template <typename DT>
static bool bc_run_dev(const DCsr& A, const int32_t* src_h, const int32_t* lane_h, int64_t ns, int64_t nsp_,
                       int64_t outs, double* bc_h, BcWork& w, hipStream_t s, bool replay);
static void bc_run_any(const DCsr& A, const int32_t* src_h, const int32_t* lane_h, int64_t ns, int64_t nsp,
                       int64_t outs, double* bc_h, BcWork& w, hipStream_t s) {
  const int mode = ctx().bc_loop;
  if (mode >= 1 && A.nrows > 0) {
    if (!bc_run_dev<uint8_t>(A, src_h, lane_h, ns, nsp, outs, bc_h, w, s, mode == 2))
      (void)bc_run_dev<uint16_t>(A, src_h, lane_h, ns, nsp, outs, bc_h, w, s, mode == 2);
    return;
  }
  if (!bc_run<uint8_t>(A, src_h, lane_h, ns, nsp, outs, bc_h, w, s))
    (void)bc_run<uint16_t>(A, src_h, lane_h, ns, nsp, outs, bc_h, w, s);
}

// A: device pattern (out-edges); src: host array of distinct sources, on
// lanes 0..ns-1.
void bc_batch_dev(const DCsr& A, const int32_t* src_h, int64_t ns, double* bc_h, BcWork& w, hipStream_t s) {
  const int64_t n = A.nrows;
  GA_REQUIRE(A.nrows == A.ncols, GA_ERR_DIM, "betweenness centrality needs a square matrix");
  GA_REQUIRE(ns >= 1 && ns <= 4096, GA_ERR_LIMIT, "between 1 and 4096 sources per batch");
  std::vector<char> seen(n, 0);
  std::vector<int32_t> lane(ns);
  for (int64_t j = 0; j < ns; ++j) {
    GA_REQUIRE(src_h[j] >= 0 && src_h[j] < n, GA_ERR_ARG, "source out of range");
    GA_REQUIRE(!seen[src_h[j]], GA_ERR_ARG, "duplicate source");
    seen[src_h[j]] = 1;
    lane[j] = (int32_t)j;
  }
  bc_run_any(A, src_h, lane.data(), ns, ceil_div(ns, 64) * 64, 1, bc_h, w, s);
}

// Several independent source sets in one pass: set g (at most 64 distinct
// sources) takes lane group g and its scores go to bc_out[g] (n doubles).
// Every (vertex, lane) value and every per-vertex sum is formed exactly as
// in a pass of that set alone, so the scores are bit-identical to it.
void bc_groups_dev(const DCsr& A, const std::vector<std::vector<int32_t>>& sets, double* const* bc_out, BcWork& w,
                   hipStream_t s) {
  const int64_t n = A.nrows;
  GA_REQUIRE(A.nrows == A.ncols, GA_ERR_DIM, "betweenness centrality needs a square matrix");
  const int64_t G = (int64_t)sets.size();
  GA_REQUIRE(G >= 1 && G <= 64, GA_ERR_LIMIT, "between 1 and 64 source sets per pass");
  std::vector<int32_t> src, lane;
  for (int64_t g = 0; g < G; ++g) {
    GA_REQUIRE(!sets[g].empty() && sets[g].size() <= 64, GA_ERR_LIMIT, "1 to 64 sources per set");
    std::vector<char> seen(n, 0);
    for (size_t i = 0; i < sets[g].size(); ++i) {
      const int32_t v = sets[g][i];
      GA_REQUIRE(v >= 0 && v < n, GA_ERR_ARG, "source out of range");
      GA_REQUIRE(!seen[v], GA_ERR_ARG, "duplicate source");
      seen[v] = 1;
      src.push_back(v);
      lane.push_back((int32_t)(64 * g + i));
    }
  }
  std::vector<double> bc((size_t)(G * std::max<int64_t>(n, 1)));
  bc_run_any(A, src.data(), lane.data(), (int64_t)src.size(), 64 * G, G, bc.data(), w, s);
  for (int64_t g = 0; g < G; ++g)
    if (n) std::memcpy(bc_out[g], bc.data() + g * n, sizeof(double) * n);
}

// false: the BFS needs more levels than DT can hold (the pass is abandoned)
template <typename DT>
static bool bc_run(const DCsr& A, const int32_t* src_h, const int32_t* lane_h, int64_t ns, int64_t nsp_, int64_t outs,
                   double* bc_h, BcWork& w, hipStream_t s) {
  const int64_t n = A.nrows;
  const BcSizes z = bc_sizes(n, A.nnz, ns, nsp_, outs);
  const int nsp = (int)z.nsp;
  const int64_t total = z.total;
  const int64_t bc_stride = outs > 1 ? n : 0;
  bc_reserve(w, n, z, ns);
  GA_HIP(hipMemcpyAsync(w.src.p, src_h, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  GA_HIP(hipMemcpyAsync(w.src.as<int32_t>() + ns, lane_h, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_bc_init<DT>, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, w.depth.as<DT>(),
                     w.sigma.as<double>(), total);
  fill_u64(w.bc.p, 0ull, n * outs, s);
  fill_u32(w.flag.p, 0u, ceil_div(std::max<int64_t>(n, 1), 4), s);
  hipLaunchKernelGGL(k_bc_sources<DT>, dim3(grid_for(ns, 256)), dim3(256), 0, s, (const int32_t*)w.src.as<int32_t>(),
                     (const int32_t*)(w.src.as<int32_t>() + ns), (int)ns, nsp, w.depth.as<DT>(),
                     w.sigma.as<double>(), w.flag.as<uint8_t>());
  GA_HIP(hipGetLastError());
  int32_t* count_h = static_cast<int32_t*>(pinned_scratch());
  // forward frontier lists ping-pong between the two halves of w.levels;
  // level 0 = the source vertices
  int32_t* lists[2] = {w.levels.as<int32_t>(), w.levels.as<int32_t>() + std::max<int64_t>(n, 1)};
  frontier_from_flags(w.flag.as<uint8_t>(), n, w, lists[0], count_h, s);
  int32_t nf = *count_h;
  int d = 0;
  Context& cx = ctx();
  // event pairs around the device work of each step (the host waits for the
  // size of every new level list between them)
  hipEvent_t ev0 = nullptr;
  auto ev_open = [&] {
    if (!cx.prof_events_on) return;
    ev0 = bc_event(cx);
    GA_HIP(hipEventRecord(ev0, s));
  };
  auto ev_close = [&] {
    if (!cx.prof_events_on) return;
    hipEvent_t ev1 = bc_event(cx);
    GA_HIP(hipEventRecord(ev1, s));
    cx.prof_events.emplace_back(ev0, ev1);
    cx.prof_launches += 1;
  };
  int64_t timed = 0;  // event pairs of this pass
  // direction-optimising forward: top-down (discovery push + sigma pull)
  // while the level list has at most 1/alpha of the edges of the vertices
  // with unvisited lanes, bottom-up from then on (the candidates only shrink)
  int32_t* cand = w.levels.as<int32_t>() + 2 * std::max<int64_t>(n, 1);
  auto* edges_h = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(count_h) + 64);
  const int64_t ng = ceil_div(std::max<int64_t>(n, 1), BC_FLAG_CHUNK);
  bool bottom_up = false;
  int32_t nc = 0;
  // candidates of a bottom-up step into `cand` (nc entries); with `costs`,
  // also the edges of the level list and of the candidates (edges_h[0..1])
  auto candidates = [&](const int32_t* front, int32_t nfront, bool costs) {
    ev_open();
    hipLaunchKernelGGL(k_unvisited_flags<DT>, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.rp,
                       (const DT*)w.depth.as<DT>(), n, nsp, w.flag.as<uint8_t>(), (const BcCtl*)nullptr, 0);
    frontier_from_flags(w.flag.as<uint8_t>(), n, w, cand, count_h, s, false);
    if (costs) {
      fill_u64(w.dsum.p, 0ull, 2, s);
      hipLaunchKernelGGL(k_list_edges, dim3(grid_for(std::max<int64_t>(n, 1), 256, 1024)), dim3(256), 0, s, front,
                         nfront, (const int32_t*)cand, (const int32_t*)(w.off.as<int32_t>() + ng), A.rp,
                         w.dsum.as<unsigned long long>(), (const BcCtl*)nullptr, 0);
      GA_HIP(hipMemcpyAsync(edges_h, w.dsum.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    }
    GA_HIP(hipGetLastError());
    ev_close();
    ++timed;
    GA_HIP(hipStreamSynchronize(s));
    nc = *count_h;
  };
  ev_open();
  bc_items(A, lists[0], nf, w, s);
  while (true) {
    // level d + 1 must stay below the unvisited marker
    if (d + 1 >= (int)bc_unvisited<DT>()) {
      GA_REQUIRE(sizeof(DT) < sizeof(uint16_t), GA_ERR_LIMIT, "BFS deeper than 65533 levels");
      GA_HIP(hipStreamSynchronize(s));
      return false;
    }
    int32_t* next = lists[(d + 1) & 1];
    if (!bottom_up && d >= 1 && cx.bc_bottom_up) {
      candidates(lists[d & 1], nf, true);
      bottom_up = (double)edges_h[0] * cx.bc_alpha > (double)edges_h[1];
    } else if (bottom_up) {
      candidates(nullptr, 0, false);
    }
    if (bottom_up) {
      if (nc == 0) break;  // no vertex can be reached any more
      ev_open();
      bc_items(A, cand, nc, w, s);
      hipLaunchKernelGGL(k_bc_bottom_up_items<DT>, dim3(grid_for(nc + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0,
                         s, A.rp, A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                         (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(w.istart.as<int32_t>() + nc), nsp,
                         d, w.depth.as<DT>(), w.sigma.as<double>(), w.partial.as<double>(), w.flag.as<uint8_t>(),
                         (const BcCtl*)nullptr, 0);
      hipLaunchKernelGGL(k_bc_bottom_up_combine<DT>, dim3(grid_for(nc, 4, 65536)), dim3(256), 0, s,
                         (const int32_t*)cand, nc, A.rp, (const int32_t*)w.hstart.as<int32_t>(), nsp, d,
                         w.depth.as<DT>(), w.sigma.as<double>(), (const double*)w.partial.as<double>(),
                         w.flag.as<uint8_t>(), (const int32_t*)nullptr, (const BcCtl*)nullptr, 0);
      GA_HIP(hipGetLastError());
      ev_close();
      ++timed;
      frontier_from_flags(w.flag.as<uint8_t>(), n, w, next, count_h, s);
      const int32_t nn = *count_h;
      if (nn == 0) break;
      nf = nn;
      ++d;
      continue;
    }
    // top-down: the items of level d are in w.it_* (count at istart[nf])
    if (d > 0) ev_open();
    hipLaunchKernelGGL(k_bc_discover_items<DT>, dim3(grid_for(nf + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0, s,
                       A.rp, A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                       (const int32_t*)(w.istart.as<int32_t>() + nf), nsp, d, w.depth.as<DT>(),
                       w.flag.as<uint8_t>(), (const BcCtl*)nullptr, 0);
    GA_HIP(hipGetLastError());
    ev_close();
    ++timed;
    frontier_from_flags(w.flag.as<uint8_t>(), n, w, next, count_h, s);
    const int32_t nn = *count_h;
    if (nn == 0) break;
    ev_open();
    bc_items(A, next, nn, w, s);
    hipLaunchKernelGGL(k_bc_sigma_items<DT>, dim3(grid_for(nn + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0, s, A.rp,
                       A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                       (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(w.istart.as<int32_t>() + nn), nsp,
                       d + 1, (const DT*)w.depth.as<DT>(), w.sigma.as<double>(), w.partial.as<double>(),
                       (const BcCtl*)nullptr, 0);
    hipLaunchKernelGGL(k_bc_sigma_combine<DT>, dim3(grid_for(nn, 4, 65536)), dim3(256), 0, s, (const int32_t*)next, nn,
                       A.rp, (const int32_t*)w.hstart.as<int32_t>(), nsp, d + 1, (const DT*)w.depth.as<DT>(),
                       w.sigma.as<double>(), (const double*)w.partial.as<double>(), (const int32_t*)nullptr,
                       (const BcCtl*)nullptr, 0);
    GA_HIP(hipGetLastError());
    ev_close();
    ++timed;
    nf = nn;
    ++d;
  }
  // d = deepest level that holds vertices; its list is lists[d & 1] (nf entries)
  const int dmax = d;
  if (cx.prof_bytes) {
    hipLaunchKernelGGL(k_bc_account<DT>, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.rp,
                       (const DT*)w.depth.as<DT>(), n, nsp, cx.prof_dev.as<unsigned long long>());
    cx.prof_bytes_host += (int64_t)(dmax + 1) * (4 * (n + 1) + 4 * A.nnz);
    cx.prof_bytes_launches += timed + 1;  // the event pairs of the pass, backward sweep included
  }
  ev_open();
  if (dmax >= 1) {
    hipLaunchKernelGGL(k_bc_last<DT>, dim3(grid_for(nf, 4, 65536)), dim3(256), 0, s, (const int32_t*)lists[dmax & 1], nf,
                       nsp, dmax, (const DT*)w.depth.as<DT>(), (const double*)w.sigma.as<double>(),
                       w.q.as<double>(), (const BcCtl*)nullptr);
    GA_HIP(hipGetLastError());
    for (int k = dmax - 1; k >= 1; --k) {
      // the level-k list again, from the depths (vertices with a lane at depth k)
      hipLaunchKernelGGL(k_level_flags<DT>, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s,
                         (const DT*)w.depth.as<DT>(), n, nsp, k, w.flag.as<uint8_t>(), (const BcCtl*)nullptr);
      int32_t* front = lists[0];
      frontier_from_flags(w.flag.as<uint8_t>(), n, w, front, count_h, s);
      const int32_t nk = *count_h;
      if (nk == 0) continue;
      bc_items(A, front, nk, w, s);
      hipLaunchKernelGGL(k_bc_backward_items<DT>, dim3(grid_for(nk + A.nnz / BC_CH + 1, 4, 65536)), dim3(256), 0, s,
                         A.rp, A.col, (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                         (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(w.istart.as<int32_t>() + nk), nsp,
                         k, (const DT*)w.depth.as<DT>(), (const double*)w.sigma.as<double>(),
                         w.q.as<double>(), w.bc.as<double>(), w.partial.as<double>(), bc_stride,
                         (const BcCtl*)nullptr);
      hipLaunchKernelGGL(k_bc_combine<DT>, dim3(grid_for(nk, 4, 65536)), dim3(256), 0, s, (const int32_t*)front, nk, A.rp,
                         (const int32_t*)w.hstart.as<int32_t>(), nsp, k, (const DT*)w.depth.as<DT>(),
                         (const double*)w.sigma.as<double>(), w.q.as<double>(), w.bc.as<double>(),
                         (const double*)w.partial.as<double>(), bc_stride, (const int32_t*)nullptr,
                         (const BcCtl*)nullptr);
      GA_HIP(hipGetLastError());
    }
  }
  ev_close();
  if (n) GA_HIP(hipMemcpyAsync(bc_h, w.bc.p, sizeof(double) * n * outs, hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  return true;
}

// The pass with level lists sized on the device (GA_BC_LOOP 1 and 2).  The
// same kernels as bc_run, in the same order within a level, on the same
// lists (ascending vertex order) and items, so the scores are bit-identical
// to it; what changes is who knows the sizes.  A forward step is one fixed
// sequence of launches: candidates and the direction decision, the
// bottom-up sweep or the top-down discovery (each skipped by the other
// direction), the next level's list, the path counts; the kernels read the
// level and the list sizes from the control block, and a step queued after
// the last level does nothing.  The host keeps two steps queued ahead of the
// sequence number the last kernel of each step writes to pinned host memory,
// and waits once, for the end of the sweep (it needs the deepest level to
// size the backward loop).  The backward steps then run with no host wait.
// With `replay`, each step is captured once as a HIP graph (kept in the
// workspace while the launch arguments stay the same) and replayed per level.
template <typename DT>
static bool bc_run_dev(const DCsr& A, const int32_t* src_h, const int32_t* lane_h, int64_t ns, int64_t nsp_,
                       int64_t outs, double* bc_h, BcWork& w, hipStream_t s, bool replay) {
  const int64_t n = A.nrows;
  const int64_t n1 = std::max<int64_t>(n, 1);
  const BcSizes z = bc_sizes(n, A.nnz, ns, nsp_, outs);
  const int nsp = (int)z.nsp;
  const int64_t total = z.total;
  const int64_t bc_stride = outs > 1 ? n : 0;
  bc_reserve(w, n, z, ns);
  Context& cx = ctx();
  GA_HIP(hipMemcpyAsync(w.src.p, src_h, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  GA_HIP(hipMemcpyAsync(w.src.as<int32_t>() + ns, lane_h, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_bc_init<DT>, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, w.depth.as<DT>(),
                     w.sigma.as<double>(), total);
  fill_u64(w.bc.p, 0ull, n * outs, s);
  fill_u32(w.flag.p, 0u, ceil_div(n1, 4), s);
  hipLaunchKernelGGL(k_bc_sources<DT>, dim3(grid_for(ns, 256)), dim3(256), 0, s, (const int32_t*)w.src.as<int32_t>(),
                     (const int32_t*)(w.src.as<int32_t>() + ns), (int)ns, nsp, w.depth.as<DT>(),
                     w.sigma.as<double>(), w.flag.as<uint8_t>());
  GA_HIP(hipGetLastError());
  BcCtl* ctl = w.ctl.as<BcCtl>();
  int32_t* lists[2] = {w.levels.as<int32_t>(), w.levels.as<int32_t>() + n1};
  int32_t* cand = w.levels.as<int32_t>() + 2 * n1;
  long long* hw_dev = nullptr;
  long long* hw = progress_words(&hw_dev) + 24;  // words 24..27: d, dmax, over, 2 x steps + done
  hw_dev += 24;
  const int limit = (int)bc_unvisited<DT>();
  // grids: sized for the largest list (one item per wavefront, as bc_run
  // sizes them for the list at hand): the dispatcher hands workgroups to
  // CUs as they free up, which balances the uneven hub items; workgroups
  // past the list's items exit at once
  const unsigned g_items = grid_for(n1 + A.nnz / BC_CH + 1, 4, 65536);
  const unsigned g_vert = grid_for(n1, 4, 16384);
  // list builder: groups of 64 vertices, one per wavefront, at most
  // BCB_MAX_BLOCKS workgroups each walking a contiguous range
  const int64_t ngroups = ceil_div(n1, BC_FLAG_CHUNK);
  const int64_t g_build = std::min<int64_t>(BCB_MAX_BLOCKS, ceil_div(ngroups, BCB_T / 64));
  const int64_t per_block = ceil_div(ngroups, g_build);
  auto build = [&](uint8_t* flag, int32_t* list, int32_t* hst, int32_t* iv, int32_t* it0, int32_t* isl, int32_t* hp,
                   int32_t* tot, const BcCtl* c, int guard, int to_next) {
    hipLaunchKernelGGL(k_bc_build_count, dim3((unsigned)g_build), dim3(BCB_T), 0, s, (const uint8_t*)flag, n, A.rp,
                       w.gcnt.as<int32_t>(), w.gsum.as<long long>(), per_block, c, guard);
    hipLaunchKernelGGL(k_bc_build_emit, dim3((unsigned)g_build), dim3(BCB_T), 0, s, flag, n, A.rp,
                       (const int32_t*)w.gcnt.as<int32_t>(), (const long long*)w.gsum.as<long long>(), per_block,
                       list, hst, iv, it0, isl, hp, tot, c, guard, to_next);
  };
  auto op = [&](int code, int lim, long long* host) {
    hipLaunchKernelGGL(k_bc_ctl, dim3(1), dim3(64), 0, s, ctl, code, (int32_t*)nullptr, (int32_t*)nullptr, 0, 0.0,
                       lim, host);
  };
  // combine passes: one wavefront per hub vertex (more than BC_CH
  // neighbours, so at most nnz / (BC_CH + 1) of them)
  const unsigned g_hub = grid_for(std::min<int64_t>(n1, A.nnz / (BC_CH + 1) + 1), 4, 65536);
  auto fwd_step = [&]() {
    // candidates of a bottom-up step and, while top-down, the decision
    hipLaunchKernelGGL(k_unvisited_flags<DT>, dim3(g_vert), dim3(256), 0, s, A.rp, (const DT*)w.depth.as<DT>(), n,
                       nsp, w.cflag.as<uint8_t>(), (const BcCtl*)ctl, (int)BG_CAND);
    build(w.cflag.as<uint8_t>(), cand, w.chstart.as<int32_t>(), w.cit_v.as<int32_t>(), w.cit_t0.as<int32_t>(),
          w.cit_slot.as<int32_t>(), w.chpos.as<int32_t>(), ctl->cand_tot, ctl, BG_CAND, 0);
    hipLaunchKernelGGL(k_list_edges, dim3(grid_for(n1, 256, 1024)), dim3(256), 0, s, (const int32_t*)nullptr, 0,
                       (const int32_t*)cand, (const int32_t*)ctl->cand_tot, A.rp, ctl->edges, (const BcCtl*)ctl,
                       (int)BG_DECIDE);
    op(BC_OP_DECIDE, 0, nullptr);
    // bottom-up: the candidates pull from level d; top-down: discovery from
    // the level-d list (its items were built with it)
    hipLaunchKernelGGL(k_bc_forward_items<DT>, dim3(g_items), dim3(256), 0, s, A.rp, A.col,
                       (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                       (const int32_t*)w.cit_v.as<int32_t>(), (const int32_t*)w.cit_t0.as<int32_t>(),
                       (const int32_t*)w.cit_slot.as<int32_t>(), nsp, w.depth.as<DT>(), w.sigma.as<double>(),
                       w.partial.as<double>(), w.flag.as<uint8_t>(), (const BcCtl*)ctl);
    hipLaunchKernelGGL(k_bc_bottom_up_combine<DT>, dim3(g_hub), dim3(256), 0, s, (const int32_t*)cand, 0, A.rp,
                       (const int32_t*)w.chstart.as<int32_t>(), nsp, 0, w.depth.as<DT>(), w.sigma.as<double>(),
                       (const double*)w.partial.as<double>(), w.flag.as<uint8_t>(),
                       (const int32_t*)w.chpos.as<int32_t>(), (const BcCtl*)ctl, (int)BG_BU);
    // the level-(d+1) list and its items
    build(w.flag.as<uint8_t>(), nullptr, w.hstart.as<int32_t>(), w.it_v.as<int32_t>(), w.it_t0.as<int32_t>(),
          w.it_slot.as<int32_t>(), w.hpos.as<int32_t>(), ctl->next_tot, ctl, BG_ANY, 1);
    op(BC_OP_NEXT, 0, nullptr);
    // top-down: path counts of level d + 1
    hipLaunchKernelGGL(k_bc_sigma_items<DT>, dim3(g_items), dim3(256), 0, s, A.rp, A.col,
                       (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                       (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(ctl->next_tot + 1), nsp, 0,
                       (const DT*)w.depth.as<DT>(), w.sigma.as<double>(), w.partial.as<double>(), (const BcCtl*)ctl,
                       (int)BG_TD);
    hipLaunchKernelGGL(k_bc_sigma_combine<DT>, dim3(g_hub), dim3(256), 0, s, (const int32_t*)nullptr, 0, A.rp,
                       (const int32_t*)w.hstart.as<int32_t>(), nsp, 0, (const DT*)w.depth.as<DT>(),
                       w.sigma.as<double>(), (const double*)w.partial.as<double>(),
                       (const int32_t*)w.hpos.as<int32_t>(), (const BcCtl*)ctl, (int)BG_TD);
    op(BC_OP_END, limit, hw_dev);
  };
  auto bwd_step = [&]() {
    hipLaunchKernelGGL(k_level_flags<DT>, dim3(g_vert), dim3(256), 0, s, (const DT*)w.depth.as<DT>(), n, nsp, 0,
                       w.flag.as<uint8_t>(), (const BcCtl*)ctl);
    build(w.flag.as<uint8_t>(), lists[0], w.hstart.as<int32_t>(), w.it_v.as<int32_t>(), w.it_t0.as<int32_t>(),
          w.it_slot.as<int32_t>(), w.hpos.as<int32_t>(), ctl->back_tot, nullptr, BG_ANY, 0);
    hipLaunchKernelGGL(k_bc_backward_items<DT>, dim3(g_items), dim3(256), 0, s, A.rp, A.col,
                       (const int32_t*)w.it_v.as<int32_t>(), (const int32_t*)w.it_t0.as<int32_t>(),
                       (const int32_t*)w.it_slot.as<int32_t>(), (const int32_t*)(ctl->back_tot + 1), nsp, 0,
                       (const DT*)w.depth.as<DT>(), (const double*)w.sigma.as<double>(), w.q.as<double>(),
                       w.bc.as<double>(), w.partial.as<double>(), bc_stride, (const BcCtl*)ctl);
    hipLaunchKernelGGL(k_bc_combine<DT>, dim3(g_hub), dim3(256), 0, s, (const int32_t*)lists[0], 0, A.rp,
                       (const int32_t*)w.hstart.as<int32_t>(), nsp, 0, (const DT*)w.depth.as<DT>(),
                       (const double*)w.sigma.as<double>(), w.q.as<double>(), w.bc.as<double>(),
                       (const double*)w.partial.as<double>(), bc_stride, (const int32_t*)w.hpos.as<int32_t>(),
                       (const BcCtl*)ctl);
    op(BC_OP_BACK, 0, nullptr);
  };
  // captured steps, kept while every launch argument they hold is the same
  if (replay) {
    const std::vector<uintptr_t> key = {
        (uintptr_t)A.rp, (uintptr_t)A.col, (uintptr_t)n, (uintptr_t)A.nnz, (uintptr_t)nsp, (uintptr_t)bc_stride,
        sizeof(DT), (uintptr_t)w.depth.p, (uintptr_t)w.sigma.p, (uintptr_t)w.q.p, (uintptr_t)w.bc.p,
        (uintptr_t)w.flag.p, (uintptr_t)w.cflag.p, (uintptr_t)w.levels.p, (uintptr_t)w.it_v.p, (uintptr_t)w.it_t0.p,
        (uintptr_t)w.it_slot.p, (uintptr_t)w.hstart.p, (uintptr_t)w.cit_v.p, (uintptr_t)w.cit_t0.p,
        (uintptr_t)w.cit_slot.p, (uintptr_t)w.chstart.p, (uintptr_t)w.partial.p, (uintptr_t)w.gcnt.p,
        (uintptr_t)w.gsum.p, (uintptr_t)w.hpos.p, (uintptr_t)w.chpos.p, (uintptr_t)ctl, (uintptr_t)hw_dev,
        (uintptr_t)cx.num_cu};
    if (w.graphs.key != key || !w.graphs.fwd) {
      w.graphs.reset();
      auto capture = [&](auto&& step) {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        GA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        step();
        GA_HIP(hipStreamEndCapture(s, &g));
        const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GA_HIP(ie);
        return e;
      };
      w.graphs.fwd = capture(fwd_step);
      w.graphs.bwd = capture(bwd_step);
      w.graphs.key = key;
    }
  }
  hipEvent_t ev0 = nullptr;
  auto ev_open = [&] {
    if (!cx.prof_events_on) return;
    ev0 = bc_event(cx);
    GA_HIP(hipEventRecord(ev0, s));
  };
  auto ev_close = [&] {
    if (!cx.prof_events_on) return;
    hipEvent_t ev1 = bc_event(cx);
    GA_HIP(hipEventRecord(ev1, s));
    cx.prof_events.emplace_back(ev0, ev1);
    cx.prof_launches += 1;
  };
  // level 0: the sources
  hipLaunchKernelGGL(k_bc_ctl, dim3(1), dim3(64), 0, s, ctl, (int)BC_OP_INIT, lists[0], lists[1],
                     cx.bc_bottom_up ? 1 : 0, cx.bc_alpha, limit, (long long*)nullptr);
  build(w.flag.as<uint8_t>(), nullptr, w.hstart.as<int32_t>(), w.it_v.as<int32_t>(), w.it_t0.as<int32_t>(),
        w.it_slot.as<int32_t>(), w.hpos.as<int32_t>(), ctl->next_tot, ctl, BG_ANY, 1);
  op(BC_OP_START, 0, nullptr);
  GA_HIP(hipGetLastError());
  // forward sweep: steps queued two ahead of the published sequence number
  for (int i = 0; i < 4; ++i) __atomic_store_n(&hw[i], 0ll, __ATOMIC_RELAXED);
  ev_open();
  long long queued = 0, dmax = 0, over = 0;
  const long long max_steps = (long long)limit + 2;
  while (true) {
    const long long word = __atomic_load_n(&hw[3], __ATOMIC_ACQUIRE);
    const long long seq = word >> 1;
    if (word & 1) {  // the sweep is over; dmax and over were stored before this word
      dmax = __atomic_load_n(&hw[1], __ATOMIC_RELAXED);
      over = __atomic_load_n(&hw[2], __ATOMIC_RELAXED);
      break;
    }
    if (queued - seq < 2) {
      GA_REQUIRE(queued < max_steps, GA_ERR_STATE, "BC forward sweep did not finish within its level bound");
      if (replay) GA_HIP(hipGraphLaunch(w.graphs.fwd, s));
      else fwd_step();
      GA_HIP(hipGetLastError());
      ++queued;
      continue;
    }
    // wait for the device; a failed launch surfaces here instead of a hang
    const hipError_t q = hipStreamQuery(s);
    if (q != hipSuccess && q != hipErrorNotReady) GA_HIP(q);
  }
  ev_close();
  if (over) {  // deeper than DT can hold: the caller redoes the pass with 16-bit depths
    GA_REQUIRE(sizeof(DT) < sizeof(uint16_t), GA_ERR_LIMIT, "BFS deeper than 65533 levels");
    GA_HIP(hipStreamSynchronize(s));
    return false;
  }
  if (cx.prof_bytes) {
    hipLaunchKernelGGL(k_bc_account<DT>, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, A.rp,
                       (const DT*)w.depth.as<DT>(), n, nsp, cx.prof_dev.as<unsigned long long>());
    cx.prof_bytes_host += (dmax + 1) * (4 * (n + 1) + 4 * A.nnz);
    cx.prof_bytes_launches += 2;  // the forward sweep and the backward sweep
  }
  ev_open();
  if (dmax >= 1) {
    hipLaunchKernelGGL(k_bc_last<DT>, dim3(g_items), dim3(256), 0, s, (const int32_t*)nullptr, 0, nsp, 0,
                       (const DT*)w.depth.as<DT>(), (const double*)w.sigma.as<double>(), w.q.as<double>(),
                       (const BcCtl*)ctl);
    if (dmax >= 2) {
      op(BC_OP_BACK, (int)dmax - 1, nullptr);  // first backward level
      for (long long k = dmax - 1; k >= 1; --k) {
        if (replay) GA_HIP(hipGraphLaunch(w.graphs.bwd, s));
        else bwd_step();
      }
    }
    GA_HIP(hipGetLastError());
  }
  ev_close();
  if (n) GA_HIP(hipMemcpyAsync(bc_h, w.bc.p, sizeof(double) * n * outs, hipMemcpyDeviceToHost, s));
  GA_HIP(hipStreamSynchronize(s));
  return true;
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

// Serving: a request loop over graphs resident in HBM.
//
// The reference answers one query per process run (src/main.cpp:185-360:
// read a graph, run one algorithm, exit).  Here graphs stay uploaded and
// queries -- triangle count (src/main.cpp:65-86), k-truss for one k
// (:89-146), batched BC (include/betweenness_centrality.h:125-237) -- arrive
// as a stream.
//
// Memory.  The server reserves one device arena (hbm_budget bytes) at
// creation.  Each resident graph is a block of it; each admitted query gets
// one block of exactly its footprint, the sum of the buffer sizes its path
// takes (mxm_workspace_bytes, ktruss_workspace_bytes, bc_workspace_bytes).
// While the query runs its worker points the thread-local lease at that
// block, so every workspace buffer of the path comes from the block: no
// hipMalloc, and no hipFree (which would synchronise the whole device) while
// other queries are in flight.  A footprint that can never fit beside the
// resident graphs is rejected at submit (GA_ERR_OOM).
//
// Scheduling.  `slots` workers, each a host thread with a HIP stream of its
// own (default priority).  A free worker takes the oldest queued query as
// soon as the arena can give it its block; the query's kernels and host
// steering then run on that worker's stream, interleaved on the device with
// the other slots' queries (continuous batching at the granularity of the
// paths' launches: a short triangle count proceeds while a k-truss waits on
// its filter step, and slots refill the moment a query ends).  Admission is
// in arrival order (no overtaking), so a large query is never starved.  A
// worker that admits a query also takes the queued queries of the same kind
// on the same graph: triangle counts share one count, k-truss queries one
// peeling run over their k values in ascending order, BC queries of at most
// 64 sources one multi-source pass.  The one exception to arrival order
// takes nothing from the others: a queued k-truss query joins a peeling run
// on its graph that is already going when that run reaches a k at or below
// the query's (it uses the run's memory and slot), and every k-truss query
// is answered as soon as its own k is reached (served runs peel with the
// host reading each round's size, GA_OPT_TRUSS_LOOP 4, so that it can).
//
// Graphs may be added while queries run: a query holds a pointer to its
// graph (taken under the server lock at submit), never an index into the
// growing graph table.
#include "common.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace ga {

// First-fit allocator over the arena's byte range (host-side bookkeeping).
class Arena {
 public:
  static constexpr size_t kAlign = 4096;
  static constexpr size_t kNone = ~size_t(0);
  void reset(size_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = 0;
  }
  size_t alloc(size_t bytes) {
    const size_t b = (std::max<size_t>(bytes, 1) + kAlign - 1) & ~(kAlign - 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < b) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > b) free_[off + b] = len - b;
      used_ += b;
      return off;
    }
    return kNone;
  }
  void release(size_t off, size_t bytes) {
    const size_t b = (std::max<size_t>(bytes, 1) + kAlign - 1) & ~(kAlign - 1);
    used_ -= b;
    auto it = free_.emplace(off, b).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t cap() const { return cap_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0;
};

struct ServerGraph {
  int64_t n = 0, nnz = 0, lnnz = 0, max_row = 0;
  OwnedCsr A, L;
  size_t off = 0, bytes = 0;
};

struct Job {
  int64_t ticket = 0;
  ga_query q{};
  // the query's graph, taken under the server lock at submit: workers read it
  // without the lock while ga_server_add_graph may grow the graph table
  const ServerGraph* g = nullptr;
  std::vector<int32_t> src;
  size_t footprint = 0;
  size_t off = Arena::kNone;
  bool done = false;
  std::string err;
  ga_query_result res{};
};

constexpr int kMaxJoined = 64;  // queries one k-truss run answers at most (its own and joined ones)

size_t graph_footprint(int64_t n, int64_t nnz, int64_t lnnz) {
  return 2 * lease_bytes(sizeof(int32_t) * (n + 1)) + lease_bytes(sizeof(int32_t) * std::max<int64_t>(nnz, 1)) +
         lease_bytes(sizeof(int32_t) * std::max<int64_t>(n, 1)) + scan_tmp_bytes(n) +
         lease_bytes(sizeof(int32_t) * std::max<int64_t>(lnnz, 1));
}

}  // namespace ga

using namespace ga;

struct ga_server {
  int device = -1;
  int slots = 0;
  char* base = nullptr;
  Arena arena;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<std::unique_ptr<ServerGraph>> graphs;
  std::vector<std::thread> workers;
  bool stop = false;
  int64_t next_ticket = 1;
  int inflight = 0;
  size_t resident = 0;
  ga_server_stats st{};

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  int max_batch = 1;
  void worker(int slot);
  // early[i] = 1: set[i] was completed (and may already be freed by its
  // waiter) before the run ended; a k-truss run appends the queries that
  // join it to set (and early)
  void run_batch(std::vector<Job*>& set, std::vector<char>& early, int slot, size_t off, size_t fp, hipStream_t s);
  size_t footprint(const ga_query& q) const;
  size_t batch_footprint(const std::vector<Job*>& set) const;
  std::vector<Job*> merge_set(Job* h) const;
};

size_t ga_server::footprint(const ga_query& q) const {
  GA_REQUIRE(q.graph >= 0 && q.graph < (int)graphs.size(), GA_ERR_ARG, "unknown graph id");
  const ServerGraph& g = *graphs[q.graph];
  switch (q.kind) {
    case GA_Q_TRIANGLES: return lease_bytes(64) + mxm_workspace_bytes(g.n, g.lnnz, false);
    case GA_Q_KTRUSS: return ktruss_workspace_bytes(g.n, g.nnz);
    case GA_Q_BC:
      GA_REQUIRE(q.nsources >= 1 && q.nsources <= 4096, GA_ERR_LIMIT, "between 1 and 4096 sources per query");
      return bc_workspace_bytes(g.n, g.nnz, q.nsources);
  }
  throw Error{GA_ERR_ARG, "unknown query kind " + std::to_string(q.kind)};
}

size_t ga_server::batch_footprint(const std::vector<Job*>& set) const {
  if (set.size() == 1 || set[0]->q.kind == GA_Q_TRIANGLES) return set[0]->footprint;  // one count serves all
  const ServerGraph& g = *set[0]->g;
  if (set[0]->q.kind == GA_Q_BC) {
    std::vector<int64_t> sizes;
    for (const Job* j : set) sizes.push_back((int64_t)j->src.size());
    return bc_groups_workspace_bytes(g.n, g.nnz, sizes);
  }
  return ktruss_workspace_bytes(g.n, g.nnz);  // one peeling run serves every k
}

// Queued jobs that can share the head's run: same graph and kind, BC sets of
// at most 64 sources (one lane group each), any k-truss, any triangle count
// (one count answers them all).
std::vector<Job*> ga_server::merge_set(Job* h) const {
  std::vector<Job*> set{h};
  const bool bc = h->q.kind == GA_Q_BC && h->src.size() <= 64;
  if (max_batch <= 1 || !(bc || h->q.kind == GA_Q_KTRUSS || h->q.kind == GA_Q_TRIANGLES)) return set;
  for (size_t i = 1; i < queue.size() && (int)set.size() < max_batch; ++i) {
    Job* c = queue[i];
    if (c->q.graph != h->q.graph || c->q.kind != h->q.kind) continue;
    if (bc && c->src.size() > 64) continue;
    set.push_back(c);
  }
  return set;
}

void ga_server::run_batch(std::vector<Job*>& set, std::vector<char>& early, int slot, size_t off, size_t fp,
                          hipStream_t s) {
  Context& c = ctx();
  GA_REQUIRE(!c.prof_events_on && !c.prof_bytes, GA_ERR_STATE,
             "kernel timing hooks are process-wide: disable them while a server runs queries");
  Job& j = *set[0];
  const ServerGraph& g = *j.g;
  Lease lease{base + off, fp, 0};
  t_lease = &lease;
  struct Unset {
    ~Unset() { t_lease = nullptr; }
  } unset;
  switch (j.q.kind) {
    case GA_Q_TRIANGLES: {
      DevBuf sums;
      sums.ensure(64);
      fill_u64(sums.p, 0ull, 1, s);
      MxmWork w;
      w.max_mask_row = w.max_b_row = g.max_row;
      const DCsr vL = g.L.view();
      mxm_masked_dev(vL, vL, vL, GA_INT64, 1, nullptr, sums.as<unsigned long long>(), nullptr, w, s);
      const int64_t t = (int64_t)publish_sum_and_wait(w, sums.as<unsigned long long>(), s);
      for (Job* x : set) x->res.value = t;  // queued counts of the same graph share this one
      break;
    }
    case GA_Q_KTRUSS: {
      std::vector<int32_t> ks;
      for (const Job* x : set) ks.push_back(x->q.k);
      std::sort(ks.begin(), ks.end());
      ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
      std::vector<int64_t> nnz(ks.size(), 0);
      std::vector<int32_t> passes(ks.size(), 0);
      TrussWork tw;
      tw.mw.max_mask_row = tw.mw.max_b_row = g.max_row;
      tw.no_graph = true;  // a workspace of one query: no capture to amortise
      // A query is answered as soon as its k-truss is known: the run goes on
      // to the larger k of the other queries without it.  While the run can
      // still go on, queued k-truss queries on the same graph join it: one
      // with the k just reached is answered at once, one with a larger k is
      // added to the run's list (no memory or slot of its own, so joining
      // delays no other query).
      const int gid = j.q.graph;
      const int join = c.serve_join;
      int32_t k_top = ks.back();  // the run's largest k
      const KDone on_k = [&](int32_t k, int64_t nz, int32_t ps, std::vector<int32_t>* more) {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_ms();
        if (more && max_batch > 1 && join > 0) {
          for (auto it = queue.begin(); it != queue.end() && (int)set.size() < kMaxJoined;) {
            Job* c = *it;
            if (c->q.kind != GA_Q_KTRUSS || c->q.graph != gid || c->q.k < k || (join == 1 && c->q.k > k_top)) {
              ++it;
              continue;
            }
            k_top = std::max(k_top, c->q.k);
            it = queue.erase(it);
            c->res.t_start_ms = t;
            c->res.slot = slot;
            c->res.footprint = (int64_t)fp;
            c->res.batched = (int32_t)set.size() + 1;
            set.push_back(c);
            early.push_back(0);
            st.joined += 1;
            if (c->q.k > k) more->push_back(c->q.k);
          }
        }
        for (size_t q = 0; q < set.size(); ++q) {
          Job* x = set[q];
          if (early[q] || x->q.k != k) continue;
          x->res.value = nz;
          x->res.passes = ps;
          x->res.status = GA_OK;
          x->res.t_end_ms = t;
          x->done = true;
          early[q] = 1;
          st.completed += 1;
        }
        cv_done.notify_all();
      };
      ktruss_multi_dev(g.A.view(), ks.data(), (int)ks.size(), tw, s, nnz.data(), passes.data(), &on_k);
      for (size_t q = 0; q < set.size(); ++q) {
        if (early[q]) continue;
        const size_t i = std::lower_bound(ks.begin(), ks.end(), set[q]->q.k) - ks.begin();
        set[q]->res.value = nnz[i];
        set[q]->res.passes = passes[i];
      }
      break;
    }
    case GA_Q_BC: {
      BcWork w;
      if (set.size() == 1) {
        bc_batch_dev(g.A.view(), j.src.data(), (int64_t)j.src.size(), j.q.bc, w, s);
      } else {
        std::vector<std::vector<int32_t>> sets;
        std::vector<double*> outs;
        for (Job* x : set) {
          sets.push_back(x->src);
          outs.push_back(x->q.bc);
        }
        bc_groups_dev(g.A.view(), sets, outs.data(), w, s);
      }
      for (Job* x : set) x->res.value = 0;
      break;
    }
  }
  GA_HIP(hipStreamSynchronize(s));
}

void ga_server::worker(int slot) {
  hipStream_t s = nullptr;
  bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  if (ok) ok = guarded([] { pinned_scratch(); }) == GA_OK;
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    std::vector<Job*> set;
    size_t off = Arena::kNone, fp = 0;
    while (true) {
      if (!queue.empty()) {
        Job* h = queue.front();
        set = merge_set(h);
        fp = batch_footprint(set);
        off = arena.alloc(fp);
        if (off == Arena::kNone && set.size() > 1) {
          set.assign(1, h);
          fp = h->footprint;
          off = arena.alloc(fp);
        }
        if (off != Arena::kNone) {
          for (Job* x : set) queue.erase(std::find(queue.begin(), queue.end(), x));
          break;
        }
        if (inflight == 0) {  // cannot fit even with every query finished (fragmentation by graphs)
          queue.pop_front();
          h->res.status = GA_ERR_OOM;
          h->err = "query footprint does not fit in the arena beside the resident graphs";
          h->res.slot = -1;
          h->done = true;
          st.rejected += 1;
          cv_done.notify_all();
          continue;
        }
      } else if (stop) {
        lk.unlock();
        if (s) (void)hipStreamDestroy(s);
        return;
      }
      cv_work.wait(lk);
    }
    inflight += 1;
    st.peak_inflight = std::max<int32_t>(st.peak_inflight, inflight);
    st.peak_bytes = std::max<int64_t>(st.peak_bytes, (int64_t)arena.used());
    st.batches += 1;
    const double t_start = now_ms();
    for (Job* x : set) {
      x->res.t_start_ms = t_start;
      x->res.slot = slot;
      x->res.footprint = (int64_t)fp;
      x->res.batched = (int32_t)set.size();
    }
    std::vector<char> early(set.size(), 0);
    lk.unlock();
    int code = GA_OK;
    std::string msg;
    if (!ok) {
      code = GA_ERR_HIP;
      msg = "worker stream could not be created";
    } else {
      code = guarded([&] { run_batch(set, early, slot, off, fp, s); });
      if (code != GA_OK) {
        msg = ga_last_error();
        (void)hipStreamSynchronize(s);  // the block is reused only after its kernels ended
      }
    }
    const double t_end = now_ms();
    lk.lock();
    arena.release(off, fp);
    inflight -= 1;
    for (size_t q = 0; q < set.size(); ++q) {
      if (early[q]) continue;  // answered during the run (possibly freed by now)
      Job* x = set[q];
      x->res.status = code;
      x->err = msg;
      x->res.t_end_ms = t_end;
      x->done = true;
      st.completed += 1;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

ga_status ga_server_create(const ga_server_config* cfg, ga_server** out) {
  return guarded([&] {
    GA_REQUIRE(cfg && out, GA_ERR_ARG, "null argument");
    GA_REQUIRE(cfg->slots >= 1 && cfg->slots <= 64, GA_ERR_ARG, "slots must be 1..64");
    GA_REQUIRE(cfg->hbm_budget > 0, GA_ERR_ARG, "hbm_budget must be positive");
    Context& c = ctx();
    auto srv = std::make_unique<ga_server>();
    srv->device = c.device;
    srv->slots = cfg->slots;
    GA_REQUIRE(cfg->max_batch >= 0 && cfg->max_batch <= 64, GA_ERR_ARG, "max_batch must be 0..64");
    srv->max_batch = cfg->max_batch == 0 ? 16 : cfg->max_batch;
    const size_t cap = (size_t)cfg->hbm_budget & ~(Arena::kAlign - 1);
    GA_REQUIRE(cap > 0, GA_ERR_ARG, "hbm_budget below one page");
    if (hipMalloc((void**)&srv->base, cap) != hipSuccess) {
      (void)hipGetLastError();
      throw Error{GA_ERR_OOM, "cannot reserve a server arena of " + std::to_string(cap) + " bytes"};
    }
    srv->arena.reset(cap);
    srv->st.arena_bytes = (int64_t)cap;
    srv->st.slots = cfg->slots;
    srv->t0 = std::chrono::steady_clock::now();
    try {
      for (int i = 0; i < cfg->slots; ++i) srv->workers.emplace_back(&ga_server::worker, srv.get(), i);
    } catch (...) {
      {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
      }
      srv->cv_work.notify_all();
      for (auto& t : srv->workers) t.join();
      (void)hipFree(srv->base);
      throw;
    }
    *out = srv.release();
  });
}

ga_status ga_server_destroy(ga_server* srv) {
  return guarded([&] {
    if (!srv) return;
    {
      std::lock_guard<std::mutex> g(srv->mu);
      srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& t : srv->workers) t.join();
    GA_HIP(hipSetDevice(srv->device));
    GA_HIP(hipDeviceSynchronize());
    srv->graphs.clear();  // graph buffers are arena blocks (not hipMalloc'd)
    (void)hipFree(srv->base);
    delete srv;
  });
}

ga_status ga_server_add_graph(ga_server* srv, const ga_csr* A, int32_t* graph_id) {
  return guarded([&] {
    GA_REQUIRE(srv && graph_id, GA_ERR_ARG, "null argument");
    validate_host_csr(A, "A");
    GA_REQUIRE(A->nrows == A->ncols, GA_ERR_DIM, "graphs are square");
    GA_REQUIRE(host_csr_symmetric(A), GA_ERR_ARG, "served graphs need a symmetric pattern");
    auto g = std::make_unique<ServerGraph>();
    g->n = A->nrows;
    g->nnz = A->nnz;
    g->max_row = host_max_row(A);
    for (int64_t i = 0; i < A->nrows; ++i)
      for (int32_t t = A->rowptr[i]; t < A->rowptr[i + 1] && A->col[t] < i; ++t) ++g->lnnz;
    g->bytes = graph_footprint(g->n, g->nnz, g->lnnz);
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      g->off = srv->arena.alloc(g->bytes);
      GA_REQUIRE(g->off != Arena::kNone, GA_ERR_OOM,
                 "the server arena cannot hold a graph of " + std::to_string(g->bytes) + " bytes");
      srv->resident += g->bytes;
      srv->st.resident_bytes = (int64_t)srv->resident;
      srv->st.peak_bytes = std::max<int64_t>(srv->st.peak_bytes, (int64_t)srv->arena.used());
    }
    try {
      Context& c = ctx();
      hipStream_t s = c.stream;
      Lease lease{srv->base + g->off, g->bytes, 0};
      t_lease = &lease;
      struct Unset {
        ~Unset() { t_lease = nullptr; }
      } unset;
      ga_csr pat = *A;
      pat.val = nullptr;
      upload_csr(&pat, GA_INT32, g->A, s);
      DevBuf tmp;
      lower_triangle_dev(g->A.view(), g->L, tmp, s);
      GA_HIP(hipStreamSynchronize(s));
      GA_REQUIRE(g->L.nnz == g->lnnz, GA_ERR_STATE, "lower triangle size differs from the host count");
    } catch (...) {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->arena.release(g->off, g->bytes);
      srv->resident -= g->bytes;
      srv->st.resident_bytes = (int64_t)srv->resident;
      throw;
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->graphs.push_back(std::move(g));
    *graph_id = (int32_t)(srv->graphs.size() - 1);
  });
}

ga_status ga_server_query_footprint(ga_server* srv, const ga_query* q, int64_t* bytes) {
  return guarded([&] {
    GA_REQUIRE(srv && q && bytes, GA_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *bytes = (int64_t)srv->footprint(*q);
  });
}

ga_status ga_server_submit(ga_server* srv, const ga_query* q, int64_t* ticket) {
  return guarded([&] {
    GA_REQUIRE(srv && q && ticket, GA_ERR_ARG, "null argument");
    auto j = std::make_unique<Job>();
    j->q = *q;
    std::lock_guard<std::mutex> lk(srv->mu);
    GA_REQUIRE(!srv->stop, GA_ERR_STATE, "server is shutting down");
    j->footprint = srv->footprint(*q);  // validates graph id, kind and source count
    j->g = srv->graphs[q->graph].get();
    const ServerGraph& g = *j->g;
    if (q->kind == GA_Q_KTRUSS) GA_REQUIRE(q->k >= 3, GA_ERR_ARG, "k must be >= 3");
    if (q->kind == GA_Q_BC) {
      GA_REQUIRE(q->sources && q->bc, GA_ERR_ARG, "BC query needs sources and an output buffer");
      std::vector<char> seen(g.n, 0);
      j->src.assign(q->sources, q->sources + q->nsources);
      for (int32_t v : j->src) {
        GA_REQUIRE(v >= 0 && v < g.n, GA_ERR_ARG, "source out of range");
        GA_REQUIRE(!seen[v], GA_ERR_ARG, "duplicate source");
        seen[v] = 1;
      }
      j->q.sources = nullptr;
    }
    const size_t room = srv->arena.cap() - srv->resident;
    if (j->footprint + Arena::kAlign > room) {
      srv->st.rejected += 1;
      throw Error{GA_ERR_OOM, "query footprint of " + std::to_string(j->footprint) +
                                  " bytes exceeds the arena left beside the resident graphs (" +
                                  std::to_string(room) + " bytes)"};
    }
    j->ticket = srv->next_ticket++;
    j->res.t_submit_ms = srv->now_ms();
    j->res.slot = -1;
    *ticket = j->ticket;
    srv->queue.push_back(j.get());
    srv->jobs.emplace(j->ticket, std::move(j));
    srv->st.submitted += 1;
    srv->cv_work.notify_one();
  });
}

ga_status ga_server_wait(ga_server* srv, int64_t ticket, ga_query_result* out) {
  std::string err;
  int code = guarded([&] {
    GA_REQUIRE(srv != nullptr, GA_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    GA_REQUIRE(it != srv->jobs.end(), GA_ERR_ARG, "unknown ticket " + std::to_string(ticket));
    Job* j = it->second.get();
    srv->cv_done.wait(lk, [&] { return j->done; });
    if (out) *out = j->res;
    err = j->err;
    const int status = j->res.status;
    srv->jobs.erase(it);
    if (status != GA_OK) throw Error{status, err};
  });
  return code;
}

ga_status ga_server_stats_read(ga_server* srv, ga_server_stats* out) {
  return guarded([&] {
    GA_REQUIRE(srv && out, GA_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->st;
  });
}

ga_status ga_server_replay(ga_server* srv, int64_t nq, const ga_query* q, const double* arrival_ms,
                           ga_query_result* out) {
  return guarded([&] {
    GA_REQUIRE(srv && (nq == 0 || (q && arrival_ms && out)), GA_ERR_ARG, "null argument");
    for (int64_t i = 1; i < nq; ++i) GA_REQUIRE(arrival_ms[i] >= arrival_ms[i - 1], GA_ERR_ARG, "arrivals unsorted");
    std::vector<int64_t> tickets(nq, 0);
    std::vector<int> sub(nq, GA_OK);
    std::vector<std::string> msg(nq);
    const auto start = std::chrono::steady_clock::now();
    const double base_ms = srv->now_ms();
    for (int64_t i = 0; i < nq; ++i) {
      std::this_thread::sleep_until(start + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                std::chrono::duration<double, std::milli>(arrival_ms[i])));
      sub[i] = ga_server_submit(srv, &q[i], &tickets[i]);
      if (sub[i] != GA_OK) msg[i] = ga_last_error();
    }
    int first_err = GA_OK;
    std::string first_msg;
    for (int64_t i = 0; i < nq; ++i) {
      ga_query_result r{};
      r.slot = -1;
      int code = sub[i];
      if (code == GA_OK) {
        code = ga_server_wait(srv, tickets[i], &r);
        if (code != GA_OK) msg[i] = ga_last_error();
        r.t_submit_ms -= base_ms;
        r.t_start_ms -= base_ms;
        r.t_end_ms -= base_ms;
      }
      r.status = code;
      out[i] = r;
      if (code != GA_OK && first_err == GA_OK) {
        first_err = code;
        first_msg = "query " + std::to_string(i) + ": " + msg[i];
      }
    }
    if (first_err != GA_OK) throw Error{first_err, first_msg};
  });
}

ga_status ga_plain_replay(int64_t ngraphs, ga_graph* const* graphs, int64_t nq, const ga_query* q,
                          const double* arrival_ms, ga_query_result* out) {
  return guarded([&] {
    GA_REQUIRE(graphs && (nq == 0 || (q && arrival_ms && out)), GA_ERR_ARG, "null argument");
    for (int64_t i = 1; i < nq; ++i) GA_REQUIRE(arrival_ms[i] >= arrival_ms[i - 1], GA_ERR_ARG, "arrivals unsorted");
    const auto start = std::chrono::steady_clock::now();
    auto ms = [&] {
      return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    };
    int first_err = GA_OK;
    std::string first_msg;
    for (int64_t i = 0; i < nq; ++i) {
      std::this_thread::sleep_until(start + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                std::chrono::duration<double, std::milli>(arrival_ms[i])));
      ga_query_result r{};
      r.slot = 0;
      r.t_submit_ms = arrival_ms[i];
      r.t_start_ms = ms();
      int code = GA_OK;
      if (q[i].graph < 0 || q[i].graph >= ngraphs) {
        code = GA_ERR_ARG;
        set_error(code, "unknown graph id");
      } else {
        ga_graph* g = graphs[q[i].graph];
        switch (q[i].kind) {
          case GA_Q_TRIANGLES: code = ga_graph_triangle_count(g, &r.value); break;
          case GA_Q_KTRUSS: code = ga_graph_ktruss(g, q[i].k, &r.value, &r.passes); break;
          case GA_Q_BC: code = ga_graph_bc(g, q[i].sources, q[i].nsources, q[i].bc); break;
          default: code = GA_ERR_ARG; set_error(code, "unknown query kind");
        }
      }
      r.t_end_ms = ms();
      r.status = code;
      out[i] = r;
      if (code != GA_OK && first_err == GA_OK) {
        first_err = code;
        first_msg = "query " + std::to_string(i) + ": " + ga_last_error();
      }
    }
    if (first_err != GA_OK) throw Error{first_err, first_msg};
  });
}

}  // extern "C"

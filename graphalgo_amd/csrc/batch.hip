// Batched triangle counting: thousands of small graphs in one persistent
// launch.
//
// The strict lower triangles of all graphs are stored back to back as one
// block-diagonal CSR whose column indices stay graph-local; a per-row graph
// id and per-graph row base let the masked-SpGEMM kernel address B rows of
// the right graph (k_mxm_items with BATCHED = true).  One work list covers
// every row of every graph, so the persistent grid never drains between
// graphs and one launch replaces ngraphs launches of the per-graph path
// (triangle_counting_masked_sandia, src/main.cpp:65-86, once per graph).
// The launch shape -- workgroup width of the long-row items (A entries per
// work item), rows per claim of the short-row kernel (the rows in flight on
// an XCD, whose graphs share its L2) and the product count up to which a row
// goes to the short-row kernel (LDS bitmap per row, rows packed per
// wavefront) rather than to the item kernel (LDS mask accumulator of the
// workgroup): the per-row choice of accumulator that replaces the
// reference's global msa / mca / heap switch (src/main.cpp:206-218) -- is
// chosen by timing the candidates on the resident batch (ga_batch_autotune).
#include "common.hpp"

#include <algorithm>
#include <vector>

using namespace ga;

struct ga_batch {
  int device = -1;
  int64_t ngraphs = 0, rows = 0, max_n = 0;
  OwnedCsr L;
  DevBuf row_group, gbase, sums;
  MxmWork mw;
  int threads = 0;  // 0 = not tuned (256)
  int claim = 0;    // rows per claim of the short-row kernel (0 = automatic)
  long long wave_flops = -2;  // short-row threshold (products per row); -2 = not tuned (library option)
};

extern "C" {

ga_status ga_batch_create(int64_t ngraphs, const ga_csr* graphs, ga_batch** out) {
  return guarded([&] {
    GA_REQUIRE(out && (ngraphs == 0 || graphs), GA_ERR_ARG, "null argument");
    GA_REQUIRE(ngraphs >= 0 && ngraphs < INT32_MAX, GA_ERR_ARG, "bad graph count");
    int64_t rows = 0, nnz = 0, max_n = 0;
    for (int64_t g = 0; g < ngraphs; ++g) {
      validate_host_csr(&graphs[g], "graph");
      GA_REQUIRE(graphs[g].nrows == graphs[g].ncols, GA_ERR_DIM, "graphs are square");
      rows += graphs[g].nrows;
      max_n = std::max(max_n, graphs[g].nrows);
      for (int64_t i = 0; i < graphs[g].nrows; ++i)
        for (int32_t t = graphs[g].rowptr[i]; t < graphs[g].rowptr[i + 1] && graphs[g].col[t] < i; ++t) ++nnz;
    }
    GA_REQUIRE(rows < INT32_MAX && nnz < INT32_MAX, GA_ERR_LIMIT, "batch exceeds int32 CSR indexing");
    std::vector<int32_t> rp(rows + 1), col(std::max<int64_t>(nnz, 1)), grp(std::max<int64_t>(rows, 1)),
        base(std::max<int64_t>(ngraphs, 1));
    int64_t r = 0, p = 0;
    rp[0] = 0;
    for (int64_t g = 0; g < ngraphs; ++g) {
      base[g] = (int32_t)r;
      const ga_csr& G = graphs[g];
      for (int64_t i = 0; i < G.nrows; ++i, ++r) {
        for (int32_t t = G.rowptr[i]; t < G.rowptr[i + 1] && G.col[t] < i; ++t) col[p++] = G.col[t];
        rp[r + 1] = (int32_t)p;
        grp[r] = (int32_t)g;
      }
    }
    Context& c = ctx();
    auto* b = new ga_batch();
    try {
      b->device = c.device;
      b->ngraphs = ngraphs;
      b->rows = rows;
      b->max_n = max_n;
      b->mw.max_mask_row = b->mw.max_b_row = max_n;
      b->mw.reuse_plan = true;  // the batch is immutable: bins and items are built once per width
      ga_csr h{rows, max_n, nnz, rp.data(), col.data(), nullptr};
      upload_csr(&h, GA_INT32, b->L, c.stream);
      b->row_group.ensure(sizeof(int32_t) * grp.size());
      b->gbase.ensure(sizeof(int32_t) * base.size());
      b->sums.ensure(sizeof(uint64_t) * std::max<int64_t>(ngraphs, 1));
      GA_HIP(hipMemcpyAsync(b->row_group.p, grp.data(), sizeof(int32_t) * grp.size(), hipMemcpyHostToDevice,
                            c.stream));
      GA_HIP(hipMemcpyAsync(b->gbase.p, base.data(), sizeof(int32_t) * base.size(), hipMemcpyHostToDevice,
                            c.stream));
      GA_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

ga_status ga_batch_destroy(ga_batch* b) {
  return guarded([&] {
    if (!b) return;
    GA_HIP(hipStreamSynchronize(ctx().stream));
    delete b;
  });
}

static void batch_run(ga_batch* b, int threads, int claim, long long wave_flops, hipStream_t s) {
  fill_u64(b->sums.p, 0ull, b->ngraphs, s);
  const DCsr L = b->L.view();
  b->mw.threads = threads;
  b->mw.row_chunk = claim;
  b->mw.wave_flops = wave_flops >= 0 ? wave_flops : -1;  // -1: the library option
  mxm_batched_sum_dev(L, b->row_group.as<int32_t>(), b->gbase.as<int32_t>(), b->max_n,
                      b->sums.as<unsigned long long>(), b->mw, s);
}

ga_status ga_batch_triangle_count(ga_batch* b, int threads, int64_t* counts) {
  return guarded([&] {
    GA_REQUIRE(b != nullptr, GA_ERR_ARG, "null batch");
    if (threads == 0) threads = b->threads ? b->threads : 256;
    GA_REQUIRE(threads == 64 || threads == 128 || threads == 256, GA_ERR_ARG, "threads must be 64, 128 or 256");
    hipStream_t s = ctx().stream;
    batch_run(b, threads, b->claim, b->wave_flops, s);
    if (counts && b->ngraphs)
      GA_HIP(hipMemcpyAsync(counts, b->sums.p, sizeof(int64_t) * b->ngraphs, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
  });
}

ga_status ga_batch_autotune(ga_batch* b, int* best_threads, int* best_claim, double* ms_per_candidate) {
  return guarded([&] {
    GA_REQUIRE(b != nullptr, GA_ERR_ARG, "null batch");
    hipStream_t s = ctx().stream;
    const int widths[2] = {128, 256};
    const int claims[4] = {8, 16, 32, 64};
    const long long flops[4] = {2048, 16384, 65536, 1 << 20};
    static_assert(GA_BATCH_TUNE_CANDIDATES == 12, "candidate grid");
    hipEvent_t e0, e1;
    GA_HIP(hipEventCreate(&e0));
    GA_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int arg_t = 256, arg_c = 0;
    long long arg_f = -1;
    try {
      // stage 1: width x claim at the library's threshold; stage 2: the
      // threshold at the best of stage 1
      for (int ci = 0; ci < GA_BATCH_TUNE_CANDIDATES; ++ci) {
        const bool st2 = ci >= 8;
        const int th = st2 ? arg_t : widths[ci / 4], cl = st2 ? arg_c : claims[ci % 4];
        const long long wf = st2 ? flops[ci - 8] : -1;
        batch_run(b, th, cl, wf, s);  // warm-up
        GA_HIP(hipEventRecord(e0, s));
        const int reps = 3;
        for (int r = 0; r < reps; ++r) batch_run(b, th, cl, wf, s);
        GA_HIP(hipEventRecord(e1, s));
        GA_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        GA_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / reps;
        if (ms_per_candidate) ms_per_candidate[ci] = per;
        if (per < best) {
          best = per;
          arg_t = th;
          arg_c = cl;
          arg_f = wf;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    GA_HIP(hipEventDestroy(e0));
    GA_HIP(hipEventDestroy(e1));
    b->threads = arg_t;
    b->claim = arg_c;
    b->wave_flops = arg_f;
    if (best_threads) *best_threads = arg_t;
    if (best_claim) *best_claim = arg_c;
  });
}

ga_status ga_batch_shape(const ga_batch* b, int* threads, int* claim, int64_t* wave_flops) {
  return guarded([&] {
    GA_REQUIRE(b != nullptr, GA_ERR_ARG, "null batch");
    if (threads) *threads = b->threads;
    if (claim) *claim = b->claim;
    if (wave_flops) *wave_flops = b->wave_flops >= 0 ? b->wave_flops : ctx().wave_flops;
  });
}

}  // extern "C"

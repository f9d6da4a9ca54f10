// Drop-in of the reference's include/betweenness_centrality.h.
//
//   betweenness_centrality_batch(isParallel, A, batchSize)
//       include/betweenness_centrality.h:125-237: sources 0..batch-1, one
//       batched pass -> ga_bc_batch.  FP64 on the GPU; the float vector of the
//       reference's signature is the FP64 result rounded once.
//   betweenness_centrality(isParallel, A, blockSize)
//       include/betweenness_centrality.h:5-123: every vertex as a source, in
//       blocks of blockSize -> ga_graph_bc per block on an HBM-resident graph,
//       summed in block order.
// The *_f64 variants return the FP64 values and take explicit source lists.
#pragma once

#include <vector>

#include "matrix_la.h"

inline std::vector<double> betweenness_centrality_batch_f64(const spMtx<int>& A, const std::vector<int32_t>& sources) {
  graphalgo_detail::ensure_init();
  if (A.m != A.n) throw "non-square matrix; BC is only for square matrices";
  std::vector<double> bc(A.m, 0.0);
  const ga_csr a = A.view(false);
  graphalgo_detail::check(ga_bc_batch(&a, sources.data(), (int64_t)sources.size(), bc.data()));
  return bc;
}

inline std::vector<float> betweenness_centrality_batch(bool /*isParallel*/, const spMtx<int>& A, size_t batchSize) {
  if (A.m != A.n) throw "non-square matrix; BC is only for square matrices";
  const size_t ns = std::min(A.n, batchSize);
  std::vector<int32_t> src(ns);
  for (size_t j = 0; j < ns; ++j) src[j] = (int32_t)j;
  std::vector<double> d = betweenness_centrality_batch_f64(A, src);
  return std::vector<float>(d.begin(), d.end());
}

inline std::vector<double> betweenness_centrality_f64(const spMtx<int>& A, size_t blockSize) {
  graphalgo_detail::ensure_init();
  if (A.m != A.n) throw "non-square matrix; BC is only for square matrices";
  if (blockSize == 0) throw "block size must be positive";
  const ga_csr a = A.view(false);
  ga_graph* g = nullptr;
  graphalgo_detail::check(ga_graph_create(&a, &g));
  std::vector<double> total(A.m, 0.0), part(A.m);
  std::vector<int32_t> src;
  for (size_t i = 0; i < A.n; i += blockSize) {
    const size_t nb = std::min(A.n - i, blockSize);
    src.resize(nb);
    for (size_t j = 0; j < nb; ++j) src[j] = (int32_t)(i + j);
    ga_status st = ga_graph_bc(g, src.data(), (int64_t)nb, part.data());
    if (st != GA_OK) {
      ga_graph_destroy(g);
      graphalgo_detail::check(st);
    }
    for (size_t v = 0; v < A.m; ++v) total[v] += part[v];
  }
  graphalgo_detail::check(ga_graph_destroy(g));
  return total;
}

inline std::vector<float> betweenness_centrality(bool /*isParallel*/, const spMtx<int>& A, size_t blockSize) {
  std::vector<double> d = betweenness_centrality_f64(A, blockSize);
  return std::vector<float>(d.begin(), d.end());
}

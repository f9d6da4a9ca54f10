// Reference-style client of the drop-in headers: the flow of the reference's
// CLI (src/main.cpp:188-277) -- read an .mtx, build the adjacency matrix,
// symmetrise its lower triangle, bind an mxmOp<int>, run triangle counting,
// k-truss, batched BC, full-vertex BC and the element-wise operators -- compiled against include/graphalgo/*.h and
// linked with libgraphalgo.so.  Prints one JSON object for the test to check.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "graphalgo/algorithms.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: dropin <graph.mtx>\n");
    return 2;
  }
  try {
    spMtx<int> gr(argv[1], "mtx");
    spMtx<int> TestMtx = build_symm_from_lower(extract_lower_triangle(build_adjacency_matrix(gr)));
    mxmOp<int> op = mxmm_msa<int, int>;
    const int64_t tc = triangle_counting_masked_sandia(TestMtx, op, true);
    const int64_t tc_lu = triangle_counting_masked_lu(TestMtx, mxmm_mca<int, int>, true);
    int it3 = 0, it4 = 0;
    spMtx<int> T3 = k_truss(TestMtx, 3, op, true, &it3);
    spMtx<int> T4 = k_truss(TestMtx, 4, mxmm_heap<int>, true, &it4);
    // C<A> = A*A through the plain operator
    spMtx<int> SQ;
    op(true, TestMtx, TestMtx, TestMtx, SQ);
    long long sq_sum = 0;
    for (size_t j = 0; j < SQ.nz; ++j) sq_sum += SQ.Val[j];
    std::vector<float> bc = betweenness_centrality_batch(true, build_adjacency_matrix(gr), 16);
    double bc_sum = 0;
    for (float x : bc) bc_sum += x;
    int32_t kmax = 0;
    std::vector<int32_t> truss = truss_decomposition(TestMtx, &kmax);
    // every vertex as a source, in blocks of 64 (include/betweenness_centrality.h:5-123)
    std::vector<double> bcf = betweenness_centrality_f64(build_adjacency_matrix(gr), 64);
    // element-wise operators (include/matrix_la.h:272-541) on A and C<A> = A*A
    auto vsum = [](const spMtx<int>& X) {
      long long t = 0;
      for (size_t j = 0; j < X.nz; ++j) t += X.Val[j];
      return t;
    };
    spMtx<int> half = multScalar(SQ, 0.5);
    spMtx<int> ew_add = eWiseAdd(TestMtx, half);
    spMtx<int> ew_mult = eWiseMult(SQ, SQ, TestMtx);
    spMtx<int> ew_cancel = eWiseAdd(SQ, multScalar(SQ, -1), TestMtx);  // every sum is zero
    spMtx<int> low = extract_lower_triangle(TestMtx);
    spMtx<int> ew_disj = add_nointersect(low, transpose(low));
    int disj_err = 0;
    try {
      add_nointersect(TestMtx, low);  // shared positions: refused
    } catch (const char*) {
      disj_err = 1;
    }
    std::printf("{\"n\": %zu, \"nnz\": %zu, \"tc\": %lld, \"tc_lu\": %lld, \"k3_nnz\": %zu, \"k3_it\": %d, "
                "\"k4_nnz\": %zu, \"k4_it\": %d, \"sq_sum\": %lld, \"bc_sum\": %.6f, \"kmax\": %d, "
                "\"half_sum\": %lld, \"add_nnz\": %zu, \"add_sum\": %lld, \"mult_nnz\": %zu, \"mult_sum\": %lld, "
                "\"cancel_nnz\": %zu, \"disj_nnz\": %zu, \"disj_same\": %d, \"disj_err\": %d, \"bc_full\": [",
                TestMtx.m, TestMtx.nz, (long long)tc, (long long)tc_lu, T3.nz, it3, T4.nz, it4, sq_sum, bc_sum,
                kmax, vsum(half), ew_add.nz, vsum(ew_add), ew_mult.nz, vsum(ew_mult), ew_cancel.nz, ew_disj.nz,
                (int)(ew_disj.nz == TestMtx.nz &&
                      std::equal(TestMtx.Col, TestMtx.Col + TestMtx.nz, ew_disj.Col) &&
                      std::equal(TestMtx.Rst, TestMtx.Rst + TestMtx.m + 1, ew_disj.Rst)),
                disj_err);
    for (size_t v = 0; v < bcf.size(); ++v) std::printf(v ? ", %.17g" : "%.17g", bcf[v]);
    std::printf("]}\n");
  } catch (const char* e) {
    std::fprintf(stderr, "error: %s\n", e);
    return 1;
  }
  return 0;
}

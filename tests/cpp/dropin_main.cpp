// Reference-style client of the drop-in headers: the flow of the reference's
// CLI (src/main.cpp:188-277) -- read an .mtx, build the adjacency matrix,
// symmetrise its lower triangle, bind an mxmOp<int>, run triangle counting,
// k-truss, batched BC, full-vertex BC and the element-wise operators, the float instantiations of the BC operators
// (include/betweenness_centrality.h:137, :210) and the dense ones (include/matrix_la.h:192-219) -- compiled against
// include/graphalgo/*.h and linked with libgraphalgo.so.  Prints one JSON object for the test to check.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "graphalgo/algorithms.h"

// --bin <in.bin> <tmp.bin> <truss.out>: the reference's binary CSR cache
// (include/matrix.h:259-277, :441-482): read a graph with spMtx(path, "bin"),
// write it back with write_crs_to_bin, read that copy, and run the triangle
// count and the truss decomposition on it; the truss numbers go to
// truss.out as int32.
static int bin_roundtrip(const char* in, const char* tmp, const char* truss_out) {
  spMtx<int> A(in, "bin");
  if (A.write_crs_to_bin(tmp) != 0) throw "cannot write the bin copy";
  spMtx<int> B(tmp, "bin");
  const int same = A == B ? 1 : 0;
  const int64_t tc = triangle_counting_masked_sandia(B, mxmm_msa<int, int>, true);
  int32_t kmax = 0;
  std::vector<int32_t> truss = truss_decomposition(B, &kmax);
  FILE* f = std::fopen(truss_out, "wb");
  if (!f) throw "cannot write the truss numbers";
  std::fwrite(truss.data(), sizeof(int32_t), truss.size(), f);
  std::fclose(f);
  std::printf("{\"same\": %d, \"n\": %zu, \"nnz\": %zu, \"tc\": %lld, \"kmax\": %d}\n", same, B.m, B.nz,
              (long long)tc, kmax);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: dropin <graph.mtx> | dropin --bin <in.bin> <tmp.bin> <truss.out>\n");
    return 2;
  }
  try {
    if (std::string(argv[1]) == "--bin") {
      if (argc < 5) return 2;
      return bin_roundtrip(argv[2], argv[3], argv[4]);
    }
    spMtx<int> gr(argv[1], "mtx");
    spMtx<int> TestMtx = build_symm_from_lower(extract_lower_triangle(build_adjacency_matrix(gr)));
    mxmOp<int> op = mxmm_msa<int, int>;
    const int64_t tc = triangle_counting_masked_sandia(TestMtx, op, true);
    const int64_t tc_lu = triangle_counting_masked_lu(TestMtx, mxmm_mca<int, int>, true);
    int it3 = 0, it4 = 0;
    spMtx<int> T3 = k_truss(TestMtx, 3, op, true, &it3);
    spMtx<int> T4 = k_truss(TestMtx, 4, mxmm_heap<int>, true, &it4);
    // the reference's own function type (src/main.cpp:90)
    spMtx<int> (*kt)(const spMtx<int>&, int, mxmOp<int>, bool) = k_truss;
    spMtx<int> T5 = kt(TestMtx, 5, op, true);
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
    // float instantiations: C<A> = A*A in FP32, the masked sparse x dense
    // product mxmm_spd<float> and the dense element-wise operators on n x 8
    // arrays (values 1 / (1 + ((7i + 3j) mod 11)) and 1 / (1 + ((5i + j) mod 13)))
    const size_t n = TestMtx.m, nc = 8;
    spMtx<float> Af = convertType<float>(TestMtx), SQf;
    mxmm_msa<float, float>(true, Af, Af, Af, SQf);
    double sqf_sum = 0;
    for (size_t j = 0; j < SQf.nz; ++j) sqf_sum += SQf.Val[j];
    spMtx<int> lvl(n, nc, 2 * n);  // row i: columns i mod 8 and (i / 8) mod 8
    size_t pos = 0;
    for (size_t i = 0; i < n; ++i) {
      const int c1 = (int)(i % nc), c2 = (int)((i / nc) % nc);
      lvl.Rst[i] = (int)pos;
      lvl.Col[pos++] = std::min(c1, c2);
      if (c1 != c2) lvl.Col[pos++] = std::max(c1, c2);
    }
    lvl.Rst[n] = (int)pos;
    lvl.nz = pos;
    for (size_t j = 0; j < pos; ++j) lvl.Val[j] = 1;
    denseMtx<float> Bd(n, nc), Xd(n, nc), Cs, Cbuf;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < nc; ++j) {
        Bd.Val[i * nc + j] = 1.0f / (float)(1 + (7 * i + 3 * j) % 11);
        Xd.Val[i * nc + j] = 1.0f / (float)(1 + (5 * i + j) % 13);
      }
    mxmm_spd(Af, Bd, lvl, Cs, Cbuf);
    denseMtx<float> Wm(n, nc), Acc(Xd);
    eWiseMult(Bd, Xd, lvl, Wm);
    fuseEWiseMultAdd(Wm, Bd, Acc);
    denseMtx<double> Bd64(n, nc), Acc64(n, nc);
    for (size_t i = 0; i < n * nc; ++i) {
      Bd64.Val[i] = Bd.Val[i];
      Acc64.Val[i] = Xd.Val[i];
    }
    fuseEWiseMultAdd(Bd64, Bd64, Acc64);
    auto dump = [](const char* key, const float* v, size_t cnt) {
      std::printf("\"%s\": [", key);
      for (size_t i = 0; i < cnt; ++i) std::printf(i ? ", %.9g" : "%.9g", (double)v[i]);
      std::printf("], ");
    };
    std::printf("{");
    dump("spd_f32", Cs.Val, n * nc);
    dump("wmul_f32", Wm.Val, n * nc);
    dump("fma_f32", Acc.Val, n * nc);
    std::printf("\"fma_f64\": [");
    for (size_t i = 0; i < n * nc; ++i) std::printf(i ? ", %.17g" : "%.17g", Acc64.Val[i]);
    std::printf("], \"k5_nnz\": %zu, \"sqf_sum\": %.1f, ", T5.nz, sqf_sum);
    std::printf("\"n\": %zu, \"nnz\": %zu, \"tc\": %lld, \"tc_lu\": %lld, \"k3_nnz\": %zu, \"k3_it\": %d, "
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

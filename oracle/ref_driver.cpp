// ORACLE — test infrastructure only.  A harness that compiles the
// reference's own headers (include/matrix.h, include/matrix_la.h,
// include/betweenness_centrality.h) and its Matrix Market reader
// (src/mmio.cpp) straight from the reference tree, to produce golden vectors
// and to time the reference's parallel MSA path.  Built into oracle/_ref/ by
// oracle/Makefile; nothing of the reference is copied into this repository.
//
// The reference's CLI (src/main.cpp) cannot be compiled as shipped (it
// includes a header that does not exist, src/main.cpp:8), so this harness
// re-expresses the small loops that live there around the reference's
// library calls:
//   tc      <graph> msa|mca   triangle_counting_masked_sandia, src/main.cpp:65-86
//   ktruss  <graph> k         k_truss, src/main.cpp:89-146 (one k, from scratch)
//   ktall   <graph> out.bin   ktruss for k = 3.. until empty; writes the truss
//                             number of every entry (int32) to out.bin
//   bc      <graph> batch out.bin  betweenness_centrality_batch (float32)
//   mxm     <graph> out.bin   C<A> = A*A with mxmm_msa (int32 values)
//   ewadd   <A> <B> out.bin   eWiseAdd(A, B) (unmasked, values 1), writes
//                             rowptr, col, val (int32) of C
//   ewdisj  <A> <B> out.bin   add_nointersect(A, B), same output
// (ewadd / ewdisj load the operands as they are, without symmetrising)
// <graph> is a Matrix Market file (read with the reference's reader, then
// build_adjacency_matrix, and for tc/ktruss/mxm the symmetrised lower
// triangle, exactly as launch_test does, src/main.cpp:221-222, :335) or a
// binary CSR "*.csr" (int64 n, int64 nnz, int32 rowptr[n+1], int32 col[nnz]).
#include <sstream>
#include <chrono>
#include <cstdio>
#include <vector>
#include "matrix.h"
#include "matrix_la.h"
#include "betweenness_centrality.h"

static spMtx<int> load(const std::string& path) {
  if (path.size() > 4 && path.substr(path.size() - 4) == ".csr") {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw "cannot open csr";
    int64_t n, nnz;
    if (fread(&n, 8, 1, f) != 1 || fread(&nnz, 8, 1, f) != 1) throw "bad csr";
    spMtx<int> A(n, n, nnz);
    if (fread(A.Rst, 4, n + 1, f) != (size_t)(n + 1) || fread(A.Col, 4, nnz, f) != (size_t)nnz) throw "bad csr";
    fclose(f);
    for (int64_t t = 0; t < nnz; ++t) A.Val[t] = 1;
    A.matcode[0] = 'M'; A.matcode[1] = 'C'; A.matcode[2] = 'I'; A.matcode[3] = 'G';
    return A;
  }
  std::string fmt = "mtx";
  // the reference picks int or double from the banner (src/main.cpp:279-310)
  std::ifstream is(path);
  std::string w, type;
  is >> w >> w >> w >> type;
  is.close();
  if (type == "real") {
    spMtx<double> g(path.c_str(), fmt);
    return build_adjacency_matrix(g);
  }
  spMtx<int> g(path.c_str(), fmt);
  return build_adjacency_matrix(g);
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static spMtx<int> ktruss_one(const spMtx<int>& A, int k, int* iters) {
  spMtx<int> C = A, Tmp;
  int n = (int)A.m;
  std::vector<int> rst(n + 1, 0);
  int it = 0;
  for (int t = 0; t < n; ++t) {
    mxmm_msa<int, int>(true, C, C, C, Tmp);
    int pos = 0;
    for (int i = 0; i < n; ++i) {
      for (int j = Tmp.Rst[i]; j < Tmp.Rst[i + 1]; ++j)
        if (Tmp.Val[j] >= k - 2) { Tmp.Col[pos] = Tmp.Col[j]; Tmp.Val[pos++] = 1; }
      rst[i + 1] = pos;
    }
    memcpy(Tmp.Rst, rst.data(), (n + 1) * sizeof(int));
    Tmp.nz = Tmp.Rst[n];
    ++it;
    if (Tmp.nz == C.nz) break;
    std::swap(C, Tmp);
  }
  *iters = it;
  return C;
}

int main(int argc, char** argv) {
  if (argc < 3) { fprintf(stderr, "usage: ref_driver tc|ktruss|ktall|bc|mxm <graph> ...\n"); return 2; }
  std::string mode = argv[1];
  try {
    if (mode == "ewadd" || mode == "ewdisj") {
      if (argc < 5) { fprintf(stderr, "usage: ref_driver ewadd|ewdisj <A> <B> out.bin\n"); return 2; }
      spMtx<int> A = load(argv[2]), B = load(argv[3]);
      spMtx<int> C = mode == "ewadd" ? eWiseAdd(A, B) : add_nointersect(A, B);
      FILE* f = fopen(argv[4], "wb");
      fwrite(C.Rst, sizeof(int), C.m + 1, f);
      fwrite(C.Col, sizeof(int), C.Rst[C.m], f);
      fwrite(C.Val, sizeof(int), C.Rst[C.m], f);
      fclose(f);
      printf("{\"mode\":\"%s\",\"n\":%zu,\"nnz\":%d}\n", mode.c_str(), C.m, C.Rst[C.m]);
      return 0;
    }
    spMtx<int> G = load(argv[2]);
    if (mode == "bc") {
      size_t batch = (size_t)atoll(argv[3]);
      auto t0 = std::chrono::steady_clock::now();
      std::vector<float> bc = betweenness_centrality_batch(true, G, batch);
      double ms = ms_since(t0);
      FILE* f = fopen(argv[4], "wb");
      fwrite(bc.data(), sizeof(float), bc.size(), f);
      fclose(f);
      printf("{\"mode\":\"bc\",\"n\":%zu,\"batch\":%zu,\"ms\":%.3f}\n", G.m, batch, ms);
      return 0;
    }
    spMtx<int> A = build_symm_from_lower(extract_lower_triangle(G));
    if (mode == "tc") {
      std::string alg = argc > 3 ? argv[3] : "msa";
      spMtx<int> L = extract_lower_triangle(A), C;
      auto t0 = std::chrono::steady_clock::now();
      if (alg == "mca") mxmm_mca<int, int>(true, L, L, L, C);
      else mxmm_msa<int, int>(true, L, L, L, C);
      int64_t T = 0;
#pragma omp parallel for reduction(+ : T)
      for (int j = 0; j < C.Rst[C.m]; ++j) T += C.Val[j];
      double ms = ms_since(t0);
      printf("{\"mode\":\"tc\",\"alg\":\"%s\",\"n\":%zu,\"nnz\":%zu,\"triangles\":%lld,\"ms\":%.3f}\n", alg.c_str(),
             A.m, A.nz, (long long)T, ms);
    } else if (mode == "ktruss") {
      int k = atoi(argv[3]), it = 0;
      auto t0 = std::chrono::steady_clock::now();
      spMtx<int> C = ktruss_one(A, k, &it);
      double ms = ms_since(t0);
      printf("{\"mode\":\"ktruss\",\"k\":%d,\"nnz\":%zu,\"iterations\":%d,\"ms\":%.3f}\n", k, C.nz, it, ms);
    } else if (mode == "ktall") {
      // truss number per entry of A: max k whose k-truss holds the entry
      std::vector<int> truss(A.nz, 2);
      double total = 0;
      printf("{\"mode\":\"ktall\",\"n\":%zu,\"nnz\":%zu,\"per_k\":[", A.m, A.nz);
      for (int k = 3;; ++k) {
        int it = 0;
        auto t0 = std::chrono::steady_clock::now();
        spMtx<int> C = ktruss_one(A, k, &it);
        total += ms_since(t0);
        printf("%s[%d,%zu,%d]", k == 3 ? "" : ",", k, C.nz, it);
        if (C.nz == 0) break;
        for (size_t i = 0; i < A.m; ++i) {
          int p = A.Rst[i];
          for (int j = C.Rst[i]; j < C.Rst[i + 1]; ++j) {
            while (A.Col[p] != C.Col[j]) ++p;
            truss[p] = k;
          }
        }
      }
      printf("],\"ms\":%.3f}\n", total);
      FILE* f = fopen(argv[3], "wb");
      fwrite(truss.data(), sizeof(int), truss.size(), f);
      fclose(f);
    } else if (mode == "mxm") {
      spMtx<int> C;
      auto t0 = std::chrono::steady_clock::now();
      mxmm_msa<int, int>(true, A, A, A, C);
      double ms = ms_since(t0);
      FILE* f = fopen(argv[3], "wb");
      fwrite(C.Val, sizeof(int), C.nz, f);
      fclose(f);
      printf("{\"mode\":\"mxm\",\"nnz\":%zu,\"ms\":%.3f}\n", C.nz, ms);
    } else {
      fprintf(stderr, "unknown mode\n");
      return 2;
    }
  } catch (const char* e) {
    fprintf(stderr, "error: %s\n", e);
    return 1;
  }
  return 0;
}

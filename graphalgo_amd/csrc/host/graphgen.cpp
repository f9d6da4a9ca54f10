// Host-side graph substrate: deterministic generators, CSR preparation and a
// Matrix Market reader.  Everything here runs on the CPU and feeds the GPU
// path with symmetric, deduplicated, loop-free CSR graphs whose rows are
// sorted (the input contract of the masked SpGEMM kernels).
//
// Determinism: every random draw is a pure function of (seed, stream,
// counter) through a SplitMix64 finaliser, so a graph is reproduced bit for
// bit regardless of thread count or machine.  The library is built without
// -march=native and with -ffp-contract=off so FP64 arithmetic (kNN
// distances, Box-Muller) rounds identically wherever the .so runs.
//
// Reference input semantics followed:
//   include/matrix.h:280-417  read_mtx_to_crs: diagonal dropped, symmetric
//                             entries mirrored, pattern values set to 1
//   include/matrix.h:736-763  extract_lower_triangle (strict lower part)
//   include/matrix.h:700-734  build_symm_from_lower   (L + L^T)
//   include/matrix_la.h:166-190 transpose (counting sort, rows sorted)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>
#include <zlib.h>

namespace {

thread_local std::string g_err;

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based stream: draw number `ctr` of stream `stream` under `seed`.
struct Rng {
  uint64_t key;
  Rng(uint64_t seed, uint64_t stream) : key(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull))) {}
  inline uint64_t bits(uint64_t ctr) const { return mix64(key + ctr * 0xD1B54A32D192ED03ull); }
  // uniform in [0,1) with 53 random bits
  inline double uni(uint64_t ctr) const { return (double)(bits(ctr) >> 11) * (1.0 / 9007199254740992.0); }
};

// Build a symmetric CSR from directed (u,v) pairs: mirror, drop loops,
// sort each row, drop duplicates.  Output arrays are malloc'd.
int build_symmetric_csr(int64_t n, const std::vector<uint64_t>& packed_pairs,
                        int32_t** rowptr_out, int32_t** col_out, int64_t* nnz_out) {
  // packed pair = (u << 32) | v ; mirror into a second list
  std::vector<uint64_t> e;
  e.reserve(packed_pairs.size() * 2);
  for (uint64_t p : packed_pairs) {
    uint32_t u = (uint32_t)(p >> 32), v = (uint32_t)p;
    if (u == v) continue;
    e.push_back(((uint64_t)u << 32) | v);
    e.push_back(((uint64_t)v << 32) | u);
  }
  // LSD radix sort on the 64-bit keys (two 32-bit halves, 16-bit digits)
  {
    std::vector<uint64_t> tmp(e.size());
    for (int pass = 0; pass < 4; ++pass) {
      int shift = 16 * pass;
      std::vector<size_t> cnt(65537, 0);
      for (uint64_t x : e) cnt[((x >> shift) & 0xFFFF) + 1]++;
      for (int i = 0; i < 65536; ++i) cnt[i + 1] += cnt[i];
      for (uint64_t x : e) tmp[cnt[(x >> shift) & 0xFFFF]++] = x;
      e.swap(tmp);
    }
  }
  e.erase(std::unique(e.begin(), e.end()), e.end());
  int64_t nnz = (int64_t)e.size();
  if (nnz > INT32_MAX) { g_err = "graph too large for int32 CSR"; return 6; }
  int32_t* rp = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  int32_t* cl = (int32_t*)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  if (!rp || !cl) { free(rp); free(cl); g_err = "host allocation failed"; return 4; }
  std::fill(rp, rp + n + 1, 0);
  for (int64_t t = 0; t < nnz; ++t) {
    rp[(e[t] >> 32) + 1]++;
    cl[t] = (int32_t)(uint32_t)e[t];
  }
  for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  *rowptr_out = rp;
  *col_out = cl;
  *nnz_out = nnz;
  return 0;
}

}  // namespace

extern "C" {

const char* gah_last_error(void) { return g_err.c_str(); }

void gah_free(void* p) { free(p); }

// Deterministic raw draws, exposed so tests can pin the stream.
uint64_t gah_rng_bits(uint64_t seed, uint64_t stream, uint64_t ctr) { return Rng(seed, stream).bits(ctr); }

// G(n, m): m distinct unordered pairs {u, v}, u != v, drawn uniformly by
// rejection from the counter stream; result is the symmetric CSR (2m nnz).
int gah_gen_er(int64_t n, int64_t m, uint64_t seed, int32_t** rowptr, int32_t** col, int64_t* nnz) {
  if (n < 2 || m < 0 || m > n * (n - 1) / 2 || n > INT32_MAX) { g_err = "gah_gen_er: bad n/m"; return 1; }
  Rng r(seed, 1);
  std::vector<uint64_t> pairs;
  pairs.reserve(m);
  // open-addressing set of unordered pair keys
  size_t cap = 1;
  while (cap < (size_t)(2 * m + 16)) cap <<= 1;
  std::vector<uint64_t> table(cap, ~0ull);
  uint64_t ctr = 0;
  while ((int64_t)pairs.size() < m) {
    uint64_t u = (uint64_t)(r.uni(ctr++) * (double)n);
    uint64_t v = (uint64_t)(r.uni(ctr++) * (double)n);
    if (u == v) continue;
    uint64_t lo = std::min(u, v), hi = std::max(u, v);
    uint64_t key = (hi << 32) | lo;
    size_t h = (size_t)mix64(key) & (cap - 1);
    bool dup = false;
    while (table[h] != ~0ull) {
      if (table[h] == key) { dup = true; break; }
      h = (h + 1) & (cap - 1);
    }
    if (dup) continue;
    table[h] = key;
    pairs.push_back(key);
  }
  return build_symmetric_csr(n, pairs, rowptr, col, nnz);
}

// R-MAT(scale, edge_factor, a, b, c, d = 1-a-b-c): edge_factor * 2^scale
// directed draws, each descending `scale` quadrant levels with one uniform
// per level; no noise and no label permutation.  The result is symmetrised,
// deduplicated and loop-free.
int gah_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 int32_t** rowptr, int32_t** col, int64_t* n_out, int64_t* nnz) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) {
    g_err = "gah_gen_rmat: bad parameters";
    return 1;
  }
  int64_t n = (int64_t)1 << scale;
  int64_t ne = (int64_t)edge_factor * n;
  Rng r(seed, 2);
  std::vector<uint64_t> pairs(ne);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; ++e) {
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      double x = r.uni((uint64_t)e * (uint64_t)scale + (uint64_t)l);
      uint64_t bu, bv;
      if (x < a) { bu = 0; bv = 0; }
      else if (x < ab) { bu = 0; bv = 1; }
      else if (x < abc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    pairs[e] = (u << 32) | v;
  }
  *n_out = n;
  return build_symmetric_csr(n, pairs, rowptr, col, nnz);
}

// Parameters of kNN graph number `g` of a batch seeded with `base_seed`:
// n ~ U{100..2000}, dim ~ U{4..8}, ncomp ~ U{3..10}.
void gah_knn_params(uint64_t base_seed, int64_t g, int64_t* n, int* dim, int* ncomp) {
  Rng r(base_seed + (uint64_t)g, 3);
  *n = 100 + (int64_t)(r.uni(0) * 1901.0);
  *dim = 4 + (int)(r.uni(1) * 5.0);
  *ncomp = 3 + (int)(r.uni(2) * 8.0);
}

// Points of a Gaussian mixture: component means ~ U(-10, 10)^dim, isotropic
// standard deviation ~ U(0.5, 2); each point picks a component uniformly and
// adds sd * N(0,1) per coordinate (Box-Muller).  Then the exact FP64
// Euclidean k-nearest-neighbour graph (self excluded, ties to lower index),
// symmetrised by union.
int gah_gen_knn(uint64_t seed, int64_t n, int dim, int ncomp, int k,
                int32_t** rowptr, int32_t** col, int64_t* nnz, double* points_out) {
  if (n < 2 || dim < 1 || ncomp < 1 || k < 1 || k >= n) { g_err = "gah_gen_knn: bad parameters"; return 1; }
  Rng r(seed, 4);
  uint64_t ctr = 16;
  std::vector<double> mean((size_t)ncomp * dim), sd(ncomp);
  for (int j = 0; j < ncomp; ++j) {
    for (int d = 0; d < dim; ++d) mean[(size_t)j * dim + d] = -10.0 + 20.0 * r.uni(ctr++);
    sd[j] = 0.5 + 1.5 * r.uni(ctr++);
  }
  std::vector<double> pts((size_t)n * dim);
  for (int64_t p = 0; p < n; ++p) {
    int j = (int)(r.uni(ctr++) * ncomp);
    if (j >= ncomp) j = ncomp - 1;
    for (int d = 0; d < dim; ++d) {
      double u1 = r.uni(ctr++), u2 = r.uni(ctr++);
      double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
      pts[(size_t)p * dim + d] = mean[(size_t)j * dim + d] + sd[j] * z;
    }
  }
  if (points_out) std::memcpy(points_out, pts.data(), sizeof(double) * pts.size());
  std::vector<uint64_t> pairs((size_t)n * k);
  std::vector<std::pair<double, int32_t>> cand(n);
  for (int64_t p = 0; p < n; ++p) {
    int64_t m = 0;
    for (int64_t q = 0; q < n; ++q) {
      if (q == p) continue;
      double s = 0.0;
      for (int d = 0; d < dim; ++d) {
        double df = pts[(size_t)p * dim + d] - pts[(size_t)q * dim + d];
        s += df * df;
      }
      cand[m++] = {s, (int32_t)q};
    }
    std::partial_sort(cand.begin(), cand.begin() + k, cand.begin() + m);
    for (int t = 0; t < k; ++t) pairs[(size_t)p * k + t] = ((uint64_t)p << 32) | (uint32_t)cand[t].second;
  }
  return build_symmetric_csr(n, pairs, rowptr, col, nnz);
}

// Matrix Market coordinate reader with the reference's semantics: 1-based
// indices, diagonal entries dropped, symmetric/skew/hermitian entries
// mirrored, values discarded (pattern = 1).  Rows are sorted and duplicate
// entries merged, which leaves the reference's row order unchanged for files
// that list entries column-major (the SuiteSparse convention).  Reads plain
// or gzip-compressed files.
int gah_read_mtx(const char* path, int64_t* n_out, int32_t** rowptr, int32_t** col, int64_t* nnz,
                 int* symmetric_out) {
  gzFile f = gzopen(path, "rb");
  if (!f) { g_err = std::string("cannot open ") + path; return 1; }
  char line[4096];
  if (!gzgets(f, line, sizeof line)) { gzclose(f); g_err = "empty file"; return 1; }
  std::string banner(line);
  for (auto& ch : banner) ch = (char)tolower(ch);
  if (banner.rfind("%%matrixmarket", 0) != 0 || banner.find("coordinate") == std::string::npos) {
    gzclose(f); g_err = "not a Matrix Market coordinate file"; return 1;
  }
  if (banner.find("complex") != std::string::npos) { gzclose(f); g_err = "complex matrices unsupported"; return 1; }
  bool pattern = banner.find("pattern") != std::string::npos;
  bool symm = banner.find("symmetric") != std::string::npos || banner.find("skew") != std::string::npos ||
              banner.find("hermitian") != std::string::npos;
  long long m = 0, n = 0, ne = 0;
  while (gzgets(f, line, sizeof line)) {
    if (line[0] == '%') continue;
    if (sscanf(line, "%lld %lld %lld", &m, &n, &ne) == 3) break;
  }
  if (m <= 0 || m != n) { gzclose(f); g_err = "not a square matrix"; return 1; }
  std::vector<uint64_t> pairs;
  pairs.reserve(symm ? 2 * ne : ne);
  for (long long t = 0; t < ne; ++t) {
    if (!gzgets(f, line, sizeof line)) { gzclose(f); g_err = "truncated file"; return 1; }
    long long r_, c_;
    if (sscanf(line, "%lld %lld", &r_, &c_) != 2) { --t; continue; }
    (void)pattern;
    uint64_t r0 = (uint64_t)(r_ - 1), c0 = (uint64_t)(c_ - 1);
    if (r0 == c0) continue;
    pairs.push_back((r0 << 32) | c0);
    if (symm) pairs.push_back((c0 << 32) | r0);
  }
  gzclose(f);
  // sort + dedupe, no forced symmetrisation
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  int64_t nz = (int64_t)pairs.size();
  int32_t* rp = (int32_t*)calloc(n + 1, sizeof(int32_t));
  int32_t* cl = (int32_t*)malloc(sizeof(int32_t) * (nz > 0 ? nz : 1));
  for (int64_t t = 0; t < nz; ++t) {
    rp[(pairs[t] >> 32) + 1]++;
    cl[t] = (int32_t)(uint32_t)pairs[t];
  }
  for (long long i = 0; i < n; ++i) rp[i + 1] += rp[i];
  *n_out = n;
  *rowptr = rp;
  *col = cl;
  *nnz = nz;
  if (symmetric_out) *symmetric_out = symm ? 1 : 0;
  return 0;
}

// Symmetrise any pattern: A + A^T with loops dropped and rows sorted.
int gah_symmetrize(int64_t n, const int32_t* rowptr, const int32_t* col, int32_t** rp_out, int32_t** col_out,
                   int64_t* nnz) {
  std::vector<uint64_t> pairs;
  pairs.reserve(rowptr[n]);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t t = rowptr[i]; t < rowptr[i + 1]; ++t) pairs.push_back(((uint64_t)i << 32) | (uint32_t)col[t]);
  return build_symmetric_csr(n, pairs, rp_out, col_out, nnz);
}

}  // extern "C"

// Drop-in container surface of the reference's include/matrix.h for the
// gfx950 library.  spMtx<ValT> keeps the reference's public fields
// (m, n, nz, capacity, matcode, Rst, Col, Val -- include/matrix.h:24-33) and
// ownership rules (owning new[] arrays, callee-resized out-parameters), so
// code written against the reference compiles unchanged; the numerical work
// behind include/graphalgo/matrix_la.h runs on the GPU through the C ABI
// (include/graphalgo.h).
//
// Readers: "mtx" (Matrix Market; diagonal dropped, symmetric entries
// mirrored, pattern values 1 -- include/matrix.h:280-417, values of real
// files kept) and "bin" (the reference's binary CSR cache,
// include/matrix.h:259-277 / :441-482).  Errors are thrown as const char*,
// as the reference does (include/matrix.h:35-51).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../graphalgo.h"

typedef char MM_typecode[4];

template <typename ValT>
class spMtx {
 public:
  size_t m = 0;
  size_t n = 0;
  size_t nz = 0;
  size_t capacity = 0;
  MM_typecode matcode = {'M', 'C', 'I', 'G'};
  int* Rst = nullptr;
  int* Col = nullptr;
  ValT* Val = nullptr;

  spMtx() {}
  spMtx(size_t _m, size_t _n) : m(_m), n(_n) { Rst = new int[m + 1](); }
  spMtx(size_t _m, size_t _n, size_t _nz) : m(_m), n(_n), nz(_nz), capacity(_nz) {
    Rst = new int[m + 1]();
    Col = new int[nz];
    Val = new ValT[nz]();
  }
  spMtx(const char* filename, const std::string& format) {
    int rc = 1;
    if (format == "mtx") rc = read_mtx(filename);
    else if (format == "bin") rc = read_bin(filename);
    else throw "unsupported format";
    if (rc) throw (format == "mtx" ? "Can't read MTX from file" : "Can't read BIN from file");
    capacity = nz;
  }
  spMtx(const spMtx& o) : m(o.m), n(o.n), nz(o.nz), capacity(o.nz) {
    std::memcpy(matcode, o.matcode, 4);
    Rst = new int[m + 1];
    std::memcpy(Rst, o.Rst, (m + 1) * sizeof(int));
    Col = new int[nz > 0 ? nz : 1];
    if (nz) std::memcpy(Col, o.Col, nz * sizeof(int));
    if (o.Val) {
      Val = new ValT[nz > 0 ? nz : 1];
      if (nz) std::memcpy(Val, o.Val, nz * sizeof(ValT));
    }
  }
  spMtx(spMtx&& o) noexcept : m(o.m), n(o.n), nz(o.nz), capacity(o.capacity), Rst(o.Rst), Col(o.Col), Val(o.Val) {
    std::memcpy(matcode, o.matcode, 4);
    o.Rst = nullptr;
    o.Col = nullptr;
    o.Val = nullptr;
  }
  ~spMtx() {
    delete[] Col;
    delete[] Rst;
    delete[] Val;
  }
  spMtx& operator=(const spMtx& o) {
    if (this == &o) return *this;
    spMtx tmp(o);
    swap_with(tmp);
    return *this;
  }
  spMtx& operator=(spMtx&& o) noexcept {
    if (this != &o) swap_with(o);
    return *this;
  }
  void swap_with(spMtx& o) noexcept {
    std::swap(m, o.m);
    std::swap(n, o.n);
    std::swap(nz, o.nz);
    std::swap(capacity, o.capacity);
    for (int i = 0; i < 4; ++i) std::swap(matcode[i], o.matcode[i]);
    std::swap(Rst, o.Rst);
    std::swap(Col, o.Col);
    std::swap(Val, o.Val);
  }

  // Growth without preserving contents, as include/matrix.h:107-127.
  void resizeVals(size_t newNz) {
    if (newNz > capacity || Col == nullptr) {
      delete[] Col;
      delete[] Val;
      Col = new int[newNz > 0 ? newNz : 1];
      Val = new ValT[newNz > 0 ? newNz : 1];
      capacity = newNz;
    }
    nz = newNz;
  }
  void resizeRows(size_t newM) {
    if (m != newM || Rst == nullptr) {
      delete[] Rst;
      Rst = new int[newM + 1]();
      m = newM;
    }
  }
  template <typename ValT2>
  void copyPattern(const spMtx<ValT2>& s) {
    resizeRows(s.m);
    std::memcpy(Rst, s.Rst, (m + 1) * sizeof(int));
    n = s.n;
    resizeVals(s.nz);
    if (nz) std::memcpy(Col, s.Col, nz * sizeof(int));
  }
  spMtx extractRows(size_t begin, size_t end) const {
    spMtx r(end - begin, n, Rst[end] - Rst[begin]);
    for (size_t i = 0; i <= end - begin; ++i) r.Rst[i] = Rst[i + begin] - Rst[begin];
    std::memcpy(r.Col, Col + Rst[begin], r.nz * sizeof(int));
    if (Val) std::memcpy(r.Val, Val + Rst[begin], r.nz * sizeof(ValT));
    std::memcpy(r.matcode, matcode, 4);
    return r;
  }
  bool operator==(const spMtx& o) const {
    if (m != o.m || n != o.n || nz != o.nz) return false;
    for (size_t i = 0; i <= m; ++i)
      if (Rst[i] != o.Rst[i]) return false;
    for (size_t j = 0; j < nz; ++j)
      if (Col[j] != o.Col[j] || (Val && o.Val && Val[j] != o.Val[j])) return false;
    return true;
  }
  // A C ABI view (pattern view when values are not requested).
  ga_csr view(bool with_values) const {
    return ga_csr{(int64_t)m, (int64_t)n, (int64_t)nz, Rst, Col, with_values ? (const void*)Val : nullptr};
  }

  int write_crs_to_bin(const char* filename) const {
    FILE* fp = std::fopen(filename, "wb");
    if (!fp) return -1;
    std::fwrite(matcode, 1, 4, fp);
    std::fwrite(&m, sizeof(size_t), 1, fp);
    std::fwrite(&n, sizeof(size_t), 1, fp);
    std::fwrite(&nz, sizeof(size_t), 1, fp);
    std::fwrite(Rst, sizeof(int), m + 1, fp);
    std::fwrite(Col, sizeof(int), nz, fp);
    std::fwrite(Val, sizeof(ValT), nz, fp);
    std::fclose(fp);
    return 0;
  }

 private:
  int read_bin(const char* filename) {
    FILE* fp = std::fopen(filename, "rb");
    if (!fp) return -1;
    bool ok = std::fread(matcode, 1, 4, fp) == 4 && std::fread(&m, sizeof(size_t), 1, fp) == 1 &&
              std::fread(&n, sizeof(size_t), 1, fp) == 1 && std::fread(&nz, sizeof(size_t), 1, fp) == 1;
    if (ok) {
      Rst = new int[m + 1];
      Col = new int[nz > 0 ? nz : 1];
      Val = new ValT[nz > 0 ? nz : 1];
      ok = std::fread(Rst, sizeof(int), m + 1, fp) == m + 1 && std::fread(Col, sizeof(int), nz, fp) == nz &&
           std::fread(Val, sizeof(ValT), nz, fp) == nz;
    }
    std::fclose(fp);
    return ok ? 0 : -3;
  }
  int read_mtx(const char* filename) {
    std::ifstream in(filename);
    if (!in) return 1;
    std::string line;
    std::getline(in, line);
    std::string low = line;
    for (auto& ch : low) ch = (char)std::tolower(ch);
    if (low.rfind("%%matrixmarket", 0) != 0 || low.find("coordinate") == std::string::npos) return 1;
    const bool pattern = low.find("pattern") != std::string::npos;
    const bool symm = low.find("symmetric") != std::string::npos || low.find("hermitian") != std::string::npos ||
                      low.find("skew") != std::string::npos;
    matcode[0] = 'M';
    matcode[1] = 'C';
    matcode[2] = pattern ? 'P' : (low.find("real") != std::string::npos ? 'R' : 'I');
    matcode[3] = symm ? 'S' : 'G';
    while (std::getline(in, line) && !line.empty() && line[0] == '%') {
    }
    long long mm = 0, nn = 0, ne = 0;
    std::istringstream hdr(line);
    hdr >> mm >> nn >> ne;
    if (mm <= 0 || mm != nn) return 1;
    std::vector<std::pair<long long, std::pair<int, ValT>>> e;
    e.reserve(symm ? 2 * ne : ne);
    for (long long t = 0; t < ne; ++t) {
      long long r, c;
      double v = 1.0;
      if (!(in >> r >> c)) return 1;
      if (!pattern) in >> v;
      if (r == c) continue;  // loops are dropped
      e.push_back({(r - 1) * nn + (c - 1), {0, pattern ? ValT(1) : (ValT)v}});
      if (symm) e.push_back({(c - 1) * nn + (r - 1), {0, pattern ? ValT(1) : (ValT)v}});
    }
    std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    m = (size_t)mm;
    n = (size_t)nn;
    nz = e.size();
    Rst = new int[m + 1]();
    Col = new int[nz > 0 ? nz : 1];
    Val = new ValT[nz > 0 ? nz : 1];
    for (size_t t = 0; t < nz; ++t) {
      Rst[e[t].first / nn + 1]++;
      Col[t] = (int)(e[t].first % nn);
      Val[t] = e[t].second.second;
    }
    for (size_t i = 0; i < m; ++i) Rst[i + 1] += Rst[i];
    return 0;
  }
};

template <typename ValT>
class denseMtx {
 public:
  size_t m = 0;
  size_t n = 0;
  size_t capacity = 0;
  ValT* Val = nullptr;

  denseMtx() {}
  denseMtx(size_t _m, size_t _n) : m(_m), n(_n), capacity(_m * _n) { Val = new ValT[capacity > 0 ? capacity : 1](); }
  denseMtx(const denseMtx& o) : m(o.m), n(o.n), capacity(o.m * o.n) {
    Val = new ValT[capacity > 0 ? capacity : 1];
    std::memcpy(Val, o.Val, capacity * sizeof(ValT));
  }
  denseMtx(denseMtx&& o) noexcept : m(o.m), n(o.n), capacity(o.capacity), Val(o.Val) { o.Val = nullptr; }
  // Zero-filled densification (the reference leaves unset entries
  // uninitialised in operator=, include/matrix.h:613-631).
  template <typename ValT2>
  explicit denseMtx(const spMtx<ValT2>& s) : m(s.m), n(s.n), capacity(s.m * s.n) {
    Val = new ValT[capacity > 0 ? capacity : 1]();
    for (size_t i = 0; i < s.m; ++i)
      for (int k = s.Rst[i]; k < s.Rst[i + 1]; ++k) Val[n * i + s.Col[k]] = s.Val ? (ValT)s.Val[k] : ValT(1);
  }
  denseMtx& operator=(const denseMtx& o) {
    if (this != &o) {
      denseMtx t(o);
      std::swap(m, t.m);
      std::swap(n, t.n);
      std::swap(capacity, t.capacity);
      std::swap(Val, t.Val);
    }
    return *this;
  }
  denseMtx& operator=(denseMtx&& o) noexcept {
    std::swap(m, o.m);
    std::swap(n, o.n);
    std::swap(capacity, o.capacity);
    std::swap(Val, o.Val);
    return *this;
  }
  ~denseMtx() { delete[] Val; }
  void resize(size_t newM, size_t newN) {
    m = newM;
    n = newN;
    if (capacity < m * n) {
      delete[] Val;
      capacity = m * n;
      Val = new ValT[capacity]();
    }
  }
};

template <typename ValT2, typename ValT1>
spMtx<ValT2> convertType(const spMtx<ValT1>& s) {
  spMtx<ValT2> r;
  r.copyPattern(s);
  for (size_t j = 0; j < s.nz; ++j) r.Val[j] = s.Val ? (ValT2)s.Val[j] : ValT2(1);
  return r;
}

// Pattern of Gr with values 1 (include/matrix.h:679-698).
template <typename T>
spMtx<int> build_adjacency_matrix(const spMtx<T>& Gr) {
  spMtx<int> r(Gr.m, Gr.n, Gr.nz);
  std::memcpy(r.Rst, Gr.Rst, (Gr.m + 1) * sizeof(int));
  if (Gr.nz) std::memcpy(r.Col, Gr.Col, Gr.nz * sizeof(int));
  std::memcpy(r.matcode, Gr.matcode, 4);
  r.matcode[2] = 'I';
  for (size_t j = 0; j < Gr.nz; ++j) r.Val[j] = 1;
  return r;
}

// Strict lower triangle of a matrix with sorted rows (include/matrix.h:736-763).
template <typename T>
spMtx<T> extract_lower_triangle(const spMtx<T>& Gr) {
  spMtx<T> r(Gr.m, Gr.n);
  for (size_t i = 0; i < Gr.m; ++i) {
    int t = Gr.Rst[i];
    while (t < Gr.Rst[i + 1] && (size_t)Gr.Col[t] < i) ++t;
    r.Rst[i + 1] = r.Rst[i] + (t - Gr.Rst[i]);
  }
  r.resizeVals(r.Rst[r.m]);
  for (size_t i = 0; i < Gr.m; ++i) {
    const int len = r.Rst[i + 1] - r.Rst[i];
    std::memcpy(r.Col + r.Rst[i], Gr.Col + Gr.Rst[i], len * sizeof(int));
    for (int q = 0; q < len; ++q) r.Val[r.Rst[i] + q] = Gr.Val ? Gr.Val[Gr.Rst[i] + q] : T(1);
  }
  return r;
}

// Strict upper triangle (include/matrix.h:765-793).
template <typename T>
spMtx<T> extract_upper_triangle(const spMtx<T>& Gr) {
  spMtx<T> r(Gr.m, Gr.n);
  std::vector<int> first(Gr.m);
  for (size_t i = 0; i < Gr.m; ++i) {
    int t = Gr.Rst[i];
    while (t < Gr.Rst[i + 1] && (size_t)Gr.Col[t] <= i) ++t;
    first[i] = t;
    r.Rst[i + 1] = r.Rst[i] + (Gr.Rst[i + 1] - t);
  }
  r.resizeVals(r.Rst[r.m]);
  for (size_t i = 0; i < Gr.m; ++i) {
    const int len = r.Rst[i + 1] - r.Rst[i];
    std::memcpy(r.Col + r.Rst[i], Gr.Col + first[i], len * sizeof(int));
    for (int q = 0; q < len; ++q) r.Val[r.Rst[i] + q] = Gr.Val ? Gr.Val[first[i] + q] : T(1);
  }
  return r;
}

// Transpose with sorted rows (include/matrix_la.h:166-190).
template <typename T>
spMtx<T> transpose(const spMtx<T>& A) {
  spMtx<T> AT(A.n, A.m, A.nz);
  for (size_t t = 0; t < A.nz; ++t) ++AT.Rst[A.Col[t] + 1];
  for (size_t i = 0; i < AT.m; ++i) AT.Rst[i + 1] += AT.Rst[i];
  std::vector<int> pos(AT.Rst, AT.Rst + AT.m);
  for (size_t i = 0; i < A.m; ++i)
    for (int t = A.Rst[i]; t < A.Rst[i + 1]; ++t) {
      const int p = pos[A.Col[t]]++;
      AT.Col[p] = (int)i;
      AT.Val[p] = A.Val ? A.Val[t] : T(1);
    }
  return AT;
}

// L + L^T (include/matrix.h:700-734).
template <typename T>
spMtx<T> build_symm_from_lower(const spMtx<T>& Low) {
  spMtx<T> Upp = transpose(Low);
  spMtx<T> r(Low.m, Low.n, 2 * Low.nz);
  size_t jl = 0, ju = 0, jr = 0;
  for (size_t i = 0; i < Low.m; ++i) {
    const size_t xl = Low.Rst[i + 1], xu = Upp.Rst[i + 1];
    while (jl < xl || ju < xu) {
      if (ju >= xu || (jl < xl && Low.Col[jl] < Upp.Col[ju])) {
        r.Col[jr] = Low.Col[jl];
        r.Val[jr++] = Low.Val ? Low.Val[jl] : T(1);
        ++jl;
      } else {
        r.Col[jr] = Upp.Col[ju];
        r.Val[jr++] = Upp.Val[ju++];
      }
    }
    r.Rst[i + 1] = (int)jr;
  }
  r.nz = jr;
  return r;
}

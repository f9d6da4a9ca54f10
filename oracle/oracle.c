/*
 * ORACLE — test infrastructure only.  CPU restatement (C11 + OpenMP) of the
 * reference algorithms on the masked-SpGEMM path.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load this library, and only as
 * the checker / CPU baseline; the product path never calls it.
 *
 * Every function names the reference lines it restates.  Deliberate
 * departures from the reference (its defects D1-D6) are listed per function:
 * parity is with the mathematically correct result, which is what the
 * reference's parallel MSA and MCA paths compute.
 *
 * Conventions: CSR with int32 row pointers and column indices; rows sorted
 * ascending; a NULL value pointer means a pattern matrix (all values 1).
 * Integer products are accumulated in int64 (the reference accumulates in the
 * matrix value type, int, and sums triangle counts into int64).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_OOM 4

static int32_t *dup_i32(const int32_t *src, int64_t n) {
  int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  if (d && n > 0) memcpy(d, src, sizeof(int32_t) * n);
  return d;
}

void orc_free(void *p) { free(p); }

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* Masked SpGEMM C<M> = A*B, structural mask, MSA accumulator.
 * Restates _mxmm_msa_parallel, include/matrix_la.h:715-749 (and the wrapper
 * mxmm_msa :700-713: C takes M's pattern, explicit zeros kept).  Per thread a
 * dense accumulator of length B.n; per row: zero acc at the mask columns
 * (:728-729), scatter-add a*b over every B row selected by A's row
 * (:732-740), gather acc at the mask columns (:743-746).
 * int64 values; Aval/Bval may be NULL (pattern = 1). */
int orc_mxm_masked_i64(int64_t m, int64_t bn,
                       const int32_t *Arp, const int32_t *Acol, const int32_t *Aval,
                       const int32_t *Brp, const int32_t *Bcol, const int32_t *Bval,
                       const int32_t *Mrp, const int32_t *Mcol, int64_t *Cval) {
  int err = ORC_OK;
#pragma omp parallel
  {
    int64_t *acc = (int64_t *)malloc(sizeof(int64_t) * (bn > 0 ? bn : 1));
    if (!acc) {
#pragma omp atomic write
      err = ORC_ERR_OOM;
    }
#pragma omp for schedule(dynamic, 32)
    for (int64_t i = 0; i < m; ++i) {
      if (!acc) continue;
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) acc[Mcol[j]] = 0;
      for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) {
        int32_t k = Acol[t];
        int64_t a = Aval ? Aval[t] : 1;
        for (int32_t j = Brp[k]; j < Brp[k + 1]; ++j) acc[Bcol[j]] += a * (Bval ? Bval[j] : 1);
      }
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) Cval[j] = acc[Mcol[j]];
    }
    free(acc);
  }
  return err;
}

/* Same restatement with FP64 values (A and B values required or NULL). */
int orc_mxm_masked_f64(int64_t m, int64_t bn,
                       const int32_t *Arp, const int32_t *Acol, const double *Aval,
                       const int32_t *Brp, const int32_t *Bcol, const double *Bval,
                       const int32_t *Mrp, const int32_t *Mcol, double *Cval) {
  int err = ORC_OK;
#pragma omp parallel
  {
    double *acc = (double *)malloc(sizeof(double) * (bn > 0 ? bn : 1));
    if (!acc) {
#pragma omp atomic write
      err = ORC_ERR_OOM;
    }
#pragma omp for schedule(dynamic, 32)
    for (int64_t i = 0; i < m; ++i) {
      if (!acc) continue;
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) acc[Mcol[j]] = 0.0;
      for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) {
        int32_t k = Acol[t];
        double a = Aval ? Aval[t] : 1.0;
        for (int32_t j = Brp[k]; j < Brp[k + 1]; ++j) acc[Bcol[j]] += a * (Bval ? Bval[j] : 1.0);
      }
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) Cval[j] = acc[Mcol[j]];
    }
    free(acc);
  }
  return err;
}

/* Wedge count W = sum_{(i,k) in A} nnz(B_k): the unit of work of Gustavson's
 * row-wise product (the hot loop include/matrix_la.h:732-740 runs W times). */
int64_t orc_wedges(int64_t m, const int32_t *Arp, const int32_t *Acol, const int32_t *Brp) {
  int64_t w = 0;
#pragma omp parallel for reduction(+ : w) schedule(static)
  for (int64_t i = 0; i < m; ++i)
    for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) w += Brp[Acol[t] + 1] - Brp[Acol[t]];
  return w;
}

/* Strict lower triangle, include/matrix.h:736-763 (assumes sorted rows,
 * :747).  Output arrays malloc'd. */
int orc_lower(int64_t n, const int32_t *rp, const int32_t *col, int32_t **Lrp, int32_t **Lcol, int64_t *Lnnz) {
  int32_t *r = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  if (!r) return ORC_ERR_OOM;
  r[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t t = rp[i];
    while (t < rp[i + 1] && col[t] < i) ++t;
    r[i + 1] = r[i] + (t - rp[i]);
  }
  int32_t *c = (int32_t *)malloc(sizeof(int32_t) * (r[n] > 0 ? r[n] : 1));
  if (!c) { free(r); return ORC_ERR_OOM; }
  for (int64_t i = 0; i < n; ++i) memcpy(c + r[i], col + rp[i], sizeof(int32_t) * (r[i + 1] - r[i]));
  *Lrp = r;
  *Lcol = c;
  *Lnnz = r[n];
  return ORC_OK;
}

/* Transpose of a pattern, include/matrix_la.h:166-190 (counting sort, so
 * the rows of the result are sorted). */
int orc_transpose(int64_t m, int64_t n, const int32_t *rp, const int32_t *col, int32_t **Trp, int32_t **Tcol) {
  int64_t nnz = rp[m];
  int32_t *r = (int32_t *)calloc(n + 1, sizeof(int32_t));
  int32_t *c = (int32_t *)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  if (!r || !c) { free(r); free(c); return ORC_ERR_OOM; }
  for (int64_t t = 0; t < nnz; ++t) r[col[t] + 1]++;
  for (int64_t i = 0; i < n; ++i) r[i + 1] += r[i];
  int32_t *pos = dup_i32(r, n);
  for (int64_t i = 0; i < m; ++i)
    for (int32_t t = rp[i]; t < rp[i + 1]; ++t) c[pos[col[t]]++] = (int32_t)i;
  free(pos);
  *Trp = r;
  *Tcol = c;
  return ORC_OK;
}

/* Triangle count, triangle_counting_masked_sandia, src/main.cpp:65-86:
 * L = tril(A,-1) (:67), C<L> = L*L (:73), T = sum C (:76-78) in int64.
 * Input: symmetric pattern with sorted rows.  Returns -1 on error. */
int64_t orc_tc_sandia(int64_t n, const int32_t *rp, const int32_t *col) {
  int32_t *Lrp, *Lcol;
  int64_t Lnnz;
  if (orc_lower(n, rp, col, &Lrp, &Lcol, &Lnnz)) return -1;
  int64_t *C = (int64_t *)malloc(sizeof(int64_t) * (Lnnz > 0 ? Lnnz : 1));
  if (!C) { free(Lrp); free(Lcol); return -1; }
  int64_t T = -1;
  if (orc_mxm_masked_i64(n, n, Lrp, Lcol, NULL, Lrp, Lcol, NULL, Lrp, Lcol, C) == ORC_OK) {
    T = 0;
#pragma omp parallel for reduction(+ : T)
    for (int64_t j = 0; j < Lnnz; ++j) T += C[j];
  }
  free(C); free(Lrp); free(Lcol);
  return T;
}

/* triangle_counting_masked_lu, src/main.cpp:40-62: C<A> = L*U with U = L^T,
 * T = (sum C) / 2. */
int64_t orc_tc_lu(int64_t n, const int32_t *rp, const int32_t *col) {
  int32_t *Lrp, *Lcol, *Urp, *Ucol;
  int64_t Lnnz;
  if (orc_lower(n, rp, col, &Lrp, &Lcol, &Lnnz)) return -1;
  if (orc_transpose(n, n, Lrp, Lcol, &Urp, &Ucol)) { free(Lrp); free(Lcol); return -1; }
  int64_t nnz = rp[n];
  int64_t *C = (int64_t *)malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  int64_t T = -1;
  if (C && orc_mxm_masked_i64(n, n, Lrp, Lcol, NULL, Urp, Ucol, NULL, rp, col, C) == ORC_OK) {
    T = 0;
    for (int64_t j = 0; j < nnz; ++j) T += C[j];
    T >>= 1;
  }
  free(C); free(Lrp); free(Lcol); free(Urp); free(Ucol);
  return T;
}

/* triangle_counting_vertex, src/main.cpp:14-37: SQ<A> = A*A, per-vertex
 * triangles = (row sum of SQ) / 2. */
int orc_tc_vertex(int64_t n, const int32_t *rp, const int32_t *col, int64_t *per_vertex) {
  int64_t nnz = rp[n];
  int64_t *C = (int64_t *)malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  if (!C) return ORC_ERR_OOM;
  int e = orc_mxm_masked_i64(n, n, rp, col, NULL, rp, col, NULL, rp, col, C);
  if (!e) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      int64_t s = 0;
      for (int32_t j = rp[i]; j < rp[i + 1]; ++j) s += C[j];
      per_vertex[i] = s >> 1;
    }
  }
  free(C);
  return e;
}

/* Keep the entries whose support is >= thr (values become 1), compacting
 * rows; optional edge-id array carried along.  src/main.cpp:106-117. */
static int64_t filter_support(int64_t n, int32_t *rp, int32_t *col, int32_t *eid, const int64_t *sup, int64_t thr) {
  int64_t pos = 0;
  int32_t start = rp[0];
  for (int64_t i = 0; i < n; ++i) {
    int32_t end = rp[i + 1];
    for (int32_t j = start; j < end; ++j) {
      if (sup[j] >= thr) {
        col[pos] = col[j];
        if (eid) eid[pos] = eid[j];
        ++pos;
      }
    }
    start = end;
    rp[i + 1] = (int32_t)pos;
  }
  return pos;
}

/* k-truss for one k, src/main.cpp:89-146: C = A; repeat { Tmp<C> = C*C;
 * keep Tmp >= k-2 as 1; stop when nnz is unchanged }.  Output: the k-truss
 * pattern (malloc'd) and the number of passes (totalIterationNum). */
int orc_ktruss(int64_t n, const int32_t *rp, const int32_t *col, int k,
               int32_t **out_rp, int32_t **out_col, int64_t *out_nnz, int *iterations) {
  int32_t *Crp = dup_i32(rp, n + 1);
  int32_t *Ccol = dup_i32(col, rp[n]);
  int64_t *sup = (int64_t *)malloc(sizeof(int64_t) * (rp[n] > 0 ? rp[n] : 1));
  if (!Crp || !Ccol || !sup) { free(Crp); free(Ccol); free(sup); return ORC_ERR_OOM; }
  int it = 0;
  for (int64_t t = 0; t < n; ++t) {
    int64_t nz = Crp[n];
    int e = orc_mxm_masked_i64(n, n, Crp, Ccol, NULL, Crp, Ccol, NULL, Crp, Ccol, sup);
    if (e) { free(Crp); free(Ccol); free(sup); return e; }
    int64_t nz2 = filter_support(n, Crp, Ccol, NULL, sup, (int64_t)k - 2);
    ++it;
    if (nz2 == nz) break;
  }
  free(sup);
  *out_rp = Crp;
  *out_col = Ccol;
  *out_nnz = Crp[n];
  if (iterations) *iterations = it;
  return ORC_OK;
}

/* Truss decomposition: truss number of every stored entry of the symmetric
 * pattern A, i.e. the largest k whose k-truss (as computed by
 * src/main.cpp:89-146) contains the edge, or 2 if none.  The k-truss is the
 * unique maximal subgraph in which every edge closes >= k-2 triangles, so
 * peeling for k may start from the (k-1)-truss instead of from A (the
 * reference restarts from A for every k, src/main.cpp:91); the support of an
 * unchanged graph is reused for the next k.
 * Outputs: truss[nnz(A)], *kmax (largest truss number), *passes (support
 * products computed), pass_nnz/pass_wedges (per pass, caller-sized arrays of
 * length max_passes, may be NULL). */
int orc_truss_decomposition(int64_t n, const int32_t *rp, const int32_t *col, int32_t *truss, int32_t *kmax,
                            int64_t *passes, int64_t *pass_nnz, int64_t *pass_wedges, int64_t max_passes) {
  int64_t nnz0 = rp[n];
  int32_t *Crp = dup_i32(rp, n + 1);
  int32_t *Ccol = dup_i32(col, nnz0);
  int32_t *eid = (int32_t *)malloc(sizeof(int32_t) * (nnz0 > 0 ? nnz0 : 1));
  int64_t *sup = (int64_t *)malloc(sizeof(int64_t) * (nnz0 > 0 ? nnz0 : 1));
  if (!Crp || !Ccol || !eid || !sup) { free(Crp); free(Ccol); free(eid); free(sup); return ORC_ERR_OOM; }
  for (int64_t t = 0; t < nnz0; ++t) { eid[t] = (int32_t)t; truss[t] = 2; }
  int64_t np = 0;
  int k = 3;
  int valid = 0;
  int32_t kbest = 2;
  while (Crp[n] > 0) {
    int64_t nz = Crp[n];
    if (!valid) {
      if (pass_nnz && np < max_passes) pass_nnz[np] = nz;
      if (pass_wedges && np < max_passes) pass_wedges[np] = orc_wedges(n, Crp, Ccol, Crp);
      int e = orc_mxm_masked_i64(n, n, Crp, Ccol, NULL, Crp, Ccol, NULL, Crp, Ccol, sup);
      if (e) { free(Crp); free(Ccol); free(eid); free(sup); return e; }
      ++np;
      valid = 1;
    }
    /* count survivors first so an unchanged graph keeps its support */
    int64_t keep = 0;
#pragma omp parallel for reduction(+ : keep)
    for (int64_t j = 0; j < nz; ++j) keep += sup[j] >= (int64_t)k - 2;
    if (keep == nz) {
      for (int64_t j = 0; j < nz; ++j) truss[eid[j]] = k;
      kbest = k;
      ++k;
      continue;
    }
    filter_support(n, Crp, Ccol, eid, sup, (int64_t)k - 2);
    valid = 0;
  }
  free(Crp); free(Ccol); free(eid); free(sup);
  if (kmax) *kmax = kbest;
  if (passes) *passes = np;
  return ORC_OK;
}

/* Complement-masked SpGEMM C<!M> = A*B with FP64 values, two phases as in
 * _mxmm_msa_cmask_parallel, include/matrix_la.h:800-891: a symbolic pass
 * counts, per row, the distinct product columns not in M (:808-837), a prefix
 * sum sizes C (:838-846), and a numeric pass accumulates and emits the
 * columns in ascending order (:852-889).  C arrays are malloc'd.
 * Bval may be NULL (pattern); Aval may be NULL. */
int orc_mxm_cmask_f64(int64_t m, int64_t bn,
                      const int32_t *Arp, const int32_t *Acol, const double *Aval,
                      const int32_t *Brp, const int32_t *Bcol, const double *Bval,
                      const int32_t *Mrp, const int32_t *Mcol,
                      int32_t **Crp_out, int32_t **Ccol_out, double **Cval_out, int64_t *Cnnz) {
  int32_t *Crp = (int32_t *)calloc(m + 1, sizeof(int32_t));
  if (!Crp) return ORC_ERR_OOM;
  int err = ORC_OK;
#pragma omp parallel
  {
    /* state: 0 = free, 1 = masked (in M), 2 = seen */
    char *state = (char *)calloc(bn > 0 ? bn : 1, 1);
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (bn > 0 ? bn : 1));
    if (!state || !touched) {
#pragma omp atomic write
      err = ORC_ERR_OOM;
    }
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; ++i) {
      if (!state || !touched) continue;
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) state[Mcol[j]] = 1;
      int32_t cnt = 0;
      for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) {
        int32_t k = Acol[t];
        for (int32_t j = Brp[k]; j < Brp[k + 1]; ++j)
          if (state[Bcol[j]] == 0) { state[Bcol[j]] = 2; touched[cnt++] = Bcol[j]; }
      }
      for (int32_t q = 0; q < cnt; ++q) state[touched[q]] = 0;
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) state[Mcol[j]] = 0;
      Crp[i + 1] = cnt;
    }
    free(state);
    free(touched);
  }
  if (err) { free(Crp); return err; }
  for (int64_t i = 0; i < m; ++i) Crp[i + 1] += Crp[i];
  int64_t nz = Crp[m];
  int32_t *Ccol = (int32_t *)malloc(sizeof(int32_t) * (nz > 0 ? nz : 1));
  double *Cval = (double *)malloc(sizeof(double) * (nz > 0 ? nz : 1));
  if (!Ccol || !Cval) { free(Crp); free(Ccol); free(Cval); return ORC_ERR_OOM; }
#pragma omp parallel
  {
    char *state = (char *)calloc(bn > 0 ? bn : 1, 1);
    double *acc = (double *)calloc(bn > 0 ? bn : 1, sizeof(double));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (bn > 0 ? bn : 1));
    if (!state || !acc || !touched) {
#pragma omp atomic write
      err = ORC_ERR_OOM;
    }
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < m; ++i) {
      if (!state || !acc || !touched) continue;
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) state[Mcol[j]] = 1;
      int32_t cnt = 0;
      for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) {
        int32_t k = Acol[t];
        double a = Aval ? Aval[t] : 1.0;
        for (int32_t j = Brp[k]; j < Brp[k + 1]; ++j) {
          int32_t c = Bcol[j];
          double v = a * (Bval ? Bval[j] : 1.0);
          if (state[c] == 0) { state[c] = 2; touched[cnt++] = c; acc[c] = v; }
          else if (state[c] == 2) acc[c] += v;
        }
      }
      /* ascending column order, as the reference's sort (:881) */
      qsort(touched, (size_t)cnt, sizeof(int32_t), cmp_i32);
      int32_t pos = Crp[i];
      for (int32_t q = 0; q < cnt; ++q) {
        Ccol[pos] = touched[q];
        Cval[pos++] = acc[touched[q]];
        acc[touched[q]] = 0.0;
        state[touched[q]] = 0;
      }
      for (int32_t j = Mrp[i]; j < Mrp[i + 1]; ++j) state[Mcol[j]] = 0;
    }
    free(state); free(acc); free(touched);
  }
  if (err) { free(Crp); free(Ccol); free(Cval); return err; }
  *Crp_out = Crp;
  *Ccol_out = Ccol;
  *Cval_out = Cval;
  *Cnnz = nz;
  return ORC_OK;
}

/* Masked SpMM, mxmm_spd, include/matrix_la.h:239-270: B and C dense
 * row-major (rows x ncols); C row i is zeroed, then for every A entry (i,j)
 * c[M.Col] += a * B[j][M.Col] over the mask columns of row i.  Aval may be
 * NULL. */
void orc_mxm_spd_f64(int64_t m, int64_t ncols,
                     const int32_t *Arp, const int32_t *Acol, const double *Aval,
                     const double *B, const int32_t *Mrp, const int32_t *Mcol, double *C) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; ++i) {
    double *c = C + (size_t)i * ncols;
    for (int64_t q = 0; q < ncols; ++q) c[q] = 0.0;
    for (int32_t t = Arp[i]; t < Arp[i + 1]; ++t) {
      double a = Aval ? Aval[t] : 1.0;
      const double *b = B + (size_t)Acol[t] * ncols;
      for (int32_t q = Mrp[i]; q < Mrp[i + 1]; ++q) c[Mcol[q]] += a * b[Mcol[q]];
    }
  }
}

/* Sparse union of two CSR matrices with disjoint patterns (add_nointersect,
 * include/matrix_la.h:337-372), FP64 values.  Output malloc'd. */
static int union_disjoint(int64_t m, const int32_t *Arp, const int32_t *Acol, const double *Aval,
                          const int32_t *Brp, const int32_t *Bcol, const double *Bval,
                          int32_t **Crp_o, int32_t **Ccol_o, double **Cval_o) {
  int32_t *Crp = (int32_t *)malloc(sizeof(int32_t) * (m + 1));
  if (!Crp) return ORC_ERR_OOM;
  Crp[0] = 0;
  for (int64_t i = 0; i < m; ++i) Crp[i + 1] = Crp[i] + (Arp[i + 1] - Arp[i]) + (Brp[i + 1] - Brp[i]);
  int64_t nz = Crp[m];
  int32_t *Ccol = (int32_t *)malloc(sizeof(int32_t) * (nz > 0 ? nz : 1));
  double *Cval = (double *)malloc(sizeof(double) * (nz > 0 ? nz : 1));
  if (!Ccol || !Cval) { free(Crp); free(Ccol); free(Cval); return ORC_ERR_OOM; }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; ++i) {
    int32_t a = Arp[i], b = Brp[i], c = Crp[i];
    while (a < Arp[i + 1] || b < Brp[i + 1]) {
      if (b >= Brp[i + 1] || (a < Arp[i + 1] && Acol[a] < Bcol[b])) { Ccol[c] = Acol[a]; Cval[c++] = Aval[a++]; }
      else { Ccol[c] = Bcol[b]; Cval[c++] = Bval[b++]; }
    }
  }
  *Crp_o = Crp; *Ccol_o = Ccol; *Cval_o = Cval;
  return ORC_OK;
}

/* Batched betweenness centrality, FP64 restatement of
 * betweenness_centrality_batch, include/betweenness_centrality.h:125-237:
 *   AT = A^T (:133); frontier F (n x ns) starts as A's source rows,
 *   transposed (:173); the visited set Numsp starts as (s_j, j) = 1
 *   (:166-172); forward: Sigmas[d] = F; Numsp = Numsp U F;
 *   F = (AT * F)<!Numsp> until F is empty (:177-192);
 *   backward for k = d-1 .. 1: W<Sigmas[k]> = (1 + delta) / sigma,
 *   W' <Sigmas[k-1]> = A * W (mxmm_spd), delta += W' .* sigma (:203-218);
 *   bc[v] = sum_j delta(v, j) (:221-228).
 * Departures: sources are an explicit list (the reference uses 0..ns-1);
 * sigma and delta are FP64 (reference: int32 and float32, defect D5); pairs
 * never reached have sigma = delta = 0 (defect D4: the reference reads
 * uninitialised memory there); delta is kept separately instead of as
 * Bcu = 1 + delta, and bc = sum(delta) instead of sum(1 + delta) - ns, which
 * is the same quantity without the cancellation.
 * Output bc[n]; *depth_out = number of BFS levels d (may be NULL). */
int orc_bc_batch(int64_t n, const int32_t *rp, const int32_t *col, const int32_t *sources, int64_t ns,
                 double *bc, int32_t *depth_out) {
  int err = ORC_OK;
  for (int64_t j = 0; j < ns; ++j)
    if (sources[j] < 0 || sources[j] >= n) return ORC_ERR_ARG;
  int32_t *ATrp, *ATcol;
  if (orc_transpose(n, n, rp, col, &ATrp, &ATcol)) return ORC_ERR_OOM;
  /* frontier F(v, j) = A(s_j, v): transpose of the selected rows */
  int32_t *Frp = (int32_t *)calloc(n + 1, sizeof(int32_t));
  int64_t fnz = 0;
  for (int64_t j = 0; j < ns; ++j) fnz += rp[sources[j] + 1] - rp[sources[j]];
  int32_t *Fcol = (int32_t *)malloc(sizeof(int32_t) * (fnz > 0 ? fnz : 1));
  double *Fval = (double *)malloc(sizeof(double) * (fnz > 0 ? fnz : 1));
  for (int64_t j = 0; j < ns; ++j)
    for (int32_t t = rp[sources[j]]; t < rp[sources[j] + 1]; ++t) Frp[col[t] + 1]++;
  for (int64_t i = 0; i < n; ++i) Frp[i + 1] += Frp[i];
  {
    int32_t *pos = dup_i32(Frp, n);
    for (int64_t j = 0; j < ns; ++j)
      for (int32_t t = rp[sources[j]]; t < rp[sources[j] + 1]; ++t) {
        Fcol[pos[col[t]]] = (int32_t)j;
        Fval[pos[col[t]]++] = 1.0;
      }
    free(pos);
  }
  /* visited set with sigma values: (s_j, j) = 1 (columns sorted per row) */
  int32_t *Vrp = (int32_t *)calloc(n + 1, sizeof(int32_t));
  int32_t *Vcol = (int32_t *)malloc(sizeof(int32_t) * (ns > 0 ? ns : 1));
  double *Vval = (double *)malloc(sizeof(double) * (ns > 0 ? ns : 1));
  for (int64_t j = 0; j < ns; ++j) Vrp[sources[j] + 1]++;
  for (int64_t i = 0; i < n; ++i) Vrp[i + 1] += Vrp[i];
  {
    int32_t *pos = dup_i32(Vrp, n);
    for (int64_t j = 0; j < ns; ++j) { Vcol[pos[sources[j]]] = (int32_t)j; Vval[pos[sources[j]]++] = 1.0; }
    free(pos);
  }
  /* per-level frontier patterns */
  int64_t cap = 16, d = 0;
  int32_t **Srp = (int32_t **)malloc(sizeof(int32_t *) * cap);
  int32_t **Scol = (int32_t **)malloc(sizeof(int32_t *) * cap);
  while (1) {
    if (d == cap) {
      cap *= 2;
      Srp = (int32_t **)realloc(Srp, sizeof(int32_t *) * cap);
      Scol = (int32_t **)realloc(Scol, sizeof(int32_t *) * cap);
    }
    Srp[d] = Frp;
    Scol[d] = Fcol;
    int32_t *Urp, *Ucol;
    double *Uval;
    err = union_disjoint(n, Vrp, Vcol, Vval, Frp, Fcol, Fval, &Urp, &Ucol, &Uval);
    if (err) break;
    free(Vrp); free(Vcol); free(Vval);
    Vrp = Urp; Vcol = Ucol; Vval = Uval;
    int32_t *Nrp, *Ncol;
    double *Nval;
    int64_t nnz_next;
    err = orc_mxm_cmask_f64(n, ns, ATrp, ATcol, NULL, Frp, Fcol, Fval, Vrp, Vcol, &Nrp, &Ncol, &Nval, &nnz_next);
    if (err) break;
    free(Fval);
    ++d;
    Frp = Nrp; Fcol = Ncol; Fval = Nval;
    if (nnz_next == 0) break;
  }
  free(Frp); free(Fcol); free(Fval);
  if (!err) {
    size_t mn = (size_t)n * (size_t)ns;
    double *sig = (double *)calloc(mn, sizeof(double));
    double *del = (double *)calloc(mn, sizeof(double));
    double *W = (double *)calloc(mn, sizeof(double));
    double *W2 = (double *)malloc(sizeof(double) * (mn > 0 ? mn : 1));
    if (!sig || !del || !W || !W2) err = ORC_ERR_OOM;
    if (!err) {
#pragma omp parallel for schedule(dynamic, 64)
      for (int64_t i = 0; i < n; ++i)
        for (int32_t t = Vrp[i]; t < Vrp[i + 1]; ++t) sig[(size_t)i * ns + Vcol[t]] = Vval[t];
      for (int64_t k = d - 1; k > 0; --k) {
        /* W<Sigmas[k]> = (1 + delta) / sigma, zero elsewhere */
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < mn; ++q) W[q] = 0.0;
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i)
          for (int32_t t = Srp[k][i]; t < Srp[k][i + 1]; ++t) {
            size_t q = (size_t)i * ns + Scol[k][t];
            W[q] = (1.0 + del[q]) / sig[q];
          }
        orc_mxm_spd_f64(n, ns, rp, col, NULL, W, Srp[k - 1], Scol[k - 1], W2);
        /* delta += W2 .* sigma on the level-(k-1) pattern */
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i)
          for (int32_t t = Srp[k - 1][i]; t < Srp[k - 1][i + 1]; ++t) {
            size_t q = (size_t)i * ns + Scol[k - 1][t];
            del[q] += W2[q] * sig[q];
          }
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < ns; ++j) s += del[(size_t)i * ns + j];
        bc[i] = s;
      }
    }
    free(sig); free(del); free(W); free(W2);
  }
  for (int64_t q = 0; q < d; ++q) { free(Srp[q]); free(Scol[q]); }
  free(Srp); free(Scol);
  free(Vrp); free(Vcol); free(Vval);
  free(ATrp); free(ATcol);
  if (depth_out) *depth_out = (int32_t)d;
  return err;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

/*
 * graphalgo.h — C ABI of the MI355X (gfx950) masked-SpGEMM library.
 *
 * The reference (UAlex322/GraphAlgo) is a header-only C++ template library;
 * its hot path is the masked sparse product C<M> = A*B over CSR and the three
 * algorithms built on it.  Every entry point below names the reference
 * interface it replaces.  The C++ drop-in surface with the reference's own
 * names and signatures (spMtx, mxmm_msa, k_truss, ...) is in
 * include/graphalgo/ and forwards to these functions.
 *
 * Conventions
 *  - CSR with int32 row pointers and int32 column indices (the reference's
 *    `int* Rst`, `int* Col`, include/matrix.h:31-32); rows sorted ascending
 *    with no duplicate columns.  `val == NULL` means a pattern matrix whose
 *    values are all 1.
 *  - Every function returns a ga_status; 0 is success.  On failure
 *    ga_last_error() returns a message for the calling thread.
 *  - Host-pointer entry points are synchronous: inputs are copied to the GPU,
 *    outputs copied back before return.  ga_graph_* entry points work on
 *    graphs kept resident in HBM.
 *  - The `isParallel` flag of the reference has no counterpart: there is one
 *    GPU path.
 */
#ifndef GRAPHALGO_H
#define GRAPHALGO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ga_status;
#define GA_OK 0
#define GA_ERR_ARG 1       /* bad argument (null pointer, negative size, bad enum) */
#define GA_ERR_DIM 2       /* dimension mismatch between operands */
#define GA_ERR_UNSORTED 3  /* CSR rows not sorted / duplicate or out-of-range column */
#define GA_ERR_OOM 4       /* device or host allocation failed / admission refused */
#define GA_ERR_HIP 5       /* HIP runtime error */
#define GA_ERR_LIMIT 6     /* size beyond a documented limit */
#define GA_ERR_STATE 7     /* library not initialised / handle misuse */

/* Value types of the operators.  GA_FLOAT32 is the reference's `float`
 * (spMtx<float> / denseMtx<float> in include/betweenness_centrality.h:137,
 * :141-143); floating-point sums are formed in the kernels' own order. */
typedef enum { GA_INT32 = 0, GA_INT64 = 1, GA_FLOAT64 = 2, GA_FLOAT32 = 3 } ga_dtype;

/* A CSR matrix view (host or device pointers depending on the call). */
typedef struct {
  int64_t nrows;
  int64_t ncols;
  int64_t nnz;
  const int32_t* rowptr; /* nrows + 1 */
  const int32_t* col;    /* nnz */
  const void* val;       /* nnz values of the call's dtype, or NULL = all ones */
} ga_csr;

/* ---- library state ----------------------------------------------------- */
const char* ga_last_error(void);
/* Selects the HIP device and creates the library's stream; idempotent. */
ga_status ga_init(int device);
/* Device name, compute-unit count and HBM size of the initialised device. */
ga_status ga_device_info(char* name, int name_len, int* compute_units, int64_t* hbm_bytes);
/* Waits for all work the library queued. */
ga_status ga_synchronize(void);
/* Checks a host CSR for the input contract (sorted rows, columns in range).
 * Returns GA_OK or GA_ERR_UNSORTED / GA_ERR_ARG. */
ga_status ga_check_csr(const ga_csr* A);

/* ---- masked SpGEMM ----------------------------------------------------- */
/* C<M> = A*B with a structural mask: C has exactly M's pattern (explicit zeros
 * kept) and Cval[j] = sum_k A(i,k) B(k, M.col[j]) for every stored j of row i.
 * Replaces mxmm_msa / mxmm_mca / mxmm_heap (include/matrix_la.h:700-713,
 * :586-599, :1013-1028, bound through mxmOp<T>, src/main.cpp:11-12, :206-218).
 * dtype selects the value type of A, B and C (plus-times semiring);
 * integer products accumulate exactly (int32 wraps like the reference's int).
 * Cval: M.nnz elements, host memory. */
ga_status ga_mxm_masked(const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, void* Cval);

/* C<!M> = A*B with a complement mask: C holds the product entries whose
 * column is not stored in M's row, columns ascending.  Replaces
 * mxmm_msa_cmask (include/matrix_la.h:788-891).  The library allocates the
 * three output arrays with malloc (free with ga_free_host):
 * *Crowptr (nrows+1), *Ccol (*Cnnz), *Cval (*Cnnz values of dtype). */
ga_status ga_mxm_cmask(const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, int32_t** Crowptr,
                       int32_t** Ccol, void** Cval, int64_t* Cnnz);

/* Element-wise operators (include/matrix_la.h:272-541, defects fixed).
 * op GA_EW_ADD: C = A + B over the union of the patterns (eWiseAdd, :375-436),
 *   or with a mask M: C<M> = A + B on M's pattern, entries summing to zero
 *   dropped (:272-335);
 * op GA_EW_MULT: C = A .* B over the intersection (eWiseMult, :438-484), with a
 *   mask restricted to M's pattern (:486-533);
 * op GA_EW_ADD_DISJOINT: the union of two patterns that share no position
 *   (add_nointersect, :337-373); GA_ERR_ARG if they do; no mask.
 * Same shapes, sorted rows, values of dtype (NULL values = ones); M may be
 * NULL.  Output arrays are allocated with malloc (free with ga_free_host):
 * *Crowptr (nrows+1), *Ccol (*Cnnz), *Cval (*Cnnz values of dtype). */
enum { GA_EW_ADD = 0, GA_EW_MULT = 1, GA_EW_ADD_DISJOINT = 2 };
ga_status ga_ewise(int op, const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, int32_t** Crowptr,
                   int32_t** Ccol, void** Cval, int64_t* Cnnz);
/* multScalar (include/matrix_la.h:535-541) without its in-place write through
 * a const operand: Cval[t] = A.val[t] * alpha in dtype (integers converted
 * back toward zero, as `v *= alpha` does). */
ga_status ga_scale(const ga_csr* A, ga_dtype dtype, double alpha, void* Cval);

/* Masked sparse x dense: C<M> = A*B with B and C dense row-major
 * (A.ncols x ncols and A.nrows x ncols); row i of C is zero outside M's
 * row i.  Replaces mxmm_spd (include/matrix_la.h:239-270).  FP64; the
 * _f32 form is the reference's float instantiation
 * (include/betweenness_centrality.h:210).  Each C entry sums A's row in CSR
 * order, as the reference does. */
ga_status ga_mxm_spd(const ga_csr* A, const double* B, int64_t ncols, const ga_csr* M, double* C);
ga_status ga_mxm_spd_f32(const ga_csr* A, const float* B, int64_t ncols, const ga_csr* M, float* C);

/* Dense element-wise operators of the reference's BC backward sweep, on
 * count = m * n row-major values of dtype (host memory, synchronous).
 * ga_dense_fuse_mult_add: C += A .* B (fuseEWiseMultAdd,
 *   include/matrix_la.h:192-198);
 * ga_dense_mult_masked: C<M> = A .* B -- C is zero outside M's pattern and
 *   A(i,j) * B(i,j) at M's entries (the dense eWiseMult,
 *   include/matrix_la.h:200-219); M is m x n.
 * Floating-point products and sums are rounded separately (no fused
 * multiply-add), as the reference's scalar loop computes them. */
ga_status ga_dense_fuse_mult_add(int64_t count, ga_dtype dtype, const void* A, const void* B, void* C);
ga_status ga_dense_mult_masked(const ga_csr* M, ga_dtype dtype, const void* A, const void* B, void* C);

void ga_free_host(void* p);

/* Page-locked host memory for result buffers a caller keeps across calls:
 * device-to-host copies into it run at DMA speed with no page faults (a truss
 * decomposition of R-MAT 16 reads back 7 MB of truss numbers: 0.5 ms into
 * fresh pageable memory).  No counterpart in the reference, whose results are
 * host arrays already; for callers that read results back often (the Python
 * mirror's resident graphs).  Free with ga_host_free. */
ga_status ga_host_alloc(int64_t bytes, void** ptr);
ga_status ga_host_free(void* ptr);

/* ---- triangle counting ------------------------------------------------- */
/* A: symmetric pattern, sorted rows.  T = sum of C<L> = L*L, L = tril(A,-1).
 * Replaces triangle_counting_masked_sandia (src/main.cpp:65-86). */
ga_status ga_triangle_count(const ga_csr* A, int64_t* triangles);
/* T = sum(C<A> = L*L^T) / 2.  Replaces triangle_counting_masked_lu
 * (src/main.cpp:40-62). */
ga_status ga_triangle_count_lu(const ga_csr* A, int64_t* triangles);
/* per_vertex[i] = (sum_j (A*A)<A>(i,j)) / 2.  Replaces
 * triangle_counting_vertex (src/main.cpp:14-37); int64 instead of int. */
ga_status ga_triangle_count_vertex(const ga_csr* A, int64_t* per_vertex);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ga_triangle_count's result computed without any
 * allocation.  L is the strict lower triangle of the symmetric pattern (the
 * matrix the reference multiplies, src/main.cpp:67-73) in DEVICE memory;
 * max_row bounds its longest row; *triangles is a DEVICE int64 (8-byte
 * aligned) that receives T = sum of C<L> = L*L.  Runs on `stream` (a
 * hipStream_t; NULL = the library stream) and returns without waiting.
 * Every work buffer comes from one fixed device array of
 * GA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a graph
 * whose buffers exceed it (about 2 million rows) is refused with
 * GA_ERR_LIMIT before anything runs.  Calls that overlap on one GPU share
 * that array and race; calls in stream order do not. */
#define GA_NOALLOC_SCRATCH_BYTES (64ll << 20)
ga_status ga_triangle_count_noalloc(const ga_csr* L, int64_t max_row, int64_t* triangles, void* stream);

/* ---- k-truss ----------------------------------------------------------- */
/* The k-truss of the symmetric pattern A: repeat { S<C> = C*C; keep S >= k-2 }
 * until nnz is unchanged.  Replaces k_truss (src/main.cpp:89-146).
 * out_rowptr (nrows+1) and out_col (capacity A.nnz) receive the pattern;
 * *iterations = number of support passes (the reference's
 * totalIterationNum). */
ga_status ga_ktruss(const ga_csr* A, int k, int32_t* out_rowptr, int32_t* out_col, int64_t* out_nnz,
                    int32_t* iterations);

typedef struct {
  int64_t passes;         /* support products computed */
  int64_t steps;          /* filter steps (one per k and per removal round) */
  int64_t wedges;         /* sum over passes of sum_{(i,k) in C} nnz(C_k) */
  int64_t bytes_alg;      /* algorithmic bytes of the support passes */
  double support_ms;      /* GPU time in the support kernels (when profiling) */
  double total_ms;        /* wall time of the whole decomposition */
} ga_truss_stats;

/* (ga_ktruss, ga_truss_decomposition and ga_graph_truss reject a pattern that
 * is not symmetric with GA_ERR_ARG: the support pass relies on S = S^T.) */
/* Truss number of every stored entry of the symmetric pattern A: the largest
 * k whose k-truss holds the entry, 2 if none (the caller-side loop over k of
 * the reference's k_truss).  truss: A.nnz int32 (host).  stats may be NULL. */
ga_status ga_truss_decomposition(const ga_csr* A, int32_t* truss, int32_t* kmax, ga_truss_stats* stats);

/* ---- betweenness centrality -------------------------------------------- */
/* bc[v] = sum over the given sources s of the Brandes dependency delta_s(v)
 * (unnormalised, not halved), FP64 path counts and dependencies.  Replaces
 * betweenness_centrality_batch (include/betweenness_centrality.h:125-237),
 * whose sources are 0..batch-1; here they are an explicit list of distinct
 * vertices.  A: square pattern (directed edges as stored). */
ga_status ga_bc_batch(const ga_csr* A, const int32_t* sources, int64_t nsources, double* bc);

/* ---- resident graphs --------------------------------------------------- */
typedef struct ga_graph ga_graph;
/* Uploads a symmetric pattern and keeps it in HBM with its workspaces. */
ga_status ga_graph_create(const ga_csr* A, ga_graph** out);
ga_status ga_graph_destroy(ga_graph* g);
ga_status ga_graph_triangle_count(ga_graph* g, int64_t* triangles);
ga_status ga_graph_truss(ga_graph* g, int32_t* truss, int32_t* kmax, ga_truss_stats* stats);
ga_status ga_graph_bc(ga_graph* g, const int32_t* sources, int64_t nsources, double* bc);
/* k-truss for one k of the resident graph (k_truss, src/main.cpp:89-146, on
 * a graph already in HBM): *nnz = stored entries of the k-truss (both
 * triangles, the reference's C.nz), *passes = support passes. */
ga_status ga_graph_ktruss(ga_graph* g, int k, int64_t* nnz, int32_t* passes);

/* ---- batched triangle counting ----------------------------------------- */
/* Many small graphs resident together and counted in one persistent launch.
 * The reference counts one graph per call (triangle_counting_masked_sandia,
 * src/main.cpp:65-86); a batch replaces ngraphs such calls. */
typedef struct ga_batch ga_batch;
/* graphs[g]: symmetric pattern, sorted rows.  Uploads the strict lower
 * triangles back to back (graph-local column indices). */
ga_status ga_batch_create(int64_t ngraphs, const ga_csr* graphs, ga_batch** out);
ga_status ga_batch_destroy(ga_batch* b);
/* counts[g] = triangles of graph g.  threads: workgroup width 64, 128 or 256
 * lanes (A entries per work item of a long row), 0 = the autotuned choice
 * (256 before tuning); the short-row kernel takes the autotuned rows per
 * claim (automatic before tuning). */
ga_status ga_batch_triangle_count(ga_batch* b, int threads, int64_t* counts);
/* Times the launch shapes on the resident batch and keeps the fastest for
 * later calls: first item width 128 or 256 x rows per claim of the short-row
 * kernel 8, 16, 32 or 64 (the rows in flight on an XCD share its L2), then,
 * at the best of those, the product count up to which a row takes the
 * short-row kernel (LDS bitmap, rows packed per wavefront) instead of the
 * item kernel (LDS mask accumulator): 2048, 16384, 65536 or 2^20 -- the
 * per-row accumulator choice that replaces the reference's global
 * msa / mca / heap switch (src/main.cpp:206-218).  ms_per_candidate (may be
 * NULL) receives GA_BATCH_TUNE_CANDIDATES times: (128, 8), (128, 16), ...,
 * (256, 64) width-major, then the four thresholds. */
#define GA_BATCH_TUNE_CANDIDATES 12
ga_status ga_batch_autotune(ga_batch* b, int* best_threads, int* best_claim, double* ms_per_candidate);
/* The launch shape the batch uses: item width (0 = not tuned, 256), rows per
 * claim (0 = automatic) and the short-row product threshold. */
ga_status ga_batch_shape(const ga_batch* b, int* threads, int* claim, int64_t* wave_flops);

/* ---- serving ----------------------------------------------------------- */
/* A request loop over graphs resident in HBM.  The reference has no serving
 * layer: each query is the reference's own call (triangle_counting_masked_
 * sandia src/main.cpp:65-86, k_truss :89-146, betweenness_centrality_batch
 * include/betweenness_centrality.h:125-237) on a graph that stays uploaded.
 *
 * The server owns one device arena of hbm_budget bytes, reserved once.
 * Resident graphs and every admitted query's workspace are blocks of that
 * arena; nothing is allocated while queries run.  A query's footprint is the
 * exact size of the workspace its path uses (ga_server_query_footprint).
 * Queries are admitted in arrival order when one of `slots` workers is free
 * and the arena holds a block of the footprint; each worker runs its query
 * on a HIP stream of its own, so queries of every kind and size overlap on
 * the device and a new query starts as soon as a slot and memory free up.
 * When a worker admits a query it also takes the queued queries of the same
 * kind on the same graph (up to max_batch): triangle counts share one count;
 * k-truss queries share one peeling run that stops at each requested k in
 * ascending order; BC queries of at most 64 sources share one multi-source
 * pass, one 64-lane group per query.  Results do not depend on what runs beside or
 * with a query (BC scores are bit-identical to the query run alone). */
typedef struct ga_server ga_server;
typedef struct {
  int64_t hbm_budget; /* bytes of the device arena (graphs + query workspaces) */
  int32_t slots;      /* runs in flight at most (one stream each), 1..64 */
  int32_t max_batch;  /* queued queries one run may take (0 = 16, 1 = no batching) */
} ga_server_config;
enum { GA_Q_TRIANGLES = 0, GA_Q_KTRUSS = 1, GA_Q_BC = 2 };
typedef struct {
  int32_t kind;            /* GA_Q_* */
  int32_t graph;           /* id from ga_server_add_graph */
  int32_t k;               /* GA_Q_KTRUSS: k >= 3 */
  int32_t nsources;        /* GA_Q_BC: 1..4096 distinct sources */
  const int32_t* sources;  /* GA_Q_BC: copied at submit */
  double* bc;              /* GA_Q_BC: n doubles, written before the query completes */
} ga_query;
typedef struct {
  int32_t status;      /* ga_status of the query */
  int32_t passes;      /* GA_Q_KTRUSS: support passes of its run up to its k */
  int64_t value;       /* triangles | k-truss entries (both triangles) | 0 for BC */
  int64_t footprint;   /* workspace bytes the query held */
  double t_submit_ms;  /* submit, admission, completion (ms; server clock, */
  double t_start_ms;   /* or the replay's clock for ga_*_replay) */
  double t_end_ms;
  int32_t slot;        /* worker that ran it (-1 if rejected) */
  int32_t batched;     /* queries that shared its run (1 = alone) */
} ga_query_result;
typedef struct {
  int64_t arena_bytes, resident_bytes; /* arena size; bytes held by graphs */
  int64_t peak_bytes;                  /* peak of graphs + leased workspaces */
  int64_t submitted, completed, rejected;
  int64_t batches;                     /* runs started (queries / batches = mean batch) */
  int32_t peak_inflight, slots;
  int64_t joined;                      /* k-truss queries that joined a run already going */
} ga_server_stats;

ga_status ga_server_create(const ga_server_config* cfg, ga_server** out);
/* Waits for every submitted query, then frees the arena. */
ga_status ga_server_destroy(ga_server* srv);
/* Symmetric pattern, sorted rows; kept in the arena with its strict lower
 * triangle.  GA_ERR_OOM when the arena cannot hold it. */
ga_status ga_server_add_graph(ga_server* srv, const ga_csr* A, int32_t* graph_id);
ga_status ga_server_query_footprint(ga_server* srv, const ga_query* q, int64_t* bytes);
/* Queues the query and returns at once.  GA_ERR_OOM when its footprint can
 * never fit beside the resident graphs; argument errors are reported here. */
ga_status ga_server_submit(ga_server* srv, const ga_query* q, int64_t* ticket);
/* Blocks until the ticket completes; returns the query's status. */
ga_status ga_server_wait(ga_server* srv, int64_t ticket, ga_query_result* out);
ga_status ga_server_stats_read(ga_server* srv, ga_server_stats* out);
/* Replays a query stream: query i is submitted at arrival_ms[i] after the
 * call starts (sorted ascending), then every result is collected; times in
 * out[] are ms since the call started. */
ga_status ga_server_replay(ga_server* srv, int64_t nq, const ga_query* q, const double* arrival_ms,
                           ga_query_result* out);
/* The same stream on the plain path: one query at a time in arrival order on
 * the library stream, through ga_graph_triangle_count / ga_graph_ktruss /
 * ga_graph_bc on graphs[q.graph] (the one-graph-per-call baseline). */
ga_status ga_plain_replay(int64_t ngraphs, ga_graph* const* graphs, int64_t nq, const ga_query* q,
                          const double* arrival_ms, ga_query_result* out);

/* ---- kernel timing ----------------------------------------------------- */
/* flags bit 0: bracket every launch of the masked-SpGEMM kernels with HIP
 * events on the stream they run on (launch count, GPU milliseconds);
 * bit 1: account the algorithmic wedges and bytes of every launch (runs an
 * extra counting kernel per launch, so keep it out of timed runs). */
typedef struct {
  int64_t launches;
  double kernel_ms;
  int64_t wedges;            /* algorithmic: sum_{(i,k) in A} nnz(B_k) per launch, summed */
  int64_t bytes_alg;         /* algorithmic bytes of the launches (full Gustavson product) */
  int64_t wedges_processed;  /* products the kernel actually formed (mask-pruned) */
  int64_t accounted_launches;/* launches covered by wedges / bytes_alg */
} ga_kernel_prof;
ga_status ga_prof_enable(int flags);
/* Diagnostic: a plain device copy of `bytes` (multiple of 16) repeated reps
 * times; *ms_per_copy = GPU time per copy.  Used as the calibration reference
 * of HBM byte counters and as the achievable-bandwidth probe.  No reference
 * counterpart. */
ga_status ga_stream_copy(int64_t bytes, int reps, double* ms_per_copy);
ga_status ga_prof_read(ga_kernel_prof* out, int reset);

/* Process-wide tuning options; the GA_* environment variables of the same
 * names set them at ga_init.  Every setting gives the same results (tests
 * compare them); only the speed differs.  Set them while no other call of
 * the library is running.  No reference counterpart. */
enum {
  GA_OPT_BC_BOTTOM_UP = 1,      /* 0/1: bottom-up BC forward steps (GA_BC_BOTTOM_UP) */
  GA_OPT_BC_ALPHA = 2,          /* > 0: go bottom-up once the level list holds more than 1/alpha
                                   of the unvisited vertices' edges (GA_BC_ALPHA) */
  GA_OPT_ROW_CLAIM = 3,         /* >= 0: rows per claim of the short-row kernels, 0 = automatic
                                   (GA_ROW_CHUNK) */
  GA_OPT_SUM_SUB = 4,           /* short rows of pattern sums (GA_SUM_SUB): 0 = rows packed per
                                   wavefront, 512 wedges per trip; 2 = the same, 256 per trip;
                                   1 = 16-lane wedge runs; 8/16/32/64 = lanes per row walking its
                                   entries */
  GA_OPT_ITEM_WEDGES = 5,       /* >= 1: target wedges per work item of repeated products
                                   (GA_ITEM_WEDGES) */
  GA_OPT_SELECTIVE_SUPPORT = 6, /* k-truss passes after the first recompute only the supports an
                                   endpoint change can affect (GA_SELECTIVE_SUPPORT): 1 = on,
                                   0 = off, -1 (default) = on for graphs of at least 2^22 stored
                                   entries when GA_OPT_SUPPORT_TRI is 0 */
  GA_OPT_TRUSS_LOOP = 7,        /* k-truss peeling loop (GA_TRUSS_LOOP): 0 = every filter step
                                   recomputes the supports, the host reads the survivor count
                                   after each; 1 = the same as device-driven steps (state in
                                   device memory, a few steps queued ahead); 2 = those steps
                                   replayed from a captured HIP graph; 3 = supports computed
                                   once and then updated: each round lists the triangles
                                   through the entries it removes and decrements the supports
                                   of their surviving edges, rounds queued on the device;
                                   4 = the same, the host reading each round's size (always
                                   the case for served runs that answer each k as it is
                                   reached); -1 (default) = 3.  Selective support passes run
                                   with 0. */
  GA_OPT_SUPPORT_TRI = 8,       /* k-truss support passes as triangle listing, C<L> = L*L over
                                   the strict lower half with each hit added to its three edges
                                   (GA_SUPPORT_TRI): 1 = on, 0 = off, -1 (default) = on for
                                   every pass whose graph holds at least 2^20 stored entries */
  GA_OPT_SERVE_JOIN = 9,        /* served k-truss queries joining a peeling run on their graph
                                   that is already going (GA_SERVE_JOIN): 0 = never, 1 = when
                                   the query's k lies within the run's k range, 2 = also above
                                   it (the run goes on to it) */
  GA_OPT_BC_LOOP = 10,          /* BC level loop (GA_BC_LOOP): 0 = host-steered, the host reading
                                   the size of every level list; 1 (default) = level lists sized
                                   on the device, the host waiting once per pass (end of the
                                   forward sweep); 2 = the same steps replayed from captured HIP
                                   graphs */
  GA_OPT_WAVE_FLOPS = 11        /* masked products on matrices of at most 4096 columns: rows with a
                                   mask of at most 128 entries and at most this many products go
                                   to the short-row kernels, the rest to the item kernel
                                   (GA_WAVE_FLOPS; default 65536; -1 = every row to the item
                                   kernel); a batch's autotuner sets its own */
};
ga_status ga_set_option(int option, double value);
ga_status ga_get_option(int option, double* value);

#ifdef __cplusplus
}
#endif
#endif /* GRAPHALGO_H */

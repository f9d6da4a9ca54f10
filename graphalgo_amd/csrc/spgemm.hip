// Masked SpGEMM C<M> = A*B with a structural mask (C has M's pattern).
//
// Row-wise Gustavson product as in the reference's MSA path
// (_mxmm_msa_parallel, include/matrix_la.h:715-749), re-laid out for CDNA4:
//
//  * Work items.  Row i is split into items of at most MXM_TCAP entries of
//    A's row (hub rows become many items, short rows one).  A persistent grid
//    sized to the chip's occupancy pulls items from an atomic counter, so the
//    skewed R-MAT rows balance dynamically instead of through static chunks.
//  * Flattened wedge space.  Inside an item the B rows selected by A's
//    entries are concatenated (LDS prefix sum of their lengths) and the 256
//    threads stride over that wedge space, so consecutive lanes read
//    consecutive B columns (coalesced) whether B rows are long or short.
//  * Mask membership.  M's row is staged in LDS as a sorted column list plus
//    an n-bit bitmap (for n <= 2^18 columns); a wedge whose column misses the
//    bitmap costs one LDS read.  Hits locate their mask position by binary
//    search in the LDS list and accumulate into a mask-indexed LDS
//    accumulator (the reference's MCA layout, include/matrix_la.h:543-569),
//    flushed once per item.  Mask rows longer than MXM_MCAP entries search
//    M in global memory and accumulate with global atomics.
//  * Integer accumulation is exact in any order, so results are bit-identical
//    to the CPU path; FP64 sums of rows split across items are summed in
//    arrival order.
#include "common.hpp"

#include <algorithm>
#include <type_traits>

namespace ga {

constexpr int64_t MXM_BITMAP_MAX_COLS = int64_t(1) << 18;  // 32 KB bitmap
#ifndef GA_MXM_G
#define GA_MXM_G 16
#endif
#ifndef GA_MXM_K
#define GA_MXM_K 8
#endif
constexpr int MXM_G = GA_MXM_G;  // lanes that read adjacent B columns (one 64-byte run)
constexpr int MXM_K = GA_MXM_K;  // wedges per lane per trip (gathers in flight)
constexpr int MXM_KV = 6;        // the same for valued products (no spill at 64 VGPRs, where 8 spills ten)
constexpr int MXM_RANK_MAX_WORDS = 4096;                     // rank lookup up to 2^17 columns
constexpr int MXM_RANK_MAX_WORDS_POS = 8192;                 // products that need positions: up to 2^18
// sums (triangle counts) need membership only: an exact bitmap up to 2^20
// columns (128 KB of LDS; no rank bases)
constexpr int64_t MXM_SUM_BITMAP_MAX_COLS = int64_t(1) << 20;
// wider matrices: a filter of this many bits (column c sets bit c mod it),
// small enough to keep eight wavefronts per SIMD; a search of the mask row
// confirms its hits
constexpr int64_t MXM_FILTER_COLS = int64_t(1) << 16;
// Items of the support pass and of pattern products: 512 A entries and an
// LDS mask accumulator of 4096 entries.  On R-MAT 16 the support kernel takes
// 344 us per pass at 256 / 1024, 233 us at 512 / 2048 and 212 us at
// 512 / 4096: mask rows beyond the LDS capacity accumulate with global
// atomics, and fewer, larger items pay the per-item staging less often.
constexpr int MXM_WIDE = 512;
constexpr int MXM_SPLIT_MAX = 8;  // a row gets at most this many times its entry-based item count
constexpr int MXM_NSEG = 32;                                 // claim segments (one counter line each)
#ifndef GA_MXM_CTR_STRIDE
#define GA_MXM_CTR_STRIDE 32
#endif
constexpr int MXM_CTR_STRIDE = GA_MXM_CTR_STRIDE;            // u32 words between segment counters
constexpr int MXM_CTR_WORDS = 2 * MXM_NSEG * MXM_CTR_STRIDE + 32;  // item segments, row segments, cursor

// THREADS = workgroup size = A entries per item (one entry per thread in the
// entry scan); MCAP = mask entries staged in LDS.  BATCHED: the operands are
// many graphs stored back to back with graph-local column indices; row i
// belongs to graph row_group[i] whose rows start at gbase[row_group[i]].
// (products with 4-byte or no accumulators: at most 64 VGPRs, eight
// wavefronts per SIMD to hide the dependent loads; valued ones fit with six
// wedges per trip: C<A> = A*A of R-MAT 16 with int32 values 2.43 ms at 81
// VGPRs and five wavefronts, 1.63 ms at 63 and eight; 8-byte accumulators
// are bound by LDS instead)
template <typename T, bool PAT_A, bool PAT_B, int MODE, int THREADS, bool BATCHED, bool RANK>
__global__ __launch_bounds__(THREADS)
__attribute__((amdgpu_waves_per_eu(MODE == 2 ? 6 : ((MODE == 1 || sizeof(T) == 4) ? 8 : 1)))) void k_mxm_items(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
    const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol, const T* __restrict__ Bval,
    const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol, T* __restrict__ Cval,
    unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ item_row, const int32_t* __restrict__ item_t0,
    const int32_t* __restrict__ n_items_ptr, unsigned int* __restrict__ work_counter,
    unsigned long long* __restrict__ wedge_counter, int bitmap_words, int bm_and, const int32_t* __restrict__ Mend,
    int trunc, int chunk, int direct_rows, const int32_t* __restrict__ mirror, const int32_t* __restrict__ opos,
    unsigned long long* __restrict__ row_work, const int32_t* __restrict__ skip) {
  // skip: a device flag that turns the launch into a no-op (device-driven
  // peeling steps that need no support pass)
  if (skip && *skip) return;
  constexpr int MXM_THREADS = THREADS;
  constexpr int MXM_TCAP = THREADS;
  constexpr int MXM_MCAP = 8 * THREADS;  // mask rows up to this long accumulate in LDS
  // dynamic LDS: the n-bit mask bitmap (RANK: plus one word that stays zero,
  // tested by dropped wedges), then (RANK) per bitmap word the position in
  // the mask row of the word's first mask column
  extern __shared__ uint32_t s_bitmap[];
  constexpr int BM_EXTRA = RANK ? 1 : 0;
  uint16_t* s_wpre = reinterpret_cast<uint16_t*>(s_bitmap + bitmap_words + BM_EXTRA);
  const int drop_col = RANK ? 32 * bitmap_words : -1;  // column of a dropped wedge
  __shared__ int32_t s_mcol[RANK ? 1 : MXM_MCAP];
  __shared__ T s_acc[MODE != 1 ? MXM_MCAP : 1];
  // MODE 2 (triangle support): hits per A entry of the item
  __shared__ int32_t s_ecnt[MODE == 2 ? MXM_TCAP : 1];
  // entry e of the item: wedges [s_eoff[e], s_eoff[e+1]) live at B.col[s_jb[e] + w]
  // (the host guarantees an item's wedge count fits int32)
  __shared__ int32_t s_jb[MXM_TCAP];
  __shared__ int32_t s_eoff[MXM_TCAP + 1];
  __shared__ T s_aval[PAT_A ? 1 : MXM_TCAP];
  __shared__ long long s_wsum[MXM_THREADS / 64];
  __shared__ int s_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const bool use_bm = RANK || bitmap_words > 0;
  // bm_and = -1: the bitmap holds a bit per column (exact membership);
  // otherwise column c owns bit c & bm_and of a smaller bitmap, a filter
  // whose hits a search of the mask row confirms
  const bool bm_exact = bm_and == -1;
  for (int q = tid; q < bitmap_words + BM_EXTRA; q += MXM_THREADS) s_bitmap[q] = 0u;
  // direct mode: items 0..n_list-1 are the listed extra chunks of long rows
  // (the heaviest work goes first), then item n_list + r is row r's first
  // chunk for r < direct_rows
  const int n_list = *n_items_ptr;
  const int n_items = direct_rows + n_list;
  // Segmented claiming: the item range is cut into MXM_NSEG segments, each
  // with its own counter on its own 128-byte line, so claims do not serialise
  // on one address.  A workgroup drains its home segment, then moves on to
  // the next ones (work stealing); `chunk` items are claimed per atomic.
  const int seg_len = (n_items + MXM_NSEG - 1) / MXM_NSEG;
  int seg = blockIdx.x % MXM_NSEG;
  bool drained = false;   // wavefront 0: no segment has items left
  int tries_seq = 0;      // sum mode: segments found drained in order
  int next = 0, end = 0;  // thread 0's claimed range
  // sum mode: per-thread partial sums run on across the items of one target
  // (one graph of a batch, or the single total) and are reduced and added
  // with one atomic when the target changes and at exit, not once per item
  long long lsum = 0;
  int cur_g = -1;
  auto flush_sum = [&]() {
    const long long r = wave_reduce_sum(lsum);
    __syncthreads();
    if (lane == 0) s_wsum[wv] = r;
    __syncthreads();
    if (tid == 0) {
      long long t = 0;
#pragma unroll
      for (int k = 0; k < MXM_THREADS / 64; ++k) t += s_wsum[k];
      if (t) atomicAdd(&sums[cur_g], (unsigned long long)t);
    }
    __syncthreads();  // s_wsum is reused by the next item's entry scan
    lsum = 0;
  };

  while (true) {
    __syncthreads();  // previous item fully retired (LDS reuse)
    if (MODE == 1) {
      // sum mode (many short items): thread 0 walks the segments in order
      if (tid == 0) {
        int item = n_items;
        if (next < end) {
          item = next++;
        } else {
          while (!drained) {
            const int lo = seg * seg_len;
            const int hi = min(lo + seg_len, n_items);
            if (lo < hi) {
              const int k = (int)atomicAdd(&work_counter[seg * MXM_CTR_STRIDE], (unsigned)chunk);
              if (lo + k < hi) {
                item = lo + k;
                next = item + 1;
                end = min(lo + k + chunk, hi);
                break;
              }
            }
            seg = (seg + 1) % MXM_NSEG;
            if (++tries_seq >= MXM_NSEG) drained = true;
          }
        }
        s_item = item;
      }
    } else if (wv == 0) {
      // value mode (the support pass): wavefront 0 claims, the current
      // segment first (one atomic); when it is drained, the 32 segment
      // counters are read at once (one lane each) and an atomic goes to the
      // next segment with items left, so the end of the launch costs two
      // round trips per claim, not one per segment (4% on the support pass;
      // the sum-mode triangle count measured slower with it, so it keeps
      // the ordered walk above)
      int item = n_items;
      int have = 0;
      if (lane == 0) {
        if (next < end) {
          item = next++;
          have = 1;
        } else if (!drained) {
          const int lo = seg * seg_len;
          const int hi = min(lo + seg_len, n_items);
          if (lo < hi) {
            const int k = (int)atomicAdd(&work_counter[seg * MXM_CTR_STRIDE], (unsigned)chunk);
            if (lo + k < hi) {
              item = lo + k;
              next = item + 1;
              end = min(lo + k + chunk, hi);
              have = 1;
            }
          }
        }
      }
      have = __shfl(have, 0, 64);
      while (!have && !drained) {
        const int sc = __shfl(seg, 0, 64);
        bool avail = false;
        if (lane < MXM_NSEG) {
          const int sj = (sc + 1 + lane) % MXM_NSEG;
          const int lo = sj * seg_len;
          const int hi = min(lo + seg_len, n_items);
          if (lo < hi) {
            const unsigned taken =
                __hip_atomic_load(&work_counter[sj * MXM_CTR_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            avail = lo + (long long)taken < hi;
          }
        }
        const unsigned long long m = __ballot(avail);
        if (m == 0ull) {
          drained = true;  // every segment handed out (wave-uniform)
          break;
        }
        const int sj = (sc + 1 + (__ffsll((long long)m) - 1)) % MXM_NSEG;
        if (lane == 0) {
          seg = sj;
          const int lo = sj * seg_len;
          const int hi = min(lo + seg_len, n_items);
          const int k = (int)atomicAdd(&work_counter[sj * MXM_CTR_STRIDE], (unsigned)chunk);
          if (lo + k < hi) {
            item = lo + k;
            next = item + 1;
            end = min(lo + k + chunk, hi);
            have = 1;
          }
        }
        have = __shfl(have, 0, 64);
      }
      if (lane == 0) s_item = item;
    }
    __syncthreads();
    if (s_item >= n_items) {
      if (MODE == 1 && cur_g >= 0) flush_sum();
      break;
    }
    // listed items are taken from the last row to the first: generators and
    // most graph files label hubs first, so the rows whose lower-half cut is
    // loosest (the heaviest items of the support pass) come last in row order
    // and would otherwise form the tail of the launch (4% on R-MAT 16)
    const int q = direct_rows ? s_item : n_items - 1 - s_item;

    const bool listed = q < n_list;
    const int i = listed ? item_row[q] : q - n_list;
    const int a_beg = Arp[i];
    const int a_end = MODE == 2 ? Mend[i] : Arp[i + 1];  // MODE 2: A is the lower half too
    const int t0 = listed ? item_t0[q] : a_beg;
    // a listed item ends where the next item of its row starts; a direct
    // item (or the last of a row) takes up to MXM_TCAP entries
    const int t1 = (listed && !direct_rows && q + 1 < n_list && item_row[q + 1] == i) ? item_t0[q + 1]
                                                                                      : min(t0 + MXM_TCAP, a_end);
    const bool multi = t0 > a_beg || t1 < a_end;  // the row is shared with other items
    const int ms = Mrp[i];
    const int mnz = (Mend ? Mend[i] : Mrp[i + 1]) - ms;  // Mend: only a prefix of M's row is the mask
    const bool local = mnz <= MXM_MCAP;
    const int bbase = BATCHED ? gbase[row_group[i]] : 0;
    if (a_end == a_beg || mnz <= 0) continue;  // direct items include empty rows
    if (MODE == 1) {
      const int g_item = row_group ? row_group[i] : 0;  // uniform across the workgroup
      if (cur_g >= 0 && g_item != cur_g) flush_sum();
      cur_g = g_item;
    }

    // stage the mask row: bitmap bits (and the per-word rank base, or the
    // sorted column list for the binary-search variant)
    for (int p = tid; p < mnz; p += MXM_THREADS) {
      const int c = Mcol[ms + p];
      if (RANK) {
        if (MODE != 1 && (p == 0 || (Mcol[ms + p - 1] >> 5) != (c >> 5))) s_wpre[c >> 5] = (uint16_t)p;
      } else if (local) {
        s_mcol[p] = c;
      }
      if (MODE != 1 && local) s_acc[p] = T(0);
      if (use_bm) atomicOr(&s_bitmap[(c & bm_and) >> 5], 1u << (c & 31));
    }
    // stage the item's entries: B row start and length, A value
    const int ne = t1 - t0;
    int len = 0, b0 = 0;
    if (tid < ne) {
      const int k = bbase + Acol[t0 + tid];
      b0 = Brp[k];
      int b1 = MODE == 2 ? Mend[k] : Brp[k + 1];  // MODE 2: B is the lower half (columns < k < i)
      if (MODE == 2) s_ecnt[tid] = 0;
      if (MODE != 2 && trunc && mnz > 0) {
        // products beyond the last mask column are discarded by the mask:
        // stop B's row at the upper bound of that column
        const int maxc = Mcol[ms + mnz - 1];
        int lo = b0, hi = b1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (Bcol[mid] <= maxc) lo = mid + 1;
          else hi = mid;
        }
        b1 = lo;
      }
      len = b1 - b0;
      if (!PAT_A) s_aval[tid] = Aval[t0 + tid];
    }
    const int incl = wave_incl_scan(len);
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < MXM_THREADS / 64; ++k) {
      if (k < wv) base += (int)s_wsum[k];
      total += (int)s_wsum[k];
    }
    if (tid < ne) {
      const int off = base + incl - len;
      s_eoff[tid] = off;
      s_jb[tid] = b0 - off;
    }
    if (tid == 0) {
      s_eoff[ne] = total;
      if (row_work) atomicAdd(&row_work[i], (unsigned long long)total);
    }
    __syncthreads();

    // Wedge mapping: lanes form groups of MXM_G, and group g owns the
    // contiguous wedge range [g0, g1) of span = ceil(total / NG) rounded up to
    // whole trips (small items leave the last groups idle rather than run
    // every group on a partly filled trip).  A trip of the group covers its
    // next MXM_G * KW wedges, lane r taking every MXM_G-th of them: the
    // MXM_G lanes read adjacent B columns (64-byte runs) and a lane's wedges
    // are MXM_G apart, so its entry only moves forward by a short linear
    // walk.  One binary search per lane and item finds the entry of its
    // first wedge; there is none per trip.
    const int grp = tid / MXM_G, rl = tid % MXM_G;
    constexpr int MXM_NG = MXM_THREADS / MXM_G;
    // valued products: fewer wedges per trip (registers for the values)
    constexpr int KW = (PAT_A && PAT_B) ? MXM_K : MXM_KV;
    constexpr int MXM_RUN = MXM_G * KW;
    const long long span = (((long long)total + MXM_NG - 1) / MXM_NG + MXM_RUN - 1) / MXM_RUN * MXM_RUN;
    const int g0 = (int)min<long long>(total, span * grp);
    const int g1 = (int)min<long long>(total, span * (grp + 1));
    // the loop is instantiated twice: mask row staged in LDS (sorted list and
    // LDS accumulator) or not (search and accumulate in global memory), so
    // the hot loop has no branch on it
    auto wedge_loop = [&](auto local_c) {
      constexpr bool LOCAL = decltype(local_c)::value;
      int e = 0;
      if (g0 + rl < g1) {  // entry holding the lane's first wedge: largest e with s_eoff[e] <= g0 + rl
        int lo = 0, hi = ne - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_eoff[mid] <= g0 + rl) lo = mid;
          else hi = mid - 1;
        }
        e = lo;
      }
      // a trip; FULL: all MXM_RUN wedges lie in the range (every trip but a
      // group's last one when its range is not whole trips), so no lane
      // checks its wedges against the range end
      auto trip = [&](auto full_c, const int p0) {
        constexpr bool FULL = decltype(full_c)::value;
        int jb = s_jb[e];
        int nb = s_eoff[e + 1];
        int cc[KW], jj[KW], ee[KW];
#pragma unroll
        for (int u = 0; u < KW; ++u) {
          // wedges past the range re-read its last one (in bounds) and are dropped
          const bool ok = FULL || p0 + u * MXM_G < g1;
          const int w = ok ? p0 + u * MXM_G : g1 - 1;
          while (w >= nb) {  // next non-empty entry (zero-length entries are skipped)
            ++e;
            nb = s_eoff[e + 1];
            jb = s_jb[e];
          }
          ee[u] = e;
          jj[u] = jb + w;
          const int c = Bcol[(unsigned)jj[u]];
          cc[u] = ok ? c : drop_col;
        }
        auto accumulate = [&](int p, T v) {
          if (MODE == 1) {
            lsum += (long long)v;
          } else if (LOCAL) {
            atomic_add_val(&s_acc[p], v);
          } else {
            const int o = opos ? opos[ms + p] : ms + p;
            atomic_add_val(&Cval[o], v);
            if (mirror) {
              const int mt = mirror[o];
              if (mt >= 0) atomic_add_val(&Cval[mt], v);
            }
          }
        };
        if (RANK) {
          // all bitmap words, then all rank bases, then the hits: the LDS
          // reads of the KW wedges are in flight together
          uint32_t wd[KW];
          int wi[KW], pre[KW];
#pragma unroll
          for (int u = 0; u < KW; ++u) {
            wi[u] = cc[u] >> 5;  // a dropped wedge reads the zero word
            wd[u] = s_bitmap[wi[u]];
          }
#pragma unroll
          for (int u = 0; u < KW; ++u) pre[u] = MODE == 1 ? 0 : s_wpre[wi[u]];  // a sum needs no positions
#pragma unroll
          for (int u = 0; u < KW; ++u) {
            const uint32_t bit = 1u << (cc[u] & 31);
            if (wd[u] & bit) {
              accumulate(pre[u] + __popc(wd[u] & (bit - 1u)),
                         (PAT_A ? T(1) : s_aval[ee[u]]) * (PAT_B ? T(1) : Bval[jj[u]]));
              if (MODE == 2) {  // the triangle's other two edges: the A entry and the B entry
                atomicAdd(&s_ecnt[ee[u]], 1);
                atomicAdd(reinterpret_cast<int*>(&Cval[jj[u]]), 1);
              }
            }
          }
        } else {
#pragma unroll
          for (int u = 0; u < KW; ++u) {
            const int c = cc[u];
            if (c < 0) continue;
            if (use_bm && !((s_bitmap[(c & bm_and) >> 5] >> (c & 31)) & 1u)) continue;
            if (MODE == 1 && use_bm && bm_exact) {  // a sum needs only membership, which the bit is
              accumulate(0, (PAT_A ? T(1) : s_aval[ee[u]]) * (PAT_B ? T(1) : Bval[jj[u]]));
              continue;
            }
            int lo = 0, hi = mnz;
            bool found;
            if (LOCAL) {
              while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_mcol[mid] < c) lo = mid + 1;
                else hi = mid;
              }
              found = lo < mnz && s_mcol[lo] == c;
            } else {
              const int32_t* mrow = Mcol + ms;
              while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (mrow[mid] < c) lo = mid + 1;
                else hi = mid;
              }
              found = lo < mnz && mrow[lo] == c;
            }
            if (found) {
              accumulate(lo, (PAT_A ? T(1) : s_aval[ee[u]]) * (PAT_B ? T(1) : Bval[jj[u]]));
              if (MODE == 2) {
                atomicAdd(&s_ecnt[ee[u]], 1);
                atomicAdd(reinterpret_cast<int*>(&Cval[jj[u]]), 1);
              }
            }
          }
        }
      };
      for (int pos = g0; pos < g1; pos += MXM_RUN) {
        const int p0 = pos + rl;
        if (p0 >= g1) break;  // later trips start even further out
        if (pos + MXM_RUN <= g1) trip(std::true_type{}, p0);
        else trip(std::false_type{}, p0);
      }
    };
    if (local) wedge_loop(std::true_type{});
    else wedge_loop(std::false_type{});
    __syncthreads();

    if (MODE == 2) {
      // every lower entry takes hits as a mask entry, an A entry and a B
      // entry of other items and rows: all added atomically (the transposed
      // entries are copied after the launch)
      if (local)
        for (int p = tid; p < mnz; p += MXM_THREADS) {
          const T v = s_acc[p];
          if (v != T(0)) atomic_add_val(&Cval[ms + p], v);
        }
      if (tid < ne && s_ecnt[tid] != 0) atomicAdd(reinterpret_cast<int*>(&Cval[t0 + tid]), s_ecnt[tid]);
    }
    if (MODE == 0) {
      if (local) {
        // mirror: C is symmetric and only its lower half is computed; each
        // value is also written to the transposed entry mirror[ms + p]
        // opos: the mask is a selection of C's entries; entry p of the mask
        // row is C entry opos[ms + p]
        for (int p = tid; p < mnz; p += MXM_THREADS) {
          const T v = s_acc[p];
          const int o = opos ? opos[ms + p] : ms + p;
          const int mt = mirror ? mirror[o] : -1;  // negative = no transposed entry
          if (multi) {
            if (v != T(0)) {
              atomic_add_val(&Cval[o], v);
              if (mt >= 0) atomic_add_val(&Cval[mt], v);
            }
          } else {
            Cval[o] = v;
            if (mt >= 0) Cval[mt] = v;
          }
        }
      }
    }
    if (use_bm) {
      for (int p = tid; p < mnz; p += MXM_THREADS) {
        const int c = (!RANK && local) ? s_mcol[p] : Mcol[ms + p];
        s_bitmap[(c & bm_and) >> 5] = 0u;
      }
    }
    if (tid == 0 && wedge_counter) atomicAdd(wedge_counter, (unsigned long long)total);
  }
}  // k_mxm_items

// ---- short rows: one wavefront per row -------------------------------------
// For matrices with at most WR_MAXCOLS columns (the small graphs of a batch),
// a row whose mask has at most WR_MCAP entries and whose product count is
// below the binning threshold is owned by a single wavefront: the mask row is
// staged as a per-wave LDS bitmap plus per-word rank bases, B rows are
// gathered over the flattened wedge space of up to 64 A entries at a time,
// and hits accumulate into a per-wave mask-indexed accumulator written out
// once.  No workgroup barrier is involved, so the four waves of a workgroup
// run four independent rows and overlap their dependent-load chains.
constexpr int WR_WAVES = 4;
constexpr int WR_MCAP = 128;
constexpr int WR_MAXCOLS = 4096;
constexpr int WR_BMW = WR_MAXCOLS / 32;

template <typename T, bool PAT_A, bool PAT_B, int MODE, bool BATCHED>
__global__ __launch_bounds__(64 * WR_WAVES) void k_mxm_rows(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const T* __restrict__ Aval,
    const int32_t* __restrict__ Brp, const int32_t* __restrict__ Bcol, const T* __restrict__ Bval,
    const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol, const int32_t* __restrict__ Mend,
    T* __restrict__ Cval, unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ rows, const int32_t* __restrict__ n_rows_ptr,
    unsigned int* __restrict__ counter, unsigned long long* __restrict__ wedge_counter, int trunc, int chunk,
    const int32_t* __restrict__ mirror, const int32_t* __restrict__ opos) {
  __shared__ uint32_t s_bm[WR_WAVES][WR_BMW];
  __shared__ uint16_t s_wpre[WR_WAVES][WR_BMW];
  __shared__ T s_acc[WR_WAVES][MODE == 0 ? WR_MCAP : 1];
  __shared__ int32_t s_jb[WR_WAVES][64];
  __shared__ int32_t s_eoff[WR_WAVES][65];
  __shared__ T s_aval[WR_WAVES][PAT_A ? 1 : 64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  uint32_t* bm = s_bm[wv];
  uint16_t* wpre = s_wpre[wv];
  int32_t* jbase = s_jb[wv];
  int32_t* eoff = s_eoff[wv];
  for (int q = lane; q < WR_BMW; q += 64) bm[q] = 0u;
  const int n_rows = *n_rows_ptr;
  unsigned long long wedges = 0;
  // Rows are claimed `chunk` at a time from MXM_NSEG segments of the row list
  // (one counter line each).  A wave starts in the segment of its workgroup
  // (blockIdx % MXM_NSEG: workgroups are dealt round-robin to the XCDs, so a
  // segment is drained by one XCD and the rows of neighbouring rows -- the
  // same graph of a batch -- share that XCD's L2), then steals from the next.
  const int seg_len = (n_rows + MXM_NSEG - 1) / MXM_NSEG;
  int seg = blockIdx.x % MXM_NSEG, tries = 0;
  int next = 0, end = 0;  // lane 0's claimed range
  long long wave_sum = 0;  // sum mode: this wave's running total for graph wave_g
  int wave_g = 0;
  while (true) {
    int r = n_rows;
    if (lane == 0) {
      if (next < end) {
        r = next++;
      } else {
        while (tries < MXM_NSEG) {
          const int lo = seg * seg_len;
          const int hi = min(lo + seg_len, n_rows);
          if (lo < hi) {
            const int k = (int)atomicAdd(&counter[seg * MXM_CTR_STRIDE], (unsigned)chunk);
            if (lo + k < hi) {
              r = lo + k;
              next = r + 1;
              end = min(lo + k + chunk, hi);
              break;
            }
          }
          seg = (seg + 1) % MXM_NSEG;
          ++tries;
        }
      }
    }
    r = __shfl(r, 0, 64);
    if (r >= n_rows) {
      if (MODE == 1 && lane == 0 && wave_sum) atomicAdd(&sums[wave_g], (unsigned long long)wave_sum);
      break;
    }
    const int i = rows[r];
    const int ms = Mrp[i];
    const int mnz = (Mend ? Mend[i] : Mrp[i + 1]) - ms;  // <= WR_MCAP by binning
    for (int p = lane; p < mnz; p += 64) {
      const int c = Mcol[ms + p];
      if (p == 0 || (Mcol[ms + p - 1] >> 5) != (c >> 5)) wpre[c >> 5] = (uint16_t)p;
      atomicOr(&bm[c >> 5], 1u << (c & 31));
      if (MODE == 0) s_acc[wv][p] = T(0);
    }
    const int maxc = (trunc && mnz > 0) ? Mcol[ms + mnz - 1] : 0x7FFFFFFF;
    const int bbase = BATCHED ? gbase[row_group[i]] : 0;
    __threadfence_block();
    long long lsum = 0;
    const int a_end = Arp[i + 1];
    for (int t0 = Arp[i]; t0 < a_end; t0 += 64) {
      const int ne = min(64, a_end - t0);
      int len = 0, b0 = 0;
      if (lane < ne) {
        const int k = bbase + Acol[t0 + lane];
        b0 = Brp[k];
        int b1 = Brp[k + 1];
        if (trunc) {
          int lo = b0, hi = b1;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (Bcol[mid] <= maxc) lo = mid + 1;
            else hi = mid;
          }
          b1 = lo;
        }
        len = b1 - b0;
        if (!PAT_A) s_aval[wv][lane] = Aval[t0 + lane];
      }
      const int incl = wave_incl_scan(len);
      if (lane < ne) {
        eoff[lane] = incl - len;
        jbase[lane] = b0 - (incl - len);
      }
      const int total = __shfl(incl, 63, 64);
      if (lane == 0) eoff[ne] = total;
      __threadfence_block();
      if (lane == 0) wedges += (unsigned long long)total;
      // same wedge mapping as the item kernel: groups of MXM_G lanes, MXM_K
      // wedges per lane per trip, one binary search per trip plus a walk
      const int grp = lane / MXM_G, rl = lane % MXM_G;
      int e = 0;
      for (int base = 0; base < total; base += 64 * MXM_K) {
        const int p0 = base + grp * (MXM_G * MXM_K) + rl;
        if (p0 >= total) break;
        {
          int lo = e, hi = ne - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (eoff[mid] <= p0) lo = mid;
            else hi = mid - 1;
          }
          e = lo;
        }
        int jb = jbase[e];
        int nb = eoff[e + 1];
        int cc[MXM_K], jj[MXM_K], ee[MXM_K];
#pragma unroll
        for (int u = 0; u < MXM_K; ++u) {
          const bool ok = p0 + u * MXM_G < total;
          const int w = ok ? p0 + u * MXM_G : total - 1;
          while (w >= nb) {
            ++e;
            nb = eoff[e + 1];
            jb = jbase[e];
          }
          ee[u] = e;
          jj[u] = jb + w;
          const int c = Bcol[(unsigned)jj[u]];
          cc[u] = ok ? c : -1;
        }
        uint32_t wd[MXM_K];
        int wi[MXM_K], pre[MXM_K];
#pragma unroll
        for (int u = 0; u < MXM_K; ++u) {
          wi[u] = cc[u] >= 0 ? (cc[u] >> 5) : 0;
          wd[u] = bm[wi[u]];
        }
#pragma unroll
        for (int u = 0; u < MXM_K; ++u) pre[u] = wpre[wi[u]];
#pragma unroll
        for (int u = 0; u < MXM_K; ++u) {
          const uint32_t bit = 1u << (cc[u] & 31);
          if (cc[u] < 0 || !(wd[u] & bit)) continue;
          const T v = (PAT_A ? T(1) : s_aval[wv][ee[u]]) * (PAT_B ? T(1) : Bval[jj[u]]);
          if (MODE == 1) lsum += (long long)v;
          else atomic_add_val(&s_acc[wv][pre[u] + __popc(wd[u] & (bit - 1u))], v);
        }
      }
      __threadfence_block();  // entry slots are rewritten by the next chunk
    }
    if (MODE == 0) {
      for (int p = lane; p < mnz; p += 64) {
        const T v = s_acc[wv][p];
        const int o = opos ? opos[ms + p] : ms + p;  // selected mask: C entry of mask entry p
        Cval[o] = v;
        if (mirror) {  // symmetric C: also the transposed entry (negative = none)
          const int mt = mirror[o];
          if (mt >= 0) Cval[mt] = v;
        }
      }
    } else {
      // the wave's total runs on while its rows belong to one graph and is
      // added once per run of rows (rows are claimed in chunks, in order)
      const long long tot = wave_reduce_sum(lsum);
      const int g_row = row_group ? row_group[i] : 0;
      if (g_row != wave_g) {
        if (lane == 0 && wave_sum) atomicAdd(&sums[wave_g], (unsigned long long)wave_sum);
        wave_sum = 0;
        wave_g = g_row;
      }
      wave_sum += tot;
    }
    for (int p = lane; p < mnz; p += 64) bm[Mcol[ms + p] >> 5] = 0u;
    __threadfence_block();
  }
  if (wedge_counter) {
    wedges = wave_reduce_sum(wedges);
    if (lane == 0 && wedges) atomicAdd(wedge_counter, wedges);
  }
}

// ---- short rows of pattern sum-mode products: 16 lanes per row ------------
// The rows of a batch of small graphs (triangle counts) have a few dozen
// entries and a few hundred products, so a row is a chain of dependent loads
// (row bounds, mask, entries, B bounds, gathers) with little work.  Here a
// 16-lane group owns a row and the four groups of a wavefront run four rows
// at once, overlapping their chains.  Same per-row structure as k_mxm_rows:
// an exact per-group bitmap (<= WR_MAXCOLS columns), entries in chunks of
// 32 (two per lane, so most rows of such graphs are one chunk) with a 16-lane
// scan, 16-lane wedge runs with a short walk.
constexpr int SR_SUB = 16;
constexpr int SR_CH = 2 * SR_SUB;                     // entries per chunk
constexpr int SR_GROUPS = WR_WAVES * (64 / SR_SUB);  // groups per workgroup
constexpr int SR_K = 8;                               // wedges per lane per trip

// (eight wavefronts per SIMD: the kernel is bound by the latency of its
// dependent load chains, so occupancy matters more than a few registers)
template <bool BATCHED>
__global__ __launch_bounds__(64 * WR_WAVES) __attribute__((amdgpu_waves_per_eu(8))) void k_sum_subrows(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const int32_t* __restrict__ Brp,
    const int32_t* __restrict__ Bcol, const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol,
    const int32_t* __restrict__ Mend, unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ rows, const int32_t* __restrict__ n_rows_ptr,
    unsigned int* __restrict__ counter, unsigned long long* __restrict__ wedge_counter, int trunc, int chunk) {
  __shared__ uint32_t s_bm[SR_GROUPS][WR_BMW + 1];  // + one word that stays zero (dropped wedges)
  __shared__ int32_t s_jb[SR_GROUPS][SR_CH];
  __shared__ int32_t s_eoff[SR_GROUPS][SR_CH + 1];
  const int sg = threadIdx.x / SR_SUB;  // group within the workgroup
  const int sl = threadIdx.x % SR_SUB;  // lane within the group
  uint32_t* bm = s_bm[sg];
  int32_t* jbase = s_jb[sg];
  int32_t* eoff = s_eoff[sg];
  for (int q = sl; q <= WR_BMW; q += SR_SUB) bm[q] = 0u;
  const int n_rows = *n_rows_ptr;
  // segmented claiming as in k_mxm_rows: the group's lane 0 takes `chunk`
  // rows at a time from its current segment (one atomic); when that segment
  // is drained the group reads the 32 segment counters at once (two per
  // lane) and moves to the next segment with rows left, so the end of the
  // launch costs two round trips per claim instead of one per segment
  const int seg_len = (n_rows + MXM_NSEG - 1) / MXM_NSEG;
  int seg = blockIdx.x % MXM_NSEG;
  bool drained = false;
  int next = 0, end = 0;  // the group's current claim (group-uniform)
  long long lsum = 0;  // this lane's hits for graph cur_g
  int cur_g = -1;
  unsigned long long wedges = 0;
  auto flush = [&]() {  // group-uniform
    long long t = lsum;
#pragma unroll
    for (int d = SR_SUB / 2; d >= 1; d >>= 1) t += __shfl_xor(t, d, SR_SUB);
    if (sl == 0 && t) atomicAdd(&sums[cur_g], (unsigned long long)t);
    lsum = 0;
  };
  auto try_claim = [&](int sj) {  // group-uniform; true when rows were taken
    const int lo = sj * seg_len;
    const int hi = min(lo + seg_len, n_rows);
    if (lo >= hi) return false;
    int k = 0;
    if (sl == 0) k = (int)atomicAdd(&counter[sj * MXM_CTR_STRIDE], (unsigned)chunk);
    k = __shfl(k, 0, SR_SUB);
    if (lo + k >= hi) return false;
    next = lo + k;
    end = min(lo + k + chunk, hi);
    return true;
  };
  while (true) {
    if (next >= end && !drained && !try_claim(seg)) {
      while (true) {
        uint32_t live = 0;
        for (int j = sl; j < MXM_NSEG; j += SR_SUB) {
          const int sj = (seg + 1 + j) % MXM_NSEG;
          const int lo = sj * seg_len;
          const int hi = min(lo + seg_len, n_rows);
          if (lo < hi) {
            const unsigned taken =
                __hip_atomic_load(&counter[sj * MXM_CTR_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lo + (long long)taken < hi) live |= 1u << j;
          }
        }
#pragma unroll
        for (int d = SR_SUB / 2; d >= 1; d >>= 1) live |= __shfl_xor(live, d, SR_SUB);
        if (live == 0u) {
          drained = true;
          break;
        }
        seg = (seg + 1 + (__ffs((int)live) - 1)) % MXM_NSEG;
        if (try_claim(seg)) break;
      }
    }
    if (next >= end) break;  // every segment handed out
    const int r = next++;
    const int i = rows[r];
    const int ms = Mrp[i];
    const int mnz = (Mend ? Mend[i] : Mrp[i + 1]) - ms;  // <= WR_MCAP by binning
    const int g_row = row_group ? row_group[i] : 0;
    if (g_row != cur_g) {
      if (cur_g >= 0) flush();
      cur_g = g_row;
    }
    for (int p = sl; p < mnz; p += SR_SUB) {
      const int c = Mcol[ms + p];
      atomicOr(&bm[c >> 5], 1u << (c & 31));
    }
    const int maxc = (trunc && mnz > 0) ? Mcol[ms + mnz - 1] : 0x7FFFFFFF;
    const int bbase = BATCHED ? gbase[row_group[i]] : 0;
    __threadfence_block();
    const int a_end = Arp[i + 1];
    unsigned int hits = 0;  // this row's hits on this lane (32-bit adds)
    for (int t0 = Arp[i]; t0 < a_end; t0 += SR_CH) {
      const int ne = min(SR_CH, a_end - t0);
      // lane sl stages entries sl and SR_SUB + sl of the chunk
      int len[2] = {0, 0}, b0[2] = {0, 0};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int q = h * SR_SUB + sl;
        if (q < ne) {
          const int k = bbase + Acol[t0 + q];
          b0[h] = Brp[k];
          int b1 = Brp[k + 1];
          if (trunc) {
            int lo = b0[h], hi = b1;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (Bcol[mid] <= maxc) lo = mid + 1;
              else hi = mid;
            }
            b1 = lo;
          }
          len[h] = b1 - b0[h];
        }
      }
      int incl[2] = {len[0], len[1]};
#pragma unroll
      for (int d = 1; d < SR_SUB; d <<= 1) {
        const int t0s = __shfl_up(incl[0], d, SR_SUB);
        const int t1s = __shfl_up(incl[1], d, SR_SUB);
        if (sl >= d) {
          incl[0] += t0s;
          incl[1] += t1s;
        }
      }
      incl[1] += __shfl(incl[0], SR_SUB - 1, SR_SUB);  // the second half follows the first
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int q = h * SR_SUB + sl;
        if (q < ne) {
          eoff[q] = incl[h] - len[h];
          jbase[q] = b0[h] - (incl[h] - len[h]);
        }
      }
      const int total = __shfl(incl[1], SR_SUB - 1, SR_SUB);
      if (sl == 0) {
        eoff[ne] = total;
        wedges += (unsigned long long)total;
      }
      __threadfence_block();
      // the group's trips cover consecutive runs of the chunk's wedges, so a
      // lane's entry only moves forward: a linear walk, no search
      int e = 0;
      for (int base = 0; base < total; base += SR_SUB * SR_K) {
        const int p0 = base + sl;
        if (p0 >= total) break;
        int jb = jbase[e];
        int nb = eoff[e + 1];
        int cc[SR_K];
#pragma unroll
        for (int u = 0; u < SR_K; ++u) {
          const bool ok = p0 + u * SR_SUB < total;
          const int w = ok ? p0 + u * SR_SUB : total - 1;
          while (w >= nb) {
            ++e;
            nb = eoff[e + 1];
            jb = jbase[e];
          }
          const int c = Bcol[(unsigned)(jb + w)];
          cc[u] = ok ? c : 32 * WR_BMW;  // a dropped wedge tests the zero word
        }
        uint32_t wd[SR_K];
#pragma unroll
        for (int u = 0; u < SR_K; ++u) wd[u] = bm[cc[u] >> 5];
#pragma unroll
        for (int u = 0; u < SR_K; ++u) hits += (wd[u] >> (cc[u] & 31)) & 1u;
      }
      __threadfence_block();  // entry slots are rewritten by the next chunk
    }
    lsum += hits;
    for (int p = sl; p < mnz; p += SR_SUB) bm[Mcol[ms + p] >> 5] = 0u;
    __threadfence_block();
  }
  if (cur_g >= 0) flush();
  if (wedge_counter && sl == 0 && wedges) atomicAdd(wedge_counter, wedges);
}

// ---- short rows of pattern sum-mode products, entry by entry -------------
// For rows whose B rows are short (a few dozen columns: kNN and other
// near-regular graphs), a SUB-lane group walks the A entries of its row one
// at a time and reads each entry's B row with SUB adjacent lanes (one
// coalesced run per SUB columns).  No wedge-to-entry search is needed: the
// search of k_sum_subrows is a chain of dependent LDS reads per trip, which
// is what bounds it on such rows.  Four entries are read at once so that
// their gathers are in flight together.  The mask row is an exact per-group
// bitmap (<= WR_MAXCOLS columns).
template <bool BATCHED, int SUB>
__global__ __launch_bounds__(64 * WR_WAVES) void k_sum_entries(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const int32_t* __restrict__ Brp,
    const int32_t* __restrict__ Bcol, const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol,
    const int32_t* __restrict__ Mend, unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ rows, const int32_t* __restrict__ n_rows_ptr,
    unsigned int* __restrict__ counter, unsigned long long* __restrict__ wedge_counter, int trunc, int chunk) {
  static_assert(SUB == 8 || SUB == 16 || SUB == 32 || SUB == 64, "group width");
  constexpr int NG = 64 * WR_WAVES / SUB;  // groups per workgroup
  constexpr int EU = 4;                    // entries whose B rows are read at once
  __shared__ uint32_t s_bm[NG][WR_BMW];
  const int sg = threadIdx.x / SUB;  // group within the workgroup
  const int sl = threadIdx.x % SUB;  // lane within the group
  uint32_t* bm = s_bm[sg];
  for (int q = sl; q < WR_BMW; q += SUB) bm[q] = 0u;
  const int n_rows = *n_rows_ptr;
  // segmented claims, as in k_sum_subrows (the group's lane 0 takes `chunk`
  // rows of its current segment; a drained segment sends the group to a
  // probe of all segment counters)
  const int seg_len = (n_rows + MXM_NSEG - 1) / MXM_NSEG;
  int seg = blockIdx.x % MXM_NSEG;
  bool drained = false;
  int next = 0, end = 0;
  unsigned long long lsum = 0;  // this lane's hits for graph cur_g
  int cur_g = -1;
  unsigned long long wedges = 0;
  auto flush = [&]() {  // group-uniform
    unsigned long long t = lsum;
#pragma unroll
    for (int d = SUB / 2; d >= 1; d >>= 1) t += __shfl_xor(t, d, SUB);
    if (sl == 0 && t) atomicAdd(&sums[cur_g], t);
    lsum = 0;
  };
  auto try_claim = [&](int sj) {
    const int lo = sj * seg_len;
    const int hi = min(lo + seg_len, n_rows);
    if (lo >= hi) return false;
    int k = 0;
    if (sl == 0) k = (int)atomicAdd(&counter[sj * MXM_CTR_STRIDE], (unsigned)chunk);
    k = __shfl(k, 0, SUB);
    if (lo + k >= hi) return false;
    next = lo + k;
    end = min(lo + k + chunk, hi);
    return true;
  };
  while (true) {
    if (next >= end && !drained && !try_claim(seg)) {
      while (true) {
        uint32_t live = 0;
        for (int j = sl; j < MXM_NSEG; j += SUB) {
          const int sj = (seg + 1 + j) % MXM_NSEG;
          const int lo = sj * seg_len;
          const int hi = min(lo + seg_len, n_rows);
          if (lo < hi) {
            const unsigned taken =
                __hip_atomic_load(&counter[sj * MXM_CTR_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lo + (long long)taken < hi) live |= 1u << j;
          }
        }
#pragma unroll
        for (int d = SUB / 2; d >= 1; d >>= 1) live |= __shfl_xor(live, d, SUB);
        if (live == 0u) {
          drained = true;
          break;
        }
        seg = (seg + 1 + (__ffs((int)live) - 1)) % MXM_NSEG;
        if (try_claim(seg)) break;
      }
    }
    if (next >= end) break;  // every segment handed out
    const int r = next++;
    const int i = rows[r];
    const int ms = Mrp[i];
    const int mnz = (Mend ? Mend[i] : Mrp[i + 1]) - ms;  // <= WR_MCAP by binning
    const int g_row = row_group ? row_group[i] : 0;
    if (g_row != cur_g) {
      if (cur_g >= 0) flush();
      cur_g = g_row;
    }
    for (int p = sl; p < mnz; p += SUB) {
      const int c = Mcol[ms + p];
      atomicOr(&bm[c >> 5], 1u << (c & 31));
    }
    const int maxc = (trunc && mnz > 0) ? Mcol[ms + mnz - 1] : 0x7FFFFFFF;
    const int bbase = BATCHED ? gbase[row_group[i]] : 0;
    __threadfence_block();
    const int a_end = Arp[i + 1];
    unsigned int hits = 0;  // this row's hits on this lane (32-bit adds)
    for (int t0 = Arp[i]; t0 < a_end; t0 += SUB) {
      const int ne = min(SUB, a_end - t0);
      int b0 = 0, b1 = 0;  // lane sl: B row of entry sl of this chunk
      if (sl < ne) {
        const int k = bbase + Acol[t0 + sl];
        b0 = Brp[k];
        b1 = Brp[k + 1];
        if (trunc) {
          int lo = b0, hi = b1;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (Bcol[mid] <= maxc) lo = mid + 1;
            else hi = mid;
          }
          b1 = lo;
        }
        wedges += (unsigned long long)(b1 - b0);
      }
      for (int e0 = 0; e0 < ne; e0 += EU) {
        int bs[EU], be[EU], cc[EU];
#pragma unroll
        for (int u = 0; u < EU; ++u) {
          const int e = e0 + u;
          const int x = __shfl(b0, e < ne ? e : 0, SUB);
          const int y = __shfl(b1, e < ne ? e : 0, SUB);
          bs[u] = e < ne ? x : 0;
          be[u] = e < ne ? y : 0;
        }
#pragma unroll
        for (int u = 0; u < EU; ++u) {
          const int p = bs[u] + sl;
          cc[u] = p < be[u] ? Bcol[p] : -1;
        }
#pragma unroll
        for (int u = 0; u < EU; ++u)
          if (cc[u] >= 0) hits += (bm[cc[u] >> 5] >> (cc[u] & 31)) & 1u;
        // B rows longer than SUB columns: the rest, SUB columns at a time
#pragma unroll
        for (int u = 0; u < EU; ++u) {
          for (int p = bs[u] + SUB + sl; p < be[u]; p += SUB) {
            const int c = Bcol[p];
            hits += (bm[c >> 5] >> (c & 31)) & 1u;
          }
        }
      }
    }
    lsum += hits;
    for (int p = sl; p < mnz; p += SUB) bm[Mcol[ms + p] >> 5] = 0u;
    __threadfence_block();
  }
  if (cur_g >= 0) flush();
  if (wedge_counter && wedges) atomicAdd(wedge_counter, wedges);
}

// ---- short rows of pattern sum-mode products, rows packed per wavefront ---
// A wavefront takes consecutive claimed rows of one graph -- up to SW_SLOTS
// rows holding at most SW_CH A entries together, or a single longer row --
// and runs their A entries as flattened chunks of SW_CH: lane L stages
// entries 2L and 2L+1 (their B row bounds, truncated at their own row's last
// mask column), a wave scan lays the entries' wedges end to end, and trips
// of 64 x K wedges cover the chunk, lane L taking wedge W0 + 64u + L at step
// u.  A wedge finds its entry by rank, not by a search or a walk: the entries
// of non-zero length start at distinct wedge positions, so with S the 64-bit
// mask of entry starts inside the step's 64-wedge window (start bits are set
// in LDS once per 32 x SW_WINW wedges), the wedge belongs to non-empty entry
// number (starts before the window) + (starts at or below the lane), which is
// two mbcnt instructions on a wave-uniform mask.  The entry's record (B
// offset, bitmap slot of its row) is one LDS read.  Each row's mask is an
// exact bitmap in its own slot.  Every lane runs the same trips, so rows of
// different lengths do not leave lanes behind a divergent group, and only the
// last trip of a chunk checks for wedges past its end (those lanes read the
// chunk's last wedge and test an all-zero bitmap).
constexpr int SW_SLOTS = 6;   // rows per batch (one bitmap each)
constexpr int SW_CH = 128;    // A entries per chunk (two per lane)
constexpr int SW_WINW = 64;   // words of entry-start bits per wave (2048 wedges, one word per lane)

template <bool BATCHED, int K, bool TRUNC>
__global__ __launch_bounds__(64 * WR_WAVES) __attribute__((amdgpu_waves_per_eu(8))) void k_sum_wave(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol, const int32_t* __restrict__ Brp,
    const int32_t* __restrict__ Bcol, const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mcol,
    const int32_t* __restrict__ Mend, unsigned long long* __restrict__ sums, const int32_t* __restrict__ row_group,
    const int32_t* __restrict__ gbase, const int32_t* __restrict__ rows, const int32_t* __restrict__ n_rows_ptr,
    unsigned int* __restrict__ counter, unsigned long long* __restrict__ wedge_counter, int chunk,
    int static_rows, long long* __restrict__ pub_host, long long pub_seq) {
  // trips of 64 K wedges tile a window of 32 SW_WINW exactly (a trip reads
  // 2 K start-bit words of the window as 16-byte groups)
  static_assert((32 * SW_WINW) % (64 * K) == 0 && (2 * K) % 4 == 0, "K must divide the start-bit window");
  // the slot bitmaps of every wave, then WR_BMW words that are never set
  // (tested by lanes past a chunk's last wedge); records address them by
  // word offset from the array's start
  constexpr int ZERO_OFF = WR_WAVES * SW_SLOTS * WR_BMW;
  __shared__ uint32_t s_bm[ZERO_OFF + WR_BMW];
  __shared__ int2 s_rec[WR_WAVES][SW_CH + 1];  // non-empty entries from 1: (B position - first wedge, bitmap offset)
  __shared__ __align__(16) uint32_t s_win[WR_WAVES][SW_WINW];  // entry starts among 32 x SW_WINW wedges
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int bm_off = wv * SW_SLOTS * WR_BMW;
  uint32_t* bm = s_bm + bm_off;
  int2* rec = s_rec[wv];
  uint32_t* win = s_win[wv];
  for (int q = threadIdx.x; q < ZERO_OFF + WR_BMW; q += blockDim.x) s_bm[q] = 0u;
  __syncthreads();
  const int n_rows = *n_rows_ptr;
  // segmented claims as in k_mxm_rows: `chunk` rows at a time from the
  // wave's current segment, the next live segment once it is drained
  const int seg_len = (n_rows + MXM_NSEG - 1) / MXM_NSEG;
  int seg = blockIdx.x % MXM_NSEG;
  bool drained = false;
  int c_lo = 0, c_hi = 0;
  long long lsum = 0;  // this lane's hits for graph cur_g
  int cur_g = -1;
  unsigned long long wedges = 0;
  auto flush = [&]() {
    const long long t = wave_reduce_sum(lsum);
    if (lane == 0 && t) atomicAdd(&sums[cur_g], (unsigned long long)t);
    lsum = 0;
  };
  auto try_claim = [&](int sj) {  // wave-uniform
    const int lo = sj * seg_len;
    const int hi = min(lo + seg_len, n_rows);
    if (lo >= hi) return false;
    int k = 0;
    if (lane == 0) k = (int)atomicAdd(&counter[sj * MXM_CTR_STRIDE], (unsigned)chunk);
    k = __shfl(k, 0, 64);
    if (lo + k >= hi) return false;
    c_lo = lo + k;
    c_hi = min(lo + k + chunk, hi);
    return true;
  };
  // static_rows: the launch has a wavefront for every `chunk` rows, so wave
  // number g takes rows g*chunk .. and no claim counter is touched
  if (static_rows) {
    c_lo = (int)min((int64_t)n_rows, ((int64_t)blockIdx.x * WR_WAVES + wv) * chunk);
    c_hi = (int)min((int64_t)n_rows, (int64_t)c_lo + chunk);
  }
  while (true) {
    bool got;
    if (static_rows) {
      got = !drained && c_lo < c_hi;
      drained = true;
    } else {
      got = !drained && try_claim(seg);
    }
    while (!got && !drained) {
      bool live = false;
      if (lane < MXM_NSEG) {
        const int sj = (seg + 1 + lane) % MXM_NSEG;
        const int lo = sj * seg_len;
        const int hi = min(lo + seg_len, n_rows);
        if (lo < hi) {
          const unsigned taken =
              __hip_atomic_load(&counter[sj * MXM_CTR_STRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          live = lo + (long long)taken < hi;
        }
      }
      const unsigned long long lm = __ballot(live);
      if (lm == 0ull) {
        drained = true;
        break;
      }
      seg = (seg + 1 + (__ffsll((long long)lm) - 1)) % MXM_NSEG;
      got = try_claim(seg);
    }
    if (!got) break;
    // the claim's rows, one per lane (chunk <= 64)
    const int nq = c_hi - c_lo;
    int qal = 0, qmnz = 0, qg = 0, qmaxc = 0x7FFFFFFF, qa_adj = 0, qm_adj = 0;
    if (lane < nq) {
      const int qi = rows[c_lo + lane];
      const int a0 = Arp[qi];
      qal = Arp[qi + 1] - a0;
      const int ms = Mrp[qi];
      qmnz = (Mend ? Mend[qi] : Mrp[qi + 1]) - ms;
      qg = row_group ? row_group[qi] : 0;
      if (TRUNC && qmnz > 0) qmaxc = Mcol[ms + qmnz - 1];
      qa_adj = a0;
      qm_adj = ms;
    }
    const int qa_incl = wave_incl_scan(qal);
    const int qm_incl = wave_incl_scan(qmnz);
    // flattened index f of a batch row's entries = position in the claim's
    // concatenation - the batch's first; qa_adj + f' is the entry's A index
    qa_adj -= qa_incl - qal;
    qm_adj -= qm_incl - qmnz;
    for (int q0 = 0; q0 < nq;) {
      const int g0 = __builtin_amdgcn_readlane(qg, q0);
      const int a_before = q0 ? __builtin_amdgcn_readlane(qa_incl, q0 - 1) : 0;
      const int m_before = q0 ? __builtin_amdgcn_readlane(qm_incl, q0 - 1) : 0;
      const bool ok = lane >= q0 && lane < nq && lane - q0 < SW_SLOTS && qg == g0 &&
                      (lane == q0 || qa_incl - a_before <= SW_CH);
      const unsigned long long bad = __ballot(!ok && lane > q0);
      const int q1 = bad ? (__ffsll((long long)bad) - 1) : 64;
      const int nb = q1 - q0;
      const int ea = __builtin_amdgcn_readlane(qa_incl, q1 - 1) - a_before;  // A entries of the batch
      const int em = __builtin_amdgcn_readlane(qm_incl, q1 - 1) - m_before;  // mask entries of the batch
      if (g0 != cur_g) {
        if (cur_g >= 0) flush();
        cur_g = g0;
      }
      // batch row holding flattened index f (of the entries or of the mask)
      auto slot_of = [&](int f, int incl, int before) {
        int sl = 0;
        for (int t = 1; t < nb; ++t)
          if (f >= __builtin_amdgcn_readlane(incl, q0 + t - 1) - before) sl = t;
        return sl;
      };
      if (ea > 0) {
        for (int f0 = 0; f0 < em; f0 += 64) {  // mask bitmaps
          const int f = f0 + lane;
          const int sl = slot_of(min(f, em - 1), qm_incl, m_before);
          const int mb = __shfl(qm_adj, q0 + sl, 64);
          if (f < em) {
            const int c = Mcol[mb + m_before + f];
            atomicOr(&bm[sl * WR_BMW + (c >> 5)], 1u << (c & 31));
          }
        }
        const int bbase = BATCHED ? gbase[g0] : 0;
        uint32_t hits = 0;
        for (int F = 0; F < ea; F += SW_CH) {
          // lane L stages entries F + 2L and F + 2L + 1
          int len[2] = {0, 0}, b0[2] = {0, 0}, sl[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int f = F + 2 * lane + h;
            sl[h] = slot_of(min(f, ea - 1), qa_incl, a_before);
            const int ab = __shfl(qa_adj, q0 + sl[h], 64);
            const int maxc = TRUNC ? __shfl(qmaxc, q0 + sl[h], 64) : 0;
            if (f < ea) {
              const int k = bbase + Acol[ab + a_before + f];
              b0[h] = Brp[k];
              int b1 = Brp[k + 1];
              if (TRUNC) {
                int lo = b0[h], hi = b1;
                while (lo < hi) {
                  const int mid = (lo + hi) >> 1;
                  if (Bcol[mid] <= maxc) lo = mid + 1;
                  else hi = mid;
                }
                b1 = lo;
              }
              len[h] = b1 - b0[h];
            }
          }
          const int incl = wave_incl_scan(len[0] + len[1]);
          const int excl[2] = {incl - len[0] - len[1], incl - len[1]};
          const int T = __builtin_amdgcn_readlane(incl, 63);
          wedges += (unsigned long long)T;
          if (T == 0) continue;
          // records of the non-empty entries, numbered from 1 in entry order
          const unsigned long long ne0 = __ballot(len[0] > 0), ne1 = __ballot(len[1] > 0);
          const int rk = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ne0 >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)ne0, 0u)) +
                         (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ne1 >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)ne1, 0u));
          if (len[0] > 0) rec[rk + 1] = make_int2(b0[0] - excl[0], bm_off + sl[0] * WR_BMW);
          if (len[1] > 0) rec[rk + 1 + (len[0] > 0)] = make_int2(b0[1] - excl[1], bm_off + sl[1] * WR_BMW);
          for (int Wb = 0; Wb < T; Wb += 32 * SW_WINW) {
            // entry starts among wedges Wb .. Wb + 32 * SW_WINW - 1
            win[lane] = 0u;
            __threadfence_block();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int rel = excl[h] - Wb;
              if (len[h] > 0 && rel >= 0 && rel < 32 * SW_WINW) atomicOr(&win[rel >> 5], 1u << (rel & 31));
            }
            __threadfence_block();
            // non-empty entries that start before the window
            int rb = __popcll(__ballot(len[0] > 0 && excl[0] < Wb)) + __popcll(__ballot(len[1] > 0 && excl[1] < Wb));
            const int wend = min(T, Wb + 32 * SW_WINW);
            for (int W0 = Wb; W0 < wend; W0 += 64 * K) {
              auto trip = [&](auto full_c) {
                constexpr bool FULL = decltype(full_c)::value;
                uint32_t wl[2 * K];
#pragma unroll
                for (int j = 0; j < 2 * K; j += 4) {
                  const uint4 v = *reinterpret_cast<const uint4*>(win + ((W0 - Wb) >> 5) + j);
                  wl[j] = v.x;
                  wl[j + 1] = v.y;
                  wl[j + 2] = v.z;
                  wl[j + 3] = v.w;
                }
                const int2* rq[K];
                int ri[K];
#pragma unroll
                for (int u = 0; u < K; ++u) {
                  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)wl[2 * u]);
                  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)wl[2 * u + 1]);
                  const unsigned long long S1 = (((unsigned long long)hi << 32) | lo) >> 1;
                  // entries starting at positions 0..lane of the step's
                  // window: bit 0 of S (a uniform part of the base) plus the
                  // bits of S >> 1 below the lane
                  rq[u] = rec + (rb + (int)(lo & 1u));
                  ri[u] = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(S1 >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)S1, 0u));
                  rb += __popc(lo) + __popc(hi);
                }
                int2 r[K];
#pragma unroll
                for (int u = 0; u < K; ++u) r[u] = rq[u][ri[u]];
                int cc[K];
#pragma unroll
                for (int u = 0; u < K; ++u) {
                  const int w = W0 + 64 * u + lane;
                  // 32-bit byte offset from the array base (B has < 2^29
                  // entries): one address register per gather
                  const uint32_t boff = 4u * (uint32_t)(r[u].x + (FULL ? w : min(w, T - 1)));
                  cc[u] = *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(Bcol) + boff);
                }
#pragma unroll
                for (int u = 0; u < K; ++u) {
                  int off = r[u].y;
                  if (!FULL && W0 + 64 * u + lane >= T) off = ZERO_OFF;
                  hits += __builtin_amdgcn_ubfe(s_bm[off + (cc[u] >> 5)], (uint32_t)cc[u], 1u);
                }
              };
              if (W0 + 64 * K <= T) trip(std::true_type{});
              else trip(std::false_type{});
            }
            __threadfence_block();  // the window is rewritten next
          }
        }
        lsum += hits;
        for (int f0 = 0; f0 < em; f0 += 64) {  // clear the bitmaps
          const int f = f0 + lane;
          const int sl = slot_of(min(f, em - 1), qm_incl, m_before);
          const int mb = __shfl(qm_adj, q0 + sl, 64);
          if (f < em) bm[sl * WR_BMW + (Mcol[mb + m_before + f] >> 5)] = 0u;
        }
        __threadfence_block();
      }
      q0 = q1;
    }
  }
  if (BATCHED || row_group) {  // several targets (graphs of a batch, or per-row groups)
    if (cur_g >= 0) flush();
  } else {
    // one target: the workgroup's waves add up in LDS and one atomic goes to
    // sums[0] (atomics on one address serialise, ~11 ns each: a small graph's
    // count had one per wavefront)
    __shared__ unsigned long long s_red[WR_WAVES];
    const long long t = wave_reduce_sum(lsum);
    if (lane == 0) s_red[wv] = (unsigned long long)t;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long tot = 0;
      for (int w = 0; w < WR_WAVES; ++w) tot += s_red[w];
      if (!pub_host) {
        if (tot) atomicAdd(&sums[0], tot);
      } else {
        // the launch publishes its own total.  Device-scope atomics complete
        // at the one point every workgroup sees; the arrival ticket's add
        // takes the sum add's returned value as an operand, so it is issued
        // only once the sum add has completed, and the last arrival reads
        // the total after every add (no release fence: those cost more than
        // the publish launch they replace)
        const unsigned long long old = atomicAdd(&sums[0], tot);
        const unsigned long long inc = 1ull + (old == ~0ull ? 1ull : 0ull);  // depends on the add's result
        if (atomicAdd(&sums[SUM_PUB_TICKET], inc) + 1 == (unsigned long long)gridDim.x) {
          const unsigned long long total =
              __hip_atomic_exchange(&sums[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&sums[SUM_PUB_TICKET], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&pub_host[0], (long long)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          __hip_atomic_store(&pub_host[1], pub_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
    }
  }
  if (wedge_counter && lane == 0 && wedges) atomicAdd(wedge_counter, wedges);
}

// Row binning: rows with an empty A row or empty mask get no work; rows
// whose mask fits WR_MCAP and whose (untruncated) product count is at most
// wave_flops go to the wavefront kernel; the rest are split into items of
// tcap A entries for the workgroup kernel.
__global__ void k_bin_rows(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol,
                           const int32_t* __restrict__ Brp, const int32_t* __restrict__ Mrp,
                           const int32_t* __restrict__ Mend, const int32_t* __restrict__ row_group,
                           const int32_t* __restrict__ gbase, int64_t m, int tcap, long long wave_flops,
                           int32_t* __restrict__ short_flag, int32_t* __restrict__ item_cnt) {
  // wave_flops < 0: no wavefront bin (matrix too wide for the per-wave bitmap)
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * wpb) {
    const int ab = Arp[i], ae = Arp[i + 1];
    const int mz = (Mend ? Mend[i] : Mrp[i + 1]) - Mrp[i];
    const int bbase = gbase ? gbase[row_group[i]] : 0;
    long long f = 0;
    if (mz > 0 && mz <= WR_MCAP)
      for (int t = ab + lane; t < ae; t += 64) f += Brp[bbase + Acol[t] + 1] - Brp[bbase + Acol[t]];
    f = wave_reduce_sum(f);
    if (lane == 0) {
      const int a = ae - ab;
      const bool none = a == 0 || mz == 0;
      const bool shrt = !none && wave_flops >= 0 && mz <= WR_MCAP && f <= wave_flops;
      short_flag[i] = shrt ? 1 : 0;
      item_cnt[i] = (none || shrt) ? 0 : (a + tcap - 1) / tcap;
    }
  }
}

__global__ void k_emit_rows(const int32_t* __restrict__ flag, const int32_t* __restrict__ off, int64_t m,
                            int32_t* __restrict__ rows) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    if (flag[i]) rows[off[i]] = (int32_t)i;
}

// Items per row: ceil(entries / tcap), or -- when the previous launch's
// wedges per row are known -- enough items of about `target` wedges each
// (entries split evenly), capped at MXM_SPLIT_MAX times the entry-based count
// and at one entry per item.
// Item planning of the listed path in two launches.  k_item_plan_count (one
// thread per row, PLAN_T rows per workgroup) counts each row's items -- one
// per tcap entries, raised to the previous launch's wedges / target when
// work_prev is given (at most MXM_SPLIT_MAX times the entry-based count) --
// and leaves per-workgroup totals.  It also clears what the launch that
// follows accumulates into (the work counters, this launch's per-row wedge
// counts) and, with split_col, computes the mask prefix Mend[i] = first
// entry of row i with column >= i (the strict lower half), so none of these
// needs a launch of its own.  k_item_plan_fill scans the totals (each
// workgroup adds the totals before it, or reads offsets prepared by
// scan_tile_offsets beyond PLAN_FUSED_TILES workgroups), places each row's
// items as k_item_fill does and writes the item count to start[m].
constexpr int PLAN_T = 256;
constexpr int PLAN_FUSED_TILES = 1024;

__global__ __launch_bounds__(PLAN_T) void k_item_plan_count(
    const int32_t* __restrict__ Arp, const int32_t* __restrict__ Mrp, const int32_t* __restrict__ Mend, int64_t m,
    int tcap, int32_t* __restrict__ cnt, const unsigned long long* work_prev, long long target,
    uint32_t* __restrict__ ctr_clear, int ctr_words, unsigned long long* work_clear,
    const int32_t* __restrict__ split_col, int32_t* __restrict__ split_out, long long* __restrict__ bsum,
    const int32_t* __restrict__ skip, int a_lower) {
  __shared__ long long red[PLAN_T / 64];
  if (skip && *skip) return;
  if (ctr_clear && blockIdx.x == 0)
    for (int q = threadIdx.x; q < ctr_words; q += PLAN_T) ctr_clear[q] = 0u;
  const int64_t i = (int64_t)blockIdx.x * PLAN_T + threadIdx.x;
  int c = 0;
  if (i < m) {
    // the previous launch's wedges of row i (work_prev may be work_clear itself)
    const unsigned long long wprev = work_prev ? work_prev[i] : 0ull;
    if (work_clear) work_clear[i] = 0ull;
    int a = Arp[i + 1] - Arp[i];
    int mend;
    if (split_col) {
      int lo = Mrp[i], hi = Mrp[i + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (split_col[mid] < (int)i) lo = mid + 1;
        else hi = mid;
      }
      split_out[i] = lo;
      mend = lo;
      if (a_lower) a = lo - Arp[i];  // A is the lower half as well (Arp == Mrp)
    } else {
      mend = Mend ? Mend[i] : Mrp[i + 1];
    }
    const int mz = mend - Mrp[i];
    c = (a > 0 && mz > 0) ? (a + tcap - 1) / tcap : 0;
    if (c > 0 && work_prev) {
      const long long byw = (long long)((wprev + (unsigned long long)target - 1) / (unsigned long long)target);
      c = (int)min<long long>(max<long long>(c, byw), min<long long>((long long)c * MXM_SPLIT_MAX, a));
    }
    cnt[i] = c;
  }
  long long t = wave_reduce_sum((long long)c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long b = 0;
    for (int w = 0; w < PLAN_T / 64; ++w) b += red[w];
    bsum[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(PLAN_T) void k_item_plan_fill(const int32_t* __restrict__ Arp,
                                                           const int32_t* __restrict__ cnt, int64_t m,
                                                           const long long* __restrict__ bsum, int offsets_ready,
                                                           int32_t* __restrict__ start, int32_t* __restrict__ item_row,
                                                           int32_t* __restrict__ item_t0,
                                                           const int32_t* __restrict__ skip,
                                                           const int32_t* __restrict__ Aend) {
  __shared__ long long wtot[PLAN_T / 64];
  if (skip && *skip) return;
  __shared__ long long prior[PLAN_T / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  long long before = 0;
  if (offsets_ready) {
    before = tid == 0 ? bsum[blockIdx.x] : 0;
  } else {
    for (int b = tid; b < (int)blockIdx.x; b += PLAN_T) before += bsum[b];
  }
  before = wave_reduce_sum(before);
  if (lane == 0) prior[wv] = before;
  const int64_t i = (int64_t)blockIdx.x * PLAN_T + tid;
  const int c = i < m ? cnt[i] : 0;
  const long long incl = wave_incl_scan((long long)c);
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  long long s0 = incl - c;
  for (int k = 0; k < PLAN_T / 64; ++k) {
    s0 += prior[k];
    if (k < wv) s0 += wtot[k];
  }
  if (i < m) {
    start[i] = (int32_t)s0;
    const int a0 = Arp[i];
    const int64_t a = (Aend ? Aend[i] : Arp[i + 1]) - a0;
    for (int q = 0; q < c; ++q) {
      item_row[s0 + q] = (int32_t)i;
      item_t0[s0 + q] = a0 + (int32_t)((int64_t)q * a / c);
    }
  }
  if (blockIdx.x == gridDim.x - 1 && tid == PLAN_T - 1) start[m] = (int32_t)(s0 + c);
}

// Extra items of rows longer than one item (chunks 1..), appended in any
// order through an atomic cursor; chunk 0 of every row is a direct item.
__global__ void k_extra_items(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Mrp,
                              const int32_t* __restrict__ Mend, int64_t m, int tcap,
                              int32_t* __restrict__ item_row, int32_t* __restrict__ item_t0,
                              int32_t* __restrict__ n_extra) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = Arp[i + 1] - Arp[i];
    const int mz = (Mend ? Mend[i] : Mrp[i + 1]) - Mrp[i];
    if (a <= tcap || mz <= 0) continue;
    const int c = (a + tcap - 1) / tcap - 1;
    const int base = atomicAdd(n_extra, c);
    for (int q = 0; q < c; ++q) {
      item_row[base + q] = (int32_t)i;
      item_t0[base + q] = Arp[i] + (q + 1) * tcap;
    }
  }
}

// Item q of a row with c items starts at entry a0 + floor(q * a / c): at most
// ceil(a / c) <= tcap entries each; an item ends where the row's next one
// starts (or at the row end).
__global__ void k_item_fill(const int32_t* __restrict__ Arp, const int32_t* __restrict__ cnt,
                            const int32_t* __restrict__ start, int64_t m, int32_t* __restrict__ item_row,
                            int32_t* __restrict__ item_t0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = cnt[i], s0 = start[i], a0 = Arp[i];
    const int64_t a = Arp[i + 1] - a0;
    for (int q = 0; q < c; ++q) {
      item_row[s0 + q] = (int32_t)i;
      item_t0[s0 + q] = a0 + (int32_t)((int64_t)q * a / c);
    }
  }
}

// first entry of row i with column >= i (the end of the strict lower half)
__global__ void k_row_split(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                            int32_t* __restrict__ split) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = rp[i], hi = rp[i + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    split[i] = lo;
  }
}

static hipEvent_t take_event(Context& c) {
  hipEvent_t ev;
  if (!c.event_pool.empty()) {
    ev = c.event_pool.back();
    c.event_pool.pop_back();
  } else {
    GA_HIP(hipEventCreate(&ev));
  }
  return ev;
}

// Resident workgroups per CU of a kernel at a launch shape.  The answer
// depends only on the kernel's resources and the device, so it is asked of
// the runtime once per (kernel, block, dynamic LDS, device) and kept per host
// thread; small products (a triangle count of a 1000-vertex graph is ~15 us
// of GPU time) are launched often enough for the query to show.
static int occupancy_of(const void* kern, int block, size_t dyn, int dev) {
  struct Entry {
    const void* kern;
    int block;
    size_t dyn;
    int device, occ;
  };
  thread_local std::vector<Entry> cache;
  for (const Entry& e : cache)
    if (e.kern == kern && e.block == block && e.dyn == dyn && e.device == dev) return e.occ;
  int occ = 0;
  GA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, dyn));
  cache.push_back(Entry{kern, block, dyn, dev, occ});
  return occ;
}
template <typename K>
static int occupancy_of(K* kern, int block, size_t dyn, int dev) {
  return occupancy_of(reinterpret_cast<const void*>(kern), block, dyn, dev);
}

// Bins the rows, then runs the wavefront kernel on the short rows and the
// workgroup kernel on the items of the long rows.  With profiling on, the
// pair is bracketed by one event pair (one "launch" of the masked product).
template <typename T, bool PA, bool PB, int MODE, int THREADS, bool BATCHED>
static void launch_binned(const DCsr& A, const DCsr& B, const DCsr& M, void* Cval, unsigned long long* sums,
                          const int32_t* row_group, const int32_t* gbase, MxmWork& w, int bitmap_words,
                          hipStream_t s, const int32_t* Mend, int trunc, const int32_t* mirror = nullptr,
                          const int32_t* opos = nullptr, int32_t* lower_out = nullptr) {
  // lower_out: the mask is the strict lower half of M; its row ends are
  // computed here into lower_out (by the item count when it runs) and used as Mend
  constexpr int TCAP = THREADS;
  const int64_t m = A.nrows;
  const int64_t max_items = (w.balance ? MXM_SPLIT_MAX : 1) * (m + ceil_div(A.nnz, TCAP)) + 1;
  // an item's wedge count (<= TCAP entries x longest B row) must fit int32
  GA_REQUIRE(std::min<int64_t>(w.max_b_row, B.nnz) * TCAP < INT32_MAX, GA_ERR_LIMIT,
             "B rows too long for 32-bit wedge indexing");
  GA_REQUIRE(max_items < INT32_MAX, GA_ERR_LIMIT, "too many work items");
  w.cnt.ensure(sizeof(int32_t) * m);
  w.start.ensure(sizeof(int32_t) * (m + 1));
  w.flag.ensure(sizeof(int32_t) * m);
  w.rstart.ensure(sizeof(int32_t) * (m + 1));
  w.rows.ensure(sizeof(int32_t) * m);
  w.item_row.ensure(sizeof(int32_t) * max_items);
  w.item_t0.ensure(sizeof(int32_t) * max_items);
  // counters: MXM_NSEG item-segment lines (32 u32 each), MXM_NSEG row-segment
  // lines, then the extra-item cursor on a line of its own
  const int ctr_words = MXM_CTR_WORDS;
  w.counter.ensure(sizeof(uint32_t) * ctr_words);
  w.scan_tmp.ensure(sizeof(long long) * std::max<int64_t>(ceil_div(m, PLAN_T), 1));  // item plan and row scans
  // MODE 2 (triangle support) runs on the listed items only
  const long long wave_flops = (M.ncols <= WR_MAXCOLS && !w.fixed_args && MODE != 2)
                                   ? (w.wave_flops >= 0 ? w.wave_flops : ctx().wave_flops)
                                   : -1;
  // direct mode (no wavefront bin): items 0..m-1 are the rows' first chunks and
  // only the extra chunks of long rows are listed -- two small launches
  // instead of binning, two scans and two compactions
  const bool direct = wave_flops < 0 && ctx().direct_items && !w.fixed_args && MODE != 2;
  GA_REQUIRE(MODE != 2 || lower_out != nullptr, GA_ERR_STATE, "triangle support needs the lower-half split");
  GA_REQUIRE(!w.fixed_args || (w.balance && !w.reuse_plan), GA_ERR_STATE,
             "fixed-argument launches are for work-balanced repeated products");
  GA_REQUIRE(w.skip == nullptr || w.fixed_args, GA_ERR_STATE, "a skip flag needs fixed-argument launches");
  // an immutable operand set (reuse_plan) keeps the row bins and item lists
  // of its first launch; work-balanced lists change every launch
  const bool plan_hit = w.reuse_plan && !w.balance && !direct && w.plan_valid && w.plan_rp == A.rp &&
                        w.plan_m == m && w.plan_nnz == A.nnz && w.plan_tcap == TCAP && w.plan_wf == wave_flops;
  const bool keep_plan = w.reuse_plan && !w.balance && !direct;
  if (keep_plan && !plan_hit) w.plan_rows = w.plan_items = -1;
  if (keep_plan) {
    w.plan_valid = true;
    w.plan_rp = A.rp;
    w.plan_m = m;
    w.plan_nnz = A.nnz;
    w.plan_tcap = TCAP;
    w.plan_wf = wave_flops;
  }
  // the listed path's item count also clears the counters and the per-row
  // work and computes the lower-half row ends; other paths launch them
  const bool listed = !direct && !plan_hit && wave_flops < 0;
  if (!listed) {
    if (!w.counters_clean) fill_u32(w.counter.p, 0u, ctr_words, s);
    if (lower_out) {
      hipLaunchKernelGGL(k_row_split, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, M.rp, M.col, m, lower_out);
      GA_HIP(hipGetLastError());
    }
  }
  w.counters_clean = false;  // the launches below claim through them
  if (lower_out) Mend = lower_out;
  // work-balanced items (listed path of repeated products): the previous
  // launch's wedges per row set this launch's split
  unsigned long long* rw_cur = nullptr;
  const unsigned long long* rw_prev = nullptr;
  if (w.fixed_args) {
    w.rwork[0].ensure(sizeof(unsigned long long) * m);
    rw_cur = w.rwork[0].as<unsigned long long>();
    rw_prev = rw_cur;
  } else if (w.balance && wave_flops < 0 && !direct) {
    w.rwork[0].ensure(sizeof(unsigned long long) * m);
    w.rwork[1].ensure(sizeof(unsigned long long) * m);
    rw_cur = w.rwork[w.rwork_cur].as<unsigned long long>();
    rw_prev = w.rwork_have ? w.rwork[w.rwork_cur ^ 1].as<unsigned long long>() : nullptr;
    if (!listed) fill_u64(rw_cur, 0ull, m, s);
    w.rwork_cur ^= 1;
    w.rwork_have = true;
  }
  int32_t* n_extra = reinterpret_cast<int32_t*>(w.counter.as<unsigned int>() + 2 * MXM_NSEG * MXM_CTR_STRIDE + 1);
  if (direct) {
    hipLaunchKernelGGL(k_extra_items, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, A.rp, M.rp, Mend, m, TCAP,
                       w.item_row.as<int32_t>(), w.item_t0.as<int32_t>(), n_extra);
    GA_HIP(hipGetLastError());
  } else if (plan_hit) {
    // same operands as the previous launch: its row bins and items stand
  } else if (wave_flops < 0) {
    // no wavefront bin: items of every non-empty row, listed in row order
    const int64_t nbp = ceil_div(m, PLAN_T);
    long long* bsum = w.scan_tmp.as<long long>();
    hipLaunchKernelGGL(k_item_plan_count, dim3((unsigned)nbp), dim3(PLAN_T), 0, s, A.rp, M.rp,
                       lower_out ? nullptr : Mend, m, TCAP, w.cnt.as<int32_t>(), rw_prev, ctx().item_wedges,
                       w.counter.as<uint32_t>(), ctr_words, rw_cur, lower_out ? M.col : nullptr, lower_out, bsum,
                       w.skip, MODE == 2 ? 1 : 0);
    const bool offsets = nbp > PLAN_FUSED_TILES;
    if (offsets) scan_tile_offsets(bsum, nbp, w.start.as<int32_t>() + m, s);
    hipLaunchKernelGGL(k_item_plan_fill, dim3((unsigned)nbp), dim3(PLAN_T), 0, s, A.rp,
                       (const int32_t*)w.cnt.as<int32_t>(), m, (const long long*)bsum, offsets ? 1 : 0,
                       w.start.as<int32_t>(), w.item_row.as<int32_t>(), w.item_t0.as<int32_t>(), w.skip,
                       MODE == 2 ? (const int32_t*)lower_out : (const int32_t*)nullptr);
    GA_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_bin_rows, dim3(grid_for(m, 4, 16384)), dim3(256), 0, s, A.rp, A.col, B.rp, M.rp, Mend,
                       row_group, BATCHED ? gbase : nullptr, m, TCAP, wave_flops, w.flag.as<int32_t>(),
                       w.cnt.as<int32_t>());
    GA_HIP(hipGetLastError());
    scan_exclusive_i32(w.flag.as<int32_t>(), w.rstart.as<int32_t>(), m, w.scan_tmp, s);
    hipLaunchKernelGGL(k_emit_rows, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s,
                       (const int32_t*)w.flag.as<int32_t>(), (const int32_t*)w.rstart.as<int32_t>(), m,
                       w.rows.as<int32_t>());
    scan_exclusive_i32(w.cnt.as<int32_t>(), w.start.as<int32_t>(), m, w.scan_tmp, s);
    hipLaunchKernelGGL(k_item_fill, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, A.rp,
                       (const int32_t*)w.cnt.as<int32_t>(), (const int32_t*)w.start.as<int32_t>(), m,
                       w.item_row.as<int32_t>(), w.item_t0.as<int32_t>());
    GA_HIP(hipGetLastError());
  }
  const bool wave_on = wave_flops >= 0;
  if (keep_plan && !plan_hit) {
    // the sizes of the kept plan, once (later launches skip an empty kernel)
    auto* h = static_cast<int32_t*>(pinned_scratch());
    h[0] = 0;
    GA_HIP(hipMemcpyAsync(h + 1, w.start.as<int32_t>() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (wave_on) GA_HIP(hipMemcpyAsync(h, w.rstart.as<int32_t>() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
    w.plan_rows = h[0];
    w.plan_items = h[1];
  }
  const bool run_rows = wave_on && !(plan_hit && w.plan_rows == 0);
  const bool run_items = !(plan_hit && w.plan_items == 0);

  auto krows = k_mxm_rows<T, PA, PB, MODE, BATCHED>;
  // rank lookup when the bitmap is at most 16 KB and every mask row fits
  // the 16-bit word rank bases
  // a bitmap narrower than the mask's columns (MXM_BITMAP_MAX_COLS bits for
  // wider matrices) is a filter: column c sets bit c & bm_and
  const int bm_and = (int64_t)M.ncols <= 32 * (int64_t)bitmap_words ? -1 : 32 * bitmap_words - 1;
  // (a sum needs no positions: it tests the exact bitmap alone, in less LDS;
  // products with 4-byte accumulators -- supports -- take the rank lookup up
  // to 2^18 columns: R-MAT 18 k-truss k = 3 12.1 -> 7.1 ms)
  // a sum takes the rank instantiation for any exact bitmap: its bitmap
  // test is the same, without rank bases
  // (1024-entry items of 4-byte supports: up to 2^19 columns, 96 KB of
  // bitmap and rank bases beside their 44 KB of arrays)
  const int rank_words = MODE == 1 ? (int)(MXM_SUM_BITMAP_MAX_COLS / 32)
                                   : (sizeof(T) == 4 ? (THREADS >= 1024 ? 2 * MXM_RANK_MAX_WORDS_POS
                                                                        : MXM_RANK_MAX_WORDS_POS)
                                                     : MXM_RANK_MAX_WORDS);
  const bool rank = bitmap_words > 0 && bitmap_words <= rank_words && (MODE == 1 || w.max_mask_row < 65536) &&
                    bm_and == -1;
  auto kitems = rank ? k_mxm_items<T, PA, PB, MODE, THREADS, BATCHED, true>
                     : k_mxm_items<T, PA, PB, MODE, THREADS, BATCHED, false>;
  const int bm_words = bitmap_words + (rank ? 1 : 0);  // rank: one zero word for dropped wedges
  const size_t dyn =
      sizeof(uint32_t) * (size_t)bm_words + (rank && MODE != 1 ? sizeof(uint16_t) * ((bm_words + 1) & ~1) : 0);
  Context& c = ctx();
  const int occ_r = occupancy_of(krows, 64 * WR_WAVES, 0, c.device);
  const int occ_i = occupancy_of(kitems, THREADS, dyn, c.device);
  const int64_t grid_r = std::max<int64_t>(1, std::min<int64_t>((int64_t)c.num_cu * std::max(occ_r, 1),
                                                               ceil_div(m, WR_WAVES)));
  const int64_t grid_i = std::max<int64_t>(1, std::min<int64_t>((int64_t)c.num_cu * std::max(occ_i, 1), max_items));
  // claim sizes: short rows about 8 claims per resident wave (at most 32
  // rows per claim); items one at a time -- they are long and uneven, and
  // claiming several at once leaves a tail of heavy items on few workgroups
  const int chunk_fix = w.row_chunk > 0 ? w.row_chunk : c.row_chunk;
  const int chunk_r = chunk_fix > 0 ? chunk_fix
                                    : (int)std::max<int64_t>(1, std::min<int64_t>(32, m / (grid_r * WR_WAVES * 8)));
  const int chunk_i = 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c.prof_events_on) {
    e0 = take_event(c);
    e1 = take_event(c);
    GA_HIP(hipEventRecord(e0, s));
  }
  unsigned long long* wc = c.prof_bytes ? c.prof_dev.as<unsigned long long>() : nullptr;
  const int sum_sub = w.sum_sub >= 0 ? w.sum_sub : c.sum_sub;
  if (run_rows && MODE == 1 && PA && PB && (sum_sub == 0 || sum_sub == 2)) {
    // pattern sum mode: rows packed per wavefront, 64 x K wedges per trip
    auto launch_wave = [&](auto kern) {  // (trunc is compiled into the kernel)
      const int occ = occupancy_of(kern, 64 * WR_WAVES, 0, c.device);
      const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)c.num_cu * std::max(occ, 1),
                                                                 ceil_div(m, WR_WAVES)));
      // rows per claim: at most 64 (one lane each); by default about four
      // claims per resident wave, 2 to 16 rows (a wavefront that packs two
      // rows of a small graph finishes sooner than two waves of one row each
      // that then drain the claim segments: Erdos-Renyi n = 1000 triangle
      // count 18.8 -> 14.6 us per launch)
      const int chunk = std::min(64, chunk_fix > 0 ? chunk_fix
                                                   : (int)std::max<int64_t>(
                                                         2, std::min<int64_t>(16, m / (grid * WR_WAVES * 4))));
      // a kept plan knows its short-row count: when every wavefront of the
      // launch gets at most one claim anyway, hand the claims out statically
      // (no claim atomics, no drain of the segment counters: the latency of a
      // small graph's count)
      const int static_rows =
          (keep_plan && w.plan_rows >= 0 && w.plan_rows <= grid * WR_WAVES * (int64_t)chunk) ? 1 : 0;
      // one statically split kernel is the whole launch: it publishes the
      // total itself (no claim counter touched, nothing left to clear)
      const bool pub = w.pub_host && static_rows && !run_items && !BATCHED && !row_group && !wc;
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * WR_WAVES), 0, s, A.rp, A.col, B.rp, B.col, M.rp,
                         M.col, Mend, sums, row_group, gbase, (const int32_t*)w.rows.as<int32_t>(),
                         (const int32_t*)(w.rstart.as<int32_t>() + m),
                         w.counter.as<unsigned int>() + MXM_NSEG * MXM_CTR_STRIDE, wc, chunk, static_rows,
                         pub ? w.pub_host : (long long*)nullptr, w.pub_seq);
      if (pub) {
        w.pub_done = true;
        w.counters_clean = true;  // cleared before the launch, untouched by it
      }
    };
    if (sum_sub == 2) {
      if (trunc) launch_wave(k_sum_wave<BATCHED, 4, true>);
      else launch_wave(k_sum_wave<BATCHED, 4, false>);
    } else {
      if (trunc) launch_wave(k_sum_wave<BATCHED, 8, true>);
      else launch_wave(k_sum_wave<BATCHED, 8, false>);
    }
  } else if (run_rows && MODE == 1 && PA && PB && sum_sub >= 8) {
    // pattern sum mode, entry by entry: sum_sub lanes per short row
    auto launch_entries = [&](auto kern, int groups) {
      const int occ = occupancy_of(kern, 64 * WR_WAVES, 0, c.device);
      const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)c.num_cu * std::max(occ, 1),
                                                                 ceil_div(m, groups)));
      const int chunk = chunk_fix > 0 ? chunk_fix
                                      : (int)std::max<int64_t>(1, std::min<int64_t>(32, m / (grid * groups * 8)));
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * WR_WAVES), 0, s, A.rp, A.col, B.rp, B.col, M.rp,
                         M.col, Mend, sums, row_group, gbase, (const int32_t*)w.rows.as<int32_t>(),
                         (const int32_t*)(w.rstart.as<int32_t>() + m),
                         w.counter.as<unsigned int>() + MXM_NSEG * MXM_CTR_STRIDE, wc, trunc, chunk);
    };
    switch (sum_sub) {
      case 8: launch_entries(k_sum_entries<BATCHED, 8>, 64 * WR_WAVES / 8); break;
      case 16: launch_entries(k_sum_entries<BATCHED, 16>, 64 * WR_WAVES / 16); break;
      case 32: launch_entries(k_sum_entries<BATCHED, 32>, 64 * WR_WAVES / 32); break;
      default: launch_entries(k_sum_entries<BATCHED, 64>, 64 * WR_WAVES / 64); break;
    }
  } else if (run_rows && MODE == 1 && PA && PB) {
    // pattern sum mode: 16 lanes per short row, four rows per wavefront
    auto ksub = k_sum_subrows<BATCHED>;
    const int occ_s = occupancy_of(ksub, 64 * WR_WAVES, 0, c.device);
    const int64_t grid_s = std::max<int64_t>(1, std::min<int64_t>((int64_t)c.num_cu * std::max(occ_s, 1),
                                                                 ceil_div(m, SR_GROUPS)));
    // rows per claim: about 8 claims per resident group, at most 4: the rows
    // in flight on an XCD are its groups x the claim size, and their graphs'
    // arrays share the XCD's L2 (4096 kNN graphs: 2.52 ms at 16 rows per
    // claim, 2.28 ms at 4)
    const int chunk_s = chunk_fix > 0 ? chunk_fix
                                      : (int)std::max<int64_t>(1, std::min<int64_t>(4, m / (grid_s * SR_GROUPS * 8)));
    hipLaunchKernelGGL(ksub, dim3((unsigned)grid_s), dim3(64 * WR_WAVES), 0, s, A.rp, A.col, B.rp, B.col, M.rp,
                       M.col, Mend, sums, row_group, gbase, (const int32_t*)w.rows.as<int32_t>(),
                       (const int32_t*)(w.rstart.as<int32_t>() + m),
                       w.counter.as<unsigned int>() + MXM_NSEG * MXM_CTR_STRIDE, wc, trunc, chunk_s);
  } else if (run_rows)
    hipLaunchKernelGGL(krows, dim3((unsigned)grid_r), dim3(64 * WR_WAVES), 0, s, A.rp, A.col,
                     static_cast<const T*>(A.val), B.rp, B.col, static_cast<const T*>(B.val), M.rp, M.col, Mend,
                     static_cast<T*>(Cval), sums, row_group, gbase, (const int32_t*)w.rows.as<int32_t>(),
                     (const int32_t*)(w.rstart.as<int32_t>() + m), w.counter.as<unsigned int>() + MXM_NSEG * MXM_CTR_STRIDE, wc,
                     trunc, chunk_r, mirror, opos);
  GA_HIP(hipGetLastError());
  if (run_items)
    hipLaunchKernelGGL(kitems, dim3((unsigned)grid_i), dim3(THREADS), dyn, s, A.rp, A.col,
                     static_cast<const T*>(A.val), B.rp, B.col, static_cast<const T*>(B.val), M.rp, M.col,
                     static_cast<T*>(Cval), sums, row_group, gbase, w.item_row.as<int32_t>(), w.item_t0.as<int32_t>(),
                     direct ? (const int32_t*)n_extra : (const int32_t*)(w.start.as<int32_t>() + m),
                     w.counter.as<unsigned int>(), wc, bitmap_words, bm_and, Mend, trunc, chunk_i, direct ? (int)m : 0,
                     mirror, opos, rw_cur, w.skip);
  GA_HIP(hipGetLastError());
  if (c.prof_events_on) {
    GA_HIP(hipEventRecord(e1, s));
    c.prof_events.emplace_back(e0, e1);
    c.prof_launches += 1;
  }
  if (c.prof_bytes) {
    // host-known algorithmic bytes: A and M streamed, B row bounds, output
    const int64_t out = MODE != 1 ? (int64_t)sizeof(T) * M.nnz : 8;
    c.prof_bytes_host += 4 * (A.nrows + 1) + 4 * A.nnz + 8 * A.nnz + 4 * (M.nrows + 1) + 4 * M.nnz + out;
    c.prof_bytes_launches += 1;
  }
}

template <typename T, int MODE>
static void dispatch_pat(const DCsr& A, const DCsr& B, const DCsr& M, void* Cval, unsigned long long* sums,
                         const int32_t* row_group, MxmWork& w, int bm, hipStream_t s) {
  const bool pa = A.val == nullptr, pb = B.val == nullptr;
  // pattern products with 4-byte (or no) accumulators use the wide items;
  // 8-byte valued accumulators keep 256 to bound the LDS
  constexpr int TH = (MODE == 1 || sizeof(T) == 4) ? MXM_WIDE : 256;
  if (pa && pb) launch_binned<T, true, true, MODE, TH, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s, nullptr, 0);
  else if (pa) launch_binned<T, true, false, MODE, TH, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s, nullptr, 0);
  else if (pb) launch_binned<T, false, true, MODE, TH, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s, nullptr, 0);
  else launch_binned<T, false, false, MODE, TH, false>(A, B, M, Cval, sums, row_group, nullptr, w, bm, s, nullptr, 0);
}

// One workgroup: the total to host memory (value, then the sequence number
// with release order), then sums[0] and the claim counters cleared.
__global__ void k_sum_publish(unsigned long long* __restrict__ sums, uint32_t* __restrict__ ctr, int ctr_words,
                              long long* __restrict__ host, long long seq) {
  for (int q = threadIdx.x; q < ctr_words; q += blockDim.x) ctr[q] = 0u;
  if (threadIdx.x == 0) {
    const unsigned long long v = sums[0];
    sums[0] = 0ull;
    __hip_atomic_store(&host[0], (long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[1], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

namespace {
thread_local long long t_publish_seq = 0;
}

long long next_host_seq() { return ++t_publish_seq; }

void wait_host_word(const long long* w, long long seq, hipStream_t s) {
  // every 1024 polls ask the stream whether it failed (a fault must end the
  // wait) or drained (then the word is there)
  for (int64_t it = 1;; ++it) {
    if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == seq) return;
    if ((it & 1023) == 0) {
      const hipError_t e = hipStreamQuery(s);
      if (e == hipSuccess) {
        GA_REQUIRE(__atomic_load_n(w, __ATOMIC_ACQUIRE) == seq, GA_ERR_HIP, "the device result did not arrive");
        return;
      }
      if (e != hipErrorNotReady) GA_HIP(e);
    }
  }
}

unsigned long long publish_sum_and_wait(MxmWork& w, unsigned long long* sums, hipStream_t s) {
  long long* hdev = nullptr;
  long long* h = progress_words(&hdev) + 4;  // words 4 (value) and 5 (sequence)
  const long long seq = next_host_seq();
  w.counter.ensure(sizeof(uint32_t) * MXM_CTR_WORDS);
  hipLaunchKernelGGL(k_sum_publish, dim3(1), dim3(256), 0, s, sums, w.counter.as<uint32_t>(), (int)MXM_CTR_WORDS,
                     hdev + 4, seq);
  GA_HIP(hipGetLastError());
  w.counters_clean = true;
  wait_host_word(&h[1], seq, s);
  return (unsigned long long)__atomic_load_n(&h[0], __ATOMIC_RELAXED);
}

__global__ void k_count_wedges(const int32_t* __restrict__ Arp, const int32_t* __restrict__ Acol,
                               const int32_t* __restrict__ Brp, int64_t m, unsigned long long* __restrict__ out,
                               const int32_t* __restrict__ row_group = nullptr,
                               const int32_t* __restrict__ gbase = nullptr) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  unsigned long long acc = 0;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * wpb) {
    const int bb = gbase ? gbase[row_group[i]] : 0;
    for (int t = Arp[i] + lane; t < Arp[i + 1]; t += 64) acc += Brp[bb + Acol[t] + 1] - Brp[bb + Acol[t]];
  }
  acc = wave_reduce_sum(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

void mxm_batched_sum_dev(const DCsr& L, const int32_t* row_group, const int32_t* gbase, int64_t max_cols,
                         unsigned long long* sums, MxmWork& w, hipStream_t s) {
  if (L.nrows == 0 || L.nnz == 0) return;
  const int bm = (max_cols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(max_cols, 32) : (int)(MXM_FILTER_COLS / 32);
  if (ctx().prof_bytes)
    hipLaunchKernelGGL(k_count_wedges, dim3(grid_for(L.nrows, 4, 4096)), dim3(256), 0, s, L.rp, L.col, L.rp,
                       L.nrows, ctx().prof_dev.as<unsigned long long>() + 1, row_group, gbase);
  switch (w.threads) {
    case 64:
      launch_binned<int64_t, true, true, 1, 64, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s, nullptr, 0);
      break;
    case 128:
      launch_binned<int64_t, true, true, 1, 128, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s, nullptr, 0);
      break;
    case 256:
      launch_binned<int64_t, true, true, 1, 256, true>(L, L, L, nullptr, sums, row_group, gbase, w, bm, s, nullptr, 0);
      break;
    default: throw Error{GA_ERR_ARG, "workgroup size must be 64, 128 or 256"};
  }
}

// ---- symmetric support: S<C> = C*C for a symmetric pattern C ---------------
// S is symmetric, so only the strict lower half of the mask is computed
// (mask prefix Mend[i] = first entry of row i with column >= i, B rows cut
// at the last lower mask column) and the values are mirrored to the upper
// half.  Same values as the full product, about half the wedges.

// mirror[t] = position of the transposed entry of entry t (symmetric pattern)
__global__ void k_mirror_index(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                               int32_t* __restrict__ mirror) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * wpb) {
    for (int t = rp[i] + lane; t < rp[i + 1]; t += 64) {
      const int j = col[t];
      int lo = rp[j], hi = rp[j + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      mirror[t] = (lo < rp[j + 1] && col[lo] == i) ? lo : -1;
    }
  }
}

// Entry-parallel: each lower entry t = (i, j), j < i, finds its row by a
// binary search of the row pointers and its transposed entry in row j, and
// writes both directions' mirror indices (a wavefront per row serialised the
// searches of hub rows: 0.70 ms on R-MAT 16 against 0.04 ms).
__global__ void k_mirror_lower(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                               int64_t nnz, int32_t* __restrict__ mirror) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nnz; t += (int64_t)gridDim.x * blockDim.x) {
    const int j = col[t];
    int64_t lo = 0, hi = n;  // row of t: the last i with rp[i] <= t
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rp[mid + 1] <= t) lo = mid + 1;
      else hi = mid;
    }
    const int i = (int)lo;
    if (j >= i) continue;
    int a = rp[j], b = rp[j + 1];
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (col[mid] < i) a = mid + 1;
      else b = mid;
    }
    if (a < rp[j + 1] && col[a] == i) {
      mirror[t] = a;
      mirror[a] = (int32_t)t;
    }
  }
}

void mirror_index_dev(const DCsr& C, int32_t* mirror, hipStream_t s) {
  if (C.nrows == 0 || C.nnz == 0) return;
  fill_u32(mirror, 0xffffffffu, C.nnz, s);  // -1 where a pattern has no transposed entry
  hipLaunchKernelGGL(k_mirror_lower, dim3(grid_for(C.nnz, 256, 8192)), dim3(256), 0, s, C.rp, C.col, C.nrows, C.nnz,
                     mirror);
  GA_HIP(hipGetLastError());
}

size_t mxm_workspace_bytes(int64_t m, int64_t nnz_a, bool with_mend) {
  // launch_binned's buffers at the 64-lane item width (the most items)
  // with_mend: the symmetric support pass, which splits rows by measured work
  const int64_t max_items = (with_mend ? MXM_SPLIT_MAX : 1) * (m + ceil_div(nnz_a, 64)) + 1;
  size_t b = 3 * lease_bytes(sizeof(int32_t) * m) + 2 * lease_bytes(sizeof(int32_t) * (m + 1));
  b += 2 * lease_bytes(sizeof(int32_t) * max_items);
  b += lease_bytes(sizeof(uint32_t) * MXM_CTR_WORDS);
  b += lease_bytes(sizeof(long long) * std::max<int64_t>(ceil_div(m, PLAN_T), 1));  // >= scan_tmp_bytes(m)
  if (with_mend) b += lease_bytes(sizeof(int32_t) * m) + 2 * lease_bytes(sizeof(unsigned long long) * m);
  return b;
}

// Upper entries take the support of their transposed (lower) entry: the
// transposed entry of p lies in a later row exactly when p is upper.  The
// current entry count is rp[n] (device-driven steps size the grid for the
// first graph); skip: as in k_mxm_items.
__global__ void k_sup_mirror(const int32_t* __restrict__ rp, int64_t n, const int32_t* __restrict__ mirror,
                             int32_t* __restrict__ S, const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t nnz = rp[n];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
    const int m = mirror[p];
    if (m > p) S[p] = S[m];
  }
}

void support_symmetric_dev(const DCsr& C, const int32_t* mirror, int32_t* S, MxmWork& w, hipStream_t s,
                           bool s_zeroed) {
  const int64_t n = C.nrows;
  GA_REQUIRE(C.nrows == C.ncols, GA_ERR_DIM, "symmetric support needs a square matrix");
  if (C.nnz > 0 && !s_zeroed) fill_u32(S, 0u, C.nnz, s);  // s_zeroed: the caller's compaction cleared S
  if (n == 0 || C.nnz == 0) return;
  w.mend.ensure(sizeof(int32_t) * n);
  int32_t* split = w.mend.as<int32_t>();
  const int bm = (C.ncols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(C.ncols, 32) : (int)(MXM_FILTER_COLS / 32);
  Context& c = ctx();
  if (c.prof_bytes)
    hipLaunchKernelGGL(k_count_wedges, dim3(grid_for(n, 4, 4096)), dim3(256), 0, s, C.rp, C.col, C.rp, n,
                       c.prof_dev.as<unsigned long long>() + 1);
  Context& cx = ctx();
  const int64_t tri_min = w.tri_min_nnz >= 0 ? w.tri_min_nnz : cx.tri_list_min_nnz;
  if (cx.support_tri > 0 || (cx.support_tri < 0 && C.nnz >= tri_min)) {
    // each triangle i > k > j once, from the wedge (i, k, j) of C<L> = L*L
    // (L = the strict lower half): +1 to (i, j) (mask), (i, k) (A) and
    // (k, j) (B); then the upper half from the lower.  Up to 2^19 columns
    // the mask positions come from an exact bitmap with rank bases (1024-entry
    // items past 2^18 columns: R-MAT 19 k-truss); wider graphs filter and
    // search
    const int64_t words = ceil_div(C.ncols, 32);
    if (C.ncols > MXM_BITMAP_MAX_COLS && words <= 2 * MXM_RANK_MAX_WORDS_POS && w.max_mask_row < 65536 &&
        6 * (words + 2) + 48 * 1024 <= cx.lds_per_block)
      launch_binned<int32_t, true, true, 2, 1024, false>(C, C, C, S, nullptr, nullptr, nullptr, w, (int)words, s,
                                                         nullptr, 0, nullptr, nullptr, split);
    else
      launch_binned<int32_t, true, true, 2, MXM_WIDE, false>(C, C, C, S, nullptr, nullptr, nullptr, w, bm, s,
                                                              nullptr, 0, nullptr, nullptr, split);
    const unsigned g = (unsigned)std::min<int64_t>(grid_for(C.nnz, 256, 8192), 4 * (int64_t)cx.num_cu);
    hipLaunchKernelGGL(k_sup_mirror, dim3(g), dim3(256), 0, s, C.rp, n, mirror, S, w.skip);
    GA_HIP(hipGetLastError());
    return;
  }
  launch_binned<int32_t, true, true, 0, MXM_WIDE, false>(C, C, C, S, nullptr, nullptr, nullptr, w, bm, s, nullptr,
                                                          1, mirror, nullptr, split);
}

// ---- support of the affected entries only ----------------------------------
// After a compaction only the entries with a "dirty" endpoint (a row that
// lost entries) can change support: S(i,j) = |N(i) & N(j)| is unchanged when
// neither N(i) nor N(j) changed.  Each affected undirected edge {i,j}, i > j,
// is computed once, by row i as a lower entry, as in the full symmetric pass:
// row i's mask is the part of its lower half with a dirty endpoint, so B rows
// are cut at the largest such column (never later than in the full pass).
// (Giving the edges of a clean row to its dirty partner's upper half forms
// fewer products, but puts the long rows of hub vertices on the slow global
// accumulator and runs slower.)  The selection becomes a mask CSR M'
// (columns + the C position of each entry), the other supports stay as
// carried over from the previous pass, and the product writes each value to
// its C entry and the transposed one.
__global__ void k_sel_count(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const uint8_t* __restrict__ dirty, int64_t n, int32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * wpb) {
    const bool di = dirty[i] != 0;
    const int b = rp[i], e = rp[i + 1];
    int c = 0;
    for (int t = b + lane; t < e; t += 64) {
      const int j = col[t];
      c += (j < i && (di || dirty[j])) ? 1 : 0;
    }
    c = wave_reduce_sum(c);
    if (lane == 0) cnt[i] = c;
  }
}

__global__ void k_sel_fill(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                           const uint8_t* __restrict__ dirty, const int32_t* __restrict__ mirror, int64_t n,
                           const int32_t* __restrict__ mrp, int32_t* __restrict__ mcol, int32_t* __restrict__ opos,
                           int32_t* __restrict__ S) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * wpb) {
    int dst = mrp[i];
    if (mrp[i + 1] == dst) continue;
    const bool di = dirty[i] != 0;
    const int b = rp[i], e = rp[i + 1];
    for (int t = b + lane; t - lane < e; t += 64) {
      bool sel = false;
      int j = 0;
      if (t < e) {
        j = col[t];
        sel = j < i && (di || dirty[j]);
      }
      const unsigned long long m = __ballot(sel);
      if (sel) {
        const int r = dst + __popcll(m & lt);
        mcol[r] = j;
        opos[r] = t;
        S[t] = 0;  // accumulated into (multi-item rows) or overwritten
        const int mt = mirror[t];
        if (mt >= 0) S[mt] = 0;
      }
      dst += __popcll(m);
    }
  }
}

void support_selected_dev(const DCsr& C, const int32_t* mirror, const uint8_t* dirty, int32_t* S, SelWork& sw,
                          MxmWork& w, hipStream_t s) {
  const int64_t n = C.nrows;
  GA_REQUIRE(C.nrows == C.ncols, GA_ERR_DIM, "symmetric support needs a square matrix");
  if (n == 0 || C.nnz == 0) return;
  sw.cnt.ensure(sizeof(int32_t) * n);
  sw.mrp.ensure(sizeof(int32_t) * (n + 1));
  sw.mcol.ensure(sizeof(int32_t) * C.nnz);
  sw.opos.ensure(sizeof(int32_t) * C.nnz);
  hipLaunchKernelGGL(k_sel_count, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, C.rp, C.col, dirty, n,
                     sw.cnt.as<int32_t>());
  scan_exclusive_i32(sw.cnt.as<int32_t>(), sw.mrp.as<int32_t>(), n, sw.scan_tmp, s);
  hipLaunchKernelGGL(k_sel_fill, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, C.rp, C.col, dirty, mirror, n,
                     (const int32_t*)sw.mrp.as<int32_t>(), sw.mcol.as<int32_t>(), sw.opos.as<int32_t>(), S);
  GA_HIP(hipGetLastError());
  // M'.nnz is only a bound here (the exact count stays on the device); the
  // algorithmic accounting is the reference's full product, M = C
  const DCsr M{n, n, C.nnz, sw.mrp.as<int32_t>(), sw.mcol.as<int32_t>(), nullptr};
  const int bm = (C.ncols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(C.ncols, 32) : (int)(MXM_FILTER_COLS / 32);
  Context& c = ctx();
  if (c.prof_bytes)
    hipLaunchKernelGGL(k_count_wedges, dim3(grid_for(n, 4, 4096)), dim3(256), 0, s, C.rp, C.col, C.rp, n,
                       c.prof_dev.as<unsigned long long>() + 1);
  launch_binned<int32_t, true, true, 0, MXM_WIDE, false>(C, C, M, S, nullptr, nullptr, nullptr, w, bm, s, nullptr,
                                                          1, mirror, sw.opos.as<int32_t>());
}

size_t sel_workspace_bytes(int64_t n, int64_t nnz) {
  return lease_bytes(sizeof(int32_t) * n) + lease_bytes(sizeof(int32_t) * (n + 1)) +
         2 * lease_bytes(sizeof(int32_t) * nnz) + scan_tmp_bytes(n);
}

void mxm_masked_dev(const DCsr& A, const DCsr& B, const DCsr& M, int dtype, int mode, void* Cval,
                    unsigned long long* sums, const int32_t* row_group, MxmWork& w, hipStream_t s) {
  GA_REQUIRE(A.nrows == M.nrows, GA_ERR_DIM, "A.nrows != M.nrows");
  GA_REQUIRE(A.ncols == B.nrows, GA_ERR_DIM, "A.ncols != B.nrows");
  GA_REQUIRE(B.ncols == M.ncols, GA_ERR_DIM, "B.ncols != M.ncols");
  const int64_t m = A.nrows;
  if (mode == 0 && M.nnz > 0) {
    const size_t es = dtype_size(dtype);
    if (es == 4) fill_u32(Cval, 0u, M.nnz, s);
    else fill_u64(Cval, 0ull, M.nnz, s);
  }
  if (m == 0 || A.nnz == 0 || M.nnz == 0) return;
  if (ctx().prof_bytes)
    hipLaunchKernelGGL(k_count_wedges, dim3(grid_for(m, 4, 4096)), dim3(256), 0, s, A.rp, A.col, B.rp, m,
                       ctx().prof_dev.as<unsigned long long>() + 1);
  const int bm = (M.ncols <= MXM_BITMAP_MAX_COLS) ? (int)ceil_div(M.ncols, 32) : (int)(MXM_FILTER_COLS / 32);
  if (mode == 1) {
    // (the exact bitmap while it fits the workgroup's LDS beside ~8 KB of
    // the kernel's own arrays)
    const int64_t words = ceil_div(M.ncols, 32);
    const bool fits = M.ncols <= MXM_SUM_BITMAP_MAX_COLS && 4 * (words + 1) + 16384 <= ctx().lds_per_block;
    const int bm_sum = fits ? (int)words : bm;
    if (bm_sum > 8192 && A.val == nullptr && B.val == nullptr) {
      // a bitmap of more than 32 KB leaves room for one or two workgroups per
      // CU: 1024-entry items keep 16 wavefronts on each (R-MAT 19 count 4.51
      // -> 3.99 ms against 512-entry items; R-MAT 20's 128 KB bitmap)
      launch_binned<int64_t, true, true, 1, 1024, false>(A, B, M, nullptr, sums, row_group, nullptr, w, bm_sum, s,
                                                         nullptr, 0);
      return;
    }
    if (fits && bm_sum <= 4096 && A.val == nullptr && B.val == nullptr) {
      // an exact bitmap of at most 16 KB: 256-entry items, twice the
      // workgroups of 512 (short rows leave most of a wide item idle: R-MAT
      // 16 count 258 -> 205 us, the served pool's scale-17 graph 0.61 -> 0.46
      // ms; with the 32 KB bitmap of R-MAT 18 the LDS holds too few of them,
      // 1.06 -> 1.21 ms, so wider graphs keep 512)
      launch_binned<int64_t, true, true, 1, 256, false>(A, B, M, nullptr, sums, row_group, nullptr, w, bm_sum, s,
                                                        nullptr, 0);
      return;
    }
    dispatch_pat<int64_t, 1>(A, B, M, nullptr, sums, row_group, w, bm_sum, s);
    return;
  }
  switch (dtype) {
    case GA_INT32: dispatch_pat<int32_t, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    case GA_INT64: dispatch_pat<int64_t, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    case GA_FLOAT64: dispatch_pat<double, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    case GA_FLOAT32: dispatch_pat<float, 0>(A, B, M, Cval, sums, row_group, w, bm, s); break;
    default: throw Error{GA_ERR_ARG, "unknown dtype"};
  }
}

}  // namespace ga

using namespace ga;

extern "C" ga_status ga_mxm_masked(const ga_csr* A, const ga_csr* B, const ga_csr* M, ga_dtype dtype, void* Cval) {
  return guarded([&] {
    dtype_size(dtype);
    validate_host_csr(A, "A");
    validate_host_csr(B, "B");
    validate_host_csr(M, "M");
    GA_REQUIRE(A->nrows == M->nrows && A->ncols == B->nrows && B->ncols == M->ncols, GA_ERR_DIM,
               "operand shapes do not match (A: m x k, B: k x n, M: m x n)");
    GA_REQUIRE(M->nnz == 0 || Cval != nullptr, GA_ERR_ARG, "null Cval");
    Context& c = ctx();
    hipStream_t s = c.stream;
    OwnedCsr dA, dB, dM;
    upload_csr(A, dtype, dA, s);
    upload_csr(B, dtype, dB, s);
    upload_csr(M, dtype, dM, s);
    DevBuf dC;
    const size_t es = dtype_size(dtype);
    dC.ensure(es * std::max<int64_t>(M->nnz, 1));
    MxmWork w;
    w.max_mask_row = host_max_row(M);
    w.max_b_row = host_max_row(B);
    DCsr vM = dM.view();
    vM.val = nullptr;
    mxm_masked_dev(dA.view(), dB.view(), vM, dtype, 0, dC.p, nullptr, nullptr, w, s);
    if (M->nnz) GA_HIP(hipMemcpyAsync(Cval, dC.p, es * M->nnz, hipMemcpyDeviceToHost, s));
    GA_HIP(hipStreamSynchronize(s));
  });
}

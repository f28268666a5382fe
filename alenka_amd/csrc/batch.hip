// Batched filter + GROUP BY: many queries over one table in a single persistent launch.
//
// Each query q keeps rows with lo_q <= pred_q <= hi_q and groups them by key_q / div_q
// with SUM((a0 + a1*x) * (b0 + b1*y)) and COUNT(*) - the shape of the reference's
// FILTER + SELECT ... GROUP BY (emit_filter operators.cu:2306, emit_select 1801-2084),
// evaluated for all queries in one pass over the table (a shared scan):
//   * a persistent grid walks the table's 4096-row tiles; the packed words of the shared
//     predicate and key columns are staged in LDS once per tile for every query;
//   * groups are dense over [min(key)/div, max(key)/div] (the zone maps give the range), so
//     a group's slot is its index and output is in key order, as one-segment GROUP BY;
//   * the first `hot` groups of every query (where skewed keys concentrate) are aggregated
//     in LDS and flushed once at the end; the others use global atomics.  The queries are
//     split into groups whose LDS tables fit one block, and each block evaluates one query
//     group over its share of the tiles (each group re-reads the table: the scan is cheap
//     next to the aggregation, and larger tables keep more adds out of global atomics);
//   * when every query sums the same expression, it is computed once per row.
// A prepared batch owns its tables; aln_batch_run replays the launch sequence (table
// fills + scan) from a captured HIP graph when asked to; aln_batch_autotune times the
// grid / hot-set shapes on the real data and returns the fastest.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "aln_internal.h"
#include "stage.h"

namespace aln {

constexpr int BQ_MAX = ALN_BATCH_MAX;
// LDS aggregation: the first `hot` groups of each query (at most BQ_HOT_MAX) have a slot
// (8-byte sum, 4-byte count) in the block's LDS; the queries are split into groups of `qg`
// whose tables fit BQ_HOT_BYTES together, and each block evaluates one query group.
constexpr int BQ_HOT_MAX = 4096;
constexpr int BQ_HOT_BYTES = 48 * 1024;
constexpr int BQ_HOT_DEFAULT = 256;
constexpr int BQ_BLOCK = 512;
constexpr int BQ_ITEMS = STG_TILE / BQ_BLOCK;  // rows per thread of a 4096-row tile

constexpr int BQ_SLOTS = 4;  // distinct columns a batch's queries read (else the general loop)
constexpr int BQ_CHUNK = 4;  // rows per thread decoded together

// A group's accumulators: its sum and count share one 16-byte slot, so a row's two
// atomics touch one cache line of the group table.
struct __align__(16) BSlot {
  unsigned long long sum;
  unsigned int cnt, pad;
};

struct BQuery {
  int32_t pred_col, key_col, x_col, y_col;
  int32_t ps, ks, xs, ys;  // the same columns as slots of BArgs::slot_col
  int64_t lo, hi;
  int64_t div;       // >= 1
  int64_t gbase;     // floor(min(key) / div)
  uint64_t base;     // gbase * div (wrapping): key - base >= 0 for every key of the column
  uint64_t magic;    // ceil(2^64 / div) for div >= 2: (key - base) / div of a 32-bit difference
  uint64_t ngroups;  // dense group range
  int64_t a0, a1, b0, b1;
  BSlot *acc;  // [ngroups]
};

struct BArgs {
  const uint32_t *words[ALN_MAX_COLS];
  const aln_seg *segs[ALN_MAX_COLS];
  const BQuery *q;  // device [nq] (the kernel reads it through a restrict parameter)
  int nq;
  int hot;
  int qg, ngq;  // queries per query group, query groups (block b evaluates group b % ngq)
  int shared_v; // every query sums the same (a0 + a1*x)(b0 + b1*y): computed once per row
  int stage_cols[2];
  int nstage;
  // columns every query shares (-1: per query)
  int pred_col, key_col, x_col, y_col;
  // distinct columns of the batch when there are at most BQ_SLOTS (nslots 0: more)
  int slot_col[BQ_SLOTS];
  int nslots;
  uint32_t seg_rows, tps;
  uint64_t ntiles;
};

// floor(key / div) - gbase = (key - base) / div with key - base >= 0.  A difference below
// 2^32 divides by one multiply-high with magic = ceil(2^64 / div): the estimate exceeds
// x / div by less than x / 2^64 < 1 / div, so its floor is exact (x * div < 2^64).
__device__ __forceinline__ uint64_t group_of(const BQuery &q, int64_t key) {
  const uint64_t x = (uint64_t)key - q.base;
  if (q.div == 1) return x;
  if (x >> 32) return x / (uint64_t)q.div;
  return __umul64hi(x, q.magic);
}

__device__ __forceinline__ int64_t pick(int s, int64_t v0, int64_t v1, int64_t v2, int64_t v3) {
  return s == 0 ? v0 : s == 1 ? v1 : s == 2 ? v2 : v3;
}

__device__ __forceinline__ int64_t bq_value(const BQuery &q, int64_t x, int64_t y) {
  return wmul((int64_t)((uint64_t)q.a0 + (uint64_t)wmul(q.a1, x)),
              (int64_t)((uint64_t)q.b0 + (uint64_t)wmul(q.b1, y)));
}

// Adds (v, 1) of one (query, row) to its group: groups below `hot` go to the query's LDS
// table (hsum/hcnt); group 0 - where Zipf ranks concentrate - is first summed over the
// wave when many lanes hit it, so one lane issues the atomic for all of them.
__device__ __forceinline__ void bq_add(const BQuery &q, int hot, uint64_t g, int64_t v, bool on,
                                       unsigned long long *hsum, unsigned int *hcnt) {
  const uint64_t z = __ballot(on && g == 0);
  if (__popcll(z) >= 8) {
    int64_t s = on && g == 0 ? v : 0;
    for (int o = 32; o > 0; o >>= 1) s = (int64_t)((uint64_t)s + (uint64_t)__shfl_xor(s, o, 64));
    if (__lane_id() == (unsigned)(__ffsll((unsigned long long)z) - 1)) {
      if (hot > 0) {
        atomicAdd(&hsum[0], (unsigned long long)s);
        atomicAdd(&hcnt[0], (unsigned int)__popcll(z));
      } else if (q.ngroups) {
        atomicAdd(&q.acc[0].sum, (unsigned long long)s);
        atomicAdd(&q.acc[0].cnt, (unsigned int)__popcll(z));
      }
    }
    on = on && g != 0;
  }
  if (!on) return;
  if (g < (uint64_t)hot) {
    atomicAdd(&hsum[g], (unsigned long long)v);
    atomicAdd(&hcnt[g], 1u);
  } else if (g < q.ngroups) {
    atomicAdd(&q.acc[g].sum, (unsigned long long)v);
    atomicAdd(&q.acc[g].cnt, 1u);
  }
}

// Persistent grid: block b evaluates query group b % ngq over the tiles b / ngq, + the
// group's block count, ...; its LDS tables (dynamic: qg * hot sums, then qg * hot counts)
// are added to the global ones at the end.
__global__ void __launch_bounds__(BQ_BLOCK, 4) k_batch(BArgs a, const BQuery *__restrict__ qv) {
  __shared__ TileStage ts;
  __shared__ int want[2];
  extern __shared__ __align__(16) unsigned long long bq_lds[];
  const uint32_t ngq = (uint32_t)a.ngq;
  const int q0 = (int)(blockIdx.x % ngq) * a.qg;
  const int q1 = q0 + a.qg < a.nq ? q0 + a.qg : a.nq;
  const int hot = a.hot;
  unsigned long long *hsum = bq_lds;
  unsigned int *hcnt = reinterpret_cast<unsigned int *>(bq_lds + (size_t)a.qg * hot);
  for (int i = threadIdx.x; i < (q1 - q0) * hot; i += BQ_BLOCK) {
    hsum[i] = 0;
    hcnt[i] = 0;
  }
  if (threadIdx.x < 2) want[threadIdx.x] = a.stage_cols[threadIdx.x];
  __syncthreads();
  const uint64_t step = gridDim.x / ngq;  // blocks of this query group (gridDim.x is a multiple of ngq)
  for (uint64_t tile = blockIdx.x / ngq; tile < a.ntiles; tile += step) {
    const uint32_t seg = (uint32_t)(tile / a.tps), t = (uint32_t)(tile % a.tps);
    const uint32_t nrows = a.segs[0][seg].nrows, r0 = t * STG_TILE;
    if (r0 >= nrows) continue;  // uniform across the block
    const uint32_t nt = nrows - r0 < (uint32_t)STG_TILE ? nrows - r0 : (uint32_t)STG_TILE;
    __syncthreads();  // the previous tile's LDS reads are done
    stage_tile<BQ_BLOCK>(ts, a.words, a.segs, want, a.nstage, seg, r0, nt);
    __syncthreads();
    if (a.nslots > 0) {
      // every distinct column decoded once per row (HBM loads of BQ_CHUNK rows in flight),
      // the shared value once per row, then each query's parameters read once per chunk
      for (int ch = 0; ch < BQ_ITEMS / BQ_CHUNK; ++ch) {
        int64_t v0[BQ_CHUNK], v1[BQ_CHUNK], v2[BQ_CHUNK], v3[BQ_CHUNK], vs[BQ_CHUNK];
        bool ok[BQ_CHUNK];
#pragma unroll
        for (int u = 0; u < BQ_CHUNK; ++u) {
          const uint32_t i = (ch * BQ_CHUNK + u) * BQ_BLOCK + threadIdx.x;
          ok[u] = i < nt;
          v0[u] = ok[u] ? tile_get(ts, a.words, a.segs, a.slot_col[0], seg, r0, i) : 0;
          v1[u] = ok[u] && a.nslots > 1 ? tile_get(ts, a.words, a.segs, a.slot_col[1], seg, r0, i) : 0;
          v2[u] = ok[u] && a.nslots > 2 ? tile_get(ts, a.words, a.segs, a.slot_col[2], seg, r0, i) : 0;
          v3[u] = ok[u] && a.nslots > 3 ? tile_get(ts, a.words, a.segs, a.slot_col[3], seg, r0, i) : 0;
          vs[u] = 0;
        }
        if (a.shared_v) {
          const BQuery &q = qv[0];
#pragma unroll
          for (int u = 0; u < BQ_CHUNK; ++u)
            vs[u] = bq_value(q, pick(q.xs, v0[u], v1[u], v2[u], v3[u]), pick(q.ys, v0[u], v1[u], v2[u], v3[u]));
        }
        for (int qi = q0; qi < q1; ++qi) {
          const BQuery &q = qv[qi];
          unsigned long long *qs = hsum + (size_t)(qi - q0) * hot;
          unsigned int *qc = hcnt + (size_t)(qi - q0) * hot;
#pragma unroll
          for (int u = 0; u < BQ_CHUNK; ++u) {
            const int64_t p = pick(q.ps, v0[u], v1[u], v2[u], v3[u]);
            const bool on = ok[u] && p >= q.lo && p <= q.hi;
            const uint64_t g = group_of(q, pick(q.ks, v0[u], v1[u], v2[u], v3[u]));
            const int64_t v = a.shared_v ? vs[u]
                                         : bq_value(q, pick(q.xs, v0[u], v1[u], v2[u], v3[u]),
                                                    pick(q.ys, v0[u], v1[u], v2[u], v3[u]));
            bq_add(q, hot, g, v, on, qs, qc);
          }
        }
      }
      continue;
    }
    for (int j = 0; j < BQ_ITEMS; ++j) {
      const uint32_t i = j * BQ_BLOCK + threadIdx.x;
      const bool row = i < nt;
      for (int qi = q0; qi < q1; ++qi) {
        const BQuery &q = qv[qi];
        bool on = false;
        uint64_t g = 0;
        int64_t v = 0;
        if (row) {
          const int64_t p = tile_get(ts, a.words, a.segs, q.pred_col, seg, r0, i);
          if (p >= q.lo && p <= q.hi) {
            g = group_of(q, tile_get(ts, a.words, a.segs, q.key_col, seg, r0, i));
            v = bq_value(q, tile_get(ts, a.words, a.segs, q.x_col, seg, r0, i),
                         tile_get(ts, a.words, a.segs, q.y_col, seg, r0, i));
            on = true;
          }
        }
        bq_add(q, hot, g, v, on, hsum + (size_t)(qi - q0) * hot, hcnt + (size_t)(qi - q0) * hot);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < (q1 - q0) * hot; i += BQ_BLOCK) {
    const int qi = q0 + i / hot, g = i % hot;
    const unsigned int c = hcnt[i];
    if (c && (uint64_t)g < qv[qi].ngroups) {
      atomicAdd(&qv[qi].acc[g].sum, hsum[i]);
      atomicAdd(&qv[qi].acc[g].cnt, c);
    }
  }
}

__global__ void k_batch_clear(const BQuery *q, int nq) {
  for (int qi = 0; qi < nq; ++qi) {
    const uint64_t n = q[qi].ngroups;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
      q[qi].acc[i].sum = 0;
      q[qi].acc[i].cnt = 0;
    }
  }
}

__global__ void k_batch_flags(const BSlot *acc, uint64_t n, uint32_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    flag[i] = acc[i].cnt != 0;
}
__global__ void k_batch_emit(const uint32_t *flag, const uint64_t *pos, uint64_t n, int64_t gbase,
                             const BSlot *acc, int64_t *okey, int64_t *osum, int64_t *ocnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    if (flag[i]) {
      const uint64_t o = pos[i];
      okey[o] = gbase + (int64_t)i;
      osum[o] = (int64_t)acc[i].sum;
      ocnt[o] = acc[i].cnt;
    }
}

}  // namespace aln

struct aln_batch {
  aln_ctx *ctx = nullptr;
  aln::BArgs args;
  std::vector<aln::BQuery> hq;
  aln::BQuery *dq = nullptr;
  std::vector<void *> allocs;
  char *pool = nullptr;  // caller-owned region the tables are carved from (else hipMalloc)
  size_t pool_size = 0, pool_used = 0;
  uint32_t blocks = 0;
  uint64_t max_groups = 0;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace aln;

namespace {

int batch_free(aln_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  return ALN_OK;
}

int batch_alloc(aln_batch *b, size_t bytes, void **out) {
  if (b->pool) {
    const size_t a = (bytes + 255) & ~(size_t)255;
    if (b->pool_used + a > b->pool_size)
      return set_error(ALN_E_NOMEM, "aln_batch: region of %zu bytes is too small", b->pool_size);
    *out = b->pool + b->pool_used;
    b->pool_used += a;
    return ALN_OK;
  }
  if (hipMalloc(out, bytes ? bytes : 8) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(ALN_E_NOMEM, "aln_batch: allocation of %zu bytes failed", bytes);
  }
  b->allocs.push_back(*out);
  return ALN_OK;
}

int batch_enqueue(aln_batch *b, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_clear, dim3(grid_for(b->max_groups, 256, 8192)), dim3(256), 0, st, b->dq,
                     b->args.nq);
  const size_t lds = (size_t)b->args.qg * b->args.hot * (sizeof(unsigned long long) + sizeof(unsigned int));
  if (b->args.ntiles) hipLaunchKernelGGL(k_batch, dim3(b->blocks), dim3(BQ_BLOCK), lds, st, b->args, b->dq);
  return check_launch("batch_scan");
}

int set_tuning(aln_batch *b, const aln_batch_tuning *tune) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->ctx->device) != hipSuccess || cus <= 0)
    cus = 256;
  const int bpc = tune && tune->blocks_per_cu > 0 ? tune->blocks_per_cu : 2;
  const int hot = tune && tune->hot_groups >= 0 ? std::min(tune->hot_groups, BQ_HOT_MAX) : BQ_HOT_DEFAULT;
  // query groups whose LDS tables fit BQ_HOT_BYTES; every group gets the same number of
  // blocks (the grid is a multiple of the group count), at most one per tile
  const int nq = b->args.nq;
  const int qmax = hot > 0 ? std::max(1, std::min(nq, BQ_HOT_BYTES / (hot * 12))) : nq;
  const int ngq = (nq + qmax - 1) / qmax;
  const int qg = (nq + ngq - 1) / ngq;  // the queries spread evenly over the groups
  uint64_t per = (uint64_t)bpc * (uint64_t)cus / (uint64_t)ngq;
  if (per > b->args.ntiles) per = b->args.ntiles;
  if (per < 1) per = 1;
  b->args.hot = hot;
  b->args.qg = qg;
  b->args.ngq = ngq;
  b->blocks = (uint32_t)(per * (uint64_t)ngq);
  const size_t lds = (size_t)qg * hot * (sizeof(unsigned long long) + sizeof(unsigned int));
  // the kernel's dynamic LDS limit, raised once per process to the largest size asked for
  static std::mutex mu;
  static size_t lds_set = 0;
  std::lock_guard<std::mutex> g(mu);
  if (lds > lds_set) {
    ALN_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_batch), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)BQ_HOT_BYTES));
    lds_set = BQ_HOT_BYTES;
  }
  b->use_graph = tune && tune->use_graph;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return ALN_OK;
}

}  // namespace

namespace aln {

int64_t floor_div(int64_t a, int64_t b) {
  const int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

uint64_t dense_groups(int64_t mn, int64_t mx, int64_t div) {
  if (mx < mn) return 0;
  return (uint64_t)(floor_div(mx, div) - floor_div(mn, div)) + 1;
}

size_t batch_pool_bytes(const uint64_t *ngroups, int nq) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t t = al(sizeof(BQuery) * (size_t)nq);
  for (int i = 0; i < nq; ++i) {
    const size_t g = ngroups[i] ? ngroups[i] : 1;
    t += al(sizeof(BSlot) * g);
  }
  return t;
}

int batch_create(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows, const aln_batch_query *qs,
                 int nq, const aln_batch_tuning *tune, void *pool, size_t pool_bytes, aln_batch **out,
                 const aln_seg *const *hsegs, const int64_t *hmin, const int64_t *hmax) {
  ALN_REQUIRE(c && cols && qs && out && seg_rows > 0, "aln_batch_create: bad arguments");
  ALN_REQUIRE(ncols >= 1 && ncols <= ALN_MAX_COLS, "aln_batch_create: %d columns", ncols);
  ALN_REQUIRE(nq >= 1 && nq <= BQ_MAX, "aln_batch_create: %d queries (1..%d)", nq, BQ_MAX);
  const uint32_t nsegs = cols[0].nsegs;
  for (int i = 0; i < ncols; ++i)
    ALN_REQUIRE(cols[i].nsegs == nsegs, "aln_batch_create: columns differ in segments");
  aln_batch *b = new aln_batch();
  b->ctx = c;
  b->pool = (char *)pool;
  b->pool_size = pool ? pool_bytes : 0;
  auto bail = [&](int rc) {
    batch_free(b);
    delete b;
    return rc;
  };
  // zone maps give each key column's range
  std::vector<std::vector<aln_seg>> hs(ncols);
  if (hsegs)
    for (int col = 0; col < ncols; ++col) hs[col].assign(hsegs[col], hsegs[col] + nsegs);
  auto range_of = [&](int col, int64_t *mn, int64_t *mx) -> int {
    if (hs[col].empty() && nsegs) {
      hs[col].resize(nsegs);
      ALN_TRY(copy_to_host(c, hs[col].data(), cols[col].segs, sizeof(aln_seg) * nsegs));
    }
    if (hmin && hmax) {  // the caller's ranges (a server computes them once per table)
      *mn = hmin[col];
      *mx = hmax[col];
      return ALN_OK;
    }
    *mn = INT64_MAX;
    *mx = INT64_MIN;
    bool open_ended = false;
    uint64_t rows = 0;
    for (auto &s : hs[col])
      if (s.nrows) {
        *mn = s.vmin < *mn ? s.vmin : *mn;
        *mx = s.vmax > *mx ? s.vmax : *mx;
        open_ended = open_ended || (s.bits == 64 && s.vmin == INT64_MIN && s.vmax == INT64_MAX);
        rows += s.nrows;
      }
    // a plain 64-bit column carries no zone map: its words are the int64 values in row order
    if (open_ended) {
      for (uint32_t s = 0; s < nsegs; ++s)
        if (hs[col][s].bits != 64 || hs[col][s].word_off != 2ull * s * seg_rows)
          return set_error(ALN_E_ARG, "aln_batch_create: column %d has no zone map", col);
      ALN_TRY(minmax_i64(c, (const int64_t *)cols[col].words, rows, mn, mx));
    }
    return ALN_OK;
  };
  // every column must split into the same segments (row counts)
  std::vector<int64_t> cmin(ncols), cmax(ncols);
  for (int col = 0; col < ncols; ++col) {  // column 0 first: the others are checked against it
    int rc = range_of(col, &cmin[col], &cmax[col]);
    if (rc) return bail(rc);
    for (uint32_t s = 0; s < nsegs; ++s)
      if (hs[col][s].nrows != hs[0][s].nrows || hs[0][s].nrows > seg_rows)
        return bail(set_error(ALN_E_ARG, "aln_batch_create: column %d segment %u has %u rows", col, s,
                              hs[col][s].nrows));
  }
  b->hq.resize(nq);
  std::vector<int> stage;
  memset(&b->args, 0, sizeof(b->args));
  b->args.pred_col = qs[0].pred_col;
  b->args.key_col = qs[0].key_col;
  b->args.x_col = qs[0].x_col;
  b->args.y_col = qs[0].y_col;
  for (int i = 0; i < nq; ++i) {
    const aln_batch_query &q = qs[i];
    if (!(q.pred_col >= 0 && q.pred_col < ncols && q.key_col >= 0 && q.key_col < ncols && q.x_col >= 0 &&
          q.x_col < ncols && q.y_col >= 0 && q.y_col < ncols))
      return bail(set_error(ALN_E_ARG, "aln_batch_create: query %d column slot", i));
    if (q.key_div < 1) return bail(set_error(ALN_E_ARG, "aln_batch_create: query %d key_div", i));
    const int64_t mn = cmin[q.key_col], mx = cmax[q.key_col];
    BQuery &bq = b->hq[i];
    bq.pred_col = q.pred_col;
    bq.key_col = q.key_col;
    bq.x_col = q.x_col;
    bq.y_col = q.y_col;
    bq.lo = q.lo;
    bq.hi = q.hi;
    bq.div = q.key_div;
    bq.a0 = q.a0;
    bq.a1 = q.a1;
    bq.b0 = q.b0;
    bq.b1 = q.b1;
    bq.gbase = mx < mn ? 0 : floor_div(mn, q.key_div);
    bq.base = (uint64_t)bq.gbase * (uint64_t)q.key_div;
    bq.magic = q.key_div >= 2 ? ~0ull / (uint64_t)q.key_div + 1 : 0;
    bq.ngroups = dense_groups(mn, mx, q.key_div);
    if (bq.ngroups > (1ull << 31))
      return bail(set_error(ALN_E_UNSUPPORTED, "aln_batch_create: query %d has %llu dense groups", i,
                            (unsigned long long)bq.ngroups));
    b->max_groups = std::max(b->max_groups, bq.ngroups);
    if (q.pred_col != b->args.pred_col) b->args.pred_col = -1;
    if (q.key_col != b->args.key_col) b->args.key_col = -1;
    if (q.x_col != b->args.x_col || q.y_col != b->args.y_col) b->args.x_col = b->args.y_col = -1;
    for (int col : {q.pred_col, q.key_col})
      if (std::find(stage.begin(), stage.end(), col) == stage.end()) stage.push_back(col);
    void *p;
    const int r1 = batch_alloc(b, sizeof(BSlot) * (bq.ngroups ? bq.ngroups : 1), &p);
    if (r1) return bail(r1);
    bq.acc = (BSlot *)p;
  }
  // distinct columns: the staged ones first, then the rest in first-use order
  std::vector<int> distinct(stage.begin(), stage.end());
  for (int i = 0; i < nq; ++i)
    for (int col : {qs[i].pred_col, qs[i].key_col, qs[i].x_col, qs[i].y_col})
      if (std::find(distinct.begin(), distinct.end(), col) == distinct.end()) distinct.push_back(col);
  if (distinct.size() <= (size_t)BQ_SLOTS) {
    b->args.nslots = (int)distinct.size();
    for (int k = 0; k < BQ_SLOTS; ++k) b->args.slot_col[k] = k < b->args.nslots ? distinct[k] : distinct[0];
    auto slot = [&](int col) { return (int32_t)(std::find(distinct.begin(), distinct.end(), col) - distinct.begin()); };
    for (int i = 0; i < nq; ++i) {
      b->hq[i].ps = slot(qs[i].pred_col);
      b->hq[i].ks = slot(qs[i].key_col);
      b->hq[i].xs = slot(qs[i].x_col);
      b->hq[i].ys = slot(qs[i].y_col);
    }
  }
  void *p;
  int r2 = batch_alloc(b, sizeof(BQuery) * nq, &p);
  if (r2) return bail(r2);
  b->dq = (BQuery *)p;
  {
    const int rc = copy_to_device(c, b->dq, b->hq.data(), sizeof(BQuery) * nq);
    if (rc) return bail(rc);
  }
  for (int i = 0; i < ncols; ++i) {
    b->args.words[i] = cols[i].words;
    b->args.segs[i] = cols[i].segs;
  }
  b->args.q = b->dq;
  b->args.nq = nq;
  b->args.shared_v = b->args.nslots > 0;
  for (int i = 1; i < nq; ++i)
    if (qs[i].x_col != qs[0].x_col || qs[i].y_col != qs[0].y_col || qs[i].a0 != qs[0].a0 || qs[i].a1 != qs[0].a1 ||
        qs[i].b0 != qs[0].b0 || qs[i].b1 != qs[0].b1)
      b->args.shared_v = 0;
  b->args.nstage = (int)std::min<size_t>(stage.size(), 2);
  for (int i = 0; i < b->args.nstage; ++i) b->args.stage_cols[i] = stage[i];
  b->args.seg_rows = seg_rows;
  b->args.tps = (seg_rows + STG_TILE - 1) / STG_TILE;
  b->args.ntiles = (uint64_t)nsegs * b->args.tps;
  set_tuning(b, tune);
  *out = b;
  return ALN_OK;
}

}  // namespace aln

extern "C" {

int aln_batch_create(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows,
                     const aln_batch_query *qs, int nq, const aln_batch_tuning *tune, aln_batch **out) {
  return batch_create(c, cols, ncols, seg_rows, qs, nq, tune, nullptr, 0, out, nullptr, nullptr, nullptr);
}

int aln_batch_set_tuning(aln_batch *b, const aln_batch_tuning *tune) {
  ALN_REQUIRE(b, "aln_batch_set_tuning: NULL batch");
  return set_tuning(b, tune);
}

// Enqueues one run (clear tables + scan) on the context stream; no synchronisation.
int aln_batch_run(aln_batch *b) {
  ALN_REQUIRE(b, "aln_batch_run: NULL batch");
  aln_ctx *c = b->ctx;
  if (!b->use_graph) {
    ProfScope p(c, "batch_scan");
    return batch_enqueue(b, c->stream);
  }
  if (!b->exec) {
    // Capture the launch sequence once (fill kernels only, no memset nodes) on a private
    // stream: the context stream may be the legacy default stream, which cannot capture.
    hipStream_t cs;
    ALN_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = ALN_OK;
    hipGraph_t g = nullptr;
    if (e == hipSuccess) {
      rc = batch_enqueue(b, cs);
      e = hipStreamEndCapture(cs, &g);
    }
    (void)hipStreamDestroy(cs);
    if (rc != ALN_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_error(ALN_E_HIP, "aln_batch_run: capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    ALN_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  ProfScope p(c, "batch_scan");
  ALN_HIP(hipGraphLaunch(b->exec, c->stream));
  return ALN_OK;
}

}  // extern "C"

// Groups of query q with at least one row, in key order (synchronises).  With host
// pointers the groups are emitted into workspace and copied out.
static int batch_results(aln_batch *b, int q, int64_t *keys, int64_t *sums, int64_t *counts, uint64_t cap,
                         uint64_t *h_ngroups, bool host, std::vector<int64_t> *vec = nullptr) {
  ALN_REQUIRE(b && h_ngroups && q >= 0 && q < b->args.nq, "aln_batch_results: bad arguments");
  aln_ctx *c = b->ctx;
  ws_reset(c);
  const BQuery &bq = b->hq[q];
  const uint64_t n = bq.ngroups;
  *h_ngroups = 0;
  if (!n) return ALN_OK;
  uint32_t *flag;
  uint64_t *pos;
  ALN_TRY(ws(c, n, &flag));
  ALN_TRY(ws(c, n, &pos));
  hipLaunchKernelGGL(k_batch_flags, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, bq.acc, n, flag);
  ALN_TRY(check_launch("batch_flags"));
  uint64_t g = 0;
  ALN_TRY(exclusive_scan_u32(c, flag, n, pos, &g));
  *h_ngroups = g;
  if (vec) {  // host vectors sized to the result
    for (int i = 0; i < 3; ++i) vec[i].resize(g);
    keys = vec[0].data();
    sums = vec[1].data();
    counts = vec[2].data();
    cap = g;
  }
  if (g > cap)
    return set_error(ALN_E_CAPACITY, "aln_batch_results: %llu groups, capacity %llu", (unsigned long long)g,
                     (unsigned long long)cap);
  if (!g) return ALN_OK;
  ALN_REQUIRE(keys && sums && counts, "aln_batch_results: no output buffers");
  int64_t *dk = keys, *ds = sums, *dc = counts;
  if (host) {
    ALN_TRY(ws(c, g, &dk));
    ALN_TRY(ws(c, g, &ds));
    ALN_TRY(ws(c, g, &dc));
  }
  hipLaunchKernelGGL(k_batch_emit, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, flag, pos, n, bq.gbase,
                     bq.acc, dk, ds, dc);
  ALN_TRY(check_launch("batch_emit"));
  if (host) {
    ALN_TRY(copy_to_host(c, keys, dk, 8 * g));
    ALN_TRY(copy_to_host(c, sums, ds, 8 * g));
    ALN_TRY(copy_to_host(c, counts, dc, 8 * g));
  }
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

namespace aln {
int batch_fetch(aln_batch *b, int q, std::vector<int64_t> *keys_sums_counts) {
  uint64_t g = 0;
  return batch_results(b, q, nullptr, nullptr, nullptr, 0, &g, true, keys_sums_counts);
}
}  // namespace aln

extern "C" {

int aln_batch_results(aln_batch *b, int q, int64_t *d_keys, int64_t *d_sums, int64_t *d_counts,
                      uint64_t cap, uint64_t *h_ngroups) {
  return batch_results(b, q, d_keys, d_sums, d_counts, cap, h_ngroups, false);
}

int aln_batch_results_host(aln_batch *b, int q, int64_t *h_keys, int64_t *h_sums, int64_t *h_counts,
                           uint64_t cap, uint64_t *h_ngroups) {
  return batch_results(b, q, h_keys, h_sums, h_counts, cap, h_ngroups, true);
}

void aln_batch_destroy(aln_batch *b) {
  if (!b) return;
  if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
  delete b;
}

// Times each candidate shape (grid size x hot-set size) on the batch and keeps the fastest.
int aln_batch_autotune(aln_batch *b, int reps, aln_batch_tuning *best, double *h_ms) {
  ALN_REQUIRE(b && best, "aln_batch_autotune: bad arguments");
  aln_ctx *c = b->ctx;
  if (reps < 1) reps = 1;
  const int bpcs[] = {2, 4};
  const int hots[] = {64, 128, 256, 512, 1024};
  hipEvent_t e0, e1;
  ALN_HIP(hipEventCreate(&e0));
  ALN_HIP(hipEventCreate(&e1));
  double best_ms = 1e30;
  aln_batch_tuning bt;
  memset(&bt, 0, sizeof(bt));
  for (int bpc : bpcs)
    for (int hot : hots) {
      aln_batch_tuning t;
      memset(&t, 0, sizeof(t));
      t.blocks_per_cu = bpc;
      t.hot_groups = hot;
      t.use_graph = 0;
      set_tuning(b, &t);
      ALN_TRY(batch_enqueue(b, c->stream));  // warm
      ALN_HIP(hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r) ALN_TRY(batch_enqueue(b, c->stream));
      ALN_HIP(hipEventRecord(e1, c->stream));
      ALN_HIP(hipEventSynchronize(e1));
      float ms = 0;
      ALN_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best_ms) {
        best_ms = ms / reps;
        bt = t;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  bt.use_graph = best->use_graph;
  *best = bt;
  set_tuning(b, best);
  if (h_ms) *h_ms = best_ms;
  return ALN_OK;
}

}  // extern "C"

// Fused star pipeline: filter -> join -> join -> group-by -> top-k, the TPC-H q3 shape
// the reference runs as FILTER x3, a two-way JOIN, a GROUP BY SELECT and ORDER ... LIMIT
// (readme.md:48-80; emit_filter operators.cu:2306-2408, emit_multijoin 940-1582,
// emit_select 1801-2084, emit_order 1641-1798, Store(limit) cm.cu:1493-1498).
//
// The group key (m_key, m_g1, m_g2) contains the mid table's unique join key, so every
// group is one passing mid row, and the fact side aggregates straight into per-mid-row
// accumulators found through the join's hash table.  Kernels:
//   D  dim scan     : decode + range predicate; passing keys set bits in a bitmap over the
//                     key range (a perfect hash of the dense dimension key).
//   M  mid scan     : decode + ranges + bitmap probe of m_fk; passing rows are compacted
//                     (key, g1, g2) with one atomic per tile.
//   B  build        : open-addressing table, 8-byte slots (group index << 32 | key code),
//                     MurmurHash64A slot hash, linear probing, capacity ~1.5x entries.
//   F  fact scan    : decode + ranges + probe + SUM((a0+a1*x)*(b0+b1*y)), COUNT(*).  Fact
//                     rows clustered on the join key hit the same group in adjacent lanes;
//                     a segmented wave scan leaves one atomic pair per run.
//   T  top-k        : bitonic sort of 1024-candidate tiles, keep `limit` per tile, repeat.
// Zone maps: a segment whose min/max fails a range is skipped ('N'); one that lies inside
// every range skips predicate evaluation ('A').
#include <string.h>

#include <vector>

#include "aln_internal.h"
#include "stage.h"

namespace aln {

constexpr int ST_BLOCK = 256;
constexpr int ST_ITEMS = 16;
constexpr int ST_TILE = ST_BLOCK * ST_ITEMS;
constexpr uint64_t ST_EMPTY = ~0ull;
constexpr uint32_t ST_SEED = 0x1b873593u;

struct ScanDesc {
  const uint32_t *words[ALN_MAX_COLS];
  const aln_seg *segs[ALN_MAX_COLS];
  int32_t rcol[ALN_MAX_RANGES];
  int64_t rlo[ALN_MAX_RANGES], rhi[ALN_MAX_RANGES];
  int nr;
  uint32_t seg_rows, tps;  // tiles per segment
  const char *modes;       // device, or null
};

__device__ __forceinline__ int64_t sd_decode(const ScanDesc &d, int c, uint32_t seg,
                                             uint32_t local) {
  return decode_row(d.words[c], d.segs[c][seg], local);
}

// Zone-map mode of a segment: 'N' skip, 'A' all rows pass the ranges, 'R' evaluate.
__device__ __forceinline__ char sd_mode(const ScanDesc &d, uint32_t seg) {
  if (d.modes) return d.modes[seg];
  char m = 'A';
  for (int k = 0; k < d.nr; ++k) {
    const aln_seg &s = d.segs[d.rcol[k]][seg];
    if (s.vmax < d.rlo[k] || s.vmin > d.rhi[k]) return 'N';
    if (!(s.vmin >= d.rlo[k] && s.vmax <= d.rhi[k])) m = 'R';
  }
  return m;
}

__device__ __forceinline__ bool sd_pass(const ScanDesc &d, char mode, uint32_t seg,
                                        uint32_t local) {
  if (mode == 'A') return true;
  bool p = true;
  for (int k = 0; k < d.nr; ++k) {
    int64_t v = sd_decode(d, d.rcol[k], seg, local);
    p = p && v >= d.rlo[k] && v <= d.rhi[k];
  }
  return p;
}

// Block prologue of every scan kernel: stage the range columns (when the zone map says the
// tile needs evaluating) and `key` into LDS.
__device__ __forceinline__ void star_stage(TileStage &ts, int *want, const ScanDesc &d, char mode,
                                           int key, uint32_t seg, uint32_t r0, uint32_t nt) {
  __shared__ int nwant;
  if (threadIdx.x == 0) {
    int n = 0;
    if (mode == 'R')
      for (int k = 0; k < d.nr; ++k) want[n++] = d.rcol[k];
    if (key >= 0) want[n++] = key;
    nwant = n;
  }
  __syncthreads();
  stage_tile(ts, d.words, d.segs, want, nwant, seg, r0, nt);
  __syncthreads();
}

__device__ __forceinline__ bool tile_pass(const ScanDesc &d, const TileStage &ts, char mode,
                                          uint32_t seg, uint32_t r0, uint32_t i) {
  if (mode == 'A') return true;
  bool p = true;
  for (int k = 0; k < d.nr; ++k) {
    int64_t v = tile_get(ts, d.words, d.segs, d.rcol[k], seg, r0, i);
    p = p && v >= d.rlo[k] && v <= d.rhi[k];
  }
  return p;
}

// ---- D: dimension scan into a key bitmap --------------------------------------------------
// Per-tile counts go to arrays indexed by block (summed afterwards): a counter shared by
// every block would serialise its atomics across the XCDs.
__device__ __forceinline__ void block_count(uint32_t v, uint32_t *out) {
  __shared__ uint32_t ws_[ST_BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if (__lane_id() == 0) ws_[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < ST_BLOCK / 64; ++w) t += ws_[w];
    out[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(ST_BLOCK) k_star_dim(ScanDesc d, int kcol, int64_t kmin,
                                                       uint64_t range, uint32_t *bitmap,
                                                       uint32_t *npass, uint32_t *dup) {
  __shared__ TileStage ts;
  __shared__ int want[ALN_MAX_RANGES + 1];
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const uint32_t nrows = d.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const char mode = sd_mode(d, seg);
  if (r0 >= nrows || mode == 'N') {
    if (threadIdx.x == 0) npass[blockIdx.x] = 0;
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  star_stage(ts, want, d, mode, kcol, seg, r0, nt);
  uint32_t cnt = 0;
  for (int j = 0; j < ST_ITEMS; ++j) {
    uint32_t i = j * ST_BLOCK + threadIdx.x;
    if (i < nt && tile_pass(d, ts, mode, seg, r0, i)) {
      uint64_t b = (uint64_t)tile_get(ts, d.words, d.segs, kcol, seg, r0, i) - (uint64_t)kmin;
      if (b < range) {
        uint32_t bit = 1u << (b & 31);
        uint32_t old = atomicOr(&bitmap[b >> 5], bit);
        if (old & bit) atomicExch(dup, 1u);
        ++cnt;
      }
    }
  }
  block_count(cnt, npass);
}

// ---- M: mid scan, probe the dimension bitmap, compact passing rows ------------------------
__global__ void __launch_bounds__(ST_BLOCK) k_star_mid(ScanDesc d, int kcol, int fkcol, int g1col,
                                                       int g2col, const uint32_t *__restrict__ bitmap,
                                                       int64_t dkmin, uint64_t drange,
                                                       uint32_t *__restrict__ counts,
                                                       uint32_t *__restrict__ probed,
                                                       int64_t *__restrict__ okey,
                                                       int64_t *__restrict__ og1,
                                                       int64_t *__restrict__ og2,
                                                       uint32_t *__restrict__ mbits, int64_t mkmin,
                                                       uint32_t *dup) {
  // passing rows of tile b are written to entries [b * ST_TILE, b * ST_TILE + counts[b])
  __shared__ uint32_t wc[ST_BLOCK / 64];
  __shared__ TileStage ts;
  __shared__ int want[ALN_MAX_RANGES + 1];
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const uint32_t nrows = d.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const char mode = sd_mode(d, seg);
  if (r0 >= nrows || mode == 'N') {
    if (threadIdx.x == 0) {
      counts[blockIdx.x] = 0;
      probed[blockIdx.x] = 0;
    }
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  star_stage(ts, want, d, mode, fkcol, seg, r0, nt);
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint32_t passmask = 0;
  uint32_t mine = 0, nprobe = 0;
  for (int j = 0; j < ST_ITEMS; ++j) {
    uint32_t i = j * ST_BLOCK + threadIdx.x;
    bool p = i < nt && tile_pass(d, ts, mode, seg, r0, i);
    if (p) {
      ++nprobe;
      uint64_t b = (uint64_t)tile_get(ts, d.words, d.segs, fkcol, seg, r0, i) - (uint64_t)dkmin;
      p = b < drange && ((bitmap[b >> 5] >> (b & 31)) & 1);
    }
    if (p && mbits) {
      uint64_t code = (uint64_t)tile_get(ts, d.words, d.segs, kcol, seg, r0, i) - (uint64_t)mkmin;
      uint32_t bit = 1u << (code & 31);
      if (atomicOr(&mbits[code >> 5], bit) & bit) atomicExch(dup, 1u);
    }
    passmask |= (uint32_t)p << j;
    mine += p;
  }
  block_count(nprobe, probed);
  uint32_t wsum = mine;
  for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o, 64);
  if (lane == 0) wc[wid] = wsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < ST_BLOCK / 64; ++w) t += wc[w];
    counts[blockIdx.x] = t;
  }
  uint64_t pos = (uint64_t)blockIdx.x * ST_TILE;
  for (int w = 0; w < (int)wid; ++w) pos += wc[w];
  // order inside the tile: (item, lane) within the wave; fine, output order is free
  uint32_t x = mine;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  pos += x - mine;
  for (int j = 0; j < ST_ITEMS; ++j) {
    if (!((passmask >> j) & 1)) continue;
    uint32_t i = j * ST_BLOCK + threadIdx.x;
    okey[pos] = tile_get(ts, d.words, d.segs, kcol, seg, r0, i);
    og1[pos] = g1col >= 0 ? tile_get(ts, d.words, d.segs, g1col, seg, r0, i) : 0;
    og2[pos] = g2col >= 0 ? tile_get(ts, d.words, d.segs, g2col, seg, r0, i) : 0;
    ++pos;
  }
}

__device__ __forceinline__ uint64_t st_slot(uint32_t code, uint64_t cap) {
  return __umul64hi(murmur64a_u64(code, ST_SEED), cap);
}

// ---- B: build the join table over the compacted mid rows ----------------------------------
// One block per mid tile region: entry i of tile b becomes group offs[b] + i; its key and
// group columns move to dense arrays and the key goes into the table.
__global__ void __launch_bounds__(256) k_star_build(const int64_t *__restrict__ rkey,
                                                    const int64_t *__restrict__ rg1,
                                                    const int64_t *__restrict__ rg2,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint64_t *__restrict__ offs,
                                                    int64_t kmin, uint64_t *table, uint64_t cap,
                                                    int64_t *__restrict__ okey,
                                                    int64_t *__restrict__ og1,
                                                    int64_t *__restrict__ og2, uint32_t *dup) {
  const uint32_t n = counts[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const uint64_t src = (uint64_t)blockIdx.x * ST_TILE + i, j = offs[blockIdx.x] + i;
    okey[j] = rkey[src];
    og1[j] = rg1[src];
    og2[j] = rg2[src];
    uint32_t code = (uint32_t)((uint64_t)rkey[src] - (uint64_t)kmin);
    uint64_t entry = ((uint64_t)j << 32) | code;
    uint64_t h = st_slot(code, cap);
    for (uint64_t probes = 0; probes < cap; ++probes) {
      unsigned long long old = atomicCAS((unsigned long long *)&table[h], (unsigned long long)ST_EMPTY,
                                         (unsigned long long)entry);
      if (old == ST_EMPTY) break;
      if ((uint32_t)old == code) {
        atomicExch(dup, 1u);
        break;
      }
      h = h + 1 == cap ? 0 : h + 1;
    }
  }
}

__device__ __forceinline__ uint32_t st_find(const uint64_t *__restrict__ table, uint64_t cap,
                                            uint32_t code) {
  uint64_t h = st_slot(code, cap);
  for (uint64_t probes = 0; probes < cap; ++probes) {
    uint64_t e = table[h];
    if (e == ST_EMPTY) return 0xFFFFFFFFu;
    if ((uint32_t)e == code) return (uint32_t)(e >> 32);
    h = h + 1 == cap ? 0 : h + 1;
  }
  return 0xFFFFFFFFu;
}

// ---- F: fact scan, probe, aggregate -------------------------------------------------------
__global__ void __launch_bounds__(ST_BLOCK) k_star_fact(ScanDesc d, int fkcol, int xcol, int ycol,
                                                        int64_t a0, int64_t a1, int64_t b0,
                                                        int64_t b1, const uint64_t *__restrict__ table,
                                                        uint64_t cap, int64_t mkmin, uint64_t mrange,
                                                        unsigned long long *__restrict__ sum,
                                                        unsigned int *__restrict__ cnt,
                                                        uint32_t *__restrict__ matched,
                                                        uint32_t *__restrict__ probed) {
  __shared__ TileStage ts;
  __shared__ int want[ALN_MAX_RANGES + 1];
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const uint32_t nrows = d.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const char mode = sd_mode(d, seg);
  if (r0 >= nrows || mode == 'N') {
    if (threadIdx.x == 0) {
      matched[blockIdx.x] = 0;
      probed[blockIdx.x] = 0;
    }
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  star_stage(ts, want, d, mode, fkcol, seg, r0, nt);
  const unsigned lane = __lane_id();
  uint32_t nmatch = 0, nprobe = 0;
  for (int j = 0; j < ST_ITEMS; ++j) {
    uint32_t i = j * ST_BLOCK + threadIdx.x;
    uint32_t local = r0 + i;
    uint32_t g = 0xFFFFFFFFu;
    int64_t v = 0;
    if (i < nt && tile_pass(d, ts, mode, seg, r0, i)) {
      ++nprobe;
      uint64_t code = (uint64_t)tile_get(ts, d.words, d.segs, fkcol, seg, r0, i) - (uint64_t)mkmin;
      if (code < mrange) g = st_find(table, cap, (uint32_t)code);
      if (g != 0xFFFFFFFFu) {
        int64_t x = sd_decode(d, xcol, seg, local), y = sd_decode(d, ycol, seg, local);
        v = wmul((int64_t)((uint64_t)a0 + (uint64_t)wmul(a1, x)),
                 (int64_t)((uint64_t)b0 + (uint64_t)wmul(b1, y)));
      }
    }
    // segmented reduction over runs of equal group index in adjacent lanes (skipped when no
    // lane of the wave joined, the common case)
    if (__ballot(g != 0xFFFFFFFFu) == 0) continue;
    uint32_t prev = __shfl_up(g, 1, 64);
    bool f = lane == 0 || prev != g;
    uint32_t c = g != 0xFFFFFFFFu;
    for (int o = 1; o < 64; o <<= 1) {
      int64_t vu = __shfl_up(v, o, 64);
      uint32_t cu = __shfl_up(c, o, 64);
      bool fu = __shfl_up((int)f, o, 64) != 0;
      if (lane >= (unsigned)o && !f) {
        v = (int64_t)((uint64_t)v + (uint64_t)vu);
        c += cu;
      }
      if (lane >= (unsigned)o) f = f || fu;
    }
    uint32_t next = __shfl_down(g, 1, 64);
    bool tail = g != 0xFFFFFFFFu && (lane == 63 || next != g);
    if (tail) {
      atomicAdd(&sum[g], (unsigned long long)v);
      atomicAdd(&cnt[g], c);
      nmatch += c;
    }
  }
  block_count(nmatch, matched);
  block_count(nprobe, probed);
}

// ---- bitmap + rank join (dense key ranges) -----------------------------------------------
// The passing mid keys set bits of a bitmap over [mkmin, mkmax]; a directory of set-bit
// counts per 512-bit block (one 64-byte line) makes rank(key) - the key's position among
// the passing keys - cost one directory read plus one line of the bitmap.  The group index
// is that rank, so groups come out in key order, and probes from a tile of fact rows that
// are clustered on the key touch a few contiguous lines instead of random slots.
constexpr int BM_BLOCK_WORDS = 16;  // 512 bits

__global__ void __launch_bounds__(256) k_bm_block_counts(const uint32_t *__restrict__ bits,
                                                        uint64_t nwords, uint64_t nblk,
                                                        uint32_t *__restrict__ cnt) {
  for (uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x; b < nblk;
       b += (uint64_t)gridDim.x * 256) {
    uint32_t c = 0;
    uint64_t w0 = b * BM_BLOCK_WORDS;
    for (int k = 0; k < BM_BLOCK_WORDS; ++k)
      if (w0 + k < nwords) c += __popc(bits[w0 + k]);
    cnt[b] = c;
  }
}

__device__ __forceinline__ uint64_t bm_rank(const uint32_t *__restrict__ bits,
                                            const uint64_t *__restrict__ rank, uint64_t code) {
  uint64_t w = code >> 5, blk = code >> 9;
  uint64_t r = rank[blk];
  for (uint64_t k = blk * BM_BLOCK_WORDS; k < w; ++k) r += __popc(bits[k]);
  return r + __popc(bits[w] & ((1u << (code & 31)) - 1u));
}

// One block per mid tile region (see k_star_mid); entry i moves to position rank(key).
__global__ void __launch_bounds__(256) k_bm_scatter(const int64_t *__restrict__ rkey,
                                                   const int64_t *__restrict__ rg1,
                                                   const int64_t *__restrict__ rg2,
                                                   const uint32_t *__restrict__ counts,
                                                   const uint32_t *__restrict__ bits,
                                                   const uint64_t *__restrict__ rank, int64_t kmin,
                                                   int64_t *__restrict__ skey,
                                                   int64_t *__restrict__ sg1,
                                                   int64_t *__restrict__ sg2) {
  const uint32_t n = counts[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const uint64_t src = (uint64_t)blockIdx.x * ST_TILE + i;
    const uint64_t j = bm_rank(bits, rank, (uint64_t)rkey[src] - (uint64_t)kmin);
    skey[j] = rkey[src];
    sg1[j] = rg1[src];
    sg2[j] = rg2[src];
  }
}

__global__ void __launch_bounds__(ST_BLOCK) k_star_fact_bm(ScanDesc d, int fkcol, int xcol, int ycol,
                                                           int64_t a0, int64_t a1, int64_t b0,
                                                           int64_t b1, const uint32_t *__restrict__ bits,
                                                           const uint64_t *__restrict__ rank,
                                                           int64_t mkmin, uint64_t mrange,
                                                           unsigned long long *__restrict__ sum,
                                                           unsigned int *__restrict__ cnt,
                                                           uint32_t *__restrict__ matched,
                                                           uint32_t *__restrict__ probed) {
  __shared__ TileStage ts;
  __shared__ int want[ALN_MAX_RANGES + 1];
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const uint32_t nrows = d.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const char mode = sd_mode(d, seg);
  if (r0 >= nrows || mode == 'N') {
    if (threadIdx.x == 0) {
      matched[blockIdx.x] = 0;
      probed[blockIdx.x] = 0;
    }
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  star_stage(ts, want, d, mode, fkcol, seg, r0, nt);
  const unsigned lane = __lane_id();
  uint32_t nmatch = 0, nprobe = 0;
  for (int j = 0; j < ST_ITEMS; ++j) {
    uint32_t i = j * ST_BLOCK + threadIdx.x;
    uint32_t local = r0 + i;
    uint32_t g = 0xFFFFFFFFu;
    int64_t v = 0;
    if (i < nt && tile_pass(d, ts, mode, seg, r0, i)) {
      ++nprobe;
      uint64_t code = (uint64_t)tile_get(ts, d.words, d.segs, fkcol, seg, r0, i) - (uint64_t)mkmin;
      if (code < mrange && ((bits[code >> 5] >> (code & 31)) & 1)) {
        g = (uint32_t)bm_rank(bits, rank, code);
        int64_t x = sd_decode(d, xcol, seg, local), y = sd_decode(d, ycol, seg, local);
        v = wmul((int64_t)((uint64_t)a0 + (uint64_t)wmul(a1, x)),
                 (int64_t)((uint64_t)b0 + (uint64_t)wmul(b1, y)));
      }
    }
    if (__ballot(g != 0xFFFFFFFFu) == 0) continue;
    uint32_t prev = __shfl_up(g, 1, 64);
    bool f = lane == 0 || prev != g;
    uint32_t c = g != 0xFFFFFFFFu;
    for (int o = 1; o < 64; o <<= 1) {
      int64_t vu = __shfl_up(v, o, 64);
      uint32_t cu = __shfl_up(c, o, 64);
      bool fu = __shfl_up((int)f, o, 64) != 0;
      if (lane >= (unsigned)o && !f) {
        v = (int64_t)((uint64_t)v + (uint64_t)vu);
        c += cu;
      }
      if (lane >= (unsigned)o) f = f || fu;
    }
    uint32_t next = __shfl_down(g, 1, 64);
    bool tail = g != 0xFFFFFFFFu && (lane == 63 || next != g);
    if (tail) {
      atomicAdd(&sum[g], (unsigned long long)v);
      atomicAdd(&cnt[g], c);
      nmatch += c;
    }
  }
  block_count(nmatch, matched);
  block_count(nprobe, probed);
}

// ---- T: top-k over groups (sum DESC, g1 ASC, key ASC, g2 ASC) -----------------------------
struct Cand {
  int64_t sum, g1, key, g2;
  uint32_t j, valid;
  uint32_t count, pad;
};

__device__ __forceinline__ bool cand_better(const Cand &a, const Cand &b) {
  if (a.valid != b.valid) return a.valid > b.valid;
  if (a.sum != b.sum) return a.sum > b.sum;
  if (a.g1 != b.g1) return a.g1 < b.g1;
  if (a.key != b.key) return a.key < b.key;
  if (a.g2 != b.g2) return a.g2 < b.g2;
  return a.j < b.j;
}

constexpr int TK_TILE2 = 1024;

__global__ void __launch_bounds__(256) k_star_topk_tile(const Cand *__restrict__ in, uint64_t n,
                                                        const unsigned long long *__restrict__ sum,
                                                        const unsigned int *__restrict__ cnt,
                                                        const int64_t *__restrict__ okey,
                                                        const int64_t *__restrict__ og1,
                                                        const int64_t *__restrict__ og2,
                                                        uint32_t k, Cand *__restrict__ out) {
  __shared__ Cand s[TK_TILE2];
  uint64_t base = (uint64_t)blockIdx.x * TK_TILE2;
  for (int t = threadIdx.x; t < TK_TILE2; t += 256) {
    uint64_t i = base + t;
    Cand c;
    if (in) {
      if (i < n) c = in[i];
      else c.valid = 0;
    } else if (i < n && cnt[i] > 0) {
      c.sum = (int64_t)sum[i];
      c.g1 = og1[i];
      c.key = okey[i];
      c.g2 = og2[i];
      c.j = (uint32_t)i;
      c.valid = 1;
      c.count = cnt[i];
    } else {
      c.valid = 0;
    }
    if (!c.valid) {
      c.sum = c.g1 = c.key = c.g2 = 0;
      c.j = 0xFFFFFFFFu;
      c.count = 0;
    }
    s[t] = c;
  }
  __syncthreads();
  for (int size = 2; size <= TK_TILE2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < TK_TILE2 / 2; t += 256) {
        int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        bool up = (lo & size) == 0;
        Cand a = s[lo], b = s[hi];
        bool sw = up ? cand_better(b, a) : cand_better(a, b);
        if (sw) {
          s[lo] = b;
          s[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t t = threadIdx.x; t < k; t += 256) out[(uint64_t)blockIdx.x * k + t] = s[t];
}

__global__ void k_star_group_flags(const unsigned int *cnt, uint64_t n, uint32_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    flag[i] = cnt[i] > 0;
}

__global__ void k_star_group_codes(const uint32_t *flag, const uint64_t *pos, uint64_t n,
                                   const int64_t *okey, int64_t kmin, uint64_t *code,
                                   uint32_t *idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    if (flag[i]) {
      code[pos[i]] = (uint64_t)okey[i] - (uint64_t)kmin;
      idx[pos[i]] = (uint32_t)i;
    }
}

__global__ void k_star_group_emit(const uint32_t *idx, uint64_t g, const int64_t *okey,
                                  const int64_t *og1, const int64_t *og2,
                                  const unsigned long long *sum, const unsigned int *cnt,
                                  aln_star_row *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    uint32_t j = idx[i];
    aln_star_row r;
    r.key = okey[j];
    r.g1 = og1[j];
    r.g2 = og2[j];
    r.sum = (int64_t)sum[j];
    r.count = cnt[j];
    out[i] = r;
  }
}

// Reads the segment descriptors of one column to the host and returns min/max over them.
static int col_range(aln_ctx *c, const aln_col &col, int64_t *mn, int64_t *mx) {
  std::vector<aln_seg> h(col.nsegs);
  ALN_TRY(copy_to_host(c, h.data(), col.segs, sizeof(aln_seg) * col.nsegs));
  *mn = INT64_MAX;
  *mx = INT64_MIN;
  for (auto &s : h) {
    if (s.nrows == 0) continue;
    *mn = s.vmin < *mn ? s.vmin : *mn;
    *mx = s.vmax > *mx ? s.vmax : *mx;
  }
  return ALN_OK;
}

static int make_scan(aln_ctx *c, const aln_table_scan &t, int need_col, ScanDesc *d,
                     uint64_t *nblocks) {
  ALN_REQUIRE(t.cols && t.ncols >= 1 && t.ncols <= ALN_MAX_COLS && t.seg_rows > 0,
              "aln_star_agg: bad table scan");
  ALN_REQUIRE(t.nranges >= 0 && t.nranges <= ALN_MAX_RANGES, "aln_star_agg: %d ranges", t.nranges);
  ALN_REQUIRE(need_col < t.ncols, "aln_star_agg: column slot %d of %d", need_col, t.ncols);
  memset(d, 0, sizeof(*d));
  for (int i = 0; i < t.ncols; ++i) {
    ALN_REQUIRE(t.cols[i].nsegs == t.cols[0].nsegs, "aln_star_agg: columns differ in segments");
    d->words[i] = t.cols[i].words;
    d->segs[i] = t.cols[i].segs;
  }
  for (int k = 0; k < t.nranges; ++k) {
    ALN_REQUIRE(t.ranges[k].col >= 0 && t.ranges[k].col < t.ncols, "aln_star_agg: range column");
    d->rcol[k] = t.ranges[k].col;
    d->rlo[k] = t.ranges[k].lo;
    d->rhi[k] = t.ranges[k].hi;
  }
  d->nr = t.nranges;
  d->seg_rows = t.seg_rows;
  d->tps = (t.seg_rows + ST_TILE - 1) / ST_TILE;
  if (t.seg_modes) {
    char *dm;
    ALN_TRY(ws(c, t.cols[0].nsegs, &dm));
    ALN_HIP(hipMemcpyAsync(dm, t.seg_modes, t.cols[0].nsegs, hipMemcpyHostToDevice, c->stream));
    d->modes = dm;
  }
  *nblocks = (uint64_t)t.cols[0].nsegs * d->tps;
  ALN_REQUIRE(*nblocks < (1ull << 31), "aln_star_agg: grid too large");
  return ALN_OK;
}

struct PhaseTimer {
  aln_ctx *c;
  hipEvent_t e[5];
  bool on;
  explicit PhaseTimer(aln_ctx *c_) : c(c_), on(c_->profiling) {
    for (auto &x : e) x = nullptr;
    if (on)
      for (auto &x : e)
        if (hipEventCreate(&x) != hipSuccess) on = false;
  }
  void mark(int i) {
    if (on) (void)hipEventRecord(e[i], c->stream);
  }
  double ms(int a, int b) {
    float t = 0;
    if (!on || hipEventElapsedTime(&t, e[a], e[b]) != hipSuccess) return 0.0;
    return t;
  }
  ~PhaseTimer() {
    for (auto &x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

}  // namespace aln

using namespace aln;

extern "C" int aln_star_agg(aln_ctx *c, const aln_star_args *a, aln_star_row *h_top,
                            uint32_t *h_ntop, aln_star_row *d_groups, uint64_t groups_cap,
                            aln_star_stats *h_stats) {
  ALN_REQUIRE(c && a && h_ntop, "aln_star_agg: bad arguments");
  ALN_REQUIRE(a->limit <= TK_TILE2 / 2, "aln_star_agg: limit %u above %d", a->limit, TK_TILE2 / 2);
  ALN_REQUIRE(a->limit == 0 || h_top, "aln_star_agg: no top-k buffer");
  ws_reset(c);
  *h_ntop = 0;
  aln_star_stats st;
  memset(&st, 0, sizeof(st));
  PhaseTimer tm(c);

  ScanDesc dd, md, fd;
  uint64_t dblocks, mblocks, fblocks;
  ALN_TRY(make_scan(c, a->dim, a->d_key, &dd, &dblocks));
  ALN_TRY(make_scan(c, a->mid, a->m_key > a->m_fk ? a->m_key : a->m_fk, &md, &mblocks));
  ALN_REQUIRE(a->m_g1 < a->mid.ncols && a->m_g2 < a->mid.ncols, "aln_star_agg: group columns");
  ALN_TRY(make_scan(c, a->fact, a->f_fk, &fd, &fblocks));
  ALN_REQUIRE(a->f_x >= 0 && a->f_x < a->fact.ncols && a->f_y >= 0 && a->f_y < a->fact.ncols,
              "aln_star_agg: aggregate columns");

  // key ranges from the zone maps
  int64_t dkmin, dkmax, mkmin, mkmax;
  ALN_TRY(col_range(c, a->dim.cols[a->d_key], &dkmin, &dkmax));
  ALN_TRY(col_range(c, a->mid.cols[a->m_key], &mkmin, &mkmax));
  if (a->dim.cols[0].nsegs == 0 || a->mid.cols[0].nsegs == 0 || dkmax < dkmin || mkmax < mkmin) {
    if (h_stats) *h_stats = st;
    return ALN_OK;
  }
  uint64_t drange = (uint64_t)dkmax - (uint64_t)dkmin + 1;
  uint64_t mrange = (uint64_t)mkmax - (uint64_t)mkmin + 1;
  if (drange == 0 || drange > (1ull << 34))
    return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: dimension key range %llu too wide for a bitmap",
                     (unsigned long long)drange);
  if (mrange == 0 || mrange >= 0xFFFFFFFFull)
    return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: mid key range %llu wider than 32 bits",
                     (unsigned long long)mrange);
  // join structure for mid -> fact: bitmap + rank over dense key ranges (up to a 256 MB
  // bitmap), MurmurHash open addressing otherwise or when flags bit 1 asks for it
  const bool use_bm = !(a->flags & 2) && mrange <= (1ull << 31);
  uint64_t mwords = use_bm ? (mrange + 31) / 32 : 0;
  uint64_t mblk = use_bm ? (mwords + BM_BLOCK_WORDS - 1) / BM_BLOCK_WORDS : 0;
  uint32_t *mbits = nullptr;
  if (use_bm) {
    ALN_TRY(ws(c, mwords, &mbits));
    ALN_TRY(fill_u32(c, mbits, mwords, 0));
  }
  uint64_t bwords = (drange + 31) / 32;
  uint32_t *bitmap, *dup;
  uint64_t *ctr;  // device sums: 0 dim pass, 1 mid pass, 2 mid probe, 3 fact pass, 4 fact probe
  uint32_t *dcnt, *mcnt, *mprobe, *fmatch, *fprobe;
  ALN_TRY(ws(c, bwords, &bitmap));
  ALN_TRY(ws(c, 2, &dup));
  ALN_TRY(ws(c, 8, &ctr));
  ALN_TRY(ws(c, dblocks, &dcnt));
  ALN_TRY(ws(c, mblocks, &mcnt));
  ALN_TRY(ws(c, mblocks, &mprobe));
  ALN_TRY(ws(c, fblocks, &fmatch));
  ALN_TRY(ws(c, fblocks, &fprobe));
  ALN_TRY(fill_u32(c, bitmap, bwords, 0));
  ALN_TRY(fill_u32(c, dup, 2, 0));
  ALN_TRY(fill_u64(c, ctr, 8, 0));

  tm.mark(0);
  {
    ProfScope p(c, "star_dim");
    hipLaunchKernelGGL(k_star_dim, dim3((unsigned)dblocks), dim3(ST_BLOCK), 0, c->stream, dd, a->d_key,
                       dkmin, drange, bitmap, dcnt, dup + 0);
  }
  ALN_TRY(check_launch("star_dim"));
  tm.mark(1);

  // mid scan: passing rows of tile b land in entries [b * ST_TILE, ...) of the region arrays
  int64_t *rkey, *rg1, *rg2;
  ALN_TRY(ws(c, mblocks * ST_TILE, &rkey));
  ALN_TRY(ws(c, mblocks * ST_TILE, &rg1));
  ALN_TRY(ws(c, mblocks * ST_TILE, &rg2));
  {
    ProfScope p(c, "star_mid");
    hipLaunchKernelGGL(k_star_mid, dim3((unsigned)mblocks), dim3(ST_BLOCK), 0, c->stream, md, a->m_key,
                       a->m_fk, a->m_g1, a->m_g2, bitmap, dkmin, drange, mcnt, mprobe, rkey, rg1, rg2,
                       mbits, mkmin, dup + 1);
  }
  ALN_TRY(check_launch("star_mid"));
  ALN_TRY(sum_u32(c, dcnt, dblocks, ctr + 0));
  ALN_TRY(sum_u32(c, mcnt, mblocks, ctr + 1));
  ALN_TRY(sum_u32(c, mprobe, mblocks, ctr + 2));
  uint64_t hc[8];
  uint32_t hdup[2];
  ALN_TRY(copy_to_host(c, hc, ctr, 24));
  ALN_TRY(copy_to_host(c, hdup, dup, 8));
  if (hdup[0]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate dimension keys");
  if (hdup[1]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate mid keys");
  st.dim_pass = hc[0];
  const uint64_t m = hc[1];
  st.mid_pass = m;
  st.mid_probe = hc[2];

  unsigned long long *sum;
  unsigned int *cnt;
  int64_t *okey, *og1, *og2;
  ALN_TRY(ws(c, m ? m : 1, &sum));
  ALN_TRY(ws(c, m ? m : 1, &cnt));
  ALN_TRY(ws(c, m ? m : 1, &okey));
  ALN_TRY(ws(c, m ? m : 1, &og1));
  ALN_TRY(ws(c, m ? m : 1, &og2));
  ALN_TRY(fill_u64(c, (uint64_t *)sum, m, 0));
  ALN_TRY(fill_u32(c, cnt, m, 0));
  uint64_t *table = nullptr, *rank = nullptr;
  uint64_t tcap = 0;
  if (use_bm) {
    // rank directory, then the group attributes in rank (= key) order
    uint32_t *bcnt;
    ALN_TRY(ws(c, mblk, &bcnt));
    ALN_TRY(ws(c, mblk, &rank));
    {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_bm_block_counts, dim3(grid_for(mblk, 256, 65536)), dim3(256), 0, c->stream,
                         mbits, mwords, mblk, bcnt);
    }
    ALN_TRY(check_launch("bm_block_counts"));
    ALN_TRY(exclusive_scan_u32(c, bcnt, mblk, rank, nullptr));
    if (m) {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_bm_scatter, dim3((unsigned)mblocks), dim3(256), 0, c->stream, rkey, rg1, rg2,
                         mcnt, mbits, rank, mkmin, okey, og1, og2);
    }
    ALN_TRY(check_launch("bm_scatter"));
  } else {
    uint64_t *moff;
    ALN_TRY(ws(c, mblocks, &moff));
    ALN_TRY(exclusive_scan_u32(c, mcnt, mblocks, moff, nullptr));
    tcap = m + m / 2 + 1024;
    ALN_TRY(ws(c, tcap, &table));
    ALN_TRY(fill_u64(c, table, tcap, ST_EMPTY));
    if (m) {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_star_build, dim3((unsigned)mblocks), dim3(256), 0, c->stream, rkey, rg1, rg2,
                         mcnt, moff, mkmin, table, tcap, okey, og1, og2, dup + 1);
    }
    ALN_TRY(check_launch("star_build"));
  }
  tm.mark(2);
  if (m) {
    ProfScope p(c, "star_fact");
    if (use_bm)
      hipLaunchKernelGGL(k_star_fact_bm, dim3((unsigned)fblocks), dim3(ST_BLOCK), 0, c->stream, fd,
                         a->f_fk, a->f_x, a->f_y, a->a0, a->a1, a->b0, a->b1, mbits, rank, mkmin, mrange,
                         sum, cnt, fmatch, fprobe);
    else
      hipLaunchKernelGGL(k_star_fact, dim3((unsigned)fblocks), dim3(ST_BLOCK), 0, c->stream, fd, a->f_fk,
                         a->f_x, a->f_y, a->a0, a->a1, a->b0, a->b1, table, tcap, mkmin, mrange, sum, cnt,
                         fmatch, fprobe);
    ALN_TRY(check_launch("star_fact"));
    ALN_TRY(sum_u32(c, fmatch, fblocks, ctr + 3));
    ALN_TRY(sum_u32(c, fprobe, fblocks, ctr + 4));
  }
  tm.mark(3);

  // top-k
  uint32_t k = a->limit;
  if (k && m) {
    const Cand *in = nullptr;
    uint64_t n = m;
    while (true) {
      uint64_t nt = (n + TK_TILE2 - 1) / TK_TILE2;
      Cand *out;
      ALN_TRY(ws(c, nt * k, &out));
      {
        ProfScope p(c, "star_topk");
        hipLaunchKernelGGL(k_star_topk_tile, dim3((unsigned)nt), dim3(256), 0, c->stream, in, n, sum, cnt,
                           okey, og1, og2, k, out);
      }
      ALN_TRY(check_launch("star_topk"));
      in = out;
      n = nt * k;
      if (nt == 1) break;
    }
    std::vector<Cand> hcand(k);
    ALN_TRY(copy_to_host(c, hcand.data(), in, sizeof(Cand) * k));
    uint32_t nt = 0;
    for (uint32_t i = 0; i < k; ++i) {
      if (!hcand[i].valid) break;
      h_top[nt].key = hcand[i].key;
      h_top[nt].g1 = hcand[i].g1;
      h_top[nt].g2 = hcand[i].g2;
      h_top[nt].sum = hcand[i].sum;
      h_top[nt].count = hcand[i].count;
      ++nt;
    }
    *h_ntop = nt;
  }
  tm.mark(4);
  ALN_TRY(copy_to_host(c, hc, ctr, 40));
  uint32_t bdup = 0;
  ALN_TRY(copy_to_host(c, &bdup, dup + 1, 4));
  if (bdup) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate mid keys");
  st.fact_pass = hc[3];
  st.fact_probe = hc[4];

  // every group, in ascending key order
  if ((a->flags & 1) && m) {
    uint32_t *flag, *idx;
    uint64_t *pos, *code;
    ALN_TRY(ws(c, m, &flag));
    ALN_TRY(ws(c, m, &pos));
    hipLaunchKernelGGL(k_star_group_flags, dim3(grid_for(m, 256, 65536)), dim3(256), 0, c->stream, cnt, m,
                       flag);
    ALN_TRY(check_launch("star_group_flags"));
    uint64_t g = 0;
    ALN_TRY(exclusive_scan_u32(c, flag, m, pos, &g));
    st.groups = g;
    if (g > groups_cap) {
      if (h_stats) *h_stats = st;
      return set_error(ALN_E_CAPACITY, "aln_star_agg: %llu groups, capacity %llu",
                       (unsigned long long)g, (unsigned long long)groups_cap);
    }
    if (g) {
      ALN_REQUIRE(d_groups, "aln_star_agg: no group buffer");
      ALN_TRY(ws(c, g, &code));
      ALN_TRY(ws(c, g, &idx));
      hipLaunchKernelGGL(k_star_group_codes, dim3(grid_for(m, 256, 65536)), dim3(256), 0, c->stream, flag,
                         pos, m, okey, mkmin, code, idx);
      ALN_TRY(check_launch("star_group_codes"));
      if (!use_bm) ALN_TRY(radix_sort_pairs(c, code, idx, g, bits_for_range_u(mrange - 1)));
      hipLaunchKernelGGL(k_star_group_emit, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, idx, g,
                         okey, og1, og2, sum, cnt, d_groups);
      ALN_TRY(check_launch("star_group_emit"));
      ALN_HIP(hipStreamSynchronize(c->stream));
    }
  }
  ALN_HIP(hipStreamSynchronize(c->stream));
  st.ms_dim = tm.ms(0, 1);
  st.ms_mid = tm.ms(1, 2);
  st.ms_fact = tm.ms(2, 3);
  st.ms_topk = tm.ms(3, 4);
  if (h_stats) *h_stats = st;
  return ALN_OK;
}

// LSD radix sort, stable multi-key ORDER BY and top-k.
//
// The reference sorts with one thrust::stable_sort_by_key per key, last key first
// (update_permutation, sorts.cu:17-70; emit_order, operators.cu:1675-1763).  Here each key
// is first mapped to an order-preserving unsigned code of just the bits its range needs
// (v - min, or max - v for DESC); when all codes fit in 64 bits they are concatenated and
// sorted once, otherwise one stable radix sort per key runs from the last key to the first.
//
// Radix sort: 8-bit digits, least significant first, each pass stable.  A first kernel
// counts pass 0's digits per block of 16384 items and finds the bits in which the codes
// differ; a pass whose digit every item shares is skipped.  Each pass scans its per-(digit,
// block) counts in digit-major order, which gives every 4096-item tile the first output
// position of each of its digits, and one kernel scatters: a block reorders its tile by
// digit in LDS (ranks from digit matching through LDS lane masks, in memory order) and
// stores every digit's run at consecutive addresses; each XCD takes a contiguous range of
// tiles, so neighbouring runs meet in one L2.  The passes after the first count their own
// digit from the keys the previous pass wrote.  A permutation sort stores only the key
// bytes the later passes read (8 or 4 per key), and its last pass stores none.  The first
// pass of an ORDER BY computes the codes from the key columns itself.  No kernel waits on
// another workgroup.
#include <string.h>

#include <algorithm>
#include <vector>

#include "aln_internal.h"

namespace aln {

struct KeySet {
  const int64_t *k[ALN_MAX_KEYS];
  int64_t ref[ALN_MAX_KEYS];  // min (ASC) or max (DESC)
  uint32_t shift[ALN_MAX_KEYS];
  int32_t desc[ALN_MAX_KEYS];
  int n;
};

// Composite order-preserving code: key 0 in the most significant bits.
__global__ void k_codes(KeySet ks, const uint32_t *__restrict__ perm, uint64_t n,
                        uint64_t *__restrict__ code, uint32_t *__restrict__ idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    uint64_t r = perm ? perm[i] : i;
    uint64_t c = 0;
    for (int k = 0; k < ks.n; ++k) {
      int64_t v = ks.k[k][r];
      uint64_t f = ks.desc[k] ? (uint64_t)ks.ref[k] - (uint64_t)v : (uint64_t)v - (uint64_t)ks.ref[k];
      if (ks.shift[k] < 64) c |= f << ks.shift[k];
    }
    code[i] = c;
    if (idx) idx[i] = (uint32_t)r;
  }
}

__device__ __forceinline__ uint64_t ks_code(const KeySet &ks, uint64_t r) {
  uint64_t c = 0;
  for (int k = 0; k < ks.n; ++k) {
    int64_t v = ks.k[k][r];
    uint64_t f = ks.desc[k] ? (uint64_t)ks.ref[k] - (uint64_t)v : (uint64_t)v - (uint64_t)ks.ref[k];
    if (ks.shift[k] < 64) c |= f << ks.shift[k];
  }
  return c;
}

constexpr int RS_BITS = 8;
constexpr int RS_BINS = 1 << RS_BITS;
constexpr int RS_BLOCK = 512;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;  // items one scatter block reorders in LDS
constexpr int RS_WAVE_ITEMS = 64 * RS_ITEMS;
constexpr int RS_SUB = 4;
constexpr int RS_SUPER = RS_TILE * RS_SUB;  // items one count block counts
constexpr int RS_MAX_PASSES = 8;
constexpr int RS_SCAN_BLOCK = 256;
constexpr uint32_t RS_XCDS = 8;
// From this size on the host reads which bits vary between the codes and skips the passes
// whose digit every item shares; smaller sorts run every pass without a synchronisation.
constexpr uint64_t RS_SYNC_MIN = 1ull << 20;
static_assert(RS_BINS * 2 == RS_BLOCK, "threads 0..255 own one digit each");
static_assert(RS_SUPER % (RS_BLOCK * 4) == 0, "count loop");

template <int B>
struct KeyWord;
template <>
struct KeyWord<4> {
  typedef uint32_t T;
};
template <>
struct KeyWord<8> {
  typedef uint64_t T;
};

// Bytes a key of `bits` bits is stored in: 4 or 8 (scattered stores of narrower words
// cost as much per item as 4-byte ones).
static inline int key_word(uint32_t bits) { return bits <= 32 ? 4 : 8; }

// The items of one pass.  KI (bytes per stored key) 4, 8: keys and values from arrays;
// KI 0: codes computed from the key columns at row r = vals ? vals[i] : i, with value r.
struct RsIn {
  const void *keys;
  const uint32_t *vals;
  KeySet ks;
};

template <int KI>
__device__ __forceinline__ uint64_t rs_key(const RsIn &in, uint64_t i) {
  if constexpr (KI == 0) {
    return ks_code(in.ks, in.vals ? in.vals[i] : i);
  } else {
    typedef typename KeyWord<KI>::T T;
    return (uint64_t)__builtin_nontemporal_load(static_cast<const T *>(in.keys) + i);
  }
}
template <int KI>
__device__ __forceinline__ uint32_t rs_val(const RsIn &in, uint64_t i) {
  if (KI == 0 && !in.vals) return (uint32_t)i;
  return __builtin_nontemporal_load(in.vals + i);
}

// Codes (and, VALS, values) of the N items i0 + j * stride; items at n and beyond get
// code and value 0.  From key columns, the rows are read first and then each column's N
// values, so the loads of all N items are in flight together.
template <int KI, int N, bool VALS>
__device__ __forceinline__ void rs_load(const RsIn &in, uint64_t n, uint64_t i0, uint32_t stride,
                                        uint64_t *key, uint32_t *val) {
  if constexpr (KI == 0) {
    uint64_t r[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      uint64_t i = i0 + (uint64_t)j * stride;
      r[j] = i < n ? (in.vals ? (uint64_t)__builtin_nontemporal_load(in.vals + i) : i) : 0;
      key[j] = 0;
      if (VALS) val[j] = (uint32_t)r[j];
    }
    for (int k = 0; k < in.ks.n; ++k) {
      const int64_t *col = in.ks.k[k];
      const uint64_t ref = (uint64_t)in.ks.ref[k];
      const uint32_t sh = in.ks.shift[k];
      const bool desc = in.ks.desc[k];
      int64_t v[N];
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] = col[r[j]];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        uint64_t f = desc ? ref - (uint64_t)v[j] : (uint64_t)v[j] - ref;
        if (sh < 64) key[j] |= f << sh;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      uint64_t i = i0 + (uint64_t)j * stride;
      key[j] = i < n ? rs_key<KI>(in, i) : 0;
      if (VALS) val[j] = i < n ? rs_val<KI>(in, i) : 0;
    }
  }
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const unsigned lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (unsigned)d) x += y;
  }
  return x;
}

// Digit counts of one pass, block s over items [s * RS_SUPER, +RS_SUPER): cnt[digit *
// nsuper + s] (digit-major, for the scan) and, per tile of the block, the count of each
// digit in the block's earlier tiles: sub[(s * RS_SUB + tile) * 256 + digit].  VARY: also
// ORs code ^ (code of item 0) into *vary, the bits in which the codes differ.
template <int KI, bool VARY>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_count(RsIn in, uint64_t n, uint32_t sh, uint32_t mask,
                                                       uint32_t nsuper, uint32_t *__restrict__ cnt,
                                                       uint16_t *__restrict__ sub,
                                                       unsigned long long *__restrict__ vary) {
  __shared__ uint32_t h[RS_SUB][RS_BINS];
  __shared__ uint64_t wv[RS_WAVES];
  const uint32_t t = threadIdx.x, s = blockIdx.x;
  for (uint32_t i = t; i < RS_SUB * RS_BINS; i += RS_BLOCK) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)s * RS_SUPER;
  uint64_t v = 0;
  const uint64_t c0 = VARY ? rs_key<KI>(in, 0) : 0;
  for (int r = 0; r < RS_SUPER / (RS_BLOCK * 4); ++r) {
    uint64_t k[4];
    const uint64_t i0 = lo + (uint64_t)r * 4 * RS_BLOCK + t;
    rs_load<KI, 4, false>(in, n, i0, RS_BLOCK, k, nullptr);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      uint64_t i = i0 + (uint64_t)u * RS_BLOCK;
      if (i < n) {
        atomicAdd(&h[(uint32_t)((i - lo) / RS_TILE)][(uint32_t)(k[u] >> sh) & mask], 1u);
        if (VARY) v |= k[u] ^ c0;
      }
    }
  }
  if (VARY) {
    for (int d = 32; d > 0; d >>= 1) v |= __shfl_xor(v, d, 64);
    if (__lane_id() == 0) wv[t / 64] = v;
  }
  __syncthreads();
  if (t < RS_BINS) {
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < RS_SUB; ++j) {
      sub[((uint64_t)s * RS_SUB + j) * RS_BINS + t] = (uint16_t)run;
      run += h[j][t];
    }
    cnt[(uint64_t)t * nsuper + s] = run;
  }
  if (VARY && t == 0) {
    uint64_t all = 0;
    for (int w = 0; w < RS_WAVES; ++w) all |= wv[w];
    if (all) atomicOr(vary, (unsigned long long)all);
  }
}

// Exclusive scan of each digit's row of block counts, written transposed for the scatter
// blocks: excl[s * 256 + d]; tot[d] = the row's sum (one block per digit).
__global__ void __launch_bounds__(RS_SCAN_BLOCK) k_rs_rowscan(const uint32_t *__restrict__ cnt, uint32_t nsuper,
                                                              uint32_t *__restrict__ excl,
                                                              uint32_t *__restrict__ tot) {
  __shared__ uint32_t wsum[RS_SCAN_BLOCK / 64];
  const uint32_t d = blockIdx.x, t = threadIdx.x, lane = __lane_id(), wid = t / 64;
  const uint32_t per = (nsuper + RS_SCAN_BLOCK - 1) / RS_SCAN_BLOCK;
  const uint32_t *row = cnt + (uint64_t)d * nsuper;
  const uint32_t b = t * per, e = b + per < nsuper ? b + per : nsuper;
  uint32_t local = 0;
  for (uint32_t i = b; i < e; ++i) local += row[i];
  uint32_t incl = wave_incl_scan(local);
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t before = incl - local;
  for (uint32_t w = 0; w < wid; ++w) before += wsum[w];
  for (uint32_t i = b; i < e; ++i) {
    uint32_t x = row[i];
    excl[(uint64_t)i * RS_BINS + d] = before;
    before += x;
  }
  if (t == RS_SCAN_BLOCK - 1) tot[d] = before;
}

// One pass: digit = (stored key >> sh) & mask; keys are written as (stored key >> oshift)
// in KO bytes (KO 0: not written).  Tile b of count block s = b / RS_SUB stores digit d
// from position base[d] + excl[s * 256 + d] + sub[b * 256 + d], base[d] = the items of
// the digits below d (an exclusive scan of tot, made by every block).
struct RsPass {
  RsIn in;
  uint64_t n;
  const uint32_t *tot, *excl;
  const uint16_t *sub;
  void *kout;
  uint32_t *vout;
  uint32_t sh, mask, oshift, nsuper;
};

// One tile of RS_TILE items per block (8 waves x 64 lanes x 8 items; wave, then (item,
// lane) order is memory order): LDS lane masks match the digits of each wave row, per-wave
// digit counts place the waves, the tile is reordered by digit in LDS and each digit's run
// is stored at consecutive addresses.
template <int KI, int KO>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(RsPass a) {
  __shared__ uint64_t sk[RS_TILE];
  __shared__ uint32_t sv[RS_TILE];
  __shared__ uint16_t wc[RS_WAVES][RS_BINS];
  __shared__ uint32_t gdel[RS_BINS];
  __shared__ uint32_t wsum[RS_BINS / 64], bsum[RS_BINS / 64];
  const unsigned t = threadIdx.x, lane = __lane_id(), wid = t / 64;
  // Workgroups go to the 8 XCDs round-robin: give each XCD a contiguous range of tiles, so
  // neighbouring tiles' runs of a digit (adjacent in the output) meet in the same L2.
  const uint32_t q = gridDim.x / RS_XCDS, rm = gridDim.x % RS_XCDS;
  const uint32_t x = blockIdx.x % RS_XCDS;
  const uint32_t tile = x * q + (x < rm ? x : rm) + blockIdx.x / RS_XCDS;
  const uint64_t n = a.n;
  const uint32_t sh = a.sh, mask = a.mask;
  // during ranking, sk holds one 64-bit lane mask per (wave, digit)
  uint64_t *pm = sk;
  static_assert(RS_WAVES * RS_BINS <= RS_TILE, "masks fit in the key buffer");
  for (int w = t; w < RS_WAVES * RS_BINS; w += RS_BLOCK) {
    (&wc[0][0])[w] = 0;
    pm[w] = 0;
  }
  const uint64_t tbase = (uint64_t)tile * RS_TILE;
  const uint64_t base = tbase + (uint64_t)wid * RS_WAVE_ITEMS;
  // this tile's first output position of digit t, read before the items so that its
  // latency overlaps theirs
  uint32_t first = 0, dtot = 0;
  if (t < RS_BINS) {
    first = a.excl[(uint64_t)(tile / RS_SUB) * RS_BINS + t] + a.sub[(uint64_t)tile * RS_BINS + t];
    dtot = a.tot[t];
  }
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS], rank[RS_ITEMS];
  rs_load<KI, RS_ITEMS, true>(a.in, n, base + lane, 64, key, val);
  __syncthreads();
  // Lanes of one wave row with equal digits: each lane ORs its bit into the row's mask of
  // its digit and reads the mask back (a wave's LDS operations complete in order); the
  // lowest such lane then advances the digit's count and clears the mask.
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    bool valid = base + (uint64_t)j * 64 + lane < n;
    uint32_t d = (uint32_t)(key[j] >> sh) & mask;
    uint64_t *slot = pm + wid * RS_BINS + d;
    if (valid) __hip_atomic_fetch_or(slot, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    uint64_t peers = valid ? __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
    uint32_t before = valid ? wc[wid][d] : 0;
    const uint32_t below = below_count(peers);  // peers in lower lanes
    rank[j] = before + below;
    if (valid && below == 0) {
      wc[wid][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
      __hip_atomic_store(slot, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  uint32_t cnt = 0, incl = 0, bincl = 0;
  if (t < RS_BINS) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      uint32_t x = wc[w][t];
      wc[w][t] = (uint16_t)cnt;
      cnt += x;
    }
    incl = wave_incl_scan(cnt);
    bincl = wave_incl_scan(dtot);
    if (lane == 63) {
      wsum[wid] = incl;
      bsum[wid] = bincl;
    }
  }
  __syncthreads();
  if (t < RS_BINS) {
    uint32_t tstart = incl - cnt, dbase = bincl - dtot;
    for (unsigned w = 0; w < wid; ++w) {
      tstart += wsum[w];
      dbase += bsum[w];
    }
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wc[w][t] = (uint16_t)(wc[w][t] + tstart);
    gdel[t] = dbase + first - tstart;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    if (base + (uint64_t)j * 64 + lane < n) {
      uint32_t d = (uint32_t)(key[j] >> sh) & mask;
      uint32_t pos = wc[wid][d] + rank[j];
      sk[pos] = key[j];
      sv[pos] = val[j];
    }
  }
  __syncthreads();
  const uint32_t nvalid = n - tbase < (uint64_t)RS_TILE ? (uint32_t)(n - tbase) : (uint32_t)RS_TILE;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint32_t i = (uint32_t)j * RS_BLOCK + t;
    if (i < nvalid) {
      uint64_t k = sk[i];
      uint32_t dest = gdel[(uint32_t)(k >> sh) & mask] + i;
      if constexpr (KO != 0) {
        typedef typename KeyWord<KO>::T T;
        static_cast<T *>(a.kout)[dest] = (T)(k >> a.oshift);
      }
      a.vout[dest] = sv[i];
    }
  }
}

// ---- bucket sort after two top-digit passes ------------------------------------------------
// Once two stable passes on the two most significant varying digits have put the rows in
// order of their codes' high part (code >> topshift, "buckets"), each block of k_rs_local
// owns the buckets that start in its window of RL_W positions and sorts them, with their
// continuation past the window, in LDS; only the permutation is stored.  A bucket that does
// not end within RS_TILE positions of its window sets *overflow (the host then sorts by the
// digit passes alone).
//
// In the block each code's high part is replaced by its bucket's ordinal, the rows of
// [lo, hi) are packed to LDS positions 0 .. m-1, and three or four stable LDS passes rank
// them by the next two varying digits and then the ordinal.  Rows that still tie on every
// bit above the remaining digits stand in runs of consecutive positions (short for spread
// keys); each row's place in its run is the number of rows of the run with a smaller code,
// or an equal code and an earlier position.  A run longer than RL_RUN makes the block rank
// all remaining digits by LDS passes instead, from the order it has (which keeps equal
// codes in input order, so the result is the same).
constexpr uint32_t RL_W = RS_TILE / 2;
constexpr uint32_t RL_RUN = 32;

struct RsLocal {
  const uint64_t *keys;
  const uint32_t *vals;
  uint32_t *vout;
  uint32_t *overflow;
  uint64_t n;
  uint32_t nblk, topshift, nlow;  // nlow >= 2 remaining varying digits:
  uint32_t sh[RS_MAX_PASSES], mask[RS_MAX_PASSES];  // least significant first
};

// 6 waves per SIMD: three blocks per CU, as many as the LDS holds
__global__ void __launch_bounds__(RS_BLOCK) __attribute__((amdgpu_waves_per_eu(6, 6))) k_rs_local(RsLocal a) {
  __shared__ uint64_t sk[RS_TILE];
  __shared__ uint32_t sv[RS_TILE];
  __shared__ uint16_t wc[RS_WAVES][RS_BINS];
  __shared__ uint32_t wsum[RS_WAVES];
  __shared__ uint32_t s_lo, s_hi, s_long;
  const unsigned t = threadIdx.x, lane = __lane_id(), wid = t / 64;
  // each XCD takes a contiguous range of windows: neighbours share RL_W codes and one L2
  const uint32_t q = a.nblk / RS_XCDS, rm = a.nblk % RS_XCDS, x = blockIdx.x % RS_XCDS;
  const uint64_t n = a.n, w0 = (uint64_t)(x * q + min(x, rm) + blockIdx.x / RS_XCDS) * RL_W;
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS], rank[RS_ITEMS];
  if (t == 0) {
    s_lo = 0xFFFFFFFFu;
    s_hi = 0xFFFFFFFFu;
    s_long = 0;
  }
  // the codes and values of positions w0 + i, i = wave * 512 + row * 64 + lane (memory
  // order); the values past the block's rows are mostly its neighbour's, in the same L2
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t i = wid * RS_WAVE_ITEMS + j * 64 + lane;
    key[j] = w0 + i < n ? __builtin_nontemporal_load(a.keys + w0 + i) : ~0ull;
    val[j] = w0 + i < n ? __builtin_nontemporal_load(a.vals + w0 + i) : 0u;
    sk[i] = key[j] >> a.topshift;
  }
  const uint64_t tprev = w0 > 0 ? a.keys[w0 - 1] >> a.topshift : ~0ull;
  __syncthreads();
  // bucket starts: the first in the window (lo) and the first at or past its end (hi)
  bool st[RS_ITEMS];
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t i = wid * RS_WAVE_ITEMS + j * 64 + lane;
    st[j] = w0 + i < n && (w0 + i == 0 || (i == 0 ? tprev : sk[i - 1]) != sk[i]);
    if (st[j]) atomicMin(i < RL_W ? &s_lo : &s_hi, i);
  }
  __syncthreads();
  const uint32_t lo = s_lo;
  if (lo == 0xFFFFFFFFu) return;  // no bucket starts here (uniform across the block)
  uint32_t hi = s_hi;
  if (hi == 0xFFFFFFFFu) {
    if (w0 + RS_TILE < n) {  // the last bucket runs past the block's capacity
      if (t == 0) atomicOr(a.overflow, 1u);
      return;
    }
    hi = (uint32_t)(n - w0);
  }
  const uint32_t m = hi - lo;
  // bucket ordinals (starts in (lo, i]) by a scan in memory order
  uint32_t carry = 0;
  uint32_t ord[RS_ITEMS];
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t i = wid * RS_WAVE_ITEMS + j * 64 + lane;
    const uint32_t inc = wave_incl_scan((st[j] && i > lo && i < hi) ? 1u : 0u);
    ord[j] = carry + inc;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) wsum[wid] = carry;
  __syncthreads();  // also: every read of the high parts in sk is done
  uint32_t obase = 0, nb = 1;  // nb: buckets in [lo, hi)
  for (unsigned w = 0; w < RS_WAVES; ++w) {
    if (w < wid) obase += wsum[w];
    nb += wsum[w];
  }
  const uint64_t lowmask = (1ull << a.topshift) - 1;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t i = wid * RS_WAVE_ITEMS + j * 64 + lane;
    if (i >= lo && i < hi) {
      sk[i - lo] = ((uint64_t)(obase + ord[j]) << a.topshift) | (key[j] & lowmask);
      sv[i - lo] = val[j];
    }
  }
  // from here on a wave holds rows wid * S + j * 64 + lane, j < J: every wave has work
  const uint32_t J = (m + RS_BLOCK - 1) / RS_BLOCK, S = 64 * J;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t i = wid * S + j * 64 + lane;
    if (j < (int)J) {
      key[j] = i < m ? sk[i] : 0;
      val[j] = i < m ? sv[i] : 0;
    }
  }
  // Stable LDS passes, each on the digit (key >> sh) & mask with ranks in position order:
  // first the two digits above the cut and the ordinal (1 or 2 passes: nb <= RS_TILE), then,
  // only after a long run, every remaining digit and the ordinal again.
  const uint32_t nl = a.nlow, cut = a.sh[nl - 2];
  const uint32_t nord = (nb > 1 ? 1u : 0u) + (nb > RS_BINS ? 1u : 0u), npre = 2 + nord;
  uint32_t npass = npre;
  for (uint32_t ps = 0; ps < npass; ++ps) {
    uint32_t sh, mask;
    if (ps < npre) {
      sh = ps < 2 ? a.sh[nl - 2 + ps] : a.topshift + RS_BITS * (ps - 2);
      mask = ps < 2 ? a.mask[nl - 2 + ps] : 0xFFu;
    } else {
      const uint32_t r = ps - npre;
      sh = r < nl ? a.sh[r] : a.topshift + RS_BITS * (r - nl);
      mask = r < nl ? a.mask[r] : 0xFFu;
    }
    uint64_t *pm = sk;  // lane masks of [wave][digit]
    __syncthreads();    // every read of sk and sv before this pass is done
    for (int w = t; w < RS_WAVES * RS_BINS; w += RS_BLOCK) {
      (&wc[0][0])[w] = 0;
      pm[w] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      if (j >= (int)J) break;
      const uint32_t i = wid * S + j * 64 + lane;
      const bool valid = i < m;
      const uint32_t d = (uint32_t)(key[j] >> sh) & mask;
      uint64_t *slot = pm + wid * RS_BINS + d;
      if (valid) __hip_atomic_fetch_or(slot, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const uint64_t peers = valid ? __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
      const uint32_t before = valid ? wc[wid][d] : 0;
      const uint32_t below = below_count(peers);  // peers in lower lanes
      rank[j] = before + below;
      if (valid && below == 0) {
        wc[wid][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
        __hip_atomic_store(slot, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    uint32_t cnt = 0, incl = 0;
    if (t < RS_BINS) {
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) {
        const uint32_t c = wc[w][t];
        wc[w][t] = (uint16_t)cnt;
        cnt += c;
      }
      incl = wave_incl_scan(cnt);
      if (lane == 63) wsum[wid] = incl;
    }
    __syncthreads();
    if (t < RS_BINS) {
      uint32_t tstart = incl - cnt;
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) tstart += w < (int)wid ? wsum[w] : 0u;
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) wc[w][t] = (uint16_t)(wc[w][t] + tstart);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      if (j >= (int)J) break;
      const uint32_t i = wid * S + j * 64 + lane;
      if (i < m) {
        const uint32_t pos = wc[wid][(uint32_t)(key[j] >> sh) & mask] + rank[j];
        sk[pos] = key[j];
        sv[pos] = val[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      if (j >= (int)J) break;
      const uint32_t i = wid * S + j * 64 + lane;
      key[j] = i < m ? sk[i] : 0;
      val[j] = i < m ? sv[i] : 0;
    }
    if (ps + 1 != npre || nl == 2) continue;
    // rows that tie on key >> cut stand in runs; a run longer than RL_RUN holds two rows
    // RL_RUN apart
    bool lng = false;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      if (j >= (int)J) break;
      const uint32_t i = wid * S + j * 64 + lane;
      lng |= i + RL_RUN < m && (sk[i + RL_RUN] >> cut) == (key[j] >> cut);
    }
    if (lng) s_long = 1;
    __syncthreads();
    if (s_long) npass += nl + nord;
  }
  if (nl == 2 || s_long) {  // the rows are in order
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      if (j >= (int)J) break;
      const uint32_t i = wid * S + j * 64 + lane;
      if (i < m) a.vout[w0 + lo + i] = val[j];
    }
    return;
  }
  // each row's place in its run: the rows of the run with a smaller code, or an equal code
  // and an earlier position.  Runs of 2 from the neighbours' codes; longer runs (rare for
  // spread keys) by a scan of the run.
  uint32_t slow = 0;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    if (j >= (int)J) break;
    const uint32_t i = wid * S + j * 64 + lane;
    if (i >= m) continue;
    const uint64_t p = key[j] >> cut;
    const uint64_t km1 = i >= 1 ? sk[i - 1] : 0, kp1 = i + 1 < m ? sk[i + 1] : 0;
    const uint64_t km2 = i >= 2 ? sk[i - 2] : 0, kp2 = i + 2 < m ? sk[i + 2] : 0;
    const bool em1 = i >= 1 && (km1 >> cut) == p, ep1 = i + 1 < m && (kp1 >> cut) == p;
    const bool em2 = em1 && i >= 2 && (km2 >> cut) == p, ep2 = ep1 && i + 2 < m && (kp2 >> cut) == p;
    uint32_t pos = i;
    if (em1 && ep1 || em2 || ep2) {
      slow |= 1u << j;
      continue;
    }
    if (em1) pos = km1 <= key[j] ? i : i - 1;  // run [i-1, i]
    if (ep1) pos = kp1 < key[j] ? i + 1 : i;   // run [i, i+1]
    a.vout[w0 + lo + pos] = val[j];
  }
#pragma unroll 1
  for (uint32_t j = 0; j < J; ++j) {
    if (!((slow >> j) & 1)) continue;
    const uint32_t i = wid * S + j * 64 + lane;
    const uint64_t kk = sk[i], p = kk >> cut;
    uint32_t rs = i, re = i + 1;
    while (rs > 0 && (sk[rs - 1] >> cut) == p) --rs;
    while (re < m && (sk[re] >> cut) == p) ++re;
    uint32_t pos = rs;
    for (uint32_t k = rs; k < re; ++k) {
      const uint64_t c = sk[k];
      pos += c < kk || (c == kk && k < i);
    }
    a.vout[w0 + lo + pos] = sv[i];
  }
}

__global__ void k_rs_vals(const uint32_t *__restrict__ vals, uint64_t n, uint32_t *__restrict__ vout) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    vout[i] = vals ? vals[i] : (uint32_t)i;
}

__global__ void k_copy_pairs(const uint64_t *ks, const uint32_t *vs, uint64_t n, uint64_t *kd,
                             uint32_t *vd) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

template <int KI, int KO>
static void rs_scatter_launch(aln_ctx *c, uint32_t ntiles, const RsPass &a) {
  hipLaunchKernelGGL((k_rs_scatter<KI, KO>), dim3(ntiles), dim3(RS_BLOCK), 0, c->stream, a);
}

static int rs_scatter(aln_ctx *c, int ki, int ko, uint32_t ntiles, const RsPass &a) {
  switch (ki * 16 + ko) {
#define RS_CASE(x, y)                      \
  case x * 16 + y:                         \
    rs_scatter_launch<x, y>(c, ntiles, a); \
    break;
    RS_CASE(0, 0) RS_CASE(0, 4) RS_CASE(0, 8)
    RS_CASE(8, 0) RS_CASE(8, 4) RS_CASE(8, 8)
    RS_CASE(4, 0) RS_CASE(4, 4)
#undef RS_CASE
    default:
      ALN_REQUIRE(false, "radix sort: key widths %d -> %d", ki, ko);
  }
  return check_launch("radix_scatter");
}

template <int KI, bool VARY>
static void rs_count_launch(aln_ctx *c, uint32_t nsuper, const RsIn &in, uint64_t n, uint32_t sh,
                            uint32_t mask, uint32_t *cnt, uint16_t *sub, unsigned long long *vary) {
  hipLaunchKernelGGL((k_rs_count<KI, VARY>), dim3(nsuper), dim3(RS_BLOCK), 0, c->stream, in, n, sh, mask,
                     nsuper, cnt, sub, vary);
}

static int rs_count(aln_ctx *c, int ki, bool vary, uint32_t nsuper, const RsIn &in, uint64_t n,
                    uint32_t sh, uint32_t mask, uint32_t *cnt, uint16_t *sub, unsigned long long *d_vary) {
  switch (ki * 2 + (vary ? 1 : 0)) {
    case 0: rs_count_launch<0, false>(c, nsuper, in, n, sh, mask, cnt, sub, d_vary); break;
    case 1: rs_count_launch<0, true>(c, nsuper, in, n, sh, mask, cnt, sub, d_vary); break;
    case 8: rs_count_launch<4, false>(c, nsuper, in, n, sh, mask, cnt, sub, d_vary); break;
    case 16: rs_count_launch<8, false>(c, nsuper, in, n, sh, mask, cnt, sub, d_vary); break;
    case 17: rs_count_launch<8, true>(c, nsuper, in, n, sh, mask, cnt, sub, d_vary); break;
    default:
      ALN_REQUIRE(false, "radix sort: key width %d", ki);
  }
  return check_launch("radix_count");
}

// The bucket path of rs_sort (see k_rs_local): *done = false when a bucket overflowed.
// kt and vt: two code arrays of 8n bytes and two row arrays of n words.
static int rs_hybrid(aln_ctx *c, const RsIn &in, bool from_keys, uint64_t n, uint32_t bits, const uint32_t *run,
                     uint32_t nrun, int counted, uint32_t *cnt, uint32_t *excl, uint32_t *tot,
                     uint16_t *sub, uint64_t *const kt[2], uint32_t *const vt[2], uint32_t *vout,
                     bool *done) {
  *done = false;
  const uint32_t nsuper = (uint32_t)((n + RS_SUPER - 1) / RS_SUPER);
  const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  auto pass_mask = [&](uint32_t p) {
    uint32_t pb = bits - RS_BITS * p < RS_BITS ? bits - RS_BITS * p : RS_BITS;
    return (1u << pb) - 1;
  };
  uint32_t *ovf;
  ALN_TRY(ws(c, 1, &ovf));
  ALN_TRY(fill_u32(c, ovf, 1, 0));
  // two passes: the second most significant varying digit, then the most significant
  for (uint32_t r = 0; r < 2; ++r) {
    const uint32_t p = run[nrun - 2 + r];
    RsPass a;
    a.in = in;
    if (r > 0) {
      a.in.keys = kt[0];
      a.in.vals = vt[0];
    }
    a.n = n;
    a.tot = tot;
    a.excl = excl;
    a.sub = sub;
    a.kout = kt[r];
    a.vout = vt[r];
    a.sh = RS_BITS * p;
    a.mask = pass_mask(p);
    a.oshift = 0;
    a.nsuper = nsuper;
    const int ki = r == 0 ? (from_keys ? 0 : 8) : 8;
    if (r > 0 || (int)p != counted) {
      ProfScope ps(c, "radix_count");
      ALN_TRY(rs_count(c, ki, false, nsuper, a.in, n, a.sh, a.mask, cnt, sub, nullptr));
    }
    {
      ProfScope ps(c, "radix_scan");
      hipLaunchKernelGGL(k_rs_rowscan, dim3(RS_BINS), dim3(RS_SCAN_BLOCK), 0, c->stream, cnt, nsuper, excl, tot);
    }
    ALN_TRY(check_launch("radix_scan"));
    {
      ProfScope ps(c, "radix_scatter");
      ALN_TRY(rs_scatter(c, ki, 8, ntiles, a));
    }
  }
  // the permutation goes to vout unless vout is also the input's row list, which the digit
  // passes would read again after an overflow: then to vt[0] (read by the second pass, now
  // free) and copied over on success
  const bool alias = in.vals && in.vals == vout;
  RsLocal l;
  memset(&l, 0, sizeof(l));
  l.keys = kt[1];
  l.vals = vt[1];
  l.vout = alias ? vt[0] : vout;
  l.overflow = ovf;
  l.n = n;
  l.nblk = (uint32_t)((n + RL_W - 1) / RL_W);
  l.topshift = RS_BITS * run[nrun - 2];
  l.nlow = nrun - 2;
  for (uint32_t r = 0; r + 2 < nrun; ++r) {
    l.sh[r] = RS_BITS * run[r];
    l.mask[r] = pass_mask(run[r]);
  }
  {
    ProfScope ps(c, "radix_local");
    hipLaunchKernelGGL(k_rs_local, dim3(l.nblk), dim3(RS_BLOCK), 0, c->stream, l);
  }
  ALN_TRY(check_launch("radix_local"));
  uint32_t o = 0;
  ALN_TRY(copy_to_host(c, &o, ovf, 4));
  *done = o == 0;
  if (*done && alias) ALN_HIP(hipMemcpyAsync(vout, vt[0], 4ull * n, hipMemcpyDeviceToDevice, c->stream));
  return ALN_OK;
}

// The bits in which the codes of the first RS_SAMPLE positions differ (a subset of those in
// which all codes differ), computed on the host.
constexpr uint64_t RS_SAMPLE = 4096;
static int sample_vary(aln_ctx *c, const RsIn &in, bool from_keys, uint64_t n, uint64_t *vary) {
  const uint64_t m = std::min<uint64_t>(n, RS_SAMPLE);
  std::vector<uint64_t> code(m, 0);
  if (!from_keys) {
    ALN_TRY(copy_to_host(c, code.data(), in.keys, 8 * m));
  } else {  // the first rows' keys: a sample of the rows whatever order in.vals gives them
    std::vector<int64_t> v(m);
    for (int k = 0; k < in.ks.n; ++k) {
      ALN_TRY(copy_to_host(c, v.data(), in.ks.k[k], 8 * m));
      for (uint64_t r = 0; r < m; ++r) {
        const uint64_t f = in.ks.desc[k] ? (uint64_t)in.ks.ref[k] - (uint64_t)v[r] : (uint64_t)v[r] - (uint64_t)in.ks.ref[k];
        if (in.ks.shift[k] < 64) code[r] |= f << in.ks.shift[k];
      }
    }
  }
  *vary = 0;
  for (uint64_t r = 0; r < m; ++r) *vary |= code[r] ^ code[0];
  return ALN_OK;
}

// Stable LSD sort of the items of `in` by the low `bits` bits of their codes.  from_keys:
// the codes come from in.ks (first pass), else from the 8-byte array in.keys.  The values
// land in vout (may be in.vals).  kout given: the codes land there, whole (may be in.keys);
// otherwise each pass stores only the key bytes the later passes read.
static int rs_sort(aln_ctx *c, RsIn in, bool from_keys, uint64_t n, uint32_t bits, uint64_t *kout,
                   uint32_t *vout) {
  if (n == 0) return ALN_OK;
  ALN_REQUIRE(bits <= 64, "radix sort: %u key bits", bits);
  ALN_REQUIRE(n < (1ull << 32), "radix sort: more than 2^32 items");
  const bool keys_out = kout != nullptr;
  const uint32_t passes = (bits + RS_BITS - 1) / RS_BITS;
  const uint32_t nsuper = (uint32_t)((n + RS_SUPER - 1) / RS_SUPER);
  const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  const uint64_t ncnt = (uint64_t)RS_BINS * nsuper;
  const int ki0 = from_keys ? 0 : 8;
  auto pass_mask = [&](uint32_t p) {
    uint32_t pb = bits - RS_BITS * p < RS_BITS ? bits - RS_BITS * p : RS_BITS;
    return (1u << pb) - 1;
  };
  uint32_t *cnt = nullptr, *excl = nullptr, *tot = nullptr;
  uint16_t *sub = nullptr;
  unsigned long long *d_vary = nullptr;
  if (passes > 0) {
    ALN_TRY(ws(c, ncnt, &cnt));
    ALN_TRY(ws(c, ncnt, &excl));
    ALN_TRY(ws(c, (size_t)RS_BINS, &tot));
    ALN_TRY(ws(c, (size_t)nsuper * RS_SUB * RS_BINS, &sub));
    ALN_TRY(ws(c, 1, &d_vary));
  }
  // the bits in which the codes differ, and the counts of one digit of the input: the first
  // the bucket path sorts by when it may run (the second most significant), else digit 0
  const bool sync = n >= RS_SYNC_MIN;
  const bool bucket = !keys_out && sync && passes >= 4;
  int counted = -1;  // the digit whose counts of the input are in cnt and sub
  if (passes > 0) {
    counted = 0;
    if (bucket) {  // the second most significant digit in which a sample of the codes differ
      uint64_t sv = 0;
      ALN_TRY(sample_vary(c, in, from_keys, n, &sv));
      int top = -1;
      for (int p = (int)passes - 1; p >= 0; --p)
        if ((sv >> (RS_BITS * p)) & pass_mask(p)) {
          if (top >= 0) {
            counted = p;
            break;
          }
          top = p;
        }
    }
    if (sync) ALN_TRY(fill_u64(c, reinterpret_cast<uint64_t *>(d_vary), 1, 0));
    ProfScope ps(c, "radix_count");
    ALN_TRY(rs_count(c, ki0, sync, nsuper, in, n, RS_BITS * counted, pass_mask(counted), cnt, sub, d_vary));
  }
  uint64_t vary = ~0ull;
  if (sync && passes > 0) ALN_TRY(copy_to_host(c, &vary, d_vary, 8));
  uint32_t run[RS_MAX_PASSES], nrun = 0;
  for (uint32_t p = 0; p < passes; ++p)
    if ((vary >> (RS_BITS * p)) & pass_mask(p)) run[nrun++] = p;
  if (nrun == 0) {  // every code equal: the input order is sorted
    if (from_keys || in.vals != vout) {
      hipLaunchKernelGGL(k_rs_vals, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, in.vals, n, vout);
      ALN_TRY(check_launch("radix_vals"));
    }
    ALN_REQUIRE(!keys_out || !from_keys, "radix sort: codes of the key columns are not kept");
    if (keys_out && kout != in.keys)
      ALN_HIP(hipMemcpyAsync(kout, in.keys, 8ull * n, hipMemcpyDeviceToDevice, c->stream));
    return ALN_OK;
  }
  const bool alias1 = nrun == 1 && ((in.vals && in.vals == vout) || (keys_out && in.keys == kout));
  void *kt[2] = {nullptr, nullptr};
  uint32_t *vt[2] = {nullptr, nullptr};
  if (nrun > 1 || alias1) {
    ALN_TRY(ws(c, 8 * n, reinterpret_cast<uint8_t **>(&kt[0])));
    ALN_TRY(ws(c, 8 * n, reinterpret_cast<uint8_t **>(&kt[1])));
    ALN_TRY(ws(c, n, &vt[0]));
    ALN_TRY(ws(c, n, &vt[1]));
  }
  // A permutation sort of many rows with at least 4 varying digits: two passes on the two
  // most significant ones, then every bucket sorted on the rest in LDS (k_rs_local); if a
  // bucket is too large for that, the digit passes below run instead, from the start, and
  // count pass 0 again (the bucket path's passes used the count buffers).
  if (bucket && nrun >= 4) {
    bool done = false;
    uint64_t *const kt64[2] = {static_cast<uint64_t *>(kt[0]), static_cast<uint64_t *>(kt[1])};
    ALN_TRY(rs_hybrid(c, in, from_keys, n, bits, run, nrun, counted, cnt, excl, tot, sub, kt64, vt, vout, &done));
    if (done) return ALN_OK;
    counted = -1;
  }
  // Bytes each pass reads (ki) and writes (ko) per key: the bytes below the next pass's
  // digit are dropped and the rest is stored in the smallest word that holds it.
  int ki[RS_MAX_PASSES], ko[RS_MAX_PASSES];
  uint32_t qin[RS_MAX_PASSES], qout[RS_MAX_PASSES];
  for (uint32_t r = 0; r < nrun; ++r) {
    qin[r] = r == 0 ? 0 : qout[r - 1];
    ki[r] = r == 0 ? ki0 : ko[r - 1];
    if (keys_out) {
      qout[r] = 0;
      ko[r] = 8;
    } else if (r + 1 < nrun) {
      qout[r] = run[r + 1];
      ko[r] = key_word(bits - RS_BITS * qout[r]);
    } else {
      qout[r] = qin[r];
      ko[r] = 0;
    }
  }
  const void *kprev = nullptr;
  const uint32_t *vprev = nullptr;
  for (uint32_t r = 0; r < nrun; ++r) {
    const uint32_t p = run[r];
    const bool direct = r + 1 == nrun && !alias1;
    RsPass a;
    a.in = in;
    if (r > 0) {
      a.in.keys = kprev;
      a.in.vals = vprev;
    }
    a.n = n;
    a.tot = tot;
    a.excl = excl;
    a.sub = sub;
    a.kout = ko[r] == 0 ? nullptr : direct ? static_cast<void *>(kout) : kt[(nrun - 1 - r) & 1];
    a.vout = direct ? vout : vt[(nrun - 1 - r) & 1];
    a.sh = RS_BITS * (p - qin[r]);
    a.mask = pass_mask(p);
    a.oshift = RS_BITS * (qout[r] - qin[r]);
    a.nsuper = nsuper;
    if (r > 0 || (int)p != counted) {
      ProfScope ps(c, "radix_count");
      ALN_TRY(rs_count(c, ki[r], false, nsuper, a.in, n, a.sh, a.mask, cnt, sub, nullptr));
    }
    {
      ProfScope ps(c, "radix_scan");
      hipLaunchKernelGGL(k_rs_rowscan, dim3(RS_BINS), dim3(RS_SCAN_BLOCK), 0, c->stream, cnt, nsuper, excl, tot);
    }
    ALN_TRY(check_launch("radix_scan"));
    {
      ProfScope ps(c, "radix_scatter");
      ALN_TRY(rs_scatter(c, ki[r], ko[r], ntiles, a));
    }
    kprev = a.kout;
    vprev = a.vout;
  }
  if (alias1) {
    if (keys_out) {
      hipLaunchKernelGGL(k_copy_pairs, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream,
                         static_cast<const uint64_t *>(kprev), vprev, n, kout, vout);
      ALN_TRY(check_launch("radix_copy"));
    } else {
      ALN_HIP(hipMemcpyAsync(vout, vprev, 4ull * n, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  return ALN_OK;
}

int radix_sort_pairs(aln_ctx *c, uint64_t *d_keys, uint32_t *d_vals, uint64_t n, uint32_t bits) {
  if (n <= 1 || bits == 0) return ALN_OK;
  RsIn in;
  memset(&in, 0, sizeof(in));
  in.keys = d_keys;
  in.vals = d_vals;
  return rs_sort(c, in, false, n, bits, d_keys, d_vals);
}

// Range of every key column: one min/max pass each.
struct KeyRanges {
  int64_t mn[ALN_MAX_KEYS], mx[ALN_MAX_KEYS];
  uint32_t bits[ALN_MAX_KEYS], total;
};

static int key_ranges(aln_ctx *c, const int64_t *const *keys, int nkeys, uint64_t n, KeyRanges *r) {
  r->total = 0;
  for (int k = 0; k < nkeys; ++k) {
    ALN_TRY(minmax_i64(c, keys[k], n, &r->mn[k], &r->mx[k]));
    r->bits[k] = bits_for_range_u((uint64_t)r->mx[k] - (uint64_t)r->mn[k]);
    r->total += r->bits[k];
  }
  return ALN_OK;
}

// The code layout of keys [k0, k1) (k0 most significant).  Returns the total bits.
static uint32_t key_set(const KeyRanges &r, const int64_t *const *keys, const int32_t *desc, int k0,
                        int k1, KeySet *ks) {
  memset(ks, 0, sizeof(*ks));
  uint32_t total = 0;
  for (int k = k0; k < k1; ++k) {
    ks->ref[k - k0] = (desc && desc[k]) ? r.mx[k] : r.mn[k];
    ks->desc[k - k0] = desc ? desc[k] : 0;
    ks->k[k - k0] = keys[k];
    total += r.bits[k];
  }
  ks->n = k1 - k0;
  uint32_t sh = total;
  for (int k = 0; k < ks->n; ++k) {
    sh -= r.bits[k0 + k];
    ks->shift[k] = sh;
  }
  return total;
}

static int sort_with_ranges(aln_ctx *c, const KeyRanges &r, const int64_t *const *keys,
                            const int32_t *desc, int nkeys, uint64_t n, uint32_t *d_perm) {
  RsIn in;
  memset(&in, 0, sizeof(in));
  if (r.total <= 64) {  // one sort of the concatenated codes, computed by its first pass
    key_set(r, keys, desc, 0, nkeys, &in.ks);
    return rs_sort(c, in, true, n, r.total, nullptr, d_perm);
  }
  // LSD over keys, last key first, each sort stable (sorts.cu:17-70); each reads its key at
  // the rows of the permutation so far
  for (int k = nkeys - 1; k >= 0; --k) {
    uint32_t b = key_set(r, keys, desc, k, k + 1, &in.ks);
    in.vals = k == nkeys - 1 ? nullptr : d_perm;
    ALN_TRY(rs_sort(c, in, true, n, b, nullptr, d_perm));
  }
  return ALN_OK;
}

int sort_keys_impl(aln_ctx *c, const int64_t *const *keys, const int32_t *desc, int nkeys,
                   uint64_t n, uint32_t *d_perm) {
  ALN_REQUIRE(nkeys >= 1 && nkeys <= ALN_MAX_KEYS, "sort: %d keys", nkeys);
  ALN_REQUIRE(n < (1ull << 32), "sort: more than 2^32 rows");
  if (n == 0) return ALN_OK;
  KeyRanges r;
  bool whole = false;
  if (nkeys == 1 && n >= RS_SYNC_MIN) {
    // One key: codes over the whole int64 range, no min/max pass; the count pass finds the
    // digits every key shares and those passes are skipped.  Unless values of a narrow range
    // differ in many high bits (they cross zero or another power of two): checked on the
    // first rows, where such keys get the min/max pass and codes of just their range.
    std::vector<int64_t> v(std::min<uint64_t>(n, RS_SAMPLE));
    ALN_TRY(copy_to_host(c, v.data(), keys[0], 8 * v.size()));
    uint64_t vary = 0;
    int64_t lo = v[0], hi = v[0];
    for (int64_t x : v) {
      vary |= (uint64_t)x ^ (uint64_t)v[0];
      lo = std::min(lo, x);
      hi = std::max(hi, x);
    }
    whole = bits_for_range_u(vary) <= bits_for_range_u((uint64_t)hi - (uint64_t)lo) + RS_BITS;
  }
  if (whole) {
    r.mn[0] = INT64_MIN;
    r.mx[0] = INT64_MAX;
    r.bits[0] = r.total = 64;
  } else {
    ALN_TRY(key_ranges(c, keys, nkeys, n, &r));
  }
  return sort_with_ranges(c, r, keys, desc, nkeys, n, d_perm);
}

// ---- top-k: tile-wise bitonic sort of (code, index), keep the first k of each tile -----
constexpr int TK_BLOCK = 256;
constexpr int TK_TILE = 1024;

__device__ __forceinline__ bool tk_less(uint64_t ca, uint32_t ia, uint64_t cb, uint32_t ib) {
  return ca < cb || (ca == cb && ia < ib);
}

__global__ void __launch_bounds__(TK_BLOCK) k_tile_topk(const uint64_t *__restrict__ code,
                                                        const uint32_t *__restrict__ idx,
                                                        uint64_t n, uint32_t k,
                                                        uint64_t *__restrict__ ocode,
                                                        uint32_t *__restrict__ oidx) {
  __shared__ uint64_t sc[TK_TILE];
  __shared__ uint32_t si[TK_TILE];
  uint64_t base = (uint64_t)blockIdx.x * TK_TILE;
  for (int t = threadIdx.x; t < TK_TILE; t += TK_BLOCK) {
    uint64_t i = base + t;
    sc[t] = i < n ? code[i] : ~0ull;
    si[t] = i < n ? idx[i] : 0xFFFFFFFFu;
  }
  __syncthreads();
  for (int size = 2; size <= TK_TILE; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < TK_TILE / 2; t += TK_BLOCK) {
        int lo = 2 * t - (t & (stride - 1));
        int hi = lo + stride;
        bool up = (lo & size) == 0;
        uint64_t a = sc[lo], b = sc[hi];
        uint32_t ia = si[lo], ib = si[hi];
        bool swap = up ? tk_less(b, ib, a, ia) : tk_less(a, ia, b, ib);
        if (swap) {
          sc[lo] = b;
          sc[hi] = a;
          si[lo] = ib;
          si[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  uint64_t obase = (uint64_t)blockIdx.x * k;
  for (uint32_t t = threadIdx.x; t < k; t += TK_BLOCK) {
    ocode[obase + t] = sc[t];
    oidx[obase + t] = si[t];
  }
}

int topk_codes(aln_ctx *c, uint64_t *code, uint32_t *idx, uint64_t n, uint32_t k,
               uint32_t *d_out_idx, uint32_t *h_nout) {
  // repeatedly reduce each tile to its k best until one tile remains
  const uint64_t n_real = n;
  while (true) {
    uint64_t nt = (n + TK_TILE - 1) / TK_TILE;
    uint64_t *oc;
    uint32_t *oi;
    ALN_TRY(ws(c, nt * k, &oc));
    ALN_TRY(ws(c, nt * k, &oi));
    {
      ProfScope ps(c, "tile_topk");
      hipLaunchKernelGGL(k_tile_topk, dim3((unsigned)nt), dim3(TK_BLOCK), 0, c->stream, code, idx, n,
                         k, oc, oi);
    }
    ALN_TRY(check_launch("tile_topk"));
    if (nt == 1) {
      uint32_t kk = (uint32_t)(n_real < k ? n_real : k);
      ALN_HIP(hipMemcpyAsync(d_out_idx, oi, 4ull * kk, hipMemcpyDeviceToDevice, c->stream));
      ALN_HIP(hipStreamSynchronize(c->stream));
      *h_nout = kk;
      return ALN_OK;
    }
    code = oc;
    idx = oi;
    n = nt * k;  // padding entries (~0, 0xFFFFFFFF) sort last and are never among the k best
  }
}

}  // namespace aln

using namespace aln;

extern "C" {

int aln_radix_sort_pairs(aln_ctx *c, uint64_t *d_keys, uint32_t *d_vals, uint64_t n,
                         uint32_t bits) {
  ALN_REQUIRE(c && (n == 0 || (d_keys && d_vals)), "aln_radix_sort_pairs: bad arguments");
  ws_reset(c);
  ALN_TRY(radix_sort_pairs(c, d_keys, d_vals, n, bits));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_sort_keys(aln_ctx *c, const int64_t *const *d_keys, const int32_t *desc, int nkeys,
                  uint64_t n, uint32_t *d_perm) {
  ALN_REQUIRE(c && d_keys && (n == 0 || d_perm), "aln_sort_keys: bad arguments");
  ws_reset(c);
  ALN_TRY(sort_keys_impl(c, d_keys, desc, nkeys, n, d_perm));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_topk(aln_ctx *c, const int64_t *const *d_keys, const int32_t *desc, int nkeys, uint64_t n,
             uint32_t k, uint32_t *d_out_idx, uint32_t *h_nout) {
  ALN_REQUIRE(c && d_keys && h_nout && nkeys >= 1 && nkeys <= ALN_MAX_KEYS,
              "aln_topk: bad arguments");
  ALN_REQUIRE(n < (1ull << 32), "aln_topk: more than 2^32 rows");
  ws_reset(c);
  *h_nout = 0;
  if (n == 0 || k == 0) return ALN_OK;
  ALN_REQUIRE(d_out_idx, "aln_topk: no output buffer");
  KeyRanges r;
  ALN_TRY(key_ranges(c, d_keys, nkeys, n, &r));
  if (r.total > 64 || k > TK_TILE / 2) {
    // wide keys or large k: full stable sort, then the prefix (emit_order + LIMIT)
    uint32_t *perm;
    ALN_TRY(ws(c, n, &perm));
    ALN_TRY(sort_with_ranges(c, r, d_keys, desc, nkeys, n, perm));
    uint32_t kk = (uint32_t)(n < k ? n : k);
    ALN_HIP(hipMemcpyAsync(d_out_idx, perm, 4ull * kk, hipMemcpyDeviceToDevice, c->stream));
    ALN_HIP(hipStreamSynchronize(c->stream));
    *h_nout = kk;
    return ALN_OK;
  }
  uint64_t *code;
  uint32_t *idx;
  ALN_TRY(ws(c, n, &code));
  ALN_TRY(ws(c, n, &idx));
  KeySet ks;
  key_set(r, d_keys, desc, 0, nkeys, &ks);
  hipLaunchKernelGGL(k_codes, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, ks,
                     (const uint32_t *)nullptr, n, code, idx);
  ALN_TRY(check_launch("codes"));
  return topk_codes(c, code, idx, n, k, d_out_idx, h_nout);
}

}  // extern "C"

// Tile staging for scans over the bit-packed layout.  A block handles a tile of 4096 rows
// of one segment; the packed words of up to two columns the tile reads for every row
// (predicate and key columns) are copied into LDS with coalesced dword loads issued up
// front, then decoded from LDS.  Columns wider than 31 bits, and columns read only for the
// few rows that survive a filter or join, are decoded straight from global memory.
#pragma once

#include "aln_internal.h"

namespace aln {

constexpr int STG_BLOCK = 256;
constexpr int STG_ITEMS = 16;
constexpr int STG_TILE = STG_BLOCK * STG_ITEMS;
constexpr int STG_SLOTS = 2;
constexpr int STG_MAX_BITS = 31;
// Words a slot can hold: the 16-byte blocks covering one tile's codes at the widest staged
// width (ceil(4096 * 31 / 32) + 1 words, plus up to 3 words of the first block before them).
// (+4: stage_code reads the word after a code's first word even when the code ends inside it)
constexpr int STG_SLOT_WORDS = ((STG_TILE * STG_MAX_BITS + 31) / 32 + 1 + 3 + 3) / 4 * 4 + 4;

struct StageSlot {
  int64_t base;
  uint32_t bits;
  uint32_t bit0;  // bit offset of the tile's first code inside the slot's staged words
};

struct __align__(16) TileStage {
  uint32_t words[STG_SLOTS][STG_SLOT_WORDS];
  StageSlot slot[STG_SLOTS];
  int8_t slot_of[ALN_MAX_COLS];
};

__device__ __forceinline__ int64_t stage_get(const uint32_t *lds, const StageSlot &s, uint32_t i) {
  if (s.bits == 0) return s.base;
  const uint32_t pos = s.bit0 + i * s.bits;
  const uint32_t wi = pos >> 5, sh = pos & 31;
  uint64_t lo = lds[wi];
  if (sh + s.bits > 32) lo |= (uint64_t)lds[wi + 1] << 32;
  return (int64_t)((uint64_t)s.base + ((lo >> sh) & ((1ull << s.bits) - 1)));
}

// Raw code (value - base) of row i of a staged slot: one funnel shift of two LDS words.
__device__ __forceinline__ uint32_t stage_code(const uint32_t *lds, const StageSlot &s, uint32_t i) {
  const uint32_t pos = s.bit0 + i * s.bits;
  const uint32_t wi = pos >> 5;
  return __builtin_amdgcn_alignbit(lds[wi + 1], lds[wi], pos & 31) & ((1u << s.bits) - 1u);
}

// A range lo <= v <= hi on a staged column, restated on the tile's codes: v - base in
// [clo, clo + cspan], tested as one unsigned compare (code - clo <= cspan); `none` when no
// code of the tile's width can fall inside.
struct CodeRange {
  uint32_t clo, cspan;
  bool none;
};
__device__ __forceinline__ CodeRange code_range(int64_t base, uint32_t bits, int64_t lo, int64_t hi) {
  CodeRange r;
  const uint64_t maxc = bits ? (1ull << bits) - 1 : 0;
  const uint64_t L = lo <= base ? 0 : (uint64_t)lo - (uint64_t)base;
  const bool below = hi < base;
  const uint64_t H = below ? 0 : ((uint64_t)hi - (uint64_t)base < maxc ? (uint64_t)hi - (uint64_t)base : maxc);
  r.none = below || L > H;
  r.clo = r.none ? 0 : (uint32_t)L;
  r.cspan = r.none ? 0 : (uint32_t)(H - L);
  return r;
}
__device__ __forceinline__ CodeRange code_range(const StageSlot &s, int64_t lo, int64_t hi) {
  return code_range(s.base, s.bits, lo, hi);
}

// Stages the listed columns (first STG_SLOTS with at most 31 bits); every thread of the
// block must call it, followed by __syncthreads().  Each slot receives the aligned 16-byte
// blocks that cover the tile's codes (an aligned block never crosses a page, so the few
// words outside the column's range it may carry are readable), copied with all of a thread's
// independent 16-byte loads (8 with 256 threads) in flight before the first store: the block keeps
// tens of KiB of loads outstanding instead of one dword per thread.
template <int BLOCK = STG_BLOCK>
__device__ __forceinline__ void stage_tile(TileStage &ts, const uint32_t *const *words,
                                           const aln_seg *const *segs, const int *want, int nwant,
                                           uint32_t seg, uint32_t r0, uint32_t nt) {
  constexpr int VEC = (2 * (STG_SLOT_WORDS / 4) + BLOCK - 1) / BLOCK;  // loads per thread
  if (threadIdx.x < ALN_MAX_COLS) ts.slot_of[threadIdx.x] = -1;
  __syncthreads();
  static_assert(STG_SLOTS == 2, "the copy loop below addresses two slots");
  // slot sources as scalars: a private array indexed by the slot number would live in
  // scratch memory
  const uint4 *src0 = nullptr, *src1 = nullptr;
  uint32_t n0 = 0, n1 = 0;
  int used = 0;
  for (int k = 0; k < nwant && used < STG_SLOTS; ++k) {
    const int c = want[k];
    bool dup = false;
    for (int q = 0; q < k; ++q) dup = dup || want[q] == c;
    if (dup) continue;
    const aln_seg sg = segs[c][seg];
    if (sg.bits > STG_MAX_BITS) continue;
    const uint64_t b0 = (uint64_t)r0 * sg.bits;
    const uint64_t wbeg = b0 >> 5, wend = ((uint64_t)(r0 + nt) * sg.bits + 31) >> 5;
    const uint32_t *p = words[c] + sg.word_off + wbeg;
    const uint32_t a = (uint32_t)(((uintptr_t)p >> 2) & 3);  // words of the first block before p
    const uint32_t nb = sg.bits ? (uint32_t)((a + (wend - wbeg) + 3) >> 2) : 0;
    if (used == 0) {
      src0 = (const uint4 *)(p - a);
      n0 = nb;
    } else {
      src1 = (const uint4 *)(p - a);
      n1 = nb;
    }
    if (threadIdx.x == 0) {
      ts.slot[used].base = sg.base;
      ts.slot[used].bits = sg.bits;
      ts.slot[used].bit0 = (uint32_t)(b0 & 31) + 32 * a;
      ts.slot_of[c] = (int8_t)used;
    }
    ++used;
  }
  // one index space over the blocks of both slots: v < n0 is slot 0, then slot 1; two
  // slots of at most STG_SLOT_WORDS / 4 blocks fit one round of VEC loads per thread
  static_assert(2 * (STG_SLOT_WORDS / 4) <= VEC * BLOCK, "one round of loads covers both slots");
  uint4 *dst0 = reinterpret_cast<uint4 *>(ts.words[0]), *dst1 = reinterpret_cast<uint4 *>(ts.words[1]);
  const uint32_t ntot = n0 + n1;
  uint4 r[VEC];
#pragma unroll
  for (int k = 0; k < VEC; ++k) {
    const uint32_t v = k * BLOCK + threadIdx.x;
    r[k] = v < n0 ? src0[v] : (v < ntot ? src1[v - n0] : make_uint4(0, 0, 0, 0));
  }
#pragma unroll
  for (int k = 0; k < VEC; ++k) {
    const uint32_t v = k * BLOCK + threadIdx.x;
    if (v < n0)
      dst0[v] = r[k];
    else if (v < ntot)
      dst1[v - n0] = r[k];
  }
}

// Value of column c at row r0 + i of segment seg: from LDS when staged, else from global.
__device__ __forceinline__ int64_t tile_get(const TileStage &ts, const uint32_t *const *words,
                                            const aln_seg *const *segs, int c, uint32_t seg,
                                            uint32_t r0, uint32_t i) {
  const int s = ts.slot_of[c];
  if (s >= 0) return stage_get(ts.words[s], ts.slot[s], i);
  return decode_row(words[c], segs[c][seg], r0 + i);
}

}  // namespace aln

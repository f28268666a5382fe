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
constexpr int STG_SLOT_WORDS = (STG_TILE * STG_MAX_BITS + 31) / 32 + 2;

struct StageSlot {
  int64_t base;
  uint32_t bits;
  uint32_t bit0;  // bit offset of the tile's first code inside the first staged word
};

struct TileStage {
  uint32_t words[STG_SLOTS][STG_SLOT_WORDS];
  StageSlot slot[STG_SLOTS];
  int8_t slot_of[ALN_MAX_COLS];
};

// Copies the words holding rows [r0, r0 + nt) of one segment into lds.
__device__ __forceinline__ void stage_copy(const uint32_t *__restrict__ words, const aln_seg &sg,
                                           uint32_t r0, uint32_t nt, uint32_t *lds) {
  const uint64_t b0 = (uint64_t)r0 * sg.bits;
  const uint64_t wbeg = b0 >> 5, wend = ((uint64_t)(r0 + nt) * sg.bits + 31) >> 5;
  const uint32_t *p = words + sg.word_off + wbeg;
  const uint32_t n = (uint32_t)(wend - wbeg);
  for (uint32_t i = threadIdx.x; i < n; i += STG_BLOCK) lds[i] = p[i];
}

__device__ __forceinline__ int64_t stage_get(const uint32_t *lds, const StageSlot &s, uint32_t i) {
  if (s.bits == 0) return s.base;
  const uint32_t pos = s.bit0 + i * s.bits;
  const uint32_t wi = pos >> 5, sh = pos & 31;
  uint64_t lo = lds[wi];
  if (sh + s.bits > 32) lo |= (uint64_t)lds[wi + 1] << 32;
  return (int64_t)((uint64_t)s.base + ((lo >> sh) & ((1ull << s.bits) - 1)));
}

// Stages the listed columns (first STG_SLOTS with at most 31 bits); every thread of the
// block must call it, followed by __syncthreads().
__device__ __forceinline__ void stage_tile(TileStage &ts, const uint32_t *const *words,
                                           const aln_seg *const *segs, const int *want, int nwant,
                                           uint32_t seg, uint32_t r0, uint32_t nt) {
  if (threadIdx.x < ALN_MAX_COLS) ts.slot_of[threadIdx.x] = -1;
  __syncthreads();
  int used = 0;
  for (int k = 0; k < nwant && used < STG_SLOTS; ++k) {
    const int c = want[k];
    bool dup = false;
    for (int q = 0; q < k; ++q) dup = dup || want[q] == c;
    if (dup) continue;
    const aln_seg sg = segs[c][seg];
    if (sg.bits > STG_MAX_BITS) continue;
    stage_copy(words[c], sg, r0, nt, ts.words[used]);
    if (threadIdx.x == 0) {
      ts.slot[used].base = sg.base;
      ts.slot[used].bits = sg.bits;
      ts.slot[used].bit0 = (uint32_t)(((uint64_t)r0 * sg.bits) & 31);
      ts.slot_of[c] = (int8_t)used;
    }
    ++used;
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

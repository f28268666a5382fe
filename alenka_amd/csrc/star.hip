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

#include <algorithm>
#include <vector>

#include "aln_internal.h"
#include "stage.h"

namespace aln {

// 512 threads x 8 rows per 4096-row tile: eight waves share one tile's 32 KB of staging, so
// the LDS that caps a CU at four tiles still admits 32 waves.
constexpr int ST_BLOCK = 512;
constexpr int ST_ITEMS = 8;
constexpr int ST_TILE = ST_BLOCK * ST_ITEMS;
static_assert(ST_TILE == STG_TILE, "a block stages one tile");
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

// The tile's ranges in code space (ranges on staged columns), built by thread 0 after
// staging; tile_pass_fast then tests a staged range with one funnel shift and one compare.
struct TilePred {
  int8_t slot[ALN_MAX_RANGES];  // staged slot of range k, -1: decode from HBM
  CodeRange cr[ALN_MAX_RANGES];
  bool none;                    // a staged range admits no code of the tile
};

// Block prologue of every scan kernel: stage the range columns (when the zone map says the
// tile needs evaluating) and `key` into LDS, and (tp) the ranges in code space.  Thread 0
// writes the column list and the slot descriptors, so it builds tp without a barrier of its
// own: two barriers in all (the one inside stage_tile and the last one here).
__device__ __forceinline__ void star_stage(TileStage &ts, int *want, const ScanDesc &d, char mode,
                                           int key, uint32_t seg, uint32_t r0, uint32_t nt,
                                           TilePred *tp = nullptr) {
  const int nwant = (mode == 'R' ? d.nr : 0) + (key >= 0 ? 1 : 0);
  if (threadIdx.x == 0) {
    int n = 0;
    if (mode == 'R')
      for (int k = 0; k < d.nr; ++k) want[n++] = d.rcol[k];
    if (key >= 0) want[n++] = key;
  }
  stage_tile<ST_BLOCK>(ts, d.words, d.segs, want, nwant, seg, r0, nt);  // barrier before reading want
  if (tp && threadIdx.x == 0) {
    tp->none = false;
    for (int k = 0; k < d.nr; ++k) {
      const int s = ts.slot_of[d.rcol[k]];
      tp->slot[k] = (int8_t)s;
      if (s >= 0) {
        tp->cr[k] = code_range(ts.slot[s], d.rlo[k], d.rhi[k]);
        tp->none = tp->none || tp->cr[k].none;
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ bool tile_pass_fast(const ScanDesc &d, const TileStage &ts, const TilePred &tp,
                                               char mode, uint32_t seg, uint32_t r0, uint32_t i) {
  if (mode == 'A') return true;
  if (tp.none) return false;
  bool p = true;
  for (int k = 0; k < d.nr; ++k) {
    const int s = tp.slot[k];
    if (s >= 0) {
      p = p && (stage_code(ts.words[s], ts.slot[s], i) - tp.cr[k].clo <= tp.cr[k].cspan);
    } else {
      const int64_t v = decode_row(d.words[d.rcol[k]], d.segs[d.rcol[k]][seg], r0 + i);
      p = p && v >= d.rlo[k] && v <= d.rhi[k];
    }
  }
  return p;
}

// The common tile shape - at most one range, on a staged column, and a staged key - decoded
// with the slot parameters held in registers (no descriptor reads from LDS per row).
struct FastTile {
  const uint32_t *rw, *kw;
  uint32_t rbit0, rbits, rmask, clo, cspan, kbit0, kbits, kmask;
  bool range;
};

__device__ __forceinline__ bool fast_tile(const ScanDesc &d, const TileStage &ts, const TilePred &tp, char mode,
                                          int keycol, FastTile &f) {
  const int ks = ts.slot_of[keycol];
  if (ks < 0 || ts.slot[ks].bits > 31) return false;
  f.range = mode == 'R';
  if (f.range) {
    if (d.nr != 1 || tp.slot[0] < 0 || tp.none) return false;
    const int rs = tp.slot[0];
    f.rw = ts.words[rs];
    f.rbit0 = ts.slot[rs].bit0;
    f.rbits = ts.slot[rs].bits;
    f.rmask = (1u << f.rbits) - 1u;
    f.clo = tp.cr[0].clo;
    f.cspan = tp.cr[0].cspan;
  }
  f.kw = ts.words[ks];
  f.kbit0 = ts.slot[ks].bit0;
  f.kbits = ts.slot[ks].bits;
  f.kmask = (1u << f.kbits) - 1u;
  return true;
}

__device__ __forceinline__ uint32_t fast_code(const uint32_t *w, uint32_t bit0, uint32_t bits, uint32_t mask,
                                              uint32_t i) {
  const uint32_t pos = bit0 + i * bits;
  const uint32_t wi = pos >> 5;
  return __builtin_amdgcn_alignbit(w[wi + 1], w[wi], pos & 31) & mask;
}

// Straight-line decode of a tile's rows for the fast shape: per row, the range test (RANGE)
// and the key shifted by `off` in 32 bits; code[j] = key code, or ~0 when the row does not
// probe (fails the range, lies past the tile, or the key falls outside [0, lim)).  Row
// j * 512 + t sits 512 * bits bits (16 * bits words) after row (j - 1) * 512 + t, so a
// thread's rows share one funnel-shift amount and step through LDS by a constant.  Every
// read stays inside the slots, which span a whole tile.  Returns the wave's probe count.
template <bool RANGE, bool FULL>
__device__ __forceinline__ uint32_t fast_probe_t(const FastTile &f, uint32_t nt, uint32_t off, uint32_t lim,
                                                 uint32_t (&code)[ST_ITEMS]) {
  const uint32_t kp0 = f.kbit0 + threadIdx.x * f.kbits, rp0 = f.rbit0 + threadIdx.x * f.rbits;
  const uint32_t *kw = f.kw + (kp0 >> 5), *rw = f.rw + (rp0 >> 5);
  const uint32_t ksh = kp0 & 31, rsh = rp0 & 31, kstep = 16 * f.kbits, rstep = 16 * f.rbits;
  uint32_t n = 0;
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    // the probing rows as a lane mask (FULL: a whole tile, every row exists), turned back
    // into the lane's condition with inverse_ballot so that it stays a mask
    uint64_t pm = FULL ? ~0ull : __builtin_amdgcn_ballot_w64(j * ST_BLOCK + threadIdx.x < nt);
    if (RANGE) {
      const uint32_t *q = rw + j * rstep;
      const uint32_t rc = __builtin_amdgcn_alignbit(q[1], q[0], rsh) & f.rmask;
      pm &= __builtin_amdgcn_ballot_w64(rc - f.clo <= f.cspan);
    }
    const uint32_t *q = kw + j * kstep;
    const uint32_t c = (__builtin_amdgcn_alignbit(q[1], q[0], ksh) & f.kmask) + off;
    n += (uint32_t)__popcll(pm);
    code[j] = (__builtin_amdgcn_inverse_ballot_w64(pm) & (c < lim)) ? c : 0xFFFFFFFFu;
  }
  return n;
}
template <bool RANGE>
__device__ __forceinline__ uint32_t fast_probe(const FastTile &f, uint32_t nt, uint32_t off, uint32_t lim,
                                               uint32_t (&code)[ST_ITEMS]) {
  return nt == (uint32_t)ST_TILE ? fast_probe_t<RANGE, true>(f, nt, off, lim, code)
                                 : fast_probe_t<RANGE, false>(f, nt, off, lim, code);
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

// Exclusive scan, in place, of the ST_ITEMS x (ST_BLOCK / 64) per-(item, wave) counts of a
// tile (one entry per lane of the first wave); the first wave calls it, *total gets the sum.
static_assert(ST_ITEMS * (ST_BLOCK / 64) == 64, "one wave scans the tile's counts");
__device__ __forceinline__ void scan_tile_counts(uint32_t *counts, uint32_t *total) {
  const unsigned lane = __lane_id();
  const uint32_t n = counts[lane];
  uint32_t x = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  counts[lane] = x - n;
  if (lane == 63) *total = x;
}

// ---- per-segment scan plans ------------------------------------------------------------------
// What a tile needs from its segment's descriptors and the query's range, resolved once per
// segment by k_scan_plan.  A block reads its segment's record with scalar loads and derives
// its staging addresses from it, so no descriptor read sits between the block's start and
// its first data load, and no thread builds per-tile state for the others behind a barrier.
// Roles: 0 the range column (staged in mode 'R'), 1 the probe key, 2..4 per kernel.
constexpr int PL_ROLES = 5;
constexpr int PL_OWN = 3;                // staged columns per tile
constexpr int PL_STAGE_WORDS = 9216;     // up to 72 staged bits per row over a 4096-row tile
constexpr int PL_VEC = (PL_STAGE_WORDS / 4 + ST_BLOCK - 1) / ST_BLOCK;  // 16-byte loads per thread
constexpr int PL_WAVE_ROWS = 64 * ST_ITEMS;  // rows of a tile one wave owns (item-major)

struct PlanCol {
  const uint32_t *w;  // the segment's packed words
  int64_t base;
  uint32_t bits;
  int32_t slot;       // LDS word offset of the column's staged blocks, -1: not staged
};

struct StageSrc {     // one staged column
  const uint32_t *w;
  uint32_t bits;      // 0: unused entry
  int32_t slot;
};

struct __align__(16) SegPlan {
  PlanCol col[PL_ROLES];
  StageSrc st[PL_OWN];
  uint32_t nrows;
  uint32_t clo, cspan;  // mode 'R': a row passes when code(col[0]) - clo <= cspan
  int32_t zmode;        // the zone map's verdict: 'N', 'A' or 'R'
  int32_t mode;         // 'N' skip, 'A' every row passes, 'R' one staged range, 'G' generic path
};

// LDS words a staged column of `bits` needs: the 16-byte blocks covering a tile's codes
// (ceil(4096 * bits / 32) + 1 words plus up to 3 words of the first block before them), and
// 4 words after them for the funnel-shift read of the last code.
__host__ __device__ constexpr uint32_t plan_slot_words(uint32_t bits) {
  return ((ST_TILE * bits + 31) / 32 + 1 + 3 + 3) / 4 * 4 + 4;
}
static_assert(plan_slot_words(STG_MAX_BITS) == STG_SLOT_WORDS, "slot sizes agree with stage.h");

// One table's plan request.  rc[1..4]: the columns of roles 1..4 (-1: none); roles whose bit
// is set in `stage` are staged when their column is not constant.  kmin: the probe key's
// code is key - kmin, and the fast path needs base - kmin in 32 bits.  fast = 0 plans every
// live segment 'G'; a segment needing more than max_own (<= PL_OWN) staged columns is 'G'.
struct PlanJob {
  int rc[PL_ROLES];
  uint32_t stage;
  int64_t kmin;
  int fast;
  uint32_t max_own, nsegs;
  SegPlan *out;
};

__device__ void plan_segment(const ScanDesc &d, const PlanJob &j, uint32_t seg) {
  SegPlan P;
  const char zm = sd_mode(d, seg);
  const int rc1 = j.rc[1], stage = (int)j.stage;
  const int64_t kmin = j.kmin;
  const int rc[PL_ROLES] = {d.nr == 1 ? d.rcol[0] : -1, j.rc[1], j.rc[2], j.rc[3], j.rc[4]};
  bool generic = !j.fast || d.nr > 1;
  uint32_t words = 0, nown = 0;
#pragma unroll
  for (int k = 0; k < PL_OWN; ++k) P.st[k] = StageSrc{nullptr, 0u, 0};
#pragma unroll
  for (int r = 0; r < PL_ROLES; ++r) {
    PlanCol &pc = P.col[r];
    pc = PlanCol{nullptr, 0, 0u, -1};
    const int c = rc[r];
    if (c < 0) continue;
    const aln_seg sg = d.segs[c][seg];
    pc.w = d.words[c] + sg.word_off;
    pc.base = sg.base;
    pc.bits = sg.bits;
    const bool want = r == 0 ? zm == 'R' : (((stage >> r) & 1) && sg.bits > 0);
    if (!want) continue;
    if (sg.bits > (uint32_t)STG_MAX_BITS || sg.bits == 0) {
      generic = true;
      continue;
    }
    int32_t slot = -1;
#pragma unroll
    for (int q = 0; q < r; ++q)
      if (rc[q] == c && P.col[q].slot >= 0) slot = P.col[q].slot;
    if (slot < 0) {
      slot = (int32_t)words;
      words += plan_slot_words(sg.bits);
#pragma unroll
      for (int k = 0; k < PL_OWN; ++k)
        if (nown == (uint32_t)k) P.st[k] = StageSrc{pc.w, sg.bits, slot};
      ++nown;
    }
    pc.slot = slot;
  }
  if (words > (uint32_t)PL_STAGE_WORDS || nown > j.max_own) generic = true;
  // the fast probe reads a staged, non-constant key and shifts it in 32 bits
  if (rc1 >= 0) {
    const int64_t off = (int64_t)((uint64_t)P.col[1].base - (uint64_t)kmin);
    if (P.col[1].slot < 0 || off < INT32_MIN || off > INT32_MAX) generic = true;
  }
  P.nrows = d.segs[0][seg].nrows;
  P.clo = P.cspan = 0;
  P.zmode = zm;
  char mode = zm == 'N' ? 'N' : (generic ? 'G' : zm);
  if (mode == 'R') {
    const CodeRange cr = code_range(P.col[0].base, P.col[0].bits, d.rlo[0], d.rhi[0]);
    if (cr.none) mode = 'N';
    P.clo = cr.clo;
    P.cspan = cr.cspan;
  }
  P.mode = mode;
  j.out[seg] = P;
}

// Zero fills of up to 4 word arrays, with the plans, in one launch.
struct FillJob {
  uint32_t *p[4];
  uint64_t n[4];
};

// Per-query setup of the star pipeline in one launch: the first plan_blocks blocks plan the
// segments of the three scans (one thread per segment), the others zero the listed arrays.
__global__ void __launch_bounds__(128) k_star_setup(ScanDesc d0, ScanDesc d1, ScanDesc d2, PlanJob j0, PlanJob j1,
                                                    PlanJob j2, uint32_t plan_blocks, FillJob f) {
  if (blockIdx.x < plan_blocks) {
    uint32_t i = blockIdx.x * 128 + threadIdx.x;
    if (i < j0.nsegs) {
      plan_segment(d0, j0, i);
    } else if ((i -= j0.nsegs) < j1.nsegs) {
      plan_segment(d1, j1, i);
    } else if ((i -= j1.nsegs) < j2.nsegs) {
      plan_segment(d2, j2, i);
    }
    return;
  }
  const uint64_t t0 = (uint64_t)(blockIdx.x - plan_blocks) * 128 + threadIdx.x;
  const uint64_t stride = (uint64_t)(gridDim.x - plan_blocks) * 128;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (uint64_t i = t0; i < f.n[k]; i += stride) f.p[k][i] = 0;
}

// A pointer read from a plan, marked as global memory: the compiler cannot tell where a
// pointer loaded from memory points and would otherwise use flat (generic) loads.
template <class T>
__device__ __forceinline__ T gload(const T *p) {
#if __HIP_DEVICE_COMPILE__
  return *(const __attribute__((address_space(1))) T *)p;
#else
  return *p;  // host pass: never executed
#endif
}
// uint4 is a class type whose copy would go through a generic reference: load the words as
// a plain vector instead
__device__ __forceinline__ uint4 gload4(const uint4 *p) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  const v4u q = gload(reinterpret_cast<const v4u *>(p));
  return make_uint4(q.x, q.y, q.z, q.w);
}

// The same load, nontemporal: for staging, whose words are read once (streamed past the
// caches, the mid scan runs ~3 % faster than with ordinary loads).
__device__ __forceinline__ uint4 gload4_nt(const uint4 *p) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
#if __HIP_DEVICE_COMPILE__
  const v4u q = __builtin_nontemporal_load((const __attribute__((address_space(1))) v4u *)(const v4u *)p);
#else
  const v4u q = *reinterpret_cast<const v4u *>(p);  // host pass: never executed
#endif
  return make_uint4(q.x, q.y, q.z, q.w);
}

// Reads single rows of a segment's column from global memory without branches: the three
// word loads always issue (a word a row does not need is read from the row's first word,
// and a constant column reads `any`, some readable global address), so the loads of several
// readers go out together.
struct ColReader {
  const uint32_t *w;
  int64_t base;
  uint32_t bits;
  uint64_t mask;
};
__device__ __forceinline__ ColReader col_reader(const PlanCol &pc, const void *any) {
  const bool c = pc.bits == 0;
  return ColReader{c ? static_cast<const uint32_t *>(any) : pc.w, pc.base, pc.bits,
                   pc.bits >= 64 ? ~0ull : (1ull << pc.bits) - 1};
}
__device__ __forceinline__ int64_t col_read(const ColReader &r, uint64_t local) {
  const uint32_t *w = r.w;
  const uint64_t pos = local * r.bits, wi = pos >> 5;
  const uint32_t sh = (uint32_t)(pos & 31);
  const bool two = sh + r.bits > 32, three = sh + r.bits > 64;
  uint64_t i1 = two ? wi + 1 : wi, i2 = three ? wi + 2 : wi;
  // opaque to the optimiser, which would otherwise turn the loads of words that may equal
  // word 0 back into conditional loads, each behind a branch and a wait
  asm volatile("" : "+v"(i1), "+v"(i2));
  const uint32_t w0 = gload(w + wi), w1 = gload(w + i1), w2 = gload(w + i2);
  const uint64_t v = ((((uint64_t)w1 << 32) | w0) >> sh) | (((uint64_t)w2 << ((64 - sh) & 63)) & (three ? ~0ull : 0ull));
  return (int64_t)((uint64_t)r.base + (v & r.mask));
}
// Value of a column at row `local` of the segment, from global memory.
__device__ __forceinline__ int64_t plan_get(const PlanCol &pc, uint64_t local) {
  if (pc.bits == 0) return pc.base;
  return col_read(col_reader(pc, pc.w), local);
}

// A staged column of the tile starting at row r0: its LDS words and the bit offset of row r0.
struct PlanSlot {
  const uint32_t *w;
  uint32_t bit0, bits, mask;
};
__device__ __forceinline__ PlanSlot plan_slot(const uint32_t *lds, const PlanCol &pc, uint32_t r0) {
  const uint64_t b0 = (uint64_t)r0 * pc.bits;
  const uint32_t a = (uint32_t)(((uintptr_t)(pc.w + (b0 >> 5)) >> 2) & 3);
  return PlanSlot{lds + (pc.slot >= 0 ? pc.slot : 0), (uint32_t)(b0 & 31) + 32 * a, pc.bits,
                  (1u << pc.bits) - 1u};
}
__device__ __forceinline__ uint32_t slot_code(const PlanSlot &s, uint32_t i) {
  const uint32_t pos = s.bit0 + i * s.bits, wi = pos >> 5;
  return __builtin_amdgcn_alignbit(s.w[wi + 1], s.w[wi], pos & 31) & s.mask;
}

// The plan's staged columns for rows [r0, r0 + nt), in two steps: issue() loads the aligned
// 16-byte blocks covering each column's codes into registers (an aligned block never
// crosses a page, so the words outside the column it may carry are readable), all of a
// thread's loads (PL_CVEC per column) in flight together; store() copies them into LDS.
// Every thread calls both; a barrier must follow store().  NOWN: staged columns the plan
// may hold.
constexpr int PL_CVEC = (STG_SLOT_WORDS / 4 + ST_BLOCK - 1) / ST_BLOCK;  // loads per column
template <int NOWN>
struct StageRegs {
  uint4 r[NOWN][PL_CVEC];
  uint32_t nb[NOWN];
  int32_t slot[NOWN];
  __device__ __forceinline__ void issue(const SegPlan &P, uint32_t r0, uint32_t nt) {
#pragma unroll
    for (int k = 0; k < NOWN; ++k) {
      const StageSrc S = P.st[k];
      const uint64_t b0 = (uint64_t)r0 * S.bits;
      const uint64_t wbeg = b0 >> 5, wend = ((uint64_t)(r0 + nt) * S.bits + 31) >> 5;
      const uint32_t *p = S.w + wbeg;
      const uint32_t a = (uint32_t)(((uintptr_t)p >> 2) & 3);
      const uint4 *src = reinterpret_cast<const uint4 *>(p - a);
      nb[k] = S.bits ? (uint32_t)((a + (wend - wbeg) + 3) >> 2) : 0u;
      slot[k] = S.slot / 4;
#pragma unroll
      for (int t = 0; t < PL_CVEC; ++t) {
        const uint32_t v = t * ST_BLOCK + threadIdx.x;
        if (v < nb[k]) r[k][t] = gload4_nt(src + v);  // (store() skips the rest)
      }
    }
  }
  __device__ __forceinline__ void store(uint32_t *lds) const {
    uint4 *l4 = reinterpret_cast<uint4 *>(lds);
#pragma unroll
    for (int k = 0; k < NOWN; ++k)
#pragma unroll
      for (int t = 0; t < PL_CVEC; ++t) {
        const uint32_t v = t * ST_BLOCK + threadIdx.x;
        if (v < nb[k]) l4[slot[k] + (int32_t)v] = r[k][t];
      }
  }
};
template <int NOWN = PL_OWN>
__device__ __forceinline__ void stage_plan(uint32_t *lds, const SegPlan &P, uint32_t r0, uint32_t nt) {
  StageRegs<NOWN> s;
  s.issue(P, r0, nt);
  s.store(lds);
}

// The fast decoder's view of the range (role 0, mode 'R') and probe-key (role 1) slots.
__device__ __forceinline__ FastTile plan_fast(const uint32_t *lds, const SegPlan &P, uint32_t r0) {
  FastTile f;
  f.range = P.mode == 'R';
  const PlanSlot k = plan_slot(lds, P.col[1], r0);
  f.kw = k.w;
  f.kbit0 = k.bit0;
  f.kbits = k.bits;
  f.kmask = k.mask;
  f.rw = k.w;
  f.rbit0 = f.rbits = f.rmask = f.clo = f.cspan = 0;
  if (f.range) {
    const PlanSlot q = plan_slot(lds, P.col[0], r0);
    f.rw = q.w;
    f.rbit0 = q.bit0;
    f.rbits = q.bits;
    f.rmask = q.mask;
    f.clo = P.clo;
    f.cspan = P.cspan;
  }
  return f;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- D: dimension scan into a key bitmap --------------------------------------------------
// Per-tile counts go to arrays indexed by block (summed afterwards): a counter shared by
// every block would serialise its atomics across the XCDs.
// Two per-block counts with one reduction (two barriers).
__device__ __forceinline__ void block_count2(uint32_t a, uint32_t *out_a, uint32_t b, uint32_t *out_b) {
  __shared__ uint32_t wa[ST_BLOCK / 64], wb[ST_BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  __syncthreads();
  if (__lane_id() == 0) {
    wa[threadIdx.x / 64] = a;
    wb[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ta = 0, tb = 0;
    for (int w = 0; w < ST_BLOCK / 64; ++w) {
      ta += wa[w];
      tb += wb[w];
    }
    out_a[blockIdx.x] = ta;
    out_b[blockIdx.x] = tb;
  }
}

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

// Plan roles: 0 range, 1 key (staged).  Passing keys set their bit without a returned value;
// a key seen twice shows as popcount(bitmap) < passing rows.  Counts are per wave:
// npass[b * 8 + w].
constexpr uint32_t DIM_STAGE = 1u << 1;
constexpr int DIM_OWN = 2;

__global__ void __launch_bounds__(ST_BLOCK, 8) k_star_dim(ScanDesc d, const SegPlan *__restrict__ plan, int kcol,
                                                          int64_t kmin, uint64_t range, uint32_t *__restrict__ bitmap,
                                                          uint32_t *__restrict__ npass) {
  static_assert(sizeof(TileStage) <= PL_STAGE_WORDS * sizeof(uint32_t), "generic staging fits the LDS area");
  __shared__ __align__(16) uint32_t lds[PL_STAGE_WORDS];
  __shared__ int want[ALN_MAX_RANGES + 1];
  __shared__ TilePred tp;
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const SegPlan &P = plan[seg];
  const uint32_t r0 = tile * ST_TILE, nrows = P.nrows;
  const uint32_t region = blockIdx.x * (ST_BLOCK / 64) + threadIdx.x / 64;
  const int mode = P.mode;
  if (r0 >= nrows || mode == 'N') {
    if (__lane_id() == 0) npass[region] = 0;
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  uint32_t n = 0;
  if (mode != 'G') {
    stage_plan<DIM_OWN>(lds, P, r0, nt);
    __syncthreads();
    const FastTile ft = plan_fast(lds, P, r0);
    const uint32_t off = (uint32_t)(int32_t)((uint64_t)P.col[1].base - (uint64_t)kmin), lim = (uint32_t)range;
    // each thread takes ST_ITEMS consecutive rows: keys of adjacent rows are often adjacent
    // (a dimension table is usually stored in key order), so a thread ORs the bits of its
    // rows that fall in one bitmap word itself and sets them with one atomic per word
    uint32_t cw = 0xFFFFFFFFu, cb = 0;
#pragma unroll
    for (int u = 0; u < ST_ITEMS; ++u) {
      const uint32_t i = threadIdx.x * ST_ITEMS + u;
      bool p = i < nt;
      if (ft.range) {
        const uint32_t rp = ft.rbit0 + i * ft.rbits;
        const uint32_t *q = ft.rw + (rp >> 5);
        p = p & (((__builtin_amdgcn_alignbit(q[1], q[0], rp & 31) & ft.rmask) - ft.clo) <= ft.cspan);
      }
      const uint32_t kp = ft.kbit0 + i * ft.kbits;
      const uint32_t *q = ft.kw + (kp >> 5);
      const uint32_t c = (__builtin_amdgcn_alignbit(q[1], q[0], kp & 31) & ft.kmask) + off;
      const bool v = p & (c < lim);
      n += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(v));
      if (v) {
        if ((c >> 5) != cw) {
          if (cb) atomicOr(&bitmap[cw], cb);
          cw = c >> 5;
          cb = 0;
        }
        cb |= 1u << (c & 31);
      }
    }
    if (cb) atomicOr(&bitmap[cw], cb);
  } else {
    TileStage &ts = *reinterpret_cast<TileStage *>(lds);
    const char zm = (char)P.zmode;
    star_stage(ts, want, d, zm, kcol, seg, r0, nt, &tp);
    for (int j = 0; j < ST_ITEMS; ++j) {
      const uint32_t i = j * ST_BLOCK + threadIdx.x;
      if (i < nt && tile_pass_fast(d, ts, tp, zm, seg, r0, i)) {
        const uint64_t b = (uint64_t)tile_get(ts, d.words, d.segs, kcol, seg, r0, i) - (uint64_t)kmin;
        if (b < range) {
          atomicOr(&bitmap[b >> 5], 1u << (b & 31));
          ++n;
        }
      }
    }
    n = wave_sum(n);
  }
  if (__lane_id() == 0) npass[region] = n;
}

// Sets the bits of the block's passing mid keys (code[j] != ~0) in the mid bitmap: keys of
// neighbouring rows are usually close (a table stored in key order), so the block ORs them
// into an LDS window of words over [its smallest word, its largest] first and then ORs the
// window's non-zero words into the bitmap, consecutive words in one wave instruction;
// keys beyond the window's PL_STAGE_WORDS words go to the bitmap directly.  The bits are
// set without a returned value: a duplicate key shows as popcount(bitmap) < passing rows.
// Reuses the tile's staging area (`win`), so every read of it must be done.
// Wave minimum / maximum through DPP row shifts and row broadcasts (lane 63 ends with the
// whole wave's; lanes a step has no source for take the neutral value), read from lane 63.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_minmax_u32(uint32_t v) {
  constexpr int NEUTRAL = MAX ? 0 : -1;
  auto step = [&](uint32_t x) __attribute__((always_inline)) { v = MAX ? max(v, x) : min(v, x); };
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x111, 0xF, 0xF, false));  // row_shr:1
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x112, 0xF, 0xF, false));  // row_shr:2
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x114, 0xF, 0xF, false));  // row_shr:4
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x118, 0xF, 0xF, false));  // row_shr:8
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
  step((uint32_t)__builtin_amdgcn_update_dpp(NEUTRAL, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ void mid_bits_window(uint32_t *win, const uint32_t *code, uint32_t *__restrict__ mbits,
                                                uint2 *__restrict__ trange) {
  __shared__ uint32_t s_lo, s_hi;
  if (threadIdx.x == 0) {
    s_lo = 0xFFFFFFFFu;
    s_hi = 0;
  }
  // the lowest code, and the highest plus one (an absent code, ~0, is the largest value
  // and wraps to 0 plus one), then their words
  uint32_t lo = 0xFFFFFFFFu, hi1 = 0;
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    lo = min(lo, code[j]);
    hi1 = max(hi1, code[j] + 1u);
  }
  const uint32_t clo = wave_minmax_u32<false>(lo), hi = (wave_minmax_u32<true>(hi1) - 1u) >> 5;
  lo = clo >> 5;
  __syncthreads();  // s_lo / s_hi set; the staging area is no longer read
  if (__lane_id() == 0 && clo != 0xFFFFFFFFu) {  // (a code is never ~0)
    atomicMin(&s_lo, lo);
    atomicMax(&s_hi, hi);
  }
  __syncthreads();
  const uint32_t wlo = s_lo, whi = s_hi;
  if (threadIdx.x == 0) trange[blockIdx.x] = make_uint2(wlo, whi);  // (empty: wlo > whi)
  if (wlo > whi) return;  // no passing row in the block (uniform)
  const uint32_t span = whi - wlo + 1 < (uint32_t)PL_STAGE_WORDS ? whi - wlo + 1 : (uint32_t)PL_STAGE_WORDS;
  for (uint32_t w = threadIdx.x; w < span; w += ST_BLOCK) win[w] = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j)
    if (code[j] != 0xFFFFFFFFu) {
      const uint32_t wi = (code[j] >> 5) - wlo, bit = 1u << (code[j] & 31);
      if (wi < span)
        atomicOr(&win[wi], bit);
      else
        atomicOr(&mbits[code[j] >> 5], bit);
    }
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < span; w += ST_BLOCK) {
    const uint32_t v = win[w];
    if (v) atomicOr(&mbits[wlo + w], v);
  }
}

// ---- M: mid scan, probe the dimension bitmap, compact passing rows ------------------------
// Plan roles: 0 range, 1 foreign key (probes the dimension bitmap, staged), 2 own key
// (staged), 3 and 4 the group columns (decoded later, for the groups that need them, from
// the row).  Each passing row leaves {key - mkmin, row}: wave w of tile b writes region
// b * 8 + w (PL_WAVE_ROWS entries) and its count to counts[b * 8 + w].
constexpr uint32_t MID_STAGE = (1u << 1) | (1u << 2);

__global__ void __launch_bounds__(ST_BLOCK, 8) k_star_mid(ScanDesc d, const SegPlan *__restrict__ plan, int kcol,
                                                       int fkcol, const uint32_t *__restrict__ bitmap,
                                                       int64_t dkmin, uint64_t drange,
                                                       uint32_t *__restrict__ counts,
                                                       uint32_t *__restrict__ probed, uint2 *__restrict__ rent,
                                                       uint32_t *__restrict__ mbits, int64_t mkmin,
                                                       uint64_t mrange, uint2 *__restrict__ trange) {
  static_assert(sizeof(TileStage) <= PL_STAGE_WORDS * sizeof(uint32_t), "generic staging fits the LDS area");
  __shared__ __align__(16) uint32_t lds[PL_STAGE_WORDS];
  __shared__ int want[ALN_MAX_RANGES + 1];
  __shared__ TilePred tp;
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const SegPlan &P = plan[seg];
  const uint32_t r0 = tile * ST_TILE, nrows = P.nrows;
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  const uint32_t region = blockIdx.x * (ST_BLOCK / 64) + wid;
  const int mode = P.mode;
  if (r0 >= nrows || mode == 'N') {
    if (lane == 0) {
      counts[region] = 0;
      probed[region] = 0;
    }
    if (mbits && threadIdx.x == 0) trange[blockIdx.x] = make_uint2(0xFFFFFFFFu, 0u);  // no passing row
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  TileStage &ts = *reinterpret_cast<TileStage *>(lds);
  const bool fast = mode != 'G';
  uint64_t dc[ST_ITEMS];
  uint32_t nprobe = 0;
  if (fast) {
    stage_plan(lds, P, r0, nt);
    __syncthreads();
    const FastTile ft = plan_fast(lds, P, r0);
    const uint32_t off = (uint32_t)(int32_t)((uint64_t)P.col[1].base - (uint64_t)dkmin), dr = (uint32_t)drange;
    uint32_t c32[ST_ITEMS];
    nprobe = ft.range ? fast_probe<true>(ft, nt, off, dr, c32) : fast_probe<false>(ft, nt, off, dr, c32);
#pragma unroll
    for (int j = 0; j < ST_ITEMS; ++j) dc[j] = c32[j] != 0xFFFFFFFFu ? c32[j] : ~0ull;
  } else {  // per-lane count, summed over the wave below
    const char zm = (char)P.zmode;
    star_stage(ts, want, d, zm, fkcol, seg, r0, nt, &tp);
    const int fs = ts.slot_of[fkcol];
    const int64_t foff = fs >= 0 ? (int64_t)((uint64_t)ts.slot[fs].base - (uint64_t)dkmin) : 0;
#pragma unroll
    for (int j = 0; j < ST_ITEMS; ++j) {
      const uint32_t i = j * ST_BLOCK + threadIdx.x;
      dc[j] = ~0ull;
      if (i < nt && tile_pass_fast(d, ts, tp, zm, seg, r0, i)) {
        ++nprobe;
        const uint64_t b = fs >= 0 ? (uint64_t)((int64_t)stage_code(ts.words[fs], ts.slot[fs], i) + foff)
                                   : (uint64_t)tile_get(ts, d.words, d.segs, fkcol, seg, r0, i) - (uint64_t)dkmin;
        if (b < drange) dc[j] = b;
      }
    }
    nprobe = wave_sum(nprobe);
  }
  // the probes in 32 bits: word index (~0: no probe; < 2^29 for any dimension range) and bit
  uint32_t pw[ST_ITEMS], dw[ST_ITEMS];
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) pw[j] = dc[j] != ~0ull ? (uint32_t)(dc[j] >> 5) : 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {  // only probing lanes load (fewer addresses per load)
    dw[j] = 0;
    if (pw[j] != 0xFFFFFFFFu) dw[j] = gload(bitmap + pw[j]);
  }
  // the passing rows' keys: staged (fast; code = kbase + the slot's code, its bit position
  // a per-lane start plus a per-item step) or decoded (generic)
  const PlanSlot ks = plan_slot(lds, P.col[2], r0);
  const uint32_t kbase = (uint32_t)((uint64_t)P.col[2].base - (uint64_t)mkmin);
  const uint32_t kp0 = ks.bit0 + threadIdx.x * ks.bits, kstep = ST_BLOCK * ks.bits;
  const uint32_t row0 = seg * d.seg_rows + r0;
  // passing rows of the wave in item-major order (consecutive lanes write consecutive
  // entries); no barrier: each wave owns its region
  uint2 *const reg = rent + (uint64_t)region * PL_WAVE_ROWS;
  uint32_t nout = 0;
  uint32_t mcode[ST_ITEMS];
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    // the passing rows as a lane mask (kept a mask, not a lane value): the bit test alone,
    // as a lane that does not probe holds a zero word
    const uint64_t b = __builtin_amdgcn_ballot_w64(((dw[j] >> ((uint32_t)dc[j] & 31)) & 1) != 0);
    mcode[j] = 0xFFFFFFFFu;
    uint32_t code = 0;  // key - mkmin (key range < 2^32)
    if (__builtin_amdgcn_inverse_ballot_w64(b)) {
      if (fast) {
        const uint32_t pos = kp0 + j * kstep, wi = pos >> 5;
        code = kbase + (__builtin_amdgcn_alignbit(ks.w[wi + 1], ks.w[wi], pos & 31) & ks.mask);
      } else {
        code = (uint32_t)((uint64_t)tile_get(ts, d.words, d.segs, kcol, seg, r0, j * ST_BLOCK + threadIdx.x) -
                          (uint64_t)mkmin);
      }
    }
    // a key outside [mkmin, mkmin + mrange) can only come from ranges the caller reused
    // after the columns changed (the run is then repeated): it is left out, so that no
    // bitmap or group index ever leaves its array
    const uint64_t b2 = __builtin_amdgcn_ballot_w64(__builtin_amdgcn_inverse_ballot_w64(b) && code < mrange);
    if (__builtin_amdgcn_inverse_ballot_w64(b2)) {
      mcode[j] = code;
      reg[nout + below_count(b2)] = make_uint2(code, row0 + j * ST_BLOCK + threadIdx.x);
    }
    nout += (uint32_t)__popcll(b2);
  }
  if (lane == 0) {
    counts[region] = nout;
    probed[region] = nprobe;
  }
  if (mbits) mid_bits_window(lds, mcode, mbits, trange);
}

__device__ __forceinline__ uint64_t st_slot(uint32_t code, uint64_t cap) {
  return __umul64hi(murmur64a_u64(code, ST_SEED), cap);
}

// ---- B: build the join table over the compacted mid rows ----------------------------------
// One block per mid tile, one wave per wave region: entry i of region b becomes group
// offs[b] + i; its mid row goes to grow[group], its key into the table, and its
// accumulators are zeroed.
__global__ void __launch_bounds__(ST_BLOCK) k_star_build(const uint2 *__restrict__ rent,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint64_t *__restrict__ offs, uint64_t *table,
                                                    uint64_t cap, uint32_t *__restrict__ grow,
                                                    unsigned long long *__restrict__ sum,
                                                    unsigned int *__restrict__ cnt, uint32_t *dup) {
  const uint32_t region = blockIdx.x * (ST_BLOCK / 64) + threadIdx.x / 64;
  const uint32_t n = counts[region];
  for (uint32_t i = __lane_id(); i < n; i += 64) {
    const uint64_t j = offs[region] + i;
    const uint2 e = rent[(uint64_t)region * PL_WAVE_ROWS + i];
    grow[j] = e.y;
    sum[j] = 0;  // every group index below the passing-row count is written exactly once
    cnt[j] = 0;
    const uint32_t code = e.x;
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
  block_count2(nmatch, matched, nprobe, probed);
}

// ---- bitmap + rank join (dense key ranges) -----------------------------------------------
// The passing mid keys set bits of a bitmap over [mkmin, mkmax]; a directory of set-bit
// counts per 512-bit block (one 64-byte line) makes rank(key) - the key's position among
// the passing keys - cost one directory read plus one line of the bitmap.  The group index
// is that rank, so groups come out in key order, and probes from a tile of fact rows that
// are clustered on the key touch a few contiguous lines instead of random slots.
constexpr int BM_BLOCK_WORDS = 16;  // 512 bits

// Set bits per 512-bit block (cnt) and, when rel != nullptr, a rank9 sub-directory (Vigna,
// "Broadword implementation of rank/select queries", 2008): for the block's 64-bit words
// i = 1..7, the set bits of words 0..i-1 as a 9-bit field at bit 9 (i - 1) of rel[b].  With
// the exclusive scan of cnt, rank(code) = scan[b] + field(word) + popcount of the bits below
// code in its 64-bit word: three 8-byte reads.
__global__ void __launch_bounds__(256) k_bm_block_counts(const uint32_t *__restrict__ bits,
                                                        uint64_t nwords, uint64_t nblk,
                                                        uint32_t *__restrict__ cnt,
                                                        uint64_t *__restrict__ rel = nullptr) {
  for (uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x; b < nblk;
       b += (uint64_t)gridDim.x * 256) {
    // the bitmap is allocated (and zero-filled) in whole blocks: one 64-byte line
    const uint4 *q = (const uint4 *)(bits + b * BM_BLOCK_WORDS);
    const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const uint32_t p[8] = {__popc(q0.x) + __popc(q0.y), __popc(q0.z) + __popc(q0.w), __popc(q1.x) + __popc(q1.y),
                           __popc(q1.z) + __popc(q1.w), __popc(q2.x) + __popc(q2.y), __popc(q2.z) + __popc(q2.w),
                           __popc(q3.x) + __popc(q3.y), __popc(q3.z) + __popc(q3.w)};
    uint32_t c = p[0];
    uint64_t r = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
      r |= (uint64_t)c << (9 * (i - 1));
      c += p[i];
    }
    cnt[b] = c;
    if (rel) rel[b] = r;
  }
}

// rank9 lookup (see k_bm_block_counts): the number of set bits below `code`.
__device__ __forceinline__ uint32_t bm_rank9(const uint32_t *__restrict__ bits, const uint64_t *__restrict__ rank,
                                             const uint64_t *__restrict__ rel, uint32_t code) {
  const uint32_t blk = code >> 9, w = (code >> 6) & 7;
  const uint64_t word = reinterpret_cast<const uint64_t *>(bits)[code >> 6], r = rel[blk];
  // word 0 shifts by 63: bit 63 of rel is never set (7 fields of 9 bits), so the field is 0;
  // selecting the shift rather than the value keeps the load of rel unconditional
  const uint32_t sub = (uint32_t)(r >> (w ? 9 * w - 9 : 63)) & 511u;
  return (uint32_t)rank[blk] + sub + (uint32_t)__popcll(word & ((1ull << (code & 63)) - 1ull));
}

// Rank of a set bit: the block's directory entry plus the bits below it in its 512-bit block,
// read as four independent 16-byte loads (the bitmap is allocated in whole blocks).
__device__ __forceinline__ uint64_t bm_rank(const uint32_t *__restrict__ bits,
                                            const uint64_t *__restrict__ rank, uint64_t code) {
  const uint64_t blk = code >> 9;
  const uint4 *q = (const uint4 *)(bits + blk * BM_BLOCK_WORDS);
  const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
  const uint32_t w[BM_BLOCK_WORDS] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                      q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const uint32_t wi = (uint32_t)(code >> 5) & (BM_BLOCK_WORDS - 1), sh = (uint32_t)code & 31;
  uint64_t r = rank[blk];
#pragma unroll
  for (uint32_t k = 0; k < (uint32_t)BM_BLOCK_WORDS; ++k)
    r += __popc(w[k] & (k < wi ? 0xFFFFFFFFu : (k == wi ? (1u << sh) - 1u : 0u)));
  return r;
}

// Code (bit position) of set bit number g of the mid bitmap, by one wave: a 64-way search of
// the rank directory (rank[b] = set bits before 512-bit block b, non-decreasing) for the
// last block with rank[b] <= g, then the block's 16 words.  Every lane returns the code.
__device__ uint32_t bm_select_wave(const uint32_t *__restrict__ bits, const uint64_t *__restrict__ rank,
                                   uint64_t nblk, uint32_t g) {
  const unsigned lane = __lane_id();
  uint64_t lo = 0, hi = nblk;  // rank[lo] <= g throughout (rank[0] = 0)
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 63) / 64, i = lo + lane * step;
    const uint64_t b = __builtin_amdgcn_ballot_w64(i < hi && rank[i] <= g);  // a prefix of the lanes
    lo += (uint64_t)(63 - __builtin_clzll(b)) * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  const uint32_t r = g - (uint32_t)rank[lo];  // the bit's number inside block lo
  const uint32_t w = lane < (unsigned)BM_BLOCK_WORDS ? bits[lo * BM_BLOCK_WORDS + lane] : 0u, pc = __popc(w);
  uint32_t x = pc;  // inclusive scan of the words' counts
  for (int o = 1; o < BM_BLOCK_WORDS; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  const uint64_t hb = __builtin_amdgcn_ballot_w64(lane < (unsigned)BM_BLOCK_WORDS && x > r);
  const int wl = __builtin_ctzll(hb);  // the word holding it
  uint32_t v = __shfl(w, wl, 64);
  for (uint32_t k = r - __shfl(x - pc, wl, 64); k; --k) v &= v - 1;  // drop the set bits below it
  return (uint32_t)(lo * 512 + (uint64_t)wl * 32 + (uint32_t)__builtin_ctz(v));
}

// The passing mid rows as k_star_mid leaves them, with the mid bitmap's rank directory.
struct MidRows {
  const uint2 *rent;       // PL_WAVE_ROWS entries {code, row} per mid wave region
  const uint32_t *counts;  // entries of each region
  uint64_t ntiles;         // mid tiles (ST_BLOCK / 64 regions each)
  const uint32_t *bits;    // mid bitmap, its rank9 directory and its block count
  const uint64_t *rank, *rel;
  uint64_t nblk;
  const uint2 *trange;     // per mid tile: the bitmap words its passing keys span (or null)
  uint32_t *grow;          // out: the mid row of group g
};

// Candidate mode: the top-k's candidate groups cj[0 .. n-1], n = min(*count, cap), and their
// codes ccode[i] (k_topk_codes).  Up to TR_MAX of them go into an LDS hash in each block;
// more, or full != 0: every group's row instead.
constexpr uint32_t TR_BITS = 12, TR_HASH = 1u << TR_BITS, TR_MAX = TR_HASH / 2;

// ccode[i] = code of candidate group cj[i], one wave per candidate (nothing when n > TR_MAX).
__global__ void __launch_bounds__(256) k_topk_codes(MidRows M, const uint32_t *__restrict__ count, uint32_t cap,
                                                   const uint32_t *__restrict__ cj, uint32_t *__restrict__ ccode) {
  const uint32_t n = *count < cap ? *count : cap, i = blockIdx.x * 4 + threadIdx.x / 64;
  if (n > TR_MAX || i >= n) return;  // (wave-uniform)
  const uint32_t code = bm_select_wave(M.bits, M.rank, M.nblk, cj[i]);
  if (__lane_id() == 0) ccode[i] = code;
}

// grow[g] for the groups that need it, from the passing mid rows: a block takes ST_BLOCK / 64
// mid tiles, one per wave, whose regions it reads as one list, TR_Q entries per lane with all
// their reads in flight (the first batch issued before the block builds its hash).
// Candidate mode: a row whose code is in the hash is its candidate's row.  Full mode: every
// row's group is the rank of its code.
constexpr int TR_Q = 8;
__global__ void __launch_bounds__(ST_BLOCK) k_topk_rows(MidRows M, const uint32_t *__restrict__ count, uint32_t cap,
                                                       const uint32_t *__restrict__ cj,
                                                       const uint32_t *__restrict__ ccode, int full) {
  __shared__ uint32_t hk[TR_HASH], hg[TR_HASH], cc[TR_MAX];
  const uint32_t n = full ? 0u : (*count < cap ? *count : cap);
  const bool all = full || n > TR_MAX;
  constexpr int R = ST_BLOCK / 64;
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  const uint64_t tile = (uint64_t)blockIdx.x * R + wid;
  bool mine = tile < M.ntiles;
  // candidate mode with tile ranges: a tile whose passing keys' words hold no candidate's
  // key is skipped (a table stored in key order leaves a few tiles to read)
  const uint2 tr = mine && !all && M.trange ? M.trange[tile] : make_uint2(0u, 0xFFFFFFFFu);
  uint32_t off[R + 1];
  off[0] = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) off[r + 1] = off[r] + (mine ? M.counts[tile * R + r] : 0u);
  const uint32_t total = off[R];
  uint2 e[TR_Q];
  auto load = [&](uint32_t base) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < TR_Q; ++q) {
      const uint32_t idx = base + q * 64 + lane;
      uint32_t r = 0;
#pragma unroll
      for (int k = 1; k < R; ++k) r += idx >= off[k];
      e[q] = make_uint2(0xFFFFFFFFu, 0u);
      if (idx < total) e[q] = M.rent[(tile * R + r) * PL_WAVE_ROWS + (idx - off[r])];
    }
  };
  if (all) load(0);  // (candidate mode: after the tile test)
  if (!all) {
    for (uint32_t q = threadIdx.x; q < TR_HASH; q += ST_BLOCK) hk[q] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += ST_BLOCK) {  // n <= TR_MAX: a free slot exists
      const uint32_t g = cj[i], code = ccode[i];
      cc[i] = code;
      uint32_t q = (code * 0x9E3779B1u) >> (32 - TR_BITS);
      while (true) {
        const uint32_t prev = atomicCAS(&hk[q], 0xFFFFFFFFu, code);
        if (prev == 0xFFFFFFFFu) {
          hg[q] = g;
          break;
        }
        if (prev == code) break;  // (distinct groups have distinct codes)
        q = (q + 1) & (TR_HASH - 1);
      }
    }
    __syncthreads();
    bool any = false;
    if (mine && tr.x <= tr.y)
      for (uint32_t i = lane; i < n; i += 64) any |= (cc[i] >> 5) - tr.x <= tr.y - tr.x;
    mine = mine && __builtin_amdgcn_ballot_w64(any) != 0;
    if (!mine) return;  // (wave-uniform; no barrier follows)
    load(0);
  }
  for (uint32_t base = 0; base < total; base += 64 * TR_Q) {
    if (base > 0) load(base);
    if (all) {
      uint32_t g[TR_Q];
#pragma unroll
      for (int q = 0; q < TR_Q; ++q) g[q] = bm_rank9(M.bits, M.rank, M.rel, e[q].x != 0xFFFFFFFFu ? e[q].x : 0u);
#pragma unroll
      for (int q = 0; q < TR_Q; ++q)
        if (e[q].x != 0xFFFFFFFFu) M.grow[g[q]] = e[q].y;
    } else {
#pragma unroll
      for (int q = 0; q < TR_Q; ++q) {
        if (e[q].x == 0xFFFFFFFFu) continue;
        uint32_t h = (e[q].x * 0x9E3779B1u) >> (32 - TR_BITS);
        while (true) {  // the table holds at most half its slots: an empty one ends the probe
          const uint32_t k = hk[h];
          if (k == e[q].x) {
            M.grow[hg[h]] = e[q].y;
            break;
          }
          if (k == 0xFFFFFFFFu) break;
          h = (h + 1) & (TR_HASH - 1);
        }
      }
    }
  }
}

// SUM and COUNT of groups 0 .. *d_m - 1 and their bits in held zeroed (the arrays are
// allocated for more groups).
__global__ void __launch_bounds__(256) k_zero_groups(unsigned long long *__restrict__ sum, unsigned int *__restrict__ cnt,
                                                     uint32_t *__restrict__ held,
                                                     const unsigned long long *__restrict__ d_m) {
  const uint64_t m = *d_m, stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += stride) {
    sum[i] = 0;
    cnt[i] = 0;
    if (i < (m + 31) / 32) held[i] = 0;
  }
}

// Plan roles: 0 range, 1 foreign key (probes the mid bitmap), 2 and 3 the aggregate's x
// and y, read from global memory for the few rows that join.  Counts are per wave
// (index b * 8 + w), so the block needs one barrier, after staging.
constexpr uint32_t FACT_STAGE = 1u << 1;
constexpr int FACT_OWN = 2;  // staged columns: the range and the key

// The fact scan's per-query constants.
struct FactArgs {
  int fkcol;
  int64_t a0, a1, b0, b1;
  const uint32_t *bits;  // mid bitmap and its rank9 directory
  const uint64_t *rank, *rel;
  int64_t mkmin;
  uint64_t mrange;
  unsigned long long *sum;
  unsigned int *cnt;
  uint32_t *matched, *probed;  // per wave: index tile * 8 + wave
  uint2 *hitbuf;               // per wave FACT_HIT_CAP slots {code, row in segment}, or null
  uint32_t *held;              // bit g set once group g holds a row (or null)
};
// A wave's first FACT_HIT_CAP joined rows are left in its slots for k_fact_resolve (the
// scan does not wait on their payload reads); the rest (rare) are resolved in place.
constexpr uint32_t FACT_HIT_CAP = 16;

// Predicate and key of a staged fast tile (after the barrier that follows staging), as codes
// into the mid bitmap (~0: no probe); returns the wave's probe count.
__device__ __forceinline__ uint32_t fact_fast_codes(const uint32_t *lds, const SegPlan &P, const FactArgs &A,
                                                    uint32_t r0, uint32_t nt, uint32_t (&code)[ST_ITEMS]) {
  const FastTile ft = plan_fast(lds, P, r0);
  // code + off lies in (-2^31, 2^32): any value outside [0, mrange), mrange <= 2^31, maps to
  // >= 2^31 in 32 bits
  const uint32_t off = (uint32_t)(int32_t)((uint64_t)P.col[1].base - (uint64_t)A.mkmin), mr = (uint32_t)A.mrange;
  return ft.range ? fast_probe<true>(ft, nt, off, mr, code) : fast_probe<false>(ft, nt, off, mr, code);
}

// The same for a tile the plan could not take the fast way ('G'): generic staging (with its
// own barriers) and decoding; returns the wave's probe count.
__device__ __forceinline__ uint32_t fact_generic_codes(uint32_t *lds, int *want, TilePred &tp, const ScanDesc &d,
                                                       const SegPlan &P, const FactArgs &A, uint32_t seg,
                                                       uint32_t r0, uint32_t nt, uint32_t (&code)[ST_ITEMS]) {
  TileStage &ts = *reinterpret_cast<TileStage *>(lds);
  const char zm = (char)P.zmode;
  star_stage(ts, want, d, zm, A.fkcol, seg, r0, nt, &tp);
  const int fs = ts.slot_of[A.fkcol];
  const int64_t foff = fs >= 0 ? (int64_t)((uint64_t)ts.slot[fs].base - (uint64_t)A.mkmin) : 0;
  uint32_t nprobe = 0;
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    const uint32_t i = j * ST_BLOCK + threadIdx.x;
    code[j] = 0xFFFFFFFFu;
    if (i < nt && tile_pass_fast(d, ts, tp, zm, seg, r0, i)) {
      ++nprobe;
      const uint64_t c = fs >= 0 ? (uint64_t)((int64_t)stage_code(ts.words[fs], ts.slot[fs], i) + foff)
                                 : (uint64_t)tile_get(ts, d.words, d.segs, A.fkcol, seg, r0, i) - (uint64_t)A.mkmin;
      if (c < A.mrange) code[j] = (uint32_t)c;  // mrange <= 2^31 on this path
    }
  }
  return wave_sum(nprobe);
}

// Probe of the mid bitmap for a tile's codes, then one hit per lane per round (rarely more
// than one round): its group index (the rank of the key), payload words and the two
// atomics, the loads of a round issued together; lane 0 writes the wave's counts.
__device__ __forceinline__ void fact_probe_hits(const SegPlan &P, const void *any, const FactArgs &A, uint32_t r0,
                                                const uint32_t (&code)[ST_ITEMS], uint32_t nprobe, uint32_t cidx) {
  // the bitmap words, loaded by the probing lanes only (fewer addresses per load)
  uint32_t word[ST_ITEMS];
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    word[j] = 0;
    if (code[j] != 0xFFFFFFFFu) word[j] = gload(A.bits + (code[j] >> 5));
  }
  uint32_t hits = 0, nmatch = 0;
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    // the joined rows as a lane mask (kept a mask, not a lane value): the bit test alone,
    // as a lane that does not probe holds a zero word
    const uint64_t b = __builtin_amdgcn_ballot_w64(((word[j] >> (code[j] & 31)) & 1) != 0);
    if (b == 0) continue;  // (wave-uniform) no joined row in this row of the wave: most of them
    const bool h = __builtin_amdgcn_inverse_ballot_w64(b);
    if (A.hitbuf) {  // slot of this row among the wave's joined rows; the first slots are deferred
      const uint32_t pos = nmatch + below_count(b);
      if (h && pos < FACT_HIT_CAP) {
        A.hitbuf[(uint64_t)cidx * FACT_HIT_CAP + pos] = make_uint2(code[j], r0 + j * ST_BLOCK + threadIdx.x);
      } else {
        hits |= (uint32_t)h << j;
      }
    } else {
      hits |= (uint32_t)h << j;
    }
    nmatch += (uint32_t)__popcll(b);
  }
  const ColReader rx = col_reader(P.col[2], any), ry = col_reader(P.col[3], any);
  while (hits) {
    const int jh = __builtin_ctz(hits);
    hits &= hits - 1;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < ST_ITEMS; ++j) c = j == jh ? code[j] : c;
    const uint32_t local = r0 + (uint32_t)jh * ST_BLOCK + threadIdx.x;
    const uint32_t g = bm_rank9(A.bits, A.rank, A.rel, c);
    const int64_t x = col_read(rx, local), y = col_read(ry, local);
    const int64_t v = wmul((int64_t)((uint64_t)A.a0 + (uint64_t)wmul(A.a1, x)),
                           (int64_t)((uint64_t)A.b0 + (uint64_t)wmul(A.b1, y)));
    atomicAdd(&A.sum[g], (unsigned long long)v);
    atomicAdd(&A.cnt[g], 1u);
    if (A.held) atomicOr(&A.held[g >> 5], 1u << (g & 31));
  }
  if (__lane_id() == 0) {
    A.matched[cidx] = nmatch;
    A.probed[cidx] = nprobe;
  }
}

// The deferred joined rows: one wave per RS_TILES adjacent fact tiles of a segment.  Their
// deferred rows (the first min(matched[w], FACT_HIT_CAP) slots of each fact wave w) are
// packed onto the lanes, 64 at a time (a group of tiles rarely has more): group index =
// rank of the key, the payload columns of the row, then SUM and COUNT.  Every lane's reads
// are independent, so the payload's HBM latency overlaps across the whole grid instead of
// stalling the scan, and a lane idles only past the group's last row.  The rows of one fact
// wave come in row order, and a group's rows are adjacent (the fact table is clustered on
// the key), so runs of equal group index in adjacent lanes are summed with a segmented wave
// scan first: one atomic pair per run.  (One tile per wave: with three, the kernel took
// 130 us against 125 us on the 600M-row q3, its time being the rows' atomics and reads.)
constexpr uint32_t RS_TILES = 1;
__global__ void __launch_bounds__(64) k_fact_resolve(const uint2 *__restrict__ hitbuf, uint64_t nsegs,
                                                     uint32_t tps, const SegPlan *__restrict__ plan, FactArgs A) {
  constexpr int NW = ST_BLOCK / 64;  // fact waves per tile
  constexpr int NK = NW * RS_TILES;  // fact waves per resolve wave
  const uint32_t gps = (tps + RS_TILES - 1) / RS_TILES;
  const uint64_t seg = blockIdx.x / gps;
  if (seg >= nsegs) return;
  const uint32_t t0 = (blockIdx.x % gps) * RS_TILES;
  const uint32_t nt = tps - t0 < RS_TILES ? tps - t0 : RS_TILES;  // tiles of this group
  const uint64_t w0 = (seg * tps + t0) * NW;                      // its first fact wave
  // deferred rows before each fact wave of the group (wave-uniform loads)
  uint32_t pre[NK + 1];
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    const uint32_t m = (uint32_t)k < nt * NW ? A.matched[w0 + k] : 0u;
    pre[k + 1] = pre[k] + (m < FACT_HIT_CAP ? m : FACT_HIT_CAP);
  }
  const uint32_t total = pre[NK];
  if (total == 0) return;
  const SegPlan &P = plan[seg];
  const ColReader rx = col_reader(P.col[2], plan), ry = col_reader(P.col[3], plan);
  const unsigned lane = __lane_id();
  for (uint32_t b0 = 0; b0 < total; b0 += 64) {
    // lane q of the round takes deferred row q: slot q - pre[k] of the fact wave k with
    // pre[k] <= q < pre[k + 1]; lanes past the last row load nothing and add nothing
    const uint32_t q = b0 + lane;
    const bool used = q < total;
    uint32_t k = 0, first = 0;
#pragma unroll
    for (int kk = 1; kk < NK; ++kk)
      if (q >= pre[kk]) {
        k = (uint32_t)kk;
        first = pre[kk];
      }
    uint2 e = make_uint2(0u, 0u);
    if (used) e = hitbuf[(w0 + k) * FACT_HIT_CAP + (q - first)];
    const uint32_t code = e.x, row = e.y;
    // only the lanes holding a row load (their rank and payload reads are issued together)
    uint32_t g = 0xFFFFFFFFu;
    int64_t x = 0, y = 0;
    if (used) {
      g = bm_rank9(A.bits, A.rank, A.rel, code);
      x = col_read(rx, row);
      y = col_read(ry, row);
    }
    int64_t v = wmul((int64_t)((uint64_t)A.a0 + (uint64_t)wmul(A.a1, x)),
                     (int64_t)((uint64_t)A.b0 + (uint64_t)wmul(A.b1, y)));
    // segmented inclusive scan over runs of equal g (unused lanes, g = ~0, form their own run)
    uint32_t c = used;
    const uint32_t prev = __shfl_up(g, 1, 64);  // every lane takes part (not inside a short circuit)
    bool f = lane == 0 || prev != g;
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t vu = __shfl_up(v, o, 64);
      const uint32_t cu = __shfl_up(c, o, 64);
      const bool fu = __shfl_up((int)f, o, 64) != 0;
      if (lane >= (unsigned)o && !f) {
        v = (int64_t)((uint64_t)v + (uint64_t)vu);
        c += cu;
      }
      if (lane >= (unsigned)o) f = f || fu;
    }
    const uint32_t next = __shfl_down(g, 1, 64);
    if (used && (lane == 63 || next != g)) {
      atomicAdd(&A.sum[g], (unsigned long long)v);
      atomicAdd(&A.cnt[g], c);
      if (A.held) atomicOr(&A.held[g >> 5], 1u << (g & 31));
    }
  }
}

// One block per tile.  (Second bound: eight waves per SIMD, four blocks per CU, i.e. at
// most 64 VGPRs.)
__global__ void __launch_bounds__(ST_BLOCK, 8) k_star_fact_bm(ScanDesc d, const SegPlan *__restrict__ plan,
                                                           FactArgs A) {
  static_assert(sizeof(TileStage) <= PL_STAGE_WORDS * sizeof(uint32_t), "generic staging fits the LDS area");
  __shared__ __align__(16) uint32_t lds[PL_STAGE_WORDS];
  __shared__ int want[ALN_MAX_RANGES + 1];
  __shared__ TilePred tp;
  const uint32_t seg = blockIdx.x / d.tps, tile = blockIdx.x % d.tps;
  const SegPlan &P = plan[seg];
  const uint32_t r0 = tile * ST_TILE, nrows = P.nrows;
  const uint32_t cidx = blockIdx.x * (ST_BLOCK / 64) + threadIdx.x / 64;
  const int mode = P.mode;
  if (r0 >= nrows || mode == 'N') {
    if (__lane_id() == 0) {
      A.matched[cidx] = 0;
      A.probed[cidx] = 0;
    }
    return;
  }
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  uint32_t code[ST_ITEMS];
  uint32_t nprobe;
  if (mode != 'G') {
    stage_plan<FACT_OWN>(lds, P, r0, nt);
    __syncthreads();
    nprobe = fact_fast_codes(lds, P, A, r0, nt, code);
  } else {
    nprobe = fact_generic_codes(lds, want, tp, d, P, A, seg, r0, nt, code);
  }
  fact_probe_hits(P, plan, A, r0, code, nprobe, cidx);
}

// ---- T: top-k over groups (sum DESC, g1 ASC, key ASC, g2 ASC) -----------------------------
// A group's (key, g1, g2): arrays indexed by group, or (row != nullptr) decoded from the
// group's mid row row[j] through the mid scan's plan (roles 2, 3, 4; a role without a column
// reads 0), for the few groups that need them.
struct GroupAttr {
  const int64_t *key, *g1, *g2;
  const uint32_t *row;
  const SegPlan *plan;
  uint32_t seg_rows;
};
__device__ __forceinline__ int64_t ga_get(const GroupAttr &a, uint64_t j, int which) {
  if (!a.row) return which == 0 ? a.key[j] : (which == 1 ? a.g1[j] : a.g2[j]);
  const uint32_t r = a.row[j], seg = r / a.seg_rows;
  return plan_get(a.plan[seg].col[2 + which], r - seg * a.seg_rows);
}
__device__ __forceinline__ int64_t ga_key(const GroupAttr &a, uint64_t j) { return ga_get(a, j, 0); }
__device__ __forceinline__ int64_t ga_g1(const GroupAttr &a, uint64_t j) { return ga_get(a, j, 1); }
__device__ __forceinline__ int64_t ga_g2(const GroupAttr &a, uint64_t j) { return ga_get(a, j, 2); }

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
                                                        GroupAttr ga, uint32_t k, Cand *__restrict__ out) {
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
      c.g1 = ga_g1(ga, i);
      c.key = ga_key(ga, i);
      c.g2 = ga_g2(ga, i);
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
                                   GroupAttr ga, int64_t kmin, uint64_t *code,
                                   uint32_t *idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    if (flag[i]) {
      code[pos[i]] = (uint64_t)ga_key(ga, i) - (uint64_t)kmin;
      idx[pos[i]] = (uint32_t)i;
    }
}

__global__ void k_star_group_emit(const uint32_t *idx, uint64_t g, GroupAttr ga,
                                  const unsigned long long *sum, const unsigned int *cnt,
                                  aln_star_row *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    uint32_t j = idx[i];
    aln_star_row r;
    r.key = ga_key(ga, j);
    r.g1 = ga_g1(ga, j);
    r.g2 = ga_g2(ga, j);
    r.sum = (int64_t)sum[j];
    r.count = cnt[j];
    out[i] = r;
  }
}

// min of vmin / max of vmax over the non-empty segments of each column: one block per column.
struct ColsRange {  // passed by value: no host-to-device copy before the launch
  const aln_seg *segs[8];
  uint32_t nsegs[8];
};
__global__ void k_cols_range(ColsRange cr, int64_t *out) {
  __shared__ int64_t smn[256], smx[256];
  const aln_seg *sg = cr.segs[blockIdx.x];
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint32_t i = threadIdx.x; i < cr.nsegs[blockIdx.x]; i += 256)
    if (sg[i].nrows) {
      mn = sg[i].vmin < mn ? sg[i].vmin : mn;
      mx = sg[i].vmax > mx ? sg[i].vmax : mx;
    }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) {
      smn[threadIdx.x] = smn[threadIdx.x + o] < smn[threadIdx.x] ? smn[threadIdx.x + o] : smn[threadIdx.x];
      smx[threadIdx.x] = smx[threadIdx.x + o] > smx[threadIdx.x] ? smx[threadIdx.x + o] : smx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = smn[0];
    out[2 * blockIdx.x + 1] = smx[0];
  }
}

// Sets *stale when the ranges the kernel found (out, 2 per column) differ from the given ones.
struct RangePair {
  int64_t v[16];
};
__global__ void k_range_check(const int64_t *__restrict__ out, RangePair want, uint32_t n,
                              unsigned long long *__restrict__ stale) {
  const uint32_t i = threadIdx.x;
  if (i < 2 * n && out[i] != want.v[i]) atomicOr(stale, 1ull);
}

// Zone-map ranges of up to 8 columns with one launch and one copy (nullptr entries: 0, 0).
static int cols_range(aln_ctx *c, const aln_col *const *cols, int n, int64_t *mn, int64_t *mx) {
  ALN_REQUIRE(n <= 8, "cols_range: %d columns", n);
  ColsRange cr;
  int used = 0;
  for (int i = 0; i < n; ++i)
    if (cols[i]) {
      cr.segs[used] = cols[i]->segs;
      cr.nsegs[used] = cols[i]->nsegs;
      ++used;
    }
  int64_t res[16];
  if (used) {
    int64_t *dout;
    ALN_TRY(ws(c, 2 * used, &dout));
    hipLaunchKernelGGL(k_cols_range, dim3(used), dim3(256), 0, c->stream, cr, dout);
    ALN_TRY(check_launch("cols_range"));
    ALN_TRY(copy_to_host(c, res, dout, 16 * used));
  }
  for (int i = 0, u = 0; i < n; ++i) {
    mn[i] = cols[i] ? res[2 * u] : 0;
    mx[i] = cols[i] ? res[2 * u + 1] : 0;
    if (cols[i]) ++u;
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

// ---- fast top-k for small limits ------------------------------------------------------
// The order (sum DESC, g1 ASC, key ASC) is a 128-bit key: hi = sum with its sign bit flipped
// and complemented (ascending hi = descending sum), lo = (g1 - g1min) << 32 | (key - kmin).
// Keys are unique per group, so the order is total and g2 never decides.  Every thread keeps
// its best TS_K candidates sorted in registers (unrolled insertion: static indices only); a
// block then merges its threads' lists by k rounds of a block-wide minimum over the list
// heads.  Pass 1 reduces the groups to k per block, pass 2 (one block) to the final k.
constexpr int TS_K = 10;
constexpr int TS_BLOCK = 128;

struct TopList {
  uint64_t hi[TS_K], lo[TS_K];
  uint32_t j[TS_K];
};

__device__ __forceinline__ bool ts_less(uint64_t h, uint64_t l, uint64_t H, uint64_t L) {
  return h < H || (h == H && l < L);
}

__device__ __forceinline__ void ts_insert(TopList &t, uint64_t h, uint64_t l, uint32_t j) {
  if (!ts_less(h, l, t.hi[TS_K - 1], t.lo[TS_K - 1])) return;
#pragma unroll
  for (int k = TS_K - 1; k > 0; --k) {
    if (ts_less(h, l, t.hi[k - 1], t.lo[k - 1])) {
      t.hi[k] = t.hi[k - 1];
      t.lo[k] = t.lo[k - 1];
      t.j[k] = t.j[k - 1];
    } else if (ts_less(h, l, t.hi[k], t.lo[k])) {
      t.hi[k] = h;
      t.lo[k] = l;
      t.j[k] = j;
    }
  }
  if (ts_less(h, l, t.hi[0], t.lo[0])) {
    t.hi[0] = h;
    t.lo[0] = l;
    t.j[0] = j;
  }
}

// Candidates: the groups (ihi == nullptr: group j valid when cnt[j] > 0) or the (hi, lo, j)
// triples of a previous pass.  Writes the block's best k triples (hi = ~0: none).
__global__ void __launch_bounds__(TS_BLOCK) k_topk_sel(uint64_t n, const uint32_t *__restrict__ d_n,
                                                       const unsigned long long *__restrict__ sum,
                                                       const unsigned int *__restrict__ cnt, GroupAttr ga,
                                                       int64_t g1min, int64_t kmin,
                                                       const uint64_t *__restrict__ ihi,
                                                       const uint64_t *__restrict__ ilo,
                                                       const uint32_t *__restrict__ ij, uint32_t k,
                                                       uint64_t *__restrict__ ohi, uint64_t *__restrict__ olo,
                                                       uint32_t *__restrict__ oj) {
  __shared__ uint64_t shi[TS_BLOCK * TS_K], slo[TS_BLOCK * TS_K];
  __shared__ uint32_t sj[TS_BLOCK * TS_K];
  __shared__ uint64_t whi[TS_BLOCK / 64], wlo[TS_BLOCK / 64];
  __shared__ uint32_t wt[TS_BLOCK / 64];
  if (d_n) n = *d_n < n ? *d_n : n;  // a candidate buffer's fill, capped at its size
  TopList t;
#pragma unroll
  for (int q = 0; q < TS_K; ++q) {
    t.hi[q] = ~0ull;
    t.lo[q] = ~0ull;
    t.j[q] = 0xFFFFFFFFu;
  }
  const uint64_t stride = (uint64_t)gridDim.x * TS_BLOCK;
  if (ihi) {
    for (uint64_t i = (uint64_t)blockIdx.x * TS_BLOCK + threadIdx.x; i < n; i += stride)
      ts_insert(t, ihi[i], ilo[i], ij[i]);
  } else {
    // four candidates' loads in flight before the first insertion
    constexpr int U = 4;
    for (uint64_t i0 = (uint64_t)blockIdx.x * TS_BLOCK + threadIdx.x; i0 < n; i0 += U * stride) {
      uint64_t h[U], l[U];
      bool v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t i = i0 + u * stride;
        v[u] = i < n && cnt[i] > 0;
        h[u] = v[u] ? ~((uint64_t)sum[i] ^ 0x8000000000000000ull) : ~0ull;
        l[u] = v[u] ? ((uint64_t)(ga_g1(ga, i) - g1min) << 32) | (uint32_t)((uint64_t)ga_key(ga, i) - (uint64_t)kmin)
                    : ~0ull;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (v[u]) ts_insert(t, h[u], l[u], (uint32_t)(i0 + u * stride));
    }
  }
#pragma unroll
  for (int q = 0; q < TS_K; ++q) {
    shi[threadIdx.x * TS_K + q] = t.hi[q];
    slo[threadIdx.x * TS_K + q] = t.lo[q];
    sj[threadIdx.x * TS_K + q] = t.j[q];
  }
  __syncthreads();
  uint32_t head = 0;  // next entry of this thread's list
  const unsigned lane = __lane_id(), wave = threadIdx.x / 64;
  for (uint32_t r = 0; r < k; ++r) {
    uint64_t h = head < TS_K ? shi[threadIdx.x * TS_K + head] : ~0ull;
    uint64_t l = head < TS_K ? slo[threadIdx.x * TS_K + head] : ~0ull;
    uint32_t w = threadIdx.x;
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t h2 = __shfl_xor(h, o, 64), l2 = __shfl_xor(l, o, 64);
      const uint32_t w2 = __shfl_xor(w, o, 64);
      if (ts_less(h2, l2, h, l) || (h2 == h && l2 == l && w2 < w)) {
        h = h2;
        l = l2;
        w = w2;
      }
    }
    if (lane == 0) {
      whi[wave] = h;
      wlo[wave] = l;
      wt[wave] = w;
    }
    __syncthreads();
    uint64_t bh = whi[0], bl = wlo[0];
    uint32_t bw = wt[0];
    for (int v = 1; v < TS_BLOCK / 64; ++v)
      if (ts_less(whi[v], wlo[v], bh, bl) || (whi[v] == bh && wlo[v] == bl && wt[v] < bw)) {
        bh = whi[v];
        bl = wlo[v];
        bw = wt[v];
      }
    const uint64_t o = (uint64_t)blockIdx.x * k + r;
    if (threadIdx.x == 0) {
      ohi[o] = bh;
      olo[o] = bl;
    }
    if (threadIdx.x == bw) {  // the winner owns the entry: emit its group, advance its head
      oj[o] = head < TS_K ? sj[threadIdx.x * TS_K + head] : 0xFFFFFFFFu;
      if (head < TS_K) ++head;
    }
    __syncthreads();
  }
}

// Threshold pass for the top-k: the k-th largest of the blocks' maximum sums is a lower
// bound of the k-th largest sum overall (each block maximum is a real group), so the groups
// at or above it hold the answer - usually a few dozen.  Sums compare as sum ^ 2^63.
constexpr int TB_BLOCK = 256, TB_ITEMS = 16;

__global__ void __launch_bounds__(TB_BLOCK) k_topk_bmax(uint64_t n, const unsigned long long *__restrict__ sum,
                                                        const unsigned int *__restrict__ cnt,
                                                        uint64_t *__restrict__ bmax) {
  __shared__ uint64_t wm[TB_BLOCK / 64];
  const uint64_t base = (uint64_t)blockIdx.x * TB_BLOCK * TB_ITEMS;
  uint64_t m = 0;
#pragma unroll
  for (int u = 0; u < TB_ITEMS; ++u) {
    const uint64_t i = base + (uint64_t)u * TB_BLOCK + threadIdx.x;
    if (i < n && cnt[i] > 0) {
      const uint64_t o = (uint64_t)sum[i] ^ 0x8000000000000000ull;
      m = o > m ? o : m;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t x = __shfl_xor(m, o, 64);
    m = x > m ? x : m;
  }
  if (__lane_id() == 0) wm[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < TB_BLOCK / 64; ++w) m = wm[w] > m ? wm[w] : m;
    bmax[blockIdx.x] = m;
  }
}

// One block: the k-th largest of nb block maxima (k <= TS_K).  Each thread keeps its best
// TS_K values sorted in registers (one read of bmax), then k rounds of a block-wide maximum
// over the list heads, the winner popping its head.  0 stands for "none": with fewer than k
// maxima the result is 0 and every group passes.
__global__ void __launch_bounds__(TB_BLOCK) k_topk_kth(const uint64_t *__restrict__ bmax, uint32_t nb, uint32_t k,
                                                       uint64_t *__restrict__ thr) {
  __shared__ uint64_t wtop[TB_BLOCK / 64][TS_K + 1];  // + 1: a merged wave's list ends in 0
  uint64_t l[TS_K];
#pragma unroll
  for (int q = 0; q < TS_K; ++q) l[q] = 0;
  // eight maxima's loads in flight before the first insertion (one round trip per eight
  // instead of one per value)
  constexpr int U = 8;
  for (uint32_t t0 = threadIdx.x; t0 < nb; t0 += U * TB_BLOCK) {
    uint64_t vs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t t = t0 + u * TB_BLOCK;
      vs[u] = 0;
      if (t < nb) vs[u] = bmax[t];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t v = vs[u];
#pragma unroll
      for (int q = 0; q < TS_K; ++q) {  // insertion into the descending list
        const uint64_t hi = v > l[q] ? v : l[q], lo = v > l[q] ? l[q] : v;
        l[q] = hi;
        v = lo;
      }
    }
  }
  // each wave: the k largest of its lanes' lists, by k rounds of a wave maximum (shuffles
  // only, no block barrier); then one thread merges the waves' descending lists
  const unsigned lane = __lane_id(), wv = threadIdx.x / 64;
  for (uint32_t r = 0; r < k; ++r) {
    uint64_t m = l[0];
    uint32_t who = lane;
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t m2 = __shfl_xor(m, o, 64);
      const uint32_t w2 = __shfl_xor(who, o, 64);
      if (m2 > m || (m2 == m && w2 < who)) {
        m = m2;
        who = w2;
      }
    }
    if (lane == 0) wtop[wv][r] = m;
    if (lane == who) {
#pragma unroll
      for (int q = 0; q + 1 < TS_K; ++q) l[q] = l[q + 1];
      l[TS_K - 1] = 0;
    }
  }
  if (lane == 0) wtop[wv][k] = 0;  // end of the wave's list
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t pos[TB_BLOCK / 64];
#pragma unroll
    for (int w = 0; w < TB_BLOCK / 64; ++w) pos[w] = 0;
    uint64_t last = 0;
    for (uint32_t r = 0; r < k; ++r) {
      uint64_t best = 0;
      int bw = 0;
#pragma unroll
      for (int w = 0; w < TB_BLOCK / 64; ++w) {
        const uint64_t v = wtop[w][pos[w]];
        if (w == 0 || v > best) {
          best = v;
          bw = w;
        }
      }
      last = best;
#pragma unroll
      for (int w = 0; w < TB_BLOCK / 64; ++w) pos[w] += (uint32_t)(w == bw);
    }
    *thr = last;
  }
}

// The groups with sum >= threshold, as (hi, lo, j) candidates; *count may exceed cap.
__global__ void __launch_bounds__(TB_BLOCK) k_topk_collect(uint64_t n, const unsigned long long *__restrict__ sum,
                                                           const unsigned int *__restrict__ cnt, GroupAttr ga,
                                                           int64_t g1min, int64_t kmin,
                                                           const uint64_t *__restrict__ thr,
                                                           uint32_t cap, uint32_t *__restrict__ count,
                                                           uint64_t *__restrict__ chi, uint64_t *__restrict__ clo,
                                                           uint32_t *__restrict__ cj) {
  const uint64_t t = *thr;
  const uint64_t base = (uint64_t)blockIdx.x * TB_BLOCK * TB_ITEMS;
#pragma unroll
  for (int u = 0; u < TB_ITEMS; ++u) {
    const uint64_t i = base + (uint64_t)u * TB_BLOCK + threadIdx.x;
    if (i < n && cnt[i] > 0) {
      const uint64_t o = (uint64_t)sum[i] ^ 0x8000000000000000ull;
      if (o >= t) {
        const uint32_t pos = atomicAdd(count, 1u);
        if (pos < cap) {
          chi[pos] = ~o;
          clo[pos] = ((uint64_t)(ga_g1(ga, i) - g1min) << 32) | (uint32_t)((uint64_t)ga_key(ga, i) - (uint64_t)kmin);
          cj[pos] = (uint32_t)i;
        }
      }
    }
  }
}

constexpr uint32_t TL_BLOCKS = 1024;
// The same two passes over the bit set of the groups that hold rows (one bit per group,
// ~5 % set): a thread per 32 groups reads one word and the sums of its set bits only.
// d_m (or null): the number of groups on the device, which bounds nw (allocated for more)
__global__ void __launch_bounds__(TB_BLOCK) k_topk_bmax_held(const uint32_t *__restrict__ held, uint64_t nw,
                                                             const unsigned long long *__restrict__ sum,
                                                             uint64_t *__restrict__ bmax,
                                                             const unsigned long long *__restrict__ d_m) {
  __shared__ uint64_t wm[TB_BLOCK / 64];
  if (d_m) nw = (*d_m + 31) / 32 < nw ? (*d_m + 31) / 32 : nw;
  uint64_t m = 0;
  for (uint64_t w = (uint64_t)blockIdx.x * TB_BLOCK + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * TB_BLOCK) {
    uint32_t b = held[w];
    while (b) {
      const uint64_t g = w * 32 + (uint64_t)__builtin_ctz(b);
      b &= b - 1;
      const uint64_t o = (uint64_t)sum[g] ^ 0x8000000000000000ull;
      m = o > m ? o : m;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t x = __shfl_xor(m, o, 64);
    m = x > m ? x : m;
  }
  if (__lane_id() == 0) wm[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < TB_BLOCK / 64; ++w) m = wm[w] > m ? wm[w] : m;
    bmax[blockIdx.x] = m;
  }
}

// attr == 0: the candidates' group attributes are not readable yet (their mid rows are
// found afterwards, k_topk_rows) and clo is left to k_topk_clo.
__global__ void __launch_bounds__(TB_BLOCK) k_topk_collect_held(const uint32_t *__restrict__ held, uint64_t nw,
                                                                const unsigned long long *__restrict__ sum,
                                                                GroupAttr ga, int64_t g1min, int64_t kmin,
                                                                const uint64_t *__restrict__ thr, uint32_t cap,
                                                                uint32_t *__restrict__ count,
                                                                uint64_t *__restrict__ chi,
                                                                uint64_t *__restrict__ clo,
                                                                uint32_t *__restrict__ cj, int attr,
                                                                const unsigned long long *__restrict__ d_m) {
  const uint64_t t = *thr;
  if (d_m) nw = (*d_m + 31) / 32 < nw ? (*d_m + 31) / 32 : nw;
  for (uint64_t w = (uint64_t)blockIdx.x * TB_BLOCK + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * TB_BLOCK) {
    uint32_t b = held[w];
    while (b) {
      const uint64_t g = w * 32 + (uint64_t)__builtin_ctz(b);
      b &= b - 1;
      const uint64_t o = (uint64_t)sum[g] ^ 0x8000000000000000ull;
      if (o >= t) {
        const uint32_t pos = atomicAdd(count, 1u);
        if (pos < cap) {
          chi[pos] = ~o;
          if (attr)
            clo[pos] = ((uint64_t)(ga_g1(ga, g) - g1min) << 32) | (uint32_t)((uint64_t)ga_key(ga, g) - (uint64_t)kmin);
          cj[pos] = (uint32_t)g;
        }
      }
    }
  }
}

// The second half of the candidates' keys (g1, then key) once their mid rows are known.
__global__ void __launch_bounds__(TB_BLOCK) k_topk_clo(const uint32_t *__restrict__ count, uint32_t cap,
                                                       const uint32_t *__restrict__ cj, GroupAttr ga, int64_t g1min,
                                                       int64_t kmin, uint64_t *__restrict__ clo) {
  const uint32_t n = *count < cap ? *count : cap;
  for (uint32_t i = blockIdx.x * TB_BLOCK + threadIdx.x; i < n; i += gridDim.x * TB_BLOCK) {
    const uint32_t g = cj[i];
    clo[i] = ((uint64_t)(ga_g1(ga, g) - g1min) << 32) | (uint32_t)((uint64_t)ga_key(ga, g) - (uint64_t)kmin);
  }
}

__global__ void k_topk_emit(const uint32_t *__restrict__ j, uint32_t k,
                            const unsigned long long *__restrict__ sum, const unsigned int *__restrict__ cnt,
                            GroupAttr ga, aln_star_row *__restrict__ out, uint32_t *__restrict__ nvalid) {
  const uint32_t i = threadIdx.x;
  const bool ok = i < k && j[i] != 0xFFFFFFFFu;  // empty entries carry no group
  if (ok) {
    const uint32_t g = j[i];
    aln_star_row r;
    r.key = ga_key(ga, g);
    r.g1 = ga_g1(ga, g);
    r.g2 = ga_g2(ga, g);
    r.sum = (int64_t)sum[g];
    r.count = cnt[g];
    out[i] = r;
  }
  const uint64_t b = __ballot(ok);  // k <= TS_K < 64: one wave; valid entries come first
  if (i == 0) *nvalid = (uint32_t)__popcll(b);
}

// The first k groups by (sum DESC, g1 ASC, key ASC, g2 ASC) (synchronises).  Small k with a
// g1 range below 2^32 take the register selection above; otherwise tiles of 1024
// candidates are sorted and cut to k, repeatedly, until one tile remains.
// With ctr (nctr device counters followed by room for the result block), the threshold
// path reads the counters and its result with one copy into h_ctr and sets *ctr_read.
// Every passing mid row's group into rows->grow (k_topk_rows, full mode).
static int mid_rows_all(aln_ctx *c, const MidRows &rows) {
  if (rows.ntiles == 0) return ALN_OK;
  hipLaunchKernelGGL(k_topk_rows, dim3((unsigned)((rows.ntiles + ST_BLOCK / 64 - 1) / (ST_BLOCK / 64))),
                     dim3(ST_BLOCK), 0, c->stream, rows, (const uint32_t *)nullptr, 0u, (const uint32_t *)nullptr,
                     (const uint32_t *)nullptr, 1);
  return check_launch("topk_rows");
}

// d_m (with held): m only bounds the groups, whose number is *d_m on the device (the
// threshold path must be taken then).
// mrows != nullptr: ga reads the groups' mid rows from mrows->grow, which holds nothing yet;
// the rows of the groups that need them are found here (only the candidates', when the
// threshold path can take them).
static int star_topk(aln_ctx *c, uint64_t m, const unsigned long long *sum, const unsigned int *cnt,
                     const GroupAttr &ga, uint32_t k, int64_t g1min, int64_t g1max, int64_t kmin,
                     aln_star_row *h_top, uint32_t *h_ntop, uint64_t *ctr = nullptr, uint32_t nctr = 0,
                     uint64_t *h_ctr = nullptr, bool *ctr_read = nullptr, const uint32_t *held = nullptr,
                     const MidRows *mrows = nullptr, const unsigned long long *d_m = nullptr) {
  const bool thr_path = k <= (uint32_t)TS_K && g1max >= g1min && (uint64_t)g1max - (uint64_t)g1min < (1ull << 32) &&
                        m < (1ull << 32);
  const bool lazy = mrows && held && thr_path;  // (with held, the path below has one result block)
  if (mrows && !lazy) ALN_TRY(mid_rows_all(c, *mrows));
  if (thr_path) {
    // threshold path: block maxima -> k-th largest -> the groups at or above it (over the
    // bit set of groups that hold rows when there is one, else over every group slot)
    {
      const uint64_t hw = (m + 31) / 32;
      const uint64_t nb = held ? std::min<uint64_t>(TL_BLOCKS, (hw + TB_BLOCK - 1) / TB_BLOCK)
                               : (m + TB_BLOCK * TB_ITEMS - 1) / (TB_BLOCK * TB_ITEMS);
      constexpr uint32_t CAP = 1u << 16;
      if (nb <= 64ull * TB_BLOCK) {
        // one result block: [u32 collected][u32 rows valid][aln_star_row rows[k]], read with
        // a single copy
        uint64_t *bmax, *thr, *chi, *clo, *oh, *ol;
        uint32_t *cj, *oj, *out;
        ALN_TRY(ws(c, nb, &bmax));
        ALN_TRY(ws(c, 1, &thr));
        ALN_TRY(ws(c, CAP, &chi));
        ALN_TRY(ws(c, CAP, &clo));
        ALN_TRY(ws(c, CAP, &cj));
        ALN_TRY(ws(c, k, &oh));
        ALN_TRY(ws(c, k, &ol));
        ALN_TRY(ws(c, k, &oj));
        if (ctr)
          out = reinterpret_cast<uint32_t *>(ctr + nctr);
        else
          ALN_TRY(ws(c, 2 + k * sizeof(aln_star_row) / 4, &out));
        uint32_t *count = out, *nv = out + 1;
        aln_star_row *rows = reinterpret_cast<aln_star_row *>(out + 2);
        ALN_TRY(fill_u32(c, out, 2, 0));
        {
          ProfScope p(c, "star_topk");
          if (held)
            hipLaunchKernelGGL(k_topk_bmax_held, dim3((unsigned)nb), dim3(TB_BLOCK), 0, c->stream, held, hw, sum,
                               bmax, d_m);
          else
            hipLaunchKernelGGL(k_topk_bmax, dim3((unsigned)nb), dim3(TB_BLOCK), 0, c->stream, m, sum, cnt, bmax);
          hipLaunchKernelGGL(k_topk_kth, dim3(1), dim3(TB_BLOCK), 0, c->stream, bmax, (uint32_t)nb, k, thr);
          if (held)
            hipLaunchKernelGGL(k_topk_collect_held, dim3((unsigned)nb), dim3(TB_BLOCK), 0, c->stream, held, hw, sum, ga,
                               g1min, kmin, thr, CAP, count, chi, clo, cj, lazy ? 0 : 1, d_m);
          else
            hipLaunchKernelGGL(k_topk_collect, dim3((unsigned)nb), dim3(TB_BLOCK), 0, c->stream, m, sum, cnt, ga,
                               g1min, kmin, thr, CAP, count, chi, clo, cj);
          if (lazy) {  // the candidates' codes and mid rows, then their (g1, key)
            uint32_t *ccode;
            ALN_TRY(ws(c, TR_MAX, &ccode));
            hipLaunchKernelGGL(k_topk_codes, dim3(TR_MAX / 4), dim3(256), 0, c->stream, *mrows,
                               (const uint32_t *)count, CAP, (const uint32_t *)cj, ccode);
            hipLaunchKernelGGL(k_topk_rows, dim3((unsigned)((mrows->ntiles + ST_BLOCK / 64 - 1) / (ST_BLOCK / 64))),
                               dim3(ST_BLOCK), 0, c->stream, *mrows, (const uint32_t *)count, CAP,
                               (const uint32_t *)cj, (const uint32_t *)ccode, 0);
            hipLaunchKernelGGL(k_topk_clo, dim3(CAP / TB_BLOCK / 4), dim3(TB_BLOCK), 0, c->stream,
                               (const uint32_t *)count, CAP, (const uint32_t *)cj, ga, g1min, kmin, clo);
          }
          hipLaunchKernelGGL(k_topk_sel, dim3(1), dim3(TS_BLOCK), 0, c->stream, (uint64_t)CAP, count, sum, cnt, ga,
                             g1min, kmin, chi, clo, cj, k, oh, ol, oj);
          hipLaunchKernelGGL(k_topk_emit, dim3(1), dim3(64), 0, c->stream, oj, k, sum, cnt, ga, rows, nv);
        }
        ALN_TRY(check_launch("star_topk"));
        std::vector<uint32_t> hres(2 + k * sizeof(aln_star_row) / 4);
        if (ctr) {  // the counters and the result with one copy
          std::vector<uint64_t> hb(nctr + (4 * hres.size() + 7) / 8);
          ALN_TRY(copy_to_host(c, hb.data(), ctr, 8 * hb.size()));
          memcpy(h_ctr, hb.data(), 8 * nctr);
          memcpy(hres.data(), hb.data() + nctr, 4 * hres.size());
          *ctr_read = true;
        } else {
          ALN_TRY(copy_to_host(c, hres.data(), out, 4 * hres.size()));
        }
        if (hres[0] <= CAP) {
          memcpy(h_top, hres.data() + 2, sizeof(aln_star_row) * hres[1]);
          *h_ntop = hres[1];
          return ALN_OK;
        }
        // more groups tie at the threshold than the buffer holds: the passes below, which
        // read every group's attributes (and, with d_m, need the exact group count: the
        // counters came back with the result)
        if (d_m) {
          ALN_REQUIRE(ctr, "star_topk: a device group count needs the counters");
          m = h_ctr[5];
        }
        if (lazy) ALN_TRY(mid_rows_all(c, *mrows));
      }
    }
    // passes of at most ~16 candidates per thread until one block holds the result
    uint64_t n = m;
    const uint64_t *ih = nullptr, *il = nullptr;
    const uint32_t *ij = nullptr;
    uint32_t *j2 = nullptr, *nv;
    aln_star_row *rows;
    ALN_TRY(ws(c, 1, &nv));
    ALN_TRY(ws(c, k, &rows));
    {
      ProfScope p(c, "star_topk");
      while (true) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>((n + TS_BLOCK * 16 - 1) / (TS_BLOCK * 16), 4096);
        uint64_t *oh, *ol;
        uint32_t *oj;
        ALN_TRY(ws(c, (size_t)nb * k, &oh));
        ALN_TRY(ws(c, (size_t)nb * k, &ol));
        ALN_TRY(ws(c, (size_t)nb * k, &oj));
        hipLaunchKernelGGL(k_topk_sel, dim3(nb), dim3(TS_BLOCK), 0, c->stream, n, (const uint32_t *)nullptr, sum, cnt,
                           ga, g1min, kmin, ih, il, ij, k, oh, ol, oj);
        ih = oh;
        il = ol;
        ij = oj;
        n = (uint64_t)nb * k;
        if (nb == 1) break;
      }
      j2 = const_cast<uint32_t *>(ij);
      hipLaunchKernelGGL(k_topk_emit, dim3(1), dim3(64), 0, c->stream, j2, k, sum, cnt, ga, rows, nv);
    }
    ALN_TRY(check_launch("star_topk"));
    uint32_t nvalid = 0;
    ALN_TRY(copy_to_host(c, &nvalid, nv, 4));
    if (nvalid) ALN_TRY(copy_to_host(c, h_top, rows, sizeof(aln_star_row) * nvalid));
    *h_ntop = nvalid;
    return ALN_OK;
  }
  const Cand *in = nullptr;
  uint64_t n = m;
  while (true) {
    uint64_t nt = (n + TK_TILE2 - 1) / TK_TILE2;
    Cand *out;
    ALN_TRY(ws(c, nt * k, &out));
    {
      ProfScope p(c, "star_topk");
      hipLaunchKernelGGL(k_star_topk_tile, dim3((unsigned)nt), dim3(256), 0, c->stream, in, n, sum, cnt, ga, k,
                         out);
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
  ALN_REQUIRE(a->d_key >= 0 && a->m_key >= 0 && a->m_fk >= 0 && a->f_fk >= 0, "aln_star_agg: key columns");
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
  ALN_REQUIRE((uint64_t)a->mid.cols[0].nsegs * a->mid.seg_rows <= 0xFFFFFFFFull,
              "aln_star_agg: mid table above 2^32 rows");
  ALN_TRY(make_scan(c, a->fact, a->f_fk, &fd, &fblocks));
  ALN_REQUIRE(a->f_x >= 0 && a->f_x < a->fact.ncols && a->f_y >= 0 && a->f_y < a->fact.ncols,
              "aln_star_agg: aggregate columns");

  // key ranges from the zone maps: those of the last call when its key columns are these
  // (checked on the device below; the run is repeated with fresh ones if they changed),
  // else read back
  int64_t dkmin, dkmax, mkmin, mkmax, g1min, g1max;
  const aln_col *rc[3] = {&a->dim.cols[a->d_key], &a->mid.cols[a->m_key],
                          a->m_g1 >= 0 ? &a->mid.cols[a->m_g1] : nullptr};
  aln_ctx::StarRanges &R = c->star_rng;
  bool spec = R.valid;
  for (int i = 0; i < 3; ++i)
    spec = spec && R.segs[i] == (rc[i] ? (const void *)rc[i]->segs : nullptr) &&
           R.nsegs[i] == (rc[i] ? rc[i]->nsegs : 0u);
  int64_t mn[3], mx[3];
  int64_t *dfresh = nullptr;  // the fresh ranges, 2 per column, on the device (spec)
  RangePair want;
  uint32_t nused = 0;
  if (spec) {
    ColsRange cr;
    for (int i = 0; i < 3; ++i) {
      mn[i] = R.mn[i];
      mx[i] = R.mx[i];
      if (!rc[i]) continue;
      cr.segs[nused] = rc[i]->segs;
      cr.nsegs[nused] = rc[i]->nsegs;
      want.v[2 * nused] = mn[i];
      want.v[2 * nused + 1] = mx[i];
      ++nused;
    }
    ALN_TRY(ws(c, 2 * nused, &dfresh));
    hipLaunchKernelGGL(k_cols_range, dim3(nused), dim3(256), 0, c->stream, cr, dfresh);
    ALN_TRY(check_launch("cols_range"));
  } else {
    R.valid = false;
    ALN_TRY(cols_range(c, rc, 3, mn, mx));
  }
  dkmin = mn[0], dkmax = mx[0], mkmin = mn[1], mkmax = mx[1], g1min = mn[2], g1max = mx[2];
  const uint64_t drange = (uint64_t)dkmax - (uint64_t)dkmin + 1;
  const uint64_t mrange = (uint64_t)mkmax - (uint64_t)mkmin + 1;
  if (a->dim.cols[0].nsegs == 0 || a->mid.cols[0].nsegs == 0 || dkmax < dkmin || mkmax < mkmin ||
      drange == 0 || drange > (1ull << 34) || mrange == 0 || mrange >= 0xFFFFFFFFull) {
    // (reused ranges are never of this kind: they are kept only when these checks pass)
    if (a->dim.cols[0].nsegs == 0 || a->mid.cols[0].nsegs == 0 || dkmax < dkmin || mkmax < mkmin) {
      if (h_stats) *h_stats = st;
      return ALN_OK;
    }
    if (drange == 0 || drange > (1ull << 34))
      return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: dimension key range %llu too wide for a bitmap",
                       (unsigned long long)drange);
    return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: mid key range %llu wider than 32 bits",
                     (unsigned long long)mrange);
  }
  if (!spec) {
    R.valid = true;
    for (int i = 0; i < 3; ++i) {
      R.segs[i] = rc[i] ? (const void *)rc[i]->segs : nullptr;
      R.nsegs[i] = rc[i] ? rc[i]->nsegs : 0u;
      R.mn[i] = mn[i];
      R.mx[i] = mx[i];
    }
  }
  // join structure for mid -> fact: bitmap + rank over dense key ranges (up to a 256 MB
  // bitmap), MurmurHash open addressing otherwise or when flags bit 1 asks for it
  const bool use_bm = !(a->flags & 2) && mrange <= (1ull << 31);
  uint64_t mwords = use_bm ? (mrange + 31) / 32 : 0;
  uint64_t mblk = use_bm ? (mwords + BM_BLOCK_WORDS - 1) / BM_BLOCK_WORDS : 0;
  uint32_t *mbits = nullptr;
  if (use_bm)  // whole 512-bit blocks (zeroed by the setup launch below)
    ALN_TRY(ws(c, mblk * BM_BLOCK_WORDS, &mbits));
  uint64_t bwords = (drange + 31) / 32;
  uint32_t *bitmap, *dup;
  // device counters, read back with one copy each time: 0 dim pass, 1 mid pass, 2 mid probe,
  // 3 fact pass, 4 fact probe, 5 distinct passing mid keys, 6 duplicate flags (two u32; the
  // hash join's build sets the second), 7 distinct passing dimension keys, 8 reused key
  // ranges found stale
  constexpr uint32_t NCTR = 9;
  uint64_t *ctr;
  uint32_t *dcnt, *mcnt, *mprobe, *fmatch, *fprobe;
  ALN_TRY(ws(c, bwords, &bitmap));
  // (room after the 8 counters for the top-k result block: 2 u32 + limit rows)
  ALN_TRY(ws(c, NCTR + (8 + (uint64_t)a->limit * sizeof(aln_star_row) + 7) / 8, &ctr));
  dup = reinterpret_cast<uint32_t *>(ctr + 6);
  const uint64_t dregions = dblocks * (ST_BLOCK / 64);
  ALN_TRY(ws(c, dregions, &dcnt));
  const uint64_t mregions = mblocks * (ST_BLOCK / 64), fwaves = fblocks * (ST_BLOCK / 64);
  ALN_TRY(ws(c, mregions, &mcnt));
  ALN_TRY(ws(c, mregions, &mprobe));
  ALN_TRY(ws(c, fwaves, &fmatch));
  ALN_TRY(ws(c, fwaves, &fprobe));
  uint32_t *held = nullptr;  // bit set of the groups that hold rows (bitmap join)
  // one launch: per-segment plans of the three scans, and zeroed bitmaps and counters
  SegPlan *dplan, *mplan, *fplan = nullptr;
  {
    const uint32_t ds = a->dim.cols[0].nsegs, ms = a->mid.cols[0].nsegs;
    const uint32_t fs = use_bm ? a->fact.cols[0].nsegs : 0u;
    const int dfast = (int)(drange <= (1ull << 31));
    ALN_TRY(ws(c, ds, &dplan));
    ALN_TRY(ws(c, ms, &mplan));
    ALN_TRY(ws(c, fs ? fs : 1, &fplan));
    const PlanJob jd{{-1, a->d_key, -1, -1, -1}, DIM_STAGE, dkmin, dfast, (uint32_t)DIM_OWN, ds, dplan};
    const PlanJob jm{{-1, a->m_fk, a->m_key, a->m_g1, a->m_g2}, MID_STAGE, dkmin, dfast, (uint32_t)PL_OWN, ms, mplan};
    const PlanJob jf{{-1, a->f_fk, a->f_x, a->f_y, -1}, FACT_STAGE, mkmin, 1, (uint32_t)FACT_OWN, fs, fplan};
    const FillJob fj{{bitmap, reinterpret_cast<uint32_t *>(ctr), mbits, nullptr},
                     {bwords, 2 * NCTR, use_bm ? mblk * BM_BLOCK_WORDS : 0, 0}};
    const uint64_t most = std::max<uint64_t>(bwords, use_bm ? mblk * BM_BLOCK_WORDS : 0);
    const uint32_t pb = (ds + ms + fs + 127) / 128, fb = (uint32_t)std::min<uint64_t>((most + 1023) / 1024, 16384);
    hipLaunchKernelGGL(k_star_setup, dim3(pb + (fb ? fb : 1)), dim3(128), 0, c->stream, dd, md, fd, jd, jm, jf, pb,
                       fj);
    ALN_TRY(check_launch("star_setup"));
  }
  if (spec) {  // the reused ranges against the fresh ones (counter 8, zeroed by the setup)
    hipLaunchKernelGGL(k_range_check, dim3(1), dim3(64), 0, c->stream, (const int64_t *)dfresh, want, nused,
                       reinterpret_cast<unsigned long long *>(ctr + 8));
    ALN_TRY(check_launch("range_check"));
  }

  tm.mark(0);
  {
    ProfScope p(c, "star_dim");
    hipLaunchKernelGGL(k_star_dim, dim3((unsigned)dblocks), dim3(ST_BLOCK), 0, c->stream, dd, dplan, a->d_key,
                       dkmin, drange, bitmap, dcnt);
  }
  ALN_TRY(check_launch("star_dim"));
  tm.mark(1);

  // mid scan: passing rows of wave region b land in entries [b * PL_WAVE_ROWS, ...) of the
  // region arrays
  uint2 *rent, *trange = nullptr;
  ALN_TRY(ws(c, mregions * PL_WAVE_ROWS, &rent));
  if (use_bm) ALN_TRY(ws(c, mblocks, &trange));
  {
    ProfScope p(c, "star_mid");
    hipLaunchKernelGGL(k_star_mid, dim3((unsigned)mblocks), dim3(ST_BLOCK), 0, c->stream, md, mplan, a->m_key,
                       a->m_fk, bitmap, dkmin, drange, mcnt, mprobe, rent, mbits, mkmin, mrange, trange);
  }
  ALN_TRY(check_launch("star_mid"));
  // bitmap join: the rank directory now (its total = distinct passing keys), before the
  // counters are summed
  uint64_t *rank = nullptr, *rel = nullptr;
  uint32_t *bcnt = nullptr;
  if (use_bm) {
    ALN_TRY(ws(c, mblk, &bcnt));
    ALN_TRY(ws(c, mblk, &rank));
    ALN_TRY(ws(c, mblk, &rel));
    {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_bm_block_counts, dim3(grid_for(mblk, 256, 65536)), dim3(256), 0, c->stream,
                         mbits, mwords, mblk, bcnt, rel);
    }
    ALN_TRY(check_launch("bm_block_counts"));
    ALN_TRY(exclusive_scan_u32(c, bcnt, mblk, rank, nullptr));
  }
  {  // the counters: dim pass, mid pass, mid probe, distinct dim keys (set bits of the
     // dimension bitmap) and (bitmap join) distinct mid keys
    const uint32_t *in[5] = {dcnt, mcnt, mprobe, bitmap, bcnt};
    const uint64_t n[5] = {dregions, mregions, mregions, bwords, mblk};
    uint64_t *out[5] = {ctr + 0, ctr + 1, ctr + 2, ctr + 7, ctr + 5};
    ALN_TRY(sum_u32_multi(c, use_bm ? 5 : 4, in, n, out, 1u << 3));
  }
  // No synchronisation here on the common path (bitmap join, threshold top-k, no group
  // list): the group arrays are allocated for an upper bound of the group count (distinct
  // passing mid keys <= mid rows, <= key range), zeroed up to the device's count of distinct
  // passing keys (counter 5), and the counters are checked after the final readback.
  const uint64_t m_ub = std::min<uint64_t>((uint64_t)a->mid.cols[0].nsegs * a->mid.seg_rows, mrange);
  const bool nosync = use_bm && !(a->flags & 1) && a->limit > 0 && a->limit <= (uint32_t)TS_K &&
                      g1max >= g1min && (uint64_t)g1max - (uint64_t)g1min < (1ull << 32) &&
                      m_ub <= (1ull << 28);  // (16 B of group arrays per possible group)
  uint64_t hc[NCTR];
  auto check_counters = [&]() -> int {  // after a readback of the counters into hc
    const uint32_t *hdup = reinterpret_cast<const uint32_t *>(hc + 6);
    if (hc[7] != hc[0]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate dimension keys");
    if (hdup[1] || (use_bm && hc[5] != hc[1]))
      return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate mid keys");
    st.dim_pass = hc[0];
    st.mid_pass = hc[1];
    st.mid_probe = hc[2];
    return ALN_OK;
  };
  uint64_t m;
  if (nosync) {
    m = m_ub;
  } else {
    ALN_TRY(copy_to_host(c, hc, ctr, sizeof(hc)));
    if (hc[8]) {  // the columns changed since their ranges were taken: run again, reading them
      R.valid = false;
      return aln_star_agg(c, a, h_top, h_ntop, d_groups, groups_cap, h_stats);
    }
    ALN_TRY(check_counters());
    m = hc[1];
  }

  unsigned long long *sum;
  unsigned int *cnt, *grow;
  ALN_TRY(ws(c, m ? m : 1, &sum));
  ALN_TRY(ws(c, m ? m : 1, &cnt));
  ALN_TRY(ws(c, m ? m : 1, &grow));
  const GroupAttr ga{nullptr, nullptr, nullptr, grow, mplan, a->mid.seg_rows};
  uint64_t *table = nullptr;
  uint64_t tcap = 0;
  const MidRows rows{rent, mcnt, mblocks, mbits, rank, rel, mblk, trange, grow};
  if (use_bm) {
    // groups are ranks (= key order): zeroed accumulators now; the mid row of a group is
    // looked up later, for the groups the top-k or the group list needs
    ALN_TRY(ws(c, (m + 31) / 32, &held));  // bit set of the groups that hold rows
    if (m) {
      ProfScope p(c, "star_build");
      if (nosync) {
        hipLaunchKernelGGL(k_zero_groups, dim3(grid_for(m, 256, 8192)), dim3(256), 0, c->stream, sum, cnt, held,
                           reinterpret_cast<const unsigned long long *>(ctr + 5));
        ALN_TRY(check_launch("zero_groups"));
      } else {
        ALN_TRY(fill_u64(c, reinterpret_cast<uint64_t *>(sum), m, 0));
        ALN_TRY(fill_u32(c, cnt, m, 0));
        ALN_TRY(fill_u32(c, held, (m + 31) / 32, 0));
      }
    }
  } else {
    uint64_t *moff;
    ALN_TRY(ws(c, mregions, &moff));
    ALN_TRY(exclusive_scan_u32(c, mcnt, mregions, moff, nullptr));
    tcap = m + m / 2 + 1024;
    ALN_TRY(ws(c, tcap, &table));
    ALN_TRY(fill_u64(c, table, tcap, ST_EMPTY));
    if (m) {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_star_build, dim3((unsigned)mblocks), dim3(ST_BLOCK), 0, c->stream, rent, mcnt, moff,
                         table, tcap, grow, sum, cnt, dup + 1);
    }
    ALN_TRY(check_launch("star_build"));
  }
  tm.mark(2);
  if (m && fblocks) {
    {
      ProfScope p(c, "star_fact");  // the scan and, for the bitmap join, the deferred rows
      if (use_bm) {
        uint2 *hitbuf;
        ALN_TRY(ws(c, fwaves * FACT_HIT_CAP, &hitbuf));
        const FactArgs fa{a->f_fk, a->a0, a->a1, a->b0, a->b1, mbits, rank, rel, mkmin, mrange, sum, cnt,
                          fmatch, fprobe, hitbuf, held};
        hipLaunchKernelGGL(k_star_fact_bm, dim3((unsigned)fblocks), dim3(ST_BLOCK), 0, c->stream, fd, fplan, fa);
        ALN_TRY(check_launch("star_fact"));
        const uint64_t fsegs = fblocks / fd.tps, rwaves = fsegs * ((fd.tps + RS_TILES - 1) / RS_TILES);
        hipLaunchKernelGGL(k_fact_resolve, dim3((unsigned)rwaves), dim3(64), 0, c->stream, hitbuf, fsegs, fd.tps,
                           fplan, fa);
      } else {  // per-block counts
        hipLaunchKernelGGL(k_star_fact, dim3((unsigned)fblocks), dim3(ST_BLOCK), 0, c->stream, fd, a->f_fk,
                           a->f_x, a->f_y, a->a0, a->a1, a->b0, a->b1, table, tcap, mkmin, mrange, sum, cnt,
                           fmatch, fprobe);
      }
      ALN_TRY(check_launch("star_fact"));
    }
    const uint64_t nfc = use_bm ? fwaves : fblocks;
    const uint32_t *in[2] = {fmatch, fprobe};
    const uint64_t n[2] = {nfc, nfc};
    uint64_t *out[2] = {ctr + 3, ctr + 4};
    ALN_TRY(sum_u32_multi(c, 2, in, n, out));
  }
  tm.mark(3);

  // top-k
  bool ctr_read = false;
  if (a->limit && m)
    ALN_TRY(star_topk(c, m, sum, cnt, ga, a->limit, g1min, g1max, mkmin, h_top, h_ntop, ctr, NCTR, hc, &ctr_read,
                      held, use_bm ? &rows : nullptr,
                      nosync ? reinterpret_cast<const unsigned long long *>(ctr + 5) : nullptr));
  tm.mark(4);
  if (!ctr_read) ALN_TRY(copy_to_host(c, hc, ctr, sizeof(hc)));
  if (nosync) {
    if (hc[8]) {  // the columns changed since their ranges were taken: run again, reading them
      R.valid = false;
      return aln_star_agg(c, a, h_top, h_ntop, d_groups, groups_cap, h_stats);
    }
    ALN_TRY(check_counters());
  }
  if (reinterpret_cast<const uint32_t *>(hc + 6)[1]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg: duplicate mid keys");
  st.fact_pass = hc[3];
  st.fact_probe = hc[4];

  // every group, in ascending key order
  if ((a->flags & 1) && m) {
    if (use_bm) ALN_TRY(mid_rows_all(c, rows));
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
                         pos, m, ga, mkmin, code, idx);
      ALN_TRY(check_launch("star_group_codes"));
      if (!use_bm) ALN_TRY(radix_sort_pairs(c, code, idx, g, bits_for_range_u(mrange - 1)));
      hipLaunchKernelGGL(k_star_group_emit, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, idx, g, ga,
                         sum, cnt, d_groups);
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

// ---- batched star pipelines ----------------------------------------------------------------
// Up to ALN_STAR_BATCH_MAX q3-shaped queries over the same three tables, differing in their
// range predicates and aggregate coefficients (TPC-H q3's substitution parameters), answered
// with one scan of each table: every tile is staged once and evaluated for every query.
// Each query keeps its own dimension bitmap, mid bitmap + rank directory and accumulators;
// the group attributes are written by a second mid scan once the ranks are known.
namespace aln {

constexpr int SB_MAX = ALN_STAR_BATCH_MAX;
constexpr int SB_CHUNK = 4;  // items per hit-list pass of the fact kernel (1024 rows)

struct SBRanges {
  int32_t nr;
  int32_t col[ALN_MAX_RANGES];
  int64_t lo[ALN_MAX_RANGES], hi[ALN_MAX_RANGES];
};

struct SBQuery {
  SBRanges r[3];  // dimension, mid, fact
  int64_t a0, a1, b0, b1;
  uint32_t *dbm;    // dimension bitmap over [dkmin, dkmin + drange)
  uint32_t *mbits;  // mid bitmap, whole 512-bit blocks
  uint64_t *rank;   // per-block directory of mbits
  unsigned long long *sum;
  unsigned int *cnt;
  int64_t *okey, *og1, *og2;
};

struct SBScan {
  const uint32_t *words[ALN_MAX_COLS];
  const aln_seg *segs[ALN_MAX_COLS];
  uint32_t tps;
  int want[2];  // staged columns
  int nwant;
};

__device__ __forceinline__ char sb_mode(const SBScan &t, const SBRanges &r, uint32_t seg) {
  char m = 'A';
  for (int k = 0; k < r.nr; ++k) {
    const aln_seg &s = t.segs[r.col[k]][seg];
    if (s.vmax < r.lo[k] || s.vmin > r.hi[k]) return 'N';
    if (!(s.vmin >= r.lo[k] && s.vmax <= r.hi[k])) m = 'R';
  }
  return m;
}

// A query's ranges on one table restated on the staged tile's codes (see TilePred).
struct SBPred {
  int8_t slot[ALN_MAX_RANGES];
  CodeRange cr[ALN_MAX_RANGES];
  bool none;
};

__device__ __forceinline__ bool sb_pass_fast(const SBScan &t, const TileStage &ts, const SBRanges &r,
                                             const SBPred &sp, char mode, uint32_t seg, uint32_t r0, uint32_t i) {
  if (mode == 'A') return true;
  if (sp.none) return false;
  bool p = true;
  for (int k = 0; k < r.nr; ++k) {
    const int s = sp.slot[k];
    if (s >= 0) {
      p = p && (stage_code(ts.words[s], ts.slot[s], i) - sp.cr[k].clo <= sp.cr[k].cspan);
    } else {
      const int64_t v = decode_row(t.words[r.col[k]], t.segs[r.col[k]][seg], r0 + i);
      p = p && v >= r.lo[k] && v <= r.hi[k];
    }
  }
  return p;
}

// Block prologue: zone-map mode of the segment for every query and the tile staged in LDS
// (skipped when no query reads the segment), then each query's ranges in code space.
// Returns false when the block has no work.
__device__ __forceinline__ bool sb_prologue(const SBScan &t, const SBQuery *q, int nq, int table,
                                            TileStage &ts, char *qmode, SBPred *sp, uint32_t seg, uint32_t r0,
                                            uint32_t nt) {
  __shared__ int any;
  __shared__ int want[2];
  if (threadIdx.x == 0) any = 0;
  if (threadIdx.x < 2) want[threadIdx.x] = t.want[threadIdx.x];
  __syncthreads();
  if (threadIdx.x < (unsigned)nq) {
    qmode[threadIdx.x] = sb_mode(t, q[threadIdx.x].r[table], seg);
    if (qmode[threadIdx.x] != 'N') atomicOr(&any, 1);
  }
  __syncthreads();
  if (!any) return false;
  stage_tile<ST_BLOCK>(ts, t.words, t.segs, want, t.nwant, seg, r0, nt);
  __syncthreads();
  if (threadIdx.x < (unsigned)nq) {
    const SBRanges &r = q[threadIdx.x].r[table];
    SBPred &p = sp[threadIdx.x];
    p.none = false;
    for (int k = 0; k < r.nr; ++k) {
      const int s = ts.slot_of[r.col[k]];
      p.slot[k] = (int8_t)s;
      if (s >= 0) {
        p.cr[k] = code_range(ts.slot[s], r.lo[k], r.hi[k]);
        p.none = p.none || p.cr[k].none;
      }
    }
  }
  __syncthreads();
  return true;
}

__global__ void __launch_bounds__(ST_BLOCK) k_sb_dim(SBScan t, const SBQuery *__restrict__ q, int nq, int kcol,
                                                     int64_t kmin, uint64_t range, uint32_t *npass,
                                                     uint32_t *dup) {
  __shared__ TileStage ts;
  __shared__ char qmode[SB_MAX];
  __shared__ SBPred sp[SB_MAX];
  const uint32_t seg = blockIdx.x / t.tps, tile = blockIdx.x % t.tps;
  const uint32_t nrows = t.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const uint32_t nt = r0 < nrows ? (nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE) : 0;
  if (!nt || !sb_prologue(t, q, nq, 0, ts, qmode, sp, seg, r0, nt)) {
    if (threadIdx.x < (unsigned)nq) npass[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = 0;
    return;
  }
  for (int qi = 0; qi < nq; ++qi) {
    const char mode = qmode[qi];
    uint32_t cnt = 0;
    bool twice = false;
    if (mode != 'N') {
      const SBQuery &qq = q[qi];
#pragma unroll 4
      for (int j = 0; j < ST_ITEMS; ++j) {
        const uint32_t i = j * ST_BLOCK + threadIdx.x;
        if (i < nt && sb_pass_fast(t, ts, qq.r[0], sp[qi], mode, seg, r0, i)) {
          const uint64_t b = (uint64_t)tile_get(ts, t.words, t.segs, kcol, seg, r0, i) - (uint64_t)kmin;
          if (b < range) {
            const uint32_t bit = 1u << (b & 31);
            twice = twice || (atomicOr(&qq.dbm[b >> 5], bit) & bit);
            ++cnt;
          }
        }
      }
    }
    if (twice) atomicExch(&dup[qi], 1u);
    block_count(cnt, npass + (uint64_t)qi * gridDim.x);
  }
}

// Mid scan 1: ranges + dimension-bitmap probe per query; passing keys set the query's mid
// bitmap.  Counts per query: probes (rows passing the ranges) and passes.
__global__ void __launch_bounds__(ST_BLOCK) k_sb_mid(SBScan t, const SBQuery *__restrict__ q, int nq, int kcol,
                                                     int fkcol, int64_t dkmin, uint64_t drange, int64_t mkmin,
                                                     uint64_t mrange, uint32_t *nprobe_out, uint32_t *npass_out,
                                                     uint32_t *dup) {
  __shared__ TileStage ts;
  __shared__ char qmode[SB_MAX];
  __shared__ SBPred sp[SB_MAX];
  const uint32_t seg = blockIdx.x / t.tps, tile = blockIdx.x % t.tps;
  const uint32_t nrows = t.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const uint32_t nt = r0 < nrows ? (nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE) : 0;
  if (!nt || !sb_prologue(t, q, nq, 1, ts, qmode, sp, seg, r0, nt)) {
    if (threadIdx.x < (unsigned)nq) {
      nprobe_out[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = 0;
      npass_out[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = 0;
    }
    return;
  }
  // foreign key (LDS) and key (HBM; only the rows of a tile, coalesced) of every row, once
  uint64_t fkc[ST_ITEMS];
  int64_t key[ST_ITEMS];
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    const uint32_t i = j * ST_BLOCK + threadIdx.x;
    fkc[j] = i < nt ? (uint64_t)tile_get(ts, t.words, t.segs, fkcol, seg, r0, i) - (uint64_t)dkmin : ~0ull;
    key[j] = i < nt ? tile_get(ts, t.words, t.segs, kcol, seg, r0, i) : 0;
  }
  for (int qi = 0; qi < nq; ++qi) {
    const char mode = qmode[qi];
    uint32_t np = 0, npass = 0;
    bool twice = false;
    if (mode != 'N') {
      const SBQuery &qq = q[qi];
      uint32_t w[ST_ITEMS];
      uint32_t probe = 0;
#pragma unroll
      for (int j = 0; j < ST_ITEMS; ++j) {
        const uint32_t i = j * ST_BLOCK + threadIdx.x;
        const bool p = i < nt && sb_pass_fast(t, ts, qq.r[1], sp[qi], mode, seg, r0, i);
        np += p;
        probe |= (uint32_t)(p && fkc[j] < drange) << j;
      }
#pragma unroll
      for (int j = 0; j < ST_ITEMS; ++j) w[j] = ((probe >> j) & 1) ? qq.dbm[fkc[j] >> 5] : 0u;
#pragma unroll
      for (int j = 0; j < ST_ITEMS; ++j)
        if (((probe >> j) & 1) && ((w[j] >> (fkc[j] & 31)) & 1)) {
          const uint64_t code = (uint64_t)key[j] - (uint64_t)mkmin;
          if (code < mrange) {
            const uint32_t bit = 1u << (code & 31);
            twice = twice || (atomicOr(&qq.mbits[code >> 5], bit) & bit);
            ++npass;
          }
        }
    }
    if (twice) atomicExch(&dup[qi], 1u);
    block_count(np, nprobe_out + (uint64_t)qi * gridDim.x);
    block_count(npass, npass_out + (uint64_t)qi * gridDim.x);
  }
}

// Mid scan 2: the rows whose key bit is set in a query's mid bitmap are that query's groups;
// their key and group columns go to position rank(key).
__global__ void __launch_bounds__(ST_BLOCK) k_sb_attr(SBScan t, const SBQuery *__restrict__ q, int nq, int kcol,
                                                      int g1col, int g2col, int64_t mkmin, uint64_t mrange) {
  __shared__ TileStage ts;
  __shared__ int want[2];
  const uint32_t seg = blockIdx.x / t.tps, tile = blockIdx.x % t.tps;
  const uint32_t nrows = t.segs[0][seg].nrows, r0 = tile * ST_TILE;
  if (r0 >= nrows) return;
  const uint32_t nt = nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE;
  if (threadIdx.x < 2) want[threadIdx.x] = t.want[threadIdx.x];
  __syncthreads();
  stage_tile<ST_BLOCK>(ts, t.words, t.segs, want, t.nwant, seg, r0, nt);
  __syncthreads();
  // items in chunks (the rank of a hit reads a 64-byte line: a whole tile of hits in flight
  // would need more registers than a wave has)
  for (int ch = 0; ch < ST_ITEMS / SB_CHUNK; ++ch) {
    uint32_t code[SB_CHUNK];
#pragma unroll
    for (int u = 0; u < SB_CHUNK; ++u) {
      const uint32_t i = (ch * SB_CHUNK + u) * ST_BLOCK + threadIdx.x;
      const uint64_t c = i < nt ? (uint64_t)tile_get(ts, t.words, t.segs, kcol, seg, r0, i) - (uint64_t)mkmin : ~0ull;
      code[u] = c < mrange ? (uint32_t)c : 0xFFFFFFFFu;
    }
    for (int qi = 0; qi < nq; ++qi) {
      const SBQuery &qq = q[qi];
      uint32_t w[SB_CHUNK];
#pragma unroll
      for (int u = 0; u < SB_CHUNK; ++u) w[u] = code[u] != 0xFFFFFFFFu ? qq.mbits[code[u] >> 5] : 0u;
#pragma unroll
      for (int u = 0; u < SB_CHUNK; ++u)
        if (code[u] != 0xFFFFFFFFu && ((w[u] >> (code[u] & 31)) & 1)) {
          const uint32_t i = (ch * SB_CHUNK + u) * ST_BLOCK + threadIdx.x;
          const uint64_t g = bm_rank(qq.mbits, qq.rank, code[u]);
          qq.okey[g] = (int64_t)((uint64_t)code[u] + (uint64_t)mkmin);
          qq.og1[g] = g1col >= 0 ? tile_get(ts, t.words, t.segs, g1col, seg, r0, i) : 0;
          qq.og2[g] = g2col >= 0 ? tile_get(ts, t.words, t.segs, g2col, seg, r0, i) : 0;
        }
    }
  }
}

// Fact scan: per query, the tile's rows passing its ranges probe its mid bitmap; hits are
// compacted (row order) into an LDS list 1024 rows at a time and aggregated from there.
__global__ void __launch_bounds__(ST_BLOCK) k_sb_fact(SBScan t, const SBQuery *__restrict__ q, int nq, int fkcol,
                                                      int xcol, int ycol, int64_t mkmin, uint64_t mrange,
                                                      uint32_t *nprobe_out, uint32_t *nmatch_out) {
  __shared__ TileStage ts;
  __shared__ char qmode[SB_MAX];
  __shared__ SBPred sp[SB_MAX];
  __shared__ uint32_t hcode[SB_CHUNK * ST_BLOCK];
  __shared__ uint16_t hrow[SB_CHUNK * ST_BLOCK];
  __shared__ uint32_t woff[SB_CHUNK][ST_BLOCK / 64];
  __shared__ uint32_t nhit_s;
  const uint32_t seg = blockIdx.x / t.tps, tile = blockIdx.x % t.tps;
  const uint32_t nrows = t.segs[0][seg].nrows, r0 = tile * ST_TILE;
  const uint32_t nt = r0 < nrows ? (nrows - r0 < (uint32_t)ST_TILE ? nrows - r0 : (uint32_t)ST_TILE) : 0;
  if (!nt || !sb_prologue(t, q, nq, 2, ts, qmode, sp, seg, r0, nt)) {
    if (threadIdx.x < (unsigned)nq) {
      nprobe_out[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = 0;
      nmatch_out[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = 0;
    }
    return;
  }
  const unsigned lane = __lane_id(), wave = threadIdx.x / 64;
  uint32_t fkc[ST_ITEMS];
#pragma unroll
  for (int j = 0; j < ST_ITEMS; ++j) {
    const uint32_t i = j * ST_BLOCK + threadIdx.x;
    const uint64_t c = i < nt ? (uint64_t)tile_get(ts, t.words, t.segs, fkcol, seg, r0, i) - (uint64_t)mkmin : ~0ull;
    fkc[j] = c < mrange ? (uint32_t)c : 0xFFFFFFFFu;  // mrange <= 2^31 on this path
  }
  for (int qi = 0; qi < nq; ++qi) {
    const char mode = qmode[qi];
    uint32_t nprobe = 0, nmatch = 0;
    if (mode != 'N') {
      const SBQuery &qq = q[qi];
      for (int ch = 0; ch < ST_ITEMS / SB_CHUNK; ++ch) {
        uint32_t code[SB_CHUNK], w[SB_CHUNK];
#pragma unroll
        for (int u = 0; u < SB_CHUNK; ++u) {
          const int j = ch * SB_CHUNK + u;
          const uint32_t i = j * ST_BLOCK + threadIdx.x;
          const bool p = i < nt && sb_pass_fast(t, ts, qq.r[2], sp[qi], mode, seg, r0, i);
          nprobe += p;
          code[u] = p ? fkc[j] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < SB_CHUNK; ++u) w[u] = code[u] != 0xFFFFFFFFu ? qq.mbits[code[u] >> 5] : 0u;
        uint32_t hits = 0;
#pragma unroll
        for (int u = 0; u < SB_CHUNK; ++u)
          hits |= (uint32_t)(code[u] != 0xFFFFFFFFu && ((w[u] >> (code[u] & 31)) & 1)) << u;
#pragma unroll
        for (int u = 0; u < SB_CHUNK; ++u) {
          const uint64_t b = __ballot((hits >> u) & 1);
          if (lane == 0) woff[u][wave] = (uint32_t)__popcll(b);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          uint32_t tot = 0;
          for (int u = 0; u < SB_CHUNK; ++u)
            for (int v = 0; v < ST_BLOCK / 64; ++v) {
              const uint32_t c = woff[u][v];
              woff[u][v] = tot;
              tot += c;
            }
          nhit_s = tot;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SB_CHUNK; ++u) {
          const bool h = (hits >> u) & 1;
          const uint64_t b = __ballot(h);
          if (h) {
            const uint32_t pos = woff[u][wave] + below_count(b);
            hcode[pos] = code[u];
            hrow[pos] = (uint16_t)((ch * SB_CHUNK + u) * ST_BLOCK + threadIdx.x);
          }
        }
        __syncthreads();
        const uint32_t nh = nhit_s;
        for (uint32_t e0 = 0; e0 < nh; e0 += ST_BLOCK) {
          const uint32_t e = e0 + threadIdx.x;
          uint32_t g = 0xFFFFFFFFu;
          int64_t v = 0;
          if (e < nh) {
            const uint32_t local = r0 + hrow[e];
            g = (uint32_t)bm_rank(qq.mbits, qq.rank, hcode[e]);
            const int64_t x = decode_row(t.words[xcol], t.segs[xcol][seg], local);
            const int64_t y = decode_row(t.words[ycol], t.segs[ycol][seg], local);
            v = wmul((int64_t)((uint64_t)qq.a0 + (uint64_t)wmul(qq.a1, x)),
                     (int64_t)((uint64_t)qq.b0 + (uint64_t)wmul(qq.b1, y)));
          }
          if (__ballot(g != 0xFFFFFFFFu) == 0) continue;
          const uint32_t prev = __shfl_up(g, 1, 64);
          bool f = lane == 0 || prev != g;
          uint32_t c = g != 0xFFFFFFFFu;
          for (int o = 1; o < 64; o <<= 1) {
            const int64_t vu = __shfl_up(v, o, 64);
            const uint32_t cu = __shfl_up(c, o, 64);
            const bool fu = __shfl_up((int)f, o, 64) != 0;
            if (lane >= (unsigned)o && !f) {
              v = (int64_t)((uint64_t)v + (uint64_t)vu);
              c += cu;
            }
            if (lane >= (unsigned)o) f = f || fu;
          }
          const uint32_t next = __shfl_down(g, 1, 64);
          if (g != 0xFFFFFFFFu && (lane == 63 || next != g)) {
            atomicAdd(&qq.sum[g], (unsigned long long)v);
            atomicAdd(&qq.cnt[g], c);
            nmatch += c;
          }
        }
        __syncthreads();  // the list is rewritten by the next chunk
      }
    }
    block_count(nprobe, nprobe_out + (uint64_t)qi * gridDim.x);
    block_count(nmatch, nmatch_out + (uint64_t)qi * gridDim.x);
  }
}

// Union of the range columns of one table over the queries, then `key`; the first two with
// at most 31 bits are staged by the kernels (the rest are read from HBM).
static void sb_want(const aln_star_args *a, int nq, int table, int key, SBScan *s) {
  std::vector<int> w;
  for (int i = 0; i < nq; ++i) {
    const aln_table_scan &t = table == 0 ? a[i].dim : table == 1 ? a[i].mid : a[i].fact;
    for (int k = 0; k < t.nranges; ++k)
      if (std::find(w.begin(), w.end(), t.ranges[k].col) == w.end()) w.push_back(t.ranges[k].col);
  }
  if (key >= 0 && std::find(w.begin(), w.end(), key) == w.end()) w.insert(w.begin() + (w.empty() ? 0 : 1), key);
  s->nwant = (int)std::min<size_t>(w.size(), 2);
  for (int i = 0; i < s->nwant; ++i) s->want[i] = w[i];
}

static int sb_scan(const aln_table_scan &t, SBScan *s, uint64_t *nblocks) {
  ALN_REQUIRE(t.cols && t.ncols >= 1 && t.ncols <= ALN_MAX_COLS && t.seg_rows > 0,
              "aln_star_agg_batch: bad table scan");
  memset(s, 0, sizeof(*s));
  for (int i = 0; i < t.ncols; ++i) {
    ALN_REQUIRE(t.cols[i].nsegs == t.cols[0].nsegs, "aln_star_agg_batch: columns differ in segments");
    s->words[i] = t.cols[i].words;
    s->segs[i] = t.cols[i].segs;
  }
  s->tps = (t.seg_rows + ST_TILE - 1) / ST_TILE;
  *nblocks = (uint64_t)t.cols[0].nsegs * s->tps;
  ALN_REQUIRE(*nblocks < (1ull << 31), "aln_star_agg_batch: grid too large");
  return ALN_OK;
}

static bool same_table(const aln_table_scan &x, const aln_table_scan &y) {
  if (x.ncols != y.ncols || x.seg_rows != y.seg_rows || x.seg_modes || y.seg_modes) return false;
  for (int i = 0; i < x.ncols; ++i)
    if (x.cols[i].words != y.cols[i].words || x.cols[i].segs != y.cols[i].segs ||
        x.cols[i].nsegs != y.cols[i].nsegs)
      return false;
  return true;
}

}  // namespace aln

extern "C" int aln_star_agg_batch(aln_ctx *c, const aln_star_args *a, int nq, aln_star_row *const *h_top,
                                  uint32_t *h_ntop, aln_star_stats *h_stats) {
  ALN_REQUIRE(c && a && h_ntop && h_top, "aln_star_agg_batch: bad arguments");
  ALN_REQUIRE(nq >= 1 && nq <= SB_MAX, "aln_star_agg_batch: %d queries (1..%d)", nq, SB_MAX);
  const aln_star_args &a0 = a[0];
  for (int i = 0; i < nq; ++i) {
    const aln_star_args &x = a[i];
    if (!same_table(x.dim, a0.dim) || !same_table(x.mid, a0.mid) || !same_table(x.fact, a0.fact) ||
        x.d_key != a0.d_key || x.m_key != a0.m_key || x.m_fk != a0.m_fk || x.m_g1 != a0.m_g1 ||
        x.m_g2 != a0.m_g2 || x.f_fk != a0.f_fk || x.f_x != a0.f_x || x.f_y != a0.f_y || (x.flags & 3))
      return set_error(ALN_E_UNSUPPORTED, "aln_star_agg_batch: query %d differs in tables or columns", i);
    ALN_REQUIRE(x.limit <= TK_TILE2 / 2 && (x.limit == 0 || h_top[i]), "aln_star_agg_batch: query %d limit", i);
    const aln_table_scan *ts[3] = {&x.dim, &x.mid, &x.fact};
    for (int k = 0; k < 3; ++k) {
      ALN_REQUIRE(ts[k]->nranges >= 0 && ts[k]->nranges <= ALN_MAX_RANGES, "aln_star_agg_batch: ranges");
      for (int r = 0; r < ts[k]->nranges; ++r)
        ALN_REQUIRE(ts[k]->ranges[r].col >= 0 && ts[k]->ranges[r].col < ts[k]->ncols,
                    "aln_star_agg_batch: range column");
    }
  }
  ALN_REQUIRE(a0.d_key >= 0 && a0.d_key < a0.dim.ncols && a0.m_key >= 0 && a0.m_key < a0.mid.ncols &&
                  a0.m_fk >= 0 && a0.m_fk < a0.mid.ncols && a0.m_g1 < a0.mid.ncols && a0.m_g2 < a0.mid.ncols &&
                  a0.f_fk >= 0 && a0.f_fk < a0.fact.ncols && a0.f_x >= 0 && a0.f_x < a0.fact.ncols &&
                  a0.f_y >= 0 && a0.f_y < a0.fact.ncols,
              "aln_star_agg_batch: column slots");
  ws_reset(c);
  for (int i = 0; i < nq; ++i) h_ntop[i] = 0;
  std::vector<aln_star_stats> st(nq);
  memset(st.data(), 0, sizeof(aln_star_stats) * nq);
  SBScan sd, sm, sa, sf;
  uint64_t dblocks, mblocks, fblocks;
  ALN_TRY(sb_scan(a0.dim, &sd, &dblocks));
  ALN_TRY(sb_scan(a0.mid, &sm, &mblocks));
  ALN_TRY(sb_scan(a0.fact, &sf, &fblocks));
  sa = sm;
  sb_want(a, nq, 0, a0.d_key, &sd);
  sb_want(a, nq, 1, a0.m_fk, &sm);
  sb_want(a, nq, 2, a0.f_fk, &sf);
  sa.nwant = 0;
  for (int col : {a0.m_key, a0.m_g1})
    if (col >= 0 && (sa.nwant == 0 || sa.want[0] != col)) sa.want[sa.nwant++] = col;

  int64_t dkmin, dkmax, mkmin, mkmax, g1min, g1max;
  {
    const aln_col *rc[3] = {&a0.dim.cols[a0.d_key], &a0.mid.cols[a0.m_key],
                            a0.m_g1 >= 0 ? &a0.mid.cols[a0.m_g1] : nullptr};
    int64_t mn[3], mx[3];
    ALN_TRY(cols_range(c, rc, 3, mn, mx));
    dkmin = mn[0], dkmax = mx[0], mkmin = mn[1], mkmax = mx[1], g1min = mn[2], g1max = mx[2];
  }
  if (a0.dim.cols[0].nsegs == 0 || a0.mid.cols[0].nsegs == 0 || dkmax < dkmin || mkmax < mkmin) {
    if (h_stats) memcpy(h_stats, st.data(), sizeof(aln_star_stats) * nq);
    return ALN_OK;
  }
  const uint64_t drange = (uint64_t)dkmax - (uint64_t)dkmin + 1;
  const uint64_t mrange = (uint64_t)mkmax - (uint64_t)mkmin + 1;
  if (drange == 0 || drange > (1ull << 34) || mrange == 0 || mrange > (1ull << 31))
    return set_error(ALN_E_UNSUPPORTED, "aln_star_agg_batch: key ranges too wide for bitmaps");
  const uint64_t dwords = (drange + 31) / 32;
  const uint64_t mwords = (mrange + 31) / 32, mblk = (mwords + BM_BLOCK_WORDS - 1) / BM_BLOCK_WORDS;

  std::vector<SBQuery> hq(nq);
  for (int i = 0; i < nq; ++i) {
    SBQuery &q = hq[i];
    memset(&q, 0, sizeof(q));
    const aln_table_scan *ts[3] = {&a[i].dim, &a[i].mid, &a[i].fact};
    for (int k = 0; k < 3; ++k) {
      q.r[k].nr = ts[k]->nranges;
      for (int r = 0; r < ts[k]->nranges; ++r) {
        q.r[k].col[r] = ts[k]->ranges[r].col;
        q.r[k].lo[r] = ts[k]->ranges[r].lo;
        q.r[k].hi[r] = ts[k]->ranges[r].hi;
      }
    }
    q.a0 = a[i].a0;
    q.a1 = a[i].a1;
    q.b0 = a[i].b0;
    q.b1 = a[i].b1;
    ALN_TRY(ws(c, dwords, &q.dbm));
    ALN_TRY(fill_u32(c, q.dbm, dwords, 0));
    ALN_TRY(ws(c, mblk * BM_BLOCK_WORDS, &q.mbits));
    ALN_TRY(fill_u32(c, q.mbits, mblk * BM_BLOCK_WORDS, 0));
    ALN_TRY(ws(c, mblk, &q.rank));
  }
  SBQuery *dq;
  uint32_t *dcnt, *mprobe, *mpass, *fprobe, *fmatch, *dup;
  uint64_t *ctr;
  ALN_TRY(ws(c, nq, &dq));
  ALN_TRY(ws(c, dblocks * nq, &dcnt));
  ALN_TRY(ws(c, mblocks * nq, &mprobe));
  ALN_TRY(ws(c, mblocks * nq, &mpass));
  ALN_TRY(ws(c, fblocks * nq, &fprobe));
  ALN_TRY(ws(c, fblocks * nq, &fmatch));
  ALN_TRY(ws(c, 2 * SB_MAX, &dup));
  ALN_TRY(ws(c, 5 * SB_MAX, &ctr));
  ALN_TRY(fill_u32(c, dup, 2 * SB_MAX, 0));
  ALN_TRY(fill_u64(c, ctr, 5 * SB_MAX, 0));
  ALN_TRY(copy_to_device(c, dq, hq.data(), sizeof(SBQuery) * nq));
  {
    ProfScope p(c, "star_dim");
    hipLaunchKernelGGL(k_sb_dim, dim3((unsigned)dblocks), dim3(ST_BLOCK), 0, c->stream, sd, dq, nq, a0.d_key, dkmin,
                       drange, dcnt, dup);
  }
  ALN_TRY(check_launch("sb_dim"));
  {
    ProfScope p(c, "star_mid");
    hipLaunchKernelGGL(k_sb_mid, dim3((unsigned)mblocks), dim3(ST_BLOCK), 0, c->stream, sm, dq, nq, a0.m_key,
                       a0.m_fk, dkmin, drange, mkmin, mrange, mprobe, mpass, dup + SB_MAX);
  }
  ALN_TRY(check_launch("sb_mid"));
  for (int i = 0; i < nq; ++i) {
    ALN_TRY(sum_u32(c, dcnt + (uint64_t)i * dblocks, dblocks, ctr + 5 * i + 0));
    ALN_TRY(sum_u32(c, mprobe + (uint64_t)i * mblocks, mblocks, ctr + 5 * i + 1));
    ALN_TRY(sum_u32(c, mpass + (uint64_t)i * mblocks, mblocks, ctr + 5 * i + 2));
  }
  std::vector<uint64_t> hc(5 * SB_MAX);
  uint32_t hdup[2 * SB_MAX];
  ALN_TRY(copy_to_host(c, hc.data(), ctr, 8 * 5 * SB_MAX));
  ALN_TRY(copy_to_host(c, hdup, dup, 4 * 2 * SB_MAX));
  for (int i = 0; i < nq; ++i) {
    if (hdup[i]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg_batch: duplicate dimension keys");
    if (hdup[SB_MAX + i]) return set_error(ALN_E_UNSUPPORTED, "aln_star_agg_batch: duplicate mid keys");
    st[i].dim_pass = hc[5 * i + 0];
    st[i].mid_probe = hc[5 * i + 1];
    st[i].mid_pass = hc[5 * i + 2];
  }
  // per query: rank directory and group arrays sized to its passing mid rows
  uint32_t *bcnt;
  ALN_TRY(ws(c, mblk, &bcnt));
  uint64_t mmax = 0;
  for (int i = 0; i < nq; ++i) {
    SBQuery &q = hq[i];
    const uint64_t m = st[i].mid_pass ? st[i].mid_pass : 1;
    mmax = std::max(mmax, m);
    {
      ProfScope p(c, "star_build");
      hipLaunchKernelGGL(k_bm_block_counts, dim3(grid_for(mblk, 256, 65536)), dim3(256), 0, c->stream, q.mbits,
                         mwords, mblk, bcnt);
    }
    ALN_TRY(check_launch("bm_block_counts"));
    ALN_TRY(exclusive_scan_u32(c, bcnt, mblk, q.rank, nullptr));
    ALN_TRY(ws(c, m, &q.sum));
    ALN_TRY(ws(c, m, &q.cnt));
    ALN_TRY(ws(c, m, &q.okey));
    ALN_TRY(ws(c, m, &q.og1));
    ALN_TRY(ws(c, m, &q.og2));
    ALN_TRY(fill_u64(c, (uint64_t *)q.sum, m, 0));
    ALN_TRY(fill_u32(c, q.cnt, m, 0));
  }
  ALN_TRY(copy_to_device(c, dq, hq.data(), sizeof(SBQuery) * nq));
  {
    ProfScope p(c, "star_build");
    hipLaunchKernelGGL(k_sb_attr, dim3((unsigned)mblocks), dim3(ST_BLOCK), 0, c->stream, sa, dq, nq, a0.m_key,
                       a0.m_g1, a0.m_g2, mkmin, mrange);
  }
  ALN_TRY(check_launch("sb_attr"));
  {
    ProfScope p(c, "star_fact");
    hipLaunchKernelGGL(k_sb_fact, dim3((unsigned)fblocks), dim3(ST_BLOCK), 0, c->stream, sf, dq, nq, a0.f_fk,
                       a0.f_x, a0.f_y, mkmin, mrange, fprobe, fmatch);
  }
  ALN_TRY(check_launch("sb_fact"));
  for (int i = 0; i < nq; ++i) {
    ALN_TRY(sum_u32(c, fprobe + (uint64_t)i * fblocks, fblocks, ctr + 5 * i + 3));
    ALN_TRY(sum_u32(c, fmatch + (uint64_t)i * fblocks, fblocks, ctr + 5 * i + 4));
  }
  for (int i = 0; i < nq; ++i)
    if (a[i].limit && st[i].mid_pass)
      ALN_TRY(star_topk(c, st[i].mid_pass, hq[i].sum, hq[i].cnt,
                        GroupAttr{hq[i].okey, hq[i].og1, hq[i].og2, nullptr, nullptr, 0}, a[i].limit, g1min, g1max,
                        mkmin, h_top[i], &h_ntop[i]));
  ALN_TRY(copy_to_host(c, hc.data(), ctr, 8 * 5 * SB_MAX));
  for (int i = 0; i < nq; ++i) {
    st[i].fact_probe = hc[5 * i + 3];
    st[i].fact_pass = hc[5 * i + 4];
  }
  ALN_HIP(hipStreamSynchronize(c->stream));
  if (h_stats) memcpy(h_stats, st.data(), sizeof(aln_star_stats) * nq);
  return ALN_OK;
}

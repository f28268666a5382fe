// Column codec and row-wise kernels over the segmented, bit-packed column layout:
// encoding (pfor_compress, compress.cu:411-527, as dense bit-packing), decoding
// (pfor_decompress, compress.cu:180-267), gathers (mygather, cm.cu:3350-3413), expression
// evaluation (select.cu:703-790) and the fused scan: decode + predicate + stream
// compaction (filter_op, cm.cu:3609-3666; filter, filter.cu:133-643).
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "aln_internal.h"
#include "vm.h"

namespace aln {

// ---- encoding -----------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_seg_minmax(const int64_t *__restrict__ v, uint64_t nrows,
                                                    uint32_t seg_rows, uint32_t blocks_per_seg,
                                                    unsigned long long *mn,
                                                    unsigned long long *mx) {
  uint32_t seg = blockIdx.x / blocks_per_seg, part = blockIdx.x % blocks_per_seg;
  uint64_t s0 = (uint64_t)seg * seg_rows;
  uint64_t s1 = s0 + seg_rows < nrows ? s0 + seg_rows : nrows;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  for (uint64_t i = s0 + (uint64_t)part * 256 + threadIdx.x; i < s1;
       i += (uint64_t)blocks_per_seg * 256) {
    int64_t x = v[i];
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    int64_t a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (__lane_id() == 0) {
    atomicMin(&mn[seg], (unsigned long long)lo ^ 0x8000000000000000ull);
    atomicMax(&mx[seg], (unsigned long long)hi ^ 0x8000000000000000ull);
  }
}

// One thread per output word: gathers the bits of every code overlapping the word.
__global__ void __launch_bounds__(256) k_pack(const int64_t *__restrict__ v, uint64_t nrows,
                                              uint32_t seg_rows, uint32_t seg0,
                                              const aln_seg *__restrict__ segs,
                                              uint32_t *__restrict__ words) {
  uint32_t seg = seg0 + blockIdx.y;
  aln_seg sg = segs[seg];
  if (sg.bits == 0) return;
  uint64_t nw = ((uint64_t)sg.nrows * sg.bits + 31) / 32;
  const int64_t *sv = v + (uint64_t)seg * seg_rows;
  for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < nw;
       w += (uint64_t)gridDim.x * 256) {
    uint64_t b0 = w * 32, b1 = b0 + 32;
    uint64_t r = b0 / sg.bits;
    uint32_t acc = 0;
    for (; r < sg.nrows && r * sg.bits < b1; ++r) {
      uint64_t code = (uint64_t)sv[r] - (uint64_t)sg.base;
      int64_t sh = (int64_t)(r * sg.bits) - (int64_t)b0;  // bit position of code bit 0
      if (sh >= 0)
        acc |= (uint32_t)(code << sh);
      else
        acc |= (uint32_t)(code >> (-sh));
    }
    words[sg.word_off + w] = acc;
  }
}

// ---- row-wise kernels ---------------------------------------------------------------------

// Rows outside the column (a segment past nsegs, or past a short last segment) read as 0.
__global__ void __launch_bounds__(256) k_decode(const uint32_t *__restrict__ words,
                                                const aln_seg *__restrict__ segs, uint32_t nsegs,
                                                uint32_t seg_rows, uint64_t r0, uint64_t n,
                                                int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    uint64_t r = r0 + i;
    uint64_t seg = r / seg_rows, local = r - seg * seg_rows;
    out[i] = (seg < nsegs && local < segs[seg].nrows) ? decode_row(words, segs[seg], local) : 0;
  }
}

__global__ void __launch_bounds__(256) k_gather(const uint32_t *__restrict__ words,
                                                const aln_seg *__restrict__ segs, uint32_t nsegs,
                                                uint32_t seg_rows,
                                                const uint32_t *__restrict__ rows, uint64_t n,
                                                int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    uint32_t r = rows ? rows[i] : (uint32_t)i;
    uint32_t seg = r / seg_rows, local = r - seg * seg_rows;
    out[i] = (seg < nsegs && local < segs[seg].nrows) ? decode_row(words, segs[seg], local) : 0;
  }
}

__global__ void __launch_bounds__(VM_BLOCK) k_eval(const ColTable *__restrict__ ct,
                                                   const aln_insn *__restrict__ prog, int nprog,
                                                   uint32_t nsegs, uint32_t seg_rows,
                                                   const uint32_t *__restrict__ rows, uint64_t n,
                                                   int64_t *__restrict__ out) {
  __shared__ int64_t stk[ALN_MAX_STACK * VM_BLOCK];
  for (uint64_t i0 = (uint64_t)blockIdx.x * VM_BLOCK; i0 < n; i0 += (uint64_t)gridDim.x * VM_BLOCK) {
    uint64_t i = i0 + threadIdx.x;
    if (i < n) {
      uint32_t r = rows ? rows[i] : (uint32_t)i;
      uint32_t seg = r / seg_rows, local = r - seg * seg_rows;
      bool inside = seg < nsegs && local < ct->segs[0][seg < nsegs ? seg : 0].nrows;
      auto load = [&](int c) -> int64_t {
        return inside ? decode_row(ct->words[c], ct->segs[c][seg], local) : 0;
      };
      out[i] = vm_run(prog, nprog, load, stk + threadIdx.x);
    }
  }
}

__global__ void __launch_bounds__(256) k_apply_perm(const int64_t *__restrict__ in,
                                                    const uint32_t *__restrict__ perm, uint64_t n,
                                                    int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    out[i] = in[perm[i]];
}

// out[i] = table[codes[i]], -1 for a code outside the table.
__global__ void __launch_bounds__(256) k_dict_lookup(const int64_t *__restrict__ codes, uint64_t n,
                                                     const int64_t *__restrict__ table, uint64_t ntable,
                                                     int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const int64_t c = codes[i];
    out[i] = (c >= 0 && (uint64_t)c < ntable) ? table[c] : -1;
  }
}

// LIKE on one dictionary entry (the reference's gpu_regex, filter.cu:67-129): '%' matches any
// run, retrying from the character after the last attempt's start on a mismatch; '?' any
// character but '.'; others themselves.  The entry is `len` bytes, NUL-padded, and the
// position counter is shared by all attempts and bounded by len, as in the reference.
__device__ bool like_entry(const char *value, uint32_t len, const char *pattern) {
  bool after_pct = false;
  uint32_t used = 0;
  const char *from = value, *pat = pattern;
  for (;;) {
    const char *s = from, *p = pat;
    int stop = 0;  // 0: the value or the budget ran out, 1: '%' consumed, 2: mismatch
    for (; used < len && *s; ++s, ++p, ++used) {
      if (*p == '%') {
        after_pct = true;
        from = s;
        pat = p;
        while (*++pat == '%') {
        }
        if (!*pat) return true;
        stop = 1;
        break;
      }
      if (*p == '?' ? *s == '.' : *s != *p) {
        stop = 2;
        break;
      }
    }
    if (stop == 1) continue;
    if (stop == 2) {
      if (!after_pct) return false;
      ++from;
      ++used;
      continue;
    }
    while (*p == '%') ++p;
    return !*p;
  }
}

// One thread per entry of a fixed-width dictionary, matched where it lies: the matcher reads
// a value character only while its position counter is below the width, and the counter is
// never behind the character's offset, so no read leaves the entry.  A wave covers 64
// entries, so its two bit-set words are written whole from one ballot.
constexpr uint32_t LIKE_MAX_WIDTH = 256;
__global__ void __launch_bounds__(64) k_like_dict(const char *__restrict__ entries, uint64_t n, uint32_t width,
                                                  const char *__restrict__ pattern, uint32_t *__restrict__ bits) {
  const uint64_t i = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  bool m = false;
  if (i < n) m = like_entry(entries + i * width, width, pattern);
  const uint64_t b = __builtin_amdgcn_ballot_w64(m);
  const uint64_t w = (uint64_t)blockIdx.x * 2;
  if (threadIdx.x == 0) bits[w] = (uint32_t)b;
  if (threadIdx.x == 32) bits[w + 1] = (uint32_t)(b >> 32);
}

// ---- fused scan: decode + predicate + compaction -------------------------------------------
constexpr int SC_BLOCK = VM_BLOCK;  // 256
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;

// One term of a conjunction the scan's fast path evaluates on column col: lo <= v <= hi, or
// (set != 0) membership of v in the bit set at device address lo over codes 0..hi-1.
struct RangeC {
  int32_t col;
  int32_t set;
  int64_t lo, hi;
};

// PASS 0: count passing rows per tile.  PASS 1: write ascending global row ids.
template <int PASS, bool FAST>
__global__ void __launch_bounds__(SC_BLOCK) k_scan(const ColTable *__restrict__ ct,
                                                   const aln_insn *__restrict__ prog, int nprog,
                                                   const RangeC *__restrict__ rg, int nrg,
                                                   uint32_t seg_rows, uint32_t tiles_per_seg,
                                                   const char *__restrict__ modes,
                                                   uint32_t *__restrict__ counts,
                                                   const uint64_t *__restrict__ offs,
                                                   uint32_t *__restrict__ out) {
  __shared__ int64_t stk[FAST ? 1 : ALN_MAX_STACK * SC_BLOCK];
  __shared__ uint32_t wc[SC_BLOCK / 64];
  const uint32_t seg = blockIdx.x / tiles_per_seg, tile = blockIdx.x % tiles_per_seg;
  const uint32_t nrows = ct->segs[0][seg].nrows;
  const uint32_t r0 = tile * SC_TILE;
  const char mode = modes ? modes[seg] : 'R';
  if (r0 >= nrows || mode == 'N') {
    if (PASS == 0 && threadIdx.x == 0) counts[blockIdx.x] = 0;
    return;
  }
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  const uint64_t gbase = (uint64_t)seg * seg_rows;
  uint64_t run = PASS ? offs[blockIdx.x] : 0;
  uint32_t total = 0;
  for (int j = 0; j < SC_ITEMS; ++j) {
    uint32_t local = r0 + j * SC_BLOCK + threadIdx.x;
    bool pass = false;
    if (local < nrows) {
      if (mode == 'A') {
        pass = true;
      } else if (FAST) {
        pass = true;
        for (int k = 0; k < nrg; ++k) {
          const int c = rg[k].col;
          int64_t v = decode_row(ct->words[c], ct->segs[c][seg], local);
          if (rg[k].set) {
            const uint32_t *bs = reinterpret_cast<const uint32_t *>(rg[k].lo);
            pass = pass && v >= 0 && v < rg[k].hi && ((bs[v >> 5] >> (v & 31)) & 1u);
          } else {
            pass = pass && v >= rg[k].lo && v <= rg[k].hi;
          }
        }
      } else {
        auto load = [&](int c) -> int64_t { return decode_row(ct->words[c], ct->segs[c][seg], local); };
        pass = vm_run(prog, nprog, load, stk + threadIdx.x) != 0;
      }
    }
    uint64_t bal = __ballot(pass);
    if (PASS == 0) {
      total += __popcll(bal);
    } else {
      if (lane == 0) wc[wid] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t before = 0, all = 0;
      for (int w = 0; w < SC_BLOCK / 64; ++w) {
        uint32_t x = wc[w];
        before += w < (int)wid ? x : 0;
        all += x;
      }
      if (pass) out[run + before + below_count(bal)] = (uint32_t)(gbase + local);
      run += all;
      __syncthreads();
    }
  }
  if (PASS == 0) {
    if (lane == 0) wc[wid] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int w = 0; w < SC_BLOCK / 64; ++w) s += wc[w];
      counts[blockIdx.x] = s;
    }
  }
}

// Recognise a conjunction of "column OP constant" terms (multiplier 1) and code-set tests
// (LIKE) and turn it into one closed interval per column plus the set tests.  Returns false
// when the program has another shape.
// At most one interval per column plus one set test per two instructions.
constexpr int MAX_RANGES_C = ALN_MAX_COLS + ALN_MAX_INSN / 2;
static bool as_intervals(const aln_insn *p, int n, RangeC *out, int *nout, bool *empty) {
  *empty = false;
  *nout = 0;
  int64_t lo[ALN_MAX_COLS], hi[ALN_MAX_COLS];
  bool used[ALN_MAX_COLS] = {false};
  RangeC sets[ALN_MAX_INSN / 2];
  int nsets = 0;
  int i = 0;
  int pending = 0;  // predicates on the evaluation stack
  while (i < n) {
    if (p[i].op == ALN_I_AND) {
      if (pending < 2) return false;
      --pending;
      ++i;
      continue;
    }
    if (i + 1 < n && p[i].op == ALN_I_COL && p[i].k == 1 && p[i + 1].op == ALN_I_INSET) {
      if (nsets == ALN_MAX_INSN / 2 || p[i].a < 0 || p[i].a >= ALN_MAX_COLS) return false;
      sets[nsets++] = RangeC{p[i].a, 1, p[i + 1].k, (int64_t)p[i + 1].a};
      ++pending;
      i += 2;
      continue;
    }
    if (i + 2 >= n) return false;
    if (!(p[i].op == ALN_I_COL && p[i].k == 1 && p[i + 1].op == ALN_I_CONST &&
          p[i + 2].op == ALN_I_CMP))
      return false;
    int c = p[i].a;
    if (c < 0 || c >= ALN_MAX_COLS) return false;
    int64_t v = p[i + 1].k, l = INT64_MIN, h = INT64_MAX;
    switch (p[i + 2].a) {
      case 1:
        if (v == INT64_MIN) *empty = true; else h = v - 1;
        break;
      case 2:
        if (v == INT64_MAX) *empty = true; else l = v + 1;
        break;
      case 4: l = h = v; break;
      case 5: h = v; break;
      case 6: l = v; break;
      default: return false;
    }
    if (!used[c]) {
      used[c] = true;
      lo[c] = l;
      hi[c] = h;
    } else {
      lo[c] = l > lo[c] ? l : lo[c];
      hi[c] = h < hi[c] ? h : hi[c];
    }
    ++pending;
    i += 3;
  }
  if (pending != 1) return false;
  int k = 0;
  for (int c = 0; c < ALN_MAX_COLS; ++c)
    if (used[c]) {
      if (lo[c] > hi[c]) *empty = true;
      out[k++] = RangeC{c, 0, lo[c], hi[c]};
    }
  for (int j = 0; j < nsets; ++j) out[k++] = sets[j];
  *nout = k;
  return true;
}

static int fill_cols(const aln_col *cols, int ncols, ColTable *h) {
  ALN_REQUIRE(ncols >= 1 && ncols <= ALN_MAX_COLS, "column count %d outside 1..%d", ncols,
              ALN_MAX_COLS);
  memset(h, 0, sizeof(*h));
  for (int i = 0; i < ncols; ++i) {
    ALN_REQUIRE(cols[i].words || cols[i].nsegs == 0, "column %d has no words", i);
    ALN_REQUIRE(cols[i].segs, "column %d has no segment descriptors", i);
    ALN_REQUIRE(cols[i].nsegs == cols[0].nsegs, "columns differ in segment count");
    h->words[i] = cols[i].words;
    h->segs[i] = cols[i].segs;
  }
  return ALN_OK;
}

static int check_prog(const aln_insn *prog, int nprog) {
  ALN_REQUIRE(prog && nprog >= 1 && nprog <= ALN_MAX_INSN, "program length %d", nprog);
  int depth = 0, mx = 0;
  for (int i = 0; i < nprog; ++i) {
    int op = prog[i].op;
    ALN_REQUIRE(op >= ALN_I_COL && op <= ALN_I_FCMP, "bad instruction %d", op);
    if (op == ALN_I_COL) ALN_REQUIRE(prog[i].a >= 0 && prog[i].a < ALN_MAX_COLS, "bad column %d", prog[i].a);
    if (op == ALN_I_INSET)
      ALN_REQUIRE(prog[i].a >= 0 && (prog[i].a == 0 || prog[i].k != 0), "set test without a bit set");
    if (op == ALN_I_I2F) ALN_REQUIRE((prog[i].a == 0 || prog[i].a == 1) && depth > prog[i].a, "bad conversion");
    depth += (op == ALN_I_COL || op == ALN_I_CONST) ? 1 : (op == ALN_I_INSET || op == ALN_I_I2F) ? 0 : -1;
    ALN_REQUIRE(depth >= 1, "program underflows its stack");
    mx = depth > mx ? depth : mx;
  }
  ALN_REQUIRE(depth == 1 && mx <= ALN_MAX_STACK, "program leaves %d values (max depth %d)", depth, mx);
  return ALN_OK;
}

int upload_cols(aln_ctx *c, const aln_col *cols, int ncols, ColTable **out) {
  ColTable h;
  ALN_TRY(fill_cols(cols, ncols, &h));
  ColTable *d;
  ALN_TRY(ws(c, 1, &d));
  ALN_HIP(hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  *out = d;
  return ALN_OK;
}

int upload_prog(aln_ctx *c, const aln_insn *prog, int nprog, aln_insn **out) {
  ALN_TRY(check_prog(prog, nprog));
  aln_insn *d;
  ALN_TRY(ws(c, nprog, &d));
  ALN_HIP(hipMemcpyAsync(d, prog, sizeof(aln_insn) * nprog, hipMemcpyHostToDevice, c->stream));
  *out = d;
  return ALN_OK;
}

static int check_prog_cols(const aln_insn *prog, int nprog, int ncols) {
  for (int i = 0; i < nprog; ++i)
    if (prog[i].op == ALN_I_COL && prog[i].a >= ncols)
      return set_error(ALN_E_ARG, "program reads column %d of %d", prog[i].a, ncols);
  return ALN_OK;
}

// Segment of the source row behind each index: r = rows ? rows[idx[i]] : idx[i]; the last
// segment start <= r when starts are given (uneven file segments), else r / seg_rows.
__global__ void __launch_bounds__(256) k_row_segments(const uint32_t *__restrict__ idx, uint64_t n,
                                                      const uint32_t *__restrict__ rows,
                                                      const uint64_t *__restrict__ starts, uint32_t nstarts,
                                                      uint32_t seg_rows, int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t r = rows ? rows[idx[i]] : idx[i];
    int64_t seg;
    if (starts) {  // upper bound of r in starts[0, nstarts), minus one
      uint32_t lo = 0, hi = nstarts;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= r) lo = mid + 1; else hi = mid;
      }
      seg = (int64_t)lo - 1;
    } else {
      seg = (int64_t)(r / seg_rows);
    }
    out[i] = seg;
  }
}

// ---- fixed scratch of the allocation-free scan -----------------------------------------------
// aln_scan_filter_noalloc keeps its temporaries in this array instead of the context's
// workspace: the column table, program and intervals it uploads, the per-segment zone-map
// verdicts, the per-tile counts, their exclusive scan and the scan's partial sums and total.
// Every byte a call reads is written earlier in the same call, so nothing is cleared.  Each
// GPU holds its own copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.
constexpr uint64_t NA_TILES = ALN_SCAN_NOALLOC_MAX_TILES;
constexpr size_t NA_CT = 0;                                               // ColTable
constexpr size_t NA_PROG = 256;                                           // ALN_MAX_INSN insns
constexpr size_t NA_RANGES = NA_PROG + sizeof(aln_insn) * ALN_MAX_INSN;   // MAX_RANGES_C
constexpr size_t NA_TOTAL = 2560;                                         // u64 scan total
constexpr size_t NA_MODES = 4096;                                         // u8 per segment
constexpr size_t NA_COUNTS = NA_MODES + NA_TILES;                         // u32 per tile
constexpr size_t NA_OFFS = NA_COUNTS + 4 * NA_TILES;                      // u64 per tile
constexpr size_t NA_SUMS = NA_OFFS + 8 * NA_TILES;                        // u64 per scan tile
constexpr size_t NA_BYTES = NA_SUMS + 8 * ((NA_TILES + 4095) / 4096 + 32);
static_assert(sizeof(ColTable) <= NA_PROG, "column table overlaps the program");
static_assert(NA_RANGES + sizeof(RangeC) * MAX_RANGES_C <= NA_TOTAL, "intervals overlap the total");
static_assert(NA_TILES % 256 == 0, "scratch regions lose their alignment");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static std::atomic<char *> g_na_ptr[NA_MAX_DEVICES];

int scan_scratch_attach(int device) {
  ALN_REQUIRE(device >= 0 && device < NA_MAX_DEVICES, "device %d outside 0..%d", device,
              NA_MAX_DEVICES - 1);
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (g_na_ptr[device].load()) return ALN_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return set_error(ALN_E_HIP, "hipGetSymbolAddress of the scan scratch on device %d failed: %s", device,
                     hipGetErrorString(e));
  }
  g_na_ptr[device].store(static_cast<char *>(p));
  return ALN_OK;
}

// The scratch address on the context's device, looked up now (with that device current for
// the call) if the context's creation could not.
static int scan_scratch(aln_ctx *c, char **out) {
  ALN_REQUIRE(c->device >= 0 && c->device < NA_MAX_DEVICES, "device %d", c->device);
  char *p = g_na_ptr[c->device].load();
  if (!p) {
    int cur = -1;
    ALN_HIP(hipGetDevice(&cur));
    if (cur != c->device) ALN_HIP(hipSetDevice(c->device));
    int r = scan_scratch_attach(c->device);
    if (cur != c->device) ALN_HIP(hipSetDevice(cur));
    ALN_TRY(r);
    p = g_na_ptr[c->device].load();
  }
  *out = p;
  return ALN_OK;
}

}  // namespace aln

using namespace aln;

extern "C" {

int aln_encode_plan(aln_ctx *c, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                    uint32_t force_bits, aln_seg *h_segs, uint64_t *h_words) {
  ALN_REQUIRE(c && h_segs && h_words && seg_rows > 0, "aln_encode_plan: bad arguments");
  ALN_REQUIRE(nrows == 0 || d_vals, "aln_encode_plan: no values");
  ALN_REQUIRE(force_bits <= 64, "aln_encode_plan: force_bits %u", force_bits);
  ws_reset(c);
  uint32_t nsegs = (uint32_t)((nrows + seg_rows - 1) / seg_rows);
  *h_words = 0;
  if (nsegs == 0) return ALN_OK;
  unsigned long long *mm;
  ALN_TRY(ws(c, 2 * (size_t)nsegs, &mm));
  std::vector<unsigned long long> init(2 * (size_t)nsegs);
  for (uint32_t s = 0; s < nsegs; ++s) {
    init[s] = ~0ull;
    init[nsegs + s] = 0;
  }
  ALN_HIP(hipMemcpyAsync(mm, init.data(), 16 * (size_t)nsegs, hipMemcpyHostToDevice, c->stream));
  uint32_t bps = (uint32_t)grid_for(seg_rows, 256 * 16, 256);
  {
    ProfScope p(c, "seg_minmax");
    hipLaunchKernelGGL(k_seg_minmax, dim3(nsegs * bps), dim3(256), 0, c->stream, d_vals, nrows,
                       seg_rows, bps, mm, mm + nsegs);
  }
  ALN_TRY(check_launch("seg_minmax"));
  ALN_TRY(copy_to_host(c, init.data(), mm, 16 * (size_t)nsegs));
  uint64_t off = 0;
  for (uint32_t s = 0; s < nsegs; ++s) {
    aln_seg &g = h_segs[s];
    uint64_t r0 = (uint64_t)s * seg_rows;
    g.nrows = (uint32_t)((r0 + seg_rows <= nrows) ? seg_rows : nrows - r0);
    g.vmin = (int64_t)(init[s] ^ 0x8000000000000000ull);
    g.vmax = (int64_t)(init[nsegs + s] ^ 0x8000000000000000ull);
    g.base = g.vmin;
    g.bits = bits_for_range_u((uint64_t)g.vmax - (uint64_t)g.vmin);
    if (force_bits) {
      ALN_REQUIRE(force_bits >= g.bits, "aln_encode_plan: segment %u needs %u bits > %u", s,
                  g.bits, force_bits);
      g.bits = force_bits;
      if (force_bits == 64) g.base = 0;
    }
    g.word_off = off;
    off += ((uint64_t)g.nrows * g.bits + 31) / 32;
  }
  *h_words = off;
  return ALN_OK;
}

int aln_encode_pack(aln_ctx *c, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                    const aln_seg *h_segs, uint32_t *d_words, aln_seg *d_segs) {
  ALN_REQUIRE(c && h_segs && d_segs && seg_rows > 0, "aln_encode_pack: bad arguments");
  ws_reset(c);
  uint32_t nsegs = (uint32_t)((nrows + seg_rows - 1) / seg_rows);
  if (nsegs == 0) return ALN_OK;
  uint64_t maxw = 0;
  for (uint32_t s = 0; s < nsegs; ++s) {
    uint64_t r0 = (uint64_t)s * seg_rows;
    uint32_t nr = (uint32_t)((r0 + seg_rows <= nrows) ? seg_rows : nrows - r0);
    ALN_REQUIRE(h_segs[s].nrows == nr && h_segs[s].bits <= 64,
                "aln_encode_pack: descriptor %u does not match the plan", s);
    uint64_t w = ((uint64_t)nr * h_segs[s].bits + 31) / 32;
    maxw = w > maxw ? w : maxw;
  }
  ALN_REQUIRE(maxw == 0 || d_words, "aln_encode_pack: no word buffer");
  ALN_HIP(hipMemcpyAsync(d_segs, h_segs, sizeof(aln_seg) * nsegs, hipMemcpyHostToDevice, c->stream));
  if (maxw) {
    ProfScope p(c, "pack");
    for (uint32_t s0 = 0; s0 < nsegs; s0 += 65535) {
      uint32_t ns = nsegs - s0 < 65535 ? nsegs - s0 : 65535;
      hipLaunchKernelGGL(k_pack, dim3(grid_for(maxw, 256, 4096), ns), dim3(256), 0, c->stream,
                         d_vals, nrows, seg_rows, s0, d_segs, d_words);
    }
  }
  ALN_TRY(check_launch("pack"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_plain_col(aln_ctx *c, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                  int32_t decimals, aln_seg *d_segs, aln_col *out) {
  ALN_REQUIRE(c && d_segs && out && seg_rows > 0, "aln_plain_col: bad arguments");
  uint32_t nsegs = (uint32_t)((nrows + seg_rows - 1) / seg_rows);
  std::vector<aln_seg> h(nsegs);
  for (uint32_t s = 0; s < nsegs; ++s) {
    uint64_t r0 = (uint64_t)s * seg_rows;
    h[s].nrows = (uint32_t)((r0 + seg_rows <= nrows) ? seg_rows : nrows - r0);
    h[s].word_off = 2 * r0;
    h[s].base = 0;
    h[s].bits = 64;
    h[s].vmin = INT64_MIN;
    h[s].vmax = INT64_MAX;
  }
  if (nsegs) ALN_TRY(copy_to_device(c, d_segs, h.data(), sizeof(aln_seg) * nsegs));
  out->words = (const uint32_t *)d_vals;
  out->segs = d_segs;
  out->nsegs = nsegs;
  out->decimals = decimals;
  return ALN_OK;
}

int aln_decode(aln_ctx *c, const aln_col *col, uint32_t seg_rows, uint64_t row_begin,
               uint64_t row_end, int64_t *d_out) {
  ALN_REQUIRE(c && col && seg_rows > 0 && row_end >= row_begin, "aln_decode: bad arguments");
  ALN_REQUIRE(row_end <= (uint64_t)col->nsegs * seg_rows, "aln_decode: rows beyond the column");
  uint64_t n = row_end - row_begin;
  if (!n) return ALN_OK;
  ProfScope p(c, "decode");
  hipLaunchKernelGGL(k_decode, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, col->words,
                     col->segs, col->nsegs, seg_rows, row_begin, n, d_out);
  return check_launch("decode");
}

int aln_gather(aln_ctx *c, const aln_col *col, uint32_t seg_rows, const uint32_t *d_rows,
               uint64_t n, int64_t *d_out) {
  ALN_REQUIRE(c && col && seg_rows > 0, "aln_gather: bad arguments");
  if (!n) return ALN_OK;
  ProfScope p(c, "gather");
  if (col->nsegs == 0) return fill_i64(c, d_out, n, 0);
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, col->words,
                     col->segs, col->nsegs, seg_rows, d_rows, n, d_out);
  return check_launch("gather");
}

int aln_eval(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows,
             const aln_insn *prog, int nprog, const uint32_t *d_rows, uint64_t n,
             int64_t *d_out) {
  ALN_REQUIRE(c && cols && seg_rows > 0, "aln_eval: bad arguments");
  ws_reset(c);
  if (!n) return ALN_OK;
  ALN_TRY(check_prog_cols(prog, nprog, ncols));
  ColTable *ct;
  aln_insn *dp;
  ALN_TRY(upload_cols(c, cols, ncols, &ct));
  ALN_TRY(upload_prog(c, prog, nprog, &dp));
  ProfScope p(c, "eval");
  if (cols[0].nsegs == 0) return fill_i64(c, d_out, n, 0);
  hipLaunchKernelGGL(k_eval, dim3(grid_for(n, VM_BLOCK, 65536)), dim3(VM_BLOCK), 0, c->stream, ct,
                     dp, nprog, cols[0].nsegs, seg_rows, d_rows, n, d_out);
  return check_launch("eval");
}

int aln_apply_perm(aln_ctx *c, const int64_t *d_in, const uint32_t *d_perm, uint64_t n,
                   int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || (d_in && d_perm && d_out)), "aln_apply_perm: bad arguments");
  if (!n) return ALN_OK;
  ProfScope p(c, "apply_perm");
  hipLaunchKernelGGL(k_apply_perm, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_in,
                     d_perm, n, d_out);
  return check_launch("apply_perm");
}

int aln_like_dict(aln_ctx *c, const char *d_entries, uint64_t n, uint32_t width, const char *h_pattern,
                  uint32_t *d_bits) {
  ALN_REQUIRE(c && h_pattern && (n == 0 || (d_entries && d_bits)), "aln_like_dict: bad arguments");
  ALN_REQUIRE(width >= 1 && width <= LIKE_MAX_WIDTH, "aln_like_dict: width %u outside 1..%u", width, LIKE_MAX_WIDTH);
  if (!n) return ALN_OK;
  ALN_REQUIRE(n < (1ull << 37), "aln_like_dict: %llu entries", (unsigned long long)n);
  ws_reset(c);
  const size_t plen = strlen(h_pattern) + 1;
  char *dp;
  ALN_TRY(ws(c, plen, &dp));
  ALN_TRY(copy_to_device(c, dp, h_pattern, plen));
  ProfScope p(c, "like_dict");
  hipLaunchKernelGGL(k_like_dict, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, c->stream, d_entries, n, width,
                     (const char *)dp, d_bits);
  ALN_TRY(check_launch("like_dict"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_dict_lookup(aln_ctx *c, const int64_t *d_codes, uint64_t n, const int64_t *d_table, uint64_t ntable,
                    int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || (d_codes && d_out)) && (ntable == 0 || d_table), "aln_dict_lookup: bad arguments");
  if (!n) return ALN_OK;
  ProfScope p(c, "dict_lookup");
  hipLaunchKernelGGL(k_dict_lookup, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_codes, n, d_table,
                     ntable, d_out);
  return check_launch("dict_lookup");
}

int aln_scan_filter(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows,
                    const aln_insn *prog, int nprog, const char *seg_modes, uint32_t *d_row_ids,
                    uint64_t cap, uint64_t *h_count) {
  ALN_REQUIRE(c && cols && h_count && seg_rows > 0, "aln_scan_filter: bad arguments");
  ws_reset(c);
  ALN_TRY(check_prog_cols(prog, nprog, ncols));
  uint32_t nsegs = cols[0].nsegs;
  ALN_REQUIRE((uint64_t)nsegs * seg_rows <= (1ull << 32), "aln_scan_filter: more than 2^32 rows");
  *h_count = 0;
  if (nsegs == 0) return ALN_OK;
  ColTable *ct;
  aln_insn *dp;
  ALN_TRY(upload_cols(c, cols, ncols, &ct));
  ALN_TRY(upload_prog(c, prog, nprog, &dp));
  RangeC rgs[MAX_RANGES_C];
  int nrg = 0;
  bool empty = false;
  bool fast = as_intervals(prog, nprog, rgs, &nrg, &empty);
  if (fast && empty) return ALN_OK;
  RangeC *drg = nullptr;
  if (fast && nrg > 0) {
    ALN_TRY(ws(c, nrg, &drg));
    ALN_HIP(hipMemcpyAsync(drg, rgs, sizeof(RangeC) * nrg, hipMemcpyHostToDevice, c->stream));
  }
  char *dmodes = nullptr;
  if (seg_modes) {
    ALN_TRY(ws(c, nsegs, &dmodes));
    ALN_HIP(hipMemcpyAsync(dmodes, seg_modes, nsegs, hipMemcpyHostToDevice, c->stream));
  }
  uint32_t tps = (seg_rows + SC_TILE - 1) / SC_TILE;
  uint64_t nblocks = (uint64_t)nsegs * tps;
  ALN_REQUIRE(nblocks < (1ull << 31), "aln_scan_filter: grid too large");
  uint32_t *counts;
  uint64_t *offs;
  ALN_TRY(ws(c, nblocks, &counts));
  ALN_TRY(ws(c, nblocks, &offs));
  {
    ProfScope p(c, "scan_count");
    if (fast)
      hipLaunchKernelGGL((k_scan<0, true>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream,
                         ct, dp, nprog, drg, nrg, seg_rows, tps, dmodes, counts,
                         (const uint64_t *)nullptr, (uint32_t *)nullptr);
    else
      hipLaunchKernelGGL((k_scan<0, false>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream,
                         ct, dp, nprog, drg, 0, seg_rows, tps, dmodes, counts,
                         (const uint64_t *)nullptr, (uint32_t *)nullptr);
  }
  ALN_TRY(check_launch("scan_count"));
  uint64_t total = 0;
  ALN_TRY(exclusive_scan_u32(c, counts, nblocks, offs, &total));
  *h_count = total;
  if (total > cap) return set_error(ALN_E_CAPACITY, "aln_scan_filter: %llu rows pass, capacity %llu",
                                    (unsigned long long)total, (unsigned long long)cap);
  if (total == 0) return ALN_OK;
  ALN_REQUIRE(d_row_ids, "aln_scan_filter: no output buffer");
  {
    ProfScope p(c, "scan_write");
    if (fast)
      hipLaunchKernelGGL((k_scan<1, true>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream,
                         ct, dp, nprog, drg, nrg, seg_rows, tps, dmodes, counts, offs,
                         d_row_ids);
    else
      hipLaunchKernelGGL((k_scan<1, false>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream,
                         ct, dp, nprog, drg, 0, seg_rows, tps, dmodes, counts, offs, d_row_ids);
  }
  ALN_TRY(check_launch("scan_write"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_row_segments(aln_ctx *c, const uint32_t *d_idx, uint64_t n, const uint32_t *d_rows,
                     const uint64_t *d_starts, uint32_t nstarts, uint32_t seg_rows, int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || (d_idx && d_out)), "aln_row_segments: bad arguments");
  ALN_REQUIRE(d_starts ? nstarts > 0 : seg_rows > 0, "aln_row_segments: no segmentation");
  if (n == 0) return ALN_OK;
  hipLaunchKernelGGL(k_row_segments, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_idx, n, d_rows,
                     d_starts, nstarts, seg_rows, d_out);
  return check_launch("row_segments");
}

int aln_scan_filter_noalloc(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows,
                            const aln_insn *prog, int nprog, const char *seg_modes,
                            uint32_t *d_row_ids, uint64_t cap, uint64_t *h_count) {
  ALN_REQUIRE(c && cols && h_count && seg_rows > 0, "aln_scan_filter_noalloc: bad arguments");
  ALN_TRY(check_prog_cols(prog, nprog, ncols));
  ColTable hct;
  ALN_TRY(fill_cols(cols, ncols, &hct));
  ALN_TRY(check_prog(prog, nprog));
  const uint32_t nsegs = cols[0].nsegs;
  ALN_REQUIRE((uint64_t)nsegs * seg_rows <= (1ull << 32), "aln_scan_filter_noalloc: more than 2^32 rows");
  *h_count = 0;
  if (nsegs == 0) return ALN_OK;
  const uint32_t tps = (seg_rows + SC_TILE - 1) / SC_TILE;
  const uint64_t nblocks = (uint64_t)nsegs * tps;
  if (nblocks > NA_TILES)
    return set_error(ALN_E_ARG,
                     "aln_scan_filter_noalloc: %u segments of %u rows are %llu tiles of %d rows; the fixed "
                     "scratch holds %llu tiles (use aln_scan_filter)",
                     nsegs, seg_rows, (unsigned long long)nblocks, SC_TILE, (unsigned long long)NA_TILES);
  RangeC rgs[MAX_RANGES_C];
  int nrg = 0;
  bool empty = false;
  const bool fast = as_intervals(prog, nprog, rgs, &nrg, &empty);
  if (fast && empty) return ALN_OK;
  char *base;
  ALN_TRY(scan_scratch(c, &base));
  ColTable *ct = reinterpret_cast<ColTable *>(base + NA_CT);
  aln_insn *dp = reinterpret_cast<aln_insn *>(base + NA_PROG);
  RangeC *drg = fast && nrg > 0 ? reinterpret_cast<RangeC *>(base + NA_RANGES) : nullptr;
  uint64_t *tot = reinterpret_cast<uint64_t *>(base + NA_TOTAL);
  char *dmodes = seg_modes ? base + NA_MODES : nullptr;
  uint32_t *counts = reinterpret_cast<uint32_t *>(base + NA_COUNTS);
  uint64_t *offs = reinterpret_cast<uint64_t *>(base + NA_OFFS);
  uint64_t *sums = reinterpret_cast<uint64_t *>(base + NA_SUMS);
  ALN_HIP(hipMemcpyAsync(ct, &hct, sizeof(hct), hipMemcpyHostToDevice, c->stream));
  ALN_HIP(hipMemcpyAsync(dp, prog, sizeof(aln_insn) * nprog, hipMemcpyHostToDevice, c->stream));
  if (drg) ALN_HIP(hipMemcpyAsync(drg, rgs, sizeof(RangeC) * nrg, hipMemcpyHostToDevice, c->stream));
  if (dmodes) ALN_HIP(hipMemcpyAsync(dmodes, seg_modes, nsegs, hipMemcpyHostToDevice, c->stream));
  if (fast)
    hipLaunchKernelGGL((k_scan<0, true>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream, ct, dp,
                       nprog, drg, nrg, seg_rows, tps, dmodes, counts, (const uint64_t *)nullptr,
                       (uint32_t *)nullptr);
  else
    hipLaunchKernelGGL((k_scan<0, false>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream, ct, dp,
                       nprog, drg, 0, seg_rows, tps, dmodes, counts, (const uint64_t *)nullptr,
                       (uint32_t *)nullptr);
  ALN_TRY(check_launch("scan_count"));
  ALN_TRY(exclusive_scan_u32_buf(c->stream, counts, nblocks, offs, sums, tot));
  uint64_t total = 0;
  ALN_HIP(hipMemcpyAsync(&total, tot, sizeof(total), hipMemcpyDeviceToHost, c->stream));
  ALN_HIP(hipStreamSynchronize(c->stream));
  *h_count = total;
  if (total > cap)
    return set_error(ALN_E_CAPACITY, "aln_scan_filter_noalloc: %llu rows pass, capacity %llu",
                     (unsigned long long)total, (unsigned long long)cap);
  if (total == 0) return ALN_OK;
  ALN_REQUIRE(d_row_ids, "aln_scan_filter_noalloc: no output buffer");
  if (fast)
    hipLaunchKernelGGL((k_scan<1, true>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream, ct, dp,
                       nprog, drg, nrg, seg_rows, tps, dmodes, counts, offs, d_row_ids);
  else
    hipLaunchKernelGGL((k_scan<1, false>), dim3((unsigned)nblocks), dim3(SC_BLOCK), 0, c->stream, ct, dp,
                       nprog, drg, 0, seg_rows, tps, dmodes, counts, offs, d_row_ids);
  ALN_TRY(check_launch("scan_write"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

}  // extern "C"

// Internal declarations of libalenka_hip: context, workspace arena, error reporting,
// launch timing and the device helpers every kernel shares (bit-unpacking, hashing).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/alenka_hip.h"

namespace aln {

// ---- errors ---------------------------------------------------------------------------
int set_error(int code, const char *fmt, ...);
#define ALN_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return ::aln::set_error(ALN_E_HIP, "%s failed: %s (%s:%d)", #expr,                  \
                              hipGetErrorString(_e), __FILE__, __LINE__);                 \
  } while (0)
#define ALN_TRY(expr)                                                                     \
  do {                                                                                    \
    int _r = (expr);                                                                      \
    if (_r != ALN_OK) return _r;                                                          \
  } while (0)
#define ALN_REQUIRE(cond, ...)                                                            \
  do {                                                                                    \
    if (!(cond)) return ::aln::set_error(ALN_E_ARG, __VA_ARGS__);                         \
  } while (0)

}  // namespace aln

// ---- context ----------------------------------------------------------------------------
struct aln_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // workspace arena: chunks are kept for the life of the context; reset() rewinds them.
  struct Chunk {
    char *ptr;
    size_t size;
  };
  std::vector<Chunk> chunks;
  size_t cur = 0, off = 0;
  // pinned host staging for readbacks
  void *pinned = nullptr;
  size_t pinned_size = 0;
  // small readbacks: a kernel writes the bytes and then a sequence number into coherent
  // host memory, and the host polls the number (no stream synchronisation)
  void *rb = nullptr;
  uint32_t *rb_dev = nullptr;
  uint32_t rb_seq = 0;
  bool rb_failed = false;
  // launch timing
  bool profiling = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, uint64_t>> prof;
  // the star pipeline's key ranges from the zone maps of the last call's key columns
  // (segment descriptors and counts): reused by the next call on the same columns, which
  // checks them on the device and runs again with fresh ranges if they changed
  struct StarRanges {
    bool valid = false;
    const void *segs[3];
    uint32_t nsegs[3];
    int64_t mn[3], mx[3];
  } star_rng;
};

namespace aln {

// Workspace allocation (device, 256-byte aligned) valid until the next ws_reset().
int ws_alloc(aln_ctx *c, size_t bytes, void **out);
void ws_reset(aln_ctx *c);
template <typename T>
int ws(aln_ctx *c, size_t n, T **out) {
  return ws_alloc(c, n * sizeof(T) + 16, reinterpret_cast<void **>(out));
}
int pinned(aln_ctx *c, size_t bytes, void **out);

// Launch timing: bracket a launch with events when profiling is on.
void prof_begin(aln_ctx *c, const char *name, int *slot);
void prof_end(aln_ctx *c, int slot);
int prof_collect(aln_ctx *c);

struct ProfScope {
  aln_ctx *c;
  int slot;
  ProfScope(aln_ctx *c_, const char *name) : c(c_), slot(-1) { prof_begin(c, name, &slot); }
  ~ProfScope() { prof_end(c, slot); }
};

int check_launch(const char *name);
int copy_to_host(aln_ctx *c, void *h, const void *d, size_t bytes);  // synchronous
int copy_to_device(aln_ctx *c, void *d, const void *h, size_t bytes);

// ---- shared utilities (host) -------------------------------------------------------------
int exclusive_scan_u32(aln_ctx *c, const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                       uint64_t *h_total);
int exclusive_scan_u64(aln_ctx *c, const uint64_t *d_in, uint64_t n, uint64_t *d_out,
                       uint64_t *h_total);
// The same scan on caller-provided temporaries: d_sums holds scan_tiles(n) entries, d_total
// one; enqueued on s, no readback.
uint64_t scan_tiles(uint64_t n);
int exclusive_scan_u32_buf(hipStream_t s, const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                           uint64_t *d_sums, uint64_t *d_total);
// Looks up this library's fixed device scratch (aln_scan_filter_noalloc) on `device`, which
// must be current; once per device, later calls return the recorded status.
int scan_scratch_attach(int device);
int minmax_i64(aln_ctx *c, const int64_t *d, uint64_t n, int64_t *h_min, int64_t *h_max);
int sum_u32(aln_ctx *c, const uint32_t *d_in, uint64_t n, uint64_t *d_out);  // no sync
// *out[i] += sum of in[i][0, n[i]) for up to SUM_MULTI_MAX arrays, one launch (no sync);
// for the arrays whose bit is set in popc, the number of set bits of the words instead
constexpr int SUM_MULTI_MAX = 8;
int sum_u32_multi(aln_ctx *c, int k, const uint32_t *const *in, const uint64_t *n, uint64_t *const *out,
                  uint32_t popc = 0);
int radix_sort_pairs(aln_ctx *c, uint64_t *d_keys, uint32_t *d_vals, uint64_t n,
                     uint32_t bits);
// Batched queries (batch.hip): a batch whose device tables are carved from a caller-owned
// region instead of hipMalloc (pool NULL: hipMalloc), and the bytes such a region needs.
// hsegs (per column, host copies of the segment headers) and hmin/hmax (per column value
// ranges) may be given to spare the device reads; NULL: read from the columns.
int batch_create(aln_ctx *c, const aln_col *cols, int ncols, uint32_t seg_rows,
                 const aln_batch_query *qs, int nq, const aln_batch_tuning *tune, void *pool,
                 size_t pool_bytes, aln_batch **out, const aln_seg *const *hsegs,
                 const int64_t *hmin, const int64_t *hmax);
size_t batch_pool_bytes(const uint64_t *ngroups, int nq);
// Groups of query q into three host vectors (keys, sums, counts), resized to fit.
int batch_fetch(aln_batch *b, int q, std::vector<int64_t> *keys_sums_counts);
// Dense group count of keys in [mn, mx] grouped by floor(key / div) (0 if mn > mx).
uint64_t dense_groups(int64_t mn, int64_t mx, int64_t div);
int64_t floor_div(int64_t a, int64_t b);
int fill_u64(aln_ctx *c, uint64_t *d, uint64_t n, uint64_t v);
int fill_u32(aln_ctx *c, uint32_t *d, uint64_t n, uint32_t v);
int fill_i64(aln_ctx *c, int64_t *d, uint64_t n, int64_t v);
int iota_u32(aln_ctx *c, uint32_t *d, uint64_t n);

inline uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}
inline uint32_t bits_for_range_u(uint64_t r) {
  return r == 0 ? 0u : (uint32_t)(64 - __builtin_clzll(r));
}
inline unsigned grid_for(uint64_t n, unsigned per_block, unsigned cap = 1u << 30) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

// Column set passed to kernels by value.
struct DevCols {
  const uint32_t *words[ALN_MAX_COLS];
  const aln_seg *segs[ALN_MAX_COLS];
  int n;
};
struct DevProg {
  aln_insn ins[ALN_MAX_INSN];
  int n;
};

}  // namespace aln

// ---- device helpers ---------------------------------------------------------------------
namespace aln {

// Code of row `i` of a segment of width `bits` starting at word `w`.  Reads only the
// words that hold the code, so a segment's last row never touches the next word.
__device__ __forceinline__ uint64_t unpack_code(const uint32_t *__restrict__ w, uint64_t i,
                                                uint32_t bits) {
  if (bits == 0) return 0;
  uint64_t pos = i * bits;
  uint64_t wi = pos >> 5;
  uint32_t s = (uint32_t)(pos & 31);
  uint64_t lo = w[wi];
  if (s + bits > 32) lo |= (uint64_t)w[wi + 1] << 32;
  uint64_t v = lo >> s;
  if (s + bits > 64) v |= (uint64_t)w[wi + 2] << (64 - s);
  return bits == 64 ? v : (v & ((1ull << bits) - 1));
}

__device__ __forceinline__ int64_t decode_row(const uint32_t *__restrict__ words,
                                              const aln_seg &sg, uint64_t local) {
  return (int64_t)((uint64_t)sg.base + unpack_code(words + sg.word_off, local, sg.bits));
}

// MurmurHash64A of one 8-byte key (MurmurHash2_64.cu:63-104 with len = 8).
__host__ __device__ __forceinline__ uint64_t murmur64a_u64(uint64_t k, uint32_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  uint64_t h = (uint64_t)seed ^ (8ull * m);
  k *= m;
  k ^= k >> 47;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}

__device__ __forceinline__ int64_t wmul(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a * (uint64_t)b);
}
__device__ __forceinline__ int cmp_code(int code, int64_t l, int64_t r) {
  switch (code) {
    case 1: return l < r;
    case 2: return l > r;
    case 3: return l != r;
    case 4: return l == r;
    case 5: return l <= r;
    case 6: return l >= r;
    default: return 0;
  }
}

// f64 values travel as their IEEE-754 bits in int64 slots.
__host__ __device__ __forceinline__ int64_t f64_bits(double x) {
  int64_t b;
  memcpy(&b, &x, 8);
  return b;
}
__host__ __device__ __forceinline__ double bits_f64(int64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}
// The EPSILON comparisons of CudaSet's float overloads (f_less, f_greater, f_equal_to,
// f_greater_equal_to, f_less_equal, f_not_equal_to; cm.h:17, cm.cu:81-134): x is the left
// operand, y the right.
constexpr double ALN_EPSILON = 1.0E-8;
__host__ __device__ __forceinline__ int fcmp_code(int code, double x, double y) {
  const bool eq = ((x - y) < ALN_EPSILON) && ((x - y) > -ALN_EPSILON);
  switch (code) {
    case 1: return (y - x) > ALN_EPSILON;
    case 2: return (x - y) > ALN_EPSILON;
    case 3: return ((x - y) > ALN_EPSILON) || ((x - y) < -ALN_EPSILON);
    case 4: return eq;
    case 5: return ((y - x) > ALN_EPSILON) || eq;
    case 6: return ((x - y) > ALN_EPSILON) || eq;
    default: return 0;
  }
}
// Order-preserving int64 of a double's bits (-0.0 as +0.0).
__host__ __device__ __forceinline__ int64_t f64_order_key(int64_t b) {
  if (b == INT64_MIN) b = 0;  // -0.0
  return b < 0 ? (int64_t)((uint64_t)b ^ 0x7FFFFFFFFFFFFFFFull) : b;
}
__host__ __device__ __forceinline__ int64_t f64_unorder_key(int64_t k) {
  return k < 0 ? (int64_t)((uint64_t)k ^ 0x7FFFFFFFFFFFFFFFull) : k;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
// Set bits of a wave mask in the lanes below this one: two mbcnt instructions (the same as
// popcount(b & lanemask_lt()) without building the mask).
__device__ __forceinline__ uint32_t below_count(uint64_t b) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

}  // namespace aln

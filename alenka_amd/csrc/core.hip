// Context, workspace arena, errors, launch timing and the device-wide primitives
// (scans, min/max, fills) the operators are built from.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "aln_internal.h"

namespace aln {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(ALN_E_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
  return ALN_OK;
}

int ws_alloc(aln_ctx *c, size_t bytes, void **out) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  while (c->cur < c->chunks.size()) {
    auto &ch = c->chunks[c->cur];
    if (c->off + bytes <= ch.size) {
      *out = ch.ptr + c->off;
      c->off += bytes;
      return ALN_OK;
    }
    ++c->cur;
    c->off = 0;
  }
  size_t want = bytes;
  if (!c->chunks.empty() && c->chunks.back().size * 2 > want) want = c->chunks.back().size * 2;
  if (want < ((size_t)64 << 20)) want = (size_t)64 << 20;
  void *p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    want = bytes;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(ALN_E_NOMEM, "workspace allocation of %zu bytes failed", bytes);
    }
  }
  c->chunks.push_back({(char *)p, want});
  c->cur = c->chunks.size() - 1;
  c->off = bytes;
  *out = p;
  return ALN_OK;
}

void ws_reset(aln_ctx *c) {
  c->cur = 0;
  c->off = 0;
}

int pinned(aln_ctx *c, size_t bytes, void **out) {
  if (c->pinned_size < bytes) {
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    size_t want = bytes < 65536 ? 65536 : bytes;
    if (hipHostMalloc(&c->pinned, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      c->pinned_size = 0;
      return set_error(ALN_E_NOMEM, "pinned allocation of %zu bytes failed", want);
    }
    c->pinned_size = want;
  }
  *out = c->pinned;
  return ALN_OK;
}

// Small readbacks without a stream synchronisation: one block copies the words into
// coherent (fine-grained) host memory, makes them visible system-wide, then stores `seq`
// after them; the host polls that word.  The wait on a blocking stream synchronisation
// (the host thread's wake-up after the GPU finishes) is what this avoids.
constexpr size_t RB_BYTES = 65536;
__global__ void __launch_bounds__(256) k_readback(const uint32_t *__restrict__ src, uint32_t nwords,
                                                 uint32_t *__restrict__ dst, uint32_t *flag, uint32_t seq) {
  for (uint32_t i = threadIdx.x; i < nwords; i += 256) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

static bool readback_ready(aln_ctx *c) {
  if (c->rb) return true;
  if (c->rb_failed) return false;
  void *p = nullptr;
  uint32_t *dp = nullptr;
  if (hipHostMalloc(&p, RB_BYTES + 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
    (void)hipGetLastError();
    c->rb_failed = true;
    return false;
  }
  if (hipHostGetDevicePointer((void **)&dp, p, 0) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipHostFree(p);
    c->rb_failed = true;
    return false;
  }
  memset(p, 0, RB_BYTES + 64);
  c->rb = p;
  c->rb_dev = dp;
  return true;
}

static int readback_small(aln_ctx *c, void *h, const void *d, size_t bytes) {
  uint32_t seq = ++c->rb_seq;
  if (seq == 0) seq = ++c->rb_seq;  // 0 is the buffer's initial value
  volatile uint32_t *flag = reinterpret_cast<volatile uint32_t *>(static_cast<char *>(c->rb) + RB_BYTES);
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(256), 0, c->stream, static_cast<const uint32_t *>(d),
                     (uint32_t)((bytes + 3) / 4), c->rb_dev, c->rb_dev + RB_BYTES / 4, seq);
  ALN_TRY(check_launch("readback"));
  // poll; after a while give the thread up between polls, and after 200 ms fall back to a
  // stream synchronisation (which also reports a failed launch)
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t spins = 0; *flag != seq; ++spins) {
    if ((spins & 1023) == 1023) {
      const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
      if (us > 200000.0) {
        ALN_HIP(hipStreamSynchronize(c->stream));
        if (*flag != seq) return set_error(ALN_E_HIP, "readback: the copy kernel did not complete");
        break;
      }
      if (us > 50.0) std::this_thread::yield();
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  memcpy(h, c->rb, bytes);
  return ALN_OK;
}

int copy_to_host(aln_ctx *c, void *h, const void *d, size_t bytes) {
  if (bytes == 0) return ALN_OK;
  if (bytes <= RB_BYTES && ((uintptr_t)d & 3) == 0 && readback_ready(c)) return readback_small(c, h, d, bytes);
  void *p;
  ALN_TRY(pinned(c, bytes, &p));
  ALN_HIP(hipMemcpyAsync(p, d, bytes, hipMemcpyDeviceToHost, c->stream));
  ALN_HIP(hipStreamSynchronize(c->stream));
  memcpy(h, p, bytes);
  return ALN_OK;
}

int copy_to_device(aln_ctx *c, void *d, const void *h, size_t bytes) {
  if (bytes == 0) return ALN_OK;
  ALN_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

void prof_begin(aln_ctx *c, const char *name, int *slot) {
  *slot = -1;
  if (!c->profiling) return;
  aln_ctx::Pending p;
  p.name = name;
  for (int i = 0; i < 2; ++i) {
    hipEvent_t e;
    if (!c->event_pool.empty()) {
      e = c->event_pool.back();
      c->event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    (i == 0 ? p.a : p.b) = e;
  }
  if (hipEventRecord(p.a, c->stream) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  c->pending.push_back(p);
  *slot = (int)c->pending.size() - 1;
}

void prof_end(aln_ctx *c, int slot) {
  if (slot < 0 || slot >= (int)c->pending.size()) return;
  if (hipEventRecord(c->pending[slot].b, c->stream) != hipSuccess) (void)hipGetLastError();
}

int prof_collect(aln_ctx *c) {
  if (c->pending.empty()) return ALN_OK;
  ALN_HIP(hipStreamSynchronize(c->stream));
  for (auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto &acc = c->prof[p.name];
      acc.first += ms;
      acc.second += 1;
    } else {
      (void)hipGetLastError();
    }
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
  return ALN_OK;
}

// ---- device-wide primitives ----------------------------------------------------------------

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename T>
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t *total) {
  // block-wide exclusive scan of one value per thread (SCAN_BLOCK threads, wave 64)
  __shared__ uint64_t wsum[SCAN_BLOCK / 64];
  unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint64_t base = 0, tot = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    uint64_t s = wsum[w];
    if (w < (int)wid) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_tile_sums(const T *__restrict__ in, uint64_t n,
                                                          uint64_t *__restrict__ sums) {
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t tot;
  (void)block_excl_scan<T>(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans `sums` (m entries) in place sequentially by one block, chunk by chunk.
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_sums(uint64_t *sums, uint64_t m,
                                                          uint64_t *total) {
  uint64_t carry = 0;
  for (uint64_t b = 0; b < m; b += SCAN_BLOCK) {
    uint64_t i = b + threadIdx.x;
    uint64_t v = i < m ? sums[i] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_scan<uint64_t>(v, &tot);
    if (i < m) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

// Each wave scans 16 rows of 64 consecutive items (a row per load, so loads and stores are
// coalesced), carrying the row totals across the rows; then the waves' totals and the
// tile's offset are added.
template <typename T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_tile_scan(const T *__restrict__ in, uint64_t n,
                                                          const uint64_t *__restrict__ offs,
                                                          uint64_t *__restrict__ out) {
  __shared__ uint64_t wsum[SCAN_BLOCK / 64];
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)wid * (64 * SCAN_ITEMS) + lane;
  uint64_t v[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64;
    v[j] = i < n ? (uint64_t)in[i] : 0;
  }
  uint64_t carry = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t x = v[j];
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(x, o, 64);
      if (lane >= (unsigned)o) x += y;
    }
    const uint64_t row = __shfl(x, 63, 64);
    v[j] = carry + x - v[j];  // exclusive within the wave's part of the tile
    carry += row;
  }
  if (lane == 0) wsum[wid] = carry;
  __syncthreads();
  uint64_t off = offs[blockIdx.x];
  for (unsigned w = 0; w < wid; ++w) off += wsum[w];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64;
    if (i < n) out[i] = off + v[j];
  }
}

template <typename T>
static int exclusive_scan_launch(hipStream_t s, const T *d_in, uint64_t n, uint64_t *d_out,
                                 uint64_t *sums, uint64_t *tot) {
  const uint64_t nt = scan_tiles(n);
  hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, s, d_in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nt, tot);
  hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, s, d_in, n, sums, d_out);
  return check_launch("exclusive_scan");
}

uint64_t scan_tiles(uint64_t n) {
  const uint64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  return nt == 0 ? 1 : nt;
}

int exclusive_scan_u32_buf(hipStream_t s, const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                           uint64_t *d_sums, uint64_t *d_total) {
  return exclusive_scan_launch<uint32_t>(s, d_in, n, d_out, d_sums, d_total);
}

template <typename T>
static int exclusive_scan_impl(aln_ctx *c, const T *d_in, uint64_t n, uint64_t *d_out,
                               uint64_t *h_total) {
  uint64_t *sums, *tot;
  ALN_TRY(ws(c, scan_tiles(n), &sums));
  ALN_TRY(ws(c, 1, &tot));
  {
    ProfScope p(c, "scan");
    ALN_TRY(exclusive_scan_launch<T>(c->stream, d_in, n, d_out, sums, tot));
  }
  if (h_total) ALN_TRY(copy_to_host(c, h_total, tot, 8));
  return ALN_OK;
}

int exclusive_scan_u32(aln_ctx *c, const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                       uint64_t *h_total) {
  return exclusive_scan_impl<uint32_t>(c, d_in, n, d_out, h_total);
}
int exclusive_scan_u64(aln_ctx *c, const uint64_t *d_in, uint64_t n, uint64_t *d_out,
                       uint64_t *h_total) {
  return exclusive_scan_impl<uint64_t>(c, d_in, n, d_out, h_total);
}

// Sum of a u32 array into *d_out (device), without atomics: per-block partial sums, then
// one block adds the partials.
__global__ void __launch_bounds__(256) k_sum_partials(const uint32_t *__restrict__ in, uint64_t n,
                                                      uint64_t *__restrict__ part) {
  uint64_t s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    s += in[i];
  uint64_t tot;
  (void)block_excl_scan<uint64_t>(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(256) k_sum_final(const uint64_t *__restrict__ part, uint32_t np,
                                                   uint64_t *__restrict__ out) {
  uint64_t s = 0;
  for (uint32_t i = threadIdx.x; i < np; i += 256) s += part[i];
  uint64_t tot;
  (void)block_excl_scan<uint64_t>(s, &tot);
  if (threadIdx.x == 0) *out = tot;
}

int sum_u32(aln_ctx *c, const uint32_t *d_in, uint64_t n, uint64_t *d_out) {
  const unsigned nb = grid_for(n, 256 * 16, 1024);
  uint64_t *part;
  ALN_TRY(ws(c, nb, &part));
  hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(256), 0, c->stream, d_in, n, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, c->stream, part, nb, d_out);
  return check_launch("sum_u32");
}

// Several sums in one launch: block (x, y) adds a strided share of array y and adds its
// total to *out[y] (at most 128 atomics per output; the caller zeroes the outputs).
struct SumSet {
  const uint32_t *in[SUM_MULTI_MAX];
  uint64_t n[SUM_MULTI_MAX];
  uint64_t *out[SUM_MULTI_MAX];
  uint32_t popc;  // bit i: add the set bits of array i's words instead of their values
};
__global__ void __launch_bounds__(256) k_sum_multi(SumSet s) {
  const uint32_t *in = s.in[blockIdx.y];
  const uint64_t n = s.n[blockIdx.y];
  const bool bits = (s.popc >> blockIdx.y) & 1;
  // eight loads in flight per thread and round
  uint64_t v = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i0 = (uint64_t)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 8 * stride) {
    uint32_t x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = i0 + u * stride < n ? in[i0 + u * stride] : 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) v += bits ? (uint32_t)__popc(x[u]) : x[u];
  }
  uint64_t tot;
  (void)block_excl_scan<uint64_t>(v, &tot);
  if (threadIdx.x == 0 && tot) atomicAdd((unsigned long long *)s.out[blockIdx.y], (unsigned long long)tot);
}

int sum_u32_multi(aln_ctx *c, int k, const uint32_t *const *in, const uint64_t *n, uint64_t *const *out,
                  uint32_t popc) {
  ALN_REQUIRE(k >= 1 && k <= SUM_MULTI_MAX, "sum_u32_multi: %d arrays", k);
  SumSet s;
  s.popc = popc;
  uint64_t most = 0;
  for (int i = 0; i < k; ++i) {
    s.in[i] = in[i];
    s.n[i] = n[i];
    s.out[i] = out[i];
    most = n[i] > most ? n[i] : most;
  }
  // few blocks (the outputs usually share a cache line, whose atomics serialise), each
  // thread with eight loads in flight
  const unsigned nb = grid_for(most ? most : 1, 256 * 32, 128);
  hipLaunchKernelGGL(k_sum_multi, dim3(nb, k), dim3(256), 0, c->stream, s);
  return check_launch("sum_u32_multi");
}

// Min/max of an int64 array: per-block partials, then one block folds them (no atomics on
// a shared address).
__global__ void __launch_bounds__(256) k_minmax_part(const int64_t *__restrict__ d, uint64_t n,
                                                     int64_t *__restrict__ pmin,
                                                     int64_t *__restrict__ pmax) {
  __shared__ int64_t smn[4], smx[4];
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  // eight independent coalesced loads in flight per thread
  for (; i + 7 * stride < n; i += 8 * stride) {
    int64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(d + i + k * stride);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      lo = v[k] < lo ? v[k] : lo;
      hi = v[k] > hi ? v[k] : hi;
    }
  }
  for (; i < n; i += stride) {
    const int64_t v = d[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    int64_t a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (__lane_id() == 0) {
    smn[threadIdx.x / 64] = lo;
    smx[threadIdx.x / 64] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = smn[w] < lo ? smn[w] : lo;
      hi = smx[w] > hi ? smx[w] : hi;
    }
    lo = smn[0] < lo ? smn[0] : lo;
    hi = smx[0] > hi ? smx[0] : hi;
    pmin[blockIdx.x] = lo;
    pmax[blockIdx.x] = hi;
  }
}

int minmax_i64(aln_ctx *c, const int64_t *d, uint64_t n, int64_t *h_min, int64_t *h_max) {
  const unsigned nb = grid_for(n, 256 * 16, 2048);  // 8 blocks per CU; 32 KB of partials
  int64_t *part;
  ALN_TRY(ws(c, 2 * (size_t)nb, &part));
  if (n) {
    ProfScope p(c, "minmax");
    hipLaunchKernelGGL(k_minmax_part, dim3(nb), dim3(256), 0, c->stream, d, n, part, part + nb);
  }
  ALN_TRY(check_launch("minmax"));
  std::vector<int64_t> h(2 * (size_t)nb);
  *h_min = INT64_MAX;
  *h_max = INT64_MIN;
  if (!n) return ALN_OK;
  ALN_TRY(copy_to_host(c, h.data(), part, 16ull * nb));
  for (unsigned i = 0; i < nb; ++i) {
    *h_min = h[i] < *h_min ? h[i] : *h_min;
    *h_max = h[nb + i] > *h_max ? h[nb + i] : *h_max;
  }
  return ALN_OK;
}

template <typename T>
__global__ void k_fill(T *d, uint64_t n, T v) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    d[i] = v;
}
__global__ void k_iota(uint32_t *d, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    d[i] = (uint32_t)i;
}

int fill_u64(aln_ctx *c, uint64_t *d, uint64_t n, uint64_t v) {
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, d,
                     n, v);
  return check_launch("fill_u64");
}
int fill_u32(aln_ctx *c, uint32_t *d, uint64_t n, uint32_t v) {
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, d,
                     n, v);
  return check_launch("fill_u32");
}
int fill_i64(aln_ctx *c, int64_t *d, uint64_t n, int64_t v) {
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_fill<int64_t>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, d,
                     n, v);
  return check_launch("fill_i64");
}
int iota_u32(aln_ctx *c, uint32_t *d, uint64_t n) {
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, d, n);
  return check_launch("iota");
}

}  // namespace aln

using namespace aln;

extern "C" {

const char *aln_last_error(void) { return g_err; }
const char *aln_version(void) { return "alenka_amd 0.1 (gfx950)"; }

int aln_create(int device, void *stream, aln_ctx **out) {
  ALN_REQUIRE(out, "aln_create: out is NULL");
  int n = 0;
  ALN_HIP(hipGetDeviceCount(&n));
  ALN_REQUIRE(device >= 0 && device < n, "aln_create: device %d of %d", device, n);
  ALN_HIP(hipSetDevice(device));
  aln_ctx *c = new aln_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  // the allocation-free scan's scratch address on this device (current here); a failure
  // is left for that entry point's first call, which looks it up again and reports it
  if (aln::scan_scratch_attach(device) != ALN_OK) (void)hipGetLastError();
  *out = c;
  return ALN_OK;
}

void aln_destroy(aln_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &ch : c->chunks) (void)hipFree(ch.ptr);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->rb) (void)hipHostFree(c->rb);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  delete c;
}

int aln_set_stream(aln_ctx *c, void *stream) {
  ALN_REQUIRE(c, "aln_set_stream: ctx is NULL");
  c->stream = (hipStream_t)stream;
  return ALN_OK;
}

int aln_sync(aln_ctx *c) {
  ALN_REQUIRE(c, "aln_sync: ctx is NULL");
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_profile_enable(aln_ctx *c, int on) {
  ALN_REQUIRE(c, "aln_profile_enable: ctx is NULL");
  c->profiling = on != 0;
  return ALN_OK;
}

int aln_profile_reset(aln_ctx *c) {
  ALN_REQUIRE(c, "aln_profile_reset: ctx is NULL");
  ALN_TRY(prof_collect(c));
  c->prof.clear();
  return ALN_OK;
}

int aln_profile_get(aln_ctx *c, const char *kernel, double *h_ms, uint64_t *h_launches) {
  ALN_REQUIRE(c && kernel && h_ms && h_launches, "aln_profile_get: NULL argument");
  ALN_TRY(prof_collect(c));
  auto it = c->prof.find(kernel);
  *h_ms = it == c->prof.end() ? 0.0 : it->second.first;
  *h_launches = it == c->prof.end() ? 0 : it->second.second;
  return ALN_OK;
}

// MurmurHash64A (MurmurHash2_64.cu:63-104), host side.
uint64_t aln_murmur64a(const void *key, int len, uint32_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  uint64_t h = (uint64_t)seed ^ ((uint64_t)(int64_t)len * m);
  const unsigned char *p = (const unsigned char *)key;
  int nb = len / 8;
  for (int b = 0; b < nb; ++b) {
    uint64_t k;
    memcpy(&k, p + 8 * b, 8);
    k *= m;
    k ^= k >> 47;
    k *= m;
    h ^= k;
    h *= m;
  }
  const unsigned char *t = p + 8 * nb;
  uint64_t tail = 0;
  int rem = len & 7;
  for (int i = rem - 1; i >= 0; --i) tail = (tail << 8) | t[i];
  if (rem) {
    h ^= tail;
    h *= m;
  }
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}

}  // extern "C"

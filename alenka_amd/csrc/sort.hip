// LSD radix sort, stable multi-key ORDER BY and top-k.
//
// The reference sorts with one thrust::stable_sort_by_key per key, last key first
// (update_permutation, sorts.cu:17-70; emit_order, operators.cu:1675-1763).  Here each key
// is first mapped to an order-preserving unsigned code of just the bits its range needs
// (v - min, or max - v for DESC); when all codes fit in 64 bits they are concatenated and
// sorted once, otherwise one stable radix sort per key runs from the last key to the first.
//
// Radix pass: 8-bit digits, 4096-item tiles (4 waves x 64 lanes x 16 items, striped so
// that wave order then (item, lane) order is memory order).  Ranks inside a wave come from
// ballot-based digit matching; a digit-major histogram scan gives each (digit, tile) its
// global offset, which keeps every pass stable.
#include <string.h>

#include <algorithm>
#include <vector>

#include "aln_internal.h"

namespace aln {

constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVE_ITEMS = 64 * RS_ITEMS;

__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const uint64_t *__restrict__ keys, uint64_t n,
                                                      uint32_t shift, uint32_t mask,
                                                      uint32_t nblocks, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * RS_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const uint64_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin, uint64_t n,
                                                         uint32_t shift, uint32_t mask,
                                                         uint32_t nblocks,
                                                         const uint64_t *__restrict__ offs,
                                                         uint64_t *__restrict__ kout,
                                                         uint32_t *__restrict__ vout) {
  __shared__ uint32_t wc[RS_WAVES][256];
  __shared__ uint64_t goff[256];
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  for (int w = 0; w < RS_WAVES; ++w) wc[w][threadIdx.x] = 0;
  goff[threadIdx.x] = offs[(uint64_t)threadIdx.x * nblocks + blockIdx.x];
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)wid * RS_WAVE_ITEMS;
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS], rank[RS_ITEMS];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * 64 + lane;
    bool valid = i < n;
    key[j] = valid ? kin[i] : 0;
    val[j] = valid ? vin[i] : 0;
    uint32_t d = (uint32_t)(key[j] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t before = valid ? wc[wid][d] : 0;
    uint64_t below = peers & lt;
    rank[j] = before + (uint32_t)__popcll(below);
    if (valid && below == 0) wc[wid][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  {
    uint32_t run = 0;
    for (int w = 0; w < RS_WAVES; ++w) {
      uint32_t t = wc[w][threadIdx.x];
      wc[w][threadIdx.x] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * 64 + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[j] >> shift) & mask;
      uint64_t dest = goff[d] + wc[wid][d] + rank[j];
      kout[dest] = key[j];
      vout[dest] = val[j];
    }
  }
}

__global__ void k_copy_pairs(const uint64_t *ks, const uint32_t *vs, uint64_t n, uint64_t *kd,
                             uint32_t *vd) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

int radix_sort_pairs(aln_ctx *c, uint64_t *d_keys, uint32_t *d_vals, uint64_t n, uint32_t bits) {
  if (n <= 1 || bits == 0) return ALN_OK;
  ALN_REQUIRE(bits <= 64, "radix sort: %u key bits", bits);
  uint64_t nb64 = (n + RS_TILE - 1) / RS_TILE;
  ALN_REQUIRE(nb64 < (1u << 31) / 256, "radix sort: %llu items is too many",
              (unsigned long long)n);
  uint32_t nblocks = (uint32_t)nb64;
  uint64_t *k2;
  uint32_t *v2, *hist;
  uint64_t *offs;
  ALN_TRY(ws(c, n, &k2));
  ALN_TRY(ws(c, n, &v2));
  ALN_TRY(ws(c, (size_t)256 * nblocks, &hist));
  ALN_TRY(ws(c, (size_t)256 * nblocks, &offs));
  uint64_t *ka = d_keys, *kb = k2;
  uint32_t *va = d_vals, *vb = v2;
  uint32_t passes = (bits + 7) / 8;
  for (uint32_t p = 0; p < passes; ++p) {
    uint32_t shift = 8 * p;
    uint32_t pb = bits - shift < 8 ? bits - shift : 8;
    uint32_t mask = (1u << pb) - 1;
    {
      ProfScope ps(c, "radix_hist");
      hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(RS_BLOCK), 0, c->stream, ka, n, shift, mask,
                         nblocks, hist);
    }
    ALN_TRY(check_launch("radix_hist"));
    ALN_TRY(exclusive_scan_u32(c, hist, (uint64_t)256 * nblocks, offs, nullptr));
    {
      ProfScope ps(c, "radix_scatter");
      hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(RS_BLOCK), 0, c->stream, ka, va, n,
                         shift, mask, nblocks, offs, kb, vb);
    }
    ALN_TRY(check_launch("radix_scatter"));
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != d_keys) {
    hipLaunchKernelGGL(k_copy_pairs, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, ka, va,
                       n, d_keys, d_vals);
    ALN_TRY(check_launch("copy_pairs"));
  }
  return ALN_OK;
}

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

// Builds the codes of keys [k0, k1) (k0 most significant).  Returns total bits.
static int key_codes(aln_ctx *c, const int64_t *const *keys, const int32_t *desc, int k0, int k1,
                     uint64_t n, const uint32_t *perm, uint64_t *code, uint32_t *idx,
                     uint32_t *bits_out) {
  KeySet ks;
  memset(&ks, 0, sizeof(ks));
  uint32_t total = 0;
  uint32_t bits[ALN_MAX_KEYS];
  for (int k = k0; k < k1; ++k) {
    int64_t mn, mx;
    ALN_TRY(minmax_i64(c, keys[k], n, &mn, &mx));
    bits[k - k0] = bits_for_range_u((uint64_t)mx - (uint64_t)mn);
    ks.ref[k - k0] = (desc && desc[k]) ? mx : mn;
    ks.desc[k - k0] = desc ? desc[k] : 0;
    ks.k[k - k0] = keys[k];
    total += bits[k - k0];
  }
  ks.n = k1 - k0;
  uint32_t sh = total;
  for (int k = 0; k < ks.n; ++k) {
    sh -= bits[k];
    ks.shift[k] = sh;
  }
  ALN_REQUIRE(total <= 64, "key codes: %u bits", total);
  hipLaunchKernelGGL(k_codes, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, ks, perm, n,
                     code, idx);
  ALN_TRY(check_launch("codes"));
  *bits_out = total;
  return ALN_OK;
}

static int key_bits(aln_ctx *c, const int64_t *const *keys, int nkeys, uint64_t n,
                    uint32_t *bits) {
  for (int k = 0; k < nkeys; ++k) {
    int64_t mn, mx;
    ALN_TRY(minmax_i64(c, keys[k], n, &mn, &mx));
    bits[k] = bits_for_range_u((uint64_t)mx - (uint64_t)mn);
  }
  return ALN_OK;
}

int sort_keys_impl(aln_ctx *c, const int64_t *const *keys, const int32_t *desc, int nkeys,
                   uint64_t n, uint32_t *d_perm) {
  ALN_REQUIRE(nkeys >= 1 && nkeys <= ALN_MAX_KEYS, "sort: %d keys", nkeys);
  ALN_REQUIRE(n < (1ull << 32), "sort: more than 2^32 rows");
  if (n == 0) return ALN_OK;
  uint32_t bits[ALN_MAX_KEYS], total = 0;
  ALN_TRY(key_bits(c, keys, nkeys, n, bits));
  for (int k = 0; k < nkeys; ++k) total += bits[k];
  uint64_t *code;
  ALN_TRY(ws(c, n, &code));
  if (total <= 64) {
    uint32_t b;
    ALN_TRY(key_codes(c, keys, desc, 0, nkeys, n, nullptr, code, d_perm, &b));
    return radix_sort_pairs(c, code, d_perm, n, b);
  }
  // LSD over keys, last key first, each pass stable (sorts.cu:17-70)
  ALN_TRY(iota_u32(c, d_perm, n));
  for (int k = nkeys - 1; k >= 0; --k) {
    uint32_t b;
    ALN_TRY(key_codes(c, keys, desc, k, k + 1, n, d_perm, code, nullptr, &b));
    ALN_TRY(radix_sort_pairs(c, code, d_perm, n, b));
  }
  return ALN_OK;
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
  uint32_t bits[ALN_MAX_KEYS], total = 0;
  ALN_TRY(key_bits(c, d_keys, nkeys, n, bits));
  for (int i = 0; i < nkeys; ++i) total += bits[i];
  if (total > 64 || k > TK_TILE / 2) {
    // wide keys or large k: full stable sort, then the prefix (emit_order + LIMIT)
    uint32_t *perm;
    ALN_TRY(ws(c, n, &perm));
    ALN_TRY(sort_keys_impl(c, d_keys, desc, nkeys, n, perm));
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
  uint32_t b;
  ALN_TRY(key_codes(c, d_keys, desc, 0, nkeys, n, nullptr, code, idx, &b));
  return topk_codes(c, code, idx, n, k, d_out_idx, h_nout);
}

}  // extern "C"

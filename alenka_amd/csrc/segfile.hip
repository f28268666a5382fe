// Column segments in the reference's on-disk format, read and written on the device, so
// data an Alenka installation wrote can be loaded (and this build's data written back):
//   [u32 cnt][i64 lower][i64 upper][u32 comp_type][payload: cnt bytes]
//   [u32 cnt][u32 rec_count][u32 bits][i64 orig_lower][u32 fit_count][i64 start_val]
// (pfor_compress compress.cu:413-527, pfor_delta_compress 270-408, pfor_decompress 180-267).
//   comp_type 0, FOR      : codes value - lower as unsigned 8/16/32/64-bit integers.
//   comp_type 1, FOR-DELTA: first differences (the first replaced by the second), minus
//                           orig_lower, packed MSB-first `fit_count = 64 / bits` per 64-bit
//                           word; value[0] = start_val, then an inclusive prefix sum.
// lower/upper are the segment's zone map (min/max for FOR, first/last value for FOR-DELTA).
#include <string.h>

#include <vector>

#include "aln_internal.h"

namespace aln {

constexpr uint64_t SF_HEAD = 24, SF_TAIL = 32;

__global__ void k_sf_for_decode(const uint8_t *__restrict__ p, uint32_t width, uint64_t n, int64_t lower,
                                int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint64_t c;
    switch (width) {
      case 1: c = p[i]; break;
      case 2: c = ((const uint16_t *)p)[i]; break;
      case 4: c = ((const uint32_t *)p)[i]; break;
      default: c = ((const uint64_t *)p)[i]; break;
    }
    out[i] = (int64_t)(c + (uint64_t)lower);
  }
}

// Deltas of FOR-DELTA (row 0 gets start_val); the prefix sum follows.
__global__ void k_sf_delta_decode(const uint64_t *__restrict__ w, uint32_t bits, uint32_t fit, uint64_t n,
                                  int64_t lower, int64_t start, int64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    if (i == 0) {
      out[0] = start;
      continue;
    }
    const uint32_t sh = 64 - bits - (uint32_t)(i % fit) * bits;
    uint64_t c = w[i / fit] >> sh;
    if (bits < 64) c &= (1ull << bits) - 1;
    out[i] = (int64_t)(c + (uint64_t)lower);
  }
}

__global__ void k_sf_add(const uint64_t *__restrict__ ex, uint64_t n, int64_t *__restrict__ v) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    v[i] = (int64_t)(ex[i] + (uint64_t)v[i]);
}

__global__ void k_sf_for_encode(const int64_t *__restrict__ v, uint64_t n, int64_t lower, uint32_t width,
                                uint8_t *__restrict__ p) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t c = (uint64_t)v[i] - (uint64_t)lower;
    switch (width) {
      case 1: p[i] = (uint8_t)c; break;
      case 2: ((uint16_t *)p)[i] = (uint16_t)c; break;
      case 4: ((uint32_t *)p)[i] = (uint32_t)c; break;
      default: ((uint64_t *)p)[i] = c; break;
    }
  }
}

// The reference's code width: ceil(log2(range + 1)) rounded up to 8, 16, 32 or 64 bits.
static uint32_t for_bits(int64_t lo, int64_t hi) {
  const uint64_t r = (uint64_t)hi - (uint64_t)lo;
  const uint32_t b = r == ~0ull ? 64 : bits_for_range_u(r);
  return b <= 8 ? 8 : b <= 16 ? 16 : b <= 32 ? 32 : 64;
}

}  // namespace aln

using namespace aln;

extern "C" {

int aln_segfile_decode(aln_ctx *c, const void *h_file, uint64_t len, int64_t *d_out, uint64_t cap,
                       uint64_t *h_rows, int64_t *h_lower, int64_t *h_upper) {
  ALN_REQUIRE(c && h_file && h_rows, "aln_segfile_decode: bad arguments");
  ALN_REQUIRE(len >= SF_HEAD + SF_TAIL, "aln_segfile_decode: %llu bytes is shorter than a segment header",
              (unsigned long long)len);
  const uint8_t *f = (const uint8_t *)h_file;
  uint32_t cnt, comp, cnt2, rec, bits, fit;
  int64_t lower, upper, olo, start;
  memcpy(&cnt, f, 4);
  memcpy(&lower, f + 4, 8);
  memcpy(&upper, f + 12, 8);
  memcpy(&comp, f + 20, 4);
  ALN_REQUIRE(len >= SF_HEAD + (uint64_t)cnt + SF_TAIL, "aln_segfile_decode: payload of %u bytes exceeds the file",
              cnt);
  const uint8_t *t = f + SF_HEAD + cnt;
  memcpy(&cnt2, t, 4);
  memcpy(&rec, t + 4, 4);
  memcpy(&bits, t + 8, 4);
  memcpy(&olo, t + 12, 8);
  memcpy(&fit, t + 20, 4);
  memcpy(&start, t + 24, 8);
  ALN_REQUIRE(cnt2 == cnt, "aln_segfile_decode: header and trailer sizes differ (%u, %u)", cnt, cnt2);
  *h_rows = rec;
  if (h_lower) *h_lower = lower;
  if (h_upper) *h_upper = upper;
  if (rec > cap) return set_error(ALN_E_CAPACITY, "aln_segfile_decode: %u rows, capacity %llu", rec,
                                  (unsigned long long)cap);
  if (!rec) return ALN_OK;
  ALN_REQUIRE(d_out, "aln_segfile_decode: no output buffer");
  ws_reset(c);
  uint8_t *dp;
  ALN_TRY(ws(c, (size_t)cnt + 8, &dp));
  ALN_TRY(copy_to_device(c, dp, f + SF_HEAD, cnt));
  if (comp == 0) {
    ALN_REQUIRE(bits == 8 || bits == 16 || bits == 32 || bits == 64, "aln_segfile_decode: FOR width %u", bits);
    ALN_REQUIRE((uint64_t)rec * (bits / 8) <= cnt, "aln_segfile_decode: %u rows of %u bits exceed the payload",
                rec, bits);
    hipLaunchKernelGGL(k_sf_for_decode, dim3(grid_for(rec, 256, 65536)), dim3(256), 0, c->stream, dp, bits / 8,
                       (uint64_t)rec, olo, d_out);
    ALN_TRY(check_launch("segfile_for_decode"));
  } else if (comp == 1) {
    ALN_REQUIRE(bits >= 1 && bits <= 64 && fit == 64 / bits, "aln_segfile_decode: FOR-DELTA bits %u fit %u", bits,
                fit);
    ALN_REQUIRE(((uint64_t)rec + fit - 1) / fit * 8 <= cnt, "aln_segfile_decode: %u rows exceed the payload", rec);
    hipLaunchKernelGGL(k_sf_delta_decode, dim3(grid_for(rec, 256, 65536)), dim3(256), 0, c->stream,
                       (const uint64_t *)dp, bits, fit, (uint64_t)rec, olo, start, d_out);
    ALN_TRY(check_launch("segfile_delta_decode"));
    // inclusive prefix sum: exclusive scan of the deltas shifted by one
    uint64_t *ex;
    ALN_TRY(ws(c, rec, &ex));
    ALN_TRY(exclusive_scan_u64(c, (const uint64_t *)d_out, rec, ex, nullptr));
    hipLaunchKernelGGL(k_sf_add, dim3(grid_for(rec, 256, 65536)), dim3(256), 0, c->stream, ex, (uint64_t)rec, d_out);
    ALN_TRY(check_launch("segfile_prefix"));
  } else {
    return set_error(ALN_E_UNSUPPORTED, "aln_segfile_decode: compression type %u", comp);
  }
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_segfile_encode(aln_ctx *c, const int64_t *d_vals, uint64_t n, void *h_file, uint64_t cap,
                       uint64_t *h_len) {
  ALN_REQUIRE(c && h_len && (n == 0 || d_vals), "aln_segfile_encode: bad arguments");
  ALN_REQUIRE(n >= 1 && n <= 0xFFFFFFFFull, "aln_segfile_encode: %llu rows (1..2^32-1)", (unsigned long long)n);
  ws_reset(c);
  int64_t lo, hi;
  ALN_TRY(minmax_i64(c, d_vals, n, &lo, &hi));
  const uint32_t bits = for_bits(lo, hi), width = bits / 8;
  ALN_REQUIRE((uint64_t)n * width <= 0xFFFFFFFFull, "aln_segfile_encode: payload above 4 GiB");
  const uint32_t cnt = (uint32_t)(n * width), rec = (uint32_t)n, comp = 0, fit = 64 / bits;
  const int64_t start = 0;
  *h_len = SF_HEAD + (uint64_t)cnt + SF_TAIL;
  if (!h_file) return ALN_OK;
  if (*h_len > cap) return set_error(ALN_E_CAPACITY, "aln_segfile_encode: %llu bytes, capacity %llu",
                                     (unsigned long long)*h_len, (unsigned long long)cap);
  uint8_t *dp;
  ALN_TRY(ws(c, (size_t)cnt + 8, &dp));
  hipLaunchKernelGGL(k_sf_for_encode, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_vals, n, lo, width,
                     dp);
  ALN_TRY(check_launch("segfile_for_encode"));
  uint8_t *f = (uint8_t *)h_file;
  memcpy(f, &cnt, 4);
  memcpy(f + 4, &lo, 8);
  memcpy(f + 12, &hi, 8);
  memcpy(f + 20, &comp, 4);
  ALN_TRY(copy_to_host(c, f + SF_HEAD, dp, cnt));
  uint8_t *t = f + SF_HEAD + cnt;
  memcpy(t, &cnt, 4);
  memcpy(t + 4, &rec, 4);
  memcpy(t + 8, &bits, 4);
  memcpy(t + 12, &lo, 8);
  memcpy(t + 20, &fit, 4);
  memcpy(t + 24, &start, 8);
  return ALN_OK;
}

}  // extern "C"

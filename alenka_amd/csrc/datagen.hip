// Seeded synthetic tables generated on the device (counter-based SplitMix64, one stream
// per column), so the 600M-row workloads are built in HBM in seconds.  The streams, ranges
// and date arithmetic are a fixed specification shared with the CPU checker:
//   flat  : key U[0,250000), date = yyyymmdd of U[0,2557) days from 1992-01-01,
//           price = 90000 + U[0,10410001) cents, discount U[0,11) cents
//   dim   : key = row + 1, segment string index U[0,5)
//   orders: key = block*32 + slot + 1 with 4 distinct slots of each 32-key block (Floyd),
//           cust = 1 + U[0,ndim), date as above, priority 0
//   fact  : 1 + U[0,7) rows per order in order-key order; price, discount as above,
//           shipdate = orderdate + 1 + U[0,121) days
// Dictionary codes follow first-seen order (compress_char, cm.cu:1808-1821).
#include <algorithm>
#include <utility>
#include <vector>

#include "aln_internal.h"

namespace aln {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ULL;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += GOLDEN;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint32_t stream) {
  return splitmix64(seed ^ ((uint64_t)stream * 0xD1B54A32D192ED03ULL));
}
__device__ __forceinline__ uint64_t uniform(uint64_t sk, uint64_t i, uint64_t n) {
  return __umul64hi(splitmix64(sk + i * GOLDEN), n);
}
__device__ __forceinline__ int64_t day_to_ymd(int32_t day) {
  int64_t z = (int64_t)day + 8035 + 719468;
  int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  uint64_t doe = (uint64_t)(z - era * 146097);
  uint64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  int64_t y = (int64_t)yoe + era * 400;
  uint64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  uint64_t mp = (5 * doy + 2) / 153;
  uint64_t d = doy - (153 * mp + 2) / 5 + 1;
  uint64_t mo = mp < 10 ? mp + 3 : mp - 9;
  y += (mo <= 2);
  return y * 10000 + (int64_t)mo * 100 + (int64_t)d;
}

enum {
  ST_FLAT_KEY = 1, ST_FLAT_DATE, ST_FLAT_PRICE, ST_FLAT_DISC,
  ST_DIM_SEG = 10,
  ST_ORD_SLOT = 20, ST_ORD_CUST, ST_ORD_DATE,
  ST_FACT_CNT = 30, ST_FACT_PRICE, ST_FACT_DISC, ST_FACT_SHIP,
  ST_ZIPF = 40, ST_SORT_KEY = 41
};
constexpr uint64_t NDAYS = 2557, PRICE_LO = 90000, PRICE_SPAN = 10410001;

__device__ __forceinline__ int64_t order_key(uint64_t kslot, uint64_t i) {
  uint64_t b = i >> 2;
  uint32_t pos = (uint32_t)(i & 3), chosen = 0;
  for (uint32_t j = 28; j < 32; ++j) {
    uint32_t t = (uint32_t)uniform(kslot, b * 4 + (j - 28), j + 1);
    chosen |= (chosen & (1u << t)) ? (1u << j) : (1u << t);
  }
  for (uint32_t p = 0; p < pos; ++p) chosen &= chosen - 1;
  return (int64_t)(b * 32 + (uint64_t)__builtin_ctz(chosen) + 1);
}

#define GRID_LOOP(i, n)                                                          \
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (uint64_t)gridDim.x * blockDim.x)

// Zipf(s) over ranks 1..N by rejection-inversion (W. Hormann, G. Derflinger, "Rejection-
// inversion to generate variates from monotone discrete distributions", ACM TOMACS 6(3),
// 1996): invert the integral H of the hat h(x) = x^-s, round to the nearest rank and accept
// unless the rank falls in the thin sliver between the hat and the histogram.  Trial t of
// row i draws its uniform from counter 8*i + t; after 8 rejections the last rank is kept.
struct ZipfParams {
  double s, hx1, hn, sconst;
  uint64_t n;
};
__host__ __device__ __forceinline__ double zf_expm1_over(double t) {  // (e^t - 1) / t
  return fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5 * (1.0 + t / 3.0 * (1.0 + 0.25 * t));
}
__host__ __device__ __forceinline__ double zf_log1p_over(double t) {  // log(1 + t) / t
  return fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * (0.5 - t * (1.0 / 3.0 - 0.25 * t));
}
__host__ __device__ __forceinline__ double zf_H(double s, double x) {  // integral of x^-s
  const double lx = log(x);
  return zf_expm1_over((1.0 - s) * lx) * lx;
}
__host__ __device__ __forceinline__ double zf_Hinv(double s, double y) {
  double t = y * (1.0 - s);
  if (t < -1.0) t = -1.0;  // rounding at the lower end
  return exp(zf_log1p_over(t) * y);
}
__host__ __device__ __forceinline__ double zf_h(double s, double x) { return exp(-s * log(x)); }

__global__ void k_gen_zipf(uint64_t seed, uint64_t n, ZipfParams z, int64_t *out) {
  const uint64_t sk = stream_key(seed, ST_ZIPF);
  GRID_LOOP(i, n) {
    double k = 1.0;
    for (int t = 0; t < 8; ++t) {
      const uint64_t r = splitmix64(sk + (8 * i + t) * GOLDEN);
      const double u01 = (double)(r >> 11) * (1.0 / 9007199254740992.0);
      const double u = z.hn + u01 * (z.hx1 - z.hn);
      const double x = zf_Hinv(z.s, u);
      k = floor(x + 0.5);
      if (k < 1.0) k = 1.0;
      if (k > (double)z.n) k = (double)z.n;
      if (k - x <= z.sconst || u >= zf_H(z.s, k + 0.5) - zf_h(z.s, k)) break;
    }
    out[i] = (int64_t)k - 1;
  }
}

// Uniform 64-bit keys and their row numbers as payload (the ORDER BY input of the batch run).
__global__ void k_gen_sort_input(uint64_t seed, uint64_t n, int64_t *keys, uint32_t *payload) {
  const uint64_t sk = stream_key(seed, ST_SORT_KEY);
  GRID_LOOP(i, n) {
    keys[i] = (int64_t)splitmix64(sk + i * GOLDEN);
    if (payload) payload[i] = (uint32_t)i;
  }
}

__global__ void k_gen_flat(uint64_t seed, uint64_t n, int64_t *key, int64_t *date, int64_t *price,
                           int64_t *disc) {
  uint64_t k1 = stream_key(seed, ST_FLAT_KEY), k2 = stream_key(seed, ST_FLAT_DATE);
  uint64_t k3 = stream_key(seed, ST_FLAT_PRICE), k4 = stream_key(seed, ST_FLAT_DISC);
  GRID_LOOP(i, n) {
    if (key) key[i] = (int64_t)uniform(k1, i, 250000);
    if (date) date[i] = day_to_ymd((int32_t)uniform(k2, i, NDAYS));
    if (price) price[i] = (int64_t)(PRICE_LO + uniform(k3, i, PRICE_SPAN));
    if (disc) disc[i] = (int64_t)uniform(k4, i, 11);
  }
}

__global__ void k_gen_dim(uint64_t seed, uint64_t n, int64_t *ckey, int64_t *cseg) {
  uint64_t ks = stream_key(seed, ST_DIM_SEG);
  GRID_LOOP(i, n) {
    if (ckey) ckey[i] = (int64_t)i + 1;
    if (cseg) cseg[i] = (int64_t)uniform(ks, i, 5);
  }
}

__global__ void k_gen_orders(uint64_t seed, uint64_t n, uint64_t ndim, int64_t *okey,
                             int64_t *ocust, int64_t *odate, int64_t *oprio) {
  uint64_t ks = stream_key(seed, ST_ORD_SLOT), kc = stream_key(seed, ST_ORD_CUST);
  uint64_t kd = stream_key(seed, ST_ORD_DATE);
  GRID_LOOP(i, n) {
    if (okey) okey[i] = order_key(ks, i);
    if (ocust) ocust[i] = 1 + (int64_t)uniform(kc, i, ndim);
    if (odate) odate[i] = day_to_ymd((int32_t)uniform(kd, i, NDAYS));
    if (oprio) oprio[i] = 0;
  }
}

__global__ void k_gen_fact_counts(uint64_t seed, uint64_t norders, uint32_t *cnt) {
  uint64_t kc = stream_key(seed, ST_FACT_CNT);
  GRID_LOOP(i, norders) cnt[i] = 1 + (uint32_t)uniform(kc, i, 7);
}

__global__ void k_gen_fact(uint64_t seed, uint64_t norders, const uint32_t *cnt,
                           const uint64_t *off, int64_t *lokey, int64_t *lprice, int64_t *ldisc,
                           int64_t *lship) {
  uint64_t kp = stream_key(seed, ST_FACT_PRICE), kd = stream_key(seed, ST_FACT_DISC);
  uint64_t ksh = stream_key(seed, ST_FACT_SHIP), ks = stream_key(seed, ST_ORD_SLOT);
  uint64_t kod = stream_key(seed, ST_ORD_DATE);
  GRID_LOOP(i, norders) {
    uint32_t c = cnt[i];
    uint64_t r = off[i];
    int64_t okey = order_key(ks, i);
    int32_t oday = (int32_t)uniform(kod, i, NDAYS);
    for (uint32_t j = 0; j < c; ++j, ++r) {
      if (lokey) lokey[r] = okey;
      if (lprice) lprice[r] = (int64_t)(PRICE_LO + uniform(kp, r, PRICE_SPAN));
      if (ldisc) ldisc[r] = (int64_t)uniform(kd, r, 11);
      if (lship) lship[r] = day_to_ymd(oday + 1 + (int32_t)uniform(ksh, r, 121));
    }
  }
}

// First row of each value: per-block minima (LDS), one atomic per block and value.
__global__ void k_first_seen(const int64_t *v, uint64_t n, int64_t nvals,
                             unsigned long long *first, uint32_t *bad) {
  __shared__ unsigned long long bmin[64];
  const bool small = nvals <= 64;
  if (small && threadIdx.x < 64) bmin[threadIdx.x] = ~0ull;
  __syncthreads();
  GRID_LOOP(i, n) {
    int64_t x = v[i];
    if (x < 0 || x >= nvals)
      atomicExch(bad, 1u);
    else if (small)
      atomicMin(&bmin[x], (unsigned long long)i);
    else
      atomicMin(&first[x], (unsigned long long)i);
  }
  __syncthreads();
  if (small && threadIdx.x < nvals && bmin[threadIdx.x] != ~0ull)
    atomicMin(&first[threadIdx.x], bmin[threadIdx.x]);
}
__global__ void k_map_codes(const int64_t *v, uint64_t n, const int64_t *code_of, int64_t *out) {
  GRID_LOOP(i, n) out[i] = code_of[v[i]];
}

}  // namespace aln

using namespace aln;

extern "C" {

int aln_gen_flat(aln_ctx *c, uint64_t seed, uint64_t n, int64_t *d_key, int64_t *d_date,
                 int64_t *d_price, int64_t *d_disc) {
  ALN_REQUIRE(c, "aln_gen_flat: ctx is NULL");
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_gen_flat, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, seed, n, d_key,
                     d_date, d_price, d_disc);
  ALN_TRY(check_launch("gen_flat"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_gen_dim(aln_ctx *c, uint64_t seed, uint64_t n, int64_t *d_key, int64_t *d_seg_idx) {
  ALN_REQUIRE(c, "aln_gen_dim: ctx is NULL");
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_gen_dim, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, seed, n, d_key,
                     d_seg_idx);
  ALN_TRY(check_launch("gen_dim"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_gen_orders(aln_ctx *c, uint64_t seed, uint64_t n, uint64_t ndim, int64_t *d_key,
                   int64_t *d_cust, int64_t *d_date, int64_t *d_prio) {
  ALN_REQUIRE(c && ndim > 0, "aln_gen_orders: bad arguments");
  ALN_REQUIRE(n / 4 * 32 + 32 < (1ull << 62), "aln_gen_orders: too many orders");
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_gen_orders, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, seed, n, ndim,
                     d_key, d_cust, d_date, d_prio);
  ALN_TRY(check_launch("gen_orders"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

// Row count of the fact table for `norders` orders (and, if d_off is non-NULL, each
// order's first fact row).
static int fact_offsets(aln_ctx *c, uint64_t seed, uint64_t norders, uint32_t **cnt,
                        uint64_t **off, uint64_t *total) {
  ALN_TRY(ws(c, norders, cnt));
  ALN_TRY(ws(c, norders, off));
  hipLaunchKernelGGL(k_gen_fact_counts, dim3(grid_for(norders, 256, 65536)), dim3(256), 0, c->stream,
                     seed, norders, *cnt);
  ALN_TRY(check_launch("gen_fact_counts"));
  return exclusive_scan_u32(c, *cnt, norders, *off, total);
}

int aln_gen_fact_rows(aln_ctx *c, uint64_t seed, uint64_t norders, uint64_t *h_nfact) {
  ALN_REQUIRE(c && h_nfact, "aln_gen_fact_rows: bad arguments");
  ws_reset(c);
  *h_nfact = 0;
  if (!norders) return ALN_OK;
  uint32_t *cnt;
  uint64_t *off;
  return fact_offsets(c, seed, norders, &cnt, &off, h_nfact);
}

int aln_gen_fact(aln_ctx *c, uint64_t seed, uint64_t norders, int64_t *d_okey, int64_t *d_price,
                 int64_t *d_disc, int64_t *d_ship) {
  ALN_REQUIRE(c, "aln_gen_fact: ctx is NULL");
  ws_reset(c);
  if (!norders) return ALN_OK;
  uint32_t *cnt;
  uint64_t *off, total;
  ALN_TRY(fact_offsets(c, seed, norders, &cnt, &off, &total));
  hipLaunchKernelGGL(k_gen_fact, dim3(grid_for(norders, 256, 65536)), dim3(256), 0, c->stream, seed,
                     norders, cnt, off, d_okey, d_price, d_disc, d_ship);
  ALN_TRY(check_launch("gen_fact"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_dict_first_seen(aln_ctx *c, const int64_t *d_vals, uint64_t n, int64_t nvals,
                        int64_t *h_code_of_val, int64_t *d_codes) {
  ALN_REQUIRE(c && h_code_of_val && nvals > 0 && nvals <= (1 << 20), "aln_dict_first_seen: bad arguments");
  ws_reset(c);
  unsigned long long *first;
  uint32_t *bad;
  int64_t *code_of;
  ALN_TRY(ws(c, (size_t)nvals, &first));
  ALN_TRY(ws(c, (size_t)nvals, &code_of));
  ALN_TRY(ws(c, 1, &bad));
  ALN_TRY(fill_u64(c, (uint64_t *)first, (uint64_t)nvals, ~0ull));
  ALN_TRY(fill_u32(c, bad, 1, 0));
  if (n) {
    hipLaunchKernelGGL(k_first_seen, dim3(grid_for(n, 256 * 64, 4096)), dim3(256), 0, c->stream, d_vals, n,
                       nvals, first, bad);
    ALN_TRY(check_launch("first_seen"));
  }
  std::vector<unsigned long long> hf((size_t)nvals);
  uint32_t hb = 0;
  ALN_TRY(copy_to_host(c, hf.data(), first, 8 * (size_t)nvals));
  ALN_TRY(copy_to_host(c, &hb, bad, 4));
  ALN_REQUIRE(!hb, "aln_dict_first_seen: value outside [0, %lld)", (long long)nvals);
  // rank the values by first occurrence
  std::vector<std::pair<unsigned long long, int64_t>> order;
  for (int64_t v = 0; v < nvals; ++v)
    if (hf[v] != ~0ull) order.push_back({hf[v], v});
  std::sort(order.begin(), order.end());
  for (int64_t v = 0; v < nvals; ++v) h_code_of_val[v] = -1;
  for (size_t r = 0; r < order.size(); ++r) h_code_of_val[order[r].second] = (int64_t)r;
  if (d_codes && n) {
    ALN_HIP(hipMemcpyAsync(code_of, h_code_of_val, 8 * (size_t)nvals, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_map_codes, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_vals, n,
                       code_of, d_codes);
    ALN_TRY(check_launch("map_codes"));
    ALN_HIP(hipStreamSynchronize(c->stream));
  }
  return ALN_OK;
}

int aln_gen_zipf(aln_ctx *c, uint64_t seed, uint64_t n, uint64_t nkeys, double s, int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || d_out), "aln_gen_zipf: bad arguments");
  ALN_REQUIRE(nkeys >= 1 && nkeys < (1ull << 53) && s > 0.0 && s < 64.0, "aln_gen_zipf: nkeys %llu, s %g",
              (unsigned long long)nkeys, s);
  if (!n) return ALN_OK;
  ZipfParams z;
  z.s = s;
  z.n = nkeys;
  z.hx1 = zf_H(s, 1.5) - 1.0;
  z.hn = zf_H(s, (double)nkeys + 0.5);
  z.sconst = 2.0 - zf_Hinv(s, zf_H(s, 2.5) - zf_h(s, 2.0));
  hipLaunchKernelGGL(k_gen_zipf, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, seed, n, z, d_out);
  ALN_TRY(check_launch("gen_zipf"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_gen_sort_input(aln_ctx *c, uint64_t seed, uint64_t n, int64_t *d_keys, uint32_t *d_payload) {
  ALN_REQUIRE(c && (n == 0 || d_keys), "aln_gen_sort_input: bad arguments");
  if (!n) return ALN_OK;
  hipLaunchKernelGGL(k_gen_sort_input, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, seed, n,
                     d_keys, d_payload);
  ALN_TRY(check_launch("gen_sort_input"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

}  // extern "C"

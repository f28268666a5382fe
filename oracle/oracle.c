/*
 * oracle.c - CPU restatement of the reference (antonmks/Alenka) relational operators.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product links it.
 *
 * Parity pins (what keeps each restatement honest; see tests/test_oracle.py):
 *   orc_murmur64a        - the reference's own MurmurHash2_64.cu compiled by
 *                          oracle/build_ref.sh into oracle/_ref (bit-equal on random keys
 *                          of every length 0..64), and the vectors in tests/golden, which
                          tests/golden/make_murmur_vectors.py wrote from that same build.
 *   orc_day_to_ymd       - Python datetime over every day of the generator's range.
 *   orc_pack/unpack      - round trip + a hand-laid 3-bit example.
 *   orc_pfor_*           - byte layout vs compress.cu:388-399/511-522 fields read back by
 *                          struct offsets, and round trips (FOR, FOR-DELTA).
 *   orc_filter/orc_eval  - known-answer tables for every operand pairing, derived by hand
 *                          from filter.cu/cm.cu (including the reference's scaling quirks),
 *                          plus brute force with Python integers for the quirk-free forms.
 *   orc_zone_map_check   - soundness invariant: 'A' => every row in [min,max] passes,
 *                          'N' => none does, for quirk-free predicates; hand cases for the
 *                          quirks.
 *   orc_groupby_*        - brute-force dictionary aggregation (Python) + ordering rules.
 *   orc_join_sortmerge   - nested-loop multiset + the documented output order.
 *   orc_str_sort         - Python's stable sorted() on the values' bytes, and the
 *                          widths the reference leaves unsorted.
 *   orc_semi_anti        - Python set membership over all key pairs (brute force), and the
 *                          stable key order of the right side.
 *   orc_order            - Python's stable sorted() with a composite key.
 *   orc_compact          - numpy's flatnonzero(mask) + row0 on random masks of every
 *                          density, the empty mask, an all-true mask and last-row-only.
 *   orc_kahan_sum_f64    - math.fsum (exactly rounded) within Kahan's error bound
 *                          2u*sum|x|; inputs where a naive sum is off by far more than that
 *                          bound, and cancelling sums whose exact value is known.
 *   orc_like             - Python's re.fullmatch with '%' -> '.*' and '?' -> '[^.]' on
 *                          values and patterns too short for the restart counter to bind,
 *                          plus hand-derived cases where it does.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* MurmurHash64A, restated from MurmurHash2_64.cu:63-104 (Austin Appleby's public-domain
 * algorithm, as the reference ships it).  Reads the key as little-endian 64-bit words.  */
uint64_t orc_murmur64a(const void *key, int len, uint32_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  const int r = 47;
  uint64_t h = (uint64_t)seed ^ ((uint64_t)(int64_t)len * m);
  const unsigned char *p = (const unsigned char *)key;
  int nblocks = len / 8;
  for (int b = 0; b < nblocks; ++b) {
    uint64_t k;
    memcpy(&k, p + 8 * b, 8);
    k *= m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
  }
  const unsigned char *tail = p + 8 * nblocks;
  switch (len & 7) {
  case 7: h ^= (uint64_t)tail[6] << 48; /* fall through */
  case 6: h ^= (uint64_t)tail[5] << 40; /* fall through */
  case 5: h ^= (uint64_t)tail[4] << 32; /* fall through */
  case 4: h ^= (uint64_t)tail[3] << 24; /* fall through */
  case 3: h ^= (uint64_t)tail[2] << 16; /* fall through */
  case 2: h ^= (uint64_t)tail[1] << 8;  /* fall through */
  case 1:
    h ^= (uint64_t)tail[0];
    h *= m;
  }
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

/* ------------------------------------------------------------------------------------ */
/* Counter-based SplitMix64 streams for the synthetic workloads (one stream per column).  */
#define GOLDEN 0x9E3779B97F4A7C15ULL
uint64_t orc_splitmix64(uint64_t x) {
  x += GOLDEN;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
uint64_t orc_stream_key(uint64_t seed, uint32_t stream) {
  return orc_splitmix64(seed ^ ((uint64_t)stream * 0xD1B54A32D192ED03ULL));
}
uint64_t orc_rnd(uint64_t sk, uint64_t i) { return orc_splitmix64(sk + i * GOLDEN); }
uint64_t orc_uniform(uint64_t sk, uint64_t i, uint64_t n) {
  return (uint64_t)(((unsigned __int128)orc_rnd(sk, i) * n) >> 64);
}

/* day 0 = 1992-01-01 -> yyyymmdd (proleptic Gregorian civil-from-days). */
int32_t orc_day_to_ymd(int32_t day) {
  int64_t z = (int64_t)day + 8035 + 719468; /* 8035 = days 1970-01-01 .. 1992-01-01 */
  int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  uint64_t doe = (uint64_t)(z - era * 146097);
  uint64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  int64_t y = (int64_t)yoe + era * 400;
  uint64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  uint64_t mp = (5 * doy + 2) / 153;
  uint64_t d = doy - (153 * mp + 2) / 5 + 1;
  uint64_t mo = mp < 10 ? mp + 3 : mp - 9;
  y += (mo <= 2);
  return (int32_t)(y * 10000 + (int64_t)mo * 100 + (int64_t)d);
}

enum {
  ST_FLAT_KEY = 1, ST_FLAT_DATE, ST_FLAT_PRICE, ST_FLAT_DISC,
  ST_DIM_SEG = 10,
  ST_ORD_SLOT = 20, ST_ORD_CUST, ST_ORD_DATE,
  ST_FACT_CNT = 30, ST_FACT_PRICE, ST_FACT_DISC, ST_FACT_SHIP
};
#define NDAYS 2557            /* 1992-01-01 .. 1998-12-31 */
#define PRICE_LO 90000        /* 900.00 in cents   */
#define PRICE_SPAN 10410001   /* .. 105000.00      */

/* flat filter+group-by table: key U[0,250000), date, price, discount */
void orc_gen_flat(uint64_t seed, uint64_t n, int64_t *key, int64_t *date, int64_t *price,
                  int64_t *disc) {
  uint64_t k1 = orc_stream_key(seed, ST_FLAT_KEY), k2 = orc_stream_key(seed, ST_FLAT_DATE);
  uint64_t k3 = orc_stream_key(seed, ST_FLAT_PRICE), k4 = orc_stream_key(seed, ST_FLAT_DISC);
  for (uint64_t i = 0; i < n; ++i) {
    key[i] = (int64_t)orc_uniform(k1, i, 250000);
    date[i] = orc_day_to_ymd((int32_t)orc_uniform(k2, i, NDAYS));
    price[i] = PRICE_LO + (int64_t)orc_uniform(k3, i, PRICE_SPAN);
    disc[i] = (int64_t)orc_uniform(k4, i, 11);
  }
}

void orc_gen_dim(uint64_t seed, uint64_t ndim, int64_t *c_key, int64_t *c_seg_idx) {
  uint64_t ks = orc_stream_key(seed, ST_DIM_SEG);
  for (uint64_t i = 0; i < ndim; ++i) {
    c_key[i] = (int64_t)i + 1;
    c_seg_idx[i] = (int64_t)orc_uniform(ks, i, 5);
  }
}

/* order key: 4 used slots of every 32-key block, chosen by Floyd's sampling */
static int64_t order_key(uint64_t kslot, uint64_t i) {
  uint64_t b = i >> 2;
  uint32_t pos = (uint32_t)(i & 3), chosen = 0;
  for (uint32_t j = 28; j < 32; ++j) {
    uint32_t t = (uint32_t)orc_uniform(kslot, b * 4 + (j - 28), j + 1);
    chosen |= (chosen & (1u << t)) ? (1u << j) : (1u << t);
  }
  for (uint32_t p = 0; p < pos; ++p) chosen &= chosen - 1;
  return (int64_t)(b * 32 + (uint64_t)__builtin_ctz(chosen) + 1);
}

void orc_gen_orders(uint64_t seed, uint64_t norders, uint64_t ndim, int64_t *o_key,
                    int64_t *o_cust, int64_t *o_date, int64_t *o_prio) {
  uint64_t ks = orc_stream_key(seed, ST_ORD_SLOT), kc = orc_stream_key(seed, ST_ORD_CUST);
  uint64_t kd = orc_stream_key(seed, ST_ORD_DATE);
  for (uint64_t i = 0; i < norders; ++i) {
    o_key[i] = order_key(ks, i);
    o_cust[i] = 1 + (int64_t)orc_uniform(kc, i, ndim);
    o_date[i] = orc_day_to_ymd((int32_t)orc_uniform(kd, i, NDAYS));
    o_prio[i] = 0;
  }
}

uint64_t orc_gen_fact_rows(uint64_t seed, uint64_t norders) {
  uint64_t kc = orc_stream_key(seed, ST_FACT_CNT), n = 0;
  for (uint64_t i = 0; i < norders; ++i) n += 1 + orc_uniform(kc, i, 7);
  return n;
}

void orc_gen_fact(uint64_t seed, uint64_t norders, int64_t *l_okey, int64_t *l_price,
                  int64_t *l_disc, int64_t *l_ship) {
  uint64_t kc = orc_stream_key(seed, ST_FACT_CNT), kp = orc_stream_key(seed, ST_FACT_PRICE);
  uint64_t kd = orc_stream_key(seed, ST_FACT_DISC), ksh = orc_stream_key(seed, ST_FACT_SHIP);
  uint64_t ks = orc_stream_key(seed, ST_ORD_SLOT), kod = orc_stream_key(seed, ST_ORD_DATE);
  uint64_t r = 0;
  for (uint64_t i = 0; i < norders; ++i) {
    uint64_t cnt = 1 + orc_uniform(kc, i, 7);
    int64_t okey = order_key(ks, i);
    int32_t oday = (int32_t)orc_uniform(kod, i, NDAYS);
    for (uint64_t j = 0; j < cnt; ++j, ++r) {
      l_okey[r] = okey;
      l_price[r] = PRICE_LO + (int64_t)orc_uniform(kp, r, PRICE_SPAN);
      l_disc[r] = (int64_t)orc_uniform(kd, r, 11);
      l_ship[r] = orc_day_to_ymd(oday + 1 + (int32_t)orc_uniform(ksh, r, 121));
    }
  }
}

/* Dictionary codes in first-seen order (compress_char, cm.cu:1808-1821). */
void orc_dict_first_seen(const int64_t *vals, uint64_t n, int64_t nvals, int64_t *code_of_val,
                         int64_t *codes) {
  int64_t next = 0;
  for (int64_t v = 0; v < nvals; ++v) code_of_val[v] = -1;
  for (uint64_t i = 0; i < n; ++i) {
    int64_t v = vals[i];
    if (code_of_val[v] < 0) code_of_val[v] = next++;
    codes[i] = code_of_val[v];
  }
}

/* ------------------------------------------------------------------------------------ */
/* LIKE, restated from the reference's gpu_regex functor (filter.cu:67-129): '%' matches
 * any run (backtracking to the last '%'), '?' any character except '.', anything else
 * itself.  Kept as the reference has it: the position counter j is not reset when the
 * match restarts after a '%', and a '%' at the end of the value is only skipped after the
 * loop.  The value is read as at most len bytes, stopping at a NUL.                        */
int orc_like(const char *s0, uint32_t len, const char *pat0) {
  int star = 0;
  uint32_t j = 0;
  const char *s, *p, *str = s0, *pat = pat0;
restart:
  for (s = str, p = pat; j < len && *s; ++s, ++p, ++j) {
    switch (*p) {
      case '?':
        if (*s == '.') goto mismatch;
        break;
      case '%':
        star = 1;
        str = s;
        pat = p;
        do {
          ++pat;
        } while (*pat == '%');
        if (!*pat) return 1;
        goto restart;
      default:
        if (*s != *p) goto mismatch;
        break;
    }
  }
  while (*p == '%') ++p;
  return !*p;
mismatch:
  if (!star) return 0;
  str++;
  j++;
  goto restart;
}

/* ------------------------------------------------------------------------------------ */
/* The build's bit-packed frame-of-reference layout (dense, LSB-first, 32-bit words).   */
uint32_t orc_bits_for_range(int64_t lo, int64_t hi) {
  uint64_t r = (uint64_t)hi - (uint64_t)lo;
  return r == 0 ? 0u : (uint32_t)(64 - __builtin_clzll(r));
}

void orc_pack(const int64_t *v, uint64_t n, int64_t base, uint32_t bits, uint32_t *words) {
  uint64_t nw = (n * bits + 31) / 32;
  memset(words, 0, nw * 4);
  for (uint64_t i = 0; i < n && bits; ++i) {
    uint64_t code = (uint64_t)v[i] - (uint64_t)base;
    uint64_t p = i * bits;
    for (uint32_t b = 0; b < bits; ++b, ++p)
      if ((code >> b) & 1) words[p >> 5] |= 1u << (p & 31);
  }
}

int64_t orc_unpack(const uint32_t *words, uint64_t i, int64_t base, uint32_t bits) {
  uint64_t code = 0, p = i * bits;
  for (uint32_t b = 0; b < bits; ++b, ++p)
    code |= (uint64_t)((words[p >> 5] >> (p & 31)) & 1) << b;
  return (int64_t)((uint64_t)base + code);
}

/* ------------------------------------------------------------------------------------ */
/* The reference's segment file format (compress.cu:411-527 writer, 180-267 reader):
 *   [u32 cnt][i64 lower][i64 upper][u32 comp_type][payload: cnt bytes]
 *   [u32 cnt][u32 recCount][u32 bits][i64 orig_lower][u32 fit_count][i64 start_val]    */
static void put(uint8_t **p, const void *v, int n) { memcpy(*p, v, n); *p += n; }

static uint32_t ceil_log2_range(int64_t lo, int64_t hi) {
  return (uint32_t)ceil(log2((double)((hi - lo) + 1)));
}

/* FOR with byte-aligned widths (pfor_compress, compress.cu:411-527, delta off). */
uint64_t orc_pfor_encode(const int64_t *v, uint32_t n, uint8_t *out, uint64_t cap) {
  int64_t lo = v[0], hi = v[0];
  for (uint32_t i = 1; i < n; ++i) { if (v[i] < lo) lo = v[i]; if (v[i] > hi) hi = v[i]; }
  uint32_t bits = ceil_log2_range(lo, hi);
  if (bits != 8 && bits != 16 && bits != 32 && bits != 64)
    bits = bits < 8 ? 8 : bits < 16 ? 16 : bits < 32 ? 32 : 64;
  uint32_t cnt = n * (bits / 8), comp = 0, fit = 64 / bits;
  int64_t start = 0;
  uint64_t total = 56 + (uint64_t)cnt;
  if (total > cap) return total;
  uint8_t *p = out;
  put(&p, &cnt, 4); put(&p, &lo, 8); put(&p, &hi, 8); put(&p, &comp, 4);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t c = (uint64_t)v[i] - (uint64_t)lo;
    put(&p, &c, (int)(bits / 8));
  }
  put(&p, &cnt, 4); put(&p, &n, 4); put(&p, &bits, 4); put(&p, &lo, 8); put(&p, &fit, 4);
  put(&p, &start, 8);
  return total;
}

/* FOR-DELTA, MSB-first within 64-bit words (pfor_delta_compress, compress.cu:270-408). */
uint64_t orc_pfor_delta_encode(const int64_t *v, uint32_t n, uint8_t *out, uint64_t cap) {
  int64_t *d = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  d[0] = v[0];
  for (uint32_t i = 1; i < n; ++i) d[i] = v[i] - v[i - 1];
  int64_t start = d[0];
  if (n > 1) d[0] = d[1];
  int64_t lo = d[0], hi = d[0];
  for (uint32_t i = 1; i < n; ++i) { if (d[i] < lo) lo = d[i]; if (d[i] > hi) hi = d[i]; }
  int64_t rlo = v[0], rhi = v[n - 1];
  uint32_t bits = ceil_log2_range(lo, hi);
  if (bits == 0) bits = 1;
  uint32_t fit = 64 / bits, nw = (n + fit - 1) / fit, cnt = nw * 8, comp = 1;
  uint64_t total = 56 + (uint64_t)cnt;
  if (total > cap) { free(d); return total; }
  uint64_t *w = (uint64_t *)calloc(nw ? nw : 1, 8);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t val = (uint64_t)(d[i] - lo);
    uint32_t sh = 64 - bits - (i % fit) * bits;
    w[i / fit] += val << sh;
  }
  uint8_t *p = out;
  put(&p, &cnt, 4); put(&p, &rlo, 8); put(&p, &rhi, 8); put(&p, &comp, 4);
  put(&p, w, (int)cnt);
  put(&p, &cnt, 4); put(&p, &n, 4); put(&p, &bits, 4); put(&p, &lo, 8); put(&p, &fit, 4);
  put(&p, &start, 8);
  free(w);
  free(d);
  return total;
}

/* Decoder (pfor_decompress, compress.cu:180-267, phase_copy off).  Returns recCount. */
int64_t orc_pfor_decode(const uint8_t *buf, uint64_t len, int64_t *out, uint64_t cap) {
  uint32_t cnt, comp, rec, bits, fit;
  int64_t olo, start;
  if (len < 56) return -1;
  memcpy(&cnt, buf, 4);
  memcpy(&comp, buf + 20, 4);
  if (len < 56 + (uint64_t)cnt) return -1;
  const uint8_t *t = buf + 24 + cnt;
  memcpy(&rec, t + 4, 4);
  memcpy(&bits, t + 8, 4);
  memcpy(&olo, t + 12, 8);
  memcpy(&fit, t + 20, 4);
  memcpy(&start, t + 24, 8);
  if (rec > cap) return -1;
  const uint8_t *pl = buf + 24;
  if (comp == 1) {
    for (uint32_t i = 0; i < rec; ++i) {
      uint64_t word;
      memcpy(&word, pl + 8 * (i / fit), 8);
      uint64_t tmp = word >> (64 - bits - (i % fit) * bits);
      if (bits < 64) { tmp <<= (64 - bits); tmp >>= (64 - bits); }
      out[i] = (int64_t)(tmp + (uint64_t)olo);
    }
    if (rec) out[0] = start;
    for (uint32_t i = 1; i < rec; ++i) out[i] += out[i - 1];
  } else {
    for (uint32_t i = 0; i < rec; ++i) {
      uint64_t c = 0;
      memcpy(&c, pl + (uint64_t)i * (bits / 8), bits / 8);
      out[i] = (int64_t)(c + (uint64_t)olo);
    }
  }
  return rec;
}

/* ------------------------------------------------------------------------------------ */
/* RPN evaluation, restating filter.cu:133-643 (predicates) and select.cu:703-790
 * (expressions) with CudaSet::op (cm.cu:2480-2655), CudaSet::compare (cm.cu:2132-2379),
 * precision_func (zone_map.cu:668-712), get_vec/get_decimals (cm.cu:4338-4379) and
 * reverse_op (cm.cu:3166-3181).  Stack pops are kept in the reference's order, so its
 * decimal-scaling behaviour is reproduced, including:
 *   - NAME op NUMBER with a column operand of nonzero scale (cm.cu:2488-2510) scales the
 *     constant twice for ADD/MINUS/DIV;
 *   - NUMBER CMP NAME scales by the raw precisions, not their difference (filter.cu:598);
 *   - constant folding NUMBER NUMBER MINUS/DIV uses the top of stack as left operand.   */

enum { T_NAME = 1, T_NUMBER = 2, T_STRING = 3, T_FLOAT = 5 };
#define SMAX 64

typedef struct {
  const int64_t *const *cols;
  const int32_t *dec;
  uint64_t n;
  int type[SMAX]; int ntype;
  int value[SMAX]; int nvalue;          /* column slot, or -1 for a computed vector     */
  int64_t nums[SMAX]; int nnums;
  uint32_t prec[SMAX]; int nprec;
  int64_t *vecs[SMAX]; int nvecs;
  int64_t strs[SMAX]; int nstrs;        /* dictionary codes of STRING literals          */
  uint8_t *bools[SMAX]; int nbools;
  double fnums[SMAX]; int nfnums;       /* FLOAT operands                                */
  int vfloat[SMAX];                     /* per vecs entry: holds double bits            */
  int64_t *owned[4 * SMAX]; int nowned;
  int mode; /* 0: filter.cu, 1: select.cu */
  int err;
} rpn_state;

static uint32_t precision_func(uint32_t *p1, uint32_t *p2, int op) {
  if (op != ALN_RPN_DIV) {
    if (op != ALN_RPN_MUL) {
      uint32_t res;
      if (*p1 > *p2) { res = *p1; *p2 = *p1 - *p2; *p1 = 0; }
      else { res = *p1; *p1 = *p2 - *p1; *p2 = 0; }
      return res;
    }
    uint32_t res = *p1 + *p2;
    *p1 = 0; *p2 = 0;
    return res;
  }
  if (*p1 == *p2) { *p1 = *p1 + 4; *p2 = 0; return *p1; }
  if (*p1 > *p2) { *p1 = *p1 + (*p1 - *p2) + 4; *p2 = 0; return *p1; }
  *p2 = *p2 + (*p2 - *p1) + 4; *p1 = 0; return *p2;
}

static int pw10(uint32_t p, int64_t *out) {
  static const int64_t t[10] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000,
                                100000000, 1000000000};
  if (p > 9) return -1;
  *out = t[p];
  return 0;
}
static int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
static int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static int64_t wdiv(int64_t a, int64_t b) {
  if (b == 0) return 0;
  if (a == INT64_MIN && b == -1) return INT64_MIN;
  return a / b;
}
static int64_t arith(int op, int64_t l, int64_t r) {
  switch (op) {
  case ALN_RPN_ADD: return wadd(l, r);
  case ALN_RPN_MINUS: return wsub(l, r);
  case ALN_RPN_MUL: return wmul(l, r);
  default: return wdiv(l, r);
  }
}
static int cmpv(int code, int64_t l, int64_t r) {
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

static int64_t *new_vec(rpn_state *s) {
  int64_t *v = (int64_t *)malloc(sizeof(int64_t) * (s->n ? s->n : 1));
  s->owned[s->nowned++] = v;
  return v;
}
static uint8_t *new_bool(rpn_state *s) {
  uint8_t *v = (uint8_t *)malloc(s->n ? s->n : 1);
  s->owned[s->nowned++] = (int64_t *)v;
  return v;
}
#define POP(arr, cnt) ((cnt) > 0 ? (arr)[--(cnt)] : (s->err = ALN_E_ARG, (arr)[0]))

static const int64_t *get_vec(rpn_state *s, int v) {
  if (v >= 0) return s->cols[v];
  if (s->nvecs == 0) { s->err = ALN_E_ARG; return NULL; }
  return s->vecs[--s->nvecs];
}

/* ---- double precision: CudaSet's float overloads (cm.cu:2178-2437, 2440-2773) ---------
 * An operation with a double operand - a FLOAT (emit_float, operators.cu:114-118), an f64
 * column (col_dec ALN_DEC_F64) or a vector computed from one - is evaluated in doubles:
 * an int column converts by a plain cast of the stored integer (long_to_float_type,
 * cm.cu:137-144), a NUMBER by digits / 10^precision; + - * / are IEEE double; comparisons
 * are the EPSILON functors (cm.h:17; f_less .. f_not_equal_to, cm.cu:81-134) applied as
 * (left = deeper operand, right = top). */
static double dbits(int64_t b) { double x; memcpy(&x, &b, 8); return x; }
static int64_t bitsd(double x) { int64_t b; memcpy(&b, &x, 8); return b; }
static int is_f64_col(const rpn_state *s, int v) { return v >= 0 && s->dec[v] == ALN_DEC_F64; }

static int f_compare(int code, double x, double y) {
  const double e = 1.0E-8;
  switch (code) {
  case 1: return (y - x) > e;                                     /* f_less             */
  case 2: return (x - y) > e;                                     /* f_greater          */
  case 3: return (x - y) > e || (x - y) < -e;                     /* f_not_equal_to     */
  case 4: return (x - y) < e && (x - y) > -e;                     /* f_equal_to         */
  case 5: return (y - x) > e || ((x - y) < e && (x - y) > -e);    /* f_less_equal       */
  case 6: return (x - y) > e || ((x - y) < e && (x - y) > -e);    /* f_greater_equal_to */
  default: return 0;
  }
}

/* whether either of the two top operands is a double, without popping */
static int float_operands(const rpn_state *s) {
  if (s->ntype < 2) return 0;
  int vi = s->nvalue, ci = s->nvecs;
  for (int k = 0; k < 2; ++k) {
    int t = s->type[s->ntype - 1 - k];
    if (t == T_FLOAT) return 1;
    if (t == T_NAME || t == T_STRING) {
      if (vi == 0) return 0;
      int v = s->value[--vi];
      if (t != T_NAME) continue;
      if (v >= 0) { if (is_f64_col(s, v)) return 1; }
      else if (ci > 0 && s->vfloat[--ci]) return 1;
    }
  }
  return 0;
}

typedef struct { const int64_t *vec; int is_float, is_const; double c; } f_operand;

static f_operand f_pop(rpn_state *s, int t) {
  f_operand o = {NULL, 0, 0, 0.0};
  if (t == T_FLOAT) {
    o.is_const = 1;
    o.c = POP(s->fnums, s->nfnums);
  } else if (t == T_NUMBER) {
    int64_t n = POP(s->nums, s->nnums);
    uint32_t p = POP(s->prec, s->nprec);
    double d = 1.0;
    if (p > 18) { s->err = ALN_E_UNSUPPORTED; return o; }
    for (uint32_t i = 0; i < p; ++i) d *= 10.0;
    o.is_const = 1;
    o.c = (double)n / d;
  } else if (t == T_NAME) {
    int v = POP(s->value, s->nvalue);
    if (v >= 0) {
      o.vec = s->cols[v];
      o.is_float = is_f64_col(s, v);
    } else {
      if (s->nvecs == 0) { s->err = ALN_E_UNSUPPORTED; return o; }
      o.is_float = s->vfloat[s->nvecs - 1];
      o.vec = get_vec(s, v);
      (void)POP(s->prec, s->nprec);
    }
  } else {
    s->err = ALN_E_UNSUPPORTED;
  }
  return o;
}

static double f_at(const f_operand *o, uint64_t i) {
  if (o->is_const) return o->c;
  return o->is_float ? dbits(o->vec[i]) : (double)o->vec[i];
}

static double f_op(int op, double l, double r) {
  switch (op) {
  case ALN_RPN_ADD: return l + r;
  case ALN_RPN_MINUS: return l - r;
  case ALN_RPN_MUL: return l * r;
  default: return l / r;
  }
}

static void f_arith(rpn_state *s, int op) {
  int s1 = POP(s->type, s->ntype), s2 = POP(s->type, s->ntype);
  f_operand r = f_pop(s, s1), l = f_pop(s, s2);
  if (s->err) return;
  if (l.is_const && r.is_const) {
    s->fnums[s->nfnums++] = f_op(op, l.c, r.c);
    s->type[s->ntype++] = T_FLOAT;
    return;
  }
  int64_t *out = new_vec(s);
  for (uint64_t i = 0; i < s->n; ++i) out[i] = bitsd(f_op(op, f_at(&l, i), f_at(&r, i)));
  s->prec[s->nprec++] = 0;
  s->type[s->ntype++] = T_NAME;
  s->value[s->nvalue++] = -1;
  s->vfloat[s->nvecs] = 1;
  s->vecs[s->nvecs++] = out;
}

static void f_cmp(rpn_state *s, int code) {
  int s1 = POP(s->type, s->ntype), s2 = POP(s->type, s->ntype);
  f_operand r = f_pop(s, s1), l = f_pop(s, s2);
  if (s->err) return;
  uint8_t *out = new_bool(s);
  for (uint64_t i = 0; i < s->n; ++i) out[i] = (uint8_t)f_compare(code, f_at(&l, i), f_at(&r, i));
  s->type[s->ntype++] = T_NAME;
  s->bools[s->nbools++] = out;
}
static uint32_t get_decimals(rpn_state *s, int v) {
  if (v >= 0) return (uint32_t)s->dec[v];
  if (s->nprec == 0) { s->err = ALN_E_ARG; return 0; }
  return s->prec[--s->nprec];
}

/* CudaSet::op(int_type*, int_type d, ...) cm.cu:2480-2546 */
static void op_col_const(rpn_state *s, const int64_t *col, int64_t d, int op, int reverse,
                         uint32_t p1, uint32_t p2, int64_t *out) {
  int64_t d1 = d, s1 = 1, s2 = 1;
  if (pw10(p1, &s1) || pw10(p2, &s2)) { s->err = ALN_E_UNSUPPORTED; return; }
  if (p2) d = wmul(d, s2);
  for (uint64_t i = 0; i < s->n; ++i) {
    int64_t c = col[i], r;
    if (!reverse) {
      if (op == ALN_RPN_MUL) r = wmul(c, d1);
      else if (!p1) r = arith(op, c, wmul(d, s2));
      else r = arith(op, wmul(c, s1), d);
    } else {
      r = arith(op, d, p1 ? wmul(c, s1) : c);
    }
    out[i] = r;
  }
}

/* CudaSet::op(int_type*, int_type*, ..., reverse = 0) cm.cu:2552-2605 */
static void op_col_col(rpn_state *s, const int64_t *c1, const int64_t *c2, int op, uint32_t p1,
                       uint32_t p2, int64_t *out) {
  int64_t s1 = 1, s2 = 1;
  if (pw10(p1, &s1) || pw10(p2, &s2)) { s->err = ALN_E_UNSUPPORTED; return; }
  for (uint64_t i = 0; i < s->n; ++i) {
    if (op == ALN_RPN_MUL) { out[i] = wmul(c1[i], c2[i]); continue; }
    int64_t l = p1 ? wmul(c1[i], s1) : c1[i];
    int64_t r = p2 ? wmul(c2[i], s2) : c2[i];
    out[i] = arith(op, l, r);
  }
}

static void rpn_arith(rpn_state *s, int op) {
  if (float_operands(s)) { f_arith(s, op); return; }
  int s1 = POP(s->type, s->ntype), s2 = POP(s->type, s->ntype);
  if (s1 == T_NUMBER && s2 == T_NUMBER) {
    int64_t n1 = POP(s->nums, s->nnums), n2 = POP(s->nums, s->nnums);
    uint32_t p1 = POP(s->prec, s->nprec), p2 = POP(s->prec, s->nprec);
    uint32_t pres = precision_func(&p1, &p2, op);
    int64_t a, b;
    if (pw10(p1, &a) || pw10(p2, &b)) { s->err = ALN_E_UNSUPPORTED; return; }
    s->prec[s->nprec++] = pres;
    if (p1) n1 = wmul(n1, a);
    if (p2) n2 = wmul(n2, b);
    int64_t res = arith(op, n1, n2), *v = new_vec(s);
    for (uint64_t i = 0; i < s->n; ++i) v[i] = res;
    s->type[s->ntype++] = T_NAME;
    s->value[s->nvalue++] = -1;
    s->vfloat[s->nvecs] = 0;
    s->vecs[s->nvecs++] = v;
  } else if (s1 == T_NAME && s2 == T_NUMBER) {
    int v1 = POP(s->value, s->nvalue);
    int64_t n1 = POP(s->nums, s->nnums);
    uint32_t p1 = POP(s->prec, s->nprec);
    uint32_t p2 = get_decimals(s, v1);
    const int64_t *t = get_vec(s, v1);
    /* filter.cu:285-289 passes (p1, p2); select.cu:754-760 passes (p2, p1) */
    uint32_t pres = s->mode == 0 ? precision_func(&p1, &p2, op) : precision_func(&p2, &p1, op);
    s->prec[s->nprec++] = pres;
    s->type[s->ntype++] = T_NAME;
    s->value[s->nvalue++] = -1;
    int64_t *out = new_vec(s);
    if (t) {
      if (s->mode == 0) op_col_const(s, t, n1, op, 1, p1, p2, out);
      else op_col_const(s, t, n1, op, 1, p2, p1, out);
    }
    s->vfloat[s->nvecs] = 0;
    s->vecs[s->nvecs++] = out;
  } else if (s1 == T_NUMBER && s2 == T_NAME) {
    int64_t n1 = POP(s->nums, s->nnums);
    int v2 = POP(s->value, s->nvalue);
    uint32_t p1 = POP(s->prec, s->nprec);
    uint32_t p2 = get_decimals(s, v2);
    const int64_t *t = get_vec(s, v2);
    uint32_t pres = precision_func(&p2, &p1, op);
    s->prec[s->nprec++] = pres;
    s->type[s->ntype++] = T_NAME;
    s->value[s->nvalue++] = -1;
    int64_t *out = new_vec(s);
    if (t) op_col_const(s, t, n1, op, 0, p2, p1, out);
    s->vfloat[s->nvecs] = 0;
    s->vecs[s->nvecs++] = out;
  } else if (s1 == T_NAME && s2 == T_NAME) {
    int v1 = POP(s->value, s->nvalue), v2 = POP(s->value, s->nvalue);
    const int64_t *t1 = get_vec(s, v1), *t = get_vec(s, v2);
    uint32_t p1 = get_decimals(s, v1), p2 = get_decimals(s, v2);
    uint32_t pres = precision_func(&p1, &p2, op);
    s->prec[s->nprec++] = pres;
    s->type[s->ntype++] = T_NAME;
    s->value[s->nvalue++] = -1;
    int64_t *out = new_vec(s);
    if (t && t1) op_col_col(s, t, t1, op, p2, p1, out);
    s->vfloat[s->nvecs] = 0;
    s->vecs[s->nvecs++] = out;
  } else {
    s->err = ALN_E_UNSUPPORTED; /* arithmetic on strings: no result in the reference */
  }
}

/* CudaSet::compare(int_type* column1, int_type d, op, p1, p2) cm.cu:2250-2295 */
static void cmp_col_const(rpn_state *s, const int64_t *col, int64_t d, int code, uint32_t p1,
                          uint32_t p2, uint8_t *out) {
  int64_t s1 = 1, s2 = 1;
  if (pw10(p1, &s1) || pw10(p2, &s2)) { s->err = ALN_E_UNSUPPORTED; return; }
  if (p2) d = wmul(d, s2);
  for (uint64_t i = 0; i < s->n; ++i) out[i] = (uint8_t)cmpv(code, p1 ? wmul(col[i], s1) : col[i], d);
}

static void rpn_cmp(rpn_state *s, int code) {
  if (code >= 1 && code <= 6 && float_operands(s)) { f_cmp(s, code); return; }
  int s1 = POP(s->type, s->ntype), s2 = POP(s->type, s->ntype);
  if (s1 == T_FLOAT || s2 == T_FLOAT) { s->err = ALN_E_UNSUPPORTED; return; }
  uint8_t *out = new_bool(s);
  if (s1 == T_NUMBER && s2 == T_NUMBER) {
    int64_t n1 = POP(s->nums, s->nnums), n2 = POP(s->nums, s->nnums);
    s->type[s->ntype++] = T_NAME;
    s->value[s->nvalue++] = -1;
    uint32_t p1 = POP(s->prec, s->nprec), p2 = POP(s->prec, s->nprec);
    uint32_t pres = precision_func(&p1, &p2, 0);
    int64_t a, b;
    if (pw10(p1, &a) || pw10(p2, &b)) { s->err = ALN_E_UNSUPPORTED; return; }
    s->prec[s->nprec++] = pres;
    if (p1) n1 = wmul(n1, a);
    if (p2) n2 = wmul(n2, b);
    uint8_t r = (uint8_t)cmpv(code, n2, n1); /* compare(s=n1, d=n2): d OP s, cm.cu:2132 */
    for (uint64_t i = 0; i < s->n; ++i) out[i] = r;
  } else if ((s1 == T_STRING && s2 == T_NAME) || (s1 == T_NAME && s2 == T_STRING)) {
    int v1 = POP(s->value, s->nvalue), v2 = POP(s->value, s->nvalue);
    int col = (s1 == T_NAME) ? v1 : v2;
    /* a STRING's value entry is -2 - index into strs[] (its dictionary code) */
    int sv = (s1 == T_STRING) ? v1 : v2;
    int64_t lit = (sv <= -2 && -2 - sv < s->nstrs) ? s->strs[-2 - sv] : -1;
    if (code != 3 && code != 4) { s->err = ALN_E_UNSUPPORTED; return; }
    if (col < 0) { s->err = ALN_E_UNSUPPORTED; return; }
    for (uint64_t i = 0; i < s->n; ++i)
      out[i] = (uint8_t)(code == 4 ? s->cols[col][i] == lit : s->cols[col][i] != lit);
    s->type[s->ntype++] = T_NAME;
  } else if (s1 == T_NUMBER && s2 == T_NAME) {
    int64_t n1 = POP(s->nums, s->nnums);
    int v = POP(s->value, s->nvalue);
    const int64_t *t = get_vec(s, v);
    uint32_t p2 = POP(s->prec, s->nprec);
    uint32_t p1 = get_decimals(s, v);
    uint32_t pres = p1 > p2 ? p1 : p2;
    s->prec[s->nprec++] = pres;
    s->type[s->ntype++] = T_NAME;
    if (t) cmp_col_const(s, t, n1, code, pres - p1, pres - p2, out);
  } else if (s1 == T_NAME && s2 == T_NUMBER) {
    static const int rev[7] = {0, 2, 1, 3, 4, 6, 5};
    code = (code >= 1 && code <= 6) ? rev[code] : code;
    int64_t n1 = POP(s->nums, s->nnums);
    int v = POP(s->value, s->nvalue);
    const int64_t *t = get_vec(s, v);
    uint32_t p2 = POP(s->prec, s->nprec);
    uint32_t p1 = get_decimals(s, v);
    uint32_t pres = p1 > p2 ? p1 : p2;
    s->prec[s->nprec++] = pres;
    s->type[s->ntype++] = T_NAME;
    if (t) cmp_col_const(s, t, n1, code, p1, p2, out);
  } else if (s1 == T_NAME && s2 == T_NAME) {
    int v1 = POP(s->value, s->nvalue), v2 = POP(s->value, s->nvalue);
    s->type[s->ntype++] = T_NAME;
    const int64_t *t = get_vec(s, v1), *t1 = get_vec(s, v2);
    uint32_t p1 = get_decimals(s, v1), p2 = get_decimals(s, v2);
    uint32_t pres = p1 > p2 ? p1 : p2;
    s->prec[s->nprec++] = pres;
    /* compare(t1, t, op, p2, p1), cm.cu:2299-2379: a scaled right operand is read past
     * its end in the reference (undefined), so it is rejected here. */
    if (p1) { s->err = ALN_E_UNSUPPORTED; return; }
    int64_t a;
    if (pw10(p2, &a)) { s->err = ALN_E_UNSUPPORTED; return; }
    if (t && t1)
      for (uint64_t i = 0; i < s->n; ++i)
        out[i] = (uint8_t)cmpv(code, p2 ? wmul(t1[i], a) : t1[i], t[i]);
  } else {
    s->err = ALN_E_UNSUPPORTED;
    return;
  }
  s->bools[s->nbools++] = out;
}

static int rpn_run(rpn_state *s, const aln_rpn *rpn, int nrpn, int mode) {
  s->mode = mode;
  for (int i = 0; i < nrpn && !s->err; ++i) {
    const aln_rpn *o = &rpn[i];
    switch (o->type) {
    case ALN_RPN_NAME:
      s->value[s->nvalue++] = o->arg;
      s->type[s->ntype++] = T_NAME;
      break;
    case ALN_RPN_NUMBER:
      s->nums[s->nnums++] = o->val;
      s->type[s->ntype++] = T_NUMBER;
      s->prec[s->nprec++] = (uint32_t)o->arg;
      break;
    case ALN_RPN_FLOAT:
      s->fnums[s->nfnums++] = dbits(o->val);
      s->type[s->ntype++] = T_FLOAT;
      break;
    case ALN_RPN_STRING:
      s->strs[s->nstrs] = o->val;
      s->value[s->nvalue++] = -2 - s->nstrs;
      s->nstrs++;
      s->type[s->ntype++] = T_STRING;
      break;
    case ALN_RPN_ADD: case ALN_RPN_MINUS: case ALN_RPN_MUL: case ALN_RPN_DIV:
      rpn_arith(s, o->type);
      break;
    case ALN_RPN_CMP:
      if (mode != 0) { s->err = ALN_E_UNSUPPORTED; break; }
      rpn_cmp(s, o->arg);
      break;
    case ALN_RPN_AND: case ALN_RPN_OR: {
      if (mode != 0 || s->nbools < 2) { s->err = ALN_E_ARG; break; }
      uint8_t *b3 = s->bools[--s->nbools], *b2 = s->bools[s->nbools - 1];
      s->type[s->ntype++] = T_NAME;
      for (uint64_t k = 0; k < s->n; ++k)
        b2[k] = (uint8_t)(o->type == ALN_RPN_AND ? (b2[k] && b3[k]) : (b2[k] || b3[k]));
      break;
    }
    default:
      s->err = ALN_E_ARG;
    }
    if (s->ntype >= SMAX - 1 || s->nvalue >= SMAX - 1 || s->nnums >= SMAX - 1 ||
        s->nstrs >= SMAX - 1 || s->nfnums >= SMAX - 1 ||
        s->nprec >= SMAX - 1 || s->nvecs >= SMAX - 1 || s->nbools >= SMAX - 1)
      s->err = ALN_E_ARG;
  }
  return s->err;
}

static void rpn_free(rpn_state *s) {
  for (int i = 0; i < s->nowned; ++i) free(s->owned[i]);
}

int orc_filter(const aln_rpn *rpn, int nrpn, const int64_t *const *cols, const int32_t *col_dec,
               uint64_t n, uint8_t *mask) {
  rpn_state *s = (rpn_state *)calloc(1, sizeof(rpn_state));
  s->cols = cols; s->dec = col_dec; s->n = n;
  int err = rpn_run(s, rpn, nrpn, 0);
  if (!err && s->nbools == 0) err = ALN_E_ARG;
  if (!err) memcpy(mask, s->bools[s->nbools - 1], n);
  rpn_free(s);
  free(s);
  return err;
}

/* select() expression value (select.cu:703-790, 830-900): a lone column, a constant, or
 * the last computed vector; *out_prec is its decimal scale. */
int orc_eval(const aln_rpn *rpn, int nrpn, const int64_t *const *cols, const int32_t *col_dec,
             uint64_t n, int64_t *out, int32_t *out_prec) {
  rpn_state *s = (rpn_state *)calloc(1, sizeof(rpn_state));
  s->cols = cols; s->dec = col_dec; s->n = n;
  int err = rpn_run(s, rpn, nrpn, 1);
  if (!err) {
    if (s->nfnums > 0 && s->ntype > 0 && s->type[s->ntype - 1] == T_FLOAT) {
      for (uint64_t i = 0; i < n; ++i) out[i] = bitsd(s->fnums[s->nfnums - 1]);
      *out_prec = ALN_DEC_F64;
    } else if (s->nnums > 0 && s->ntype > 0 && s->type[s->ntype - 1] == T_NUMBER) {
      for (uint64_t i = 0; i < n; ++i) out[i] = s->nums[s->nnums - 1];
      *out_prec = (int32_t)s->prec[s->nprec - 1];
    } else if (s->nvalue > 0 && s->value[s->nvalue - 1] >= 0) {
      int c = s->value[s->nvalue - 1];
      memcpy(out, cols[c], n * 8);
      *out_prec = col_dec[c];
    } else if (s->nvecs > 0) {
      memcpy(out, s->vecs[s->nvecs - 1], n * 8);
      *out_prec = s->vfloat[s->nvecs - 1] ? ALN_DEC_F64 : (int32_t)s->prec[s->nprec - 1];
    } else {
      err = ALN_E_ARG;
    }
  }
  rpn_free(s);
  free(s);
  return err;
}

/* ------------------------------------------------------------------------------------ */
/* zone_map_check (zone_map.cu:720-1135) for int/decimal columns: host_compare
 * (zone_map.cu:78-174), host_logical_and/or (48-74).  Min/max are scaled in place as the
 * reference's get_host_vec does.  A predicate with arithmetic yields 'R' here: the
 * reference's interval arithmetic (host_op, zone_map.cu:513-560) can misclassify. */
static char hc_const(int64_t s, int64_t d, int op) {
  char res = 'N';
  if (op == 2 && d > s) res = 'A';
  else if (op == 1 && d < s) res = 'A';
  else if (op == 6 && d >= s) res = 'A';
  else if (op == 5 && d <= s) res = 'A';
  else if (op == 4 && d == s) res = 'A';
  else if (d != s) res = 'A';
  return res;
}
static char hc_range(const int64_t *c, int64_t d, int op) {
  char res = 'R';
  if (op == 2) { if (c[1] <= d) res = 'N'; else if (c[0] > d) res = 'A'; }
  else if (op == 1) { if (c[0] >= d) res = 'N'; else if (c[1] < d) res = 'A'; }
  else if (op == 6) { if (c[1] < d) res = 'N'; else if (c[0] >= d) res = 'A'; }
  else if (op == 5) { if (c[0] > d) res = 'N'; else if (c[1] <= d) res = 'A'; }
  else if (op == 4 && c[0] == d && c[1] == d) res = 'A';
  return res;
}
static char hl_and(char a, char b) { return (a == 'A' && b == 'A') ? 'A' : (a == 'N' || b == 'N') ? 'N' : 'R'; }
static char hl_or(char a, char b) { return (a == 'A' && b == 'A') ? 'A' : (a == 'N' && b == 'N') ? 'N' : 'R'; }

int orc_zone_map_check(const aln_rpn *rpn, int nrpn, const int64_t *cmin, const int64_t *cmax,
                       const int32_t *col_dec, int not_compressed, char *out) {
  if (not_compressed) { *out = 'R'; return 0; }
  int maxc = 0;
  for (int i = 0; i < nrpn; ++i) {
    if (rpn[i].type == ALN_RPN_NAME && rpn[i].arg + 1 > maxc) maxc = rpn[i].arg + 1;
    if (rpn[i].type >= ALN_RPN_ADD && rpn[i].type <= ALN_RPN_DIV) { *out = 'R'; return 0; }
    /* double operands: every segment is evaluated (no float zone maps are kept) */
    if (rpn[i].type == ALN_RPN_FLOAT ||
        (rpn[i].type == ALN_RPN_NAME && col_dec && col_dec[rpn[i].arg] == ALN_DEC_F64)) { *out = 'R'; return 0; }
  }
  int64_t mm[ALN_MAX_COLS][2];
  if (maxc > ALN_MAX_COLS) return ALN_E_ARG;
  for (int c = 0; c < maxc; ++c) { mm[c][0] = cmin[c]; mm[c][1] = cmax[c]; }
  int type[SMAX], nt = 0, val[SMAX], nv = 0, np = 0, nb = 0, nn = 0;
  int64_t nums[SMAX];
  uint32_t prec[SMAX];
  char bools[SMAX];
  for (int i = 0; i < nrpn; ++i) {
    const aln_rpn *o = &rpn[i];
    if (nt >= SMAX - 1 || nv >= SMAX - 1 || np >= SMAX - 1 || nb >= SMAX - 1) return ALN_E_ARG;
    if (o->type == ALN_RPN_NAME) { type[nt++] = T_NAME; val[nv++] = o->arg; }
    else if (o->type == ALN_RPN_NUMBER) { type[nt++] = T_NUMBER; nums[nn++] = o->val; prec[np++] = (uint32_t)o->arg; }
    else if (o->type == ALN_RPN_STRING) { type[nt++] = T_STRING; val[nv++] = -1; }
    else if (o->type == ALN_RPN_CMP) {
      int code = o->arg;
      if (nt < 2) return ALN_E_ARG;
      int s1 = type[--nt], s2 = type[--nt];
      if (s1 == T_NUMBER && s2 == T_NUMBER) {
        if (nn < 2 || np < 2) return ALN_E_ARG;
        int64_t n1 = nums[--nn], n2 = nums[--nn];
        uint32_t p1 = prec[--np], p2 = prec[--np];
        uint32_t pres = p1 > p2 ? p1 : p2;
        int64_t a, b;
        if (pw10(pres - p1, &a) || pw10(pres - p2, &b)) return ALN_E_UNSUPPORTED;
        prec[np++] = pres;
        type[nt++] = T_NAME; val[nv++] = -1;
        if (p1) n1 = wmul(n1, a);
        if (p2) n2 = wmul(n2, b);
        bools[nb++] = hc_const(n1, n2, code);
      } else if ((s1 == T_STRING && s2 == T_NAME) || (s1 == T_NAME && s2 == T_STRING)) {
        if (nv < 2) return ALN_E_ARG;
        nv -= 2;
        bools[nb++] = 'R';
        type[nt++] = T_NAME; val[nv++] = -1;
      } else if ((s1 == T_NUMBER && s2 == T_NAME) || (s1 == T_NAME && s2 == T_NUMBER)) {
        if (s1 == T_NAME) {
          static const int rev[7] = {0, 2, 1, 3, 4, 6, 5};
          if (code >= 1 && code <= 6) code = rev[code];
        }
        if (nn < 1 || nv < 1 || np < 1) return ALN_E_ARG;
        int64_t n1 = nums[--nn];
        int c = val[--nv];
        uint32_t p2 = prec[--np];
        if (c < 0) { *out = 'R'; return 0; }
        uint32_t p1 = (uint32_t)col_dec[c];
        uint32_t pres = p1 > p2 ? p1 : p2;
        int64_t a, b;
        if (pw10(pres - p1, &a) || pw10(pres - p2, &b)) return ALN_E_UNSUPPORTED;
        prec[np++] = pres;
        if (p1) { mm[c][0] = wmul(mm[c][0], a); mm[c][1] = wmul(mm[c][1], a); }
        if (p2) n1 = wmul(n1, b);
        type[nt++] = T_NAME; val[nv++] = -1;
        bools[nb++] = hc_range(mm[c], n1, code);
      } else {
        *out = 'R';
        return 0;
      }
    } else if (o->type == ALN_RPN_AND || o->type == ALN_RPN_OR) {
      if (nb < 2) return ALN_E_ARG;
      char b3 = bools[--nb], b2 = bools[--nb];
      type[nt++] = T_NAME;
      bools[nb++] = o->type == ALN_RPN_AND ? hl_and(b2, b3) : hl_or(b2, b3);
    } else {
      return ALN_E_ARG;
    }
  }
  if (nb == 0) return ALN_E_ARG;
  *out = bools[nb - 1];
  return 0;
}

/* thrust::copy_if(counting_iterator) over the mask (cm.cu:3648-3650). */
uint64_t orc_compact(const uint8_t *mask, uint64_t n, uint32_t row0, uint32_t *ids) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (mask[i]) ids[k++] = row0 + (uint32_t)i;
  return k;
}

/* ------------------------------------------------------------------------------------ */
/* stable merge sort of row indices under a comparator */
typedef struct {
  const int64_t *const *keys;
  const int32_t *desc;
  int nk;
} cmp_ctx;

static int cmp_rows(const cmp_ctx *c, uint32_t a, uint32_t b) {
  for (int k = 0; k < c->nk; ++k) {
    int64_t x = c->keys[k][a], y = c->keys[k][b];
    if (x != y) {
      int lt = x < y;
      if (c->desc && c->desc[k]) lt = !lt;
      return lt ? -1 : 1;
    }
  }
  return 0;
}

static void msort(const cmp_ctx *c, uint32_t *a, uint32_t *tmp, uint64_t n) {
  if (n < 2) return;
  uint64_t h = n / 2;
  msort(c, a, tmp, h);
  msort(c, a + h, tmp, n - h);
  uint64_t i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = (cmp_rows(c, a[j], a[i]) < 0) ? a[j++] : a[i++];
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, n * 4);
}

/* Stable multi-key ORDER BY (emit_order operators.cu:1675-1782: one stable_sort_by_key
 * per key, last key first, sorts.cu:17-70), which equals one stable sort by the
 * composite key. */
void orc_order(const int64_t *const *keys, const int32_t *desc, int nk, uint64_t n,
               uint32_t *perm) {
  cmp_ctx c = {keys, desc, nk};
  uint32_t *tmp = (uint32_t *)malloc(4 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
  msort(&c, perm, tmp, n);
  free(tmp);
}

/* One-segment GROUP BY: order_inplace (operators.cu:367-406, stable ASC sort on the keys),
 * CudaSet::GroupBy (cm.cu:910-956, run heads) and segreduce SUM/COUNT/MIN/MAX
 * (select.cu:481-684).  Output in ascending key order; key values of each run head. */
uint64_t orc_groupby_sorted(const int64_t *const *keys, int nk, const int64_t *const *vals,
                            const int32_t *aggs, int nv, uint64_t n, int64_t *const *out_keys,
                            int64_t *const *out_vals) {
  uint32_t *perm = (uint32_t *)malloc(4 * (n ? n : 1));
  orc_order(keys, NULL, nk, n, perm);
  cmp_ctx c = {keys, NULL, nk};
  uint64_t g = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r = perm[i];
    int head = (i == 0) || cmp_rows(&c, perm[i - 1], r) != 0;
    if (head) {
      for (int k = 0; k < nk; ++k) out_keys[k][g] = keys[k][r];
      for (int v = 0; v < nv; ++v) {
        int a = aggs[v];
        out_vals[v][g] = a == ALN_AGG_COUNT ? 1 : vals[v][r];
      }
      ++g;
    } else {
      for (int v = 0; v < nv; ++v) {
        int a = aggs[v];
        int64_t *o = &out_vals[v][g - 1];
        if (a == ALN_AGG_COUNT) *o += 1;
        else if (a == ALN_AGG_SUM) *o = wadd(*o, vals[v][r]);
        else if (a == ALN_AGG_MIN) { if (vals[v][r] < *o) *o = vals[v][r]; }
        else if (a == ALN_AGG_MAX) { if (vals[v][r] > *o) *o = vals[v][r]; }
      }
    }
  }
  free(perm);
  return g;
}

typedef struct {
  uint64_t hash;
  uint64_t seq;
} hentry;
static int hcmp(const void *a, const void *b) {
  const hentry *x = (const hentry *)a, *y = (const hentry *)b;
  if (x->hash != y->hash) return x->hash < y->hash ? -1 : 1;
  return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* Multi-segment GROUP BY: per-segment groups, then add() (merge.cu:92-268) identifies
 * each group by MurmurHash64A over its keys as int64[nk] (8*nk bytes, seed hash_seed),
 * merges in ascending hash order, and count_avg() (merge.cu:314-496) reduces equal hashes
 * (SUM/COUNT add; key columns and MAX take max; MIN takes min). */
uint64_t orc_groupby_segmented(const int64_t *const *keys, int nk, const int64_t *const *vals,
                               const int32_t *aggs, int nv, uint64_t n, const uint64_t *seg_bounds,
                               int nsegs, uint32_t seed, int64_t *const *out_keys,
                               int64_t *const *out_vals) {
  if (nsegs <= 1) return orc_groupby_sorted(keys, nk, vals, aggs, nv, n, out_keys, out_vals);
  int64_t *pk[ALN_MAX_KEYS], *pv[ALN_MAX_AGGS];
  for (int k = 0; k < nk; ++k) pk[k] = (int64_t *)malloc(8 * (n ? n : 1));
  for (int v = 0; v < nv; ++v) pv[v] = (int64_t *)malloc(8 * (n ? n : 1));
  uint64_t np = 0;
  for (int s = 0; s < nsegs; ++s) {
    uint64_t b = seg_bounds[s], e = seg_bounds[s + 1];
    const int64_t *sk[ALN_MAX_KEYS], *sv[ALN_MAX_AGGS];
    int64_t *ok[ALN_MAX_KEYS], *ov[ALN_MAX_AGGS];
    for (int k = 0; k < nk; ++k) { sk[k] = keys[k] + b; ok[k] = pk[k] + np; }
    for (int v = 0; v < nv; ++v) { sv[v] = vals[v] + b; ov[v] = pv[v] + np; }
    np += orc_groupby_sorted(sk, nk, sv, aggs, nv, e - b, ok, ov);
  }
  hentry *h = (hentry *)malloc(sizeof(hentry) * (np ? np : 1));
  for (uint64_t i = 0; i < np; ++i) {
    int64_t tup[ALN_MAX_KEYS];
    for (int k = 0; k < nk; ++k) tup[k] = pk[k][i];
    h[i].hash = orc_murmur64a(tup, 8 * nk, seed);
    h[i].seq = i;
  }
  qsort(h, np, sizeof(hentry), hcmp);
  uint64_t g = 0;
  for (uint64_t i = 0; i < np; ++i) {
    uint64_t r = h[i].seq;
    if (i == 0 || h[i].hash != h[i - 1].hash) {
      for (int k = 0; k < nk; ++k) out_keys[k][g] = pk[k][r];
      for (int v = 0; v < nv; ++v) out_vals[v][g] = pv[v][r];
      ++g;
    } else {
      for (int k = 0; k < nk; ++k)
        if (pk[k][r] > out_keys[k][g - 1]) out_keys[k][g - 1] = pk[k][r];
      for (int v = 0; v < nv; ++v) {
        int64_t *o = &out_vals[v][g - 1];
        if (aggs[v] == ALN_AGG_MIN) { if (pv[v][r] < *o) *o = pv[v][r]; }
        else if (aggs[v] == ALN_AGG_MAX) { if (pv[v][r] > *o) *o = pv[v][r]; }
        else *o = wadd(*o, pv[v][r]);
      }
    }
  }
  free(h);
  for (int k = 0; k < nk; ++k) free(pk[k]);
  for (int v = 0; v < nv; ++v) free(pv[v]);
  return g;
}

/* Inner equi-join as emit_multijoin does it (operators.cu:1117-1284): the right side is
 * stably sorted by key; each left segment is sorted by key carrying its row ids
 * (operators.cu:1184-1199); the merge join (moderngpu inner_join semantics) emits, for
 * each left row in that order, every equal right row in right-sorted order.  Pairs are
 * (left row, right row), global indices, concatenated over left segments.  Returns the
 * pair count (only the first cap pairs are written). */
uint64_t orc_join_sortmerge(const int64_t *lkeys, uint64_t nl, const int64_t *rkeys, uint64_t nr,
                            const uint64_t *lseg_bounds, int nlsegs, uint32_t *out_l,
                            uint32_t *out_r, uint64_t cap) {
  uint32_t *rp = (uint32_t *)malloc(4 * (nr ? nr : 1));
  const int64_t *rk[1] = {rkeys};
  orc_order(rk, NULL, 1, nr, rp);
  uint64_t k = 0;
  uint64_t one[2] = {0, nl};
  if (nlsegs <= 0) { lseg_bounds = one; nlsegs = 1; }
  for (int s = 0; s < nlsegs; ++s) {
    uint64_t b = lseg_bounds[s], e = lseg_bounds[s + 1];
    uint32_t *lp = (uint32_t *)malloc(4 * ((e - b) ? (e - b) : 1));
    const int64_t *lk[1] = {lkeys + b};
    orc_order(lk, NULL, 1, e - b, lp);
    uint64_t j = 0;
    for (uint64_t i = 0; i < e - b; ++i) {
      int64_t key = lkeys[b + lp[i]];
      while (j < nr && rkeys[rp[j]] < key) ++j;
      for (uint64_t t = j; t < nr && rkeys[rp[t]] == key; ++t) {
        if (k < cap) { out_l[k] = (uint32_t)(b + lp[i]); out_r[k] = rp[t]; }
        ++k;
      }
    }
    free(lp);
  }
  free(rp);
  return k;
}

/* One string key of ORDER BY, restated from str_sort (strings_sort_device.cu:21-31, 125-137)
 * and Str::comparison (strings_type.h:46-115): perm (row ids, in their current order) is
 * stably sorted by the rows' width-byte values vals[row * width ..], compared as unsigned
 * bytes, descending when desc.  Widths the reference has no sort for (above 101, other than
 * 255) leave perm unchanged. */
static const char *g_str_vals;
static uint32_t g_str_width;
static int g_str_desc;
static int cmp_str_rows(const void *a, const void *b) {
  const uint32_t x = ((const uint32_t *)a)[0], y = ((const uint32_t *)b)[0];
  int r = memcmp(g_str_vals + (uint64_t)x * g_str_width, g_str_vals + (uint64_t)y * g_str_width, g_str_width);
  if (g_str_desc) r = -r;
  if (r) return r;
  /* the second word is the input position: ties keep their order (a stable sort) */
  const uint32_t px = ((const uint32_t *)a)[1], py = ((const uint32_t *)b)[1];
  return px < py ? -1 : px > py;
}

void orc_str_sort(const char *vals, uint32_t width, uint64_t n, int desc, uint32_t *perm) {
  if (!((width >= 1 && width <= 101) || width == 255)) return;
  uint32_t *pairs = (uint32_t *)malloc(8 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    pairs[2 * i] = perm[i];
    pairs[2 * i + 1] = (uint32_t)i;
  }
  g_str_vals = vals;
  g_str_width = width;
  g_str_desc = desc;
  qsort(pairs, n, 8, cmp_str_rows);
  for (uint64_t i = 0; i < n; ++i) perm[i] = pairs[2 * i];
  free(pairs);
}

/* Semi and anti joins, restated from operators.cu:1357-1399 and 1473-1530.  The pairs of
 * the sort-merge join on the first key, narrowed to those where the other nkeys-1 keys are
 * equal too (operators.cu:1288-1348); then
 *   '1' LEFT SEMI   sort + unique of the pairs' left rows;
 *   '3' LEFT ANTI   0..nl-1 minus those rows (set difference);
 *   '2' RIGHT SEMI  sort + unique of the pairs' right positions, a position being the
 *                   right row's rank in the right side's stable sort by its first key;
 *   '4' RIGHT ANTI  0..nr-1 minus those positions;
 * right positions are written as right rows.  Writes the rows to out (capacity nl or nr)
 * and returns their count. */
static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

uint64_t orc_semi_anti(const int64_t *const *lkeys, const int64_t *const *rkeys, int nkeys, uint64_t nl,
                       uint64_t nr, const uint64_t *lseg_bounds, int nlsegs, char kind, uint32_t *out) {
  uint64_t np = orc_join_sortmerge(lkeys[0], nl, rkeys[0], nr, lseg_bounds, nlsegs, NULL, NULL, 0);
  uint32_t *pl = (uint32_t *)malloc(4 * (np ? np : 1)), *pr = (uint32_t *)malloc(4 * (np ? np : 1));
  orc_join_sortmerge(lkeys[0], nl, rkeys[0], nr, lseg_bounds, nlsegs, pl, pr, np);
  uint64_t k = 0;
  for (uint64_t i = 0; i < np; ++i) {
    int eq = 1;
    for (int c = 1; c < nkeys; ++c) eq = eq && lkeys[c][pl[i]] == rkeys[c][pr[i]];
    if (eq) { pl[k] = pl[i]; pr[k] = pr[i]; ++k; }
  }
  const int left = kind == '1' || kind == '3';
  const uint64_t n = left ? nl : nr;
  uint32_t *rp = NULL, *v = left ? pl : pr;
  if (!left) {  /* right rows -> positions in the stable key order */
    rp = (uint32_t *)malloc(4 * (nr ? nr : 1));
    uint32_t *pos = (uint32_t *)malloc(4 * (nr ? nr : 1));
    const int64_t *rk[1] = {rkeys[0]};
    orc_order(rk, NULL, 1, nr, rp);
    for (uint64_t i = 0; i < nr; ++i) pos[rp[i]] = (uint32_t)i;
    for (uint64_t i = 0; i < k; ++i) v[i] = pos[v[i]];
    free(pos);
  }
  qsort(v, k, 4, cmp_u32);
  uint64_t u = 0;
  for (uint64_t i = 0; i < k; ++i)
    if (u == 0 || v[u - 1] != v[i]) v[u++] = v[i];
  uint64_t m = 0;
  if (kind == '1' || kind == '2') {
    for (uint64_t i = 0; i < u; ++i) out[m++] = v[i];
  } else {
    uint64_t j = 0;
    for (uint64_t i = 0; i < n; ++i) {
      while (j < u && v[j] < i) ++j;
      if (j < u && v[j] == i) continue;
      out[m++] = (uint32_t)i;
    }
  }
  if (!left)
    for (uint64_t i = 0; i < m; ++i) out[i] = rp[out[i]];
  free(rp);
  free(pl);
  free(pr);
  return m;
}

/* Kahan-compensated FP64 sum (the FP64 oracle for SUM aggregates). */
uint64_t orc_kahan_sum_f64(const double *v, uint64_t n, double *out) {
  double s = 0.0, c = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    double y = v[i] - c, t = s + y;
    c = (t - s) - y;
    s = t;
  }
  *out = s;
  return n;
}

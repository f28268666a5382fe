// GROUP BY with SUM / COUNT / MIN / MAX, and row hashing.
//
// The reference groups by sorting (order_inplace, operators.cu:367-406), finding run heads
// (CudaSet::GroupBy, cm.cu:910-956) and reducing each run (moderngpu segreduce,
// select.cu:481-684); across segments it merges partial groups by MurmurHash64A of the key
// tuple on the host (add / count_avg, merge.cu:92-496).  Integer sums are exact, so the
// result does not depend on the order rows are added in, and this kernel aggregates in an
// table of accumulators instead:
//   * keys are packed into one 63-bit code (key_i - min_i in bits_i bits, key 0 most
//     significant); codes of at most GB_DENSE_BITS bits index the accumulators directly,
//     wider ones go through an open-addressing table (64-bit CAS, linear probing), first
//     one sized for 2^20 groups and, if more turn up, one of 2n slots;
//   * runs of equal codes in adjacent lanes (clustered input) are pre-reduced with a
//     segmented wave scan, and each block adds its rows into an LDS table first, so hot
//     keys and small group counts aggregate in LDS and reach the global table once per
//     block;
//   * no GROUP BY (or keys that never vary) is one group: a block reduction;
//   * the groups are compacted and put in the reference's output order: ascending key
//     tuple (one segment) or ascending MurmurHash64A(keys, 8*nkeys, seed) (the
//     multi-segment merge order).
// Key tuples wider than 63 bits fall back to the reference's own scheme: stable sort,
// run heads, per-run reduction.
#include <string.h>

#include <algorithm>
#include <vector>

#include "aln_internal.h"

namespace aln {

int sort_keys_impl(aln_ctx *c, const int64_t *const *keys, const int32_t *desc, int nkeys,
                   uint64_t n, uint32_t *d_perm);

constexpr uint64_t GB_EMPTY = ~0ull;
// Packed key codes of at most this many bits (and at most 2n codes) use direct-indexed
// accumulators instead of a hash table.
constexpr uint32_t GB_DENSE_BITS = 26;
// The first hash table holds this many groups (at most half full) before the rows are
// aggregated again in a table of 2n slots.
constexpr uint64_t GB_HASH_START = 1ull << 20;
constexpr uint64_t GB_FIRST_PROBES = 64;

struct GbArgs {
  const int64_t *keys[ALN_MAX_KEYS];
  int64_t kmin[ALN_MAX_KEYS];
  uint32_t shift[ALN_MAX_KEYS];
  uint64_t fmask[ALN_MAX_KEYS];
  const int64_t *vals[ALN_MAX_AGGS];
  int32_t aggs[ALN_MAX_AGGS];
  int64_t *tab[ALN_MAX_AGGS];
  int nk, nv;
};

__device__ __forceinline__ int64_t agg_combine(int a, int64_t x, int64_t y) {
  if (a == ALN_AGG_MIN) return x < y ? x : y;
  if (a == ALN_AGG_MAX) return x > y ? x : y;
  if (a == ALN_AGG_FSUM) return f64_bits(bits_f64(x) + bits_f64(y));
  if (a == ALN_AGG_FMIN) return bits_f64(y) < bits_f64(x) ? y : x;
  if (a == ALN_AGG_FMAX) return bits_f64(y) > bits_f64(x) ? y : x;
  return (int64_t)((uint64_t)x + (uint64_t)y);  // SUM, COUNT
}

// The starting value of an aggregate's accumulator.
__host__ __device__ __forceinline__ int64_t agg_identity(int a) {
  switch (a) {
    case ALN_AGG_MIN: return INT64_MAX;
    case ALN_AGG_MAX: return INT64_MIN;
    case ALN_AGG_FMIN: return 0x7FF0000000000000ll;                // +inf
    case ALN_AGG_FMAX: return (int64_t)0xFFF0000000000000ull;      // -inf
    default: return 0;                                             // 0, 0.0
  }
}

// Adds x into the accumulator at t.
__device__ __forceinline__ void agg_atomic(int a, int64_t *t, int64_t x) {
  if (a == ALN_AGG_MIN) {
    atomicMin((long long *)t, (long long)x);
  } else if (a == ALN_AGG_MAX) {
    atomicMax((long long *)t, (long long)x);
  } else if (a == ALN_AGG_FSUM) {
    atomicAdd(reinterpret_cast<double *>(t), bits_f64(x));
  } else if (a == ALN_AGG_FMIN || a == ALN_AGG_FMAX) {
    unsigned long long *p = reinterpret_cast<unsigned long long *>(t);
    unsigned long long old = *p, assumed;
    do {
      assumed = old;
      const int64_t nv = agg_combine(a, (int64_t)assumed, x);
      if (nv == (int64_t)assumed) break;
      old = atomicCAS(p, assumed, (unsigned long long)nv);
    } while (old != assumed);
  } else {
    atomicAdd((unsigned long long *)t, (unsigned long long)x);
  }
}

__device__ __forceinline__ uint64_t murmur64a_tuple(const int64_t *v, int nk, uint32_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  uint64_t h = (uint64_t)seed ^ ((uint64_t)(8 * nk) * m);
  for (int i = 0; i < nk; ++i) {
    uint64_t k = (uint64_t)v[i];
    k *= m;
    k ^= k >> 47;
    k *= m;
    h ^= k;
    h *= m;
  }
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}

__global__ void __launch_bounds__(256) k_gb_init(uint64_t *tkey, GbArgs a, uint64_t cap) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * 256) {
    tkey[i] = GB_EMPTY;
    for (int v = 0; v < a.nv; ++v) a.tab[v][i] = agg_identity(a.aggs[v]);
  }
}

struct GbOut {
  int64_t *keys[ALN_MAX_KEYS];
  int64_t *vals[ALN_MAX_AGGS];
};

// Adjacent lanes with equal codes (clustered input): the last lane of each run gets the
// run's combined values and row count and returns true.
__device__ __forceinline__ bool gb_wave_runs(const GbArgs &a, bool valid, uint64_t code, int64_t *val,
                                             uint32_t *cnt) {
  const unsigned lane = __lane_id();
  const uint64_t prev = __shfl_up(code, 1, 64);
  const bool pvalid = __shfl_up((int)valid, 1, 64) != 0;
  bool f = lane == 0 || !pvalid || prev != code;
  uint32_t c = valid ? 1u : 0u;
  for (int o = 1; o < 64; o <<= 1) {
    const bool fu = __shfl_up((int)f, o, 64) != 0;
    const uint32_t cu = __shfl_up(c, o, 64);
    for (int v = 0; v < a.nv; ++v) {
      const int64_t vu = __shfl_up(val[v], o, 64);
      if (lane >= (unsigned)o && !f) val[v] = agg_combine(a.aggs[v], vu, val[v]);
    }
    if (lane >= (unsigned)o && !f) c += cu;
    if (lane >= (unsigned)o) f = f || fu;
  }
  const uint64_t nextc = __shfl_down(code, 1, 64);
  const bool nvalid = __shfl_down((int)valid, 1, 64) != 0;
  *cnt = c;
  return valid && (lane == 63 || !nvalid || nextc != code);
}

__device__ __forceinline__ uint64_t gb_code(const GbArgs &a, uint64_t i) {
  uint64_t code = 0;
  for (int k = 0; k < a.nk; ++k)
    code |= (((uint64_t)a.keys[k][i] - (uint64_t)a.kmin[k]) & a.fmask[k]) << a.shift[k];
  return code;
}

// The global table: DENSE, slot = code (row count in gcnt, when there is no COUNT
// aggregate to tell the groups); otherwise open addressing over tkey
// (cap a power of two) by linear probing.  A new key counts into *ngroups; past glimit
// keys, a probe run past GB_FIRST_PROBES in such a limited table, or a full table,
// *overflow is set (the host retries with a larger table).
template <bool DENSE>
__device__ __forceinline__ void gb_global_add(const GbArgs &a, uint64_t code, const int64_t *val, uint32_t cnt,
                                              uint64_t *tkey, uint64_t cap, uint32_t *gcnt, uint32_t *ngroups,
                                              uint32_t glimit, uint32_t *overflow) {
  uint64_t h;
  if (DENSE) {
    h = code;
    if (gcnt) atomicAdd(&gcnt[h], cnt);
  } else {
    // a first, small table that has overflowed is given up: its rows are aggregated again
    const bool first = glimit != 0xFFFFFFFFu;
    if (first && __hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint64_t mask = cap - 1, limit = first ? GB_FIRST_PROBES : cap;
    h = murmur64a_u64(code, 0x9747b28cu) & mask;
    uint64_t probes = 0;
    while (true) {
      uint64_t cur = tkey[h];
      if (cur == GB_EMPTY) {
        const unsigned long long old = atomicCAS((unsigned long long *)&tkey[h], (unsigned long long)GB_EMPTY,
                                                 (unsigned long long)code);
        if (old == GB_EMPTY) {
          if (atomicAdd(ngroups, 1u) >= glimit) atomicExch(overflow, 1u);
          cur = code;
        } else {
          cur = old;
        }
      }
      if (cur == code) break;
      h = (h + 1) & mask;
      if (++probes > limit) {
        atomicExch(overflow, 1u);
        return;
      }
    }
  }
  for (int v = 0; v < a.nv; ++v) agg_atomic(a.aggs[v], &a.tab[v][h], val[v]);
}

// Aggregation with a per-block LDS table in front of the global one: the block's rows
// (grid-stride) add into GB_LT LDS slots (a hash of the code, GB_LT_PROBES probes; DENSE with
// at most GB_LT codes: slot = code), so hot keys and small group counts aggregate in LDS;
// rows whose key finds no slot go to the global table at once, and the LDS slots are added
// to it at the end.  LDS (dynamic): keys[GB_LT] u64, counts[GB_LT] u32, accumulators
// [nv][GB_LT] i64.
constexpr int GB_LT = 1024;
constexpr int GB_LT_PROBES = 8;

__host__ __device__ constexpr size_t gb_lds_bytes(int nv) { return (size_t)GB_LT * (8 + 4 + 8 * (size_t)nv); }

template <bool DENSE>
__global__ void __launch_bounds__(256) k_gb_agg(GbArgs a, uint64_t n, uint64_t *tkey, uint64_t cap,
                                                uint32_t *gcnt, uint32_t *ngroups, uint32_t glimit,
                                                uint32_t *overflow) {
  extern __shared__ __align__(16) int64_t gb_lds[];
  uint64_t *lkey = reinterpret_cast<uint64_t *>(gb_lds);
  int64_t *lacc = gb_lds + GB_LT;
  uint32_t *lcnt = reinterpret_cast<uint32_t *>(gb_lds + GB_LT + (size_t)a.nv * GB_LT);
  const bool direct = DENSE && cap <= (uint64_t)GB_LT;
  for (int s = threadIdx.x; s < GB_LT; s += 256) {
    lkey[s] = direct ? (uint64_t)s : GB_EMPTY;
    lcnt[s] = 0;
    for (int v = 0; v < a.nv; ++v) lacc[(size_t)v * GB_LT + s] = agg_identity(a.aggs[v]);
  }
  __syncthreads();
  for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < n; i0 += (uint64_t)gridDim.x * 256) {
    const uint64_t i = i0 + threadIdx.x;
    const bool valid = i < n;
    const uint64_t code = valid ? gb_code(a, i) : 0;
    int64_t val[ALN_MAX_AGGS];
    for (int v = 0; v < a.nv; ++v) val[v] = !valid ? 0 : a.aggs[v] == ALN_AGG_COUNT ? 1 : a.vals[v][i];
    uint32_t cnt;
    if (!gb_wave_runs(a, valid, code, val, &cnt)) continue;
    int slot = -1;
    if (direct) {
      slot = (int)code;
    } else {
      uint32_t h = (uint32_t)murmur64a_u64(code, 0x85ebca6bu) & (GB_LT - 1);
      for (int p = 0; p < GB_LT_PROBES; ++p) {
        uint64_t k = lkey[h];
        if (k == GB_EMPTY) {
          const unsigned long long old =
              atomicCAS((unsigned long long *)&lkey[h], (unsigned long long)GB_EMPTY, (unsigned long long)code);
          k = old == GB_EMPTY ? code : old;
        }
        if (k == code) {
          slot = (int)h;
          break;
        }
        h = (h + 1) & (GB_LT - 1);
      }
    }
    if (slot >= 0) {
      atomicAdd(&lcnt[slot], cnt);
      for (int v = 0; v < a.nv; ++v) agg_atomic(a.aggs[v], &lacc[(size_t)v * GB_LT + slot], val[v]);
    } else {
      gb_global_add<DENSE>(a, code, val, cnt, tkey, cap, gcnt, ngroups, glimit, overflow);
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < GB_LT; s += 256) {
    const uint32_t c = lcnt[s];
    if (c == 0) continue;
    int64_t val[ALN_MAX_AGGS];
    for (int v = 0; v < a.nv; ++v) val[v] = lacc[(size_t)v * GB_LT + s];
    gb_global_add<DENSE>(a, lkey[s], val, c, tkey, cap, gcnt, ngroups, glimit, overflow);
  }
}

// No key that varies: one group, a reduction.  Block partials part[block][v], each block's
// rows grid-strided, combined in a fixed order so the result does not depend on timing.
__global__ void __launch_bounds__(256) k_gb_reduce(GbArgs a, uint64_t n, int64_t *__restrict__ part) {
  __shared__ int64_t red[4][ALN_MAX_AGGS];
  int64_t acc[ALN_MAX_AGGS];
  for (int v = 0; v < a.nv; ++v) acc[v] = agg_identity(a.aggs[v]);
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    for (int v = 0; v < a.nv; ++v)
      acc[v] = agg_combine(a.aggs[v], acc[v], a.aggs[v] == ALN_AGG_COUNT ? 1 : a.vals[v][i]);
  for (int v = 0; v < a.nv; ++v) {
    for (int o = 32; o > 0; o >>= 1)
      acc[v] = agg_combine(a.aggs[v], acc[v], (int64_t)__shfl_xor((long long)acc[v], o, 64));
    if (__lane_id() == 0) red[threadIdx.x / 64][v] = acc[v];
  }
  __syncthreads();
  if ((int)threadIdx.x < a.nv) {
    const int v = threadIdx.x;
    int64_t r = red[0][v];
    for (int w = 1; w < 4; ++w) r = agg_combine(a.aggs[v], r, red[w][v]);
    part[(uint64_t)blockIdx.x * a.nv + v] = r;
  }
}

__global__ void k_gb_reduce_final(GbArgs a, const int64_t *__restrict__ part, uint32_t nparts, GbOut o) {
  const int t = threadIdx.x;
  if (t < a.nv) {
    int64_t r = part[t];
    for (uint32_t b = 1; b < nparts; ++b) r = agg_combine(a.aggs[t], r, part[(uint64_t)b * a.nv + t]);
    o.vals[t][0] = r;
  }
  if (t < a.nk) o.keys[t][0] = a.kmin[t];
}

__global__ void k_gb_dense_init(GbArgs a, uint64_t ncodes, uint32_t *gcnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < ncodes; i += (uint64_t)gridDim.x * 256) {
    if (gcnt) gcnt[i] = 0;
    for (int v = 0; v < a.nv; ++v) a.tab[v][i] = agg_identity(a.aggs[v]);
  }
}

// Occupied codes: a non-zero row count, from gcnt or (gcnt NULL) from COUNT aggregate cv.
__global__ void k_gb_dense_flags(GbArgs a, int cv, const uint32_t *gcnt, uint64_t ncodes, uint32_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < ncodes; i += (uint64_t)gridDim.x * 256)
    flag[i] = gcnt ? gcnt[i] != 0 : a.tab[cv][i] != 0;
}

__global__ void k_gb_dense_compact(const uint32_t *flag, uint64_t ncodes, const uint64_t *pos, uint64_t *gcode,
                                   uint32_t *gslot) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < ncodes; i += (uint64_t)gridDim.x * 256)
    if (flag[i]) {
      gcode[pos[i]] = i;
      gslot[pos[i]] = (uint32_t)i;
    }
}

__global__ void k_gb_flags(const uint64_t *tkey, uint64_t cap, uint32_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * 256)
    flag[i] = tkey[i] != GB_EMPTY;
}

__global__ void k_gb_compact(const uint64_t *tkey, uint64_t cap, const uint32_t *flag,
                             const uint64_t *pos, uint64_t *gcode, uint32_t *gslot) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * 256)
    if (flag[i]) {
      gcode[pos[i]] = tkey[i];
      gslot[pos[i]] = (uint32_t)i;
    }
}

// Decode group keys from codes, compute the merge-order hash of each group.
__global__ void k_gb_hash(GbArgs a, const uint64_t *gcode, uint64_t g, uint32_t seed,
                          uint64_t *hash) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    int64_t kv[ALN_MAX_KEYS];
    for (int k = 0; k < a.nk; ++k)
      kv[k] = (int64_t)((uint64_t)a.kmin[k] + ((gcode[i] >> a.shift[k]) & a.fmask[k]));
    hash[i] = murmur64a_tuple(kv, a.nk, seed);
  }
}

__global__ void k_gb_emit(GbArgs a, GbOut o, const uint64_t *gcode, const uint32_t *gslot,
                          uint64_t g) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    uint32_t s = gslot[i];
    for (int k = 0; k < a.nk; ++k)
      o.keys[k][i] = (int64_t)((uint64_t)a.kmin[k] + ((gcode[i] >> a.shift[k]) & a.fmask[k]));
    for (int v = 0; v < a.nv; ++v) o.vals[v][i] = a.tab[v][s];
  }
}

// ---- sort-based fallback ------------------------------------------------------------------
__global__ void k_gb_heads(GbArgs a, const uint32_t *perm, uint64_t n, uint32_t *head) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    bool h = i == 0;
    if (!h)
      for (int k = 0; k < a.nk; ++k)
        if (a.keys[k][perm[i]] != a.keys[k][perm[i - 1]]) h = true;
    head[i] = h;
  }
}

__global__ void k_gb_sorted_agg(GbArgs a, GbOut o, const uint32_t *perm, const uint32_t *head,
                                const uint64_t *gid_ex, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    uint64_t g = gid_ex[i] + head[i] - 1;  // group index of sorted row i
    uint32_t r = perm[i];
    if (head[i])
      for (int k = 0; k < a.nk; ++k) o.keys[k][g] = a.keys[k][r];
    for (int v = 0; v < a.nv; ++v) {
      const int ag = a.aggs[v];
      agg_atomic(ag, &o.vals[v][g], ag == ALN_AGG_COUNT ? 1 : a.vals[v][r]);
    }
  }
}

// Reorders the sort-path groups into hash order (in place through temporaries).
__global__ void k_gb_hash_keys(GbOut o, int nk, uint64_t g, uint32_t seed, uint64_t *hash,
                               uint32_t *idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    int64_t kv[ALN_MAX_KEYS];
    for (int k = 0; k < nk; ++k) kv[k] = o.keys[k][i];
    hash[i] = murmur64a_tuple(kv, nk, seed);
    idx[i] = (uint32_t)i;
  }
}

__global__ void k_gb_gather(const uint64_t *gc, const uint32_t *gs, const uint32_t *gi, uint64_t g,
                            uint64_t *oc, uint32_t *os) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < g;
       i += (uint64_t)gridDim.x * 256) {
    oc[i] = gc[gi[i]];
    os[i] = gs[gi[i]];
  }
}

__global__ void k_permute_i64(const int64_t *in, const uint32_t *perm, uint64_t n, int64_t *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    out[i] = in[perm[i]];
}

__global__ void k_hash_rows(GbArgs a, uint64_t n, uint32_t seed, uint64_t *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    int64_t kv[ALN_MAX_KEYS];
    for (int k = 0; k < a.nk; ++k) kv[k] = a.keys[k][i];
    out[i] = murmur64a_tuple(kv, a.nk, seed);
  }
}

__global__ void k_f64_keys(const int64_t *in, uint64_t n, int back, int64_t *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    out[i] = back ? f64_unorder_key(in[i]) : f64_order_key(in[i]);
}

static int groupby_sorted(aln_ctx *c, GbArgs &a, uint64_t n, int order, uint32_t seed, GbOut &o,
                          uint64_t cap, uint64_t *h_ngroups) {
  uint32_t *perm, *head;
  uint64_t *gid;
  ALN_TRY(ws(c, n, &perm));
  ALN_TRY(ws(c, n, &head));
  ALN_TRY(ws(c, n, &gid));
  const int64_t *keys[ALN_MAX_KEYS];
  for (int k = 0; k < a.nk; ++k) keys[k] = a.keys[k];
  ALN_TRY(sort_keys_impl(c, keys, nullptr, a.nk, n, perm));
  hipLaunchKernelGGL(k_gb_heads, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, a, perm, n,
                     head);
  ALN_TRY(check_launch("gb_heads"));
  uint64_t g = 0;
  ALN_TRY(exclusive_scan_u32(c, head, n, gid, &g));
  *h_ngroups = g;
  if (g > cap)
    return set_error(ALN_E_CAPACITY, "aln_groupby: %llu groups, capacity %llu",
                     (unsigned long long)g, (unsigned long long)cap);
  for (int v = 0; v < a.nv; ++v) ALN_TRY(fill_i64(c, o.vals[v], g, agg_identity(a.aggs[v])));
  {
    ProfScope p(c, "groupby_sorted_agg");
    hipLaunchKernelGGL(k_gb_sorted_agg, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, a,
                       o, perm, head, gid, n);
  }
  ALN_TRY(check_launch("gb_sorted_agg"));
  if (order == ALN_ORDER_HASH && g > 1) {
    uint64_t *hash;
    uint32_t *idx;
    int64_t *tmp;
    ALN_TRY(ws(c, g, &hash));
    ALN_TRY(ws(c, g, &idx));
    ALN_TRY(ws(c, g, &tmp));
    hipLaunchKernelGGL(k_gb_hash_keys, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, o,
                       a.nk, g, seed, hash, idx);
    ALN_TRY(check_launch("gb_hash_keys"));
    ALN_TRY(radix_sort_pairs(c, hash, idx, g, 64));
    auto permute = [&](int64_t *col) -> int {
      hipLaunchKernelGGL(k_permute_i64, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, col,
                         idx, g, tmp);
      ALN_TRY(check_launch("permute"));
      ALN_HIP(hipMemcpyAsync(col, tmp, 8 * g, hipMemcpyDeviceToDevice, c->stream));
      return ALN_OK;
    };
    for (int k = 0; k < a.nk; ++k) ALN_TRY(permute(o.keys[k]));
    for (int v = 0; v < a.nv; ++v) ALN_TRY(permute(o.vals[v]));
  }
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

}  // namespace aln

using namespace aln;

extern "C" {

int aln_groupby(aln_ctx *c, const int64_t *const *d_keys, int nkeys, const int64_t *const *d_vals,
                const int32_t *aggs, int nvals, uint64_t n, int order, uint32_t seed,
                int64_t *const *d_out_keys, int64_t *const *d_out_vals, uint64_t cap,
                uint64_t *h_ngroups) {
  ALN_REQUIRE(c && h_ngroups && nkeys >= 0 && nkeys <= ALN_MAX_KEYS && (nkeys == 0 || (d_keys && d_out_keys)),
              "aln_groupby: bad key arguments");
  ALN_REQUIRE(nvals >= 0 && nvals <= ALN_MAX_AGGS && (nvals == 0 || (aggs && d_vals && d_out_vals)),
              "aln_groupby: bad aggregate arguments");
  ALN_REQUIRE(order >= ALN_ORDER_KEYS && order <= ALN_ORDER_NONE, "aln_groupby: order %d", order);
  ALN_REQUIRE(n < (1ull << 32), "aln_groupby: more than 2^32 rows");
  ws_reset(c);
  *h_ngroups = 0;
  if (n == 0) return ALN_OK;
  GbArgs a;
  memset(&a, 0, sizeof(a));
  a.nk = nkeys;
  a.nv = nvals;
  uint32_t bits[ALN_MAX_KEYS], total = 0;
  for (int k = 0; k < nkeys; ++k) {
    ALN_REQUIRE(d_keys[k], "aln_groupby: key %d is NULL", k);
    int64_t mn, mx;
    ALN_TRY(minmax_i64(c, d_keys[k], n, &mn, &mx));
    bits[k] = bits_for_range_u((uint64_t)mx - (uint64_t)mn);
    a.keys[k] = d_keys[k];
    a.kmin[k] = mn;
    a.fmask[k] = bits[k] == 64 ? ~0ull : ((1ull << bits[k]) - 1);
    total += bits[k];
  }
  for (int v = 0; v < nvals; ++v) {
    ALN_REQUIRE(aggs[v] == ALN_AGG_COUNT || aggs[v] == ALN_AGG_SUM || aggs[v] == ALN_AGG_MIN ||
                    aggs[v] == ALN_AGG_MAX || aggs[v] == ALN_AGG_FSUM || aggs[v] == ALN_AGG_FMIN ||
                    aggs[v] == ALN_AGG_FMAX,
                "aln_groupby: aggregate %d", aggs[v]);
    ALN_REQUIRE(aggs[v] == ALN_AGG_COUNT || d_vals[v], "aln_groupby: value %d is NULL", v);
    a.vals[v] = d_vals[v];
    a.aggs[v] = aggs[v];
  }
  GbOut o;
  memset(&o, 0, sizeof(o));
  for (int k = 0; k < nkeys; ++k) o.keys[k] = d_out_keys[k];
  for (int v = 0; v < nvals; ++v) o.vals[v] = d_out_vals[v];

  if (total == 0) {
    // no GROUP BY, or keys that never vary: one group, a reduction (reduce, select.cu:549)
    *h_ngroups = 1;
    if (cap < 1) return set_error(ALN_E_CAPACITY, "aln_groupby: 1 group, capacity 0");
    const uint32_t nb = grid_for(n, 256 * 16, 1024);
    int64_t *part;
    ALN_TRY(ws(c, (size_t)nb * (nvals ? nvals : 1), &part));
    if (nvals) {
      ProfScope p(c, "groupby_reduce");
      hipLaunchKernelGGL(k_gb_reduce, dim3(nb), dim3(256), 0, c->stream, a, n, part);
    }
    ALN_TRY(check_launch("gb_reduce"));
    hipLaunchKernelGGL(k_gb_reduce_final, dim3(1), dim3(64), 0, c->stream, a, part, nb, o);
    ALN_TRY(check_launch("gb_reduce_final"));
    ALN_HIP(hipStreamSynchronize(c->stream));
    return ALN_OK;
  }
  if (total > 63) {
    for (int k = 0; k < nkeys; ++k) a.shift[k] = 0;
    return groupby_sorted(c, a, n, order, seed, o, cap, h_ngroups);
  }
  uint32_t sh = total;
  for (int k = 0; k < nkeys; ++k) {
    sh -= bits[k];
    a.shift[k] = sh;
  }
  // LDS-cached aggregation kernel: blocks of at least 32 rows per thread, its LDS size
  const size_t lds = gb_lds_bytes(nvals);
  const uint32_t nb = grid_for(n, 256 * 32, 2048);
  if (lds > 65536) {
    ALN_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_gb_agg<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ALN_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_gb_agg<false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  uint32_t *ctl;  // [0] overflow, [1] keys inserted
  ALN_TRY(ws(c, 2, &ctl));
  uint64_t g = 0, *gcode;
  uint32_t *gslot;
  const uint64_t ncodes = 1ull << total;
  const bool dense = total <= GB_DENSE_BITS && ncodes <= std::max<uint64_t>(2 * n, (uint64_t)GB_LT);
  if (dense) {
    // direct-indexed accumulators: slot = code, groups come out in key order
    uint32_t *gcnt = nullptr, *flag;
    uint64_t *pos;
    int cv = -1;  // a COUNT aggregate, whose table tells the occupied codes
    for (int v = 0; v < nvals && cv < 0; ++v)
      if (aggs[v] == ALN_AGG_COUNT) cv = v;
    if (cv < 0) ALN_TRY(ws(c, ncodes, &gcnt));
    for (int v = 0; v < nvals; ++v) ALN_TRY(ws(c, ncodes, &a.tab[v]));
    hipLaunchKernelGGL(k_gb_dense_init, dim3(grid_for(ncodes, 256, 65536)), dim3(256), 0, c->stream, a, ncodes,
                       gcnt);
    ALN_TRY(check_launch("gb_dense_init"));
    {
      ProfScope p(c, "groupby_insert");
      hipLaunchKernelGGL(k_gb_agg<true>, dim3(nb), dim3(256), lds, c->stream, a, n, (uint64_t *)nullptr, ncodes,
                         gcnt, ctl + 1, 0xFFFFFFFFu, ctl);
    }
    ALN_TRY(check_launch("gb_agg"));
    ALN_TRY(ws(c, ncodes, &flag));
    ALN_TRY(ws(c, ncodes, &pos));
    hipLaunchKernelGGL(k_gb_dense_flags, dim3(grid_for(ncodes, 256, 65536)), dim3(256), 0, c->stream, a, cv, gcnt,
                       ncodes, flag);
    ALN_TRY(check_launch("gb_flags"));
    ALN_TRY(exclusive_scan_u32(c, flag, ncodes, pos, &g));
    *h_ngroups = g;
    if (g > cap)
      return set_error(ALN_E_CAPACITY, "aln_groupby: %llu groups, capacity %llu", (unsigned long long)g,
                       (unsigned long long)cap);
    ALN_TRY(ws(c, g, &gcode));
    ALN_TRY(ws(c, g, &gslot));
    hipLaunchKernelGGL(k_gb_dense_compact, dim3(grid_for(ncodes, 256, 65536)), dim3(256), 0, c->stream, flag,
                       ncodes, pos, gcode, gslot);
    ALN_TRY(check_launch("gb_compact"));
  } else {
    // open addressing: first a table for up to GB_HASH_START keys (half full at most); more
    // keys than that set the overflow flag and the rows are aggregated again in a table of
    // 2n slots
    const uint64_t full = next_pow2(2 * n < 1024 ? 1024 : 2 * n);
    uint64_t tcap = std::min<uint64_t>(full, next_pow2(2 * std::min<uint64_t>(n, GB_HASH_START) < 1024
                                                           ? 1024
                                                           : 2 * std::min<uint64_t>(n, GB_HASH_START)));
    uint64_t *tkey;
    uint32_t *flag;
    uint64_t *pos;
    while (true) {
      const uint32_t glimit = tcap < full ? (uint32_t)(tcap / 2) : 0xFFFFFFFFu;
      ALN_TRY(ws(c, tcap, &tkey));
      for (int v = 0; v < nvals; ++v) ALN_TRY(ws(c, tcap, &a.tab[v]));
      ALN_TRY(fill_u32(c, ctl, 2, 0));
      hipLaunchKernelGGL(k_gb_init, dim3(grid_for(tcap, 256, 65536)), dim3(256), 0, c->stream, tkey, a, tcap);
      ALN_TRY(check_launch("gb_init"));
      {
        ProfScope p(c, "groupby_insert");
        hipLaunchKernelGGL(k_gb_agg<false>, dim3(nb), dim3(256), lds, c->stream, a, n, tkey, tcap,
                           (uint32_t *)nullptr, ctl + 1, glimit, ctl);
      }
      ALN_TRY(check_launch("gb_agg"));
      uint32_t hov = 0;
      ALN_TRY(copy_to_host(c, &hov, ctl, sizeof(uint32_t)));
      if (!hov) break;
      if (tcap >= full) return set_error(ALN_E_CAPACITY, "aln_groupby: hash table overflow");
      tcap = full;
    }
    ALN_TRY(ws(c, tcap, &flag));
    ALN_TRY(ws(c, tcap, &pos));
    hipLaunchKernelGGL(k_gb_flags, dim3(grid_for(tcap, 256, 65536)), dim3(256), 0, c->stream, tkey, tcap, flag);
    ALN_TRY(check_launch("gb_flags"));
    ALN_TRY(exclusive_scan_u32(c, flag, tcap, pos, &g));
    *h_ngroups = g;
    if (g > cap)
      return set_error(ALN_E_CAPACITY, "aln_groupby: %llu groups, capacity %llu", (unsigned long long)g,
                       (unsigned long long)cap);
    ALN_TRY(ws(c, g, &gcode));
    ALN_TRY(ws(c, g, &gslot));
    hipLaunchKernelGGL(k_gb_compact, dim3(grid_for(tcap, 256, 65536)), dim3(256), 0, c->stream, tkey, tcap, flag,
                       pos, gcode, gslot);
    ALN_TRY(check_launch("gb_compact"));
    if (order == ALN_ORDER_KEYS) ALN_TRY(radix_sort_pairs(c, gcode, gslot, g, total));
  }
  if (order == ALN_ORDER_HASH) {
    uint64_t *hash;
    ALN_TRY(ws(c, g, &hash));
    hipLaunchKernelGGL(k_gb_hash, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, a, gcode,
                       g, seed, hash);
    ALN_TRY(check_launch("gb_hash"));
    // sort (hash, group index) then carry the codes along through the index
    uint32_t *gi;
    uint64_t *gcode2;
    uint32_t *gslot2;
    ALN_TRY(ws(c, g, &gi));
    ALN_TRY(ws(c, g, &gcode2));
    ALN_TRY(ws(c, g, &gslot2));
    ALN_TRY(iota_u32(c, gi, g));
    ALN_TRY(radix_sort_pairs(c, hash, gi, g, 64));
    hipLaunchKernelGGL(k_gb_gather, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, gcode,
                       gslot, gi, g, gcode2, gslot2);
    ALN_TRY(check_launch("gb_gather"));
    gcode = gcode2;
    gslot = gslot2;
  }
  {
    ProfScope p(c, "groupby_emit");
    hipLaunchKernelGGL(k_gb_emit, dim3(grid_for(g, 256, 65536)), dim3(256), 0, c->stream, a, o,
                       gcode, gslot, g);
  }
  ALN_TRY(check_launch("gb_emit"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

int aln_f64_order_keys(aln_ctx *c, const int64_t *d_in, uint64_t n, int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || (d_in && d_out)), "aln_f64_order_keys: bad arguments");
  if (n == 0) return ALN_OK;
  hipLaunchKernelGGL(k_f64_keys, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_in, n, 0, d_out);
  return check_launch("f64_order_keys");
}

int aln_f64_unorder_keys(aln_ctx *c, const int64_t *d_in, uint64_t n, int64_t *d_out) {
  ALN_REQUIRE(c && (n == 0 || (d_in && d_out)), "aln_f64_unorder_keys: bad arguments");
  if (n == 0) return ALN_OK;
  hipLaunchKernelGGL(k_f64_keys, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, d_in, n, 1, d_out);
  return check_launch("f64_unorder_keys");
}

int aln_hash_rows(aln_ctx *c, const int64_t *const *d_keys, int nkeys, uint64_t n, uint32_t seed,
                  uint64_t *d_out) {
  ALN_REQUIRE(c && d_keys && nkeys >= 1 && nkeys <= ALN_MAX_KEYS && (n == 0 || d_out),
              "aln_hash_rows: bad arguments");
  if (n == 0) return ALN_OK;
  GbArgs a;
  memset(&a, 0, sizeof(a));
  a.nk = nkeys;
  for (int k = 0; k < nkeys; ++k) a.keys[k] = d_keys[k];
  hipLaunchKernelGGL(k_hash_rows, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream, a, n, seed,
                     d_out);
  return check_launch("hash_rows");
}

}  // extern "C"

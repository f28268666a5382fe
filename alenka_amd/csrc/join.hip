// Inner equi-join on int64 keys.
//
// The reference joins by sorting: the right (build) side is stably sorted by key
// (operators.cu:1117-1157), each left (probe) segment is sorted carrying its row ids
// (operators.cu:1184-1199), and moderngpu inner_join merges them into (left, right) pairs
// (operators.cu:1239-1284).  Here each probe row is looked up directly in a structure over
// the build side:
//   * unique build keys whose range is at most 64 per build row (the primary-key case): a
//     table with one 8-byte entry per 32 keys (bits present, rank base) and the build rows
//     in key order;
//   * other unique keys: an open-addressing table keyed by MurmurHash64A (linear probing), a
//     slot holding the build row;
//   * duplicate build keys: the build side is radix-sorted by key (stable, so equal keys
//     keep ascending row order) and a slot holds a run of equal keys.
// Pairs come out in probe-row order and, for one probe row, in ascending build-row order;
// the count pass records each probe row's match so the write pass does not probe again.
#include <string.h>

#include "aln_internal.h"

namespace aln {

constexpr uint32_t J_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t J_SEED = 0x5bd1e995u;
constexpr int J_BLOCK = 256;
constexpr int J_ITEMS = 8;
constexpr int J_TILE = J_BLOCK * J_ITEMS;

// Build: slot value = index into `keys` (a build row, or a run of equal keys).
__global__ void __launch_bounds__(256) k_join_insert(const int64_t *__restrict__ keys, uint64_t n,
                                                     uint32_t *slot, int64_t *tk, uint64_t cap,
                                                     uint32_t *dup) {
  const uint64_t mask = cap - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    int64_t key = keys[i];
    uint32_t id = (uint32_t)i;
    uint64_t h = murmur64a_u64((uint64_t)key, J_SEED) & mask;
    for (uint64_t probes = 0; probes <= cap; ++probes) {
      uint32_t cur = slot[h];
      if (cur == J_EMPTY) {
        cur = atomicCAS(&slot[h], J_EMPTY, id);
        if (cur == J_EMPTY) {
          tk[h] = key;
          break;
        }
      }
      // occupied: compare with the key of the row (or run) that owns it
      int64_t other = keys[cur];
      if (other == key) {
        atomicExch(dup, 1u);
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__device__ __forceinline__ uint32_t join_find(const uint32_t *__restrict__ slot,
                                              const int64_t *__restrict__ tk, uint64_t mask,
                                              int64_t key) {
  uint64_t h = murmur64a_u64((uint64_t)key, J_SEED) & mask;
  for (uint64_t probes = 0; probes <= mask; ++probes) {
    uint32_t s = slot[h];
    if (s == J_EMPTY) return J_EMPTY;
    if (tk[h] == key) return s;
    h = (h + 1) & mask;
  }
  return J_EMPTY;
}

// PASS 0: per probe row, the matching slot value and per-tile pair counts.
__global__ void __launch_bounds__(J_BLOCK) k_join_probe_count(
    const int64_t *__restrict__ pk, uint64_t np, const uint32_t *__restrict__ slot,
    const int64_t *__restrict__ tk, uint64_t mask, const uint32_t *__restrict__ rlen,
    uint32_t *__restrict__ match, uint32_t *__restrict__ tcount) {
  uint64_t base = (uint64_t)blockIdx.x * J_TILE;
  uint32_t s = 0;
  for (int j = 0; j < J_ITEMS; ++j) {
    uint64_t p = base + (uint64_t)j * J_BLOCK + threadIdx.x;
    if (p < np) {
      uint32_t m = join_find(slot, tk, mask, pk[p]);
      match[p] = m;
      if (m != J_EMPTY) s += rlen ? rlen[m] : 1;
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ uint32_t wsum[J_BLOCK / 64];
  if (__lane_id() == 0) wsum[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < J_BLOCK / 64; ++w) t += wsum[w];
    tcount[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(J_BLOCK) k_join_probe_write(
    uint64_t np, const uint32_t *__restrict__ match, const uint32_t *__restrict__ rstart,
    const uint32_t *__restrict__ rlen, const uint32_t *__restrict__ srow,
    const uint64_t *__restrict__ toff, uint32_t *__restrict__ out_p, uint32_t *__restrict__ out_b) {
  __shared__ uint64_t wsum[J_BLOCK / 64];
  // blocked: thread t owns probe rows base + t*ITEMS .. +ITEMS-1 (keeps probe-row order)
  uint64_t base = (uint64_t)blockIdx.x * J_TILE + (uint64_t)threadIdx.x * J_ITEMS;
  uint32_t cnt[J_ITEMS], m[J_ITEMS];
  uint64_t s = 0;
  for (int j = 0; j < J_ITEMS; ++j) {
    uint64_t p = base + j;
    m[j] = p < np ? match[p] : J_EMPTY;
    cnt[j] = m[j] == J_EMPTY ? 0 : (rlen ? rlen[m[j]] : 1);
    s += cnt[j];
  }
  unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint64_t x = s;
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint64_t before = 0;
  for (int w = 0; w < (int)wid; ++w) before += wsum[w];
  uint64_t pos = toff[blockIdx.x] + before + x - s;
  for (int j = 0; j < J_ITEMS; ++j) {
    if (!cnt[j]) continue;
    uint32_t p = (uint32_t)(base + j);
    if (!rlen) {
      out_p[pos] = p;
      out_b[pos] = m[j];
      ++pos;
    } else {
      uint32_t st = rstart[m[j]];
      for (uint32_t t = 0; t < cnt[j]; ++t, ++pos) {
        out_p[pos] = p;
        out_b[pos] = srow[st + t];
      }
    }
  }
}

// Runs mode helpers
__global__ void k_join_codes(const int64_t *k, uint64_t n, int64_t kmin, uint64_t *code,
                             uint32_t *row) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    code[i] = (uint64_t)k[i] - (uint64_t)kmin;
    row[i] = (uint32_t)i;
  }
}
__global__ void k_join_heads(const uint64_t *code, uint64_t n, uint32_t *head) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    head[i] = i == 0 || code[i] != code[i - 1];
}
__global__ void k_join_runs(const uint64_t *code, const uint32_t *head, const uint64_t *pos,
                            uint64_t n, int64_t kmin, uint32_t *rstart, uint32_t *rlen,
                            int64_t *rkey) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    if (!head[i]) continue;
    uint64_t r = pos[i];
    uint64_t e = i + 1;
    while (e < n && !head[e]) ++e;
    rstart[r] = (uint32_t)i;
    rlen[r] = (uint32_t)(e - i);
    rkey[r] = (int64_t)((uint64_t)kmin + code[i]);
  }
}

// ---- dense unique build keys: bit + rank table -----------------------------------------------
// When the build keys are unique and their range is at most JD_DENSITY per build row, the
// build side becomes a table over the key range with one 32-byte group per 224 keys: seven
// words with the bits of the keys present and the number of keys before the group (its
// rank base), plus the build rows in key order (rowid[rank]).  At 10M build keys over a 20M
// range the table is 2.9 MB, which can stay in an XCD's 4 MB L2 (a probe reads one aligned
// 32-byte group, inside one cache line) as long as no other random reads share the pass.
//   build: a byte per key (plain byte stores, no atomics: every XCD's L2 sees them merged at
//          write-back), packed into the groups with a block scan of the counts, then the row
//          ids scattered to their ranks;
//   probe: pass 0 streams the probe keys and writes each row's rank (or none) and the tile's
//          pair count; pass 1 streams those, gathers the hits' row ids and writes the pairs
//          of a tile through LDS, so the stores to the output are whole, contiguous lines.
constexpr uint64_t JD_DENSITY = 64;  // key range <= 64 x build rows (and <= 2^31)
constexpr int JD_BLOCK = 256;        // groups per block of the packing pass
constexpr uint32_t JD_GWORDS = 7;    // bit words per group (the eighth word is the rank base)
constexpr uint32_t JD_GKEYS = 32 * JD_GWORDS;

__global__ void __launch_bounds__(256) k_jd_bytes(const int64_t *__restrict__ keys, uint64_t n, int64_t kmin,
                                                  uint8_t *bytes) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    bytes[(uint64_t)keys[i] - (uint64_t)kmin] = 1;
}

// Bits of 32 keys from 32 bytes that are each 0 or 1 (bit 0 of each byte, in order).
__device__ __forceinline__ uint32_t jd_byte_bits(uint4 x, uint4 y) {
  const uint32_t v[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
  uint32_t b = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    b |= (((v[k] & 1u) | ((v[k] >> 7) & 2u) | ((v[k] >> 14) & 4u) | ((v[k] >> 21) & 8u))) << (4 * k);
  return b;
}

// Group g: the bits of keys 224g .. 224g+223 from the byte map; pass 0 writes the keys per
// block, pass 1 the groups with their rank bases (block offset + keys of earlier groups).
template <int PASS>
__global__ void __launch_bounds__(JD_BLOCK) k_jd_pack(const uint8_t *__restrict__ bytes, uint64_t ngrp,
                                                      uint32_t *__restrict__ bcount, const uint64_t *__restrict__ boff,
                                                      uint4 *__restrict__ grp) {
  __shared__ uint32_t wsum[JD_BLOCK / 64];
  const uint64_t g = (uint64_t)blockIdx.x * JD_BLOCK + threadIdx.x;
  uint32_t w[JD_GWORDS] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t n = 0;
  if (g < ngrp) {
    const uint4 *q = reinterpret_cast<const uint4 *>(bytes + g * JD_GKEYS);
#pragma unroll
    for (uint32_t k = 0; k < JD_GWORDS; ++k) {
      w[k] = jd_byte_bits(q[2 * k], q[2 * k + 1]);
      n += (uint32_t)__popc(w[k]);
    }
  }
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint32_t x = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (unsigned)o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (PASS == 0) {
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int k = 0; k < JD_BLOCK / 64; ++k) t += wsum[k];
      bcount[blockIdx.x] = t;
    }
    return;
  }
  uint32_t before = 0;
  for (int k = 0; k < (int)wid; ++k) before += wsum[k];
  if (g < ngrp) {
    grp[2 * g] = make_uint4(w[0], w[1], w[2], w[3]);
    grp[2 * g + 1] = make_uint4(w[4], w[5], w[6], (uint32_t)boff[blockIdx.x] + before + x - n);
  }
}

// Rank of key code c among the build keys (*hit: whether c is one of them), from its group.
__device__ __forceinline__ uint32_t jd_rank(uint4 lo, uint4 hi, uint32_t c, bool *hit) {
  const uint32_t wi = c >> 5, k = wi - (wi / JD_GWORDS) * JD_GWORDS, bit = 1u << (c & 31);
  const uint32_t w[JD_GWORDS] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z};
  uint32_t r = hi.w, word = 0;
#pragma unroll
  for (uint32_t j = 0; j < JD_GWORDS; ++j) {
    if (j < k) r += (uint32_t)__popc(w[j]);
    if (j == k) word = w[j];
  }
  *hit = (word & bit) != 0;
  return r + (uint32_t)__popc(word & (bit - 1u));
}

__device__ __forceinline__ uint64_t jd_group(uint32_t c) { return (uint64_t)((c >> 5) / JD_GWORDS); }

__global__ void __launch_bounds__(256) k_jd_rows(const int64_t *__restrict__ keys, uint64_t n, int64_t kmin,
                                                 const uint4 *__restrict__ grp, uint32_t *rowid) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint32_t c = (uint32_t)((uint64_t)keys[i] - (uint64_t)kmin);
    const uint64_t g = jd_group(c);
    bool hit;
    rowid[jd_rank(grp[2 * g], grp[2 * g + 1], c, &hit)] = (uint32_t)i;
  }
}

// Pass 0: each probe row's build-key rank into match (coalesced), and the tile's pair
// count.  A thread's loads go out stage by stage for all of its rows at once - the keys,
// then their groups - so a thread waits two round trips, not one per row and level (rows
// outside the key range read group 0 and are masked afterwards).  The only random reads
// are the groups, so the table keeps its place in L2 (the row ids are gathered by pass 1).
__global__ void __launch_bounds__(J_BLOCK) k_jd_match(const int64_t *__restrict__ pk, uint64_t np,
                                                      const uint4 *__restrict__ grp, int64_t kmin,
                                                      uint64_t range, uint32_t *__restrict__ match,
                                                      uint32_t *__restrict__ tcount) {
  __shared__ uint32_t wsum[J_BLOCK / 64];
  const uint64_t t0 = (uint64_t)blockIdx.x * J_TILE;
  int64_t key[J_ITEMS];
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {  // striped: row t0 + j * J_BLOCK + thread (coalesced)
    const uint64_t p = t0 + (uint64_t)j * J_BLOCK + threadIdx.x;
    key[j] = p < np ? __builtin_nontemporal_load(pk + p) : kmin - 1;
  }
  uint32_t code[J_ITEMS];
  uint4 lo[J_ITEMS], hi[J_ITEMS];
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {
    const uint64_t c = (uint64_t)key[j] - (uint64_t)kmin;
    code[j] = c < range ? (uint32_t)c : 0xFFFFFFFFu;
    const uint64_t g = code[j] != 0xFFFFFFFFu ? jd_group(code[j]) : 0u;
    lo[j] = grp[2 * g];
    hi[j] = grp[2 * g + 1];
  }
  uint32_t m[J_ITEMS], n = 0;
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {
    bool h;
    const uint32_t r = jd_rank(lo[j], hi[j], code[j] != 0xFFFFFFFFu ? code[j] : 0u, &h);
    h = h && code[j] != 0xFFFFFFFFu;
    m[j] = h ? r : J_EMPTY;
    n += h;
  }
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {
    const uint64_t p = t0 + (uint64_t)j * J_BLOCK + threadIdx.x;
    if (p < np) __builtin_nontemporal_store(m[j], match + p);
  }
  for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
  if (__lane_id() == 0) wsum[threadIdx.x / 64] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < J_BLOCK / 64; ++k) t += wsum[k];
    tcount[blockIdx.x] = t;
  }
}

// Pass 1: a tile's pairs in probe-row order, compacted in LDS, then written as contiguous
// runs of the two output arrays.  rowid given: match holds ranks, and the build row of a
// hit is rowid[rank] (all of a thread's gathers in flight at once).
__global__ void __launch_bounds__(J_BLOCK) k_jd_write(const uint32_t *__restrict__ match, uint64_t np,
                                                      const uint32_t *__restrict__ rowid,
                                                      const uint64_t *__restrict__ toff, uint32_t *__restrict__ out_p,
                                                      uint32_t *__restrict__ out_b) {
  __shared__ uint32_t sp[J_TILE], sb[J_TILE];
  __shared__ uint32_t wsum[J_BLOCK / 64];
  const uint64_t t0 = (uint64_t)blockIdx.x * J_TILE;
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint32_t mv[J_ITEMS];
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {
    const uint64_t p = t0 + (uint64_t)j * J_BLOCK + threadIdx.x;
    mv[j] = p < np ? __builtin_nontemporal_load(match + p) : J_EMPTY;
  }
  if (rowid) {
#pragma unroll
    for (int j = 0; j < J_ITEMS; ++j)
      if (mv[j] != J_EMPTY) mv[j] = rowid[mv[j]];
  }
  uint32_t base = 0;  // pairs of the tile before this round
#pragma unroll
  for (int j = 0; j < J_ITEMS; ++j) {  // round j: rows t0 + j * J_BLOCK .. + J_BLOCK - 1, in order
    const uint64_t p = t0 + (uint64_t)j * J_BLOCK + threadIdx.x;
    const uint32_t m = mv[j];
    const bool h = m != J_EMPTY;
    const uint64_t bal = __builtin_amdgcn_ballot_w64(h);
    if (lane == 0) wsum[wid] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < J_BLOCK / 64; ++k) {
      const uint32_t c = wsum[k];
      before += k < (int)wid ? c : 0u;
      all += c;
    }
    if (h) {
      const uint32_t q = base + before + below_count(bal);
      sp[q] = (uint32_t)p;
      sb[q] = m;
    }
    base += all;
    __syncthreads();
  }
  const uint64_t o = toff[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < base; i += J_BLOCK) {
    out_p[o + i] = sp[i];
    out_b[o + i] = sb[i];
  }
}

// The pairs from each probe row's build row (match) and the tiles' pair counts.
// rowid: as for k_jd_write.
static int join_write_pairs(aln_ctx *c, const uint32_t *match, const uint32_t *rowid, uint64_t np,
                            const uint32_t *tcount, uint64_t *toff, uint64_t ntiles, uint32_t *d_probe_idx,
                            uint32_t *d_build_idx, uint64_t cap, uint64_t *h_npairs) {
  uint64_t total = 0;
  ALN_TRY(exclusive_scan_u32(c, tcount, ntiles, toff, &total));
  *h_npairs = total;
  if (total > cap)
    return set_error(ALN_E_CAPACITY, "aln_hash_join: %llu pairs, capacity %llu", (unsigned long long)total,
                     (unsigned long long)cap);
  if (total == 0) return ALN_OK;
  ALN_REQUIRE(d_probe_idx && d_build_idx, "aln_hash_join: no output buffers");
  {
    ProfScope p(c, "join_write");
    hipLaunchKernelGGL(k_jd_write, dim3((unsigned)ntiles), dim3(J_BLOCK), 0, c->stream, match, np, rowid,
                       (const uint64_t *)toff, d_probe_idx, d_build_idx);
  }
  ALN_TRY(check_launch("join_write"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

// The dense path; *taken = false (and nothing else done) when the build keys repeat.
static int join_dense(aln_ctx *c, const int64_t *bk, uint64_t nb, const int64_t *pk, uint64_t np, int64_t kmin,
                      uint64_t range, uint32_t *d_probe_idx, uint32_t *d_build_idx, uint64_t cap,
                      uint64_t *h_npairs, bool *taken) {
  *taken = false;
  const uint64_t ntiles = (np + J_TILE - 1) / J_TILE;
  uint32_t *match, *tcount;
  uint64_t *toff;
  const uint64_t ngrp = (range + JD_GKEYS - 1) / JD_GKEYS, nblk = (ngrp + JD_BLOCK - 1) / JD_BLOCK;
  uint8_t *bytes;
  uint32_t *bcount, *rowid;
  uint64_t *boff;
  uint4 *grp;
  ALN_TRY(ws(c, ngrp * JD_GKEYS, &bytes));  // whole groups: the packing pass reads JD_GKEYS bytes each
  ALN_TRY(ws(c, nblk, &bcount));
  ALN_TRY(ws(c, nblk, &boff));
  ALN_TRY(ws(c, 2 * ngrp, &grp));
  ALN_TRY(fill_u32(c, reinterpret_cast<uint32_t *>(bytes), ngrp * (JD_GKEYS / 4), 0));
  uint64_t distinct = 0;
  {
    ProfScope p(c, "join_build");
    hipLaunchKernelGGL(k_jd_bytes, dim3(grid_for(nb, 256, 65536)), dim3(256), 0, c->stream, bk, nb, kmin, bytes);
    hipLaunchKernelGGL(k_jd_pack<0>, dim3((unsigned)nblk), dim3(JD_BLOCK), 0, c->stream, (const uint8_t *)bytes,
                       ngrp, bcount, (const uint64_t *)nullptr, (uint4 *)nullptr);
  }
  ALN_TRY(check_launch("join_dense_bytes"));
  ALN_TRY(exclusive_scan_u32(c, bcount, nblk, boff, &distinct));
  if (distinct != nb) return ALN_OK;  // repeated build keys
  *taken = true;
  ALN_TRY(ws(c, nb, &rowid));
  {
    ProfScope p(c, "join_build");
    hipLaunchKernelGGL(k_jd_pack<1>, dim3((unsigned)nblk), dim3(JD_BLOCK), 0, c->stream, (const uint8_t *)bytes,
                       ngrp, bcount, (const uint64_t *)boff, grp);
    hipLaunchKernelGGL(k_jd_rows, dim3(grid_for(nb, 256, 65536)), dim3(256), 0, c->stream, bk, nb, kmin,
                       (const uint4 *)grp, rowid);
  }
  ALN_TRY(check_launch("join_dense_rows"));
  ALN_TRY(ws(c, np, &match));
  ALN_TRY(ws(c, ntiles, &tcount));
  ALN_TRY(ws(c, ntiles, &toff));
  {
    ProfScope p(c, "join_probe");
    hipLaunchKernelGGL(k_jd_match, dim3((unsigned)ntiles), dim3(J_BLOCK), 0, c->stream, pk, np, (const uint4 *)grp,
                       kmin, range, match, tcount);
  }
  ALN_TRY(check_launch("join_dense_match"));
  return join_write_pairs(c, match, rowid, np, tcount, toff, ntiles, d_probe_idx, d_build_idx, cap, h_npairs);
}

}  // namespace aln

using namespace aln;

extern "C" int aln_hash_join(aln_ctx *c, const int64_t *d_build_keys, uint64_t nbuild,
                             const int64_t *d_probe_keys, uint64_t nprobe, uint32_t *d_probe_idx,
                             uint32_t *d_build_idx, uint64_t cap, uint64_t *h_npairs) {
  ALN_REQUIRE(c && h_npairs, "aln_hash_join: bad arguments");
  ALN_REQUIRE(nbuild < 0xFFFFFFFFull && nprobe < 0xFFFFFFFFull, "aln_hash_join: more than 2^32-1 rows");
  ALN_REQUIRE((nbuild == 0 || d_build_keys) && (nprobe == 0 || d_probe_keys),
              "aln_hash_join: NULL keys");
  ws_reset(c);
  *h_npairs = 0;
  if (nbuild == 0 || nprobe == 0) return ALN_OK;
  // dense unique build keys: the bitmap + rank index
  int64_t bmin, bmax;
  ALN_TRY(minmax_i64(c, d_build_keys, nbuild, &bmin, &bmax));
  const uint64_t brange = (uint64_t)bmax - (uint64_t)bmin + 1;
  if (brange != 0 && brange <= (1ull << 31) && brange <= JD_DENSITY * nbuild) {
    bool taken = false;
    ALN_TRY(join_dense(c, d_build_keys, nbuild, d_probe_keys, nprobe, bmin, brange, d_probe_idx, d_build_idx,
                       cap, h_npairs, &taken));
    if (taken) return ALN_OK;
    ws_reset(c);
  }
  uint64_t tcap = next_pow2(2 * nbuild < 1024 ? 1024 : 2 * nbuild);
  uint32_t *slot, *dup;
  int64_t *tk;
  ALN_TRY(ws(c, tcap, &slot));
  ALN_TRY(ws(c, tcap, &tk));
  ALN_TRY(ws(c, 1, &dup));
  ALN_TRY(fill_u32(c, slot, tcap, J_EMPTY));
  ALN_TRY(fill_u32(c, dup, 1, 0));
  {
    ProfScope p(c, "join_build");
    hipLaunchKernelGGL(k_join_insert, dim3(grid_for(nbuild, 256, 65536)), dim3(256), 0, c->stream,
                       d_build_keys, nbuild, slot, tk, tcap, dup);
  }
  ALN_TRY(check_launch("join_build"));
  uint32_t hdup = 0;
  ALN_TRY(copy_to_host(c, &hdup, dup, 4));
  uint32_t *rstart = nullptr, *rlen = nullptr, *srow = nullptr;
  if (hdup) {
    // runs mode: stable sort of the build side by key, one slot per run of equal keys
    int64_t kmin, kmax;
    ALN_TRY(minmax_i64(c, d_build_keys, nbuild, &kmin, &kmax));
    uint64_t *code;
    uint32_t *head;
    uint64_t *pos;
    ALN_TRY(ws(c, nbuild, &code));
    ALN_TRY(ws(c, nbuild, &srow));
    ALN_TRY(ws(c, nbuild, &head));
    ALN_TRY(ws(c, nbuild, &pos));
    hipLaunchKernelGGL(k_join_codes, dim3(grid_for(nbuild, 256, 65536)), dim3(256), 0, c->stream,
                       d_build_keys, nbuild, kmin, code, srow);
    ALN_TRY(check_launch("join_codes"));
    ALN_TRY(radix_sort_pairs(c, code, srow, nbuild, bits_for_range_u((uint64_t)kmax - (uint64_t)kmin)));
    hipLaunchKernelGGL(k_join_heads, dim3(grid_for(nbuild, 256, 65536)), dim3(256), 0, c->stream, code,
                       nbuild, head);
    ALN_TRY(check_launch("join_heads"));
    uint64_t nruns = 0;
    ALN_TRY(exclusive_scan_u32(c, head, nbuild, pos, &nruns));
    int64_t *rkey;
    ALN_TRY(ws(c, nruns, &rstart));
    ALN_TRY(ws(c, nruns, &rlen));
    ALN_TRY(ws(c, nruns, &rkey));
    hipLaunchKernelGGL(k_join_runs, dim3(grid_for(nbuild, 256, 65536)), dim3(256), 0, c->stream, code,
                       head, pos, nbuild, kmin, rstart, rlen, rkey);
    ALN_TRY(check_launch("join_runs"));
    ALN_TRY(fill_u32(c, slot, tcap, J_EMPTY));
    {
      ProfScope p(c, "join_build");
      hipLaunchKernelGGL(k_join_insert, dim3(grid_for(nruns, 256, 65536)), dim3(256), 0, c->stream,
                         (const int64_t *)rkey, nruns, slot, tk, tcap, dup);
    }
    ALN_TRY(check_launch("join_build_runs"));
  }
  uint64_t ntiles = (nprobe + J_TILE - 1) / J_TILE;
  uint32_t *match, *tcount;
  uint64_t *toff;
  ALN_TRY(ws(c, nprobe, &match));
  ALN_TRY(ws(c, ntiles, &tcount));
  ALN_TRY(ws(c, ntiles, &toff));
  {
    ProfScope p(c, "join_probe");
    hipLaunchKernelGGL(k_join_probe_count, dim3((unsigned)ntiles), dim3(J_BLOCK), 0, c->stream,
                       d_probe_keys, nprobe, slot, tk, tcap - 1, rlen, match, tcount);
  }
  ALN_TRY(check_launch("join_probe_count"));
  // unique keys: match holds the build row itself, written out through LDS
  if (!rlen)
    return join_write_pairs(c, match, nullptr, nprobe, tcount, toff, ntiles, d_probe_idx, d_build_idx, cap,
                            h_npairs);
  uint64_t total = 0;
  ALN_TRY(exclusive_scan_u32(c, tcount, ntiles, toff, &total));
  *h_npairs = total;
  if (total > cap)
    return set_error(ALN_E_CAPACITY, "aln_hash_join: %llu pairs, capacity %llu",
                     (unsigned long long)total, (unsigned long long)cap);
  if (total == 0) return ALN_OK;
  ALN_REQUIRE(d_probe_idx && d_build_idx, "aln_hash_join: no output buffers");
  {
    ProfScope p(c, "join_write");
    hipLaunchKernelGGL(k_join_probe_write, dim3((unsigned)ntiles), dim3(J_BLOCK), 0, c->stream, nprobe,
                       match, rstart, rlen, srow, toff, d_probe_idx, d_build_idx);
  }
  ALN_TRY(check_launch("join_probe_write"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

namespace aln {

// ---- semi / anti reduction ---------------------------------------------------------------
// The reference narrows a join's pairs to one side's rows with sort + unique (semi) or a
// set difference against 0..n-1 (anti), operators.cu:1357-1399 and 1473-1480.  Here the
// rows a side contributes are marked in a bitmap (one bit per row, atomicOr), and one
// compaction pass writes the set (semi) or clear (anti) rows in ascending order: one
// thread per 32-row word, 8192 rows per block.
constexpr int JR_BLOCK = 256;

// Each thread takes 8 consecutive indices: a join's pairs come in probe-row order, so they
// are mostly rows of one bitmap word, ORed locally and set with one atomic per word.
__global__ void __launch_bounds__(256) k_rows_mark(const uint32_t *__restrict__ idx, uint64_t n,
                                                   uint64_t nrows, uint32_t *bm) {
  const uint64_t nchunks = (n + 7) / 8;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < nchunks; t += (uint64_t)gridDim.x * 256) {
    uint32_t r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint64_t i = t * 8 + u;
      r[u] = 0xFFFFFFFFu;
      if (i < n) r[u] = idx[i];
    }
    uint32_t cw = 0xFFFFFFFFu, cb = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (r[u] >= nrows) continue;  // past the end, or not a row
      if ((r[u] >> 5) != cw) {
        if (cb) atomicOr(&bm[cw], cb);
        cw = r[u] >> 5;
        cb = 0;
      }
      cb |= 1u << (r[u] & 31);
    }
    if (cb) atomicOr(&bm[cw], cb);
  }
}

// Rows of word w that are kept: its marked rows (semi) or unmarked rows (anti), clipped to
// the row count.
__device__ __forceinline__ uint32_t kept_rows(const uint32_t *__restrict__ bm, uint64_t w, uint64_t nrows,
                                              int anti) {
  if (w >= (nrows + 31) >> 5) return 0;
  uint32_t x = bm[w];
  if (anti) x = ~x;
  const uint64_t rem = nrows - (w << 5);
  if (rem < 32) x &= (1u << rem) - 1u;
  return x;
}

__global__ void __launch_bounds__(JR_BLOCK) k_rows_count(const uint32_t *__restrict__ bm, uint64_t nrows,
                                                         int anti, uint32_t *counts) {
  __shared__ uint32_t wsum[JR_BLOCK / 64];
  const uint64_t w = (uint64_t)blockIdx.x * JR_BLOCK + threadIdx.x;
  uint32_t n = __popc(kept_rows(bm, w, nrows, anti));
  for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
  if (__lane_id() == 0) wsum[threadIdx.x / 64] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int k = 0; k < JR_BLOCK / 64; ++k) s += wsum[k];
    counts[blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(JR_BLOCK) k_rows_write(const uint32_t *__restrict__ bm, uint64_t nrows,
                                                         int anti, const uint64_t *__restrict__ off,
                                                         uint32_t *__restrict__ out) {
  __shared__ uint32_t wsum[JR_BLOCK / 64];
  const uint64_t w = (uint64_t)blockIdx.x * JR_BLOCK + threadIdx.x;
  uint32_t x = kept_rows(bm, w, nrows, anti);
  const uint32_t n = __popc(x);
  const unsigned lane = __lane_id(), wid = threadIdx.x / 64;
  uint32_t incl = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= (unsigned)o) incl += y;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (unsigned k = 0; k < wid; ++k) before += wsum[k];
  uint64_t pos = off[blockIdx.x] + before + incl - n;
  while (x) {
    out[pos++] = (uint32_t)((w << 5) + (uint32_t)(__ffs(x) - 1));
    x &= x - 1;
  }
}

}  // namespace aln

extern "C" int aln_join_rows(aln_ctx *c, const uint32_t *d_idx, uint64_t nidx, uint64_t nrows, int anti,
                             uint32_t *d_rows, uint64_t cap, uint64_t *h_count) {
  using namespace aln;
  ALN_REQUIRE(c && h_count, "aln_join_rows: bad arguments");
  ALN_REQUIRE(nrows < 0xFFFFFFFFull, "aln_join_rows: more than 2^32-1 rows");
  ALN_REQUIRE(nidx == 0 || d_idx, "aln_join_rows: NULL row indices");
  ws_reset(c);
  *h_count = 0;
  if (nrows == 0) return ALN_OK;
  const uint64_t nw = (nrows + 31) / 32, nblocks = (nw + JR_BLOCK - 1) / JR_BLOCK;
  uint32_t *bm, *counts;
  uint64_t *off;
  ALN_TRY(ws(c, nw, &bm));
  ALN_TRY(ws(c, nblocks, &counts));
  ALN_TRY(ws(c, nblocks, &off));
  ALN_TRY(fill_u32(c, bm, nw, 0));
  if (nidx) {
    hipLaunchKernelGGL(k_rows_mark, dim3(grid_for((nidx + 7) / 8, 256, 65536)), dim3(256), 0, c->stream, d_idx,
                       nidx, nrows, bm);
    ALN_TRY(check_launch("join_rows_mark"));
  }
  hipLaunchKernelGGL(k_rows_count, dim3((unsigned)nblocks), dim3(JR_BLOCK), 0, c->stream, (const uint32_t *)bm,
                     nrows, anti ? 1 : 0, counts);
  ALN_TRY(check_launch("join_rows_count"));
  uint64_t total = 0;
  ALN_TRY(exclusive_scan_u32(c, counts, nblocks, off, &total));
  *h_count = total;
  if (total > cap)
    return set_error(ALN_E_CAPACITY, "aln_join_rows: %llu rows, capacity %llu", (unsigned long long)total,
                     (unsigned long long)cap);
  if (total == 0) return ALN_OK;
  ALN_REQUIRE(d_rows, "aln_join_rows: no output buffer");
  hipLaunchKernelGGL(k_rows_write, dim3((unsigned)nblocks), dim3(JR_BLOCK), 0, c->stream, (const uint32_t *)bm,
                     nrows, anti ? 1 : 0, (const uint64_t *)off, d_rows);
  ALN_TRY(check_launch("join_rows_write"));
  ALN_HIP(hipStreamSynchronize(c->stream));
  return ALN_OK;
}

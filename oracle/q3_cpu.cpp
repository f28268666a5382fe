// Multi-threaded CPU run of the TPC-H q3 script over host arrays: the CPU baseline the
// benchmark times beside the GPU (test infrastructure, like the rest of oracle/).
//
// It follows the reference's operator algorithms, parallelised with OpenMP and the GNU
// parallel-mode sorts instead of Thrust's OpenMP backend:
//   FILTER          count + copy_if compaction of the passing row ids (filter_op, cm.cu:3609-3666)
//   JOIN            sort both sides by key, merge join (emit_multijoin, operators.cu:1184-1240)
//   SELECT/GROUP BY stable sort by the group columns, run heads, segmented SUM/COUNT
//                   (order_inplace operators.cu:1969, GroupBy cm.cu:910-956, select.cu:481-580)
//   ORDER ... LIMIT stable sort by (sum DESC, orderdate ASC), first rows (sorts.cu:17-125)
// Revenue is SUM(price * (100 - discount)) at decimal scale 4, exact int64.
#include <omp.h>
#include <parallel/algorithm>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

// Ascending ids of the rows where pred(i) holds: per-thread counts, prefix, then writes.
template <class Pred>
std::vector<uint32_t> compact(uint64_t n, Pred pred) {
  const int T = omp_get_max_threads();
  std::vector<uint64_t> cnt(T + 1, 0);
  std::vector<uint32_t> out;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
    uint64_t c = 0;
    for (uint64_t i = b; i < e; ++i) c += pred(i) ? 1 : 0;
    cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
    {
      for (int k = 0; k < nt; ++k) cnt[k + 1] += cnt[k];
      out.resize(cnt[nt]);
    }
    uint64_t o = cnt[t];
    for (uint64_t i = b; i < e; ++i)
      if (pred(i)) out[o++] = (uint32_t)i;
  }
  return out;
}

struct KeyIdx {
  int64_t key;
  uint32_t idx;
};

// Inner join of left rows (lkey[li]) with right rows (rkey[ri]) on equal keys, both sides
// sorted by key first; pairs come out in left-key order.
void merge_join(const int64_t *lkey, const std::vector<uint32_t> &lids, const int64_t *rkey,
                const std::vector<uint32_t> &rids, std::vector<uint32_t> &lout, std::vector<uint32_t> &rout) {
  std::vector<KeyIdx> L(lids.size()), R(rids.size());
#pragma omp parallel for
  for (size_t i = 0; i < lids.size(); ++i) L[i] = {lkey[lids[i]], lids[i]};
#pragma omp parallel for
  for (size_t i = 0; i < rids.size(); ++i) R[i] = {rkey[rids[i]], rids[i]};
  auto by_key = [](const KeyIdx &a, const KeyIdx &b) { return a.key < b.key; };
  __gnu_parallel::stable_sort(L.begin(), L.end(), by_key);
  __gnu_parallel::stable_sort(R.begin(), R.end(), by_key);
  // matches per left row (binary search on the sorted right side), prefix, then pairs
  const size_t n = L.size();
  std::vector<uint64_t> off(n + 1, 0);
#pragma omp parallel for
  for (size_t i = 0; i < n; ++i) {
    auto lo = std::lower_bound(R.begin(), R.end(), L[i], by_key);
    auto hi = std::upper_bound(lo, R.end(), L[i], by_key);
    off[i + 1] = (uint64_t)(hi - lo);
  }
  for (size_t i = 0; i < n; ++i) off[i + 1] += off[i];
  lout.resize(off[n]);
  rout.resize(off[n]);
#pragma omp parallel for
  for (size_t i = 0; i < n; ++i) {
    auto lo = std::lower_bound(R.begin(), R.end(), L[i], by_key);
    uint64_t o = off[i];
    for (auto it = lo; it != R.end() && it->key == L[i].key; ++it, ++o) {
      lout[o] = L[i].idx;
      rout[o] = it->idx;
    }
  }
}

}  // namespace

extern "C" {

// Returns the number of result rows written to top (limit x 5: key, date, prio, sum, count),
// or -1 on bad arguments.  threads <= 0 uses OpenMP's default.
int q3cpu_run(const int64_t *c_key, const int64_t *c_seg, uint64_t nd, const int64_t *o_key,
              const int64_t *o_cust, const int64_t *o_date, const int64_t *o_prio, uint64_t no,
              const int64_t *l_okey, const int64_t *l_price, const int64_t *l_disc, const int64_t *l_ship,
              uint64_t nl, int64_t seg_code, int64_t date, int threads, int limit, int64_t *top) {
  if (!top || limit < 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
  // FILTERs
  std::vector<uint32_t> cf = compact(nd, [&](uint64_t i) { return c_seg[i] == seg_code; });
  std::vector<uint32_t> ofi = compact(no, [&](uint64_t i) { return o_date[i] < date; });
  std::vector<uint32_t> lf = compact(nl, [&](uint64_t i) { return l_ship[i] > date; });
  // OLC1 := OFI JOIN CF ON o_custkey = c_custkey; OLC := LF JOIN OLC1 ON orderkey = o_orderkey
  std::vector<uint32_t> o1, c1;
  merge_join(o_cust, ofi, c_key, cf, o1, c1);
  std::vector<uint32_t> lrow, orow;
  merge_join(l_okey, lf, o_key, o1, lrow, orow);
  const size_t m = lrow.size();
  // GROUP BY o_orderkey, o_orderdate, o_shippriority: stable sort of the joined rows
  std::vector<uint32_t> perm(m);
  std::iota(perm.begin(), perm.end(), 0u);
  __gnu_parallel::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
    const uint32_t x = orow[a], y = orow[b];
    if (o_key[x] != o_key[y]) return o_key[x] < o_key[y];
    if (o_date[x] != o_date[y]) return o_date[x] < o_date[y];
    return o_prio[x] < o_prio[y];
  });
  // run heads, then one SUM/COUNT per run
  std::vector<uint32_t> heads = compact(m, [&](uint64_t i) {
    if (i == 0) return true;
    const uint32_t x = orow[perm[i]], y = orow[perm[i - 1]];
    return o_key[x] != o_key[y] || o_date[x] != o_date[y] || o_prio[x] != o_prio[y];
  });
  const size_t g = heads.size();
  struct Group {
    int64_t key, date, prio, sum, count;
  };
  std::vector<Group> groups(g);
#pragma omp parallel for
  for (size_t k = 0; k < g; ++k) {
    const uint64_t b = heads[k], e = k + 1 < g ? heads[k + 1] : m;
    int64_t s = 0;
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t r = lrow[perm[i]];
      s += l_price[r] * (100 - l_disc[r]);
    }
    const uint32_t x = orow[perm[b]];
    groups[k] = {o_key[x], o_date[x], o_prio[x], s, (int64_t)(e - b)};
  }
  // ORDER BY sum_revenue DESC, orderdate ASC (stable)
  __gnu_parallel::stable_sort(groups.begin(), groups.end(), [](const Group &a, const Group &b) {
    if (a.sum != b.sum) return a.sum > b.sum;
    return a.date < b.date;
  });
  const int n = (int)std::min<size_t>((size_t)limit, g);
  for (int i = 0; i < n; ++i) {
    top[5 * i + 0] = groups[i].key;
    top[5 * i + 1] = groups[i].date;
    top[5 * i + 2] = groups[i].prio;
    top[5 * i + 3] = groups[i].sum;
    top[5 * i + 4] = groups[i].count;
  }
  return n;
}

int q3cpu_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

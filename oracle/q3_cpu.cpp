// CPU baseline of the benchmark workloads over host arrays, timed beside the GPU (test
// infrastructure, like the rest of oracle/): the reference's operator algorithms on
// Thrust's OpenMP backend (rocThrust, THRUST_DEVICE_SYSTEM_OMP), which is the reference's
// own Thrust code path run on host cores.
//
//   FILTER          count + copy_if of the passing row ids (filter_op, cm.cu:3609-3666)
//   JOIN            sort_by_key of the right side, then of the left keys, merge join of the
//                   two sorted key lists (emit_multijoin, operators.cu:1117-1240)
//   SELECT/GROUP BY one stable_sort_by_key per group column, last column first, on a
//                   permutation (order_inplace operators.cu:367-406, update_permutation
//                   sorts.cu:17-70), run heads (GroupBy cm.cu:910-956), reduce_by_key
//                   SUM/COUNT per run (select.cu:481-580)
//   ORDER ... LIMIT one stable_sort_by_key per ORDER BY column, last first (emit_order,
//                   operators.cu:1641-1798), first `limit` rows
// Revenue is SUM(price * (100 - discount)) at decimal scale 4, exact int64.
#include <stddef.h>  // rocThrust headers use ::size_t

#include <thrust/binary_search.h>
#include <thrust/copy.h>
#include <thrust/execution_policy.h>
#include <thrust/functional.h>
#include <thrust/gather.h>
#include <thrust/iterator/constant_iterator.h>
#include <thrust/iterator/counting_iterator.h>
#include <thrust/iterator/discard_iterator.h>
#include <thrust/iterator/permutation_iterator.h>
#include <thrust/iterator/transform_iterator.h>
#include <thrust/iterator/zip_iterator.h>
#include <thrust/reduce.h>
#include <thrust/scan.h>
#include <thrust/sequence.h>
#include <thrust/sort.h>
#include <thrust/transform.h>

#include <omp.h>
#include <stdint.h>

#include <vector>

namespace {

auto P = thrust::omp::par;
using u32 = uint32_t;
using i64 = int64_t;

// Ascending ids of the rows 0..n-1 where pred(i) holds (count + copy_if).
template <class Pred>
std::vector<u32> compact(uint64_t n, Pred pred) {
  auto first = thrust::counting_iterator<u32>(0);
  const uint64_t k = thrust::count_if(P, first, first + n, pred);
  std::vector<u32> out(k);
  thrust::copy_if(P, first, first + n, out.begin(), pred);
  return out;
}

// ids sorted by key[id] (stable), keys gathered alongside.
void sort_by_col(const i64 *col, const std::vector<u32> &ids, std::vector<i64> &keys, std::vector<u32> &sorted) {
  keys.resize(ids.size());
  sorted = ids;
  thrust::gather(P, ids.begin(), ids.end(), col, keys.begin());
  thrust::stable_sort_by_key(P, keys.begin(), keys.end(), sorted.begin());
}

// Inner join on lcol[lids] == rcol[rids]: both sides sorted by key, then for each left row
// the equal range of the right side (lower/upper bound), pairs in left-key order.
void merge_join(const i64 *lcol, const std::vector<u32> &lids, const i64 *rcol, const std::vector<u32> &rids,
                std::vector<u32> &lout, std::vector<u32> &rout) {
  std::vector<i64> lk, rk;
  std::vector<u32> ls, rs;
  sort_by_col(rcol, rids, rk, rs);
  sort_by_col(lcol, lids, lk, ls);
  const size_t n = lk.size();
  std::vector<u32> lo(n), hi(n);
  thrust::lower_bound(P, rk.begin(), rk.end(), lk.begin(), lk.end(), lo.begin());
  thrust::upper_bound(P, rk.begin(), rk.end(), lk.begin(), lk.end(), hi.begin());
  std::vector<uint64_t> off(n + 1, 0);
  thrust::transform(P, hi.begin(), hi.end(), lo.begin(), off.begin() + 1, thrust::minus<uint64_t>());
  thrust::inclusive_scan(P, off.begin() + 1, off.end(), off.begin() + 1);
  lout.resize(off[n]);
  rout.resize(off[n]);
  u32 *lo_ = lo.data(), *hi_ = hi.data(), *ls_ = ls.data(), *rs_ = rs.data(), *lout_ = lout.data(),
      *rout_ = rout.data();
  const uint64_t *off_ = off.data();
  thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(n), [=](size_t i) {
    uint64_t o = off_[i];
    for (u32 j = lo_[i]; j < hi_[i]; ++j, ++o) {
      lout_[o] = ls_[i];
      rout_[o] = rs_[j];
    }
  });
}

// Stable LSD sort of perm by the columns cols[c][rowmap[perm[i]]], last column first,
// ascending (desc[c] = descending), as order_inplace / update_permutation do.
void order_perm(std::vector<u32> &perm, const std::vector<const i64 *> &cols, const std::vector<int> &desc,
                const u32 *rowmap) {
  std::vector<i64> tmp(perm.size());
  for (int c = (int)cols.size() - 1; c >= 0; --c) {
    const i64 *col = cols[c];
    i64 *t = tmp.data();
    const u32 *pm = perm.data();
    thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(perm.size()),
                     [=](size_t i) { t[i] = col[rowmap ? rowmap[pm[i]] : pm[i]]; });
    if (desc[c])
      thrust::stable_sort_by_key(P, tmp.begin(), tmp.end(), perm.begin(), thrust::greater<i64>());
    else
      thrust::stable_sort_by_key(P, tmp.begin(), tmp.end(), perm.begin());
  }
}

// Run heads of rows perm[0..m) over the key columns (GroupBy, cm.cu:910-956): segment ids,
// one per row, 0-based.
std::vector<u32> run_ids(const std::vector<u32> &perm, const std::vector<const i64 *> &keys, const u32 *rowmap) {
  const size_t m = perm.size();
  std::vector<u32> head(m);
  const u32 *pm = perm.data();
  u32 *h = head.data();
  const std::vector<const i64 *> *kp = &keys;
  thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(m), [=](size_t i) {
    if (i == 0) {
      h[i] = 0;
      return;
    }
    const u32 x = rowmap ? rowmap[pm[i]] : pm[i], y = rowmap ? rowmap[pm[i - 1]] : pm[i - 1];
    u32 d = 0;
    for (const i64 *k : *kp) d |= k[x] != k[y];
    h[i] = d;
  });
  thrust::inclusive_scan(P, head.begin(), head.end(), head.begin());
  return head;
}

}  // namespace

extern "C" {

// q3 over host columns.  Returns the number of result rows written to top (limit x 5: key,
// date, prio, sum, count), or -1 on bad arguments; stats (may be null) receives
// {dimension rows passing, orders rows joined to it, fact rows joined, groups}.
// threads <= 0 uses OpenMP's default.
int q3cpu_run_stats(const i64 *c_key, const i64 *c_seg, uint64_t nd, const i64 *o_key, const i64 *o_cust,
                    const i64 *o_date, const i64 *o_prio, uint64_t no, const i64 *l_okey, const i64 *l_price,
                    const i64 *l_disc, const i64 *l_ship, uint64_t nl, i64 seg_code, i64 date, int threads,
                    int limit, i64 *top, i64 *stats) {
  if (!top || limit < 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
  // CF, OFI, LF
  std::vector<u32> cf = compact(nd, [=](u32 i) { return c_seg[i] == seg_code; });
  std::vector<u32> ofi = compact(no, [=](u32 i) { return o_date[i] < date; });
  std::vector<u32> lf = compact(nl, [=](u32 i) { return l_ship[i] > date; });
  // OLC1 := OFI JOIN CF ON o_custkey = c_custkey; OLC := LF JOIN OLC1 ON orderkey = o_orderkey
  std::vector<u32> o1, c1, lrow, orow;
  merge_join(o_cust, ofi, c_key, cf, o1, c1);
  merge_join(l_okey, lf, o_key, o1, lrow, orow);
  const size_t m = lrow.size();
  // GROUP BY o_orderkey, o_orderdate, o_shippriority
  std::vector<u32> perm(m);
  thrust::sequence(P, perm.begin(), perm.end());
  const std::vector<const i64 *> gcols = {o_key, o_date, o_prio};
  order_perm(perm, gcols, {0, 0, 0}, orow.data());
  std::vector<u32> rid = run_ids(perm, gcols, orow.data());
  const size_t g = m ? rid[m - 1] + 1 : 0;
  // SUM(price * (100 - discount)), COUNT per run
  std::vector<i64> rev(m), sum(g), cnt(g);
  {
    const u32 *pm = perm.data(), *lr = lrow.data();
    i64 *rv = rev.data();
    thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(m), [=](size_t i) {
      const u32 r = lr[pm[i]];
      rv[i] = l_price[r] * (100 - l_disc[r]);
    });
  }
  thrust::reduce_by_key(P, rid.begin(), rid.end(), rev.begin(), thrust::make_discard_iterator(), sum.begin());
  thrust::reduce_by_key(P, rid.begin(), rid.end(), thrust::make_constant_iterator<i64>(1),
                        thrust::make_discard_iterator(), cnt.begin());
  // each group's first row gives its key columns
  std::vector<u32> first(g);
  {
    const u32 *r = rid.data(), *pm = perm.data(), *orw = orow.data();
    u32 *f = first.data();
    thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(m), [=](size_t i) {
      if (i == 0 || r[i] != r[i - 1]) f[r[i]] = orw[pm[i]];
    });
  }
  std::vector<i64> gdate(g);
  thrust::gather(P, first.begin(), first.end(), o_date, gdate.begin());
  // ORDER BY sum_revenue DESC, orderdate ASC
  std::vector<u32> gperm(g);
  thrust::sequence(P, gperm.begin(), gperm.end());
  order_perm(gperm, {sum.data(), gdate.data()}, {1, 0}, nullptr);
  const int n = (int)(g < (size_t)limit ? g : (size_t)limit);
  for (int i = 0; i < n; ++i) {
    const u32 k = gperm[i], r = first[k];
    top[5 * i + 0] = o_key[r];
    top[5 * i + 1] = o_date[r];
    top[5 * i + 2] = o_prio[r];
    top[5 * i + 3] = sum[k];
    top[5 * i + 4] = cnt[k];
  }
  if (stats) {
    stats[0] = (i64)cf.size();
    stats[1] = (i64)o1.size();
    stats[2] = (i64)m;
    stats[3] = (i64)g;
  }
  return n;
}

int q3cpu_run(const i64 *c_key, const i64 *c_seg, uint64_t nd, const i64 *o_key, const i64 *o_cust,
              const i64 *o_date, const i64 *o_prio, uint64_t no, const i64 *l_okey, const i64 *l_price,
              const i64 *l_disc, const i64 *l_ship, uint64_t nl, i64 seg_code, i64 date, int threads, int limit,
              i64 *top) {
  return q3cpu_run_stats(c_key, c_seg, nd, o_key, o_cust, o_date, o_prio, no, l_okey, l_price, l_disc, l_ship, nl,
                         seg_code, date, threads, limit, top, nullptr);
}

// The flat query over one segment: WHERE date < date_lt GROUP BY key,
// SUM(price * (100 - discount)), COUNT(*); groups in key order.  Writes up to cap groups
// and returns the group count (-1 on bad arguments).
int64_t flatcpu_run(const i64 *key, const i64 *date, const i64 *price, const i64 *disc, uint64_t n, i64 date_lt,
                    int threads, i64 *out_key, i64 *out_sum, i64 *out_cnt, uint64_t cap) {
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<u32> ids = compact(n, [=](u32 i) { return date[i] < date_lt; });
  const size_t m = ids.size();
  std::vector<u32> perm(m);
  thrust::sequence(P, perm.begin(), perm.end());
  order_perm(perm, {key}, {0}, ids.data());
  std::vector<i64> k(m), rev(m);
  {
    const u32 *pm = perm.data(), *id = ids.data();
    i64 *kk = k.data(), *rv = rev.data();
    thrust::for_each(P, thrust::counting_iterator<size_t>(0), thrust::counting_iterator<size_t>(m), [=](size_t i) {
      const u32 r = id[pm[i]];
      kk[i] = key[r];
      rv[i] = price[r] * (100 - disc[r]);
    });
  }
  std::vector<i64> gk(m), gs(m), gc(m);
  auto e = thrust::reduce_by_key(P, k.begin(), k.end(), rev.begin(), gk.begin(), gs.begin());
  const size_t g = e.first - gk.begin();
  thrust::reduce_by_key(P, k.begin(), k.end(), thrust::make_constant_iterator<i64>(1), thrust::make_discard_iterator(),
                        gc.begin());
  if (!out_key || !out_sum || !out_cnt) return (int64_t)g;
  const size_t w = g < cap ? g : cap;
  for (size_t i = 0; i < w; ++i) {
    out_key[i] = gk[i];
    out_sum[i] = gs[i];
    out_cnt[i] = gc[i];
  }
  return (int64_t)g;
}

int q3cpu_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

/*
 * oracle.h - CPU restatement of the reference's relational operator semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker; the product path
 * (alenka_amd/, libalenka_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.  Parity pins
 * (which check guards which function) are listed in the header of oracle.c.
 */
#ifndef ALENKA_ORACLE_H
#define ALENKA_ORACLE_H

#include <stdint.h>

#include "../include/alenka_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hashing / random streams */
uint64_t orc_murmur64a(const void *key, int len, uint32_t seed);
uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_stream_key(uint64_t seed, uint32_t stream);
uint64_t orc_rnd(uint64_t sk, uint64_t i);
uint64_t orc_uniform(uint64_t sk, uint64_t i, uint64_t n);
int32_t orc_day_to_ymd(int32_t day);

/* synthetic tables (the bench workloads, any scale) */
void orc_gen_flat(uint64_t seed, uint64_t n, int64_t *key, int64_t *date, int64_t *price,
                  int64_t *disc);
uint64_t orc_gen_fact_rows(uint64_t seed, uint64_t norders);
void orc_gen_dim(uint64_t seed, uint64_t ndim, int64_t *c_key, int64_t *c_seg_idx);
void orc_gen_orders(uint64_t seed, uint64_t norders, uint64_t ndim, int64_t *o_key,
                    int64_t *o_cust, int64_t *o_date, int64_t *o_prio);
void orc_gen_fact(uint64_t seed, uint64_t norders, int64_t *l_okey, int64_t *l_price,
                  int64_t *l_disc, int64_t *l_ship);
void orc_dict_first_seen(const int64_t *vals, uint64_t n, int64_t nvals, int64_t *code_of_val,
                         int64_t *codes);
/* One char[width] ORDER BY key (strings_sort_device.cu:125-137): stable sort of the row ids
 * in perm by their values, unsigned byte order, descending when desc */
void orc_str_sort(const char *vals, uint32_t width, uint64_t n, int desc, uint32_t *perm);
/* Semi ('1' left, '2' right) and anti ('3' left, '4' right) joins on nkeys equal keys
 * (operators.cu:1357-1399, 1473-1530); returns the kept rows of that side in out */
uint64_t orc_semi_anti(const int64_t *const *lkeys, const int64_t *const *rkeys, int nkeys, uint64_t nl,
                       uint64_t nr, const uint64_t *lseg_bounds, int nlsegs, char kind, uint32_t *out);
/* LIKE on one fixed-width char value (filter.cu:67-129): 1 when `pat` matches the first
 * `len` bytes of `s` (NUL-padded) */
int orc_like(const char *s, uint32_t len, const char *pat);

/* codecs */
uint32_t orc_bits_for_range(int64_t lo, int64_t hi);
void orc_pack(const int64_t *v, uint64_t n, int64_t base, uint32_t bits, uint32_t *words);
int64_t orc_unpack(const uint32_t *words, uint64_t i, int64_t base, uint32_t bits);
uint64_t orc_pfor_encode(const int64_t *v, uint32_t n, uint8_t *out, uint64_t cap);
int64_t orc_pfor_decode(const uint8_t *buf, uint64_t len, int64_t *out, uint64_t cap);
uint64_t orc_pfor_delta_encode(const int64_t *v, uint32_t n, uint8_t *out, uint64_t cap);

/* predicates and expressions */
int orc_filter(const aln_rpn *rpn, int nrpn, const int64_t *const *cols, const int32_t *col_dec,
               uint64_t n, uint8_t *mask);
int orc_eval(const aln_rpn *rpn, int nrpn, const int64_t *const *cols, const int32_t *col_dec,
             uint64_t n, int64_t *out, int32_t *out_prec);
int orc_zone_map_check(const aln_rpn *rpn, int nrpn, const int64_t *cmin, const int64_t *cmax,
                       const int32_t *col_dec, int not_compressed, char *out);
uint64_t orc_compact(const uint8_t *mask, uint64_t n, uint32_t row0, uint32_t *ids);

/* group-by, join, order */
uint64_t orc_groupby_sorted(const int64_t *const *keys, int nk, const int64_t *const *vals,
                            const int32_t *aggs, int nv, uint64_t n, int64_t *const *out_keys,
                            int64_t *const *out_vals);
uint64_t orc_groupby_segmented(const int64_t *const *keys, int nk, const int64_t *const *vals,
                               const int32_t *aggs, int nv, uint64_t n, const uint64_t *seg_bounds,
                               int nsegs, uint32_t seed, int64_t *const *out_keys,
                               int64_t *const *out_vals);
uint64_t orc_join_sortmerge(const int64_t *lkeys, uint64_t nl, const int64_t *rkeys, uint64_t nr,
                            const uint64_t *lseg_bounds, int nlsegs, uint32_t *out_l,
                            uint32_t *out_r, uint64_t cap);
void orc_order(const int64_t *const *keys, const int32_t *desc, int nk, uint64_t n,
               uint32_t *perm);
uint64_t orc_kahan_sum_f64(const double *v, uint64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif

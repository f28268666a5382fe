/*
 * alenka_hip.h - C ABI of the MI355X relational operator hot path
 * (filter -> decode -> join -> group-by -> order/top-k).
 *
 * Conventions
 *   - every entry point returns an int status: ALN_OK (0) or a negative ALN_E_* code;
 *     aln_last_error() gives the message of the calling thread's last failure.  Nothing
 *     calls exit(); the host layer turns a status into the reference's process_error().
 *   - "d_" arguments are device pointers (HBM), "h_" arguments host pointers.
 *   - all device work is enqueued on the context's stream; entry points that return a
 *     count through an h_ pointer synchronise that stream before returning.
 *   - an aln_ctx is not re-entrant: use one per host thread / stream.
 *
 * Reference interfaces replaced (file:line in antonmks/Alenka):
 *   filter()          filter.h:18-19, filter.cu:133-643      -> aln_compile_rpn + aln_scan_filter
 *   zone_map_check()  zone_map.h:17, zone_map.cu:720-1135    -> aln_zone_map_check
 *   filter_op()       cm.cu:3609-3666                        -> aln_scan_filter
 *   pfor_decompress() compress.cu:180-267                    -> aln_decode / fused decode in scans
 *   pfor_compress()   compress.cu:411-527                    -> aln_encode_plan + aln_encode_pack
 *   mygather/copyColumns cm.cu:3305-3413                     -> aln_gather, aln_eval
 *   emit_multijoin + mgpu::inner_join operators.cu:940-1582  -> aln_hash_join
 *   CudaSet::GroupBy + select() segreduce cm.cu:910-956, select.cu:481-580 -> aln_groupby
 *   add()/count_avg() merge.cu:92-496 (hash-ordered merge)   -> aln_groupby(order=ALN_ORDER_HASH)
 *   semi/anti joins: unique / set_difference operators.cu:1357-1399 -> aln_join_rows
 *   update_permutation/apply_permutation sorts.cu:17-125     -> aln_sort_keys, aln_apply_perm
 *   emit_order + Store(limit) operators.cu:1641-1798, cm.cu:1493-1498 -> aln_topk
 *   MurmurHash64A     MurmurHash2_64.cu:63-104               -> aln_murmur64a, aln_hash_rows
 */
#ifndef ALENKA_HIP_H
#define ALENKA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
#define ALN_OK 0
#define ALN_E_ARG (-1)         /* invalid argument / shape                                */
#define ALN_E_HIP (-2)         /* a HIP runtime call failed                               */
#define ALN_E_NOMEM (-3)       /* device allocation failed                                */
#define ALN_E_UNSUPPORTED (-4) /* construct the reference also rejects or leaves undefined */
#define ALN_E_CAPACITY (-5)    /* output capacity too small: *h_count holds the need      */

const char *aln_last_error(void);
const char *aln_version(void);

/* ---- context ----------------------------------------------------------------------- */
typedef struct aln_ctx aln_ctx;
/* stream: a hipStream_t (NULL = the device's null stream). */
int aln_create(int device, void *stream, aln_ctx **out);
void aln_destroy(aln_ctx *ctx);
int aln_set_stream(aln_ctx *ctx, void *stream);
int aln_sync(aln_ctx *ctx);
/* Per-kernel timing with HIP events recorded on the context stream around every launch
 * (off by default).  aln_profile_get returns the summed milliseconds and launch count of
 * one kernel name since the last aln_profile_reset. */
int aln_profile_enable(aln_ctx *ctx, int on);
int aln_profile_reset(aln_ctx *ctx);
int aln_profile_get(aln_ctx *ctx, const char *kernel, double *h_ms, uint64_t *h_launches);

/* ---- RPN as the reference's parser queues carry it (operators.cu:68-233) ------------- */
enum {
  ALN_RPN_NAME = 1,   /* arg = column slot                                              */
  ALN_RPN_NUMBER = 2, /* val = integer digits, arg = precision (emit_decimal)          */
  ALN_RPN_STRING = 3, /* val = dictionary code of the literal in the column it is
                         compared with, -1 if the literal is not in the dictionary      */
  ALN_RPN_CMP = 4,    /* arg = comparison code, fl.l:121-127: 1 <, 2 >, 3 !=, 4 ==, 5 <=, 6 >= */
  ALN_RPN_AND = 5,
  ALN_RPN_OR = 6,
  ALN_RPN_ADD = 7,
  ALN_RPN_MINUS = 8,
  ALN_RPN_MUL = 9,
  ALN_RPN_DIV = 10,
  ALN_RPN_CODESET = 11, /* LIKE operand (filter.cu:452-499): val = device address of a bit set
                          over the column's dictionary codes (bit c set when entry c matches
                          the pattern), arg = number of codes it covers; compared with the
                          column by CMP 7 */
  ALN_RPN_FLOAT = 12   /* emit_float (operators.cu:114-118): val = the IEEE-754 bits of a
                          double */
};
/* col_dec entry of an f64 (type 1, d_columns_float) column slot: its values are the bits of
 * doubles.  An operation with an f64 operand (an f64 column, a FLOAT, or a value computed
 * from one) is evaluated in double precision, as CudaSet's float overloads do
 * (cm.cu:2178-2437, 2440-2773): an int column is converted by a plain cast of its stored
 * integer (long_to_float_type, cm.cu:137-144), a NUMBER by its decimal value (digits /
 * 10^precision), arithmetic is IEEE double, comparisons use the EPSILON (1e-8) functors
 * (cm.h:17, cm.cu:81-134) with the left operand deeper in the stack.  A float expression
 * compiles to a program whose result holds double bits; its out_prec is ALN_DEC_F64. */
#define ALN_DEC_F64 (-1)
typedef struct {
  int32_t type;
  int32_t arg;
  int64_t val;
} aln_rpn;

/* ---- resolved device program (decimal scaling already folded in) -------------------- */
enum {
  ALN_I_COL = 1,   /* push col[a] * k                         */
  ALN_I_CONST = 2, /* push k                                  */
  ALN_I_ADD = 3,   /* r = pop, l = pop, push l + r           */
  ALN_I_SUB = 4,
  ALN_I_MUL = 5,
  ALN_I_DIV = 6, /* l / r, truncating; r == 0 yields 0        */
  ALN_I_CMP = 7, /* a = comparison code, push (l op r)        */
  ALN_I_AND = 8,
  ALN_I_OR = 9,
  ALN_I_INSET = 10, /* v = pop, push (0 <= v < a and bit v of the uint32 bit set at device
                      address k is set) */
  /* double precision (values are IEEE-754 bits in the int64 stack slots) */
  ALN_I_I2F = 11,  /* the entry a deep (0 = top, 1 = below it) from int64 to double */
  ALN_I_FADD = 12, /* r = pop, l = pop, push l + r (double) */
  ALN_I_FSUB = 13,
  ALN_I_FMUL = 14,
  ALN_I_FDIV = 15,
  ALN_I_FCMP = 16  /* a = comparison code, push the EPSILON comparison of l with r */
};
typedef struct {
  int32_t op;
  int32_t a;
  int64_t k;
} aln_insn;
#define ALN_MAX_INSN 64
#define ALN_MAX_STACK 16
#define ALN_MAX_COLS 16

/* Compile a reference RPN predicate (mode 0, filter.cu:133-643 semantics) or arithmetic
 * expression (mode 1, select.cu:703-790 semantics) into a device program.  col_dec holds
 * decimal_zeroes of every column slot.  *out_prec receives the result precision (mode 1). */
int aln_compile_rpn(const aln_rpn *rpn, int nrpn, const int32_t *col_dec, int ncols, int mode,
                    aln_insn *out, int *nout, int32_t *out_prec);

/* Zone-map classification of one segment (zone_map.cu:720-1135): 'A' all rows pass,
 * 'N' none, 'R' evaluate.  seg_min/seg_max: per column slot, decoded values. */
int aln_zone_map_check(const aln_rpn *rpn, int nrpn, const int64_t *seg_min,
                       const int64_t *seg_max, const int32_t *col_dec, int ncols,
                       int not_compressed, char *out);

/* ---- compressed column layout (the CudaSet columns of cm.h:681-789, resident in HBM) -- */
/* One segment of one column: frame-of-reference codes (value - base), `bits` wide, packed
 * LSB-first and densely into 32-bit words starting at words[word_off].  bits == 0 means
 * every value equals base; a plain int64 array is bits == 64, base == 0. */
typedef struct {
  uint64_t word_off;
  int64_t base;
  int64_t vmin; /* zone map: smallest decoded value of the segment */
  int64_t vmax; /* zone map: largest decoded value of the segment  */
  uint32_t bits;
  uint32_t nrows;
} aln_seg;

typedef struct {
  const uint32_t *words; /* device */
  const aln_seg *segs;   /* device, nsegs entries */
  uint32_t nsegs;
  int32_t decimals; /* decimal_zeroes (cm.cu:2941-2946) */
} aln_col;

/* Encoding: plan computes per-segment min/max/bits/word offsets (h_segs, nsegs entries,
 * nsegs = ceil(nrows/seg_rows)) and the total word count; pack writes the words and a
 * device copy of the descriptors.  force_bits > 0 forces that code width. */
int aln_encode_plan(aln_ctx *ctx, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                    uint32_t force_bits, aln_seg *h_segs, uint64_t *h_words);
int aln_encode_pack(aln_ctx *ctx, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                    const aln_seg *h_segs, uint32_t *d_words, aln_seg *d_segs);
/* Describe a plain device int64 array as a column (fills d_segs, nsegs entries). */
int aln_plain_col(aln_ctx *ctx, const int64_t *d_vals, uint64_t nrows, uint32_t seg_rows,
                  int32_t decimals, aln_seg *d_segs, aln_col *out);
/* Decode rows [row_begin, row_end) to int64. */
int aln_decode(aln_ctx *ctx, const aln_col *col, uint32_t seg_rows, uint64_t row_begin,
               uint64_t row_end, int64_t *d_out);

/* ---- the reference's segment files (compress.cu:180-527) --------------------------------
 * One column segment in the on-disk layout pfor_compress / pfor_delta_compress write:
 *   [u32 cnt][i64 lower][i64 upper][u32 comp_type][cnt payload bytes]
 *   [u32 cnt][u32 rec_count][u32 bits][i64 orig_lower][u32 fit_count][i64 start_val]
 * aln_segfile_decode (pfor_decompress, compress.cu:180-267) reads FOR (comp_type 0) and
 * FOR-DELTA (comp_type 1) files given as host bytes into int64 values on the device and
 * returns the header's zone map; aln_segfile_encode writes the FOR form of n device values
 * (h_file NULL: only *h_len, the file size).  Encode with aln_encode_* to keep the values
 * resident in this build's bit-packed layout. */
int aln_segfile_decode(aln_ctx *ctx, const void *h_file, uint64_t len, int64_t *d_out, uint64_t cap,
                       uint64_t *h_rows, int64_t *h_lower, int64_t *h_upper);
int aln_segfile_encode(aln_ctx *ctx, const int64_t *d_vals, uint64_t n, void *h_file, uint64_t cap,
                       uint64_t *h_len);

/* ---- scan: fused decode + predicate + stream compaction (filter_op, cm.cu:3609-3666) --
 * cols: host array of ncols column descriptors sharing one segmentation (seg_rows rows per
 * segment).  seg_modes: host array of nsegs zone-map results ('A','N','R'; NULL = all 'R').
 * Writes the ascending global row ids of passing rows. */
int aln_scan_filter(aln_ctx *ctx, const aln_col *cols, int ncols, uint32_t seg_rows,
                    const aln_insn *prog, int nprog, const char *seg_modes, uint32_t *d_row_ids,
                    uint64_t cap, uint64_t *h_count);

/* aln_scan_filter without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (work under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * aln_scan_filter, but its temporaries live in a fixed device array of the library (about
 * 3.3 MB per GPU, held until the process ends) instead of the context's workspace, and it
 * calls no hipMalloc, hipFree or other allocating routine.  Limit: the scan covers at most
 * ALN_SCAN_NOALLOC_MAX_TILES tiles of 4096 rows (nsegs * ceil(seg_rows / 4096) tiles, e.g.
 * 1.07e9 rows in full segments); a larger table returns ALN_E_ARG.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several host threads or streams)
 * race and must be serialised by the caller. */
#define ALN_SCAN_NOALLOC_MAX_TILES 262144
int aln_scan_filter_noalloc(aln_ctx *ctx, const aln_col *cols, int ncols, uint32_t seg_rows,
                            const aln_insn *prog, int nprog, const char *seg_modes,
                            uint32_t *d_row_ids, uint64_t cap, uint64_t *h_count);

/* Segment of the source row behind each of n indices (the order key of a join's pairs: the
 * reference joins left segment by left segment, operators.cu:1168-1240): d_out[i] = segment
 * of row r, r = d_rows ? d_rows[d_idx[i]] : d_idx[i] (d_rows: the selection a filtered view
 * reads through); the segment is the index of the last of nstarts ascending first rows
 * d_starts that is <= r (uneven file segments), or r / seg_rows when d_starts is NULL. */
int aln_row_segments(aln_ctx *ctx, const uint32_t *d_idx, uint64_t n, const uint32_t *d_rows,
                     const uint64_t *d_starts, uint32_t nstarts, uint32_t seg_rows, int64_t *d_out);

/* Gather decoded values of rows d_rows[0..n) (d_rows NULL = rows 0..n-1). */
int aln_gather(aln_ctx *ctx, const aln_col *col, uint32_t seg_rows, const uint32_t *d_rows,
               uint64_t n, int64_t *d_out);
/* Evaluate an expression program over rows d_rows[0..n) (NULL = 0..n-1). */
int aln_eval(aln_ctx *ctx, const aln_col *cols, int ncols, uint32_t seg_rows,
             const aln_insn *prog, int nprog, const uint32_t *d_rows, uint64_t n,
             int64_t *d_out);
/* d_out[i] = d_in[d_perm[i]] (apply_permutation, sorts.cu:86-125). */
int aln_apply_perm(aln_ctx *ctx, const int64_t *d_in, const uint32_t *d_perm, uint64_t n,
                   int64_t *d_out);
/* LIKE over a dictionary (gpu_regex over the dictionary file, filter.cu:67-129, 452-485):
 * d_entries holds n NUL-padded entries of width bytes (1..256); bit i of the uint32 bit set
 * d_bits is set when entry i matches h_pattern ('%' any run, '?' one character but '.').
 * d_bits must hold ceil(n / 64) * 2 words. */
int aln_like_dict(aln_ctx *ctx, const char *d_entries, uint64_t n, uint32_t width, const char *h_pattern,
                  uint32_t *d_bits);
/* d_out[i] = d_table[d_codes[i]] (-1 for a code outside 0..ntable-1).  ORDER BY on a char
 * column maps each row's dictionary code to the rank of its string in byte order and sorts
 * the ranks; it replaces the per-row string fetch and string sort of
 * update_char_permutation + str_sort (cm.cu:4218-4260, strings_sort_device.cu:125-137). */
int aln_dict_lookup(aln_ctx *ctx, const int64_t *d_codes, uint64_t n, const int64_t *d_table,
                    uint64_t ntable, int64_t *d_out);

/* ---- hash join (replaces the sort-merge join of operators.cu:1117-1284) -------------
 * Inner equi-join on int64 keys.  Pairs are written in probe-row order and, within one
 * probe row, in ascending build-row order. */
int aln_hash_join(aln_ctx *ctx, const int64_t *d_build_keys, uint64_t nbuild,
                  const int64_t *d_probe_keys, uint64_t nprobe, uint32_t *d_probe_idx,
                  uint32_t *d_build_idx, uint64_t cap, uint64_t *h_npairs);

/* Semi / anti reduction of a join (replaces sort + unique and set_difference in
 * operators.cu:1357-1399 and 1473-1480): d_idx holds the row indices (< nrows) one side
 * contributes to the join's pairs, duplicates allowed.  Writes the rows of that side that
 * occur in d_idx (anti == 0: LEFT/RIGHT SEMI) or that do not (anti != 0: LEFT/RIGHT ANTI),
 * each once, in ascending order.  *h_count receives the row count; ALN_E_CAPACITY when it
 * exceeds cap. */
int aln_join_rows(aln_ctx *ctx, const uint32_t *d_idx, uint64_t nidx, uint64_t nrows, int anti,
                  uint32_t *d_rows, uint64_t cap, uint64_t *h_count);

/* ---- group-by SUM/COUNT/MIN/MAX ------------------------------------------------------ */
/* F* aggregates take and return f64 bits (SUM in double precision, accumulated in no fixed
 * order: within 1e-12 relative of a Kahan sum for sums without cancellation). */
enum { ALN_AGG_COUNT = 0, ALN_AGG_SUM = 2, ALN_AGG_MIN = 4, ALN_AGG_MAX = 5, ALN_AGG_FSUM = 6,
       ALN_AGG_FMIN = 7, ALN_AGG_FMAX = 8 };
enum {
  ALN_ORDER_KEYS = 0, /* ascending key tuple: one-segment reference output (cm.cu:910-956) */
  ALN_ORDER_HASH = 1, /* ascending MurmurHash64A(keys, 8*nkeys, seed): multi-segment merge
                         order (merge.cu:138-203)                                         */
  ALN_ORDER_NONE = 2
};
#define ALN_MAX_KEYS 4
#define ALN_MAX_AGGS 8
/* d_keys/d_vals/d_out_*: host arrays of device pointers.  COUNT ignores its value column.
 * nkeys 0 (no GROUP BY): one group over all n rows, a reduction (reduce, select.cu:549). */
int aln_groupby(aln_ctx *ctx, const int64_t *const *d_keys, int nkeys,
                const int64_t *const *d_vals, const int32_t *aggs, int nvals, uint64_t n,
                int order, uint32_t seed, int64_t *const *d_out_keys,
                int64_t *const *d_out_vals, uint64_t cap, uint64_t *h_ngroups);

/* ---- ORDER BY: stable multi-key sort (emit_order, operators.cu:1641-1798) ------------
 * Writes the stable permutation sorting rows by keys[0] (most significant) .. keys[n-1],
 * desc[i] != 0 for DESC.  Ties keep input order. */
int aln_sort_keys(aln_ctx *ctx, const int64_t *const *d_keys, const int32_t *desc, int nkeys,
                  uint64_t n, uint32_t *d_perm);
/* First k rows of that stable order (ORDER BY ... LIMIT k). */
int aln_topk(aln_ctx *ctx, const int64_t *const *d_keys, const int32_t *desc, int nkeys,
             uint64_t n, uint32_t k, uint32_t *d_out_idx, uint32_t *h_nout);
/* Order-preserving int64 of each double (f64 bits in d_in): a < b as doubles iff
 * out(a) < out(b) as int64, with -0.0 taken as +0.0 (equal doubles, equal keys); NaNs
 * order after +inf (sign bit clear) or before -inf (set).  ORDER BY / GROUP BY on f64
 * columns (update_permutation on d_columns_float, sorts.cu:17-86) sort these keys;
 * aln_f64_unorder_keys maps them back to the doubles' bits. */
int aln_f64_order_keys(aln_ctx *ctx, const int64_t *d_in, uint64_t n, int64_t *d_out);
int aln_f64_unorder_keys(aln_ctx *ctx, const int64_t *d_in, uint64_t n, int64_t *d_out);
/* Stable radix sort of (u64 key, u32 value) pairs by the low `bits` bits of key. */
int aln_radix_sort_pairs(aln_ctx *ctx, uint64_t *d_keys, uint32_t *d_vals, uint64_t n,
                         uint32_t bits);

/* ---- hashing -------------------------------------------------------------------------- */
uint64_t aln_murmur64a(const void *key, int len, uint32_t seed);
/* d_out[i] = MurmurHash64A of the row's nkeys int64 values laid out contiguously. */
int aln_hash_rows(aln_ctx *ctx, const int64_t *const *d_keys, int nkeys, uint64_t n,
                  uint32_t seed, uint64_t *d_out);

/* ---- fused star pipeline: filter -> join -> join -> group-by -> top-k ----------------
 * The TPC-H q3 shape of the reference's readme (readme.md:48-80):
 *   dim  D: rows passing d_ranges; key d_key
 *   mid  M: rows passing m_ranges whose m_fk is a passing D key; unique key m_key;
 *           group columns m_g1, m_g2 (functionally dependent on m_key)
 *   fact F: rows passing f_ranges whose f_fk is a passing M key; aggregate per M row
 *           SUM((a0 + a1*f_x) * (b0 + b1*f_y)) and COUNT(*)
 * then ORDER BY sum DESC, g1 ASC, m_key ASC LIMIT k.  A range is lo <= value <= hi on the
 * decoded value of column slot `col` of its table; ranges of one table are ANDed. */
typedef struct {
  int32_t col;
  int32_t pad;
  int64_t lo;
  int64_t hi;
} aln_range;
#define ALN_MAX_RANGES 8
typedef struct {
  const aln_col *cols; /* host array */
  int32_t ncols;
  uint32_t seg_rows;
  const char *seg_modes; /* host, may be NULL */
  const aln_range *ranges;
  int32_t nranges;
} aln_table_scan;
typedef struct {
  aln_table_scan dim, mid, fact;
  int32_t d_key;
  int32_t m_key, m_fk, m_g1, m_g2;
  int32_t f_fk, f_x, f_y;
  int64_t a0, a1, b0, b1;
  uint32_t limit;
  uint32_t flags; /* bit 0: also return every group (d_groups);
                     bit 1: join mid -> fact through the MurmurHash64A open-addressing table
                     even when the mid key range is dense enough for the bitmap + rank index */
} aln_star_args;
typedef struct {
  int64_t key, g1, g2, sum, count;
} aln_star_row;
/* Row counts of each stage; the ms_* phase times are filled when profiling is enabled. */
typedef struct {
  uint64_t dim_pass;   /* dimension rows passing (bitmap inserts)           */
  uint64_t mid_probe;  /* mid rows passing the mid predicate (probes)       */
  uint64_t mid_pass;   /* mid rows passing and joined (hash-table entries)  */
  uint64_t fact_probe; /* fact rows passing the fact predicate (probes)     */
  uint64_t fact_pass;  /* fact rows joined and aggregated                   */
  uint64_t groups;     /* groups returned (flags bit 0)                     */
  double ms_dim, ms_mid, ms_fact, ms_topk;
} aln_star_stats;
/* Join structures: the dimension keys of passing rows set bits of a bitmap over the key
 * range; the passing mid keys go into a bitmap with a per-512-bit rank directory when the
 * mid key range is at most 2^31 (order-preserving: a group's index is the rank of its key),
 * else into a MurmurHash64A open-addressing table.  Duplicate keys on either build side
 * return ALN_E_UNSUPPORTED (the host layer then runs the unfused operators).
 * Groups (if flags & 1) are written to d_groups (cap rows) in ascending key order.
 * A call whose key columns (segment descriptor arrays and counts) are those of the
 * context's previous call reuses their zone-map key ranges instead of reading them back;
 * the device checks them against the zone maps and, if the columns changed in place, the
 * call runs again with fresh ranges (same answer, one more pass). */
int aln_star_agg(aln_ctx *ctx, const aln_star_args *args, aln_star_row *h_top, uint32_t *h_ntop,
                 aln_star_row *d_groups, uint64_t groups_cap, aln_star_stats *h_stats);

/* Several star pipelines over the same three tables (same column arrays and slots, no
 * seg_modes, flags 0) that differ in their ranges, coefficients and limit - TPC-H q3 with
 * different substitution parameters - answered with one scan of each table.  h_top[i]
 * receives query i's top rows (a[i].limit of them), h_ntop[i] their number, h_stats[i]
 * (optional) its counts.  Bitmap joins only: ALN_E_UNSUPPORTED when a key range needs the
 * hash table or a build side has duplicate keys (run the queries one by one then). */
#define ALN_STAR_BATCH_MAX 8
int aln_star_agg_batch(aln_ctx *ctx, const aln_star_args *a, int nq, aln_star_row *const *h_top,
                       uint32_t *h_ntop, aln_star_stats *h_stats);

/* ---- batched filter + GROUP BY (shared scan) -------------------------------------------
 * Up to ALN_BATCH_MAX queries over one table, each of the shape
 *   F := FILTER T BY pred >= lo AND pred <= hi;            (emit_filter, operators.cu:2306)
 *   R := SELECT key/key_div, SUM((a0 + a1*x) * (b0 + b1*y)), COUNT(*) FROM F GROUP BY ..;
 *                                                           (emit_select, operators.cu:1801)
 * answered by one scan of the table that reads every shared column once for all queries.
 * Values are the columns' stored integers (decimals at their column scale); key/key_div is
 * floor division; groups come back in key order, as one-segment GROUP BY returns them.
 * There is no single reference counterpart: the reference runs each query as its own
 * statements (a scan per query); aln_batch_* is the served form of those statements.
 * A batch holds device tables sized from the key columns' zone maps (dense over
 * [min/key_div, max/key_div], at most 2^31 groups per query) until aln_batch_destroy. */
#define ALN_BATCH_MAX 64
typedef struct {
  int32_t pred_col, key_col, x_col, y_col; /* slots in the cols array */
  int64_t lo, hi;                          /* inclusive bounds on pred */
  int64_t key_div;                         /* >= 1 */
  int64_t a0, a1, b0, b1;
} aln_batch_query;
/* Tuning: the queries are split into groups whose LDS tables (hot_groups slots of 12 bytes
 * per query) fit 48 KB; each block evaluates one query group over its share of the tiles. */
typedef struct {
  int32_t hot_groups;    /* groups [0, hot) of each query aggregated in LDS (0..4096; <0: 256) */
  int32_t blocks_per_cu; /* persistent grid size per compute unit (<=0: 2)                */
  int32_t use_graph;     /* replay the launch sequence from a captured HIP graph            */
} aln_batch_tuning;
typedef struct aln_batch aln_batch;
int aln_batch_create(aln_ctx *ctx, const aln_col *cols, int ncols, uint32_t seg_rows,
                     const aln_batch_query *qs, int nq, const aln_batch_tuning *tune,
                     aln_batch **out);
int aln_batch_set_tuning(aln_batch *b, const aln_batch_tuning *tune);
/* Enqueues one evaluation of every query on the context stream (no synchronisation). */
int aln_batch_run(aln_batch *b);
/* Groups of query q with at least one row (synchronises); ALN_E_CAPACITY if > cap. */
int aln_batch_results(aln_batch *b, int q, int64_t *d_keys, int64_t *d_sums, int64_t *d_counts,
                      uint64_t cap, uint64_t *h_ngroups);
/* The same into host arrays. */
int aln_batch_results_host(aln_batch *b, int q, int64_t *h_keys, int64_t *h_sums, int64_t *h_counts,
                           uint64_t cap, uint64_t *h_ngroups);
/* Times the candidate grid / hot-set shapes on the batch, applies and returns the fastest
 * (best->use_graph is kept as passed in). */
int aln_batch_autotune(aln_batch *b, int reps, aln_batch_tuning *best, double *h_ms);
void aln_batch_destroy(aln_batch *b);

/* ---- served form: request loop with continuous batching -------------------------------
 * No single reference counterpart: the reference answers one script at a time (bison.cu
 * main loop, main.cu:36-44 for its --QPS-test timing); this is the serving loop around the
 * same operators.  `streams` worker threads each own a context on their own HIP stream;
 * a worker that comes free takes the oldest waiting request and (unless `plain`) every
 * other waiting query on the same table that fits `max_batch` and the memory budget, and
 * answers them with one shared scan (aln_batch_*).  Requests whose device memory would
 * exceed what the budget has free wait until running work releases it; a request larger
 * than the whole budget fails with ALN_E_NOMEM.  In `plain` mode each query runs alone
 * through aln_scan_filter -> aln_eval -> aln_groupby.  Tables (device columns) must stay
 * alive while the server runs.  All entry points are thread-safe. */
typedef struct {
  int32_t streams;   /* concurrent workers (1..8)                            */
  int32_t max_batch; /* queries per shared scan (<= ALN_BATCH_MAX; <=0: max) */
  int32_t plain;     /* 1: no co-batching, operators one query at a time     */
  int32_t reserved;
  uint64_t budget_bytes; /* device bytes admitted requests may hold together  */
  aln_batch_tuning tune; /* batch kernel shape (blocks_per_cu <= 0: 1 here)   */
} aln_server_config;
typedef struct {
  int32_t status; /* ALN_OK or the request's error code */
  int32_t kind;   /* 0 query, 1 star pipeline           */
  int32_t batch_size;
  int32_t reserved;
  uint64_t batch; /* id of the batch that answered it   */
  uint64_t ngroups; /* groups (query) or top rows (star) */
  double queue_ms, run_ms, latency_ms; /* submit -> start -> done */
} aln_server_result;
typedef struct {
  uint64_t submitted, completed, batches, batched_queries, max_concurrent, pending, mem_in_use;
} aln_server_totals;
typedef struct aln_server aln_server;
int aln_server_create(int device, const aln_server_config *cfg, aln_server **out);
/* Registers a table (the column descriptors are copied; the device data is not). */
int aln_server_add_table(aln_server *s, const aln_col *cols, int ncols, uint32_t seg_rows,
                         int32_t *h_table);
int aln_server_submit(aln_server *s, int32_t table, const aln_batch_query *q, uint64_t *h_id);
int aln_server_submit_star(aln_server *s, const aln_star_args *args, uint64_t *h_id);
/* Blocks until the request is done; returns its status (message in aln_last_error). */
int aln_server_wait(aln_server *s, uint64_t id, aln_server_result *h_res);
int aln_server_fetch(aln_server *s, uint64_t id, int64_t *h_keys, int64_t *h_sums,
                     int64_t *h_counts, uint64_t cap);
int aln_server_fetch_star(aln_server *s, uint64_t id, aln_star_row *h_top, uint32_t cap,
                          uint32_t *h_ntop);
/* Forgets a finished request (waits for it first). */
int aln_server_release(aln_server *s, uint64_t id);
int aln_server_stats(aln_server *s, aln_server_totals *h_stats);
/* Finishes every submitted request, then stops the workers. */
void aln_server_destroy(aln_server *s);

/* ---- synthetic workloads (no reference counterpart: the bench's seeded inputs) --------
 * Counter-based SplitMix64 streams; see datagen.hip for the exact specification.  All
 * output pointers are device int64 arrays and may be NULL to skip a column. */
int aln_gen_flat(aln_ctx *ctx, uint64_t seed, uint64_t n, int64_t *d_key, int64_t *d_date,
                 int64_t *d_price, int64_t *d_disc);
int aln_gen_dim(aln_ctx *ctx, uint64_t seed, uint64_t n, int64_t *d_key, int64_t *d_seg_idx);
int aln_gen_orders(aln_ctx *ctx, uint64_t seed, uint64_t n, uint64_t ndim, int64_t *d_key,
                   int64_t *d_cust, int64_t *d_date, int64_t *d_prio);
int aln_gen_fact_rows(aln_ctx *ctx, uint64_t seed, uint64_t norders, uint64_t *h_nfact);
int aln_gen_fact(aln_ctx *ctx, uint64_t seed, uint64_t norders, int64_t *d_okey,
                 int64_t *d_price, int64_t *d_disc, int64_t *d_ship);
/* Zipf(s) ranks in [0, nkeys) (rank 0 most frequent), by rejection-inversion. */
int aln_gen_zipf(aln_ctx *ctx, uint64_t seed, uint64_t n, uint64_t nkeys, double s, int64_t *d_out);
/* Uniform 64-bit keys and, when d_payload is not NULL, payload[i] = i. */
int aln_gen_sort_input(aln_ctx *ctx, uint64_t seed, uint64_t n, int64_t *d_keys, uint32_t *d_payload);
/* Dictionary codes in first-seen order (compress_char, cm.cu:1808-1821) for values that
 * index a string table [0, nvals): h_code_of_val[v] = code or -1; d_codes (optional) gets
 * the code of every row. */
int aln_dict_first_seen(aln_ctx *ctx, const int64_t *d_vals, uint64_t n, int64_t nvals,
                        int64_t *h_code_of_val, int64_t *d_codes);

#ifdef __cplusplus
}
#endif
#endif /* ALENKA_HIP_H */

/*
 * alenka_ops.h - the reference's operator entry points (operators.h) over HBM-resident
 * tables, as a C ABI.
 *
 * A statement is issued the way the reference's parser issues it: the expression and
 * select-list calls (emit_name, emit_number, emit_cmp, emit_sum, emit_sel_name, ...) fill
 * the RPN queues (operators.cu:68-326), then one statement call consumes them:
 *
 *   emit_filter  operators.h:50  (operators.cu:2306-2408)  lazy: records the predicate
 *   emit_select  operators.h:55  (operators.cu:1801-2084)  projection / GROUP BY
 *   emit_join    operators.h:56  (operators.cu:797-928)    inner equi-join chain
 *   emit_order   operators.h:53  (operators.cu:1641-1798)  stable multi-key ORDER BY
 *   emit_store   operators.h:48  (operators.cu:2410-2443)  text output with LIMIT
 *   emit_display operators.h:59  (operators.cu:2446-2479)  text output to stdout
 *   emit_store_binary / emit_load_binary (operators.cu:2446-2562)  the segment files
 *
 * Errors: the reference reports through process_error (operators.cu:2752-2765), which
 * exits the process for severity > 1.  Here process_error records the message, every
 * statement call leaves its status in alenka_status(), and nothing exits.
 *
 * Tables are registered from host or device arrays, or read from the reference's segment
 * files (emit_load_binary operators.cu:2518-2562, and implicitly for every table the loaded
 * data dictionary lists, as load_vars operators.cu:2708-2747 does).  Int/decimal columns are
 * stored frame-of-reference +
 * bit-packed in segments of seg_rows rows with min/max zone maps; char(n) columns are
 * dictionary-encoded in first-seen order (compress_char, cm.cu:1765-1844).
 */
#ifndef ALENKA_OPS_H
#define ALENKA_OPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- session --------------------------------------------------------------------------- */
int alenka_init(int device);             /* alenkaInit (alenka.h:30) without the script  */
int alenka_close(void);                  /* frees every variable and the device context  */
int alenka_status(void);                 /* status of the last statement (ALN_OK = 0)     */
const char *alenka_last_error(void);
void alenka_set_hash_seed(uint32_t seed); /* hash_seed, 100 in the file driver (bison.y:304) */
/* process_count: rows per segment that emit_store_binary writes (-l, bison.y:299-308; 10^9) */
void alenka_set_process_count(uint64_t rows);
/* load_col_data / save_col_data (cm.cu:4057-4114): read or write the data dictionary file
 * (table -> column -> type, length) the segment files are described by; the reference reads
 * "data.dictionary" at start-up (bison.y:333) and writes it after a binary STORE. */
int alenka_load_col_data(const char *file);
int alenka_save_col_data(const char *file);

/* ---- tables ---------------------------------------------------------------------------- */
int alenka_table_create(const char *name, uint64_t nrows, uint32_t seg_rows);
/* int/decimal column from host int64 values (decimals = scale); compress != 0 stores it
 * FOR + bit-packed, else as plain int64. */
int alenka_table_add_int(const char *table, const char *col, const int64_t *h_vals,
                         int32_t decimals, int compress);
/* the same from a device int64 array (copied) */
int alenka_table_add_int_device(const char *table, const char *col, const int64_t *d_vals,
                                int32_t decimals, int compress);
/* f64 (type 1, d_columns_float) column from host doubles, kept uncompressed as the
 * reference stores non-decimal floats (comp_type 3, cm.cu:1150-1180) */
int alenka_table_add_float(const char *table, const char *col, const double *h_vals);
/* char(width) column from nrows fixed-width, NUL-padded host strings */
int alenka_table_add_char(const char *table, const char *col, const char *h_chars,
                          uint32_t width);
int alenka_rows(const char *var, uint64_t *h_rows);
int alenka_column_count(const char *var, int *h_n);
/* name of column i (into buf of len bytes), its type (0 int/decimal, 2 char) and scale */
int alenka_column_info(const char *var, int i, char *buf, int len, int *h_type, int *h_decimals);
/* decoded values of an int/decimal column (or dictionary codes of a char column) */
int alenka_fetch_int(const char *var, const char *col, int64_t *h_out, uint64_t cap);
/* fixed-width strings of a char column, width bytes per row */
int alenka_fetch_f64(const char *var, const char *col, double *h_out, uint64_t cap);
int alenka_fetch_char(const char *var, const char *col, char *h_out, uint64_t cap_rows,
                      uint32_t width);
int alenka_drop(const char *var);

/* ---- RPN queue emitters (operators.cu:68-326) -------------------------------------------- */
void emit_name(const char *name);
void emit_number(int64_t val);
void emit_decimal(const char *str);
void emit_float(double val); /* operators.h:39: an APPROXNUM (1.5E3) literal */
void emit_string(const char *str);
void emit_cmp(int val);
void emit_and(void);
void emit_or(void);
void emit_add(void);
void emit_minus(void);
void emit_mul(void);
void emit_div(void);
void emit_eq(void);
void emit_count(void);
void emit_sum(void);
void emit_average(void);
void emit_min(void);
void emit_max(void);
void emit_sel_name(const char *name);
void emit_var_asc(const char *name);
void emit_var_desc(const char *name);
void emit_join_tab(const char *name, char tp);
void emit_limit(int val);
void clean_queues(void);
void process_error(int severity, const char *err);

/* ---- statements ------------------------------------------------------------------------ */
void emit_filter(char *s, char *f);
void emit_select(const char *s, const char *f, int grp_cnt);
void emit_join(const char *s, const char *j1, int grp, int start_seg, int end_seg);
void emit_order(const char *s, const char *f, int e, int ll);
void emit_store(const char *s, const char *f, const char *sep);
void emit_display(const char *s, const char *sep);
/* Deferred statements: like emit_filter (operators.cu:2306-2408), a JOIN of three tables in
 * the star shape (fact JOIN mid ON f = m_key JOIN dim ON m_fk = d_key, inner, one column pair
 * each), a GROUP BY over such a join and an ORDER BY of that GROUP BY record their work and
 * run it when a consumer first reads the variable.  STORE / DISPLAY ... LIMIT k of such an
 * ORDER BY, whose keys give ORDER BY SUM DESC, <group column> ASC, runs the whole chain as
 * one fused pipeline (aln_star_agg's top-k); any other consumer runs the statements one by
 * one, a GROUP BY over the join through aln_star_agg's group output when its shape allows,
 * and the unfused operators otherwise (or when the fused kernels refuse the data, e.g.
 * duplicate build keys).  Results are the same either way.  alenka_last_path() names the
 * path the last deferred chain took: "aln_star_agg top-k", "aln_star_agg groups" or
 * "unfused" (no reference counterpart: a diagnostic). */
const char *alenka_last_path(void);
/* emit_store_binary (operators.h:49, operators.cu:2446-2516): the variable's columns as the
 * reference's segment files <f>.<col>.<seg> (char: dictionary <f>.<col>, .idx and .hash per
 * segment) with a .header per column, and its entry in the data dictionary (save it with
 * alenka_save_col_data).  append != 0 is ALN_E_UNSUPPORTED. */
void emit_store_binary(const char *s, const char *f, int append);
/* emit_load_binary (operators.h:47, operators.cu:2518-2562): variable s from the segment files
 * with prefix f, described by the data dictionary's entry for f; .sort/.presort are read. */
void emit_load_binary(const char *s, const char *f, int d);

/* A script in the reference's statement language (alenkaExecute, alenka.h:31; grammar
 * bison.y:127-287): FILTER, SELECT [JOIN ... ON] [GROUP BY], ORDER, STORE ... USING,
 * DISPLAY and DROP TABLE over the registered tables and variables, each statement turned
 * into the emit_* calls the reference's parser makes.  Returns ALN_OK, or the status of
 * the first statement that failed (its message in alenka_last_error); other statements
 * (LOAD, binary STORE, outer joins, ...) return ALN_E_UNSUPPORTED. */
int alenkaExecute(char *s);

#ifdef __cplusplus
}
#endif
#endif /* ALENKA_OPS_H */

// The reference's operator layer (operators.h / operators.cu) over HBM-resident tables:
// the RPN queue emitters and the statement calls emit_filter / emit_select / emit_join /
// emit_order / emit_store / emit_display, each running on the C-ABI kernels of
// alenka_hip.h.  Variables live in a varNames catalog (cm.h:791) of CudaSet-shaped sets:
// named columns, each either FOR + bit-packed in segments with zone maps (loaded tables)
// or a plain int64 device array (operator results, `not_compressed`).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/alenka_hip.h"
#include "../../include/alenka_ops.h"

namespace {

// ---- device memory -----------------------------------------------------------------------
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};
using DBufP = std::shared_ptr<DBuf>;

DBufP dalloc(size_t bytes) {
  auto b = std::make_shared<DBuf>();
  if (bytes == 0) bytes = 8;
  if (hipMalloc(&b->p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  b->bytes = bytes;
  return b;
}

struct Col {
  int type = 0;       // 0 int/decimal, 1 f64 (the doubles' bits, plain), 2 char (dictionary code)
  int decimals = 0;   // decimal_zeroes
  DBufP words;        // packed words, or the int64 values of a plain column
  DBufP dsegs;        // device aln_seg[nsegs]
  std::vector<aln_seg> hsegs;
  std::shared_ptr<std::vector<std::string>> dict;  // char: code -> string
  uint32_t width = 0;
  aln_col desc() const {
    aln_col c;
    c.words = (const uint32_t *)words->p;
    c.segs = (const aln_seg *)dsegs->p;
    c.nsegs = (uint32_t)hsegs.size();
    c.decimals = decimals;
    return c;
  }
};

// The compiler's col_dec entry of a column: its decimal scale, or ALN_DEC_F64 for doubles.
inline int32_t col_dec(const Col &c) { return c.type == 1 ? ALN_DEC_F64 : c.decimals; }

struct Tok {
  std::string type;  // NAME NUMBER STRING CMP AND OR ADD MINUS MUL DIV JOIN COUNT SUM AVG MIN MAX SEL ORDER
  std::string value;
  int64_t num = 0;
  uint32_t prec = 0;
};

struct Set {
  std::vector<std::string> names;
  std::map<std::string, Col> cols;
  uint64_t nrows = 0;
  uint32_t seg_rows = 1;
  bool not_compressed = false;
  bool filtered = false;
  std::string source;
  std::vector<Tok> pred;
  // cached selection of a filtered set
  DBufP sel;  // uint32 row ids of the source
  uint64_t nsel = 0;
  bool sel_ready = false;
  // a table read from the reference's segment files keeps those files' segment boundaries
  // (first row of each segment, then nrows; empty: every segment has seg_rows rows), which
  // decide the group-by merge order and the join's pair order
  std::vector<uint64_t> seg_starts;
  std::vector<std::string> sorted_fields, presorted_fields;  // the .sort / .presort lists
  // a statement whose work waits for a consumer (JOIN, GROUP BY over a join, ORDER BY of that
  // GROUP BY); until it runs, the set has no columns
  std::shared_ptr<struct Deferred> deferred;
  uint32_t nsegs() const { return (uint32_t)((nrows + seg_rows - 1) / seg_rows); }
  uint32_t file_segs() const { return seg_starts.empty() ? nsegs() : (uint32_t)(seg_starts.size() - 1); }
  int64_t seg_of(uint64_t row) const {
    if (seg_starts.empty()) return (int64_t)(row / seg_rows);
    return (int64_t)(std::upper_bound(seg_starts.begin(), seg_starts.end(), row) - seg_starts.begin()) - 1;
  }
};

// ---- deferred statements ---------------------------------------------------------------------
// The reference defers a FILTER until a consumer reads it (emit_filter, operators.cu:2306-2408).
// A JOIN of three tables in the star shape, a GROUP BY over such a join and an ORDER BY of that
// GROUP BY are deferred here the same way, so that a STORE/DISPLAY ... LIMIT k of the ORDER BY
// can run the whole chain as one fused pipeline (aln_star_agg); every other consumer runs the
// statements one by one, as the reference does, when it first reads the set.
struct StarTable {
  std::shared_ptr<Set> base;      // the stored table the scan reads
  std::vector<std::string> cols;  // its column slots
  std::vector<aln_range> ranges;  // the table's filter as ANDed closed ranges
  int slot(const std::string &n) {
    auto it = std::find(cols.begin(), cols.end(), n);
    if (it != cols.end()) return (int)(it - cols.begin());
    cols.push_back(n);
    return (int)cols.size() - 1;
  }
};
struct StarOut {  // one output column of the GROUP BY, in select-list order
  std::string alias;
  int field = 0;  // 0 key, 1 g1, 2 g2, 3 sum, 4 count (aln_star_row order)
  int decimals = 0, type = 0;
  std::shared_ptr<std::vector<std::string>> dict;
  uint32_t width = 0;
};
struct StarPlan {
  StarTable dim, mid, fact;
  int d_key = 0, m_key = 0, m_fk = 0, m_g1 = 0, m_g2 = 0, f_fk = 0, f_x = 0, f_y = 0;
  int64_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
  std::vector<StarOut> out;
};
struct Deferred {
  enum { JOIN = 1, STAR_GROUP = 2, ORDER = 3 };
  int kind = 0;
  std::vector<Tok> tok;            // the statement's queue
  std::vector<std::string> join;   // JOIN: op_join, innermost first
  std::vector<char> join_kind;
  std::string src;                 // JOIN: the left variable (j1); else the FROM variable
  int grp_cnt = 0;                 // STAR_GROUP
  // the variables the statement reads, as they were when it was issued (a later statement
  // may assign their names again before this one runs)
  std::vector<std::pair<std::string, std::shared_ptr<Set>>> pins;
  std::shared_ptr<StarPlan> star;  // STAR_GROUP, and an ORDER BY of one
  std::vector<std::pair<std::string, int>> order_keys;  // ORDER: (column, desc)
};

// ---- session state (the reference's globals, cm.h:66-117, operators.cu:20-50) -------------
aln_ctx *g_ctx = nullptr;
int g_device = 0;
std::map<std::string, std::shared_ptr<Set>> varNames;
std::vector<Tok> g_tok;
std::vector<std::string> g_join;  // op_join
std::vector<char> g_join_kind;    // join_type: 'I' inner, '1'/'2' left/right semi, '3'/'4' anti
int g_limit = 0;
bool g_has_limit = false;
bool g_join_bad = false;  // an unsupported join kind was queued for this statement
uint32_t g_hash_seed = 100;
int g_status = ALN_OK;
char g_err[512] = "";

// When a deferred statement fails inside the statement that consumes it, its error is the
// consumer's error (not the "couldn't find" that follows from it).
int g_deferred_rc = ALN_OK;
char g_deferred_err[512] = "";
// how the last deferred chain ran (alenka_last_path)
std::string g_path = "";

int fail(int code, const char *fmt, ...) {
  if (g_deferred_rc != ALN_OK) {
    snprintf(g_err, sizeof(g_err), "%s", g_deferred_err);
    g_status = g_deferred_rc;
    return g_deferred_rc;
  }
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_status = code;
  return code;
}
int fail_aln(int rc, const char *what) { return fail(rc, "%s: %s", what, aln_last_error()); }
#define TRY(x)                    \
  do {                            \
    int _r = (x);                 \
    if (_r != ALN_OK) return _r;  \
  } while (0)
#define ALN(x, what)                                   \
  do {                                                 \
    int _r = (x);                                      \
    if (_r != ALN_OK) return fail_aln(_r, what);       \
  } while (0)

int ensure_ctx() {
  if (g_ctx) return ALN_OK;
  int rc = aln_create(g_device, nullptr, &g_ctx);
  if (rc != ALN_OK) return fail_aln(rc, "alenka_init");
  return ALN_OK;
}

// data_dict (cm.h, load_col_data / save_col_data cm.cu:4057-4114): table -> column ->
// {col_type 0 int/decimal/timestamp, 1 float, 2 char; col_length decimal digits, UINT_MAX for a
// timestamp, or the char width}
struct ColData {
  uint32_t type = 0, length = 0;
};
std::map<std::string, std::map<std::string, ColData>> g_data_dict;
int load_binary(const std::string &name, const std::string &file);

// A variable by name; a table that is not a variable yet but is in data_dict is read from its
// segment files first, as load_vars (operators.cu:2708-2747) loads every table a script uses.
std::shared_ptr<Set> peek_var(const std::string &n) {
  auto it = varNames.find(n);
  if (it != varNames.end()) return it->second;
  if (!n.empty() && g_data_dict.count(n) && ensure_ctx() == ALN_OK && load_binary(n, n) == ALN_OK)
    return varNames[n];
  return nullptr;
}
int run_deferred(const std::shared_ptr<Set> &s);
bool defer_join(const char *s, const char *j1);
bool defer_star_group(const char *s, const char *f, int grp_cnt, const std::shared_ptr<Set> &src);
bool defer_order(const char *s, const char *f, const std::shared_ptr<Set> &src);
int store_star_topk(const std::shared_ptr<Set> &v, FILE *fp, const char *sep, int limit, bool *done);
// A variable by name with its deferred statement (if any) run first: every consumer but the
// fused pipeline's reads sets this way.
std::shared_ptr<Set> find_var(const std::string &n) {
  auto s = peek_var(n);
  if (s && s->deferred && run_deferred(s) != ALN_OK) return nullptr;
  return s;
}

// ---- column construction ---------------------------------------------------------------
int make_plain(const DBufP &vals, uint64_t n, uint32_t seg_rows, int decimals, Col *out) {
  uint32_t nsegs = (uint32_t)((n + seg_rows - 1) / seg_rows);
  DBufP ds = dalloc(sizeof(aln_seg) * (nsegs ? nsegs : 1));
  if (!ds) return fail(ALN_E_NOMEM, "device allocation failed");
  aln_col c;
  ALN(aln_plain_col(g_ctx, (const int64_t *)vals->p, n, seg_rows, decimals, (aln_seg *)ds->p, &c),
      "aln_plain_col");
  out->words = vals;
  out->dsegs = ds;
  out->decimals = decimals;
  out->hsegs.resize(nsegs);
  for (uint32_t s = 0; s < nsegs; ++s) {
    uint64_t r0 = (uint64_t)s * seg_rows;
    aln_seg &g = out->hsegs[s];
    g.word_off = 2 * r0;
    g.base = 0;
    g.bits = 64;
    g.nrows = (uint32_t)(r0 + seg_rows <= n ? seg_rows : n - r0);
    g.vmin = INT64_MIN;
    g.vmax = INT64_MAX;
  }
  return ALN_OK;
}

int make_packed(const int64_t *d_vals, uint64_t n, uint32_t seg_rows, int decimals, Col *out) {
  uint32_t nsegs = (uint32_t)((n + seg_rows - 1) / seg_rows);
  out->hsegs.resize(nsegs);
  if (nsegs == 0) {  // an empty column: no segments to encode
    out->words = dalloc(8);
    out->dsegs = dalloc(sizeof(aln_seg));
    if (!out->words || !out->dsegs) return fail(ALN_E_NOMEM, "device allocation failed");
    out->decimals = decimals;
    return ALN_OK;
  }
  uint64_t words = 0;
  ALN(aln_encode_plan(g_ctx, d_vals, n, seg_rows, 0, out->hsegs.data(), &words), "aln_encode_plan");
  out->words = dalloc(words * 4);
  out->dsegs = dalloc(sizeof(aln_seg) * (nsegs ? nsegs : 1));
  if (!out->words || !out->dsegs) return fail(ALN_E_NOMEM, "device allocation failed");
  ALN(aln_encode_pack(g_ctx, d_vals, n, seg_rows, out->hsegs.data(), (uint32_t *)out->words->p,
                      (aln_seg *)out->dsegs->p),
      "aln_encode_pack");
  out->decimals = decimals;
  return ALN_OK;
}

// ---- views: a set's columns, resolved through a lazy filter ------------------------------
struct View {
  std::map<std::string, Col> cols;  // columns usable with seg_rows
  uint64_t nrows = 0;
  uint32_t seg_rows = 1;
  uint32_t cycles = 1;  // the reference's segment loop count (group-by merge order)
  DBufP rowids;         // uint32: source row id of each row, NULL = row i (join ordering)
};

// Bit sets of the dictionary codes a LIKE pattern matches, cached per dictionary and
// pattern (the reference keeps the matching codes per column in map_res, filter.cu:452-499).
// The dictionary goes to the device once, as width-byte NUL-padded entries (the layout of
// the reference's dictionary file), and aln_like_dict matches every entry there.
struct LikeCache {
  std::weak_ptr<std::vector<std::string>> dict;
  DBufP entries;
  uint32_t width = 0;
  std::map<std::string, DBufP> sets;
};
std::map<const void *, LikeCache> g_like;

int like_codes(const Col &c, const std::string &pattern, aln_rpn *r) {
  const auto &d = *c.dict;
  if (d.size() > 0x7FFFFFFFu) return fail(ALN_E_UNSUPPORTED, "dictionary too large for LIKE");
  if (c.width < 1 || c.width > 256) return fail(ALN_E_UNSUPPORTED, "LIKE on a char(%u) column", c.width);
  auto &e = g_like[(const void *)c.dict.get()];
  if (e.dict.expired() || e.dict.lock() != c.dict || e.width != c.width) {
    e.dict = c.dict;
    e.width = c.width;
    e.sets.clear();
    std::vector<char> h(d.size() * c.width + 1, 0);
    for (size_t i = 0; i < d.size(); ++i)
      memcpy(h.data() + i * c.width, d[i].data(), std::min<size_t>(d[i].size(), c.width));
    e.entries = dalloc(h.size());
    if (!e.entries) return fail(ALN_E_NOMEM, "device allocation failed");
    if (hipMemcpy(e.entries->p, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(ALN_E_HIP, "copy of a dictionary failed");
  }
  DBufP &set = e.sets[pattern];
  if (!set) {
    set = dalloc(4 * ((d.size() + 63) / 64 * 2 + 1));
    if (!set) return fail(ALN_E_NOMEM, "device allocation failed");
    ALN(aln_like_dict(g_ctx, (const char *)e.entries->p, d.size(), c.width, pattern.c_str(), (uint32_t *)set->p),
        "LIKE");
  }
  r->type = ALN_RPN_CODESET;
  r->arg = (int32_t)d.size();
  r->val = (int64_t)(uintptr_t)set->p;
  return ALN_OK;
}

int resolve_string_codes(const Set &src, const std::vector<Tok> &t, std::vector<aln_rpn> *rpn,
                         const std::map<std::string, int> &slot) {
  for (size_t i = 0; i < t.size(); ++i) {
    aln_rpn r;
    memset(&r, 0, sizeof(r));
    const Tok &k = t[i];
    if (k.type == "STRING" && i > 0 && t[i - 1].type == "NAME" && i + 1 < t.size() &&
        t[i + 1].type == "CMP" && t[i + 1].num == 7) {
      // column LIKE pattern: the pattern becomes the set of matching dictionary codes
      auto it = src.cols.find(t[i - 1].value);
      if (it == src.cols.end() || !it->second.dict)
        return fail(ALN_E_UNSUPPORTED, "LIKE on %s, which is not a char column", t[i - 1].value.c_str());
      TRY(like_codes(it->second, k.value, &r));
    } else if (k.type == "NAME") {
      r.type = ALN_RPN_NAME;
      r.arg = slot.at(k.value);
    } else if (k.type == "NUMBER") {
      r.type = ALN_RPN_NUMBER;
      r.val = k.num;
      r.arg = (int32_t)k.prec;
    } else if (k.type == "FLOAT") {
      r.type = ALN_RPN_FLOAT;
      r.val = k.num;  // the double's bits
    } else if (k.type == "STRING") {
      r.type = ALN_RPN_STRING;
      r.val = -1;
      // the literal is compared with the column named next to it (NAME STRING CMP or
      // STRING NAME CMP); it becomes that column's dictionary code
      std::string colname;
      if (i > 0 && t[i - 1].type == "NAME") colname = t[i - 1].value;
      if (i + 1 < t.size() && t[i + 1].type == "NAME") colname = t[i + 1].value;
      auto it = src.cols.find(colname);
      if (it != src.cols.end() && it->second.dict) {
        auto &d = *it->second.dict;
        auto f = std::find(d.begin(), d.end(), k.value);
        if (f != d.end()) r.val = (int64_t)(f - d.begin());
      }
    } else if (k.type == "CMP") {
      r.type = ALN_RPN_CMP;
      r.arg = (int32_t)k.num;
    } else if (k.type == "AND") {
      r.type = ALN_RPN_AND;
    } else if (k.type == "OR") {
      r.type = ALN_RPN_OR;
    } else if (k.type == "ADD") {
      r.type = ALN_RPN_ADD;
    } else if (k.type == "MINUS") {
      r.type = ALN_RPN_MINUS;
    } else if (k.type == "MUL") {
      r.type = ALN_RPN_MUL;
    } else if (k.type == "DIV") {
      r.type = ALN_RPN_DIV;
    } else {
      return fail(ALN_E_UNSUPPORTED, "operation %s is not valid here", k.type.c_str());
    }
    rpn->push_back(r);
  }
  return ALN_OK;
}

// filter_op (cm.cu:3609-3666): zone maps per segment, then the fused scan.
int run_filter(Set &f) {
  if (f.sel_ready) return ALN_OK;
  auto src = find_var(f.source);
  if (!src) return fail(ALN_E_ARG, "Couldn't find %s", f.source.c_str());
  std::map<std::string, int> slot;
  std::vector<std::string> names;
  for (auto &k : f.pred)
    if (k.type == "NAME" && !slot.count(k.value)) {
      if (!src->cols.count(k.value)) return fail(ALN_E_ARG, "Couldn't find column %s", k.value.c_str());
      slot[k.value] = (int)names.size();
      names.push_back(k.value);
    }
  if (names.empty()) {  // constant predicate: evaluate against any column
    if (src->names.empty()) return fail(ALN_E_ARG, "filter on an empty set");
    names.push_back(src->names[0]);
  }
  if (names.size() > ALN_MAX_COLS) return fail(ALN_E_UNSUPPORTED, "predicate reads too many columns");
  std::vector<aln_rpn> rpn;
  TRY(resolve_string_codes(*src, f.pred, &rpn, slot));
  std::vector<aln_col> cols;
  std::vector<int32_t> dec;
  for (auto &n : names) {
    cols.push_back(src->cols[n].desc());
    dec.push_back(col_dec(src->cols[n]));
  }
  aln_insn prog[ALN_MAX_INSN];
  int nprog = 0;
  int32_t prec = 0;
  ALN(aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 0, prog, &nprog, &prec),
      "filter");
  uint32_t nsegs = src->nsegs();
  std::vector<char> modes(nsegs ? nsegs : 1, 'R');
  std::vector<int64_t> mn(names.size()), mx(names.size());
  for (uint32_t s = 0; s < nsegs; ++s) {
    for (size_t c = 0; c < names.size(); ++c) {
      mn[c] = src->cols[names[c]].hsegs[s].vmin;
      mx[c] = src->cols[names[c]].hsegs[s].vmax;
    }
    ALN(aln_zone_map_check(rpn.data(), (int)rpn.size(), mn.data(), mx.data(), dec.data(),
                           (int)dec.size(), src->not_compressed, &modes[s]),
        "zone_map_check");
  }
  f.sel = dalloc(4 * (src->nrows ? src->nrows : 1));
  if (!f.sel) return fail(ALN_E_NOMEM, "device allocation failed");
  uint64_t cnt = 0;
  ALN(aln_scan_filter(g_ctx, cols.data(), (int)cols.size(), src->seg_rows, prog, nprog, modes.data(),
                      (uint32_t *)f.sel->p, src->nrows, &cnt),
      "filter");
  f.nsel = cnt;
  f.sel_ready = true;
  return ALN_OK;
}

// Columns `need` of variable `name`, decoded through its filter when it has one.
int make_view(const std::string &name, const std::set<std::string> &need, View *v) {
  auto a = find_var(name);
  if (!a) return fail(ALN_E_ARG, "Couldn't find %s", name.c_str());
  if (!a->filtered) {
    for (auto &n : need) {
      if (!a->cols.count(n)) return fail(ALN_E_ARG, "Couldn't find column %s", n.c_str());
      v->cols[n] = a->cols[n];
    }
    v->nrows = a->nrows;
    v->seg_rows = a->seg_rows;
    v->cycles = a->not_compressed ? 1 : a->file_segs();
    v->rowids = nullptr;
    return ALN_OK;
  }
  TRY(run_filter(*a));
  auto src = find_var(a->source);
  const uint64_t k = a->nsel;
  v->nrows = k;
  v->seg_rows = (uint32_t)(k ? (k < 0xFFFFFFFFull ? k : 0xFFFFFFFFull) : 1);
  v->cycles = src->not_compressed ? 1 : src->file_segs();
  for (auto &n : need) {
    if (!src->cols.count(n)) return fail(ALN_E_ARG, "Couldn't find column %s", n.c_str());
    const Col &sc = src->cols[n];
    DBufP vals = dalloc(8 * (k ? k : 1));
    if (!vals) return fail(ALN_E_NOMEM, "device allocation failed");
    aln_col d = sc.desc();
    ALN(aln_gather(g_ctx, &d, src->seg_rows, (const uint32_t *)a->sel->p, k, (int64_t *)vals->p),
        "gather");
    Col c;
    TRY(make_plain(vals, k, v->seg_rows, sc.decimals, &c));
    c.type = sc.type;
    c.dict = sc.dict;
    c.width = sc.width;
    v->cols[n] = c;
  }
  // the source row of each row (join output order needs the left segment)
  v->rowids = a->sel;
  return ALN_OK;
}

int fetch_col(const Col &c, uint64_t n, uint32_t seg_rows, std::vector<int64_t> *out) {
  out->resize(n);
  if (!n) return ALN_OK;
  DBufP tmp = dalloc(8 * n);
  if (!tmp) return fail(ALN_E_NOMEM, "device allocation failed");
  aln_col d = c.desc();
  ALN(aln_decode(g_ctx, &d, seg_rows, 0, n, (int64_t *)tmp->p), "decode");
  ALN(aln_sync(g_ctx), "sync");
  if (hipMemcpy(out->data(), tmp->p, 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(ALN_E_HIP, "device to host copy failed");
  return ALN_OK;
}

// A plain int64 copy of a column (decoded).
int plain_copy(const Col &c, uint64_t n, uint32_t seg_rows, DBufP *out) {
  *out = dalloc(8 * (n ? n : 1));
  if (!*out) return fail(ALN_E_NOMEM, "device allocation failed");
  if (!n) return ALN_OK;
  aln_col d = c.desc();
  ALN(aln_decode(g_ctx, &d, seg_rows, 0, n, (int64_t *)(*out)->p), "decode");
  return ALN_OK;
}

// ---- select-list parsing --------------------------------------------------------------------
struct SelItem {
  std::vector<Tok> expr;
  std::string agg;  // "" (NULL), COUNT, SUM, AVG, MIN, MAX
  std::string alias;
};

void parse_select(const std::vector<Tok> &t, std::vector<SelItem> *items, std::vector<Tok> *rest) {
  SelItem cur;
  for (auto &k : t) {
    if (k.type == "SEL") {
      cur.alias = k.value;
      items->push_back(cur);
      cur = SelItem();
    } else if (k.type == "COUNT" || k.type == "SUM" || k.type == "AVG" || k.type == "MIN" ||
               k.type == "MAX") {
      cur.agg = k.type;
    } else {
      cur.expr.push_back(k);
    }
  }
  *rest = cur.expr;  // trailing tokens: GROUP BY names, join conditions
}

int eval_expr(const View &v, const std::vector<Tok> &expr, DBufP *out, int *prec) {
  std::map<std::string, int> slot;
  std::vector<std::string> names;
  for (auto &k : expr)
    if (k.type == "NAME" && !slot.count(k.value)) {
      if (!v.cols.count(k.value)) return fail(ALN_E_ARG, "Couldn't find column %s", k.value.c_str());
      slot[k.value] = (int)names.size();
      names.push_back(k.value);
    }
  Set dummy;
  for (auto &n : names) dummy.cols[n] = v.cols.at(n);
  std::vector<aln_rpn> rpn;
  TRY(resolve_string_codes(dummy, expr, &rpn, slot));
  std::vector<aln_col> cols;
  std::vector<int32_t> dec;
  for (auto &n : names) {
    cols.push_back(v.cols.at(n).desc());
    dec.push_back(col_dec(v.cols.at(n)));
  }
  *out = dalloc(8 * (v.nrows ? v.nrows : 1));
  if (!*out) return fail(ALN_E_NOMEM, "device allocation failed");
  if (names.empty()) {  // a constant
    if (rpn.size() != 1 || (rpn[0].type != ALN_RPN_NUMBER && rpn[0].type != ALN_RPN_FLOAT))
      return fail(ALN_E_UNSUPPORTED, "expression without columns");
    std::vector<int64_t> h(v.nrows, rpn[0].val);
    *prec = rpn[0].type == ALN_RPN_FLOAT ? ALN_DEC_F64 : rpn[0].arg;
    if (v.nrows && hipMemcpy((*out)->p, h.data(), 8 * v.nrows, hipMemcpyHostToDevice) != hipSuccess)
      return fail(ALN_E_HIP, "host to device copy failed");
    return ALN_OK;
  }
  aln_insn prog[ALN_MAX_INSN];
  int nprog = 0;
  int32_t p = 0;
  ALN(aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 1, prog, &nprog, &p),
      "select expression");
  ALN(aln_eval(g_ctx, cols.data(), (int)cols.size(), v.seg_rows, prog, nprog, nullptr, v.nrows,
               (int64_t *)(*out)->p),
      "select expression");
  *prec = p;
  return ALN_OK;
}

std::set<std::string> names_in(const std::vector<Tok> &t) {
  std::set<std::string> s;
  for (auto &k : t)
    if (k.type == "NAME") s.insert(k.value);
  return s;
}

void finish_statement() {
  g_deferred_rc = ALN_OK;
  g_join_bad = false;
  g_tok.clear();
  g_join.clear();
  g_join_kind.clear();
  g_has_limit = false;
  g_limit = 0;
}

}  // namespace

// =============================================================================================
extern "C" {

int alenka_init(int device) {
  g_device = device;
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ALN_E_HIP, "alenka_init: device %d", device);
  }
  g_status = ALN_OK;
  return ensure_ctx();
}

int alenka_close(void) {
  varNames.clear();
  g_like.clear();
  finish_statement();
  if (g_ctx) aln_destroy(g_ctx);
  g_ctx = nullptr;
  return ALN_OK;
}

int alenka_status(void) { return g_status; }
const char *alenka_last_error(void) { return g_err; }
void alenka_set_hash_seed(uint32_t seed) { g_hash_seed = seed; }

int alenka_table_create(const char *name, uint64_t nrows, uint32_t seg_rows) {
  g_status = ALN_OK;
  TRY(ensure_ctx());
  if (!name || !seg_rows) return fail(ALN_E_ARG, "alenka_table_create: bad arguments");
  auto s = std::make_shared<Set>();
  s->nrows = nrows;
  s->seg_rows = seg_rows;
  varNames[name] = s;
  return ALN_OK;
}

static int add_int(const char *table, const char *col, const int64_t *vals, bool device,
                   int32_t decimals, int compress) {
  g_status = ALN_OK;
  TRY(ensure_ctx());
  auto s = find_var(table ? table : "");
  if (!s || !col) return fail(ALN_E_ARG, "alenka_table_add_int: unknown table");
  DBufP d = dalloc(8 * (s->nrows ? s->nrows : 1));
  if (!d) return fail(ALN_E_NOMEM, "device allocation failed");
  if (s->nrows && hipMemcpy(d->p, vals, 8 * s->nrows, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice) != hipSuccess)
    return fail(ALN_E_HIP, "copy of column %s failed", col);
  Col c;
  if (compress)
    TRY(make_packed((const int64_t *)d->p, s->nrows, s->seg_rows, decimals, &c));
  else
    TRY(make_plain(d, s->nrows, s->seg_rows, decimals, &c));
  if (!s->cols.count(col)) s->names.push_back(col);
  s->cols[col] = c;
  return ALN_OK;
}

int alenka_table_add_int(const char *table, const char *col, const int64_t *h_vals, int32_t decimals,
                         int compress) {
  return add_int(table, col, h_vals, false, decimals, compress);
}
int alenka_table_add_int_device(const char *table, const char *col, const int64_t *d_vals,
                                int32_t decimals, int compress) {
  return add_int(table, col, d_vals, true, decimals, compress);
}

int alenka_table_add_float(const char *table, const char *col, const double *h_vals) {
  int rc = add_int(table, col, reinterpret_cast<const int64_t *>(h_vals), false, 0, 0);
  if (rc == ALN_OK) find_var(table)->cols[col].type = 1;
  return rc;
}

int alenka_table_add_char(const char *table, const char *col, const char *h_chars, uint32_t width) {
  g_status = ALN_OK;
  TRY(ensure_ctx());
  auto s = find_var(table ? table : "");
  if (!s || !col || !width) return fail(ALN_E_ARG, "alenka_table_add_char: bad arguments");
  // dictionary in first-seen order (compress_char, cm.cu:1808-1821)
  auto dict = std::make_shared<std::vector<std::string>>();
  std::map<std::string, int64_t> code_of;
  std::vector<int64_t> codes(s->nrows);
  for (uint64_t i = 0; i < s->nrows; ++i) {
    const char *p = h_chars + i * width;
    std::string v(p, strnlen(p, width));
    auto it = code_of.find(v);
    if (it == code_of.end()) {
      it = code_of.emplace(v, (int64_t)dict->size()).first;
      dict->push_back(v);
    }
    codes[i] = it->second;
  }
  TRY(add_int(table, col, codes.data(), false, 0, 1));
  s->cols[col].type = 2;
  s->cols[col].dict = dict;
  s->cols[col].width = width;
  return ALN_OK;
}

int alenka_rows(const char *var, uint64_t *h_rows) {
  auto s = find_var(var ? var : "");
  if (!s || !h_rows) return fail(ALN_E_ARG, "alenka_rows: unknown variable");
  if (s->filtered) {
    TRY(run_filter(*s));
    *h_rows = s->nsel;
  } else {
    *h_rows = s->nrows;
  }
  return ALN_OK;
}

int alenka_column_count(const char *var, int *h_n) {
  auto s = find_var(var ? var : "");
  if (!s || !h_n) return fail(ALN_E_ARG, "alenka_column_count: unknown variable");
  auto base = s->filtered ? find_var(s->source) : s;
  *h_n = (int)base->names.size();
  return ALN_OK;
}

int alenka_column_info(const char *var, int i, char *buf, int len, int *h_type, int *h_decimals) {
  auto s = find_var(var ? var : "");
  if (!s) return fail(ALN_E_ARG, "alenka_column_info: unknown variable");
  auto base = s->filtered ? find_var(s->source) : s;
  if (i < 0 || i >= (int)base->names.size()) return fail(ALN_E_ARG, "alenka_column_info: index");
  const std::string &n = base->names[i];
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s", n.c_str());
  if (h_type) *h_type = base->cols[n].type;
  if (h_decimals) *h_decimals = base->cols[n].decimals;
  return ALN_OK;
}

int alenka_fetch_int(const char *var, const char *col, int64_t *h_out, uint64_t cap) {
  g_status = ALN_OK;
  View v;
  TRY(make_view(var ? var : "", {col ? col : ""}, &v));
  if (v.nrows > cap) return fail(ALN_E_CAPACITY, "alenka_fetch_int: %llu rows", (unsigned long long)v.nrows);
  std::vector<int64_t> h;
  TRY(fetch_col(v.cols[col], v.nrows, v.seg_rows, &h));
  if (!h.empty()) memcpy(h_out, h.data(), 8 * h.size());
  return ALN_OK;
}

int alenka_fetch_f64(const char *var, const char *col, double *h_out, uint64_t cap) {
  g_status = ALN_OK;
  View v;
  TRY(make_view(var ? var : "", {col ? col : ""}, &v));
  if (v.cols[col].type != 1) return fail(ALN_E_ARG, "alenka_fetch_f64: %s is not a float column", col);
  if (v.nrows > cap) return fail(ALN_E_CAPACITY, "alenka_fetch_f64: %llu rows", (unsigned long long)v.nrows);
  std::vector<int64_t> h;
  TRY(fetch_col(v.cols[col], v.nrows, v.seg_rows, &h));
  if (!h.empty()) memcpy(h_out, h.data(), 8 * h.size());
  return ALN_OK;
}

int alenka_fetch_char(const char *var, const char *col, char *h_out, uint64_t cap_rows, uint32_t width) {
  g_status = ALN_OK;
  View v;
  TRY(make_view(var ? var : "", {col ? col : ""}, &v));
  const Col &c = v.cols[col];
  if (!c.dict) return fail(ALN_E_ARG, "alenka_fetch_char: %s is not a char column", col);
  if (v.nrows > cap_rows) return fail(ALN_E_CAPACITY, "alenka_fetch_char: %llu rows", (unsigned long long)v.nrows);
  std::vector<int64_t> h;
  TRY(fetch_col(c, v.nrows, v.seg_rows, &h));
  memset(h_out, 0, (size_t)v.nrows * width);
  for (uint64_t i = 0; i < v.nrows; ++i) {
    const std::string &sv = (*c.dict)[(size_t)h[i]];
    memcpy(h_out + i * width, sv.data(), std::min<size_t>(sv.size(), width));
  }
  return ALN_OK;
}

int alenka_drop(const char *var) {
  varNames.erase(var ? var : "");
  return ALN_OK;
}

// ---- RPN queue emitters ---------------------------------------------------------------------
static void push(const char *type, const std::string &value = "", int64_t num = 0, uint32_t prec = 0) {
  Tok t;
  t.type = type;
  t.value = value;
  t.num = num;
  t.prec = prec;
  g_tok.push_back(t);
}

void emit_name(const char *name) { push("NAME", name); }
void emit_number(int64_t val) { push("NUMBER", "", val, 0); }
// emit_decimal (operators.cu:120-134): digits without the point, precision = digits after it
void emit_decimal(const char *str) {
  std::string s(str);
  uint32_t precision = 0;
  auto pos = s.find('.');
  if (pos != std::string::npos) {
    precision = (uint32_t)(s.length() - pos - 1);
    s.erase(pos, 1);
  }
  push("NUMBER", "", (int64_t)std::stoll(s), precision);
}
// emit_string (operators.cu:82-94): surrounding double quotes are removed
void emit_string(const char *str) {
  std::string s(str);
  if (!s.empty() && s[0] == '"') s = s.substr(1, s.size() >= 2 ? s.size() - 2 : 0);
  push("STRING", s);
}
// emit_float (operators.cu:114-118): the double travels as its bits
void emit_float(double val) {
  int64_t b;
  memcpy(&b, &val, 8);
  push("FLOAT", "", b);
}
void emit_cmp(int val) { push("CMP", "", val); }
void emit_and(void) { push("AND"); }
void emit_or(void) { push("OR"); }
void emit_add(void) { push("ADD"); }
void emit_minus(void) { push("MINUS"); }
void emit_mul(void) { push("MUL"); }
void emit_div(void) { push("DIV"); }
void emit_eq(void) { push("JOIN"); }
void emit_count(void) { push("COUNT"); }
void emit_sum(void) { push("SUM"); }
void emit_average(void) { push("AVG"); }
void emit_min(void) { push("MIN"); }
void emit_max(void) { push("MAX"); }
void emit_sel_name(const char *name) { push("SEL", name); }
void emit_var_asc(const char *name) { push("ORDER", name, 0); }
void emit_var_desc(const char *name) { push("ORDER", name, 1); }
void emit_join_tab(const char *name, char tp) {
  if (tp != 'I' && tp != '1' && tp != '2' && tp != '3' && tp != '4') g_join_bad = true;
  g_join.push_back(name);
  g_join_kind.push_back(tp);
}
void emit_limit(int val) {
  g_limit = val;
  g_has_limit = true;
}
void clean_queues(void) { finish_statement(); }
// process_error (operators.cu:2752-2765): recorded; nothing exits
void process_error(int severity, const char *err) {
  fail(severity > 1 ? ALN_E_ARG : ALN_OK, "%s", err);
}

// ---- emit_filter (operators.cu:2306-2408) ---------------------------------------------------
void emit_filter(char *s, char *f) {
  g_status = ALN_OK;
  auto a = find_var(f ? f : "");
  if (!a) {
    fail(ALN_E_ARG, "Filter : couldn't find variable %s", f ? f : "");
    finish_statement();
    return;
  }
  auto b = std::make_shared<Set>();
  b->filtered = true;
  b->pred = g_tok;
  if (a->filtered) {  // AND-chain with the source filter (operators.cu:2359-2392)
    b->source = a->source;
    b->pred.insert(b->pred.end(), a->pred.begin(), a->pred.end());
    Tok t;
    t.type = "AND";
    b->pred.push_back(t);
  } else {
    b->source = f;
  }
  varNames[s] = b;
  finish_statement();
}

// ---- emit_select (operators.cu:1801-2084, select.cu:359-952, merge.cu) -----------------------
static int do_select(const char *s, const char *f, int grp_cnt) {
  std::vector<SelItem> items;
  std::vector<Tok> rest;
  parse_select(g_tok, &items, &rest);
  if (items.empty()) return fail(ALN_E_ARG, "Select : no output columns");
  std::vector<std::string> groups;
  for (auto &k : rest)
    if (k.type == "NAME") groups.push_back(k.value);
  if ((int)groups.size() < grp_cnt) return fail(ALN_E_ARG, "Select : GROUP BY list");
  groups.erase(groups.begin(), groups.end() - grp_cnt);
  std::set<std::string> need(groups.begin(), groups.end());
  for (auto &it : items)
    for (auto &n : names_in(it.expr)) need.insert(n);
  View v;
  TRY(make_view(f, need, &v));
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  const bool any_agg = std::any_of(items.begin(), items.end(), [](const SelItem &i) { return !i.agg.empty(); });

  if (grp_cnt == 0 && !any_agg) {  // projection
    for (auto &it : items) {
      DBufP vals;
      int prec = 0;
      TRY(eval_expr(v, it.expr, &vals, &prec));
      Col c;
      TRY(make_plain(vals, v.nrows, (uint32_t)(v.nrows ? v.nrows : 1), prec == ALN_DEC_F64 ? 0 : prec, &c));
      if (prec == ALN_DEC_F64) c.type = 1;
      if (it.expr.size() == 1 && it.expr[0].type == "NAME") {
        const Col &srcc = v.cols[it.expr[0].value];
        c.type = srcc.type;
        c.dict = srcc.dict;
        c.width = srcc.width;
        c.decimals = srcc.decimals;
      }
      out->names.push_back(it.alias);
      out->cols[it.alias] = c;
    }
    out->nrows = v.nrows;
    out->seg_rows = (uint32_t)(v.nrows ? v.nrows : 1);
    varNames[s] = out;
    return ALN_OK;
  }

  // one source segment: key order (order_inplace + GroupBy); several: the hash-ordered
  // merge of add()/count_avg() (merge.cu:92-496)
  const int order = (grp_cnt > 0 && v.cycles > 1) ? ALN_ORDER_HASH : ALN_ORDER_KEYS;
  // keys: the GROUP BY columns (none: one output row, a reduction).  f64
  // keys group by value: in key order they are sorted as doubles (order-preserving keys,
  // mapped back afterwards), in hash order the hash covers the doubles' 8 bytes as the
  // reference's merge does (merge.cu:131-133).
  std::vector<DBufP> keybufs;
  std::vector<const int64_t *> keys;
  std::vector<int> keydec;
  std::vector<bool> keyf64;
  if (grp_cnt > 0) {
    if (grp_cnt > ALN_MAX_KEYS) return fail(ALN_E_UNSUPPORTED, "GROUP BY on more than %d columns", ALN_MAX_KEYS);
    for (auto &g : groups) {
      DBufP d;
      TRY(plain_copy(v.cols[g], v.nrows, v.seg_rows, &d));
      const bool f64 = v.cols[g].type == 1;
      if (f64 && order == ALN_ORDER_KEYS && v.nrows)
        ALN(aln_f64_order_keys(g_ctx, (const int64_t *)d->p, v.nrows, (int64_t *)d->p), "group by");
      keybufs.push_back(d);
      keys.push_back((const int64_t *)d->p);
      keydec.push_back(v.cols[g].decimals);
      keyf64.push_back(f64);
    }
  }
  // aggregate inputs
  struct Out {
    int kind;      // 0 key, 1 aggregate, 2 AVG (sum / count)
    int key = -1;  // key index
    int agg = -1, agg2 = -1;
    int prec = 0;
    bool f64 = false;  // the value is a double
  };
  std::vector<Out> outs;
  std::vector<DBufP> valbufs;
  std::vector<const int64_t *> vals;
  std::vector<int32_t> aggs;
  for (auto &it : items) {
    Out o;
    if (it.agg.empty()) {
      if (it.expr.size() != 1 || it.expr[0].type != "NAME")
        return fail(ALN_E_UNSUPPORTED, "Select : non-aggregate expression %s in a GROUP BY", it.alias.c_str());
      auto g = std::find(groups.begin(), groups.end(), it.expr[0].value);
      if (g == groups.end())
        return fail(ALN_E_UNSUPPORTED, "Select : %s is neither grouped nor aggregated", it.expr[0].value.c_str());
      o.kind = 0;
      o.key = (int)(g - groups.begin());
      o.prec = keydec[o.key];
      o.f64 = keyf64[o.key];
    } else {
      DBufP vb;
      int prec = 0;
      TRY(eval_expr(v, it.expr, &vb, &prec));
      valbufs.push_back(vb);
      const bool f64 = prec == ALN_DEC_F64 && it.agg != "COUNT";
      o.f64 = f64;
      o.prec = it.agg == "COUNT" || f64 ? 0 : prec;
      int code = it.agg == "COUNT" ? ALN_AGG_COUNT
               : it.agg == "MIN"   ? (f64 ? ALN_AGG_FMIN : ALN_AGG_MIN)
               : it.agg == "MAX"   ? (f64 ? ALN_AGG_FMAX : ALN_AGG_MAX)
                                   : (f64 ? ALN_AGG_FSUM : ALN_AGG_SUM);
      o.kind = it.agg == "AVG" ? 2 : 1;
      o.agg = (int)aggs.size();
      vals.push_back((const int64_t *)vb->p);
      aggs.push_back(code);
      if (o.kind == 2) {
        o.agg2 = (int)aggs.size();
        vals.push_back((const int64_t *)vb->p);
        aggs.push_back(ALN_AGG_COUNT);
      }
    }
    outs.push_back(o);
  }
  if (aggs.size() > ALN_MAX_AGGS) return fail(ALN_E_UNSUPPORTED, "more than %d aggregates", ALN_MAX_AGGS);
  const uint64_t cap = v.nrows ? v.nrows : 1;
  std::vector<DBufP> okb, ovb;
  std::vector<int64_t *> ok, ov;
  for (size_t i = 0; i < keys.size(); ++i) {
    okb.push_back(dalloc(8 * cap));
    ok.push_back((int64_t *)okb.back()->p);
  }
  for (size_t i = 0; i < aggs.size(); ++i) {
    ovb.push_back(dalloc(8 * cap));
    ov.push_back((int64_t *)ovb.back()->p);
  }
  uint64_t ng = 0;
  if (v.nrows) {
    ALN(aln_groupby(g_ctx, keys.data(), (int)keys.size(), vals.data(), aggs.data(), (int)aggs.size(),
                    v.nrows, order, g_hash_seed, ok.data(), ov.data(), cap, &ng),
        "group by");
  }
  if (grp_cnt == 0 && ng == 0) {
    // aggregates over no rows: the reference returns an empty set (emit_select, no c_set)
    varNames[s] = out;
    return ALN_OK;
  }
  for (size_t i = 0; i < items.size(); ++i) {
    const Out &o = outs[i];
    Col c;
    if (o.kind == 0) {
      if (o.f64 && order == ALN_ORDER_KEYS && ng)
        ALN(aln_f64_unorder_keys(g_ctx, okb[o.key]->p ? (const int64_t *)okb[o.key]->p : nullptr, ng,
                                 (int64_t *)okb[o.key]->p),
            "group by");
      TRY(make_plain(okb[o.key], ng, (uint32_t)(ng ? ng : 1), o.prec, &c));
      const Col &srcc = v.cols[groups[o.key]];
      c.type = srcc.type;
      c.dict = srcc.dict;
      c.width = srcc.width;
    } else if (o.kind == 1) {
      TRY(make_plain(ovb[o.agg], ng, (uint32_t)(ng ? ng : 1), o.prec, &c));
      if (o.f64) c.type = 1;
    } else if (o.f64) {
      // AVG of doubles: sum / count (float_avg, merge.cu:31-36)
      std::vector<int64_t> sum(ng), cnt(ng);
      if (hipMemcpy(sum.data(), ovb[o.agg]->p, 8 * ng, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(cnt.data(), ovb[o.agg2]->p, 8 * ng, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ALN_E_HIP, "device to host copy failed");
      for (uint64_t g = 0; g < ng; ++g) {
        double x;
        memcpy(&x, &sum[g], 8);
        x /= (double)cnt[g];
        memcpy(&sum[g], &x, 8);
      }
      DBufP d = dalloc(8 * (ng ? ng : 1));
      if (!d) return fail(ALN_E_NOMEM, "device allocation failed");
      if (ng && hipMemcpy(d->p, sum.data(), 8 * ng, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "host to device copy failed");
      TRY(make_plain(d, ng, (uint32_t)(ng ? ng : 1), 0, &c));
      c.type = 1;
    } else {
      // count_avg (merge.cu:446-460): scale 2 or less gains two digits, (sum*100)/count
      std::vector<int64_t> sum(ng), cnt(ng);
      if (hipMemcpy(sum.data(), ovb[o.agg]->p, 8 * ng, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(cnt.data(), ovb[o.agg2]->p, 8 * ng, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ALN_E_HIP, "device to host copy failed");
      int prec = o.prec;
      for (uint64_t g = 0; g < ng; ++g) sum[g] = prec <= 2 ? (sum[g] * 100) / cnt[g] : sum[g] / cnt[g];
      if (prec <= 2) prec += 2;
      DBufP d = dalloc(8 * (ng ? ng : 1));
      if (ng && hipMemcpy(d->p, sum.data(), 8 * ng, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "host to device copy failed");
      TRY(make_plain(d, ng, (uint32_t)(ng ? ng : 1), prec, &c));
    }
    out->names.push_back(items[i].alias);
    out->cols[items[i].alias] = c;
  }
  out->nrows = ng;
  out->seg_rows = (uint32_t)(ng ? ng : 1);
  varNames[s] = out;
  return ALN_OK;
}

void emit_select(const char *s, const char *f, int grp_cnt) {
  g_status = ALN_OK;
  auto src = peek_var(f ? f : "");
  if (!src) {
    fail(ALN_E_ARG, "Couldn't find(1) %s", f ? f : "");
  } else if (!(grp_cnt > 0 && src->deferred && src->deferred->kind == Deferred::JOIN &&
               defer_star_group(s, f, grp_cnt, src))) {
    if (find_var(f)) do_select(s, f, grp_cnt);
  }
  finish_statement();
}

// ---- emit_join (operators.cu:797-928, emit_multijoin 940-1582) -------------------------------
struct Cond {
  std::vector<std::pair<std::string, std::string>> eqs;  // a = b, in order
};

// Splits the condition tokens into one condition per join, in parse order, counting them as
// emit_and / emit_eq do (operators.cu:153-172): the ANDs seen since the last condition give
// the number of extra column pairs of the next one, whose equalities then close it.  '=' has
// the grammar's lowest precedence (bison.y:42-49), so `ON a = b AND c = d` arrives as
// a b c AND d = =, and a condition's names pair up in text order: (a, b), (c, d).
static int parse_conditions(const std::vector<Tok> &t, std::vector<Cond> *out) {
  std::vector<std::string> names;
  std::vector<int> extra;  // extra column pairs of each condition
  int and_cnt = 0, misses = 0;
  for (auto &k : t) {
    if (k.type == "NAME") {
      names.push_back(k.value);
    } else if (k.type == "AND") {
      ++and_cnt;
    } else if (k.type == "JOIN") {
      if (misses == 0) {
        extra.push_back(and_cnt);
        misses = and_cnt;
        and_cnt = 0;
      } else {
        --misses;
      }
    } else {
      return fail(ALN_E_UNSUPPORTED, "Join : operation %s in a join condition", k.type.c_str());
    }
  }
  if (and_cnt || misses) return fail(ALN_E_ARG, "Join : bad condition");
  size_t p = 0;
  for (int x : extra) {
    // the reference compares the third and fourth names for every extra pair
    // (operators.cu:1288-1296), so it checks only one extra equality
    if (x > 1) return fail(ALN_E_UNSUPPORTED, "Join : more than two column pairs in one condition");
    Cond c;
    for (int q = 0; q <= x; ++q, p += 2) {
      if (p + 2 > names.size()) return fail(ALN_E_ARG, "Join : bad condition");
      c.eqs.push_back({names[p], names[p + 1]});
    }
    out->push_back(c);
  }
  if (p != names.size()) return fail(ALN_E_ARG, "Join : bad condition");
  return ALN_OK;
}

static bool has_col(const std::string &var, const std::string &col) {
  auto s = find_var(var);
  if (!s) return false;
  auto base = s->filtered ? find_var(s->source) : s;
  return base && base->cols.count(col);
}

static std::vector<std::string> all_cols(const std::string &var) {
  auto s = find_var(var);
  auto base = s->filtered ? find_var(s->source) : s;
  return base->names;
}

static int semi_anti(const std::string &name, char kind, const View &lv, const View &rv,
                     const std::vector<std::pair<std::string, std::string>> &eqs, DBufP pi, DBufP bi,
                     uint64_t np, const std::set<std::string> &lneed, const std::set<std::string> &rneed);

// left JOIN right ON cond -> result `name` with the columns of both sides in `keep`; for the
// semi and anti kinds, the columns of the side that is kept.
static int join_pair(const std::string &name, const std::string &left, const std::string &right,
                     const Cond &cond, const std::set<std::string> &keep, char kind) {
  // orient each equality: the left column belongs to the left variable
  std::vector<std::pair<std::string, std::string>> eqs;
  for (auto &e : cond.eqs) {
    if (has_col(left, e.first) && has_col(right, e.second))
      eqs.push_back(e);
    else if (has_col(left, e.second) && has_col(right, e.first))
      eqs.push_back({e.second, e.first});
    else
      return fail(ALN_E_ARG, "Join : couldn't find %s / %s", e.first.c_str(), e.second.c_str());
  }
  std::set<std::string> lneed, rneed;
  for (auto &c : all_cols(left))
    if (keep.count(c)) lneed.insert(c);
  for (auto &c : all_cols(right))
    if (keep.count(c) && !lneed.count(c)) rneed.insert(c);
  for (auto &e : eqs) {
    lneed.insert(e.first);
    rneed.insert(e.second);
  }
  View lv, rv;
  TRY(make_view(left, lneed, &lv));
  TRY(make_view(right, rneed, &rv));
  for (auto &e : eqs) {
    if (lv.cols[e.first].type == 2 || rv.cols[e.second].type == 2)
      return fail(ALN_E_UNSUPPORTED, "Join : string join keys (operators.cu:1030-1033)");
    // the reference joins float keys only when both are decimal floats (operators.cu:1031),
    // which are ints here
    if (lv.cols[e.first].type == 1 || rv.cols[e.second].type == 1)
      return fail(ALN_E_UNSUPPORTED, "Join : float join keys (operators.cu:1030-1033)");
  }
  DBufP lk, rk;
  TRY(plain_copy(lv.cols[eqs[0].first], lv.nrows, lv.seg_rows, &lk));
  TRY(plain_copy(rv.cols[eqs[0].second], rv.nrows, rv.seg_rows, &rk));
  uint64_t cap = lv.nrows ? lv.nrows : 1, np = 0;
  DBufP pi, bi;
  for (;;) {
    pi = dalloc(4 * cap);
    bi = dalloc(4 * cap);
    if (!pi || !bi) return fail(ALN_E_NOMEM, "device allocation failed");
    int rc = aln_hash_join(g_ctx, (const int64_t *)rk->p, rv.nrows, (const int64_t *)lk->p, lv.nrows,
                           (uint32_t *)pi->p, (uint32_t *)bi->p, cap, &np);
    if (rc == ALN_E_CAPACITY && np > cap) {
      cap = np;
      continue;
    }
    ALN(rc, "join");
    break;
  }
  if (kind != 'I') return semi_anti(name, kind, lv, rv, eqs, pi, bi, np, lneed, rneed);
  // gather every kept column through the pairs
  auto gather_all = [&](const View &v, const std::set<std::string> &cols, const DBufP &idx, uint64_t n,
                        std::map<std::string, Col> *dst) -> int {
    for (auto &cn : cols) {
      const Col &sc = v.cols.at(cn);
      DBufP vals = dalloc(8 * (n ? n : 1));
      if (!vals) return fail(ALN_E_NOMEM, "device allocation failed");
      aln_col d = sc.desc();
      if (n) ALN(aln_gather(g_ctx, &d, v.seg_rows, (const uint32_t *)idx->p, n, (int64_t *)vals->p), "join gather");
      Col c;
      TRY(make_plain(vals, n, (uint32_t)(n ? n : 1), sc.decimals, &c));
      c.type = sc.type;
      c.dict = sc.dict;
      c.width = sc.width;
      (*dst)[cn] = c;
    }
    return ALN_OK;
  };
  std::map<std::string, Col> lc, rc;
  std::set<std::string> lall = lneed, rall = rneed;
  TRY(gather_all(lv, lall, pi, np, &lc));
  TRY(gather_all(rv, rall, bi, np, &rc));
  // pair order of the reference: per left segment, ascending left key, then left row, then
  // right row (operators.cu:1184-1240).  Pairs arrive in left-row order; sort them stably
  // by (left segment, left key).
  uint64_t n = np;
  std::vector<uint32_t> perm;
  if (n > 1) {
    // the left source segment of each pair, on the device
    auto lsrc = find_var(left);
    auto lbase = lsrc->filtered ? find_var(lsrc->source) : lsrc;
    DBufP dseg = dalloc(8 * n);
    if (!dseg) return fail(ALN_E_NOMEM, "device allocation failed");
    if (lbase->not_compressed) {  // an operator result is one segment
      ALN(aln_row_segments(g_ctx, (const uint32_t *)pi->p, n, nullptr, nullptr, 0, 0xFFFFFFFFu,
                           (int64_t *)dseg->p),
          "join order");
    } else {
      DBufP starts;
      if (!lbase->seg_starts.empty()) {
        starts = dalloc(8 * lbase->seg_starts.size());
        if (!starts) return fail(ALN_E_NOMEM, "device allocation failed");
        if (hipMemcpy(starts->p, lbase->seg_starts.data(), 8 * lbase->seg_starts.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
          return fail(ALN_E_HIP, "host to device copy failed");
      }
      ALN(aln_row_segments(g_ctx, (const uint32_t *)pi->p, n, lv.rowids ? (const uint32_t *)lv.rowids->p : nullptr,
                           starts ? (const uint64_t *)starts->p : nullptr, (uint32_t)lbase->seg_starts.size(),
                           lbase->seg_rows, (int64_t *)dseg->p),
          "join order");
    }
    DBufP key;
    TRY(plain_copy(lc[eqs[0].first], n, (uint32_t)n, &key));
    const int64_t *ks[2] = {(const int64_t *)dseg->p, (const int64_t *)key->p};
    int32_t desc[2] = {0, 0};
    DBufP dperm = dalloc(4 * n);
    ALN(aln_sort_keys(g_ctx, ks, desc, 2, n, (uint32_t *)dperm->p), "join order");
    auto permute = [&](std::map<std::string, Col> &m) -> int {
      for (auto &kv : m) {
        DBufP src;
        TRY(plain_copy(kv.second, n, (uint32_t)n, &src));
        DBufP dst = dalloc(8 * n);
        ALN(aln_apply_perm(g_ctx, (const int64_t *)src->p, (const uint32_t *)dperm->p, n, (int64_t *)dst->p),
            "join order");
        Col c;
        TRY(make_plain(dst, n, (uint32_t)n, kv.second.decimals, &c));
        c.type = kv.second.type;
        c.dict = kv.second.dict;
        c.width = kv.second.width;
        kv.second = c;
      }
      return ALN_OK;
    };
    TRY(permute(lc));
    TRY(permute(rc));
  }
  // extra equalities (AND a = b, operators.cu:1288-1348) keep the pairs where they hold
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  out->nrows = n;
  out->seg_rows = (uint32_t)(n ? n : 1);
  for (auto &kv : lc) {
    out->names.push_back(kv.first);
    out->cols[kv.first] = kv.second;
  }
  for (auto &kv : rc)
    if (!out->cols.count(kv.first)) {
      out->names.push_back(kv.first);
      out->cols[kv.first] = kv.second;
    }
  if (eqs.size() > 1 && n) {
    std::vector<aln_col> cols;
    std::vector<aln_rpn> rpn;
    std::vector<int32_t> dec;
    for (size_t e = 1; e < eqs.size(); ++e) {
      int a = (int)cols.size();
      cols.push_back(out->cols[eqs[e].first].desc());
      cols.push_back(out->cols[eqs[e].second].desc());
      dec.push_back(0);
      dec.push_back(0);
      rpn.push_back(aln_rpn{ALN_RPN_NAME, a, 0});
      rpn.push_back(aln_rpn{ALN_RPN_NAME, a + 1, 0});
      rpn.push_back(aln_rpn{ALN_RPN_CMP, 4, 0});
      if (e > 1) rpn.push_back(aln_rpn{ALN_RPN_AND, 0, 0});
    }
    aln_insn prog[ALN_MAX_INSN];
    int nprog = 0;
    ALN(aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 0, prog, &nprog, nullptr),
        "join condition");
    DBufP sel = dalloc(4 * n);
    uint64_t k = 0;
    ALN(aln_scan_filter(g_ctx, cols.data(), (int)cols.size(), out->seg_rows, prog, nprog, nullptr,
                        (uint32_t *)sel->p, n, &k),
        "join condition");
    for (auto &kv : out->cols) {
      DBufP vals = dalloc(8 * (k ? k : 1));
      aln_col d = kv.second.desc();
      if (k) ALN(aln_gather(g_ctx, &d, out->seg_rows, (const uint32_t *)sel->p, k, (int64_t *)vals->p), "gather");
      Col c;
      TRY(make_plain(vals, k, (uint32_t)(k ? k : 1), kv.second.decimals, &c));
      c.type = kv.second.type;
      c.dict = kv.second.dict;
      c.width = kv.second.width;
      kv.second = c;
    }
    out->nrows = k;
    out->seg_rows = (uint32_t)(k ? k : 1);
  }
  varNames[name] = out;
  return ALN_OK;
}

// A column of `n` rows gathered from view `v` through the row ids `idx`, as a plain column.
static int gather_plain(const View &v, const std::string &cn, const DBufP &idx, uint64_t n, Col *out) {
  const Col &sc = v.cols.at(cn);
  DBufP vals = dalloc(8 * (n ? n : 1));
  if (!vals) return fail(ALN_E_NOMEM, "device allocation failed");
  aln_col d = sc.desc();
  if (n) ALN(aln_gather(g_ctx, &d, v.seg_rows, (const uint32_t *)idx->p, n, (int64_t *)vals->p), "join gather");
  TRY(make_plain(vals, n, (uint32_t)(n ? n : 1), sc.decimals, out));
  out->type = sc.type;
  out->dict = sc.dict;
  out->width = sc.width;
  return ALN_OK;
}

// LEFT/RIGHT SEMI and ANTI joins (operators.cu:1357-1399, 1473-1530): the inner join's pairs,
// narrowed by the extra equalities, reduced to the rows of one side that take part (semi) or
// do not (anti), each once.  The left kinds keep the left's row order (the reference works
// per left segment, in row order within each); the right kinds return the right's rows in
// the order of its stable sort by the join key, the order the reference's right side is in
// when it gathers them.  Only the kept side's columns are in the result: the reference
// gathers the other side's through indices that no longer line up with the rows.
static int semi_anti(const std::string &name, char kind, const View &lv, const View &rv,
                     const std::vector<std::pair<std::string, std::string>> &eqs, DBufP pi, DBufP bi,
                     uint64_t np, const std::set<std::string> &lneed, const std::set<std::string> &rneed) {
  if (eqs.size() > 1 && np) {
    // the pairs where every extra equality holds (operators.cu:1288-1348)
    std::vector<Col> tmp;
    std::vector<aln_col> cols;
    std::vector<aln_rpn> rpn;
    std::vector<int32_t> dec;
    for (size_t e = 1; e < eqs.size(); ++e) {
      Col a, b;
      TRY(gather_plain(lv, eqs[e].first, pi, np, &a));
      TRY(gather_plain(rv, eqs[e].second, bi, np, &b));
      if (a.type == 2 || b.type == 2) return fail(ALN_E_UNSUPPORTED, "Join : string join keys");
      if (a.type == 1 || b.type == 1) return fail(ALN_E_UNSUPPORTED, "Join : float join keys");
      const int s = (int)cols.size();
      tmp.push_back(a);
      tmp.push_back(b);
      cols.push_back(a.desc());
      cols.push_back(b.desc());
      dec.push_back(0);
      dec.push_back(0);
      rpn.push_back(aln_rpn{ALN_RPN_NAME, s, 0});
      rpn.push_back(aln_rpn{ALN_RPN_NAME, s + 1, 0});
      rpn.push_back(aln_rpn{ALN_RPN_CMP, 4, 0});
      if (e > 1) rpn.push_back(aln_rpn{ALN_RPN_AND, 0, 0});
    }
    aln_insn prog[ALN_MAX_INSN];
    int nprog = 0;
    ALN(aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 0, prog, &nprog, nullptr),
        "join condition");
    DBufP sel = dalloc(4 * np);
    if (!sel) return fail(ALN_E_NOMEM, "device allocation failed");
    uint64_t k = 0;
    ALN(aln_scan_filter(g_ctx, cols.data(), (int)cols.size(), (uint32_t)np, prog, nprog, nullptr,
                        (uint32_t *)sel->p, np, &k),
        "join condition");
    std::vector<uint32_t> hs(k), hp(np), hb(np);
    if ((k && hipMemcpy(hs.data(), sel->p, 4 * k, hipMemcpyDeviceToHost) != hipSuccess) ||
        hipMemcpy(hp.data(), pi->p, 4 * np, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hb.data(), bi->p, 4 * np, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(ALN_E_HIP, "device to host copy failed");
    for (uint64_t j = 0; j < k; ++j) {
      hp[j] = hp[hs[j]];
      hb[j] = hb[hs[j]];
    }
    if (k && (hipMemcpy(pi->p, hp.data(), 4 * k, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(bi->p, hb.data(), 4 * k, hipMemcpyHostToDevice) != hipSuccess))
      return fail(ALN_E_HIP, "host to device copy failed");
    np = k;
  }
  const bool left_side = kind == '1' || kind == '3';
  const bool anti = kind == '3' || kind == '4';
  const View &v = left_side ? lv : rv;
  DBufP rows = dalloc(4 * (v.nrows ? v.nrows : 1));
  if (!rows) return fail(ALN_E_NOMEM, "device allocation failed");
  uint64_t n = 0;
  ALN(aln_join_rows(g_ctx, (const uint32_t *)(left_side ? pi : bi)->p, np, v.nrows, anti ? 1 : 0,
                    (uint32_t *)rows->p, v.nrows, &n),
      "join");
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  out->nrows = n;
  out->seg_rows = (uint32_t)(n ? n : 1);
  for (auto &cn : left_side ? lneed : rneed) {
    Col c;
    TRY(gather_plain(v, cn, rows, n, &c));
    out->names.push_back(cn);
    out->cols[cn] = c;
  }
  if (!left_side && n > 1) {
    // the right side's order: stable by its join key
    const int64_t *ks[1] = {(const int64_t *)out->cols[eqs[0].second].words->p};
    int32_t desc[1] = {0};
    DBufP dperm = dalloc(4 * n);
    if (!dperm) return fail(ALN_E_NOMEM, "device allocation failed");
    ALN(aln_sort_keys(g_ctx, ks, desc, 1, n, (uint32_t *)dperm->p), "join order");
    for (auto &kv : out->cols) {
      DBufP dst = dalloc(8 * n);
      if (!dst) return fail(ALN_E_NOMEM, "device allocation failed");
      ALN(aln_apply_perm(g_ctx, (const int64_t *)kv.second.words->p, (const uint32_t *)dperm->p, n,
                         (int64_t *)dst->p),
          "join order");
      Col c;
      TRY(make_plain(dst, n, (uint32_t)n, kv.second.decimals, &c));
      c.type = kv.second.type;
      c.dict = kv.second.dict;
      c.width = kv.second.width;
      kv.second = c;
    }
  }
  varNames[name] = out;
  return ALN_OK;
}

static int do_join(const char *s, const char *j1) {
  std::vector<SelItem> items;
  std::vector<Tok> rest;
  parse_select(g_tok, &items, &rest);
  std::vector<Cond> conds;
  TRY(parse_conditions(rest, &conds));
  const size_t nt = g_join.size();
  if (nt == 0 || conds.size() != nt) return fail(ALN_E_ARG, "Join : %zu tables, %zu conditions", nt, conds.size());
  for (auto &it : items)
    if (!it.agg.empty() || it.expr.size() != 1 || it.expr[0].type != "NAME")
      return fail(ALN_E_UNSUPPORTED, "Join : select list entries must be columns");
  std::set<std::string> keep;
  for (auto &it : items) keep.insert(it.expr[0].value);
  for (auto &c : conds)
    for (auto &e : c.eqs) {
      keep.insert(e.first);
      keep.insert(e.second);
    }
  // the chain of emit_join (operators.cu:832-860): op_join holds the tables innermost
  // first; the i-th join (1-based) uses the condition parsed (n+1-i)-th
  std::string prev;
  for (size_t i = 1; i <= nt; ++i) {
    std::string name = i == nt ? std::string(s) + "\x01" : std::string(s) + std::to_string(i);
    std::string left, right;
    if (i == 1) {
      right = g_join[0];
      left = nt > 1 ? g_join[1] : std::string(j1);
    } else {
      left = i < nt ? g_join[i] : std::string(j1);
      right = prev;
    }
    TRY(join_pair(name, left, right, conds[nt - i], keep, g_join_kind[i - 1]));
    if (!prev.empty()) varNames.erase(prev);
    prev = name;
  }
  // the result keeps the selected columns under their aliases
  auto r = find_var(prev);
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  out->nrows = r->nrows;
  out->seg_rows = r->seg_rows;
  for (auto &it : items) {
    const std::string &c = it.expr[0].value;
    if (!r->cols.count(c)) return fail(ALN_E_ARG, "Join : couldn't find column %s", c.c_str());
    out->names.push_back(it.alias);
    out->cols[it.alias] = r->cols[c];
  }
  varNames.erase(prev);
  varNames[s] = out;
  return ALN_OK;
}

void emit_join(const char *s, const char *j1, int grp, int start_seg, int end_seg) {
  (void)grp;
  (void)start_seg;
  (void)end_seg;
  g_status = ALN_OK;
  if (g_join_bad) {
    fail(ALN_E_UNSUPPORTED, "Join : only inner, semi and anti joins are supported");
  } else {
    if (!peek_var(j1 ? j1 : ""))
      fail(ALN_E_ARG, "Join : couldn't find variable %s", j1 ? j1 : "");
    else if (!defer_join(s, j1) && find_var(j1))
      do_join(s, j1);
  }
  finish_statement();
}

// ---- emit_order (operators.cu:1641-1798, sorts.cu:17-125) ----------------------------------
static int do_order(const char *s, const char *f) {
  std::vector<std::string> keys;
  std::vector<int32_t> desc;
  for (auto &k : g_tok) {
    if (k.type == "NAME") {
      keys.push_back(k.value);
      desc.push_back(0);
    } else if (k.type == "ORDER") {
      keys.push_back(k.value);
      desc.push_back((int32_t)k.num);
    }
  }
  if (keys.empty()) return fail(ALN_E_ARG, "Order : no keys");
  if (keys.size() > ALN_MAX_KEYS) return fail(ALN_E_UNSUPPORTED, "ORDER BY on more than %d keys", ALN_MAX_KEYS);
  auto names = all_cols(f);
  std::set<std::string> need(names.begin(), names.end());
  for (auto &k : keys)
    if (!need.count(k)) return fail(ALN_E_ARG, "Couldn't find name %s", k.c_str());
  View v;
  TRY(make_view(f, need, &v));
  const uint64_t n = v.nrows;
  std::vector<DBufP> kb;
  std::vector<const int64_t *> kp;
  std::vector<int32_t> kdesc;
  for (size_t i = 0; i < keys.size(); ++i) {
    const Col &kc = v.cols[keys[i]];
    DBufP d;
    TRY(plain_copy(kc, n, v.seg_rows, &d));
    if (kc.type == 1 && n) {
      // f64 key: doubles sort by value (update_permutation on d_columns_float, sorts.cu:17-86)
      ALN(aln_f64_order_keys(g_ctx, (const int64_t *)d->p, n, (int64_t *)d->p), "order");
    }
    if (kc.dict) {
      // char key: str_sort handles widths 1..101 and 255 and leaves the order as it is for
      // any other width (strings_sort_device.cu:125-137), so such a key does not sort
      if (!((kc.width >= 1 && kc.width <= 101) || kc.width == 255)) continue;
      // rows sort by their string's rank in the dictionary's byte order (the strings are
      // NUL-padded to the width and compared as unsigned bytes, strings_type.h:46-115)
      const auto &dict = *kc.dict;
      std::vector<uint32_t> ord(dict.size());
      for (size_t j = 0; j < ord.size(); ++j) ord[j] = (uint32_t)j;
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return dict[a] < dict[b]; });
      std::vector<int64_t> rank(dict.size());
      for (size_t j = 0; j < ord.size(); ++j) rank[ord[j]] = (int64_t)j;
      DBufP table = dalloc(8 * (rank.size() ? rank.size() : 1));
      DBufP r = dalloc(8 * (n ? n : 1));
      if (!table || !r) return fail(ALN_E_NOMEM, "device allocation failed");
      if (!rank.empty() && hipMemcpy(table->p, rank.data(), 8 * rank.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "host to device copy failed");
      ALN(aln_dict_lookup(g_ctx, (const int64_t *)d->p, n, (const int64_t *)table->p, rank.size(), (int64_t *)r->p),
          "order");
      ALN(aln_sync(g_ctx), "order");
      d = r;
    }
    kb.push_back(d);
    kp.push_back((const int64_t *)d->p);
    kdesc.push_back(desc[i]);
  }
  DBufP perm = dalloc(4 * (n ? n : 1));
  if (n && kp.empty()) {
    std::vector<uint32_t> id(n);
    for (uint64_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
    if (hipMemcpy(perm->p, id.data(), 4 * n, hipMemcpyHostToDevice) != hipSuccess)
      return fail(ALN_E_HIP, "host to device copy failed");
  } else if (n) {
    ALN(aln_sort_keys(g_ctx, kp.data(), kdesc.data(), (int)kp.size(), n, (uint32_t *)perm->p), "order");
  }
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  out->nrows = n;
  out->seg_rows = (uint32_t)(n ? n : 1);
  for (auto &cn : names) {
    DBufP src;
    TRY(plain_copy(v.cols[cn], n, v.seg_rows, &src));
    DBufP dst = dalloc(8 * (n ? n : 1));
    if (n) ALN(aln_apply_perm(g_ctx, (const int64_t *)src->p, (const uint32_t *)perm->p, n, (int64_t *)dst->p), "order");
    Col c;
    TRY(make_plain(dst, n, out->seg_rows, v.cols[cn].decimals, &c));
    c.type = v.cols[cn].type;
    c.dict = v.cols[cn].dict;
    c.width = v.cols[cn].width;
    out->names.push_back(cn);
    out->cols[cn] = c;
  }
  varNames[s] = out;
  return ALN_OK;
}

void emit_order(const char *s, const char *f, int e, int ll) {
  (void)e;
  (void)ll;
  g_status = ALN_OK;
  auto src = peek_var(f ? f : "");
  if (!src)
    fail(ALN_E_ARG, "Order : couldn't find variable %s", f ? f : "");
  else if (!(src->deferred && src->deferred->kind == Deferred::STAR_GROUP && defer_order(s, f, src)) &&
           find_var(f))
    do_order(s, f);
  finish_statement();
}

// ---- emit_store / emit_display: CudaSet::Store text output (cm.cu:1479-1762) ----------------
static void print_rows(FILE *fp, const char *sep, const std::vector<const Col *> &cols,
                       const std::vector<std::vector<int64_t>> &data, uint64_t m) {
  for (uint64_t i = 0; i < m; ++i) {
    for (size_t j = 0; j < cols.size(); ++j) {
      const Col &c = *cols[j];
      int64_t x = data[j][i];
      if (c.dict) {
        const std::string &sv = (*c.dict)[(size_t)x];
        fprintf(fp, "%.*s", (int)c.width, sv.c_str());
      } else if (c.type == 1) {  // float: two digits (cm.cu:1569)
        double d;
        memcpy(&d, &x, 8);
        fprintf(fp, "%.2f", d);
      } else if (c.decimals) {
        std::string str = std::to_string(x);
        while (str.length() <= (size_t)c.decimals) str = '0' + str;
        str.insert(str.length() - c.decimals, ".");
        fprintf(fp, "%s", str.c_str());
      } else {
        fprintf(fp, "%lld", (long long)x);
      }
      fputs(sep, fp);
    }
    if (i != m - 1) fputs("\n", fp);
  }
}

static int store_text(const char *var, FILE *fp, const char *sep, int limit) {
  auto names = all_cols(var);
  View v;
  TRY(make_view(var, std::set<std::string>(names.begin(), names.end()), &v));
  uint64_t m = (limit != 0 && (uint64_t)limit < v.nrows) ? (uint64_t)limit : v.nrows;
  std::vector<std::vector<int64_t>> data(names.size());
  std::vector<const Col *> cols;
  for (size_t j = 0; j < names.size(); ++j) {  // the printed rows only
    TRY(fetch_col(v.cols[names[j]], m, v.seg_rows, &data[j]));
    cols.push_back(&v.cols[names[j]]);
  }
  print_rows(fp, sep, cols, data, m);
  return ALN_OK;
}

void emit_store(const char *s, const char *f, const char *sep) {
  g_status = ALN_OK;
  int limit = g_has_limit ? g_limit : 0;
  auto v = peek_var(s ? s : "");
  if (!v) {
    fail(ALN_E_ARG, "Store : couldn't find variable %s", s ? s : "");
  } else {
    FILE *fp = fopen(f, "w");
    if (!fp) {
      fail(ALN_E_ARG, "Could not open file %s", f);
    } else {
      bool done = false;
      if (store_star_topk(v, fp, sep, limit, &done) == ALN_OK && !done) {
        if (find_var(s)) store_text(s, fp, sep, limit);
      }
      fclose(fp);
    }
  }
  finish_statement();
}

void emit_display(const char *s, const char *sep) {
  g_status = ALN_OK;
  int limit = g_has_limit ? g_limit : 0;
  auto v = peek_var(s ? s : "");
  if (!v) {
    fail(ALN_E_ARG, "Display : couldn't find variable %s", s ? s : "");
  } else {
    bool done = false;
    if (store_star_topk(v, stdout, sep, limit, &done) == ALN_OK && !done) {
      if (find_var(s)) store_text(s, stdout, sep, limit);
    }
  }
  fflush(stdout);
  finish_statement();
}

const char *alenka_last_path(void) { return g_path.c_str(); }

}  // extern "C"

// ---- deferred statements: the star pipeline behind emit_join / emit_select / emit_order ------
namespace {

uint64_t g_tmp_seq = 0;

std::shared_ptr<Set> pinned(const Deferred &d, const std::string &n) {
  for (auto &p : d.pins)
    if (p.first == n) return p.second;
  return nullptr;
}

void pin(Deferred *d, const std::string &n, const std::shared_ptr<Set> &v) {
  if (!pinned(*d, n)) d->pins.push_back({n, v});
}

// The stored table a star scan reads for variable v: v itself, or the source of a filter
// (pinned under its name); NULL when v is the result of another statement.
std::shared_ptr<Set> star_base(const Deferred &d, const std::shared_ptr<Set> &v) {
  if (!v || v->deferred) return nullptr;
  if (!v->filtered) return v;
  auto b = pinned(d, v->source);
  if (!b || b->deferred || b->filtered) return nullptr;
  return b;
}

// A filter as ANDed closed ranges on the table's columns: the compiled program must be a
// conjunction of `column OP constant` terms on unscaled columns (the shapes the scan's
// interval path takes).  False for any other predicate.
bool star_ranges(const Set &var, StarTable *t) {
  if (!var.filtered) return true;
  std::map<std::string, int> slot;
  std::vector<std::string> names;
  for (auto &k : var.pred)
    if (k.type == "NAME" && !slot.count(k.value)) {
      if (!t->base->cols.count(k.value)) return false;
      slot[k.value] = (int)names.size();
      names.push_back(k.value);
    }
  if (names.empty() || names.size() > ALN_MAX_COLS) return false;
  std::vector<aln_rpn> rpn;
  if (resolve_string_codes(*t->base, var.pred, &rpn, slot) != ALN_OK) return false;
  std::vector<int32_t> dec;
  for (auto &n : names) dec.push_back(col_dec(t->base->cols[n]));
  aln_insn prog[ALN_MAX_INSN];
  int np = 0;
  if (aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 0, prog, &np, nullptr) != ALN_OK)
    return false;
  std::vector<int64_t> lo(names.size(), INT64_MIN), hi(names.size(), INT64_MAX);
  std::vector<bool> used(names.size(), false);
  int pending = 0;
  for (int i = 0; i < np;) {
    if (prog[i].op == ALN_I_AND) {
      if (pending < 2) return false;
      --pending;
      ++i;
      continue;
    }
    if (i + 2 >= np || prog[i].op != ALN_I_COL || prog[i].k != 1 || prog[i + 1].op != ALN_I_CONST ||
        prog[i + 2].op != ALN_I_CMP)
      return false;
    const int c = prog[i].a;
    const int64_t v = prog[i + 1].k;
    int64_t l = INT64_MIN, h = INT64_MAX;
    switch (prog[i + 2].a) {
      case 1: if (v == INT64_MIN) return false; h = v - 1; break;
      case 2: if (v == INT64_MAX) return false; l = v + 1; break;
      case 4: l = h = v; break;
      case 5: h = v; break;
      case 6: l = v; break;
      default: return false;
    }
    if (c < 0 || c >= (int)names.size()) return false;
    used[c] = true;
    lo[c] = std::max(lo[c], l);
    hi[c] = std::min(hi[c], h);
    ++pending;
    i += 3;
  }
  if (pending != 1) return false;
  for (size_t c = 0; c < names.size(); ++c) {
    if (!used[c]) continue;
    if (lo[c] > hi[c]) return false;  // an empty filter: the statements run one by one
    aln_range r;
    memset(&r, 0, sizeof(r));
    r.col = t->slot(names[c]);
    r.lo = lo[c];
    r.hi = hi[c];
    t->ranges.push_back(r);
  }
  return t->ranges.size() <= ALN_MAX_RANGES;
}

// SUM's expression as (a0 + a1*x) * (b0 + b1*y): each value on the evaluation stack is an
// affine form in at most one column, or (only as the final value) the product of two.
// Integer arithmetic wraps (mod 2^64) here as in the kernels, so the expanded form gives the
// interpreted program's exact results.
bool sum_form(const aln_insn *prog, int np, int *cx, int64_t *a0, int64_t *a1, int *cy, int64_t *b0,
              int64_t *b1) {
  struct Aff {
    int col = -1;
    uint64_t a = 0, c = 0;
  };
  struct Val {
    bool prod = false;
    Aff l, r;
  };
  std::vector<Val> st;
  for (int i = 0; i < np; ++i) {
    const aln_insn &in = prog[i];
    if (in.op == ALN_I_COL || in.op == ALN_I_CONST) {
      Val v;
      if (in.op == ALN_I_COL) {
        v.l.col = in.a;
        v.l.a = (uint64_t)in.k;
      } else {
        v.l.c = (uint64_t)in.k;
      }
      st.push_back(v);
      continue;
    }
    if (st.size() < 2) return false;
    Val r = st.back();
    st.pop_back();
    Val l = st.back();
    st.pop_back();
    if (l.prod || r.prod) return false;
    Val o;
    if (in.op == ALN_I_ADD || in.op == ALN_I_SUB) {
      if (l.l.col >= 0 && r.l.col >= 0 && l.l.col != r.l.col) return false;
      o.l.col = l.l.col >= 0 ? l.l.col : r.l.col;
      o.l.a = in.op == ALN_I_ADD ? l.l.a + r.l.a : l.l.a - r.l.a;
      o.l.c = in.op == ALN_I_ADD ? l.l.c + r.l.c : l.l.c - r.l.c;
    } else if (in.op == ALN_I_MUL) {
      if (l.l.col < 0 || r.l.col < 0) {
        const Aff &k = l.l.col < 0 ? l.l : r.l, &x = l.l.col < 0 ? r.l : l.l;
        o.l.col = x.col;
        o.l.a = x.a * k.c;
        o.l.c = x.c * k.c;
      } else {
        o.prod = true;
        o.l = l.l;
        o.r = r.l;
      }
    } else {
      return false;
    }
    st.push_back(o);
  }
  if (st.size() != 1) return false;
  const Val &v = st[0];
  Aff x = v.l, y;
  if (v.prod) {
    y = v.r;
  } else {
    y.c = 1;
  }
  *cx = x.col;
  *a0 = (int64_t)x.c;
  *a1 = (int64_t)x.a;
  *cy = y.col;
  *b0 = (int64_t)y.c;
  *b1 = (int64_t)y.a;
  return true;
}

// Plans SELECT ... GROUP BY over a deferred join as one aln_star_agg call: the join must be
// fact JOIN mid ON f_fk = m_key JOIN dim ON m_fk = d_key (inner, one column pair each, the
// three inputs stored tables or filters of them with interval predicates), the GROUP BY's
// first column the mid table's join key and the others mid columns, and the select list
// the group columns, COUNTs and at most one SUM of an expression of the form
// (a0 + a1*x) * (b0 + b1*y) over fact columns.  False (and no error) for any other shape.
bool plan_star(const Deferred &dj, const std::vector<Tok> &sel_tok, int grp_cnt, StarPlan *p) {
  std::vector<SelItem> jitems;
  std::vector<Tok> jrest;
  parse_select(dj.tok, &jitems, &jrest);
  std::vector<Cond> conds;
  if (parse_conditions(jrest, &conds) != ALN_OK) return false;
  if (dj.join.size() != 2 || conds.size() != 2) return false;
  for (char k : dj.join_kind)
    if (k != 'I') return false;
  for (auto &c : conds)
    if (c.eqs.size() != 1) return false;
  auto vd = pinned(dj, dj.join[0]), vm = pinned(dj, dj.join[1]), vf = pinned(dj, dj.src);
  auto bd = star_base(dj, vd), bm = star_base(dj, vm), bf = star_base(dj, vf);
  if (!bd || !bm || !bf) return false;
  auto has = [](const std::shared_ptr<Set> &b, const std::string &c) { return b->cols.count(c) > 0; };
  // the first join: mid (left) with dim (right) on the last condition (operators.cu:832-860)
  std::string mfk, dkey;
  {
    const auto &e = conds[1].eqs[0];
    if (has(bm, e.first) && has(bd, e.second)) {
      mfk = e.first;
      dkey = e.second;
    } else if (has(bm, e.second) && has(bd, e.first)) {
      mfk = e.second;
      dkey = e.first;
    } else {
      return false;
    }
  }
  std::set<std::string> keep;
  std::map<std::string, std::string> alias;  // join output name -> source column
  for (auto &it : jitems) {
    if (!it.agg.empty() || it.expr.size() != 1 || it.expr[0].type != "NAME") return false;
    keep.insert(it.expr[0].value);
    alias[it.alias] = it.expr[0].value;
  }
  for (auto &c : conds)
    for (auto &e : c.eqs) {
      keep.insert(e.first);
      keep.insert(e.second);
    }
  auto j1has = [&](const std::string &c) { return keep.count(c) && (has(bm, c) || has(bd, c)); };
  // the second join: fact (left) with the first join's result on the first condition
  std::string ffk, mkey;
  {
    const auto &e = conds[0].eqs[0];
    if (has(bf, e.first) && j1has(e.second)) {
      ffk = e.first;
      mkey = e.second;
    } else if (has(bf, e.second) && j1has(e.first)) {
      ffk = e.second;
      mkey = e.first;
    } else {
      return false;
    }
    if (!has(bm, mkey)) return false;  // the fact table joins the dimension directly
  }
  for (const std::string *k : {&mfk, &dkey, &ffk, &mkey}) {
    const auto &b = k == &dkey ? bd : k == &ffk ? bf : bm;
    if (b->cols.at(*k).type != 0) return false;  // string / float keys: the join's own error
  }
  // a join output column comes from the fact table, else the mid, else the dimension table
  auto provider = [&](const std::string &a, std::string *col) -> int {
    auto it = alias.find(a);
    if (it == alias.end()) return -1;
    *col = it->second;
    return has(bf, *col) ? 2 : has(bm, *col) ? 1 : has(bd, *col) ? 0 : -1;
  };
  std::vector<SelItem> items;
  std::vector<Tok> rest;
  parse_select(sel_tok, &items, &rest);
  if (items.empty() || grp_cnt < 1 || grp_cnt > 3) return false;
  std::vector<std::string> groups;
  for (auto &k : rest)
    if (k.type == "NAME") groups.push_back(k.value);
  if ((int)groups.size() < grp_cnt) return false;
  groups.erase(groups.begin(), groups.end() - grp_cnt);
  std::vector<std::string> gcol(grp_cnt);
  for (int i = 0; i < grp_cnt; ++i)
    if (provider(groups[i], &gcol[i]) != 1 || bm->cols.at(gcol[i]).type == 1) return false;
  if (gcol[0] != mkey) return false;
  p->dim.base = bd;
  p->mid.base = bm;
  p->fact.base = bf;
  p->d_key = p->dim.slot(dkey);
  p->m_key = p->mid.slot(mkey);
  p->m_fk = p->mid.slot(mfk);
  p->m_g1 = p->mid.slot(grp_cnt > 1 ? gcol[1] : mkey);
  p->m_g2 = p->mid.slot(grp_cnt > 2 ? gcol[2] : grp_cnt > 1 ? gcol[1] : mkey);
  p->f_fk = p->fact.slot(ffk);
  p->f_x = p->f_y = p->f_fk;
  bool have_sum = false;
  for (auto &it : items) {
    StarOut o;
    o.alias = it.alias;
    if (it.agg.empty()) {
      if (it.expr.size() != 1 || it.expr[0].type != "NAME") return false;
      auto g = std::find(groups.begin(), groups.end(), it.expr[0].value);
      if (g == groups.end()) return false;
      o.field = (int)(g - groups.begin());
      const Col &sc = bm->cols.at(gcol[o.field]);
      o.decimals = sc.decimals;
      o.type = sc.type;
      o.dict = sc.dict;
      o.width = sc.width;
    } else if (it.agg == "COUNT") {
      for (auto &k : it.expr) {
        std::string c;
        if (k.type == "NAME" && provider(k.value, &c) < 0) return false;
      }
      o.field = 4;
    } else if (it.agg == "SUM") {
      if (have_sum) return false;
      have_sum = true;
      std::map<std::string, int> slot;
      std::vector<std::string> names;  // fact columns, in first-use order
      Set view;
      for (auto &k : it.expr)
        if (k.type == "NAME" && !slot.count(k.value)) {
          std::string c;
          if (provider(k.value, &c) != 2) return false;
          slot[k.value] = (int)names.size();
          names.push_back(c);
          view.cols[k.value] = bf->cols.at(c);
        }
      if (names.empty()) return false;
      std::vector<aln_rpn> rpn;
      if (resolve_string_codes(view, it.expr, &rpn, slot) != ALN_OK) return false;
      std::vector<int32_t> dec;
      for (auto &n : names) dec.push_back(col_dec(bf->cols.at(n)));
      aln_insn prog[ALN_MAX_INSN];
      int np = 0;
      int32_t prec = 0;
      if (aln_compile_rpn(rpn.data(), (int)rpn.size(), dec.data(), (int)dec.size(), 1, prog, &np, &prec) != ALN_OK)
        return false;
      int cx = -1, cy = -1;
      if (!sum_form(prog, np, &cx, &p->a0, &p->a1, &cy, &p->b0, &p->b1)) return false;
      if (cx >= 0) p->f_x = p->fact.slot(names[cx]);
      if (cy >= 0) p->f_y = p->fact.slot(names[cy]);
      o.field = 3;
      o.decimals = prec;
    } else {
      return false;  // AVG, MIN, MAX
    }
    p->out.push_back(o);
  }
  return star_ranges(*vd, &p->dim) && star_ranges(*vm, &p->mid) && star_ranges(*vf, &p->fact);
}

// One aln_star_agg call of a plan: the first `limit` rows of ORDER BY sum DESC, g1 ASC, key ASC
// into *top, or (limit 0) every group in key order, one int64 array per field, into *fields.
// ALN_E_UNSUPPORTED / ALN_E_ARG (shapes the fused kernels do not take) come back without an
// error being recorded; the caller then runs the statements one by one.
int run_star(const StarPlan &p, uint32_t limit, std::vector<aln_star_row> *top, std::vector<DBufP> *fields,
             uint64_t *ngroups) {
  auto cols_of = [](const StarTable &t, std::vector<aln_col> *v) {
    for (auto &n : t.cols) v->push_back(t.base->cols.at(n).desc());
  };
  std::vector<aln_col> dc, mc, fc;
  cols_of(p.dim, &dc);
  cols_of(p.mid, &mc);
  cols_of(p.fact, &fc);
  aln_star_args a;
  memset(&a, 0, sizeof(a));
  auto scan = [](const StarTable &t, const std::vector<aln_col> &c) {
    aln_table_scan s;
    memset(&s, 0, sizeof(s));
    s.cols = c.data();
    s.ncols = (int32_t)c.size();
    s.seg_rows = t.base->seg_rows;
    s.seg_modes = nullptr;
    s.ranges = t.ranges.data();
    s.nranges = (int32_t)t.ranges.size();
    return s;
  };
  a.dim = scan(p.dim, dc);
  a.mid = scan(p.mid, mc);
  a.fact = scan(p.fact, fc);
  a.d_key = p.d_key;
  a.m_key = p.m_key;
  a.m_fk = p.m_fk;
  a.m_g1 = p.m_g1;
  a.m_g2 = p.m_g2;
  a.f_fk = p.f_fk;
  a.f_x = p.f_x;
  a.f_y = p.f_y;
  a.a0 = p.a0;
  a.a1 = p.a1;
  a.b0 = p.b0;
  a.b1 = p.b1;
  auto soft = [](int rc) { return rc == ALN_E_UNSUPPORTED || rc == ALN_E_ARG; };
  if (limit) {
    a.limit = limit;
    top->assign(limit, aln_star_row());
    uint32_t nt = 0;
    int rc = aln_star_agg(g_ctx, &a, top->data(), &nt, nullptr, 0, nullptr);
    if (soft(rc)) return rc;
    ALN(rc, "star pipeline");
    top->resize(nt);
    return ALN_OK;
  }
  a.flags = 1;
  uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(p.fact.base->nrows, 1u << 20));
  for (;;) {
    DBufP rows = dalloc(sizeof(aln_star_row) * cap);
    if (!rows) return fail(ALN_E_NOMEM, "device allocation failed");
    aln_star_stats st;
    memset(&st, 0, sizeof(st));
    uint32_t nt = 0;
    int rc = aln_star_agg(g_ctx, &a, nullptr, &nt, (aln_star_row *)rows->p, cap, &st);
    if (rc == ALN_E_CAPACITY && st.groups > cap) {
      cap = st.groups;
      continue;
    }
    if (soft(rc)) return rc;
    ALN(rc, "star pipeline");
    const uint64_t g = st.groups;
    fields->clear();
    for (int f = 0; f < 5; ++f) {  // row layout -> one array per field
      DBufP d = dalloc(8 * (g ? g : 1));
      if (!d) return fail(ALN_E_NOMEM, "device allocation failed");
      if (g && hipMemcpy2D(d->p, 8, (const char *)rows->p + 8 * f, sizeof(aln_star_row), 8, g,
                           hipMemcpyDeviceToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "device copy failed");
      fields->push_back(d);
    }
    *ngroups = g;
    return ALN_OK;
  }
}

// The GROUP BY's result set from the fused pipeline: every group in key order (the order a
// one-segment GROUP BY over the join's rows gives, cm.cu:910-956).
int star_groups(const StarPlan &p, const std::string &name) {
  std::vector<DBufP> fields;
  uint64_t g = 0;
  TRY(run_star(p, 0, nullptr, &fields, &g));
  auto out = std::make_shared<Set>();
  out->not_compressed = true;
  out->nrows = g;
  out->seg_rows = (uint32_t)(g ? g : 1);
  for (auto &o : p.out) {
    Col c;
    TRY(make_plain(fields[o.field], g, out->seg_rows, o.decimals, &c));
    c.type = o.type;
    c.dict = o.dict;
    c.width = o.width;
    out->names.push_back(o.alias);
    out->cols[o.alias] = c;
  }
  varNames[name] = out;
  g_path = "aln_star_agg groups";
  return ALN_OK;
}

}  // namespace

namespace {

// Runs a deferred statement in place: the consumer's own statement state is set aside, the
// variables the statement read are put back under their names while it runs, and its result
// replaces the set's contents.  On failure the set stays deferred and the error becomes the
// consumer's.
int run_deferred(const std::shared_ptr<Set> &s) {
  std::shared_ptr<Deferred> d = s->deferred;
  if (!d) return ALN_OK;
  s->deferred = nullptr;
  const auto tok = g_tok;
  const auto join = g_join;
  const auto jk = g_join_kind;
  const int lim = g_limit;
  const bool hl = g_has_limit, jb = g_join_bad;
  std::vector<std::pair<std::string, std::shared_ptr<Set>>> saved;
  for (auto &pn : d->pins) {
    auto it = varNames.find(pn.first);
    saved.push_back({pn.first, it == varNames.end() ? nullptr : it->second});
    varNames[pn.first] = pn.second;
  }
  const std::string tmp = std::string("\x02") + std::to_string(++g_tmp_seq);
  g_tok = d->tok;
  g_join = d->join;
  g_join_kind = d->join_kind;
  g_limit = 0;
  g_has_limit = false;
  g_join_bad = false;
  int rc = ALN_OK;
  if (d->kind == Deferred::JOIN) {
    rc = do_join(tmp.c_str(), d->src.c_str());
    if (rc == ALN_OK) g_path = "unfused";
  } else if (d->kind == Deferred::STAR_GROUP) {
    rc = star_groups(*d->star, tmp);
    if (rc == ALN_E_UNSUPPORTED || rc == ALN_E_ARG) {
      g_tok = d->tok;
      rc = find_var(d->src) ? do_select(tmp.c_str(), d->src.c_str(), d->grp_cnt) : g_status;
      if (rc == ALN_OK) g_path = "unfused";
    }
  } else {
    rc = find_var(d->src) ? do_order(tmp.c_str(), d->src.c_str()) : g_status;
  }
  for (auto it = saved.rbegin(); it != saved.rend(); ++it) {
    if (it->second)
      varNames[it->first] = it->second;
    else
      varNames.erase(it->first);
  }
  g_tok = tok;
  g_join = join;
  g_join_kind = jk;
  g_limit = lim;
  g_has_limit = hl;
  g_join_bad = jb;
  std::shared_ptr<Set> r;
  if (rc == ALN_OK) {
    auto it = varNames.find(tmp);
    if (it != varNames.end()) r = it->second;
    if (!r) rc = fail(ALN_E_ARG, "deferred statement left no result");
  }
  // its intermediate variables (do_join names them after the statement)
  for (auto it = varNames.begin(); it != varNames.end();) {
    if (it->first.compare(0, tmp.size(), tmp) == 0)
      it = varNames.erase(it);
    else
      ++it;
  }
  if (rc != ALN_OK) {
    s->deferred = d;
    if (g_deferred_rc == ALN_OK) {
      g_deferred_rc = rc == ALN_OK ? ALN_E_ARG : rc;
      snprintf(g_deferred_err, sizeof(g_deferred_err), "%s", g_err);
    }
    return rc;
  }
  *s = *r;
  return ALN_OK;
}

bool defer_join(const char *s, const char *j1) {
  if (g_join.size() != 2 || g_join_bad) return false;
  for (char k : g_join_kind)
    if (k != 'I') return false;
  auto d = std::make_shared<Deferred>();
  d->kind = Deferred::JOIN;
  d->tok = g_tok;
  d->join = g_join;
  d->join_kind = g_join_kind;
  d->src = j1;
  for (const std::string &n : {std::string(j1), g_join[0], g_join[1]}) {
    auto v = peek_var(n);
    if (!v || v->deferred) return false;
    pin(d.get(), n, v);
    if (v->filtered) {
      auto b = peek_var(v->source);
      if (!b || b->deferred || b->filtered) return false;
      pin(d.get(), v->source, b);
    }
  }
  // the statement must be one the star plan can take apart (its columns and conditions
  // resolve); anything else runs now, with the reference's errors
  const int st = g_status;
  std::string err = g_err;
  std::vector<SelItem> items;
  std::vector<Tok> rest;
  parse_select(g_tok, &items, &rest);
  std::vector<Cond> conds;
  const bool ok = parse_conditions(rest, &conds) == ALN_OK && conds.size() == 2 && conds[0].eqs.size() == 1 &&
                  conds[1].eqs.size() == 1 && !items.empty();
  g_status = st;
  snprintf(g_err, sizeof(g_err), "%s", err.c_str());
  if (!ok) return false;
  auto bd = star_base(*d, pinned(*d, g_join[0])), bm = star_base(*d, pinned(*d, g_join[1])),
       bf = star_base(*d, pinned(*d, j1));
  if (!bd || !bm || !bf) return false;
  for (auto &it : items) {
    if (!it.agg.empty() || it.expr.size() != 1 || it.expr[0].type != "NAME") return false;
    const std::string &c = it.expr[0].value;
    if (!bd->cols.count(c) && !bm->cols.count(c) && !bf->cols.count(c)) return false;
  }
  for (auto &c : conds)
    for (auto &e : c.eqs)
      for (const std::string *n : {&e.first, &e.second}) {
        auto in = [&](const std::shared_ptr<Set> &b) { return b->cols.count(*n) && b->cols.at(*n).type == 0; };
        if (!in(bd) && !in(bm) && !in(bf)) return false;
      }
  auto b = std::make_shared<Set>();
  b->deferred = d;
  varNames[s] = b;
  return true;
}

bool defer_star_group(const char *s, const char *f, int grp_cnt, const std::shared_ptr<Set> &src) {
  auto plan = std::make_shared<StarPlan>();
  const int st = g_status;
  std::string err = g_err;
  const bool ok = plan_star(*src->deferred, g_tok, grp_cnt, plan.get());
  g_status = st;
  snprintf(g_err, sizeof(g_err), "%s", err.c_str());
  if (!ok) return false;
  auto d = std::make_shared<Deferred>();
  d->kind = Deferred::STAR_GROUP;
  d->tok = g_tok;
  d->src = f;
  d->grp_cnt = grp_cnt;
  d->star = plan;
  pin(d.get(), f, src);
  auto b = std::make_shared<Set>();
  b->deferred = d;
  varNames[s] = b;
  return true;
}

bool defer_order(const char *s, const char *f, const std::shared_ptr<Set> &src) {
  auto d = std::make_shared<Deferred>();
  d->kind = Deferred::ORDER;
  d->tok = g_tok;
  d->src = f;
  d->star = src->deferred->star;
  for (auto &k : g_tok) {
    if (k.type == "NAME")
      d->order_keys.push_back({k.value, 0});
    else if (k.type == "ORDER")
      d->order_keys.push_back({k.value, (int)k.num});
  }
  if (d->order_keys.empty() || d->order_keys.size() > ALN_MAX_KEYS) return false;
  for (auto &k : d->order_keys) {
    bool found = false;
    for (auto &o : d->star->out) found = found || o.alias == k.first;
    if (!found) return false;  // the statement runs now, with the reference's error
  }
  pin(d.get(), f, src);
  auto b = std::make_shared<Set>();
  b->deferred = d;
  varNames[s] = b;
  return true;
}

// STORE / DISPLAY ... LIMIT k of a deferred ORDER BY over a deferred star GROUP BY: when the
// ORDER BY's keys give the fused top-k's order (SUM DESC, then the first group column after
// the key ASC, then the key: a one-segment GROUP BY returns groups in key order and the sort
// is stable, so the key breaks the remaining ties), the first k rows come from one
// aln_star_agg call.  *done is false when the chain has another shape or the fused kernels
// refuse it; the caller then runs the statements one by one.
int store_star_topk(const std::shared_ptr<Set> &v, FILE *fp, const char *sep, int limit, bool *done) {
  *done = false;
  if (!v->deferred || v->deferred->kind != Deferred::ORDER || !v->deferred->star) return ALN_OK;
  if (limit <= 0 || limit > 512) return ALN_OK;
  const Deferred &d = *v->deferred;
  const StarPlan &p = *d.star;
  // the output column each alias names (the last item of that name, as do_select stores it)
  std::map<std::string, const StarOut *> by_alias;
  for (auto &o : p.out) by_alias[o.alias] = &o;
  const bool g1_is_key = p.m_g1 == p.m_key;
  std::vector<std::pair<int, int>> want = {{3, 1}};
  if (!g1_is_key) want.push_back({1, 0});
  want.push_back({0, 0});
  std::vector<std::pair<int, int>> got;
  for (auto &k : d.order_keys) {
    const StarOut *o = by_alias.at(k.first);
    int field = o->field;
    if (field == 1 && g1_is_key) field = 0;
    if (field == 2 && p.m_g2 == p.m_key) field = 0;
    if (field == 2 && p.m_g2 == p.m_g1) field = g1_is_key ? 0 : 1;
    if (field == 1 && o->dict) return ALN_OK;  // char keys sort by their strings
    bool seen = false;
    for (auto &g : got) seen = seen || g.first == field;
    if (seen) continue;
    got.push_back({field, k.second});
    if (field == 0) break;  // the key is unique: later keys change nothing
  }
  if (got.empty() || got.back().first != 0) got.push_back({0, 0});
  if (got != want) return ALN_OK;
  std::vector<aln_star_row> top;
  int rc = run_star(p, (uint32_t)limit, &top, nullptr, nullptr);
  if (rc == ALN_E_UNSUPPORTED || rc == ALN_E_ARG) return ALN_OK;
  TRY(rc);
  // the GROUP BY's columns in select-list order, each alias printing its last item's column
  std::vector<Col> metas;
  std::vector<std::vector<int64_t>> data;
  for (auto &o : p.out) {
    const StarOut &src = *by_alias.at(o.alias);
    Col c;
    c.decimals = src.decimals;
    c.type = src.type;
    c.dict = src.dict;
    c.width = src.width;
    metas.push_back(c);
    std::vector<int64_t> col;
    for (auto &r : top) {
      const int64_t f[5] = {r.key, r.g1, r.g2, r.sum, r.count};
      col.push_back(f[src.field]);
    }
    data.push_back(col);
  }
  std::vector<const Col *> cols;
  for (auto &c : metas) cols.push_back(&c);
  print_rows(fp, sep, cols, data, top.size());
  g_path = "aln_star_agg top-k";
  *done = true;
  return ALN_OK;
}

}  // namespace

// ---- binary tables: the reference's on-disk format ------------------------------------------
// <f>.<col>.header       u64 rows, u32 segments, u32 most rows in a segment, u32 largest
//                        payload (writeHeader, cm.cu:1251-1263)
// <f>.<col>.<seg>        int / decimal segment, FOR or FOR-DELTA (compress.cu:388-399, 511-522)
// <f>.<col>              char dictionary: width-byte NUL-padded entries, code = entry index
// <f>.<col>.<seg>.idx    char codes of a segment, FOR; <f>.<col>.<seg>.hash: u32 rows, then
//                        MurmurHash64A(entry, width, hash_seed) / 2 per row (compress_char,
//                        cm.cu:1765-1844)
// <f>.sort, <f>.presort  u32 count, then (u32 length, name) per sorted column (cm.cu:1279-1324)
namespace {

bool read_file(const std::string &path, std::vector<char> *out) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) return false;
  out->clear();
  char buf[1 << 16];
  size_t got;
  while ((got = fread(buf, 1, sizeof(buf), fp)) > 0) out->insert(out->end(), buf, buf + got);
  fclose(fp);
  return true;
}

bool write_file(const std::string &path, const void *data, size_t len) {
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) return false;
  const bool ok = len == 0 || fwrite(data, 1, len, fp) == len;
  return fclose(fp) == 0 && ok;
}

int read_name_list(const std::string &path, std::vector<std::string> *out) {
  std::vector<char> b;
  out->clear();
  if (!read_file(path, &b)) return ALN_OK;  // no such list
  uint32_t n = 0;
  if (b.size() < 4) return fail(ALN_E_ARG, "%s is truncated", path.c_str());
  memcpy(&n, b.data(), 4);
  size_t p = 4;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t len = 0;
    if (p + 4 > b.size()) return fail(ALN_E_ARG, "%s is truncated", path.c_str());
    memcpy(&len, b.data() + p, 4);
    p += 4;
    if (p + len > b.size()) return fail(ALN_E_ARG, "%s is truncated", path.c_str());
    out->emplace_back(b.data() + p, len);
    p += len;
  }
  return ALN_OK;
}

// emit_load_binary (operators.cu:2518-2562) with CudaSet::initialize (cm.cu:2849-2962): the
// columns data_dict lists for `file`, in its order; the row and segment counts from the first
// column's header; every segment decoded on the device (aln_segfile_decode) and re-encoded
// into this build's bit-packed layout with the files' largest segment as seg_rows, the file
// segment boundaries kept.
int load_binary(const std::string &name, const std::string &file) {
  auto dd = g_data_dict.find(file);
  if (dd == g_data_dict.end() || dd->second.empty())
    return fail(ALN_E_ARG, "Couldn't find %s in the data dictionary", file.c_str());
  const std::string first = dd->second.begin()->first;
  std::vector<char> hb;
  if (!read_file(file + "." + first + ".header", &hb) || hb.size() < 16)
    return fail(ALN_E_ARG, "Couldn't open file %s.%s.header", file.c_str(), first.c_str());
  uint64_t total = 0;
  uint32_t nseg = 0, max_recs = 0;
  memcpy(&total, hb.data(), 8);
  memcpy(&nseg, hb.data() + 8, 4);
  memcpy(&max_recs, hb.data() + 12, 4);
  auto s = std::make_shared<Set>();
  s->nrows = total;
  s->seg_rows = max_recs ? max_recs : 1;
  TRY(read_name_list(file + ".sort", &s->sorted_fields));
  TRY(read_name_list(file + ".presort", &s->presorted_fields));
  DBufP vals = dalloc(8 * (total ? total : 1));
  if (!vals) return fail(ALN_E_NOMEM, "device allocation failed");
  for (auto &kv : dd->second) {
    const std::string &cn = kv.first;
    const ColData cd = kv.second;
    if (cd.type == 2 && cd.length == 0) return fail(ALN_E_ARG, "char column %s without a width", cn.c_str());
    if (cd.type == 1) {
      // float segments: [u32 cnt][cnt doubles][u32 comp_type 3] (cm.cu:1150-1180), or a
      // decimal float stored as value*100 in a FOR file, read back as x / 100.0 (long_to_float,
      // cm.h:300-307; cm.cu:767-787)
      std::vector<int64_t> h;
      std::vector<uint64_t> starts;
      std::vector<char> fb;
      for (uint32_t g = 0; g < nseg; ++g) {
        const std::string path = file + "." + cn + "." + std::to_string(g);
        if (!read_file(path, &fb) || fb.size() < 8) return fail(ALN_E_ARG, "Error opening %s file", path.c_str());
        uint32_t cnt = 0, ct = 0;
        memcpy(&cnt, fb.data(), 4);
        memcpy(&ct, fb.data() + fb.size() - 4, 4);
        starts.push_back(h.size());
        if (ct == 3 && fb.size() == 8 + 8ull * cnt) {
          const size_t at = h.size();
          h.resize(at + cnt);
          memcpy(h.data() + at, fb.data() + 4, 8ull * cnt);
        } else {
          uint64_t rows = 0;
          int64_t lo = 0, hi = 0;
          ALN(aln_segfile_decode(g_ctx, fb.data(), fb.size(), (int64_t *)vals->p, total, &rows, &lo, &hi),
              path.c_str());
          ALN(aln_sync(g_ctx), "load");
          std::vector<int64_t> iv(rows);
          if (rows && hipMemcpy(iv.data(), vals->p, 8 * rows, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(ALN_E_HIP, "device to host copy failed");
          for (int64_t x : iv) {
            const double d = (double)x / 100.0;
            int64_t b;
            memcpy(&b, &d, 8);
            h.push_back(b);
          }
        }
      }
      if (h.size() != total)
        return fail(ALN_E_ARG, "%s: segments hold %llu rows, header %llu", cn.c_str(), (unsigned long long)h.size(),
                    (unsigned long long)total);
      starts.push_back(total);
      if (s->seg_starts.empty())
        s->seg_starts = starts;
      else if (s->seg_starts != starts)
        return fail(ALN_E_ARG, "column %s is segmented differently", cn.c_str());
      DBufP fv = dalloc(8 * (total ? total : 1));
      if (!fv) return fail(ALN_E_NOMEM, "device allocation failed");
      if (total && hipMemcpy(fv->p, h.data(), 8 * total, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "host to device copy failed");
      Col c;
      TRY(make_plain(fv, total, s->seg_rows, 0, &c));
      c.type = 1;
      s->names.push_back(cn);
      s->cols[cn] = c;
      continue;
    }
    // the segments, back to back on the device
    std::vector<uint64_t> starts;
    uint64_t off = 0;
    std::vector<char> fb;
    for (uint32_t g = 0; g < nseg; ++g) {
      const std::string path = file + "." + cn + "." + std::to_string(g) + (cd.type == 2 ? ".idx" : "");
      if (!read_file(path, &fb)) return fail(ALN_E_ARG, "Error opening %s file", path.c_str());
      uint64_t rows = 0;
      int64_t lo = 0, hi = 0;
      ALN(aln_segfile_decode(g_ctx, fb.data(), fb.size(), (int64_t *)vals->p + off, total - off, &rows, &lo, &hi),
          path.c_str());
      starts.push_back(off);
      off += rows;
    }
    if (off != total)
      return fail(ALN_E_ARG, "%s: segments hold %llu rows, header %llu", cn.c_str(), (unsigned long long)off,
                  (unsigned long long)total);
    starts.push_back(total);
    if (s->seg_starts.empty())
      s->seg_starts = starts;
    else if (s->seg_starts != starts)
      return fail(ALN_E_ARG, "column %s is segmented differently", cn.c_str());
    ALN(aln_sync(g_ctx), "load");
    const int decimals = (cd.type == 2 || cd.length == 0xFFFFFFFFu) ? 0 : (int)cd.length;
    Col c;
    TRY(make_packed((const int64_t *)vals->p, total, s->seg_rows, decimals, &c));
    if (cd.type == 2) {
      std::vector<char> db;
      if (!read_file(file + "." + cn, &db)) return fail(ALN_E_ARG, "Couldn't open file %s.%s", file.c_str(), cn.c_str());
      auto dict = std::make_shared<std::vector<std::string>>();
      for (size_t e = 0; e + cd.length <= db.size(); e += cd.length)
        dict->emplace_back(db.data() + e, strnlen(db.data() + e, cd.length));
      c.type = 2;
      c.dict = dict;
      c.width = cd.length;
    }
    s->names.push_back(cn);
    s->cols[cn] = c;
  }
  varNames[name] = s;
  return ALN_OK;
}

// rows per stored segment: -l (bison.y:299-308), 10^9 unless set
uint64_t g_process_count = 1000000000ull;

// emit_store_binary (operators.cu:2446-2516) for a variable held on the device: data_dict is
// updated (CudaSet::Store, cm.cu:1685-1701) and each column is written in segments of
// g_process_count rows the way the text-source path of CudaSet::compress writes them
// (cm.cu:1034-1240): FOR segment files, the char dictionary rebuilt in first-seen order of the
// stored rows with .idx and .hash files, a header per column; .sort and .presort are removed.
int store_binary(const std::string &var, const std::string &file) {
  auto names = all_cols(var);
  View v;
  TRY(make_view(var, std::set<std::string>(names.begin(), names.end()), &v));
  const uint64_t n = v.nrows, pc = g_process_count;
  if (n > 0xFFFFFFFFull * pc) return fail(ALN_E_UNSUPPORTED, "Store : too many rows");
  const uint32_t nseg = (uint32_t)((n + pc - 1) / pc);
  const uint32_t most = (uint32_t)std::min<uint64_t>(n, pc);
  std::map<std::string, ColData> cols;
  std::vector<char> fb;
  for (auto &cn : names) {
    const Col &c = v.cols[cn];
    DBufP d;
    TRY(plain_copy(c, n, v.seg_rows, &d));
    ColData cd;
    cd.type = c.dict ? 2 : c.type == 1 ? 1 : 0;
    cd.length = c.dict ? c.width : c.type == 1 ? 0 : (uint32_t)c.decimals;
    if (c.type == 1) {
      // doubles, uncompressed: [u32 cnt][cnt doubles][u32 comp_type 3] (cm.cu:1150-1180)
      std::vector<int64_t> h(n);
      if (n && hipMemcpy(h.data(), d->p, 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ALN_E_HIP, "device to host copy failed");
      uint32_t largest = 0;
      for (uint32_t g = 0; g < nseg; ++g) {
        const uint64_t r0 = (uint64_t)g * pc;
        const uint32_t r = (uint32_t)std::min<uint64_t>(pc, n - r0), ct = 3;
        fb.assign(8 + 8ull * r, 0);
        memcpy(fb.data(), &r, 4);
        memcpy(fb.data() + 4, h.data() + r0, 8ull * r);
        memcpy(fb.data() + 4 + 8ull * r, &ct, 4);
        largest = std::max(largest, 8 * r);
        const std::string sp = file + "." + cn + "." + std::to_string(g);
        if (!write_file(sp, fb.data(), fb.size())) return fail(ALN_E_ARG, "Could not write %s", sp.c_str());
      }
      char hdr[20];
      memcpy(hdr, &n, 8);
      memcpy(hdr + 8, &nseg, 4);
      memcpy(hdr + 12, &most, 4);
      memcpy(hdr + 16, &largest, 4);
      if (!write_file(file + "." + cn + ".header", hdr, sizeof(hdr)))
        return fail(ALN_E_ARG, "Could not write %s.%s.header", file.c_str(), cn.c_str());
      cols[cn] = cd;
      continue;
    }
    if (c.dict) {
      // the stored rows' own dictionary, first-seen order (compress_char, cm.cu:1808-1821)
      std::vector<int64_t> h(n);
      if (n && hipMemcpy(h.data(), d->p, 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ALN_E_HIP, "device to host copy failed");
      std::map<int64_t, int64_t> recode;
      std::vector<char> dfile, hfile;
      for (uint64_t i = 0; i < n; ++i) {
        auto it = recode.find(h[i]);
        if (it == recode.end()) {
          it = recode.emplace(h[i], (int64_t)recode.size()).first;
          std::string e = (h[i] >= 0 && (size_t)h[i] < c.dict->size()) ? (*c.dict)[(size_t)h[i]] : std::string();
          e.resize(c.width, '\0');
          dfile.insert(dfile.end(), e.begin(), e.end());
        }
        h[i] = it->second;
      }
      if (!write_file(file + "." + cn, dfile.data(), dfile.size()))
        return fail(ALN_E_ARG, "Could not write %s.%s", file.c_str(), cn.c_str());
      if (n && hipMemcpy(d->p, h.data(), 8 * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ALN_E_HIP, "host to device copy failed");
      for (uint32_t g = 0; g < nseg; ++g) {
        const uint64_t r0 = (uint64_t)g * pc, r = std::min<uint64_t>(pc, n - r0);
        const uint32_t r32 = (uint32_t)r;
        hfile.assign(4 + 8 * r, 0);
        memcpy(hfile.data(), &r32, 4);
        for (uint64_t i = 0; i < r; ++i) {
          const char *e = dfile.data() + (size_t)h[r0 + i] * c.width;
          const uint64_t hv = aln_murmur64a(e, (int)c.width, g_hash_seed) / 2;
          memcpy(hfile.data() + 4 + 8 * i, &hv, 8);
        }
        const std::string hp = file + "." + cn + "." + std::to_string(g) + ".hash";
        if (!write_file(hp, hfile.data(), hfile.size())) return fail(ALN_E_ARG, "Could not write %s", hp.c_str());
      }
    }
    uint32_t largest = 0;
    for (uint32_t g = 0; g < nseg; ++g) {
      const uint64_t r0 = (uint64_t)g * pc, r = std::min<uint64_t>(pc, n - r0);
      uint64_t len = 0;
      ALN(aln_segfile_encode(g_ctx, (const int64_t *)d->p + r0, r, nullptr, 0, &len), "store");
      fb.resize(len);
      ALN(aln_segfile_encode(g_ctx, (const int64_t *)d->p + r0, r, fb.data(), fb.size(), &len), "store");
      uint32_t cnt = 0;
      memcpy(&cnt, fb.data(), 4);
      largest = std::max(largest, cnt);
      const std::string sp = file + "." + cn + "." + std::to_string(g) + (c.dict ? ".idx" : "");
      if (!write_file(sp, fb.data(), fb.size())) return fail(ALN_E_ARG, "Could not write %s", sp.c_str());
    }
    char hdr[20];
    memcpy(hdr, &n, 8);
    memcpy(hdr + 8, &nseg, 4);
    memcpy(hdr + 12, &most, 4);
    memcpy(hdr + 16, &largest, 4);
    if (!write_file(file + "." + cn + ".header", hdr, sizeof(hdr)))
      return fail(ALN_E_ARG, "Could not write %s.%s.header", file.c_str(), cn.c_str());
    cols[cn] = cd;
  }
  g_data_dict[file] = cols;
  remove((file + ".sort").c_str());
  remove((file + ".presort").c_str());
  return ALN_OK;
}

}  // namespace

extern "C" {

void emit_load_binary(const char *s, const char *f, int d) {
  (void)d;
  g_status = ALN_OK;
  if (ensure_ctx() == ALN_OK) {
    if (!s || !f)
      fail(ALN_E_ARG, "emit_load_binary: bad arguments");
    else
      load_binary(s, f);
  }
  finish_statement();
}

void emit_store_binary(const char *s, const char *f, int append) {
  g_status = ALN_OK;
  if (append)
    fail(ALN_E_UNSUPPORTED, "Store : APPEND is not supported");
  else if (!f || !find_var(s ? s : ""))
    fail(ALN_E_ARG, "Store : couldn't find variable %s", s ? s : "");
  else
    store_binary(s, f);
  finish_statement();
}

void alenka_set_process_count(uint64_t rows) { g_process_count = rows ? rows : 1; }

int alenka_load_col_data(const char *file) {
  g_status = ALN_OK;
  std::vector<char> b;
  if (!file || !read_file(file, &b)) return fail(ALN_E_ARG, "Couldn't open data dictionary");
  size_t p = 0;
  auto rd = [&](void *dst, size_t len) {
    if (p + len > b.size()) return false;
    memcpy(dst, b.data() + p, len);
    p += len;
    return true;
  };
  auto rds = [&](std::string *out) {
    uint64_t len = 0;
    if (!rd(&len, 8) || p + len > b.size()) return false;
    out->assign(b.data() + p, (size_t)len);
    p += (size_t)len;
    return true;
  };
  uint64_t ntab = 0;
  if (!rd(&ntab, 8)) return fail(ALN_E_ARG, "data dictionary %s is truncated", file);
  for (uint64_t t = 0; t < ntab; ++t) {
    std::string tab;
    uint64_t ncol = 0;
    if (!rds(&tab) || !rd(&ncol, 8)) return fail(ALN_E_ARG, "data dictionary %s is truncated", file);
    for (uint64_t c = 0; c < ncol; ++c) {
      std::string col;
      ColData cd;
      if (!rds(&col) || !rd(&cd.type, 4) || !rd(&cd.length, 4))
        return fail(ALN_E_ARG, "data dictionary %s is truncated", file);
      g_data_dict[tab][col] = cd;
    }
  }
  return ALN_OK;
}

int alenka_save_col_data(const char *file) {
  g_status = ALN_OK;
  std::vector<char> b;
  auto put = [&](const void *src, size_t len) { b.insert(b.end(), (const char *)src, (const char *)src + len); };
  auto put_str = [&](const std::string &str) {
    const uint64_t len = str.size();
    put(&len, 8);
    put(str.data(), str.size());
  };
  const uint64_t ntab = g_data_dict.size();
  put(&ntab, 8);
  for (auto &t : g_data_dict) {
    put_str(t.first);
    const uint64_t ncol = t.second.size();
    put(&ncol, 8);
    for (auto &c : t.second) {
      put_str(c.first);
      put(&c.second.type, 4);
      put(&c.second.length, 4);
    }
  }
  if (!file || !write_file(file, b.data(), b.size())) return fail(ALN_E_ARG, "Could not write the data dictionary");
  return ALN_OK;
}

}  // extern "C"


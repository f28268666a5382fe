// Script front end: the reference's statement language (grammar bison.y:127-287, tokens
// fl.l:36-133) parsed by recursive descent and turned into the same emit_* calls, in the
// same order, that the reference's parser actions make, so that scripts such as TPC-H q3
// (tpch_queries.zip q3.sql) run end to end over tables registered with alenka_table_*.
//
// Statements on the hot path are accepted:
//   V := FILTER T BY expr;
//   V := SELECT expr AS name, ... FROM T [JOIN U ON expr ...] [GROUP BY expr, ...];
//   V := ORDER T BY name [ASC|DESC], ...;
//   STORE V INTO 'file' USING ('sep') [LIMIT n];   DISPLAY V USING ('sep') [LIMIT n];
//   STORE V INTO 'file' [LIMIT n] BINARY;   (tables in the data dictionary load by name)
//   DROP TABLE V;
// Others (text LOAD, APPEND, sorted binary STORE, DESCRIBE, INSERT, DELETE, SHOW, CREATE
// INDEX, outer joins) fail with ALN_E_UNSUPPORTED.  Operator precedence is the grammar's (bison.y:41-57):
// '=' binds loosest, then '<>', OR, XOR, AND, LIKE, NOT, comparisons, + -, * / %.  Numbers
// are lexed as the reference lexes them: "-?[0-9]+" is one token, so "a-1" is NAME INTNUM.
#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/alenka_hip.h"
#include "../../include/alenka_ops.h"

namespace {

// The value of an APPROXNUM literal as the reference's lexer computes it (atoff, atof.cu:16-90,
// Tom Van Baak's published fast atof): the mantissa digits accumulated in doubles, the
// fraction digit by digit over growing powers of ten, the exponent as a product of 1e50, 1e8
// and 10 factors (capped at 308), then one division or multiplication.  This is not
// correctly rounded, and the literal's value decides which rows a comparison keeps, so the
// same steps are taken.
double approx_value(const char *p) {
  double sign = 1.0, value = 0.0, scale = 1.0;
  if (*p == '-') {
    sign = -1.0;
    ++p;
  } else if (*p == '+') {
    ++p;
  }
  for (; *p >= '0' && *p <= '9'; ++p) value = value * 10.0 + (*p - '0');
  if (*p == '.') {
    double pw = 10.0;
    for (++p; *p >= '0' && *p <= '9'; ++p) {
      value += (*p - '0') / pw;
      pw *= 10.0;
    }
  }
  bool neg = false;
  if (*p == 'e' || *p == 'E') {
    ++p;
    if (*p == '-') {
      neg = true;
      ++p;
    } else if (*p == '+') {
      ++p;
    }
    unsigned e = 0;
    for (; *p >= '0' && *p <= '9'; ++p) e = e * 10 + (unsigned)(*p - '0');
    if (e > 308) e = 308;
    for (; e >= 50; e -= 50) scale *= 1E50;
    for (; e >= 8; e -= 8) scale *= 1E8;
    for (; e > 0; --e) scale *= 10.0;
  }
  return sign * (neg ? value / scale : value * scale);
}

enum Tok {
  T_END, T_NAME, T_INT, T_DEC, T_FLOAT, T_STR, T_FILE, T_ASSIGN, T_EQUAL, T_NONEQ, T_CMP, T_OR, T_XOR,
  T_AND, T_NOT, T_LIKE, T_PUNCT, T_KW
};

struct Token {
  Tok t;
  std::string s;  // text (keywords upper-cased)
  long long i = 0;
  int sub = 0;    // comparison code
  char c = 0;     // punctuation
};

const char *KEYWORDS[] = {"LOAD", "FILTER", "DESC", "ASC", "ORDER", "BY", "JOIN", "STORE", "INTO", "GROUP",
                          "AS", "SELECT", "USING", "COUNT", "SUM", "AVG", "MIN", "MAX", "LIMIT", "ON", "BINARY",
                          "DISTINCT", "LEFT", "RIGHT", "OUTER", "SEMI", "ANTI", "SORT", "SEGMENTS", "PRESORTED",
                          "INSERT", "WHERE", "DISPLAY", "CASE", "WHEN", "THEN", "ELSE", "END", "SHOW", "TABLES",
                          "TABLE", "DESCRIBE", "DROP", "DELETE", "FROM", "CREATE", "INDEX", "YEAR", "MONTH",
                          "DAY", "CAST_TO_INT", "APPEND", "NO", "ENCODING"};

struct Lexer {
  const char *p;
  std::string err;

  bool next(Token &k) {
    for (;;) {
      while (*p && isspace((unsigned char)*p)) ++p;
      if (p[0] == '-' && p[1] == '-') {  // line comment
        while (*p && *p != '\n') ++p;
        continue;
      }
      break;
    }
    k = Token();
    if (!*p) {
      k.t = T_END;
      return true;
    }
    const char *q = p;
    // numbers, with the optional leading '-' that is part of the token
    const char *d = q + (*q == '-' ? 1 : 0);
    if (isdigit((unsigned char)*d) || (*d == '.' && isdigit((unsigned char)d[1]))) {
      const char *e = d;
      while (isdigit((unsigned char)*e)) ++e;
      bool dot = false, ex = false;
      if (*e == '.') {
        dot = true;
        ++e;
        while (isdigit((unsigned char)*e)) ++e;
      }
      if ((*e == 'e' || *e == 'E') && (isdigit((unsigned char)e[1]) ||
                                       ((e[1] == '+' || e[1] == '-') && isdigit((unsigned char)e[2])))) {
        ex = true;
        e += 2;
        while (isdigit((unsigned char)*e)) ++e;
      }
      k.s.assign(q, e - q);
      k.t = ex ? T_FLOAT : dot ? T_DEC : T_INT;
      if (k.t == T_INT) k.i = atoll(k.s.c_str());
      p = e;
      return true;
    }
    if (*q == '"') {  // string literal, quotes kept (emit_string strips them)
      const char *e = q + 1;
      while (*e && *e != '"' && *e != '\n') e += (*e == '\\' && e[1]) ? 2 : 1;
      if (*e != '"') return fail("unterminated string");
      k.t = T_STR;
      k.s.assign(q, e + 1 - q);
      p = e + 1;
      return true;
    }
    if (*q == '\'') {  // file name or separator
      const char *e = q + 1;
      while (*e && *e != '\'' && *e != '\n') ++e;
      if (*e != '\'') return fail("unterminated file name");
      k.t = T_FILE;
      k.s.assign(q + 1, e - q - 1);
      p = e + 1;
      return true;
    }
    if (isalpha((unsigned char)*q)) {
      const char *e = q;
      while (isalnum((unsigned char)*e) || *e == '_') ++e;
      std::string w(q, e - q), up = w;
      for (auto &ch : up) ch = (char)toupper((unsigned char)ch);
      p = e;
      if (up == "AND") { k.t = T_AND; return true; }
      if (up == "OR") { k.t = T_OR; return true; }
      if (up == "XOR") { k.t = T_XOR; return true; }
      if (up == "NOT") { k.t = T_NOT; return true; }
      if (up == "LIKE") { k.t = T_LIKE; return true; }
      if (up == "MOD") { k.t = T_PUNCT; k.c = '%'; return true; }
      for (const char *kw : KEYWORDS)
        if (up == kw) {
          k.t = T_KW;
          k.s = up;
          return true;
        }
      k.t = T_NAME;
      k.s = w;
      return true;
    }
    static const struct { const char *op; Tok t; int sub; } OPS[] = {
        {":=", T_ASSIGN, 0}, {"&&", T_AND, 0}, {"||", T_OR, 0}, {"<=>", T_CMP, 12}, {"==", T_CMP, 4},
        {">=", T_CMP, 6},    {"<=", T_CMP, 5}, {"!=", T_CMP, 3}, {"<>", T_NONEQ, 0}, {">", T_CMP, 2},
        {"<", T_CMP, 1},     {"=", T_EQUAL, 0}};
    for (const auto &o : OPS)
      if (!strncmp(q, o.op, strlen(o.op))) {
        k.t = o.t;
        k.sub = o.sub;
        p = q + strlen(o.op);
        return true;
      }
    if (strchr("-+&~|^/%*(),.;:!{}", *q)) {
      k.t = T_PUNCT;
      k.c = *q;
      p = q + 1;
      return true;
    }
    ++p;  // the reference's lexer skips any other character
    return next(k);
  }

  bool fail(const char *m) {
    err = m;
    return false;
  }
};

struct Parser {
  Lexer lx;
  Token cur;
  std::string err;
  bool ok = true;

  explicit Parser(const char *s) { lx.p = s; }

  bool advance() {
    if (!lx.next(cur)) {
      error("%s", lx.err.c_str());
      return false;
    }
    return true;
  }
  void error(const char *fmt, ...) {
    if (!ok) return;
    ok = false;
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
  }
  bool kw(const char *w) const { return cur.t == T_KW && cur.s == w; }
  bool punct(char c) const { return cur.t == T_PUNCT && cur.c == c; }
  bool expect_kw(const char *w) {
    if (!kw(w)) {
      error("syntax error: expected %s", w);
      return false;
    }
    return advance();
  }
  bool expect_punct(char c) {
    if (!punct(c)) {
      error("syntax error: expected '%c'", c);
      return false;
    }
    return advance();
  }
  bool name(std::string &out) {
    if (cur.t != T_NAME) {
      error("syntax error: expected a name");
      return false;
    }
    out = cur.s;
    return advance();
  }

  // ---- expressions: precedence climbing over the grammar's table ----------------------
  static int prec(const Token &k, bool *right) {
    *right = false;
    switch (k.t) {
      case T_EQUAL: *right = true; return 1;
      case T_NONEQ: *right = true; return 2;
      case T_OR: return 3;
      case T_XOR: return 4;
      case T_AND: return 5;
      case T_LIKE: return 6;
      case T_CMP: return 8;
      case T_PUNCT:
        if (k.c == '+' || k.c == '-') return 9;
        if (k.c == '*' || k.c == '/' || k.c == '%') return 10;
        return 0;
      default: return 0;
    }
  }

  bool emit_binary(const Token &op) {
    switch (op.t) {
      case T_EQUAL: emit_eq(); return true;
      case T_OR: emit_or(); return true;
      case T_AND: emit_and(); return true;
      case T_CMP:
        if (op.sub == 12) break;
        emit_cmp(op.sub);
        return true;
      case T_LIKE: emit_cmp(7); return true;
      case T_PUNCT:
        if (op.c == '+') { emit_add(); return true; }
        if (op.c == '-') { emit_minus(); return true; }
        if (op.c == '*') { emit_mul(); return true; }
        if (op.c == '/') { emit_div(); return true; }
        break;
      default: break;
    }
    error("operator not supported by this build");
    return false;
  }

  bool primary() {
    if (cur.t == T_NOT || punct('!')) {
      error("NOT is not supported by this build");
      return false;
    }
    if (punct('(')) {
      if (!advance() || !expr(1)) return false;
      return expect_punct(')');
    }
    if (cur.t == T_INT) {
      emit_number(cur.i);
      return advance();
    }
    if (cur.t == T_DEC) {
      emit_decimal(cur.s.c_str());
      return advance();
    }
    if (cur.t == T_STR) {
      emit_string(cur.s.c_str());
      return advance();
    }
    if (cur.t == T_FLOAT) {
      emit_float(approx_value(cur.s.c_str()));
      return advance();
    }
    if (cur.t == T_KW && (cur.s == "COUNT" || cur.s == "SUM" || cur.s == "AVG" || cur.s == "MIN" || cur.s == "MAX")) {
      const std::string f = cur.s;
      if (!advance() || !expect_punct('(') || !expr(1) || !expect_punct(')')) return false;
      if (f == "COUNT") emit_count();
      else if (f == "SUM") emit_sum();
      else if (f == "AVG") emit_average();
      else if (f == "MIN") emit_min();
      else emit_max();
      return true;
    }
    if (cur.t == T_NAME) {
      const std::string n = cur.s;
      if (!advance()) return false;
      if (punct('.')) {
        error("table.column names are not supported by this build");
        return false;
      }
      if (kw("ASC")) {
        emit_var_asc(n.c_str());
        return advance();
      }
      if (kw("DESC")) {
        emit_var_desc(n.c_str());
        return advance();
      }
      emit_name(n.c_str());
      return true;
    }
    error("syntax error in expression");
    return false;
  }

  bool expr(int minp) {
    if (!primary()) return false;
    for (;;) {
      bool right;
      const int p = prec(cur, &right);
      if (p == 0 || p < minp) return true;
      const Token op = cur;
      if (!advance() || !expr(right ? p : p + 1)) return false;
      if (!emit_binary(op)) return false;
    }
  }

  // ---- statements ------------------------------------------------------------------------
  int opt_limit() {
    if (!kw("LIMIT")) return 1;
    if (!advance()) return 0;
    if (cur.t != T_INT) {
      error("syntax error: LIMIT needs a number");
      return 0;
    }
    emit_limit((int)cur.i);
    return advance() ? 1 : 0;
  }

  bool using_sep(std::string &sep) {
    if (!expect_kw("USING") || !expect_punct('(')) return false;
    if (cur.t != T_FILE) {
      error("syntax error: expected a separator");
      return false;
    }
    sep = cur.s;
    return advance() && expect_punct(')');
  }

  // one statement; false on error (message in err or in alenka_last_error)
  bool statement() {
    if (kw("STORE")) {
      std::string v, file, sep;
      if (!advance() || !name(v) || !expect_kw("INTO")) return false;
      if (cur.t != T_FILE) {
        error("syntax error: expected a file name");
        return false;
      }
      file = cur.s;
      if (!advance()) return false;
      if (!kw("USING")) {
        // STORE V INTO 'f' [APPEND] [LIMIT n] BINARY [sort_def] (bison.y:148-151)
        bool append = false;
        if (kw("APPEND")) {
          append = true;
          if (!advance()) return false;
        }
        if (!opt_limit() || !expect_kw("BINARY")) return false;
        if (kw("SORT") || kw("PRESORTED")) {
          error("binary STORE with a sort order is not supported by this build");
          return false;
        }
        emit_store_binary(v.c_str(), file.c_str(), append ? 1 : 0);
        return true;
      }
      if (!using_sep(sep) || !opt_limit()) return false;
      emit_store(v.c_str(), file.c_str(), sep.c_str());
      return true;
    }
    if (kw("DISPLAY")) {
      std::string v, sep;
      if (!advance() || !name(v) || !using_sep(sep) || !opt_limit()) return false;
      emit_display(v.c_str(), sep.c_str());
      return true;
    }
    if (kw("DROP")) {
      std::string v;
      if (!advance() || !expect_kw("TABLE") || !name(v)) return false;
      alenka_drop(v.c_str());
      return true;
    }
    if (cur.t != T_NAME) {
      error("statement not supported by this build");
      return false;
    }
    std::string v;
    if (!name(v)) return false;
    if (cur.t != T_ASSIGN) {
      error("syntax error: expected :=");
      return false;
    }
    if (!advance()) return false;
    if (kw("FILTER")) {
      std::string t;
      if (!advance() || !name(t) || !expect_kw("BY") || !expr(1)) return false;
      emit_filter((char *)v.c_str(), (char *)t.c_str());
      return true;
    }
    if (kw("ORDER")) {
      std::string t;
      if (!advance() || !name(t) || !expect_kw("BY")) return false;
      int n = 0;
      while (ok) {
        if (!expr(1)) return false;
        ++n;
        if (!punct(',')) break;
        if (!advance()) return false;
      }
      emit_order(v.c_str(), t.c_str(), n, 0);
      return true;
    }
    if (kw("SELECT")) {
      if (!advance()) return false;
      if (punct('*')) {
        error("SELECT * is not supported by this build");
        return false;
      }
      while (ok) {
        std::string as;
        if (!expr(1) || !expect_kw("AS") || !name(as)) return false;
        emit_sel_name(as.c_str());
        if (!punct(',')) break;
        if (!advance()) return false;
      }
      std::string t;
      if (!expect_kw("FROM") || !name(t)) return false;
      // joins: the grammar's join_list is right-recursive, so the tables are handed over
      // innermost (last) first, after all the ON expressions
      std::vector<std::string> joins;
      std::vector<std::string> kinds;  // the words before JOIN: "", "LEFT SEMI", "RIGHT", ...
      while (kw("JOIN") || kw("LEFT") || kw("RIGHT") || kw("OUTER")) {
        std::string k;
        while (!kw("JOIN")) {
          if (!(kw("LEFT") || kw("RIGHT") || kw("OUTER") || kw("SEMI") || kw("ANTI"))) {
            error("JOIN expected");
            return false;
          }
          k += (k.empty() ? "" : " ") + cur.s;
          if (!advance()) return false;
        }
        std::string u;
        if (!advance() || !name(u) || !expect_kw("ON") || !expr(1)) return false;
        joins.push_back(u);
        kinds.push_back(k);
      }
      // join kinds as bison.y:260-275 queues them; a join followed by another one takes the
      // grammar's second set of rules, which queue LEFT as 'L' and RIGHT SEMI as 'R'
      for (size_t k = joins.size(); k-- > 0;) {
        const bool last = k + 1 == joins.size();
        const std::string &w = kinds[k];
        char tp;
        if (w.empty()) tp = 'I';
        else if (w == "LEFT ANTI") tp = '3';
        else if (w == "RIGHT ANTI") tp = '4';
        else if (w == "LEFT SEMI") tp = '1';
        else if (w == "RIGHT SEMI") tp = last ? '2' : 'R';
        else if (w == "LEFT") tp = last ? 'S' : 'L';
        else if (w == "RIGHT") tp = 'R';
        else if (w == "OUTER") tp = 'O';
        else {
          error("unknown join kind %s", w.c_str());
          return false;
        }
        emit_join_tab(joins[k].c_str(), tp);
      }
      int grp = 0;
      if (kw("GROUP")) {
        if (!advance() || !expect_kw("BY")) return false;
        while (ok) {
          if (!expr(1)) return false;
          ++grp;
          if (!punct(',')) break;
          if (!advance()) return false;
        }
      }
      if (joins.empty())
        emit_select(v.c_str(), t.c_str(), grp);
      else  // the grammar passes the join list's value (1) here, not the group count
        emit_join(v.c_str(), t.c_str(), 1, 0, -1);
      return true;
    }
    error("statement not supported by this build");
    return false;
  }
};

}  // namespace

extern "C" int alenkaExecute(char *s) {
  if (!s) return ALN_E_ARG;
  Parser ps(s);
  if (!ps.advance()) {
    process_error(2, ps.err.c_str());
    return ALN_E_ARG;
  }
  int n = 0;
  while (ps.cur.t != T_END) {
    clean_queues();
    const bool good = ps.statement();
    if (!good || !ps.ok) {
      char buf[320];
      snprintf(buf, sizeof(buf), "statement %d: %s", n + 1, ps.err.c_str());
      clean_queues();
      process_error(2, buf);
      return strstr(ps.err.c_str(), "not supported") ? ALN_E_UNSUPPORTED : ALN_E_ARG;
    }
    if (alenka_status() != ALN_OK) return alenka_status();  // the statement itself failed
    if (!ps.expect_punct(';')) {
      process_error(2, ps.err.c_str());
      return ALN_E_ARG;
    }
    ++n;
  }
  return ALN_OK;
}

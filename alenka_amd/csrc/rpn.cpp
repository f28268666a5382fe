// Host-side compilation of the reference's RPN queues into device programs, and the
// zone-map segment classifier.
//
// The compiler walks the RPN with the same stacks and pop order as filter.cu:133-643
// (predicates) and select.cu:703-790 (expressions), but where the reference materialises
// a device vector per operator it emits a program fragment.  The decimal scaling that
// CudaSet::op (cm.cu:2480-2655) and CudaSet::compare (cm.cu:2250-2379) apply, chosen by
// precision_func (zone_map.cu:668-712), is folded into constants and column multipliers,
// so the device evaluates each row in one pass with the reference's results.
#include <string.h>

#include <utility>
#include <vector>

#include "aln_internal.h"

namespace {

using aln::bits_f64;
using aln::f64_bits;
using Frag = std::vector<aln_insn>;
enum { T_NAME = 1, T_NUMBER = 2, T_STRING = 3, T_SET = 4, T_FLOAT = 5 };

int pw10(uint32_t p, int64_t *out) {
  static const int64_t t[10] = {1,      10,      100,      1000,      10000,
                                100000, 1000000, 10000000, 100000000, 1000000000};
  if (p > 9) return -1;
  *out = t[p];
  return 0;
}
int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
int64_t wdiv(int64_t a, int64_t b) {
  if (b == 0) return 0;
  if (a == INT64_MIN && b == -1) return INT64_MIN;
  return a / b;
}

uint32_t precision_func(uint32_t *p1, uint32_t *p2, int op) {
  if (op != ALN_RPN_DIV) {
    if (op != ALN_RPN_MUL) {
      uint32_t res;
      if (*p1 > *p2) {
        res = *p1;
        *p2 = *p1 - *p2;
        *p1 = 0;
      } else {
        res = *p1;
        *p1 = *p2 - *p1;
        *p2 = 0;
      }
      return res;
    }
    uint32_t res = *p1 + *p2;
    *p1 = 0;
    *p2 = 0;
    return res;
  }
  if (*p1 == *p2) {
    *p1 = *p1 + 4;
    *p2 = 0;
    return *p1;
  }
  if (*p1 > *p2) {
    *p1 = *p1 + (*p1 - *p2) + 4;
    *p2 = 0;
    return *p1;
  }
  *p2 = *p2 + (*p2 - *p1) + 4;
  *p1 = 0;
  return *p2;
}

int arith_insn(int rpn_op) {
  switch (rpn_op) {
    case ALN_RPN_ADD: return ALN_I_ADD;
    case ALN_RPN_MINUS: return ALN_I_SUB;
    case ALN_RPN_MUL: return ALN_I_MUL;
    default: return ALN_I_DIV;
  }
}
int64_t arith_fold(int rpn_op, int64_t l, int64_t r) {
  switch (rpn_op) {
    case ALN_RPN_ADD: return wadd(l, r);
    case ALN_RPN_MINUS: return wsub(l, r);
    case ALN_RPN_MUL: return wmul(l, r);
    default: return wdiv(l, r);
  }
}

aln_insn I(int op, int a, int64_t k) {
  aln_insn x;
  x.op = op;
  x.a = a;
  x.k = k;
  return x;
}

// value * s, folded into a lone COL instruction when possible
Frag scaled(const Frag &f, int64_t s) {
  Frag r = f;
  if (s == 1) return r;
  if (r.size() == 1 && r[0].op == ALN_I_COL) {
    r[0].k = wmul(r[0].k, s);
    return r;
  }
  if (r.size() == 1 && r[0].op == ALN_I_CONST) {
    r[0].k = wmul(r[0].k, s);
    return r;
  }
  r.push_back(I(ALN_I_CONST, 0, s));
  r.push_back(I(ALN_I_MUL, 0, 0));
  return r;
}
Frag cat(const Frag &a, const Frag &b, aln_insn op) {
  Frag r = a;
  r.insert(r.end(), b.begin(), b.end());
  r.push_back(op);
  return r;
}

struct Compiler {
  const int32_t *dec;
  int ncols;
  int mode = 0;  // 0: predicate (filter.cu), 1: expression (select.cu)
  std::vector<int> type, value;
  std::vector<int64_t> nums, strs;
  std::vector<std::pair<int64_t, int32_t>> sets;  // LIKE operands: bit set address, codes
  std::vector<uint32_t> prec;
  std::vector<Frag> vecs, bools;
  std::vector<bool> vfloat;   // per vecs entry: the vector holds doubles
  std::vector<double> fnums;  // FLOAT operands
  int err = ALN_OK;
  const char *why = "";

  template <typename T>
  T pop(std::vector<T> &v) {
    if (v.empty()) {
      err = ALN_E_ARG;
      why = "RPN stack underflow";
      return T();
    }
    T x = v.back();
    v.pop_back();
    return x;
  }
  Frag get_vec(int v) {
    if (v >= 0) return Frag{I(ALN_I_COL, v, 1)};
    if (!vfloat.empty()) vfloat.pop_back();
    return pop(vecs);
  }
  void push_vec(const Frag &f, bool fl) {
    vecs.push_back(f);
    vfloat.push_back(fl);
  }
  bool float_col(int v) const { return v >= 0 && v < ncols && dec[v] == ALN_DEC_F64; }

  // ---- double precision: CudaSet's float overloads (cm.cu:2178-2437, 2440-2773) ----------
  // Whether either of the top two operands is a double (a FLOAT, an f64 column or a value
  // computed from one), read without popping.
  bool float_operands() const {
    const size_t nt = type.size();
    if (nt < 2) return false;
    size_t vi = value.size(), ci = vecs.size();
    for (int k = 0; k < 2; ++k) {
      const int t = type[nt - 1 - k];
      if (t == T_FLOAT) return true;
      if (t == T_NAME || t == T_STRING || t == T_SET) {
        if (vi == 0) return false;
        const int v = value[--vi];
        if (t != T_NAME) continue;
        if (v >= 0) {
          if (float_col(v)) return true;
        } else if (ci > 0 && vfloat[--ci]) {
          return true;
        }
      }
    }
    return false;
  }
  struct FOp {
    Frag f;
    bool is_const = false;
    double c = 0;
  };
  // One operand as a double: an int column by a plain cast of its stored integer
  // (long_to_float_type, cm.cu:137-144), a NUMBER by its decimal value.
  FOp fpop(int t) {
    FOp o;
    if (t == T_FLOAT) {
      o.is_const = true;
      o.c = pop(fnums);
    } else if (t == T_NUMBER) {
      const int64_t n = pop(nums);
      const uint32_t p = pop(prec);
      if (p > 18) {
        fail(ALN_E_UNSUPPORTED, "decimal scale above 10^18");
        return o;
      }
      double d = 1;
      for (uint32_t i = 0; i < p; ++i) d *= 10;
      o.is_const = true;
      o.c = (double)n / d;
    } else if (t == T_NAME) {
      const int v = pop(value);
      if (v >= 0) {
        o.f = Frag{I(ALN_I_COL, v, 1)};
        if (!float_col(v)) o.f.push_back(I(ALN_I_I2F, 0, 0));
      } else {
        if (vecs.empty()) {
          fail(ALN_E_UNSUPPORTED, "a comparison result used as a value");
          return o;
        }
        const bool fl = vfloat.back();
        o.f = get_vec(v);
        (void)pop(prec);
        if (!fl) o.f.push_back(I(ALN_I_I2F, 0, 0));
      }
      return o;
    } else {
      fail(ALN_E_UNSUPPORTED, "a string operand in double arithmetic");
      return o;
    }
    o.f = Frag{I(ALN_I_CONST, 0, f64_bits(o.c))};
    return o;
  }
  void farith(int op) {
    const int s1 = pop(type), s2 = pop(type);
    FOp r = fpop(s1), l = fpop(s2);
    if (err) return;
    if (l.is_const && r.is_const) {  // folded, as the reference folds NUMBER NUMBER
      double v = op == ALN_RPN_ADD ? l.c + r.c : op == ALN_RPN_MINUS ? l.c - r.c
               : op == ALN_RPN_MUL ? l.c * r.c : l.c / r.c;
      fnums.push_back(v);
      type.push_back(T_FLOAT);
      return;
    }
    const int io = op == ALN_RPN_ADD ? ALN_I_FADD : op == ALN_RPN_MINUS ? ALN_I_FSUB
                 : op == ALN_RPN_MUL ? ALN_I_FMUL : ALN_I_FDIV;
    prec.push_back(0);
    type.push_back(T_NAME);
    value.push_back(-1);
    push_vec(cat(l.f, r.f, I(io, 0, 0)), true);
  }
  void fcmp(int code) {
    const int s1 = pop(type), s2 = pop(type);
    FOp r = fpop(s1), l = fpop(s2);
    if (err) return;
    type.push_back(T_NAME);
    bools.push_back(cat(l.f, r.f, I(ALN_I_FCMP, code, 0)));
  }
  uint32_t get_decimals(int v) {
    if (v >= 0) return (uint32_t)dec[v];
    return pop(prec);
  }
  void fail(int code, const char *w) {
    if (err == ALN_OK) {
      err = code;
      why = w;
    }
  }

  // CudaSet::op(int_type*, int_type d, op, reverse, p1, p2), cm.cu:2480-2546
  Frag op_col_const(const Frag &t, int64_t d, int op, int reverse, uint32_t p1, uint32_t p2) {
    int64_t d1 = d, s1 = 1, s2 = 1;
    if (pw10(p1, &s1) || pw10(p2, &s2)) {
      fail(ALN_E_UNSUPPORTED, "decimal scale above 10^9");
      return t;
    }
    if (p2) d = wmul(d, s2);
    int io = arith_insn(op);
    if (!reverse) {
      if (op == ALN_RPN_MUL) return cat(t, Frag{I(ALN_I_CONST, 0, d1)}, I(io, 0, 0));
      if (!p1) return cat(t, Frag{I(ALN_I_CONST, 0, wmul(d, s2))}, I(io, 0, 0));
      return cat(scaled(t, s1), Frag{I(ALN_I_CONST, 0, d)}, I(io, 0, 0));
    }
    return cat(Frag{I(ALN_I_CONST, 0, d)}, p1 ? scaled(t, s1) : t, I(io, 0, 0));
  }

  // CudaSet::op(int_type*, int_type*, op, 0, p1, p2), cm.cu:2552-2605
  Frag op_col_col(const Frag &c1, const Frag &c2, int op, uint32_t p1, uint32_t p2) {
    int64_t s1 = 1, s2 = 1;
    if (pw10(p1, &s1) || pw10(p2, &s2)) {
      fail(ALN_E_UNSUPPORTED, "decimal scale above 10^9");
      return c1;
    }
    int io = arith_insn(op);
    if (op == ALN_RPN_MUL) return cat(c1, c2, I(io, 0, 0));
    return cat(p1 ? scaled(c1, s1) : c1, p2 ? scaled(c2, s2) : c2, I(io, 0, 0));
  }

  void arith(int op) {
    if (float_operands()) return farith(op);
    int s1 = pop(type), s2 = pop(type);
    if (err) return;
    if (s1 == T_NUMBER && s2 == T_NUMBER) {
      int64_t n1 = pop(nums), n2 = pop(nums);
      uint32_t p1 = pop(prec), p2 = pop(prec);
      uint32_t pres = precision_func(&p1, &p2, op);
      int64_t a, b;
      if (pw10(p1, &a) || pw10(p2, &b)) return fail(ALN_E_UNSUPPORTED, "decimal scale");
      prec.push_back(pres);
      if (p1) n1 = wmul(n1, a);
      if (p2) n2 = wmul(n2, b);
      type.push_back(T_NAME);
      value.push_back(-1);
      push_vec(Frag{I(ALN_I_CONST, 0, arith_fold(op, n1, n2))}, false);
    } else if (s1 == T_NAME && s2 == T_NUMBER) {
      int v1 = pop(value);
      int64_t n1 = pop(nums);
      uint32_t p1 = pop(prec);
      uint32_t p2 = get_decimals(v1);
      Frag t = get_vec(v1);
      // filter.cu:285-289 passes (p1, p2); select.cu:754-760 passes (p2, p1)
      uint32_t pres = mode == 0 ? precision_func(&p1, &p2, op) : precision_func(&p2, &p1, op);
      prec.push_back(pres);
      type.push_back(T_NAME);
      value.push_back(-1);
      push_vec(mode == 0 ? op_col_const(t, n1, op, 1, p1, p2) : op_col_const(t, n1, op, 1, p2, p1), false);
    } else if (s1 == T_NUMBER && s2 == T_NAME) {
      int64_t n1 = pop(nums);
      int v2 = pop(value);
      uint32_t p1 = pop(prec);
      uint32_t p2 = get_decimals(v2);
      Frag t = get_vec(v2);
      uint32_t pres = precision_func(&p2, &p1, op);
      prec.push_back(pres);
      type.push_back(T_NAME);
      value.push_back(-1);
      push_vec(op_col_const(t, n1, op, 0, p2, p1), false);
    } else if (s1 == T_NAME && s2 == T_NAME) {
      int v1 = pop(value), v2 = pop(value);
      Frag t1 = get_vec(v1), t = get_vec(v2);
      uint32_t p1 = get_decimals(v1), p2 = get_decimals(v2);
      uint32_t pres = precision_func(&p1, &p2, op);
      prec.push_back(pres);
      type.push_back(T_NAME);
      value.push_back(-1);
      push_vec(op_col_col(t, t1, op, p2, p1), false);
    } else {
      fail(ALN_E_UNSUPPORTED, "arithmetic on a string operand");
    }
  }

  // CudaSet::compare(int_type* column1, d, op, p1, p2), cm.cu:2250-2295
  Frag cmp_col_const(const Frag &t, int64_t d, int code, uint32_t p1, uint32_t p2) {
    int64_t s1 = 1, s2 = 1;
    if (pw10(p1, &s1) || pw10(p2, &s2)) {
      fail(ALN_E_UNSUPPORTED, "decimal scale above 10^9");
      return t;
    }
    if (p2) d = wmul(d, s2);
    return cat(p1 ? scaled(t, s1) : t, Frag{I(ALN_I_CONST, 0, d)}, I(ALN_I_CMP, code, 0));
  }

  void cmp(int code) {
    if (code >= 1 && code <= 6 && float_operands()) return fcmp(code);
    int s1 = pop(type), s2 = pop(type);
    if (err) return;
    if (code == 7 || s1 == T_SET || s2 == T_SET) {
      // LIKE (filter.cu:452-499): the pattern was matched against the column's dictionary
      // when the query was bound; the row test is membership of the row's code in that set
      if (code != 7 || !((s1 == T_SET && s2 == T_NAME) || (s1 == T_NAME && s2 == T_SET)))
        return fail(ALN_E_UNSUPPORTED, "LIKE needs a char column and a pattern");
      int v1 = pop(value), v2 = pop(value);
      int col = (s1 == T_NAME) ? v1 : v2;
      int sv = (s1 == T_SET) ? v1 : v2;
      if (col < 0) return fail(ALN_E_UNSUPPORTED, "LIKE on a computed value");
      if (sv > -2 || -2 - sv >= (int)sets.size()) return fail(ALN_E_ARG, "LIKE operand missing");
      const auto &st = sets[-2 - sv];
      bools.push_back(Frag{I(ALN_I_COL, col, 1), I(ALN_I_INSET, st.second, st.first)});
      type.push_back(T_NAME);
      return;
    }
    if (code < 1 || code > 6) return fail(ALN_E_UNSUPPORTED, "comparison code outside 1..6");
    if (s1 == T_FLOAT || s2 == T_FLOAT) return fail(ALN_E_UNSUPPORTED, "FLOAT compared with a string");
    if (s1 == T_NUMBER && s2 == T_NUMBER) {
      int64_t n1 = pop(nums), n2 = pop(nums);
      type.push_back(T_NAME);
      value.push_back(-1);
      uint32_t p1 = pop(prec), p2 = pop(prec);
      uint32_t pres = precision_func(&p1, &p2, 0);
      int64_t a, b;
      if (pw10(p1, &a) || pw10(p2, &b)) return fail(ALN_E_UNSUPPORTED, "decimal scale");
      prec.push_back(pres);
      if (p1) n1 = wmul(n1, a);
      if (p2) n2 = wmul(n2, b);
      // CudaSet::compare(s = n1, d = n2) tests d OP s (cm.cu:2132-2175)
      bools.push_back(Frag{I(ALN_I_CONST, 0, n2), I(ALN_I_CONST, 0, n1), I(ALN_I_CMP, code, 0)});
    } else if ((s1 == T_STRING && s2 == T_NAME) || (s1 == T_NAME && s2 == T_STRING)) {
      int v1 = pop(value), v2 = pop(value);
      int col = (s1 == T_NAME) ? v1 : v2;
      int sv = (s1 == T_STRING) ? v1 : v2;
      int64_t lit = (sv <= -2 && -2 - sv < (int)strs.size()) ? strs[-2 - sv] : -1;
      if (code != 3 && code != 4) return fail(ALN_E_UNSUPPORTED, "string comparison other than == / !=");
      if (col < 0) return fail(ALN_E_UNSUPPORTED, "string compared with a computed value");
      bools.push_back(Frag{I(ALN_I_COL, col, 1), I(ALN_I_CONST, 0, lit), I(ALN_I_CMP, code, 0)});
      type.push_back(T_NAME);
    } else if (s1 == T_NUMBER && s2 == T_NAME) {
      int64_t n1 = pop(nums);
      int v = pop(value);
      Frag t = get_vec(v);
      uint32_t p2 = pop(prec);
      uint32_t p1 = get_decimals(v);
      uint32_t pres = p1 > p2 ? p1 : p2;
      prec.push_back(pres);
      type.push_back(T_NAME);
      bools.push_back(cmp_col_const(t, n1, code, pres - p1, pres - p2));
    } else if (s1 == T_NAME && s2 == T_NUMBER) {
      static const int rev[7] = {0, 2, 1, 3, 4, 6, 5};
      code = rev[code];
      int64_t n1 = pop(nums);
      int v = pop(value);
      Frag t = get_vec(v);
      uint32_t p2 = pop(prec);
      uint32_t p1 = get_decimals(v);
      uint32_t pres = p1 > p2 ? p1 : p2;
      prec.push_back(pres);
      type.push_back(T_NAME);
      bools.push_back(cmp_col_const(t, n1, code, p1, p2));
    } else if (s1 == T_NAME && s2 == T_NAME) {
      int v1 = pop(value), v2 = pop(value);
      type.push_back(T_NAME);
      Frag t = get_vec(v1), t1 = get_vec(v2);
      uint32_t p1 = get_decimals(v1), p2 = get_decimals(v2);
      uint32_t pres = p1 > p2 ? p1 : p2;
      prec.push_back(pres);
      // compare(t1, t, op, p2, p1) (cm.cu:2299-2379) reads a scaled right operand past its
      // end, so a right operand with decimals is rejected.
      if (p1) return fail(ALN_E_UNSUPPORTED, "column comparison with a scaled right operand");
      int64_t a;
      if (pw10(p2, &a)) return fail(ALN_E_UNSUPPORTED, "decimal scale");
      bools.push_back(cat(p2 ? scaled(t1, a) : t1, t, I(ALN_I_CMP, code, 0)));
    } else {
      fail(ALN_E_UNSUPPORTED, "unsupported comparison operands");
    }
  }

  int run(const aln_rpn *rpn, int nrpn, int mode_) {
    mode = mode_;
    for (int i = 0; i < nrpn && !err; ++i) {
      const aln_rpn &o = rpn[i];
      switch (o.type) {
        case ALN_RPN_NAME:
          if (o.arg < 0 || o.arg >= ncols) {
            fail(ALN_E_ARG, "column slot out of range");
            break;
          }
          value.push_back(o.arg);
          type.push_back(T_NAME);
          break;
        case ALN_RPN_NUMBER:
          nums.push_back(o.val);
          type.push_back(T_NUMBER);
          prec.push_back((uint32_t)o.arg);
          break;
        case ALN_RPN_FLOAT:
          fnums.push_back(bits_f64(o.val));
          type.push_back(T_FLOAT);
          break;
        case ALN_RPN_STRING:
          strs.push_back(o.val);
          value.push_back(-2 - (int)(strs.size() - 1));
          type.push_back(T_STRING);
          break;
        case ALN_RPN_CODESET:
          if (o.arg < 0 || (o.arg > 0 && o.val == 0)) {
            fail(ALN_E_ARG, "LIKE operand without a bit set");
            break;
          }
          sets.emplace_back(o.val, o.arg);
          value.push_back(-2 - (int)(sets.size() - 1));
          type.push_back(T_SET);
          break;
        case ALN_RPN_ADD:
        case ALN_RPN_MINUS:
        case ALN_RPN_MUL:
        case ALN_RPN_DIV:
          arith(o.type);
          break;
        case ALN_RPN_CMP:
          if (mode != 0) {
            fail(ALN_E_UNSUPPORTED, "comparison in an expression");
            break;
          }
          cmp(o.arg);
          break;
        case ALN_RPN_AND:
        case ALN_RPN_OR: {
          if (mode != 0 || bools.size() < 2) {
            fail(ALN_E_ARG, "AND/OR needs two predicates");
            break;
          }
          Frag b3 = pop(bools), b2 = pop(bools);
          type.push_back(T_NAME);
          bools.push_back(cat(b2, b3, I(o.type == ALN_RPN_AND ? ALN_I_AND : ALN_I_OR, 0, 0)));
          break;
        }
        default:
          fail(ALN_E_ARG, "unknown RPN node type");
      }
    }
    return err;
  }
};

int stack_depth(const Frag &f) {
  int d = 0, mx = 0;
  for (auto &x : f) {
    if (x.op == ALN_I_COL || x.op == ALN_I_CONST)
      ++d;
    else if (x.op != ALN_I_INSET && x.op != ALN_I_I2F)
      --d;
    if (d > mx) mx = d;
    if (d < 1) return -1;
  }
  return d == 1 ? mx : -1;
}

}  // namespace

using namespace aln;

extern "C" int aln_compile_rpn(const aln_rpn *rpn, int nrpn, const int32_t *col_dec, int ncols,
                               int mode, aln_insn *out, int *nout, int32_t *out_prec) {
  ALN_REQUIRE(rpn && out && nout && nrpn > 0, "aln_compile_rpn: bad arguments");
  ALN_REQUIRE(ncols >= 0 && ncols <= ALN_MAX_COLS && (ncols == 0 || col_dec),
              "aln_compile_rpn: ncols %d", ncols);
  Compiler c;
  c.dec = col_dec;
  c.ncols = ncols;
  int r = c.run(rpn, nrpn, mode);
  if (r != ALN_OK) return set_error(r, "aln_compile_rpn: %s", c.why);
  Frag res;
  int32_t prec = 0;
  if (mode == 0) {
    ALN_REQUIRE(!c.bools.empty(), "aln_compile_rpn: predicate yields no boolean");
    res = c.bools.back();
  } else {
    if (!c.type.empty() && c.type.back() == T_FLOAT && !c.fnums.empty()) {
      res = Frag{I(ALN_I_CONST, 0, f64_bits(c.fnums.back()))};
      prec = ALN_DEC_F64;
    } else if (!c.type.empty() && c.type.back() == T_NUMBER && !c.nums.empty()) {
      res = Frag{I(ALN_I_CONST, 0, c.nums.back())};
      prec = c.prec.empty() ? 0 : (int32_t)c.prec.back();
    } else if (!c.value.empty() && c.value.back() >= 0) {
      res = Frag{I(ALN_I_COL, c.value.back(), 1)};
      prec = col_dec[c.value.back()];
    } else if (!c.vecs.empty()) {
      res = c.vecs.back();
      prec = c.vfloat.back() ? ALN_DEC_F64 : c.prec.empty() ? 0 : (int32_t)c.prec.back();
    } else {
      return set_error(ALN_E_ARG, "aln_compile_rpn: expression yields no value");
    }
  }
  int depth = stack_depth(res);
  if ((int)res.size() > ALN_MAX_INSN || depth < 0 || depth > ALN_MAX_STACK)
    return set_error(ALN_E_UNSUPPORTED, "aln_compile_rpn: program too large (%zu insns, depth %d)",
                     res.size(), depth);
  memcpy(out, res.data(), res.size() * sizeof(aln_insn));
  *nout = (int)res.size();
  if (out_prec) *out_prec = prec;
  return ALN_OK;
}

// ---- zone maps: zone_map_check (zone_map.cu:720-1135) ----------------------------------
namespace {
char hc_const(int64_t s, int64_t d, int op) {  // host_compare(int_type, int_type, op)
  char res = 'N';
  if (op == 2 && d > s)
    res = 'A';
  else if (op == 1 && d < s)
    res = 'A';
  else if (op == 6 && d >= s)
    res = 'A';
  else if (op == 5 && d <= s)
    res = 'A';
  else if (op == 4 && d == s)
    res = 'A';
  else if (d != s)
    res = 'A';
  return res;
}
char hc_range(const int64_t *c, int64_t d, int op) {  // host_compare(int_type*, int_type, op)
  char res = 'R';
  if (op == 2) {
    if (c[1] <= d)
      res = 'N';
    else if (c[0] > d)
      res = 'A';
  } else if (op == 1) {
    if (c[0] >= d)
      res = 'N';
    else if (c[1] < d)
      res = 'A';
  } else if (op == 6) {
    if (c[1] < d)
      res = 'N';
    else if (c[0] >= d)
      res = 'A';
  } else if (op == 5) {
    if (c[0] > d)
      res = 'N';
    else if (c[1] <= d)
      res = 'A';
  } else if (op == 4 && c[0] == d && c[1] == d) {
    res = 'A';
  }
  return res;
}
}  // namespace

// For int/decimal columns.  Min/max are scaled in place across the predicate as
// get_host_vec does in the reference.  Predicates with arithmetic classify as 'R': the
// reference's interval arithmetic (host_op, zone_map.cu:513-560) can misclassify them.
extern "C" int aln_zone_map_check(const aln_rpn *rpn, int nrpn, const int64_t *seg_min,
                                  const int64_t *seg_max, const int32_t *col_dec, int ncols,
                                  int not_compressed, char *out) {
  ALN_REQUIRE(rpn && out && nrpn > 0, "aln_zone_map_check: bad arguments");
  ALN_REQUIRE(ncols >= 0 && ncols <= ALN_MAX_COLS, "aln_zone_map_check: ncols %d", ncols);
  if (not_compressed) {
    *out = 'R';
    return ALN_OK;
  }
  for (int i = 0; i < nrpn; ++i) {
    // arithmetic, FLOAT operands and f64 columns: evaluated on every segment
    if ((rpn[i].type >= ALN_RPN_ADD && rpn[i].type <= ALN_RPN_DIV) || rpn[i].type == ALN_RPN_FLOAT ||
        (rpn[i].type == ALN_RPN_NAME && rpn[i].arg >= 0 && rpn[i].arg < ncols && col_dec &&
         col_dec[rpn[i].arg] == ALN_DEC_F64)) {
      *out = 'R';
      return ALN_OK;
    }
    if (rpn[i].type == ALN_RPN_NAME && (rpn[i].arg < 0 || rpn[i].arg >= ncols))
      return set_error(ALN_E_ARG, "aln_zone_map_check: column slot %d", rpn[i].arg);
  }
  int64_t mm[ALN_MAX_COLS][2];
  for (int c = 0; c < ncols; ++c) {
    mm[c][0] = seg_min[c];
    mm[c][1] = seg_max[c];
  }
  std::vector<int> type, val;
  std::vector<int64_t> nums;
  std::vector<uint32_t> prec;
  std::vector<char> bools;
  auto under = [&]() { return set_error(ALN_E_ARG, "aln_zone_map_check: stack underflow"); };
  for (int i = 0; i < nrpn; ++i) {
    const aln_rpn &o = rpn[i];
    if (o.type == ALN_RPN_NAME) {
      type.push_back(T_NAME);
      val.push_back(o.arg);
    } else if (o.type == ALN_RPN_NUMBER) {
      type.push_back(T_NUMBER);
      nums.push_back(o.val);
      prec.push_back((uint32_t)o.arg);
    } else if (o.type == ALN_RPN_STRING || o.type == ALN_RPN_CODESET) {
      // a char comparison or LIKE is evaluated on every segment (zone_map.cu:1018-1044)
      type.push_back(T_STRING);
      val.push_back(-1);
    } else if (o.type == ALN_RPN_CMP) {
      int code = o.arg;
      if (type.size() < 2) return under();
      int s1 = type.back();
      type.pop_back();
      int s2 = type.back();
      type.pop_back();
      if (s1 == T_NUMBER && s2 == T_NUMBER) {
        if (nums.size() < 2 || prec.size() < 2) return under();
        int64_t n1 = nums.back();
        nums.pop_back();
        int64_t n2 = nums.back();
        nums.pop_back();
        uint32_t p1 = prec.back();
        prec.pop_back();
        uint32_t p2 = prec.back();
        prec.pop_back();
        uint32_t pres = p1 > p2 ? p1 : p2;
        int64_t a, b;
        if (pw10(pres - p1, &a) || pw10(pres - p2, &b))
          return set_error(ALN_E_UNSUPPORTED, "aln_zone_map_check: decimal scale");
        prec.push_back(pres);
        type.push_back(T_NAME);
        val.push_back(-1);
        if (p1) n1 = wmul(n1, a);
        if (p2) n2 = wmul(n2, b);
        bools.push_back(hc_const(n1, n2, code));
      } else if ((s1 == T_STRING && s2 == T_NAME) || (s1 == T_NAME && s2 == T_STRING)) {
        if (val.size() < 2) return under();
        val.pop_back();
        val.pop_back();
        bools.push_back('R');
        type.push_back(T_NAME);
        val.push_back(-1);
      } else if ((s1 == T_NUMBER && s2 == T_NAME) || (s1 == T_NAME && s2 == T_NUMBER)) {
        if (s1 == T_NAME) {
          static const int rev[7] = {0, 2, 1, 3, 4, 6, 5};
          if (code >= 1 && code <= 6) code = rev[code];
        }
        if (nums.empty() || val.empty() || prec.empty()) return under();
        int64_t n1 = nums.back();
        nums.pop_back();
        int c = val.back();
        val.pop_back();
        uint32_t p2 = prec.back();
        prec.pop_back();
        if (c < 0) {
          *out = 'R';
          return ALN_OK;
        }
        uint32_t p1 = (uint32_t)col_dec[c];
        uint32_t pres = p1 > p2 ? p1 : p2;
        int64_t a, b;
        if (pw10(pres - p1, &a) || pw10(pres - p2, &b))
          return set_error(ALN_E_UNSUPPORTED, "aln_zone_map_check: decimal scale");
        prec.push_back(pres);
        if (p1) {
          mm[c][0] = wmul(mm[c][0], a);
          mm[c][1] = wmul(mm[c][1], a);
        }
        if (p2) n1 = wmul(n1, b);
        type.push_back(T_NAME);
        val.push_back(-1);
        bools.push_back(hc_range(mm[c], n1, code));
      } else {
        *out = 'R';
        return ALN_OK;
      }
    } else if (o.type == ALN_RPN_AND || o.type == ALN_RPN_OR) {
      if (bools.size() < 2) return under();
      char b3 = bools.back();
      bools.pop_back();
      char b2 = bools.back();
      bools.pop_back();
      type.push_back(T_NAME);
      char r;
      if (o.type == ALN_RPN_AND)
        r = (b2 == 'A' && b3 == 'A') ? 'A' : (b2 == 'N' || b3 == 'N') ? 'N' : 'R';
      else
        r = (b2 == 'A' && b3 == 'A') ? 'A' : (b2 == 'N' && b3 == 'N') ? 'N' : 'R';
      bools.push_back(r);
    } else {
      return set_error(ALN_E_ARG, "aln_zone_map_check: operation %d is not valid", o.type);
    }
  }
  if (bools.empty()) return set_error(ALN_E_ARG, "aln_zone_map_check: no predicate");
  *out = bools.back();
  return ALN_OK;
}

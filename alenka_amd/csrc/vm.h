// Device interpreter of compiled RPN programs (aln_insn).  The program is read from
// device memory at a wave-uniform index, so instructions come through scalar loads; the
// value stack lives in LDS, one column of VM_BLOCK entries per depth level.
#pragma once

#include "aln_internal.h"

namespace aln {

constexpr int VM_BLOCK = 256;

struct ColTable {
  const uint32_t *words[ALN_MAX_COLS];
  const aln_seg *segs[ALN_MAX_COLS];
};

__device__ __forceinline__ int64_t vm_div(int64_t l, int64_t r) {
  if (r == 0) return 0;
  if (l == INT64_MIN && r == -1) return INT64_MIN;
  return l / r;
}

// stk points at this thread's entry of depth 0; depth d is stk[d * VM_BLOCK].
template <typename Load>
__device__ __forceinline__ int64_t vm_run(const aln_insn *__restrict__ prog, int n, Load load,
                                          int64_t *stk) {
  int sp = 0;
  for (int i = 0; i < n; ++i) {
    const int op = prog[i].op;
    if (op == ALN_I_COL) {
      stk[sp * VM_BLOCK] = wmul(load(prog[i].a), prog[i].k);
      ++sp;
    } else if (op == ALN_I_CONST) {
      stk[sp * VM_BLOCK] = prog[i].k;
      ++sp;
    } else if (op == ALN_I_INSET) {
      const int64_t v = stk[(sp - 1) * VM_BLOCK];
      const uint32_t *set = reinterpret_cast<const uint32_t *>(prog[i].k);
      stk[(sp - 1) * VM_BLOCK] = (v >= 0 && v < prog[i].a) ? (set[v >> 5] >> (v & 31)) & 1u : 0;
    } else if (op == ALN_I_I2F) {
      int64_t *e = &stk[(sp - 1 - prog[i].a) * VM_BLOCK];
      *e = f64_bits((double)*e);
    } else if (op >= ALN_I_FADD) {
      const double r = bits_f64(stk[(sp - 1) * VM_BLOCK]), l = bits_f64(stk[(sp - 2) * VM_BLOCK]);
      int64_t v;
      switch (op) {
        case ALN_I_FADD: v = f64_bits(l + r); break;
        case ALN_I_FSUB: v = f64_bits(l - r); break;
        case ALN_I_FMUL: v = f64_bits(l * r); break;
        case ALN_I_FDIV: v = f64_bits(l / r); break;
        default: v = fcmp_code(prog[i].a, l, r); break;
      }
      stk[(sp - 2) * VM_BLOCK] = v;
      --sp;
    } else {
      int64_t r = stk[(sp - 1) * VM_BLOCK], l = stk[(sp - 2) * VM_BLOCK], v;
      switch (op) {
        case ALN_I_ADD: v = (int64_t)((uint64_t)l + (uint64_t)r); break;
        case ALN_I_SUB: v = (int64_t)((uint64_t)l - (uint64_t)r); break;
        case ALN_I_MUL: v = wmul(l, r); break;
        case ALN_I_DIV: v = vm_div(l, r); break;
        case ALN_I_CMP: v = cmp_code(prog[i].a, l, r); break;
        case ALN_I_AND: v = (l != 0) && (r != 0); break;
        default: v = (l != 0) || (r != 0); break;
      }
      stk[(sp - 2) * VM_BLOCK] = v;
      --sp;
    }
  }
  return stk[0];
}

}  // namespace aln

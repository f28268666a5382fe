"""CPU tests of the C ABI: the library loads, exports every entry point the header
declares, and its host-side entry points (RPN compiler, zone maps, hashing) agree with the
oracle.  Compiled programs are run by a small Python interpreter of the instruction set."""
import os
import re

import math
import struct

import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "alenka_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(aln_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)
    assert b"gfx950" in lib.aln_version()


def test_no_oracle_in_product():
    """The product library neither links nor names the oracle."""
    data = open(L.LIB_PATH, "rb").read()
    assert b"liboracle" not in data and b"orc_" not in data


def f2b(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def b2f(b):
    return struct.unpack("<d", struct.pack("<q", b))[0]


def fcmp(code, x, y):
    e = 1e-8
    eq = (x - y) < e and (x - y) > -e
    return {1: (y - x) > e, 2: (x - y) > e, 3: (x - y) > e or (x - y) < -e, 4: eq,
            5: (y - x) > e or eq, 6: (x - y) > e or eq}[code]


def run_prog(prog, cols):
    """Reference interpreter of aln_insn programs (int64 wrap-around semantics; double
    instructions on the IEEE bits held in the stack slots)."""
    M = (1 << 64) - 1

    def wrap(x):
        x &= M
        return x - (1 << 64) if x >> 63 else x

    n = len(cols[0])
    out = np.empty(n, np.int64)
    for i in range(n):
        st = []
        for op, a, k in prog:
            if op == L.I_COL:
                st.append(wrap(int(cols[a][i]) * k))
            elif op == L.I_CONST:
                st.append(k)
            elif op == L.I_I2F:
                st[-1 - a] = f2b(float(st[-1 - a]))
            elif op >= L.I_FADD:
                r, l = b2f(st.pop()), b2f(st.pop())
                if op == L.I_FADD:
                    st.append(f2b(l + r))
                elif op == L.I_FSUB:
                    st.append(f2b(l - r))
                elif op == L.I_FMUL:
                    st.append(f2b(l * r))
                elif op == L.I_FDIV:
                    st.append(f2b(l / r if r != 0 else (math.copysign(math.inf, l) * math.copysign(1, r)
                                                        if l != 0 and not math.isnan(l) else math.nan)))
                else:
                    st.append(int(fcmp(a, l, r)))
            else:
                r, l = st.pop(), st.pop()
                if op == L.I_ADD:
                    v = wrap(l + r)
                elif op == L.I_SUB:
                    v = wrap(l - r)
                elif op == L.I_MUL:
                    v = wrap(l * r)
                elif op == L.I_DIV:
                    if r == 0:
                        v = 0
                    elif l == -(1 << 63) and r == -1:
                        v = l
                    else:  # C truncating division
                        v = (abs(l) // abs(r)) * (1 if (l < 0) == (r < 0) else -1)
                elif op == L.I_CMP:
                    v = int({1: l < r, 2: l > r, 3: l != r, 4: l == r, 5: l <= r, 6: l >= r}[a])
                elif op == L.I_AND:
                    v = int(bool(l) and bool(r))
                else:
                    v = int(bool(l) or bool(r))
                st.append(v)
        out[i] = st[-1]
    return out


NAME, NUM, STR, CMP, AND, OR = L.RPN_NAME, L.RPN_NUMBER, L.RPN_STRING, L.RPN_CMP, L.RPN_AND, L.RPN_OR
ADD, MINUS, MUL, DIV = L.RPN_ADD, L.RPN_MINUS, L.RPN_MUL, L.RPN_DIV


def random_rpn(r, ncols, depth=0):
    """A random well-formed predicate RPN over columns and literals."""
    def operand(d):
        c = r.random()
        if c < 0.45 or d > 1:
            return [(NAME, int(r.integers(0, ncols)), 0)]
        if c < 0.8:
            return [(NUM, int(r.integers(0, 3)), int(r.integers(-200, 200)))]
        return operand(d + 1) + operand(d + 1) + [(int(r.choice([ADD, MINUS, MUL, DIV])), 0, 0)]

    def pred(d):
        if d < 2 and r.random() < 0.4:
            return pred(d + 1) + pred(d + 1) + [(int(r.choice([AND, OR])), 0, 0)]
        return operand(0) + operand(0) + [(CMP, int(r.integers(1, 7)), 0)]

    return pred(depth)


def test_compiled_predicates_match_oracle():
    r = np.random.default_rng(11)
    cols = [r.integers(-300, 300, 64) for _ in range(3)]
    dec = [0, 2, 1]
    checked = 0
    for _ in range(400):
        ops = random_rpn(r, 3)
        try:
            want = O.filter_mask(ops, cols, dec)
        except L.AlnError as e:
            with pytest.raises(L.AlnError):
                L.compile_rpn(ops, dec)
            continue
        prog, _ = L.compile_rpn(ops, dec)
        got = run_prog(prog, cols).astype(bool)
        assert np.array_equal(got, want), ops
        checked += 1
    assert checked > 200


def test_compiled_expressions_match_oracle():
    r = np.random.default_rng(12)
    cols = [r.integers(-1000, 1000, 64) for _ in range(3)]
    dec = [2, 0, 2]
    for _ in range(300):
        a = [(NAME, int(r.integers(0, 3)), 0)]
        b = [(NUM, int(r.integers(0, 3)), int(r.integers(-50, 50)))]
        ops = (a + b if r.random() < 0.5 else b + a) + [(int(r.choice([ADD, MINUS, MUL, DIV])), 0, 0)]
        if r.random() < 0.5:
            ops = ops + [(NAME, int(r.integers(0, 3)), 0), (int(r.choice([ADD, MINUS, MUL])), 0, 0)]
        want, wp = O.eval_expr(ops, cols, dec)
        prog, p = L.compile_rpn(ops, dec, mode=1)
        assert p == wp
        assert np.array_equal(run_prog(prog, cols), want), ops


def test_zone_map_matches_oracle_random():
    r = np.random.default_rng(13)
    for _ in range(500):
        ops = random_rpn(r, 2)
        mins = [int(r.integers(-300, 300)) for _ in range(2)]
        maxs = [m + int(r.integers(0, 200)) for m in mins]
        dec = [int(r.integers(0, 3)) for _ in range(2)]
        try:
            want = O.zone_map_check(ops, mins, maxs, dec)
        except L.AlnError:
            with pytest.raises(L.AlnError):
                L.zone_map_check(ops, mins, maxs, dec)
            continue
        assert L.zone_map_check(ops, mins, maxs, dec) == want, ops


def test_fast_interval_shape_detection_is_a_conjunction():
    """The scan's interval fast path applies only to AND-chains of column-vs-constant terms."""
    prog, _ = L.compile_rpn([(NAME, 0, 0), (NUM, 0, 5), (CMP, 1, 0), (NAME, 1, 0), (NUM, 0, 2),
                             (CMP, 6, 0), (AND, 0, 0)], [0, 0])
    assert prog == [(L.I_COL, 0, 1), (L.I_CONST, 0, 5), (L.I_CMP, 1, 0), (L.I_COL, 1, 1),
                    (L.I_CONST, 0, 2), (L.I_CMP, 6, 0), (L.I_AND, 0, 0)]


def test_errors_are_status_codes():
    with pytest.raises(L.AlnError) as e:
        L.compile_rpn([(NAME, 0, 0), (CMP, 1, 0)], [0])
    assert e.value.code == L.ALN_E_ARG
    with pytest.raises(L.AlnError) as e:
        L.compile_rpn([(NAME, 0, 0), (STR, 0, 1), (CMP, 1, 0)], [0])  # '<' on a string
    assert e.value.code == L.ALN_E_UNSUPPORTED
    with pytest.raises(L.AlnError):
        L.compile_rpn([(NAME, 5, 0), (NUM, 0, 1), (CMP, 1, 0)], [0])  # unknown column slot


def test_like_compiles_to_a_code_set_test():
    """column LIKE pattern arrives as NAME, CODESET (the dictionary codes the pattern matches,
    as a device bit set), CMP 7 and becomes one set-membership instruction; the zone map
    evaluates such a segment ('R'), as the reference does for every char comparison."""
    SET = L.RPN_CODESET
    prog, _ = L.compile_rpn([(NAME, 1, 0), (SET, 150, 0x7F0000001000), (CMP, 7, 0), (NAME, 0, 0),
                             (NUM, 0, 15), (CMP, 4, 0), (AND, 0, 0)], [0, 0])
    assert prog == [(L.I_COL, 1, 1), (L.I_INSET, 150, 0x7F0000001000), (L.I_COL, 0, 1),
                    (L.I_CONST, 0, 15), (L.I_CMP, 4, 0), (L.I_AND, 0, 0)]
    assert L.zone_map_check([(NAME, 0, 0), (SET, 150, 0x1000), (CMP, 7, 0)], [0], [9], [0]) == "R"
    for bad in ([(NAME, 0, 0), (STR, 0, 3), (CMP, 7, 0)],      # pattern not matched to codes
                [(NAME, 0, 0), (SET, 150, 0x1000), (CMP, 4, 0)],  # a set compared with ==
                [(NUM, 0, 1), (SET, 150, 0x1000), (CMP, 7, 0)]):  # LIKE on a number
        with pytest.raises(L.AlnError) as e:
            L.compile_rpn(bad, [0])
        assert e.value.code == L.ALN_E_UNSUPPORTED
    with pytest.raises(L.AlnError) as e:
        L.compile_rpn([(NAME, 0, 0), (SET, 150, 0), (CMP, 7, 0)], [0])  # no bit set
    assert e.value.code == L.ALN_E_ARG


def test_host_murmur_matches_oracle():
    r = np.random.default_rng(14)
    for n in range(0, 40):
        b = r.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert L.murmur64a(b, 100) == O.murmur64a(b, 100)


def test_operator_layer_exports_every_declared_symbol():
    """include/alenka_ops.h: the reference's emit_* entry points, with C linkage."""
    from alenka_amd import ops

    text = open(os.path.join(ROOT, "include", "alenka_ops.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*((?:alenka|emit|clean|process)_\w+|alenkaExecute)\s*\(",
                                  text, re.M)))
    lib = L.load()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(ops._SIG), set(names) ^ set(ops._SIG)
    for n in ("emit_filter", "emit_select", "emit_join", "emit_order", "emit_store", "emit_name",
              "emit_number", "emit_cmp", "emit_sum", "emit_count", "emit_sel_name"):
        assert n in names


FLT = L.RPN_FLOAT


def random_f64_rpn(r, ncols, fcols):
    """A random predicate mixing f64 columns (slots in fcols), int columns, FLOATs and
    decimal NUMBERs."""
    def operand(d):
        c = r.random()
        if c < 0.4 or d > 1:
            return [(NAME, int(r.integers(0, ncols)), 0)]
        if c < 0.6:
            return [(FLT, 0, f2b(float(r.choice([1.5, -2.25, 0.0, 3e-9, 100.125, 7.0]))))]
        if c < 0.8:
            return [(NUM, int(r.integers(0, 3)), int(r.integers(-200, 200)))]
        return operand(d + 1) + operand(d + 1) + [(int(r.choice([ADD, MINUS, MUL, DIV])), 0, 0)]

    def pred(d):
        if d < 2 and r.random() < 0.4:
            return pred(d + 1) + pred(d + 1) + [(int(r.choice([AND, OR])), 0, 0)]
        a, b = operand(0), operand(0)
        if not any(t == FLT or (t == NAME and x in fcols) for t, x, _ in a + b):
            b = [(FLT, 0, f2b(0.5))]
        return a + b + [(CMP, int(r.integers(1, 7)), 0)]

    return pred(0)


def test_compiled_f64_predicates_match_oracle():
    """Programs with double operands, compiled by the C ABI and run by the interpreter above,
    give the oracle's masks (EPSILON comparisons, plain casts of int columns)."""
    r = np.random.default_rng(21)
    f = r.choice([1.5, 1.5 + 5e-9, 1.5 - 2e-8, -2.25, 0.0, 7.0, 100.125], 64) + r.normal(0, 1, 64) * (r.random(64) < 0.5)
    cols = [np.array([f2b(x) for x in f], np.int64), r.integers(-300, 300, 64),
            np.array([f2b(x) for x in r.normal(0, 50, 64)], np.int64), r.integers(-300, 300, 64)]
    dec = [L.DEC_F64, 0, L.DEC_F64, 2]
    checked = 0
    for _ in range(400):
        ops = random_f64_rpn(r, 4, (0, 2))
        try:
            want = O.filter_mask(ops, cols, dec)
        except L.AlnError:
            with pytest.raises(L.AlnError):
                L.compile_rpn(ops, dec)
            continue
        prog, _ = L.compile_rpn(ops, dec)
        got = run_prog(prog, cols).astype(bool)
        assert np.array_equal(got, want), ops
        checked += 1
    assert checked > 200


def test_compiled_f64_expressions_match_oracle():
    r = np.random.default_rng(22)
    cols = [np.array([f2b(x) for x in r.normal(0, 100, 64)], np.int64), r.integers(-1000, 1000, 64)]
    dec = [L.DEC_F64, 2]
    for _ in range(300):
        a = [(NAME, int(r.integers(0, 2)), 0)]
        b = [(FLT, 0, f2b(float(r.normal(0, 10))))] if r.random() < 0.6 else [(NUM, int(r.integers(0, 3)), int(r.integers(-50, 50)))]
        ops = (a + b if r.random() < 0.5 else b + a) + [(int(r.choice([ADD, MINUS, MUL, DIV])), 0, 0)]
        if r.random() < 0.5:
            ops = ops + [(NAME, 0, 0), (int(r.choice([ADD, MINUS, MUL])), 0, 0)]
        want, wp = O.eval_expr(ops, cols, dec)
        prog, p = L.compile_rpn(ops, dec, mode=1)
        assert p == wp
        got = run_prog(prog, cols)
        if p == L.DEC_F64:  # NaN bit patterns may differ; compare as doubles
            assert np.array_equal(got.view(np.float64), want.view(np.float64), equal_nan=True), ops
        else:
            assert np.array_equal(got, want), ops


def test_f64_predicates_are_not_zone_mapped():
    dec = [L.DEC_F64, 0]
    assert L.zone_map_check([(NAME, 0, 0), (FLT, 0, f2b(2.0)), (CMP, 1, 0)], [0, 0], [1, 1], dec) == "R"
    assert L.zone_map_check([(NAME, 1, 0), (FLT, 0, f2b(2.0)), (CMP, 1, 0)], [0, 0], [1, 1], dec) == "R"
    assert O.zone_map_check([(NAME, 1, 0), (NUM, 0, 2), (CMP, 1, 0)], [0, 0], [1, 1], dec) == "A"
    assert O.zone_map_check([(NAME, 0, 0), (FLT, 0, f2b(2.0)), (CMP, 1, 0)], [0, 0], [1, 1], dec) == "R"

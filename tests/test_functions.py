"""Boundary and source functions: the expression reader and the postfix
programs handed to the library (dgx_function), checked on the CPU.

The programs are pinned by running them through an independent stack
interpreter written here from the instruction set of include/dgx.h and
comparing with the expression trees evaluated directly; the trees themselves
are pinned against hand-written formulas in test_oracle_golden.py.
"""
import math

import numpy as np
import pytest

from dgm_amd import formats
from fixture_runs import GOLDEN

_UN = {10: lambda a: -a, 11: math.sin, 12: math.cos, 13: math.tan, 14: math.asin, 15: math.acos,
       16: math.atan, 17: math.sinh, 18: math.cosh, 19: math.tanh, 20: math.exp, 21: math.log,
       22: math.log10, 23: math.sqrt, 24: abs}
_BIN = {5: lambda a, b: a + b, 6: lambda a, b: a - b, 7: lambda a, b: a * b, 8: lambda a, b: a / b,
        9: lambda a, b: a ** b, 25: math.atan2}


def run_program(code, b, e, consts, x, y, t):
    st = []
    for ins in code[b:e]:
        op = int(ins) & 255
        if op == 0:
            st.append(consts[int(ins) >> 8])
        elif op <= 4:
            st.append((x, y, 0.0, t)[op - 1])
        elif op in _BIN:
            bb = st.pop()
            st[-1] = _BIN[op](st[-1], bb)
        else:
            st[-1] = _UN[op](st[-1])
        assert len(st) <= 16
    assert len(st) == 1
    return st[0]


@pytest.mark.parametrize("path,name", [("shallow_mms.bct", "dansMMSFunction"), ("shallow_mms.src", "funcMMS")])
def test_programs_match_the_trees(path, name):
    comps = formats.read_functions(f"{GOLDEN}/{path}")[name]
    start, code, consts = formats.compile_function(comps)
    assert len(start) == len(comps) + 1 and start[0] == 0
    rng = np.random.default_rng(3)
    f = formats.field_function(comps)
    for _ in range(20):
        x, y, t = rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0, 1)
        want = f(t, np.array([x]), np.array([y]))[:, 0]
        got = [run_program(code, start[c], start[c + 1], consts, x, y, t) for c in range(len(comps))]
        assert np.allclose(got, want, rtol=1e-14, atol=1e-14)


def test_every_operator():
    text = ("function f { a := sin(x) + cos(y) * tan(0.3) - asin(0.2) / acos(0.1) ^ atan(t);"
            " b := -sinh(x) + cosh(y) - tanh(t) + exp(0.5) * log(2.0) + log10(3.0) - sqrt(4.0) + abs(-x);"
            " c := atan2(y, x) + pow(2.0, t) + fabs(y) - PI; }")
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.src")
        with open(p, "w") as fh:
            fh.write(text)
        comps = formats.read_functions(p)["f"]
    start, code, consts = formats.compile_function(comps)
    ops = {int(i) & 255 for i in code}
    assert ops == set(range(26)) - {3}          # every instruction but z
    x, y, t = 0.7, -0.4, 0.25
    want = [math.sin(x) + math.cos(y) * math.tan(0.3) - math.asin(0.2) / math.acos(0.1) ** math.atan(t),
            -math.sinh(x) + math.cosh(y) - math.tanh(t) + math.exp(0.5) * math.log(2.0) + math.log10(3.0)
            - math.sqrt(4.0) + abs(-x),
            math.atan2(y, x) + 2.0 ** t + abs(y) - math.pi]
    got = [run_program(code, start[c], start[c + 1], consts, x, y, t) for c in range(3)]
    assert np.allclose(got, want, rtol=1e-15, atol=1e-15)


def test_reader_rejects_what_it_does_not_know():
    for bad in ("a := foo(x);", "a := q + 1;", "a := (x + 1;", "a := x ! 2;"):
        toks = None
        with pytest.raises(ValueError):
            toks = formats._tokens(bad.split(":=")[1].rstrip(";"))
            p = formats._ExprParser(toks)
            p.expr()
            if p.i != len(toks):
                raise ValueError("trailing tokens")

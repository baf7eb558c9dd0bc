"""The algorithmic flop count of one interior element (e3 at the 4-point
rule, residual and 20x20 tangent), frozen.

oracle/count_flops.cpp compiles the oracle's element routines, which follow
e3.f and the routines it calls operation for operation, with a counting
number type.  Operations with a structurally zero operand (the exact-zero
A-matrix entries, zeroed accumulators) are counted apart and are not
arithmetic work.  These figures price the assembly kernels' FLOP rate.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_build", "count_flops")

# per element: (residual only, residual + tangent)
WORK = (4600, 38268)        # add + mul + div + sqrt with nonzero operands
EXECUTED = (5556, 47532)    # every operation the oracle performs


def counts():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    rows = {}
    for ln in out.splitlines():
        tok = ln.split()
        lhs = int(tok[0].split("=")[1])
        rows[lhs] = {tok[i]: int(tok[i + 1]) for i in range(1, len(tok) - 1, 2)}
    return rows


def test_e3_flop_count_is_frozen():
    rows = counts()
    for lhs in (0, 1):
        r = rows[lhs]
        assert r["add"] + r["mul"] + r["div"] + r["sqrt"] == r["work"]
        assert r["work"] + r["zero"] == r["executed"]
        assert r["work"] == WORK[lhs], (lhs, r)
        assert r["executed"] == EXECUTED[lhs], (lhs, r)
    # one square root per point (tau), four points
    assert rows[1]["sqrt"] == 4 and rows[0]["sqrt"] == 4


def test_flop_count_sees_the_tangent_loops():
    """The tangent adds the 20x20 products of e3conv, e3ls, e3massl and
    e3wmlt: at least 400 multiply-adds per point (a missing loop nest would
    fall well below)."""
    rows = counts()
    extra = rows[1]["work"] - rows[0]["work"]
    assert extra > 4 * 2 * 400 * 3

"""The binomial checkpointing schedule of the adjoint (dgx_revolve_count runs
it on the host, no GPU): every step is reversed once, in reverse order, from
its own forward state; it never holds more states than slots; and it advances
no more than binomial checkpointing allows.

With s free slots and r advances per step, binomial checkpointing reverses
C(s + r, s) steps (Griewank and Walther's revolve).  Here the reversal of step
k also advances state_k to state_{k+1} once, so the bound is (r + 1) n.
"""
from __future__ import annotations

from math import comb

import pytest

from dgm_amd import acoustic


def _rep(n, s):
    r = 1
    while comb(s + r, s) < n:
        r += 1
    return r


@pytest.mark.parametrize("n,slots", [(1, 1), (2, 1), (7, 1), (10, 2), (100, 3), (100, 11), (2000, 46),
                                     (2000, 10), (999, 4), (64, 64), (500, 500)])
def test_schedule(n, slots):
    adv, used, ok = acoustic.revolve_count(n, slots)
    assert ok
    assert used <= slots
    if slots == 1:
        assert adv == n * (n + 1) // 2                 # recompute from state 0 every time
    else:
        r = _rep(n, slots - 1)
        assert adv <= (r + 1) * n, (adv, r)
    assert adv >= n                                    # at least one advance per step


def test_more_slots_fewer_advances():
    a = [acoustic.revolve_count(2000, s)[0] for s in (3, 6, 12, 24, 46)]
    assert all(x >= y for x, y in zip(a, a[1:])), a
    # sqrt(N) slots: about three advances per step at N = 2000
    assert a[-1] <= 4 * 2000

"""The binomial (revolve-style) checkpointing schedule of the adjoint's long
solves, modelled on the CPU: revolve_plan and adj_reverse_range in
lbm_amd/csrc/capi.cpp, restated step for step.  The GPU tests check the
C++ code's result bitwise against other memory budgets; this checks the
schedule itself: every block of a solve is reversed exactly once, last first,
from a snapshot holding that block's start, no snapshot is overwritten while
a later block still needs it, and the recompute cost stays within d + 1
passes for C(F + d, d) blocks."""
from math import comb

import pytest


def capacity(f, d):
    return comb(f + d, d)


def plan(lo, hi, F):
    d = 1
    while capacity(F, d) < hi - lo:
        d += 1
    b = [lo]
    i = 0
    while b[-1] < hi:
        ln = capacity(F - i, d - 1) if i <= F else hi - b[-1]
        b.append(min(hi, b[-1] + ln))
        i += 1
    return b


def reverse(nblocks, F_total, top_from_forward=False):
    """Returns (order of reversed blocks, recomputed blocks)."""
    pool = {}
    order, cost = [], [0]

    def forward(start_block, marks):
        for blk, slot in marks:
            assert blk > start_block
            pool[slot] = blk
        if marks:
            cost[0] += marks[-1][0] - start_block

    def block(j, start):
        assert (start is None) if j == 0 else pool[start] == j, (j, start)
        order.append(j)
        cost[0] += 1

    def segments(b, start, fb):
        for i in range(len(b) - 2, -1, -1):
            rng(b[i], b[i + 1], start if i == 0 else fb + i - 1, fb + i)

    def rng(lo, hi, start, fb):
        if hi - lo == 1:
            block(lo, start)
            return
        F = F_total - fb if fb < F_total else 0
        if F == 0:
            for j in range(hi - 1, lo - 1, -1):
                sj = start
                if j > lo:
                    forward(lo, [(j, "scratch")])
                    sj = "scratch"
                block(j, sj)
            return
        b = plan(lo, hi, F)
        assert len(b) - 2 <= F
        forward(lo, [(b[i], fb + i - 1) for i in range(1, len(b) - 1)])
        segments(b, start, fb)

    if top_from_forward and nblocks - 1 > F_total > 0:
        b = plan(0, nblocks, F_total)
        for i in range(1, len(b) - 1):
            pool[i - 1] = b[i]
        segments(b, None, 0)
    else:
        rng(0, nblocks, None, 0)
    return order, cost[0]


@pytest.mark.parametrize("nblocks,F", [(1, 0), (2, 0), (5, 0), (7, 1), (3, 1), (100, 2), (552, 29),
                                       (552, 34), (789, 69), (2000, 3), (40, 39)])
@pytest.mark.parametrize("top", [False, True])
def test_every_block_reversed_once_in_order_from_its_start(nblocks, F, top):
    order, cost = reverse(nblocks, F, top)
    assert order == list(range(nblocks - 1, -1, -1))
    if F > 0 and nblocks > 1:
        d = 1
        while capacity(F, d) < nblocks:
            d += 1
        # d forward passes (the first saved when the forward pass planned
        # it) plus the blocks themselves
        assert cost <= (d + (0 if top and nblocks - 1 > F else 1)) * nblocks, (cost, d)


def test_planning_saves_the_top_pass():
    _, plain = reverse(552, 34)
    _, planned = reverse(552, 34, top_from_forward=True)
    assert planned < plain - 0.8 * 552

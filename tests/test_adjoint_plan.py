"""Checkpoint schedules of the adjoint (culbm_adjoint_plan; host only).

The library checks every plan symbolically (each op reads the state it
expects, reverse steps in order N-1..0) and fails the call otherwise; these
tests pin the forward-step counts against closed forms and against a
brute-force search over binary splits (Griewank's binomial bound)."""
from __future__ import annotations

from functools import lru_cache

import pytest

from culbm_amd import CulbmError
from culbm_amd.lattice import adjoint_plan


@lru_cache(None)
def best(l: int, extra: int) -> int:
    """Fewest forward steps to reverse l steps from a held state with
    `extra` more slots (exhaustive over the first split)."""
    if l == 1:
        return 0
    if extra == 0:
        return l * (l - 1) // 2
    return min(m + best(l - m, extra - 1) + best(m, extra) for m in range(1, l))


@pytest.mark.parametrize("n,k", [(300, 20), (10, 3), (7, 7), (7, 100), (1, 1), (1000, 50), (24, 5)])
def test_every_k_recomputes_each_segment_but_the_last_once(n, k):
    fwd, slots = adjoint_plan(n, k)
    segs = -(-n // k)
    # segment j spans min(k, n - jk) steps and is recomputed into its slots; the
    # last one's states come from the first sweep
    assert fwd == n + sum(min(k, n - j * k) - 1 for j in range(segs - 1))
    assert slots == segs + min(k, n) - 1


@pytest.mark.parametrize("n", [1, 2, 5, 24, 300])
def test_binomial_with_a_slot_per_step_recomputes_nothing(n):
    assert adjoint_plan(n, 0, n) == (n, n)
    assert adjoint_plan(n, 0, 10 * n) == (n, n)


@pytest.mark.parametrize("n", [1, 2, 7, 50])
def test_binomial_with_one_slot_is_quadratic(n):
    assert adjoint_plan(n, 0, 1) == (n + n * (n - 1) // 2, 1)


@pytest.mark.parametrize("n,s", [(10, 2), (30, 3), (50, 5), (120, 4), (200, 8), (300, 16)])
def test_binomial_within_the_optimal_bound(n, s):
    fwd, used = adjoint_plan(n, 0, s)
    assert used == s
    # the first sweep runs all n steps and stores the splits on its path, so the
    # total stays within n + (the optimum of the reverse sweep alone)
    assert n <= fwd <= n + best(n, s - 1)


def test_binomial_for_the_large_run():
    # 1000 steps with 33 slots: revolve's repetition number is 3 (C(35, 3) >= 1000 > C(34, 2))
    fwd, used = adjoint_plan(1000, 0, 33)
    assert used == 33
    assert fwd < 3 * 1000


def test_plan_rejects_bad_arguments():
    with pytest.raises(CulbmError):
        adjoint_plan(0, 5)
    with pytest.raises(CulbmError):
        adjoint_plan(10, 0, 0)
    with pytest.raises(CulbmError):
        adjoint_plan(10, -1)

"""The binomial checkpointing recurrence the revolve schedule minimises, pinned by Griewank's
closed form (Griewank & Walther, ACM TOMS 26 (2000) 19-45, Prop. 1): reversing l steps with c
checkpoints (the initial state one of them) takes at least

    t(l, c) = r l - beta(c + 1, r - 1),   beta(c, r) = C(c + r, c),   beta(c, r - 1) < l <= beta(c, r)

forward advances.  The recurrence T(n, s) = min_k [k + T(n - k, s - 1) + T(k, s)], T(1, s) = 0,
T(n, 0) = n (n - 1) / 2 counts the same advances with s snapshot buffers besides the start state,
so T(n, s) = t(n, s + 1).  tests/test_gpu_revolve.py checks the library's recomputation count
against this recurrence."""
import functools
from math import comb


@functools.lru_cache(maxsize=None)
def _T(n, s):
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(k + _T(n - k, s - 1) + _T(k, s) for k in range(1, n))


def _griewank(l, c):
    r = 0
    while comb(c + r, c) < l:
        r += 1
    return r * l - comb(c + 1 + r - 1, c + 1)


def test_recurrence_equals_closed_form():
    for s in range(0, 6):
        for n in range(1, 160):
            assert _T(n, s) == _griewank(n, s + 1), (n, s)


def test_closed_form_special_cases():
    # one checkpoint (the start): step i is reached by i advances
    assert all(_griewank(l, 1) == l * (l - 1) // 2 for l in range(1, 50))
    # as many snapshots as steps: one sweep to place them, nothing recomputed beyond it
    assert all(_T(n, n) == n - 1 for n in range(1, 40))

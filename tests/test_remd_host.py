"""The REMD exchange pass (sdmd_remd_rearrange, Server.c:145-192) on the
host, pinned by what its rule fixes rather than by a second copy of it:

  * a pass only permutes the {num, T} records;
  * downhill swaps (dE <= 0) are taken without a random draw;
  * a swap so far uphill that exp(-dE) underflows is never taken and costs
    exactly one draw per attempt;
  * the acceptance frequency of an uphill swap is exp(-dE);
  * the server's pairing: every replica but #1 tries the one a step below,
    and #1 tries #2, so the (1,2) pair is tried twice when it is not swapped.
"""
import ctypes as C

import numpy as np
import pytest

from sdmd_amd import _lib
from sdmd_amd._lib import RandState, RemdTemp
from sdmd_amd.remd import BOLTZMANN, ladder, rearrange


def temps(values):
    t = (RemdTemp * len(values))()
    for k, v in enumerate(values):
        t[k].num = k + 1
        t[k].T = v
    return t


def rng(seed=7):
    r = RandState()
    _lib.lib().sdmd_rand_seed(C.byref(r), seed)
    return r


def test_ladder_is_geometric_in_reduced_units():
    t = np.array(ladder(64))
    assert t[0] == pytest.approx(280 * BOLTZMANN) and t[-1] == pytest.approx(450 * BOLTZMANN)
    r = t[1:] / t[:-1]
    assert np.allclose(r, r[0])
    assert ladder(1) == [280 * BOLTZMANN]


def test_pass_permutes_records():
    T = temps(ladder(16))
    before = sorted((t.num, t.T) for t in T)
    g = rng(3)
    e = np.random.default_rng(0).normal(0, 50, 16)
    for _ in range(20):
        rearrange(e, T, g)
        assert sorted((t.num, t.T) for t in T) == before


def test_level_swaps_take_no_draw():
    # equal energies: every tried pair has dE = 0, which is accepted without
    # a random draw
    n = 6
    T = temps(ladder(n))
    g = rng(11)
    counts = np.zeros(2 * n, np.int64)
    rearrange(np.full(n, -42.0), T, g, counts)
    assert g.calls == 0
    assert counts[0::2].sum() == counts[1::2].sum() == n


def test_far_uphill_swaps_never_taken_one_draw_each():
    n = 5
    T = temps(ladder(n))
    e = np.arange(n, dtype=float) * 1e6          # colder replicas far lower in energy
    g = rng(5)
    counts = np.zeros(2 * n, np.int64)
    rearrange(e, T, g, counts)
    assert [t.num for t in T] == list(range(1, n + 1))
    assert counts[1::2].sum() == 0
    # replicas #2..#n try the one below, #1 tries #2: n attempts, n draws
    assert g.calls == n
    assert counts[0] == 2 and counts[2:2 * (n - 1):2].tolist() == [1] * (n - 2)


def test_uphill_acceptance_frequency_is_boltzmann():
    # two replicas: #1 tries #2 uphill (dE = 0.8) and, when that is taken,
    # the swapped pair's reverse try is downhill and swaps back; when it is
    # refused, #2 tries #1 with the same dE.  So the pass ends swapped with
    # probability p (1 - p), p = exp(-0.8), and draws 2 - p times on average.
    t1, t2 = 0.6, 0.7
    dE = 0.8
    e = np.array([0.0, dE / (1 / t1 - 1 / t2)])
    g = rng(123)
    swapped = 0
    trials = 40000
    for _ in range(trials):
        T = temps([t1, t2])
        rearrange(e, T, g)
        swapped += int(T[0].num == 2)
    p = np.exp(-dE)
    assert abs(swapped / trials - p * (1 - p)) < 0.01
    assert abs(g.calls / trials - (2 - p)) < 0.01


def test_finished_replicas_are_skipped():
    T = temps([0.6, 0.0, 0.7])
    T[1].T = 0.0
    g = rng(9)
    rearrange(np.array([0.0, -1e9, 0.0]), T, g)
    assert T[1].num == 2 and T[1].T == 0.0


def test_rejects_bad_arguments():
    T = temps([0.6])
    assert _lib.lib().sdmd_remd_rearrange(None, T, 1, None, None) == -1

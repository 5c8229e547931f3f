"""Smooth cylinder wall in the fast engine, the device energy, the
per-system temperature and the REMD loop, through the C ABI.

(The plain engine's wall runs are in test_gpu_parity.py: the cylinder golden
cases replay the compiled reference's event sequence bit for bit.)"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from conftest import golden_system
from sdmd_amd import _lib

pytestmark = pytest.mark.gpu


def host_energy(arr):
    d = arr.desc()
    ke, pe = C.c_double(), C.c_double()
    assert _lib.lib().sdmd_host_energy(C.byref(d), C.byref(ke), C.byref(pe)) == 0
    return ke.value, pe.value


def _nve_state_near_wall(name, pops):
    """The oracle's state, thermostat off, at the first pop after `pops`
    where some atom's pending event is a wall event, as a System the engine
    can open."""
    from sdmd_amd.engine import System

    s, _ = golden_system(name)
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    tr, _ = o.run(pops, trace=True)
    # pops after which a re-predicted atom holds a wall event
    cand = np.flatnonzero((tr["r"][:, 0]["type"] == 5) | (tr["r"][:, 1]["type"] == 5))
    assert len(cand) > 0
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    o.run(int(cand[len(cand) // 2]) + 1, trace=False)
    st = o.state()
    st.scalars["thermostat"] = 0
    return System(None, st)


def test_fast_bulk_prediction_with_wall_matches_oracle(require_gpu):
    from conftest import fast_bulk_vs_oracle

    sysm = _nve_state_near_wall("ala10cyl", 60000)
    fast_bulk_vs_oracle(sysm)
    # a run this close to the wall: some atoms' next event is a wall event
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    assert np.count_nonzero(o.state().arr["ev_type"][1:] == 5) > 0


@pytest.mark.parametrize("name", ["ala10cyl", "cyl8x20"])
def test_fast_wall_run_screens_agree_with_fp64(require_gpu, name):
    from sdmd_amd.engine import GpuEngine

    sysm, _ = golden_system(name)
    eng = GpuEngine([sysm], block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run(60000 if name == "ala10cyl" else 150000)
    st, fs = eng.stats(0), eng.fast_stats(0)
    assert st["error"] == 0
    assert fs["agree"] == fs["validated"] and fs["validated"] > 10000
    assert fs["causal"] == 0
    if name == "ala10cyl":
        assert st["per_type"][5] > 0          # wall events executed
    # every atom stays inside the cylinder, clear of the wall's core
    a = eng.state(0).arr
    sc = sysm.arrays.scalars
    R = sc["box"][1] / 2
    r = np.hypot(a["y"][1:] - sc["wall_pos"][1], a["z"][1:] - sc["wall_pos"][2])
    assert r.max() < R - 1.0
    eng.close()


@pytest.mark.parametrize("mode", ["plain", "fast"])
def test_device_energy_matches_host(require_gpu, mode):
    from sdmd_amd.engine import GpuEngine

    s1, _ = golden_system("cyl8x20")
    s2, _ = golden_system("mix4x20")
    kw = dict(max_pops=20000) if mode == "plain" else {}
    eng = GpuEngine([s1, s2], block=256, mode=mode, **kw)
    eng.prepare()
    for _ in range(2):
        ke, pe = eng.energy()
        for k in range(2):
            hk, hp = host_energy(eng.state(k))
            assert ke[k] == pytest.approx(hk, rel=1e-12, abs=1e-9)
            # per-atom partial sums in another order than the host loop
            assert pe[k] == pytest.approx(hp, rel=1e-10, abs=1e-8)
        eng.run(8000)
    eng.close()


def test_run_events_counts_executed_events(require_gpu):
    from sdmd_amd.engine import GpuEngine

    s1, _ = golden_system("mix4x20")
    s2, _ = golden_system("ala10ext")
    for mode in ("plain", "fast"):
        kw = dict(max_pops=40000) if mode == "plain" else {}
        eng = GpuEngine([s1, s2], block=256, mode=mode, **kw)
        eng.prepare()
        f0 = [eng.stats(k)["frame"] for k in range(2)]
        eng.run_events(5000)
        for k in range(2):
            d = eng.stats(k)["frame"] - f0[k]
            if mode == "plain":
                assert d == 5000, (mode, k, d)
            else:
                assert 5000 <= d < 5000 + 4096, (mode, k, d)     # plus the closing drain
        eng.close()


def test_set_temperature_moves_the_thermostat(require_gpu):
    """Andersen draws at variance T/m: after raising T the instantaneous
    temperature 2 KE / 3N follows it."""
    from sdmd_amd.engine import GpuEngine

    s, _ = golden_system("mix4x20")
    T0 = s.arrays.scalars["temperature"]
    eng = GpuEngine([s, s], block=256, mode="fast")
    eng.prepare()
    eng.set_temperature(1, 2 * T0)
    eng.run_events(150000)
    samples = []
    for _ in range(10):
        eng.run_events(20000)
        ke, _ = eng.energy()
        samples.append(2 * ke / (3 * s.natom))
    t = np.mean(samples, axis=0)
    assert abs(t[0] / T0 - 1) < 0.15, t
    assert abs(t[1] / (2 * T0) - 1) < 0.15, t
    with pytest.raises(_lib.SdmdError):
        eng.set_temperature(0, -1.0)
    eng.close()


def test_remd_loop_exchanges_temperatures(require_gpu):
    from sdmd_amd.remd import Remd, ladder

    s, _ = golden_system("cyl8x20")
    n = 4
    temps = ladder(n, 280.0, 330.0)
    r = Remd([s] * n, temps, seed=3, interval=4000, block=256)
    r.prepare()
    for _ in range(6):
        pe = r.cycle()
        assert np.all(np.isfinite(pe))
    T, nums = r.temperatures()
    assert sorted(nums.tolist()) == list(range(1, n + 1))
    assert np.allclose(np.sort(T), temps)
    for k in range(n):                     # each system runs at the temperature it holds
        assert r.eng.stats(k)["error"] == 0
    assert r.counts[0::2].sum() == 6 * n
    assert r.counts[1::2].sum() > 0        # neighbouring temperatures this close do swap
    r.close()


def test_fast_wall_schedule_invariance(require_gpu):
    """In the cylinder (wall events, periodic along x only) one event per
    round and full concurrency reach bit-identical states."""
    from conftest import assert_same_schedule_result, fast_schedule_states

    s, _ = golden_system("cyl8x20")
    assert_same_schedule_result(fast_schedule_states(s, 30000), 15000)


def test_remd_device_exchange_matches_host_pass(require_gpu):
    """The exchange pass run on the device between event windows
    (sdmd_remd_run: no host round trip per cycle) makes, on the energies it
    used, the same swaps, the same attempt/acceptance counts and the same
    rand() draws as the server's pass on the host (sdmd_remd_rearrange,
    Server.c:145-192) replayed on those energies."""
    from sdmd_amd._lib import RandState, RemdTemp
    from sdmd_amd.remd import Remd, ladder, rearrange

    s, _ = golden_system("cyl8x20")
    n, cycles = 8, 6
    temps = ladder(n, 280.0, 320.0)
    r = Remd([s] * n, temps, seed=3, interval=3000, block=256)
    r.prepare()
    T0 = [(t.num, t.T) for t in r.T]
    rng0 = RandState.from_buffer_copy(bytes(r.rng))
    ep, tab = r.run(cycles, history=True)
    assert ep.shape == (cycles, n) and np.all(np.isfinite(ep))
    T = (RemdTemp * n)()
    for k, (num, t) in enumerate(T0):
        T[k].num, T[k].T = num, t
    counts = np.zeros(2 * n, np.int64)
    for c in range(cycles):
        rearrange(ep[c], T, rng0, counts)
        assert [(t.num, t.T) for t in T] == tab[c], c
    assert np.array_equal(counts, r.counts)
    assert bytes(rng0) == bytes(r.rng)                  # same draws, same stream position
    assert counts[0::2].sum() == cycles * n and counts[1::2].sum() > 0
    for k in range(n):
        assert r.eng.stats(k)["error"] == 0 and r.eng.fast_stats(k)["causal"] == 0
    r.close()

"""The benchmark-scale configurations under test, through the C ABI.

  * the batching form at its full size: the seeded batch of 48 peptide
    systems (10-80 residues x 1-32 chains, about 455k atoms) that the
    benchmark times, run in one persistent launch with every screened
    prediction re-solved by a full FP64 scan: 100 % agreement, no causal
    read and no error in any system;
  * the fast engine's predictions after it has evolved a state on its own:
    the smallest, median and largest systems of the batch, thermostat off,
    run for 2*10^4 events; on the snapshot each closing drain leaves, the
    fast engine's bulk prediction equals the oracle's (partner, type and
    subtype for every atom; times to 1e-12 relative where the oracle solved
    the pair itself, 1e-9 where it inherited the partner's time);
  * REMD at its stated size: 64 replicas of the 8 x 20-residue cylinder
    system, 280-450 K, five exchange cycles of 10^4 events per replica, the
    exchange passes on the device.
"""
import numpy as np
import pytest

from conftest import fast_bulk_vs_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch48(tmp_path_factory):
    import bench

    d = str(tmp_path_factory.mktemp("batch48"))
    dirs, _ = bench.build_batch(d)
    systems, _ = bench.load_systems(dirs)
    return systems


def test_config4_batch_screens_agree_with_fp64(require_gpu, batch48):
    from sdmd_amd.engine import GpuEngine

    systems = batch48
    assert len(systems) == 48 and sum(s.natom for s in systems) > 400000
    eng = GpuEngine(systems, block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run_for(1 << 40, 300.0)
    for k in range(len(systems)):
        st, fs = eng.stats(k), eng.fast_stats(k)
        assert st["error"] == 0, (k, st["error"])
        assert fs["causal"] == 0, k
        assert st["frame"] > 500, (k, st["frame"])
        assert fs["validated"] > 0 and fs["agree"] == fs["validated"], (k, fs)
    eng.close()


def test_fast_predictions_on_evolved_states_match_oracle(require_gpu, batch48):
    from sdmd_amd.engine import GpuEngine, System

    order = sorted(range(len(batch48)), key=lambda k: batch48[k].natom)
    picks = [order[0], order[len(order) // 2], order[-1]]
    nve = []
    for k in picks:
        arr = batch48[k].arrays.copy()
        arr.scalars["thermostat"] = 0
        nve.append(System(None, arr))
    eng = GpuEngine(nve, block=512, mode="fast")
    eng.prepare()
    eng.run_events(20000)
    snaps = []
    for q in range(len(nve)):
        assert eng.stats(q)["error"] == 0 and eng.fast_stats(q)["causal"] == 0
        assert eng.stats(q)["frame"] >= 20000
        snaps.append(System(None, eng.state(q)))
    eng.close()
    for q, s in enumerate(snaps):
        e_solved, e_inh, n_inh = fast_bulk_vs_oracle(s)
        print("system %d (%d atoms): max rel time error %.3g solved, %.3g over %d inherited"
              % (picks[q], s.natom, e_solved, e_inh, n_inh))


def test_config5_remd_64_replicas(require_gpu, tmp_path):
    from sdmd_amd import builder
    from sdmd_amd.remd import Remd, ladder, replica_systems

    n = 64
    temps = ladder(n)
    sysm = builder.cylinder_peptides(8, 20, seed=5, radius=40.0, box_x=160.0)
    systems = replica_systems(sysm, temps, str(tmp_path))
    r = Remd(systems, temps, seed=5, interval=10000, block=512)
    r.prepare()
    f0 = [r.eng.stats(k)["frame"] for k in range(n)]
    ep, tab = r.run(5, history=True)                  # exchange passes on the device
    assert ep.shape == (5, n) and np.all(np.isfinite(ep))
    for c in range(5):
        nums = sorted(num for num, _ in tab[c])
        assert nums == list(range(1, n + 1))          # a permutation of the ladder
        assert sorted(t for _, t in tab[c]) == sorted(temps)
    for k in range(n):
        st, fs = r.eng.stats(k), r.eng.fast_stats(k)
        assert st["error"] == 0 and fs["causal"] == 0, k
        assert st["frame"] - f0[k] >= 5 * 10000, k
    assert r.counts[0::2].sum() == 5 * n            # every replica attempts once per pass (Server.c:149-158)
    assert r.counts[1::2].sum() > 0                                 # some neighbours swapped
    r.close()

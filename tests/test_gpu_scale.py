"""GPU path vs the CPU oracle beyond the golden cases: multi-chain systems
at 5 mM with cut-off 7.25 A (many cells, periodic images, all residue
types) -- 12 x 30 residues, and 64 x 40 residues at full size (26 830
atoms, 10^5 pops) -- the multi-system batch, and the fine-grained C-ABI
entry points driven from the host."""
import numpy as np
import pytest

import oracle_lib
from conftest import golden_system

pytestmark = pytest.mark.gpu

FIELDS = ["target", "partner", "type", "sub", "exec", "dt", "pot"]


@pytest.fixture(scope="module")
def midsize(tmp_path_factory):
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    d = tmp_path_factory.mktemp("mid")
    s = builder.random_peptides(12, 30, seed=11, mM=5.0)
    builder.write_input_dir(s, str(d), cutoff=7.25)
    return System.load(str(d))


def test_midsize_trace_matches_oracle(require_gpu, midsize):
    from sdmd_amd.engine import GpuEngine

    pops = 6000
    eng = GpuEngine([midsize], block=256, trace=True, trace_capacity=pops, max_pops=pops)
    eng.prepare()
    eng.run(pops)
    tr = eng.trace(0)
    o = oracle_lib.Oracle(midsize.arrays)
    o.bulk()
    ref, _ = o.run(pops)
    for f in FIELDS:
        bad = np.flatnonzero(tr[f] != ref[f])
        assert len(bad) == 0, (f, int(bad[0]))
    g, c = eng.state(0), o.state()
    for f in ("x", "y", "z", "vx", "vy", "vz", "ev_time", "celllist", "cellx", "hb_bond"):
        assert np.array_equal(g.arr[f], c.arr[f]), f
    st = eng.stats(0)
    oc = o.counters()
    assert st["pair_evals"] > 0
    assert np.array_equal(st["per_type"], oc["per_type"])


@pytest.mark.parametrize("block", [64, 512, 1024])
def test_block_sizes_agree(require_gpu, block):
    from sdmd_amd.engine import GpuEngine

    sysm, g = golden_system("ala10helix")
    eng = GpuEngine([sysm], block=block, trace=True, trace_capacity=20000, max_pops=20000)
    eng.prepare()
    eng.run(20000)
    tr = eng.trace(0)
    n = len(g["trace"]["target"])
    for f in FIELDS:
        assert np.array_equal(tr[f][:n], g["trace"][f].astype(tr[f].dtype)), f


def test_batch_of_systems_is_independent(require_gpu):
    """Several systems in one launch (one workgroup each) give the same
    trajectories as each system alone."""
    from sdmd_amd.engine import GpuEngine

    names = ["ala10ext", "ala10helix", "mix4x20", "ala10nve"]
    systems = [golden_system(n)[0] for n in names]
    eng = GpuEngine(systems, block=128, trace=True, trace_capacity=5000, hash_every=1000, max_pops=5000)
    eng.prepare()
    eng.run(5000)
    for k, n in enumerate(names):
        g = golden_system(n)[1]
        tr = eng.trace(k)
        m = min(len(tr), len(g["trace"]["target"]))
        for f in FIELDS:
            assert np.array_equal(tr[f][:m], g["trace"][f][:m].astype(tr[f].dtype)), (n, f)
        hs = eng.hashes(k)
        assert np.array_equal(hs[:, 1], g["trace"]["hashes"][: len(hs), 1]), n


def test_entry_points_drive_the_reference_loop(require_gpu):
    """The event loop of SGThread.c:75-151 driven from the host through
    sdmd_gpu_pop / advance / do_event / predict / update_tree reproduces the
    golden pops of the helix run (HB events included)."""
    from sdmd_amd.engine import GpuEngine

    sysm, g = golden_system("ala10helix")
    eng = GpuEngine([sysm], max_pops=400)
    eng.prepare()
    tr = g["trace"]
    for k in range(300):
        ev = eng.pop(0)
        assert ev["target"] == tr["target"][k] and ev["time"] == tr["dt"][k], k
        partner = ev["partner"] if ev["partner"] > 0 else 0
        eng.advance(0, ev["time"])
        assert bool(tr["exec"][k]) == (not ev["stale"]), k
        if not ev["stale"]:
            eng.do_event(0, ev["target"], partner)
            renew = [ev["target"]] + ([partner] if partner else [])
        else:
            renew = [ev["target"]]
        eng.predict(0, renew)
        eng.update_tree(0, renew)
        s = eng.state(0)
        for q in range(int(tr["nrenew"][k])):
            a = tr["renew"][k][q]
            i = a["atom"]
            assert (s.arr["ev_partner"][i], s.arr["ev_time"][i], s.arr["ev_counter"][i]) == \
                (a["partner"], a["time"], a["counter"]), (k, q)


@pytest.fixture(scope="module")
def large(tmp_path_factory):
    """64 chains of 40 residues, uniform-random sequences, random placement
    at 5 mM, cut-off 7.25 A, 300 K Andersen (26 830 atoms)."""
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    d = tmp_path_factory.mktemp("large")
    builder.write_input_dir(builder.random_peptides(64, 40, seed=2, mM=5.0), str(d), cutoff=7.25)
    s = System.load(str(d))
    print("large system: %d atoms" % s.natom, flush=True)
    return s


def test_large_system_bulk_and_1e5_pops_match_oracle(require_gpu, large):
    """Full size: the bulk prediction of every atom, then 10^5 consecutive
    pops (bulk re-prediction plus incremental updates), pop for pop."""
    from sdmd_amd.engine import GpuEngine

    pops = 100000
    eng = GpuEngine([large], block=1024, trace=True, trace_capacity=pops, max_pops=pops)
    eng.prepare()
    b = eng.state(0)
    o = oracle_lib.Oracle(large.arrays)
    o.bulk()
    ob = o.state()
    for f in ("ev_partner", "ev_type", "ev_sub", "ev_time", "ev_pot", "ev_pcounter"):
        assert np.array_equal(b.arr[f], ob.arr[f]), f
    eng.run(pops)
    tr = eng.trace(0)
    ref, _ = o.run(pops)
    assert len(tr) == pops
    for f in FIELDS:
        bad = np.flatnonzero(tr[f] != ref[f])
        assert len(bad) == 0, (f, int(bad[0]))
    g, c = eng.state(0), o.state()
    for f in ("x", "y", "z", "vx", "vy", "vz", "ev_partner", "ev_type", "ev_sub", "ev_time", "celllist",
              "cellx", "hb_bond"):
        assert np.array_equal(g.arr[f], c.arr[f]), f
    assert eng.stats(0)["error"] == 0
    eng.close()


def test_large_system_fast_screen_agrees_with_fp64(require_gpu, large):
    """The fast engine at full size: every screened prediction re-solved by
    a full FP64 scan agrees (partner, subtype, time)."""
    from sdmd_amd.engine import GpuEngine

    eng = GpuEngine([large], block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run(100000)
    st, fs = eng.stats(0), eng.fast_stats(0)
    assert st["error"] == 0 and fs["causal"] == 0
    assert fs["validated"] > 100000
    assert fs["agree"] == fs["validated"], fs
    eng.close()


def test_large_system_fast_schedule_invariance(require_gpu, large):
    """At full size (about 24 concurrent events per round) one event per
    round and full concurrency reach bit-identical states."""
    from conftest import assert_same_schedule_result, fast_schedule_states

    assert_same_schedule_result(fast_schedule_states(large, 40000), 20000, min_per_round=10)

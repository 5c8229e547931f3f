"""Fast mode (local-time state, FP32 screen + FP64 confirm, concurrent
independent events) against the CPU oracle and against itself.

The fast engine follows its own trajectory (per-atom local time, its own
thermostat stream), so the checks are the ones that trajectory allows:

  * bulk prediction on the oracle's start state: the same partner, event
    type and subtype for every atom and the same time to 1e-12 relative
    (NVE systems, where no random draw enters the prediction); 1e-9 only
    for atoms whose oracle time was inherited through CheckDuplication,
    i.e. solved from the mirrored pair;
  * every screened prediction re-solved against a full FP64 scan
    (validate mode): 100 % agreement;
  * schedule invariance: running to the same time with one event per round
    and with up to 64 concurrent events gives bit-identical states (the
    events executed together commute);
  * batching: systems launched together evolve exactly as each alone;
  * cell arrays: every atom sits in the cell its position indexes;
  * NVE energy drift of the same order as the reference's own, and total
    momentum conserved to rounding.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from conftest import fast_bulk_vs_oracle, golden_system
from sdmd_amd import _lib

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ev_partner", "ev_type", "ev_sub", "ev_time", "cellx", "celly", "cellz")


def energy(arr):
    d = arr.desc()
    ke, pe = C.c_double(), C.c_double()
    assert _lib.lib().sdmd_host_energy(C.byref(d), C.byref(ke), C.byref(pe)) == 0
    return ke.value + pe.value


@pytest.fixture(scope="module")
def mid_nve(tmp_path_factory):
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    d = tmp_path_factory.mktemp("midnve")
    builder.write_input_dir(builder.random_peptides(12, 30, seed=11, mM=5.0), str(d), cutoff=7.25,
                            thermostat="no")
    return System.load(str(d))


@pytest.fixture(scope="module")
def mid_nvt(tmp_path_factory):
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    d = tmp_path_factory.mktemp("midnvt")
    builder.write_input_dir(builder.random_peptides(12, 30, seed=11, mM=5.0), str(d), cutoff=7.25)
    return System.load(str(d))


def _bulk_check(system):
    e_solved, e_inh, n_inh = fast_bulk_vs_oracle(system)
    return e_solved, e_inh, n_inh


def test_fast_bulk_prediction_matches_oracle_ala10(require_gpu):
    s, _ = golden_system("ala10nve")
    _bulk_check(s)


def test_fast_bulk_prediction_matches_oracle_multichain(require_gpu, mid_nve):
    _bulk_check(mid_nve)


def test_fast_screen_boundary_cases_match_oracle(require_gpu, tmp_path):
    """Pairs placed on the screen's decision boundaries (separations at
    every shell radius of the pair times 1 +- eps, eps down to 1e-12;
    approaching, receding, b = 0, grazing at impact parameters on the
    radii; just beyond the last step; tests/screen_cases.py): the fast
    bulk prediction equals the oracle's for every atom, and a run with
    every prediction re-solved by a full FP64 scan agrees throughout."""
    import screen_cases
    from sdmd_amd.engine import GpuEngine

    sysm, _, _ = screen_cases.boundary_system(tmp_path)
    _bulk_check(sysm)
    eng = GpuEngine([sysm], block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run(20000)
    fs, st = eng.fast_stats(0), eng.stats(0)
    eng.close()
    assert st["error"] == 0 and fs["causal"] == 0
    assert fs["validated"] > 20000 and fs["agree"] == fs["validated"], fs


def test_fast_bulk_prediction_with_hbonds_matches_oracle(require_gpu):
    """A state with H bonds formed (the oracle's state after 1.5*10^5 pops
    of the reference's ala10.gro run, thermostat then switched off so no
    random draw enters): bonded H-bond partners (HBTime's bonded branch)
    and H-bond neighbour slots (HBNeighborTime) enter the bulk prediction."""
    from sdmd_amd.engine import System

    s, _ = golden_system("ala10helix")
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    o.run(150000, trace=False)
    arr = o.state()
    o.close()
    arr.scalars["thermostat"] = 0
    assert (arr.arr["hb_bond"][1:] > 0).sum() >= 4 and (arr.arr["nb_status"][1:] > 0).sum() >= 8
    _bulk_check(System(None, arr))


def test_fast_screen_agrees_with_full_fp64_scan(require_gpu, mid_nvt):
    from sdmd_amd.engine import GpuEngine

    eng = GpuEngine([mid_nvt], block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run(30000)
    fs = eng.fast_stats(0)
    st = eng.stats(0)
    assert st["error"] == 0
    assert fs["validated"] > 30000
    assert fs["agree"] == fs["validated"], fs
    # the FP32 time of a winner can be far off when the pair is about to
    # touch (r^2 - d^2 cancels); its error bound sends it to FP64
    assert fs["max_screen_rel_err"] < 1.0
    assert fs["screens"] > 10 * fs["confirms"]


def test_fast_schedule_invariance(require_gpu, mid_nvt):
    """One event per round and up to 64 concurrent events reach the same state."""
    from sdmd_amd.engine import GpuEngine

    states = []
    for mb in (1, 0):
        eng = GpuEngine([mid_nvt], block=256, mode="fast", max_batch=mb)
        eng.prepare()
        eng.run_until(0.5)
        states.append((eng.state(0), eng.stats(0), eng.fast_stats(0)))
        eng.close()
    (a, sa, fa), (b, sb, fb) = states
    assert fa["causal"] == 0 and fb["causal"] == 0
    assert sa["frame"] == sb["frame"] and sa["frame"] > 1000
    assert fb["batches"] < sb["pops"]            # events did run concurrently
    for f in FIELDS:
        assert np.array_equal(a.arr[f], b.arr[f]), f


def test_fast_batch_of_systems_is_independent(require_gpu, mid_nvt):
    from sdmd_amd.engine import GpuEngine

    s2, _ = golden_system("ala10ext")
    both = GpuEngine([mid_nvt, s2], block=256, mode="fast")
    both.prepare()
    both.run_until(0.3)
    alone = GpuEngine([mid_nvt], block=256, mode="fast")
    alone.prepare()
    alone.run_until(0.3)
    a, b = both.state(0), alone.state(0)
    for f in FIELDS:
        assert np.array_equal(a.arr[f], b.arr[f]), f


def test_fast_cells_track_positions(require_gpu, mid_nvt):
    from sdmd_amd.engine import GpuEngine

    eng = GpuEngine([mid_nvt], block=1024, mode="fast")
    eng.prepare()
    eng.run(20000)
    st = eng.state(0)
    a = st.arr
    sc = mid_nvt.arrays.scalars
    cs = np.array(sc["cellsize"][:3])
    N = mid_nvt.natom
    pos = np.stack([a["x"][1:], a["y"][1:], a["z"][1:]], 1)
    cell = np.stack([a["cellx"][1:], a["celly"][1:], a["cellz"][1:]], 1)
    lo = cell * cs
    assert np.all(pos >= lo - 1e-6) and np.all(pos < lo + cs + 1e-6)
    # linked lists rebuilt from the cell arrays hold every atom once
    seen = np.zeros(N + 1, int)
    cl = a["celllist"]
    for c in range(1, len(cl) - N):
        j = cl[N + c]
        while j:
            seen[j] += 1
            j = cl[j]
    assert np.all(seen[1:] == 1)


def test_fast_nve_energy_drift_within_reference(require_gpu):
    from sdmd_amd.engine import GpuEngine

    s, _ = golden_system("ala10nve")
    eng = GpuEngine([s], block=256, mode="fast")
    eng.prepare()
    e0 = energy(eng.state(0))
    fast = []
    for _ in range(10):
        eng.run(5000)
        fast.append(energy(eng.state(0)) - e0)
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    ref = []
    for _ in range(10):
        o.run(5000, trace=False)
        ref.append(energy(o.state()) - e0)
    # the reference's own NVE runs move the total energy in discrete
    # H-bond bookkeeping steps; the fast trajectory must stay within the
    # same envelope
    bound = 2 * max(abs(x) for x in ref) + 1.0
    assert max(abs(x) for x in fast) <= bound, (fast, ref)


def test_fast_nve_momentum_conserved(require_gpu, mid_nve):
    """Pair events exchange equal and opposite impulses: over 10^5 events of
    concurrent execution the total momentum moves by rounding only."""
    from sdmd_amd.engine import GpuEngine

    def momentum(a):
        m = a["mass"][1:]
        return np.array([(m * a[k][1:]).sum() for k in ("vx", "vy", "vz")]), (m * np.abs(a["vx"][1:])).sum()

    eng = GpuEngine([mid_nve], block=512, mode="fast")
    eng.prepare()
    p0, scale = momentum(eng.state(0).arr)
    eng.run_events(100000)
    st = eng.stats(0)
    p1, _ = momentum(eng.state(0).arr)
    eng.close()
    assert st["error"] == 0 and st["frame"] >= 100000
    assert st["per_type"][0] > 1000 and st["per_type"][1] > 10000     # collisions and bond events
    assert np.abs(p1 - p0).max() < 1e-10 * scale, (p0, p1, scale)


def _run_to(system, t_end, shape=None, graph=None):
    from sdmd_amd.engine import GpuEngine

    eng = GpuEngine([system], block=512, mode="fast")
    eng.prepare()
    picked = eng.autotune(20.0) if shape == "auto" else None
    if graph:
        eng.run_graph(*graph)            # bounded launches replayed from a HIP graph, several times
        eng.run_graph(*graph)
    eng.run_until(t_end)
    st, stats, fs = eng.state(0), eng.stats(0), eng.fast_stats(0)
    eng.close()
    return st, stats, fs, picked


def test_fast_graph_replay_and_autotune_keep_the_trajectory(require_gpu, mid_nvt):
    """The trajectory does not depend on how the run is cut into launches
    (HIP-graph replays of device-time-bounded launches, each ending with its
    closing drain) nor on the workgroup shape the autotuner picks: all reach
    the same state at the same time, bit for bit."""
    # late enough that the autotuner's trial launches (8 x 20 ms) end before it
    T = 10.0
    ref, sref, fref, _ = _run_to(mid_nvt, T)
    assert sref["error"] == 0 and sref["frame"] > 1000
    for kw in (dict(graph=(3, 2.0)), dict(shape="auto")):
        st, s, fs, picked = _run_to(mid_nvt, T, **kw)
        assert s["error"] == 0 and fs["causal"] == 0, kw
        assert s["frame"] == sref["frame"], kw
        for f in FIELDS:
            assert np.array_equal(st.arr[f], ref.arr[f]), (kw, f)
        if picked:
            assert picked[0] in (256, 512, 768, 1024) and picked[1] in (2, 4)


def test_fast_helper_workgroups_keep_the_trajectory(require_gpu, mid_nvt):
    """Re-predictions shared with helper workgroups on other CUs (published
    per round through a job word in global memory, agent-scope
    release/acquire) reach the same state at the same time, bit for bit, as
    the owner workgroup alone: a round's predictions read only state the
    round's events left, whoever runs them."""
    from sdmd_amd.engine import GpuEngine

    s2, _ = golden_system("mix4x20")
    runs = []
    for helpers in (-1, 0, 1):
        eng = GpuEngine([mid_nvt, s2], block=512, mode="fast", helpers=helpers)
        eng.prepare()
        eng.run_until(3.0)
        fs = [eng.fast_stats(k) for k in range(2)]
        runs.append(([eng.state(k) for k in range(2)], [eng.stats(k) for k in range(2)], fs))
        eng.close()
    assert runs[0][2][0]["helpers"] == 0 and runs[1][2][0]["helpers"] >= 1 and runs[2][2][0]["helpers"] == 1
    base_states, base_stats, _ = runs[0]
    for states, stats, fs in runs[1:]:
        for k in range(2):
            assert stats[k]["error"] == 0 and fs[k]["causal"] == 0
            assert stats[k]["frame"] == base_stats[k]["frame"] > 1000
            assert stats[k]["pair_evals"] == base_stats[k]["pair_evals"]
            for f in FIELDS:
                assert np.array_equal(states[k].arr[f], base_states[k].arr[f]), (k, f)


@pytest.fixture(scope="module")
def dense_nve(tmp_path_factory):
    """The dense single chain (76 atoms in the fullest 12 A cell), NVE."""
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    d = tmp_path_factory.mktemp("densenve")
    builder.write_input_dir(builder.single_chain(["TRP", "ARG", "LYS", "PHE", "TYR"] * 6, seed=8, box=72.0), str(d), cutoff=12.0,
                            thermostat="no")
    return System.load(str(d))


def test_fast_dense_cells(require_gpu, dense_nve):
    """Cells fuller than 64 atoms: the fast engine sizes its cell slots from
    the fullest cell; bulk prediction matches the oracle, screened
    predictions agree with the full FP64 scan, cells track positions."""
    from sdmd_amd.engine import GpuEngine

    _bulk_check(dense_nve)
    eng = GpuEngine([dense_nve], block=512, mode="fast", validate=True)
    eng.prepare()
    eng.run(20000)
    fs, st = eng.fast_stats(0), eng.stats(0)
    assert st["error"] == 0
    assert fs["validated"] > 20000 and fs["agree"] == fs["validated"], fs
    a = eng.state(0).arr
    eng.close()
    cs = np.array(dense_nve.arrays.scalars["cellsize"][:3])
    pos = np.stack([a["x"][1:], a["y"][1:], a["z"][1:]], 1)
    cell = np.stack([a["cellx"][1:], a["celly"][1:], a["cellz"][1:]], 1)
    lo = cell * cs
    assert np.all(pos >= lo - 1e-6) and np.all(pos < lo + cs + 1e-6)
    N = dense_nve.natom
    cl = a["celllist"]
    counts = []
    for c in range(1, len(cl) - N):
        n, j = 0, cl[N + c]
        while j:
            n, j = n + 1, cl[j]
        counts.append(n)
    assert sum(counts) == N and max(counts) > 64

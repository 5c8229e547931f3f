import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
for p in (ROOT, HERE, GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)

CASES = ["ala10ext", "ala10helix", "mix4x20", "ala10nve", "ala10cyl", "cyl8x20", "dense30"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def golden(name):
    out = {}
    for part in ("setup", "topo", "trace", "input"):
        with np.load(os.path.join(GOLDEN, "%s_%s.npz" % (name, part)), allow_pickle=False) as z:
            out[part] = {k: z[k] for k in z.files}
    return out


_SYSTEMS = {}


def golden_system(name):
    """Host-prepared System for a golden case (from the stored input text)."""
    from sdmd_amd.engine import System

    g = golden(name)
    seed = int(g["setup"]["seed"])
    wall = str(g["input"]["wall"]) if "wall" in g["input"] else None
    return System.from_text(str(g["input"]["gro"]), str(g["input"]["parameter"]), seed=seed, wall_text=wall), g


def gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("no HIP device visible: GPU tests must run on an MI355X")


STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ev_partner", "ev_type", "ev_sub", "ev_time", "cellx", "celly",
                "cellz")


def fast_schedule_states(system, events, block=512):
    """The fast engine run to one time twice: one event per round, and with
    its full concurrency.  Returns (state, stats, fast_stats) of each run;
    concurrent events commute, so the states must be bit-identical."""
    from sdmd_amd.engine import GpuEngine

    eng = GpuEngine([system], block=block, mode="fast")
    eng.prepare()
    eng.run_events(events)
    t_end = 0.9 * eng.fast_stats(0)["clock"]
    eng.close()
    out = []
    for mb in (1, 0):
        eng = GpuEngine([system], block=block, mode="fast", max_batch=mb)
        eng.prepare()
        eng.run_until(t_end)
        out.append((eng.state(0), eng.stats(0), eng.fast_stats(0)))
        eng.close()
    return out


def assert_same_schedule_result(runs, min_frames, min_per_round=2):
    (a, sa, fa), (b, sb, fb) = runs
    assert sa["error"] == 0 and sb["error"] == 0
    assert fa["causal"] == 0 and fb["causal"] == 0
    assert sa["frame"] == sb["frame"] and sa["frame"] > min_frames, (sa["frame"], sb["frame"])
    assert fb["batches"] * min_per_round < sb["pops"]     # events per round ran concurrently
    for f in STATE_FIELDS:
        assert np.array_equal(a.arr[f], b.arr[f]), f


def fast_bulk_vs_oracle(system, block=256, tol_solved=1e-12, tol_inherited=1e-9):
    """The fast engine's bulk prediction against the oracle's on the same
    state, with the clock moved to 0 (event times are relative, so the
    predictions do not change, and the engine's absolute times then carry
    no rounding of a large clock).  Partner, event type and subtype must be
    equal for every atom; times within tol_solved (relative) where the
    oracle solved the pair for the atom itself, within tol_inherited where
    it copied the partner's time through CheckDuplication
    (TimePrediction.c:1101-1117): that time was solved from the mirrored
    pair (r_ji instead of r_ij), whose naive quadratic rounds differently.
    Returns (max error over solved atoms, max over inherited ones, number
    of inherited atoms)."""
    import oracle_lib
    from sdmd_amd.engine import GpuEngine, System

    arr = system.arrays.copy()
    arr.scalars["currenttime"] = 0.0
    sysm = System(None, arr)
    eng = GpuEngine([sysm], block=block, mode="fast")
    eng.prepare()
    st = eng.state(0)
    eng.close()
    o = oracle_lib.Oracle(arr)
    assert o.bulk() == 0
    ref = o.state().arr
    inh = o.inherited()[1:]
    a = st.arr
    T = st.scalars["currenttime"]          # the engine's snapshot time: its earliest pending event
    for f in ("ev_partner", "ev_type", "ev_sub"):
        bad = np.flatnonzero(a[f][1:] != ref[f][1:])
        assert len(bad) == 0, (f, int(bad[0]) + 1 if len(bad) else None)
    rel = np.abs((a["ev_time"][1:] + T) - ref["ev_time"][1:]) / ref["ev_time"][1:]
    e_solved = float(rel[~inh].max()) if (~inh).any() else 0.0
    e_inh = float(rel[inh].max()) if inh.any() else 0.0
    assert e_solved <= tol_solved, ("solved", e_solved, int(np.argmax(np.where(inh, 0, rel))) + 1)
    assert e_inh <= tol_inherited, ("inherited", e_inh)
    return e_solved, e_inh, int(inh.sum())

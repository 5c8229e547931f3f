"""Parity of the HIP plain-FP64 path with the reference (golden traces from
the compiled reference) and with the CPU oracle, through the C ABI."""
import numpy as np
import pytest

from conftest import CASES, golden_system

pytestmark = pytest.mark.gpu

EVENT_FIELDS = ["ev_partner", "ev_type", "ev_sub", "ev_time", "ev_pot", "ev_pcounter"]
POP_FIELDS = ["target", "partner", "type", "sub", "exec", "dt", "pot", "rand_calls"]
RENEW_FIELDS = ["atom", "partner", "type", "sub", "counter", "pcounter", "time", "pot"]


def first_mismatch(mine, g, n):
    for f in POP_FIELDS:
        a = mine[f][:n]
        b = g["trace"][f][:n].astype(a.dtype)
        bad = np.flatnonzero(a != b)
        if len(bad):
            return (int(bad[0]), f, a[bad[0]], b[bad[0]])
    nr = g["trace"]["nrenew"][:n]
    if not np.array_equal(mine["nrenew"][:n], nr):
        return ("nrenew",)
    for q in range(2):
        sel = nr > q
        for f in RENEW_FIELDS:
            a = mine["r"][:n, q][f][sel]
            b = g["trace"]["renew"][:n, q][f][sel].astype(a.dtype)
            bad = np.flatnonzero(a != b)
            if len(bad):
                return (int(np.flatnonzero(sel)[bad[0]]), "r%d.%s" % (q, f), a[bad[0]], b[bad[0]])
    return None


@pytest.mark.parametrize("name", CASES)
def test_bulk_predict_matches_reference(require_gpu, name):
    from sdmd_amd.engine import GpuEngine

    sysm, g = golden_system(name)
    eng = GpuEngine([sysm], trace=False, max_pops=10)
    eng.prepare()
    st = eng.state(0)
    for f in EVENT_FIELDS:
        a = st.arr[f][1:]
        b = g["setup"][f][1:].astype(a.dtype)
        assert np.array_equal(a, b), (name, f, np.flatnonzero(a != b)[:5])
    node, leaf = eng.tree(0)
    assert np.array_equal(node[1:], g["setup"]["tree_node"][1:])
    assert np.array_equal(leaf[1:], g["setup"]["tree_leaf"][1:])
    eng.close()


@pytest.mark.parametrize("name", CASES)
def test_event_sequence_matches_reference(require_gpu, name):
    """Every popped event of the golden run: full records for the stored
    prefix, target/partner/type/executed flag of every pop, FNV hashes of
    positions/velocities/times/counters every 1000 pops for the whole run
    (1.7*10^5 pops -- more than 10^5 executed events -- for the ALA10
    extended chain; 2*10^5 for the helix)."""
    from sdmd_amd.engine import GpuEngine

    sysm, g = golden_system(name)
    npops = int(g["trace"]["npops"])
    nfull = len(g["trace"]["target"])
    eng = GpuEngine([sysm], trace=True, trace_capacity=npops, hash_every=1000, max_pops=npops)
    eng.prepare()
    eng.run(npops)
    st = eng.stats(0)
    assert st["error"] == 0 and st["pops"] == npops
    tr = eng.trace(0)
    assert len(tr) == npops
    mm = first_mismatch(tr, g, nfull)
    assert mm is None, (name, mm)
    for f in ("target", "partner", "type", "exec"):          # every pop of the run
        ref_seq = g["trace"]["seq_" + f]
        bad = np.flatnonzero(tr[f] != ref_seq.astype(tr[f].dtype))
        assert len(bad) == 0, (name, f, "first mismatch at pop", int(bad[0]) if len(bad) else None)
    hs = eng.hashes(0)
    ref = g["trace"]["hashes"]
    assert hs.shape == ref.shape
    bad = np.flatnonzero(hs[:, 1] != ref[:, 1])
    assert len(bad) == 0, (name, "first hash mismatch at pop", int(hs[bad[0], 0]))
    fin = eng.state(0)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(fin.arr[f][1:], g["trace"]["final_" + f][1:]), f
    assert fin.scalars["currenttime"] == float(g["trace"]["final_time"])
    counts = g["trace"]["counts_type"]
    assert np.array_equal(st["per_type"][: len(counts)], counts)
    assert st["recal"] == int(g["trace"]["recal"])
    eng.close()


@pytest.mark.parametrize("name,window", [("dense30", 64), ("dense30", 5), ("ala10helix", 7)])
def test_windowed_scan_matches_reference(require_gpu, name, window):
    """The neighbour scan in windows of `window` candidates (the staged
    cells bypassed, the cell lists re-walked per window, the running minimum
    and CheckDuplication carried across windows) pops the reference's
    events: full records for the stored prefix and the state hashes.
    dense30 holds 76 atoms in its fullest cell, more than the 64 staged per
    neighbour cell, so its default run takes the same path."""
    from sdmd_amd.engine import GpuEngine

    sysm, g = golden_system(name)
    nfull = len(g["trace"]["target"])
    npops = 20000
    eng = GpuEngine([sysm], trace=True, trace_capacity=npops, hash_every=1000, max_pops=npops, window=window)
    eng.prepare()
    st = eng.state(0)
    for f in EVENT_FIELDS:
        assert np.array_equal(st.arr[f][1:], g["setup"][f][1:].astype(st.arr[f].dtype)), f
    eng.run(npops)
    assert eng.stats(0)["error"] == 0
    tr = eng.trace(0)
    assert first_mismatch(tr, g, nfull) is None
    for f in ("target", "partner", "type", "exec"):
        assert np.array_equal(tr[f], g["trace"]["seq_" + f][:npops].astype(tr[f].dtype)), f
    hs = eng.hashes(0)
    ref = g["trace"]["hashes"]
    assert np.array_equal(hs[:, 1], ref[: len(hs), 1])
    eng.close()


def test_window_option_is_checked(require_gpu):
    from sdmd_amd._lib import SdmdError
    from sdmd_amd.engine import GpuEngine

    sysm, _ = golden_system("ala10ext")
    for bad in (-1, 27 * 64 + 1):
        with pytest.raises(SdmdError):
            GpuEngine([sysm], window=bad)

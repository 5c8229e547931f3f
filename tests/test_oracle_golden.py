"""The CPU oracle (the checker) against the compiled reference's golden
vectors: bulk prediction, tree, every popped event and state hashes."""
import numpy as np
import pytest

import oracle_lib
from conftest import CASES, golden_system

EVENT_FIELDS = ["ev_partner", "ev_type", "ev_sub", "ev_time", "ev_pot", "ev_pcounter"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_bulk_predict(name):
    sysm, g = golden_system(name)
    o = oracle_lib.Oracle(sysm.arrays)
    assert o.bulk() == 0
    st = o.state()
    for f in EVENT_FIELDS:
        a = st.arr[f][1:]
        assert np.array_equal(a, g["setup"][f][1:].astype(a.dtype)), f
    assert o.rand_calls() == int(g["setup"]["rand_calls"])
    node, leaf = o.tree()
    assert np.array_equal(node[1:], g["setup"]["tree_node"][1:])
    assert np.array_equal(leaf[1:], g["setup"]["tree_leaf"][1:])


@pytest.mark.parametrize("name", CASES)
def test_oracle_event_sequence(name):
    sysm, g = golden_system(name)
    tr = g["trace"]
    npops = int(tr["npops"])
    n = len(tr["target"])
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    mine, hs = o.run(npops, trace=True, hash_every=1000)
    for f in ["target", "partner", "type", "sub", "exec", "dt", "pot", "rand_calls"]:
        a = mine[f][:n]
        assert np.array_equal(a, tr[f].astype(a.dtype)), f
    sel = tr["nrenew"] > 0
    for f in ["atom", "partner", "type", "sub", "counter", "pcounter", "time", "pot"]:
        a = mine["r"][:n, 0][f][sel]
        assert np.array_equal(a, tr["renew"][:, 0][f][sel].astype(a.dtype)), f
    for f in ("target", "partner", "type", "exec"):          # every pop of the run
        assert np.array_equal(mine[f], tr["seq_" + f].astype(mine[f].dtype)), f
    assert np.array_equal(hs, tr["hashes"])
    st = o.state()
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(st.arr[f][1:], tr["final_" + f][1:])
    c = o.counters()
    assert c["error"] == 0
    assert np.array_equal(c["per_type"][: len(tr["counts_type"])], tr["counts_type"])
    assert c["recal"] == int(tr["recal"])


def test_helix_exercises_hb_events():
    _, g = golden_system("ala10helix")
    counts = g["trace"]["counts_type"]
    assert counts[2] > 100 and counts[3] > 100     # HB and HB-neighbour events


def test_config1_golden_run_covers_1e5_events():
    """The ALA10 extended-chain golden run pins more than 10^5 consecutive
    executed events of the reference."""
    _, g = golden_system("ala10ext")
    tr = g["trace"]
    assert int(tr["seq_exec"].sum()) >= 100000
    assert int(tr["counts_type"].sum()) == int(tr["seq_exec"].sum())


def test_nve_run_has_no_thermostat_events():
    _, g = golden_system("ala10nve")
    assert g["trace"]["counts_type"][4] == 0


def cell_counts(setup, N):
    """Atoms per cell from a cell list (heads at N + cell, Events.c:158-200)."""
    cl = setup["celllist"]
    out = []
    for head in cl[N + 1:]:
        n, j = 0, int(head)
        while j:
            n, j = n + 1, int(cl[j])
        out.append(n)
    return np.array(out)


def test_dense_case_fills_a_cell_beyond_the_staging_area():
    """dense30 puts more atoms in one cell than the plain engine stages per
    neighbour cell (64), so its predictions take the windowed scan."""
    sysm, g = golden_system("dense30")
    counts = cell_counts(g["setup"], sysm.natom)
    assert counts.sum() == sysm.natom
    assert counts.max() > 64, counts.max()

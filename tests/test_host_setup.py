"""C host setup (table/template/gro readers, SDEnergyMin, velocities, cell
list) against the reference's own state dumped at CreateCBT."""
import numpy as np
import pytest

from conftest import CASES, golden, golden_system

STATE_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "mass", "type", "aatype", "protein", "aanum",
                "cellx", "celly", "cellz", "hb_neighbor", "hb_role"]


@pytest.mark.parametrize("name", CASES)
def test_setup_state_bit_exact(name):
    sysm, g = golden_system(name)
    a = sysm.arrays
    st = g["setup"]
    assert a.scalars["natom"] == int(st["natom"])
    for f in STATE_FIELDS:
        mine = a.arr[f][1:]
        ref = st[f][1:].astype(mine.dtype)
        assert np.array_equal(mine, ref), (f, np.flatnonzero(mine != ref)[:5])
    assert np.array_equal(a.arr["celllist"], st["celllist"])
    assert list(a.scalars["cellnum"]) == list(st["cellnum"])
    assert np.array_equal(np.array(a.scalars["cellsize"][:]), st["cellsize"])
    assert np.array_equal(np.array(a.scalars["box"][:]), st["box"])


@pytest.mark.parametrize("name", CASES)
def test_topology_matches_reference(name):
    sysm, g = golden_system(name)
    a = sysm.arrays
    t = g["topo"]
    N = a.scalars["natom"]
    bonds = []
    for i in range(1, N + 1):
        for q in range(a.arr["bond_off"][i], a.arr["bond_off"][i + 1]):
            bonds.append((i, a.arr["bond_j"][q], a.arr["bond_dmin"][q], a.arr["bond_dmax"][q]))
    assert np.array_equal(np.array(bonds, dtype=float), t["bonds"])
    cons = []
    for i in range(1, N + 1):
        for q in range(a.arr["con_off"][i], a.arr["con_off"][i + 1]):
            c = a.arr["con"][q]
            cons.append((i, a.arr["con_j"][q], c["dmin"], c["dmax"], c["nstep"]))
    assert np.array_equal(np.array(cons, dtype=float), t["constr"])
    cm = set()
    for i in range(1, N + 1):
        for q in range(a.arr["cmap_off"][i], a.arr["cmap_off"][i + 1]):
            cm.add((i, int(a.arr["cmap_j"][q]), int(a.arr["cmap_bits"][q])))
    assert cm == set(map(tuple, t["cmap"].tolist()))


def test_tables_cover_the_force_field():
    sysm, _ = golden_system("ala10ext")
    pair = sysm.arrays.arr["pair"].reshape(33, 33)
    present = pair["dmin"] > 0
    assert np.array_equal(present, present.T)
    assert present.sum() >= 2 * 400          # 452 table rows, symmetric
    assert np.all(pair["nstep"][present] <= 13)
    d = sysm.arrays.arr["step_d"]
    for a in range(33):
        for b in range(33):
            p = pair[a, b]
            if p["nstep"] > 1:
                s = d[p["step0"]:p["step0"] + p["nstep"]]
                assert np.all(np.diff(s) > 0) and s[0] > p["dmin"]

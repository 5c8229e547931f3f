"""The reference's output and continue files against the compiled reference
(tests/golden/*_output.npz, written by its own SaveData and continue path),
on the CPU with the oracle's state (bit-exact with the reference's):

  * trajectory and connection-map frames (SaveGRO / SaveConnectionMap,
    DataSave.c:616-661) at the reference's own output pops: byte for byte;
  * savedData.dat (SaveDAT, DataSave.c:512-585): every state field equal to
    the reference's file (heap pointers, padding and the unused index-0
    entries excepted; the reference itself continues from the written file
    exactly as from its own, checked when the fixture is made);
  * reading the reference's savedData.dat (ReadBreakPoint,
    Initialization.c:360-509): the state equals the oracle's at that pop,
    the topology and tables equal a new run's, and continuing from it
    (CreateCBT, srand(saved seed)) pops the reference's continue run event
    for event;
  * a save / load / set-tree resume is bit-exact with an uninterrupted run.
"""
import os

import numpy as np
import pytest

import dat_format
import oracle_lib
from conftest import GOLDEN

CASES = ["ala10ext", "ala10helix"]
STATE = ("x", "y", "z", "vx", "vy", "vz", "type", "cellx", "celly", "cellz", "hb_bond", "hb_itype", "nb_partner",
         "nb_partnernum", "nb_status", "ev_partner", "ev_counter", "ev_pcounter", "ev_time", "ev_pot", "ev_type",
         "ev_sub", "celllist")
TOPO = ("mass", "aatype", "protein", "aanum", "hb_role", "hb_neighbor", "bond_off", "bond_j", "bond_dmin",
        "bond_dmax", "con_off", "con_j", "cmap_off", "cmap_j", "cmap_bits", "pair", "hbpair", "step_d", "step_e",
        "step_acc")


def fixture(name):
    with np.load(os.path.join(GOLDEN, name + "_output.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def new_system(g):
    from sdmd_amd.engine import System

    return System.from_text(str(g["gro"]), str(g["parameter"]))


def continue_dir(g, tmp, dat_bytes):
    """An input directory for a continue run (parameter.txt says continue,
    Time as the fixture's continue run) and the savedData.dat path."""
    from sdmd_amd.builder import LIB_DIR

    d = os.path.join(tmp, "cont_in")
    os.makedirs(d, exist_ok=True)
    lines = str(g["parameter"]).splitlines()
    lines[0] = "%-18s %s" % ("Continue?", "continue")
    lines[1] = "%-18s %r" % ("Time", float(g["continue_time"]))
    with open(os.path.join(d, "parameter.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    if not os.path.lexists(os.path.join(d, "Library_Ding")):
        os.symlink(LIB_DIR, os.path.join(d, "Library_Ding"))
    dat = os.path.join(tmp, "savedData.dat")
    with open(dat, "wb") as fh:
        fh.write(dat_bytes)
    return d, dat


def dyn_map(a, N):
    """{(i, j): bits} of the dynamic slots."""
    out = {}
    dj, db = a["dyn_j"].reshape(N + 1, -1), a["dyn_bits"].reshape(N + 1, -1)
    for i in range(1, N + 1):
        for j, b in zip(dj[i], db[i]):
            if j:
                out[(i, int(j))] = int(b)
    return out


@pytest.mark.parametrize("name", CASES)
def test_trajectory_and_connection_map_frames_match_reference(name, tmp_path):
    from sdmd_amd import output

    g = fixture(name)
    sysm = new_system(g)
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    done = 0
    trj, cnt = str(tmp_path / "out_trj.gro"), str(tmp_path / "out_cnt.txt")
    for p in g["dump_pops"]:
        o.run(int(p) - done, trace=False)
        done = int(p)
        st = o.state()
        output.write_gro(sysm, st, trj)
        output.write_cnt(st, cnt)
    assert open(trj).read() == str(g["trj"])
    assert open(cnt).read() == str(g["cnt"])


@pytest.mark.parametrize("name", CASES)
def test_saved_dat_fields_match_reference(name, tmp_path):
    from sdmd_amd import output

    g = fixture(name)
    sysm = new_system(g)
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    o.run(int(g["dump_pops"][-1]), trace=False)
    st = o.state()
    c = o.counters()
    info = output.dat_info(sysm, st, frame=int(c["per_type"].sum()), per_type=c["per_type"])
    info.timestep, info.outputrate = float(g["time"]), float(g["dump_rate"])
    path = str(tmp_path / "ours.dat")
    output.save_dat(sysm, st, path, info)
    ours = dat_format.state_fields(dat_format.read(open(path, "rb").read()))
    ref = dat_format.state_fields(dat_format.read(g["dat"].tobytes()))
    assert ours.keys() == ref.keys()
    for k in ref:
        assert np.array_equal(np.asarray(ours[k]), np.asarray(ref[k])), k
    if name == "ala10helix":          # H bonds formed at the save: dynamic map words written
        assert np.any(ref["cmap"] & 12) and ref["hb"][0] > 0


@pytest.mark.parametrize("name", CASES)
def test_load_reference_dat_gives_the_reference_state(name, tmp_path):
    from sdmd_amd import output

    g = fixture(name)
    sysm = new_system(g)
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    o.run(int(g["dump_pops"][-1]), trace=False)
    ref = o.state()
    d, dat = continue_dir(g, str(tmp_path), g["dat"].tobytes())
    loaded, info = output.load_dat(d, dat)
    a, b = loaded.arrays.arr, ref.arr
    for f in STATE:
        assert np.array_equal(a[f], b[f]), f
    for f in TOPO:
        assert np.array_equal(a[f], sysm.arrays.arr[f]), f
    N = sysm.natom
    assert dyn_map(a, N) == dyn_map(b, N)
    sc = loaded.arrays.scalars
    assert sc["currenttime"] == ref.scalars["currenttime"]
    assert sc["rand_calls"] == 0 and sc["seed"] == sysm.arrays.scalars["seed"]
    c = o.counters()
    assert info.frame == int(c["per_type"].sum())
    assert [info.event_sums[k] for k in range(7)] == [int(c["per_type"][t]) for t in output.EVENT_SUM_TYPES]


@pytest.mark.parametrize("name", CASES)
def test_continue_from_reference_dat_matches_reference(name, tmp_path):
    from sdmd_amd import output

    g = fixture(name)
    d, dat = continue_dir(g, str(tmp_path), g["dat"].tobytes())
    loaded, _ = output.load_dat(d, dat)
    o = oracle_lib.Oracle(loaded.arrays)          # srand(saved seed): no rand() call yet
    o.create_cbt()                                 # CreateCBT, SGThread.c:62
    node, leaf = o.tree()
    assert np.array_equal(node[1:], g["cont_tree_node"][1:]) and np.array_equal(leaf[1:], g["cont_tree_leaf"][1:])
    n = len(g["cont_target"])
    tr, hs = o.run(n, trace=True, hash_every=1000)
    for f in ("target", "partner", "type", "sub", "exec", "dt", "pot", "rand_calls"):
        assert np.array_equal(tr[f], g["cont_" + f].astype(tr[f].dtype)), f
    assert np.array_equal(hs, g["cont_hashes"])
    st = o.state()
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(st.arr[f][1:], g["cont_final_" + f][1:]), f


@pytest.mark.parametrize("name", CASES)
def test_save_load_resume_is_bit_exact(name, tmp_path):
    """2P pops in one run = P pops, savedData.dat + resume record, load, P pops."""
    from sdmd_amd import output

    g = fixture(name)
    sysm = new_system(g)
    P = 6000
    full = oracle_lib.Oracle(sysm.arrays)
    full.bulk()
    ref_tr, ref_hs = full.run(2 * P, trace=True, hash_every=1000)
    first = oracle_lib.Oracle(sysm.arrays)
    first.bulk()
    first.run(P, trace=False)
    st = first.state()
    c = first.counters()
    info = output.dat_info(sysm, st, frame=int(c["per_type"].sum()), per_type=c["per_type"])
    info.timestep, info.outputrate = float(g["time"]), float(g["dump_rate"])
    dat, rsm = str(tmp_path / "savedData.dat"), str(tmp_path / "savedData.resume")
    output.save_dat(sysm, st, dat, info)
    node, leaf = first.tree()
    output.save_resume(rsm, node, leaf, st.scalars["seed"], first.rand_calls(), P, c["recal"])
    d, _ = continue_dir(g, str(tmp_path), open(dat, "rb").read())
    loaded, _ = output.load_dat(d, dat)
    r = output.load_resume(rsm, sysm.natom)
    second = oracle_lib.Oracle(loaded.arrays, seed=r["seed"], rand_skip=r["rand_calls"])
    second.set_tree(r["node"], r["leaf"])
    tr, _ = second.run(P, trace=True)
    for f in ("target", "partner", "type", "sub", "exec", "dt", "pot"):
        assert np.array_equal(tr[f], ref_tr[f][P:]), f
    a, b = second.state().arr, full.state().arr
    for f in STATE:
        assert np.array_equal(a[f], b[f]), f

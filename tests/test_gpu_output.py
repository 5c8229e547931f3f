"""The reference's output and continue files through the plain engine on
the GPU (C ABI), against the compiled reference's own files
(tests/golden/*_output.npz):

  * SingleThreadRun's output cadence (RunOutput: sdmd_gpu_run_to_output,
    then SaveGRO / SaveConnectionMap / SaveDAT): the same output pops, the
    trajectory and connection-map files byte for byte, savedData.dat's state
    fields equal;
  * a continue run from the reference's savedData.dat (loaded, tree built
    by CreateCBT, RNG from srand(saved seed)): the reference's continue run
    pop for pop, state hashes and final state;
  * save / load / set-tree resume of the engine: bit-exact with one
    uninterrupted run.
"""
import numpy as np
import pytest

import dat_format
from test_output import CASES, STATE, continue_dir, fixture, new_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CASES)
def test_run_output_matches_reference_files(require_gpu, name, tmp_path):
    from sdmd_amd.engine import GpuEngine
    from sdmd_amd.output import RunOutput

    g = fixture(name)
    sysm = new_system(g)
    cap = int(g["dump_pops"][-1]) + 20000
    eng = GpuEngine([sysm], block=256, max_pops=cap)
    eng.prepare()
    out = RunOutput(eng, sysm, str(tmp_path / "out"), float(g["dump_rate"]), float(g["time"]))
    dumps = out.run(max_pops=cap)
    assert [d[0] for d in dumps] == [int(p) for p in g["dump_pops"]]
    assert [d[1] for d in dumps] == [int(f) for f in g["dump_frames"]]
    assert open(out.path("out_trj.gro")).read() == str(g["trj"])
    assert open(out.path("out_cnt.txt")).read() == str(g["cnt"])
    ours = dat_format.state_fields(dat_format.read(open(out.path("savedData.dat"), "rb").read()))
    ref = dat_format.state_fields(dat_format.read(g["dat"].tobytes()))
    for k in ref:
        assert np.array_equal(np.asarray(ours[k]), np.asarray(ref[k])), k
    eng.close()


@pytest.mark.parametrize("name", CASES)
def test_continue_from_reference_dat_on_gpu(require_gpu, name, tmp_path):
    from sdmd_amd import output
    from sdmd_amd.engine import GpuEngine

    g = fixture(name)
    d, dat = continue_dir(g, str(tmp_path), g["dat"].tobytes())
    loaded, _ = output.load_dat(d, dat)
    n = len(g["cont_target"])
    eng = GpuEngine([loaded], block=256, trace=True, trace_capacity=n, hash_every=1000, max_pops=n)
    eng.create_tree(0)                               # CreateCBT on the saved event records
    node, leaf = eng.tree(0)
    assert np.array_equal(node[1:], g["cont_tree_node"][1:]) and np.array_equal(leaf[1:], g["cont_tree_leaf"][1:])
    eng.run(n)
    tr = eng.trace(0)
    for f in ("target", "partner", "type", "sub", "exec", "dt", "pot", "rand_calls"):
        assert np.array_equal(tr[f], g["cont_" + f].astype(tr[f].dtype)), f
    assert np.array_equal(eng.hashes(0), g["cont_hashes"])
    st = eng.state(0)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(st.arr[f][1:], g["cont_final_" + f][1:]), f
    assert st.scalars["currenttime"] == float(g["cont_final_time"])
    eng.close()


@pytest.mark.parametrize("name", CASES)
def test_engine_resume_is_bit_exact(require_gpu, name, tmp_path):
    """2P pops in one run = P pops, savedData.dat + resume record, a new
    context from the loaded state with the saved tree, P pops."""
    from sdmd_amd import output
    from sdmd_amd.engine import GpuEngine

    g = fixture(name)
    sysm = new_system(g)
    P = 10000
    full = GpuEngine([sysm], block=256, trace=True, trace_capacity=2 * P, max_pops=2 * P)
    full.prepare()
    full.run(2 * P)
    ref_tr, ref = full.trace(0), full.state(0)
    full.close()
    first = GpuEngine([sysm], block=256, max_pops=P)
    first.prepare()
    first.run(P)
    st, stats = first.state(0), first.stats(0)
    info = output.dat_info(sysm, st, frame=stats["frame"], per_type=stats["per_type"])
    dat, rsm = str(tmp_path / "savedData.dat"), str(tmp_path / "savedData.resume")
    output.save_dat(sysm, st, dat, info)
    node, leaf = first.tree(0)
    output.save_resume(rsm, node, leaf, st.scalars["seed"], first.rand_position(0), stats["pops"], stats["recal"])
    first.close()
    d, _ = continue_dir(g, str(tmp_path), open(dat, "rb").read())
    loaded, _ = output.load_dat(d, dat)
    r = output.load_resume(rsm, sysm.natom)
    loaded.arrays.scalars["rand_calls"] = r["rand_calls"]        # the thermostat stream continues
    second = GpuEngine([loaded], block=256, trace=True, trace_capacity=P, max_pops=P)
    second.set_tree(0, r["node"], r["leaf"])
    second.run(P)
    tr = second.trace(0)
    for f in ("target", "partner", "type", "sub", "exec", "dt", "pot"):
        assert np.array_equal(tr[f], ref_tr[f][P:]), f
    a = second.state(0)
    for f in STATE:
        assert np.array_equal(a.arr[f], ref.arr[f]), f
    assert a.scalars["currenttime"] == ref.scalars["currenttime"]
    second.close()

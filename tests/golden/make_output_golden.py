"""Generate the output/continue fixtures under tests/golden/ from the
compiled reference (only where /root/reference exists).

For each case the reference (oracle/_ref/sdmd_ref, hooks with SDMD_SAVE=1
so SaveData writes its files) runs a new simulation to `Time` with output
every Dump_Rate: out_trj.gro, out_cnt.txt and the last savedData.dat, and
the pop count of every output (dumps.txt).  Then it continues from that
savedData.dat ("Continue? continue", ReadBreakPoint + CreateCBT) for a
bounded number of pops, traced by the hooks.

Cross-check at generation time: this library's writer (sdmd_save_dat) is
given the state the CPU oracle reaches at the last output, the reference
continues from that file as well, and both continue runs must pop the same
events bit for bit -- so the reference itself accepts the written file as
its own.

  <name>_output.npz  input texts, the reference's output files (text and
                     bytes), output pops, and the continue run: its state
                     after reading (setup), tree, per-pop records, hashes
                     and final positions/velocities
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import make_golden  # noqa: E402
import refdump  # noqa: E402

REF_BIN = make_golden.REF_BIN

CASES = {
    # name: (input kind, parameters, continue pops)
    "ala10ext": ("ala10ext", dict(temperature=0.5962, cutoff=9.0, time=20.0, dump_rate=5.0), 4000),
    "ala10helix": ("helix", dict(temperature=0.5962, cutoff=9.0, time=60.0, dump_rate=15.0), 4000),
}


def set_continue(inp, dst, time):
    """dst: a copy of input dir inp whose parameter.txt says continue, Time = time."""
    os.makedirs(dst, exist_ok=True)
    for name in os.listdir(inp):
        if name != "parameter.txt":
            os.symlink(os.path.join(inp, name), os.path.join(dst, name))
    lines = open(os.path.join(inp, "parameter.txt")).read().splitlines()
    lines[0] = "%-18s %s" % ("Continue?", "continue")
    lines[1] = "%-18s %r" % ("Time", float(time))
    with open(os.path.join(dst, "parameter.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


def run(inp, out, dump, env_extra):
    os.makedirs(out, exist_ok=True)
    os.makedirs(dump, exist_ok=True)
    env = dict(os.environ, SDMD_DUMP=dump, SDMD_HASH_EVERY="1000", **env_extra)
    with open(os.path.join(dump, "stdout.txt"), "w") as fh:
        subprocess.run([REF_BIN, "-f", "coordinate.gro", "-i", inp, "-o", out], env=env, stdout=fh,
                       stderr=subprocess.STDOUT, check=True)


def ours_at(inp, pops, timestep, outputrate, path):
    """This library's savedData.dat of the oracle's state after `pops` pops."""
    import oracle_lib
    from sdmd_amd import output
    from sdmd_amd.engine import System

    sysm = System.load(inp)
    o = oracle_lib.Oracle(sysm.arrays)
    o.bulk()
    o.run(pops, trace=False)
    st = o.state()
    c = o.counters()
    info = output.dat_info(sysm, st, frame=int(c["per_type"].sum()), per_type=c["per_type"])
    info.timestep = timestep
    info.outputrate = outputrate
    output.save_dat(sysm, st, path, info)


def main():
    if not os.path.isdir(make_golden.REF):
        sys.exit("reference tree not present; fixtures are generated only where it is")
    make_golden.build_ref()
    tmp = tempfile.mkdtemp(prefix="sdmd_outgold_")
    try:
        for name, (kind, params, cpops) in CASES.items():
            inp = os.path.join(tmp, name + "_in")
            make_golden.make_input(kind, inp, params)
            out, dump = os.path.join(tmp, name + "_out"), os.path.join(tmp, name + "_dump")
            run(inp, out, dump, {"SDMD_SAVE": "1"})
            dumps = np.array([[int(a), int(b)] for a, b, _ in (l.split() for l in open(os.path.join(dump, "dumps.txt")))])
            times = np.array([float.fromhex(l.split()[2]) for l in open(os.path.join(dump, "dumps.txt"))])
            dat = open(os.path.join(out, "savedData.dat"), "rb").read()
            t_cont = times[-1] + 10 * params["dump_rate"]
            # the reference continuing from its own file and from this library's
            traces = []
            for tag, src in (("ref", None), ("ours", os.path.join(tmp, name + "_ours.dat"))):
                cin, cout, cdump = (os.path.join(tmp, "%s_%s_%s" % (name, tag, k)) for k in ("cin", "cout", "cdump"))
                set_continue(inp, cin, t_cont)
                os.makedirs(cout, exist_ok=True)
                if src is None:
                    shutil.copy(os.path.join(out, "savedData.dat"), os.path.join(cout, "savedData.dat"))
                else:
                    ours_at(inp, int(dumps[-1, 0]), params["time"], params["dump_rate"], src)
                    shutil.copy(src, os.path.join(cout, "savedData.dat"))
                run(cin, cout, cdump, {"SDMD_MAX_POPS": str(cpops)})
                traces.append(refdump.load_dir(cdump))
            a, b = traces
            assert np.array_equal(a["trace"], b["trace"]), name + ": the reference continues differently from our file"
            assert np.array_equal(a["hashes"], b["hashes"]), name
            r = a
            tr = r["trace"]
            with open(os.path.join(inp, "coordinate.gro")) as fh:
                gro = fh.read()
            with open(os.path.join(inp, "parameter.txt")) as fh:
                par = fh.read()
            np.savez_compressed(
                os.path.join(HERE, name + "_output.npz"),
                gro=np.array(gro), parameter=np.array(par), time=np.array(params["time"]),
                dump_rate=np.array(params["dump_rate"]), continue_time=np.array(t_cont),
                trj=np.array(open(os.path.join(out, "out_trj.gro")).read()),
                cnt=np.array(open(os.path.join(out, "out_cnt.txt")).read()),
                dat=np.frombuffer(dat, np.uint8), dump_pops=dumps[:, 0], dump_frames=dumps[:, 1], dump_times=times,
                cont_tree_node=r["tree_node"], cont_tree_leaf=r["tree_leaf"],
                cont_target=tr["target"].astype(np.int32), cont_partner=tr["partner"].astype(np.int32),
                cont_type=tr["type"].astype(np.int8), cont_sub=tr["sub"].astype(np.int8),
                cont_exec=tr["exec"].astype(np.int8), cont_dt=tr["dt"], cont_pot=tr["pot"],
                cont_rand_calls=tr["rand_calls"], cont_hashes=r["hashes"],
                cont_final_x=r["final"]["x"], cont_final_y=r["final"]["y"], cont_final_z=r["final"]["z"],
                cont_final_vx=r["final"]["vx"], cont_final_vy=r["final"]["vy"], cont_final_vz=r["final"]["vz"],
                cont_final_time=r["final"]["currenttime"])
            print("wrote", name, "outputs at pops", dumps[:, 0].tolist(), "; continue", len(tr), "pops")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()

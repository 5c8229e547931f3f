"""Generate the golden fixtures under tests/golden/ from the compiled reference.

Runs only in a container where /root/reference exists: builds oracle/_ref
(the reference sDMD from its own sources + the --wrap hooks of
oracle/ref_hooks.c), writes seeded inputs with sdmd_amd.builder (or takes
the reference's own input/ala10.gro), runs the reference, and stores
compact .npz fixtures:

  <name>_setup.npz  state after the reference's bulk Predict (positions,
                    velocities, cells, cell list, HB roles, event records,
                    RNG position) plus the CBT arrays
  <name>_topo.npz   bonds / constraints / connection map / pair tables
  <name>_trace.npz  every popped event (target, partner, type, subtype,
                    executed flag, dt, potential, RNG position) and the
                    re-predicted atoms' records for the first pops; target,
                    partner, type and executed flag of every pop of the run
                    (seq_*); FNV state hashes every 1000 pops for the whole
                    run

The fixtures are data (inputs and expected outputs); no reference source
is copied.
"""
import argparse
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

import refdump  # noqa: E402
from sdmd_amd import builder  # noqa: E402

REF = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "sdmd_ref")

CASES = {
    # name: (builder, parameter overrides, pops, full-record pops)
    # 170 000 pops: more than 10^5 executed events (about 0.6 executed per pop)
    "ala10ext": ("ala10ext", dict(temperature=0.5962, cutoff=9.0), 170000, 5000),
    "ala10helix": ("helix", dict(temperature=0.5962, cutoff=9.0), 200000, 5000),
    "mix4x20": ("mix", dict(temperature=0.5962, cutoff=9.0), 100000, 3000),
    "ala10nve": ("ala10ext", dict(temperature=0.5962, cutoff=9.0, thermostat="no"), 100000, 2000),
    "ala10cyl": ("ala10cyl", dict(temperature=0.5962, cutoff=9.0, wall="smooth", wall_type="cylinder"),
                 100000, 5000),
    "cyl8x20": ("cyl8x20", dict(temperature=0.5962, cutoff=7.25, wall="smooth", wall_type="cylinder"),
                300000, 3000),
    # 76 atoms in the fullest 12 A cell: the windowed neighbour scan
    "dense30": ("dense", dict(temperature=0.5962, cutoff=12.0), 100000, 3000),
}


DENSE_SEQ = ["TRP", "ARG", "LYS", "PHE", "TYR"] * 6


def build_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + REF], check=True)


def make_input(kind, d, params):
    if kind == "ala10ext":
        builder.write_input_dir(builder.ala10_extended(seed=1), d, **params)
    elif kind == "mix":
        builder.write_input_dir(builder.random_peptides(4, 20, seed=7, box=80.0), d, **params)
    elif kind == "ala10cyl":
        builder.write_input_dir(builder.ala10_cylinder(seed=1), d, **params)
    elif kind == "cyl8x20":
        builder.write_input_dir(builder.cylinder_peptides(8, 20, seed=5, radius=40.0, box_x=160.0, margin=4.0), d, **params)
    elif kind == "dense":
        builder.write_input_dir(builder.single_chain(DENSE_SEQ, seed=8, box=72.0), d, **params)
    elif kind == "helix":
        os.makedirs(d, exist_ok=True)
        shutil.copy(os.path.join(REF, "input", "ala10.gro"), os.path.join(d, "coordinate.gro"))
        builder.write_parameter(os.path.join(d, "parameter.txt"), **params)
        os.symlink(builder.LIB_DIR, os.path.join(d, "Library_Ding"))
    else:
        raise ValueError(kind)


def run_ref(inp, dump, pops, seed=None):
    os.makedirs(dump, exist_ok=True)
    out = os.path.join(dump, "out")
    os.makedirs(out, exist_ok=True)
    env = dict(os.environ, SDMD_DUMP=dump, SDMD_MAX_POPS=str(pops), SDMD_HASH_EVERY="1000")
    if seed is not None:
        env["SDMD_SEED"] = str(seed)
    with open(os.path.join(dump, "stdout.txt"), "w") as fh:
        subprocess.run([REF_BIN, "-f", "coordinate.gro", "-i", inp, "-o", out], env=env,
                       stdout=fh, stderr=subprocess.STDOUT, check=True)


def save_case(name, dump, full, outdir, inp):
    r = refdump.load_dir(dump)
    st = r["setup"]
    arrays = {k: v for k, v in st.items() if isinstance(v, np.ndarray)}
    scal = {k: np.array(v) for k, v in st.items() if not isinstance(v, np.ndarray)}
    np.savez_compressed(os.path.join(outdir, name + "_setup.npz"), tree_node=r["tree_node"],
                        tree_leaf=r["tree_leaf"], **arrays, **scal)
    topo = r["topo"]
    bonds = [(i, j, a, b) for i, l in topo["bonds"].items() for (j, a, b) in l]
    cons = [(i, j, a, b, len(s)) for i, l in topo["constr"].items() for (j, a, b, s) in l]
    cmap = [(i, j, v) for i, row in topo["cmap"].items() for j, v in row.items()]
    np.savez_compressed(os.path.join(outdir, name + "_topo.npz"),
                        bonds=np.array(bonds, dtype=float), constr=np.array(cons, dtype=float),
                        cmap=np.array(cmap, dtype=np.int64))
    tr = r["trace"]
    f = tr[:full]
    np.savez_compressed(os.path.join(outdir, name + "_trace.npz"), npops=np.array(len(tr)),
                        target=f["target"].astype(np.int32), partner=f["partner"].astype(np.int32),
                        type=f["type"].astype(np.int8), sub=f["sub"].astype(np.int8),
                        seq_target=tr["target"].astype(np.int32), seq_partner=tr["partner"].astype(np.int32),
                        seq_type=tr["type"].astype(np.int8), seq_exec=tr["exec"].astype(np.int8),
                        exec=f["exec"].astype(np.int8), dt=f["dt"], pot=f["pot"],
                        rand_calls=f["rand_calls"], renew=f["r"], nrenew=f["nrenew"],
                        hashes=r["hashes"], counts_type=np.bincount(tr["type"][tr["exec"] == 1], minlength=16),
                        recal=np.array(int((tr["exec"] == 0).sum())),
                        final_x=r["final"]["x"], final_y=r["final"]["y"], final_z=r["final"]["z"],
                        final_vx=r["final"]["vx"], final_vy=r["final"]["vy"], final_vz=r["final"]["vz"],
                        final_time=r["final"]["currenttime"],
                        energy_setup=r["energy_setup"], energy_final=r["energy_final"])
    with open(os.path.join(inp, "coordinate.gro")) as fh:
        gro = fh.read()
    with open(os.path.join(inp, "parameter.txt")) as fh:
        par = fh.read()
    extra = {}
    wall = os.path.join(inp, "Library_Ding", "Wall.txt")
    if not os.path.islink(os.path.join(inp, "Library_Ding")) and os.path.exists(wall):
        with open(wall) as fh:
            extra["wall"] = np.array(fh.read())
    np.savez_compressed(os.path.join(outdir, name + "_input.npz"), gro=np.array(gro), parameter=np.array(par),
                        **extra)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", nargs="*", default=list(CASES))
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are generated only where it is")
    build_ref()
    tmp = tempfile.mkdtemp(prefix="sdmd_golden_")
    try:
        for name in args.cases:
            kind, params, pops, full = CASES[name]
            inp = os.path.join(tmp, name + "_in")
            dump = os.path.join(tmp, name + "_dump")
            make_input(kind, inp, params)
            run_ref(inp, dump, pops)
            save_case(name, dump, full, args.out, inp)
            print("wrote", name)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()

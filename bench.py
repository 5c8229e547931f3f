"""Benchmark: aggregate events/s and pair-time evaluations/s of the sDMD hot
path on one MI355X.

Workload (N=1): 64 chains of 40 residues, each a uniform-random sequence over
the 20 amino acids (seeded), all-atom Ding force field (26830 atoms), random
non-overlapping placement in a 277 A periodic box (5 mM per chain), cut-off
7.25 A, 300 K (T* 0.5962) with the Andersen thermostat, plain FP64 path
(bit-exact with the reference).  A step is one call of the device event loop
over `--pops` pops; bulk prediction (Predict(1..N) + CreateCBT) runs before
the timed region.

Prints one JSON line.  `roofline` prices the dominant kernel (the persistent
event-loop kernel) with algorithmic bytes: per pop N*88 B for the position
and clock sweep plus (n_nb+1)*56 B + 40 B + 2*ceil(log2 N)*12 B per
re-predicted atom.  `cpu_baseline` times the reference built from its own
sources (oracle/_ref/sdmd_ref) on a bounded sample of the same input on one
host core, or the CPU oracle when that binary is absent.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def build_workload(tmp, chains=64, length=40, seed=2, mM=5.0, cutoff=7.25, temperature=0.5962):
    from sdmd_amd import builder

    t0 = time.time()
    system = builder.random_peptides(chains, length, seed=seed, mM=mM)
    builder.write_input_dir(system, tmp, cutoff=cutoff, temperature=temperature)
    return system, time.time() - t0


def cpu_baseline(indir, pops, natom):
    ref = os.path.join(ROOT, "oracle", "_ref", "sdmd_ref")
    out = tempfile.mkdtemp(prefix="sdmd_cpu_")
    if os.path.exists(ref):
        env = dict(os.environ, SDMD_LIGHT="1", SDMD_DUMP=out, SDMD_MAX_POPS=str(pops))
        cmd = [ref, "-f", "coordinate.gro", "-i", indir, "-o", out]
        if subprocess.run(["which", "taskset"], capture_output=True).returncode == 0:
            cmd = ["taskset", "-c", "0"] + cmd
        with open(os.path.join(out, "stdout.txt"), "w") as fh:
            subprocess.run(cmd, env=env, stdout=fh, stderr=subprocess.STDOUT, check=True)
        tok = open(os.path.join(out, "timing.txt")).read().split()
        vals = dict(zip(tok[0::2], tok[1::2]))
        ev, sec = int(vals["events"]), float(vals["seconds"])
        return {"value": ev / sec, "unit": "events/s", "cores": 1, "kind": "reference",
                "sample": "reference sDMD (own sources, -O2), same input, first %d pops (%d events) "
                          "of its event loop, %.2f s, one core" % (pops, ev, sec)}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from sdmd_amd.engine import System

    s = System.load(indir)
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    t0 = time.perf_counter()
    o.run(pops, trace=False)
    sec = time.perf_counter() - t0
    ev = int(o.counters()["per_type"].sum())
    return {"value": ev / sec, "unit": "events/s", "cores": 1, "kind": "port",
            "sample": "CPU oracle, same input, first %d pops (%d events), %.2f s, one core" % (pops, ev, sec)}


def calibrate_copy(nbytes=1 << 30):
    """A plain device copy of known size, for scaling rocprofv3 FETCH/WRITE_SIZE."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pops", type=int, default=20000, help="event-loop pops per step")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-pops", type=int, default=20000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the boxes here have a single MI355X; "
              "systems shard across GPUs by whole systems (see DESIGN.md), N=%d is not run" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    from sdmd_amd.engine import GpuEngine, System

    tmp = tempfile.mkdtemp(prefix="sdmd_bench_")
    system, tbuild = build_workload(tmp)
    t0 = time.time()
    sysm = System.load(tmp)
    thost = time.time() - t0
    N = sysm.natom
    total_pops = (args.steps + args.warmup) * args.pops
    eng = GpuEngine([sysm], block=args.block, max_pops=total_pops + 100)
    eng.prepare()
    bulk_ms = eng.last_kernel_ms()
    for _ in range(args.warmup):
        eng.run(args.pops)

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    s0 = eng.stats(0)
    sync()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(args.steps):
        eng.run(args.pops)
        kernel_ms += eng.last_kernel_ms()
    sync()
    elapsed = time.perf_counter() - t0
    s1 = eng.stats(0)
    assert s1["error"] == 0
    events = s1["frame"] - s0["frame"]
    pops = s1["pops"] - s0["pops"]
    recal = s1["recal"] - s0["recal"]
    evals = s1["pair_evals"] - s0["pair_evals"]
    cands = s1["candidates"] - s0["candidates"]
    renewed = 2 * events + recal - (s1["per_type"] - s0["per_type"])[[4, 10]].sum()   # one-body events renew 1 atom
    n_nb = cands / max(renewed, 1)
    logn = math.ceil(math.log2(N))
    bytes_total = pops * N * 88 + renewed * ((n_nb + 1) * 56 + 40 + 2 * logn * 12)
    achieved = bytes_total / (kernel_ms / 1e3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "plain_event_loop_cfg2_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    line = {
        "metric": "aggregate events/s + pair-time evals/s; batched FP32-screen vs plain FP64 GPU",
        "value": events / elapsed,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded builder: random sequences, random placement)",
        "config": {"workload": "64 x 40-residue random peptides, %d atoms, 5 mM, %.1f A box, cutoff 7.25 A, "
                               "300 K Andersen, plain FP64 path" % (N, system.box[0]),
                   "natom": N, "pops_per_step": args.pops, "block": args.block},
        "pair_evals_per_s": evals / elapsed,
        "pops_per_s": pops / elapsed,
        "events": events,
        "n_nb": n_nb,
        "bulk_predict_ms": bulk_ms,
        "setup_s": {"build": tbuild, "host": thost},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_total / args.steps,
                     "kernel_ms_per_launch": kernel_ms / args.steps},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(tmp, args.cpu_pops, N)
    eng.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

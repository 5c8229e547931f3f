"""Benchmark: aggregate events/s and pair-time evaluations/s of the sDMD hot
path on one MI355X, batched FP32-screen engine against the plain FP64 engine.

Workload (N=1): a seeded batch of 48 independent peptide systems of uneven
size (chains of 10-80 residues, 1-32 chains per system, uniform-random
sequences, about 455k atoms in all, 2.6k-23k per system), each in its own
cubic periodic box at 5 mM per chain, Ding all-atom force field, cut-off
7.25 A, 300 K (T* 0.5962) with the Andersen thermostat.  All 48 systems run
in one persistent launch (one workgroup per system).

A step is one launch of the device event loop on every system for a fixed
window of device time (--step-ms); the launch ends with a drain that leaves
every system at one consistent time.  `value` is the fast engine's executed
events (stale pops excluded) summed over the systems, divided by the wall
time of the timed steps.  The plain FP64 engine (bit-exact with the
reference) runs the same systems for the same windows; `plain` and
`vs_plain` report it.

`roofline` prices the dominant kernel (the fast event loop) with the
algorithmic bytes of the hot-path scope: every re-predicted atom reads its
candidates' state, (n_nb + 1) x 56 B (position, velocity, type and
exclusion word of each of the n_nb candidates scanned and of the atom
itself); every executed event writes a 40-B event record and walks a tree
path of 2 ceil(log2 N) x 12 B.  The counts behind it (events, predictions,
candidates, confirms, one-body events, stale pops, launches) are printed, so
the figure can be recomputed from the line.  `cpu_baseline` runs the
reference built from its own sources (oracle/_ref/sdmd_ref) as P = min(#systems,
host cores) concurrent single-threaded processes pinned to one core each,
one system of the batch per process (spread over the size range), each for a
bounded number of pops, and reports the sum of their event rates.
"""
import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fast_event_loop_batch48_traffic.json")


def build_batch(tmp, nsys=48, seed=4, cutoff=7.25, temperature=0.5962):
    from sdmd_amd import builder

    t0 = time.time()
    workers = max(1, min(16, os.cpu_count() or 1))
    def progress(done, total):
        if done % 50 == 0 or done == total:
            print("building chains: %d / %d" % (done, total), file=sys.stderr, flush=True)

    systems = builder.batch_peptides(nsys=nsys, seed=seed, workers=workers, progress=progress)
    dirs = []
    for k, s in enumerate(systems):
        d = os.path.join(tmp, "s%02d" % k)
        builder.write_input_dir(s, d, cutoff=cutoff, temperature=temperature)
        dirs.append(d)
    return dirs, time.time() - t0


def load_systems(dirs):
    from sdmd_amd.engine import System

    t0 = time.time()
    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        systems = list(ex.map(System.load, dirs))
    return systems, time.time() - t0


def host_cpus():
    """(cores this process may use, CPU model)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_baseline(dirs, natoms, seconds, procs):
    """The reference (own sources, -O2) as `procs` concurrent processes, one
    system each, pinned to one core each, each for about `seconds` of its
    event loop; the sum of their loop rates."""
    ref = os.path.join(ROOT, "oracle", "_ref", "sdmd_ref")
    if not os.path.exists(ref):
        raise SystemExit("bench.py: oracle/_ref/sdmd_ref (the reference built from its own sources by "
                         "`make -C oracle ref`) is missing; the CPU baseline needs it")
    cores, model = host_cpus()
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(cores))
    P = max(1, min(procs, len(dirs), len(allowed)))
    order = sorted(range(len(dirs)), key=lambda k: natoms[k])
    pick = [order[int(round(q * (len(order) - 1) / max(P - 1, 1)))] for q in range(P)]
    have_taskset = subprocess.run(["which", "taskset"], capture_output=True).returncode == 0
    runs = []
    barrier = tempfile.mkdtemp(prefix="sdmd_cpu_barrier_")
    for q, k in enumerate(pick):
        out = tempfile.mkdtemp(prefix="sdmd_cpu_")
        # every loop waits at its start for the others (SDMD_BARRIER) and runs
        # `seconds` of its event loop (SDMD_MAX_SECONDS); the pop cap is a
        # safety bound only (a pop costs at least ~1.5 us + 2.4 ns per atom)
        pops = int(seconds / 1e-6)
        env = dict(os.environ, SDMD_LIGHT="1", SDMD_DUMP=out, SDMD_MAX_POPS=str(pops),
                   SDMD_MAX_SECONDS=repr(float(seconds)), SDMD_BARRIER="%s:%d" % (barrier, P))
        cmd = [ref, "-f", "coordinate.gro", "-i", dirs[k], "-o", out]
        if have_taskset:
            cmd = ["taskset", "-c", str(allowed[q])] + cmd
        fh = open(os.path.join(out, "stdout.txt"), "w")
        runs.append((k, out, subprocess.Popen(cmd, env=env, stdout=fh, stderr=subprocess.STDOUT), fh))
    rates, events, t_lo, t_hi, npops = [], 0, [], [], 0
    # a process that fails would leave the others waiting at the start barrier
    while any(p.poll() is None for _, _, p, _ in runs):
        bad = [(k, out) for k, out, p, _ in runs if p.poll() not in (None, 0)]
        if bad:
            for _, _, p, _ in runs:
                if p.poll() is None:
                    p.kill()
            raise SystemExit("bench.py: the reference failed on %s (see %s)" % (dirs[bad[0][0]], bad[0][1]))
        time.sleep(0.05)
    for k, out, p, fh in runs:
        if p.wait() != 0:
            raise SystemExit("bench.py: the reference failed on %s (see %s)" % (dirs[k], out))
        fh.close()
        tok = open(os.path.join(out, "timing.txt")).read().split()
        v = dict(zip(tok[0::2], tok[1::2]))
        rates.append(int(v["events"]) / float(v["seconds"]))
        events += int(v["events"])
        npops += int(v["pops"])
        t_lo.append(float(v["start"]))
        t_hi.append(float(v["end"]))
    overlap = max(0.0, min(t_hi) - max(t_lo))
    for d in [barrier] + [out for _, out, _, _ in runs]:
        shutil.rmtree(d, ignore_errors=True)
    return {"value": sum(rates), "unit": "events/s", "cores": P, "kind": "reference",
            "host_cores": cores, "cpu_model": model,
            "sample": "reference sDMD (own sources, -O2), %d concurrent single-threaded processes pinned to one "
                      "core each, one batch system each (%d-%d atoms), started together, %.0f s of each event loop "
                      "(%d pops, %d events in all); value = sum of the %d loop rates (all loops ran together for "
                      "%.1f s)"
                      % (P, min(natoms[k] for k in pick), max(natoms[k] for k in pick), seconds, npops, events, P,
                         overlap),
            "per_process_events_per_s": [round(r, 1) for r in rates]}


def totals(eng, n, fast):
    t = {"frame": 0, "pops": 0, "recal": 0, "pair_evals": 0, "candidates": 0, "onebody": 0,
         "screens": 0, "confirms": 0, "batches": 0, "causal": 0, "errors": set(), "frame_per_sys": [0] * n}
    for k in range(n):
        st = eng.stats(k)
        t["frame"] += st["frame"]
        t["frame_per_sys"][k] = st["frame"]
        t["pops"] += st["pops"]
        t["recal"] += st["recal"]
        t["pair_evals"] += st["pair_evals"]
        t["candidates"] += st["candidates"]
        t["onebody"] += int(st["per_type"][4] + st["per_type"][5] + st["per_type"][10])   # thermostat, wall, cell crossing
        t["errors"].add(st["error"])
        if fast:
            fs = eng.fast_stats(k)
            for f in ("screens", "confirms", "batches", "causal"):
                t[f] += fs[f]
    return t


def diff(a, b):
    out = {k: (a[k] - b[k]) for k in a if k not in ("errors", "frame_per_sys")}
    out["frame_per_sys"] = [x - y for x, y in zip(a["frame_per_sys"], b["frame_per_sys"])]
    return out


def timed(eng, nsys, fast, steps, warmup, step_ms, graph=0):
    """warmup + steps windows of step_ms device time; with graph > 0 a window
    is `graph` bounded launches replayed from a captured HIP graph."""
    def step():
        if graph > 0:
            eng.run_graph(graph, step_ms / graph)
        else:
            eng.run_for(1 << 40, step_ms)
    for _ in range(warmup):
        step()
    t0s = totals(eng, nsys, fast)
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(steps):
        step()
        kernel_ms += eng.last_kernel_ms()
    elapsed = time.perf_counter() - t0
    t1s = totals(eng, nsys, fast)
    d = diff(t1s, t0s)
    d["errors"] = t1s["errors"]
    return d, elapsed, kernel_ms


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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--step-ms", type=float, default=1000.0, help="device-time window of one step")
    ap.add_argument("--block", type=int, default=512, help="fast engine threads per system (before autotuning)")
    ap.add_argument("--autotune-ms", type=float, default=150.0,
                    help="device time per candidate shape for the autotuner (0: keep --block)")
    ap.add_argument("--graph", type=int, default=4, help="bounded launches per step replayed from a HIP graph "
                                                         "(0: one launch per step)")
    ap.add_argument("--plain-block", type=int, default=512, help="plain engine threads per system (512 measured fastest of 256/512/1024)")
    ap.add_argument("--systems", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="event-loop time per reference process")
    ap.add_argument("--cpu-procs", type=int, default=48, help="most concurrent reference processes "
                                                              "(capped by the batch and the usable cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--workdir", default=None, help="build the inputs here, or reuse them when present")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the batch shards across GPUs by whole systems (one process "
              "per GPU), which this benchmark does not launch; N=%d is not run" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    from sdmd_amd.engine import GpuEngine

    tmp = args.workdir or tempfile.mkdtemp(prefix="sdmd_bench_")
    dirs = [os.path.join(tmp, "s%02d" % k) for k in range(args.systems)]
    if all(os.path.exists(os.path.join(d, "coordinate.gro")) for d in dirs):
        tbuild = 0.0                                   # inputs already built in --workdir
    else:
        os.makedirs(tmp, exist_ok=True)
        dirs, tbuild = build_batch(tmp, nsys=args.systems)
    systems, thost = load_systems(dirs)
    nsys = len(systems)
    natoms = [s.natom for s in systems]

    # fast engine: FP32 screen + FP64 confirm, concurrent independent events
    eng = GpuEngine(systems, block=args.block, mode="fast")
    eng.prepare()
    bulk_ms = eng.last_kernel_ms()
    depth = None
    if args.autotune_ms > 0:
        args.block, depth = eng.autotune(args.autotune_ms)
    f, elapsed, kernel_ms = timed(eng, nsys, True, args.steps, args.warmup, args.step_ms, graph=args.graph)
    assert f["errors"] == {0}, f["errors"]
    assert f["causal"] == 0, f["causal"]             # no read of an atom later than the reader
    helpers = eng.fast_stats(0)["helpers"]
    eng.close()
    events = f["frame"]
    predictions = 2 * events + f["recal"] - f["onebody"]
    n_nb = f["candidates"] / max(predictions, 1)
    launches = args.steps * max(args.graph, 1)
    # algorithmic bytes: (n_nb + 1) x 56 B per prediction, 40 B + 2 ceil(log2 N) x 12 B per executed event
    tree_bytes = sum(f["frame_per_sys"][k] * 2 * math.ceil(math.log2(natoms[k])) * 12 for k in range(nsys))
    bytes_launch = (f["candidates"] * 56 + predictions * 56 + events * 40 + tree_bytes) / launches
    kernel_ms_launch = kernel_ms / launches
    achieved = bytes_launch / (kernel_ms_launch / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            per_event = json.load(fh).get("hbm_bytes_per_event")
        if per_event:
            traffic = per_event * events / launches

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
        "dtype": "f32/f64",
        "data": "synthetic (seeded builder: random sequences, random placement)",
        "config": {"workload": "batch of %d peptide systems, 10-80 residues x 1-32 chains each, %d atoms "
                               "(%d-%d per system), 5 mM per chain, cut-off 7.25 A, 300 K Andersen; "
                               "fast engine (FP32 screen + FP64 confirm)" % (nsys, sum(natoms), min(natoms),
                                                                             max(natoms)),
                   "systems": nsys, "natom_total": sum(natoms), "block": args.block, "staging_depth": depth,
                   "autotuned": args.autotune_ms > 0, "graph_launches_per_step": args.graph,
                   "step_ms": args.step_ms},
        "pair_evals_per_s": (f["screens"] + f["pair_evals"]) / elapsed,
        "fp32_screens_per_s": f["screens"] / elapsed,
        "fp64_pair_evals_per_s": f["pair_evals"] / elapsed,
        "pops_per_s": f["pops"] / elapsed,
        "events_per_round": f["pops"] / max(f["batches"], 1),
        "n_nb": n_nb,
        "causal_reads": f["causal"],
        "bulk_predict_ms": bulk_ms,
        "setup_s": {"build": tbuild, "host": thost},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_launch, "kernel_ms_per_launch": kernel_ms_launch,
                     "formula": "((candidates + predictions) * 56 + events * 40 + sum_s events_s * 24 * "
                                "ceil(log2 N_s)) / launches",
                     "counts": {"events": events, "predictions": predictions, "candidates": f["candidates"],
                                "confirms": f["confirms"], "onebody": f["onebody"], "stale": f["recal"],
                                "launches": launches, "tree_bytes": tree_bytes}},
        "helpers_per_system": helpers,
    }
    if not args.no_plain:
        pops_cap = int(40000 * (args.steps + args.warmup) * args.step_ms / 1000.0) + 1000
        eng = GpuEngine(systems, block=args.plain_block, mode="plain", max_pops=pops_cap)
        eng.prepare()
        p, pel, pkms = timed(eng, nsys, False, args.steps, args.warmup, args.step_ms)
        assert p["errors"] == {0}, p["errors"]
        eng.close()
        line["plain"] = {"value": p["frame"] / pel, "unit": "events/s", "pair_evals_per_s": p["pair_evals"] / pel,
                         "pops_per_s": p["pops"] / pel, "block": args.plain_block}
        line["vs_plain"] = line["value"] / line["plain"]["value"]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(dirs, natoms, args.cpu_seconds, args.cpu_procs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

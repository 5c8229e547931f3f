"""One large system on both engines: the plain FP64 engine (bit-exact with
the reference) and the fast engine on the same input.

    python -m sdmd_amd.single [--chains 64] [--length 40] [--seconds 2]

Default workload: 64 chains of 40 residues with uniform-random sequences
(seeded), random non-overlapping placement in a periodic box at 5 mM,
cut-off 7.25 A (0.9 x cut-off covers the 6.5 A outermost well), 300 K
Andersen.  Each engine runs `--seconds` of device time after its bulk
prediction; prints one JSON object (events/s, pops/s, pair-time
evaluations/s per engine: for the fast engine FP32 screens plus FP64
solves, as bench.py counts them).
"""
import argparse
import json
import sys
import tempfile
import time

from . import builder
from .engine import GpuEngine, System


def measure(eng, seconds, fast):
    eng.prepare()
    eng.run_for(1 << 40, 200.0)                        # warm-up
    s0 = eng.stats(0)
    f0 = eng.fast_stats(0) if fast else None
    t0 = time.perf_counter()
    eng.run_for(1 << 40, seconds * 1e3)
    wall = time.perf_counter() - t0
    s1 = eng.stats(0)
    out = {"events_per_s": (s1["frame"] - s0["frame"]) / wall, "pops_per_s": (s1["pops"] - s0["pops"]) / wall,
           "pair_evals_per_s": (s1["pair_evals"] - s0["pair_evals"]) / wall, "seconds": wall,
           "error": s1["error"]}
    if fast:
        f1 = eng.fast_stats(0)
        out["fp32_screens_per_s"] = (f1["screens"] - f0["screens"]) / wall
        out["fp64_pair_evals_per_s"] = out["pair_evals_per_s"]
        out["pair_evals_per_s"] += out["fp32_screens_per_s"]
        out["events_per_round"] = (s1["pops"] - s0["pops"]) / max(f1["batches"] - f0["batches"], 1)
        out["helpers"] = f1["helpers"]
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--length", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--seconds", type=float, default=2.0)
    ap.add_argument("--plain-block", type=int, default=1024)
    ap.add_argument("--fast-block", type=int, default=512)
    args = ap.parse_args()
    t0 = time.perf_counter()
    d = tempfile.mkdtemp(prefix="sdmd_single_")
    sysm = builder.random_peptides(args.chains, args.length, seed=args.seed, mM=5.0)
    builder.write_input_dir(sysm, d, cutoff=7.25)
    s = System.load(d)
    print("single: %d atoms loaded in %.1f s" % (s.natom, time.perf_counter() - t0), file=sys.stderr, flush=True)
    rate = 200000.0 * (args.seconds + 0.5)          # pops cap for the plain thermostat draws
    plain = measure(GpuEngine([s], block=args.plain_block, mode="plain", max_pops=int(rate) + 1000),
                    args.seconds, False)
    print("single: plain done", file=sys.stderr, flush=True)
    fast = measure(GpuEngine([s], block=args.fast_block, mode="fast"), args.seconds, True)
    out = {"workload": "%d x %d-residue peptides, %d atoms, box %.1f A, 5 mM, cut-off 7.25 A, 300 K Andersen"
                       % (args.chains, args.length, s.natom, sysm.box[0]),
           "plain": dict(plain, block=args.plain_block), "fast": dict(fast, block=args.fast_block),
           "fast_over_plain": fast["events_per_s"] / plain["events_per_s"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

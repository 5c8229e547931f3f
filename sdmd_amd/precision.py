"""Precision-form report: the FP32-screen / FP64-confirm engine against FP64.

    python -m sdmd_amd.precision [--events 1000000] [--checkpoints 20] [--only fast|plain]

System: the ALA10 extended chain (59 atoms, 50 A box, cut-off 9 A, T* 0.5962)
in NVE (no thermostat), so total energy is the conserved quantity.

* The fast engine runs with validation on: every prediction's collision
  winner is also found by a full FP64 scan of the same neighbourhood, which
  gives the partner-agreement rate and, for disagreements, the time error;
  the largest relative error of the FP32 screen time of a confirmed winner
  is reported as well.
* Energy drift: total energy (CalKinetE + CalPotenE, ToolFunctions.c:626-752)
  at evenly spaced checkpoints, for the fast engine and for the plain FP64
  engine, which is bit-exact with the reference, over the same number of
  executed events.  The reference's NVE energy moves in steps when H-bond
  bookkeeping (HBEnergyChange, Events.c:1027-1095) sums shell energies in
  cell-list order; both engines are measured against that envelope.

Prints one JSON object.
"""
import argparse
import ctypes as C
import json
import sys
import tempfile
import time

from . import _lib, builder
from .engine import GpuEngine, System


def energy(arr):
    d = arr.desc()
    ke, pe = C.c_double(), C.c_double()
    _lib.check(_lib.lib().sdmd_host_energy(C.byref(d), C.byref(ke), C.byref(pe)), "sdmd_host_energy")
    return ke.value + pe.value


def nve_system():
    d = tempfile.mkdtemp(prefix="sdmd_prec_")
    builder.write_input_dir(builder.ala10_extended(seed=1), d, temperature=0.5962, cutoff=9.0, thermostat="no")
    return System.load(d)


def run_fast(system, events, checkpoints):
    eng = GpuEngine([system], block=256, mode="fast", validate=True)
    eng.prepare()
    e0 = energy(eng.state(0))
    series = []
    per = max(1, events // checkpoints)
    done = 0
    while done < events:
        target = done + per
        while eng.stats(0)["frame"] < target:
            eng.run(2 * per)
        done = eng.stats(0)["frame"]
        series.append((done, energy(eng.state(0)) - e0))
        print("fast: %d events" % done, file=sys.stderr, flush=True)
    st, fs = eng.stats(0), eng.fast_stats(0)
    eng.close()
    return e0, series, st, fs


def run_plain(system, events, checkpoints):
    per = max(1, events // checkpoints)
    pops_est = int(events * 1.8) + 10 * per
    eng = GpuEngine([system], block=256, mode="plain", max_pops=pops_est, normals_per_pop=0.0)
    eng.prepare()
    e0 = energy(eng.state(0))
    series = []
    done = 0
    while done < events:
        target = done + per
        while eng.stats(0)["frame"] < target:
            eng.run(max(1, int((target - eng.stats(0)["frame"]) * 1.7)))
        done = eng.stats(0)["frame"]
        series.append((done, energy(eng.state(0)) - e0))
        print("plain: %d events" % done, file=sys.stderr, flush=True)
    st = eng.stats(0)
    eng.close()
    return e0, series, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--events", type=int, default=1000000)
    ap.add_argument("--checkpoints", type=int, default=20)
    ap.add_argument("--only", choices=("fast", "plain"), default=None,
                    help="run one engine (long runs are split across invocations)")
    args = ap.parse_args()
    s = nve_system()
    out = {"system": "ALA10 extended chain, %d atoms, NVE (thermostat off), T* 0.5962, cut-off 9 A" % s.natom,
           "events": {}}
    if args.only in (None, "fast"):
        t0 = time.perf_counter()
        e0, fser, fst, ffs = run_fast(s, args.events, args.checkpoints)
        out["events"]["fast"] = fst["frame"]
        out.update({
            "fast_seconds": time.perf_counter() - t0,
            "partner_agreement": ffs["agree"] / max(ffs["validated"], 1),
            "validated_predictions": ffs["validated"],
            "max_time_rel_err_vs_fp64": ffs["max_time_rel_err"],
            "max_fp32_screen_rel_err": ffs["max_screen_rel_err"],
            "fp32_screens": ffs["screens"], "fp64_confirms": ffs["confirms"],
            "causal_reads": ffs["causal"],
            "energy0": e0,
            "fast_energy_drift": fser,
            "fast_max_abs_drift": max(abs(x) for _, x in fser),
        })
    if args.only in (None, "plain"):
        t0 = time.perf_counter()
        p0, pser, pst = run_plain(s, args.events, args.checkpoints)
        out["events"]["plain"] = pst["frame"]
        out.update({"plain_seconds": time.perf_counter() - t0, "energy0": p0, "plain_energy_drift": pser,
                    "plain_max_abs_drift": max(abs(x) for _, x in pser)})
    print(json.dumps(out))


if __name__ == "__main__":
    main()

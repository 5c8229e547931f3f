"""Replica exchange on one GPU.

The reference runs REMD as N `sDMD -REMD` processes and one `sServer`
exchanging (E, T) over localhost TCP (REMDExec.c:111-144, REMD.c:135-192,
Server.c:92-211).  Here the replicas are the systems of one engine context
(batching form, normally the fast engine): every `interval` executed events
per replica each replica's potential energy (CalPotenE) is computed, the
server's exchange pass (Server.c rearrange) runs on it and the new
temperatures are set per system (REMD.c:97-101).  Configurations stay with
their system; the temperatures move.

    Remd.run(cycles)   all on the device (sdmd_remd_run): event windows,
                       energies, exchange pass and temperature updates
                       enqueued back to back, one host synchronisation
    Remd.cycle()       one cycle with the pass on the host
                       (sdmd_remd_rearrange): energies copied back, new
                       temperatures sent per replica

    python -m sdmd_amd.remd [--replicas 64] [--cycles 20] [--interval 10000]

prints one JSON object (events/s over the whole loop including exchanges,
acceptance per neighbour pair, exchange overhead).
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import _lib, builder
from ._lib import RandState, RemdTemp, check
from .engine import GpuEngine, System

BOLTZMANN = 0.0019872041        # REMDExec.c:19, K -> kcal/mol


def ladder(n, t_lo=280.0, t_hi=450.0):
    """Geometric temperature ladder in reduced units, T_k = t_lo (t_hi/t_lo)^(k/(n-1)) K."""
    if n == 1:
        return [BOLTZMANN * t_lo]
    return [BOLTZMANN * t_lo * (t_hi / t_lo) ** (k / (n - 1)) for k in range(n)]


def rearrange(epot, temps, rng, counts=None):
    """One exchange pass (Server.c:145-192) over ctypes RemdTemp[n] in place."""
    n = len(temps)
    e = np.ascontiguousarray(epot, dtype=np.float64)
    if len(e) != n:
        raise ValueError("one energy per replica")
    cp = counts.ctypes.data if counts is not None else None
    check(_lib.lib().sdmd_remd_rearrange(e.ctypes.data, temps, n, C.byref(rng), cp), "sdmd_remd_rearrange")


def replica_systems(system, temps, workdir, cutoff=7.25, seed=1575263273, wall=True):
    """Load one System per temperature from the same configuration (each
    with its own parameter.txt temperature, so its velocities are drawn at
    it, and its own RNG seed)."""
    out = []
    for k, T in enumerate(temps):
        d = os.path.join(workdir, "r%02d" % k)
        kw = dict(temperature=T, cutoff=cutoff)
        if wall:
            kw.update(wall="smooth", wall_type="cylinder")
        builder.write_input_dir(system, d, **kw)
        out.append(System.load(d, seed=seed + k))
    return out


class Remd:
    """REMD over the systems of one engine context."""

    def __init__(self, systems, temps, seed=1, mode="fast", block=512, interval=10000, **engine_kw):
        if len(systems) != len(temps):
            raise ValueError("one temperature per replica")
        self.n = len(systems)
        self.eng = GpuEngine(systems, block=block, mode=mode, **engine_kw)
        self.interval = int(interval)
        self.T = (RemdTemp * self.n)()
        for k, t in enumerate(temps):
            self.T[k].num = k + 1
            self.T[k].T = t
        self.rng = RandState()
        _lib.lib().sdmd_rand_seed(C.byref(self.rng), seed)          # srand(seed), Server.c:58
        self.counts = np.zeros(2 * self.n, np.int64)
        self.exchange_s = 0.0
        self.cycles = 0

    def prepare(self):
        for k in range(self.n):
            self.eng.set_temperature(k, self.T[k].T)
        self.eng.prepare()

    def temperatures(self):
        return np.array([t.T for t in self.T]), np.array([t.num for t in self.T])

    def cycle(self):
        """`interval` events per replica, then one exchange pass; returns the
        potential energies the pass used."""
        self.eng.run_events(self.interval)
        t0 = time.perf_counter()
        _, pe = self.eng.energy()
        before = [t.T for t in self.T]
        rearrange(pe, self.T, self.rng, self.counts)
        for k in range(self.n):
            if self.T[k].T != before[k]:
                self.eng.set_temperature(k, self.T[k].T)
        self.exchange_s += time.perf_counter() - t0
        self.cycles += 1
        return pe

    def run(self, cycles, history=False):
        """`cycles` exchange cycles on the device; returns the energies each
        pass used (cycles x n) and, with history, the {num, T} table after
        each pass."""
        n = self.n
        ep = np.zeros((cycles, n))
        hist = (RemdTemp * (cycles * n))() if history else None
        t0 = time.perf_counter()
        check(_lib.lib().sdmd_remd_run(self.eng._h, self.T, C.byref(self.rng), self.counts.ctypes.data, int(cycles),
                                       self.interval, hist, ep.ctypes.data), "sdmd_remd_run")
        self.run_s = getattr(self, "run_s", 0.0) + time.perf_counter() - t0
        self.cycles += cycles
        if history:
            tab = [[(hist[c * n + k].num, hist[c * n + k].T) for k in range(n)] for c in range(cycles)]
            return ep, tab
        return ep

    def close(self):
        self.eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--chains", type=int, default=8)
    ap.add_argument("--length", type=int, default=20)
    ap.add_argument("--radius", type=float, default=40.0)
    ap.add_argument("--box-x", type=float, default=160.0)
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--interval", type=int, default=10000)
    ap.add_argument("--block", type=int, default=512)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--host-exchange", action="store_true", help="exchange pass on the host, one cycle per call")
    args = ap.parse_args()
    t0 = time.perf_counter()
    sysm = builder.cylinder_peptides(args.chains, args.length, seed=args.seed, radius=args.radius,
                                     box_x=args.box_x)
    temps = ladder(args.replicas)
    work = tempfile.mkdtemp(prefix="sdmd_remd_")
    systems = replica_systems(sysm, temps, work)
    setup_s = time.perf_counter() - t0
    print("remd: %d replicas of %d atoms loaded in %.1f s" % (len(systems), systems[0].natom, setup_s),
          file=sys.stderr, flush=True)
    r = Remd(systems, temps, seed=args.seed, interval=args.interval, block=args.block)
    r.prepare()
    f0 = sum(r.eng.stats(k)["frame"] for k in range(r.n))
    t1 = time.perf_counter()
    if args.host_exchange:
        for c in range(args.cycles):
            r.cycle()
            if (c + 1) % 5 == 0:
                print("remd: cycle %d" % (c + 1), file=sys.stderr, flush=True)
    else:
        r.run(args.cycles)
    wall = time.perf_counter() - t1
    st = [r.eng.stats(k) for k in range(r.n)]
    events = sum(s["frame"] for s in st) - f0
    temps_now, nums = r.temperatures()
    acc = r.counts[1::2] / np.maximum(r.counts[0::2], 1)
    out = {
        "workload": "%d REMD replicas, %d x %d-residue peptides (%d atoms) in a smooth cylinder wall r=%g A, "
                    "box %g x %g x %g A, 280-450 K geometric, exchange every %d events per replica"
                    % (r.n, args.chains, args.length, systems[0].natom, args.radius, args.box_x,
                       2 * args.radius, 2 * args.radius, args.interval),
        "events": int(events), "seconds": wall, "events_per_s": events / wall,
        "exchange": "host pass per cycle" if args.host_exchange else "device (one host synchronisation)",
        "exchange_seconds": r.exchange_s if args.host_exchange else None, "cycles": r.cycles,
        "acceptance": [float(a) for a in acc[: r.n - 1]],
        "attempts": [int(a) for a in r.counts[0:2 * (r.n - 1):2]],
        "wall_events": int(sum(s["per_type"][5] for s in st)),
        "errors": int(sum(s["error"] != 0 for s in st)),
        "causal_reads": int(sum(r.eng.fast_stats(k)["causal"] for k in range(r.n))),
        "setup_s": setup_s,
        "temperature_index_of_replica": [int(x) for x in nums],
    }
    r.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Python mirror of the reference's run interface over the native library.

    System.load(datadir, gro, seed)   InputData (Initialization.c:47-357, "new")
    System.prepare()                  Prepare up to the bulk Predict
                                      (SGThread.c:21-45)
    GpuEngine(systems).prepare()      bulk Predict(1..N) + CreateCBT on device
                                      (SGThread.c:47-62)
    GpuEngine.run(pops)               SingleThreadRun (SGThread.c:67-156) on
                                      device, `pops` iterations per system

Errors raise SdmdError with the library's message (the reference exits).
"""
import ctypes as C
import os
import tempfile

import numpy as np

from . import _lib
from ._ctypes_types import DescArrays, SystemDesc
from ._lib import FastStats, GpuEvent, GpuOpts, GpuStats, PopRec, check

MODES = {"plain": 0, "fast": 1}

POP_DTYPE = np.dtype([
    ("target", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"), ("exec", "<i4"),
    ("nrenew", "<i4"), ("dt", "<f8"), ("pot", "<f8"), ("rand_calls", "<i8"),
    ("r", [("atom", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"),
           ("counter", "<i8"), ("pcounter", "<i8"), ("time", "<f8"), ("pot", "<f8")], (2,)),
])
assert POP_DTYPE.itemsize == C.sizeof(PopRec)


class System:
    """One system prepared by the C host; `arrays` holds a private copy of
    its flattened state (numpy), `desc()` a C view of it."""

    def __init__(self, handle, arrays):
        self._h = handle
        self.arrays = arrays

    @classmethod
    def load(cls, datadir, gro="coordinate.gro", seed=1575263273, prepare=True, max_min_steps=20000):
        L = _lib.lib()
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.sdmd_host_load(os.fsencode(datadir), os.fsencode(gro), seed, C.byref(h), err, 1024)
        if rc != 0:
            raise _lib.SdmdError("sdmd_host_load: %s (%s)" % (err.value.decode(), L.sdmd_strerror(rc).decode()))
        if prepare:
            rc = L.sdmd_host_prepare(h, max_min_steps)
            if rc != 0:
                L.sdmd_host_free(h)
                check(rc, "sdmd_host_prepare")
        arrays = DescArrays(L.sdmd_host_desc(h).contents)
        return cls(h, arrays)

    @classmethod
    def from_text(cls, gro_text, parameter_text, seed=1575263273, wall_text=None, **kw):
        """Load from in-memory .gro and parameter.txt contents (and Wall.txt
        for a run with a smooth wall)."""
        from .builder import LIB_DIR, write_library
        d = tempfile.mkdtemp(prefix="sdmd_in_")
        with open(os.path.join(d, "coordinate.gro"), "w") as fh:
            fh.write(gro_text)
        with open(os.path.join(d, "parameter.txt"), "w") as fh:
            fh.write(parameter_text)
        if wall_text:
            lib = write_library(d)
            with open(os.path.join(lib, "Wall.txt"), "w") as fh:
                fh.write(wall_text)
        else:
            os.symlink(LIB_DIR, os.path.join(d, "Library_Ding"))
        return cls.load(d, "coordinate.gro", seed=seed, **kw)

    @property
    def natom(self):
        return self.arrays.scalars["natom"]

    def desc(self):
        return self.arrays.desc()

    def normals(self, n):
        """The first n thermostat draws of the event loop (the reference's
        stream from the RNG position reached by setup) and the rand() call
        count after each.  Stateless: every call returns the same stream."""
        z = np.zeros(n, np.float64)
        calls = np.zeros(n, np.int64)
        sc = self.arrays.scalars
        check(_lib.lib().sdmd_normal_stream(sc["seed"], sc["rand_calls"], z.ctypes.data, n, calls.ctypes.data),
              "sdmd_normal_stream")
        return z, calls

    def close(self):
        if self._h:
            _lib.lib().sdmd_host_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuEngine:
    """Device-resident systems on one HIP stream (one workgroup per system).

    mode "plain": the reference's swept FP64 state, bit-exact event sequence.
    mode "fast":  local-time state, FP32 screen + FP64 confirm, concurrent
                  independent events (one per wavefront); `validate` re-solves
                  every collision candidate in FP64 and counts agreement;
                  `helpers` workgroups per system take re-predictions on
                  other CUs (0: as many as fit, -1: none).
    `window` (plain): neighbour candidates per pass of a prediction's
    windowed scan (0: 1728); a smaller window takes more passes and gives
    the same result."""

    PHASES = ["pop", "sweep", "event", "stage", "gather", "bondsel", "eval", "select", "hb", "tail", "tree"]
    FAST_PHASES = ["select", "exec", "stage", "screen", "confirm", "other", "tree", "event", "bond", "scan", "collect"]

    def __init__(self, systems, block=256, trace=False, trace_capacity=0, hash_every=0,
                 normals_per_pop=5.0, max_pops=0, profile=False, mode="plain", validate=False,
                 max_batch=0, seed=0, helpers=0, window=0):
        self.systems = list(systems)
        self.mode = mode
        L = _lib.lib()
        self._descs = [s.desc() for s in self.systems]
        arr = (C.POINTER(SystemDesc) * len(self._descs))(*[C.pointer(d) for d in self._descs])
        opts = GpuOpts(MODES[mode], block, 1 if trace else 0, 1 if profile else 0, trace_capacity, hash_every,
                       1 if validate else 0, max_batch, seed, helpers, window)
        h = C.c_void_p()
        check(L.sdmd_gpu_open(arr, len(self._descs), C.byref(opts), C.byref(h)), "sdmd_gpu_open")
        self._h = h
        self.block = block
        self.trace_capacity = trace_capacity
        self.hash_every = hash_every
        self._calls = [None] * len(self.systems)
        self._rand0 = [s.arrays.scalars["rand_calls"] for s in self.systems]
        if max_pops and mode == "plain":
            self.feed_normals(int(normals_per_pop * max_pops) + 4 * max(s.natom for s in self.systems))

    def feed_normals(self, n):
        """Give every system n fresh thermostat draws (replaces unread ones)."""
        L = _lib.lib()
        for k, s in enumerate(self.systems):
            z, calls = s.normals(n)
            self._calls[k] = calls
            check(L.sdmd_gpu_set_normals(self._h, k, z.ctypes.data, n), "sdmd_gpu_set_normals")

    def prepare(self):
        check(_lib.lib().sdmd_gpu_prepare(self._h), "sdmd_gpu_prepare")

    def run(self, pops):
        check(_lib.lib().sdmd_gpu_run(self._h, int(pops)), "sdmd_gpu_run")

    def run_for(self, max_pops, max_ms):
        """Run until every system did max_pops pops or max_ms of device time passed."""
        check(_lib.lib().sdmd_gpu_run_for(self._h, int(max_pops), float(max_ms)), "sdmd_gpu_run_for")

    def run_until(self, t_end, max_pops=1 << 40):
        """Run until the clock passes t_end (at most max_pops pops per system)."""
        check(_lib.lib().sdmd_gpu_run_until(self._h, float(t_end), int(max_pops)), "sdmd_gpu_run_until")

    def run_to_output(self, t_out, t_end=1e300, max_pops=1 << 40):
        """Plain mode: pop while the clock is at most t_end and stop after
        the first executed event that leaves it at or past t_out (where
        SingleThreadRun writes its output files, SGThread.c:130-139)."""
        check(_lib.lib().sdmd_gpu_run_to_output(self._h, float(t_out), float(t_end), int(max_pops)),
              "sdmd_gpu_run_to_output")

    def set_tree(self, sys, node, leaf):
        """Plain mode: replace a system's event tree (a resumed run's)."""
        node = np.ascontiguousarray(node, dtype=np.int32)
        leaf = np.ascontiguousarray(leaf, dtype=np.int32)
        check(_lib.lib().sdmd_gpu_set_tree(self._h, sys, node.ctypes.data, leaf.ctypes.data), "sdmd_gpu_set_tree")

    def rand_position(self, sys=0):
        """Plain mode: rand() calls since srand(seed) consumed so far by the
        thermostat draws (the stream position a resumed run starts from)."""
        used = self.stats(sys)["normals_used"]
        calls = self._calls[sys]
        if used == 0 or calls is None:
            return int(self._rand0[sys])
        return int(calls[used - 1])

    def run_events(self, events, max_ms=0.0):
        """Run until every system executed `events` more events (or max_ms of
        device time passed)."""
        check(_lib.lib().sdmd_gpu_run_events(self._h, int(events), float(max_ms)), "sdmd_gpu_run_events")

    def energy(self):
        """(kinetic, potential) arrays over the systems, computed on the device."""
        n = len(self.systems)
        ke, pe = np.zeros(n), np.zeros(n)
        check(_lib.lib().sdmd_gpu_energy(self._h, ke.ctypes.data, pe.ctypes.data), "sdmd_gpu_energy")
        return ke, pe

    def set_temperature(self, sys, temperature):
        """Thermostat temperature of one system (reduced units)."""
        check(_lib.lib().sdmd_gpu_set_temperature(self._h, sys, float(temperature)), "sdmd_gpu_set_temperature")

    def autotune(self, trial_ms=200.0):
        """Fast mode: pick the fastest workgroup shape on the live systems;
        returns (block, staging depth)."""
        b, q = C.c_int32(), C.c_int32()
        check(_lib.lib().sdmd_gpu_autotune(self._h, float(trial_ms), C.byref(b), C.byref(q)), "sdmd_gpu_autotune")
        self.block = b.value
        return b.value, q.value

    def run_graph(self, launches, max_ms):
        """Fast mode: `launches` bounded launches replayed from a captured HIP graph."""
        check(_lib.lib().sdmd_gpu_run_graph(self._h, int(launches), float(max_ms)), "sdmd_gpu_run_graph")

    def fast_stats(self, sys=0):
        st = FastStats()
        check(_lib.lib().sdmd_gpu_fast_stats_get(self._h, sys, C.byref(st)), "sdmd_gpu_fast_stats_get")
        return {f: getattr(st, f) for f, _ in FastStats._fields_}

    def phase_times(self, sys=0):
        names = self.FAST_PHASES if self.mode == "fast" else self.PHASES
        out = np.zeros(len(names))
        check(_lib.lib().sdmd_gpu_phase_times(self._h, sys, out.ctypes.data, len(out)), "sdmd_gpu_phase_times")
        return dict(zip(names, out))

    def last_kernel_ms(self):
        ms = C.c_double()
        check(_lib.lib().sdmd_gpu_last_kernel_ms(self._h, C.byref(ms)), "sdmd_gpu_last_kernel_ms")
        return ms.value

    def predict(self, sys, atoms):
        renew = np.zeros(len(atoms) + 2, np.int32)
        renew[0] = len(atoms)
        renew[1:1 + len(atoms)] = atoms
        check(_lib.lib().sdmd_gpu_predict(self._h, sys, renew.ctypes.data), "sdmd_gpu_predict")

    def create_tree(self, sys):
        check(_lib.lib().sdmd_gpu_create_tree(self._h, sys), "sdmd_gpu_create_tree")

    def update_tree(self, sys, atoms):
        renew = np.zeros(len(atoms) + 2, np.int32)
        renew[0] = len(atoms)
        renew[1:1 + len(atoms)] = atoms
        check(_lib.lib().sdmd_gpu_update_tree(self._h, sys, renew.ctypes.data), "sdmd_gpu_update_tree")

    def pop(self, sys):
        ev = GpuEvent()
        check(_lib.lib().sdmd_gpu_pop(self._h, sys, C.byref(ev)), "sdmd_gpu_pop")
        return {"target": ev.target, "partner": ev.partner, "type": ev.type, "sub": ev.sub,
                "time": ev.time, "potential": ev.potential, "stale": bool(ev.stale)}

    def advance(self, sys, dt):
        check(_lib.lib().sdmd_gpu_advance(self._h, sys, dt), "sdmd_gpu_advance")

    def do_event(self, sys, target, partner):
        check(_lib.lib().sdmd_gpu_do_event(self._h, sys, target, partner), "sdmd_gpu_do_event")

    def cell_move(self, sys, atom):
        check(_lib.lib().sdmd_gpu_cell_move(self._h, sys, atom), "sdmd_gpu_cell_move")

    def state(self, sys=0):
        out = self.systems[sys].arrays.copy()
        d = out.desc()
        check(_lib.lib().sdmd_gpu_download(self._h, sys, C.byref(d)), "sdmd_gpu_download")
        out.scalars["currenttime"] = d.currenttime
        return out

    def tree(self, sys=0):
        N = self.systems[sys].natom
        node = np.zeros(2 * N, np.int32)
        leaf = np.zeros(N + 1, np.int32)
        check(_lib.lib().sdmd_gpu_get_tree(self._h, sys, node.ctypes.data, leaf.ctypes.data), "sdmd_gpu_get_tree")
        return node, leaf

    def stats(self, sys=0):
        st = GpuStats()
        check(_lib.lib().sdmd_gpu_stats_get(self._h, sys, C.byref(st)), "sdmd_gpu_stats_get")
        return {"pops": st.pops, "frame": st.frame, "recal": st.recal,
                "per_type": np.array(st.per_type[:]), "pair_evals": st.pair_evals,
                "candidates": st.candidates, "normals_used": st.normals_used,
                "currenttime": st.currenttime, "error": st.error, "last_exec": bool(st.last_exec)}

    def trace(self, sys=0):
        n = C.c_int64()
        out = np.zeros(max(self.trace_capacity, 1), POP_DTYPE)
        check(_lib.lib().sdmd_gpu_get_trace(self._h, sys, out.ctypes.data, len(out), C.byref(n)), "sdmd_gpu_get_trace")
        tr = out[:n.value].copy()
        # normal-draw position -> rand() call count of the reference
        calls = self._calls[sys]
        if calls is not None:
            pos = tr["rand_calls"]
            rc = np.where(pos > 0, calls[np.maximum(pos - 1, 0)], self._rand0[sys])
            tr["rand_calls"] = rc
        return tr

    def hashes(self, sys=0):
        n = C.c_int64()
        cap = max(self.trace_capacity // max(self.hash_every, 1) + 1, 1)
        out = np.zeros((cap, 2), np.int64)
        check(_lib.lib().sdmd_gpu_get_hashes(self._h, sys, out.ctypes.data, cap, C.byref(n)), "sdmd_gpu_get_hashes")
        return out[:n.value].copy()

    def close(self):
        if self._h:
            _lib.lib().sdmd_gpu_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

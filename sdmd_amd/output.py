"""The reference's output and continue files over the native library
(include/sdmd_io.h).

    write_gro(system, state, path)    SaveGRO, DataSave.c:616-642
    write_cnt(state, path)            SaveConnectionMap, DataSave.c:644-661
    save_dat(system, state, path)     SaveDAT, DataSave.c:512-585
    load_dat(datadir, path)           InputData "continue", Initialization.c:337-356
                                      + ReadBreakPoint 360-509
    save_resume / load_resume         the event tree and random-stream position
                                      a bit-exact resume of this engine needs
    RunOutput                         SingleThreadRun's output cadence
                                      (SGThread.c:130-139): every file marked
                                      for output (trajectory, connection map,
                                      continue file) written after the first
                                      executed event at or past each multiple
                                      of Dump_Rate

`system` is the System a state was set up from (it holds the atom, residue
and chain names); `state` is a DescArrays (GpuEngine.state(), Oracle.state()).
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import DatInfo, Resume, check

# executed-event counters of savedData.dat, in its order, as EEventType
# indices (DMD.h:162-179): collision, HB, bond, HB neighbour, thermostat,
# PBC / cell crossing, wall
EVENT_SUM_TYPES = (0, 2, 1, 3, 4, 10, 5)


def dat_info(system, state, frame=0, per_type=None):
    """The run-level values of savedData.dat: defaults for a run outside
    REMD, H-bond counts from the state, `frame` executed events and the
    per-type executed counts (16 entries, EEventType order)."""
    info = DatInfo()
    d = state.desc()
    _lib.lib().sdmd_dat_info_init(C.byref(info), system._h, C.byref(d))
    info.frame = int(frame)
    if per_type is not None:
        for k, t in enumerate(EVENT_SUM_TYPES):
            info.event_sums[k] = int(per_type[t])
    return info


def write_gro(system, state, path, append=True):
    d = state.desc()
    check(_lib.lib().sdmd_write_gro(system._h, C.byref(d), os.fsencode(path), 1 if append else 0), "sdmd_write_gro")


def write_cnt(state, path, append=True):
    d = state.desc()
    check(_lib.lib().sdmd_write_cnt(C.byref(d), os.fsencode(path), 1 if append else 0), "sdmd_write_cnt")


def save_dat(system, state, path, info=None):
    d = state.desc()
    if info is None:
        info = dat_info(system, state)
    check(_lib.lib().sdmd_save_dat(system._h, C.byref(d), C.byref(info), os.fsencode(path)), "sdmd_save_dat")


def load_dat(datadir, path):
    """A continue run: (System, info) from datadir (parameter.txt,
    Library_Ding) and the savedData.dat at path.  The state's event records
    are as saved; build the tree from them (sdmd_gpu_create_tree) or set a
    saved one (GpuEngine.set_tree)."""
    from ._ctypes_types import DescArrays
    from .engine import System

    L = _lib.lib()
    h = C.c_void_p()
    info = DatInfo()
    err = C.create_string_buffer(1024)
    rc = L.sdmd_host_load_dat(os.fsencode(datadir), os.fsencode(path), C.byref(h), C.byref(info), err, 1024)
    if rc != 0:
        raise _lib.SdmdError("sdmd_host_load_dat: %s (%s)" % (err.value.decode(), L.sdmd_strerror(rc).decode()))
    return System(h, DescArrays(L.sdmd_host_desc(h).contents)), info


def save_resume(path, node, leaf, seed, rand_calls, pops=0, recal=0):
    node = np.ascontiguousarray(node, dtype=np.int32)
    leaf = np.ascontiguousarray(leaf, dtype=np.int32)
    r = Resume(len(leaf) - 1, seed, rand_calls, pops, recal, node.ctypes.data, leaf.ctypes.data)
    check(_lib.lib().sdmd_save_resume(C.byref(r), os.fsencode(path)), "sdmd_save_resume")


def load_resume(path, natom):
    node = np.zeros(2 * natom, np.int32)
    leaf = np.zeros(natom + 1, np.int32)
    r = Resume(natom, 0, 0, 0, 0, node.ctypes.data, leaf.ctypes.data)
    check(_lib.lib().sdmd_load_resume(C.byref(r), os.fsencode(path)), "sdmd_load_resume")
    return {"node": node, "leaf": leaf, "seed": r.seed, "rand_calls": r.rand_calls, "pops": r.pops,
            "recal": r.recal}


class RunOutput:
    """SingleThreadRun with its output files for one system of a plain
    engine: pop while the clock is at most t_end; after the first executed
    event at or past the next output time write a trajectory frame
    (out_trj.gro), a connection-map frame (out_cnt.txt) and the continue
    file (savedData.dat), then move the output time on by Dump_Rate.  A new
    run's first output time is 0 (the first executed event writes); a
    continued run's is its clock plus Dump_Rate (SGThread.c:54-55)."""

    def __init__(self, engine, system, outdir, outputrate, timestep, sys=0, first=0.0):
        self.eng, self.system, self.sys = engine, system, sys
        self.outdir = outdir
        self.outputrate = float(outputrate)
        self.timestep = float(timestep)
        self.record = float(first)
        self.dumps = []                 # (pops, executed events, clock) of each output
        os.makedirs(outdir, exist_ok=True)

    def path(self, name):
        return os.path.join(self.outdir, name)

    def write(self):
        st = self.eng.stats(self.sys)
        state = self.eng.state(self.sys)
        write_gro(self.system, state, self.path("out_trj.gro"), append=True)
        write_cnt(state, self.path("out_cnt.txt"), append=True)
        info = dat_info(self.system, state, frame=st["frame"], per_type=st["per_type"])
        info.timestep = self.timestep
        info.outputrate = self.outputrate
        save_dat(self.system, state, self.path("savedData.dat"), info)
        self.dumps.append((st["pops"], st["frame"], st["currenttime"]))

    def run(self, max_pops=1 << 40):
        pops0 = self.eng.stats(self.sys)["pops"]
        while True:
            left = max_pops - (self.eng.stats(self.sys)["pops"] - pops0)
            if left <= 0:
                break
            before = self.eng.stats(self.sys)["pops"]
            self.eng.run_to_output(self.record, self.timestep, left)
            st = self.eng.stats(self.sys)
            if st["pops"] > before and st["last_exec"] and st["currenttime"] >= self.record:
                self.write()
                self.record += self.outputrate
                if st["currenttime"] > self.timestep:
                    break
            else:
                break                   # the clock passed t_end (or the pop bound) first
        return self.dumps

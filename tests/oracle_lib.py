"""ctypes binding of oracle/liboracle.so (the CPU restatement used as checker).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from sdmd_amd._ctypes_types import SystemDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

POP = np.dtype([
    ("target", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"), ("exec", "<i4"),
    ("nrenew", "<i4"), ("dt", "<f8"), ("pot", "<f8"), ("rand_calls", "<i8"),
    ("r", [("atom", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"),
           ("counter", "<i8"), ("pcounter", "<i8"), ("time", "<f8"), ("pot", "<f8")], (2,)),
])

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "lib"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_open.restype = P
        L.oracle_open.argtypes = [C.POINTER(SystemDesc)]
        L.oracle_close.argtypes = [P]
        L.oracle_set_rng.argtypes = [P, C.c_uint32, C.c_int64]
        L.oracle_rand_calls.restype = C.c_int64
        L.oracle_rand_calls.argtypes = [P]
        L.oracle_predict.argtypes = [P, C.POINTER(C.c_int32)]
        L.oracle_bulk.argtypes = [P]
        L.oracle_inherited.argtypes = [P, C.c_void_p]
        L.oracle_create_cbt.argtypes = [P]
        L.oracle_run.restype = C.c_long
        L.oracle_run.argtypes = [P, C.c_long, C.c_void_p, C.c_void_p, C.c_long]
        L.oracle_get_state.argtypes = [P, C.POINTER(SystemDesc)]
        L.oracle_get_tree.argtypes = [P, C.c_void_p, C.c_void_p]
        L.oracle_set_tree.argtypes = [P, C.c_void_p, C.c_void_p]
        L.oracle_state_hash.restype = C.c_uint64
        L.oracle_state_hash.argtypes = [P]
        L.oracle_currenttime.restype = C.c_double
        L.oracle_currenttime.argtypes = [P]
        L.oracle_counters.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_find_pair_types.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]
        L.oracle_wall_time.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
        _lib = L
    return _lib


class Oracle:
    """One CPU-restated system; `arrays` is a DescArrays giving the start state."""

    def __init__(self, arrays, seed=None, rand_skip=None):
        self.arrays = arrays
        self._desc = arrays.desc()
        self.h = lib().oracle_open(C.byref(self._desc))
        if seed is None:
            seed = arrays.scalars["seed"]
        if rand_skip is None:
            rand_skip = arrays.scalars["rand_calls"]
        lib().oracle_set_rng(self.h, seed, rand_skip)

    def close(self):
        if self.h:
            lib().oracle_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def bulk(self):
        return lib().oracle_bulk(self.h)

    def inherited(self):
        """Boolean per atom (index 0 dummy): the current event record was
        copied from the partner's by CheckDuplication, not solved."""
        out = np.zeros(self.arrays.scalars["natom"] + 1, np.int8)
        lib().oracle_inherited(self.h, out.ctypes.data)
        return out.astype(bool)

    def predict(self, atoms):
        renew = np.zeros(len(atoms) + 2, np.int32)
        renew[0] = len(atoms)
        renew[1:1 + len(atoms)] = atoms
        return lib().oracle_predict(self.h, renew.ctypes.data_as(C.POINTER(C.c_int32)))

    def create_cbt(self):
        return lib().oracle_create_cbt(self.h)

    def run(self, npops, trace=True, hash_every=0):
        tr = np.zeros(npops, POP) if trace else None
        hs = np.zeros((max(npops // hash_every, 1) if hash_every else 1, 2), np.int64)
        n = lib().oracle_run(self.h, npops, tr.ctypes.data if trace else None,
                             hs.ctypes.data if hash_every else None, hash_every)
        if n < 0:
            raise RuntimeError("oracle error %d" % n)
        return tr, (hs[: npops // hash_every] if hash_every else None)

    def state(self):
        """Current state as a DescArrays copy."""
        out = self.arrays.copy()
        d = out.desc()
        lib().oracle_get_state(self.h, C.byref(d))
        out.scalars["currenttime"] = d.currenttime
        out.scalars["rand_calls"] = d.rand_calls
        return out

    def tree(self):
        N = self.arrays.scalars["natom"]
        node = np.zeros(2 * N, np.int32)
        leaf = np.zeros(N + 1, np.int32)
        lib().oracle_get_tree(self.h, node.ctypes.data, leaf.ctypes.data)
        return node, leaf

    def set_tree(self, node, leaf):
        node = np.ascontiguousarray(node, dtype=np.int32)
        leaf = np.ascontiguousarray(leaf, dtype=np.int32)
        lib().oracle_set_tree(self.h, node.ctypes.data, leaf.ctypes.data)

    def state_hash(self):
        return lib().oracle_state_hash(self.h)

    def counters(self):
        per = np.zeros(16, np.int64)
        v = np.zeros(3, np.int64)
        err = np.zeros(1, np.int32)
        lib().oracle_counters(self.h, per.ctypes.data, v[0:].ctypes.data, v[1:].ctypes.data,
                              v[2:].ctypes.data, err.ctypes.data)
        return {"per_type": per, "recal": int(v[0]), "pair_evals": int(v[1]),
                "candidates": int(v[2]), "error": int(err[0])}

    def rand_calls(self):
        return lib().oracle_rand_calls(self.h)

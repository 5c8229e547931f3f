"""ctypes mirror of include/sdmd_types.h (sdmd_system_desc, sdmd_pair)."""
import ctypes as C

import numpy as np

NTYPE = 33
NHBCLASS = 12
NBSLOTS = 5
DYNSLOTS = 8
INFINIT = 1e8
INVALID = -111

# enum EEventType / EInteractionType (DMD.h:162-191)
EVENT_NAMES = ["Coli", "Bond", "HB", "HBNe", "Ther", "Wall", "Obst", "Tunl", "Chrg", "SphO",
               "PCC", "CGSu", "Lagv", "TBD", "Cacl", "Invd"]
SUB_NAMES = ["CoreColli", "WellBounc", "WellCaptr", "WellEscap", "TBDCaptr", "TBDEscap",
             "TBDHBForm", "TBDHBBrik", "Invalid_Inter"]


class Pair(C.Structure):
    _fields_ = [("dmin", C.c_double), ("dmax", C.c_double), ("step0", C.c_int32),
                ("nstep", C.c_int32)]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)
_B = C.POINTER(C.c_int8)


class SystemDesc(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("cellnum", C.c_int32 * 3), ("cellsize", C.c_double * 3),
        ("box", C.c_double * 3), ("cutoffr", C.c_double), ("temperature", C.c_double),
        ("thermo_freq", C.c_double), ("thermostat", C.c_int32), ("wall", C.c_int32),
        ("wall_type", C.c_int32), ("wall_mass", C.c_double), ("nprotein", C.c_int32),
        ("x", _D), ("y", _D), ("z", _D), ("vx", _D), ("vy", _D), ("vz", _D), ("mass", _D),
        ("type", _I), ("aatype", _I), ("protein", _I), ("aanum", _I),
        ("cellx", _I), ("celly", _I), ("cellz", _I),
        ("hb_role", _B), ("hb_neighbor", _I), ("hb_bond", _I), ("hb_itype", _I),
        ("nb_partner", _I), ("nb_partnernum", _I), ("nb_status", _I),
        ("dyn_j", _I), ("dyn_bits", _I),
        ("ev_partner", _I), ("ev_counter", _L), ("ev_pcounter", _L), ("ev_time", _D),
        ("ev_pot", _D), ("ev_type", _I), ("ev_sub", _I),
        ("celllist", _I), ("ncelllist", C.c_int32),
        ("bond_off", _I), ("bond_j", _I), ("bond_dmin", _D), ("bond_dmax", _D),
        ("con_off", _I), ("con_j", _I), ("con", C.POINTER(Pair)),
        ("cmap_off", _I), ("cmap_j", _I), ("cmap_bits", _I),
        ("pair", C.POINTER(Pair)), ("hbpair", C.POINTER(Pair)),
        ("nstep_pool", C.c_int32), ("step_d", _D), ("step_e", _D), ("step_acc", _D),
        ("hb_pot", C.c_double * 6), ("wall_pos", C.c_double * 3),
        ("currenttime", C.c_double), ("seed", C.c_uint32), ("rand_calls", C.c_int64),
    ]


# per-atom arrays: name -> (numpy dtype, elements per atom)
ATOM_ARRAYS = {
    "x": (np.float64, 1), "y": (np.float64, 1), "z": (np.float64, 1),
    "vx": (np.float64, 1), "vy": (np.float64, 1), "vz": (np.float64, 1),
    "mass": (np.float64, 1), "type": (np.int32, 1), "aatype": (np.int32, 1),
    "protein": (np.int32, 1), "aanum": (np.int32, 1), "cellx": (np.int32, 1),
    "celly": (np.int32, 1), "cellz": (np.int32, 1), "hb_role": (np.int8, 1),
    "hb_neighbor": (np.int32, 1), "hb_bond": (np.int32, 1), "hb_itype": (np.int32, 1),
    "nb_partner": (np.int32, NBSLOTS), "nb_partnernum": (np.int32, 1),
    "nb_status": (np.int32, 1), "dyn_j": (np.int32, DYNSLOTS), "dyn_bits": (np.int32, DYNSLOTS),
    "ev_partner": (np.int32, 1), "ev_counter": (np.int64, 1), "ev_pcounter": (np.int64, 1),
    "ev_time": (np.float64, 1), "ev_pot": (np.float64, 1), "ev_type": (np.int32, 1),
    "ev_sub": (np.int32, 1),
}


def view(ptr, n, dtype):
    """numpy view (no copy) of n elements behind a ctypes pointer."""
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                 shape=(n,))


def _own(v):
    """An independent copy of a ctypes array value (other values as they are)."""
    if isinstance(v, C.Array):
        return type(v)(*v)
    return v


class DescArrays:
    """Owns numpy copies of every array of a desc and can rebuild a desc
    pointing at them (so tests and the oracle can edit state freely)."""

    def __init__(self, desc: SystemDesc):
        d = desc
        # fixed-size fields (cellnum, box, ...) are copied out of the struct:
        # a ctypes array read from a field shares the struct's memory, which
        # the desc's owner may free
        self.scalars = {f: _own(getattr(d, f)) for f, _ in SystemDesc._fields_
                        if not isinstance(getattr(d, f), (C._Pointer,))}
        N1 = d.natom + 1
        self.arr = {}
        for name, (dt, per) in ATOM_ARRAYS.items():
            self.arr[name] = view(getattr(d, name), N1 * per, dt).copy()
        self.arr["celllist"] = view(d.celllist, d.ncelllist, np.int32).copy()
        nb = view(d.bond_off, d.natom + 2, np.int32)[-1]
        nc = view(d.con_off, d.natom + 2, np.int32)[-1]
        ncm = view(d.cmap_off, d.natom + 2, np.int32)[-1]
        for f, n in (("bond_off", d.natom + 2), ("con_off", d.natom + 2), ("cmap_off", d.natom + 2)):
            self.arr[f] = view(getattr(d, f), n, np.int32).copy()
        self.arr["bond_j"] = view(d.bond_j, nb, np.int32).copy()
        self.arr["bond_dmin"] = view(d.bond_dmin, nb, np.float64).copy()
        self.arr["bond_dmax"] = view(d.bond_dmax, nb, np.float64).copy()
        self.arr["con_j"] = view(d.con_j, nc, np.int32).copy()
        self.arr["con"] = np.ctypeslib.as_array(d.con, shape=(max(nc, 1),)).copy()[:nc] if nc else np.zeros(0, dtype=np.dtype([("dmin", "<f8"), ("dmax", "<f8"), ("step0", "<i4"), ("nstep", "<i4")]))
        self.arr["cmap_j"] = view(d.cmap_j, ncm, np.int32).copy()
        self.arr["cmap_bits"] = view(d.cmap_bits, ncm, np.int32).copy()
        self.arr["pair"] = np.ctypeslib.as_array(d.pair, shape=(NTYPE * NTYPE,)).copy()
        self.arr["hbpair"] = np.ctypeslib.as_array(d.hbpair, shape=(NHBCLASS * NTYPE * NTYPE,)).copy()
        npool = d.nstep_pool
        for f in ("step_d", "step_e", "step_acc"):
            self.arr[f] = view(getattr(d, f), npool + 1, np.float64).copy()
        self._desc = None

    def desc(self) -> SystemDesc:
        d = SystemDesc()
        for k, v in self.scalars.items():
            setattr(d, k, v)
        for name, a in self.arr.items():
            a = np.ascontiguousarray(a)
            self.arr[name] = a
            if name in ("pair", "hbpair", "con"):
                if len(a) == 0:
                    a = np.zeros(1, dtype=self.arr["pair"].dtype)
                    self.arr[name] = a[:0]
                    self._keep_con = a
                setattr(d, name, a.ctypes.data_as(C.POINTER(Pair)))
            else:
                ct = np.ctypeslib.as_ctypes_type(a.dtype)
                setattr(d, name, a.ctypes.data_as(C.POINTER(ct)))
        self._desc = d
        return d

    def copy(self):
        other = object.__new__(DescArrays)
        other.scalars = {k: _own(v) for k, v in self.scalars.items()}
        other.arr = {k: v.copy() for k, v in self.arr.items()}
        other._desc = None
        return other

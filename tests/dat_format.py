"""Reader of the reference's savedData.dat (SaveDAT, DataSave.c:512-585):
a raw dump of its x86-64 structs (DMD.h).  Test infrastructure: parses both
the reference's files and this library's, so they can be compared field by
field."""
import numpy as np

PROP = np.dtype([("name", "S8"), ("extra", "S16", (2,)), ("aa", "S8"), ("num", "<i4"), ("type", "<i4"),
                 ("aatype", "<i4"), ("pad0", "<i4"), ("charge", "<f8"), ("mass", "<f8"), ("color", "<f8", (3,)),
                 ("seq", "<i4", (3,)), ("pad1", "<i4"), ("bond", "<u8"), ("constr", "<u8")])
DYN = np.dtype([("cell", "<i4", (4,)), ("coord", "<f8", (4,)), ("vel", "<f8", (4,)), ("hbbond", "<i4"),
                ("hbnb", "<i4"), ("role", "i1"), ("pad", "i1", (3,)), ("hbitype", "<i4"), ("nbp", "<i4", (5,)),
                ("nbnum", "<i4"), ("nbstat", "<i4"), ("pad2", "<i4"), ("evp", "<i4"), ("pad3", "<i4"),
                ("evc", "<i8"), ("evpc", "<i8"), ("evt", "<f8"), ("evpot", "<f8"), ("evtype", "<i4"),
                ("evsub", "<i4")])
AA = np.dtype([("name", "S5"), ("pad", "S3"), ("type", "<i4"), ("start", "<i4"), ("end", "<i4"),
               ("protein", "<i4"), ("mass", "<f8")])
PEP = np.dtype([("start_atom", "<i4"), ("end_atom", "<i4"), ("start_aa", "<i4"), ("end_aa", "<i4"),
                ("mass", "<f8")])
REMD = np.dtype([("flag", "<i4"), ("port", "<i4"), ("output_rate", "<i4"), ("server", "S50"), ("extra", "S50"),
                 ("num", "<i4"), ("pad", "<i4"), ("T", "<f8")])
assert PROP.itemsize == 136 and DYN.itemsize == 176 and AA.itemsize == 32 and PEP.itemsize == 24
assert REMD.itemsize == 128

# fields that carry no state: heap pointers, padding, the index-0 entries
# of the residue/chain/atom arrays, the index-0 slots of the 1-based vectors
NON_STATE_PROP = ("pad0", "pad1", "bond", "constr")
NON_STATE_DYN = ("pad", "pad2", "pad3")


def read(buf):
    """dict of the file's fields (numpy arrays / scalars)."""
    if isinstance(buf, (str, bytes)) and not isinstance(buf, bytes):
        buf = open(buf, "rb").read()
    o = 0

    def take(dt, n=1):
        nonlocal o
        a = np.frombuffer(buf, dt, n, o)
        o += a.nbytes
        return a

    N, nprot, naa = (int(v) for v in take("<i4", 3))
    out = {"natom": N, "nprot": nprot, "naa": naa}
    out["cmap"] = take("<i4", (N + 1) ** 2).reshape(N + 1, N + 1)
    out["frame"] = int(take("<i8")[0])
    out["timestep"], out["currenttime"], out["cutoffr"], out["temperature"], out["outputrate"] = take("<f8", 5)
    out["method"] = take("S20")[0]
    out["box"] = take("<f8", 4)
    out["aa"] = take(AA, naa + 1)
    out["pep"] = take(PEP, nprot + 1)
    atoms = np.frombuffer(buf, np.dtype([("p", PROP), ("d", DYN)]), N + 1, o)
    o += atoms.nbytes
    out["prop"], out["dyn"] = atoms["p"], atoms["d"]
    out["wall_int"] = take("<i4", 2)
    out["wall_dbl"] = take("<f8", 5)
    out["event_sums"] = take("<i8", 7)
    out["cellnum"] = take("<i4", 4)
    cn = out["cellnum"]
    out["celllist"] = take("<i4", N + 1 + int(cn[1]) * int(cn[2]) * int(cn[3]))
    out["cellsize"] = take("<f8", 4)
    out["hb"] = take("<i4", 2)
    out["seed"] = int(take("<u4")[0])
    out["remd"] = take(REMD)[0]
    assert o == len(buf), (o, len(buf))
    return out


def state_fields(d):
    """The fields that hold state, as comparable arrays: residues, chains
    and atoms from index 1, the 1-based vectors from slot 1, names as C
    strings."""
    cstr = lambda a: np.array([x.split(b"\0")[0] for x in a.ravel()]).reshape(a.shape)
    p, y = d["prop"][1:], d["dyn"][1:]
    f = {
        "header": np.array([d["natom"], d["nprot"], d["naa"]]), "cmap": d["cmap"], "frame": d["frame"],
        "run": np.array([d["timestep"], d["currenttime"], d["cutoffr"], d["temperature"], d["outputrate"]]),
        "method": cstr(np.array([d["method"]])), "box": d["box"][1:],
        "aa_name": cstr(d["aa"]["name"][1:]), "pep": d["pep"][1:],
        "wall_int": d["wall_int"], "wall_dbl": d["wall_dbl"], "event_sums": d["event_sums"],
        "cellnum": d["cellnum"], "celllist": d["celllist"], "cellsize": d["cellsize"][1:], "hb": d["hb"],
        "seed": d["seed"],
    }
    for k in ("type", "start", "end", "protein", "mass"):
        f["aa_" + k] = d["aa"][k][1:]
    for k in ("name", "aa"):
        f["prop_" + k] = cstr(p[k])
    f["prop_extra"] = cstr(p["extra"])
    for k in ("num", "type", "aatype", "charge", "mass", "color", "seq"):
        f["prop_" + k] = p[k]
    f["dyn_cell"], f["dyn_coord"], f["dyn_vel"] = y["cell"][:, 1:], y["coord"][:, 1:], y["vel"][:, 1:]
    for k in ("hbbond", "hbnb", "role", "hbitype", "nbp", "nbnum", "nbstat", "evp", "evc", "evpc", "evt", "evpot",
              "evtype", "evsub"):
        f["dyn_" + k] = y[k]
    f["atom0_event"] = np.array([d["dyn"][0]["evp"], d["dyn"][0]["evt"]])
    r = d["remd"]
    f["remd"] = np.array([r["flag"], r["port"], r["output_rate"], r["num"], r["T"]])
    f["remd_names"] = cstr(np.array([r["server"], r["extra"]]))
    return f

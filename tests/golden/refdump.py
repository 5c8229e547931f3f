"""Readers for the binary dumps written by oracle/ref_hooks.c (test-only)."""
import os

import numpy as np

STATE_ATOM = np.dtype([("dv", "<f8", 9), ("iv", "<i4", 22), ("lv", "<i8", 2)])

POP = np.dtype([
    ("target", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"), ("exec", "<i4"),
    ("nrenew", "<i4"), ("dt", "<f8"), ("pot", "<f8"), ("rand_calls", "<i8"),
    ("r", [("atom", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"),
           ("counter", "<i8"), ("pcounter", "<i8"), ("time", "<f8"), ("pot", "<f8")], (2,)),
])


def read_state(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, "<i4", 8, 0)
    N = int(hdr[1])
    off = 32
    g = np.frombuffer(raw, "<f8", 9, off); off += 72
    misc = np.frombuffer(raw, "<i8", 4, off); off += 32
    ct = np.frombuffer(raw, "<f8", 1, off)[0]; off += 8
    atoms = np.frombuffer(raw, STATE_ATOM, N, off); off += STATE_ATOM.itemsize * N
    ncl = int(np.frombuffer(raw, "<i4", 1, off)[0]); off += 4
    cl = np.frombuffer(raw, "<i4", ncl, off)
    dv, iv, lv = atoms["dv"], atoms["iv"], atoms["lv"]
    st = {
        "natom": N, "cellnum": hdr[2:5].copy(), "cellsize": g[0:3].copy(), "box": g[3:6].copy(),
        "cutoffr": g[6], "temperature": g[7], "thermo_freq": g[8],
        "rand_calls": int(misc[0]), "seed": int(misc[1]), "pops": int(misc[2]), "frame": int(misc[3]),
        "currenttime": ct, "celllist": cl.copy(),
    }
    names = ["x", "y", "z", "vx", "vy", "vz", "mass", "ev_time", "ev_pot"]
    for k, n in enumerate(names):
        st[n] = np.concatenate([[0.0], dv[:, k]])
    inames = ["cellx", "celly", "cellz", "type", "aatype", "protein", "aanum", "ev_partner",
              "ev_type", "ev_sub", "hb_bond", "hb_neighbor", "hb_role", "hb_itype"]
    for k, n in enumerate(inames):
        st[n] = np.concatenate([[0], iv[:, k]]).astype(np.int32)
    st["nb_partner"] = np.concatenate([np.zeros((1, 5), np.int32), iv[:, 14:19]])
    st["nb_partnernum"] = np.concatenate([[0], iv[:, 19]]).astype(np.int32)
    st["nb_status"] = np.concatenate([[0], iv[:, 20]]).astype(np.int32)
    st["ev_counter"] = np.concatenate([[0], lv[:, 0]]).astype(np.int64)
    st["ev_pcounter"] = np.concatenate([[0], lv[:, 1]]).astype(np.int64)
    return st


def read_trace(path):
    return np.fromfile(path, POP)


def read_hashes(path):
    return np.fromfile(path, "<i8").reshape(-1, 2)


def read_topology(path):
    """topo.txt -> dict of bonds/constr/cmap per atom and tables (floats from %a)."""
    bonds, constr, cmap, atoms = {}, {}, {}, {}
    pairs, hbpairs = {}, {}
    hbpot = None
    for ln in open(path):
        t = ln.split()
        if not t:
            continue
        k = t[0]
        if k == "atom":
            atoms[int(t[1])] = t[2:]
        elif k == "bond":
            bonds.setdefault(int(t[1]), []).append((int(t[2]), float.fromhex(t[3]), float.fromhex(t[4])))
        elif k == "constr":
            n = int(t[5])
            steps = [tuple(float.fromhex(v) for v in t[6 + 3 * q: 9 + 3 * q]) for q in range(n)]
            constr.setdefault(int(t[1]), []).append((int(t[2]), float.fromhex(t[3]), float.fromhex(t[4]), steps))
        elif k == "cmap":
            cmap.setdefault(int(t[1]), {})[int(t[2])] = int(t[3])
        elif k == "pair":
            n = int(t[5])
            steps = [tuple(float.fromhex(v) for v in t[6 + 3 * q: 9 + 3 * q]) for q in range(n)]
            pairs[(int(t[1]), int(t[2]))] = (float.fromhex(t[3]), float.fromhex(t[4]), steps)
        elif k == "hbpair":
            n = int(t[6])
            steps = [tuple(float.fromhex(v) for v in t[7 + 3 * q: 10 + 3 * q]) for q in range(n)]
            hbpairs[(int(t[1]), int(t[2]), int(t[3]))] = (float.fromhex(t[4]), float.fromhex(t[5]), steps)
        elif k == "hbpot":
            hbpot = [float.fromhex(v) for v in t[1:7]]
    return {"atoms": atoms, "bonds": bonds, "constr": constr, "cmap": cmap, "pairs": pairs,
            "hbpairs": hbpairs, "hbpot": hbpot}


def load_dir(d):
    out = {"setup": read_state(os.path.join(d, "setup.bin")),
           "topo": read_topology(os.path.join(d, "topo.txt"))}
    if os.path.exists(os.path.join(d, "trace.bin")):
        out["trace"] = read_trace(os.path.join(d, "trace.bin"))
        out["hashes"] = read_hashes(os.path.join(d, "hash.bin"))
    if os.path.exists(os.path.join(d, "final.bin")):
        out["final"] = read_state(os.path.join(d, "final.bin"))
    en = os.path.join(d, "energy.txt")
    if os.path.exists(en):
        with open(en) as fh:
            for ln in fh:
                tok = ln.split()
                out["energy_" + tok[0]] = np.array([float.fromhex(tok[1]), float.fromhex(tok[2])])
    tree = os.path.join(d, "tree.bin")
    if os.path.exists(tree):
        raw = np.fromfile(tree, "<i4")
        N = out["setup"]["natom"]
        out["tree_node"], out["tree_leaf"] = raw[:2 * N], raw[2 * N:]
    return out

"""Seeded synthetic peptide systems in the reference's input formats.

The builder writes what the reference's setup reads: a GROMACS ``.gro``
coordinate file (nm, 3 decimals; read by Initialization.c:877-989) and the
16-line positional ``parameter.txt`` (read by Initialization.c:512-730).
Atom names, types and masses come from the Ding residue templates
(``Library_Ding/AA/*.txt``, parsed by Initialization.c:1098-1230), so every
atom written here survives the reference's template filter
(Initialization.c:910-944, CheckValidAtom 1342-1357).

Geometry: each chain starts from a seeded CA random walk (or a straight
zig-zag for extended chains) and is relaxed by iterative distance
projection until every BOND/ANGLE/INTER/CONSTRn range of the templates and
every non-bonded hard core of the pair table holds with a margin.  Chains
are then placed with random rotations/translations without inter-chain
contacts closer than ``min_gap`` (minimum image).  Coordinates stay
unwrapped per chain, as the reference's energy minimiser measures
distances without periodic images (ToolFunctions.c:378-382) and wraps only
afterwards (SGThread.c:44).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "Library_Ding")

AMINO_ACIDS = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
               "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"]

# Models.c:129-209 (AtomModel): atom type name -> type index.
ATOM_TYPES = {
    "OZB": 26, "NZB": 18, "CRB": 6, "CA": 9, "CA1": 10, "CA2": 11, "CA3": 12,
    "CA1P": 14, "CA2P": 15, "CM1": 13, "CM1P": 16, "CR": 5, "CRN": 7, "CRNP": 8,
    "NCK": 21, "NCR": 22, "NR": 23, "NZNP": 19, "NZNQ": 20, "NZ": 17, "OC": 28,
    "OW": 29, "OZH": 25, "OZ": 24, "SGNP": 30, "SG": 31, "OZB_HB": 27, "HB": 2,
    "HO": 1, "HN": 1, "HS": 1, "HW": 1, "H": 1, "H_HB": 3, "HB_HB": 4, "GEL": 32,
}

BOLTZMANN_KCAL = 0.0019872041  # kcal/(mol K), REMDExec.c:19


@dataclass
class Residue:
    name: str
    atoms: list            # [(name, type_name, mass)]
    bonds: list            # [(a, b, length, delta)]  BOND/ANGLE/INTER
    constr: list           # [(a, b, dmin, dmax)]     CONSTRn hard limits


def _parse_template(path: str) -> Residue:
    atoms, bonds, constr = [], [], []
    with open(path) as fh:
        lines = fh.read().split("\n")
    name = lines[0].split()[1] if len(lines[0].split()) > 1 else os.path.basename(path)[:-4]
    for ln in lines[1:]:
        tok = ln.split()
        if not tok:
            continue
        kind = tok[0]
        if kind == "ATOM":
            atoms.append((tok[1], tok[2], float(tok[3])))
        elif kind in ("BOND", "ANGLE", "INTER"):
            bonds.append((tok[1], tok[2], float(tok[3]), float(tok[4])))
        elif kind.startswith("CONSTR"):
            nstep = int(kind[6])
            vals = [float(v) for v in tok[3:]]
            dmin, dmax = vals[0], vals[1 + 2 * nstep]
            constr.append((tok[1], tok[2], dmin, dmax))
    return Residue(name, atoms, bonds, constr)


_TEMPLATE_CACHE: dict = {}


def templates(lib_dir: str = LIB_DIR) -> dict:
    if lib_dir not in _TEMPLATE_CACHE:
        out = {}
        for aa in AMINO_ACIDS + ["PRO_INTER"]:
            out[aa] = _parse_template(os.path.join(lib_dir, "AA", aa + ".txt"))
        _TEMPLATE_CACHE[lib_dir] = out
    return _TEMPLATE_CACHE[lib_dir]


def pair_core_table(lib_dir: str = LIB_DIR) -> np.ndarray:
    """Hard-core distance (A) per type pair, with the H-bond type fallback of
    ToolFunctions.c:278-292 (RightPair) folded in."""
    core = np.zeros((33, 33))
    with open(os.path.join(lib_dir, "InteractionPotentialTable.txt")) as fh:
        for ln in fh:
            if ln.startswith("#") or not ln.strip():
                continue
            tok = ln.split()
            a, b = ATOM_TYPES[tok[0]], ATOM_TYPES[tok[1]]
            core[a, b] = core[b, a] = float(tok[2])
    back = {3: 1, 4: 2, 27: 26}
    for a in range(33):
        for b in range(33):
            if core[a, b] == 0:
                for a2, b2 in ((back.get(a, a), b), (a, back.get(b, b)), (back.get(a, a), back.get(b, b))):
                    if core[a2, b2] > 0:
                        core[a, b] = core[a2, b2]
                        break
    return core


@dataclass
class Chain:
    sequence: list
    names: list = field(default_factory=list)
    resid: list = field(default_factory=list)     # 1-based residue index in chain
    resname: list = field(default_factory=list)
    types: list = field(default_factory=list)
    xyz: np.ndarray = None


def _chain_atoms(seq, tpl):
    names, resid, resname, types = [], [], [], []
    for r, aa in enumerate(seq, start=1):
        for (nm, ty, _m) in tpl[aa].atoms:
            if nm == "H" and r == 1:
                continue          # N-terminal amide H is not present (ala10.gro carries H1-H3)
            names.append(nm)
            resid.append(r)
            resname.append(aa)
            types.append(ATOM_TYPES[ty])
    return names, resid, resname, types


def _restraints(seq, names, resid, tpl):
    """Distance ranges (i, j, lo, hi) implied by the templates for one chain."""
    index = {(r, n): k for k, (r, n) in enumerate(zip(resid, names))}
    rng = {}

    def add(r, a, b, lo, hi, override):
        ra, rb = r, r
        if a.startswith("+"):
            ra, a = r + 1, a[1:]
        if b.startswith("+"):
            rb, b = r + 1, b[1:]
        ia, ib = index.get((ra, a)), index.get((rb, b))
        if ia is None or ib is None:
            return
        key = (min(ia, ib), max(ia, ib))
        if override or key not in rng:
            rng[key] = (lo, hi)

    for r, aa in enumerate(seq, start=1):
        t = tpl[aa]
        for (a, b, L, d) in t.bonds:
            add(r, a, b, L * (1 - d), L * (1 + d), False)
        for (a, b, lo, hi) in t.constr:
            add(r, a, b, lo, hi, False)
        if aa not in ("ALA", "PRO"):
            ala = tpl["ALA"]
            for (a, b, L, d) in ala.bonds:
                add(r, a, b, L * (1 - d), L * (1 + d), False)
            for (a, b, lo, hi) in ala.constr:
                add(r, a, b, lo, hi, False)
        if r < len(seq) and seq[r] == "PRO":
            pi = tpl["PRO_INTER"]
            for (a, b, L, d) in pi.bonds:
                add(r, a, b, L * (1 - d), L * (1 + d), True)
            for (a, b, lo, hi) in pi.constr:
                add(r, a, b, lo, hi, False)
    out = np.array([(i, j, lo, hi) for (i, j), (lo, hi) in rng.items()], dtype=float)
    return out


def _relax(xyz, restr, types, core, rng, rounds=12, margin=0.03, squeeze=0.0):
    """Minimise a flat-bottom penalty: squared violation of every distance
    range (shrunk by a small margin) and of every non-bonded hard core."""
    from scipy.optimize import minimize
    from scipy.spatial import cKDTree

    n = len(xyz)
    ii = restr[:, 0].astype(np.int64)
    jj = restr[:, 1].astype(np.int64)
    width = restr[:, 3] - restr[:, 2]
    shrink = np.minimum(0.25 * width, 0.005)
    lo = restr[:, 2] + shrink
    hi = restr[:, 3] - shrink
    bonded_keys = np.sort(ii * n + jj)
    types = np.asarray(types)

    def nonbonded_pairs(x):
        pairs = cKDTree(x[:, :3]).query_pairs(5.0 if x.shape[1] == 3 else 8.0, output_type="ndarray")
        if len(pairs):
            keys = pairs[:, 0].astype(np.int64) * n + pairs[:, 1]
            pos = np.searchsorted(bonded_keys, keys)
            pos[pos >= len(bonded_keys)] = 0
            pairs = pairs[bonded_keys[pos] != keys]
        return pairs

    def violations(x, pairs, tol=0.0):
        d = np.linalg.norm(x[jj] - x[ii], axis=1)
        nv = int(np.sum((d < restr[:, 2] + tol) | (d > restr[:, 3] - tol)))
        if len(pairs):
            a, b = pairs[:, 0], pairs[:, 1]
            r = np.linalg.norm(x[b] - x[a], axis=1)
            nv += int(np.sum(r < core[types[a], types[b]] + tol))
        return nv

    for rnd in range(rounds):
        pairs = nonbonded_pairs(xyz)
        if len(pairs):
            pa, pb = pairs[:, 0], pairs[:, 1]
            pmin = core[types[pa], types[pb]] + margin
        else:
            pa = pb = np.zeros(0, np.int64)
            pmin = np.zeros(0)

        dim = xyz.shape[1]

        def f(flat):
            x = flat.reshape(n, dim)
            g = np.zeros_like(x)
            d = x[jj] - x[ii]
            r = np.linalg.norm(d, axis=1) + 1e-12
            ex = np.where(r < lo, r - lo, np.where(r > hi, r - hi, 0.0))
            e = np.sum(ex * ex)
            gr = (2 * ex / r)[:, None] * d
            for c in range(dim):
                g[:, c] += np.bincount(jj, gr[:, c], n) - np.bincount(ii, gr[:, c], n)
            if len(pa):
                d2 = x[pb] - x[pa]
                r2 = np.linalg.norm(d2, axis=1) + 1e-12
                ex2 = np.minimum(r2 - pmin, 0.0)
                e += np.sum(ex2 * ex2)
                gr2 = (2 * ex2 / r2)[:, None] * d2
                for c in range(dim):
                    g[:, c] += np.bincount(pb, gr2[:, c], n) - np.bincount(pa, gr2[:, c], n)
            if dim > 3 and squeeze > 0:
                e += squeeze * np.sum(x[:, 3:] ** 2)
                g[:, 3:] += 2 * squeeze * x[:, 3:]
            return e, g.ravel()

        res = minimize(f, xyz.ravel(), jac=True, method="L-BFGS-B",
                       options={"maxiter": 800 if squeeze == 0.0 else 3000, "gtol": 1e-10, "ftol": 1e-14})
        xyz = res.x.reshape(n, dim)
        if dim > 3:
            return xyz, -1
        nv = violations(xyz, nonbonded_pairs(xyz))
        if nv == 0:
            return xyz, 0
        if rnd + 1 < rounds:
            xyz = xyz + rng.normal(scale=0.05, size=xyz.shape)
    return xyz, nv


def _embed(n, restr, types, core, rng, init=None):
    """Distance-geometry start: triangle-smoothed bounds, random distances
    within them, metric-matrix embedding in 3-D."""
    L = core[np.asarray(types)[:, None], np.asarray(types)[None, :]].copy()
    U = np.full((n, n), np.inf)
    ii = restr[:, 0].astype(np.int64)
    jj = restr[:, 1].astype(np.int64)
    L[ii, jj] = L[jj, ii] = restr[:, 2]
    U[ii, jj] = U[jj, ii] = restr[:, 3]
    np.fill_diagonal(U, 0.0)
    np.fill_diagonal(L, 0.0)
    for k in range(n):                      # Floyd-Warshall on upper bounds
        U = np.minimum(U, U[:, k:k + 1] + U[k:k + 1, :])
    for k in range(n):                      # one lower-bound smoothing pass
        L = np.maximum(L, np.maximum(L[:, k:k + 1] - U[k:k + 1, :], L[k:k + 1, :] - U[:, k:k + 1]))
    L = np.minimum(L, U)
    if init is not None:
        from scipy.spatial.distance import cdist
        D = np.clip(cdist(init, init), L, U)
    else:
        D = L + (U - L) * rng.uniform(size=(n, n))
    D = np.triu(D, 1)
    D = D + D.T
    J = np.eye(n) - 1.0 / n
    G = -0.5 * J @ (D * D) @ J
    w, v = np.linalg.eigh(G)
    idx = np.argsort(w)[::-1][:4]
    return v[:, idx] * np.sqrt(np.maximum(w[idx], 1e-6))


def build_chain_dg(seq, rng, extended=False, lib_dir=LIB_DIR):
    tpl = templates(lib_dir)
    core = pair_core_table(lib_dir)
    ch = Chain(list(seq))
    ch.names, ch.resid, ch.resname, ch.types = _chain_atoms(seq, tpl)
    nres = len(seq)
    restr = _restraints(seq, ch.names, ch.resid, tpl)
    n = len(ch.names)
    init = None
    if extended:
        # straight zig-zag backbone guide: residue k near x = 3.6 k
        init = np.zeros((n, 3))
        for k, r in enumerate(ch.resid):
            init[k] = (3.6 * (r - 1), 0.9 * ((r % 2) - 0.5), 0.0)
        init += rng.normal(scale=0.5, size=init.shape)
    nviol = -1
    for attempt in range(8):
        x4 = _embed(n, restr, ch.types, core, rng, init=init)
        # refine in 4-D, then squeeze the fourth coordinate out (avoids
        # mirror-image traps of fused rings), then finish in 3-D
        for sq in (0.0, 0.001, 0.01, 0.1, 1.0, 10.0, 100.0):
            x4, _ = _relax(x4, restr, ch.types, core, rng, rounds=1, squeeze=sq)
        xyz, nviol = _relax(np.ascontiguousarray(x4[:, :3]), restr, ch.types, core, rng, rounds=4)
        if nviol == 0:
            break
    if nviol:
        raise RuntimeError(f"chain relaxation did not converge ({nviol} violations)")
    ch.xyz = xyz - xyz.mean(axis=0)
    return ch


def _restraints_ok(seq, ch, tpl):
    restr = _restraints(seq, ch.names, ch.resid, tpl)
    return restr


_CONFORMERS: dict = {}


def _conformer(aa, lib_dir=LIB_DIR):
    """Relaxed coordinates of residue ``aa`` (built inside ALA-aa-ALA) and of
    the peptide unit CA, C, O, +N, +H, +CA (built inside ALA-ALA-ALA)."""
    key = (aa, lib_dir)
    if key not in _CONFORMERS:
        rng = np.random.default_rng(1234 + AMINO_ACIDS.index(aa) if aa in AMINO_ACIDS else 99)
        ch = build_chain_dg(["ALA", aa, "ALA"], rng, lib_dir=lib_dir)
        res = {n: ch.xyz[k] for k, (n, r) in enumerate(zip(ch.names, ch.resid)) if r == 2}
        unit = {}
        for k, (n, r) in enumerate(zip(ch.names, ch.resid)):
            if r == 1 and n in ("CA", "C", "O"):
                unit[n] = ch.xyz[k]
            if r == 2 and n in ("N", "H", "CA"):
                unit["+" + n] = ch.xyz[k]
        _CONFORMERS[key] = (res, unit)
    return _CONFORMERS[key]


def _rotate(v, axis, ang):
    axis = axis / np.linalg.norm(axis)
    return (v * math.cos(ang) + np.cross(axis, v) * math.sin(ang)
            + axis * np.dot(axis, v) * (1 - math.cos(ang)))


def _frame(a, b, c):
    e1 = b - a
    e1 = e1 / np.linalg.norm(e1)
    u = c - b
    e2 = u - np.dot(u, e1) * e1
    e2 = e2 / np.linalg.norm(e2)
    return np.stack([e1, e2, np.cross(e1, e2)])


def _place(coords, src_a, src_b, src_c, dst_a, dst_b, ref, torsion):
    """Rigidly move ``coords`` so that src_a->dst_a, the src_a->src_b axis
    lies along dst_a->dst_b, and dihedral(ref, dst_a, dst_b, src_c') = torsion."""
    E = _frame(src_a, src_b, src_c)
    f1 = dst_b - dst_a
    f1 = f1 / np.linalg.norm(f1)
    u = ref - dst_a
    u = u - np.dot(u, f1) * f1
    u = u / np.linalg.norm(u)
    f2 = _rotate(u, f1, torsion)
    F = np.stack([f1, f2, np.cross(f1, f2)])
    shift = dst_a - src_a @ E.T @ F
    return {k: v @ E.T @ F + shift for k, v in coords.items()}


_PHI_PSI = [(-120.0, 130.0), (-75.0, 145.0), (-65.0, -40.0), (-140.0, 155.0)]


def build_chain(seq, rng, extended=False, lib_dir=LIB_DIR, attempts=4):
    """Fragment assembly: residue conformers placed along the chain with
    chosen backbone torsions (all 180 deg for an extended chain, otherwise
    seeded draws from common basins with steric rejection), then polished
    by the restraint/hard-core minimiser.  A polish that leaves violations
    is retried with fresh torsions (up to `attempts` assemblies) before the
    slower full distance-geometry build."""
    tpl = templates(lib_dir)
    core = pair_core_table(lib_dir)
    ch = Chain(list(seq))
    ch.names, ch.resid, ch.resname, ch.types = _chain_atoms(seq, tpl)
    restr = _restraints(seq, ch.names, ch.resid, tpl)
    for _ in range(1 if extended else attempts):
        xyz = _assemble(seq, ch, rng, extended, lib_dir)
        xyz, nviol = _relax(xyz, restr, ch.types, core, rng, rounds=6)
        if not nviol:
            ch.xyz = xyz - xyz.mean(axis=0)
            return ch
    return build_chain_dg(seq, rng, extended=extended, lib_dir=lib_dir)


def _assemble(seq, ch, rng, extended, lib_dir):
    from scipy.spatial import cKDTree

    placed = {}
    prev = None           # (N, CA, C) of previous residue, and unit positions for this one
    for r, aa in enumerate(seq, start=1):
        res, _ = _conformer(aa, lib_dir)
        tries = 1 if extended else 30
        for t in range(tries):
            if extended:
                phi, psi = math.pi, math.pi
            else:
                b = _PHI_PSI[int(rng.integers(len(_PHI_PSI)))]
                phi = math.radians(b[0] + rng.normal(scale=15.0))
                psi = math.radians(b[1] + rng.normal(scale=15.0))
            if prev is None:
                cur = {k: v.copy() for k, v in res.items()}
            else:
                n_t, ca_t, cprev = prev["+N"], prev["+CA"], prev["C"]
                cur = _place(res, res["N"], res["CA"], res["C"], n_t, ca_t, cprev, phi)
                if "H" in cur and "+H" in prev:
                    cur["H"] = prev["+H"]
            if r == 1 and "H" in cur:
                del cur["H"]
            # peptide unit to the next residue, rotated about CA-C by psi
            _, unit = _conformer("ALA", lib_dir)
            nxt = None
            if r < len(seq):
                nxt = _place(unit, unit["CA"], unit["C"], unit["+N"], cur["CA"], cur["C"], cur["N"], psi)
                cur["O"] = nxt["O"]
            if extended or not placed:
                break
            pts = np.array([v for k, v in placed.items() if k[0] < r - 1])
            if len(pts) == 0:
                break
            new = np.array(list(cur.values()))
            if cKDTree(pts).query_ball_point(new, 3.0, return_length=True).sum() == 0:
                break
        for k, v in cur.items():
            placed[(r, k)] = v
        prev = {"C": cur["C"]}
        if nxt is not None:
            prev.update(nxt)
    return np.array([placed[(r, n)] for r, n in zip(ch.resid, ch.names)])


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@dataclass
class System:
    chains: list
    box: tuple
    title: str = "sdmd synthetic system"

    @property
    def natom(self):
        return sum(len(c.names) for c in self.chains)


def place_chains(chains, box, rng, min_gap=4.0, max_tries=2000, cylinder=None):
    """Random rotation and position for each chain with no atom closer than
    min_gap to an atom of a placed chain (periodic distances).  With
    cylinder=rmax every atom stays within rmax of the box's x axis line
    (y = box_y/2, z = box_z/2), the smooth cylinder wall's axis."""
    from scipy.spatial import cKDTree

    box = np.asarray(box, dtype=float)
    placed = []
    for ch in chains:
        ext = np.max(np.linalg.norm(ch.xyz, axis=1))
        if 2 * ext >= box.min():
            raise RuntimeError("chain too large for the box")
        if cylinder is not None and ext >= cylinder:
            raise RuntimeError("chain too large for the cylinder")
        for _ in range(max_tries):
            R = _random_rotation(rng)
            # any centre in the box: coordinates stay within (-L/2, 3L/2), which a
            # single periodic wrap (Events.c:220-227) brings back into [0, L)
            c = rng.uniform(0.0, box)
            if cylinder is not None:
                # centre uniform over the disc of radius rmax; the exact
                # radial test below rejects placements that cross it
                rr = cylinder * np.sqrt(rng.uniform())
                ph = rng.uniform(0.0, 2 * np.pi)
                c[1] = box[1] / 2 + rr * np.cos(ph)
                c[2] = box[2] / 2 + rr * np.sin(ph)
            xyz = ch.xyz @ R.T + c
            if cylinder is not None:
                rad = np.hypot(xyz[:, 1] - box[1] / 2, xyz[:, 2] - box[2] / 2)
                if rad.max() > cylinder:
                    continue
            ok = True
            if placed:
                wrapped = np.mod(xyz, box)
                for other in placed:
                    ow = np.mod(other, box)
                    tree = cKDTree(ow, boxsize=box)
                    if tree.query_ball_point(wrapped, min_gap, return_length=True).any():
                        ok = False
                        break
            if ok:
                ch.xyz = xyz
                placed.append(xyz)
                break
        else:
            raise RuntimeError("could not place chain without contacts")
    return chains


def write_gro(system: System, path: str):
    """GROMACS layout %5d%-5s%5s%5d%8.3f%8.3f%8.3f; residue numbers restart at 1
    per chain, which is how the reader detects a new chain (Initialization.c:949)."""
    lines = [system.title, "%5d" % system.natom]
    n = 0
    for ch in system.chains:
        for k in range(len(ch.names)):
            n += 1
            x, y, z = ch.xyz[k] / 10.0
            lines.append("%5d%-5s%5s%5d%8.3f%8.3f%8.3f" % (
                ch.resid[k] % 100000, ch.resname[k], ch.names[k], n % 100000, x, y, z))
    lines.append("%10.5f%10.5f%10.5f" % tuple(b / 10.0 for b in system.box))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_parameter(path, time=1e9, temperature=0.5962, dump_rate=1e9, cutoff=9.0,
                    thermostat="Andersen", thermo_freq=6.0, wall="no", wall_type="no"):
    """The 16 positional lines read by Initialization.c:527-726."""
    rows = [("Continue?", "new"), ("Time", repr(float(time))),
            ("T(reduced)", repr(float(temperature))), ("Dump_Rate", repr(float(dump_rate))),
            ("CutOff_R(A)", repr(float(cutoff))), ("Force_Field", "Ding"),
            ("Thermostat", thermostat), ("Thermo_Freq", repr(float(thermo_freq))),
            ("CGModel", "no"), ("WallExist?", wall), ("WallType", wall_type),
            ("WallDyn", "no"), ("Flow", "no"), ("Obstruct?", "no"), ("Tunnel?", "no"),
            ("SphObst?", "no")]
    with open(path, "w") as fh:
        for k, v in rows:
            fh.write("%-18s %s\n" % (k, v))


def box_for_concentration(nchains: int, mM: float) -> float:
    """Cubic box edge (A) holding nchains at the given molar concentration."""
    litres = nchains / (mM * 1e-3 * 6.02214076e23)
    return (litres * 1e-3) ** (1.0 / 3.0) * 1e10


def ala10_extended(seed=1, box=50.0):
    """ALA10 extended chain in a cubic box (59 atoms)."""
    ch = build_chains([["ALA"] * 10], seed, extended=True, workers=1)[0]
    ch.xyz = ch.xyz + np.array([box / 2, box / 2, box / 2])
    return System([ch], (box, box, box), "ALA10 extended chain")


def single_chain(seq, seed, box):
    """One chain of the given sequence (fragment assembly with seeded
    backbone torsions) centred in a cubic box of edge `box` (A).  Bulky
    residues with a large cut-off give dense cells: TRP-ARG-LYS-PHE-TYR x 6
    (431 atoms, seed 8) in a 72 A box with a 12 A cut-off puts 76 atoms in
    its fullest cell."""
    ch = build_chains([list(seq)], seed, workers=1)[0]
    if 2 * np.max(np.linalg.norm(ch.xyz, axis=1)) >= box:
        raise RuntimeError("chain too large for the box")
    ch.xyz = ch.xyz + np.array([box / 2, box / 2, box / 2])
    return System([ch], (box, box, box), "single chain seed %d" % seed)


def ala10_cylinder(seed=1, box=(50.0, 24.0, 24.0)):
    """ALA10 extended chain inside a smooth cylinder wall of radius box_y/2
    along x: the chain's long axis is turned onto x and centred on the
    cylinder axis (59 atoms, closest approach to the wall about 8 A)."""
    ch = build_chains([["ALA"] * 10], seed, extended=True, workers=1)[0]
    xyz = ch.xyz - ch.xyz.mean(axis=0)
    _, _, vt = np.linalg.svd(xyz, full_matrices=False)
    if np.linalg.det(vt) < 0:
        vt[2] = -vt[2]
    ch.xyz = xyz @ vt.T + np.array([box[0] / 2, box[1] / 2, box[2] / 2])
    return System([ch], tuple(box), "ALA10 extended chain in a cylinder")


def cylinder_peptides(nchains, length, seed, radius=40.0, box_x=160.0, margin=6.0, sequences=None,
                      min_gap=4.0, workers=None, progress=None):
    """nchains chains of uniform-random sequence (or given sequences) inside a
    smooth cylinder wall of the given radius along x, periodic along x only:
    box (box_x, 2*radius, 2*radius); every atom starts at least `margin` A
    from the wall."""
    if sequences is None:
        sequences = random_sequences(nchains, length, seed)
    box = (float(box_x), 2.0 * radius, 2.0 * radius)
    chains = build_chains(sequences, seed, workers=workers, progress=progress)
    place_chains(chains, box, np.random.default_rng([seed, 130363]), min_gap=min_gap,
                 cylinder=radius - margin)
    return System(chains, box, "peptides in a cylinder seed %d" % seed)


def _build_one(args):
    seq, seed, extended = args
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1):
            return build_chain(seq, np.random.default_rng(seed), extended=extended)
    except ImportError:
        return build_chain(seq, np.random.default_rng(seed), extended=extended)


def build_chains(seqs, seed, extended=False, workers=None, progress=None):
    """Build chains in parallel; chain k uses its own seed (seed, k), so the
    result does not depend on worker scheduling.  `progress(done, total)` is
    called as chains complete."""
    for aa in set(a for sq in seqs for a in sq) | {"ALA"}:
        _conformer(aa)
    jobs = [(list(sq), [seed, k], extended) for k, sq in enumerate(seqs)]
    if workers is None:
        workers = min(len(jobs), max(1, min(16, os.cpu_count() or 1)))
    if workers <= 1 or len(jobs) == 1:
        out = []
        for j in jobs:
            out.append(_build_one(j))
            if progress:
                progress(len(out), len(jobs))
        return out
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        out = []
        for ch in pool.imap(_build_one, jobs, chunksize=1):
            out.append(ch)
            if progress:
                progress(len(out), len(jobs))
        return out


def random_sequences(nchains, length, seed):
    rng = np.random.default_rng([seed, 7919])
    seqs = []
    for _ in range(nchains):
        L = length if np.isscalar(length) else int(rng.integers(length[0], length[1] + 1))
        seqs.append([AMINO_ACIDS[k] for k in rng.integers(0, 20, size=L)])
    return seqs


def random_peptides(nchains, length, seed, mM=5.0, box=None, sequences=None, min_gap=4.0,
                    workers=None):
    """nchains chains of uniform-random sequence (or given sequences) placed
    at random in a cubic box sized for the molar concentration mM."""
    if sequences is None:
        sequences = random_sequences(nchains, length, seed)
    if box is None:
        box = box_for_concentration(len(sequences), mM)
    chains = build_chains(sequences, seed, workers=workers)
    place_chains(chains, (box, box, box), np.random.default_rng([seed, 104729]), min_gap=min_gap)
    return System(chains, (box, box, box), "random peptides seed %d" % seed)


def batch_specs(nsys=48, seed=4, length=(10, 80), chains=(1, 32), min_atoms=2000,
                atoms_per_residue=10.45):
    """(nchains, length) of each system of a seeded batch of uneven peptide
    systems: chain length and chain count uniform in the given ranges,
    redrawn until the expected atom count reaches min_atoms."""
    rng = np.random.default_rng([seed, 31337])
    out = []
    while len(out) < nsys:
        nc = int(rng.integers(chains[0], chains[1] + 1))
        L = int(rng.integers(length[0], length[1] + 1))
        if nc * L * atoms_per_residue >= min_atoms:
            out.append((nc, L))
    return out


def batch_peptides(nsys=48, seed=4, mM=5.0, workers=None, progress=None, **kw):
    """A batch of independent peptide systems of uneven size (each its own
    cubic periodic box at mM per chain, widened where a chain's extent needs
    it: edge >= 2*extent + 8 A).  Every chain of every system is built in
    one worker pool; system k uses seed (seed, k)."""
    specs = batch_specs(nsys, seed, **kw)
    seqs, owner = [], []
    for k, (nc, L) in enumerate(specs):
        s = random_sequences(nc, L, seed * 1000 + k)
        seqs += s
        owner += [k] * len(s)
    chains = build_chains(seqs, seed, workers=workers, progress=progress)
    systems = []
    for k, (nc, L) in enumerate(specs):
        mine = [c for c, o in zip(chains, owner) if o == k]
        ext = max(np.max(np.linalg.norm(c.xyz, axis=1)) for c in mine)
        box = max(box_for_concentration(nc, mM), 2 * ext + 8.0)
        place_chains(mine, (box, box, box), np.random.default_rng([seed, k, 104729]))
        systems.append(System(mine, (box, box, box), "batch seed %d system %d" % (seed, k)))
    return systems


def write_library(outdir, wall_atom="CA", wall_mass=12.0):
    """Library_Ding for a run with a smooth wall: the shipped tables and
    residue templates (symlinked) and a Wall.txt whose pseudo-atom type is a
    valid atom type (ReadWall, Initialization.c:1915-1920, maps it with
    AtomModel; the shipped file names PHO, which AtomModel does not know)."""
    lib = os.path.join(outdir, "Library_Ding")
    os.makedirs(lib, exist_ok=True)
    for name in os.listdir(LIB_DIR):
        if name == "Wall.txt":
            continue
        dst = os.path.join(lib, name)
        if not os.path.lexists(dst):
            os.symlink(os.path.join(LIB_DIR, name), dst)
    with open(os.path.join(lib, "Wall.txt"), "w") as fh:
        fh.write("Wall\n   ATOM %-5s %-5s %5.1f  N N\n" % (wall_atom, wall_atom, wall_mass))
    return lib


def write_input_dir(system: System, outdir: str, gro_name="coordinate.gro", wall_atom="CA", **param):
    """Lay out an input folder the reference and this package both read:
    parameter.txt, the .gro file, and Library_Ding (symlinked; with a wall,
    a folder of symlinks plus a Wall.txt naming wall_atom)."""
    os.makedirs(outdir, exist_ok=True)
    write_gro(system, os.path.join(outdir, gro_name))
    write_parameter(os.path.join(outdir, "parameter.txt"), **param)
    link = os.path.join(outdir, "Library_Ding")
    if param.get("wall", "no") != "no":
        write_library(outdir, wall_atom)
    elif not os.path.exists(link):
        os.symlink(LIB_DIR, link)
    return outdir

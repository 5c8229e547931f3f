"""Adversarial pair geometries for the fast engine's FP32 screen.

Each experiment is a pair of rigid extended ALA-ALA chains (11 atoms) far
from every other experiment (22 A apart, more than twice the cut-off).  An
end atom of chain A (the N-terminal N or the C-terminal O) and one of chain
B are placed at a designed separation D on a random axis u, each chain
turned so the chosen atom points at the other; chain A is at rest and
chain B moves rigidly, so no bond of either chain changes and the designed
pair is the nearest contact.  The separations and velocities sit on the decisions
the screen makes with error bands:

  * |r^2 - d^2| ~ band: D = r_k (1 +- eps) for every shell radius r_k of
    the pair (core, steps, well edge), eps from 0 to 1e-3, approaching,
    receding and moving perpendicular to u (b = 0);
  * disc ~ 0: grazing flights whose impact parameter is r_k (1 +- eps);
  * beyond the last step: D just outside the outermost radius.

The positions are set exactly after the host's set-up (the .gro text holds
0.01 A), and the cell indices and cell list are rebuilt as the reference's
LinkList("full") builds them (Events.c:121-156).  Thermostat off: no random
draw enters a prediction.
"""
import os

import numpy as np

NT = 33
PAIRS = [(("N", 1), ("N", 1)), (("O", 2), ("O", 2)), (("N", 1), ("O", 2))]
EPS = [0.0, 1e-12, 1e-9, 1e-7, 1e-5, 1e-3]
SPACING = 22.0


def shell_radii(arr, t1, t2):
    """Shell radii (A) of the collision pair of atom types t1, t2."""
    a, b = min(t1, t2), max(t1, t2)
    p = arr["pair"][a * NT + b]
    sd = arr["step_d"]
    r2 = [float(p["dmin"])] + [float(sd[p["step0"] + k]) for k in range(int(p["nstep"]))]
    if p["dmax"] > 0:
        r2.append(float(p["dmax"]))
    return sorted(set(np.sqrt([x for x in r2 if x > 0])))


def _rotation_to(src, dst):
    """Rotation matrix taking unit vector src onto unit vector dst."""
    v = np.cross(src, dst)
    c = float(src @ dst)
    if np.linalg.norm(v) < 1e-12:
        if c > 0:
            return np.eye(3)
        p = np.array([1.0, 0, 0]) if abs(src[0]) < 0.9 else np.array([0, 1.0, 0])
        p = p - src * (p @ src)
        p /= np.linalg.norm(p)
        return 2 * np.outer(p, p) - np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * (1 / (1 + c))


def _experiments(arr_types_by_name, radii_of, rng):
    out = []
    for na, nb in PAIRS:
        radii = radii_of(arr_types_by_name[na], arr_types_by_name[nb])
        for k, r in enumerate(radii):
            for eps in EPS:
                for sgn in ((1,) if eps == 0 else (1, -1)):
                    D = r * (1 + sgn * eps)
                    for mode in ("in", "out", "perp"):
                        out.append((na, nb, D, mode, None, r))
                    out.append((na, nb, r + 0.8, "graze", r * (1 + sgn * eps), r))
        out.append((na, nb, radii[-1] * (1 + 1e-6), "in", None, radii[-1]))
        out.append((na, nb, radii[-1] * 1.05, "in", None, radii[-1]))
    return out


def boundary_system(tmpdir, seed=17, speed=0.25):
    """(System, experiments, atom index pairs) with the designed geometry."""
    from conftest import golden_system
    from sdmd_amd import builder
    from sdmd_amd.engine import System

    ref, _ = golden_system("ala10nve")                 # the force-field tables
    tab = ref.arrays.arr
    proto = builder.build_chains([["ALA", "ALA"]], seed, extended=True, workers=1)[0]
    names = list(zip(proto.names, [int(r) for r in proto.resid]))
    tbn = dict(zip(names, [int(t) for t in proto.types]))
    rng = np.random.default_rng(seed)
    exps = _experiments(tbn, lambda t1, t2: shell_radii(tab, t1, t2), rng)
    n = len(exps)
    g = int(np.ceil(n ** (1 / 3)))
    box = g * SPACING
    xyz0 = proto.xyz - proto.xyz.mean(axis=0)
    chains, pos, vel, pairs = [], [], [], []
    for e, (na, nb, D, mode, bimp, r) in enumerate(exps):
        ix, iy, iz = e % g, (e // g) % g, e // (g * g)
        centre = (np.array([ix, iy, iz]) + 0.5) * SPACING
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        p = np.cross(u, rng.normal(size=3))
        p /= np.linalg.norm(p)
        ia, ib = names.index(na), names.index(nb)
        va = xyz0[ia] / np.linalg.norm(xyz0[ia])
        vb = xyz0[ib] / np.linalg.norm(xyz0[ib])
        A = xyz0 @ _rotation_to(va, u).T
        B = xyz0 @ _rotation_to(vb, -u).T
        a_at = centre - u * (D / 2)
        A = A - A[ia] + a_at
        B = B - B[ib] + a_at + u * D
        if mode == "in":
            v = -u * speed
        elif mode == "out":
            v = u * speed
        elif mode == "perp":
            v = p * speed
        else:                                           # impact parameter bimp from distance D
            s = bimp / D
            v = (-u * np.sqrt(1 - s * s) + p * s) * speed
        for X, V in ((A, np.zeros(3)), (B, v)):
            ch = builder.Chain(["ALA", "ALA"])
            ch.names, ch.resid, ch.resname, ch.types = proto.names, proto.resid, proto.resname, proto.types
            ch.xyz = X.copy()
            chains.append(ch)
            pos.append(X)
            vel.append(np.tile(V, (len(names), 1)))
        base = 2 * e * len(names)
        pairs.append((base + ia + 1, base + len(names) + ib + 1))
    d = os.path.join(str(tmpdir), "screen_cases")
    builder.write_input_dir(builder.System(chains, (box, box, box), "screen boundary cases"), d, cutoff=9.0,
                            thermostat="no")
    sysm = System.load(d)
    arr = sysm.arrays
    a = arr.arr
    N = sysm.natom
    P = np.concatenate(pos)
    V = np.concatenate(vel)
    assert P.shape[0] == N
    assert np.all(a["type"][1:] == np.tile([tbn[x] for x in names], 2 * n))
    a["x"][1:], a["y"][1:], a["z"][1:] = P[:, 0], P[:, 1], P[:, 2]
    a["vx"][1:], a["vy"][1:], a["vz"][1:] = V[:, 0], V[:, 1], V[:, 2]
    relink(arr)
    return sysm, exps, pairs


def relink(arr):
    """Cell indices and cell list for the current positions, as
    LinkList("full") builds them (Events.c:121-156): atoms 1..N each pushed
    on the front of its cell's list."""
    a, sc = arr.arr, arr.scalars
    N = int(sc["natom"])
    cn = [int(sc["cellnum"][k]) for k in range(3)]
    cs = [float(sc["cellsize"][k]) for k in range(3)]
    cl = a["celllist"]
    cl[:] = 0
    for i in range(1, N + 1):
        c = []
        for k, f in enumerate(("x", "y", "z")):
            v = int(a[f][i] / cs[k])
            if v == cn[k]:
                v -= 1
            assert 0 <= v < cn[k], (i, f, a[f][i])
            c.append(v)
        a["cellx"][i], a["celly"][i], a["cellz"][i] = c
        cell = c[2] * cn[0] * cn[1] + c[1] * cn[0] + c[0] + 1
        cl[i] = cl[N + cell]
        cl[N + cell] = i

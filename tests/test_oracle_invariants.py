"""The CPU oracle pinned by what the mathematics and the reference's own
energy routines fix, beyond the golden event traces:

  * energies: the host's CalKinetE/CalPotenE restatement equals the values
    the compiled reference printed for the same states (set-up state and the
    state after the golden run);
  * contact times: advancing a predicted colliding pair by its event time
    puts it exactly on the shell radius the event names (the textbook
    quadratic root; a wrong sign or a dropped term moves it off);
  * FindPair: the shell bracketing a squared distance, by brute force over
    the step table;
  * momentum: collisions conserve total momentum (NVE run);
  * the event tree: every internal node holds the winner of its children
    under the right-wins tie rule, and the root holds the earliest time;
  * cell lists: every atom sits in exactly one list, the one its position
    indexes."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from conftest import CASES, golden, golden_system
from sdmd_amd import _lib

SHIFT_TOL = 1e-7


def energy(arr):
    d = arr.desc()
    ke, pe = C.c_double(), C.c_double()
    assert _lib.lib().sdmd_host_energy(C.byref(d), C.byref(ke), C.byref(pe)) == 0
    return np.array([ke.value, pe.value])


@pytest.mark.parametrize("name", CASES)
def test_energy_matches_reference(name):
    s, g = golden_system(name)
    assert np.allclose(energy(s.arrays), g["trace"]["energy_setup"], rtol=1e-12, atol=1e-9)
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    o.run(int(g["trace"]["npops"]), trace=False)
    assert np.allclose(energy(o.state()), g["trace"]["energy_final"], rtol=1e-12, atol=1e-9)


def _min_image(d, box):
    return d - box * np.round(d / box)


def test_contact_times_land_on_shell_radii():
    s, _ = golden_system("mix4x20")
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    st = o.state().arr
    sc = s.arrays.scalars
    box = np.array(sc["box"][:3])
    L = oracle_lib.lib()
    checked = 0
    for i in range(1, s.natom + 1):
        j = int(st["ev_partner"][i])
        if st["ev_type"][i] != 0 or j <= 0:       # SDMD_COLI
            continue
        cons = st["con_j"][st["con_off"][i]:st["con_off"][i + 1]]
        if j in cons:
            continue                                # constraint pair: its own table
        r = np.array([st["x"][i] - st["x"][j], st["y"][i] - st["y"][j], st["z"][i] - st["z"][j]])
        r = _min_image(r, box)
        v = np.array([st["vx"][i] - st["vx"][j], st["vy"][i] - st["vy"][j], st["vz"][i] - st["vz"][j]])
        d2 = r @ r
        b = r @ v
        out = np.zeros(5)
        L.oracle_find_pair_types(o.h, int(st["type"][i]), int(st["type"][j]), b, d2, out.ctypes.data)
        assert out[0] < 1e7
        t = st["ev_time"][i]
        rt = r + v * t
        d2t = rt @ rt
        target = [x for x in (out[0], out[1]) if x < 1e7]
        err = min(abs(d2t - x) / x for x in target)
        assert err < 1e-9, (i, j, d2, d2t, target)
        checked += 1
    assert checked > 20


def test_find_pair_brackets_distance_by_brute_force():
    s, _ = golden_system("mix4x20")
    o = oracle_lib.Oracle(s.arrays)
    L = oracle_lib.lib()
    sc = s.arrays
    pair = sc.arr["pair"]
    sd = sc.arr["step_d"]
    rng = np.random.default_rng(5)
    types = sorted(set(int(t) for t in sc.arr["type"][1:]))
    n = 0
    for _ in range(400):
        t1, t2 = (int(x) for x in rng.choice(types, 2))
        p = pair[t1 * 33 + t2]
        if p["dmin"] <= 0:
            continue
        radii = [p["dmin"]] + [sd[p["step0"] + k] for k in range(p["nstep"])]
        if p["dmax"] > 0:
            radii.append(p["dmax"])
        d2 = rng.uniform(radii[0] * 0.8, radii[-1] * 1.2)
        if min(abs(d2 - x) for x in radii) < 1e-6:
            continue
        out = np.zeros(5)
        L.oracle_find_pair_types(o.h, t1, t2, -1.0, d2, out.ctypes.data)
        if d2 < radii[0]:
            assert out[0] >= 1e7 and out[1] >= 1e7
            continue
        inner = max(x for x in radii if x <= d2)
        outer = [x for x in radii if x > d2]
        assert out[0] == inner
        if outer:
            assert out[1] == min(outer)
        else:
            assert out[1] >= 1e7
        n += 1
    assert n > 100


def test_momentum_conserved_in_nve():
    s, _ = golden_system("ala10nve")
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()

    def momentum(a):
        m = a["mass"][1:]
        return np.array([(m * a[k][1:]).sum() for k in ("vx", "vy", "vz")]), (m * np.abs(a["vx"][1:])).sum()

    p0, scale = momentum(o.state().arr)
    o.run(20000, trace=False)
    p1, _ = momentum(o.state().arr)
    assert np.abs(p1 - p0).max() < 1e-10 * scale


def test_tree_and_cell_invariants_after_run():
    s, _ = golden_system("mix4x20")
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    o.run(3000, trace=False)
    st = o.state()
    a = st.arr
    N = s.natom
    node, leaf = o.tree()
    t = a["ev_time"]
    for k in range(N - 1, 0, -1):
        left, right = node[2 * k], node[2 * k + 1]
        if left == 0 or right == 0:
            continue
        assert node[k] == (left if t[left] < t[right] else right), k
    assert t[node[1]] == t[1:].min()
    sc = s.arrays.scalars
    cn = np.array(sc["cellnum"][:3])
    cs = np.array(sc["cellsize"][:3])
    seen = np.zeros(N + 1, int)
    cl = a["celllist"]
    for cz in range(cn[2]):
        for cy in range(cn[1]):
            for cx in range(cn[0]):
                cell = cz * cn[0] * cn[1] + cy * cn[0] + cx + 1
                j = cl[N + cell]
                while j:
                    seen[j] += 1
                    pos = np.array([a["x"][j], a["y"][j], a["z"][j]])
                    lo = np.array([cx, cy, cz]) * cs
                    assert np.all(pos >= lo - SHIFT_TOL) and np.all(pos < lo + cs + SHIFT_TOL), j
                    assert (a["cellx"][j], a["celly"][j], a["cellz"][j]) == (cx, cy, cz)
                    j = cl[j]
    assert np.all(seen[1:] == 1)


# ---------------------------------------------------------------- smooth cylinder wall
def _wall_setup(name, pops):
    s, _ = golden_system(name)
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    if pops:
        o.run(pops, trace=False)
    sc = s.arrays.scalars
    box = np.array(sc["box"][:3], float)
    centre = np.array(sc["wall_pos"][:3], float)
    return s, o, box[1] / 2, centre, int(sc["wall_type"])


@pytest.mark.parametrize("name,pops", [("ala10cyl", 30000), ("cyl8x20", 0)])
def test_wall_times_land_on_shell_limits(name, pops):
    """WallTime's naive quadratic in the radial plane: at the offered time the
    atom's distance to the wall surface is the square root of the shell limit
    the event names -- the lower limit of its shell when it moves out (or
    does not reach the inner radius), the upper one on TBDCaptr."""
    s, o, R, c, wtype = _wall_setup(name, pops)
    st = o.state().arr
    L = oracle_lib.lib()
    checked = 0
    for i in range(1, s.natom + 1):
        t, pot, sub = C.c_double(), C.c_double(), C.c_int()
        if not L.oracle_wall_time(o.h, i, C.byref(t), C.byref(pot), C.byref(sub)):
            continue
        if not t.value < 1e10:
            continue
        y, z, vy, vz = st["y"][i], st["z"][i], st["vy"][i], st["vz"][i]
        r0 = np.hypot(y - c[1], z - c[2])
        w0 = R - r0
        b = (y - c[1]) * vy + (z - c[2]) * vz
        lim = np.zeros(5)
        L.oracle_find_pair_types(o.h, int(st["type"][i]), wtype, -b, w0 * w0, lim.ctypes.data)
        lo, hi, lopot, hipot = lim[:4]
        yt, zt = y + vy * t.value, z + vz * t.value
        wt = R - np.hypot(yt - c[1], zt - c[2])
        if sub.value == 4:                     # TBDCaptr: leaves through the upper limit
            assert abs(wt - np.sqrt(hi)) < 1e-6, (i, wt, hi)
            assert pot.value == hipot
        else:
            assert sub.value in (1, 5)         # WellBounc, TBDEscap
            assert abs(wt - np.sqrt(lo)) < 1e-6, (i, wt, lo)
            assert pot.value == lopot
        assert t.value > 0
        checked += 1
    assert checked >= 10


def test_wall_events_reflect_or_cross_radially():
    """WallEvent's impulse is radial: x velocity and the tangential part in
    the y-z plane are unchanged; a bounce negates the radial speed, a
    crossing of a step of height U changes its square by -2U/m."""
    s, g = golden_system("ala10cyl")
    o = oracle_lib.Oracle(s.arrays)
    o.bulk()
    tr, _ = o.run(int(g["trace"]["npops"]), trace=True)
    pops = np.flatnonzero((tr["type"] == 5) & (tr["exec"] == 1))
    assert len(pops) >= 20
    sc = s.arrays.scalars
    c = np.array(sc["wall_pos"][:3], float)
    o2 = oracle_lib.Oracle(s.arrays)
    o2.bulk()
    done = 0
    kinds = set()
    for k in pops:
        o2.run(int(k) - done, trace=False)
        before = o2.state().arr
        o2.run(1, trace=False)
        after = o2.state().arr
        done = int(k) + 1
        i = int(tr["target"][k])
        m = before["mass"][i]
        U = tr["pot"][k]
        n = np.array([after["y"][i] - c[1], after["z"][i] - c[2]])
        n /= np.linalg.norm(n)
        v0 = np.array([before["vy"][i], before["vz"][i]])
        v1 = np.array([after["vy"][i], after["vz"][i]])
        assert after["vx"][i] == before["vx"][i]
        t0, t1 = v0 - (v0 @ n) * n, v1 - (v1 @ n) * n
        assert np.allclose(t0, t1, rtol=0, atol=1e-12)
        vr0, vr1 = v0 @ n, v1 @ n
        bounce = abs(vr1 + vr0) < 1e-12 * max(1.0, abs(vr0))
        cross = abs(vr1 * vr1 - (vr0 * vr0 - 2 * U / m)) < 1e-9 and np.sign(vr1) == np.sign(vr0)
        assert bounce or cross, (k, tr["sub"][k], vr0, vr1, U)
        if tr["sub"][k] == 1:
            assert bounce
        kinds.add(("bounce" if bounce else "cross", int(tr["sub"][k])))
    # this run has TBDCaptr and TBDEscap events, both crossing steps
    assert {k[1] for k in kinds} >= {4, 5}


def test_screen_boundary_cases_are_built_as_designed(tmp_path):
    """The adversarial pair geometries (tests/screen_cases.py): every
    experiment's chains keep their bonds (no bond event), and the designed
    pair is the predicted contact of one of its atoms in most experiments,
    so the near-boundary decisions are the ones that choose the winners."""
    import screen_cases

    sysm, exps, pairs = screen_cases.boundary_system(tmp_path)
    o = oracle_lib.Oracle(sysm.arrays)
    assert o.bulk() == 0
    st = o.state().arr
    assert np.all(st["ev_type"][1:] != 2)                    # SDMD_BOND
    won = [st["ev_partner"][ia] == ib or st["ev_partner"][ib] == ia for ia, ib in pairs]
    assert len(exps) > 700 and np.mean(won) > 0.4

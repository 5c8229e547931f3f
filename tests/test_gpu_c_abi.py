"""A C caller of the library (tests/c_abi_loop.c: the reference's
SingleThreadRun written against include/*.h only, one fine-grained entry
point per reference call) reproduces the compiled reference's event
sequence, pop for pop, and reports the cost of a pop through the
fine-grained entry points."""
import os
import subprocess

import numpy as np
import pytest

from conftest import HERE, golden

pytestmark = pytest.mark.gpu

REC = np.dtype([("target", "<i4"), ("partner", "<i4"), ("type", "<i4"), ("sub", "<i4"), ("exec", "<i4"),
                ("pad", "<i4"), ("dt", "<f8")])


def test_c_abi_loop_matches_reference(require_gpu, tmp_path):
    from sdmd_amd.builder import LIB_DIR

    exe = os.path.join(HERE, "c_abi_loop.bin")
    assert os.path.exists(exe), "tests/c_abi_loop.bin is built by __graft_entry__.build()"
    g = golden("ala10ext")
    d = tmp_path / "in"
    d.mkdir()
    (d / "coordinate.gro").write_text(str(g["input"]["gro"]))
    (d / "parameter.txt").write_text(str(g["input"]["parameter"]))
    os.symlink(LIB_DIR, str(d / "Library_Ding"))
    pops = 3000
    out = tmp_path / "trace.bin"
    r = subprocess.run([exe, str(d), str(int(g["setup"]["seed"])), str(pops), str(out)], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    tr = np.fromfile(str(out), REC)
    ref = g["trace"]
    assert len(tr) == pops
    for f in ("target", "partner", "type", "sub", "exec", "dt"):
        bad = np.flatnonzero(tr[f] != ref[f][:pops].astype(tr[f].dtype))
        assert len(bad) == 0, (f, int(bad[0]) if len(bad) else None)

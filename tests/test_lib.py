"""The C-ABI library loads and exports every function include/*.h declares,
the headers build as C99, and bench.py refuses more than one GPU (no device
calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT


def declared_functions():
    names = set()
    for h in ("sdmd_host.h", "sdmd_gpu.h", "sdmd_remd.h", "sdmd_io.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sdmd_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from sdmd_amd import _lib

    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.HOST_SYMBOLS + _lib.GPU_SYMBOLS) == set(names)


def test_gpu_open_rejects_bad_arguments_without_device():
    from sdmd_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.sdmd_gpu_open(None, 0, None, ctypes.byref(h)) == -1
    assert L.sdmd_gpu_close(None) == -1
    assert L.sdmd_gpu_run(None, 10) == -1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from sdmd_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        _lib.lib()


def test_headers_are_plain_c(tmp_path):
    """include/*.h is the boundary a C99 caller (the reference is C99)
    compiles against: the headers must build as C with -std=c99 -pedantic."""
    import shutil

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use.c"
    src.write_text('#include "sdmd_types.h"\n#include "sdmd_host.h"\n#include "sdmd_gpu.h"\n'
                   '#include "sdmd_remd.h"\n'
                   'int main(void) { sdmd_gpu_opts o = {0}; o.mode = SDMD_MODE_FAST; '
                   'sdmd_remd_temp t = {1, 0.5}; (void)t; return o.mode == 1 ? 0 : 1; }\n')
    r = subprocess.run([gcc, "-std=c99", "-pedantic", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_bench_runs_on_one_gpu_only():
    """bench.py --gpus N with N > 1 exits with a message, before touching a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in r.stderr and "N=2" in r.stderr

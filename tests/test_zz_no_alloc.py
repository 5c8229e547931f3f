"""The allocation-free energy entry point against the allocating one,
through the C ABI: bit-identical kinetic and potential energies on two
calls in a row (the second on a smaller context), in both engine modes."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden_system
from sdmd_amd import _lib

pytestmark = pytest.mark.gpu


def _both(eng):
    L = _lib.lib()
    n = len(eng.systems)
    ke, pe = np.zeros(n), np.zeros(n)
    ke0, pe0 = np.full(n, np.nan), np.full(n, np.nan)
    assert L.sdmd_gpu_energy(eng._h, ke.ctypes.data, pe.ctypes.data) == 0
    assert L.sdmd_gpu_energy_noalloc(eng._h, ke0.ctypes.data, pe0.ctypes.data) == 0
    return ke, pe, ke0, pe0


@pytest.mark.parametrize("mode", ["plain", "fast"])
def test_energy_noalloc_matches_energy(require_gpu, mode):
    from sdmd_amd.engine import GpuEngine

    s1, _ = golden_system("mix4x20")
    s2, _ = golden_system("ala10ext")
    s3, _ = golden_system("cyl8x20")
    kw = dict(max_pops=3000) if mode == "plain" else {}
    for systems in ([s1, s2, s3], [s2]):                # second call: a smaller context
        eng = GpuEngine(systems, block=256, mode=mode, **kw)
        eng.prepare()
        eng.run(2000)
        ke, pe, ke0, pe0 = _both(eng)
        assert np.all(np.isfinite(ke)) and np.all(ke > 0)
        assert np.array_equal(ke, ke0) and np.array_equal(pe, pe0), (mode, len(systems))
        eng.close()

"""flusi_energy_k_noalloc, the entry point that allocates no device memory,
called through the C ABI: on two contexts in a row, the second with a smaller
grid, its value equals flusi_energy_k's on the same state to the bit (the
same kernels in the same order; only where the partial sums live differs),
in both precisions.  The allocation-free call comes first on each context,
before flusi_energy_k has allocated its own partials."""
import ctypes
import os

import pytest

import flusi_amd
from tests._util import GOLDEN, rewrite_ini

pytestmark = pytest.mark.gpu


def _ring(tmp_path, n, name):
    ini = rewrite_ini(os.path.join(GOLDEN, "vortex_ring1_AB2", "vortexring1_AB2.ini"), str(tmp_path), name=name,
                      Resolution__nx=n, Resolution__ny=n, Resolution__nz=n, Time__dt_fixed="1.0e-3",
                      Saving__tintegral=100.0, Saving__tsave=100.0)
    return ini


@pytest.mark.parametrize("prec", [flusi_amd.FP64, flusi_amd.FP32_MIXED])
def test_energy_noalloc_equals_energy_k_twice(tmp_path, prec):
    lib = flusi_amd.load_library()
    values = []
    for n in (64, 32):   # the second call with the smaller input
        print(f"[noalloc] grid {n}^3 precision {prec}", flush=True)
        with flusi_amd.Flusi(_ring(tmp_path, n, f"r{n}.ini"), prec) as s:
            s.init_fields()
            s.advance(2)
            e = ctypes.c_double(-1.0)
            rc = lib.flusi_energy_k_noalloc(s._ctx, ctypes.byref(e))
            assert rc == 0, lib.flusi_last_error(s._ctx)
            ref = s.energy_k()
            print(f"[noalloc] grid {n}^3: {e.value!r} vs energy_k {ref!r}", flush=True)
            assert e.value == ref and ref > 0.0
            # the array keeps the last call's contents; a repeated call agrees
            assert s.energy_k_noalloc() == ref
            values.append(ref)
    assert values[0] != values[1]

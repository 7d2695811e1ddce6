"""The allocation-free standalone solve (vdi_mg_solve_noalloc) against the
allocating one (vdi_mg_create + vdi_mg_solve) through the C ABI: two calls in
a row on device arrays the caller owns, the second on a smaller system, each
bit-identical to the allocating solve (the same V-cycles on the same tiles),
and a system too large for the scratch array rejected with VDI_ERR_ARG.

The caller's device arrays are those of a second standalone system
(vdi_mg_device_arrays): b set from the host, z zeroed, the result read back."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def allocating_solve(vdi, kind, m, n, cx, cy, b):
    g = vdi.MGSystem(kind, m, n, cx, cy, 1)
    g.set_b(b)
    g.set_z(np.zeros(m * n))
    c = g.solve(1e-10, 100)
    z = g.get_z()
    g.close()
    return z, c


def noalloc_solve(vdi, kind, m, n, cx, cy, b):
    """The no-allocation solve on device arrays held by another system."""
    holder = vdi.MGSystem(kind, m, n, cx, cy, 1)
    holder.set_b(b)
    holder.set_z(np.zeros(m * n))
    z_dev, b_dev = holder.device_arrays()
    lib = vdi.load_library()
    cyc = ctypes.c_int(0)
    rc = lib.vdi_mg_solve_noalloc(kind, m, n, cx, cy, 1, b_dev, z_dev, 1e-10, 100, ctypes.byref(cyc))
    z = holder.get_z()
    holder.close()
    return rc, z, cyc.value


@pytest.mark.parametrize("kind,sizes", [(0, [(512, 256), (256, 128)]), (1, [(513, 257), (129, 65)])])
def test_gpu_noalloc_solve_matches_allocating_solve(vdi, kind, sizes):
    rng = np.random.default_rng(41 + kind)
    for m, n in sizes:
        say("noalloc solve", kind, m, n)
        cx, cy = (1.0, 1.3) if kind == 0 else (0.01, 0.013)
        b = rng.standard_normal(m * n)
        z_ref, c_ref = allocating_solve(vdi, kind, m, n, cx, cy, b)
        rc, z, c = noalloc_solve(vdi, kind, m, n, cx, cy, b)
        assert rc == vdi.OK
        assert c == c_ref and c > 0
        np.testing.assert_array_equal(z, z_ref)


def test_gpu_noalloc_rejects_oversized_system(vdi):
    say("noalloc oversized")
    m, n = 2048, 1024
    rc, _z, _c = noalloc_solve(vdi, 0, m, n, 1.0, 1.0, np.ones(m * n))
    assert rc == vdi.ERR_ARG

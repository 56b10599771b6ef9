"""oe_fields_energy_noalloc (the energy of a caller-owned device field block,
computed without allocating) against oe_engine_energy on the same values,
through the C ABI: two calls in a row, the second on a smaller block, each
bit-identical to the engine's sums and within 1e-12 of numpy's float64 sums.
The field blocks and the output pair are the caller's device memory,
allocated here with the HIP runtime the library itself is linked against;
the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from openems_amd import HipEngine, Operator
from openems_amd._ffi import OEError, check, lib
from tests.parity_util import say

pytestmark = pytest.mark.gpu


def _engine_energy(shape, F):
    eng = HipEngine(Operator.uniform(*shape, 1e-3))
    eng.write_fields(F[:3], F[3:])
    return eng.energy()


class _Hip:
    """The HIP runtime liboe.so is linked against (loaded after it, so the
    same instance), for the caller's own device buffers."""

    def __init__(self):
        lib()
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0  # device to host

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _noalloc_energy(F):
    hip = _Hip()
    host = np.ascontiguousarray(F.reshape(6, -1))
    dev = hip.alloc(host.nbytes)
    out = hip.alloc(16)
    try:
        hip.put(dev, host)
        hip.put(out, np.full(2, -1.0))
        check(lib().oe_fields_energy_noalloc(dev, F[0].size, out, 0, None))
        hip.sync()
        got = np.zeros(2)
        hip.get(got, out)
    finally:
        hip.free(dev)
        hip.free(out)
    return tuple(got.tolist())


def test_noalloc_energy_matches_engine_energy():
    rng = np.random.default_rng(3)
    for shape in ((40, 36, 28), (12, 10, 8)):  # the second call is smaller
        F = rng.standard_normal((6,) + shape).astype(np.float32)
        want = _engine_energy(shape, F)
        got = _noalloc_energy(F)
        ref = (float((F[:3].astype(np.float64) ** 2).sum()), float((F[3:].astype(np.float64) ** 2).sum()))
        say(f"[noalloc] {shape}: {got} engine {want} numpy {ref}")
        assert got == want
        assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_noalloc_energy_rejects_bad_sizes():
    hip = _Hip()
    dev = hip.alloc(6 * 4)
    out = hip.alloc(16)
    try:
        for n in (0, (1 << 34) + 1):
            with pytest.raises(OEError) as ei:
                check(lib().oe_fields_energy_noalloc(dev, n, out, 0, None))
            assert ei.value.code == 3  # OE_ERR_GRID
        with pytest.raises(OEError):
            check(lib().oe_fields_energy_noalloc(None, 1, out, 0, None))
    finally:
        hip.free(dev)
        hip.free(out)

"""CPU-side checks of the C ABI library (no GPU compute is invoked).

* liboe.so loads and exports every function include/oe/oe.h declares;
* the host-side waveform sampler (oe_waveform_sample, the counterpart of
  ScheduledExcitation::from_excitation) matches the oracle restatement
  bit-for-bit, including the reference's own known answers;
* argument validation returns the reference's error kinds, and engine
  creation without a HIP device fails loudly (no CPU fallback).
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from oracle import excitation as X
from openems_amd import _ffi
from openems_amd.engine import (EXC_CUSTOM, EXC_DIRAC, EXC_GAUSSIAN_MODULATED, EXC_SINUSOIDAL, EXC_STEP, Operator,
                                Waveform)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oe", "oe.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(oe_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_functions_exported():
    names = header_functions()
    assert len(names) >= 30
    L = C.CDLL(_ffi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers the whole header
    assert sorted(_ffi.SIGNATURES) == names


def test_version_and_error_string():
    L = _ffi.lib()
    assert b"gfx950" in L.oe_version()
    h = C.c_void_p()
    st = L.oe_operator_create(0, 4, 4, (C.c_double * 3)(1e-3, 1e-3, 1e-3), C.byref(h))
    assert st == 3  # OE_ERR_GRID (crate::Error::Grid)
    assert b"empty" in L.oe_last_error()


def test_operator_validation_and_dt():
    op = Operator.uniform(100, 100, 100, 1e-3)
    assert op.dt == pytest.approx(0.99 * 1e-3 / (299792458.0 * math.sqrt(3)), rel=1e-15)
    with pytest.raises(_ffi.OEError) as ei:
        op.set_material(-1.0, 1.0, 0.0, 0.0, (0, 1, 0, 1, 0, 1))
    assert ei.value.code == 4  # MATERIAL
    with pytest.raises(_ffi.OEError):
        op.set_pec((5, 2, 0, 1, 0, 1))
    # mesh-line constructor uses the minimum spacing (src/geometry/mod.rs:113-118)
    L = _ffi.lib()
    x = np.array([0.0, 1e-3, 1.5e-3, 3e-3])
    h = C.c_void_p()
    dp = C.POINTER(C.c_double)
    _ffi.check(L.oe_operator_create_lines(x.ctypes.data_as(dp), 4, x.ctypes.data_as(dp), 4, x.ctypes.data_as(dp), 4,
                                          C.byref(h)))
    dims = (C.c_size_t * 3)()
    L.oe_operator_dims(h, dims)
    assert list(dims) == [3, 3, 3]
    assert L.oe_operator_timestep(h) == pytest.approx(0.99 * 0.5e-3 / (299792458.0 * math.sqrt(3)), rel=1e-15)
    L.oe_operator_destroy(h)


def test_waveform_known_answers():
    # reference unit tests (src/fdtd/excitation.rs:358-418)
    cust = Waveform(EXC_CUSTOM, times=[0, 1, 2, 3], values=[0, 1, 0.5, 0])
    assert cust.evaluate(0.5) == pytest.approx(0.5, abs=1e-12)
    assert cust.evaluate(1.5) == pytest.approx(0.75, abs=1e-12)
    assert cust.evaluate(-1.0) == 0.0 and cust.evaluate(5.0) == 0.0
    assert Waveform(EXC_DIRAC).evaluate(0.0) == 1.0 and Waveform(EXC_DIRAC).evaluate(1e-9) == 0.0
    assert Waveform(EXC_STEP, amplitude=3.0).evaluate(0.0) == 3.0
    g = Waveform.reference_gaussian(1e9, 0.5, amplitude=2.5)
    assert g.evaluate(g.t0) == pytest.approx(2.5, abs=1e-12)
    assert Waveform(EXC_SINUSOIDAL, frequency=1e9).evaluate(0.25e-9) == pytest.approx(1.0, abs=1e-10)


@pytest.mark.parametrize("start", [0, 137])
def test_waveform_sampling_matches_oracle(start):
    dt = 9.532874347655029e-13
    n = 600
    t0, tau = X.openems_gaussian_params(15e9)
    cases = [
        (Waveform.openems_gaussian(15e9), lambda t: X.gaussian(t, t0, tau)),
        (Waveform(EXC_SINUSOIDAL, frequency=3e9, amplitude=0.7), lambda t: X.sinusoidal(t, 3e9, 0.7)),
        (Waveform(EXC_GAUSSIAN_MODULATED, frequency=5e9, t0=2e-10, tau=5e-11),
         lambda t: X.gaussian_modulated(t, 5e9, 2e-10, 5e-11)),
        (Waveform(EXC_DIRAC), lambda t: X.dirac(t)),
        (Waveform(EXC_STEP, amplitude=2.0), lambda t: X.step(t, 2.0)),
    ]
    for w, fn in cases:
        got = w.sample(start, n, dt)
        ref = X.schedule(fn, start, n, dt)
        assert np.array_equal(got, ref)
    times = [0.0, 1e-10, 2e-10, 4e-10]
    vals = [0.0, 1.0, -0.5, 0.25]
    got = Waveform(EXC_CUSTOM, times=times, values=vals).sample(start, n, dt)
    ref = np.array([np.float32(X.custom((start + q) * dt, times, vals)) for q in range(n)])
    assert np.array_equal(got, ref)


def test_engine_creation_fails_loudly_without_device():
    """With no HIP device the engine refuses to exist (OE_ERR_DEVICE); with one
    it is created on the GPU.  There is no CPU path to fall back to."""
    op = Operator.uniform(8, 8, 8, 1e-3)
    L = _ffi.lib()
    h = C.c_void_p()
    st = L.oe_engine_create(op._h, 0, C.byref(h))
    if st == 0:
        L.oe_engine_destroy(h)
        pytest.skip("a HIP device is present")
    assert st == 7, L.oe_last_error()
    assert b"device" in L.oe_last_error().lower()


def test_null_arguments_rejected():
    L = _ffi.lib()
    assert L.oe_engine_run_batch(None, None, None) == 1
    assert L.oe_engine_read_fields(None, None) == 1
    assert L.oe_ext_probe_read(None, None, 0, None) == 1


def test_destructors_skip_the_library_at_interpreter_exit(monkeypatch):
    """A handle still alive at interpreter exit is left to the process's end:
    the destructor must not reach for the library, whose module state may
    already be torn down."""
    from openems_amd import engine as E

    op = Operator.uniform(6, 6, 6, 1e-3)

    def gone():
        raise AssertionError("library called during interpreter exit")

    monkeypatch.setattr(E.sys, "is_finalizing", lambda: True)
    monkeypatch.setattr(E, "lib", gone)
    op.__del__()  # returns without calling lib()
    monkeypatch.undo()
    op.__del__()  # a normal release afterwards
    assert not op._h.value

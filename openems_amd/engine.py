"""Python mirror of the reference's engine-level API over liboe.so.

Names follow the Rust reference so tests read like its own harness
(src/fdtd/engine_testing.rs:401-475):

    op = Operator.uniform(nx, ny, nz, delta)          # Grid::uniform + Operator::new
    op.set_material(eps_r, mu_r, sigma_e, sigma_m, region)
    eng = HipEngine(op)                                # EngineImpl::new
    cpml = eng.cpml(CpmlBoundaries(...))               # device-resident extension
    res = eng.run_batch(EngineBatch(num_steps=..., excitations=[...], extensions=[cpml]))
    E, H = eng.read_fields()

All compute happens in liboe.so on the GPU; this module only marshals
arguments.  Excitation waveforms are sampled by the library's C++ host code
(oe_waveform_sample, the counterpart of ScheduledExcitation::from_excitation).
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _ffi
from ._ffi import check, lib

MODE_COMPAT = 0
MODE_PHYSICAL = 1
ADE_LORENTZ, ADE_DRUDE, ADE_DEBYE = 0, 1, 2

EXC_GAUSSIAN, EXC_SINUSOIDAL, EXC_GAUSSIAN_MODULATED, EXC_DIRAC, EXC_STEP, EXC_CUSTOM = range(6)


# ------------------------------------------------------------------ operator
class Operator:
    """Operator (src/fdtd/operator.rs:40-271): grid, dt and material boxes."""

    def __init__(self, handle: int, shape: Tuple[int, int, int], delta: Tuple[float, float, float]):
        self._h = C.c_void_p(handle)
        self.shape = shape
        self.delta = delta

    @classmethod
    def uniform(cls, nx: int, ny: int, nz: int, delta) -> "Operator":
        if np.isscalar(delta):
            delta = (float(delta),) * 3
        d = (C.c_double * 3)(*delta)
        h = C.c_void_p()
        check(lib().oe_operator_create(nx, ny, nz, d, C.byref(h)))
        return cls(h.value, (nx, ny, nz), tuple(map(float, delta)))

    @classmethod
    def from_lines(cls, x, y, z) -> "Operator":
        """Grid::cartesian + Operator::new (src/geometry/mod.rs:69-71,
        src/fdtd/operator.rs:59-77): mesh lines in metres, cells = lines - 1
        per axis; dt and the coefficients use the smallest spacing per axis."""
        ls = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (x, y, z)]
        dp = C.POINTER(C.c_double)
        h = C.c_void_p()
        check(lib().oe_operator_create_lines(ls[0].ctypes.data_as(dp), ls[0].size, ls[1].ctypes.data_as(dp),
                                             ls[1].size, ls[2].ctypes.data_as(dp), ls[2].size, C.byref(h)))
        delta = tuple(float(np.abs(np.diff(l)).min()) if l.size > 1 else 0.0 for l in ls)
        return cls(h.value, tuple(max(l.size - 1, 1) for l in ls), delta)

    def __del__(self):
        # at interpreter exit the library's module state may already be
        # gone; the process releases the device memory then
        if sys.is_finalizing():
            return
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_operator_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def dt(self) -> float:
        return lib().oe_operator_timestep(self._h)

    def set_material(self, eps_r, mu_r, sigma_e, sigma_m, region):
        r = (C.c_size_t * 6)(*map(int, region))
        check(lib().oe_operator_set_material(self._h, eps_r, mu_r, sigma_e, sigma_m, r))

    def set_pec(self, region):
        r = (C.c_size_t * 6)(*map(int, region))
        check(lib().oe_operator_set_pec(self._h, r))

    def add_lumped_element(self, direction: int, region, R: float):
        """openEMS LumpedElement resistor on the `direction` edges of region."""
        r = (C.c_size_t * 6)(*map(int, region))
        check(lib().oe_operator_add_lumped_element(self._h, int(direction), r, float(R)))


# -------------------------------------------------------------- waveforms
@dataclass
class Waveform:
    """ExcitationType + amplitude (src/fdtd/excitation.rs:9-57)."""

    type: int
    t0: float = 0.0
    tau: float = 0.0
    frequency: float = 0.0
    times: Optional[Sequence[float]] = None
    values: Optional[Sequence[float]] = None
    amplitude: float = 1.0

    @staticmethod
    def gaussian(t0: float, tau: float, amplitude: float = 1.0) -> "Waveform":
        return Waveform(EXC_GAUSSIAN, t0=t0, tau=tau, amplitude=amplitude)

    @staticmethod
    def openems_gaussian(fc: float, amplitude: float = 1.0) -> "Waveform":
        """openEMS Gaussian with f0 = 0 and cutoff fc, written as the reference's
        Gaussian{t0, tau}: tau = 3/(2 pi fc), t0 = 9/(2 pi fc)
        (matlab/examples/other/gauss_excitation_test.m:14-21)."""
        return Waveform(EXC_GAUSSIAN, t0=9.0 / (2 * math.pi * fc), tau=3.0 / (2 * math.pi * fc), amplitude=amplitude)

    @staticmethod
    def reference_gaussian(center_freq: float, bandwidth: float, amplitude: float = 1.0) -> "Waveform":
        """Excitation::gaussian (src/fdtd/excitation.rs:67-85)."""
        f_max = center_freq * (1.0 + 0.5 * bandwidth)
        tau = 1.0 / (math.pi * f_max)
        return Waveform(EXC_GAUSSIAN, t0=4.0 * tau, tau=tau, amplitude=amplitude)

    def _c(self):
        w = _ffi.Waveform()
        w.type = self.type
        w.t0, w.tau, w.frequency, w.amplitude = self.t0, self.tau, self.frequency, self.amplitude
        keep = []
        if self.type == EXC_CUSTOM:
            t = np.ascontiguousarray(self.times, dtype=np.float64)
            v = np.ascontiguousarray(self.values, dtype=np.float64)
            keep = [t, v]
            w.times = t.ctypes.data_as(C.POINTER(C.c_double))
            w.values = v.ctypes.data_as(C.POINTER(C.c_double))
            w.n_points = len(t)
        return w, keep

    def evaluate(self, t: float) -> float:
        w, keep = self._c()
        out = C.c_double()
        check(lib().oe_waveform_evaluate(C.byref(w), float(t), C.byref(out)))
        return out.value

    def sample(self, start_timestep: int, n: int, dt: float) -> np.ndarray:
        w, keep = self._c()
        out = np.empty(n, np.float32)
        check(lib().oe_waveform_sample(C.byref(w), start_timestep, n, dt, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out


@dataclass
class ScheduledExcitation:
    """ScheduledExcitation with a Sampled waveform (src/fdtd/batch.rs:71-100)."""

    position: Tuple[int, int, int]
    direction: int
    samples: np.ndarray
    soft_source: bool = True


# ------------------------------------------------------------- extensions
@dataclass
class CpmlBoundaries:
    """CpmlBoundaries (src/extensions/cpml.rs:27-51) + the mode switch."""

    layers: Sequence[int] = (8, 8, 8, 8, 8, 8)
    grading_order: float = 3.0
    reflection_coeff: float = 1e-6
    alpha_max: float = 0.0
    kappa_max: float = 1.0
    mode: int = MODE_COMPAT


class Extension:
    def __init__(self, eng: "HipEngine", handle: int, kind: str):
        self._eng = eng
        self._h = C.c_void_p(handle)
        self.kind = kind

    def __del__(self):
        # at interpreter exit the library's module state may already be
        # gone; the process releases the device memory then
        if sys.is_finalizing():
            return
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_extension_destroy(self._h)
            self._h = C.c_void_p()

    def reset(self):
        check(lib().oe_extension_reset(self._h))

    # CPML
    def psi(self, face: int, field: str, which: str) -> Optional[np.ndarray]:
        f = 0 if field == "E" else 1
        w = 0 if which == "pos" else 1
        n = C.c_size_t()
        check(lib().oe_ext_cpml_read_psi(self._h, face, f, w, None, 0, C.byref(n)))
        if n.value == 0:
            return None
        out = np.empty(n.value, np.float32)
        check(lib().oe_ext_cpml_read_psi(self._h, face, f, w, out.ctypes.data_as(C.POINTER(C.c_float)), n.value, C.byref(n)))
        return out

    # UPML
    UPML_ARRAYS = ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn", "volt_flux", "curr_flux")

    def upml_array(self, region: int, name: str, component: int) -> np.ndarray:
        w = self.UPML_ARRAYS.index(name)
        n = C.c_size_t()
        st = (C.c_size_t * 3)()
        sz = (C.c_size_t * 3)()
        check(lib().oe_ext_upml_read(self._h, region, w, component, None, 0, C.byref(n), st, sz))
        out = np.empty(n.value, np.float32)
        check(lib().oe_ext_upml_read(self._h, region, w, component, out.ctypes.data_as(C.POINTER(C.c_float)),
                                     n.value, C.byref(n), st, sz))
        return out.reshape(tuple(sz))

    # ADE
    def state(self) -> np.ndarray:
        n = C.c_size_t()
        check(lib().oe_ext_ade_read_state(self._h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.float32)
        check(lib().oe_ext_ade_read_state(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n.value, C.byref(n)))
        return out.reshape(3, -1)

    # DFT / lumped port
    def dft_values(self) -> np.ndarray:
        """[n_freqs, n_signals] complex: 2 dt sum x exp(-j 2 pi f t)."""
        nf, ns = self._dft_shape
        re = np.empty(nf * ns)
        im = np.empty(nf * ns)
        n = C.c_uint64()
        check(lib().oe_ext_dft_read(self._h, re.ctypes.data_as(C.POINTER(C.c_double)),
                                    im.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
        return (re + 1j * im).reshape(nf, ns)

    def series(self) -> np.ndarray:
        nf, ns = self._dft_shape
        n = C.c_size_t()
        check(lib().oe_ext_dft_read_series(self._h, None, 0, C.byref(n)))
        out = np.empty((n.value, ns), np.float64)
        if n.value:
            check(lib().oe_ext_dft_read_series(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return out

    def s11(self, z_ref: float) -> np.ndarray:
        nf, _ = self._dft_shape
        re = np.empty(nf)
        im = np.empty(nf)
        check(lib().oe_ext_port_s11(self._h, float(z_ref), re.ctypes.data_as(C.POINTER(C.c_double)),
                                    im.ctypes.data_as(C.POINTER(C.c_double))))
        return re + 1j * im

    # NF2FF box
    def far_field(self, freq_index: int, theta, phi, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> "FarFieldResult":
        """Nf2ff::calculate on the recorded surface phasors of freqs[freq_index]."""
        return _far_field(lambda th, nt, ph, np_, r, c, out: lib().oe_ext_nf2ff_calculate(
            self._h, int(freq_index), th, nt, ph, np_, r, c, out), self._freqs[freq_index], theta, phi, radius, center)

    # host callback
    def stop_reason(self) -> str:
        """The reason the last check_termination gave for stopping ("" if none)."""
        n = C.c_size_t()
        check(lib().oe_ext_host_callback_reason(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().oe_ext_host_callback_reason(self._h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode(errors="replace")

    # probes
    def data(self) -> np.ndarray:
        n = C.c_size_t()
        check(lib().oe_ext_probe_read(self._h, None, 0, C.byref(n)))
        out = np.empty((n.value, 6), np.float64)
        if n.value:
            check(lib().oe_ext_probe_read(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return out


@dataclass
class FarFieldResult:
    """FarFieldResult (src/nf2ff/mod.rs:11-24) plus the radiated power, the
    maximum directivity and the device timing of the transform."""
    theta: np.ndarray
    phi: np.ndarray
    e_theta: np.ndarray  # [n_theta, n_phi] complex
    e_phi: np.ndarray
    frequency: float
    prad: float
    dmax: float
    gemm_ms: float
    total_ms: float
    gemm_flop: float
    gemm_split_ms: float = 0.0     # the contractions run on the FP16 matrix cores (split operands)
    gemm_split_flop: float = 0.0

    def directivity(self) -> np.ndarray:
        """|E|^2 normalised to its maximum (FarFieldResult::directivity, src/nf2ff/mod.rs:28-55)."""
        p = np.abs(self.e_theta) ** 2 + np.abs(self.e_phi) ** 2
        m = p.max()
        return p / m if m > 0 else np.zeros_like(p)

    def max_directivity(self) -> float:
        return self.dmax


def _far_field(call, frequency, theta, phi, radius, center, split: bool = False) -> FarFieldResult:
    """The C call writes E_theta and E_phi as interleaved complex straight
    into the result arrays; split=True asks for the separate real and
    imaginary arrays instead (the other output form of the C-ABI)."""
    th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    ph = np.ascontiguousarray(phi, dtype=np.float64).ravel()
    shape = (th.size, ph.size)
    dp = C.POINTER(C.c_double)
    out = _ffi.FarField()
    if split:
        bufs = [np.empty(shape, np.float64) for _ in range(4)]
        out.e_theta_re, out.e_theta_im, out.e_phi_re, out.e_phi_im = (b.ctypes.data_as(dp) for b in bufs)
    else:
        et = np.empty(shape, np.complex128)
        ep = np.empty(shape, np.complex128)
        out.e_theta = et.view(np.float64).ctypes.data_as(dp)
        out.e_phi = ep.view(np.float64).ctypes.data_as(dp)
    c = (C.c_double * 3)(*map(float, center))
    check(call(th.ctypes.data_as(dp), th.size, ph.ctypes.data_as(dp), ph.size, float(radius), c, C.byref(out)))
    if split:
        et, ep = bufs[0] + 1j * bufs[1], bufs[2] + 1j * bufs[3]
    return FarFieldResult(th, ph, et, ep, float(frequency), out.prad, out.dmax, out.gemm_ms, out.total_ms,
                          out.gemm_flop, out.gemm_split_ms, out.gemm_split_flop)


class Nf2ff:
    """Nf2ff (src/nf2ff/mod.rs:72-154) on a HIP device: surfaces of tangential
    E/H phasors in, far field out."""

    def __init__(self, frequency: float, device: int = 0):
        h = C.c_void_p()
        check(lib().oe_nf2ff_create(int(device), float(frequency), C.byref(h)))
        self._h = h
        self.frequency = float(frequency)

    def __del__(self):
        # at interpreter exit the library's module state may already be
        # gone; the process releases the device memory then
        if sys.is_finalizing():
            return
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_nf2ff_destroy(self._h)
            self._h = C.c_void_p()

    def add_surface(self, normal_dir: int, sign: int, position: float, c1, c2, e_t, h_t):
        """e_t, h_t: complex [2, len(c1), len(c2)], the phasors along the
        in-plane axes (normal_dir+1)%3 and (normal_dir+2)%3."""
        c1 = np.ascontiguousarray(c1, dtype=np.float64)
        c2 = np.ascontiguousarray(c2, dtype=np.float64)
        shape = (2, c1.size, c2.size)
        e = np.ascontiguousarray(np.broadcast_to(e_t, shape), dtype=np.complex128)
        hh = np.ascontiguousarray(np.broadcast_to(h_t, shape), dtype=np.complex128)
        dp = C.POINTER(C.c_double)
        check(lib().oe_nf2ff_add_surface(self._h, int(normal_dir), int(sign), float(position), c1.ctypes.data_as(dp),
                                         c1.size, c2.ctypes.data_as(dp), c2.size,
                                         e.view(np.float64).ctypes.data_as(dp), hh.view(np.float64).ctypes.data_as(dp)))

    def calculate(self, theta, phi, radius: float = 1.0, center=(0.0, 0.0, 0.0), split: bool = False) -> FarFieldResult:
        return _far_field(lambda th, nt, ph, np_, r, c, out: lib().oe_nf2ff_calculate(self._h, th, nt, ph, np_, r, c, out),
                          self.frequency, theta, phi, radius, center, split)


@dataclass
class SweepResult:
    timesteps_executed: int
    termination_reason: int
    final_decay_db: float
    peak_energy: float
    final_energy: float
    elapsed_seconds: float


class Sweep:
    """Many structures stepped together (oe_sweep_*): each is a HipEngine with
    its CPML, lumped-port DFT and excitations; all half-steps of all active
    structures run as one launch per kernel variant, replayed from a captured
    HIP graph between energy checks."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().oe_sweep_create(int(device), C.byref(h)))
        self._h = h
        self._keep = []

    def __del__(self):
        # at interpreter exit the library's module state may already be
        # gone; the process releases the device memory then
        if sys.is_finalizing():
            return
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_sweep_destroy(self._h)
            self._h = C.c_void_p()

    def add(self, eng: "HipEngine", cpml: Optional[Extension] = None, port: Optional[Extension] = None,
            excitations: Sequence[ScheduledExcitation] = ()) -> int:
        n = len(excitations)
        ex = (_ffi.Excitation * max(n, 1))()
        arrays = []
        for q, x in enumerate(excitations):
            a = np.ascontiguousarray(x.samples, dtype=np.float32)
            arrays.append(a)
            i, j, k = x.position
            ex[q] = _ffi.Excitation(i, j, k, x.direction, int(bool(x.soft_source)),
                                    a.ctypes.data_as(C.POINTER(C.c_float)), a.size)
        idx = C.c_size_t()
        check(lib().oe_sweep_add(self._h, eng._h, cpml._h if cpml else None, port._h if port else None, ex, n,
                                 C.byref(idx)))
        self._keep.append((eng, cpml, port))
        return idx.value

    def run(self, max_steps: int, energy_decay_db: Optional[float] = None, check_interval: int = 100,
            use_graph: bool = True, autotune: bool = True) -> List[SweepResult]:
        cfg = _ffi.SweepConfig(int(max_steps), int(energy_decay_db is not None), float(energy_decay_db or 0.0),
                               int(check_interval), int(use_graph), int(autotune))
        out = (_ffi.SweepEntry * len(self._keep))()
        check(lib().oe_sweep_run(self._h, C.byref(cfg), out))
        return [SweepResult(e.timesteps_executed, e.termination_reason, e.final_decay_db, e.peak_energy,
                            e.final_energy, e.elapsed_seconds) for e in out]

    def tuning(self):
        ci = C.c_int()
        ms = (C.c_double * 4)()
        cand = (C.c_int * 4)()
        check(lib().oe_sweep_tuning(self._h, C.byref(ci), ms, cand))
        return {"ci": ci.value, "candidates": list(cand), "ms_per_step": list(ms)}

    def set_ci(self, ci: int):
        check(lib().oe_sweep_set_ci(self._h, int(ci)))

    def energy_history(self, index: int):
        """(timesteps, energies) of structure `index` at every check of the last run."""
        n = C.c_size_t()
        check(lib().oe_sweep_energy_history(self._h, int(index), None, None, 0, C.byref(n)))
        ts = np.zeros(n.value, np.uint64)
        en = np.zeros(n.value, np.float64)
        if n.value:
            check(lib().oe_sweep_energy_history(self._h, int(index), ts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                en.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return ts.astype(np.int64), en


def port_excitations(direction: int, region, samples: np.ndarray, sign: int = 1, amplitude: float = 1.0):
    """The excitation box of a lumped port: a soft source of
    -sign * amplitude * samples on every `direction` edge of region
    (python/openEMS/ports.py:183-188)."""
    v = (-sign * amplitude * np.asarray(samples, np.float64)).astype(np.float32)
    out = []
    for i in range(region[0], region[1]):
        for j in range(region[2], region[3]):
            for k in range(region[4], region[5]):
                out.append(ScheduledExcitation((i, j, k), direction, v, True))
    return out


@dataclass
class EngineBatch:
    """EngineBatch + TerminationConfig + EnergyMonitorConfig (src/fdtd/batch.rs)."""

    num_steps: Optional[int] = None
    excitations: List[ScheduledExcitation] = field(default_factory=list)
    extensions: List[Extension] = field(default_factory=list)
    max_timesteps: Optional[int] = None
    energy_decay_db: Optional[float] = None
    check_interval: int = 100
    sample_interval: int = 0
    track_peak: bool = False
    persistent_peak: bool = False
    energy_capacity: int = 100000


@dataclass
class BatchResult:
    timesteps_executed: int
    termination_reason: int
    final_decay_db: float
    energy_samples: np.ndarray  # [n, 3] timestep, e, h
    elapsed_seconds: float


# ----------------------------------------------------------------- engine
class HipEngine:
    """EngineImpl on one MI355X (src/fdtd/engine_impl.rs:24-71)."""

    def __init__(self, op: Operator, device: int = 0):
        h = C.c_void_p()
        check(lib().oe_engine_create(op._h, device, C.byref(h)))
        self._h = h
        self.shape = op.shape
        self.N = int(np.prod(op.shape))
        self.dt = op.dt
        self.delta = op.delta

    @classmethod
    def from_arrays(cls, shape, delta, dt: float, coeff: np.ndarray, device: int = 0) -> "HipEngine":
        """EngineImpl::new on an Operator's cloned arrays (src/fdtd/basic_engine.rs:42-54):
        coeff is [12, nx, ny, nz] float32 (Ca, Cb, Da, Db of x, y, z; Db negative)."""
        nx, ny, nz = map(int, shape)
        a = np.ascontiguousarray(coeff, dtype=np.float32).reshape(12, nx * ny * nz)
        ptrs = (C.POINTER(C.c_float) * 12)(*[a[w].ctypes.data_as(C.POINTER(C.c_float)) for w in range(12)])
        d = (C.c_double * 3)(*([float(delta)] * 3 if np.isscalar(delta) else map(float, delta)))
        h = C.c_void_p()
        check(lib().oe_engine_create_from_arrays(nx, ny, nz, d, float(dt), ptrs, device, C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.shape = (nx, ny, nz)
        self.N = nx * ny * nz
        self.dt = float(dt)
        self.delta = tuple(d)
        return self

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if not sys.is_finalizing():  # see Operator.__del__
            self.close()

    @property
    def timestep(self) -> int:
        return lib().oe_engine_timestep(self._h)

    def read_fields(self) -> Tuple[np.ndarray, np.ndarray]:
        out = np.empty((6,) + tuple(self.shape), np.float32)
        ptrs = (C.POINTER(C.c_float) * 6)(*[out[d].ctypes.data_as(C.POINTER(C.c_float)) for d in range(6)])
        check(lib().oe_engine_read_fields(self._h, ptrs))
        return out[:3], out[3:]

    def write_fields(self, E: Optional[np.ndarray] = None, H: Optional[np.ndarray] = None):
        arrs = [None] * 6
        keep = []
        for base, F in ((0, E), (3, H)):
            if F is None:
                continue
            for d in range(3):
                a = np.ascontiguousarray(F[d], dtype=np.float32)
                keep.append(a)
                arrs[base + d] = a.ctypes.data_as(C.POINTER(C.c_float))
        ptrs = (C.POINTER(C.c_float) * 6)(*[p if p is not None else C.POINTER(C.c_float)() for p in arrs])
        check(lib().oe_engine_write_fields(self._h, ptrs))

    def coefficients(self) -> np.ndarray:
        out = np.empty((12,) + tuple(self.shape), np.float32)
        for w in range(12):
            check(lib().oe_engine_read_coefficients(self._h, w, out[w].ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def reset(self):
        check(lib().oe_engine_reset(self._h))

    def energy(self) -> Tuple[float, float]:
        e, h = C.c_double(), C.c_double()
        check(lib().oe_engine_energy(self._h, C.byref(e), C.byref(h)))
        return e.value, h.value

    def set_profiling(self, period: int):
        """0 off; n > 0 times every launch of every n-th step."""
        check(lib().oe_engine_set_profiling(self._h, int(period)))

    def kernel_time(self, name: str) -> Tuple[int, float]:
        n, ms = C.c_uint64(), C.c_double()
        check(lib().oe_engine_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def material_table(self, enable: int = -1) -> Tuple[int, int]:
        """Material-table rows of the (E, H) half-steps; 0 for a half that
        reads the 12 coefficient arrays.  enable 1 builds, 0 drops, -1 queries."""
        r = (C.c_int * 2)()
        check(lib().oe_engine_material_table(self._h, int(enable), r))
        return r[0], r[1]

    def graph_steps(self, steps: int = -1) -> int:
        """Steps per captured HIP graph (0: launch every step directly;
        default 16); -1 only queries."""
        cur = C.c_int()
        check(lib().oe_engine_graph_steps(self._h, int(steps), C.byref(cur)))
        return cur.value

    def fused_steps(self, mode: int = -1) -> int:
        """0: two in-place half-steps per timestep (default); 1: one fused
        pass per timestep where the batch allows it (nz from 128 to 256);
        2: also for nz below 128; -1 only queries."""
        cur = C.c_int()
        check(lib().oe_engine_fused_steps(self._h, int(mode), C.byref(cur)))
        return cur.value

    def set_tiles(self, bx=0, by=0, ci=0, vec=0):
        check(lib().oe_engine_set_tiles(self._h, bx, by, ci, vec))

    def tiles(self):
        t = (C.c_int * 4)()
        check(lib().oe_engine_get_tiles(self._h, t))
        return dict(bx=t[0], by=t[1], ci=t[2], vec=t[3])

    # extensions
    def cpml(self, cfg: CpmlBoundaries) -> Extension:
        c = _ffi.CpmlConfig()
        for q in range(6):
            c.layers[q] = int(cfg.layers[q])
        c.grading_order, c.reflection_coeff = cfg.grading_order, cfg.reflection_coeff
        c.alpha_max, c.kappa_max, c.mode = cfg.alpha_max, cfg.kappa_max, cfg.mode
        h = C.c_void_p()
        check(lib().oe_ext_cpml_create(self._h, C.byref(c), C.byref(h)))
        return Extension(self, h.value, "cpml")

    def upml(self, layers: Sequence[int] = (8,) * 6, grading_order: float = 3.0,
             reflection_coeff: float = 1e-6) -> Extension:
        """Upml with PmlBoundaries (src/extensions/pml.rs), the reference's formulation."""
        arr = (C.c_size_t * 6)(*map(int, layers))
        h = C.c_void_p()
        check(lib().oe_ext_upml_create(self._h, arr, float(grading_order), float(reflection_coeff), C.byref(h)))
        return Extension(self, h.value, "upml")

    def mur(self, enabled: Sequence[bool] = (True,) * 6, phase_velocity: float = 0.0) -> Extension:
        en = (C.c_int * 6)(*[int(bool(x)) for x in enabled])
        h = C.c_void_p()
        check(lib().oe_ext_mur_create(self._h, en, phase_velocity, C.byref(h)))
        return Extension(self, h.value, "mur")

    def ade(self, kind: int, params: Sequence[float], positions, mode: int = MODE_COMPAT) -> Extension:
        p = np.ascontiguousarray(np.asarray(positions, dtype=np.uint32).reshape(-1, 3))
        prm = (C.c_double * 3)(*(list(params) + [0.0] * (3 - len(params))))
        h = C.c_void_p()
        check(lib().oe_ext_ade_create(self._h, kind, prm, p.ctypes.data_as(C.POINTER(C.c_uint32)), p.shape[0], mode,
                                      C.byref(h)))
        return Extension(self, h.value, "ade")

    def dft(self, terms, freqs, precision: int = 0, record_steps: int = 0, time_stride: int = 1) -> Extension:
        """Running DFT of signals; terms: per signal a list of (component 0..5, i, j, k, weight)."""
        nx, ny, nz = self.shape
        off = np.zeros(len(terms) + 1, np.uint64)
        comp, cell, w = [], [], []
        for m, tl in enumerate(terms):
            for (c, i, j, k, wt) in tl:
                comp.append(c)
                cell.append((i * ny + j) * nz + k)
                w.append(wt)
            off[m + 1] = len(comp)
        comp = np.asarray(comp, np.uint32)
        cell = np.asarray(cell, np.uint64)
        w = np.asarray(w, np.float64)
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        cfg = _ffi.DftConfig(len(terms), off.ctypes.data_as(C.POINTER(C.c_uint64)),
                             comp.ctypes.data_as(C.POINTER(C.c_uint32)), cell.ctypes.data_as(C.POINTER(C.c_uint64)),
                             w.ctypes.data_as(C.POINTER(C.c_double)), fr.ctypes.data_as(C.POINTER(C.c_double)),
                             len(fr), int(precision), int(record_steps), int(time_stride))
        h = C.c_void_p()
        check(lib().oe_ext_dft_create(self._h, C.byref(cfg), C.byref(h)))
        x = Extension(self, h.value, "dft")
        x._dft_shape = (len(fr), len(terms))
        return x

    def lumped_port(self, direction: int, region, freqs, sign: int = 1, precision: int = 0,
                    record_steps: int = 0) -> Extension:
        """U/I probes and their running DFT for a lumped port (openEMS LumpedPort)."""
        r = (C.c_size_t * 6)(*map(int, region))
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        h = C.c_void_p()
        check(lib().oe_ext_lumped_port_create(self._h, int(direction), r, int(sign),
                                              fr.ctypes.data_as(C.POINTER(C.c_double)), len(fr), int(precision),
                                              int(record_steps), C.byref(h)))
        x = Extension(self, h.value, "port")
        x._dft_shape = (len(fr), 2)
        return x

    def nf2ff_box(self, box, freqs, directions: int = 0, precision: int = 0, time_stride: int = 1) -> Extension:
        """NF2FF recording box through the mesh nodes box = (i0, i1, j0, j1, k0, k1),
        inclusive (openEMS CreateNF2FFBox with frequency-domain recording)."""
        b = (C.c_size_t * 6)(*map(int, box))
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        h = C.c_void_p()
        check(lib().oe_ext_nf2ff_create(self._h, b, int(directions), fr.ctypes.data_as(C.POINTER(C.c_double)), len(fr),
                                        int(precision), int(time_stride), C.byref(h)))
        x = Extension(self, h.value, "nf2ff")
        lo, hi = np.array(box[0::2]), np.array(box[1::2])
        nodes = sum(int(np.prod([(hi[a] - lo[a] + 1) for a in ((ax + 1) % 3, (ax + 2) % 3)]))
                    for ax in range(3) for side in range(2) if (directions or 0x3F) >> (2 * ax + side) & 1)
        x._dft_shape = (len(fr), 4 * nodes)
        x._freqs = fr
        return x

    HOOKS = ("pre_update_h", "post_update_h", "pre_update_e", "post_update_e", "pre_batch", "post_batch")

    def host_callback(self, hook=None, hooks: Sequence = (), check_termination=None) -> Extension:
        """An Extension without a device form, run on the host
        (src/extensions/extension_trait.rs:32-166): hook(which, step, fields)
        is called at the hook points named in `hooks` (names of HOOKS or their
        indices) with the six field arrays, each shaped like the grid, to change
        in place (H at the H hooks, E at the E hooks, all at the batch hooks);
        it may return a nonzero oe_status to fail the batch.
        check_termination() returns None or a reason string, as the trait's
        check_termination does.  An exception in either ends the batch and is
        raised again by run_batch."""
        shape, n = tuple(self.shape), self.N
        mask = 0
        for h in hooks:
            mask |= 1 << (self.HOOKS.index(h) if isinstance(h, str) else int(h))
        x_ref = {}

        def _hook(user, which, step, fields):
            try:
                views = [np.ctypeslib.as_array(fields[d], (n,)).reshape(shape) for d in range(6)]
                return int(hook(int(which), int(step), views) or 0)
            except BaseException as exc:  # raised again by run_batch
                x_ref["x"]._host_error = exc
                return 5

        def _check(user, reason, cap):
            try:
                why = check_termination()
            except BaseException as exc:
                x_ref["x"]._host_error = exc
                why = "exception in check_termination"
            if why is None:
                return 0
            b = str(why).encode()[: max(int(cap) - 1, 0)]
            C.memmove(reason, b, len(b))
            return 1

        cfg = _ffi.HostCallbackConfig()
        cfg.hook = _ffi.HOST_HOOK_FN(_hook) if hook else _ffi.HOST_HOOK_FN()
        cfg.check_termination = _ffi.HOST_CHECK_FN(_check) if check_termination else _ffi.HOST_CHECK_FN()
        cfg.hook_mask = mask
        h = C.c_void_p()
        check(lib().oe_ext_host_callback_create(self._h, C.byref(cfg), C.byref(h)))
        x = Extension(self, h.value, "host")
        x._host_keep = (cfg.hook, cfg.check_termination)
        x._host_error = None
        x_ref["x"] = x
        return x

    def field_probe(self, pos) -> Extension:
        h = C.c_void_p()
        check(lib().oe_ext_field_probe_create(self._h, *map(int, pos), C.byref(h)))
        return Extension(self, h.value, "probe")

    def run_batch(self, batch: EngineBatch) -> BatchResult:
        n_exc = len(batch.excitations)
        ex = (_ffi.Excitation * max(n_exc, 1))()
        keep = []
        for q, s in enumerate(batch.excitations):
            a = np.ascontiguousarray(s.samples, dtype=np.float32)
            keep.append(a)
            i, j, k = s.position
            ex[q] = _ffi.Excitation(i, j, k, s.direction, int(bool(s.soft_source)),
                                    a.ctypes.data_as(C.POINTER(C.c_float)), a.size)
        n_ext = len(batch.extensions)
        exts = (C.c_void_p * max(n_ext, 1))(*[x._h.value for x in batch.extensions])
        # room for the samples this batch can produce: one per sample_interval
        # steps (allocating the full default capacity, 3.2 MB, costs tens of
        # milliseconds when the pages are fresh)
        cap = batch.energy_capacity
        if batch.num_steps is not None:
            cap = min(cap, batch.num_steps // batch.sample_interval + 1 if batch.sample_interval > 0 else 0)
        samples = (_ffi.EnergySample * max(cap, 1))()
        d = _ffi.BatchDesc()
        d.has_num_steps = int(batch.num_steps is not None)
        d.num_steps = int(batch.num_steps or 0)
        d.excitations = ex
        d.n_excitations = n_exc
        d.extensions = exts
        d.n_extensions = n_ext
        d.has_max_timesteps = int(batch.max_timesteps is not None)
        d.max_timesteps = int(batch.max_timesteps or 0)
        d.has_energy_decay_db = int(batch.energy_decay_db is not None)
        d.energy_decay_db = float(batch.energy_decay_db or 0.0)
        d.check_interval = batch.check_interval
        d.sample_interval = batch.sample_interval
        d.track_peak = int(batch.track_peak)
        d.persistent_peak = int(batch.persistent_peak)
        d.energy_samples = samples
        d.energy_samples_capacity = cap
        r = _ffi.BatchResultC()
        for x in batch.extensions:
            if x.kind == "host":
                x._host_error = None
        status = lib().oe_engine_run_batch(self._h, C.byref(d), C.byref(r))
        for x in batch.extensions:
            err = getattr(x, "_host_error", None)
            if err is not None:
                x._host_error = None
                raise err
        check(status)
        n = min(r.n_energy_samples, cap)
        es = np.array([(samples[q].timestep, samples[q].e_energy, samples[q].h_energy) for q in range(n)],
                      dtype=np.float64).reshape(-1, 3)
        return BatchResult(r.timesteps_executed, r.termination_reason, r.final_decay_db, es, r.elapsed_seconds)

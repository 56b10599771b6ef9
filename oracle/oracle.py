"""ctypes front-end of the CPU oracle (liboe_oracle_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.

The classes mirror the reference's test harness
(src/fdtd/engine_testing.rs:401-475): an operator is built from a uniform grid,
materials are applied in order, extensions are attached, and run_batch()
executes the per-step loop of BasicEngine::run_batch
(src/fdtd/basic_engine.rs:56-163).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

MODE_COMPAT = 0
MODE_PHYSICAL = 1
ADE_LORENTZ, ADE_DRUDE, ADE_DEBYE = 0, 1, 2


def _host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd"))
    except OSError:
        pass
    return False


def library_path() -> str:
    name = "liboe_oracle_v4.so" if _host_has_avx512() else "liboe_oracle_v3.so"
    return os.path.join(_HERE, name)


class _Exc(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32),
                ("dir", C.c_uint32), ("soft", C.c_uint32)]


class _Batch(C.Structure):
    _fields_ = [
        ("num_steps", C.c_uint64), ("max_timesteps", C.c_uint64),
        ("exc", C.POINTER(_Exc)), ("n_exc", C.c_size_t),
        ("samples", C.POINTER(C.c_double)), ("n_samples", C.c_size_t),
        ("check_interval", C.c_uint64), ("energy_decay_db", C.c_double),
        ("sample_interval", C.c_uint64), ("track_peak", C.c_int),
        ("energy_out", C.POINTER(C.c_double)), ("energy_cap", C.c_size_t),
        ("n_energy", C.c_size_t), ("steps_executed", C.c_uint64),
        ("reason", C.c_int), ("final_decay_db", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        vp, sz, d, u32, u64 = C.c_void_p, C.c_size_t, C.c_double, C.c_uint32, C.c_uint64
        dp = C.POINTER(C.c_double)
        L.oro_create.restype = vp
        L.oro_create.argtypes = [C.c_int, sz, sz, sz, d, d, d]
        L.oro_destroy.argtypes = [vp]
        L.oro_set_threads.argtypes = [C.c_int]
        L.oro_max_threads.restype = C.c_int
        L.oro_dt.restype = d
        L.oro_dt.argtypes = [vp]
        L.oro_timestep.restype = u64
        L.oro_timestep.argtypes = [vp]
        L.oro_set_material.argtypes = [vp, d, d, d, d, sz, sz, sz, sz, sz, sz]
        L.oro_set_pec.argtypes = [vp, sz, sz, sz, sz, sz, sz]
        L.oro_get_coeff.argtypes = [vp, C.c_int, dp]
        L.oro_get_coeff_f32.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
        L.oro_set_coeff_f32.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
        L.oro_get_fields.argtypes = [vp, dp]
        L.oro_get_fields_f32.argtypes = [vp, C.POINTER(C.c_float)]
        L.oro_set_fields.argtypes = [vp, dp]
        L.oro_reset.argtypes = [vp]
        L.oro_add_cpml.argtypes = [vp, C.POINTER(sz), d, d, d, d, C.c_int]
        L.oro_cpml_psi.restype = sz
        L.oro_cpml_psi.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.oro_add_mur.argtypes = [vp, C.POINTER(C.c_int), d]
        L.oro_add_upml.argtypes = [vp, C.POINTER(sz), d, d]
        L.oro_upml_regions.restype = sz
        L.oro_upml_regions.argtypes = [vp]
        L.oro_upml_array.restype = sz
        L.oro_upml_array.argtypes = [vp, sz, C.c_int, C.c_int, dp, C.POINTER(sz), C.POINTER(sz)]
        L.oro_add_ade.argtypes = [vp, C.c_int, dp, C.POINTER(u32), sz, C.c_int]
        L.oro_ade_state.restype = sz
        L.oro_ade_state.argtypes = [vp, sz, dp]
        L.oro_add_field_probe.restype = sz
        L.oro_add_field_probe.argtypes = [vp, u32, u32, u32]
        L.oro_probe_data.restype = sz
        L.oro_probe_data.argtypes = [vp, sz, dp]
        L.oro_add_lumped.argtypes = [vp, C.c_int, C.POINTER(sz), d]
        L.oro_add_signals.restype = sz
        L.oro_add_signals.argtypes = [vp, sz, C.POINTER(u64), C.POINTER(u32), C.POINTER(u64), dp]
        L.oro_signal_series.restype = sz
        L.oro_signal_series.argtypes = [vp, sz, dp]
        L.oro_energy.restype = d
        L.oro_energy.argtypes = [vp, C.c_int]
        L.oro_yee_steps.argtypes = [vp, u64]
        L.oro_set_hook.restype = None
        L.oro_set_hook.argtypes = [vp, _HOOK_FN, _TERM_FN, vp, C.c_uint]
        L.oro_run_batch.restype = C.c_int
        L.oro_run_batch.argtypes = [vp, C.POINTER(_Batch)]
        _lib = L
    return _lib


_HOOK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.c_void_p))
_TERM_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class BatchResult:
    timesteps_executed: int
    termination_reason: int  # 0 steps completed, 1 energy decay, 2 extension stop
    final_decay_db: float
    energy_samples: np.ndarray  # [n, 3]: timestep, e_energy, h_energy


class OracleSim:
    """One simulated grid; prec is 32 (parity oracle) or 64 (error oracle)."""

    def __init__(self, nx: int, ny: int, nz: int, delta, prec: int = 32):
        if np.isscalar(delta):
            delta = (float(delta),) * 3
        self.shape = (int(nx), int(ny), int(nz))
        self.N = nx * ny * nz
        self.prec = prec
        self._h = lib().oro_create(prec, nx, ny, nz, *map(float, delta))
        self.delta = tuple(map(float, delta))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oro_destroy(h)
            self._h = None

    @property
    def dt(self) -> float:
        return lib().oro_dt(self._h)

    @property
    def timestep(self) -> int:
        return lib().oro_timestep(self._h)

    def set_material(self, eps_r, mu_r, sigma_e, sigma_m, region):
        i0, i1, j0, j1, k0, k1 = region
        lib().oro_set_material(self._h, eps_r, mu_r, sigma_e, sigma_m, i0, i1, j0, j1, k0, k1)

    def set_pec(self, region):
        lib().oro_set_pec(self._h, *region)

    def coefficients(self) -> np.ndarray:
        """[12, nx, ny, nz]: ca(x,y,z), cb(x,y,z), da(x,y,z), db(x,y,z)."""
        if self.prec == 32:
            out = np.empty((12, self.N), np.float32)
            for w in range(12):
                lib().oro_get_coeff_f32(self._h, w, out[w].ctypes.data_as(C.POINTER(C.c_float)))
        else:
            out = np.empty((12, self.N), np.float64)
            for w in range(12):
                lib().oro_get_coeff(self._h, w, _dptr(out[w]))
        return out.reshape((12,) + self.shape)

    def set_coefficients(self, coeff: np.ndarray):
        """Replace the 12 coefficient arrays ([12, nx, ny, nz] float32; FP32 oracle)."""
        assert self.prec == 32
        a = np.ascontiguousarray(coeff, dtype=np.float32).reshape(12, self.N)
        for w in range(12):
            lib().oro_set_coeff_f32(self._h, w, a[w].ctypes.data_as(C.POINTER(C.c_float)))

    def fields(self) -> np.ndarray:
        """[6, nx, ny, nz]: Ex, Ey, Ez, Hx, Hy, Hz (float32 or float64)."""
        if self.prec == 32:
            out = np.empty(6 * self.N, np.float32)
            lib().oro_get_fields_f32(self._h, out.ctypes.data_as(C.POINTER(C.c_float)))
        else:
            out = np.empty(6 * self.N, np.float64)
            lib().oro_get_fields(self._h, _dptr(out))
        return out.reshape((6,) + self.shape)

    def set_fields(self, f: np.ndarray):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert a.size == 6 * self.N
        lib().oro_set_fields(self._h, _dptr(a))

    def reset(self):
        lib().oro_reset(self._h)

    def add_cpml(self, layers: Sequence[int] = (8,) * 6, grading_order=3.0, reflection_coeff=1e-6,
                 alpha_max=0.0, kappa_max=1.0, mode=MODE_COMPAT):
        arr = (C.c_size_t * 6)(*layers)
        lib().oro_add_cpml(self._h, arr, grading_order, reflection_coeff, alpha_max, kappa_max, mode)

    def cpml_array(self, face: int, field: str, which: str) -> Optional[np.ndarray]:
        f = 0 if field == "E" else 1
        w = {"pos": 0, "neg": 1, "b": 2, "a": 3, "km": 4}[which]
        n = lib().oro_cpml_psi(self._h, face, f, w, None)
        if n == 0:
            return None
        out = np.empty(n, np.float64)
        lib().oro_cpml_psi(self._h, face, f, w, _dptr(out))
        return out

    UPML_ARRAYS = ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn", "volt_flux", "curr_flux")

    def add_upml(self, layers: Sequence[int] = (8,) * 6, grading_order: float = 3.0, reflection_coeff: float = 1e-6):
        arr = (C.c_size_t * 6)(*layers)
        lib().oro_add_upml(self._h, arr, grading_order, reflection_coeff)

    def upml_regions(self):
        """[(start, size)] of the UPML regions in build order."""
        out = []
        for q in range(lib().oro_upml_regions(self._h)):
            st = (C.c_size_t * 3)()
            sz = (C.c_size_t * 3)()
            lib().oro_upml_array(self._h, q, 0, 0, None, st, sz)
            out.append((tuple(st), tuple(sz)))
        return out

    def upml_array(self, region: int, name: str, component: int) -> np.ndarray:
        w = self.UPML_ARRAYS.index(name)
        st = (C.c_size_t * 3)()
        sz = (C.c_size_t * 3)()
        n = lib().oro_upml_array(self._h, region, w, component, None, st, sz)
        out = np.empty(n, np.float64)
        lib().oro_upml_array(self._h, region, w, component, _dptr(out), st, sz)
        return out.reshape(tuple(sz))

    def add_mur(self, enabled: Sequence[bool] = (True,) * 6, phase_velocity: float = 0.0):
        arr = (C.c_int * 6)(*[int(bool(x)) for x in enabled])
        lib().oro_add_mur(self._h, arr, phase_velocity)

    def add_ade(self, kind: int, params: Sequence[float], positions: np.ndarray, mode=MODE_COMPAT):
        p = np.ascontiguousarray(np.asarray(positions, dtype=np.uint32).reshape(-1, 3))
        prm = np.zeros(3, np.float64)
        prm[: len(params)] = params
        lib().oro_add_ade(self._h, kind, _dptr(prm), p.ctypes.data_as(C.POINTER(C.c_uint32)), p.shape[0], mode)

    def ade_state(self, idx: int) -> np.ndarray:
        n = lib().oro_ade_state(self._h, idx, None)
        out = np.empty(3 * n, np.float64)
        lib().oro_ade_state(self._h, idx, _dptr(out))
        return out.reshape(3, n)

    def add_field_probe(self, pos) -> int:
        return lib().oro_add_field_probe(self._h, *map(int, pos))

    def probe_data(self, idx: int) -> np.ndarray:
        n = lib().oro_probe_data(self._h, idx, None)
        out = np.empty(6 * n, np.float64)
        lib().oro_probe_data(self._h, idx, _dptr(out))
        return out.reshape(n, 6)

    def add_lumped(self, direction: int, region, R: float):
        arr = (C.c_size_t * 6)(*map(int, region))
        lib().oro_add_lumped(self._h, direction, arr, float(R))

    def add_signals(self, terms) -> int:
        """terms: list (one per signal) of lists of (component 0..5, i, j, k, weight)."""
        off = np.zeros(len(terms) + 1, np.uint64)
        comp, cell, w = [], [], []
        nx, ny, nz = self.shape
        for m, tl in enumerate(terms):
            for (c, i, j, k, wt) in tl:
                comp.append(c)
                cell.append((i * ny + j) * nz + k)
                w.append(wt)
            off[m + 1] = len(comp)
        comp = np.asarray(comp, np.uint32)
        cell = np.asarray(cell, np.uint64)
        w = np.asarray(w, np.float64)
        return lib().oro_add_signals(self._h, len(terms), off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     comp.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     cell.ctypes.data_as(C.POINTER(C.c_uint64)), _dptr(w))

    def signal_series(self, idx: int, n_signals: int) -> np.ndarray:
        steps = lib().oro_signal_series(self._h, idx, None)
        out = np.empty(steps * n_signals, np.float64)
        lib().oro_signal_series(self._h, idx, _dptr(out))
        return out.reshape(steps, n_signals)

    def set_hook(self, hook=None, hooks=(), check_termination=None):
        """Extension hooks on the oracle's fields, with the engine's host-callback
        contract: hook(which, step, fields) gets six arrays of the grid's shape
        (float32 or float64 as the oracle's precision) to change in place."""
        dt = np.float32 if self.prec == 32 else np.float64
        ct = C.c_float if self.prec == 32 else C.c_double
        shape, n = self.shape, self.N

        def _hook(user, which, step, fields):
            views = [np.ctypeslib.as_array(C.cast(fields[d], C.POINTER(ct)), (n,)).reshape(shape) for d in range(6)]
            assert views[0].dtype == dt
            return int(hook(int(which), int(step), views) or 0)

        def _term(user):
            return int(bool(check_termination()))

        self._hook_keep = (_HOOK_FN(_hook) if hook else _HOOK_FN(),
                           _TERM_FN(_term) if check_termination else _TERM_FN())
        lib().oro_set_hook(self._h, self._hook_keep[0], self._hook_keep[1], None,
                           sum(1 << int(h) for h in hooks))

    def energy(self):
        return lib().oro_energy(self._h, 0), lib().oro_energy(self._h, 1)

    def yee_steps(self, n: int):
        lib().oro_yee_steps(self._h, n)

    def run_batch(self, num_steps: Optional[int], excitations=(), samples: Optional[np.ndarray] = None,
                  max_timesteps: Optional[int] = None, check_interval: int = 100,
                  energy_decay_db: Optional[float] = None, sample_interval: int = 0,
                  track_peak: bool = False, energy_cap: int = 100000) -> BatchResult:
        """excitations: list of (i, j, k, dir, soft); samples: [n_exc, n_samples] float64."""
        n_exc = len(excitations)
        ex = (_Exc * max(n_exc, 1))()
        for e, (i, j, k, d, soft) in enumerate(excitations):
            ex[e] = _Exc(i, j, k, d, int(bool(soft)))
        if samples is None:
            samples = np.zeros((max(n_exc, 1), 0))
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        eout = np.zeros((energy_cap, 3), np.float64)
        b = _Batch()
        b.num_steps = int(num_steps or 0)
        b.max_timesteps = int(max_timesteps or 0)
        b.exc = ex
        b.n_exc = n_exc
        b.samples = _dptr(samples)
        b.n_samples = samples.shape[1] if samples.ndim == 2 else 0
        b.check_interval = check_interval
        b.energy_decay_db = float(energy_decay_db) if energy_decay_db else 0.0
        b.sample_interval = sample_interval
        b.track_peak = int(track_peak)
        b.energy_out = _dptr(eout)
        b.energy_cap = energy_cap
        rc = lib().oro_run_batch(self._h, C.byref(b))
        if rc:
            raise RuntimeError(f"oracle batch: a hook returned status {rc}")
        n = min(b.n_energy, energy_cap)
        return BatchResult(b.steps_executed, b.reason, b.final_decay_db, eout[:n].copy())


def set_threads(n: int):
    lib().oro_set_threads(int(n))


def max_threads() -> int:
    return lib().oro_max_threads()

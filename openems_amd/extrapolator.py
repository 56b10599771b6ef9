"""Learned ringdown extrapolator over liboe.so (oe_extrap_*).

The network and its training run in the library's HIP kernels (BF16 matrix
cores, FP32 accumulation); this module marshals arrays and holds the data
pipeline that turns port V/I series into training pairs:

  window   the first `window` steps of V and I, decimated by `decimate`
           (every n-th sample) and divided by the window's peak |V| and |I|
  target   the following `tail * decimate` steps, decimated and divided by
           the same peaks

Layouts follow include/oe/oe.h: inputs [B][L0][2], targets [B][2][tail].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import numpy as np

from . import _ffi
from ._ffi import check, lib

ORDER = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4", "W5", "b5"]
KERNEL = 9


@dataclass
class ExtrapolatorDims:
    L0: int = 256
    c1: int = 64
    c2: int = 128
    c3: int = 128
    hidden: int = 384
    modes: int = 32
    batch: int = 256
    tail: int = 2250

    def shapes(self) -> Dict[str, Tuple[int, ...]]:
        L3 = self.L0 // 8
        return {"W1": (self.c1, 2, KERNEL), "b1": (self.c1,), "W2": (self.c2, self.c1, KERNEL), "b2": (self.c2,),
                "W3": (self.c3, self.c2, KERNEL), "b3": (self.c3,), "W4": (self.hidden, L3 * self.c3),
                "b4": (self.hidden,), "W5": (6 * self.modes, self.hidden), "b5": (6 * self.modes,)}


def pack(params: Dict[str, np.ndarray], dims: ExtrapolatorDims) -> np.ndarray:
    return np.concatenate([np.asarray(params[k], np.float32).reshape(-1) for k in ORDER])


def unpack(flat: np.ndarray, dims: ExtrapolatorDims) -> Dict[str, np.ndarray]:
    out, o = {}, 0
    for k in ORDER:
        sh = dims.shapes()[k]
        n = int(np.prod(sh))
        out[k] = np.asarray(flat[o:o + n]).reshape(sh)
        o += n
    return out


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Extrapolator:
    def __init__(self, dims: ExtrapolatorDims, params: Optional[np.ndarray] = None, seed: int = 0, device: int = 0):
        self.dims = dims
        d = _ffi.ExtrapDims(dims.L0, dims.c1, dims.c2, dims.c3, dims.hidden, dims.modes, dims.batch, dims.tail)
        h = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        check(lib().oe_extrap_create(int(device), C.byref(d), None if p is None else _fp(p), int(seed), C.byref(h)))
        self._h = h
        self.n_params = lib().oe_extrap_param_count(h)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().oe_extrap_destroy(self._h)
            self._h = C.c_void_p()

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(lib().oe_extrap_params(self._h, _fp(out), out.size))
        return out

    def set_params(self, flat: np.ndarray):
        a = np.ascontiguousarray(flat, np.float32)
        check(lib().oe_extrap_set_params(self._h, _fp(a), a.size))

    @staticmethod
    def _inputs(x: np.ndarray) -> np.ndarray:
        """[B, 2, L0] -> device layout [B, L0, 2]."""
        return np.ascontiguousarray(np.asarray(x, np.float32).transpose(0, 2, 1))

    def forward(self, x: np.ndarray, tails: bool = True):
        xi = self._inputs(x)
        B = xi.shape[0]
        modes = np.empty((B, 6 * self.dims.modes), np.float32)
        t = np.empty((B, 2, self.dims.tail), np.float32) if tails else None
        check(lib().oe_extrap_forward(self._h, _fp(xi), B, _fp(modes), None if t is None else _fp(t)))
        return modes, t

    def gradients(self, x: np.ndarray, target: np.ndarray):
        xi = self._inputs(x)
        tg = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        g = np.empty(self.n_params, np.float32)
        check(lib().oe_extrap_gradients(self._h, _fp(xi), _fp(tg), xi.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, x: np.ndarray, target: np.ndarray, lr: float) -> float:
        xi = self._inputs(x)
        tg = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        check(lib().oe_extrap_train_step(self._h, _fp(xi), _fp(tg), xi.shape[0], float(lr), C.byref(loss)))
        return loss.value


def training_pair(v: np.ndarray, i: np.ndarray, window: int, decimate: int, tail: int):
    """(input [2, window/decimate], target [2, tail], scales (sV, sI)) from
    full V, I series; samples past the end of the series count as zero."""
    L0 = window // decimate
    need = window + tail * decimate
    vv = np.zeros(need)
    ii = np.zeros(need)
    n = min(len(v), need)
    vv[:n] = v[:n]
    ii[:n] = i[:n]
    sv = max(np.abs(vv[:window]).max(), 1e-30)
    si = max(np.abs(ii[:window]).max(), 1e-30)
    x = np.stack([vv[:window:decimate][:L0] / sv, ii[:window:decimate][:L0] / si])
    t = np.stack([vv[window::decimate][:tail] / sv, ii[window::decimate][:tail] / si])
    return x.astype(np.float32), t.astype(np.float32), (sv, si)

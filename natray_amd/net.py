"""The CG warm-start network (include/natray.h, natray_net_*): a ctypes view of the C ABI.

A 4-layer 3D convolution net on 16^3 tiles, BF16 MFMA kernels with FP32 accumulation;
inputs per voxel (scaled initial residual, cell kind), output the scaled pressure correction.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

PARAMS = 15169
VOXELS = 4096
# Offsets of the layers' weights and biases in the parameter vector.
LAYERS = [
    # (cin, cout, relu, w_off, b_off)
    (2, 16, True, 0, 864),
    (16, 16, True, 880, 880 + 6912),
    (16, 16, True, 880 + 6928, 880 + 6928 + 6912),
    (16, 1, False, 880 + 2 * 6928, 880 + 2 * 6928 + 432),
]

_P = C.c_void_p
_SIGS = {
    "natray_net_create": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(_P)]),
    "natray_net_destroy": (None, [_P]),
    "natray_net_params": (C.c_int, [_P, _P]),
    "natray_net_set_params": (C.c_int, [_P, _P]),
    "natray_net_forward": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "natray_net_train_step": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "natray_net_grads": (C.c_int, [_P, _P]),
    "natray_net_train_step_residual": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "natray_net_layer_check": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, _P, _P, _P, _P, C.c_uint32, _P, _P, _P, _P]),
    "natray_net_last_error": (C.c_char_p, []),
}
_bound = False


def lib() -> C.CDLL:
    global _bound
    L = abi.lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def check(code: int) -> None:
    if code != abi.OK:
        raise abi.NatrayError(code, lib().natray_net_last_error().decode())


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Net:
    def __init__(self, max_tiles: int = 64, seed: int = 1):
        h = _P()
        check(lib().natray_net_create(max_tiles, seed, C.byref(h)))
        self._h = h
        self.max_tiles = max_tiles

    def close(self) -> None:
        if self._h:
            lib().natray_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def params(self) -> np.ndarray:
        out = np.empty(PARAMS, np.float32)
        check(lib().natray_net_params(self._h, abi.ptr(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = _f32(p)
        assert p.shape == (PARAMS,)
        check(lib().natray_net_set_params(self._h, abi.ptr(p)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        """x[tiles][4096][2] -> [tiles][4096]."""
        x = _f32(x)
        t = x.shape[0]
        assert x.shape == (t, VOXELS, 2)
        out = np.empty((t, VOXELS), np.float32)
        check(lib().natray_net_forward(self._h, abi.ptr(x), t, abi.ptr(out)))
        return out

    def train_step(self, x: np.ndarray, target: np.ndarray, mask: np.ndarray, lr: float) -> float:
        x, target, mask = _f32(x), _f32(target), _f32(mask)
        t = x.shape[0]
        loss = C.c_float()
        check(lib().natray_net_train_step(self._h, abi.ptr(x), abi.ptr(target), abi.ptr(mask), t, lr, C.byref(loss)))
        return float(loss.value)

    def train_step_residual(self, x: np.ndarray, diag: np.ndarray, nbm: np.ndarray, lr: float) -> float:
        """One Adam step on the scaled residual x - A o (natray_net_train_step_residual)."""
        x, diag = _f32(x), _f32(diag)
        nbm = np.ascontiguousarray(nbm, dtype=np.uint32)
        t = x.shape[0]
        assert x.shape == (t, VOXELS, 2) and diag.shape == (t, VOXELS) and nbm.shape == (t, VOXELS)
        loss = C.c_float()
        check(lib().natray_net_train_step_residual(self._h, abi.ptr(x), abi.ptr(diag), abi.ptr(nbm), t, lr, C.byref(loss)))
        return float(loss.value)

    def grads(self) -> np.ndarray:
        out = np.empty(PARAMS, np.float32)
        check(lib().natray_net_grads(self._h, abi.ptr(out)))
        return out


def layer_check(cin: int, cout: int, relu: bool, X: np.ndarray, W: np.ndarray, b: np.ndarray, dY: np.ndarray):
    """One layer on the GPU: (Y, dX, dW, db); see natray_net_layer_check."""
    X, W, b, dY = _f32(X), _f32(W), _f32(b), _f32(dY)
    t = X.shape[0]
    Y = np.empty((t, VOXELS, cout), np.float32)
    dX = np.empty((t, VOXELS, cin), np.float32)
    dW = np.empty((27, cin, cout), np.float32)
    db = np.empty(cout, np.float32)
    check(lib().natray_net_layer_check(cin, cout, 1 if relu else 0, abi.ptr(X), abi.ptr(W), abi.ptr(b), abi.ptr(dY), t,
                                       abi.ptr(Y), abi.ptr(dX), abi.ptr(dW), abi.ptr(db)))
    return Y, dX, dW, db

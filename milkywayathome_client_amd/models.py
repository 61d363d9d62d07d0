"""Initial conditions (host side) over ``libnbmodels.so``.

``DSFMT`` mirrors the Lua ``DSFMT.create(seed)`` object
(milkyway/src/milkyway_lua_dsfmt.c) and ``plummer`` mirrors
``predefinedModels.plummer`` (nbody/src/nbody_plummer.c:109-158).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
MODELS_LIB_PATH = os.path.join(_PKG_DIR, "libnbmodels.so")

_LIB = None


class _Dsfmt(C.Structure):
    _fields_ = [("status", C.c_uint64 * (2 * 192)), ("idx", C.c_int)]


def load_models() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(MODELS_LIB_PATH):
        raise RuntimeError(f"{MODELS_LIB_PATH} is missing: build it with "
                           "`python -m milkywayathome_client_amd.build`")
    lib = C.CDLL(MODELS_LIB_PATH)
    dp = C.POINTER(C.c_double)
    sp = C.POINTER(_Dsfmt)
    lib.nbm_dsfmt_init.argtypes = [sp, C.c_uint32]
    lib.nbm_dsfmt_init.restype = None
    for name in ("nbm_dsfmt_close1_open2", "nbm_dsfmt_close_open", "nbm_dsfmt_open_close",
                 "nbm_dsfmt_open_open"):
        f = getattr(lib, name)
        f.argtypes = [sp]
        f.restype = C.c_double
    lib.nbm_dsfmt_fill.argtypes = [sp, C.c_int, C.c_int, dp]
    lib.nbm_dsfmt_fill.restype = None
    lib.nbm_plummer.argtypes = [sp, C.c_int, C.c_double, C.c_double, dp, dp, C.c_int, dp, dp, dp,
                                C.POINTER(C.c_int)]
    lib.nbm_plummer.restype = C.c_int
    _LIB = lib
    return lib


class DSFMT:
    """dSFMT-19937 stream seeded like ``DSFMT.create(seed)``."""

    KINDS = {"close1_open2": 0, "close_open": 1, "open_close": 2, "open_open": 3}

    def __init__(self, seed: int):
        self._lib = load_models()
        self._st = _Dsfmt()
        self._lib.nbm_dsfmt_init(C.byref(self._st), int(seed) & 0xFFFFFFFF)

    @staticmethod
    def create(seed: int) -> "DSFMT":
        return DSFMT(seed)

    def draw(self, n: int, kind: str = "close_open") -> np.ndarray:
        out = np.empty(n)
        self._lib.nbm_dsfmt_fill(C.byref(self._st), self.KINDS[kind], n,
                                 out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def genrandOpenOpen(self) -> float:
        return self._lib.nbm_dsfmt_open_open(C.byref(self._st))

    def genrandCloseOpen(self) -> float:
        return self._lib.nbm_dsfmt_close_open(C.byref(self._st))


def plummer(nbody: int, prng: DSFMT, mass: float, scaleRadius: float,
            position: Sequence[float] = (0.0, 0.0, 0.0),
            velocity: Sequence[float] = (0.0, 0.0, 0.0), ignore: bool = False):
    """Equal-mass Plummer sphere; returns (pos (n,3), vel (n,3), mass (n,), type (n,))."""
    lib = load_models()
    pos = np.empty((nbody, 3))
    vel = np.empty((nbody, 3))
    m = np.empty(nbody)
    typ = np.empty(nbody, dtype=np.int32)
    rs = np.ascontiguousarray(position, dtype=np.float64)
    vs = np.ascontiguousarray(velocity, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    rc = lib.nbm_plummer(C.byref(prng._st), nbody, mass, scaleRadius, rs.ctypes.data_as(dp),
                         vs.ctypes.data_as(dp), int(bool(ignore)), pos.ctypes.data_as(dp),
                         vel.ctypes.data_as(dp), m.ctypes.data_as(dp),
                         typ.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise ValueError(f"plummer: bad arguments (nbody={nbody})")
    return pos, vel, m, typ

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
    lib.nbm_ext_accel.argtypes = [dp, dp, dp]
    lib.nbm_ext_accel.restype = None
    lib.nbm_lbr_to_cartesian.argtypes = [dp, C.c_double, dp]
    lib.nbm_lbr_to_cartesian.restype = None
    lib.nbm_reverse_orbit.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp, dp]
    lib.nbm_reverse_orbit.restype = None
    lib.nbm_virial_radius.argtypes = [C.c_double] * 4
    lib.nbm_virial_radius.restype = C.c_double
    lib.nbm_calculate_eps2.argtypes = [C.c_double] * 5
    lib.nbm_calculate_eps2.restype = C.c_double
    lib.nbm_calculate_timestep.argtypes = [C.c_double] * 2
    lib.nbm_calculate_timestep.restype = C.c_double
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


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def potential_array(potential) -> np.ndarray:
    """Flatten a nbody.Potential into the 15-double layout of nbm_ext_accel."""
    p = potential
    return np.array([p.spherical.type, p.spherical.mass, p.spherical.scale,
                     p.disk.type, p.disk.mass, p.disk.scaleLength, p.disk.scaleHeight,
                     p.disk2.type, p.disk2.mass, p.disk2.scaleLength, p.disk2.scaleHeight,
                     p.halo.type, p.halo.vhalo, p.halo.scaleLength, p.halo.flattenZ],
                    dtype=np.float64)


def lbr_to_cartesian(lbr, sun_gc_dist: float = 8.0) -> np.ndarray:
    """lbrToCartesian (nbody/src/nbody_coordinates.c:68-71); degrees, kpc."""
    lib = load_models()
    a = np.ascontiguousarray(lbr, dtype=np.float64)
    out = np.empty(3)
    lib.nbm_lbr_to_cartesian(_dp(a), sun_gc_dist, _dp(out))
    return out


def reverse_orbit(potential, position, velocity, tstop: float, dt: float):
    """reverseOrbit{potential, position, velocity, tstop, dt}
    (nbody/src/nbody_orbit_integrator.c:44-86)."""
    lib = load_models()
    P = potential_array(potential)
    x = np.ascontiguousarray(position, dtype=np.float64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    fx, fv = np.empty(3), np.empty(3)
    lib.nbm_reverse_orbit(_dp(P), _dp(x), _dp(v), tstop, dt, _dp(fx), _dp(fv))
    return fx, fv


def ext_accel(potential, pos) -> np.ndarray:
    lib = load_models()
    P = potential_array(potential)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(pos)
    for i in range(pos.shape[0]):
        lib.nbm_ext_accel(_dp(P), _dp(pos[i]), _dp(out[i]))
    return out


def virial_radius(a_b: float, a_d: float, m_b: float, m_d: float) -> float:
    """nbCalculateVirial (nbody/src/nbody_virial.c:38-102): the virial radius
    of a two-Plummer dwarf."""
    return float(load_models().nbm_virial_radius(a_b, a_d, m_b, m_d))


def calculate_eps2(nbody: float, a_b: float, a_d: float, m_b: float, m_d: float) -> float:
    """calculateEps2(nbody, a_b, a_d, M_b, M_d, 0) (nbody/src/nbody_lua_models.c:158-169)."""
    return load_models().nbm_calculate_eps2(float(nbody), a_b, a_d, m_b, m_d)


def calculate_timestep(mass: float, r0: float) -> float:
    """calculateTimestep (nbody/src/nbody_lua_models.c:139-142)."""
    return load_models().nbm_calculate_timestep(mass, r0)

"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  The product (libnbhip) never
uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

CRITERIA = {"treecode": 0, "sw93": 1, "bh86": 2, "exact": 3}


class OrcPotential(C.Structure):
    _fields_ = [
        ("sphere_type", C.c_int), ("sphere_mass", C.c_double), ("sphere_scale", C.c_double),
        ("disk_type", C.c_int), ("disk_mass", C.c_double), ("disk_a", C.c_double),
        ("disk_b", C.c_double),
        ("disk2_type", C.c_int), ("disk2_mass", C.c_double), ("disk2_a", C.c_double),
        ("disk2_b", C.c_double),
        ("halo_type", C.c_int), ("halo_vhalo", C.c_double), ("halo_d", C.c_double),
        ("halo_q", C.c_double),
    ]


class OrcCtx(C.Structure):
    _fields_ = [
        ("eps2", C.c_double), ("theta", C.c_double), ("timestep", C.c_double),
        ("tree_rsize", C.c_double), ("criterion", C.c_int), ("use_quad", C.c_int),
        ("allow_incest", C.c_int), ("use_potential", C.c_int), ("pot", OrcPotential),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)),
        ("mass", C.POINTER(C.c_double)), ("acc", C.POINTER(C.c_double)),
        ("rsize", C.c_double), ("step", C.c_uint), ("tree_incest", C.c_int),
        ("tree", C.c_void_p),
    ]


_LIB = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sp = C.POINTER(OrcState)
    cp = C.POINTER(OrcCtx)
    lib.orc_state_new.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double]
    lib.orc_state_new.restype = sp
    lib.orc_state_free.argtypes = [sp]
    lib.orc_state_free.restype = None
    lib.orc_grav_map.argtypes = [cp, sp]
    lib.orc_grav_map.restype = C.c_int
    lib.orc_step.argtypes = [cp, sp]
    lib.orc_step.restype = C.c_int
    lib.orc_run.argtypes = [cp, sp, C.c_uint]
    lib.orc_run.restype = C.c_int
    lib.orc_direct_accel.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _dp]
    lib.orc_direct_accel.restype = None
    lib.orc_ext_accel.argtypes = [C.POINTER(OrcPotential), _dp, _dp]
    lib.orc_ext_accel.restype = None
    lib.orc_walk_bodies.argtypes = [cp, sp, C.c_int, _ip, _dp, _ip]
    lib.orc_walk_bodies.restype = None
    lib.orc_walk_counts.argtypes = [cp, sp, C.c_int, _ip, C.POINTER(C.c_longlong)]
    lib.orc_walk_counts.restype = None
    lib.orc_tree_ncell.argtypes = [C.c_void_p]
    lib.orc_tree_ncell.restype = C.c_int
    lib.orc_tree_maxdepth.argtypes = [C.c_void_p]
    lib.orc_tree_maxdepth.restype = C.c_int
    lib.orc_tree_nnode.argtypes = [C.c_void_p]
    lib.orc_tree_nnode.restype = C.c_int
    lib.orc_tree_preorder.argtypes = [C.c_void_p, _dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp,
                                      _dp, _dp]
    lib.orc_tree_preorder.restype = None
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_set_threads.restype = None
    lib.orc_get_threads.argtypes = []
    lib.orc_get_threads.restype = C.c_int
    _LIB = lib
    return lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def make_potential(sphere=(0, 0.0, 0.0), disk=(0, 0.0, 0.0, 0.0), disk2=(0, 0.0, 0.0, 0.0),
                   halo=(0, 0.0, 0.0, 1.0)) -> OrcPotential:
    p = OrcPotential()
    p.sphere_type, p.sphere_mass, p.sphere_scale = sphere
    p.disk_type, p.disk_mass, p.disk_a, p.disk_b = disk
    p.disk2_type, p.disk2_mass, p.disk2_a, p.disk2_b = disk2
    p.halo_type, p.halo_vhalo, p.halo_d, p.halo_q = halo
    return p


def milky_way() -> OrcPotential:
    """Hernquist bulge + Miyamoto-Nagai disk + logarithmic halo of the sample
    workunits (nbody/sample_workunits/for_developers.lua:156-161)."""
    return make_potential(sphere=(1, 1.52954402e5, 0.7), disk=(1, 4.45865888e5, 6.5, 0.26),
                          halo=(1, 74.61, 12.0, 1.0))


def make_ctx(eps2: float, theta: float, timestep: float = 1e-3, criterion: str = "bh86",
             use_quad: bool = True, tree_rsize: float = 4.0, allow_incest: bool = False,
             potential: Optional[OrcPotential] = None) -> OrcCtx:
    c = OrcCtx()
    c.eps2, c.theta, c.timestep, c.tree_rsize = eps2, theta, timestep, tree_rsize
    c.criterion = CRITERIA[criterion.lower()]
    c.use_quad = int(bool(use_quad))
    c.allow_incest = int(bool(allow_incest))
    c.use_potential = 0 if potential is None else 1
    c.pot = potential if potential is not None else make_potential()
    return c


class Oracle:
    """One CPU simulation: the reference algorithm restated in C."""

    def __init__(self, ctx: OrcCtx, pos, vel, mass, threads: int = 0):
        self.lib = load()
        if threads:
            self.lib.orc_set_threads(threads)
        self.ctx = ctx
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.n = mass.shape[0]
        self.st = self.lib.orc_state_new(self.n, _p(pos), _p(vel), _p(mass), ctx.tree_rsize)

    def __del__(self):
        try:
            self.lib.orc_state_free(self.st)
        except Exception:
            pass

    def gravity(self) -> int:
        return self.lib.orc_grav_map(C.byref(self.ctx), self.st)

    def step(self) -> int:
        return self.lib.orc_step(C.byref(self.ctx), self.st)

    def run(self, nsteps: int) -> int:
        return self.lib.orc_run(C.byref(self.ctx), self.st, nsteps)

    def _arr(self, ptr, k):
        return np.ctypeslib.as_array(ptr, shape=(self.n * k,)).copy()

    @property
    def pos(self):
        return self._arr(self.st.contents.pos, 3).reshape(self.n, 3)

    @property
    def vel(self):
        return self._arr(self.st.contents.vel, 3).reshape(self.n, 3)

    @property
    def acc(self):
        return self._arr(self.st.contents.acc, 3).reshape(self.n, 3)

    @property
    def rsize(self) -> float:
        return self.st.contents.rsize

    def tree_summary(self):
        t = self.st.contents.tree
        return (self.lib.orc_tree_ncell(t), self.lib.orc_tree_maxdepth(t),
                self.lib.orc_tree_nnode(t))

    def tree(self) -> dict:
        t = self.st.contents.tree
        nn = self.lib.orc_tree_nnode(t)
        out = {
            "kind": np.empty(nn, np.int32), "ref": np.empty(nn, np.int32),
            "skip": np.empty(nn, np.int32), "level": np.empty(nn, np.int32),
            "pos": np.empty((nn, 3)), "mass": np.empty(nn), "rcrit2": np.empty(nn),
            "quad": np.empty((nn, 6)), "mid": np.empty((nn, 3)),
        }
        st = self.st.contents
        self.lib.orc_tree_preorder(t, st.pos, st.mass, _p(out["kind"], C.c_int),
                                   _p(out["ref"], C.c_int), _p(out["skip"], C.c_int),
                                   _p(out["level"], C.c_int), _p(out["pos"]), _p(out["mass"]),
                                   _p(out["rcrit2"]), _p(out["quad"]), _p(out["mid"]))
        return out

    def walk(self, idx) -> tuple:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.empty((idx.shape[0], 3))
        inc = C.c_int(0)
        self.lib.orc_walk_bodies(C.byref(self.ctx), self.st, idx.shape[0], _p(idx, C.c_int),
                                 _p(acc), C.byref(inc))
        return acc, inc.value

    def walk_counts(self, idx) -> tuple:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        c = (C.c_longlong * 3)()
        self.lib.orc_walk_counts(C.byref(self.ctx), self.st, idx.shape[0], _p(idx, C.c_int), c)
        return int(c[0]), int(c[1]), int(c[2])


def direct_accel(pos, mass, eps2: float, idx=None) -> np.ndarray:
    lib = load()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = mass.shape[0]
    idx = np.arange(n, dtype=np.int32) if idx is None else np.ascontiguousarray(idx, np.int32)
    acc = np.empty((idx.shape[0], 3))
    lib.orc_direct_accel(n, _p(pos), _p(mass), eps2, idx.shape[0], _p(idx, C.c_int), _p(acc))
    return acc


def ext_accel(pot: OrcPotential, pos) -> np.ndarray:
    lib = load()
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(pos)
    for i in range(pos.shape[0]):
        lib.orc_ext_accel(C.byref(pot), _p(pos[i]), _p(out[i]))
    return out

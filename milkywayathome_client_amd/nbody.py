"""Host-side mirror of the reference's nbody run interface over the nbhip C ABI.

The classes follow the Lua-facing API that the reference's tests drive
(``NBodyCtx.create``, ``Potential.create``, ``NBodyState.create``,
``st:step(ctx)``; nbody/src/nbody_lua_types/*.c), with the force step and the
integrator running in ``libnbhip.so`` on the GPU.  Nothing here computes
physics on the CPU: if the library is missing the calls fail.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libnbhip.so")

# NBodyStatus bits (nbody/include/nbody_types.h:541-577)
NBODY_SUCCESS = 0
NBODY_ERROR = 1 << 1
NBODY_TREE_STRUCTURE_ERROR = 1 << 3
NBODY_TREE_INCEST_FATAL = 1 << 4
NBODY_GPU_ERROR = 1 << 7
NBODY_CAPABILITY_ERROR = 1 << 8
NBODY_UNSUPPORTED = 1 << 11
NBODY_USER_ERROR = 1 << 12
NBODY_MAX_DEPTH_ERROR = 1 << 16
NBODY_CELL_OVERFLOW_ERROR = 1 << 17
NBODY_TREE_INCEST_NONFATAL = 1 << 24
NBODY_STATUS_ALL_WARNINGS = 0xFF000000

CRITERIA = {"treecode": 0, "sw93": 1, "bh86": 2, "exact": 3}
PRECISION = {"fp64": 0, "mixed": 1, "mixed_coherent": 2, "mixed_group": 3}


def status_is_fatal(rc: int) -> bool:
    return (rc & ~NBODY_STATUS_ALL_WARNINGS) != 0


class NBodyError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"nbhip status 0x{rc:x}: {msg}")
        self.status = rc


class _Potential(C.Structure):
    _fields_ = [
        ("sphere_type", C.c_int), ("sphere_mass", C.c_double), ("sphere_scale", C.c_double),
        ("disk_type", C.c_int), ("disk_mass", C.c_double),
        ("disk_scale_length", C.c_double), ("disk_scale_height", C.c_double),
        ("disk2_type", C.c_int), ("disk2_mass", C.c_double),
        ("disk2_scale_length", C.c_double), ("disk2_scale_height", C.c_double),
        ("halo_type", C.c_int), ("halo_vhalo", C.c_double),
        ("halo_scale_length", C.c_double), ("halo_flatten_z", C.c_double),
    ]


class _Ctx(C.Structure):
    _fields_ = [
        ("eps2", C.c_double), ("theta", C.c_double), ("timestep", C.c_double),
        ("tree_rsize", C.c_double), ("criterion", C.c_int), ("use_quad", C.c_int),
        ("allow_incest", C.c_int), ("use_potential", C.c_int), ("precision", C.c_int),
        ("lmc", C.c_int), ("pot", _Potential),
    ]


class _TreeInfo(C.Structure):
    _fields_ = [
        ("rsize", C.c_double), ("ncell", C.c_int), ("max_depth", C.c_int),
        ("nnode", C.c_int), ("nmassive", C.c_int), ("step", C.c_uint), ("status", C.c_uint),
    ]


class _DispParams(C.Structure):
    _fields_ = [("iter_max", C.c_int), ("beta_sigma", C.c_double), ("vel_sigma", C.c_double),
                ("dist_sigma", C.c_double), ("pm_sigma", C.c_double),
                ("beta_correct", C.c_double), ("vel_correct", C.c_double),
                ("dist_correct", C.c_double), ("pm_correct", C.c_double),
                ("sun_vel", C.c_double * 3), ("ngp_dec", C.c_double), ("ngp_ra", C.c_double),
                ("l_ncp", C.c_double)]


# The proper-motion frame's defaults (nbody/include/nbody_defaults.h:40-45):
# the Sun's velocity in the model's units and the NGP / NCP angles in degrees.
SUN_VEL = (10.3, 229.2, 6.9)
NGP_DEC, NGP_RA, L_NCP = 27.4, 192.0, 123.0


class _HistParams(C.Structure):
    _fields_ = [("phi", C.c_double), ("theta", C.c_double), ("psi", C.c_double),
                ("lambda_start", C.c_double), ("lambda_end", C.c_double),
                ("beta_start", C.c_double), ("beta_end", C.c_double),
                ("lambda_bins", C.c_int), ("beta_bins", C.c_int), ("sun_gc_dist", C.c_double)]


class _Tune(C.Structure):
    _fields_ = [("block", C.c_int), ("batch", C.c_int), ("ms", C.c_double)]


_LIKE_FN = C.CFUNCTYPE(C.c_uint, C.c_void_p, C.c_void_p, C.c_uint, C.POINTER(C.c_double))
_CKPT_FN = C.CFUNCTYPE(C.c_uint, C.c_void_p, C.c_void_p, C.c_uint)


class _RunHooks(C.Structure):
    _fields_ = [("nstep", C.c_uint), ("use_best_like", C.c_int),
                ("best_like_start", C.c_double), ("time_evolve", C.c_double),
                ("keep_best_bodies", C.c_int), ("chunk", C.c_uint),
                ("likelihood", _LIKE_FN), ("checkpoint", _CKPT_FN), ("user", C.c_void_p)]


class _BestLike(C.Structure):
    _fields_ = [("lk", C.c_double * 10), ("time", C.c_double), ("step", C.c_uint),
                ("count", C.c_uint), ("evaluated", C.c_uint)]


class _CkptHeader(C.Structure):
    _fields_ = [("major_version", C.c_uint), ("minor_version", C.c_uint), ("nbody", C.c_uint),
                ("step", C.c_uint), ("real_size", C.c_uint), ("ptr_size", C.c_uint),
                ("n_orbit_trace", C.c_uint), ("n_shift_lmc", C.c_uint),
                ("tree_incest", C.c_uint), ("rsize", C.c_double)]


class _CkptIO(C.Structure):
    _fields_ = [("major_version", C.c_uint), ("minor_version", C.c_uint),
                ("ctx", C.c_void_p), ("ctx_size", C.c_size_t),
                ("state", C.c_void_p), ("state_size", C.c_size_t),
                ("type", C.POINTER(C.c_short)), ("id", C.POINTER(C.c_uint)),
                ("best_pos", C.POINTER(C.c_double)), ("best_vel", C.POINTER(C.c_double)),
                ("orbit_trace", C.POINTER(C.c_double)), ("n_orbit_trace", C.c_uint),
                ("shift_lmc", C.POINTER(C.c_double)), ("n_shift_lmc", C.c_uint),
                ("lmc_pos", C.c_double * 4), ("lmc_vel", C.c_double * 4)]


# The reference's NBodyCtx as an x86-64 build lays it out
# (nbody/include/nbody_types.h:450-521 with Potential,
# nbody/include/nbody_potential_types.h:42-175; mwbool is a char, enums are
# ints, and the MW_ALIGN_TYPE structs are 16-byte aligned, which the explicit
# pads reproduce).  A checkpoint file carries it raw in its header.
class _RefSpherical(C.Structure):
    _fields_ = [("type", C.c_int), ("mass", C.c_double), ("scale", C.c_double),
                ("_pad", C.c_char * 8)]


class _RefDisk(C.Structure):
    _fields_ = [("type", C.c_int), ("mass", C.c_double), ("scaleLength", C.c_double),
                ("scaleHeight", C.c_double), ("patternSpeed", C.c_double),
                ("startAngle", C.c_double)]


class _RefHalo(C.Structure):
    _fields_ = [("type", C.c_int)] + [(k, C.c_double) for k in (
        "vhalo", "scaleLength", "flattenZ", "flattenY", "flattenX", "triaxAngle", "c1", "c2",
        "c3", "mass", "gamma", "lambda_", "rho0")]


class _RefPotential(C.Structure):
    _fields_ = [("sphere", _RefSpherical), ("disk", _RefDisk), ("disk2", _RefDisk),
                ("halo", _RefHalo), ("rings", C.c_void_p), ("_pad", C.c_char * 8)]


class _RefNBodyCtx(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "eps2", "theta", "timestep", "timeEvolve", "timeBack", "treeRSize", "sunGCDist",
        "sunVelx", "sunVely", "sunVelz", "NGPdec", "NGPra", "lNCP", "b", "r", "vx", "vy",
        "vz")] + [("criterion", C.c_int), ("potentialType", C.c_int)] + [(k, C.c_char) for k in (
        "Nstep_control", "useBestLike", "useVelDisp", "useBetaDisp", "useBetaComp", "useVlos",
        "useDist", "usePropMot", "MultiOutput", "InitialOutput", "SimpleOutput", "useQuad",
        "allowIncest", "quietErrors")] + [(k, C.c_double) for k in (
        "BestLikeStart", "OutputFreq", "BetaSigma", "VelSigma", "DistSigma", "PMSigma",
        "IterMax", "BetaCorrect", "VelCorrect", "DistCorrect", "PMCorrect")] + [
        ("LMC", C.c_char), ("LMCmass", C.c_double), ("LMCscale", C.c_double),
        ("LMCDynaFric", C.c_char), ("coulomb_log", C.c_double), ("calibrationRuns", C.c_uint),
        ("Ntsteps", C.c_double), ("checkpointT", C.c_int64), ("nStep", C.c_uint),
        ("_pad", C.c_char * 12), ("pot", _RefPotential)]


assert C.sizeof(_RefSpherical) == 32 and C.sizeof(_RefPotential) == 256
assert _RefNBodyCtx.pot.offset == 336 and C.sizeof(_RefNBodyCtx) == 592

class _DwarfParams(C.Structure):
    _fields_ = [("light_type", C.c_int), ("dark_type", C.c_int),
                ("light_mass", C.c_double), ("light_scale", C.c_double),
                ("dark_mass", C.c_double), ("dark_scale", C.c_double),
                ("nbody", C.c_uint), ("nbody_baryon", C.c_uint), ("seed", C.c_ulonglong),
                ("pos", C.c_double * 3), ("vel", C.c_double * 3)]


# dwarf_t (nbody/include/nbody_potential_types.h:139-152) of the supported components
DWARF_TYPES = {"plummer": 0, "nfw": 1, "general_hernquist": 2}


# NBODY_VERSION_MAJOR / _MINOR of the reference build (nbody/CMakeLists.txt:24-27)
NBODY_VERSION = (1, 93)
NBODY_CHECKPOINT_ERROR = 1 << 6
# sizeof(NBodyState) of an x86-64 build without OpenCL
NBODY_STATE_SIZE = 384

EXPORTED = [
    "nbhip_create", "nbhip_upload", "nbhip_download", "nbhip_gravity", "nbhip_step",
    "nbhip_run", "nbhip_steps", "nbhip_tree_info_get", "nbhip_tree_download",
    "nbhip_direct_accel", "nbhip_walk_accel", "nbhip_walk_counters", "nbhip_set_profiling",
    "nbhip_phase_times", "nbhip_sync", "nbhip_release", "nbhip_last_error",
    "nbhip_batch_create", "nbhip_batch_upload", "nbhip_batch_download", "nbhip_batch_status",
    "nbhip_batch_tree_info", "nbhip_set_graph", "nbhip_autotune", "nbhip_set_tune",
    "nbhip_adjoint_forward", "nbhip_adjoint_backward", "nbhip_gravity_vjp", "nbhip_set_ignore",
    "nbhip_histogram", "nbhip_orbit_tangent", "nbhip_restore", "nbhip_histogram_dispersion",
    "nbhip_set_open_scale", "nbhip_batch_tree_download", "nbhip_histogram_noalloc",
    "nbhip_run_system", "nbhip_best_bodies", "nbhip_checkpoint_pack", "nbhip_checkpoint_unpack",
    "nbhip_checkpoint_write", "nbhip_checkpoint_read", "nbhip_mixed_dwarf",
    "nbhip_mixed_dwarf_state", "nbhip_dwarf_df", "nbhip_philox4x32_10",
    "nbhip_set_group_expansion", "nbhip_walk_counters5",
]

_LIB: Optional[C.CDLL] = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def load_library() -> C.CDLL:
    """Load the in-tree libnbhip.so; never builds, never falls back."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m milkywayathome_client_amd.build`")
    from . import build as _build
    stamp = open(_build.STAMP_PATH).read().strip() if os.path.exists(_build.STAMP_PATH) else ""
    if stamp != _build.source_hash():
        raise RuntimeError(
            f"{LIB_PATH} was not built from the current sources (digest "
            f"{stamp[:12] or 'missing'}): rebuild with `python -m milkywayathome_client_amd.build`")
    lib = C.CDLL(LIB_PATH)
    u = C.c_uint
    st = C.c_void_p
    sig = {
        "nbhip_create": (u, [C.POINTER(_Ctx), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
        "nbhip_upload": (u, [st, _dp, _dp, _dp]),
        "nbhip_download": (u, [st, _dp, _dp, _dp]),
        "nbhip_gravity": (u, [st]),
        "nbhip_step": (u, [st]),
        "nbhip_run": (u, [st, C.c_uint]),
        "nbhip_steps": (u, [st, C.c_uint]),
        "nbhip_tree_info_get": (u, [st, C.POINTER(_TreeInfo)]),
        "nbhip_tree_download": (u, [st, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
        "nbhip_batch_tree_download": (u, [st, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
        "nbhip_direct_accel": (u, [st, C.c_int, _ip, _dp]),
        "nbhip_walk_accel": (u, [st, _dp]),
        "nbhip_walk_counters": (u, [st, C.POINTER(C.c_ulonglong)]),
        "nbhip_set_profiling": (u, [st, C.c_int]),
        "nbhip_phase_times": (u, [st, _dp]),
        "nbhip_sync": (u, [st]),
        "nbhip_release": (None, [st]),
        "nbhip_last_error": (C.c_char_p, []),
        "nbhip_batch_create": (u, [C.POINTER(_Ctx), C.c_int, C.c_int, _ip, _dp,
                                   C.POINTER(C.c_void_p)]),
        "nbhip_batch_upload": (u, [st, C.c_int, _dp, _dp, _dp]),
        "nbhip_batch_download": (u, [st, C.c_int, _dp, _dp, _dp]),
        "nbhip_batch_status": (u, [st, C.POINTER(C.c_uint)]),
        "nbhip_batch_tree_info": (u, [st, C.c_int, C.POINTER(_TreeInfo)]),
        "nbhip_set_graph": (u, [st, C.c_int]),
        "nbhip_set_open_scale": (u, [st, C.c_double]),
        "nbhip_autotune": (u, [st, C.POINTER(_Tune), C.POINTER(_Tune), C.c_int,
                               C.POINTER(C.c_int)]),
        "nbhip_set_tune": (u, [st, C.POINTER(_Tune)]),
        "nbhip_set_ignore": (u, [st, C.POINTER(C.c_ubyte)]),
        "nbhip_histogram": (u, [st, C.POINTER(_HistParams), C.POINTER(C.c_uint),
                                C.POINTER(C.c_uint)]),
        "nbhip_run_system": (u, [st, C.POINTER(_RunHooks), C.POINTER(_BestLike)]),
        "nbhip_best_bodies": (u, [st, _dp, _dp]),
        "nbhip_histogram_noalloc": (u, [st, C.POINTER(_HistParams), C.POINTER(C.c_uint),
                                        C.POINTER(C.c_uint)]),
        "nbhip_adjoint_forward": (u, [st, C.c_uint, C.c_uint]),
        "nbhip_gravity_vjp": (u, [st, _dp, _dp, _dp]),
        "nbhip_adjoint_backward": (u, [st, _dp, _dp, _dp, _dp, _dp, _dp]),
        "nbhip_orbit_tangent": (u, [C.POINTER(_Potential), _dp, _dp, C.c_double, C.c_int, _dp,
                                    _dp, _dp, _dp]),
        "nbhip_restore": (u, [st, C.c_int, C.c_uint, C.c_double, C.c_uint]),
        "nbhip_histogram_dispersion": (u, [st, C.c_int, C.POINTER(_HistParams),
                                           C.POINTER(_DispParams), _dp]),
        "nbhip_checkpoint_pack": (u, [C.c_char_p, C.POINTER(_CkptHeader), C.POINTER(_CkptIO),
                                      _dp, _dp, _dp]),
        "nbhip_checkpoint_unpack": (u, [C.c_char_p, C.POINTER(_CkptHeader), C.POINTER(_CkptIO),
                                        _dp, _dp, _dp]),
        "nbhip_checkpoint_write": (u, [st, C.c_char_p, C.POINTER(_CkptIO)]),
        "nbhip_checkpoint_read": (u, [st, C.c_char_p, C.POINTER(_CkptIO)]),
        "nbhip_set_group_expansion": (u, [st, C.c_double]),
        "nbhip_walk_counters5": (u, [st, C.POINTER(C.c_ulonglong)]),
        "nbhip_mixed_dwarf": (u, [C.POINTER(_DwarfParams), C.c_int, _dp, _dp, _dp, _dp]),
        "nbhip_mixed_dwarf_state": (u, [st, C.c_int, C.POINTER(_DwarfParams), _dp]),
        "nbhip_dwarf_df": (u, [C.POINTER(_DwarfParams), C.c_int, C.c_int, C.c_int, _dp, _dp]),
        "nbhip_philox4x32_10": (u, [C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_uint)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


def _ptr(a: Optional[np.ndarray], ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


# ---------------------------------------------------------------------------
# Potential (nbody/src/nbody_lua_types/nbody_lua_potential.c and friends)
# ---------------------------------------------------------------------------
@dataclass
class Spherical:
    type: int = 0
    mass: float = 0.0
    scale: float = 0.0

    @staticmethod
    def hernquist(mass: float, scale: float) -> "Spherical":
        return Spherical(1, mass, scale)

    @staticmethod
    def plummer(mass: float, scale: float) -> "Spherical":
        return Spherical(2, mass, scale)

    @staticmethod
    def none(mass: float = 0.0) -> "Spherical":
        return Spherical(0, mass, 0.0)


@dataclass
class Disk:
    type: int = 0
    mass: float = 0.0
    scaleLength: float = 0.0
    scaleHeight: float = 0.0

    @staticmethod
    def miyamotoNagai(mass: float, scaleLength: float, scaleHeight: float) -> "Disk":
        return Disk(1, mass, scaleLength, scaleHeight)

    @staticmethod
    def none(mass: float = 0.0) -> "Disk":
        return Disk(0, mass)


@dataclass
class Halo:
    type: int = 0
    vhalo: float = 0.0
    scaleLength: float = 0.0
    flattenZ: float = 1.0

    @staticmethod
    def logarithmic(vhalo: float, scaleLength: float, flattenZ: float) -> "Halo":
        return Halo(1, vhalo, scaleLength, flattenZ)

    @staticmethod
    def nfw(vhalo: float, scaleLength: float) -> "Halo":
        return Halo(2, vhalo, scaleLength, 1.0)

    @staticmethod
    def none(mass: float = 0.0) -> "Halo":
        return Halo(0)


@dataclass
class Potential:
    spherical: Spherical = field(default_factory=Spherical)
    disk: Disk = field(default_factory=Disk)
    disk2: Disk = field(default_factory=Disk)
    halo: Halo = field(default_factory=Halo)

    @staticmethod
    def create(spherical=None, disk=None, disk2=None, halo=None) -> "Potential":
        return Potential(spherical or Spherical(), disk or Disk(), disk2 or Disk(), halo or Halo())

    def to_c(self) -> _Potential:
        p = _Potential()
        p.sphere_type, p.sphere_mass, p.sphere_scale = (
            self.spherical.type, self.spherical.mass, self.spherical.scale)
        p.disk_type, p.disk_mass = self.disk.type, self.disk.mass
        p.disk_scale_length, p.disk_scale_height = self.disk.scaleLength, self.disk.scaleHeight
        p.disk2_type, p.disk2_mass = self.disk2.type, self.disk2.mass
        p.disk2_scale_length, p.disk2_scale_height = self.disk2.scaleLength, self.disk2.scaleHeight
        p.halo_type, p.halo_vhalo = self.halo.type, self.halo.vhalo
        p.halo_scale_length, p.halo_flatten_z = self.halo.scaleLength, self.halo.flattenZ
        return p


def orbit_tangent(potential: Potential, pos, vel, t_rev: float, nsteps: int):
    """Reverse orbit with a fixed step count (step t_rev / nsteps) and its
    derivative with respect to t_rev: returns (pos, vel, dpos, dvel), each
    of 3 (nbhip_orbit_tangent; host code, no device needed)."""
    lib = load_library()
    p = potential.to_c()
    x = np.ascontiguousarray(pos, dtype=np.float64).reshape(3)
    v = np.ascontiguousarray(vel, dtype=np.float64).reshape(3)
    out = [np.empty(3) for _ in range(4)]
    rc = lib.nbhip_orbit_tangent(C.byref(p), _ptr(x), _ptr(v), float(t_rev), int(nsteps),
                                 *[_ptr(o) for o in out])
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())
    return tuple(out)


def milky_way_potential() -> Potential:
    """The sample workunits' Milky Way (nbody/sample_workunits/for_developers.lua:156-161):
    Hernquist bulge, Miyamoto-Nagai disk, logarithmic halo (vhalo 74.61 kpc/Gyr = 73 km/s)."""
    return Potential.create(
        spherical=Spherical.hernquist(mass=1.52954402e5, scale=0.7),
        disk=Disk.miyamotoNagai(mass=4.45865888e5, scaleLength=6.5, scaleHeight=0.26),
        disk2=Disk.none(mass=3.0e5),
        halo=Halo.logarithmic(vhalo=74.61, scaleLength=12.0, flattenZ=1.0),
    )


# ---------------------------------------------------------------------------
# NBodyCtx (nbody/src/nbody_lua_types/nbody_lua_nbodyctx.c:84-221)
# ---------------------------------------------------------------------------
@dataclass
class NBodyCtx:
    timestep: float
    timeEvolve: float
    eps2: float
    theta: float = 0.5
    treeRSize: float = 4.0
    criterion: str = "bh86"
    useQuad: bool = True
    allowIncest: bool = False
    nStep: int = 0

    @staticmethod
    def create(timestep: float, timeEvolve: float, eps2: float, theta: float = -1.0,
               treeRSize: float = 4.0, criterion: str = "BH86", useQuad: bool = True,
               allowIncest: bool = False) -> "NBodyCtx":
        crit = criterion.lower()
        if crit not in CRITERIA:
            raise ValueError(f"unknown criterion {criterion!r}")
        if crit != "exact" and theta < 0.0:
            raise ValueError("Theta argument required for criterion != 'Exact'")
        if crit == "exact":
            theta, useQuad = 0.0, False
        n_step = math.ceil(timeEvolve / timestep)
        if n_step >= 2 ** 32 - 1:
            raise ValueError("Number of timesteps exceeds UINT_MAX")
        # nbCorrectTimestep (nbody/src/nbody_util.c:26-30)
        corrected = timeEvolve / math.ceil(timeEvolve / timestep) if n_step > 0 else timestep
        return NBodyCtx(timestep=corrected, timeEvolve=timeEvolve, eps2=eps2, theta=theta,
                        treeRSize=treeRSize, criterion=crit, useQuad=useQuad,
                        allowIncest=allowIncest, nStep=int(n_step))

    def reference_image(self, potential: Optional[Potential] = None) -> bytes:
        """This context as the reference's NBodyCtx bytes (x86-64 layout, 592
        bytes), as a checkpoint header carries it: the fields this mirror
        holds and the potential; the rest (output switches, likelihood
        settings, LMC) are zero."""
        r = _RefNBodyCtx()
        r.eps2, r.theta, r.timestep = self.eps2, self.theta, self.timestep
        r.timeEvolve, r.treeRSize = self.timeEvolve, self.treeRSize
        r.criterion = CRITERIA[self.criterion]
        r.potentialType = 1 if potential is None else 0     # EXTERNAL_POTENTIAL_NONE / _DEFAULT
        r.useQuad = bytes([int(bool(self.useQuad))])
        r.allowIncest = bytes([int(bool(self.allowIncest))])
        r.nStep = self.nStep
        r.checkpointT = -1
        if potential is not None:
            sp, dk, d2, ha = potential.spherical, potential.disk, potential.disk2, potential.halo
            r.pot.sphere.type, r.pot.sphere.mass, r.pot.sphere.scale = sp.type, sp.mass, sp.scale
            r.pot.disk.type, r.pot.disk.mass = dk.type, dk.mass
            r.pot.disk.scaleLength, r.pot.disk.scaleHeight = dk.scaleLength, dk.scaleHeight
            r.pot.disk2.type, r.pot.disk2.mass = d2.type, d2.mass
            r.pot.disk2.scaleLength, r.pot.disk2.scaleHeight = d2.scaleLength, d2.scaleHeight
            r.pot.halo.type, r.pot.halo.vhalo = ha.type, ha.vhalo
            r.pot.halo.scaleLength, r.pot.halo.flattenZ = ha.scaleLength, ha.flattenZ
        return bytes(r)

    def to_c(self, potential: Optional[Potential], precision: str = "fp64") -> _Ctx:
        c = _Ctx()
        c.eps2, c.theta, c.timestep, c.tree_rsize = (self.eps2, self.theta, self.timestep,
                                                     self.treeRSize)
        c.criterion = CRITERIA[self.criterion]
        c.use_quad = int(bool(self.useQuad))
        c.allow_incest = int(bool(self.allowIncest))
        c.use_potential = 0 if potential is None else 1
        c.precision = PRECISION[precision]
        c.lmc = 0
        c.pot = (potential or Potential()).to_c()
        return c


class _CkptBuffers:
    """Keeps the arrays an nbhip_checkpoint_io points to alive."""

    def __init__(self, ctx_image, state_image, version, type_=None, id_=None, best=None,
                 orbit_trace=None, shift_lmc=None, lmc_posvel=None):
        self.ctx = C.create_string_buffer(bytes(ctx_image), len(ctx_image))
        self.state = C.create_string_buffer(bytes(state_image), max(len(state_image), 1))
        io = _CkptIO()
        io.major_version, io.minor_version = version
        io.ctx, io.ctx_size = C.cast(self.ctx, C.c_void_p), len(ctx_image)
        io.state, io.state_size = C.cast(self.state, C.c_void_p), len(state_image)
        self.keep = []

        def arr(a, dt, shape=None):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            if shape is not None:
                a = a.reshape(shape)
            self.keep.append(a)
            return a

        t = arr(type_, np.int16)
        i = arr(id_, np.uint32)
        io.type = _ptr(t, C.c_short)
        io.id = _ptr(i, C.c_uint)
        if best is not None:
            io.best_pos = _ptr(arr(best[0], np.float64))
            io.best_vel = _ptr(arr(best[1], np.float64))
        tr = arr(orbit_trace, np.float64, (-1, 4))
        if tr is not None:
            io.orbit_trace, io.n_orbit_trace = _ptr(tr), tr.shape[0]
        sh = arr(shift_lmc, np.float64, (-1, 4))
        if sh is not None:
            io.shift_lmc, io.n_shift_lmc = _ptr(sh), sh.shape[0]
            lp, lv = lmc_posvel
            io.lmc_pos[:] = [float(x) for x in lp]
            io.lmc_vel[:] = [float(x) for x in lv]
        self.io = io


def checkpoint_pack(path: str, nbody: int, step: int, rsize: float, tree_incest: int,
                    pos, vel, mass, ctx_image: bytes, state_image: bytes = b"",
                    version=NBODY_VERSION, **kw) -> None:
    """Writes a checkpoint file from host arrays (nbhip_checkpoint_pack)."""
    lib = load_library()
    h = _CkptHeader()
    h.nbody, h.step, h.rsize, h.tree_incest = nbody, step, rsize, tree_incest
    b = _CkptBuffers(ctx_image, state_image, version, **kw)
    x = np.ascontiguousarray(pos, dtype=np.float64)
    v = np.ascontiguousarray(vel, dtype=np.float64)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    rc = lib.nbhip_checkpoint_pack(os.fsencode(path), C.byref(h), C.byref(b.io), _ptr(x),
                                   _ptr(v), _ptr(m))
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())


def checkpoint_unpack(path: str, ctx_size: int = 592, state_size: int = 0,
                      version=NBODY_VERSION, max_vectors: int = 1 << 20) -> dict:
    """Reads a checkpoint file on the host (nbhip_checkpoint_unpack): header
    fields, the NBodyCtx / NBodyState images, bodies and vectors."""
    lib = load_library()
    h = _CkptHeader()
    b = _CkptBuffers(bytes(ctx_size), bytes(state_size), version)
    rc = lib.nbhip_checkpoint_unpack(os.fsencode(path), C.byref(h), C.byref(b.io), None, None,
                                     None)
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())
    n = h.nbody
    out = {k: getattr(h, k) for k, _ in _CkptHeader._fields_}
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    bpos, bvel = np.empty((n, 3)), np.empty((n, 3))
    typ, ids = np.empty(n, np.int16), np.empty(n, np.uint32)
    tr = np.empty((max(h.n_orbit_trace, 1), 4))
    sh = np.empty((max(h.n_shift_lmc, 1), 4))
    b = _CkptBuffers(bytes(ctx_size), bytes(state_size), version)
    b.io.type, b.io.id = _ptr(typ, C.c_short), _ptr(ids, C.c_uint)
    b.io.best_pos, b.io.best_vel = _ptr(bpos), _ptr(bvel)
    b.io.orbit_trace, b.io.n_orbit_trace = _ptr(tr), h.n_orbit_trace
    b.io.shift_lmc, b.io.n_shift_lmc = _ptr(sh), h.n_shift_lmc
    rc = lib.nbhip_checkpoint_unpack(os.fsencode(path), C.byref(h), C.byref(b.io), _ptr(pos),
                                     _ptr(vel), _ptr(mass))
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())
    out.update(pos=pos, vel=vel, mass=mass, best_pos=bpos, best_vel=bvel, type=typ, id=ids,
               ctx=b.ctx.raw[:ctx_size], state=b.state.raw[:b.io.state_size],
               orbit_trace=tr[:h.n_orbit_trace], shift_lmc=sh[:h.n_shift_lmc],
               lmc_pos=np.array(b.io.lmc_pos[:]), lmc_vel=np.array(b.io.lmc_vel[:]))
    return out


def dwarf_params(nbody: int, nbody_baryon: int, light=("plummer", 12.0, 0.2),
                 dark=("nfw", 60.0, 0.8), seed: int = 34086709, position=(0.0, 0.0, 0.0),
                 velocity=(0.0, 0.0, 0.0)) -> _DwarfParams:
    """predefinedModels.mixeddwarf's arguments: each component as (type,
    mass, scale radius), an NFW mass being M200."""
    p = _DwarfParams()
    p.light_type, p.light_mass, p.light_scale = DWARF_TYPES[light[0]], light[1], light[2]
    p.dark_type, p.dark_mass, p.dark_scale = DWARF_TYPES[dark[0]], dark[1], dark[2]
    p.nbody, p.nbody_baryon, p.seed = int(nbody), int(nbody_baryon), int(seed)
    p.pos[:] = [float(x) for x in position]
    p.vel[:] = [float(x) for x in velocity]
    return p


def mixed_dwarf(nbody: int, nbody_baryon: int, device: int = 0, **kw):
    """Two-component dwarf sampled on the GPU (nbhip_mixed_dwarf): returns
    (pos (n,3), vel (n,3), mass (n,), type (n,), component masses (2,));
    baryons first (type 1), dark matter after (type -1)."""
    lib = load_library()
    p = dwarf_params(nbody, nbody_baryon, **kw)
    pos, vel, m, cm = np.empty((nbody, 3)), np.empty((nbody, 3)), np.empty(nbody), np.empty(2)
    rc = lib.nbhip_mixed_dwarf(C.byref(p), device, _ptr(pos), _ptr(vel), _ptr(m), _ptr(cm))
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())
    typ = np.where(np.arange(nbody) < nbody_baryon, 1, -1).astype(np.int32)
    return pos, vel, m, typ, cm


def dwarf_df(energy, comp: int, device: int = 0, **kw) -> np.ndarray:
    """The sampler's tabulated distribution function of component comp
    (0 light, 1 dark) at the given energies (nbhip_dwarf_df)."""
    lib = load_library()
    p = dwarf_params(kw.pop("nbody", 0), kw.pop("nbody_baryon", 0), **kw)
    e = np.ascontiguousarray(energy, dtype=np.float64).ravel()
    f = np.empty_like(e)
    rc = lib.nbhip_dwarf_df(C.byref(p), device, comp, e.size, _ptr(e), _ptr(f))
    if rc:
        raise NBodyError(rc, lib.nbhip_last_error().decode())
    return f


def philox4x32_10(ctr, key) -> tuple:
    """One Philox4x32-10 block from the library (host code)."""
    lib = load_library()
    c = (C.c_uint * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint * 4)()
    lib.nbhip_philox4x32_10(c, k, o)
    return tuple(o)


@dataclass
class TreeInfo:
    rsize: float
    ncell: int
    max_depth: int
    nnode: int
    nmassive: int
    step: int
    status: int


class NBodyState:
    """A simulation resident on one GPU (mirror of the Lua NBodyState,
    nbody/src/nbody_lua_types/nbody_lua_nbodystate.c)."""

    def __init__(self, ctx: NBodyCtx, potential: Optional[Potential], pos, vel, mass,
                 device: int = 0, precision: str = "fp64", prestep: bool = True):
        self.lib = load_library()
        self.ctx = ctx
        self.potential = potential
        self.precision = precision
        self.n = int(np.asarray(mass).shape[0])
        h = C.c_void_p()
        cctx = ctx.to_c(potential, precision)
        rc = self.lib.nbhip_create(C.byref(cctx), device, self.n, C.byref(h))
        if rc:
            raise NBodyError(rc, self.lib.nbhip_last_error().decode())
        self._h = h
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(self.n)
        self._check(self.lib.nbhip_upload(h, _ptr(pos), _ptr(vel), _ptr(mass)))
        self.last_status = 0
        if prestep:
            # NBodyState.create runs one force evaluation (nbody_lua_nbodystate.c:123-147)
            self.last_status = self._check(self.lib.nbhip_gravity(h), allow_warning=True)

    @staticmethod
    def create(ctx: NBodyCtx, potential: Optional[Potential], bodies, **kw) -> "NBodyState":
        pos, vel, mass = bodies
        return NBodyState(ctx, potential, pos, vel, mass, **kw)

    def _check(self, rc: int, allow_warning: bool = True) -> int:
        if rc and (status_is_fatal(rc) or not allow_warning):
            raise NBodyError(rc, self.lib.nbhip_last_error().decode())
        return rc

    def sample_mixed_dwarf(self, sim: int = 0, **kw):
        """Samples a two-component dwarf straight into simulation `sim`
        (nbhip_mixed_dwarf_state; dark bodies flagged as ignored).  Takes
        dwarf_params' keywords; returns the component masses."""
        nsim_n = self.n if getattr(self, "nsim", 1) == 1 else self.counts[sim]
        nb = kw.pop("nbody_baryon", nsim_n // 2)
        p = dwarf_params(nsim_n, nb, **kw)
        cm = np.empty(2)
        self._check(self.lib.nbhip_mixed_dwarf_state(self._h, int(sim), C.byref(p), _ptr(cm)),
                    allow_warning=False)
        return cm

    def write_checkpoint(self, path: str, state_image: bytes = bytes(NBODY_STATE_SIZE),
                         version=NBODY_VERSION, **kw) -> None:
        """nbWriteCheckpoint of this GPU state (nbhip_checkpoint_write): the
        context's NBodyCtx image, `state_image` as the raw NBodyState."""
        b = _CkptBuffers(self.ctx.reference_image(self.potential), state_image, version, **kw)
        self._check(self.lib.nbhip_checkpoint_write(self._h, os.fsencode(path), C.byref(b.io)),
                    allow_warning=False)

    @staticmethod
    def from_checkpoint(path: str, ctx: NBodyCtx, potential: Optional[Potential],
                        device: int = 0, precision: str = "fp64", version=NBODY_VERSION,
                        state_size: int = NBODY_STATE_SIZE) -> "NBodyState":
        """nbReadCheckpoint into a new GPU state (nbhip_checkpoint_read); no
        force evaluation: the next run() starts with one, as on resume.  The
        file's NBodyCtx and NBodyState images are kept as .checkpoint_ctx and
        .checkpoint_state."""
        head = checkpoint_unpack(path, state_size=state_size, version=version)
        n = int(head["nbody"])
        st = NBodyState(ctx, potential, np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n),
                        device=device, precision=precision, prestep=False)
        b = _CkptBuffers(bytes(592), bytes(state_size), version)
        st._check(st.lib.nbhip_checkpoint_read(st._h, os.fsencode(path), C.byref(b.io)),
                  allow_warning=False)
        st.checkpoint_ctx = b.ctx.raw[:592]
        st.checkpoint_state = b.state.raw[:state_size]
        return st

    def gravity(self) -> int:
        self.last_status = self._check(self.lib.nbhip_gravity(self._h))
        return self.last_status

    def step(self) -> int:
        self.last_status = self._check(self.lib.nbhip_step(self._h))
        return self.last_status

    def steps(self, nsteps: int) -> int:
        self.last_status = self._check(self.lib.nbhip_steps(self._h, nsteps))
        return self.last_status

    def run(self, nsteps: Optional[int] = None) -> int:
        """nbRunSystem: one force evaluation then nStep KDK steps."""
        k = self.ctx.nStep if nsteps is None else nsteps
        self.last_status = self._check(self.lib.nbhip_run(self._h, k))
        return self.last_status

    def run_system(self, likelihood=None, checkpoint=None, nstep: Optional[int] = None,
                   use_best_like: bool = True, best_like_start: float = 0.95,
                   keep_best_bodies: bool = False, chunk: int = 64):
        """nbRunSystem with its hooks (nbhip_run_system): the force, then KDK
        steps to nstep (ctx.nStep by default).  likelihood(step) returns the
        10-entry nbSystemLikelihood array of the current state (NaN where
        unknown); checkpoint(step) may write a checkpoint and returns a status.
        Returns (status, best) with best = dict(lk, time, step, count,
        evaluated)."""
        failure = []

        def _like(user, h, step, out):
            try:
                arr = likelihood(int(step))
                for i in range(10):
                    out[i] = float(arr[i]) if i < len(arr) else float("nan")
                return 0
            except Exception as e:          # surfaced after the run
                failure.append(e)
                return NBODY_ERROR

        def _ckpt(user, h, step):
            try:
                return int(checkpoint(int(step)) or 0)
            except Exception as e:
                failure.append(e)
                return NBODY_ERROR

        hooks = _RunHooks(int(self.ctx.nStep if nstep is None else nstep), int(use_best_like),
                          float(best_like_start), float(self.ctx.timeEvolve),
                          int(keep_best_bodies), int(chunk),
                          _LIKE_FN(_like) if likelihood else _LIKE_FN(),
                          _CKPT_FN(_ckpt) if checkpoint else _CKPT_FN(), None)
        best = _BestLike()
        rc = self.lib.nbhip_run_system(self._h, C.byref(hooks), C.byref(best))
        if failure:
            raise failure[0]
        self.last_status = self._check(rc)
        return rc, {"lk": list(best.lk), "time": best.time, "step": best.step,
                    "count": best.count, "evaluated": best.evaluated}

    def best_bodies(self):
        """(pos, vel) of the best-likelihood step kept by run_system."""
        pos, vel = np.empty((self.n, 3)), np.empty((self.n, 3))
        self._check(self.lib.nbhip_best_bodies(self._h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def bodies(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        acc = np.empty((self.n, 3))
        self._check(self.lib.nbhip_download(self._h, _ptr(pos), _ptr(vel), _ptr(acc)))
        return pos, vel, acc

    def accelerations(self) -> np.ndarray:
        return self.bodies()[2]

    def restore(self, step: int, rsize: float, status: int = 0) -> None:
        """Resume from a checkpoint (nbReadCheckpointHeader): after uploading
        the checkpointed bodies, set the step counter, the persistent root
        size and sticky status bits; the next run() continues bit for bit."""
        self._check(self.lib.nbhip_restore(self._h, 0, int(step), float(rsize), int(status)))

    def tree_info(self) -> TreeInfo:
        ti = _TreeInfo()
        self._check(self.lib.nbhip_tree_info_get(self._h, C.byref(ti)))
        return TreeInfo(ti.rsize, ti.ncell, ti.max_depth, ti.nnode, ti.nmassive, ti.step,
                        ti.status)

    def tree(self) -> dict:
        nn = self.tree_info().nnode
        out = {
            "kind": np.empty(nn, np.int32), "ref": np.empty(nn, np.int32),
            "skip": np.empty(nn, np.int32), "level": np.empty(nn, np.int32),
            "pos": np.empty((nn, 3)), "mass": np.empty(nn), "rcrit2": np.empty(nn),
            "quad": np.empty((nn, 6)),
        }
        self._check(self.lib.nbhip_tree_download(
            self._h, _ptr(out["kind"], C.c_int), _ptr(out["ref"], C.c_int),
            _ptr(out["skip"], C.c_int), _ptr(out["level"], C.c_int), _ptr(out["pos"]),
            _ptr(out["mass"]), _ptr(out["rcrit2"]), _ptr(out["quad"])))
        return out

    def direct_accel(self, idx) -> np.ndarray:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.empty((idx.shape[0], 3))
        self._check(self.lib.nbhip_direct_accel(self._h, idx.shape[0], _ptr(idx, C.c_int),
                                                _ptr(acc)))
        return acc

    def walk_accel(self) -> np.ndarray:
        acc = np.empty((self.n, 3))
        self._check(self.lib.nbhip_walk_accel(self._h, _ptr(acc)))
        return acc

    def walk_counters(self):
        """(body-body, body-cell, opened, wave node steps) of a counting walk."""
        c = (C.c_ulonglong * 4)()
        self._check(self.lib.nbhip_walk_counters(self._h, c))
        return int(c[0]), int(c[1]), int(c[2]), int(c[3])

    def walk_counters5(self):
        """walk_counters() and the group walk's far-field (cell, group) expansions."""
        c = (C.c_ulonglong * 5)()
        self._check(self.lib.nbhip_walk_counters5(self._h, c))
        return tuple(int(x) for x in c)

    def set_group_expansion(self, ratio: float) -> None:
        """Group walk only: evaluate an accepted cell once per group, as a
        Taylor expansion about the group's centre, when the group's half
        box diagonal is at most `ratio` times its distance (0: off)."""
        self._check(self.lib.nbhip_set_group_expansion(self._h, float(ratio)))

    def set_graph(self, enable: bool = True) -> None:
        """Replay steps as a captured hipGraph (no host sync inside a step)."""
        self._check(self.lib.nbhip_set_graph(self._h, int(enable)))

    def autotune(self):
        """Time every walk setting on the current tree; keep the fastest.
        Returns (best (block, batch, ms), [all trials])."""
        best = _Tune()
        trials = (_Tune * 32)()
        cnt = C.c_int(0)
        self._check(self.lib.nbhip_autotune(self._h, C.byref(best), trials, 32, C.byref(cnt)))
        return ((best.block, best.batch, best.ms),
                [(trials[i].block, trials[i].batch, trials[i].ms) for i in range(cnt.value)])

    def set_open_scale(self, scale: float) -> None:
        """Coherent mixed walk only: scale every cell's squared critical
        radius by `scale` (>= 1), so the walk opens more and each body's list
        refines its reference list further; from the next tree build."""
        self._check(self.lib.nbhip_set_open_scale(self._h, float(scale)))

    def set_tune(self, block: int, batch: int = 1024) -> None:
        t = _Tune(block, batch, 0.0)
        self._check(self.lib.nbhip_set_tune(self._h, C.byref(t)))

    def set_ignore(self, ignore) -> None:
        """Per-body flags (upload order, all simulations): True = left out
        of histograms (the dark matter of a mixed dwarf); None clears."""
        if ignore is None:
            self._check(self.lib.nbhip_set_ignore(self._h, None))
            return
        a = np.ascontiguousarray(np.asarray(ignore).reshape(self.n) != 0, dtype=np.uint8)
        self._check(self.lib.nbhip_set_ignore(self._h, _ptr(a, C.c_ubyte)))

    def histogram(self, hp):
        """Raw lambda-beta counts on the GPU for each simulation: returns
        (counts [nsim, lambdaBins * betaBins], totals [nsim]); hp is a
        histogram.HistogramParams."""
        h = _HistParams(hp.phi, hp.theta, hp.psi, hp.lambdaStart, hp.lambdaEnd, hp.betaStart,
                        hp.betaEnd, hp.lambdaBins, hp.betaBins, hp.sunGCDist)
        nsim = getattr(self, "nsim", 1)
        counts = np.zeros((nsim, hp.lambdaBins * hp.betaBins), dtype=np.uint32)
        totals = np.zeros(nsim, dtype=np.uint32)
        self._check(self.lib.nbhip_histogram(self._h, C.byref(h), _ptr(counts, C.c_uint),
                                             _ptr(totals, C.c_uint)))
        return counts, totals

    def histogram_dispersion(self, hp, sim: int = 0, iter_max: int = 6, beta_sigma: float = 2.5,
                             vel_sigma: float = 2.5, dist_sigma: float = 2.5,
                             pm_sigma: float = 2.5, beta_correct: float = 1.111,
                             vel_correct: float = 1.111, dist_correct: float = 1.111,
                             pm_correct: float = 1.111, sun_vel=SUN_VEL, ngp_dec: float = NGP_DEC,
                             ngp_ra: float = NGP_RA, l_ncp: float = L_NCP) -> dict:
        """nbCreateHistogram's histograms 1-7 for simulation `sim` after the
        outlier rejection: {1: beta dispersion, 2: v_los dispersion, 3: v_los
        average, 4: beta average, 5: distance average, 6: mu_dec average,
        7: mu_ra average}, each (value, error) per bin.  The sigmas and
        corrections default to the sample workunit's (SigmaCutoff 2.5,
        SigmaIter 6, Correction 1.111, nbody/sample_workunits/
        for_developers.lua:93-95); the angles are in degrees."""
        h = _HistParams(hp.phi, hp.theta, hp.psi, hp.lambdaStart, hp.lambdaEnd, hp.betaStart,
                        hp.betaEnd, hp.lambdaBins, hp.betaBins, hp.sunGCDist)
        d2r = math.pi / 180.0
        d = _DispParams(int(iter_max), float(beta_sigma), float(vel_sigma), float(dist_sigma),
                        float(pm_sigma), float(beta_correct), float(vel_correct),
                        float(dist_correct), float(pm_correct),
                        (C.c_double * 3)(*[float(v) for v in sun_vel]), ngp_dec * d2r,
                        ngp_ra * d2r, l_ncp * d2r)
        nbin = hp.lambdaBins * hp.betaBins
        out = np.empty(14 * nbin)
        self._check(self.lib.nbhip_histogram_dispersion(self._h, int(sim), C.byref(h),
                                                        C.byref(d), _ptr(out)))
        out = out.reshape(14, nbin)
        return {k: (out[2 * (k - 1)].copy(), out[2 * (k - 1) + 1].copy()) for k in range(1, 8)}

    def gravity_vjp(self, gacc):
        """(dL/dpos [n,3], dL/dmass [n]) of one force evaluation at the current
        positions, tree frozen, from dL/dacc [n,3]."""
        g = np.ascontiguousarray(gacc, dtype=np.float64).reshape(self.n, 3)
        gp, gm = np.empty((self.n, 3)), np.empty(self.n)
        self._check(self.lib.nbhip_gravity_vjp(self._h, _ptr(g), _ptr(gp), _ptr(gm)))
        return gp, gm

    def adjoint_forward(self, nsteps: int, every: int = 50) -> int:
        """Force evaluation + nsteps KDK steps (as run()), checkpointing the
        state every `every` steps for adjoint_backward."""
        self.last_status = self._check(self.lib.nbhip_adjoint_forward(self._h, nsteps, every))
        return self.last_status

    def adjoint_backward(self, gpos, gvel):
        """dL/d(initial pos, vel, mass, timestep) from dL/d(final pos, vel).
        Returns (gpos0 [n,3], gvel0 [n,3], gmass [n], gdt [nsim])."""
        gpos = np.ascontiguousarray(gpos, dtype=np.float64).reshape(self.n, 3)
        gvel = np.ascontiguousarray(gvel, dtype=np.float64).reshape(self.n, 3)
        gp0, gv0 = np.empty((self.n, 3)), np.empty((self.n, 3))
        gm = np.empty(self.n)
        gdt = np.empty(getattr(self, "nsim", 1))
        self._check(self.lib.nbhip_adjoint_backward(self._h, _ptr(gpos), _ptr(gvel), _ptr(gp0),
                                                    _ptr(gv0), _ptr(gm), _ptr(gdt)))
        return gp0, gv0, gm, gdt

    def set_profiling(self, enable: bool = True) -> None:
        self._check(self.lib.nbhip_set_profiling(self._h, int(enable)))

    def phase_times(self) -> np.ndarray:
        ms = np.zeros(5)
        self._check(self.lib.nbhip_phase_times(self._h, _ptr(ms)))
        return ms

    def sync(self) -> None:
        self._check(self.lib.nbhip_sync(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.nbhip_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class NBodyBatch(NBodyState):
    """Many independent simulations in one state (same context and potential,
    per-simulation softening): every phase of a step is one launch."""

    def __init__(self, ctx: NBodyCtx, potential: Optional[Potential], bodies, eps2=None,
                 device: int = 0, precision: str = "fp64", prestep: bool = True):
        self.lib = load_library()
        self.ctx = ctx
        self.potential = potential
        self.precision = precision
        self.counts = [int(np.asarray(b[2]).shape[0]) for b in bodies]
        self.nsim = len(bodies)
        self.n = sum(self.counts)
        cnt = np.asarray(self.counts, dtype=np.int32)
        e2 = None if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64)
        h = C.c_void_p()
        cctx = ctx.to_c(potential, precision)
        rc = self.lib.nbhip_batch_create(C.byref(cctx), device, self.nsim,
                                         _ptr(cnt, C.c_int), _ptr(e2), C.byref(h))
        if rc:
            raise NBodyError(rc, self.lib.nbhip_last_error().decode())
        self._h = h
        for k, (pos, vel, mass) in enumerate(bodies):
            n = self.counts[k]
            pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
            vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
            mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(n)
            self._check(self.lib.nbhip_batch_upload(h, k, _ptr(pos), _ptr(vel), _ptr(mass)))
        self.last_status = 0
        if prestep:
            self.last_status = self._check(self.lib.nbhip_gravity(h), allow_warning=True)

    def sim_bodies(self, k: int):
        n = self.counts[k]
        pos, vel, acc = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        self._check(self.lib.nbhip_batch_download(self._h, k, _ptr(pos), _ptr(vel), _ptr(acc)))
        return pos, vel, acc

    def restore_sim(self, k: int, step: int, rsize: float, status: int = 0) -> None:
        """restore() for simulation k (the step counter is shared)."""
        self._check(self.lib.nbhip_restore(self._h, int(k), int(step), float(rsize), int(status)))

    def statuses(self):
        out = (C.c_uint * self.nsim)()
        self._check(self.lib.nbhip_batch_status(self._h, out))
        return [int(x) for x in out]

    def tree(self, sim: int = 0) -> dict:
        """The pre-order tree of simulation `sim` alone (skip and ref counted
        from its own root and first body, as in its single run)."""
        nn = self.sim_tree_info(sim).nnode
        out = {
            "kind": np.empty(nn, np.int32), "ref": np.empty(nn, np.int32),
            "skip": np.empty(nn, np.int32), "level": np.empty(nn, np.int32),
            "pos": np.empty((nn, 3)), "mass": np.empty(nn), "rcrit2": np.empty(nn),
            "quad": np.empty((nn, 6)),
        }
        self._check(self.lib.nbhip_batch_tree_download(
            self._h, int(sim), _ptr(out["kind"], C.c_int), _ptr(out["ref"], C.c_int),
            _ptr(out["skip"], C.c_int), _ptr(out["level"], C.c_int), _ptr(out["pos"]),
            _ptr(out["mass"]), _ptr(out["rcrit2"]), _ptr(out["quad"])))
        return out

    def sim_tree_info(self, k: int) -> TreeInfo:
        ti = _TreeInfo()
        self._check(self.lib.nbhip_batch_tree_info(self._h, k, C.byref(ti)))
        return TreeInfo(ti.rsize, ti.ncell, ti.max_depth, ti.nnode, ti.nmassive, ti.step,
                        ti.status)

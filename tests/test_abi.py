"""The C-ABI libraries load and export every symbol their headers declare
(no GPU compute: only argument validation that precedes device access)."""
import ctypes
import os
import re

import pytest

from milkywayathome_client_amd import models, nbody as nb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:nbhip|nbm)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("nbhip.h", nb.load_library),
                                           ("nbmodels.h", models.load_models)])
def test_library_exports_every_declared_symbol(header, loader):
    lib = loader()
    names = _declared(header)
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(nb.EXPORTED) == set(_declared("nbhip.h"))


def test_context_validation_happens_before_device_access():
    lib = nb.load_library()
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5)
    h = ctypes.c_void_p()
    c = ctx.to_c(None)
    c.tree_rsize = 3.0                      # not a power of two
    assert lib.nbhip_create(ctypes.byref(c), 0, 10, ctypes.byref(h)) == nb.NBODY_UNSUPPORTED
    assert b"power of two" in lib.nbhip_last_error()
    c = ctx.to_c(None)
    c.criterion = 7
    assert lib.nbhip_create(ctypes.byref(c), 0, 10, ctypes.byref(h)) == nb.NBODY_USER_ERROR
    c = ctx.to_c(nb.milky_way_potential())
    c.pot.halo_type = 3                     # triaxial halo: not supported
    assert lib.nbhip_create(ctypes.byref(c), 0, 10, ctypes.byref(h)) == nb.NBODY_UNSUPPORTED
    assert lib.nbhip_create(None, 0, 10, ctypes.byref(h)) == nb.NBODY_USER_ERROR


def test_status_codes_match_nbodystatus():
    # nbody/include/nbody_types.h:541-577
    assert nb.NBODY_TREE_STRUCTURE_ERROR == 8
    assert nb.NBODY_TREE_INCEST_FATAL == 16
    assert nb.NBODY_CELL_OVERFLOW_ERROR == 1 << 17
    assert nb.NBODY_TREE_INCEST_NONFATAL == 1 << 24
    assert not nb.status_is_fatal(nb.NBODY_TREE_INCEST_NONFATAL)
    assert nb.status_is_fatal(nb.NBODY_TREE_INCEST_FATAL)


def test_lmc_context_is_rejected_before_device_access():
    lib = nb.load_library()
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5)
    h = ctypes.c_void_p()
    c = ctx.to_c(nb.milky_way_potential())
    c.lmc = 1                               # NBodyCtx.LMC: not supported on the GPU
    assert lib.nbhip_create(ctypes.byref(c), 0, 10, ctypes.byref(h)) == nb.NBODY_UNSUPPORTED
    assert b"LMC" in lib.nbhip_last_error()

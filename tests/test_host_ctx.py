"""Host-side context rules (NBodyCtx.create, nbody/src/nbody_lua_types/
nbody_lua_nbodyctx.c:84-221; nbCorrectTimestep, nbody/src/nbody_util.c:26-30)."""
import math

import pytest

from milkywayathome_client_amd import nbody as nb


@pytest.mark.parametrize("dt,evolve,nstep", [(1e-4, 0.02, 200), (1e-3, 4.0, 4000),
                                             (1e-3, 2.0, 2000), (1e-3, 1.0, 1000)])
def test_step_counts_of_the_configs(dt, evolve, nstep):
    c = nb.NBodyCtx.create(timestep=dt, timeEvolve=evolve, eps2=1e-4, theta=0.5)
    assert c.nStep == nstep
    assert c.timestep == evolve / nstep


def test_timestep_is_corrected_to_cover_evolve_time():
    c = nb.NBodyCtx.create(timestep=0.3, timeEvolve=1.0, eps2=1e-4, theta=0.5)
    assert c.nStep == 4 and math.isclose(c.timestep * c.nStep, 1.0)


def test_exact_criterion_forces_theta_zero_no_quad():
    c = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, criterion="Exact")
    assert c.theta == 0.0 and c.useQuad is False


def test_theta_required():
    with pytest.raises(ValueError):
        nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4)


def _batch_create(counts, ctx=None):
    """nbhip_batch_create through the C ABI; the size checks come before
    any device call, so this runs without a GPU for rejected requests."""
    import ctypes as C
    import numpy as np
    lib = nb.load_library()
    c = (ctx or nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4,
                                   theta=0.5)).to_c(None)
    arr = np.ascontiguousarray(counts, dtype=np.int32)
    h = C.c_void_p()
    rc = lib.nbhip_batch_create(C.byref(c), 0, len(arr),
                                arr.ctypes.data_as(C.POINTER(C.c_int)), None, C.byref(h))
    return rc, lib.nbhip_last_error().decode(), h


def test_state_size_limits_are_rejected_before_any_device_call():
    rc, msg, h = _batch_create([1 << 24, 1])
    assert rc == nb.NBODY_USER_ERROR and "2^24" in msg and not h.value
    rc, msg, _ = _batch_create([10] * 4097)
    assert rc == nb.NBODY_USER_ERROR and "4096" in msg
    rc, msg, _ = _batch_create([5, -1])
    assert rc == nb.NBODY_USER_ERROR and "negative" in msg


def test_unsupported_context_is_rejected_before_any_device_call():
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5,
                             treeRSize=3.0)
    rc, msg, _ = _batch_create([100], ctx)
    assert rc == nb.NBODY_UNSUPPORTED and "power of two" in msg

"""Host-side parts of the GPU dwarf sampler (no GPU): the Philox4x32-10
generator against its published known-answer vectors (Salmon et al., SC'11;
the Random123 distribution's kat_vectors), and parameter checks that come
before any device access."""
import ctypes as C

import pytest

from milkywayathome_client_amd import nbody as nb


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, want):
    assert nb.philox4x32_10(ctr, key) == want


@pytest.mark.parametrize("kw,status", [
    (dict(light=("plummer", 12.0, 0.2), dark=("nfw", 60.0, 0.8), nbody=10, nbody_baryon=11),
     nb.NBODY_USER_ERROR),
    (dict(light=("plummer", 12.0, 0.0), dark=("nfw", 60.0, 0.8), nbody=10, nbody_baryon=5),
     nb.NBODY_USER_ERROR),
])
def test_bad_parameters_are_rejected_before_the_device(kw, status):
    lib = nb.load_library()
    n, nbar = kw.pop("nbody"), kw.pop("nbody_baryon")
    p = nb.dwarf_params(n, nbar, **kw)
    out = (C.c_double * (3 * max(n, 1)))()
    rc = lib.nbhip_mixed_dwarf(C.byref(p), 0, out, out, out, None)
    assert rc == status


def test_unsupported_component_type():
    lib = nb.load_library()
    p = nb.dwarf_params(10, 5)
    p.dark_type = 4                       # Cored: not supported
    out = (C.c_double * 30)()
    assert lib.nbhip_mixed_dwarf(C.byref(p), 0, out, out, out, None) == nb.NBODY_UNSUPPORTED

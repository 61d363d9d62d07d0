"""dSFMT-19937 stream against the reference's known-answer file
(dSFMT/dSFMT.19937.out.txt, extracted to tests/golden/dsfmt_19937_kat.json)."""
import json
import os

import numpy as np
import pytest

from milkywayathome_client_amd.models import DSFMT

with open(os.path.join(os.path.dirname(__file__), "golden", "dsfmt_19937_kat.json")) as _f:
    KAT = json.load(_f)


@pytest.mark.parametrize("block", KAT["blocks"], ids=lambda b: f"seed{b['seed']}-{b['kind']}")
def test_dsfmt_known_answers(block):
    got = DSFMT(block["seed"]).draw(len(block["values"]), block["kind"])
    # the file prints 15 decimals
    np.testing.assert_allclose(got, block["values"], rtol=0, atol=6e-16)


def test_kat_has_long_stream():
    assert max(len(b["values"]) for b in KAT["blocks"]) == 1000


def test_ranges():
    d = DSFMT(12345)
    a = d.draw(20000, "open_open")
    assert a.min() > 0.0 and a.max() < 1.0
    b = DSFMT(12345).draw(20000, "close_open")
    assert b.min() >= 0.0 and b.max() < 1.0
    # open_open sets the lowest mantissa bit of the same [1,2) draw
    assert np.all(np.abs(a - b) < 1e-15)

"""The drop-in from a C client's side: tests/c/run_system_client.c, a plain C
program built by gcc against include/nbhip.h, creates a state, uploads the
bodies, runs nbRunSystemPlain's loop through nbhip_run_system (best
likelihood from BestLikeStart on, checkpoint hook, graph-replayed chunks)
and downloads the bodies (nbody/src/nbody_plain.c:366-461,
nbody/src/nbody.c:417-497).  Its `gpu` mode checks the final bodies against
the oracle bit for bit and the best likelihood, its step, time, count and
kept bodies against a step-by-step loop; its `cpu` mode checks the context
validation that needs no device (LMC, treeRSize, potential, precision)."""
import os
import subprocess

import pytest

from conftest import C_CLIENT


def _run(mode, timeout):
    assert os.path.exists(C_CLIENT), "build it with `make -C tests/c`"
    res = subprocess.run([C_CLIENT, mode], capture_output=True, text=True, timeout=timeout)
    print(res.stdout, res.stderr, flush=True)
    return res


def test_c_client_context_checks():
    res = _run("cpu", 60)
    assert res.returncode == 0, res.stdout + res.stderr


@pytest.mark.gpu
def test_c_client_run_system_matches_oracle():
    print("C client: create, upload, run_system (best likelihood), download", flush=True)
    res = _run("gpu", 180)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "gpu: 0 failure(s)" in res.stdout

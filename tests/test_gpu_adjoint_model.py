"""The stream-likelihood gradient through the torch.autograd.Function over
the HIP adjoint (milkywayathome_client_amd/adjoint.py), against central
differences (h = 1e-5 relative) of the same model's forward runs, for the
six workunit parameters.  With the Exact criterion the forward map is smooth
in the parameters, so the two must agree closely; the full-size workload
(20 000 bodies, 1000 steps) is measured by `bench.py --workload adjoint`."""
import numpy as np
import pytest
import torch

from milkywayathome_client_amd import adjoint as adj

pytestmark = pytest.mark.gpu


def test_stream_gradient_matches_central_differences():
    p_star = np.array([0.2, 1.0, 0.2, 0.2, 12.0, 0.2])
    model = adj.make_model(n=5000, p_ref=p_star, criterion="exact", every=25)
    with torch.no_grad():
        target = model.histogram(torch.tensor(p_star), grad=False)
    rng = np.random.default_rng(11)
    p = p_star + rng.normal(size=6) * 0.02 * np.abs(p_star)
    L, g = model.value_and_grad(p, target)
    fd = np.zeros(6)
    for k in range(6):
        h = 1e-5 * abs(p[k])
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        with torch.no_grad():
            fd[k] = (float(model.loss(torch.tensor(a), target, grad=False))
                     - float(model.loss(torch.tensor(b), target, grad=False))) / (2 * h)
    rel = np.abs(g - fd) / np.abs(fd)
    print("loss", L, "\nadjoint", g, "\ncentral", fd, "\nrel", rel)
    assert L > 0 and np.all(np.isfinite(g))
    assert rel.max() <= 1e-3

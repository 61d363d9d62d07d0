"""Shared test helpers: workloads and error metrics."""
import numpy as np

from milkywayathome_client_amd.models import DSFMT, plummer


def plummer_bodies(n, seed=42, mass=12.0, radius=0.2, position=(0.0, 0.0, 0.0),
                   velocity=(0.0, 0.0, 0.0)):
    pos, vel, m, _ = plummer(n, DSFMT(seed), mass, radius, position, velocity)
    return pos, vel, m


def rel_err(a, ref, floor_frac=1e-3):
    """Per-body |a - ref| / max(|ref|, floor_frac * median |ref|).  The floor
    keeps bodies whose net self-gravity nearly cancels from dominating."""
    na = np.linalg.norm(ref, axis=1)
    den = np.maximum(na, floor_frac * np.median(na))
    return np.linalg.norm(a - ref, axis=1) / den

"""MI355X-native Barnes-Hut force step and KDK leapfrog for the Milkyway@Home
nbody client.  The compute path is ``libnbhip.so`` (HIP, gfx950) behind a
plain C ABI (``include/nbhip.h``); this package is its host-side mirror of the
reference's run interface plus the initial-condition layer."""
from .nbody import (NBodyCtx, NBodyState, Potential, Spherical, Disk, Halo,  # noqa: F401
                    milky_way_potential, load_library, NBodyError, status_is_fatal)
from .models import DSFMT, plummer  # noqa: F401

"""heatsim2_amd: heatsim2's 3-D ADI Crank-Nicolson heat-conduction step on
AMD Instinct MI355X (gfx950), in FP64.

Drop-in surface of the upstream package (heatsim2/__init__.py): the kind
constants, grid builders, zero_elements, setup(), run_adi_steps(), the four
boundary face-law modules, surface_temperature and hs2_indexing.  The step
runs in hand-written HIP kernels (libhs2.so, C ABI in
include/heatsim2_amd.h); there is no CPU execution path.
"""
from heatsim2_amd.constants import (IMPULSE_POINT_SOURCE_JOULES, IMPULSE_SOURCE, NO_SOURCE,
                                    SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE, STEPPED_SOURCE,
                                    TEMPERATURE_COMPUTE, TEMPERATURE_FIXED)
from heatsim2_amd.grids import (build_grid, build_grid_min_step, build_grid_min_step_edge,
                                zero_elements)
import heatsim2_amd.expression as expression
import heatsim2_amd.boundary_conducting as boundary_conducting
import heatsim2_amd.boundary_conducting_anisotropic as boundary_conducting_anisotropic
import heatsim2_amd.boundary_insulating as boundary_insulating
import heatsim2_amd.boundary_thininsulatinglayer as boundary_thininsulatinglayer
import heatsim2_amd.surface_temperature as surface_temperature
import heatsim2_amd.hs2_indexing as hs2_indexing
import heatsim2_amd.crank_nicolson as crank_nicolson
import heatsim2_amd.alternatingdirection as alternatingdirection
import heatsim2_amd.alternatingdirection_c_pyx as alternatingdirection_c_pyx
import heatsim2_amd.tridiag as tridiag
import heatsim2_amd.autotune as autotune
import heatsim2_amd.adjoint as adjoint
import heatsim2_amd.batch as batch

from heatsim2_amd._lib import LIBRARY_VERSION as __version__

setup = crank_nicolson.setup
run_adi_steps = alternatingdirection.run_adi_steps
run_adi_steps_many = alternatingdirection.run_adi_steps_many


# the upstream's module names, for plugins written against it
_UPSTREAM_MODULES = ("expression", "boundary_conducting", "boundary_conducting_anisotropic",
                     "boundary_insulating", "boundary_thininsulatinglayer", "surface_temperature",
                     "hs2_indexing", "crank_nicolson", "alternatingdirection_c_pyx", "tridiag")


def alias_as_heatsim2(install=True):
    """Opt-in alias for boundary plugins written for the upstream package.

    Such plugins do ``from heatsim2.expression import group``
    (boundary_conducting.py:4 of the upstream).  alias_as_heatsim2()
    registers this package under the name ``heatsim2`` (and its modules under
    ``heatsim2.expression``, ``heatsim2.boundary_conducting``, ...), so that
    those imports bind heatsim2_amd's expression objects; import the plugins
    after calling it.  It refuses when a different ``heatsim2`` is already
    imported.  alias_as_heatsim2(False) removes the alias again (modules
    imported meanwhile keep the objects they bound).
    """
    import sys
    me = sys.modules[__name__]
    cur = sys.modules.get("heatsim2")
    if not install:
        if cur is me:
            del sys.modules["heatsim2"]
            for name in _UPSTREAM_MODULES:
                sys.modules.pop("heatsim2." + name, None)
        return
    if cur is not None and cur is not me:
        raise RuntimeError("another module named heatsim2 is already imported (%s); the alias "
                           "would shadow it" % getattr(cur, "__file__", cur))
    sys.modules["heatsim2"] = me
    for name in _UPSTREAM_MODULES:
        sys.modules["heatsim2." + name] = getattr(me, name)

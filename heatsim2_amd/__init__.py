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

__version__ = "0.1.0"

setup = crank_nicolson.setup
run_adi_steps = alternatingdirection.run_adi_steps
run_adi_steps_many = alternatingdirection.run_adi_steps_many

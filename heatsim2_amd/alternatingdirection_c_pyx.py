"""The step module under the upstream's own name.  heatsim2/__init__.py:15-16
imports heatsim2.alternatingdirection_c_pyx both as itself and as
`alternatingdirection`; here both names hold the same objects:
run_adi_steps, adi_params, pyadi_step (plus run_adi_steps_many and
source_tables).  The upstream's symbolic assembly helpers (adi_setup,
add_equation_to_adi_matrices, adi_expressions, :418-572) have no
counterpart: setup() derives the coefficients in closed form."""
from heatsim2_amd.alternatingdirection import (adi_params, pyadi_step, run_adi_steps,  # noqa: F401
                                               run_adi_steps_many, source_tables)

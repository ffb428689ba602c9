"""Isotropic conducting face: q = -(k_m + k_p)/2 * (T_p - T_m)/d.

Face-law plugin with the upstream's argument order (boundary_conducting.py:
7-71): conductivities of the minus / plus cells, the three spacings, the
minus / plus temperatures and twelve off-axis neighbour temperatures.  The
normal difference sits in ``group()`` so the integrator treats it
implicitly.
"""
from heatsim2_amd.expression import group


def qz(kmatm55, kmatp55, dz, dy, dx, Tm55, Tp55, Tm45, Tp45, Tm65, Tp65, Tm54, Tp54,
       Tm56, Tp56, Tm46, Tp46, Tm64, Tp64):
    return -((kmatm55 + kmatp55) * 0.5) * group((Tp55 - Tm55) * (1.0 / dz))


def qy(kmat5m5, kmat5p5, dz, dy, dx, T5m5, T5p5, T4m5, T4p5, T6m5, T6p5, T5m4, T5p4,
       T5m6, T5p6, T4m6, T4p6, T6m4, T6p4):
    return -((kmat5m5 + kmat5p5) * 0.5) * group((T5p5 - T5m5) * (1.0 / dy))


def qx(kmat55m, kmat55p, dz, dy, dx, T55m, T55p, T45m, T45p, T65m, T65p, T54m, T54p,
       T56m, T56p, T46m, T46p, T64m, T64p):
    return -((kmat55m + kmat55p) * 0.5) * group((T55p - T55m) * (1.0 / dx))

"""Tensor-conductivity face: q_a = -sum_b K_ab dT/db with K the mean of the
two cells' 3x3 conductivities in (z, y, x) order.

Same face law as the upstream boundary_conducting_anisotropic.py:7-67: the
normal derivative is a two-point difference, the tangential derivatives are
four-point averages over both cells.  With a diagonal K only the normal term
survives.  Off-diagonal K needs the 15-substep unaligned mode, whose
upstream setup fails an assert before any step exists; setup() here fails
the same way (an AssertionError subclass).
"""
from heatsim2_amd.expression import group


def qz(kmatm55, kmatp55, dz, dy, dx, Tm55, Tp55, Tm45, Tp45, Tm65, Tp65, Tm54, Tp54,
       Tm56, Tp56, Tm46, Tp46, Tm64, Tp64):
    K = (kmatm55 + kmatp55) * 0.5
    dTdz = group((Tp55 - Tm55) * (1.0 / dz))
    dTdy = group((Tp65 - Tp45) * (0.25 / dy)) + group((Tm65 - Tm45) * (0.25 / dy))
    dTdx = group((Tp56 - Tp54) * (0.25 / dx)) + group((Tm56 - Tm54) * (0.25 / dx))
    return -K[0, 0] * dTdz - K[0, 1] * dTdy - K[0, 2] * dTdx


def qy(kmat5m5, kmat5p5, dz, dy, dx, T5m5, T5p5, T4m5, T4p5, T6m5, T6p5, T5m4, T5p4,
       T5m6, T5p6, T4m6, T4p6, T6m4, T6p4):
    K = (kmat5m5 + kmat5p5) * 0.5
    dTdz = group((T6p5 - T4p5) * (0.25 / dz)) + group((T6m5 - T4m5) * (0.25 / dz))
    dTdy = group((T5p5 - T5m5) * (1.0 / dy))
    dTdx = group((T5p6 - T5p4) * (0.25 / dx)) + group((T5m6 - T5m4) * (0.25 / dx))
    return -K[1, 0] * dTdz - K[1, 1] * dTdy - K[1, 2] * dTdx


def qx(kmat55m, kmat55p, dz, dy, dx, T55m, T55p, T45m, T45p, T65m, T65p, T54m, T54p,
       T56m, T56p, T46m, T46p, T64m, T64p):
    K = (kmat55m + kmat55p) * 0.5
    dTdz = group((T65p - T45p) * (0.25 / dz)) + group((T65m - T45m) * (0.25 / dz))
    # the minus cell's y difference is T55m - T54m upstream, not T56m - T54m
    # (boundary_conducting_anisotropic.py:60); kept, since it is the face law
    # a plugin caller sees, and K[2, 1] = 0 makes it drop out of every case
    # setup() accepts
    dTdy = group((T56p - T54p) * (0.25 / dy)) + group((T55m - T54m) * (0.25 / dy))
    dTdx = group((T55p - T55m) * (1.0 / dx))
    return -K[2, 0] * dTdz - K[2, 1] * dTdy - K[2, 2] * dTdx

"""A third-party face law written against the upstream package (it imports
heatsim2.expression, as the upstream's own boundary modules do): a
conducting face whose conductivity is the harmonic mean of the two cells',
q = -2 k_m k_p / (k_m + k_p) (T_p - T_m) / d."""
from heatsim2.expression import group


def _kh(km, kp):
    return 2.0 * km * kp / (km + kp)


def qz(kmatm55, kmatp55, dz, dy, dx, Tm55, Tp55, *neighbours):
    return -_kh(kmatm55, kmatp55) * group((Tp55 - Tm55) * (1.0 / dz))


def qy(kmat5m5, kmat5p5, dz, dy, dx, T5m5, T5p5, *neighbours):
    return -_kh(kmat5m5, kmat5p5) * group((T5p5 - T5m5) * (1.0 / dy))


def qx(kmat55m, kmat55p, dz, dy, dx, T55m, T55p, *neighbours):
    return -_kh(kmat55m, kmat55p) * group((T55p - T55m) * (1.0 / dx))

"""Insulating face: no flux (boundary_insulating.py:2-34).  Every outer face
of a COMPUTE cell must use this law."""


def qz(*args):
    return 0.0


def qy(*args):
    return 0.0


def qx(*args):
    return 0.0

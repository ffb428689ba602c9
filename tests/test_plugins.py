"""Boundary plugins that are not the four shipped face laws.

- A third-party law written against the upstream package (it does
  ``from heatsim2.expression import group``, as the upstream's own modules
  do) imports through heatsim2_amd.alias_as_heatsim2() and gives a harmonic
  mean conductance; on the GPU its trajectory matches the oracle's step on
  per-cell rows built from the same law.
- A law with its difference outside group() (explicitly integrated by the
  upstream, expression.py:712-716) is refused with a clear error, and so is
  a plugin that returns an object of another expression type.
"""
import importlib
import os
import sys

import numpy as np
import pytest

import heatsim2_amd as hs
from heatsim2_amd import crank_nicolson
import oracle
from hs2_cases import oracle_coef

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture
def upstream_alias():
    sys.path.insert(0, os.path.join(HERE, "plugins"))
    hs.alias_as_heatsim2()
    try:
        yield
    finally:
        hs.alias_as_heatsim2(False)
        sys.modules.pop("harmonic_conductor", None)
        sys.path.remove(os.path.join(HERE, "plugins"))


def _harmonic_case():
    """Two materials side by side in x (k 40 and 0.5), two layers in z with
    a third k, harmonic faces inside, insulating outside, a flash on top."""
    nz, ny, nx = 6, 5, 8
    d = (4e-4, 5e-4, 6e-4)
    m = np.zeros((nz, ny, nx), np.uint8)
    m[:, :, 4:] = 1
    m[3:] = 2
    bz = np.zeros((nz + 1, ny, nx), np.uint8)
    by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    for b in (bz[0], bz[-1], by[:, 0], by[:, -1], bx[:, :, 0], bx[:, :, -1]):
        b[...] = 1
    v = np.zeros((nz, ny, nx), np.uint8)
    v[0] = 1
    ks = (40.0, 0.5, 3.0)
    rhoc = (7850.0 * 470.0, 1200.0 * 1500.0, 2700.0 * 900.0)
    return dict(shape=(nz, ny, nx), d=d, dt=2e-3, maps=(m, bz, by, bx, v), k=ks, rhoc=rhoc)


def _harmonic_rows(c):
    """Per-cell oracle rows of the harmonic law, restated directly."""
    m = c["maps"][0]
    k = np.array(c["k"])[m]
    M = np.array(c["rhoc"])[m] / c["dt"]
    g = []
    for a in range(3):
        ga = np.zeros(c["shape"])
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, c["shape"][a] - 1)
        hi[a] = slice(1, None)
        km, kp = k[tuple(lo)], k[tuple(hi)]
        ga[tuple(hi)] = 2.0 * km * kp / (km + kp) / c["d"][a] ** 2
        g.append(ga)
    return oracle_coef(M, np.ones(c["shape"]), *g)


def _setup_args(c, plugin):
    nz, ny, nx = c["shape"]
    dz, dy, dx = c["d"]
    mats = tuple((hs.TEMPERATURE_COMPUTE, k, rc, 1.0) for k, rc in zip(c["k"], c["rhoc"]))
    bnds = ((plugin,), (hs.boundary_insulating,))
    vol = ((hs.NO_SOURCE,), (hs.IMPULSE_SOURCE, 0.0, 1e4 / dz))
    return (0.0, 0.0, 0.0, dz, dy, dx, nz, ny, nx, c["dt"], mats, bnds, vol) + tuple(c["maps"])


def test_upstream_style_plugin_imports_through_the_alias(upstream_alias):
    import heatsim2
    import heatsim2.expression as ex
    assert heatsim2 is hs and ex is hs.expression
    plugin = importlib.import_module("harmonic_conductor")
    assert plugin.group is hs.expression.group
    c = _harmonic_case()
    m, bz, by, bx, _ = c["maps"]
    mats = tuple((hs.TEMPERATURE_COMPUTE, k, rc, 1.0) for k, rc in zip(c["k"], c["rhoc"]))
    gtab, mtab, comp = crank_nicolson.build_tables(c["d"], c["dt"], mats,
                                                   ((plugin,), (hs.boundary_insulating,)),
                                                   m, bz, by, bx)
    for a in range(3):
        for ma, mb in ((0, 1), (0, 2), (1, 2), (0, 0)):
            ka, kb = c["k"][ma], c["k"][mb]
            want = 2.0 * ka * kb / (ka + kb) / c["d"][a] ** 2
            got = gtab[0, a, ma, mb]
            assert got == 0.0 or abs(got - want) <= 1e-15 * want, (a, ma, mb, got, want)
    # the x faces between materials 0 and 1 occur, so that entry is set
    want = 2 * 40 * 0.5 / 40.5 / c["d"][2] ** 2
    assert abs(gtab[0, 2, 0, 1] - want) <= 1e-15 * want


def test_alias_refuses_to_shadow_another_heatsim2():
    fake = type(sys)("heatsim2")
    sys.modules["heatsim2"] = fake
    try:
        with pytest.raises(RuntimeError, match="already imported"):
            hs.alias_as_heatsim2()
    finally:
        del sys.modules["heatsim2"]
    hs.alias_as_heatsim2()
    hs.alias_as_heatsim2(False)
    assert "heatsim2" not in sys.modules and "heatsim2.expression" not in sys.modules


def test_flux_outside_group_is_refused():
    sys.path.insert(0, os.path.join(HERE, "plugins"))
    try:
        plugin = importlib.import_module("explicit_flux")
    finally:
        sys.path.remove(os.path.join(HERE, "plugins"))
    c = _harmonic_case()
    m, bz, by, bx, _ = c["maps"]
    mats = ((hs.TEMPERATURE_COMPUTE, 1.0, 1e6, 1.0),) * 3
    with pytest.raises(ValueError, match="outside group"):
        crank_nicolson.build_tables(c["d"], c["dt"], mats, ((plugin,), (hs.boundary_insulating,)),
                                    m, bz, by, bx)


def test_foreign_expression_objects_are_refused():
    class Foreign(object):
        def __mul__(self, o):
            return self
        __rmul__ = __mul__
        __neg__ = lambda self: self

    class Plugin(object):
        @staticmethod
        def qz(*a):
            return Foreign()
        qy = qz
        qx = qz

    c = _harmonic_case()
    m, bz, by, bx, _ = c["maps"]
    mats = ((hs.TEMPERATURE_COMPUTE, 1.0, 1e6, 1.0),) * 3
    with pytest.raises(TypeError, match="alias_as_heatsim2"):
        crank_nicolson.build_tables(c["d"], c["dt"], mats, ((Plugin,), (hs.boundary_insulating,)),
                                    m, bz, by, bx)


@pytest.mark.gpu
def test_harmonic_plugin_trajectory_matches_oracle(upstream_alias):
    import torch
    plugin = importlib.import_module("harmonic_conductor")
    c = _harmonic_case()
    P, S = hs.setup(*_setup_args(c, plugin))
    nsteps = 40
    T0 = np.full(c["shape"], 293.0)
    T, hist = hs.run_adi_steps_many(P, 0.0, c["dt"], nsteps, torch.tensor(T0).cuda(),
                                    c["maps"][4], ((hs.NO_SOURCE,),
                                                   (hs.IMPULSE_SOURCE, 0.0, 1e4 / c["d"][0])),
                                    surface_history=True)
    S0 = np.where(c["maps"][4] == 1, 1e4 / c["d"][0] / c["dt"], 0.0)
    To, ho = oracle.fast_run(_harmonic_rows(c), T0, nsteps, S0, 1, c["d"][0])
    rise = float(ho.max() - 293.0)
    assert np.abs(T.cpu().numpy() - To).max() <= 1e-10 * rise
    assert np.abs(hist.cpu().numpy() - ho).max() <= 1e-10 * rise

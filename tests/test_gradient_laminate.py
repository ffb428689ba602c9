"""Gradients at the single-instance inversion size: a 64x128x128
anisotropic laminate (kx = ky = 5, rho*c = 1550*1200, thin insulating
delaminations, a 1e4 J/m^2 flash at step 0, 0.5 mm cells, dt = 2 ms,
500 steps, checkpoints every 50 steps).

The true kz field is 0.6 W/m/K x (1 + 0.3 g), g white noise (seed 3)
Gaussian-filtered with sigma = 8 cells and re-standardised; the observation
is its surface-temperature history plus N(0, 0.01^2) K noise (seed 4); the
loss is the L2 misfit at the homogeneous guess kz = 0.6.  Gradients with
respect to per-voxel kz, ky, kx, rho*c and the flash amplitude E are checked
against Richardson-extrapolated central differences of the device forward
at 32 random voxels (seed 5) of the top four layers per parameter.  The
surface history senses a voxel through z diffusion, whose length over the
1 s window is about 1.6 cells, so deeper voxels carry gradients at
round-off level.  The loss difference of a central difference is formed as
sum((h+ - h-) (h+ + h- - 2 obs)), which keeps its round-off at the scale of
the history change instead of the loss.
Bar for kz, rho*c and E: |g - g_fd| <= 1e-6 max(|g_fd|, 1e-3 max|g_fd|),
with steps of 3e-3 of the value.
The lateral conductivities matter only where the field varies laterally
(delamination edges, the kz texture), and their gradients are 1e-6 of kz's:
a central difference of this 1e7 loss resolves them to about 5e-9 (its
loss difference carries ~1e-9 of round-off from the two forward runs).
Their 32 voxels are drawn among the top-four-layer voxels whose gradient is
at least a quarter of the largest, with steps of 2e-2 of the value, and the
bar is 1e-5 max(|g_fd|, 1e-3 max|g_fd|).  (kx and ky meet the 1e-6 bar on
the smaller grids of tests/test_adjoint.py, against the upstream solver's
and the device's differences.)
"""
import numpy as np
import pytest
import torch

import heatsim2_amd as hs
from heatsim2_amd import adjoint
from hs2_cases import cases

NSTEPS = 500
K = 50


def _setup(dev):
    from scipy.ndimage import gaussian_filter
    case = cases.case_laminate(64, 128, 128, nsteps=NSTEPS)
    shape = case["shape"]
    g = gaussian_filter(np.random.default_rng(3).standard_normal(shape), sigma=8.0)
    g = (g - g.mean()) / g.std()
    kz_true = 0.6 * (1.0 + 0.3 * g)
    _, bz, by, bx, _ = case["maps"]
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, case["boundaries"][2][1]))
    fixed, kz0, ky0, kx0, rho, c, _t = case["materials"][0]
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    base = dict(kz=t(np.full(shape, kz0)), ky=t(np.full(shape, ky0)), kx=t(np.full(shape, kx0)),
                rhoc=t(np.full(shape, rho * c)), E=t(1e4))
    mask = torch.zeros(shape, dtype=torch.float64, device=dev)
    mask[0] = 1.0
    env = dict(shape=shape, d=case["d"], dt=case["dt"], bnds=bnds, maps=(bz, by, bx), mask=mask,
               gplan=adjoint.GradPlan(shape, case["d"], case["dt"]))
    T0 = torch.zeros(shape, dtype=torch.float64, device=dev)  # relative to ambient
    env["T0"] = T0
    with torch.no_grad():
        truth = dict(base, kz=t(kz_true))
        h = _history(env, truth)
        noise = np.random.default_rng(4).standard_normal(tuple(h.shape))
        env["obs"] = h + 0.01 * t(noise)
    return env, base


def _history(env, p):
    coef = adjoint.per_voxel_coefficients(p["kz"], p["ky"], p["kx"], p["rhoc"], env["dt"],
                                          env["d"], env["bnds"], *env["maps"])
    S0 = p["E"] * env["mask"] / (env["d"][0] * env["dt"])
    return adjoint.surface_history(env["gplan"], coef, env["T0"], NSTEPS, sources=[(0, S0)],
                                   ckpt_every=K)


def _loss(env, p):
    return ((_history(env, p) - env["obs"]) ** 2).sum()


def _richardson(hist_at, obs, x0, rel, h=None):
    """(4 D(h) - D(2h)) / 3 of the loss, D(h) = (L(x0+h) - L(x0-h)) / 2h with
    L(a) - L(b) = sum((h_a - h_b) (h_a + h_b - 2 obs)); h = rel |x0| unless
    given."""
    def D(h):
        a, b = hist_at(x0 + h), hist_at(x0 - h)
        return float(((a - b) * (a + b - 2.0 * obs)).sum()) / (2 * h)
    if h is None:
        h = rel * abs(x0)
    return (4 * D(h) - D(2 * h)) / 3


def _grad_ok(g, fd, bar=1e-6):
    g, fd = np.asarray(g, float), np.asarray(fd, float)
    tol = bar * np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
    return bool(np.all(np.abs(g - fd) <= tol)), float(np.max(np.abs(g - fd) / tol))


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_laminate_gradients_match_central_differences():
    dev = torch.device("cuda")
    env, base = _setup(dev)
    p = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    loss = _loss(env, p)
    loss.backward()
    nz, ny, nx = env["shape"]
    top = 4 * ny * nx
    rng = np.random.default_rng(5)
    for name in ("kz", "rhoc", "ky", "kx"):
        gall = p[name].grad.detach().cpu().numpy().ravel()
        if name in ("kz", "rhoc"):
            vox = rng.choice(top, size=32, replace=False)
            rel, bar = 3e-3, 1e-6
        else:
            live = np.flatnonzero(np.abs(gall[:top]) >= 0.25 * np.abs(gall[:top]).max())
            vox = rng.choice(live, size=min(32, live.size), replace=False)
            rel, bar = 2e-2, 1e-5
        g = gall[vox]
        fd = []
        for v in vox:
            def hist_at(val, v=v):
                q = {k: t.detach().clone() for k, t in p.items()}
                q[name].view(-1)[int(v)] = val
                with torch.no_grad():
                    return _history(env, q)
            fd.append(_richardson(hist_at, env["obs"], float(base[name].view(-1)[int(v)]), rel))
        ok, worst = _grad_ok(g, fd, bar)
        print("d/d%s at %d voxels: |g| %.3e .. %.3e, worst %.3f of the tolerance"
              % (name, len(vox), np.abs(g).min(), np.abs(g).max(), worst), flush=True)
        assert ok, (name, vox, g, fd, worst)

    def hist_E(val):
        q = {k: t.detach().clone() for k, t in p.items()}
        q["E"] = torch.tensor(val, dtype=torch.float64, device=dev)
        with torch.no_grad():
            return _history(env, q)
    fdE = _richardson(hist_E, env["obs"], float(base["E"]), 3e-3)
    ok, worst = _grad_ok([float(p["E"].grad)], [fdE])
    print("d/dE: adjoint %.9e, fd %.9e" % (float(p["E"].grad), fdE), flush=True)
    assert ok, (float(p["E"].grad), fdE, worst)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_laminate_lateral_gradients_along_smooth_directions():
    """The lateral conductivities' gradients over every voxel at once: g . v
    of the adjoint's dJ/dkx (and dJ/dky) against the Richardson-extrapolated
    central difference of the loss along v, a smooth random field
    (Gaussian-filtered noise, sigma = 4 cells, seed 6) scaled to 32 % of the
    conductivity at its largest, with steps of 1 and 2 v.  Along v the
    history moves by ~1e-8 K per value, so the differences' own round-off
    limits them to about 2e-6: Richardson estimates from steps of 4 % to 64 %
    agree with each other, and with g . v, to that (measured: 4.9e-7 for kx,
    1.2e-6 for ky at these steps).  Bar: |g . v - fd| <= 5e-6 |fd|."""
    from scipy.ndimage import gaussian_filter
    dev = torch.device("cuda")
    env, base = _setup(dev)
    p = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    _loss(env, p).backward()
    rng = np.random.default_rng(6)
    for name in ("kx", "ky"):
        v = gaussian_filter(rng.standard_normal(env["shape"]), sigma=4.0)
        v = torch.tensor(v / np.abs(v).max(), dtype=torch.float64, device=dev)
        v = v * 0.32 * float(base[name].abs().max())
        gv = float((p[name].grad * v).sum())

        def hist_at(eps, name=name, v=v):
            q = {k: t.detach().clone() for k, t in p.items()}
            q[name] = q[name] + eps * v
            with torch.no_grad():
                return _history(env, q)
        fd = _richardson(hist_at, env["obs"], 0.0, None, h=1.0)
        print("d/d%s along v: adjoint %.12e, fd %.12e, rel %.2e"
              % (name, gv, fd, abs(gv - fd) / abs(fd)), flush=True)
        assert abs(gv - fd) <= 5e-6 * abs(fd), (name, gv, fd)

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
at 32 random voxels (seed 5) of the top four layers -- the surface history
senses a voxel through z diffusion, whose length over the 1 s window is
about 1.6 cells, so deeper voxels carry gradients at round-off level.
Bar: |g - g_fd| <= 1e-6 max(|g_fd|, 1e-3 max|g_fd|).
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


def _richardson(f, x0, rel=3e-3):
    h = rel * abs(x0)
    d1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    d2 = (f(x0 + 2 * h) - f(x0 - 2 * h)) / (4 * h)
    return (4 * d1 - d2) / 3


def _grad_ok(g, fd):
    g, fd = np.asarray(g, float), np.asarray(fd, float)
    tol = 1e-6 * np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
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
    vox = np.random.default_rng(5).choice(4 * ny * nx, size=32, replace=False)
    for name in ("kz", "ky", "kx", "rhoc"):
        g = p[name].grad.detach().cpu().numpy().ravel()[vox]
        fd = []
        for v in vox:
            def f(val, v=v):
                q = {k: t.detach().clone() for k, t in p.items()}
                q[name].view(-1)[int(v)] = val
                with torch.no_grad():
                    return float(_loss(env, q))
            fd.append(_richardson(f, float(base[name].view(-1)[int(v)])))
        ok, worst = _grad_ok(g, fd)
        print("d/d%s at 32 voxels: max |g| %.3e, worst %.3f of the tolerance"
              % (name, np.abs(g).max(), worst), flush=True)
        assert ok, (name, vox, g, fd, worst)

    def fE(val):
        q = {k: t.detach().clone() for k, t in p.items()}
        q["E"] = torch.tensor(val, dtype=torch.float64, device=dev)
        with torch.no_grad():
            return float(_loss(env, q))
    fdE = _richardson(fE, float(base["E"]))
    ok, worst = _grad_ok([float(p["E"].grad)], [fdE])
    print("d/dE: adjoint %.9e, fd %.9e" % (float(p["E"].grad), fdE), flush=True)
    assert ok, (float(p["E"].grad), fdE, worst)

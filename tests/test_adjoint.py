"""Gradients through the ADI step (heatsim2_amd.adjoint).

CPU: the differentiable per-voxel coefficient builder against the oracle's
assembly (itself pinned to the upstream's assembled operators).
GPU: the hand-written adjoint against
  - finite differences of the real upstream solver (grad_tiny fixture:
    per-voxel kz, ky, kx, rho*c at 10 voxels each and the flash amplitude),
  - central finite differences of the device forward on a case with FIXED
    cells, thin layers and several sources,
  - itself across checkpoint intervals,
and the per-voxel forward against the class-map forward.
Gradient tolerance: |g - g_fd| <= 1e-6 * max(|g_fd|, 1e-3 * max|g_fd|).
Finite differences are Richardson-extrapolated central differences,
(4 D(h) - D(2h)) / 3 with h = 3e-3 * |value|, on temperatures relative to
ambient (the step is affine in T, so gradients are unchanged while the
round-off in T stays at the scale of the rise).
"""
import numpy as np
import pytest
import torch

import heatsim2_amd as hs
from heatsim2_amd import adjoint
from hs2_cases import cases, load_golden

import oracle

BMOD = {cases.BND_CONDUCTING: hs.boundary_conducting,
        cases.BND_ANISO: hs.boundary_conducting_anisotropic,
        cases.BND_INSULATING: hs.boundary_insulating,
        cases.BND_THIN: hs.boundary_thininsulatinglayer}


def _boundaries(case_bnds):
    return tuple((BMOD[k], h) if k == cases.BND_THIN else (BMOD[k],) for k, h in case_bnds)


def _fields_of_case(case):
    """Per-voxel kz, ky, kx, rho*c and FIXED mask of a material-map case."""
    m = case["maps"][0]
    mats = np.array([[kz, ky, kx, rho * c, fixed] for (fixed, kz, ky, kx, rho, c, _t)
                     in case["materials"]], dtype=np.float64)
    f = mats[m]
    t = [torch.from_numpy(np.ascontiguousarray(f[..., i])) for i in range(4)]
    return t[0], t[1], t[2], t[3], torch.from_numpy(f[..., 4] > 0)


def _coef_from_fields(case, kz, ky, kx, rhoc, fixed):
    return adjoint.per_voxel_coefficients(kz, ky, kx, rhoc, case["dt"], case["d"],
                                          _boundaries(case["boundaries"]), case["maps"][1],
                                          case["maps"][2], case["maps"][3], fixed=fixed)


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso"])
def test_per_voxel_coefficients_match_oracle(name):
    case = load_golden(name)
    kz, ky, kx, rhoc, fixed = _fields_of_case(case)
    coef = _coef_from_fields(case, kz, ky, kx, rhoc, fixed)
    ref = oracle.assemble_case(case)
    nz, ny, nx = case["shape"]
    w = coef["w"].numpy().ravel().astype(bool)
    np.testing.assert_allclose(coef["M"].numpy().ravel()[w], ref[w, 0], rtol=1e-15)
    assert np.all(coef["M"].numpy().ravel()[~w] == 1.0)
    for a, key in enumerate(("gz", "gy", "gx")):
        g = coef[key].numpy()
        # minus face as seen by the COMPUTE cell itself
        np.testing.assert_allclose(g.ravel()[w], ref[w, 2 + 2 * a], rtol=1e-15, atol=0)
        # plus face of each COMPUTE cell = minus face of its neighbour
        gp = np.zeros_like(g)
        sl = [slice(None)] * 3
        sl_hi = list(sl); sl_hi[a] = slice(0, g.shape[a] - 1)
        sl_lo = list(sl); sl_lo[a] = slice(1, None)
        gp[tuple(sl_hi)] = g[tuple(sl_lo)]
        np.testing.assert_allclose(gp.ravel()[w], ref[w, 3 + 2 * a], rtol=1e-15, atol=0)


def test_per_voxel_coefficients_reject_open_outer_face():
    case = load_golden("tiny_mixed")
    kz, ky, kx, rhoc, fixed = _fields_of_case(case)
    bz = case["maps"][1].copy()
    c = np.argwhere(~fixed.numpy()[0])[0]
    bz[0, c[0], c[1]] = 0  # conducting class on the top outer face of a COMPUTE cell
    with pytest.raises(ValueError):
        adjoint.per_voxel_coefficients(kz, ky, kx, rhoc, case["dt"], case["d"],
                                       _boundaries(case["boundaries"]), bz, case["maps"][2],
                                       case["maps"][3], fixed=fixed)


def _grad_ok(g, fd):
    g, fd = np.asarray(g, float), np.asarray(fd, float)
    tol = 1e-6 * np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
    return np.abs(g - fd) <= tol, np.max(np.abs(g - fd) / tol)


def _richardson(f, x0, rel=3e-3, h=None):
    h = rel * abs(x0) if h is None else h
    d1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    d2 = (f(x0 + 2 * h) - f(x0 - 2 * h)) / (4 * h)
    return (4 * d1 - d2) / 3


def _grad_tiny_inputs(z, dev):
    kz = torch.tensor(z["kz"], device=dev, requires_grad=True)
    ky = torch.tensor(z["ky"], device=dev, requires_grad=True)
    kx = torch.tensor(z["kx"], device=dev, requires_grad=True)
    rhoc = torch.tensor(z["rhoc"], device=dev, requires_grad=True)
    E = torch.tensor(float(z["E"]), dtype=torch.float64, device=dev, requires_grad=True)
    return kz, ky, kx, rhoc, E


def _grad_tiny_loss(z, gplan, kz, ky, kx, rhoc, E, ckpt_every):
    dev = kz.device
    d = tuple(float(v) for v in z["d"])
    dt = float(z["dt"])
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, float(z["h"])))
    coef = adjoint.per_voxel_coefficients(kz, ky, kx, rhoc, dt, d, bnds, z["boundary_z_elements"],
                                          z["boundary_y_elements"], z["boundary_x_elements"])
    top = torch.tensor(z["volumetric_elements"] == 1, device=dev)
    S0 = E * top / (d[0] * dt)
    T0 = torch.full(tuple(kz.shape), float(z["T0"]), dtype=torch.float64, device=dev)
    hist = adjoint.surface_history(gplan, coef, T0, int(z["nsteps"]), sources=[(0, S0)],
                                   ckpt_every=ckpt_every)
    obs = torch.tensor(z["obs"], device=dev)
    return ((hist - obs) ** 2).sum(), hist


@pytest.mark.gpu
@pytest.mark.parametrize("ckpt_every", [7, 30])
def test_gradients_match_upstream_finite_differences(ckpt_every):
    z = load_golden_raw("grad_tiny")
    dev = torch.device("cuda")
    kz, ky, kx, rhoc, E = _grad_tiny_inputs(z, dev)
    gplan = adjoint.GradPlan(tuple(z["shape"]), tuple(z["d"]), float(z["dt"]))
    loss, hist = _grad_tiny_loss(z, gplan, kz, ky, kx, rhoc, E, ckpt_every)
    # forward parity with the upstream trajectory that produced the fixture
    rise = float(np.max(z["hist"]) - float(z["T0"]))
    assert np.max(np.abs(hist.detach().cpu().numpy() - z["hist"])) <= 1e-10 * rise
    assert abs(loss.item() - float(z["loss"])) <= 1e-9 * float(z["loss"])
    loss.backward()
    for name, t in (("kz", kz), ("ky", ky), ("kx", kx), ("rhoc", rhoc)):
        g = t.grad.detach().cpu().numpy().ravel()[z["fd_%s_vox" % name]]
        ok, worst = _grad_ok(g, z["fd_%s" % name])
        assert ok.all(), (name, g, z["fd_%s" % name], worst)
    ok, worst = _grad_ok([E.grad.item()], [float(z["fd_E"])])
    assert ok.all(), (E.grad.item(), float(z["fd_E"]), worst)


def load_golden_raw(name):
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return np.load(os.path.join(here, "golden", name + ".npz"), allow_pickle=False)


def _fd_case(dev):
    """8x7x9 grid: two materials, a FIXED cell block, thin layers on two
    axes, insulating shell; sources at steps 0 and 3."""
    rng = np.random.default_rng(5)
    nz, ny, nx = 8, 7, 9
    d = (3e-4, 4e-4, 5e-4)
    dt = 1e-3
    kz = 0.8 * (1 + 0.3 * rng.random((nz, ny, nx)))
    ky = 3.0 * (1 + 0.3 * rng.random((nz, ny, nx)))
    kx = 2.0 * (1 + 0.3 * rng.random((nz, ny, nx)))
    rhoc = 2e6 * (1 + 0.2 * rng.random((nz, ny, nx)))
    fixed = np.zeros((nz, ny, nx), bool)
    fixed[5:7, 2:4, 6:8] = True
    bz = np.zeros((nz + 1, ny, nx), np.uint8)
    by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    bz[0] = bz[-1] = 1
    by[:, 0] = by[:, -1] = 1
    bx[:, :, 0] = bx[:, :, -1] = 1
    bz[3, 1:5, 2:7] = 2
    bx[2:6, 1:6, 4] = 2
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, 400.0))
    T0 = rng.random((nz, ny, nx))
    src_mask = np.zeros((nz, ny, nx)); src_mask[0] = 1.0
    src2 = rng.random((nz, ny, nx)) * 1e6
    noise = rng.standard_normal((12, ny, nx))
    t = lambda a: torch.tensor(a, device=dev)
    c = dict(shape=(nz, ny, nx), d=d, dt=dt, kz=t(kz), ky=t(ky), kx=t(kx), rhoc=t(rhoc),
             fixed=t(fixed), maps=(bz, by, bx), bnds=bnds, T0=t(T0), mask=t(src_mask),
             src2=t(src2), nsteps=12, E=2e3)
    # observations: the history of a perturbed "true" field plus 0.01 K noise,
    # so that the misfit (and its round-off) stays small
    wig = lambda a: a * (1 + 0.1 * torch.sin(torch.arange(a.numel(), device=dev,
                                                           dtype=torch.float64).view(a.shape)))
    c["obs"] = torch.zeros((12, ny, nx), dtype=torch.float64, device=dev)
    with torch.no_grad():
        gplan = adjoint.GradPlan(c["shape"], d, dt)
        coef = adjoint.per_voxel_coefficients(wig(c["kz"]), wig(c["ky"]), wig(c["kx"]),
                                              wig(c["rhoc"]), dt, d, bnds, bz, by, bx,
                                              fixed=c["fixed"])
        S0 = 1.05 * c["E"] * c["mask"] / (d[0] * dt)
        hist = adjoint.surface_history(gplan, coef, c["T0"], 12, sources=[(0, S0), (3, c["src2"])])
    c["obs"] = hist + 0.01 * t(noise)
    return c


def _fd_loss(c, gplan, kz, ky, kx, rhoc, E, T0, ckpt_every=5):
    coef = adjoint.per_voxel_coefficients(kz, ky, kx, rhoc, c["dt"], c["d"], c["bnds"],
                                          *c["maps"], fixed=c["fixed"])
    S0 = E * c["mask"] / (c["d"][0] * c["dt"])
    hist = adjoint.surface_history(gplan, coef, T0, c["nsteps"],
                                   sources=[(0, S0), (3, c["src2"])], ckpt_every=ckpt_every)
    return ((hist - c["obs"]) ** 2).sum()


@pytest.mark.gpu
def test_gradients_match_device_central_differences():
    dev = torch.device("cuda")
    c = _fd_case(dev)
    gplan = adjoint.GradPlan(c["shape"], c["d"], c["dt"])
    params = {k: c[k].clone().requires_grad_(True) for k in ("kz", "ky", "kx", "rhoc", "T0")}
    E = torch.tensor(c["E"], dtype=torch.float64, device=dev, requires_grad=True)
    loss = _fd_loss(c, gplan, params["kz"], params["ky"], params["kx"], params["rhoc"], E,
                    params["T0"])
    loss.backward()
    rng = np.random.default_rng(6)
    n = int(np.prod(c["shape"]))
    fixed = c["fixed"].cpu().numpy().ravel()
    for name in ("kz", "ky", "kx", "rhoc", "T0"):
        # the two most sensitive voxels plus four random ones, so that the
        # tolerance floor (1e-3 of the largest checked |g_fd|) has a scale
        gall = params[name].grad.detach().cpu().numpy().ravel()
        free = np.flatnonzero(~fixed)
        top2 = free[np.argsort(-np.abs(gall[free]))[:2]]
        rest = rng.choice(np.setdiff1d(free, top2), size=4, replace=False)
        vox = np.concatenate((top2, rest))
        g = gall[vox]
        fd = []
        for v in vox:
            def f(val):
                p = {k: t.detach().clone() for k, t in params.items()}
                p[name].view(-1)[v] = val
                with torch.no_grad():
                    return _fd_loss(c, gplan, p["kz"], p["ky"], p["kx"], p["rhoc"], E.detach(),
                                    p["T0"]).item()
            x0 = float(params[name].detach().view(-1)[v])
            fd.append(_richardson(f, x0, h=3e-3 if name == "T0" else None))
        ok, worst = _grad_ok(g, fd)
        assert ok.all(), (name, vox, g, fd, worst)
    def fE(val):
        with torch.no_grad():
            return _fd_loss(c, gplan, params["kz"].detach(), params["ky"].detach(),
                            params["kx"].detach(), params["rhoc"].detach(),
                            torch.tensor(val, dtype=torch.float64, device=dev),
                            params["T0"].detach()).item()
    fdE = _richardson(fE, E.item())
    ok, worst = _grad_ok([E.grad.item()], [fdE])
    assert ok.all(), (E.grad.item(), fdE, worst)
    # FIXED cells: no gradient reaches their (unused) parameters
    assert np.abs(params["rhoc"].grad.cpu().numpy().ravel()[fixed]).max() == 0.0


@pytest.mark.gpu
def test_gradients_independent_of_checkpoint_interval():
    dev = torch.device("cuda")
    c = _fd_case(dev)
    gplan = adjoint.GradPlan(c["shape"], c["d"], c["dt"])
    grads = []
    for K in (1, 4, 12):
        kz = c["kz"].clone().requires_grad_(True)
        E = torch.tensor(c["E"], dtype=torch.float64, device=dev, requires_grad=True)
        loss = _fd_loss(c, gplan, kz, c["ky"], c["kx"], c["rhoc"], E, c["T0"], ckpt_every=K)
        loss.backward()
        grads.append((loss.item(), kz.grad.cpu().numpy(), E.grad.item()))
    for L, g, gE in grads[1:]:
        assert L == grads[0][0]
        assert np.array_equal(g, grads[0][1])
        assert gE == grads[0][2]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso"])
def test_per_voxel_forward_matches_class_forward(name):
    """The per-voxel plan reproduces the class-map plan's trajectory."""
    case = load_golden(name)
    dev = torch.device("cuda")
    kz, ky, kx, rhoc, fixed = _fields_of_case(case)
    coef = _coef_from_fields(case, kz, ky, kx, rhoc, fixed)
    gplan = adjoint.GradPlan(case["shape"], case["d"], case["dt"])
    T0 = torch.tensor(case["T0"], device=dev)
    nsteps = 6
    hist = adjoint.surface_history(gplan, {k: v.to(dev) for k, v in coef.items()}, T0, nsteps)
    from hs2_cases import to_package_args
    pa = to_package_args(hs, case)
    P, S = hs.setup(*pa["setup"])
    _, hist_ref = hs.run_adi_steps_many(P, 1e9, case["dt"], nsteps, T0, pa["vmap"],
                                        pa["volumetric"], surface_history=True)
    np.testing.assert_allclose(hist.cpu().numpy(), hist_ref.cpu().numpy(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso"])
def test_oracle_rows_of_per_voxel_coefficients(name):
    """hs2_cases.oracle_coef turns per-voxel coefficients into the oracle's
    per-cell rows: equal to the oracle's own assembly of the same case,
    COMPUTE and FIXED rows alike (the batched-path tests rely on it)."""
    from hs2_cases import oracle_coef
    case = load_golden(name)
    kz, ky, kx, rhoc, fixed = _fields_of_case(case)
    c = _coef_from_fields(case, kz, ky, kx, rhoc, fixed)
    rows = oracle_coef(*(c[k].numpy() for k in ("M", "w", "gz", "gy", "gx")))
    np.testing.assert_allclose(rows, oracle.assemble_case(case), rtol=1e-15, atol=0)

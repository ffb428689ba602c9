"""The batched path (heatsim2_amd.batch) on instances of the batched-gradient
workload's range (axes 32-128 and extreme aspect ratios, 100-160 steps)
against the oracle, the single-instance adjoint and device central
differences.

Instances: isotropic or anisotropic diagonal k drawn log-uniformly from
[0.1, 60] W/m/K, rho*c from [1e5, 5e6] J/m^3/K, 0-4 thin insulating
patches at random z-faces, a flash of about 1e4 J/m^2 on the top layer at
step 0, 0.5 mm cells, dt = 2 ms.  At k = 60 and rho*c = 1e5 the lines are
strongly coupled (k dt / (rho c dx^2) ~ 5), so every PCR level runs.

Bars:
  forward: final field and surface history within 1e-10 of the peak rise
    of the oracle's delta-form step on the same per-voxel coefficients
    (hs2_cases.oracle_coef, itself checked against the oracle's assembly);
  gradients: within 1e-9 of each field's largest entry of the
    single-instance adjoint, and |g - g_fd| <= 1e-6 max(|g_fd|, 1e-3
    max|g_fd|) against Richardson-extrapolated central differences of the
    batched forward.
"""
import numpy as np
import pytest
import torch

import heatsim2_amd as hs
import oracle
from heatsim2_amd import adjoint, batch
from hs2_cases import oracle_coef

DT = 2e-3
D = (5e-4, 5e-4, 5e-4)
H_THIN = 100.0
# cfg-range instances (every axis reaches 128; most lengths are not
# multiples of 8) and three extreme aspect ratios
SHAPES = [(128, 128, 128), (37, 128, 70), (128, 33, 99), (64, 100, 128), (45, 57, 121),
          (33, 127, 34), (100, 41, 64), (77, 90, 115), (128, 3, 5), (4, 100, 128), (3, 128, 70)]
NSTEPS = [100, 137, 160, 121, 149, 103, 155, 100, 111, 130, 100]


def _instances(seed=61):
    rng = np.random.default_rng(seed)
    out = []
    for i, sh in enumerate(SHAPES):
        nz, ny, nx = sh
        aniso = i % 2 == 1
        lk = lambda: float(np.exp(rng.uniform(np.log(0.1), np.log(60.0))))
        k = (lk(), lk(), lk()) if aniso else (lk(),) * 3
        rhoc = float(np.exp(rng.uniform(np.log(1e5), np.log(5e6))))
        if i == 0:  # the strongly coupled corner of the range
            k, rhoc = (60.0,) * 3, 1e5
        bz = np.zeros((nz + 1, ny, nx), np.uint8)
        by = np.zeros((nz, ny + 1, nx), np.uint8)
        bx = np.zeros((nz, ny, nx + 1), np.uint8)
        bz[0] = bz[-1] = 1
        by[:, 0] = by[:, -1] = 1
        bx[:, :, 0] = bx[:, :, -1] = 1
        npatch = i % 5
        for _p in range(npatch if nz > 1 else 0):
            kz = int(rng.integers(1, nz))
            sy, sx = int(rng.integers(1, ny + 1)), int(rng.integers(1, nx + 1))
            y0, x0 = int(rng.integers(0, ny - sy + 1)), int(rng.integers(0, nx - sx + 1))
            bz[kz, y0:y0 + sy, x0:x0 + sx] = 2
        E = 1e4 * (1 + 0.2 * rng.uniform(-1, 1))
        out.append(dict(shape=sh, nsteps=NSTEPS[i], k=k, rhoc=rhoc, aniso=aniso,
                        maps=(bz, by, bx), E=E))
    return out


def _bnds(x):
    law = hs.boundary_conducting_anisotropic if x["aniso"] else hs.boundary_conducting
    return ((law,), (hs.boundary_insulating,), (hs.boundary_thininsulatinglayer, H_THIN))


def _params(x, dev):
    sh = x["shape"]
    p = {n: torch.full(sh, v, dtype=torch.float64, device=dev)
         for n, v in (("kz", x["k"][0]), ("ky", x["k"][1]), ("kx", x["k"][2]), ("rhoc", x["rhoc"]))}
    p["E"] = torch.tensor(x["E"], dtype=torch.float64, device=dev)
    return p


def _coef(x, p):
    return adjoint.per_voxel_coefficients(p["kz"], p["ky"], p["kx"], p["rhoc"], DT, D, _bnds(x),
                                          *x["maps"])


def _src(x, p):
    mask = torch.zeros(x["shape"], dtype=torch.float64, device=p["kz"].device)
    mask[0] = 1.0
    return p["E"] * mask / (D[0] * DT)


def _oracle(x, coef, src):
    rows = oracle_coef(*(coef[k].detach().cpu().numpy() for k in ("M", "w", "gz", "gy", "gx")))
    T0 = np.zeros(x["shape"])
    return oracle.fast_run(rows, T0, x["nsteps"], src.detach().cpu().numpy(), 1, D[0])


@pytest.mark.gpu
def test_batch_matches_oracle_on_workload_range():
    dev = torch.device("cuda")
    insts = _instances()
    params = [_params(x, dev) for x in insts]
    coefs = [_coef(x, p) for x, p in zip(insts, params)]
    srcs = [_src(x, p) for x, p in zip(insts, params)]
    T0s = [torch.zeros(x["shape"], dtype=torch.float64, device=dev) for x in insts]
    gb = batch.GradBatch([x["shape"] for x in insts], D[0], NSTEPS, [0] * len(insts),
                         ckpt_every=20)
    finals, hists = gb.run(coefs, T0s, srcs)
    for i, x in enumerate(insts):
        To, ho = _oracle(x, coefs[i], srcs[i])
        rise = float(ho.max())
        assert rise > 0
        ef = float(np.abs(finals[i].cpu().numpy() - To).max()) / rise
        eh = float(np.abs(hists[i].cpu().numpy() - ho).max()) / rise
        print("instance %d %s x %d steps (k %s, rho c %.3g): final %.2e, history %.2e of %.4g K"
              % (i, x["shape"], x["nsteps"], tuple(round(v, 3) for v in x["k"]), x["rhoc"], ef,
                 eh, rise), flush=True)
        assert ef <= 1e-10 and eh <= 1e-10, (i, ef, eh)


def _loss(hist, obs):
    return ((hist - obs) ** 2).sum()


def _grad_ok(g, fd):
    g, fd = np.asarray(g, float), np.asarray(fd, float)
    tol = 1e-6 * np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
    return bool(np.all(np.abs(g - fd) <= tol)), float(np.max(np.abs(g - fd) / tol))


def _richardson(f, x0, rel=3e-3):
    h = rel * abs(x0)
    d1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    d2 = (f(x0 + 2 * h) - f(x0 - 2 * h)) / (4 * h)
    return (4 * d1 - d2) / 3


@pytest.mark.gpu
def test_batch_gradients_match_single_instances_and_central_differences():
    """Gradients of a surface-history misfit: the batched adjoint against the
    single-instance adjoint (three instances) and against central
    differences of the batched forward (kz and rho*c at the most sensitive
    voxel and two random top-layer voxels of two instances, and E)."""
    dev = torch.device("cuda")
    pick = [0, 1, 8]  # strongly coupled 128^3, anisotropic 37x128x70, 128x3x5
    insts = [_instances()[i] for i in pick]
    params = [{k: v.clone().requires_grad_(True) for k, v in _params(x, dev).items()}
              for x in insts]
    rng = np.random.default_rng(7)
    T0s = [torch.zeros(x["shape"], dtype=torch.float64, device=dev) for x in insts]
    # observations: a 10 % stiffer, 5 % brighter truth plus 0.01 K noise
    gb = batch.GradBatch([x["shape"] for x in insts], D[0], [x["nsteps"] for x in insts],
                         [0] * len(insts), ckpt_every=16)
    with torch.no_grad():
        truth = [{k: (v * 1.1 if k != "E" else v * 1.05) for k, v in p.items()} for p in params]
        _, h_true = gb.run([_coef(x, t) for x, t in zip(insts, truth)], T0s,
                           [_src(x, t) for x, t in zip(insts, truth)])
        obs = [h + 0.01 * torch.tensor(rng.standard_normal(tuple(h.shape)), device=dev)
               for h in h_true]
    hists = gb.surface_history([_coef(x, p) for x, p in zip(insts, params)], T0s,
                               [_src(x, p) for x, p in zip(insts, params)])
    sum(_loss(h, o) for h, o in zip(hists, obs)).backward()
    gbat = [{k: v.grad.detach().clone() for k, v in p.items()} for p in params]

    # single-instance adjoint on the same instances
    for i, x in enumerate(insts):
        p = {k: v.detach().clone().requires_grad_(True) for k, v in params[i].items()}
        gp = adjoint.GradPlan(x["shape"], D, DT)
        h = adjoint.surface_history(gp, _coef(x, p), T0s[i], x["nsteps"],
                                    sources=[(0, _src(x, p))], ckpt_every=16)
        _loss(h, obs[i]).backward()
        kscale = max(float(p[n].grad.abs().max()) for n in ("kz", "ky", "kx"))
        for n in ("kz", "ky", "kx", "rhoc"):
            scale = kscale if n != "rhoc" else float(p[n].grad.abs().max())
            err = float((gbat[i][n] - p[n].grad).abs().max())
            assert err <= 1e-9 * scale, (pick[i], n, err, scale)
        assert abs(float(gbat[i]["E"]) - float(p["E"].grad)) <= 1e-9 * abs(float(p["E"].grad))

    # central differences of the batched forward
    def loss_of(i, name, vox, val):
        ps = [{k: v.detach().clone() for k, v in p.items()} for p in params]
        if name == "E":
            ps[i]["E"] = torch.tensor(val, dtype=torch.float64, device=dev)
        else:
            ps[i][name].view(-1)[vox] = val
        with torch.no_grad():
            _, hh = gb.run([_coef(x, p) for x, p in zip(insts, ps)], T0s,
                           [_src(x, p) for x, p in zip(insts, ps)])
        return float(_loss(hh[i], obs[i]))

    for i in (0, 1):
        ny, nx = insts[i]["shape"][1:]
        for name in ("kz", "rhoc"):
            gall = gbat[i][name].cpu().numpy().ravel()
            top = int(np.argmax(np.abs(gall)))
            vox = [top] + [int(v) for v in rng.integers(0, 2 * ny * nx, size=2)]  # top 2 layers
            g = gall[vox]
            fd = [_richardson(lambda v, vx=vx: loss_of(i, name, vx, v),
                              float(params[i][name].view(-1)[vx])) for vx in vox]
            ok, worst = _grad_ok(g, fd)
            print("instance %d d/d%s at %s: adjoint %s, fd %s (worst %.2f of tol)"
                  % (pick[i], name, vox, g, np.array(fd), worst), flush=True)
            assert ok, (pick[i], name, vox, g, fd, worst)
        fdE = _richardson(lambda v: loss_of(i, "E", None, v), float(params[i]["E"]))
        ok, worst = _grad_ok([float(gbat[i]["E"])], [fdE])
        assert ok, (pick[i], float(gbat[i]["E"]), fdE, worst)

"""Batched forward + adjoint (heatsim2_amd.batch) against the same
instances run one at a time through heatsim2_amd.adjoint.

The batched kernels use one fixed launch shape for every instance while a
single-instance plan picks its own, so results agree to round-off, not
bitwise: histories within 1e-11 K, gradients within 1e-9 of each field's
largest entry.  Graph replay and host-driven launches are bitwise equal.
"""
import numpy as np
import pytest
import torch

import heatsim2_amd as hs
from heatsim2_amd import adjoint, batch

SHAPES = [(6, 5, 7), (9, 4, 3), (4, 8, 6), (3, 3, 3), (12, 10, 9), (2, 17, 33)]
NSTEPS = [7, 12, 3, 20, 9, 5]
SRC = [0, 0, -1, 2, 0, 0]
DT = 1e-3
BNDS = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
        (hs.boundary_thininsulatinglayer, 300.0))


def _instance(i, dev):
    rng = np.random.default_rng(100 + i)
    nz, ny, nx = SHAPES[i]
    d = (3e-4 + 1e-4 * i, 4e-4, 5e-4)
    f = lambda base, s: torch.tensor(base * np.exp(s * rng.standard_normal((nz, ny, nx))),
                                     device=dev, requires_grad=True)
    p = dict(kz=f(0.8, 0.3), ky=f(3.0, 0.3), kx=f(2.0, 0.3), rhoc=f(2e6, 0.2))
    bz = np.zeros((nz + 1, ny, nx), np.uint8)
    by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    bz[0] = bz[-1] = 1
    by[:, 0] = by[:, -1] = 1
    bx[:, :, 0] = bx[:, :, -1] = 1
    if nz > 2:
        bz[nz // 2, :ny // 2 + 1, 1:] = 2
    fixed = torch.zeros((nz, ny, nx), dtype=torch.bool, device=dev)
    if i == 4:
        fixed[8:10, 3:5, 2:4] = True
    mask = torch.zeros((nz, ny, nx), dtype=torch.float64, device=dev)
    mask[0] = 1.0
    E = torch.tensor(1e4 * (1 + 0.1 * i), dtype=torch.float64, device=dev, requires_grad=True)
    T0 = torch.tensor(rng.random((nz, ny, nx)), device=dev)
    obs = torch.tensor(rng.random((NSTEPS[i], ny, nx)), device=dev)
    return dict(p=p, maps=(bz, by, bx), fixed=fixed, mask=mask, E=E, T0=T0, obs=obs, d=d)


def _coef(inst):
    p = inst["p"]
    return adjoint.per_voxel_coefficients(p["kz"], p["ky"], p["kx"], p["rhoc"], DT, inst["d"],
                                          BNDS, *inst["maps"], fixed=inst["fixed"])


def _src(inst, i):
    return inst["E"] * inst["mask"] / (inst["d"][0] * DT)


def _grads(insts):
    out = []
    for inst in insts:
        g = {k: v.grad.detach().cpu().numpy() for k, v in inst["p"].items()}
        g["E"] = None if inst["E"].grad is None else inst["E"].grad.item()
        out.append(g)
    return out


def _run_batch(dev, use_graph, K=4):
    insts = [_instance(i, dev) for i in range(len(SHAPES))]
    gb = batch.GradBatch(SHAPES, [x["d"][0] for x in insts], NSTEPS, SRC, ckpt_every=K,
                         use_graph=use_graph, chunk=3)
    coefs = [_coef(x) for x in insts]
    srcs = [_src(x, i) for i, x in enumerate(insts)]
    hists = gb.surface_history(coefs, [x["T0"] for x in insts], srcs)
    loss = sum(((h - x["obs"]) ** 2).sum() for h, x in zip(hists, insts))
    loss.backward()
    return [h.detach().cpu().numpy() for h in hists], _grads(insts)


def _run_single(dev, K=4):
    insts = [_instance(i, dev) for i in range(len(SHAPES))]
    hists = []
    for i, x in enumerate(insts):
        gp = adjoint.GradPlan(SHAPES[i], x["d"], DT)
        sources = [(SRC[i], _src(x, i))] if SRC[i] >= 0 else []
        h = adjoint.surface_history(gp, _coef(x), x["T0"], NSTEPS[i], sources=sources,
                                    ckpt_every=K)
        ((h - x["obs"]) ** 2).sum().backward()
        hists.append(h.detach().cpu().numpy())
    return hists, _grads(insts)


@pytest.mark.gpu
def test_batch_matches_single_instances():
    dev = torch.device("cuda")
    hb, gb = _run_batch(dev, use_graph=True)
    hs_, gs = _run_single(dev)
    for i in range(len(SHAPES)):
        np.testing.assert_allclose(hb[i], hs_[i], rtol=0, atol=1e-11, err_msg="instance %d" % i)
        for k in ("kz", "ky", "kx", "rhoc"):
            scale = np.abs(gs[i][k]).max()
            assert np.abs(gb[i][k] - gs[i][k]).max() <= 1e-9 * scale, (i, k)
        if SRC[i] >= 0:
            assert abs(gb[i]["E"] - gs[i]["E"]) <= 1e-9 * abs(gs[i]["E"]), i
        else:
            assert gb[i]["E"] in (None, 0.0), i


@pytest.mark.gpu
def test_batch_graph_equals_direct_launches():
    dev = torch.device("cuda")
    hg, gg = _run_batch(dev, use_graph=True)
    hd, gd = _run_batch(dev, use_graph=False)
    for i in range(len(SHAPES)):
        assert np.array_equal(hg[i], hd[i])
        for k in ("kz", "ky", "kx", "rhoc"):
            assert np.array_equal(gg[i][k], gd[i][k])


@pytest.mark.gpu
def test_batch_replays_are_stable():
    """The captured graphs are replayed across calls: a second forward +
    backward on the same batch object gives the same numbers."""
    dev = torch.device("cuda")
    insts = [_instance(i, dev) for i in range(len(SHAPES))]
    gb = batch.GradBatch(SHAPES, [x["d"][0] for x in insts], NSTEPS, SRC, ckpt_every=4)
    results = []
    for rep in range(3):
        for x in insts:
            for v in x["p"].values():
                v.grad = None
        coefs = [_coef(x) for x in insts]
        srcs = [_src(x, i) for i, x in enumerate(insts)]
        hists = gb.surface_history(coefs, [x["T0"] for x in insts], srcs)
        loss = sum(((h - x["obs"]) ** 2).sum() for h, x in zip(hists, insts))
        loss.backward()
        torch.cuda.synchronize()
        results.append((loss.item(), [x["p"]["kz"].grad.cpu().numpy() for x in insts]))
    for L, g in results[1:]:
        assert L == results[0][0]
        for a, b in zip(g, results[0][1]):
            assert np.array_equal(a, b)

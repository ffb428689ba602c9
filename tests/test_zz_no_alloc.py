"""hs2_step_noalloc, the step that allocates nothing, against hs2_step on a
plan given the same per-voxel coefficients (hs2_plan_set_dense): the same
kernels and launch shapes, so the results must be bitwise equal.  Two calls
in a row, the second on a smaller grid (the shared workspace keeps the first
call's contents), then an in-place call; source field and surface readout
included.  Also checked against the oracle's step, and the size limit is
checked on the host."""
import ctypes

import numpy as np
import pytest
import torch

import heatsim2_amd as hs
from heatsim2_amd import _lib, adjoint
import oracle
from hs2_cases import oracle_coef

MAX_CELLS = 1 << 21


def _problem(shape, seed, dev):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    d = (3e-4, 4e-4, 5e-4)
    dt = 2e-3
    f = lambda base, s: torch.tensor(base * np.exp(s * rng.standard_normal(shape)), device=dev)
    kz, ky, kx, rhoc = f(0.8, 0.3), f(4.0, 0.3), f(2.0, 0.3), f(2e6, 0.2)
    bz = np.zeros((nz + 1, ny, nx), np.uint8)
    by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    bz[0] = bz[-1] = 1
    by[:, 0] = by[:, -1] = 1
    bx[:, :, 0] = bx[:, :, -1] = 1
    bz[nz // 2, : ny // 2, 2:] = 2
    fixed = torch.zeros(shape, dtype=torch.bool, device=dev)
    fixed[nz - 2:, 1:3, 1:4] = True
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, 300.0))
    c = adjoint.per_voxel_coefficients(kz, ky, kx, rhoc, dt, d, bnds, bz, by, bx, fixed=fixed)
    coef = [c[k].contiguous() for k in ("M", "w", "gz", "gy", "gx")]
    T = torch.tensor(293.0 + rng.random(shape), device=dev)
    S = torch.tensor(rng.random(shape) * 1e7, device=dev)
    return dict(shape=shape, d=d, dt=dt, coef=coef, T=T, S=S)


def _grid(p):
    nz, ny, nx = p["shape"]
    return _lib.GridDesc(nz, ny, nx, p["d"][0], p["d"][1], p["d"][2], p["dt"])


def _noalloc(p, Tin, Tout, S, surf):
    L = _lib.load()
    _lib.check("hs2_step_noalloc", L.hs2_step_noalloc(
        ctypes.byref(_grid(p)), *[_lib.ptr(t) for t in p["coef"]], _lib.ptr(Tin),
        _lib.ptr(Tout), _lib.ptr(S), _lib.ptr(surf), _lib.stream_handle()))


def _planned(p, Tin, S):
    L = _lib.load()
    gp = adjoint.GradPlan(p["shape"], p["d"], p["dt"])
    M, w, gz, gy, gx = p["coef"]
    _lib.check("hs2_plan_set_dense", L.hs2_plan_set_dense(
        gp.plan, _lib.ptr(M), _lib.ptr(w), _lib.ptr(gz), _lib.ptr(gy), _lib.ptr(gx)))
    out = torch.empty_like(Tin)
    surf = torch.empty(p["shape"][1:], dtype=torch.float64, device=Tin.device)
    _lib.check("hs2_step", L.hs2_step(gp.plan, _lib.ptr(Tin), _lib.ptr(out), None, None, 0,
                                      _lib.ptr(S), _lib.ptr(surf), _lib.stream_handle()))
    torch.cuda.synchronize()
    return out, surf


@pytest.mark.gpu
def test_noalloc_step_equals_planned_step_twice_then_in_place():
    dev = torch.device("cuda")
    for shape, seed in (((48, 72, 100), 1), ((9, 17, 30), 2)):  # second call smaller
        p = _problem(shape, seed, dev)
        out = torch.full_like(p["T"], np.nan)
        surf = torch.full(shape[1:], np.nan, dtype=torch.float64, device=dev)
        _noalloc(p, p["T"], out, p["S"], surf)
        torch.cuda.synchronize()
        ref, ref_surf = _planned(p, p["T"], p["S"])
        print("%s: max |noalloc - hs2_step| = %g" % (shape, float((out - ref).abs().max())),
              flush=True)
        assert torch.equal(out, ref) and torch.equal(surf, ref_surf), shape
        # the oracle's step on the same coefficients
        rows = oracle_coef(*(t.cpu().numpy() for t in p["coef"]))
        To = oracle.fast_step(rows, p["T"].cpu().numpy(), p["S"].cpu().numpy())
        rise = float(np.abs(To - p["T"].cpu().numpy()).max())
        assert float(np.abs(out.cpu().numpy() - To).max()) <= 1e-12 * max(rise, 1.0)
    # in place, no source, no readout: still the planned step's numbers
    T = p["T"].clone()
    _noalloc(p, T, T, None, None)
    ref, _ = _planned(p, p["T"], torch.zeros_like(p["T"]))
    torch.cuda.synchronize()
    assert torch.equal(T, ref)


def test_noalloc_rejects_grids_beyond_its_workspace():
    """Checked on the host before any HIP call (runs without a GPU)."""
    L = _lib.load()
    g = _lib.GridDesc(128, 128, 129, 1e-3, 1e-3, 1e-3, 1e-3)  # 2^21 + 2^14 cells
    dummy = ctypes.c_void_p(256)
    rc = L.hs2_step_noalloc(ctypes.byref(g), dummy, dummy, dummy, dummy, dummy, dummy, dummy,
                            None, None, None)
    assert rc == 2 and "exceeds" in L.hs2_last_error().decode()
    assert 128 * 128 * 128 == MAX_CELLS

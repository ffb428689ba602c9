"""Gradients of a surface-temperature history through the ADI step.

There is no upstream counterpart: heatsim2 has no adjoint.  The forward
model is the same Douglas ADI Crank-Nicolson step as run_adi_steps, on
per-voxel coefficients:

    coef = per_voxel_coefficients(kz, ky, kx, rhoc, dt, (dz, dy, dx), boundaries,
                                  boundary_z_elements, boundary_y_elements,
                                  boundary_x_elements)
    plan = GradPlan((nz, ny, nx), (dz, dy, dx), dt)
    S0 = E * top_mask / (dz * dt)               # a flash: IMPULSE_SOURCE of E/dz J/m^3
    hist = surface_history(plan, coef, T0, nsteps, sources=[(0, S0)], ckpt_every=50)
    loss = ((hist - obs) ** 2).sum()
    loss.backward()                              # kz.grad, ky.grad, kx.grad, rhoc.grad, E.grad

hist[n] is the insulating-top surface temperature (surface_temperature.py:31-36)
after step n.  The coefficients are plain torch expressions of the material
fields, so autograd carries dJ/dM and dJ/dg (from the HIP adjoint) on to k,
rho*c and anything upstream of them.  The adjoint itself is hand-written
HIP (hs2_grad_forward / hs2_grad_backward): transposed line solves in
reverse axis order plus the adjoint stencil, with the forward states
recomputed from checkpoints every ckpt_every steps.
"""
import ctypes

import numpy as np
import torch

from heatsim2_amd import _lib

# face-law kinds of the four boundary modules (boundary_*.py)
_CONDUCTING, _INSULATING, _THIN = 0, 1, 2
_KIND_OF_MODULE = {
    "boundary_conducting": _CONDUCTING,
    "boundary_conducting_anisotropic": _CONDUCTING,
    "boundary_insulating": _INSULATING,
    "boundary_thininsulatinglayer": _THIN,
}


def _boundary_kinds(boundaries):
    kinds, h = [], []
    for b in boundaries:
        name = getattr(b[0], "__name__", "").rsplit(".", 1)[-1]
        if name not in _KIND_OF_MODULE:
            raise NotImplementedError(
                "per-voxel coefficients know the four heatsim2 boundary modules; got %r" % (b[0],))
        kind = _KIND_OF_MODULE[name]
        kinds.append(kind)
        h.append(float(b[1]) if kind == _THIN else 0.0)
    return kinds, h


def per_voxel_coefficients(kz, ky, kx, rhoc, dt, d, boundaries, boundary_z_elements,
                           boundary_y_elements, boundary_x_elements, fixed=None):
    """Differentiable per-voxel ADI coefficients.

    kz, ky, kx, rhoc: (nz, ny, nx) float64 tensors (conductivity per axis in
    W/m/K, volumetric heat capacity in J/m^3/K); `fixed`: optional bool
    tensor of TEMPERATURE_FIXED cells.  Boundary maps and `boundaries` follow
    setup()'s conventions.  Returns dict(M, w, gz, gy, gx) with M = rho*c/dt
    (1 for FIXED), w (uint8, 1 COMPUTE) and g_a = conductance per volume of
    each cell's minus face on axis a (0 on the outer faces):
      conducting / conducting_anisotropic: (k_m + k_p) / (2 d^2) with the
        axis's k, a FIXED neighbour's k replaced by the cell's own
        (crank_nicolson.pyx:298-338; boundary_conducting.py,
        boundary_conducting_anisotropic.py);
      thin insulating layer: h / d (boundary_thininsulatinglayer.py);
      insulating: 0 (boundary_insulating.py).
    Outer faces of COMPUTE cells must be insulating, as in setup().
    """
    shape = tuple(kz.shape)
    dev = kz.device
    kinds, hvals = _boundary_kinds(boundaries)
    kinds_t = torch.tensor(kinds, dtype=torch.int64, device=dev)
    h_t = torch.tensor(hvals, dtype=torch.float64, device=dev)
    if fixed is None:
        fixed = torch.zeros(shape, dtype=torch.bool, device=dev)
    fixed = fixed.to(dev)
    maps = []
    for a, (m, want) in enumerate(((boundary_z_elements, (shape[0] + 1, shape[1], shape[2])),
                                   (boundary_y_elements, (shape[0], shape[1] + 1, shape[2])),
                                   (boundary_x_elements, (shape[0], shape[1], shape[2] + 1)))):
        m = torch.as_tensor(np.asarray(m) if not isinstance(m, torch.Tensor) else m)
        if tuple(m.shape) != want or m.dtype != torch.uint8:
            raise ValueError("boundary map %d must be uint8 of shape %s" % (a, want))
        if int(m.max()) >= len(boundaries):
            raise IndexError("boundary map %d references class %d of %d" %
                             (a, int(m.max()), len(boundaries)))
        maps.append(m.to(dev).long())
    compute = ~fixed
    # outer faces of COMPUTE cells carry no flux (the upstream exits otherwise,
    # alternatingdirection_c.c:158-163)
    for a in range(3):
        kind = kinds_t[maps[a]]
        lo = kind.narrow(a, 0, 1)
        hi = kind.narrow(a, kind.shape[a] - 1, 1)
        c_lo = compute.narrow(a, 0, 1)
        c_hi = compute.narrow(a, shape[a] - 1, 1)
        if bool(((lo != _INSULATING) & c_lo).any()) or bool(((hi != _INSULATING) & c_hi).any()):
            raise ValueError("outer faces of COMPUTE cells must use boundary_insulating (axis %d)" % a)
    out = {}
    M = torch.where(fixed, torch.ones_like(rhoc), rhoc / dt)
    out["M"] = M
    out["w"] = compute.to(torch.uint8)
    for a, (name, k) in enumerate((("gz", kz), ("gy", ky), ("gx", kx))):
        n_a = shape[a]
        km = k.narrow(a, 0, n_a - 1)       # minus-side cell of interior faces
        kp = k.narrow(a, 1, n_a - 1)       # plus-side cell (the face's owner)
        fm = fixed.narrow(a, 0, n_a - 1)
        fp = fixed.narrow(a, 1, n_a - 1)
        km_eff = torch.where(fm, kp, km)
        kp_eff = torch.where(fp, km, kp)
        kind = kinds_t[maps[a].narrow(a, 1, n_a - 1)]
        hh = h_t[maps[a].narrow(a, 1, n_a - 1)]
        da = float(d[a])
        g_int = torch.where(kind == _CONDUCTING, (km_eff + kp_eff) / (2.0 * da * da),
                            torch.where(kind == _THIN, hh / da, torch.zeros_like(km)))
        g_int = torch.where(fm & fp, torch.zeros_like(g_int), g_int)
        zero = torch.zeros_like(k.narrow(a, 0, 1))
        out[name] = torch.cat((zero, g_int), dim=a)
    return out


class GradPlan(object):
    """A device plan in per-voxel coefficient mode for surface_history."""

    def __init__(self, shape, d, dt):
        L = _lib.load()
        self.shape = tuple(int(s) for s in shape)
        self.d = tuple(float(x) for x in d)
        self.dt = float(dt)
        self.n = int(np.prod(self.shape))
        grid = _lib.GridDesc(self.shape[0], self.shape[1], self.shape[2], self.d[0], self.d[1],
                             self.d[2], self.dt)
        h = ctypes.c_void_p()
        _lib.check("hs2_plan_create", L.hs2_plan_create(ctypes.byref(grid), ctypes.byref(h)))
        self.plan = h
        self._coef = None

    def set_coefficients(self, coef):
        """coef: (M, w, gz, gy, gx) device tensors (kept alive by the plan)."""
        coef = tuple(t.detach().contiguous().clone() for t in coef)
        for t in coef:
            if tuple(t.shape) != self.shape:
                raise ValueError("coefficient shape %s != grid %s" % (tuple(t.shape), self.shape))
        L = _lib.load()
        M, w, gz, gy, gx = coef
        _lib.check("hs2_plan_set_dense", L.hs2_plan_set_dense(
            self.plan, _lib.ptr(M), _lib.ptr(w), _lib.ptr(gz), _lib.ptr(gy), _lib.ptr(gx)))
        self._coef = coef
        return coef

    def __del__(self):
        if getattr(self, "plan", None) is not None and _lib._lib is not None:
            _lib._lib.hs2_plan_destroy(self.plan)
            self.plan = None


class _SurfaceHistory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gplan, nsteps, ckpt_every, src_steps, w, M, gz, gy, gx, T0, *src):
        L = _lib.load()
        dev = _lib.device()
        nz, ny, nx = gplan.shape
        coef = gplan.set_coefficients((M, w.to(torch.uint8), gz, gy, gx))
        T = T0.detach().to(dev, torch.float64).contiguous().clone()
        nck = (nsteps + ckpt_every - 1) // ckpt_every
        ckpt = torch.empty((nck,) + gplan.shape, dtype=torch.float64, device=dev)
        srcs = (torch.stack([s.detach().to(dev, torch.float64) for s in src]).contiguous()
                if src else None)
        steps = (ctypes.c_int32 * max(len(src_steps), 1))(*src_steps)
        hist = torch.empty((nsteps, ny, nx), dtype=torch.float64, device=dev)
        _lib.check("hs2_grad_forward", L.hs2_grad_forward(
            gplan.plan, _lib.ptr(T), nsteps, ckpt_every, _lib.ptr(ckpt), _lib.ptr(srcs),
            steps if src_steps else None, len(src_steps), _lib.ptr(hist), _lib.stream_handle()))
        ctx.gplan, ctx.nsteps, ctx.ckpt_every, ctx.src_steps = gplan, nsteps, ckpt_every, src_steps
        ctx.coef, ctx.ckpt, ctx.srcs = coef, ckpt, srcs
        return hist

    @staticmethod
    def backward(ctx, ghist):
        L = _lib.load()
        gplan = ctx.gplan
        dev = ctx.ckpt.device
        if gplan._coef is not ctx.coef:  # the plan ran other coefficients since
            gplan.set_coefficients(ctx.coef)
            ctx.coef = gplan._coef
        ghist = ghist.to(dev, torch.float64).contiguous()
        shape = gplan.shape
        work_n = int(L.hs2_grad_work_doubles(gplan.plan, ctx.ckpt_every))
        work = torch.empty(work_n, dtype=torch.float64, device=dev)
        gM, gz, gy, gx, gT0 = (torch.empty(shape, dtype=torch.float64, device=dev)
                               for _ in range(5))
        nsrc = len(ctx.src_steps)
        gsrc = torch.empty((nsrc,) + shape, dtype=torch.float64, device=dev) if nsrc else None
        steps = (ctypes.c_int32 * max(nsrc, 1))(*ctx.src_steps)
        _lib.check("hs2_grad_backward", L.hs2_grad_backward(
            gplan.plan, ctx.nsteps, ctx.ckpt_every, _lib.ptr(ctx.ckpt), _lib.ptr(ctx.srcs),
            steps if nsrc else None, nsrc, _lib.ptr(ghist), _lib.ptr(gM), _lib.ptr(gz),
            _lib.ptr(gy), _lib.ptr(gx), _lib.ptr(gsrc), _lib.ptr(gT0), _lib.ptr(work),
            ctypes.c_size_t(work_n), _lib.stream_handle()))
        gsrcs = tuple(gsrc[i] for i in range(nsrc)) if nsrc else ()
        return (None, None, None, None, None, gM, gz, gy, gx, gT0) + gsrcs


def surface_history(gplan, coef, T0, nsteps, sources=(), ckpt_every=50):
    """Differentiable surface-temperature history (nsteps, ny, nx).

    gplan: GradPlan; coef: dict from per_voxel_coefficients (or any tensors
    M, w, gz, gy, gx of the grid's shape); T0: initial field; sources:
    sequence of (step, field) with field the volumetric power density
    (W/m^3) added during that step (at most one per step); ckpt_every:
    steps between checkpoints (memory: ceil(nsteps/ckpt_every) +
    3*ckpt_every + 6 fields).  Gradients flow to M, g, T0 and the source fields.
    """
    nsteps, ckpt_every = int(nsteps), int(ckpt_every)
    if nsteps < 1 or ckpt_every < 1:
        raise ValueError("nsteps and ckpt_every must be >= 1")
    steps = tuple(int(s) for s, _ in sources)
    fields = [f for _, f in sources]
    dev = _lib.device()
    args = [coef["M"], coef["gz"], coef["gy"], coef["gx"]]
    args = [a.to(dev) for a in args]
    fields = [f.to(dev) for f in fields]
    return _SurfaceHistory.apply(gplan, nsteps, ckpt_every, steps, coef["w"].to(dev), *args,
                                 T0.to(dev), *fields)

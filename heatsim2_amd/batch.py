"""Batched forward + adjoint of many independent problems (no upstream
counterpart).

    gb = GradBatch(shapes, dz, nsteps, src_steps, ckpt_every=50)
    hists = gb.surface_history(coefs, T0s, srcs)     # list of (nsteps_i, ny_i, nx_i)
    loss = sum(((h - o) ** 2).sum() for h, o in zip(hists, obs))
    loss.backward()

Every instance has its own grid (axes up to 128 cells), step count,
per-voxel coefficients (per_voxel_coefficients of heatsim2_amd.adjoint) and
source field applied at its src_step.  The HIP library runs all of them in
one launch per sweep, with the time loop captured as a HIP graph; the
results equal what adjoint.surface_history gives for each instance alone.
"""
import ctypes

import numpy as np
import torch

from heatsim2_amd import _lib


class hs2_batch_inst(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("nsteps", ctypes.c_int32), ("src_step", ctypes.c_int32),
                ("dz", ctypes.c_double)]


class GradBatch(object):
    """A batch of instances: shapes [(nz, ny, nx)], dz [float] (z spacing,
    for the readout), nsteps [int], src_steps [int or -1]."""

    def __init__(self, shapes, dz, nsteps, src_steps=None, ckpt_every=50, use_graph=True,
                 chunk=8):
        L = _lib.load()
        self.shapes = [tuple(int(s) for s in sh) for sh in shapes]
        m = len(self.shapes)
        dz = list(dz) if np.ndim(dz) else [float(dz)] * m
        self.nsteps = [int(n) for n in nsteps]
        src_steps = [0] * m if src_steps is None else [int(s) for s in src_steps]
        arr = (hs2_batch_inst * m)()
        for i in range(m):
            nz, ny, nx = self.shapes[i]
            arr[i].nz, arr[i].ny, arr[i].nx = nz, ny, nx
            arr[i].nsteps, arr[i].src_step, arr[i].dz = self.nsteps[i], src_steps[i], float(dz[i])
        h = ctypes.c_void_p()
        _lib.check("hs2_batch_create", L.hs2_batch_create(arr, m, int(ckpt_every), ctypes.byref(h)))
        self.handle = h
        _lib.check("hs2_batch_set_options", L.hs2_batch_set_options(h, int(bool(use_graph)),
                                                                    int(chunk)))
        coff = (ctypes.c_int64 * (m + 1))()
        hoff = (ctypes.c_int64 * (m + 1))()
        sizes = (ctypes.c_int64 * 5)()
        _lib.check("hs2_batch_layout", L.hs2_batch_layout(h, coff, hoff, sizes))
        self.cell_off = list(coff)
        self.hist_off = list(hoff)
        self.N, self.H, self.work_doubles, self.K, self.maxsteps = [int(v) for v in sizes]
        self._work = None
        self._coef = None
        self._generation = 0

    def __len__(self):
        return len(self.shapes)

    def pack(self, fields, dtype=torch.float64):
        """Concatenate per-instance (nz, ny, nx) tensors into the batch layout
        (differentiable; padding cells are zero)."""
        dev = _lib.device()
        parts = []
        for i, f in enumerate(fields):
            f = f.to(dev, dtype).reshape(-1)
            parts.append(f)
            pad = self.cell_off[i + 1] - self.cell_off[i] - f.numel()
            if pad:
                parts.append(torch.zeros(pad, dtype=dtype, device=dev))
        return torch.cat(parts)

    def unpack(self, flat):
        return [flat[self.cell_off[i]:self.cell_off[i] + int(np.prod(self.shapes[i]))]
                .view(self.shapes[i]) for i in range(len(self))]

    def unpack_hist(self, flat):
        return [flat[self.hist_off[i]:self.hist_off[i + 1]]
                .view(self.nsteps[i], self.shapes[i][1], self.shapes[i][2])
                for i in range(len(self))]

    def work(self):
        if self._work is None:
            self._work = torch.empty(self.work_doubles, dtype=torch.float64,
                                     device=_lib.device())
        return self._work

    def run(self, coefs, T0s, srcs=None):
        """Forward only, no gradients: the final fields and the surface
        histories of every instance (two lists of per-instance tensors)."""
        packed = self._packed(coefs, T0s, srcs)
        with torch.no_grad():
            T, hist = _forward(self, *packed)
        return self.unpack(T), self.unpack_hist(hist)

    def _packed(self, coefs, T0s, srcs):
        M = self.pack([c["M"] for c in coefs])
        gz = self.pack([c["gz"] for c in coefs])
        gy = self.pack([c["gy"] for c in coefs])
        gx = self.pack([c["gx"] for c in coefs])
        w = self.pack([c["w"] for c in coefs], dtype=torch.uint8)
        T0 = self.pack(T0s)
        src = self.pack(srcs) if srcs is not None else None
        return w, M, gz, gy, gx, T0, src

    def surface_history(self, coefs, T0s, srcs=None):
        """coefs: per-instance dicts (M, w, gz, gy, gx); T0s: initial fields;
        srcs: per-instance source fields (W/m^3) or None.  Returns the list
        of per-instance surface histories; gradients flow to the
        coefficients and the source fields."""
        return list(_BatchHistory.apply(self, *self._packed(coefs, T0s, srcs)))

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib._lib is not None:
            _lib._lib.hs2_batch_destroy(self.handle)
            self.handle = None


def _forward(gb, w, M, gz, gy, gx, T0, src):
    """hs2_batch_forward on packed fields: returns (final fields, histories)
    and keeps the checkpoints in the batch's work buffer."""
    L = _lib.load()
    dev = _lib.device()
    coef = tuple(t.detach().contiguous() for t in (M, w, gz, gy, gx))
    _lib.check("hs2_batch_set_dense", L.hs2_batch_set_dense(gb.handle, *[_lib.ptr(t) for t in coef]))
    gb._coef = coef
    gb._generation += 1
    T = T0.detach().contiguous().clone()
    srcd = src.detach().contiguous() if src is not None else None
    gb._src = srcd
    hist = torch.empty(gb.H, dtype=torch.float64, device=dev)
    work = gb.work()
    _lib.check("hs2_batch_forward", L.hs2_batch_forward(
        gb.handle, _lib.ptr(T), _lib.ptr(srcd), _lib.ptr(hist), _lib.ptr(work),
        ctypes.c_size_t(gb.work_doubles), _lib.stream_handle()))
    return T, hist


class _BatchHistory(torch.autograd.Function):
    """Per-instance histories as separate outputs: their gradients arrive as
    one tuple and are concatenated once (slices of one returned tensor would
    each scatter a gradient the size of the whole history)."""

    @staticmethod
    def forward(ctx, gb, w, M, gz, gy, gx, T0, src):
        _, hist = _forward(gb, w, M, gz, gy, gx, T0, src)
        ctx.gb, ctx.coef, ctx.src, ctx.generation = gb, gb._coef, gb._src, gb._generation
        ctx.want_src = src is not None
        return tuple(h.clone() for h in gb.unpack_hist(hist))

    @staticmethod
    def backward(ctx, *ghists):
        gb = ctx.gb
        if gb._generation != ctx.generation:
            raise RuntimeError("GradBatch: another forward ran since this history was computed; "
                               "backward needs the checkpoints of its own forward")
        L = _lib.load()
        dev = _lib.device()
        ghist = torch.cat([(g if g is not None else
                            torch.zeros((gb.nsteps[i],) + gb.shapes[i][1:], dtype=torch.float64,
                                        device=dev)).reshape(-1).to(torch.float64)
                           for i, g in enumerate(ghists)])
        gM, gz, gy, gx = (torch.empty(gb.N, dtype=torch.float64, device=dev) for _ in range(4))
        gsrc = torch.empty(gb.N, dtype=torch.float64, device=dev) if ctx.want_src else None
        work = gb.work()
        _lib.check("hs2_batch_backward", L.hs2_batch_backward(
            gb.handle, _lib.ptr(ctx.src), _lib.ptr(ghist), _lib.ptr(gM), _lib.ptr(gz),
            _lib.ptr(gy), _lib.ptr(gx), _lib.ptr(gsrc), _lib.ptr(work),
            ctypes.c_size_t(gb.work_doubles), _lib.stream_handle()))
        return None, None, gM, gz, gy, gx, None, gsrc

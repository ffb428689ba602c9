"""Benchmark: FP64 ADI cell-steps/s of the plain step at 256^3 on one GPU.

Workload: the two-layer stack -- 256x256x256 cells of 0.5 mm, steel
(k=50, rho=7850, c=470) in the first 32 z-layers over foam (k=0.03, rho=30,
c=1500), a thin insulating layer (h=250 W m^-2 K^-1) between them,
insulating sides, a flash of 1e4*(1+0.05 U(-1,1)) J/m^2 (seed 1) into the
top layer at t=0, dt=0.5 ms, FP64, surface readout captured every step.
Synthetic data of that shape; the uint8 maps and coefficient tables are
built by heatsim2_amd.setup exactly as a user would.

A step is one full Douglas ADI step (x, y, z substeps) over all cells.
`value` = cells x K / wall time of K steps run by hs2_run with the field and
maps resident in HBM, bracketed by torch.cuda.synchronize().

roofline: the dominant sweep kernel's algorithmic bytes per launch (the FP64
field passes: x 16 B per cell, z 24 B; y 24 B, or 16 B when the x sweep
hands it its right-hand side, as on this workload -- 56 B per cell-step in
all, against the canonical 64 B of the delta-form step) over its mean
launch duration, from HIP
events recorded on the launch stream around every substep kernel of
--timing-steps more steps run right after the timed region (hs2_run_timed,
events created with hipEventDisableSystemFence: an event with the default
system-scope fence writes back and invalidates the caches, which lengthens
the kernel it brackets by a few microseconds; those timings are reported
beside them), less the dispatch gap each fence-free pair still adds (the
pairs' sum over the step time of a run with no events between kernels,
split in three: event_pair_overhead_ms); peak = 8 TB/s.  traffic: the HBM bytes per launch of that kernel read from
the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
(profiles/hbm_traffic_two_layer_256.txt, copy-calibrated).  copy_GBps: a
device-to-device copy of one field on the same GPU in the same run (read + write bytes per
copy, events on torch's stream); step_frac_of_copy puts the step's canonical
64 B per cell-step against it.

single_gradient: one gradient evaluation of the single-instance inversion
workload (64x128x128 laminate, 500 steps, checkpoints every 50).
batched: gradients/s of 256 uneven instances through the batched forward +
adjoint against the same instances one after another.

cpu_baseline: the oracle's reference-structured step (assembled permuted
CSR operators, global Thomas with precomputed LU, transposes -- the
upstream's structure) on one host thread, over a few steps of the same
256^3 workload, reported as cell-steps/s, with the host CPUs available to
the process beside the one it uses.  Also: the same one-thread step over the
whole 40^3 steel-cube CPU workload (200 steps), and the oracle's OpenMP
delta-form step on all host cores at 256^3 (4 steps).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = (16.0, 24.0, 24.0)  # x, y, z sweep algorithmic bytes (y: 16 from x's R_y)
CANONICAL_STEP_BYTES = 64.0  # the delta-form step's field passes, the metric's yardstick
# committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries per workload
TRAFFIC_SUMMARY = {"two_layer": "hbm_traffic_two_layer_256.txt",
                   "laminate": "hbm_traffic_laminate_128x512x512.txt"}
METRIC = "FP64 ADI cell-steps/s @256³ + batched fwd+adjoint gradients/s vs plain"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_copy_gbs(T, reps=20):
    """Bandwidth of a device-to-device copy of one field (bytes read plus
    bytes written per copy), timed with events on torch's stream."""
    import torch
    a = torch.empty_like(T)
    a.copy_(T)
    b = torch.empty_like(T)
    b.copy_(a)  # warm
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * a.numel() * a.element_size()
    del a, b
    return nbytes / (ms * 1e-3) / 1e9


def _ref_step_rate(case, nsteps):
    """cell-steps/s of the oracle's reference-structured step (one thread)
    over nsteps source-free steps of a case, and the untimed assembly time"""
    import oracle
    t0 = time.time()
    plan = oracle.RefPlan(case["shape"], oracle.assemble_case(case))
    t_setup = time.time() - t0
    T = np.array(case["T0"])
    t0 = time.time()
    for _ in range(nsteps):
        T = plan.step(T)
    dt = time.time() - t0
    del plan
    return int(np.prod(case["shape"])) * nsteps / dt, dt, t_setup


def cpu_baseline(case, nsteps, laminate_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from hs2_cases import cases, load_golden
    value, dt, t_setup = _ref_step_rate(case, nsteps)
    n = int(np.prod(case["shape"]))
    # the same step over the whole CPU reference workload (40^3 steel cube,
    # 200 steps), with the upstream's own wall time for it from the golden
    # fixture: both were timed one after the other on the fixture-writing
    # host, which gives the port's speed relative to the upstream solver
    c0 = cases.case_steel_cube(40, 200)
    p0 = oracle.RefPlan(c0["shape"], oracle.assemble_case(c0))
    T0 = np.array(c0["T0"])
    t1 = time.time()
    for _ in range(c0["nsteps"]):
        T0 = p0.step(T0)
    cube = int(np.prod(c0["shape"])) * c0["nsteps"] / (time.time() - t1)
    del p0
    fx = load_golden("steel_cube40")["fixture"]
    port_over_ref = float(fx["ref_run_s"]) / float(fx["port_run_s"])
    # the anisotropic laminate, 128 x 512 x 512, a few steps
    lam = cases.case_laminate(128, 512, 512, nsteps=laminate_steps)
    lam_rate, lam_dt, lam_setup = _ref_step_rate(lam, laminate_steps)
    # the oracle's OpenMP delta-form step on all host cores
    coef = oracle.assemble_case(case)
    t1 = time.time()
    oracle.fast_run(coef, np.array(case["T0"]), 4)
    omp = n * 4 / (time.time() - t1)
    return dict(value=value, unit="cell-steps/s", cores=1, kind="port",
                host_cpus_available=len(os.sched_getaffinity(0)), host_cpus=os.cpu_count(),
                steel_cube_40_200_steps=cube,
                port_speed_over_reference=round(port_over_ref, 4),
                reference_equivalent_value=value / port_over_ref,
                laminate_128x512x512=lam_rate,
                omp_all_cores=omp,
                omp_threads=int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0)))),
                sample="%d source-free steps of the same %dx%dx%d two-layer workload through the "
                       "oracle's reference-structured step (permuted CSR B/C operators, global "
                       "Thomas with precomputed LU, transposes), one thread; %.1f s timed, "
                       "assembly %.1f s untimed. steel_cube_40_200_steps: the same step over "
                       "all 200 steps of the 40x40x40 steel cube, one thread. "
                       "port_speed_over_reference: the upstream solver's own time over the port's "
                       "on that 40^3 x 200-step case, both timed on the fixture-writing host "
                       "(%.3f s vs %.3f s; the upstream cannot run on this host); "
                       "reference_equivalent_value = value / port_speed_over_reference. "
                       "laminate_128x512x512: %d source-free steps of the 128x512x512 laminate "
                       "through the same one-thread step (%.1f s timed, assembly %.1f s). "
                       "omp_all_cores: the oracle's OpenMP delta-form step, 4 steps of the "
                       "%dx%dx%d workload on omp_threads threads"
                       % ((nsteps,) + tuple(case["shape"]) + (dt, t_setup) +
                          (float(fx["ref_run_s"]), float(fx["port_run_s"]), laminate_steps,
                           lam_dt, lam_setup) + tuple(case["shape"])))


def single_gradient(hs, log, reps=3):
    """The single-instance inversion workload: one gradient evaluation
    (per-voxel coefficients, forward with checkpoints, loss, hand-written
    adjoint, chain rule to kx, ky, kz, rho*c and the flash amplitude) of the
    64x128x128 anisotropic laminate, 500 steps at dt = 2 ms, checkpoints every
    50 steps, at the homogeneous guess against the history of a smooth
    random kz field (seed 3, sigma 8 cells) plus 0.01 K noise (seed 4)."""
    import torch
    from scipy.ndimage import gaussian_filter
    from hs2_cases import cases
    from heatsim2_amd import _lib
    dev = _lib.device()
    nsteps, K = 500, 50
    case = cases.case_laminate(64, 128, 128, nsteps=nsteps)
    shape = case["shape"]
    g = gaussian_filter(np.random.default_rng(3).standard_normal(shape), sigma=8.0)
    g = (g - g.mean()) / g.std()
    _, bz, by, bx, _ = case["maps"]
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, case["boundaries"][2][1]))
    _fixed, kz0, ky0, kx0, rho, c, _t = case["materials"][0]
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    base = dict(kz=t(np.full(shape, kz0)), ky=t(np.full(shape, ky0)), kx=t(np.full(shape, kx0)),
                rhoc=t(np.full(shape, rho * c)), E=t(1e4))
    mask = torch.zeros(shape, dtype=torch.float64, device=dev)
    mask[0] = 1.0
    plan = hs.adjoint.GradPlan(shape, case["d"], case["dt"])
    T0 = torch.zeros(shape, dtype=torch.float64, device=dev)

    def history(p):
        coef = hs.adjoint.per_voxel_coefficients(p["kz"], p["ky"], p["kx"], p["rhoc"],
                                                 case["dt"], case["d"], bnds, bz, by, bx)
        S0 = p["E"] * mask / (case["d"][0] * case["dt"])
        return hs.adjoint.surface_history(plan, coef, T0, nsteps, sources=[(0, S0)],
                                          ckpt_every=K)
    with torch.no_grad():
        h = history(dict(base, kz=t(kz0 * (1.0 + 0.3 * g))))
        obs = h + 0.01 * t(np.random.default_rng(4).standard_normal(tuple(h.shape)))
    times = []
    for rep in range(reps + 1):  # the first evaluation warms the code objects
        p = {k: v.clone().requires_grad_(True) for k, v in base.items()}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = ((history(p) - obs) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    cells = int(np.prod(shape))
    log("single-instance gradient (64x128x128 laminate, 500 steps, K=50): %.3f s per evaluation"
        % sec)
    return dict(workload="64x128x128 anisotropic laminate, 500 steps at dt = 2 ms, checkpoints "
                         "every 50 steps; gradients w.r.t. per-voxel kx, ky, kz, rho*c and the "
                         "flash amplitude of an L2 surface-history misfit, FP64",
                seconds=sec, gradients_per_s=1.0 / sec,
                cell_steps_per_s=cells * nsteps / sec, evaluations_timed=reps)


def strongly_coupled(hs, size, steps, log):
    """The 256^3 step where the line solvers' early exit never fires: a steel
    block (k=50, rho c=7850*470, 0.5 mm cells) at dt = 50 ms, k dt / (rho c
    dx^2) = 2.7 on every axis, so every line's couplings after the first PCR
    level stay far above 2^-64 and every level runs."""
    import ctypes
    import torch
    from heatsim2_amd import _lib
    from hs2_cases import cases, to_package_args
    case = cases.case_steel_block(*size, nsteps=steps, dt=0.05)
    pa = to_package_args(hs, case)
    P, S = hs.setup(*pa["setup"])
    dev = _lib.device()
    T = torch.from_numpy(case["T0"]).to(dev)
    cfg = hs.autotune.tune(P, T, use_cache=False)
    T, _ = hs.run_adi_steps_many(P, 0.0, case["dt"], 5, T, pa["vmap"], pa["volumetric"])
    L = _lib.load()
    stream = _lib.stream_handle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check("hs2_run", L.hs2_run(P.plan, _lib.ptr(T), steps, None, None, 1, None, stream))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    arr = (ctypes.c_double * 4)()
    _lib.check("hs2_run_timed", L.hs2_run_timed(P.plan, _lib.ptr(T), steps, 1, arr, stream))
    torch.cuda.synchronize()
    assert torch.isfinite(T).all().item()
    n = int(np.prod(size))
    info = _lib.plan_info(P.plan)
    P.release()
    res = dict(workload="steel block %dx%dx%d, 0.5 mm cells, dt = 50 ms (k dt / (rho c dx^2) = "
                        "2.7: every PCR level runs on every line)" % tuple(size),
               value=n * steps / wall, unit="cell-steps/s", ms_per_step=1e3 * wall / steps,
               steps=steps, kernel_ms=[round(v, 5) for v in arr][:3],
               launch={k: info[k] for k in ("x_cpl", "x_lanes", "x_lines", "x_lines_y", "y_seg",
                                            "y_lines", "z_seg", "z_lines", "y_from_x")},
               step_frac=round(CANONICAL_STEP_BYTES * n * steps / wall / 1e9 / HBM_PEAK_GBS, 4))
    log("strongly coupled: %.4f ms/step, %.3e cell-steps/s, kernels %s (tuned %s)"
        % (res["ms_per_step"], res["value"], res["kernel_ms"], cfg))
    return res


# ---------------------------------------------------------------- batched
BATCH_DT = 2e-3
BATCH_D = (5e-4, 5e-4, 5e-4)
BATCH_H = 100.0


def batch_workload(ninst=256, seed=6, lo=32, hi=128, smin=100, smax=1000):
    """Uneven instances: axes U{lo..hi}, steps U{smin..smax}, isotropic or
    anisotropic (p = 1/2) k log-uniform on [0.1, 60] W/m/K per diagonal
    component, rho*c log-uniform on [1e5, 5e6] J/m^3/K, 0-4 thin-layer
    patches (h = 100 W/m^2/K) at random z-faces, flash E = 1e4 (1 + 0.2 U)
    J/m^2 on the top layer at step 0; 0.5 mm cells, dt = 2 ms."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ninst):
        nz, ny, nx = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        nsteps = int(rng.integers(smin, smax + 1))
        aniso = bool(rng.random() < 0.5)
        lk = lambda: float(np.exp(rng.uniform(np.log(0.1), np.log(60.0))))
        k = (lk(), lk(), lk()) if aniso else (lk(),) * 3
        rhoc = float(np.exp(rng.uniform(np.log(1e5), np.log(5e6))))
        bz = np.zeros((nz + 1, ny, nx), np.uint8)
        by = np.zeros((nz, ny + 1, nx), np.uint8)
        bx = np.zeros((nz, ny, nx + 1), np.uint8)
        bz[0] = bz[-1] = 1
        by[:, 0] = by[:, -1] = 1
        bx[:, :, 0] = bx[:, :, -1] = 1
        for _p in range(int(rng.integers(0, 5))):
            kz = int(rng.integers(1, nz))
            sy, sx = int(rng.integers(1, ny + 1)), int(rng.integers(1, nx + 1))
            y0, x0 = int(rng.integers(0, ny - sy + 1)), int(rng.integers(0, nx - sx + 1))
            bz[kz, y0:y0 + sy, x0:x0 + sx] = 2
        E = 1e4 * (1 + 0.2 * rng.uniform(-1, 1))
        out.append(dict(shape=(nz, ny, nx), nsteps=nsteps, k=k, rhoc=rhoc, aniso=aniso,
                        maps=(bz, by, bx), E=E))
    return out


class _BatchProblem(object):
    """Per-instance leaf parameters (homogeneous initial guess), sources and
    observations of one batch workload."""

    def __init__(self, hs, insts, dev):
        import torch
        self.hs, self.insts, self.dev = hs, insts, dev
        self.params = []
        for x in insts:
            sh = x["shape"]
            p = {n: torch.full(sh, v, dtype=torch.float64, device=dev, requires_grad=True)
                 for n, v in (("kz", x["k"][0]), ("ky", x["k"][1]), ("kx", x["k"][2]),
                              ("rhoc", x["rhoc"]))}
            p["E"] = torch.tensor(x["E"], dtype=torch.float64, device=dev, requires_grad=True)
            mask = torch.zeros(sh, dtype=torch.float64, device=dev)
            mask[0] = 1.0
            p["mask"] = mask
            p["T0"] = torch.zeros(sh, dtype=torch.float64, device=dev)
            self.params.append(p)
        self.obs = None

    def boundaries(self, x):
        hs = self.hs
        law = hs.boundary_conducting_anisotropic if x["aniso"] else hs.boundary_conducting
        return ((law,), (hs.boundary_insulating,), (hs.boundary_thininsulatinglayer, BATCH_H))

    def coef(self, i, scale=None):
        x, p = self.insts[i], self.params[i]
        f = (lambda t: t) if scale is None else scale
        return self.hs.adjoint.per_voxel_coefficients(f(p["kz"]), f(p["ky"]), f(p["kx"]),
                                                      f(p["rhoc"]), BATCH_DT, BATCH_D,
                                                      self.boundaries(x), *x["maps"])

    def src(self, i, factor=1.0):
        p = self.params[i]
        return factor * p["E"] * p["mask"] / (BATCH_D[0] * BATCH_DT)

    def zero_grad(self):
        for p in self.params:
            for n in ("kz", "ky", "kx", "rhoc", "E"):
                p[n].grad = None


def batched_kernels(hs, ninst, seed, ckpt_every, log, use_graph=True, reps=1):
    """The batched library calls alone (hs2_batch_forward, then
    hs2_batch_backward of a fixed history gradient), with no autograd or
    torch arithmetic in between: the run that per-kernel counters are
    collected on.  Returns seconds per forward and per backward."""
    import ctypes
    import torch
    from heatsim2_amd import _lib, batch
    dev = _lib.device()
    L = _lib.load()
    insts = batch_workload(ninst, seed)
    prob = _BatchProblem(hs, insts, dev)
    gb = hs.batch.GradBatch([x["shape"] for x in insts], BATCH_D[0],
                            [x["nsteps"] for x in insts], [0] * ninst, ckpt_every=ckpt_every,
                            use_graph=use_graph)
    with torch.no_grad():
        packed = gb._packed([prob.coef(i) for i in range(ninst)],
                            [p["T0"] for p in prob.params],
                            [prob.src(i) for i in range(ninst)])
        ghist = torch.full((gb.H,), 1e-3, dtype=torch.float64, device=dev)
        grads = [torch.empty(gb.N, dtype=torch.float64, device=dev) for _ in range(5)]
        tf = tb = 0.0
        for rep in range(reps + 1):  # the first pass captures the graphs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            batch._forward(gb, *packed)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _lib.check("hs2_batch_backward", L.hs2_batch_backward(
                gb.handle, _lib.ptr(gb._src), _lib.ptr(ghist),
                *[_lib.ptr(g) for g in grads], _lib.ptr(gb.work()),
                ctypes.c_size_t(gb.work_doubles), _lib.stream_handle()))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep:
                tf += (t1 - t0) / reps
                tb += (t2 - t1) / reps
    log("batched library calls: forward %.3f s, backward %.3f s" % (tf, tb))
    return {"forward_s": tf, "backward_s": tb, "instances": ninst}


def batched_gradients(hs, ninst, seed, ckpt_every, plain_sample, log, use_graph=True):
    """Gradient evaluations per second (forward + loss + adjoint for every
    instance) of the batched path and of the plain per-instance path."""
    import torch
    from heatsim2_amd import _lib
    dev = _lib.device()
    insts = batch_workload(ninst, seed)
    prob = _BatchProblem(hs, insts, dev)
    cells = sum(int(np.prod(x["shape"])) for x in insts)
    cell_steps = sum(int(np.prod(x["shape"])) * x["nsteps"] for x in insts)
    log("batch: %d instances, %.3g cells, %.3g cell-steps per evaluation"
        % (ninst, cells, cell_steps))
    gb = hs.batch.GradBatch([x["shape"] for x in insts], BATCH_D[0],
                            [x["nsteps"] for x in insts], [0] * ninst, ckpt_every=ckpt_every,
                            use_graph=use_graph)
    T0s = [p["T0"] for p in prob.params]
    # observations: the history of a perturbed "true" field plus 0.01 K noise
    rng = np.random.default_rng(seed + 1)
    with torch.no_grad():
        truth = lambda t: t * 1.1
        hist = gb.surface_history([prob.coef(i, truth) for i in range(ninst)], T0s,
                                  [prob.src(i, 1.05) for i in range(ninst)])
        prob.obs = [h + 0.01 * torch.tensor(rng.standard_normal(tuple(h.shape)), device=dev)
                    for h in hist]

    def batched_eval():
        prob.zero_grad()
        coefs = [prob.coef(i) for i in range(ninst)]
        srcs = [prob.src(i) for i in range(ninst)]
        hists = gb.surface_history(coefs, T0s, srcs)
        loss = sum(((h - o) ** 2).sum() for h, o in zip(hists, prob.obs))
        loss.backward()
        return loss

    batched_eval()  # captures the graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss_b = batched_eval().item()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log("batched: %.3f s per evaluation, %.2f gradients/s" % (t_batch, ninst / t_batch))

    # plain: the same instances one after another through the single-instance path
    idx = list(range(ninst)) if plain_sample <= 0 else \
        sorted(np.random.default_rng(seed + 2).choice(ninst, min(plain_sample, ninst),
                                                      replace=False).tolist())

    def plain_eval(i):
        x, p = insts[i], prob.params[i]
        gp = hs.adjoint.GradPlan(x["shape"], BATCH_D, BATCH_DT)
        h = hs.adjoint.surface_history(gp, prob.coef(i), p["T0"], x["nsteps"],
                                       sources=[(0, prob.src(i))], ckpt_every=ckpt_every)
        loss = ((h - prob.obs[i]) ** 2).sum()
        loss.backward()
        return loss

    gk = {i: {n: prob.params[i][n].grad.clone() for n in ("kz", "ky", "kx", "rhoc")}
          for i in idx[:4]}
    prob.zero_grad()
    plain_eval(idx[0])  # warm-up (code objects, allocator)
    prob.zero_grad()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss_p = 0.0
    for i in idx:
        loss_p += plain_eval(i).item()
    torch.cuda.synchronize()
    t_plain_sample = time.perf_counter() - t0
    t_plain = t_plain_sample * ninst / len(idx)
    log("plain: %.3f s for %d instances -> %.3f s per evaluation of all %d, %.2f gradients/s"
        % (t_plain_sample, len(idx), t_plain, ninst, ninst / t_plain))
    # the two paths agree (same instances, same loss terms)
    # the conductivity gradients share one scale (same units; a lateral one is
    # zero up to round-off when nothing breaks the lateral symmetry)
    dev_rel = 0.0
    for i, gi in gk.items():
        kscale = max(float(gi[n].abs().max()) for n in ("kz", "ky", "kx"))
        for n, g in gi.items():
            scale = kscale if n != "rhoc" else float(g.abs().max())
            gp = prob.params[i][n].grad
            dev_rel = max(dev_rel, float((gp - g).abs().max()) / scale)
    return dict(metric="batched fwd+adjoint gradients/s vs plain", instances=ninst,
                gradients_per_s=ninst / t_batch, plain_gradients_per_s=ninst / t_plain,
                speedup=t_plain / t_batch, seconds_batched=round(t_batch, 3),
                seconds_plain=round(t_plain, 3),
                plain_sample=("all" if len(idx) == ninst else
                              "%d of %d instances, time scaled by %d/%d" % (len(idx), ninst,
                                                                            ninst, len(idx))),
                cells=cells, cell_steps=cell_steps, ckpt_every=ckpt_every,
                loss=loss_b, max_rel_grad_diff_vs_plain=dev_rel,
                workload="%d uneven instances (axes 32-128, 100-1000 steps, seed %d), per-voxel "
                         "kz/ky/kx/rho*c and flash-amplitude gradients of an L2 surface-history "
                         "misfit, FP64" % (ninst, seed))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--cpu-laminate-steps", type=int, default=3)
    ap.add_argument("--timing-steps", type=int, default=500,
                    help="steps per run of the per-kernel timing (hs2_run_timed)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("two_layer", "laminate"), default="two_layer",
                    help="two_layer: the 256^3 steel/foam stack (the metric's); laminate: the "
                         "128x512x512 anisotropic laminate")
    ap.add_argument("--size", type=int, nargs=3, default=None)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--config", default=None,
                    help="force launch shapes, e.g. 'x=4,64,0;y=16,16;z=16,16' (skips the tuner)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-gradient measurement")
    ap.add_argument("--no-strong", action="store_true", help="skip the strongly coupled step")
    ap.add_argument("--strong-steps", type=int, default=200)
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--batch-kernels", action="store_true",
                    help="only the batched forward and backward library calls (for counters)")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-seed", type=int, default=6)
    ap.add_argument("--ckpt-every", type=int, default=50)
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="launch the batched steps from the host instead of replaying graphs "
                         "(for counter collection)")
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="time the plain path on this many instances and scale (0: all)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256^3 float64 field 20 times (scales HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.rand(256 ** 3, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 8, "copies": 20}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import heatsim2_amd as hs
    from heatsim2_amd import _lib
    from hs2_cases import cases, to_package_args

    if args.batch_kernels:
        res = batched_kernels(hs, args.batch_instances, args.batch_seed, args.ckpt_every, log,
                              use_graph=not args.batch_no_graph)
        print(json.dumps(res), flush=True)
        return
    if args.batch_only:
        res = batched_gradients(hs, args.batch_instances, args.batch_seed, args.ckpt_every,
                                args.plain_sample, log, use_graph=not args.batch_no_graph)
        print(json.dumps(res), flush=True)
        return

    if args.size is None:
        args.size = (256, 256, 256) if args.workload == "two_layer" else (128, 512, 512)
    nz, ny, nx = args.size
    if args.workload == "two_layer":
        case = cases.case_two_layer(nz=nz, ny=ny, nx=nx, nsteps=1000)
        wl = ("two-layer steel/foam stack %dx%dx%d (z,y,x), dx=0.5 mm, thin layer h=250 W/m^2K, "
              "flash %.1f J/m^2, dt=0.5 ms, FP64, surface readout every step"
              % (nz, ny, nx, case["flash_J_m2"]))
    else:
        case = cases.case_laminate(nz, ny, nx, nsteps=2000)
        wl = ("anisotropic laminate %dx%dx%d (z,y,x), kz=0.6, kx=ky=5 W/mK, rho c=1550*1200, "
              "%d thin-layer delaminations (h=100 W/m^2K), flash 1e4 J/m^2, dt=2 ms, FP64, "
              "surface readout every step" % (nz, ny, nx, case["npatch"]))
    pa = to_package_args(hs, case)
    t0 = time.time()
    P, S = hs.setup(*pa["setup"])
    log("setup %.2f s" % (time.time() - t0))
    dev = _lib.device()
    n = nz * ny * nx
    dt = case["dt"]
    T = torch.from_numpy(case["T0"]).to(dev)
    if args.config:
        cfg = {}
        for part in args.config.split(";"):
            k, v = part.split("=")
            cfg[k.strip()] = [int(t) for t in v.split(",")]
        hs.autotune._apply(_lib.load(), P.plan, cfg)
    elif not args.no_tune:
        t0 = time.time()
        cfg = hs.autotune.tune(P, T, use_cache=False, verbose=args.verbose)
        log("autotune %.2f s: %s" % (time.time() - t0, cfg))
    tuned = _lib.plan_info(P.plan)
    # warmup: the flash step plus W-1 more
    T, _ = hs.run_adi_steps_many(P, 0.0, dt, max(args.warmup, 1), T, pa["vmap"], pa["volumetric"])
    K = args.steps
    hist = torch.empty((K, ny, nx), dtype=torch.float64, device=dev)
    L = _lib.load()
    import ctypes
    stream = _lib.stream_handle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check("hs2_run", L.hs2_run(P.plan, _lib.ptr(T), K, None, None, 1,
                                    ctypes.c_void_p(hist.data_ptr()), stream))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = n * K / wall
    log("timed %d steps: %.3f s, %.3f ms/step, %.3e cell-steps/s" % (K, wall, 1e3 * wall / K, value))
    assert torch.isfinite(T).all().item() and torch.isfinite(hist).all().item()

    # per-kernel device time from HIP events on the launch stream around every
    # substep kernel (hs2_run_timed): fence-free events give `kernel_ms`;
    # events with the default system-scope fence are reported beside them
    by_method = {}
    for method, name in ((1, "event_pairs_no_fence"), (0, "event_pairs_default"),
                         (2, "differential")):
        arr = (ctypes.c_double * 4)()
        _lib.check("hs2_run_timed", L.hs2_run_timed(P.plan, _lib.ptr(T), args.timing_steps,
                                                    method, arr, stream))
        by_method[name] = [round(v, 5) for v in arr]
    torch.cuda.synchronize()
    log("kernel ms (x, y, z, step) by timing method: %s" % by_method)
    # each fence-free event pair still adds its own dispatch gap to the kernel
    # it brackets: the three pairs' sum exceeds the step time of a run with no
    # events between its kernels (the differential method's plain run, one
    # event pair around all its steps) by that overhead, taken off in equal
    # parts
    ev = by_method["event_pairs_no_fence"]
    plain_step = by_method["differential"][3]
    pair_overhead = max(0.0, (ev[0] + ev[1] + ev[2] - plain_step) / 3.0)
    kms = [v - pair_overhead for v in ev[:3]]
    names = ("sweep_x", "sweep_y", "sweep_z")
    bpc = list(BYTES_PER_CELL)
    if tuned["y_from_x"]:
        bpc[1] = 16.0  # y reads R_y = M T* - (Ly/2) T^n and writes T**
    gbs = [bpc[i] * n / (kms[i] * 1e-3) / 1e9 for i in range(3)]
    dom = int(np.argmax(kms))
    for i in range(3):
        log("  %s: %.4f ms/launch, %.0f GB/s algorithmic (%.1f%% of 8 TB/s)"
            % (names[i], kms[i], gbs[i], 100 * gbs[i] / HBM_PEAK_GBS))
    step_gbs = CANONICAL_STEP_BYTES * value / 1e9
    # a plain device copy of one field on the same GPU, as a practical
    # ceiling next to the 8 TB/s peak (read + write bytes per copy)
    copy_gbs = device_copy_gbs(T)
    log("device copy: %.0f GB/s; the step reaches %.1f%% of it on 64 B per cell-step"
        % (copy_gbs, 100 * step_gbs / copy_gbs))
    # measured HBM bytes per launch: committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # summary of this workload (copy-calibrated), when it is for this size
    from heatsim2_amd.rocprof_summary import read_traffic
    tr = read_traffic(os.path.join(ROOT, "profiles", TRAFFIC_SUMMARY[args.workload]))
    traffic = None
    if tr and tr.get("cells") == n:
        traffic = tr["traffic_bytes"].get(names[dom])
    roofline = dict(bound="hbm", kernel=names[dom], achieved=round(gbs[dom], 1),
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs[dom] / HBM_PEAK_GBS, 4),
                    traffic=traffic,
                    traffic_source=TRAFFIC_SUMMARY[args.workload] if traffic else None,
                    bytes_per_launch=bpc[dom] * n,
                    kernel_ms=[round(x, 5) for x in kms], kernel_ms_by_method=by_method,
                    event_pair_overhead_ms=round(pair_overhead, 5),
                    timing_steps=args.timing_steps,
                    step_achieved=round(step_gbs, 1), step_frac=round(step_gbs / HBM_PEAK_GBS, 4),
                    step_bytes_per_cell=sum(bpc),
                    step_frac_own_bytes=round(sum(bpc) * value / 1e9 / HBM_PEAK_GBS, 4),
                    copy_GBps=round(copy_gbs, 1), step_frac_of_copy=round(step_gbs / copy_gbs, 4))

    out = {"metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / K, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": wl,
                      "cells": n, "parallelism": "single GPU",
                      "launch": {k: tuned[k] for k in ("x_cpl", "x_lanes", "x_lines",
                                                       "x_lines_y", "y_seg", "y_lines", "z_seg",
                                                       "z_lines", "mode", "y_from_x",
                                                       "x_combined")}},
           "roofline": roofline}
    del T, hist
    P.release()
    torch.cuda.empty_cache()
    out["strongly_coupled"] = None
    if not args.no_strong:
        log("strongly coupled step ...")
        out["strongly_coupled"] = strongly_coupled(hs, (nz, ny, nx), args.strong_steps, log)
        out["strongly_coupled"]["slowdown_vs_value"] = round(value / out["strongly_coupled"]["value"], 4)
        torch.cuda.empty_cache()
    if not args.no_batch:
        log("single-instance gradient ...")
        out["single_gradient"] = single_gradient(hs, log)
        torch.cuda.empty_cache()
        log("batched gradients ...")
        out["batched"] = batched_gradients(hs, args.batch_instances, args.batch_seed,
                                           args.ckpt_every, args.plain_sample, log)
        torch.cuda.empty_cache()
    else:
        out["batched"] = out["single_gradient"] = None
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps, args.cpu_laminate_steps)
        log("cpu baseline %.3e cell-steps/s" % out["cpu_baseline"]["value"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

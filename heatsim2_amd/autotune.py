"""Launch-shape autotuner for the three sweep kernels of a plan.

For the x sweep it chooses (cells per lane, lanes per line, lines per block,
rows of the block's line tile); for the y and z sweeps (cells per thread
segment, lines per block) -- i.e. how much of each line is eliminated
Thomas-style in registers and how much goes to the parallel cyclic
reduction across lanes / threads.  Candidates are timed on
the device with HIP events (hs2_run_timed) on a scratch copy of the field;
the fastest of each sweep is set on the plan.  Results are cached per
(device name, grid shape, coefficient mode) in memory and, if writable, in
$HS2_TUNE_CACHE (default ~/.cache/heatsim2_amd/tune.json).

There is no upstream counterpart: the upstream solves every line with one
serial Thomas pass (tridiag.pyx:46-69).
"""
import ctypes
import json
import os

import torch

from heatsim2_amd import _lib

# (cells per lane, lanes per line, lines per block of the two-phase kernel or 0)
X_COMBOS = ((1, 16, 0), (2, 16, 0), (4, 16, 0), (8, 16, 0), (16, 16, 0), (2, 32, 0), (4, 32, 0),
            (8, 32, 0), (16, 32, 0), (2, 64, 0), (4, 64, 0), (8, 64, 0), (16, 64, 0),
            (1, 16, 8), (2, 16, 8), (4, 16, 8), (8, 16, 8), (16, 16, 8), (4, 16, 16),
            (8, 16, 16), (16, 16, 4), (16, 16, 16), (8, 32, 4), (16, 32, 4), (16, 32, 8),
            (16, 32, 2), (16, 64, 2), (8, 64, 2), (16, 64, 4))
SEGS = (2, 4, 8, 16, 32)
LINES = (8, 16, 32, 64)
MAX_BLOCK = 512

_memo = {}


def x_tiles(lines):
    """Row counts of a two-phase block's line tile: 0 (lines in order) and
    the proper divisors of `lines` up to 8."""
    if lines == 0:
        return (0,)
    return (0,) + tuple(t for t in (1, 2, 4, 8) if t < lines and lines % t == 0)


def x_candidates(nx):
    """x-sweep shapes worth timing: (cells per lane, lanes, lines, tile rows)."""
    out = [c for c in X_COMBOS if c[0] * c[1] >= nx and c[0] * c[1] < 2 * nx + 16]
    out = out or [c for c in X_COMBOS if c[0] * c[1] >= nx][:1]
    return [c + (t,) for c in out for t in x_tiles(c[2])]


def strided_candidates(L):
    out = []
    for seg in SEGS:
        P = -(-L // seg)
        if seg > 2 * L and seg > 2:
            continue
        for cw in LINES:
            if P * cw <= MAX_BLOCK and P * cw >= 64:
                out.append((seg, cw))
    return out


def _cache_path():
    return os.environ.get("HS2_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "heatsim2_amd",
                                       "tune.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, cfg):
    path = _cache_path()
    try:
        data = _load_cache()
        data[key] = cfg
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    except OSError:
        pass


def _time(L, plan, T, steps):
    """Per-launch x, y, z times from fence-free event pairs around every
    kernel (hs2_run_timed method 1)."""
    ms = (ctypes.c_double * 4)()
    _lib.check("hs2_run_timed", L.hs2_run_timed(plan, _lib.ptr(T), steps, 1, ms,
                                                _lib.stream_handle()))
    return list(ms)[:3]


def tune(ADI_params, T=None, steps=4, use_cache=True, verbose=False):
    """Pick and set the fastest launch shapes for this plan; returns
    {"x": (cpl, lanes, lines, tile rows), "y": (seg, lines), "z": (seg, lines),
    "ms": [x, y, z]}."""
    L = _lib.load()
    plan = ADI_params.plan
    nz, ny, nx = ADI_params.shape
    info = _lib.plan_info(plan)
    key = "%s|%s|%dx%dx%d|mode%d" % (L.hs2_version().decode(), torch.cuda.get_device_name(),
                                     nz, ny, nx, info["mode"])
    if use_cache:
        cfg = _memo.get(key) or _load_cache().get(key)
        if cfg and _apply(L, plan, cfg, strict=False):
            _memo[key] = cfg
            return cfg
    dev = _lib.device()
    if T is None:
        T = torch.full(ADI_params.shape, 293.0, dtype=torch.float64, device=dev)
    scratch = T.detach().clone()
    cur = {"x": (info["x_cpl"], info["x_lanes"], info["x_lines"], info["x_lines_y"]),
           "y": (info["y_seg"], info["y_lines"]),
           "z": (info["z_seg"], info["z_lines"])}
    best_ms = [None, None, None]
    for axis_name, idx, cands in (("x", 0, x_candidates(nx)), ("y", 1, strided_candidates(ny)),
                                  ("z", 2, strided_candidates(nz))):
        results = []
        for cand in cands:
            trial = dict(cur)
            trial[axis_name] = cand
            if not _apply(L, plan, trial, strict=False):
                continue
            _time(L, plan, scratch, 1)  # warm the code object
            t = _time(L, plan, scratch, steps)
            # the x shape also decides whether y reads x's right-hand side
            # (hs2_plan_info "y_from_x"), so x is scored with the y sweep
            ms = t[0] + t[1] if axis_name == "x" else t[idx]
            results.append((ms, cand))
            if verbose:
                print("autotune %s %s: %.4f ms%s" % (axis_name, cand, ms,
                                                     " (x + y)" if axis_name == "x" else ""),
                      flush=True)
        if results:
            ms, cand = min(results)
            cur[axis_name] = cand
        _apply(L, plan, cur)
        best_ms[idx] = _time(L, plan, scratch, steps)[idx] if results else None
    cfg = {"x": list(cur["x"]), "y": list(cur["y"]), "z": list(cur["z"]), "ms": best_ms}
    _memo[key] = cfg
    if use_cache:
        _store_cache(key, cfg)
    torch.cuda.synchronize()
    return cfg


def _apply(L, plan, cfg, strict=True):
    x = list(cfg["x"]) + [0] * (4 - len(cfg["x"]))
    rc = L.hs2_plan_set_config(plan, int(x[0]), int(x[1]), int(x[2]), int(x[3]),
                               int(cfg["y"][0]), int(cfg["y"][1]), int(cfg["z"][0]),
                               int(cfg["z"][1]))
    if strict:
        _lib.check("hs2_plan_set_config", rc)
    return rc == 0

"""Shared helpers for the tests, smoke() and bench.py: load the golden
fixtures as plain-data cases and turn a case into heatsim2_amd arguments."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
import cases  # noqa: E402,F401  (tests/golden/cases.py)


def load_golden(name):
    """Case dict (as in tests/golden/cases.py) plus the fixture arrays."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["case_json"]))
    case = dict(meta)
    case["shape"] = tuple(meta["shape"])
    case["d"] = tuple(meta["d"])
    case["materials"] = [tuple(m) for m in meta["materials"]]
    case["boundaries"] = [tuple(b) for b in meta["boundaries"]]
    vol = []
    for e in meta["volumetric"]:
        e = list(e)
        if e[0] == cases.SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE:
            e[2] = tuple(e[2])
        vol.append(tuple(e))
    case["volumetric"] = vol
    case["maps"] = (z["material_elements"], z["boundary_z_elements"], z["boundary_y_elements"],
                    z["boundary_x_elements"], z["volumetric_elements"])
    case["T0"] = z["T0"]
    if "curv_y" in z.files:
        case["curvature"] = (z["curv_y"], z["curv_x"])
    case["fixture"] = z
    return case


def zgrid(case):
    nz, ny, nx = case["shape"]
    z = (np.arange(nz) + 0.5) * case["d"][0]
    return np.broadcast_to(z[:, None, None], (nz, ny, nx)).copy()


def to_package_args(hs, case):
    """setup() argument tuple and run-time arguments for package `hs`
    (heatsim2_amd, or any module with the upstream surface)."""
    bmods = {cases.BND_CONDUCTING: hs.boundary_conducting,
             cases.BND_ANISO: hs.boundary_conducting_anisotropic,
             cases.BND_INSULATING: hs.boundary_insulating,
             cases.BND_THIN: hs.boundary_thininsulatinglayer}
    materials = []
    for (fixed, kz, ky, kx, rho, c, tensor) in case["materials"]:
        if fixed:
            materials.append((hs.TEMPERATURE_FIXED,))
        elif tensor:
            materials.append((hs.TEMPERATURE_COMPUTE, np.diag(np.array((kz, ky, kx), "d")), rho, c))
        else:
            materials.append((hs.TEMPERATURE_COMPUTE, kz, rho, c))
    boundaries = []
    for (kind, h) in case["boundaries"]:
        boundaries.append((bmods[kind], h) if kind == cases.BND_THIN else (bmods[kind],))
    volumetric = []
    for e in case["volumetric"]:
        if e[0] == cases.SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE:
            e = list(e)
            e[2] = np.array(e[2], "d")
            e[4] = zgrid(case)
            e = tuple(e)
        volumetric.append(tuple(e))
    nz, ny, nx = case["shape"]
    dz, dy, dx = case["d"]
    m, bz, by, bx, v = case["maps"]
    setup_args = (0.0, 0.0, 0.0, dz, dy, dx, nz, ny, nx, case["dt"], tuple(materials),
                  tuple(boundaries), tuple(volumetric), m, bz, by, bx, v)
    if case.get("curvature") is not None:
        setup_args += (np.array(case["curvature"][0], "d"), np.array(case["curvature"][1], "d"))
    return dict(setup=setup_args, vmap=v, volumetric=tuple(volumetric))


def oracle_coef(M, w, gz, gy, gx):
    """The oracle's per-cell rows (n x 8: M, D, gzm, gzp, gym, gyp, gxm, gxp)
    of per-voxel coefficients (M, w and each cell's minus-face conductance
    per axis, as heatsim2_amd.adjoint.per_voxel_coefficients gives them): a
    COMPUTE cell's plus face is its neighbour's minus face, a FIXED row (w=0)
    has no couplings."""
    M, w, gz, gy, gx = (np.asarray(a, dtype=np.float64) for a in (M, w, gz, gy, gx))
    coef = np.zeros(M.shape + (8,))
    coef[..., 0] = M
    coef[..., 1] = w
    for a, g in enumerate((gz, gy, gx)):
        coef[..., 2 + 2 * a] = w * g
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, M.shape[a] - 1)
        hi[a] = slice(1, None)
        coef[tuple(lo) + (3 + 2 * a,)] = w[tuple(lo)] * g[tuple(hi)]
    return np.ascontiguousarray(coef.reshape(-1, 8))

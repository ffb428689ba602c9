"""Write the golden fixtures in tests/golden/ from the upstream solver.

Runs only in the container that has the upstream sources mounted, after
``make -C oracle ref``.  It imports the upstream package (see _refload.py) and
records, for the cases in cases.py:

  * the assembled per-substep operators of heatsim2.setup (Amat as n x 3,
    Bmat and Cmats as CSR, Dvec) -- the numeric output of
    crank_nicolson.pyx:128-498 and alternatingdirection_c.c:102-199, which
    only the upstream code fixes (neighbour-k fallback, thin-layer and FIXED
    conventions);
  * end-to-end heatsim2.run_adi_steps trajectories
    (alternatingdirection_c_pyx.pyx:287-416), including every volumetric
    source kind and two of the upstream's own demos at their 10 ms step
    (steel on foam, the tutorial), with the wall time of the upstream's run
    and of the oracle's reference-structured step over the same steps;
  * central finite differences of a surface-history misfit computed with the
    upstream solver on a tiny grid where every voxel has its own material
    class, for the gradient checks.

Usage:  python tests/golden/make_golden.py [case ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import _refload  # noqa: E402
import cases  # noqa: E402


def ref_tuples(hs, case):
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
    for (kind, param) in case["boundaries"]:
        boundaries.append((bmods[kind], param) if kind == cases.BND_THIN else (bmods[kind],))
    volumetric = []
    for entry in case["volumetric"]:
        if entry[0] == cases.SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE:
            e = list(entry)
            e[2] = np.array(e[2], "d")
            e[4] = cases.zgrid_for(case)
            volumetric.append(tuple(e))
        else:
            volumetric.append(tuple(entry))
    return materials, boundaries, volumetric


def ref_setup(hs, case, materials=None):
    mats, bnds, vol = ref_tuples(hs, case)
    if materials is not None:
        mats = materials
    nz, ny, nx = case["shape"]
    dz, dy, dx = case["d"]
    m, bz, by, bx, v = case["maps"]
    curv = case.get("curvature")
    kw = {} if curv is None else dict(top_surface_y_curvatures=np.array(curv[0], "d"),
                                      top_surface_x_curvatures=np.array(curv[1], "d"))
    with _refload.quiet():
        P, S = hs.setup(0.0, 0.0, 0.0, dz, dy, dx, nz, ny, nx, case["dt"], mats, bnds, vol,
                        m, bz, by, bx, v, **kw)
    return P, S, vol


def case_json(case):
    vol = []
    for e in case["volumetric"]:
        vol.append([list(x) if isinstance(x, tuple) else x for x in e])
    keep = dict(name=case["name"], shape=list(case["shape"]), d=list(case["d"]),
                dt=case["dt"], t0=case["t0"], nsteps=case["nsteps"],
                materials=[list(x) for x in case["materials"]],
                boundaries=[list(x) for x in case["boundaries"]], volumetric=vol)
    for k in ("flash_J_m2", "npatch"):
        if k in case:
            keep[k] = case[k]
    return json.dumps(keep)


def maps_dict(case):
    m, bz, by, bx, v = case["maps"]
    out = dict(material_elements=m, boundary_z_elements=bz, boundary_y_elements=by,
               boundary_x_elements=bx, volumetric_elements=v)
    if case.get("curvature") is not None:
        out["curv_y"], out["curv_x"] = case["curvature"]
    return out


def gen_matrix_case(hs, case):
    P, S, vol = ref_setup(hs, case)
    out = dict(case_json=np.array(case_json(case)), T0=case["T0"], **maps_dict(case))
    for s, st in enumerate(S):
        out["A%d" % s] = np.array(st.Amat)
        B = st.Bmat.tocsr()
        B.sum_duplicates()
        out["B%d_data" % s], out["B%d_indices" % s], out["B%d_indptr" % s] = \
            B.data, B.indices.astype(np.int64), B.indptr.astype(np.int64)
        for c, Cm in enumerate(st.Cmats):
            Cm = Cm.tocsr()
            Cm.sum_duplicates()
            out["C%d_%d_data" % (s, c)], out["C%d_%d_indices" % (s, c)], \
                out["C%d_%d_indptr" % (s, c)] = Cm.data, Cm.indices.astype(np.int64), \
                Cm.indptr.astype(np.int64)
        out["D%d" % s] = np.array(st.Dvec)
    T = case["T0"].copy()
    traj = []
    for n in range(case["nsteps"]):
        t = case["t0"] + n * case["dt"]
        m, bz, by, bx, v = case["maps"]
        T = np.array(hs.run_adi_steps(P, S, t, case["dt"], T, v, vol))
        traj.append(T.copy())
    out["traj"] = np.array(traj)
    return out


def gen_trajectory_case(hs, case, zprofile_at=None, top_stride=4):
    t_setup = time.time()
    P, S, vol = ref_setup(hs, case)
    t_setup = time.time() - t_setup
    m, bz, by, bx, v = case["maps"]
    T = case["T0"].copy()
    nz, ny, nx = case["shape"]
    top, zprof, surf_mean = [], [], []
    t_run = time.time()
    for n in range(case["nsteps"]):
        t = case["t0"] + n * case["dt"]
        T = np.array(hs.run_adi_steps(P, S, t, case["dt"], T, v, vol))
        top.append(T[0, ::top_stride, ::top_stride].copy())
        if zprofile_at is not None:
            zprof.append(T[:, zprofile_at[0], zprofile_at[1]].copy())
        surf_mean.append(T[0].mean())
    t_run = time.time() - t_run
    out = dict(case_json=np.array(case_json(case)), T0=case["T0"], final=T,
               top_history=np.array(top), top_stride=np.array(top_stride),
               surf_mean=np.array(surf_mean), ref_setup_s=np.array(t_setup),
               ref_run_s=np.array(t_run), port_run_s=np.array(port_run_seconds(case)),
               **maps_dict(case))
    if zprofile_at is not None:
        out["zprofile"] = np.array(zprof)
        out["zprofile_at"] = np.array(zprofile_at)
    return out


def port_run_seconds(case):
    """Wall time of the oracle's reference-structured step (one thread) over
    the same steps, sources evaluated per step as the upstream does, right
    after the upstream's own run on the same host: the port's speed relative
    to the upstream's."""
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import oracle
    plan = oracle.RefPlan(case["shape"], oracle.assemble_case(case))
    vol = oracle.case_volume(case)
    zg = oracle.zgrid(case["shape"], case["d"][0])
    T = np.array(case["T0"], dtype=np.float64)
    t0 = time.time()
    for n in range(case["nsteps"]):
        t = case["t0"] + n * case["dt"]
        S = oracle.source_field(case["volumetric"], case["maps"][4], t, case["dt"], vol, zg)
        T = plan.step(T, S)
    return time.time() - t0


# ---------------------------------------------------------------- gradients
def case_grad_tiny(seed=11):
    """Tiny laminate where every voxel is its own material class, so a
    per-voxel parameter can be perturbed through the upstream material table
    (the upstream maps are uint8, so at most 256 classes)."""
    nz, ny, nx = 4, 5, 6
    rng = np.random.default_rng(seed)
    d = (4e-4, 5e-4, 5e-4)
    dt = 2e-3
    m = np.arange(nz * ny * nx, dtype=np.uint8).reshape(nz, ny, nx)
    bz = np.zeros((nz + 1, ny, nx), np.uint8)
    by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    bx[:, :, 0] = bx[:, :, -1] = 1
    by[:, 0, :] = by[:, -1, :] = 1
    bz[0] = bz[-1] = 1
    bz[2, 1:4, 2:5] = 2
    v = np.zeros((nz, ny, nx), np.uint8)
    v[0] = 1
    kz = 0.6 * (1 + 0.3 * rng.standard_normal((nz, ny, nx)))
    ky = 5.0 * (1 + 0.2 * rng.standard_normal((nz, ny, nx)))
    kx = 5.0 * (1 + 0.2 * rng.standard_normal((nz, ny, nx)))
    rhoc = 1550.0 * 1200.0 * (1 + 0.1 * rng.standard_normal((nz, ny, nx)))
    E = 1e4
    nsteps = 30
    obs_noise = rng.standard_normal((nsteps, ny, nx))
    return dict(name="grad_tiny", shape=(nz, ny, nx), d=d, dt=dt, t0=0.0, nsteps=nsteps,
                maps=(m, bz, by, bx, v), kz=kz, ky=ky, kx=kx, rhoc=rhoc, E=E,
                h=300.0, obs_noise=obs_noise)


def grad_loss_ref(hs, g, kz, ky, kx, rhoc, E, obs, T0=0.0):
    nz, ny, nx = g["shape"]
    dz, dy, dx = g["d"]
    m, bz, by, bx, v = g["maps"]
    mats = []
    for idx in range(nz * ny * nx):
        k, j, i = np.unravel_index(idx, (nz, ny, nx))
        mats.append((hs.TEMPERATURE_COMPUTE,
                     np.diag(np.array((kz[k, j, i], ky[k, j, i], kx[k, j, i]), "d")),
                     rhoc[k, j, i], 1.0))
    bnds = ((hs.boundary_conducting_anisotropic,), (hs.boundary_insulating,),
            (hs.boundary_thininsulatinglayer, g["h"]))
    vol = ((hs.NO_SOURCE,), (hs.IMPULSE_SOURCE, 0.0, E / dz))
    with _refload.quiet():
        P, S = hs.setup(0.0, 0.0, 0.0, dz, dy, dx, nz, ny, nx, g["dt"], mats, bnds, vol,
                        m, bz, by, bx, v)
    T = np.full((nz, ny, nx), T0)
    from_st = __import__("heatsim2.surface_temperature", fromlist=["x"])
    hist = []
    for n in range(g["nsteps"]):
        T = np.array(hs.run_adi_steps(P, S, n * g["dt"], g["dt"], T, v, vol))
        hist.append(from_st.insulating_z_min_surface_temperature(T, dz))
    hist = np.array(hist)
    if obs is None:
        return hist
    return float(np.sum((hist - obs) ** 2)), hist


def gen_grad_case(hs, nvox=10, rel=3e-3):
    """Finite differences of the upstream solver's loss.  Temperatures are
    relative to ambient (T0 = 0): the step is affine in T, so this changes
    no gradient but keeps round-off in T at the scale of the rise rather
    than of 293 K.  Each derivative is a Richardson-extrapolated central
    difference, (4 D(h) - D(2h)) / 3 with h = rel * |value|."""
    g = case_grad_tiny()
    base = dict(kz=g["kz"], ky=g["ky"], kx=g["kx"], rhoc=g["rhoc"])
    # observations: the forward history at a perturbed "true" field plus noise
    true = {k: v * (1 + 0.1 * np.sin(np.arange(v.size).reshape(v.shape))) for k, v in base.items()}
    hist_true = grad_loss_ref(hs, g, true["kz"], true["ky"], true["kx"], true["rhoc"], g["E"] * 1.05, None)
    obs = hist_true + 0.01 * g["obs_noise"]
    L0, hist0 = grad_loss_ref(hs, g, base["kz"], base["ky"], base["kx"], base["rhoc"], g["E"], obs)
    rng = np.random.default_rng(12)
    n = g["kz"].size
    out = dict(shape=np.array(g["shape"]), d=np.array(g["d"]), dt=np.array(g["dt"]),
               nsteps=np.array(g["nsteps"]), h=np.array(g["h"]), E=np.array(g["E"]),
               T0=np.array(0.0),
               material_elements=g["maps"][0], boundary_z_elements=g["maps"][1],
               boundary_y_elements=g["maps"][2], boundary_x_elements=g["maps"][3],
               volumetric_elements=g["maps"][4], obs=obs, loss=np.array(L0), hist=hist0,
               fd_rel_step=np.array(rel), **base)

    def richardson(f, x0):
        h = rel * abs(x0)
        d1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
        d2 = (f(x0 + 2 * h) - f(x0 - 2 * h)) / (4 * h)
        return (4 * d1 - d2) / 3

    for name in ("kz", "ky", "kx", "rhoc"):
        vox = rng.choice(n, size=nvox, replace=False)
        fd = []
        for idx in vox:
            def f(val):
                p = {k: v.copy() for k, v in base.items()}
                p[name].flat[idx] = val
                return grad_loss_ref(hs, g, p["kz"], p["ky"], p["kx"], p["rhoc"], g["E"], obs)[0]
            fd.append(richardson(f, base[name].flat[idx]))
        out["fd_%s_vox" % name] = vox.astype(np.int64)
        out["fd_%s" % name] = np.array(fd)
    out["fd_E"] = np.array(richardson(
        lambda e: grad_loss_ref(hs, g, base["kz"], base["ky"], base["kx"], base["rhoc"], e, obs)[0],
        g["E"]))
    return out


def main(argv):
    hs = _refload.load_reference()
    todo = argv or ["tiny_mixed", "tiny_aniso", "tiny_curved", "steel_cube40",
                    "two_layer_mini", "laminate_mini", "steelonfoam", "tutorial", "grad_tiny"]
    for name in todo:
        t = time.time()
        if name == "tiny_mixed":
            out = gen_matrix_case(hs, cases.case_tiny_mixed())
        elif name == "tiny_aniso":
            out = gen_matrix_case(hs, cases.case_tiny_aniso())
        elif name == "tiny_curved":
            out = gen_matrix_case(hs, cases.case_tiny_curved())
        elif name == "steel_cube40":
            out = gen_trajectory_case(hs, cases.case_steel_cube(40, 200), zprofile_at=(20, 20))
        elif name == "two_layer_mini":
            c = cases.case_two_layer(nz=20, ny=12, nx=10, nsteps=120, steel_layers=6)
            out = gen_trajectory_case(hs, c, zprofile_at=(5, 4), top_stride=1)
            out["traj_final_all"] = out["final"]
        elif name == "laminate_mini":
            c = cases.case_laminate(nz=12, ny=24, nx=20, nsteps=80, patch_min=3, patch_max=10)
            out = gen_trajectory_case(hs, c, zprofile_at=(10, 9), top_stride=1)
        elif name == "steelonfoam":
            out = gen_trajectory_case(hs, cases.case_steelonfoam(), zprofile_at=(2, 2))
        elif name == "tutorial":
            out = gen_trajectory_case(hs, cases.case_tutorial(), zprofile_at=(20, 21),
                                      top_stride=2)
        elif name == "grad_tiny":
            out = gen_grad_case(hs)
        else:
            raise SystemExit("unknown case %s" % name)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%-16s %7.1f s  -> %s (%d KiB)" % (name, time.time() - t, path,
                                                os.path.getsize(path) // 1024), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

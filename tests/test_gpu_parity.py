"""HIP path against the upstream fixtures and the oracle (GPU)."""
import numpy as np
import pytest
import torch

import oracle
from hs2_cases import cases, load_golden, to_package_args

pytestmark = pytest.mark.gpu

TOL = 1e-10  # max |T_gpu - T_ref| / peak temperature rise


@pytest.fixture(scope="module")
def hs():
    import heatsim2_amd
    return heatsim2_amd


def _rel(a, b, rise):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / rise


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso"])
@pytest.mark.parametrize("on_device", [False, True])
def test_small_cases_match_upstream(hs, name, on_device):
    case = load_golden(name)
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = case["fixture"]["traj"]
    rise = np.abs(ref - case["T0"]).max()
    T = torch.from_numpy(case["T0"]).cuda() if on_device else case["T0"].copy()
    T_in0 = T.clone() if on_device else T.copy()
    for n in range(case["nsteps"]):
        t = case["t0"] + n * case["dt"]
        T_new = hs.run_adi_steps(P, S, t, case["dt"], T, args["vmap"], args["volumetric"])
        if n == 0:  # the input is never modified
            assert (torch.equal(T, T_in0) if on_device else np.array_equal(T, T_in0))
        T = T_new
        got = T.cpu().numpy() if on_device else T
        assert _rel(got, ref[n], rise) < TOL, n
    assert isinstance(T, torch.Tensor) == on_device


@pytest.mark.parametrize("name", ["steel_cube40", "two_layer_mini", "laminate_mini"])
def test_trajectories_match_upstream(hs, name):
    case = load_golden(name)
    z = case["fixture"]
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    st = int(z["top_stride"])
    T, hist = hs.run_adi_steps_many(P, case["t0"], case["dt"], case["nsteps"],
                                    torch.from_numpy(case["T0"]).cuda(), args["vmap"],
                                    args["volumetric"], surface_history=True)
    rise = np.abs(z["top_history"] - case["T0"][0, ::st, ::st]).max()
    assert _rel(T.cpu().numpy(), z["final"], rise) < TOL
    # readout history equals the readout of the upstream top layers
    if "zprofile" in z.files:
        pass
    T_host = T.cpu().numpy()
    assert np.allclose(hist[-1].cpu().numpy(), oracle.surface_readout(T_host, case["d"][0]),
                       rtol=0, atol=1e-12 * rise)


def test_steel_cube_history_matches_upstream(hs):
    """The 40^3 steel flash case, step by step: the top-layer history and the
    z profile at the centre over all 200 steps."""
    case = load_golden("steel_cube40")
    z = case["fixture"]
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    T = torch.from_numpy(case["T0"]).cuda()
    st = int(z["top_stride"])
    rise = np.abs(z["top_history"] - 293.0).max()
    worst = 0.0
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"], args["volumetric"])
        Th = T.cpu().numpy()
        worst = max(worst, _rel(Th[0, ::st, ::st], z["top_history"][n], rise),
                    _rel(Th[:, 20, 20], z["zprofile"][n], rise))
    assert worst < TOL
    # first-frame energy check: the flash lands in the top layer (E/(rho c dz))
    assert abs(rise - 1e4 / (7850 * 470 * 5e-4)) / rise < 0.5


def test_expand_matches_oracle_coefficients(hs):
    """The device-built per-cell coefficients (classes from the maps) equal the
    oracle's per-cell restatement of crank_nicolson.pyx."""
    from heatsim2_amd import _lib
    for name in ("tiny_mixed", "tiny_aniso", "laminate_mini"):
        case = load_golden(name)
        args = to_package_args(hs, case)
        P, S = hs.setup(*args["setup"])
        assert _lib.plan_info(P.plan)["mode"] == 1
        nz, ny, nx = case["shape"]
        f = [torch.zeros((nz, ny, nx), dtype=torch.float64, device="cuda") for _ in range(5)]
        _lib.check("hs2_plan_expand", _lib.load().hs2_plan_expand(
            P.plan, *[_lib.ptr(t) for t in f], _lib.stream_handle()))
        M, w, gz, gy, gx = [t.cpu().numpy() for t in f]
        ref = oracle.assemble_case(case).reshape(nz, ny, nx, 8)

        def plus(g, axis):
            out = np.zeros_like(g)
            sl = [slice(None)] * 3
            sl2 = [slice(None)] * 3
            sl[axis] = slice(0, -1)
            sl2[axis] = slice(1, None)
            out[tuple(sl)] = g[tuple(sl2)]
            return out
        got = np.stack([M, w, gz * w, plus(gz, 0) * w, gy * w, plus(gy, 1) * w, gx * w,
                        plus(gx, 2) * w], axis=-1)
        assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-15, name


def _random_case(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    m = rng.integers(0, 3, size=shape).astype(np.uint8)
    m[rng.random(shape) < 0.05] = 3  # FIXED
    bz = rng.integers(0, 2, size=(nz + 1, ny, nx)).astype(np.uint8) * 2
    by = rng.integers(0, 2, size=(nz, ny + 1, nx)).astype(np.uint8) * 2
    bx = rng.integers(0, 2, size=(nz, ny, nx + 1)).astype(np.uint8) * 2
    bz[rng.random(bz.shape) < 0.1] = 3
    by[rng.random(by.shape) < 0.1] = 3
    bx[rng.random(bx.shape) < 0.1] = 3
    for a, sl in ((bz, (0,)), (bz, (-1,)), (by, (slice(None), 0)), (by, (slice(None), -1)),
                  (bx, (Ellipsis, 0)), (bx, (Ellipsis, -1))):
        a[sl] = 1
    v = np.zeros(shape, np.uint8)
    v[0] = 1
    v[rng.random(shape) < 0.1] = 2
    d = (3e-4 + 2e-4 * rng.random(), 3e-4 + 2e-4 * rng.random(), 3e-4 + 2e-4 * rng.random())
    materials = [(0, 50.0, 50.0, 50.0, 7850.0, 470.0, 0), (0, 0.4, 0.4, 0.4, 1200.0, 1300.0, 0),
                 (0, 5.0, 5.0, 5.0, 1550.0, 1200.0, 0), (1, 0, 0, 0, 0, 0, 0)]
    boundaries = [(cases.BND_CONDUCTING, 0.0), (cases.BND_INSULATING, 0.0),
                  (cases.BND_CONDUCTING, 0.0), (cases.BND_THIN, 800.0)]
    volumetric = [(0,), (1, 0.0, 1e4 / d[0]), (2, 0.0, 3e-3, 2e7)]
    T0 = 293.0 + rng.random(shape)
    T0[m == 3] = 290.0
    return dict(shape=shape, d=d, dt=1e-3, t0=0.0, nsteps=4, materials=materials,
                boundaries=boundaries, volumetric=volumetric, maps=(m, bz, by, bx, v), T0=T0)


SHAPES = [(3, 5, 7), (2, 1, 1), (1, 4, 9), (5, 3, 64), (4, 6, 65), (3, 70, 130), (70, 3, 5),
          (6, 4, 256), (9, 17, 200), (130, 2, 33), (4, 300, 3), (260, 3, 2), (2, 2, 513),
          (3, 512, 2), (517, 2, 3), (2, 1000, 2), (1024, 1, 1), (2, 3, 1024),
          (2, 1024, 8), (1024, 2, 8)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_random_problems_match_oracle(hs, shape):
    """Random materials (incl. FIXED), conducting / thin / insulating faces and
    sources on grids whose line lengths exercise every kernel variant
    (1..1024 cells on every axis, non-multiples of 64; 1024-cell y and z lines
    whose 16-cell segments and 8-column blocks tile the grid exactly)."""
    case = _random_case(shape, seed=sum(shape))
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    T = torch.from_numpy(case["T0"]).cuda()
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"], args["volumetric"])
        assert _rel(T.cpu().numpy(), ref[n], rise) < TOL, n


@pytest.mark.parametrize("on_device", [False, True])
def test_curved_case_matches_upstream(hs, on_device):
    """Curvature mode against the upstream's own trajectory (tiny_curved:
    seeded curvatures of both signs, thin z/y/x faces, FIXED cells, impulse,
    stepped and volume-scaled point sources)."""
    case = load_golden("tiny_curved")
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    assert _lib_mode(P) == 3
    ref = case["fixture"]["traj"]
    rise = np.abs(ref - case["T0"]).max()
    T = torch.from_numpy(case["T0"]).cuda() if on_device else case["T0"].copy()
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, case["t0"] + n * case["dt"], case["dt"], T, args["vmap"],
                             args["volumetric"])
        got = T.cpu().numpy() if on_device else T
        assert _rel(got, ref[n], rise) < TOL, n


def _lib_mode(P):
    from heatsim2_amd import _lib
    return _lib.plan_info(P.plan)["mode"]


def _curved_case(shape, seed):
    case = _random_case(shape, seed)
    rng = np.random.default_rng(seed + 1)
    nz, ny, nx = shape
    kmax = 0.5 / (nz * case["d"][0])  # edges change by at most half over the depth
    case["curvature"] = (rng.uniform(-kmax, kmax, (ny, nx)), rng.uniform(-kmax, kmax, (ny, nx)))
    v = case["maps"][4]
    v[rng.random(shape) < 0.05] = 3
    case["volumetric"] = list(case["volumetric"]) + [(3, 2 * case["dt"], 0.05)]
    case["nsteps"] = 5
    return case


@pytest.mark.parametrize("shape", [(5, 16, 64), (7, 12, 256), (9, 20, 33), (40, 3, 5),
                                   (3, 130, 17)],
                         ids=lambda s: "x".join(map(str, s)))
def test_curved_random_problems_match_oracle(hs, shape):
    """Curvature mode on random problems whose lines take the full-line and
    the masked x kernels and every strided segment length, against the
    oracle's curvature assembly (pinned to the upstream) and its
    reference-structured step."""
    case = _curved_case(shape, seed=31 + sum(shape))
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    T = torch.from_numpy(case["T0"]).cuda()
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"], args["volumetric"])
        assert _rel(T.cpu().numpy(), ref[n], rise) < TOL, n
    T2, _ = hs.run_adi_steps_many(P, 0.0, case["dt"], case["nsteps"],
                                  torch.from_numpy(case["T0"]).cuda(), args["vmap"],
                                  args["volumetric"])
    assert torch.equal(T, T2)


def test_y_from_x_path_matches_oracle(hs):
    """Full-line x shapes on a class-map plan hand the y sweep its right-hand
    side M T* - (Ly/2) T^n (hs2_plan_info "y_from_x"); steps with sources
    and other shapes keep T*.  Both agree with the oracle, step by step."""
    from heatsim2_amd import _lib, autotune
    case = _random_case((8, 16, 128), seed=5)
    case["nsteps"] = 6  # the stepped source ends after step 3
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    info = _lib.plan_info(P.plan)
    assert info["y_from_x"] == 1
    base = {"y": (info["y_seg"], info["y_lines"]), "z": (info["z_seg"], info["z_lines"])}
    for x, expect in (((8, 16, 8, 2), 1), ((8, 16, 8, 0), 1), ((4, 32, 0, 0), 0),
                      ((16, 16, 8, 0), 0)):
        autotune._apply(_lib.load(), P.plan, dict(base, x=x))
        assert _lib.plan_info(P.plan)["y_from_x"] == expect, x
        T = torch.from_numpy(case["T0"]).cuda()
        for n in range(case["nsteps"]):
            T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"],
                                 args["volumetric"])
            assert _rel(T.cpu().numpy(), ref[n], rise) < TOL, (x, n)


@pytest.mark.parametrize("shape,x", [((16, 10, 256), (16, 16, 16)), ((16, 10, 256), (16, 16, 8)),
                                     ((4, 6, 512), (16, 32, 4)), ((8, 6, 512), (16, 32, 8)),
                                     ((4, 6, 1024), (16, 64, 4))],
                         ids=["256", "256-8", "512", "512-8", "1024-4"])
def test_x_plane_pairs_match_oracle(hs, shape, x):
    """x lines of 16-cell lanes in blocks of 4, 8 or 16 lines, tiled as
    LPB/2 planes x 2 rows, run with the lines ordered so that each thread's
    quads are neighbours in y and z and share those rows
    (source-free steps 4-5; steps 0-3 carry sources and run the source
    kernel).  Against the oracle step by step, and bit for bit against
    1-row tiles, which take the same arithmetic per cell."""
    from heatsim2_amd import _lib, autotune
    case = _random_case(shape, seed=23)
    case["nsteps"] = 6
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    out = {}
    for ty in (2, 1):
        autotune._apply(_lib.load(), P.plan, {"x": x + (ty,), "y": (16, 8), "z": (4, 64)})
        assert _lib.plan_info(P.plan)["y_from_x"] == 1
        T = torch.from_numpy(case["T0"]).cuda()
        steps = []
        for n in range(case["nsteps"]):
            T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"],
                                 args["volumetric"])
            steps.append(T.cpu().numpy())
            assert _rel(steps[-1], ref[n], rise) < TOL, (ty, n)
        out[ty] = steps
    for a, b in zip(out[2], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("maker,x", [
    (lambda: cases.case_laminate(16, 10, 256, nsteps=6, patch_min=2, patch_max=8), (16, 16, 16, 2)),
    (lambda: cases.case_laminate(16, 10, 256, nsteps=6, patch_min=2, patch_max=8), (16, 16, 16, 0)),
    (lambda: cases.case_two_layer(nz=4, ny=6, nx=512, nsteps=6, steel_layers=2), (16, 32, 4, 2)),
    (lambda: cases.case_two_layer(nz=8, ny=6, nx=512, nsteps=6, steel_layers=3), (16, 32, 8, 2)),
    (lambda: cases.case_two_layer(nz=6, ny=8, nx=128, nsteps=6, steel_layers=3), (8, 16, 8, 0))],
    ids=["256-pairs", "256-rows", "512-pairs", "512-pairs8", "128"])
def test_x_combined_class_plane_matches_three_planes(hs, maker, x):
    """Source-free steps on full x lines read one combined class byte per
    cell (its z, y, x class triple) instead of three planes: bit for bit the
    same steps as the three planes (HS2_OPT_COMBINED_X off), and the oracle's."""
    from heatsim2_amd import _lib, autotune
    case = maker()
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    L = _lib.load()
    out = {}
    for comb in (1, 0):
        _lib.check("hs2_plan_set_option", L.hs2_plan_set_option(P.plan, _lib.HS2_OPT_COMBINED_X,
                                                                comb))
        autotune._apply(L, P.plan, {"x": x, "y": (8, 16), "z": (4, 32)})
        info = _lib.plan_info(P.plan)
        assert (info["x_combined"] > 0) == bool(comb), info
        T = torch.from_numpy(case["T0"]).cuda()
        steps = []
        for n in range(case["nsteps"]):
            T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"],
                                 args["volumetric"])
            steps.append(T.cpu().numpy())
            assert _rel(steps[-1], ref[n], rise) < TOL, (comb, n)
        out[comb] = steps
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)


def test_x_combined_plane_needs_few_triples(hs):
    """Random maps give more than 256 class triples: no combined plane, the
    x sweep reads the three planes (and still matches the oracle in the
    random-problem tests)."""
    from heatsim2_amd import _lib
    case = _random_case((8, 16, 128), seed=5)
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    info = _lib.plan_info(P.plan)
    assert info["mode"] == 1 and info["x_combined"] == 0


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso", "tiny_curved"])
def test_step_operators_match_upstream(hs, name):
    """ADI_steps[s].Amat/Bmat/Cmats/Dvec of a device plan (per-cell
    coefficients read back from the device) equal the upstream's assembled
    matrices, and Lmat/Umat multiply back to Amat.  1e-14: curvature mode
    keeps its coefficients per unit volume and multiplies the volume back."""
    import scipy.sparse as sp
    case = load_golden(name)
    z = case["fixture"]
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    n = int(np.prod(case["shape"]))
    for s, st in enumerate(S):
        for mine, key in ((st.Amat, "A%d" % s), (st.Dvec, "D%d" % s)):
            assert _rel(mine, z[key], np.abs(z[key]).max()) < 1e-14, key
        assert len(st.Cmats) == s
        for which, mine in enumerate([st.Bmat] + list(st.Cmats)):
            key = "B%d" % s if which == 0 else "C%d_%d" % (s, which - 1)
            ref = sp.csr_matrix((z[key + "_data"], z[key + "_indices"], z[key + "_indptr"]),
                                shape=(n, n)).toarray()
            assert _rel(mine.toarray(), ref, np.abs(ref).max()) < 1e-14, key
        L, U = st.Lmat, st.Umat
        prev = np.concatenate([[0.0], U[:-1, 2]])
        LU = np.stack([L[:, 0], L[:, 0] * prev + L[:, 1], L[:, 1] * U[:, 2]], axis=1)
        assert _rel(LU, st.Amat, np.abs(st.Amat).max()) < 1e-14, s


def test_many_classes_use_per_cell_coefficients(hs):
    """40 materials of distinct k and rho*c, placed at random, give more than
    256 coefficient classes on an axis: setup falls back to per-cell float64
    coefficients (plan mode 2).  Trajectory against the oracle, and the
    step operators against the oracle's reference-structured assembly."""
    import scipy.sparse as sp
    from heatsim2_amd import _lib
    case = _random_case((6, 12, 14), seed=31)
    rng = np.random.default_rng(32)
    mats = [(0,) + (float(k),) * 3 + (float(r), float(c), 0)
            for k, r, c in zip(rng.uniform(0.5, 60.0, 40), rng.uniform(1e3, 8e3, 40),
                               rng.uniform(300.0, 1500.0, 40))]
    case["materials"] = mats + [(1, 0, 0, 0, 0, 0, 0)]
    m = rng.integers(0, 40, size=case["shape"]).astype(np.uint8)
    m[rng.random(case["shape"]) < 0.05] = 40  # FIXED
    case["maps"] = (m,) + tuple(case["maps"][1:])
    case["T0"] = np.where(m == 40, 290.0, case["T0"])
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    assert _lib.plan_info(P.plan)["mode"] == 2
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    T = torch.from_numpy(case["T0"]).cuda()
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"], args["volumetric"])
        assert _rel(T.cpu().numpy(), ref[n], rise) < TOL, n
    R = oracle.RefPlan(case["shape"], oracle.assemble_case(case))
    n = R.n
    for s, st in enumerate(S):
        A, D = R.A(s)
        assert _rel(st.Amat, A, np.abs(A).max()) < 1e-14, s
        assert _rel(st.Dvec, D, 1.0) < 1e-14, s
        for which, mine in enumerate([st.Bmat] + list(st.Cmats)):
            want = sp.csr_matrix(R.csr(s, which), shape=(n, n)).toarray()
            assert _rel(mine.toarray(), want, np.abs(want).max()) < 1e-14, (s, which)


def test_run_many_equals_stepwise(hs):
    case = load_golden("tiny_mixed")
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    T = torch.from_numpy(case["T0"]).cuda()
    hist = []
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"], args["volumetric"])
        hist.append(hs.surface_temperature.insulating_z_min_surface_temperature(T, case["d"][0]))
    T2, h2 = hs.run_adi_steps_many(P, 0.0, case["dt"], case["nsteps"],
                                   torch.from_numpy(case["T0"]).cuda(), args["vmap"],
                                   args["volumetric"], surface_history=True)
    assert torch.equal(T, T2)
    assert torch.equal(torch.stack(hist), h2)


def test_zdecay_source_matches_upstream(hs):
    case = load_golden("tiny_aniso")  # carries a z-decaying impulse at t=0
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = case["fixture"]["traj"]
    T, _ = hs.run_adi_steps_many(P, 0.0, case["dt"], case["nsteps"],
                                 torch.from_numpy(case["T0"]).cuda(), args["vmap"],
                                 args["volumetric"])
    rise = np.abs(ref - case["T0"]).max()
    assert _rel(T.cpu().numpy(), ref[-1], rise) < TOL


def test_deterministic(hs):
    """No atomics in the step: two runs are bitwise equal."""
    case = load_golden("laminate_mini")
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    outs = []
    for _ in range(2):
        T, _h = hs.run_adi_steps_many(P, 0.0, case["dt"], 30, torch.from_numpy(case["T0"]).cuda(),
                                      args["vmap"], args["volumetric"])
        outs.append(T)
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("shape", [(20, 48, 200), (130, 9, 40), (9, 300, 12), (4, 8, 256),
                                   (6, 4, 64), (2, 16, 32), (13, 24, 128), (7, 16, 64),
                                   (5, 8, 512)],
                         ids=lambda s: "x".join(map(str, s)))
def test_every_launch_shape_matches_oracle(hs, shape):
    """Every launch shape the autotuner may pick (x: cells per lane x lanes per
    line x lines per block x rows of the block's line tile; y, z: segment
    length x lines per block) gives the oracle's result, over steps with and
    without source terms.  From (4, 8, 256) on, the grids fill the x lanes
    exactly (the unmasked x path where the tiles also cover the planes and
    rows; ragged tiles elsewhere, e.g. 13 planes)."""
    case = _random_case(shape, seed=7 + sum(shape))
    # the stepped source lasts through t = 3 ms: steps 4 and 5 have none
    case["nsteps"] = 6
    _every_launch_shape(hs, case)


@pytest.mark.parametrize("shape", [(20, 48, 200), (4, 8, 256), (6, 4, 64), (2, 16, 32),
                                   (13, 24, 128)],
                         ids=lambda s: "x".join(map(str, s)))
def test_every_launch_shape_curved(hs, shape):
    """The same over the curvature mode's per-side coefficients (each kernel
    variant takes a cell's plus faces from the cell itself)."""
    case = _curved_case(shape, seed=11 + sum(shape))
    case["nsteps"] = 6
    _every_launch_shape(hs, case)


def _every_launch_shape(hs, case):
    from heatsim2_amd import _lib, autotune
    shape = case["shape"]
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    ref = oracle.run_ref(case)
    rise = max(np.abs(np.array(ref) - case["T0"]).max(), 1e-300)
    info = _lib.plan_info(P.plan)
    base = {"x": (info["x_cpl"], info["x_lanes"], info["x_lines"], info["x_lines_y"]),
            "y": (info["y_seg"], info["y_lines"]),
            "z": (info["z_seg"], info["z_lines"])}
    nz, ny, nx = shape
    trials = [("x", c) for c in autotune.x_candidates(nx)] + \
             [("y", c) for c in autotune.strided_candidates(ny)] + \
             [("z", c) for c in autotune.strided_candidates(nz)]
    for axis, cand in trials:
        cfg = dict(base)
        cfg[axis] = cand
        autotune._apply(_lib.load(), P.plan, cfg)
        T = torch.from_numpy(case["T0"]).cuda()
        for n in range(case["nsteps"]):
            T = hs.run_adi_steps(P, S, n * case["dt"], case["dt"], T, args["vmap"],
                                 args["volumetric"])
            assert _rel(T.cpu().numpy(), ref[n], rise) < TOL, (axis, cand, n)


def test_run_timed_advances_like_run(hs):
    """hs2_run_timed times the same kernels hs2_run launches: methods 0 and 1
    advance the field by nsteps and method 2 by 4 nsteps (its extra launches
    write a scratch field only), each bit for bit as hs2_run; the times are
    positive, and an unknown method is refused."""
    import ctypes
    from heatsim2_amd import _lib
    case = cases.case_two_layer(nz=16, ny=32, nx=64, nsteps=1, steel_layers=4)
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    L = _lib.load()
    T0, _h = hs.run_adi_steps_many(P, 0.0, case["dt"], 1, torch.from_numpy(case["T0"]).cuda(),
                                   args["vmap"], args["volumetric"])
    n = 5
    for method, steps in ((0, n), (1, n), (2, 4 * n)):
        A, B = T0.clone(), T0.clone()
        _lib.check("hs2_run", L.hs2_run(P.plan, _lib.ptr(A), steps, None, None, 0, None,
                                        _lib.stream_handle()))
        ms = (ctypes.c_double * 4)()
        _lib.check("hs2_run_timed", L.hs2_run_timed(P.plan, _lib.ptr(B), n, method, ms,
                                                    _lib.stream_handle()))
        torch.cuda.synchronize()
        assert torch.equal(A, B), method
        assert all(v > 0 for v in ms), (method, list(ms))
    ms = (ctypes.c_double * 4)()
    assert L.hs2_run_timed(P.plan, _lib.ptr(T0), n, 3, ms, _lib.stream_handle()) != 0

"""The oracle pinned: against the upstream's own operators and trajectories
(golden fixtures written by tests/golden/make_golden.py), against dense
linear algebra, and against what the mathematics fixes (energy balance,
constant fields, exact 1-D Crank-Nicolson, closed-form readout)."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

import oracle
from hs2_cases import cases, load_golden

MATRIX_CASES = ("tiny_mixed", "tiny_aniso", "tiny_curved")
TRAJ_CASES = ("steel_cube40", "two_layer_mini", "laminate_mini")


def rel_err(a, b, scale):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / scale


@pytest.mark.parametrize("name", MATRIX_CASES)
def test_operators_match_upstream(name):
    """A, B, C[], D of every substep equal the upstream's assembled
    matrices (alternatingdirection_c.c:102-199 via crank_nicolson.pyx)."""
    case = load_golden(name)
    z = case["fixture"]
    P = oracle.RefPlan(case["shape"], oracle.assemble_case(case))
    n = P.n
    for s in range(3):
        A, D = P.A(s)
        assert rel_err(A, z["A%d" % s], np.abs(z["A%d" % s]).max()) < 1e-15
        assert np.array_equal(D, z["D%d" % s])
        for which in range(s + 1):
            key = "B%d" % s if which == 0 else "C%d_%d" % (s, which - 1)
            mine = sp.csr_matrix(P.csr(s, which), shape=(n, n)).toarray()
            ref = sp.csr_matrix((z[key + "_data"], z[key + "_indices"], z[key + "_indptr"]),
                                shape=(n, n)).toarray()
            assert rel_err(mine, ref, np.abs(ref).max()) < 1e-15, key


@pytest.mark.parametrize("name", MATRIX_CASES)
@pytest.mark.parametrize("structured", [True, False])
def test_trajectory_matches_upstream_small(name, structured):
    case = load_golden(name)
    ref = case["fixture"]["traj"]
    mine = oracle.run_ref(case, structured=structured)
    rise = np.abs(ref - case["T0"]).max()
    assert rel_err(mine, ref, rise) < 1e-12


@pytest.mark.parametrize("name", TRAJ_CASES)
def test_trajectory_matches_upstream(name):
    case = load_golden(name)
    z = case["fixture"]
    coef = oracle.assemble_case(case)
    vol = case["d"][0] * case["d"][1] * case["d"][2]
    zg = oracle.zgrid(case["shape"], case["d"][0])
    T = case["T0"].copy()
    st = int(z["top_stride"])
    worst = 0.0
    rise = np.abs(z["top_history"] - case["T0"][0, ::st, ::st]).max()
    for n in range(case["nsteps"]):
        t = case["t0"] + n * case["dt"]
        S = oracle.source_field(case["volumetric"], case["maps"][4], t, case["dt"], vol, zg)
        T = oracle.fast_step(coef, T, S)
        worst = max(worst, rel_err(T[0, ::st, ::st], z["top_history"][n], rise))
    assert worst < 1e-11
    assert rel_err(T, z["final"], rise) < 1e-11


def test_zero_curvature_is_the_uncurved_assembly_times_the_volume():
    """Curvature mode with both curvatures 0: every edge keeps its nominal
    length, so each cell's equation is the uncurved one multiplied by
    dz*dy*dx (crank_nicolson.pyx:409-457 against :359-378)."""
    case = cases.case_tiny_curved()
    ny, nx = case["shape"][1:]
    case["curvature"] = (np.zeros((ny, nx)), np.zeros((ny, nx)))
    curved, vol = oracle.assemble_curved(case["shape"], case["d"], case["dt"],
                                         case["materials"], case["boundaries"], case["maps"],
                                         *case["curvature"])
    flat = oracle.assemble(case["shape"], case["d"], case["dt"], case["materials"],
                           case["boundaries"], case["maps"])
    V = case["d"][0] * case["d"][1] * case["d"][2]
    assert np.allclose(vol, V, rtol=1e-15, atol=0)
    fixed = flat[:, 1] == 0.0  # FIXED cells keep M = 1 unscaled
    scaled = flat * V
    scaled[fixed] = flat[fixed]
    assert rel_err(curved, scaled, np.abs(scaled).max()) < 1e-15


def test_thomas_matches_dense_and_upstream_diagnostic():
    """tridiag.pyx:71-112's own check, seeded: LU solve of a random
    4x4 tridiagonal system equals inv(A) b; plus a long diagonally dominant
    system against scipy's banded solver."""
    rng = np.random.default_rng(0)
    for n in (4, 257):
        a = rng.random(n); b = rng.random(n) + (0 if n == 4 else 3.0); c = rng.random(n)
        a[0] = 0.0; c[-1] = 0.0
        r = rng.random(n)
        x = oracle.thomas(a, b, c, r)
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.linalg.norm(x - np.linalg.solve(A, r)) / np.linalg.norm(x) < 1e-10
        ab = np.zeros((3, n)); ab[0, 1:] = c[:-1]; ab[1] = b; ab[2, :-1] = a[1:]
        assert np.allclose(x, scipy.linalg.solve_banded((1, 1), ab, r), rtol=1e-12, atol=0)


def _uniform_case(shape=(6, 7, 8), k=(3.0, 5.0, 7.0), tensor=1):
    nz, ny, nx = shape
    m = np.zeros(shape, np.uint8)
    bz = np.zeros((nz + 1, ny, nx), np.uint8); by = np.zeros((nz, ny + 1, nx), np.uint8)
    bx = np.zeros((nz, ny, nx + 1), np.uint8)
    for a, sl in ((bz, (0,)), (bz, (-1,)), (by, (slice(None), 0)), (by, (slice(None), -1)),
                  (bx, (Ellipsis, 0)), (bx, (Ellipsis, -1))):
        a[sl] = 1
    return dict(shape=shape, d=(4e-4, 5e-4, 6e-4), dt=2e-3,
                materials=[(0, k[0], k[1], k[2], 2000.0, 900.0, tensor)],
                boundaries=[(cases.BND_ANISO if tensor else cases.BND_CONDUCTING, 0.0),
                            (cases.BND_INSULATING, 0.0)],
                maps=(m, bz, by, bx, np.zeros(shape, np.uint8)))


@pytest.mark.parametrize("structured", [True, False])
def test_energy_conservation(structured):
    """Insulated box, no FIXED cells: sum rho c V (T1 - T0) = dt V sum S
    (the face terms cancel pairwise)."""
    case = _uniform_case()
    coef = oracle.assemble_case(case)
    rng = np.random.default_rng(1)
    T0 = 293 + rng.random(case["shape"])
    S = rng.random(case["shape"]) * 1e7
    T1 = oracle.RefPlan(case["shape"], coef).step(T0, S) if structured else oracle.fast_step(coef, T0, S)
    M = coef[:, 0].reshape(case["shape"])  # rho c / dt
    lhs = np.sum(M * (T1 - T0))
    assert abs(lhs - S.sum()) / S.sum() < 1e-10


def test_constant_field_is_steady():
    case = _uniform_case()
    coef = oracle.assemble_case(case)
    T = np.full(case["shape"], 311.25)
    assert np.array_equal(oracle.RefPlan(case["shape"], coef).step(T), T) or \
        np.abs(oracle.RefPlan(case["shape"], coef).step(T) - T).max() < 1e-12
    assert np.abs(oracle.fast_step(coef, T) - T).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_one_dimensional_reduction_is_crank_nicolson(axis):
    """A field varying along one axis only evolves by textbook 1-D
    Crank-Nicolson (M - L/2) T1 = (M + L/2) T0 along that axis, solved
    densely; this fixes the 1/2 weights, M and g of the split."""
    shape = [5, 5, 5]
    shape[axis] = 11
    case = _uniform_case(tuple(shape))
    coef = oracle.assemble_case(case)
    L = shape[axis]
    prof = 293 + np.sin(np.linspace(0, 3, L)) * 4
    sh = [1, 1, 1]; sh[axis] = L
    T0 = np.broadcast_to(prof.reshape(sh), case["shape"]).copy()
    T1 = oracle.RefPlan(case["shape"], coef).step(T0)
    rc = case["materials"][0][4] * case["materials"][0][5]
    Mv = rc / case["dt"]
    g = case["materials"][0][1 + axis] / case["d"][axis] ** 2
    Lop = np.zeros((L, L))
    for i in range(L):
        if i > 0: Lop[i, i - 1] += g; Lop[i, i] -= g
        if i < L - 1: Lop[i, i + 1] += g; Lop[i, i] -= g
    I = np.eye(L)
    expect = np.linalg.solve(Mv * I - 0.5 * Lop, (Mv * I + 0.5 * Lop) @ prof)
    got = np.moveaxis(T1, axis, 0)[:, 0, 0]
    assert np.abs(got - expect).max() / np.abs(prof - 293).max() < 1e-12
    # the other two directions stay uniform
    assert np.ptp(np.moveaxis(T1, axis, 0), axis=(1, 2)).max() < 1e-11


def test_surface_readout_closed_form():
    rng = np.random.default_rng(2)
    T = 300 + rng.random((3, 4, 5))
    assert np.allclose(oracle.surface_readout(T, 5e-4), T[0] - 5.0 / 16.0 * (T[1] - T[0]),
                       rtol=0, atol=1e-12)


def test_source_field_semantics():
    """Impulse fires only at its exact time; stepped is inclusive; point
    source divides by the cell volume; the z-decay integrates
    exp(-d/l)/l over each cell's depth interval."""
    v = np.zeros((3, 2, 2), np.uint8); v[0] = 1; v[1] = 2; v[2] = 3
    dt, V = 1e-3, 2e-10
    vol = [(0,), (1, 0.002, 5.0), (2, 0.001, 0.003, 7.0), (3, 0.002, 3.0)]
    assert not oracle.source_field(vol, v, 0.0, dt, V).any()
    S = oracle.source_field(vol, v, 0.002, dt, V)
    assert np.all(S[0] == 5.0 / dt) and np.all(S[1] == 7.0) and np.allclose(S[2], 3.0 / (V * dt))
    assert np.all(oracle.source_field(vol, v, 0.003, dt, V)[1] == 7.0)
    assert not oracle.source_field(vol, v, 0.0031, dt, V)[1].any()
    zg = (np.arange(3) + 0.5)[:, None, None] * np.ones((3, 2, 2)) * 1e-3
    vz = np.ones((3, 2, 2), np.uint8)
    S = oracle.source_field([(0,), (4, 0.0, (1.0, 0, 0), 0.0, zg, 1e-3, 1e4, 2e-3)], vz, 0.0, dt, V)
    edges = np.arange(4) * 1e-3
    expect = (np.exp(-edges[:-1] / 2e-3) - np.exp(-edges[1:] / 2e-3)) * 1e4 / (1e-3 * dt)
    assert np.allclose(S[:, 0, 0], expect, rtol=1e-13)
    # energy: the column integral equals the deposited fraction
    assert abs(S[:, 0, 0].sum() * 1e-3 * dt / 1e4 - (1 - np.exp(-1.5))) < 1e-12


def test_assembly_errors():
    case = _uniform_case(tensor=0)
    case["boundaries"] = [(cases.BND_ANISO, 0.0), (cases.BND_INSULATING, 0.0)]
    with pytest.raises(ValueError, match="conductivity type"):
        oracle.assemble_case(case)
    case = _uniform_case()
    case["maps"][1][0, 1, 1] = 0  # conducting outer z face
    with pytest.raises(ValueError, match="outer face"):
        oracle.assemble_case(case)

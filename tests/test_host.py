"""Host-side logic that runs without a GPU: the C-ABI library's exports, the
compact coefficient tables of setup(), setup()'s input validation, the
upstream's assembled operators and Thomas utilities, the module names, the
grid builders, hs2_indexing, the launch-shape candidate lists and the rocprof
summary reader."""
import ctypes
import os
import re

import numpy as np
import pytest

import heatsim2_amd as hs
from heatsim2_amd import _lib, autotune, crank_nicolson, hs2_indexing, rocprof_summary, tridiag
from hs2_cases import load_golden, to_package_args

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "heatsim2_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hs2_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = _header_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names
    L = _lib.load()
    assert L.hs2_version().decode().startswith("heatsim2_amd %s " % hs.__version__)


@pytest.mark.parametrize("dims,spacing,code,msg", [
    ((2, 2, 1025), 1e-3, 2, "nx=1025"),
    ((2, 1025, 2), 1e-3, 2, "ny=1025"),
    ((1025, 2, 2), 1e-3, 2, "nz=1025"),
    ((1024, 1024, 2048), 1e-3, 2, "2^31"),
    ((0, 4, 4), 1e-3, 1, "must be positive"),
    ((4, 4, 4), 0.0, 1, "must be positive"),
])
def test_plan_create_rejects_before_touching_the_device(dims, spacing, code, msg):
    """hs2_plan_create validates the grid (axis lengths up to 1024, fewer than
    2^31 indices, positive sizes and spacings) before its first HIP call, so
    these errors come back on a host without a GPU, with no plan handed out."""
    L = _lib.load()
    g = _lib.GridDesc(dims[0], dims[1], dims[2], spacing, 1e-3, 1e-3, 1e-3)
    h = ctypes.c_void_p(123)
    assert L.hs2_plan_create(ctypes.byref(g), ctypes.byref(h)) == code
    assert h.value is None
    assert msg in L.hs2_last_error().decode()


def _compact_tables(case):
    pa = to_package_args(hs, case)
    a = pa["setup"]
    m, bz, by, bx = a[13], a[14], a[15], a[16]
    d = (a[3], a[4], a[5])
    return crank_nicolson.build_tables(d, a[9], a[10], a[11], m, bz, by, bx), (m, bz, by, bx)


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso", "steel_cube40"])
def test_compact_tables_match_oracle_assembly(name):
    """Every COMPUTE cell's minus-face conductance and rho*c/dt, looked up in
    setup()'s compact tables, equals the oracle's per-cell assembly (which
    is pinned to the upstream's assembled operators)."""
    case = load_golden(name)
    (gtab, mtab, compute), (m, bz, by, bx) = _compact_tables(case)
    ref = oracle.assemble_case(case).reshape(case["shape"] + (8,))
    nz, ny, nx = case["shape"]
    nmat = len(case["materials"])
    comp = compute[m].astype(bool)
    np.testing.assert_allclose(mtab[m][comp], ref[..., 0][comp], rtol=1e-15)
    faces = ((0, bz[1:-1], (slice(1, None),)), (1, by[:, 1:-1], (slice(None), slice(1, None))),
             (2, bx[:, :, 1:-1], (slice(None), slice(None), slice(1, None))))
    for axis, bint, sl in faces:
        lo = [slice(None)] * 3
        lo[axis] = slice(0, -1)
        hi = [slice(None)] * 3
        hi[axis] = slice(1, None)
        g = gtab[bint.astype(int), axis, m[tuple(lo)], m[tuple(hi)]]
        # as seen by the plus-side cell (its minus face) where that cell computes
        c_hi = comp[tuple(hi)]
        np.testing.assert_allclose(g[c_hi], ref[tuple(hi)][..., 2 + 2 * axis][c_hi], rtol=1e-14,
                                   atol=0)
        # and by the minus-side cell (its plus face)
        c_lo = comp[tuple(lo)]
        np.testing.assert_allclose(g[c_lo], ref[tuple(lo)][..., 3 + 2 * axis][c_lo], rtol=1e-14,
                                   atol=0)


def _setup_args(name="tiny_mixed"):
    return list(to_package_args(hs, load_golden(name))["setup"])


def test_setup_rejects_bad_maps():
    a = _setup_args()
    a[13] = a[13].astype(np.int32)  # material map must be uint8
    with pytest.raises(ValueError):
        hs.setup(*a)
    a = _setup_args()
    a[14] = a[14][:-1]  # bz must be (nz+1, ny, nx)
    with pytest.raises(ValueError):
        hs.setup(*a)
    a = _setup_args()
    a[13] = a[13].copy()
    a[13][0, 0, 0] = 9  # only 3 materials
    with pytest.raises(IndexError):
        hs.setup(*a)


def test_setup_rejects_flux_through_outer_face():
    a = _setup_args()
    a[14] = a[14].copy()
    a[14][0, 2, 2] = 0  # conducting class on the top outer face of a COMPUTE cell
    with pytest.raises(ValueError, match="outer faces must be insulating"):
        hs.setup(*a)


def test_setup_rejects_wrong_conductivity_type():
    a = _setup_args("tiny_aniso")
    mats = list(a[10])
    mats[0] = (hs.TEMPERATURE_COMPUTE, 5.0, mats[0][2], mats[0][3])  # scalar k, anisotropic law
    a[10] = tuple(mats)
    with pytest.raises((TypeError, ValueError)):
        hs.setup(*a)


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso"])
def test_unaligned_anisotropic_fails_as_upstream(name):
    """The upstream's setup cannot complete in this mode: adi_setup builds 3
    steps, adi_expressions returns 15, and add_equation_to_adi_matrices
    asserts they agree (alternatingdirection_c_pyx.pyx:418-452, 511-564).
    The same call fails here with an AssertionError, before any device work."""
    a = _setup_args(name)
    with pytest.raises(AssertionError, match="15 substep"):
        hs.setup(*a, unaligned_anisotropic=True)
    with pytest.raises(NotImplementedError):
        hs.setup(*a, unaligned_anisotropic=True)


@pytest.mark.parametrize("entry", [(0, 1), (0, 2), (1, 2), (2, 1)])
def test_off_diagonal_conductivity_fails_as_upstream(entry):
    """Off-diagonal K without the unaligned mode leaves tangential groups in
    the last substep, and the upstream fails `assert no_groups`
    (alternatingdirection_c_pyx.pyx:570).  Each off-diagonal entry alone is
    caught, (2, 1) included, whose qx term is T56p - T54p and T55m - T54m
    upstream (boundary_conducting_anisotropic.py:60)."""
    a = _setup_args("tiny_aniso")
    mats = list(a[10])
    k = np.array(mats[0][1], dtype=np.float64).copy()
    k[entry] = 0.1 * k[entry[0], entry[0]]
    mats[0] = (mats[0][0], k) + tuple(mats[0][2:])
    a[10] = tuple(mats)
    with pytest.raises(AssertionError, match="off-diagonal"):
        hs.setup(*a)
    with pytest.raises(ValueError):
        hs.setup(*a)


def _rel_dense(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(float(np.abs(b).max()), 1e-300)


@pytest.mark.parametrize("name", ["tiny_mixed", "tiny_aniso", "tiny_curved"])
def test_reference_operators_match_upstream(name):
    """reference_operators (behind ADI_steps[s].Amat/Bmat/Cmats/Dvec), fed the
    oracle's per-cell equation coefficients, gives the upstream's assembled
    matrices of every substep (fixtures written by the real reference), and
    tridiaglu's factors multiply back to Amat."""
    import scipy.sparse as sp
    case = load_golden(name)
    z = case["fixture"]
    shape = case["shape"]
    coef = oracle.assemble_case(case).reshape(*shape, 8)
    M, D = coef[..., 0], coef[..., 1]
    gm = [coef[..., 2], coef[..., 4], coef[..., 6]]
    gp = [coef[..., 3], coef[..., 5], coef[..., 7]]
    n = M.size
    for s in range(3):
        A, B, C, Dv = crank_nicolson.reference_operators(shape, s, M, D, gm, gp)
        assert _rel_dense(A, z["A%d" % s]) < 1e-15, s
        assert _rel_dense(Dv, z["D%d" % s]) < 1e-15, s
        assert len(C) == s
        for which, mine in enumerate([B] + C):
            key = "B%d" % s if which == 0 else "C%d_%d" % (s, which - 1)
            ref = sp.csr_matrix((z[key + "_data"], z[key + "_indices"], z[key + "_indptr"]),
                                shape=(n, n))
            assert _rel_dense(mine.toarray(), ref.toarray()) < 1e-15, key
        L, U = tridiag.tridiaglu(A)
        assert (U[:, 1] == 1.0).all() and (U[:, 0] == 0.0).all() and (L[:, 2] == 0.0).all()
        # (L U)[r, r-1] = L[r, 0]; (L U)[r, r] = L[r, 0] U[r-1, 2] + L[r, 1];
        # (L U)[r, r+1] = L[r, 1] U[r, 2]
        prev = np.concatenate([[0.0], U[:-1, 2]])
        LU = np.stack([L[:, 0], L[:, 0] * prev + L[:, 1], L[:, 1] * U[:, 2]], axis=1)
        assert _rel_dense(LU, A) < 1e-14, s


def _tridiag_system(rng, nlines, line, cuts=0.1):
    """Diagonally dominant lines of `line` rows, a = 0 at each line's first
    row and c = 0 at its last, plus some zero couplings inside lines."""
    n = nlines * line
    A = np.empty((n, 3))
    A[:, 0] = -rng.random(n)
    A[:, 2] = -rng.random(n)
    A[:, 1] = 2.5 + rng.random(n)
    A[::line, 0] = 0.0
    A[line - 1::line, 2] = 0.0
    A[rng.random(n) < cuts, 0] = 0.0
    A[rng.random(n) < cuts, 2] = 0.0
    return A


def _dense(A):
    n = A.shape[0]
    return np.diag(A[:, 1]) + np.diag(A[1:, 0], -1) + np.diag(A[:-1, 2], 1)


def test_tridiag_solves_and_factors():
    """tridiag.tridiaglu / tridiagsolve (tridiag.pyx:9-69): L U equals A and
    the solve equals a dense solve, on many lines with zero couplings inside
    them; and the upstream's 4 x 4 diagnostic (tridiag.pyx:71-112), seeded."""
    rng = np.random.default_rng(3)
    for nlines, line in ((1, 1), (1, 7), (5, 16), (37, 33)):
        A = _tridiag_system(rng, nlines, line)
        L, U = tridiag.tridiaglu(A)
        dense_L = np.diag(L[:, 1]) + np.diag(L[1:, 0], -1)
        dense_U = np.diag(U[:, 1]) + np.diag(U[:-1, 2], 1)
        assert np.abs(dense_L @ dense_U - _dense(A)).max() < 1e-13
        b = rng.random(A.shape[0])
        x = tridiag.tridiagsolve(L, U, b)
        assert np.abs(x - np.linalg.solve(_dense(A), b)).max() < 1e-13 * np.abs(x).max()
    A = rng.random((4, 3))
    A[0, 0] = A[-1, 2] = 0.0
    L, U = tridiag.tridiaglu(A)
    dense_A = _dense(A)
    LU = (np.diag(L[:, 1]) + np.diag(L[1:, 0], -1)) @ (np.diag(U[:, 1]) + np.diag(U[:-1, 2], 1))
    assert np.linalg.norm(dense_A - LU) / np.linalg.norm(dense_A) < 1e-10
    b = rng.random(4)
    x = tridiag.tridiagsolve(L, U, b)
    ref = np.linalg.inv(dense_A) @ b
    assert np.linalg.norm(ref - x) / np.linalg.norm(ref) < 1e-10
    with pytest.raises(AssertionError):
        tridiag.tridiaglu(np.ones((3, 3)))


def test_tridiag_runs_equal_one_recurrence():
    """Splitting rows into runs at zero couplings gives bit for bit what one
    recurrence over all rows gives (the upstream's serial pass)."""
    rng = np.random.default_rng(4)
    A = _tridiag_system(rng, 9, 20, cuts=0.2)
    n = A.shape[0]
    d = np.empty(n)
    cp = np.empty(n)
    for r in range(n):
        d[r] = A[r, 1] if r == 0 else A[r, 1] - cp[r - 1] * A[r, 0]
        cp[r] = A[r, 2] / d[r]
    L, U = tridiag.tridiaglu(A)
    assert np.array_equal(L[:, 1], d) and np.array_equal(U[:, 2], cp)
    b = rng.random(n)
    y = np.empty(n)
    for r in range(n):
        y[r] = b[r] / d[r] if r == 0 else (b[r] - A[r, 0] * y[r - 1]) / d[r]
    for r in range(n - 2, -1, -1):
        y[r] = y[r] - cp[r] * y[r + 1]
    assert np.array_equal(tridiag.tridiagsolve(L, U, b), y)


def test_upstream_module_names():
    """The names heatsim2/__init__.py:15-17 imports exist under the same
    module names."""
    import heatsim2_amd.alternatingdirection_c_pyx as adc
    assert adc.run_adi_steps is hs.run_adi_steps
    assert hs.alternatingdirection.pyadi_step is crank_nicolson.adi_step
    assert hs.alternatingdirection_c_pyx.adi_params is crank_nicolson.adi_params
    assert hs.setup is crank_nicolson.setup


def test_curvature_setup_arguments():
    """Both curvature arrays or none (the upstream indexes both,
    crank_nicolson.pyx:399); (ny, nx) shapes; edges stay positive."""
    a = _setup_args()
    ny, nx = a[7], a[8]
    with pytest.raises(TypeError):
        hs.setup(*a, top_surface_y_curvatures=np.zeros((ny, nx)))
    with pytest.raises(ValueError, match="shape"):
        hs.setup(*a, top_surface_y_curvatures=np.zeros((ny, nx)),
                 top_surface_x_curvatures=np.zeros((ny, nx + 1)))
    big = np.full((ny, nx), -1.0 / (a[3] * a[6]))  # edges shrink to zero at the bottom
    with pytest.raises(ValueError, match="not positive"):
        hs.setup(*a, top_surface_y_curvatures=big, top_surface_x_curvatures=np.zeros((ny, nx)))


def test_curved_coefficients_match_oracle():
    """The curvature-mode coefficients setup() uploads (per unit volume, each
    cell's own minus and plus faces) equal the oracle's volume-multiplied
    assembly divided by the volume; the oracle is pinned to the upstream's
    assembled operators for this case (test_oracle)."""
    from heatsim2_amd import crank_nicolson as cn
    case = load_golden("tiny_curved")
    a = list(to_package_args(hs, case)["setup"])
    d = tuple(case["d"])
    m, bz, by, bx = case["maps"][:4]
    nz = case["shape"][0]
    gtab, mtab, compute = cn.build_tables(d, case["dt"], a[10], a[11], m, bz, by, bx)
    geo = cn.curved_geometry(d, nz, *case["curvature"])
    M, w, gm, gp = cn.curved_coefficients(d, a[10], a[11], m, bz, by, bx, mtab, compute, geo)
    ref, vol = oracle.assemble_curved(case["shape"], d, case["dt"], case["materials"],
                                      case["boundaries"], case["maps"], *case["curvature"])
    assert np.allclose(geo["vol"].ravel(), vol.ravel(), rtol=1e-15, atol=0)
    comp = ref[:, 1] != 0.0
    V = np.where(comp, vol.ravel(), 1.0)
    got = np.stack([M.ravel(), w.ravel(), gm[0].ravel(), gp[0].ravel(), gm[1].ravel(),
                    gp[1].ravel(), gm[2].ravel(), gp[2].ravel()], axis=-1)
    want = ref / V[:, None]
    want[:, 1] = comp
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)) < 1e-14
    # the case has cells whose y / x faces differ from the two sides
    assert np.abs(gp[2][:, :, :-1] - gm[2][:, :, 1:]).max() > 1e-3 * gm[2].max()


def test_grid_builders_closed_forms():
    g = hs.build_grid(0.0, 1.0, 4, -1.0, 1.0, 5, 0.0, 3.0, 6)
    dz, dy, dx, z, y, x, zg, yg, xg, zb = g[:10]
    assert (dz, dy, dx) == (0.25, 0.4, 0.5)
    np.testing.assert_allclose(z, [0.125, 0.375, 0.625, 0.875])
    np.testing.assert_allclose(y, -1.0 + 0.2 + 0.4 * np.arange(5), atol=1e-15)
    assert zg.shape == (4, 5, 6) and np.all(zg[:, 0, 0] == z) and np.all(xg[0, 0, :] == x)
    np.testing.assert_allclose(g[-2], np.sqrt(xg ** 2 + yg ** 2 + zg ** 2))
    g2 = hs.build_grid_min_step(0.125, 0.25, 4, -0.8, 0.4, 5, 0.25, 0.5, 6)
    np.testing.assert_allclose(g2[0], z)
    g3 = hs.build_grid_min_step_edge(0.0, 0.25, 4, -1.0, 0.4, 5, 0.0, 0.5, 6)
    np.testing.assert_allclose(g3[6], zb)
    maps = hs.zero_elements(2, 3, 4)
    assert [mm.shape for mm in maps] == [(2, 3, 4), (3, 3, 4), (2, 4, 4), (2, 3, 5), (2, 3, 4)]
    assert all(mm.dtype == np.uint8 for mm in maps)


def test_hs2_indexing_round_trip():
    s = [0, 1, 1, 2]
    shape = (4, 8, 16)
    total = sum((shape[1] >> q) * (shape[2] >> q) for q in s)
    seen = set()
    for flat in range(total):
        ci = hs2_indexing.compressed_index_from_single(flat, s, shape)
        assert hs2_indexing.compressed_single_index(ci, s, shape) == flat
        (k0, k1), (j0, j1), (i0, i1) = hs2_indexing.uncompressed_index_range(ci, s)
        for j in range(j0, j1):
            for i in range(i0, i1):
                assert hs2_indexing.compressed_index((k0, j, i), s) == ci
                seen.add((k0, j, i))
    assert len(seen) == shape[0] * shape[1] * shape[2]
    with pytest.raises(IndexError):
        hs2_indexing.compressed_index_from_single(total, s, shape)


def test_launch_shape_candidates_fit_the_kernels():
    for nx in (1, 3, 16, 17, 100, 256, 300, 512, 1000, 1024):
        cands = autotune.x_candidates(nx)
        assert cands and all(c[0] * c[1] >= nx for c in cands)
        assert all(c[:3] in autotune.X_COMBOS for c in cands)
        # tile rows: 0, or a proper divisor of the two-phase kernel's lines
        assert all(c[3] == 0 or (c[2] > 0 and c[3] < c[2] and c[2] % c[3] == 0) for c in cands)
        assert {c[:3] for c in cands if c[3]} == {c[:3] for c in cands if c[2] > 1}
    for L in (1, 2, 5, 64, 100, 256, 517, 1024):
        for seg, cw in autotune.strided_candidates(L):
            assert -(-L // seg) * cw <= autotune.MAX_BLOCK


def test_rocprof_summary_attribution_and_committed_traffic():
    assert rocprof_summary.sweep_of("void hs2::k_sweep_strided<hs2::ClassView, 16, 1, true>(x)") == "y"
    assert rocprof_summary.sweep_of("void hs2::k_sweep_strided<hs2::ClassView, 8, 0, false>(x)") == "z"
    assert rocprof_summary.sweep_of("void hs2::k_sweep_x2<hs2::ClassView, 16, 16, 4, false>") == "x"
    assert rocprof_summary.sweep_of("hs2::k_classify") is None
    ry = "void hs2::k_sweep_strided<hs2::ClassView, 16, 1, true, true>(x)"
    assert rocprof_summary.sweep_of(ry) == "y"
    assert rocprof_summary.algo_bytes_per_cell("y", ry) == 16.0
    assert rocprof_summary.algo_bytes_per_cell("y", ry.replace("true>", "false>")) == 24.0
    assert rocprof_summary.algo_bytes_per_cell("x", "void hs2::k_sweep_x2<hs2::ClassView, 16, "
                                                    "16, 16, false, true, true>") == 16.0
    tr = rocprof_summary.read_traffic(os.path.join(ROOT, "profiles",
                                                   "hbm_traffic_two_layer_256.txt"))
    assert tr["cells"] == 256 ** 3
    # the profiled configuration hands y its right-hand side (16 B per cell)
    for k, algo in (("sweep_x", 16), ("sweep_y", 16), ("sweep_z", 24)):
        assert algo * 256 ** 3 <= tr["traffic_bytes"][k] < 2 * algo * 256 ** 3


def test_volumetric_map_change_detection():
    """The source map is uploaded again exactly when a byte changes (eight
    bytes compared at a time, odd sizes byte by byte)."""
    from heatsim2_amd.alternatingdirection import _same_map
    rng = np.random.default_rng(3)
    for shape in ((4, 6, 8), (3, 5, 7)):
        a = rng.integers(0, 4, size=shape).astype(np.uint8)
        assert _same_map(a, a.copy())
        for flat in (0, a.size // 2, a.size - 1):
            b = a.copy()
            b.reshape(-1)[flat] ^= 1
            assert not _same_map(a, b)
        assert not _same_map(a, a[:-1])
        assert _same_map(a, np.asfortranarray(a))


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without libhs2.so the package raises on first use; it never falls
    back to a CPU path."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libhs2.so"))
    with pytest.raises(ImportError, match="not built"):
        _lib.load()
    with pytest.raises(ImportError, match="not built"):
        hs.setup(0.0, 0.0, 0.0, 1e-3, 1e-3, 1e-3, 2, 2, 2, 1e-3,
                 ((hs.TEMPERATURE_COMPUTE, 1.0, 1.0, 1.0),), ((hs.boundary_insulating,),),
                 ((hs.NO_SOURCE,),), np.zeros((2, 2, 2), np.uint8),
                 np.zeros((3, 2, 2), np.uint8), np.zeros((2, 3, 2), np.uint8),
                 np.zeros((2, 2, 3), np.uint8), np.zeros((2, 2, 2), np.uint8))


def test_trace_summary_timed_run(tmp_path):
    """The trace summary's per-sweep means: over every launch of the most
    launched instantiation, and over launches skip .. skip + count - 1 in
    dispatch order (a bench's timed run, without its warm-up and timing
    passes), on a synthetic kernel trace."""
    import csv
    names = {"x": "void hs2::k_sweep_x2<hs2::CombView, 16, 16, 16, false, true, true, true>(a)",
             "y": "void hs2::k_sweep_strided<hs2::ClassView, 16, 1, true, true>(a)",
             "z": "void hs2::k_sweep_strided<hs2::ClassView, 16, 0, true, false>(a)"}
    dur = {"x": 70.0, "y": 56.0, "z": 74.0}
    rows, d, t = [], 0, 0
    for n in range(30):  # 5 warm-up, 20 timed, 5 slower timing-pass launches
        for s in "xyz":
            us = dur[s] + (10.0 if n >= 25 else 0.0) + (5.0 if n < 5 else 0.0)
            rows.append({"Dispatch_Id": d, "Kernel_Name": names[s], "Start_Timestamp": t,
                         "End_Timestamp": t + int(us * 1000)})
            d += 1
            t += int(us * 1000) + 1000
    with open(tmp_path / "kt_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)
    with open(tmp_path / "kt_kernel_stats.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs",
                                          "MinNs", "MaxNs"])
        w.writeheader()
        for s in "xyz":
            w.writerow({"Name": names[s], "Calls": 30, "TotalDurationNs": 1, "AverageNs": 1,
                        "MinNs": 1, "MaxNs": 1})
    text, means = rocprof_summary.trace_summary(str(tmp_path), skip=5, count=20)
    assert means == {s: dur[s] for s in "xyz"}
    assert "timed run: launches 5 .. 24 of each" in text
    _, all_means = rocprof_summary.trace_summary(str(tmp_path))
    for s in "xyz":
        assert abs(all_means[s] - (dur[s] + (5 * 10.0 + 5 * 5.0) / 30)) < 1e-9


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before
    touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

"""The upstream's own regime: its demos run at dt = 10-50 ms
(demos/steelonfoam.py:21, demos/testheatsim2.py:82, demos/ktest.py:147),
where lines are strongly coupled and the line solvers' early exit (once
every coupling left after the first reduction level is below 2^-64) cannot
fire, so every PCR level runs.

- Golden trajectories of two upstream demos at their 10 ms step, written by
  the upstream solver itself (tests/golden/make_golden.py): steel on foam
  (80x40x48, strongly coupled z lines) and the tutorial (5x50x55 titanium,
  strongly coupled x and y lines), stepped on the GPU with the untuned
  default launch shapes.
- Oracle parity at dt = 50 ms on shapes whose default shapes are 16-cell
  segments on every axis: (4, 96, 256) and (96, 4, 64) steel slabs.
Each case first shows, with a host restatement of the device's segment
elimination and first PCR level, that the early exit does not fire.
Bar: 1e-10 of the peak rise.
"""
import numpy as np
import pytest
import torch

import heatsim2_amd as hs
import oracle
from heatsim2_amd import _lib
from hs2_cases import cases, load_golden, to_package_args

TOL = 1e-10
NEGLIGIBLE = 2.0 ** -64


def _line_rows(coef, shape, axis, line):
    """Tridiagonal rows (a, b, c) of the implicit substep along one line:
    (M - L/2) with the oracle's per-cell rows (M, D, gzm, gzp, gym, gyp,
    gxm, gxp); `line` indexes the other two axes."""
    nz, ny, nx = shape
    C = coef.reshape(nz, ny, nx, 8)
    sl = list(line)
    sl.insert(axis, slice(None))
    r = C[tuple(sl)]
    gm, gp = r[:, 2 + 2 * axis], r[:, 3 + 2 * axis]
    return -0.5 * gm, r[:, 0] + 0.5 * (gm + gp), -0.5 * gp


def _coupling_after_pairing(a, b, c, seg):
    """The device's early-exit test for one line: normalise the rows, reduce
    every seg-cell segment to its two end rows (segment elimination), pair
    each start row with the previous segment's end row and each end row with
    the next segment's start row (the first PCR level), and return the
    largest coupling left."""
    a, c, r = a / b, c / b, np.zeros_like(b)
    L = len(b)
    ends = []
    for s0 in range(0, L, seg):
        idx = range(s0, min(s0 + seg, L))
        # end rows by dense elimination of the segment's interior
        n = len(idx)
        if n == 1:
            ends.append((a[s0], c[s0], a[s0], c[s0]))
            continue
        A = np.zeros((n, n + 2))  # columns: x_before, x_0..x_{n-1}, x_after
        for q, i in enumerate(idx):
            A[q, q] = a[i]
            A[q, q + 1] = 1.0
            A[q, q + 2] = c[i]
        # express the end rows in x_before, x_0, x_{n-1}, x_after
        keep = [0, 1, n, n + 1]
        inner = [k for k in range(n + 2) if k not in keep]
        rows_end = [0, n - 1]
        if inner:
            Ai = A[1:n - 1][:, inner]
            sol = np.linalg.solve(Ai, -A[1:n - 1][:, keep])
            full = A[:, keep] + A[:, inner] @ sol
        else:
            full = A[:, keep]
        s_row, e_row = full[rows_end[0]], full[rows_end[1]]
        s_row, e_row = s_row / s_row[1], e_row / e_row[2]
        ends.append((s_row[0], s_row[2], e_row[1], e_row[3]))
    worst = 0.0
    zero = (0.0, 0.0, 0.0, 0.0)
    for p, (As, Cs, Ae, Ce) in enumerate(ends):
        prev = ends[p - 1] if p > 0 else zero
        nxt = ends[p + 1] if p + 1 < len(ends) else zero
        # s_p couples e_{p-1} and e_p: eliminating them leaves s_{p-1}, s_{p+1}
        k = 1.0 / (1.0 - As * prev[3] - Cs * Ae)
        worst = max(worst, abs(As * prev[2] * k), abs(Cs * Ce * k))
        # e_p couples s_p and s_{p+1}: eliminating them leaves e_{p-1}, e_{p+1}
        k = 1.0 / (1.0 - Ae * Cs - Ce * nxt[0])
        worst = max(worst, abs(Ae * As * k), abs(Ce * nxt[1] * k))
    return worst


def _defaults(P):
    info = _lib.plan_info(P.plan)
    return {0: info["z_seg"], 1: info["y_seg"], 2: info["x_cpl"]}


def _assert_full_reduction(case, P, axis, line):
    coef = oracle.assemble_case(case)
    seg = _defaults(P)[axis]
    a, b, c = _line_rows(coef, case["shape"], axis, line)
    left = _coupling_after_pairing(a, b, c, seg)
    print("%s: axis %d, %d-cell segments: coupling after the first PCR level %.3e"
          % (case.get("name"), axis, seg, left), flush=True)
    assert left > NEGLIGIBLE * 1e6, left


@pytest.mark.gpu
@pytest.mark.parametrize("name,axis,line", [("steelonfoam", 0, (20, 30)),
                                            ("tutorial", 2, (0, 25)),
                                            ("tutorial", 1, (0, 27))])
def test_demo_trajectory_matches_upstream(name, axis, line):
    case = load_golden(name)
    z = case["fixture"]
    args = to_package_args(hs, case)
    P, S = hs.setup(*args["setup"])
    _assert_full_reduction(case, P, axis, line)
    st = int(z["top_stride"])
    zy, zx = (int(v) for v in z["zprofile_at"])
    rise = float(np.abs(z["top_history"]).max())
    T = torch.from_numpy(case["T0"]).cuda()
    worst = 0.0
    for n in range(case["nsteps"]):
        T = hs.run_adi_steps(P, S, case["t0"] + n * case["dt"], case["dt"], T, args["vmap"],
                             args["volumetric"])
        Th = T.cpu().numpy()
        worst = max(worst, float(np.abs(Th[0, ::st, ::st] - z["top_history"][n]).max()),
                    float(np.abs(Th[:, zy, zx] - z["zprofile"][n]).max()))
    worst = max(worst, float(np.abs(T.cpu().numpy() - z["final"]).max()))
    print("%s x %d steps: %.3e of the peak rise %.4g K" % (name, case["nsteps"], worst / rise,
                                                           rise), flush=True)
    assert worst <= TOL * rise


def _steel_slab(shape, dt=0.05):
    return cases.case_steel_block(*shape, nsteps=60, dt=dt)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,axes", [((4, 96, 256), (1, 2)), ((96, 4, 64), (0,))],
                         ids=["4x96x256", "96x4x64"])
def test_strongly_coupled_default_shapes_match_oracle(shape, axes):
    case = _steel_slab(shape)
    pa = to_package_args(hs, case)
    P, S = hs.setup(*pa["setup"])
    for a in axes:
        line = [s // 2 for i, s in enumerate(shape) if i != a]
        _assert_full_reduction(case, P, a, tuple(line))
    T, hist = hs.run_adi_steps_many(P, 0.0, case["dt"], case["nsteps"],
                                    torch.from_numpy(case["T0"]).cuda(), pa["vmap"],
                                    pa["volumetric"], surface_history=True)
    coef = oracle.assemble_case(case)
    S0 = oracle.source_field(case["volumetric"], case["maps"][4], 0.0, case["dt"],
                             float(np.prod(case["d"])))
    To, ho = oracle.fast_run(coef, case["T0"], case["nsteps"], S0, 1, case["d"][0])
    rise = float(ho.max() - case["T0"].max())
    ef = float(np.abs(T.cpu().numpy() - To).max()) / rise
    eh = float(np.abs(hist.cpu().numpy() - ho).max()) / rise
    print("%s: final %.3e, history %.3e" % (case["name"], ef, eh), flush=True)
    assert ef <= TOL and eh <= TOL


def test_exit_rule_restatement_against_closed_form():
    """The host restatement of the early-exit test: on a uniform line the
    coupling after pairing falls like rho^seg with rho the decay rate of
    the constant-coefficient recurrence; weak coupling gives a negligible
    value, strong coupling a large one."""
    for r, expect_big in ((0.01, False), (3.0, True)):
        L, seg = 64, 16
        g = r  # M = 1, g / M = r
        a = np.full(L, -0.5 * g)
        c = np.full(L, -0.5 * g)
        a[0] = c[-1] = 0.0
        b = 1.0 + 0.5 * g * ((np.arange(L) > 0).astype(float) + (np.arange(L) < L - 1))
        left = _coupling_after_pairing(a, b, c, seg)
        t = 0.5 * g / (1.0 + g)
        rho = (1.0 - np.sqrt(1.0 - 4.0 * t * t)) / (2.0 * t)
        assert (left > NEGLIGIBLE) == expect_big, (r, left)
        assert 0.01 < left / rho ** seg < 100.0, (r, left, rho ** seg)

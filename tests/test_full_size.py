"""Long runs against the oracle at the benchmark configurations' sizes:
the 256^3 two-layer stack x 1000 steps (and 128^3), the 64x128x128 and
32x64x64 laminates x 2000 steps, and the 128x512x512 laminate x 200 steps
(all 2000 steps of it when HS2_FULL_SIZE=1: about 5 min of host time for
the oracle).

Bar: max |T_gpu - T_oracle| over the final field and over the surface
history <= 1e-10 of the peak temperature rise (the history's maximum above
the initial field).  The oracle (OpenMP delta-form step on the host cores)
is pinned to the upstream solver by the trajectory fixtures
(tests/test_oracle.py).
"""
import os

import numpy as np
import pytest
import torch

import heatsim2_amd as hs
import oracle
from hs2_cases import cases, to_package_args

FULL = bool(os.environ.get("HS2_FULL_SIZE"))
TOL = 1e-10


def _gpu_run(case):
    pa = to_package_args(hs, case)
    P, S = hs.setup(*pa["setup"])
    T, hist = hs.run_adi_steps_many(P, 0.0, case["dt"], case["nsteps"],
                                    torch.from_numpy(case["T0"]).cuda(), pa["vmap"],
                                    pa["volumetric"], surface_history=True)
    return T.cpu().numpy(), hist.cpu().numpy()


def _oracle_run(case, chunk=100):
    """The oracle's steps in chunks (the source fires on the first step
    only), printing progress so that a long run is visibly alive."""
    coef = oracle.assemble_case(case)
    S = oracle.source_field(case["volumetric"], case["maps"][4], 0.0, case["dt"],
                            float(np.prod(case["d"])))
    T, hists = case["T0"], []
    for n0 in range(0, case["nsteps"], chunk):
        k = min(chunk, case["nsteps"] - n0)
        T, h = oracle.fast_run(coef, T, k, S if n0 == 0 else None, 1, case["d"][0])
        hists.append(h)
        print("  oracle %s: %d / %d steps" % (case["name"], n0 + k, case["nsteps"]), flush=True)
    return T, np.concatenate(hists)


def _check(case):
    Tg, hg = _gpu_run(case)
    To, ho = _oracle_run(case)
    rise = float(ho.max() - case["T0"].max())
    assert rise > 0
    e_final = float(np.abs(Tg - To).max()) / rise
    e_hist = float(np.abs(hg - ho).max()) / rise
    print("%s %s x %d steps: final %.3e, history %.3e (of peak rise %.4g K)"
          % (case["name"], case["shape"], case["nsteps"], e_final, e_hist, rise))
    assert e_final <= TOL and e_hist <= TOL


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [128, 256])
def test_two_layer_1000_steps_matches_oracle(n):
    _check(cases.case_two_layer(nz=n, ny=n, nx=n, nsteps=1000))


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("shape", [(32, 64, 64), (64, 128, 128)],
                         ids=lambda s: "x".join(map(str, s)))
def test_laminate_2000_steps_matches_oracle(shape):
    _check(cases.case_laminate(*shape, nsteps=2000))


@pytest.mark.gpu
@pytest.mark.timeout(900)
@pytest.mark.parametrize("nsteps", [200, pytest.param(2000, marks=pytest.mark.skipif(
    not FULL, reason="all 2000 steps need HS2_FULL_SIZE=1 (about 5 min of host time)"))])
def test_laminate_128x512x512_matches_oracle(nsteps):
    _check(cases.case_laminate(128, 512, 512, nsteps=nsteps))

"""Thomas factorisation and solve on the upstream's (n, 3) tridiagonal layout.

Same functions and layout as heatsim2's tridiag (tridiag.pyx:9-69): row r of
Amat holds (sub-diagonal a_r, diagonal b_r, super-diagonal c_r), with
a_0 = c_{n-1} = 0.  These are host utilities for the assembled operators
(ADI_steps[s].Amat etc.); the device step solves its lines in the sweep
kernels and never calls them.

The upstream runs one recurrence over all n rows.  Wherever a_r = 0 exactly,
its factor recurrence restarts exactly (d'_r = b_r - c'_{r-1} * 0 = b_r), and
so does the forward solve; wherever c_r = 0 the back substitution does.  Here
the rows are split into those runs and the recurrences step through all runs
at once, with the same arithmetic per row as a serial pass.
"""
import numpy as np


def _runs(breaks, n):
    """Start row and length of each run; a run starts at row 0 and at every
    row where breaks[r] holds."""
    starts = np.flatnonzero(breaks)
    if starts.size == 0 or starts[0] != 0:
        starts = np.concatenate([[0], starts])
    lengths = np.diff(np.concatenate([starts, [n]]))
    return starts, lengths


def _by_position(starts, lengths):
    """For q = 0, 1, ...: the rows at position q of every run long enough."""
    order = np.argsort(-lengths, kind="stable")
    starts, lengths = starts[order], lengths[order]
    for q in range(int(lengths.max()) if lengths.size else 0):
        k = np.searchsorted(-lengths, -q, side="left")  # runs with length > q
        yield q, starts[:k] + q


def tridiaglu(Amat):
    """(Lmat, Umat) as tridiag.tridiaglu returns them (tridiag.pyx:9-43):
    Lmat[:, 0] = a, Lmat[:, 1] = d' (d'_0 = b_0, d'_r = b_r - a_r c'_{r-1}),
    Umat[:, 1] = 1, Umat[:, 2] = c' = c / d'.  Requires a_0 = c_{n-1} = 0,
    as upstream (which asserts it)."""
    A = np.asarray(Amat, dtype=np.float64)
    if A.ndim != 2 or A.shape[1] != 3:
        raise ValueError("Amat must be (n, 3)")
    n = A.shape[0]
    if n == 0:
        return np.zeros((0, 3)), np.zeros((0, 3))
    if A[0, 0] != 0.0 or A[-1, 2] != 0.0:
        raise AssertionError("tridiaglu: Amat[0,0] and Amat[-1,2] must be 0")
    a, b, c = A[:, 0], A[:, 1], A[:, 2]
    d = np.empty(n)
    cp = np.empty(n)
    starts, lengths = _runs(a == 0.0, n)
    for q, rows in _by_position(starts, lengths):
        d[rows] = b[rows] if q == 0 else b[rows] - cp[rows - 1] * a[rows]
        cp[rows] = c[rows] / d[rows]
    Lmat = np.zeros((n, 3))
    Lmat[1:, 0] = a[1:]
    Lmat[:, 1] = d
    Umat = np.zeros((n, 3))
    Umat[:, 1] = 1.0
    Umat[:, 2] = cp
    return Lmat, Umat


def tridiagsolve(Lmat, Umat, bvec):
    """x with (L U) x = bvec, as tridiag.tridiagsolve (tridiag.pyx:46-69):
    y_0 = b_0 / d'_0, y_r = (b_r - a_r y_{r-1}) / d'_r, then
    x_r = y_r - c'_r x_{r+1} from the last row up."""
    Lm = np.asarray(Lmat, dtype=np.float64)
    Um = np.asarray(Umat, dtype=np.float64)
    bv = np.asarray(bvec, dtype=np.float64)
    n = bv.shape[0]
    if Lm.shape != (n, 3) or Um.shape != (n, 3):
        raise ValueError("Lmat and Umat must be (n, 3) for n = len(bvec)")
    x = np.zeros(n)
    if n == 0:
        return x
    a, d, cp = Lm[:, 0], Lm[:, 1], Um[:, 2]
    starts, lengths = _runs(a == 0.0, n)
    for q, rows in _by_position(starts, lengths):
        x[rows] = bv[rows] / d[rows] if q == 0 else (bv[rows] - a[rows] * x[rows - 1]) / d[rows]
    # back substitution runs bottom-up; a row with c' = 0 keeps its y and
    # ends the run above it
    last = np.flatnonzero(np.concatenate([cp[:-1] == 0.0, [True]]))
    first = np.concatenate([[0], last[:-1] + 1])
    lengths = last - first + 1
    for q in range(1, int(lengths.max())):
        rows = last[lengths > q] - q
        x[rows] = x[rows] - cp[rows] * x[rows + 1]
    return x

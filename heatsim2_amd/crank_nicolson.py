"""setup(): build a device plan for the ADI Crank-Nicolson step.

Drop-in for heatsim2's crank_nicolson.setup (crank_nicolson.pyx:128-498),
same signature and argument conventions.  Where the upstream derives each
cell's equation symbolically and accumulates assembled matrices through C
(crank_nicolson.pyx:272-491, alternatingdirection_c.c:102-199), this
version:

1. evaluates every boundary plugin's face law once per (face class, axis,
   material pair) that occurs in the maps, applying the upstream's
   neighbour-k rule -- a neighbour's k counts only if it is in range and
   COMPUTE, else the cell's own k is used (crank_nicolson.pyx:298-338) --
   and reduces it to one conductance per volume g = -alpha/d for a flux
   q = alpha (T_p - T_m) (crank_nicolson.pyx:359-363);
2. M = rho*c/dt per COMPUTE material (crank_nicolson.pyx:378), M = 1 and
   no coupling for FIXED (crank_nicolson.pyx:473-486);
3. hands the uint8 maps (kept on the device) and these small tables to the
   HIP library, whose sweep kernels derive every coefficient on the fly
   (where an axis has more than 256 distinct coefficient classes -- many
   materials -- the library keeps per-cell float64 coefficients instead).

Validation happens here, before any launch, with the upstream's failure
modes turned into exceptions: IndexError for map entries beyond the tables,
TypeError / ValueError for a conductivity type that does not fit the face
law (expression.py:311-313, :737), ValueError for flux through an outer face
of a COMPUTE cell (the upstream exits, alternatingdirection_c.c:158-163).
"""
import numbers

import numpy as np
import torch

from heatsim2_amd import _lib, expression, tridiag
from heatsim2_amd.constants import TEMPERATURE_COMPUTE, TEMPERATURE_FIXED

# placeholder names of each face law's temperature arguments; the first two
# are the minus / plus cells across the face (crank_nicolson.pyx:220-250)
_T_ARGS = (
    ("Tm55", "Tp55", "Tm45", "Tp45", "Tm65", "Tp65", "Tm54", "Tp54", "Tm56", "Tp56", "Tm46",
     "Tp46", "Tm64", "Tp64"),
    ("T5m5", "T5p5", "T4m5", "T4p5", "T6m5", "T6p5", "T5m4", "T5p4", "T5m6", "T5p6", "T4m6",
     "T4p6", "T6m4", "T6p4"),
    ("T55m", "T55p", "T45m", "T45p", "T65m", "T65p", "T54m", "T54p", "T56m", "T56p", "T46m",
     "T46p", "T64m", "T64p"),
)
_AXIS_NAME = ("z", "y", "x")


class UnalignedAnisotropicError(NotImplementedError, AssertionError):
    """setup(..., unaligned_anisotropic=True).  The upstream fails this mode
    with an AssertionError before any step exists: adi_expressions returns 15
    substep expressions (alternatingdirection_c_pyx.pyx:511-564) while
    adi_setup always builds 3 steps (:418-445), so
    add_equation_to_adi_matrices fails `len(ADI_steps) ==
    len(spatial_expressions)` at the first cell (:452).  Substeps 3-14 would
    also have no line to solve on: their unknowns (e.g. T656p3, T654p3) are
    shifted on two axes, which add_equation asserts against
    (alternatingdirection_c.c:146), and sit in another plane than the
    substep's own T555p3.  Raised as both AssertionError, as upstream, and
    NotImplementedError."""


class OffDiagonalConductivityError(ValueError, AssertionError):
    """A face flux that depends on a tangential difference (off-diagonal
    conductivity) without the unaligned mode.  The upstream fails
    `assert no_groups(...)` here (alternatingdirection_c_pyx.pyx:570)."""


class adi_params(object):
    """Plan for one grid (alternatingdirection_c_pyx.pyx:75-85 fields plus
    the device state)."""
    shape = None
    volume_array = None

    def __init__(self, **kwargs):
        for key, val in kwargs.items():
            setattr(self, key, val)
        self._plan = None

    @property
    def plan(self):
        if self._plan is None:
            raise RuntimeError("plan released")
        return self._plan

    def release(self):
        if getattr(self, "_plan", None) is not None:
            _lib.load().hs2_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class adi_step(object):
    """Descriptor of one Douglas substep (alternatingdirection_c_pyx.pyx:
    120-167): the axis it solves and the upstream's permutation.  The
    operators live on the device as per-cell coefficients; Amat, Bmat, Cmats,
    Dvec, Lmat and Umat give them in the upstream's assembled form on first
    access (alternatingdirection_c_pyx.pyx:212-282), for inspection and
    comparison.  The step never forms them."""

    def __init__(self, ADI_params, permuteorder, stepnum):
        self.ADI_params = ADI_params
        self.permuteorder = tuple(permuteorder)
        self.stepnum = stepnum
        self.permutedshape = [ADI_params.shape[i] for i in permuteorder]
        inv = sorted((p, i) for i, p in enumerate(permuteorder))
        self.invpermuteorder = tuple(i for _, i in inv)
        self.axis = ("x", "y", "z")[stepnum]
        self._ops = None
        self._lu = None

    def _operators(self):
        if self._ops is None:
            params = self.ADI_params
            if getattr(params, "_eqcoef", None) is None:  # shared by the three substeps
                params._eqcoef = _equation_coefficients(params)
            M, D, gm, gp = params._eqcoef
            self._ops = reference_operators(self.ADI_params.shape, self.stepnum, M, D, gm, gp)
        return self._ops

    @property
    def Amat(self):
        """(n, 3): row r of this substep's tridiagonal matrix, entries for the
        line neighbour before, the cell, the neighbour after."""
        return self._operators()[0]

    @property
    def Bmat(self):
        """CSR (n, n) on T^n, in this substep's row and column order."""
        return self._operators()[1]

    @property
    def Cmats(self):
        """CSR (n, n) on each earlier substep's result."""
        return self._operators()[2]

    @property
    def Dvec(self):
        """(n,): weight of the volumetric source in each row."""
        return self._operators()[3]

    @property
    def Lmat(self):
        """Thomas factor L of Amat (tridiag.tridiaglu's layout)."""
        if self._lu is None:
            self._lu = tridiag.tridiaglu(self.Amat)
        return self._lu[0]

    @property
    def Umat(self):
        """Thomas factor U of Amat (tridiag.tridiaglu's layout)."""
        if self._lu is None:
            self._lu = tridiag.tridiaglu(self.Amat)
        return self._lu[1]


# array axis (z, y, x) -> the substep whose Crank-Nicolson split makes it
# implicit (adi_expressions: x first, then y, then z)
_SUBSTEP_OF_AXIS = (2, 1, 0)
_PERMUTATIONS = ((0, 1, 2), (2, 0, 1), (1, 2, 0))


def reference_operators(shape, stepnum, M, D, gm, gp):
    """A (n x 3), B (CSR), [C] (CSR), D of Douglas substep `stepnum` in the
    upstream's assembled form, from per-cell equation coefficients.

    M is each row's time-term weight (rho c V/dt, 1 for a FIXED cell), D its
    source weight, gm[a] / gp[a] the weight of its minus / plus face along
    array axis a (z, y, x), zero on outer faces and in FIXED rows.  In the
    upstream, a face term g (T_nb - T_c) of an axis already split by this or
    an earlier substep s' reads g/2 (T_nb,p(s') + T_nb,m - T_c,p(s') - T_c,m)
    (expression.py:753-827); later axes keep g (T_nb - T_c) at the old level
    (expression.py:712-716).  Unknowns of this substep go to A with the sign
    flipped, old-level terms to B, terms in an earlier substep's unknowns to
    C[s'], and the source weight to D (alternatingdirection_c.c:102-199).
    Rows and columns follow the substep's permutation
    (alternatingdirection_c_pyx.pyx:418-445, alternatingdirection_c.c:114).
    The upstream's ADI_steps[s] hold the same operators."""
    import scipy.sparse as sp
    shape = tuple(int(v) for v in shape)
    n = int(np.prod(shape))
    perm = _PERMUTATIONS[stepnum]
    ps = [shape[p] for p in perm]
    row = np.arange(n).reshape(ps).transpose(np.argsort(perm))  # equation of cell (k, j, i)
    M = np.broadcast_to(np.asarray(M, dtype=np.float64), shape)
    D = np.broadcast_to(np.asarray(D, dtype=np.float64), shape)
    A = np.zeros((n, 3))
    Dvec = np.zeros(n)
    r_all = row.ravel()
    A[r_all, 1] = M.ravel()
    Dvec[r_all] = D.ravel()
    B = [(r_all, r_all, M.ravel())]
    C = [[] for _ in range(stepnum)]
    for a in range(3):
        sub = _SUBSTEP_OF_AXIS[a]
        for side, g in ((-1, gm[a]), (1, gp[a])):
            g = np.broadcast_to(np.asarray(g, dtype=np.float64), shape)
            own = [slice(None)] * 3
            nb = [slice(None)] * 3
            own[a] = slice(1, None) if side < 0 else slice(0, -1)
            nb[a] = slice(0, -1) if side < 0 else slice(1, None)
            gf = g[tuple(own)]
            keep = gf != 0.0
            r = row[tuple(own)][keep]
            cn = row[tuple(nb)][keep]
            gf = gf[keep]
            if sub > stepnum:  # not split yet: old level in full
                B += [(r, r, -gf), (r, cn, gf)]
                continue
            h = 0.5 * gf
            B += [(r, r, -h), (r, cn, h)]
            if sub == stepnum:
                A[r, 1] += h
                A[r, 1 + side] = -h
            else:
                C[sub] += [(r, r, -h), (r, cn, h)]

    def csr(parts):
        if not parts:
            return sp.csr_matrix((n, n))
        rows, cols, vals = (np.concatenate(x) for x in zip(*parts))
        return sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()

    return A, csr(B), [csr(c) for c in C], Dvec


def _equation_coefficients(params):
    """Per-cell (M, D, gm[3], gp[3]) of a plan at equation level, read back
    from the device: what each cell's equation holds, multiplied through by
    its volume in curvature mode as the upstream does (crank_nicolson.pyx:
    436-453)."""
    shape = params.shape
    if getattr(params, "curvatures", None) is not None:
        c = [t.cpu().numpy() for t in params.d_coef]
        comp = params.d_w.cpu().numpy() != 0
        vol = np.broadcast_to(params.volume_array, shape)
        M = np.where(comp, c[0] * vol, 1.0)
        D = np.where(comp, vol, 0.0)
        gm = [np.where(comp, c[1 + a] * vol, 0.0) for a in range(3)]
        gp = [np.where(comp, c[4 + a] * vol, 0.0) for a in range(3)]
        return M, D, gm, gp
    f = [torch.zeros(shape, dtype=torch.float64, device=_lib.device()) for _ in range(5)]
    _lib.check("hs2_plan_expand", _lib.load().hs2_plan_expand(
        params.plan, *[_lib.ptr(t) for t in f], _lib.stream_handle()))
    M, w, gz, gy, gx = [t.cpu().numpy() for t in f]
    gm, gp = [], []
    for a, g in enumerate((gz, gy, gx)):
        up = np.zeros_like(g)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[a] = slice(1, None)
        dst[a] = slice(0, -1)
        up[tuple(dst)] = g[tuple(src)]
        gm.append(g * w)
        gp.append(up * w)
    return M, w.astype(np.float64), gm, gp


def _material_table(materials, dt):
    fixed, kvals, M = [], [], []
    for idx, mat in enumerate(materials):
        if mat[0] == TEMPERATURE_COMPUTE:
            (_, k, rho, c) = mat
            fixed.append(False)
            kvals.append(k)
            M.append(rho * c * (1.0 / dt))
        elif mat[0] == TEMPERATURE_FIXED:
            fixed.append(True)
            kvals.append(None)
            M.append(1.0)
        else:
            raise AssertionError("material %d: first member must be TEMPERATURE_COMPUTE or "
                                 "TEMPERATURE_FIXED" % idx)
    return fixed, kvals, M


def _checked_law(q, b, axis):
    """A face law's result must be a number or one of this package's
    expressions; a plugin that built it from another package's expression
    type (an upstream plugin importing the upstream's heatsim2.expression)
    cannot be read."""
    if isinstance(q, (numbers.Number, expression.linear_expression)):
        return q
    raise TypeError(
        "boundary %d (%s faces): the face law returned a %s.%s, not a heatsim2_amd expression; "
        "a plugin written for the upstream package imports heatsim2.expression -- call "
        "heatsim2_amd.alias_as_heatsim2() before importing it"
        % (b, _AXIS_NAME[axis], type(q).__module__, type(q).__name__))


class _FaceLaws(object):
    """Face conductance per volume for (class, axis, k_minus, k_plus)."""

    def __init__(self, boundaries, d):
        self.boundaries = boundaries
        self.d = d  # (dz, dy, dx)
        self._expr = {}
        self._cache = {}

    def expr(self, b, axis):
        key = (b, axis)
        if key not in self._expr:
            bnd = self.boundaries[b]
            plugin, extra = bnd[0], tuple(bnd[1:])
            fn = (plugin.qz, plugin.qy, plugin.qx)[axis]
            KM = expression.linear_expression("kmat_m")
            KP = expression.linear_expression("kmat_p")
            Ts = [expression.linear_expression(n) for n in _T_ARGS[axis]]
            dz, dy, dx = self.d
            self._expr[key] = _checked_law(fn(KM, KP, dz, dy, dx, *Ts, *extra), b, axis)
        return self._expr[key]

    def _alpha(self, expr, b, axis, km, kp):
        """alpha of the flux alpha*(T_plus - T_minus) that expr states"""
        where = "boundary %d (%s faces)" % (b, _AXIS_NAME[axis])
        const, terms = expression.linear_form(expr, {"kmat_m": km, "kmat_p": kp})
        if const != 0.0:
            raise ValueError("%s: the flux has a constant term %g; only fluxes linear in the "
                             "temperatures are supported (the upstream asserts on it)"
                             % (where, const))
        tm, tp = _T_ARGS[axis][0], _T_ARGS[axis][1]
        alpha = beta = 0.0
        for (name, gid), coef in terms.items():
            if gid is None:
                raise ValueError("%s: the term in %s lies outside group(); such terms are "
                                 "integrated explicitly, which this build does not support"
                                 % (where, name))
            if name == tp:
                alpha += coef
            elif name == tm:
                beta += coef
            else:
                raise OffDiagonalConductivityError(
                    "%s: the flux depends on %s (off-diagonal conductivity); that needs the "
                    "unaligned_anisotropic mode, which the upstream cannot run either"
                    % (where, name))
        if abs(alpha + beta) > 1e-12 * max(abs(alpha), abs(beta)):
            raise ValueError("%s: flux %g*T_plus + %g*T_minus is not a two-point difference"
                             % (where, alpha, beta))
        return alpha

    def alpha_in_spacings(self, b, axis, km, kp):
        """(A, B, C) with alpha(dy, dx) = A + B/dx + C/dy for the face law
        evaluated at other y and x spacings (curvature mode substitutes each
        cell's mean spacings, crank_nicolson.pyx:433-434).  Probed at powers
        of two, where the shipped laws (-k/d, -h, 0) are exact, and checked
        at a fourth point; a law of another form is rejected."""
        bnd = self.boundaries[b]
        plugin, extra = bnd[0], tuple(bnd[1:])
        fn = (plugin.qz, plugin.qy, plugin.qx)[axis]
        KM = expression.linear_expression("kmat_m")
        KP = expression.linear_expression("kmat_p")
        Ts = [expression.linear_expression(n) for n in _T_ARGS[axis]]
        dz = self.d[0]

        def at(dy, dx):
            return self._alpha(_checked_law(fn(KM, KP, dz, dy, dx, *Ts, *extra), b, axis), b,
                               axis, km, kp)

        a1, a2, a3 = at(1.0, 1.0), at(2.0, 1.0), at(1.0, 2.0)
        C = 2.0 * (a1 - a2)
        B = 2.0 * (a1 - a3)
        A = a1 - B - C
        a4 = at(4.0, 8.0)
        if abs(A + B / 8.0 + C / 4.0 - a4) > 1e-12 * max(abs(a1), abs(a4), 1e-300):
            raise NotImplementedError(
                "boundary %d (%s faces): curvature mode needs a face law of the form "
                "a + b/dx + c/dy" % (b, _AXIS_NAME[axis]))
        return A, B, C

    def g(self, b, axis, km, kp, kkey):
        key = (b, axis, kkey)
        if key in self._cache:
            return self._cache[key]
        where = "boundary %d (%s faces)" % (b, _AXIS_NAME[axis])
        alpha = self._alpha(self.expr(b, axis), b, axis, km, kp)
        g = -alpha / self.d[axis]
        if g < 0.0:
            raise ValueError("%s: negative face conductance %g" % (where, g))
        self._cache[key] = g
        return g


def _check_map(name, arr, shape):
    if not isinstance(arr, np.ndarray) or arr.dtype != np.uint8 or arr.ndim != 3:
        raise ValueError("Buffer dtype mismatch: %s must be a 3-D uint8 ndarray" % name)
    if tuple(arr.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, arr.shape, tuple(shape)))
    return np.ascontiguousarray(arr)


def _pair_counts(b, mm, mp, nmat):
    key = (b.astype(np.int32) * nmat + mm) * nmat + mp
    return np.bincount(key.ravel(), minlength=1)


def build_tables(d, dt, materials, boundaries, mat, bz, by, bx):
    """Compact tables (gtab[nb,3,nmat,nmat], mtab[nmat], compute[nmat]) for
    the maps, evaluated only for the (class, axis, material pair) that occur.
    Raises like the upstream setup on invalid input."""
    nmat, nb = len(materials), len(boundaries)
    if nmat == 0 or nb == 0:
        raise ValueError("need at least one material and one boundary class")
    if nmat > 256 or nb > 256:
        raise ValueError("at most 256 materials / boundary classes (uint8 maps)")
    for name, a, lim in (("material_elements", mat, nmat), ("boundary_z_elements", bz, nb),
                         ("boundary_y_elements", by, nb), ("boundary_x_elements", bx, nb)):
        if a.size and int(a.max()) >= lim:
            raise IndexError("%s holds index %d but only %d entries are defined"
                             % (name, int(a.max()), lim))
    fixed, kvals, M = _material_table(materials, dt)
    laws = _FaceLaws(boundaries, d)
    gtab = np.zeros((nb, 3, nmat, nmat))
    used = np.bincount(mat.ravel(), minlength=nmat) > 0

    def kk(m):
        return kvals[m]

    faces = (
        (0, bz[1:-1], mat[:-1], mat[1:], (bz[0], mat[0]), (bz[-1], mat[-1])),
        (1, by[:, 1:-1], mat[:, :-1], mat[:, 1:], (by[:, 0], mat[:, 0]), (by[:, -1], mat[:, -1])),
        (2, bx[:, :, 1:-1], mat[:, :, :-1], mat[:, :, 1:], (bx[:, :, 0], mat[:, :, 0]),
         (bx[:, :, -1], mat[:, :, -1])),
    )
    for axis, bint, mm, mp, outer0, outer1 in faces:
        if bint.size:
            counts = _pair_counts(bint, mm, mp, nmat)
            for key in np.nonzero(counts)[0]:
                b, rem = divmod(int(key), nmat * nmat)
                ma, mb = divmod(rem, nmat)
                if fixed[ma] and fixed[mb]:
                    continue
                km_i = ma if not fixed[ma] else mb
                kp_i = mb if not fixed[mb] else ma
                gtab[b, axis, ma, mb] = laws.g(b, axis, kk(km_i), kk(kp_i), (km_i, kp_i))
        for bo, mo in (outer0, outer1):
            counts = np.bincount((bo.astype(np.int32) * nmat + mo).ravel(), minlength=1)
            for key in np.nonzero(counts)[0]:
                b, m = divmod(int(key), nmat)
                if fixed[m]:
                    continue
                if laws.g(b, axis, kk(m), kk(m), (m, m)) != 0.0:
                    raise ValueError(
                        "Equation exceeds bounds of domain: boundary class %d on an outer %s "
                        "face of a COMPUTE cell carries flux. Are external boundaries set "
                        "correctly? (outer faces must be insulating)" % (b, _AXIS_NAME[axis]))
    mtab = np.array([M[m] if (used[m] or fixed[m]) else 1.0 for m in range(nmat)], dtype=np.float64)
    for m in range(nmat):
        if used[m] and not fixed[m] and not (mtab[m] > 0.0 and np.isfinite(mtab[m])):
            raise ValueError("material %d: rho*c/dt = %g must be positive" % (m, mtab[m]))
    compute = np.array([0 if f else 1 for f in fixed], dtype=np.uint8)
    return gtab, mtab, compute


def curved_geometry(d, nz, curv_y, curv_x):
    """Edge lengths and volumes of curvature mode (crank_nicolson.pyx:
    409-424): under a top surface of curvature kappa, an edge k*dz deep is
    (1 + k*kappa*dz) times its nominal length; a cell takes the mean of its
    top and bottom edges, and the volume |dx_mean * dy_mean * dz|."""
    dz, dy, dx = d
    k = np.arange(nz, dtype=np.float64)[:, None, None]
    cy = np.asarray(curv_y, dtype=np.float64)[None]
    cx = np.asarray(curv_x, dtype=np.float64)[None]
    geo = dict(dy_top=(1.0 + k * cy * dz) * dy, dy_bot=(1.0 + (k + 1.0) * cy * dz) * dy,
               dx_top=(1.0 + k * cx * dz) * dx, dx_bot=(1.0 + (k + 1.0) * cx * dz) * dx)
    for key, val in geo.items():
        if not (val > 0.0).all():
            raise ValueError("curvature mode: %s edges are not positive everywhere (a "
                             "curvature of magnitude >= 1/depth)" % key.split("_")[0][1:])
    geo["dy_mean"] = (geo["dy_top"] + geo["dy_bot"]) / 2.0
    geo["dx_mean"] = (geo["dx_top"] + geo["dx_bot"]) / 2.0
    geo["vol"] = np.abs(geo["dx_mean"] * geo["dy_mean"] * dz)
    return geo


def curved_coefficients(d, materials, boundaries, mat, bz, by, bx, mtab, compute, geo):
    """Per-cell coefficients of curvature mode, per unit volume.

    Each cell's balance (crank_nicolson.pyx:436-440) weighs a face by the
    cell's own area -- dx_top*dy_top and dx_bot*dy_bot for the z faces,
    dx_mean*dz for y, dy_mean*dz for x -- and takes the face law with its own
    mean spacings; divided by the volume it reads
      sum_faces G_f (T_nb - T_c),  G_f = -alpha_f * area_f / V.
    Two cells meet a y or x face with different G, so every cell gets its own
    minus- and plus-face values.  The neighbour-k fallback is the uncurved
    one.  Returns M, w, gm[3], gp[3] as (nz, ny, nx) arrays."""
    nmat = len(materials)
    fixed, kvals, _ = _material_table(materials, 1.0)
    fixed = np.array(fixed)
    laws = _FaceLaws(boundaries, d)
    comp = compute[mat].astype(bool)
    shape = mat.shape
    dz = d[0]
    area = ((geo["dx_top"] * geo["dy_top"], geo["dx_bot"] * geo["dy_bot"]),
            (geo["dx_mean"] * dz,) * 2, (geo["dy_mean"] * dz,) * 2)
    faces = ((bz[:-1], bz[1:]), (by[:, :-1], by[:, 1:]), (bx[:, :, :-1], bx[:, :, 1:]))
    gm, gp = [], []
    for axis in range(3):
        lo = np.roll(mat, 1, axis=axis)
        hi = np.roll(mat, -1, axis=axis)
        idx = np.arange(shape[axis]).reshape([-1 if a == axis else 1 for a in range(3)])
        lo = np.where((idx > 0) & ~fixed[lo], lo, mat)
        hi = np.where((idx < shape[axis] - 1) & ~fixed[hi], hi, mat)
        out = []
        for side, (b, ma, mb) in enumerate(((faces[axis][0], lo, mat), (faces[axis][1], mat, hi))):
            key = (b.astype(np.int64) * nmat + ma) * nmat + mb
            ukeys, inv = np.unique(np.where(comp, key, -1), return_inverse=True)
            coef = np.zeros((len(ukeys), 3))
            for t, kk in enumerate(ukeys):
                if kk < 0:
                    continue
                bb, rem = divmod(int(kk), nmat * nmat)
                m1, m2 = divmod(rem, nmat)
                coef[t] = laws.alpha_in_spacings(bb, axis, kvals[m1], kvals[m2])
            c = coef[inv.reshape(shape)]
            alpha = c[..., 0] + c[..., 1] / geo["dx_mean"] + c[..., 2] / geo["dy_mean"]
            G = np.where(comp, -alpha * area[axis][side] / geo["vol"], 0.0)
            if (G < 0.0).any():
                raise ValueError("curvature mode: negative %s face conductance" % _AXIS_NAME[axis])
            out.append(G)
        gm.append(out[0])
        gp.append(out[1])
    M = mtab[mat]
    w = compute[mat].astype(np.uint8)
    return M, w, gm, gp


def setup(z0, y0, x0,
          dz, dy, dx,
          nz, ny, nx,
          dt,
          materials,
          boundaries,
          volumetric,
          material_elements,
          boundary_z_elements,
          boundary_y_elements,
          boundary_x_elements,
          volumetric_elements,
          top_surface_y_curvatures=None,
          top_surface_x_curvatures=None,
          unaligned_anisotropic=False):
    """Prepare the ADI step for a grid; returns (ADI_params, ADI_steps).

    Same arguments as the upstream setup (crank_nicolson.pyx:128-142).
    z0, y0, x0 are accepted and unused, as upstream.  With both top-surface
    curvature arrays (ny, nx) given, the curvature mode of
    crank_nicolson.pyx:388-457 (per-cell areas and volumes; the volumes also
    scale point sources).  unaligned_anisotropic=True raises
    UnalignedAnisotropicError, an AssertionError as upstream, whose setup
    cannot complete in that mode.
    """
    curved = top_surface_y_curvatures is not None or top_surface_x_curvatures is not None
    if curved and (top_surface_y_curvatures is None or top_surface_x_curvatures is None):
        raise TypeError("curvature mode needs both top_surface_y_curvatures and "
                        "top_surface_x_curvatures (the upstream indexes both)")
    if unaligned_anisotropic:
        raise UnalignedAnisotropicError(
            "unaligned_anisotropic mode: the upstream builds 3 ADI steps for 15 substep "
            "expressions and fails its assert at the first cell; there is no result to "
            "reproduce")
    for v in (nz, ny, nx):
        if not isinstance(v, (numbers.Integral, np.integer)) or v < 1:
            raise ValueError("nz, ny, nx must be positive integers")
    nz, ny, nx = int(nz), int(ny), int(nx)
    mat = _check_map("material_elements", material_elements, (nz, ny, nx))
    bz = _check_map("boundary_z_elements", boundary_z_elements, (nz + 1, ny, nx))
    by = _check_map("boundary_y_elements", boundary_y_elements, (nz, ny + 1, nx))
    bx = _check_map("boundary_x_elements", boundary_x_elements, (nz, ny, nx + 1))
    _check_map("volumetric_elements", volumetric_elements, (nz, ny, nx))
    d = (float(dz), float(dy), float(dx))
    gtab, mtab, compute = build_tables(d, float(dt), materials, boundaries, mat, bz, by, bx)
    nmat, nb = len(materials), len(boundaries)
    if curved:
        return _setup_curved(d, float(dt), (nz, ny, nx), materials, boundaries, mat, bz, by, bx,
                             mtab, compute, top_surface_y_curvatures, top_surface_x_curvatures)
    L = _lib.load()
    dev = _lib.device()
    params = adi_params(shape=(nz, ny, nx), volume_array=dz * dy * dx, d=d, dt=float(dt),
                        nmat=nmat, nb=nb, gtab=gtab, mtab=mtab, compute=compute)
    params.d_mat = torch.from_numpy(mat).to(dev)
    params.d_bz = torch.from_numpy(bz).to(dev)
    params.d_by = torch.from_numpy(by).to(dev)
    params.d_bx = torch.from_numpy(bx).to(dev)
    import ctypes
    grid = _lib.GridDesc(nz, ny, nx, d[0], d[1], d[2], float(dt))
    h = ctypes.c_void_p()
    _lib.check("hs2_plan_create", L.hs2_plan_create(ctypes.byref(grid), ctypes.byref(h)))
    params._plan = h
    gt = np.ascontiguousarray(gtab)
    _lib.check("hs2_plan_set_compact",
               L.hs2_plan_set_compact(h, _lib.ptr(params.d_mat), _lib.ptr(params.d_bz),
                                      _lib.ptr(params.d_by), _lib.ptr(params.d_bx), nmat, nb,
                                      _lib.host_ptr(gt), _lib.host_ptr(mtab),
                                      _lib.host_ptr(compute)))
    torch.cuda.synchronize()
    steps = [adi_step(params, (0, 1, 2), 0), adi_step(params, (2, 0, 1), 1),
             adi_step(params, (1, 2, 0), 2)]
    return params, steps


def _setup_curved(d, dt, shape, materials, boundaries, mat, bz, by, bx, mtab, compute, cy, cx):
    nz, ny, nx = shape
    cy = np.asarray(cy, dtype=np.float64)
    cx = np.asarray(cx, dtype=np.float64)
    for name, c in (("top_surface_y_curvatures", cy), ("top_surface_x_curvatures", cx)):
        if c.shape != (ny, nx):
            raise ValueError("%s has shape %s, expected %s" % (name, c.shape, (ny, nx)))
        if not np.isfinite(c).all():
            raise ValueError("%s must be finite" % name)
    geo = curved_geometry(d, nz, cy, cx)
    M, w, gm, gp = curved_coefficients(d, materials, boundaries, mat, bz, by, bx, mtab, compute,
                                       geo)
    L = _lib.load()
    dev = _lib.device()
    params = adi_params(shape=(nz, ny, nx), volume_array=np.ascontiguousarray(geo["vol"]),
                        d=d, dt=dt, nmat=len(materials), nb=len(boundaries), mtab=mtab,
                        compute=compute, curvatures=(cy.copy(), cx.copy()))
    params.d_coef = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
                     for a in [M] + gm + gp]
    params.d_w = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    import ctypes
    grid = _lib.GridDesc(nz, ny, nx, d[0], d[1], d[2], dt)
    h = ctypes.c_void_p()
    _lib.check("hs2_plan_create", L.hs2_plan_create(ctypes.byref(grid), ctypes.byref(h)))
    params._plan = h
    c = params.d_coef
    _lib.check("hs2_plan_set_dense_asym",
               L.hs2_plan_set_dense_asym(h, _lib.ptr(c[0]), _lib.ptr(params.d_w), _lib.ptr(c[1]),
                                         _lib.ptr(c[2]), _lib.ptr(c[3]), _lib.ptr(c[4]),
                                         _lib.ptr(c[5]), _lib.ptr(c[6])))
    torch.cuda.synchronize()
    steps = [adi_step(params, (0, 1, 2), 0), adi_step(params, (2, 0, 1), 1),
             adi_step(params, (1, 2, 0), 2)]
    return params, steps

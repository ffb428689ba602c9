"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the checker for the HIP path: tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py use it.  The product package
(heatsim2_amd) never imports it.

* liboracle.so (hs2_oracle.c): per-cell coefficient assembly
  (crank_nicolson.pyx:272-491), the reference-structured step
  (alternatingdirection_c_pyx.pyx:388-416 with tridiag.pyx:10-69) and a
  delta-form OpenMP step.
* numpy restatements here: volumetric source evaluation
  (alternatingdirection_c_pyx.pyx:294-386) and the surface readout
  (surface_temperature.py:31-36).

Problem descriptions are the plain-data cases of tests/golden/cases.py
(kind codes for boundaries, (fixed,kz,ky,kx,rho,c,tensor) materials).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BND_CONDUCTING, BND_ANISO, BND_INSULATING, BND_THIN = 0, 1, 2, 3
NO_SOURCE, IMPULSE_SOURCE, STEPPED_SOURCE, IMPULSE_POINT_SOURCE_JOULES, \
    SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE = 0, 1, 2, 3, 4

ERRORS = {1: "material index out of range", 2: "boundary index out of range",
          3: "conductivity type does not match the boundary law",
          4: "non-insulating outer face on a COMPUTE cell", 5: "allocation failed",
          6: "bad argument"}


class _Material(ctypes.Structure):
    _fields_ = [("fixed", ctypes.c_int), ("tensor", ctypes.c_int), ("k", ctypes.c_double * 3),
                ("rho", ctypes.c_double), ("c", ctypes.c_double)]


class _Boundary(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("h", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        i64 = ctypes.POINTER(ctypes.c_int64)
        I, D = ctypes.c_int, ctypes.c_double
        L.orc_assemble_cells.argtypes = [I, I, I, D, D, D, D, ctypes.POINTER(_Material), I,
                                         ctypes.POINTER(_Boundary), I, u8, u8, u8, u8, dp]
        L.orc_assemble_cells.restype = I
        L.orc_ref_create.argtypes = [I, I, I, dp]
        L.orc_ref_create.restype = P
        L.orc_ref_destroy.argtypes = [P]
        L.orc_ref_step.argtypes = [P, dp, dp, dp]
        L.orc_ref_export_A.argtypes = [P, I, dp, dp]
        L.orc_ref_export_A.restype = I
        L.orc_ref_export_csr.argtypes = [P, I, I, i64, i64, dp]
        L.orc_ref_export_csr.restype = ctypes.c_int64
        L.orc_fast_step.argtypes = [I, I, I, dp, dp, dp, dp, dp, I]
        L.orc_fast_run.argtypes = [I, I, I, dp, dp, I, dp, I, D, dp, I]
        L.orc_thomas.argtypes = [I, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


def assemble(shape, d, dt, materials, boundaries, maps):
    """Per-cell coefficients (n x 8: M, D, gzm, gzp, gym, gyp, gxm, gxp).
    Restates crank_nicolson.pyx:272-491 (in hs2_oracle.c)."""
    nz, ny, nx = shape
    m, bz, by, bx = [np.ascontiguousarray(a, dtype=np.uint8) for a in maps[:4]]
    mats = (_Material * len(materials))()
    for t, (fixed, kz, ky, kx, rho, c, tensor) in enumerate(materials):
        mats[t].fixed, mats[t].tensor = int(fixed), int(tensor)
        mats[t].k[0], mats[t].k[1], mats[t].k[2] = kz, ky, kx
        mats[t].rho, mats[t].c = rho, c
    bnds = (_Boundary * len(boundaries))()
    for t, (kind, h) in enumerate(boundaries):
        bnds[t].kind, bnds[t].h = int(kind), float(h)
    coef = np.zeros((nz * ny * nx, 8))
    rc = lib().orc_assemble_cells(nz, ny, nx, d[0], d[1], d[2], dt, mats, len(materials), bnds,
                                  len(boundaries), _p(m, ctypes.c_uint8), _p(bz, ctypes.c_uint8),
                                  _p(by, ctypes.c_uint8), _p(bx, ctypes.c_uint8), _p(coef))
    if rc:
        raise ValueError("oracle assembly: " + ERRORS.get(rc, str(rc)))
    return coef


def _curved_geometry(shape, d, curv_y, curv_x):
    """Cell edge lengths and volumes of curvature mode, crank_nicolson.pyx:
    409-424: with top-surface curvature kappa, an edge at depth k*dz is
    (1 + k*kappa*dz) times its nominal length; a cell uses the mean of its top
    and bottom edges and the volume |dx_mean * dy_mean * dz|."""
    nz = shape[0]
    dz, dy, dx = d
    k = np.arange(nz, dtype=np.float64)[:, None, None]
    cy = np.asarray(curv_y, dtype=np.float64)[None]
    cx = np.asarray(curv_x, dtype=np.float64)[None]
    dy_top = (1.0 + k * cy * dz) * dy
    dy_bot = (1.0 + (k + 1.0) * cy * dz) * dy
    dx_top = (1.0 + k * cx * dz) * dx
    dx_bot = (1.0 + (k + 1.0) * cx * dz) * dx
    dy_mean = (dy_top + dy_bot) / 2.0
    dx_mean = (dx_top + dx_bot) / 2.0
    vol = np.abs(dx_mean * dy_mean * dz)
    return dict(dy_top=dy_top, dy_bot=dy_bot, dx_top=dx_top, dx_bot=dx_bot, dy_mean=dy_mean,
                dx_mean=dx_mean, vol=vol)


def assemble_curved(shape, d, dt, materials, boundaries, maps, curv_y, curv_x):
    """Per-cell coefficients of curvature mode (n x 8: M, D, gzm, gzp, gym,
    gyp, gxm, gxp) and the cell volumes, multiplied through by each cell's
    volume as the upstream writes them (crank_nicolson.pyx:388-457):

      heat flow = qz_m*dx_top*dy_top - qz_p*dx_bot*dy_bot
                  + (qy_m - qy_p)*dx_mean*dz + (qx_m - qx_p)*dy_mean*dz + S*V,
      time term = rho*c*V/dt,

    the face laws evaluated with the cell's own dx_mean, dy_mean in place of
    dx, dy (:433-434) and the neighbour-k fallback of the uncurved path.
    A cell's faces use its own geometry, so the coupling of two cells across
    a y or x face differs from each side.  The uncurved assembly (in C)
    validates the maps and face laws first."""
    nz, ny, nx = shape
    dz, dy, dx = d
    assemble(shape, d, dt, materials, boundaries, maps)  # validation only
    m, bz, by, bx = [np.asarray(a, dtype=np.int64) for a in maps[:4]]
    geo = _curved_geometry(shape, d, curv_y, curv_x)
    fixed = np.array([bool(mt[0]) for mt in materials])
    kvec = np.array([[mt[1], mt[2], mt[3]] for mt in materials], dtype=np.float64)
    rhoc = np.array([mt[4] * mt[5] for mt in materials], dtype=np.float64)
    kind = np.array([b[0] for b in boundaries])
    hval = np.array([b[1] for b in boundaries], dtype=np.float64)
    comp = ~fixed[m]

    def alpha(b, mm, mp, axis, dlen):
        """flux law q = alpha*(T_p - T_m) of face class b between the
        (fallback) materials mm, mp, with spacing dlen"""
        a = np.zeros(b.shape)
        cond = (kind[b] == BND_CONDUCTING) | (kind[b] == BND_ANISO)
        thin = kind[b] == BND_THIN
        kmean = (kvec[mm, axis] + kvec[mp, axis]) * 0.5
        a[cond] = -(kmean * (1.0 / dlen))[cond]
        a[thin] = -hval[b][thin]
        return a

    def faces(axis, bmin, bplus, area_m, area_p, dlen):
        # neighbour materials with the fallback to the cell's own
        lo = np.roll(m, 1, axis=axis)
        hi = np.roll(m, -1, axis=axis)
        idx = np.arange(shape[axis]).reshape([-1 if a == axis else 1 for a in range(3)])
        lo_ok = (idx > 0) & ~fixed[lo]
        hi_ok = (idx < shape[axis] - 1) & ~fixed[hi]
        mm = np.where(lo_ok, lo, m)
        mp = np.where(hi_ok, hi, m)
        gm = -alpha(bmin, mm, m, axis, dlen) * area_m
        gp = -alpha(bplus, m, mp, axis, dlen) * area_p
        return np.where(comp, gm, 0.0), np.where(comp, gp, 0.0)

    dzv = np.full(shape, dz)
    gzm, gzp = faces(0, bz[:-1], bz[1:], geo["dx_top"] * geo["dy_top"],
                     geo["dx_bot"] * geo["dy_bot"], dzv)
    gym, gyp = faces(1, by[:, :-1], by[:, 1:], geo["dx_mean"] * dz, geo["dx_mean"] * dz,
                     geo["dy_mean"])
    gxm, gxp = faces(2, bx[:, :, :-1], bx[:, :, 1:], geo["dy_mean"] * dz, geo["dy_mean"] * dz,
                     geo["dx_mean"])
    vol = np.broadcast_to(geo["vol"], shape)
    M = np.where(comp, rhoc[m] * vol * (1.0 / dt), 1.0)
    D = np.where(comp, vol, 0.0)
    coef = np.stack([M, D, gzm, gzp, gym, gyp, gxm, gxp], axis=-1).reshape(-1, 8)
    return np.ascontiguousarray(coef), np.ascontiguousarray(vol)


def case_volume(case):
    """The cells' volumes: dz*dy*dx, or per cell in curvature mode."""
    if case.get("curvature") is not None:
        return _curved_geometry(case["shape"], case["d"], *case["curvature"])["vol"] * \
            np.ones(case["shape"])
    return case["d"][0] * case["d"][1] * case["d"][2]


def assemble_case(case):
    if case.get("curvature") is not None:
        return assemble_curved(case["shape"], case["d"], case["dt"], case["materials"],
                               case["boundaries"], case["maps"], *case["curvature"])[0]
    return assemble(case["shape"], case["d"], case["dt"], case["materials"], case["boundaries"],
                    case["maps"])


class RefPlan:
    """Reference-structured operators (permuted CSR + global Thomas LU)."""

    def __init__(self, shape, coef):
        self.shape = tuple(shape)
        self.n = int(np.prod(shape))
        self._coef = np.ascontiguousarray(coef)
        self.h = lib().orc_ref_create(shape[0], shape[1], shape[2], _p(self._coef))
        if not self.h:
            raise MemoryError("orc_ref_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ref_destroy(self.h)
            self.h = None

    def step(self, T, S=None):
        T = np.ascontiguousarray(T, dtype=np.float64)
        out = np.empty(self.shape)
        Sc = None if S is None else np.ascontiguousarray(S, dtype=np.float64)
        lib().orc_ref_step(self.h, _p(T), _p(Sc), _p(out))
        return out

    def A(self, s):
        A = np.empty((self.n, 3))
        D = np.empty(self.n)
        lib().orc_ref_export_A(self.h, s, _p(A), _p(D))
        return A, D

    def csr(self, s, which):
        nnz = lib().orc_ref_export_csr(self.h, s, which, None, None, None)
        indptr = np.empty(self.n + 1, np.int64)
        indices = np.empty(nnz, np.int64)
        data = np.empty(nnz)
        lib().orc_ref_export_csr(self.h, s, which, _p(indptr, ctypes.c_int64),
                                 _p(indices, ctypes.c_int64), _p(data))
        return data, indices, indptr


def fast_step(coef, T, S=None, nthreads=0):
    shape = T.shape
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.empty(shape)
    work = np.empty(2 * T.size)
    Sc = None if S is None else np.ascontiguousarray(S, dtype=np.float64)
    lib().orc_fast_step(shape[0], shape[1], shape[2], _p(np.ascontiguousarray(coef)), _p(T),
                        _p(Sc), _p(out), _p(work), int(nthreads))
    return out


def fast_run(coef, T0, nsteps, S_first=None, src_steps=1, dz=None, nthreads=0):
    """nsteps delta-form steps; S_first applied on the first src_steps steps.
    Returns (T_final, surface_history or None)."""
    nz, ny, nx = T0.shape
    T = np.array(T0, dtype=np.float64, order="C")
    hist = np.empty((nsteps, ny, nx)) if dz is not None else None
    Sc = None if S_first is None else np.ascontiguousarray(S_first, dtype=np.float64)
    lib().orc_fast_run(nz, ny, nx, _p(np.ascontiguousarray(coef)), _p(T), int(nsteps), _p(Sc),
                       int(src_steps), float(dz or 0.0), _p(hist), int(nthreads))
    return T, hist


def thomas(a, b, c, r):
    n = len(b)
    x = np.empty(n)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r)]
    lib().orc_thomas(n, *[_p(v) for v in arrs], _p(x))
    return x


def source_field(volumetric, vmap, t, dt, volume, zgrid=None):
    """Volumetric power density S (W/m^3) at call time t.
    Restates alternatingdirection_c_pyx.pyx:294-386 (volume: the scalar
    dz*dy*dx of the uncurved grid, crank_nicolson.pyx:261, or the per-cell
    volumes of curvature mode, :310-316)."""
    S = np.zeros(vmap.shape)
    for idx, entry in enumerate(volumetric):
        kind = entry[0]
        sel = vmap == idx
        if kind == IMPULSE_SOURCE:
            if t == entry[1]:
                S[sel] = entry[2] / dt
        elif kind == STEPPED_SOURCE:
            _, t0, t1, w = entry
            if t >= t0 and t <= t1:
                S[sel] = w
        elif kind == IMPULSE_POINT_SOURCE_JOULES:
            if t == entry[1]:
                if np.ndim(volume):
                    S[sel] = entry[2] / (np.asarray(volume)[sel] * dt)
                else:
                    S[sel] = entry[2] / (volume * dt)
        elif kind == SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE:
            (_, t_imp, direc, offset, zg, ddz, jpm2, ell) = entry
            if t == t_imp:
                zg = zgrid if isinstance(zg, str) else zg
                direc = np.asarray(direc, dtype=float)
                centre = zg * direc[0] - float(offset)
                left = centre - abs(float(ddz)) / 2.0
                right = centre + abs(float(ddz)) / 2.0
                use = sel & (right > 0.0)
                r_ = right[use]
                l_ = left[use]
                l_[l_ < 0.0] = 0.0
                zint = -np.exp(-r_ / float(ell)) + np.exp(-l_ / float(ell))
                S[use] = zint * float(jpm2) / (abs(float(ddz)) * dt)
    return S


def surface_readout(T, dz):
    """surface_temperature.py:31-36 restated."""
    a = (T[1] - T[0]) / (2.0 * dz ** 2.0)
    c = T[0] - .25 * a * dz ** 2.0
    dT = T[1] - T[0]
    lin = T[0] - dT / 2.0
    return (lin + c) / 2.0


def zgrid(shape, dz):
    nz, ny, nx = shape
    z = (np.arange(nz) + 0.5) * dz
    return np.broadcast_to(z[:, None, None], shape).copy()


def run_ref(case, nsteps=None, structured=True):
    """Trajectory of a case through the reference-structured (or delta-form)
    step, with sources evaluated at t = t0 + n*dt.  Returns list of fields."""
    coef = assemble_case(case)
    shape = case["shape"]
    vol = case_volume(case)
    zg = zgrid(shape, case["d"][0])
    plan = RefPlan(shape, coef) if structured else None
    T = np.array(case["T0"], dtype=np.float64)
    out = []
    for n in range(case["nsteps"] if nsteps is None else nsteps):
        t = case["t0"] + n * case["dt"]
        S = source_field(case["volumetric"], case["maps"][4], t, case["dt"], vol, zg)
        T = plan.step(T, S) if structured else fast_step(coef, T, S)
        out.append(T)
    return out

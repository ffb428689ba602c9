"""Problem definitions shared by the fixture generator and the tests.

Each case is plain data (numbers, uint8 maps, kind codes), so the same case
can be fed to the upstream solver (fixture generation, this container only),
to the C oracle and to the HIP path.  Boundary kinds are codes, turned into
modules by whichever package runs the case:

  BND_CONDUCTING=0, BND_ANISO=1, BND_INSULATING=2, BND_THIN=3 (param h)

Materials: (fixed, kz, ky, kx, rho, c, is_tensor).  is_tensor=1 means k is
given as a diagonal 3x3 matrix in (z,y,x) order (for the anisotropic
boundary); 0 means a scalar k (kz=ky=kx).

Volumetric: tuples in the upstream layout, with the z-decay entry's z grid
given as the string "zgrid" (replaced by the cell-centre z array).
"""
import numpy as np

BND_CONDUCTING, BND_ANISO, BND_INSULATING, BND_THIN = 0, 1, 2, 3

NO_SOURCE, IMPULSE_SOURCE, STEPPED_SOURCE, IMPULSE_POINT_SOURCE_JOULES, \
    SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE = 0, 1, 2, 3, 4


def _zero(nz, ny, nx):
    return (np.zeros((nz, ny, nx), np.uint8), np.zeros((nz + 1, ny, nx), np.uint8),
            np.zeros((nz, ny + 1, nx), np.uint8), np.zeros((nz, ny, nx + 1), np.uint8),
            np.zeros((nz, ny, nx), np.uint8))


def _outer(bz, by, bx, cls):
    bx[:, :, 0] = cls
    bx[:, :, -1] = cls
    by[:, 0, :] = cls
    by[:, -1, :] = cls
    bz[0] = cls
    bz[-1] = cls


def case_tiny_mixed():
    """Three materials (two conducting + FIXED), conducting / insulating /
    thin-layer faces, impulse + stepped + point sources."""
    nz, ny, nx = 4, 5, 6
    d = (4e-4, 5e-4, 6e-4)
    dt = 2e-3
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    m[2:, :, :] = 1            # lower layers: second material
    m[3, 0:2, 0:2] = 2         # FIXED corner block
    m[1, 4, 5] = 2             # isolated FIXED cell touching the outer faces
    bz[2, 1:4, 1:5] = 2        # thin layer patch between layers 1 and 2
    bx[0, 2, 3] = 2            # one thin x-face
    by[3, 2, 1:3] = 2          # thin y-faces
    v[0] = 1
    v[1, 2, 2] = 2
    v[2, 1:3, 3:5] = 3
    materials = [(0, 50.0, 50.0, 50.0, 7850.0, 470.0, 0),
                 (0, 0.5, 0.5, 0.5, 1200.0, 1500.0, 0),
                 (1, 0, 0, 0, 0, 0, 0)]
    boundaries = [(BND_CONDUCTING, 0.0), (BND_INSULATING, 0.0), (BND_THIN, 2000.0)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 1e4 / d[0]),
                  (IMPULSE_POINT_SOURCE_JOULES, 3 * dt, 0.02),
                  (STEPPED_SOURCE, 2 * dt, 5 * dt, 3e7)]
    T0 = 293.0 + 0.01 * np.arange(nz * ny * nx, dtype=np.float64).reshape(nz, ny, nx)
    T0[m == 2] = 300.0
    return dict(name="tiny_mixed", shape=(nz, ny, nx), d=d, dt=dt, t0=0.0, nsteps=12,
                materials=materials, boundaries=boundaries, volumetric=volumetric,
                maps=(m, bz, by, bx, v), T0=T0, matrices=True)


def case_tiny_curved(seed=13):
    """Curvature mode (crank_nicolson.pyx:388-457): the tiny_mixed layout
    (two conducting materials + FIXED, thin z, y and x faces) with seeded
    top-surface curvatures of both signs in y and x, up to +-60 /m so that the
    edges change by up to a fifth over the depth, and impulse, stepped and
    point sources (the point source uses the per-cell volumes)."""
    c = case_tiny_mixed()
    ny, nx = c["shape"][1], c["shape"][2]
    rng = np.random.default_rng(seed)
    cy = rng.uniform(-60.0, 60.0, (ny, nx))
    cx = rng.uniform(-60.0, 60.0, (ny, nx))
    cy[0, 0] = cx[0, 0] = 0.0
    c.update(name="tiny_curved", curvature=(cy, cx))
    return c


def case_tiny_aniso():
    """Diagonal-tensor k with the anisotropic face law, thin layers, FIXED
    cells, a scalar-k material on isotropic faces elsewhere is not mixed in
    (the upstream rejects scalar k on anisotropic faces)."""
    nz, ny, nx = 5, 6, 4
    d = (3e-4, 5e-4, 4e-4)
    dt = 1e-3
    m, bz, by, bx, v = _zero(nz, ny, nx)
    bz[:], by[:], bx[:] = 0, 0, 0
    _outer(bz, by, bx, 1)
    m[3:, :, :] = 1
    m[4, 5, 0:2] = 2
    bz[2, 2:5, 1:3] = 2
    bx[1, 1:4, 2] = 2
    v[0] = 1
    materials = [(0, 0.6, 5.0, 3.0, 1550.0, 1200.0, 1),
                 (0, 20.0, 15.0, 25.0, 2700.0, 900.0, 1),
                 (1, 0, 0, 0, 0, 0, 0)]
    boundaries = [(BND_ANISO, 0.0), (BND_INSULATING, 0.0), (BND_THIN, 500.0)]
    volumetric = [(NO_SOURCE,),
                  (SPATIALLY_Z_DECAYING_TEMPORAL_IMPULSE, 0.0, (1.0, 0.0, 0.0), 0.0,
                   "zgrid", d[0], 1e4, 4e-4)]
    T0 = np.full((nz, ny, nx), 293.0)
    T0[m == 2] = 280.0
    return dict(name="tiny_aniso", shape=(nz, ny, nx), d=d, dt=dt, t0=0.0, nsteps=10,
                materials=materials, boundaries=boundaries, volumetric=volumetric,
                maps=(m, bz, by, bx, v), T0=T0, matrices=True)


def case_steel_cube(n=40, nsteps=200):
    """40x40x40 steel cube, insulating faces, 1e4 J/m^2 flash into the top
    z-layer at t=0, dt=1 ms, 200 steps (the CPU reference workload)."""
    nz = ny = nx = n
    d = (5e-4, 5e-4, 5e-4)
    dt = 1e-3
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    v[0] = 1
    materials = [(0, 50.0, 50.0, 50.0, 7850.0, 470.0, 0)]
    boundaries = [(BND_CONDUCTING, 0.0), (BND_INSULATING, 0.0)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 1e4 / d[0])]
    T0 = np.full((nz, ny, nx), 293.0)
    return dict(name="steel_cube%d" % n, shape=(nz, ny, nx), d=d, dt=dt, t0=0.0,
                nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=T0, matrices=False)


def case_two_layer(nz=256, ny=256, nx=256, nsteps=1000, steel_layers=32, seed=1,
                   h=250.0, dxyz=5e-4, dt=5e-4):
    """Steel layer (first `steel_layers` z-cells) over foam, thin insulating
    layer between them, insulating sides, flash 1e4*(1+0.05*U(-1,1)) J/m^2."""
    rng = np.random.default_rng(seed)
    E = 1e4 * (1.0 + 0.05 * rng.uniform(-1.0, 1.0))
    d = (dxyz, dxyz, dxyz)
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    m[steel_layers:] = 1
    bz[steel_layers] = 2
    v[0] = 1
    materials = [(0, 50.0, 50.0, 50.0, 7850.0, 470.0, 0),
                 (0, 0.03, 0.03, 0.03, 30.0, 1500.0, 0)]
    boundaries = [(BND_CONDUCTING, 0.0), (BND_INSULATING, 0.0), (BND_THIN, h)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, E / d[0])]
    T0 = np.full((nz, ny, nx), 293.0)
    return dict(name="two_layer_%dx%dx%d" % (nz, ny, nx), shape=(nz, ny, nx), d=d, dt=dt,
                t0=0.0, nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=T0, matrices=False,
                flash_J_m2=E)


def case_laminate(nz=128, ny=512, nx=512, nsteps=2000, seed=2, h=100.0, dxyz=5e-4,
                  dt=2e-3, patch_min=10, patch_max=80):
    """Anisotropic laminate kz=0.6, ky=kx=5, rho=1550, c=1200 on anisotropic
    faces, 1..8 rectangular delaminations as thin layers at random depths,
    flash 1e4 J/m^2 on top."""
    rng = np.random.default_rng(seed)
    d = (dxyz, dxyz, dxyz)
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    npatch = int(rng.integers(1, 9))
    for _ in range(npatch):
        sy = int(rng.integers(patch_min, min(patch_max, ny) + 1))
        sx = int(rng.integers(patch_min, min(patch_max, nx) + 1))
        y0 = int(rng.integers(0, ny - sy + 1))
        x0 = int(rng.integers(0, nx - sx + 1))
        kz = int(rng.integers(1, nz))
        bz[kz, y0:y0 + sy, x0:x0 + sx] = 2
    v[0] = 1
    materials = [(0, 0.6, 5.0, 5.0, 1550.0, 1200.0, 1)]
    boundaries = [(BND_ANISO, 0.0), (BND_INSULATING, 0.0), (BND_THIN, h)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 1e4 / d[0])]
    T0 = np.full((nz, ny, nx), 293.0)
    return dict(name="laminate_%dx%dx%d" % (nz, ny, nx), shape=(nz, ny, nx), d=d, dt=dt,
                t0=0.0, nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=T0, matrices=False,
                npatch=npatch)


def zgrid_for(case):
    nz, ny, nx = case["shape"]
    dz = case["d"][0]
    z = (np.arange(nz) + 0.5) * dz
    return np.broadcast_to(z[:, None, None], (nz, ny, nx)).copy()


def source_times(case):
    return [case["t0"] + n * case["dt"] for n in range(case["nsteps"])]


def _grid_axis(lo, hi, n):
    """Cell centres and spacing as the upstream's build_grid makes them
    (__init__.py:58-65)."""
    bnd, d = np.linspace(lo, hi, num=n + 1, retstep=True)
    return bnd, bnd[:-1] + d / 2.0, float(d)


def case_steelonfoam(nsteps=300):
    """The upstream's steel-on-foam demo (demos/steelonfoam.py:10-112):
    80x40x48 cells over 10.8 x 100 x 120 mm, 3.175 mm of steel (k=40,
    rho=7750, c=466) on foam (k=0.4, rho=40, c=1500), an insulating gap of
    10 x 10 mm in the steel/foam interface, insulating outer faces, a
    1e4 J/m^2 flash into the top layer at t=0, T starting at 0, dt = 10 ms.
    At this dt the steel's z lines are strongly coupled (k dt / (rho c dz^2)
    about 6)."""
    nz, ny, nx = 80, 40, 48
    z_bnd, z, dz = _grid_axis(0.0, 10.8e-3, nz)
    y_bnd, y, dy = _grid_axis(-0.05, 0.05, ny)
    x_bnd, x, dx = _grid_axis(-0.06, 0.06, nx)
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    iface = int(np.argmin(np.abs(3.175e-3 - z_bnd)))
    m[iface:] = 1
    yg, xg = np.meshgrid(y, x, indexing="ij")
    gap = (yg > 0.0) & (yg < 10e-3) & (xg > 0.0) & (xg < 10e-3)
    bz[iface][gap] = 1
    v[0] = 1
    materials = [(0, 40.0, 40.0, 40.0, 7.75e3, 466.0, 0),
                 (0, 0.4, 0.4, 0.4, 40.0, 1500.0, 0),
                 (1, 0, 0, 0, 0, 0, 0)]
    boundaries = [(BND_CONDUCTING, 0.0), (BND_INSULATING, 0.0)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 10e3 / dz)]
    return dict(name="steelonfoam", shape=(nz, ny, nx), d=(dz, dy, dx), dt=0.01, t0=0.0,
                nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=np.zeros((nz, ny, nx)),
                matrices=False)


def case_tutorial(nsteps=200):
    """The upstream's tutorial (demos/testheatsim2.py:7-99): 5x50x55 cells
    over 1.4 x 3.2 x 3.0 mm of titanium (k = diag(6.7, 6.7, 6.7) on the
    anisotropic face law, rho=4430, c=526.3), an insulating delamination
    at 0.4 mm depth over 2 x 1.2 mm, insulating outer faces, a 1e4 J/m^2
    flash into the top layer at t=0, T starting at 0, dt = 10 ms.  At this
    dt the x and y lines are strongly coupled (k dt / (rho c dx^2) about
    10)."""
    nz, ny, nx = 5, 50, 55
    z_bnd, z, dz = _grid_axis(0.0, 1.4e-3, nz)
    y_bnd, y, dy = _grid_axis(-1.6e-3, 1.6e-3, ny)
    x_bnd, x, dx = _grid_axis(-1.5e-3, 1.5e-3, nx)
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    zz, yy, xx = np.meshgrid(z_bnd, y, x, indexing="ij")
    delam = ((xx > -1e-3) & (xx < 1e-3) & (yy > -0.6e-3) & (yy < 0.6e-3) &
             (zz == z_bnd[np.abs(z_bnd - 0.4e-3).argmin()]))
    bz[delam] = 1
    v[0] = 1
    materials = [(0, 6.7, 6.7, 6.7, 4.43e3, 526.3, 1), (1, 0, 0, 0, 0, 0, 0)]
    boundaries = [(BND_ANISO, 0.0), (BND_INSULATING, 0.0)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 10e3 / dz)]
    return dict(name="tutorial", shape=(nz, ny, nx), d=(dz, dy, dx), dt=0.010, t0=0.0,
                nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=np.zeros((nz, ny, nx)),
                matrices=False)


def case_steel_block(nz, ny, nx, nsteps, dt, dxyz=5e-4):
    """A steel block (k=50, rho=7850, c=470) of 0.5 mm cells, insulating
    faces, a 1e4 J/m^2 flash into the top layer at t=0, T0 = 293 K."""
    d = (dxyz, dxyz, dxyz)
    m, bz, by, bx, v = _zero(nz, ny, nx)
    _outer(bz, by, bx, 1)
    v[0] = 1
    materials = [(0, 50.0, 50.0, 50.0, 7850.0, 470.0, 0)]
    boundaries = [(BND_CONDUCTING, 0.0), (BND_INSULATING, 0.0)]
    volumetric = [(NO_SOURCE,), (IMPULSE_SOURCE, 0.0, 1e4 / d[0])]
    return dict(name="steel_%dx%dx%d_dt%g" % (nz, ny, nx, dt), shape=(nz, ny, nx), d=d, dt=dt,
                t0=0.0, nsteps=nsteps, materials=materials, boundaries=boundaries,
                volumetric=volumetric, maps=(m, bz, by, bx, v), T0=np.full((nz, ny, nx), 293.0),
                matrices=False)

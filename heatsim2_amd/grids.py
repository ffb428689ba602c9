"""Grid builders and map allocation (heatsim2/__init__.py:53-197).

Cells are (k, j, i) = (z, y, x) in C order.  Boundary coordinates are the
faces; centres lie half a step inside.  These are host-side helpers with the
upstream's return layout.
"""
import numpy as np


def _grids(z_bnd, y_bnd, x_bnd, dz, dy, dx):
    z = z_bnd[:-1] + dz / 2.0
    y = y_bnd[:-1] + dy / 2.0
    x = x_bnd[:-1] + dx / 2.0
    zb = np.meshgrid(z_bnd, y, x, indexing="ij")
    yb = np.meshgrid(z, y_bnd, x, indexing="ij")
    xb = np.meshgrid(z, y, x_bnd, indexing="ij")
    zgrid, ygrid, xgrid = np.meshgrid(z, y, x, indexing="ij")
    r3d = np.sqrt(xgrid ** 2 + ygrid ** 2 + zgrid ** 2)
    r2d = np.sqrt(xgrid ** 2 + ygrid ** 2)
    return (z, y, x, zgrid, ygrid, xgrid, z_bnd, y_bnd, x_bnd,
            zb[0], zb[1], zb[2], yb[0], yb[1], yb[2], xb[0], xb[1], xb[2], r3d, r2d)


def build_grid(minz, maxz, nz, miny, maxy, ny, minx, maxx, nx):
    """Faces at min..max per axis; returns the 23-tuple (dz, dy, dx, z, y, x,
    zgrid, ygrid, xgrid, z_bnd, y_bnd, x_bnd, z_bnd_z, z_bnd_y, z_bnd_x,
    y_bnd_z, y_bnd_y, y_bnd_x, x_bnd_z, x_bnd_y, x_bnd_x, r3d, r2d)."""
    z_bnd, dz = np.linspace(minz, maxz, num=nz + 1, retstep=True)
    y_bnd, dy = np.linspace(miny, maxy, num=ny + 1, retstep=True)
    x_bnd, dx = np.linspace(minx, maxx, num=nx + 1, retstep=True)
    return (dz, dy, dx) + _grids(z_bnd, y_bnd, x_bnd, dz, dy, dx)


def build_grid_min_step(minzcenter, dz, nz, minycenter, dy, ny, minxcenter, dx, nx):
    """First cell centred at min*center; returns the 20-tuple of build_grid
    without the spacings."""
    z_bnd = minzcenter - dz / 2.0 + np.arange(nz + 1, dtype="d") * dz
    y_bnd = minycenter - dy / 2.0 + np.arange(ny + 1, dtype="d") * dy
    x_bnd = minxcenter - dx / 2.0 + np.arange(nx + 1, dtype="d") * dx
    return _grids(z_bnd, y_bnd, x_bnd, dz, dy, dx)


def build_grid_min_step_edge(minzedge, dz, nz, minyedge, dy, ny, minxedge, dx, nx):
    """First face at min*edge; returns the 20-tuple of build_grid without
    the spacings."""
    z_bnd = minzedge + np.arange(nz + 1, dtype="d") * dz
    y_bnd = minyedge + np.arange(ny + 1, dtype="d") * dy
    x_bnd = minxedge + np.arange(nx + 1, dtype="d") * dx
    return _grids(z_bnd, y_bnd, x_bnd, dz, dy, dx)


def zero_elements(nz, ny, nx):
    """(material, boundary_z, boundary_y, boundary_x, volumetric) uint8 maps,
    all zero, shaped (nz,ny,nx), (nz+1,ny,nx), (nz,ny+1,nx), (nz,ny,nx+1),
    (nz,ny,nx)."""
    return (np.zeros((nz, ny, nx), dtype="u1"), np.zeros((nz + 1, ny, nx), dtype="u1"),
            np.zeros((nz, ny + 1, nx), dtype="u1"), np.zeros((nz, ny, nx + 1), dtype="u1"),
            np.zeros((nz, ny, nx), dtype="u1"))

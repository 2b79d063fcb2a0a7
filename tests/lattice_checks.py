"""Invariants of lattice states shared by the CPU and GPU tests."""
import numpy as np

from thesis_amd import lattices as LT
from thesis_amd import lbm as L

C9 = LT.D2Q9_C[:, :2].astype(int)


def plane_fluxes(f, cls):
    """Lattice mass flux through every plane x + 1/2, 1 <= x <= nx-3, of a D2Q9
    state f (9, ny, nx) after streaming: the populations that crossed the
    plane from a non-solid source (pull: population i at (x, y) left
    (x - cx, y - cy)), rightward minus leftward.  Streaming and collision
    conserve mass, so once the flow is steady this is the same for every
    plane; the columns next to the inlet and outlet, whose incoming
    populations the boundary conditions replace, are not used."""
    ny, nx = cls.shape
    solid = cls == L.SOLID
    planes = np.arange(1, nx - 2)
    flux = np.zeros(planes.size)
    y = np.arange(ny)
    for i in range(1, 9):
        cx, cy = C9[i]
        if cx == 0:
            continue
        xd = planes + 1 if cx > 0 else planes
        xs = xd - cx
        ys = y - cy
        rows = (ys >= 0) & (ys < ny)
        ok = np.zeros((ny, planes.size), bool)
        ok[rows] = ~solid[np.ix_(ys[rows], xs)]
        ok &= ~solid[:, xd]
        flux += cx * np.where(ok, f[i][:, xd], 0.0).sum(axis=0)
    return flux


def mass_balance_error(f1, rho0, rho1, cls):
    """Discrete mass balance of one step, exact for any (not only steady)
    state: with f1 the streamed populations of step t -> t+1 and rho0, rho1
    the densities before and after, the flux through the first inner plane
    (x = 1 + 1/2) minus the flux through plane p equals the mass gained by
    columns 2..p.  Returns the largest violation relative to the flux."""
    fx = plane_fluxes(f1, cls)
    fluid = cls != L.SOLID
    dm = np.where(fluid, rho1 - rho0, 0.0).sum(axis=0)          # per column
    gained = np.cumsum(dm[2:cls.shape[1] - 1])                  # columns 2..p for p = 2 .. nx-2
    lhs = fx[0] - fx[1:]                                          # planes p = 2 .. nx-3
    return float(np.abs(lhs - gained[:lhs.size]).max() / abs(fx[0]))

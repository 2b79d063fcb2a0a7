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


# Ghia, Ghia & Shin (1982), J. Comput. Phys. 48:387-411, Table I, Re = 100:
# u / U_lid along the vertical centre line at height y / L.
GHIA_RE100_U = ((0.9766, 0.84123), (0.9688, 0.78871), (0.9609, 0.73722), (0.9531, 0.68717),
                (0.8516, 0.23151), (0.7344, 0.00332), (0.6172, -0.13641), (0.5000, -0.20581),
                (0.4531, -0.21090), (0.2813, -0.15662), (0.1719, -0.10150), (0.1016, -0.06434),
                (0.0703, -0.04775), (0.0625, -0.04192), (0.0547, -0.03717))


def ghia_centreline_error(u, u_lid):
    """Largest |u/U - Ghia| on the vertical centre line of an n x n cavity
    with walls on nodes 0 and n-1 (side n - 1), u indexed [y, x]."""
    n = u.shape[-1]
    col = 0.5 * (u[:, (n - 1) // 2] + u[:, n // 2]) / u_lid
    ys = np.arange(n) / (n - 1)
    return max(abs(float(np.interp(y, ys, col)) - g) for y, g in GHIA_RE100_U)


def square_duct_series(n, terms=99, points=None):
    """Series of fully developed laminar flow in a square duct of half-width
    a = (n-1)/2 (walls on nodes 0 and n-1; White, Viscous Fluid Flow,
    eq. 3-48): u = 16 a^2 g / (nu pi^3) * S with
    S(y, z) = sum over odd k of (-1)^((k-1)/2) / k^3
              [1 - cosh(k pi z / 2a) / cosh(k pi / 2)] cos(k pi y / 2a),
    y, z measured from the axis; returns S on the n x n nodes (or on `points`
    per axis)."""
    a = (n - 1) / 2.0
    y = np.arange(n) - a if points is None else points
    Y, Z = np.meshgrid(y, y, indexing="ij")
    S = np.zeros_like(Y, dtype=float)
    for k in range(1, 2 * terms, 2):
        S += (-1) ** ((k - 1) // 2) / k ** 3 * (1 - np.cosh(k * np.pi * Z / (2 * a)) / np.cosh(k * np.pi / 2)) \
            * np.cos(k * np.pi * Y / (2 * a))
    return S


def square_duct_profile(n, terms=99):
    """The series profile on the nodes normalised to mean 1 over the continuous
    cross-section (u_max / u_mean = 2.0962 for a square)."""
    a = (n - 1) / 2.0
    fine = np.linspace(-a, a, 801)
    Sf = square_duct_series(n, terms, fine)
    mean = np.trapezoid(np.trapezoid(Sf, fine, axis=1), fine) / (2 * a) ** 2
    return square_duct_series(n, terms) / mean


def sphere_wall_q(shape, centre, r, q=9):
    """Solid nodes inside a circle (2D, shape (ny, nx)) or sphere (3D, shape
    (nz, ny, nx)) of radius r, the wall links of the fluid nodes into them and
    each link's wall fraction: the smaller root t in (0, 1] of |p + t c| = r
    with p the node's offset from the centre.  Returns (solid, bb, wall_q)
    with wall_q shaped shape + (19,)."""
    c = LT.velocities(q).astype(float)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    pos = [g.astype(float) for g in grids[::-1]]          # x, y[, z]
    ctr = list(centre)
    rel = [pos[a] - ctr[a] for a in range(len(shape))]
    solid = sum(p * p for p in rel) <= r * r
    bb = np.zeros(shape, np.uint32)
    wq = np.zeros(shape + (19,))
    for d in range(1, q):
        cd = c[d, :len(shape)]
        nb = tuple((grids[::-1][a] + int(cd[a])) for a in range(len(shape)))
        inside = np.zeros(shape, bool)
        ok = np.ones(shape, bool)
        for a in range(len(shape)):
            ok &= (nb[a] >= 0) & (nb[a] < shape[::-1][a])
        idx = tuple(np.clip(nb[a], 0, shape[::-1][a] - 1) for a in range(len(shape)))[::-1]
        inside[ok] = solid[idx][ok]
        link = ~solid & inside
        cc = float(cd @ cd)
        pc = sum(rel[a] * cd[a] for a in range(len(shape)))
        pp = sum(rel[a] * rel[a] for a in range(len(shape)))
        t = (-pc - np.sqrt(np.maximum(pc * pc - cc * (pp - r * r), 0.0))) / cc
        bb |= np.where(link, np.uint32(1 << d), np.uint32(0))
        wq[..., d] = np.where(link, t, 0.0)
    return solid, bb, wq

"""Lattice builders in the solver's node description.

A lattice is `node_class` (uint8 per node: SOLID, FLUID, VEL_W, PRES_E, VEL_N) plus
`bb` (uint32 per node, bit d set when the link from the node in direction d
crosses a wall -- the content of a latticegen2 _BC.dat WALL line,
L/FluidTracking.cpp:129-156), in the reference's node order
ind = i + j*n + k*n*m (S/include/Iterate.h:10-35).
"""
import numpy as np

from .lbm import SOLID, FLUID, VEL_W, PRES_E, VEL_N

# reference direction numbering (S/GpuInit.cu:94-99, :822-842)
D2Q9_C = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0],
                   [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]])
D3Q19_C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                    [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def velocities(q):
    return D2Q9_C if q == 9 else D3Q19_C


def weights(q):
    return D2Q9_W if q == 9 else D3Q19_W


def opposite(q):
    c = velocities(q)
    return np.array([int(np.where((c == -c[i]).all(axis=1))[0][0]) for i in range(q)])


def bb_from_solid(solid, q, periodic=(True, True, True), outside_is_wall=(False, False, False)):
    """Bounce-back link masks of the non-solid nodes of a (nz, ny, nx) or (ny, nx)
    boolean solid field: bit d set when node + c_d is solid, or lies outside
    the box on an axis flagged in outside_is_wall."""
    s3 = solid[None] if solid.ndim == 2 else solid
    nz, ny, nx = s3.shape
    bb = np.zeros(s3.shape, np.uint32)
    c = velocities(q)
    for d in range(1, q):
        dx, dy, dz = (int(v) for v in c[d])
        nb = s3
        # neighbour at +c_d: roll by -c_d
        nb = np.roll(nb, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
        hit = nb.copy()
        for ax, (dd, n) in enumerate(((dx, nx), (dy, ny), (dz, nz))):
            if dd == 0:
                continue
            axis = 2 - ax
            edge = [slice(None)] * 3
            edge[axis] = slice(n - 1, n) if dd > 0 else slice(0, 1)
            if not periodic[ax]:
                hit[tuple(edge)] = bool(outside_is_wall[ax])
            elif outside_is_wall[ax]:
                hit[tuple(edge)] = True
        bb |= (hit & ~s3).astype(np.uint32) << np.uint32(d)
    return bb[0] if solid.ndim == 2 else bb


def channel2d(nx, ny, u_max, obstacles=(), rho_out=1.0):
    """D2Q9 channel: half-way walls at y = -1/2 and y = ny - 1/2, Zou-He
    parabolic velocity inlet on x = 0 (u = 4 u_max (y+1/2)(ny-y-1/2)/ny^2),
    Zou-He pressure outlet on x = nx-1, optional circular obstacles
    (cx, cy, r): nodes with |x - c| <= r are solid."""
    solid = np.zeros((ny, nx), bool)
    yy, xx = np.mgrid[0:ny, 0:nx]
    for (ox, oy, r) in obstacles:
        solid |= (xx - ox) ** 2 + (yy - oy) ** 2 <= r * r
    cls = np.where(solid, SOLID, FLUID).astype(np.uint8)
    cls[:, 0] = np.where(solid[:, 0], SOLID, VEL_W)
    cls[:, -1] = np.where(solid[:, -1], SOLID, PRES_E)
    bb = bb_from_solid(solid, 9, periodic=(False, False, True), outside_is_wall=(False, True, False))
    y = np.arange(ny) + 0.5
    prof = 4.0 * u_max * y * (ny - y) / ny ** 2
    ubc = np.zeros((ny, nx))
    ubc[:, 0] = prof
    return dict(node_class=cls, bb=bb, ubc=ubc, vbc=np.zeros((ny, nx)), solid=solid, profile=prof,
                periodic=(0, 0, 0), rho_out=rho_out)


def cavity2d(n, u_lid):
    """D2Q9 lid-driven cavity of n x n nodes in the reference's layout for a
    latticegen2 box (S/GpuBoundaries.cu:21-41 for the lid): every node fluid,
    wall links on every link leaving the west, east and south faces (for
    on-node bounce-back, walls at x = 0, n-1 and y = 0), the top row a moving
    lid (VEL_N) at u_lid.  The two top corner nodes carry the wall links of
    their side wall; the cavity side is n - 1."""
    c = velocities(9)
    yy, xx = np.mgrid[0:n, 0:n]
    bb = np.zeros((n, n), np.uint32)
    for d in range(1, 9):
        sx, sy = xx + c[d, 0], yy + c[d, 1]
        bb |= np.where((sx < 0) | (sx >= n) | (sy < 0), np.uint32(1 << d), np.uint32(0))
    cls = np.full((n, n), FLUID, np.uint8)
    cls[-1, :] = VEL_N
    ubc = np.where(cls == VEL_N, float(u_lid), 0.0)
    return dict(node_class=cls, bb=bb, ubc=ubc, vbc=np.zeros((n, n)), periodic=(0, 0, 0))


def duct3d(nx, n, u_in):
    """D3Q19 square duct along x with wet-node walls on the planes y, z = 0
    and n-1 (the reference's HHmodel layout: every node fluid, the wall
    nodes carry a wall link on every link leaving the box, so faces have 5,
    edges 9 and corners 12 of them), a plug velocity inlet (VEL_W, u_in along
    x) on the inner nodes of x = 0 and an outlet (PRES_E) on the inner nodes of
    x = nx-1.  The duct side is n - 1."""
    c = velocities(19)
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:nx]
    wall = (yy == 0) | (yy == n - 1) | (zz == 0) | (zz == n - 1)
    bb = np.zeros((n, n, nx), np.uint32)
    for d in range(1, 19):
        sx, sy, sz = xx + c[d, 0], yy + c[d, 1], zz + c[d, 2]
        out = (sx < 0) | (sx >= nx) | (sy < 0) | (sy >= n) | (sz < 0) | (sz >= n)
        bb |= np.where(wall & out, np.uint32(1 << d), np.uint32(0))
    cls = np.full((n, n, nx), FLUID, np.uint8)
    cls[~wall & (xx == 0)] = VEL_W
    cls[~wall & (xx == nx - 1)] = PRES_E
    ubc = np.where(cls == VEL_W, float(u_in), 0.0)
    zero = np.zeros((n, n, nx))
    return dict(node_class=cls, bb=bb, ubc=ubc, vbc=zero, wbc=zero.copy(), periodic=(0, 0, 0))


def channel_sweep(count=64, seed=7, nx_range=(128, 1024), ny_range=(32, 256), re_range=(10.0, 100.0),
                  obstacles=(1, 5), r_range=(4.0, 20.0), trt_magic=3.0 / 16.0, u_cap=0.1, drho=0.03):
    """Seeded family of D2Q9 channels for a batched parametric sweep.

    Per channel, from numpy.random.Generator(PCG64(seed)): nx, ny uniform
    integers in the given ranges with ny <= nx/4 (aspect ratio at least 4);
    Re ~ U[re_range] on the hydraulic diameter D_h = 2 ny, Re = u_max D_h / nu.
    The peak inlet speed is the largest (up to u_cap) for which the Poiseuille
    density drop along the channel, 3 dp/rho = 24 nu u_max nx / ny^2
    = 48 u_max^2 nx / (ny Re), stays at `drho`, so the flow stays weakly
    compressible; then nu = 2 u_max ny / Re and omega = 1 / (3 nu + 1/2).
    1-5 non-overlapping circular obstacles with radius ~ U[4, min(20, ny/8)]
    (cylinder Reynolds number 2 r u_max / nu <= Re/8, far below vortex
    shedding) are kept out of the inlet strip (ny/2 nodes) and the outlet strip
    (ny nodes) and two nodes clear of the walls.  Each entry is channel2d's
    dict plus omega, trt_magic, Re, nu, u_max, obstacles and the shape."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], max(ny_range[0], min(ny_range[1], nx // 4)) + 1))
        re = float(rng.uniform(*re_range))
        u_max = min(u_cap, float(np.sqrt(drho * re * ny / (48.0 * nx))))
        nu = 2.0 * u_max * ny / re
        rmax = max(r_range[0], min(r_range[1], ny / 8.0))
        nobs = int(rng.integers(obstacles[0], obstacles[1] + 1))
        obs = []
        x0, x1 = ny / 2.0, nx - ny
        for _attempt in range(200):
            if len(obs) == nobs:
                break
            r = float(rng.uniform(r_range[0], rmax))
            lo_x, hi_x = x0 + r, x1 - r
            lo_y, hi_y = r + 2.0, ny - 3.0 - r
            if hi_x <= lo_x or hi_y <= lo_y:
                continue
            cx, cy = float(rng.uniform(lo_x, hi_x)), float(rng.uniform(lo_y, hi_y))
            if all((cx - ox) ** 2 + (cy - oy) ** 2 > (r + orr + 2.0) ** 2 for ox, oy, orr in obs):
                obs.append((cx, cy, r))
        d = channel2d(nx, ny, u_max, obstacles=obs)
        d.update(omega=1.0 / (3.0 * nu + 0.5), trt_magic=trt_magic, Re=re, nu=nu, obstacles=obs, nx=nx, ny=ny,
                 u_max=u_max)
        out.append(d)
    return out


def surrogate_channels(count, seed, nx=512, ny=128, obstacles=(1, 6), r_range=(4.0, 16.0), u_range=(0.04, 0.1),
                       tau=1.0):
    """Seeded D2Q9 channels of one size for the learned warm start.

    Per channel, from numpy.random.Generator(PCG64(seed)): peak inlet speed
    u_max ~ U[u_range] at fixed tau (nu = (tau - 1/2)/3, so Re = 2 u_max ny / nu
    on the hydraulic diameter runs over about 60-150 for the defaults), and
    1-6 non-overlapping circular obstacles with radius ~ U[r_range], kept two
    nodes clear of the walls and out of the inlet strip (ny/2 nodes) and the
    outlet strip (ny nodes).  Each entry is channel2d's dict plus omega, nu,
    u_max, Re and the obstacles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nu = (tau - 0.5) / 3.0
    out = []
    for _ in range(count):
        u_max = float(rng.uniform(*u_range))
        nobs = int(rng.integers(obstacles[0], obstacles[1] + 1))
        obs = []
        for _attempt in range(400):
            if len(obs) == nobs:
                break
            r = float(rng.uniform(*r_range))
            cx = float(rng.uniform(ny / 2.0 + r, nx - ny - r))
            cy = float(rng.uniform(r + 2.0, ny - 3.0 - r))
            if all((cx - ox) ** 2 + (cy - oy) ** 2 > (r + orr + 2.0) ** 2 for ox, oy, orr in obs):
                obs.append((cx, cy, r))
        d = channel2d(nx, ny, u_max, obstacles=obs)
        d.update(omega=1.0 / tau, trt_magic=3.0 / 16.0, nu=nu, u_max=u_max, Re=2.0 * u_max * ny / nu,
                 obstacles=obs, nx=nx, ny=ny)
        out.append(d)
    return out


def poiseuille_start(d):
    """Fully developed channel flow of d's inlet in every column: the inlet
    profile, zero in solid nodes, and the Poiseuille pressure drop
    rho = 1 + 3 G (nx - 1 - x) with G = 8 nu u_max / ny^2 down to the outlet
    density 1 (walls at y = -1/2 and ny - 1/2).  The analytic start the
    learned one has to beat."""
    cls = d["node_class"]
    ny, nx = cls.shape
    nu = (1.0 / d["omega"] - 0.5) / 3.0
    G = 8.0 * nu * d["u_max"] / ny ** 2
    solid = cls == SOLID
    u = np.where(solid, 0.0, d["profile"][:, None])
    rho = np.where(solid, 1.0, 1.0 + 3.0 * G * (nx - 1.0 - np.arange(nx))[None, :])
    return rho, u, np.zeros_like(u)


def porous3d(n, seed=2024, rmin=6.0, rmax=12.0, solid_fraction=0.30, ny=None, nz=None):
    """Periodic overlapping-sphere porous medium on an n x ny x nz box.

    Sphere centres ~ U[0, n)^3 and radii ~ U[rmin, rmax] from
    numpy.random.Generator(PCG64(seed)); a node is solid when its periodic
    minimum-image distance to a centre is <= r.  Spheres are added until the
    solid fraction reaches `solid_fraction`.  Returns (node_class, bb, info)."""
    ny = ny or n
    nz = nz or n
    rng = np.random.Generator(np.random.PCG64(seed))
    solid = np.zeros((nz, ny, nx := n), bool)
    total = nx * ny * nz
    count = 0
    spheres = 0
    while count < solid_fraction * total:
        c = rng.uniform(0.0, 1.0, 3) * np.array([nx, ny, nz])
        r = rng.uniform(rmin, rmax)
        ri = int(np.ceil(r))
        ix = np.arange(int(np.floor(c[0])) - ri, int(np.floor(c[0])) + ri + 2)
        iy = np.arange(int(np.floor(c[1])) - ri, int(np.floor(c[1])) + ri + 2)
        iz = np.arange(int(np.floor(c[2])) - ri, int(np.floor(c[2])) + ri + 2)
        dz = (iz - c[2])[:, None, None]
        dy = (iy - c[1])[None, :, None]
        dx = (ix - c[0])[None, None, :]
        inside = dx * dx + dy * dy + dz * dz <= r * r
        sub = np.ix_(iz % nz, iy % ny, ix % nx)
        before = solid[sub]
        # duplicated wrapped indices only occur when 2r+2 > n; handled by |=
        solid[sub] = before | inside
        count = int(solid.sum()) if (2 * ri + 2 > min(nx, ny, nz)) else count + int((inside & ~before).sum())
        spheres += 1
    count = int(solid.sum())
    cls = np.where(solid, SOLID, FLUID).astype(np.uint8)
    bb = bb_from_solid(solid, 19, periodic=(True, True, True))
    info = dict(seed=seed, spheres=spheres, solid_fraction=count / total, n_fluid=total - count)
    return cls, bb, info


def droplet2d(nx, ny, radius, cx=None, cy=None):
    """Colour densities of a disc of radius `radius` (red inside, blue
    outside), centred in an nx x ny periodic box."""
    cx = (nx - 1) / 2.0 if cx is None else cx
    cy = (ny - 1) / 2.0 if cy is None else cy
    yy, xx = np.mgrid[0:ny, 0:nx]
    inside = (xx - cx) ** 2 + (yy - cy) ** 2 <= radius ** 2
    r = inside.astype(np.float64)
    return r, 1.0 - r


def droplet3d(nx, ny, nz, radius, cx=None, cy=None, cz=None):
    """Colour densities (nz, ny, nx) of a sphere of radius `radius` (red
    inside, blue outside) in an nx x ny x nz periodic box."""
    cx = (nx - 1) / 2.0 if cx is None else cx
    cy = (ny - 1) / 2.0 if cy is None else cy
    cz = (nz - 1) / 2.0 if cz is None else cz
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    inside = (xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2 <= radius ** 2
    r = inside.astype(np.float64)
    return r, 1.0 - r


def bubble3d_populations(r_rho, b_rho):
    """Colour populations (19, nz, ny, nx) of the reference's steady-bubble
    start (S/GpuInit.cu:543-592): f_q = rho_k phi_q with phi = 1/12 on the
    axes, 1/24 on the diagonals and f_0 = 0, whatever the rest fraction."""
    phi = np.array([0.0] + [1 / 12] * 6 + [1 / 24] * 12)
    r = np.asarray(r_rho, np.float64)
    b = np.asarray(b_rho, np.float64)
    return phi[:, None, None, None] * r[None], phi[:, None, None, None] * b[None]


def write_latticegen2(path_prefix, node_class, bb=None, delta=1.0, inlet_dirs=None, outlet_dirs=None, wall_q=None,
                      cyclic_axes=()):
    """Write <prefix>_Node.dat and <prefix>_BC.dat in the latticegen2 layout
    (L/main.cpp:59-66, L/FluidTracking.cpp:229-236): 2D lines "i j x y fluid"
    and "i j dir type x y id", 3D with the k column.  WALL links come from
    `bb`; VEL_W / PRES_E / VEL_N nodes of a 2D lattice get INLE / OUTL / INLE
    links in the directions leaving the west / east / north face (wall links
    excepted).  A WALL line's point lies at the wall fraction wall_q[..., d]
    of the link (1/2 when wall_q is None or 0 there).  With cyclic_axes
    (0 x, 1 y, 2 z) every link that leaves the lattice through faces of those
    axes only is a CYCL link, and one that also leaves through a face of
    another axis an INLE link (the box's remaining faces)."""
    cls = np.asarray(node_class)
    two_d = cls.ndim == 2
    c3 = cls[None] if two_d else cls
    nz, ny, nx = c3.shape
    b3 = np.zeros(c3.shape, np.uint32) if bb is None else (np.asarray(bb)[None] if two_d else np.asarray(bb))
    q = 9 if two_d else 19
    c = velocities(q)
    inlet_dirs = [d for d in range(1, q) if c[d, 0] < 0] if inlet_dirs is None else inlet_dirs
    outlet_dirs = [d for d in range(1, q) if c[d, 0] > 0] if outlet_dirs is None else outlet_dirs
    with open(path_prefix + "_Node.dat", "w") as fh:
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    fl = 1 if c3[k, j, i] != SOLID else 0
                    if two_d:
                        fh.write(f"{i}\t{j}\t{i * delta:.6f}\t{j * delta:.6f}\t{fl}\n")
                    else:
                        fh.write(f"{i}\t{j}\t{k}\t{i * delta:.6f}\t{j * delta:.6f}\t{k * delta:.6f}\t{fl}\n")
    with open(path_prefix + "_BC.dat", "w") as fh:
        wq = None if wall_q is None else (np.asarray(wall_q)[None] if two_d else np.asarray(wall_q))

        def link(i, j, k, d, t):
            f = 0.5
            if t == 1 and wq is not None and wq[k, j, i, d] > 0:
                f = float(wq[k, j, i, d])
            x, y, z = (i + f * c[d, 0]) * delta, (j + f * c[d, 1]) * delta, (k + f * c[d, 2]) * delta
            if two_d:
                fh.write(f"{i}\t{j}\t{d}\t{t}\t{x:.6f}\t{y:.6f}\t1\n")
            else:
                fh.write(f"{i}\t{j}\t{k}\t{d}\t{t}\t{x:.6f}\t{y:.6f}\t{z:.6f}\t1\n")
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    if c3[k, j, i] == SOLID:
                        continue
                    for d in range(1, q):
                        if (int(b3[k, j, i]) >> d) & 1:
                            link(i, j, k, d, 1)
                    if c3[k, j, i] == VEL_W:
                        for d in inlet_dirs:
                            if not (int(b3[k, j, i]) >> d) & 1:
                                link(i, j, k, d, 2)
                    if c3[k, j, i] == VEL_N:
                        for d in range(1, q):
                            if c[d, 1] > 0 and not (int(b3[k, j, i]) >> d) & 1:
                                link(i, j, k, d, 2)
                    if c3[k, j, i] == PRES_E:
                        for d in outlet_dirs:
                            if not (int(b3[k, j, i]) >> d) & 1:
                                link(i, j, k, d, 3)
                    if cyclic_axes:
                        pos, size = (i, j, k), (nx, ny, nz)
                        for d in range(1, q):
                            if (int(b3[k, j, i]) >> d) & 1:
                                continue
                            out = [a for a in range(3) if size[a] > 1 and
                                   ((c[d, a] < 0 and pos[a] == 0) or (c[d, a] > 0 and pos[a] == size[a] - 1))]
                            if out:
                                link(i, j, k, d, 4 if all(a in cyclic_axes for a in out) else 2)

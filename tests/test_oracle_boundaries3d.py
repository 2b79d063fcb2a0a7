"""Pins the oracle's D3Q19 boundaries: velocity faces (Hecht-Harting form of
S/GpuBoundaries.cu:73-399), wet-node walls on faces, edges and corners
(:492-1560), outlets (:1655-1874) and symmetry planes (:2086-2167).
Node-local conditions are checked against what they must impose on random
populations (moments, normal non-equilibrium bounce-back, zero wall
velocity, rest equilibria as fixed points); the assembled duct against the
series solution of fully developed square-duct flow.  CPU only."""
import itertools

import numpy as np
import pytest

from thesis_amd import lattices as LT
from lattice_checks import square_duct_profile

C = LT.velocities(19).astype(float)
W = LT.weights(19)
FACES = {  # class name: (axis, inward sign)
    "VEL_W": (0, 1), "VEL_E": (0, -1), "VEL_S": (1, 1), "VEL_N": (1, -1), "VEL_B": (2, 1), "VEL_T": (2, -1),
}


def node(oracle, cls_value, bb=0, u=(0.0, 0.0, 0.0), **kw):
    """3 x 3 x 3 lattice with the node under test in the middle (index 13)."""
    cls = np.ones((3, 3, 3), np.uint8)
    cls.ravel()[13] = cls_value
    bbs = np.zeros((3, 3, 3), np.uint32)
    bbs.ravel()[13] = bb
    U = [np.full((3, 3, 3), float(c)) for c in u]
    s = oracle.Sim(19, (3, 3, 3), cls, bbs, ubc=U[0], vbc=U[1], wbc=U[2], **kw)
    return s, 13


def mom(f):
    rho = f.sum()
    return rho, C.T @ f / rho


def rand_f(rng):
    return W * (1.0 + 0.2 * rng.standard_normal(19))


def eq(oracle, rho, u):
    return oracle.equilibrium(19, rho, tuple(u))


def dir_of(c):
    return int(np.where((C == np.asarray(c, float)).all(axis=1))[0][0])


@pytest.mark.parametrize("face", list(FACES))
def test_velocity_face_imposes_u_and_normal_bounce_back(oracle, face):
    a, g = FACES[face]
    rng = np.random.default_rng(a * 2 + (g > 0))
    for _ in range(4):
        U = 0.05 * rng.standard_normal(3)
        s, ind = node(oracle, getattr(oracle, face), u=U)
        f = rand_f(rng)
        h = s.open_boundary(ind, f)
        known = g * C[:, a] <= 0
        assert (h[known] == f[known]).all()
        rho, u = mom(h)
        assert np.abs(u - U).max() < 1e-14
        e = eq(oracle, rho, U)
        nin, nout = dir_of(g * np.eye(3)[a]), dir_of(-g * np.eye(3)[a])
        assert abs((h[nin] - e[nin]) - (h[nout] - e[nout])) < 1e-15


@pytest.mark.parametrize("face", ["VEL_W", "VEL_N", "VEL_T"])
def test_frozen_inlet_collides_with_initial_macros(oracle, face):
    """UpdateInltOutl = no (S/GpuUpdateMacro.cu:150-151): the 3D macro update
    skips velocity-face nodes, which keep their initial macros and collide
    with them.  BGK at omega = 1 makes the post-collision populations
    f_eq(rho, u) of the macros the collision reads, and streaming leaves the
    rest population in place, so after the second step f_0 = w_0 rho
    (1 - 1.5 |u|^2) with (rho, u) the initial macros under the flag and the
    face's own moments of the first step (rho_1, u_bc) without it."""
    rho0, u0, ubc = 1.02, np.array([0.01, -0.02, 0.015]), np.array([0.04, 0.01, -0.02])
    init = dict(rho=np.full(27, rho0), u=np.full(27, u0[0]), v=np.full(27, u0[1]), w=np.full(27, u0[2]))
    res = {}
    for frozen in (True, False):
        s, ind = node(oracle, getattr(oracle, face), u=ubc, collision=0, omega=1.0,
                      parity=oracle.PARITY_INLET_FROZEN if frozen else 0)
        s.init(0, **init)
        s.step(1)
        m1 = [a.ravel()[ind] for a in s.macros()]
        s.step(1)
        m2 = [a.ravel()[ind] for a in s.macros()]
        res[frozen] = (m1, m2, s.get_f()[0].ravel()[ind])
    m1, m2, f0 = res[True]
    assert m1 == m2 == [rho0, *u0]
    assert abs(f0 - W[0] * rho0 * (1 - 1.5 * u0 @ u0)) < 1e-15
    m1, m2, f0 = res[False]
    assert np.abs(np.array(m1[1:]) - ubc).max() < 1e-14 and abs(m1[0] - rho0) > 1e-5
    assert abs(f0 - W[0] * m1[0] * (1 - 1.5 * ubc @ ubc)) < 1e-15


@pytest.mark.parametrize("face", list(FACES))
def test_reference_west_east_density_flip(oracle, face):
    """With the parity flag the W and E faces divide by 1 + g u_a, so the node
    moves at U / (1 + 2 g U_a); the other faces are unchanged."""
    a, g = FACES[face]
    rng = np.random.default_rng(7)
    U = np.array([0.04, -0.02, 0.03])
    f = rand_f(rng)
    s, ind = node(oracle, getattr(oracle, face), u=U, parity=oracle.PARITY_INLET3D_FLIP)
    _, u = mom(s.open_boundary(ind, f))
    want = U / (1 + 2 * g * U[a]) if a == 0 else U
    assert np.abs(u - want).max() < 1e-14


def wall_bits(normals):
    """bb of a box node whose walls have the given inward normals (axis: sign)."""
    bb = 0
    for d in range(1, 19):
        if any(C[d, a] == -g for a, g in normals.items()):
            bb |= 1 << d
    return bb


def wall_cases():
    out = []
    for a in range(3):
        for g in (1, -1):
            out.append({a: g})
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for ga, gb in itertools.product((1, -1), repeat=2):
            out.append({a: ga, b: gb})
    for gs in itertools.product((1, -1), repeat=3):
        out.append(dict(enumerate(gs)))
    return out


@pytest.mark.parametrize("normals", wall_cases(), ids=lambda n: "".join(f"{'xyz'[a]}{'+' if g > 0 else '-'}" for a, g in n.items()))
def test_wet_node_walls_stop_the_node_and_keep_rest(oracle, normals):
    bb = wall_bits(normals)
    assert bin(bb).count("1") == {1: 5, 2: 9, 3: 12}[len(normals)]
    s, ind = node(oracle, oracle.FLUID, bb=bb, wall_model=oracle.WALL_HH)
    rng = np.random.default_rng(len(normals))
    for _ in range(3):
        f = rand_f(rng)
        h = s.wall_node(ind, f)
        rho, u = mom(h)
        assert np.abs(u).max() < 1e-15
        known = np.ones(19, bool)
        known[0] = False
        for a, g in normals.items():
            known &= g * C[:, a] <= 0
        assert (h[known] == f[known]).all()
        r = 1 + 0.1 * rng.random()
        e = eq(oracle, r, (0, 0, 0))
        assert np.abs(s.wall_node(ind, e) - e).max() < 4e-16
    # a wall node of another shape keeps its values
    s, ind = node(oracle, oracle.FLUID, bb=(1 << 2) | (1 << 8), wall_model=oracle.WALL_HH)
    f = rand_f(rng)
    assert (s.wall_node(ind, f) == f).all()


@pytest.mark.parametrize("face", ["SYM_W", "SYM_E", "SYM_S", "SYM_N", "SYM_B", "SYM_T"])
def test_symmetry_plane_mirrors(oracle, face):
    a, g = FACES["VEL_" + face[-1]]
    s, ind = node(oracle, getattr(oracle, face))
    rng = np.random.default_rng(3)
    f = rand_f(rng)
    h = s.open_boundary(ind, f)
    known = g * C[:, a] <= 0
    assert (h[known] == f[known]).all()
    assert abs(C[:, a] @ h) < 1e-16
    U = np.array([0.03, -0.02, 0.05])
    U[a] = 0
    e = eq(oracle, 1.02, U)
    assert np.abs(s.open_boundary(ind, e) - e).max() < 4e-16


def test_outlets_3d(oracle):
    """East outlet of a 3D channel: the unknowns (c_x < 0) are f(x-1),
    2 f(x-1) - f(x-2) of the same step's upstream nodes, or f_eq of the
    node's macros stored by the previous step; other populations are those
    streamed in."""
    nz, ny, nx = 4, 5, 10
    cls = np.ones((nz, ny, nx), np.uint8)
    cls[..., 0] = oracle.VEL_W
    cls[..., -1] = oracle.PRES_E
    rng = np.random.default_rng(2)
    u = 0.03 + 0.01 * rng.random((nz, ny, nx))
    v = 0.01 * rng.standard_normal((nz, ny, nx))
    U = np.full((nz, ny, nx), 0.03)
    unk = C[:, 0] < 0
    res = {}
    for model in (oracle.OUTLET_EXTRAP1, oracle.OUTLET_EXTRAP2, oracle.OUTLET_FEQ):
        s = oracle.Sim(19, (nz, ny, nx), cls, None, collision=1, omega=1.2, periodic=(0, 1, 1), ubc=U,
                       vbc=np.zeros_like(U), wbc=np.zeros_like(U), outlet_model=model)
        s.init(0, u=u, v=v)
        s.step(2)
        r0, u0, v0, w0 = s.macros()
        s.step(1)
        f = s.get_f()
        res[model] = f
        if model == oracle.OUTLET_FEQ:
            for k in range(nz):
                for j in range(ny):
                    e = eq(oracle, r0[k, j, -1], (u0[k, j, -1], v0[k, j, -1], w0[k, j, -1]))
                    assert np.abs(f[unk, k, j, -1] - e[unk]).max() < 1e-15
        else:
            a, b = f[unk][..., -2], f[unk][..., -3]
            want = a if model == oracle.OUTLET_EXTRAP1 else 2 * a - b
            assert np.abs(f[unk][..., -1] - want).max() < 1e-15
    assert np.abs(res[oracle.OUTLET_EXTRAP1][unk][..., -1] - res[oracle.OUTLET_EXTRAP2][unk][..., -1]).max() > 1e-8


@pytest.mark.parametrize("model", ["OUTLET_EXTRAP1", "OUTLET_EXTRAP2", "OUTLET_FEQ"])
def test_uniform_stream_through_3d_box_stays_uniform(oracle, model):
    nz, ny, nx = 3, 4, 12
    cls = np.ones((nz, ny, nx), np.uint8)
    cls[..., 0] = oracle.VEL_W
    cls[..., -1] = oracle.PRES_E
    U = 0.04
    z = np.zeros((nz, ny, nx))
    s = oracle.Sim(19, (nz, ny, nx), cls, None, collision=1, omega=1.3, periodic=(0, 1, 1), ubc=z + U, vbc=z,
                   wbc=z, outlet_model=getattr(oracle, model))
    s.init(0, u=z + U)
    s.step(100)
    rho, u, v, w = s.macros()
    assert np.abs(u - U).max() < 1e-13 and np.abs(v).max() < 1e-13 and np.abs(rho - 1).max() < 1e-13


def test_duct_flow_matches_series_solution(oracle):
    """16 x 16 square duct, 40 long: wet-node walls, plug velocity inlet, the
    pressure outlet, TRT, to a steady state.  The node-sum mass flux is the
    same on every plane of the developed part, and the mid-length profile's
    shape is within 0.5 % (of the maximum) of the series solution."""
    n, nx, U = 16, 40, 0.02
    d = LT.duct3d(nx, n, U)
    s = oracle.Sim(19, (n, n, nx), d["node_class"], d["bb"], collision=1, omega=1 / 0.8, trt_magic=0.25,
                   wall_model=oracle.WALL_HH, outlet_model=oracle.OUTLET_PRESSURE, rho_out=1.0, ubc=d["ubc"],
                   vbc=d["vbc"], wbc=d["wbc"], nthreads=8)
    s.init(0)
    prev = None
    for _ in range(40):
        s.step(500)
        u = s.macros()[1]
        if prev is not None and np.abs(u - prev).max() < 1e-10:
            break
        prev = u.copy()
    assert np.abs(u - prev).max() < 1e-10
    rho = s.macros()[0]
    flux = (rho * u).sum(axis=(0, 1))
    dev = flux[nx // 3:-3]                      # developed part: node sums are plane fluxes there
    assert np.ptp(dev) < 1e-4 * dev.mean()
    mid = u[:, :, nx // 2]
    ref = square_duct_profile(n)
    scale = (ref[1:-1, 1:-1] * mid[1:-1, 1:-1]).sum() / (mid[1:-1, 1:-1] ** 2).sum()
    err = np.abs(scale * mid - ref).max() / ref.max()
    print(f"duct shape error {err:.4f}")
    assert err < 0.005


def duct_walls_only(n, nx):
    """bb of wet-node walls on y, z = 0 and n-1 of an x-periodic duct."""
    c = LT.velocities(19)
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:nx]
    wall = (yy == 0) | (yy == n - 1) | (zz == 0) | (zz == n - 1)
    bb = np.zeros((n, n, nx), np.uint32)
    for d in range(1, 19):
        sy, sz = yy + c[d, 1], zz + c[d, 2]
        bb |= np.where(wall & ((sy < 0) | (sy >= n) | (sz < 0) | (sz >= n)), np.uint32(1 << d), np.uint32(0))
    return bb


def test_forced_periodic_duct_matches_series_magnitude(oracle):
    """x-periodic 16 x 16 duct driven by a Guo body force g, wet-node walls:
    the steady velocity equals 16 a^2 g / (nu pi^3) times the series to 0.5 %
    of its maximum (the wall lies on the wall nodes)."""
    from lattice_checks import square_duct_series
    n, nx, g, tau = 16, 6, 1e-5, 0.8
    s = oracle.Sim(19, (n, n, nx), np.ones((n, n, nx), np.uint8), duct_walls_only(n, nx), collision=1,
                   omega=1 / tau, trt_magic=0.25, wall_model=oracle.WALL_HH, force_model=oracle.FORCE_GUO,
                   g=(g, 0, 0), periodic=(1, 0, 0), nthreads=8)
    s.init(0)
    s.step(40000)
    u = s.macros()[1][:, :, 2]
    nu, a = (tau - 0.5) / 3, (n - 1) / 2
    want = 16 * a * a * g / (nu * np.pi ** 3) * square_duct_series(n)
    err = np.abs(u - want).max() / want.max()
    print(f"forced duct error {err:.4f}")
    assert err < 0.005

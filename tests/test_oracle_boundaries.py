"""Pins the oracle's 2D open boundaries beyond the west inlet and the Zou-He
pressure outlet: the north lid (textbook Zou-He and the reference's form,
S/GpuBoundaries.cu:21-41), the reference's velocity outlet (:1571-1605) and
its extrapolating outlets (:1609-1640).  Node-local conditions are checked
against what they must impose (moments, non-equilibrium bounce-back, their
rest and equilibrium limits), the assembled lid against the Ghia et al.
Re = 100 cavity data.  CPU only."""
import numpy as np
import pytest

from thesis_amd import lattices as LT
from lattice_checks import ghia_centreline_error

C = LT.velocities(9)[:, :2].astype(float)


def one_node(oracle, cls_value, bb=0, inlet_model=0, outlet_model=0, u0=0.0, v0=0.0):
    """3 x 3 lattice whose node 7 (x = 1, y = 2, the top row) or node 5
    (x = 2, y = 1, the east column) carries the class under test."""
    cls = np.ones((3, 3), np.uint8)
    ind = 7 if cls_value == oracle.VEL_N else 5
    cls.ravel()[ind] = cls_value
    bbs = np.zeros((3, 3), np.uint32)
    bbs.ravel()[ind] = bb
    ubc = np.full((3, 3), u0)
    vbc = np.full((3, 3), v0)
    s = oracle.Sim(9, (3, 3), cls, bbs, inlet_model=inlet_model, outlet_model=outlet_model, ubc=ubc, vbc=vbc)
    return s, ind


def moments(f):
    rho = f.sum()
    return rho, (C[:, 0] * f).sum() / rho, (C[:, 1] * f).sum() / rho


def random_f(rng):
    return LT.weights(9) * (1.0 + 0.2 * rng.standard_normal(9))


@pytest.mark.parametrize("seed", range(5))
def test_zouhe_lid_imposes_velocity_and_normal_nonequilibrium_bounce_back(oracle, seed):
    rng = np.random.default_rng(seed)
    u0, v0 = 0.1 * rng.standard_normal(2)
    s, ind = one_node(oracle, oracle.VEL_N, u0=u0, v0=v0)
    f = random_f(rng)
    g = s.open_boundary(ind, f)
    known = [0, 1, 2, 3, 5, 6]
    assert (g[known] == f[known]).all()
    rho, ux, uy = moments(g)
    assert abs(ux - u0) < 1e-14 and abs(uy - v0) < 1e-14
    feq = oracle.equilibrium(9, rho, (u0, v0))
    assert abs((g[4] - feq[4]) - (g[2] - feq[2])) < 1e-15


def test_reference_lid_limits(oracle):
    rng = np.random.default_rng(3)
    f = random_f(rng)
    # at rest it is on-node bounce-back of the three links leaving the face
    s, ind = one_node(oracle, oracle.VEL_N, inlet_model=oracle.INLET_REF2D, u0=0.0)
    g = s.open_boundary(ind, f)
    assert g[4] == f[2] and g[7] == f[5] and g[8] == f[6]
    # an equilibrium moving with the lid is kept (the omitted transverse
    # terms cancel there), and a non-equilibrium state is not given u0
    u0 = 0.08
    s, ind = one_node(oracle, oracle.VEL_N, inlet_model=oracle.INLET_REF2D, u0=u0)
    feq = oracle.equilibrium(9, 1.03, (u0, 0.0))
    assert np.abs(s.open_boundary(ind, feq) - feq).max() < 1e-16
    g = s.open_boundary(ind, f)
    assert g[4] == f[2] and abs(g[7] + g[8] - f[5] - f[6]) < 1e-16
    assert abs(moments(g)[1] - u0) > 1e-3
    # corner nodes (any wall link) are left to the walls
    s, ind = one_node(oracle, oracle.VEL_N, bb=1 << 3, inlet_model=oracle.INLET_REF2D, u0=u0)
    assert (s.open_boundary(ind, f) == f).all()


@pytest.mark.parametrize("seed", range(3))
def test_reference_velocity_outlet(oracle, seed):
    rng = np.random.default_rng(seed)
    f = random_f(rng)
    s, ind = one_node(oracle, oracle.PRES_E, outlet_model=oracle.OUTLET_REF_VELOCITY, u0=0.0)
    g = s.open_boundary(ind, f)
    assert g[3] == f[1] and g[7] == f[5] and g[6] == f[8]
    u0 = 0.05 + 0.05 * rng.random()
    s, ind = one_node(oracle, oracle.PRES_E, outlet_model=oracle.OUTLET_REF_VELOCITY, u0=u0)
    g = s.open_boundary(ind, f)
    known = [0, 1, 2, 4, 5, 8]
    assert (g[known] == f[known]).all()
    # the (1 - u0) density denominator makes the node move at u0 / (1 - 2 u0)
    # with density (1 - 2 u0) / (1 - u0) times the textbook closure's sum
    rho, ux, _ = moments(g)
    assert abs(ux - u0 / (1 - 2 * u0)) < 1e-14
    known_sum = f[0] + f[2] + f[4] + 2 * (f[1] + f[5] + f[8])
    assert abs(rho - known_sum * (1 - 2 * u0) / (1 - u0)) < 1e-14


def test_extrapolating_outlets_copy_upstream_nodes(oracle):
    """One step from the same state under the pressure, first- and
    second-order outlets: only the outlet nodes' f1, f5, f8 differ, and they
    are f(x-1) or 2 f(x-1) - f(x-2) of the same step's upstream nodes."""
    ny, nx = 10, 24
    ch = LT.channel2d(nx, ny, 0.05, obstacles=[(12.0, 5.0, 2.5)])
    rng = np.random.default_rng(11)
    u = 0.03 * rng.random((ny, nx))
    v = 0.01 * rng.standard_normal((ny, nx))
    rho = 1 + 0.01 * rng.standard_normal((ny, nx))
    out = {}
    for model in (oracle.OUTLET_PRESSURE, oracle.OUTLET_EXTRAP1, oracle.OUTLET_EXTRAP2):
        s = oracle.Sim(9, (ny, nx), ch["node_class"], ch["bb"], collision=1, omega=1 / 0.8, ubc=ch["ubc"],
                       vbc=ch["vbc"], outlet_model=model)
        s.init(0, rho=rho, u=u, v=v)
        s.step(3)
        out[model] = (s.get_f()[:, 0], s.macros())
    fp = out[oracle.OUTLET_PRESSURE][0]
    f1 = out[oracle.OUTLET_EXTRAP1][0]
    f2 = out[oracle.OUTLET_EXTRAP2][0]
    fluid = ch["node_class"][:, -1] != oracle.SOLID
    # the first two steps differ only in the outlet column's state; after the
    # third, compare this step's outlet against its own upstream nodes
    for f, k in ((f1, 1), (f2, 2)):
        for q in (1, 5, 8):
            a, b = f[q, :, -2], f[q, :, -3]
            want = a if k == 1 else 2 * a - b
            assert np.abs((f[q, :, -1] - want)[fluid]).max() < 1e-15
    assert np.abs(fp[[1, 5, 8], :, -1] - f1[[1, 5, 8], :, -1]).max() > 1e-6
    # stored macros are the moments of the final populations
    for f, (r, ux, uy, _) in (out[oracle.OUTLET_EXTRAP1], out[oracle.OUTLET_EXTRAP2]):
        m = ch["node_class"] != oracle.SOLID
        rr = f.sum(0)[m]
        assert np.abs(rr - r[0][m]).max() < 1e-14
        assert np.abs((C[:, 0] @ f[:, m]) / rr - ux[0][m]).max() < 1e-14


def test_extrapolating_outlets_keep_a_uniform_flow(oracle):
    """A uniform stream through a periodic-in-y box with an extrapolating
    outlet and a velocity inlet stays uniform: x-invariant states are fixed
    points of both extrapolations."""
    ny, nx = 6, 16
    cls = np.ones((ny, nx), np.uint8)
    cls[:, 0] = oracle.VEL_W
    cls[:, -1] = oracle.PRES_E
    U = 0.04
    for model in (oracle.OUTLET_EXTRAP1, oracle.OUTLET_EXTRAP2):
        s = oracle.Sim(9, (ny, nx), cls, None, collision=1, omega=1.2, periodic=(0, 1, 0),
                       ubc=np.full((ny, nx), U), vbc=np.zeros((ny, nx)), outlet_model=model)
        s.init(0, u=np.full((ny, nx), U))
        s.step(200)
        rho, u, v, _ = s.macros()
        assert np.abs(u - U).max() < 1e-13 and np.abs(v).max() < 1e-13 and np.abs(rho - 1).max() < 1e-13


def test_cavity_re100_matches_ghia(oracle):
    """Lid-driven cavity, 64 x 64 nodes, Re = U (n-1) / nu = 100, on-node
    walls and the Zou-He lid, run to a steady state: the vertical centre-line
    velocity is within 0.02 U of Ghia et al. (1982)."""
    n, U = 64, 0.1
    cav = LT.cavity2d(n, U)
    nu = U * (n - 1) / 100.0
    s = oracle.Sim(9, (n, n), cav["node_class"], cav["bb"], collision=1, omega=1 / (3 * nu + 0.5),
                   wall_model=oracle.WALL_ONNODE, ubc=cav["ubc"], vbc=cav["vbc"])
    s.init(0)
    prev = None
    for _ in range(60):
        s.step(1000)
        u = s.macros()[1][0]
        if prev is not None and np.abs(u - prev).max() < 1e-7:
            break
        prev = u.copy()
    assert np.abs(u - prev).max() < 1e-7
    assert ghia_centreline_error(u, U) < 0.02
    # the reference's lid, exact on equilibria only, reproduces the data as well
    s = oracle.Sim(9, (n, n), cav["node_class"], cav["bb"], collision=1, omega=1 / (3 * nu + 0.5),
                   wall_model=oracle.WALL_ONNODE, inlet_model=oracle.INLET_REF2D, ubc=cav["ubc"],
                   vbc=cav["vbc"])
    s.init(0)
    s.step(15000)
    assert ghia_centreline_error(s.macros()[1][0], U) < 0.02


def test_reader_classifies_lid_and_corner_links(tmp_path):
    """latticegen2 files of the cavity: INLE links in direction 2 on the top
    row make lid nodes; the top corners, with wall and inlet links, get their
    diagonal links turned into walls (S/GpuInit.cu:1073-1116)."""
    from thesis_amd import lbm as L
    n = 8
    cav = LT.cavity2d(n, 0.1)
    LT.write_latticegen2(str(tmp_path / "cav"), cav["node_class"], cav["bb"])
    r = L.read_lattice(str(tmp_path / "cav_Node.dat"), str(tmp_path / "cav_BC.dat"))
    cls, bb = r["node_class"][0], r["bb_links"][0]
    assert (cls == cav["node_class"]).all()
    # inner lid nodes keep only their own (none) wall links; corners keep the side walls
    assert (bb[-1, 1:-1] == 0).all()
    assert bb[-1, 0] == cav["bb"][-1, 0] and bb[-1, -1] == cav["bb"][-1, -1]
    # a lid node whose north-west diagonal was written as an inlet link
    (tmp_path / "c2_Node.dat").write_text("".join(f"{i} {j} {i}.0 {j}.0 1\n" for j in range(3) for i in range(3)))
    (tmp_path / "c2_BC.dat").write_text("0 2 3 1 -0.5 2.0 1\n0 2 2 2 0.0 2.5 1\n0 2 6 2 -0.5 2.5 1\n"
                                        "1 2 2 2 1.0 2.5 1\n")
    r = L.read_lattice(str(tmp_path / "c2_Node.dat"), str(tmp_path / "c2_BC.dat"))
    assert r["node_class"][0, 2, 1] == L.VEL_N and r["node_class"][0, 2, 0] == L.VEL_N
    assert r["bb_links"][0, 2, 0] == (1 << 3) | (1 << 6) and r["bb_links"][0, 2, 1] == 0

"""HIP path against the FP64 oracle for curved (Bouzidi) walls: a channel
with walls off the half-way positions, a cylinder and a sphere with the wall
fractions of their exact surfaces, FP32 storage, and lbmsolver with
Walls = CURVED reading the fractions from latticegen2 coordinates."""
import os
import subprocess

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L
from lattice_checks import sphere_wall_q
from test_gpu_parity import compare, pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "thesis_amd", "lbmsolver")


def shifted_walls(ny, nx, d1, d2):
    c = LT.velocities(9)
    bb = np.zeros((ny, nx), np.uint32)
    q = np.zeros((ny, nx, 19))
    for d in range(1, 9):
        if c[d, 1] < 0:
            bb[0] |= 1 << d
            q[0, :, d] = d1
        if c[d, 1] > 0:
            bb[-1] |= 1 << d
            q[-1, :, d] = d2
    return np.ones((ny, nx), np.uint8), bb, q


@pytest.mark.parametrize("coll", [L.BGK, L.TRT, L.MRT])
def test_shifted_channel_f64(oracle, coll):
    cls, bb, q = shifted_walls(12, 8, 0.3, 0.7)
    g, o = pair(oracle, 9, (12, 8), cls, bb, collision=coll, omega=1 / 0.8, force_model=L.FORCE_GUO,
                g=(1e-5, 0, 0), periodic=(1, 0, 0), wall_model=L.WALL_CURVED, wall_q=q)
    g.init(0)
    o.init(0)
    compare(g, o, 400)


def cylinder_case():
    ny, nx = 24, 48
    solid, bb, q = sphere_wall_q((ny, nx), (15.3, 11.7), 5.4)
    cls = np.where(solid, L.SOLID, L.FLUID).astype(np.uint8)
    wall = LT.bb_from_solid(np.zeros((ny, nx), bool), 9, periodic=(True, False, True),
                            outside_is_wall=(False, True, False))
    return cls, bb | wall, q


@pytest.mark.parametrize("coll", [L.BGK, L.TRT])
def test_cylinder_in_channel_f64(oracle, coll):
    """Force-driven x-periodic channel (half-way walls at y = -1/2, ny - 1/2,
    fraction 0 meaning 1/2) past a cylinder off the lattice nodes."""
    cls, bb, q = cylinder_case()
    g, o = pair(oracle, 9, cls.shape, cls, bb, collision=coll, omega=1 / 0.7, force_model=L.FORCE_GUO,
                g=(2e-5, 0, 0), periodic=(1, 0, 0), wall_model=L.WALL_CURVED, wall_q=q)
    g.init(0)
    o.init(0)
    compare(g, o, 300)


def test_sphere_in_periodic_box_f64(oracle):
    n = 16
    solid, bb, q = sphere_wall_q((n, n, n), (7.4, 8.2, 7.9), 4.3, q=19)
    cls = np.where(solid, L.SOLID, L.FLUID).astype(np.uint8)
    g, o = pair(oracle, 19, (n, n, n), cls, bb, collision=L.TRT, omega=1 / 0.8, trt_magic=0.25,
                force_model=L.FORCE_GUO, g=(1e-5, 0, 0), periodic=(1, 1, 1), wall_model=L.WALL_CURVED, wall_q=q)
    g.init(0)
    o.init(0)
    compare(g, o, 120)


def test_cylinder_f32_close_to_f64(oracle):
    cls, bb, q = cylinder_case()
    g, o = pair(oracle, 9, cls.shape, cls, bb, storage=L.STORE_F32, collision=L.TRT, omega=1 / 0.7,
                force_model=L.FORCE_GUO, g=(2e-5, 0, 0), periodic=(1, 0, 0), wall_model=L.WALL_CURVED, wall_q=q)
    g.init(0)
    o.init(0)
    g.step(300)
    o.step(300)
    fl = cls != L.SOLID
    gu, ou = g.macros()[1], o.macros()[1][0]
    e = np.abs(gu[fl] - ou[fl]).max() / np.abs(ou[fl]).max()
    print(f"f32 curved: relative velocity error {e:.3e}", flush=True)
    assert e < 5e-5


INI = """Lattice = "cyl"
U= 0.02
V= 0.0
W= 0.0
Density = 1.0
viscosity = 0.1
InletProfile = INLET_PROF
CollisionModel = TRT
Walls = CURVED
BCwallModel = BBack
BCoutlet = HEmodel
Number of iterations = 300
AutosaveAfter = 0
AutosaveEach = 0
MinMacroDiffs(u_v_w_rho)(StopCond) = 0.0 0.0 0.0 0.0
OutputFormat = CSV
InitialConditionFromFile = no
InitialFile = "none"
ResidualModel = L2
ResidualsAndMacroDiffsAfter = 100
Force = 0
DragAndLift = 0
UpdateInltOutl = yes
[mi355x]
init = equilibrium
"""


def test_lbmsolver_curved_walls(tmp_path):
    """Walls = CURVED: lbmsolver reads the wall fractions from the BC file's
    coordinates (S/GpuInit.cu:1088) and equals the ABI run with the same
    fractions."""
    print("case: lbmsolver CURVED", flush=True)
    ny, nx, U = 24, 64, 0.02
    solid, bbc, q = sphere_wall_q((ny, nx), (20.3, 11.6), 4.6)
    ch = LT.channel2d(nx, ny, 1.5 * U)
    cls = np.where(solid, L.SOLID, ch["node_class"]).astype(np.uint8)
    bb = LT.bb_from_solid(solid, 9, periodic=(False, False, True), outside_is_wall=(False, True, False))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "cyl"), cls, bb, wall_q=q)
    (tmp_path / "SetUpData.ini").write_text(INI)
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    node, bcf = str(tmp_path / "Mesh" / "cyl_Node.dat"), str(tmp_path / "Mesh" / "cyl_BC.dat")
    lat = L.read_lattice(node, bcf)
    wq = L.read_wall_q(node, bcf)[0]
    assert np.abs(wq[..., 1:9][q[..., 1:9] > 0] - q[..., 1:9][q[..., 1:9] > 0]).max() < 1e-5
    lcls = lat["node_class"][0]
    ys = np.arange(ny)
    ubc = np.zeros((ny, nx))
    ubc[:, 0] = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
    s = L.Solver(9, (ny, nx), lcls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * 0.1 + 0.5),
                 wall_model=L.WALL_CURVED, wall_q=wq, ubc=np.where(lcls == L.VEL_W, ubc, 0.0),
                 vbc=np.zeros((ny, nx)), rho_out=1.0)
    s.init(0, rho=np.ones((ny, nx)), u=np.where(lcls == L.VEL_W, ubc, U))
    s.step(300)
    rho, u, v, w = s.macros()
    assert np.abs(data[:, 3] - u.ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - rho.ravel()).max() < 1e-9

"""HIP path against the FP64 oracle for the 2D north lid and the reference's
east outlets (velocity outlet, second- and first-order extrapolation), the
lid-driven cavity against Ghia et al. (1982), and lbmsolver's reading of
BCoutlet.  FP64 storage at the 1e-12 bar of the other parity tests."""
import os
import subprocess

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L
from lattice_checks import ghia_centreline_error
from test_gpu_parity import compare, pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "thesis_amd", "lbmsolver")


@pytest.mark.parametrize("inlet", [L.INLET_ZOUHE, L.INLET_REF2D])
@pytest.mark.parametrize("coll", [L.BGK, L.TRT, L.MRT])
@pytest.mark.parametrize("path", ["rank", "step"])
def test_lid_cavity_f64(oracle, monkeypatch, inlet, coll, path):
    """24 x 24 cavity, on-node walls, lid at 0.1; both step kernels."""
    if path == "step":
        monkeypatch.setenv("LBM_RANK", "0")
    n = 24
    cav = LT.cavity2d(n, 0.1)
    g, o = pair(oracle, 9, (n, n), cav["node_class"], cav["bb"], collision=coll, omega=1 / 0.7,
                wall_model=L.WALL_ONNODE, inlet_model=inlet, ubc=cav["ubc"], vbc=cav["vbc"])
    g.init(0)
    o.init(0)
    compare(g, o, 300)


def test_lid_with_normal_velocity_halfway_walls_f64(oracle):
    """Textbook lid with a wall-normal component and half-way walls."""
    n = 20
    cav = LT.cavity2d(n, 0.08)
    vbc = np.where(cav["node_class"] == L.VEL_N, -0.01, 0.0)
    g, o = pair(oracle, 9, (n, n), cav["node_class"], cav["bb"], collision=L.TRT, omega=1 / 0.8,
                ubc=cav["ubc"], vbc=vbc)
    g.init(0)
    o.init(0)
    compare(g, o, 100)


def outlet_channel(model):
    ny, nx = 16, 40
    ch = LT.channel2d(nx, ny, 0.05, obstacles=((14, 7, 3),))
    ubc = ch["ubc"].copy()
    if model == L.OUTLET_REF_VELOCITY:
        ubc[:, -1] = ch["profile"]
    return ch, ubc, (ny, nx)


@pytest.mark.parametrize("model", [L.OUTLET_REF_VELOCITY, L.OUTLET_EXTRAP2, L.OUTLET_EXTRAP1])
@pytest.mark.parametrize("wall", [L.WALL_HALFWAY, L.WALL_ONNODE])
def test_outlet_models_f64(oracle, model, wall):
    ch, ubc, shape = outlet_channel(model)
    g, o = pair(oracle, 9, shape, ch["node_class"], ch["bb"], collision=L.TRT, omega=1 / 0.8, ubc=ubc,
                vbc=ch["vbc"], wall_model=wall, outlet_model=model)
    rng = np.random.default_rng(5)
    u = 0.03 + 0.01 * rng.random(shape)
    g.init(0, u=u)
    o.init(0, u=u)
    compare(g, o, 300)


@pytest.mark.parametrize("model", [L.OUTLET_EXTRAP2, L.OUTLET_EXTRAP1])
def test_outlet_models_f32_close_to_f64(oracle, model):
    ch, ubc, shape = outlet_channel(model)
    g, o = pair(oracle, 9, shape, ch["node_class"], ch["bb"], storage=L.STORE_F32, collision=L.TRT,
                omega=1 / 0.8, ubc=ubc, vbc=ch["vbc"], outlet_model=model)
    g.init(0)
    o.init(0)
    g.step(300)
    o.step(300)
    fl = ch["node_class"] != L.SOLID
    gu, ou = g.macros()[1], o.macros()[1][0]
    e = np.abs(gu[fl] - ou[fl]).max() / np.abs(ou[fl]).max()
    print(f"f32 outlet {model}: relative velocity error {e:.3e}", flush=True)
    assert e < 5e-5


def test_cavity_re100_matches_ghia_128():
    """128 x 128 cavity at Re = 100 to a steady state on the GPU: the
    centre-line velocity is within 0.015 U of Ghia et al. (1982) (the
    oracle's 128 x 128 state is 0.0094 U off, its 64 x 64 one 0.015 U)."""
    print("case: cavity 128 Re 100", flush=True)
    n, U = 128, 0.1
    cav = LT.cavity2d(n, U)
    nu = U * (n - 1) / 100.0
    s = L.Solver(9, (n, n), cav["node_class"], cav["bb"], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                 wall_model=L.WALL_ONNODE, ubc=cav["ubc"], vbc=cav["vbc"])
    s.init(0)
    steps, res = s.run_to_residual(1e-9, 1000, 200000)
    u = s.macros()[1]
    err = ghia_centreline_error(u, U)
    print(f"steps {steps} residual {res:.2e} Ghia error {err:.4f}", flush=True)
    assert res < 1e-9 and err < 0.015


INI = """Lattice = "{name}"
U= {U}
V= 0.0
W= 0.0
Density = 1.0
viscosity = {nu}
InletProfile = {profile}
CollisionModel = TRT
Walls = STRAIGHT
BCwallModel = BBack
BCoutlet = {outlet}
Number of iterations = {iters}
AutosaveAfter = 0
AutosaveEach = 0
MinMacroDiffs(u_v_w_rho)(StopCond) = 0.0 0.0 0.0 0.0
OutputFormat = CSV
InitialConditionFromFile = no
InitialFile = "none"
ResidualModel = L2
ResidualsAndMacroDiffsAfter = 1000
Force = 0
DragAndLift = 0
UpdateInltOutl = yes
[mi355x]
init = equilibrium
"""


@pytest.mark.parametrize("token,model", [("Vout", L.OUTLET_REF_VELOCITY), ("1stE", L.OUTLET_EXTRAP2),
                                         ("2ndE", L.OUTLET_EXTRAP1)])
def test_lbmsolver_bcoutlet_maps_to_reference_outlets(tmp_path, token, model):
    """BCoutlet tokens as the reference reads them (S/FilesReading.cu:180-190:
    1stE selects OUTLET_SECOND, 2ndE OUTLET_FIRST); the velocity outlet's u0
    is the inlet profile evaluated on the outlet nodes (S/GpuInit.cu:994-1006)."""
    print(f"case: lbmsolver BCoutlet {token}", flush=True)
    nx, ny, U, nu, iters = 48, 16, 0.02, 0.1, 400
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((18, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    (tmp_path / "SetUpData.ini").write_text(INI.format(name="chan", U=U, nu=nu, iters=iters, outlet=token,
                                                       profile="INLET_PROF"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
    cls, bb = lat["node_class"][0], lat["bb_links"][0]
    ys = np.arange(ny)[:, None] * np.ones((1, nx))
    prof = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
    ubc = np.where((cls == L.VEL_W) | (cls == L.PRES_E), prof, 0.0)
    s = L.Solver(9, (ny, nx), cls, bb, collision=L.TRT, omega=1 / (3 * nu + 0.5), wall_model=L.WALL_ONNODE,
                 ubc=ubc, vbc=np.zeros((ny, nx)), outlet_model=model)
    s.init(0, rho=np.ones((ny, nx)), u=np.where(cls == L.VEL_W, ubc, U))
    s.step(iters)
    rho, u, v, w = s.macros()
    assert np.abs(data[:, 3] - u.ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - rho.ravel()).max() < 1e-9


def test_lbmsolver_lid_driven_cavity(tmp_path):
    """A latticegen2 cavity (INLE links on the top row) through lbmsolver with
    a constant inlet profile reaches Ghia et al.'s Re = 100 centre line."""
    print("case: lbmsolver cavity 64", flush=True)
    n, U = 64, 0.1
    cav = LT.cavity2d(n, U)
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "cav"), cav["node_class"], cav["bb"])
    nu = U * (n - 1) / 100.0
    (tmp_path / "SetUpData.ini").write_text(INI.format(name="cav", U=U, nu=nu, iters=30000, outlet="HEmodel",
                                                       profile="INLET_CTE"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    u = data[:, 3].reshape(n, n)
    err = ghia_centreline_error(u, U)
    print(f"Ghia error {err:.4f}", flush=True)
    assert err < 0.02

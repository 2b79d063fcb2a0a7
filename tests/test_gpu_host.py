"""End-to-end drop-in: the C++ lbmsolver reads SetUpData.ini and a latticegen2
lattice and its FinalData output equals the same run through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "thesis_amd", "lbmsolver")

INI = """Lattice = "chan"
U= {U}
V= 0.0
W= 0.0
Density = 1.0
viscosity = {nu}
InletProfile = INLET_PROF
CollisionModel = TRT
Walls = STRAIGHT
BCwallModel = BBack
BCoutlet = Vout
Number of iterations = {iters}
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
walls = halfway
init = equilibrium
"""


def test_lbmsolver_matches_abi_run(tmp_path):
    print("case: lbmsolver 2D channel 64x16 TRT", flush=True)
    nx, ny, U, nu, iters = 64, 16, 0.02, 0.1, 500
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((20, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    (tmp_path / "SetUpData.ini").write_text(INI.format(U=U, nu=nu, iters=iters))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    assert data.shape == (nx * ny, 9)
    assert (tmp_path / "Results" / "residuals.dat").exists()
    # same lattice through the ABI: reread the written files, same profile rule
    lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
    cls = lat["node_class"][0]
    assert (cls == ch["node_class"]).all()
    ys = np.arange(ny)
    ubc = np.zeros((ny, nx))
    lo, hi = -0.5, ny - 0.5
    ubc[:, 0] = 6 * U * (ys - lo) * (hi - ys) / (hi - lo) ** 2
    s = L.Solver(9, (ny, nx), cls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                 ubc=np.where(cls == L.VEL_W, ubc, 0.0), vbc=np.zeros((ny, nx)), rho_out=1.0)
    u0 = np.where(cls == L.VEL_W, ubc, U)
    s.init(0, rho=np.ones((ny, nx)), u=u0)
    s.step(iters)
    rho, u, v, w = s.macros()
    assert np.abs(data[:, 3] - u.ravel()).max() < 1e-9     # CSV keeps 10 decimals
    assert np.abs(data[:, 7] - rho.ravel()).max() < 1e-9

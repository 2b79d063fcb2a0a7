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
outlet = pressure
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
    # runtimes.dat and lbmsolver.log in the reference's layout (S/LogWriter.cu:5-188)
    keys = [l.rsplit(" ", 1)[0] for l in (tmp_path / "Results" / "runtimes.dat").read_text().splitlines()]
    assert keys == ["Overall", "Iteration", "Initialisation", "Collision", "Streaming", "Boundaries", "Macroscopic",
                    "Residuals", "File writing"]
    log = (tmp_path / "Results" / "lbmsolver.log").read_text()
    for line in (">>> CollisionModel    : TRT", ">>> # of nodes in x (n)      = 64", ">>> NumInletNodes            = 16",
                 ":::: Start Iterations ::::", "Main while loop took", ":::: Iterations done! ::::"):
        assert line in log, line
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


@pytest.mark.parametrize("fmt,ext", [("CSV", "csv"), ("DAT", "dat"), ("VTI", "vti")])
def test_lbmsolver_restart_from_results_file(tmp_path, fmt, ext):
    """InitialConditionFromFile (S/FilesReading.cu:224-268): a run restarted
    from its own FinalData file with 0 further iterations writes the same
    fields back, in each output format; a restart that then continues equals
    the uninterrupted run from the same fields through the C ABI (f = 0 start
    with those macros, the reference's restart)."""
    print(f"case: lbmsolver restart {fmt}", flush=True)
    nx, ny, U, nu = 64, 16, 0.02, 0.1
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((20, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    ini = INI.format(U=U, nu=nu, iters=300).replace("OutputFormat = CSV", f"OutputFormat = {fmt}")
    (tmp_path / "SetUpData.ini").write_text(ini)
    r = subprocess.run([CLI, "-f", "SetUpData.ini", "-o", "Run1/"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    first = tmp_path / "Run1" / f"FinalData.{ext}"
    ini2 = (ini.replace("InitialConditionFromFile = no", "InitialConditionFromFile = yes")
               .replace('InitialFile = "none"', f'InitialFile = "{first}"')
               .replace("Number of iterations = 300", "Number of iterations = 0")
               .replace("init = equilibrium\n", ""))
    (tmp_path / "restart.ini").write_text(ini2)
    r = subprocess.run([CLI, "-f", "restart.ini", "-o", "Run2/"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and "Initial conditions loaded" in r.stdout
    r = subprocess.run([CLI, "compare", str(first), str(tmp_path / "Run2" / f"FinalData.{ext}")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    if fmt == "CSV":
        (tmp_path / "restart200.ini").write_text(ini2.replace("Number of iterations = 0", "Number of iterations = 200"))
        r = subprocess.run([CLI, "-f", "restart200.ini", "-o", "Run3/"], cwd=tmp_path, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0
        d0 = np.loadtxt(first, delimiter=",", skiprows=1)
        d3 = np.loadtxt(tmp_path / "Run3" / "FinalData.csv", delimiter=",", skiprows=1)
        lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
        cls = lat["node_class"][0]
        ys = np.arange(ny)
        ubc = np.zeros((ny, nx))
        ubc[:, 0] = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
        s = L.Solver(9, (ny, nx), cls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                     ubc=np.where(cls == L.VEL_W, ubc, 0.0), vbc=np.zeros((ny, nx)), rho_out=1.0)
        s.init(1, rho=d0[:, 7].reshape(ny, nx), u=d0[:, 3].reshape(ny, nx), v=d0[:, 4].reshape(ny, nx))
        s.step(200)
        rho, u, v, w = s.macros()
        assert np.abs(d3[:, 3] - u.ravel()).max() < 1e-9 and np.abs(d3[:, 7] - rho.ravel()).max() < 1e-9
    # a wrong file falls back to the uniform start with a message
    (tmp_path / "bad.ini").write_text(ini2.replace(str(first), str(tmp_path / "missing.csv")))
    r = subprocess.run([CLI, "-f", "bad.ini", "-o", "Run4/"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "uniform start" in r.stderr


def test_lbmsolver_macro_diff_stop_condition(tmp_path):
    """ResidualModel = MaxMacroDiff with per-field thresholds
    (S/Iterate3D.cu:696-756): lbmsolver stops at the first check where the
    largest changes of u, v, w and rho since the previous check are all below
    them, logs every check in MacroDiffs.txt (S/LogWriter.cu:215-221), and the
    state it writes equals the ABI run of that many steps."""
    print("case: lbmsolver MaxMacroDiff", flush=True)
    nx, ny, U, nu = 64, 16, 0.02, 0.1
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((20, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    ini = (INI.format(U=U, nu=nu, iters=40000)
           .replace("MinMacroDiffs(u_v_w_rho)(StopCond) = 0.0 0.0 0.0 0.0",
                    "MinMacroDiffs(u_v_w_rho)(StopCond) = 1.0e-7 1.0e-7 1.0e-7 1.0e-7")
           .replace("ResidualModel = L2", "ResidualModel = MaxMacroDiff"))
    (tmp_path / "SetUpData.ini").write_text(ini)
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and "simulation converged!" in r.stdout
    lines = (tmp_path / "Results" / "MacroDiffs.txt").read_text().splitlines()
    assert lines[0] == "iteration\tMax u diff\tMax v diff\tMax w diff\tMax rho diff"   # S/main.cu:125-127
    lines = lines[1:]
    import re
    assert all(re.fullmatch(r"\d+\t\t\d+\.\d{10}\t\d+\.\d{10}\t\d+\.\d{10}\t\d+\.\d{10}", ln) for ln in lines)
    last = lines[-1].split()
    steps = int(last[0])
    assert steps < 40000 and steps % 100 == 0
    assert all(float(x) < 1e-7 for x in last[1:])
    assert any(float(x) >= 1e-7 for x in lines[-2].split()[1:])
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
    cls = lat["node_class"][0]
    ys = np.arange(ny)
    ubc = np.zeros((ny, nx))
    ubc[:, 0] = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
    s = L.Solver(9, (ny, nx), cls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                 ubc=np.where(cls == L.VEL_W, ubc, 0.0), vbc=np.zeros((ny, nx)), rho_out=1.0)
    s.init(0, rho=np.ones((ny, nx)), u=np.where(cls == L.VEL_W, ubc, U))
    s.step(steps)
    rho, u, v, w = s.macros()
    assert np.abs(data[:, 3] - u.ravel()).max() < 1e-9


def test_lbmsolver_fdreldiff_stop_condition(tmp_path):
    """ResidualModel = FdRelDiff (S/Iterate3D.cu:677-693): lbmsolver logs
    max |f - f_prev| / f1 at every check in residuals.dat (1 at the first
    check), stops when it falls below the first stop value, and the values
    equal those of the ABI run checked at the same steps."""
    print("case: lbmsolver FdRelDiff", flush=True)
    nx, ny, U, nu = 64, 16, 0.02, 0.1
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((20, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    ini = (INI.format(U=U, nu=nu, iters=40000)
           .replace("MinMacroDiffs(u_v_w_rho)(StopCond) = 0.0 0.0 0.0 0.0",
                    "MinMacroDiffs(u_v_w_rho)(StopCond) = 1.0e-6 0.0 0.0 0.0")
           .replace("ResidualModel = L2", "ResidualModel = FdRelDiff"))
    (tmp_path / "SetUpData.ini").write_text(ini)
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and "simulation converged!" in r.stdout
    rows = np.loadtxt(tmp_path / "Results" / "residuals.dat", skiprows=1)
    assert rows[0, 1] == 1.0 and rows[-1, 1] < 1e-6 <= rows[-2, 1]
    lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
    cls = lat["node_class"][0]
    ys = np.arange(ny)
    ubc = np.zeros((ny, nx))
    ubc[:, 0] = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
    s = L.Solver(9, (ny, nx), cls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                 ubc=np.where(cls == L.VEL_W, ubc, 0.0), vbc=np.zeros((ny, nx)), rho_out=1.0)
    s.init(0, rho=np.ones((ny, nx)), u=np.where(cls == L.VEL_W, ubc, U))
    every = int(rows[1, 0] - rows[0, 0])
    got = []
    for _ in range(len(rows)):
        s.step(every)
        got.append(s.residual_fdrel())
    assert np.allclose(rows[:, 1], got, rtol=1e-4, atol=0)


def test_lbmsolver_binary_lattice(tmp_path):
    """With Mesh/<Lattice>.lbl present lbmsolver reads the binary copy; the
    results equal those from the text pair."""
    print("case: lbmsolver binary lattice", flush=True)
    nx, ny, U, nu = 48, 14, 0.02, 0.1
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((15, 7, 2.5),))
    out = {}
    for mode in ("text", "binary"):
        d = tmp_path / mode
        os.makedirs(d / "Mesh")
        LT.write_latticegen2(str(d / "Mesh" / "chan"), ch["node_class"], ch["bb"])
        if mode == "binary":
            L.write_lattice_binary(str(d / "Mesh" / "chan_Node.dat"), str(d / "Mesh" / "chan_BC.dat"),
                                   str(d / "Mesh" / "chan.lbl"))
            os.remove(d / "Mesh" / "chan_BC.dat")      # the binary copy alone must do
        (d / "SetUpData.ini").write_text(INI.format(U=U, nu=nu, iters=200))
        r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=d, capture_output=True, text=True, timeout=120)
        print(r.stdout[-200:], r.stderr[-300:])
        assert r.returncode == 0
        out[mode] = (d / "Results" / "FinalData.csv").read_bytes()
    assert out["text"] == out["binary"]


def read_appended_vti(path, n):
    """Velocity (n x 3) and density (n) of a binary-appended VTI file."""
    raw = open(path, "rb").read()
    p = raw.index(b'<AppendedData encoding="raw">')
    p = raw.index(b"_", p) + 1
    out = []
    for count in (3 * n, n, n):
        nb = int(np.frombuffer(raw[p:p + 4], np.uint32)[0])
        assert nb == 4 * count
        out.append(np.frombuffer(raw[p + 4:p + 4 + nb], np.float32).astype(np.float64))
        p += 4 + nb
    return out[0].reshape(n, 3), out[1]


def test_lbmsolver_autosave_and_appended_vti(tmp_path):
    """AutosaveAfter / AutosaveEach (S/Iterate.cu:590-621): a 2D ParaView run
    writes Video/FinalData_<k>.vti at every multiple of AutosaveEach beyond
    AutosaveAfter; with vti = appended all VTI files carry Float32 binary
    data, the final one equals the ABI run to Float32 rounding, and a restart
    from it with no further iterations writes the same fields back."""
    print("case: lbmsolver autosave + appended VTI", flush=True)
    nx, ny, U, nu = 64, 16, 0.02, 0.1
    ch = LT.channel2d(nx, ny, 1.5 * U, obstacles=((20, 8, 3),))
    os.makedirs(tmp_path / "Mesh")
    LT.write_latticegen2(str(tmp_path / "Mesh" / "chan"), ch["node_class"], ch["bb"])
    ini = (INI.format(U=U, nu=nu, iters=600).replace("OutputFormat = CSV", "OutputFormat = VTI")
           .replace("AutosaveAfter = 0", "AutosaveAfter = 100").replace("AutosaveEach = 0", "AutosaveEach = 200")
           + "vti = appended\n")
    (tmp_path / "SetUpData.ini").write_text(ini)
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and r.stdout.count("autosave") == 3
    video = sorted(os.listdir(tmp_path / "Results" / "Video"))
    assert video == ["FinalData_2.vti", "FinalData_3.vti", "FinalData_4.vti"]
    n = nx * ny
    vel, rho = read_appended_vti(tmp_path / "Results" / "FinalData.vti", n)
    vel4, rho4 = read_appended_vti(tmp_path / "Results" / "Video" / "FinalData_4.vti", n)
    assert (vel == vel4).all() and (rho == rho4).all()
    lat = L.read_lattice(str(tmp_path / "Mesh" / "chan_Node.dat"), str(tmp_path / "Mesh" / "chan_BC.dat"))
    cls = lat["node_class"][0]
    ys = np.arange(ny)
    ubc = np.zeros((ny, nx))
    ubc[:, 0] = 6 * U * (ys + 0.5) * (ny - 0.5 - ys) / ny ** 2
    s = L.Solver(9, (ny, nx), cls, lat["bb_links"][0], collision=L.TRT, omega=1 / (3 * nu + 0.5),
                 ubc=np.where(cls == L.VEL_W, ubc, 0.0), vbc=np.zeros((ny, nx)), rho_out=1.0)
    s.init(0, rho=np.ones((ny, nx)), u=np.where(cls == L.VEL_W, ubc, U))
    s.step(600)
    rr, uu, vv, ww = s.macros()
    assert np.abs(vel[:, 0] - uu.ravel()).max() < 1e-8 and np.abs(rho - rr.ravel()).max() < 1e-6
    first = tmp_path / "Results" / "FinalData.vti"
    ini2 = (ini.replace("InitialConditionFromFile = no", "InitialConditionFromFile = yes")
               .replace('InitialFile = "none"', f'InitialFile = "{first}"')
               .replace("Number of iterations = 600", "Number of iterations = 0"))
    (tmp_path / "restart.ini").write_text(ini2)
    r = subprocess.run([CLI, "-f", "restart.ini", "-o", "Run2/"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "Initial conditions loaded" in r.stdout, r.stderr
    vel2, rho2 = read_appended_vti(tmp_path / "Run2" / "FinalData.vti", n)
    assert (vel2 == vel).all() and (rho2 == rho).all()


INI_MP = INI.replace('Lattice = "chan"', 'Lattice = "box"').replace("BCoutlet = Vout", "BCoutlet = 1stE") + \
    "multiphase = yes\ncg_bubble_radius = {R}\n"


@pytest.mark.parametrize("three_d", [True, False])
def test_lbmsolver_multiphase_bubble(tmp_path, three_d):
    """[mi355x] multiphase = yes: the reference's default colour-gradient
    steady bubble (S/main.cu:52-72) on the lattice's periodic box equals the
    ABI run from the same start, and the surface-tension error file holds one
    value per iteration."""
    print(f"case: lbmsolver multiphase 3D={three_d}", flush=True)
    os.makedirs(tmp_path / "Mesh")
    iters, R = 120, 5.0
    if three_d:
        # the reference's periodic cube (CYCL on all faces), 17^3 nodes: period 16
        star = os.path.join(ROOT, "tests", "golden", "starcd")
        dims, _ = L.latticegen(os.path.join(star, "periodicCube.vrt"), os.path.join(star, "periodicCube.bnd"), 17,
                               (0.5, 0.5, 0.5), str(tmp_path / "Mesh" / "box"))
        n = 16
        shape = (n, n, n)
    else:
        n = 32
        LT.write_latticegen2(str(tmp_path / "Mesh" / "box"), np.ones((n + 1, n + 1), np.uint8),
                             np.zeros((n + 1, n + 1), np.uint32), delta=1.0 / n)
        shape = (n, n)
    (tmp_path / "SetUpData.ini").write_text(INI_MP.format(U=0.0, nu=0.1, iters=iters, R=R)
                                            .replace("OutputFormat = CSV", "OutputFormat = CSV"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-300:])
    assert r.returncode == 0 and "Suface tension error" in r.stdout
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    if three_d:      # WriteArray: one value per line
        err = np.loadtxt(tmp_path / "Results" / "surface tension.txt")
    else:            # writeOscilatingSolution: "i value" per line
        rows = np.loadtxt(tmp_path / "Results" / "Oscilating_frequency.txt")
        assert (rows[:, 0] == np.arange(iters)).all()
        err = rows[:, 1]
        assert "Pressure inside" in r.stdout
    assert err.shape == (iters,) and np.isfinite(err).all()
    # the same run through the C ABI
    grids = np.meshgrid(*[np.arange(k) / n for k in shape], indexing="ij")
    d2 = sum((g - 0.5) ** 2 for g in grids)
    inside = np.sqrt(d2) <= R / (n + 1)
    r_rho = np.where(inside, 1.0, 0.0)
    b_rho = 1.0 - r_rho
    kw = dict(r_alpha=4 / 9, b_alpha=4 / 9, r_nu=0.5, b_nu=0.5, beta=0.99, A=1e-4, high_order=1, kappa=0.9)
    if three_d:
        g = L.CgSolver(n, n, nz=n, **kw)
        g.init_f(*LT.bubble3d_populations(r_rho, b_rho))
    else:
        g = L.CgSolver(n, n, **kw)
        g.init(r_rho, b_rho)
    g.step(iters)
    m = g.macros()
    red = g.reduce()
    g.close()
    full = tuple(k + 1 for k in shape)
    idx = np.ix_(*[np.arange(k) % (k - 1) for k in full])
    rho = m["rho"][idx].ravel()
    u = m["u"][idx].ravel()
    assert np.abs(data[:, 7] - rho).max() < 1e-9
    assert np.abs(data[:, 3] - u).max() < 1e-9
    om = 1 / (3 * 0.5 + 0.5)
    sigma = 4 * 1e-4 / 9 / om
    pf = 0.5 if three_d else 0.6
    pin = pf * (5 / 9) * red["pin_sum"] / red["pin_n"]
    pout = pf * (5 / 9) * red["pout_sum"] / red["pout_n"]
    assert abs(err[-1] - abs(sigma - R * (pin - pout)) / sigma * 100) < 1e-6


def _mp_case_ini(case, U, iters, extra):
    return (INI.replace('Lattice = "chan"', 'Lattice = "box"').replace("BCoutlet = Vout", "BCoutlet = 1stE")
            .format(U=U, nu=0.1, iters=iters) + f"multiphase = yes\ncg_test_case = {case}\n" + extra)


@pytest.mark.parametrize("case", [2, 3, 5, 6])
def test_lbmsolver_multiphase_test_cases(tmp_path, case):
    """[mi355x] cg_test_case: the reference's 2D two-phase cases through
    lbmsolver equal the ABI run from the same start, and the validators'
    files and prints hold what thesis_amd.cg_cases computes from it:
    Profile_Couette.txt (case 2), the radius of the relaxed square (3), the
    interface traces Oscilating_frequency.txt (5) and RT_Interface.txt (6)."""
    from thesis_amd import cg_cases as CC
    print(f"case: lbmsolver two-phase test case {case}", flush=True)
    os.makedirs(tmp_path / "Mesh")
    n = 32 if case in (2, 6) else 64
    m = 4 * n if case == 6 else n          # the Rayleigh-Taylor box is 1 x 4
    LT.write_latticegen2(str(tmp_path / "Mesh" / "box"), np.ones((m + 1, n + 1), np.uint8),
                         np.zeros((m + 1, n + 1), np.uint32), delta=1.0 / n)
    nu_r, nu_b, U, g, gamma, A = 1 / 6, 1 / 6, 0.0, 0.0, 1.0, 1e-3
    iters = {2: 3000, 3: 3000, 5: 2500, 6: 2000}[case]
    if case == 2:
        nu_b, U = 1 / 24, 0.01
    if case == 5:
        nu_r = nu_b = 0.02
        A = 0.01
    if case == 6:
        g, gamma = -5e-5, 3.0
    extra = (f"cg_r_viscosity = {nu_r!r}\ncg_b_viscosity = {nu_b!r}\ncg_A = {A!r}\ncg_beta = 0.7\n"
             f"cg_gamma = {gamma!r}\ncg_high_order = 0\n")
    (tmp_path / "SetUpData.ini").write_text(_mp_case_ini(case, U, iters, extra).replace("Force = 0", f"Force = {g!r}"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0
    # the same run through the C ABI: x periodic with period n, all m + 1
    # rows for the wall cases, period m otherwise
    ny = m + 1 if case in (2, 6) else m
    b_alpha = 4 / 9
    kw = dict(r_alpha=1 - (1 - b_alpha) / gamma, b_alpha=b_alpha, r_nu=nu_r, b_nu=nu_b, beta=0.7, A=A,
              high_order=0, kappa=0.9, walls={2: 1, 6: 2}.get(case, 0), u_wall=U, g=g, external_force=0)
    y, x = np.mgrid[0:ny, 0:n]
    xs, ys = x / n, y / n
    N, M = n + 1, m + 1                    # the reference's length and depth
    red = {2: ys > 0.5,
           3: (x < 0.75 * N) & (x > 0.25 * N) & (y < 0.75 * M) & (y > 0.25 * M),
           5: ((x - 0.5 * N) / (0.125 * N)) ** 2 + ((y - 0.5 * M) / (0.1875 * M)) ** 2 <= 1,
           6: ys > 2.0 + 0.1 * np.cos(2 * np.pi * xs)}[case].astype(float)
    s = L.CgSolver(n, ny, **kw)
    s.init(red, (1 - red) / gamma)
    full = np.ix_(np.arange(M) % ny, np.arange(N) % n)
    trace = []
    for _ in range(iters if case in (5, 6) else 1):
        s.step(1 if case in (5, 6) else iters)
        if case in (5, 6):
            mm = s.macros()
            fn = CC.max_y_oscillating if case == 5 else CC.min_y_rt
            trace.append(fn(mm["r_rho"][full], mm["b_rho"][full], np.arange(M) / n))
    mm = s.macros()
    s.close()
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    assert np.abs(data[:, 3] - mm["u"][full].ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - mm["rho"][full].ravel()).max() < 1e-9
    if case == 2:
        prof = np.loadtxt(tmp_path / "Results" / "Profile_Couette.txt", skiprows=1)
        assert prof.shape == (M, 3)
        yy = np.arange(M) / n
        assert np.abs(prof[:, 0] - np.round(yy, 4)).max() < 1e-9
        assert np.abs(prof[:, 1] - mm["u"][full][:, N // 2 if N % 2 == 0 else (N + 1) // 2]).max() < 1e-9
        assert np.abs(prof[:, 2] - CC.analytical_couette((nu_r / nu_b), yy, U)).max() < 1e-9
    elif case == 3:
        got, want = CC.deforming_radius(mm["r_rho"][full], mm["b_rho"][full])
        assert f"Predicted radius: {want:.6f} Final radius: {got:.6f}" in r.stdout
    else:
        name = "Oscilating_frequency" if case == 5 else "RT_Interface"
        rows = np.loadtxt(tmp_path / "Results" / f"{name}.txt")
        assert (rows[:, 0] == np.arange(iters)).all()
        assert np.abs(rows[:, 1] - np.array(trace)).max() < 1e-9
        if case == 5:
            assert "Theoretical" in r.stdout


@pytest.mark.parametrize("case", [2, 3, 4, 5])
def test_lbmsolver_multiphase_test_cases_3d(tmp_path, case):
    """[mi355x] cg_test_case on a 3D lattice: the reference's 3D two-phase
    cases (initCGBubble3D, S/GpuInit.cu:536-720) through lbmsolver equal the
    ABI run from the same start: the cube relaxing to a sphere (2, with the
    3D deformingBubbleValid print), two layers between inlet faces on y (4,
    Couette, Profile_Couette.txt of the middle column of the middle plane) and
    a heavy layer over a light one between bounce-back faces under gravity (5,
    Rayleigh-Taylor); two spheres merging (3, the 3D validateCoalescenceCase
    print).  The lattice is cyclic along x and z (and y for cases 2 and 3)."""
    from thesis_amd import cg_cases as CC
    print(f"case: lbmsolver two-phase 3D test case {case}", flush=True)
    os.makedirs(tmp_path / "Mesh")
    n = 16
    M = {2: n + 1, 3: n + 1, 4: 17, 5: 33}[case]
    cyc = (0, 1, 2) if case in (2, 3) else (0, 2)
    LT.write_latticegen2(str(tmp_path / "Mesh" / "box"), np.ones((n + 1, M, n + 1), np.uint8), delta=1.0 / n,
                         cyclic_axes=cyc)
    nu_r, nu_b, U, g, gamma, A = 1 / 6, 1 / 6, 0.0, 0.0, 1.0, 1e-3
    iters = {2: 400, 3: 400, 4: 1500, 5: 600}[case]
    if case == 4:
        nu_b, U = 1 / 24, 0.01
    if case == 5:
        g, gamma = -5e-5, 3.0
    extra = (f"cg_r_viscosity = {nu_r!r}\ncg_b_viscosity = {nu_b!r}\ncg_A = {A!r}\ncg_beta = 0.7\n"
             f"cg_gamma = {gamma!r}\ncg_high_order = 0\ncg_bubble_radius = 5.0\n")
    (tmp_path / "SetUpData.ini").write_text(_mp_case_ini(case, U, iters, extra).replace("Force = 0", f"Force = {g!r}"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0
    # the same run through the C ABI: x and z periodic with period n, y with
    # period n (case 2) or every plane between the walls
    ny = n if case in (2, 3) else M
    b_alpha = 4 / 9
    kw = dict(r_alpha=1 - (1 - b_alpha) / gamma, b_alpha=b_alpha, r_nu=nu_r, b_nu=nu_b, beta=0.7, A=A,
              high_order=0, kappa=0.9, walls={4: 1, 5: 2}.get(case, 0), u_wall=U, g=g, external_force=0)
    z, y, x = np.mgrid[0:n, 0:ny, 0:n]
    N = n + 1
    R = 5.0 / N                                # cg_bubble_radius 5 over the lattice width
    c2 = (x / n - 0.5) ** 2 + (z / n - 0.5) ** 2
    red = {2: (x < 0.75 * N) & (x > 0.25 * N) & (y < 0.75 * M) & (y > 0.25 * M) & (z < 0.75 * N) & (z > 0.25 * N),
           3: (np.sqrt(c2 + (y / n - 0.5 + R) ** 2) <= R) | (np.sqrt(c2 + (y / n - 0.5 - R) ** 2) <= R),
           4: y / n > 0.5,
           5: y / n > 1.0 + 0.1 * np.cos(2 * np.pi * x / n)}[case].astype(float)
    s = L.CgSolver(n, ny, nz=n, **kw)
    s.init_f(*LT.bubble3d_populations(red, (1 - red) / gamma))
    s.step(iters)
    mm = s.macros()
    s.close()
    full = np.ix_(np.arange(N) % n, np.arange(M) % ny, np.arange(N) % n)
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    assert np.abs(data[:, 3] - mm["u"][full].ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - mm["rho"][full].ravel()).max() < 1e-9
    mid = N // 2 if N % 2 == 0 else (N + 1) // 2
    if case == 2:
        got, want = CC.deforming_radius(mm["r_rho"][full], mm["b_rho"][full])
        assert abs(want - N / 2 * 0.6203504908994) < 1e-9 * N  # (3 / (4 pi))^(1/3) of the lattice's half width
        assert f"Predicted radius: {want:.6f} Final radius: {got:.6f}" in r.stdout
    elif case == 3:
        got, want = CC.coalescence_radius(mm["r_rho"][full], mm["b_rho"][full], R)
        assert abs(want - R * 2 ** (1 / 3)) < 1e-12
        assert f"Predicted radius: {want:.6f} Final radius: {got:.6f}  (physical units)" in r.stdout
    elif case == 4:
        prof = np.loadtxt(tmp_path / "Results" / "Profile_Couette.txt", skiprows=1)
        assert prof.shape == (M, 3)
        yy = np.arange(M) / n
        assert np.abs(prof[:, 0] - np.round(yy, 4)).max() < 1e-9
        assert np.abs(prof[:, 1] - mm["u"][full][mid, :, mid]).max() < 1e-9
        assert np.abs(prof[:, 2] - CC.analytical_couette(nu_r / nu_b, yy, U)).max() < 1e-9
    else:
        # gravity sets the layers in motion; no plane is left untouched
        assert np.isfinite(mm["v"]).all() and np.abs(mm["v"]).max() > 1e-6

"""HIP path against the FP64 oracle for the D3Q19 boundaries: velocity faces,
wet-node walls (faces, edges, corners), outlets and symmetry planes; the
reference's own 30^3 lid-driven cavity lattice (L/lidvalidation) with
wet-node walls; the square duct against its series solution; lbmsolver on
the reference lattice.  FP64 storage at the 1e-12 bar of the other parity
tests."""
import gzip
import os
import subprocess

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L
from lattice_checks import square_duct_profile
from test_gpu_parity import compare, pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "thesis_amd", "lbmsolver")

# (velocity face, opposite outlet, axis, inward sign of the velocity face)
PAIRS = [(L.VEL_W, L.PRES_E, 0, 1), (L.VEL_E, L.OUT_W, 0, -1), (L.VEL_S, L.OUT_N, 1, 1),
         (L.VEL_N, L.OUT_S, 1, -1), (L.VEL_B, L.OUT_T, 2, 1), (L.VEL_T, L.OUT_B, 2, -1)]


def open_box(vel, out, axis, g, shape=(8, 9, 10)):
    """Box (nz, ny, nx) periodic across the two tangential axes with a
    velocity face and the opposite outlet along `axis`; the face velocity has
    a normal part into the box and tangential parts."""
    cls = np.ones(shape, np.uint8)
    sl_in = [slice(None)] * 3
    sl_out = [slice(None)] * 3
    ax = 2 - axis                                   # numpy axis of the lattice axis
    sl_in[ax] = 0 if g > 0 else -1
    sl_out[ax] = -1 if g > 0 else 0
    cls[tuple(sl_in)] = vel
    cls[tuple(sl_out)] = out
    U = np.zeros((3,) + shape)
    U[axis] = 0.03 * g
    U[(axis + 1) % 3] = 0.01
    U[(axis + 2) % 3] = -0.005
    per = [1, 1, 1]
    per[axis] = 0
    return cls, U, tuple(per)


@pytest.mark.parametrize("k", range(6))
@pytest.mark.parametrize("outlet", [L.OUTLET_PRESSURE, L.OUTLET_EXTRAP2, L.OUTLET_EXTRAP1, L.OUTLET_FEQ])
def test_velocity_faces_and_outlets_f64(oracle, k, outlet):
    vel, out, axis, g = PAIRS[k]
    cls, U, per = open_box(vel, out, axis, g)
    shape = cls.shape
    rng = np.random.default_rng(k)
    u0 = [0.02 * rng.standard_normal(shape) for _ in range(3)]
    gg, o = pair(oracle, 19, shape, cls, None, collision=L.TRT, omega=1 / 0.8, trt_magic=0.25, periodic=per,
                 ubc=U[0], vbc=U[1], wbc=U[2], outlet_model=outlet, rho_out=1.0)
    gg.init(0, u=u0[0], v=u0[1], w=u0[2])
    o.init(0, u=u0[0], v=u0[1], w=u0[2])
    compare(gg, o, 60)


@pytest.mark.parametrize("k", [0, 1])
def test_reference_west_east_flip_f64(oracle, k):
    vel, out, axis, g = PAIRS[k]
    cls, U, per = open_box(vel, out, axis, g)
    gg, o = pair(oracle, 19, cls.shape, cls, None, collision=L.BGK, omega=1 / 0.7, periodic=per, ubc=U[0],
                 vbc=U[1], wbc=U[2], outlet_model=L.OUTLET_EXTRAP1, parity=L.PARITY_INLET3D_FLIP)
    gg.init(0)
    o.init(0)
    compare(gg, o, 60)


@pytest.mark.parametrize("k", [0, 3, 5])
def test_frozen_inlet_macros_f64(oracle, k):
    """UpdateInltOutl = no (LBM_PARITY_INLET_FROZEN): velocity-face nodes keep
    the macros lbm_init set and collide with them; also through lbm_set_f,
    which leaves them frozen."""
    vel, out, axis, g = PAIRS[k]
    cls, U, per = open_box(vel, out, axis, g)
    shape = cls.shape
    rng = np.random.default_rng(20 + k)
    r0 = 1 + 0.01 * rng.standard_normal(shape)
    m0 = [0.02 * rng.standard_normal(shape) for _ in range(3)]
    kw = dict(collision=L.TRT, omega=1 / 0.8, trt_magic=0.25, periodic=per, ubc=U[0], vbc=U[1], wbc=U[2],
              outlet_model=L.OUTLET_EXTRAP1)
    gg, o = pair(oracle, 19, shape, cls, None, parity=L.PARITY_INLET_FROZEN, **kw)
    for s in (gg, o):
        s.init(0, rho=r0, u=m0[0], v=m0[1], w=m0[2])
    compare(gg, o, 40)
    face = cls == vel
    got = gg.macros()
    for a, b in zip(got, [r0] + m0):
        assert np.array_equal(a[face], b[face])
    f = o.get_f()
    gg.set_f(f)
    o.set_f(f)
    compare(gg, o, 20)
    # without the flag the face's macros are its moments (u = u_bc)
    free = L.Solver(19, shape, cls, None, **kw)
    free.init(0, rho=r0, u=m0[0], v=m0[1], w=m0[2])
    free.step(40)
    assert np.abs(free.macros()[1][face] - U[0][face]).max() < 1e-13
    assert np.abs(got[1][face] - U[0][face]).max() > 1e-3


def lid_cavity():
    paths = []
    import tempfile
    d = tempfile.mkdtemp()
    for nm in ("lidvalidation_Node.dat", "lidvalidation_BC.dat"):
        p = os.path.join(d, nm)
        with open(p, "wb") as fh:
            fh.write(gzip.decompress(open(os.path.join(GOLDEN, nm + ".gz"), "rb").read()))
        paths.append(p)
    r = L.read_lattice(*paths)
    cls, bb = r["node_class"], r["bb_links"]
    U = np.zeros((3,) + cls.shape)
    U[0][cls == L.VEL_T] = 0.05
    return cls, bb, U, paths


@pytest.mark.parametrize("coll", [L.BGK, L.TRT, L.MRT])
def test_reference_lid_cavity_wet_node_walls_f64(oracle, coll):
    """L/lidvalidation (30^3, the reference's 3D validation lattice): lid at
    0.05 on the top face, wet-node walls everywhere else."""
    cls, bb, U, _ = lid_cavity()
    gg, o = pair(oracle, 19, cls.shape, cls, bb, collision=coll, omega=1 / 0.8, wall_model=L.WALL_HH,
                 ubc=U[0], vbc=U[1], wbc=U[2])
    gg.init(0)
    o.init(0)
    compare(gg, o, 150)


def test_symmetry_planes_and_on_node_walls_f64(oracle):
    """Force-driven channel along x: symmetry planes on y = 0, ny-1, wet-node
    walls on z = 0, nz-1, periodic x."""
    nz, ny, nx = 10, 8, 6
    cls = np.ones((nz, ny, nx), np.uint8)
    cls[1:-1, 0, :] = L.SYM_S
    cls[1:-1, -1, :] = L.SYM_N
    c = LT.velocities(19)
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    bb = np.zeros(cls.shape, np.uint32)
    wall = (zz == 0) | (zz == nz - 1)
    for d in range(1, 19):
        sz, sy = zz + c[d, 2], yy + c[d, 1]
        bb |= np.where(wall & ((sz < 0) | (sz >= nz) | (sy < 0) | (sy >= ny)), np.uint32(1 << d), np.uint32(0))
    gg, o = pair(oracle, 19, cls.shape, cls, bb, collision=L.TRT, omega=1 / 0.8, wall_model=L.WALL_HH,
                 force_model=L.FORCE_GUO, g=(1e-5, 0, 0), periodic=(1, 0, 0))
    gg.init(0)
    o.init(0)
    compare(gg, o, 200)


def test_duct_parity_and_series_solution(oracle):
    """Square duct, wet-node walls, plug inlet, pressure outlet: 200 steps
    against the oracle, then the steady mid-length profile against the series
    solution (shape within 0.5 % of the maximum)."""
    n, nx, U = 16, 40, 0.02
    d = LT.duct3d(nx, n, U)
    kw = dict(collision=L.TRT, omega=1 / 0.8, trt_magic=0.25, wall_model=L.WALL_HH,
              outlet_model=L.OUTLET_PRESSURE, rho_out=1.0, ubc=d["ubc"], vbc=d["vbc"], wbc=d["wbc"])
    gg, o = pair(oracle, 19, (n, n, nx), d["node_class"], d["bb"], **kw)
    gg.init(0)
    o.init(0)
    compare(gg, o, 200)
    steps, res = gg.run_to_residual(1e-11, 500, 100000)
    u = gg.macros()[1]
    mid = u[:, :, nx // 2]
    ref = square_duct_profile(n)
    scale = (ref[1:-1, 1:-1] * mid[1:-1, 1:-1]).sum() / (mid[1:-1, 1:-1] ** 2).sum()
    err = np.abs(scale * mid - ref).max() / ref.max()
    print(f"steps {steps} residual {res:.2e} duct shape error {err:.4f}", flush=True)
    assert res < 1e-11 and err < 0.005


@pytest.mark.parametrize("outlet", [L.OUTLET_EXTRAP2, L.OUTLET_FEQ])
def test_duct_f32_close_to_f64(oracle, outlet):
    n, nx, U = 12, 24, 0.02
    d = LT.duct3d(nx, n, U)
    kw = dict(collision=L.TRT, omega=1 / 0.8, wall_model=L.WALL_HH, outlet_model=outlet, ubc=d["ubc"],
              vbc=d["vbc"], wbc=d["wbc"])
    gg, o = pair(oracle, 19, (n, n, nx), d["node_class"], d["bb"], storage=L.STORE_F32, **kw)
    gg.init(0)
    o.init(0)
    gg.step(200)
    o.step(200)
    fl = d["node_class"] != L.SOLID
    gu, ou = gg.macros()[1], o.macros()[1]
    e = np.abs(gu[fl] - ou[fl]).max() / np.abs(ou[fl]).max()
    print(f"f32 duct outlet {outlet}: relative velocity error {e:.3e}", flush=True)
    assert e < 5e-5


INI3 = """Lattice = "lidvalidation"
U= 0.05
V= 0.0
W= 0.0
Density = 1.0
viscosity = 0.1
InletProfile = INLET_CTE
CollisionModel = BGKW
Walls = STRAIGHT
BCwallModel = HHmodel
BCoutlet = 1stE
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
"""


def test_lbmsolver_reference_lid_cavity(tmp_path):
    """The reference's 3D validation lattice through lbmsolver with its
    SetUpData keys (HHmodel walls, U = 0.05 on the lid) equals the ABI run,
    and autosaves every 100 iterations."""
    print("case: lbmsolver lidvalidation HHmodel", flush=True)
    cls, bb, U, paths = lid_cavity()
    os.makedirs(tmp_path / "Mesh")
    for p in paths:
        os.replace(p, tmp_path / "Mesh" / os.path.basename(p))
    (tmp_path / "SetUpData.ini").write_text(INI3.replace("AutosaveEach = 0", "AutosaveEach = 100"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    s = L.Solver(19, cls.shape, cls, bb, collision=L.BGK, omega=1 / (3 * 0.1 + 0.5), wall_model=L.WALL_HH,
                 ubc=U[0], vbc=U[1], wbc=U[2])
    n = cls.size
    s.init(1, rho=np.ones(cls.shape), u=np.where(cls == L.VEL_T, 0.05, 0.05), v=np.zeros(cls.shape),
           w=np.zeros(cls.shape))
    s.step(300)
    rho, u, v, w = s.macros()
    assert data.shape[0] == n
    assert np.abs(data[:, 3] - u.ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - rho.ravel()).max() < 1e-9
    # 3D autosave files (S/Iterate3D.cu:803-837): every 100 iterations
    saves = sorted(f for f in os.listdir(tmp_path / "Results") if f.startswith("autosave_iter"))
    assert saves == ["autosave_iter00100.csv", "autosave_iter00200.csv", "autosave_iter00300.csv"]
    last = np.loadtxt(tmp_path / "Results" / saves[-1], delimiter=",", skiprows=1)
    assert (last == data).all()
    # the 3D line and plane files (S/FilesWriting.cu:260-337)
    lines = (tmp_path / "Results" / "midLines.txt").read_text().splitlines()
    assert lines[0] == "CASE: Ulid = 0.050000 Nx = 30" and lines[1] == "U/Ulid\tZ\tX\tW/Ulid"
    ml = np.loadtxt(lines[2:])
    u3 = data[:, 3].reshape(cls.shape)
    w3 = data[:, 5].reshape(cls.shape)
    assert ml.shape == (30, 4)
    assert np.abs(ml[:, 0] - u3[:, 15, 15] / 0.05).max() < 1e-7
    assert np.abs(ml[:, 3] - w3[15, 15, :] / 0.05).max() < 1e-7
    plane = np.loadtxt(tmp_path / "Results" / "channelPlane.txt", skiprows=2)
    assert plane.shape == (900, 6) and np.abs(plane[:, 3] - u3[:, :, 22].ravel()).max() < 1e-9


def test_lbmsolver_update_inlet_outlet_no(tmp_path):
    """UpdateInltOutl = no freezes the lid's macros (S/GpuUpdateMacro.cu:150-151):
    lbmsolver equals the ABI run with LBM_PARITY_INLET_FROZEN and differs
    from the run without it."""
    print("case: lbmsolver lidvalidation UpdateInltOutl = no", flush=True)
    cls, bb, U, paths = lid_cavity()
    os.makedirs(tmp_path / "Mesh")
    for p in paths:
        os.replace(p, tmp_path / "Mesh" / os.path.basename(p))
    (tmp_path / "SetUpData.ini").write_text(INI3.replace("UpdateInltOutl = yes", "UpdateInltOutl = no"))
    r = subprocess.run([CLI, "-f", "SetUpData.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    data = np.loadtxt(tmp_path / "Results" / "FinalData.csv", delimiter=",", skiprows=1)
    us = {}
    for parity in (L.PARITY_INLET_FROZEN, 0):
        s = L.Solver(19, cls.shape, cls, bb, collision=L.BGK, omega=1 / (3 * 0.1 + 0.5), wall_model=L.WALL_HH,
                     ubc=U[0], vbc=U[1], wbc=U[2], parity=parity)
        s.init(1, rho=np.ones(cls.shape), u=np.full(cls.shape, 0.05), v=np.zeros(cls.shape), w=np.zeros(cls.shape))
        s.step(300)
        us[parity] = s.macros()[1].ravel()
    assert np.abs(data[:, 3] - us[L.PARITY_INLET_FROZEN]).max() < 1e-9
    assert np.abs(data[:, 3] - us[0]).max() > 1e-6

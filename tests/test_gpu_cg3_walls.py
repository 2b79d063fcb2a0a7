"""The D3Q19 colour-gradient box with walls on the y faces and a body force
(the reference's single-phase 3D faces per colour, S/Iterate3D.cu:576-598, in
its collide / stream / wall / macro order) on the GPU: FP64 parity with the
oracle, FP32 storage close to it, and the two-layer Couette profile of the
reference's 3D test case 4 against analyticalCouette."""
import numpy as np
import pytest

from thesis_amd import cg_cases as CC
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu


def compare(g, o, tol=1e-12):
    gm, om = g.macros(), o.get()
    for k in ("r_rho", "b_rho", "rho"):
        assert np.abs(gm[k] - om[k]).max() <= tol * np.abs(om[k]).max(), k
    du = np.sqrt((gm["u"] - om["u"]) ** 2 + (gm["v"] - om["v"]) ** 2 + (gm["w"] - om["w"]) ** 2).max()
    speed = np.sqrt(om["u"] ** 2 + om["v"] ** 2 + om["w"] ** 2).max()
    assert du <= max(tol * speed, 1e-15), (du, speed)
    gr, gb = g.get_f()
    scale = max(np.abs(om["r_f"]).max(), np.abs(om["b_f"]).max())
    assert np.abs(gr.ravel() - om["r_f"].ravel()).max() <= tol * scale
    assert np.abs(gb.ravel() - om["b_f"].ravel()).max() <= tol * scale


def layers3(nx, ny, nz):
    return np.broadcast_to(CC.layers(nx, ny), (nz, ny, nx)).copy()


@pytest.mark.parametrize("walls,ho,enh", [(1, 0, 0), (1, 1, 0), (1, 1, 1), (2, 0, 0), (2, 1, 1)])
def test_walls_f64_match_oracle(oracle, walls, ho, enh):
    nx, ny, nz = 20, 17, 12
    print(f"case: cg3 walls={walls} high_order={ho} enhanced={enh} {nx}x{ny}x{nz}", flush=True)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    red = ((y > ny / 2) ^ (np.hypot(x - 9.3, z - 6.1) < 3.5)).astype(float)
    b = (1 - red) * (0.7 if enh else 1.0)
    kw = dict(beta=0.7, A=1e-3, r_nu=1 / 6, b_nu=1 / 24, r_alpha=1 / 3, b_alpha=1 / 3, high_order=ho, enhanced=enh,
              walls=walls, u_wall=0.01 if walls == 1 else 0.0, g=-2e-6 if walls == 2 else 0.0)
    g = L.CgSolver(nx, ny, nz=nz, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(nx, ny, nz=nz, nthreads=4, **kw)
    g.init(red, b)
    o.init(red, b)
    g.step(150)
    o.step(150)
    compare(g, o)
    g.close()


@pytest.mark.parametrize("walls,ext", [(0, 1), (0, 0), (1, 1)])
def test_body_force_f64_matches_oracle(oracle, walls, ext):
    nx, ny, nz = 16, 14, 10
    print(f"case: cg3 force walls={walls} external_force={ext}", flush=True)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    red = (np.sqrt((x - 7.6) ** 2 + (y - 6.4) ** 2 + (z - 4.8) ** 2) < 4).astype(float)
    kw = dict(beta=0.7, A=1e-3, r_nu=1 / 6, b_nu=0.1, r_alpha=1 / 3, b_alpha=1 / 3, walls=walls, g=2e-6,
              external_force=ext)
    g = L.CgSolver(nx, ny, nz=nz, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(nx, ny, nz=nz, nthreads=4, **kw)
    g.init(red, 1 - red)
    o.init(red, 1 - red)
    g.step(150)
    o.step(150)
    compare(g, o)
    g.close()


def test_walls_f32_storage_close_to_oracle(oracle):
    nx, ny, nz = 12, 17, 8
    kw = dict(beta=0.7, A=1e-3, r_nu=1 / 6, b_nu=1 / 24, r_alpha=1 / 3, b_alpha=1 / 3, walls=1, u_wall=0.01)
    red = layers3(nx, ny, nz)
    g = L.CgSolver(nx, ny, nz=nz, storage=L.STORE_F32, **kw)
    o = oracle.CgSim(nx, ny, nz=nz, nthreads=4, **kw)
    g.init(red, 1 - red)
    o.init(red, 1 - red)
    g.step(400)
    o.step(400)
    gm, om = g.macros(), o.get()
    g.close()
    err = np.abs(gm["u"] - om["u"]).max() / 0.01
    print("FP32 storage: max |du| / U", err)
    assert err < 1e-4
    assert np.abs(gm["rho"] - om["rho"]).max() < 1e-5


def test_two_layer_couette_profile_3d_on_gpu():
    """The reference's 3D Couette validation (test case 4, analyticalCouette
    on the mid column of the mid plane): layers of viscosity ratio 4, the
    north face at U; after 30 000 steps the profile is piecewise linear away
    from the diffuse interface and the same in every column."""
    nx, ny, nz, U = 8, 33, 8, 0.01
    print("case: cg3 walls Couette profile", flush=True)
    nr, nb = 1 / 6, 1 / 24
    s = L.CgSolver(nx, ny, nz=nz, storage=L.STORE_F64, beta=0.7, A=1e-3, r_nu=nr, b_nu=nb, r_alpha=1 / 3,
                   b_alpha=1 / 3, walls=1, u_wall=U)
    red = layers3(nx, ny, nz)
    s.init(red, 1 - red)
    s.step(30000)
    m = s.macros()
    s.close()
    y = np.arange(ny) / (ny - 1)
    col = m["u"][nz // 2, :, nx // 2]
    err = np.abs(col - CC.analytical_couette(nr / nb, y, U)) / U
    print("max error / U", err.max(), "away from the interface", err[np.abs(y - 0.5) > 0.2].max())
    assert err.max() < 0.035 and err[np.abs(y - 0.5) > 0.2].max() < 0.01
    assert np.abs(m["u"] - m["u"][:1, :, :1]).max() < 1e-12

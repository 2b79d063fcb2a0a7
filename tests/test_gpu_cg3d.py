"""D3Q19 colour-gradient step on the GPU: FP64 parity with the oracle (the
reference's 3D GPU formulation on a periodic box) for both gradient stencils,
the reference's steady-bubble start, FP32 / FP16 storage against FP64, and the
Laplace law of a sphere, Delta p = 2 sigma / R."""
import numpy as np
import pytest

from thesis_amd import lattices as LT
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
    assert np.abs(gr - om["r_f"]).max() <= tol
    assert np.abs(gb - om["b_f"]).max() <= tol
    return gm, om


@pytest.mark.parametrize("ho,enh", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_cg3d_f64_matches_oracle(oracle, ho, enh):
    nx, ny, nz = 70, 13, 37            # partial 64 x 4 blocks
    print(f"case: cg3d parity {nx}x{ny}x{nz} high_order={ho} enhanced={enh}", flush=True)
    r, b = LT.droplet3d(nx, ny, nz, 4.6, cx=30.3, cy=6.7, cz=17.2)
    b = 0.6 * b                        # density contrast for the enhanced terms
    kw = dict(beta=0.7, A=0.01, r_nu=1 / 6, b_nu=0.1, r_alpha=0.3, b_alpha=0.4, high_order=ho, enhanced=enh)
    g = L.CgSolver(nx, ny, nz=nz, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(nx, ny, nz=nz, **kw)
    g.init(r, b)
    o.init(r, b)
    g.step(80)
    o.step(80)
    compare(g, o)
    red = g.reduce()
    pin, nin, pout, nout, umax = o.laplace()
    assert abs(red["pin_sum"] - pin) < 1e-9 and red["pin_n"] == nin and red["pout_n"] == nout
    assert abs(red["umax"] - umax) <= 1e-12 * max(umax, 1e-300) + 1e-16
    g.close()


def test_cg3d_reference_bubble_start_f64(oracle):
    """The reference's steady bubble (S/GpuInit.cu:543-592) with its default
    colour parameters (S/main.cu:52-72): rest fractions 4/9, beta 0.99,
    A 1e-4, viscosities 0.5, 104-neighbour gradient."""
    n = 24
    print(f"case: cg3d reference bubble {n}^3", flush=True)
    r, b = LT.droplet3d(n, n, n, 6.0)
    rf, bf = LT.bubble3d_populations(r, b)
    kw = dict(beta=0.99, A=1e-4, r_nu=0.5, b_nu=0.5, r_alpha=4 / 9, b_alpha=4 / 9, high_order=1, kappa=0.9)
    g = L.CgSolver(n, n, nz=n, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(n, n, nz=n, **kw)
    g.init_f(rf, bf)
    o.init_f(rf, bf)
    g.step(60)
    o.step(60)
    compare(g, o)
    g.close()


# 32^3 runs the two-launch step, 64^3 the single pass (whole 64 x 8 row tiles)
@pytest.mark.parametrize("storage,tol", [(L.STORE_F32, 1e-5), (L.STORE_F16, 2e-3)])
@pytest.mark.parametrize("n", [32, 64])
def test_cg3d_reduced_storage_close_to_f64(storage, tol, n):
    print(f"case: cg3d storage {storage} {n}^3", flush=True)
    r, b = LT.droplet3d(n, n, n, 8.3 * n / 32, cx=15.2 * n / 32, cy=16.1 * n / 32, cz=14.7 * n / 32)
    kw = dict(beta=0.7, A=0.01125, r_alpha=1 / 3, b_alpha=1 / 3, high_order=0)
    out = {}
    for st in (L.STORE_F64, storage):
        g = L.CgSolver(n, n, nz=n, storage=st, **kw)
        g.init(r, b)
        g.step(300)
        out[st] = (g.macros(), g.reduce())
        g.close()
    m64, r64 = out[L.STORE_F64]
    m, red = out[storage]
    e = np.abs(m["rho"] - m64["rho"]).max()
    print(f"storage {storage}: max |drho| {e:.2e}, red mass {red['mass_r']:.10g} vs {r64['mass_r']:.10g}", flush=True)
    assert red["nonfinite"] == 0
    assert e < tol
    assert abs(red["mass_r"] - r64["mass_r"]) < 10 * tol * r64["mass_r"]


def test_cg3d_laplace_law_sphere():
    """48^3 box, R = 12, tau = 1, sigma = 0.005: R Delta p / (2 sigma) within
    5 % of the oracle's 1 + b/R^2 trend (1.04 at R = 12), FP32 storage."""
    n, R, A = 48, 12, 0.01125
    print(f"case: cg3d laplace {n}^3 R={R}", flush=True)
    g = L.CgSolver(n, n, nz=n, beta=0.7, A=A, r_alpha=1 / 3, b_alpha=1 / 3, storage=L.STORE_F32)
    r, b = LT.droplet3d(n, n, n, R)
    g.init(r, b)
    m0 = g.reduce()
    g.step(3000)
    red = g.reduce()
    dp = (red["rho_in_sum"] / red["rho_in_n"] - red["rho_out_sum"] / red["rho_out_n"]) / 3.0
    ratio = dp * R / (2 * 4 * A / 9)
    ms = g.time_steps(200)
    print(f"R dp/(2 sigma) {ratio:.4f}, umax {red['umax']:.2e}, {n ** 3 * 200 / ms / 1e3:.1f} MLUPS", flush=True)
    g.close()
    assert abs(ratio - 1.0) < 0.08
    assert red["umax"] < 1e-3
    assert abs(red["mass_r"] - m0["mass_r"]) < 1e-4 * m0["mass_r"]


@pytest.mark.parametrize("ho", [0, 1])
def test_cg3d_throughput_256(ho):
    """Throughput at 256^3 (FP32 and FP16 storage): the single-pass step with
    the 18-neighbour gradient, the two-launch step with the 104-neighbour one."""
    n = 256
    r, b = LT.droplet3d(n, n, n, 40)
    for st in (L.STORE_F32, L.STORE_F16):
        print(f"case: cg3d throughput {n}^3 storage {st} high_order {ho}", flush=True)
        g = L.CgSolver(n, n, nz=n, storage=st, A=0.01125, r_alpha=1 / 3, b_alpha=1 / 3, high_order=ho)
        g.init(r, b)
        g.step(5)
        ms = g.time_steps(20)
        print(f"storage {st}: {ms / 20:.3f} ms/step, {n ** 3 * 20 / ms / 1e3:.0f} MLUPS", flush=True)
        red = g.reduce()
        g.close()
        assert red["nonfinite"] == 0

"""Colour-gradient two-phase step on the GPU: FP64 parity with the oracle
(reference GPU formulation, periodic stencil) and the static-droplet config:
512^2, R in {32, 48, 64, 96}, density ratio 1, sigma 0.005, beta 0.7, tau 1,
30 000 steps; Laplace law and spurious currents in FP32 compute with FP16
storage against FP64."""
import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ho,enh", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_cg_f64_matches_oracle(oracle, ho, enh):
    nx, ny = 70, 45          # partial 32x16 tiles on both axes
    print(f"case: cg parity {nx}x{ny} high_order={ho} enhanced={enh}", flush=True)
    r, b = LT.droplet2d(nx, ny, 12, cx=30.3, cy=20.7)
    if enh:
        b = 0.6 * b          # density contrast for the enhanced terms
    kw = dict(beta=0.7, A=0.01125, r_nu=1 / 6, b_nu=0.1, r_alpha=4 / 9, b_alpha=0.2, high_order=ho, enhanced=enh)
    g = L.CgSolver(nx, ny, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(nx, ny, **kw)
    g.init(r, b)
    o.init(r, b)
    g.step(150)
    o.step(150)
    gm, om = g.macros(), o.get()
    for k in ("r_rho", "b_rho", "rho"):
        assert np.abs(gm[k] - om[k]).max() <= 1e-12 * np.abs(om[k]).max(), k
    du = np.sqrt((gm["u"] - om["u"]) ** 2 + (gm["v"] - om["v"]) ** 2).max()
    speed = np.sqrt(om["u"] ** 2 + om["v"] ** 2).max()
    assert du <= max(1e-12 * speed, 1e-15), (du, speed)
    gr, gb = g.get_f()
    assert np.abs(gr - om["r_f"]).max() <= 1e-12
    assert np.abs(gb - om["b_f"]).max() <= 1e-12
    red = g.reduce()
    pin, nin, pout, nout, umax = o.laplace()
    assert abs(red["pin_sum"] - pin) < 1e-9 and red["pin_n"] == nin and red["pout_n"] == nout
    assert abs(red["umax"] - umax) <= 1e-12 * max(umax, 1e-300) + 1e-16


def run_droplet(R, storage, steps=30000, n=512, sigma=0.005, avg_from=24000, every=100):
    """Static droplet; the pressure jump is averaged over steps avg_from..steps
    (the box keeps a slowly damped breathing mode, period ~ thousands of steps)."""
    A = 9 * sigma * 1.0 / 4          # sigma = 4 A / (9 omega), tau = 1
    s = L.CgSolver(n, n, beta=0.7, A=A, r_nu=1 / 6, b_nu=1 / 6, storage=storage)
    r, b = LT.droplet2d(n, n, R)
    s.init(r, b)
    m0 = s.reduce()
    s.step(avg_from)
    dps, umax = [], 0.0
    for _ in range((steps - avg_from) // every):
        s.step(every)
        red = s.reduce()
        dps.append((red["rho_in_sum"] / red["rho_in_n"] - red["rho_out_sum"] / red["rho_out_n"]) / 3.0)
        umax = max(umax, red["umax"])
    dp = float(np.mean(dps))
    mac = s.macros()
    s.close()
    return dict(dp=dp, laplace=dp * R / sigma, umax=umax, mass_r=(m0["mass_r"], red["mass_r"]),
                u=mac["u"], v=mac["v"], rho=mac["rho"], nonfinite=red["nonfinite"])


def test_cg_static_droplet_config():
    rows = []
    for R in (32, 48, 64, 96):
        print(f"case: droplet R={R} f64 and f16, 512^2, 30000 steps", flush=True)
        d64 = run_droplet(R, L.STORE_F64)
        d16 = run_droplet(R, L.STORE_F16)
        err_dp = abs(d16["dp"] - d64["dp"]) / abs(d64["dp"])
        err_rho = np.abs(d16["rho"] - d64["rho"]).max()
        m0, m1 = d64["mass_r"]
        rows.append(dict(R=R, lap64=d64["laplace"], lap16=d16["laplace"], u64=d64["umax"], u16=d16["umax"],
                         err_dp=err_dp, err_rho=err_rho, dmass=abs(m1 - m0) / m0,
                         bad=d64["nonfinite"] + d16["nonfinite"]))
    print("R  dpR/sigma(f64) dpR/sigma(f16) umax(f64) umax(f16) rel.err dp f16 max|drho| f16 red-mass drift")
    for r in rows:
        print("%(R)3d %(lap64).4f %(lap16).4f %(u64).3e %(u16).3e %(err_dp).2e %(err_rho).2e %(dmass).1e" % r, flush=True)
    for r in rows:
        assert r["bad"] == 0
        # Laplace law Delta p = sigma / R (time-averaged jump)
        assert abs(r["lap64"] - 1) < 0.1 and abs(r["lap16"] - 1) < 0.1
        assert r["u64"] < 1e-3 and r["u16"] < 1e-3          # spurious currents
        assert r["dmass"] < 1e-10                             # red mass, FP64 storage
        assert r["err_dp"] < 0.1                              # FP16 storage vs FP64


def test_cg_init_f_f64_matches_oracle(oracle):
    """Start from given colour populations (perturbed rest states)."""
    nx, ny = 40, 36
    print(f"case: cg init_f parity {nx}x{ny}", flush=True)
    r, b = LT.droplet2d(nx, ny, 9, cx=20.2, cy=17.6)
    phi = np.array([0.0, 0.2, 0.2, 0.2, 0.2, 0.05, 0.05, 0.05, 0.05])
    teta = np.array([1.0, -0.2, -0.2, -0.2, -0.2, -0.05, -0.05, -0.05, -0.05])
    rest = phi + teta * 0.3
    rng = np.random.default_rng(7)
    rf = rest[:, None, None] * r[None] * (1 + 0.05 * rng.random((9, ny, nx)))
    bf = rest[:, None, None] * b[None] * (1 + 0.05 * rng.random((9, ny, nx)))
    kw = dict(beta=0.7, A=0.01125, r_nu=1 / 6, b_nu=0.1, r_alpha=0.3, b_alpha=0.3, high_order=1)
    g = L.CgSolver(nx, ny, storage=L.STORE_F64, **kw)
    o = oracle.CgSim(nx, ny, **kw)
    g.init_f(rf, bf)
    o.init_f(rf, bf)
    g.step(50)
    o.step(50)
    gm, om = g.macros(), o.get()
    for k in ("r_rho", "b_rho", "rho", "u", "v"):
        assert np.abs(gm[k] - om[k]).max() <= 1e-12 * max(np.abs(om[k]).max(), 1e-3), k
    gr, gb = g.get_f()
    assert np.abs(gr - om["r_f"]).max() <= 1e-12 and np.abs(gb - om["b_f"]).max() <= 1e-12


def test_cg_throughput():
    """Whole-step time of the D2Q9 colour-gradient step (droplet, A = 0.01125,
    beta = 0.7, tau = 1, 8-point gradient): 512^2 (the static-droplet box)
    and 2048^2, FP64, FP32 and FP16 storage."""
    for n in (512, 2048):
        r, b = LT.droplet2d(n, n, n / 8)
        for st in (L.STORE_F64, L.STORE_F32, L.STORE_F16):
            print(f"case: cg throughput {n}^2 storage {st}", flush=True)
            s = L.CgSolver(n, n, beta=0.7, A=0.01125, r_nu=1 / 6, b_nu=1 / 6, storage=st)
            s.init(r, b)
            s.step(20)
            ms = min(s.time_steps(200) / 200 for _ in range(2))
            red = s.reduce()
            s.close()
            print(f"{n}^2 storage {st}: {ms * 1e3:.1f} us/step, {n * n / ms / 1e6:.2f} GLUPS", flush=True)
            assert red["nonfinite"] == 0

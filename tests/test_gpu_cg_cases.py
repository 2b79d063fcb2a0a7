"""The reference's periodic two-phase test cases on the GPU with its
validators (thesis_amd/cg_cases.py restating S/Multiphase.cu:1104-1340):
a square relaxing to the disc of the same area (case 3), two touching discs
merging into one of sqrt(2) times the radius (case 4), and an elliptic bubble
oscillating at the 2D mode-2 frequency of Lamb (case 5).

The validators measure a radius as half the sum of the positive phase field
along the middle row, which misses about one node of each diffuse interface:
their error falls as 1/n, and the tests ask for that fall as well as for a
bound at n = 256.  The oscillation frequency comes out 17-20 % below Lamb's
inviscid small-amplitude value at these sizes (amplitude 40 %, viscous
damping): the bound is 25 %."""
import numpy as np
import pytest

from thesis_amd import cg_cases as CC
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu


def relax(n, shape, steps, nu=1 / 6, A=0.01):
    s = L.CgSolver(n, n, beta=0.7, A=A, r_nu=nu, b_nu=nu, storage=L.STORE_F64)
    red = shape(n, n)
    s.init(red, 1 - red)
    s.step(steps)
    m = s.macros()
    s.close()
    return m["r_rho"], m["b_rho"]


def test_deforming_square_becomes_disc_of_equal_area():
    errs = {}
    for n, steps in ((128, 20000), (256, 60000)):
        print(f"case: two-phase case 3, n = {n}", flush=True)
        got, want = CC.deforming_radius(*relax(n, CC.square, steps))
        errs[n] = abs(got - want) / want
        print(f"n {n}: radius {got:.3f} predicted {want:.3f} error {100 * errs[n]:.2f} %")
    assert errs[256] < 0.035 and errs[256] < 0.6 * errs[128]


def test_coalescence_gives_sqrt2_radius():
    R = 0.15
    errs = {}
    for n, steps in ((128, 20000), (256, 60000)):
        print(f"case: two-phase case 4, n = {n}", flush=True)
        got, want = CC.coalescence_radius(*relax(n, lambda a, b: CC.two_bubbles(a, b, R), steps), R)
        errs[n] = abs(got - want) / want
        print(f"n {n}: radius {got:.4f} predicted {want:.4f} error {100 * errs[n]:.2f} %")
    assert errs[256] < 0.035 and errs[256] < 0.6 * errs[128]


def test_oscillating_bubble_frequency():
    n, nu, A, every = 256, 0.01, 0.01, 50
    print("case: two-phase case 5, n = 256", flush=True)
    s = L.CgSolver(n, n, beta=0.7, A=A, r_nu=nu, b_nu=nu, storage=L.STORE_F64)
    red = CC.ellipse(n, n)
    s.init(red, 1 - red)
    trace = []
    for _ in range(60000 // every):
        s.step(every)
        m = s.macros()
        trace.append(CC.max_y_oscillating(m["r_rho"], m["b_rho"]))
    radius = CC.mid_row_red_width(m["r_rho"], m["b_rho"]) / 2
    sigma = 4 * A / (9 * (1 / (3 * nu + 0.5)))        # sigma = 4 A / (9 omega)
    theo, num, err = CC.oscillation_frequency(trace, sigma, 1.0, 1.0, radius, every=every)
    print(f"radius {radius:.2f}: Lamb {theo:.4e}, measured {num:.4e}, error {err:.1f} %")
    assert np.isfinite(num) and err < 25.0
    # the oscillation decays: later extremes lie closer to the mean
    tr = np.asarray(trace[20:])
    assert np.abs(tr[-200:] - tr[-200:].mean()).max() < 0.5 * np.abs(tr[:200] - tr[-200:].mean()).max()

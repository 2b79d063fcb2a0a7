"""Pins of the oracle's D2Q9 colour-gradient walls (walls on the y rows, x
periodic; gpuBcPeriodic2D, S/GpuBoundaries.cu:1876-1968) and body force
(S/GpuUpdateMacro.cu:118-119).  CPU only.

* Zou-He velocity walls (the reference's Couette case): one fluid between a
  wall at rest and one moving at U has the exact linear profile U y/(ny-1)
  (the wet-node Zou-He walls sit on the end rows, whose populations carry
  exactly the wall velocity), with either gradient stencil; two layers of viscosity ratio kappa = mu_r/mu_b have the piecewise
  linear profile of analyticalCouette (S/Multiphase.cu:1136-1160) away from
  the diffuse interface.
* Wall rows take the mirrored 8-point gradient with x components only
  (S/GpuCollision.cu:135-156): for a phase field that varies along x only,
  the wall-row gradient equals the one of an interior row, and for one that
  varies along y only it is zero.
* Bounce-back walls (the Rayleigh-Taylor case): each colour's three incoming
  links on a wall row equal the opposite outgoing ones after the step, and a
  heavy layer over a light one under gravity sinks where its cosine
  perturbation dips.
* The reference's force shift on a Zou-He wall row: the wall keeps the
  velocity the macroscopic update stored, so with a force along x it moves
  by g/(rho omega_eff) more every step (its two-phase Poiseuille case).
"""
import numpy as np
import pytest

from thesis_amd import cg_cases as CC


@pytest.mark.parametrize("ho", [0, 1])
def test_single_fluid_couette_is_exactly_linear(oracle, ho):
    nx, ny, U = 6, 17, 0.01
    s = oracle.CgSim(nx, ny, r_nu=1 / 6, b_nu=1 / 6, A=0.0, high_order=ho, walls=1, u_wall=U, nthreads=2)
    s.init(np.ones((ny, nx)), np.zeros((ny, nx)))
    s.step(6000)
    d = s.get()
    y = np.arange(ny) / (ny - 1)
    assert np.abs(d["u"] - U * y[:, None]).max() < 1e-10 * U
    assert np.abs(d["v"]).max() < 1e-15
    # the Zou-He rows carry exactly the wall velocity in each colour's populations
    c = np.array([[0, 1, 0, -1, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1]], float)
    f = d["r_f"]
    for row, uw in ((-1, U), (0, 0.0)):
        rho = f[:, row].sum(0)
        j = np.einsum("ak,kx->ax", c, f[:, row])
        assert np.abs(j[0] / rho - uw).max() < 1e-16 and np.abs(j[1] / rho).max() < 1e-16


def test_two_layer_couette_matches_piecewise_profile(oracle):
    nx, ny, U = 6, 33, 0.01
    nr, nb = 1 / 6, 1 / 24
    s = oracle.CgSim(nx, ny, r_nu=nr, b_nu=nb, beta=0.7, A=1e-3, walls=1, u_wall=U, nthreads=2)
    red = CC.layers(nx, ny)
    s.init(red, 1 - red)
    s.step(20000)
    d = s.get()
    y = np.arange(ny) / (ny - 1)
    ana = CC.analytical_couette(nr / nb, y, U)
    err = np.abs(d["u"][:, 0] - ana) / U
    # the diffuse interface (red from the first row above y = 1/2, as the
    # reference starts it) shifts the kink by about half a node
    assert err.max() < 0.035
    far = np.abs(y - 0.5) > 0.2
    assert err[far].max() < 0.01
    # slopes away from the interface in the ratio of the viscosities
    lo = np.polyfit(y[y < 0.3], d["u"][y < 0.3, 0], 1)[0]
    hi = np.polyfit(y[y > 0.7], d["u"][y > 0.7, 0], 1)[0]
    assert abs(lo / hi / (nr / nb) - 1) < 1e-3


def test_wall_rows_take_the_mirrored_gradient(oracle):
    nx, ny = 16, 9
    x = np.arange(nx)
    phase_x = 0.5 + 0.4 * np.sin(2 * np.pi * x / nx)[None, :] * np.ones((ny, 1))
    s = oracle.CgSim(nx, ny, walls=2)
    s.init(phase_x, 1 - phase_x)
    gx, gy = s.gradient()[:2]
    assert np.abs(gx[0] - gx[4]).max() < 1e-15 and np.abs(gx[-1] - gx[4]).max() < 1e-15
    assert np.abs(gy).max() < 1e-15
    y = np.arange(ny)
    phase_y = 0.5 + 0.3 * np.cos(np.pi * y / (ny - 1))[:, None] * np.ones((1, nx))
    s2 = oracle.CgSim(nx, ny, walls=2)
    s2.init(phase_y, 1 - phase_y)
    gx, gy = s2.gradient()[:2]
    assert np.abs(gx[[0, -1]]).max() < 1e-15 and np.abs(gy[[0, -1]]).max() < 1e-15
    assert np.abs(gy[1:-1]).min() > 1e-3


def test_rayleigh_taylor_bounce_back_walls_and_the_dip_sinks(oracle):
    nx, ny, gamma = 32, 64, 3.0           # density ratio 3 (S/main.cu:70-71 form of the rest fractions)
    b_alpha = 4 / 9
    s = oracle.CgSim(nx, ny, r_alpha=1 - (1 - b_alpha) / gamma, b_alpha=b_alpha, r_nu=1 / 6, b_nu=1 / 6, beta=0.7,
                     A=1e-3, walls=2, g=-5e-5, external_force=0, nthreads=2)
    red = CC.rayleigh_taylor(nx, ny)
    s.init(red, (1 - red) / gamma)
    m0 = red.sum() + (1 - red).sum() / gamma
    depth = [CC.min_y_rt(*_colours(s))]
    for _ in range(8):
        s.step(500)
        depth.append(CC.min_y_rt(*_colours(s)))
    d = s.get()
    # on-node bounce-back of the three incoming links on both wall rows, per colour
    for f in (d["r_f"], d["b_f"]):
        assert np.array_equal(f[4, -1], f[2, -1]) and np.array_equal(f[7, -1], f[5, -1])
        assert np.array_equal(f[8, -1], f[6, -1])
        assert np.array_equal(f[2, 0], f[4, 0]) and np.array_equal(f[5, 0], f[7, 0])
        assert np.array_equal(f[6, 0], f[8, 0])
    # the heavy layer sinks where the perturbation dips; the walls (which
    # drop what leaves through them and double what arrives) move little mass
    assert all(b < a for a, b in zip(depth[1:], depth[2:]))
    assert depth[-1] < depth[0] - 3
    assert abs(d["rho"].sum() / m0 - 1) < 2e-3


def _colours(s):
    d = s.get()
    return d["r_rho"], d["b_rho"]


def test_force_shift_moves_a_zou_he_wall(oracle):
    nx, ny, g = 4, 9, 1e-5
    s = oracle.CgSim(nx, ny, r_nu=1 / 6, b_nu=1 / 6, A=0.0, walls=1, u_wall=0.0, g=g, external_force=1, nthreads=1)
    s.init(np.ones((ny, nx)), np.zeros((ny, nx)))
    s.step(100)
    u100 = s.get()["u"][0].mean()
    s.step(100)
    u200 = s.get()["u"][0].mean()
    omega = 1 / (3 * (1 / 6) + 0.5)
    assert abs(u100 / (100 * g / omega) - 1) < 2e-3
    assert abs(u200 / u100 - 2) < 2e-3

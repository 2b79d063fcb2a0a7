"""Pins of two oracle functions that conservation alone does not fix.

1. The D3Q19 MRT equilibrium moments (oracle/lbm_oracle.c mrt_meq, restating
   S/GpuCollision.cu:1065-1083).  The oracle's m_eq is read back through its
   own collision: with S = I and no force, out = f - M^-1 (M f - m_eq) =
   M^-1 m_eq, so m_eq = M out, with M itself pinned against the reference's
   compiled MRTInitializer3D (tests/test_oracle.py).  At rho = 1 every moment
   whose equilibrium is the projection of the second-order f_eq must equal
   (M f_eq)(1, u); the three free ones (epsilon, 3 pi_xx, pi_ww) must follow
   the published weights w_eps = 0, w_epsj = -475/63, w_xx = 0 of
   d'Humieres, Ginzburg, Krafczyk, Lallemand & Luo, Phil. Trans. R. Soc. A 360
   (2002) 437, eqs. (3.5)-(3.6); and the quadratic terms are functions of
   j = rho u alone (the incompressible rho_0 = 1 form).  A force-driven channel
   with rates that make the MRT a TRT of magic parameter 3/16 must give the
   exact parabola (to the O(u^2) effect of the free equilibria).

2. The recolouring strength beta (oracle/lbm_oracle.c, Latva-Kokko recolouring
   of S/GpuCollision.cu:552-622 and :709-779).  For a flat interface at rest,
   the recolouring r_q = (rho_r/rho) f*_q + beta (rho_r rho_b / rho) cos(theta_q)
   f_eq,q(1, 0) followed by streaming has, to second order in the lattice
   spacing, the steady phase field N = (rho_r - rho_b)/rho = tanh(k x) with
   k = 3 beta K, K = sum_q w_q c_qx^2 / |c_q| (K = (2 + 1/sqrt 2)/9 for D2Q9,
   (1 + sqrt 2)/9 for D3Q19): the diffusion N''/12 of the streamed colour
   balances the recoloured flux beta K N N'/2 (Halliday, Hollis & Care,
   Phys. Rev. E 76 (2007) 026708 derive the same tanh profile).  k is
   independent of the mean density; the host formulation's
   beta rho_r rho_b / rho^2 (S/Multiphase.cu:6-138) would halve it at
   rho = 2, and any slip in beta scales it.  CPU only.
"""
import numpy as np
import pytest

from thesis_amd import lattices as LT

# moment rows of the reference's D3Q19 basis (S/CellFunctions.cu:73-114):
# 0 rho, 1 e, 2 eps, 3 jx, 4 qx, 5 jy, 6 qy, 7 jz, 8 qz, 9 3pxx, 10 3pixx,
# 11 pww, 12 piww, 13 pxy, 14 pyz, 15 pxz, 16 mx, 17 my, 18 mz
PROJECTED = [0, 1, 3, 4, 5, 6, 7, 8, 9, 11, 13, 14, 15, 16, 17, 18]
W_EPS, W_EPSJ, W_XX = 0.0, -475.0 / 63.0, 0.0


def _node(oracle, **kw):
    return oracle.Sim(19, (1, 1, 1), np.ones(1, np.uint8), periodic=(1, 1, 1), **kw)


def oracle_meq(oracle, rho, u, seed=0):
    """The oracle's m_eq(rho, u), read through its collision with S = I."""
    M, _, _ = oracle.mrt_matrices(19, 1.0)
    rng = np.random.default_rng(seed)
    f = oracle.equilibrium(19, 1.0, [0.01, 0.02, -0.01]) * (1 + 0.2 * rng.random(19))
    out = _node(oracle, collision=2, omega=1.0, mrt_s=[1.0] * 19).collide_node(f, rho, u)
    return M @ out


@pytest.mark.parametrize("u", [(0.05, -0.03, 0.02), (0.0, 0.07, 0.0), (-0.04, 0.01, 0.06)])
def test_mrt3d_equilibrium_is_projection_at_unit_density(oracle, u):
    M, _, _ = oracle.mrt_matrices(19, 1.0)
    meq = oracle_meq(oracle, 1.0, u)
    proj = M @ oracle.equilibrium(19, 1.0, u)
    assert np.abs(meq[PROJECTED] - proj[PROJECTED]).max() < 1e-14
    # the free moments are not the projection (the projection of a
    # second-order f_eq is not what the reference relaxes towards)
    assert abs(meq[2] - proj[2]) > 1e-4


@pytest.mark.parametrize("u", [(0.05, -0.03, 0.02), (0.03, 0.04, -0.05)])
def test_mrt3d_free_equilibria_follow_published_weights(oracle, u):
    for rho in (1.0, 1.3):
        meq = oracle_meq(oracle, rho, u, seed=1)
        j = rho * np.asarray(u)
        jj = float(j @ j)
        assert abs(meq[2] - (W_EPS * rho + W_EPSJ * jj)) < 1e-14
        assert abs(meq[10] - W_XX * meq[9]) < 1e-15
        assert abs(meq[12] - W_XX * meq[11]) < 1e-15


def test_mrt3d_equilibrium_quadratic_terms_depend_on_j_only(oracle):
    """Incompressible form (rho_0 = 1): at fixed j the non-conserved
    equilibria change only through their linear rho terms (e = -11 rho +
    19 j.j is the only one), so m_eq(rho, j/rho) - m_eq(1, j) is -11 (rho - 1)
    in e and the change of the conserved moments."""
    j = np.array([0.04, -0.03, 0.05])
    a = oracle_meq(oracle, 1.0, j)
    for rho in (0.8, 1.25, 2.0):
        b = oracle_meq(oracle, rho, j / rho)
        d = b - a
        assert abs(d[0] - (rho - 1)) < 1e-14
        assert abs(d[1] - (-11.0) * (rho - 1)) < 1e-13
        rest = [r for r in range(19) if r not in (0, 1)]
        assert np.abs(d[rest]).max() < 1e-14


def test_mrt3d_force_driven_channel_trt_rates_is_exact_parabola(oracle):
    """Rates that turn the MRT into a TRT with Lambda = 3/16 (even
    non-conserved moments at omega, odd at omega_minus), half-way walls,
    Guo forcing in moment space: u = g/(2 nu) y (H - y) to the O(u^2) part
    of the free equilibria; the reference's own rates slip by a few per cent."""
    nx, nz, H = 2, 2, 10
    tau = 0.9
    om = 1 / tau
    om_m = 1 / (0.5 + (3 / 16) / (tau - 0.5))
    nu, g = (tau - 0.5) / 3, 2e-6
    solid = np.zeros((nz, H, nx), bool)
    bb = LT.bb_from_solid(solid, 19, periodic=(True, False, True), outside_is_wall=(False, True, False))
    odd = {4, 6, 8, 16, 17, 18}
    s_trt = [0.0 if r in (0, 3, 5, 7) else (om_m if r in odd else om) for r in range(19)]
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y)
    errs = {}
    for name, rates in (("trt", s_trt), ("reference", None)):
        sim = oracle.Sim(19, (nz, H, nx), np.ones(nx * H * nz, np.uint8), bb, collision=2, omega=om,
                         mrt_s=rates, force_model=oracle.FORCE_GUO, g=(g, 0, 0), periodic=(1, 0, 1))
        sim.init(0)
        sim.step(6000)
        rho, u, v, w = sim.macros()
        prof = u[0, :, 0]
        assert np.abs(u - prof[None, :, None]).max() < 1e-15      # x and z invariant
        assert np.abs(v).max() < 1e-15 and np.abs(w).max() < 1e-15
        errs[name] = np.abs(prof - ua).max() / ua.max()
    # u_max = 1.2e-4: the free equilibria's j^2 terms enter at ~1e-8 relative
    assert errs["trt"] < 1e-6
    assert 1e-5 < errs["reference"] < 0.05


K2 = (2 + 1 / np.sqrt(2)) / 9
K3 = (1 + np.sqrt(2)) / 9


def flat_interface_slope(oracle, q, beta, rho0, nx=96, steps=4000):
    """Steady phase field of two flat interfaces in a periodic box (A = 0:
    recolouring only), slope of atanh(N) across the first interface."""
    n = 4
    x = np.arange(nx)
    N = np.where((x >= nx // 4) & (x < 3 * nx // 4), 1.0, -1.0)
    if q == 9:
        N2 = np.tile(N, (n, 1))
        s = oracle.CgSim(nx, n, beta=beta, A=0.0, nthreads=2)
    else:
        N2 = np.tile(N, (n, n, 1))
        s = oracle.CgSim(nx, n, nz=n, beta=beta, A=0.0, nthreads=2)
    s.init(rho0 * (1 + N2) / 2, rho0 * (1 - N2) / 2)
    s.step(steps)
    d = s.get()
    assert np.abs(d["r_rho"] + d["b_rho"] - rho0).max() < 1e-12       # at rest, uniform density
    prof = ((d["r_rho"] - d["b_rho"]) / d["rho"]).reshape(-1, nx)
    assert np.abs(prof - prof[0]).max() < 1e-13                      # flat
    p = prof[0]
    xs = np.arange(nx // 4 - 6, nx // 4 + 6)
    m = np.abs(p[xs]) < 0.9
    return np.polyfit(xs[m], np.arctanh(p[xs][m]), 1)[0]


@pytest.mark.parametrize("q,betas,steps", [(9, (0.1, 0.2, 0.4), 4000), (19, (0.2, 0.4), 2000)])
def test_recolouring_flat_interface_width_scales_with_beta(oracle, q, betas, steps):
    K = K2 if q == 9 else K3
    nx = 96 if q == 9 else 64
    ks = {}
    for beta in betas:
        for rho0 in (1.0, 2.0):
            k = flat_interface_slope(oracle, q, beta, rho0, nx=nx, steps=steps)
            ks[(beta, rho0)] = k
            assert abs(k / (3 * beta * K) - 1) < 0.01, (beta, rho0, k, 3 * beta * K)
        # no dependence on the mean density (beta rho_r rho_b / rho, not / rho^2)
        assert abs(ks[(beta, 2.0)] / ks[(beta, 1.0)] - 1) < 1e-9
    b0, b1 = betas[-2], betas[-1]
    assert abs(ks[(b1, 1.0)] / ks[(b0, 1.0)] - b1 / b0) < 0.01 * b1 / b0

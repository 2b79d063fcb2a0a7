"""The enhanced colour-gradient equilibrium (gpuCollEnhancedBgkwGC2D / 3D,
S/GpuCollision.cu:624-707, :781-880) in the oracle, pinned by its moments:
the correction nu (chi_q u.grad rho + psi_q G:c c) carries no mass and no
momentum, and its second moment is nu (delta u.grad rho + u grad rho +
grad rho u) because sum_q psi_q c c c c = (delta delta + ...)/2 on both
lattices.  Checked on one step from a uniform stream over a linear density
field (exact stencil gradient), against the same step without the
correction; at uniform density the two models agree."""
import numpy as np
import pytest

from thesis_amd import lattices as LT

PHI = {9: np.array([0.0] + [0.2] * 4 + [0.05] * 4), 19: np.array([0.0] + [1 / 12] * 6 + [1 / 24] * 12)}
TETA = {9: np.array([1.0] + [-0.2] * 4 + [-0.05] * 4), 19: np.array([1.0] + [-1 / 12] * 6 + [-1 / 24] * 12)}
W = {9: np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4), 19: np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)}


def stream_over_slope(q, shape, u, grad, alpha):
    """Red-only populations: the colour-gradient equilibrium of density
    1 + grad.x and uniform velocity u."""
    c = LT.velocities(q)[:, : len(u)].astype(float)
    idx = np.indices(shape)[::-1]                      # x, y(, z)
    rho = 1.0 + sum(g * (i - s / 2) for g, i, s in zip(grad, idx, shape[::-1]))
    cu = c @ np.asarray(u)
    uu = float(np.dot(u, u))
    feq = (PHI[q] + TETA[q] * alpha + W[q] * (3 * cu + 4.5 * cu ** 2 - 1.5 * uu))
    f = feq.reshape((q,) + (1,) * len(shape)) * rho[None]
    return f, np.zeros_like(f)


@pytest.mark.parametrize("three_d", [False, True])
def test_enhanced_correction_moments(oracle, three_d):
    nu, alpha = 0.1, 0.3
    if three_d:
        q, shape, u, grad = 19, (12, 12, 12), (0.03, -0.02, 0.015), (0.004, 0.003, -0.002)
    else:
        q, shape, u, grad = 9, (14, 14), (0.03, -0.02), (0.004, 0.003)
    rf, bf = stream_over_slope(q, shape, u, grad, alpha)
    kw = dict(r_nu=nu, b_nu=nu, r_alpha=alpha, b_alpha=alpha, beta=0.7, A=0.01)
    nz = shape[0] if three_d else 1
    ny, nx = shape[-2], shape[-1]
    runs = []
    for enh in (0, 1):
        s = oracle.CgSim(nx, ny, nz=nz, enhanced=enh, **kw)
        s.init_f(rf, bf)
        s.step(1)
        d = s.get()
        runs.append(d["r_f"] + d["b_f"])
    diff = runs[1] - runs[0]
    inner = (slice(None),) + (slice(3, -3),) * len(shape)
    dq = diff[inner].reshape(q, -1)
    c = LT.velocities(q)[:, : len(u)].astype(float)
    om = 1 / (3 * nu + 0.5)
    ug = float(np.dot(u, grad))
    expect = om * nu * (ug * np.eye(len(u)) + np.outer(u, grad) + np.outer(grad, u))
    scale = om * nu * abs(ug) + 1e-30
    assert np.abs(dq.sum(0)).max() < 1e-15                       # no mass
    assert np.abs(c.T @ dq).max() < 1e-15                        # no momentum
    m2 = np.einsum("qa,qb,qn->abn", c, c, dq)
    assert np.abs(m2 - expect[:, :, None]).max() < 1e-9 * scale + 1e-16
    assert np.abs(expect).max() > 1e-6


def test_enhanced_equals_standard_at_uniform_density(oracle):
    n = 10
    r, b = LT.droplet3d(n, n, n, 3.1)
    rf, bf = LT.bubble3d_populations(r, b)            # rho = 1 everywhere
    out = []
    for enh in (0, 1):
        s = oracle.CgSim(n, n, nz=n, enhanced=enh, A=0.01, r_alpha=0.0, b_alpha=0.0)
        s.init_f(rf, bf)
        s.step(1)
        out.append(s.get()["r_f"])
    assert np.abs(out[1] - out[0]).max() < 1e-17


@pytest.mark.parametrize("three_d", [False, True])
def test_enhanced_conserves_colour_mass_and_momentum(oracle, three_d):
    n = 12
    kw = dict(enhanced=1, beta=0.7, A=0.01, r_alpha=0.3, b_alpha=0.4, r_nu=0.1, b_nu=0.2)
    if three_d:
        r, b = LT.droplet3d(n, n, n, 3.7, cx=5.2, cy=6.1, cz=5.5)
        b = 0.5 * b
        s = oracle.CgSim(n, n, nz=n, **kw)
    else:
        r, b = LT.droplet2d(n, n, 3.7, cx=5.2, cy=6.1)
        b = 0.5 * b
        s = oracle.CgSim(n, n, **kw)
    s.init(r, b)
    s.step(1)
    d0 = s.get()
    s.step(60)
    d = s.get()
    for k in ("r_rho", "b_rho"):
        assert abs(d[k].sum() - d0[k].sum()) < 1e-11 * d0[k].sum()
    q = 19 if three_d else 9
    c = LT.velocities(q)[:, : (3 if three_d else 2)].astype(float)
    f = (d["r_f"] + d["b_f"]).reshape(q, -1)
    assert np.abs((c.T @ f).sum(axis=1)).max() < 1e-12             # total momentum stays zero
    assert np.abs(d["u"]).max() > 1e-6

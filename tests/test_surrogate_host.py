"""Host-side logic of the learned warm start (CPU): input features, the
residual targets and their inversion, the analytic start, the parameter
layout and the data assembly (with the solver replaced by the analytic
start)."""
import numpy as np

from thesis_amd import lattices as LT
from thesis_amd import lbm as L
from thesis_amd import surrogate as S


def test_layout_and_size():
    wo, bo, nw, n = S.layout()
    assert 0.9e6 < n < 1.2e6                      # about one million parameters
    assert all(o % 32 == 0 for o in wo) and nw == bo[0]
    p = S.init_params(seed=1)
    for li, (_, cin, cout, k, _) in enumerate(S.ARCH):
        kp = S.kpad(cin, k)
        blk = p[wo[li]:wo[li] + cout * kp].reshape(cout, kp)
        assert (blk[:, cin * k * k:] == 0).all()  # padded im2col columns stay zero
    assert S.ARCH[0][1] == S.features(LT.surrogate_channels(1, 3)[0]).shape[-1]


def test_analytic_start_is_channel_flow():
    d = LT.surrogate_channels(1, 5)[0]
    rho, u, v = LT.poiseuille_start(d)
    ny, nx = d["node_class"].shape
    assert np.allclose(rho[:, -1][d["node_class"][:, -1] != L.SOLID], 1.0)
    nu = (1 / d["omega"] - 0.5) / 3
    drop = 3 * 8 * nu * d["u_max"] / ny ** 2 * (nx - 1)
    assert abs(rho[0, 0] - 1 - drop) < 1e-12
    assert (u[d["node_class"] == L.SOLID] == 0).all() and (v == 0).all()
    assert np.allclose(u[:, 0], d["ubc"][:, 0])


def test_targets_invert_to_fields():
    d = LT.surrogate_channels(1, 7)[0]
    r = np.random.Generator(np.random.PCG64(2))
    rho = 1 + 0.01 * r.random(d["node_class"].shape)
    u = 0.05 * r.random(d["node_class"].shape)
    v = 0.01 * r.standard_normal(d["node_class"].shape)
    t, m = S.targets(d, rho, u, v)
    r2, u2, v2 = S.fields(d, t)
    fl = (d["node_class"] == L.FLUID) | (d["node_class"] == L.PRES_E)
    assert np.abs(r2 - rho)[fl].max() < 1e-6 and np.abs(u2 - u)[fl].max() < 1e-7
    assert np.abs(v2 - v)[fl].max() < 1e-7
    assert (t[m == 0] == 0).all()


def test_dataset_assembly(monkeypatch):
    def fake_solve(ds, **kw):
        return [LT.poiseuille_start(d) + (0, 0.0) for d in ds], 0.0
    monkeypatch.setattr(S, "solve", fake_solve)
    X, Y, M = S.dataset(3, 11, log=lambda *a: None)
    assert X.shape == (3, 128, 512, S.ARCH[0][1]) and Y.shape == (3, 128, 512, 3) and M.shape == (3, 128, 512)
    assert np.abs(Y).max() < 1e-6                  # the analytic start has zero deviation
    assert (X[..., 0] == 1 - M).all()

"""Pins the FP64 oracle of the warm-start network (oracle/nn_oracle.py):
adjoint identities, a loop convolution, and central finite differences of
the whole network.  CPU only."""
import numpy as np
import pytest

from oracle import nn_oracle as O
from thesis_amd import surrogate as S


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def test_im2col_matches_loop_convolution():
    r = rng(1)
    x = r.standard_normal((2, 5, 7, 3))
    kp = O.kpad(3, 3)
    w = np.zeros((4, kp))
    w[:, :27] = r.standard_normal((4, 27))
    b = r.standard_normal(4)
    y = (O.im2col(x, kp) @ w.T + b).reshape(2, 5, 7, 4)
    assert np.abs(y - O.conv_direct(x, w, b, 3)).max() < 1e-12
    assert (O.im2col(x, kp)[:, 27:] == 0).all()


@pytest.mark.parametrize("op", ["im2col", "pool", "up", "xint", "ymix"])
def test_adjoints(op):
    r = rng(2)
    x = r.standard_normal((2, 4, 8, 3))
    if op == "im2col":
        fx, kp = O.im2col(x, 32), 32
        y = r.standard_normal(fx.shape)
        ty = O.col2im(y, x.shape)
    elif op == "pool":
        fx = O.avgpool2(x)
        y = r.standard_normal(fx.shape)
        ty = O.avgpool2_bwd(y)
    elif op == "up":
        fx = O.up2(x)
        y = r.standard_normal(fx.shape)
        ty = O.up2_bwd(y)
    elif op == "xint":
        fx = O.xint(x)
        y = r.standard_normal(fx.shape)
        ty = O.xint_bwd(y)
    else:
        fx = O.ymix(x)
        y = r.standard_normal(fx.shape)
        ty = O.ymix(y)
    assert abs((fx * y).sum() - (x * ty).sum()) < 1e-10 * (1 + abs((fx * y).sum()))


def test_xint_is_the_running_sum_from_the_outlet():
    x = np.zeros((1, 1, 4, 1))
    x[0, 0, :, 0] = [1.0, 2.0, 3.0, 4.0]
    assert np.allclose(O.xint(x)[0, 0, :, 0], [1 + 10 / 4, 2 + 9 / 4, 3 + 7 / 4, 4 + 4 / 4])


def test_network_gradient_by_central_differences():
    """Directional derivatives of the masked loss along a random direction in
    every layer's weights and biases against the backward pass (FP64, no
    rounding emulation; a step of 1e-7 keeps the ReLU pattern fixed)."""
    r = rng(3)
    p = S.init_params(seed=5).astype(np.float64)
    wo, bo, nw, n = S.layout()
    p[nw:] = 0.05 * r.standard_normal(n - nw)        # non-zero biases exercise every ReLU branch
    net = O.Net(S.ARCH, p)
    x = r.standard_normal((1, 16, 32, S.ARCH[0][1]))
    t = r.standard_normal((1, 16, 32, 3))
    m = (r.random((1, 16, 32)) > 0.2).astype(float)
    out, c = net.forward(x)
    loss, dout = O.mse(out, t, m)
    g = net.backward(dout, c)
    eps = 1e-7
    for li, (name, cin, cout, k, _) in enumerate(S.ARCH):
        d = np.zeros(n)
        kp = S.kpad(cin, k)
        blk = r.standard_normal((cout, kp))
        blk[:, cin * k * k:] = 0.0
        d[wo[li]:wo[li] + cout * kp] = blk.ravel()
        d[bo[li]:bo[li] + cout] = r.standard_normal(cout)
        lp = O.mse(O.Net(S.ARCH, p + eps * d).forward(x)[0], t, m)[0]
        lm = O.mse(O.Net(S.ARCH, p - eps * d).forward(x)[0], t, m)[0]
        fd = (lp - lm) / (2 * eps)
        an = float(g @ d)
        assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-4), (name, fd, an)


def test_gradient_check_catches_a_wrong_skip():
    """The same check fails when the backward drops one skip connection."""
    r = rng(4)
    p = S.init_params(seed=6).astype(np.float64)
    net = O.Net(S.ARCH, p)
    x = r.standard_normal((1, 16, 32, S.ARCH[0][1]))
    t = r.standard_normal((1, 16, 32, 3))
    m = np.ones((1, 16, 32))
    out, c = net.forward(x)
    loss, dout = O.mse(out, t, m)

    class Broken(O.Net):
        def backward(self, dout, c):
            g = np.zeros(self.n)
            d = self.conv_bwd(dout, "out", c, g)
            d = self.conv_bwd(d, "d1b", c, g)
            d = O.up2_bwd(self.conv_bwd(d, "d1a", c, g))
            d = self.conv_bwd(d, "d2b", c, g)
            ds2 = d
            d = O.up2_bwd(self.conv_bwd(d, "d2a", c, g))
            d = self.conv_bwd(d, "d3b", c, g)
            ds3 = d
            d = O.up2_bwd(self.conv_bwd(d, "d3a", c, g))
            d = O.xint_bwd(O.ymix(self.conv_bwd(d, "b2", c, g)))
            d = O.avgpool2_bwd(self.conv_bwd(d, "b1", c, g)) + ds3
            d = self.conv_bwd(self.conv_bwd(d, "e3b", c, g), "e3a", c, g)
            d = O.avgpool2_bwd(d) + ds2
            d = self.conv_bwd(self.conv_bwd(d, "e2b", c, g), "e2a", c, g)
            d = O.avgpool2_bwd(d)                  # skip s1 dropped
            self.conv_bwd(self.conv_bwd(d, "e1b", c, g), "e1a", c, g)
            return g

    good = net.backward(dout, c)
    bad = Broken(S.ARCH, p).backward(dout, c)
    wo = S.layout()[0]
    e1b = slice(wo[1], wo[2])
    assert np.abs(good[e1b] - bad[e1b]).max() > 1e-3 * np.abs(good[e1b]).max()


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float32)
    assert list(O.bf16(v)) == [1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5]

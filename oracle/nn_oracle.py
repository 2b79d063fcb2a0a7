"""FP64 numpy restatement of the warm-start network (test infrastructure).

Only tests, __graft_entry__.smoke() and bench.py's CPU baseline may use this
module, as the checker of the HIP kernels in thesis_amd/csrc/nn_kernels.hip
and of the forward/backward sequence in thesis_amd/surrogate.py; nothing in
the product imports it.  The reference has no surrogate, so the pins are
mathematical: adjoint identities of every data-movement operator, a direct
(loop) convolution, and central finite differences of the whole network
(tests/test_nn_oracle.py).

With emulate=True the operands the GPU rounds to bf16 (im2col columns,
weights, the backward's output gradients) are rounded the same way (to
nearest even) before FP64 products, so the GPU's FP32 results can be compared
tightly.
"""
import numpy as np


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def kpad(cin, k):
    return (cin * k * k + 31) // 32 * 32


def layout(arch):
    off, wo, bo = 0, [], []
    for (_, cin, cout, k, _) in arch:
        wo.append(off)
        off += cout * kpad(cin, k)
    nw = off
    for (_, _, cout, _, _) in arch:
        bo.append(off)
        off += cout
    return wo, bo, nw, off


def im2col(x, kp):
    """x (B, H, W, C) -> (B*H*W, kp); column (ky*3 + kx)*C + c holds
    x[b, y+ky-1, x+kx-1, c] (zero outside), columns >= 9C are zero."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    col = np.zeros((B, H, W, kp))
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            col[..., t * C:(t + 1) * C] = xp[:, ky:ky + H, kx:kx + W]
    return col.reshape(B * H * W, kp)


def col2im(dcol, shape):
    """Adjoint of im2col."""
    B, H, W, C = shape
    d = dcol.reshape(B, H, W, -1)
    xp = np.zeros((B, H + 2, W + 2, C))
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            xp[:, ky:ky + H, kx:kx + W] += d[..., t * C:(t + 1) * C]
    return xp[:, 1:-1, 1:-1]


def conv_direct(x, w, b, cin):
    """3x3 same convolution by loops over taps (a second statement of the
    im2col ordering): y[p, o] = b[o] + sum w[o, (ky*3+kx)*cin + c] x[p + (ky-1, kx-1), c]."""
    B, H, W, C = x.shape
    cout = w.shape[0]
    y = np.zeros((B, H, W, cout)) + b
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            for yy in range(H):
                ys = yy + ky - 1
                if ys < 0 or ys >= H:
                    continue
                for xx in range(W):
                    xs = xx + kx - 1
                    if xs < 0 or xs >= W:
                        continue
                    y[:, yy, xx] += x[:, ys, xs] @ w[:, t * cin:(t + 1) * cin].T
    return y


def avgpool2(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def avgpool2_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(dy):
    B, H, W, C = dy.shape
    return dy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def xint(x):
    """x + (1/W) sum_{x' >= x} x along the width axis."""
    W = x.shape[2]
    return x + np.flip(np.cumsum(np.flip(x, 2), axis=2), 2) / W


def xint_bwd(dy):
    return dy + np.cumsum(dy, axis=2) / dy.shape[2]


def ymix(x):
    """x + the mean over the height axis (symmetric, its own adjoint)."""
    return x + x.mean(axis=1, keepdims=True)


class Net:
    """Forward/backward of the encoder-decoder for a flat parameter vector."""

    def __init__(self, arch, params, emulate=False):
        self.arch = arch
        self.index = {a[0]: i for i, a in enumerate(arch)}
        self.wo, self.bo, self.nw, self.n = layout(arch)
        self.p = np.asarray(params, np.float64)
        self.emulate = emulate

    def _r(self, a):
        return bf16(a) if self.emulate else a

    def W(self, li):
        _, cin, cout, k, _ = self.arch[li]
        w = self.p[self.wo[li]:self.wo[li] + cout * kpad(cin, k)].reshape(cout, kpad(cin, k))
        return self._r(w)

    def b(self, li):
        cout = self.arch[li][2]
        return self.p[self.bo[li]:self.bo[li] + cout]

    def conv(self, x, name, c):
        li = self.index[name]
        _, cin, cout, k, relu = self.arch[li]
        B, H, Wd, _ = x.shape
        col = im2col(x, kpad(cin, k)) if k == 3 else x.reshape(-1, cin)
        col = self._r(col)
        y = col @ self.W(li).T + self.b(li)
        if relu:
            y = np.maximum(y, 0.0)
        y = y.reshape(B, H, Wd, cout)
        c[name] = (col, y, x.shape)
        return y

    def conv_bwd(self, dy, name, c, g):
        li = self.index[name]
        _, cin, cout, k, relu = self.arch[li]
        col, y, xshape = c[name]
        dz = np.where(y > 0, dy, 0.0) if relu else dy
        dz = dz.reshape(-1, cout)
        g[self.bo[li]:self.bo[li] + cout] += dz.sum(axis=0)
        dzr = self._r(dz)
        g[self.wo[li]:self.wo[li] + cout * kpad(cin, k)] += (dzr.T @ col).ravel()
        dcol = dzr @ self.W(li)
        return col2im(dcol, xshape) if k == 3 else dcol.reshape(xshape)

    def forward(self, x):
        c = {}
        a = self.conv(x, "e1a", c)
        s1 = self.conv(a, "e1b", c)
        a = self.conv(avgpool2(s1), "e2a", c)
        s2 = self.conv(a, "e2b", c)
        a = self.conv(avgpool2(s2), "e3a", c)
        s3 = self.conv(a, "e3b", c)
        a = self.conv(avgpool2(s3), "b1", c)
        a = self.conv(ymix(xint(a)), "b2", c)
        a = self.conv(up2(a), "d3a", c) + s3
        a = self.conv(a, "d3b", c)
        a = self.conv(up2(a), "d2a", c) + s2
        a = self.conv(a, "d2b", c)
        a = self.conv(up2(a), "d1a", c) + s1
        a = self.conv(a, "d1b", c)
        return self.conv(a, "out", c), c

    def backward(self, dout, c):
        g = np.zeros(self.n)
        d = self.conv_bwd(dout, "out", c, g)
        d = self.conv_bwd(d, "d1b", c, g)
        ds1 = d
        d = up2_bwd(self.conv_bwd(d, "d1a", c, g))
        d = self.conv_bwd(d, "d2b", c, g)
        ds2 = d
        d = up2_bwd(self.conv_bwd(d, "d2a", c, g))
        d = self.conv_bwd(d, "d3b", c, g)
        ds3 = d
        d = up2_bwd(self.conv_bwd(d, "d3a", c, g))
        d = xint_bwd(ymix(self.conv_bwd(d, "b2", c, g)))
        d = avgpool2_bwd(self.conv_bwd(d, "b1", c, g)) + ds3
        d = self.conv_bwd(self.conv_bwd(d, "e3b", c, g), "e3a", c, g)
        d = avgpool2_bwd(d) + ds2
        d = self.conv_bwd(self.conv_bwd(d, "e2b", c, g), "e2a", c, g)
        d = avgpool2_bwd(d) + ds1
        self.conv_bwd(self.conv_bwd(d, "e1b", c, g), "e1a", c, g)
        return g


def mse(pred, target, mask, wch=(1.0, 1.0, 1.0)):
    """sum_c w_c (pred - target)^2 over mask > 0 pixels, divided by their count;
    returns (loss, dloss/dpred)."""
    m = (mask > 0)[..., None]
    w = np.asarray(wch)
    d = np.where(m, pred - target, 0.0)
    cnt = max(float(m.sum()), 1.0)
    return float((w * d * d).sum() / cnt), 2.0 * w * d / cnt

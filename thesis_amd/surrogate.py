"""Learned warm start for D2Q9 channels: a 4-level convolutional
encoder-decoder trained on the GPU with the library's own HIP kernels.

The reference starts every run from uniform fields or from a macro file
(S/Iterate.cu:211-275, S/FilesReading.cu:224-268); here a network maps the
obstacle mask and inlet velocity of a 512 x 128 channel to its steady
(u, v, rho), and the solver starts from f = f_eq(rho, u) of that prediction.

Network (about 1.0 M parameters, bf16 weights, FP32 accumulation and
activations, NHWC): 3x3 convolutions with ReLU, two per level at 32, 64, 128
and 128 channels, 2x2 average pooling down, nearest x2 upsampling up with
additive skip connections, and at the coarsest level a running integral
along the channel axis (x + (1/W) sum_{x' >= x} x), which carries the
downstream obstacles' pressure drop upstream, followed by a mean over the
channel width added back (x + mean_y x), which couples the two walls.  A 1x1 convolution gives the
deviations ((u - u_p), v, (rho - rho_p)) / u_max.  Inputs: solid mask, the
inlet profile p(y), x/(nx-1) and u_max/0.1.

Device memory is torch's (plumbing); every layer runs a kernel of
include/lbm_nn.h.  Training data comes from the batched solver
(lbm_batch_*), run from rest to a 1e-6 residual.
"""
import json
import math
import time

import numpy as np
import torch

from . import lattices as LT
from . import lbm as L
from . import nn

# name, input channels, output channels, kernel size, ReLU
ARCH = [
    ("e1a", 4, 32, 3, True), ("e1b", 32, 32, 3, True),
    ("e2a", 32, 64, 3, True), ("e2b", 64, 64, 3, True),
    ("e3a", 64, 128, 3, True), ("e3b", 128, 128, 3, True),
    ("b1", 128, 128, 3, True), ("b2", 128, 128, 3, True),
    ("d3a", 128, 128, 3, True), ("d3b", 128, 128, 3, True),
    ("d2a", 128, 64, 3, True), ("d2b", 64, 64, 3, True),
    ("d1a", 64, 32, 3, True), ("d1b", 32, 32, 3, True),
    ("out", 32, 3, 1, False),
]
U_SCALE = 0.1
INDEX = {a[0]: i for i, a in enumerate(ARCH)}


def kpad(cin, k):
    """Columns of a layer's im2col matrix: cin*k*k rounded up to 32."""
    return (cin * k * k + 31) // 32 * 32


def layout():
    """Offsets of every layer's weight block [cout x Kp] and bias in the flat
    parameter vector: all weight blocks first (the bf16 region), then biases."""
    off, wo = 0, []
    for (_, cin, cout, k, _) in ARCH:
        wo.append(off)
        off += cout * kpad(cin, k)
    nw = off
    bo = []
    for (_, _, cout, _, _) in ARCH:
        bo.append(off)
        off += cout
    return wo, bo, nw, off


def init_params(seed=0):
    """He-normal weights for ReLU layers (zero in the padded columns), a small
    output layer, zero biases; FP32 numpy vector in layout() order."""
    wo, bo, nw, n = layout()
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(n, np.float32)
    for li, (_, cin, cout, k, relu) in enumerate(ARCH):
        kp, fan = kpad(cin, k), cin * k * k
        std = math.sqrt(2.0 / fan) if relu else 0.1 / math.sqrt(fan)
        w = np.zeros((cout, kp), np.float32)
        w[:, :fan] = rng.standard_normal((cout, fan)) * std
        p[wo[li]:wo[li] + cout * kp] = w.ravel()
    return p


def features(d):
    """Network input (ny, nx, 4) of a channel dict from lattices.surrogate_channels:
    solid mask, inlet profile p(y) = 4 (y+1/2)(ny-y-1/2)/ny^2 broadcast along x,
    x/(nx-1), and u_max/0.1."""
    cls = d["node_class"]
    ny, nx = cls.shape
    x = np.zeros((ny, nx, 4), np.float32)
    y = np.arange(ny) + 0.5
    x[..., 0] = cls == L.SOLID
    x[..., 1] = (4.0 * y * (ny - y) / ny ** 2)[:, None]
    x[..., 2] = (np.arange(nx) / (nx - 1.0))[None, :]
    x[..., 3] = d["u_max"] / U_SCALE
    return x


def targets(d, rho, u, v):
    """Normalised deviation (ny, nx, 3) of the steady fields from the analytic
    channel start (lattices.poiseuille_start), ((u - u_p), v, (rho - rho_p)) / u_max,
    and the fluid mask (ny, nx)."""
    um = d["u_max"]
    rp, up, _ = LT.poiseuille_start(d)
    t = np.stack([(u - up) / um, v / um, (rho - rp) / um], axis=-1).astype(np.float32)
    mask = (d["node_class"] != L.SOLID).astype(np.float32)
    return np.where(mask[..., None] > 0, t, 0.0).astype(np.float32), mask


def fields(d, pred):
    """Physical (rho, u, v) from a normalised prediction (ny, nx, 3): the
    analytic start plus the predicted deviation, zero velocity in solids, the
    prescribed profile on the inlet column."""
    um = d["u_max"]
    rp, up, _ = LT.poiseuille_start(d)
    solid = d["node_class"] == L.SOLID
    u = np.where(solid, 0.0, up + pred[..., 0] * um)
    v = np.where(solid, 0.0, pred[..., 1] * um)
    rho = np.where(solid, 1.0, rp + pred[..., 2] * um)
    inlet = d["node_class"] == L.VEL_W
    u = np.where(inlet, d["ubc"], u)
    v = np.where(inlet, 0.0, v)
    return rho, u, v


class Net:
    """Parameters, Adam state and the explicit forward/backward passes."""

    def __init__(self, params=None, device="cuda"):
        nn.load()
        self.dev = torch.device(device)
        self.wo, self.bo, self.nw, self.n = layout()
        p = init_params() if params is None else np.asarray(params, np.float32)
        assert p.size == self.n
        self.w = torch.from_numpy(p.copy()).to(self.dev)
        self.wbf = torch.empty(self.nw, dtype=torch.int16, device=self.dev)
        nn.to_bf16(self.w[:self.nw], self.wbf)
        self.g = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.g)
        self.v = torch.zeros_like(self.g)
        self.t = 0

    # -- layers -----------------------------------------------------------
    def _conv(self, x, Bn, H, W, li, cache):
        _, cin, cout, k, relu = ARCH[li]
        kp = kpad(cin, k)
        M = Bn * H * W
        col = torch.empty(M * kp, dtype=torch.int16, device=self.dev)
        if k == 3:
            nn.im2col(x, Bn, H, W, cin, kp, col)
        else:
            assert kp == cin
            nn.to_bf16(x, col)
        y = torch.empty(M * cout, dtype=torch.float32, device=self.dev)
        wo, bo = self.wo[li], self.bo[li]
        nn.gemm(M, cout, kp, col, kp, 0, self.wbf[wo:wo + cout * kp], kp, 0, y, cout,
                self.w[bo:bo + cout], nn.RELU if relu else 0, 1)
        cache[li] = (col, y, Bn, H, W)
        return y

    def _conv_bwd(self, dy, li, cache, need_dx=True):
        _, cin, cout, k, relu = ARCH[li]
        kp = kpad(cin, k)
        col, y, Bn, H, W = cache[li]
        M = Bn * H * W
        wo, bo = self.wo[li], self.bo[li]
        dz = torch.empty(M * cout, dtype=torch.int16, device=self.dev)
        nn.grad_bf16(dy, y, relu, dz)
        nn.colsum(dy, y, relu, M, cout, self.g[bo:bo + cout])
        bm = 256 if cout <= 32 else 128
        tiles = -(-kp // bm) * -(-cout // (32 if cout <= 32 else 64 if cout <= 64 else 128))
        splits = max(1, min(2048 // max(tiles, 1), M // 2048))
        nn.gemm(kp, cout, M, col, kp, 1, dz, cout, 1, self.g[wo:wo + cout * kp], kp, None,
                nn.ATOMIC | nn.TRANSC, splits)
        if not need_dx:
            return None
        dcol = torch.empty(M * kp, dtype=torch.float32, device=self.dev)
        nn.gemm(M, kp, cout, dz, cout, 0, self.wbf[wo:wo + cout * kp], kp, 1, dcol, kp, None, 0, 1)
        if k == 1:
            return dcol
        dx = torch.empty(M * cin, dtype=torch.float32, device=self.dev)
        nn.col2im(dcol, Bn, H, W, cin, kp, dx)
        return dx

    def _new(self, n):
        return torch.empty(n, dtype=torch.float32, device=self.dev)

    # -- network ----------------------------------------------------------
    def forward(self, x, Bn, H, W):
        """x: FP32 device tensor (Bn*H*W*3) NHWC; returns (out, cache)."""
        c = {}
        I = INDEX
        a = self._conv(x, Bn, H, W, I["e1a"], c)
        s1 = self._conv(a, Bn, H, W, I["e1b"], c)
        H2, W2 = H // 2, W // 2
        p = self._new(Bn * H2 * W2 * 32)
        nn.avgpool2(s1, Bn, H, W, 32, p)
        a = self._conv(p, Bn, H2, W2, I["e2a"], c)
        s2 = self._conv(a, Bn, H2, W2, I["e2b"], c)
        H3, W3 = H2 // 2, W2 // 2
        p = self._new(Bn * H3 * W3 * 64)
        nn.avgpool2(s2, Bn, H2, W2, 64, p)
        a = self._conv(p, Bn, H3, W3, I["e3a"], c)
        s3 = self._conv(a, Bn, H3, W3, I["e3b"], c)
        H4, W4 = H3 // 2, W3 // 2
        p = self._new(Bn * H4 * W4 * 128)
        nn.avgpool2(s3, Bn, H3, W3, 128, p)
        a = self._conv(p, Bn, H4, W4, I["b1"], c)
        xi = self._new(a.numel())
        nn.xint(a, Bn, H4, W4, 128, 0, xi)
        ym = self._new(a.numel())
        nn.ymix(xi, Bn, H4, W4, 128, ym)
        a = self._conv(ym, Bn, H4, W4, I["b2"], c)
        u = self._new(Bn * H3 * W3 * 128)
        nn.up2(a, Bn, H4, W4, 128, u)
        a = self._conv(u, Bn, H3, W3, I["d3a"], c)
        t = self._new(a.numel())
        nn.add(a, s3, t)        # the convolution output stays intact for its ReLU mask
        a = t
        a = self._conv(a, Bn, H3, W3, I["d3b"], c)
        u = self._new(Bn * H2 * W2 * 128)
        nn.up2(a, Bn, H3, W3, 128, u)
        a = self._conv(u, Bn, H2, W2, I["d2a"], c)
        t = self._new(a.numel())
        nn.add(a, s2, t)        # the convolution output stays intact for its ReLU mask
        a = t
        a = self._conv(a, Bn, H2, W2, I["d2b"], c)
        u = self._new(Bn * H * W * 64)
        nn.up2(a, Bn, H2, W2, 64, u)
        a = self._conv(u, Bn, H, W, I["d1a"], c)
        t = self._new(a.numel())
        nn.add(a, s1, t)        # the convolution output stays intact for its ReLU mask
        a = t
        a = self._conv(a, Bn, H, W, I["d1b"], c)
        out = self._conv(a, Bn, H, W, I["out"], c)
        c["dims"] = (Bn, H, W)
        return out, c

    def backward(self, dout, c):
        """Accumulates parameter gradients of the loss whose output gradient
        is dout into self.g (zero it first)."""
        I = INDEX
        Bn, H, W = c["dims"]
        H2, W2, H3, W3, H4, W4 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        d = self._conv_bwd(dout, I["out"], c)
        d = self._conv_bwd(d, I["d1b"], c)            # gradient of relu(d1a) + s1
        ds1 = d
        d = self._conv_bwd(d, I["d1a"], c)
        da = self._new(Bn * H2 * W2 * 64)
        nn.up2_bwd(d, Bn, H2, W2, 64, da)
        d = self._conv_bwd(da, I["d2b"], c)
        ds2 = d
        d = self._conv_bwd(d, I["d2a"], c)
        da = self._new(Bn * H3 * W3 * 128)
        nn.up2_bwd(d, Bn, H3, W3, 128, da)
        d = self._conv_bwd(da, I["d3b"], c)
        ds3 = d
        d = self._conv_bwd(d, I["d3a"], c)
        da = self._new(Bn * H4 * W4 * 128)
        nn.up2_bwd(d, Bn, H4, W4, 128, da)
        d = self._conv_bwd(da, I["b2"], c)
        dym = self._new(d.numel())
        nn.ymix(d, Bn, H4, W4, 128, dym)
        dxi = self._new(d.numel())
        nn.xint(dym, Bn, H4, W4, 128, 1, dxi)
        d = self._conv_bwd(dxi, I["b1"], c)
        dp = self._new(Bn * H3 * W3 * 128)
        nn.avgpool2_bwd(d, Bn, H3, W3, 128, dp)
        nn.axpy(1.0, ds3, dp)
        d = self._conv_bwd(dp, I["e3b"], c)
        d = self._conv_bwd(d, I["e3a"], c)
        dp = self._new(Bn * H2 * W2 * 64)
        nn.avgpool2_bwd(d, Bn, H2, W2, 64, dp)
        nn.axpy(1.0, ds2, dp)
        d = self._conv_bwd(dp, I["e2b"], c)
        d = self._conv_bwd(d, I["e2a"], c)
        dp = self._new(Bn * H * W * 32)
        nn.avgpool2_bwd(d, Bn, H, W, 32, dp)
        nn.axpy(1.0, ds1, dp)
        d = self._conv_bwd(dp, I["e1b"], c)
        self._conv_bwd(d, I["e1a"], c, need_dx=False)

    def loss_grad(self, out, target, mask, wch):
        """Masked MSE (per fluid pixel, summed over channels) and its gradient."""
        npix = mask.numel()
        grad = torch.empty_like(out)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        inv = 1.0 / max(float(mask.sum().item()), 1.0)
        nn.mse(out, target, mask, wch, npix, 3, inv, grad, loss)
        return loss, grad

    def step(self, x, target, mask, Bn, H, W, lr, wch):
        nn.fill(0.0, self.g)
        out, c = self.forward(x, Bn, H, W)
        loss, grad = self.loss_grad(out, target, mask, wch)
        self.backward(grad, c)
        self.t += 1
        nn.adam(self.w, self.g, self.m, self.v, lr, 0.9, 0.999, 1e-8, self.t, self.wbf, self.nw)
        return loss

    def predict(self, X, Bn, H, W):
        out, _ = self.forward(X, Bn, H, W)
        return out

    # -- persistence --------------------------------------------------------
    def save(self, path, meta=None):
        w = self.w.detach().cpu().numpy()
        np.savez(path, wbf=self.wbf.detach().cpu().numpy().view(np.uint16), bias=w[self.nw:],
                 meta=np.frombuffer(json.dumps(meta or {}).encode(), dtype=np.uint8))

    @classmethod
    def load(cls, path, device="cuda"):
        z = np.load(path, allow_pickle=False)
        wbf = z["wbf"].astype(np.uint32) << 16
        p = np.concatenate([wbf.view(np.float32), z["bias"].astype(np.float32)])
        return cls(p, device=device)


# ---------------------------------------------------------------------------
# data, training and the warm-start comparison


def solve(ds, tol=1e-6, every=100, max_steps=400000, starts=None, chunk=64, then=None):
    """Runs channels to the residual with the batched solver (FP64), 64 per
    persistent launch; returns per channel (rho, u, v, steps, residual) and
    the device milliseconds of the launches.  With `then`, each batch then
    continues from its exact state to that tighter residual, and a second
    list (the references) is returned too."""
    out, refs, total_ms = [], [], 0.0
    for k in range(0, len(ds), chunk):
        part = ds[k:k + chunk]
        b = L.Batch(part, collision=L.TRT, storage="f64")
        if starts is not None:
            for i, s in enumerate(starts[k:k + chunk]):
                b.init(i, *s)
        steps, res, ms = b.run(tol, every=every, max_steps=max_steps)
        total_ms += ms
        for i in range(len(part)):
            rho, u, v = b.macros(i)
            out.append((rho, u, v, int(steps[i]), float(res[i])))
        if then is not None:
            steps2, res2, _ = b.run(then, every=every, max_steps=max_steps)
            for i in range(len(part)):
                rho, u, v = b.macros(i)
                refs.append((rho, u, v, int(steps[i] + steps2[i]), float(res2[i])))
        b.close()
    return (out, total_ms, refs) if then is not None else (out, total_ms)


def dataset(count, seed, log=print):
    """Input features, normalised targets and masks (numpy float32) for
    `count` channels of lattices.surrogate_channels(count, seed)."""
    ds = LT.surrogate_channels(count, seed)
    ny, nx = ds[0]["node_class"].shape
    X = np.zeros((count, ny, nx, ARCH[0][1]), np.float32)
    Y = np.zeros((count, ny, nx, 3), np.float32)
    Mk = np.zeros((count, ny, nx), np.float32)
    t0 = time.time()
    for k in range(0, count, 64):
        res, ms = solve(ds[k:k + 64])
        for i, (rho, u, v, _, _) in enumerate(res):
            d = ds[k + i]
            X[k + i] = features(d)
            Y[k + i], Mk[k + i] = targets(d, rho, u, v)
        log(f"data {min(k + 64, count)}/{count} ({time.time() - t0:.0f} s)")
    return X, Y, Mk


def train(X, Y, Mk, epochs=30, batch=16, lr=2e-3, seed=0, log=print, net=None, rho_weight=1.0, flip=False):
    """Adam with a cosine learning-rate decay over the epochs; rho_weight
    weighs the density channel of the loss against the two velocity channels.
    flip: each sample is mirrored across the channel's centre line with
    probability 1/2 (the walls, the inlet profile and the analytic start are
    symmetric, so the mirrored obstacles' steady state is the mirrored state
    with v negated)."""
    dev = torch.device("cuda")
    net = net or Net()
    n, H, W, _ = X.shape
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.from_numpy(Y).to(dev)
    Md = torch.from_numpy(Mk).to(dev)
    wch = torch.tensor([1.0, 1.0, float(rho_weight)], dtype=torch.float32, device=dev)
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    history = []
    t0 = time.time()
    it = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(dev)
        acc = torch.zeros(1, dtype=torch.float32, device=dev)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            xs_, ys_, ms_ = Xd.index_select(0, idx), Yd.index_select(0, idx), Md.index_select(0, idx)
            if flip:
                f = (torch.rand(batch, generator=g) < 0.5).to(dev)
                sgn = torch.tensor([1.0, -1.0, 1.0], dtype=torch.float32, device=dev)
                xs_ = torch.where(f[:, None, None, None], torch.flip(xs_, dims=[1]), xs_)
                ys_ = torch.where(f[:, None, None, None], torch.flip(ys_, dims=[1]) * sgn, ys_)
                ms_ = torch.where(f[:, None, None], torch.flip(ms_, dims=[1]), ms_)
            xb = xs_.contiguous().view(-1)
            yb = ys_.contiguous().view(-1)
            mb = ms_.contiguous().view(-1)
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * it / total))
            acc += net.step(xb, yb, mb, batch, H, W, lr_t, wch)
            it += 1
        mean = float(acc.item()) / steps_per_epoch
        history.append(mean)
        log(f"epoch {ep + 1}/{epochs} loss {mean:.3e} ({time.time() - t0:.0f} s)")
    return net, history


def predict_fields(net, ds, batch=16):
    """Physical (rho, u, v) predictions for channel dicts."""
    dev = torch.device("cuda")
    ny, nx = ds[0]["node_class"].shape
    out = []
    for k in range(0, len(ds), batch):
        part = ds[k:k + batch]
        X = torch.from_numpy(np.stack([features(d) for d in part])).to(dev).view(-1)
        P = net.predict(X, len(part), ny, nx).view(len(part), ny, nx, 3).cpu().numpy()
        out += [fields(d, P[i]) for i, d in enumerate(part)]
    return out


def warm_start_single(net, d, tol=1e-6, every=100, max_steps=400000):
    """Wall time of one warm-started single-lattice run, inference and
    injection included: returns (steps, seconds, solver)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rho, u, v = predict_fields(net, [d])[0]
    s = L.Solver(9, d["node_class"].shape, d["node_class"], d["bb"], collision=L.TRT, omega=d["omega"],
                 trt_magic=d["trt_magic"], ubc=d["ubc"], vbc=d["vbc"], rho_out=d["rho_out"], storage="f64")
    s.init(0, rho=rho, u=u, v=v)
    n, _ = s.run_to_residual(tol, every=every, max_steps=max_steps)
    return n, time.perf_counter() - t0, s


def rest_start_single(d, tol=1e-6, every=100, max_steps=400000):
    t0 = time.perf_counter()
    s = L.Solver(9, d["node_class"].shape, d["node_class"], d["bb"], collision=L.TRT, omega=d["omega"],
                 trt_magic=d["trt_magic"], ubc=d["ubc"], vbc=d["vbc"], rho_out=d["rho_out"], storage="f64")
    s.init(0)
    n, _ = s.run_to_residual(tol, every=every, max_steps=max_steps)
    return n, time.perf_counter() - t0, s


def rel_err(a, b, mask):
    """(max, L2) velocity error of field a against b over mask, relative to b."""
    ua, va = a
    ub, vb = b
    d = np.sqrt((ua - ub) ** 2 + (va - vb) ** 2)[mask]
    s = np.sqrt(ub ** 2 + vb ** 2)[mask]
    return float(d.max() / s.max()), float(np.sqrt((d ** 2).sum() / (s ** 2).sum()))


def evaluate(net, test, log=print, single=16, tol=1e-6):
    """Rest, analytic (Poiseuille) and learned starts of the held-out channels
    to the residual: steps, device time of the batched launches, single-lattice
    wall time (inference and injection included) on the first `single`
    channels, and errors against a 1e-8 reference."""
    rep = {}
    t0 = time.time()
    rest, ms_rest, ref = solve(test, tol=tol, then=1e-8)
    log(f"rest starts and their 1e-8 continuations done ({time.time() - t0:.0f} s)")
    pois, ms_pois = solve(test, tol=tol, starts=[LT.poiseuille_start(d) for d in test])
    torch.cuda.synchronize()
    ti = time.perf_counter()
    pred = predict_fields(net, test)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - ti
    warm, ms_warm = solve(test, tol=tol, starts=pred)
    log(f"warm starts done ({time.time() - t0:.0f} s)")
    rows = []
    for i, d in enumerate(test):
        fl = d["node_class"] != L.SOLID
        rr = (ref[i][1], ref[i][2])
        rows.append(dict(
            steps_rest=rest[i][3], steps_pois=pois[i][3], steps_warm=warm[i][3],
            err_rest=rel_err((rest[i][1], rest[i][2]), rr, fl), err_pois=rel_err((pois[i][1], pois[i][2]), rr, fl),
            err_warm=rel_err((warm[i][1], warm[i][2]), rr, fl), err_pred=rel_err((pred[i][1], pred[i][2]), rr, fl),
            err_pois_start=rel_err(LT.poiseuille_start(d)[1:], rr, fl),
            err_rho_pred=float(np.abs(pred[i][0] - ref[i][0])[fl].max() / np.abs(ref[i][0] - 1.0)[fl].max())))
    rep["per_channel"] = rows
    sr = np.array([r["steps_rest"] for r in rows], float)
    sw = np.array([r["steps_warm"] for r in rows], float)
    sp = np.array([r["steps_pois"] for r in rows], float)
    rep["steps"] = dict(rest_mean=sr.mean(), warm_mean=sw.mean(), pois_mean=sp.mean(),
                        ratio_mean=float((sr / sw).mean()), ratio_total=float(sr.sum() / sw.sum()))
    rep["batched_ms"] = dict(rest=ms_rest, pois=ms_pois, warm=ms_warm, inference_s=t_inf,
                             speedup=ms_rest / (ms_warm + 1e3 * t_inf))
    e = np.array([r["err_rho_pred"] for r in rows])
    rep["err_rho_pred"] = dict(mean=float(e.mean()), max=float(e.max()))
    for key in ("err_rest", "err_pois", "err_warm", "err_pred", "err_pois_start"):
        e = np.array([r[key] for r in rows])
        rep[key] = dict(max_mean=float(e[:, 0].mean()), max_max=float(e[:, 0].max()), l2_mean=float(e[:, 1].mean()),
                        l2_max=float(e[:, 1].max()))
    # single-lattice wall time, inference and injection included
    wr, ww = [], []
    for d in test[:single]:
        n_r, t_r, s = rest_start_single(d, tol=tol)
        s.close()
        n_w, t_w, s = warm_start_single(net, d, tol=tol)
        s.close()
        wr.append(t_r)
        ww.append(t_w)
    rep["single_wall_s"] = dict(rest_mean=float(np.mean(wr)), warm_mean=float(np.mean(ww)),
                                speedup=float(np.sum(wr) / np.sum(ww)), n=len(wr))
    log(f"single-lattice runs done ({time.time() - t0:.0f} s)")
    return rep


def start_field_ablation(net, test, log=print):
    """Which predicted field limits the warm start: steps to the residual of
    the batched solver from starts that mix the network's density and
    velocity with the converged fields (a rest run continued to 1e-8), and
    with the analytic density.  Returns {start: (mean steps, rest/start step ratio)}."""
    rest, _, ref = solve(test, then=1e-8)
    pred = predict_fields(net, test)
    sr = np.array([r[3] for r in rest], float)
    starts = {
        "network rho, network u": pred,
        "converged rho, network u": [(r[0], p[1], p[2]) for p, r in zip(pred, ref)],
        "network rho, converged u": [(p[0], r[1], r[2]) for p, r in zip(pred, ref)],
        "converged rho, converged u": [(r[0], r[1], r[2]) for r in ref],
        "analytic rho, network u": [(LT.poiseuille_start(d)[0], p[1], p[2]) for d, p in zip(test, pred)],
    }
    out = {"rest": (float(sr.mean()), 1.0)}
    for k, st in starts.items():
        res, _ = solve(test, starts=st)
        sw = np.array([r[3] for r in res], float)
        out[k] = (float(sw.mean()), float(sr.sum() / sw.sum()))
        log(f"{k}: mean steps {sw.mean():.0f}, {sr.sum() / sw.sum():.3f} x fewer than from rest")
    return out


def main(argv=None):
    import argparse
    import os
    ap = argparse.ArgumentParser(description="train and evaluate the learned warm start on the GPU")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--train-seed", type=int, default=1)
    ap.add_argument("--test-seed", type=int, default=9001)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--rho-weight", type=float, default=1.0, help="loss weight of the density channel")
    ap.add_argument("--flip", action="store_true", help="mirror training samples across the centre line")
    ap.add_argument("--out", default="surrogate_out")
    ap.add_argument("--weights", default=None, help="skip training, evaluate these weights")
    ap.add_argument("--ablate", type=int, default=0,
                    help="with --weights: only the start-field ablation on this many held-out channels")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    t0 = time.time()

    def log(msg):
        print(f"[{time.time() - t0:7.1f} s] {msg}", flush=True)

    report = dict(args=vars(a))
    if a.weights and a.ablate:
        report["ablation"] = start_field_ablation(Net.load(a.weights), LT.surrogate_channels(a.ablate, a.test_seed),
                                                  log=log)
        with open(os.path.join(a.out, "ablation.json"), "w") as f:
            json.dump(report, f, indent=1, default=float)
        return
    if a.weights:
        net = Net.load(a.weights)
    else:
        X, Y, Mk = dataset(a.train, a.train_seed, log=log)
        report["data_s"] = time.time() - t0
        t1 = time.time()
        net, hist = train(X, Y, Mk, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, rho_weight=a.rho_weight,
                          flip=a.flip)
        report["train_s"] = time.time() - t1
        report["loss_history"] = hist
        del X, Y, Mk
        net.save(os.path.join(a.out, "weights.npz"),
                 meta=dict(train=a.train, seed=a.train_seed, epochs=a.epochs, rho_weight=a.rho_weight, flip=a.flip))
    if a.test == 0:
        # training only (e.g. to profile the kernels of the training steps)
        with open(os.path.join(a.out, "report.json"), "w") as f:
            json.dump(report, f, indent=1, default=float)
        return
    test = LT.surrogate_channels(a.test, a.test_seed)
    report["eval"] = evaluate(net, test, log=log)
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1, default=float)
    ev = report["eval"]
    log(f"steps rest {ev['steps']['rest_mean']:.0f} warm {ev['steps']['warm_mean']:.0f} "
        f"poiseuille {ev['steps']['pois_mean']:.0f}; batched speedup {ev['batched_ms']['speedup']:.2f}; "
        f"single-lattice wall speedup {ev['single_wall_s']['speedup']:.2f}")
    log(f"errors vs 1e-8 reference (max-norm mean): rest {ev['err_rest']['max_mean']:.2e} "
        f"warm {ev['err_warm']['max_mean']:.2e} poiseuille {ev['err_pois']['max_mean']:.2e} "
        f"prediction {ev['err_pred']['max_mean']:.2e} (L2 {ev['err_pred']['l2_mean']:.2e}; analytic start "
        f"{ev['err_pois_start']['max_mean']:.2e}, L2 {ev['err_pois_start']['l2_mean']:.2e})")


if __name__ == "__main__":
    main()

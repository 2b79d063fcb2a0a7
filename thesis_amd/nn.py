"""ctypes binding of the warm-start kernels (include/lbm_nn.h).

Device buffers are torch tensors (memory and stream plumbing only); every
operation is one of the library's HIP kernels.  bf16 operands are held in
int16 tensors (raw 16-bit words).  Missing library or device raises; there
is no CPU path."""
import ctypes as C

import torch

from . import lbm as L

RELU, ATOMIC, TRANSC = 0x1, 0x2, 0x4

EXPORTED = [
    "nn_gemm_bf16", "nn_im2col3x3", "nn_col2im3x3", "nn_grad_bf16", "nn_colsum", "nn_to_bf16", "nn_avgpool2",
    "nn_avgpool2_bwd", "nn_up2", "nn_up2_bwd", "nn_xint", "nn_ymix", "nn_axpy", "nn_add", "nn_fill", "nn_mse", "nn_adam",
]

_lib = None
vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float


def load():
    global _lib
    if _lib is not None:
        return _lib
    lib = L.load()
    sig = {
        "nn_gemm_bf16": [i32, i32, i32, vp, i32, i32, vp, i32, i32, vp, i32, vp, i32, i32, vp],
        "nn_im2col3x3": [vp, i32, i32, i32, i32, i32, vp, vp],
        "nn_col2im3x3": [vp, i32, i32, i32, i32, i32, vp, vp],
        "nn_grad_bf16": [vp, vp, i64, i32, vp, vp],
        "nn_colsum": [vp, vp, i32, i32, i32, vp, vp],
        "nn_to_bf16": [vp, i64, vp, vp],
        "nn_avgpool2": [vp, i32, i32, i32, i32, vp, vp],
        "nn_avgpool2_bwd": [vp, i32, i32, i32, i32, vp, vp],
        "nn_up2": [vp, i32, i32, i32, i32, vp, vp],
        "nn_up2_bwd": [vp, i32, i32, i32, i32, vp, vp],
        "nn_xint": [vp, i32, i32, i32, i32, i32, vp, vp],
        "nn_ymix": [vp, i32, i32, i32, i32, vp, vp],
        "nn_axpy": [f32, vp, i64, vp, vp],
        "nn_add": [vp, vp, i64, vp, vp],
        "nn_fill": [f32, i64, vp, vp],
        "nn_mse": [vp, vp, vp, vp, i64, i32, f32, vp, vp, vp],
        "nn_adam": [vp, vp, vp, vp, i64, f32, f32, f32, f32, i64, vp, i64, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(rc, name):
    if rc:
        raise L.LbmError(4 if rc == 4 else 3, f"{name} failed ({rc})")


def gemm(M, N, R, A, lda, ta, B, ldb, tb, Cm, ldc, bias=None, flags=0, splits=1):
    _check(load().nn_gemm_bf16(M, N, R, _p(A), lda, int(ta), _p(B), ldb, int(tb), _p(Cm), ldc, _p(bias), flags,
                               splits, _stream()), "nn_gemm_bf16")


def im2col(x, Bn, H, W, Cc, Kp, col):
    _check(load().nn_im2col3x3(_p(x), Bn, H, W, Cc, Kp, _p(col), _stream()), "nn_im2col3x3")


def col2im(dcol, Bn, H, W, Cc, Kp, dx):
    _check(load().nn_col2im3x3(_p(dcol), Bn, H, W, Cc, Kp, _p(dx), _stream()), "nn_col2im3x3")


def grad_bf16(dy, y, relu, dz):
    _check(load().nn_grad_bf16(_p(dy), _p(y), dy.numel(), int(relu), _p(dz), _stream()), "nn_grad_bf16")


def colsum(dy, y, relu, M, N, db):
    _check(load().nn_colsum(_p(dy), _p(y), int(relu), M, N, _p(db), _stream()), "nn_colsum")


def to_bf16(x, y):
    _check(load().nn_to_bf16(_p(x), x.numel(), _p(y), _stream()), "nn_to_bf16")


def avgpool2(x, Bn, H, W, Cc, y):
    _check(load().nn_avgpool2(_p(x), Bn, H, W, Cc, _p(y), _stream()), "nn_avgpool2")


def avgpool2_bwd(dy, Bn, H, W, Cc, dx):
    _check(load().nn_avgpool2_bwd(_p(dy), Bn, H, W, Cc, _p(dx), _stream()), "nn_avgpool2_bwd")


def up2(x, Bn, H, W, Cc, y):
    _check(load().nn_up2(_p(x), Bn, H, W, Cc, _p(y), _stream()), "nn_up2")


def up2_bwd(dy, Bn, H, W, Cc, dx):
    _check(load().nn_up2_bwd(_p(dy), Bn, H, W, Cc, _p(dx), _stream()), "nn_up2_bwd")


def xint(x, Bn, H, W, Cc, bwd, y):
    _check(load().nn_xint(_p(x), Bn, H, W, Cc, int(bwd), _p(y), _stream()), "nn_xint")


def ymix(x, Bn, H, W, Cc, y):
    _check(load().nn_ymix(_p(x), Bn, H, W, Cc, _p(y), _stream()), "nn_ymix")


def axpy(a, x, y):
    _check(load().nn_axpy(float(a), _p(x), x.numel(), _p(y), _stream()), "nn_axpy")


def add(x, y, z):
    _check(load().nn_add(_p(x), _p(y), x.numel(), _p(z), _stream()), "nn_add")


def fill(v, y):
    _check(load().nn_fill(float(v), y.numel(), _p(y), _stream()), "nn_fill")


def mse(pred, target, mask, wch, npix, Cc, inv_count, grad, loss):
    _check(load().nn_mse(_p(pred), _p(target), _p(mask), _p(wch), npix, Cc, float(inv_count), _p(grad), _p(loss),
                         _stream()), "nn_mse")


def adam(w, g, m, v, lr, b1, b2, eps, t, wbf, n_bf):
    _check(load().nn_adam(_p(w), _p(g), _p(m), _p(v), w.numel(), float(lr), float(b1), float(b2), float(eps), int(t),
                          _p(wbf), int(n_bf), _stream()), "nn_adam")

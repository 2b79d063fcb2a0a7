"""lbm_residual_l2_noalloc (no device allocation: populations rebuilt in
registers, partial sums in the library's static device scratch) against
lbm_residual_l2 (dense FP64 staging array), through the C ABI: two calls in a
row on different contexts, the second on a smaller lattice, and repeated
calls on one context, FP64 and 16-bit storage, periodic and open lattices."""
import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu


def _channel(storage):
    ch = LT.channel2d(48, 20, 0.05, obstacles=[(20, 9, 4)])
    s = L.Solver(9, (20, 48), ch["node_class"], ch["bb"], collision=L.TRT, omega=1 / 0.8, ubc=ch["ubc"],
                 vbc=ch["vbc"], rho_out=1.0, storage=storage)
    return s


def _porous(n, storage):
    cls, bb, _ = LT.porous3d(n, seed=7, rmin=2.0, rmax=3.5)
    return L.Solver(19, (n, n, n), cls, bb, collision=L.TRT, omega=1 / 0.9, trt_magic=0.25,
                    force_model=L.FORCE_GUO, g=(1e-4, 0, 0), periodic=(1, 1, 1), storage=storage)


def _check(s, steps, tol):
    s.init(0)
    s.step(steps)
    a = s.residual_l2_noalloc()
    b = s.residual_l2()
    print(f"noalloc {a:.17e} staged {b:.17e}", flush=True)
    assert np.isfinite(a) and a > 0
    assert abs(a - b) <= tol * b


@pytest.mark.parametrize("storage", [L.STORE_F64, L.STORE_F16])
def test_noalloc_l2_two_calls_second_smaller(storage):
    if L.device_count() < 1:
        pytest.skip("no HIP device")
    tol = 1e-13
    print(f"case: noalloc L2 porous 32^3 then 12^3, storage {storage}", flush=True)
    big = _porous(32, storage)
    _check(big, 25, tol)
    small = _porous(12, storage)
    _check(small, 9, tol)
    big.close()
    small.close()


def test_noalloc_l2_open_channel_repeated():
    if L.device_count() < 1:
        pytest.skip("no HIP device")
    print("case: noalloc L2 open channel 48x20, repeated calls", flush=True)
    s = _channel(L.STORE_F32)
    s.init(0)
    for k in range(3):
        s.step(11)
        a = s.residual_l2_noalloc()
        b = s.residual_l2()
        print(f"call {k}: noalloc {a:.17e} staged {b:.17e}", flush=True)
        assert abs(a - b) <= 1e-13 * b
    s.close()


def test_noalloc_l2_needs_a_step():
    if L.device_count() < 1:
        pytest.skip("no HIP device")
    s = _porous(8, L.STORE_F64)
    s.init(0)
    with pytest.raises(L.LbmError):
        s.residual_l2_noalloc()
    s.close()

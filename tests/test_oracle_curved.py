"""Pins the oracle's curved walls (Bouzidi linear interpolation, the
textbook form of S/GpuBoundaries.cu:414-431): q = 1/2 is half-way
bounce-back exactly, a force-driven channel with walls off the half-way
positions converges at second order to the parabola through the true walls,
and the wall fractions read from latticegen2 coordinates.  CPU only."""
import numpy as np
import pytest

from thesis_amd import lattices as LT

C9 = LT.velocities(9)


def shifted_channel(oracle, ny, d1, d2, wall_model, nx=4, tau=0.8, g=1e-5, steps=None):
    cls = np.ones((ny, nx), np.uint8)
    bb = np.zeros((ny, nx), np.uint32)
    q = np.zeros((ny, nx, 19))
    for d in range(1, 9):
        if C9[d, 1] < 0:
            bb[0] |= 1 << d
            q[0, :, d] = d1
        if C9[d, 1] > 0:
            bb[-1] |= 1 << d
            q[-1, :, d] = d2
    s = oracle.Sim(9, (ny, nx), cls, bb, collision=1, omega=1 / tau, trt_magic=3 / 16,
                   force_model=oracle.FORCE_GUO, g=(g, 0, 0), periodic=(1, 0, 0), wall_model=wall_model)
    s.set_wall_q(q)
    s.init(0)
    s.step(steps if steps else int(30000 * (ny / 12) ** 2))
    u = s.macros()[1][0][:, 0]
    nu = (tau - 0.5) / 3
    y = np.arange(ny)
    want = g / (2 * nu) * (y + d1) * (ny - 1 + d2 - y)
    return s, np.abs(u - want).max() / want.max()


def test_half_fraction_is_half_way_bounce_back(oracle):
    a, _ = shifted_channel(oracle, 10, 0.5, 0.5, oracle.WALL_CURVED, steps=300)
    b, _ = shifted_channel(oracle, 10, 0.5, 0.5, oracle.WALL_HALFWAY, steps=300)
    assert (a.get_f() == b.get_f()).all()




@pytest.mark.parametrize("d1,d2", [(0.3, 0.7), (0.2, 0.2), (0.8, 0.6)])
def test_shifted_walls_converge_at_second_order(oracle, d1, d2):
    _, e12 = shifted_channel(oracle, 12, d1, d2, oracle.WALL_CURVED)
    _, e24 = shifted_channel(oracle, 24, d1, d2, oracle.WALL_CURVED)
    _, h12 = shifted_channel(oracle, 12, d1, d2, oracle.WALL_HALFWAY)
    print(d1, d2, e12, e24, h12)
    assert e12 < 0.015 and e12 / e24 > 3.5 and h12 > 4 * e12


def test_read_wall_q_from_coordinates(tmp_path):
    from thesis_amd import lbm as L
    node = tmp_path / "c_Node.dat"
    bc = tmp_path / "c_BC.dat"
    node.write_text("".join(f"{i} {j} {0.5 * i:.6f} {0.5 * j:.6f} 1\n" for j in range(3) for i in range(4)))
    # node (1, 0): wall below at 0.3 of the link (dir 4), the diagonal dir 7 at 0.6
    bc.write_text("1 0 4 1 0.500000 -0.150000 1\n1 0 7 1 0.200000 -0.300000 1\n2 2 2 2 1.0 1.25 1\n")
    q = L.read_wall_q(str(node), str(bc))
    assert q.shape == (1, 3, 4, 19)
    assert abs(q[0, 0, 1, 4] - 0.3) < 1e-12 and abs(q[0, 0, 1, 7] - 0.6) < 1e-12
    assert (q != 0).sum() == 2

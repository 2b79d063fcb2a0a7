"""The oracle's D3Q19 colour-gradient box with walls on the y faces (the
reference's single-phase 3D faces applied per colour, S/Iterate3D.cu:576-598,
its mirrored face gradient, S/GpuCollision.cu:306-333, and the force shift of
gpuUpdateMacro3DCG, S/GpuUpdateMacro.cu:340-341), pinned by closed forms and
node-level identities.  CPU only.

* One fluid between a face at rest and one moving at U (the 3D inlet
  condition per colour): the steady profile is exactly linear, and every
  colour's populations on the faces carry exactly the face velocity.
* One fluid driven by a force g along x between faces at rest: the steady
  velocity is g/omega + g/(2 nu) y (H - y) exactly (the shift of the macro
  update on top of the parabola whose second difference is -g/nu).
* Two layers of viscosity ratio 4 follow analyticalCouette.
* On-node bounce-back faces: each unknown link equals its opposite.
* The face gradient: x- and z-varying fields give the interior's components
  on the faces, a y-varying one none.
"""
import numpy as np
import pytest

from thesis_amd import cg_cases as CC


def box(oracle, nx, ny, nz, **kw):
    return oracle.CgSim(nx, ny, nz=nz, nthreads=4, **kw)


@pytest.mark.parametrize("ho", [0, 1])
def test_single_fluid_couette_is_exactly_linear(oracle, ho):
    nx, ny, nz, U = 5, 17, 6, 0.01
    s = box(oracle, nx, ny, nz, A=0.0, walls=1, u_wall=U, high_order=ho)
    s.init(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx)))
    s.step(6000)
    d = s.get()
    y = np.arange(ny) / (ny - 1)
    assert np.abs(d["u"] - U * y[None, :, None]).max() < 1e-10 * U
    assert np.abs(d["v"]).max() < 1e-12 * U and np.abs(d["w"]).max() < 1e-12 * U


def test_each_colour_carries_the_face_velocity(oracle):
    """After a step from a random two-colour state the north face's red and
    blue populations each have momentum rho_k (U, 0, 0), the south face's zero."""
    nx, ny, nz, U = 6, 9, 5, 0.02
    rng = np.random.default_rng(3)
    red = rng.uniform(0.1, 0.9, (nz, ny, nx))
    s = box(oracle, nx, ny, nz, beta=0.7, A=1e-3, walls=1, u_wall=U)
    s.init(red, 1 - red)
    s.step(3)
    d = s.get()
    cx = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
    cy = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
    cz = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
    for key in ("r_f", "b_f"):
        f = d[key].reshape(19, nz, ny, nx)
        for j, Uf in ((ny - 1, U), (0, 0.0)):
            fj = f[:, :, j, :]
            rho = fj.sum(0)
            for c, want in ((cx, Uf), (cy, 0.0), (cz, 0.0)):
                mom = np.tensordot(c, fj, axes=(0, 0))
                assert np.abs(mom - want * rho).max() < 1e-15, (key, j)


def test_forced_channel_closed_form(oracle):
    nx, ny, nz, g = 4, 17, 4, 1e-6
    nu = 1 / 6
    om = 1 / (3 * nu + 0.5)
    s = box(oracle, nx, ny, nz, A=0.0, walls=1, u_wall=0.0, g=g, external_force=1, r_nu=nu, b_nu=nu)
    s.init(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx)))
    s.step(8000)
    u = s.get()["u"]
    y = np.arange(ny)
    want = g / om + g / (2 * nu) * y * (ny - 1 - y)
    assert np.abs(u - want[None, :, None]).max() < 1e-9 * want.max()


def test_two_layer_couette_follows_analytical_profile(oracle):
    nx, ny, nz, U = 4, 33, 4, 0.01
    nr, nb = 1 / 6, 1 / 24
    s = box(oracle, nx, ny, nz, beta=0.7, A=1e-3, r_nu=nr, b_nu=nb, walls=1, u_wall=U)
    red = np.broadcast_to(CC.layers(nx, ny), (nz, ny, nx)).copy()
    s.init(red, 1 - red)
    s.step(20000)
    col = s.get()["u"][nz // 2, :, nx // 2]
    y = np.arange(ny) / (ny - 1)
    err = np.abs(col - CC.analytical_couette(nr / nb, y, U)) / U
    assert err.max() < 0.035 and err[np.abs(y - 0.5) > 0.2].max() < 0.01
    # the profile does not depend on x or z
    u = s.get()["u"]
    assert np.abs(u - u[:1, :, :1]).max() < 1e-12


def test_bounce_back_faces_reflect_every_unknown_link(oracle):
    nx, ny, nz = 6, 10, 5
    rng = np.random.default_rng(5)
    red = rng.uniform(0.2, 0.8, (nz, ny, nx))
    s = box(oracle, nx, ny, nz, beta=0.7, A=1e-3, walls=2, g=-2e-5, external_force=0, r_alpha=0.6)
    s.init(red, 1 - red)
    s.step(4)
    d = s.get()
    cx = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    cy = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    cz = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    opp = [next(j for j in range(19) if (cx[j], cy[j], cz[j]) == (-cx[i], -cy[i], -cz[i])) for i in range(19)]
    for key in ("r_f", "b_f"):
        f = d[key].reshape(19, nz, ny, nx)
        for j, sign in ((ny - 1, -1), (0, 1)):
            for k in range(19):
                if cy[k] == sign:
                    assert np.array_equal(f[k, :, j, :], f[opp[k], :, j, :]), (key, j, k)


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_face_gradient_mirrors_the_plane_inside(oracle, axis):
    nx, ny, nz = 8, 9, 7
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    t = {"x": x / nx, "y": y / ny, "z": z / nz}[axis]
    red = 0.5 + 0.3 * np.sin(2 * np.pi * t)
    s = box(oracle, nx, ny, nz, walls=1)
    s.init(red, 1 - red)
    gx, gy, gz = s.gradient()
    faces = (slice(None), [0, ny - 1], slice(None))
    inner = (slice(None), [1, ny - 2], slice(None))
    if axis == "y":
        # the mirrored plane makes every y-varying field look even about the face
        assert np.abs(gx[faces]).max() == 0 and np.abs(gy[faces]).max() == 0 and np.abs(gz[faces]).max() == 0
        assert np.abs(gy[inner]).max() > 1e-3
    else:
        g = gx if axis == "x" else gz
        # a field constant along y has the same stencil sums on the face and inside
        assert np.abs(g[faces] - g[inner]).max() < 1e-15
        assert np.abs(g[faces]).max() > 1e-3 and np.abs(gy[faces]).max() == 0

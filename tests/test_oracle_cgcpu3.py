"""The reference's D3Q19 CPU colour-gradient path (S/Multiphase.cu: mp3DColl,
streamMP3D, peridicBoundaries3D, updateMacroMP3D; the #define CUDA 0 branch of
S/Iterate3D.cu, compiled out in the shipped file) restated in the oracle as a
second opinion, pinned against the GPU formulation and closed forms.  CPU only.

* One step away from the faces is the GPU formulation's step with three
  differences, each checked through the GPU formulation's own outputs: half the
  perturbation amplitude, the enhanced terms weighted by the mean alpha instead
  of the mean viscosity (so the CPU path's enhanced part is the GPU path's
  times mean_alpha / mean_nu), and blue recoloured with red's sign.  With K the
  GPU path's recolouring term (k_b r - k_r b of its collided colours) and F its
  summed populations, the CPU path's colours are k_r F + K and k_b F + K.
* The faces are never streamed or updated: their densities and velocities stay
  those of the start, and the periodic copies move only face populations, with
  the reference's slips (a blue link copied into the red array).
* The gradient on the faces: the normal component is zero; the front and back
  faces sample i + |c_x| (so an x-varying field has no x gradient there); the x
  and y faces on the front and back planes leave out the links beyond the box
  (a linear field gives 5/6 of its slope); a node on an x and a y face gets none.
* The consequence of blue's sign: blue gathers in a red bubble's core, where the
  GPU formulation keeps it out.
"""
import numpy as np
import pytest

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def random_state(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    red = 0.5 + 0.4 * np.tanh(3.0 - np.sqrt((x - nx / 2 + 0.4) ** 2 + (y - ny / 2 - 0.2) ** 2 + (z - nz / 2 + 0.3) ** 2))
    blue = 0.6 * (1 - red) + 0.05
    u = 0.02 * rng.standard_normal((3, nz, ny, nx))
    cu = np.tensordot(CX, u[0], 0) + np.tensordot(CY, u[1], 0) + np.tensordot(CZ, u[2], 0)
    shape = W[:, None, None, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(0))
    rf = red * shape * (1 + 0.01 * rng.standard_normal(shape.shape))
    bf = blue * shape * (1 + 0.01 * rng.standard_normal(shape.shape))
    return rf, bf


def from_source(a, d):
    """the value at x - c_d for every node x (periodic shift; used away from the faces)"""
    return np.roll(a, (CZ[d], CY[d], CX[d]), axis=(0, 1, 2))


def test_one_step_is_the_gpu_formulation_with_its_three_differences(oracle):
    nx, ny, nz, A = 10, 9, 8, 2e-3
    rf, bf = random_state(nx, ny, nz, 7)
    kw = dict(r_alpha=0.3, b_alpha=0.2, r_nu=1 / 6, b_nu=0.1, beta=0.7)
    cpu = oracle.CgCpuSim(nx, ny, nz=nz, A=A, **kw)
    g0 = oracle.CgSim(nx, ny, nz=nz, A=A / 2, enhanced=0, **kw)
    g1 = oracle.CgSim(nx, ny, nz=nz, A=A / 2, enhanced=1, **kw)
    for s in (cpu, g0, g1):
        s.init_f(rf, bf)
        s.step(1)
    a, b0, b1 = cpu.get(), g0.get(), g1.get()
    R, B = rf.sum(0), bf.sum(0)
    rho = R + B
    kr, kb = R / rho, B / rho
    ratio = (0.3 * R + 0.2 * B) / rho * (R / (rho / 6) + B / (rho * 0.1))   # mean_alpha / mean_nu
    want_r, want_b = np.zeros_like(rf), np.zeros_like(bf)
    for d in range(19):
        F0 = b0["r_f"][d] + b0["b_f"][d]
        F = F0 + from_source(ratio, d) * (b1["r_f"][d] + b1["b_f"][d] - F0)
        K = from_source(kb, d) * b0["r_f"][d] - from_source(kr, d) * b0["b_f"][d]
        want_r[d] = from_source(kr, d) * F + K
        want_b[d] = from_source(kb, d) * F + K
    inner = (slice(None), slice(2, nz - 2), slice(2, ny - 2), slice(2, nx - 2))
    scale = np.abs(want_r[inner]).max()
    assert np.abs(a["r_f"][inner] - want_r[inner]).max() < 1e-13 * scale
    assert np.abs(a["b_f"][inner] - want_b[inner]).max() < 1e-13 * scale
    # and the three differences are not small: the GPU path's own colours are off
    assert np.abs(a["b_f"][inner] - b1["b_f"][inner]).max() > 1e-2 * scale
    assert np.abs(a["r_f"][inner] - b1["r_f"][inner]).max() > 1e-3 * scale
    # the interior macros follow from the streamed colours
    i3 = inner[1:]
    assert np.allclose(a["r_rho"][i3], a["r_f"][inner].sum(0), rtol=1e-14, atol=0)
    jz = np.tensordot(CZ, a["r_f"] + a["b_f"], 1)
    assert np.allclose(a["w"][i3], (jz / a["rho"])[i3], rtol=1e-12, atol=1e-17)


def test_faces_keep_their_start_and_the_copies_keep_the_slips(oracle):
    nx, ny, nz = 7, 6, 5
    rf, bf = random_state(nx, ny, nz, 11)
    s = oracle.CgCpuSim(nx, ny, nz=nz, A=1e-3, r_alpha=0.3, b_alpha=0.3)
    s.init_f(rf, bf)
    a0 = s.get()
    s.step(12)
    a = s.get()
    face = np.ones((nz, ny, nx), bool)
    face[1:-1, 1:-1, 1:-1] = False
    for key in ("r_rho", "b_rho", "rho", "u", "v", "w"):
        assert np.array_equal(a[key][face], a0[key][face]), key
    assert np.abs(a["r_rho"][~face] - a0["r_rho"][~face]).max() > 1e-4
    # every face population is the start of the same link at its own node or at
    # one of its images across the faces it lies on (copies chain at edges)
    def across(c, n):
        return np.where((c == 0) | (c == n - 1), n - 1 - c, c)
    k, j, i = np.nonzero(face)
    images = [(kk, jj, ii) for kk in (k, across(k, nz)) for jj in (j, across(j, ny)) for ii in (i, across(i, nx))]
    for key, f0 in (("r_f", rf), ("b_f", bf)):
        for d in range(19):
            fd = a[key][d][k, j, i]
            assert np.all(np.any(np.stack([fd == f0[d][im] for im in images]), 0)), (key, d)
    # north face (below the front plane), link 18 (c = (0, -1, -1)): the red one
    # comes from the south face, the blue one stays (copied into the red array)
    assert np.array_equal(a["r_f"][18, :-1, -1, :], rf[18, :-1, 0, :])
    assert np.array_equal(a["b_f"][18, :-1, -1, :], bf[18, :-1, -1, :])
    assert not np.array_equal(bf[18, :-1, -1, :], bf[18, :-1, 0, :])
    # and link 4 (c = (0, -1, 0)) is copied in both colours
    assert np.array_equal(a["b_f"][4, :-1, -1, :], bf[4, :-1, 0, :])


@pytest.mark.parametrize("axis", ["x", "y"])
def test_face_gradient_branches(oracle, axis):
    nx, ny, nz, slope = 8, 9, 7, 0.04
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    red = 0.5 + 0.5 * slope * (x if axis == "x" else y)      # rho = 1, phase field linear with the slope
    s = oracle.CgCpuSim(nx, ny, nz=nz, A=1e-3)
    s.init(red, 1 - red)
    s.step(1)
    gx, gy, gz = s.collided()[2:]
    ii, jj = slice(1, nx - 1), slice(1, ny - 1)
    kk = slice(1, nz - 1)
    if axis == "x":
        # interior: the slope; north and south faces (y mirrored): the slope
        assert np.allclose(gx[kk, jj, ii], slope, rtol=0, atol=1e-15)
        assert np.allclose(gx[kk, [0, ny - 1], ii], slope, rtol=0, atol=1e-15)
        assert np.abs(gy[:, [0, ny - 1], :]).max() == 0
        # east and west faces: no x component
        assert np.abs(gx[:, jj, [0, nx - 1]]).max() == 0
        # front and back faces sample i + |c_x|: no x gradient survives
        assert np.abs(gx[[0, nz - 1], jj, ii]).max() < 1e-16
        # the links beyond the box left out on the y faces' front and back rows
        assert np.allclose(gx[[0, nz - 1], 0, ii], slope * 5 / 6, rtol=0, atol=1e-15)
    else:
        assert np.allclose(gy[kk, jj, ii], slope, rtol=0, atol=1e-15)
        # west and east faces, inside the box and on its front and back rows
        assert np.allclose(gy[kk, jj, 0], slope, rtol=0, atol=1e-15)
        assert np.allclose(gy[[0, nz - 1], jj, 0], slope * 5 / 6, rtol=0, atol=1e-15)
        assert np.allclose(gy[[0, nz - 1], jj, nx - 1], slope * 5 / 6, rtol=0, atol=1e-15)
        # front and back: the y component is the interior's, the z one zero
        assert np.allclose(gy[[0, nz - 1], jj, ii], slope, rtol=0, atol=1e-15)
        assert np.abs(gz[[0, nz - 1]][:, jj, ii]).max() == 0
    # a node on an x face and a y face has no gradient
    for g in (gx, gy, gz):
        assert np.abs(g[:, [0, 0, ny - 1, ny - 1], [0, nx - 1, 0, nx - 1]]).max() == 0


def test_blue_gathers_in_the_bubble_core(oracle):
    """createBubble3D's sharp bubble (S/Multiphase.cu:415-449): with blue
    recoloured like red, blue is drawn into the red core; the GPU formulation
    (with the CPU path's amplitude) keeps the core red."""
    n = 20
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    red = (np.sqrt((x - 9.5) ** 2 + (y - 9.5) ** 2 + (z - 9.5) ** 2) <= 5).astype(float)
    kw = dict(r_alpha=4 / 9, b_alpha=4 / 9, r_nu=1 / 6, b_nu=1 / 6, beta=0.7)
    cpu = oracle.CgCpuSim(n, n, nz=n, A=1e-3, **kw)
    gpu = oracle.CgSim(n, n, nz=n, A=5e-4, enhanced=1, nthreads=4, **kw)
    cpu.init(red, 1 - red)
    gpu.init(red, 1 - red)
    cpu.step(60)
    gpu.step(60)
    c = n // 2
    a, b = cpu.get(), gpu.get()
    # after 60 steps: about 2.0 against 0.011
    assert a["b_rho"][c, c, c] > 1.0 and b["b_rho"][c, c, c] < 0.05

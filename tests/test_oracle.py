"""Pins the FP64 oracle (oracle/lbm_oracle.c) before anything is compared
against it: closed forms, invariants, the reference's own compiled MRT tables
and the reference's lattice fixture.  CPU only."""
import numpy as np
import pytest

from thesis_amd import lattices as LT

pytestmark = []


def test_lattice_tables_match_reference_numbering(oracle):
    # S/GpuInit.cu:94-99 (D2Q9) and :822-842 (D3Q19)
    cx, cy, cz, opp, w = oracle.lattice(9)
    assert list(cx) == [0, 1, 0, -1, 0, 1, -1, -1, 1]
    assert list(cy) == [0, 0, 1, 0, -1, 1, 1, -1, -1]
    assert list(opp) == [0, 3, 4, 1, 2, 7, 8, 5, 6]
    assert np.allclose(w, [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    cx, cy, cz, opp, w = oracle.lattice(19)
    assert list(cx) == [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    assert list(cy) == [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    assert list(cz) == [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    pairs = {(1, 2), (3, 4), (5, 6), (7, 10), (8, 9), (11, 14), (12, 13), (15, 18), (16, 17)}
    for a, b in pairs:
        assert opp[a] == b and opp[b] == a
    # the product's Python-side tables agree
    for q in (9, 19):
        c = LT.velocities(q)
        cx, cy, cz, opp, w = oracle.lattice(q)
        assert (c[:, 0] == cx).all() and (c[:, 1] == cy).all() and (c[:, 2] == cz).all()
        assert (LT.opposite(q) == opp).all() and np.allclose(LT.weights(q), w)


@pytest.mark.parametrize("q", [9, 19])
def test_mrt_matrices(oracle, q):
    omega = 1 / 0.9
    M, Mi, s = oracle.mrt_matrices(q, omega)
    assert np.abs(M @ Mi - np.eye(q)).max() < 1e-14
    # the reference tabulates Minv with denominators 36 (2D) and 47880 (3D)
    den = 36 if q == 9 else 47880
    assert np.abs(Mi * den - np.round(Mi * den)).max() < 1e-9
    if q == 19:
        # spot values of S/CellFunctions.cu:73-76 (Minv row 1: 2520, -220, -760, 4788, -4788, ..., 2660, -2660)
        assert np.allclose(Mi[1, :5] * den, [2520, -220, -760, 4788, -4788])
        assert np.allclose(Mi[1, 9:11] * den, [2660, -2660])
        assert np.allclose(s, [0, 1.19, 1.4, 0, 1.2, 0, 1.2, 0, 1.2, omega, 1.4, omega, 1.4,
                               omega, omega, omega, 1.98, 1.98, 1.98])
    else:
        assert np.allclose(Mi[5] * den, [4, 2, 1, 6, 3, 6, 3, 0, 9])   # S/CellFunctions.cu:21
        assert np.allclose(s, [1, 1.4, 1.4, 1, 1.2, 1, 1.2, omega, omega])


@pytest.mark.parametrize("q", [9, 19])
def test_mrt_against_reference_build(oracle, q):
    """M and Minv*S from the reference's own MRTInitializer2D/3D (oracle/_ref)."""
    omega = 1 / 0.8
    ref = oracle.ref_mrt_tables(q, omega)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    M, Mi, s = oracle.mrt_matrices(q, omega)
    assert np.array_equal(ref[0], M)
    assert np.abs(ref[1] - Mi * s[None, :]).max() < 1e-15


@pytest.mark.parametrize("q", [9, 19])
def test_equilibrium_moments(oracle, q):
    c = LT.velocities(q).astype(float)
    rho, u = 1.07, np.array([0.03, -0.02, 0.01 if q == 19 else 0.0])
    feq = oracle.equilibrium(q, rho, u)
    d = 2 if q == 9 else 3
    assert abs(feq.sum() - rho) < 1e-15
    assert np.allclose(c.T @ feq, rho * u, atol=1e-16)
    P = np.einsum("i,ia,ib->ab", feq, c, c)[:d, :d]
    assert np.allclose(P, rho / 3 * np.eye(d) + rho * np.outer(u[:d], u[:d]), atol=1e-15)


def _single_node_sim(oracle, q, **kw):
    shape = (1, 1) if q == 9 else (1, 1, 1)
    return oracle.Sim(q, shape, np.ones(1, np.uint8), periodic=(1, 1, 1), **kw)


@pytest.mark.parametrize("q", [9, 19])
@pytest.mark.parametrize("coll", [0, 1, 2])
def test_collision_conserves_mass_and_momentum(oracle, q, coll):
    rng = np.random.default_rng(1)
    s = _single_node_sim(oracle, q, collision=coll, omega=1 / 0.7, trt_magic=0.25)
    f = oracle.equilibrium(q, 1.0, [0.02, 0.01, -0.01]) * (1 + 0.05 * rng.standard_normal(q))
    c = LT.velocities(q)
    rho = f.sum()
    u = (c.T @ f) / rho
    out = s.collide_node(f, rho, u)
    assert abs(out.sum() - rho) < 1e-14
    assert np.abs(c.T @ out - c.T @ f).max() < 1e-14


@pytest.mark.parametrize("q", [9, 19])
def test_trt_with_equal_rates_is_bgk(oracle, q):
    rng = np.random.default_rng(2)
    om = 1 / 0.8
    lam = (1 / om - 0.5) ** 2   # Lambda that gives omega_minus == omega_plus
    f = oracle.equilibrium(q, 1.0, [0.05, 0.0, 0.02]) * (1 + 0.1 * rng.random(q))
    b = _single_node_sim(oracle, q, collision=0, omega=om).collide_node(f, 1.02, [0.01, 0.02, 0.0])
    t = _single_node_sim(oracle, q, collision=1, omega=om, trt_magic=lam).collide_node(f, 1.02, [0.01, 0.02, 0.0])
    assert np.abs(b - t).max() < 1e-15


def test_mrt2d_textbook_with_uniform_rates_is_bgk(oracle):
    """D2Q9: with m_eq = M f_eq (textbook, no rho^2) and S = omega I, MRT == BGK."""
    rng = np.random.default_rng(3)
    om = 1 / 0.75
    f = oracle.equilibrium(9, 1.0, [0.05, -0.03]) * (1 + 0.1 * rng.random(9))
    b = _single_node_sim(oracle, 9, collision=0, omega=om).collide_node(f, 1.01, [0.03, -0.02])
    m = _single_node_sim(oracle, 9, collision=2, omega=om, mrt_s=[om] * 9).collide_node(f, 1.01, [0.03, -0.02])
    assert np.abs(b - m).max() < 1e-15
    # the reference's rho^2 quirk (S/GpuCollision.cu:987-988) changes it when rho != 1
    mq = _single_node_sim(oracle, 9, collision=2, omega=om, mrt_s=[om] * 9,
                          parity=oracle.PARITY_MRT_RHO2).collide_node(f, 1.01, [0.03, -0.02])
    assert np.abs(mq - m).max() > 1e-8


def test_mrt3d_zero_rates_is_identity(oracle):
    f = oracle.equilibrium(19, 1.0, [0.01, 0.02, 0.03]) * 1.03
    out = _single_node_sim(oracle, 19, collision=2, omega=1.0, mrt_s=[1e-300] * 19).collide_node(f, 0.9, [0.1, 0, 0])
    assert np.abs(out - f).max() < 1e-15


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_periodic_stream_round_trip_is_identity(oracle, coll):
    """Streaming is a permutation: with a collision that does nothing, L steps
    around an L-periodic box return every population bitwise."""
    n = 6
    rng = np.random.default_rng(4)
    s = oracle.Sim(19, (n, n, n), np.ones(n ** 3, np.uint8), collision=2, omega=1.0,
                   mrt_s=[1e-300] * 19, periodic=(1, 1, 1))
    s.init(0)
    f0 = 0.05 + rng.random((19, n, n, n))
    s.set_f(f0)
    s.step(n)
    assert np.array_equal(s.get_f(), f0)


def test_force_driven_poiseuille_trt_exact(oracle):
    """TRT with Lambda = 3/16, half-way bounce-back and Guo forcing gives the
    parabola u = g/(2 nu) y (H - y) with walls at y = -1/2, H - 1/2, to round-off."""
    nx, H = 3, 12
    solid = np.zeros((H, nx), bool)
    bb = LT.bb_from_solid(solid, 9, periodic=(True, False, True), outside_is_wall=(False, True, False))
    tau = 0.9
    nu, g = (tau - 0.5) / 3, 1e-5
    s = oracle.Sim(9, (H, nx), np.ones(H * nx, np.uint8), bb, collision=1, omega=1 / tau, trt_magic=3 / 16,
                   force_model=oracle.FORCE_GUO, g=(g, 0, 0), periodic=(1, 0, 0))
    s.init(0)
    s.step(8000)
    rho, u, v, w = s.macros()
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y)
    assert np.abs(u[0, :, 1] - ua).max() / ua.max() < 1e-9
    assert np.abs(v).max() < 1e-15


def test_bgk_poiseuille_has_known_slip(oracle):
    """BGK + half-way bounce-back: the profile is the parabola plus a constant
    slip u_s = g (8 (tau-1/2)^2 - ... ) -- only checked to be small and uniform."""
    nx, H = 3, 12
    bb = LT.bb_from_solid(np.zeros((H, nx), bool), 9, periodic=(True, False, True), outside_is_wall=(False, True, False))
    tau = 0.8
    nu, g = (tau - 0.5) / 3, 1e-5
    s = oracle.Sim(9, (H, nx), np.ones(H * nx, np.uint8), bb, collision=0, omega=1 / tau,
                   force_model=oracle.FORCE_GUO, g=(g, 0, 0), periodic=(1, 0, 0))
    s.init(0)
    s.step(8000)
    u = s.macros()[1][0, :, 1]
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y)
    slip = u - ua
    assert np.ptp(slip) < 1e-9 * ua.max() + 1e-3 * abs(slip).max()
    assert abs(slip).max() < 0.02 * ua.max()


def test_zouhe_prescribes_inlet_velocity_and_outlet_density(oracle):
    ch = LT.channel2d(24, 10, 0.05)
    s = oracle.Sim(9, (10, 24), ch["node_class"], ch["bb"], collision=1, omega=1 / 0.8, ubc=ch["ubc"],
                   vbc=ch["vbc"], rho_out=1.0)
    s.init(0)
    s.step(50)
    f = s.get_f()[:, 0]
    c = LT.velocities(9)
    rho = f.sum(0)
    ux = np.tensordot(c[:, 0], f, 1) / rho
    uy = np.tensordot(c[:, 1], f, 1) / rho
    assert np.abs(ux[:, 0] - ch["profile"]).max() < 1e-14
    assert np.abs(uy[:, 0]).max() < 1e-14
    assert np.abs(rho[:, -1] - 1.0).max() < 1e-14


def test_steady_channel_with_obstacles_has_equal_plane_fluxes(oracle):
    """Open channel with two obstacles run to a steady state: the lattice mass
    flux through every inner plane is the same (mass conservation of
    streaming, collision and bounce-back), and it stays within the density
    drop of the prescribed inlet flux sum(u_in) (the Zou-He inlet fixes u,
    not rho u)."""
    from lattice_checks import plane_fluxes
    ch = LT.channel2d(96, 20, 0.05, obstacles=[(30.0, 9.0, 3.0), (60.0, 12.0, 2.5)])
    s = oracle.Sim(9, (20, 96), ch["node_class"], ch["bb"], collision=1, omega=1 / 0.8, trt_magic=3 / 16,
                   ubc=ch["ubc"], vbc=ch["vbc"], rho_out=1.0)
    s.init(0)
    uprev = np.zeros((20, 96))
    for _ in range(200):
        s.step(100)
        u = s.macros()[1][0]
        res = np.sqrt(((u - uprev) ** 2).sum() / (u ** 2).sum())
        uprev = u
        if res < 1e-9:
            break
    assert res < 1e-9
    fx = plane_fluxes(s.get_f()[:, 0], ch["node_class"])
    assert (fx.max() - fx.min()) / fx.mean() < 1e-7, (fx.min(), fx.max())
    rho_in = s.macros()[0][0][:, 0].mean()
    assert abs(fx.mean() / ch["profile"].sum() - 1) < 2 * abs(rho_in - 1) + 1e-3
    # the same sum over the raw column, counting bounced populations as if they
    # had crossed, is not conserved: the check tells the two apart
    f = s.get_f()[:, 0]
    c = LT.velocities(9)
    col = np.tensordot(c[:, 0], f, 1)[:, 10:80].sum(0)
    assert (col.max() - col.min()) / col.mean() > 1e-4


def test_one_step_mass_balance_through_planes(oracle):
    """For any state, the flux through one plane minus the flux through a
    later plane equals the one-step mass gain of the columns between them
    (exact to rounding); counting bounced populations as crossing, or using
    the wrong step's density, breaks it at O(1)."""
    from lattice_checks import mass_balance_error
    ch = LT.channel2d(96, 20, 0.05, obstacles=[(30.0, 9.0, 3.0), (60.0, 12.0, 2.5)])
    s = oracle.Sim(9, (20, 96), ch["node_class"], ch["bb"], collision=1, omega=1 / 0.8, trt_magic=3 / 16,
                   ubc=ch["ubc"], vbc=ch["vbc"], rho_out=1.0)
    s.init(0)
    s.step(37)
    r0 = s.macros()[0][0]
    s.step(1)
    r1 = s.macros()[0][0]
    f1 = s.get_f()[:, 0]
    assert mass_balance_error(f1, r0, r1, ch["node_class"]) < 1e-12
    assert mass_balance_error(f1, r0, r0, ch["node_class"]) > 0.1


def test_mass_conserved_in_closed_porous_box(oracle):
    cls, bb, info = LT.porous3d(16, seed=7, rmin=2.0, rmax=3.5)
    for coll in (0, 1, 2):
        s = oracle.Sim(19, (16, 16, 16), cls, bb, collision=coll, omega=1 / 0.9, trt_magic=0.25,
                       force_model=oracle.FORCE_GUO, g=(1e-5, 0, 0), periodic=(1, 1, 1))
        s.init(0)
        m0 = s.mass()
        s.step(40)
        assert abs(s.mass() - m0) / m0 < 1e-13


def test_lattice_reader_matches_reference_fixture(tmp_path):
    """L/lidvalidation_{Node,BC}.dat: 30^3 nodes, 26 640 links (22 290 WALL,
    4 350 INLE); 900 links in each of dirs 1-6 and 1 770 in each of 7-18."""
    import gzip
    import os
    from thesis_amd import lbm as L
    here = os.path.join(os.path.dirname(__file__), "golden")
    paths = []
    for nm in ("lidvalidation_Node.dat", "lidvalidation_BC.dat"):
        p = tmp_path / nm
        p.write_bytes(gzip.decompress(open(os.path.join(here, nm + ".gz"), "rb").read()))
        paths.append(str(p))
    r = L.read_lattice(*paths)
    assert r["dims"] == (30, 30, 30)
    cls = r["node_class"]
    assert (cls != L.SOLID).sum() == 27000
    # the lid: INLE links leave the top face; its 28 x 28 inner nodes are a
    # velocity face, its rim (wall and inlet links: corners) is all wall
    assert (cls == L.VEL_T).sum() == 28 * 28 and (cls[-1, 1:-1, 1:-1] == L.VEL_T).all()
    assert (cls[:-1] == L.FLUID).all() and (cls[-1, 0] == L.FLUID).all()
    rim = r["bb_links"][-1, 0, 5]
    assert bin(int(rim)).count("1") == 9           # an edge of the box: 9 wall links
    assert r["nlinks"] == 26640
    assert list(r["type_counts"][:3]) == [0, 22290, 4350]
    assert list(r["dir_counts"][1:7]) == [900] * 6 and list(r["dir_counts"][7:]) == [1770] * 12
    # every WALL link of the fixture points out of the box (on-node walls on the faces)
    c = LT.velocities(19)
    bb = r["bb_links"]
    for d in range(1, 19):
        k, j, i = np.nonzero((bb >> d) & 1)
        nxt = np.stack([i + c[d, 0], j + c[d, 1], k + c[d, 2]])
        assert ((nxt < 0) | (nxt > 29)).any(axis=0).all()


def test_cg_colour_masses_conserved_and_laplace_law(oracle):
    """Per-colour mass is conserved exactly; the bulk pressure jump of a
    static droplet approaches sigma/R with sigma = 4A/(9 omega) (S/Iterate.cu:163-168)."""
    n, R, A = 48, 12, 0.01125
    r, b = LT.droplet2d(n, n, R)
    s = oracle.CgSim(n, n, beta=0.7, A=A, r_nu=1 / 6, b_nu=1 / 6)
    s.init(r, b)
    d0 = s.get()
    s.step(1500)
    d = s.get()
    assert abs(d["r_rho"].sum() - d0["r_rho"].sum()) < 1e-10
    assert abs(d["b_rho"].sum() - d0["b_rho"].sum()) < 1e-10
    chi = (d["r_rho"] - d["b_rho"]) / d["rho"]
    dp = (d["rho"][chi > 0.99].mean() - d["rho"][chi < -0.99].mean()) / 3.0
    sigma = 4 * A / 9.0
    assert abs(dp * R - sigma) / sigma < 0.1
    # the reference's own estimator (S/GpuUpdateMacro.cu:124-133) stays available
    pin, nin, pout, nout, umax = s.laplace()
    assert nin > 0 and nout > 0 and umax < 1e-3


@pytest.mark.parametrize("ho", [0, 1])
def test_cg_gradient_exact_for_linear_field(oracle, ho):
    """sum_i w_i c_i c_i = I for both stencils, so the gradient of a linear phase
    field is exact away from the periodic seam."""
    n = 32
    s = oracle.CgSim(n, n, high_order=ho)
    x = np.arange(n)[None, :] * np.ones((n, 1))
    y = np.arange(n)[:, None] * np.ones((1, n))
    phi = 0.01 * (x - 16) + 0.02 * (y - 16)     # (r - b)/rho with rho = 1
    r = 0.5 * (1 + phi)
    s.init(r, 1 - r)
    gx, gy = s.gradient()
    sl = slice(3, n - 3)
    assert np.abs(gx[sl, sl] - 0.01).max() < 1e-15
    assert np.abs(gy[sl, sl] - 0.02).max() < 1e-15


def test_reference_l2_residual_is_the_streaming_difference(oracle):
    """orc_residual_l2 = sqrt(sum (f - fColl)^2) of the last step
    (S/ComputeResiduals.cu:13-27), restated here from the node collision and
    periodic streaming: fColl from collide_node on the state before the step,
    f from get_f after it."""
    ny, nx = 5, 6
    cls = np.ones((ny, nx), np.uint8)
    rng = np.random.default_rng(4)
    u = 0.05 * rng.standard_normal((ny, nx))
    v = 0.05 * rng.standard_normal((ny, nx))
    rho = 1 + 0.02 * rng.standard_normal((ny, nx))
    s = oracle.Sim(9, (ny, nx), cls, None, collision=1, omega=1 / 0.8, periodic=(1, 1, 0))
    s.init(0, rho=rho, u=u, v=v)
    s.step(2)
    f0 = s.get_f()[:, 0]
    r0, u0, v0, _ = (a[0] for a in s.macros())
    fc = np.zeros_like(f0)
    for y in range(ny):
        for x in range(nx):
            fc[:, y, x] = s.collide_node(f0[:, y, x], r0[y, x], (u0[y, x], v0[y, x]))
    s.step(1)
    f1 = s.get_f()[:, 0]
    c = LT.velocities(9)
    for i in range(9):                      # streaming: f1_i(x) = fc_i(x - c_i)
        assert np.allclose(f1[i], np.roll(np.roll(fc[i], c[i, 0], axis=1), c[i, 1], axis=0), atol=1e-15)
    want = np.sqrt(((f1 - fc) ** 2).sum())
    assert abs(s.residual_l2() - want) < 1e-14 * want


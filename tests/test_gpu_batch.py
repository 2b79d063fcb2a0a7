"""Batched channels: many D2Q9 lattices in one persistent launch.

Every instance is checked on its own against the FP64 oracle after a fixed
number of steps (<= 1e-12 relative, as the single-lattice FP64 path), and the
run to a 1e-6 residual is checked against single-lattice runs of the same
instances (same step count, same fields) and against steady-state mass
conservation.  Each case prints its name first."""
import numpy as np
import pytest

from lattice_checks import mass_balance_error, plane_fluxes
from thesis_amd import lattices as LT
from thesis_amd import lbm as L

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
ABS_FLOOR = 1e-15


def small_sweep(count=8, seed=11):
    return LT.channel_sweep(count=count, seed=seed, nx_range=(64, 200), ny_range=(16, 48), re_range=(10, 60),
                            r_range=(2.0, 6.0))


def oracle_of(oracle, d, collision):
    return oracle.Sim(9, d["node_class"].shape, d["node_class"], d["bb"], collision=collision, omega=d["omega"],
                      trt_magic=d["trt_magic"] if collision == L.TRT else 0.0, ubc=d["ubc"], vbc=d["vbc"],
                      rho_out=d["rho_out"])


def check_instance(b, i, o, d, tol=TOL64):
    fluid = d["node_class"] != L.SOLID
    gr, gu, gv = b.macros(i)
    orr, ou, ov, _ = (a[0] for a in o.macros())
    dr = np.abs(gr - orr)[fluid].max()
    du = np.sqrt((gu - ou) ** 2 + (gv - ov) ** 2)[fluid].max()
    speed = np.sqrt(ou ** 2 + ov ** 2)[fluid].max()
    gf = b.get_f(i)[:, fluid]
    of = o.get_f()[:, 0][:, fluid]
    df = np.abs(gf - of).max()
    errs = dict(i=i, rho=dr, u=du / speed, f=df / np.abs(of).max())
    assert dr <= max(tol * np.abs(orr[fluid]).max(), ABS_FLOOR), errs
    assert du <= max(tol * speed, ABS_FLOOR), errs
    assert df <= max(tol * np.abs(of).max(), ABS_FLOOR), errs
    return errs


@pytest.mark.parametrize("collision", [L.TRT, L.BGK])
def test_batch_small_f64_matches_oracle(oracle, collision):
    sweep = small_sweep()
    print(f"case batch small f64 collision={collision} shapes={[d['node_class'].shape for d in sweep]}", flush=True)
    b = L.Batch(sweep, collision=collision, storage="f64")
    steps, res, _ = b.run(0.0, every=50, max_steps=300)
    assert (steps == 300).all()
    for i, d in enumerate(sweep):
        o = oracle_of(oracle, d, collision)
        o.init(0)
        o.step(300)
        check_instance(b, i, o, d)
    b.close()


def test_batch_continue_and_reinit_one_instance(oracle):
    """Two runs in a row continue each instance from its own newest buffer
    (instances retire on different parities), and re-initialising one
    instance leaves the others untouched."""
    sweep = small_sweep(count=5, seed=3)
    print("case batch continue/reinit", flush=True)
    probe = L.Batch(sweep, collision=L.TRT, storage="f64")
    _, r203, _ = probe.run(0.0, every=7, max_steps=203)
    probe.close()
    b = L.Batch(sweep, collision=L.TRT, storage="f64")
    # the median residual at step 203 as tolerance retires instances at
    # different check steps (multiples of 7, so on both buffer parities)
    steps1, _, _ = b.run(float(np.median(r203)), every=7, max_steps=400)
    print("retired at", steps1.tolist(), flush=True)
    steps2, _, _ = b.run(0.0, every=13, max_steps=91)
    assert len(set(steps1.tolist())) > 1
    for i, d in enumerate(sweep):
        o = oracle_of(oracle, d, L.TRT)
        o.init(0)
        o.step(int(steps1[i] + steps2[i]))
        check_instance(b, i, o, d)
    ny, nx = sweep[2]["node_class"].shape
    u0 = np.where(sweep[2]["node_class"] != L.SOLID, 0.01, 0.0)
    b.init(2, u=u0)
    b.run(0.0, every=10, max_steps=40)
    for i, d in enumerate(sweep):
        o = oracle_of(oracle, d, L.TRT)
        if i == 2:
            o.init(0, u=u0)
            o.step(40)
        else:
            o.init(0)
            o.step(int(steps1[i] + steps2[i] + 40))
        check_instance(b, i, o, d)
    b.close()


def test_batch_full_sweep_64_first_steps_match_oracle(oracle):
    """The full 64-channel sweep (128x32 .. 1024x256, 1-5 obstacles, Re 10-100),
    FP64, every instance against the oracle after 60 steps."""
    sweep = LT.channel_sweep()
    print(f"case batch full sweep 64, nodes={sum(d['nx'] * d['ny'] for d in sweep)}", flush=True)
    b = L.Batch(sweep, collision=L.TRT, storage="f64")
    steps, _, _ = b.run(0.0, every=20, max_steps=60)
    assert (steps == 60).all()
    worst = 0.0
    for i, d in enumerate(sweep):
        o = oracle_of(oracle, d, L.TRT)
        o.init(0)
        o.step(60)
        worst = max(worst, check_instance(b, i, o, d)["u"])
        del o
    print("worst relative u error", worst, flush=True)
    b.close()


def test_batch_full_sweep_64_to_residual_f64():
    """The sweep run to a 1e-6 residual in one launch: every instance
    converges, carries the same lattice mass flux through every inner plane,
    and selected instances match single-lattice runs to the same residual
    (same step count, same fields)."""
    sweep = LT.channel_sweep()
    print("case batch full sweep 64 to 1e-6 f64", flush=True)
    b = L.Batch(sweep, collision=L.TRT, storage="f64")
    steps, res, ms = b.run(1e-6, every=100, max_steps=400000)
    tot, fl = b.nodes()
    print(f"steps min {steps.min()} max {steps.max()} sum(node*steps) {sum(steps[i] * sweep[i]['nx'] * sweep[i]['ny'] for i in range(64)):.3e} ms {ms:.1f}", flush=True)
    assert (res < 1e-6).all(), res
    assert (steps < 400000).all()
    worst = 0.0
    rho0 = [b.macros(i)[0] for i in range(len(sweep))]
    for i, d in enumerate(sweep):
        fx = plane_fluxes(b.get_f(i), d["node_class"])
        spread = (fx.max() - fx.min()) / fx.mean()
        worst = max(worst, spread)
        assert spread <= 3e-4, (i, fx.min(), fx.max())
    print("worst plane-flux spread at the 1e-6 residual", worst, flush=True)
    # one more step: the exact discrete mass balance between every pair of planes
    b.run(0.0, every=1, max_steps=1)
    worst = max(mass_balance_error(b.get_f(i), rho0[i], b.macros(i)[0], d["node_class"])
                for i, d in enumerate(sweep))
    print("worst one-step mass-balance error", worst, flush=True)
    assert worst < 1e-10
    steps = steps + 1
    order = np.argsort([d["nx"] * d["ny"] for d in sweep])
    for i in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        d = sweep[i]
        print(f"single-lattice check of instance {i} {d['node_class'].shape}", flush=True)
        s = L.Solver(9, d["node_class"].shape, d["node_class"], d["bb"], collision=L.TRT, omega=d["omega"],
                     trt_magic=d["trt_magic"], ubc=d["ubc"], vbc=d["vbc"], rho_out=d["rho_out"], storage="f64")
        s.init(0)
        n1, r1 = s.run_to_residual(1e-6, every=100, max_steps=400000)
        assert n1 + 1 == steps[i], (i, n1, steps[i], r1, res[i])
        s.step(1)
        gr, gu, gv = b.macros(i)
        sr, su, sv, _ = (a.reshape(gr.shape) for a in s.macros())
        speed = np.sqrt(su ** 2 + sv ** 2).max()
        assert np.sqrt((gu - su) ** 2 + (gv - sv) ** 2).max() <= 1e-10 * speed
        assert np.abs(gr - sr).max() <= 1e-10
        s.close()
    b.close()


@pytest.mark.skip(reason="FP32 batch storage is switched off: its first cooperative launch did not finish "
                         "(the FP64 launches do); the cause is not found yet")
def test_batch_full_sweep_f32_against_f64():
    """FP32 storage converges on the same sweep; its steady velocity agrees
    with the FP64 steady state to 1e-4 of the peak speed."""
    sweep = LT.channel_sweep()
    print("case batch full sweep 64 f32 vs f64", flush=True)
    b64 = L.Batch(sweep, collision=L.TRT, storage="f64")
    s64, r64, _ = b64.run(1e-6, every=100, max_steps=400000)
    b32 = L.Batch(sweep, collision=L.TRT, storage="f32")
    s32, r32, _ = b32.run(1e-6, every=100, max_steps=400000)
    print("f32 residuals max", r32.max(), "steps", s32.min(), s32.max(), flush=True)
    worst = 0.0
    for i in range(len(sweep)):
        _, u6, v6 = b64.macros(i)
        _, u3, v3 = b32.macros(i)
        speed = np.sqrt(u6 ** 2 + v6 ** 2).max()
        worst = max(worst, np.sqrt((u6 - u3) ** 2 + (v6 - v3) ** 2).max() / speed)
    print("worst f32 relative velocity error", worst, flush=True)
    assert worst < 1e-4
    b64.close()
    b32.close()

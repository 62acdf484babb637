"""GPU parity of the integrator drivers against the oracle.

Criteria (north star): over the first 100 steps, positions agree within
1e-9 AU; the accepted step sizes agree within 1e-12 relative (the same
accept/reject decisions); the final BodyData arrays follow the reference's
pointer swap.
"""
import numpy as np
import pytest

from oracle.cpu_ref import Oracle
from solaris_amd import _abi, runtime, scenarios as S

pytestmark = pytest.mark.gpu

POS_TOL_AU = 1e-9


def run_pair(scn, integrator, eps, steps, h0):
    ctx = runtime.Context(scn)
    orc = Oracle(scn)
    ctx.set_state(scn.y0)
    orc.set_state(scn.y0)
    tg = _abi.timeline(0.0, h0, 1e12)
    to = _abi.timeline(0.0, h0, 1e12)
    hs = []
    for k in range(steps):
        rg, bhg = ctx.driver(integrator, eps, tg)
        ro, bho = orc.driver(integrator, eps, to)
        assert rg == ro == 0, (k, rg, ro, ctx.last_error())
        hs.append((tg.hDid, to.hDid, bhg, bho))
    h = np.array(hs)
    np.testing.assert_allclose(h[:, 0], h[:, 1], rtol=1e-12)
    np.testing.assert_allclose(h[:, 2], h[:, 3], rtol=1e-12)
    assert abs(tg.time - to.time) <= 1e-12 * abs(to.time)
    yg, yo = ctx.get_array(runtime.Y0), orc.get_array(runtime.Y0)
    dpos = np.abs(yg[:, :3] - yo[:, :3]).max()
    assert dpos <= POS_TOL_AU, dpos
    # previous state lands in Y (pointer swap), accel holds f0 of the last step
    np.testing.assert_allclose(ctx.get_array(runtime.Y)[:, :3], orc.get_array(runtime.Y)[:, :3],
                               rtol=0, atol=POS_TOL_AU)
    return ctx, orc, dpos


def test_rkf78_c1_100_steps():
    scn = S.make_c1()
    print("[rkf78 c1]", flush=True)
    ctx, orc, d = run_pair(scn, S.RUNGE_KUTTA_FEHLBERG78, 1e-10, 100, scn.h_next)
    # side outputs of the last evaluation (stage 12 of the accepted attempt)
    a, b = ctx.side_outputs(), orc.side_outputs()
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12)
    assert np.array_equal(a[1], b[1])


def test_rkf78_c2_drag_migration_100_steps():
    scn = S.make_c2(n_pl=1024)
    print("[rkf78 c2]", flush=True)
    ctx, orc, d = run_pair(scn, S.RUNGE_KUTTA_FEHLBERG78, 1e-10, 100, scn.h_next)
    assert np.array_equal(ctx.side_outputs()[3], orc.side_outputs()[3])


def test_rk4_c5_100_steps():
    scn = S.make_c5(n_tp=256)
    print("[rk4 c5]", flush=True)
    run_pair(scn, S.RUNGE_KUTTA4, 1e-10, 100, 1.0)


def test_dormand_prince_c2_50_steps():
    scn = S.make_c2(n_pl=512)
    print("[dp c2]", flush=True)
    run_pair(scn, S.DORMAND_PRINCE, 1e-10, 50, 5.0)


def test_step_and_error_max_match():
    scn = S.make_c1(n_tp=64)
    ctx, orc = runtime.Context(scn), Oracle(scn)
    ctx.set_state(scn.y0); orc.set_state(scn.y0)
    f0 = orc.compute(0.0, scn.y0)
    ctx.set_accel(f0); orc.set_accel(f0)
    for integ in (S.RUNGE_KUTTA_FEHLBERG78, S.RUNGE_KUTTA4, S.DORMAND_PRINCE):
        ctx.step(integ, 0.0, 3.0); orc.step(integ, 0.0, 3.0)
        np.testing.assert_allclose(ctx.get_array(runtime.Y), orc.get_array(runtime.Y), rtol=1e-13, atol=1e-15)
        eg, eo = ctx.get_array(runtime.ERROR), orc.get_array(runtime.ERROR)
        np.testing.assert_allclose(eg, eo, rtol=1e-6, atol=1e-22)


def test_integrate_clips_to_length():
    scn = S.make_c1(n_tp=32)
    ctx, orc = runtime.Context(scn), Oracle(scn)
    ctx.set_state(scn.y0); orc.set_state(scn.y0)
    tg = _abi.timeline(0.0, scn.h_next, 400.0, output=150.0)
    to = _abi.timeline(0.0, scn.h_next, 400.0, output=150.0)
    rg, sg = ctx.integrate(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, tg)
    ro, so = orc.integrate(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, to)
    assert rg == ro == 0
    assert sg.accepted_steps == so.accepted_steps
    assert tg.time == pytest.approx(400.0, abs=1e-9) and to.time == pytest.approx(400.0, abs=1e-9)
    np.testing.assert_allclose(ctx.get_array(runtime.Y0)[:, :3], orc.get_array(runtime.Y0)[:, :3],
                               rtol=0, atol=POS_TOL_AU)
    np.testing.assert_allclose(ctx.integrals(), orc.integrals(), rtol=1e-12, atol=1e-20)

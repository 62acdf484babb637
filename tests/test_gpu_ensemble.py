"""Batched ensemble (sol_ensemble_*) against the oracle, system by system.

Each system of the ensemble is an ordinary reference run (RKF78 1e-10
with the caller loop's clipping to the end of its time line).  Free-running
runs choose their own steps, which differ from the oracle's at the rounding
level (see test_gpu_drivers.py), and many C4 systems are chaotic: a 1e-15
relative change of the initial positions of system 0 moves its bodies by
3.5e-3 AU after three years.  Hence two criteria:

* 30 days, 24 systems: status DONE at the end time, positions within
  1e-8 AU of the oracle's run (a 1e-15 perturbation moves them by at most
  9e-11 AU over that span);
* one year, 16 systems with n = 23..183: status DONE, accepted steps within
  2% + 1 of the oracle's, and barycentric energy and angular-momentum drift
  no larger than three times the oracle's (plus 1e-11).
"""
import numpy as np
import pytest

from oracle.cpu_ref import Oracle
from solaris_amd import _abi, runtime, scenarios as S

from invariants import drifts

pytestmark = pytest.mark.gpu

# the 16 systems of the one-year check (the cheapest for the oracle)
LONG = [23, 9, 18, 13, 21, 11, 20, 3, 12, 16, 1, 0, 14, 6, 7, 19]


def _run_ensemble(scns, T):
    ens = runtime.Ensemble(scns)
    for k, s in enumerate(scns):
        ens.set_system(k, s, _abi.timeline(0.0, s.h_next, T))
    st = ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10)
    print(f"  {len(scns)} systems to {T} d: {st.accepted_steps} steps, {st.interactions:.3e} interactions, "
          f"{st.device_ms:.1f} ms", flush=True)
    return ens


def _run_oracle(s, T):
    o = Oracle(s)
    o.set_state(s.y0)
    to = _abi.timeline(0.0, s.h_next, T)
    r, so = o.integrate(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, to)
    assert r == 0
    return o.get_array(0), so


def test_ensemble_matches_oracle_short_span():
    scns = [S.make_c4_system(k) for k in range(24)]
    T = 30.0
    ens = _run_ensemble(scns, T)
    worst = 0.0
    for k, s in enumerate(scns):
        y, tl, status, steps = ens.get_system(k)
        assert status == 1 and tl.time == T, (k, status, tl.time)
        yo, so = _run_oracle(s, T)
        assert abs(steps - so.accepted_steps) <= 0.02 * so.accepted_steps + 1, (k, steps, so.accepted_steps)
        d = np.abs(y[:, :3] - yo[:, :3]).max()
        worst = max(worst, d)
        assert d <= 1e-8, (k, s.n, d)
    print(f"  worst position difference at {T} d: {worst:.2e} AU", flush=True)


def test_ensemble_one_year_steps_and_drift():
    scns = [S.make_c4_system(k) for k in LONG]
    T = 365.25
    ens = _run_ensemble(scns, T)
    for k, s in enumerate(scns):
        y, tl, status, steps = ens.get_system(k)
        assert status == 1 and tl.time == T, (LONG[k], status, tl.time)
        yo, so = _run_oracle(s, T)
        assert abs(steps - so.accepted_steps) <= 0.02 * so.accepted_steps + 1, (LONG[k], steps, so.accepted_steps)
        dEg, dLg = drifts(s.y0, y, s.mass)
        dEo, dLo = drifts(s.y0, yo, s.mass)
        print(f"  system {LONG[k]} (n={s.n}): steps {steps}/{so.accepted_steps}, energy drift {dEg:.2e}/{dEo:.2e}, "
              f"AM drift {dLg:.2e}/{dLo:.2e}", flush=True)
        assert dEg <= 3 * dEo + 1e-11 and dLg <= 3 * dLo + 1e-11, (LONG[k], dEg, dEo, dLg, dLo)


def test_ensemble_step_budget_and_restart():
    scns = [S.make_c4_system(k) for k in range(100, 140)]
    ens = runtime.Ensemble(scns)
    for k, s in enumerate(scns):
        ens.set_system(k, s, _abi.timeline(0.0, s.h_next, 1e9))
    ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, max_steps=7)
    for k in range(len(scns)):
        _, tl, status, steps = ens.get_system(k)
        assert status == 0 and steps == 7
    ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, max_steps=5)
    assert all(ens.get_system(k)[3] == 12 for k in range(len(scns)))


def test_ensemble_isolates_a_failing_system():
    """A system that fails (RKF78 step-size underflow, t + h == t) stops
    with its own status and the others finish.  The failure is forced from
    the host: two protoplanets 1e-9 AU apart at t = 1e7 d, where the step
    the pair needs is below the spacing of doubles at t (the oracle fails
    the same way at its first step)."""
    scns = [S.make_c4_system(k) for k in range(8)]
    bad = scns[3]
    bad.y0[2, :3] = bad.y0[1, :3] + 1e-9
    bad.y0[2, 3:] = bad.y0[1, 3:]
    o = Oracle(bad)
    o.set_state(bad.y0)
    to = _abi.timeline(1.0e7, bad.h_next, 200.0)
    assert o.integrate(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, to, max_steps=2000)[0] == 4
    T = 30.0
    ens = runtime.Ensemble(scns)
    for k, s in enumerate(scns):
        tl = _abi.timeline(1.0e7, s.h_next, 200.0) if k == 3 else _abi.timeline(0.0, s.h_next, T)
        ens.set_system(k, s, tl)
    ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, max_steps=2000)
    statuses = [ens.get_system(k)[2] for k in range(len(scns))]
    assert statuses[3] == 2, statuses
    for k, s in enumerate(scns):
        if k == 3:
            continue
        y, tl, status, _ = ens.get_system(k)
        assert status == 1 and tl.time == T, (k, status, tl.time)
        yo, _ = _run_oracle(s, T)
        assert np.abs(y[:, :3] - yo[:, :3]).max() <= 1e-8, k

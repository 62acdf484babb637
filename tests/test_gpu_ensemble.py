"""Batched ensemble (sol_ensemble_*) against the oracle, system by system.

Each system of the ensemble is an ordinary reference run (RKF78 1e-10
with the caller loop's clipping to the end of its time line).  Criteria:
every system reaches its end time with status DONE, the same number of
accepted steps as the oracle's run to within 2% + 1, and positions at the
common end time within 1e-8 AU (free-running runs: the step sequences
differ at the rounding level, see test_gpu_drivers.py).
"""
import numpy as np
import pytest

from oracle.cpu_ref import Oracle
from solaris_amd import _abi, runtime, scenarios as S

pytestmark = pytest.mark.gpu


def test_ensemble_matches_oracle_per_system():
    scns = [S.make_c4_system(k) for k in range(24)]
    T = 3.0 * 365.25
    ens = runtime.Ensemble(scns)
    for k, s in enumerate(scns):
        ens.set_system(k, s, _abi.timeline(0.0, s.h_next, T))
    st = ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10)
    print(f"  {len(scns)} systems, {st.accepted_steps} steps, {st.interactions:.3e} interactions, "
          f"{st.device_ms:.1f} ms", flush=True)
    worst = 0.0
    for k, s in enumerate(scns):
        y, tl, status, steps = ens.get_system(k)
        assert status == 1 and tl.time == T, (k, status, tl.time)
        o = Oracle(s)
        o.set_state(s.y0)
        to = _abi.timeline(0.0, s.h_next, T)
        r, so = o.integrate(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, to)
        assert r == 0
        assert abs(steps - so.accepted_steps) <= 0.02 * so.accepted_steps + 1, (k, steps, so.accepted_steps)
        d = np.abs(y[:, :3] - o.get_array(0)[:, :3]).max()
        worst = max(worst, d)
        assert d <= 1e-8, (k, s.n, d)
    print(f"  worst position difference at T: {worst:.2e} AU", flush=True)


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
    """A system whose tolerance cannot be met (a near collision that
    drives h to underflow) stops with its own status; the others finish.
    The failure is forced from the host: an extremely close pair."""
    scns = [S.make_c4_system(k) for k in range(8)]
    bad = scns[3]
    bad.y0[2, :3] = bad.y0[1, :3] + 1e-13      # two protoplanets almost on top of each other
    bad.y0[2, 3:] = bad.y0[1, 3:]
    ens = runtime.Ensemble(scns)
    for k, s in enumerate(scns):
        ens.set_system(k, s, _abi.timeline(0.0, s.h_next, 200.0))
    ens.run(S.RUNGE_KUTTA_FEHLBERG78, 1e-10, max_steps=2000)
    statuses = [ens.get_system(k)[2] for k in range(len(scns))]
    assert statuses[3] in (2, 0), statuses       # underflow (or still creeping)
    assert all(st == 1 for k, st in enumerate(statuses) if k != 3), statuses

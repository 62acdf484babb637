"""Pins the oracle's forward-mode (dual-number) RK4 derivative,
cr_rk4_tangent, before it is trusted as the reference gradient of the
device adjoint.

* Its value part is the oracle's own RK4 (Compute + Step, which the golden
  vectors pin to the compiled reference).
* Its derivative part agrees with central finite differences of that
  independent code path, along random directions in the initial state and
  in the masses, and along single masses (the central body's included):
  a dropped Jacobian term (the indirect term, a Kepler mu term, a
  pair-mass term) fails one of these.
"""
import numpy as np
import pytest

from oracle.cpu_ref import Oracle

from adjoint_common import central_difference, oracle_rk4, small_c5, with_masses

H = 1.0
STEPS = 120


@pytest.fixture(scope="module")
def scn():
    return small_c5()


def test_tangent_value_is_oracle_rk4(scn):
    o = Oracle(scn)
    yT, _ = o.rk4_tangent(scn.y0, H, STEPS)
    ref = oracle_rk4(scn, scn.y0, H, STEPS)
    assert np.max(np.abs(yT - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_tangent_value_with_step_sequence(scn):
    hs = 0.5 + (np.arange(40) % 3) * 0.25
    o = Oracle(scn)
    yT, _ = o.rk4_tangent(scn.y0, 0.0, len(hs), h_seq=hs)
    ref = oracle_rk4(scn, scn.y0, hs, len(hs))
    assert np.max(np.abs(yT - ref)) <= 1e-13 * np.max(np.abs(ref))


def _check_direction(scn, dy0, dm, delta, rtol):
    o = Oracle(scn)
    _, dyT = o.rk4_tangent(scn.y0, H, STEPS, dy0=dy0, dm=dm)
    w = np.random.default_rng(7).normal(size=dyT.shape)
    w[0] = 0.0

    def f(e):
        y0 = scn.y0 + (e * dy0 if dy0 is not None else 0.0)
        s = with_masses(scn, scn.mass + e * dm) if dm is not None else scn
        return float(np.sum(w * oracle_rk4(s, y0, H, STEPS)))

    fd = central_difference(f, delta)
    ad = float(np.sum(w * dyT))
    assert abs(ad - fd) <= rtol * abs(fd), (ad, fd)


def test_tangent_state_direction_matches_fd(scn):
    rng = np.random.default_rng(11)
    dy0 = rng.normal(size=scn.y0.shape)
    dy0[0] = 0.0
    dy0[:, 3:] *= 0.01           # velocities in AU/day
    _check_direction(scn, dy0, None, 1e-5, 1e-7)


def test_tangent_mass_direction_matches_fd(scn):
    rng = np.random.default_rng(12)
    dm = np.zeros(scn.n)
    npl = scn.counts["proto_planet"]
    dm[1:1 + npl] = rng.uniform(0.5, 1.5, npl) * scn.mass[1:1 + npl]
    _check_direction(scn, None, dm, 0.3, 1e-7)


# The loss moves little with a small planet's mass, so the difference step
# is a large fraction of it (the map is close to linear in the masses).
@pytest.mark.parametrize("body", [0, 1, 2])
def test_tangent_single_mass_matches_fd(scn, body):
    dm = np.zeros(scn.n)
    dm[body] = scn.mass[body]
    _check_direction(scn, None, dm, 1e-4 if body == 0 else 0.3, 1e-7)

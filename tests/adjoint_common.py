"""Shared pieces of the adjoint tests: the loss, oracle RK4 stepping and
finite differences."""
import dataclasses

import numpy as np

from oracle.cpu_ref import Oracle
from solaris_amd import scenarios as S


def loss_and_grad(yT, targets):
    """L = sum_{i>=1} |r_i(T) - r_i*|^2 and dL/dy(T) (positions only)."""
    d = yT[:, :3] - targets
    d[0] = 0.0
    g = np.zeros_like(yT)
    g[:, :3] = 2.0 * d
    return float(np.sum(d * d)), g


def with_masses(scn, mass):
    return dataclasses.replace(scn, mass=np.array(mass, dtype=np.float64))


def oracle_rk4(scn, y0, h, nsteps, t0=0.0):
    """nsteps RK4 steps with the oracle's Compute + Step, the Driver's order
    (RungeKutta4::Driver, f0 evaluated at the step's start state)."""
    o = Oracle(scn)
    y = np.array(y0, dtype=np.float64)
    t = t0
    hs = np.broadcast_to(np.asarray(h, dtype=np.float64), (nsteps,))
    for k in range(nsteps):
        o.set_state(y)
        o.set_accel(o.compute(t, y))
        assert o.step(S.RUNGE_KUTTA4, t, float(hs[k])) == 0
        y = o.get_array(1)
        t += float(hs[k])
    o.close()
    return y


def central_difference(f, delta):
    """Fourth-order central difference of a scalar function of one
    variable at 0: [8(f(d) - f(-d)) - (f(2d) - f(-2d))] / (12 d)."""
    return (8.0 * (f(delta) - f(-delta)) - (f(2 * delta) - f(-2 * delta))) / (12.0 * delta)


def small_c5(n_tp=24, n_pl=4):
    return S.make_c5(n_tp=n_tp, n_pl=n_pl)

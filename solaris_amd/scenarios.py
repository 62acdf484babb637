"""Seeded synthetic inputs for the force + integration path.

Bodies are produced type-sorted (central, giant, rocky, proto,
super-planetesimal, planetesimal, test particle), as
``Simulator::BodyListToBodyData`` lays them out
(Solaris/Simulator.cpp:449-495, :525-593), in the astrocentric frame with
the star's state exactly zero.

Randomness comes from a SplitMix64 stream of our own (53-bit uniforms), so
a seed fixes the input on every machine.  Orbital elements are turned into
Cartesian states with mu = k^2 (M_star + m), or k^2 M_star for test
particles (Solaris/Simulation.cpp:143-146), through a Newton solution of
Kepler's equation iterated to |dE| < 1e-15.

Configurations (sizes are the workload's own):
  c1  star + Jupiter + Saturn (giants) + 512 test particles, RKF78, 1e-10
  c2  star + 4 type-I migrating protoplanets + 16,384 planetesimals in a
      gas disk with Stokes drag (Sigma ~ r^-1.5, H/r = 0.05, Cd = 0.44,
      2 g/cm^3), RKF78
  c3  star + 65,536 equal-mass (1e-10 Msun) self-gravitating protoplanets
      in a disk a~U[0.5,5], e~Rayleigh(0.02), i~Rayleigh(0.01), RKF78
  c4  one member of the 4,096-system ensemble: star + n~U{2..256}
      protoplanets, masses log-U[1e-7,1e-3], a log-U[0.3,30], e~U[0,0.3]
  c5  star + 8 protoplanets + 1,024 test particles, RK4, h = 1 day
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# Solaris/Constants.h, the same defining expressions.
GAUSS = 1.720209895e-2
GAUSS2 = 2.959122082855911025e-4
PI = 3.14159265358979323846
DEG = 1.745329251994329509e-2
YEAR_TO_DAY = 365.25
SOLAR_TO_GRAM = 1.0e3 * 1.98911e30
GRAM_TO_SOLAR = 1.0 / SOLAR_TO_GRAM
METER_TO_AU = 1.0 / 1.495978707e11
GCM3_TO_SOLAR_AU3 = GRAM_TO_SOLAR / ((1.0e-2 * METER_TO_AU) ** 3)
SOLAR_RADIUS_TO_AU = 1.0 / 215.094
PROTON_MASS_CMU = 1.672621777e-27 * (1.0 / 1.98911e30)

CENTRAL, GIANT, ROCKY, PROTO, SUPERPL, PL, TP = 1, 2, 3, 4, 5, 6, 7
MIG_NO, MIG_I, MIG_II = 0, 1, 2
DORMAND_PRINCE, RUNGE_KUTTA4, RUNGE_KUTTA_FEHLBERG78 = 0, 1, 3

SEEDS = {"c1": 0x501A0001, "c2": 0x501A0002, "c3": 0x501A0003,
         "c4": 0x501A0004, "c5": 0x501A0005}


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014) with 53-bit uniform doubles."""

    _M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self._M

    def next_u64(self, n: int) -> np.ndarray:
        """The next n outputs; state k is seed + k * golden (mod 2^64)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & self._M
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        """n uniforms in [lo, hi) from the top 53 bits."""
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def log_uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        return np.exp(self.uniform(n, math.log(lo), math.log(hi)))

    def rayleigh(self, n: int, sigma: float) -> np.ndarray:
        return sigma * np.sqrt(-2.0 * np.log1p(-self.uniform(n)))

    def normal(self, n: int, sigma: float = 1.0) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return sigma * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * PI * u2)

    def power_law(self, n: int, lo: float, hi: float, slope: float) -> np.ndarray:
        """dN/dm ~ m^slope on [lo, hi] by inverse CDF (slope != -1)."""
        q = slope + 1.0
        u = self.uniform(n)
        return (lo ** q + u * (hi ** q - lo ** q)) ** (1.0 / q)


def kepler_solve(M: np.ndarray, e: np.ndarray) -> np.ndarray:
    """Eccentric anomaly from E - e sin E = M by Newton's method."""
    E = np.where(e < 0.8, M, np.full_like(M, PI))
    for _ in range(60):
        f = E - e * np.sin(E) - M
        dE = f / (1.0 - e * np.cos(E))
        E = E - dE
        if np.all(np.abs(dE) < 1e-15):
            break
    return E


def elements_to_cartesian(mu, a, e, inc, peri, node, mean) -> np.ndarray:
    """Keplerian elements -> [x, y, z, vx, vy, vz] (N x 6)."""
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), np.shape(a))
    E = kepler_solve(np.mod(mean, 2.0 * PI), e)
    cE, sE = np.cos(E), np.sin(E)
    s1e2 = np.sqrt(1.0 - e * e)
    xo = a * (cE - e)
    yo = a * s1e2 * sE
    n = np.sqrt(mu / (a * a * a))
    Edot = n / (1.0 - e * cE)
    vxo = -a * sE * Edot
    vyo = a * s1e2 * cE * Edot
    cw, sw = np.cos(peri), np.sin(peri)
    cO, sO = np.cos(node), np.sin(node)
    ci, si = np.cos(inc), np.sin(inc)
    P = np.stack([cw * cO - sw * sO * ci, cw * sO + sw * cO * ci, sw * si], axis=-1)
    Q = np.stack([-sw * cO - cw * sO * ci, -sw * sO + cw * cO * ci, cw * si], axis=-1)
    r = xo[:, None] * P + yo[:, None] * Q
    v = vxo[:, None] * P + vyo[:, None] * Q
    return np.concatenate([r, v], axis=1)


def default_gas() -> dict:
    """GasComponent() defaults (Solaris/GasComponent.cpp:9-34)."""
    g = dict(alpha=2.0e-3, mean_molecular_weight=2.3, particle_diameter=3.0e-10,
             type=0, time_scale=0.0, t0=0.0, t1=0.0,
             inner_edge=10.0 * SOLAR_RADIUS_TO_AU,
             eta=(0.0019, 0.5), tau=(2.0 / 3.0, 2.0), scale_height=(0.02, 1.25),
             density=(1.4e-9 * GCM3_TO_SOLAR_AU3, -2.75))
    g["a"] = g["density"][0] * g["inner_edge"] ** g["density"][1] / (g["inner_edge"] ** 2) ** 2
    clambda = (g["mean_molecular_weight"] * PROTON_MASS_CMU /
               (math.sqrt(2.0) * PI * (g["particle_diameter"] * METER_TO_AU) ** 2 * g["density"][0]))
    g["mean_free_path"] = (clambda, -g["density"][1])
    return g


@dataclasses.dataclass
class Scenario:
    name: str
    counts: dict                      # NBodies
    mass: np.ndarray
    radius: np.ndarray
    density: np.ndarray
    cd: np.ndarray
    gamma_stokes: np.ndarray
    gamma_epstein: np.ndarray
    mig_stop_at: np.ndarray
    type: np.ndarray
    mig_type: np.ndarray
    y0: np.ndarray                    # (N, 6)
    gas: Optional[dict] = None
    integrator: int = RUNGE_KUTTA_FEHLBERG78
    epsilon: float = 1.0e-10
    h_next: float = 0.0
    length: float = 0.0
    softening: float = 0.0
    target_offsets: Optional[np.ndarray] = None   # (N, 3), adjoint loss targets

    @property
    def n(self) -> int:
        return int(self.counts["total"])

    def n_massive(self) -> int:
        c = self.counts
        return c["central_body"] + c["giant_planet"] + c["rocky_planet"] + c["proto_planet"]

    def interactions_per_evaluation(self) -> float:
        """Sum over sinks i >= 1 of (1 + |S(i) minus i|) under the
        reference's astrocentric source ranges (Solaris/Acceleration.cpp:285-289)."""
        nm = self.n_massive()
        end_a = nm + self.counts["super_planetesimal"]
        n_a = nm - 1                       # sinks of type <= ProtoPlanet
        n_b = self.n - nm
        return float(n_a * (1 + max(end_a - 2, 0)) + n_b * (1 + max(nm - 1, 0)))

    def shortest_period(self) -> float:
        """Simulator::ShortestPeriod (Solaris/Simulator.cpp:510-523)."""
        m0 = self.mass[0]
        y = self.y0[1:]
        mu = GAUSS2 * (m0 + np.where(self.type[1:] == TP, 0.0, self.mass[1:]))
        h = 0.5 * np.sum(y[:, 3:] ** 2, axis=1) - mu / np.sqrt(np.sum(y[:, :3] ** 2, axis=1))
        a = -mu / (2.0 * h)
        p = np.where(h > 0.0, -1.0, 2.0 * PI * np.sqrt(a ** 3 / mu))
        p = p[p > 0]
        return float(min(1.0e10, p.min())) if p.size else 1.0e10


def _counts(**kw) -> dict:
    c = dict(central_body=0, giant_planet=0, rocky_planet=0, proto_planet=0,
             super_planetesimal=0, planetesimal=0, test_particle=0)
    c.update(kw)
    c["total"] = sum(c.values())
    return c


def _assemble(name, counts, groups, gas=None, **kw) -> Scenario:
    """groups: list of dicts with mass, type, elements, optional physics."""
    mass, typ, mig, stop, rad, den, cd, ys = [], [], [], [], [], [], [], []
    for g in groups:
        n = len(g["mass"])
        mass.append(g["mass"]); typ.append(np.full(n, g["type"], np.int32))
        mig.append(np.full(n, g.get("mig", MIG_NO), np.int32))
        stop.append(np.full(n, g.get("stop", 0.0)))
        rad.append(g.get("radius", np.zeros(n))); den.append(g.get("density", np.zeros(n)))
        cd.append(g.get("cd", np.zeros(n))); ys.append(g["y0"])
    mass = np.concatenate(mass); typ = np.concatenate(typ); rad = np.concatenate(rad)
    den = np.concatenate(den); cd = np.concatenate(cd)
    # BodyListToBodyData (Solaris/Simulator.cpp:548-568): test particles are
    # massless; gammas from Characteristics (Solaris/Characteristics.cpp:75-87)
    is_tp = typ == TP
    mass = np.where(is_tp, 0.0, mass)
    with np.errstate(divide="ignore", invalid="ignore"):
        ge = np.where((rad > 0) & ~is_tp, 1.0 / (den * rad), 0.0)
        gs = np.where((rad > 0) & (cd > 0) & ~is_tp, (3.0 / 8.0) * cd / (den * rad), 0.0)
    return Scenario(name=name, counts=counts, mass=mass, radius=np.where(is_tp, 0.0, rad),
                    density=np.where(is_tp, 0.0, den), cd=np.where(is_tp, 0.0, cd),
                    gamma_stokes=gs, gamma_epstein=ge, mig_stop_at=np.concatenate(stop),
                    type=typ, mig_type=np.concatenate(mig), y0=np.concatenate(ys, axis=0),
                    gas=gas, **kw)


def _star(mass=1.0):
    return dict(mass=np.array([mass]), type=CENTRAL, y0=np.zeros((1, 6)))


def _orbits(rng, n, m_star, masses, a, e, inc, massless=False):
    node = rng.uniform(n, 0.0, 2.0 * PI)
    peri = rng.uniform(n, 0.0, 2.0 * PI)
    mean = rng.uniform(n, 0.0, 2.0 * PI)
    mu = GAUSS2 * (m_star + (0.0 if massless else masses))
    return elements_to_cartesian(mu, a, e, inc, peri, node, mean)


def radius_from_density(mass, density):
    """Characteristics::CalculateRadius (Solaris/Characteristics.cpp:57-60)."""
    return np.power(3.0 / (4.0 * PI) * mass / density, 1.0 / 3.0)


def make_c1(n_tp: int = 512, seed: int = SEEDS["c1"], years: float = 1000.0) -> Scenario:
    rng = SplitMix64(seed)
    gm = np.array([9.54e-4, 2.86e-4])
    ga = np.array([5.2, 9.58])
    ge = rng.uniform(2, 0.0, 0.05)
    gi = rng.uniform(2, 0.0, 1.0 * DEG)
    giants = dict(mass=gm, type=GIANT, y0=_orbits(rng, 2, 1.0, gm, ga, ge, gi))
    ta = rng.uniform(n_tp, 3.0, 12.0)
    te = rng.uniform(n_tp, 0.0, 0.1)
    ti = rng.uniform(n_tp, 0.0, 2.0 * DEG)
    tps = dict(mass=np.zeros(n_tp), type=TP, y0=_orbits(rng, n_tp, 1.0, 0.0, ta, te, ti, massless=True))
    s = _assemble("c1", _counts(central_body=1, giant_planet=2, test_particle=n_tp),
                  [_star(), giants, tps], length=years * YEAR_TO_DAY)
    s.h_next = s.shortest_period() / 50000.0
    return s


def c2_gas() -> dict:
    """Nebula of the drag workload: Sigma ~ r^-1.5 (volume density index
    -2.5), H/r = 0.05 (scale height c = 0.05, index 1), rho0 = 1.4e-9 g/cm^3;
    eta, tau and the mean free path law keep the constructor defaults."""
    g = default_gas()
    g["scale_height"] = (0.05, 1.0)
    g["density"] = (1.4e-9 * GCM3_TO_SOLAR_AU3, -2.5)
    return g


def make_c2(n_pl: int = 16384, n_proto: int = 4, seed: int = SEEDS["c2"],
            self_gravity: bool = False) -> Scenario:
    rng = SplitMix64(seed)
    rho = 2.0 * GCM3_TO_SOLAR_AU3
    pm = rng.log_uniform(n_proto, 1.0e-4, 1.0e-3)
    pa = rng.uniform(n_proto, 5.0, 30.0)
    pe = rng.uniform(n_proto, 0.0, 0.05)
    pi_ = rng.uniform(n_proto, 0.0, 1.0 * DEG)
    protos = dict(mass=pm, type=PROTO, mig=MIG_I, stop=0.0,
                  radius=radius_from_density(pm, rho), density=np.full(n_proto, rho),
                  y0=_orbits(rng, n_proto, 1.0, pm, pa, pe, pi_))
    m = rng.power_law(n_pl, 1.0e-12, 1.0e-9, -1.8)
    a = rng.uniform(n_pl, 1.0, 40.0)
    e = np.minimum(rng.rayleigh(n_pl, 0.01), 0.9)
    i = rng.rayleigh(n_pl, 0.005)
    pls = dict(mass=m, type=PL, radius=radius_from_density(m, rho), density=np.full(n_pl, rho),
               cd=np.full(n_pl, 0.44), y0=_orbits(rng, n_pl, 1.0, m, a, e, i))
    s = _assemble("c2", _counts(central_body=1, proto_planet=n_proto, planetesimal=n_pl),
                  [_star(), protos, pls], gas=c2_gas())
    s.h_next = s.shortest_period() / 50000.0
    return s


def make_c3(n: int = 65536, seed: int = SEEDS["c3"], softening: float = 1.0e-5) -> Scenario:
    rng = SplitMix64(seed)
    m = np.full(n, 1.0e-10)
    a = rng.uniform(n, 0.5, 5.0)
    e = np.minimum(rng.rayleigh(n, 0.02), 0.9)
    i = rng.rayleigh(n, 0.01)
    bodies = dict(mass=m, type=PROTO, y0=_orbits(rng, n, 1.0, m, a, e, i))
    s = _assemble("c3", _counts(central_body=1, proto_planet=n), [_star(), bodies],
                  softening=softening)
    s.h_next = s.shortest_period() / 50000.0
    return s


def make_c4_system(k: int, seed: int = SEEDS["c4"]) -> Scenario:
    """System k of the ensemble; each system has its own SplitMix64 stream."""
    rng = SplitMix64(seed ^ (0x9E3779B97F4A7C15 * (k + 1) & ((1 << 64) - 1)))
    n = int(2 + (rng.next_u64(1)[0] % np.uint64(255)))
    m = rng.log_uniform(n, 1.0e-7, 1.0e-3)
    a = rng.log_uniform(n, 0.3, 30.0)
    e = rng.uniform(n, 0.0, 0.3)
    i = rng.uniform(n, 0.0, 1.0 * DEG)
    bodies = dict(mass=m, type=PROTO, y0=_orbits(rng, n, 1.0, m, a, e, i))
    s = _assemble(f"c4[{k}]", _counts(central_body=1, proto_planet=n), [_star(), bodies],
                  length=1.0e4 * YEAR_TO_DAY)
    s.h_next = s.shortest_period() / 50000.0
    return s


def make_c5(n_tp: int = 1024, n_pl: int = 8, seed: int = SEEDS["c5"]) -> Scenario:
    rng = SplitMix64(seed)
    pm = rng.log_uniform(n_pl, 1.0e-6, 1.0e-3)
    pa = rng.log_uniform(n_pl, 0.5, 30.0)
    pe = rng.uniform(n_pl, 0.0, 0.05)
    pi_ = rng.uniform(n_pl, 0.0, 1.0 * DEG)
    planets = dict(mass=pm, type=PROTO, y0=_orbits(rng, n_pl, 1.0, pm, pa, pe, pi_))
    ta = rng.uniform(n_tp, 0.5, 30.0)
    te = rng.uniform(n_tp, 0.0, 0.1)
    ti = rng.uniform(n_tp, 0.0, 2.0 * DEG)
    tps = dict(mass=np.zeros(n_tp), type=TP, y0=_orbits(rng, n_tp, 1.0, 0.0, ta, te, ti, massless=True))
    s = _assemble("c5", _counts(central_body=1, proto_planet=n_pl, test_particle=n_tp),
                  [_star(), planets, tps], integrator=RUNGE_KUTTA4, h_next=1.0,
                  length=2000.0)
    # loss targets r_i* = r_i(T) of the unperturbed run + these offsets
    s.target_offsets = rng.normal(s.n * 3, 1.0e-3).reshape(s.n, 3)
    s.target_offsets[0] = 0.0
    return s


def make(name: str, **kw) -> Scenario:
    return {"c1": make_c1, "c2": make_c2, "c3": make_c3, "c5": make_c5}[name](**kw)

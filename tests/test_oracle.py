"""Pins the FP64 CPU oracle (oracle/ddyn_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a wrong RK4 weight or stage velocity) fails:
  * long-double golden vectors from the textbook physics (tests/golden);
  * closed forms: two-body acceleration, Kepler orbits (RK4 global order 4),
    Wyatt-Whipple PR-drag decay, circular speed with reduced gravity;
  * invariants: Newton's third law, momentum, energy, angular momentum,
    the Jacobi integral of the circular restricted three-body problem;
  * removal rules: sublimation inside Rmin, exact hit on a massive body.
"""
import math

import numpy as np
import pytest

from diskdyn_amd import orbits, synth
from diskdyn_amd.system import FP64, Params, make_state
from diskdyn_amd.units import GRAV
from oracle import oracle as O
from tests.paritylib import golden_state

def test_golden_accelerations():
    st, prm, g = golden_state()
    a, rem = O.accel(prm, st)
    assert not rem.any()
    err = np.linalg.norm(a - g["acc"], axis=1) / np.linalg.norm(g["acc"], axis=1)
    assert err.max() < 2e-14, err.max()


def test_golden_rk4_step():
    st, prm, g = golden_state()
    O.step(prm, st, float(g["step_dt"][0]), 1)
    ep = np.linalg.norm(st.pos() - g["pos1"], axis=1) / np.linalg.norm(g["pos1"], axis=1)
    ev = np.linalg.norm(st.vel() - g["vel1"], axis=1) / np.linalg.norm(g["vel1"], axis=1)
    assert ep.max() < 1e-14 and ev.max() < 1e-14, (ep.max(), ev.max())


def test_golden_catches_planted_mistakes():
    """The golden comparison must fail for a drag sign flip or a missing Lorentz term."""
    st, prm, g = golden_state()
    from dataclasses import replace
    for bad in (replace(prm, FLQ=0), replace(prm, light=-prm.light), replace(prm, eps=0.0)):
        a, _ = O.accel(bad, st)
        err = np.linalg.norm(a - g["acc"], axis=1) / np.linalg.norm(g["acc"], axis=1)
        assert err.max() > 1e-9


def test_two_body_acceleration():
    mu0, r = GRAV, 7.3
    a = O.massive_accel(Params(n_planets=2), np.array([0.0, r * 0.6]), np.array([0.0, r * 0.8]),
                        np.array([0.0, 0.0]), np.array([mu0, 1e-3 * GRAV]))
    np.testing.assert_allclose(np.linalg.norm(a[1]), mu0 / r ** 2, rtol=1e-15)
    np.testing.assert_allclose(a[1] / np.linalg.norm(a[1]), [-0.6, -0.8, 0.0], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(a[0] * mu0, -a[1] * 1e-3 * GRAV, rtol=1e-15)


@pytest.mark.parametrize("eps", [0.0, 1e-2])
def test_newton_third_law_and_brute_force(eps):
    w = synth.config2(n_swarm=200, precision=FP64, eps=eps)
    st = w.state
    # two coincident swarm bodies: zero mutual force (f == 0 guard) without softening
    st.x[5], st.y[5], st.z[5] = st.x[6], st.y[6], st.z[6]
    a = O.massive_accel(w.params, st.x, st.y, st.z, st.mu)
    F = st.mu[:, None] * a
    assert np.abs(F.sum(0)).max() <= 1e-15 * np.abs(F).sum()
    # long-double brute force
    P = st.pos().astype(np.longdouble)
    M = st.mu.astype(np.longdouble)
    for i in (0, 1, 5, 6, 100, 200):
        acc = np.zeros(3, dtype=np.longdouble)
        for j in range(st.n_grav):
            if j == i:
                continue
            d = P[j] - P[i]
            r2 = d @ d + (np.longdouble(eps) ** 2 if (i >= 1 and j >= 1) else 0)
            if r2 > 0:
                acc += M[j] * d / (r2 * np.sqrt(r2))
        assert np.linalg.norm(a[i] - acc.astype(float)) <= 1e-14 * np.linalg.norm(acc.astype(float))


def _test_particle(a0, e0, nu0, beta=0.0):
    st = make_state(1, 1)
    st.mu[0] = GRAV
    r = orbits.coe2rv(GRAV, 0.0, a0, e0, 0.3, 0.7, 1.1, nu0, beta)
    st.x[1], st.y[1], st.z[1], st.vx[1], st.vy[1], st.vz[1] = [float(v) for v in r]
    st.mu[1] = 0.0
    st.beta_p = np.array([beta])
    st.betasw_p = np.array([0.0])
    return st


def test_rk4_global_order_four_kepler():
    """Test particle on an e = 0.3 orbit for 2 periods: global error ~ h^4."""
    prm = Params(n_planets=1)
    a0 = 1.5
    T = 2.0 * 2 * math.pi * math.sqrt(a0 ** 3 / GRAV)
    errs = []
    for nsteps in (1000, 2000, 4000):
        st = _test_particle(a0, 0.3, 0.4)
        r_ex, _ = orbits.kepler_propagate(GRAV, st.x[1], st.y[1], st.z[1], st.vx[1], st.vy[1], st.vz[1], T)
        O.step(prm, st, T / nsteps, nsteps)
        errs.append(np.linalg.norm(st.pos()[1] - r_ex))
    p1, p2 = math.log2(errs[0] / errs[1]), math.log2(errs[1] / errs[2])
    assert 3.7 < p1 < 4.3 and 3.7 < p2 < 4.3, (errs, p1, p2)


def test_rk4_order_massive_two_body():
    """Star + planet: the relative orbit is Kepler with mu = mu0 + mu1."""
    mu0, mu1 = GRAV, 0.01 * GRAV
    a0, e0 = 2.0, 0.2
    T = 2 * math.pi * math.sqrt(a0 ** 3 / (mu0 + mu1))
    errs = []
    for nsteps in (300, 600):
        st = make_state(2, 0)
        st.mu[:] = [mu0, mu1]
        r = orbits.coe2rv(mu0, mu1, a0, e0, 0.1, 0.2, 0.3, 0.5)
        st.x[1], st.y[1], st.z[1], st.vx[1], st.vy[1], st.vz[1] = [float(v) for v in r]
        rr0 = np.array([st.x[1] - st.x[0], st.y[1] - st.y[0], st.z[1] - st.z[0]])
        vv0 = np.array([st.vx[1] - st.vx[0], st.vy[1] - st.vy[0], st.vz[1] - st.vz[0]])
        r_ex, _ = orbits.kepler_propagate(mu0 + mu1, *rr0, *vv0, T)
        P0 = (st.mu[:, None] * st.vel()).sum(0)
        O.step(Params(n_planets=2), st, T / nsteps, nsteps)
        errs.append(np.linalg.norm(st.pos()[1] - st.pos()[0] - r_ex))
        # barycentre moves uniformly: total momentum conserved
        P = (st.mu[:, None] * st.vel()).sum(0)
        assert np.abs(P - P0).max() < 1e-13 * np.abs(P0).max()
    assert 3.7 < math.log2(errs[0] / errs[1]) < 4.3, errs


def test_radiation_pressure_reduces_gravity():
    """With beta and no drag term resolved over one orbit, a circular orbit at
    v = sqrt((1 - beta) mu / r) stays circular: radius constant to O(PR drift)."""
    beta = 0.3
    st = _test_particle(4.0, 0.0, 0.0, beta)
    prm = Params(n_planets=1, light=1e30)   # no PR drag
    T = 2 * math.pi * math.sqrt(4.0 ** 3 / ((1 - beta) * GRAV))
    O.step(prm, st, T / 2000, 2000)
    r = math.sqrt(st.x[1] ** 2 + st.y[1] ** 2 + st.z[1] ** 2)
    assert abs(r - 4.0) < 1e-9


@pytest.mark.parametrize("swq", [0, 1])
def test_pr_drag_wyatt_whipple(swq):
    """da/dt = -(beta mu / c)(2 + 3 e^2) / (a (1 - e^2)^(3/2)) (Wyatt & Whipple 1950);
    stellar-wind drag scales the rate by (1 + (beta_sw/v_sw)/(beta/c))."""
    beta, a0, e0 = 0.2, 3.0, 0.3
    c = 2000.0       # a slow "light" speed makes the decay large enough to measure
    vsw = 200.0
    betasw = 0.35 * beta * vsw / c if swq else 0.0
    st = _test_particle(a0, e0, 0.0, beta + betasw)
    st.betasw_p = np.array([betasw])
    prm = Params(n_planets=1, SWQ=swq, light=c, vsw=vsw if swq else 0.0)
    mu_eff = (1 - beta - betasw) * GRAV
    P = 2 * math.pi * math.sqrt(a0 ** 3 / mu_eff)
    norb, per = 4, 100
    t_hist, a_hist = [], []
    for k in range(norb * per + 1):
        a_k, _, _ = orbits.rv2coe(mu_eff, 0.0, st.x[1], st.y[1], st.z[1], st.vx[1], st.vy[1], st.vz[1])
        t_hist.append(k * P / per)
        a_hist.append(a_k)
        O.step(prm, st, P / (40 * per), 40)
    # orbit-averaged rate: least-squares slope over whole orbits
    rate = np.polyfit(np.array(t_hist[:-1]), np.array(a_hist[:-1]), 1)[0]
    scale = 1.0 + (0.35 if swq else 0.0)
    expect = -scale * beta * GRAV / c * (2 + 3 * e0 ** 2) / (a0 * (1 - e0 ** 2) ** 1.5)
    assert abs(rate / expect - 1) < 0.02, (rate, expect)
    # a sign error in either drag term flips or shrinks the rate by O(1)


def test_sublimation_and_hit_remove_grain():
    st = make_state(2, 3)
    st.mu[:2] = [GRAV, 1e-3 * GRAV]
    st.x[1], st.vy[1] = 5.0, math.sqrt(GRAV / 5.0)
    st.x[2], st.vy[2] = 0.05, 1.0             # inside Rmin = 0.1
    st.x[3], st.vy[3] = 5.0, 0.3              # exactly on the planet
    st.x[4], st.vy[4] = 8.0, math.sqrt(GRAV / 8.0)
    st.mu[2:] = 0.0
    prm = Params(n_planets=2, Rmin=0.1)
    a, rem = O.accel(prm, st)
    assert list(rem) == [1, 1, 0]
    before = st.copy()
    O.step(prm, st, 0.01, 1)
    assert list(st.N[2:]) == [0.0, 0.0, 1.0]
    assert st.x[2] == before.x[2] and st.x[3] == before.x[3]   # removed grains are not advanced
    assert st.x[4] != before.x[4]
    O.step(prm, st, 0.01, 1)                  # N == 0 grains stay frozen
    assert st.x[2] == before.x[2]


def test_energy_and_momentum_conservation():
    w = synth.config2(n_swarm=300, precision=FP64)
    st = w.state.copy()
    e0 = O.energy(w.params, st)
    O.step(w.params, st, w.dt, 50)
    e1 = O.energy(w.params, st)
    assert abs((e1[0] - e0[0]) / e0[0]) < 1e-10
    assert np.abs(e1[1:4] - e0[1:4]).max() < 1e-12 * np.linalg.norm(e0[1:4])
    assert np.abs(e1[4:] - e0[4:]).max() < 1e-14


def test_jacobi_integral():
    w = synth.config1_jacobi(n_dust=64, nsteps=500)
    st = w.state.copy()
    n = w.meta["mean_motion"]
    jacobi = lambda s: synth.jacobi_constants(s, n)  # noqa: E731

    c0 = jacobi(st)
    O.step(w.params, st, w.dt, w.nsteps)
    c1 = jacobi(st)
    assert np.abs((c1 - c0) / c0).max() < 1e-10
    # the planet stays on its circle
    r = math.hypot(st.x[1] - st.x[0], st.y[1] - st.y[0])
    assert abs(r - 5.2) < 1e-9


def test_coe2rv_matches_oracle_and_closed_forms():
    rng = np.random.default_rng(3)
    for _ in range(20):
        M, m = GRAV, rng.uniform(0, 1e-3) * GRAV
        a, e = rng.uniform(1, 50), rng.uniform(0, 0.9)
        inc, w, o, nu, beta = rng.uniform(0, 3), rng.uniform(0, 6), rng.uniform(0, 6), rng.uniform(0, 6), rng.uniform(0, 0.4)
        r_py = np.array([float(v) for v in orbits.coe2rv(M, m, a, e, inc, w, o, nu, beta)])
        r_c = O.coe2rv(M, m, a, e, inc, w, o, nu, beta)
        np.testing.assert_allclose(r_py, r_c, rtol=1e-13, atol=1e-13)
        mu = (M + m) * (1 - beta)
        R = np.linalg.norm(r_py[:3])
        np.testing.assert_allclose(R, a * (1 - e * e) / (1 + e * math.cos(nu)), rtol=1e-13)
        np.testing.assert_allclose(r_py[3:] @ r_py[3:], mu * (2 / R - 1 / a), rtol=1e-12)   # vis-viva
        a2, e2, i2 = orbits.rv2coe(M, m, *r_py, beta)
        np.testing.assert_allclose([a2, e2, i2], [a, e, inc], rtol=1e-9, atol=1e-11)


def test_recalc_com():
    w = synth.config1(n_dust=8)
    st = w.state
    sl = slice(0, 3)
    P = (st.mu[sl, None] * st.vel()[sl]).sum(0)
    X = (st.mu[sl, None] * st.pos()[sl]).sum(0)
    assert np.abs(P).max() < 1e-15 and np.abs(X).max() < 1e-14
    # oracle restatement agrees with the product's
    raw = synth.config1(n_dust=8)
    xs = [a[:3].copy() for a in (raw.state.x, raw.state.y, raw.state.z, raw.state.vx, raw.state.vy, raw.state.vz)]
    out = O.recalc_com(*xs, raw.state.mu[:3])
    np.testing.assert_allclose(np.array(out), np.array(xs), rtol=1e-13, atol=1e-15)

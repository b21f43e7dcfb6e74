"""Pins the oracle's discrete adjoint (orc_adjoint) with central finite
differences of the oracle's own forward map: directional derivatives in the
initial state, and partial derivatives in masses and grain beta."""
import numpy as np
import pytest

from diskdyn_amd import synth
from diskdyn_amd.system import FP64
from oracle import oracle as O


def _system():
    w = synth.config3(n_dust=8, n_swarm=6, precision=FP64, eps=0.05)
    return w


def _loss(w, st, lam_r, lam_v, nsteps):
    s = O.step(w.params, st.copy(), w.dt, nsteps)
    return float((lam_r * s.pos()).sum() + (lam_v * s.vel()).sum())


@pytest.mark.parametrize("nsteps", [1, 6])
def test_adjoint_state_directional(nsteps):
    w = _system()
    st = w.state
    rng = np.random.default_rng(1)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    g_r, g_v, _, _ = O.adjoint(w.params, st, w.dt, nsteps, lam_r, lam_v)
    for trial in range(3):
        dr, dv = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))

        def moved(h):
            s = st.copy()
            s.x, s.y, s.z = (np.ascontiguousarray(st.pos()[:, k] + h * dr[:, k]) for k in range(3))
            s.vx, s.vy, s.vz = (np.ascontiguousarray(st.vel()[:, k] + h * dv[:, k]) for k in range(3))
            return s

        h = 1e-6
        fd = (_loss(w, moved(h), lam_r, lam_v, nsteps) - _loss(w, moved(-h), lam_r, lam_v, nsteps)) / (2 * h)
        an = float((g_r * dr).sum() + (g_v * dv).sum())
        assert abs(an - fd) <= 1e-7 * abs(fd), (an, fd)


def test_adjoint_masses_and_beta():
    w = _system()
    st = w.state
    nsteps = 5
    rng = np.random.default_rng(2)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    _, _, g_mu, g_beta = O.adjoint(w.params, st, w.dt, nsteps, lam_r, lam_v)
    for j in range(st.n_grav):
        # star/planet: relative step; swarm bodies are tiny: absolute step (L is near-linear in them)
        h = 1e-2 * st.mu[j] if j < w.params.n_planets else 1e-6
        sp, sm = st.copy(), st.copy()
        sp.mu[j] += h
        sm.mu[j] -= h
        fd = (_loss(w, sp, lam_r, lam_v, nsteps) - _loss(w, sm, lam_r, lam_v, nsteps)) / (2 * h)
        assert abs(g_mu[j] - fd) <= 2e-5 * abs(fd) + 1e-12, (j, g_mu[j], fd)
    for g in range(st.n_dust):
        h = 1e-3 * st.beta_p[g]
        sp, sm = st.copy(), st.copy()
        sp.beta_p[g] += h
        sm.beta_p[g] -= h
        fd = (_loss(w, sp, lam_r, lam_v, nsteps) - _loss(w, sm, lam_r, lam_v, nsteps)) / (2 * h)
        assert abs(g_beta[g] - fd) <= 1e-6 * abs(fd), (g, g_beta[g], fd)


def test_adjoint_zero_steps_is_identity():
    w = _system()
    st = w.state
    rng = np.random.default_rng(3)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    g_r, g_v, g_mu, g_beta = O.adjoint(w.params, st, w.dt, 0, lam_r, lam_v)
    assert np.array_equal(g_r, lam_r) and np.array_equal(g_v, lam_v)
    assert not g_mu.any() and not g_beta.any()

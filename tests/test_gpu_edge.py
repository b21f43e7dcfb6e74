"""GPU edge cases against the FP64 CPU oracle: a star alone with grains, only
planets (more than the multi-step kernel takes), a swarm of one body, ragged
sizes that fill no tile or workgroup, zero steps, and a context re-used across
uploads of different sizes.

Bars as in test_gpu_parity: plain FP64 <= 1e-12 relative on accelerations and
<= 1e-10 on positions; FP32-force <= 1e-5 median relative acceleration error.
"""
from dataclasses import replace

import numpy as np
import pytest

from diskdyn_amd import synth
from diskdyn_amd.system import FP32_FORCE, FP64
from oracle import oracle as O
from tests.paritylib import pos_rel_err, rel_err_rows, subset

pytestmark = pytest.mark.gpu


def debris(n_swarm, n_dust, precision):
    w = synth.config3(n_dust=max(n_dust, 1), n_swarm=max(n_swarm, 1), precision=precision)
    return w


def check_accel(engine_cls, st, prm):
    with engine_cls() as eng:
        eng.upload(st.copy(), prm)
        a_gpu, rem_gpu = eng.accel()
    a_ref, rem_ref = O.accel(prm, st)
    err = rel_err_rows(a_gpu, a_ref)
    assert np.array_equal(rem_gpu, rem_ref)
    if prm.precision == FP64:
        assert err.max() <= 1e-12, err.max()
    else:
        assert np.median(err) <= 1e-5
        assert err.max() <= 1e-3
    return err


def check_steps(engine_cls, st, prm, dt, nsteps):
    s = st.copy()
    with engine_cls() as eng:
        eng.upload(s, prm)
        eng.step(dt, nsteps)
        eng.download(s)
    ref = O.step(prm, st.copy(), dt, nsteps)
    assert np.array_equal(s.N, ref.N)
    ep = pos_rel_err(s, ref)
    if prm.precision == FP64:
        assert ep.max() <= 1e-10, ep.max()
    else:
        assert np.median(ep) <= 1e-8
    return s


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE])
def test_star_alone_with_grains(engine_cls, precision):
    w = debris(4, 200, precision)
    st = subset(w.state, [0], range(200))
    prm = replace(w.params, n_planets=1)
    print(f"[edge] star alone + 200 grains, precision={precision}", flush=True)
    check_accel(engine_cls, st, prm)
    check_steps(engine_cls, st, prm, w.dt, 12)


@pytest.mark.parametrize("schedule", [0, 1])
def test_planets_only_beyond_multistep_kernel(engine_cls, schedule):
    """40 massive bodies, all unsoftened planets (no swarm), with grains: the
    per-kernel path with no pair launches, in both schedules."""
    w = debris(38, 130, FP64)
    st = w.state.copy()
    prm = replace(w.params, n_planets=st.n_grav, schedule=schedule)
    assert st.n_grav == 40
    print(f"[edge] 40 planets + 130 grains, schedule={schedule}", flush=True)
    check_accel(engine_cls, st, prm)
    check_steps(engine_cls, st, prm, w.dt, 9)


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE])
def test_swarm_of_one_body(engine_cls, precision):
    w = debris(1, 65, precision)
    st = w.state
    assert st.n_grav - w.params.n_planets == 1 and st.n_dust == 65
    print(f"[edge] one swarm body + 65 grains, precision={precision}", flush=True)
    check_accel(engine_cls, st, w.params)
    check_steps(engine_cls, st, w.params, w.dt, 7)


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE])
def test_ragged_sizes(engine_cls, precision):
    """Swarm of 300 (a partial source tile), 1,027 grains (a partial workgroup)."""
    w = debris(300, 1027, precision)
    print(f"[edge] ragged n_grav={w.state.n_grav} n_dust={w.state.n_dust}, precision={precision}", flush=True)
    check_accel(engine_cls, w.state, w.params)
    check_steps(engine_cls, w.state, w.params, w.dt, 5)


def test_zero_steps_leaves_state_unchanged(engine_cls):
    w = debris(100, 300, FP32_FORCE)
    s = w.state.copy()
    with engine_cls() as eng:
        eng.upload(s, w.params)
        eng.step(w.dt, 0)
        eng.download(s)
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(s, k), getattr(w.state, k)), k


def test_context_reused_across_sizes(engine_cls):
    """One context, three uploads (grains, none, more grains and a larger
    swarm): every run equals a fresh context bit for bit."""
    cases = [debris(200, 500, FP32_FORCE), synth.config2(n_swarm=700, precision=FP32_FORCE),
             debris(900, 2000, FP32_FORCE)]

    def run(eng, w):
        s = w.state.copy()
        eng.upload(s, w.params)
        eng.step(w.dt, 6)
        return eng.download(s)

    with engine_cls() as eng:
        reused = [run(eng, w) for w in cases]
    for w, r in zip(cases, reused):
        with engine_cls() as eng:
            fresh = run(eng, w)
        for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
            assert np.array_equal(getattr(r, k), getattr(fresh, k)), k

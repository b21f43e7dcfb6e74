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


def test_unsoftened_fp32_force_close_pair_and_exact_hit(engine_cls):
    """The reference law (eps = 0) in the FP32-force form: forces and steps are
    refused without an FP64 radius (10 eps would be 0); with one, a near-coincident
    swarm pair (1e-5 AU apart, far below the FP32 resolution of positions at
    30-40 AU) is evaluated in FP64 and a grain placed exactly on a swarm body
    is removed, as the oracle does."""
    w = synth.config3(n_dust=500, n_swarm=200, precision=FP64, eps=0.0)
    st = w.state.copy()
    npl = w.params.n_planets
    j, k, h = npl + 3, npl + 4, npl + 9
    st.x[k], st.y[k], st.z[k] = st.x[j] + 1.0e-5, st.y[j], st.z[j] - 2.0e-6
    g0 = st.n_grav
    st.x[g0], st.y[g0], st.z[g0] = st.x[h], st.y[h], st.z[h]
    prm = replace(w.params, eps=0.0, fp64_radius=0.0, precision=FP32_FORCE)
    with engine_cls() as eng:
        eng.upload(st.copy(), prm)
        with pytest.raises(Exception, match="fp64_radius"):
            eng.accel()
        with pytest.raises(Exception, match="fp64_radius"):
            eng.step(w.dt, 1)
    prm = replace(prm, fp64_radius=0.05)
    with engine_cls() as eng:
        eng.upload(st.copy(), prm)
        a_gpu, rem_gpu = eng.accel()
        eng.step(w.dt, 1)
        s1 = eng.download(st.copy())
    a_ref, rem_ref = O.accel(prm, st)
    err = rel_err_rows(a_gpu, a_ref)
    print(f"[unsoftened fp32] close pair rel err {err[j]:.2e} {err[k]:.2e}; median {np.median(err):.2e}; "
          f"removed {int(rem_gpu.sum())}")
    assert np.array_equal(rem_gpu, rem_ref) and rem_gpu[0] == 1
    assert err[j] <= 1e-9 and err[k] <= 1e-9
    assert np.median(err) <= 1e-5
    ref1 = O.step(prm, st.copy(), w.dt, 1)
    assert np.array_equal(s1.N, ref1.N) and s1.N[g0] == 0.0


def test_rsqrt_fp64_accuracy(engine_cls):
    """The FP64 reciprocal square root of every DDYN_FP64_RSQ pair law and of
    the adjoint's Jacobians (v_rsq_f64 refined by one third-order step): at
    most 1.25 x 2^-53 relative error over 2^22 inputs log-uniform across 24
    decades (even powers of two among them), against 1/sqrt in 80-bit long
    double."""
    rng = np.random.default_rng(53)
    x = 10.0 ** rng.uniform(-12.0, 12.0, 1 << 22)
    x[:64] = 4.0 ** np.arange(-32, 32)
    with engine_cls() as eng:
        y = eng.rsqrt_fp64(x)
    xl = x.astype(np.longdouble)
    ref = np.longdouble(1) / np.sqrt(xl)
    rel = np.abs((y.astype(np.longdouble) - ref) / ref).astype(np.float64)
    ulp = 2.0 ** -53
    print(f"[rsqrt_fp64] max rel err {rel.max():.3e} = {rel.max() / ulp:.3f} x 2^-53, rms {np.sqrt(np.mean(rel ** 2)):.2e}")
    assert rel.max() <= 1.25 * ulp

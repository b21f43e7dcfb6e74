"""GPU parity: the HIP library (through the C ABI) against the FP64 CPU oracle.

Bars: plain FP64 form <= 1e-12 relative on per-particle accelerations and
<= 1e-10 relative on positions after the stated steps; FP32-force form
<= 1e-5 median relative acceleration error; the FP64 form with refined hardware
reciprocal square roots in every pair law (FP64_RSQ) is held to the plain FP64
bars.
"""
from dataclasses import replace

import numpy as np
import pytest

from diskdyn_amd import synth
from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ
from oracle import oracle as O
from tests.paritylib import pos_rel_err, rel_err_rows, subset, vel_rel_err

pytestmark = pytest.mark.gpu

SMALL = {
    "cfg1": lambda: synth.config1(n_dust=4096),
    "cfg2": lambda: synth.config2(n_swarm=4096, precision=FP64),
    "cfg3": lambda: synth.config3(n_dust=4096, n_swarm=1024, precision=FP64),
    "cfg4": lambda: synth.config4(n_dust=2048),
    "cfg5": lambda: synth.config5(n_swarm=8192, precision=FP64),
}


def _gpu_accel(engine_cls, w, precision):
    prm = w.params.with_precision(precision)
    with engine_cls() as eng:
        eng.upload(w.state.copy(), prm)
        return eng.accel()


@pytest.mark.parametrize("name", sorted(SMALL))
def test_accel_fp64(engine_cls, name):
    w = SMALL[name]()
    print(f"[accel_fp64] {name} n_grav={w.state.n_grav} n_dust={w.state.n_dust}", flush=True)
    a_gpu, rem_gpu = _gpu_accel(engine_cls, w, FP64)
    a_ref, rem_ref = O.accel(w.params, w.state)
    err = rel_err_rows(a_gpu, a_ref)
    print(f"  max rel {err.max():.3e} median {np.median(err):.3e}")
    assert np.array_equal(rem_gpu, rem_ref)
    assert err.max() <= 1e-12


@pytest.mark.parametrize("name", ["cfg2", "cfg3", "cfg5"])
def test_accel_fp64_rsq(engine_cls, name):
    """Refined hardware reciprocal square roots in every pair law: the plain FP64
    bar, and not the plain form's bits (the form reaches the kernels)."""
    w = SMALL[name]()
    print(f"[accel_fp64_rsq] {name}", flush=True)
    a_gpu, rem_gpu = _gpu_accel(engine_cls, w, FP64_RSQ)
    a_plain, _ = _gpu_accel(engine_cls, w, FP64)
    a_ref, rem_ref = O.accel(w.params, w.state)
    err = rel_err_rows(a_gpu, a_ref)
    print(f"  max rel {err.max():.3e} median {np.median(err):.3e}")
    assert np.array_equal(rem_gpu, rem_ref)
    assert err.max() <= 1e-12
    assert not np.array_equal(a_gpu, a_plain)


@pytest.mark.parametrize("name", ["cfg2", "cfg3", "cfg5"])
def test_accel_fp32(engine_cls, name):
    w = SMALL[name]()
    print(f"[accel_fp32] {name}", flush=True)
    a_gpu, rem_gpu = _gpu_accel(engine_cls, w, FP32_FORCE)
    a_ref, rem_ref = O.accel(w.params, w.state)
    err = rel_err_rows(a_gpu, a_ref)
    print(f"  max rel {err.max():.3e} median {np.median(err):.3e}")
    assert np.array_equal(rem_gpu, rem_ref)
    assert np.median(err) <= 1e-5
    assert err.max() <= 1e-3


@pytest.mark.parametrize("name,nsteps", [("cfg1", 100), ("cfg2", 10), ("cfg3", 20), ("cfg4", 40),
                                         ("cfg5", 4)])
def test_step_fp64(engine_cls, name, nsteps):
    w = SMALL[name]()
    print(f"[step_fp64] {name} steps={nsteps}", flush=True)
    st_gpu = w.state.copy()
    with engine_cls() as eng:
        eng.upload(st_gpu, w.params.with_precision(FP64))
        eng.step(w.dt, nsteps)
        eng.download(st_gpu)
    st_ref = O.step(w.params, w.state.copy(), w.dt, nsteps)
    ep = pos_rel_err(st_gpu, st_ref)
    ev = vel_rel_err(st_gpu, st_ref)
    print(f"  pos max {ep.max():.3e} vel max {ev.max():.3e}")
    assert np.array_equal(st_gpu.N, st_ref.N)
    assert ep.max() <= 1e-10
    assert ev.max() <= 1e-9


@pytest.mark.parametrize("name,nsteps", [("cfg1", 100), ("cfg2", 10), ("cfg3", 20), ("cfg4", 40), ("cfg5", 4)])
def test_step_fp64_rsq(engine_cls, name, nsteps):
    w = SMALL[name]()
    print(f"[step_fp64_rsq] {name} steps={nsteps}", flush=True)
    st_gpu = w.state.copy()
    with engine_cls() as eng:
        eng.upload(st_gpu, w.params.with_precision(FP64_RSQ))
        eng.step(w.dt, nsteps)
        eng.download(st_gpu)
    st_ref = O.step(w.params, w.state.copy(), w.dt, nsteps)
    ep = pos_rel_err(st_gpu, st_ref)
    ev = vel_rel_err(st_gpu, st_ref)
    print(f"  pos max {ep.max():.3e} vel max {ev.max():.3e}")
    assert np.array_equal(st_gpu.N, st_ref.N)
    assert ep.max() <= 1e-10
    assert ev.max() <= 1e-9


@pytest.mark.parametrize("name,nsteps", [("cfg2", 10), ("cfg3", 20)])
def test_step_fp32(engine_cls, name, nsteps):
    w = SMALL[name]()
    print(f"[step_fp32] {name} steps={nsteps}", flush=True)
    st_gpu = w.state.copy()
    with engine_cls() as eng:
        eng.upload(st_gpu, w.params.with_precision(FP32_FORCE))
        eng.step(w.dt, nsteps)
        eng.download(st_gpu)
    st_ref = O.step(w.params, w.state.copy(), w.dt, nsteps)
    ep = pos_rel_err(st_gpu, st_ref)
    print(f"  pos max {ep.max():.3e} median {np.median(ep):.3e}")
    assert np.median(ep) <= 1e-8


def _planets_only(n_bodies, n_dust=700):
    """Star, planet and n_bodies - 2 former swarm bodies, all unsoftened
    planets, with n_dust ring grains under radiation pressure, PR and
    stellar-wind drag."""
    w = synth.config3(n_dust=n_dust, n_swarm=max(n_bodies - 2, 1), precision=FP64)
    st = subset(w.state, range(n_bodies), range(n_dust))
    return replace(w, state=st, params=replace(w.params, n_planets=n_bodies))


FUSED = {
    "cfg1": SMALL["cfg1"],
    "cfg4": SMALL["cfg4"],
    "star_only": lambda: replace(SMALL["cfg1"](), state=subset(synth.config1(n_dust=512).state, [0], range(512)),
                                 params=replace(SMALL["cfg1"]().params, n_planets=1)),
    "bodies4": lambda: _planets_only(4),
    "bodies6": lambda: _planets_only(6),
    "bodies9": lambda: _planets_only(9),
    "bodies10": lambda: _planets_only(10),
    # 131,072 grains or more: the split form (k_small_massive, k_small_grains)
    "ring140k": lambda: synth.config4(n_dust=140_000),
    "bodies3_140k": lambda: _planets_only(3, 140_000),
    "bodies10_140k": lambda: _planets_only(10, 140_000),
}


@pytest.mark.parametrize("name,nsteps,precision", [("cfg1", 300, FP64), ("cfg4", 50, FP64), ("star_only", 60, FP64),
                                                   ("bodies4", 40, FP64), ("bodies6", 40, FP64), ("bodies9", 40, FP64),
                                                   ("bodies10", 40, FP64), ("cfg1", 300, FP64_RSQ),
                                                   ("cfg4", 50, FP64_RSQ), ("bodies9", 40, FP64_RSQ),
                                                   ("bodies10", 40, FP64_RSQ), ("ring140k", 30, FP64),
                                                   ("ring140k", 30, FP64_RSQ), ("bodies3_140k", 20, FP64),
                                                   ("bodies10_140k", 10, FP64_RSQ)])
def test_fused_small_steps_match_per_kernel_launches(engine_cls, name, nsteps, precision):
    """Star + planets only: the multi-step kernel equals one launch per kernel
    bit for bit, for 1 to 10 massive bodies (the kernel is specialised for 1-9),
    in the plain FP64 form and in FP64_RSQ; with 140,000 grains in its split
    form (the star and planets first, then the grains alone)."""
    w = FUSED[name]()
    st = w.state.copy()
    # one grain inside Rmin, one exactly on a planet: removal must agree too
    g0 = st.n_grav
    st.x[g0], st.y[g0], st.z[g0] = st.x[0] + 0.01, st.y[0], st.z[0]
    if st.n_grav > 1:
        st.x[g0 + 1], st.y[g0 + 1], st.z[g0 + 1] = st.x[1], st.y[1], st.z[1]
    prm = replace(w.params, Rmin=0.05, precision=precision)
    out = []
    for schedule in (0, 1):
        s = st.copy()
        with engine_cls() as eng:
            eng.upload(s, replace(prm, schedule=schedule))
            eng.step(w.dt, nsteps)
            eng.download(s)
        out.append(s)
    print(f"[fused] {name}: removed {int((out[0].N == 0).sum())}")
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(out[0], k), getattr(out[1], k)), k
    ref = O.step(prm, st.copy(), w.dt, nsteps)
    assert np.array_equal(out[0].N, ref.N)
    assert pos_rel_err(out[0], ref).max() <= 1e-10


@pytest.mark.parametrize("n_dust", [120_000, 400_000])
def test_fused_small_steps_more_workgroups_than_fit_at_once(engine_cls, n_dust):
    """120,000 test particles make 1,875 workgroups of the multi-step kernel,
    more than the GPU holds at once, so some start after others have finished:
    every one must still integrate the massive set from the call's initial
    state.  400,000 take the split form (from 131,072 grains): the star and
    planets in one wave, then 1,563 workgroups of grains against their
    recorded stage positions.  Single-step calls (the shortest launches) and a
    longer call, against one launch per kernel, bit for bit, and the planets
    against the oracle."""
    w = synth.config1(n_dust=n_dust)
    calls = [1, 1, 1, 7]
    out = []
    for schedule in (0, 1):
        s = w.state.copy()
        with engine_cls() as eng:
            eng.upload(s, replace(w.params, schedule=schedule))
            for n in calls:
                eng.step(w.dt, n)
            eng.download(s)
        out.append(s)
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(out[0], k), getattr(out[1], k)), k
    massive = subset(w.state, range(w.state.n_grav), [])
    ref = O.step(w.params, massive, w.dt, sum(calls))
    assert pos_rel_err(out[0], ref, slice(0, w.state.n_grav)).max() <= 1e-10


@pytest.mark.parametrize("n_dust", [3000, 140000])
@pytest.mark.parametrize("precision", [FP64, FP32_FORCE, FP64_RSQ])
def test_overlapped_grain_stream_matches_in_order_schedule(engine_cls, precision, n_dust):
    """Swarm + grains: the grain kernel of step k beside the massive stages of
    step k+1 (double-buffered stage positions; with 131,072 grains or more
    the grain blocks split over two streams) equals one launch per kernel in
    order, bit for bit, over two calls with grain removals at the first step
    and at the fifth, and the forces at the end state agree."""
    w = synth.config3(n_dust=n_dust, n_swarm=700, precision=precision)
    st = w.state.copy()
    g0 = st.n_grav
    st.x[g0], st.y[g0], st.z[g0] = st.x[0] + 0.05, st.y[0], st.z[0]   # sublimates (Rmin) at once
    last = st.n_all - 1                                               # the last block (second stream)
    st.x[last], st.y[last], st.z[last] = st.x[0], st.y[0] - 0.05, st.z[0]
    # at rest 1 AU from the star: falls inside Rmin after about five steps
    st.x[g0 + 1], st.y[g0 + 1], st.z[g0 + 1] = st.x[0] + 1.0, st.y[0], st.z[0]
    st.vx[g0 + 1], st.vy[g0 + 1], st.vz[g0 + 1] = st.vx[0], st.vy[0], st.vz[0]
    prm = replace(w.params, Rmin=0.1)
    out = []
    for schedule in (0, 1):
        s = st.copy()
        with engine_cls() as eng:
            eng.upload(s, replace(prm, schedule=schedule))
            eng.step(w.dt, 7)
            mid = eng.download(s.copy())
            eng.step(w.dt, 4)
            eng.download(s)
            a, rem = eng.accel()
        out.append((mid, s, a, rem))
    print(f"[overlap] precision={precision} n_dust={n_dust} removed {int((out[0][1].N == 0).sum())}")
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(out[0][0], k), getattr(out[1][0], k)), k
        assert np.array_equal(getattr(out[0][1], k), getattr(out[1][1], k)), k
    assert np.array_equal(out[0][2], out[1][2])
    assert np.array_equal(out[0][3], out[1][3])
    assert out[0][1].N[g0] == 0.0 and out[0][1].N[last] == 0.0
    assert out[0][0].N[g0 + 1] == 0.0 and out[0][0].x[g0 + 1] != st.x[g0 + 1]


def test_jacobi_integral_gpu(engine_cls):
    """Circular restricted three-body problem: the Jacobi integral E - n Lz of
    every test particle is conserved to 1e-10 over 1,000 steps on the GPU."""
    w = synth.config1_jacobi(n_dust=512, nsteps=1000)
    n = w.meta["mean_motion"]
    jacobi = lambda s: synth.jacobi_constants(s, n)  # noqa: E731

    st = w.state.copy()
    c0 = jacobi(st)
    with engine_cls() as eng:
        eng.upload(st, w.params)
        eng.step(w.dt, w.nsteps)
        eng.download(st)
    drift = np.abs((jacobi(st) - c0) / c0)
    print(f"[jacobi] max |dC/C| {drift.max():.3e} median {np.median(drift):.3e}")
    assert drift.max() < 1e-10


def test_energy(engine_cls):
    w = SMALL["cfg2"]()
    with engine_cls() as eng:
        eng.upload(w.state.copy(), w.params)
        e_gpu = eng.energy()
    e_ref = O.energy(w.params, w.state)
    print("energy", e_gpu[0], e_ref[0])
    assert abs(e_gpu[0] - e_ref[0]) <= 1e-12 * abs(e_ref[0])
    assert np.allclose(e_gpu[1:4], e_ref[1:4], rtol=1e-12, atol=0)

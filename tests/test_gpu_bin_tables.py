"""The reference's own grain-coefficient path on the GPU: float per-size-bin
beta, beta_sw and charge tables indexed by each grain's bin id
(src/gpu_functions.cu:655-665; the tables are __constant__ floats promoted to
double, src/ddyn.cu:30-32), with stellar-wind drag and the Lorentz force on.

Grains are spread over 40 bins, with some at id = -1 and some at ids past the
table (both: no bin, so beta = beta_sw = q = 0).  No per-grain beta is
uploaded, so every coefficient comes from the tables.  Checked against the
oracle given the same float tables: accelerations (plain FP64 and FP64_RSQ
<= 1e-12 relative, FP32-force median <= 1e-5) and 20 RK4 steps (positions
<= 1e-10 relative; FP32-force median <= 1e-8), on the swarm-and-grains
kernels and on the star-and-planets multi-step kernel.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

from diskdyn_amd import synth
from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ
from diskdyn_amd.units import GRAV, LIGHT_AU
from oracle import oracle as O
from tests.paritylib import pos_rel_err, rel_err_rows, subset, vel_rel_err

pytestmark = pytest.mark.gpu

N_BINS = 40


def binned_ring(n_dust=3000, n_swarm=300, planets_only=False):
    w = synth.config3(n_dust=n_dust, n_swarm=n_swarm, precision=FP64)
    st = w.state
    prm = replace(w.params, FLQ=1, Wz=3.0, Bstar=1.0e-2, Rstar=0.005)
    if planets_only:   # star + planet only: the multi-step kernel
        st = subset(st, range(w.params.n_planets), range(n_dust))
    ng = st.n_grav
    rng = np.random.default_rng(655)
    s_um = np.geomspace(1.0, 1000.0, N_BINS)
    beta_b = synth.grain_beta(s_um, lum=0.1, mstar=0.5)
    m_b = (4.0 / 3.0) * math.pi * (s_um * 1e-4) ** 3 * 2.5 * 1e-3 / 1.98892e30 * GRAV
    ids = rng.integers(-1, N_BINS + 6, n_dust).astype(np.int32)
    st.id[ng:] = ids
    ok = (ids >= 0) & (ids < N_BINS)
    st.mu[ng:] = np.where(ok, m_b[np.clip(ids, 0, N_BINS - 1)], m_b[0])
    st.beta_p = None
    st.betasw_p = None
    st.beta_t = beta_b.astype(np.float32)
    st.betasw_t = (0.35 * beta_b * prm.vsw / LIGHT_AU).astype(np.float32)
    # charge of a bin: a Lorentz acceleration of about 1e-3 of the star's pull at 35 AU
    st.q_t = (1e-3 * (GRAV * 0.5 / 35.0 ** 2) / (prm.Bstar * prm.Rstar ** 3 / 35.0 ** 3) * m_b).astype(np.float32)
    assert np.all(np.isfinite(st.q_t)) and st.q_t.min() > 0
    if planets_only:
        prm = replace(prm, n_planets=ng)
    return w, st, prm


@pytest.mark.parametrize("planets_only", [False, True])
@pytest.mark.parametrize("precision", [FP64, FP64_RSQ, FP32_FORCE])
def test_bin_tables_accel(engine_cls, precision, planets_only):
    w, st, prm = binned_ring(planets_only=planets_only)
    prm = prm.with_precision(precision)
    print(f"[bins] accel precision={precision} planets_only={planets_only} n_grav={st.n_grav}", flush=True)
    with engine_cls() as eng:
        eng.upload(st.copy(), prm)
        a_gpu, rem_gpu = eng.accel()
    a_ref, rem_ref = O.accel(prm, st)
    # the tables matter: zero tables change the grains' accelerations
    st0 = st.copy()
    st0.beta_t = np.zeros_like(st.beta_t)
    st0.betasw_t = np.zeros_like(st.betasw_t)
    st0.q_t = np.zeros_like(st.q_t)
    a0, _ = O.accel(prm, st0)
    ng = st.n_grav
    binned = ng + np.nonzero((st.id[ng:] >= 0) & (st.id[ng:] < N_BINS))[0]
    unbinned = ng + np.nonzero((st.id[ng:] < 0) | (st.id[ng:] >= N_BINS))[0]
    assert len(unbinned) > 0
    assert np.all(rel_err_rows(a0[binned], a_ref[binned]) > 1e-9)
    assert np.array_equal(a0[unbinned], a_ref[unbinned])
    err = rel_err_rows(a_gpu, a_ref)
    print(f"  max rel {err.max():.3e} median {np.median(err):.3e}")
    assert np.array_equal(rem_gpu, rem_ref)
    if precision == FP32_FORCE:
        assert np.median(err) <= 1e-5
    else:
        assert err.max() <= 1e-12


@pytest.mark.parametrize("planets_only", [False, True])
@pytest.mark.parametrize("precision", [FP64, FP64_RSQ, FP32_FORCE])
def test_bin_tables_steps(engine_cls, precision, planets_only):
    w, st, prm = binned_ring(planets_only=planets_only)
    prm = prm.with_precision(precision)
    nsteps = 20
    print(f"[bins] steps precision={precision} planets_only={planets_only}", flush=True)
    s = st.copy()
    with engine_cls() as eng:
        eng.upload(s, prm)
        eng.step(w.dt, nsteps)
        eng.download(s)
    ref = O.step(prm, st.copy(), w.dt, nsteps)
    ep, ev = pos_rel_err(s, ref), vel_rel_err(s, ref)
    print(f"  position max {ep.max():.3e} median {np.median(ep):.3e}")
    assert np.array_equal(s.N, ref.N)
    if precision == FP32_FORCE:
        assert np.median(ep) <= 1e-8
    else:
        assert ep.max() <= 1e-10
        assert ev.max() <= 1e-9

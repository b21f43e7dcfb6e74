"""The GPU against the long-double golden vectors (tests/golden): a 99-body
system with a star, two planets, a softened swarm and 64 grains under
radiation pressure, PR and stellar-wind drag and the Lorentz force.

The golden accelerations and RK4 step come from the textbook force law in
80-bit arithmetic (tests/golden/make_golden.py), independently of the oracle,
so this pins the GPU's grain forces, Lorentz term and RK4 directly.
"""
from dataclasses import replace

import numpy as np
import pytest

from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ
from oracle import oracle as O
from tests.paritylib import golden_state, pos_rel_err, rel_err_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE, FP64_RSQ])
def test_gpu_golden_accelerations(engine_cls, precision):
    st, prm, g = golden_state()
    with engine_cls() as eng:
        eng.upload(st.copy(), replace(prm, precision=precision))
        a, rem = eng.accel()
    assert not rem.any()
    err = rel_err_rows(a, g["acc"])
    print(f"[golden] precision={precision} accel max rel {err.max():.2e} median {np.median(err):.2e}")
    if precision in (FP64, FP64_RSQ):
        assert err.max() <= 1e-13
    else:
        assert np.median(err) <= 1e-5
        assert err.max() <= 1e-3


@pytest.mark.parametrize("schedule,precision", [(0, FP64), (1, FP64), (0, FP64_RSQ), (1, FP64_RSQ)])
def test_gpu_golden_rk4_step(engine_cls, schedule, precision):
    st, prm, g = golden_state()
    s = st.copy()
    with engine_cls() as eng:
        eng.upload(s, replace(prm, schedule=schedule, precision=precision))
        eng.step(float(g["step_dt"][0]), 1)
        eng.download(s)
    ep = rel_err_rows(s.pos(), g["pos1"])
    ev = rel_err_rows(s.vel(), g["vel1"])
    print(f"[golden] one RK4 step: pos max {ep.max():.2e} vel max {ev.max():.2e}")
    assert ep.max() <= 1e-13 and ev.max() <= 1e-13


def test_gpu_lorentz_wind_drag_steps_match_oracle(engine_cls):
    """50 steps of the golden system (Lorentz force and wind drag on) against
    the oracle, plain FP64."""
    st, prm, g = golden_state()
    dt = float(g["step_dt"][0])
    s = st.copy()
    with engine_cls() as eng:
        eng.upload(s, prm)
        eng.step(dt, 50)
        eng.download(s)
    ref = O.step(prm, st.copy(), dt, 50)
    assert np.array_equal(s.N, ref.N)
    ep = pos_rel_err(s, ref)
    print(f"[golden] 50 steps vs oracle: pos max {ep.max():.2e}")
    assert ep.max() <= 1e-10

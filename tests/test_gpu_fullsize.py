"""Full-size GPU checks through size-independent properties.

At the benchmark sizes the FP64 oracle is too slow to step, so these tests use
what the dynamics fixes whatever the size:

- total momentum of the massive set is conserved (each pair force is summed
  from both sides, so only rounding breaks it);
- total energy and angular momentum drift only by the RK4 truncation error;
- the FP32-force and plain FP64 forms stay together (the FP32 error is a small
  perturbation of the swarm term, far below the star's pull), and the FP64 form
  with refined hardware reciprocal square roots stays with the plain one to
  rounding.
"""
import numpy as np
import pytest

from diskdyn_amd import synth
from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ
from tests.paritylib import pos_rel_err

pytestmark = pytest.mark.gpu


def run(engine_cls, w, precision, nsteps):
    st = w.state.copy()
    with engine_cls() as eng:
        eng.upload(st, w.params.with_precision(precision))
        e0 = eng.energy()
        eng.step(w.dt, nsteps)
        e1 = eng.energy()
        eng.download(st)
    return st, e0, e1


def momentum_scale(st):
    ng = st.n_grav
    return float(np.sum(st.mu[:ng] * np.sqrt(st.vx[:ng] ** 2 + st.vy[:ng] ** 2 + st.vz[:ng] ** 2)))


@pytest.mark.parametrize("precision,bound", [(FP64, 1e-13), (FP32_FORCE, 1e-9), (FP64_RSQ, 1e-13)])
def test_swarm_65537_conservation(engine_cls, precision, bound):
    w = synth.config2(precision=precision)
    print(f"[fullsize] swarm n_grav={w.state.n_grav} precision={precision}", flush=True)
    st, e0, e1 = run(engine_cls, w, precision, 10)
    scale = momentum_scale(w.state)
    dP = np.abs(e1[4:7] - e0[4:7]).max() / scale
    dE = abs((e1[0] - e0[0]) / e0[0])
    Lz = np.linalg.norm(e0[1:4])
    dL = np.linalg.norm(e1[1:4] - e0[1:4]) / Lz
    print(f"  |dP|/sum|mu v| {dP:.2e}  |dE/E| {dE:.2e}  |dL|/|L| {dL:.2e}")
    assert dP <= bound
    assert dE <= 1e-9
    assert dL <= 1e-10


def test_debris_ring_forms_agree(engine_cls):
    """262,144 grains + 1,026 massive bodies, 20 steps in all three forms."""
    w = synth.config3(precision=FP64)
    print(f"[fullsize] debris ring n_dust={w.state.n_dust}", flush=True)
    s64, e0, e1 = run(engine_cls, w, FP64, 20)
    s32, _, _ = run(engine_cls, w, FP32_FORCE, 20)
    srq, _, _ = run(engine_cls, w, FP64_RSQ, 20)
    assert np.array_equal(s64.N, s32.N) and np.array_equal(s64.N, srq.N)
    erq = pos_rel_err(srq, s64)
    print(f"  FP64_RSQ vs FP64 positions: max {erq.max():.2e}")
    assert erq.max() <= 1e-12
    err = pos_rel_err(s32, s64)
    print(f"  FP32-force vs FP64 positions: median {np.median(err):.2e} max {err.max():.2e}; "
          f"|dE/E| {abs((e1[0] - e0[0]) / e0[0]):.2e}")
    assert np.median(err) <= 1e-9
    assert err.max() <= 1e-6
    assert abs((e1[0] - e0[0]) / e0[0]) <= 1e-10


def test_disk_1048577_forms_agree(engine_cls):
    """One RK4 step of the 1,048,577-body disk in all three forms."""
    w = synth.config5(precision=FP64)
    print(f"[fullsize] disk n_grav={w.state.n_grav}", flush=True)
    s64, e0, e1 = run(engine_cls, w, FP64, 1)
    s32, _, _ = run(engine_cls, w, FP32_FORCE, 1)
    srq, _, _ = run(engine_cls, w, FP64_RSQ, 1)
    erq = pos_rel_err(srq, s64)
    print(f"  FP64_RSQ vs FP64 positions: max {erq.max():.2e}")
    assert erq.max() <= 1e-12
    err = pos_rel_err(s32, s64)
    dP = np.abs(e1[4:7] - e0[4:7]).max() / momentum_scale(w.state)
    print(f"  FP32-force vs FP64 positions: median {np.median(err):.2e} max {err.max():.2e}; FP64 |dP| {dP:.2e}")
    assert np.median(err) <= 1e-9
    assert err.max() <= 1e-6
    assert dP <= 1e-13

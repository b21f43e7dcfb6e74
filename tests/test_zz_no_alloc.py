"""ddyn_accel_noalloc (static device scratch, no allocation) against ddyn_accel.

The allocation-free variant takes the caller's device arrays; its stage-0
positions and slice partials live in the library's static scratch array.  Two
calls in a row on one context, the second with a smaller system (and the bin
tables instead of per-grain beta), must give ddyn_accel's accelerations and
removal flags bit for bit (same pair split); a system large enough that the
scratch caps the slice count must agree to summation order (1e-13 relative,
plain FP64); a system beyond the scratch is refused with DDYN_ERR_ARG.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from diskdyn_amd import engine, synth
from diskdyn_amd.system import FP32_FORCE, FP64
from tests.paritylib import rel_err_rows

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda:0", dtype=dtype)


def _noalloc(eng, st, use_bins):
    n = st.n_all
    f64, i32 = torch.float64, torch.int32
    t = {k: _dev(getattr(st, k), f64) for k in ("x", "y", "z", "vx", "vy", "vz", "mu")}
    t["id"] = _dev(st.id, i32)
    bp = None if use_bins else _dev(st.beta_p, f64)
    bsw = None if use_bins else _dev(st.betasw_p, f64)
    out = {k: torch.full((n,), np.nan, dtype=f64, device="cuda:0") for k in ("ax", "ay", "az")}
    rem = torch.full((max(st.n_dust, 1),), -7, dtype=i32, device="cuda:0")
    torch.cuda.synchronize()
    p = lambda v: None if v is None else v.data_ptr()  # noqa: E731
    eng.accel_noalloc(st.n_grav, st.n_dust, *[p(t[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "mu", "id")],
                      p(bp), p(bsw), p(out["ax"]), p(out["ay"]), p(out["az"]), p(rem))
    eng.synchronize()
    a = torch.stack([out["ax"], out["ay"], out["az"]], dim=1).cpu().numpy()
    return a, rem.cpu().numpy()[: st.n_dust]


def _binned(st, n_bins=30):
    """Per-bin float tables in place of per-grain beta (ids over the bins, some -1)."""
    s = st.copy()
    ng = s.n_grav
    rng = np.random.default_rng(8)
    ids = rng.integers(-1, n_bins, s.n_dust).astype(np.int32)
    s.id[ng:] = ids
    s.beta_t = np.linspace(0.01, 0.4, n_bins).astype(np.float32)
    s.betasw_t = (0.002 * s.beta_t).astype(np.float32)
    s.q_t = np.zeros(n_bins, dtype=np.float32)
    s.beta_p = None
    s.betasw_p = None
    return s


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE])
def test_noalloc_accel_equals_accel_twice_then_smaller(engine_cls, precision):
    big = synth.config3(n_dust=3000, n_swarm=700, precision=precision)
    small = synth.config3(n_dust=900, n_swarm=300, precision=precision, seed=77)
    cases = [(big.state, False), (_binned(small.state), True)]
    with engine_cls() as eng:
        for st, use_bins in cases:
            print(f"[noalloc] precision={precision} n_grav={st.n_grav} n_dust={st.n_dust} bins={use_bins}",
                  flush=True)
            prm = replace(big.params, precision=precision)
            eng.upload(st.copy(), prm)
            a_ref, rem_ref = eng.accel()
            a, rem = _noalloc(eng, st, use_bins)
            assert np.array_equal(rem, rem_ref)
            assert np.array_equal(a, a_ref)


def test_noalloc_capped_slices_and_refusal(engine_cls):
    """200,001 bodies: the scratch holds 4 slices of partials where the
    context plans more; the sums differ in order only.  Then a swarm larger
    than the scratch is refused with a message."""
    w = synth.config2(n_swarm=200_000, precision=FP64)
    st = w.state
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        nslice_ctx = eng.pair_plan()[0]
        a_ref, _ = eng.accel()
        a, _ = _noalloc(eng, st, use_bins=True)
        err = rel_err_rows(a, a_ref)
        print(f"[noalloc] 200,001 bodies: context slices {nslice_ctx}, max rel diff {err.max():.2e}")
        assert nslice_ctx > 4
        assert err.max() <= 1e-13
        n = 2_000_000
        with pytest.raises(engine.DDynError, match="static scratch"):
            eng.accel_noalloc(n, 0, *([1] * 7), None, None, None, 1, 1, 1, None)

"""Adjoint of the large disk's 10-step window at full size (1,048,577 bodies).

L = mean over swarm bodies of (|r_i(T)| - |r_i^Kep(T)|)^2 after 10 RK4 steps
of dt = 0.02 yr, one tape segment.  Checked here:

- the FP32-force form's gradient (FP32 pair forces, FP32 VJP tiles with FP64
  flushes) against the gradient of the FP64 form with refined hardware
  reciprocal square roots (FP64_RSQ) of the same window: <= 1e-4 relative,
  norm-wise in r and in v.  Along the random direction the bar is 1e-2: the
  loss measures deviations of about 2e-8 AU, while the FP32-force forward
  pass moves the few bodies that have a pair just outside the FP64 radius
  (10 eps) by up to about 4e-7 AU, so their loss gradients differ between the
  forms (the FP32 VJP tiles themselves agree with the FP64 adjoint to 1e-6 on
  the same trajectory, test_gpu_adjoint_fp32_tiles);
- the same with a wider FP64 radius (100 eps): the pairs that FP32 still
  evaluates are ten times farther apart, their relative force error ten
  times smaller, and the directional derivative agrees with FP64_RSQ to
  1e-3;
- the FP64_RSQ directional derivative against central finite differences of
  the FP64_RSQ forward map (h = 2.5e-7 AU and AU/yr on every coordinate):
  <= 1e-4 relative.  The step is small against the softening length
  (5e-4 AU), so the differences' truncation error stays far below the bar,
  and large enough that the rounding of the two forward runs (about 1e-14 AU
  per body, independent between the runs) moves the difference by about 1e-6
  of the signal.

The work is split over several tests that share one lazily computed record
(each test computes what it needs if it runs alone), so that no single test
runs much longer than a minute.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from diskdyn_amd import synth
from diskdyn_amd.adjoint import kepler_deviation_loss, kepler_targets
from diskdyn_amd.system import FP32_FORCE, FP64_RSQ

pytestmark = pytest.mark.gpu

NSTEPS = 10
FD_STEP = 2.5e-7
FP32_WIDE = "fp32-force, FP64 radius 100 eps"


class _Window:
    def __init__(self):
        self.w = synth.config5(precision=FP64_RSQ)
        st = self.w.state
        self.np_ = self.w.params.n_planets
        self.target = kepler_targets(st, self.np_, NSTEPS * self.w.dt)
        rng = np.random.default_rng(3)
        self.dr = rng.normal(size=(st.n_all, 3))
        self.dv = rng.normal(size=(st.n_all, 3))
        self.grad = {}
        self.lam = {}
        self.eng = {}

    def loss_and_lam(self, rT):
        r = torch.from_numpy(rT).requires_grad_(True)
        L = kepler_deviation_loss(r, self.target, self.np_)
        L.backward()
        return L.item(), r.grad.numpy().copy()

    def params(self, key):
        if key == FP32_WIDE:
            return replace(self.w.params.with_precision(FP32_FORCE), fp64_radius=100 * self.w.params.eps)
        return self.w.params.with_precision(key)

    def record(self, engine_cls, precision):
        if precision in self.lam or precision in self.grad:
            return
        eng = engine_cls()
        eng.upload(self.w.state.copy(), self.params(precision))
        eng.adjoint_record(self.w.dt, NSTEPS, NSTEPS)
        fin = eng.download(self.w.state.copy())
        self.lam[precision] = self.loss_and_lam(fin.pos())
        self.eng[precision] = eng

    def gradient(self, engine_cls, precision):
        if precision not in self.grad:
            self.record(engine_cls, precision)
            L, lam_r = self.lam[precision]
            eng = self.eng.pop(precision)
            try:
                g_r, g_v, _, _ = eng.adjoint_reverse(lam_r, np.zeros_like(lam_r))
            finally:
                eng.close()
            self.grad[precision] = (L, g_r, g_v)
        return self.grad[precision]

    def directional(self, engine_cls, precision):
        _, g_r, g_v = self.gradient(engine_cls, precision)
        return float((g_r * self.dr).sum() + (g_v * self.dv).sum())

    def loss_at(self, engine_cls, sign, h=FD_STEP):
        st = self.w.state.copy()
        p, v = st.pos() + sign * h * self.dr, st.vel() + sign * h * self.dv
        st.x, st.y, st.z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        st.vx, st.vy, st.vz = (np.ascontiguousarray(v[:, k]) for k in range(3))
        with engine_cls() as eng:
            eng.upload(st, self.w.params.with_precision(FP64_RSQ))
            eng.step(self.w.dt, NSTEPS)
            fin = eng.download(st.copy())
        r = torch.from_numpy(fin.pos())
        return kepler_deviation_loss(r, self.target, self.np_).item()


_W = {}


@pytest.fixture(scope="module")
def window():
    if "w" not in _W:
        _W["w"] = _Window()
    yield _W["w"]
    for eng in _W["w"].eng.values():
        eng.close()
    _W["w"].eng.clear()


def test_disk_window_fp64_rsq_record(engine_cls, window):
    print(f"[disk window] n_grav={window.w.state.n_grav}: FP64_RSQ forward with tape", flush=True)
    window.record(engine_cls, FP64_RSQ)
    L, lam_r = window.lam[FP64_RSQ]
    print(f"  loss {L:.6e}", flush=True)
    assert np.isfinite(L) and np.all(np.isfinite(lam_r))


def test_disk_window_fp64_rsq_reverse(engine_cls, window):
    L, g_r, g_v = window.gradient(engine_cls, FP64_RSQ)
    print(f"[disk window] FP64_RSQ loss {L:.6e}, |g_r| {np.linalg.norm(g_r):.3e}", flush=True)
    assert np.all(np.isfinite(g_r)) and np.all(np.isfinite(g_v))


def test_disk_window_fp32_gradient_matches_fp64_rsq(engine_cls, window):
    _, g64r, g64v = window.gradient(engine_cls, FP64_RSQ)
    _, g32r, g32v = window.gradient(engine_cls, FP32_FORCE)
    er, ev = np.linalg.norm(g32r - g64r) / np.linalg.norm(g64r), np.linalg.norm(g32v - g64v) / np.linalg.norm(g64v)
    d64, d32 = window.directional(engine_cls, FP64_RSQ), window.directional(engine_cls, FP32_FORCE)
    ed = abs(d32 - d64) / abs(d64)
    print(f"[disk window] FP32-force vs FP64_RSQ gradient: r {er:.2e}, v {ev:.2e}, directional {ed:.2e}")
    assert er <= 1e-4 and ev <= 1e-4
    assert ed <= 1e-2


def test_disk_window_fp32_wide_fp64_radius(engine_cls, window):
    _, g64r, g64v = window.gradient(engine_cls, FP64_RSQ)
    _, gr, gv = window.gradient(engine_cls, FP32_WIDE)
    er, ev = np.linalg.norm(gr - g64r) / np.linalg.norm(g64r), np.linalg.norm(gv - g64v) / np.linalg.norm(g64v)
    d64, d = window.directional(engine_cls, FP64_RSQ), window.directional(engine_cls, FP32_WIDE)
    ed = abs(d - d64) / abs(d64)
    print(f"[disk window] FP32-force, FP64 radius 100 eps, vs FP64_RSQ gradient: r {er:.2e}, v {ev:.2e}, "
          f"directional {ed:.2e}")
    assert er <= 1e-4 and ev <= 1e-4
    assert ed <= 1e-3


def test_disk_window_fp64_rsq_finite_difference_plus(engine_cls, window):
    window.lp = window.loss_at(engine_cls, +1.0)
    assert np.isfinite(window.lp)


def test_disk_window_fp64_rsq_finite_difference(engine_cls, window):
    lp = getattr(window, "lp", None)
    if lp is None:
        lp = window.loss_at(engine_cls, +1.0)
    lm = window.loss_at(engine_cls, -1.0)
    fd = (lp - lm) / (2 * FD_STEP)
    an = window.directional(engine_cls, FP64_RSQ)
    print(f"[disk window] FP64_RSQ directional: adjoint {an:.10e} fd {fd:.10e} rel {abs(an - fd) / abs(fd):.2e}")
    assert abs(an - fd) <= 1e-4 * abs(fd)

"""GPU discrete adjoint (ddyn_adjoint) against the FP64 CPU adjoint and central
finite differences; planet (m, a, e) gradients through the torch wrapper.

Bars: GPU adjoint == CPU adjoint to 1e-9 relative (same algebra, different
summation order); gradients within 1e-4 relative of central finite
differences with h = 1e-6 relative.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from diskdyn_amd import engine, synth
from diskdyn_amd.adjoint import Runner, Trajectory, planet_initial_state, ring_loss
from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "ring": (lambda: synth.config4(n_dust=256), 20, 6),
    "swarm": (lambda: synth.config2(n_swarm=300, precision=FP64, eps=1e-2), 10, 4),
    "debris": (lambda: synth.config3(n_dust=200, n_swarm=24, precision=FP64, eps=2e-2), 8, 3),
    # more than 64 massive bodies: the grain pulls run as a pair kernel (k_adj_pull)
    "debris_many": (lambda: synth.config3(n_dust=700, n_swarm=160, precision=FP64, eps=2e-2), 6, 3),
}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_adjoint_matches_cpu_adjoint(engine_cls, name):
    make, nsteps, ck = CASES[name]
    w = make()
    st = w.state
    print(f"[adjoint] {name} n_grav={st.n_grav} n_dust={st.n_dust} steps={nsteps} ckpt={ck}", flush=True)
    rng = np.random.default_rng(11)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        g_r, g_v, g_mu, g_beta = eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=ck)
        # the context state ends at the forward final state
        fin = eng.download(st.copy())
    c_r, c_v, c_mu, c_beta = O.adjoint(w.params, st, w.dt, nsteps, lam_r, lam_v)
    ref = O.step(w.params, st.copy(), w.dt, nsteps)
    errs = {"r": _rel(g_r, c_r), "v": _rel(g_v, c_v), "mu": _rel(g_mu, c_mu)}
    if st.n_dust:
        errs["beta"] = _rel(g_beta, c_beta)
    print("  ", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= 1e-9
    assert np.abs(fin.pos() - ref.pos()).max() <= 1e-10 * np.abs(ref.pos()).max()


@pytest.mark.parametrize("name", ["debris", "debris_many", "swarm"])
def test_gpu_adjoint_fp64_rsq_context(engine_cls, name):
    """FP64_RSQ context: the forward pass takes the refined-rsqrt pair
    laws; the gradients still meet the FP64 bar against the CPU adjoint."""
    make, nsteps, ck = CASES[name]
    w = make()
    st = w.state
    rng = np.random.default_rng(13)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params.with_precision(FP64_RSQ))
        g_r, g_v, g_mu, g_beta = eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=ck)
    c_r, c_v, c_mu, c_beta = O.adjoint(w.params, st, w.dt, nsteps, lam_r, lam_v)
    errs = {"r": _rel(g_r, c_r), "v": _rel(g_v, c_v), "mu": _rel(g_mu, c_mu)}
    if st.n_dust:
        errs["beta"] = _rel(g_beta, c_beta)
    print(f"[adjoint fp64_rsq] {name}", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= 1e-9


@pytest.mark.parametrize("precision", [FP64, FP32_FORCE])
def test_grain_pull_pairs_match_wave_sums(engine_cls, precision):
    """The debris ring's 1,026 massive bodies with 2,048 grains: the grain
    pulls on the massive bodies as a pair kernel over LDS tiles of grains
    (default schedule) against the per-wave sums of the one-launch-per-kernel
    schedule and against the CPU adjoint."""
    w = synth.config3(n_dust=2048, precision=precision)
    st = w.state
    nsteps, ck = 2, 1
    rng = np.random.default_rng(17)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    out = {}
    for sched in (0, 1):
        with engine_cls() as eng:
            eng.upload(st.copy(), replace(w.params, schedule=sched))
            out[sched] = eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=ck)
    d = {k: _rel(out[0][i], out[1][i]) for i, k in enumerate(("r", "v", "mu", "beta"))}
    print(f"[grain pulls] pairs vs wave sums, precision={precision}", {k: f"{v:.2e}" for k, v in d.items()})
    assert max(d.values()) <= 1e-12
    if precision == FP64:
        c = O.adjoint(w.params, st, w.dt, nsteps, lam_r, lam_v)
        e = {k: _rel(out[0][i], c[i]) for i, k in enumerate(("r", "v", "mu", "beta"))}
        print("   against the CPU adjoint", {k: f"{v:.2e}" for k, v in e.items()})
        assert max(e.values()) <= 1e-9


FUSED_REVERSE = {
    "swarm250": (lambda: synth.config2(n_swarm=250, precision=FP64, eps=1e-2), 6, 4),
    "debris10": (lambda: synth.config3(n_dust=300, n_swarm=8, precision=FP64, eps=2e-2), 8, 3),
}


@pytest.mark.parametrize("name,precision", [("ring", FP64), ("debris", FP64), ("debris", FP32_FORCE),
                                            ("debris", FP64_RSQ), ("swarm250", FP32_FORCE), ("swarm250", FP64),
                                            ("debris10", FP64), ("debris10", FP32_FORCE)])
def test_fused_small_massive_reverse_equals_per_stage_launches(engine_cls, name, precision):
    """At most 256 massive bodies: the single-workgroup massive reverse step
    (schedule 0) gives the gradients of the per-stage launches (schedule 1)
    bit for bit, with and without grains, in every precision form."""
    make, nsteps, ck = FUSED_REVERSE[name] if name in FUSED_REVERSE else CASES[name]
    w = make()
    st = w.state
    rng = np.random.default_rng(14)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    out = []
    for schedule in (0, 1):
        with engine_cls() as eng:
            eng.upload(st.copy(), replace(w.params, precision=precision, schedule=schedule))
            out.append(eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=ck))
    print(f"[fused reverse] {name} precision={precision} n_grav={st.n_grav} n_dust={st.n_dust}", flush=True)
    for a, b in zip(out[0], out[1]):
        assert (a is None and b is None) or np.array_equal(a, b)


def test_gpu_adjoint_fp32_tiles(engine_cls):
    """FP32-force context: FP32 VJP tiles (FP64 flush, FP64 close pairs) against the FP64 CPU adjoint."""
    from diskdyn_amd.system import FP32_FORCE
    w = synth.config2(n_swarm=2000, precision=FP32_FORCE, eps=1e-2)
    st = w.state
    nsteps = 6
    rng = np.random.default_rng(12)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        g_r, g_v, g_mu, _ = eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=3)
    c_r, c_v, c_mu, _ = O.adjoint(w.params.with_precision(FP64), st, w.dt, nsteps, lam_r, lam_v)
    errs = {"r": _rel(g_r, c_r), "v": _rel(g_v, c_v), "mu": _rel(g_mu, c_mu)}
    print("fp32 tiles:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= 1e-5


def test_segment_chained_adjoint_equals_single_call():
    """Runner with a progress report chains ddyn_adjoint window by window; the
    gradients equal one ddyn_adjoint call over all steps."""
    w = synth.config3(n_dust=300, n_swarm=16, precision=FP64, eps=2e-2)
    st = w.state
    rng = np.random.default_rng(21)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    msgs = []
    a = Runner(st, w.params, w.dt, 9, ckpt_every=4).adjoint(st.copy(), lam_r, lam_v)
    b = Runner(st, w.params, w.dt, 9, ckpt_every=4, report=msgs.append).adjoint(st.copy(), lam_r, lam_v)
    assert len(msgs) == 4   # forward + 3 windows
    for x, y in zip(a, b):
        assert _rel(np.asarray(y), np.asarray(x)) <= 1e-12
    # window-start states kept from the forward run: no second forward pass
    run = Runner(st, w.params, w.dt, 9, ckpt_every=4, report=msgs.append)
    _, ckpts = run.forward_with_checkpoints(st.copy())
    assert len(ckpts) == 3
    c = run.adjoint(st.copy(), lam_r, lam_v, ckpts)
    for x, y in zip(a, c):
        assert _rel(np.asarray(y), np.asarray(x)) <= 1e-12


@pytest.mark.parametrize("precision,ck", [(FP64, 4), (FP64, 0), (FP32_FORCE, 4)])
def test_recorded_tape_reverse_equals_single_call(engine_cls, precision, ck):
    """ddyn_adjoint_record + ddyn_adjoint_reverse give what ddyn_adjoint gives,
    bit for bit: in one segment (ck = 0) and in three (the last on the tape,
    the others recomputed); a second reverse with another lam (the last
    segment then recomputed) too.  The recorded run ends where ddyn_step ends,
    and a step drops the tape."""
    w = synth.config3(n_dust=300, n_swarm=40, precision=precision, eps=2e-2)
    st = w.state
    nsteps = 9
    rng = np.random.default_rng(31)
    lams = [(rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))) for _ in range(2)]
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        eng.adjoint_record(w.dt, nsteps, ckpt_every=ck)
        end_rec = eng.download(st.copy())
        taped = [eng.adjoint_reverse(*lam) for lam in lams]
        after = eng.download(st.copy())
        eng.step(w.dt, 1)
        with pytest.raises(engine.DDynError, match="tape"):
            eng.adjoint_reverse(*lams[0])
    for lam, t in zip(lams, taped):
        with engine_cls() as eng:
            eng.upload(st.copy(), w.params)
            ref = eng.adjoint(w.dt, nsteps, *lam, ckpt_every=ck)
        for x, y in zip(ref, t):
            assert np.array_equal(np.asarray(x), np.asarray(y))
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        eng.step(w.dt, nsteps)
        end_step = eng.download(st.copy())
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(end_rec, k), getattr(end_step, k)), k
        assert np.array_equal(getattr(after, k), getattr(end_step, k)), k


def test_trajectory_tape_mode_equals_windowed():
    """The torch wrapper with the tape (one forward pass) gives the gradients of
    the windowed Runner (forward, then every window recomputed)."""
    w = synth.config3(n_dust=200, n_swarm=24, precision=FP64, eps=2e-2)
    st = w.state
    beta = torch.from_numpy(st.beta_p.copy())
    rng = np.random.default_rng(41)
    wr = torch.from_numpy(rng.normal(size=(st.n_all, 3)))
    grads = []
    for kw in ({"tape": True}, {"report": lambda m: None}):
        run = Runner(st, w.params, w.dt, 8, ckpt_every=3, **kw)
        r0 = torch.from_numpy(st.pos()).requires_grad_(True)
        v0 = torch.from_numpy(st.vel()).requires_grad_(True)
        mu = torch.from_numpy(st.mu[: st.n_grav].copy()).requires_grad_(True)
        rT, vT = Trajectory.apply(r0, v0, mu, beta, run)
        ((rT * wr).sum() + (vT * wr).sum()).backward()
        grads.append((r0.grad.numpy(), v0.grad.numpy(), mu.grad.numpy()))
    for x, y in zip(*grads):
        assert _rel(x, y) <= 1e-12


def test_gpu_adjoint_directional_fd(engine_cls):
    w = synth.config4(n_dust=128)
    st = w.state
    nsteps = 30
    rng = np.random.default_rng(5)
    lam_r, lam_v = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))
    dr, dv = rng.normal(size=(st.n_all, 3)), rng.normal(size=(st.n_all, 3))

    def L(s0):
        with engine_cls() as eng:
            eng.upload(s0, w.params)
            eng.step(w.dt, nsteps)
            s = eng.download(s0.copy())
        return float((lam_r * s.pos()).sum() + (lam_v * s.vel()).sum())

    def moved(h):
        s = st.copy()
        s.x, s.y, s.z = (np.ascontiguousarray(s.pos()[:, k] + h * dr[:, k]) for k in range(3))
        s.vx, s.vy, s.vz = (np.ascontiguousarray(s.vel()[:, k] + h * dv[:, k]) for k in range(3))
        return s

    h = 1e-6
    fd = (L(moved(h)) - L(moved(-h))) / (2 * h)
    with engine_cls() as eng:
        eng.upload(st.copy(), w.params)
        g_r, g_v, _, _ = eng.adjoint(w.dt, nsteps, lam_r, lam_v, ckpt_every=7)
    an = float((g_r * dr).sum() + (g_v * dv).sum())
    print(f"directional: adjoint {an:.12e} fd {fd:.12e}")
    assert abs(an - fd) <= 1e-6 * abs(fd)


def test_swarm_window_adjoint_vs_finite_differences():
    """Config 5 shape at reduced N: 10-step window, checkpoints every 5,
    L = mean (|r(T)| - |r_Kepler(T)|)^2; directional derivative in the initial
    state and partials in two masses against central finite differences."""
    from diskdyn_amd.adjoint import kepler_deviation_loss, kepler_targets
    w = synth.config5(n_swarm=2048, precision=FP64, eps=5e-3)
    st = w.state
    nsteps, ck = 10, 5
    run = Runner(st, w.params, w.dt, nsteps, ckpt_every=ck)
    target = kepler_targets(st, w.params.n_planets, nsteps * w.dt)
    ng = st.n_grav

    def L_of(r0, v0, mu):
        rT, _ = Trajectory.apply(r0, v0, mu, torch.zeros(0, dtype=torch.float64), run)
        return kepler_deviation_loss(rT, target, w.params.n_planets)

    r0 = torch.from_numpy(st.pos()).requires_grad_(True)
    v0 = torch.from_numpy(st.vel()).requires_grad_(True)
    mu = torch.from_numpy(st.mu[:ng].copy()).requires_grad_(True)
    L = L_of(r0, v0, mu)
    L.backward()
    rng = np.random.default_rng(9)
    dr, dv = rng.normal(size=r0.shape), rng.normal(size=v0.shape)
    an = float((r0.grad.numpy() * dr).sum() + (v0.grad.numpy() * dv).sum())
    h = 1e-6
    with torch.no_grad():
        Lp = L_of(r0 + h * torch.from_numpy(dr), v0 + h * torch.from_numpy(dv), mu).item()
        Lm = L_of(r0 - h * torch.from_numpy(dr), v0 - h * torch.from_numpy(dv), mu).item()
    fd = (Lp - Lm) / (2 * h)
    print(f"swarm window: directional adjoint {an:.10e} fd {fd:.10e} rel {abs(an - fd) / abs(fd):.2e}")
    assert abs(an - fd) <= 1e-4 * abs(fd)
    # star mass: partial derivative; swarm masses (each far below FD resolution
    # on its own): directional derivative along dmu_j = mu_j u_j, u ~ N(0, 1)
    u = np.zeros(ng)
    u[1:] = rng.normal(size=ng - 1)
    dirs = {"mu[0]": (np.eye(ng)[0] * float(mu.detach()[0]), 1e-6), "swarm": (u * mu.detach().numpy(), 1e-3)}
    for name, (d, h) in dirs.items():
        dmu = torch.from_numpy(d)
        with torch.no_grad():
            fdm = (L_of(r0, v0, mu.detach() + h * dmu).item() - L_of(r0, v0, mu.detach() - h * dmu).item()) / (2 * h)
        g = float((mu.grad * dmu).sum())
        print(f"  dL along {name}: adjoint {g:.10e} fd {fdm:.10e}")
        assert abs(g - fdm) <= 1e-4 * abs(fdm)


def test_planet_gradients_vs_finite_differences():
    """Reduced adjoint system (config 4 shape): dL/d(m, a, e) of the planet and
    dL/dbeta of grains, L = mean |r(T) - r_target(T)|^2."""
    n_dust, nsteps, ck = 512, 40, 10
    base = synth.config4(n_dust=n_dust)
    tgt = synth.config4(n_dust=n_dust, planet_m=1.2e-4, planet_a=26.0)
    run = Runner(base.state, base.params, base.dt, nsteps, ckpt_every=ck)
    target = torch.from_numpy(run.forward(tgt.state.copy()).pos())
    ng = base.state.n_grav

    def loss_of(m, a, e, beta):
        r_m, v_m, mu = planet_initial_state(1.0, m, a, e)
        r0 = torch.cat([r_m, torch.from_numpy(base.state.pos()[ng:])])
        v0 = torch.cat([v_m, torch.from_numpy(base.state.vel()[ng:])])
        rT, _ = Trajectory.apply(r0, v0, mu, beta, run)
        return ring_loss(rT, target, ng)

    m = torch.tensor(1e-4, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(25.0, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(0.15, dtype=torch.float64, requires_grad=True)
    beta = torch.from_numpy(base.state.beta_p.copy()).requires_grad_(True)
    L = loss_of(m, a, e, beta)
    L.backward()
    grads = {"m": m.grad.item(), "a": a.grad.item(), "e": e.grad.item()}
    with torch.no_grad():
        for name, x0 in (("m", 1e-4), ("a", 25.0), ("e", 0.15)):
            h = 1e-6 * x0
            args = {"m": 1e-4, "a": 25.0, "e": 0.15}
            args[name] = x0 + h
            Lp = loss_of(torch.tensor(args["m"], dtype=torch.float64), torch.tensor(args["a"], dtype=torch.float64),
                         torch.tensor(args["e"], dtype=torch.float64), beta.detach()).item()
            args[name] = x0 - h
            Lm = loss_of(torch.tensor(args["m"], dtype=torch.float64), torch.tensor(args["a"], dtype=torch.float64),
                         torch.tensor(args["e"], dtype=torch.float64), beta.detach()).item()
            fd = (Lp - Lm) / (2 * h)
            print(f"d L / d {name}: adjoint {grads[name]:.10e} fd {fd:.10e} rel {abs(grads[name] - fd) / abs(fd):.2e}")
            assert abs(grads[name] - fd) <= 1e-4 * abs(fd)
        gb = beta.grad.numpy()
        for k in (0, 7, 300):
            h = 1e-6 * float(beta[k])
            bp, bm = beta.detach().clone(), beta.detach().clone()
            bp[k] += h
            bm[k] -= h
            fd = (loss_of(m.detach(), a.detach(), e.detach(), bp).item() -
                  loss_of(m.detach(), a.detach(), e.detach(), bm).item()) / (2 * h)
            print(f"d L / d beta[{k}]: adjoint {gb[k]:.10e} fd {fd:.10e}")
            assert abs(gb[k] - fd) <= 1e-4 * abs(fd)

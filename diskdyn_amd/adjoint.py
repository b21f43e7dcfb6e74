"""Gradients of trajectory losses through the GPU discrete adjoint.

`Trajectory.apply(r0, v0, mu, beta)` integrates nsteps RK4 steps on the GPU
(ddyn_step) and returns the final positions and velocities; its backward pass
calls ddyn_adjoint (checkpoints every `ckpt_every` steps, recomputation in
between).  Inputs are float64 CPU tensors: r0, v0 (n_all, 3), mu (n_grav,)
the massive G*M, beta (n_dust,) the per-grain radiation-pressure ratio.

`planet_initial_state` rebuilds the star + planet initial state from planet
(m, a, e) with the algebra of coe2rv (src/functions.cu:6-46) and recalc_COM
(src/read_in.cu:344-375) written in torch, so dL/d(m, a, e) follows by
autograd from the adjoint's dL/d(initial state) and dL/dmu.
"""
import time

import numpy as np
import torch

from .engine import Engine
from .units import GRAV


class Runner:
    """Fixed parts of a run: the template state (ids, N, beta_sw, grain masses),
    parameters, step size, step count and checkpoint interval.

    With tape=True a forward pass that will be differentiated records the
    adjoint's tape on the device (ddyn_adjoint_record) and keeps its context;
    the reverse pass then needs no second forward pass."""

    def __init__(self, state, params, dt, nsteps, ckpt_every=0, device=0, report=None, tape=False):
        self.state, self.params = state, params
        self.dt, self.nsteps, self.ckpt_every, self.device = dt, nsteps, ckpt_every, device
        self.report = report
        self.tape = tape

    def make_state(self, r0, v0, mu, beta):
        st = self.state.copy()
        ng = st.n_grav
        r = r0.detach().cpu().numpy()
        v = v0.detach().cpu().numpy()
        st.x, st.y, st.z = (np.ascontiguousarray(r[:, k]) for k in range(3))
        st.vx, st.vy, st.vz = (np.ascontiguousarray(v[:, k]) for k in range(3))
        st.mu[:ng] = mu.detach().cpu().numpy()
        if st.n_dust:
            st.beta_p = np.ascontiguousarray(beta.detach().cpu().numpy(), dtype=np.float64)
            if st.betasw_p is None:
                st.betasw_p = np.zeros(st.n_dust)
        return st

    @property
    def windowed(self):
        """True when the reverse sweep is chained window by window (report set)."""
        return self.report is not None and 0 < self.ckpt_every < self.nsteps

    def forward(self, st):
        return self.forward_with_checkpoints(st)[0]

    def forward_with_checkpoints(self, st, record=False):
        """Final state, and what the reverse pass can reuse: with tape and
        record, the Engine holding the recorded tape (the caller hands it to
        adjoint(), which closes it); in windowed mode the host states at the
        window starts; None otherwise."""
        if self.tape and record:
            t0 = time.perf_counter()
            eng = Engine(self.device)
            try:
                eng.upload(st, self.params)
                eng.adjoint_record(self.dt, self.nsteps, self.ckpt_every)
                out = eng.download(st.copy())
            except BaseException:
                eng.close()
                raise
            if self.report is not None:
                self.report(f"forward {self.nsteps} steps, adjoint tape recorded ({time.perf_counter() - t0:.1f} s)")
            return out, eng
        ckpts = None
        with Engine(self.device) as eng:
            eng.upload(st, self.params)
            if self.report is None or not self.ckpt_every:
                eng.step(self.dt, self.nsteps)
            else:
                ckpts = [] if self.windowed else None
                t0 = time.perf_counter()
                for k0 in range(0, self.nsteps, self.ckpt_every):
                    if ckpts is not None:
                        ckpts.append(eng.download(st.copy()))
                    eng.step(self.dt, min(self.ckpt_every, self.nsteps - k0))
                    eng.synchronize()
                    self.report(f"forward step {min(k0 + self.ckpt_every, self.nsteps)}/{self.nsteps} "
                                f"({time.perf_counter() - t0:.1f} s)")
            out = eng.download(st.copy())
        return out, ckpts

    def adjoint(self, st, lam_r, lam_v, ckpts=None):
        """dL/d(initial r, v), dL/dmu and dL/dbeta for dL/d(final r, v) = lam_r, lam_v.

        With `report` set, the reverse sweep is chained segment by segment
        (ddyn_adjoint on each checkpoint window, the adjoint state passed from
        one window to the previous) and `report` is called after each window.
        `ckpts` are the window-start states from forward_with_checkpoints;
        without them the forward pass runs again to make them.  If `ckpts` is
        the Engine of a recorded forward pass (tape mode), the gradients come
        from its tape and the Engine is closed."""
        if isinstance(ckpts, Engine):
            t0 = time.perf_counter()
            try:
                g = ckpts.adjoint_reverse(lam_r, lam_v)
            finally:
                ckpts.close()
            if self.report is not None:
                self.report(f"adjoint of {self.nsteps} steps from the tape ({time.perf_counter() - t0:.1f} s)")
            return g
        if not self.windowed:
            with Engine(self.device) as eng:
                eng.upload(st, self.params)
                return eng.adjoint(self.dt, self.nsteps, lam_r, lam_v, self.ckpt_every)
        ck = self.ckpt_every
        starts = list(range(0, self.nsteps, ck))
        t0 = time.perf_counter()
        with Engine(self.device) as eng:
            if ckpts is None:
                eng.upload(st, self.params)
                ckpts = []
                for k0 in starts:                      # forward pass, host checkpoints
                    ckpts.append(eng.download(st.copy()))
                    eng.step(self.dt, min(ck, self.nsteps - k0))
                self.report(f"forward {self.nsteps} steps done ({time.perf_counter() - t0:.1f} s)")
            g_mu = np.zeros(st.n_grav)
            g_beta = np.zeros(st.n_dust)
            for i in range(len(starts) - 1, -1, -1):
                n = min(ck, self.nsteps - starts[i])
                eng.upload(ckpts[i], self.params)
                lam_r, lam_v, gm, gb = eng.adjoint(self.dt, n, lam_r, lam_v, n)
                g_mu += gm
                g_beta += gb
                self.report(f"adjoint window {len(starts) - i}/{len(starts)} done ({time.perf_counter() - t0:.1f} s)")
        return lam_r, lam_v, g_mu, g_beta


class Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r0, v0, mu, beta, runner):
        st = runner.make_state(r0, v0, mu, beta)
        out, ckpts = runner.forward_with_checkpoints(st, record=any(ctx.needs_input_grad[:4]))
        ctx.runner = runner
        ctx.state = st
        ctx.ckpts = ckpts
        rT = torch.from_numpy(out.pos()).to(r0.dtype)
        vT = torch.from_numpy(out.vel()).to(v0.dtype)
        return rT, vT

    @staticmethod
    def backward(ctx, g_rT, g_vT):
        lam_r = np.zeros(ctx.state.pos().shape) if g_rT is None else g_rT.detach().cpu().numpy()
        lam_v = np.zeros(ctx.state.pos().shape) if g_vT is None else g_vT.detach().cpu().numpy()
        g_r, g_v, g_mu, g_beta = ctx.runner.adjoint(ctx.state, lam_r, lam_v, ctx.ckpts)
        return (torch.from_numpy(g_r), torch.from_numpy(g_v), torch.from_numpy(g_mu),
                torch.from_numpy(np.ascontiguousarray(g_beta)), None)


def coe2rv_t(M, m, a, e, inc, w, o, nu, beta=0.0):
    """Torch restatement of coe2rv (src/functions.cu:6-46); returns (r, v) 3-vectors."""
    p = a * (1.0 - e * e)
    mu = (M + m) * (1.0 - beta)
    cnu, snu = torch.cos(nu), torch.sin(nu)
    xe = p * cnu / (1.0 + e * cnu)
    ye = p * snu / (1.0 + e * cnu)
    sq = torch.sqrt(mu / p)
    vxe = -sq * snu
    vye = sq * (e + cnu)
    co, so, cw, sw, ci, si = (torch.cos(o), torch.sin(o), torch.cos(w), torch.sin(w), torch.cos(inc),
                              torch.sin(inc))
    r11, r12 = co * cw - so * sw * ci, -co * sw - so * cw * ci
    r21, r22 = so * cw + co * sw * ci, -so * sw + co * cw * ci
    r31, r32 = sw * si, cw * si
    r = torch.stack([xe * r11 + ye * r12, xe * r21 + ye * r22, xe * r31 + ye * r32])
    v = torch.stack([vxe * r11 + vye * r12, vxe * r21 + vye * r22, vxe * r31 + vye * r32])
    return r, v


def planet_initial_state(mstar, m_p, a_p, e_p, inc=0.0, w=0.0, o=0.0, nu=0.0):
    """Star + one planet after recalc_COM, as torch tensors (2, 3) and mu (2,).

    mstar, m_p in Msun (m_p, a_p, e_p may require grad)."""
    dt = torch.float64
    f = lambda v: v if torch.is_tensor(v) else torch.tensor(float(v), dtype=dt)
    m_p, a_p, e_p = f(m_p), f(a_p), f(e_p)
    mu0 = torch.tensor(mstar * GRAV, dtype=dt)
    mu1 = m_p * GRAV
    rp, vp = coe2rv_t(mu0, mu1, a_p, e_p, f(inc), f(w), f(o), f(nu))
    mu = torch.stack([mu0, mu1])
    r = torch.stack([torch.zeros(3, dtype=dt), rp])
    v = torch.stack([torch.zeros(3, dtype=dt), vp])
    # recalc_COM: shift to the barycentre, star velocity zeroes the momentum
    com = (mu[:, None] * r).sum(0) / mu.sum()
    r = r - com
    v0 = -(mu1 * vp) / mu0
    v = torch.stack([v0, vp])
    return r, v, mu


def ring_loss(rT, target, n_grav):
    """Mean squared distance of the final grain positions from a target."""
    d = rT[n_grav:] - target[n_grav:]
    return (d * d).sum(1).mean()


def kepler_targets(state, n_planets, t):
    """Final positions of the swarm bodies on their unperturbed two-body orbits
    around the star (mu = mu_star), as a fixed (detached) target."""
    from .orbits import kepler_propagate_many
    r0 = state.pos()[n_planets:] - state.pos()[0]
    v0 = state.vel()[n_planets:] - state.vel()[0]
    r, _ = kepler_propagate_many(state.mu[0], r0, v0, t)
    out = state.pos().copy()
    out[n_planets:] = r + state.pos()[0]
    return torch.from_numpy(out)


def kepler_deviation_loss(rT, target, n_planets):
    """mean over swarm bodies of (|r_i(T)| - |r_i^Kep(T)|)^2."""
    d = torch.linalg.norm(rT[n_planets:], dim=1) - torch.linalg.norm(target[n_planets:], dim=1)
    return (d * d).mean()


__all__ = ["Runner", "Trajectory", "coe2rv_t", "planet_initial_state", "ring_loss", "kepler_targets",
           "kepler_deviation_loss"]

"""Writes the golden vectors in tests/golden/*.npz (run: python tests/golden/make_golden.py).

The reference ships no expected outputs for its dynamics, so these vectors come
from an independent long-double (80-bit) evaluation of the published physics:

  * gravity: a_i = sum_j mu_j (r_j - r_i) / (|r_j - r_i|^2 + eps_ij^2)^(3/2),
    eps_ij = eps only between swarm bodies (and grain-swarm pairs);
  * grain star term (Burns, Lamy & Soter 1979, Icarus 40, 1):
    a = -(1 - beta - beta_sw) mu0 d/R^3 - (mu0/R^2)(beta/c + beta_sw/v_sw)((v.dhat) dhat + v)
    plus the Lorentz force (q/m) (v - Omega x d) x B, B = B* R*^3/R^3 zhat;
  * one classical RK4 step (Butcher tableau 0, 1/2, 1/2, 1; weights 1/6, 1/3,
    1/3, 1/6) of the whole first-order system y' = f(y), massive bodies and
    grains together.

Inputs are stored as float64 and the long-double results rounded to float64.
"""
import os

import numpy as np

LD = np.longdouble
HERE = os.path.dirname(os.path.abspath(__file__))
C_LIGHT = 63239.6717367


def system(seed=2024, n_planets=3, n_swarm=32, n_dust=64):
    rng = np.random.Generator(np.random.Philox(seed))
    ng = n_planets + n_swarm
    pos = np.zeros((ng + n_dust, 3))
    vel = np.zeros((ng + n_dust, 3))
    mu = np.zeros(ng + n_dust)
    mu[0] = 4 * np.pi ** 2 * 0.8
    pos[0] = rng.normal(0, 0.01, 3)
    vel[0] = rng.normal(0, 0.001, 3)
    for k in range(1, ng + n_dust):
        r = rng.uniform(2.0, 12.0)
        th = rng.uniform(0, 2 * np.pi)
        pos[k] = [r * np.cos(th), r * np.sin(th), rng.normal(0, 0.1)]
        v = np.sqrt(mu[0] / r)
        vel[k] = [-v * np.sin(th), v * np.cos(th), rng.normal(0, 0.02)] + rng.normal(0, 0.05, 3)
    mu[1:n_planets] = rng.uniform(1e-4, 1e-3, n_planets - 1) * 4 * np.pi ** 2
    mu[n_planets:ng] = rng.uniform(1e-9, 1e-6, n_swarm) * 4 * np.pi ** 2
    mu[ng:] = rng.uniform(1e-20, 1e-18, n_dust)
    # two swarm bodies close together so the softening matters
    pos[n_planets + 1] = pos[n_planets] + [0.03, -0.02, 0.01]
    beta = rng.uniform(0.0, 0.4, n_dust)
    betasw = 0.35 * beta * 94.93 / C_LIGHT
    q = rng.uniform(-1e-11, 1e-11, n_dust).astype(np.float32).astype(np.float64)  # float tables
    return dict(pos=pos, vel=vel, mu=mu, n_planets=n_planets, n_grav=ng, beta=beta, betasw=betasw, q=q,
                eps=0.05, vsw=94.93, Wz=3.0, Bstar=1.0e-2, Rstar=0.005, Rmin=0.3, light=C_LIGHT)


def grav_ld(pos, mu, n_planets, n_grav, eps, targets):
    """Long-double gravity from the massive bodies on the given targets."""
    P = pos.astype(LD)
    M = mu.astype(LD)
    e2 = LD(eps) * LD(eps)
    out = np.zeros((len(targets), 3), dtype=LD)
    for k, i in enumerate(targets):
        acc = np.zeros(3, dtype=LD)
        for j in range(n_grav):
            if j == i:
                continue
            d = P[j] - P[i]
            soft = (i >= n_planets) and (j >= n_planets)
            r2 = d @ d + (e2 if soft else LD(0))
            if r2 == 0:
                continue
            acc += M[j] * d / (r2 * np.sqrt(r2))
        out[k] = acc
    return out


def grain_star_ld(S, pos, vel, g_idx):
    """Star term (gravity, radiation, PR + wind drag, Lorentz) on grains, long double."""
    P, V = pos.astype(LD), vel.astype(LD)
    m0 = LD(S["mu"][0])
    c, vsw = LD(S["light"]), LD(S["vsw"])
    out = np.zeros((len(g_idx), 3), dtype=LD)
    for k, t in enumerate(g_idx):
        g = t - S["n_grav"]
        b, bsw, q = LD(S["beta"][g]), LD(S["betasw"][g]), LD(S["q"][g])
        d = P[t] - P[0]
        v = V[t]
        R2 = d @ d
        R = np.sqrt(R2)
        dhat = d / R
        a = -(1 - b - bsw) * m0 * d / (R2 * R)
        a -= m0 / R2 * (b / c + bsw / vsw) * ((v @ dhat) * dhat + v)
        Bz = LD(S["Bstar"]) * LD(S["Rstar"]) ** 3 / (R2 * R)
        Om = np.array([0, 0, LD(S["Wz"])], dtype=LD)
        vrel = v - np.cross(Om, d)
        a += q / LD(S["mu"][t]) * np.cross(vrel, np.array([0, 0, Bz], dtype=LD))
        out[k] = a
    return out


def deriv_ld(S, pos, vel):
    """y' = f(y) of the full system: massive gravity; grains: star term + massive pulls."""
    ng = S["n_grav"]
    n = len(S["mu"])
    acc = np.zeros((n, 3), dtype=LD)
    acc[:ng] = grav_ld(pos, S["mu"], S["n_planets"], ng, S["eps"], range(ng))
    gi = list(range(ng, n))
    acc[ng:] = grain_star_ld(S, pos, vel, gi)
    P = pos.astype(LD)
    M = S["mu"].astype(LD)
    e2 = LD(S["eps"]) ** 2
    for t in gi:
        for j in range(1, ng):
            d = P[j] - P[t]
            r2 = d @ d + (e2 if j >= S["n_planets"] else LD(0))
            acc[t] += M[j] * d / (r2 * np.sqrt(r2))
    return vel.astype(LD), acc


def rk4_ld(S, pos, vel, h):
    h = LD(h)
    P0, V0 = pos.astype(LD), vel.astype(LD)
    k1r, k1v = deriv_ld(S, P0, V0)
    k2r, k2v = deriv_ld(S, P0 + h / 2 * k1r, V0 + h / 2 * k1v)
    k3r, k3v = deriv_ld(S, P0 + h / 2 * k2r, V0 + h / 2 * k2v)
    k4r, k4v = deriv_ld(S, P0 + h * k3r, V0 + h * k3v)
    P1 = P0 + h / 6 * (k1r + 2 * k2r + 2 * k3r + k4r)
    V1 = V0 + h / 6 * (k1v + 2 * k2v + 2 * k3v + k4v)
    return P1, V1


def main():
    S = system()
    n = len(S["mu"])
    _, acc = deriv_ld(S, S["pos"], S["vel"])
    P1, V1 = rk4_ld(S, S["pos"], S["vel"], 0.01)
    meta = {k: S[k] for k in ("n_planets", "n_grav", "eps", "vsw", "Wz", "Bstar", "Rstar", "Rmin", "light")}
    np.savez(os.path.join(HERE, "golden_system.npz"), pos=S["pos"], vel=S["vel"], mu=S["mu"],
             beta=S["beta"], betasw=S["betasw"], q=S["q"],
             meta=np.array([float(meta[k]) for k in ("n_planets", "n_grav", "eps", "vsw", "Wz", "Bstar",
                                                       "Rstar", "Rmin", "light")]),
             acc=acc.astype(np.float64), step_dt=np.array([0.01]),
             pos1=P1.astype(np.float64), vel1=V1.astype(np.float64))
    print("wrote golden_system.npz with", n, "particles")


if __name__ == "__main__":
    main()

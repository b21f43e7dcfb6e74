"""Shared error measures and the golden-vector system for the parity tests."""
import os

import numpy as np

from diskdyn_amd.system import DiskState, Params, make_state


def rel_err_rows(a, b, floor=0.0):
    """Per-row |a - b| / |b| (norm-wise over the 3 components)."""
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(b, axis=1), floor)
    return num / np.where(den > 0, den, 1.0)


def pos_rel_err(st_a, st_b, sl=slice(None)):
    pa = np.stack([st_a.x[sl], st_a.y[sl], st_a.z[sl]], axis=1)
    pb = np.stack([st_b.x[sl], st_b.y[sl], st_b.z[sl]], axis=1)
    return rel_err_rows(pa, pb)


def vel_rel_err(st_a, st_b, sl=slice(None)):
    pa = np.stack([st_a.vx[sl], st_a.vy[sl], st_a.vz[sl]], axis=1)
    pb = np.stack([st_b.vx[sl], st_b.vy[sl], st_b.vz[sl]], axis=1)
    return rel_err_rows(pa, pb)


def subset(st, grav, dust):
    """State with the massive rows `grav` and the grain rows `dust` of st."""
    rows = np.concatenate([np.asarray(grav, dtype=np.int64), st.n_grav + np.asarray(dust, dtype=np.int64)])
    pick = lambda a: np.ascontiguousarray(a[rows])  # noqa: E731
    pd = lambda a: None if a is None else np.ascontiguousarray(a[np.asarray(dust, dtype=np.int64)])  # noqa: E731
    return DiskState(len(grav), len(dust), pick(st.x), pick(st.y), pick(st.z), pick(st.vx), pick(st.vy),
                     pick(st.vz), pick(st.mu), pick(st.N), pick(st.id), pd(st.beta_p), pd(st.betasw_p),
                     st.beta_t, st.betasw_t, st.q_t)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_system.npz")


def golden_state():
    g = np.load(GOLDEN, allow_pickle=False)
    n_planets, n_grav = int(g["meta"][0]), int(g["meta"][1])
    eps, vsw, Wz, Bstar, Rstar, Rmin, light = (float(v) for v in g["meta"][2:])
    n = len(g["mu"])
    st = make_state(n_grav, n - n_grav)
    st.x, st.y, st.z = (np.ascontiguousarray(g["pos"][:, k]) for k in range(3))
    st.vx, st.vy, st.vz = (np.ascontiguousarray(g["vel"][:, k]) for k in range(3))
    st.mu = np.ascontiguousarray(g["mu"])
    st.id = np.full(n, -1, dtype=np.int32)
    st.id[n_grav:] = np.arange(n - n_grav, dtype=np.int32)
    st.beta_p = np.ascontiguousarray(g["beta"])
    st.betasw_p = np.ascontiguousarray(g["betasw"])
    st.beta_t = np.zeros(n - n_grav, dtype=np.float32)
    st.betasw_t = np.zeros(n - n_grav, dtype=np.float32)
    st.q_t = g["q"].astype(np.float32)
    prm = Params(n_planets=n_planets, SWQ=1, FLQ=1, eps=eps, Rmin=Rmin, vsw=vsw, Wz=Wz, Bstar=Bstar,
                 Rstar=Rstar, light=light)
    return st, prm, g

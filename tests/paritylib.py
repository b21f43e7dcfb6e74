"""Shared error measures for the parity tests."""
import numpy as np


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

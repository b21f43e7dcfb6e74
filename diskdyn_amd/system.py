"""Host-side particle state and physics parameters, in the reference layout.

`DiskState` holds the SoA arrays that the reference keeps in Grid's host
coord (src/ddyn.cuh:44-58, filled by add_to_host in src/grid.cu:106-189):
order [star | planets | swarm | dust], positions in AU, velocities in AU/yr,
mu = G*M, N the tracer count (0 = removed) and id the grain size-bin index.
`Params` mirrors the C ABI's ddyn_params (include/ddyn.h).
"""
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from .units import LIGHT_AU

FP64 = 0          # every term in FP64, IEEE square roots and divisions
FP32_FORCE = 1    # swarm-sourced pairs in FP32 tiles; everything else FP64
FP64_RSQ = 2      # FP64; every pair law uses refined hardware reciprocal square roots


@dataclass
class Params:
    n_planets: int                 # star + planets (cfg->n_planets counts the star)
    SWQ: int = 0                   # stellar-wind drag on
    FLQ: int = 0                   # Lorentz force on
    precision: int = FP64          # FP64, FP32_FORCE or FP64_RSQ
    eps: float = 0.0               # Plummer softening [AU] of swarm-sourced pairs
    Rmin: float = 0.0              # sublimation radius [AU]
    vsw: float = 0.0               # stellar-wind speed [AU/yr]
    Wz: float = 0.0                # stellar rotation [1/yr]
    Bstar: float = 0.0             # stellar field (reference units)
    Rstar: float = 0.0             # stellar radius [AU]
    light: float = LIGHT_AU        # speed of light [AU/yr]
    fp64_radius: float = 0.0       # FP32_FORCE: FP64 pair radius [AU] (0 -> 10*eps)
    schedule: int = 0              # 0: fused multi-step launches, overlapped grain stream; 1: one per kernel, in order

    def with_precision(self, precision):
        return replace(self, precision=precision)


@dataclass
class DiskState:
    n_grav: int
    n_dust: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    mu: np.ndarray
    N: np.ndarray
    id: np.ndarray
    beta_p: Optional[np.ndarray] = None      # per-grain beta (n_dust), overrides bin tables
    betasw_p: Optional[np.ndarray] = None    # per-grain beta_sw (n_dust)
    beta_t: Optional[np.ndarray] = None      # per-bin float tables (src/ddyn.cu:30-32)
    betasw_t: Optional[np.ndarray] = None
    q_t: Optional[np.ndarray] = None

    @property
    def n_all(self):
        return self.n_grav + self.n_dust

    def pos(self):
        return np.stack([self.x, self.y, self.z], axis=1)

    def vel(self):
        return np.stack([self.vx, self.vy, self.vz], axis=1)

    def copy(self):
        cp = lambda a: None if a is None else np.array(a, copy=True)
        return DiskState(self.n_grav, self.n_dust, cp(self.x), cp(self.y), cp(self.z),
                         cp(self.vx), cp(self.vy), cp(self.vz), cp(self.mu), cp(self.N),
                         cp(self.id), cp(self.beta_p), cp(self.betasw_p), cp(self.beta_t),
                         cp(self.betasw_t), cp(self.q_t))

    def validate(self):
        n = self.n_all
        for name in ("x", "y", "z", "vx", "vy", "vz", "mu", "N"):
            a = getattr(self, name)
            if a.dtype != np.float64 or a.shape != (n,) or not a.flags.c_contiguous:
                raise ValueError(f"{name} must be a contiguous float64 array of length {n}")
        if self.id.dtype != np.int32 or self.id.shape != (n,):
            raise ValueError(f"id must be an int32 array of length {n}")
        for name in ("beta_p", "betasw_p"):
            a = getattr(self, name)
            if a is not None and (a.dtype != np.float64 or a.shape != (self.n_dust,)):
                raise ValueError(f"{name} must be a float64 array of length {self.n_dust}")


def make_state(n_grav, n_dust):
    n = n_grav + n_dust
    z = lambda: np.zeros(n, dtype=np.float64)
    return DiskState(n_grav, n_dust, z(), z(), z(), z(), z(), z(), z(), np.ones(n),
                     np.full(n, -1, dtype=np.int32))


@dataclass
class Workload:
    """A synthetic run: initial state, physics, step size and step count."""
    name: str
    state: DiskState
    params: Params
    dt: float
    nsteps: int
    seed: int
    meta: dict = field(default_factory=dict)

    def interactions_per_eval(self):
        """Directed pair interactions of one force evaluation:
        n_grav (n_grav - 1) + n_dust * n_grav."""
        g, d = self.state.n_grav, self.state.n_dust
        return g * (g - 1) + d * g

"""ctypes binding of the C ABI in include/ddyn.h (libddyn_hip.so).

The library is built in-tree by diskdyn_amd.build (hipcc --offload-arch=gfx950).
There is no CPU fallback: if the library or a GPU is missing, constructing an
Engine raises.
"""
import ctypes
import os

import numpy as np

from .imaging import COpticsTables, CView, OpticsTables, View
from .system import DiskState, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddyn_hip.so")

DDYN_KCLASS_PAIRS, DDYN_KCLASS_STAGE, DDYN_KCLASS_DUST = 0, 1, 2
KCLASS_NAMES = {0: "pairs", 1: "stage", 2: "dust"}

# Every entry point declared in include/ddyn.h, with its ctypes signature.
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_F = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p


class DDynParams(ctypes.Structure):
    _fields_ = [("n_planets", ctypes.c_int32), ("SWQ", ctypes.c_int32), ("FLQ", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("eps", ctypes.c_double), ("Rmin", ctypes.c_double),
                ("vsw", ctypes.c_double), ("Wz", ctypes.c_double), ("Bstar", ctypes.c_double),
                ("Rstar", ctypes.c_double), ("light", ctypes.c_double),
                ("fp64_radius", ctypes.c_double), ("schedule", ctypes.c_int32), ("reserved", ctypes.c_int32)]


SIGNATURES = {
    "ddyn_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int]),
    "ddyn_destroy": (None, [_VP]),
    "ddyn_last_error": (ctypes.c_char_p, [_VP]),
    "ddyn_set_params": (ctypes.c_int, [_VP, ctypes.POINTER(DDynParams)]),
    "ddyn_upload_state": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int64] + [_D] * 8 + [_I32]),
    "ddyn_upload_bin_tables": (ctypes.c_int, [_VP, ctypes.c_int32, _F, _F, _F]),
    "ddyn_upload_grain_beta": (ctypes.c_int, [_VP, _D, _D]),
    "ddyn_step": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int64]),
    "ddyn_accel": (ctypes.c_int, [_VP, _D, _D, _D, _I32]),
    "ddyn_accel_noalloc": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int64] + [_VP] * 14),
    "ddyn_download_state": (ctypes.c_int, [_VP] + [_D] * 7),
    "ddyn_energy": (ctypes.c_int, [_VP, _D]),
    "ddyn_adjoint": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int64, ctypes.c_int64] + [_D] * 6),
    "ddyn_adjoint_record": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int64, ctypes.c_int64]),
    "ddyn_adjoint_reverse": (ctypes.c_int, [_VP] + [_D] * 6),
    "ddyn_synchronize": (ctypes.c_int, [_VP]),
    "ddyn_set_timing": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ddyn_read_timing": (ctypes.c_int, [_VP, ctypes.c_int32, _I64, _D]),
    "ddyn_pair_plan": (ctypes.c_int, [_VP, _I32, _I32]),
    "ddyn_eval_rsqrt_fp64": (ctypes.c_int, [_VP, ctypes.c_int64, _D, _D]),
    "ddyn_upload_optics": (ctypes.c_int, [_VP, ctypes.POINTER(COpticsTables)]),
    "ddyn_calc_flux": (ctypes.c_int, [_VP, ctypes.POINTER(CView), ctypes.c_int32, _D]),
    "ddyn_calc_sed": (ctypes.c_int, [_VP, ctypes.POINTER(CView), _D]),
    "ddyn_mie": (ctypes.c_int, [_VP, ctypes.c_int32, _D, ctypes.c_int32, _D, _D, _D, ctypes.c_int32, _D, _D, _D,
                                _D]),
}

_lib = None


class DDynError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libddyn_hip.so and declare every C ABI function (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DDynError(f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32)


def _fp(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def to_c_params(p: Params) -> DDynParams:
    return DDynParams(int(p.n_planets), int(p.SWQ), int(p.FLQ), int(p.precision), float(p.eps),
                      float(p.Rmin), float(p.vsw), float(p.Wz), float(p.Bstar), float(p.Rstar),
                      float(p.light), float(p.fp64_radius), int(p.schedule), 0)


class Engine:
    """One device context: owns the device state of one DiskState."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.ctx = _VP()
        rc = self.lib.ddyn_create(ctypes.byref(self.ctx), int(device))
        if rc != 0:
            raise DDynError(f"ddyn_create(device={device}) failed with code {rc}")
        self.n_grav = self.n_dust = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.ddyn_last_error(self.ctx)
            raise DDynError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.ctx:
            self.lib.ddyn_destroy(self.ctx)
            self.ctx = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, p: Params):
        cp = to_c_params(p)
        self._check(self.lib.ddyn_set_params(self.ctx, ctypes.byref(cp)), "ddyn_set_params")

    def upload(self, st: DiskState, params: Params = None):
        st.validate()
        if params is not None:
            self.set_params(params)
        self._check(self.lib.ddyn_upload_state(self.ctx, st.n_grav, st.n_dust, _dp(st.x), _dp(st.y),
                                               _dp(st.z), _dp(st.vx), _dp(st.vy), _dp(st.vz),
                                               _dp(st.mu), _dp(st.N), _ip(st.id)),
                    "ddyn_upload_state")
        self.n_grav, self.n_dust = st.n_grav, st.n_dust
        if st.beta_t is not None:
            nb = len(st.beta_t)
            self._check(self.lib.ddyn_upload_bin_tables(
                self.ctx, nb, _fp(np.ascontiguousarray(st.beta_t, np.float32)),
                _fp(np.ascontiguousarray(st.betasw_t, np.float32)),
                _fp(np.ascontiguousarray(st.q_t, np.float32))), "ddyn_upload_bin_tables")
        self._check(self.lib.ddyn_upload_grain_beta(self.ctx, _dp(st.beta_p), _dp(st.betasw_p)),
                    "ddyn_upload_grain_beta")
        if params is not None:   # plan the pair split for the new sizes
            self.set_params(params)

    def step(self, dt, nsteps=1):
        self._check(self.lib.ddyn_step(self.ctx, float(dt), int(nsteps)), "ddyn_step")

    def synchronize(self):
        self._check(self.lib.ddyn_synchronize(self.ctx), "ddyn_synchronize")

    def accel(self):
        n = self.n_grav + self.n_dust
        ax, ay, az = np.empty(n), np.empty(n), np.empty(n)
        rem = np.zeros(max(self.n_dust, 1), dtype=np.int32)
        self._check(self.lib.ddyn_accel(self.ctx, _dp(ax), _dp(ay), _dp(az), _ip(rem)), "ddyn_accel")
        return np.stack([ax, ay, az], axis=1), rem[: self.n_dust]

    def accel_noalloc(self, n_grav, n_dust, x, y, z, vx, vy, vz, mu, id_, beta_p, betasw_p, ax, ay, az, removed):
        """ddyn_accel_noalloc on device addresses (ints, or None for the
        optional arrays); asynchronous on the context's stream."""
        self._check(self.lib.ddyn_accel_noalloc(self.ctx, int(n_grav), int(n_dust), x, y, z, vx, vy, vz, mu, id_,
                                                beta_p, betasw_p, ax, ay, az, removed), "ddyn_accel_noalloc")

    def download(self, st: DiskState = None):
        n = self.n_grav + self.n_dust
        out = [np.empty(n) for _ in range(7)]
        self._check(self.lib.ddyn_download_state(self.ctx, *[_dp(a) for a in out]),
                    "ddyn_download_state")
        if st is None:
            return out
        st.x, st.y, st.z, st.vx, st.vy, st.vz, st.N = out
        return st

    def energy(self):
        out = np.zeros(7)
        self._check(self.lib.ddyn_energy(self.ctx, _dp(out)), "ddyn_energy")
        return out

    def adjoint(self, dt, nsteps, lam_r, lam_v, ckpt_every=0):
        """Discrete adjoint from the uploaded state.  lam_r, lam_v: (n_all, 3)
        dL/d(final r, v).  Returns (g_r, g_v) as (n_all, 3), g_mu (n_grav) and
        g_beta (n_dust).  The device state ends at the state after nsteps."""
        n = self.n_grav + self.n_dust
        lr = np.ascontiguousarray(np.asarray(lam_r, dtype=np.float64).T).reshape(3 * n)
        lv = np.ascontiguousarray(np.asarray(lam_v, dtype=np.float64).T).reshape(3 * n)
        g_r, g_v = np.empty(3 * n), np.empty(3 * n)
        g_mu = np.empty(self.n_grav)
        g_beta = np.zeros(max(self.n_dust, 1))
        self._check(self.lib.ddyn_adjoint(self.ctx, float(dt), int(nsteps), int(ckpt_every), _dp(lr), _dp(lv),
                                          _dp(g_r), _dp(g_v), _dp(g_mu), _dp(g_beta)), "ddyn_adjoint")
        return (g_r.reshape(3, n).T.copy(), g_v.reshape(3, n).T.copy(), g_mu, g_beta[: self.n_dust])

    def adjoint_record(self, dt, nsteps, ckpt_every=0):
        """Forward pass of the adjoint, keeping its tape on the device; the
        device state ends at the state after nsteps."""
        self._check(self.lib.ddyn_adjoint_record(self.ctx, float(dt), int(nsteps), int(ckpt_every)),
                    "ddyn_adjoint_record")

    def adjoint_reverse(self, lam_r, lam_v):
        """Gradients of the recorded run for dL/d(final r, v) = lam_r, lam_v,
        as adjoint() returns them."""
        n = self.n_grav + self.n_dust
        lr = np.ascontiguousarray(np.asarray(lam_r, dtype=np.float64).T).reshape(3 * n)
        lv = np.ascontiguousarray(np.asarray(lam_v, dtype=np.float64).T).reshape(3 * n)
        g_r, g_v = np.empty(3 * n), np.empty(3 * n)
        g_mu = np.empty(self.n_grav)
        g_beta = np.zeros(max(self.n_dust, 1))
        self._check(self.lib.ddyn_adjoint_reverse(self.ctx, _dp(lr), _dp(lv), _dp(g_r), _dp(g_v), _dp(g_mu),
                                                  _dp(g_beta)), "ddyn_adjoint_reverse")
        return (g_r.reshape(3, n).T.copy(), g_v.reshape(3, n).T.copy(), g_mu, g_beta[: self.n_dust])

    def set_timing(self, enable=True):
        self._check(self.lib.ddyn_set_timing(self.ctx, 1 if enable else 0), "ddyn_set_timing")

    def read_timing(self, kclass):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        self._check(self.lib.ddyn_read_timing(self.ctx, int(kclass), ctypes.byref(n), ctypes.byref(ms)),
                    "ddyn_read_timing")
        return n.value, ms.value

    def pair_plan(self):
        """(nslice, slice_len) of the massive pair kernels for the uploaded state."""
        ns = ctypes.c_int32(0)
        ln = ctypes.c_int32(0)
        self._check(self.lib.ddyn_pair_plan(self.ctx, ctypes.byref(ns), ctypes.byref(ln)), "ddyn_pair_plan")
        return ns.value, ln.value


    def upload_optics(self, t: OpticsTables):
        """The dust grid's optical tables for images and SEDs."""
        c = t.c_struct()
        self._check(self.lib.ddyn_upload_optics(self.ctx, ctypes.byref(c)), "ddyn_upload_optics")
        self.n_wav = t.n_wav

    def calc_flux(self, view: View, wj):
        """Image [Jy per pixel] of the current state at wavelength index wj,
        in the layout pixel (x, y) at F[ysize * x + y]."""
        F = np.empty(int(view.xsize) * int(view.ysize))
        v = view.c_struct()
        self._check(self.lib.ddyn_calc_flux(self.ctx, ctypes.byref(v), int(wj), _dp(F)), "ddyn_calc_flux")
        return F

    def calc_sed(self, view: View):
        """Grain flux [Jy] at every wavelength of the uploaded tables."""
        out = np.empty(getattr(self, "n_wav", 0))
        v = view.c_struct()
        self._check(self.lib.ddyn_calc_sed(self.ctx, ctypes.byref(v), _dp(out) if len(out) else None),
                    "ddyn_calc_sed")
        return out

    def mie(self, s_um, lam_um, n, k, cost):
        """Mie tables on the device: Qabs, Qsca (n_sizes, n_wav) and the phase
        function (n_sizes, n_theta, n_wav) at the cosines cost."""
        s, lam, nr, ki, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (s_um, lam_um, n, k, cost))
        ns, nw, nt = len(s), len(lam), len(c)
        qa, qs, qp = np.empty(ns * nw), np.empty(ns * nw), np.empty(ns * nt * nw)
        self._check(self.lib.ddyn_mie(self.ctx, ns, _dp(s), nw, _dp(lam), _dp(nr), _dp(ki), nt, _dp(c), _dp(qa),
                                      _dp(qs), _dp(qp)), "ddyn_mie")
        return qa.reshape(ns, nw), qs.reshape(ns, nw), qp.reshape(ns, nt, nw)

    def rsqrt_fp64(self, x):
        """The library's device FP64 reciprocal square root of every x > 0."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._check(self.lib.ddyn_eval_rsqrt_fp64(self.ctx, len(x), _dp(x), _dp(y)), "ddyn_eval_rsqrt_fp64")
        return y


def run(workload, device=0, nsteps=None, precision=None):
    """Upload a workload, step it and return the final DiskState (convenience)."""
    st = workload.state.copy()
    prm = workload.params if precision is None else workload.params.with_precision(precision)
    with Engine(device) as eng:
        eng.upload(st, prm)
        eng.step(workload.dt, workload.nsteps if nsteps is None else nsteps)
        eng.download(st)
    return st

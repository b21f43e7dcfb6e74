"""ctypes wrapper of the FP64 CPU oracle (oracle/ddyn_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the reference results.  The product (diskdyn_amd) never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I32 = ctypes.POINTER(ctypes.c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = [("n_planets", ctypes.c_int32), ("SWQ", ctypes.c_int32), ("FLQ", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("eps", ctypes.c_double), ("Rmin", ctypes.c_double),
                ("vsw", ctypes.c_double), ("Wz", ctypes.c_double), ("Bstar", ctypes.c_double),
                ("Rstar", ctypes.c_double), ("light", ctypes.c_double)]


class OrcState(ctypes.Structure):
    _fields_ = [("n_grav", ctypes.c_int64), ("n_dust", ctypes.c_int64)] + \
               [(n, _D) for n in ("x", "y", "z", "vx", "vy", "vz", "mu", "N")] + \
               [("id", _I32), ("n_bins", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("beta_tab", _F), ("betasw_tab", _F), ("q_tab", _F),
                ("beta_p", _D), ("betasw_p", _D)]


class OrcOptics(ctypes.Structure):
    _fields_ = [("n_bins", ctypes.c_int32), ("n_wav", ctypes.c_int32), ("n_theta", ctypes.c_int32),
                ("n_r", ctypes.c_int32)] + [(n, _D) for n in ("w", "in_", "size", "Qabs", "Qsca", "Qpfunc", "Td")] + \
               [("Rmin", ctypes.c_double), ("Rmax", ctypes.c_double)]


class OrcView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dist", "rot", "incl", "PA", "fov")] + \
               [("xsize", ctypes.c_int32), ("ysize", ctypes.c_int32)]


_lib = None


REF_ROOT = "/root/reference"
REF_RUNMIE = os.path.join(HERE, "_ref", "runmie")


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    # the reference's own Mie tool, where its sources are present (a checker only)
    if os.path.isfile(os.path.join(REF_ROOT, "tools", "runmie.c")):
        r = subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + REF_ROOT], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference runmie build failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_massive_accel.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int64] + [_D] * 7
        L.orc_massive_accel.restype = None
        L.orc_massive_accel_targets.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int64] + [_D] * 4 + \
            [ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)] + [_D] * 3
        L.orc_massive_accel_targets.restype = None
        L.orc_dust_accel.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcState), _D, _D, _D, _I32]
        L.orc_dust_accel.restype = ctypes.c_int
        L.orc_step.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcState), ctypes.c_double,
                               ctypes.c_int64]
        L.orc_step.restype = ctypes.c_int
        L.orc_adjoint.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcState), ctypes.c_double,
                                  ctypes.c_int64] + [_D] * 6
        L.orc_adjoint.restype = ctypes.c_int
        L.orc_energy.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int64] + [_D] * 8
        L.orc_energy.restype = None
        L.orc_coe2rv.argtypes = [ctypes.c_double] * 9 + [_D]
        L.orc_coe2rv.restype = None
        L.orc_recalc_com.argtypes = [ctypes.c_int64] + [_D] * 7
        L.orc_recalc_com.restype = None
        L.orc_image.argtypes = [ctypes.POINTER(OrcOptics), ctypes.POINTER(OrcView), ctypes.c_int64, ctypes.c_int64] + \
            [_D] * 4 + [_I32, ctypes.c_int32, _D]
        L.orc_image.restype = None
        L.orc_sed.argtypes = [ctypes.POINTER(OrcOptics), ctypes.POINTER(OrcView), ctypes.c_int64, ctypes.c_int64] + \
            [_D] * 4 + [_I32, _D]
        L.orc_sed.restype = None
        L.orc_mie.argtypes = [ctypes.c_double, _D, _D, _D, ctypes.c_int, _D, ctypes.c_int, _D, _D, _D]
        L.orc_mie.restype = ctypes.c_int
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def params(p):
    """OrcParams from a diskdyn_amd.system.Params (same meaning of every field)."""
    return OrcParams(int(p.n_planets), int(p.SWQ), int(p.FLQ), 0, float(p.eps), float(p.Rmin),
                     float(p.vsw), float(p.Wz), float(p.Bstar), float(p.Rstar), float(p.light))


class _Bound:
    """Keeps numpy arrays alive while the C struct points at them."""

    def __init__(self, st):
        self.arrays = {}
        f = lambda a, dt: None if a is None else np.ascontiguousarray(a, dtype=dt)
        for n in ("x", "y", "z", "vx", "vy", "vz", "mu", "N"):
            self.arrays[n] = getattr(st, n)
            assert self.arrays[n].dtype == np.float64 and self.arrays[n].flags.c_contiguous
        self.arrays["id"] = f(st.id, np.int32)
        for n in ("beta_t", "betasw_t", "q_t"):
            self.arrays[n] = f(getattr(st, n), np.float32)
        for n in ("beta_p", "betasw_p"):
            self.arrays[n] = f(getattr(st, n), np.float64)
        A = self.arrays
        nb = 0 if A["beta_t"] is None else len(A["beta_t"])
        self.c = OrcState(st.n_grav, st.n_dust, *[_p(A[n]) for n in ("x", "y", "z", "vx", "vy", "vz", "mu", "N")],
                          _p(A["id"], _I32), nb, 0, _p(A["beta_t"], _F), _p(A["betasw_t"], _F),
                          _p(A["q_t"], _F), _p(A["beta_p"]), _p(A["betasw_p"]))


def massive_accel(p, x, y, z, mu):
    n = len(x)
    ax, ay, az = np.empty(n), np.empty(n), np.empty(n)
    c = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, mu)]
    lib().orc_massive_accel(ctypes.byref(params(p)), n, *[_p(v) for v in c], _p(ax), _p(ay), _p(az))
    return np.stack([ax, ay, az], axis=1)


def massive_accel_targets(p, x, y, z, mu, idx):
    """Accelerations of the massive targets idx (each over all sources)."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    nt = len(idx)
    ax, ay, az = np.empty(nt), np.empty(nt), np.empty(nt)
    c = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, mu)]
    lib().orc_massive_accel_targets(ctypes.byref(params(p)), len(c[0]), *[_p(v) for v in c], nt,
                                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    _p(ax), _p(ay), _p(az))
    return np.stack([ax, ay, az], axis=1)


def accel(p, st):
    """Accelerations of every particle at the current state (n_all x 3) and the
    per-grain removal flags, as the C ABI's ddyn_accel reports them."""
    a_m = massive_accel(p, st.x[: st.n_grav], st.y[: st.n_grav], st.z[: st.n_grav], st.mu[: st.n_grav])
    nd = st.n_dust
    ax, ay, az = np.zeros(nd), np.zeros(nd), np.zeros(nd)
    rem = np.zeros(max(nd, 1), dtype=np.int32)
    b = _Bound(st)
    if nd:
        lib().orc_dust_accel(ctypes.byref(params(p)), ctypes.byref(b.c), _p(ax), _p(ay), _p(az),
                             _p(rem, _I32))
    return np.concatenate([a_m, np.stack([ax, ay, az], axis=1)]), rem[:nd]


def step(p, st, dt, nsteps):
    """Advance a DiskState in place by nsteps RK4 steps."""
    b = _Bound(st)
    rc = lib().orc_step(ctypes.byref(params(p)), ctypes.byref(b.c), float(dt), int(nsteps))
    if rc != 0:
        raise MemoryError("oracle workspace allocation failed")
    return st


def adjoint(p, st, dt, nsteps, lam_r, lam_v):
    """Discrete adjoint of nsteps steps from st (unchanged).  lam_r, lam_v:
    (n_all, 3) = dL/d(final r, v).  Returns (g_r, g_v, g_mu, g_beta)."""
    na = st.n_all
    lr = np.ascontiguousarray(lam_r, dtype=np.float64).reshape(na * 3)
    lv = np.ascontiguousarray(lam_v, dtype=np.float64).reshape(na * 3)
    g_r, g_v = np.zeros(na * 3), np.zeros(na * 3)
    g_mu, g_beta = np.zeros(st.n_grav), np.zeros(max(st.n_dust, 1))
    b = _Bound(st)
    rc = lib().orc_adjoint(ctypes.byref(params(p)), ctypes.byref(b.c), float(dt), int(nsteps),
                           _p(lr), _p(lv), _p(g_r), _p(g_v), _p(g_mu), _p(g_beta))
    if rc != 0:
        raise RuntimeError(f"orc_adjoint failed ({rc})")
    return g_r.reshape(na, 3), g_v.reshape(na, 3), g_mu, g_beta[: st.n_dust]


def energy(p, st):
    n = st.n_grav
    out = np.zeros(7)
    arrs = [np.ascontiguousarray(getattr(st, k)[:n]) for k in ("x", "y", "z", "vx", "vy", "vz", "mu")]
    lib().orc_energy(ctypes.byref(params(p)), n, *[_p(a) for a in arrs], _p(out))
    return out


def coe2rv(M, m, a, e, inc, w, o, nu, beta=0.0):
    out = np.zeros(6)
    lib().orc_coe2rv(M, m, a, e, inc, w, o, nu, beta, _p(out))
    return out


def recalc_com(x, y, z, vx, vy, vz, mu):
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, vx, vy, vz, mu)]
    lib().orc_recalc_com(len(mu), *[_p(a) for a in arrs])
    return arrs[:6]


def max_threads():
    return lib().orc_max_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def _optics_view(t, v):
    """C structs of a diskdyn_amd.imaging OpticsTables and View (same fields)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (t.w, t.in_, t.size, t.Qabs, t.Qsca, t.Qpfunc, t.Td)]
    o = OrcOptics(t.Qabs.shape[0], t.Qabs.shape[1], t.Qpfunc.shape[1], t.Td.shape[1], *[_p(a) for a in arrs],
                  float(t.Rmin), float(t.Rmax))
    cv = OrcView(float(v.dist), float(v.rot), float(v.incl), float(v.PA), float(v.fov), int(v.xsize), int(v.ysize))
    return o, cv, arrs


def image(t, v, st, wj):
    """Image at wavelength index wj, pixel (x, y) at F[ysize * x + y] [Jy]."""
    o, cv, keep = _optics_view(t, v)
    F = np.zeros(int(v.xsize) * int(v.ysize))
    ids = None if st.id is None else np.ascontiguousarray(st.id, dtype=np.int32)
    lib().orc_image(ctypes.byref(o), ctypes.byref(cv), st.n_grav, st.n_dust, _p(st.x), _p(st.y), _p(st.z), _p(st.N),
                    _p(ids, _I32), int(wj), _p(F))
    return F


def sed(t, v, st):
    """Grain flux [Jy] at every wavelength of the tables."""
    o, cv, keep = _optics_view(t, v)
    out = np.zeros(t.Qabs.shape[1])
    ids = None if st.id is None else np.ascontiguousarray(st.id, dtype=np.int32)
    lib().orc_sed(ctypes.byref(o), ctypes.byref(cv), st.n_grav, st.n_dust, _p(st.x), _p(st.y), _p(st.z), _p(st.N),
                  _p(ids, _I32), _p(out))
    return out


def mie(s_um, lam_um, n, k, cost):
    """Mie tables of one sphere: Qabs, Qsca (n_wav) and the phase function
    (n_theta, n_wav) at the cosines cost (tools/mie_single.c restated)."""
    lam, nre, kim, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (lam_um, n, k, cost))
    nl, nt = len(lam), len(c)
    qa, qs, qp = np.zeros(nl), np.zeros(nl), np.zeros(nl * nt)
    rc = lib().orc_mie(float(s_um), _p(lam), _p(nre), _p(kim), nl, _p(c), nt, _p(qa), _p(qs), _p(qp))
    if rc:
        raise ValueError("size parameter beyond 7.2e6")
    return qa, qs, qp.reshape(nt, nl)

"""ctypes binding of libddyn_host.so (include/ddyn_host.h): .param reader,
det_step and the data-file writer/reader.  Host code only (no GPU)."""
import ctypes
import os

import numpy as np

from .imaging import COpticsTables, OpticsTables
from .system import DiskState, Params, make_state
from .units import AU, YEAR

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB_PATH = os.path.join(_HERE, "libddyn_host.so")
VERSION = "1.01"

MAXP, MAXD, MAXB, MAXW, STR = 64, 16, 16, 64, 256
_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


class Disk(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("Rin", "Rout", "dhdr", "sigmae", "amax", "amin", "eta", "Mtot", "mmax",
                                   "mmin", "gammas", "dmax", "dmmax", "C")] + [("Ngrav", _I64), ("Nsmall", _I64)]


class Blob(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("a", "ecc", "inc", "w", "o", "nu", "sigma", "amax", "amin", "eta", "Mtot",
                                   "mmax", "mmin", "vkepp", "dmax", "dmmax", "C")] + \
               [("Ngrav", _I64), ("Nsmall", _I64), ("collcloud", _I32), ("pad_", _I32)]


class Config(ctypes.Structure):
    _fields_ = [("verb", _I32), ("gpuid", _I32), ("gl", _I32), ("n_step", _I32)] + \
               [(n, _D) for n in ("t_gl", "t_write", "t_fits", "t_sed", "t_end")] + \
               [("sptype", ctypes.c_char * STR)] + \
               [(n, _D) for n in ("teff", "rstar", "Bstar", "Wz", "mloss", "vsw", "dist", "fov", "PA", "inc",
                                   "longitude")] + \
               [("xsize", _I32), ("ysize", _I32), ("nwavs", _I32), ("radQ", _I32), ("wavs", _D * MAXW),
                ("specfile", ctypes.c_char * STR), ("SWQ", _I32), ("FLQ", _I32), ("n_planets", _I32),
                ("n_disks", _I32)] + \
               [(n, _D * MAXP) for n in ("px", "py", "pz", "pvx", "pvy", "pvz", "pm")] + \
               [("rho", _D), ("kqe", _D), ("comp", ctypes.c_char * STR), ("n_blobs", _I32), ("n_grid", _I32),
                ("n_wav", _I32), ("n_theta", _I32), ("disks", Disk * MAXD), ("blobs", Blob * MAXB)] + \
               [(n, _D) for n in ("amin", "amax", "mmin", "mmax", "f")] + \
               [("seed", _I64), ("name_stub", ctypes.c_char * STR)]


class RunInfo(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("time", "dt", "tlastd", "tlastf", "tlasts", "Rmax")] + \
               [("n_mass", _I64), ("n_dust", _I64), ("n_all", _I64)]


_lib = None
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int32)
_PF = ctypes.POINTER(ctypes.c_float)


class HostIOError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {lib().ddyn_host_error_text(code).decode()} (code {code})")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"{HOST_LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(HOST_LIB_PATH)
        L.ddyn_host_error_text.restype = ctypes.c_char_p
        L.ddyn_host_error_text.argtypes = [ctypes.c_int]
        L.ddyn_read_param.restype = ctypes.c_int
        L.ddyn_read_param.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config)]
        L.ddyn_det_step.restype = ctypes.c_double
        L.ddyn_det_step.argtypes = [ctypes.POINTER(Config), ctypes.c_int64] + [_PD] * 4
        L.ddyn_write_datafile.restype = ctypes.c_int
        L.ddyn_write_datafile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Config),
                                          ctypes.POINTER(RunInfo)] + [_PD] * 8 + [_PI, _PF, ctypes.c_int32, _PD]
        L.ddyn_read_datafile.restype = ctypes.c_int
        L.ddyn_read_datafile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Config),
                                         ctypes.POINTER(RunInfo)] + [_PD] * 8 + [_PI, _PD]
        L.ddyn_optics_read.restype = ctypes.c_int
        L.ddyn_optics_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Config),
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.ddyn_optics_free.restype = None
        L.ddyn_optics_free.argtypes = [ctypes.c_void_p]
        L.ddyn_optics_info.restype = ctypes.c_int
        L.ddyn_optics_info.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(_I32)] * 3 + [_PD] * 2
        L.ddyn_set_dust_grid.restype = ctypes.c_int
        L.ddyn_set_dust_grid.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p]
        L.ddyn_outer_radius.restype = ctypes.c_double
        L.ddyn_outer_radius.argtypes = [ctypes.POINTER(Config)]
        L.ddyn_grain_tables.restype = ctypes.c_int
        L.ddyn_grain_tables.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_int32] + [_PF] * 3 + \
            [_PD] * 3
        L.ddyn_generate_counts.restype = ctypes.c_int
        L.ddyn_generate_counts.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]
        L.ddyn_generate.restype = ctypes.c_int
        L.ddyn_generate.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p, _PF, ctypes.c_int32, ctypes.c_int64] + \
            [_PD] * 8 + [_PI]
        L.ddyn_restart_ids.restype = ctypes.c_int
        L.ddyn_restart_ids.argtypes = [ctypes.POINTER(Config), ctypes.c_int64, _PD, _PI]
        L.ddyn_grid_optics_make.restype = ctypes.c_int
        L.ddyn_grid_optics_make.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.ddyn_grid_optics_free.restype = None
        L.ddyn_grid_optics_free.argtypes = [ctypes.c_void_p]
        L.ddyn_grid_optics_tables.restype = ctypes.c_int
        L.ddyn_grid_optics_tables.argtypes = [ctypes.c_void_p, ctypes.POINTER(COpticsTables)]
        L.ddyn_image_wavelengths.restype = ctypes.c_int
        L.ddyn_image_wavelengths.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p, _PI]
        L.ddyn_write_image.restype = ctypes.c_int
        L.ddyn_write_image.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config), ctypes.c_double, ctypes.c_double,
                                       _PD, ctypes.c_char_p, ctypes.c_int32]
        L.ddyn_write_sed.restype = ctypes.c_int
        L.ddyn_write_sed.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config), ctypes.c_double, ctypes.c_int32,
                                     _PD, _PD, _PD, ctypes.c_char_p, ctypes.c_int32]
        _lib = L
    return _lib


def _p(a, t=_PD):
    return None if a is None else a.ctypes.data_as(t)


def read_param(path):
    cfg = Config()
    rc = lib().ddyn_read_param(str(path).encode(), ctypes.byref(cfg))
    if rc:
        raise HostIOError(rc, f"read_param({path})")
    return cfg


def det_step(cfg, st):
    ng = st.n_grav
    arr = [np.ascontiguousarray(a[:ng]) for a in (st.x, st.y, st.z, st.mu)]
    return lib().ddyn_det_step(ctypes.byref(cfg), ng, *[_p(a) for a in arr])


def state_from_config(cfg):
    """Star + planets of a parsed .param as a DiskState."""
    n = cfg.n_planets
    st = make_state(n, 0)
    for j in range(n):
        st.x[j], st.y[j], st.z[j] = cfg.px[j], cfg.py[j], cfg.pz[j]
        st.vx[j], st.vy[j], st.vz[j] = cfg.pvx[j], cfg.pvy[j], cfg.pvz[j]
        st.mu[j] = cfg.pm[j]
    return st


def params_from_config(cfg, **kw):
    """Physics parameters of the C ABI from a parsed configuration."""
    return Params(n_planets=cfg.n_planets, SWQ=cfg.SWQ, FLQ=cfg.FLQ, vsw=cfg.vsw * YEAR / AU, Wz=cfg.Wz,
                  Bstar=cfg.Bstar, Rstar=cfg.rstar, **kw)


def write_datafile(path, cfg, st, run, beta_bins=None):
    st.validate()
    bb = None if beta_bins is None else np.ascontiguousarray(beta_bins, dtype=np.float32)
    rc = lib().ddyn_write_datafile(str(path).encode(), VERSION.encode(), ctypes.byref(cfg), ctypes.byref(run),
                                   *[_p(getattr(st, n)) for n in ("x", "y", "z", "vx", "vy", "vz", "N", "mu")],
                                   _p(st.id, _PI), _p(bb, _PF), 0 if bb is None else len(bb), _p(st.beta_p))
    if rc:
        raise HostIOError(rc, f"write_datafile({path})")


def read_datafile(path):
    cfg, run = Config(), RunInfo()
    rc = lib().ddyn_read_datafile(str(path).encode(), VERSION.encode(), ctypes.byref(cfg), ctypes.byref(run),
                                  *([None] * 8), None, None)
    if rc:
        raise HostIOError(rc, f"read_datafile({path})")
    n = run.n_all
    arrs = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "N", "mu", "beta")}
    ids = np.zeros(n, dtype=np.int32)
    rc = lib().ddyn_read_datafile(str(path).encode(), VERSION.encode(), ctypes.byref(cfg), ctypes.byref(run),
                                  *[_p(arrs[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "N", "mu")],
                                  _p(ids, _PI), _p(arrs["beta"]))
    if rc:
        raise HostIOError(rc, f"read_datafile({path})")
    nd = run.n_dust
    st = DiskState(n - nd, nd, arrs["x"], arrs["y"], arrs["z"], arrs["vx"], arrs["vy"], arrs["vz"], arrs["mu"],
                   arrs["N"], ids, beta_p=np.ascontiguousarray(arrs["beta"][n - nd:]) if nd else None)
    return cfg, run, st


class Optics:
    """The grain composition table and stellar spectrum of a run (ddyn_optics_read)."""

    def __init__(self, comp_path, spec_path, cfg):
        self.ptr = ctypes.c_void_p()
        rc = lib().ddyn_optics_read(str(comp_path).encode(), str(spec_path).encode(), ctypes.byref(cfg),
                                    ctypes.byref(self.ptr))
        if rc:
            raise HostIOError(rc, f"optics_read({comp_path}, {spec_path})")

    def info(self):
        nw, nt, ns = _I32(), _I32(), _I32()
        dmin, dmax = ctypes.c_double(), ctypes.c_double()
        lib().ddyn_optics_info(self.ptr, ctypes.byref(nw), ctypes.byref(nt), ctypes.byref(ns), ctypes.byref(dmin),
                               ctypes.byref(dmax))
        return {"n_wav": nw.value, "n_theta": nt.value, "n_sizes": ns.value, "dmin": dmin.value, "dmax": dmax.value}

    def close(self):
        if self.ptr:
            lib().ddyn_optics_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_dust_grid(cfg, optics):
    rc = lib().ddyn_set_dust_grid(ctypes.byref(cfg), optics.ptr)
    if rc:
        raise HostIOError(rc, "set_dust_grid")


def outer_radius(cfg):
    return lib().ddyn_outer_radius(ctypes.byref(cfg))


def grain_tables(cfg, optics):
    """Per-bin float beta, beta_sw, q, sizes [m], Rmin [AU] and L [Lsun] of the dust grid."""
    n = cfg.n_grid
    beta, betasw, q = (np.zeros(n, dtype=np.float32) for _ in range(3))
    size = np.zeros(n)
    rmin, lum = ctypes.c_double(), ctypes.c_double()
    rc = lib().ddyn_grain_tables(ctypes.byref(cfg), optics.ptr, n, _p(beta, _PF), _p(betasw, _PF), _p(q, _PF),
                                 _p(size), ctypes.byref(rmin), ctypes.byref(lum))
    if rc:
        raise HostIOError(rc, "grain_tables")
    return {"beta": beta, "betasw": betasw, "q": q, "size": size, "Rmin": rmin.value, "lum": lum.value}


def generate(cfg, optics=None, beta=None):
    """Star, planets and the particles of the disks and blobs as a DiskState."""
    nm, nd = _I64(), _I64()
    lib().ddyn_generate_counts(ctypes.byref(cfg), ctypes.byref(nm), ctypes.byref(nd))
    ng = cfg.n_planets + nm.value
    st = make_state(ng, nd.value)
    b = None if beta is None else np.ascontiguousarray(beta, dtype=np.float32)
    rc = lib().ddyn_generate(ctypes.byref(cfg), None if optics is None else optics.ptr, _p(b, _PF),
                             0 if b is None else len(b), st.n_all,
                             *[_p(getattr(st, k)) for k in ("x", "y", "z", "vx", "vy", "vz", "N", "mu")],
                             _p(st.id, _PI))
    if rc:
        raise HostIOError(rc, "generate")
    return st


def restart_ids(cfg, st):
    ids = np.zeros(st.n_all, dtype=np.int32)
    lib().ddyn_restart_ids(ctypes.byref(cfg), st.n_all, _p(np.ascontiguousarray(st.mu)), _p(ids, _PI))
    return ids


def grid_optics(cfg, optics):
    """The dust grid's optical tables (ddyn_grid_optics_make) as OpticsTables,
    with the image wavelength indices of the run (ddyn_image_wavelengths)."""
    L = lib()
    ptr = ctypes.c_void_p()
    rc = L.ddyn_grid_optics_make(ctypes.byref(cfg), optics.ptr, ctypes.byref(ptr))
    if rc:
        raise HostIOError(rc, "grid_optics_make")
    try:
        v = COpticsTables()
        L.ddyn_grid_optics_tables(ptr, ctypes.byref(v))
        nb, nw, nt, nr = v.n_bins, v.n_wav, v.n_theta, v.n_r
        arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
        t = OpticsTables(w=arr(v.w, nw), in_=arr(v.in_, nw), size=arr(v.size, nb),
                         Qabs=arr(v.Qabs, nb * nw).reshape(nb, nw), Qsca=arr(v.Qsca, nb * nw).reshape(nb, nw),
                         Qpfunc=arr(v.Qpfunc, nb * nt * nw).reshape(nb, nt, nw), Td=arr(v.Td, nb * nr).reshape(nb, nr),
                         Rmin=v.Rmin, Rmax=v.Rmax)
        wj = np.zeros(max(cfg.nwavs, 1), dtype=np.int32)
        rc = L.ddyn_image_wavelengths(ctypes.byref(cfg), ptr, _p(wj, _PI))
        if rc:
            raise HostIOError(rc, "image_wavelengths")
        return t, wj[: cfg.nwavs]
    finally:
        L.ddyn_grid_optics_free(ptr)


def write_image(cfg, time, wvl, F, version=VERSION):
    """FITS image <stub>_<wvl um>_<time>.fits; returns the file name."""
    name = ctypes.create_string_buffer(STR + 64)
    F = np.ascontiguousarray(F, dtype=np.float64)
    rc = lib().ddyn_write_image(version.encode(), ctypes.byref(cfg), float(time), float(wvl), _p(F), name, len(name))
    if rc:
        raise HostIOError(rc, "write_image")
    return name.value.decode()


def write_sed(cfg, time, tables, Fexc, version=VERSION):
    """SED file <stub>_SED_<time>.dat; returns the file name."""
    name = ctypes.create_string_buffer(STR + 64)
    Fexc = np.ascontiguousarray(Fexc, dtype=np.float64)
    rc = lib().ddyn_write_sed(version.encode(), ctypes.byref(cfg), float(time), tables.n_wav, _p(tables.w),
                              _p(tables.in_), _p(Fexc), name, len(name))
    if rc:
        raise HostIOError(rc, "write_sed")
    return name.value.decode()

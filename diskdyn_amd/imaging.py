"""Images and SEDs of the grains: the optical tables of the dust grid and the
observer, as numpy-side objects, and their C structs (ddyn_optics_tables and
ddyn_view of include/ddyn.h).  The computation runs on the GPU
(Engine.upload_optics, Engine.calc_flux, Engine.calc_sed); the tables of a
.param run come from the host library (hostio.grid_optics)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from .units import PC

_PD = ctypes.POINTER(ctypes.c_double)


class COpticsTables(ctypes.Structure):
    _fields_ = [("n_bins", ctypes.c_int32), ("n_wav", ctypes.c_int32), ("n_theta", ctypes.c_int32),
                ("n_r", ctypes.c_int32)] + [(n, _PD) for n in ("w", "in_", "size", "Qabs", "Qsca", "Qpfunc", "Td")] + \
               [("Rmin", ctypes.c_double), ("Rmax", ctypes.c_double)]


class CView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dist", "rot", "incl", "PA", "fov")] + \
               [("xsize", ctypes.c_int32), ("ysize", ctypes.c_int32)]


@dataclass
class OpticsTables:
    """w, in_ (n_wav): wavelengths [m] and stellar flux at the observer
    [W m^-3]; size (n_bins) grain radii [m]; Qabs, Qsca (n_bins, n_wav);
    Qpfunc (n_bins, n_theta, n_wav) on cos = -1 .. 1; Td (n_bins, n_r) grain
    temperatures [K] on n_r radii log-spaced from Rmin to Rmax [AU]."""
    w: np.ndarray
    in_: np.ndarray
    size: np.ndarray
    Qabs: np.ndarray
    Qsca: np.ndarray
    Qpfunc: np.ndarray
    Td: np.ndarray
    Rmin: float
    Rmax: float

    def __post_init__(self):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.w, self.in_, self.size = f(self.w), f(self.in_), f(self.size)
        self.Qabs, self.Qsca, self.Qpfunc, self.Td = f(self.Qabs), f(self.Qsca), f(self.Qpfunc), f(self.Td)
        nb, nw = self.Qabs.shape
        assert self.w.shape == (nw,) and self.in_.shape == (nw,) and self.size.shape == (nb,)
        assert self.Qsca.shape == (nb, nw) and self.Qpfunc.shape[0] == nb and self.Qpfunc.shape[2] == nw
        assert self.Td.shape[0] == nb

    @property
    def n_bins(self):
        return self.Qabs.shape[0]

    @property
    def n_wav(self):
        return self.Qabs.shape[1]

    @property
    def n_theta(self):
        return self.Qpfunc.shape[1]

    @property
    def n_r(self):
        return self.Td.shape[1]

    def c_struct(self):
        """The C struct; valid while this object lives."""
        p = lambda a: a.ctypes.data_as(_PD)
        return COpticsTables(self.n_bins, self.n_wav, self.n_theta, self.n_r, p(self.w), p(self.in_), p(self.size),
                             p(self.Qabs), p(self.Qsca), p(self.Qpfunc), p(self.Td), float(self.Rmin),
                             float(self.Rmax))


@dataclass
class View:
    """Observer at dist [m]; rot (longitude), incl, PA [rad]; a field of fov
    [AU] across the larger side of an xsize x ysize image."""
    dist: float
    rot: float = 0.0
    incl: float = 0.0
    PA: float = 0.0
    fov: float = 100.0
    xsize: int = 64
    ysize: int = 64

    def c_struct(self):
        return CView(float(self.dist), float(self.rot), float(self.incl), float(self.PA), float(self.fov),
                     int(self.xsize), int(self.ysize))


def view_from_config(cfg):
    """The .param run's observer (record 7) from a hostio.Config."""
    return View(dist=cfg.dist, rot=cfg.longitude, incl=cfg.inc, PA=cfg.PA, fov=cfg.fov, xsize=cfg.xsize,
                ysize=cfg.ysize)


def view_pc(dist_pc, **kw):
    return View(dist=dist_pc * PC, **kw)


def image_to_fits_rows(F, xsize, ysize):
    """The image as rows y = 0..ysize-1 of xsize pixels (the FITS array's
    order) from the layout pixel (x, y) at F[ysize * x + y]."""
    return np.asarray(F).reshape(xsize, ysize).T

"""The oracle's images and SEDs (oracle/ddyn_oracle.c: orc_image, orc_sed,
restating src/gpu_functions.cu:983-1156), pinned by what the physics and the
geometry fix independently of the formulas restated there:

* the thermal SED of one grain integrates over wavelength to s^2 sigma T^4 / d^2
  (the Stefan-Boltzmann law: pins the Planck function and the Jy conversion);
* scattered light falls off as 1/R^2 and grows as s^2, and the phase function
  is read where the reference's geometry puts a grain in front of, behind and
  beside the star (an exponential phase function interpolates exactly);
* a power-law temperature table is interpolated exactly (log-log);
* pixel positions follow from the rotations by hand;
* the image's total equals the SED at its wavelength when every grain is in
  the field;
* grains without a size bin, within Rmin or with N = 0 add nothing.
"""
import math

import numpy as np
import pytest

from diskdyn_amd.imaging import OpticsTables, View
from diskdyn_amd.system import make_state
from diskdyn_amd.units import AU, PC
from oracle import oracle as O

H, C, K = 6.626068e-34, 299792458.0, 1.3806503e-23       # src/ddyn.cuh:23-25
SIGMA = 2.0 * math.pi ** 5 * K ** 4 / (15.0 * H ** 3 * C ** 2)
DIST = 10.0 * PC


def tables(wav, n_bins=1, sizes=None, qabs=1.0, qsca=0.0, n_theta=21, phase_a=0.0, Td=None, n_r=100,
           Rmin=0.1, Rmax=1000.0, influx=1.0):
    nw = len(wav)
    sizes = np.full(n_bins, 1e-5) if sizes is None else np.asarray(sizes, float)
    mu = np.linspace(-1.0, 1.0, n_theta)
    qpf = np.broadcast_to(np.exp(phase_a * mu)[None, :, None], (n_bins, n_theta, nw))
    if Td is None:
        Td = np.full((n_bins, n_r), 150.0)
    return OpticsTables(w=wav, in_=np.full(nw, influx), size=sizes, Qabs=np.full((n_bins, nw), qabs),
                        Qsca=np.full((n_bins, nw), qsca), Qpfunc=qpf, Td=Td, Rmin=Rmin, Rmax=Rmax)


def grains(pos, ids=None, N=None):
    pos = np.asarray(pos, float).reshape(-1, 3)
    n = len(pos)
    st = make_state(1, n)
    st.mu[0] = 4.0 * math.pi ** 2
    st.x[1:], st.y[1:], st.z[1:] = pos[:, 0], pos[:, 1], pos[:, 2]
    st.N[1:] = 1.0 if N is None else N
    st.id[1:] = 0 if ids is None else ids
    return st


def test_thermal_sed_integrates_to_stefan_boltzmann():
    wav = np.geomspace(0.3e-6, 0.2, 6000)
    T0, s = 150.0, 1e-5
    t = tables(wav, qabs=1.0, qsca=0.0, Td=np.full((1, 100), T0))
    st = grains([30.0, 0.0, 0.0], N=[3.0])
    Fjy = O.sed(t, View(dist=DIST), st)
    F_lambda = Fjy * C / (1e26 * wav ** 2)
    total = np.trapezoid(F_lambda, wav)
    expect = 3.0 * s * s * SIGMA * T0 ** 4 / DIST ** 2
    assert abs(total / expect - 1.0) < 2e-4


def test_scattered_light_geometry_and_scaling():
    wav = np.array([0.5e-6, 1e-6])
    a, nt = 1.3, 41
    t = tables(wav, qabs=0.0, qsca=1.0, n_theta=nt, phase_a=a, influx=2.0)
    edge_on = View(dist=DIST, incl=math.pi / 2)
    # observer towards -y: a grain at -y is between star and observer, one at +y behind the star
    st = grains([[0.0, -30.0, 0.0], [0.0, 30.0, 0.0], [0.0, -60.0, 0.0]])
    sed_each = [O.sed(t, edge_on, grains([p]))[0] for p in ([0.0, -30.0, 0.0], [0.0, 30.0, 0.0],
                                                            [0.0, -60.0, 0.0])]
    front, back, front_far = sed_each
    # the reference reads the phase function at s.l = -1 in front, clamped to 1 - 2/n_theta behind
    assert front / back == pytest.approx(math.exp(-a) / math.exp(a * (1.0 - 2.0 / nt)), rel=1e-12)
    # inverse square law in the distance from the star
    assert front / front_far == pytest.approx(4.0, rel=1e-12)
    # absolute value: in pi s^2 Qsca P / R^2, in Jy
    s, R = 1e-5, 30.0 * AU
    expect = 2.0 * math.pi * s * s * math.exp(-a) / R ** 2 * 1e26 * wav[0] ** 2 / C
    assert front == pytest.approx(expect, rel=1e-12)
    assert O.sed(t, edge_on, st)[0] == pytest.approx(front + back + front_far, rel=1e-14)
    # a face-on observer sees every grain in the plane at quadrature: P(0)
    face_on = View(dist=DIST)
    q = O.sed(t, face_on, grains([[30.0, 0.0, 0.0]]))[0]
    assert q / front == pytest.approx(1.0 / math.exp(-a), rel=1e-9)
    # grain size: flux ~ s^2
    t2 = tables(wav, qabs=0.0, qsca=1.0, n_theta=nt, phase_a=a, influx=2.0, sizes=[2e-5])
    assert O.sed(t2, edge_on, grains([[0.0, -30.0, 0.0]]))[0] / front == pytest.approx(4.0, rel=1e-12)


def test_power_law_temperature_table_is_interpolated_exactly():
    wav = np.array([10e-6, 20e-6, 100e-6])
    Rmin, Rmax, n_r = 0.5, 400.0, 100
    Rg = Rmin * (Rmax / Rmin) ** (np.arange(n_r) / (n_r - 1.0))
    Td = (300.0 * Rg ** -0.5)[None, :]
    t = tables(wav, Td=Td, Rmin=Rmin, Rmax=Rmax)
    for R in (0.73, 3.3, 31.0, 77.7, 399.0):
        F = O.sed(t, View(dist=DIST), grains([[R, 0.0, 0.0]]))
        # invert the Planck function at each wavelength for the grain's temperature
        s = 1e-5
        B = F * C / (1e26 * wav ** 2) * DIST ** 2 / (s * s) / math.pi
        T = H * C / (wav * K) / np.log1p(2 * H * C * C / (wav ** 5 * B))
        assert np.allclose(T, 300.0 * R ** -0.5, rtol=1e-11, atol=0)


def test_pixel_positions_follow_the_rotations():
    wav = np.array([1e-6, 2e-6])
    t = tables(wav, qabs=1.0)
    fov, nx, ny = 100.0, 50, 40      # pixel scale 2 AU (the larger side)
    pts = [[10.3, 4.1, 0.0], [-21.0, 7.9, 0.0], [3.3, -15.5, 0.0]]

    def hit(view, p):
        F = O.image(t, view, grains([p]), 0)
        k = np.flatnonzero(F)
        assert len(k) == 1
        return divmod(int(k[0]), view.ysize)    # (x, y) of pixel F[ysize * x + y]

    def pix(X, Y, view):
        ps = fov / max(view.xsize, view.ysize)
        return (math.floor((X + 0.5 * ps * view.xsize) / ps), math.floor((Y + 0.5 * ps * view.ysize) / ps))

    v = View(dist=DIST, fov=fov, xsize=nx, ysize=ny)
    for p in pts:
        assert hit(v, p) == pix(p[0], p[1], v)
    # longitude: the system turns counter-clockwise about z
    v = View(dist=DIST, fov=fov, xsize=nx, ysize=ny, rot=math.pi / 2)
    for p in pts:
        assert hit(v, p) == pix(-p[1], p[0], v)
    # inclination: the sky y is y cos i for grains in the plane
    v = View(dist=DIST, fov=fov, xsize=nx, ysize=ny, incl=math.radians(60.0))
    for p in pts:
        assert hit(v, p) == pix(p[0], p[1] * 0.5, v)
    # position angle: the sky plane turns counter-clockwise
    v = View(dist=DIST, fov=fov, xsize=nx, ysize=ny, PA=math.pi)
    for p in pts:
        assert hit(v, p) == pix(-p[0], -p[1], v)
    # outside the field: no pixel
    F = O.image(t, View(dist=DIST, fov=fov, xsize=nx, ysize=ny), grains([[0.0, 45.0, 0.0]]), 0)
    assert not F.any()


def test_image_total_equals_sed_when_every_grain_is_in_the_field():
    rng = np.random.default_rng(5)
    n = 3000
    r = rng.uniform(5.0, 40.0, n)
    ph = rng.uniform(0.0, 2 * math.pi, n)
    pos = np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(n)], axis=1)
    nb, wav = 8, np.geomspace(1e-6, 1e-3, 12)
    sizes = np.geomspace(1e-6, 1e-4, nb)
    mu = np.linspace(-1, 1, 21)
    qpf = np.exp(0.7 * mu)[None, :, None] * np.ones((nb, 1, len(wav)))
    Rg = 0.1 * (1000.0 / 0.1) ** (np.arange(100) / 99.0)
    t = OpticsTables(w=wav, in_=np.linspace(1.0, 2.0, len(wav)), size=sizes,
                     Qabs=np.linspace(0.2, 1.0, nb * len(wav)).reshape(nb, -1),
                     Qsca=np.linspace(0.1, 0.5, nb * len(wav)).reshape(nb, -1), Qpfunc=qpf,
                     Td=np.outer(np.linspace(0.8, 1.2, nb), 300.0 * Rg ** -0.5), Rmin=0.1, Rmax=1000.0)
    st = grains(pos, ids=rng.integers(0, nb, n), N=rng.uniform(1.0, 1e20, n))
    v = View(dist=DIST, fov=100.0, xsize=64, ysize=48, incl=0.4, PA=0.3, rot=1.1)
    sed = O.sed(t, v, st)
    for wj in (0, 5, 11):
        F = O.image(t, v, st, wj)
        assert F.sum() == pytest.approx(sed[wj], rel=1e-13)


def test_grains_that_add_nothing():
    wav = np.array([1e-6, 2e-6])
    t = tables(wav, n_bins=2, qabs=1.0, qsca=0.5, Rmin=1.0)
    v = View(dist=DIST, fov=100.0, xsize=10, ysize=10)
    base = grains([[20.0, 0.0, 0.0]])
    ref_img, ref_sed = O.image(t, v, base, 0), O.sed(t, v, base)
    assert ref_img.sum() > 0 and ref_sed[0] > 0
    for st in (grains([[20.0, 0.0, 0.0]], ids=[-1]), grains([[20.0, 0.0, 0.0]], ids=[2]),
               grains([[0.9, 0.0, 0.0]]), grains([[20.0, 0.0, 0.0]], N=[0.0])):
        assert not O.image(t, v, st, 0).any()
        assert not O.sed(t, v, st).any()
    # adding them to a grain that emits changes nothing
    st = grains([[20.0, 0.0, 0.0], [20.0, 1.0, 0.0], [0.5, 0.0, 0.0]], ids=[0, -1, 0])
    assert np.array_equal(O.image(t, v, st, 0), ref_img)

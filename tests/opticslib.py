"""Optical-data and spectrum files written for the tests, in the formats the
reference reads (src/read_in.cu:377-506), and the closed forms that pin what
the host set-up computes from them.  No optical data ships with the
reference (its examples name an optical/astrosil.dat that is absent)."""
import math

import numpy as np

H, C, K = 6.626068e-34, 299792458.0, 1.3806503e-23      # src/ddyn.cuh:23-25
SIGMA_SB = 2.0 * math.pi ** 5 * K ** 4 / (15.0 * H ** 3 * C ** 2)
LSUN_W = 3.827e26


def planck(T, wvl):
    """B_lambda [W m^-3 sr^-1]."""
    return 2 * H * C * C / (wvl ** 5 * (np.exp(H * C / (wvl * K * T)) - 1.0))


def write_spectrum(path, T, wmin=1e-8, wmax=2e-3, n=4000):
    """Surface flux pi B_lambda(T) of a blackbody star, wavelengths in m."""
    w = np.geomspace(wmin, wmax, n)
    with open(path, "w") as f:
        f.write(f"# blackbody surface flux, T = {T} K: wavelength [m]  F_lambda [W m^-3]\n")
        for wi, Fi in zip(w, math.pi * planck(T, w)):
            f.write(f"{wi:.10e} {Fi:.10e}\n")


def write_composition(path, sizes_um, wav_um, qabs, qsca, phase, dmin_um=None, dmax_um=None):
    """Composition table: qabs(s, w), qsca(s, w) give (n_sizes, n_wav) arrays;
    phase(s, mu) gives (n_sizes, n_theta, n_wav) values of the phase function
    at mu = cos(theta) on n_theta points from -1 to 1."""
    sizes_um, wav_um = np.asarray(sizes_um, float), np.asarray(wav_um, float)
    qa, qs, pf = np.asarray(qabs, float), np.asarray(qsca, float), np.asarray(phase, float)
    ns, nw = qa.shape
    nt = pf.shape[1]
    assert (nt * nw) % 10 == 0, "the phase function is read ten values per line"
    dmin_um = sizes_um[0] if dmin_um is None else dmin_um
    dmax_um = sizes_um[-1] if dmax_um is None else dmax_um
    lines = ["# grain optical properties written for the tests",
             f"# {nw} {nt} {ns} {dmin_um:.10g} {dmax_um:.10g}"]
    for i, s in enumerate(sizes_um):
        lines.append(f"# {s:.10g}")
        for j, w in enumerate(wav_um):
            lines.append(f"{w:.10e} {qa[i, j]:.10e} {qs[i, j]:.10e}")
        flat = pf[i].reshape(-1)
        for k in range(0, len(flat), 10):
            lines.append(" ".join(f"{v:.10e}" for v in flat[k:k + 10]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def gray_table(path, sizes_um, wav_um, qabs=1.0, qsca=0.0, g_lin=0.0, n_theta=21, qabs_pow=0.0):
    """Grains with Qabs = qabs (s / s0)^qabs_pow, Qsca = qsca at every
    wavelength, and a phase function 1 + g_lin mu (asymmetry parameter
    g = g_lin / 3)."""
    s, w = np.asarray(sizes_um, float), np.asarray(wav_um, float)
    qa = qabs * (s[:, None] / s[0]) ** qabs_pow * np.ones((len(s), len(w)))
    qs = qsca * np.ones((len(s), len(w)))
    mu = np.linspace(-1.0, 1.0, n_theta)
    pf = np.broadcast_to((1.0 + g_lin * mu)[None, :, None], (len(s), n_theta, len(w)))
    write_composition(path, s, w, qa, qs, pf)


def trapezoid_g(g_lin, n_theta):
    """The asymmetry parameter of phase function 1 + g_lin mu by the
    trapezoid rule on n_theta points, in closed form: int mu (1 + g mu) and
    int (1 + g mu) over [-1, 1] are 2 g (1/3 + h^2/6) / 2 with h = 2/(n-1)
    (the rule integrates mu^2 to 2/3 + h^2/3 exactly)."""
    h = 2.0 / (n_theta - 1)
    return g_lin * (2.0 / 3.0 + h * h / 3.0) / 2.0


PARAM_SETUP = """# set-up test system
1
0
no
1.0 1.0 1.0
{t_end}
200
# star: M sptype Teff R B P mloss vsw
1.0 G2 5800 1.0 1.0 30.0 2e-14 450
10.0 100.0 0.0 0.0 0.0 64 64
0.5 1.0
{spec}
{physics}
begin_planets
1.30530 5.20336301 0.04839266 21.62746 100.55615 14.75385 9.54240000e-4
end_planets
2.5 {comp} 27.5
begin_disks
{disks}
end_disks
begin_blobs
{blobs}
end_blobs
setupstub
"""


def setup_param(path, comp, spec, disks, blobs="", physics="yes yes no", t_end=1.0):
    path.write_text(PARAM_SETUP.format(comp=comp, spec=spec, disks=disks, blobs=blobs, physics=physics, t_end=t_end))
    return path

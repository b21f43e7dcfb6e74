"""The reference's own Fomalhaut example (tests/Fomalhaut.param of
merope82/DiskDyn v1.01, with its refractive indices optical/astrosil.lnk and
the Kurucz spectrum its "Library" spectral type selects), copied unchanged
into tests/golden/reference_inputs/ by tests/golden/copy_reference_inputs.py,
set up in a scratch directory the way the reference lays out its files.

Changed for a test run: the end time (20 yr instead of 4 Myr, with data and
SED output every 20 yr and images every 10 yr) and the disk's grain tracers
(200,000 instead of 2,000,000, so the data files stay small).  Everything else
(the A0 star at 7.7 pc, the two planets of the Boley model, the 132-148 AU
disk of 2.77 um to 10 km bodies, the 1024 x 1024 images at 0.6 and 850 um)
is the reference's input."""
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "reference_inputs")


def read_lnk(path):
    L = np.loadtxt(path, comments="#").astype(np.float32).astype(np.float64)
    return L[:, 0], L[:, 1], L[:, 2]


def write_composition_from(path, sizes_um, lam_um, qabs, qsca, qpf, amin, amax):
    """A composition table in the layout of the runmie tool (qpf: (n_sizes,
    n_theta, n_wav))."""
    nl, nt = len(lam_um), qpf.shape[1]
    with open(path, "w") as f:
        f.write("#\tn_lnk\tnthetas\tndust\tamin\t\tamax\n")
        f.write(f"#\t{nl}\t{nt}\t{len(sizes_um)}\t{amin:.6e}\t{amax:.6e}\n")
        for i, s in enumerate(sizes_um):
            f.write(f"#s {s:.6e}\n")
            for j in range(nl):
                f.write(f"{lam_um[j]:.6e} {max(qabs[i, j], 0):.6e} {max(qsca[i, j], 0):.6e}\n")
            flat = np.maximum(qpf[i].reshape(-1), 0)
            for k in range(len(flat) // 10):
                f.write(" ".join(f"{v:.6e}" for v in flat[10 * k:10 * k + 10]) + " \n")


def setup(tmp, n_small=200000, t_end=20.0):
    """The run directory: Fomalhaut.param (changed as above), Kurucz/, JWST/
    and optical/ (the composition table is left to the caller)."""
    os.makedirs(os.path.join(tmp, "optical"), exist_ok=True)
    os.makedirs(os.path.join(tmp, "Kurucz"), exist_ok=True)
    os.makedirs(os.path.join(tmp, "JWST"), exist_ok=True)
    shutil.copyfile(os.path.join(INPUTS, "Kurucz_8500.dat"), os.path.join(tmp, "Kurucz", "Kurucz_8500.dat"))
    text = open(os.path.join(INPUTS, "Fomalhaut.param")).read()
    for old, new in (("\n100000.0 10000.0 100000.0\n", f"\n{t_end} {t_end / 2} {t_end}\n"),
                     ("\n4000000\n", f"\n{t_end}\n"),
                     ("10.0 2.77 3.66 4.0 -1.0 0 2000000\n", f"10.0 2.77 3.66 4.0 -1.0 0 {n_small}\n")):
        assert text.count(old) == 1, old
        text = text.replace(old, new)
    p = os.path.join(tmp, "Fomalhaut.param")
    with open(p, "w") as f:
        f.write(text)
    return p

"""CPU set-up of the reference's own Fomalhaut example (tests/reference_example.py):
its .param file with the "Library" spectral type, the Kurucz spectrum that
selects, a composition table made from its astrosil.lnk refractive indices by
the oracle's Mie restatement, the dust grid, the grain tables, the imaging
tables and the particle counts.  What pins them: the Stefan-Boltzmann
luminosity of the star (the Kurucz flux of an 8,500 K, 1.805 Rsun star, 5%),
the reference's rounding of Teff to the library grid, the image wavelength
rule, and grain temperatures that fall with distance."""
import math
import os

import numpy as np
import pytest

from diskdyn_amd import hostio
from diskdyn_amd.units import AU, RSUN
from oracle import oracle as O
from tests import opticslib as OL
from tests import reference_example as RE


@pytest.fixture
def run_dir(tmp_path, monkeypatch):
    p = RE.setup(str(tmp_path))
    lam, n, k = RE.read_lnk(os.path.join(RE.INPUTS, "astrosil.lnk"))
    sizes = np.geomspace(0.1, 500.0, 10).astype(np.float32).astype(np.float64)
    cost = np.linspace(-1.0, 1.0, 20).astype(np.float32).astype(np.float64)
    tabs = [O.mie(s, lam, n, k, cost) for s in sizes]
    RE.write_composition_from(tmp_path / "optical" / "astrosil.dat", sizes, lam,
                              np.array([t[0] for t in tabs]), np.array([t[1] for t in tabs]),
                              np.array([t[2] for t in tabs]), 0.1, 500.0)
    monkeypatch.chdir(tmp_path)
    return p


def test_fomalhaut_example_sets_up(run_dir):
    cfg = hostio.read_param(run_dir)
    assert cfg.n_planets == 3 and cfg.n_disks == 1 and cfg.n_blobs == 0
    assert cfg.teff == 8500.0 and cfg.specfile.decode() == "Kurucz/Kurucz_8500.dat"
    assert cfg.xsize == 1024 and cfg.ysize == 1024 and cfg.nwavs == 2
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    info = opt.info()
    assert info["n_wav"] == 1000 and info["n_theta"] == 20 and info["n_sizes"] == 10
    hostio.set_dust_grid(cfg, opt)
    tab = hostio.grain_tables(cfg, opt)
    L_sb = 4 * math.pi * (1.805 * RSUN) ** 2 * OL.SIGMA_SB * 8500.0 ** 4 / OL.LSUN_W
    assert abs(tab["lum"] / L_sb - 1.0) < 0.05
    assert np.all(tab["beta"] > 0) and tab["beta"][0] > tab["beta"][-1]
    assert 0.0 < tab["Rmin"] < 1.0
    t, wj = hostio.grid_optics(cfg, opt)
    for j, want in zip(wj, (0.6e-6, 850e-6)):
        assert t.w[j] < want <= t.w[j + 1]
    assert np.all(np.diff(t.Td, axis=1) < 0) and t.Td.min() > 4.0 and t.Td.max() <= 4000.0
    st = hostio.generate(cfg, opt, tab["beta"])
    # the disk's tracers: an equal number per mass bin, about the 200,000 asked for
    assert st.n_grav == 3 and abs(st.n_dust - 200000) < 1000
    r = np.hypot(st.x[3:], st.y[3:])
    assert np.median(r) > 100 and np.median(r) < 180

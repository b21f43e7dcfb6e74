"""The reference's own Fomalhaut example end to end on the GPU, in the
reference's workflow (tests/reference_example.py): runmie turns its
astrosil.lnk refractive indices into optical/astrosil.dat; ddyn -i
Fomalhaut.param (its "Library" spectrum, two planets, a disk of 200,000 grain
tracers, 1024 x 1024 images at 0.6 and 850 um) runs to t = 20 yr and writes
its data files, FITS images and SEDs under JWST/.  Checked: the first data
file holds the generated state (host generator, same seed); the last one the
oracle's state stepped with the same dt sequence, grain tables and
sublimation radius (to the file's %.8e); the initial images the oracle's
images of the generated state, pixel for pixel (1e-12)."""
import os
import subprocess

import numpy as np
import pytest

from diskdyn_amd import hostio
from diskdyn_amd.imaging import image_to_fits_rows, view_from_config
from oracle import oracle as O
from tests import reference_example as RE
from tests.test_gpu_cli import dt_sequence
from tests.test_imaging_host import _read_fits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNMIE = os.path.join(ROOT, "diskdyn_amd", "runmie")
DDYN = os.path.join(ROOT, "diskdyn_amd", "ddyn")


def test_fomalhaut_example_end_to_end(tmp_path, monkeypatch):
    p = RE.setup(str(tmp_path))
    r = subprocess.run([RUNMIE, "-Ntheta", "20", "-Ndust", "24", "-amin", "0.1", "-amax", "2000", "-lnkfile",
                        os.path.join(RE.INPUTS, "astrosil.lnk"), "-output", str(tmp_path / "optical" / "astrosil.dat"),
                        "--quiet"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([DDYN, "-i", p], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    monkeypatch.chdir(tmp_path)
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    hostio.set_dust_grid(cfg, opt)
    tab = hostio.grain_tables(cfg, opt)
    tables, wj = hostio.grid_optics(cfg, opt)
    st0 = hostio.generate(cfg, opt, tab["beta"])
    stub = "JWST/Fomalhaut2"
    _, run0, f0 = hostio.read_datafile(f"{stub}_0.00.dat")
    assert run0.n_dust == st0.n_dust > 190000
    assert np.abs(f0.pos() - st0.pos()).max() <= 1e-8 * np.abs(st0.pos()).max()
    # the final state against the oracle stepped with the same dt sequence
    st0.beta_t, st0.betasw_t, st0.q_t = tab["beta"], tab["betasw"], tab["q"]
    prm = hostio.params_from_config(cfg, Rmin=tab["Rmin"])
    seq = dt_sequence(cfg.t_write, cfg.t_fits, cfg.t_sed, cfg.t_end, hostio.det_step(cfg, st0))
    ref = st0.copy()
    for d in seq:
        O.step(prm, ref, d, 1)
    _, run, fin = hostio.read_datafile(f"{stub}_20.00.dat")
    assert np.array_equal(fin.N == 0, ref.N == 0)
    err = np.abs(fin.pos() - ref.pos()).max() / np.abs(ref.pos()).max()
    print(f"[fomalhaut] {len(seq)} steps, {st0.n_dust} grains, final position error {err:.2e}")
    assert err < 1e-8
    # images of the initial state, and the later outputs
    view = view_from_config(cfg)
    for j in wj:
        name = f"{stub}_{tables.w[j] * 1e6:.2f}_0.000.fits"
        cards, data = _read_fits(name)
        want = image_to_fits_rows(O.image(tables, view, st0, j), view.xsize, view.ysize)
        nz = want > 0
        assert nz.sum() > 1000 and np.array_equal(data > 0, nz)
        assert np.abs(data[nz] / want[nz] - 1.0).max() <= 1e-12
        for t in ("10.000", "20.000"):
            assert os.path.exists(f"{stub}_{tables.w[j] * 1e6:.2f}_{t}.fits")
    for t in ("0.00", "20.00"):
        sed = np.loadtxt(f"{stub}_SED_{t}.dat")
        assert sed.shape == (1000, 3) and sed[:, 2].max() > 0

"""CPU tests of the host side of images and SEDs (libddyn_host.so): the dust
grid's optical tables with the grain temperature table (ddyn_grid_optics_make,
restating calculate_dust_temp, src/grid.cu:754-834), the image wavelength rule
(src/functions.cu:329-335), and the FITS and SED writers (src/fits.cu:9-85,
src/grid.cu:675-691).  What pins them: the blackbody equilibrium temperature
T = T_star sqrt(R_star / 2R) of grains with Qabs = 1 (a closed form the
table must follow at every radius), and the FITS standard's layout read back
by an independent reader."""
import math
import os

import numpy as np
import pytest

from diskdyn_amd import hostio
from diskdyn_amd.imaging import image_to_fits_rows
from diskdyn_amd.units import AU, PC, RSUN
from tests import opticslib as OL

WAV_UM = np.geomspace(0.05, 2000.0, 800)
SIZES_UM = np.geomspace(0.1, 2000.0, 14)
DISK = "30.0 50.0 0.05 0.05 100.0 1.0 3.67 0.1 -1.0 0 0"


def _setup(tmp_path, qabs=1.0, qsca=0.0):
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "bb.dat"
    OL.gray_table(comp, SIZES_UM, WAV_UM, qabs=qabs, qsca=qsca, n_theta=21)
    p = OL.setup_param(tmp_path / "run.param", comp, spec, DISK)
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    hostio.set_dust_grid(cfg, opt)
    return cfg, opt


def test_temperature_table_of_blackbody_grains(tmp_path):
    cfg, opt = _setup(tmp_path)
    t, wj = hostio.grid_optics(cfg, opt)
    assert t.n_r == 100 and t.n_bins == cfg.n_grid and t.n_wav == len(WAV_UM) and t.n_theta == 21
    # the radial range: the sublimation radius and twice the outer radius
    assert t.Rmin == pytest.approx(0.5 * RSUN * (5800.0 / 2000.0) ** 2 / AU, rel=2e-4)
    assert t.Rmax == pytest.approx(hostio.outer_radius(cfg), rel=1e-14)
    R = t.Rmin * (t.Rmax / t.Rmin) ** (np.arange(t.n_r) / (t.n_r - 1.0))
    T_bb = 5800.0 * np.sqrt(RSUN / (2.0 * R * AU))
    assert np.allclose(t.Td, T_bb[None, :], rtol=3e-4, atol=0)
    # the grain tables' sizes, and Qabs = 1 on the grid
    g = hostio.grain_tables(cfg, opt)
    assert np.array_equal(t.size, g["size"])
    assert np.all(t.Qabs == 1.0) and np.all(t.Qsca == 0.0)
    assert np.all(t.in_ >= 0.0) and t.in_.max() > 0


def test_image_wavelength_rule(tmp_path):
    cfg, opt = _setup(tmp_path)
    t, wj = hostio.grid_optics(cfg, opt)
    assert cfg.nwavs == 2
    for k, want in enumerate((0.5e-6, 1.0e-6)):
        i = int(np.argmax(t.w >= want))        # first optical wavelength at or above the request
        assert wj[k] == i - 1 and t.w[wj[k]] < want <= t.w[wj[k] + 1]
    cfg.wavs[0] = t.w[0]
    assert hostio.grid_optics(cfg, opt)[1][0] == 0
    cfg.wavs[0] = 0.5 * t.w[0]
    with pytest.raises(hostio.HostIOError) as e:
        hostio.grid_optics(cfg, opt)
    assert e.value.code == 43


def _read_fits(path):
    """An independent reader of a primary FITS array: 80-character cards in
    2880-byte blocks up to END, then big-endian data."""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    cards, pos = {}, 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            val = card[10:].split(" / ")[0].strip()
            cards[key] = val.strip("'").strip() if val.startswith("'") else val
        else:
            cards.setdefault(key, []).append(card[8:].strip())
    data_start = (pos + 2879) // 2880 * 2880
    nx, ny = int(cards["NAXIS1"]), int(cards["NAXIS2"])
    assert int(cards["BITPIX"]) == -64 and int(cards["NAXIS"]) == 2 and cards["SIMPLE"] == "T"
    data = np.frombuffer(raw[data_start:data_start + nx * ny * 8], dtype=">f8").reshape(ny, nx)
    assert not any(raw[data_start + nx * ny * 8:])
    return cards, data


def test_fits_file_layout(tmp_path):
    cfg, _ = _setup(tmp_path)
    cfg.xsize, cfg.ysize = 7, 5
    cfg.name_stub = str(tmp_path / "img").encode()
    nx, ny = 7, 5
    F = np.array([x + 10.0 * y + 0.25 for x in range(nx) for y in range(ny)])   # pixel (x, y) at F[ny*x + y]
    name = hostio.write_image(cfg, 12.3456, 0.5e-6, F)
    assert name == str(tmp_path / "img") + "_0.50_12.346.fits" and os.path.exists(name)
    cards, data = _read_fits(name)
    assert int(cards["NAXIS1"]) == nx and int(cards["NAXIS2"]) == ny
    for x in range(nx):
        for y in range(ny):
            assert data[y, x] == x + 10.0 * y + 0.25
    assert np.array_equal(data, image_to_fits_rows(F, nx, ny))
    assert float(cards["FLUX"]) == pytest.approx(F.sum(), rel=1e-14)
    assert float(cards["WAVELENG"]) == pytest.approx(0.5, rel=1e-7)
    assert float(cards["PIXELSCA"]) == pytest.approx(cfg.fov / (cfg.dist / PC) / max(nx, ny), rel=1e-6)
    assert cards["VERSION"] == "1.01" and cards["HISTORY"] == ["Created with DiskDyn"] and "DATE" in cards
    # an existing file is replaced
    hostio.write_image(cfg, 12.3456, 0.5e-6, 2.0 * F)
    assert np.array_equal(_read_fits(name)[1], 2.0 * image_to_fits_rows(F, nx, ny))


def test_sed_file_layout(tmp_path):
    cfg, opt = _setup(tmp_path)
    t, _ = hostio.grid_optics(cfg, opt)
    cfg.name_stub = str(tmp_path / "run").encode()
    Fexc = np.linspace(1e-3, 2.0, t.n_wav)
    name = hostio.write_sed(cfg, 3.5, t, Fexc)
    assert name == str(tmp_path / "run") + "_SED_3.50.dat"
    lines = open(name).read().splitlines()
    assert lines[0] == "#Calculated with DiskDyn Version 1.01"
    assert lines[1] == "#Wav[mu]\tFlux_star[Jy]\tFlux_exc[Jy]"
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[2:]])
    assert rows.shape == (t.n_wav, 3)
    assert np.allclose(rows[:, 0], t.w * 1e6, rtol=0, atol=6e-4)
    assert np.allclose(rows[:, 1], t.in_ * 1e26 * t.w ** 2 / 299792458.0, rtol=6e-7, atol=0)
    assert np.allclose(rows[:, 2], Fexc, rtol=6e-7, atol=0)
    assert lines[2] == "%.3f %.6e %.6e" % (t.w[0] * 1e6, t.in_[0] * 1e26 * t.w[0] ** 2 / 299792458.0, Fexc[0])

"""End-to-end drop-in run of the ddyn driver on the GPU: .param in, data files
out, the evolve loop's step clipping, and a -m restart from the written file.
The final state is checked against the oracle stepped with the same dt sequence."""
import os
import subprocess

import numpy as np
import pytest

from diskdyn_amd import hostio
from diskdyn_amd.system import Params
from oracle import oracle as O
from tests.test_hostio import PARAM

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DDYN = os.path.join(ROOT, "diskdyn_amd", "ddyn")

# the test system without disks/blobs that would need particle generation
PARAM_NO_DISKS = PARAM.replace("30.0 50.0 0.05 0.05 1.0 1.0 -3.5 0.1 -1.0 0 0", "30.0 50.0 0.05 0.05 1.0 1.0 -3.5 0.1 -1.0 0 0")


def dt_sequence(t_write, t_fits, t_sed, t_end, dt, t0=0.0, tl=(0.0, 0.0, 0.0)):
    """The step sizes of Grid::evolve's clipping loop (src/grid.cu:378-438)."""
    time, (td, tf, ts) = t0, tl
    seq = []

    def clip(d):
        if time + d > td + t_write:
            d = t_write + td - time
        if time + d > tf + t_fits:
            d = t_fits + tf - time
        if time + d > ts + t_sed:
            d = t_sed + ts - time
        if time + d > t_end:
            d = t_end - time
        return d

    d = clip(dt)
    while True:
        seq.append(d)
        time += d
        reset = False
        if time >= ts + t_sed or time >= t_end:
            ts, reset = time, True
        if time >= tf + t_fits or time >= t_end:
            tf, reset = time, True
        if time >= td + t_write or time >= t_end:
            td, reset = time, True
        if reset:
            d = dt
        d = clip(d)
        if not time < t_end:
            break
    return seq


def test_param_run_and_restart(tmp_path):
    p = tmp_path / "sys.param"
    p.write_text(PARAM_NO_DISKS.replace("20.0\n200\n", "2.0\n40\n").replace("5.0 10.0 10.0", "0.5 0.75 1.0"))
    r = subprocess.run([DDYN, "-i", str(p), "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    outs = sorted(f for f in os.listdir(tmp_path) if f.startswith("teststub_"))
    assert "teststub_0.00.dat" in outs and "teststub_2.00.dat" in outs and "teststub_0.50.dat" in outs
    cfg = hostio.read_param(p)
    st = hostio.state_from_config(cfg)
    dt = hostio.det_step(cfg, st)
    seq = dt_sequence(cfg.t_write, cfg.t_fits, cfg.t_sed, cfg.t_end, dt)
    prm = Params(n_planets=cfg.n_planets)
    ref = st.copy()
    for d in seq:
        O.step(prm, ref, d, 1)
    _, run, fin = hostio.read_datafile(tmp_path / "teststub_2.00.dat")
    assert abs(run.time - 2.0) < 1e-12
    err = np.abs(fin.pos() - ref.pos()).max() / np.abs(ref.pos()).max()
    assert err < 1e-8          # file precision %.8e
    # restart from the middle file and run to the end
    (tmp_path / "restart").mkdir()
    r = subprocess.run([DDYN, "-m", str(tmp_path / "teststub_1.00.dat"), "--quiet"], cwd=tmp_path / "restart",
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    _, run2, fin2 = hostio.read_datafile(tmp_path / "restart" / "teststub_2.00.dat")
    assert np.abs(fin2.pos() - fin.pos()).max() < 1e-6 * np.abs(fin.pos()).max()


def test_param_with_particles_needs_its_optical_data(tmp_path):
    """Particles to generate need the optical data the .param names
    (optical/astrosil.dat here, absent): refused with the reference's code 39."""
    p = tmp_path / "d.param"
    p.write_text(PARAM.replace("-3.5 0.1 -1.0 0 0", "-3.5 0.1 -1.0 0 100"))
    r = subprocess.run([DDYN, "-i", str(p), "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "optical data" in r.stderr


def test_precision_flags_without_swarm(tmp_path):
    """Star and planets only: the FP32-force form is the plain FP64
    computation there, so --fp32 writes the default run's file byte for byte;
    --fp64-rsq (refined reciprocal square roots in every pair) stays within
    the file's printed precision of it."""
    p = tmp_path / "sys.param"
    p.write_text(PARAM_NO_DISKS.replace("20.0\n200\n", "1.0\n40\n").replace("5.0 10.0 10.0", "1.0 1.0 1.0"))
    texts = []
    for flag in ([], ["--fp32"], ["--fp64-rsq"]):
        d = tmp_path / ("run" + "".join(flag).replace("-", "_"))
        d.mkdir()
        r = subprocess.run([DDYN, "-i", str(p), "--quiet", *flag], cwd=d, capture_output=True, text=True, timeout=120)
        print(flag, r.returncode, r.stderr)
        assert r.returncode == 0
        texts.append(d / "teststub_1.00.dat")
    assert texts[1].read_text() == texts[0].read_text()
    _, _, a = hostio.read_datafile(texts[0])
    _, _, b = hostio.read_datafile(texts[2])
    assert np.abs(b.pos() - a.pos()).max() <= 1e-8 * np.abs(a.pos()).max()


def _write_optics(tmp_path):
    from tests import opticslib as OL
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    comp = tmp_path / "gray.dat"
    OL.gray_table(comp, np.geomspace(0.1, 2000.0, 14), np.geomspace(0.05, 2000.0, 400), qabs=0.7, qsca=0.4,
                  g_lin=0.6, n_theta=21)
    return comp, spec


def test_param_run_generates_disks_and_blobs(tmp_path):
    """A .param run whose disk and blob generate particles: 200 gravitating
    disk bodies (the reference law, no softening), 3,000 disk grains and 1,000
    blob grains from a collision, with per-bin beta, beta_sw from the optical
    data, wind drag on.  The first data file holds exactly the generated state
    (host generator, same seed); the final file matches the oracle stepped from
    it with the same dt sequence, bin tables and sublimation radius; a -m
    restart from the middle file, which rebuilds the tables from the optical
    data and the bins from the masses, reaches the same final state."""
    from tests import opticslib as OL
    comp, spec = _write_optics(tmp_path)
    disks = "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.5 -1.0 200 3000"
    blobs = "35.0 0.1 5.0 10.0 20.0 30.0 100.0 1.0 3.67 0.1 0 1000 collision 3.0"
    p = OL.setup_param(tmp_path / "gen.param", comp, spec, disks, blobs, physics="yes yes no", t_end=1.0)
    p.write_text(p.read_text().replace("1.0 1.0 1.0\n", "0.5 0.5 0.5\n", 1))
    r = subprocess.run([DDYN, "-i", str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "generated 200 gravitating bodies" in r.stdout
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    hostio.set_dust_grid(cfg, opt)
    tab = hostio.grain_tables(cfg, opt)
    st0 = hostio.generate(cfg, opt, tab["beta"])
    assert st0.n_grav == cfg.n_planets + 200 and st0.n_dust > 3500
    _, run0, f0 = hostio.read_datafile(tmp_path / "setupstub_0.00.dat")
    assert run0.n_mass == 200 and run0.n_dust == st0.n_dust
    assert np.abs(f0.pos() - st0.pos()).max() <= 1e-8 * np.abs(st0.pos()).max()
    assert np.array_equal(f0.id[st0.n_grav:], st0.id[st0.n_grav:])
    # the oracle with the same tables, dt sequence and sublimation radius
    st0.beta_t, st0.betasw_t, st0.q_t = tab["beta"], tab["betasw"], tab["q"]
    prm = hostio.params_from_config(cfg, Rmin=tab["Rmin"])
    dt = hostio.det_step(cfg, st0)
    seq = dt_sequence(cfg.t_write, cfg.t_fits, cfg.t_sed, cfg.t_end, dt)
    ref = st0.copy()
    for d in seq:
        O.step(prm, ref, d, 1)
    _, run, fin = hostio.read_datafile(tmp_path / "setupstub_1.00.dat")
    assert abs(run.time - 1.0) < 1e-12
    assert np.array_equal(fin.N == 0, ref.N == 0)
    err = np.abs(fin.pos() - ref.pos()).max() / np.abs(ref.pos()).max()
    print(f"[cli gen] {len(seq)} steps, final position error {err:.2e} (file precision %.8e)")
    assert err < 1e-8
    (tmp_path / "restart").mkdir()
    r = subprocess.run([DDYN, "-m", str(tmp_path / "setupstub_0.50.dat"), "--quiet"], cwd=tmp_path / "restart",
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    _, _, fin2 = hostio.read_datafile(tmp_path / "restart" / "setupstub_1.00.dat")
    assert np.abs(fin2.pos() - fin.pos()).max() < 1e-6 * np.abs(fin.pos()).max()


def test_param_run_writes_images_and_seds(tmp_path):
    """The same kind of run writes, at t = 0 and at every FITS and SED output
    time, one FITS image per image wavelength and one SED file.  The initial
    image is the oracle's image of the generated state (with the host's grid
    optics and the .param observer) pixel for pixel, the initial SED its SED
    to the file's %.6e; the final images carry the total flux of the oracle's
    image of the final data file's state; --no-images writes data files only."""
    from diskdyn_amd.imaging import image_to_fits_rows, view_from_config
    from tests import opticslib as OL
    from tests.test_imaging_host import _read_fits
    comp, spec = _write_optics(tmp_path)
    disks = "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.5 -1.0 50 3000"
    blobs = "35.0 0.1 5.0 10.0 20.0 30.0 100.0 1.0 3.67 0.1 0 1000 collision 3.0"
    p = OL.setup_param(tmp_path / "img.param", comp, spec, disks, blobs, physics="yes yes no", t_end=1.0)
    p.write_text(p.read_text().replace("1.0 1.0 1.0\n", "0.5 0.5 0.5\n", 1))
    r = subprocess.run([DDYN, "-i", str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    hostio.set_dust_grid(cfg, opt)
    tab = hostio.grain_tables(cfg, opt)
    tables, wj = hostio.grid_optics(cfg, opt)
    view = view_from_config(cfg)
    assert len(wj) == 2
    st0 = hostio.generate(cfg, opt, tab["beta"])
    for t in (0.0, 0.5, 1.0):
        assert os.path.exists(tmp_path / f"setupstub_SED_{t:.2f}.dat")
        for j in wj:
            assert os.path.exists(tmp_path / f"setupstub_{tables.w[j] * 1e6:.2f}_{t:.3f}.fits")
    for j in wj:
        cards, data = _read_fits(tmp_path / f"setupstub_{tables.w[j] * 1e6:.2f}_0.000.fits")
        ref = image_to_fits_rows(O.image(tables, view, st0, j), view.xsize, view.ysize)
        assert np.array_equal(data > 0, ref > 0) and (ref > 0).sum() > 100
        nz = ref > 0
        assert np.abs(data[nz] / ref[nz] - 1.0).max() <= 1e-12
        assert float(cards["FLUX"]) == pytest.approx(ref.sum(), rel=1e-12)
    rows = np.loadtxt(tmp_path / "setupstub_SED_0.00.dat")
    sed0 = O.sed(tables, view, st0)
    assert rows.shape == (tables.n_wav, 3)
    assert np.allclose(rows[:, 2], sed0, rtol=6e-7, atol=0)
    _, _, fin = hostio.read_datafile(tmp_path / "setupstub_1.00.dat")
    fin.id = st0.id
    for j in wj:
        cards, data = _read_fits(tmp_path / f"setupstub_{tables.w[j] * 1e6:.2f}_1.000.fits")
        total = O.image(tables, view, fin, j).sum()
        assert float(cards["FLUX"]) == pytest.approx(total, rel=1e-6)
    (tmp_path / "noimg").mkdir()
    r = subprocess.run([DDYN, "-i", str(p), "--no-images", "--quiet"], cwd=tmp_path / "noimg", capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0
    names = sorted(os.listdir(tmp_path / "noimg"))
    assert names == ["setupstub_0.00.dat", "setupstub_0.50.dat", "setupstub_1.00.dat"]

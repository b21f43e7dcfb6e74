"""Mie tables on the GPU (ddyn_mie through the C ABI, and the runmie tool)
against the oracle (orc_mie, restating tools/mie_single.c) and against the
reference's own tool built from its sources (oracle/_ref/runmie, when
built): efficiencies within 1e-9 relative, phase functions within 1e-9 of
their largest value, for sizes from the Rayleigh limit to size parameters of
2e4 and 1 to 4 angles per lane; the tool's file against the reference tool's
file (their single-precision output); a table made by the tool driving a
.param run with images; the error returns."""
import math
import os
import subprocess

import numpy as np
import pytest

from diskdyn_amd.engine import DDynError, Engine
from oracle import oracle as O
from tests.test_oracle_mie import read_table, synthetic_lnk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNMIE = os.path.join(ROOT, "diskdyn_amd", "runmie")
DDYN = os.path.join(ROOT, "diskdyn_amd", "ddyn")


def lnk_arrays(path):
    L = np.loadtxt(path).astype(np.float32).astype(np.float64)
    return L[:, 0], L[:, 1], L[:, 2]


@pytest.mark.parametrize("n_theta", [20, 90, 200])
def test_mie_matches_oracle(tmp_path, n_theta):
    lam, n, k = lnk_arrays(synthetic_lnk(tmp_path / "syn.lnk"))
    sizes = np.array([0.01, 0.3, 3.0, 30.0, 300.0])
    cost = np.linspace(-1.0, 1.0, n_theta)
    with Engine(0) as eng:
        qa, qs, qp = eng.mie(sizes, lam, n, k, cost)
    worst = [0.0, 0.0, 0.0]
    for i, s in enumerate(sizes):
        ra, rs, rp = O.mie(s, lam, n, k, cost)
        worst[0] = max(worst[0], np.max(np.abs(qa[i] - ra) / np.abs(ra)))
        worst[1] = max(worst[1], np.max(np.abs(qs[i] - rs) / rs))
        worst[2] = max(worst[2], np.max(np.abs(qp[i] - rp) / rp.max(axis=0)))
    print(f"[mie n_theta={n_theta}] max rel err Qabs {worst[0]:.2e} Qsca {worst[1]:.2e} phase {worst[2]:.2e}")
    assert max(worst) <= 1e-9


def test_mie_errors():
    lam = np.array([0.5, 1.0])
    nk = np.array([1.5, 1.5]), np.array([0.01, 0.01])
    with Engine(0) as eng:
        with pytest.raises(DDynError, match="code -1"):
            eng.mie([1.0], lam, *nk, np.linspace(-1, 1, 257))
        with pytest.raises(DDynError, match="code -1"):
            eng.mie([6e5], [0.5, 1e-1], *nk, np.linspace(-1, 1, 10))    # 2 pi s / lambda > 7.2e6
        with pytest.raises(DDynError, match="code -1"):
            eng.mie([0.0], lam, *nk, np.linspace(-1, 1, 10))
        qa, qs, qp = eng.mie([1.0], lam, *nk, np.linspace(-1, 1, 10))   # the context still works
        assert np.all(qs > 0)


def test_runmie_tool_writes_the_reference_tables(tmp_path):
    lnk = synthetic_lnk(tmp_path / "syn.lnk")
    args = ["-Ntheta", "30", "-Ndust", "5", "-amin", "0.05", "-amax", "40", "-lnkfile", str(lnk)]
    r = subprocess.run([RUNMIE] + args + ["-output", str(tmp_path / "ours.dat"), "--quiet"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nl, nt, ours = read_table(tmp_path / "ours.dat")
    assert (nl, nt, len(ours)) == (24, 30, 5)
    lam, n, k = lnk_arrays(lnk)
    cost = np.array([np.float32(i * 2.0 / (nt - 1) - 1.0) for i in range(nt)], dtype=np.float64)
    from tests.test_oracle_mie import runmie_sizes
    for (s, tab, pf), sz in zip(ours, runmie_sizes(0.05, 40.0, 5)):
        assert s == pytest.approx(sz, rel=6e-7)
        ra, rs, rp = O.mie(sz, lam, n, k, cost)
        assert np.allclose(tab[:, 2], rs, rtol=2e-6, atol=0)
        assert np.all(np.abs(tab[:, 1] - np.maximum(ra, 0)) <= 2e-6 * np.abs(ra) + 1e-30)
    if os.path.exists(O.REF_RUNMIE):
        subprocess.run([O.REF_RUNMIE] + args + ["-output", str(tmp_path / "ref.dat")], check=True,
                       capture_output=True, timeout=300)
        a, b = open(tmp_path / "ours.dat").read().splitlines(), open(tmp_path / "ref.dat").read().splitlines()
        assert len(a) == len(b)
        _, _, ref = read_table(tmp_path / "ref.dat")
        assert a[:2] == b[:2] and [x for x in a if x.startswith("#s")] == [x for x in b if x.startswith("#s")]
        for (s1, t1, p1), (s2, t2, p2) in zip(ours, ref):
            assert s1 == s2 and np.array_equal(t1[:, 0], t2[:, 0])
            assert np.allclose(t1[:, 2], t2[:, 2], rtol=2e-6, atol=0)
            assert np.all(np.abs(t1[:, 1] - t2[:, 1]) <= 2e-6 * t1[:, 1] + 1.2e-7 * (t1[:, 1] + t1[:, 2]))
            assert np.all(np.abs(p1 - p2) <= 2e-6 * p2 + 1e-30)
        print("[runmie] file matches the reference tool's to its single-precision output")


def test_runmie_table_drives_a_param_run_with_images(tmp_path):
    """The reference's workflow: refractive indices -> runmie -> composition
    table -> ddyn -i with disks and images."""
    from diskdyn_amd import hostio
    from tests import opticslib as OL
    lnk = synthetic_lnk(tmp_path / "syn.lnk", n_wav=60, lo=0.08, hi=3000.0)
    comp = tmp_path / "syn.dat"
    r = subprocess.run([RUNMIE, "-Ntheta", "21", "-Ndust", "12", "-amin", "0.1", "-amax", "2000", "-lnkfile", str(lnk),
                        "-output", str(comp), "--quiet"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    spec = tmp_path / "star.dat"
    OL.write_spectrum(spec, 5800.0)
    disks = "30.0 40.0 0.05 0.05 100.0 1.0 3.67 0.5 -1.0 20 2000"
    p = OL.setup_param(tmp_path / "mie.param", comp, spec, disks, "", physics="yes yes no", t_end=0.5)
    cfg = hostio.read_param(p)
    opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
    info = opt.info()
    assert info["n_wav"] == 60 and info["n_theta"] == 21 and info["n_sizes"] == 12
    r = subprocess.run([DDYN, "-i", str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert any(f.endswith(".fits") for f in os.listdir(tmp_path))
    assert os.path.exists(tmp_path / "setupstub_SED_0.50.dat") and os.path.exists(tmp_path / "setupstub_0.50.dat")
    sed = np.loadtxt(tmp_path / "setupstub_SED_0.50.dat")
    assert sed.shape == (60, 3) and np.all(np.isfinite(sed)) and sed[:, 2].max() > 0

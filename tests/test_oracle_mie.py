"""The oracle's Mie tables (orc_mie, restating tools/mie_single.c), pinned by
the closed forms and invariants of Mie theory and by the reference's own tool:

* the Rayleigh limit x -> 0: Q_sca = (8/3) x^4 |K|^2 and Q_abs = 4 x |Im K|
  with K = (m^2 - 1)/(m^2 + 2);
* a sphere that does not absorb (k = 0) has Q_abs = 0 at every size;
* the phase function integrates to 1 over the sphere;
* the extinction paradox: Q_ext -> 2 for large spheres;
* tools/runmie built from the reference's sources (oracle/_ref, when the
  reference is present) writes the same tables to its single-precision output
  (its Q_abs rounds Q_ext to single precision first: compared to 1e-7 Q_ext).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

COST = np.linspace(-1.0, 1.0, 2001)


def one(x, m, cost=COST):
    lam = 1.0
    s = x * lam / (2.0 * math.pi)
    qa, qs, qp = O.mie(s, [lam], [m.real], [-m.imag], cost)
    return qa[0], qs[0], qp[:, 0]


@pytest.mark.parametrize("m", [1.5 - 0.01j, 1.33 - 0.1j, 2.5 - 1.0j])
def test_rayleigh_limit(m):
    x = 1e-3
    qa, qs, _ = one(x, m)
    K = (m * m - 1.0) / (m * m + 2.0)
    assert qs == pytest.approx(8.0 / 3.0 * x ** 4 * abs(K) ** 2, rel=1e-4)
    assert qa == pytest.approx(4.0 * x * abs(K.imag), rel=1e-4)


@pytest.mark.parametrize("x", [0.3, 3.0, 30.0, 300.0])
def test_no_absorption_without_k(x):
    qa, qs, _ = one(x, 1.5 + 0.0j)
    assert abs(qa) <= 1e-12 * qs


@pytest.mark.parametrize("x,m", [(0.5, 1.5 - 0.01j), (2.0, 1.33 - 0.0j), (5.0, 1.7 - 0.3j)])
def test_phase_function_is_normalised(x, m):
    _, _, p = one(x, m)
    assert 2.0 * math.pi * np.trapezoid(p, COST) == pytest.approx(1.0, abs=3e-5)
    assert np.all(p > 0)


def test_extinction_paradox():
    """Q_ext -> 2 for large spheres, with the edge term ~ x^(-2/3) of the
    asymptotic theory (van de Hulst): (Q_ext - 2) x^(2/3) settles to a constant."""
    c = []
    for x in (500.0, 2000.0, 8000.0):
        qa, qs, _ = one(x, 1.5 - 0.1j, cost=np.linspace(-1, 1, 3))
        c.append((qa + qs - 2.0) * x ** (2.0 / 3.0))
        # a strongly absorbing large sphere absorbs what it does not reflect
        assert 0.8 < qa < 1.0 and 1.0 < qs < 1.2
    assert 0.5 < c[0] < 3.0 and abs(c[1] / c[0] - 1.0) < 0.1 and abs(c[2] / c[1] - 1.0) < 0.1


def synthetic_lnk(path, n_wav=24, lo=0.1, hi=500.0):
    """A smooth refractive-index table written for the tests."""
    lam = np.geomspace(lo, hi, n_wav)
    n = 1.5 + 0.8 * np.log10(lam / lo) / np.log10(hi / lo)
    k = 10.0 ** (-3.0 + 3.0 * np.sin(np.log(lam)) ** 2)
    with open(path, "w") as f:
        f.write("# refractive indices written for the tests\n# lambda (microns)    n            k\n")
        for a, b, c in zip(lam, n, k):
            f.write(f"  {a:.4f}   {b:.4E}   {c:.4E}\n")
    return path


def read_table(path):
    """A composition table as (sizes, [n_wav x 3 rows], phase values) per size."""
    blocks = open(path).read().split("#s ")
    head = blocks[0].splitlines()[1].split()
    nl, nt, nd = int(head[1]), int(head[2]), int(head[3])
    out = []
    for blk in blocks[1:]:
        rows = blk.strip().split("\n")
        s = float(rows[0])
        tab = np.array([[float(v) for v in r.split()] for r in rows[1:1 + nl]])
        pf = np.array([float(v) for r in rows[1 + nl:] for v in r.split()])
        out.append((s, tab, pf))
    assert len(out) == nd
    return nl, nt, out


def runmie_sizes(amin, amax, ndust):
    """The grain sizes of tools/runmie.c:200: single-precision amin, amax and
    result, the rest in double precision."""
    a0, a1 = np.float32(amin), np.float32(amax)
    ratio = float(np.float32(a1 / a0))
    return [float(np.float32(float(a0) * math.exp(math.log(ratio) * z / (float(np.float32(ndust)) - 1.0))))
            for z in range(ndust)]


@pytest.mark.skipif(not os.path.exists(O.REF_RUNMIE), reason="the reference's runmie is not built (oracle/_ref)")
def test_reference_tool_writes_the_oracle_tables(tmp_path):
    lnk = synthetic_lnk(tmp_path / "syn.lnk")
    out = tmp_path / "ref.dat"
    subprocess.run([O.REF_RUNMIE, "-Ntheta", "30", "-Ndust", "5", "-amin", "0.05", "-amax", "40", "-lnkfile", str(lnk),
                    "-output", str(out)], check=True, capture_output=True, timeout=300)
    nl, nt, blocks = read_table(out)
    assert (nl, nt) == (24, 30)
    cost = np.array([np.float32(i * 2.0 / (nt - 1) - 1.0) for i in range(nt)], dtype=np.float64)
    sizes = runmie_sizes(0.05, 40.0, 5)
    for (s_printed, tab, pf), s in zip(blocks, sizes):
        assert s_printed == pytest.approx(s, rel=6e-7)
        lam, qa_ref, qs_ref = tab[:, 0], tab[:, 1], tab[:, 2]
        L = np.loadtxt(lnk).astype(np.float32).astype(np.float64)
        assert np.allclose(lam, L[:, 0], rtol=6e-7)
        qa, qs, qp = O.mie(s, L[:, 0], L[:, 1], L[:, 2], cost)
        assert np.allclose(qs_ref, qs, rtol=2e-6, atol=0)
        assert np.all(np.abs(qa_ref - np.maximum(qa, 0)) <= 2e-6 * np.abs(qa) + 1.2e-7 * (qa + qs))
        want = np.maximum(qp.reshape(-1)[:len(pf)], 0)
        assert np.all(np.abs(pf - want) <= 2e-6 * want + 1e-30)

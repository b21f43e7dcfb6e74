"""Copies three input data files of the reference (merope82/DiskDyn v1.01)
into tests/golden/reference_inputs/, for the end-to-end run of its own
Fomalhaut example: the parameter file tests/Fomalhaut.param, the refractive
indices optical/astrosil.lnk (astronomical silicate, Draine & Lee 1984 with
Li & Draine 2001) and the stellar spectrum Kurucz/Kurucz_8500.dat that the
file's "Library" spectral type selects for Teff = 8590 K.  They are inputs
(data), copied unchanged; run once where the reference is present."""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_inputs")
FILES = ["tests/Fomalhaut.param", "optical/astrosil.lnk", "Kurucz/Kurucz_8500.dat"]

if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(REF, f), os.path.join(OUT, os.path.basename(f)))
        print("copied", f)

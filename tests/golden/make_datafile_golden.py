"""Expected data file of a tiny system, restated from the reference's writer.

The reference writes its data files in Grid::writegrid (src/grid.cu:441-673);
the text below follows each of its fprintf calls in order, with the same
format strings (Python's % formatting is C's printf for these conversions),
so that tests/test_hostio.py can compare the library's writer with it byte
for byte.  The system is written for this purpose: a star and a planet, one
disk and one blob (every header row has an entry), a massive body and three
grains, one of them beyond the outer radius and one inside the star (both
marked "## "), one without a size bin.

Two reference behaviours are not reproducible and are stated instead:
- the beta column of a body without a size bin (id -1) reads beta[-1] in the
  reference (src/grid.cu:668), outside its table; here it is 0;
- the second "#Number_of_dust_grids" line prints cfg->n_dust
  (src/grid.cu:496), which interpolate_dust_to_grid has set to n_grid before
  any file is written (src/functions.cu:335), so it is n_grid.

Run as a script to (re)write tests/golden/datafile_tiny.dat.
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "datafile_tiny.dat")

AU, MSUN, RSUN, YEAR, PC = 1.49598e11, 1.98892e30, 695700000.0, 3.1556926e7, 3.08567758149137e16
GRAV = 4.0 * math.pi * math.pi
PI = 3.14159265358979323846
NR = 100
VERSION = "1.01"


def tiny_system():
    """Every value the writer prints, in the reference's internal units."""
    c = dict(verb=1, gpuid=0, gl=1, t_gl=2.5, t_write=5.0, t_sed=10.0, t_fits=7.5, t_end=20.0, n_step=200,
             sptype="A0V", teff=9500.0, rstar=1.7 * RSUN / AU, Bstar=1.0e-4 * YEAR * YEAR / MSUN,
             Wz=365.2422 / 2.5, mloss=2e-14, vsw=450e3, radQ=1, SWQ=1, FLQ=0, rho=2500.0, kqe=27.5 / YEAR,
             name_stub="tiny", dist=10.0 * PC, fov=100.0, PA=30.0 * PI / 180.0, inc=60.0 * PI / 180.0,
             longitude=15.0 * PI / 180.0, xsize=128, ysize=64, specfile="Kurucz/Kurucz_9500.dat",
             amin=1.0e-6, amax=2.0e-3, mmin=1.047197551e-14, mmax=8.37758041e-5, n_grid=240, f=0.0955,
             comp="optical/astrosil.dat", n_wav=400, n_theta=21, n_planets=2, n_disks=1, n_blobs=1,
             wavs=[0.5e-6, 10.0e-6, 850.0e-6], seed=-2)
    c["disk"] = dict(Rin=30.0, Rout=50.0, dhdr=math.atan(0.05), sigmae=0.05, amax=100e3, mmax=1.047e19,
                     amin=1e-6, mmin=1.047e-14, dmax=2e-3, dmmax=8.377e-5, eta=1.89, gammas=-1.0, C=3.25e10,
                     Mtot=0.1 * 5.97e24, Ngrav=2, Nsmall=3)
    c["blob"] = dict(a=40.0, ecc=0.1, inc=5.0 * PI / 180.0, w=10.0 * PI / 180.0, o=20.0 * PI / 180.0,
                     nu=30.0 * PI / 180.0, sigma=0.5, amax=1e3, mmax=1.047e13, amin=1e-6, mmin=1.047e-14,
                     dmax=2e-3, dmmax=8.377e-5, eta=1.8833, C=1.5e8, Mtot=0.01 * 5.97e24, Ngrav=0, Nsmall=0,
                     collcloud=1, vkepp=0.0)
    c["planets"] = dict(x=[-0.0049, 5.1951], y=[0.0012, -0.0103], z=[0.0, 0.0001],
                        vx=[0.0001, -0.0123], vy=[-0.0027, 2.7551], vz=[0.0, 0.0102],
                        m=[1.5 * GRAV, 1e-3 * GRAV])
    run = dict(time=5.0, dt=0.0123456789, tlastd=5.0, tlastf=0.0, tlasts=0.0, Rmax=100.0, n_mass=1, n_dust=3)
    # [star | planet | body | grains]
    x = np.array([-0.0049, 5.1951, 35.25, 31.5, 150.0, 0.001])
    y = np.array([0.0012, -0.0103, -2.125, 4.75, 0.0, 0.0])
    z = np.array([0.0, 0.0001, 0.25, -0.125, 1.0, 0.0])
    vx = np.array([0.0001, -0.0123, 0.5, -1.0625, 0.0, 0.0])
    vy = np.array([-0.0027, 2.7551, 1.75, 1.125, 0.5, 3.0])
    vz = np.array([0.0, 0.0102, 0.0, 0.015625, 0.0, 0.0])
    N = np.array([1.0, 1.0, 1.0, 2.5e20, 1.25e19, 0.0])
    mu = np.array([1.5 * GRAV, 1e-3 * GRAV, 1e-9 * GRAV, 3.5e-41 * GRAV, 2.0e-38 * GRAV, 4.0e-39 * GRAV])
    ids = np.array([-1, -1, -1, 3, 17, -1], dtype=np.int32)
    beta_bins = (np.arange(40) * 0.01 + 0.005).astype(np.float32)
    return c, run, dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, N=N, mu=mu, id=ids), beta_bins


def expected_text(c, run, s, beta_bins):
    out = []
    w = out.append
    d, b, p = c["disk"], c["blob"], c["planets"]
    w("#\n")                                                              # grid.cu:452
    w("#Calculated_with_DiskDyn_Version %s\n" % VERSION)                  # :453
    w("#\n")
    w("#Config file parameters\n")
    w("#Verbose\t\t\t%d\n" % c["verb"])                                   # :456
    w("#GPUID\t\t\t\t%d\n" % c["gpuid"])
    w("#Graphical_Output\t\t%d\n" % c["gl"])
    w("#Time_Out_for_Graphics\t\t%.3f\n" % c["t_gl"])
    w("#Time_Out_for_Data\t\t%.3f\n" % c["t_write"])
    w("#Time_Out_for_SED\t\t%.3f\n" % c["t_sed"])                         # :461
    w("#Time_Out_for_FITS\t\t%.3f\n" % c["t_fits"])
    w("#Time_End_for_model\t\t%.3f\n" % c["t_end"])
    w("#N_step_per_orbit\t\t%d\n" % c["n_step"])
    w("#Spectral-type_of_star\t\t%s\n" % c["sptype"])
    w("#Temperature_of_star\t\t%.2f\n" % c["teff"])                       # :466
    w("#Radius_of_star\t\t\t%.5f\n" % (c["rstar"] * AU / RSUN))
    w("#Magnetic_field_of_star\t\t%.5e\n" % (c["Bstar"] * MSUN * 1.0e4 / YEAR / YEAR))
    w("#Stellar_rotation_of_star\t%.5f\n" % (365.2422 / c["Wz"]))
    w("#Stellar_mass_loss\t\t%.5e\n" % c["mloss"])
    w("#Stellar_wind_speed\t\t%.5e\n" % c["vsw"])                         # :471
    w("#Include_rad_press\t\t%d\n" % c["radQ"])
    w("#Include_SWD\t\t\t%d\n" % c["SWQ"])
    w("#Include_Lorentz\t\t%d\n" % c["FLQ"])
    w("#Bulk_density\t\t\t%.3f\n" % c["rho"])
    w("#Electric_charge_coeff\t\t%.3f\n" % (c["kqe"] * YEAR))             # :476
    w("#Name_Stub\t\t\t%s\n" % c["name_stub"])
    w("#Distance\t\t\t%.5f\n" % (c["dist"] / PC))
    w("#Field_of_View\t\t\t%.6f\n" % c["fov"])
    w("#Position_Angle\t\t\t%.5f\n" % (c["PA"] * 180.0 / PI))
    w("#Inclination_Angle\t\t%.5f\n" % (c["inc"] * 180.0 / PI))           # :481
    w("#Longitude_Angle\t\t%.5f\n" % (c["longitude"] * 180.0 / PI))
    w("#Image_x_size\t\t\t%d\n" % c["xsize"])
    w("#Image_y_size\t\t\t%d\n" % c["ysize"])
    w("#Spectral-file\t\t\t%s\n" % c["specfile"])
    w("#Minimum_particle_size\t\t%.8e\n" % c["amin"])                     # :486
    w("#Maximum_particle_size\t\t%.8e\n" % c["amax"])
    w("#Minimum_particle_mass\t\t%.8e\n" % c["mmin"])
    w("#Maximum_particle_mass\t\t%.8e\n" % c["mmax"])
    w("#Number_of_dust_grids\t\t%d\n" % c["n_grid"])
    w("#Mass_grid_ratio\t\t%.8e\n" % c["f"])                              # :491
    w("#Optical_composition_file\t%s\n" % c["comp"])
    w("#Wavelengths_in_interpolation\t%d\n" % c["n_wav"])
    w("#Number_of_thetas\t\t%d\n" % c["n_theta"])
    w("#Number_of_dust_grids\t\t%d\n" % c["n_grid"])                      # :496 (cfg->n_dust == n_grid)
    w("#Number_of_planets\t\t%d\n" % c["n_planets"])
    w("#Number_of_disks\t\t%d\n" % c["n_disks"])
    w("#Number_of_blobs\t\t%d\n" % c["n_blobs"])
    w("#Image_wavelengths\t")                                             # :500
    for v in c["wavs"]:
        w("\t%.3f" % (v * 1e6))
    w("\n")
    w("#Seed_of_model\t\t\t%d\n" % c["seed"])                             # :503

    def row(label, fmt, vals):                                            # :504-631
        w(label)
        for v in vals:
            w(fmt % v)
        w("\n")
    row("#Disk_Rin\t", "\t\t%.3f", [d["Rin"]])
    row("#Disk_Rout\t", "\t\t%.3f", [d["Rout"]])
    row("#Disk_dhdr\t", "\t\t%.4f", [d["dhdr"]])
    row("#Disk_sigmae\t", "\t\t%.4f", [d["sigmae"]])
    row("#Disk_amax\t\t", "\t%.6e", [d["amax"]])
    row("#Disk_mmax\t\t", "\t%.6e", [d["mmax"]])
    row("#Disk_amin\t\t", "\t%.6e", [d["amin"]])
    row("#Disk_mmin\t\t", "\t%.6e", [d["mmin"]])
    row("#Disk_dmax\t\t", "\t%.6e", [d["dmax"]])
    row("#Disk_dmmax\t\t", "\t%.6e", [d["dmmax"]])
    row("#Disk_eta\t", "\t\t%.4f", [d["eta"]])
    row("#Disk_gammas\t", "\t\t%.3f", [d["gammas"]])
    row("#Disk_C\t\t\t", "\t%.8e", [d["C"]])
    row("#Disk_Mtot\t\t", "\t%.8e", [d["Mtot"]])
    row("#Disk_Ngrav\t", "\t\t%d", [d["Ngrav"]])
    row("#Disk_Nsmall\t", "\t\t%d", [d["Nsmall"]])
    row("#Blob_a\t\t", "\t\t%.5f", [b["a"]])
    row("#Blob_ecc\t", "\t\t%.5f", [b["ecc"]])
    row("#Blob_inc\t", "\t\t%.5f", [b["inc"]])
    row("#Blob_w\t\t", "\t\t%.5f", [b["w"]])
    row("#Blob_o\t", "\t\t%.5f", [b["o"]])
    row("#Blob_nu\t", "\t\t%.5f", [b["nu"]])
    row("#Blob_sigma\t", "\t\t%.3f", [b["sigma"]])
    row("#Blob_amax\t\t", "\t%.6e", [b["amax"]])
    row("#Blob_mmax\t\t", "\t%.6e", [b["mmax"]])
    row("#Blob_amin\t\t", "\t%.6e", [b["amin"]])
    row("#Blob_mmin\t\t", "\t%.6e", [b["mmin"]])
    row("#Blob_dmax\t\t", "\t%.6e", [b["dmax"]])
    row("#Blob_dmmax\t\t", "\t%.6e", [b["dmmax"]])
    row("#Blob_eta\t", "\t\t%.3f", [b["eta"]])
    row("#Blob_C\t\t\t", "\t%.8e", [b["C"]])
    row("#Blob_Mtot\t\t", "\t%.8e", [b["Mtot"]])
    row("#Blob_Ngrav\t", "\t\t%d", [b["Ngrav"]])
    row("#Blob_Nsmall\t", "\t\t%d", [b["Nsmall"]])
    row("#Blob_collcloud\t", "\t\t%d", [b["collcloud"]])
    row("#Blob_vkepp\t", "\t\t%.5f", [b["vkepp"]])
    w("#\n")                                                              # :614
    w("# t=0 positions and velocities of planets and Star\n")
    row("#Planet_x\t", "\t\t%.4f", p["x"])
    row("#Planet_y\t", "\t\t%.4f", p["y"])
    row("#Planet_z\t", "\t\t%.4f", p["z"])
    row("#Planet_vx\t", "\t\t%.4f", p["vx"])
    row("#Planet_vy\t", "\t\t%.4f", p["vy"])
    row("#Planet_vz\t", "\t\t%.4f", p["vz"])
    row("#Planet_m\t\t", "\t%.6e", p["m"])
    w("#NR_used\t\t\t%d\n" % NR)                                          # :634
    w("#\n")
    w("# Remaining parameters and data are for Grid class read-in\n")
    w("#\n")
    w("#Time_of_model\t\t\t%.8f\n" % run["time"])                          # :638
    w("#dt_of_model\t\t\t%.8f\n" % run["dt"])
    w("#Last_t_of_data\t\t\t%.8f\n" % run["tlastd"])
    w("#Last_t_of_fits\t\t\t%.8f\n" % run["tlastf"])
    w("#Last_t_of_SED\t\t\t%.8f\n" % run["tlasts"])
    w("#Number_of_grav_bodies\t\t%d\n" % run["n_mass"])
    w("#Number_of_dust_particles\t%d\n" % run["n_dust"])
    w("#Number_of_all_particles\t%d\n" % len(s["x"]))
    w("#Data: x[AU] y[AU] z[AU] vx[AU/yr] vy[AU/yr] vz[AU/yr] N m[kg] r[m] beta\n")  # :648
    w("#Planets:\n")
    for i in range(c["n_planets"]):                                       # :650-653
        w("% .8e % .8e % .8e % .8e % .8e % .8e %.12e %.8e %d\n" % (
            s["x"][i], s["y"][i], s["z"][i], s["vx"][i], s["vy"][i], s["vz"][i], 1.0, s["mu"][i] * MSUN / GRAV,
            s["id"][i]))
    w("#Bodies:\n")
    for i in range(c["n_planets"], len(s["x"])):                          # :655-669
        R = s["x"][i] ** 2 + s["y"][i] ** 2 + s["z"][i] ** 2
        if R > run["Rmax"] ** 2:
            w("## ")
        if R < c["rstar"] ** 2:
            w("## ")
        idi = int(s["id"][i])
        beta = float(beta_bins[idi]) if 0 <= idi < len(beta_bins) else 0.0
        m = s["mu"][i] * MSUN / GRAV
        w("% .8e % .8e % .8e % .8e % .8e % .8e %.12e %.8e %.8e %.8e %d\n" % (
            s["x"][i], s["y"][i], s["z"][i], s["vx"][i], s["vy"][i], s["vz"][i], s["N"][i], m,
            (3.0 * m / 4.0 / PI / c["rho"]) ** (1.0 / 3.0), beta, idi))
    return "".join(out)


if __name__ == "__main__":
    text = expected_text(*tiny_system())
    with open(FIXTURE, "w") as f:
        f.write(text)
    print(f"wrote {FIXTURE} ({len(text.splitlines())} lines)")

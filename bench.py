"""Throughput of the dynamical hot path on one MI355X.

Default workload (the configuration the metric is quoted on): the dusty debris
ring, a 0.5 Msun / 0.1 Lsun star, a 3e-5 Msun planet at 20 AU, 1,024
gravitating planetesimals (1e-9 Msun, Plummer eps = 1e-3 AU) and 262,144
massless grains with radiation pressure, Poynting-Robertson and stellar-wind
drag, dt = 0.05 yr, 500 RK4 steps, in the FP32-force / FP64-state form.  The
plain FP64 form runs the same steps for the speed ratio, and so does the FP64
form with refined hardware reciprocal square roots in every pair law
(FP64_RSQ, in "fp64_rsq").  A "step" is one RK4 step (four force evaluations)
of the whole system with the state resident in HBM.  Metric: directed pairwise
interactions per second, n_grav (n_grav - 1) + n_dust n_grav per evaluation,
four evaluations per step.

The roofline entry is the dominant kernel (the grain kernel on the debris
ring, the massive swarm-pair kernel on the swarms): 20 FLOP per directed
interaction (FMA = 2, rsqrt = 1) over its mean launch time measured with HIP
events on the stream it runs on, against the FP32 vector peak.  The CPU
baseline is the FP64 oracle (OpenMP) timed on this host on a bounded sample
of whole RK4 steps.  The self-gravitating swarm (65,537 bodies), the large
disk (1,048,577 bodies) and the debris ring with 1,048,576 grains run too, at
reduced step counts, under "also", with images and SEDs of the debris ring's
grains ("also.imaging"), a table of Mie efficiencies and phase functions
("also.mie") and the reference's own Fomalhaut example at its full size of
2,000,000 grain tracers ("also.fomalhaut").

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from diskdyn_amd import synth  # noqa: E402
from diskdyn_amd.engine import DDYN_KCLASS_DUST, DDYN_KCLASS_PAIRS, DDYN_KCLASS_STAGE, Engine  # noqa: E402
from diskdyn_amd.system import FP32_FORCE, FP64, FP64_RSQ  # noqa: E402

FLOP_PER_INTERACTION = 20
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes of the state update per RK4 step.  Massive bodies (the
# four stage epilogues k_stage<0..3>): stages 1-3 read r0, v0 (48 B), read and
# write the RK4 sums (2 x 48 B) and write the stage r, v (48 B), 192 B each;
# stage 4 reads r0, v0 and the sums (96 B) and writes r, v (48 B), 144 B;
# 720 B, plus the 16 B FP32 shadow of the position and mass at each of the four
# stages.  Grains (fused into the grain kernel): read r, v (48 B), N (8 B) and
# per-grain beta, beta_sw (16 B), write r, v (48 B): 120 B.
STAGE_BYTES_PER_BODY = 720 + 4 * 16
GRAIN_BYTES_PER_GRAIN = 120


def workload(name, precision, n_dust=None):
    if name == "cfg1":
        w = synth.config1()
        desc = ("small planetary system: 1 Msun star, planets at 5.2 and 9.5 AU, "
                "4,096 massless test particles, dt=0.01 yr")
    elif name == "cfg4":
        w = synth.config4()
        desc = ("adjoint ring forward: 1 Msun star, 1e-4 Msun planet at 25 AU, 32,768 grains "
                "with radiation pressure and PR drag, dt=0.05 yr")
    elif name == "cfg2":
        w = synth.config2(precision=precision)
        desc = "self-gravitating swarm: 1 Msun star + 65,536 bodies, eps=1e-3 AU, dt=0.02 yr"
    elif name == "cfg3":
        nd = 262144 if n_dust is None else n_dust
        w = synth.config3(n_dust=nd, precision=precision)
        desc = ("dusty debris ring: 0.5 Msun star, 3e-5 Msun planet, 1,024 planetesimals, "
                f"{nd:,} grains with radiation/PR/wind drag, dt=0.05 yr")
    elif name == "cfg5":
        w = synth.config5(precision=precision)
        desc = "large self-gravitating disk: 1 Msun star + 1,048,576 bodies, eps=5e-4 AU, dt=0.02 yr"
    else:
        raise SystemExit(f"unknown workload {name}")
    return w, desc


def timed_run(w, precision, steps, warmup, in_order=False):
    """Warm up, time `steps` RK4 steps, then run `steps` more with a HIP event
    pair around every launch.

    The wall-clock figure comes from the first pass, which records no events
    (an event pair per launch adds about 20 us per launch on this stack, a
    large share of a step when launches are short) and uses the library's
    default schedule (with grains: the grain kernel of one step beside the
    massive stages of the next, on a second stream).  The per-kernel means
    come from the second pass, which (where grains overlap the massive
    stages) runs one launch per kernel in order, so that no kernel shares the
    GPU with another while it is timed.  in_order runs both passes that way.
    Also times one upload (host to HBM) and one download of the whole state,
    which the timed region excludes.  Returns (seconds, kernel timings,
    initial-state accelerations, |dE/E0| over the first pass, (upload s,
    download s))."""
    from dataclasses import replace
    prm = w.params.with_precision(precision)
    st = w.state
    overlapped = st.n_dust > 0 and not (st.n_grav - prm.n_planets == 0 and st.n_grav <= 32)
    timing_prm = replace(prm, schedule=1) if overlapped else prm
    with Engine(0) as eng:
        host = w.state.copy()
        eng.upload(host, timing_prm if in_order else prm)     # first upload: allocation + copy
        eng.synchronize()
        tu0 = time.perf_counter()
        eng.upload(host, timing_prm if in_order else prm)
        eng.synchronize()
        t_up = time.perf_counter() - tu0
        a, _ = eng.accel()          # forces at the initial state, for the error check
        if warmup:
            eng.step(w.dt, warmup)
        eng.synchronize()
        e0 = eng.energy()[0]
        eng.synchronize()
        t0 = time.perf_counter()
        eng.step(w.dt, steps)
        eng.synchronize()
        t1 = time.perf_counter()
        td0 = time.perf_counter()
        eng.download(host)
        t_down = time.perf_counter() - td0
        e1 = eng.energy()[0]
        eng.set_params(timing_prm)
        eng.set_timing(True)
        eng.step(w.dt, steps)
        kt = {k: eng.read_timing(k) for k in (DDYN_KCLASS_PAIRS, DDYN_KCLASS_STAGE, DDYN_KCLASS_DUST)}
        eng.set_timing(False)
        kt["nslice"] = eng.pair_plan()[0]
    return t1 - t0, kt, a, abs((e1 - e0) / e0), (t_up, t_down)


def state_update(w, kt, steps):
    """Achieved HBM bandwidth of the state update in algorithmic bytes: the
    massive stage epilogues (k_stage<0..3>, 784 B per body per step) and, with
    grains, the grain kernel's state reads and writes (120 B per grain per
    step), each over its HIP-event time per step.  The epilogues also read the
    pair kernels' slice partials; those bytes are the kernels' own scratch, not
    state, and are not counted."""
    out = {}
    nl, ms = kt[DDYN_KCLASS_STAGE]
    if nl:
        nbytes = w.state.n_grav * STAGE_BYTES_PER_BODY
        ms_step = ms / steps
        gbs = nbytes / (ms_step * 1e-3) / 1e9
        out["massive"] = {"kernel": "k_stage<0..3>", "bytes_per_body": STAGE_BYTES_PER_BODY,
                          "bytes_per_step": nbytes, "ms_per_step": ms_step, "achieved_GBs": gbs,
                          "frac_hbm_peak": gbs / PEAK_HBM_GBS, "nslice": kt["nslice"]}
    nl, ms = kt[DDYN_KCLASS_DUST]
    if nl and w.state.n_dust and w.state.n_grav > w.params.n_planets:
        nbytes = w.state.n_dust * GRAIN_BYTES_PER_GRAIN
        ms_step = ms / steps
        gbs = nbytes / (ms_step * 1e-3) / 1e9
        out["grains"] = {"kernel": "k_dust (fused with the grain forces)", "bytes_per_grain": GRAIN_BYTES_PER_GRAIN,
                         "bytes_per_step": nbytes, "ms_per_step": ms_step, "achieved_GBs": gbs,
                         "frac_hbm_peak": gbs / PEAK_HBM_GBS,
                         "note": "VALU-bound kernel: the state is read and written once per step"}
    return out or None


def jacobi_drift(n_dust=4096, nsteps=1000):
    """Jacobi-integral drift of the planetary system's circular restricted
    variant (one planet on a circular orbit, which the two eccentric planets of
    the benchmark system do not allow): max and median |dC/C0| over the test
    particles after nsteps plain FP64 steps on the GPU."""
    w = synth.config1_jacobi(n_dust=n_dust, nsteps=nsteps)
    n = w.meta["mean_motion"]
    st = w.state.copy()
    c0 = synth.jacobi_constants(st, n)
    with Engine(0) as eng:
        eng.upload(st, w.params)
        eng.step(w.dt, w.nsteps)
        eng.download(st)
    d = np.abs((synth.jacobi_constants(st, n) - c0) / c0)
    return {"max": float(d.max()), "median": float(np.median(d)), "steps": nsteps, "test_particles": n_dust,
            "note": "circular restricted variant: 1 Msun star, 9.5e-4 Msun planet on a circular orbit at 5.2 AU"}


def force_error(w, a_gpu, nsample=2048, seed=7):
    """Median/max relative acceleration error of the GPU against the FP64
    oracle on sampled massive targets (grains: all of a sampled subset)."""
    from oracle import oracle as O
    st = w.state
    rng = np.random.default_rng(seed)
    ng = st.n_grav
    idx = np.sort(rng.choice(ng, size=min(nsample, ng), replace=False))
    ref = O.massive_accel_targets(w.params, st.x[:ng], st.y[:ng], st.z[:ng], st.mu[:ng], idx)
    err = np.linalg.norm(a_gpu[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    if st.n_dust:
        a_ref, _ = O.accel(w.params, st)
        e2 = np.linalg.norm(a_gpu[ng:] - a_ref[ng:], axis=1) / np.linalg.norm(a_ref[ng:], axis=1)
        err = np.concatenate([err, e2])
    return float(np.median(err)), float(err.max())


def measured_traffic(kernel, workload, n_grav, n_dust):
    """HBM bytes per launch of `kernel` on `workload` at these sizes from the
    committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (each counter
    scaled by its reading for a copy kernel of known size), or None if not
    profiled."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json")):
        with open(path) as f:
            d = json.load(f)
        if (d.get("kernel") == kernel and d.get("workload") == workload and d.get("n_grav") == n_grav
                and d.get("n_dust") == n_dust):
            scale_f = d["copy_kib_true"] / d["copy_fetch_kib_raw"]
            scale_w = d["copy_kib_true"] / d["copy_write_kib_raw"]
            return int(round((d["fetch_kib_raw"] * scale_f + d["write_kib_raw"] * scale_w) * 1024))
    return None


def cpu_baseline(w, target_s=12.0):
    """FP64 oracle on this host's cores: a bounded sample of force evaluations.

    Massive targets are evaluated in chunks (each over all n_grav sources) until
    about target_s seconds have passed; the rate is interactions/s."""
    from oracle import oracle as O
    st = w.state
    ng, nd = st.n_grav, st.n_dust
    threads = O.max_threads()
    done = 0
    t0 = time.perf_counter()
    if nd == 0:
        chunk = max(256, min(ng, 4096))
        i0 = 0
        while time.perf_counter() - t0 < target_s:
            idx = np.arange(i0, min(i0 + chunk, ng))
            O.massive_accel_targets(w.params, st.x[:ng], st.y[:ng], st.z[:ng], st.mu[:ng], idx)
            done += len(idx) * (ng - 1)
            i0 = (i0 + chunk) % ng
        sample = f"massive accelerations of {done // (ng - 1)} targets x {ng} sources (FP64)"
    else:
        sub = w.state.copy()
        steps = 0
        while time.perf_counter() - t0 < target_s:
            O.step(w.params, sub, w.dt, 1)
            steps += 1
        done = steps * 4 * w.interactions_per_eval()
        sample = f"{steps} full FP64 RK4 steps"
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": sample + f" in {dt:.1f} s", "cpu": host_cpu()}


def host_cpu():
    """CPU model and logical CPU count of this host (Linux), for the baseline."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def imaging_run(n_dust=262144, xsize=512, ysize=512, calls=20):
    """Images and SEDs of the debris ring's grains (synthetic optics, 48
    wavelengths, an inclined observer at 10 pc): milliseconds per
    ddyn_calc_flux call (one 512 x 512 image, including its 2 MB copy to the
    host) and per ddyn_calc_sed call (all wavelengths), host clock over
    `calls` calls after two untimed ones; each call synchronises.  The oracle
    image and SED of the same state, single-threaded on this host, are timed
    beside them and give the largest relative difference."""
    from oracle import oracle as O
    w, _ = workload("cfg3", FP32_FORCE, n_dust=n_dust)
    t, v = synth.ring_imaging(w, xsize=xsize, ysize=ysize)
    st = w.state
    st.N[st.n_grav:] = 10.0 ** np.random.default_rng(3).uniform(10.0, 25.0, st.n_dust)
    wj = t.n_wav // 2
    with Engine(0) as eng:
        eng.upload(st.copy(), w.params)
        eng.upload_optics(t)
        for _ in range(2):
            F = eng.calc_flux(v, wj)
            S = eng.calc_sed(v)
        t0 = time.perf_counter()
        for _ in range(calls):
            F = eng.calc_flux(v, wj)
        t_img = (time.perf_counter() - t0) / calls
        t0 = time.perf_counter()
        for _ in range(calls):
            S = eng.calc_sed(v)
        t_sed = (time.perf_counter() - t0) / calls
    t0 = time.perf_counter()
    Fr = O.image(t, v, st, wj)
    c_img = time.perf_counter() - t0
    t0 = time.perf_counter()
    Sr = O.sed(t, v, st)
    c_sed = time.perf_counter() - t0
    nz = Fr > 0
    return {
        "config": {"workload": "cfg3 debris ring grains, imaging", "n_dust": n_dust, "image": [xsize, ysize],
                   "n_wav": t.n_wav, "n_theta": t.n_theta, "n_bins": t.n_bins},
        "image_ms": t_img * 1e3, "image_grains_per_s": n_dust / t_img,
        "sed_ms": t_sed * 1e3, "sed_grain_wavelengths_per_s": n_dust * t.n_wav / t_sed,
        "max_rel_err_vs_oracle": {"image": float(np.max(np.abs(F[nz] / Fr[nz] - 1.0))),
                                  "sed": float(np.max(np.abs(S / Sr - 1.0)))},
        "cpu_baseline": {"image_ms": c_img * 1e3, "sed_ms": c_sed * 1e3, "cores": 1, "kind": "port",
                         "sample": "the oracle's image and SED of the same state (whole)"},
    }


FOMALHAUT_INPUTS = os.path.join(ROOT, "tests", "golden", "reference_inputs")


def fomalhaut_run(steps=500, warmup=5, cpu=True):
    """The reference's own example tests/Fomalhaut.param (merope82/DiskDyn
    v1.01; copied unchanged with its optical/astrosil.lnk and the Kurucz
    8,500 K spectrum its "Library" line selects into
    tests/golden/reference_inputs/) at its full size: the A0 star, its two
    planets and a 132-148 AU disk of 2,000,000 grain tracers.  The reference's
    workflow: runmie makes optical/astrosil.dat from the refractive indices
    (150 radii, 0.1-2000 um, 90 angles), then the host library reads the
    .param file, the optical data and the spectrum, sets the dust grid and the
    grain tables and generates the particles (the reference's random stream).
    Timed: `steps` RK4 steps of the reference's own dt (det_step) on the GPU in
    plain FP64 (the reference's arithmetic; with no swarm the FP32-force form
    is the same computation) and in FP64_RSQ, state resident in HBM.  CPU
    baseline: the FP64 oracle stepping the same state on this host's cores."""
    import shutil
    import subprocess
    import tempfile
    from dataclasses import replace
    from diskdyn_amd import hostio
    tmp = tempfile.mkdtemp(prefix="ddyn_fomalhaut_")
    cwd = os.getcwd()
    try:
        for d in ("optical", "Kurucz", "JWST"):
            os.makedirs(os.path.join(tmp, d))
        shutil.copyfile(os.path.join(FOMALHAUT_INPUTS, "Kurucz_8500.dat"), os.path.join(tmp, "Kurucz", "Kurucz_8500.dat"))
        shutil.copyfile(os.path.join(FOMALHAUT_INPUTS, "Fomalhaut.param"), os.path.join(tmp, "Fomalhaut.param"))
        t0 = time.perf_counter()
        subprocess.run([os.path.join(ROOT, "diskdyn_amd", "runmie"), "-Ntheta", "90", "-Ndust", "150", "-amin", "0.1",
                        "-amax", "2000", "-lnkfile", os.path.join(FOMALHAUT_INPUTS, "astrosil.lnk"), "-output",
                        os.path.join(tmp, "optical", "astrosil.dat"), "--quiet"], check=True, timeout=600)
        t_mie = time.perf_counter() - t0
        os.chdir(tmp)
        t0 = time.perf_counter()
        cfg = hostio.read_param("Fomalhaut.param")
        opt = hostio.Optics(cfg.comp.decode(), cfg.specfile.decode(), cfg)
        hostio.set_dust_grid(cfg, opt)
        tab = hostio.grain_tables(cfg, opt)
        st = hostio.generate(cfg, opt, tab["beta"])
        t_setup = time.perf_counter() - t0
        st.beta_t, st.betasw_t, st.q_t = tab["beta"], tab["betasw"], tab["q"]
        prm = hostio.params_from_config(cfg, Rmin=tab["Rmin"])
        dt = hostio.det_step(cfg, st)
        per_step = 4 * (st.n_dust * st.n_grav + st.n_grav * (st.n_grav - 1))
        out = {"config": {"workload": "the reference's tests/Fomalhaut.param, full size: A0 star, two planets, "
                                      "132-148 AU disk of grain tracers", "n_grav": st.n_grav,
                          "n_dust": st.n_dust, "dt_yr": dt, "interactions_per_step": per_step},
               "runmie_s": t_mie, "host_setup_s": t_setup}
        for label, form in (("fp64", FP64), ("fp64_rsq", FP64_RSQ)):
            with Engine(0) as eng:
                eng.upload(st.copy(), replace(prm, precision=form))
                eng.step(dt, warmup)
                eng.synchronize()
                t0 = time.perf_counter()
                eng.step(dt, steps)
                eng.synchronize()
                tt = time.perf_counter() - t0
            out[label] = {"value": per_step * steps / tt, "unit": "interactions/s", "steps": steps,
                          "ms_per_step": 1e3 * tt / steps, "steps_per_s": steps / tt,
                          "grain_steps_per_s": st.n_dust * steps / tt}
        if cpu:
            class _W:   # the pieces of a workload the CPU baseline reads
                pass
            w = _W()
            w.state, w.params, w.dt = st, replace(prm, precision=FP64), dt
            w.interactions_per_eval = lambda: per_step // 4
            out["cpu_baseline"] = cpu_baseline(w, target_s=10.0)
        return out
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp, ignore_errors=True)


def mie_nstop(x):
    """Terms of the Mie series at size parameter x (the rule of tools/mie_single.c)."""
    x = np.asarray(x, dtype=np.float64)
    return np.floor(np.where(x < 0.02, 2.0, x + np.where((x > 8.0) & (x < 4200.0), 4.05, 4.0) * np.cbrt(x) + 2.0))


def mie_run(n_sizes=250, amin=0.001, amax=1000.0, n_wav=1000, n_theta=90, cpu_sizes=60):
    """Mie tables (ddyn_mie): n_sizes radii log-spaced over [amin, amax] um, a
    smooth synthetic refractive-index table of n_wav wavelengths over 0.09 to
    1e4 um (the range of the reference's astrosil.lnk), n_theta angles.  Rate:
    series terms per second (the sum over problems of the terms n_stop of the
    Mie series, each term with its n_theta angular updates).  CPU baseline: the
    reference's own runmie (built from its sources into oracle/_ref) on
    cpu_sizes radii over the same range, same wavelengths and angles, one
    thread."""
    import subprocess
    import tempfile
    lam = np.geomspace(0.09, 1.0e4, n_wav)
    nre = 1.5 + 1.9 * np.log10(lam / 0.09) / np.log10(1.0e4 / 0.09)
    kim = 10.0 ** (-3.0 + 3.0 * np.sin(np.log(lam)) ** 2)
    lam32, n32, k32 = (a.astype(np.float32).astype(np.float64) for a in (lam, nre, kim))
    sizes = amin * (amax / amin) ** (np.arange(n_sizes) / (n_sizes - 1.0))
    cost = np.linspace(-1.0, 1.0, n_theta)
    terms = lambda sz: float(mie_nstop(2.0 * np.pi * np.asarray(sz)[:, None] / lam32[None, :]).sum())
    with Engine(0) as eng:
        eng.mie(sizes[:4], lam32, n32, k32, cost)           # warm-up
        t0 = time.perf_counter()
        qa, qs, qp = eng.mie(sizes, lam32, n32, k32, cost)
        t_gpu = time.perf_counter() - t0
    out = {"config": {"workload": "Mie tables, synthetic refractive indices", "n_sizes": n_sizes,
                      "size_um": [amin, amax], "n_wav": n_wav, "wav_um": [0.09, 1.0e4], "n_theta": n_theta,
                      "max_size_parameter": float(2.0 * np.pi * amax / lam32[0])},
           "seconds": t_gpu, "terms": terms(sizes), "terms_per_s": terms(sizes) / t_gpu, "cpu_baseline": None}
    from oracle import oracle as O
    if os.path.exists(O.REF_RUNMIE):
        with tempfile.TemporaryDirectory() as d:
            lnk = os.path.join(d, "syn.lnk")
            np.savetxt(lnk, np.stack([lam, nre, kim], axis=1), fmt="%.6e", header="lambda n k")
            t0 = time.perf_counter()
            subprocess.run([O.REF_RUNMIE, "-Ntheta", str(n_theta), "-Ndust", str(cpu_sizes), "-amin", str(amin),
                            "-amax", str(amax), "-lnkfile", lnk, "-output", os.path.join(d, "t.dat")],
                           check=True, capture_output=True)
            t_cpu = time.perf_counter() - t0
        cs = amin * (amax / amin) ** (np.arange(cpu_sizes) / (cpu_sizes - 1.0))
        out["cpu_baseline"] = {"terms_per_s": terms(cs) / t_cpu, "seconds": t_cpu, "cores": 1, "kind": "reference",
                               "sample": f"tools/runmie built from the reference's sources, {cpu_sizes} radii over "
                                         "the same range, same wavelengths and angles"}
    return out


def adjoint_run(name, precision=FP32_FORCE):
    """Discrete-adjoint workloads.

    cfg4: 32,768-grain ring, 400 steps of dt = 0.05 yr, checkpoints every 20;
          L = mean |r(T) - r_target(T)|^2 with the target from planet
          m = 1.2e-4 Msun, a = 26 AU; gradients in planet (m, a, e) checked by
          central finite differences (h = 1e-6 relative).  The timed forward +
          reverse pass follows one untimed pass.
    cfg5: 1,048,577-body disk, 10-step window in one tape segment; FP32-force
          form, or the plain FP64 form (--fp64), or FP64 with refined hardware
          reciprocal square roots (--fp64-rsq); L = mean (|r(T)| - |r_Kepler(T)|)^2;
          directional derivative in the initial state checked by central
          differences.  Progress of each forward chunk and adjoint window goes
          to stderr."""
    import torch

    from diskdyn_amd.adjoint import (Runner, Trajectory, kepler_deviation_loss, kepler_targets,
                                     planet_initial_state, ring_loss)
    out = {"metric": "discrete-adjoint RK4 steps/s (forward + reverse)", "unit": "steps/s", "n_gpus": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic (seeded Philox)"}
    if name == "cfg4":
        nsteps, ck = 400, 20
        base = synth.config4()
        tgt = synth.config4(planet_m=1.2e-4, planet_a=26.0)
        run = Runner(base.state, base.params, base.dt, nsteps, ckpt_every=ck, tape=True)
        target = torch.from_numpy(run.forward(tgt.state.copy()).pos())
        ng = base.state.n_grav
        beta = torch.from_numpy(base.state.beta_p.copy())

        def loss(m, a, e):
            r_m, v_m, mu = planet_initial_state(1.0, m, a, e)
            r0 = torch.cat([r_m, torch.from_numpy(base.state.pos()[ng:])])
            v0 = torch.cat([v_m, torch.from_numpy(base.state.vel()[ng:])])
            rT, _ = Trajectory.apply(r0, v0, mu, beta, run)
            return ring_loss(rT, target, ng)

        x0 = {"m": 1e-4, "a": 25.0, "e": 0.15}
        # one untimed pass first: the first call carries one-off costs (device
        # buffers of the run, module and allocator warm-up)
        w_ = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in x0.items()}
        loss(w_["m"], w_["a"], w_["e"]).backward()
        p = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in x0.items()}
        t0 = time.perf_counter()
        L = loss(p["m"], p["a"], p["e"])
        L.backward()
        t_adj = time.perf_counter() - t0
        grads, fd_err = {}, {}
        with torch.no_grad():
            for k in x0:
                h = 1e-6 * x0[k]
                args_p = {q: torch.tensor(x0[q] + (h if q == k else 0.0), dtype=torch.float64) for q in x0}
                args_m = {q: torch.tensor(x0[q] - (h if q == k else 0.0), dtype=torch.float64) for q in x0}
                fd = (loss(**args_p).item() - loss(**args_m).item()) / (2 * h)
                grads[k] = p[k].grad.item()
                fd_err[k] = abs(grads[k] - fd) / abs(fd)
        out.update({"value": nsteps / t_adj, "seconds": t_adj, "steps": nsteps, "ckpt_every": ck,
                    "dtype": "fp64", "loss": L.item(), "grad": grads, "fd_rel_err": fd_err,
                    "config": {"workload": "cfg4_adjoint: 1 Msun star, planet m=1e-4 Msun a=25 AU e=0.15, "
                               "32,768 ring grains, beta~U(0,0.3), radiation + PR drag, dt=0.05 yr",
                               "n_grav": ng, "n_dust": base.state.n_dust}})
    else:
        # one segment: the whole window's states and stage positions stay on
        # the device tape (about 1.7 GB at 1,048,577 bodies), so the reverse
        # pass recomputes no step
        nsteps = ck = 10
        w = synth.config5(precision=precision)
        st = w.state
        report = lambda msg: print(f"[adjoint cfg5] {msg}", file=sys.stderr, flush=True)  # noqa: E731
        run = Runner(st, w.params, w.dt, nsteps, ckpt_every=ck, report=report, tape=True)
        target = kepler_targets(st, w.params.n_planets, nsteps * w.dt)
        empty = torch.zeros(0, dtype=torch.float64)

        def loss(r0, v0, mu):
            rT, _ = Trajectory.apply(r0, v0, mu, empty, run)
            return kepler_deviation_loss(rT, target, w.params.n_planets)

        r0 = torch.from_numpy(st.pos()).requires_grad_(True)
        v0 = torch.from_numpy(st.vel()).requires_grad_(True)
        mu = torch.from_numpy(st.mu.copy())
        t0 = time.perf_counter()
        L = loss(r0, v0, mu)
        L.backward()
        t_adj = time.perf_counter() - t0
        rng = np.random.default_rng(3)
        dr = torch.from_numpy(rng.normal(size=r0.shape))
        dv = torch.from_numpy(rng.normal(size=v0.shape))
        h = 1e-6
        with torch.no_grad():
            fd = (loss(r0 + h * dr, v0 + h * dv, mu).item() - loss(r0 - h * dr, v0 - h * dv, mu).item()) / (2 * h)
        an = float((r0.grad * dr).sum() + (v0.grad * dv).sum())
        out.update({"value": nsteps / t_adj, "seconds": t_adj, "steps": nsteps, "ckpt_every": ck,
                    "dtype": {FP32_FORCE: "fp32 pair forces and VJP tiles, fp64 state",
                              FP64_RSQ: "fp64, refined hardware rsqrt in every pair law"}.get(precision, "fp64"),
                    "loss": L.item(),
                    "directional": {"adjoint": an, "fd": fd, "rel_err": abs(an - fd) / abs(fd)},
                    "config": {"workload": "cfg5_large_disk 10-step adjoint window: 1 Msun star + 1,048,576 bodies, "
                               "eps=5e-4 AU, dt=0.02 yr", "n_grav": st.n_grav, "n_dust": 0}})
    print(json.dumps(out), flush=True)
    return 0


DEFAULT_STEPS = {"cfg1": 1000, "cfg2": 200, "cfg3": 500, "cfg4": 400, "cfg5": 20}


def measure(name, steps, fp64_steps, warmup, in_order=False, cpu=False, n_dust=None):
    """All three precision forms of one workload; returns the JSON fields of
    its line (value, roofline, errors, speed-ups, state update, ...)."""
    w32, desc = workload(name, FP32_FORCE, n_dust)
    per_step = 4 * w32.interactions_per_eval()
    t32, kt32, a32, drift32, io32 = timed_run(w32, FP32_FORCE, steps, warmup, in_order)
    print(f"[bench] {name} FP32-force: {steps} steps in {t32:.3f} s", file=sys.stderr, flush=True)
    k64 = fp64_steps
    t64, kt64, a64, drift64, _ = timed_run(w32, FP64, k64, min(warmup, 2), in_order)
    print(f"[bench] {name} plain FP64: {k64} steps in {t64:.3f} s", file=sys.stderr, flush=True)
    ns = w32.state.n_grav - w32.params.n_planets
    # FP64 with refined hardware reciprocal square roots in every pair law
    trq, ktrq, arq, driftrq, _ = timed_run(w32, FP64_RSQ, k64, min(warmup, 2), in_order)
    print(f"[bench] {name} FP64_RSQ: {k64} steps in {trq:.3f} s", file=sys.stderr, flush=True)

    med32, max32 = force_error(w32, a32)
    med64, max64 = force_error(w32, a64)

    value = per_step * steps / t32
    v64 = per_step * k64 / t64
    # dominant kernel: massive pairs (swarm workloads) or the grain kernel (debris ring)
    dom = DDYN_KCLASS_PAIRS if w32.state.n_dust == 0 else DDYN_KCLASS_DUST
    nl, ms = kt32[dom]
    if dom == DDYN_KCLASS_PAIRS:
        inter_launch = w32.state.n_grav * ns
        kname = "k_pairs32"
    elif ns == 0:
        # star + planets only: k_small_steps runs all steps of a call in one launch
        inter_launch = per_step * steps / max(nl, 1)
        kname = "k_small_steps"
    else:
        inter_launch = 4 * w32.state.n_dust * w32.state.n_grav
        kname = "k_dust<fp32>"
    avg_ms = ms / max(nl, 1)
    achieved = FLOP_PER_INTERACTION * inter_launch / (avg_ms * 1e-3) / 1e12
    nl64, ms64 = kt64[dom]
    inter64 = per_step * k64 / max(nl64, 1) if (ns == 0 and dom != DDYN_KCLASS_PAIRS) else inter_launch
    ach64 = FLOP_PER_INTERACTION * inter64 / (ms64 / max(nl64, 1) * 1e-3) / 1e12
    vrq = per_step * k64 / trq
    nlrq, msrq = ktrq[dom]
    interrq = per_step * k64 / max(nlrq, 1) if (ns == 0 and dom != DDYN_KCLASS_PAIRS) else inter_launch
    achrq = FLOP_PER_INTERACTION * interrq / (msrq / max(nlrq, 1) * 1e-3) / 1e12
    medrq, maxrq = force_error(w32, arq)
    rsq = {"value": vrq, "ms_per_step": 1e3 * trq / k64, "steps_per_s": k64 / trq, "steps": k64,
           "force_rel_err_median": medrq, "force_rel_err_max": maxrq,
           "kernel_tflops": achrq, "kernel_frac_fp64_peak": achrq / PEAK_FP64_TFLOPS,
           "speedup_fp32_vs_fp64_rsq": value / vrq, "speedup_vs_plain_fp64": vrq / v64, "energy_drift": driftrq,
           "note": "FP64 throughout; 1/R^3 (and the star term's 1/R^2) from v_rsq_f64 + one third-order step"}
    out = {
        "value": value,
        "steps": steps,
        "ms_per_step": 1e3 * t32 / steps,
        "steps_per_s": steps / t32,
        "config": {"workload": f"{w32.name}: {desc}", "n_grav": w32.state.n_grav,
                   "n_dust": w32.state.n_dust, "interactions_per_step": per_step,
                   "parallelism": "single GPU"},
        "force_rel_err": {"median": med32, "max": max32, "vs": "FP64 CPU oracle, sampled targets"},
        "fp64": {"value": v64, "ms_per_step": 1e3 * t64 / k64, "steps_per_s": k64 / t64, "steps": k64,
                 "force_rel_err_median": med64, "force_rel_err_max": max64,
                 "kernel_tflops": ach64, "kernel_frac_fp64_peak": ach64 / PEAK_FP64_TFLOPS},
        "speedup_fp32_vs_fp64": value / v64,
        "fp64_rsq": rsq,
        "energy_drift": {"fp32_force": drift32, "fp64": drift64,
                         "note": "|dE/E0| of the massive set over the timed steps (each form's own step count)"},
        "roofline": {"bound": "valu", "kernel": kname, "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": measured_traffic(kname, w32.name, w32.state.n_grav, w32.state.n_dust), "launches": nl, "avg_launch_ms": avg_ms,
                     "flop_per_launch": FLOP_PER_INTERACTION * inter_launch,
                     "note": "achieved = 20 FLOP x interactions per launch / mean HIP-event launch time; "
                             "traffic = calibrated FETCH_SIZE + WRITE_SIZE bytes per launch from profiles/",
                     "in_timed_schedule": {
                         "achieved": FLOP_PER_INTERACTION * per_step * steps / t32 / 1e12,
                         "frac": FLOP_PER_INTERACTION * per_step * steps / t32 / 1e12 / PEAK_FP32_TFLOPS,
                         "note": "20 FLOP x every interaction of a step / the timed step: all kernels as the "
                                 "timed schedule runs them (with grains: two grain launches per step that "
                                 "overlap each other and the massive stages)"}},
        "state_update": state_update(w32, kt32, steps),
        "host_transfer": {"upload_ms": 1e3 * io32[0], "download_ms": 1e3 * io32[1],
                          "value_incl_transfer": per_step * steps / (t32 + io32[0] + io32[1]),
                          "note": "one upload and one download of the whole state around the timed steps; "
                                  "value excludes them (state resident in HBM)"},
    }
    if name == "cfg1":
        out["jacobi_drift"] = jacobi_drift()
    if cpu:
        out["cpu_baseline"] = cpu_baseline(w32)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed RK4 steps (default: the workload's stated count)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg3", choices=["cfg1", "cfg2", "cfg3", "cfg4", "cfg5"])
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="steps for the FP64 comparison runs (default: --steps)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-also", action="store_true",
                    help="skip the swarm, large-disk, 1,048,576-grain, imaging and Mie lines that the debris-ring "
                         "run adds under 'also'")
    ap.add_argument("--mie", action="store_true",
                    help="time only the Mie tables (ddyn_mie) against the reference's runmie on the host")
    ap.add_argument("--fomalhaut", action="store_true",
                    help="time only the reference's Fomalhaut example at full size (the 'also.fomalhaut' entry)")
    ap.add_argument("--imaging", action="store_true",
                    help="time only the images and SEDs of the debris ring's grains (the 'also.imaging' entry)")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the discrete adjoint instead (workloads cfg4 and cfg5)")
    ap.add_argument("--fp64", action="store_true", help="with --adjoint cfg5: the plain FP64 form")
    ap.add_argument("--fp64-rsq", action="store_true",
                    help="with --adjoint cfg5: FP64 with refined hardware rsqrt in every pair law")
    ap.add_argument("--in-order", action="store_true",
                    help="time every pass with one launch per kernel, in order (kernel profiling runs)")
    ap.add_argument("--n-dust", type=int, default=None,
                    help="cfg3 only: grain count of the debris ring (default 262,144, the stated size; "
                         "other counts are for kernel scaling studies)")
    args = ap.parse_args()
    if args.n_dust is not None and (args.workload != "cfg3" or args.n_dust < 1):
        raise SystemExit("--n-dust takes a positive grain count with --workload cfg3")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the swarm would shard by target block, the grains by grain range)", file=sys.stderr)
        return 2
    if args.imaging:
        print(json.dumps({"imaging": imaging_run()}), flush=True)
        return 0
    if args.fomalhaut:
        print(json.dumps({"fomalhaut": fomalhaut_run(cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.mie:
        print(json.dumps({"mie": mie_run()}), flush=True)
        return 0
    if args.adjoint:
        if args.workload not in ("cfg4", "cfg5"):
            raise SystemExit("--adjoint runs the cfg4 or cfg5 workload")
        return adjoint_run(args.workload, FP64 if args.fp64 else FP64_RSQ if args.fp64_rsq else FP32_FORCE)

    steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.workload]
    k64 = args.fp64_steps if args.fp64_steps is not None else steps
    m = measure(args.workload, steps, k64, args.warmup, args.in_order, cpu=not args.no_cpu, n_dust=args.n_dust)
    out = {
        "metric": "pairwise interactions/s & RK4 steps/s, FP32-force vs FP64 at 1e-5 rel err",
        "value": m.pop("value"),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": m.pop("ms_per_step"),
        "steps_per_s": m.pop("steps_per_s"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair forces, fp64 state",
        "data": "synthetic (seeded Philox)",
    }
    m.pop("steps")
    out.update(m)
    if (args.workload == "cfg3" and not args.no_also and args.steps is None and not args.in_order
            and args.n_dust is None):
        # the swarm and the disk at reduced step counts, and the debris ring with
        # four times the grains: their own throughput, roofline and FP32-force
        # speed-up (no CPU baseline)
        also = {}
        for key, name, st32, st64, nd in (("cfg2", "cfg2", 200, 20, None), ("cfg5", "cfg5", 2, 1, None),
                                          ("cfg3_1048576_grains", "cfg3", 40, 4, 1048576)):
            a = measure(name, st32, st64, 1, n_dust=nd)
            also[key] = {k: a[k] for k in ("value", "ms_per_step", "steps_per_s", "steps", "config",
                                           "force_rel_err", "speedup_fp32_vs_fp64", "roofline", "state_update")}
            also[key]["fp64"] = {k: a["fp64"][k] for k in ("value", "ms_per_step", "steps", "kernel_frac_fp64_peak")}
            also[key]["fp64_rsq"] = {k: a["fp64_rsq"][k] for k in ("value", "ms_per_step", "steps",
                                                                    "speedup_fp32_vs_fp64_rsq")}
        # secondary entries: a failure in one is reported in its place and does
        # not take the headline line with it
        for key, fn in (("imaging", imaging_run), ("mie", mie_run), ("fomalhaut", fomalhaut_run)):
            try:
                also[key] = fn()
            except Exception as e:   # noqa: BLE001
                print(f"[bench] also.{key} failed: {e!r}", file=sys.stderr, flush=True)
                also[key] = {"error": repr(e)}
        out["also"] = also
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

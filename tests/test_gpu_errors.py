"""Error paths of the C ABI on a real context: every misuse returns its
DDYN_ERR_* code with a message, nothing exits, and the context stays usable."""
import ctypes
from dataclasses import replace

import numpy as np
import pytest

from diskdyn_amd import engine, synth
from diskdyn_amd.engine import Engine, to_c_params
from diskdyn_amd.system import FP64

pytestmark = pytest.mark.gpu

DDYN_ERR_ARG, DDYN_ERR_STATE = -1, -3


def _dp(n):
    return np.zeros(max(n, 1))


def upload_raw(eng, n_grav, n_dust, arrays=None):
    a = arrays or [_dp(4) for _ in range(8)]
    ids = np.zeros(4, dtype=np.int32)
    return eng.lib.ddyn_upload_state(eng.ctx, n_grav, n_dust, *[engine._dp(x) for x in a], engine._ip(ids))


def message(eng):
    return eng.lib.ddyn_last_error(eng.ctx).decode()


def test_bad_device_is_refused():
    lib = engine.load_library()
    ctx = ctypes.c_void_p()
    assert lib.ddyn_create(ctypes.byref(ctx), 1 << 20) == DDYN_ERR_ARG
    assert not ctx.value


def test_calls_out_of_order_and_bad_arguments_then_recovery():
    w = synth.config3(n_dust=300, n_swarm=40, precision=FP64)
    with Engine() as eng:
        lib, ctx = eng.lib, eng.ctx
        assert lib.ddyn_step(ctx, 0.01, 1) == DDYN_ERR_STATE
        assert "ddyn_set_params" in message(eng)
        ns = ctypes.c_int32()
        assert lib.ddyn_pair_plan(ctx, ctypes.byref(ns), None) == DDYN_ERR_STATE
        bad = to_c_params(replace(w.params, n_planets=0))
        assert lib.ddyn_set_params(ctx, ctypes.byref(bad)) == DDYN_ERR_ARG
        assert "n_planets" in message(eng)
        bad = to_c_params(replace(w.params, precision=7))
        assert lib.ddyn_set_params(ctx, ctypes.byref(bad)) == DDYN_ERR_ARG
        bad = to_c_params(replace(w.params, SWQ=1, vsw=0.0))
        assert lib.ddyn_set_params(ctx, ctypes.byref(bad)) == DDYN_ERR_ARG
        eng.set_params(w.params)
        assert lib.ddyn_step(ctx, 0.01, 1) == DDYN_ERR_STATE        # no state yet
        assert upload_raw(eng, 0, 0) == DDYN_ERR_ARG
        assert upload_raw(eng, (1 << 30) + 1, 0) == DDYN_ERR_ARG
        assert upload_raw(eng, 4, -1) == DDYN_ERR_ARG
        f = np.zeros(2701, dtype=np.float32)
        assert lib.ddyn_upload_bin_tables(ctx, 2701, engine._fp(f), engine._fp(f), engine._fp(f)) == DDYN_ERR_ARG
        assert "2700" in message(eng)
        # a valid upload, then arguments that must be refused on a ready context
        eng.upload(w.state.copy(), w.params)
        assert lib.ddyn_step(ctx, 0.0, 1) == DDYN_ERR_ARG
        assert lib.ddyn_step(ctx, 0.01, -1) == DDYN_ERR_ARG
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        assert lib.ddyn_read_timing(ctx, 3, ctypes.byref(n), ctypes.byref(ms)) == DDYN_ERR_ARG
        big = to_c_params(replace(w.params, n_planets=w.state.n_grav + 1))
        assert lib.ddyn_set_params(ctx, ctypes.byref(big)) == DDYN_ERR_ARG
        lor = to_c_params(replace(w.params, FLQ=1))
        assert lib.ddyn_set_params(ctx, ctypes.byref(lor)) == 0
        lam = np.zeros(3 * w.state.n_all)
        g = np.zeros(3 * w.state.n_all)
        gm = np.zeros(w.state.n_grav)
        rc = lib.ddyn_adjoint(ctx, 0.01, 2, 0, engine._dp(lam), engine._dp(lam), engine._dp(g), engine._dp(g),
                              engine._dp(gm), None)
        assert rc == DDYN_ERR_ARG and "Lorentz" in message(eng)
        # still usable: the same context steps and matches a fresh one
        eng.set_params(w.params)
        eng.step(w.dt, 3)
        a = eng.download(w.state.copy())
    with Engine() as fresh:
        fresh.upload(w.state.copy(), w.params)
        fresh.step(w.dt, 3)
        b = fresh.download(w.state.copy())
    for k in ("x", "y", "z", "vx", "vy", "vz", "N"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k

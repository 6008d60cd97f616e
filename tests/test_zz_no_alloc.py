"""tsp_cfl_noalloc (the CFL estimate on the library's fixed per-device scratch
array, no device allocation) against tsp_cfl on the same contexts: bitwise the
same lambda_max and dt, on two calls in a row, the second on a smaller mesh,
and meshes above its scratch limit are refused."""
import ctypes

import numpy as np
import pytest

from tspeed_amd import capi, workloads as W

pytestmark = pytest.mark.gpu


def _cfl(eng, fn, factor=0.5, iters=60):
    dt, lm = ctypes.c_double(), ctypes.c_double()
    rc = getattr(capi.lib(), fn)(eng.handle, factor, iters, ctypes.byref(dt), ctypes.byref(lm))
    return rc, dt.value, lm.value


def _engine(nx, ny, N, precision="fp64"):
    p = W.square_ricker(nx=nx, ny=ny, N=N, nsteps=10, nrec=2)
    eng = capi.Engine(precision)
    eng.set_tables(N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    return eng


def test_cfl_noalloc_equals_cfl_on_two_calls_in_a_row():
    big = _engine(30, 28, 5)     # 1680 triangles, 35 280 coefficients per vector
    small = _engine(6, 5, 3)     # 60 triangles: runs on the scratch the big call left behind
    for eng in (big, small):
        rc1, dt1, lm1 = _cfl(eng, "tsp_cfl_noalloc")
        rc0, dt0, lm0 = _cfl(eng, "tsp_cfl")
        print(f"ne={eng.ne} N={eng.N}: noalloc lambda_max {lm1!r} dt {dt1!r}; tsp_cfl {lm0!r} {dt0!r}", flush=True)
        assert rc1 == 0 and rc0 == 0
        assert lm1 > 0 and dt1 > 0
        assert lm1 == lm0 and dt1 == dt0
    # a third call on the big mesh after the small one: same result again
    assert _cfl(big, "tsp_cfl_noalloc") == _cfl(big, "tsp_cfl")
    big.close()
    small.close()


def test_cfl_noalloc_refuses_meshes_above_its_scratch():
    # 2 * 40 * 160 triangles at p = 8: 12 800 * 90 = 1 152 000 coefficients fit;
    # 2 * 60 * 200 triangles at p = 8: 24 000 * 90 = 2 160 000 do not
    eng = _engine(200, 60, 8)
    rc, _, _ = _cfl(eng, "tsp_cfl_noalloc", iters=2)
    assert rc == capi.TSP_ERR_ARG
    assert b"scratch" in capi.lib().tsp_last_error(eng.handle)
    rc0, dt0, _ = _cfl(eng, "tsp_cfl", iters=2)
    assert rc0 == 0 and dt0 > 0
    eng.close()

"""The operator plugin below LeapFrog (tsp_apply_operator / tsp_apply_invmass,
replacing MyMatMultiDim<MyMat>::operator* and MyMatMultiDimBlockDiag::operator*,
MyMat.hpp:330-342) against the FP64 oracle's B x and invM x.

Bar: relative L2 <= 1e-13 (the device sums each row with FMA over merged 2x2
component blocks, the oracle per component block in the reference's loop
order).  Device-pointer calls use the HIP runtime the engine itself loaded."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tspeed_amd import capi, workloads as W

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _problem(N, basis=capi.DUBINER, seed=31):
    p = W.square_ricker(nx=5, ny=4, N=N, nsteps=10, nrec=3, seed=seed)
    rng = np.random.default_rng(seed + N)
    p.lam = p.lam * rng.uniform(0.5, 1.5, p.ne)
    p.mu = p.mu * rng.uniform(0.5, 1.5, p.ne)
    p.rho = p.rho * rng.uniform(0.8, 1.2, p.ne)
    p.basis = basis
    p.__post_init__()
    return p


def _engine(p, precision="fp64", ncols=1):
    eng = capi.Engine(precision, ncols=ncols)
    eng.set_tables(p.N, capi.tables(p.N, p.basis))
    if p.basis != capi.DUBINER:
        eng.set_mass(capi.base_invmass(p.N, p.basis))
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    return eng


@pytest.mark.parametrize("basis", [capi.DUBINER, capi.BOUNDARY_ADAPTED])
@pytest.mark.parametrize("N", range(1, 9))
def test_apply_operator_and_invmass_match_oracle(N, basis):
    p = _problem(N, basis)
    say(f"plugin N={N} basis={basis} ne={p.ne}")
    eng = _engine(p)
    op = O.Operator(p.N, p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty, basis=basis)
    rng = np.random.default_rng(N)
    for trial in range(2):
        x = rng.standard_normal(op.ndof)
        eb, em = rel_l2(eng.apply_operator(x), op.apply(x)), rel_l2(eng.apply_invmass(x), op.apply_invM(x))
        say(f"  B x {eb:.2e}  invM x {em:.2e}")
        assert eb <= 1e-13 and em <= 1e-13
    eng.close()


def test_apply_on_device_pointers_and_errors():
    """device_ptrs = 1: vectors stay on the device, the call only queues work;
    invM (B x) composed on the device equals the host-side composition; an
    FP32C context of a non-orthonormal basis refuses B x (its operator is
    stored premultiplied), and overlapping vectors are refused."""
    p = _problem(4)
    eng = _engine(p)
    op = O.Operator(p.N, p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = op.ndof
    nbytes = n * 8
    ptrs = []
    for _ in range(3):
        d = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(nbytes)) == 0
        ptrs.append(d)
    dx, dy, dz = ptrs
    try:
        x = np.random.default_rng(3).standard_normal(n)
        assert hip.hipMemcpy(dx, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nbytes), 1) == 0
        L = capi.lib()
        assert L.tsp_apply_operator(eng.handle, dx, dy, 1) == 0
        assert L.tsp_apply_invmass(eng.handle, dy, dz, 1) == 0
        eng.synchronize()
        z = np.zeros(n)
        assert hip.hipMemcpy(z.ctypes.data_as(ctypes.c_void_p), dz, ctypes.c_size_t(nbytes), 2) == 0
        ref = op.apply_invM(op.apply(x))
        say(f"  device invM B x vs oracle {rel_l2(z, ref):.2e}")
        assert rel_l2(z, ref) <= 1e-13
        # in-place (overlapping) calls are refused
        assert L.tsp_apply_operator(eng.handle, dx, dx, 1) == capi.TSP_ERR_ARG
        assert L.tsp_apply_operator(eng.handle, None, dy, 1) == capi.TSP_ERR_ARG
    finally:
        for d in ptrs:
            hip.hipFree(d)
    eng.close()
    q = _problem(3, capi.BOUNDARY_ADAPTED)
    e32 = _engine(q, "fp32c")
    with pytest.raises(capi.EngineError) as ex:
        e32.apply_operator(np.zeros(2 * q.ne * O.nln(3)))
    assert ex.value.code == capi.TSP_ERR_STATE
    # invM needs no operator and works in every context
    x = np.random.default_rng(5).standard_normal(2 * q.ne * O.nln(3))
    opq = O.Operator(q.N, q.verts, q.neigh, q.neighedge, q.lam, q.mu, q.rho, q.penalty, basis=q.basis)
    assert rel_l2(e32.apply_invmass(x), opq.apply_invM(x)) <= 1e-13
    e32.release_fp64()
    with pytest.raises(capi.EngineError):
        e32.apply_operator(x)
    e32.close()


def test_shared_operator_lifetime():
    """tsp_share_operator keeps the owner and its borrowers consistent: the
    owner cannot take a new mesh while borrowed, releasing its FP64 copy takes it
    from the borrowers too, and a borrower whose owner is destroyed reports
    TSP_ERR_STATE instead of reading freed memory."""
    p = _problem(3)
    owner = _engine(p, "fp32c")
    b1 = capi.Engine(share=owner)
    blk, _ = b1.get_blocks(0, 2)
    assert np.abs(blk).max() > 0
    L = capi.lib()
    verts = np.ascontiguousarray(p.verts.reshape(-1), dtype=np.float64)
    rc = L.tsp_set_mesh(owner.handle, p.ne, verts, np.ascontiguousarray(p.neigh.reshape(-1), dtype=np.int32),
                        np.ascontiguousarray(p.neighedge.reshape(-1), dtype=np.int32),
                        np.ascontiguousarray(p.lam), np.ascontiguousarray(p.mu), np.ascontiguousarray(p.rho), 10.0)
    assert rc == capi.TSP_ERR_STATE
    say("  owner mesh change refused:", L.tsp_last_error(owner.handle).decode())
    owner.release_fp64()
    for e in (owner, b1):
        with pytest.raises(capi.EngineError) as ex:
            e.get_blocks(0, 1)
        assert ex.value.code == capi.TSP_ERR_STATE
    # the borrower still steps on the shared FP32 operator
    b1.set_dt(p.dt)
    b1.first_step()
    b1.run(3)
    b1.synchronize()
    owner.close()
    assert L.tsp_run(b1.handle, 1) == capi.TSP_ERR_STATE
    with pytest.raises(capi.EngineError) as ex:
        b1.apply_invmass(np.zeros(2 * p.ne * O.nln(3)))
    assert ex.value.code == capi.TSP_ERR_STATE
    b1.close()

"""Parity of the HIP path (through the C-ABI) with the FP64 CPU oracle.

Error measure (stated once, used throughout): relative L2 = ||x - x_ref|| /
||x_ref|| over the whole array -- the final displacement vector, or the full
seismogram array [sample, receiver, component].  Per-receiver errors are
normalised by the norm of the whole seismogram array, so that receivers the
wave has not reached (tiny signal) do not turn roundoff into large ratios.

Bars: FP64 <= 1e-12; assembled blocks <= 1e-13 (relative to the largest entry);
FP32 compensated <= 1e-4 on seismograms."""
import os
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O
from tspeed_amd import capi, workloads as W

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_problem(N, nx=6, ny=5, nsteps=200, hetero=True, seed=4, nrec=7, stride=1):
    p = W.square_ricker(nx=nx, ny=ny, N=N, nsteps=nsteps, nrec=nrec, seed=seed, stride=stride)
    if hetero:
        rng = np.random.default_rng(seed)
        f = rng.uniform(0.6, 1.4, p.ne)
        p.lam = p.lam * f * rng.uniform(0.8, 1.2, p.ne)
        p.mu = p.mu * f
        p.rho = p.rho * rng.uniform(0.8, 1.2, p.ne)
    # strong source early in the run so short runs see signal
    p.t0 = np.array([10 * p.dt])
    p.f0 = np.array([1.0 / (8 * p.dt)])
    return p


def oracle_op(p):
    return O.Operator(p.N, p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty, basis=p.basis)


def oracle_run(p, op=None, u0=None, v0=None, stride=None):
    op = op or oracle_op(p)
    return op.run(p.dt, p.nsteps, src=p.source_dict(0), rec_elem=p.rec_elem, rec_phi=p.rec_phi,
                  stride=p.stride if stride is None else stride, u0=u0, v0=v0)


def seis_err(s, sref):
    nrm = np.linalg.norm(sref)
    per_rec = np.linalg.norm((s - sref).transpose(1, 0, 2).reshape(s.shape[1], -1), axis=1) / nrm
    return rel_l2(s, sref), float(per_rec.max())


# ------------------------------------------------------------------ assembly
@pytest.mark.parametrize("N", range(1, 9))
def test_assembly_matches_oracle(N):
    p = small_problem(N, nx=4, ny=3)
    say(f"assembly N={N} ne={p.ne}")
    eng = capi.Engine("fp64")
    eng.set_tables(N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    blk, invm = eng.get_blocks()
    op = oracle_op(p)
    ref = op.blocks()
    err = np.abs(blk - ref).max() / np.abs(ref).max()
    say(f"  max block diff / max entry = {err:.3e}; bitwise equal entries: {(blk == ref).mean():.4f}")
    assert err <= 1e-13
    np.testing.assert_array_equal(invm, op.invm())
    # absent neighbours give zero blocks
    for e, f in zip(*np.nonzero(p.neigh < 0)):
        assert not blk[e, 1 + f].any()


# ------------------------------------------------------------------ BoundaryAdapted basis
def _ba_problem(N, nsteps=60, seed=50):
    p = small_problem(N, nx=4, ny=3, nsteps=nsteps, seed=seed + N, nrec=5)
    if N == 1:
        p.dt *= 0.5          # 0.2 h / (vp p^2) is above the leap-frog limit at p = 1
    p.basis = capi.BOUNDARY_ADAPTED
    p.__post_init__()
    return p


@pytest.mark.parametrize("N", [2, 4, 5])
def test_boundary_adapted_on_the_batched_and_multi_paths(N):
    """The BoundaryAdapted inverse mass is folded into the operator rows and the
    source values on upload, so the multi-column MFMA path and the set of
    instances take it unchanged: three sources as columns of an 8-column
    context match the single-context FP32 run of each source and the FP64
    oracle, and a BoundaryAdapted instance in a set is bitwise its single-context
    FP32 run."""
    import copy
    p = _ba_problem(N, nsteps=50)
    p.src_xy = np.array([[300.0, 700.0], [650.0, 420.0], [820.0, 160.0]])
    p.f0 = np.repeat(p.f0, 3)
    p.t0 = np.repeat(p.t0, 3)
    p.dirs = np.array([[0.0, -1.0], [0.6, 0.8], [1.0, 0.0]])
    p.__post_init__()
    eng, s = W.run_batched(p, 8)
    assert eng.status() == 0
    for k in range(3):
        q = copy.copy(p)
        q.src_xy, q.f0, q.t0, q.dirs = p.src_xy[k:k + 1], p.f0[k:k + 1], p.t0[k:k + 1], p.dirs[k:k + 1]
        q.__post_init__()
        _, _, s32 = W.run_engine(q, "fp32c")
        _, sref = oracle_run(q)
        e_o, e_32 = rel_l2(s[:, :, k], sref), rel_l2(s[:, :, k], s32)
        say(f"  BA p={N} column {k}: vs oracle {e_o:.2e}, vs single-column FP32 {e_32:.2e}")
        assert e_o <= 1e-3 and e_32 <= 1e-3
        if k == 0:
            m = capi.Multi(chunk=8)
            i = m.add(W.prepare_instance(q), q.nsteps)
            while m.run(13):
                pass
            np.testing.assert_array_equal(m.engine(i).seismograms(), s32)
            m.close()
    eng.close()


@pytest.mark.parametrize("N", [2, 4, 5])
def test_boundary_adapted_split_products(N):
    """The split BF16 products on the folded BoundaryAdapted operator: three
    sources in an 8-column context, 50 steps, with five products within 1e-4
    of the FP64 oracle, as the FP32 products are (measured 1e-5 - 3.5e-5 and
    2e-6 - 3.3e-5); three products (operator to 16 bits, 1e-4 - 1e-3) are
    reported beside them."""
    import copy
    p = _ba_problem(N, nsteps=50)
    p.src_xy = np.array([[300.0, 700.0], [650.0, 420.0], [820.0, 160.0]])
    p.f0 = np.repeat(p.f0, 3)
    p.t0 = np.repeat(p.t0, 3)
    p.dirs = np.array([[0.0, -1.0], [0.6, 0.8], [1.0, 0.0]])
    p.__post_init__()
    runs = {}
    for precision in ("fp32c", "bf16x5", "bf16x3"):
        eng, s = W.run_batched(p, 8, precision=precision)
        assert eng.status() == 0
        eng.close()
        runs[precision] = s
    for k in range(3):
        q = copy.copy(p)
        q.src_xy, q.f0, q.t0, q.dirs = p.src_xy[k:k + 1], p.f0[k:k + 1], p.t0[k:k + 1], p.dirs[k:k + 1]
        q.__post_init__()
        _, sref = oracle_run(q)
        errs = {prec: rel_l2(s[:, :, k], sref) for prec, s in runs.items()}
        say(f"  BA p={N} column {k} vs oracle: " + ", ".join(f"{a} {b:.2e}" for a, b in errs.items()))
        assert errs["fp32c"] <= 1e-4 and errs["bf16x5"] <= 1e-4


@pytest.mark.parametrize("N", range(1, 9))
def test_boundary_adapted_operator_and_fp64_trajectory(N):
    """Non-orthonormal basis (dense inverse mass, Matrices_imp.hpp:183-190): an
    FP64 context keeps B assembled from the BoundaryAdapted tables (equal to the
    oracle's B to 1e-13) and applies M^-1 = invm_e * Mb^-1 after the product in
    the reference's order, so its run matches the oracle's to <= 1e-12 in u and
    seismograms and its CFL estimate is lambda_max(M^-1 B); an FP32C context
    folds the inverse mass into the operator, (I2 (x) Mb^-1) B."""
    p = _ba_problem(N)
    eng, u, s = W.run_engine(p, "fp64")
    blk, invm = eng.get_blocks()
    op = oracle_op(p)
    ref = op.blocks()
    err = np.abs(blk - ref).max() / np.abs(ref).max()
    np.testing.assert_array_equal(invm, op.invm())
    uref, sref = oracle_run(p)
    e_u, (e_s, e_r) = rel_l2(u, uref), seis_err(s, sref)
    _, lm = eng.cfl(1.0, 400)
    e_l = abs(lm / op.lambda_max(400) - 1)
    say(f"  BA p={N}: blocks {err:.2e}; u {e_u:.2e}; seis {e_s:.2e} (worst receiver {e_r:.2e}); "
        f"lambda_max {e_l:.1e}")
    assert err <= 1e-13
    assert e_u <= 1e-12 and e_s <= 1e-12 and e_r <= 1e-12
    assert e_l <= 1e-6
    eng.close()
    f = capi.Engine("fp32c")
    W.set_space(f, p)
    f.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    blk32, _ = f.get_blocks()
    K = np.kron(np.eye(2), O.base_invmass(N, O.BOUNDARY_ADAPTED))
    folded = np.einsum("ij,esjk->esik", K, ref)
    assert np.abs(blk32 - folded).max() / np.abs(folded).max() <= 1e-13
    f.close()


@pytest.mark.parametrize("N", [3, 5])
def test_boundary_adapted_seismograms_equal_dubiner(N):
    """The discrete dynamics do not depend on the basis: the BoundaryAdapted and
    Dubiner FP64 runs give the same seismograms (<= 1e-10).  The FP32c
    BoundaryAdapted run stays within 1e-3 of the oracle: the non-orthonormal
    basis's mass matrix (condition number ~1e2-1e3 at p = 3..5) amplifies FP32
    rounding of the folded operator (measured 3.7e-5 at p = 3, 1.9e-4 at p = 5,
    against ~1e-5 for the orthonormal basis)."""
    p = _ba_problem(N, nsteps=200)
    q = small_problem(N, nx=4, ny=3, nsteps=200, seed=50 + N, nrec=5)
    _, _, s_ba = W.run_engine(p, "fp64")
    _, _, s_d = W.run_engine(q, "fp64")
    e = rel_l2(s_ba, s_d)
    _, _, s32 = W.run_engine(p, "fp32c")
    _, sref = oracle_run(p)
    e32 = rel_l2(s32, sref)
    say(f"  BA vs Dubiner p={N}: {e:.2e}; BA fp32c vs oracle {e32:.2e}")
    assert e <= 1e-10 and e32 <= 1e-3


# ------------------------------------------------------------------ FP64 trajectories
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
def test_fp64_trajectory_matches_oracle(N):
    p = small_problem(N, nsteps=150)
    rng = np.random.default_rng(N)
    u0 = rng.standard_normal(p.ndof) * 1e-6
    v0 = rng.standard_normal(p.ndof) * 1e-4
    say(f"fp64 trajectory N={N} ne={p.ne} steps={p.nsteps}")
    eng, u, s = W.run_engine(p, "fp64", u0=u0, v0=v0)
    uref, sref = oracle_run(p, u0=u0, v0=v0)
    e_u = rel_l2(u, uref)
    e_s, e_r = seis_err(s, sref)
    say(f"  u rel L2 {e_u:.3e}  seis rel L2 {e_s:.3e}  worst receiver {e_r:.3e}")
    assert s.shape == sref.shape
    assert e_u <= 1e-12 and e_s <= 1e-12 and e_r <= 1e-12
    assert eng.status() == 0


def test_fp64_large_operator_step_matches_oracle():
    """FP64 operators above 128 MB (here 1400 triangles at p = 6: 140 MB) take the
    step whose store of u^{n+1} goes through L2; it stores the same values, so it
    matches the oracle to the same 1e-12 as the plain-store step of small meshes."""
    p = W.layered_ricker(nx=35, ny=20, N=6, nsteps=100, nrec=16)
    p.t0 = np.array([10 * p.dt])
    p.f0 = np.array([1.0 / (8 * p.dt)])
    # receivers on a line through the source, which 100 steps reach
    p.rec_xy = np.stack([np.linspace(1600.0, 2400.0, 16), np.full(16, 1750.0)], 1)
    p.__post_init__()
    d = 2 * p.nln
    assert p.ne * 4 * d * d * 8 > 128 * 2 ** 20
    say(f"large-operator FP64 ne={p.ne} operator {p.ne * 4 * d * d * 8 / 2**20:.0f} MiB steps={p.nsteps}")
    eng, u, s = W.run_engine(p, "fp64")
    assert eng.status() == 0
    uref, sref = oracle_run(p)
    e_u, (e_s, e_r) = rel_l2(u, uref), seis_err(s, sref)
    say(f"  u {e_u:.2e}  seismograms {e_s:.2e} (worst receiver {e_r:.2e})")
    assert e_u <= 1e-12 and e_s <= 1e-12 and e_r <= 1e-12
    eng.close()


@pytest.mark.parametrize("basis", [O.DUBINER, O.BOUNDARY_ADAPTED])
@pytest.mark.parametrize("N", range(1, 9))
def test_fp64_face_form_against_block_ell_and_oracle(N, basis):
    """The FP64 face form (each interior face's block read once, from the
    lower-numbered element, and applied to both sides) against the step that
    reads every stored block, and both against the oracle: within 1e-12 of the
    oracle and of each other (they differ by the assembled blocks' rounding
    asymmetry and the order of the slot sums: ~1e-13 on the BoundaryAdapted
    basis, whose runs are themselves ~1e-13 from the oracle)."""
    p = small_problem(N, nsteps=150)
    if basis != O.DUBINER:
        p.basis = basis
        p.__post_init__()
    rng = np.random.default_rng(10 + N)
    u0 = rng.standard_normal(p.ndof) * 1e-6
    v0 = rng.standard_normal(p.ndof) * 1e-4
    runs = {}
    for on in (True, False):
        eng = capi.Engine("fp64")
        eng.set_face_form("always" if on else False)   # a small operator: forced
        eng, u, s = W.run_engine(p, engine=eng, u0=u0, v0=v0)
        assert eng.set_face_form() is on
        assert eng.status() == 0
        runs[on] = (u, s)
        eng.close()
    uref, sref = oracle_run(p, u0=u0, v0=v0)
    (uf, sf), (ue, se) = runs[True], runs[False]
    say(f"fp64 face form N={N} basis={basis}: vs oracle u {rel_l2(uf, uref):.2e} seis {rel_l2(sf, sref):.2e}; "
        f"vs block-ELL u {rel_l2(uf, ue):.2e} seis {rel_l2(sf, se):.2e}")
    assert rel_l2(uf, uref) <= 1e-12 and max(seis_err(sf, sref)) <= 1e-12
    assert rel_l2(ue, uref) <= 1e-12 and max(seis_err(se, sref)) <= 1e-12
    assert rel_l2(uf, ue) <= 1e-12 and rel_l2(sf, se) <= 1e-12


def test_fp64_face_form_availability_and_uploaded_blocks():
    """The face form is taken by default for assembled FP64 operators above
    the cache-resident size (128 MB) at every degree, on request ("always") for
    smaller ones, by FP32c contexts only on request, and for uploaded blocks only
    when they are symmetric bit for bit; on such blocks it reads the same
    entries as the block-ELL step (they differ in the order of the slot sums
    only)."""
    for N in (2, 4, 5):
        p = small_problem(N, nsteps=5)
        eng = capi.Engine("fp64")
        W.set_space(eng, p)
        eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
        # out-of-range settings are refused and leave the setting as it was
        assert eng._L.tsp_set_face_form(eng._h, 3, None) == capi.TSP_ERR_ARG
        assert eng._L.tsp_set_face_form(eng._h, -2, None) == capi.TSP_ERR_ARG
        assert eng.set_face_form() is False          # a small operator
        assert eng.set_face_form("always") is True
        assert eng.set_face_form(False) is False and eng.set_face_form("always") is True
        eng.close()
    big = W.layered_ricker(nx=35, ny=20, N=6, nsteps=5)   # 140 MB operator
    eng = capi.Engine("fp64")
    W.set_space(eng, big)
    eng.set_mesh(big.verts, big.neigh, big.neighedge, big.lam, big.mu, big.rho, big.penalty)
    assert eng.set_face_form() is True
    eng.close()
    f = capi.Engine("fp32c")
    p = small_problem(5, nsteps=5)
    W.set_space(f, p)
    f.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    assert f.set_face_form() is False            # FP32c: only when asked
    f.close()
    p = small_problem(6, nsteps=60)
    op = oracle_op(p)
    blocks = op.blocks().copy()
    # the assembled blocks are symmetric to rounding only: uploaded as they are,
    # every stored block is read
    eng = capi.Engine("fp64")
    eng.set_face_form("always")
    eng.set_tables(p.N)
    eng.upload_blocks(p.neigh, blocks, op.invm())
    assert eng.set_face_form() is False
    eng.close()
    # made symmetric bit for bit (B_{b,a} := B_{a,b}^T for a < b)
    nb = np.asarray(p.neigh).reshape(-1, 3)
    for a in range(p.ne):
        for s in range(3):
            b = nb[a, s]
            if b > a:
                sb = list(nb[b]).index(a)
                blocks[b, 1 + sb] = blocks[a, 1 + s].T
    out = {}
    for on in (True, False):
        eng = capi.Engine("fp64")
        eng.set_tables(p.N)
        eng.set_face_form("always" if on else False)
        eng.upload_blocks(p.neigh, blocks, op.invm())
        assert eng.set_face_form() is on
        eng.set_sources(p.src_elem, p.src_phi, p.wavelet())
        eng.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
        eng.set_dt(p.dt)
        eng.set_state()
        eng.first_step()
        eng.run(p.nsteps)
        out[on] = (eng.get_u(), eng.seismograms())
        assert eng.status() == 0
        eng.close()
    say(f"uploaded symmetric blocks N=6: face form vs block-ELL u {rel_l2(out[True][0], out[False][0]):.2e}")
    assert rel_l2(out[True][0], out[False][0]) <= 1e-12 and rel_l2(out[True][1], out[False][1]) <= 1e-12


def test_fp64_face_form_shared_operator_and_toggle_mid_run():
    """A context borrowing the operator (tsp_share_operator) builds its own face
    list and products; switching the face form off and on between runs keeps
    the trajectory (to 1e-12 of a run that never switches)."""
    p = small_problem(7, nsteps=80)
    e1 = capi.Engine("fp64")
    e1.set_face_form("always")
    e1, u1, s1 = W.run_engine(p, engine=e1)
    b = capi.Engine("fp64", share=e1)
    assert b.set_face_form() is True
    b.set_sources(p.src_elem, p.src_phi, p.wavelet())
    b.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
    b.set_dt(p.dt)
    b.set_state()
    b.first_step()
    b.run(30)
    assert b.set_face_form(False) is False
    b.run(20)
    assert b.set_face_form("always") is True
    b.run(p.nsteps - 50)
    u2, s2 = b.get_u(), b.seismograms()
    say(f"shared operator, face form toggled: u {rel_l2(u2, u1):.2e} seis {rel_l2(s2, s1):.2e}")
    assert rel_l2(u2, u1) <= 1e-12 and rel_l2(s2, s1) <= 1e-12
    b.close()
    e1.close()


def test_square_ricker_full_parity():
    """The 1 km square at full size (3200 triangles, p=4, 2000 steps, 32 receivers):
    final displacement and every seismogram within 1e-12 of the oracle."""
    p = W.square_ricker()
    say(f"square_ricker ne={p.ne} dof={p.ndof} dt={p.dt:.4e} steps={p.nsteps}")
    t = time.time()
    eng, u, s = W.run_engine(p, "fp64")
    say(f"  gpu done in {time.time() - t:.2f}s")
    t = time.time()
    uref, sref = oracle_run(p)
    say(f"  oracle done in {time.time() - t:.2f}s")
    e_u = rel_l2(u, uref)
    e_s, e_r = seis_err(s, sref)
    say(f"  u rel L2 {e_u:.3e}  seis rel L2 {e_s:.3e}  worst receiver {e_r:.3e}  |s|max {np.abs(sref).max():.3e}")
    assert np.abs(sref).max() > 0
    assert e_u <= 1e-12 and e_s <= 1e-12 and e_r <= 1e-12
    np.save(os.path.join(os.environ.get("TSP_OUT", "/tmp"), "square_ricker_seis_fp64.npy"), s) if os.environ.get("TSP_OUT") else None


def test_fp32c_error_vs_fp64_oracle_square_ricker():
    """FP32 blocks + compensated update stay within 1e-4 of the FP64 oracle on
    the 1 km square's seismograms over 2000 steps."""
    p = W.square_ricker()
    say(f"fp32c square_ricker steps={p.nsteps}")
    eng, u, s = W.run_engine(p, "fp32c")
    uref, sref = oracle_run(p)
    e_u = rel_l2(u, uref)
    e_s, e_r = seis_err(s, sref)
    say(f"  fp32c: u rel L2 {e_u:.3e}  seis rel L2 {e_s:.3e}  worst receiver {e_r:.3e}")
    assert e_s <= 1e-4 and e_r <= 1e-4
    assert eng.status() == 0


@pytest.mark.parametrize("N", range(1, 9))
def test_fp32c_small_many_degrees(N):
    """The single-column FP32c step of every degree (k_step32c<1..8>) against the
    oracle over 400 steps."""
    p = small_problem(N, nsteps=400)
    if N == 1:
        p.dt *= 0.5      # 0.2 h / (vp p^2) is above the leap-frog limit at p = 1
    eng, u, s = W.run_engine(p, "fp32c")
    uref, sref = oracle_run(p)
    e_s, _ = seis_err(s, sref)
    say(f"fp32c N={N}: u rel L2 {rel_l2(u, uref):.3e} seis rel L2 {e_s:.3e}")
    assert e_s <= 1e-4 and rel_l2(u, uref) <= 1e-4


@pytest.mark.parametrize("N", [5, 6, 7, 8])
def test_fp32c_face_form_against_oracle_and_block_ell(N):
    """The FP32c face form (each interior face's FP32 block read once, both
    sides' translation-free products from it), forced on a small mesh, against
    the oracle over 400 steps (the FP32c bar, 1e-4) and against the FP32c step
    reading every stored block (FP32 rounding of different sums)."""
    p = small_problem(N, nsteps=400)
    runs = {}
    for on in (True, False):
        eng = capi.Engine("fp32c")
        assert eng.set_face_form() is False            # FP32c: off unless asked
        eng.set_face_form("always" if on else False)
        eng, u, s = W.run_engine(p, engine=eng)
        assert eng.set_face_form() is on
        assert eng.status() == 0
        runs[on] = (u, s)
        eng.close()
    uref, sref = oracle_run(p)
    (uf, sf), (ue, se) = runs[True], runs[False]
    say(f"fp32c face form N={N}: vs oracle u {rel_l2(uf, uref):.2e} seis {rel_l2(sf, sref):.2e}; "
        f"block-ELL vs oracle u {rel_l2(ue, uref):.2e} seis {rel_l2(se, sref):.2e}; "
        f"face vs block-ELL seis {rel_l2(sf, se):.2e}")
    assert rel_l2(uf, uref) <= 1e-4 and max(seis_err(sf, sref)) <= 1e-4
    assert rel_l2(sf, se) <= 1e-4


def test_fp32c_face_form_availability():
    """FP32c contexts take the face form only when asked, at N >= 5, on the
    orthonormal basis (a BoundaryAdapted FP32 operator has the dense inverse
    mass folded into its rows and is not symmetric)."""
    for N, basis, want in ((4, O.DUBINER, False), (5, O.DUBINER, True), (5, O.BOUNDARY_ADAPTED, False)):
        p = small_problem(N, nsteps=5)
        if basis != O.DUBINER:
            p.basis = basis
            p.__post_init__()
        eng = capi.Engine("fp32c")
        W.set_space(eng, p)
        eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
        assert eng.set_face_form() is False
        assert eng.set_face_form("always") is want, (N, basis)
        eng.close()


# ------------------------------------------------------------------ API paths
def test_upload_blocks_path_matches_assembled_path():
    p = small_problem(4, nsteps=60)
    op = oracle_op(p)
    e1, u1, s1 = W.run_engine(p, "fp64")
    eng = capi.Engine("fp64")
    eng.set_tables(p.N)
    eng.upload_blocks(p.neigh, op.blocks(), op.invm())
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet())
    eng.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
    eng.set_dt(p.dt)
    eng.set_state()
    eng.first_step()
    eng.run(p.nsteps)
    u2, s2 = eng.get_u(), eng.seismograms()
    assert rel_l2(u2, u1) <= 1e-13 and rel_l2(s2, s1) <= 1e-13


def test_stride_and_eval_receivers():
    """Fused sampling every 3 steps equals the stride-1 record decimated, and an
    explicit eval_receivers() after the run equals the last fused sample."""
    p = small_problem(3, nsteps=30)
    _, _, s1 = W.run_engine(p, "fp64", stride=1)
    eng, _, s3 = W.run_engine(p, "fp64", stride=3)
    np.testing.assert_array_equal(s3[0], s1[0])
    np.testing.assert_allclose(s3[1:], s1[3::3], rtol=0, atol=1e-14 * np.abs(s1).max())
    # stepping in pieces equals stepping at once
    eng2 = capi.Engine("fp64")
    eng2.set_tables(p.N)
    eng2.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    eng2.set_sources(p.src_elem, p.src_phi, p.wavelet())
    eng2.set_receivers(p.rec_elem, p.rec_phi, 0, 4)
    eng2.set_dt(p.dt)
    eng2.set_state()
    eng2.first_step()
    eng2.eval_receivers()
    for chunk in (7, 13, 10):
        eng2.run(chunk)
    eng2.eval_receivers()
    s = eng2.seismograms()
    assert s.shape[0] == 2
    np.testing.assert_array_equal(s[0], s1[0])
    np.testing.assert_allclose(s[1], s1[-1], rtol=0, atol=1e-14 * np.abs(s1).max())


def test_cfl_lambda_max_matches_oracle():
    p = small_problem(3, nx=4, ny=4)
    eng = capi.Engine("fp64")
    eng.set_tables(p.N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    dt, lm = eng.cfl(1.0, 400)
    ref = oracle_op(p).lambda_max(400)
    say(f"lambda_max gpu {lm:.6e} oracle {ref:.6e}")
    assert abs(lm - ref) <= 1e-6 * ref
    assert abs(dt - 2 / np.sqrt(lm)) <= 1e-12 * dt
    assert p.dt < dt            # the workload step size is stable


def test_nonfinite_detected_and_errors_reported():
    p = small_problem(2, nsteps=400)
    p.dt *= 50                   # far above the stability limit
    eng, u, s = W.run_engine(p, "fp64")
    assert eng.status() & capi.TSP_FLAG_NONFINITE
    e = capi.Engine("fp64")
    with pytest.raises(capi.EngineError) as ex:
        e.run(1)
    assert ex.value.code == capi.TSP_ERR_STATE
    e.set_tables(2)
    e.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    e.set_sources(p.src_elem, p.src_phi, p.wavelet(5))
    e.set_dt(p.dt)
    e.set_state()
    e.first_step()
    with pytest.raises(capi.EngineError) as ex:
        e.run(10)                # wavelet table has 5 rows
    assert ex.value.code == capi.TSP_ERR_ARG
    bad = p.neigh.copy()
    bad[0, 0] = p.ne + 5
    with pytest.raises(capi.EngineError):
        e.set_mesh(p.verts, bad, p.neighedge, p.lam, p.mu, p.rho, p.penalty)


def test_state_errors_and_zero_step_runs():
    """Calls out of order or past a buffer are refused with the documented code
    and leave the context usable; a zero-step run changes nothing."""
    p = small_problem(3, nsteps=30, nrec=4)
    e = capi.Engine("fp64")
    e.set_tables(3)
    e.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    with pytest.raises(capi.EngineError) as ex:       # the operator is already assembled
        e.set_mass(np.eye(p.nln))
    assert ex.value.code == capi.TSP_ERR_STATE
    e.set_sources(p.src_elem, p.src_phi, p.wavelet(31))
    e.set_receivers(p.rec_elem, p.rec_phi, 1, 11)     # room for u1 + 10 samples
    e.set_dt(p.dt)
    e.set_state()
    e.first_step()
    e.run(0)
    assert e.steps_done() == 0
    e.run(10)
    with pytest.raises(capi.EngineError) as ex:       # the 12th sample has no slot
        e.run(1)
    assert ex.value.code == capi.TSP_ERR_ARG
    assert e.steps_done() == 10
    u = e.get_u()
    e.run(0)
    np.testing.assert_array_equal(e.get_u(), u)
    # the same 10 steps as a fresh run with a large buffer
    _, u_ref, s_ref = W.run_engine(small_problem(3, nsteps=10, nrec=4), "fp64")
    np.testing.assert_array_equal(u, u_ref)
    np.testing.assert_array_equal(e.seismograms(), s_ref)
    e.close()
    # a set takes only first-stepped FP32C contexts (FP32 products, any width)
    m = capi.Multi(chunk=4)
    for prec, ncols in (("fp64", 1), ("bf16x3", 32)):
        q = capi.Engine(prec, ncols=ncols)
        q.set_tables(3)
        q.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
        q.set_dt(p.dt)
        for c in range(ncols):
            q.set_state(col=c)
        q.first_step()
        with pytest.raises(capi.EngineError) as ex:
            m.add(q, 5)
        assert ex.value.code == capi.TSP_ERR_UNSUPPORTED
        q.close()
    q = capi.Engine("fp32c")
    q.set_tables(3)
    q.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    with pytest.raises(capi.EngineError) as ex:       # no first step yet
        m.add(q, 5)
    assert ex.value.code == capi.TSP_ERR_STATE
    q.close()
    m.close()


def test_single_element_mesh():
    """A lone triangle: every face is a boundary face, B = A."""
    pts = np.array([[0.0, 0.0], [100.0, 10.0], [20.0, 90.0]])
    tri = np.array([[0, 1, 2]], np.int32)
    p = W.Problem("one", 3, pts, tri, np.array([6e9]), np.array([5e9]), np.array([2000.0]), 10.0, 1e-5, 50,
                  np.array([[40.0, 30.0]]), np.array([50.0]), np.array([0.01]), np.array([[1.0, 0.0]]),
                  np.array([[30.0, 40.0]]))
    eng, u, s = W.run_engine(p, "fp64")
    uref, sref = oracle_run(p)
    assert rel_l2(u, uref) <= 1e-12 and rel_l2(s, sref) <= 1e-12


# ------------------------------------------------------------------ full-size layered workload
def _submesh(p, centre):
    """Elements `centre` plus their neighbours, renumbered; returns the sub-problem
    arrays and the local index of each centre element."""
    keep = set(int(e) for e in centre)
    for e in centre:
        keep.update(int(x) for x in p.neigh[e] if x >= 0)
    keep = np.array(sorted(keep))
    tri = p.tri[keep]
    used, inv = np.unique(tri, return_inverse=True)
    pts = p.pts[used]
    tri = inv.reshape(-1, 3).astype(np.int32)
    nb, ng = capi.neighbors(tri)
    loc = np.searchsorted(keep, centre)
    return pts, tri, nb, ng, keep, loc


def test_layered_full_size_blocks_and_linearity():
    """The layered workload at full size (160k triangles, p=6, 16 GB of FP64 blocks):
    (a) blocks of 64 sampled elements equal the oracle's blocks assembled on the
        sub-mesh of those elements and their neighbours;
    (b) doubling the force doubles every displacement and sample bit for bit
        (every operation of the step is linear and scaling by 2 is exact)."""
    p = W.layered_ricker(nsteps=20)
    say(f"layered_ricker ne={p.ne} dof={p.ndof}")
    eng = capi.Engine("fp64")
    t = time.time()
    eng.set_tables(p.N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    say(f"  assembled in {time.time() - t:.2f}s, device bytes {eng.device_bytes() / 1e9:.2f} GB")
    rng = np.random.default_rng(0)
    centre = np.unique(np.concatenate([rng.integers(0, p.ne, 60), [0, p.ne - 1, int(p.src_elem[0]), 1000]]))
    pts, tri, nb, ng, keep, loc = _submesh(p, centre)
    op = O.Operator(p.N, pts[tri].reshape(-1, 6), nb, ng, p.lam[keep], p.mu[keep], p.rho[keep], p.penalty)
    ref = op.blocks()
    worst = 0.0
    for e, l in zip(centre, loc):
        blk, im = eng.get_blocks(int(e), 1)
        worst = max(worst, np.abs(blk[0] - ref[l]).max() / np.abs(ref[l]).max())
        assert im[0] == op.invm()[l]
    say(f"  sampled block max rel diff {worst:.3e}")
    assert worst <= 1e-13
    # (b) linearity at full size
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet())
    eng.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
    eng.set_dt(p.dt)
    eng.set_state()
    eng.first_step()
    eng.run(p.nsteps)
    u1, s1 = eng.get_u(), eng.seismograms()
    eng.set_sources(p.src_elem, p.src_phi, 2.0 * p.wavelet())
    eng.set_state()
    eng.first_step()
    eng.run(p.nsteps)
    u2, s2 = eng.get_u(), eng.seismograms()
    assert np.abs(u1).max() > 0
    np.testing.assert_array_equal(u2, 2.0 * u1)
    np.testing.assert_array_equal(s2, 2.0 * s1)
    assert eng.status() == 0


def test_layered_full_size_face_form_against_block_ell():
    """The layered workload at full size (160k triangles, p = 6, 16 GB of FP64
    blocks), whose operator takes the face form by default: 300 steps with the
    source firing early (t0 = 20 dt, f0 = 1/(16 dt)) and 16 receivers 5-30 m
    from it, against the same context stepping every stored block (the
    block-ELL step, itself the oracle to 1e-12 on the meshes the oracle runs):
    within 1e-12 in u and in every seismogram, the wave having reached the
    receivers."""
    p = W.layered_ricker(nsteps=300)
    p.t0 = np.array([20 * p.dt])
    p.f0 = np.array([1.0 / (16 * p.dt)])
    ang = np.arange(4) * np.pi / 2 + 0.3
    p.rec_xy = np.array([[p.src_xy[0, 0] + r * np.cos(a), p.src_xy[0, 1] + r * np.sin(a)]
                         for r in (5.0, 10.0, 20.0, 30.0) for a in ang])
    p.__post_init__()
    eng = capi.Engine("fp64")
    W.set_space(eng, p)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet())
    eng.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
    eng.set_dt(p.dt)
    out = {}
    for on in (True, False):
        assert eng.set_face_form(on) is on
        eng.set_state()
        eng.first_step()
        eng.run(p.nsteps)
        out[on] = (eng.get_u(), eng.seismograms())
        assert eng.status() == 0
    eng.close()
    (uf, sf), (ue, se) = out[True], out[False]
    eu, es = rel_l2(uf, ue), seis_err(sf, se)
    say(f"  layered full size, face form vs block-ELL: u {eu:.2e}, seismograms {es[0]:.2e} "
        f"(worst receiver {es[1]:.2e}), |s|max {np.abs(se).max():.2e}")
    assert np.abs(se).max() > 1e-3 * np.abs(ue).max()
    assert eu <= 1e-12 and es[0] <= 1e-12 and es[1] <= 1e-12


@pytest.mark.parametrize("basis", [0, 1])
def test_cpp_facade_driver_matches_oracle(tmp_path, basis):
    """examples/square_driver (the C++ LeapFrog facade, reference-style loop with
    one step() and one eval_receivers() per iteration) against the oracle, with
    the Dubiner space and with FESpace<N, Dunavant<2N>, BoundaryAdapted<N>>."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples", "square_driver")
    assert os.path.exists(exe), "build the example with python -m tspeed_amd.build"
    steps = 150
    out = str(tmp_path / "c1")
    # 17 digits: the files carry every bit (the default is the reference's 6)
    r = subprocess.run([exe, str(steps), out, "0", str(basis), "17"], capture_output=True, text=True, timeout=240)
    say(r.stdout.strip(), r.stderr.strip()[-500:])
    assert r.returncode == 0
    from tspeed_amd import io
    s = io.read_receivers(out, 32)
    p = W.square_ricker(nsteps=steps)
    if basis:
        p.basis = capi.BOUNDARY_ADAPTED
        p.__post_init__()
    _, sref = oracle_run(p)
    assert s.shape == sref.shape
    e_s, e_r = seis_err(s, sref)
    say(f"  facade driver: seis rel L2 {e_s:.3e} worst receiver {e_r:.3e}")
    # both bases apply the inverse mass in the reference's order (BoundaryAdapted:
    # after the block product, dense per element), so one bar holds for both
    assert e_s <= 1e-12 and e_r <= 1e-12


# ------------------------------------------------------------------ many columns (MFMA)
def test_batched_columns_match_oracle_and_single_column_path():
    """64 sources as 64 columns of one block SpMM (FP32 on MFMA, compensated
    update): every column's seismograms within 1e-4 of the FP64 oracle run of
    that source alone, and close to the single-column FP32 path."""
    p = W.multisource(nsteps=400)
    p.t0 = np.full(64, 20 * p.dt)               # early wavelets: signal within 400 steps
    p.f0 = 1.0 / (10 * p.dt) * np.linspace(0.6, 1.0, 64)
    say(f"batched multisource ne={p.ne} cols=64 steps={p.nsteps}")
    eng, s = W.run_batched(p, 64)
    assert s.shape == (p.nsamples, 128, 64, 2)
    assert eng.status() == 0
    worst = 0.0
    for k in (0, 17, 40, 63):
        q = W.multisource_column(k, nsteps=p.nsteps)
        q.t0, q.f0 = p.t0[k:k + 1], p.f0[k:k + 1]
        _, sref = oracle_run(q)
        e_s, e_r = seis_err(s[:, :, k, :], sref)
        _, _, s1 = W.run_engine(q, "fp32c")
        e_1 = rel_l2(s[:, :, k, :], s1)
        say(f"  column {k}: vs oracle {e_s:.3e} (worst receiver {e_r:.3e}); vs single-column fp32c {e_1:.3e}")
        worst = max(worst, e_s, e_r)
        assert e_1 <= 1e-4
    assert worst <= 1e-4


@pytest.mark.parametrize("precision,tol", [("bf16x3", 2e-3), ("bf16x5", 1e-4)])
def test_batched_split_products_error_against_the_oracle(precision, tol):
    """The split BF16 product variants on the 64-source square (64 columns,
    400 steps): each sampled column's seismograms within 2e-3 (three products,
    the operator kept to 16 bits) or 1e-4 (five products, the whole FP32
    operator) of the FP64 oracle; the measured errors are printed beside the
    distance to the FP32 products."""
    p = W.multisource(nsteps=400)
    p.t0 = np.full(64, 20 * p.dt)
    p.f0 = 1.0 / (10 * p.dt) * np.linspace(0.6, 1.0, 64)
    eng, s = W.run_batched(p, 64, precision=precision)
    assert eng.status() == 0
    _, s32 = W.run_batched(p, 64)
    worst = 0.0
    for k in (0, 17, 40, 63):
        q = W.multisource_column(k, nsteps=p.nsteps)
        q.t0, q.f0 = p.t0[k:k + 1], p.f0[k:k + 1]
        _, sref = oracle_run(q)
        e_s, e_r = seis_err(s[:, :, k, :], sref)
        e_32 = rel_l2(s[:, :, k, :], s32[:, :, k, :])
        say(f"  {precision} column {k}: vs oracle {e_s:.3e} (worst receiver {e_r:.3e}); vs FP32 products {e_32:.3e}")
        worst = max(worst, e_s)
    assert worst <= tol
    eng.close()


@pytest.mark.parametrize("precision", ["fp32c", "bf16x3", "bf16x5"])
@pytest.mark.parametrize("N", [2, 4, 5, 7])
def test_narrow_and_wide_contexts_agree(N, precision):
    """8- and 16-column contexts fill part of the 32-column MFMA tile and every
    output column is summed on its own, and every width of one degree runs the
    same kernel (one slot per wave for one-tile blocks, one wave per tile
    otherwise), so a source's seismograms and final state are bit for bit the
    same at widths 8, 16, 32 and 64."""
    p = small_problem(N, nx=5, ny=4, nsteps=30, seed=50 + N, nrec=6)
    p.src_xy = np.array([[220.0, 640.0], [700.0, 500.0], [480.0, 180.0]])
    p.f0 = np.repeat(p.f0, 3)
    p.t0 = np.repeat(p.t0, 3)
    p.dirs = np.array([[0.0, -1.0], [0.6, 0.8], [1.0, 0.0]])
    p.__post_init__()
    runs = {}
    for ncols in (8, 16, 32, 64):
        eng, s = W.run_batched(p, ncols, precision=precision)
        assert eng.status() == 0
        runs[ncols] = (s[:, :, :3], [eng.get_u(c) for c in range(3)])
        assert not s[:, :, 3:].any()
        eng.close()
    for ncols in (16, 32, 64):
        np.testing.assert_array_equal(runs[ncols][0], runs[8][0])
        for x, y in zip(runs[ncols][1], runs[8][1]):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp32c", "bf16x3", "bf16x5"])
@pytest.mark.parametrize("N", [1, 2, 3, 5, 6, 7, 8])
@pytest.mark.parametrize("ncols", [8, 16, 32, 64])
def test_batched_every_degree(N, ncols, precision):
    """Every degree on the MFMA path (row tiles 1..3, padded row octets): three
    sources in columns 0, 5 and ncols-1, each within 1e-4 (FP32 products, five
    split BF16 products) or 2e-3 (three split products) of the FP64 oracle and
    of the single-column FP32 path run with that source alone; the other
    columns stay exactly zero."""
    tol = 2e-3 if precision == "bf16x3" else 1e-4
    # dt = 0.2 h / (vp p^2) is just above the leap-frog limit 2/sqrt(lambda_max)
    # at p = 1 (ratio 1.05 on this mesh): halve it there so the run is stable
    dt_scale = 0.5 if N == 1 else 1.0
    base = small_problem(N, nx=4, ny=3, nsteps=40, seed=30 + N, nrec=5)
    cols = [0, 5, ncols - 1]
    xy = np.array([[300.0, 700.0], [650.0, 420.0], [820.0, 160.0]])
    p = small_problem(N, nx=4, ny=3, nsteps=40, seed=30 + N, nrec=5)
    p.src_xy = xy
    p.f0 = np.repeat(base.f0, 3)
    p.t0 = np.repeat(base.t0, 3)
    p.dirs = np.array([[0.0, -1.0], [0.6, 0.8], [1.0, 0.0]])
    p.dt *= dt_scale
    p.__post_init__()
    eng = capi.Engine(precision, ncols=ncols)
    eng.set_tables(p.N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet(), col=np.array(cols))
    eng.set_receivers(p.rec_elem, p.rec_phi, 1, p.nsamples)
    eng.set_dt(p.dt)
    for c in range(ncols):
        eng.set_state(None, None, col=c)
    eng.first_step()
    eng.run(p.nsteps)
    s = eng.seismograms()
    assert eng.status() == 0
    others = [c for c in range(ncols) if c not in cols]
    assert not s[:, :, others].any()
    assert not eng.get_u(others[0]).any()
    for k, c in enumerate(cols):
        q = small_problem(N, nx=4, ny=3, nsteps=40, seed=30 + N, nrec=5)
        q.src_xy, q.dirs = xy[k:k + 1], p.dirs[k:k + 1]
        q.dt = p.dt
        q.__post_init__()
        uref, sref = oracle_run(q)
        _, u1, s1 = W.run_engine(q, "fp32c")
        e_o, e_1 = rel_l2(s[:, :, c], sref), rel_l2(s[:, :, c], s1)
        e_u = rel_l2(eng.get_u(c), uref)
        say(f"  {precision} p={N} cols={ncols} col {c}: vs oracle {e_o:.2e} (u {e_u:.2e}), vs fp32c {e_1:.2e}")
        assert e_o <= tol and e_u <= tol and e_1 <= tol
    eng.close()


@pytest.mark.parametrize("precision", ["fp32c", "bf16x3", "bf16x5"])
def test_batched_graph_replay_equals_step_launches(precision):
    """Multi-column runs replay a captured graph of 32 steps whose steps read
    their index from a device counter that the graph itself advances, and
    launch the rest step by step; the result is bit-identical to the same
    steps launched one at a time (runs of 3), with a sampling stride of 2:
    two replays in one run (70 = 2 x 32 + 6), one replay in each of two runs
    (40 + 40), and a fall-back after an extra eval_receivers sample."""
    p = W.multisource(nsrc=5, nx=6, ny=5, nsteps=100)
    outs = []
    for chunks in ([70, 10, 20], [40, 40, 20], [3] * 33 + [1]):
        eng = capi.Engine(precision, ncols=32)
        eng.set_tables(p.N)
        eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
        eng.set_sources(p.src_elem, p.src_phi, p.wavelet(), col=np.arange(5))
        eng.set_receivers(p.rec_elem, p.rec_phi, 2, 60)
        eng.set_dt(p.dt)
        eng.first_step()
        for n in chunks:
            eng.run(n)
        outs.append((eng.seismograms(), eng.get_u(0), eng.get_u(4), eng.steps_done()))
        eng.close()
    (s0, a0, b0, n0) = outs[-1]
    for (s1, a1, b1, n1) in outs[:-1]:
        assert n0 == n1 == 100
        np.testing.assert_array_equal(s0, s1)
        np.testing.assert_array_equal(a0, a1)
        np.testing.assert_array_equal(b0, b1)
    # an extra sample makes the slots irregular: the launches take over, samples stay in order
    eng = capi.Engine(precision, ncols=32)
    eng.set_tables(p.N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet(), col=np.arange(5))
    eng.set_receivers(p.rec_elem, p.rec_phi, 2, 60)
    eng.set_dt(p.dt)
    eng.first_step()
    eng.run(40)
    eng.eval_receivers()
    eng.run(60)
    s2 = eng.seismograms()
    assert s2.shape[0] == 52
    np.testing.assert_array_equal(s2[:21], s0[:21])
    np.testing.assert_array_equal(s2[22:], s0[21:])
    np.testing.assert_array_equal(eng.get_u(0), a0)
    eng.close()


def test_batched_state_roundtrip_and_eval_receivers():
    """set_state / get_u per column, eval_receivers, and linearity across columns:
    column 1 carries twice column 0's source and initial state (u0 and v0, so
    the first step's earlier time level u0 - dt v0 is exercised); column 0
    against the FP64 oracle from the same u0 and v0."""
    p = W.multisource(nsrc=2, nx=6, ny=5, nsteps=40)
    p.src_xy[1] = p.src_xy[0]
    p.f0[1], p.t0[1] = p.f0[0], p.t0[0]
    p.dirs[1] = 2 * p.dirs[0]
    p.__post_init__()
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal(p.ndof) * 1e-6
    v0 = rng.standard_normal(p.ndof) * 1e-3
    eng = capi.Engine("fp32c", ncols=32)
    eng.set_tables(p.N)
    eng.set_mesh(p.verts, p.neigh, p.neighedge, p.lam, p.mu, p.rho, p.penalty)
    eng.set_sources(p.src_elem, p.src_phi, p.wavelet(), col=np.array([0, 1]))
    eng.set_receivers(p.rec_elem, p.rec_phi, 0, 2)
    eng.set_dt(p.dt)
    for c in range(32):
        eng.set_state(u0 if c == 0 else (2 * u0 if c == 1 else None),
                      v0 if c == 0 else (2 * v0 if c == 1 else None), col=c)
    eng.first_step()
    eng.eval_receivers()
    eng.run(p.nsteps)
    eng.eval_receivers()
    s = eng.seismograms()
    u_0, u_1, u_5 = eng.get_u(0), eng.get_u(1), eng.get_u(5)
    assert not u_5.any()                                    # no source, zero state
    np.testing.assert_allclose(u_1, 2 * u_0, rtol=0, atol=1e-6 * np.abs(u_0).max())
    np.testing.assert_allclose(s[:, :, 1], 2 * s[:, :, 0], rtol=0, atol=1e-6 * np.abs(s[:, :, 0]).max())
    # against the FP64 oracle for column 0
    q = W.multisource_column(0, nx=6, ny=5, nsteps=p.nsteps, nrec=p.rec_xy.shape[0])
    q.src_xy, q.f0, q.t0, q.dirs = p.src_xy[:1], p.f0[:1], p.t0[:1], p.dirs[:1]
    q.__post_init__()
    uref, _ = oracle_run(q, u0=u0, v0=v0)
    say(f"  column 0 from u0, v0 after {p.nsteps} steps: u vs oracle {rel_l2(u_0, uref):.2e}")
    assert rel_l2(u_0, uref) <= 1e-4


# ------------------------------------------------------------------ sets of instances
def _mixed_set(steps=(40, 25, 60, 35, 50, 30, 45, 55), degrees=(2, 3, 4, 5, 6, 8, 7, 1)):
    """Eight meshes of different sizes, one of every degree 1..8."""
    probs = []
    for k, (n, N) in enumerate(zip(steps, degrees)):
        p = small_problem(N, nx=3 + k, ny=2 + k % 3, nsteps=n, seed=10 + k, nrec=5)
        if N == 1:
            p.dt *= 0.5      # 0.2 h / (vp p^2) is above the leap-frog limit at p = 1
        probs.append(p)
    return probs


def test_multi_set_equals_each_instance_alone():
    """Eight meshes of different size and degree stepped together in the persistent
    kernel (graph chunks of 8): each instance's seismograms and final state equal
    the same instance stepped alone by the single-context FP32 path, bit for bit,
    and stay within 1e-4 of the FP64 oracle."""
    probs = _mixed_set()
    m = capi.Multi(chunk=8)
    ids = [m.add(W.prepare_instance(p), p.nsteps) for p in probs]
    total = 0
    while m.run(7):
        total += 7
    for i, p in zip(ids, probs):
        done, n = m.info(i)
        assert done == n == p.nsteps
        e = m.engine(i)
        s, u = e.seismograms(), e.get_u()
        _, u1, s1 = W.run_engine(p, "fp32c")
        np.testing.assert_array_equal(s, s1)
        np.testing.assert_array_equal(u, u1)
        _, sref = oracle_run(p)
        say(f"  instance p={p.N} ne={p.ne} steps={p.nsteps}: vs oracle {rel_l2(s, sref):.3e}")
        assert rel_l2(s, sref) <= 1e-4
    m.close()


def test_multi_graph_and_plain_launches_agree_and_autotune():
    probs = _mixed_set()
    res = []
    for chunk in (0, 16):
        m = capi.Multi(chunk=chunk)
        ids = [m.add(W.prepare_instance(p), p.nsteps) for p in probs]
        while m.run(16):
            pass
        res.append([m.engine(i).seismograms() for i in ids])
        m.close()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)
    # autotuned mapping: results within FP32 roundoff of the default mapping
    m = capi.Multi(chunk=8)
    ids = [m.add(W.prepare_instance(p), p.nsteps) for p in probs]
    m.autotune(3)
    say("  autotuned variants: " + ", ".join(f"p{N}:{m.variant(N)}" for N in range(1, 9)) +
        f"; launch mode {m.launch()}")
    assert m.launch() in (capi.Multi.PER_DEGREE, capi.Multi.FUSED)
    while m.run(9):
        pass
    for i, p, s0 in zip(ids, probs, res[0]):
        s = m.engine(i).seismograms()
        assert rel_l2(s, s0) <= 1e-5
    m.close()


def test_multi_every_launch_variant():
    """Every launch variant of every degree (persistent or one item per wave,
    whole-wave rows for d <= 32, one element per work-group with a slot per
    wave, or with a half slot per wave for d > 32): the default (one item per
    wave), the half-slot variants and, at p = 5 and 6, the slot variants sum
    each row in the same chunk order and are bitwise the single-context FP32
    run; every other variant agrees with it to FP32 rounding."""
    probs = _mixed_set()
    degrees = sorted({p.N for p in probs})
    ref = [W.run_engine(p, "fp32c")[2] for p in probs]
    for var in range(16):
        m = capi.Multi(chunk=8)
        ids = [m.add(W.prepare_instance(p), p.nsteps) for p in probs]
        used = {}
        for N in degrees:
            wide = (N + 1) * (N + 2) > 32
            v = var if (wide or not (var & m.CHUNKS)) else var & ~m.CHUNKS
            if wide and (v & m.ROWS) and not (v & (m.SLOTS | m.CHUNKS)):
                v &= ~m.ROWS
            m.set_variant(N, v)
            assert m.variant(N) == v
            used[N] = v
        while m.run(9):
            pass
        for i, p, s1 in zip(ids, probs, ref):
            s = m.engine(i).seismograms()
            v = used[p.N]
            if v == m.ONE_SHOT or (v & m.CHUNKS) or ((v & m.SLOTS) and p.N in (5, 6)):
                np.testing.assert_array_equal(s, s1, err_msg=f"variant {v}, p = {p.N}")
            else:
                assert rel_l2(s, s1) <= 1e-5, (v, p.N)
        m.close()
    m = capi.Multi(chunk=8)
    with pytest.raises(capi.EngineError):
        m.set_variant(6, m.ROWS)          # d = 56: no whole-wave rows
    with pytest.raises(capi.EngineError):
        m.set_variant(3, 9)
    with pytest.raises(capi.EngineError):
        m.set_variant(4, m.CHUNKS)        # d = 30: no half-slot chunks
    with pytest.raises(capi.EngineError):
        m.set_variant(6, 16)
    m.close()


def test_multi_launch_modes_are_bitwise_equal():
    """The fused kernel (every degree in one launch) runs each element through
    the same step as the per-degree kernels: the eight
    mixed instances end bit for bit equal in every launch mode, and switching
    mode mid-run changes nothing either."""
    probs = _mixed_set()
    res = {}
    for mode in (capi.Multi.PER_DEGREE, capi.Multi.FUSED, "switch"):
        m = capi.Multi(chunk=8)
        ids = [m.add(W.prepare_instance(p), p.nsteps) for p in probs]
        if mode == "switch":
            k = 0
            while m.run(5):
                k += 1
                m.set_launch(k % 2)
        else:
            m.set_launch(mode)
            assert m.launch() == mode
            while m.run(11):
                pass
        res[mode] = [(m.engine(i).seismograms(), m.engine(i).get_u()) for i in ids]
        m.close()
    for mode in (capi.Multi.FUSED, "switch"):
        for (s0, u0), (s1, u1) in zip(res[capi.Multi.PER_DEGREE], res[mode]):
            np.testing.assert_array_equal(s0, s1)
            np.testing.assert_array_equal(u0, u1)
    with pytest.raises(capi.EngineError):
        capi.Multi(chunk=8).set_launch(2)     # the persistent fused form was removed


def test_multi_shared_operator_and_remove():
    """Three simulations on one mesh share one assembled operator; removing one
    mid-run leaves the others unchanged."""
    base = small_problem(4, nx=6, ny=5, nsteps=60, seed=21, nrec=6)
    probs = []
    for k, (x, y) in enumerate([(150.0, 700.0), (600.0, 300.0), (820.0, 880.0)]):
        q = small_problem(4, nx=6, ny=5, nsteps=60, seed=21, nrec=6)
        q.src_xy = np.array([[x, y]])
        q.dirs = np.array([[np.cos(k), np.sin(k)]])
        q.__post_init__()
        probs.append(q)
    owner = W.prepare_instance(probs[0])
    others = [W.prepare_instance(q, share=owner) for q in probs[1:]]
    m = capi.Multi(chunk=4)
    ids = [m.add(owner, 60)] + [m.add(e, 60) for e in others]
    m.run(20)
    with pytest.raises(capi.EngineError):
        m.remove(ids[0])          # its operator is still borrowed
    m.remove(ids[2])
    while m.run(10):
        pass
    for i, q in zip(ids[:2], probs[:2]):
        _, _, s1 = W.run_engine(q, "fp32c")
        np.testing.assert_array_equal(m.engine(i).seismograms(), s1)
    m.close()


def _columns_problem(N, nsrc, nsteps, seed, nx=5, ny=4):
    """A small mesh with nsrc Ricker sources (one per column of a multi-column context)."""
    p = small_problem(N, nx=nx, ny=ny, nsteps=nsteps, seed=seed, nrec=5)
    rng = np.random.default_rng(seed)
    p.src_xy = np.stack([rng.uniform(100.0, 900.0, nsrc), rng.uniform(100.0, 900.0, nsrc)], 1)
    p.f0 = np.repeat(p.f0, nsrc)
    p.t0 = np.repeat(p.t0, nsrc)
    ang = rng.uniform(0.0, 2 * np.pi, nsrc)
    p.dirs = np.stack([np.cos(ang), np.sin(ang)], 1)
    p.__post_init__()
    return p


def test_multi_set_with_multi_column_instances():
    """Multi-column contexts (8 to 64 columns, p = 1..8, two of the same degree
    and width) in one set beside single-column ones: each (element, 32-column
    tile) runs its context's own block-SpMM step, so every instance ends bitwise
    equal to its context stepped alone (seismograms and every column's
    displacement), with graph chunks or plain launches and in either launch
    mode of the single-column instances; a multi-column owner cannot leave the
    set before its borrower."""
    cases = [(3, 3, 8, 41), (4, 9, 16, 37), (5, 20, 32, 29), (7, 5, 8, 33), (4, 40, 64, 30), (3, 2, 8, 45),
             (8, 2, 8, 21), (6, 12, 16, 26), (1, 4, 8, 36), (2, 6, 16, 31), (7, 24, 32, 19), (8, 36, 64, 17)]
    probs = [_columns_problem(N, ns, st, seed=60 + k) for k, (N, ns, R, st) in enumerate(cases)]
    for p in probs:
        if p.N == 1:
            p.dt *= 0.5      # 0.2 h / (vp p^2) is above the leap-frog limit at p = 1
    widths = [R for (_, _, R, _) in cases]
    singles = _mixed_set()[:3]
    alone = []
    for p, R in zip(probs, widths):
        eng, s = W.run_batched(p, R)
        alone.append((s, [eng.get_u(c) for c in range(len(p.src_elem))]))
        eng.close()
    alone1 = [W.run_engine(p, "fp32c")[2] for p in singles]
    for chunk, mode in ((8, capi.Multi.PER_DEGREE), (0, capi.Multi.PER_DEGREE), (8, capi.Multi.FUSED)):
        say(f"  chunk {chunk}, launch mode {mode}")
        m = capi.Multi(chunk=chunk)
        m.set_launch(mode)
        ids = [m.add(W.prepare_instance(p, ncols=R), p.nsteps) for p, R in zip(probs, widths)]
        ids1 = [m.add(W.prepare_instance(p), p.nsteps) for p in singles]
        while m.run(7):
            pass
        for i, p, R, (s, us) in zip(ids, probs, widths, alone):
            e = m.engine(i)
            assert e.status() == 0
            done, n = m.info(i)
            assert done == n == p.nsteps
            np.testing.assert_array_equal(e.seismograms(), s, err_msg=f"p={p.N} R={R}")
            for c, u in enumerate(us):
                np.testing.assert_array_equal(e.get_u(c), u, err_msg=f"p={p.N} R={R} column {c}")
        for i, s1 in zip(ids1, alone1):
            np.testing.assert_array_equal(m.engine(i).seismograms(), s1)
        m.close()
    owner = W.prepare_instance(probs[0], ncols=8)
    other = W.prepare_instance(probs[0], share=owner)
    m = capi.Multi(chunk=4)
    a, b = m.add(owner, 20), m.add(other, 20)
    m.run(8)
    with pytest.raises(capi.EngineError):
        m.remove(a)               # its operator is still borrowed
    m.remove(b)
    m.remove(a)
    m.close()


# ------------------------------------------------------------------ request loop
def _req(k, nsteps=120, n=(8, 14), nsrc=(1, 4)):
    return W.serving_request(k, n_range=n, nsrc_range=nsrc, steps_range=(nsteps, nsteps + 40), nrec_range=(6, 12))


_MODES = {"columns": capi.Server.COLUMNS, "instances": capi.Server.INSTANCES, "batched": capi.Server.BATCHED}


@pytest.mark.parametrize("mode", ["columns", "instances", "batched"])
def test_server_result_is_the_same_alone_and_in_the_mix(mode):
    """A request's seismograms are identical whether it runs alone or admitted
    among five others with staggered arrivals, and stay within 1e-4 of the FP64
    path source by source.  Instances mode: equal to the single-context FP32
    path bit for bit; columns and batched modes: equal to the same request run
    as an 8-column context (the width the loop picks for 3 sources), and to FP32
    rounding to it run as a 32-column one."""
    import copy
    md = _MODES[mode]
    target = _req(0, nsrc=(3, 3))
    srv, ids, lat, _ = W.serve([target], rate=0, mode=md)
    alone, _ = srv.result(ids[0])
    srv.close()
    mix = [_req(k) for k in range(1, 6)]
    srv, ids, lat, wall = W.serve(mix[:2] + [target] + mix[2:], rate=40.0, seed=3, chunk=8, mode=md)
    mixed, t = srv.result(ids[2])
    say(f"  {mode}: mix of 6: latencies {np.round(lat, 3).tolist()} s, wall {wall:.2f} s")
    srv.close()
    np.testing.assert_array_equal(mixed, alone)
    if md != capi.Server.INSTANCES:
        _, sb = W.run_batched(target, 8)
        np.testing.assert_array_equal(alone, sb[..., :3, :].transpose(2, 0, 1, 3))
        _, sb = W.run_batched(target, 32)   # one slot per wave at full tiles: FP32 rounding apart
        assert rel_l2(alone, sb[..., :3, :].transpose(2, 0, 1, 3)) <= 1e-5
    for k in range(len(target.src_elem)):
        q = copy.copy(target)
        q.src_xy, q.f0, q.t0, q.dirs = target.src_xy[k:k + 1], target.f0[k:k + 1], target.t0[k:k + 1], \
            target.dirs[k:k + 1]
        q.__post_init__()
        if md == capi.Server.INSTANCES:
            _, _, s32 = W.run_engine(q, "fp32c")
            np.testing.assert_array_equal(alone[k], s32)
        _, _, s64 = W.run_engine(q, "fp64")
        err = rel_l2(alone[k], s64)
        say(f"  source {k}: vs FP64 {err:.3e}")
        assert err <= 1e-4


@pytest.mark.parametrize("mode", ["columns", "instances", "batched"])
def test_server_boundary_adapted_request(mode):
    """A request in the BoundaryAdapted basis (tsp_request.basis = 1): the loop
    sets the basis's tables and inverse mass on its contexts, so the result is
    the request run directly, bit for bit (an 8-column context in columns mode,
    each source in its own FP32 context in instances mode), and within the
    basis's FP32 tolerance of the FP64 oracle."""
    import copy
    md = _MODES[mode]
    target = _req(3, nsteps=80, nsrc=(2, 2))
    target.basis = capi.BOUNDARY_ADAPTED
    target.__post_init__()
    srv, ids, lat, _ = W.serve([target], rate=0, mode=md)
    res, _ = srv.result(ids[0])
    srv.close()
    if md != capi.Server.INSTANCES:
        _, sb = W.run_batched(target, 8)
        np.testing.assert_array_equal(res, sb[..., :2, :].transpose(2, 0, 1, 3))
    for k in range(2):
        q = copy.copy(target)
        q.src_xy, q.f0, q.t0, q.dirs = target.src_xy[k:k + 1], target.f0[k:k + 1], target.t0[k:k + 1], \
            target.dirs[k:k + 1]
        q.__post_init__()
        if md == capi.Server.INSTANCES:
            np.testing.assert_array_equal(res[k], W.run_engine(q, "fp32c")[2])
        _, sref = oracle_run(q)
        err = rel_l2(res[k], sref)
        say(f"  BA request p={target.N} source {k}: vs oracle {err:.2e}")
        assert err <= 1e-3


@pytest.mark.parametrize("mode", ["columns", "instances", "batched"])
def test_server_admission_respects_the_memory_budget(mode):
    """With a budget that holds one request at a time, three requests run one
    after the other (each admitted no earlier than the previous completion), and
    all three results equal the request run with an ample budget; a request
    larger than the budget is refused at submission."""
    from tspeed_amd import capi as C
    md = _MODES[mode]
    _, total = ctypes_mem()
    req = _req(7, nsteps=60, nsrc=(2, 2))
    res, peak = C.Server.footprint(req, mode=md)
    frac = (res + peak + res // 2) / total
    srv, ids, lat, _ = W.serve([req, req, req], rate=0, mem_fraction=frac, chunk=16, mode=md)
    times = sorted((srv.result(i)[1] for i in ids), key=lambda t: t["admitted"])
    for a, b in zip(times, times[1:]):
        assert b["admitted"] >= a["completed"]
    outs = [srv.result(i)[0] for i in ids]
    with pytest.raises(C.EngineError):
        srv.submit(_req(8, nsteps=60, n=(30, 30), nsrc=(4, 4)), srv.now())
    srv.close()
    srv, ids, _, _ = W.serve([req], rate=0, mode=md)
    ref = srv.result(ids[0])[0]
    srv.close()
    for o in outs:
        np.testing.assert_array_equal(o, ref)


def ctypes_mem():
    import ctypes as ct
    f, t = ct.c_longlong(), ct.c_longlong()
    assert capi.lib().tsp_device_memory(0, ct.byref(f), ct.byref(t)) == 0
    return f.value, t.value


# ------------------------------------------------------------------ the reference's Wedge example
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _wedge_initial_fields(dt, x0=1480.0):
    """The Rayleigh-wave initial displacement and velocity of Examples/src/Wedge.cpp:
    wedge_init_param there receives (rho, cf) = (10, 2200) in those argument
    slots, which makes the pulse ~0.1 m wide on elements tens of metres across
    (its projection is numerically zero); the evident rho = 2200, cf = 10 Hz
    gives a ~20 m surface pulse, used here.  The example starts it at
    x0 = 1000 m and needs ~5000 steps to reach its first receiver at 1500 m; a
    short test starts it at 1480 m so that the receivers record it."""
    lam, mu, rho, cf, csurf = 1.76e10, 8.8e9, 2200.0, 10.0, 1865.05
    ymax = 2000.0
    beta = np.pi ** 2 * (cf * 3) ** 2
    c2 = np.sqrt(mu / rho)
    m = mu / (lam + 2 * mu)
    k = cf * 3 / csurf
    q = k * np.sqrt(1 - m * (csurf / c2) ** 2)
    s = k * np.sqrt(1 - (csurf / c2) ** 2)

    def depth_x(y):
        return k * (np.exp(-q * (ymax - y)) - 2 * q * s / (s * s + k * k) * np.exp(-s * (ymax - y)))

    def depth_y(y):
        return q * (np.exp(-q * (ymax - y)) - 2 * k * k / (s * s + k * k) * np.exp(-s * (ymax - y)))

    def u0(x, y):
        a = (0 - (x - x0) / csurf)
        g = (1 - beta * a ** 2) * np.exp(-beta * a ** 2)
        return depth_x(y) * g, depth_y(y) * g

    def v0(x, y):
        a = dt + (x - x0) / csurf
        g = 2 * (beta * a ** 2 - 2) * beta * a * np.exp(-beta * a ** 2)
        return depth_x(y) * g, depth_y(y) * g

    return u0, v0


@pytest.mark.parametrize("precision", ["fp64", "fp64_faces", "fp32c"])
def test_reference_wedge_example(precision):
    """The reference's own Wedge example input (Examples/Meshes/wedge.msh and
    Examples/Receivers/wedge.rcv, kept as fixtures in tests/golden): an
    unstructured 2360-triangle mesh with two materials meeting at an internal
    interface, p = 5, dt = 5e-5, no force, a Rayleigh pulse as initial u and v
    (L2-projected as FESpace::transform does) started next to the first
    receiver, 100 receivers on the surface sampled every second step
    (Wedge.cpp:93-146).  FP64 matches the oracle to 1e-12 in the final
    displacement and every seismogram, in the block-ELL step and in the face
    form (forced on this small operator: an unstructured mesh's face slots);
    FP32c to 1e-4."""
    N, dt, nsteps, stride = 5, 5e-5, 300, 2
    pts, tri, reg, _ = capi.read_msh(os.path.join(GOLDEN, "wedge.msh"))
    nb, ng = capi.neighbors(tri)
    assert sorted(set(reg.tolist())) == [1, 2]
    lam = np.where(reg == 1, 6.25e9, 1.76e10)
    mu = np.where(reg == 1, 3.125e9, 8.8e9)
    rho = np.where(reg == 1, 2000.0, 2200.0)
    verts = pts[tri].reshape(-1, 6)
    rxy = np.loadtxt(os.path.join(GOLDEN, "wedge.rcv")).reshape(-1, 2)
    rel, _, rphi = capi.locate(pts, tri, rxy[:, 0], rxy[:, 1], N=N)
    assert (rel >= 0).all() and rxy.shape[0] == 100
    fu, fv = _wedge_initial_fields(dt)
    u0, v0 = capi.project(N, verts, fu), capi.project(N, verts, fv)
    assert np.abs(u0).max() > 0 and np.abs(v0).max() > 0
    op = O.Operator(N, verts, nb, ng, lam, mu, rho, 2.0)
    assert dt < 2 / np.sqrt(op.lambda_max(300))          # the example's step is stable
    uref, sref = op.run(dt, nsteps, rec_elem=rel, rec_phi=rphi, stride=stride, u0=u0, v0=v0)
    e = capi.Engine("fp64" if precision == "fp64_faces" else precision)
    e.set_face_form("always" if precision == "fp64_faces" else True)
    e.set_tables(N)
    e.set_mesh(verts, nb, ng, lam, mu, rho, 2.0)
    assert e.set_face_form() is (precision == "fp64_faces")
    e.set_receivers(rel, rphi, stride, 1 + nsteps // stride)
    e.set_dt(dt)
    e.set_state(u0, v0)
    e.first_step()
    e.run(nsteps)
    u, s = e.get_u(), e.seismograms()
    assert e.status() == 0
    e.close()
    assert np.abs(sref).max() > 1e-3 * np.abs(uref).max()   # the receivers record the pulse
    tol = 1e-4 if precision == "fp32c" else 1e-12
    eu, es = rel_l2(u, uref), seis_err(s, sref)
    say(f"  wedge {precision}: u {eu:.2e}, seismograms {es[0]:.2e} (worst receiver {es[1]:.2e})")
    assert s.shape == sref.shape
    assert eu <= tol and es[0] <= tol and es[1] <= tol


@pytest.mark.parametrize("precision", ["fp64", "fp64_faces", "fp32c"])
def test_reference_lamb_example(precision):
    """The reference's own Lamb example input (Examples/Meshes/
    Lamb_fullyunstruct3.msh and Examples/Receivers/lamb.rcv, fixtures in
    tests/golden) with its set-up (Lamb.cpp): the BoundaryAdapted basis at p = 5
    (FESpace<N, Dunavant<2N>, BoundaryAdapted<N>>), one material, dt = 5e-5 and
    a vertical Ricker force (f0 = 10 Hz, t0 = 0.1 s, sign as Lamb.cpp:37) at
    (1500, 1950).  Its two receivers are 700-1200 m away, beyond a short run,
    so a third receiver right above the source is added; 1500 steps, samples
    every second step.  FP64 (inverse mass applied after the block product, as
    the reference does) matches the oracle's BoundaryAdapted run to 1e-12, in
    the block-ELL step and in the face form (forced on this small operator);
    FP32c, which folds the inverse mass into the operator, to 1e-3 (the basis's
    mass matrix is ill-conditioned in FP32)."""
    N, dt, nsteps, stride = 5, 5e-5, 1500, 2
    BA = capi.BOUNDARY_ADAPTED
    pts, tri, reg, _ = capi.read_msh(os.path.join(GOLDEN, "lamb_fullyunstruct3.msh"))
    assert set(reg.tolist()) == {6}
    nb, ng = capi.neighbors(tri)
    ne = tri.shape[0]
    lam, mu, rho = np.full(ne, 6.8270e9), np.full(ne, 6.8265125e9), np.full(ne, 2000.0)
    verts = pts[tri].reshape(-1, 6)
    rxy = np.vstack([np.loadtxt(os.path.join(GOLDEN, "lamb.rcv")).reshape(-1, 2), [[1500.0, 2000.0]]])
    rel, rxi, _ = capi.locate(pts, tri, rxy[:, 0], rxy[:, 1])
    assert (rel >= 0).all()
    rphi = capi.basis_eval(N, BA, rxi)
    sel, sxi, _ = capi.locate(pts, tri, 1500.0, 1950.0)
    sphi = capi.basis_eval(N, BA, sxi)[0]
    f0, t0 = 10.0, 0.1
    op = O.Operator(N, verts, nb, ng, lam, mu, rho, 2.0, basis=O.BOUNDARY_ADAPTED)
    assert dt < 2 / np.sqrt(op.lambda_max(300))
    uref, sref = op.run(dt, nsteps, src=dict(elem=int(sel[0]), phi=sphi, f0=f0, t0=t0, dirx=0.0, diry=-1.0),
                        rec_elem=rel, rec_phi=rphi, stride=stride)
    assert np.abs(sref[:, 2]).max() > 1e-3 * np.abs(uref).max()   # the added receiver records the source
    e = capi.Engine("fp64" if precision == "fp64_faces" else precision)
    e.set_face_form("always" if precision == "fp64_faces" else True)
    e.set_tables(N, capi.tables(N, BA))
    e.set_mass(capi.base_invmass(N, BA))
    e.set_mesh(verts, nb, ng, lam, mu, rho, 2.0)
    assert e.set_face_form() is (precision == "fp64_faces")
    wav = capi.ricker_table(f0, t0, 0.0, -1.0, dt, nsteps + 1)
    e.set_sources(sel, sphi[None, :], wav)
    e.set_receivers(rel, rphi, stride, 1 + nsteps // stride)
    e.set_dt(dt)
    e.set_state()
    e.first_step()
    e.run(nsteps)
    u, s = e.get_u(), e.seismograms()
    assert e.status() == 0
    e.close()
    tol = 1e-3 if precision == "fp32c" else 1e-12
    eu, es = rel_l2(u, uref), seis_err(s, sref)
    say(f"  lamb {precision}: u {eu:.2e}, seismograms {es[0]:.2e} (worst receiver {es[1]:.2e})")
    assert s.shape == sref.shape
    assert eu <= tol and es[0] <= tol and es[1] <= tol

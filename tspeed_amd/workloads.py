"""Synthetic workloads: the split-square elastic problems the engine is measured
and checked on.  Everything here is input generation (meshes, materials,
sources, receivers, step size); the numerics live in libtspeed_amd.so.

Conventions (fixed and documented so runs are reproducible):
  - meshes come from ``split_square`` (host C++ layer): squares cut along the
    (i,j)-(i+1,j+1) diagonal, interior vertices jittered by U(-0.2h, 0.2h);
  - step size dt = cfl * h_min / (vp_max * p^2) with cfl = 0.2, h_min the
    shortest edge; this sits below the leap-frog limit 2/sqrt(lambda_max(M^-1 B)),
    which measures at 0.26-0.38 h_min/(vp p^2) for penalty 10 on these meshes;
  - penalty 10 (the SIPG penalty 10 p^2/h);
  - Ricker sources F = dir * ricker(f0, t0, t), t0 = 1.2/f0; the "vertical
    point force" has dir = (0, -1) as in Examples/src/Lamb.cpp:37.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import time

import numpy as np

from . import capi

CFL = 0.2
PENALTY = 10.0


@dataclass
class Problem:
    name: str
    N: int
    pts: np.ndarray
    tri: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    rho: np.ndarray
    penalty: float
    dt: float
    nsteps: int
    src_xy: np.ndarray          # (nsrc, 2)
    f0: np.ndarray
    t0: np.ndarray
    dirs: np.ndarray            # (nsrc, 2)
    rec_xy: np.ndarray          # (nrec, 2)
    stride: int = 1
    meta: dict = field(default_factory=dict)
    basis: int = capi.DUBINER   # or capi.BOUNDARY_ADAPTED

    def __post_init__(self):
        self.verts = self.pts[self.tri].reshape(-1, 6)
        self.neigh, self.neighedge = capi.neighbors(self.tri)
        self.src_elem, sxi, self.src_phi = capi.locate(self.pts, self.tri, self.src_xy[:, 0], self.src_xy[:, 1],
                                                       self.N)
        self.rec_elem, rxi, self.rec_phi = capi.locate(self.pts, self.tri, self.rec_xy[:, 0], self.rec_xy[:, 1],
                                                       self.N)
        if self.basis != capi.DUBINER:
            self.src_phi = capi.basis_eval(self.N, self.basis, sxi)
            self.rec_phi = capi.basis_eval(self.N, self.basis, rxi)
        if (self.src_elem < 0).any() or (self.rec_elem < 0).any():
            raise ValueError(f"{self.name}: a source or receiver lies outside the mesh")

    @property
    def ne(self):
        return self.tri.shape[0]

    @property
    def nln(self):
        return capi.nln(self.N)

    @property
    def ndof(self):
        """DOF per right-hand-side column: ne * 2 * nln."""
        return self.ne * 2 * self.nln

    @property
    def nsamples(self):
        return 1 + self.nsteps // self.stride

    def wavelet(self, nrows=None, src=None):
        """Force table rows 0..nrows-1 (row 0 = t=0, row n = t_{n+1})."""
        nrows = self.nsteps + 1 if nrows is None else nrows
        idx = slice(None) if src is None else [src]
        return capi.ricker_table(self.f0[idx], self.t0[idx], self.dirs[idx, 0], self.dirs[idx, 1], self.dt, nrows)

    def source_dict(self, k=0):
        """Source k in the oracle's form."""
        return dict(elem=int(self.src_elem[k]), phi=self.src_phi[k], f0=float(self.f0[k]), t0=float(self.t0[k]),
                    dirx=float(self.dirs[k, 0]), diry=float(self.dirs[k, 1]))

    def min_edge(self):
        v = self.verts.reshape(-1, 3, 2)
        return float(min(np.hypot(*(v[:, a] - v[:, b]).T).min() for a, b in ((0, 1), (1, 2), (2, 0))))


def cfl_dt(pts, tri, N, vp_max, cfl=CFL):
    v = pts[tri]
    hmin = min(np.hypot(*(v[:, a] - v[:, b]).T).min() for a, b in ((0, 1), (1, 2), (2, 0)))
    return cfl * hmin / (vp_max * N * N)


def lame(rho, vp, vs):
    mu = rho * vs * vs
    return rho * vp * vp - 2 * mu, mu


def square_ricker(nx=40, ny=40, N=4, nsteps=2000, nrec=32, seed=1, stride=1):
    """Plain method: 1000 m x 1000 m, 40x40 split squares (3200 triangles), p=4,
    homogeneous rho=2000, vp=3000, vs=1732, vertical Ricker f0=10 Hz at
    (500, 950), 32 receivers along y=990, 2000 steps.  nx/ny/nsteps shrink it
    for tests (the domain stays 1000 m x 1000 m)."""
    L = 1000.0
    pts, tri = capi.split_square(nx, ny, 0.0, 0.0, L, L, 0.2, seed)
    ne = tri.shape[0]
    lam, mu = lame(2000.0, 3000.0, 1732.0)
    rx = (np.arange(nrec) + 0.5) * L / nrec
    return Problem("square_ricker", N, pts, tri, np.full(ne, lam), np.full(ne, mu), np.full(ne, 2000.0), PENALTY,
                   cfl_dt(pts, tri, N, 3000.0), nsteps, np.array([[500.0, 950.0]]), np.array([10.0]),
                   np.array([1.2 / 10.0]), np.array([[0.0, -1.0]]), np.stack([rx, np.full(nrec, 990.0)], 1),
                   stride, dict(nx=nx, ny=ny, seed=seed))


def layered_ricker(nx=400, ny=200, N=6, nsteps=5000, nrec=128, seed=2, stride=1):
    """Larger single run: 4000 m x 2000 m, 400x200 split squares (160k triangles),
    p=6; three horizontal layers (equal thirds, top to bottom vs = 800, 1500,
    2500 m/s, vp = sqrt(3) vs, rho = 1800, 2100, 2400) with every element's
    lambda and mu scaled by one lognormal factor exp(0.1 Z); one vertical Ricker
    f0 = 8 Hz at 10% depth (2000, 1800); 128 receivers half an element below the
    surface; 5000 steps."""
    Lx, Ly = 4000.0, 2000.0
    pts, tri = capi.split_square(nx, ny, 0.0, 0.0, Lx, Ly, 0.2, seed)
    ne = tri.shape[0]
    yc = pts[tri][:, :, 1].mean(1)
    layer = np.where(yc > 2 * Ly / 3, 0, np.where(yc > Ly / 3, 1, 2))
    vs = np.array([800.0, 1500.0, 2500.0])[layer]
    rho = np.array([1800.0, 2100.0, 2400.0])[layer]
    lam, mu = lame(rho, np.sqrt(3.0) * vs, vs)
    fac = np.exp(0.1 * np.random.default_rng(seed).standard_normal(ne))
    lam, mu = lam * fac, mu * fac
    vp_max = float(np.sqrt((lam + 2 * mu) / rho).max())
    rx = (np.arange(nrec) + 0.5) * Lx / nrec
    return Problem("layered_ricker", N, pts, tri, lam, mu, rho, PENALTY, cfl_dt(pts, tri, N, vp_max), nsteps,
                   np.array([[Lx / 2, 0.9 * Ly]]), np.array([8.0]), np.array([1.2 / 8.0]), np.array([[0.0, -1.0]]),
                   np.stack([rx, np.full(nrec, Ly - 0.5 * Ly / ny)], 1), stride, dict(nx=nx, ny=ny, seed=seed))


def multisource_positions(nsrc=64, seed=3, L=1000.0):
    """Sources of the multi-source workload: positions uniform in the top 10% of
    the domain, f0 ~ U(5, 15) Hz, direction uniform on the unit circle."""
    rng = np.random.default_rng(seed)
    xy = np.stack([rng.uniform(0.01 * L, 0.99 * L, nsrc), rng.uniform(0.9 * L, 0.99 * L, nsrc)], 1)
    f0 = rng.uniform(5.0, 15.0, nsrc)
    ang = rng.uniform(0.0, 2 * np.pi, nsrc)
    return xy, f0, 1.2 / f0, np.stack([np.cos(ang), np.sin(ang)], 1)


def multisource_column(k, nx=40, ny=40, N=4, nsteps=3000, nrec=128, seed=3, stride=1):
    """Column k of the 64-source workload as a single-source problem."""
    base = square_ricker(nx, ny, N, nsteps, nrec=nrec, stride=stride)
    xy, f0, t0, dirs = multisource_positions(64, seed)
    base.name = f"multisource[{k}]"
    base.src_xy, base.f0, base.t0, base.dirs = xy[k:k + 1], f0[k:k + 1], t0[k:k + 1], dirs[k:k + 1]
    base.__post_init__()
    return base


def multisource(nsrc=64, nx=40, ny=40, N=4, nsteps=3000, nrec=128, seed=3, stride=1):
    """Multi-source batching: the square_ricker mesh and materials at p=4 with `nsrc`
    Ricker sources (positions uniform in the top 10% of the domain, f0 ~ U(5,15),
    random directions), all advanced together as nsrc right-hand-side columns;
    the same 128 receivers along y = 990 record every column; 3000 steps."""
    base = square_ricker(nx, ny, N, nsteps, nrec=nrec, stride=stride)
    xy, f0, t0, dirs = multisource_positions(nsrc, seed)
    base.name = f"multisource x{nsrc}"
    base.src_xy, base.f0, base.t0, base.dirs = xy, f0, t0, dirs
    base.__post_init__()
    return base


def run_batched(prob: Problem, ncols=64, stride=None, precision="fp32c"):
    """Run all sources of `prob` as ncols columns (source k -> column k) on the
    MFMA path (FP32 or BF16x3 products); returns (engine, seismograms
    [sample, rec, col, 2])."""
    assert len(prob.src_elem) <= ncols
    eng = capi.Engine(precision, ncols=ncols)
    set_space(eng, prob)
    eng.set_mesh(prob.verts, prob.neigh, prob.neighedge, prob.lam, prob.mu, prob.rho, prob.penalty)
    stride = prob.stride if stride is None else stride
    eng.set_sources(prob.src_elem, prob.src_phi, prob.wavelet(), col=np.arange(len(prob.src_elem)))
    eng.set_receivers(prob.rec_elem, prob.rec_phi, stride, 1 + prob.nsteps // max(stride, 1))
    eng.set_dt(prob.dt)
    for c in range(ncols):
        eng.set_state(None, None, col=c)
    eng.first_step()
    eng.run(prob.nsteps)
    return eng, eng.seismograms()


def uneven_instance(k, seed=4, ne_range=(500, 20000), p_range=(2, 8), steps_range=(500, 5000), nrec=8,
                     h=10.0):
    """Instance k of the uneven batch: ne log-uniform in ne_range (split squares of
    side h with aspect ratio U(1,4)), p uniform in p_range, homogeneous
    vs ~ U(500,3000), vp/vs ~ U(1.6,2.0), rho ~ U(1600,2800), one Ricker source
    (f0 ~ U(5,20), random direction) inside the domain, `nrec` receivers along
    the top, steps uniform in steps_range."""
    rng = np.random.default_rng([seed, k])
    ne = float(np.exp(rng.uniform(np.log(ne_range[0]), np.log(ne_range[1]))))
    ar = rng.uniform(1.0, 4.0)
    ny = max(2, int(round(np.sqrt(ne / (2 * ar)))))
    nx = max(2, int(round(ne / (2 * ny))))
    N = int(rng.integers(p_range[0], p_range[1] + 1))
    vs = rng.uniform(500.0, 3000.0)
    vp = vs * rng.uniform(1.6, 2.0)
    rho = rng.uniform(1600.0, 2800.0)
    f0 = rng.uniform(5.0, 20.0)
    ang = rng.uniform(0, 2 * np.pi)
    nsteps = int(rng.integers(steps_range[0], steps_range[1] + 1))
    Lx, Ly = nx * h, ny * h
    pts, tri = capi.split_square(nx, ny, 0.0, 0.0, Lx, Ly, 0.2, 1000 + k)
    n = tri.shape[0]
    lam, mu = lame(rho, vp, vs)
    src = np.array([[rng.uniform(0.2, 0.8) * Lx, rng.uniform(0.2, 0.8) * Ly]])
    rx = (np.arange(nrec) + 0.5) * Lx / nrec
    return Problem(f"uneven[{k}]", N, pts, tri, np.full(n, lam), np.full(n, mu), np.full(n, rho), PENALTY,
                   cfl_dt(pts, tri, N, vp), nsteps, src, np.array([f0]), np.array([1.2 / f0]),
                   np.array([[np.cos(ang), np.sin(ang)]]), np.stack([rx, np.full(nrec, Ly - 0.5 * h)], 1), 1,
                   dict(nx=nx, ny=ny, seed=seed))


def serving_request(k, seed=5, n_range=(20, 40), degrees=(3, 4, 5), nsrc_range=(1, 16), steps_range=(500, 5000),
                    nrec_range=(16, 128)):
    """Request k of the serving workload, drawn like the multi-source workload
    (1000 m x 1000 m split squares, sources in the top 10%, receivers along
    y = 990) with per-request mesh size n x n (n uniform in n_range), degree,
    homogeneous materials (vs ~ U(500,3000), vp/vs ~ U(1.6,2.0),
    rho ~ U(1600,2800)), 1..16 Ricker sources (f0 ~ U(5,15), random directions),
    step count and 16..128 receivers."""
    rng = np.random.default_rng([seed, k])
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    N = int(rng.choice(degrees))
    nsrc = int(rng.integers(nsrc_range[0], nsrc_range[1] + 1))
    nsteps = int(rng.integers(steps_range[0], steps_range[1] + 1))
    nrec = int(rng.integers(nrec_range[0], nrec_range[1] + 1))
    vs = rng.uniform(500.0, 3000.0)
    vp = vs * rng.uniform(1.6, 2.0)
    rho = rng.uniform(1600.0, 2800.0)
    L = 1000.0
    pts, tri = capi.split_square(n, n, 0.0, 0.0, L, L, 0.2, 5000 + k)
    ne = tri.shape[0]
    lam, mu = lame(rho, vp, vs)
    xy = np.stack([rng.uniform(0.01 * L, 0.99 * L, nsrc), rng.uniform(0.9 * L, 0.99 * L, nsrc)], 1)
    f0 = rng.uniform(5.0, 15.0, nsrc)
    ang = rng.uniform(0.0, 2 * np.pi, nsrc)
    rx = (np.arange(nrec) + 0.5) * L / nrec
    return Problem(f"request[{k}]", N, pts, tri, np.full(ne, lam), np.full(ne, mu), np.full(ne, rho), PENALTY,
                   cfl_dt(pts, tri, N, vp), nsteps, xy, f0, 1.2 / f0, np.stack([np.cos(ang), np.sin(ang)], 1),
                   np.stack([rx, np.full(nrec, 990.0)], 1), 1, dict(n=n, seed=seed))


def serve(requests, rate, seed=0, mem_fraction=0.8, chunk=32, stride=1, server=None, mode=None):
    """Submit `requests` (Problems) with Poisson arrivals at `rate` per second and
    run the request loop until all complete.  Returns (server, ids, latencies [s],
    wall seconds)."""
    srv = server or capi.Server(mem_fraction=mem_fraction, chunk=chunk,
                                mode=capi.Server.COLUMNS if mode is None else mode)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(requests)) if rate > 0 else np.zeros(len(requests))
    t0 = srv.now() + 0.05
    ids = [srv.submit(p, t0 + float(a), stride=stride) for p, a in zip(requests, np.cumsum(gaps))]
    pending = len(ids)
    while pending:
        c, r, q = srv.tick()
        pending -= c
        if r == 0 and q == 0 and pending:
            raise RuntimeError("request loop idle with requests outstanding")
        if r == 0 and q > 0 and c == 0:
            time.sleep(0.001)          # nothing running: wait for the next arrival
    lat = []
    for i in ids:
        _, t = srv.result(i)
        lat.append(t["completed"] - t["arrival"])
    return srv, ids, np.array(lat), srv.now() - t0


def prepare_instance(prob: Problem, share=None, stride=None, nsteps=None, ncols=1):
    """An FP32-compensated Engine for `prob` (operator assembled, or borrowed from
    `share`), sources/receivers set, first step taken: ready to be added to a
    Multi set.  ncols > 1: a multi-column context, source k in column k."""
    nsteps = prob.nsteps if nsteps is None else nsteps
    stride = prob.stride if stride is None else stride
    if share is None:
        eng = capi.Engine("fp32c", ncols=ncols)
        set_space(eng, prob)
        eng.set_mesh(prob.verts, prob.neigh, prob.neighedge, prob.lam, prob.mu, prob.rho, prob.penalty)
        eng.release_fp64()
    else:
        eng = capi.Engine(share=share)
    cols = np.arange(len(prob.src_elem)) if ncols > 1 else None
    eng.set_sources(prob.src_elem, prob.src_phi, prob.wavelet(nsteps + 1), col=cols)
    eng.set_receivers(prob.rec_elem, prob.rec_phi, stride, 1 + nsteps // max(stride, 1))
    eng.set_dt(prob.dt)
    for c in range(ncols):
        eng.set_state(None, None, col=c)
    eng.first_step()
    return eng


def set_space(eng, prob: Problem):
    """Tables of prob's basis and, for a non-orthonormal one, its inverse mass."""
    eng.set_tables(prob.N, capi.tables(prob.N, prob.basis))
    if prob.basis != capi.DUBINER:
        eng.set_mass(capi.base_invmass(prob.N, prob.basis))


def run_engine(prob: Problem, precision="fp64", engine=None, u0=None, v0=None, stride=None):
    """Assemble and run `prob` on the device; returns (engine, u_final, seismograms)."""
    eng = engine or capi.Engine(precision)
    set_space(eng, prob)
    eng.set_mesh(prob.verts, prob.neigh, prob.neighedge, prob.lam, prob.mu, prob.rho, prob.penalty)
    stride = prob.stride if stride is None else stride
    eng.set_sources(prob.src_elem, prob.src_phi, prob.wavelet())
    eng.set_receivers(prob.rec_elem, prob.rec_phi, stride, 1 + prob.nsteps // max(stride, 1))
    eng.set_dt(prob.dt)
    eng.set_state(u0, v0)
    eng.first_step()
    eng.run(prob.nsteps)
    return eng, eng.get_u(), eng.seismograms()

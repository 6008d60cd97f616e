"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of bench.py may import this module; the product package
(tspeed_amd) never does.  The C source it wraps restates tspeed's leap-frog
path (see oracle/oracle.c for the reference file:line of each function).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_OMP_PATH = os.path.join(_HERE, "liboracle_omp.so")
_libs = {}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Source(ctypes.Structure):
    _fields_ = [("elem", ctypes.c_int), ("phi", ctypes.c_void_p), ("f0", ctypes.c_double),
                ("t0", ctypes.c_double), ("dirx", ctypes.c_double), ("diry", ctypes.c_double)]


def build():
    """Compile liboracle.so with the reference's Release flags (and
    liboracle_omp.so, the same source with -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(omp=False):
    """The single-threaded oracle, or with omp=True its OpenMP build (bitwise
    equal results; only for timing a multi-core CPU baseline)."""
    if omp in _libs:
        return _libs[omp]
    path = _OMP_PATH if omp else _LIB_PATH
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_gauleg.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.orc_gauss_rule.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_edge_nodes.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
    L.orc_jacobi.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.orc_jacobi.restype = ctypes.c_double
    L.orc_dubiner.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.orc_tables.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_tables_basis.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_boundary_adapted.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.orc_base_invmass.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_assemble_basis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _dp, _dp, _dp,
                                     ctypes.c_double]
    L.orc_assemble_basis.restype = ctypes.c_void_p
    L.orc_neighbors.argtypes = [ctypes.c_int, _ip, _ip, _ip]
    L.orc_geometry.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_locate.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp]
    L.orc_assemble.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _dp, _dp, _dp, ctypes.c_double]
    L.orc_assemble.restype = ctypes.c_void_p
    L.orc_free.argtypes = [ctypes.c_void_p]
    L.orc_export_blocks.argtypes = [ctypes.c_void_p, _dp]
    L.orc_export_invm.argtypes = [ctypes.c_void_p, _dp]
    L.orc_apply_B.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_apply_invM.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Source),
                          ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, _dp, _dp]
    L.orc_ricker.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.orc_ricker.restype = ctypes.c_double
    L.orc_lambda_max.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong]
    L.orc_lambda_max.restype = ctypes.c_double
    L.orc_threads.argtypes = [ctypes.c_int]
    L.orc_threads.restype = ctypes.c_int
    _libs[omp] = L
    return L


def threads(n=0, omp=False):
    """Set (n > 0) and report the thread count of the chosen build."""
    return lib(omp).orc_threads(int(n))


def nln(N):
    return (N + 1) * (N + 2) // 2


def gauleg(n, a, b):
    x = np.zeros(n); w = np.zeros(n)
    assert lib().orc_gauleg(n, a, b, x, w) == 0
    return x, w


def gauss_rule(nq):
    w2 = np.zeros(nq * nq); xy2 = np.zeros(nq * nq * 2); w1 = np.zeros(nq); x1 = np.zeros(nq)
    assert lib().orc_gauss_rule(nq, w2, xy2, w1, x1) == 0
    return w2, xy2.reshape(-1, 2), w1, x1


def edge_nodes(nq, x1, edge):
    xy = np.zeros(nq * 2)
    lib().orc_edge_nodes(nq, np.ascontiguousarray(x1, dtype=np.float64), edge, xy)
    return xy.reshape(-1, 2)


def jacobi(n, a, b, z):
    return lib().orc_jacobi(n, a, b, float(z))


def dubiner(N, xi, eta):
    phi = np.zeros(nln(N)); gr = np.zeros(2 * nln(N))
    lib().orc_dubiner(N, float(xi), float(eta), phi, gr)
    return phi, gr.reshape(-1, 2)


DUBINER, BOUNDARY_ADAPTED = 0, 1


def boundary_adapted(N, xi, eta):
    phi = np.zeros(nln(N)); gr = np.zeros(2 * nln(N))
    lib().orc_boundary_adapted(N, float(xi), float(eta), phi, gr)
    return phi, gr.reshape(-1, 2)


def base_invmass(N, basis=DUBINER):
    out = np.zeros(nln(N) * nln(N))
    assert lib().orc_base_invmass(N, basis, out) == 0
    return out.reshape(nln(N), nln(N))


def tables(N, basis=DUBINER):
    n, nq1 = nln(N), N + 1
    nq2 = nq1 * nq1
    B = np.zeros(nq2 * n); G = np.zeros(n * 2 * nq2); Be = np.zeros(3 * nq1 * n); Ge = np.zeros(n * 3 * 2 * nq1)
    w2 = np.zeros(nq2); xy2 = np.zeros(nq2 * 2); w1 = np.zeros(nq1); x1 = np.zeros(nq1)
    assert lib().orc_tables_basis(N, basis, B, G, Be, Ge, w2, xy2, w1, x1) == 0
    return dict(B=B.reshape(nq2, n), G=G.reshape(n, 2, nq2), Bedge=Be.reshape(3, nq1, n),
                Gedge=Ge.reshape(n, 3, 2, nq1), w2=w2, xy2=xy2.reshape(-1, 2), w1=w1, x1=x1)


def neighbors(tri):
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    ne = tri.shape[0]
    nb = np.zeros(ne * 3, dtype=np.int32); ng = np.zeros(ne * 3, dtype=np.int32)
    lib().orc_neighbors(ne, tri.reshape(-1), nb, ng)
    return nb.reshape(ne, 3), ng.reshape(ne, 3)


def geometry(verts):
    verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 6)
    ne = verts.shape[0]
    J = np.zeros(ne * 4); iJ = np.zeros(ne * 4); dJ = np.zeros(ne); nr = np.zeros(ne * 6); ln = np.zeros(ne * 3)
    lib().orc_geometry(ne, verts.reshape(-1), J, iJ, dJ, nr, ln)
    return dict(J=J.reshape(ne, 2, 2), invJ=iJ.reshape(ne, 2, 2), detJ=dJ, normal=nr.reshape(ne, 3, 2),
                length=ln.reshape(ne, 3))


def locate(verts, px, py):
    verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    xi = np.zeros(2)
    e = lib().orc_locate(verts.size // 6, verts, float(px), float(py), xi)
    return e, xi


def ricker(f0, t0, t):
    return lib().orc_ricker(f0, t0, t)


class Operator:
    """B = penalty*N^2*S + I + A assembled as the reference does (Matrices_imp.hpp:192-215)."""

    def __init__(self, N, verts, neigh, neighedge, lam, mu, rho, penalty, basis=DUBINER, omp=False):
        self.N, self.nln, self.basis = N, nln(N), basis
        self._L = lib(omp)
        self.verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
        self.ne = self.verts.size // 6
        self._h = self._L.orc_assemble_basis(N, basis, self.ne, self.verts,
                                     np.ascontiguousarray(neigh, dtype=np.int32).reshape(-1),
                                     np.ascontiguousarray(neighedge, dtype=np.int32).reshape(-1),
                                     np.ascontiguousarray(lam, dtype=np.float64),
                                     np.ascontiguousarray(mu, dtype=np.float64),
                                     np.ascontiguousarray(rho, dtype=np.float64), float(penalty))

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_free(self._h)
            self._h = None

    @property
    def ndof(self):
        return 2 * self.ne * self.nln

    def blocks(self):
        d = 2 * self.nln
        out = np.zeros(self.ne * 4 * d * d)
        self._L.orc_export_blocks(self._h, out)
        return out.reshape(self.ne, 4, d, d)

    def invm(self):
        out = np.zeros(self.ne)
        self._L.orc_export_invm(self._h, out)
        return out

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._L.orc_apply_B(self._h, x, out)
        return out

    def apply_invM(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._L.orc_apply_invM(self._h, x, out)
        return out

    def lambda_max(self, iters=200, seed=7):
        return self._L.orc_lambda_max(self._h, iters, seed)

    def run(self, dt, nsteps, src=None, rec_elem=None, rec_phi=None, stride=1, u0=None, v0=None):
        """first_step + nsteps steps; src = dict(elem, phi, f0, t0, dirx, diry) or None.
        Returns (u_final component-major, seismograms [nsamples, nrec, 2])."""
        L = self._L
        s = Source(-1, None, 0.0, 0.0, 0.0, 0.0)
        keep = []
        if src is not None:
            phi = np.ascontiguousarray(src["phi"], dtype=np.float64)
            keep.append(phi)
            s = Source(int(src["elem"]), phi.ctypes.data, float(src["f0"]), float(src["t0"]),
                       float(src["dirx"]), float(src["diry"]))
        nrec = 0 if rec_elem is None else len(rec_elem)
        re = np.ascontiguousarray(rec_elem if nrec else np.zeros(1), dtype=np.int32)
        rp = np.ascontiguousarray(rec_phi if nrec else np.zeros(1), dtype=np.float64).reshape(-1)
        nsamp = 1 + (nsteps // stride if stride > 0 else 0)
        seis = np.zeros(max(1, nsamp * max(nrec, 1) * 2))
        u = np.zeros(self.ndof)
        pu0 = None if u0 is None else np.ascontiguousarray(u0, dtype=np.float64)
        pv0 = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        ns = L.orc_run(self._h, float(dt), int(nsteps), ctypes.byref(s), nrec, re, rp, int(stride),
                       None if pu0 is None else pu0.ctypes.data, None if pv0 is None else pv0.ctypes.data,
                       u, seis)
        return u, seis[: ns * nrec * 2].reshape(ns, nrec, 2)

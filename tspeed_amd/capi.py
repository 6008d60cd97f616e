"""ctypes bindings of libtspeed_amd.so: the device engine (include/tspeed_amd.h)
and the host-layer exports (include/tspeed_host.h).

The library is built in-tree by ``tspeed_amd.build``.  Loading fails loudly when
it is missing: there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TSPEED_AMD_LIB", os.path.join(HERE, "libtspeed_amd.so"))

TSP_OK, TSP_ERR_ARG, TSP_ERR_STATE, TSP_ERR_HIP, TSP_ERR_UNSUPPORTED, TSP_ERR_NOMEM = range(6)
TSP_FP64, TSP_FP32C, TSP_BF16X3, TSP_BF16X5 = 0, 1, 2, 3
TSP_FLAG_NONFINITE = 1
PRECISIONS = {"fp64": TSP_FP64, "fp32c": TSP_FP32C, "bf16x3": TSP_BF16X3, "bf16x5": TSP_BF16X5}

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# name -> (restype, argtypes) for every symbol of include/tspeed_amd.h and tspeed_host.h
SIGNATURES = {
    "tsp_device_count": (_i, [ctypes.POINTER(_i)]),
    "tsp_create": (_i, [_i, _i, _i, ctypes.POINTER(_vp)]),
    "tsp_destroy": (None, [_vp]),
    "tsp_last_error": (ctypes.c_char_p, [_vp]),
    "tsp_set_stream": (_i, [_vp, _vp]),
    "tsp_set_tables": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp]),
    "tsp_set_mesh": (_i, [_vp, _i, _dp, _ip, _ip, _dp, _dp, _dp, _d]),
    "tsp_upload_blocks": (_i, [_vp, _i, _ip, _dp, _dp]),
    "tsp_get_blocks": (_i, [_vp, _i, _i, _vp, _vp]),
    "tsp_set_state": (_i, [_vp, _i, _vp, _vp]),
    "tsp_set_sources": (_i, [_vp, _i, _ip, _ip, _dp, _i, _dp]),
    "tsp_set_wavelet": (_i, [_vp, _i, _i, _dp]),
    "tsp_set_receivers": (_i, [_vp, _i, _ip, _dp, _i, _i]),
    "tsp_set_stride": (_i, [_vp, _i]),
    "tsp_set_face_form": (_i, [_vp, _i, ctypes.POINTER(_i)]),
    "tsp_set_dt": (_i, [_vp, _d]),
    "tsp_first_step": (_i, [_vp]),
    "tsp_run": (_i, [_vp, _i]),
    "tsp_run_timed": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_float)]),
    "tsp_eval_receivers": (_i, [_vp]),
    "tsp_steps_done": (_i, [_vp]),
    "tsp_get_u": (_i, [_vp, _i, _dp]),
    "tsp_get_seismograms": (_i, [_vp, _vp, _i, ctypes.POINTER(_i)]),
    "tsp_cfl": (_i, [_vp, _d, _i, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "tsp_cfl_noalloc": (_i, [_vp, _d, _i, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "tsp_status": (_i, [_vp, ctypes.POINTER(ctypes.c_uint)]),
    "tsp_apply_operator": (_i, [_vp, _vp, _vp, _i]),
    "tsp_apply_invmass": (_i, [_vp, _vp, _vp, _i]),
    "tsp_synchronize": (_i, [_vp]),
    "tsp_device_bytes": (ctypes.c_longlong, [_vp]),
    "tsp_share_operator": (_i, [_vp, ctypes.POINTER(_vp)]),
    "tsp_release_fp64": (_i, [_vp]),
    "tsp_multi_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "tsp_multi_destroy": (None, [_vp]),
    "tsp_multi_last_error": (ctypes.c_char_p, [_vp]),
    "tsp_multi_add": (_i, [_vp, _vp, _i, ctypes.POINTER(_i)]),
    "tsp_multi_set_chunk": (_i, [_vp, _i]),
    "tsp_multi_run": (_i, [_vp, _i, ctypes.POINTER(_i)]),
    "tsp_multi_run_timed": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_float)]),
    "tsp_multi_autotune": (_i, [_vp, _i]),
    "tsp_multi_variant": (_i, [_vp, _i]),
    "tsp_multi_set_variant": (_i, [_vp, _i, _i]),
    "tsp_multi_set_launch": (_i, [_vp, _i]),
    "tsp_multi_launch": (_i, [_vp]),
    "tsp_multi_info": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tsp_multi_context": (_vp, [_vp, _i]),
    "tsp_multi_remove": (_i, [_vp, _i]),
    "tsp_multi_synchronize": (_i, [_vp]),
    "tsp_device_memory": (_i, [_i, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tsp_server_create": (_i, [_i, _d, _i, ctypes.POINTER(_vp)]),
    "tsp_server_destroy": (None, [_vp]),
    "tsp_server_last_error": (ctypes.c_char_p, [_vp]),
    "tsp_server_now": (_d, []),
    "tsp_server_footprint": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tsp_server_set_mode": (_i, [_vp, _i]),
    "tsp_get_seismograms_cols": (_i, [_vp, _i, _i, _vp, _i, ctypes.POINTER(_i), _i]),
    "tsp_stream_idle": (_i, [_vp, ctypes.POINTER(_i)]),
    "tsp_host_alloc": (_vp, [ctypes.c_size_t]),
    "tsp_host_free": (None, [_vp]),
    "tsp_server_phases": (_i, [_vp, ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "tsp_server_submit": (_i, [_vp, _vp, _d, ctypes.POINTER(_i)]),
    "tsp_server_tick": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tsp_server_result": (_i, [_vp, _i, _vp, ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_d),
                               ctypes.POINTER(_d)]),
    "tsp_server_stats": (_i, [_vp, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong),
                              ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "tsh_field_points": (_i, [_i, _dp, _dp]),
    "tsh_field_values": (_i, [_i, _i, _i, _dp, _dp]),
    "tsh_quad_points": (_i, [_i, _i, _dp, _dp]),
    "tsh_project": (_i, [_i, _i, _i, _dp, _dp, _dp]),
    "tsh_l2_error": (_i, [_i, _i, _i, _dp, _dp, _dp, ctypes.POINTER(_d)]),
    "tsh_tables_basis": (_i, [_i, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "tsh_base_invmass": (_i, [_i, _i, _dp]),
    "tsh_basis_eval": (_i, [_i, _i, _i, _dp, _dp, _vp]),
    "tsp_set_mass": (_i, [_vp, _vp]),
    "tsh_gauleg": (_i, [_i, _d, _d, _dp, _dp]),
    "tsh_dubiner": (_i, [_i, _d, _d, _dp, _vp]),
    "tsh_tables": (_i, [_i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "tsh_neighbors": (_i, [_i, _ip, _ip, _ip]),
    "tsh_split_square": (_i, [_i, _i, _d, _d, _d, _d, _d, ctypes.c_uint64, _dp, _ip]),
    "tsh_msh_counts": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tsh_msh_read": (_i, [ctypes.c_char_p, _dp, _ip, _ip, _ip]),
    "tsh_locate": (_i, [_i, _dp, _i, _ip, _i, _dp, _dp, _i, _ip, _dp, _dp]),
    "tsh_ricker_table": (_i, [_i, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "tsh_write_receivers": (_i, [ctypes.c_char_p, _i, _i, _dp, _i]),
}


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libtspeed_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tspeed_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def device_count():
    n = _i()
    lib().tsp_device_count(ctypes.byref(n))
    return n.value


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def nln(N):
    return (N + 1) * (N + 2) // 2


# ---------------------------------------------------------------- host layer
def gauleg(n, a=-1.0, b=1.0):
    x, w = np.zeros(n), np.zeros(n)
    assert lib().tsh_gauleg(n, a, b, x, w) == 0
    return x, w


def dubiner(N, xi, eta, grad=True):
    phi = np.zeros(nln(N))
    g = np.zeros(2 * nln(N)) if grad else None
    assert lib().tsh_dubiner(N, float(xi), float(eta), phi, None if g is None else g.ctypes.data) == 0
    return (phi, g.reshape(-1, 2)) if grad else phi


DUBINER, BOUNDARY_ADAPTED = 0, 1


def tables(N, basis=DUBINER):
    n, nq1 = nln(N), N + 1
    nq2 = nq1 * nq1
    B = np.zeros(nq2 * n); G = np.zeros(n * 2 * nq2); Be = np.zeros(3 * nq1 * n); Ge = np.zeros(n * 3 * 2 * nq1)
    w2 = np.zeros(nq2); xy2 = np.zeros(2 * nq2); w1 = np.zeros(nq1); x1 = np.zeros(nq1)
    assert lib().tsh_tables_basis(N, basis, B, G, Be, Ge, w2, xy2, w1, x1) == 0
    return dict(B=B, G=G, Bedge=Be, Gedge=Ge, w2=w2, xy2=xy2.reshape(-1, 2), w1=w1, x1=x1)


def neighbors(tri):
    tri = _c32i(tri).reshape(-1, 3)
    ne = tri.shape[0]
    nb = np.zeros(3 * ne, np.int32); ng = np.zeros(3 * ne, np.int32)
    if lib().tsh_neighbors(ne, tri.reshape(-1), nb, ng) != 0:
        raise ValueError("non-manifold mesh")
    return nb.reshape(ne, 3), ng.reshape(ne, 3)


def split_square(nx, ny, x0, y0, Lx, Ly, jitter, seed):
    pts = np.zeros(2 * (nx + 1) * (ny + 1)); tri = np.zeros(6 * nx * ny, np.int32)
    assert lib().tsh_split_square(nx, ny, x0, y0, Lx, Ly, jitter, seed, pts, tri) == 0
    return pts.reshape(-1, 2), tri.reshape(-1, 3)


def read_msh(path):
    np_, nt, nbd = _i(), _i(), _i()
    if lib().tsh_msh_counts(path.encode(), ctypes.byref(np_), ctypes.byref(nt), ctypes.byref(nbd)) != 0:
        raise FileNotFoundError(path)
    pts = np.zeros(2 * np_.value); tri = np.zeros(3 * nt.value, np.int32)
    reg = np.zeros(max(nt.value, 1), np.int32); bnd = np.zeros(max(2 * nbd.value, 1), np.int32)
    assert lib().tsh_msh_read(path.encode(), pts, tri, reg, bnd) == 0
    return pts.reshape(-1, 2), tri.reshape(-1, 3), reg[:nt.value], bnd[:2 * nbd.value].reshape(-1, 2)


def locate(pts, tri, qx, qy, N=0):
    pts = _c64(pts).reshape(-1); tri = _c32i(tri).reshape(-1)
    qx = _c64(np.atleast_1d(qx)); qy = _c64(np.atleast_1d(qy))
    nq = qx.size
    el = np.zeros(nq, np.int32); xi = np.zeros(2 * nq); phi = np.zeros(max(1, nq * (nln(N) if N > 0 else 1)))
    assert lib().tsh_locate(pts.size // 2, pts, tri.size // 3, tri, nq, qx, qy, N, el, xi, phi) == 0
    return el, xi.reshape(-1, 2), (phi.reshape(nq, nln(N)) if N > 0 else None)


def ricker_table(f0, t0, dirx, diry, dt, nrows):
    f0, t0, dirx, diry = (_c64(np.atleast_1d(v)) for v in (f0, t0, dirx, diry))
    ns = f0.size
    out = np.zeros(max(1, nrows * ns * 2))
    assert lib().tsh_ricker_table(ns, f0, t0, dirx, diry, float(dt), nrows, out) == 0
    return out[: nrows * ns * 2].reshape(nrows, ns, 2)


def field_points(verts):
    """Images of the four plot points in every element, (ne, 4, 2)."""
    verts = _c64(verts).reshape(-1, 6)
    out = np.zeros((max(1, verts.shape[0]), 4, 2))
    assert lib().tsh_field_points(verts.shape[0], verts, out) == 0
    return out[: verts.shape[0]]


def base_invmass(N, basis=DUBINER):
    """base_invmass of the degree-N space (identity for Dubiner)."""
    out = np.zeros(nln(N) * nln(N))
    assert lib().tsh_base_invmass(N, basis, out) == 0
    return out.reshape(nln(N), nln(N))


def basis_eval(N, basis, xi):
    """Basis values at reference points xi (n, 2): (n, nln)."""
    xi = _c64(np.atleast_2d(xi))
    n = xi.shape[0]
    phi = np.zeros(max(1, n * nln(N)))
    assert lib().tsh_basis_eval(N, basis, n, xi.reshape(-1), phi, None) == 0
    return phi[: n * nln(N)].reshape(n, nln(N))


def field_values(N, uh, basis=DUBINER):
    """uh (component-major, degree N) at the four plot points, (ne, 4, 2)."""
    uh = _c64(uh)
    ne = uh.size // (2 * nln(N))
    out = np.zeros((max(1, ne), 4, 2))
    assert lib().tsh_field_values(N, basis, ne, uh, out) == 0
    return out[:ne]


def quad_points(N, verts):
    """Quadrature nodes of the degree-N space in every element, (ne, (N+1)^2, 2)."""
    verts = _c64(verts).reshape(-1, 6)
    ne = verts.shape[0]
    out = np.zeros((max(1, ne), (N + 1) ** 2, 2))
    assert lib().tsh_quad_points(N, ne, verts, out) == 0
    return out[:ne]


def project(N, verts, fun, basis=DUBINER):
    """L2 projection of fun(x, y) -> (fx, fy) (vectorised over arrays) onto the
    degree-N space: component-major coefficients (FESpace::transform)."""
    verts = _c64(verts).reshape(-1, 6)
    ne = verts.shape[0]
    xy = quad_points(N, verts)
    fx, fy = fun(xy[..., 0], xy[..., 1])
    f = _c64(np.stack([np.broadcast_to(fx, xy.shape[:2]), np.broadcast_to(fy, xy.shape[:2])], -1))
    uh = np.zeros(max(1, 2 * ne * nln(N)))
    assert lib().tsh_project(N, basis, ne, verts, f, uh) == 0
    return uh[: 2 * ne * nln(N)]


def l2_error(N, verts, uh, fun, basis=DUBINER):
    """L2 norm of fun - uh over the mesh (FESpace::L2error)."""
    verts = _c64(verts).reshape(-1, 6)
    xy = quad_points(N, verts)
    fx, fy = fun(xy[..., 0], xy[..., 1])
    f = _c64(np.stack([np.broadcast_to(fx, xy.shape[:2]), np.broadcast_to(fy, xy.shape[:2])], -1))
    err = _d()
    assert lib().tsh_l2_error(N, basis, verts.shape[0], verts, _c64(uh), f, ctypes.byref(err)) == 0
    return err.value


# ---------------------------------------------------------------- device engine
class Engine:
    """One tsp_context: a mesh, its operator and the time-stepping state."""

    def __init__(self, precision="fp64", device=0, ncols=1, share=None, handle=None):
        """New context; share=Engine borrows that engine's operator
        (tsp_share_operator); handle wraps an existing context without owning it."""
        self._L = lib()
        self.precision = precision
        self.ncols = ncols
        self.N = None
        self.ne = 0
        self._owned = handle is None
        self._keep = share
        if handle is not None:
            self._h = handle
            return
        h = _vp()
        if share is not None:
            rc = self._L.tsp_share_operator(share._h, ctypes.byref(h))
            self.N, self.ne, self.precision, self.ncols = share.N, share.ne, share.precision, share.ncols
        else:
            rc = self._L.tsp_create(device, PRECISIONS[precision], ncols, ctypes.byref(h))
        if rc != TSP_OK:
            raise EngineError(rc, f"tsp_create(device={device}, precision={precision}) failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and getattr(self, "_owned", True):
            self._L.tsp_destroy(self._h)
        self._h = None

    def release(self):
        """Hand the context over (to a Multi set): it is no longer destroyed here."""
        self._owned = False
        return self._h

    def release_fp64(self):
        self._ck(self._L.tsp_release_fp64(self._h))

    def __del__(self):
        self.close()

    def _ck(self, rc):
        if rc != TSP_OK:
            raise EngineError(rc, self._L.tsp_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        self._ck(self._L.tsp_set_stream(self._h, stream_ptr))

    def set_tables(self, N, tab=None):
        tab = tab or tables(N)
        self._ck(self._L.tsp_set_tables(self._h, N, _c64(tab["B"]), _c64(tab["G"]), _c64(tab["Bedge"]),
                                        _c64(tab["Gedge"]), _c64(tab["w2"]), _c64(tab["w1"])))
        self.N = N

    def set_mass(self, base_invmass):
        """tsp_set_mass: None = orthonormal basis."""
        m = None if base_invmass is None else _c64(base_invmass).reshape(-1)
        self._ck(self._L.tsp_set_mass(self._h, None if m is None else m.ctypes.data))

    def set_mesh(self, verts, neigh, neighedge, lam, mu, rho, penalty):
        verts = _c64(verts).reshape(-1)
        self.ne = verts.size // 6
        self._ck(self._L.tsp_set_mesh(self._h, self.ne, verts, _c32i(neigh).reshape(-1),
                                      _c32i(neighedge).reshape(-1), _c64(lam), _c64(mu), _c64(rho), float(penalty)))

    def upload_blocks(self, neigh, blocks, invm):
        neigh = _c32i(neigh).reshape(-1)
        self.ne = neigh.size // 3
        self._ck(self._L.tsp_upload_blocks(self._h, self.ne, neigh, _c64(blocks).reshape(-1), _c64(invm)))

    def get_blocks(self, e_begin=0, e_count=None):
        e_count = self.ne - e_begin if e_count is None else e_count
        d = 2 * nln(self.N)
        b = np.zeros(max(1, e_count * 4 * d * d)); im = np.zeros(max(1, e_count))
        self._ck(self._L.tsp_get_blocks(self._h, int(e_begin), int(e_count), b.ctypes.data, im.ctypes.data))
        return b[: e_count * 4 * d * d].reshape(e_count, 4, d, d), im[:e_count]

    def set_state(self, u0=None, v0=None, col=0):
        a = None if u0 is None else _c64(u0)
        b = None if v0 is None else _c64(v0)
        self._ck(self._L.tsp_set_state(self._h, col, None if a is None else a.ctypes.data,
                                       None if b is None else b.ctypes.data))

    def set_sources(self, elem, phi, wavelet, col=None):
        elem = _c32i(np.atleast_1d(elem))
        ns = elem.size
        col = _c32i(np.zeros(ns) if col is None else col)
        wav = _c64(wavelet).reshape(-1)
        ntab = wav.size // (2 * ns) if ns else 0
        self._ck(self._L.tsp_set_sources(self._h, ns, col, elem, _c64(phi).reshape(-1), ntab,
                                         wav if ns else np.zeros(1)))

    def set_receivers(self, elem, phi, stride, max_samples):
        elem = _c32i(np.atleast_1d(elem))
        nr = elem.size
        self._ck(self._L.tsp_set_receivers(self._h, nr, elem, _c64(phi).reshape(-1) if nr else np.zeros(1),
                                           int(stride), int(max_samples)))
        self.nrec = nr

    def set_stride(self, stride):
        self._ck(self._L.tsp_set_stride(self._h, int(stride)))

    def set_face_form(self, on=None):
        """FP64 face form (each interior face's block read once for both sides):
        True = for operators above the cache-resident size (the default),
        "always" = wherever possible, False = never, None = only ask; returns
        whether the steps use it."""
        code = -1 if on is None else (2 if on == "always" else int(bool(on)))
        act = _i(0)
        self._ck(self._L.tsp_set_face_form(self._h, code, ctypes.byref(act)))
        return bool(act.value)

    def set_dt(self, dt):
        self._ck(self._L.tsp_set_dt(self._h, float(dt)))

    def first_step(self):
        self._ck(self._L.tsp_first_step(self._h))

    def run(self, nsteps):
        self._ck(self._L.tsp_run(self._h, int(nsteps)))

    def run_timed(self, nsteps):
        """Run nsteps and return the device time in milliseconds (HIP events)."""
        ms = ctypes.c_float()
        self._ck(self._L.tsp_run_timed(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def eval_receivers(self):
        self._ck(self._L.tsp_eval_receivers(self._h))

    def steps_done(self):
        return self._L.tsp_steps_done(self._h)

    def get_u(self, col=0):
        out = np.zeros(2 * self.ne * nln(self.N))
        self._ck(self._L.tsp_get_u(self._h, col, out))
        return out

    def seismograms(self):
        ns = _i()
        self._ck(self._L.tsp_get_seismograms(self._h, None, 0, ctypes.byref(ns)))
        out = np.zeros(max(1, ns.value * self.nrec * self.ncols * 2))
        self._ck(self._L.tsp_get_seismograms(self._h, out.ctypes.data, ns.value, ctypes.byref(ns)))
        s = out[: ns.value * self.nrec * self.ncols * 2].reshape(ns.value, self.nrec, self.ncols, 2)
        return s[:, :, 0, :] if self.ncols == 1 else s

    def cfl(self, factor=1.0, iters=100):
        dt, lm = _d(), _d()
        self._ck(self._L.tsp_cfl(self._h, float(factor), int(iters), ctypes.byref(dt), ctypes.byref(lm)))
        return dt.value, lm.value

    def apply_operator(self, x):
        """y = B x on a component-major host vector (tsp_apply_operator)."""
        x = _c64(x)
        y = np.zeros_like(x)
        self._ck(self._L.tsp_apply_operator(self._h, x.ctypes.data, y.ctypes.data, 0))
        return y

    def apply_invmass(self, x):
        """y = invM x on a component-major host vector (tsp_apply_invmass)."""
        x = _c64(x)
        y = np.zeros_like(x)
        self._ck(self._L.tsp_apply_invmass(self._h, x.ctypes.data, y.ctypes.data, 0))
        return y

    def status(self):
        f = ctypes.c_uint()
        self._ck(self._L.tsp_status(self._h, ctypes.byref(f)))
        return f.value

    def synchronize(self):
        self._ck(self._L.tsp_synchronize(self._h))

    def device_bytes(self):
        return int(self._L.tsp_device_bytes(self._h))


class Multi:
    """A set of independent simulations stepped together (tsp_multi_*)."""

    PER_DEGREE, FUSED = 0, 1

    def __init__(self, device=0, chunk=32):
        self._L = lib()
        h = _vp()
        rc = self._L.tsp_multi_create(device, ctypes.byref(h))
        if rc != TSP_OK:
            raise EngineError(rc, "tsp_multi_create failed")
        self._h = h
        self.engines = {}
        self.set_chunk(chunk)

    def _ck(self, rc):
        if rc != TSP_OK:
            raise EngineError(rc, self._L.tsp_multi_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            for e in self.engines.values():
                e._h = None
            self._L.tsp_multi_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_chunk(self, chunk):
        self._ck(self._L.tsp_multi_set_chunk(self._h, int(chunk)))

    def add(self, engine, nsteps):
        """Hand over an Engine after its first step; returns the instance id."""
        i = _i()
        self._ck(self._L.tsp_multi_add(self._h, engine.handle, int(nsteps), ctypes.byref(i)))
        engine.release()
        self.engines[i.value] = engine
        return i.value

    def run(self, nsteps):
        a = _i()
        self._ck(self._L.tsp_multi_run(self._h, int(nsteps), ctypes.byref(a)))
        return a.value

    def run_timed(self, nsteps):
        ms = ctypes.c_float()
        self._ck(self._L.tsp_multi_run_timed(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def autotune(self, steps_per_trial=5):
        self._ck(self._L.tsp_multi_autotune(self._h, int(steps_per_trial)))

    def variant(self, degree):
        return self._L.tsp_multi_variant(self._h, int(degree))

    ROWS, ONE_SHOT, SLOTS, CHUNKS = 1, 2, 4, 8

    def set_variant(self, degree, variant):
        """Pin a degree's launch variant (bits ROWS | ONE_SHOT | SLOTS | CHUNKS)."""
        self._ck(self._L.tsp_multi_set_variant(self._h, int(degree), int(variant)))

    def set_launch(self, mode):
        """PER_DEGREE (one kernel per degree) or FUSED (one kernel for every
        degree)."""
        self._ck(self._L.tsp_multi_set_launch(self._h, int(mode)))

    def launch(self):
        return self._L.tsp_multi_launch(self._h)

    def info(self, i):
        a, b = _i(), _i()
        self._ck(self._L.tsp_multi_info(self._h, int(i), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def engine(self, i):
        return self.engines[i]

    def remove(self, i):
        self._ck(self._L.tsp_multi_remove(self._h, int(i)))
        e = self.engines.pop(i)
        e._h = None

    def synchronize(self):
        self._ck(self._L.tsp_multi_synchronize(self._h))


class Request(ctypes.Structure):
    """tsp_request (include/tspeed_amd.h)."""
    _fields_ = [("N", _i), ("ne", _i), ("verts", _vp), ("neigh", _vp), ("neighedge", _vp), ("lam", _vp),
                ("mu", _vp), ("rho", _vp), ("penalty", _d), ("dt", _d), ("nsteps", _i), ("nsrc", _i),
                ("src_elem", _vp), ("src_phi", _vp), ("wavelet", _vp), ("nrec", _i), ("rec_elem", _vp),
                ("rec_phi", _vp), ("stride", _i), ("basis", _i)]


class Server:
    """The request loop (tsp_server_*)."""

    INSTANCES, COLUMNS, BATCHED = 0, 1, 2

    def __init__(self, device=0, mem_fraction=0.8, chunk=32, mode=COLUMNS):
        self._L = lib()
        h = _vp()
        rc = self._L.tsp_server_create(device, float(mem_fraction), int(chunk), ctypes.byref(h))
        if rc != TSP_OK:
            raise EngineError(rc, "tsp_server_create failed")
        self._h = h
        self.mode = mode
        self.meta = {}
        self._ck(self._L.tsp_server_set_mode(self._h, int(mode)))

    def _ck(self, rc):
        if rc != TSP_OK:
            raise EngineError(rc, self._L.tsp_server_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.tsp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def now(self):
        return self._L.tsp_server_now()

    def submit(self, prob, arrival, stride=1):
        """Queue Problem `prob` (all its sources: one request) arriving at `arrival`."""
        arrs = dict(verts=_c64(prob.verts), neigh=_c32i(prob.neigh), neighedge=_c32i(prob.neighedge),
                    lam=_c64(prob.lam), mu=_c64(prob.mu), rho=_c64(prob.rho), src_elem=_c32i(prob.src_elem),
                    src_phi=_c64(prob.src_phi), wavelet=_c64(prob.wavelet()), rec_elem=_c32i(prob.rec_elem),
                    rec_phi=_c64(prob.rec_phi))
        q = Request(prob.N, prob.ne, arrs["verts"].ctypes.data, arrs["neigh"].ctypes.data,
                    arrs["neighedge"].ctypes.data, arrs["lam"].ctypes.data, arrs["mu"].ctypes.data,
                    arrs["rho"].ctypes.data, prob.penalty, prob.dt, prob.nsteps, len(prob.src_elem),
                    arrs["src_elem"].ctypes.data, arrs["src_phi"].ctypes.data, arrs["wavelet"].ctypes.data,
                    len(prob.rec_elem), arrs["rec_elem"].ctypes.data, arrs["rec_phi"].ctypes.data, int(stride),
                    int(getattr(prob, "basis", DUBINER)))
        i = _i()
        self._ck(self._L.tsp_server_submit(self._h, ctypes.byref(q), float(arrival), ctypes.byref(i)))
        self.meta[i.value] = (len(prob.src_elem), len(prob.rec_elem))
        return i.value

    @staticmethod
    def footprint(prob, stride=1, mode=COLUMNS):
        """(resident, peak) device bytes the loop books for `prob`."""
        q = Request(prob.N, prob.ne, None, None, None, None, None, None, prob.penalty, prob.dt, prob.nsteps,
                    len(prob.src_elem), None, None, None, len(prob.rec_elem), None, None, int(stride))
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        rc = lib().tsp_server_footprint(ctypes.byref(q), int(mode), ctypes.byref(a), ctypes.byref(b))
        if rc != TSP_OK:
            raise EngineError(rc, "tsp_server_footprint: bad request")
        return a.value, b.value

    def tick(self):
        c, r, q = _i(), _i(), _i()
        self._ck(self._L.tsp_server_tick(self._h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(q)))
        return c.value, r.value, q.value

    def result(self, i):
        nsrc, nrec = self.meta[i]
        ns, a, b, c = _i(), _d(), _d(), _d()
        self._ck(self._L.tsp_server_result(self._h, int(i), None, ctypes.byref(ns), None, None, None))
        out = np.zeros(max(1, nsrc * ns.value * nrec * 2))
        self._ck(self._L.tsp_server_result(self._h, int(i), out.ctypes.data, ctypes.byref(ns), ctypes.byref(a),
                                           ctypes.byref(b), ctypes.byref(c)))
        return out[: nsrc * ns.value * nrec * 2].reshape(nsrc, ns.value, nrec, 2), dict(
            arrival=a.value, admitted=b.value, completed=c.value)

    def stats(self):
        b, d, u, g = _d(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self._ck(self._L.tsp_server_stats(self._h, ctypes.byref(b), ctypes.byref(d), ctypes.byref(u),
                                          ctypes.byref(g)))
        a, r, t = _d(), _d(), _d()
        self._ck(self._L.tsp_server_phases(self._h, ctypes.byref(a), ctypes.byref(r), ctypes.byref(t)))
        return dict(busy_s=b.value, dof_steps=d.value, mem_used=u.value, mem_budget=g.value,
                    admit_s=a.value, run_s=r.value, retire_s=t.value)

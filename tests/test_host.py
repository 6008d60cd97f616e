"""Host C++ layer (tspeed_amd/csrc/host) against the oracle and closed forms.
CPU only: nothing here touches the GPU."""
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tspeed_amd import capi, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MESHES = "/root/reference/Examples/Meshes"


@pytest.mark.parametrize("N", range(1, 9))
def test_tables_equal_oracle(N):
    a, b = capi.tables(N), O.tables(N)
    for k in ("B", "G", "Bedge", "Gedge", "w2", "w1", "x1"):
        np.testing.assert_array_equal(a[k].ravel(), b[k].ravel(), err_msg=k)
    np.testing.assert_array_equal(a["xy2"].ravel(), b["xy2"].ravel())


@pytest.mark.parametrize("N", [1, 4, 8])
def test_dubiner_point_values_equal_oracle(N):
    rng = np.random.default_rng(N)
    for _ in range(10):
        x, y = rng.uniform(0.0, 0.5, 2)
        p, g = capi.dubiner(N, x, y)
        po, go = O.dubiner(N, x, y)
        np.testing.assert_array_equal(p, po)
        np.testing.assert_array_equal(g, go)


@pytest.mark.parametrize("nx,ny,jit,seed", [(3, 2, 0.0, 1), (7, 5, 0.2, 2), (12, 9, 0.2, 3)])
def test_edge_hash_neighbors_equal_bruteforce(nx, ny, jit, seed):
    pts, tri = capi.split_square(nx, ny, 0.0, 0.0, 1.0, 1.0, jit, seed)
    nb, ng = capi.neighbors(tri)
    nbo, ngo = O.neighbors(tri)
    np.testing.assert_array_equal(nb, nbo)
    np.testing.assert_array_equal(ng, ngo)


def test_split_square_is_valid_at_layered_size():
    """160k triangles, jitter 0.2h: every triangle counter-clockwise, boundary
    vertices on the boundary, interior jitter within +-0.2h."""
    pts, tri = capi.split_square(400, 200, 0.0, 0.0, 4000.0, 2000.0, 0.2, 2)
    v = pts[tri]
    area = 0.5 * ((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0]))
    assert tri.shape == (160000, 3)
    assert area.min() > 0
    np.testing.assert_allclose(area.sum(), 4000.0 * 2000.0, rtol=1e-12)
    grid = np.stack(np.meshgrid(np.arange(401) * 10.0, np.arange(201) * 10.0), -1).reshape(-1, 2)
    d = np.abs(pts - grid)
    assert d.max() <= 0.2 * 10.0 + 1e-12
    onb = (grid[:, 0] == 0) | (grid[:, 0] == 4000) | (grid[:, 1] == 0) | (grid[:, 1] == 2000)
    assert d[onb].max() == 0.0


def test_locate_equals_linear_scan():
    pts, tri = capi.split_square(9, 7, 0.0, 0.0, 900.0, 700.0, 0.2, 9)
    verts = pts[tri].reshape(-1, 6)
    rng = np.random.default_rng(0)
    q = rng.uniform([0, 0], [900, 700], (300, 2))
    # include vertices and edge midpoints, where several elements pass the tolerant test
    q = np.concatenate([q, pts, 0.5 * (pts[tri[:, 0]] + pts[tri[:, 1]])])
    el, xi, phi = capi.locate(pts, tri, q[:, 0], q[:, 1], 3)
    for k, (x, y) in enumerate(q):
        eo, xo = O.locate(verts, x, y)
        assert el[k] == eo, (k, x, y)
        if eo >= 0:
            np.testing.assert_array_equal(xi[k], xo)
    el2, _, _ = capi.locate(pts, tri, [-5.0, 950.0], [10.0, 10.0], 3)
    assert (el2 == -1).all()


def test_ricker_table_times():
    """Row 0 at t = 0, row 1 at dt + dt, then repeated addition (TimeAdvance.cpp:48)."""
    dt = 1.3e-4
    tab = capi.ricker_table([10.0], [0.12], [0.0], [-1.0], dt, 50)
    t, ts = 0.0, [0.0]
    for r in range(1, 50):
        t = dt + dt if r == 1 else t + dt
        ts.append(t)
    exp = np.array([-O.ricker(10.0, 0.12, t) for t in ts])
    np.testing.assert_array_equal(tab[:, 0, 1], exp)
    np.testing.assert_array_equal(tab[:, 0, 0], 0.0 * exp)


def test_config_builders():
    p = W.square_ricker()
    assert p.ne == 3200 and p.ndof == 96000 and p.N == 4
    assert p.rec_xy.shape == (32, 2) and (p.rec_elem >= 0).all()
    lam, mu = p.lam[0], p.mu[0]
    np.testing.assert_allclose(np.sqrt((lam + 2 * mu) / 2000.0), 3000.0)
    np.testing.assert_allclose(np.sqrt(mu / 2000.0), 1732.0)
    assert 0 < p.dt < 1e-4
    small = W.layered_ricker(nx=40, ny=20)
    vs = np.sqrt(small.mu / small.rho)
    assert vs.min() > 800 * np.exp(-0.6) and vs.max() < 2500 * np.exp(0.6)
    np.testing.assert_allclose((small.lam / small.mu), 1.0, rtol=1e-12)   # vp = sqrt(3) vs


# -------- reference mesh fixtures: known counts (read when the reference is mounted)
MSH_FACTS = {
    "Lamb_fullyunstruct3.msh": dict(nodes=822, tri=1532, bnd=110),
    "wedge.msh": dict(nodes=1234, tri=2360, bnd=162),
    "Lamb.msh": dict(nodes=3729, tri=7220, bnd=236),
    "Lamb_fullyunstruct.msh": dict(tri=1178),
    "Lamb_fullyunstruct2.msh": dict(tri=2298),
    "Lamb_new_unstruct.msh": dict(tri=2686),
    "wedge2.msh": dict(tri=2970),
}


@pytest.mark.skipif(not os.path.isdir(REF_MESHES), reason="reference meshes not mounted")
@pytest.mark.parametrize("name", sorted(MSH_FACTS))
def test_msh_reader_and_neighbors_on_reference_meshes(name):
    pts, tri, reg, bnd = capi.read_msh(os.path.join(REF_MESHES, name))
    f = MSH_FACTS[name]
    assert tri.shape[0] == f["tri"]
    if "nodes" in f:
        assert pts.shape[0] == f["nodes"]
        assert bnd.shape[0] == f["bnd"]
    nb, ng = capi.neighbors(tri)
    # 2*interior + boundary = 3*ne; every boundary half-face is one of the mesh's
    # el_type-1 lines (wedge.msh also carries lines on its internal interface)
    nbnd = int((nb < 0).sum())
    assert (3 * tri.shape[0] - nbnd) % 2 == 0
    lines = {tuple(sorted(b)) for b in bnd.tolist()}
    for e, fa in zip(*np.nonzero(nb < 0)):
        assert tuple(sorted((tri[e, (fa + 1) % 3], tri[e, (fa + 2) % 3]))) in lines
    if name.startswith("Lamb") and "bnd" in f:
        assert nbnd == f["bnd"]
    if tri.shape[0] <= 3000:
        nbo, ngo = O.neighbors(tri)
        np.testing.assert_array_equal(nb, nbo)
        np.testing.assert_array_equal(ng, ngo)
    v = pts[tri]
    area = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0])
    assert (area > 0).all()                        # all counter-clockwise
    if name == "wedge.msh":
        assert sorted(set(reg.tolist())) == [1, 2] and (reg == 1).sum() == 548


def test_config_receivers_file_format_roundtrip(tmp_path):
    """Receivers::write format (Receivers.cpp:32-46): one 'x y' line per sample;
    with 17 digits the file carries every double exactly."""
    from tspeed_amd import io
    s = np.arange(24, dtype=np.float64).reshape(4, 3, 2) * 1.5e-7 + np.pi * 1e-9
    io.write_receivers(str(tmp_path / "run"), s, digits=17)
    np.testing.assert_array_equal(io.read_receivers(str(tmp_path / "run"), 3), s)


# Values whose 6-digit forms exercise %g's switches: fixed / exponent notation,
# trailing-zero removal, rounding up to a new decade, signed zero, large exponents.
_SEIS_VALUES = [0.0, -0.0, 1.0, -2.5, 1e-5, 1.234567e-5, 9.9999996e-5, 123456.7, 1234567.0, 0.1 + 0.2,
                -3.14159265358979e-12, 6.02214076e23, 1e100, 5e-324, 2.0 / 3.0, 999999.5, 1e-4, 100000.0]


def _ostream_reference(tmp_path, values):
    """The bytes `outf << x << " " << y << std::endl` writes with a default
    std::ofstream (Receivers.cpp:40-43), from a program compiled here."""
    import subprocess
    src = tmp_path / "ref.cpp"
    src.write_text('#include <fstream>\n#include <cstdio>\nint main(int c, char **v) {\n'
                   '  std::ofstream f(v[1]); double x, y;\n'
                   '  while (std::scanf("%la %la", &x, &y) == 2) f << x << " " << y << std::endl;\n}\n')
    exe = tmp_path / "ref"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    out = tmp_path / "ref.out"
    pairs = "\n".join(f"{float(a).hex()} {float(b).hex()}" for a, b in zip(values[0::2], values[1::2]))
    subprocess.run([str(exe), str(out)], input=pairs, text=True, check=True)
    return out.read_bytes()


def test_receiver_files_are_byte_identical_to_the_reference_format(tmp_path):
    """io.write_receivers and the C++ Receivers::write (through
    tsh_write_receivers) write, by default, exactly the bytes of the
    reference's default-formatted std::ostream output."""
    import ctypes
    from tspeed_amd import io
    vals = np.array(_SEIS_VALUES)
    seis = vals.reshape(-1, 1, 2)                  # [sample][receiver 0][x|y]
    want = _ostream_reference(tmp_path, vals)
    io.write_receivers(str(tmp_path / "py"), seis)
    assert (tmp_path / "py_rcv_0.out").read_bytes() == want
    L = capi.lib()
    flat = np.ascontiguousarray(seis.reshape(-1))
    assert L.tsh_write_receivers(str(tmp_path / "cxx").encode(), 1, seis.shape[0], flat, 0) == 0
    assert (tmp_path / "cxx_rcv_0.out").read_bytes() == want
    # 17 digits: both writers agree and the values come back exactly
    assert L.tsh_write_receivers(str(tmp_path / "c17").encode(), 1, seis.shape[0], flat, 17) == 0
    io.write_receivers(str(tmp_path / "p17"), seis, digits=17)
    assert (tmp_path / "c17_rcv_0.out").read_bytes() == (tmp_path / "p17_rcv_0.out").read_bytes()
    np.testing.assert_array_equal(io.read_receivers(str(tmp_path / "c17"), 1), seis)
    assert L.tsh_write_receivers(str(tmp_path / "no" / "such" / "dir").encode(), 1, 1, flat, 0) == -1


# ------------------------------------------------------------------ plot output and error norm
def _unit_square(n=3, N=3, seed=2):
    pts, tri = capi.split_square(n, n, 0.0, 0.0, 1.0, 1.0, 0.2, seed)
    return pts[tri].reshape(-1, 6), N


def test_field_points_are_the_affine_images():
    """FESpace::points_out (FESpace_imp.hpp:71-83): the four reference points
    mapped by x = v0 + (v1 - v0) xi + (v2 - v0) eta."""
    verts, _ = _unit_square()
    p = capi.field_points(verts)
    v = verts.reshape(-1, 3, 2)
    ref = np.array([[0.2, 0.2], [0.8, 0.2], [0.2, 0.8], [0.5, 0.5]])
    expect = v[:, None, 0] + ref[None, :, :1] * (v[:, None, 1] - v[:, None, 0]) + ref[None, :, 1:] * (v[:, None, 2] - v[:, None, 0])
    np.testing.assert_allclose(p, expect, rtol=0, atol=1e-15)


@pytest.mark.parametrize("N", [1, 2, 4, 6])
def test_projection_reproduces_polynomials_and_l2_error(N):
    """FESpace::transform of a degree-N polynomial field is exact: its values at
    the plot points equal the polynomial there (field_out), and its L2 error
    (FESpace::L2error) vanishes; the L2 error of uh = 0 against (1, 2) is
    sqrt(5 * area)."""
    verts, _ = _unit_square(N=N)
    def fun(x, y):
        return (1 + x - 2 * y + x ** N, 0.5 * y ** N - x * y ** (N - 1) + 3)
    uh = capi.project(N, verts, fun)
    pts = capi.field_points(verts)
    vals = capi.field_values(N, uh)
    fx, fy = fun(pts[..., 0], pts[..., 1])
    np.testing.assert_allclose(vals[..., 0], fx, rtol=0, atol=1e-11)
    np.testing.assert_allclose(vals[..., 1], fy, rtol=0, atol=1e-11)
    assert capi.l2_error(N, verts, uh, fun) <= 1e-11
    zero = np.zeros_like(uh)
    assert capi.l2_error(N, verts, zero, lambda x, y: (1.0, 2.0)) == pytest.approx(np.sqrt(5.0), rel=1e-13)
    # a degree N+1 field is not in the space: a positive projection error
    assert capi.l2_error(N, verts, capi.project(N, verts, lambda x, y: (x ** (N + 1), 0 * x)),
                         lambda x, y: (x ** (N + 1), 0 * x)) > 1e-10


# ------------------------------------------------------------------ BoundaryAdapted basis (host side)
@pytest.mark.parametrize("N", [1, 2, 4, 8])
def test_boundary_adapted_tables_match_oracle(N):
    """The engine's BoundaryAdapted tables and base_invmass equal the oracle's
    (same expressions, forward-mode gradients), and tsh_basis_eval agrees."""
    h = capi.tables(N, capi.BOUNDARY_ADAPTED)
    o = O.tables(N, O.BOUNDARY_ADAPTED)
    np.testing.assert_allclose(h["B"].reshape(o["B"].shape), o["B"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(h["G"].reshape(o["G"].shape), o["G"], rtol=1e-14, atol=1e-13)
    np.testing.assert_allclose(h["Bedge"].reshape(o["Bedge"].shape), o["Bedge"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(h["Gedge"].reshape(o["Gedge"].shape), o["Gedge"], rtol=1e-14, atol=1e-13)
    np.testing.assert_allclose(capi.base_invmass(N, capi.BOUNDARY_ADAPTED), O.base_invmass(N, O.BOUNDARY_ADAPTED),
                               rtol=1e-11, atol=1e-11)
    np.testing.assert_array_equal(capi.base_invmass(N), np.eye(capi.nln(N)))
    xi = np.array([[0.1, 0.2], [0.7, 0.25], [0.05, 0.9]])
    np.testing.assert_allclose(capi.basis_eval(N, capi.BOUNDARY_ADAPTED, xi),
                               np.array([O.boundary_adapted(N, *p)[0] for p in xi]), rtol=0, atol=1e-15)


@pytest.mark.parametrize("N", [2, 5])
def test_boundary_adapted_projection_exact(N):
    """FESpace::transform with the BoundaryAdapted basis (base_invmass applied)
    reproduces degree-N polynomials at the plot points and in L2."""
    verts, _ = _unit_square(N=N)
    def fun(x, y):
        return (1 + x - 2 * y + x ** N, 0.5 * y ** N - x * y + 3)
    ba = capi.BOUNDARY_ADAPTED
    uh = capi.project(N, verts, fun, ba)
    pts = capi.field_points(verts)
    vals = capi.field_values(N, uh, ba)
    fx, fy = fun(pts[..., 0], pts[..., 1])
    np.testing.assert_allclose(vals[..., 0], fx, rtol=0, atol=1e-11)
    np.testing.assert_allclose(vals[..., 1], fy, rtol=0, atol=1e-11)
    assert capi.l2_error(N, verts, uh, fun, ba) <= 1e-11

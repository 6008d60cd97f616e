"""Pins the FP64 CPU oracle (oracle/) against golden vectors, closed forms and
invariants, so that it can serve as the checker of the HIP path.

Each test names what a plausible mistake in the oracle would break."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tspeed_amd import capi

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ quadrature
@pytest.mark.parametrize("n", range(1, 13))
def test_gauleg_matches_numpy_leggauss(n):
    g = np.load(os.path.join(GOLD, "gauss_legendre.npz"))
    x, w = O.gauleg(n, -1.0, 1.0)
    np.testing.assert_allclose(x, g[f"x{n}"], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, g[f"w{n}"], rtol=0, atol=2e-15)
    # mapped to [0, 1]: nodes (1+x)/2, weights w/2
    x01, w01 = O.gauleg(n, 0.0, 1.0)
    np.testing.assert_allclose(x01, (1 + g[f"x{n}"]) / 2, atol=2e-15)
    np.testing.assert_allclose(w01, g[f"w{n}"] / 2, atol=2e-15)


@pytest.mark.parametrize("nq", range(2, 10))
def test_collapsed_rule_integrates_monomials(nq):
    """int_T x^a y^b = a! b! / (a+b+2)! exactly up to total degree 2*nq-2 (the Duffy
    Jacobian (1-x_i)/8 uses one degree; QuadratureRule.hpp:168-169: nq = N+1
    integrates degree 2N).  Dropping a factor or the Jacobian breaks this."""
    w2, xy, w1, x1 = O.gauss_rule(nq)
    assert abs(w2.sum() - 0.5) < 1e-15
    for a in range(2 * nq - 1):
        for b in range(2 * nq - 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = float(np.sum(w2 * xy[:, 0] ** a * xy[:, 1] ** b))
            assert abs(got - exact) <= 1e-14 * max(1.0, exact) + 1e-16, (a, b, got, exact)


def test_edge_nodes_lie_on_their_edges_in_order():
    nq = 5
    _, _, _, x1 = O.gauss_rule(nq)
    e0, e1, e2 = (O.edge_nodes(nq, x1, f) for f in range(3))
    np.testing.assert_allclose(e0.sum(1), 1.0, atol=1e-15)       # x + y = 1
    np.testing.assert_allclose(e1[:, 0], 0.0)                      # x = 0
    np.testing.assert_allclose(e2[:, 1], 0.0)                      # y = 0
    # edge f runs from vertex f+1 to f+2 (P1->P2, P2->P0, P0->P1)
    assert e0[0, 0] > e0[-1, 0] and e1[0, 1] > e1[-1, 1] and e2[0, 0] < e2[-1, 0]


# ------------------------------------------------------------------ Jacobi / Dubiner
def test_jacobi_matches_scipy():
    g = np.load(os.path.join(GOLD, "jacobi.npz"))
    for (n, a, b), ref in zip(g["cases"], g["vals"]):
        got = np.array([O.jacobi(int(n), int(a), int(b), z) for z in g["z"]])
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * max(1, np.abs(ref).max()))


@pytest.mark.parametrize("N", range(1, 9))
def test_dubiner_orthonormal(N):
    """int_T phi_i phi_j = delta_ij with an over-integrating rule; phi_00 = sqrt(2).
    A wrong c_ij, a wrong Jacobi parameter or a swapped (i, j) order fails."""
    w2, xy, _, _ = O.gauss_rule(N + 2)
    P = np.array([O.dubiner(N, x, y)[0] for x, y in xy])
    M = (P * w2[:, None]).T @ P
    np.testing.assert_allclose(M, np.eye(O.nln(N)), atol=5e-14)
    np.testing.assert_allclose(P[:, 0], math.sqrt(2.0), rtol=1e-15)


def test_dubiner_mode_order_is_j_major():
    """Mode s = (i, j) enumerated for j in 0..N, i in 0..N-j: mode 1 is (1,0), which
    depends on xi at fixed eta; mode N+1 is (0,1), which depends on eta only."""
    N = 3
    a = O.dubiner(N, 0.1, 0.3)[0]
    b = O.dubiner(N, 0.2, 0.3)[0]
    assert a[1] != b[1]
    assert a[N + 1] == b[N + 1]


@pytest.mark.parametrize("N", [1, 2, 4, 6, 8])
def test_dubiner_gradient_is_derivative(N):
    rng = np.random.default_rng(N)
    h = 1e-6
    for _ in range(20):
        x, y = rng.uniform(0.05, 0.45), rng.uniform(0.05, 0.45)
        _, g = O.dubiner(N, x, y)
        fx = (O.dubiner(N, x + h, y)[0] - O.dubiner(N, x - h, y)[0]) / (2 * h)
        fy = (O.dubiner(N, x, y + h)[0] - O.dubiner(N, x, y - h)[0]) / (2 * h)
        scale = max(1.0, np.abs(g).max())
        np.testing.assert_allclose(g[:, 0], fx, atol=2e-6 * scale)
        np.testing.assert_allclose(g[:, 1], fy, atol=2e-6 * scale)


# ------------------------------------------------------------------ geometry / mesh
def _mesh(nx=5, ny=4, jit=0.2, seed=11, L=(500.0, 400.0)):
    pts, tri = capi.split_square(nx, ny, 0.0, 0.0, L[0], L[1], jit, seed)
    return pts, tri, pts[tri].reshape(-1, 6)


def test_geometry_closed_forms():
    pts, tri, verts = _mesh()
    g = O.geometry(verts)
    v = verts.reshape(-1, 3, 2)
    area = 0.5 * ((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0]))
    np.testing.assert_allclose(g["detJ"], 2 * area, rtol=1e-14)
    np.testing.assert_allclose(g["J"] @ g["invJ"], np.broadcast_to(np.eye(2), g["J"].shape), atol=1e-14)
    for f in range(3):
        a, b = v[:, (f + 1) % 3], v[:, (f + 2) % 3]
        t = b - a
        n = g["normal"][:, f]
        np.testing.assert_allclose(np.hypot(*t.T), g["length"][:, f], rtol=1e-15)
        np.testing.assert_allclose((n * t).sum(1), 0.0, atol=1e-12)
        np.testing.assert_allclose(np.hypot(*n.T), 1.0, rtol=1e-15)
        # outward: points away from the opposite vertex
        assert ((n * (a - v[:, f])).sum(1) > 0).all()


def test_neighbors_bruteforce_invariants():
    pts, tri, _ = _mesh(6, 5)
    nb, ng = O.neighbors(tri)
    ne = tri.shape[0]
    nint = (nb >= 0).sum()
    nbnd = (nb < 0).sum()
    assert nint + nbnd == 3 * ne
    assert nbnd == 2 * (6 + 5)                      # boundary edges of the rectangle
    for e in range(ne):
        for f in range(3):
            if nb[e, f] >= 0:
                assert nb[nb[e, f], ng[e, f]] == e and ng[nb[e, f], ng[e, f]] == f


def test_locate_reference_coordinates():
    pts, tri, verts = _mesh()
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform([1, 1], [499, 399])
        e, xi = O.locate(verts, *p)
        assert e >= 0
        P = verts[e].reshape(3, 2)
        back = P[0] + (P[1] - P[0]) * xi[0] + (P[2] - P[0]) * xi[1]
        np.testing.assert_allclose(back, p, atol=1e-9)
        assert xi.min() > -1e-9 and xi.sum() < 1 + 1e-9


# ------------------------------------------------------------------ operator
def _operator(N=3, nx=4, ny=3, hetero=True, penalty=10.0, seed=3):
    pts, tri, verts = _mesh(nx, ny, seed=seed)
    nb, ng = O.neighbors(tri)
    ne = tri.shape[0]
    rng = np.random.default_rng(seed)
    rho = np.full(ne, 2000.0) * (rng.uniform(0.8, 1.2, ne) if hetero else 1)
    mu = np.full(ne, 6e9) * (rng.uniform(0.5, 1.5, ne) if hetero else 1)
    lam = np.full(ne, 6e9) * (rng.uniform(0.5, 1.5, ne) if hetero else 1)
    return O.Operator(N, verts, nb, ng, lam, mu, rho, penalty), verts, nb, (lam, mu, rho)


def _full(op, nb):
    """Dense B in the component-major ordering of the reference's vectors."""
    nl, ne = op.nln, op.ne
    d = 2 * nl
    Bk = op.blocks()
    idx = np.array([[c * ne * nl + e * nl + i for c in range(2) for i in range(nl)] for e in range(ne)])
    M = np.zeros((op.ndof, op.ndof))
    for e in range(ne):
        for s in range(4):
            c = e if s == 0 else nb[e, s - 1]
            if c >= 0:
                M[np.ix_(idx[e], idx[c])] += Bk[e, s]
    return M


def _project(op, verts, fun):
    """L2 projection onto the orthonormal basis (FESpace::transform)."""
    T = O.tables(op.N)
    g = O.geometry(verts)
    nl, ne = op.nln, op.ne
    u = np.zeros(op.ndof)
    for e in range(ne):
        P0, J, det = verts[e, :2], g["J"][e], g["detJ"][e]
        xy = P0 + T["xy2"] @ J.T
        F = np.array([fun(x, y) for x, y in xy])
        for c in range(2):
            u[c * ne * nl + e * nl:c * ne * nl + (e + 1) * nl] = (abs(det) * T["w2"] * F[:, c]) @ T["B"] / det
    return u


def test_operator_symmetric_and_psd():
    op, verts, nb, _ = _operator()
    M = _full(op, nb)
    assert np.abs(M - M.T).max() <= 1e-14 * np.abs(M).max()
    ev = np.linalg.eigvalsh(M)
    assert ev[0] > -1e-10 * ev[-1]
    assert (ev < 1e-10 * ev[-1]).sum() == 3          # exactly the three rigid-body modes


@pytest.mark.parametrize("basis", [O.DUBINER, O.BOUNDARY_ADAPTED])
@pytest.mark.parametrize("N", range(1, 9))
def test_face_blocks_are_transposes_of_each_other(N, basis):
    """B_{b,a} = B_{a,b}^T on every interior face, to the rounding of assembly
    (the consistency term enters as I + I^T and the penalty block is the same
    form seen from either side), at every degree, in both bases, with
    per-element materials: the premise of the FP64 face form, which reads one
    copy of each face block for both sides.  The element's own blocks are
    symmetric too."""
    pts, tri, verts = _mesh(4, 3, seed=5 + N)
    nb, ng = O.neighbors(tri)
    ne = tri.shape[0]
    rng = np.random.default_rng(N)
    op = O.Operator(N, verts, nb, ng, 6e9 * rng.uniform(0.5, 1.5, ne), 6e9 * rng.uniform(0.5, 1.5, ne),
                    2000.0 * rng.uniform(0.8, 1.2, ne), 10.0, basis=basis)
    B = op.blocks()
    scale = np.abs(B).max()
    worst = 0.0
    for a in range(ne):
        worst = max(worst, np.abs(B[a, 0] - B[a, 0].T).max())
        for f in range(3):
            b = nb[a, f]
            if b >= 0:
                worst = max(worst, np.abs(B[a, 1 + f] - B[b, 1 + ng[a, f]].T).max())
    assert worst <= 1e-15 * scale


@pytest.mark.parametrize("N", [1, 2, 4])
def test_rigid_body_modes_in_kernel(N):
    """All boundaries are traction-free (Matrices_imp.hpp:107,155 skip boundary
    edges), so translations and the infinitesimal rotation have zero energy."""
    op, verts, nb, _ = _operator(N=N)
    for fun in (lambda x, y: (1.0, 0.0), lambda x, y: (0.0, 1.0), lambda x, y: (-y, x)):
        u = _project(op, verts, fun)
        Bu = op.apply(u)
        assert np.abs(Bu).max() <= 1e-12 * np.abs(op.blocks()).max() * np.abs(u).max()


@pytest.mark.parametrize("N", [1, 3])
def test_energy_of_linear_fields(N):
    """For a continuous linear displacement the jump terms vanish and
    u^T B u = int sigma:eps = sum_K |K| (lambda tr(eps)^2 + 2 mu eps:eps).
    Swapping lambda and mu, or a wrong strain, fails this."""
    op, verts, nb, (lam, mu, rho) = _operator(N=N)
    g = O.geometry(verts)
    area = np.abs(g["detJ"]) / 2
    for a, b, c, d in [(1e-3, 0, 0, 0), (0, 2e-3, 0, 0), (0, 0, 0, 1e-3), (1e-3, -2e-3, 3e-4, 5e-4)]:
        u = _project(op, verts, lambda x, y: (a * x + b * y, c * x + d * y))
        eps = np.array([[a, 0.5 * (b + c)], [0.5 * (b + c), d]])
        dens = lam * np.trace(eps) ** 2 + 2 * mu * (eps * eps).sum()
        np.testing.assert_allclose(u @ op.apply(u), (area * dens).sum(), rtol=1e-11)


def test_invm_is_inverse_density_jacobian():
    op, verts, nb, (lam, mu, rho) = _operator()
    g = O.geometry(verts)
    np.testing.assert_allclose(op.invm(), 1.0 / (rho * g["detJ"]), rtol=1e-15)


def test_penalty_scales_only_the_jump_part():
    """B(pen) - B(0) is pen * N^2 * S, and S annihilates continuous fields."""
    N = 2
    op1, verts, nb, mats = _operator(N=N, penalty=10.0)
    pts, tri, _ = _mesh(4, 3, seed=3)
    nbb, ngg = O.neighbors(tri)
    op0 = O.Operator(N, verts, nbb, ngg, *mats, 0.0)
    S = (_full(op1, nb) - _full(op0, nb)) / (10.0 * N * N)
    u = _project(op1, verts, lambda x, y: (1e-3 * x * y, 2e-3 * x - 1e-3 * y))  # continuous, degree 2
    assert np.abs(S @ u).max() <= 1e-12 * np.abs(S).max() * np.abs(u).max()
    assert np.linalg.eigvalsh(S).min() > -1e-8 * np.abs(S).max()


# ------------------------------------------------------------------ SIPG penalty block, closed form
def _two_triangles():
    """Two counter-clockwise triangles sharing the edge (100,0)-(0,100)."""
    pts = np.array([[0.0, 0.0], [100.0, 0.0], [0.0, 100.0], [130.0, 120.0]])
    tri = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    return pts, tri


def _face_matrix(lam_e, lam_n, mu_e, mu_n, n, mean="harmonic", two_mu=True):
    """The penalty tensor of one face (Matrices_imp.hpp:167-173) from the face
    averages of Parameters::avg_p (Parameters.cpp:37-45).  `mean` and `two_mu`
    switch in the plausible mistakes the test must be able to see."""
    def avg(a, b):
        if mean == "arithmetic":
            return 0.5 * (a + b)
        return 0.0 if a + b == 0 else 2 * a * b / (a + b)
    al, am = avg(lam_e, lam_n), avg(mu_e, mu_n)
    k = 2.0 if two_mu else 1.0
    c00 = (al + k * am) * n[0] ** 2 + k * am * n[1] ** 2
    c11 = (al + k * am) * n[1] ** 2 + k * am * n[0] ** 2
    c01 = al * n[0] * n[1]
    return np.array([[c00, c01], [c01, c11]])


def _pw_linear_basis(verts):
    """psi_{e,c,m}: component c of element e equal to 1, (x - xc)/h or (y - yc)/h
    on e and zero elsewhere; returns (centroid, h) per element."""
    v = verts.reshape(-1, 3, 2)
    a, b = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    return v.mean(1), np.sqrt(np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]))


def _penalty_closed_form(verts, nb, lam, mu, **mutation):
    """Q[a, b] = sum over interior faces of (1/|e|) int_e [psi_a]^T C [psi_b] ds
    for the piecewise-linear basis psi of _pw_linear_basis, integrated exactly
    (a product of two linear functions on a segment: |e|/6 (2 f_A g_A + f_A g_B
    + f_B g_A + 2 f_B g_B)).  Only vertex coordinates and materials enter: no
    quadrature rule and no basis table of the oracle."""
    v = verts.reshape(-1, 3, 2)
    ne = v.shape[0]
    cen, h = _pw_linear_basis(verts)
    reverse = mutation.pop("unreversed", False)
    Q = np.zeros((6 * ne, 6 * ne))

    def values(e, X):  # [6 basis functions of element e] x [2 components] at point X
        m = np.array([1.0, (X[0] - cen[e, 0]) / h[e], (X[1] - cen[e, 1]) / h[e]])
        out = np.zeros((6, 2))
        for c in range(2):
            out[3 * c:3 * c + 3, c] = m
        return out

    for e in range(ne):
        for f in range(3):
            o = nb[e, f]
            if o < e:          # boundary (-1) or already counted from the other side
                continue
            A, B = v[e, (f + 1) % 3], v[e, (f + 2) % 3]
            L = np.hypot(*(B - A))
            n = np.array([B[1] - A[1], A[0] - B[0]]) / L
            C = _face_matrix(lam[e], lam[o], mu[e], mu[o], n, **mutation)
            idx = np.r_[6 * e:6 * e + 6, 6 * o:6 * o + 6]
            # jump [psi] = psi|_e - psi|_o at the two ends of the face
            jA = np.vstack([values(e, A), -values(o, B if reverse else A)])
            jB = np.vstack([values(e, B), -values(o, A if reverse else B)])
            Q[np.ix_(idx, idx)] += (2 * jA @ C @ jA.T + jA @ C @ jB.T + jB @ C @ jA.T + 2 * jB @ C @ jB.T) / 6
    return Q


def _pw_linear_coefficients(N, verts):
    """Columns: the oracle-space coefficients of every psi_{e,c,m} (exact
    L2 projection, FESpace::transform; exact since N >= 1)."""
    ne = verts.shape[0]
    cen, h = _pw_linear_basis(verts)
    nl = O.nln(N)
    Phi = np.zeros((2 * ne * nl, 6 * ne))
    for e in range(ne):
        for c in range(2):
            for m in range(3):
                def fun(x, y, e=e, c=c, m=m):
                    val = [np.ones_like(x), (x - cen[e, 0]) / h[e], (y - cen[e, 1]) / h[e]][m]
                    return (val if c == 0 else 0 * x, val if c == 1 else 0 * x)
                u = capi.project(N, verts, fun)
                mask = np.zeros(2 * ne * nl, bool)
                for cc in range(2):
                    mask[cc * ne * nl + e * nl:cc * ne * nl + (e + 1) * nl] = True
                Phi[mask, 6 * e + 3 * c + m] = u[mask]
    return Phi


@pytest.mark.parametrize("N,mesh", [(1, "two"), (2, "two"), (1, "grid"), (3, "grid")])
def test_penalty_block_matches_face_integral_closed_form(N, mesh):
    """Pins Matrices::stability (Matrices_imp.hpp:142-182) and the harmonic face
    averages (Parameters::avg_p, Parameters.cpp:37-45) on meshes whose elements
    all have different lambda != 0 (one lambda = 0) and mu.

    S = (B(pen) - B(0)) / (pen N^2) satisfies, for discontinuous fields u, v,
    v^T S u = sum_faces (1/|e|) int_e [v]^T C [u] ds with the jump
    [u] = u_e - u_nb and C the face tensor built from the harmonic means.  For
    every piecewise-linear pair this is integrated exactly from the vertex
    coordinates alone.  At N = 1 the piecewise-linear fields span the whole
    space, so the whole S is pinned; at N = 2, 3 its restriction to them.
    The same closed form with the arithmetic mean, with mu-bar in place of
    2 mu-bar, or with the neighbour's trace not reversed (kk = k instead of
    nq - 1 - k) is far outside the tolerance, so each of these mistakes in the
    oracle would fail the test (measured: the oracle within 0.6-3e-15 of the
    closed form; the three mistakes 0.33, 0.42 and 0.08-0.14 away)."""
    if mesh == "two":
        pts, tri = _two_triangles()
        lam, mu = np.array([7.0e9, 2.5e9]), np.array([3.0e9, 9.0e9])
    else:
        pts, tri = capi.split_square(3, 2, 0.0, 0.0, 300.0, 200.0, 0.2, 17)
        ne = tri.shape[0]
        rng = np.random.default_rng(23)
        lam, mu = rng.uniform(1e9, 9e9, ne), rng.uniform(1e9, 9e9, ne)
        lam[3] = 0.0   # avg_p's harmonic mean of 0 and x is 0 (an arithmetic mean is not)
    verts = pts[tri].reshape(-1, 6)
    nb, ng = O.neighbors(tri)
    ne = tri.shape[0]
    rho = np.full(ne, 2000.0)
    pen = 10.0
    op1 = O.Operator(N, verts, nb, ng, lam, mu, rho, pen)
    op0 = O.Operator(N, verts, nb, ng, lam, mu, rho, 0.0)
    S = (_full(op1, nb) - _full(op0, nb)) / (pen * N * N)
    Phi = _pw_linear_coefficients(N, verts)
    got = Phi.T @ S @ Phi
    want = _penalty_closed_form(verts, nb, lam, mu)
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(got - want).max() / scale
    assert err <= 1e-12, err
    for mutation in (dict(mean="arithmetic"), dict(two_mu=False), dict(unreversed=True)):
        wrong = _penalty_closed_form(verts, nb, lam, mu, **mutation)
        dist = np.abs(got - wrong).max() / scale
        assert dist > 1e-3, mutation


# ------------------------------------------------------------------ leap-frog
def test_energy_conserved_without_force():
    """E^{n+1/2} = 1/2 v^T M v + 1/2 u^{n+1}^T B u^n with v = (u^{n+1}-u^n)/dt is
    exactly conserved by leap-frog when f = 0.  A wrong factor 2, sign or
    missing invM breaks it."""
    op, verts, nb, (lam, mu, rho) = _operator(N=2)
    g = O.geometry(verts)
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal(op.ndof) * 1e-3
    lm = op.lambda_max(300)
    dt = 0.5 * 2 / np.sqrt(lm)
    M = np.concatenate([np.repeat(rho * g["detJ"], op.nln)] * 2)
    energies = []
    prev = None
    for nsteps in range(0, 12):
        u, _ = op.run(dt, nsteps, u0=u0)
        if prev is not None:
            v = (u - prev) / dt
            energies.append(0.5 * v @ (M * v) + 0.5 * u @ op.apply(prev))
        prev = u
    energies = np.array(energies)
    assert np.abs(energies - energies[0]).max() <= 1e-11 * abs(energies[0])


def test_first_step_from_rest_is_source_only():
    """u1 = invM * dt^2/2 * f(0): non-zero only on the source element, with the
    force sign and the basis values at the source point."""
    op, verts, nb, (lam, mu, rho) = _operator(N=2)
    g = O.geometry(verts)
    e, xi = O.locate(verts, 250.0, 200.0)
    phi = O.dubiner(2, *xi)[0]
    src = dict(elem=e, phi=phi, f0=10.0, t0=0.05, dirx=0.3, diry=-1.0)
    dt = 1e-4
    u, _ = op.run(dt, 0, src=src)
    r0 = O.ricker(10.0, 0.05, 0.0)
    nl, ne = op.nln, op.ne
    exp = np.zeros(op.ndof)
    invm = 1.0 / rho[e] / g["detJ"][e]
    exp[e * nl:(e + 1) * nl] = invm * dt * dt / 2 * 0.3 * r0 * phi
    exp[ne * nl + e * nl:ne * nl + (e + 1) * nl] = invm * dt * dt / 2 * (-1.0) * r0 * phi
    np.testing.assert_allclose(u, exp, rtol=1e-14, atol=1e-30)


def test_source_time_is_t_plus_dt():
    """step() evaluates the force at t_{n+1} (TimeAdvance.cpp:48-51): a source
    switched on only at t = 2*dt (the time step 1 uses) moves step 1."""
    op, verts, nb, _ = _operator(N=1)
    e, xi = O.locate(verts, 250.0, 200.0)
    phi = O.dubiner(1, *xi)[0]
    dt = 1e-3
    # narrow Ricker centred at 2*dt: non-negligible at t=2dt, ~0 at t=0
    src = dict(elem=e, phi=phi, f0=2000.0, t0=2 * dt, dirx=0.0, diry=1.0)
    u1, _ = op.run(dt, 0, src=src)
    u2, _ = op.run(dt, 1, src=src)
    assert np.abs(u1).max() < 1e-30 or np.abs(u1).max() < 1e-12 * np.abs(u2).max()
    assert np.abs(u2).max() > 0


def test_ricker_closed_form():
    for f0, t0, t in [(10.0, 0.12, 0.0), (10.0, 0.12, 0.12), (8.0, 0.15, 0.2)]:
        a = (np.pi * f0 * (t - t0)) ** 2
        assert abs(O.ricker(f0, t0, t) - (1 - 2 * a) * np.exp(-a)) <= 1e-15


def test_lambda_max_matches_dense_eigs():
    op, verts, nb, (lam, mu, rho) = _operator(N=2, nx=3, ny=2)
    M = _full(op, nb)
    g = O.geometry(verts)
    minv = np.concatenate([np.repeat(1 / (rho * g["detJ"]), op.nln)] * 2)
    ev = np.abs(np.linalg.eigvals(minv[:, None] * M)).max()
    assert abs(op.lambda_max(2000) - ev) <= 1e-6 * ev


@pytest.mark.parametrize("N", [1, 2, 3])
def test_leapfrog_converges_to_an_exact_standing_wave(N):
    """Physics check of the whole discretisation (assembly, SIPG, inverse mass,
    leap-frog with its Taylor start): with lambda = 0 every boundary of the
    square [0, L]^2 is traction-free for u = (cos(pi x / L) cos(w t), 0),
    w = pi c_p / L, c_p = sqrt(2 mu / rho), an exact solution of
    rho u_tt = div sigma. The oracle's relative L2 error at t = L / c_p
    (half a period) is small and falls by at least 0.8 * 2^(N+1) when h and dt
    halve (order N+1 in space, 2 in time); dt is half the stability limit from
    the power iteration."""
    L, rho, mu = 1000.0, 2000.0, 4e9
    cp = math.sqrt(2 * mu / rho)
    w = math.pi * cp / L
    errs = []
    for n in (6, 12):
        pts, tri = capi.split_square(n, n, 0.0, 0.0, L, L, 0.1, 5)
        verts = pts[tri].reshape(-1, 6)
        nb, ng = O.neighbors(tri)
        ne = tri.shape[0]
        op = O.Operator(N, verts, nb, ng, np.zeros(ne), np.full(ne, mu), np.full(ne, rho), 10.0)
        h = L / n
        T = L / cp
        # half the leap-frog stability limit 2 / sqrt(lambda_max(M^-1 B))
        nsteps = int(math.ceil(T / (1.0 / math.sqrt(op.lambda_max()))))
        dt = T / nsteps
        u0 = capi.project(N, verts, lambda x, y: (np.cos(math.pi * x / L), 0 * x))
        u, _ = op.run(dt, nsteps - 1, u0=u0, v0=np.zeros_like(u0))   # first step + nsteps-1 steps: t = T
        exact = lambda x, y: (np.cos(math.pi * x / L) * math.cos(w * T), 0 * x)
        err = capi.l2_error(N, verts, u, exact) / capi.l2_error(N, verts, np.zeros_like(u), exact)
        errs.append(err)
    assert errs[0] < 0.05 / N
    # measured 3.35, 7.97, 14.7: the spatial order N+1 dominates on these meshes
    assert errs[0] / errs[1] >= 0.8 * 2 ** (N + 1), errs


# ------------------------------------------------------------------ BoundaryAdapted basis
def _ba_modes(N):
    """(kind, edge set the mode is nonzero on) for the reference's mode order."""
    kinds = [("v", None)] * 3 + [("e", "eta0")] * (N - 1) + [("e", "xi0")] * (N - 1) + [("e", "hyp")] * (N - 1)
    kinds += [("b", None)] * ((N - 1) * (N - 2) // 2)
    return kinds


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8])
def test_boundary_adapted_modes_and_traces(N):
    """Vertex modes are the barycentric coordinates; an edge mode vanishes on the
    two other edges and equals xi(1-xi) P_{i-1}^{1,1}(2xi-1) on eta = 0 (resp.
    eta(1-eta) P^{1,1}_{j-1}(2eta-1) on xi = 0, xi eta P^{1,1}_{j-1}(2eta-1) on the
    hypotenuse); bubbles vanish on every edge; the count is (N+1)(N+2)/2."""
    from scipy.special import eval_jacobi
    kinds = _ba_modes(N)
    assert len(kinds) == O.nln(N)
    for xi, eta in [(0.13, 0.21), (0.6, 0.3), (0.05, 0.9)]:
        phi, _ = O.boundary_adapted(N, xi, eta)
        np.testing.assert_allclose(phi[:3], [1 - xi - eta, xi, eta], rtol=0, atol=1e-15)
    t = np.linspace(0.03, 0.97, 7)
    for s in t:
        on = {"eta0": O.boundary_adapted(N, s, 0.0)[0], "xi0": O.boundary_adapted(N, 0.0, s)[0],
              "hyp": O.boundary_adapted(N, s, 1 - s)[0]}
        ie = {"eta0": 0, "xi0": 0, "hyp": 0}
        for k, (kind, edge) in enumerate(kinds):
            if kind == "b":
                for v in on.values():
                    assert abs(v[k]) < 1e-14
            if kind == "e":
                for name, v in on.items():
                    if name != edge:
                        assert abs(v[k]) < 1e-14
                i = ie[edge] = ie[edge] + 1
                if edge == "eta0":
                    expect = s * (1 - s) * eval_jacobi(i - 1, 1, 1, 2 * s - 1)
                elif edge == "xi0":
                    expect = s * (1 - s) * eval_jacobi(i - 1, 1, 1, 2 * s - 1)
                else:
                    expect = s * (1 - s) * eval_jacobi(i - 1, 1, 1, 2 * (1 - s) - 1)
                assert on[edge][k] == pytest.approx(expect, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("N", [2, 4, 7])
def test_boundary_adapted_gradient_and_span(N):
    """The gradient is the derivative (central differences), and the modes span
    the polynomials of degree N: each is an exact combination of the Dubiner
    modes (least-squares residual at 60 points ~ roundoff)."""
    rng = np.random.default_rng(N)
    h = 1e-6
    pts = []
    while len(pts) < 60:
        x, y = rng.uniform(0.02, 0.95, 2)
        if x + y < 0.97:
            pts.append((x, y))
    for x, y in pts[:8]:
        _, g = O.boundary_adapted(N, x, y)
        fx = (O.boundary_adapted(N, x + h, y)[0] - O.boundary_adapted(N, x - h, y)[0]) / (2 * h)
        fy = (O.boundary_adapted(N, x, y + h)[0] - O.boundary_adapted(N, x, y - h)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, 0], fx, rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(g[:, 1], fy, rtol=1e-6, atol=1e-6)
    D = np.array([O.dubiner(N, x, y)[0] for x, y in pts])
    A = np.array([O.boundary_adapted(N, x, y)[0] for x, y in pts])
    T, *_ = np.linalg.lstsq(D, A, rcond=None)
    assert np.abs(D @ T - A).max() < 1e-12


@pytest.mark.parametrize("N", [1, 3, 6])
def test_boundary_adapted_mass_inverse(N):
    """base_invmass inverts the quadrature mass matrix, which is the exact
    integral of phi_i phi_j (the rule is exact to degree 2N+1): checked through
    the Dubiner change of basis, M = T^T T for phi_BA = phi_Dub T."""
    T_ = O.tables(N, O.BOUNDARY_ADAPTED)
    M = (T_["B"] * T_["w2"][:, None]).T @ T_["B"]
    Minv = O.base_invmass(N, O.BOUNDARY_ADAPTED)
    np.testing.assert_allclose(Minv @ M, np.eye(O.nln(N)), rtol=0, atol=1e-9)
    D = O.tables(N)["B"]
    Tc, *_ = np.linalg.lstsq(D, T_["B"], rcond=None)
    np.testing.assert_allclose(M, Tc.T @ Tc, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(O.base_invmass(N), np.eye(O.nln(N)), rtol=0, atol=0)


@pytest.mark.parametrize("N", [2, 4])
def test_boundary_adapted_dynamics_equal_dubiner(N):
    """The discrete dynamics do not depend on the basis: lambda_max(M^-1 B) and
    the seismograms of the same source and receivers agree between the
    BoundaryAdapted and the Dubiner spaces to roundoff."""
    op_d, verts, nb, (lam, mu, rho) = _operator(N=N, nx=4, ny=3)
    ng = O.neighbors(_mesh(4, 3, seed=3)[1])[1]
    op_b = O.Operator(N, verts, nb, ng, lam, mu, rho, 10.0, basis=O.BOUNDARY_ADAPTED)
    assert op_b.lambda_max(400) == pytest.approx(op_d.lambda_max(400), rel=1e-9)
    sx, sy = 230.0, 170.0
    es, xs = O.locate(verts, sx, sy)
    recs = [(60.0, 350.0), (420.0, 90.0), (300.0, 300.0)]
    rloc = [O.locate(verts, x, y) for x, y in recs]
    dt = 0.4 / np.sqrt(op_d.lambda_max(400))
    out = []
    for op, fn in ((op_d, O.dubiner), (op_b, O.boundary_adapted)):
        src = dict(elem=es, phi=fn(N, *xs)[0], f0=1.0 / (8 * dt), t0=10 * dt, dirx=0.3, diry=-1.0)
        rphi = np.array([fn(N, *xi)[0] for _, xi in rloc])
        _, s = op.run(dt, 120, src=src, rec_elem=np.array([e for e, _ in rloc]), rec_phi=rphi, stride=1)
        out.append(s)
    assert np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0]) < 1e-11


def test_openmp_build_is_bitwise_equal():
    """The multi-core oracle build (timed as a labelled CPU baseline) splits rows
    and vector entries over threads only, so its trajectory equals the
    single-threaded build's bit for bit."""
    pts, tri, verts = _mesh(8, 6, seed=5)
    nb, ng = O.neighbors(tri)
    ne = tri.shape[0]
    rng = np.random.default_rng(5)
    lam, mu, rho = rng.uniform(3e9, 9e9, ne), rng.uniform(3e9, 9e9, ne), rng.uniform(1600, 2400, ne)
    e, xi = O.locate(verts, 260.0, 190.0)
    src = dict(elem=e, phi=O.dubiner(3, *xi)[0], f0=12.0, t0=0.05, dirx=0.6, diry=-0.8)
    assert O.threads(omp=False) == 1
    O.threads(4, omp=True)
    assert O.threads(omp=True) == 4
    runs = []
    for omp in (False, True):
        op = O.Operator(3, verts, nb, ng, lam, mu, rho, 10.0, omp=omp)
        dt = 0.5 * 2 / np.sqrt(op.lambda_max(200))
        runs.append(op.run(dt, 40, src=src, rec_elem=np.array([e, 0], np.int32),
                           rec_phi=np.concatenate([src["phi"], O.dubiner(3, 0.3, 0.3)[0]]), stride=1))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])

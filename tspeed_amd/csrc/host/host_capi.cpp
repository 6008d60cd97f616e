// host_capi.cpp -- C exports of the host layer (include/tspeed_host.h).
#include "../../../include/tspeed_host.h"

#include "tspeed.hpp"

#include <cstring>
#include <exception>

using namespace tspeed;

namespace {
bool basis_ok(int basis) { return basis == 0 || basis == 1; }
tspeed::BasisKind kind_of(int basis) { return basis == 1 ? tspeed::BasisKind::BoundaryAdapted : tspeed::BasisKind::Dubiner; }
}  // namespace

extern "C" {

int tsh_gauleg(int n, double a, double b, double *x, double *w)
{
    if (n < 1 || !x || !w) return -1;
    gauleg(n, a, b, x, w);
    return 0;
}

int tsh_dubiner(int N, double xi, double eta, double *phi, double *grad)
{
    if (N < 0) return -1;
    Dubiner(N).eval(xi, eta, phi, grad);
    return 0;
}

int tsh_tables(int N, double *B, double *G, double *Be, double *Ge, double *w2, double *xy2, double *w1, double *x1)
{
    return tsh_tables_basis(N, 0, B, G, Be, Ge, w2, xy2, w1, x1);
}

int tsh_neighbors(int ne, const int *tri, int *neigh, int *neighedge)
{
    try {
        int np = 0;
        for (int q = 0; q < 3 * ne; ++q) np = std::max(np, tri[q] + 1);
        Mesh m(std::vector<double>(2 * (size_t)np, 0.0), std::vector<int>(tri, tri + 3 * (size_t)ne), {});
        std::memcpy(neigh, m.neigh().data(), sizeof(int) * 3 * (size_t)ne);
        std::memcpy(neighedge, m.neighedge().data(), sizeof(int) * 3 * (size_t)ne);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_split_square(int nx, int ny, double x0, double y0, double Lx, double Ly, double jitter, uint64_t seed,
                     double *pts, int *tri)
{
    if (nx < 1 || ny < 1) return -1;
    auto m = Mesh::split_square(nx, ny, x0, y0, Lx, Ly, jitter, seed);
    std::memcpy(pts, m->points().data(), sizeof(double) * m->points().size());
    std::memcpy(tri, m->triangles().data(), sizeof(int) * m->triangles().size());
    return 0;
}

int tsh_msh_counts(const char *path, int *np, int *ntri, int *nbnd)
{
    try {
        Mesh m(path);
        *np = m.np();
        *ntri = m.ne();
        *nbnd = m.nbnd();
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_msh_read(const char *path, double *points, int *tri, int *region, int *bnd)
{
    try {
        Mesh m(path);
        std::memcpy(points, m.points().data(), sizeof(double) * m.points().size());
        std::memcpy(tri, m.triangles().data(), sizeof(int) * m.triangles().size());
        std::memcpy(region, m.regions().data(), sizeof(int) * m.regions().size());
        if (bnd) std::memcpy(bnd, m.boundary_lines().data(), sizeof(int) * m.boundary_lines().size());
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_locate(int np, const double *points, int ne, const int *tri, int nq, const double *qx, const double *qy,
               int N, int *elem, double *xi, double *phi)
{
    try {
        Mesh m(std::vector<double>(points, points + 2 * (size_t)np), std::vector<int>(tri, tri + 3 * (size_t)ne),
               {});
        const int nln = N > 0 ? (N + 1) * (N + 2) / 2 : 0;
        Dubiner sh(N > 0 ? N : 0);
        for (int q = 0; q < nq; ++q) {
            double x[2] = {0, 0};
            elem[q] = m.locate(Geo::Point(qx[q], qy[q]), x);
            if (xi) {
                xi[2 * q] = x[0];
                xi[2 * q + 1] = x[1];
            }
            if (phi && nln) {
                if (elem[q] >= 0) sh.eval(x[0], x[1], phi + (size_t)q * nln, nullptr);
                else std::memset(phi + (size_t)q * nln, 0, sizeof(double) * nln);
            }
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_ricker_table(int nsrc, const double *f0, const double *t0, const double *dirx, const double *diry, double dt,
                     int nrows, double *out)
{
    if (nsrc < 0 || nrows < 0 || !(dt > 0)) return -1;
    double t = 0;
    for (int r = 0; r < nrows; ++r) {
        if (r == 1) t = dt + dt;
        else if (r > 1) t = t + dt;
        for (int s = 0; s < nsrc; ++s) {
            const double w = ricker(f0[s], t0[s], t);
            out[((size_t)r * nsrc + s) * 2] = dirx[s] * w;
            out[((size_t)r * nsrc + s) * 2 + 1] = diry[s] * w;
        }
    }
    return 0;
}


int tsh_field_points(int ne, const double *verts, double *out)
{
    if (ne < 0 || (ne && (!verts || !out))) return -1;
    tspeed::field_points(ne, verts, out);
    return 0;
}

int tsh_field_values(int N, int basis, int ne, const double *uh, double *out)
{
    if (N < 1 || N > 8 || !basis_ok(basis) || ne < 0 || (ne && (!uh || !out))) return -1;
    tspeed::field_values(N, ne, uh, out, kind_of(basis));
    return 0;
}

int tsh_tables_basis(int N, int basis, double *B, double *G, double *Be, double *Ge, double *w2, double *xy2,
                     double *w1, double *x1)
{
    if (N < 1 || N > 8 || !basis_ok(basis)) return -1;
    try {
        FESpaceRT X(nullptr, N, kind_of(basis));
        const int nq1 = N + 1, nq2 = nq1 * nq1;
        std::memcpy(B, X.B().data(), sizeof(double) * X.B().size());
        std::memcpy(G, X.G().data(), sizeof(double) * X.G().size());
        std::memcpy(Be, X.Bedge().data(), sizeof(double) * X.Bedge().size());
        std::memcpy(Ge, X.Gedge().data(), sizeof(double) * X.Gedge().size());
        for (int k = 0; k < nq2; ++k) {
            w2[k] = X.quad().iweight(k);
            if (xy2) {
                const auto p = X.quad().inode(k);
                xy2[2 * k] = p[0];
                xy2[2 * k + 1] = p[1];
            }
        }
        for (int k = 0; k < nq1; ++k) {
            w1[k] = X.quad().eweight(k);
            if (x1) x1[k] = X.quad().x1()[k];
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_base_invmass(int N, int basis, double *out)
{
    if (N < 1 || N > 8 || !basis_ok(basis) || !out) return -1;
    try {
        const auto m = tspeed::base_invmass(kind_of(basis), N);
        std::copy(m.begin(), m.end(), out);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int tsh_basis_eval(int N, int basis, int n, const double *xi, double *phi, double *grad)
{
    if (N < 0 || N > 8 || !basis_ok(basis) || n < 0 || (n && (!xi || !phi))) return -1;
    const int nln = (N + 1) * (N + 2) / 2;
    for (int q = 0; q < n; ++q)
        tspeed::eval_basis(kind_of(basis), N, xi[2 * q], xi[2 * q + 1], phi + (size_t)q * nln,
                           grad ? grad + (size_t)q * 2 * nln : nullptr);
    return 0;
}

int tsh_quad_points(int N, int ne, const double *verts, double *out)
{
    if (N < 1 || N > 8 || ne < 0 || (ne && (!verts || !out))) return -1;
    tspeed::quad_points(N, ne, verts, out);
    return 0;
}

int tsh_project(int N, int basis, int ne, const double *verts, const double *f, double *uh)
{
    if (N < 1 || N > 8 || !basis_ok(basis) || ne < 0 || (ne && (!verts || !f || !uh))) return -1;
    tspeed::project_nodes(N, ne, verts, f, uh, kind_of(basis));
    return 0;
}

int tsh_l2_error(int N, int basis, int ne, const double *verts, const double *uh, const double *uex, double *err)
{
    if (N < 1 || N > 8 || !basis_ok(basis) || ne < 0 || !err || (ne && (!verts || !uh || !uex))) return -1;
    *err = tspeed::l2_error(N, ne, verts, uh, uex, kind_of(basis));
    return 0;
}

int tsh_write_receivers(const char *fn, int nrec, int nsamples, const double *seis, int digits)
{
    if (!fn || nrec < 0 || nsamples < 0 || (nrec && nsamples && !seis)) return -1;
    try {
        Receivers::write(fn, nrec, nsamples, std::vector<double>(seis, seis + (size_t)nrec * nsamples * 2), digits);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

}  // extern "C"

// tspeed.cpp -- host C++ layer: quadrature, Dubiner basis, FESpace tables, mesh,
// parameters, point entities and the LeapFrog facade over the device engine.
#include "tspeed.hpp"

#include "../../../include/tspeed_amd.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace tspeed {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

// Newton iteration on P_n from the Chebyshev-like first guess; nodes ascend.
void gauleg(int n, double a, double b, double *x, double *w)
{
    const double half = (n + 1.0) / 2.0;
    const double xm = 0.5 * (b + a), xl = 0.5 * (b - a);
    for (int i = 1; i <= half; ++i) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
        int it = 0;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > DBL_EPSILON && ++it < 200);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

Gauss::Gauss(int nq) : nq_(nq), w2_(nq * nq), xy2_(2 * nq * nq), w1_(nq), x1_(nq)
{
    std::vector<double> x(nq), w(nq);
    gauleg(nq, 0.0, 1.0, x1_.data(), w1_.data());
    gauleg(nq, -1.0, 1.0, x.data(), w.data());
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            w2_[i * nq + j] = (1 - x[i]) * w[i] * w[j] / 8.;
            xy2_[2 * (i * nq + j)] = (1 + x[i]) / 2.;
            xy2_[2 * (i * nq + j) + 1] = (1 - x[i]) * (1 + x[j]) / 4.;
        }
}

std::array<double, 2> Gauss::edge_node(int f, int k) const
{
    const double s = x1_[k];
    if (f == 0) return {1.0 - s, s};
    if (f == 1) return {0.0, 1.0 - s};
    return {s, 0.0};
}

namespace {
// A value with its two partial derivatives (d/dxi, d/deta): forward-mode
// differentiation of the basis expressions.
struct Dual {
    double v = 0, dx = 0, dy = 0;
    Dual() = default;
    Dual(double c) : v(c) {}
    Dual(double a, double b, double c) : v(a), dx(b), dy(c) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.dx + b.dx, a.dy + b.dy}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.dx - b.dx, a.dy - b.dy}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy}; }
inline Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return {q, (a.dx - q * b.dx) / b.v, (a.dy - q * b.dy) / b.v};
}
inline Dual powi(Dual a, int n)
{
    Dual r(1.0);
    for (int k = 0; k < n; ++k) r = r * a;
    return r;
}

// jacobi_polynomial's recurrence on any arithmetic type
template <typename T>
T jacobi_t(int n, int alpha, int beta, T z)
{
    if (n == 0) return T(1.0);
    const T first = T(0.5) * (T(double(alpha - beta)) + T(alpha + beta + 2.0) * z);
    if (n == 1) return first;
    const int apb = alpha + beta;
    T poly(0.0), pm2(1.0), pm1 = first;
    for (int k = 2; k <= n; ++k) {
        const double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 /= a1;
        a3 /= a1;
        a4 /= a1;
        poly = (T(a2) + T(a3) * z) * pm1 - T(a4) * pm2;
        pm2 = pm1;
        pm1 = poly;
    }
    return poly;
}
}  // namespace

void BoundaryAdaptedBasis::eval(double xi, double eta, double *phi, double *grad) const
{
    const Dual x(xi, 1, 0), y(eta, 0, 1);
    const Dual l0 = Dual(1.0) - x - y, om = Dual(1.0) - y;
    const Dual zx = Dual(2.0) * x / om - Dual(1.0), zy = Dual(2.0) * y - Dual(1.0);
    std::vector<Dual> f;
    f.reserve(gdl());
    f.push_back(l0);  // vertices
    f.push_back(x);
    f.push_back(y);
    for (int i = 1; i < N_; ++i) f.push_back(l0 * x * powi(om, i - 1) * jacobi_t(i - 1, 1, 1, zx));  // eta = 0
    for (int j = 1; j < N_; ++j) f.push_back(l0 * y * jacobi_t(j - 1, 1, 1, zy));                   // xi = 0
    for (int j = 1; j < N_; ++j) f.push_back(x * y * jacobi_t(j - 1, 1, 1, zy));                    // xi + eta = 1
    for (int j = 1; j < N_; ++j)
        for (int i = 1; i < N_; ++i)
            if (i + j < N_)
                f.push_back(l0 * x * y * powi(om, i - 1) * jacobi_t(i - 1, 1, 1, zx) *
                            jacobi_t(j - 1, 2 * i + 1, 1, zy));
    for (size_t k = 0; k < f.size(); ++k) {
        if (phi) phi[k] = f[k].v;
        if (grad) {
            grad[2 * k] = f[k].dx;
            grad[2 * k + 1] = f[k].dy;
        }
    }
}

void eval_basis(BasisKind kind, int N, double xi, double eta, double *phi, double *grad)
{
    if (kind == BasisKind::BoundaryAdapted) BoundaryAdaptedBasis(N).eval(xi, eta, phi, grad);
    else Dubiner(N).eval(xi, eta, phi, grad);
}

std::vector<double> base_invmass(BasisKind kind, int N)
{
    const int n = (N + 1) * (N + 2) / 2;
    std::vector<double> inv((size_t)n * n, 0.0);
    if (kind == BasisKind::Dubiner) {
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
        return inv;
    }
    const Gauss q(N + 1);
    std::vector<double> phi(n), M((size_t)n * n, 0.0);
    for (int k = 0; k < q.nqn2d(); ++k) {
        const auto xy = q.inode(k);
        eval_basis(kind, N, xy[0], xy[1], phi.data(), nullptr);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) M[(size_t)i * n + j] += phi[i] * phi[j] * q.iweight(k);
    }
    // Gauss-Jordan elimination with partial pivoting on [M | I]
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)p * n + c])) p = r;
        if (M[(size_t)p * n + c] == 0.0) throw std::runtime_error("base_invmass: singular mass matrix");
        for (int j = 0; j < n; ++j) {
            std::swap(M[(size_t)c * n + j], M[(size_t)p * n + j]);
            std::swap(inv[(size_t)c * n + j], inv[(size_t)p * n + j]);
        }
        const double piv = M[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            M[(size_t)c * n + j] /= piv;
            inv[(size_t)c * n + j] /= piv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = M[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                M[(size_t)r * n + j] -= f * M[(size_t)c * n + j];
                inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
            }
        }
    }
    return inv;
}

double jacobi_polynomial(int n, int alpha, int beta, double z)
{
    if (n == 0) return 1.0;
    if (n == 1) return 0.5 * (alpha - beta + (alpha + beta + 2.0) * z);
    const int apb = alpha + beta;
    double poly = 0.0, pm2 = 1.0, pm1 = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z);
    for (int k = 2; k <= n; ++k) {
        const double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 /= a1;
        a3 /= a1;
        a4 /= a1;
        poly = (a2 + a3 * z) * pm1 - a4 * pm2;
        pm2 = pm1;
        pm1 = poly;
    }
    return poly;
}

// phi_ij = c_ij 2^i (1-eta)^i P_i^{0,0}(2xi/(1-eta)-1) P_j^{2i+1,0}(2eta-1),
// c_ij = sqrt((2i+1) 2(i+j+1) / 4^i), modes ordered j-major; gradients are the
// analytic derivatives (ShapeFunctions_imp.hpp:29-51).
void Dubiner::eval(double csi, double eta, double *phi, double *grad) const
{
    int s = 0;
    for (int j = 0; j <= N_; ++j)
        for (int i = 0; i <= N_ - j; ++i, ++s) {
            const double cij = std::sqrt((2. * i + 1) * 2. * (i + j + 1) / std::pow(4., i));
            const double a = 2 * csi / (1 - eta) - 1, b = 2 * eta - 1;
            const double Pi = jacobi_polynomial(i, 0, 0, a), Pj = jacobi_polynomial(j, 2 * i + 1, 0, b);
            if (phi) phi[s] = cij * std::pow(2., i) * std::pow(1 - eta, i) * Pi * Pj;
            if (!grad) continue;
            double gx = 0.0, gy = 0.0;
            if (i == 0 && j != 0) {
                gy = cij * (j + 2) * jacobi_polynomial(j - 1, 2, 1, b);
            } else if (i != 0 && j == 0) {
                const double Q = jacobi_polynomial(i - 1, 1, 1, a);
                gx = cij * std::pow(2., i) * std::pow(1 - eta, i - 1) * (i + 1) * Q;
                gy = cij * std::pow(2., i) *
                     (-i * std::pow(1 - eta, i - 1) * Pi + csi * std::pow(1 - eta, i - 2) * (i + 1) * Q);
            } else if (i != 0) {
                const double Q = jacobi_polynomial(i - 1, 1, 1, a), R = jacobi_polynomial(j - 1, 2 * i + 2, 1, b);
                gx = cij * std::pow(2., i) * std::pow(1 - eta, i - 1) * (i + 1) * Q * Pj;
                gy = cij * std::pow(2., i) *
                     (-i * std::pow(1 - eta, i - 1) * Pi * Pj + csi * std::pow(1 - eta, i - 2) * (i + 1) * Q * Pj +
                      std::pow(1 - eta, i) * (2 * i + j + 2) * Pi * R);
            }
            grad[2 * s] = gx;
            grad[2 * s + 1] = gy;
        }
}

double Dubiner::phi(int s, double xi, double eta) const
{
    std::vector<double> v(gdl());
    eval(xi, eta, v.data(), nullptr);
    return v[s];
}

// ---------------------------------------------------------------- Mesh
namespace {
void skip_to(std::istream &f, const std::string &s)
{
    std::string line;
    do {
        std::getline(f, line);
    } while (line.find(s) == std::string::npos && !f.eof());
}

uint64_t splitmix64(uint64_t &st)
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

Mesh::Mesh(const std::string &fileName) : fname_(fileName)
{
    std::ifstream f(fileName);
    if (!f.is_open()) throw std::runtime_error("Mesh file does not exist or is corrupted: " + fileName);
    skip_to(f, "$Nodes");
    if (f.eof() || f.fail()) throw std::runtime_error("FILE ERROR! Cannot find Nodes in " + fileName);
    int nP = 0;
    f >> nP;
    pts_.resize(2 * (size_t)nP);
    for (int i = 0; i < nP; ++i) {
        int id;
        double x, y;
        f >> id >> x >> y;
        f.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        pts_[2 * i] = x;
        pts_[2 * i + 1] = y;
    }
    skip_to(f, "$Elements");
    int nE = 0;
    f >> nE;
    for (int i = 0; i < nE; ++i) {
        int dummy, el_type, num_tags, phys, geom, p1, p2, p3;
        f >> dummy >> el_type >> num_tags >> phys >> geom;
        for (int j = 0; j < num_tags - 2; ++j) f >> dummy;
        if (el_type == 2) {
            f >> p1 >> p2 >> p3;
            tri_.insert(tri_.end(), {p1 - 1, p2 - 1, p3 - 1});
            reg_.push_back(geom);
        } else if (el_type == 1) {
            f >> p1 >> p2;
            bnd_.insert(bnd_.end(), {p1 - 1, p2 - 1});
        }
        f.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    }
    if (f.fail()) throw std::runtime_error("malformed $Elements section in " + fileName);
    build_neighbors();
}

Mesh::Mesh(std::vector<double> points, std::vector<int> tri, std::vector<int> region)
    : pts_(std::move(points)), tri_(std::move(tri)), reg_(std::move(region))
{
    if (reg_.empty()) reg_.assign(tri_.size() / 3, 0);
    build_neighbors();
}

std::shared_ptr<Mesh> Mesh::split_square(int nx, int ny, double x0, double y0, double Lx, double Ly, double jitter,
                                         uint64_t seed)
{
    const double hx = Lx / nx, hy = Ly / ny;
    std::vector<double> pts(2 * (size_t)(nx + 1) * (ny + 1));
    uint64_t st = seed;
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            double x = x0 + i * hx, y = y0 + j * hy;
            if (i > 0 && i < nx && j > 0 && j < ny && jitter > 0) {
                const double ux = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
                const double uy = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
                x += (2 * ux - 1) * jitter * hx;
                y += (2 * uy - 1) * jitter * hy;
            }
            pts[2 * ((size_t)j * (nx + 1) + i)] = x;
            pts[2 * ((size_t)j * (nx + 1) + i) + 1] = y;
        }
    std::vector<int> tri;
    tri.reserve(6 * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int a = j * (nx + 1) + i, b = a + 1, c = a + nx + 2, d = a + nx + 1;
            tri.insert(tri.end(), {a, b, c, a, c, d});
        }
    return std::make_shared<Mesh>(std::move(pts), std::move(tri), std::vector<int>());
}

// Neighbours across edges that appear reversed in the two triangles (the
// condition of Mesh.cpp:49-111), found through a hash of the undirected edge.
void Mesh::build_neighbors()
{
    const int n = ne();
    neigh_.assign(3 * (size_t)n, -1);
    neighedge_.assign(3 * (size_t)n, -1);
    std::unordered_map<uint64_t, int> open;  // undirected edge -> half-face e*3+f
    open.reserve(3 * (size_t)n);
    for (int e = 0; e < n; ++e)
        for (int f = 0; f < 3; ++f) {
            const int va = tri_[3 * e + (f + 1) % 3], vb = tri_[3 * e + (f + 2) % 3];
            const uint64_t key = ((uint64_t)(uint32_t)std::min(va, vb) << 32) | (uint32_t)std::max(va, vb);
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, 3 * e + f);
                continue;
            }
            const int h = it->second, e2 = h / 3, f2 = h % 3;
            if (h < 0) throw std::runtime_error("non-manifold edge in mesh");
            const int wa = tri_[3 * e2 + (f2 + 1) % 3], wb = tri_[3 * e2 + (f2 + 2) % 3];
            if (va == wb && vb == wa) {
                neigh_[3 * e + f] = e2;
                neighedge_[3 * e + f] = f2;
                neigh_[3 * e2 + f2] = e;
                neighedge_[3 * e2 + f2] = f;
            }
            it->second = -1;  // closed
        }
}

std::vector<double> Mesh::vertex_array() const
{
    std::vector<double> v(6 * (size_t)ne());
    for (int e = 0; e < ne(); ++e)
        for (int k = 0; k < 3; ++k) {
            v[6 * (size_t)e + 2 * k] = pts_[2 * (size_t)tri_[3 * e + k]];
            v[6 * (size_t)e + 2 * k + 1] = pts_[2 * (size_t)tri_[3 * e + k] + 1];
        }
    return v;
}

double Mesh::signed_area(int e) const
{
    const double *a = &pts_[2 * (size_t)tri_[3 * e]], *b = &pts_[2 * (size_t)tri_[3 * e + 1]],
                 *c = &pts_[2 * (size_t)tri_[3 * e + 2]];
    return 0.5 * ((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]));
}

double Mesh::min_edge() const
{
    double m = std::numeric_limits<double>::max();
    for (int e = 0; e < ne(); ++e)
        for (int f = 0; f < 3; ++f) {
            const double *a = &pts_[2 * (size_t)tri_[3 * e + f]], *b = &pts_[2 * (size_t)tri_[3 * e + (f + 1) % 3]];
            m = std::min(m, std::hypot(a[0] - b[0], a[1] - b[1]));
        }
    return m;
}

void Mesh::build_grid() const
{
    if (gnx_) return;
    const int n = ne();
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (size_t q = 0; q < pts_.size(); q += 2) {
        xmin = std::min(xmin, pts_[q]);
        xmax = std::max(xmax, pts_[q]);
        ymin = std::min(ymin, pts_[q + 1]);
        ymax = std::max(ymax, pts_[q + 1]);
    }
    const double w = std::max(xmax - xmin, 1e-300), h = std::max(ymax - ymin, 1e-300);
    const double cells = std::max(1.0, (double)n);
    gnx_ = std::max(1, (int)std::sqrt(cells * w / h));
    gny_ = std::max(1, (int)(cells / gnx_));
    const double pad = 1e-9 * std::max(w, h);
    gx0_ = xmin - pad;
    gy0_ = ymin - pad;
    gdx_ = (w + 2 * pad) / gnx_;
    gdy_ = (h + 2 * pad) / gny_;
    std::vector<std::vector<int>> cell((size_t)gnx_ * gny_);
    for (int e = 0; e < n; ++e) {
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int k = 0; k < 3; ++k) {
            const double *p = &pts_[2 * (size_t)tri_[3 * e + k]];
            x0 = std::min(x0, p[0]); x1 = std::max(x1, p[0]);
            y0 = std::min(y0, p[1]); y1 = std::max(y1, p[1]);
        }
        const int i0 = std::max(0, (int)std::floor((x0 - pad - gx0_) / gdx_));
        const int i1 = std::min(gnx_ - 1, (int)std::floor((x1 + pad - gx0_) / gdx_));
        const int j0 = std::max(0, (int)std::floor((y0 - pad - gy0_) / gdy_));
        const int j1 = std::min(gny_ - 1, (int)std::floor((y1 + pad - gy0_) / gdy_));
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) cell[(size_t)j * gnx_ + i].push_back(e);
    }
    cell_start_.assign(cell.size() + 1, 0);
    for (size_t c = 0; c < cell.size(); ++c) cell_start_[c + 1] = cell_start_[c] + (int)cell[c].size();
    cell_items_.clear();
    cell_items_.reserve(cell_start_.back());
    for (auto &v : cell) cell_items_.insert(cell_items_.end(), v.begin(), v.end());
}

int Mesh::locate(const Geo::Point &p, double *xi) const
{
    build_grid();
    const int i = (int)std::floor((p.x - gx0_) / gdx_), j = (int)std::floor((p.y - gy0_) / gdy_);
    if (i < 0 || i >= gnx_ || j < 0 || j >= gny_) return -1;
    int best = -1;
    const size_t c = (size_t)j * gnx_ + i;
    for (int q = cell_start_[c]; q < cell_start_[c + 1]; ++q) {
        const int e = cell_items_[q];
        if (best >= 0 && e >= best) continue;
        const double *P0 = &pts_[2 * (size_t)tri_[3 * e]], *P1 = &pts_[2 * (size_t)tri_[3 * e + 1]],
                     *P2 = &pts_[2 * (size_t)tri_[3 * e + 2]];
        const double x0 = P0[0], y0 = P0[1], x1 = P1[0], y1 = P1[1], x2 = P2[0], y2 = P2[1];
        // tolerant point-in-triangle test (Geometry.cpp:112-118)
        const double s = y0 * x2 - x0 * y2 + (y2 - y0) * p.x + (x0 - x2) * p.y;
        const double t = x0 * y1 - y0 * x1 + (y0 - y1) * p.x + (x1 - x0) * p.y;
        const double A = 1. / 2. * (-y1 * x2 + y0 * (-x1 + x2) + x0 * (y1 - y2) + x1 * y2);
        if (s > -1e-6 && t > -1e-6 && (s + t) < (2 + 1e-6) * A) best = e;
    }
    if (best >= 0 && xi) {
        const double *P0 = &pts_[2 * (size_t)tri_[3 * best]], *P1 = &pts_[2 * (size_t)tri_[3 * best + 1]],
                     *P2 = &pts_[2 * (size_t)tri_[3 * best + 2]];
        const double J00 = P1[0] - P0[0], J10 = P1[1] - P0[1], J01 = P2[0] - P0[0], J11 = P2[1] - P0[1];
        const double invdet = 1.0 / (J00 * J11 - J10 * J01);
        const double i00 = J11 * invdet, i10 = -J10 * invdet, i01 = -J01 * invdet, i11 = J00 * invdet;
        const double dx = p.x - P0[0], dy = p.y - P0[1];
        xi[0] = i00 * dx + i01 * dy;
        xi[1] = i10 * dx + i11 * dy;
    }
    return best;
}

// ---------------------------------------------------------------- FESpace
FESpaceRT::FESpaceRT(Mesh_ptr m, int N, BasisKind kind)
    : mesh_(std::move(m)), N_(N), nln_((N + 1) * (N + 2) / 2), kind_(kind), quad_(N + 1),
      minv_(tspeed::base_invmass(kind, N))
{
    const int nq1 = N + 1, nq2 = nq1 * nq1;
    B_.resize((size_t)nq2 * nln_);
    G_.resize((size_t)nln_ * 2 * nq2);
    Be_.resize((size_t)3 * nq1 * nln_);
    Ge_.resize((size_t)nln_ * 3 * 2 * nq1);
    std::vector<double> phi(nln_), gr(2 * nln_);
    for (int k = 0; k < nq2; ++k) {
        const auto xy = quad_.inode(k);
        basis(xy[0], xy[1], phi.data(), gr.data());
        for (int i = 0; i < nln_; ++i) {
            B_[(size_t)k * nln_ + i] = phi[i];
            G_[((size_t)i * 2) * nq2 + k] = gr[2 * i];
            G_[((size_t)i * 2 + 1) * nq2 + k] = gr[2 * i + 1];
        }
    }
    for (int f = 0; f < 3; ++f)
        for (int k = 0; k < nq1; ++k) {
            const auto xy = quad_.edge_node(f, k);
            basis(xy[0], xy[1], phi.data(), gr.data());
            for (int i = 0; i < nln_; ++i) {
                Be_[((size_t)f * nq1 + k) * nln_ + i] = phi[i];
                Ge_[(((size_t)i * 3 + f) * 2) * nq1 + k] = gr[2 * i];
                Ge_[(((size_t)i * 3 + f) * 2 + 1) * nq1 + k] = gr[2 * i + 1];
            }
        }
}

std::vector<double> FESpaceRT::w2() const
{
    std::vector<double> w(quad_.nqn2d());
    for (int k = 0; k < quad_.nqn2d(); ++k) w[k] = quad_.iweight(k);
    return w;
}
std::vector<double> FESpaceRT::w1() const
{
    std::vector<double> w(quad_.nqn1d());
    for (int k = 0; k < quad_.nqn1d(); ++k) w[k] = quad_.eweight(k);
    return w;
}

// uh_i = (1/detJ) * base_invmass * int_K f phi_i, base_invmass = I (Dubiner)
std::vector<double> FESpaceRT::transform(const std::function<std::array<double, 2>(double, double)> &fun) const
{
    const int n = ne(), nq2 = quad_.nqn2d();
    std::vector<double> uh(2 * (size_t)n * nln_, 0.0);
    const auto v = mesh_->vertex_array();
    for (int e = 0; e < n; ++e) {
        const double *P = &v[6 * (size_t)e];
        const double J00 = P[2] - P[0], J10 = P[3] - P[1], J01 = P[4] - P[0], J11 = P[5] - P[1];
        const double det = J00 * J11 - J10 * J01;
        std::vector<double> rhs(2 * nln_, 0.0);
        for (int k = 0; k < nq2; ++k) {
            const double dx = std::fabs(det) * quad_.iweight(k);
            const auto xy = quad_.inode(k);
            const double x = P[0] + J00 * xy[0] + J01 * xy[1], y = P[1] + J10 * xy[0] + J11 * xy[1];
            const auto F = fun(x, y);
            for (int i = 0; i < nln_; ++i) {
                rhs[i] += F[0] * dx * B_[(size_t)k * nln_ + i];
                rhs[nln_ + i] += F[1] * dx * B_[(size_t)k * nln_ + i];
            }
        }
        // uh = 1/detJ * base_invmass * rhs (FESpace_imp.hpp:151-152)
        for (int i = 0; i < nln_; ++i) {
            double sx = 0, sy = 0;
            for (int j = 0; j < nln_; ++j) {
                sx += minv_[(size_t)i * nln_ + j] * rhs[j];
                sy += minv_[(size_t)i * nln_ + j] * rhs[nln_ + j];
            }
            uh[(size_t)e * nln_ + i] = 1. / det * sx;
            uh[(size_t)n * nln_ + (size_t)e * nln_ + i] = 1. / det * sy;
        }
    }
    return uh;
}

// ---------------------------------------------------------------- plot output and error norm
const double kPlotPoints[4][2] = {{0.2, 0.2}, {0.8, 0.2}, {0.2, 0.8}, {0.5, 0.5}};

namespace {
// affine map of the reference triangle onto element e (Geometry.cpp:199-207)
inline void map_point(const double *P, double xi, double eta, double *x, double *y)
{
    *x = P[0] + (P[2] - P[0]) * xi + (P[4] - P[0]) * eta;
    *y = P[1] + (P[3] - P[1]) * xi + (P[5] - P[1]) * eta;
}
inline double det_of(const double *P)
{
    return (P[2] - P[0]) * (P[5] - P[1]) - (P[3] - P[1]) * (P[4] - P[0]);
}
}  // namespace

void field_points(int ne, const double *verts, double *out)
{
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 4; ++k)
            map_point(verts + 6 * (size_t)e, kPlotPoints[k][0], kPlotPoints[k][1], &out[((size_t)e * 4 + k) * 2],
                      &out[((size_t)e * 4 + k) * 2 + 1]);
}

void field_values(int N, int ne, const double *uh, double *out, BasisKind kind)
{
    const int nln = (N + 1) * (N + 2) / 2;
    std::vector<double> vb(4 * (size_t)nln);  // basis at the plot points
    for (int k = 0; k < 4; ++k) eval_basis(kind, N, kPlotPoints[k][0], kPlotPoints[k][1], &vb[(size_t)k * nln], nullptr);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 4; ++k) {
            double vx = 0, vy = 0;
            for (int i = 0; i < nln; ++i) {
                vx += vb[(size_t)k * nln + i] * uh[(size_t)e * nln + i];
                vy += vb[(size_t)k * nln + i] * uh[(size_t)ne * nln + (size_t)e * nln + i];
            }
            out[((size_t)e * 4 + k) * 2] = vx;
            out[((size_t)e * 4 + k) * 2 + 1] = vy;
        }
}

void quad_points(int N, int ne, const double *verts, double *out)
{
    const Gauss q(N + 1);
    const int nq2 = q.nqn2d();
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < nq2; ++k) {
            const auto xy = q.inode(k);
            map_point(verts + 6 * (size_t)e, xy[0], xy[1], &out[((size_t)e * nq2 + k) * 2],
                      &out[((size_t)e * nq2 + k) * 2 + 1]);
        }
}

void project_nodes(int N, int ne, const double *verts, const double *f, double *uh, BasisKind kind)
{
    const Gauss q(N + 1);
    const int nq2 = q.nqn2d(), nln = (N + 1) * (N + 2) / 2;
    const auto minv = tspeed::base_invmass(kind, N);
    std::vector<double> Bq((size_t)nq2 * nln);
    for (int k = 0; k < nq2; ++k) {
        const auto xy = q.inode(k);
        eval_basis(kind, N, xy[0], xy[1], &Bq[(size_t)k * nln], nullptr);
    }
    for (int e = 0; e < ne; ++e) {
        const double det = det_of(verts + 6 * (size_t)e);
        std::vector<double> rhs(2 * (size_t)nln, 0.0);
        for (int k = 0; k < nq2; ++k) {
            const double dx = std::fabs(det) * q.iweight(k);
            const double fx = f[((size_t)e * nq2 + k) * 2], fy = f[((size_t)e * nq2 + k) * 2 + 1];
            for (int i = 0; i < nln; ++i) {
                rhs[i] += fx * dx * Bq[(size_t)k * nln + i];
                rhs[nln + i] += fy * dx * Bq[(size_t)k * nln + i];
            }
        }
        for (int i = 0; i < nln; ++i) {
            double sx = 0, sy = 0;
            for (int j = 0; j < nln; ++j) {
                sx += minv[(size_t)i * nln + j] * rhs[j];
                sy += minv[(size_t)i * nln + j] * rhs[nln + j];
            }
            uh[(size_t)e * nln + i] = 1. / det * sx;
            uh[(size_t)ne * nln + (size_t)e * nln + i] = 1. / det * sy;
        }
    }
}

double l2_error(int N, int ne, const double *verts, const double *uh, const double *uex, BasisKind kind)
{
    const Gauss q(N + 1);
    const int nq2 = q.nqn2d(), nln = (N + 1) * (N + 2) / 2;
    std::vector<double> Bq((size_t)nq2 * nln);
    for (int k = 0; k < nq2; ++k) {
        const auto xy = q.inode(k);
        eval_basis(kind, N, xy[0], xy[1], &Bq[(size_t)k * nln], nullptr);
    }
    double error = 0;
    for (int e = 0; e < ne; ++e) {
        double error_loc = 0;
        for (int k = 0; k < nq2; ++k) {
            double ax = 0, ay = 0;
            for (int i = 0; i < nln; ++i) {
                ax += Bq[(size_t)k * nln + i] * uh[(size_t)e * nln + i];
                ay += Bq[(size_t)k * nln + i] * uh[(size_t)ne * nln + (size_t)e * nln + i];
            }
            const double dx = ax - uex[((size_t)e * nq2 + k) * 2], dy = ay - uex[((size_t)e * nq2 + k) * 2 + 1];
            error_loc += (dx * dx + dy * dy) * q.iweight(k);
        }
        error += error_loc * det_of(verts + 6 * (size_t)e);
    }
    return std::sqrt(error);
}

void FESpaceRT::points_out(const std::string &fname) const
{
    const auto v = mesh_->vertex_array();
    std::vector<double> pts(8 * (size_t)ne());
    field_points(ne(), v.data(), pts.data());
    std::ofstream out(fname);
    if (!out.is_open()) throw std::runtime_error("File " + fname + " may not be opened");
    for (size_t k = 0; k < 4 * (size_t)ne(); ++k) out << pts[2 * k] << " " << pts[2 * k + 1] << std::endl;
}

void FESpaceRT::field_out(const std::string &fname, const std::vector<double> &uh, unsigned step) const
{
    if (uh.size() != 2 * (size_t)ne() * nln_) throw std::invalid_argument("field_out: uh has the wrong size");
    std::vector<double> val(8 * (size_t)ne());
    field_values(N_, ne(), uh.data(), val.data(), kind_);
    const std::string name = fname + "u_" + std::to_string(step) + ".field";
    std::ofstream out(name);
    if (!out.is_open()) throw std::runtime_error("File " + name + " may not be opened");
    for (size_t k = 0; k < 4 * (size_t)ne(); ++k) out << val[2 * k] << " " << val[2 * k + 1] << std::endl;
}

double FESpaceRT::L2error(const std::function<std::array<double, 2>(double, double)> &uex,
                          const std::vector<double> &uh) const
{
    if (uh.size() != 2 * (size_t)ne() * nln_) throw std::invalid_argument("L2error: uh has the wrong size");
    const auto v = mesh_->vertex_array();
    const int nq2 = quad_.nqn2d();
    std::vector<double> xy(2 * (size_t)ne() * nq2), f(2 * (size_t)ne() * nq2);
    quad_points(N_, ne(), v.data(), xy.data());
    for (size_t k = 0; k < (size_t)ne() * nq2; ++k) {
        const auto F = uex(xy[2 * k], xy[2 * k + 1]);
        f[2 * k] = F[0];
        f[2 * k + 1] = F[1];
    }
    return l2_error(N_, ne(), v.data(), uh.data(), f.data(), kind_);
}

// ---------------------------------------------------------------- Parameters
Parameters::Parameters(Mesh_ptr m) : mesh_(std::move(m))
{
    const int n = mesh_->ne();
    lambda_.assign(n, 0.0);
    mu_.assign(n, 0.0);
    rho_.assign(n, 0.0);
}
std::vector<double> &Parameters::ref(const std::string &p)
{
    if (p == "lambda") return lambda_;
    if (p == "mu") return mu_;
    if (p == "rho") return rho_;
    throw std::invalid_argument("unknown parameter " + p);
}
const std::vector<double> &Parameters::ref(const std::string &p) const
{
    return const_cast<Parameters *>(this)->ref(p);
}
int Parameters::setp(const std::string &p, int lab, double value)
{
    auto &v = ref(p);
    int count = 0;
    for (int e = 0; e < mesh_->ne(); ++e)
        if (mesh_->regions()[e] == lab) {
            v[e] = value;
            ++count;
        }
    return count;
}
void Parameters::set_elem(const std::string &p, int e, double value) { ref(p).at(e) = value; }
double Parameters::avg_p(const std::string &p, int i, int j) const
{
    const double a = ref(p)[i], b = ref(p)[j];
    return (a + b == 0) ? 0 : 2 * a * b / (a + b);
}

// ---------------------------------------------------------------- points
void PointWiseEntity::add_point(const FESpaceRT &Xh, const Geo::Point &p)
{
    double xi[2];
    const int e = Xh.mesh()->locate(p, xi);
    if (e < 0) {
        std::ostringstream os;
        os << "Point (" << p.x << "," << p.y << ") not found in the mesh";
        throw std::runtime_error(os.str());
    }
    ie_.push_back(e);
    relp_.emplace_back(xi[0], xi[1]);
    std::vector<double> s(Xh.nln());
    Xh.basis(xi[0], xi[1], s.data(), nullptr);
    shape_.push_back(std::move(s));
}

Receivers::Receivers(FESpace_ptr Xh, const std::string &fname)
{
    std::ifstream f(fname);
    if (!f.is_open()) throw std::runtime_error("Unable to open receivers file " + fname);
    double x, y;
    while (f >> x >> y) add_point(*Xh, Geo::Point(x, y));
}
Receivers::Receivers(FESpace_ptr Xh, const Geo::Point &p) { add_point(*Xh, p); }
Receivers::Receivers(FESpace_ptr Xh, const std::vector<Geo::Point> &pts)
{
    for (const auto &p : pts) add_point(*Xh, p);
}

void Receivers::write(const std::string &fn, int nrec, int nsamples, const std::vector<double> &seis, int digits)
{
    for (int r = 0; r < nrec; ++r) {
        const std::string name = fn + "_rcv_" + std::to_string(r) + ".out";
        std::ofstream out(name);
        if (!out.is_open()) throw std::runtime_error("File " + name + " may not be opened");
        // digits <= 0 keeps the stream's default formatting (6 significant
        // digits), which is what the reference writes
        if (digits > 0) out.precision(digits);
        for (int s = 0; s < nsamples; ++s)
            out << seis[((size_t)s * nrec + r) * 2] << " " << seis[((size_t)s * nrec + r) * 2 + 1] << "\n";
    }
}

PointWiseForce::PointWiseForce(Fun f, const Geo::Point &p, FESpace_ptr Xh) : Force(std::move(f))
{
    add_point(*Xh, p);
}

double ricker(double f0, double t0, double t)
{
    const double f0sq = f0 * f0;
    return (1 - 2 * kPi * kPi * f0sq * (t - t0) * (t - t0)) * std::exp(-kPi * kPi * f0sq * (t - t0) * (t - t0));
}

// ---------------------------------------------------------------- LeapFrog
void LeapFrog::check(int rc) const
{
    if (rc != TSP_OK) throw EngineError(rc, std::string("tspeed_amd: ") + tsp_last_error(ctx_));
}

LeapFrog::LeapFrog(FESpace_ptr Xh, const Parameters &p, const Receivers &r, double penalty, int precision,
                   int device)
    : Xh_(std::move(Xh)), penalty_(penalty)
{
    int rc = tsp_create(device, precision, 1, &ctx_);
    if (rc != TSP_OK) throw EngineError(rc, "tspeed_amd: cannot create a context on device " + std::to_string(device));
    const auto w2 = Xh_->w2(), w1 = Xh_->w1();
    check(tsp_set_tables(ctx_, Xh_->degree(), Xh_->B().data(), Xh_->G().data(), Xh_->Bedge().data(),
                         Xh_->Gedge().data(), w2.data(), w1.data()));
    // invM = 1/rho 1/detJ base_invmass (Matrices_imp.hpp:183-190)
    if (!Xh_->orthonormal()) check(tsp_set_mass(ctx_, Xh_->base_invmass().data()));
    const Mesh &m = *Xh_->mesh();
    const auto v = m.vertex_array();
    check(tsp_set_mesh(ctx_, m.ne(), v.data(), m.neigh().data(), m.neighedge().data(), p.lambda().data(),
                       p.mu().data(), p.rho().data(), penalty_));
    nrec_ = r.size();
    rec_el_.resize(nrec_);
    rec_phi_.resize((size_t)nrec_ * Xh_->nln());
    for (int k = 0; k < nrec_; ++k) {
        rec_el_[k] = r.elem(k);
        std::copy(r.shape(k).begin(), r.shape(k).end(), rec_phi_.begin() + (size_t)k * Xh_->nln());
    }
    // receivers are registered at first_step, once tmax/dt size the sample buffer
    u0_.assign(2 * (size_t)m.ne() * Xh_->nln(), 0.0);
    v0_ = u0_;
}

LeapFrog::~LeapFrog() { tsp_destroy(ctx_); }

void LeapFrog::set_dt(double dt)
{
    dt_ = dt;
    check(tsp_set_dt(ctx_, dt));
}

void LeapFrog::set_penalty(double p)
{
    if (p != penalty_)
        throw std::logic_error("LeapFrog::set_penalty: the penalty is fixed when B is assembled; pass it to the "
                               "constructor");
}

void LeapFrog::add_force(Force_ptr f)
{
    f_ = std::move(f);
    pf_ = std::dynamic_pointer_cast<PointWiseForce>(f_);
}

void LeapFrog::set_initial_u(FESpace_ptr Xh, const std::function<std::array<double, 2>(double, double)> &fun)
{
    u0_ = Xh->transform(fun);
}
void LeapFrog::set_initial_v(FESpace_ptr Xh, const std::function<std::array<double, 2>(double, double)> &fun)
{
    v0_ = Xh->transform(fun);
}

// Rows of the force table: row 0 at t = 0 (first_step), row n at the time step()
// number n uses, t_{n+1} = t_n + dt accumulated from t_1 = dt (TimeAdvance.cpp:48).
void LeapFrog::ensure_wavelet(int rows_needed)
{
    if (!pf_ || rows_needed <= tab_rows_) return;
    int rows = std::max(rows_needed, 2 * tab_rows_);
    std::vector<double> tab(2 * (size_t)rows);
    double t = 0;
    for (int r = 0; r < rows; ++r) {
        if (r == 1) t = dt_ + dt_;
        else if (r > 1) t = t + dt_;
        const auto F = pf_->value(t);
        tab[2 * (size_t)r] = F[0];
        tab[2 * (size_t)r + 1] = F[1];
    }
    int col = 0, el = pf_->elem(0);
    check(tsp_set_sources(ctx_, 1, &col, &el, pf_->shape(0).data(), rows, tab.data()));
    tab_rows_ = rows;
}

void LeapFrog::first_step()
{
    if (first_done_) throw std::logic_error("Warning: attempt to do first step after others");
    check(tsp_set_state(ctx_, 0, u0_.data(), v0_.data()));
    const int est = (tmax_ > 0 && dt_ > 0) ? (int)std::ceil(tmax_ / dt_) + 2 : 1024;
    max_samples_ = est + 1;
    if (nrec_ > 0) {
        check(tsp_set_receivers(ctx_, nrec_, rec_el_.data(), rec_phi_.data(), 0, max_samples_));
    }
    ensure_wavelet(est + 1);
    check(tsp_first_step(ctx_));
    last_time_ = dt_;
    first_done_ = true;
}

void LeapFrog::step()
{
    run(1, 0);
}

void LeapFrog::run(int n, int stride)
{
    if (!first_done_) throw std::logic_error("LeapFrog::step before first_step");
    ensure_wavelet(tsp_steps_done(ctx_) + n + 1);
    // stride > 0: the step kernels record the receivers themselves every `stride`
    // step() calls; stride 0 leaves sampling to eval_receivers()
    check(tsp_set_stride(ctx_, nrec_ > 0 ? stride : 0));
    check(tsp_run(ctx_, n));
    for (int k = 0; k < n; ++k) {
        const double t = last_time_ + dt_;
        last_time_ = t;
        if (t >= tmax_ - dt_) completed_ = true;
    }
    state_dirty_ = true;
}

void LeapFrog::eval_receivers() { check(tsp_eval_receivers(ctx_)); }

void LeapFrog::write_receivers(const std::string &fn, int digits) const
{
    int ns = 0;
    check(tsp_get_seismograms(ctx_, nullptr, 0, &ns));
    std::vector<double> s((size_t)ns * nrec_ * 2);
    check(tsp_get_seismograms(ctx_, s.data(), ns, &ns));
    Receivers::write(fn, nrec_, ns, s, digits);
}

const std::vector<double> &LeapFrog::get_uh() const
{
    if (state_dirty_ || uh_.empty()) {
        uh_.resize(2 * (size_t)Xh_->ne() * Xh_->nln());
        check(tsp_get_u(ctx_, 0, uh_.data()));
        const_cast<LeapFrog *>(this)->state_dirty_ = false;
    }
    return uh_;
}

int LeapFrog::steps_done() const { return tsp_steps_done(ctx_); }

}  // namespace tspeed

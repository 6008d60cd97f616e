// tspeed.hpp -- host C++ layer of the MI355X engine, keeping tspeed's class and
// method names so that a driver written against the reference (e.g.
// Examples/src/Lamb.cpp, Wedge.cpp) ports by changing the include.  Eigen is not
// used: vectors are std::vector<double>.
//
// What stays on the host (unchanged in behaviour from the reference):
//   Gauss quadrature        QuadratureRule_imp.hpp:23-119
//   Dubiner basis           ShapeFunctions_imp.hpp:25-54, ShapeFunctions.cpp:25-60
//   FESpace tables          FESpace_imp.hpp:26-47
//   Mesh (Gmsh 2.2 reader)  Mesh.cpp:174-257; neighbours by an edge hash in place
//                           of the O(ne^2) scan (Mesh.cpp:33-128), same result on
//                           consistently oriented meshes
//   Parameters              FESpace.hpp:194-247, Parameters.cpp:23-45
//   point location          Receivers_imp.hpp:35-51 with a bucket grid in place of
//                           the linear scan, same element chosen (lowest index
//                           passing Geometry.cpp:112-118)
// What moves to the device: assembly and every time step (include/tspeed_amd.h).
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct tsp_context;

namespace tspeed {

// ------------------------------------------------------------- quadrature
// Gauss-Legendre nodes/weights on [a,b] (QuadratureRule_imp.hpp:23-61).
void gauleg(int n, double a, double b, double *x, double *w);

// Collapsed Gauss rule with nq^2 area nodes and nq edge nodes (Gauss<nq>,
// QuadratureRule.hpp:171-181; QuadratureRule_imp.hpp:66-119).
class Gauss {
public:
    explicit Gauss(int nq);
    int nqn1d() const { return nq_; }
    int nqn2d() const { return nq_ * nq_; }
    double iweight(int k) const { return w2_[k]; }
    double eweight(int k) const { return w1_[k]; }
    std::array<double, 2> inode(int k) const { return {xy2_[2 * k], xy2_[2 * k + 1]}; }
    // node k of reference edge f (0: (1,0)->(0,1), 1: (0,1)->(0,0), 2: (0,0)->(1,0))
    std::array<double, 2> edge_node(int f, int k) const;
    const std::vector<double> &x1() const { return x1_; }

private:
    int nq_;
    std::vector<double> w2_, xy2_, w1_, x1_;
};

// ------------------------------------------------------------- basis
double jacobi_polynomial(int n, int alpha, int beta, double z);

// Orthonormal Dubiner basis of degree N on the reference triangle.
class Dubiner {
public:
    explicit Dubiner(int N) : N_(N) {}
    int degree() const { return N_; }
    int gdl() const { return (N_ + 1) * (N_ + 2) / 2; }
    static constexpr bool is_orthonormal = true;
    // phi[gdl], grad[2*gdl] (d/dxi, d/deta interleaved); grad may be null
    void eval(double xi, double eta, double *phi, double *grad) const;
    double phi(int s, double xi, double eta) const;

private:
    int N_;
};

// The reference's non-orthonormal BoundaryAdapted basis (ShapeFunctions_imp.hpp:
// 56-118): three vertex modes, N-1 modes per edge and the interior bubbles, in
// the reference's order; gradients by forward-mode differentiation of the same
// expressions.
class BoundaryAdaptedBasis {
public:
    explicit BoundaryAdaptedBasis(int N) : N_(N) {}
    int degree() const { return N_; }
    int gdl() const { return (N_ + 1) * (N_ + 2) / 2; }
    void eval(double xi, double eta, double *phi, double *grad) const;

private:
    int N_;
};

enum class BasisKind { Dubiner = 0, BoundaryAdapted = 1 };
void eval_basis(BasisKind kind, int N, double xi, double eta, double *phi, double *grad);
// base_invmass (nln x nln, row-major): identity for Dubiner, the inverse of the
// quadrature mass matrix sum_k phi_i phi_j w_k otherwise (FESpace_imp.hpp:49-67)
std::vector<double> base_invmass(BasisKind kind, int N);

// Type tags for source compatibility with FESpace<N, Q, S> (FESpace.hpp:37):
// FESpace<N, Dunavant<2*N>, BoundaryAdapted<N>> selects the BoundaryAdapted
// basis. The quadrature is always the collapsed Gauss rule with N+1 nodes per
// direction (exact to degree 2N on the triangle, as Dunavant<2N>, and N+1 edge
// nodes exact to degree 2N+1): every integral of the operator is polynomial of
// degree <= 2N, so either rule assembles the same operator up to roundoff.
template <int M> struct BoundaryAdapted {};
template <int M> struct Dunavant {};
template <int M> struct GaussRule {};


// ------------------------------------------------------------- geometry / mesh
namespace Geo {
struct Point {
    double x = 0, y = 0;
    Point() = default;
    Point(double a, double b) : x(a), y(b) {}
};
}  // namespace Geo

class Mesh {
public:
    Mesh() = default;
    // Gmsh 2.2 ASCII reader (Mesh.cpp:174-257): triangles are el_type 2, boundary
    // lines el_type 1, the region is the elementary-entity tag.
    explicit Mesh(const std::string &fileName);
    // Structured rectangle [x0,x0+Lx] x [y0,y0+Ly] of nx*ny squares, each split along
    // its (i,j)-(i+1,j+1) diagonal into two counter-clockwise triangles; interior
    // vertices moved by U(-jitter*h, jitter*h) per coordinate from a splitmix64
    // stream seeded with `seed`.
    static std::shared_ptr<Mesh> split_square(int nx, int ny, double x0, double y0, double Lx, double Ly,
                                              double jitter, uint64_t seed);
    // Build from raw arrays (points[np*2], tri[ne*3] zero-based, region per triangle).
    Mesh(std::vector<double> points, std::vector<int> tri, std::vector<int> region);

    int ne() const { return (int)tri_.size() / 3; }
    int np() const { return (int)pts_.size() / 2; }
    int nbnd() const { return (int)bnd_.size() / 2; }
    const std::vector<double> &points() const { return pts_; }
    const std::vector<int> &triangles() const { return tri_; }
    const std::vector<int> &regions() const { return reg_; }
    const std::vector<int> &boundary_lines() const { return bnd_; }
    const std::vector<int> &neigh() const { return neigh_; }
    const std::vector<int> &neighedge() const { return neighedge_; }
    // verts[e*6 + 2v + c]
    std::vector<double> vertex_array() const;
    double min_edge() const;
    double signed_area(int e) const;

    // First element (lowest index) that contains p under the reference's tolerant
    // test; -1 if none.  xi receives invJ*(p - P0).
    int locate(const Geo::Point &p, double *xi) const;

private:
    void build_neighbors();
    void build_grid() const;
    std::vector<double> pts_;
    std::vector<int> tri_, reg_, bnd_, neigh_, neighedge_;
    std::string fname_;
    // point-location buckets
    mutable std::vector<int> cell_start_, cell_items_;
    mutable double gx0_ = 0, gy0_ = 0, gdx_ = 1, gdy_ = 1;
    mutable int gnx_ = 0, gny_ = 0;
};
using Mesh_ptr = std::shared_ptr<Mesh>;

// FESpace<N, Gauss<N+1>, Dubiner<N>> (FESpace.hpp:37): basis tables on the
// quadrature nodes, in the layout include/tspeed_amd.h documents.
class FESpaceRT {
public:
    FESpaceRT(Mesh_ptr m, int N, BasisKind kind = BasisKind::Dubiner);
    virtual ~FESpaceRT() = default;
    Mesh_ptr mesh() const { return mesh_; }
    int degree() const { return N_; }
    int nln() const { return nln_; }
    int ne() const { return mesh_ ? mesh_->ne() : 0; }
    const Gauss &quad() const { return quad_; }
    BasisKind basis_kind() const { return kind_; }
    bool orthonormal() const { return kind_ == BasisKind::Dubiner; }
    void basis(double xi, double eta, double *phi, double *grad) const { eval_basis(kind_, N_, xi, eta, phi, grad); }
    const std::vector<double> &base_invmass() const { return minv_; }
    const std::vector<double> &B() const { return B_; }
    const std::vector<double> &G() const { return G_; }
    const std::vector<double> &Bedge() const { return Be_; }
    const std::vector<double> &Gedge() const { return Ge_; }
    std::vector<double> w2() const;
    std::vector<double> w1() const;
    // L2 projection of a vector field onto the space (FESpace_imp.hpp:137-179),
    // component-major coefficients.
    std::vector<double> transform(const std::function<std::array<double, 2>(double, double)> &fun) const;
    // The reference's plot output (FESpace_imp.hpp:71-106): four points per
    // element (the images of (0.2,0.2), (0.8,0.2), (0.2,0.8), (0.5,0.5)) as "x y"
    // lines, and the field uh at those points as "ux uy" lines in
    // fname + "u_" + step + ".field".
    void points_out(const std::string &fname) const;
    void field_out(const std::string &fname, const std::vector<double> &uh, unsigned step) const;
    // L2 norm of uex - uh over the mesh with the space's quadrature
    // (FESpace_imp.hpp:108-135).
    double L2error(const std::function<std::array<double, 2>(double, double)> &uex,
                   const std::vector<double> &uh) const;

private:
    Mesh_ptr mesh_;
    int N_, nln_;
    BasisKind kind_;
    Gauss quad_;
    std::vector<double> B_, G_, Be_, Ge_, minv_;
};

template <typename S>
struct basis_of {
    static constexpr BasisKind kind = BasisKind::Dubiner;
};
template <int M>
struct basis_of<BoundaryAdapted<M>> {
    static constexpr BasisKind kind = BasisKind::BoundaryAdapted;
};

template <int N, typename Q = void, typename S = void>
class FESpace : public FESpaceRT {
public:
    explicit FESpace(Mesh_ptr m) : FESpaceRT(std::move(m), N, basis_of<S>::kind) {}
};
using FESpace_ptr = std::shared_ptr<FESpaceRT>;

// Array forms of the output and error routines (also exported as tsh_*):
// verts[e*6 + 2v + c]; uh component-major (c*ne*nln + e*nln + i).
extern const double kPlotPoints[4][2];
void field_points(int ne, const double *verts, double *out);                       // out[(e*4+k)*2+c]
void field_values(int N, int ne, const double *uh, double *out,
                  BasisKind kind = BasisKind::Dubiner);                              // out[(e*4+k)*2+c]
void quad_points(int N, int ne, const double *verts, double *out);                  // out[(e*nq2+k)*2+c]
void project_nodes(int N, int ne, const double *verts, const double *f, double *uh,
                   BasisKind kind = BasisKind::Dubiner);                             // f at quad_points
double l2_error(int N, int ne, const double *verts, const double *uh, const double *uex,
                BasisKind kind = BasisKind::Dubiner);                                // uex at quad_points

// Per-element Lame parameters and density (FESpace.hpp:194-247).
class Parameters {
public:
    explicit Parameters(Mesh_ptr m);
    // set parameter "lambda" | "mu" | "rho" on the elements of region `lab`
    // (Parameters.cpp:23-36); returns the number of elements set
    int setp(const std::string &p, int lab, double value);
    // per-element setter (the reference sets by region only)
    void set_elem(const std::string &p, int e, double value);
    double lambda(int e) const { return lambda_[e]; }
    double mu(int e) const { return mu_[e]; }
    double rho(int e) const { return rho_[e]; }
    const std::vector<double> &lambda() const { return lambda_; }
    const std::vector<double> &mu() const { return mu_; }
    const std::vector<double> &rho() const { return rho_; }
    // harmonic average (Parameters.cpp:37-45)
    double avg_p(const std::string &p, int i, int j) const;

private:
    std::vector<double> &ref(const std::string &p);
    const std::vector<double> &ref(const std::string &p) const;
    Mesh_ptr mesh_;
    std::vector<double> lambda_, mu_, rho_;
};

// Points carried with their element and basis values (PointWiseEntity,
// Receivers.hpp:33-73).  Unlike the reference, a point outside the mesh throws
// std::runtime_error instead of calling exit(1) (Receivers_imp.hpp:50).
class PointWiseEntity {
public:
    virtual ~PointWiseEntity() = default;
    const std::vector<double> &shape(int i) const { return shape_[i]; }
    const Geo::Point &point(int i) const { return relp_[i]; }
    int elem(int i) const { return ie_[i]; }
    int size() const { return (int)ie_.size(); }

protected:
    void add_point(const FESpaceRT &Xh, const Geo::Point &p);
    std::vector<int> ie_;
    std::vector<Geo::Point> relp_;
    std::vector<std::vector<double>> shape_;
};

class Receivers : public PointWiseEntity {
public:
    Receivers(FESpace_ptr Xh, const std::string &fname);   // Receivers_imp.hpp:61-82
    Receivers(FESpace_ptr Xh, const Geo::Point &p);        // :54-60
    Receivers(FESpace_ptr Xh, const std::vector<Geo::Point> &pts);
    // Write <fn>_rcv_<i>.out, one "x y" line per sample (Receivers.cpp:32-46).
    // digits <= 0: the stream's default formatting (6 significant digits), byte
    // for byte the reference's files; digits > 0 sets that precision (17 keeps
    // every double exactly).
    static void write(const std::string &fn, int nrec, int nsamples, const std::vector<double> &seis,
                      int digits = 0);
};

class Force {
public:
    using Fun = std::function<std::array<double, 2>(const double &)>;
    Force() = default;
    explicit Force(Fun f) : f_(std::move(f)) {}
    virtual ~Force() = default;
    std::array<double, 2> value(double t) const { return f_(t); }

protected:
    Fun f_;
};
using Force_ptr = std::shared_ptr<Force>;

// A time-dependent force at one point (Force.hpp:56-86).
class PointWiseForce : public Force, public PointWiseEntity {
public:
    PointWiseForce(Fun f, const Geo::Point &p, FESpace_ptr Xh);
};

// Ricker wavelet in the drivers' form (Examples/src/Lamb.cpp:37).
double ricker(double f0, double t0, double t);

// Error raised by the facade when the device engine reports a failure.
struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// LeapFrog (TimeAdvance.hpp:123-256) over the device engine.  Method names and
// their order of use follow the reference; differences:
//   - penalty and precision are constructor arguments (the reference's
//     set_penalty is a no-op for B; set_penalty here throws if called after
//     construction with a different value);
//   - get_uh() copies the device state to the host on demand;
//   - errors throw EngineError instead of exit().
class LeapFrog {
public:
    LeapFrog(FESpace_ptr Xh, const Parameters &p, const Receivers &r, double penalty = 2.0, int precision = 0,
             int device = 0);
    ~LeapFrog();
    LeapFrog(const LeapFrog &) = delete;
    LeapFrog &operator=(const LeapFrog &) = delete;

    void set_dt(double dt);
    void set_tmax(double tmax) { tmax_ = tmax; }
    void set_penalty(double p);
    void add_force(Force_ptr f);
    void set_initial_u(FESpace_ptr Xh, const std::function<std::array<double, 2>(double, double)> &fun);
    void set_initial_v(FESpace_ptr Xh, const std::function<std::array<double, 2>(double, double)> &fun);
    void first_step();
    void step();
    // advance n steps at once, sampling the receivers every `stride` steps on the
    // device (the fused path); equivalent to n x { step(); if(k%stride==0) eval_receivers(); }
    void run(int n, int stride);
    bool is_running() const { return !completed_; }
    void eval_receivers();
    void write_receivers(const std::string &fn, int digits = 0) const;   // digits as Receivers::write
    const std::vector<double> &get_uh() const;
    const std::vector<double> &u() const { return get_uh(); }
    int steps_done() const;
    tsp_context *context() { return ctx_; }

private:
    void check(int rc) const;
    void ensure_wavelet(int rows_needed);
    void upload_state();
    tsp_context *ctx_ = nullptr;
    FESpace_ptr Xh_;
    double penalty_, dt_ = 0, tmax_ = 0, last_time_ = 0;
    bool completed_ = false, first_done_ = false, state_dirty_ = true;
    Force_ptr f_;
    std::shared_ptr<PointWiseForce> pf_;
    int nrec_ = 0, tab_rows_ = 0;
    std::vector<double> u0_, v0_, rec_phi_;
    std::vector<int> rec_el_;
    mutable std::vector<double> uh_;
    int max_samples_ = 0;
};

}  // namespace tspeed

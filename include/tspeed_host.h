/*
 * tspeed_host.h -- C exports of the host layer (mesh, basis tables, point
 * location, source time functions) for callers that drive the engine through an
 * FFI instead of the C++ classes in tspeed_amd/csrc/host/tspeed.hpp.  None of
 * these touch the GPU.  Each names the reference interface it restates.
 */
#ifndef TSPEED_HOST_H
#define TSPEED_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre rule on [a,b] (QuadratureRule_imp.hpp:23-61). */
int tsh_gauleg(int n, double a, double b, double *x, double *w);

/* Dubiner basis of degree N at one point: phi[nln], grad[2*nln] (d/dxi, d/deta
 * interleaved; may be NULL) (ShapeFunctions_imp.hpp:25-54). */
int tsh_dubiner(int N, double xi, double eta, double *phi, double *grad);

/* FESpace<N> tables in the layout of tsp_set_tables, plus the area nodes
 * xy2[nq2*2] and edge abscissae x1[nq1] (FESpace_imp.hpp:26-47). */
int tsh_tables(int N, double *B, double *G, double *Bedge, double *Gedge, double *w2, double *xy2, double *w1,
               double *x1);

/* Neighbours across reversed shared edges by an edge hash (same result as the
 * O(ne^2) scan of Mesh.cpp:33-128 on consistently oriented meshes).
 * tri[ne*3] zero-based vertex ids; outputs neigh/neighedge[ne*3], -1 on the
 * boundary.  Returns 0, or -1 for a non-manifold edge. */
int tsh_neighbors(int ne, const int *tri, int *neigh, int *neighedge);

/* Structured split-square mesh (see tspeed::Mesh::split_square):
 * pts[(nx+1)*(ny+1)*2], tri[2*nx*ny*3]. */
int tsh_split_square(int nx, int ny, double x0, double y0, double Lx, double Ly, double jitter, uint64_t seed,
                     double *pts, int *tri);

/* Gmsh 2.2 reader (Mesh.cpp:174-257) in two calls: counts first, then the
 * arrays (points[np*2], tri[ntri*3] zero-based, region[ntri],
 * bnd[nbnd*2] zero-based).  Returns 0, or -1 when the file cannot be read. */
int tsh_msh_counts(const char *path, int *np, int *ntri, int *nbnd);
int tsh_msh_read(const char *path, double *points, int *tri, int *region, int *bnd);

/* Point location (Receivers_imp.hpp:35-51): for each query point the first
 * element (lowest index) passing the tolerant point-in-triangle test of
 * Geometry.cpp:112-118, its reference coordinates xi[2*q], and the basis values
 * phi[q*nln] of degree N at that point (N <= 0: no phi).  elem[q] = -1 when the
 * point is outside the mesh.  points[np*2], tri[ne*3]. */
int tsh_locate(int np, const double *points, int ne, const int *tri, int nq, const double *qx, const double *qy,
               int N, int *elem, double *xi, double *phi);

/* Ricker force table in the layout tsp_set_sources takes, for nsrc sources with
 * parameters f0[s], t0[s] and direction (dirx[s], diry[s]):
 * out[(row*nsrc+s)*2+c] = dir_c * ricker(f0, t0, t_row) with t_0 = 0,
 * t_1 = dt + dt and t_{n+1} = t_n + dt (the time LeapFrog::step evaluates the
 * force at, TimeAdvance.cpp:28,48-51); ricker as in Examples/src/Lamb.cpp:37. */
int tsh_ricker_table(int nsrc, const double *f0, const double *t0, const double *dirx, const double *diry,
                     double dt, int nrows, double *out);

/* Plot output of FESpace::points_out / field_out (FESpace_imp.hpp:71-106): the
 * images of the reference points (0.2,0.2), (0.8,0.2), (0.2,0.8), (0.5,0.5) in
 * every element, out[(e*4+k)*2+c], and the field uh (component-major
 * coefficients, degree N) at those points. */
int tsh_field_points(int ne, const double *verts, double *out);
int tsh_field_values(int N, int basis, int ne, const double *uh, double *out);

/* Quadrature nodes of the degree-N space mapped into every element,
 * out[(e*nq2+k)*2+c] (nq2 = (N+1)^2); the L2 projection of a field given by its
 * values f at those nodes (FESpace::transform, FESpace_imp.hpp:137-179); and
 * the L2 norm of uex - uh with uex given at those nodes (FESpace::L2error,
 * FESpace_imp.hpp:108-135). */
int tsh_quad_points(int N, int ne, const double *verts, double *out);
int tsh_project(int N, int basis, int ne, const double *verts, const double *f, double *uh);
int tsh_l2_error(int N, int basis, int ne, const double *verts, const double *uh, const double *uex, double *err);

/* Basis selection (`basis` above and below): 0 = the orthonormal Dubiner basis
 * (ShapeFunctions_imp.hpp:25-54), 1 = the BoundaryAdapted basis
 * (ShapeFunctions_imp.hpp:56-118).  tsh_tables_basis is tsh_tables for either;
 * tsh_base_invmass gives base_invmass (nln x nln, identity for Dubiner,
 * FESpace_imp.hpp:49-67), which tsp_set_mass takes; tsh_basis_eval evaluates
 * the basis (and its gradient, grad may be NULL) at n reference points
 * xi[q*2+c]: phi[q*nln+i], grad[(q*nln+i)*2+c]. */
int tsh_tables_basis(int N, int basis, double *B, double *G, double *Bedge, double *Gedge, double *w2, double *xy2,
                     double *w1, double *x1);
int tsh_base_invmass(int N, int basis, double *out);
int tsh_basis_eval(int N, int basis, int n, const double *xi, double *phi, double *grad);

/* Receivers::write (Receivers.cpp:32-46): <fn>_rcv_<r>.out for r < nrec, one
 * "x y" line per sample from seis[(sample*nrec + r)*2 + c].  digits <= 0 writes
 * the stream's default formatting (6 significant digits), the reference's bytes;
 * digits > 0 that many significant digits (17 keeps every double).  Returns 0,
 * or -1 when a file cannot be opened. */
int tsh_write_receivers(const char *fn, int nrec, int nsamples, const double *seis, int digits);

#ifdef __cplusplus
}
#endif
#endif

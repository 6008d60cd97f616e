/*
 * oracle.h -- FP64 CPU restatement of tspeed's leap-frog path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tspeed_amd/) may include,
 * link or call this code.  It is the checker that tests/, the smoke entry and
 * the CPU-baseline leg of bench.py compare the HIP path against.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the tspeed source tree).  Data layouts follow the reference:
 *   - coefficient vectors are component-major: idx = c*ne*nln + e*nln + i
 *     (TimeAdvance.cpp:73-74, Force_imp.hpp:28-29);
 *   - basis index order: for j in 0..N, for i in 0..N-j (ShapeFunctions_imp.hpp:29-31).
 *
 * Pinning status of each function is stated in the header of oracle.c.
 */
#ifndef TSPEED_ORACLE_H
#define TSPEED_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- quadrature and basis (QuadratureRule_imp.hpp, ShapeFunctions*.{hpp,cpp}) ---- */
int    orc_gauleg(int n, double a, double b, double *x, double *w);
/* collapsed Gauss rule with nq x nq nodes on the reference triangle; 1-D rule on [0,1] */
int    orc_gauss_rule(int nq, double *w2, double *xy2, double *w1, double *x1);
void   orc_edge_nodes(int nq, const double *x1, int edge, double *xy);
double orc_jacobi(int n, int alpha, int beta, double z);
int    orc_nln(int N);
/* phi[nln], grad[nln*2] (d/dxi, d/deta) at one point */
void   orc_dubiner(int N, double xi, double eta, double *phi, double *grad);
/* the non-orthonormal BoundaryAdapted basis (ShapeFunctions_imp.hpp:56-118): vertex,
 * edge and bubble modes in the reference's order */
void   orc_boundary_adapted(int N, double xi, double eta, double *phi, double *grad);
enum { ORC_DUBINER = 0, ORC_BOUNDARY_ADAPTED = 1 };
/* base_invmass (nln x nln): identity for Dubiner, the inverse of the quadrature mass
 * matrix otherwise (FESpace_imp.hpp:49-67) */
int    orc_base_invmass(int N, int basis, double *out);
/* Tables as FESpace builds them (FESpace_imp.hpp:26-47):
 *   B[k*nln+i], G[(i*2+a)*nq2+k], Bedge[(f*nq1+k)*nln+i], Gedge[((i*3+f)*2+a)*nq1+k],
 *   w2[nq2], xy2[nq2*2], w1[nq1], x1[nq1]  with nq1 = N+1, nq2 = nq1^2.              */
int    orc_tables(int N, double *B, double *G, double *Bedge, double *Gedge,
                  double *w2, double *xy2, double *w1, double *x1);
int    orc_tables_basis(int N, int basis, double *B, double *G, double *Bedge, double *Gedge,
                        double *w2, double *xy2, double *w1, double *x1);

/* ---- mesh and geometry (Mesh.cpp:33-128, Geometry.cpp:100-207) ---- */
/* O(ne^2) neighbour search exactly as the reference; tri[ne*3] are vertex ids. */
void   orc_neighbors(int ne, const int *tri, int *neigh, int *neighedge);
/* geo[ne*8] = J00 J01 J10 J11 detJ inv00 inv01 inv10 ... see oracle.c */
void   orc_geometry(int ne, const double *verts, double *jac, double *invjac, double *detj,
                    double *normal /* ne*3*2 */, double *length /* ne*3 */);
/* first element (in index order) containing (px,py); returns element or -1 and xi[2] */
int    orc_locate(int ne, const double *verts, double px, double py, double *xi);

/* ---- operator (Matrices_imp.hpp:22-215, MyMat.cpp, TimeAdvance_imp.hpp:32) ---- */
typedef struct orc_op orc_op;
orc_op *orc_assemble(int N, int ne, const double *verts, const int *neigh, const int *neighedge,
                     const double *lambda, const double *mu, const double *rho, double penalty);
/* the same with the basis selectable (invM blocks = 1/(rho detJ) * base_invmass) */
orc_op *orc_assemble_basis(int N, int basis, int ne, const double *verts, const int *neigh,
                           const int *neighedge, const double *lambda, const double *mu,
                           const double *rho, double penalty);
void   orc_free(orc_op *op);
/* B in canonical merged form: out[((e*4+s)*d + I)*d + J], s=0 self, s=1+f neighbour across
 * edge f (zero block when absent), I = m*nln+i, J = n*nln+j, d = 2*nln. */
void   orc_export_blocks(const orc_op *op, double *out);
/* per-element inverse-mass scalar as the reference forms it (Matrices_imp.hpp:187,189) */
void   orc_export_invm(const orc_op *op, double *out);
/* out = B*x (component-major vectors of length 2*ne*nln) */
void   orc_apply_B(const orc_op *op, const double *x, double *out);
/* y = invM x with the dense per-element blocks (MyMat.hpp:337-342) */
void   orc_apply_invM(const orc_op *op, const double *x, double *out);

/* ---- leap-frog (TimeAdvance.cpp:24-79) ---- */
typedef struct {
    int    elem;          /* source element, -1 = no source */
    const double *phi;    /* nln basis values at the source point */
    double f0, t0;        /* Ricker parameters */
    double dirx, diry;    /* force direction: F = (dirx, diry) * ricker(t) */
} orc_source;

/* Runs first_step then nsteps calls of step().  Samples receivers after first_step and
 * after every step whose 1-based index is a multiple of stride.  u0/v0 may be NULL
 * (zero).  seis[(sample*nrec + r)*2 + c]; returns the number of samples written. */
int    orc_run(const orc_op *op, double dt, int nsteps, const orc_source *src,
               int nrec, const int *rec_elem, const double *rec_phi, int stride,
               const double *u0, const double *v0, double *u_out, double *seis);
/* Ricker wavelet value as the drivers write it (Examples/src/Lamb.cpp:37) */
double orc_ricker(double f0, double t0, double t);
/* largest eigenvalue of invM*B by power iteration (FP64); iters fixed */
/* Threads the parallel loops use (0 = keep): the single-threaded build
 * (liboracle.so) always reports 1; liboracle_omp.so sets and reports its
 * OpenMP team size. */
int    orc_threads(int n);
double orc_lambda_max(const orc_op *op, int iters, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif

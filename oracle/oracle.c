/*
 * oracle.c -- FP64 CPU restatement of tspeed's leap-frog path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  The product never
 * includes, links or calls this file.  Compiled by oracle/Makefile with the
 * reference's Release flags (-O3 -DNDEBUG -msse4, CMakeLists.txt:33), single thread.
 *
 * The reference cannot be built here (it hard-requires Eigen3, which is absent),
 * so this is a restatement written from the reference's source text.  Eigen's
 * dense kernels (mat-vec, dot, 2x2 inverse/determinant) are restated as the
 * textbook loops; their summation order is unpinned.
 *
 * Pinning of each part (what fails if this restatement is wrong):
 *   orc_gauleg / orc_gauss_rule .. numpy leggauss golden vectors; monomial exactness
 *                                  a!b!/(a+b+2)! on the triangle (tests/test_oracle.py)
 *   orc_jacobi .................... scipy.special.eval_jacobi golden vectors
 *   orc_dubiner (phi) ............. orthonormality on the triangle, phi_00 = sqrt(2)
 *   orc_dubiner (grad) ............ central finite differences of phi
 *   orc_neighbors ................. mesh fixture counts (2*interior + boundary = 3*ne,
 *                                   boundary half-faces == el_type-1 lines), symmetry
 *   orc_geometry .................. closed forms (detJ = 2*area, J*invJ = I, n . edge = 0,
 *                                   n outward)
 *   orc_assemble .................. symmetry of B; rigid-body modes in ker(B) (all
 *                                   boundaries traction-free); A against the analytic
 *                                   constant-strain energy; S kills continuous fields;
 *                                   invM = 1/(rho*detJ)
 *   orc_run ....................... discrete energy conservation with f = 0; one-element
 *                                   hand step; Taylor-start identity; convergence to an
 *                                   exact standing wave at order N+1 (tests/test_oracle.py)
 *   orc_boundary_adapted .......... vertex modes = barycentric coordinates; edge/bubble
 *                                   traces vanish where they must and equal the closed
 *                                   Jacobi forms on their edge; gradient by central
 *                                   differences; span = P_N (exact Dubiner combination)
 *   orc_base_invmass .............. inverse of the quadrature mass; mass = T^T T through
 *                                   the Dubiner change of basis
 *   orc_assemble_basis ............ lambda_max(M^-1 B) and seismograms identical for the
 *                                   BoundaryAdapted and Dubiner spaces (basis-independent
 *                                   dynamics)
 *   Only running the compiled reference would pin its exact FP64 trajectory (penalty
 *   tensor with 2*mu_bar, the 1/2-weighted consistency term, Eigen's summation order).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* Built twice: liboracle.so with the reference's single-threaded Release flags,
 * and liboracle_omp.so with -fopenmp as a labelled multi-core CPU baseline.  The
 * parallel loops split rows or vector entries only, so each output entry keeps
 * the same summation order and both builds give bitwise equal results. */
#ifdef _OPENMP
#include <omp.h>
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PAR
#endif

int orc_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Gauss-Legendre nodes/weights on [a,b] by Newton iteration on P_n
 * (QuadratureRule_imp.hpp:23-61; the textbook "gauleg" recurrence).  Nodes ascend. */
int orc_gauleg(int n, double a, double b, double *x, double *w)
{
    const double m = (n + 1.0) / 2.0;
    const double xm = 0.5 * (b + a), xl = 0.5 * (b - a);
    for (int i = 1; i <= m; ++i) {
        double z = cos(ORC_PI * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
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
        } while (fabs(z - z1) > DBL_EPSILON && ++it < 200);
        if (it >= 200) return -1;
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
    return 0;
}

/* Collapsed (Duffy) Gauss rule, nq^2 nodes, weights summing to 1/2
 * (QuadratureRule_imp.hpp:93-119).  1-D edge rule on [0,1] (:100). */
int orc_gauss_rule(int nq, double *w2, double *xy2, double *w1, double *x1)
{
    double x[32], w[32];
    if (nq < 1 || nq > 32) return -1;
    if (orc_gauleg(nq, 0.0, 1.0, x1, w1)) return -1;
    if (orc_gauleg(nq, -1.0, 1.0, x, w)) return -1;
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            w2[i * nq + j] = (1 - x[i]) * w[i] * w[j] / 8.;
            xy2[(i * nq + j) * 2 + 0] = (1 + x[i]) / 2.;
            xy2[(i * nq + j) * 2 + 1] = (1 - x[i]) * (1 + x[j]) / 4.;
        }
    return 0;
}

/* Edge quadrature nodes on the reference triangle (QuadratureRule_imp.hpp:66-91):
 * edge 0 = (1,0)->(0,1), edge 1 = (0,1)->(0,0), edge 2 = (0,0)->(1,0). */
void orc_edge_nodes(int nq, const double *x1, int edge, double *xy)
{
    for (int k = 0; k < nq; ++k) {
        double s = x1[k];
        if (edge == 2) { xy[2 * k] = s;       xy[2 * k + 1] = 0.0; }
        if (edge == 0) { xy[2 * k] = 1.0 - s; xy[2 * k + 1] = s; }
        if (edge == 1) { xy[2 * k] = 0.0;     xy[2 * k + 1] = 1.0 - s; }
    }
}

/* Jacobi polynomial P_n^{(alpha,beta)}(z) by the three-term recurrence
 * (ShapeFunctions.cpp:25-60). */
double orc_jacobi(int n, int alpha, int beta, double z)
{
    if (n == 0) return 1.0;
    if (n == 1) return 0.5 * (alpha - beta + (alpha + beta + 2.0) * z);
    const int apb = alpha + beta;
    double poly = 0.0, polyn2 = 1.0, polyn1 = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z);
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 = a2 / a1;
        a3 = a3 / a1;
        a4 = a4 / a1;
        poly = (a2 + a3 * z) * polyn1 - a4 * polyn2;
        polyn2 = polyn1;
        polyn1 = poly;
    }
    return poly;
}

int orc_nln(int N) { return (N + 1) * (N + 2) / 2; }

/* Orthonormal Dubiner basis and its analytic gradient (ShapeFunctions_imp.hpp:25-54). */
void orc_dubiner(int N, double csi, double eta, double *phi, double *grad)
{
    int s = 0;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N - j; ++i, ++s) {
            const double cij = sqrt((2. * i + 1) * 2. * (i + j + 1) / pow(4., i));
            const double a = 2 * csi / (1 - eta) - 1, b = 2 * eta - 1;
            const double Pi = orc_jacobi(i, 0, 0, a), Pj = orc_jacobi(j, 2 * i + 1, 0, b);
            if (phi) phi[s] = cij * pow(2., i) * pow(1 - eta, i) * Pi * Pj;
            if (!grad) continue;
            double gx, gy;
            if (i == 0 && j == 0) {
                gx = 0.0;
                gy = 0.0;
            } else if (i == 0) {
                gx = 0.0;
                gy = cij * (j + 2) * orc_jacobi(j - 1, 2, 1, b);
            } else if (j == 0) {
                const double Q = orc_jacobi(i - 1, 1, 1, a);
                gx = cij * pow(2., i) * pow(1 - eta, i - 1) * (i + 1) * Q;
                gy = cij * pow(2., i) * (-i * pow(1 - eta, i - 1) * Pi + csi * pow(1 - eta, i - 2) * (i + 1) * Q);
            } else {
                const double Q = orc_jacobi(i - 1, 1, 1, a), R = orc_jacobi(j - 1, 2 * i + 2, 1, b);
                gx = cij * pow(2., i) * pow(1 - eta, i - 1) * (i + 1) * Q * Pj;
                gy = cij * pow(2., i) *
                     (-i * pow(1 - eta, i - 1) * Pi * Pj + csi * pow(1 - eta, i - 2) * (i + 1) * Q * Pj +
                      pow(1 - eta, i) * (2 * i + j + 2) * Pi * R);
            }
            grad[2 * s] = gx;
            grad[2 * s + 1] = gy;
        }
}

/* Basis tables exactly as FESpace tabulates them (FESpace_imp.hpp:26-47),
 * with Q = Gauss<N+1> (FESpace.hpp:37). */
/* ---- BoundaryAdapted basis (ShapeFunctions_imp.hpp:56-118), evaluated with
 * forward-mode dual numbers (value, d/dxi, d/deta) so the gradient is the exact
 * derivative of the same expression. */
typedef struct { double v, x, y; } dual;
static dual dconst(double c) { dual r = {c, 0, 0}; return r; }
static dual dadd(dual a, dual b) { dual r = {a.v + b.v, a.x + b.x, a.y + b.y}; return r; }
static dual dsub(dual a, dual b) { dual r = {a.v - b.v, a.x - b.x, a.y - b.y}; return r; }
static dual dmul(dual a, dual b) { dual r = {a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y}; return r; }
static dual dscale(dual a, double c) { dual r = {a.v * c, a.x * c, a.y * c}; return r; }
static dual ddiv(dual a, dual b)
{
    const double q = a.v / b.v;
    dual r = {q, (a.x - q * b.x) / b.v, (a.y - q * b.y) / b.v};
    return r;
}
static dual dpowi(dual a, int n)
{
    dual r = dconst(1.0);
    for (int k = 0; k < n; ++k) r = dmul(r, a);
    return r;
}
/* the three-term recurrence of orc_jacobi on a dual argument */
static dual djacobi(int n, int alpha, int beta, dual z)
{
    if (n == 0) return dconst(1.0);
    dual p1 = dscale(dadd(dconst((double)(alpha - beta)), dscale(z, alpha + beta + 2.0)), 0.5);
    if (n == 1) return p1;
    const int apb = alpha + beta;
    dual p2 = dconst(1.0), p = p1;
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 = a2 / a1;
        a3 = a3 / a1;
        a4 = a4 / a1;
        p = dsub(dmul(dadd(dconst(a2), dscale(z, a3)), p1), dscale(p2, a4));
        p2 = p1;
        p1 = p;
    }
    return p;
}

void orc_boundary_adapted(int N, double csi, double eta, double *phi, double *grad)
{
    const dual x = {csi, 1, 0}, y = {eta, 0, 1};
    const dual one = dconst(1.0);
    const dual l0 = dsub(dsub(one, x), y);            /* 1 - xi - eta */
    const dual om = dsub(one, y);                     /* 1 - eta */
    const dual zx = dsub(dscale(ddiv(x, om), 2.0), one); /* 2 xi / (1 - eta) - 1 */
    const dual zy = dsub(dscale(y, 2.0), one);        /* 2 eta - 1 */
    dual f[(8 + 1) * (8 + 2) / 2];
    int s = 0;
    f[s++] = l0;                                       /* vertex modes */
    f[s++] = x;
    f[s++] = y;
    for (int i = 1; i < N; ++i)                        /* edge eta = 0 */
        f[s++] = dmul(dmul(dmul(l0, x), dpowi(om, i - 1)), djacobi(i - 1, 1, 1, zx));
    for (int j = 1; j < N; ++j)                        /* edge xi = 0 */
        f[s++] = dmul(dmul(l0, y), djacobi(j - 1, 1, 1, zy));
    for (int j = 1; j < N; ++j)                        /* edge xi + eta = 1 */
        f[s++] = dmul(dmul(x, y), djacobi(j - 1, 1, 1, zy));
    for (int j = 1; j < N; ++j)                        /* interior bubbles */
        for (int i = 1; i < N; ++i)
            if (i + j < N)
                f[s++] = dmul(dmul(dmul(dmul(dmul(l0, x), y), dpowi(om, i - 1)), djacobi(i - 1, 1, 1, zx)),
                              djacobi(j - 1, 2 * i + 1, 1, zy));
    for (int k = 0; k < s; ++k) {
        if (phi) phi[k] = f[k].v;
        if (grad) {
            grad[2 * k] = f[k].x;
            grad[2 * k + 1] = f[k].y;
        }
    }
}

int orc_tables(int N, double *B, double *G, double *Bedge, double *Gedge,
               double *w2, double *xy2, double *w1, double *x1)
{
    return orc_tables_basis(N, ORC_DUBINER, B, G, Bedge, Gedge, w2, xy2, w1, x1);
}

static void eval_basis(int N, int basis, double xi, double eta, double *phi, double *grad)
{
    if (basis == ORC_BOUNDARY_ADAPTED) orc_boundary_adapted(N, xi, eta, phi, grad);
    else orc_dubiner(N, xi, eta, phi, grad);
}

int orc_tables_basis(int N, int basis, double *B, double *G, double *Bedge, double *Gedge,
                     double *w2, double *xy2, double *w1, double *x1)
{
    const int nln = orc_nln(N), nq1 = N + 1, nq2 = nq1 * nq1;
    double *phi = malloc(sizeof(double) * nln), *gr = malloc(sizeof(double) * 2 * nln);
    double exy[2 * 32];
    if (orc_gauss_rule(nq1, w2, xy2, w1, x1)) { free(phi); free(gr); return -1; }
    for (int k = 0; k < nq2; ++k) {
        eval_basis(N, basis, xy2[2 * k], xy2[2 * k + 1], phi, gr);
        for (int i = 0; i < nln; ++i) {
            B[k * nln + i] = phi[i];
            G[(i * 2 + 0) * nq2 + k] = gr[2 * i];
            G[(i * 2 + 1) * nq2 + k] = gr[2 * i + 1];
        }
    }
    for (int f = 0; f < 3; ++f) {
        orc_edge_nodes(nq1, x1, f, exy);
        for (int k = 0; k < nq1; ++k) {
            eval_basis(N, basis, exy[2 * k], exy[2 * k + 1], phi, gr);
            for (int i = 0; i < nln; ++i) {
                Bedge[(f * nq1 + k) * nln + i] = phi[i];
                Gedge[((i * 3 + f) * 2 + 0) * nq1 + k] = gr[2 * i];
                Gedge[((i * 3 + f) * 2 + 1) * nq1 + k] = gr[2 * i + 1];
            }
        }
    }
    free(phi);
    free(gr);
    return 0;
}

/* O(ne^2) neighbour search (Mesh.cpp:33-128).  Edge f of a triangle (v0,v1,v2) is
 * f=0:(v1,v2), f=1:(v2,v0), f=2:(v0,v1).  Two triangles are neighbours across
 * edges f/g when the edges appear reversed.  The reference's else-if chain tests
 * f in {2,0,1} and g in {2,1,0}, keeping the first match of a pair. */
void orc_neighbors(int ne, const int *tri, int *neigh, int *neighedge)
{
    static const int forder[3] = {2, 0, 1}, gorder[3] = {2, 1, 0};
    for (int e = 0; e < 3 * ne; ++e) neigh[e] = neighedge[e] = -1;
    for (int a = 0; a < ne - 1; ++a) {
        const int *v = tri + 3 * a;
        for (int b = a + 1; b < ne; ++b) {
            const int *w = tri + 3 * b;
            int done = 0;
            for (int fi = 0; fi < 3 && !done; ++fi)
                for (int gi = 0; gi < 3 && !done; ++gi) {
                    const int f = forder[fi], g = gorder[gi];
                    const int va = v[(f + 1) % 3], vb = v[(f + 2) % 3];
                    const int wa = w[(g + 1) % 3], wb = w[(g + 2) % 3];
                    if (va == wb && vb == wa) {
                        neigh[3 * a + f] = b;
                        neigh[3 * b + g] = a;
                        neighedge[3 * a + f] = g;
                        neighedge[3 * b + g] = f;
                        done = 1;
                    }
                }
        }
    }
}

/* Per-element affine geometry (Geometry.cpp:178-198): J = [P1-P0, P2-P0],
 * detJ = J00*J11 - J10*J01, invJ by the 2x2 cofactor formula; edges
 * e0=(P1,P2), e1=(P2,P0), e2=(P0,P1) (:125-132); unit normal
 * (b.y-a.y, a.x-b.x)/|.| (:100-105); length |a-b| (Geometry.hpp:287).
 * jac/invjac are row-major 2x2 per element. */
void orc_geometry(int ne, const double *verts, double *jac, double *invjac, double *detj,
                  double *normal, double *length)
{
    for (int e = 0; e < ne; ++e) {
        const double *P = verts + 6 * e;
        const double J00 = P[2] - P[0], J10 = P[3] - P[1];
        const double J01 = P[4] - P[0], J11 = P[5] - P[1];
        const double det = J00 * J11 - J10 * J01;
        const double invdet = 1.0 / det;
        if (jac) { jac[4 * e] = J00; jac[4 * e + 1] = J01; jac[4 * e + 2] = J10; jac[4 * e + 3] = J11; }
        if (invjac) {
            invjac[4 * e + 0] = J11 * invdet;
            invjac[4 * e + 2] = -J10 * invdet;
            invjac[4 * e + 1] = -J01 * invdet;
            invjac[4 * e + 3] = J00 * invdet;
        }
        if (detj) detj[e] = det;
        for (int f = 0; f < 3; ++f) {
            const double ax = P[2 * ((f + 1) % 3)], ay = P[2 * ((f + 1) % 3) + 1];
            const double bx = P[2 * ((f + 2) % 3)], by = P[2 * ((f + 2) % 3) + 1];
            const double nx = -ay + by, ny = ax - bx;
            const double nn = sqrt(nx * nx + ny * ny);
            if (normal) { normal[(3 * e + f) * 2] = nx / nn; normal[(3 * e + f) * 2 + 1] = ny / nn; }
            if (length) {
                const double dx = ax - bx, dy = ay - by;
                length[3 * e + f] = sqrt(dx * dx + dy * dy);
            }
        }
    }
}

/* Point location by linear scan (Receivers_imp.hpp:35-51) with the tolerant
 * point-in-triangle test (Geometry.cpp:112-118) and invmap (:199-202). */
int orc_locate(int ne, const double *verts, double px, double py, double *xi)
{
    for (int e = 0; e < ne; ++e) {
        const double *P = verts + 6 * e;
        const double x0 = P[0], y0 = P[1], x1 = P[2], y1 = P[3], x2 = P[4], y2 = P[5];
        const double s = y0 * x2 - x0 * y2 + (y2 - y0) * px + (x0 - x2) * py;
        const double t = x0 * y1 - y0 * x1 + (y0 - y1) * px + (x1 - x0) * py;
        const double A = 1. / 2. * (-y1 * x2 + y0 * (-x1 + x2) + x0 * (y1 - y2) + x1 * y2);
        if (s > -1e-6 && t > -1e-6 && (s + t) < (2 + 1e-6) * A) {
            double inv[4];
            orc_geometry(1, P, NULL, inv, NULL, NULL, NULL);
            const double dx = px - x0, dy = py - y0;
            xi[0] = inv[0] * dx + inv[1] * dy;
            xi[1] = inv[2] * dx + inv[3] * dy;
            return e;
        }
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* Block matrices with the reference's pattern (MyMat.cpp:110-154): row e holds
 * its diagonal block, then one block per existing neighbour in edge order 0,1,2.
 * Each block is nln x nln, stored row-major here. */
struct orc_op {
    int N, nln, ne, nblk;
    int *rowptr, *col;             /* block CSR */
    double *B[4];                  /* components (0,0),(0,1),(1,0),(1,1), nblk blocks */
    double *invM[2];               /* dense diagonal blocks, ne blocks per component */
    double *invm;                  /* the scalar each invM block is built from */
    int *neigh;                    /* copy of ne*3 */
};

static double *blk(double *base, int b, int nln) { return base + (size_t)b * nln * nln; }

static int find_block(const orc_op *op, int r, int c)
{
    for (int b = op->rowptr[r]; b < op->rowptr[r + 1]; ++b)
        if (op->col[b] == c) return b;
    return -1;
}

static double mat_dot(const double *a, const double *b) /* TimeAdvance.cpp:66; [r*2+c] */
{
    return a[0] * b[0] + a[2] * b[2] + a[1] * b[1] + a[3] * b[3];
}

/* sigma = C : eps for isotropic Hooke (TimeAdvance.cpp:60-65, Matrices_imp.hpp:64-65) */
static void hooke(const double *e, double lambda, double mu, double *s)
{
    s[0] = (lambda + 2 * mu) * e[0] + lambda * e[3];
    s[1] = 2 * mu * e[1];
    s[2] = 2 * mu * e[1];
    s[3] = (lambda + 2 * mu) * e[3] + lambda * e[0];
}

/* strain of phi_i e_n from the physical gradient g (Matrices_imp.hpp:37-38, 60-61) */
static void strain(const double *g, int n, double *e)
{
    if (n == 0) { e[0] = g[0]; e[1] = .5 * g[1]; e[2] = .5 * g[1]; e[3] = 0; }
    else        { e[0] = 0.;   e[1] = .5 * g[0]; e[2] = .5 * g[0]; e[3] = g[1]; }
}

/* harmonic average (Parameters.cpp:37-45) */
static double avg_p(double a, double b) { return (a + b == 0) ? 0 : 2 * a * b / (a + b); }

/* Matrices::Matrices (Matrices_imp.hpp:192-215) followed by
 * B = penalty*N*N*S + I + A (TimeAdvance_imp.hpp:32). */
/* base_mass(i,j) = sum_k B(k,i) B(k,j) w_k and its inverse (FESpace_imp.hpp:49-67;
 * identity for the orthonormal Dubiner basis).  Gauss-Jordan with partial pivoting
 * in place of Eigen's inverse(). */
int orc_base_invmass(int N, int basis, double *out)
{
    const int nln = orc_nln(N), nq2 = (N + 1) * (N + 1);
    if (basis != ORC_BOUNDARY_ADAPTED) {
        for (int i = 0; i < nln; ++i)
            for (int j = 0; j < nln; ++j) out[i * nln + j] = i == j ? 1.0 : 0.0;
        return 0;
    }
    double *Bt = malloc(sizeof(double) * nq2 * nln), *Gt = malloc(sizeof(double) * nln * 2 * nq2);
    double *Be = malloc(sizeof(double) * 3 * (N + 1) * nln), *Ge = malloc(sizeof(double) * nln * 3 * 2 * (N + 1));
    double w2[32 * 32], xy2[2 * 32 * 32], w1[32], x1[32];
    orc_tables_basis(N, basis, Bt, Gt, Be, Ge, w2, xy2, w1, x1);
    double *a = calloc((size_t)nln * 2 * nln, sizeof(double));   /* [M | I] */
    for (int i = 0; i < nln; ++i) {
        for (int j = 0; j < nln; ++j) {
            double m = 0.0;
            for (int k = 0; k < nq2; ++k) m += Bt[k * nln + i] * Bt[k * nln + j] * w2[k];
            a[i * 2 * nln + j] = m;
        }
        a[i * 2 * nln + nln + i] = 1.0;
    }
    int rc = 0;
    for (int c = 0; c < nln && !rc; ++c) {
        int p = c;
        for (int r = c + 1; r < nln; ++r)
            if (fabs(a[r * 2 * nln + c]) > fabs(a[p * 2 * nln + c])) p = r;
        if (a[p * 2 * nln + c] == 0.0) { rc = -1; break; }
        if (p != c)
            for (int j = 0; j < 2 * nln; ++j) {
                const double t = a[c * 2 * nln + j];
                a[c * 2 * nln + j] = a[p * 2 * nln + j];
                a[p * 2 * nln + j] = t;
            }
        const double piv = a[c * 2 * nln + c];
        for (int j = 0; j < 2 * nln; ++j) a[c * 2 * nln + j] /= piv;
        for (int r = 0; r < nln; ++r)
            if (r != c) {
                const double fct = a[r * 2 * nln + c];
                if (fct != 0.0)
                    for (int j = 0; j < 2 * nln; ++j) a[r * 2 * nln + j] -= fct * a[c * 2 * nln + j];
            }
    }
    for (int i = 0; i < nln; ++i)
        for (int j = 0; j < nln; ++j) out[i * nln + j] = a[i * 2 * nln + nln + j];
    free(a); free(Bt); free(Gt); free(Be); free(Ge);
    return rc;
}

orc_op *orc_assemble(int N, int ne, const double *verts, const int *neigh, const int *neighedge,
                     const double *lambda, const double *mu, const double *rho, double penalty)
{
    return orc_assemble_basis(N, ORC_DUBINER, ne, verts, neigh, neighedge, lambda, mu, rho, penalty);
}

orc_op *orc_assemble_basis(int N, int basis, int ne, const double *verts, const int *neigh, const int *neighedge,
                           const double *lambda, const double *mu, const double *rho, double penalty)
{
    const int nln = orc_nln(N), nq1 = N + 1, nq2 = nq1 * nq1, nn = nln * nln;
    orc_op *op = calloc(1, sizeof(orc_op));
    op->N = N; op->nln = nln; op->ne = ne;
    op->rowptr = malloc(sizeof(int) * (ne + 1));
    op->neigh = malloc(sizeof(int) * 3 * ne);
    memcpy(op->neigh, neigh, sizeof(int) * 3 * ne);
    int cnt = 0;
    for (int e = 0; e < ne; ++e) {
        op->rowptr[e] = cnt++;
        for (int f = 0; f < 3; ++f) if (neigh[3 * e + f] >= 0) ++cnt;
    }
    op->rowptr[ne] = cnt;
    op->nblk = cnt;
    op->col = malloc(sizeof(int) * cnt);
    for (int e = 0, b = 0; e < ne; ++e) {
        op->col[b++] = e;
        for (int f = 0; f < 3; ++f) if (neigh[3 * e + f] >= 0) op->col[b++] = neigh[3 * e + f];
    }
    double *S[4], *I[4], *A[4];
    for (int c = 0; c < 4; ++c) {
        S[c] = calloc((size_t)cnt * nn, sizeof(double));
        I[c] = calloc((size_t)cnt * nn, sizeof(double));
        A[c] = calloc((size_t)ne * nn, sizeof(double));
    }
    op->invM[0] = calloc((size_t)ne * nn, sizeof(double));
    op->invM[1] = calloc((size_t)ne * nn, sizeof(double));
    op->invm = malloc(sizeof(double) * ne);

    double *Bt = malloc(sizeof(double) * nq2 * nln), *Gt = malloc(sizeof(double) * nln * 2 * nq2);
    double *Be = malloc(sizeof(double) * 3 * nq1 * nln), *Ge = malloc(sizeof(double) * nln * 3 * 2 * nq1);
    double w2[32 * 32], xy2[2 * 32 * 32], w1[32], x1[32];
    orc_tables_basis(N, basis, Bt, Gt, Be, Ge, w2, xy2, w1, x1);
    double *bim = malloc(sizeof(double) * nn);
    orc_base_invmass(N, basis, bim);

    /* per-element scratch: sigma/eps at area nodes [k][i][n][4] and edge nodes [k][i][f][n][4] */
    double *sig = malloc(sizeof(double) * nq2 * nln * 2 * 4), *eps = malloc(sizeof(double) * nq2 * nln * 2 * 4);
    double *sige = malloc(sizeof(double) * nq1 * nln * 3 * 2 * 4);
    double jac[4], inv[4], det, nrm[6], len[3];

    for (int e = 0; e < ne; ++e) {
        orc_geometry(1, verts + 6 * e, jac, inv, &det, nrm, len);
        const double lam = lambda[e], m_u = mu[e];
        /* sigmaeps (Matrices_imp.hpp:49-72) */
        for (int k = 0; k < nq2; ++k)
            for (int i = 0; i < nln; ++i) {
                const double gx = Gt[(i * 2) * nq2 + k], gy = Gt[(i * 2 + 1) * nq2 + k];
                const double g[2] = {gx * inv[0] + gy * inv[2], gx * inv[1] + gy * inv[3]};
                for (int n = 0; n < 2; ++n) {
                    double *ep = eps + ((k * nln + i) * 2 + n) * 4;
                    strain(g, n, ep);
                    hooke(ep, lam, m_u, sig + ((k * nln + i) * 2 + n) * 4);
                }
            }
        /* sigmaeps_edge (Matrices_imp.hpp:23-47) */
        for (int f = 0; f < 3; ++f)
            for (int k = 0; k < nq1; ++k)
                for (int i = 0; i < nln; ++i) {
                    const double gx = Ge[((i * 3 + f) * 2) * nq1 + k], gy = Ge[((i * 3 + f) * 2 + 1) * nq1 + k];
                    const double g[2] = {gx * inv[0] + gy * inv[2], gx * inv[1] + gy * inv[3]};
                    for (int n = 0; n < 2; ++n) {
                        double ep[4];
                        strain(g, n, ep);
                        hooke(ep, lam, m_u, sige + ((((k * nln + i) * 3 + f) * 2) + n) * 4);
                    }
                }
        /* stress_loc (Matrices_imp.hpp:73-94) */
        for (int k = 0; k < nq2; ++k) {
            const double dx = fabs(det) * w2[k];
            for (int j = 0; j < nln; ++j)
                for (int i = 0; i < nln; ++i)
                    for (int m = 0; m < 2; ++m)
                        for (int n = 0; n < 2; ++n)
                            blk(A[m * 2 + n], e, nln)[i * nln + j] +=
                                mat_dot(sig + ((k * nln + j) * 2 + n) * 4, eps + ((k * nln + i) * 2 + m) * 4) * dx;
        }
        /* stability (Matrices_imp.hpp:142-182) */
        const int bee = find_block(op, e, e);
        for (int f = 0; f < 3; ++f) {
            const int nb = neigh[3 * e + f];
            if (nb < 0) continue;
            const double h_scaled = 1. / len[f], n0 = nrm[2 * f], n1 = nrm[2 * f + 1];
            const int g = neighedge[3 * e + f], ben = find_block(op, e, nb);
            const double al = avg_p(lam, lambda[nb]), am = avg_p(m_u, mu[nb]);
            const double c00 = (al + 2 * am) * n0 * n0 + 2 * am * n1 * n1;
            const double c11 = (al + 2 * am) * n1 * n1 + 2 * am * n0 * n0;
            const double c01 = al * n0 * n1;
            for (int k = 0; k < nq1; ++k) {
                const int kk = nq1 - k - 1;
                const double ds = len[f] * w1[k];
                const double *bk = Be + (f * nq1 + k) * nln, *bn = Be + (g * nq1 + kk) * nln;
                for (int i = 0; i < nln; ++i)
                    for (int j = 0; j < nln; ++j) {
                        blk(S[0], bee, nln)[i * nln + j] += c00 * h_scaled * bk[i] * bk[j] * ds;
                        blk(S[3], bee, nln)[i * nln + j] += c11 * h_scaled * bk[i] * bk[j] * ds;
                        blk(S[1], bee, nln)[i * nln + j] += c01 * h_scaled * bk[i] * bk[j] * ds;
                        blk(S[0], ben, nln)[i * nln + j] -= c00 * h_scaled * bk[i] * bn[j] * ds;
                        blk(S[3], ben, nln)[i * nln + j] -= c11 * h_scaled * bk[i] * bn[j] * ds;
                        blk(S[1], ben, nln)[i * nln + j] -= c01 * h_scaled * bk[i] * bn[j] * ds;
                    }
            }
            memcpy(blk(S[2], ben, nln), blk(S[1], ben, nln), sizeof(double) * nn);
        }
        memcpy(blk(S[2], bee, nln), blk(S[1], bee, nln), sizeof(double) * nn);
        /* interelement (Matrices_imp.hpp:95-141) */
        for (int f = 0; f < 3; ++f) {
            const int nb = neigh[3 * e + f];
            if (nb < 0) continue;
            const double n0 = nrm[2 * f], n1 = nrm[2 * f + 1];
            const int g = neighedge[3 * e + f], ben = find_block(op, e, nb);
            for (int k = 0; k < nq1; ++k) {
                const int kk = nq1 - k - 1;
                const double ds = len[f] * w1[k];
                const double *bk = Be + (f * nq1 + k) * nln, *bn = Be + (g * nq1 + kk) * nln;
                for (int i = 0; i < nln; ++i)
                    for (int j = 0; j < nln; ++j)
                        for (int m = 0; m < 2; ++m)
                            for (int n = 0; n < 2; ++n) {
                                const double *s = sige + ((((k * nln + i) * 3 + f) * 2) + m) * 4;
                                const double sn = s[n * 2 + 0] * n0 + s[n * 2 + 1] * n1;
                                blk(I[m * 2 + n], bee, nln)[i * nln + j] -= 0.5 * sn * bk[j] * ds;
                                blk(I[m * 2 + n], ben, nln)[i * nln + j] += 0.5 * sn * bn[j] * ds;
                            }
            }
        }
        /* mass (Matrices_imp.hpp:183-190): invM = 1./rho*1./detJ * base_invmass */
        const double sc = 1. / rho[e] * 1. / det;
        op->invm[e] = sc;
        for (int c = 0; c < 2; ++c)
            for (int i = 0; i < nln; ++i)
                for (int j = 0; j < nln; ++j)
                    blk(op->invM[c], e, nln)[i * nln + j] = sc * bim[i * nln + j];
    }
    /* I.symmetrize() (MyMat.hpp:321-328, MyMat.cpp:156-187): I <- I + I^T */
    for (int c = 0; c < 4; ++c) op->B[c] = malloc(sizeof(double) * (size_t)cnt * nn);
    for (int m = 0; m < 2; ++m)
        for (int n = 0; n < 2; ++n)
            for (int r = 0; r < ne; ++r)
                for (int b = op->rowptr[r]; b < op->rowptr[r + 1]; ++b) {
                    const int c = op->col[b], bt = find_block(op, c, r);
                    const double *src = blk(I[m * 2 + n], b, nln), *tr = blk(I[n * 2 + m], bt, nln);
                    double *dst = blk(op->B[m * 2 + n], b, nln);
                    for (int i = 0; i < nln; ++i)
                        for (int j = 0; j < nln; ++j) dst[i * nln + j] = src[i * nln + j] + tr[j * nln + i];
                }
    /* B = penalty*N*N*S + I + A  (MyMat.cpp:97-109, 219-232, 263-289) */
    const double cpen = penalty * N * N;
    for (int c = 0; c < 4; ++c) {
        for (size_t q = 0; q < (size_t)cnt * nn; ++q) op->B[c][q] = cpen * S[c][q] + op->B[c][q];
        for (int e = 0; e < ne; ++e) {
            double *d = blk(op->B[c], op->rowptr[e], nln);
            const double *a = blk(A[c], e, nln);
            for (int q = 0; q < nn; ++q) d[q] += a[q];
        }
        free(S[c]); free(I[c]); free(A[c]);
    }
    free(Bt); free(Gt); free(Be); free(Ge); free(sig); free(eps); free(sige); free(bim);
    return op;
}

void orc_free(orc_op *op)
{
    if (!op) return;
    for (int c = 0; c < 4; ++c) free(op->B[c]);
    free(op->invM[0]); free(op->invM[1]); free(op->invm);
    free(op->rowptr); free(op->col); free(op->neigh);
    free(op);
}

void orc_export_blocks(const orc_op *op, double *out)
{
    const int nln = op->nln, d = 2 * nln;
    memset(out, 0, sizeof(double) * (size_t)op->ne * 4 * d * d);
    for (int e = 0; e < op->ne; ++e)
        for (int s = 0; s < 4; ++s) {
            const int c = (s == 0) ? e : op->neigh[3 * e + s - 1];
            if (c < 0) continue;
            const int b = find_block(op, e, c);
            double *o = out + ((size_t)e * 4 + s) * d * d;
            for (int m = 0; m < 2; ++m)
                for (int n = 0; n < 2; ++n) {
                    const double *src = op->B[m * 2 + n] + (size_t)b * nln * nln;
                    for (int i = 0; i < nln; ++i)
                        for (int j = 0; j < nln; ++j) o[(m * nln + i) * d + n * nln + j] = src[i * nln + j];
                }
        }
}

void orc_export_invm(const orc_op *op, double *out) { memcpy(out, op->invm, sizeof(double) * op->ne); }

/* BaseMat::vecMult (MyMat.cpp:200-218) for one component matrix */
static void vecmult(const orc_op *op, const double *M, const double *x, double *out)
{
    const int nln = op->nln;
    ORC_PAR
    for (int r = 0; r < op->ne; ++r) {
        double *o = out + (size_t)r * nln;
        for (int b = op->rowptr[r]; b < op->rowptr[r + 1]; ++b) {
            const double *blkp = M + (size_t)b * nln * nln, *xs = x + (size_t)op->col[b] * nln;
            for (int i = 0; i < nln; ++i) {
                double acc = 0.0;
                for (int j = 0; j < nln; ++j) acc += blkp[i * nln + j] * xs[j];
                o[i] = (b == op->rowptr[r]) ? acc : o[i] + acc;
            }
        }
    }
}

/* MyMatMultiDim::vecMult (MyMat.hpp:311-319) */
void orc_apply_B(const orc_op *op, const double *x, double *out)
{
    const size_t hs = (size_t)op->ne * op->nln;
    double *t0 = malloc(sizeof(double) * hs), *t1 = malloc(sizeof(double) * hs);
    vecmult(op, op->B[0], x, t0);
    vecmult(op, op->B[1], x + hs, t1);
    ORC_PAR
    for (size_t q = 0; q < hs; ++q) out[q] = t0[q] + t1[q];
    vecmult(op, op->B[2], x, t0);
    vecmult(op, op->B[3], x + hs, t1);
    ORC_PAR
    for (size_t q = 0; q < hs; ++q) out[hs + q] = t0[q] + t1[q];
    free(t0);
    free(t1);
}

/* MyMatMultiDimBlockDiag::vecMult (MyMat.hpp:304-310): dense block products */
static void apply_invM(const orc_op *op, const double *x, double *out)
{
    const int nln = op->nln;
    const size_t hs = (size_t)op->ne * nln;
    for (int c = 0; c < 2; ++c) {
        ORC_PAR
        for (int e = 0; e < op->ne; ++e) {
            const double *M = op->invM[c] + (size_t)e * nln * nln, *xs = x + c * hs + (size_t)e * nln;
            double *o = out + c * hs + (size_t)e * nln;
            for (int i = 0; i < nln; ++i) {
                double acc = 0.0;
                for (int j = 0; j < nln; ++j) acc += M[i * nln + j] * xs[j];
                o[i] = acc;
            }
        }
    }
}

/* MyMatMultiDimBlockDiag<MyMat>::operator* (MyMat.hpp:337-342) */
void orc_apply_invM(const orc_op *op, const double *x, double *out) { apply_invM(op, x, out); }

double orc_ricker(double f0, double t0, double t)
{
    const double f0sq = f0 * f0;
    return (1 - 2 * ORC_PI * ORC_PI * f0sq * (t - t0) * (t - t0)) * exp(-ORC_PI * ORC_PI * f0sq * (t - t0) * (t - t0));
}

/* PointWiseForce::eval (Force.cpp:22-30): dense vector, 2*nln non-zeros */
static void force_eval(const orc_op *op, const orc_source *src, double t, double *f)
{
    const size_t hs = (size_t)op->ne * op->nln;
    memset(f, 0, sizeof(double) * 2 * hs);
    if (!src || src->elem < 0) return;
    const double r = orc_ricker(src->f0, src->t0, t);
    const double F0 = src->dirx * r, F1 = src->diry * r;
    for (int i = 0; i < op->nln; ++i) {
        f[(size_t)src->elem * op->nln + i] = F0 * src->phi[i];
        f[hs + (size_t)src->elem * op->nln + i] = F1 * src->phi[i];
    }
}

/* TimeAdvance::eval_receivers (TimeAdvance.cpp:68-79) */
static void sample(const orc_op *op, int nrec, const int *re, const double *rphi, const double *uh, double *out)
{
    const int nln = op->nln;
    const size_t hs = (size_t)op->ne * nln;
    for (int r = 0; r < nrec; ++r) {
        double x = 0, y = 0;
        for (int i = 0; i < nln; ++i) {
            x += rphi[r * nln + i] * uh[(size_t)re[r] * nln + i];
            y += rphi[r * nln + i] * uh[hs + (size_t)re[r] * nln + i];
        }
        out[2 * r] = x;
        out[2 * r + 1] = y;
    }
}

/* LeapFrog::first_step + step (TimeAdvance.cpp:24-59) with update_variables
 * (TimeAdvance.hpp:230).  The force of step() is taken at t_{n+1} = t_prev + dt,
 * with time accumulated by repeated addition (TimeAdvance.cpp:48-51). */
int orc_run(const orc_op *op, double dt, int nsteps, const orc_source *src,
            int nrec, const int *rec_elem, const double *rec_phi, int stride,
            const double *u0, const double *v0, double *u_out, double *seis)
{
    const size_t n = 2 * (size_t)op->ne * op->nln;
    double *uhold = malloc(sizeof(double) * n), *uholdold = calloc(n, sizeof(double));
    double *uh = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n);
    double *Bu = malloc(sizeof(double) * n), *w = calloc(n, sizeof(double)), *iw = malloc(sizeof(double) * n);
    if (u0) memcpy(uholdold, u0, sizeof(double) * n);
    int ns = 0;
    /* first_step */
    force_eval(op, src, 0., f);
    orc_apply_B(op, uholdold, Bu);
    for (size_t q = 0; q < n; ++q) w[q] = dt * dt / 2 * (Bu[q] - f[q]);
    apply_invM(op, w, iw);
    for (size_t q = 0; q < n; ++q) uhold[q] = uholdold[q] + dt * (v0 ? v0[q] : 0.0) - iw[q];
    double last_time = dt;
    if (nrec > 0 && seis) sample(op, nrec, rec_elem, rec_phi, uhold, seis + (size_t)(ns++) * nrec * 2);
    /* step */
    for (int s = 1; s <= nsteps; ++s) {
        const double t = last_time + dt;
        force_eval(op, src, t, f);
        orc_apply_B(op, uhold, Bu);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) w[q] = -dt * dt * (Bu[q] - f[q]);
        apply_invM(op, w, iw);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) uh[q] = 2 * uhold[q] - uholdold[q] + iw[q];
        double *tmp = uholdold;
        uholdold = uhold;
        uhold = uh;
        uh = tmp;
        last_time = t;
        if (nrec > 0 && seis && stride > 0 && s % stride == 0)
            sample(op, nrec, rec_elem, rec_phi, uhold, seis + (size_t)(ns++) * nrec * 2);
    }
    if (u_out) memcpy(u_out, uhold, sizeof(double) * n);
    free(uhold); free(uholdold); free(uh); free(f); free(Bu); free(w); free(iw);
    return ns;
}

/* Power iteration for the largest eigenvalue of invM*B (a check on the step
 * size; the reference itself has none). */
double orc_lambda_max(const orc_op *op, int iters, unsigned long long seed)
{
    const size_t n = 2 * (size_t)op->ne * op->nln;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    uint64_t st = seed ? seed : 88172645463325252ull;
    for (size_t q = 0; q < n; ++q) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        x[q] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
    }
    double lam = 0;
    for (int it = 0; it < iters; ++it) {
        double nrm = 0;
        for (size_t q = 0; q < n; ++q) nrm += x[q] * x[q];
        nrm = sqrt(nrm);
        for (size_t q = 0; q < n; ++q) x[q] /= nrm;
        orc_apply_B(op, x, y);
        apply_invM(op, y, z);
        lam = 0;
        for (size_t q = 0; q < n; ++q) lam += z[q] * z[q];
        lam = sqrt(lam); /* |invM*B*x| with |x| = 1 tends to the dominant eigenvalue */
        double *t = x; x = z; z = t;
    }
    free(x); free(y); free(z);
    return lam;
}

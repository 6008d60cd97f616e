// assemble.hip -- FP64 geometry and operator assembly on the device.
//
// Restates Matrices::Matrices (Matrices_imp.hpp:192-215) and the combination
// B = penalty*N^2*S + I + A (TimeAdvance_imp.hpp:32) as two kernels:
//   k_geometry  one thread per element: J^{-1}, detJ, inverse mass scalar, and per
//               face the unit normal, length and penalty coefficients
//               (Geometry.cpp:100-105,178-198; Parameters.cpp:37-45;
//               Matrices_imp.hpp:157-173,187).
//   k_blocks    one work-group per element; each thread owns entries of the
//               element's four d x d blocks and sums their quadrature terms in
//               the same order as the reference loops (area nodes k; then faces
//               f = 0,1,2 and edge nodes k), so that every entry is formed by the
//               same sequence of FP64 operations.  The global symmetrisation
//               I <- I + I^T (MyMat.hpp:321-328; MyMat.cpp:156-187) is folded in:
//               the thread forming B[e,c](I,J) also forms the consistency term
//               of the neighbour's row, I[c,e](J,I), from the neighbour's side.
// Floating-point contraction is off in this file: no FMA is formed behind the
// source's back, so each product and sum rounds as written.
#include "engine.hpp"

#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

__device__ inline double avg_p(double a, double b) { return (a + b == 0) ? 0 : 2 * a * b / (a + b); }

__global__ void k_geometry(AssembleArgs a)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne) return;
    const double *P = a.verts + 6 * e;
    const double J00 = P[2] - P[0], J10 = P[3] - P[1];
    const double J01 = P[4] - P[0], J11 = P[5] - P[1];
    const double det = J00 * J11 - J10 * J01;
    const double invdet = 1.0 / det;
    ElemGeo g;
    g.inv[0] = J11 * invdet;
    g.inv[2] = -J10 * invdet;
    g.inv[1] = -J01 * invdet;
    g.inv[3] = J00 * invdet;
    g.det = det;
    g.invm = 1. / a.rho[e] * 1. / det;
    a.geo[e] = g;
    a.invm[e] = g.invm;
    for (int f = 0; f < 3; ++f) {
        const double ax = P[2 * ((f + 1) % 3)], ay = P[2 * ((f + 1) % 3) + 1];
        const double bx = P[2 * ((f + 2) % 3)], by = P[2 * ((f + 2) % 3) + 1];
        const double nx = -ay + by, ny = ax - bx;
        const double nn = sqrt(nx * nx + ny * ny);
        const double dx = ax - bx, dy = ay - by;
        FaceGeo fg;
        fg.n0 = nx / nn;
        fg.n1 = ny / nn;
        fg.len = sqrt(dx * dx + dy * dy);
        const int nb = a.neigh[3 * e + f];
        if (nb >= 0) {
            const double al = avg_p(a.lambda[e], a.lambda[nb]), am = avg_p(a.mu[e], a.mu[nb]);
            const double h_scaled = 1. / fg.len;
            const double c00 = (al + 2 * am) * fg.n0 * fg.n0 + 2 * am * fg.n1 * fg.n1;
            const double c11 = (al + 2 * am) * fg.n1 * fg.n1 + 2 * am * fg.n0 * fg.n0;
            const double c01 = al * fg.n0 * fg.n1;
            fg.ch00 = c00 * h_scaled;
            fg.ch11 = c11 * h_scaled;
            fg.ch01 = c01 * h_scaled;
        } else {
            fg.ch00 = fg.ch11 = fg.ch01 = 0;
        }
        a.face[3 * e + f] = fg;
    }
}

// Physical gradient g = (grad_ref)^T * invJ (Matrices_imp.hpp:36,59)
__device__ inline void phys_grad(double gx, double gy, const double *inv, double *g)
{
    g[0] = gx * inv[0] + gy * inv[2];
    g[1] = gx * inv[1] + gy * inv[3];
}
// strain of phi e_n (Matrices_imp.hpp:37-38,60-61), 2x2 row-major
__device__ inline void strain(const double *g, int n, double *e)
{
    if (n == 0) { e[0] = g[0]; e[1] = .5 * g[1]; e[2] = .5 * g[1]; e[3] = 0; }
    else        { e[0] = 0.;   e[1] = .5 * g[0]; e[2] = .5 * g[0]; e[3] = g[1]; }
}
// isotropic Hooke tensor (TimeAdvance.cpp:60-65; Matrices_imp.hpp:64-65)
__device__ inline void hooke(const double *e, double lambda, double mu, double *s)
{
    s[0] = (lambda + 2 * mu) * e[0] + lambda * e[3];
    s[1] = 2 * mu * e[1];
    s[2] = 2 * mu * e[1];
    s[3] = (lambda + 2 * mu) * e[3] + lambda * e[0];
}
__device__ inline double mat_dot(const double *a, const double *b)  // TimeAdvance.cpp:66
{
    return a[0] * b[0] + a[2] * b[2] + a[1] * b[1] + a[3] * b[3];
}
// (sigma(phi_i e_m) n)_n at edge node k of edge f of an element (Matrices_imp.hpp:127)
template <int P>
__device__ inline double sig_n(const Tables &t, const double *inv, double lam, double mu, int k, int i, int f,
                               int m, int n, double n0, double n1)
{
    using G = Geom<P>;
    const double gx = t.Ge[((i * 3 + f) * 2) * G::NQ1 + k], gy = t.Ge[((i * 3 + f) * 2 + 1) * G::NQ1 + k];
    double g[2], ep[4], s[4];
    phys_grad(gx, gy, inv, g);
    strain(g, m, ep);
    hooke(ep, lam, mu, s);
    return s[n * 2 + 0] * n0 + s[n * 2 + 1] * n1;
}

template <int P>
__global__ __launch_bounds__(256) void k_blocks(AssembleArgs a)
{
    using G = Geom<P>;
    constexpr int NLN = G::NLN, D = G::D, DH = G::DH, NQ1 = G::NQ1, NQ2 = G::NQ2;
    const int e = blockIdx.x;
    const Tables t = a.tab;
    const ElemGeo ge = a.geo[e];
    const double lam = a.lambda[e], mu = a.mu[e];
    for (int q = threadIdx.x; q < 4 * D * D; q += blockDim.x) {
        // q enumerates the device layout: [s][jp][row][t]
        const int tt = q & 1;
        const int row = (q >> 1) % D;
        const int jp = ((q >> 1) / D) % DH;
        const int s = q / (D * D);
        const int I = row, J = 2 * jp + tt;
        const int m = I / NLN, i = I % NLN, n = J / NLN, j = J % NLN;
        double out = 0.0;
        if (s == 0) {
            // stress_loc (Matrices_imp.hpp:73-94)
            double A = 0.0;
            for (int k = 0; k < NQ2; ++k) {
                const double dx = fabs(ge.det) * t.w2[k];
                double gj[2], gi[2], ej[4], sj[4], ei[4];
                phys_grad(t.G[(j * 2) * NQ2 + k], t.G[(j * 2 + 1) * NQ2 + k], ge.inv, gj);
                phys_grad(t.G[(i * 2) * NQ2 + k], t.G[(i * 2 + 1) * NQ2 + k], ge.inv, gi);
                strain(gj, n, ej);
                hooke(ej, lam, mu, sj);
                strain(gi, m, ei);
                A += mat_dot(sj, ei) * dx;
            }
            // stability, diagonal part (Matrices_imp.hpp:167-169,179) and
            // interelement, diagonal part (:127) for (m,n,i,j) and its transpose (n,m,j,i)
            double S = 0.0, Iv = 0.0, It = 0.0;
            for (int f = 0; f < 3; ++f) {
                if (a.neigh[3 * e + f] < 0) continue;
                const FaceGeo fg = a.face[3 * e + f];
                const double ch = (m == 0 && n == 0) ? fg.ch00 : ((m == 1 && n == 1) ? fg.ch11 : fg.ch01);
                for (int k = 0; k < NQ1; ++k) {
                    const double ds = fg.len * t.w1[k];
                    const double *bk = t.Be + (f * NQ1 + k) * NLN;
                    S += ch * bk[i] * bk[j] * ds;
                    const double sv = sig_n<P>(t, ge.inv, lam, mu, k, i, f, m, n, fg.n0, fg.n1);
                    Iv -= 0.5 * sv * bk[j] * ds;
                    const double st = sig_n<P>(t, ge.inv, lam, mu, k, j, f, n, m, fg.n0, fg.n1);
                    It -= 0.5 * st * bk[i] * ds;
                }
            }
            out = a.cpen * S + (Iv + It);
            out += A;
        } else {
            const int f = s - 1;
            const int c = a.neigh[3 * e + f];
            if (c >= 0) {
                const int gedge = a.neighedge[3 * e + f];
                const FaceGeo fg = a.face[3 * e + f];
                const FaceGeo fc = a.face[3 * c + gedge];
                const ElemGeo gc = a.geo[c];
                const double lamc = a.lambda[c], muc = a.mu[c];
                const double ch = (m == 0 && n == 0) ? fg.ch00 : ((m == 1 && n == 1) ? fg.ch11 : fg.ch01);
                double S = 0.0, Iv = 0.0, It = 0.0;
                for (int k = 0; k < NQ1; ++k) {
                    const int kk = NQ1 - k - 1;
                    const double ds = fg.len * t.w1[k];
                    const double *bk = t.Be + (f * NQ1 + k) * NLN, *bn = t.Be + (gedge * NQ1 + kk) * NLN;
                    S -= ch * bk[i] * bn[j] * ds;
                    const double sv = sig_n<P>(t, ge.inv, lam, mu, k, i, f, m, n, fg.n0, fg.n1);
                    Iv += 0.5 * sv * bn[j] * ds;
                }
                // I[c,e](n,m; j,i) from the neighbour's side
                for (int k = 0; k < NQ1; ++k) {
                    const int kk = NQ1 - k - 1;
                    const double ds = fc.len * t.w1[k];
                    const double *bn = t.Be + (f * NQ1 + kk) * NLN;
                    const double st = sig_n<P>(t, gc.inv, lamc, muc, k, j, gedge, n, m, fc.n0, fc.n1);
                    It += 0.5 * st * bn[i] * ds;
                }
                out = a.cpen * S + (Iv + It);
            }
        }
        a.blk[((size_t)e * 4 + s) * D * D + q % (D * D)] = out;
    }
}

template <int P>
__global__ void k_to_canonical(int ne, const double *blk, double *canon)
{
    using G = Geom<P>;
    constexpr int D = G::D, DH = G::DH;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)ne * 4 * D * D) return;
    const size_t b = q / (D * D);
    const int r = q % (D * D);
    const int tt = r & 1, row = (r >> 1) % D, jp = (r >> 1) / D;
    (void)DH;
    canon[b * D * D + (size_t)row * D + 2 * jp + tt] = blk[q];
}

template <int P>
__global__ void k_from_canonical(int ne, const double *canon, double *blk)
{
    using G = Geom<P>;
    constexpr int D = G::D;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)ne * 4 * D * D) return;
    const size_t b = q / (D * D);
    const int r = q % (D * D);
    const int tt = r & 1, row = (r >> 1) % D, jp = (r >> 1) / D;
    blk[q] = canon[b * D * D + (size_t)row * D + 2 * jp + tt];
}

// FP32 copy of the operator in the column-quad layout (engine.hpp).
template <int P>
__global__ void k_demote_quads(int ne, const double *src, float *dst)
{
    constexpr int D = Geom<P>::D, NFULL = D / 4;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)ne * 4 * D * D) return;
    const size_t b = q / (D * D);
    const int r = q % (D * D);
    int row, col;
    if (r < NFULL * D * 4) {
        const int jq = r / (4 * D), rem = r % (4 * D);
        row = rem / 4;
        col = 4 * jq + rem % 4;
    } else {
        const int rr = r - NFULL * D * 4;
        row = rr / 2;
        col = 4 * NFULL + rr % 2;
    }
    dst[q] = (float)src[b * D * D + (size_t)(col / 2) * 2 * D + row * 2 + (col & 1)];
}

// Non-orthonormal basis: fold base_invmass into the block rows, B' = (I2 (x) Mb^-1) B,
// so the step's per-element scalar 1/(rho detJ) completes the dense inverse mass
// (Matrices_imp.hpp:183-190).  One thread per column of one block, in place.
template <int P>
__global__ void k_premul_invmass(int ne, const double *minv, double *blk)
{
    constexpr int D = Geom<P>::D, NLN = Geom<P>::NLN;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)ne * 4 * D) return;
    const size_t b = q / D;
    const int J = (int)(q % D);
    double *col = blk + b * D * D + (size_t)(J / 2) * 2 * D + (J & 1);
    double v[D];
#pragma unroll
    for (int r = 0; r < D; ++r) v[r] = col[r * 2];
    for (int r = 0; r < D; ++r) {
        const int m = r / NLN, i = r % NLN;
        double acc = 0.0;
        for (int k = 0; k < NLN; ++k) acc += minv[i * NLN + k] * v[m * NLN + k];
        col[r * 2] = acc;
    }
}

// Uploaded operators: the blocks of absent neighbours are set to zero, as the
// device assembly leaves them (the step kernels read every slot).
template <int P>
__global__ void k_zero_missing(int ne, const int *neigh, double *blk)
{
    constexpr int D = Geom<P>::D;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)ne * 3 || neigh[q] >= 0) return;
    double *b = blk + ((q / 3) * 4 + 1 + q % 3) * D * D;
    for (int k = 0; k < D * D; ++k) b[k] = 0.0;
}

template <int P>
int zero_missing_p(int ne, const int *neigh, double *blk, hipStream_t s)
{
    const size_t n = (size_t)ne * 3;
    hipLaunchKernelGGL(k_zero_missing<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, neigh, blk);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int premul_p(int ne, const double *minv, double *blk, hipStream_t s)
{
    const size_t n = (size_t)ne * 4 * Geom<P>::D;
    hipLaunchKernelGGL(k_premul_invmass<P>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, ne, minv, blk);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int assemble_p(const AssembleArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_geometry, dim3((a.ne + 127) / 128), dim3(128), 0, s, a);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_blocks<P>, dim3(a.ne), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int conv_p(bool to_canon, int ne, const double *src, double *dst, hipStream_t s)
{
    constexpr int D = Geom<P>::D;
    const size_t n = (size_t)ne * 4 * D * D;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (to_canon) hipLaunchKernelGGL(k_to_canonical<P>, dim3(nb), dim3(256), 0, s, ne, src, dst);
    else          hipLaunchKernelGGL(k_from_canonical<P>, dim3(nb), dim3(256), 0, s, ne, src, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

#define TSP_DISPATCH_P(P, CALL)                                   \
    switch (P) {                                                  \
    case 1: return CALL(1); case 2: return CALL(2);               \
    case 3: return CALL(3); case 4: return CALL(4);               \
    case 5: return CALL(5); case 6: return CALL(6);               \
    case 7: return CALL(7); case 8: return CALL(8);               \
    default: return -2;                                           \
    }

int launch_assemble(int P, const AssembleArgs &a, hipStream_t s)
{
#define CALL(p) assemble_p<p>(a, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_zero_missing(int P, int ne, const int *neigh, double *blk, hipStream_t s)
{
#define CALL(p) zero_missing_p<p>(ne, neigh, blk, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_premul_invmass(int P, int ne, const double *minv, double *blk, hipStream_t s)
{
#define CALL(p) premul_p<p>(ne, minv, blk, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_blocks_to_canonical(int P, int ne, const double *blk, double *canon, hipStream_t s)
{
#define CALL(p) conv_p<p>(true, ne, blk, canon, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_canonical_to_blocks(int P, int ne, const double *canon, double *blk, hipStream_t s)
{
#define CALL(p) conv_p<p>(false, ne, canon, blk, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

template <int P>
int demote_p(int ne, const double *src, float *dst, hipStream_t s)
{
    constexpr int D = Geom<P>::D;
    const size_t n = (size_t)ne * 4 * D * D;
    hipLaunchKernelGGL(k_demote_quads<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, src, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_demote_blocks(int P, int ne, const double *src, float *dst, hipStream_t s)
{
#define CALL(p) demote_p<p>(ne, src, dst, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

}  // namespace tspeed_dev

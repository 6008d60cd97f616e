// step.hip -- the fused leap-frog step on gfx950.
//
// One 64-lane wave per element, four elements per 256-thread work-group.  A
// step restates LeapFrog::step (TimeAdvance.cpp:41-59) for one element row:
//   r   = sum_s B_{e,s} u^n_{nb(e,s)}           block-ELL SpMV (MyMat.cpp:200-218)
//   w   = coef * (r - f)                          f = point source (Force.cpp:22-30)
//   u^{n+1} = (2u^n - u^{n-1}) + invm_e * w       invM is invm_e * I for the
//                                                 orthonormal Dubiner basis
//                                                 (Matrices_imp.hpp:187; FESpace_imp.hpp:51-54)
// followed by the receiver dot products (TimeAdvance.cpp:68-79) on the freshly
// written coefficients.  u^{n+1} overwrites u^{n-1} in place (only the element
// itself reads its own u^{n-1}), so two vectors carry the state and the
// reference's four copies per step (update_variables, TimeAdvance.hpp:230)
// become a pointer swap.
//
// SpMV lane mapping: lane = block row; the wave walks the block's column pairs,
// each lane loading its row's two entries with one 16-byte (FP64) / 8-byte (FP32)
// load, so one wave instruction reads 2*d contiguous values.  For d <= 32 the
// wave is split into 2 or 4 column groups that walk interleaved column pairs
// and are summed with cross-lane shuffles; for d > 64 each lane owns two rows.
// The neighbour coefficients are staged in LDS and read as broadcasts.
//
// Elements are dealt to work-groups so that each XCD (work-groups are
// dispatched round-robin over the 8 XCDs) processes one contiguous range of
// elements, keeping neighbour gathers within that XCD's L2.
#include "engine.hpp"

// Products and sums in the update round as written (no FMA contraction); the
// SpMV accumulates with explicit fma().
#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

constexpr int kWaves = 4;  // waves (elements) per work-group

__device__ inline int xcd_remap(int bid, int nwg)
{
    const int per = nwg / 8, rem = nwg % 8;
    const int x = bid % 8, k = bid / 8;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum of the source terms landing on row `row` of element e (Force.cpp:27-28).
__device__ inline double source_term(const StepArgs &a, int4 ex, int row, int nln)
{
    double f = 0.0;
    const int c = row / nln, i = row % nln;
    for (int k = ex.x; k < ex.y; ++k) {
        const int src = a.src_sorted[k];
        f += a.wav_row[src * 2 + c] * a.src_phi[src * nln + i];
    }
    return f;
}

// Receiver dot products for element e on the new coefficients held one (or two)
// rows per lane: seis[sample][rec] = (phi . u_x, phi . u_y) (TimeAdvance.cpp:73-74).
template <int P, typename V>
__device__ inline void record(const StepArgs &a, int4 ex, const V *unew, const int *rows, int nrows_lane)
{
    using G = Geom<P>;
    for (int k = ex.z; k < ex.w; ++k) {
        const int rec = a.rec_sorted[k];
        double sx = 0.0, sy = 0.0;
        for (int q = 0; q < nrows_lane; ++q) {
            const int row = rows[q];
            if (row < 0 || row >= G::D) continue;
            const int c = row / G::NLN, i = row % G::NLN;
            const double v = a.rec_phi[rec * G::NLN + i] * (double)unew[q];
            if (c == 0) sx += v; else sy += v;
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        if ((threadIdx.x & 63) == 0) {
            a.seis[((size_t)a.sample * a.nrec + rec) * 2 + 0] = sx;
            a.seis[((size_t)a.sample * a.nrec + rec) * 2 + 1] = sy;
        }
    }
}

// Gather u^n of the element and its neighbours into the wave's LDS slice.
template <int P, typename T>
__device__ inline void gather(T (*xs)[Geom<P>::D], const T *u, int e, int4 m, int lane, bool act)
{
    using G = Geom<P>;
    const int nbs[4] = {act ? e : -1, m.x, m.y, m.z};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (nbs[s] < 0) continue;
        for (int idx = lane; idx < G::D; idx += 64) xs[s][idx] = u[(size_t)nbs[s] * G::D + idx];
    }
}

// r = sum_s B_{e,s} x_s for the lane's rows.  Returns partial sums per row slot;
// for NG > 1 the column groups are reduced into group 0.
template <int P, typename T, typename T2>
__device__ inline void spmv(const T *blk, int e, int4 m, const T (*xs)[Geom<P>::D], int lane,
                            T (&acc)[Geom<P>::RPL])
{
    using G = Geom<P>;
    const int g = lane / G::RW, r = lane % G::RW;
    const int nbs[4] = {e, m.x, m.y, m.z};
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = T(0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (nbs[s] < 0) continue;
        const T2 *Bs = reinterpret_cast<const T2 *>(blk + ((size_t)e * 4 + s) * G::D * G::D);
#pragma unroll
        for (int it = 0; it < G::NIT; ++it) {
            const int jp = it * G::NG + g;
            if (G::NG > 1 && jp >= G::DH) continue;
            const T x0 = xs[s][2 * jp], x1 = xs[s][2 * jp + 1];
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                const int row = r + 64 * q;
                if (row < G::D) {
                    const T2 b = Bs[jp * G::D + row];
                    acc[q] = fma(b.x, x0, acc[q]);
                    acc[q] = fma(b.y, x1, acc[q]);
                }
            }
        }
    }
    if constexpr (G::NG > 1) {
#pragma unroll
        for (int o = 32; o >= G::RW; o >>= 1) acc[0] += __shfl_down(acc[0], o);
    }
}

}  // namespace

// ---------------------------------------------------------------- FP64 plain
template <int P, bool FIRST>
__global__ __launch_bounds__(256) void k_step64(StepArgs a)
{
    using G = Geom<P>;
    __shared__ double xs_all[kWaves][4][G::D];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    const bool act = e < a.ne;
    const int4 m = act ? a.meta[e] : make_int4(-1, -1, -1, -1);
    double(*xs)[G::D] = xs_all[wave];
    gather<P, double>(xs, a.ucur, e, m, lane, act);
    __syncthreads();
    if (!act) return;

    double acc[G::RPL];
    spmv<P, double, double2>(a.blk64, e, m, xs, lane, acc);

    const int g = lane / G::RW, r = lane % G::RW;
    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    const double invm = a.invm[e];
    double unew[G::RPL];
    int rows[G::RPL];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (g == 0 && row < G::D) ? row : -1;
        unew[q] = 0.0;
        if (rows[q] < 0) continue;
        const size_t idx = (size_t)e * G::D + row;
        const double f = (ex.y > ex.x) ? source_term(a, ex, row, G::NLN) : 0.0;
        const double w = a.coef * (acc[q] - f);
        double un;
        if (FIRST) {
            // u1 = u0 + dt*v0 - invM*(dt*dt/2*(B*u0 - f(0)))  (TimeAdvance.cpp:30)
            un = xs[0][row] + a.dt * a.uprev[idx] - invm * w;
        } else {
            un = 2 * xs[0][row] - a.uprev[idx] + invm * w;
        }
        a.uprev[idx] = un;
        unew[q] = un;
        bad |= !isfinite(un);
    }
    if (bad) atomicOr(a.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<P, double>(a, ex, unew, rows, G::RPL);
}

// ---------------------------------------------------- FP32 compensated update
// Velocity form of the same recurrence: w^{n+1/2} = w^{n-1/2} + sc_e*(B u^n - f)
// with sc_e = invm_e*(-dt^2), and u^{n+1} = u^n + w^{n+1/2} accumulated by Kahan
// summation (hi part uhi, compensation ucomp, true value = uhi - ucomp).  The
// block products and the increment are FP32; the compensation keeps the
// displacement's roundoff from drifting over thousands of steps.
template <int P, bool FIRST>
__global__ __launch_bounds__(256) void k_step32c(StepArgs a)
{
    using G = Geom<P>;
    __shared__ float xs_all[kWaves][4][G::D];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    const bool act = e < a.ne;
    const int4 m = act ? a.meta[e] : make_int4(-1, -1, -1, -1);
    float(*xs)[G::D] = xs_all[wave];
    gather<P, float>(xs, a.uhi, e, m, lane, act);
    __syncthreads();
    if (!act) return;

    float acc[G::RPL];
    spmv<P, float, float2>(a.blk32, e, m, xs, lane, acc);

    const int g = lane / G::RW, r = lane % G::RW;
    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    const float sc = a.sc32[e];
    float unew[G::RPL];
    int rows[G::RPL];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (g == 0 && row < G::D) ? row : -1;
        unew[q] = 0.f;
        if (rows[q] < 0) continue;
        const size_t idx = (size_t)e * G::D + row;
        const float f = (ex.y > ex.x) ? (float)source_term(a, ex, row, G::NLN) : 0.f;
        const float acl = sc * (acc[q] - f);
        // FIRST: winc holds dt*v0 (prepared on the host side of the ABI) and
        // sc is invm*(-dt^2/2): w^{1/2} = dt*v0 - invm*dt^2/2*(B u0 - f(0)).
        const float w = a.winc[idx] + acl;
        const float uh = xs[0][row];
        const float y = w - a.ucomp[idx];
        const float t = uh + y;
        a.ucomp[idx] = (t - uh) - y;
        a.uhi_out[idx] = t;
        a.winc[idx] = w;
        unew[q] = t;
        bad |= !isfinite(t);
    }
    if (bad) atomicOr(a.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<P, float>(a, ex, unew, rows, G::RPL);
}

// Receiver sample of a stored state, outside a step (eval_receivers):
// one wave per receiver.
template <int P, typename T>
__global__ __launch_bounds__(64) void k_sample(StepArgs a, const T *u, const T *comp)
{
    using G = Geom<P>;
    const int rec = blockIdx.x;
    if (rec >= a.nrec) return;
    const int lane = threadIdx.x;
    const int re = a.rec_elem[rec];
    double sx = 0.0, sy = 0.0;
    for (int row = lane; row < G::D; row += 64) {
        const int c = row / G::NLN, i = row % G::NLN;
        double uv = (double)u[(size_t)re * G::D + row];
        if (comp) uv -= (double)comp[(size_t)re * G::D + row];
        const double v = a.rec_phi[rec * G::NLN + i] * uv;
        if (c == 0) sx += v; else sy += v;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) {
        a.seis[((size_t)a.sample * a.nrec + rec) * 2 + 0] = sx;
        a.seis[((size_t)a.sample * a.nrec + rec) * 2 + 1] = sy;
    }
}

__global__ void k_split(size_t n, const double *u, float *hi, float *comp)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float h = (float)u[q];
    hi[q] = h;
    comp[q] = (float)((double)h - u[q]);
}

__global__ void k_join(size_t n, const float *hi, const float *comp, double *u)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) u[q] = (double)hi[q] - (double)comp[q];
}

__global__ void k_scale_demote(size_t n, const double *src, double coef, float *dst)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (float)(src[q] * coef);
}

// y = invM * B * x for the power iteration (FP64).
template <int P>
__global__ __launch_bounds__(256) void k_apply_minvB(int ne, const int4 *meta, const double *blk,
                                                     const double *invm, const double *x, double *y)
{
    using G = Geom<P>;
    __shared__ double xs_all[kWaves][4][G::D];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    const bool act = e < ne;
    const int4 m = act ? meta[e] : make_int4(-1, -1, -1, -1);
    double(*xs)[G::D] = xs_all[wave];
    gather<P, double>(xs, x, e, m, lane, act);
    __syncthreads();
    if (!act) return;
    double acc[G::RPL];
    spmv<P, double, double2>(blk, e, m, xs, lane, acc);
    const int g = lane / G::RW, r = lane % G::RW;
    const double iv = invm[e];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        if (g == 0 && row < G::D) y[(size_t)e * G::D + row] = iv * acc[q];
    }
}

__global__ void k_dot(size_t n, const double *x, const double *y, double *partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        s += x[q] * y[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_cm_to_em(int ne, int nln, const double *cm, double *em)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)2 * ne * nln;
    if (q >= n) return;
    const int c = (int)(q / ((size_t)ne * nln));
    const size_t rem = q % ((size_t)ne * nln);
    const size_t e = rem / nln;
    const int i = (int)(rem % nln);
    em[e * 2 * nln + c * nln + i] = cm[q];
}

__global__ void k_em_to_cm(int ne, int nln, const double *em, double *cm)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)2 * ne * nln;
    if (q >= n) return;
    const int c = (int)(q / ((size_t)ne * nln));
    const size_t rem = q % ((size_t)ne * nln);
    const size_t e = rem / nln;
    const int i = (int)(rem % nln);
    cm[q] = em[e * 2 * nln + c * nln + i];
}

namespace {
inline unsigned nwg_for(int ne) { return (unsigned)((ne + kWaves - 1) / kWaves); }

template <int P>
int step64_p(bool first, const StepArgs &a, hipStream_t s)
{
    if (first) hipLaunchKernelGGL((k_step64<P, true>), dim3(nwg_for(a.ne)), dim3(256), 0, s, a);
    else       hipLaunchKernelGGL((k_step64<P, false>), dim3(nwg_for(a.ne)), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int step32c_p(bool first, const StepArgs &a, hipStream_t s)
{
    if (first) hipLaunchKernelGGL((k_step32c<P, true>), dim3(nwg_for(a.ne)), dim3(256), 0, s, a);
    else       hipLaunchKernelGGL((k_step32c<P, false>), dim3(nwg_for(a.ne)), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int sample64_p(const StepArgs &a, const double *u, hipStream_t s)
{
    hipLaunchKernelGGL((k_sample<P, double>), dim3(a.nrec), dim3(64), 0, s, a, u, (const double *)nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int sample32_p(const StepArgs &a, const float *hi, const float *comp, hipStream_t s)
{
    hipLaunchKernelGGL((k_sample<P, float>), dim3(a.nrec), dim3(64), 0, s, a, hi, comp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int minvb_p(int ne, const int4 *meta, const double *blk, const double *invm, const double *x, double *y,
            hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_minvB<P>, dim3(nwg_for(ne)), dim3(256), 0, s, ne, meta, blk, invm, x, y);
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

int launch_step64(int P, bool first, const StepArgs &a, hipStream_t s)
{
#define CALL(p) step64_p<p>(first, a, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_step32c(int P, bool first, const StepArgs &a, hipStream_t s)
{
#define CALL(p) step32c_p<p>(first, a, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_sample64(int P, const StepArgs &a, const double *u, hipStream_t s)
{
#define CALL(p) sample64_p<p>(a, u, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_sample32(int P, const StepArgs &a, const float *hi, const float *comp, hipStream_t s)
{
#define CALL(p) sample32_p<p>(a, hi, comp, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_apply_minvB(int P, int ne, const int4 *meta, const double *blk, const double *invm, const double *x,
                       double *y, hipStream_t s)
{
#define CALL(p) minvb_p<p>(ne, meta, blk, invm, x, y, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_split_state(int ne, int D, const double *u64, float *hi, float *comp, hipStream_t s)
{
    const size_t n = (size_t)ne * D;
    hipLaunchKernelGGL(k_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, u64, hi, comp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_join_state(int ne, int D, const float *hi, const float *comp, double *u64, hipStream_t s)
{
    const size_t n = (size_t)ne * D;
    hipLaunchKernelGGL(k_join, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, hi, comp, u64);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_scale_demote(size_t n, const double *src, double coef, float *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_demote, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, coef, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_dot(size_t n, const double *x, const double *y, double *partial, int nparts, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot, dim3(nparts), dim3(256), 0, s, n, x, y, partial);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_to_elem_major(int ne, int nln, const double *cm, double *em, hipStream_t s)
{
    const size_t n = (size_t)2 * ne * nln;
    hipLaunchKernelGGL(k_cm_to_em, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, nln, cm, em);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_to_comp_major(int ne, int nln, const double *em, double *cm, hipStream_t s)
{
    const size_t n = (size_t)2 * ne * nln;
    hipLaunchKernelGGL(k_em_to_cm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, nln, em, cm);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tspeed_dev

// spmm.hip -- many right-hand sides on one mesh: the leap-frog step as a block
// SpMM on the MFMA cores (FP32, compensated update).
//
// R columns (R = 32*NT) share the operator; each column is an independent
// simulation with its own point source, i.e. the reference's LeapFrog::step
// (TimeAdvance.cpp:41-59) applied to R vectors at once.  Per element the step is
//   Y (d x R) = sum_s B_{e,s} (d x d) * U^n_{nb(e,s)} (d x R)
// computed with v_mfma_f32_32x32x2f32 in 32x32 output tiles (MT row tiles of
// the block, NT column tiles of R), followed by the same compensated update as
// the single-column FP32 path and the receiver dot products of every column.
//
// Layouts (rows padded to KP = 32*MT with zeros):
//   state     u[(e*KP + k)*R + j]            row k of element e, column j
//   operator  A-fragment order: for block (e,s), row tile mt and k-block kb
//             (8 columns = 4 MFMA k-steps) 64 lanes x 4 floats, where lane l,
//             slot t holds B_{e,s}(32*mt + l%32, 8*kb + 2t + l/32); one 16-byte
//             non-temporal load per lane feeds four MFMAs.
// MFMA operand/result layout of v_mfma_f32_32x32x2f32 (wave64): A lane l holds
// A[l%32][l/32], B lane l holds B[l/32][l%32], result register r of lane l
// holds C[8*(r/4) + 4*(l/32) + r%4][l%32].
#include "engine.hpp"

#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBWaves = 4;

__device__ inline int xcd_remap_b(int bid, int nwg)
{
    const int per = nwg / 8, rem = nwg % 8;
    const int x = bid % 8, k = bid / 8;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

template <int P>
struct BGeom {
    static constexpr int NLN = Geom<P>::NLN, D = Geom<P>::D;
    static constexpr int MT = (D + 31) / 32;  // 32-row tiles
    static constexpr int KP = 32 * MT;        // padded rows / contraction length
    static constexpr int KB = KP / 8;         // k-blocks per block
};

template <int P, int NT>
__global__ __launch_bounds__(64 * kBWaves) void k_step_b32(BatchArgs a)
{
    using G = BGeom<P>;
    constexpr int R = 32 * NT, MT = G::MT, KP = G::KP, KB = G::KB, D = G::D, NLN = G::NLN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap_b(blockIdx.x, gridDim.x) * kBWaves + wave;
    if (e >= a.ne) return;  // no block-level barrier in this kernel
    const int half = lane >> 5, jj = lane & 31;
    const int4 m = a.meta[e];
    const int nbs[4] = {e, m.x, m.y, m.z};

    f32x16 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x16{};

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (nbs[s] < 0) continue;
        const f32x4 *Af = reinterpret_cast<const f32x4 *>(a.blk + ((size_t)e * 4 + s) * MT * KB * 256);
        const float *X = a.uhi + (size_t)nbs[s] * KP * R + jj;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            f32x4 av[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) av[mt] = __builtin_nontemporal_load(Af + (mt * KB + kb) * 64 + lane);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = kb * 8 + 2 * t + half;
                float bv[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bv[nt] = k < D ? X[k * R + nt * 32] : 0.f;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt][t], bv[nt], acc[mt][nt], 0, 0, 0);
            }
        }
    }

    // compensated update of every (row, column) the lane holds
    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    const float sc = a.sc32[e];
    bool bad = false;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = mt * 32 + 8 * (r / 4) + 4 * half + (r % 4);
                const int j = nt * 32 + jj;
                if (i >= D) continue;
                const size_t idx = ((size_t)e * KP + i) * R + j;
                double f = 0.0;
                for (int k = ex.x; k < ex.y; ++k) {
                    const int src = a.src_sorted[k];
                    if (a.src_col[src] == j)
                        f += a.wav_row[src * 2 + i / NLN] * a.src_phi[src * NLN + i % NLN];
                }
                const float w = a.winc[idx] + sc * (acc[mt][nt][r] - (float)f);
                const float uh = a.uhi[idx];
                const float y = w - a.ucomp[idx];
                const float t = uh + y;
                a.ucomp[idx] = (t - uh) - y;
                a.uhi_out[idx] = t;
                a.winc[idx] = w;
                acc[mt][nt][r] = t;
                bad |= !isfinite(t);
            }
    if (bad) atomicOr(a.flags, 1u);

    // receivers: every receiver records all R columns
    if (a.sample >= 0) {
        for (int k = ex.z; k < ex.w; ++k) {
            const int rec = a.rec_sorted[k];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float sx = 0.f, sy = 0.f;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int i = mt * 32 + 8 * (r / 4) + 4 * half + (r % 4);
                        if (i >= D) continue;
                        const float v = (float)a.rec_phi[rec * NLN + i % NLN] * acc[mt][nt][r];
                        if (i < NLN) sx += v; else sy += v;
                    }
                sx += __shfl_xor(sx, 32);
                sy += __shfl_xor(sy, 32);
                if (half == 0) {
                    const size_t o = (((size_t)a.sample * a.nrec + rec) * R + nt * 32 + jj) * 2;
                    a.seis[o] = sx;
                    a.seis[o + 1] = sy;
                }
            }
        }
    }
}

// FP64 column-pair operator -> FP32 A-fragment order with zero padding.
template <int P>
__global__ void k_pack_mfma(int ne, const double *src, float *dst)
{
    using G = BGeom<P>;
    constexpr int D = G::D, MT = G::MT, KB = G::KB;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)MT * KB * 256;
    if (q >= (size_t)ne * 4 * per) return;
    const size_t b = q / per;
    const int rem = (int)(q % per);
    const int frag = rem / 256, l = (rem % 256) / 4, t = rem % 4;
    const int mt = frag / KB, kb = frag % KB;
    const int row = mt * 32 + l % 32, col = kb * 8 + 2 * t + l / 32;
    float v = 0.f;
    if (row < D && col < D) v = (float)src[b * D * D + (size_t)(col / 2) * 2 * D + row * 2 + (col & 1)];
    dst[q] = v;
}

// Receiver sample of the stored state (eval_receivers): one wave per receiver.
template <int P>
__global__ __launch_bounds__(64) void k_sample_b32(BatchArgs a, int R)
{
    using G = BGeom<P>;
    const int rec = blockIdx.x;
    if (rec >= a.nrec) return;
    const int e = a.rec_elem[rec];
    for (int j = threadIdx.x; j < R; j += 64) {
        double sx = 0.0, sy = 0.0;
        for (int i = 0; i < G::D; ++i) {
            const size_t idx = ((size_t)e * G::KP + i) * R + j;
            const double u = (double)a.uhi[idx] - (double)a.ucomp[idx];
            const double v = a.rec_phi[rec * G::NLN + i % G::NLN] * u;
            if (i < G::NLN) sx += v; else sy += v;
        }
        const size_t o = (((size_t)a.sample * a.nrec + rec) * R + j) * 2;
        a.seis[o] = sx;
        a.seis[o + 1] = sy;
    }
}

// component-major column `col` (FP64) <-> batched staging layout (FP64)
__global__ void k_col_in(int ne, int nln, int KP, int R, int col, const double *cm, double *st)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)2 * ne * nln;
    if (q >= n) return;
    const int c = (int)(q / ((size_t)ne * nln));
    const size_t rem = q % ((size_t)ne * nln);
    const size_t e = rem / nln;
    const int i = (int)(rem % nln);
    st[(e * KP + c * nln + i) * R + col] = cm ? cm[q] : 0.0;
}

__global__ void k_col_out(int ne, int nln, int KP, int R, int col, const float *hi, const float *comp, double *cm)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)2 * ne * nln;
    if (q >= n) return;
    const int c = (int)(q / ((size_t)ne * nln));
    const size_t rem = q % ((size_t)ne * nln);
    const size_t e = rem / nln;
    const int i = (int)(rem % nln);
    const size_t idx = (e * KP + c * nln + i) * R + col;
    cm[q] = (double)hi[idx] - (double)comp[idx];
}

template <int P, int NT>
int step_b32_p(const BatchArgs &a, hipStream_t s)
{
    const unsigned nwg = (unsigned)((a.ne + kBWaves - 1) / kBWaves);
    hipLaunchKernelGGL((k_step_b32<P, NT>), dim3(nwg), dim3(64 * kBWaves), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int pack_p(int ne, const double *src, float *dst, hipStream_t s)
{
    using G = BGeom<P>;
    const size_t n = (size_t)ne * 4 * G::MT * G::KB * 256;
    hipLaunchKernelGGL(k_pack_mfma<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, src, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int sample_b32_p(const BatchArgs &a, int R, hipStream_t s)
{
    hipLaunchKernelGGL(k_sample_b32<P>, dim3(a.nrec), dim3(64), 0, s, a, R);
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

int batch_kp(int P) { return 32 * ((2 * nln_of(P) + 31) / 32); }
size_t batch_block_floats(int P) { return (size_t)4 * batch_kp(P) / 32 * (batch_kp(P) / 8) * 256; }

int launch_step_b32(int P, int ncols, const BatchArgs &a, hipStream_t s)
{
    if (ncols == 32) {
#define CALL(p) step_b32_p<p, 1>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 64) {
#define CALL(p) step_b32_p<p, 2>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    return -2;
}

int launch_pack_mfma(int P, int ne, const double *src, float *dst, hipStream_t s)
{
#define CALL(p) pack_p<p>(ne, src, dst, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_sample_b32(int P, int ncols, const BatchArgs &a, hipStream_t s)
{
#define CALL(p) sample_b32_p<p>(a, ncols, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_col_in(int ne, int nln, int KP, int R, int col, const double *cm, double *st, hipStream_t s)
{
    const size_t n = (size_t)2 * ne * nln;
    hipLaunchKernelGGL(k_col_in, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, nln, KP, R, col, cm, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_col_out(int ne, int nln, int KP, int R, int col, const float *hi, const float *comp, double *cm,
                   hipStream_t s)
{
    const size_t n = (size_t)2 * ne * nln;
    hipLaunchKernelGGL(k_col_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, nln, KP, R, col, hi,
                       comp, cm);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tspeed_dev

// step_common.hpp -- device helpers shared by the single-mesh step kernels
// (step.hip) and the multi-instance kernel (multi.hip): XCD-aware element
// dealing, neighbour gather into LDS, the block-ELL SpMV over column groups,
// source injection and receiver sampling.
#pragma once
#include "engine.hpp"

// Products and sums round as written (no FMA contraction); the SpMV
// accumulates with explicit fma().
#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

constexpr int kWaves = 4;  // waves (elements) per work-group
constexpr int kThreads = 64 * kWaves;

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
template <typename G, typename V>
__device__ inline void record(const StepArgs &a, int4 ex, const V *unew, const int *rows, int nrows_lane)
{
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

// Store of the next state written through L2 at agent scope (a relaxed atomic
// store of the same value), so the step leaves no dirty L2 lines behind for the
// write-back at the end of the kernel.  On the 160 000-triangle p = 6 FP64 step
// (16 GB operator): 2.35 -> 2.29 ms (a copy without the store: 2.19 ms); on a
// 3200-triangle p = 4 one, whose operator stays in the Infinity Cache, plain
// stores are faster (14.3 against 15.1 us), and the FP32c steps gained nothing.
template <typename T>
__device__ inline void st_state(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Translation-free products (FP32 paths, orthonormal basis).  Every boundary
// is traction-free, so a rigid translation -- the same mode-0 coefficient in
// every element, the constant mode of the orthonormal basis -- is in the kernel
// of B: sum_s B_{e,s}(:, mode 0 of c) = 0 for every row.  Hence
//   sum_s B_{e,s} u_{nb(e,s)} = sum_s B_{e,s} (u_{nb(e,s)} - T_e)
// with T_e the translation carrying element e's own mode-0 coefficients
// (rows 0 and NLN).  With the operator rounded to FP32 the right-hand form is
// far more accurate for smooth fields: the rounding error of the blocks then
// multiplies differences of neighbouring means instead of the means
// themselves.  (CPU experiment with the operator rounded to FP32 and FP64
// arithmetic otherwise, against FP64: 820 triangles at p = 2, 4416 steps,
// displacement 5.7e-4 -> 7.4e-6; 3200 triangles at p = 4, 3000 steps,
// seismograms 7.3e-5 -> 6.4e-6.)  The difference of two FP32 neighbours'
// means is exact when they are within a factor 2 of each other (Sterbenz).

// Gather u^n of one element (nb < 0: zeros) into an LDS slice.  With te >= 0
// the translation of element te (its rows 0 and NLN) is subtracted from those
// rows; own (if given) receives the lane's loaded values before that.
template <typename G, typename T>
__device__ inline void gather_one(T *x, const T *u, int nb, int lane, int te = -1, T *own = nullptr)
{
    T v[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int idx = lane + 64 * q;
        v[q] = (nb >= 0 && idx < G::D) ? u[(size_t)nb * G::D + idx] : T(0);
    }
    if (own) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) own[q] = v[q];
    }
    if (te >= 0) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int idx = lane + 64 * q;
            if (idx == 0 || idx == G::NLN) v[q] = v[q] - u[(size_t)te * G::D + idx];
        }
    }
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int idx = lane + 64 * q;
        if (idx < G::D) x[idx] = v[q];
    }
}

// Gather u^n of the element and its three neighbours into the wave's LDS slice;
// every load is issued before the first store so the four fetches overlap.
// With tfree the element's translation (rows 0 and NLN of its own slot) is
// subtracted from rows 0 and NLN of every slot, its own included (which then
// holds 0 there); own (if given) receives the lane's own-slot values as loaded,
// i.e. u^n of rows lane + 64 q.
template <typename G, typename T>
__device__ inline void gather(T (*xs)[G::DP], const T *u, int e, int4 m, int lane, bool act, bool tfree = false,
                              T *own = nullptr)
{
    const int nbs[4] = {act ? e : -1, m.x, m.y, m.z};
    T v[4][G::RPL];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int idx = lane + 64 * q;
            v[s][q] = (nbs[s] >= 0 && idx < G::D) ? u[(size_t)nbs[s] * G::D + idx] : T(0);
        }
    if (own) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) own[q] = v[0][q];
    }
    if (tfree) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int idx = lane + 64 * q;
            if (idx == 0 || idx == G::NLN) {
                const T t = v[0][q];
#pragma unroll
                for (int s = 0; s < 4; ++s) v[s][q] = v[s][q] - t;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int idx = lane + 64 * q;
            if (idx < G::D) xs[s][idx] = v[s][q];
        }
}

template <typename T, int W>
using vec_t = T __attribute__((ext_vector_type(W)));

// r = sum_s B_{e,s} x_s for the lane's rows, blocks stored as column groups of
// width W (+ a trailing pair when d % W == 2).  For NG > 1 the column groups of
// the wave are reduced into group 0.  Branch-free: every slot is read (a
// missing neighbour's block is zero and its gathered columns are zero), lanes
// past the last row re-read the last row and their sums are discarded, and a
// column group past the last column reads the last one with x = 0.  With no
// divergent branch, all loads of a slot are in flight together (a per-row
// branch made the compiler wait on each load of the first slot in turn:
// 24 -> 17 us per step on a 3200-triangle mesh at p = 4).
// Slots s0 .. s0+ns-1 only (xs[k] holds slot s0+k's coefficients): the
// four-waves-per-element FP64 step gives each wave one slot.
template <typename G, typename T, int W>
__device__ inline void spmv(const T *blk, int e, const T (*xs)[G::DP], int lane, T (&acc)[G::RPL], int s0 = 0,
                            int ns = 4)
{
    constexpr int NFULL = G::D / W, TAIL = G::D % W;
    constexpr int NIT = (NFULL + G::NG - 1) / G::NG;
    constexpr int TAILG = NFULL % G::NG;  // the column group with one iteration less takes the tail
    const int g = lane / G::RW, r = lane % G::RW;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = T(0);
    // one slot per iteration: unrolling the four would keep every load of the
    // element live at once and spill
#pragma unroll 1
    for (int s = 0; s < ns; ++s) {
        const T *Bs = blk + ((size_t)e * 4 + s0 + s) * G::D * G::D;
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int jraw = it * G::NG + g;
            const bool jok = G::NG == 1 || jraw < NFULL;
            const int jg = jok ? jraw : NFULL - 1;
            vec_t<T, W> x = *reinterpret_cast<const vec_t<T, W> *>(&xs[s][W * jg]);
            if (!jok) x = vec_t<T, W>{};
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                const int row = min(r + 64 * q, G::D - 1);
                const vec_t<T, W> b = __builtin_nontemporal_load(
                    reinterpret_cast<const vec_t<T, W> *>(Bs + ((size_t)jg * G::D + row) * W));
#pragma unroll
                for (int t = 0; t < W; ++t) acc[q] = fma(b[t], x[t], acc[q]);
            }
        }
        if constexpr (TAIL == 2) {
            vec_t<T, 2> x = *reinterpret_cast<const vec_t<T, 2> *>(&xs[s][W * NFULL]);
            if (g != TAILG) x = vec_t<T, 2>{};
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                const int row = min(r + 64 * q, G::D - 1);
                const vec_t<T, 2> b = __builtin_nontemporal_load(
                    reinterpret_cast<const vec_t<T, 2> *>(Bs + (size_t)NFULL * G::D * W + row * 2));
                acc[q] = fma(b[0], x[0], acc[q]);
                acc[q] = fma(b[1], x[1], acc[q]);
            }
        }
    }
    if constexpr (G::NG > 1) {
#pragma unroll
        for (int o = 32; o >= G::RW; o >>= 1) acc[0] += __shfl_down(acc[0], o);
    }
}


// Columns [it0, it1) of one slot's block (one column group, no tail pair): the
// FP64 step gives each slot to two waves, one per half of the columns.
template <typename G, typename T, int W, int it0, int it1>
__device__ inline void spmv_cols(const T *blk, int e, int slot, const T *x, int lane, T (&acc)[G::RPL])
{
    static_assert(G::NG == 1 && G::D % W == 0, "one column group, no tail");
    const int r = lane % G::RW;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = T(0);
    const T *Bs = blk + ((size_t)e * 4 + slot) * G::D * G::D;
#pragma unroll
    for (int jg = it0; jg < it1; ++jg) {
        const vec_t<T, W> xv = *reinterpret_cast<const vec_t<T, W> *>(&x[W * jg]);
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = min(r + 64 * q, G::D - 1);
            const vec_t<T, W> b = __builtin_nontemporal_load(
                reinterpret_cast<const vec_t<T, W> *>(Bs + ((size_t)jg * G::D + row) * W));
#pragma unroll
            for (int t = 0; t < W; ++t) acc[q] = fma(b[t], xv[t], acc[q]);
        }
    }
}

// FP32 blocks one column group wide (d > 32, p >= 5).  The row sums of an
// element are cut into eight chunks, (slot s, half h) in that order: half 0
// holds the first half of the block's column quads, half 1 the second half and
// the trailing pair when d % 4 == 2.  Each chunk is an fma chain from zero and
// the chunks' sums are added in chunk order.  Every FP32c step of these
// degrees sums in this order, whether one wave walks the chunks (the
// multi-instance step) or eight waves take one chunk each (the single-mesh
// step), so both give bitwise the same state.
template <typename G>
struct Chunks {
    static constexpr int NFULL = G::D / 4, TAIL = G::D % 4;
    static constexpr int NQH = NFULL / 2;   // column quads of one half
    static_assert(G::NG == 1 && NFULL % 2 == 0 && (TAIL == 0 || TAIL == 2), "chunked rows: d > 32");
};

// A chunk's operator entries for the lane's rows (lanes past the last row
// re-read it; their sums are discarded).
template <typename G>
struct ChunkRegs {
    vec_t<float, 4> q[Chunks<G>::NQH][G::RPL];
    vec_t<float, 2> t[G::RPL];
};

template <typename G>
__device__ inline void chunk_load(const float *blk, int e, int s, int h, int lane, ChunkRegs<G> &R)
{
    using C = Chunks<G>;
    const float *Bs = blk + ((size_t)e * 4 + s) * G::D * G::D;
#pragma unroll
    for (int k = 0; k < C::NQH; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = min(lane + 64 * q, G::D - 1);
            R.q[k][q] = __builtin_nontemporal_load(
                reinterpret_cast<const vec_t<float, 4> *>(Bs + ((size_t)(h * C::NQH + k) * G::D + row) * 4));
        }
    if constexpr (C::TAIL == 2) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = min(lane + 64 * q, G::D - 1);
            R.t[q] = h == 1 ? __builtin_nontemporal_load(
                                  reinterpret_cast<const vec_t<float, 2> *>(Bs + (size_t)C::NFULL * G::D * 4 + row * 2))
                            : vec_t<float, 2>{};
        }
    }
}

// acc = chunk (s, h) of the row sums, x = slot s's gathered coefficients.
template <typename G>
__device__ inline void chunk_fma(const ChunkRegs<G> &R, const float *x, int h, float (&acc)[G::RPL])
{
    using C = Chunks<G>;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = 0.f;
#pragma unroll
    for (int k = 0; k < C::NQH; ++k) {
        const vec_t<float, 4> xv = *reinterpret_cast<const vec_t<float, 4> *>(&x[4 * (h * C::NQH + k)]);
#pragma unroll
        for (int q = 0; q < G::RPL; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[q] = fma(R.q[k][q][t], xv[t], acc[q]);
    }
    if constexpr (C::TAIL == 2) {
        if (h == 1) {
            const vec_t<float, 2> xv = *reinterpret_cast<const vec_t<float, 2> *>(&x[4 * C::NFULL]);
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                acc[q] = fma(R.t[q][0], xv[0], acc[q]);
                acc[q] = fma(R.t[q][1], xv[1], acc[q]);
            }
        }
    }
}

// One wave walking the eight chunks of element e (xs[s] = slot s's coefficients).
template <typename G>
__device__ inline void spmv_chunked(const float *blk, int e, const float (*xs)[G::DP], int lane, float (&acc)[G::RPL])
{
    // one slot per iteration, as spmv: the element's loads are not all live at once
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            ChunkRegs<G> R;
            chunk_load<G>(blk, e, s, h, lane, R);
            float c[G::RPL];
            chunk_fma<G>(R, xs[s], h, c);
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) acc[q] = (s == 0 && h == 0) ? c[q] : acc[q] + c[q];
        }
    }
}

// y = (invm_e * minv) w on one element's rows, in the reference's order: the
// entries of the dense inverse mass are formed as invm_e * minv_ij, then each
// row's sum runs over j from 0 (MyMatMultiDimBlockDiag::vecMult, MyMat.hpp:
// 304-310).  w holds the element's d rows (component-major) in LDS.
template <typename G>
__device__ inline double apply_minv_row(const double *minv, double invm, const double *w, int row)
{
    const int c = row / G::NLN, i = row % G::NLN;
    double acc = 0.0;
    for (int j = 0; j < G::NLN; ++j) acc += (invm * minv[i * G::NLN + j]) * w[c * G::NLN + j];
    return acc;
}
}  // namespace
}  // namespace tspeed_dev

// multi.hip -- many independent simulations (meshes of any size, degrees 1..8,
// FP32 compensated) advanced together.
//
// Each instance is an ordinary context (its own operator, or one shared with
// tsp_share_operator; its own sources, receivers and state).  The (instance,
// element) items of all running instances form one list sorted by degree.  One
// step of the whole set is one persistent kernel per degree present --
// k_multi32c<Geom<P>> specialised for that degree, so each keeps the register
// budget (and occupancy) of the single-mesh kernel -- whose resident waves walk
// that degree's items, each running the same element step as k_step32c
// (block-ELL SpMV over column quads + compensated update + source + receivers).
// In a captured chunk of steps each degree is its own chain of graph nodes
// (step kernel, then k_multi_advance bumping the step counters of that degree's
// instances), so the degrees run side by side and none waits for another at a
// step boundary.  Step counters live on the device, so a chunk of steps is one
// graph, built once and replayed.  Instances stop at their
// own step count; finished ones are dropped from the items at the next chunk.
//
// The other launch mode, k_multi_fused, is one kernel for every degree: one
// item list, largest degree first, one item per wave, each dispatched on its
// instance's degree to the same element step.  A kernel is allocated the
// registers of its largest path, so the small degrees run there at the
// occupancy of the largest; the per-degree kernels stay the default.  (A
// persistent form of the fused kernel, resident waves looping over the items,
// held 169 VGPRs and 3 waves per SIMD and stepped the 256-mesh set in 24.8 ms
// against 14.1 ms one-shot and 11.5 ms per degree; it was removed.)
//
// Autotuning: per degree, tsp_multi_autotune times the launch variants on real
// steps of that degree's instances and keeps the fastest: persistent loop or
// one item per wave (or work-group), one slot per wave, for d > 32 one half
// slot per wave, and for d <= 32 also the lane mapping (column groups of 16/32
// lanes reduced by shuffles, or one row per lane across the wave).  Then
// it times one graph chunk of the whole set in each launch mode (per-degree
// kernels, fused) and keeps the fastest.
#include "../../../include/tspeed_amd.h"
#include "internal.hpp"
#include "step_common.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

struct MInst {
    int P, variant, ne, nsteps, stride, nrec, nsrc, max_samples;
    int tfree;  // translation-free products (see gather)
    int steps;  // completed step() calls (advanced on the device)
    int live;   // 0 once removed
    int sel;    // 1 when the instance's elements are in the current item list
    const int4 *meta, *extra;
    const int *src_sorted, *rec_sorted;
    const double *src_phi, *wav, *rec_phi;
    double *seis;
    unsigned *flags;
    const float *blk, *sc;
    float *uhi0, *uhi1, *ulo0, *ulo1;   // the two time levels of u = hi - lo (lo BF16)
};

constexpr int kMWaves = 4;
constexpr int kMaxDP = Geom<8>::DP;

// The end of an element step: source term, compensated update (u^{n+1} over
// u^{n-1} in the ping-pong buffers the step writes), receivers.  rows[q] = the
// lane's q-th row or -1, acc = B u^n of those rows, own = u^n's hi there, cp
// its lo, wi = u^n - u^{n-1} (lf_incr).
template <typename G>
__device__ inline void finish_elem(const MInst &I, int e, int4 m, const int *rows, const float *wi, const float *cp,
                                   const float *own, const float *acc, float sc)
{
    const int n = I.steps + 1;               // 1-based index of this step() call
    float *uout = (n & 1) ? I.uhi0 : I.uhi1, *lout = (n & 1) ? I.ulo0 : I.ulo1;
    StepArgs a{};
    a.nrec = I.nrec;
    a.src_sorted = I.src_sorted;
    a.rec_sorted = I.rec_sorted;
    a.src_phi = I.src_phi;
    a.wav_row = I.nsrc ? I.wav + (size_t)n * I.nsrc * 2 : nullptr;
    a.rec_phi = I.rec_phi;
    a.seis = I.seis;
    const int slot = (I.stride > 0 && n % I.stride == 0) ? n / I.stride : -1;
    a.sample = (slot >= 0 && slot < I.max_samples) ? slot : -1;
    const int4 ex = m.w >= 0 ? I.extra[m.w] : make_int4(0, 0, 0, 0);
    float unew[G::RPL];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        unew[q] = 0.f;
        const int row = rows[q];
        if (row < 0) continue;
        const size_t idx = (size_t)e * G::D + row;
        const float f = (ex.y > ex.x) ? (float)source_term(a, ex, row, G::NLN) : 0.f;
        const float w = wi[q] + sc * (acc[q] - f);
        const float uh = own[q];
        const float y = w - cp[q];
        const float t = uh + y;
        // through L2 (st_state): the 256-mesh set steps ~1.3% faster; same values
        st_state(reinterpret_cast<unsigned short *>(lout) + idx, lo_bits((t - uh) - y));
        st_state(uout + idx, t);
        unew[q] = t;
        bad |= !isfinite(t);
    }
    if (bad) atomicOr(I.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<G, float>(a, ex, unew, rows, G::RPL);
}

template <typename G>
__device__ inline void multi_elem(const MInst &I, int e, int lane, float (*xs)[kMaxDP])
{
    const int cur = (I.steps + 1) & 1;       // after the first step u^n sits in uhi1
    const float *uin = cur ? I.uhi1 : I.uhi0, *lin = cur ? I.ulo1 : I.ulo0;
    const float *hout = cur ? I.uhi0 : I.uhi1, *lout = cur ? I.ulo0 : I.ulo1;   // u^{n-1}
    const int4 m = I.meta[e];
    const int g = lane / G::RW, r = lane % G::RW;
    float wi[G::RPL], cp[G::RPL], hp[G::RPL], cq[G::RPL];
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (g == 0 && row < G::D) ? row : -1;
        const size_t idx = (size_t)e * G::D + row;
        hp[q] = rows[q] >= 0 ? hout[idx] : 0.f;
        cp[q] = rows[q] >= 0 ? lo_ld(lin, idx) : 0.f;
        cq[q] = rows[q] >= 0 ? lo_ld(lout, idx) : 0.f;
    }
    const float sc = I.sc[e];
    float(*xv)[G::DP] = reinterpret_cast<float(*)[G::DP]>(&xs[0][0]);
    float own[G::RPL];   // u^n of the lane's rows (the update's uh)
    gather<G, float>(xv, uin, e, m, lane, true, I.tfree != 0, own);
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) wi[q] = lf_incr(own[q], hp[q], cp[q], cq[q]);
    // the wave reads only its own LDS slice: order its stores before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float acc[G::RPL];
    if constexpr (G::D > 32) spmv_chunked<G>(I.blk, e, xv, lane, acc);
    else spmv<G, float, 4>(I.blk, e, xv, lane, acc);
    finish_elem<G>(I, e, m, rows, wi, cp, own, acc, sc);
    // the next item's gather overwrites this slice
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Two launch shapes per degree: PERSIST = resident waves loop over the items
// (at least 4 waves per SIMD where that fits without spilling, 2 for the
// two-rows-per-lane degrees); otherwise one item per wave and the hardware
// dispatcher balances.  The compiler schedules the looped form more
// aggressively (more loads in flight, more registers), so which one is faster
// depends on the degree -- the autotuner decides.
template <typename G, bool PERSIST>
__global__ __launch_bounds__(64 * kMWaves, PERSIST ? (G::D > 64 ? 2 : 4) : 1) void k_multi32c(
    const MInst *__restrict__ insts, const int2 *__restrict__ items, int nitems)
{
    __shared__ __attribute__((aligned(16))) float xs_all[kMWaves][4][G::DP];
    const int lane = threadIdx.x & 63;
    // wave-uniform values held in scalar registers: the item, and through it the
    // instance record, are then fetched with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float(*xs)[kMaxDP] = reinterpret_cast<float(*)[kMaxDP]>(&xs_all[wave][0][0]);
    if (PERSIST) {
        for (int it = blockIdx.x * kMWaves + wave; it < nitems; it += gridDim.x * kMWaves) {
            const int2 t = items[it];
            const MInst I = insts[t.x];
            if (!I.live || I.steps >= I.nsteps) continue;
            multi_elem<G>(I, t.y, lane, xs);
        }
    } else {
        const int it = blockIdx.x * kMWaves + wave;
        if (it >= nitems) return;
        const int2 t = items[it];
        const MInst I = insts[t.x];
        if (!I.live || I.steps >= I.nsteps) return;
        multi_elem<G>(I, t.y, lane, xs);
    }
}

// One element per work-group, one slot per wave: wave s gathers slot s's
// coefficients and streams block (e, s); waves 1..3 leave their partial sums in
// LDS and wave 0 adds them in slot order and finishes the step as multi_elem
// does.  Shorter streams per wave, four times the waves.  For p = 5, 6 a wave
// sums its slot as the two chunks of Chunks and wave 0 adds all eight in chunk
// order, bitwise as multi_elem; for the other degrees the slot sums are added
// in another order than multi_elem's, so there this is an autotuner candidate,
// never the default (the default keeps every instance bitwise equal to its
// single-context FP32c run).
// Chunked slot sums where a lane holds one row (d = 42, 56); with two rows per
// lane (d = 72, 90) the two halves took 230-256 VGPRs against 140 for one
// chain per slot, so those degrees keep the slot sums.
template <typename G>
constexpr bool kSlotChunks = G::D > 32 && G::RPL == 1;

template <typename G>
__device__ inline void multi_elem_slots(const MInst &I, int e, int lane, int s, float *x,
                                        float (*part)[2][G::RPL][64])
{
    const int cur = (I.steps + 1) & 1;
    const float *uin = cur ? I.uhi1 : I.uhi0, *lin = cur ? I.ulo1 : I.ulo0;
    const float *hout = cur ? I.uhi0 : I.uhi1, *lout = cur ? I.ulo0 : I.ulo1;   // u^{n-1}
    const int4 m = I.meta[e];
    const int nb = s == 0 ? e : (s == 1 ? m.x : (s == 2 ? m.y : m.z));
    const int g = lane / G::RW, r = lane % G::RW;
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (s == 0 && g == 0 && row < G::D) ? row : -1;
    }
    const float sc = s == 0 ? I.sc[e] : 0.f;
    float own[G::RPL];   // wave 0: u^n of the lane's rows (the update's uh)
    gather_one<G, float>(x, uin, nb, lane, I.tfree ? e : -1, own);
    // the wave reads only its own LDS slice: order its stores before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float(*xv)[G::DP] = reinterpret_cast<const float(*)[G::DP]>(x);
    float acc[G::RPL];
    if constexpr (kSlotChunks<G>) {
        float c[2][G::RPL];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            ChunkRegs<G> R;
            chunk_load<G>(I.blk, e, s, h, lane, R);
            chunk_fma<G>(R, x, h, c[h]);
        }
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            if (s > 0) {
                part[s - 1][0][q][lane] = c[0][q];
                part[s - 1][1][q][lane] = c[1][q];
            }
            acc[q] = c[0][q] + c[1][q];
        }
    } else {
        spmv<G, float, 4>(I.blk, e, xv, lane, acc, s, 1);
        if (s > 0) {
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) part[s - 1][0][q][lane] = acc[q];
        }
    }
    __syncthreads();
    if (s > 0) return;
    // wave 0 finishes the element: the rest of its state is loaded only now,
    // so it adds nothing to the registers of the products (p = 7, 8: 66
    // instead of 90-94 VGPRs, 7 instead of 5 waves per SIMD)
    float wi[G::RPL], cp[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const size_t idx = (size_t)e * G::D + rows[q];
        const float hp = rows[q] >= 0 ? hout[idx] : 0.f;
        cp[q] = rows[q] >= 0 ? lo_ld(lin, idx) : 0.f;
        const float cq = rows[q] >= 0 ? lo_ld(lout, idx) : 0.f;
        wi[q] = lf_incr(own[q], hp, cp[q], cq);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            acc[q] += part[k][0][q][lane];
            if constexpr (kSlotChunks<G>) acc[q] += part[k][1][q][lane];
        }
    finish_elem<G>(I, e, m, rows, wi, cp, own, acc, sc);
}

template <typename G, bool PERSIST>
__global__ __launch_bounds__(64 * kMWaves) void k_multi32s(const MInst *__restrict__ insts, const int2 *__restrict__ items, int nitems)
{
    static_assert(kMWaves == 4, "one wave per slot");
    __shared__ __attribute__((aligned(16))) float xs_all[kMWaves][G::DP];
    __shared__ float part[kMWaves - 1][2][G::RPL][64];
    const int lane = threadIdx.x & 63;
    const int s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (!PERSIST) {
        // one item per work-group (the whole work-group takes it: every branch
        // below is uniform over it); without the loop the compiler keeps fewer
        // registers live (p = 7, 8: 104-124 VGPRs instead of 142-174 in the
        // half-slot kernel)
        if ((int)blockIdx.x >= nitems) return;
        const int2 t = items[blockIdx.x];
        const MInst I = insts[t.x];
        if (!I.live || I.steps >= I.nsteps) return;
        multi_elem_slots<G>(I, t.y, lane, s, xs_all[s], part);
        return;
    }
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 t = items[it];
        const MInst I = insts[t.x];
        if (!I.live || I.steps >= I.nsteps) continue;
        multi_elem_slots<G>(I, t.y, lane, s, xs_all[s], part);
        __syncthreads();  // the next item's partial sums overwrite these
    }
}

// One element per work-group of eight waves and one chunk of its row sums per
// wave (Chunks; d > 32): the single-mesh k_step32w on the items of one degree.
// A wave issues its chunk's block loads before it gathers its slot's
// coefficients; waves 1..7 leave their chunk sums in LDS and wave 0 adds them
// in chunk order, bitwise as multi_elem.
constexpr int kCWaves = 8;

template <typename G>
__device__ inline void multi_elem_chunks(const MInst &I, int e, int lane, int w, float *x,
                                         float (*part)[G::RPL][64])
{
    const int s = w >> 1, h = w & 1;
    ChunkRegs<G> R;
    chunk_load<G>(I.blk, e, s, h, lane, R);
    // keep the block loads ahead of the gather
    __builtin_amdgcn_sched_barrier(0);
    const int cur = (I.steps + 1) & 1;
    const float *uin = cur ? I.uhi1 : I.uhi0, *lin = cur ? I.ulo1 : I.ulo0;
    const float *hout = cur ? I.uhi0 : I.uhi1, *lout = cur ? I.ulo0 : I.ulo1;   // u^{n-1}
    const int4 m = I.meta[e];
    const int nb = s == 0 ? e : (s == 1 ? m.x : (s == 2 ? m.y : m.z));
    float wi[G::RPL], cp[G::RPL], hp[G::RPL], cq[G::RPL];
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = lane + 64 * q;
        rows[q] = (w == 0 && row < G::D) ? row : -1;
        const size_t idx = (size_t)e * G::D + row;
        hp[q] = rows[q] >= 0 ? hout[idx] : 0.f;
        cp[q] = rows[q] >= 0 ? lo_ld(lin, idx) : 0.f;
        cq[q] = rows[q] >= 0 ? lo_ld(lout, idx) : 0.f;
    }
    const float sc = w == 0 ? I.sc[e] : 0.f;
    float own[G::RPL];   // wave 0: u^n of the lane's rows (the update's uh)
    gather_one<G, float>(x, uin, nb, lane, I.tfree ? e : -1, own);
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) wi[q] = lf_incr(own[q], hp[q], cp[q], cq[q]);
    // the wave reads only its own LDS slice: order its stores before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float acc[G::RPL];
    chunk_fma<G>(R, x, h, acc);
    if (w > 0) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) part[w - 1][q][lane] = acc[q];
    }
    __syncthreads();
    if (w > 0) return;
#pragma unroll
    for (int k = 0; k < kCWaves - 1; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) acc[q] += part[k][q][lane];
    finish_elem<G>(I, e, m, rows, wi, cp, own, acc, sc);
}

template <typename G, bool PERSIST>
__global__ __launch_bounds__(64 * kCWaves) void k_multi32w(const MInst *__restrict__ insts, const int2 *__restrict__ items, int nitems)
{
    __shared__ __attribute__((aligned(16))) float xs_all[kCWaves][G::DP];
    __shared__ float part[kCWaves - 1][G::RPL][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (!PERSIST) {
        // one item per work-group, as k_multi32s
        if ((int)blockIdx.x >= nitems) return;
        const int2 t = items[blockIdx.x];
        const MInst I = insts[t.x];
        if (!I.live || I.steps >= I.nsteps) return;
        multi_elem_chunks<G>(I, t.y, lane, w, xs_all[w], part);
        return;
    }
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 t = items[it];
        const MInst I = insts[t.x];
        if (!I.live || I.steps >= I.nsteps) continue;
        multi_elem_chunks<G>(I, t.y, lane, w, xs_all[w], part);
        __syncthreads();  // the next item's partial sums overwrite these
    }
}

// Launch variants: bit 0 = whole-wave rows (degrees with d <= 32 only),
// bit 1 = one item per wave / per work-group (else persistent), bit 2 = one
// element per work-group and one slot per wave (k_multi32s; bit 0 ignored),
// bit 3 = one element per work-group and one chunk per wave (k_multi32w;
// d > 32 only, bits 0 and 2 ignored).
constexpr int kVarRows = 1, kVarOneShot = 2, kVarSlots = 4, kVarChunks = 8;

template <int P>
__device__ inline void fused_item(const MInst &I, int e, int lane, float (*xs)[kMaxDP])
{
    if constexpr (2 * ((P + 1) * (P + 2) / 2) <= 32) {
        if (I.variant & kVarRows) {
            multi_elem<Geom<P, 64>>(I, e, lane, xs);
            return;
        }
    }
    multi_elem<Geom<P>>(I, e, lane, xs);
}

// Every degree in one launch: the items of all running instances in one list,
// largest degree first, each dispatched (on the wave-uniform degree of its
// instance) to that degree's element step with the degree's tuned lane mapping.
// The kernel carries the register budget of its largest body, so the small
// degrees run at a lower occupancy than in their own kernels; in exchange the
// whole step is one launch and the dispatcher balances all items together.
__device__ inline void fused_dispatch(const MInst &I, int e, int lane, float (*xs)[kMaxDP])
{
    switch (I.P) {
    case 1: fused_item<1>(I, e, lane, xs); break;
    case 2: fused_item<2>(I, e, lane, xs); break;
    case 3: fused_item<3>(I, e, lane, xs); break;
    case 4: fused_item<4>(I, e, lane, xs); break;
    case 5: fused_item<5>(I, e, lane, xs); break;
    case 6: fused_item<6>(I, e, lane, xs); break;
    case 7: fused_item<7>(I, e, lane, xs); break;
    case 8: fused_item<8>(I, e, lane, xs); break;
    default: break;
    }
}

__global__ __launch_bounds__(64 * kMWaves) void k_multi_fused(const MInst *__restrict__ insts, const int2 *__restrict__ items, int nitems)
{
    __shared__ __attribute__((aligned(16))) float xs_all[kMWaves][4][kMaxDP];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float(*xs)[kMaxDP] = &xs_all[wave][0];
    const int it = blockIdx.x * kMWaves + wave;
    if (it >= nitems) return;
    const int2 t = items[it];
    const MInst I = insts[t.x];
    if (!I.live || I.steps >= I.nsteps) return;
    fused_dispatch(I, t.y, lane, xs);
}

template <int P>
const void *kern_p(int v)
{
    if constexpr (Geom<P>::D > 32) {
        if (v & kVarChunks)
            return (v & kVarOneShot) ? reinterpret_cast<const void *>(&k_multi32w<Geom<P>, false>)
                                     : reinterpret_cast<const void *>(&k_multi32w<Geom<P>, true>);
    }
    if (v & kVarChunks) return nullptr;
    if (v & kVarSlots)
        return (v & kVarOneShot) ? reinterpret_cast<const void *>(&k_multi32s<Geom<P>, false>)
                                 : reinterpret_cast<const void *>(&k_multi32s<Geom<P>, true>);
    constexpr bool narrow = 2 * ((P + 1) * (P + 2) / 2) <= 32;
    if (narrow && (v & kVarRows))
        return (v & kVarOneShot) ? reinterpret_cast<const void *>(&k_multi32c<Geom<P, 64>, false>)
                                 : reinterpret_cast<const void *>(&k_multi32c<Geom<P, 64>, true>);
    return (v & kVarOneShot) ? reinterpret_cast<const void *>(&k_multi32c<Geom<P>, false>)
                             : reinterpret_cast<const void *>(&k_multi32c<Geom<P>, true>);
}

const void *multi_kernel(int P, int v)
{
    switch (P) {
    case 1: return kern_p<1>(v);
    case 2: return kern_p<2>(v);
    case 3: return kern_p<3>(v);
    case 4: return kern_p<4>(v);
    case 5: return kern_p<5>(v);
    case 6: return kern_p<6>(v);
    case 7: return kern_p<7>(v);
    case 8: return kern_p<8>(v);
    default: return nullptr;
    }
}

// Advance the step counter of every running instance of degree P (0 = all).
__global__ void k_multi_advance(MInst *insts, int n, int P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (P == 0 || insts[i].P == P) && insts[i].sel && insts[i].live && insts[i].steps < insts[i].nsteps)
        insts[i].steps += 1;
}

// The same for the multi-column instances of degree P and width R (P = 0: all).
__global__ void k_batch_advance(BatchInst *insts, int n, int P, int R)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (P == 0 || (insts[i].P == P && insts[i].R == R)) && insts[i].sel && insts[i].live &&
        insts[i].steps < insts[i].nsteps)
        insts[i].steps += 1;
}

}  // namespace
}  // namespace tspeed_dev

using namespace tspeed_dev;

struct tsp_multi {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<tsp_context *> ctx;     // by instance id (nullptr once removed)
    // Host mirrors of the two instance tables, both indexed by instance id:
    // single-column FP32C instances in `host`, multi-column ones in `bhost`
    // (an id's entry in the other table stays dead: live = 0, R = 0).
    std::vector<MInst> host;
    std::vector<BatchInst> bhost;
    MInst *d_inst = nullptr;
    BatchInst *d_binst = nullptr;
    int d_cap = 0;
    int d_n = 0;                        // entries uploaded by the last rebuild
    bool stale = false;                 // steps launched since the host mirror was read back
    int2 *d_items = nullptr;
    int nitems = 0, items_cap = 0;
    bool dirty = true;                  // table / items must be rebuilt
    int chunk = 32;                     // steps per captured graph (0 = plain launches)
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0;
    struct Range {
        int P, R, variant, off, count, grid, block;   // R = 1: single-column instances
        const void *kern;
    };
    std::vector<Range> ranges;          // items of one degree (and width): [off, off+count)
    std::map<int, int> variant_of_p;    // autotuned lane mapping per degree
    std::vector<int> only_p;            // restrict the item list (autotuning)
    int launch = TSP_MULTI_PER_DEGREE;  // per-degree kernels or one fused kernel
};

namespace {

struct MFail {
    int code;
};

void mfail(tsp_multi *m, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    m->err = buf;
    throw MFail{code};
}

void mhip(tsp_multi *m, hipError_t e, const char *what)
{
    if (e != hipSuccess) mfail(m, TSP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

template <typename F>
int mguard(tsp_multi *m, F &&f)
{
    if (!m) return TSP_ERR_ARG;
    try {
        mhip(m, hipSetDevice(m->device), "hipSetDevice");
        f();
        m->err.clear();
        return TSP_OK;
    } catch (const MFail &x) {
        return x.code;
    }
}

void drop_graph(tsp_multi *m)
{
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
    m->graph_chunk = 0;
}

// pull the device step counters into the host mirror and the contexts
// (only the entries the device table holds: slots added since the last
// rebuild are not on the device yet, and their host record is the truth)
void sync_progress(tsp_multi *m)
{
    if (!m->stale || !m->d_inst || m->d_n == 0) return;
    std::vector<MInst> dev(m->d_n);
    std::vector<BatchInst> bdev(m->d_n);
    mhip(m, hipMemcpyAsync(dev.data(), m->d_inst, sizeof(MInst) * dev.size(), hipMemcpyDeviceToHost, m->stream),
         "progress");
    mhip(m, hipMemcpyAsync(bdev.data(), m->d_binst, sizeof(BatchInst) * bdev.size(), hipMemcpyDeviceToHost,
                           m->stream),
         "progress");
    mhip(m, hipStreamSynchronize(m->stream), "progress");
    m->stale = false;
    for (size_t i = 0; i < dev.size(); ++i) {
        if (!m->ctx[i]) continue;
        int steps, stride, max_samples;
        if (m->bhost[i].R > 1) {
            BatchInst &b = m->bhost[i];
            b.steps = bdev[i].steps;
            steps = b.steps;
            stride = b.a.stride;
            max_samples = b.a.max_samples;
        } else {
            MInst &h = m->host[i];
            h.steps = dev[i].steps;
            steps = h.steps;
            stride = h.stride;
            max_samples = h.max_samples;
        }
        const int ns = (stride > 0 ? 1 + steps / stride : 0);
        ctx_set_progress(m->ctx[i], steps, (steps + 1) & 1, std::min(ns, max_samples));
    }
}

// progress of instance i, whichever table holds it
int inst_steps(const tsp_multi *m, int i) { return m->bhost[i].R > 1 ? m->bhost[i].steps : m->host[i].steps; }
int inst_nsteps(const tsp_multi *m, int i) { return m->bhost[i].R > 1 ? m->bhost[i].nsteps : m->host[i].nsteps; }
bool inst_live(const tsp_multi *m, int i) { return m->bhost[i].R > 1 ? m->bhost[i].live : m->host[i].live; }

void rebuild(tsp_multi *m)
{
    if (!m->dirty) return;
    drop_graph(m);
    const int n = (int)m->host.size();
    for (int i = 0; i < n; ++i) {
        MInst &h = m->host[i];
        auto it = m->variant_of_p.find(h.P);
        h.variant = it == m->variant_of_p.end() ? kVarOneShot : it->second;
        h.sel = (h.live && h.steps < h.nsteps &&
                 (m->only_p.empty() || std::find(m->only_p.begin(), m->only_p.end(), h.P) != m->only_p.end()))
                    ? 1 : 0;
        // multi-column instances have no launch variants: the autotuner's
        // trials (only_p set) leave them out
        BatchInst &b = m->bhost[i];
        b.sel = (b.R > 1 && b.live && b.steps < b.nsteps && m->only_p.empty()) ? 1 : 0;
    }
    if (n > m->d_cap) {
        if (m->d_inst) (void)hipFree(m->d_inst);
        if (m->d_binst) (void)hipFree(m->d_binst);
        m->d_inst = nullptr;
        m->d_binst = nullptr;
        m->d_cap = std::max(n, 2 * m->d_cap);
        mhip(m, hipMalloc(&m->d_inst, sizeof(MInst) * m->d_cap), "instance table");
        mhip(m, hipMalloc(&m->d_binst, sizeof(BatchInst) * m->d_cap), "instance table");
    }
    if (n) {
        mhip(m, hipMemcpyAsync(m->d_inst, m->host.data(), sizeof(MInst) * n, hipMemcpyHostToDevice, m->stream),
             "instance table");
        mhip(m, hipMemcpyAsync(m->d_binst, m->bhost.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice,
                               m->stream),
             "instance table");
    }
    m->d_n = n;
    // items of running instances, grouped by degree
    std::vector<int2> items;
    std::vector<int> order;
    for (int i = 0; i < n; ++i)
        if (m->host[i].sel) order.push_back(i);
    // per-degree launches need each degree's items contiguous; the fused kernel
    // takes the largest degree (the longest items) first
    const bool fused = m->launch != TSP_MULTI_PER_DEGREE;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return fused ? m->host[a].P > m->host[b].P : m->host[a].P < m->host[b].P;
    });
    for (int i : order)
        for (int e = 0; e < m->host[i].ne; ++e) items.push_back(make_int2(i, e));
    const int nsingle = (int)items.size();
    // then the multi-column instances, grouped by (degree, width); their work
    // items are (element, 32-column tile) pairs
    std::vector<int> border;
    for (int i = 0; i < n; ++i)
        if (m->bhost[i].sel) border.push_back(i);
    std::stable_sort(border.begin(), border.end(), [&](int a, int b) {
        const BatchInst &x = m->bhost[a], &y = m->bhost[b];
        return x.P != y.P ? x.P < y.P : x.R < y.R;
    });
    for (int i : border) {
        const int w = m->bhost[i].a.ne * batch_tiles(m->bhost[i].R);
        for (int k = 0; k < w; ++k) items.push_back(make_int2(i, k));
    }
    m->nitems = (int)items.size();
    if (m->nitems > m->items_cap) {
        if (m->d_items) (void)hipFree(m->d_items);
        m->items_cap = std::max(m->nitems, 2 * m->items_cap);
        mhip(m, hipMalloc(&m->d_items, sizeof(int2) * m->items_cap), "items");
    }
    if (m->nitems)
        mhip(m, hipMemcpyAsync(m->d_items, items.data(), sizeof(int2) * m->nitems, hipMemcpyHostToDevice, m->stream),
             "items");
    mhip(m, hipStreamSynchronize(m->stream), "rebuild");
    hipDeviceProp_t prop;
    mhip(m, hipGetDeviceProperties(&prop, m->device), "properties");
    const int cus = prop.multiProcessorCount;
    m->ranges.clear();
    if (fused && nsingle) {
        tsp_multi::Range g;
        g.P = 0;
        g.R = 1;
        g.variant = 0;
        g.off = 0;
        g.count = nsingle;
        g.kern = reinterpret_cast<const void *>(&k_multi_fused);
        g.grid = (g.count + kMWaves - 1) / kMWaves;
        g.block = 64 * kMWaves;
        m->ranges.push_back(g);
    }
    for (int q = 0; q < nsingle && !fused;) {
        const int P = m->host[items[q].x].P;
        int r = q;
        while (r < nsingle && m->host[items[r].x].P == P) ++r;
        tsp_multi::Range g;
        g.P = P;
        g.R = 1;
        g.variant = m->host[items[q].x].variant;
        g.off = q;
        g.count = r - q;
        g.kern = multi_kernel(P, g.variant);
        if (!g.kern) mfail(m, TSP_ERR_UNSUPPORTED, "degree %d not built", P);
        // items per work-group: one per wave, or one in all with a slot or a
        // chunk per wave
        const bool whole = (g.variant & (kVarSlots | kVarChunks)) != 0;
        const int want = whole ? g.count : (g.count + kMWaves - 1) / kMWaves;
        g.block = 64 * ((g.variant & kVarChunks) ? kCWaves : kMWaves);
        if (g.variant & kVarOneShot) {
            g.grid = want;
        } else {
            int per_cu = 0;
            mhip(m, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, g.kern, g.block, 0), "occupancy");
            g.grid = std::max(1, std::min(want, std::max(1, per_cu) * cus));
        }
        m->ranges.push_back(g);
        q = r;
    }
    for (int q = nsingle; q < m->nitems;) {
        const BatchInst &b0 = m->bhost[items[q].x];
        const int P = b0.P, R = b0.R;
        int r = q;
        while (r < m->nitems && m->bhost[items[r].x].P == P && m->bhost[items[r].x].R == R) ++r;
        tsp_multi::Range g;
        g.P = P;
        g.R = R;
        g.variant = 0;
        g.off = q;
        g.count = r - q;
        int per_wg = 1;
        if (batch_multi_kernel(P, R, &g.kern, &per_wg) != 0)
            mfail(m, TSP_ERR_UNSUPPORTED, "degree %d at %d columns not built", P, R);
        g.grid = (g.count + per_wg - 1) / per_wg;
        g.block = 256;
        m->ranges.push_back(g);
        q = r;
    }
    m->dirty = false;
}

void launch_steps(tsp_multi *m, int nsteps)
{
    int n = (int)m->host.size();
    for (int k = 0; k < nsteps; ++k) {
        for (auto &g : m->ranges) {
            const int2 *items = m->d_items + g.off;
            void *table = g.R > 1 ? (void *)m->d_binst : (void *)m->d_inst;
            void *args[] = {&table, &items, &g.count};
            mhip(m, hipLaunchKernel(g.kern, dim3(g.grid), dim3(g.block), args, 0, m->stream), "multi step");
        }
        hipLaunchKernelGGL(k_multi_advance, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->d_inst, n, 0);
        hipLaunchKernelGGL(k_batch_advance, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->d_binst, n, 0, 0);
    }
    mhip(m, hipGetLastError(), "multi step launch");
}

// one graph for `chunk` steps: one chain per launch range (degree), each step
// of a chain followed by the counter advance of that range's instances only.
// Instances of different degrees never read each other's data, so the chains
// are independent: a degree that finishes its steps early leaves the device to
// the others instead of waiting at every step for the slowest degree.
void build_graph(tsp_multi *m)
{
    drop_graph(m);
    hipGraph_t g = nullptr;
    mhip(m, hipGraphCreate(&g, 0), "graph");
    int n = (int)m->host.size();
    for (auto &r : m->ranges) {
        const int2 *items = m->d_items + r.off;
        void *table = r.R > 1 ? (void *)m->d_binst : (void *)m->d_inst;
        void *args[] = {&table, &items, &r.count};
        int adv_p = r.P;  // 0 for the fused range: every single-column instance
        int adv_r = r.R;
        void *aargs[] = {&m->d_inst, &n, &adv_p};
        void *bargs[] = {&m->d_binst, &n, &adv_p, &adv_r};
        hipGraphNode_t prev = nullptr;
        for (int k = 0; k < m->chunk; ++k) {
            hipKernelNodeParams kp{};
            kp.func = const_cast<void *>(r.kern);
            kp.gridDim = dim3(r.grid);
            kp.blockDim = dim3(r.block);
            kp.sharedMemBytes = 0;
            kp.kernelParams = args;
            kp.extra = nullptr;
            hipGraphNode_t node;
            mhip(m, hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp), "graph node");
            hipKernelNodeParams ap{};
            ap.func = r.R > 1 ? reinterpret_cast<void *>(&k_batch_advance) : reinterpret_cast<void *>(&k_multi_advance);
            ap.gridDim = dim3((n + 255) / 256);
            ap.blockDim = dim3(256);
            ap.sharedMemBytes = 0;
            ap.kernelParams = r.R > 1 ? bargs : aargs;
            ap.extra = nullptr;
            mhip(m, hipGraphAddKernelNode(&prev, g, &node, 1, &ap), "graph node");
        }
    }
    hipError_t e = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    mhip(m, e, "instantiate");
    m->graph_chunk = m->chunk;
}

// run nsteps: whole chunks from the graph, the remainder as launches
void advance(tsp_multi *m, int nsteps)
{
    rebuild(m);
    if (!m->nitems) return;
    m->stale = true;
    int left = nsteps;
    if (m->chunk > 0 && left >= m->chunk) {
        if (!m->graph || m->graph_chunk != m->chunk) build_graph(m);
        while (left >= m->chunk) {
            mhip(m, hipGraphLaunch(m->graph, m->stream), "graph launch");
            left -= m->chunk;
        }
    }
    if (left > 0) launch_steps(m, left);
}

}  // namespace

extern "C" {

int tsp_multi_create(int device, tsp_multi **out)
{
    if (!out) return TSP_ERR_ARG;
    *out = nullptr;
    tsp_multi *m = new (std::nothrow) tsp_multi;
    if (!m) return TSP_ERR_NOMEM;
    m->device = device;
    int rc = mguard(m, [&] { mhip(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream"); });
    if (rc != TSP_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return TSP_OK;
}

void tsp_multi_destroy(tsp_multi *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    drop_graph(m);
    // (a context that borrows an operator never frees it, so the order is free)
    for (size_t i = m->ctx.size(); i-- > 0;)
        if (m->ctx[i]) tsp_destroy(m->ctx[i]);
    if (m->d_inst) (void)hipFree(m->d_inst);
    if (m->d_binst) (void)hipFree(m->d_binst);
    if (m->d_items) (void)hipFree(m->d_items);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *tsp_multi_last_error(const tsp_multi *m) { return m ? m->err.c_str() : "null set"; }

int tsp_multi_add(tsp_multi *m, tsp_context *c, int nsteps, int *id)
{
    return mguard(m, [&] {
        InstView v{};
        if (!ctx_view(c, &v)) mfail(m, TSP_ERR_ARG, "null context");
        if (v.device != m->device) mfail(m, TSP_ERR_ARG, "context on device %d, set on device %d", v.device, m->device);
        if (v.precision != TSP_FP32C)
            mfail(m, TSP_ERR_UNSUPPORTED, "instances must be TSP_FP32C contexts (FP32 products)");
        if (!v.first_done || v.steps != 0) mfail(m, TSP_ERR_STATE, "take the first step, and no other, before adding");
        if (nsteps < 0) mfail(m, TSP_ERR_ARG, "negative step count");
        if (v.nsrc && nsteps >= v.ntab)
            mfail(m, TSP_ERR_ARG, "wavelet table has %d rows, %d steps need %d", v.ntab, nsteps, nsteps + 1);
        if (v.nrec && v.stride > 0 && 1 + nsteps / v.stride > v.max_samples)
            mfail(m, TSP_ERR_ARG, "seismogram buffer of %d samples cannot hold %d", v.max_samples, 1 + nsteps / v.stride);
        // the context's own stream must be done with the first step
        mhip(m, hipStreamSynchronize(v.stream), "context sync");
        sync_progress(m);
        MInst I{};
        BatchInst B{};
        I.P = v.P;
        if (v.ncols > 1) {
            // a multi-column context: its record lives in the batched table, the
            // single-column one stays dead
            if (!ctx_batch_args(c, &B.a)) mfail(m, TSP_ERR_ARG, "not a multi-column context");
            B.steps = 0;
            B.nsteps = nsteps;
            B.live = 1;
            B.P = v.P;
            B.R = v.ncols;
        }
        I.ne = v.ne;
        I.nsteps = nsteps;
        I.stride = v.stride;
        I.nrec = v.nrec;
        I.nsrc = v.nsrc;
        I.max_samples = v.max_samples;
        I.tfree = v.tfree ? 1 : 0;
        I.steps = 0;
        I.live = v.ncols > 1 ? 0 : 1;
        I.meta = v.meta;
        I.extra = v.extra;
        I.src_sorted = v.src_sorted;
        I.rec_sorted = v.rec_sorted;
        I.src_phi = v.src_phi;
        I.wav = v.wav;
        I.rec_phi = v.rec_phi;
        I.seis = v.seis;
        I.flags = v.flags;
        I.blk = v.blk32;
        I.sc = v.sc_step;
        I.uhi0 = v.uhi[0];
        I.uhi1 = v.uhi[1];
        I.ulo0 = v.ulo[0];
        I.ulo1 = v.ulo[1];
        // reuse the slot of a removed instance, so the table stays as long as
        // the most instances ever held at once
        int slot = 0;
        while (slot < (int)m->ctx.size() && m->ctx[slot]) ++slot;
        if (slot == (int)m->ctx.size()) {
            m->host.push_back(I);
            m->bhost.push_back(B);
            m->ctx.push_back(c);
        } else {
            m->host[slot] = I;
            m->bhost[slot] = B;
            m->ctx[slot] = c;
        }
        m->dirty = true;
        if (id) *id = slot;
    });
}

int tsp_multi_set_chunk(tsp_multi *m, int chunk)
{
    return mguard(m, [&] {
        if (chunk < 0) mfail(m, TSP_ERR_ARG, "negative chunk");
        m->chunk = chunk;
        drop_graph(m);
    });
}

int tsp_multi_run(tsp_multi *m, int nsteps, int *active)
{
    return mguard(m, [&] {
        if (nsteps < 0) mfail(m, TSP_ERR_ARG, "negative step count");
        sync_progress(m);
        bool any_done = false;
        for (int i = 0; i < (int)m->ctx.size(); ++i)
            if (inst_live(m, i) && inst_steps(m, i) >= inst_nsteps(m, i) && inst_nsteps(m, i) > 0) any_done = true;
        if (any_done) m->dirty = true;  // drop finished instances from the items
        advance(m, nsteps);
        sync_progress(m);
        int act = 0;
        for (int i = 0; i < (int)m->ctx.size(); ++i)
            act += (inst_live(m, i) && inst_steps(m, i) < inst_nsteps(m, i)) ? 1 : 0;
        if (active) *active = act;
    });
}

int tsp_multi_run_timed(tsp_multi *m, int nsteps, float *ms)
{
    return mguard(m, [&] {
        if (!ms) mfail(m, TSP_ERR_ARG, "null output");
        sync_progress(m);
        rebuild(m);
        // graph instantiation is host work: keep it out of the timed span
        if (m->nitems && m->chunk > 0 && nsteps >= m->chunk && (!m->graph || m->graph_chunk != m->chunk))
            build_graph(m);
        hipEvent_t a, b;
        mhip(m, hipEventCreate(&a), "event");
        mhip(m, hipEventCreate(&b), "event");
        mhip(m, hipEventRecord(a, m->stream), "event");
        advance(m, nsteps);
        mhip(m, hipEventRecord(b, m->stream), "event");
        mhip(m, hipEventSynchronize(b), "event");
        mhip(m, hipEventElapsedTime(ms, a, b), "event");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        sync_progress(m);
    });
}

int tsp_multi_autotune(tsp_multi *m, int steps_per_trial)
{
    return mguard(m, [&] {
        if (steps_per_trial < 1) mfail(m, TSP_ERR_ARG, "steps_per_trial must be >= 1");
        sync_progress(m);
        std::vector<int> degrees;   // of the single-column instances (the others have no variants)
        for (const auto &h : m->host)
            if (h.live && h.steps + 4 * steps_per_trial <= h.nsteps) degrees.push_back(h.P);
        std::sort(degrees.begin(), degrees.end());
        degrees.erase(std::unique(degrees.begin(), degrees.end()), degrees.end());
        const int saved_chunk = m->chunk;
        m->chunk = 0;
        for (int p : degrees) {
            std::vector<int> cand = 2 * nln_of(p) <= 32 ? std::vector<int>{0, 1, 2, 3} : std::vector<int>{0, 2};
            cand.push_back(kVarSlots);
            cand.push_back(kVarSlots | kVarOneShot);
            if (2 * nln_of(p) > 32) {
                cand.push_back(kVarChunks);
                cand.push_back(kVarChunks | kVarOneShot);
            }
            float best = 0;
            int bestv = kVarOneShot;
            for (int var : cand) {
                m->variant_of_p[p] = var;
                m->only_p = {p};
                m->dirty = true;
                rebuild(m);
                hipEvent_t a, b;
                mhip(m, hipEventCreate(&a), "event");
                mhip(m, hipEventCreate(&b), "event");
                mhip(m, hipEventRecord(a, m->stream), "event");
                advance(m, steps_per_trial);
                mhip(m, hipEventRecord(b, m->stream), "event");
                mhip(m, hipEventSynchronize(b), "event");
                float t = 0;
                mhip(m, hipEventElapsedTime(&t, a, b), "event");
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                sync_progress(m);
                if (var == cand[0] || t < best) {
                    best = t;
                    bestv = var;
                }
            }
            m->variant_of_p[p] = bestv;
        }
        m->only_p.clear();
        m->chunk = saved_chunk;
        m->dirty = true;
        // launch mode, on one graph chunk of the whole set (the per-degree kernels
        // only run side by side inside a graph)
        if (m->chunk > 0) {
            bool enough = false;
            for (const auto &h : m->host)
                if (h.live && h.steps + 3 * m->chunk <= h.nsteps) enough = true;
            if (enough) {
                float best = 0;
                int bestm = TSP_MULTI_PER_DEGREE;
                for (int mode : {TSP_MULTI_PER_DEGREE, TSP_MULTI_FUSED}) {
                    m->launch = mode;
                    m->dirty = true;
                    rebuild(m);
                    build_graph(m);
                    hipEvent_t a, b;
                    mhip(m, hipEventCreate(&a), "event");
                    mhip(m, hipEventCreate(&b), "event");
                    mhip(m, hipEventRecord(a, m->stream), "event");
                    advance(m, m->chunk);
                    mhip(m, hipEventRecord(b, m->stream), "event");
                    mhip(m, hipEventSynchronize(b), "event");
                    float t = 0;
                    mhip(m, hipEventElapsedTime(&t, a, b), "event");
                    (void)hipEventDestroy(a);
                    (void)hipEventDestroy(b);
                    sync_progress(m);
                    if (mode == TSP_MULTI_PER_DEGREE || t < best) {
                        best = t;
                        bestm = mode;
                    }
                }
                m->launch = bestm;
                m->dirty = true;
            }
        }
    });
}

int tsp_multi_set_launch(tsp_multi *m, int mode)
{
    return mguard(m, [&] {
        if (mode != TSP_MULTI_PER_DEGREE && mode != TSP_MULTI_FUSED)
            mfail(m, TSP_ERR_ARG, "unknown launch mode %d", mode);
        if (mode != m->launch) {
            sync_progress(m);
            m->launch = mode;
            m->dirty = true;
        }
    });
}

int tsp_multi_launch(const tsp_multi *m) { return m ? m->launch : -1; }

int tsp_multi_set_variant(tsp_multi *m, int degree, int variant)
{
    return mguard(m, [&] {
        if (degree < 1 || degree > 8) mfail(m, TSP_ERR_ARG, "degree %d outside 1..8", degree);
        if (variant < 0 || variant > 15) mfail(m, TSP_ERR_ARG, "unknown variant %d", variant);
        if ((variant & kVarRows) && !(variant & (kVarSlots | kVarChunks)) && 2 * nln_of(degree) > 32)
            mfail(m, TSP_ERR_ARG, "whole-wave rows need d <= 32 (degree %d has d = %d)", degree,
                  2 * nln_of(degree));
        if ((variant & kVarChunks) && 2 * nln_of(degree) <= 32)
            mfail(m, TSP_ERR_ARG, "chunk per wave needs d > 32 (degree %d has d = %d)", degree,
                  2 * nln_of(degree));
        sync_progress(m);
        m->variant_of_p[degree] = variant;
        m->dirty = true;
    });
}

int tsp_multi_variant(const tsp_multi *m, int degree)
{
    if (!m) return -1;
    auto it = m->variant_of_p.find(degree);
    return it == m->variant_of_p.end() ? kVarOneShot : it->second;
}

int tsp_multi_info(tsp_multi *m, int id, int *steps_done, int *nsteps)
{
    return mguard(m, [&] {
        if (id < 0 || id >= (int)m->host.size() || !m->ctx[id]) mfail(m, TSP_ERR_ARG, "no instance %d", id);
        sync_progress(m);
        if (steps_done) *steps_done = inst_steps(m, id);
        if (nsteps) *nsteps = inst_nsteps(m, id);
    });
}

tsp_context *tsp_multi_context(tsp_multi *m, int id)
{
    if (!m || id < 0 || id >= (int)m->ctx.size()) return nullptr;
    return m->ctx[id];
}

int tsp_multi_remove(tsp_multi *m, int id)
{
    return mguard(m, [&] {
        if (id < 0 || id >= (int)m->host.size() || !m->ctx[id]) mfail(m, TSP_ERR_ARG, "no instance %d", id);
        InstView v{};
        ctx_view(m->ctx[id], &v);
        if (v.owns_operator)
            for (size_t j = 0; j < m->ctx.size(); ++j) {
                if ((int)j == id || !m->ctx[j]) continue;
                const bool uses = v.ncols > 1 ? (m->bhost[j].R > 1 && m->bhost[j].a.blk == v.blkA)
                                              : (m->bhost[j].R <= 1 && m->host[j].blk == v.blk32);
                if (uses)
                    mfail(m, TSP_ERR_STATE, "instance %zu still uses the operator of instance %d: remove it first", j,
                          id);
            }
        mhip(m, hipStreamSynchronize(m->stream), "sync");
        sync_progress(m);
        tsp_destroy(m->ctx[id]);
        m->ctx[id] = nullptr;
        m->host[id].live = 0;
        m->bhost[id].live = 0;
        m->bhost[id].R = 0;
        m->dirty = true;
    });
}

int tsp_multi_synchronize(tsp_multi *m)
{
    return mguard(m, [&] { mhip(m, hipStreamSynchronize(m->stream), "synchronize"); });
}

}  // extern "C"

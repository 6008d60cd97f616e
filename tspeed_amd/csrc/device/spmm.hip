// spmm.hip -- many right-hand sides on one mesh: the leap-frog step as a block
// SpMM on the MFMA cores (FP32, compensated update).
//
// R columns (R = 8, 16, 32 or 64) share the operator; each column is an independent
// simulation with its own point source, i.e. the reference's LeapFrog::step
// (TimeAdvance.cpp:41-59) applied to R vectors at once.  Per element the step is
//   Y (d x R) = sum_s B_{e,s} (d x d) * U^n_{nb(e,s)} (d x R)
// computed with v_mfma_f32_32x32x2f32 in 32x32 output tiles (MT row tiles of
// the block, one 32-column tile per wave), followed by the same compensated
// update as the single-column FP32 path and the receiver dot products of every
// column.
//
// MFMA operand/result layout of v_mfma_f32_32x32x2f32 (wave64): A lane l holds
// A[l%32][l/32], B lane l holds B[l/32][l%32], result register r of lane l
// holds C[8*(r/4) + 4*(l/32) + r%4][l%32].
//
// Layouts (block rows padded to 32*MT with zeros -- the MFMA output tile --,
// the contraction and the state rows only to KB = 2*ceil(d/16) octets, the
// 16-row K-chunk of the BF16 MFMA; e.g. p = 5, d = 42: 64 x 48 blocks, 48 state
// rows, instead of 64 x 64 and 64):
//   state     u[((e*KB + i/8)*R + j)*8 + i%8]    row i of element e, column j.
//             A lane's four result registers 4g..4g+3 are rows 8g+4h..8g+4h+3
//             (h = lane/32) of one column: 16 contiguous bytes.
//   contraction order: in k-octet kb, MFMA step t pairs rows 8kb+t (lanes
//             0-31) and 8kb+4+t (lanes 32-63), so the four B operands of a
//             lane over t = 0..3 are again 16 contiguous bytes of the state.
//   operator  A-fragment order: for block (e,s), row tile mt and k-octet kb,
//             64 lanes x 4 floats, lane l / slot t holding
//             B_{e,s}(32*mt + l%32, 8*kb + 4*(l/32) + t); one 16-byte
//             load per lane feeds four MFMAs.
#include "engine.hpp"

#pragma clang fp contract(off)

namespace tspeed_dev {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBWaves = 4;

// Plain (cached) loads and stores throughout: a multi-column working set
// (operator + state, a few hundred MB) stays largely in the 256 MB Infinity
// Cache from one step to the next; non-temporal operator loads measured slower.
__device__ inline f32x4 ld4(const float *p) { return *reinterpret_cast<const f32x4 *>(p); }
__device__ inline void st4(float *p, f32x4 v) { *reinterpret_cast<f32x4 *>(p) = v; }

// 16-byte store of the next state written through L2 (cache policy sc1), via a
// raw buffer resource over one element's block of a state array (a
// wave-uniform base; the fourth descriptor word is the gfx9 raw-buffer one).
// In the slot-per-wave step it measured 35.3-35.9 -> 34.0-34.8 us per step on
// the 64-source square (BF16x3 33.4-33.6 -> 31.2-31.4); the step then leaves
// fewer dirty L2 lines for the write-back at the end of the kernel.
typedef unsigned int u32x4_raw __attribute__((__vector_size__(4 * sizeof(unsigned int))));
constexpr int kRawBufferWord3 = 0x00020000;
constexpr int kCachePolicySC1 = 16;
template <int NBYTES>
__device__ inline void st4_l2(float *block, int off_bytes, f32x4 v)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(block, 0, NBYTES, kRawBufferWord3);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_raw, v), r, off_bytes, 0, kCachePolicySC1);
}

// The multi-column state keeps two time levels of the compensated
// displacement, u = hi - lo with hi in FP32 and lo in BF16 (rounded to nearest
// even, 2 bytes per entry; widened to FP32 for the update), so u carries ~32
// bits; the increment u^n - u^{n-1} is recomputed from them each step instead
// of being stored.  A step reads u^n (hi, lo) and u^{n-1} (hi, lo) and
// overwrites u^{n-1} with u^{n+1}: 18 bytes per entry and step, 6 of them
// stored, against 24 (12 stored) for an FP32 displacement, compensation and
// increment.  The stores are the step's costliest part (on the 64-source
// square a copy without them stepped in 21.6 us against 31.3); the
// compensation in 2 bytes took the step from 24 to 20 bytes (31.3 -> 28.4 us
// in that copy) and the recomputed increment from 20 to 18.  A CPU
// restatement of the FP32 update (column 63 of that workload, 3000 steps)
// gives seismograms within 6.23e-6 of FP64 with the increment recomputed and
// 6.27e-6 with it stored; on column 21, 6.5e-6 with the BF16 compensation,
// 5.3e-6 with an FP32 one and 6.0e-6 with none: the operator's FP32 rounding,
// not the update's, sets the error at these lengths.
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_raw __attribute__((__vector_size__(2 * sizeof(unsigned int))));
__device__ inline u16x4 pack_bf16(f32x4 v) { return u16x4{lo_bits(v[0]), lo_bits(v[1]), lo_bits(v[2]), lo_bits(v[3])}; }
// four compensation entries at entry index idx of a compensation array
__device__ inline f32x4 ldc4(const float *comp, size_t idx)
{
    const u16x4 c = *reinterpret_cast<const u16x4 *>(reinterpret_cast<const unsigned short *>(comp) + idx);
    return f32x4{lo_value(c[0]), lo_value(c[1]), lo_value(c[2]), lo_value(c[3])};
}
__device__ inline void stc4(float *comp, size_t idx, f32x4 v)
{
    *reinterpret_cast<u16x4 *>(reinterpret_cast<unsigned short *>(comp) + idx) = pack_bf16(v);
}
// stc4 written through L2 as st4_l2: eb = entry index of the element's block,
// off_bytes = byte offset of the entries in its FP32 layout (NBYTES = bytes of
// the element's block in FP32)
template <int NBYTES>
__device__ inline void stc4_l2(float *comp, size_t eb, int off_bytes, f32x4 v)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        reinterpret_cast<unsigned short *>(comp) + eb, 0, NBYTES / 2, kRawBufferWord3);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_raw, pack_bf16(v)), r, off_bytes / 2, 0,
                                          kCachePolicySC1);
}

__device__ inline int xcd_remap_b(int bid, int nwg)
{
    const int per = nwg / 8, rem = nwg % 8;
    const int x = bid % 8, k = bid / 8;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

template <int P>
struct BGeom {
    static constexpr int NLN = Geom<P>::NLN, D = Geom<P>::D;
    static constexpr int MT = (D + 31) / 32;  // 32-row output tiles
    static constexpr int KB = 2 * ((D + 15) / 16);  // contraction / state row octets (even)
    // Partial last tile: the last 32-row tile stores only LR rows, its real
    // rows rounded up to a multiple of 8 (lanes l with l % 32 < LR, kept
    // compactly: lane l -> LR (l / 32) + l % 32); the other lanes feed zeros
    // from registers.  p = 3: 24 x 32 entries per block instead of 32 x 32;
    // p = 5: 48 x 48 instead of 64 x 48; p = 6: 56 x 64 instead of 64 x 64.
    static constexpr int LR = 8 * ((D - 32 * (MT - 1) + 7) / 8);
    // operator block size and fragment offsets, in 16-byte units
    static constexpr int BLK4 = (MT - 1) * KB * 64 + KB * 2 * LR;
    static constexpr int BLKF = 4 * BLK4;     // floats per block
    __host__ __device__ static constexpr bool part(int mt) { return LR < 32 && mt == MT - 1; }
    // 16-byte units of one FP32 fragment of row tile mt
    __host__ __device__ static constexpr int fsize(int mt) { return mt == MT - 1 ? 2 * LR : 64; }
    // FP32 A-fragment (mt, kb)
    __host__ __device__ static constexpr int frag(int mt, int kb) { return mt * KB * 64 + kb * fsize(mt); }
    // where row NLN (mode 0 of the second component) sits: row octet, the half
    // of the wave holding it in a 16-byte piece, and its place in the piece
    static constexpr int KBN = NLN / 8, HN = (NLN % 8) / 4, QN = NLN % 4;
};

// Lane l's 16-byte piece of a fragment: full fragments hold all 64 lanes; the
// last tile's only lanes with l % 32 < LR, stored compactly, and the other
// lanes read zeros (their load address is clamped into the fragment).
template <typename G>
__device__ inline f32x4 frag_ld(const f32x4 *base, int mt, int lane)
{
    if (G::part(mt)) {
        const int r = lane & 31;
        const f32x4 v = base[(lane >> 5) * G::LR + (r < G::LR ? r : G::LR - 1)];
        return r < G::LR ? v : f32x4{};
    }
    return base[lane];
}

// Block products on the BF16 MFMA (v_mfma_f32_32x32x16_bf16), accumulating in
// FP32, from the FP32 operator and state.  Each FP32 operator entry is split
// into BF16 parts a = a_h + a_m + a_l (exact for normal values: every residual
// of an FP32 value and its BF16 rounding is again an FP32 value with fewer
// bits, and three parts carry all 24), each state entry into x = x_h + x_l
// (16 bits), and, largest terms first,
//   NP = 3:  a x ~ a_h x_h + a_h x_l + a_m x_h                  (a to 16 bits)
//   NP = 5:  a x ~ a_h x_h + a_h x_l + a_m x_h + a_m x_l + a_l x_h
// NP = 3 drops a_l entirely: the operator is then perturbed by ~2^-17 of each
// entry, a fixed error of B that accumulates as a phase error (64-source
// square, 3000 steps: seismograms ~9e-4 from FP64).  NP = 5 keeps the operator
// whole; its remaining error is the MFMA's own summation, measured on this
// GPU with single-instruction probes: the instruction aligns the accumulator
// and its 16 exact products to the largest of them and keeps 24 bits below
// it, truncating each product toward zero and flooring the accumulator
// (toward -inf) when the products dominate it; v_mfma_f32_32x32x2f32 keeps 23
// bits but rounds to nearest.  Those one-sided cuts act as a small systematic
// change of B: NP = 5 ends 1.6e-5 - 6.2e-5 from FP64 on that workload (FP32
// products 3.6e-6 - 7.2e-6; a CPU restatement with exact sums predicts
// 6.8e-6).  Largest terms first keeps the accumulator the dominant term of
// every instruction after the first, so it is seldom floored: 6.2e-5 against
// 1.25e-4 with the smallest terms first; separate accumulators for the small
// terms gave 5.2e-5 for 16 more registers.
//
// Operands: K-chunk c (16 contraction rows) is FP32 k-octets 2c and 2c + 1.
// A lane of half h holds, in its 8 BF16 slots, contraction indices
// 16c + 4h + 0..3 (from octet 2c, the FP32 fragment's four values) and
// 16c + 8 + 4h + 0..3 (octet 2c + 1); its B operand holds the same rows of
// its column, the FP32 kernels' two 16-byte pieces.  The MFMA sums over the
// 16 slots of both halves, i.e. over all 16 rows of the chunk, whatever their
// order, so the FP32 operator layout serves both product forms.
template <int NP>
__device__ inline f32x16 bsplit_mfma(f32x4 a0, f32x4 a1, f32x4 x0, f32x4 x1, f32x16 acc)
{
    static_assert(NP == 3 || NP == 5, "three or five split products");
    bf16x8 ah, am, al, xh, xl;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float av = q < 4 ? a0[q] : a1[q - 4];
        const __bf16 h = (__bf16)av;
        const float r = av - (float)h;
        const __bf16 m = (__bf16)r;
        ah[q] = h;
        am[q] = m;
        if constexpr (NP == 5) al[q] = (__bf16)(r - (float)m);
        const float xv = q < 4 ? x0[q] : x1[q - 4];
        const __bf16 g = (__bf16)xv;
        xh[q] = g;
        xl[q] = (__bf16)(xv - (float)g);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, xh, acc, 0, 0, 0);
    if constexpr (NP == 5) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, xl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, xh, acc, 0, 0, 0);
    }
    return acc;
}

// Translation-free products (the element's mode-0 coefficients, rows 0 and
// NLN of its own columns, subtracted from those rows of every slot; exact in
// the algebra because every row of B annihilates a rigid translation -- see
// gather in step_common.hpp for the accuracy this buys in FP32).  The two
// values of this lane's column, or 0 where the lane holds neither row; `base`
// points at the lane's 16-byte piece of its column in the element's first row
// octet (lanes of half h hold rows 8 kb + 4 h .. 8 kb + 4 h + 3).
template <int P, int R>
__device__ inline void b_translation(const BatchArgs &a, const float *base, int half, bool colok, float &t0,
                                     float &tn)
{
    using G = BGeom<P>;
    t0 = tn = 0.f;
    if (!a.tfree || !colok) return;
    if (half == 0) t0 = base[0];
    if (half == G::HN) tn = base[(size_t)G::KBN * R * 8 + G::QN];
}

// Waves per element and columns per wave: a wave owns one 32-column MFMA tile;
// a context of fewer than 32 columns fills only the first R columns of it (the
// other lanes read column 0 of the element, use zeros and store nothing).
template <int R>
struct BCols {
    static constexpr int WPE = R >= 32 ? R / 32 : 1;
    static constexpr int CW = R >= 32 ? 32 : R;
};

// Source term of rows i0..i0+3 of column j: the point forces of the element's
// sources that sit in column j, summed in FP64 and rounded once.
template <int P>
__device__ inline f32x4 b_source(const BatchArgs &a, int4 ex, int i0, int j)
{
    constexpr int D = BGeom<P>::D, NLN = BGeom<P>::NLN;
    double fd[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = ex.x; k < ex.y; ++k) {
        const int src = a.src_sorted[k];
        if (a.src_col[src] != j) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = i0 + q;
            if (i < D) fd[q] += a.wav_row[src * 2 + i / NLN] * a.src_phi[src * NLN + i % NLN];
        }
    }
    return f32x4{(float)fd[0], (float)fd[1], (float)fd[2], (float)fd[3]};
}

// The compensated update of four rows of one column (one 16-byte piece):
// u = hi - lo at two time levels, the increment recomputed as their difference,
//   w = (u^n - u^{n-1}) + sc (B u^n - f),   u^{n+1} = u^n + w
// with the sum compensated (Fast2Sum on hi, the new error kept in lo); uh, cp:
// u^n (hi, lo); hp, cq: u^{n-1}; y = B u^n - f.  Returns u^{n+1}'s hi in tn and
// lo in cn; `bad` collects non-finite results.
__device__ inline void b_step4(float sc, f32x4 y, f32x4 uh, f32x4 hp, f32x4 cp, f32x4 cq, f32x4 &tn, f32x4 &cn,
                               bool &bad)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float w = lf_incr(uh[q], hp[q], cp[q], cq[q]) + sc * y[q];
        const float z = w - cp[q];
        const float t = uh[q] + z;
        cn[q] = (t - uh[q]) - z;
        tn[q] = t;
        bad |= !isfinite(t);
    }
}

// Compensated update of the rows and columns a lane holds (four rows of one
// column per 16-byte access) and the NaN/Inf flag; acc holds B u^n on entry
// and u^{n+1} on return, uhv u^n's hi.  The rest of the state (u^n's lo,
// u^{n-1}) is loaded here one 32-row tile at a time, so the four state values
// of every row are never live at once (p = 8, 64 columns: 170 VGPRs and 2
// waves per SIMD, against 226 and 1 with all of it loaded up front; 297 against
// 412 us per step on the 3200-triangle square).  fsrc(mt, nt, g, i0, j) gives
// the source term of the four rows i0.. of column j.
template <int P, int NT, int R, class FSrc>
__device__ inline void b_update(const BatchArgs &a, int e, int c0, int half, int jj, bool colok, size_t ebase,
                                f32x16 (&acc)[BGeom<P>::MT][NT], const f32x4 (&uhv)[BGeom<P>::KB][NT], FSrc fsrc)
{
    using G = BGeom<P>;
    constexpr int MT = G::MT, D = G::D;
    const float sc = a.sc32[e];
    bool bad = false;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        f32x4 hp[4][NT], cp[4][NT], cq[4][NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const size_t idx = ebase + (size_t)(mt * 4 + g) * R * 8 + nt * 256;
                const bool ld = colok && mt * 32 + 8 * g < D;
                hp[g][nt] = ld ? ld4(a.uhi_out + idx) : f32x4{};
                cp[g][nt] = ld ? ldc4(a.ulo, idx) : f32x4{};
                cq[g][nt] = ld ? ldc4(a.ulo_out, idx) : f32x4{};
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int i0 = mt * 32 + 8 * g + 4 * half;
                if (mt * 32 + 8 * g >= D) continue;  // octet entirely padding
                const int j = c0 + nt * 32 + jj;
                const int kb = mt * 4 + g;
                const size_t idx = ebase + (size_t)kb * R * 8 + nt * 256;
                const f32x4 f = fsrc(mt, nt, g, i0, j);
                f32x4 y, tn, cn;
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] = acc[mt][nt][4 * g + q] - f[q];
                b_step4(sc, y, uhv[kb][nt], hp[g][nt], cp[g][nt], cq[g][nt], tn, cn, bad);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[mt][nt][4 * g + q] = tn[q];
                if (colok) {
                    if constexpr (R == 64) {
                        // through L2 at 64 columns: p = 7 measured 393 -> 290 us per
                        // step on the 3200-triangle square (BF16x3 268 -> 219), p = 8
                        // 3% slower; 8-32 columns unchanged, so they keep plain stores
                        constexpr int NB = G::KB * R * 8 * 4;
                        const size_t eb = (size_t)e * G::KB * R * 8;
                        const int off = (int)(idx - eb) * 4;  // within element e's block
                        stc4_l2<NB>(a.ulo_out, eb, off, cn);
                        st4_l2<NB>(a.uhi_out + eb, off, tn);
                    } else {
                        stc4(a.ulo_out, idx, cn);
                        st4(a.uhi_out + idx, tn);
                    }
                }
            }
        // one row tile's state at a time: keep the next tile's loads below
        __builtin_amdgcn_sched_barrier(0);
    }
    if (bad) atomicOr(a.flags, 1u);
}

// Receiver samples of this wave's NT column tiles for receivers ex.z + k0,
// ex.z + k0 + kstep, ...; tval(mt, nt, r) is u^{n+1} in result-register layout.
template <int P, int NT, int R, class TVal>
__device__ inline void b_receivers(const BatchArgs &a, int4 ex, int k0, int kstep, int c0, int half, int jj, bool colok,
                                   TVal tval)
{
    using G = BGeom<P>;
    constexpr int MT = G::MT, D = G::D, NLN = G::NLN;
    for (int k = ex.z + k0; k < ex.w; k += kstep) {
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
                    const float v = (float)a.rec_phi[rec * NLN + i % NLN] * tval(mt, nt, r);
                    if (i < NLN) sx += v; else sy += v;
                }
            sx += __shfl_xor(sx, 32);
            sy += __shfl_xor(sy, 32);
            if (half == 0 && colok) {
                const size_t o = (((size_t)a.sample * a.nrec + rec) * R + c0 + nt * 32 + jj) * 2;
                a.seis[o] = sx;
                a.seis[o + 1] = sy;
            }
        }
    }
}

// The update, the source term and every receiver of the element, by one wave.
template <int P, int NT, int R>
__device__ inline void b_finish(const BatchArgs &a, int e, int4 m, int c0, int half, int jj, bool colok, size_t ebase,
                                f32x16 (&acc)[BGeom<P>::MT][NT], const f32x4 (&uhv)[BGeom<P>::KB][NT])
{
    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    b_update<P, NT, R>(a, e, c0, half, jj, colok, ebase, acc, uhv,
                       [&](int, int, int, int i0, int j) { return b_source<P>(a, ex, i0, j); });
    if (a.sample >= 0)
        b_receivers<P, NT, R>(a, ex, 0, 1, c0, half, jj, colok,
                              [&](int mt, int nt, int r) { return acc[mt][nt][r]; });
}

// Per-step values of step n (1-based index of the step() call) from the
// replay fields of BatchArgs (wav, stride, max_samples, uhi_pair, ulo_pair).
__device__ inline void resolve_n(BatchArgs &a, int n)
{
    a.wav_row = a.nsrc ? a.wav + (size_t)n * a.nsrc * 2 : nullptr;
    a.sample = (a.stride > 0 && n % a.stride == 0 && n / a.stride < a.max_samples) ? n / a.stride : -1;
    const bool odd = n & 1;  // selects, not an index: keeps the arguments out of scratch
    a.uhi = odd ? a.uhi_pair[1] : a.uhi_pair[0];
    a.uhi_out = odd ? a.uhi_pair[0] : a.uhi_pair[1];
    a.ulo = odd ? a.ulo_pair[1] : a.ulo_pair[0];
    a.ulo_out = odd ? a.ulo_pair[0] : a.ulo_pair[1];
}

// Per-step values of a graph-replayed step (see BatchArgs::ctr).
__device__ inline void resolve_step(BatchArgs &a)
{
    if (a.ctr) resolve_n(a, *a.ctr + 1 + a.koff);
}

__global__ void k_ctr_set(int *ctr, int value) { *ctr = value; }
__global__ void k_ctr_add(int *ctr, int value) { *ctr += value; }

// Blocks of three 32-row tiles (p = 7, 8): one wave computes one
// element's rows for one 32-column tile (R/32 waves share an element when
// R >= 32; below that the tile is partly empty), walking the four slots.  A
// missing neighbour's block is zero in the packed operator, so its slot
// multiplies the element's own columns by zero and the slot loop needs no
// branch.  Padded rows stay zero: their operator rows and columns are zero and
// they carry no source.
// gw = the wave's work item: element gw / WPE, column tile gw % WPE.
template <int P, int R, int PD, int NP>
__device__ inline void step_walk(const BatchArgs &a, int gw)
{
    using G = BGeom<P>;
    constexpr int MT = G::MT, KB = G::KB, NT = 1;
    constexpr int WPE = BCols<R>::WPE, CW = BCols<R>::CW;
    const int lane = threadIdx.x & 63;
    const int e = gw / WPE, c0 = (gw % WPE) * 32;
    if (e >= a.ne) return;  // no block-level barrier in this step
    const int half = lane >> 5, jj = lane & 31;
    const bool colok = CW == 32 || jj < CW;
    const int4 m = a.meta[e];
    const int nbs[4] = {e, m.x >= 0 ? m.x : e, m.y >= 0 ? m.y : e, m.z >= 0 ? m.z : e};
    // this lane's 16-byte piece of column c0 + jj in any row octet
    const int lofs = (c0 + (colok ? jj : 0)) * 8 + 4 * half;

    f32x16 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x16{};

    // operands of PD slots ahead are in flight while a slot's products run
    constexpr int NB = PD + 1;
    f32x4 av[NB][KB][MT];
    f32x4 bv[NB][KB][NT];
    // the element's own columns (slot 0) are also u^n's hi of the update: kept;
    // the rest of the update's state is loaded by the update (b_update)
    f32x4 own[KB][NT];
    const size_t ebase = (size_t)e * KB * R * 8 + lofs;
    auto load = [&](int s, int buf) {
        const f32x4 *Af = reinterpret_cast<const f32x4 *>(a.blk + ((size_t)e * 4 + s) * G::BLKF);
        const float *X = a.uhi + (size_t)nbs[s] * KB * R * 8 + lofs;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) av[buf][kb][mt] = frag_ld<G>(Af + G::frag(mt, kb), mt, lane);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                f32x4 x = *reinterpret_cast<const f32x4 *>(X + (size_t)kb * R * 8 + nt * 256);
                if (!colok) x = f32x4{};
                if (s == 0) own[kb][nt] = x; else bv[buf][kb][nt] = x;
            }
    };
    // translation of this lane's column (own element, rows 0 and NLN)
    float t0 = 0.f, tn = 0.f;
    b_translation<P, R>(a, a.uhi + ebase, half, colok, t0, tn);
    auto product = [&](int s, int buf) {
        // the slot's columns, less the translation (rows 0 and NLN)
        f32x4 xs[KB][NT];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                xs[kb][nt] = s == 0 ? own[kb][nt] : bv[buf][kb][nt];
                if (kb == 0) xs[kb][nt][0] = xs[kb][nt][0] - t0;
                if (kb == G::KBN) xs[kb][nt][G::QN] = xs[kb][nt][G::QN] - tn;
            }
        if constexpr (NP == 0) {
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[buf][kb][mt][t], xs[kb][nt][t],
                                                                               acc[mt][nt], 0, 0, 0);
        } else {
#pragma unroll
            for (int c = 0; c < KB / 2; ++c)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[mt][nt] = bsplit_mfma<NP>(av[buf][2 * c][mt], av[buf][2 * c + 1][mt], xs[2 * c][nt],
                                                      xs[2 * c + 1][nt], acc[mt][nt]);
        }
    };
#pragma unroll
    for (int s = 0; s < PD && s < 4; ++s) load(s, s % NB);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s + PD < 4) load(s + PD, (s + PD) % NB);
        // keep the issue order: the scheduler would otherwise sink the loads
        // next to their first use
        if constexpr (PD > 0) __builtin_amdgcn_sched_barrier(0);
        product(s, s % NB);
        if constexpr (PD > 0) __builtin_amdgcn_sched_barrier(0);
    }

    b_finish<P, NT, R>(a, e, m, c0, half, jj, colok, ebase, acc, own);
}

template <int P, int R, int PD>
__global__ __launch_bounds__(64 * kBWaves) void k_step_b32(BatchArgs a)
{
    resolve_step(a);
    const int wave = threadIdx.x >> 6;
    step_walk<P, R, PD, 0>(a, xcd_remap_b(blockIdx.x, gridDim.x) * kBWaves + wave);
}

// The same step with split BF16 products (bsplit_mfma, NP = 3 or 5).
template <int P, int R, int PD, int NP>
__global__ __launch_bounds__(64 * kBWaves) void k_step_b16(BatchArgs a)
{
    resolve_step(a);
    const int wave = threadIdx.x >> 6;
    step_walk<P, R, PD, NP>(a, xcd_remap_b(blockIdx.x, gridDim.x) * kBWaves + wave);
}

// One slot per wave, for blocks of one or two 32-row tiles (p <= 6), at every
// width: a work-group of four waves takes one (element, 32-column tile).  Wave
// s multiplies block (e, s) into its own accumulators (FP32 products when
// NP = 0, else NP split BF16 products, bsplit_mfma) and leaves them in LDS;
// then wave g updates row octets g and g + 4 of the tile, adding the four
// partial sums in slot order, so no wave holds
// the whole tile's update operands (p <= 4: 6-7 waves per SIMD instead of 3;
// p = 5, 6: 2-3 instead of 1).  Wave 1 evaluates the element's source term
// first; elements with receivers take a second barrier and the four waves
// share the receivers.  On the 3200-triangle square (p = 4) a step takes 35.4
// us at 64 columns against 39.5 in the one-wave-per-tile form, 15.2 against
// 21.2 at 8; at p = 5, 8 columns, 45.6 against 59.6.  Blocks of three row
// tiles (p = 7, 8) keep the one-wave-per-tile kernels.
template <int P, int R, int NP>
__device__ inline void step_slots(const BatchArgs &a, int item)
{
    using G = BGeom<P>;
    constexpr int MT = G::MT, KB = G::KB, NO = 4 * MT;  // NO row octets per tile column
    constexpr int WPE = BCols<R>::WPE, CW = BCols<R>::CW;
    static_assert(kBWaves == 4 && MT <= 2, "one wave per slot, MT row octets per wave");
    // part[s][o][lane]: registers 4g..4g+3 of slot s's accumulator of row tile
    // o / 4 (g = o % 4), i.e. row octet o; the same layout holds the source term
    // (fsrc) and, for the receivers, u^{n+1} (part[0])
    __shared__ f32x4 part[4][NO][64];
    __shared__ f32x4 fsrc[NO][64];
    const int lane = threadIdx.x & 63;
    const int s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // this wave's slot; it updates octets s + 4 mt
    const int e = item / WPE, c0 = (item % WPE) * 32;
    const bool act = e < a.ne;
    const int half = lane >> 5, jj = lane & 31;
    const bool colok = CW == 32 || jj < CW;
    const int4 m = act ? a.meta[e] : make_int4(-1, -1, -1, -1);
    // a missing neighbour's block is zero: its slot reads the element's own columns
    const int nb = s == 0 ? e : (s == 1 ? (m.x >= 0 ? m.x : e) : (s == 2 ? (m.y >= 0 ? m.y : e) : (m.z >= 0 ? m.z : e)));
    const int lofs = (c0 + (colok ? jj : 0)) * 8 + 4 * half;
    // the element's sources and receivers: the same for the whole work-group
    const int4 ex = act && m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    const bool src_here = ex.x < ex.y, rec_here = a.sample >= 0 && ex.z < ex.w;
    // wave 1 evaluates the source term (most elements have none) before it
    // loads its own operands, so the FP64 sums add no register pressure
    if (s == 1 && src_here) {
#pragma unroll
        for (int o = 0; o < NO; ++o)
            fsrc[o][lane] = 8 * o < G::D ? b_source<P>(a, ex, 8 * o + 4 * half, c0 + jj) : f32x4{};
    }
    const size_t ebase = (size_t)e * KB * R * 8 + lofs;
    f32x4 uh[MT], hp[MT], cp[MT], cq[MT];
    f32x16 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        uh[mt] = hp[mt] = cp[mt] = cq[mt] = f32x4{};
        acc[mt] = f32x16{};
    }
    // the rows this wave updates: octets s + 4 mt
    auto upd = [&](int mt) { return act && 8 * (s + 4 * mt) < G::D; };
    auto load_update_operands = [&]() {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
            if (upd(mt) && colok) {
                const size_t idx = ebase + (size_t)(s + 4 * mt) * R * 8;
                uh[mt] = *reinterpret_cast<const f32x4 *>(a.uhi + idx);
                hp[mt] = ld4(a.uhi_out + idx);
                cp[mt] = ldc4(a.ulo, idx);
                cq[mt] = ldc4(a.ulo_out, idx);
            }
    };
    if (act) {
        const f32x4 *Af = reinterpret_cast<const f32x4 *>(a.blk + ((size_t)e * 4 + s) * G::BLKF);
        const float *X = a.uhi + (size_t)nb * KB * R * 8 + lofs;
        f32x4 av[KB][MT], bv[KB];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) av[kb][mt] = frag_ld<G>(Af + G::frag(mt, kb), mt, lane);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            const f32x4 x = *reinterpret_cast<const f32x4 *>(X + (size_t)kb * R * 8);
            bv[kb] = colok ? x : f32x4{};
        }
        float t0, tn;
        b_translation<P, R>(a, a.uhi + ebase, half, colok, t0, tn);
        bv[0][0] = bv[0][0] - t0;
        bv[G::KBN][G::QN] = bv[G::KBN][G::QN] - tn;
        if constexpr (NP == 0) {
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kb][mt][t], bv[kb][t], acc[mt], 0, 0, 0);
        } else {
#pragma unroll
            for (int c = 0; c < KB / 2; ++c)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[mt] = bsplit_mfma<NP>(av[2 * c][mt], av[2 * c + 1][mt], bv[2 * c], bv[2 * c + 1], acc[mt]);
        }
        // the update's operands are loaded after the products: their
        // registers are then not live across the MFMAs (8 waves per SIMD
        // at one row tile instead of 6 with FP32 products)
        load_update_operands();
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int g = 0; g < 4; ++g)
                part[s][mt * 4 + g][lane] =
                    f32x4{acc[mt][4 * g], acc[mt][4 * g + 1], acc[mt][4 * g + 2], acc[mt][4 * g + 3]};
    }
    __syncthreads();
    const float sc = act ? a.sc32[e] : 0.f;
    bool bad = false;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        if (!upd(mt)) continue;
        const int o = s + 4 * mt;
        // B u^n in slot order, then - f (x - 0 is x exactly), then the update
        f32x4 y = part[0][o][lane];
#pragma unroll
        for (int k = 1; k < 4; ++k) y += part[k][o][lane];
        if (src_here) y -= fsrc[o][lane];
        f32x4 tn, cn;
        b_step4(sc, y, uh[mt], hp[mt], cp[mt], cq[mt], tn, cn, bad);
        if (colok) {
            // byte offset within element e's block (lofs < R * 8, o < KB)
            constexpr int NB = KB * R * 8 * 4;
            const size_t eb = (size_t)e * KB * R * 8;
            const int off = (int)(lofs + (size_t)o * R * 8) * 4;
            stc4_l2<NB>(a.ulo_out, eb, off, cn);
            st4_l2<NB>(a.uhi_out + eb, off, tn);
        }
        if (rec_here) part[0][o][lane] = tn;
    }
    if (bad) atomicOr(a.flags, 1u);
    if (!rec_here) return;
    // the four waves share the element's receivers, reading u^{n+1} from LDS
    __syncthreads();
    b_receivers<P, 1, R>(a, ex, s, kBWaves, c0, half, jj, colok,
                         [&](int mt, int, int r) { return part[0][mt * 4 + r / 4][lane][r % 4]; });
}

// One row tile (p <= 4): 8 waves per SIMD fit without spilling.
template <int P>
constexpr int slot_min_waves() { return BGeom<P>::MT == 1 ? 8 : 1; }

template <int P, int R>
__global__ __launch_bounds__(64 * kBWaves, slot_min_waves<P>()) void k_step_b32s(BatchArgs a)
{
    resolve_step(a);
    step_slots<P, R, 0>(a, xcd_remap_b(blockIdx.x, gridDim.x));
}

// The same step with split BF16 products (bsplit_mfma, NP = 3 or 5).
template <int P, int R, int NP>
__global__ __launch_bounds__(64 * kBWaves, slot_min_waves<P>()) void k_step_b16s(BatchArgs a)
{
    resolve_step(a);
    step_slots<P, R, NP>(a, xcd_remap_b(blockIdx.x, gridDim.x));
}

// FP64 column-pair operator -> FP32 A-fragment order with zero padding; the
// blocks of missing neighbours are written as zeros.
template <int P>
__global__ void k_pack_mfma(int ne, const int4 *meta, const double *src, float *dst)
{
    using G = BGeom<P>;
    constexpr int D = G::D, MT = G::MT, KB = G::KB;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)MT * KB * 256;   // floats of a full-layout block
    if (q >= (size_t)ne * 4 * per) return;
    const size_t b = q / per;
    const int rem = (int)(q % per);
    const int frag = rem / 256, l = (rem % 256) / 4, t = rem % 4;
    const int mt = frag / KB, kb = frag % KB;
    if (G::part(mt) && l % 32 >= G::LR) return;  // rows past LR of the last tile are not stored
    const int row = mt * 32 + l % 32, col = kb * 8 + 4 * (l / 32) + t;
    const int slot = (int)(b % 4);
    const int4 mm = meta[b / 4];
    const int nb = slot == 0 ? 0 : (slot == 1 ? mm.x : (slot == 2 ? mm.y : mm.z));
    float v = 0.f;
    if (row < D && col < D && nb >= 0) v = (float)src[b * D * D + (size_t)(col / 2) * 2 * D + row * 2 + (col & 1)];
    const int lc = G::part(mt) ? (l / 32) * G::LR + l % 32 : l;
    dst[b * G::BLKF + ((size_t)G::frag(mt, kb) + lc) * 4 + t] = v;
}

__device__ inline size_t oct_index(int e, int i, int j, int KB, int R)
{
    return (((size_t)e * KB + i / 8) * R + j) * 8 + i % 8;
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
            const size_t idx = oct_index(e, i, j, G::KB, R);
            const double u = (double)a.uhi[idx] - (double)lo_ld(a.ulo, idx);
            const double v = a.rec_phi[rec * G::NLN + i % G::NLN] * u;
            if (i < G::NLN) sx += v; else sy += v;
        }
        const size_t o = (((size_t)a.sample * a.nrec + rec) * R + j) * 2;
        a.seis[o] = sx;
        a.seis[o + 1] = sy;
    }
}

// u^{-1} = u0 - dt v0 over the v0 staging (FP64): with it as the earlier time
// level, the first step's increment u0 - u^{-1} is dt v0
__global__ void k_prev_b(size_t n, const double *u0, double *v, double dt)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    v[q] = u0[q] - dt * v[q];
}

// FP64 staging -> FP32 hi + BF16 compensation (true value = hi - comp)
__global__ void k_split_b(size_t n, const double *u, float *hi, float *comp)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float h = (float)u[q];
    hi[q] = h;
    reinterpret_cast<unsigned short *>(comp)[q] = lo_bits((float)((double)h - u[q]));
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
    st[oct_index((int)e, c * nln + i, col, KP / 8, R)] = cm ? cm[q] : 0.0;
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
    const size_t idx = oct_index((int)e, c * nln + i, col, KP / 8, R);
    cm[q] = (double)hi[idx] - (double)lo_ld(comp, idx);
}

// Launch shape per degree: one 32-column tile per wave (R/32 waves per
// element); the operands of the next slot are prefetched while a slot's
// products run when the registers allow it (up to two row tiles).

template <int P, int R>
int step_b32_p(const BatchArgs &a, hipStream_t s)
{
    const size_t items = (size_t)a.ne * BCols<R>::WPE;  // (element, 32-column tile)
    if constexpr (BGeom<P>::MT <= 2) {
        hipLaunchKernelGGL((k_step_b32s<P, R>), dim3((unsigned)items), dim3(64 * kBWaves), 0, s, a);
    } else {
        constexpr int PD = BGeom<P>::MT <= 2 ? 1 : 0;
        const unsigned nwg = (unsigned)((items + kBWaves - 1) / kBWaves);
        hipLaunchKernelGGL((k_step_b32<P, R, PD>), dim3(nwg), dim3(64 * kBWaves), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P, int R, int NP>
int step_b16_p(const BatchArgs &a, hipStream_t s)
{
    const size_t items = (size_t)a.ne * BCols<R>::WPE;
    if constexpr (BGeom<P>::MT <= 2) {
        hipLaunchKernelGGL((k_step_b16s<P, R, NP>), dim3((unsigned)items), dim3(64 * kBWaves), 0, s, a);
    } else {
        const unsigned nwg = (unsigned)((items + kBWaves - 1) / kBWaves);
        hipLaunchKernelGGL((k_step_b16<P, R, 0, NP>), dim3(nwg), dim3(64 * kBWaves), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int pack_p(int ne, const int4 *meta, const double *src, float *dst, hipStream_t s)
{
    using G = BGeom<P>;
    const size_t n = (size_t)ne * 4 * G::MT * G::KB * 256;
    hipLaunchKernelGGL(k_pack_mfma<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ne, meta, src, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P>
int sample_b32_p(const BatchArgs &a, int R, hipStream_t s)
{
    hipLaunchKernelGGL(k_sample_b32<P>, dim3(a.nrec), dim3(64), 0, s, a, R);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Many multi-column contexts in one launch (the multi-instance set,
// multi.hip): items[k] = (instance, work item of that instance), every instance
// of the launch of degree P and width R.  Each item runs the same element step
// as its context's own kernel (bitwise the same results), with the step index
// taken from the instance's counter.  An instance past its step count is
// skipped.  Work items are dealt to the XCDs in contiguous ranges as in the
// single-context kernels.
template <int P, int R>
__global__ __launch_bounds__(64 * kBWaves, slot_min_waves<P>()) void k_bmulti_slots(
    const BatchInst *__restrict__ insts, const int2 *__restrict__ items, int nitems)
{
    const int it = xcd_remap_b(blockIdx.x, gridDim.x);   // grid = nitems
    if (it >= nitems) return;
    const int2 t = items[it];
    const BatchInst &I = insts[t.x];
    if (!I.live || I.steps >= I.nsteps) return;   // uniform over the work-group
    BatchArgs a = I.a;
    resolve_n(a, I.steps + 1);
    step_slots<P, R, 0>(a, t.y);
}

template <int P, int R>
__global__ __launch_bounds__(64 * kBWaves) void k_bmulti_walk(const BatchInst *__restrict__ insts,
                                                              const int2 *__restrict__ items, int nitems)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int it = xcd_remap_b(blockIdx.x, gridDim.x) * kBWaves + wave;
    if (it >= nitems) return;   // no block-level barrier in this step
    const int2 t = items[it];
    const BatchInst &I = insts[t.x];
    if (!I.live || I.steps >= I.nsteps) return;
    BatchArgs a = I.a;
    resolve_n(a, I.steps + 1);
    step_walk<P, R, 0, 0>(a, t.y);
}

template <int P, int R>
int bmulti_p(const void **kern, int *per_wg)
{
    if constexpr (BGeom<P>::MT <= 2) {
        *kern = reinterpret_cast<const void *>(&k_bmulti_slots<P, R>);
        *per_wg = 1;
    } else {
        *kern = reinterpret_cast<const void *>(&k_bmulti_walk<P, R>);
        *per_wg = kBWaves;
    }
    return 0;
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

int launch_ctr_set(int *ctr, int value, hipStream_t s)
{
    hipLaunchKernelGGL(k_ctr_set, dim3(1), dim3(1), 0, s, ctr, value);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_ctr_add(int *ctr, int value, hipStream_t s)
{
    hipLaunchKernelGGL(k_ctr_add, dim3(1), dim3(1), 0, s, ctr, value);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int batch_kp(int P) { return 16 * ((2 * nln_of(P) + 15) / 16); }
int batch_tiles(int R) { return R >= 32 ? R / 32 : 1; }

int batch_multi_kernel(int P, int R, const void **kern, int *items_per_wg)
{
    if (R == 8) {
#define CALL(p) bmulti_p<p, 8>(kern, items_per_wg)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (R == 16) {
#define CALL(p) bmulti_p<p, 16>(kern, items_per_wg)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (R == 32) {
#define CALL(p) bmulti_p<p, 32>(kern, items_per_wg)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (R == 64) {
#define CALL(p) bmulti_p<p, 64>(kern, items_per_wg)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    return -2;
}
size_t batch_block_floats(int P)
{
    switch (P) {
    case 1: return (size_t)4 * BGeom<1>::BLKF; case 2: return (size_t)4 * BGeom<2>::BLKF;
    case 3: return (size_t)4 * BGeom<3>::BLKF; case 4: return (size_t)4 * BGeom<4>::BLKF;
    case 5: return (size_t)4 * BGeom<5>::BLKF; case 6: return (size_t)4 * BGeom<6>::BLKF;
    case 7: return (size_t)4 * BGeom<7>::BLKF; case 8: return (size_t)4 * BGeom<8>::BLKF;
    default: return 0;
    }
}

template <int NP>
int launch_step_b16_np(int P, int ncols, const BatchArgs &a, hipStream_t s)
{
    if (ncols == 8) {
#define CALL(p) step_b16_p<p, 8, NP>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 16) {
#define CALL(p) step_b16_p<p, 16, NP>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 32) {
#define CALL(p) step_b16_p<p, 32, NP>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 64) {
#define CALL(p) step_b16_p<p, 64, NP>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    return -2;
}

int launch_step_b16(int P, int ncols, int nprod, const BatchArgs &a, hipStream_t s)
{
    if (nprod == 3) return launch_step_b16_np<3>(P, ncols, a, s);
    if (nprod == 5) return launch_step_b16_np<5>(P, ncols, a, s);
    return -2;
}

int launch_step_b32(int P, int ncols, const BatchArgs &a, hipStream_t s)
{
    if (ncols == 8) {
#define CALL(p) step_b32_p<p, 8>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 16) {
#define CALL(p) step_b32_p<p, 16>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 32) {
#define CALL(p) step_b32_p<p, 32>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    if (ncols == 64) {
#define CALL(p) step_b32_p<p, 64>(a, s)
        TSP_DISPATCH_P(P, CALL)
#undef CALL
    }
    return -2;
}

int launch_pack_mfma(int P, int ne, const int4 *meta, const double *src, float *dst, hipStream_t s)
{
#define CALL(p) pack_p<p>(ne, meta, src, dst, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_sample_b32(int P, int ncols, const BatchArgs &a, hipStream_t s)
{
#define CALL(p) sample_b32_p<p>(a, ncols, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_prev_b(size_t n, const double *u0, double *v, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_prev_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, u0, v, dt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_split_b(size_t n, const double *u64, float *hi, float *comp, hipStream_t s)
{
    hipLaunchKernelGGL(k_split_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, u64, hi, comp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
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

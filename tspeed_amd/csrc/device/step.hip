// step.hip -- the fused leap-frog step on gfx950.
//
// A step restates LeapFrog::step (TimeAdvance.cpp:41-59) for one element row:
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
// Work mapping: FP64, one element per work-group and one or two waves per slot
// (block); FP32c, one element per wave walking its four slots (d <= 32) or one
// element per work-group and one half slot per wave (d > 32).
// SpMV lane mapping: lane = block row; the wave walks the block's column
// groups (pairs of FP64 columns, quads of FP32 columns -- see engine.hpp), each
// lane fetching its row's entries of one group with a single 16-byte
// non-temporal load (the operator is streamed once per step; keeping it out of
// the caches leaves them to the neighbour gathers).  For d <= 32 the wave is
// split into 2 or 4 column groups reduced with cross-lane shuffles; for d > 64
// each lane owns two rows.  The element's and its neighbours' coefficients are
// gathered into LDS (all loads issued before the first wait) and read as
// broadcasts; the row state the update needs is fetched before the barrier so
// its latency hides behind the block stream.
//
// Elements are dealt to work-groups so that each XCD (work-groups are
// dispatched round-robin over the 8 XCDs) processes one contiguous range of
// elements, keeping neighbour gathers within that XCD's L2.
#include "step_common.hpp"

// Products and sums in the update round as written (no FMA contraction); the
// SpMV accumulates with explicit fma().
#pragma clang fp contract(off)

namespace tspeed_dev {

// ---------------------------------------------------------------- FP64 plain
// One element per work-group; each of the four slots (blocks) goes to its own
// wave, or to two waves (one per half of the block's columns) when the block
// is one column group wide (d > 32, p >= 5).  A wave gathers the coefficients
// of its slot's element into its LDS slice and streams its part of the block;
// the other waves leave their partial sums in LDS and wave 0 (slot 0, first
// half: the element itself) adds them in wave order and finishes the update.
// Short streams per wave keep more loads in flight than one wave walking all
// four blocks: on the 160 000-triangle p = 6 mesh 2.38 ms per step (85% of HBM)
// against 2.42 with one wave per slot and 2.48 with one wave per element; on
// 3200 triangles at p = 4, 14.4 us against 17.2.
// With the face form (FACES) the work-group holds only slot 0's waves and
// adds the three face products the face pass left in zf, in slot order.
template <int P, bool FACES = false>
struct StepShape {
    static constexpr int H = Geom<P>::NG == 1 ? 2 : 1;  // waves per slot
    static constexpr int WAVES = FACES ? H : 4 * H;
};

// Columns [C0, C1) of block `slot` of element e (one column group, no tail),
// with the block loads issued before the gather of x (u of element nb into the
// wave's LDS slice), so the gather's latency hides behind the stream; sums as
// spmv_cols.
template <typename G, int C0, int C1>
__device__ inline void spmv_cols_loads_first(const double *blk, int e, int slot, const double *u, int nb, double *x,
                                             int lane, double (&acc)[G::RPL])
{
    static_assert(G::NG == 1 && G::D % 2 == 0, "one column group, no tail");
    const double *Bs = blk + ((size_t)e * 4 + slot) * G::D * G::D;
    vec_t<double, 2> bl[C1 - C0][G::RPL];
#pragma unroll
    for (int k = 0; k < C1 - C0; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = min(lane + 64 * q, G::D - 1);
            bl[k][q] = __builtin_nontemporal_load(
                reinterpret_cast<const vec_t<double, 2> *>(Bs + ((size_t)(C0 + k) * G::D + row) * 2));
        }
    gather_one<G, double>(x, u, nb, lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = 0.0;
#pragma unroll
    for (int k = 0; k < C1 - C0; ++k) {
        const vec_t<double, 2> xv = *reinterpret_cast<const vec_t<double, 2> *>(&x[2 * (C0 + k)]);
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            acc[q] = fma(bl[k][q][0], xv[0], acc[q]);
            acc[q] = fma(bl[k][q][1], xv[1], acc[q]);
        }
    }
}

// WT: the new state is stored through L2 (see st_state); taken for operators
// larger than the Infinity Cache.  The face form's element step issues its
// block loads before the gather where a lane holds one row (d <= 64), and then
// asks for four waves per SIMD (at most 128 VGPRs, no spill).
template <int P, bool FACES>
constexpr bool elem_loads_first() { return FACES && Geom<P>::NG == 1 && Geom<P>::RPL == 1; }
template <int P, bool FIRST, bool WT, bool FACES = false>
__global__ __launch_bounds__((64 * StepShape<P, FACES>::WAVES), (elem_loads_first<P, FACES>() ? 4 : 1))
void k_step64(StepArgs a)
{
    using G = Geom<P>;
    constexpr int H = StepShape<P>::H, NW = StepShape<P, FACES>::WAVES;
    __shared__ __attribute__((aligned(16))) double xs_all[NW][G::DP];
    __shared__ double part[NW > 1 ? NW - 1 : 1][G::RPL][64];
    __shared__ double wrow[G::DP];   // dense inverse mass: the element's w rows
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = w / H, h = w % H;  // slot, half of the block's columns
    const int e = xcd_remap(blockIdx.x, gridDim.x);
    const bool act = e < a.ne;
    const int4 m = act ? a.meta[e] : make_int4(-1, -1, -1, -1);
    const int nb = !act ? -1 : (s == 0 ? e : (s == 1 ? m.x : (s == 2 ? m.y : m.z)));
    const int g = lane / G::RW, r = lane % G::RW;
    // row state of the update (wave 0), fetched before the barrier
    double up[G::RPL];
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (act && w == 0 && g == 0 && row < G::D) ? row : -1;
        up[q] = rows[q] >= 0 ? a.uprev[(size_t)e * G::D + row] : 0.0;
    }
    const double invm = (act && w == 0) ? a.invm[e] : 0.0;
    // face products of slots 1..3 for the rows of wave 0 (read once), fetched before the barrier
    double zs[3][G::RPL];
    if constexpr (FACES) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q)
                zs[k][q] = rows[q] >= 0 ? __builtin_nontemporal_load(a.zf + ((size_t)e * 3 + k) * G::ZS + rows[q])
                                        : 0.0;
    }
    double(*xs)[G::DP] = &xs_all[w];
    double acc[G::RPL];
    if constexpr (elem_loads_first<P, FACES>()) {
        constexpr int NIT = G::D / 2, HALF = NIT / 2;
        if (h == 0) spmv_cols_loads_first<G, 0, HALF>(a.blk64, e, 0, a.ucur, nb, xs[0], lane, acc);
        else        spmv_cols_loads_first<G, HALF, NIT>(a.blk64, e, 0, a.ucur, nb, xs[0], lane, acc);
    } else {
        gather_one<G, double>(xs[0], a.ucur, nb, lane);
        // the wave reads only its own LDS slice: order its stores before its loads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if constexpr (H == 1) {
            spmv<G, double, 2>(a.blk64, e, xs, lane, acc, s, 1);
        } else {
            constexpr int NIT = G::D / 2, HALF = NIT / 2;
            if (h == 0) spmv_cols<G, double, 2, 0, HALF>(a.blk64, e, s, xs[0], lane, acc);
            else        spmv_cols<G, double, 2, HALF, NIT>(a.blk64, e, s, xs[0], lane, acc);
        }
    }
    if (w > 0) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) part[w - 1][q][lane] = acc[q];
    }
    __syncthreads();
    if (!act || w > 0) return;
#pragma unroll
    for (int k = 0; k < NW - 1; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) acc[q] += part[k][q][lane];
    if constexpr (FACES) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) acc[q] += zs[k][q];
    }

    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    double wv[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = rows[q];
        const double f = (row >= 0 && ex.y > ex.x) ? source_term(a, ex, row, G::NLN) : 0.0;
        wv[q] = a.coef * (acc[q] - f);
    }
    double iw[G::RPL];
    if (a.minv) {
        // non-orthonormal basis: iw = M^-1 w with the element's dense inverse mass
#pragma unroll
        for (int q = 0; q < G::RPL; ++q)
            if (rows[q] >= 0) wrow[rows[q]] = wv[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int q = 0; q < G::RPL; ++q)
            iw[q] = rows[q] >= 0 ? apply_minv_row<G>(a.minv, invm, wrow, rows[q]) : 0.0;
    } else {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) iw[q] = invm * wv[q];
    }
    double unew[G::RPL];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        unew[q] = 0.0;
        const int row = rows[q];
        if (row < 0) continue;
        double un;
        if (FIRST) {
            // u1 = u0 + dt*v0 - invM*(dt*dt/2*(B*u0 - f(0)))  (TimeAdvance.cpp:30)
            un = xs[0][row] + a.dt * up[q] - iw[q];
        } else {
            un = 2 * xs[0][row] - up[q] + iw[q];
        }
        if constexpr (WT) st_state(a.uprev + (size_t)e * G::D + row, un);
        else a.uprev[(size_t)e * G::D + row] = un;
        unew[q] = un;
        bad |= !isfinite(un);
    }
    if (bad) atomicOr(a.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<G, double>(a, ex, unew, rows, G::RPL);
}

// ------------------------------------------------------- FP64 face pass
// B is symmetric: the consistency term enters as I + I^T and the penalty block
// is the same form seen from either side, so B_{eb,ea} = B_{ea,eb}^T (to the
// rounding of assembly, ~2e-16 of the largest entry).  The face form reads the
// lower-numbered element's copy F = B_{ea,eb} of each interior face once and
// applies it to both sides:
//   zf(ea, slot of eb) = F x_eb     lane = row, as the block-ELL SpMV
//   zf(eb, slot of ea) = F^T x_ea   each lane forms F(row, j) x_ea(row) and a
//                                   reduce-scatter butterfly sums each column
//                                   over the lanes
// The operator stream per step falls from ne + #half-faces blocks to ne +
// #faces (3.99 -> 2.50 blocks per element on a 160 000-triangle mesh).  Each
// wave of a face takes a contiguous range of the block's column pairs (3, 4 or
// 5 waves by degree, so that a wave's block entries stay in registers): the
// transposed sums of its columns are complete within the wave, the row sums of
// the ranges are added in LDS in wave order.

// Lane exchanges of the transposed sums.  DPP moves (quad permutations, the
// half-row mirror, the 8-lane rotation of a 16-lane row) stay in the VALU;
// the exchanges across 16-lane rows go through the LDS crossbar (ds_swizzle
// within 32 lanes, ds_bpermute across the halves).
template <int CTRL>
__device__ inline double lane_dpp(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ inline float lane_dpp(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kDppHalfMirror = 0x141; // lane ^ 7 within 8 lanes
constexpr int kDppRor8 = 0x128;       // lane ^ 8 within 16 lanes
__device__ inline double swz_xor16(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    // bitmask mode: and 0x1F, or 0, xor 0x10 (lane ^ 16 within 32 lanes)
    const int lo = __builtin_amdgcn_ds_swizzle((int)(unsigned)u, 0x401F);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(unsigned)(u >> 32), 0x401F);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline float swz_xor16(float x)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x401F));
}

// One halving step of the reduce-scatter: the lane whose bit is set keeps the
// upper half of its n values and adds its partner's copy of that half.
template <int CTRL, int N, typename T>
__device__ inline void halve(T (&v)[8], bool hi)
{
#pragma unroll
    for (int k = 0; k < N / 2; ++k) {
        const T keep = hi ? v[k + N / 2] : v[k];
        const T send = hi ? v[k] : v[k + N / 2];
        v[k] = keep + lane_dpp<CTRL>(send);
    }
}

// Sum over an aligned group of GL lanes (16, 32 or the wave's 64) of NV values
// per lane (NV = 2, 4, 8).  Halving steps on lane bit 2 (partner lane ^ 7),
// bit 1 (lane ^ 2) and bit 0 (lane ^ 1), as many as NV needs, then plain
// butterfly steps over the group's other bits.  Returns the full sum of value
// k: k = lane & 7 (NV = 8), (lane >> 1) & 3 (NV = 4), (lane >> 2) & 1
// (NV = 2).  The order of the additions depends only on the lane bits.
template <int NV, int GL = 64, typename T>
__device__ inline T fold_lanes(const T (&v0)[8], int lane)
{
    static_assert(GL == 16 || GL == 32 || GL == 64, "groups of 16, 32 or 64 lanes");
    static_assert(NV == 2 || NV == 4 || NV == 8, "NV = 2, 4 or 8");
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = v0[k];
    halve<kDppHalfMirror, NV>(v, (lane & 4) != 0);
    if constexpr (NV >= 4) halve<kDppXor2, NV / 2>(v, (lane & 2) != 0);
    if constexpr (NV >= 8) halve<kDppXor1, NV / 4>(v, (lane & 1) != 0);
    T t = v[0];
    if constexpr (NV <= 4) t += lane_dpp<kDppXor1>(t);
    if constexpr (NV <= 2) t += lane_dpp<kDppXor2>(t);
    t += lane_dpp<kDppRor8>(t);
    if constexpr (GL >= 32) t += swz_xor16(t);
    if constexpr (GL == 64) t += __shfl_xor(t, 32);
    return t;
}

// Chunk CI (up to four column pairs) of the NPH column pairs a wave holds:
// the transposed sums of its columns.  bl holds the wave's block entries.
template <typename G, int NPH, int CI, int GL = 64>
__device__ inline void face_chunk(const vec_t<double, 2> (&bl)[NPH][G::RPL], const double (&xa)[G::RPL], int lane,
                                  double (&t)[(NPH + 3) / 4])
{
    if constexpr (4 * CI < NPH) {
        constexpr int NP = NPH - 4 * CI < 4 ? NPH - 4 * CI : 4;
        constexpr int NV = NP == 1 ? 2 : (NP == 2 ? 4 : 8);
        double pv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) pv[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                // lanes past the last row hold xa = 0
                pv[2 * k] = fma(bl[4 * CI + k][q][0], xa[q], pv[2 * k]);
                pv[2 * k + 1] = fma(bl[4 * CI + k][q][1], xa[q], pv[2 * k + 1]);
            }
        t[CI] = fold_lanes<NV, GL>(pv, lane);
        face_chunk<G, NPH, CI + 1, GL>(bl, xa, lane, t);
    }
}

// Which transposed sum lane `lane` stores: chunk CI = lane >> 3, and within it
// one lane per column (the lanes of a chunk's 8 hold its up to 8 columns).
template <int NPH, int CI>
__device__ inline void face_pick(const double (&t)[(NPH + 3) / 4], int lane, double &val, int &col)
{
    if constexpr (4 * CI < NPH) {
        constexpr int NP = NPH - 4 * CI < 4 ? NPH - 4 * CI : 4;
        constexpr int NV = NP == 1 ? 2 : (NP == 2 ? 4 : 8);
        const int k = NV == 8 ? (lane & 7) : (NV == 4 ? (lane >> 1) & 3 : (lane >> 2) & 1);
        const bool writer = NV == 8 ? true : (NV == 4 ? (lane & 1) == 0 : (lane & 3) == 0);
        if ((lane >> 3) == CI && writer && k < 2 * NP) {
            val = t[CI];
            col = 8 * CI + k;
        }
        face_pick<NPH, CI + 1>(t, lane, val, col);
    }
}

// Column pairs [C0, C1) of the face block F of face fc: all of the wave's
// block loads are issued first, then the face's first wave (gather) puts u of
// both elements into the face's LDS rows (x_eb at xb, x_ea after it) and the
// work-group meets at a barrier, so the gathers' latency hides behind the
// stream and each face gathers once (p = 6: 1.78 -> 1.76 ms per step against a
// gather per wave); then the row sums of F x_eb into acc (fma chain in column
// order, as spmv_cols) and the column sums of F^T x_ea, written to the face's
// LDS row zrow[col] (one write per lane).
template <typename G, int C0, int C1>
__device__ inline void face_half(const double *F, const double *u, int4 fc, double *xb, int lane,
                                 double (&acc)[G::RPL], double *zrow, bool gather)
{
    constexpr int NPH = C1 - C0, NCH = (NPH + 3) / 4;
    static_assert(NCH <= 8, "at most 8 chunks per wave");
    vec_t<double, 2> bl[NPH][G::RPL];
#pragma unroll
    for (int k = 0; k < NPH; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = min(lane + 64 * q, G::D - 1);
            bl[k][q] = __builtin_nontemporal_load(
                reinterpret_cast<const vec_t<double, 2> *>(F + ((size_t)(C0 + k) * G::D + row) * 2));
        }
    if (gather) {
        gather_one<G, double>(xb, u, fc.z, lane);
        gather_one<G, double>(xb + G::DP, u, fc.x, lane);
    }
    __syncthreads();
    double xa[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int idx = lane + 64 * q;
        xa[q] = idx < G::D ? xb[G::DP + idx] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NPH; ++k) {
        const vec_t<double, 2> xv = *reinterpret_cast<const vec_t<double, 2> *>(&xb[2 * (C0 + k)]);
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            acc[q] = fma(bl[k][q][0], xv[0], acc[q]);
            acc[q] = fma(bl[k][q][1], xv[1], acc[q]);
        }
    }
    double t[NCH];
    face_chunk<G, NPH, 0>(bl, xa, lane, t);
    double val = 0.0;
    int col = -1;
    face_pick<NPH, 0>(t, lane, val, col);
    if (col >= 0) zrow[2 * C0 + col] = val;
}

// Waves per face (each takes a contiguous range of the block's column pairs,
// sized so that its loads stay in registers) and faces per work-group.
template <int P>
struct FaceShape {
    static constexpr int HF = P <= 5 ? 3 : (P <= 7 ? 4 : 5);
    static constexpr int FPW = HF == 2 ? 2 : 1;
    static constexpr int THREADS = 64 * HF * FPW;
};

// Wave h of a face: column pairs [NIT*h/HF, NIT*(h+1)/HF).
template <typename G, int HF, int H>
__device__ inline void face_wave(int h, const double *F, const double *u, int4 fc, double *xb, int lane,
                                 double (&acc)[G::RPL], double *zrow)
{
    if constexpr (H < HF) {
        constexpr int NIT = G::DH;
        if (h == H) face_half<G, NIT * H / HF, NIT * (H + 1) / HF>(F, u, fc, xb, lane, acc, zrow, H == 0);
        else        face_wave<G, HF, H + 1>(h, F, u, fc, xb, lane, acc, zrow);
    }
}

template <int P>
__global__ __launch_bounds__(FaceShape<P>::THREADS) void k_face64(StepArgs a)
{
    using G = Geom<P>;
    using S = FaceShape<P>;
    static_assert(G::NG == 1 && G::D % 2 == 0, "face form: d > 32");
    __shared__ __attribute__((aligned(16))) double xs_all[S::HF * S::FPW][G::DP];
    __shared__ double part[S::FPW][S::HF - 1][G::RPL][64];
    __shared__ double zrow[S::FPW][G::ZS];   // the face's transposed sums, assembled from its waves
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int fl = w / S::HF, h = w % S::HF;   // face of the work-group, column range of the wave
    const int fraw = xcd_remap(blockIdx.x, gridDim.x) * S::FPW + fl;
    const bool act = fraw < a.nfaces;
    // a spare wave group of the last work-group recomputes the last face and stores nothing
    const int4 fc = a.faces[act ? fraw : a.nfaces - 1];
    const double *F = a.blk64 + ((size_t)fc.x * 4 + fc.y) * G::D * G::D;
    double acc[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) acc[q] = 0.0;
    if (h == 0 && lane < G::ZS - G::D) zrow[fl][G::D + lane] = 0.0;   // the row's padding
    face_wave<G, S::HF, 0>(h, F, a.ucur, fc, xs_all[fl * S::HF], lane, acc, zrow[fl]);
    if (h > 0) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) part[fl][h - 1][q][lane] = acc[q];
    }
    __syncthreads();
    if (h > 1 || !act) return;
    // each product row is written whole, padding included, with non-temporal
    // stores: wave 0 the row sums, wave 1 the transposed sums.  Writes mixed
    // into the block stream cost far more than their bytes (a block stream
    // with 896 bytes written per 25 KB block: 7.15 -> 5.85 TB/s of blocks), and
    // the non-temporal form is the cheapest measured here (p = 6: 1.85 ->
    // 1.78 ms per step; stores through L2 1.82)
    if (h == 1) {
        double *zb = a.zf + ((size_t)fc.z * 3 + fc.w - 1) * G::ZS;
#pragma unroll
        for (int q = 0; q < (G::ZS + 63) / 64; ++q) {
            const int row = lane + 64 * q;
            if (row < G::ZS) __builtin_nontemporal_store(zrow[fl][row], zb + row);
        }
        return;
    }
    double *za = a.zf + ((size_t)fc.x * 3 + fc.y - 1) * G::ZS;
#pragma unroll
    for (int k = 0; k < S::HF - 1; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) acc[q] += part[fl][k][q][lane];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = lane + 64 * q;
        if (row < G::ZS) __builtin_nontemporal_store(row < G::D ? acc[q] : 0.0, za + row);
    }
}

// The face pass for blocks of at most 32 rows (d <= 32, p <= 4): each group
// of RW = 16 or 32 lanes (lane = row) takes one face whole: its block entries
// in registers (loads first), the row sums, the transposed sums folded over
// the group's lanes, and the two product rows written whole by the group.
template <int P>
__global__ __launch_bounds__(256) void k_face64n(StepArgs a)
{
    using G = Geom<P>;
    static_assert(G::NG > 1 && G::RPL == 1 && G::D % 2 == 0, "narrow face pass: d <= 32");
    constexpr int NF = 4 * G::NG, NPH = G::DH, NCH = (NPH + 3) / 4;
    static_assert(NCH <= G::RW / 8 && G::ZS <= G::RW, "a group holds its chunks and its product rows");
    __shared__ __attribute__((aligned(16))) double xs_all[NF][G::DP];
    __shared__ double zrow[NF][G::ZS];
    const int lane = threadIdx.x & 63;
    const int g = lane / G::RW, r = lane % G::RW;
    const int fl = (threadIdx.x >> 6) * G::NG + g;   // face of the work-group
    const int fraw = xcd_remap(blockIdx.x, gridDim.x) * NF + fl;
    const bool act = fraw < a.nfaces;
    // a spare group of the last work-group recomputes the last face and stores nothing
    const int4 fc = a.faces[act ? fraw : a.nfaces - 1];
    const double *F = a.blk64 + ((size_t)fc.x * 4 + fc.y) * G::D * G::D;
    const int row = min(r, G::D - 1);
    vec_t<double, 2> bl[NPH][1];
#pragma unroll
    for (int k = 0; k < NPH; ++k)
        bl[k][0] = __builtin_nontemporal_load(
            reinterpret_cast<const vec_t<double, 2> *>(F + ((size_t)k * G::D + row) * 2));
    double *xb = xs_all[fl];
    const double xbv = r < G::D ? a.ucur[(size_t)fc.z * G::D + r] : 0.0;
    double xa[1] = {r < G::D ? a.ucur[(size_t)fc.x * G::D + r] : 0.0};
    if (r < G::D) xb[r] = xbv;
    if (r < G::ZS - G::D) zrow[fl][G::D + r] = 0.0;   // the row's padding
    // the group reads only its own LDS slices: order its stores before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NPH; ++k) {
        const vec_t<double, 2> xv = *reinterpret_cast<const vec_t<double, 2> *>(&xb[2 * k]);
        acc = fma(bl[k][0][0], xv[0], acc);
        acc = fma(bl[k][0][1], xv[1], acc);
    }
    double t[NCH];
    face_chunk<G, NPH, 0, G::RW>(bl, xa, lane, t);
    double val = 0.0;
    int col = -1;
    face_pick<NPH, 0>(t, r, val, col);
    if (col >= 0) zrow[fl][col] = val;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!act || r >= G::ZS) return;
    double *za = a.zf + ((size_t)fc.x * 3 + fc.y - 1) * G::ZS;
    double *zb = a.zf + ((size_t)fc.z * 3 + fc.w - 1) * G::ZS;
    __builtin_nontemporal_store(r < G::D ? acc : 0.0, za + r);
    __builtin_nontemporal_store(zrow[fl][r], zb + r);
}

// ---------------------------------------------------- FP32 compensated update
// Velocity form of the same recurrence: w^{n+1/2} = w^{n-1/2} + sc_e*(B u^n - f)
// with sc_e = invm_e*(-dt^2), and u^{n+1} = u^n + w^{n+1/2} accumulated by Kahan
// summation (u = hi - lo, lo in BF16), the increment w^{n-1/2} recomputed as
// u^n - u^{n-1} from the two stored time levels (lf_incr): a step reads u^n and
// u^{n-1} and writes u^{n+1} over u^{n-1}, 6 bytes stored per entry instead of
// 12 (FP32 hi, compensation and increment).  The block products and the
// increment are FP32; the compensation keeps the displacement's roundoff from
// drifting over thousands of steps.  The first step uses the same kernel with
// u^{-1} = u0 - dt v0 (so the increment is dt v0) and sc_e = invm_e*(-dt^2/2).
template <int P>
__global__ __launch_bounds__(kThreads) void k_step32c(StepArgs a)
{
    using G = Geom<P>;
    __shared__ __attribute__((aligned(16))) float xs_all[kWaves][4][G::DP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    const bool act = e < a.ne;
    const int4 m = act ? a.meta[e] : make_int4(-1, -1, -1, -1);
    const int g = lane / G::RW, r = lane % G::RW;
    float wi[G::RPL], cp[G::RPL], hp[G::RPL], cq[G::RPL];
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        rows[q] = (act && g == 0 && row < G::D) ? row : -1;
        const size_t idx = (size_t)e * G::D + row;
        hp[q] = rows[q] >= 0 ? a.uhi_out[idx] : 0.f;
        cp[q] = rows[q] >= 0 ? lo_ld(a.ulo, idx) : 0.f;
        cq[q] = rows[q] >= 0 ? lo_ld(a.ulo_out, idx) : 0.f;
    }
    const float sc = act ? a.sc32[e] : 0.f;
    float(*xs)[G::DP] = xs_all[wave];
    float own[G::RPL];   // u^n of the lane's rows (the update's uh)
    gather<G, float>(xs, a.uhi, e, m, lane, act, a.tfree != 0, own);
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) wi[q] = lf_incr(own[q], hp[q], cp[q], cq[q]);
    __syncthreads();
    if (!act) return;

    float acc[G::RPL];
    spmv<G, float, 4>(a.blk32, e, xs, lane, acc);

    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
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
        lo_st(a.ulo_out, idx, (t - uh) - y);
        a.uhi_out[idx] = t;
        unew[q] = t;
        bad |= !isfinite(t);
    }
    if (bad) atomicOr(a.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<G, float>(a, ex, unew, rows, G::RPL);
}

// The same step for blocks one column group wide (d > 32, p >= 5): one element
// per work-group and one chunk of its row sums per wave (slot s = w / 2, half
// h = w % 2; see Chunks).  A wave issues its chunk's block loads first and then
// gathers its slot's coefficients, so the gather's latency hides behind the
// stream; waves 1..7 leave their chunk sums in LDS and wave 0 adds them in
// chunk order and finishes the step as k_step32c does, storing the state
// through L2.  On the 160 000-triangle p = 6 mesh (synthetic operator, same
// shapes): 1.246 ms per step against 1.323 with one wave walking the element,
// 1.274 with the gather first, 1.293 with one wave per slot.
//
// FACES: the FP32c face form's element step: slot 0's two chunks only, then
// the three face products of its slots (zf32) added in slot order.
template <int P, bool FACES = false>
__global__ __launch_bounds__(64 * (FACES ? 2 : 8)) void k_step32w(StepArgs a)
{
    using G = Geom<P>;
    constexpr int NW = FACES ? 2 : 8;
    __shared__ __attribute__((aligned(16))) float xs_all[NW][G::DP];
    __shared__ float part[NW - 1][G::RPL][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = w >> 1, h = w & 1;
    const int e = xcd_remap(blockIdx.x, gridDim.x);   // grid = ne
    ChunkRegs<G> R;
    chunk_load<G>(a.blk32, e, s, h, lane, R);
    // keep the block loads ahead of the gather
    __builtin_amdgcn_sched_barrier(0);
    const int4 m = a.meta[e];
    const int nb = s == 0 ? e : (s == 1 ? m.x : (s == 2 ? m.y : m.z));
    float wi[G::RPL], cp[G::RPL], hp[G::RPL], cq[G::RPL];
    int rows[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = lane + 64 * q;
        rows[q] = (w == 0 && row < G::D) ? row : -1;
        const size_t idx = (size_t)e * G::D + row;
        hp[q] = rows[q] >= 0 ? a.uhi_out[idx] : 0.f;
        cp[q] = rows[q] >= 0 ? lo_ld(a.ulo, idx) : 0.f;
        cq[q] = rows[q] >= 0 ? lo_ld(a.ulo_out, idx) : 0.f;
    }
    const float sc = w == 0 ? a.sc32[e] : 0.f;
    float zs[3][G::RPL];   // face form: the products of slots 1..3 (wave 0's rows, read once)
    if constexpr (FACES) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q)
                zs[k][q] = rows[q] >= 0 ? __builtin_nontemporal_load(a.zf32 + ((size_t)e * 3 + k) * G::ZS32 + rows[q])
                                        : 0.f;
    }
    float own[G::RPL];   // wave 0: u^n of the lane's rows (the update's uh)
    gather_one<G, float>(xs_all[w], a.uhi, nb, lane, a.tfree ? e : -1, own);
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) wi[q] = lf_incr(own[q], hp[q], cp[q], cq[q]);
    // the wave reads only its own LDS slice: order its stores before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float acc[G::RPL];
    chunk_fma<G>(R, xs_all[w], h, acc);
    if (w > 0) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) part[w - 1][q][lane] = acc[q];
    }
    __syncthreads();
    if (w > 0) return;
#pragma unroll
    for (int k = 0; k < NW - 1; ++k)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) acc[q] += part[k][q][lane];
    if constexpr (FACES) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) acc[q] += zs[k][q];
    }

    const int4 ex = m.w >= 0 ? a.extra[m.w] : make_int4(0, 0, 0, 0);
    float unew[G::RPL];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        unew[q] = 0.f;
        const int row = rows[q];
        if (row < 0) continue;
        const size_t idx = (size_t)e * G::D + row;
        const float f = (ex.y > ex.x) ? (float)source_term(a, ex, row, G::NLN) : 0.f;
        const float wn = wi[q] + sc * (acc[q] - f);
        const float uh = own[q];
        const float y = wn - cp[q];
        const float t = uh + y;
        st_state(reinterpret_cast<unsigned short *>(a.ulo_out) + idx, lo_bits((t - uh) - y));
        st_state(a.uhi_out + idx, t);
        unew[q] = t;
        bad |= !isfinite(t);
    }
    if (bad) atomicOr(a.flags, 1u);
    if (a.sample >= 0 && ex.w > ex.z) record<G, float>(a, ex, unew, rows, G::RPL);
}

// ------------------------------------------------------- FP32c face pass
// The FP64 face form for the FP32c step (orthonormal basis): each interior
// face's FP32 block, read once from the lower-numbered element's copy, gives
// both product rows with the translation-free inputs each side's block-ELL
// step uses, zf32(ea) = F (x_eb - T_ea) and zf32(eb) = F^T (x_ea - T_eb) (T_e:
// element e's own mode-0 coefficients).  Two waves per face, one per half of
// the block's column quads (the trailing pair with half 1) as k_step32w splits
// a block; transposed sums folded over the lanes two quads at a time.

// Transposed sums of the column quads of half h, chunk CI = quads 2CI, 2CI+1
// (one quad in the last chunk when the half has an odd count); lanes
// 8CI .. 8CI+7 write the chunk's columns to zrow.
template <typename G, int CI>
__device__ inline void face32_quads(const ChunkRegs<G> &R, const float (&xa)[G::RPL], int h, int lane, float *zrow)
{
    using C = Chunks<G>;
    constexpr int NQC = (C::NQH + 1) / 2;
    static_assert(NQC + 1 <= 8, "at most 8 chunks per wave");
    if constexpr (CI < NQC) {
        constexpr int NQ = C::NQH - 2 * CI < 2 ? C::NQH - 2 * CI : 2;
        constexpr int NV = NQ == 2 ? 8 : 4;
        float pv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) pv[k] = 0.f;
#pragma unroll
        for (int k = 0; k < NQ; ++k)
#pragma unroll
            for (int q = 0; q < G::RPL; ++q)
#pragma unroll
                for (int t = 0; t < 4; ++t) pv[4 * k + t] = fma(R.q[2 * CI + k][q][t], xa[q], pv[4 * k + t]);
        const float tsum = fold_lanes<NV, 64>(pv, lane);
        const int k = NV == 8 ? (lane & 7) : (lane >> 1) & 3;
        const bool writer = NV == 8 ? true : (lane & 1) == 0;
        if ((lane >> 3) == CI && writer) zrow[4 * (h * C::NQH + 2 * CI) + k] = tsum;
        face32_quads<G, CI + 1>(R, xa, h, lane, zrow);
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_face32(StepArgs a)
{
    using G = Geom<P>;
    using C = Chunks<G>;
    static_assert(G::NG == 1, "FP32 face pass: d > 32");
    __shared__ __attribute__((aligned(16))) float xs_all[4][G::DP];
    __shared__ float part[2][G::RPL][64];
    __shared__ float zrow[2][G::ZS32];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int fl = w >> 1, h = w & 1;   // face of the work-group, half of the column quads
    const int fraw = xcd_remap(blockIdx.x, gridDim.x) * 2 + fl;
    const bool act = fraw < a.nfaces;
    // a spare wave pair of the last work-group recomputes the last face and stores nothing
    const int4 fc = a.faces[act ? fraw : a.nfaces - 1];
    ChunkRegs<G> R;
    chunk_load<G>(a.blk32, fc.x, fc.y, h, lane, R);
    // keep the block loads ahead of the gathers
    __builtin_amdgcn_sched_barrier(0);
    // the face's first wave gathers both rows into the face's LDS slices:
    // x_eb - T_ea, then x_ea - T_eb; the work-group meets before they are read
    float *xb = xs_all[2 * fl], *xav = xs_all[2 * fl + 1];
    if (h == 0) {
        gather_one<G, float>(xb, a.uhi, fc.z, lane, a.tfree ? fc.x : -1);
        gather_one<G, float>(xav, a.uhi, fc.x, lane, a.tfree ? fc.z : -1);
        if (lane < G::ZS32 - G::D) zrow[fl][G::D + lane] = 0.f;   // the row's padding
    }
    __syncthreads();
    float xa[G::RPL];
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = lane + 64 * q;
        xa[q] = row < G::D ? xav[row] : 0.f;
    }
    float acc[G::RPL];
    chunk_fma<G>(R, xb, h, acc);
    face32_quads<G, 0>(R, xa, h, lane, zrow[fl]);
    if constexpr (C::TAIL == 2) {
        if (h == 1) {
            float pv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) pv[k] = 0.f;
#pragma unroll
            for (int q = 0; q < G::RPL; ++q) {
                pv[0] = fma(R.t[q][0], xa[q], pv[0]);
                pv[1] = fma(R.t[q][1], xa[q], pv[1]);
            }
            const float tsum = fold_lanes<2, 64>(pv, lane);
            constexpr int CT = (C::NQH + 1) / 2;   // the chunk after the quads
            if ((lane >> 3) == CT && (lane & 3) == 0) zrow[fl][4 * C::NFULL + ((lane >> 2) & 1)] = tsum;
        }
    }
    if (h == 1) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) part[fl][q][lane] = acc[q];
    }
    __syncthreads();
    if (!act) return;
    if (h == 1) {
        float *zb = a.zf32 + ((size_t)fc.z * 3 + fc.w - 1) * G::ZS32;
#pragma unroll
        for (int q = 0; q < (G::ZS32 + 63) / 64; ++q) {
            const int row = lane + 64 * q;
            if (row < G::ZS32) __builtin_nontemporal_store(zrow[fl][row], zb + row);
        }
        return;
    }
    float *za = a.zf32 + ((size_t)fc.x * 3 + fc.y - 1) * G::ZS32;
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = lane + 64 * q;
        if (row < G::ZS32) __builtin_nontemporal_store(row < G::D ? acc[q] + part[fl][q][lane] : 0.f, za + row);
    }
}

// Receiver sample of a stored state, outside a step (eval_receivers):
// one wave per receiver; lo (FP32 state: its BF16 compensation) or null.
template <int P, typename T>
__global__ __launch_bounds__(64) void k_sample(StepArgs a, const T *u, const float *lo)
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
        if (lo) uv -= (double)lo_ld(lo, (size_t)re * G::D + row);
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

__global__ void k_join(size_t n, const float *hi, const float *lo, double *u)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) u[q] = (double)hi[q] - (double)lo_ld(lo, q);
}

__global__ void k_scale_demote(size_t n, const double *src, double coef, float *dst)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (float)(src[q] * coef);
}

// y = invM * B * x for the power iteration (FP64).
template <int P>
__global__ __launch_bounds__(kThreads) void k_apply_minvB(int ne, const int4 *meta, const double *blk,
                                                          const double *invm, const double *minv,
                                                          const double *x, double *y)
{
    using G = Geom<P>;
    __shared__ __attribute__((aligned(16))) double xs_all[kWaves][4][G::DP];
    __shared__ double trow[kWaves][G::DP];   // dense inverse mass: the element's (B x) rows
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    const bool act = e < ne;
    const int4 m = act ? meta[e] : make_int4(-1, -1, -1, -1);
    double(*xs)[G::DP] = xs_all[wave];
    gather<G, double>(xs, x, e, m, lane, act);
    __syncthreads();
    if (!act) return;
    double acc[G::RPL];
    spmv<G, double, 2>(blk, e, xs, lane, acc);
    const int g = lane / G::RW, r = lane % G::RW;
    const double iv = invm[e];
    if (minv) {
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = r + 64 * q;
            if (g == 0 && row < G::D) trow[wave][row] = acc[q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int q = 0; q < G::RPL; ++q) {
        const int row = r + 64 * q;
        if (g == 0 && row < G::D)
            y[(size_t)e * G::D + row] = minv ? apply_minv_row<G>(minv, iv, trow[wave], row) : iv * acc[q];
    }
}

// The operator plugin on vectors in the reference's component-major layout
// x[c*ne*nln + e*nln + i]:
//   MODE 0: y = B x        (MyMatMultiDim<MyMat>::operator*, MyMat.hpp:330-335)
//   MODE 1: y = invM x     (MyMatMultiDimBlockDiag<MyMat>::operator*, MyMat.hpp:337-342;
//                           invM_e = invm_e * base_invmass, Matrices_imp.hpp:187)
// One wave per element; the element's (MODE 0: and its neighbours') entries are
// gathered from the two component halves into LDS, so no layout conversion or
// staging buffer is needed.
template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void k_apply_cm(int ne, const int4 *meta, const double *blk,
                                                       const double *invm, const double *minv,
                                                       const double *x, double *y)
{
    using G = Geom<P>;
    __shared__ __attribute__((aligned(16))) double xs_all[kWaves][4][G::DP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = xcd_remap(blockIdx.x, gridDim.x) * kWaves + wave;
    if (e >= ne) return;   // no block-wide barrier below: each wave works alone
    const size_t half = (size_t)ne * G::NLN;
    const int4 m = meta[e];
    const int nbs[4] = {e, m.x, m.y, m.z};
    double(*xs)[G::DP] = xs_all[wave];
    constexpr int NS = MODE == 0 ? 4 : 1;
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int idx = lane + 64 * q;
            if (idx < G::D)
                xs[s][idx] = nbs[s] >= 0 ? x[(idx / G::NLN) * half + (size_t)nbs[s] * G::NLN + idx % G::NLN] : 0.0;
        }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (MODE == 0) {
        double acc[G::RPL];
        spmv<G, double, 2>(blk, e, xs, lane, acc);
        const int g = lane / G::RW, r = lane % G::RW;
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = r + 64 * q;
            if (g == 0 && row < G::D) y[(row / G::NLN) * half + (size_t)e * G::NLN + row % G::NLN] = acc[q];
        }
    } else {
        const double iv = invm[e];
#pragma unroll
        for (int q = 0; q < G::RPL; ++q) {
            const int row = lane + 64 * q;
            if (row < G::D) {
                // orthonormal basis: the dense block is invm_e * I, whose row sum is exactly invm_e * x_i
                const double v = minv ? apply_minv_row<G>(minv, iv, xs[0], row) : iv * xs[0][row];
                y[(row / G::NLN) * half + (size_t)e * G::NLN + row % G::NLN] = v;
            }
        }
    }
}

// Start vector of the power iteration: x[q] uniform in [-0.5, 0.5) from a
// counter-based hash of q (splitmix64's finaliser), so no host buffer is needed.
__global__ void k_fill_start(size_t n, double *x)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    unsigned long long z = (unsigned long long)q + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[q] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

__global__ void k_scale(size_t n, double *x, double s)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] *= s;
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
    // one element per work-group
    {
        if (a.faces) {
            if (a.nfaces > 0) {
                if constexpr (Geom<P>::NG == 1)
                    hipLaunchKernelGGL(k_face64<P>,
                                       dim3((unsigned)((a.nfaces + FaceShape<P>::FPW - 1) / FaceShape<P>::FPW)),
                                       dim3(FaceShape<P>::THREADS), 0, s, a);
                else
                    hipLaunchKernelGGL(k_face64n<P>,
                                       dim3((unsigned)((a.nfaces + 4 * Geom<P>::NG - 1) / (4 * Geom<P>::NG))),
                                       dim3(256), 0, s, a);
            }
            const dim3 block(64 * StepShape<P, true>::WAVES), grid((unsigned)a.ne);
            if (a.wt) {
                if (first) hipLaunchKernelGGL((k_step64<P, true, true, true>), grid, block, 0, s, a);
                else       hipLaunchKernelGGL((k_step64<P, false, true, true>), grid, block, 0, s, a);
            } else {
                if (first) hipLaunchKernelGGL((k_step64<P, true, false, true>), grid, block, 0, s, a);
                else       hipLaunchKernelGGL((k_step64<P, false, false, true>), grid, block, 0, s, a);
            }
            return hipGetLastError() == hipSuccess ? 0 : -1;
        }
    }
    const dim3 block(64 * StepShape<P>::WAVES);
    const dim3 grid((unsigned)a.ne);
    if (a.wt) {
        if (first) hipLaunchKernelGGL((k_step64<P, true, true>), grid, block, 0, s, a);
        else       hipLaunchKernelGGL((k_step64<P, false, true>), grid, block, 0, s, a);
    } else {
        if (first) hipLaunchKernelGGL((k_step64<P, true, false>), grid, block, 0, s, a);
        else       hipLaunchKernelGGL((k_step64<P, false, false>), grid, block, 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int step32c_p(bool first, const StepArgs &a, hipStream_t s)
{
    (void)first;  // the first step differs only in the data the context prepares
    if constexpr (Geom<P>::NG == 1) {
        if (a.zf32) {
            if (a.nfaces > 0)
                hipLaunchKernelGGL(k_face32<P>, dim3((unsigned)((a.nfaces + 1) / 2)), dim3(256), 0, s, a);
            hipLaunchKernelGGL((k_step32w<P, true>), dim3((unsigned)a.ne), dim3(64 * 2), 0, s, a);
            return hipGetLastError() == hipSuccess ? 0 : -1;
        }
    }
    if (a.zf32) return -3;   // no FP32 face form at this degree
    if constexpr (Geom<P>::D > 32)
        hipLaunchKernelGGL(k_step32w<P>, dim3((unsigned)a.ne), dim3(64 * 8), 0, s, a);
    else
        hipLaunchKernelGGL(k_step32c<P>, dim3(nwg_for(a.ne)), dim3(kThreads), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int sample64_p(const StepArgs &a, const double *u, hipStream_t s)
{
    hipLaunchKernelGGL((k_sample<P, double>), dim3(a.nrec), dim3(64), 0, s, a, u, (const float *)nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int sample32_p(const StepArgs &a, const float *hi, const float *lo, hipStream_t s)
{
    hipLaunchKernelGGL((k_sample<P, float>), dim3(a.nrec), dim3(64), 0, s, a, hi, lo);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int minvb_p(int ne, const int4 *meta, const double *blk, const double *invm, const double *minv, const double *x,
            double *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_minvB<P>, dim3(nwg_for(ne)), dim3(kThreads), 0, s, ne, meta, blk, invm, minv, x, y);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int P>
int apply_cm_p(int mode, int ne, const int4 *meta, const double *blk, const double *invm, const double *minv,
               const double *x, double *y, hipStream_t s)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_apply_cm<P, 0>), dim3(nwg_for(ne)), dim3(kThreads), 0, s, ne, meta, blk, invm, minv, x, y);
    else
        hipLaunchKernelGGL((k_apply_cm<P, 1>), dim3(nwg_for(ne)), dim3(kThreads), 0, s, ne, meta, blk, invm, minv, x, y);
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
int launch_sample32(int P, const StepArgs &a, const float *hi, const float *lo, hipStream_t s)
{
#define CALL(p) sample32_p<p>(a, hi, lo, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_apply_minvB(int P, int ne, const int4 *meta, const double *blk, const double *invm,
                       const double *minv, const double *x, double *y, hipStream_t s)
{
#define CALL(p) minvb_p<p>(ne, meta, blk, invm, minv, x, y, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}
int launch_apply_cm(int P, int mode, int ne, const int4 *meta, const double *blk, const double *invm,
                    const double *minv, const double *x, double *y, hipStream_t s)
{
#define CALL(p) apply_cm_p<p>(mode, ne, meta, blk, invm, minv, x, y, s)
    TSP_DISPATCH_P(P, CALL)
#undef CALL
}

int launch_join_state(size_t n, const float *hi, const float *lo, double *u64, hipStream_t s)
{
    hipLaunchKernelGGL(k_join, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, hi, lo, u64);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_scale_demote(size_t n, const double *src, double coef, float *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_demote, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, coef, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_fill_start(size_t n, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_start, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_scale(size_t n, double *x, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, v);
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

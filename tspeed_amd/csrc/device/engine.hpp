// engine.hpp -- shapes, device layouts and kernel-launch entry points of the
// gfx950 leap-frog engine.  Host-visible declarations only; kernels live in
// assemble.hip and step.hip.
//
// Device layouts (element-major, chosen for coalesced streaming):
//   operator  FP64: blk[((e*4 + s)*DH + jp)*2D + row*2 + t] = B_{e,s}(row, 2*jp + t)
//             s = 0 own element, s = 1+f neighbour across edge f; each d x d block
//             is stored as column pairs so one lane (= one row) loads two columns
//             with a single 16-byte load, and a wave's load of one column pair
//             covers 2*d contiguous values.
//             FP32: the same blocks as column quads, [jq][row][4] for the first
//             4*(d/4) columns, followed by [row][2] for the last two columns when
//             d % 4 == 2: again one 16-byte load per lane per group.
//   vectors   u[e*D + c*NLN + i]  (the ABI converts from the reference's
//             component-major layout c*ne*nln + e*nln + i)
//   meta      int4 {nb0, nb1, nb2, aux}  aux = row of `extra` for elements that
//             hold a source or receiver, else -1
//   extra     int4 {src_begin, src_end, rec_begin, rec_end} into the per-element
//             sorted source / receiver index lists
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace tspeed_dev {

constexpr int kMaxDegree = 8;

// Shape of one element's work.  RWF forces the lanes per column group (0 = the
// narrowest power of two >= d, capped at 64).
template <int P, int RWF = 0>
struct Geom {
    static constexpr int NLN = (P + 1) * (P + 2) / 2;  // modes per component
    static constexpr int D = 2 * NLN;                  // rows of a merged block
    static constexpr int DH = D / 2;                   // column pairs per block
    static constexpr int NQ1 = P + 1;                  // Gauss<N+1> edge nodes
    static constexpr int NQ2 = NQ1 * NQ1;              // collapsed area nodes
    // lanes of a wave that hold one column group (one row per lane)
    static constexpr int RW = RWF ? RWF : (D <= 16 ? 16 : (D <= 32 ? 32 : 64));
    static constexpr int NG = 64 / RW;                  // column groups per wave
    static constexpr int RPL = (D + 63) / 64;           // rows per lane when D > 64
    static constexpr int DP = (D + 3) / 4 * 4;          // LDS row stride (16-byte aligned groups)
    static constexpr int ZS = (D + 15) / 16 * 16;       // face-product row stride: whole 128-byte lines
    static constexpr int ZS32 = (D + 31) / 32 * 32;     // the same for FP32 rows
};

inline int nln_of(int P) { return (P + 1) * (P + 2) / 2; }
// Degrees whose FP64 step has the face form: all of them (the face pass splits
// a block's columns over waves for d > 32 and gives a lane group one face for
// d <= 32).
inline bool face_form_degree(int P) { return P >= 1 && P <= kMaxDegree; }
inline int face_zstride(int d) { return (d + 15) / 16 * 16; }
inline int face_zstride32(int d) { return (d + 31) / 32 * 32; }
// Degrees whose FP32c step has the face form: blocks one column group wide
// (d > 32), each half of a block's column quads to a wave as in k_step32w.
inline bool face_form_degree32(int P) { return P >= 5 && P <= kMaxDegree; }

// The FP32 paths keep the displacement as u = hi - lo at two time levels
// (u^n and u^{n-1}), hi in FP32 and lo, its compensation, in BF16: 2 bytes per
// entry (float-sized allocations of half the entries), rounded to nearest even
// on store and widened exactly on load, so u carries ~32 bits.  The increment
// u^n - u^{n-1} is recomputed from the two levels rather than stored.
__device__ inline unsigned short lo_bits(float x)
{
    unsigned b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (unsigned short)(b >> 16);
}
__device__ inline float lo_value(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline float lo_ld(const float *lo, size_t i)
{
    return lo_value(reinterpret_cast<const unsigned short *>(lo)[i]);
}
__device__ inline void lo_st(float *lo, size_t i, float v) { reinterpret_cast<unsigned short *>(lo)[i] = lo_bits(v); }
// u^n - u^{n-1} from (hi, lo) of both levels: uh, cp at step n, hp, cq at n-1.
// Each difference is exact in FP32 when its two values are within a factor 2
// of each other (Sterbenz), as neighbouring time levels are except near a zero
// crossing, where the values, and so the rounding, are small.
__device__ inline float lf_incr(float uh, float hp, float cp, float cq) { return (uh - hp) - (cp - cq); }

// Per-element FP64 geometry produced by the geometry kernel.
struct ElemGeo {
    double inv[4];   // invJ row-major
    double det;      // detJ (signed, as the reference uses it)
    double invm;     // ((1/rho)*1)/detJ
};
// Per half-face data: unit outward normal, 1/length * penalty coefficients.
struct FaceGeo {
    double n0, n1, len;
    double ch00, ch11, ch01;  // c_mn * (1/len) with harmonic-mean Lame parameters
};

struct Tables {
    const double *B, *G, *Be, *Ge, *w2, *w1;
};

struct AssembleArgs {
    int ne;
    const double *verts;
    const int *neigh, *neighedge;
    const double *lambda, *mu, *rho;
    double cpen;        // penalty * N * N
    Tables tab;
    ElemGeo *geo;       // out (and in for the block kernel)
    FaceGeo *face;      // out
    double *blk;        // out: device operator layout, FP64
    double *invm;       // out
};

int launch_assemble(int P, const AssembleArgs &a, hipStream_t s);

// Canonical <-> device block layout conversion and FP32 demotion.
int launch_blocks_to_canonical(int P, int ne, const double *blk, double *canon, hipStream_t s);
int launch_canonical_to_blocks(int P, int ne, const double *canon, double *blk, hipStream_t s);
int launch_premul_invmass(int P, int ne, const double *minv, double *blk, hipStream_t s);
int launch_zero_missing(int P, int ne, const int *neigh, double *blk, hipStream_t s);
int launch_demote_blocks(int P, int ne, const double *src, float *dst, hipStream_t s);

struct StepArgs {
    int ne;
    int nsrc, nrec;
    const int4 *meta;
    const int4 *extra;
    const int *src_sorted, *rec_sorted;
    const int *rec_elem;      // element of each receiver
    const double *src_phi;    // nsrc * NLN
    const double *wav_row;    // wavelet row for this step: nsrc * 2
    const double *rec_phi;    // nrec * NLN
    double *seis;             // [sample][rec][2]
    int sample;               // sample slot to write, -1 = none
    unsigned *flags;
    // FP64 plain
    const double *blk64;
    const double *invm;
    double coef;              // first step: dt*dt/2 ; step: (-dt)*dt
    double dt;                // first step only
    const double *ucur;       // u^n (first step: u0)
    double *uprev;            // u^{n-1} in, u^{n+1} out (first step: v0 in, u1 out)
    const double *minv;       // non-orthonormal basis: base inverse mass (nln x nln),
                              // applied after the block product as the reference does
                              // (M^-1 (B u - f), M^-1 = invm_e * minv); null otherwise
    int wt;                   // 1: FP64 state stored through L2 (operator larger than
                              // the Infinity Cache); 0: plain stores
    // Face form of the FP64 product (face_form_degree, symmetric operator):
    // the face pass (k_face64) applies each interior face's block, read once
    // from the lower-numbered element's copy, to both sides and leaves the
    // products in zf; the element step adds them to its own block's product.
    // faces null: every block is read by the element step (block-ELL).
    const int4 *faces;        // {ea, slot of eb in ea, eb, slot of ea in eb}, ea < eb, slots 1..3
    int nfaces;
    double *zf;               // zf[(e*3 + slot-1)*zs + row], zs = face_zstride(d): rows padded
                              // to whole 128-byte lines, written whole; boundary slots stay zero
    float *zf32;              // FP32c face form (translation-free products): the same rows in
                              // FP32, zs = face_zstride32(d)
    // FP32 compensated: u^n in (uhi, ulo), u^{n-1} in (uhi_out, ulo_out), which
    // the step overwrites with u^{n+1}; lo arrays hold BF16 (lo_ld / lo_st)
    int tfree;                // 1: translation-free products (orthonormal basis; see gather)
    const float *blk32;
    const float *sc32;        // per element (float)(invm * coef)
    const float *uhi;         // gathered displacement (hi part)
    const float *ulo;
    float *uhi_out, *ulo_out;
};

// FP64 step: with a.faces set, the face pass and then the element step.
int launch_step64(int P, bool first, const StepArgs &a, hipStream_t s);
int launch_step32c(int P, bool first, const StepArgs &a, hipStream_t s);

// FP32-compensated state helpers (lo in BF16; splitting is launch_split_b)
int launch_join_state(size_t n, const float *hi, const float *lo, double *u64, hipStream_t s);
int launch_scale_demote(size_t n, const double *src, double coef, float *dst, hipStream_t s);

// Receiver sample of the current displacement (eval_receivers).
int launch_sample64(int P, const StepArgs &a, const double *u, hipStream_t s);
int launch_sample32(int P, const StepArgs &a, const float *hi, const float *lo, hipStream_t s);

// Power iteration pieces for the step-size check: y = invM * B * x (FP64).
int launch_apply_minvB(int P, int ne, const int4 *meta, const double *blk, const double *invm,
                       const double *minv, const double *x, double *y, hipStream_t s);
int launch_dot(size_t n, const double *x, const double *y, double *partial, int nparts, hipStream_t s);
int launch_fill_start(size_t n, double *x, hipStream_t s);   // power-iteration start vector
int launch_scale(size_t n, double *x, double v, hipStream_t s);
// Operator plugin on component-major vectors: mode 0 y = B x, mode 1 y = invM x.
int launch_apply_cm(int P, int mode, int ne, const int4 *meta, const double *blk, const double *invm,
                    const double *minv, const double *x, double *y, hipStream_t s);

// Many right-hand sides (spmm.hip): R = ncols columns, state in row octets
// u[((e*KB + k/8)*R + j)*8 + k%8], KB = KP/8; the lo arrays hold BF16 entries
// at the same indices (2 bytes each).
struct BatchArgs {
    int ne, nsrc, nrec;
    const int4 *meta;
    const int4 *extra;
    const int *src_sorted, *rec_sorted;
    const int *src_col, *rec_elem;
    const double *src_phi, *wav_row, *rec_phi;
    double *seis;             // [sample][rec][col][2]
    int sample;
    unsigned *flags;
    const float *blk;         // A-fragment operator
    int tfree;                // 1: translation-free products (orthonormal basis)
    const float *sc32;
    // Two time levels of the compensated displacement, u = hi - lo (hi FP32,
    // lo BF16 stored in 2 bytes per entry): u^n in (uhi, ulo); u^{n-1} in
    // (uhi_out, ulo_out), which the step overwrites with u^{n+1}.  The
    // increment u^n - u^{n-1} is recomputed from them, not stored.
    const float *uhi, *ulo;
    float *uhi_out, *ulo_out;
    // Graph replay: when ctr is set, the step reads its 1-based index
    // n = *ctr + 1 + koff (*ctr = steps done before the replay, koff = the step's
    // place in the captured run) and derives the force row, the sample slot
    // (n / stride) and the ping-pong buffers (u^n in uhi_pair[n & 1],
    // ulo_pair[n & 1]) on the device; wav_row, sample and the four state
    // pointers above are then ignored.
    const int *ctr;
    int koff;
    const double *wav;
    int stride, max_samples;
    float *uhi_pair[2], *ulo_pair[2];
};
// One multi-column context inside a multi-instance set (multi.hip): the
// arguments of its graph-replayed step (ctr null: the step index comes from
// `steps`) and its progress.
struct BatchInst {
    BatchArgs a;
    int steps, nsteps;        // completed step() calls, steps to run in the set
    int live, sel;            // 0 once removed / not in the current item list
    int P, R;
};
// The step kernel of every multi-column instance of degree P and width R in one
// launch, arguments (const BatchInst *, const int2 *items, int nitems), items =
// (instance, work item); work items per instance = ne * batch_tiles(R), and
// each work-group of 256 threads takes items_per_wg items.  0, or -2 if the
// degree or width is not built.
int batch_multi_kernel(int P, int R, const void **kern, int *items_per_wg);
int batch_tiles(int R);              // 32-column tiles per element

int launch_ctr_set(int *ctr, int value, hipStream_t s);
int launch_ctr_add(int *ctr, int value, hipStream_t s);
int batch_kp(int P);                 // state rows per element (d padded to a multiple of 16)
size_t batch_block_floats(int P);    // floats of one element's four packed blocks
int launch_step_b32(int P, int ncols, const BatchArgs &a, hipStream_t s);
int launch_pack_mfma(int P, int ne, const int4 *meta, const double *src, float *dst, hipStream_t s);
// split BF16 products on the FP32 operator layout: nprod = 3 or 5
int launch_step_b16(int P, int ncols, int nprod, const BatchArgs &a, hipStream_t s);
int launch_sample_b32(int P, int ncols, const BatchArgs &a, hipStream_t s);
// multi-column state: FP64 staging -> FP32 hi + BF16 compensation (spmm.hip)
int launch_split_b(size_t n, const double *u64, float *hi, float *comp, hipStream_t s);
// u_prev = u0 - dt * v, in place over v (FP64): the multi-column state's u^{-1}
int launch_prev_b(size_t n, const double *u0, double *v, double dt, hipStream_t s);
int launch_col_in(int ne, int nln, int KP, int R, int col, const double *cm, double *st, hipStream_t s);
int launch_col_out(int ne, int nln, int KP, int R, int col, const float *hi, const float *comp, double *cm,
                   hipStream_t s);

// Layout helpers (element-major <-> component-major), column-agnostic.
int launch_to_elem_major(int ne, int nln, const double *cm, double *em, hipStream_t s);
int launch_to_comp_major(int ne, int nln, const double *em, double *cm, hipStream_t s);

}  // namespace tspeed_dev

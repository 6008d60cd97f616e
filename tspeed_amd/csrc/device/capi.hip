// capi.hip -- implementation of the C-ABI declared in include/tspeed_amd.h.
//
// The context owns every device buffer of one simulation (or of ncols columns
// sharing one mesh).  Nothing here computes on the host: set-up data is copied
// in, assembled and stepped by the kernels of assemble.hip / step.hip, and
// results are copied out on request.  Errors never throw across the ABI: each
// entry point catches, records a message and returns a status code.
#include "engine.hpp"
#include "internal.hpp"
#include "../../../include/tspeed_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace tspeed_dev;

struct tsp_context {
    int device = 0, precision = TSP_FP64, ncols = 1;
    hipStream_t own_stream = nullptr, stream = nullptr;
    std::string err;
    long long bytes = 0;
    std::vector<void *> allocs;

    // tables
    int P = 0, nln = 0, D = 0, nq1 = 0, nq2 = 0;
    double *t_B = nullptr, *t_G = nullptr, *t_Be = nullptr, *t_Ge = nullptr, *t_w2 = nullptr, *t_w1 = nullptr;

    // mesh / operator
    int ne = 0;
    std::vector<int> h_neigh;
    int4 *meta = nullptr;
    double *blk64 = nullptr, *invm = nullptr;
    float *blk32 = nullptr, *sc_first = nullptr, *sc_step = nullptr;
    // FP64 face form (StepArgs::faces): taken when the operator is symmetric
    // (assembled by tsp_set_mesh, or uploaded blocks symmetric bit for bit),
    // the degree has it, every neighbour relation is mutual and the caller has
    // not turned it off (tsp_set_face_form)
    bool symmetric = false;
    int face_form = 1;   // tsp_set_face_form: 0 never, 1 above the cache-resident size, 2 always
    int4 *faces = nullptr;
    int nfaces = 0;
    double *zf = nullptr;
    float *zf32 = nullptr;   // the FP32c face form's product rows

    // state: FP64 two vectors; FP32 (single- and multi-column) two time levels of
    // u = hi - lo, hi FP32 and lo BF16, both ping-pong
    double *u64[2] = {nullptr, nullptr};
    int cur = 0;
    float *uhi[2] = {nullptr, nullptr};
    float *ulo[2] = {nullptr, nullptr};   // 2 bytes per entry (BF16)
    double *scratch64 = nullptr;  // 2*ne*nln staging for layout conversion
    bool have_state = false;

    // sources / receivers
    int nsrc = 0, ntab = 0;
    std::vector<int> h_src_elem, h_src_col;
    double *src_phi = nullptr, *wav = nullptr;
    int nrec = 0, stride = 0, max_samples = 0, nsamples = 0;
    std::vector<int> h_rec_elem;
    double *rec_phi = nullptr, *seis = nullptr;
    int *rec_elem = nullptr;
    int4 *extra = nullptr;
    int *src_sorted = nullptr, *rec_sorted = nullptr, *src_col = nullptr;

    // many columns (ncols > 1): packed MFMA operator, state rows padded to KP
    float *blkA = nullptr;
    int KP = 0;
    // multi-column stepping replayed from a captured graph of graph_len steps;
    // the device step counter feeds the per-step values (BatchArgs::ctr)
    int *d_ctr = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_len = 0;
    // non-orthonormal basis: base_invmass (nln x nln), folded into the operator
    // rows after assembly and applied to each source's basis values
    std::vector<double> h_minv;
    double *d_minv = nullptr;
    bool batched() const { return ncols > 1; }
    // non-orthonormal basis: FP64 contexts apply the dense inverse mass after
    // the block product (the reference's order); FP32C and multi-column
    // contexts fold it into the operator rows and the sources
    bool minv_after() const { return precision == TSP_FP64 && !h_minv.empty(); }
    size_t state_len() const { return batched() ? (size_t)ne * KP * ncols : (size_t)ne * D; }

    unsigned *flags = nullptr;
    double dt = 0;
    int steps = 0;        // completed step() calls
    bool first_done = false;

    // tsp_share_operator: the context whose operator this one borrows, and the
    // contexts borrowing this one's (kept so that neither side is left holding
    // pointers into buffers the other has freed)
    tsp_context *lender = nullptr;
    std::vector<tsp_context *> borrowers;
};

namespace {

struct Fail {
    int code;
};

void fail(tsp_context *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
    throw Fail{code};
}

void hipcheck(tsp_context *c, hipError_t e, const char *what)
{
    if (e != hipSuccess) fail(c, TSP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

void kcheck(tsp_context *c, int rc, const char *what)
{
    if (rc == -2) fail(c, TSP_ERR_UNSUPPORTED, "%s: degree %d not built", what, c->P);
    if (rc == -3) fail(c, TSP_ERR_STATE, "%s: no face form at degree %d", what, c->P);
    if (rc != 0) fail(c, TSP_ERR_HIP, "%s: launch failed: %s", what, hipGetErrorString(hipGetLastError()));
}

template <typename T>
T *dalloc(tsp_context *c, size_t n, const char *what)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(c, TSP_ERR_NOMEM, "%s: cannot allocate %zu bytes on device %d", what, n * sizeof(T), c->device);
    }
    c->allocs.push_back(p);
    c->bytes += (long long)(n * sizeof(T));
    return static_cast<T *>(p);
}

template <typename T>
void dfree(tsp_context *c, T *&p)
{
    if (!p) return;
    // pointers borrowed from another context (tsp_share_operator) are not ours to free
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)p);
    if (it != c->allocs.end()) {
        c->allocs.erase(it);
        (void)hipFree(p);
    }
    p = nullptr;
}

// A borrower leaves its lender's list (it keeps nothing it has to free).
void detach_lender(tsp_context *c)
{
    if (!c->lender) return;
    auto &b = c->lender->borrowers;
    b.erase(std::remove(b.begin(), b.end(), c), b.end());
    c->lender = nullptr;
}

// The lender's buffers are about to be freed: every borrower drops its copies of
// those pointers, so later calls on it report TSP_ERR_STATE instead of reading
// freed memory.
void orphan_borrowers(tsp_context *c)
{
    for (tsp_context *b : c->borrowers) {
        b->lender = nullptr;
        b->blk64 = nullptr;
        b->blk32 = nullptr;
        b->blkA = nullptr;
        b->invm = nullptr;
        b->d_minv = nullptr;
    }
    c->borrowers.clear();
}

template <typename T>
void h2d(tsp_context *c, T *dst, const T *src, size_t n, const char *what)
{
    hipcheck(c, hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream), what);
}

template <typename T>
void d2h(tsp_context *c, T *dst, const T *src, size_t n, const char *what)
{
    hipcheck(c, hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream), what);
    hipcheck(c, hipStreamSynchronize(c->stream), what);
}

void set_device(tsp_context *c) { hipcheck(c, hipSetDevice(c->device), "hipSetDevice"); }

template <typename F>
int guarded(tsp_context *c, F &&f)
{
    if (!c) return TSP_ERR_ARG;
    try {
        set_device(c);
        f();
        c->err.clear();
        return TSP_OK;
    } catch (const Fail &x) {
        return x.code;
    } catch (const std::exception &x) {
        c->err = x.what();
        return TSP_ERR_ARG;
    }
}

// Rebuild the per-element source/receiver lists and the aux column of meta.
// a captured graph holds buffer pointers, strides and sample limits: any call
// that changes them drops it
void drop_graph(tsp_context *c)
{
    if (c->graph) {
        // tsp_run only queues work: a replay of this graph may still be running
        (void)hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(c->graph);
    }
    c->graph = nullptr;
    c->graph_len = 0;
}

void rebuild_extra(tsp_context *c)
{
    drop_graph(c);
    if (!c->ne) return;
    std::vector<std::vector<int>> s(c->ne), r(c->ne);
    for (int k = 0; k < c->nsrc; ++k) s[c->h_src_elem[k]].push_back(k);
    for (int k = 0; k < c->nrec; ++k) r[c->h_rec_elem[k]].push_back(k);
    std::vector<int4> ex;
    std::vector<int> ss, rs;
    std::vector<int4> meta(c->ne);
    for (int e = 0; e < c->ne; ++e) {
        meta[e] = make_int4(c->h_neigh[3 * e], c->h_neigh[3 * e + 1], c->h_neigh[3 * e + 2], -1);
        if (s[e].empty() && r[e].empty()) continue;
        int4 x;
        x.x = (int)ss.size();
        ss.insert(ss.end(), s[e].begin(), s[e].end());
        x.y = (int)ss.size();
        x.z = (int)rs.size();
        rs.insert(rs.end(), r[e].begin(), r[e].end());
        x.w = (int)rs.size();
        meta[e].w = (int)ex.size();
        ex.push_back(x);
    }
    dfree(c, c->extra);
    dfree(c, c->src_sorted);
    dfree(c, c->rec_sorted);
    dfree(c, c->src_col);
    c->src_col = dalloc<int>(c, c->nsrc, "src columns");
    if (c->nsrc) h2d(c, c->src_col, c->h_src_col.data(), c->nsrc, "src columns");
    c->extra = dalloc<int4>(c, ex.size(), "extra");
    c->src_sorted = dalloc<int>(c, ss.size(), "src list");
    c->rec_sorted = dalloc<int>(c, rs.size(), "rec list");
    if (!ex.empty()) h2d(c, c->extra, ex.data(), ex.size(), "extra");
    if (!ss.empty()) h2d(c, c->src_sorted, ss.data(), ss.size(), "src list");
    if (!rs.empty()) h2d(c, c->rec_sorted, rs.data(), rs.size(), "rec list");
    h2d(c, c->meta, meta.data(), meta.size(), "meta");
    hipcheck(c, hipStreamSynchronize(c->stream), "rebuild_extra");
}

void require_mesh(tsp_context *c)
{
    if (!c->ne || !(c->blk64 || c->blk32 || c->blkA))
        fail(c, TSP_ERR_STATE, "no operator: call tsp_set_mesh or tsp_upload_blocks first");
}

void require_fp64_operator(tsp_context *c)
{
    require_mesh(c);
    if (!c->blk64) fail(c, TSP_ERR_STATE, "the FP64 operator was released (tsp_release_fp64)");
}

void alloc_state(tsp_context *c)
{
    const size_t n = c->state_len();
    if (!c->scratch64) c->scratch64 = dalloc<double>(c, (size_t)c->ne * c->D, "scratch");
    if (c->precision == TSP_FP64) {
        for (int k = 0; k < 2; ++k)
            if (!c->u64[k]) c->u64[k] = dalloc<double>(c, n, "state");
    } else {
        const bool fresh = !c->uhi[0];
        // n is even (d = 2 nln rows per element, or KP, a multiple of 16):
        // n / 2 floats hold n BF16
        for (int k = 0; k < 2; ++k) {
            if (!c->uhi[k]) c->uhi[k] = dalloc<float>(c, n, "state hi");
            if (!c->ulo[k]) c->ulo[k] = dalloc<float>(c, n / 2, "state lo");
        }
        if (!c->u64[0]) c->u64[0] = dalloc<double>(c, n, "state in");
        if (!c->u64[1]) c->u64[1] = dalloc<double>(c, n, "state in");
        if (fresh && c->batched()) {
            // the multi-column step reads and writes whole row octets: the
            // padded rows must start (and so stay) zero in every buffer
            for (int k = 0; k < 2; ++k) {
                hipcheck(c, hipMemsetAsync(c->uhi[k], 0, n * sizeof(float), c->stream), "memset");
                hipcheck(c, hipMemsetAsync(c->ulo[k], 0, n / 2 * sizeof(float), c->stream), "memset");
                hipcheck(c, hipMemsetAsync(c->u64[k], 0, n * sizeof(double), c->stream), "memset");
            }
        }
    }
}

BatchArgs batch_args(tsp_context *c);

int launch_batched(tsp_context *c, const BatchArgs &b)
{
    if (c->precision == TSP_BF16X3) return launch_step_b16(c->P, c->ncols, 3, b, c->stream);
    if (c->precision == TSP_BF16X5) return launch_step_b16(c->P, c->ncols, 5, b, c->stream);
    return launch_step_b32(c->P, c->ncols, b, c->stream);
}

// Multi-column runs replay one captured graph of kGraphSteps steps per
// context: each captured step reads its index from a device counter (set once
// per tsp_run) plus its place in the graph, and the graph ends by advancing the
// counter, so replays follow each other with no host work between them.  The
// rest of a run (and runs shorter than the graph) are launched step by step;
// the kernels are the same, so the results are too.  The graph is captured on
// a context's first long run and kept until a set-up call changes what it
// holds: a running request is never re-captured and no graph is destroyed
// while others run.  (Single-column steps are launched directly: replaying
// them measured no faster.)
constexpr int kGraphSteps = 32;

// samples so far are exactly the stride-driven ones (the graph derives the
// sample slot of step n as n / stride)
bool regular_samples(const tsp_context *c)
{
    if (c->nrec == 0 || c->stride == 0) return true;
    return c->nsamples == std::min(c->max_samples, 1 + c->steps / c->stride);
}

// Capture the context's graph of kGraphSteps multi-column steps (once).
void ensure_graph(tsp_context *c)
{
    if (c->graph) return;
    if (!c->d_ctr) c->d_ctr = dalloc<int>(c, 1, "step counter");
    BatchArgs b = batch_args(c);
    b.sc32 = c->sc_step;
    b.ctr = c->d_ctr;
    b.wav = c->wav;
    b.stride = c->nrec ? c->stride : 0;
    b.max_samples = c->max_samples;
    b.uhi_pair[0] = c->uhi[0];
    b.uhi_pair[1] = c->uhi[1];
    b.ulo_pair[0] = c->ulo[0];
    b.ulo_pair[1] = c->ulo[1];
    hipcheck(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "capture");
    int rc = 0;
    for (int k = 0; k < kGraphSteps && !rc; ++k) {
        b.koff = k;  // the step's place in the graph: no counter kernel between steps
        rc = launch_batched(c, b);
    }
    if (!rc) rc = launch_ctr_add(c->d_ctr, kGraphSteps, c->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        fail(c, TSP_ERR_HIP, "graph capture of %d steps failed", kGraphSteps);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hipcheck(c, ei, "graph instantiate");
    c->graph = x;
    c->graph_len = kGraphSteps;
}

// Whole graphs of the run; returns the steps left for single launches.
int run_graph(tsp_context *c, int nsteps)
{
    ensure_graph(c);
    const int reps = nsteps / c->graph_len;
    kcheck(c, launch_ctr_set(c->d_ctr, c->steps, c->stream), "step counter");
    for (int r = 0; r < reps; ++r) hipcheck(c, hipGraphLaunch(c->graph, c->stream), "graph launch");
    c->steps += reps * c->graph_len;
    c->cur = (c->steps + 1) & 1;
    if (c->nrec > 0 && c->stride > 0) c->nsamples = std::min(c->max_samples, 1 + c->steps / c->stride);
    return nsteps - reps * c->graph_len;
}

// FP64 operators larger than this (the 256 MB Infinity Cache less room for
// the state) take the store-through-L2 step and, by default, the face form;
// smaller ones stay cache-resident between steps and do better with plain
// stores and with every block read by one launch (3200 triangles at p = 4:
// 14.5 us per step against 15.9 with the face form's two launches; 1600 at
// p = 6, a 160 MB operator: 27.7 against 20.7).
constexpr double kCacheResidentBytes = 128.0 * (1 << 20);

// Face list of the FP64 face form: one entry per interior face, from the
// lower-numbered element's side.  Left empty (the block-ELL step) unless every
// neighbour relation is mutual and single, so that each (element, slot) with a
// neighbour is written by exactly one face.
void build_faces(tsp_context *c)
{
    drop_graph(c);
    dfree(c, c->faces);
    dfree(c, c->zf);
    dfree(c, c->zf32);
    c->nfaces = 0;
    // FP64, or FP32c on the orthonormal basis (its FP32 operator is the FP64
    // one rounded; with a dense inverse mass folded in it is not symmetric)
    const bool fp32 = c->precision == TSP_FP32C;
    if (c->batched() || !c->symmetric || !c->ne) return;
    if (!(c->precision == TSP_FP64 && face_form_degree(c->P)) &&
        !(fp32 && c->h_minv.empty() && c->blk32 && face_form_degree32(c->P)))
        return;
    const double ell_bytes = (double)c->ne * 4 * c->D * c->D * (fp32 ? sizeof(float) : sizeof(double));
    if (c->face_form == 0 || (c->face_form == 1 && ell_bytes <= kCacheResidentBytes)) return;
    const std::vector<int> &nb = c->h_neigh;
    auto count = [&](int e, int x) {
        return (nb[3 * e] == x) + (nb[3 * e + 1] == x) + (nb[3 * e + 2] == x);
    };
    std::vector<int4> f;
    f.reserve((size_t)c->ne * 3 / 2 + 1);
    for (int e = 0; e < c->ne; ++e)
        for (int s = 0; s < 3; ++s) {
            const int x = nb[3 * e + s];
            if (x < 0) continue;
            if (x == e || x >= c->ne || count(e, x) != 1 || count(x, e) != 1) return;
            if (x < e) continue;
            const int sx = nb[3 * x] == e ? 0 : (nb[3 * x + 1] == e ? 1 : 2);
            f.push_back(make_int4(e, 1 + s, x, 1 + sx));
        }
    // boundary slots are never written: they read zero
    if (fp32) {
        const size_t nz = (size_t)c->ne * 3 * face_zstride32(c->D);
        c->zf32 = dalloc<float>(c, nz, "face products");
        hipcheck(c, hipMemsetAsync(c->zf32, 0, sizeof(float) * nz, c->stream), "memset");
    } else {
        const size_t nz = (size_t)c->ne * 3 * face_zstride(c->D);
        c->zf = dalloc<double>(c, nz, "face products");
        hipcheck(c, hipMemsetAsync(c->zf, 0, sizeof(double) * nz, c->stream), "memset");
    }
    c->faces = dalloc<int4>(c, f.size(), "faces");
    if (!f.empty()) h2d(c, c->faces, f.data(), f.size(), "faces");
    c->nfaces = (int)f.size();
    hipcheck(c, hipStreamSynchronize(c->stream), "faces");
}

// Blocks in the canonical form (tsp_upload_blocks) with B_{b,a} = B_{a,b}^T bit
// for bit on every face whose neighbour relation is mutual.
bool blocks_symmetric(int ne, int d, const int *neigh, const double *blocks)
{
    const size_t bs = (size_t)d * d;
    for (int a = 0; a < ne; ++a)
        for (int s = 0; s < 3; ++s) {
            const int b = neigh[3 * a + s];
            if (b <= a || b >= ne) continue;
            int sb = -1;
            for (int k = 0; k < 3; ++k)
                if (neigh[3 * b + k] == a) sb = k;
            if (sb < 0) return false;
            const double *F = blocks + ((size_t)a * 4 + 1 + s) * bs;
            const double *Fb = blocks + ((size_t)b * 4 + 1 + sb) * bs;
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j)
                    if (F[(size_t)i * d + j] != Fb[(size_t)j * d + i]) return false;
        }
    return true;
}

void operator_ready(tsp_context *c)
{
    // neighbour table on the device first: the multi-column packing reads it
    rebuild_extra(c);
    // FP32 copy of the operator for the compensated path
    if (c->batched()) {
        c->KP = batch_kp(c->P);
        if (!c->blkA) c->blkA = dalloc<float>(c, (size_t)c->ne * batch_block_floats(c->P), "blocks mfma");
        // FP32 and split BF16 products read the same FP32 operator layout
        kcheck(c, launch_pack_mfma(c->P, c->ne, c->meta, c->blk64, c->blkA, c->stream), "pack");
    } else if (c->precision == TSP_FP32C) {
        const size_t n = (size_t)c->ne * 4 * c->D * c->D;
        if (!c->blk32) c->blk32 = dalloc<float>(c, n, "blocks fp32");
        kcheck(c, launch_demote_blocks(c->P, c->ne, c->blk64, c->blk32, c->stream), "demote");
    }
    if (c->precision != TSP_FP64) {
        if (!c->sc_first) c->sc_first = dalloc<float>(c, c->ne, "sc");
        if (!c->sc_step) c->sc_step = dalloc<float>(c, c->ne, "sc");
    }
    alloc_state(c);
    hipcheck(c, hipMemsetAsync(c->u64[0], 0, sizeof(double) * c->state_len(), c->stream), "memset");
    hipcheck(c, hipMemsetAsync(c->u64[1], 0, sizeof(double) * c->state_len(), c->stream), "memset");
    c->have_state = true;
    c->steps = 0;
    c->first_done = false;
    c->nsamples = 0;
    hipcheck(c, hipStreamSynchronize(c->stream), "operator");
    build_faces(c);
}


StepArgs base_args(tsp_context *c)
{
    StepArgs a{};
    a.ne = c->ne;
    a.nsrc = c->nsrc;
    a.nrec = c->nrec;
    a.meta = c->meta;
    a.extra = c->extra;
    a.src_sorted = c->src_sorted;
    a.rec_sorted = c->rec_sorted;
    a.rec_elem = c->rec_elem;
    a.src_phi = c->src_phi;
    a.rec_phi = c->rec_phi;
    a.seis = c->seis;
    a.sample = -1;
    a.flags = c->flags;
    a.blk64 = c->blk64;
    a.invm = c->invm;
    a.minv = c->minv_after() ? c->d_minv : nullptr;
    a.wt = (double)c->ne * 4 * c->D * c->D * sizeof(double) > kCacheResidentBytes;
    a.faces = c->faces;
    a.nfaces = c->nfaces;
    a.zf = c->zf;
    a.zf32 = c->zf32;
    a.blk32 = c->blk32;
    // FP32 products drop the element's translation (see gather); the
    // BoundaryAdapted basis has no single constant mode, so it keeps the plain form
    a.tfree = c->h_minv.empty() ? 1 : 0;
    return a;
}

BatchArgs batch_args(tsp_context *c)
{
    BatchArgs a{};
    a.ne = c->ne;
    a.nsrc = c->nsrc;
    a.nrec = c->nrec;
    a.meta = c->meta;
    a.extra = c->extra;
    a.src_sorted = c->src_sorted;
    a.rec_sorted = c->rec_sorted;
    a.src_col = c->src_col;
    a.rec_elem = c->rec_elem;
    a.src_phi = c->src_phi;
    a.rec_phi = c->rec_phi;
    a.seis = c->seis;
    a.sample = -1;
    a.flags = c->flags;
    a.blk = c->blkA;
    a.tfree = c->h_minv.empty() ? 1 : 0;
    a.uhi = c->uhi[c->cur];
    a.ulo = c->ulo[c->cur];
    a.uhi_out = c->uhi[1 - c->cur];
    a.ulo_out = c->ulo[1 - c->cur];
    return a;
}

const double *wav_row(tsp_context *c, int row)
{
    if (c->nsrc == 0) return nullptr;
    if (row >= c->ntab)
        fail(c, TSP_ERR_ARG, "wavelet table has %d rows; step %d needs row %d (extend it with tsp_set_wavelet)",
             c->ntab, row, row);
    return c->wav + (size_t)row * c->nsrc * 2;
}

int take_sample_slot(tsp_context *c)
{
    if (c->nrec == 0) return -1;
    if (c->nsamples >= c->max_samples) return -1;
    return c->nsamples++;
}

}  // namespace

extern "C" {

int tsp_device_count(int *n)
{
    if (!n) return TSP_ERR_ARG;
    *n = 0;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) {
        (void)hipGetLastError();
        return TSP_OK;
    }
    *n = k;
    return TSP_OK;
}

int tsp_device_memory(int device, long long *free_bytes, long long *total_bytes)
{
    if (!free_bytes || !total_bytes) return TSP_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return TSP_ERR_ARG;
    }
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) {
        (void)hipGetLastError();
        return TSP_ERR_HIP;
    }
    *free_bytes = (long long)f;
    *total_bytes = (long long)t;
    return TSP_OK;
}

int tsp_create(int device, int precision, int ncols, tsp_context **out)
{
    if (!out) return TSP_ERR_ARG;
    *out = nullptr;
    const bool split = precision == TSP_BF16X3 || precision == TSP_BF16X5;
    if (precision != TSP_FP64 && precision != TSP_FP32C && !split) return TSP_ERR_UNSUPPORTED;
    if (ncols != 1 && ncols != 8 && ncols != 16 && ncols != 32 && ncols != 64) return TSP_ERR_UNSUPPORTED;
    if (ncols > 1 && precision == TSP_FP64) return TSP_ERR_UNSUPPORTED;
    if (ncols == 1 && split) return TSP_ERR_UNSUPPORTED;
    tsp_context *c = new (std::nothrow) tsp_context;
    if (!c) return TSP_ERR_NOMEM;
    c->device = device;
    c->precision = precision;
    c->ncols = ncols;
    // the face form by default for FP64 only: an FP32c context steps bitwise as
    // it does inside a multi-instance set (tsp_multi_add) unless asked otherwise
    c->face_form = precision == TSP_FP64 ? 1 : 0;
    int rc = guarded(c, [&] {
        int n = 0;
        hipcheck(c, hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) fail(c, TSP_ERR_ARG, "device %d out of range (%d devices)", device, n);
        set_device(c);
        hipcheck(c, hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking), "hipStreamCreate");
        c->stream = c->own_stream;
        c->flags = dalloc<unsigned>(c, 1, "flags");
        hipcheck(c, hipMemsetAsync(c->flags, 0, sizeof(unsigned), c->stream), "memset");
    });
    if (rc != TSP_OK) {
        tsp_destroy(c);
        return rc;
    }
    *out = c;
    return TSP_OK;
}

void tsp_destroy(tsp_context *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (tsp_context *b : c->borrowers)   // their queued steps may read this operator
        if (b->stream) (void)hipStreamSynchronize(b->stream);
    orphan_borrowers(c);
    detach_lender(c);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

const char *tsp_last_error(const tsp_context *c) { return c ? c->err.c_str() : "null context"; }

int tsp_set_stream(tsp_context *c, void *s)
{
    return guarded(c, [&] { c->stream = s ? static_cast<hipStream_t>(s) : c->own_stream; });
}

int tsp_set_tables(tsp_context *c, int N, const double *B, const double *G, const double *Be, const double *Ge,
                   const double *w2, const double *w1)
{
    return guarded(c, [&] {
        if (N < 1 || N > kMaxDegree) fail(c, TSP_ERR_UNSUPPORTED, "degree %d outside 1..%d", N, kMaxDegree);
        if (!B || !G || !Be || !Ge || !w2 || !w1) fail(c, TSP_ERR_ARG, "null table pointer");
        if (c->P && c->P != N && c->ne) fail(c, TSP_ERR_STATE, "degree change after the mesh was set");
        if (c->P != N) {
            c->h_minv.clear();
            dfree(c, c->d_minv);
        }
        c->P = N;
        c->nln = nln_of(N);
        c->D = 2 * c->nln;
        c->nq1 = N + 1;
        c->nq2 = c->nq1 * c->nq1;
        dfree(c, c->t_B); dfree(c, c->t_G); dfree(c, c->t_Be); dfree(c, c->t_Ge); dfree(c, c->t_w2); dfree(c, c->t_w1);
        c->t_B = dalloc<double>(c, (size_t)c->nq2 * c->nln, "tables");
        c->t_G = dalloc<double>(c, (size_t)c->nln * 2 * c->nq2, "tables");
        c->t_Be = dalloc<double>(c, (size_t)3 * c->nq1 * c->nln, "tables");
        c->t_Ge = dalloc<double>(c, (size_t)c->nln * 3 * 2 * c->nq1, "tables");
        c->t_w2 = dalloc<double>(c, c->nq2, "tables");
        c->t_w1 = dalloc<double>(c, c->nq1, "tables");
        h2d(c, c->t_B, B, (size_t)c->nq2 * c->nln, "tables");
        h2d(c, c->t_G, G, (size_t)c->nln * 2 * c->nq2, "tables");
        h2d(c, c->t_Be, Be, (size_t)3 * c->nq1 * c->nln, "tables");
        h2d(c, c->t_Ge, Ge, (size_t)c->nln * 3 * 2 * c->nq1, "tables");
        h2d(c, c->t_w2, w2, c->nq2, "tables");
        h2d(c, c->t_w1, w1, c->nq1, "tables");
        hipcheck(c, hipStreamSynchronize(c->stream), "tables");
    });
}

int tsp_set_mass(tsp_context *c, const double *base_invmass)
{
    return guarded(c, [&] {
        if (!c->P) fail(c, TSP_ERR_STATE, "call tsp_set_tables before tsp_set_mass");
        if (c->ne) fail(c, TSP_ERR_STATE, "call tsp_set_mass before tsp_set_mesh / tsp_upload_blocks");
        c->h_minv.clear();
        dfree(c, c->d_minv);
        if (!base_invmass) return;  // back to the orthonormal (identity) case
        for (int k = 0; k < c->nln * c->nln; ++k)
            if (!std::isfinite(base_invmass[k])) fail(c, TSP_ERR_ARG, "non-finite base_invmass entry %d", k);
        c->h_minv.assign(base_invmass, base_invmass + (size_t)c->nln * c->nln);
        c->d_minv = dalloc<double>(c, (size_t)c->nln * c->nln, "invmass");
        h2d(c, c->d_minv, base_invmass, (size_t)c->nln * c->nln, "invmass");
        hipcheck(c, hipStreamSynchronize(c->stream), "invmass");
    });
}

static void check_mesh_args(tsp_context *c, int ne, const int *neigh, const int *neighedge)
{
    if (ne <= 0) fail(c, TSP_ERR_ARG, "ne must be positive (got %d)", ne);
    for (int q = 0; q < 3 * ne; ++q) {
        const int nb = neigh[q];
        if (nb < -1 || nb >= ne) fail(c, TSP_ERR_ARG, "neigh[%d] = %d out of range", q, nb);
        if (nb == q / 3) fail(c, TSP_ERR_ARG, "element %d is its own neighbour", q / 3);
        if (nb >= 0 && neighedge) {
            const int g = neighedge[q];
            if (g < 0 || g > 2) fail(c, TSP_ERR_ARG, "neighedge[%d] = %d out of range", q, g);
            if (neigh[3 * nb + g] != q / 3)
                fail(c, TSP_ERR_ARG, "neighbour relation not symmetric at element %d edge %d", q / 3, q % 3);
        }
    }
}

static void reset_mesh(tsp_context *c, int ne)
{
    if (!c->borrowers.empty())
        fail(c, TSP_ERR_STATE, "%zu context(s) borrow this context's operator (tsp_share_operator): destroy them first",
             c->borrowers.size());
    detach_lender(c);   // a borrower that sets its own mesh stops borrowing
    drop_graph(c);
    dfree(c, c->meta); dfree(c, c->blk64); dfree(c, c->invm); dfree(c, c->blk32); dfree(c, c->blkA);
    dfree(c, c->sc_first); dfree(c, c->sc_step);
    dfree(c, c->faces); dfree(c, c->zf); dfree(c, c->zf32);
    c->nfaces = 0;
    c->symmetric = false;
    for (int k = 0; k < 2; ++k) { dfree(c, c->u64[k]); dfree(c, c->uhi[k]); dfree(c, c->ulo[k]); }
    dfree(c, c->scratch64);
    c->ne = ne;
    c->nsrc = 0; c->h_src_elem.clear(); c->h_src_col.clear(); c->ntab = 0;
    dfree(c, c->src_phi); dfree(c, c->wav);
    c->nrec = 0; c->h_rec_elem.clear(); dfree(c, c->rec_phi); dfree(c, c->seis); dfree(c, c->rec_elem);
    c->meta = dalloc<int4>(c, ne, "meta");
    c->blk64 = dalloc<double>(c, (size_t)ne * 4 * c->D * c->D, "blocks");
    c->invm = dalloc<double>(c, ne, "invm");
}

int tsp_set_mesh(tsp_context *c, int ne, const double *verts, const int *neigh, const int *neighedge,
                 const double *lambda, const double *mu, const double *rho, double penalty)
{
    return guarded(c, [&] {
        if (!c->P) fail(c, TSP_ERR_STATE, "call tsp_set_tables before tsp_set_mesh");
        if (!verts || !neigh || !neighedge || !lambda || !mu || !rho) fail(c, TSP_ERR_ARG, "null mesh pointer");
        check_mesh_args(c, ne, neigh, neighedge);
        for (int e = 0; e < ne; ++e)
            if (!(rho[e] > 0) || !std::isfinite(lambda[e]) || !std::isfinite(mu[e]))
                fail(c, TSP_ERR_ARG, "bad material at element %d", e);
        reset_mesh(c, ne);
        c->h_neigh.assign(neigh, neigh + 3 * ne);
        double *d_verts = dalloc<double>(c, (size_t)ne * 6, "verts");
        int *d_nb = dalloc<int>(c, (size_t)ne * 3, "neigh");
        int *d_ng = dalloc<int>(c, (size_t)ne * 3, "neighedge");
        double *d_mat = dalloc<double>(c, (size_t)ne * 3, "materials");
        ElemGeo *d_geo = dalloc<ElemGeo>(c, ne, "geometry");
        FaceGeo *d_face = dalloc<FaceGeo>(c, (size_t)ne * 3, "faces");
        h2d(c, d_verts, verts, (size_t)ne * 6, "verts");
        h2d(c, d_nb, neigh, (size_t)ne * 3, "neigh");
        h2d(c, d_ng, neighedge, (size_t)ne * 3, "neighedge");
        h2d(c, d_mat, lambda, ne, "lambda");
        h2d(c, d_mat + ne, mu, ne, "mu");
        h2d(c, d_mat + 2 * ne, rho, ne, "rho");
        AssembleArgs a{};
        a.ne = ne;
        a.verts = d_verts;
        a.neigh = d_nb;
        a.neighedge = d_ng;
        a.lambda = d_mat;
        a.mu = d_mat + ne;
        a.rho = d_mat + 2 * ne;
        a.cpen = penalty * c->P * c->P;
        a.tab = Tables{c->t_B, c->t_G, c->t_Be, c->t_Ge, c->t_w2, c->t_w1};
        a.geo = d_geo;
        a.face = d_face;
        a.blk = c->blk64;
        a.invm = c->invm;
        kcheck(c, launch_assemble(c->P, a, c->stream), "assemble");
        if (c->d_minv && !c->minv_after())
            kcheck(c, launch_premul_invmass(c->P, ne, c->d_minv, c->blk64, c->stream), "invmass");
        hipcheck(c, hipStreamSynchronize(c->stream), "assemble");
        dfree(c, d_verts); dfree(c, d_nb); dfree(c, d_ng); dfree(c, d_mat); dfree(c, d_geo); dfree(c, d_face);
        // SIPG: symmetric by construction (to the rounding of assembly)
        c->symmetric = true;
        operator_ready(c);
    });
}

int tsp_upload_blocks(tsp_context *c, int ne, const int *neigh, const double *blocks, const double *invm)
{
    return guarded(c, [&] {
        if (!c->P) fail(c, TSP_ERR_STATE, "call tsp_set_tables first (it fixes the degree)");
        if (!neigh || !blocks || !invm) fail(c, TSP_ERR_ARG, "null pointer");
        check_mesh_args(c, ne, neigh, nullptr);
        reset_mesh(c, ne);
        c->h_neigh.assign(neigh, neigh + 3 * ne);
        const size_t n = (size_t)ne * 4 * c->D * c->D;
        double *canon = dalloc<double>(c, n, "canonical blocks");
        h2d(c, canon, blocks, n, "blocks");
        kcheck(c, launch_canonical_to_blocks(c->P, ne, canon, c->blk64, c->stream), "layout");
        int *d_nb = dalloc<int>(c, (size_t)ne * 3, "neigh");
        h2d(c, d_nb, neigh, (size_t)ne * 3, "neigh");
        kcheck(c, launch_zero_missing(c->P, ne, d_nb, c->blk64, c->stream), "zero blocks");
        if (c->d_minv && !c->minv_after())
            kcheck(c, launch_premul_invmass(c->P, ne, c->d_minv, c->blk64, c->stream), "invmass");
        h2d(c, c->invm, invm, ne, "invm");
        hipcheck(c, hipStreamSynchronize(c->stream), "upload");
        dfree(c, canon);
        dfree(c, d_nb);
        // only a single-column FP64 or FP32c context can take the face form
        c->symmetric = (c->precision == TSP_FP64 || c->precision == TSP_FP32C) && !c->batched() &&
                       blocks_symmetric(ne, c->D, neigh, blocks);
        operator_ready(c);
    });
}

int tsp_get_blocks(tsp_context *c, int e_begin, int e_count, double *blocks, double *invm)
{
    return guarded(c, [&] {
        require_fp64_operator(c);
        if (e_begin < 0 || e_count < 0 || e_begin + e_count > c->ne)
            fail(c, TSP_ERR_ARG, "element range [%d, %d) outside 0..%d", e_begin, e_begin + e_count, c->ne);
        if (blocks && e_count) {
            const size_t per = (size_t)4 * c->D * c->D;
            double *canon = dalloc<double>(c, per * e_count, "canonical blocks");
            kcheck(c, launch_blocks_to_canonical(c->P, e_count, c->blk64 + per * e_begin, canon, c->stream), "layout");
            d2h(c, blocks, canon, per * e_count, "blocks");
            dfree(c, canon);
        }
        if (invm && e_count) d2h(c, invm, c->invm + e_begin, e_count, "invm");
    });
}

int tsp_set_state(tsp_context *c, int col, const double *u0, const double *v0)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (col < 0 || col >= c->ncols) fail(c, TSP_ERR_ARG, "column %d out of range (ncols = %d)", col, c->ncols);
        const size_t n = (size_t)c->ne * c->D;
        // u64[0] <- u0, u64[1] <- v0 (element-major, or the padded multi-column layout)
        for (int k = 0; k < 2; ++k) {
            const double *src = k == 0 ? u0 : v0;
            if (src) h2d(c, c->scratch64, src, n, "state");
            if (c->batched())
                kcheck(c, launch_col_in(c->ne, c->nln, c->KP, c->ncols, col, src ? c->scratch64 : nullptr, c->u64[k],
                                        c->stream), "layout");
            else if (src)
                kcheck(c, launch_to_elem_major(c->ne, c->nln, c->scratch64, c->u64[k], c->stream), "layout");
            else
                hipcheck(c, hipMemsetAsync(c->u64[k], 0, n * sizeof(double), c->stream), "memset");
        }
        c->cur = 0;
        c->steps = 0;
        c->first_done = false;
        c->nsamples = 0;
        hipcheck(c, hipMemsetAsync(c->flags, 0, sizeof(unsigned), c->stream), "memset");
        hipcheck(c, hipStreamSynchronize(c->stream), "state");
    });
}

int tsp_set_sources(tsp_context *c, int nsrc, const int *col, const int *elem, const double *phi, int ntab,
                    const double *wavelet)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (nsrc < 0 || (nsrc > 0 && (!col || !elem || !phi))) fail(c, TSP_ERR_ARG, "bad source arrays");
        if (nsrc > 0 && (ntab < 1 || !wavelet)) fail(c, TSP_ERR_ARG, "sources need a wavelet table");
        for (int k = 0; k < nsrc; ++k) {
            if (col[k] < 0 || col[k] >= c->ncols) fail(c, TSP_ERR_ARG, "source %d: column %d out of range", k, col[k]);
            if (elem[k] < 0 || elem[k] >= c->ne) fail(c, TSP_ERR_ARG, "source %d: element %d out of range", k, elem[k]);
        }
        dfree(c, c->src_phi);
        dfree(c, c->wav);
        c->nsrc = nsrc;
        c->ntab = nsrc ? ntab : 0;
        c->h_src_elem.assign(elem, elem + nsrc);
        c->h_src_col.assign(col, col + nsrc);
        c->src_phi = dalloc<double>(c, (size_t)nsrc * c->nln, "source phi");
        c->wav = dalloc<double>(c, (size_t)c->ntab * nsrc * 2, "wavelet");
        std::vector<double> mphi;
        if (nsrc && !c->h_minv.empty() && !c->minv_after()) {
            // M^-1 f for f = F(t) phi: the folded operator needs the source in the same rows
            mphi.assign((size_t)nsrc * c->nln, 0.0);
            for (int k = 0; k < nsrc; ++k)
                for (int i = 0; i < c->nln; ++i) {
                    double acc = 0.0;
                    for (int j = 0; j < c->nln; ++j)
                        acc += c->h_minv[(size_t)i * c->nln + j] * phi[(size_t)k * c->nln + j];
                    mphi[(size_t)k * c->nln + i] = acc;
                }
            phi = mphi.data();
        }
        if (nsrc) {
            h2d(c, c->src_phi, phi, (size_t)nsrc * c->nln, "source phi");
            h2d(c, c->wav, wavelet, (size_t)c->ntab * nsrc * 2, "wavelet");
        }
        rebuild_extra(c);
    });
}

int tsp_set_wavelet(tsp_context *c, int row0, int nrows, const double *wavelet)
{
    return guarded(c, [&] {
        if (!c->nsrc) fail(c, TSP_ERR_STATE, "no sources");
        if (row0 < 0 || nrows < 0 || row0 + nrows > c->ntab || (nrows && !wavelet))
            fail(c, TSP_ERR_ARG, "rows [%d, %d) outside the table of %d rows", row0, row0 + nrows, c->ntab);
        if (nrows) {
            h2d(c, c->wav + (size_t)row0 * c->nsrc * 2, wavelet, (size_t)nrows * c->nsrc * 2, "wavelet");
            hipcheck(c, hipStreamSynchronize(c->stream), "wavelet");
        }
    });
}

int tsp_set_receivers(tsp_context *c, int nrec, const int *elem, const double *phi, int stride, int max_samples)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (nrec < 0 || (nrec > 0 && (!elem || !phi))) fail(c, TSP_ERR_ARG, "bad receiver arrays");
        if (stride < 0 || max_samples < 0) fail(c, TSP_ERR_ARG, "negative stride or sample count");
        for (int k = 0; k < nrec; ++k)
            if (elem[k] < 0 || elem[k] >= c->ne) fail(c, TSP_ERR_ARG, "receiver %d: element %d out of range", k, elem[k]);
        dfree(c, c->rec_phi);
        dfree(c, c->seis);
        dfree(c, c->rec_elem);
        c->nrec = nrec;
        c->stride = stride;
        c->max_samples = max_samples;
        c->nsamples = 0;
        c->h_rec_elem.assign(elem, elem + nrec);
        c->rec_phi = dalloc<double>(c, (size_t)nrec * c->nln, "receiver phi");
        c->rec_elem = dalloc<int>(c, nrec, "receiver elem");
        c->seis = dalloc<double>(c, (size_t)max_samples * nrec * c->ncols * 2, "seismograms");
        if (nrec) {
            h2d(c, c->rec_phi, phi, (size_t)nrec * c->nln, "receiver phi");
            h2d(c, c->rec_elem, elem, nrec, "receiver elem");
        }
        rebuild_extra(c);
    });
}

int tsp_set_stride(tsp_context *c, int stride)
{
    return guarded(c, [&] {
        if (stride < 0) fail(c, TSP_ERR_ARG, "negative stride");
        if (stride != c->stride) drop_graph(c);
        c->stride = stride;
    });
}

int tsp_set_face_form(tsp_context *c, int on, int *active)
{
    return guarded(c, [&] {
        if (on < -1 || on > 2) fail(c, TSP_ERR_ARG, "on must be 1, 2, 0 or -1 (query)");
        if (on >= 0 && on != c->face_form) {
            c->face_form = on;
            if (c->ne) {
                // queued steps may still read the face buffers
                hipcheck(c, hipStreamSynchronize(c->stream), "synchronize");
                build_faces(c);
            }
        }
        if (active) *active = c->zf != nullptr || c->zf32 != nullptr;
    });
}

int tsp_set_dt(tsp_context *c, double dt)
{
    return guarded(c, [&] {
        if (!(dt > 0) || !std::isfinite(dt)) fail(c, TSP_ERR_ARG, "dt must be positive and finite");
        c->dt = dt;
    });
}

int tsp_first_step(tsp_context *c)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (!(c->dt > 0)) fail(c, TSP_ERR_STATE, "call tsp_set_dt first");
        if (c->first_done) fail(c, TSP_ERR_STATE, "first step already taken (reset with tsp_set_state)");
        StepArgs a = base_args(c);
        a.wav_row = wav_row(c, 0);
        // with a sampling stride the first sample is u1 (the driver's eval_receivers
        // after first_step); stride 0 leaves every sample to tsp_eval_receivers
        a.sample = c->stride > 0 ? take_sample_slot(c) : -1;
        const double dt = c->dt;
        if (c->precision == TSP_FP64) {
            a.coef = dt * dt / 2;
            a.dt = dt;
            a.ucur = c->u64[0];   // u0
            a.uprev = c->u64[1];  // v0 in, u1 out
            kcheck(c, launch_step64(c->P, true, a, c->stream), "first step");
            c->cur = 1;           // u^n = u1 in u64[1]; u^{n-1} = u0 in u64[0]
        } else if (c->batched()) {
            // two time levels: u0 in buffers 0, u^{-1} = u0 - dt v0 (FP64, so the
            // step's increment u0 - u^{-1} is dt v0) in buffers 1, which the
            // first step overwrites with u1 (the first-step scale -dt^2/2 invM)
            const size_t n = c->state_len();
            kcheck(c, launch_prev_b(n, c->u64[0], c->u64[1], dt, c->stream), "u-1");
            kcheck(c, launch_split_b(n, c->u64[0], c->uhi[0], c->ulo[0], c->stream), "split");
            kcheck(c, launch_split_b(n, c->u64[1], c->uhi[1], c->ulo[1], c->stream), "split");
            kcheck(c, launch_scale_demote(c->ne, c->invm, -(dt * dt / 2), c->sc_first, c->stream), "sc");
            kcheck(c, launch_scale_demote(c->ne, c->invm, (-dt) * dt, c->sc_step, c->stream), "sc");
            c->cur = 0;
            BatchArgs b = batch_args(c);
            b.wav_row = a.wav_row;
            b.sample = a.sample;
            b.sc32 = c->sc_first;
            kcheck(c, launch_batched(c, b), "first step");
            c->cur = 1;
        } else {
            // two time levels as in the multi-column branch: u0 in buffers 0,
            // u^{-1} = u0 - dt v0 in buffers 1, overwritten with u1
            const size_t n = (size_t)c->ne * c->D;
            kcheck(c, launch_prev_b(n, c->u64[0], c->u64[1], dt, c->stream), "u-1");
            kcheck(c, launch_split_b(n, c->u64[0], c->uhi[0], c->ulo[0], c->stream), "split");
            kcheck(c, launch_split_b(n, c->u64[1], c->uhi[1], c->ulo[1], c->stream), "split");
            kcheck(c, launch_scale_demote(c->ne, c->invm, -(dt * dt / 2), c->sc_first, c->stream), "sc");
            kcheck(c, launch_scale_demote(c->ne, c->invm, (-dt) * dt, c->sc_step, c->stream), "sc");
            a.sc32 = c->sc_first;
            a.uhi = c->uhi[0];
            a.ulo = c->ulo[0];
            a.uhi_out = c->uhi[1];
            a.ulo_out = c->ulo[1];
            kcheck(c, launch_step32c(c->P, true, a, c->stream), "first step");
            c->cur = 1;
        }
        c->first_done = true;
    });
}

int tsp_run(tsp_context *c, int nsteps)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (!c->first_done) fail(c, TSP_ERR_STATE, "call tsp_first_step before tsp_run");
        if (nsteps < 0) fail(c, TSP_ERR_ARG, "negative step count");
        if (c->nsrc && c->steps + nsteps >= c->ntab)
            fail(c, TSP_ERR_ARG, "wavelet table has %d rows; %d more steps need %d", c->ntab, nsteps,
                 c->steps + nsteps + 1);
        if (c->nrec > 0 && c->stride > 0) {
            const int need = (c->steps + nsteps) / c->stride - c->steps / c->stride;
            if (c->nsamples + need > c->max_samples)
                fail(c, TSP_ERR_ARG, "seismogram buffer of %d samples cannot hold %d more", c->max_samples, need);
        }
        if (c->batched() && nsteps >= kGraphSteps && regular_samples(c)) nsteps = run_graph(c, nsteps);
        StepArgs a = base_args(c);
        const double dt = c->dt;
        a.coef = (-dt) * dt;
        a.sc32 = c->sc_step;
        for (int k = 0; k < nsteps; ++k) {
            const int n = c->steps + 1;  // 1-based index of this step() call
            a.wav_row = c->nsrc ? c->wav + (size_t)n * c->nsrc * 2 : nullptr;
            a.sample = (c->stride > 0 && n % c->stride == 0) ? take_sample_slot(c) : -1;
            if (c->precision == TSP_FP64) {
                a.ucur = c->u64[c->cur];
                a.uprev = c->u64[1 - c->cur];
                kcheck(c, launch_step64(c->P, false, a, c->stream), "step");
            } else if (c->batched()) {
                BatchArgs b = batch_args(c);   // u^n in buffers cur, u^{n-1} in 1 - cur
                b.wav_row = a.wav_row;
                b.sample = a.sample;
                b.sc32 = c->sc_step;
                kcheck(c, launch_batched(c, b), "step");
            } else {
                a.uhi = c->uhi[c->cur];
                a.ulo = c->ulo[c->cur];
                a.uhi_out = c->uhi[1 - c->cur];
                a.ulo_out = c->ulo[1 - c->cur];
                kcheck(c, launch_step32c(c->P, false, a, c->stream), "step");
            }
            c->cur = 1 - c->cur;
            c->steps = n;
        }
    });
}

int tsp_run_timed(tsp_context *c, int nsteps, float *ms)
{
    if (!c) return TSP_ERR_ARG;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = guarded(c, [&] {
        if (!ms) fail(c, TSP_ERR_ARG, "null output");
        // graph capture is host work: keep it out of the timed span
        if (c->batched() && c->first_done && nsteps >= kGraphSteps && regular_samples(c)) ensure_graph(c);
        hipcheck(c, hipEventCreate(&a), "hipEventCreate");
        hipcheck(c, hipEventCreate(&b), "hipEventCreate");
        hipcheck(c, hipEventRecord(a, c->stream), "hipEventRecord");
    });
    if (rc == TSP_OK) rc = tsp_run(c, nsteps);
    if (rc == TSP_OK)
        rc = guarded(c, [&] {
            hipcheck(c, hipEventRecord(b, c->stream), "hipEventRecord");
            hipcheck(c, hipEventSynchronize(b), "hipEventSynchronize");
            hipcheck(c, hipEventElapsedTime(ms, a, b), "hipEventElapsedTime");
        });
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return rc;
}

int tsp_eval_receivers(tsp_context *c)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (!c->first_done) fail(c, TSP_ERR_STATE, "no displacement yet: call tsp_first_step");
        if (c->nrec == 0) return;
        StepArgs a = base_args(c);
        a.sample = take_sample_slot(c);
        if (a.sample < 0) fail(c, TSP_ERR_ARG, "seismogram buffer full (%d samples)", c->max_samples);
        if (c->batched()) {
            BatchArgs b = batch_args(c);
            b.sample = a.sample;
            kcheck(c, launch_sample_b32(c->P, c->ncols, b, c->stream), "sample");
        } else if (c->precision == TSP_FP64)
            kcheck(c, launch_sample64(c->P, a, c->u64[c->cur], c->stream), "sample");
        else
            kcheck(c, launch_sample32(c->P, a, c->uhi[c->cur], c->ulo[c->cur], c->stream), "sample");
    });
}

int tsp_steps_done(const tsp_context *c) { return c ? c->steps : -1; }

int tsp_get_u(tsp_context *c, int col, double *out)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (col < 0 || col >= c->ncols) fail(c, TSP_ERR_ARG, "column %d out of range", col);
        if (!out) fail(c, TSP_ERR_ARG, "null output");
        const size_t n = (size_t)c->ne * c->D;
        if (c->batched()) {
            if (!c->first_done) fail(c, TSP_ERR_STATE, "no displacement yet: call tsp_first_step");
            kcheck(c, launch_col_out(c->ne, c->nln, c->KP, c->ncols, col, c->uhi[c->cur], c->ulo[c->cur],
                                     c->scratch64, c->stream), "layout");
            d2h(c, out, c->scratch64, n, "get_u");
            return;
        }
        const double *src;
        if (!c->first_done) {
            src = c->u64[0];
        } else if (c->precision == TSP_FP64) {
            src = c->u64[c->cur];
        } else {
            // u0/v0 staging is free once the first step has run: join hi - comp there
            kcheck(c, launch_join_state((size_t)c->ne * c->D, c->uhi[c->cur], c->ulo[c->cur], c->u64[0], c->stream),
                   "join");
            src = c->u64[0];
        }
        kcheck(c, launch_to_comp_major(c->ne, c->nln, src, c->scratch64, c->stream), "layout");
        d2h(c, out, c->scratch64, n, "get_u");
    });
}

int tsp_get_seismograms(tsp_context *c, double *out, int max_samples, int *nsamples)
{
    return guarded(c, [&] {
        const int ns = std::min(c->nsamples, std::max(0, max_samples));
        if (nsamples) *nsamples = c->nsamples;
        if (out && ns > 0 && c->nrec > 0) d2h(c, out, c->seis, (size_t)ns * c->nrec * c->ncols * 2, "seismograms");
    });
}

int tsp_get_seismograms_cols(tsp_context *c, int col0, int ncols_out, double *out, int max_samples, int *nsamples,
                             int async)
{
    return guarded(c, [&] {
        if (col0 < 0 || ncols_out < 1 || col0 + ncols_out > c->ncols)
            fail(c, TSP_ERR_ARG, "columns [%d, %d) outside 0..%d", col0, col0 + ncols_out, c->ncols);
        const int ns = std::min(c->nsamples, std::max(0, max_samples));
        if (nsamples) *nsamples = c->nsamples;
        if (!out || ns <= 0 || c->nrec <= 0) return;
        const size_t width = sizeof(double) * 2 * ncols_out, pitch = sizeof(double) * 2 * c->ncols;
        hipcheck(c, hipMemcpy2DAsync(out, width, c->seis + (size_t)col0 * 2, pitch, width, (size_t)ns * c->nrec,
                                     hipMemcpyDeviceToHost, c->stream), "seismograms");
        if (!async) hipcheck(c, hipStreamSynchronize(c->stream), "seismograms");
    });
}

int tsp_stream_idle(tsp_context *c, int *idle)
{
    return guarded(c, [&] {
        if (!idle) fail(c, TSP_ERR_ARG, "null output");
        const hipError_t e = hipStreamQuery(c->stream);
        if (e == hipErrorNotReady) {
            *idle = 0;
            return;
        }
        hipcheck(c, e, "stream query");
        *idle = 1;
    });
}

void *tsp_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void tsp_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

// Power iteration for lambda_max(invM B) on caller-provided device vectors
// x, y (ne*D doubles each) and kCflParts partial sums; every buffer is fully
// written before it is read, so none needs clearing.  tsp_cfl and
// tsp_cfl_noalloc run exactly this, so their results are bitwise the same.
constexpr int kCflParts = 256;

static void cfl_core(tsp_context *c, double factor, int iters, double *x, double *y, double *part, double *dt_out,
                     double *lmax_out)
{
    require_fp64_operator(c);
    if (iters < 1) fail(c, TSP_ERR_ARG, "iters must be >= 1");
    const size_t n = (size_t)c->ne * c->D;
    double hp[kCflParts];
    kcheck(c, launch_fill_start(n, x, c->stream), "cfl start");
    auto norm = [&](const double *v) {
        kcheck(c, launch_dot(n, v, v, part, kCflParts, c->stream), "dot");
        d2h(c, hp, part, kCflParts, "dot");
        double s = 0;
        for (double p : hp) s += p;
        return std::sqrt(s);
    };
    double lam = 0, nx = norm(x);
    for (int it = 0; it < iters; ++it) {
        kcheck(c, launch_apply_minvB(c->P, c->ne, c->meta, c->blk64, c->invm, c->minv_after() ? c->d_minv : nullptr,
                                     x, y, c->stream), "apply");
        const double ny = norm(y);
        lam = ny / nx;
        std::swap(x, y);
        nx = ny;
        if (!(nx > 0) || !std::isfinite(nx)) break;
        if (nx > 1e100 || nx < 1e-100) {
            // rescale to keep the iterate in range
            kcheck(c, launch_scale(n, x, 1.0 / nx, c->stream), "cfl scale");
            nx = norm(x);
        }
    }
    if (lmax_out) *lmax_out = lam;
    if (dt_out) *dt_out = lam > 0 ? factor * 2.0 / std::sqrt(lam) : 0.0;
}

int tsp_cfl(tsp_context *c, double factor, int iters, double *dt_out, double *lmax_out)
{
    return guarded(c, [&] {
        require_fp64_operator(c);
        const size_t n = (size_t)c->ne * c->D;
        double *x = dalloc<double>(c, n, "cfl"), *y = nullptr, *part = nullptr;
        try {
            y = dalloc<double>(c, n, "cfl");
            part = dalloc<double>(c, kCflParts, "cfl");
            cfl_core(c, factor, iters, x, y, part, dt_out, lmax_out);
        } catch (...) {
            (void)hipStreamSynchronize(c->stream);
            dfree(c, x); dfree(c, y); dfree(c, part);
            throw;
        }
        dfree(c, x); dfree(c, y); dfree(c, part);
    });
}

// ---------------------------------------------------------------- no-alloc CFL
// tsp_cfl_noalloc runs cfl_core on a file-scope device array instead of
// buffers it allocates.  Each device holds its own copy of `scratch` (from the
// first load of this code object on that device until the process ends); its
// address on a device comes from hipGetSymbolAddress with that device current,
// taken once per device on the entry point's first call there and kept in
// g_scratch_ptr.  Host code never passes `scratch` itself to a kernel: in host
// code the name stands for a host-side shadow, not a device address.
constexpr size_t kNoAllocMaxDof = 2000000;   // ne*2*nln per vector: 16 MB
constexpr size_t kNoAllocDoubles = 2 * kNoAllocMaxDof + kCflParts;
static __device__ __align__(256) double scratch[kNoAllocDoubles];   // 32 MB per device

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static double *scratch_ptr(tsp_context *c)
{
    if (c->device < 0 || c->device >= kMaxDevices) fail(c, TSP_ERR_ARG, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void *p = nullptr;
        hipcheck(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[c->device] = static_cast<double *>(p);
    }
    return g_scratch_ptr[c->device];
}

int tsp_cfl_noalloc(tsp_context *c, double factor, int iters, double *dt_out, double *lmax_out)
{
    return guarded(c, [&] {
        require_fp64_operator(c);
        const size_t n = (size_t)c->ne * c->D;
        if (n > kNoAllocMaxDof)
            fail(c, TSP_ERR_ARG, "tsp_cfl_noalloc: %zu coefficients per vector exceed its scratch of %zu (use tsp_cfl)",
                 n, kNoAllocMaxDof);
        double *base = scratch_ptr(c);   // set_device(c) made c's device current
        cfl_core(c, factor, iters, base, base + kNoAllocMaxDof, base + 2 * kNoAllocMaxDof, dt_out, lmax_out);
    });
}

// y = B x (mode 0) or y = invM x (mode 1) on component-major vectors.
static void apply_plugin(tsp_context *c, int mode, const double *x, double *y, int device_ptrs)
{
    require_mesh(c);
    if (!x || !y) fail(c, TSP_ERR_ARG, "null vector");
    if (mode == 0) {
        require_fp64_operator(c);
        if (c->d_minv && !c->minv_after())
            fail(c, TSP_ERR_STATE,
                 "this context stores the operator premultiplied by the inverse mass (non-orthonormal basis "
                 "in a TSP_FP32C or multi-column context): apply B from a TSP_FP64 context");
    }
    const size_t n = (size_t)c->ne * c->D;
    const double *minv = c->d_minv;   // invM_e = invm_e * base_invmass (identity when null)
    if (device_ptrs) {
        const uintptr_t xa = (uintptr_t)x, ya = (uintptr_t)y, nb = n * sizeof(double);
        if (xa < ya + nb && ya < xa + nb) fail(c, TSP_ERR_ARG, "x and y overlap");
        kcheck(c, launch_apply_cm(c->P, mode, c->ne, c->meta, c->blk64, c->invm, minv, x, y, c->stream), "apply");
        return;
    }
    double *dx = dalloc<double>(c, n, "apply"), *dy = nullptr;
    try {
        dy = dalloc<double>(c, n, "apply");
        h2d(c, dx, x, n, "apply");
        kcheck(c, launch_apply_cm(c->P, mode, c->ne, c->meta, c->blk64, c->invm, minv, dx, dy, c->stream), "apply");
        d2h(c, y, dy, n, "apply");
    } catch (...) {
        (void)hipStreamSynchronize(c->stream);
        dfree(c, dx);
        dfree(c, dy);
        throw;
    }
    dfree(c, dx);
    dfree(c, dy);
}

int tsp_apply_operator(tsp_context *c, const double *x, double *y, int device_ptrs)
{
    return guarded(c, [&] { apply_plugin(c, 0, x, y, device_ptrs); });
}

int tsp_apply_invmass(tsp_context *c, const double *x, double *y, int device_ptrs)
{
    return guarded(c, [&] { apply_plugin(c, 1, x, y, device_ptrs); });
}

int tsp_status(tsp_context *c, unsigned *flags)
{
    return guarded(c, [&] {
        if (!flags) fail(c, TSP_ERR_ARG, "null output");
        d2h(c, flags, c->flags, 1, "flags");
    });
}

int tsp_synchronize(tsp_context *c)
{
    return guarded(c, [&] { hipcheck(c, hipStreamSynchronize(c->stream), "synchronize"); });
}

long long tsp_device_bytes(const tsp_context *c) { return c ? c->bytes : 0; }

int tsp_share_operator(tsp_context *owner, tsp_context **out)
{
    if (!owner || !out) return TSP_ERR_ARG;
    *out = nullptr;
    tsp_context *c = nullptr;
    int rc = tsp_create(owner->device, owner->precision, owner->ncols, &c);
    if (rc != TSP_OK) return rc;
    rc = guarded(c, [&] {
        if (!owner->ne) fail(c, TSP_ERR_STATE, "the owner has no operator");
        c->P = owner->P; c->nln = owner->nln; c->D = owner->D; c->nq1 = owner->nq1; c->nq2 = owner->nq2;
        c->ne = owner->ne;
        c->h_neigh = owner->h_neigh;
        // borrowed (not in c->allocs, never freed by c)
        c->blk64 = owner->blk64;
        c->blk32 = owner->blk32;
        c->blkA = owner->blkA;
        c->invm = owner->invm;
        c->KP = owner->KP;
        c->h_minv = owner->h_minv;   // sources of a borrower get the same M^-1
        c->d_minv = owner->d_minv;   // borrowed, as the operator
        c->meta = dalloc<int4>(c, c->ne, "meta");
        if (c->precision != TSP_FP64) {
            c->sc_first = dalloc<float>(c, c->ne, "sc");
            c->sc_step = dalloc<float>(c, c->ne, "sc");
        }
        alloc_state(c);
        hipcheck(c, hipMemsetAsync(c->u64[0], 0, sizeof(double) * c->state_len(), c->stream), "memset");
        hipcheck(c, hipMemsetAsync(c->u64[1], 0, sizeof(double) * c->state_len(), c->stream), "memset");
        c->have_state = true;
        rebuild_extra(c);
        c->symmetric = owner->symmetric;
        c->face_form = owner->face_form;
        build_faces(c);
        c->lender = owner;
        owner->borrowers.push_back(c);
    });
    if (rc != TSP_OK) {
        if (owner) owner->err = c->err;
        tsp_destroy(c);
        return rc;
    }
    *out = c;
    return TSP_OK;
}

int tsp_release_fp64(tsp_context *c)
{
    return guarded(c, [&] {
        require_mesh(c);
        if (c->precision == TSP_FP64) fail(c, TSP_ERR_STATE, "the FP64 context steps with its FP64 operator");
        hipcheck(c, hipStreamSynchronize(c->stream), "synchronize");
        // borrowers share this FP64 copy: they lose it too (their steps use the FP32 one)
        for (tsp_context *b : c->borrowers) {
            if (b->blk64 == c->blk64) {
                hipcheck(c, hipStreamSynchronize(b->stream), "synchronize");
                b->blk64 = nullptr;
            }
        }
        dfree(c, c->blk64);
    });
}

}  // extern "C"

namespace tspeed_dev {

bool ctx_view(tsp_context *c, InstView *v)
{
    if (!c || !v) return false;
    v->device = c->device;
    v->precision = c->precision;
    v->ncols = c->ncols;
    v->P = c->P;
    v->ne = c->ne;
    v->nsrc = c->nsrc;
    v->nrec = c->nrec;
    v->ntab = c->ntab;
    v->stride = c->stride;
    v->max_samples = c->max_samples;
    v->nsamples = c->nsamples;
    v->steps = c->steps;
    v->cur = c->cur;
    v->first_done = c->first_done;
    const float *op = c->batched() ? c->blkA : c->blk32;
    v->owns_operator = op && std::find(c->allocs.begin(), c->allocs.end(), (const void *)op) != c->allocs.end();
    v->tfree = c->h_minv.empty();
    v->stream = c->stream;
    v->meta = c->meta;
    v->extra = c->extra;
    v->src_sorted = c->src_sorted;
    v->rec_sorted = c->rec_sorted;
    v->src_phi = c->src_phi;
    v->wav = c->wav;
    v->rec_phi = c->rec_phi;
    v->seis = c->seis;
    v->flags = c->flags;
    v->blk32 = c->blk32;
    v->blkA = c->blkA;
    v->sc_step = c->sc_step;
    v->uhi[0] = c->uhi[0];
    v->uhi[1] = c->uhi[1];
    v->ulo[0] = c->ulo[0];
    v->ulo[1] = c->ulo[1];
    return true;
}

bool ctx_batch_args(tsp_context *c, BatchArgs *out)
{
    if (!c || !out || !c->batched()) return false;
    BatchArgs b = batch_args(c);
    b.sc32 = c->sc_step;
    b.ctr = nullptr;
    b.koff = 0;
    b.wav = c->wav;
    b.stride = c->nrec ? c->stride : 0;
    b.max_samples = c->max_samples;
    b.uhi_pair[0] = c->uhi[0];
    b.uhi_pair[1] = c->uhi[1];
    b.ulo_pair[0] = c->ulo[0];
    b.ulo_pair[1] = c->ulo[1];
    *out = b;
    return true;
}

void ctx_set_progress(tsp_context *c, int steps, int cur, int nsamples)
{
    c->steps = steps;
    c->cur = cur;
    c->nsamples = nsamples;
}

}  // namespace tspeed_dev

// internal.hpp -- library-internal view of a context, used by the
// multi-instance engine (multi.hip) to step many contexts in one kernel.
#pragma once
#include "engine.hpp"

struct tsp_context;

namespace tspeed_dev {

struct InstView {
    int device, precision, ncols;
    int P, ne, nsrc, nrec, ntab, stride, max_samples, nsamples, steps, cur;
    bool first_done;
    bool owns_operator;   // false for a context made by tsp_share_operator
    bool tfree;           // translation-free products (orthonormal basis)
    hipStream_t stream;
    const int4 *meta, *extra;
    const int *src_sorted, *rec_sorted;
    const double *src_phi, *wav, *rec_phi;
    double *seis;
    unsigned *flags;
    const float *blk32, *blkA, *sc_step;   // single-column / multi-column operator
    float *uhi[2], *ulo[2];   // FP32 state: two time levels of u = hi - lo (lo BF16)
};

bool ctx_view(tsp_context *c, InstView *v);
// a multi-column context's step arguments for replay from a step counter
// (BatchArgs::wav, stride, max_samples, uhi_pair set; ctr null)
bool ctx_batch_args(tsp_context *c, BatchArgs *out);
// record progress made by kernels launched outside the context
void ctx_set_progress(tsp_context *c, int steps, int cur, int nsamples);

}  // namespace tspeed_dev

/*
 * tspeed_amd.h -- C-ABI of the MI355X (gfx950) leap-frog engine.
 *
 * This is the drop-in boundary under tspeed's `LeapFrog` class.  Each entry point
 * names the reference interface it replaces (paths relative to the tspeed source
 * tree).  Conventions:
 *   - every function returns an int status (TSP_OK = 0) and never throws;
 *     tsp_last_error() gives the message of the last failure on a context;
 *   - host buffers belong to the caller, device buffers to the library;
 *   - coefficient vectors cross the ABI in the reference's component-major
 *     layout idx = c*ne*nln + e*nln + i (TimeAdvance.cpp:73-74), one vector per
 *     right-hand-side column; the device keeps its own element-major layout;
 *   - basis tables follow FESpace's layout (FESpace_imp.hpp:26-47), see
 *     tsp_set_tables;
 *   - a context is not thread-safe: one context per host thread, or lock it.
 */
#ifndef TSPEED_AMD_H
#define TSPEED_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tsp_context tsp_context;

enum tsp_status_code {
    TSP_OK = 0,
    TSP_ERR_ARG = 1,         /* bad argument or size                          */
    TSP_ERR_STATE = 2,       /* call out of order (e.g. step before mesh)     */
    TSP_ERR_HIP = 3,         /* HIP runtime error                             */
    TSP_ERR_UNSUPPORTED = 4, /* degree / precision / column count not built   */
    TSP_ERR_NOMEM = 5        /* device allocation failed                      */
};

enum tsp_precision {
    TSP_FP64 = 0,   /* plain FP64 blocks and update: the correctness path            */
    TSP_FP32C = 1,  /* FP32 blocks and products, Kahan-compensated displacement kept
                       at two time levels as FP32 hi + BF16 compensation (the
                       increment u^n - u^{n-1} recomputed, not stored)         */
    TSP_BF16X3 = 2, /* as TSP_FP32C, with each block product split into three BF16
                       products accumulated in FP32 (multi-column contexts only):
                       operator and state each kept to 16 bits (two BF16 parts),
                       a_h x_h + a_h x_l + a_m x_h; the operator's 16 bits set the
                       error: seismograms drift to ~4-9e-4 of FP64 over 3000 steps */
    TSP_BF16X5 = 3  /* five BF16 products on the same MFMA path: the FP32 operator
                       in three BF16 parts (all 24 bits), the state in two,
                       a_h x_h + a_h x_l + a_m x_h + a_m x_l + a_l x_h; the BF16
                       MFMA truncates its internal sums (24 bits below the largest
                       term), so seismograms end 2-6e-5 from FP64 over 3000 steps
                       (TSP_FP32C's FP32 products: 4-7e-6), ~6% faster          */
};

/* Status bits reported by tsp_status (no counterpart in the reference, which has
 * no failure detection: a blown-up run just prints Inf/NaN). */
#define TSP_FLAG_NONFINITE 1u

/* Number of visible HIP devices (0 when no GPU driver is present). */
int tsp_device_count(int *n);

/* Free and total device memory in bytes. */
int tsp_device_memory(int device, long long *free_bytes, long long *total_bytes);

/* Context lifetime.  Replaces the construction of TimeAdvance/LeapFrog
 * (TimeAdvance_imp.hpp:26-40).  `device` is the HIP ordinal; `precision` one
 * of tsp_precision; `ncols` the number of right-hand-side columns: 1 = the
 * reference's single simulation; 8, 16, 32 or 64 = that many simulations on
 * one mesh (one point source each), stepped together as a block SpMM on the
 * MFMA cores (TSP_FP32C, TSP_BF16X3 or TSP_BF16X5).  The MFMA tile is 32
 * columns wide, so 8 and 16 cost the operator stream and products of 32 but
 * carry (and move)
 * only their own columns' state.  Every column is summed on its own and every
 * width of a degree runs the same kernel, so a source's results are bitwise
 * the same at every width. */
int  tsp_create(int device, int precision, int ncols, tsp_context **out);
void tsp_destroy(tsp_context *ctx);
const char *tsp_last_error(const tsp_context *ctx);

/* Optional: launch on a caller-owned hipStream_t (default: a stream the
 * context creates).  Pass NULL to return to the context's own stream. */
int tsp_set_stream(tsp_context *ctx, void *hip_stream);

/* Basis/quadrature tables of FESpace<N, Gauss<N+1>, Dubiner<N>>
 * (FESpace_imp.hpp:26-47; FESpace.hpp:89,98,108), nln = (N+1)(N+2)/2,
 * nq1 = N+1, nq2 = nq1^2:
 *   B[k*nln+i]                  phi_i at area node k          (M_B)
 *   G[(i*2+a)*nq2+k]            d phi_i / d xi_a at node k    (M_G)
 *   Bedge[(f*nq1+k)*nln+i]      phi_i at node k of edge f     (M_Bedge)
 *   Gedge[((i*3+f)*2+a)*nq1+k]  gradient on edge f            (M_Gedge)
 *   w2[nq2], w1[nq1]            area / edge weights           (iweight, eweight)
 * Degrees N = 1..8 are built. */
int tsp_set_tables(tsp_context *ctx, int N, const double *B, const double *G, const double *Bedge,
                   const double *Gedge, const double *w2, const double *w1);

/* Non-orthonormal bases (FESpace<N,Q,BoundaryAdapted<N>>, Matrices_imp.hpp:183-190:
 * invM = 1/rho * 1/detJ * base_invmass).  base_invmass is nln x nln row-major;
 * NULL restores the orthonormal case (identity, the default).  Call after
 * tsp_set_tables and before tsp_set_mesh / tsp_upload_blocks.  A TSP_FP64
 * context keeps B and applies invM after the block product in the reference's
 * order, M^-1 (B u - f); TSP_FP32C and multi-column contexts store the operator
 * rows premultiplied, (I2 (x) base_invmass) B, keep the per-element scalar
 * 1/(rho detJ) and apply base_invmass to every source's basis values in
 * tsp_set_sources.  tsp_get_blocks returns the blocks as stored. */
int tsp_set_mass(tsp_context *ctx, const double *base_invmass);

/* Mesh and materials; assembles B = penalty*N^2*S + I + A and the inverse mass
 * on the device in FP64 (Matrices::Matrices, Matrices_imp.hpp:192-215;
 * TimeAdvance_imp.hpp:32).  Unlike the reference, whose set_penalty is a no-op
 * for B (TimeAdvance.hpp:165 vs TimeAdvance_imp.hpp:27,32), the penalty is an
 * argument here; penalty = 2 reproduces the reference's B.
 *   verts[e*6 + 2*v + {0,1}]   vertex v of triangle e (counter-clockwise)
 *   neigh[e*3+f], neighedge[e*3+f]  neighbour across edge f and its edge index,
 *                              -1 on the boundary (Mesh.cpp:33-128; edge f is
 *                              opposite vertex f)
 *   lambda, mu, rho[ne]        per-element Lame parameters and density
 *                              (Parameters, FESpace.hpp:194-247) */
int tsp_set_mesh(tsp_context *ctx, int ne, const double *verts, const int *neigh, const int *neighedge,
                 const double *lambda, const double *mu, const double *rho, double penalty);

/* Alternative operator input: already assembled blocks in the canonical form
 * blocks[((e*4+s)*d + I)*d + J] (s = 0 own element, s = 1+f neighbour across
 * edge f; I = m*nln+i, J = n*nln+j; d = 2*nln) and the per-element inverse-mass
 * scalar invm[e] (Matrices_imp.hpp:187).  Replaces filling MyMatMultiDim by
 * hand (MyMat.hpp:193-237). */
int tsp_upload_blocks(tsp_context *ctx, int ne, const int *neigh, const double *blocks, const double *invm);
/* Download the FP64 operator rows of elements [e_begin, e_begin+e_count) in the
 * same canonical form (either pointer may be NULL). */
int tsp_get_blocks(tsp_context *ctx, int e_begin, int e_count, double *blocks, double *invm);

/* Initial state of column `col`: u0 (set_initial_u, TimeAdvance_imp.hpp:46-50)
 * and v0 (set_initial_v, :41-45), component-major, 2*ne*nln each; NULL = zero. */
int tsp_set_state(tsp_context *ctx, int col, const double *u0, const double *v0);

/* Point sources (PointWiseForce, Force.hpp:56-86; Force_imp.hpp:24-31):
 *   col[s], elem[s]            column and element of source s
 *   phi[s*nln+i]               basis values at the source point
 *   wavelet[(row*nsrc+s)*2+c]  force component c of source s for time row `row`:
 *                              row 0 is F(0) used by first_step, row n >= 1 the
 *                              value step n uses, F(t_{n+1}) with t accumulated by
 *                              repeated addition (TimeAdvance.cpp:28,48-51).
 * ntab rows are stored; tsp_set_wavelet rewrites rows later. */
int tsp_set_sources(tsp_context *ctx, int nsrc, const int *col, const int *elem, const double *phi,
                    int ntab, const double *wavelet);
int tsp_set_wavelet(tsp_context *ctx, int row0, int nrows, const double *wavelet);

/* Receivers (Receivers.hpp:77-121; PointWiseEntity::M_add, Receivers_imp.hpp:35-51):
 * elem[r], phi[r*nln+i]; every receiver records every column.  Samples are
 * recorded on the device after first_step and after every step whose 1-based
 * index is a multiple of `stride` (0 = only on tsp_eval_receivers), up to
 * max_samples samples. */
int tsp_set_receivers(tsp_context *ctx, int nrec, const int *elem, const double *phi, int stride,
                      int max_samples);

/* Change the receiver sampling stride without touching recorded samples
 * (the driver loop's `if(step%2 == 0) eval_receivers()`, Examples/src/Lamb.cpp:56-57,
 * moved into the step kernels).  0 = sample only on tsp_eval_receivers. */
int tsp_set_stride(tsp_context *ctx, int stride);

/* Engine option with no counterpart in the reference (whose B*u reads every
 * stored block, MyMat.cpp:200-218): the face form of the TSP_FP64 step.  B is
 * symmetric, so the two off-diagonal blocks of an interior face carry the same
 * entries (transposed); the face form reads only the lower-numbered element's
 * copy and applies it to both sides, streaming ne + #faces blocks per step
 * instead of ne + 2 #faces.  It is possible when the operator comes from
 * tsp_set_mesh, or from tsp_upload_blocks with B_{b,a} = B_{a,b}^T bit for bit,
 * and every neighbour relation is mutual; it changes the result by the
 * rounding asymmetry of the assembled blocks (~2e-16 of the largest entry).
 * on = 1 (the default): use it where possible for operators above 128 MB (a
 * smaller one stays in the Infinity Cache, where one launch reading every
 * block is faster); 2: wherever possible; 0: never (every block read as
 * stored); -1: leave the setting.  *active (may be NULL) receives whether the
 * steps use it.  A TSP_FP32C context (single column, N >= 5, orthonormal
 * basis) has the same form with translation-free FP32 products on both sides,
 * but only on request (its default is 0): by default it steps bitwise as it
 * does inside a multi-instance set (tsp_multi_add). */
int tsp_set_face_form(tsp_context *ctx, int on, int *active);

/* Step size (TimeAdvance::set_dt, TimeAdvance.hpp:153). */
int tsp_set_dt(tsp_context *ctx, double dt);

/* LeapFrog::first_step (TimeAdvance.cpp:24-39):
 *   u1 = u0 + dt*v0 - invM*(dt^2/2*(B*u0 - f(0))). */
int tsp_first_step(tsp_context *ctx);

/* nsteps calls of LeapFrog::step (TimeAdvance.cpp:41-59), fused with the point
 * source and the receiver sampling:
 *   u^{n+1} = 2u^n - u^{n-1} + invM*(-dt^2*(B*u^n - f(t_{n+1}))). */
int tsp_run(tsp_context *ctx, int nsteps);

/* tsp_run bracketed by HIP events recorded on the context's stream; *ms gets
 * the elapsed device time of the nsteps launches (measurement helper, no
 * counterpart in the reference). */
int tsp_run_timed(tsp_context *ctx, int nsteps, float *ms);

/* TimeAdvance::eval_receivers (TimeAdvance.cpp:68-79): record one sample of the
 * current displacement now. */
int tsp_eval_receivers(tsp_context *ctx);

/* Number of steps taken so far (M_last_step, TimeAdvance.hpp:233). */
int tsp_steps_done(const tsp_context *ctx);

/* Current displacement of column `col` (get_uh, TimeAdvance.hpp:202),
 * component-major 2*ne*nln doubles. */
int tsp_get_u(tsp_context *ctx, int col, double *out);

/* Seismograms recorded so far: out[((sample*nrec + r)*ncols + col)*2 + c], up to
 * max_samples samples (Receivers::M_val, Receivers.hpp:115).  *nsamples gets
 * the count. */
int tsp_get_seismograms(tsp_context *ctx, double *out, int max_samples, int *nsamples);

/* Columns col0 .. col0+ncols_out-1 of the seismograms, compacted:
 * out[((sample*nrec + r)*ncols_out + k)*2 + c] for column col0 + k.  With
 * async != 0 the copy is only queued on the context's stream: `out` must stay
 * valid (page-locked memory from tsp_host_alloc for a truly asynchronous copy)
 * until tsp_stream_idle reports the stream idle. */
int tsp_get_seismograms_cols(tsp_context *ctx, int col0, int ncols_out, double *out, int max_samples,
                             int *nsamples, int async);
/* *idle = 1 when every operation queued on the context's stream has finished. */
int tsp_stream_idle(tsp_context *ctx, int *idle);
/* Page-locked host memory (hipHostMalloc) for asynchronous downloads. */
void *tsp_host_alloc(size_t bytes);
void tsp_host_free(void *p);

/* Step-size estimate and stability check (no counterpart in the reference,
 * which hard-codes dt, Examples/src/Lamb.cpp:18): lambda_max of invM*B by FP64
 * power iteration on the device (`iters` iterations), and
 * dt = factor * 2/sqrt(lambda_max), the leap-frog limit scaled by `factor`. */
int tsp_cfl(tsp_context *ctx, double factor, int iters, double *dt_out, double *lambda_max_out);

/* The operator plugin below LeapFrog, on one vector in the reference's
 * component-major layout (2*ne*nln doubles), for callers that apply the
 * operators themselves (an energy check, another time integrator):
 *   tsp_apply_operator:  y = B x     (MyMatMultiDim<MyMat>::operator*, MyMat.hpp:330-335,
 *                                     the FP64 B = penalty*N^2*S + I + A)
 *   tsp_apply_invmass:   y = invM x  (MyMatMultiDimBlockDiag<MyMat>::operator*, MyMat.hpp:337-342,
 *                                     invM_e = 1/rho_e * 1/detJ_e * base_invmass)
 * device_ptrs = 0: x and y are host arrays and the call returns with y filled;
 * device_ptrs != 0: x and y are device arrays on the context's device, the work
 * is only queued on the context's stream, and the call allocates nothing.  x and
 * y must not overlap.  tsp_apply_operator needs the FP64 operator
 * (TSP_ERR_STATE after tsp_release_fp64, and in a TSP_FP32C or multi-column
 * context of a non-orthonormal basis, whose operator is stored premultiplied by
 * base_invmass). */
int tsp_apply_operator(tsp_context *ctx, const double *x, double *y, int device_ptrs);
int tsp_apply_invmass(tsp_context *ctx, const double *x, double *y, int device_ptrs);

/* tsp_cfl without any device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory, such as code under HIP graph
 * capture or a serving loop that must not call hipMalloc.  The power-iteration
 * vectors and partial sums live in a fixed 32 MB array per device, built into
 * the library, so meshes with more than 2,000,000 coefficients per vector
 * (ne*2*nln) are refused with TSP_ERR_ARG.  The results are bitwise those of
 * tsp_cfl.  Every call on one device shares that array: calls that overlap on
 * one device (from several host threads) race and must be serialised by the
 * caller. */
int tsp_cfl_noalloc(tsp_context *ctx, double factor, int iters, double *dt_out, double *lambda_max_out);

/* OR of TSP_FLAG_* seen on the current displacement. */
int tsp_status(tsp_context *ctx, unsigned *flags);

/* Wait for all queued work of the context. */
int tsp_synchronize(tsp_context *ctx);

/* Device bytes held by the context (operator + state + tables). */
long long tsp_device_bytes(const tsp_context *ctx);

/* A new context on the same mesh that borrows owner's assembled operator (no
 * second assembly, no second copy) and has its own sources, receivers and
 * state: several simulations on one Mesh/Parameters, as a driver would build
 * several LeapFrog objects from one FESpace (TimeAdvance_imp.hpp:30).  The
 * owner keeps a list of its borrowers: while it has any, tsp_set_mesh and
 * tsp_upload_blocks on it return TSP_ERR_STATE; tsp_release_fp64 on it drops
 * the FP64 copy for the borrowers too; destroying it first leaves each borrower
 * without an operator (its calls then return TSP_ERR_STATE).  A borrower that
 * sets its own mesh stops borrowing.  A borrower handed to a tsp_multi set needs
 * its owner in the same set (which removes owners last) or alive until the set
 * is destroyed. */
int tsp_share_operator(tsp_context *owner, tsp_context **out);

/* Free the FP64 operator of a TSP_FP32C context once its FP32 copy exists
 * (tsp_get_blocks and tsp_cfl then report TSP_ERR_STATE). */
int tsp_release_fp64(tsp_context *ctx);

/* ------------------------------------------------------------------------
 * Sets of independent simulations advanced together (no counterpart in the
 * reference, which runs one LeapFrog per process).  Instances are ordinary
 * TSP_FP32C contexts of any width (1, 8, 16, 32 or 64 columns), built with the
 * calls above (mesh or tsp_share_operator, sources, receivers with a stride,
 * dt, state) and stopped after tsp_first_step; the set takes ownership of
 * each.  A step of the set covers every element of every running instance,
 * whatever its size, degree and width: the single-column instances as one
 * launch per degree present (side by side in the graph) or as one launch of a
 * single kernel that dispatches each element to its degree's step
 * (tsp_multi_set_launch), the multi-column ones as one launch per (degree,
 * width) present, each (element, 32-column tile) running its context's own
 * block-SpMM step.  Chunks of steps replay a captured HIP graph.  Every
 * instance's result is bitwise that of its context stepped alone in its
 * default form (a single-column FP32c context asked for the face form,
 * tsp_set_face_form, steps differently alone; in a set it reads every stored
 * block).  Results are
 * read through tsp_multi_context + tsp_get_u / tsp_get_seismograms.
 * ------------------------------------------------------------------------ */
typedef struct tsp_multi tsp_multi;

int  tsp_multi_create(int device, tsp_multi **out);
void tsp_multi_destroy(tsp_multi *m);
const char *tsp_multi_last_error(const tsp_multi *m);
/* Add a context that will run `nsteps` step() calls; *id gets its index (the
 * index of a removed instance may be given again). */
int  tsp_multi_add(tsp_multi *m, tsp_context *ctx, int nsteps, int *id);
/* Steps per captured graph (default 32; 0 = launch every step directly). */
int  tsp_multi_set_chunk(tsp_multi *m, int chunk);
/* Advance every running instance by up to nsteps steps; *active gets the number
 * of instances still running. */
int  tsp_multi_run(tsp_multi *m, int nsteps, int *active);
/* tsp_multi_run timed with HIP events on the set's stream (milliseconds). */
int  tsp_multi_run_timed(tsp_multi *m, int nsteps, float *ms);
/* How a step is launched:
 *   TSP_MULTI_PER_DEGREE (default): one kernel per degree present, each with the
 *     register budget of its own degree, side by side in the captured graph;
 *   TSP_MULTI_FUSED: one kernel for all degrees, one element per wave, largest
 *     degree first (one launch per step, at the occupancy of the largest
 *     degree's registers).
 * Results are bitwise the same in every mode. */
enum { TSP_MULTI_PER_DEGREE = 0, TSP_MULTI_FUSED = 1 };
int  tsp_multi_set_launch(tsp_multi *m, int mode);
int  tsp_multi_launch(const tsp_multi *m);
/* Choose, per degree, the fastest launch variant of the single-column element
 * step by timing each on `steps_per_trial` real steps of that degree's
 * instances; then, with a graph chunk set, the fastest launch mode on one
 * chunk of the whole set. */
int  tsp_multi_autotune(tsp_multi *m, int steps_per_trial);
/* The launch variant in use for a degree: bit 0 = one row per lane across the
 * wave (else column groups of 16/32 lanes, d <= 32 only), bit 1 = one item per
 * wave or work-group (else a persistent loop), bit 2 = one element per
 * work-group with one slot per wave (bit 0 then unused), bit 3 = one element
 * per work-group of eight waves with one half slot per wave (d > 32, p >= 5
 * only; bits 0 and 2 then unused).  The default (2) keeps every instance
 * bitwise equal to its single-context FP32C run, and so do bit 3 and, at
 * p = 5 and 6, bit 2; the others agree with it to FP32 rounding. */
int  tsp_multi_variant(const tsp_multi *m, int degree);
/* Pin the launch variant of a degree (as tsp_multi_autotune would). */
int  tsp_multi_set_variant(tsp_multi *m, int degree, int variant);
int  tsp_multi_info(tsp_multi *m, int id, int *steps_done, int *nsteps);
/* The instance's context (owned by the set; do not destroy). */
tsp_context *tsp_multi_context(tsp_multi *m, int id);
/* Destroy an instance.  TSP_ERR_STATE if it owns an operator that another
 * instance in the set still borrows (remove the borrowers first). */
int  tsp_multi_remove(tsp_multi *m, int id);
int  tsp_multi_synchronize(tsp_multi *m);

/* ------------------------------------------------------------------------
 * Request loop (no counterpart in the reference).  Simulations arrive over
 * time; a request whose arrival time has passed is admitted as soon as its
 * device-memory footprint fits mem_fraction of the device, then runs in the
 * shared set together with every other admitted request (admission and
 * retirement between chunks of `chunk` steps; fields stay resident meanwhile);
 * on completion its seismograms are copied out and its memory released.  A
 * request is one mesh with 1..16 point sources (one operator assembled, shared
 * by one simulation per source).  Times are seconds of tsp_server_now().
 * ------------------------------------------------------------------------ */
typedef struct tsp_server tsp_server;

typedef struct {
    int N, ne;                         /* degree, triangles                         */
    const double *verts;               /* ne*6, as tsp_set_mesh                     */
    const int *neigh, *neighedge;      /* ne*3                                      */
    const double *lambda, *mu, *rho;   /* ne                                        */
    double penalty, dt;
    int nsteps;                        /* step() calls after the first step         */
    int nsrc;                          /* 1..16                                     */
    const int *src_elem;               /* nsrc                                      */
    const double *src_phi;             /* nsrc*nln                                  */
    const double *wavelet;             /* (nsteps+1)*nsrc*2: [row][source][x|y]     */
    int nrec;
    const int *rec_elem;               /* nrec                                      */
    const double *rec_phi;             /* nrec*nln                                  */
    int stride;                        /* receiver sampling stride (>= 1)           */
    int basis;                         /* 0 Dubiner, 1 BoundaryAdapted (src_phi and
                                          rec_phi in that basis); 0 when zeroed     */
} tsp_request;

int    tsp_server_create(int device, double mem_fraction, int chunk, tsp_server **out);
void   tsp_server_destroy(tsp_server *s);
/* How requests run (before the first submit):
 *   TSP_SERVE_COLUMNS (default): one 8/16/32/64-column context per request, the
 *     narrowest that holds its sources (sources as
 *     columns of a block SpMM on MFMA), each request's chunk on its own stream;
 *   TSP_SERVE_BATCHED: the same multi-column context per request, every
 *     running request in one tsp_multi set (one launch per degree and width
 *     present per step, chunks of steps from a captured graph);
 *   TSP_SERVE_INSTANCES: one single-column context per source (sharing the
 *     request's operator) in one tsp_multi set. */
enum { TSP_SERVE_INSTANCES = 0, TSP_SERVE_COLUMNS = 1, TSP_SERVE_BATCHED = 2 };
int    tsp_server_set_mode(tsp_server *s, int mode);
const char *tsp_server_last_error(const tsp_server *s);
double tsp_server_now(void);
/* Device bytes the loop books for a request while it runs (resident) and the
 * extra it needs while its operator is assembled (peak).  A request is
 * admitted when used + resident + peak <= budget. */
int    tsp_server_footprint(const tsp_request *req, int mode, long long *resident, long long *peak);
/* Queue a request (its arrays are copied) that arrives at time `arrival`.
 * TSP_ERR_NOMEM if resident + peak exceeds the whole budget. */
int    tsp_server_submit(tsp_server *s, const tsp_request *req, double arrival, int *id);
/* One pass of the loop: admit, run, retire.  Instances mode runs one chunk
 * of the set and waits for it; columns mode does not wait: it queues the next
 * two chunks on every request whose stream has run dry (and the result
 * download behind a request's last chunk), so call it in a loop.  Batched
 * mode runs one chunk of the set like instances mode.  Reports how many
 * requests finished in this pass (completed, or failed at admission: then
 * tsp_server_result returns TSP_ERR_STATE with the reason in
 * tsp_server_last_error) and how many are running / queued. */
int    tsp_server_tick(tsp_server *s, int *completed, int *running, int *queued);
/* Result of a completed request: seis[((src*nsamples + sample)*nrec + r)*2 + c]
 * and its arrival / admission / completion times. */
int    tsp_server_result(tsp_server *s, int id, double *seis, int *nsamples, double *arrival, double *admitted,
                         double *completed);
/* Span from the first admission to the last completion, DOF*steps completed,
 * device bytes held by running requests, and the budget. */
int    tsp_server_stats(const tsp_server *s, double *busy_seconds, long long *dof_steps, long long *mem_used,
                        long long *mem_budget);
/* Host seconds tsp_server_tick has spent admitting (context set-up, assembly,
 * first step), running chunks (instances mode: queueing and waiting for the
 * device; columns mode: polling the requests' streams and queueing the next
 * chunks) and retiring (seismogram download, release). */
int    tsp_server_phases(const tsp_server *s, double *admit_seconds, double *run_seconds, double *retire_seconds);

#ifdef __cplusplus
}
#endif
#endif

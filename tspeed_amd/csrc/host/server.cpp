// server.cpp -- a request loop over the multi-instance engine: simulations
// arrive over time, are admitted while their device-memory footprint fits a
// budget, run together with everything already on the device (continuous
// batching: admissions and retirements happen between chunks of steps, while
// the others keep their fields resident), and leave with their seismograms.
//
// A request is one mesh with 1..16 point sources.  Three ways to run it:
//   columns (default): one multi-column context of the narrowest width (8, 16,
//     32 or 64) that holds the sources, one source per column, stepped as a
//     block SpMM on the MFMA cores; every running request's chunk of steps is
//     queued on its own stream, so requests overlap on the device;
//   batched: the same multi-column context, added to one tsp_multi set: a step
//     of everything running is one launch per (degree, width) present;
//   instances: one owner context that assembles the operator and (nsrc-1)
//     contexts sharing it, all added to one tsp_multi set (one launch per degree
//     per step for everything running).
// Either way a request's arithmetic does not depend on what else runs, so its
// result is the same alone and in a mix.
//
// Uses only the public C-ABI (include/tspeed_amd.h, include/tspeed_host.h).
#include "../../../include/tspeed_amd.h"
#include "../../../include/tspeed_host.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace {

double clock_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Tables {
    std::vector<double> B, G, Be, Ge, w2, xy2, w1, x1;
    std::vector<double> minv;  // base inverse mass of a non-orthonormal basis (else empty)
};

struct Request {
    int id = 0;
    double arrival = 0, admitted = -1, completed = -1;
    int N = 0, ne = 0, nsteps = 0, nsrc = 0, nrec = 0, stride = 1, basis = 0;
    double penalty = 0, dt = 0;
    std::vector<double> verts, lambda, mu, rho, src_phi, wavelet, rec_phi;
    std::vector<int> neigh, neighedge, src_elem, rec_elem;
    long long footprint = 0;
    std::vector<int> inst;        // multi-set ids, owner first (instances and batched modes)
    tsp_context *ctx = nullptr;   // the multi-column context (columns mode; batched until added)
    int ncols = 0, done = 0;
    std::vector<double> seis;     // [src][sample][rec][2] (instances mode)
    double *pinned = nullptr;     // [sample][rec][src][2] (columns and batched modes, page-locked)
    int nsamples = 0;
    int state = 0;                // 0 queued, 1 running, 2 done, 3 failed, 4 results in flight
    std::string error;
};

// Columns of the context a request runs in: the narrowest width that holds
// its sources (a column's state is streamed every step whether it carries a
// source or not).
int columns_for(int nsrc) { return nsrc <= 8 ? 8 : nsrc <= 16 ? 16 : nsrc <= 32 ? 32 : 64; }

long long footprint_columns(const Request &r)
{
    const long long nln = (long long)(r.N + 1) * (r.N + 2) / 2, d = 2 * nln;
    const long long mp = 32 * ((d + 31) / 32), kp = 16 * ((d + 15) / 16), R = columns_for(r.nsrc);
    const long long ns = 1 + r.nsteps / std::max(r.stride, 1);
    return (long long)r.ne * 4 * mp * kp * 4          // MFMA operator (rows to 32, columns to 16)
           + (long long)r.ne * kp * R * (2 * 4 + 2 * 2 + 2 * 8)  // hi x2, BF16 lo x2, FP64 staging x2
           + ns * r.nrec * R * 16                     // seismograms of every column
           + (long long)(r.nsteps + 1) * r.nsrc * 16  // force table
           + (long long)r.ne * (d * 8 + 64);          // staging vector, per-element data
}

long long footprint_of(const Request &r, int mode)
{
    if (mode != TSP_SERVE_INSTANCES) return footprint_columns(r);
    const long long nln = (long long)(r.N + 1) * (r.N + 2) / 2, d = 2 * nln;
    const long long op32 = (long long)r.ne * 4 * d * d * 4;       // FP32 column-quad operator
    const long long per_col = (long long)r.ne * d * (2 * 4 + 2 * 2)  // hi x2, BF16 lo x2
                              + (long long)r.ne * d * 8 * 3        // staging + layout scratch
                              + (long long)(r.nsteps + 1) * 16     // force rows
                              + (long long)(1 + r.nsteps / std::max(r.stride, 1)) * r.nrec * 16;
    return op32 + r.nsrc * per_col + (long long)r.ne * 64;
}

long long assembly_peak(const Request &r)
{
    const long long nln = (long long)(r.N + 1) * (r.N + 2) / 2, d = 2 * nln;
    return (long long)r.ne * 4 * d * d * 8 + (long long)r.ne * 200;  // FP64 operator + geometry
}

}  // namespace

struct tsp_server {
    int device = 0;
    int mode = TSP_SERVE_COLUMNS;
    tsp_multi *multi = nullptr;
    long long budget = 0, used = 0;
    int chunk = 32;
    std::map<int, Tables> tables;
    std::vector<Request> req;
    std::deque<int> queue;   // ids waiting, arrival order
    std::vector<int> running;
    std::vector<int> draining;  // columns mode: finished, results in flight
    std::string err;
    double t_first = -1, t_last = -1;
    long long dof_steps = 0;
    int failed_pending = 0;
    double t_admit = 0, t_run = 0, t_retire = 0;  // host seconds spent in each phase of tick()
};

namespace {

const Tables &tables_for(tsp_server *s, int N, int basis)
{
    const int key = 2 * N + basis;
    auto it = s->tables.find(key);
    if (it != s->tables.end()) return it->second;
    Tables t;
    const int nln = (N + 1) * (N + 2) / 2, nq1 = N + 1, nq2 = nq1 * nq1;
    t.B.resize((size_t)nq2 * nln);
    t.G.resize((size_t)nln * 2 * nq2);
    t.Be.resize((size_t)3 * nq1 * nln);
    t.Ge.resize((size_t)nln * 3 * 2 * nq1);
    t.w2.resize(nq2);
    t.xy2.resize(2 * nq2);
    t.w1.resize(nq1);
    t.x1.resize(nq1);
    tsh_tables_basis(N, basis, t.B.data(), t.G.data(), t.Be.data(), t.Ge.data(), t.w2.data(), t.xy2.data(),
                     t.w1.data(), t.x1.data());
    if (basis != 0) {
        t.minv.resize((size_t)nln * nln);
        tsh_base_invmass(N, basis, t.minv.data());
    }
    return s->tables.emplace(key, std::move(t)).first->second;
}

// tables, and for a non-orthonormal basis its inverse mass, on a new context
int set_space(tsp_context *c, const Request &r, const Tables &t)
{
    int rc = tsp_set_tables(c, r.N, t.B.data(), t.G.data(), t.Be.data(), t.Ge.data(), t.w2.data(), t.w1.data());
    if (rc == TSP_OK && !t.minv.empty()) rc = tsp_set_mass(c, t.minv.data());
    return rc;
}

// columns mode: one multi-column context, first step taken, steps queued by tick
bool admit_columns(tsp_server *s, Request &r)
{
    const Tables &t = tables_for(s, r.N, r.basis);
    r.ncols = columns_for(r.nsrc);
    tsp_context *c = nullptr;
    auto bail = [&](const char *what) {
        r.error = std::string(what) + ": " + (c ? tsp_last_error(c) : "");
        if (c) tsp_destroy(c);
        r.ctx = nullptr;
        r.state = 3;
        return false;
    };
    if (tsp_create(s->device, TSP_FP32C, r.ncols, &c) != TSP_OK) return bail("create");
    if (set_space(c, r, t)) return bail("tables");
    if (tsp_set_mesh(c, r.ne, r.verts.data(), r.neigh.data(), r.neighedge.data(), r.lambda.data(), r.mu.data(),
                     r.rho.data(), r.penalty))
        return bail("mesh");
    if (tsp_release_fp64(c)) return bail("release");
    std::vector<int> cols(r.nsrc);
    for (int k = 0; k < r.nsrc; ++k) cols[k] = k;
    if (tsp_set_sources(c, r.nsrc, cols.data(), r.src_elem.data(), r.src_phi.data(), r.nsteps + 1, r.wavelet.data()))
        return bail("sources");
    if (tsp_set_receivers(c, r.nrec, r.rec_elem.data(), r.rec_phi.data(), r.stride,
                          1 + r.nsteps / std::max(r.stride, 1)))
        return bail("receivers");
    if (tsp_set_dt(c, r.dt) || tsp_first_step(c)) return bail("first step");
    r.ctx = c;
    r.done = 0;
    r.state = 1;
    return true;
}

// columns mode: queue the download of the request's own columns into
// page-locked memory behind its last steps; the context is released once the
// copy has landed (finish_columns)
bool start_retire_columns(Request &r)
{
    r.nsamples = 1 + r.nsteps / std::max(r.stride, 1);
    r.pinned = static_cast<double *>(tsp_host_alloc(sizeof(double) * (size_t)r.nsamples * r.nrec * r.nsrc * 2));
    int ns = 0;
    if (!r.pinned || tsp_get_seismograms_cols(r.ctx, 0, r.nsrc, r.pinned, r.nsamples, &ns, 1) != TSP_OK) {
        r.error = std::string("seismograms: ") + tsp_last_error(r.ctx);
        return false;
    }
    r.state = 4;
    return true;
}

bool copy_landed(Request &r)
{
    int idle = 0;
    return tsp_stream_idle(r.ctx, &idle) != TSP_OK || idle;
}

void finish_columns(Request &r)
{
    tsp_destroy(r.ctx);
    r.ctx = nullptr;
    r.state = 2;
}

// batched mode: the request's multi-column context joins the set
bool admit_batched(tsp_server *s, Request &r)
{
    if (!admit_columns(s, r)) return false;
    int id = -1;
    if (tsp_multi_add(s->multi, r.ctx, r.nsteps, &id) != TSP_OK) {
        r.error = std::string("add: ") + tsp_multi_last_error(s->multi);
        tsp_destroy(r.ctx);
        r.ctx = nullptr;
        r.state = 3;
        return false;
    }
    r.ctx = nullptr;   // owned by the set now
    r.inst.assign(1, id);
    return true;
}

// build the contexts of one request and add them to the set; on failure every
// context made so far is destroyed and the request is marked failed
bool admit(tsp_server *s, Request &r)
{
    if (s->mode == TSP_SERVE_COLUMNS) return admit_columns(s, r);
    if (s->mode == TSP_SERVE_BATCHED) return admit_batched(s, r);
    const Tables &t = tables_for(s, r.N, r.basis);
    const int nln = (r.N + 1) * (r.N + 2) / 2;
    std::vector<tsp_context *> made;
    auto bail = [&](tsp_context *c, const char *what) {
        r.error = std::string(what) + ": " + (c ? tsp_last_error(c) : "");
        for (auto it = made.rbegin(); it != made.rend(); ++it) tsp_destroy(*it);
        r.state = 3;
        return false;
    };
    tsp_context *owner = nullptr;
    if (tsp_create(s->device, TSP_FP32C, 1, &owner) != TSP_OK) return bail(nullptr, "create");
    made.push_back(owner);
    if (set_space(owner, r, t)) return bail(owner, "tables");
    if (tsp_set_mesh(owner, r.ne, r.verts.data(), r.neigh.data(), r.neighedge.data(), r.lambda.data(), r.mu.data(),
                     r.rho.data(), r.penalty))
        return bail(owner, "mesh");
    if (tsp_release_fp64(owner)) return bail(owner, "release");
    for (int k = 1; k < r.nsrc; ++k) {
        tsp_context *c = nullptr;
        if (tsp_share_operator(owner, &c) != TSP_OK) return bail(owner, "share");
        made.push_back(c);
    }
    const int rows = r.nsteps + 1;
    std::vector<double> wav((size_t)rows * 2);
    for (int k = 0; k < r.nsrc; ++k) {
        tsp_context *c = made[k];
        for (int row = 0; row < rows; ++row) {
            wav[2 * (size_t)row] = r.wavelet[((size_t)row * r.nsrc + k) * 2];
            wav[2 * (size_t)row + 1] = r.wavelet[((size_t)row * r.nsrc + k) * 2 + 1];
        }
        const int col = 0;
        if (tsp_set_sources(c, 1, &col, &r.src_elem[k], r.src_phi.data() + (size_t)k * nln, rows, wav.data()))
            return bail(c, "sources");
        if (tsp_set_receivers(c, r.nrec, r.rec_elem.data(), r.rec_phi.data(), r.stride,
                              1 + r.nsteps / std::max(r.stride, 1)))
            return bail(c, "receivers");
        if (tsp_set_dt(c, r.dt) || tsp_set_state(c, 0, nullptr, nullptr) || tsp_first_step(c))
            return bail(c, "first step");
    }
    for (int k = 0; k < r.nsrc; ++k) {
        int id = -1;
        if (tsp_multi_add(s->multi, made[k], r.nsteps, &id) != TSP_OK) {
            r.error = std::string("add: ") + tsp_multi_last_error(s->multi);
            // contexts already handed over are removed, the rest destroyed here
            for (int j = (int)r.inst.size(); j-- > 0;) tsp_multi_remove(s->multi, r.inst[j]);
            for (int j = (int)made.size(); j-- > k;) tsp_destroy(made[j]);
            r.inst.clear();
            r.state = 3;
            return false;
        }
        r.inst.push_back(id);
    }
    r.state = 1;
    return true;
}

// batched mode: the set has synchronised its stream (tsp_multi_run), so the
// request's seismograms are complete; copy them out and drop the instance
bool retire_batched(tsp_server *s, Request &r)
{
    r.nsamples = 1 + r.nsteps / std::max(r.stride, 1);
    r.pinned = static_cast<double *>(tsp_host_alloc(sizeof(double) * (size_t)r.nsamples * r.nrec * r.nsrc * 2));
    tsp_context *c = tsp_multi_context(s->multi, r.inst[0]);
    int ns = 0;
    const bool ok = r.pinned && tsp_get_seismograms_cols(c, 0, r.nsrc, r.pinned, r.nsamples, &ns, 0) == TSP_OK;
    if (!ok) r.error = std::string("seismograms: ") + tsp_last_error(c);
    tsp_multi_remove(s->multi, r.inst[0]);
    r.inst.clear();
    r.state = ok ? 2 : 3;
    return ok;
}

void retire(tsp_server *s, Request &r)
{
    r.nsamples = 1 + r.nsteps / std::max(r.stride, 1);
    const size_t per = (size_t)r.nsamples * r.nrec * 2;
    r.seis.assign(per * r.nsrc, 0.0);
    for (int k = 0; k < r.nsrc; ++k) {
        int ns = 0;
        tsp_get_seismograms(tsp_multi_context(s->multi, r.inst[k]), r.seis.data() + per * k, r.nsamples, &ns);
    }
    // borrowers before the owner
    for (int k = r.nsrc; k-- > 0;) tsp_multi_remove(s->multi, r.inst[k]);
    r.inst.clear();
    r.state = 2;
}

}  // namespace

extern "C" {

int tsp_server_create(int device, double mem_fraction, int chunk, tsp_server **out)
{
    if (!out || !(mem_fraction > 0 && mem_fraction <= 1) || chunk < 1) return TSP_ERR_ARG;
    *out = nullptr;
    long long fr = 0, tot = 0;
    int rc = tsp_device_memory(device, &fr, &tot);
    if (rc != TSP_OK) return rc;
    tsp_server *s = new tsp_server;
    s->device = device;
    s->budget = (long long)(mem_fraction * (double)tot);
    s->chunk = chunk;
    rc = tsp_multi_create(device, &s->multi);
    if (rc != TSP_OK) {
        delete s;
        return rc;
    }
    tsp_multi_set_chunk(s->multi, chunk);
    *out = s;
    return TSP_OK;
}

void tsp_server_destroy(tsp_server *s)
{
    if (!s) return;
    for (auto &r : s->req) {
        if (r.ctx) tsp_destroy(r.ctx);
        tsp_host_free(r.pinned);
    }
    tsp_multi_destroy(s->multi);
    delete s;
}

int tsp_server_set_mode(tsp_server *s, int mode)
{
    if (!s || (mode != TSP_SERVE_INSTANCES && mode != TSP_SERVE_COLUMNS && mode != TSP_SERVE_BATCHED))
        return TSP_ERR_ARG;
    if (!s->req.empty()) {
        s->err = "the mode is fixed once requests were submitted";
        return TSP_ERR_STATE;
    }
    s->mode = mode;
    return TSP_OK;
}

const char *tsp_server_last_error(const tsp_server *s) { return s ? s->err.c_str() : "null server"; }

double tsp_server_now(void) { return clock_s(); }

namespace {
bool request_ok(const tsp_request *q)
{
    return q && q->N >= 1 && q->N <= 8 && q->ne >= 1 && q->nsteps >= 0 && q->nsrc >= 1 && q->nsrc <= 16 &&
           q->nrec >= 0 && q->stride >= 1 && q->dt > 0 && (q->basis == 0 || q->basis == 1);
}
Request sizes_of(const tsp_request *q)
{
    Request r;
    r.N = q->N;
    r.ne = q->ne;
    r.nsteps = q->nsteps;
    r.nsrc = q->nsrc;
    r.nrec = q->nrec;
    r.stride = q->stride;
    r.basis = q->basis;
    return r;
}
}  // namespace

int tsp_server_footprint(const tsp_request *q, int mode, long long *resident, long long *peak)
{
    if (!request_ok(q) || (mode != TSP_SERVE_INSTANCES && mode != TSP_SERVE_COLUMNS && mode != TSP_SERVE_BATCHED))
        return TSP_ERR_ARG;
    const Request r = sizes_of(q);
    if (resident) *resident = footprint_of(r, mode);
    if (peak) *peak = assembly_peak(r);
    return TSP_OK;
}

int tsp_server_submit(tsp_server *s, const tsp_request *q, double arrival, int *id)
{
    if (!s || !q) return TSP_ERR_ARG;
    const int nln = (q->N + 1) * (q->N + 2) / 2;
    if (!request_ok(q)) {
        s->err = "bad request";
        return TSP_ERR_ARG;
    }
    Request r = sizes_of(q);
    r.id = (int)s->req.size();
    r.arrival = arrival;
    r.penalty = q->penalty;
    r.dt = q->dt;
    r.verts.assign(q->verts, q->verts + 6 * (size_t)q->ne);
    r.neigh.assign(q->neigh, q->neigh + 3 * (size_t)q->ne);
    r.neighedge.assign(q->neighedge, q->neighedge + 3 * (size_t)q->ne);
    r.lambda.assign(q->lambda, q->lambda + q->ne);
    r.mu.assign(q->mu, q->mu + q->ne);
    r.rho.assign(q->rho, q->rho + q->ne);
    r.src_elem.assign(q->src_elem, q->src_elem + q->nsrc);
    r.src_phi.assign(q->src_phi, q->src_phi + (size_t)q->nsrc * nln);
    r.wavelet.assign(q->wavelet, q->wavelet + (size_t)(q->nsteps + 1) * q->nsrc * 2);
    r.rec_elem.assign(q->rec_elem, q->rec_elem + q->nrec);
    r.rec_phi.assign(q->rec_phi, q->rec_phi + (size_t)q->nrec * nln);
    r.footprint = footprint_of(r, s->mode);
    if (r.footprint + assembly_peak(r) > s->budget) {
        s->err = "request does not fit the device-memory budget";
        return TSP_ERR_NOMEM;
    }
    s->req.push_back(std::move(r));
    // keep the queue in arrival order
    auto pos = std::upper_bound(s->queue.begin(), s->queue.end(), arrival,
                                [&](double a, int i) { return a < s->req[i].arrival; });
    s->queue.insert(pos, (int)s->req.size() - 1);
    if (id) *id = (int)s->req.size() - 1;
    return TSP_OK;
}

int tsp_server_tick(tsp_server *s, int *completed, int *running, int *queued)
{
    if (!s) return TSP_ERR_ARG;
    const double now = clock_s();
    const double t0 = now;
    // admit every arrived request that fits (first fit in arrival order)
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request &r = s->req[*it];
        if (r.arrival > now) break;
        if (s->used + r.footprint + assembly_peak(r) > s->budget) {
            ++it;
            continue;
        }
        int failed = 0;
        if (admit(s, r)) {
            s->used += r.footprint;
            r.admitted = clock_s();
            s->running.push_back(r.id);
            if (s->t_first < 0) s->t_first = r.admitted;
        } else {
            r.completed = clock_s();   // leaves with its error (tsp_server_result reports it)
            failed = 1;
        }
        it = s->queue.erase(it);
        s->failed_pending += failed;
    }
    int done = s->failed_pending;
    s->failed_pending = 0;
    const double t1 = clock_s();
    s->t_admit += t1 - t0;
    bool nap = false;
    if ((!s->running.empty() || !s->draining.empty()) && s->mode == TSP_SERVE_COLUMNS) {
        // No wait on the whole set: a request whose stream has run dry gets its
        // next two chunks queued (so it never waits for a slower one), and the
        // download of its results is queued right behind its last chunk.
        bool moved = false;
        for (auto it = s->running.begin(); it != s->running.end();) {
            Request &r = s->req[*it];
            int idle = 0;
            if (tsp_stream_idle(r.ctx, &idle) != TSP_OK) {
                s->err = tsp_last_error(r.ctx);
                return TSP_ERR_HIP;
            }
            if (!idle) {
                ++it;
                continue;
            }
            moved = true;
            for (int k = 0; k < 2 && r.done < r.nsteps; ++k) {
                const int n = std::min(s->chunk, r.nsteps - r.done);
                if (tsp_run(r.ctx, n) != TSP_OK) {
                    s->err = tsp_last_error(r.ctx);
                    return TSP_ERR_HIP;
                }
                r.done += n;
            }
            if (r.done < r.nsteps) {
                ++it;
                continue;
            }
            if (!start_retire_columns(r)) {
                tsp_destroy(r.ctx);
                r.ctx = nullptr;
                r.state = 3;
                r.completed = clock_s();
                s->used -= r.footprint;
                ++done;
            } else {
                s->draining.push_back(*it);
            }
            it = s->running.erase(it);
        }
        const double t = clock_s();
        s->t_run += t - t1;
        // results that have landed: release their contexts
        for (auto it = s->draining.begin(); it != s->draining.end();) {
            Request &r = s->req[*it];
            if (copy_landed(r)) {
                finish_columns(r);
                r.completed = t;
                s->used -= r.footprint;
                s->t_last = t;
                s->dof_steps += (long long)r.ne * (r.N + 1) * (r.N + 2) * r.nsrc * r.nsteps;
                ++done;
                moved = true;
                it = s->draining.erase(it);
            } else {
                ++it;
            }
        }
        nap = !moved && done == 0;
    } else if (!s->running.empty()) {
        int active = 0;
        if (tsp_multi_run(s->multi, s->chunk, &active) != TSP_OK) {
            s->err = tsp_multi_last_error(s->multi);
            return TSP_ERR_HIP;
        }
        const double t = clock_s();
        s->t_run += t - t1;
        for (auto it = s->running.begin(); it != s->running.end();) {
            Request &r = s->req[*it];
            int steps = 0, n = 0;
            tsp_multi_info(s->multi, r.inst[0], &steps, &n);
            if (steps >= n) {
                bool ok = true;
                if (s->mode == TSP_SERVE_BATCHED) ok = retire_batched(s, r);
                else retire(s, r);
                r.completed = t;
                if (!ok) {
                    s->used -= r.footprint;
                    ++done;
                    it = s->running.erase(it);
                    continue;
                }
                s->used -= r.footprint;
                s->t_last = t;
                s->dof_steps += (long long)r.ne * (r.N + 1) * (r.N + 2) * r.nsrc * r.nsteps;
                ++done;
                it = s->running.erase(it);
            } else {
                ++it;
            }
        }
    }
    s->t_retire += clock_s() - t1;
    // columns mode with nothing to do until some stream runs dry: yield the core briefly
    if (nap) std::this_thread::sleep_for(std::chrono::microseconds(20));
    if (completed) *completed = done;
    if (running) *running = (int)(s->running.size() + s->draining.size());
    if (queued) *queued = (int)s->queue.size();
    return TSP_OK;
}

int tsp_server_result(tsp_server *s, int id, double *seis, int *nsamples, double *arrival, double *admitted,
                      double *completed)
{
    if (!s || id < 0 || id >= (int)s->req.size()) return TSP_ERR_ARG;
    const Request &r = s->req[id];
    if (r.state == 3) {
        s->err = r.error;
        return TSP_ERR_STATE;
    }
    if (r.state != 2) return TSP_ERR_STATE;
    if (nsamples) *nsamples = r.nsamples;
    if (seis && r.pinned) {
        // [sample][rec][src][2] as downloaded -> [src][sample][rec][2]
        const size_t per = (size_t)r.nsamples * r.nrec * 2;
        for (int smp = 0; smp < r.nsamples; ++smp)
            for (int rec = 0; rec < r.nrec; ++rec) {
                const double *in = r.pinned + ((size_t)smp * r.nrec + rec) * r.nsrc * 2;
                for (int k = 0; k < r.nsrc; ++k) {
                    double *o = seis + per * k + ((size_t)smp * r.nrec + rec) * 2;
                    o[0] = in[2 * k];
                    o[1] = in[2 * k + 1];
                }
            }
    } else if (seis) {
        std::memcpy(seis, r.seis.data(), sizeof(double) * r.seis.size());
    }
    if (arrival) *arrival = r.arrival;
    if (admitted) *admitted = r.admitted;
    if (completed) *completed = r.completed;
    return TSP_OK;
}

int tsp_server_phases(const tsp_server *s, double *admit_seconds, double *run_seconds, double *retire_seconds)
{
    if (!s) return TSP_ERR_ARG;
    if (admit_seconds) *admit_seconds = s->t_admit;
    if (run_seconds) *run_seconds = s->t_run;
    if (retire_seconds) *retire_seconds = s->t_retire - s->t_run;
    return TSP_OK;
}

int tsp_server_stats(const tsp_server *s, double *busy_seconds, long long *dof_steps, long long *mem_used,
                     long long *mem_budget)
{
    if (!s) return TSP_ERR_ARG;
    if (busy_seconds) *busy_seconds = (s->t_first >= 0 && s->t_last >= 0) ? s->t_last - s->t_first : 0.0;
    if (dof_steps) *dof_steps = s->dof_steps;
    if (mem_used) *mem_used = s->used;
    if (mem_budget) *mem_budget = s->budget;
    return TSP_OK;
}

}  // extern "C"
